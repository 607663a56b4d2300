// Evaluation on a double-build tree: the reference compiled with USE_DOUBLE
// (include/utils.h:13-29).  Brute force, deterministic Barnes-Hut and the
// stochastic control-variate estimator in float64, with the operation order of
// the float build's exact mode, so that each result equals the reference
// double build's for the same tree, queries and seed.
//
// These are the precision forms, not the throughput path: one lane per
// query.  The Barnes-Hut walk keeps the union-walk structure of the fp32
// kernel (one rope walk per wavefront over 64-byte depth-first records, node
// index wave-uniform); the estimator walks each lane's paths on their own,
// with a per-lane memo of its level-1 nodes' diffs.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sbh_device.hpp"
#include "sbh_eval.hpp"
#include "sbh_runtime.hpp"

namespace sbh {
namespace {

constexpr int kThreads = 256;
constexpr int kTile = 256;  // brute force: points staged per LDS tile

inline uint32_t nblk(uint64_t n, uint32_t t) { return (uint32_t)((n + t - 1) / t); }

struct TreeViewD {
  const dvec4* pts;
  const dvec4* nrm;
  const dvec4* nodeC;
  const dvec4* nodeN;
  const double* diam;
  const uint4* nodeL;  // {fp32 diameter bits, first child | first point, rope, info}
  const uint32_t* n_start;
  const uint32_t* n_end;
  const uint32_t* n_skip;
  const WalkRecordD* dfsR;  // depth-first walk records (Barnes-Hut)
  const dvec4* dfsN;
  double alpha;
};

TreeViewD view_double(const DeviceTree& t) {
  return TreeViewD{t.pts64.get(),  t.nrm64.get(),   t.nodeC64.get(), t.nodeN64.get(),
                   t.n_diam64.get(), t.nodeL.get(), t.n_start.get(), t.n_end.get(),
                   t.n_skip.get(),  t.dfsR64.get(),  t.dfsN64.get(),  t.alpha64};
}

struct LaneD {
  double value;
  int32_t points;
  int32_t nodes;
};

__device__ __forceinline__ void store_result(uint32_t i, uint32_t nq, const LaneD& r, double* out,
                                             int32_t* counts, int split) {
  out[i] = r.value;
  if (counts) {
    if (split) {
      counts[i] = r.nodes;
      counts[nq + i] = r.points;
    } else {
      counts[i] = r.points + r.nodes;
    }
  }
}

// Adds node v's terms to acc: its points in order if v is a leaf, else its
// moment.  Returns the number of point terms (0: one node term).
template <int D, int K>
__device__ __forceinline__ uint32_t node_accumulate_d(const TreeViewD& t, uint32_t v, double x,
                                                      double y, double z, double& acc) {
#pragma clang fp contract(off)
  const uint4 l = t.nodeL[v];
  if (l.w & kLeafBit) {
    const uint32_t np = l.w & ~kLeafBit;
    for (uint32_t j = l.y; j < l.y + np; ++j)
      acc = acc + kernel_double<D, K>(t.pts[j], t.nrm[j], t.alpha, x, y, z);
    return np;
  }
  acc = acc + kernel_double<D, K>(t.nodeC[v], t.nodeN[v], t.alpha, x, y, z);
  return 0;
}

// Term counts from node_accumulate_d's return value (counters stay in registers).
__device__ __forceinline__ void count_terms(uint32_t np, int32_t& pts, int32_t& nodes) {
  pts += (int32_t)np;
  nodes += np == 0 ? 1 : 0;
}

// ---------------------------------------------------------------------------
// Brute force (include/brute_force_gpu.h:45-63): every point in tree order,
// Kahan-compensated, one query per lane; points staged through LDS tiles.
// ---------------------------------------------------------------------------
template <int D, int K>
__global__ __launch_bounds__(kThreads) void k_brute_double(TreeViewD t, uint32_t n,
                                                           const double* qx, const double* qy,
                                                           const double* qz, uint32_t nq,
                                                           double* out) {
#pragma clang fp contract(off)
  constexpr bool kWN = (K == WINDING_NUMBER);
  __shared__ dvec4 sp[kTile];
  __shared__ dvec4 sn[kWN ? kTile : 1];
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  const bool live = i < nq;
  const double x = live ? qx[i] : 0.0;
  const double y = live ? qy[i] : 0.0;
  const double z = (live && D == 3) ? qz[i] : 0.0;
  double total = 0.0, c = 0.0;
  for (uint32_t base = 0; base < n; base += kTile) {
    const uint32_t m = min((uint32_t)kTile, n - base);
    __syncthreads();
    if (threadIdx.x < m) {
      sp[threadIdx.x] = t.pts[base + threadIdx.x];
      if constexpr (kWN) sn[threadIdx.x] = t.nrm[base + threadIdx.x];
    }
    __syncthreads();
    // unrolled so that the terms of neighbouring points overlap; the
    // compensated additions stay in point order
#pragma unroll 4
    for (uint32_t j = 0; j < m; ++j) {
      const double kv = kernel_double<D, K>(sp[j], kWN ? sn[j] : sp[j], t.alpha, x, y, z);
      const double yk = kv - c;
      const double tt = total + yk;
      c = (tt - total) - yk;
      total = tt;
    }
  }
  if (live) out[i] = total;
}

// ---------------------------------------------------------------------------
// Deterministic Barnes-Hut (include/barnes_hut.h:23-78 per query,
// include/barnes_hut_gpu.h:19-72 vote) over the depth-first records and
// their ropes: the wavefront walks the union of its lanes' traversals; a lane
// that accepts a node while the wave descends adds its term and sleeps until
// the wave reaches the node's rope target.  Each lane adds exactly the terms
// of its own walk, in the same order.  The node index is wave-uniform: each
// step is one 64-byte scalar load, issued as soon as the next node is known.
// ---------------------------------------------------------------------------
template <int D, int K, int GROUP>
__global__ __launch_bounds__(kThreads) void k_bh_double(TreeViewD t, double beta,
                                                        const double* qx, const double* qy,
                                                        const double* qz, uint32_t nq,
                                                        const uint32_t* order, double* out,
                                                        int32_t* counts, int split) {
#pragma clang fp contract(off)
  constexpr bool kWN = (K == WINDING_NUMBER);
  const uint32_t slot = blockIdx.x * kThreads + threadIdx.x;
  const bool live = slot < nq;
  const uint32_t i = live ? (order ? order[slot] : slot) : 0u;
  const double x = live ? qx[i] : 0.0;
  const double y = live ? qy[i] : 0.0;
  const double z = (live && D == 3) ? qz[i] : 0.0;
  const uint32_t lane = threadIdx.x & 63;
  uint64_t my_group;
  if constexpr (GROUP == 64)
    my_group = ~0ull;
  else if constexpr (GROUP == 32)
    my_group = lane < 32 ? 0x00000000FFFFFFFFull : 0xFFFFFFFF00000000ull;
  else
    my_group = 1ull << lane;
  const WalkRecordD* __restrict__ rec = t.dfsR;
  double acc = 0.0;
  int32_t npts = 0, nnodes = 0;
  bool active = live;
  uint32_t resume = kEnd;
  uint32_t v = 0;
  WalkRecordD r = rec[0];
  while (v != kEnd) {
    v = uniform_u32(v);
    active = active || (live && resume == v);
    const bool leaf = (r.info & kLeafBit) != 0;
    bool reject = false;
    if (!leaf) {
      // ||q - c|| / diam >= beta, decided as the correctly rounded quotient
      // decides it: squared distances more than 1e-12 (relative) from
      // (beta diam)^2 settle it without the square root and the divide,
      // whose roundings are ~1e-16; the rest (and a (beta diam)^2 outside
      // the normal range) take the exact formula.
      const double dx = x - r.c.x, dy = y - r.c.y, dz = (D == 3) ? (z - r.c.z) : 0.0;
      const double d2 = sqnorm_exact<D>(dx, dy, dz);
      const double b2 = (beta * r.diam) * (beta * r.diam);
      const bool in_range = b2 >= 1e-290 && b2 <= 1e290;
      const bool far = in_range && d2 >= b2 * (1.0 + 1e-12);
      const bool near = in_range && d2 <= b2 * (1.0 - 1e-12);
      bool open_ok = far;
      if (!far && !near) open_ok = __builtin_sqrt(d2) / r.diam >= beta;
      reject = active && !open_ok;
    }
    const uint64_t rej = __ballot(reject);
    const bool accept = active && (GROUP == 1 ? !reject : (rej & my_group) == 0);
    const uint32_t np = leaf ? (r.info & ~kLeafBit) : 0u;
    if (accept) {
      if (np > 1) {
        for (uint32_t j = r.first; j < r.first + np; ++j)
          acc = acc + kernel_double<D, K>(t.pts[j], t.nrm[j], t.alpha, x, y, z);
      } else {
        // an internal node's moment, or a one-point leaf's point (its record holds it)
        const dvec4 nv = kWN ? t.dfsN[v] : r.c;
        acc = acc + kernel_double<D, K>(r.c, nv, t.alpha, x, y, z);
      }
      npts += (int32_t)np;
      nnodes += leaf ? 0 : 1;
    }
    uint32_t next;
    if (rej == 0) {
      next = r.rope;
    } else {
      resume = accept ? r.rope : resume;
      active = active && !accept;
      next = r.first;
    }
    next = uniform_u32(next);
    if (next != kEnd) r = rec[next];
    v = next;
  }
  if (live) store_result(i, nq, LaneD{acc, npts, nnodes}, out, counts, split);
}

// ---------------------------------------------------------------------------
// Stochastic control-variate estimator (include/tree_sampler.h:21-113,
// include/tree_sampler_gpu.h:25-140).  The path of sample `ss` in subdomain
// `si` of group g comes from Philox block {ss, si, g, b} under key = seed:
// words 0,1 give the sampled point index, word 2 + l the roulette uniform at
// level l (all 32 bits: rr_uniform_double).  g is the query's 64-query
// wavefront (the paths the wave shares) or the query itself (per-query
// paths).  Level 0 of a subdomain is the same for every sample (its roulette
// keeps every path, weight 1) and is evaluated once.
//
// Level 1 is memoised: the path node there is one of the subdomain's <= 64
// children, and its diff (children's sum minus its own term) depends only on
// that node and the query, so each lane keeps the diff of every level-1 node
// it has evaluated, with the term counts, in a scratch column of its own
// (memo.diff[j * stride + slot], j = the child's index in the subdomain; a
// wavefront's lanes have consecutive slots, so these accesses coalesce).  A
// reused diff is the recomputed one bit for bit, and every sample still adds
// its own weight in sample order.
// ---------------------------------------------------------------------------
// Index of the child of a node (children f .. f + nch - 1, in point order)
// whose point range holds point idx: the last child starting at or before
// idx, by bisection.
__device__ __forceinline__ uint32_t child_holding(const uint32_t* n_start, uint32_t f,
                                                  uint32_t nch, uint32_t idx) {
  uint32_t lo = 0, hi = nch;  // n_start[f + lo] <= idx
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (n_start[f + mid] <= idx)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

struct Memo {
  double* diff;  // nullptr: no memo
  int32_t* cnt;  // optional: {point terms, node terms} per entry
  uint32_t stride;
  uint32_t slot;
};

template <int D, int K>
__device__ LaneD stochastic_double(const TreeViewD& t, int spl, uint32_t key0, uint32_t key1,
                                   uint32_t g, double x, double y, double z, const Memo& memo) {
#pragma clang fp contract(off)
  int32_t npts = 0, nnodes = 0;
  const uint4 rl = t.nodeL[0];
  if (rl.w & kLeafBit) {
    double v = 0.0;
    count_terms(node_accumulate_d<D, K>(t, 0, x, y, z, v), npts, nnodes);
    return {v, npts, nnodes};
  }
  double bh = 0.0, mlcv = 0.0;
  const uint32_t nsub = rl.w, first = rl.y;
  for (uint32_t si = 0; si < nsub; ++si) {
    const uint32_t S = first + si;
    count_terms(node_accumulate_d<D, K>(t, S, x, y, z, bh), npts, nnodes);
    const uint4 Sl = t.nodeL[S];
    if (Sl.w & kLeafBit) continue;  // the subdomain's points were summed exactly
    const uint32_t sstart = t.n_start[S];
    const uint32_t nS = t.n_end[S] - sstart;
    const double cS = ratio_double<D>(t.nodeC[S], t.diam[S], x, y, z);
    const uint32_t f0 = Sl.y, nch0 = Sl.w;
    double r0sum = 0.0;
    int32_t p0 = 0, n0 = 0;
    for (uint32_t j = 0; j < nch0; ++j)
      count_terms(node_accumulate_d<D, K>(t, f0 + j, x, y, z, r0sum), p0, n0);
    const double prev0 = kernel_double<D, K>(t.nodeC[S], t.nodeN[S], t.alpha, x, y, z);
    const double wd0 = (double)nS / (1.0 * (double)nS) * (r0sum - prev0);
    const bool use_memo = memo.diff != nullptr && nch0 <= 64;
    uint64_t memo_valid = 0;  // level-1 nodes of this subdomain whose diff is stored
    double sub_total = 0.0;
    for (int ss = 0; ss < spl; ++ss) {
      const Philox4 r0 = philox4x32_10((uint32_t)ss, si, g, 0u, key0, key1);
      const uint32_t idx = sstart + scale_draw(r0.x, r0.y, nS);
      Philox4 blk = r0;
      uint32_t blk_id = 0;
      double event_prob = 1.0;
      double wd_all = 0.0;
      double cur_beta = fmax(cS, fmax(1.0, cS));
      const uint32_t pick = child_holding(t.n_start, f0, nch0, idx);
      uint32_t cur = t.n_skip[f0 + pick];
      wd_all = wd_all + wd0;
      npts += p0;
      nnodes += n0;
      uint32_t level = 1;
      uint4 curL = t.nodeL[cur];
      bool at1 = true;  // the path node is the level-1 node S's child `pick` leads to
      while (!(curL.w & kLeafBit)) {
        const uint32_t wi = level + 2;
        if ((wi >> 2) != blk_id) {
          blk_id = wi >> 2;
          blk = philox4x32_10((uint32_t)ss, si, g, blk_id, key0, key1);
        }
        const uint32_t wsel = wi & 3;
        const uint32_t word = wsel == 0 ? blk.x : wsel == 1 ? blk.y : wsel == 2 ? blk.z : blk.w;
        const double p = rr_uniform_double(word);
        ++level;
        const double ffr = ratio_double<D>(t.nodeC[cur], t.diam[cur], x, y, z);
        const double tp = fmin(cur_beta / ffr, 1.0);
        if (p > tp) {
          event_prob = event_prob * (1.0 - tp);
          break;
        }
        event_prob = event_prob * tp;
        const uint32_t prev = cur;
        const uint32_t f = curL.y, nch = curL.w;
        const uint32_t pk = child_holding(t.n_start, f, nch, idx);
        double diff;
        const bool memo_here = use_memo && at1;
        const size_t me = (size_t)pick * memo.stride + memo.slot;
        if (memo_here && ((memo_valid >> pick) & 1ull)) {
          diff = memo.diff[me];
          if (memo.cnt) {
            npts += memo.cnt[2 * me];
            nnodes += memo.cnt[2 * me + 1];
          }
        } else {
          double cur_result = 0.0;
          int32_t pc = 0, nc = 0;
          for (uint32_t j = 0; j < nch; ++j)
            count_terms(node_accumulate_d<D, K>(t, f + j, x, y, z, cur_result), pc, nc);
          const double prev_result =
              kernel_double<D, K>(t.nodeC[prev], t.nodeN[prev], t.alpha, x, y, z);
          diff = cur_result - prev_result;
          npts += pc;
          nnodes += nc;
          if (memo_here) {
            memo.diff[me] = diff;
            if (memo.cnt) {
              memo.cnt[2 * me] = pc;
              memo.cnt[2 * me + 1] = nc;
            }
            memo_valid |= 1ull << pick;
          }
        }
        at1 = false;
        cur = t.n_skip[f + pk];
        curL = t.nodeL[cur];
        cur_beta = fmax(cur_beta, ffr);
        const double n_prev = (double)(t.n_end[prev] - t.n_start[prev]);
        wd_all = wd_all + (double)nS / (event_prob * n_prev) * diff;
      }
      sub_total = sub_total + wd_all;
    }
    mlcv = mlcv + sub_total / (double)spl;
  }
  return {bh + mlcv, npts, nnodes};
}

template <int D, int K>
__global__ __launch_bounds__(kThreads) void k_stochastic_double(TreeViewD t, int spl,
                                                                uint32_t key0, uint32_t key1,
                                                                int per_query, const double* qx,
                                                                const double* qy,
                                                                const double* qz, uint32_t nq,
                                                                const uint32_t* order,
                                                                double* out, int32_t* counts,
                                                                int split, double* memo_diff,
                                                                int32_t* memo_cnt) {
  const uint32_t slot = blockIdx.x * kThreads + threadIdx.x;
  if (slot >= nq) return;
  const uint32_t i = order ? order[slot] : slot;
  const double x = qx[i], y = qy[i], z = D == 3 ? qz[i] : 0.0;
  // paths: the query's own, or those of its wavefront in launch order
  const uint32_t g = per_query ? i : slot / 64u;
  const Memo memo{memo_diff, memo_cnt, nq, slot};
  store_result(i, nq, stochastic_double<D, K>(t, spl, key0, key1, g, x, y, z, memo), out, counts,
               split);
}

// float copy of the queries, for their spatial order
__global__ void k_queries_to_float(const double* q, float* qf, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) qf[i] = (float)q[i];
}

__global__ void k_smooth_distance_double(double* w, uint32_t n, double alpha) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) w[i] = -log(w[i]) / alpha;
}

// Dispatch over (dim, kernel).
template <template <int, int> class L, typename... Args>
void dispatch(int dim, int kernel, Args&&... args) {
  auto by_kernel = [&](auto dtag) {
    constexpr int D = decltype(dtag)::value;
    switch (kernel) {
      case GRAVITY: L<D, GRAVITY>::run(args...); break;
      case WINDING_NUMBER: L<D, WINDING_NUMBER>::run(args...); break;
      case SMOOTH_DIST: L<D, SMOOTH_DIST>::run(args...); break;
      case HELMHOLTZ_X: L<D, HELMHOLTZ_X>::run(args...); break;
      case HELMHOLTZ_Y: L<D, HELMHOLTZ_Y>::run(args...); break;
      case GRAVITY_GRAD_X: L<D, GRAVITY_GRAD_X>::run(args...); break;
      case GRAVITY_GRAD_Y: L<D, GRAVITY_GRAD_Y>::run(args...); break;
      case GRAVITY_GRAD_Z:
        if constexpr (D == 3) L<D, GRAVITY_GRAD_Z>::run(args...);
        break;
      default: throw Error(kErrInvalid, "unknown kernel");
    }
  };
  if (dim == 2)
    by_kernel(std::integral_constant<int, 2>{});
  else
    by_kernel(std::integral_constant<int, 3>{});
}

template <int D, int K>
struct BruteL {
  static void run(const DeviceTree& t, const double* q, uint32_t nq, double* out,
                  hipStream_t s) {
    k_brute_double<D, K><<<nblk(nq, kThreads), kThreads, 0, s>>>(
        view_double(t), t.n, q, q + nq, D == 3 ? q + 2ull * nq : nullptr, nq, out);
  }
};

template <int D, int K>
struct BhL {
  static void run(const DeviceTree& t, double beta, int group, const double* q, uint32_t nq,
                  const uint32_t* order, double* out, int32_t* counts, int split,
                  hipStream_t s) {
    const double* qz = D == 3 ? q + 2ull * nq : nullptr;
    const uint32_t nb = nblk(nq, kThreads);
    if (group == 64)
      k_bh_double<D, K, 64><<<nb, kThreads, 0, s>>>(view_double(t), beta, q, q + nq, qz, nq,
                                                   order, out, counts, split);
    else if (group == 32)
      k_bh_double<D, K, 32><<<nb, kThreads, 0, s>>>(view_double(t), beta, q, q + nq, qz, nq,
                                                   order, out, counts, split);
    else
      k_bh_double<D, K, 1><<<nb, kThreads, 0, s>>>(view_double(t), beta, q, q + nq, qz, nq,
                                                  order, out, counts, split);
  }
};

template <int D, int K>
struct StochL {
  static void run(const DeviceTree& t, int spl, uint64_t seed, bool per_query, const double* q,
                  uint32_t nq, const uint32_t* order, double* out, int32_t* counts, int split,
                  double* memo_diff, int32_t* memo_cnt, hipStream_t s) {
    k_stochastic_double<D, K><<<nblk(nq, kThreads), kThreads, 0, s>>>(
        view_double(t), spl, (uint32_t)seed, (uint32_t)(seed >> 32), per_query ? 1 : 0, q,
        q + nq, D == 3 ? q + 2ull * nq : nullptr, nq, order, out, counts, split, memo_diff,
        memo_cnt);
  }
};

}  // namespace

void run_brute_double(const DeviceTree& t, int kernel, const double* q, uint32_t nq, double* out,
                      hipStream_t s) {
  dispatch<BruteL>(t.dim, kernel, t, q, nq, out, s);
  SBH_HIP(hipGetLastError());
}

void run_bh_double(const DeviceTree& t, int kernel, double beta, int group, const double* q,
                   uint32_t nq, const uint32_t* order, double* out, int32_t* counts, int split,
                   hipStream_t s) {
  dispatch<BhL>(t.dim, kernel, t, beta, group, q, nq, order, out, counts, split, s);
  SBH_HIP(hipGetLastError());
}

void run_stochastic_double(const DeviceTree& t, int kernel, int spl, uint64_t seed,
                           bool per_query, const double* q, uint32_t nq, const uint32_t* order,
                           double* out, int32_t* counts, int split, double* memo_diff,
                           int32_t* memo_cnt, hipStream_t s) {
  dispatch<StochL>(t.dim, kernel, t, spl, seed, per_query, q, nq, order, out, counts, split,
                   memo_diff, memo_cnt, s);
  SBH_HIP(hipGetLastError());
}

void query_order_double(const DeviceTree& t, const double* q, uint32_t nq, float* qf,
                        uint32_t* order, uint32_t* scratch, hipStream_t s) {
  const uint32_t n = nq * (uint32_t)t.dim;
  k_queries_to_float<<<nblk(n, 256), 256, 0, s>>>(q, qf, n);
  SBH_HIP(hipGetLastError());
  query_order(t, qf, nq, order, scratch, s);
}

void run_smooth_distance_double(double* w, uint32_t n, double alpha, hipStream_t s) {
  k_smooth_distance_double<<<nblk(n, 256), 256, 0, s>>>(w, n, alpha);
  SBH_HIP(hipGetLastError());
}

}  // namespace sbh
