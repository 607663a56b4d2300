// Evaluation kernels: brute force, deterministic Barnes-Hut and the
// stochastic control-variate estimator, for the gravity, winding-number and
// smooth-distance kernels in 2D and 3D.
//
// Queries arrive as SoA columns [x..., y..., (z...)] (the reference's
// column-major query matrix, src/sbh_gpu.cu:36-41 and :217).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "sbh_device.hpp"
#include "sbh_eval.hpp"
#include "sbh_runtime.hpp"

namespace sbh {
namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ void load_query(const float* qx, const float* qy, const float* qz,
                                           uint32_t i, int dim, float& x, float& y, float& z) {
  x = qx[i];
  y = qy[i];
  z = dim == 3 ? qz[i] : 0.0f;
}

// ---------------------------------------------------------------------------
// Brute force (include/brute_force_gpu.h:18-66): every thread owns kQpt
// queries; the workgroup streams source tiles through LDS and every lane reads
// each source as a broadcast.  Tile sums are kept in fp32 and folded into a
// float64 accumulator per tile.  blockIdx.y splits the sources when there are
// too few query blocks to fill the device; partial sums are then reduced by
// k_reduce_splits in split order.
// ---------------------------------------------------------------------------
constexpr int kQpt = 2;
constexpr int kTile = kThreads;

template <int D, int K>
__global__ __launch_bounds__(kThreads) void k_brute(const float4* __restrict__ pts,
                                                     const float4* __restrict__ nrm, uint32_t n,
                                                     float alpha_k, const float* __restrict__ qx,
                                                     const float* __restrict__ qy,
                                                     const float* __restrict__ qz, uint32_t nq,
                                                     uint32_t per_split, double* __restrict__ partial,
                                                     float* __restrict__ out) {
  __shared__ float4 tile_p[kTile];
  __shared__ float4 tile_n[(K == WINDING_NUMBER) ? kTile : 1];
  const uint32_t tid = threadIdx.x;
  float x[kQpt], y[kQpt], z[kQpt];
  uint32_t qi[kQpt];
#pragma unroll
  for (int j = 0; j < kQpt; ++j) {
    qi[j] = blockIdx.x * (kThreads * kQpt) + j * kThreads + tid;
    uint32_t c = qi[j] < nq ? qi[j] : 0;
    load_query(qx, qy, qz, c, D, x[j], y[j], z[j]);
  }
  const uint32_t s0 = blockIdx.y * per_split;
  const uint32_t s1 = min(n, s0 + per_split);
  double accd[kQpt];
#pragma unroll
  for (int j = 0; j < kQpt; ++j) accd[j] = 0.0;
  for (uint32_t base = s0; base < s1; base += kTile) {
    const uint32_t cnt = min((uint32_t)kTile, s1 - base);
    __syncthreads();
    if (tid < cnt) {
      tile_p[tid] = pts[base + tid];
      if constexpr (K == WINDING_NUMBER) tile_n[tid] = nrm[base + tid];
    }
    __syncthreads();
    float acc[kQpt];
#pragma unroll
    for (int j = 0; j < kQpt; ++j) acc[j] = 0.0f;
    if (cnt == kTile) {
#pragma unroll 8
      for (int s = 0; s < kTile; ++s) {
        float4 sp = tile_p[s];
        float4 sn = (K == WINDING_NUMBER) ? tile_n[s] : sp;
#pragma unroll
        for (int j = 0; j < kQpt; ++j) acc[j] += kernel_fast<D, K>(sp, sn, alpha_k, x[j], y[j], z[j]);
      }
    } else {
      for (uint32_t s = 0; s < cnt; ++s) {
        float4 sp = tile_p[s];
        float4 sn = (K == WINDING_NUMBER) ? tile_n[s] : sp;
#pragma unroll
        for (int j = 0; j < kQpt; ++j) acc[j] += kernel_fast<D, K>(sp, sn, alpha_k, x[j], y[j], z[j]);
      }
    }
#pragma unroll
    for (int j = 0; j < kQpt; ++j) accd[j] += (double)acc[j];
  }
#pragma unroll
  for (int j = 0; j < kQpt; ++j) {
    if (qi[j] < nq) {
      if (partial)
        partial[(size_t)blockIdx.y * nq + qi[j]] = accd[j];
      else
        out[qi[j]] = (float)accd[j];
    }
  }
}

__global__ void k_reduce_splits(const double* partial, uint32_t nsplit, uint32_t nq, float* out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  double s = 0.0;
  for (uint32_t k = 0; k < nsplit; ++k) s += partial[(size_t)k * nq + i];
  out[i] = (float)s;
}

// float64 all-pairs sums (error reference).
template <int D, int K>
__global__ __launch_bounds__(kThreads) void k_brute_f64(const float4* __restrict__ pts,
                                                         const float4* __restrict__ nrm, uint32_t n,
                                                         float alpha, const float* __restrict__ qx,
                                                         const float* __restrict__ qy,
                                                         const float* __restrict__ qz, uint32_t nq,
                                                         uint32_t per_split, double* __restrict__ partial) {
  __shared__ float4 tile_p[kTile];
  __shared__ float4 tile_n[kTile];
  const uint32_t tid = threadIdx.x;
  const uint32_t q = blockIdx.x * kThreads + tid;
  float x, y, z;
  load_query(qx, qy, qz, q < nq ? q : 0, D, x, y, z);
  const uint32_t s0 = blockIdx.y * per_split;
  const uint32_t s1 = min(n, s0 + per_split);
  double acc = 0.0;
  for (uint32_t base = s0; base < s1; base += kTile) {
    const uint32_t cnt = min((uint32_t)kTile, s1 - base);
    __syncthreads();
    if (tid < cnt) {
      tile_p[tid] = pts[base + tid];
      tile_n[tid] = nrm[base + tid];
    }
    __syncthreads();
    for (uint32_t s = 0; s < cnt; ++s) acc += kernel_f64<D, K>(tile_p[s], tile_n[s], alpha, x, y, z);
  }
  if (q < nq) partial[(size_t)blockIdx.y * nq + q] = acc;
}

__global__ void k_reduce_splits_f64(const double* partial, uint32_t nsplit, uint32_t nq,
                                    double* out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  double s = 0.0;
  for (uint32_t k = 0; k < nsplit; ++k) s += partial[(size_t)k * nq + i];
  out[i] = s;
}

// One term of a sum: exact (correctly rounded, reference operation order) or
// fast (hardware sqrt / reciprocal / exp2; alpha_k = alpha * log2(e)).
template <int D, int K, bool EXACT>
__device__ __forceinline__ float term(float4 src, float4 nrm, float alpha, float alpha_k,
                                      float x, float y, float z) {
  if constexpr (EXACT)
    return kernel_exact<D, K>(src, nrm, alpha, x, y, z);
  else
    return kernel_fast<D, K>(src, nrm, alpha_k, x, y, z);
}

// ---------------------------------------------------------------------------
// Deterministic Barnes-Hut (include/barnes_hut.h:23-78 per query,
// include/barnes_hut_gpu.h:19-72 vote).
//
// One wavefront walks the union of its lanes' traversals along the tree's
// ropes (depth-first order, children in cell order): at node v the wave
// descends to v's first child if any active lane (GROUP == 1) or any lane of
// an undecided vote group (GROUP > 1) rejects v, and otherwise follows v's
// rope.  A lane that accepts v while the wave descends adds v's term and
// sleeps until the wave reaches v's rope target, which is where its own walk
// resumes.  Each lane therefore adds exactly the terms of its own walk, in the
// same order.  The node index is wave-uniform, so node records are scalar
// loads, and the records are stored in depth-first order so that the walk
// reads forward through memory.  The opening test is always evaluated with
// correctly rounded operations, so every lane opens exactly the nodes the
// reference walk opens.
// ---------------------------------------------------------------------------
template <int D, int K, int GROUP, bool EXACT>
__global__ __launch_bounds__(kThreads) void k_bh(const float4* __restrict__ nodeC,
                                                  const uint4* __restrict__ nodeL,
                                                  const float4* __restrict__ nodeN,
                                                  const float4* __restrict__ pts,
                                                  const float4* __restrict__ nrm, float alpha,
                                                  float alpha_k, float beta,
                                                  const float* __restrict__ qx,
                                                  const float* __restrict__ qy,
                                                  const float* __restrict__ qz, uint32_t nq,
                                                  const uint32_t* __restrict__ order,
                                                  float* __restrict__ out,
                                                  int32_t* __restrict__ counts, int split_counts) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t slot = blockIdx.x * kThreads + threadIdx.x;
  const bool live = slot < nq;
  uint32_t qi = live ? (order ? order[slot] : slot) : 0;
  float x, y, z;
  load_query(qx, qy, qz, qi, D, x, y, z);
  uint64_t my_group;
  if constexpr (GROUP == 64) {
    my_group = ~0ull;
  } else if constexpr (GROUP == 32) {
    my_group = lane < 32 ? 0x00000000FFFFFFFFull : 0xFFFFFFFF00000000ull;
  } else {
    my_group = 1ull << lane;
  }
  float acc = 0.0f;
  int32_t cnt = 0, cnt_nodes = 0;
  bool active = live;
  uint32_t resume = kEnd;
  uint32_t v = 0;
  while (v != kEnd) {
    v = uniform_u32(v);
    if (!active && live && resume == v) active = true;
    const float4 c = nodeC[v];
    const uint4 l = nodeL[v];
    const bool leaf = (l.w & kLeafBit) != 0;
    bool reject = false;
    if (active && !leaf) reject = !opens_far<D>(c, __uint_as_float(l.x), beta, x, y, z);
    const uint64_t rej = __ballot(reject);
    const bool accept = active && ((rej & my_group) == 0);
    if (accept) {
      if (leaf) {
        const uint32_t p0 = l.y, p1 = l.y + (l.w & ~kLeafBit);
        for (uint32_t j = p0; j < p1; ++j)
          acc += term<D, K, EXACT>(pts[j], nrm[j], alpha, alpha_k, x, y, z);
        cnt += (int32_t)(p1 - p0);
      } else {
        acc += term<D, K, EXACT>(c, nodeN[v], alpha, alpha_k, x, y, z);
        cnt_nodes += 1;
      }
    }
    if (rej == 0) {
      v = l.z;
    } else {
      if (accept) {
        active = false;
        resume = l.z;
      }
      v = l.y;
    }
  }
  if (live) {
    out[qi] = acc;
    if (counts) {
      if (split_counts) {
        counts[qi] = cnt_nodes;
        counts[nq + qi] = cnt;
      } else {
        counts[qi] = cnt + cnt_nodes;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Stochastic control-variate estimator (include/tree_sampler.h:21-113,
// include/tree_sampler_gpu.h:25-140).  The 64 queries of a wavefront share
// their random sample paths (the reference shares them per thread block); the
// path of sample `ss` in subdomain `si` of wave `g` is drawn from Philox block
// {ss, si, g, b} under key = seed: words 0,1 give the sampled point index,
// word 2 + l the Russian-roulette uniform at descent level l.  Each lane runs
// its own roulette against the shared uniforms; the roulette, the weights and
// the opening ratios use the reference's fp32 operations exactly.
//
// At every level the wavefront stages the children of the current node in
// LDS with one coalesced load per lane (children are contiguous in
// breadth-first order) and every lane then reads them as broadcasts.
// ---------------------------------------------------------------------------
constexpr uint32_t kMulti = 1u;  // leaf with more than one point

template <int D, int K, bool EXACT>
__device__ __forceinline__ float node_term(const float4* __restrict__ nodeC,
                                           const uint4* __restrict__ nodeL,
                                           const float4* __restrict__ nodeN,
                                           const float4* __restrict__ pts,
                                           const float4* __restrict__ nrm, float alpha,
                                           float alpha_k, uint32_t v, float x, float y, float z,
                                           int32_t& cnt, int32_t& cnt_nodes) {
  const uint4 l = nodeL[v];
  float s = 0.0f;
  if (l.w & kLeafBit) {
    const uint32_t p0 = l.y, p1 = l.y + (l.w & ~kLeafBit);
    for (uint32_t j = p0; j < p1; ++j) s += term<D, K, EXACT>(pts[j], nrm[j], alpha, alpha_k, x, y, z);
    cnt += (int32_t)(p1 - p0);
  } else {
    s = term<D, K, EXACT>(nodeC[v], nodeN[v], alpha, alpha_k, x, y, z);
    cnt_nodes += 1;
  }
  return s;
}

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int D, int K, bool EXACT>
__global__ __launch_bounds__(kThreads) void k_stochastic(
    const float4* __restrict__ nodeC, const uint4* __restrict__ nodeL,
    const float4* __restrict__ nodeN, const uint32_t* __restrict__ n_start,
    const uint32_t* __restrict__ n_end, const uint32_t* __restrict__ n_skip,
    const float4* __restrict__ pts, const float4* __restrict__ nrm, float alpha, float alpha_k,
    int spl, uint32_t key0, uint32_t key1, const float* __restrict__ qx,
    const float* __restrict__ qy, const float* __restrict__ qz, uint32_t nq,
    const uint32_t* __restrict__ order, float* __restrict__ out, int32_t* __restrict__ counts,
    int split_counts) {
#pragma clang fp contract(off)
  constexpr bool kWN = (K == WINDING_NUMBER);
  __shared__ float4 sSrc[kThreads];
  __shared__ float4 sNrm[kWN ? kThreads : 1];
  __shared__ uint4 sInfo[kThreads];  // {flags, first point, point count, 0}
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wbase = threadIdx.x & ~63u;
  const uint32_t slot = blockIdx.x * kThreads + threadIdx.x;
  const bool live = slot < nq;
  const uint32_t g = uniform_u32(slot >> 6);
  uint32_t qi = live ? (order ? order[slot] : slot) : 0;
  float x, y, z;
  load_query(qx, qy, qz, qi, D, x, y, z);
  int32_t cnt = 0, cnt_nodes = 0;
  const uint4 rl = nodeL[0];
  float result;
  if (rl.w & kLeafBit) {
    result = node_term<D, K, EXACT>(nodeC, nodeL, nodeN, pts, nrm, alpha, alpha_k, 0, x, y, z,
                                    cnt, cnt_nodes);
  } else {
    float bh = 0.0f, mlcv = 0.0f;
    const uint32_t nsub = rl.w, first = rl.y;
    for (uint32_t si = 0; si < nsub; ++si) {
      const uint32_t S = uniform_u32(first + si);
      bh += node_term<D, K, EXACT>(nodeC, nodeL, nodeN, pts, nrm, alpha, alpha_k, S, x, y, z,
                                   cnt, cnt_nodes);
      const uint4 Sl = nodeL[S];
      const uint32_t sstart = n_start[S];
      const uint32_t nS = n_end[S] - sstart;
      float sub_total = 0.0f;
      if (!(Sl.w & kLeafBit)) {
        const float cS = ratio_exact<D>(nodeC[S], __uint_as_float(Sl.x), x, y, z);
        for (int ss = 0; ss < spl; ++ss) {
          Philox4 r0 = philox4x32_10((uint32_t)ss, si, g, 0u, key0, key1);
          const uint32_t idx = uniform_u32(sstart + scale_draw(r0.x, r0.y, nS));
          Philox4 blk = r0;
          uint32_t blk_id = 0;
          float event_prob = 1.0f;
          float wd_all = 0.0f;
          float cur_beta = fmaxf(1.0f, cS);
          bool alive = live;
          uint32_t cur = S;
          uint32_t level = 0;
          uint4 curL = Sl;
          float4 curC = nodeC[S];
          while (!(curL.w & kLeafBit)) {
            if (__ballot(alive) == 0) break;
            const uint32_t wi = level + 2;
            if ((wi >> 2) != blk_id) {
              blk_id = wi >> 2;
              blk = philox4x32_10((uint32_t)ss, si, g, blk_id, key0, key1);
            }
            const uint32_t wsel = wi & 3;
            const uint32_t word = wsel == 0 ? blk.x : wsel == 1 ? blk.y : wsel == 2 ? blk.z : blk.w;
            const float p = rr_uniform(word);
            ++level;
            const float ffr = ratio_exact<D>(curC, __uint_as_float(curL.x), x, y, z);
            const float tp = fminf(cur_beta / ffr, 1.0f);
            if (alive) {
              if (p > tp) {
                event_prob *= 1.0f - tp;
                alive = false;
              } else {
                event_prob *= tp;
              }
            }
            if (__ballot(alive) == 0) break;
            const uint32_t prev = cur;
            const uint32_t f = curL.y, nch = curL.w;
            float cur_result = 0.0f;
            uint32_t pick = 0;
            for (uint32_t b = 0; b < nch; b += 64) {
              // stage children b .. b+63 of prev
              const uint32_t cj = b + lane;
              const bool valid = cj < nch;
              uint32_t cstart = 0xFFFFFFFFu;
              if (valid) {
                const uint32_t c = f + cj;
                const uint4 l = nodeL[c];
                cstart = n_start[c];
                float4 src, nv = make_float4(0.f, 0.f, 0.f, 0.f);
                uint4 info = make_uint4(0u, 0u, 0u, 0u);
                if (l.w & kLeafBit) {
                  const uint32_t np = l.w & ~kLeafBit;
                  src = pts[l.y];
                  if (kWN) nv = nrm[l.y];
                  info = make_uint4(np > 1 ? kMulti : 0u, l.y, np, 0u);  // z: points
                } else {
                  src = nodeC[c];
                  if (kWN) nv = nodeN[c];
                  info = make_uint4(0u, 0u, 0u, 0u);
                }
                sSrc[threadIdx.x] = src;
                if (kWN) sNrm[threadIdx.x] = nv;
                sInfo[threadIdx.x] = info;
              }
              pick += (uint32_t)__popcll(__ballot(valid && cstart <= idx));
              const uint32_t np_lane = valid ? sInfo[threadIdx.x].z : 0u;
              const uint64_t multi = __ballot(np_lane > 1);
              const uint32_t m = min(64u, nch - b);
              wave_lds_sync();
              if (alive) {
                if (multi == 0) {
                  // every child is one source: a node moment or a one-point leaf
#pragma unroll 4
                  for (uint32_t j = 0; j < m; ++j) {
                    const float4 src = sSrc[wbase + j];
                    const float4 nv = kWN ? sNrm[wbase + j] : src;
                    cur_result += term<D, K, EXACT>(src, nv, alpha, alpha_k, x, y, z);
                  }
                } else {
                  for (uint32_t j = 0; j < m; ++j) {
                    const uint4 info = sInfo[wbase + j];
                    if (info.x & kMulti) {
                      for (uint32_t pp = info.y; pp < info.y + info.z; ++pp)
                        cur_result += term<D, K, EXACT>(pts[pp], nrm[pp], alpha, alpha_k, x, y, z);
                    } else {
                      const float4 src = sSrc[wbase + j];
                      const float4 nv = kWN ? sNrm[wbase + j] : src;
                      cur_result += term<D, K, EXACT>(src, nv, alpha, alpha_k, x, y, z);
                    }
                  }
                }
              }
              // term counts of this batch, the same for every lane that evaluated it
              {
                uint32_t pt = np_lane, nd = (valid && np_lane == 0) ? 1u : 0u;
#pragma unroll
                for (int off = 32; off >= 1; off >>= 1) {
                  pt += __shfl_xor(pt, off, 64);
                  nd += __shfl_xor(nd, off, 64);
                }
                if (alive) {
                  cnt += (int32_t)pt;
                  cnt_nodes += (int32_t)nd;
                }
              }
              wave_lds_sync();
            }
            // child holding the sample, then down any single-child chain
            cur = uniform_u32(n_skip[f + pick - 1]);
            curL = nodeL[cur];
            curC = nodeC[cur];
            cur_beta = fmaxf(cur_beta, ffr);
            if (alive) {
              const float prev_result =
                  term<D, K, EXACT>(nodeC[prev], nodeN[prev], alpha, alpha_k, x, y, z);
              const float diff = cur_result - prev_result;
              const float wd = (float)nS / (event_prob * (float)(n_end[prev] - n_start[prev])) * diff;
              wd_all += wd;
            }
          }
          sub_total += wd_all;
        }
      }
      mlcv += sub_total / (float)spl;
    }
    result = bh + mlcv;
  }
  if (live) {
    out[qi] = result;
    if (counts) {
      if (split_counts) {
        counts[qi] = cnt_nodes;
        counts[nq + qi] = cnt;
      } else {
        counts[qi] = cnt + cnt_nodes;
      }
    }
  }
}

// Morton keys of the queries inside the tree's root box (10 bits per axis in
// 3D, 16 in 2D), used to order queries so that wavefronts hold nearby ones.
__device__ __forceinline__ uint32_t spread3(uint32_t v) {
  v &= 0x3FF;
  v = (v | (v << 16)) & 0x030000FF;
  v = (v | (v << 8)) & 0x0300F00F;
  v = (v | (v << 4)) & 0x030C30C3;
  v = (v | (v << 2)) & 0x09249249;
  return v;
}
__device__ __forceinline__ uint32_t spread2(uint32_t v) {
  v &= 0xFFFF;
  v = (v | (v << 8)) & 0x00FF00FF;
  v = (v | (v << 4)) & 0x0F0F0F0F;
  v = (v | (v << 2)) & 0x33333333;
  v = (v | (v << 1)) & 0x55555555;
  return v;
}
__device__ __forceinline__ uint32_t grid_coord(float v, float lo, float hi, float cells) {
  float t = (v - lo) / (hi - lo) * cells;
  t = fminf(fmaxf(t, 0.0f), cells - 1.0f);
  return (uint32_t)t;
}

__global__ void k_morton(const float* qx, const float* qy, const float* qz, uint32_t nq, int dim,
                         float4 lo, float4 hi, uint32_t* keys, uint32_t* idx) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  uint32_t key;
  if (dim == 3) {
    key = (spread3(grid_coord(qz[i], lo.z, hi.z, 1024.f)) << 2) |
          (spread3(grid_coord(qy[i], lo.y, hi.y, 1024.f)) << 1) |
          spread3(grid_coord(qx[i], lo.x, hi.x, 1024.f));
  } else {
    key = (spread2(grid_coord(qy[i], lo.y, hi.y, 65536.f)) << 1) |
          spread2(grid_coord(qx[i], lo.x, hi.x, 65536.f));
  }
  keys[i] = key;
  idx[i] = i;
}

// LSD radix sort of (key, index) pairs, 8 bits per pass: per-block digit
// histograms, a column-major scan of them, then a stable scatter in which each
// wavefront ranks its 256 elements in order.
constexpr int kSortThreads = 256;
constexpr int kSortItems = 4;  // per thread
constexpr int kSortTile = kSortThreads * kSortItems;

__global__ __launch_bounds__(kSortThreads) void k_sort_hist(const uint32_t* keys, uint32_t n,
                                                            int shift, uint32_t nblocks,
                                                            uint32_t* hist) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  uint32_t base = blockIdx.x * kSortTile;
  for (int j = 0; j < kSortItems; ++j) {
    uint32_t i = base + j * kSortThreads + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 0xFF], 1u);
  }
  __syncthreads();
  hist[threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];  // digit-major
}

__global__ __launch_bounds__(kSortThreads) void k_sort_scatter(const uint32_t* keys_in,
                                                               const uint32_t* vals_in,
                                                               uint32_t* keys_out,
                                                               uint32_t* vals_out, uint32_t n,
                                                               int shift, uint32_t nblocks,
                                                               const uint32_t* offsets) {
  // Serial within the block to keep the scatter stable: each of the 256
  // threads owns one digit and walks the tile in order.
  __shared__ uint32_t tk[kSortTile];
  __shared__ uint32_t tv[kSortTile];
  uint32_t base = blockIdx.x * kSortTile;
  uint32_t cnt = min((uint32_t)kSortTile, n - base);
  for (uint32_t j = threadIdx.x; j < cnt; j += kSortThreads) {
    tk[j] = keys_in[base + j];
    tv[j] = vals_in[base + j];
  }
  __syncthreads();
  const uint32_t d = threadIdx.x;
  uint32_t pos = offsets[d * nblocks + blockIdx.x];
  for (uint32_t j = 0; j < cnt; ++j) {
    uint32_t k = tk[j];
    if (((k >> shift) & 0xFF) == d) {
      keys_out[pos] = k;
      vals_out[pos] = tv[j];
      ++pos;
    }
  }
}

inline uint32_t nblk(uint64_t n, uint32_t t) { return (uint32_t)((n + t - 1) / t); }

template <int D, int K>
void launch_brute(const DeviceTree& t, const float* qx, const float* qy, const float* qz,
                  uint32_t nq, float* out, double* partial, uint32_t nsplit, uint32_t per_split,
                  hipStream_t s) {
  dim3 grid(nblk(nq, kThreads * kQpt), nsplit);
  float alpha_k = t.alpha * 1.4426950408889634f;
  k_brute<D, K><<<grid, kThreads, 0, s>>>(t.pts.get(), t.nrm.get(), t.n, alpha_k, qx, qy, qz, nq,
                                          per_split, nsplit > 1 ? partial : nullptr, out);
  SBH_HIP(hipGetLastError());
  if (nsplit > 1) {
    k_reduce_splits<<<nblk(nq, 256), 256, 0, s>>>(partial, nsplit, nq, out);
    SBH_HIP(hipGetLastError());
  }
}

template <int D, int K>
void launch_brute_f64(const DeviceTree& t, const float* qx, const float* qy, const float* qz,
                      uint32_t nq, double* out, double* partial, uint32_t nsplit,
                      uint32_t per_split, hipStream_t s) {
  dim3 grid(nblk(nq, kThreads), nsplit);
  k_brute_f64<D, K><<<grid, kThreads, 0, s>>>(t.pts.get(), t.nrm.get(), t.n, t.alpha, qx, qy, qz,
                                              nq, per_split, partial);
  SBH_HIP(hipGetLastError());
  k_reduce_splits_f64<<<nblk(nq, 256), 256, 0, s>>>(partial, nsplit, nq, out);
  SBH_HIP(hipGetLastError());
}

template <int D, int K, bool EXACT>
void launch_bh(const DeviceTree& t, float beta, int group, const float* qx, const float* qy,
               const float* qz, uint32_t nq, const uint32_t* order, float* out, int32_t* counts,
               int split, hipStream_t s) {
  uint32_t grid = nblk(nq, kThreads);
  const float alpha_k = t.alpha * 1.4426950408889634f;
#define SBH_LAUNCH_BH(G)                                                                      \
  k_bh<D, K, G, EXACT><<<grid, kThreads, 0, s>>>(t.dfsC.get(), t.dfsL.get(), t.dfsN.get(),   \
                                                  t.pts.get(), t.nrm.get(), t.alpha, alpha_k, \
                                                  beta, qx, qy, qz, nq, order, out, counts,   \
                                                  split)
  if (group == 64)
    SBH_LAUNCH_BH(64);
  else if (group == 32)
    SBH_LAUNCH_BH(32);
  else
    SBH_LAUNCH_BH(1);
#undef SBH_LAUNCH_BH
  SBH_HIP(hipGetLastError());
}

template <int D, int K, bool EXACT>
void launch_stoch(const DeviceTree& t, int spl, uint64_t seed, const float* qx, const float* qy,
                  const float* qz, uint32_t nq, const uint32_t* order, float* out, int32_t* counts,
                  int split, hipStream_t s) {
  const float alpha_k = t.alpha * 1.4426950408889634f;
  k_stochastic<D, K, EXACT><<<nblk(nq, kThreads), kThreads, 0, s>>>(
      t.nodeC.get(), t.nodeL.get(), t.nodeN.get(), t.n_start.get(), t.n_end.get(),
      t.n_skip.get(), t.pts.get(), t.nrm.get(), t.alpha, alpha_k, spl, (uint32_t)seed,
      (uint32_t)(seed >> 32), qx, qy, qz, nq, order, out, counts, split);
  SBH_HIP(hipGetLastError());
}

template <int D>
void dispatch_kernel(int kernel, auto&& fn) {
  switch (kernel) {
    case GRAVITY:
      fn.template operator()<D, GRAVITY>();
      break;
    case WINDING_NUMBER:
      fn.template operator()<D, WINDING_NUMBER>();
      break;
    case SMOOTH_DIST:
      fn.template operator()<D, SMOOTH_DIST>();
      break;
    default:
      throw Error(kErrInvalid, "unknown kernel id");
  }
}

void dispatch(int dim, int kernel, auto&& fn) {
  if (dim == 2)
    dispatch_kernel<2>(kernel, fn);
  else
    dispatch_kernel<3>(kernel, fn);
}

}  // namespace

uint32_t brute_split_count(uint32_t n, uint32_t nq) {
  // Enough workgroups to cover every CU several times, splits of at least
  // 8 tiles.
  uint32_t qblocks = nblk(nq, kThreads * kQpt);
  uint32_t want = 4096;
  uint32_t ns = std::max(1u, want / std::max(1u, qblocks));
  uint32_t max_split = std::max(1u, n / (8 * kTile));
  return std::min(ns, max_split);
}

void run_brute(const DeviceTree& t, int kernel, const float* q, uint32_t nq, float* out,
               double* partial, uint32_t nsplit, hipStream_t s) {
  const float* qx = q;
  const float* qy = q + nq;
  const float* qz = t.dim == 3 ? q + 2ull * nq : nullptr;
  uint32_t per_split = nblk(nblk(t.n, nsplit), kTile) * kTile;
  dispatch(t.dim, kernel, [&]<int D, int K>() {
    launch_brute<D, K>(t, qx, qy, qz, nq, out, partial, nsplit, per_split, s);
  });
}

void run_brute_f64(const DeviceTree& t, int kernel, const float* q, uint32_t nq, double* out,
                   double* partial, uint32_t nsplit, hipStream_t s) {
  const float* qx = q;
  const float* qy = q + nq;
  const float* qz = t.dim == 3 ? q + 2ull * nq : nullptr;
  uint32_t per_split = nblk(nblk(t.n, nsplit), kTile) * kTile;
  dispatch(t.dim, kernel, [&]<int D, int K>() {
    launch_brute_f64<D, K>(t, qx, qy, qz, nq, out, partial, nsplit, per_split, s);
  });
}

void run_bh(const DeviceTree& t, int kernel, float beta, int group, const float* q, uint32_t nq,
            const uint32_t* order, float* out, int32_t* counts, int split, bool exact,
            hipStream_t s) {
  const float* qx = q;
  const float* qy = q + nq;
  const float* qz = t.dim == 3 ? q + 2ull * nq : nullptr;
  dispatch(t.dim, kernel, [&]<int D, int K>() {
    if (exact)
      launch_bh<D, K, true>(t, beta, group, qx, qy, qz, nq, order, out, counts, split, s);
    else
      launch_bh<D, K, false>(t, beta, group, qx, qy, qz, nq, order, out, counts, split, s);
  });
}

void run_stochastic(const DeviceTree& t, int kernel, int spl, uint64_t seed, const float* q,
                    uint32_t nq, const uint32_t* order, float* out, int32_t* counts, int split,
                    bool exact, hipStream_t s) {
  const float* qx = q;
  const float* qy = q + nq;
  const float* qz = t.dim == 3 ? q + 2ull * nq : nullptr;
  dispatch(t.dim, kernel, [&]<int D, int K>() {
    if (exact)
      launch_stoch<D, K, true>(t, spl, seed, qx, qy, qz, nq, order, out, counts, split, s);
    else
      launch_stoch<D, K, false>(t, spl, seed, qx, qy, qz, nq, order, out, counts, split, s);
  });
}

size_t sort_scratch_u32(uint32_t nq) {
  uint32_t nb = nblk(nq, kSortTile);
  return 4ull * nq + 2ull * 256 * nb + exclusive_scan_scratch(256u * nb) + 8;
}

void query_order(const DeviceTree& t, const float* q, uint32_t nq, uint32_t* order,
                 uint32_t* scratch, hipStream_t s) {
  const float* qx = q;
  const float* qy = q + nq;
  const float* qz = t.dim == 3 ? q + 2ull * nq : nullptr;
  uint32_t* k0 = scratch;
  uint32_t* k1 = scratch + nq;
  uint32_t* v1 = scratch + 2ull * nq;
  uint32_t nb = nblk(nq, kSortTile);
  uint32_t* hist = scratch + 4ull * nq;
  uint32_t* offs = hist + 256ull * nb;
  uint32_t* scan_scratch = offs + 256ull * nb;
  float4 lo = make_float4(t.lo[0], t.lo[1], t.lo[2], 0.f);
  float4 hi = make_float4(t.hi[0], t.hi[1], t.dim == 3 ? t.hi[2] : 1.f, 0.f);
  k_morton<<<nblk(nq, 256), 256, 0, s>>>(qx, qy, qz, nq, t.dim, lo, hi, k0, order);
  SBH_HIP(hipGetLastError());
  const int passes = 4;  // 32-bit keys
  uint32_t* kin = k0;
  uint32_t* vin = order;
  uint32_t* kout = k1;
  uint32_t* vout = v1;
  for (int pss = 0; pss < passes; ++pss) {
    k_sort_hist<<<nb, kSortThreads, 0, s>>>(kin, nq, 8 * pss, nb, hist);
    SBH_HIP(hipGetLastError());
    exclusive_scan(hist, offs, 256u * nb, nullptr, scan_scratch, s);
    k_sort_scatter<<<nb, kSortThreads, 0, s>>>(kin, vin, kout, vout, nq, 8 * pss, nb, offs);
    SBH_HIP(hipGetLastError());
    std::swap(kin, kout);
    std::swap(vin, vout);
  }
  // After an even number of passes the sorted indices are back in `order`.
}

}  // namespace sbh
