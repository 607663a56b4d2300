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
#include "sbh_walk.hpp"

namespace sbh {
namespace {

constexpr int kThreads = 256;


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
#pragma unroll 8  // loads in flight ahead of the (ordered) additions
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
#pragma unroll 8  // loads in flight ahead of the (ordered) additions
  for (uint32_t k = 0; k < nsplit; ++k) s += partial[(size_t)k * nq + i];
  out[i] = s;
}

template <int D, int K, int GROUP, bool EXACT, bool COUNT>
__global__ __launch_bounds__(kThreads) void k_bh(TreeView t, float beta,
                                                  const float* __restrict__ qx,
                                                  const float* __restrict__ qy,
                                                  const float* __restrict__ qz, uint32_t nq,
                                                  const uint32_t* __restrict__ order,
                                                  float* __restrict__ out,
                                                  int32_t* __restrict__ counts, int split_counts) {
  const uint32_t slot = blockIdx.x * kThreads + threadIdx.x;
  const bool live = slot < nq;
  const uint32_t qi = live ? (order ? order[slot] : slot) : 0;
  float x, y, z;
  load_query(qx, qy, qz, qi, D, x, y, z);
  const LaneResult r =
      bh_walk<D, K, GROUP, EXACT, COUNT, GROUP == 1 && !EXACT>(t, beta, live, x, y, z);
  if (live) {
    out[qi] = r.value;
    if (COUNT && counts) {
      if (split_counts) {
        counts[qi] = r.nodes;
        counts[nq + qi] = r.points;
      } else {
        counts[qi] = r.points + r.nodes;
      }
    }
  }
}

// Barnes-Hut per query for launches too small to fill the GPU: the walk is
// split over a frontier of disjoint subtrees (blockIdx.y = task, heaviest
// first), each task walking one subtree for 64 queries, and the partial sums
// are added in depth-first order afterwards.  A query whose opening
// test accepts an ancestor of a subtree takes that ancestor's moment (in the
// first task below it) and nothing else, as the whole walk would.  Fast terms
// only: the association of the sum changes.  With COUNT the tasks also count
// their terms (pcnt: {nodes, points} per task).
template <int D, int K, bool COUNT>
__global__ __launch_bounds__(kThreads) void k_bh_split(TreeView t,
                                                        const WalkTask* __restrict__ tasks,
                                                        float beta,
                                                        const float* __restrict__ qx,
                                                        const float* __restrict__ qy,
                                                        const float* __restrict__ qz, uint32_t nq,
                                                        const uint32_t* __restrict__ order,
                                                        float* __restrict__ partial,
                                                        int2* __restrict__ pcnt) {
  const uint32_t slot = blockIdx.x * kThreads + threadIdx.x;
  const bool live = slot < nq;
  const uint32_t qi = live ? (order ? order[slot] : slot) : 0;
  float x, y, z;
  load_query(qx, qy, qz, qi, D, x, y, z);
  const LaneResult r = bh_walk_task<D, K, COUNT>(t, tasks, blockIdx.y, beta, live, x, y, z);
  if (live) {
    const size_t e = (size_t)load_invariant(&tasks[blockIdx.y].part, 0) * nq + slot;
    partial[e] = r.value;
    if constexpr (COUNT) pcnt[e] = make_int2(r.nodes, r.points);
  }
}

__global__ void k_reduce_bh_split(const float* __restrict__ partial,
                                  const int2* __restrict__ pcnt, uint32_t nsplit, uint32_t nq,
                                  const uint32_t* __restrict__ order, float* __restrict__ out,
                                  int32_t* __restrict__ counts, int split_counts) {
  const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= nq) return;
  const uint32_t qi = order ? order[slot] : slot;
  out[qi] = sum_bh_parts(partial, nsplit, nq, slot);
  if (counts) {
    int32_t nodes = 0, points = 0;
#pragma unroll 8
    for (uint32_t k = 0; k < nsplit; ++k) {
      const int2 c = pcnt[(size_t)k * nq + slot];
      nodes += c.x;
      points += c.y;
    }
    if (split_counts) {
      counts[qi] = nodes;
      counts[nq + qi] = points;
    } else {
      counts[qi] = nodes + points;
    }
  }
}

template <int D, int K, bool EXACT>
__global__ __launch_bounds__(kThreads) void k_stochastic(TreeView t, int spl, uint32_t key0,
                                                          uint32_t key1,
                                                          const float* __restrict__ qx,
                                                          const float* __restrict__ qy,
                                                          const float* __restrict__ qz,
                                                          uint32_t nq,
                                                          const uint32_t* __restrict__ order,
                                                          float* __restrict__ out,
                                                          int32_t* __restrict__ counts,
                                                          int split_counts) {
  __shared__ float4 sSrc[kThreads];
  __shared__ float4 sNrm[K == WINDING_NUMBER ? kThreads : 1];  // normals: winding number only
  __shared__ uint4 sRec[kThreads];
  __shared__ uint2 sAux[kThreads];
  const uint32_t wbase = threadIdx.x & ~63u;
  const uint32_t slot = blockIdx.x * kThreads + threadIdx.x;
  const bool live = slot < nq;
  const uint32_t g = uniform_u32(slot >> 6);
  const uint32_t qi = live ? (order ? order[slot] : slot) : 0;
  float x, y, z;
  load_query(qx, qy, qz, qi, D, x, y, z);
  const SamplerLds lds{sSrc + wbase, sNrm + (K == WINDING_NUMBER ? wbase : 0), sRec + wbase,
                       sAux + wbase};
  const LaneResult r =
      stochastic_estimate<D, K, EXACT>(t, spl, key0, key1, g, live, x, y, z, lds, counts != nullptr);
  if (live) {
    out[qi] = r.value;
    if (counts) {
      if (split_counts) {
        counts[qi] = r.nodes;
        counts[nq + qi] = r.points;
      } else {
        counts[qi] = r.points + r.nodes;
      }
    }
  }
}

// Fast-mode estimator with sorted samples and the per-level node cache, for
// trees whose nodes have at most 64 children.  WORK: the counts are the work
// counts of stochastic_estimate_sorted (kernel terms, path levels).
template <int D, int K, bool COUNT, bool WORK = false>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(7, 7))) void k_stochastic_sorted(
    TreeView t, int spl, uint32_t key0, uint32_t key1, const float* __restrict__ qx,
    const float* __restrict__ qy, const float* __restrict__ qz, uint32_t nq,
    const uint32_t* __restrict__ order, float* __restrict__ out, int32_t* __restrict__ counts,
    int split_counts) {
  constexpr uint32_t kWaves = kThreads / 64;
  __shared__ float4 sSrc[kThreads];
  __shared__ float4 sNrm[K == WINDING_NUMBER ? kThreads : 1];  // normals: winding number only
  __shared__ uint2 sKeys[kWaves * chunk_for<K>()];
  __shared__ uint4 sNext[kWaves * (kCacheLevels + 1) * 3];
  const uint32_t wave = threadIdx.x >> 6;
  const uint32_t wbase = threadIdx.x & ~63u;
  const uint32_t slot = blockIdx.x * kThreads + threadIdx.x;
  const bool live = slot < nq;
  const uint32_t g = uniform_u32(slot >> 6);
  const uint32_t qi = live ? (order ? order[slot] : slot) : 0;
  float x, y, z;
  load_query(qx, qy, qz, qi, D, x, y, z);
  const SortedLds lds{sSrc + wbase, sNrm + (K == WINDING_NUMBER ? wbase : 0),
                      sKeys + wave * chunk_for<K>(),
                      sNext + wave * (kCacheLevels + 1) * 3};
  const LaneResult r = stochastic_estimate_sorted<D, K, COUNT, false, WORK>(t, spl, key0, key1,
                                                                          g, live, x, y, z, lds);
  if (live) {
    out[qi] = r.value;
    if (WORK) {
      counts[qi] = r.points;
      counts[nq + qi] = r.nodes;
    } else if (COUNT && counts) {
      if (split_counts) {
        counts[qi] = r.nodes;
        counts[nq + qi] = r.points;
      } else {
        counts[qi] = r.points + r.nodes;
      }
    }
  }
}

// The sorted-sample estimator split into tasks: one per (64 queries, range
// of root subdomains) (blockIdx.y = the range's task in sub_tasks, heaviest
// first), each writing its Barnes-Hut and control-variate parts to the
// range's slot; k_reduce_stoch_split adds the parts in subdomain order.  Same
// sample paths (keyed by the wavefront's slot block and the subdomain) and
// terms; the fast terms' fused accumulation rounds differently across the
// split, so values agree with the whole estimate to fp32 rounding.
template <int D, int K, bool COUNT>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(7, 7))) void
k_stochastic_sorted_split(TreeView t, int spl, uint32_t key0, uint32_t key1,
                          const float* __restrict__ qx, const float* __restrict__ qy,
                          const float* __restrict__ qz, uint32_t nq,
                          const uint32_t* __restrict__ order, const uint4* __restrict__ tasks,
                          float2* __restrict__ partial, int2* __restrict__ pcnt) {
  constexpr uint32_t kWaves = kThreads / 64;
  __shared__ float4 sSrc[kThreads];
  __shared__ float4 sNrm[K == WINDING_NUMBER ? kThreads : 1];  // normals: winding number only
  __shared__ uint2 sKeys[kWaves * chunk_for<K>()];
  __shared__ uint4 sNext[kWaves * (kCacheLevels + 1) * 3];
  const uint32_t wave = threadIdx.x >> 6;
  const uint32_t wbase = threadIdx.x & ~63u;
  const uint32_t slot = blockIdx.x * kThreads + threadIdx.x;
  const bool live = slot < nq;
  const uint32_t g = uniform_u32(slot >> 6);
  const uint32_t qi = live ? (order ? order[slot] : slot) : 0;
  float x, y, z;
  load_query(qx, qy, qz, qi, D, x, y, z);
  const SortedLds lds{sSrc + wbase, sNrm + (K == WINDING_NUMBER ? wbase : 0),
                      sKeys + wave * chunk_for<K>(),
                      sNext + wave * (kCacheLevels + 1) * 3};
  float bh = 0.0f;
  const uint4 tk = load_invariant(tasks, blockIdx.y);
  const LaneResult r = stochastic_estimate_sorted<D, K, COUNT, true>(
      t, spl, key0, key1, g, live, x, y, z, lds, tk.x, tk.y, &bh);
  if (live) {
    const size_t e = (size_t)tk.z * nq + slot;
    partial[e] = make_float2(bh, r.value);
    if constexpr (COUNT) pcnt[e] = make_int2(r.nodes, r.points);
  }
}

__global__ void k_reduce_stoch_split(const float2* __restrict__ partial,
                                     const int2* __restrict__ pcnt, uint32_t nsplit,
                                     uint32_t nq, const uint32_t* __restrict__ order,
                                     float* __restrict__ out, int32_t* __restrict__ counts,
                                     int split_counts) {
  const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= nq) return;
  const uint32_t qi = order ? order[slot] : slot;
  out[qi] = sum_stochastic_parts(partial, nsplit, nq, slot);
  if (counts) {
    int32_t nodes = 0, points = 0;
#pragma unroll 8
    for (uint32_t k = 0; k < nsplit; ++k) {
      const int2 c = pcnt[(size_t)k * nq + slot];
      nodes += c.x;
      points += c.y;
    }
    if (split_counts) {
      counts[qi] = nodes;
      counts[nq + qi] = points;
    } else {
      counts[qi] = nodes + points;
    }
  }
}

// One sample path per query: thread = query, path key = the query's index in
// the caller's array.
template <int D, int K, bool EXACT>
__global__ __launch_bounds__(kThreads) void k_stochastic_query(
    TreeView t, int spl, uint32_t key0, uint32_t key1, const float* __restrict__ qx,
    const float* __restrict__ qy, const float* __restrict__ qz, uint32_t nq,
    const uint32_t* __restrict__ order, float* __restrict__ out, int32_t* __restrict__ counts,
    int split_counts) {
  const uint32_t slot = blockIdx.x * kThreads + threadIdx.x;
  if (slot >= nq) return;
  const uint32_t qi = order ? order[slot] : slot;
  float x, y, z;
  load_query(qx, qy, qz, qi, D, x, y, z);
  const LaneResult r = stochastic_estimate_query<D, K, EXACT>(t, spl, key0, key1, qi, x, y, z);
  out[qi] = r.value;
  if (counts) {
    if (split_counts) {
      counts[qi] = r.nodes;
      counts[nq + qi] = r.points;
    } else {
      counts[qi] = r.points + r.nodes;
    }
  }
}

// Hilbert-curve keys of the queries inside the tree's root box: 21 bits per
// axis in 3D, 31 in 2D (63- and 62-bit keys), so that even a root box
// thousands of times wider than the dense core still orders the core's
// queries.  Along a Hilbert curve consecutive cells are face neighbours, so
// the 64 queries of a wavefront stay together in space more often than along
// a Morton curve (on the 2^20-query Plummer workload the union of a
// wavefront's Barnes-Hut walks is ~10 % shorter).  Skilling's transpose form
// (J. Skilling, "Programming the Hilbert curve", AIP Conf. Proc. 707, 2004):
// undo the excess work of the recursive reflections and rotations, Gray-code
// the result, then interleave the axes' bits, axis 0 first.
template <int D>
__device__ __forceinline__ uint64_t hilbert_key(uint32_t* X, int bits) {
  const uint32_t M = 1u << (bits - 1);
  for (uint32_t Q = M; Q > 1; Q >>= 1) {
    const uint32_t P = Q - 1;
#pragma unroll
    for (int i = 0; i < D; ++i) {
      if (X[i] & Q) {
        X[0] ^= P;  // invert
      } else {      // exchange the low bits of X[0] and X[i]
        const uint32_t t = (X[0] ^ X[i]) & P;
        X[0] ^= t;
        X[i] ^= t;
      }
    }
  }
#pragma unroll
  for (int i = 1; i < D; ++i) X[i] ^= X[i - 1];
  uint32_t t = 0;
  for (uint32_t Q = M; Q > 1; Q >>= 1)
    if (X[D - 1] & Q) t ^= Q - 1;
#pragma unroll
  for (int i = 0; i < D; ++i) X[i] ^= t;
  uint64_t key = 0;
  for (int b = bits - 1; b >= 0; --b) {
#pragma unroll
    for (int i = 0; i < D; ++i) key = (key << 1) | ((X[i] >> b) & 1u);
  }
  return key;
}

__device__ __forceinline__ uint32_t grid_coord(float v, float lo, float hi, double cells) {
  double t = ((double)v - (double)lo) / ((double)hi - (double)lo) * cells;
  t = fmin(fmax(t, 0.0), cells - 1.0);
  return (uint32_t)t;
}

__device__ __forceinline__ uint64_t curve_key(const float* qx, const float* qy, const float* qz,
                                              uint32_t i, int dim, float4 lo, float4 hi) {
  if (dim == 3) {
    const double cells = 2097152.0;
    uint32_t X[3] = {grid_coord(qx[i], lo.x, hi.x, cells), grid_coord(qy[i], lo.y, hi.y, cells),
                     grid_coord(qz[i], lo.z, hi.z, cells)};
    return hilbert_key<3>(X, 21);
  }
  const double cells = 2147483648.0;
  uint32_t X[2] = {grid_coord(qx[i], lo.x, hi.x, cells), grid_coord(qy[i], lo.y, hi.y, cells)};
  return hilbert_key<2>(X, 31);
}

__global__ void k_curve_keys(const float* qx, const float* qy, const float* qz, uint32_t nq,
                             int dim, float4 lo, float4 hi, uint64_t* keys, uint32_t* idx) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  keys[i] = curve_key(qx, qy, qz, i, dim, lo, hi);
  idx[i] = i;
}

// The launch order is the queries sorted by (curve key, input index): a
// total order, so every correct sort yields the same permutation (the LSD
// radix passes below are stable, and the keys start in input order).
//
// Up to kFusedSortMax queries, where launch latency dominates, the sort takes
// ten launches: the histograms are zeroed, the keys computed (one per thread,
// over the whole GPU) with their first digit's per-tile histogram, then eight
// digit passes.  A pass's workgroup derives its tile's output offsets from
// the pass's per-tile histograms (every digit's total before it plus the same
// digit in earlier tiles), ranks its elements stably (per wavefront, chunk by
// chunk, with running digit counts in LDS), scatters them, and adds each
// element's next digit to the next pass's histogram of the tile it lands in.
constexpr uint32_t kFrTile = 4096;
constexpr uint32_t kFusedSortMax = 1u << 18;

__global__ void k_zero_u32n(uint32_t* p, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    p[i] = 0;
}

// Keys and input indices, and the first pass's histograms hist[0][digit][tile].
__global__ __launch_bounds__(256) void k_fr_keys(const float* qx, const float* qy,
                                                 const float* qz, uint32_t nq, int dim,
                                                 float4 lo, float4 hi, uint64_t* keys,
                                                 uint32_t* idx, uint32_t* hist0, uint32_t nb) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < nq) {
    const uint64_t key = curve_key(qx, qy, qz, i, dim, lo, hi);
    keys[i] = key;
    idx[i] = i;
    atomicAdd(&h[(uint32_t)key & 0xFFu], 1u);
  }
  __syncthreads();
  // a 256-key block lies in one tile
  const uint32_t c = h[threadIdx.x];
  if (c) atomicAdd(&hist0[threadIdx.x * nb + blockIdx.x * 256 / kFrTile], c);
}

// THREADS threads per tile (the order does not depend on it): 1024 for a
// launch of its own, 256 where the sort shares the GPU with running work
// (serving), whose workgroups then fit in the CUs' gaps.  Dynamic LDS: the
// pass's histograms, 1 KB per tile.
template <int THREADS>
__global__ __launch_bounds__(THREADS) void k_fr_pass(const uint64_t* __restrict__ kin,
                                                        const uint32_t* __restrict__ vin,
                                                        uint64_t* __restrict__ kout,
                                                        uint32_t* __restrict__ vout, uint32_t n,
                                                        int shift, uint32_t nb,
                                                        const uint32_t* __restrict__ hist,
                                                        uint32_t* __restrict__ hist_next) {
  constexpr int kItems = (int)kFrTile / THREADS;
  constexpr int kChunks = kItems;  // 64-element chunks per wavefront
  static_assert(THREADS >= 256 && kItems * THREADS == (int)kFrTile, "tile shape");
  __shared__ uint32_t wcnt[THREADS / 64][256];  // running, then exclusive, per-wave counts
  __shared__ uint32_t boff[256];
  __shared__ uint32_t scan[4];  // digit totals of wavefronts 0..3
  extern __shared__ uint32_t hn[];  // the pass's, then the next pass's [digit][tile]
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t b = blockIdx.x;
  // the tile: chunks of 64 consecutive elements, wavefront w owning chunks
  // [w kChunks, (w + 1) kChunks) in order (loaded first, so that their
  // latency overlaps the offsets below)
  uint64_t key[kItems];
  uint32_t val[kItems], rank[kItems];
#pragma unroll
  for (int m = 0; m < kItems; ++m) {
    const uint32_t i = b * kFrTile + (wave * kChunks + m) * 64 + lane;
    const bool ok = i < n;
    key[m] = ok ? kin[i] : ~0ull;
    val[m] = ok ? vin[i] : 0u;
  }
  // this tile's output start for digit tid, from the pass's histograms
  // (staged in hn with one round of coalesced loads)
  for (uint32_t e = tid; e < 256 * nb; e += THREADS) hn[e] = hist[e];
  for (uint32_t e = tid; e < (THREADS / 64) * 256; e += THREADS) (&wcnt[0][0])[e] = 0;
  __syncthreads();
  const bool digit_thread = tid < 256;  // wavefronts 0..3 own a digit per thread
  uint32_t tot = 0, pre = 0, incl = 0;
  if (digit_thread) {
    for (uint32_t j = 0; j < nb; ++j) {
      const uint32_t c = hn[tid * nb + j];
      tot += c;
      pre += j < b ? c : 0u;
    }
    // inclusive scan of the digit totals within the wavefront
    incl = tot;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t v = (uint32_t)__shfl_up((int)incl, off, 64);
      incl += lane >= (uint32_t)off ? v : 0u;
    }
    if (lane == 63) scan[wave] = incl;  // the wavefront's total
  }
  __syncthreads();
  if (hist_next)
    for (uint32_t e = tid; e < 256 * nb; e += THREADS) hn[e] = 0;
  if (digit_thread) {
    uint32_t before = 0;  // totals of the wavefronts of smaller digits
    for (uint32_t w = 0; w < wave; ++w) before += scan[w];
    boff[tid] = pre + before + incl - tot;  // + totals of the smaller digits
  }
  // stable ranks: a lane's rank among the chunk's lanes holding its digit
  // (eight ballots), after the wavefront's earlier chunks (running counts)
  const uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
  for (int m = 0; m < kItems; ++m) {
    const uint32_t i = b * kFrTile + (wave * kChunks + m) * 64 + lane;
    const bool ok = i < n;
    const uint32_t d = (uint32_t)(key[m] >> shift) & 0xFFu;
    uint64_t peers = __ballot(ok);
#pragma unroll
    for (int bit = 0; bit < 8; ++bit) {
      const bool on = (d >> bit) & 1u;
      const uint64_t set = __ballot(on);
      peers &= on ? set : ~set;
    }
    const uint32_t below = (uint32_t)__popcll(peers & lt);
    const uint32_t before = wcnt[wave][d];
    rank[m] = before + below;
    if (ok && below == 0) wcnt[wave][d] = before + (uint32_t)__popcll(peers);
  }
  __syncthreads();
  if (digit_thread) {
    uint32_t run = 0;
#pragma unroll
    for (int w = 0; w < THREADS / 64; ++w) {
      const uint32_t c = wcnt[w][tid];
      wcnt[w][tid] = run;
      run += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < kItems; ++m) {
    const uint32_t i = b * kFrTile + (wave * kChunks + m) * 64 + lane;
    if (i >= n) continue;
    const uint32_t d = (uint32_t)(key[m] >> shift) & 0xFFu;
    const uint32_t pos = boff[d] + wcnt[wave][d] + rank[m];
    kout[pos] = key[m];
    vout[pos] = val[m];
    if (hist_next) {
      const uint32_t dn = (uint32_t)(key[m] >> (shift + 8)) & 0xFFu;
      atomicAdd(&hn[dn * nb + pos / kFrTile], 1u);
    }
  }
  if (hist_next) {
    __syncthreads();
    for (uint32_t e = tid; e < 256 * nb; e += THREADS) {
      const uint32_t c = hn[e];
      if (c) atomicAdd(&hist_next[e], c);
    }
  }
}

// LSD radix sort of (key, index) pairs, 8 bits per pass: per-block digit
// histograms, a digit-major scan of them, then a stable scatter.  In the
// scatter each wavefront ranks its 64-element chunks with eight ballots (a
// lane's rank among the lanes holding the same digit), the block adds the
// chunks' digit counts in chunk order, and every element writes to its
// digit's block offset + the counts of earlier chunks + its rank.
constexpr int kSortThreads = 256;
constexpr int kSortItems = 4;  // per thread
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kSortChunks = kSortTile / 64;
static_assert(kSortThreads == 256, "the scatter's prefix step takes one thread per digit");

__global__ __launch_bounds__(kSortThreads) void k_sort_hist(const uint64_t* keys, uint32_t n,
                                                            int shift, uint32_t nblocks,
                                                            uint32_t* hist) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  uint32_t base = blockIdx.x * kSortTile;
  for (int j = 0; j < kSortItems; ++j) {
    uint32_t i = base + j * kSortThreads + threadIdx.x;
    if (i < n) atomicAdd(&h[(uint32_t)(keys[i] >> shift) & 0xFFu], 1u);
  }
  __syncthreads();
  hist[threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];  // digit-major
}

__global__ __launch_bounds__(kSortThreads) void k_sort_scatter(const uint64_t* keys_in,
                                                               const uint32_t* vals_in,
                                                               uint64_t* keys_out,
                                                               uint32_t* vals_out, uint32_t n,
                                                               int shift, uint32_t nblocks,
                                                               const uint32_t* offsets) {
  __shared__ uint32_t cnt[kSortChunks * 256];   // [chunk][digit] counts
  __shared__ uint32_t start[kSortChunks * 256];  // [chunk][digit] output starts
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (uint32_t i = threadIdx.x; i < kSortChunks * 256; i += kSortThreads) cnt[i] = 0;
  __syncthreads();
  const uint32_t tile = blockIdx.x * kSortTile;
  uint64_t key[kSortItems];
  uint32_t val[kSortItems], dig[kSortItems], rank[kSortItems];
  bool ok[kSortItems];
#pragma unroll
  for (int m = 0; m < kSortItems; ++m) {
    const uint32_t c = wave + (kSortThreads / 64) * m;  // chunks in tile order
    const uint32_t i = tile + c * 64 + lane;
    ok[m] = i < n;
    key[m] = ok[m] ? keys_in[i] : 0ull;
    val[m] = ok[m] ? vals_in[i] : 0u;
    const uint32_t d = (uint32_t)(key[m] >> shift) & 0xFFu;
    dig[m] = d;
    uint64_t peers = __ballot(ok[m]);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t set = __ballot(bit);
      peers &= bit ? set : ~set;
    }
    rank[m] = (uint32_t)__popcll(peers & ((1ull << lane) - 1ull));
    if (ok[m] && rank[m] == 0) cnt[c * 256 + d] = (uint32_t)__popcll(peers);
  }
  __syncthreads();
  {
    const uint32_t d = threadIdx.x;  // one thread per digit
    uint32_t run = offsets[d * nblocks + blockIdx.x];
    for (uint32_t c = 0; c < (uint32_t)kSortChunks; ++c) {
      start[c * 256 + d] = run;
      run += cnt[c * 256 + d];
    }
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < kSortItems; ++m) {
    if (!ok[m]) continue;
    const uint32_t c = wave + (kSortThreads / 64) * m;
    const uint32_t pos = start[c * 256 + dig[m]] + rank[m];
    keys_out[pos] = key[m];
    vals_out[pos] = val[m];
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
               int split, float* partial, WalkSplit ws, hipStream_t s) {
  uint32_t grid = nblk(nq, kThreads);
  const TreeView tv = view_of(t);
  const uint32_t nsplit = ws.n;
  if (!EXACT && nsplit > 1) {
    // partial: nsplit * nq floats, then (with counts) nsplit * nq int2
    int2* pcnt = reinterpret_cast<int2*>(partial + (((size_t)nsplit * nq + 1) & ~(size_t)1));
    if (counts)
      k_bh_split<D, K, true><<<dim3(grid, nsplit), kThreads, 0, s>>>(
          tv, ws.tasks, beta, qx, qy, qz, nq, order, partial, pcnt);
    else
      k_bh_split<D, K, false><<<dim3(grid, nsplit), kThreads, 0, s>>>(
          tv, ws.tasks, beta, qx, qy, qz, nq, order, partial, pcnt);
    SBH_HIP(hipGetLastError());
    k_reduce_bh_split<<<nblk(nq, 256), 256, 0, s>>>(partial, pcnt, nsplit, nq, order, out,
                                                     counts, split);
    SBH_HIP(hipGetLastError());
    return;
  }
#define SBH_LAUNCH_BH(G, COUNT)                                                      \
  k_bh<D, K, G, EXACT, COUNT><<<grid, kThreads, 0, s>>>(tv, beta, qx, qy, qz, nq, order, out, \
                                                         counts, split)
  if (group == 64)
    SBH_LAUNCH_BH(64, true);
  else if (group == 32)
    SBH_LAUNCH_BH(32, true);
  else if (counts || EXACT)
    SBH_LAUNCH_BH(1, true);
  else  // the fast per-query walk without counts (evaluation and serving)
    SBH_LAUNCH_BH(1, false);
#undef SBH_LAUNCH_BH
  SBH_HIP(hipGetLastError());
}

template <int D, int K, bool EXACT>
void launch_stoch(const DeviceTree& t, int spl, uint64_t seed, bool per_query, const float* qx,
                  const float* qy, const float* qz, uint32_t nq, const uint32_t* order,
                  float* out, int32_t* counts, int split, float* partial, uint32_t nsplit,
                  hipStream_t s, bool work = false) {
  // nodes have at most 2^(dim * split_log) children
  const bool small_fanout = t.dim * t.split_log <= 6;
  if (work) {
    if constexpr (!EXACT && K <= SMOOTH_DIST) {
      if (small_fanout && !per_query && counts) {
        k_stochastic_sorted<D, K, true, true><<<nblk(nq, kThreads), kThreads, 0, s>>>(
            view_of(t), spl, (uint32_t)seed, (uint32_t)(seed >> 32), qx, qy, qz, nq, order, out,
            counts, 1);
        SBH_HIP(hipGetLastError());
        return;
      }
    }
    throw Error(kErrInvalid,
                "work counts cover the fast sorted-sample estimator (shared paths, at most 64 "
                "children per node, gravity / winding number / smooth distance) with counts");
  }
  if (!EXACT && small_fanout && !per_query && nsplit > 1) {
    // partial: nsplit * nq float2, then (with counts) nsplit * nq int2
    float2* part = reinterpret_cast<float2*>(partial);
    int2* pcnt = reinterpret_cast<int2*>(part + (size_t)nsplit * nq);
    const dim3 grid(nblk(nq, kThreads), nsplit);
    SBH_REQUIRE(nsplit == t.n_sub_tasks, "split estimator: task count changed");
    const uint4* tasks = t.sub_tasks.get();
    if (counts)
      k_stochastic_sorted_split<D, K, true><<<grid, kThreads, 0, s>>>(
          view_of(t), spl, (uint32_t)seed, (uint32_t)(seed >> 32), qx, qy, qz, nq, order, tasks,
          part, pcnt);
    else
      k_stochastic_sorted_split<D, K, false><<<grid, kThreads, 0, s>>>(
          view_of(t), spl, (uint32_t)seed, (uint32_t)(seed >> 32), qx, qy, qz, nq, order, tasks,
          part, pcnt);
    SBH_HIP(hipGetLastError());
    k_reduce_stoch_split<<<nblk(nq, 256), 256, 0, s>>>(part, pcnt, nsplit, nq, order, out,
                                                        counts, split);
    SBH_HIP(hipGetLastError());
    return;
  }
  if (per_query)
    k_stochastic_query<D, K, EXACT><<<nblk(nq, kThreads), kThreads, 0, s>>>(
        view_of(t), spl, (uint32_t)seed, (uint32_t)(seed >> 32), qx, qy, qz, nq, order, out,
        counts, split);
  else if (!EXACT && small_fanout && counts)
    k_stochastic_sorted<D, K, true><<<nblk(nq, kThreads), kThreads, 0, s>>>(
        view_of(t), spl, (uint32_t)seed, (uint32_t)(seed >> 32), qx, qy, qz, nq, order, out,
        counts, split);
  else if (!EXACT && small_fanout)
    k_stochastic_sorted<D, K, false><<<nblk(nq, kThreads), kThreads, 0, s>>>(
        view_of(t), spl, (uint32_t)seed, (uint32_t)(seed >> 32), qx, qy, qz, nq, order, out,
        counts, split);
  else
    k_stochastic<D, K, EXACT><<<nblk(nq, kThreads), kThreads, 0, s>>>(
        view_of(t), spl, (uint32_t)seed, (uint32_t)(seed >> 32), qx, qy, qz, nq, order, out,
        counts, split);
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
    case HELMHOLTZ_X:
      fn.template operator()<D, HELMHOLTZ_X>();
      break;
    case HELMHOLTZ_Y:
      fn.template operator()<D, HELMHOLTZ_Y>();
      break;
    case GRAVITY_GRAD_X:
      fn.template operator()<D, GRAVITY_GRAD_X>();
      break;
    case GRAVITY_GRAD_Y:
      fn.template operator()<D, GRAVITY_GRAD_Y>();
      break;
    case GRAVITY_GRAD_Z:
      if constexpr (D == 3) {
        fn.template operator()<D, GRAVITY_GRAD_Z>();
        break;
      }
      [[fallthrough]];
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

// Split Barnes-Hut launches: wavefronts wanted (task count x 64-query blocks)
// and the most partial sums (task count x queries, 64 MiB of floats).
constexpr uint32_t kSplitWaves = 16u * 1024u;
constexpr size_t kSplitParts = size_t(1) << 24;

WalkSplit bh_split(const DeviceTree& t, uint32_t nq, int group, bool exact) {
  // fewer than four wavefronts per SIMD (1024 SIMDs): split the walk over the
  // smallest frontier that makes kSplitWaves wavefronts, else the largest one
  // whose partial sums stay within kSplitParts values.  Counted launches split
  // the same way (the values must not depend on it).
  WalkSplit ws;
  if (exact || group != 1 || t.is_double || t.walk_frontiers.empty() || nq >= 4u * 1024u * 64u)
    return ws;
  const uint32_t waves = nblk(nq, kWave);
  for (const auto& f : t.walk_frontiers) {
    if (ws.n > 1 && (size_t)f.second * nq > kSplitParts) break;
    ws.n = f.second;
    ws.tasks = t.walk_tasks.get() + f.first;
    if ((size_t)f.second * waves >= kSplitWaves) break;
  }
  return ws;
}

uint32_t bh_split_count(const DeviceTree& t, uint32_t nq, int group, bool exact, bool counts) {
  (void)counts;
  return bh_split(t, nq, group, exact).n;
}

void run_bh(const DeviceTree& t, int kernel, float beta, int group, const float* q, uint32_t nq,
            const uint32_t* order, float* out, int32_t* counts, int split, bool exact,
            float* partial, hipStream_t s) {
  const float* qx = q;
  const float* qy = q + nq;
  const float* qz = t.dim == 3 ? q + 2ull * nq : nullptr;
  const WalkSplit ws = partial ? bh_split(t, nq, group, exact) : WalkSplit{};
  dispatch(t.dim, kernel, [&]<int D, int K>() {
    if (exact)
      launch_bh<D, K, true>(t, beta, group, qx, qy, qz, nq, order, out, counts, split, partial,
                            WalkSplit{}, s);
    else
      launch_bh<D, K, false>(t, beta, group, qx, qy, qz, nq, order, out, counts, split, partial,
                             ws, s);
  });
}

uint32_t stochastic_split_count(const DeviceTree& t, uint32_t nq, bool per_query, bool exact) {
  // the sorted-sample estimator: one task per 64-query block and range of
  // root subdomains holding one internal subdomain (the tree's sub_tasks).
  // Smaller tasks, heaviest dispatched first, leave a shorter tail of
  // wavefronts than whole walks, whose lengths vary with the queries'
  // positions.
  (void)nq;
  if (exact || per_query || t.is_double || t.dim * t.split_log > 6 || t.n_sub_tasks < 2)
    return 1;
  return t.n_sub_tasks;
}

void run_stochastic(const DeviceTree& t, int kernel, int spl, uint64_t seed, bool per_query,
                    const float* q, uint32_t nq, const uint32_t* order, float* out,
                    int32_t* counts, int split, bool exact, float* partial, hipStream_t s,
                    bool work) {
  const float* qx = q;
  const float* qy = q + nq;
  const float* qz = t.dim == 3 ? q + 2ull * nq : nullptr;
  const uint32_t ns = partial && !work ? stochastic_split_count(t, nq, per_query, exact) : 1;
  dispatch(t.dim, kernel, [&]<int D, int K>() {
    if (exact)
      launch_stoch<D, K, true>(t, spl, seed, per_query, qx, qy, qz, nq, order, out, counts,
                               split, partial, 1, s, work);
    else
      launch_stoch<D, K, false>(t, spl, seed, per_query, qx, qy, qz, nq, order, out, counts,
                                split, partial, ns, s, work);
  });
}

namespace {
__global__ void k_fill_i32(int32_t* p, uint32_t n, int32_t v) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// -log(w) / alpha in place (python/example.py:79).
__global__ void k_smooth_distance(float* w, uint32_t n, float alpha) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) w[i] = -logf(w[i]) / alpha;
}
}  // namespace

void fill_i32(int32_t* p, uint32_t n, int32_t v, hipStream_t s) {
  if (n == 0) return;
  k_fill_i32<<<nblk(n, 256), 256, 0, s>>>(p, n, v);
  SBH_HIP(hipGetLastError());
}

void run_smooth_distance(float* w, uint32_t n, float alpha, hipStream_t s) {
  k_smooth_distance<<<nblk(n, 256), 256, 0, s>>>(w, n, alpha);
  SBH_HIP(hipGetLastError());
}

size_t sort_scratch_u32(uint32_t nq) {
  // two 64-bit key buffers, one value buffer, then histograms and offsets
  // and scan scratch (radix passes) or the per-tile histograms of the eight
  // passes (fused passes)
  const uint32_t nb = nblk(nq, kSortTile);
  const size_t radix = 2ull * 256 * nb + exclusive_scan_scratch(256u * nb);
  const size_t fused = nq <= kFusedSortMax ? 8ull * 256 * nblk(nq, kFrTile) : 0;
  return 4ull * nq + (size_t)nq + std::max(radix, fused) + 8;
}

void query_order(const DeviceTree& t, const float* q, uint32_t nq, uint32_t* order,
                 uint32_t* scratch, hipStream_t s, bool compact) {
  const float* qx = q;
  const float* qy = q + nq;
  const float* qz = t.dim == 3 ? q + 2ull * nq : nullptr;
  uint64_t* k0 = reinterpret_cast<uint64_t*>(scratch);
  uint64_t* k1 = k0 + nq;
  uint32_t* v1 = scratch + 4ull * nq;
  uint32_t nb = nblk(nq, kSortTile);
  uint32_t* hist = v1 + nq;
  uint32_t* offs = hist + 256ull * nb;
  uint32_t* scan_scratch = offs + 256ull * nb;
  float4 lo = make_float4(t.lo[0], t.lo[1], t.lo[2], 0.f);
  float4 hi = make_float4(t.hi[0], t.hi[1], t.dim == 3 ? t.hi[2] : 1.f, 0.f);
  if (nq <= kFusedSortMax) {
    const uint32_t fb = nblk(nq, kFrTile);
    uint32_t* fh = v1 + nq;  // hist[pass][digit][tile]
    const uint32_t nh = 8u * 256u * fb;
    k_zero_u32n<<<std::min(nblk(nh, 256), 256u), 256, 0, s>>>(fh, nh);
    SBH_HIP(hipGetLastError());
    k_fr_keys<<<nblk(nq, 256), 256, 0, s>>>(qx, qy, qz, nq, t.dim, lo, hi, k0, order, fh, fb);
    SBH_HIP(hipGetLastError());
    uint64_t* kin = k0;
    uint32_t* vin = order;
    uint64_t* kout = k1;
    uint32_t* vout = v1;
    for (int pss = 0; pss < 8; ++pss) {
      const uint32_t* h = fh + (size_t)pss * 256 * fb;
      uint32_t* hn = pss < 7 ? fh + (size_t)(pss + 1) * 256 * fb : nullptr;
      const size_t lds = 256u * fb * sizeof(uint32_t);
      if (compact)
        k_fr_pass<256><<<fb, 256, lds, s>>>(kin, vin, kout, vout, nq, 8 * pss, fb, h, hn);
      else
        k_fr_pass<1024><<<fb, 1024, lds, s>>>(kin, vin, kout, vout, nq, 8 * pss, fb, h, hn);
      SBH_HIP(hipGetLastError());
      std::swap(kin, kout);
      std::swap(vin, vout);
    }
    return;  // eight passes: the sorted indices are in `order`
  }
  k_curve_keys<<<nblk(nq, 256), 256, 0, s>>>(qx, qy, qz, nq, t.dim, lo, hi, k0, order);
  SBH_HIP(hipGetLastError());
  const int passes = 8;  // 64-bit keys, 8 bits per pass
  uint64_t* kin = k0;
  uint32_t* vin = order;
  uint64_t* kout = k1;
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
