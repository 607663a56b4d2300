// Level-synchronous device tree build.
//
// Follows the reference build (include/tree3d.h:238-371, include/tree2d.h:
// 237-365) exactly: every internal node runs one two-ended std::partition per
// child cell, cells in raster order, each against the node's own fp32 bounds;
// children are numbered breadth first; moments are aggregated bottom up.  The
// leaf order it produces is therefore the reference's leaf order, including
// the order inside multi-point leaves.
//
// One two-ended partition of [s, e) by predicate P moves nothing but this: with
// t = #P, the i-th element of [s, s+t) failing P swaps with the i-th element
// of [s+t, e) satisfying P counted from the right.  Each pass is therefore a
// pair of prefix ranks and a swap, which is what the kernels below compute:
//   * nodes of up to kSmallMax points: one workgroup, everything in LDS;
//   * larger nodes: chunked device-wide passes (count, offsets, rank, swap).
#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "sbh_device.hpp"
#include "sbh_eval.hpp"
#include "sbh_runtime.hpp"

namespace sbh {
namespace {

constexpr int kSmallMax = 4096;
constexpr int kSmallThreads = 256;
constexpr int kTinyMax = 512;  // nodes partitioned by one wavefront each
constexpr int kChunk = 1024;
constexpr int kChunkThreads = 256;
constexpr int kMaxCells = 512;  // 3D with split_log 3

// Point / bound types of the float build and of the double build.
template <typename T>
struct VecOf;
template <>
struct VecOf<float> {
  using type = float4;
};
template <>
struct VecOf<double> {
  using type = dvec4;
};
__host__ __device__ inline float4 make_vec(float x, float y, float z, float w) {
  return make_float4(x, y, z, w);
}
__host__ __device__ inline dvec4 make_vec(double x, double y, double z, double w) {
  return dvec4{x, y, z, w};
}
__device__ __forceinline__ float sqrt_rn(float v) { return __builtin_sqrtf(v); }
__device__ __forceinline__ double sqrt_rn(double v) { return __builtin_sqrt(v); }

// float / double -> int with x86 cvttss2si / cvttsd2si semantics (out of
// range / NaN -> INT_MIN), the conversion behind the reference's cast<int>()
// (include/aabb.h:34-41).
__device__ __forceinline__ int trunc_x86(float v) {
  return (v >= -2147483648.0f && v < 2147483648.0f) ? (int)v : INT_MIN;
}
__device__ __forceinline__ int trunc_x86(double v) {
  return (v >= -2147483648.0 && v < 2147483648.0) ? (int)v : INT_MIN;
}

template <typename T>
__device__ __forceinline__ int quant_axis(T p, T lo, T hi, int dsplit) {
#pragma clang fp contract(off)
  T num = (T)dsplit * (p - lo);
  T den = hi - lo;
  int v = trunc_x86(num / den);
  v = v < dsplit - 1 ? v : dsplit - 1;
  v = v > 0 ? v : 0;
  return v;
}

// Child cell of point p in box [lo, hi): raster index ((z << k) + y) << k + x
// (include/tree3d.h:62-65).
template <int D, typename V>
__device__ __forceinline__ uint32_t cell_of(V p, V lo, V hi, int k) {
  int ds = 1 << k;
  uint32_t x = quant_axis(p.x, lo.x, hi.x, ds);
  uint32_t y = quant_axis(p.y, lo.y, hi.y, ds);
  if constexpr (D == 3) {
    uint32_t z = quant_axis(p.z, lo.z, hi.z, ds);
    return (((z << k) + y) << k) + x;
  } else {
    return (y << k) + x;
  }
}

// Sub-box of cell `cell` (include/tree3d.h:269-280): min + coord * inc, + inc.
template <int D, typename V, typename T>
__device__ __forceinline__ void cell_box(V lo, V hi, int k, uint32_t cell, V& clo, V& chi,
                                         T& diam) {
#pragma clang fp contract(off)
  int ds = 1 << k;
  T fds = (T)ds;
  uint32_t mask = (uint32_t)ds - 1;
  T ix = (T)(cell & mask);
  T iy = (T)((cell >> k) & mask);
  T iz = (T)((cell >> (2 * k)) & mask);
  T incx = (hi.x - lo.x) / fds;
  T incy = (hi.y - lo.y) / fds;
  T incz = (D == 3) ? (hi.z - lo.z) / fds : (T)0;
  clo.x = lo.x + ix * incx;
  clo.y = lo.y + iy * incy;
  clo.z = (D == 3) ? lo.z + iz * incz : (T)0;
  clo.w = (T)0;
  chi.x = clo.x + incx;
  chi.y = clo.y + incy;
  chi.z = (D == 3) ? clo.z + incz : (T)0;
  chi.w = (T)0;
  T dx = chi.x - clo.x, dy = chi.y - clo.y, dz = chi.z - clo.z;
  diam = sqrt_rn(sqnorm_exact<D>(dx, dy, dz));
}

// Exclusive scan of one value per thread over a block of NT threads.
template <int NT>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* wsum,
                                                         uint32_t& total) {
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  uint32_t base = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < NT / 64; ++i) {
    uint32_t s = wsum[i];
    base += (i < w) ? s : 0u;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return base + x - v;
}

template <int NT>
__device__ __forceinline__ uint32_t block_sum(uint32_t v, uint32_t* wsum) {
  uint32_t total;
  (void)block_exclusive_scan<NT>(v, wsum, total);
  return total;
}

// ---------------------------------------------------------------------------
// Level bookkeeping
// ---------------------------------------------------------------------------
// Internal nodes of the level by size: tiny (<= kTinyMax points, one
// wavefront each), small (<= kSmallMax, one workgroup each) and large;
// counters = {small, large, tiny}.
__global__ void k_classify(uint32_t lb, uint32_t count, const uint32_t* n_start,
                           const uint32_t* n_end, const int32_t* n_depth, int leaf_size,
                           int max_depth, uint32_t* n_nchild, uint32_t* n_first,
                           uint32_t* small_list, uint32_t* large_list, uint32_t* tiny_list,
                           uint32_t* counters) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t lane = threadIdx.x & 63;
  bool small = false, large = false, tiny = false;
  uint32_t v = lb + i;
  if (i < count) {
    uint32_t cnt = n_end[v] - n_start[v];
    // include/tree3d.h:34-36
    bool leaf = n_depth[v] >= max_depth || cnt <= (uint32_t)leaf_size;
    n_nchild[v] = 0;
    n_first[v] = kEnd;
    tiny = !leaf && cnt <= (uint32_t)kTinyMax;
    small = !leaf && cnt > (uint32_t)kTinyMax && cnt <= (uint32_t)kSmallMax;
    large = !leaf && cnt > (uint32_t)kSmallMax;
  }
  // one atomic per wavefront and list
  const uint64_t ms = __ballot(small), ml = __ballot(large), mt = __ballot(tiny);
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint32_t bs = 0, bl = 0, bt = 0;
  if (lane == 0) {
    if (ms) bs = atomicAdd(&counters[0], (uint32_t)__popcll(ms));
    if (ml) bl = atomicAdd(&counters[1], (uint32_t)__popcll(ml));
    if (mt) bt = atomicAdd(&counters[2], (uint32_t)__popcll(mt));
  }
  bs = __shfl(bs, 0, 64);
  bl = __shfl(bl, 0, 64);
  bt = __shfl(bt, 0, 64);
  if (small) small_list[bs + __popcll(ms & below)] = v;
  if (large) large_list[bl + __popcll(ml & below)] = v;
  if (tiny) tiny_list[bt + __popcll(mt & below)] = v;
}

// Zone ranking inside one LDS-resident node: element at position p of [a, b)
// is flagged when (dig[p] == c) == want_eq; flagged positions are written to
// out[] in position order (or reversed).  Returns the number flagged.
template <int NT>
__device__ uint32_t rank_zone(const uint16_t* dig, uint32_t a, uint32_t b, uint32_t c,
                              bool want_eq, bool reverse, uint16_t* out, uint32_t* wsum) {
  const uint32_t tid = threadIdx.x;
  uint32_t len = b - a;
  uint32_t per = (len + NT - 1) / NT;
  uint32_t p0 = a + min(len, tid * per);
  uint32_t p1 = a + min(len, (tid + 1) * per);
  uint32_t cnt = 0;
  for (uint32_t p = p0; p < p1; ++p) cnt += ((dig[p] == c) == want_eq) ? 1u : 0u;
  uint32_t total;
  uint32_t r = block_exclusive_scan<NT>(cnt, wsum, total);
  for (uint32_t p = p0; p < p1; ++p) {
    if ((dig[p] == c) == want_eq) {
      out[reverse ? (total - 1 - r) : r] = (uint16_t)p;
      ++r;
    }
  }
  return total;
}

// One workgroup per internal node of at most kSmallMax points.
// counters[0] = the number of small nodes in `list`.
template <int D, typename V>
__global__ __launch_bounds__(kSmallThreads) void k_partition_small(
    const uint32_t* list, const uint32_t* counters, const uint32_t* n_start,
    const uint32_t* n_end, const V* n_lo, const V* n_hi, uint32_t* perm, const V* src_pts, int k,
    uint32_t* n_nchild, uint4* temp_child) {
  __shared__ uint32_t lperm[kSmallMax];
  __shared__ uint16_t ldig[kSmallMax];
  __shared__ uint16_t L[kSmallMax / 2];
  __shared__ uint16_t R[kSmallMax / 2];
  __shared__ uint32_t hist[kMaxCells];
  __shared__ uint32_t wsum[kSmallThreads / 64];

  const uint32_t tid = threadIdx.x;
  if (blockIdx.x >= counters[0]) return;  // uniform per block
  const uint32_t node = list[blockIdx.x];
  const uint32_t s0 = n_start[node];
  const uint32_t n = n_end[node] - s0;
  const V lo = n_lo[node], hi = n_hi[node];
  const uint32_t B = 1u << (D * k);

  for (uint32_t i = tid; i < B; i += kSmallThreads) hist[i] = 0;
  __syncthreads();
  for (uint32_t i = tid; i < n; i += kSmallThreads) {
    uint32_t pi = perm[s0 + i];
    lperm[i] = pi;
    uint32_t d = cell_of<D>(src_pts[pi], lo, hi, k);
    ldig[i] = (uint16_t)d;
    atomicAdd(&hist[d], 1u);
  }
  __syncthreads();

  uint32_t s = 0, nch = 0;
  for (uint32_t c = 0; c < B && s < n; ++c) {
    uint32_t t = hist[c];
    if (t == 0) continue;
    if (tid == 0) temp_child[s0 + nch] = make_uint4(c, s0 + s, s0 + s + t, 0);
    ++nch;
    if (s + t < n) {
      uint32_t m = rank_zone<kSmallThreads>(ldig, s, s + t, c, false, false, L, wsum);
      if (m > 0) {
        (void)rank_zone<kSmallThreads>(ldig, s + t, n, c, true, true, R, wsum);
        __syncthreads();
        for (uint32_t i = tid; i < m; i += kSmallThreads) {
          uint32_t a = L[i], b = R[i];
          uint32_t tp = lperm[a];
          lperm[a] = lperm[b];
          lperm[b] = tp;
          uint16_t td = ldig[a];
          ldig[a] = ldig[b];
          ldig[b] = td;
        }
        __syncthreads();
      }
    }
    s += t;
  }
  for (uint32_t i = tid; i < n; i += kSmallThreads) perm[s0 + i] = lperm[i];
  if (tid == 0) n_nchild[node] = nch;
}

// One wavefront per internal node of at most kTinyMax points (counters[2]
// of them in `list`), four wavefronts a block.  The same passes as
// k_partition_small: for each non-empty cell c in order, the remainder
// [s, n) is split at s + t (t = the cell's count) and the i-th point of
// [s, s + t) outside c swaps with the i-th point of [s + t, n) in c counted
// from the right.  Ranks come from wavefront ballots over 64-position slabs,
// so a pass needs no block barrier.
template <int D, typename V>
__global__ __launch_bounds__(256) void k_partition_tiny(
    const uint32_t* list, const uint32_t* counters, const uint32_t* n_start,
    const uint32_t* n_end, const V* n_lo, const V* n_hi, uint32_t* perm, const V* src_pts, int k,
    uint32_t* n_nchild, uint4* temp_child) {
  constexpr int kWaves = 4;
  __shared__ uint32_t s_perm[kWaves][kTinyMax];
  __shared__ uint16_t s_dig[kWaves][kTinyMax];
  __shared__ uint16_t s_L[kWaves][kTinyMax / 2];
  __shared__ uint16_t s_R[kWaves][kTinyMax / 2];
  __shared__ uint32_t s_hist[kWaves][kMaxCells];
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t idx = blockIdx.x * kWaves + wave;
  if (idx >= counters[2]) return;  // uniform per wavefront; no block barrier follows
  uint32_t* lperm = s_perm[wave];
  uint16_t* ldig = s_dig[wave];
  uint16_t* L = s_L[wave];
  uint16_t* R = s_R[wave];
  uint32_t* hist = s_hist[wave];
  const uint32_t node = list[idx];
  const uint32_t s0 = n_start[node];
  const uint32_t n = n_end[node] - s0;
  const V lo = n_lo[node], hi = n_hi[node];
  const uint32_t B = 1u << (D * k);

  for (uint32_t i = lane; i < B; i += 64) hist[i] = 0;
  wave_lds_sync();
  for (uint32_t i = lane; i < n; i += 64) {
    const uint32_t pi = perm[s0 + i];
    lperm[i] = pi;
    const uint32_t d = cell_of<D>(src_pts[pi], lo, hi, k);
    ldig[i] = (uint16_t)d;
    atomicAdd(&hist[d], 1u);
  }
  wave_lds_sync();

  uint32_t s = 0, nch = 0;
  for (uint32_t cb = 0; cb < B && s < n; cb += 64) {
    const uint32_t h = cb + lane < B ? hist[cb + lane] : 0u;
    uint64_t present = __ballot(h != 0);
    while (present != 0 && s < n) {
      const uint32_t j = (uint32_t)__builtin_ctzll(present);
      present &= present - 1;
      const uint32_t c = cb + j;
      const uint32_t t = (uint32_t)__builtin_amdgcn_readlane((int)h, (int)j);
      if (lane == 0) temp_child[s0 + nch] = make_uint4(c, s0 + s, s0 + s + t, 0);
      ++nch;
      if (s + t < n) {
        // positions of the misplaced points, left zone ascending, right zone ascending
        uint32_t nl = 0, nr = 0;
        for (uint32_t p0 = s; p0 < n; p0 += 64) {
          const uint32_t p = p0 + lane;
          const uint32_t d = p < n ? (uint32_t)ldig[p] : 0xFFFFFFFFu;
          const bool fl = p < s + t && d != c;
          const bool fr = p >= s + t && d == c;
          const uint64_t ml = __ballot(fl), mr = __ballot(fr);
          if (fl) L[nl + (uint32_t)__builtin_amdgcn_mbcnt_hi((uint32_t)(ml >> 32),
                                    __builtin_amdgcn_mbcnt_lo((uint32_t)ml, 0u))] = (uint16_t)p;
          if (fr) R[nr + (uint32_t)__builtin_amdgcn_mbcnt_hi((uint32_t)(mr >> 32),
                                    __builtin_amdgcn_mbcnt_lo((uint32_t)mr, 0u))] = (uint16_t)p;
          nl += (uint32_t)__popcll(ml);
          nr += (uint32_t)__popcll(mr);
        }
        wave_lds_sync();
        // nl == nr: the i-th misplaced point on the left pairs with the i-th
        // misplaced point on the right counted from the right
        for (uint32_t i = lane; i < nl; i += 64) {
          const uint32_t a = L[i], b = R[nl - 1 - i];
          const uint32_t tp = lperm[a];
          lperm[a] = lperm[b];
          lperm[b] = tp;
          const uint16_t td = ldig[a];
          ldig[a] = ldig[b];
          ldig[b] = td;
        }
        wave_lds_sync();
      }
      s += t;
    }
  }
  for (uint32_t i = lane; i < n; i += 64) perm[s0 + i] = lperm[i];
  if (lane == 0) n_nchild[node] = nch;
}

// ---------------------------------------------------------------------------
// Large nodes: chunk table entries are {large index, begin, end, node}.
// ---------------------------------------------------------------------------
struct LargeCtx {
  const uint4* chunks;
  const uint32_t* l_start;  // per large node
  const uint32_t* l_end;
  uint32_t* hist;        // [nlarge * B]
  uint32_t* cellstart;   // [nlarge * B]
  uint32_t* mcount;      // [nlarge]
  uint2* ccnt;           // [nchunks]
  const uint32_t* l_chunk0;  // first chunk of each large node
  const uint32_t* l_nchunk;
  uint16_t* digit;       // [N] position-indexed
  uint32_t* Lg;          // [N]
  uint32_t* Rg;          // [N]
};

template <int D, typename V>
__global__ __launch_bounds__(kChunkThreads) void k_large_digits(LargeCtx cx, const V* n_lo,
                                                                 const V* n_hi,
                                                                 const uint32_t* perm,
                                                                 const V* src_pts, int k) {
  __shared__ uint32_t lh[kMaxCells];
  const uint4 ch = cx.chunks[blockIdx.x];
  const uint32_t B = 1u << (D * k);
  for (uint32_t i = threadIdx.x; i < B; i += kChunkThreads) lh[i] = 0;
  __syncthreads();
  const V lo = n_lo[ch.w], hi = n_hi[ch.w];
  for (uint32_t p = ch.y + threadIdx.x; p < ch.z; p += kChunkThreads) {
    uint32_t d = cell_of<D>(src_pts[perm[p]], lo, hi, k);
    cx.digit[p] = (uint16_t)d;
    atomicAdd(&lh[d], 1u);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < B; i += kChunkThreads)
    if (lh[i]) atomicAdd(&cx.hist[ch.x * B + i], lh[i]);
}

// The level's large nodes with their point ranges: out = [nodes | starts | ends].
__global__ void k_gather_large(const uint32_t* large_list, uint32_t nlarge,
                               const uint32_t* n_start, const uint32_t* n_end, uint32_t* out) {
  const uint32_t li = blockIdx.x * blockDim.x + threadIdx.x;
  if (li >= nlarge) return;
  const uint32_t v = large_list[li];
  out[li] = v;
  out[nlarge + li] = n_start[v];
  out[2 * nlarge + li] = n_end[v];
}

// Per large node: child ranges from the histogram (children of a partitioned
// node are the runs of equal cell, in cell order).
__global__ void k_large_cells(LargeCtx cx, const uint32_t* large_list, uint32_t nlarge, uint32_t B,
                              uint32_t* n_nchild, uint4* temp_child) {
  uint32_t li = blockIdx.x * blockDim.x + threadIdx.x;
  if (li >= nlarge) return;
  uint32_t run = cx.l_start[li], nch = 0;
  for (uint32_t c = 0; c < B; ++c) {
    uint32_t h = cx.hist[li * B + c];
    cx.cellstart[li * B + c] = run;
    if (h) {
      temp_child[cx.l_start[li] + nch] = make_uint4(c, run, run + h, 0);
      ++nch;
    }
    run += h;
  }
  n_nchild[large_list[li]] = nch;
}

__device__ __forceinline__ void large_flags(const LargeCtx& cx, uint32_t li, uint32_t B, uint32_t c,
                                            uint32_t p, bool& lf, bool& rt) {
  uint32_t s = cx.cellstart[li * B + c];
  uint32_t t = cx.hist[li * B + c];
  uint32_t e = cx.l_end[li];
  uint32_t d = cx.digit[p];
  lf = (p >= s) && (p < s + t) && (d != c);
  rt = (p >= s + t) && (p < e) && (d == c);
}

__global__ __launch_bounds__(kChunkThreads) void k_large_count(LargeCtx cx, uint32_t B, uint32_t c) {
  __shared__ uint32_t wsum[kChunkThreads / 64];
  const uint4 ch = cx.chunks[blockIdx.x];
  uint32_t nl = 0, nr = 0;
  if (cx.hist[ch.x * B + c] != 0) {
    for (uint32_t p = ch.y + threadIdx.x; p < ch.z; p += kChunkThreads) {
      bool lf, rt;
      large_flags(cx, ch.x, B, c, p, lf, rt);
      nl += lf;
      nr += rt;
    }
  }
  uint32_t tot = block_sum<kChunkThreads>(nl | (nr << 16), wsum);
  if (threadIdx.x == 0) cx.ccnt[blockIdx.x] = make_uint2(tot & 0xFFFFu, tot >> 16);
}

// Chunk of kChunk positions, 4 consecutive positions per thread.  The
// chunk's offsets (left-false and right-true counts of the node's earlier
// chunks) and the node's left-false total m come from the chunk counts: each
// block adds them up itself (a node has at most a few thousand chunks).
__global__ __launch_bounds__(kChunkThreads) void k_large_rank(LargeCtx cx, uint32_t B, uint32_t c) {
  __shared__ uint32_t wsum[kChunkThreads / 64];
  const uint4 ch = cx.chunks[blockIdx.x];
  if (cx.hist[ch.x * B + c] == 0) return;  // uniform per block
  const uint32_t s = cx.cellstart[ch.x * B + c];
  const uint32_t c0 = cx.l_chunk0[ch.x], nc = cx.l_nchunk[ch.x], me = blockIdx.x - c0;
  uint32_t bl = 0, br = 0, ml = 0;
  for (uint32_t j = threadIdx.x; j < nc; j += kChunkThreads) {
    const uint2 cc = cx.ccnt[c0 + j];
    bl += j < me ? cc.x : 0u;
    br += j < me ? cc.y : 0u;
    ml += cc.x;
  }
  const uint2 off = make_uint2(block_sum<kChunkThreads>(bl, wsum),
                               block_sum<kChunkThreads>(br, wsum));
  const uint32_t m = block_sum<kChunkThreads>(ml, wsum);
  if (me == 0 && threadIdx.x == 0) cx.mcount[ch.x] = m;  // for the swap pass
  uint32_t p0 = ch.y + threadIdx.x * 4;
  bool lf[4], rt[4];
  uint32_t nl = 0, nr = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    uint32_t p = p0 + j;
    lf[j] = rt[j] = false;
    if (p < ch.z) large_flags(cx, ch.x, B, c, p, lf[j], rt[j]);
    nl += lf[j];
    nr += rt[j];
  }
  uint32_t total;
  uint32_t pre = block_exclusive_scan<kChunkThreads>(nl | (nr << 16), wsum, total);
  uint32_t rl = off.x + (pre & 0xFFFFu);
  uint32_t rr = off.y + (pre >> 16);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    uint32_t p = p0 + j;
    if (lf[j]) cx.Lg[s + rl++] = p;
    if (rt[j]) cx.Rg[s + (m - 1 - rr++)] = p;
  }
}

__global__ __launch_bounds__(kChunkThreads) void k_large_swap(LargeCtx cx, uint32_t B, uint32_t c,
                                                              uint32_t* perm) {
  const uint4 ch = cx.chunks[blockIdx.x];
  if (cx.hist[ch.x * B + c] == 0) return;
  const uint32_t s = cx.cellstart[ch.x * B + c];
  const uint32_t m = cx.mcount[ch.x];
  for (uint32_t p = ch.y + threadIdx.x; p < ch.z; p += kChunkThreads) {
    if (p >= s && p < s + m) {
      uint32_t a = cx.Lg[p], b = cx.Rg[p];
      uint32_t tp = perm[a];
      perm[a] = perm[b];
      perm[b] = tp;
      uint16_t td = cx.digit[a];
      cx.digit[a] = cx.digit[b];
      cx.digit[b] = td;
    }
  }
}

// Children of the level's internal nodes get consecutive breadth-first
// numbers: node order, then cell order (include/tree3d.h:400-411).
template <int D, typename V, typename T>
__global__ void k_emit(uint32_t lb, uint32_t count, uint32_t next_base, const uint32_t* cbase, int k,
                       uint32_t* n_start, uint32_t* n_end, uint32_t* n_parent, uint32_t* n_first,
                       const uint32_t* n_nchild, uint32_t* n_next, int32_t* n_depth, V* n_lo,
                       V* n_hi, T* n_diam, const uint4* temp_child) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  uint32_t v = lb + i;
  uint32_t nch = n_nchild[v];
  if (nch == 0) return;
  uint32_t base = next_base + cbase[i];
  n_first[v] = base;
  V lo = n_lo[v], hi = n_hi[v];
  int32_t dep = n_depth[v] + 1;
  uint32_t vnext = n_next[v];
  uint32_t s0 = n_start[v];
  for (uint32_t j = 0; j < nch; ++j) {
    uint4 tc = temp_child[s0 + j];
    uint32_t cidx = base + j;
    n_start[cidx] = tc.y;
    n_end[cidx] = tc.z;
    n_parent[cidx] = v;
    n_depth[cidx] = dep;
    V clo, chi;
    T diam;
    cell_box<D>(lo, hi, k, tc.x, clo, chi, diam);
    n_lo[cidx] = clo;
    n_hi[cidx] = chi;
    n_diam[cidx] = diam;
    n_next[cidx] = (j + 1 < nch) ? cidx + 1 : vnext;
  }
}

template <typename V>
__global__ void k_gather_points(uint32_t n, const uint32_t* perm, const V* src_pts,
                                const V* src_nrm, V* pts, V* nrm) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t j = perm[i];
  pts[i] = src_pts[j];
  nrm[i] = src_nrm[j];
}

// Node moments (include/tree3d.h:255-264 for leaves, :310-327 for internal
// nodes): fp32, array / bitmask order, one rounding per operation.  The
// float64 aggregation of the same quantities is kept beside them.
template <int D>
__global__ void k_moments(uint32_t lb, uint32_t count, const uint32_t* n_start,
                          const uint32_t* n_end, const uint32_t* n_first,
                          const uint32_t* n_nchild, const float4* pts, const float4* nrm,
                          float4* nodeC, float4* nodeN, double* m64, uint32_t* n_skip,
                          uint32_t* n_size, int use64) {
#pragma clang fp contract(off)
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  uint32_t v = lb + i;
  uint32_t nch = n_nchild[v];
  float mass = 0.f, cx = 0.f, cy = 0.f, cz = 0.f, nx = 0.f, ny = 0.f, nz = 0.f;
  double dm = 0.0, dcx = 0.0, dcy = 0.0, dcz = 0.0, dnx = 0.0, dny = 0.0, dnz = 0.0;
  if (nch == 0) {
    for (uint32_t j = n_start[v]; j < n_end[v]; ++j) {
      float4 p = pts[j];
      float4 q = nrm[j];
      mass += p.w;
      cx += p.w * p.x;
      cy += p.w * p.y;
      cz += p.w * p.z;
      nx += p.w * q.x;
      ny += p.w * q.y;
      nz += p.w * q.z;
      double w = p.w;
      dm += w;
      dcx += w * p.x;
      dcy += w * p.y;
      dcz += w * p.z;
      dnx += w * q.x;
      dny += w * q.y;
      dnz += w * q.z;
    }
    n_skip[v] = v;
    n_size[v] = 1;
  } else {
    uint32_t f = n_first[v];
    for (uint32_t j = 0; j < nch; ++j) {
      float4 c = nodeC[f + j];
      float4 q = nodeN[f + j];
      mass += c.w;
      cx += c.w * c.x;
      cy += c.w * c.y;
      cz += c.w * c.z;
      nx += c.w * q.x;
      ny += c.w * q.y;
      nz += c.w * q.z;
      const double* cm = m64 + 8ull * (f + j);
      dm += cm[3];
      dcx += cm[3] * cm[0];
      dcy += cm[3] * cm[1];
      dcz += cm[3] * cm[2];
      dnx += cm[3] * cm[4];
      dny += cm[3] * cm[5];
      dnz += cm[3] * cm[6];
    }
    n_skip[v] = (nch == 1) ? n_skip[f] : v;
    uint32_t sz = 1;
    for (uint32_t j = 0; j < nch; ++j) sz += n_size[f + j];
    n_size[v] = sz;
  }
  cx /= mass;
  cy /= mass;
  cz /= mass;
  nx /= mass;
  ny /= mass;
  nz /= mass;
  double* o = m64 + 8ull * v;
  o[0] = dcx / dm;
  o[1] = dcy / dm;
  o[2] = dcz / dm;
  o[3] = dm;
  o[4] = dnx / dm;
  o[5] = dny / dm;
  o[6] = dnz / dm;
  o[7] = 0.0;
  if (use64) {
    nodeC[v] = make_float4((float)o[0], (float)o[1], (float)o[2], (float)dm);
    nodeN[v] = make_float4((float)o[4], (float)o[5], (float)o[6], 0.f);
  } else {
    nodeC[v] = make_float4(cx, cy, cz, mass);
    nodeN[v] = make_float4(nx, ny, nz, 0.f);
  }
}

// Node moments of the double build: the fp32 recurrence above in float64
// (include/tree3d.h:255-264, :310-327 with FLOAT = double).  The same values
// go to n_m64, the float64 moments every tree exports.
template <int D>
__global__ void k_moments_double(uint32_t lb, uint32_t count, const uint32_t* n_start,
                                 const uint32_t* n_end, const uint32_t* n_first,
                                 const uint32_t* n_nchild, const dvec4* pts, const dvec4* nrm,
                                 dvec4* nodeC, dvec4* nodeN, double* m64, uint32_t* n_skip,
                                 uint32_t* n_size) {
#pragma clang fp contract(off)
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  uint32_t v = lb + i;
  uint32_t nch = n_nchild[v];
  double mass = 0.0, cx = 0.0, cy = 0.0, cz = 0.0, nx = 0.0, ny = 0.0, nz = 0.0;
  if (nch == 0) {
    for (uint32_t j = n_start[v]; j < n_end[v]; ++j) {
      const dvec4 p = pts[j];
      const dvec4 q = nrm[j];
      mass += p.w;
      cx += p.w * p.x;
      cy += p.w * p.y;
      cz += p.w * p.z;
      nx += p.w * q.x;
      ny += p.w * q.y;
      nz += p.w * q.z;
    }
    n_skip[v] = v;
    n_size[v] = 1;
  } else {
    uint32_t f = n_first[v];
    for (uint32_t j = 0; j < nch; ++j) {
      const dvec4 c = nodeC[f + j];
      const dvec4 q = nodeN[f + j];
      mass += c.w;
      cx += c.w * c.x;
      cy += c.w * c.y;
      cz += c.w * c.z;
      nx += c.w * q.x;
      ny += c.w * q.y;
      nz += c.w * q.z;
    }
    n_skip[v] = (nch == 1) ? n_skip[f] : v;
    uint32_t sz = 1;
    for (uint32_t j = 0; j < nch; ++j) sz += n_size[f + j];
    n_size[v] = sz;
  }
  cx /= mass;
  cy /= mass;
  cz /= mass;
  nx /= mass;
  ny /= mass;
  nz /= mass;
  nodeC[v] = dvec4{cx, cy, cz, mass};
  nodeN[v] = dvec4{nx, ny, nz, 0.0};
  double* o = m64 + 8ull * v;
  o[0] = cx;
  o[1] = cy;
  o[2] = cz;
  o[3] = mass;
  o[4] = nx;
  o[5] = ny;
  o[6] = nz;
  o[7] = 0.0;
}

// Depth-first records of the double build, as k_dfs_scatter writes them for
// the float build (a one-point leaf's record holds its point).
__global__ void k_dfs_scatter_double(uint32_t nn, const uint32_t* n_dfs, const uint32_t* n_size,
                                     const dvec4* nodeC, const dvec4* nodeN, const uint4* nodeL,
                                     const double* n_diam, const dvec4* pts, const dvec4* nrm,
                                     WalkRecordD* dfsR, dvec4* dfsN) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nn) return;
  const uint32_t d = n_dfs[v];
  const uint4 l = nodeL[v];
  const bool one_point = l.w == (kLeafBit | 1u);
  const uint32_t rope = d + n_size[v];
  WalkRecordD r;
  r.c = one_point ? pts[l.y] : nodeC[v];
  r.diam = n_diam[v];
  r.first = (l.w & kLeafBit) ? l.y : d + 1;
  r.rope = rope < nn ? rope : kEnd;
  r.info = l.w;
  r.pad = 0;
  dfsR[d] = r;
  dfsN[d] = one_point ? nrm[l.y] : nodeN[v];
}

// The walk record {diameter bits, first child | first point, rope, info}; a
// double tree keeps its diameters in n_diam64 and records their fp32 rounding.
template <typename T>
__global__ void k_finalize(uint32_t nn, const uint32_t* n_start, const uint32_t* n_end,
                           const uint32_t* n_first, const uint32_t* n_nchild,
                           const uint32_t* n_next, const T* n_diam, uint4* nodeL) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nn) return;
  uint32_t nch = n_nchild[v];
  uint4 r;
  r.x = __float_as_uint((float)n_diam[v]);
  r.z = n_next[v];
  if (nch == 0) {
    r.y = n_start[v];
    r.w = kLeafBit | (n_end[v] - n_start[v]);
  } else {
    r.y = n_first[v];
    r.w = nch;
  }
  nodeL[v] = r;
}

// Depth-first positions, one level at a time from the root: a node's first
// child follows it, each further child follows its elder sibling's subtree.
__global__ void k_dfs_positions(uint32_t lb, uint32_t count, const uint32_t* n_first,
                                const uint32_t* n_nchild, const uint32_t* n_size,
                                uint32_t* n_dfs) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  uint32_t v = lb + i;
  uint32_t nch = n_nchild[v];
  if (nch == 0) return;
  uint32_t p = n_dfs[v] + 1;
  uint32_t f = n_first[v];
  for (uint32_t j = 0; j < nch; ++j) {
    n_dfs[f + j] = p;
    p += n_size[f + j];
  }
}

// A sampled path sums a one-point leaf child through its point; the child
// records give every child's term source in one load.
__global__ void k_child_sources(uint32_t nn, const float4* nodeC, const float4* nodeN,
                                const uint4* nodeL, const float4* pts, const float4* nrm,
                                const float4* dfsC, const uint4* dfsL, const uint32_t* n_start,
                                const uint32_t* n_end, const uint32_t* n_skip, float4* kidC,
                                float4* kidN, WalkRecord* dfsR, WalkRecord* bfsR,
                                uint4* stepR) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nn) return;
  const uint4 l = nodeL[v];
  const bool one_point = l.w == (kLeafBit | 1u);
  kidC[v] = one_point ? pts[l.y] : nodeC[v];
  kidN[v] = one_point ? nrm[l.y] : nodeN[v];
  dfsR[v] = WalkRecord{dfsC[v], dfsL[v]};
  // the fast sampler's opening ratios use the hardware reciprocal of the
  // diameter (never the diameter itself): its records hold the reciprocal
  bfsR[v] = WalkRecord{nodeC[v],
                       make_uint4(__float_as_uint(__builtin_amdgcn_rcpf(__uint_as_float(l.x))), l.y,
                                  l.z, l.w)};
  stepR[v] = make_uint4(n_start[v], n_end[v], n_skip[v], 0u);
}

// The walk never reads a leaf's moment (leaves are summed point by point), so
// a one-point leaf's depth-first record holds the point itself and the walk
// sums it without a further load.
__global__ void k_dfs_scatter(uint32_t nn, const uint32_t* n_dfs, const uint32_t* n_size,
                              const float4* nodeC, const float4* nodeN, const uint4* nodeL,
                              const float4* pts, const float4* nrm, float4* dfsC, float4* dfsN,
                              uint4* dfsL) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nn) return;
  uint32_t d = n_dfs[v];
  uint4 l = nodeL[v];
  const bool one_point = l.w == (kLeafBit | 1u);
  if (!(l.w & kLeafBit)) l.y = d + 1;
  uint32_t rope = d + n_size[v];
  l.z = rope < nn ? rope : kEnd;
  dfsC[d] = one_point ? pts[l.y] : nodeC[v];
  dfsN[d] = one_point ? nrm[l.y] : nodeN[v];
  dfsL[d] = l;
}

template <int D, typename T, typename V>
__global__ void k_pack_inputs(uint32_t n, const T* pts, const T* m, const T* nrm, V* out_pts,
                              V* out_nrm, uint32_t* perm) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T* p = pts + (size_t)i * D;
  out_pts[i] = make_vec(p[0], p[1], D == 3 ? p[2] : (T)0, m[i]);
  if (nrm) {
    const T* q = nrm + (size_t)i * D;
    out_nrm[i] = make_vec(q[0], q[1], D == 3 ? q[2] : (T)0, (T)0);
  } else {
    out_nrm[i] = make_vec((T)0, (T)0, (T)0, (T)0);
  }
  perm[i] = i;
}

__global__ void k_zero_u32(uint32_t* p) { *p = 0; }
__global__ void k_zero_u32x4(uint32_t* p) { p[threadIdx.x] = 0; }
__global__ void k_zero_u32s(uint32_t* p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

template <int D, typename V, typename T>
__global__ void k_set_root(uint32_t n, V lo, V hi, uint32_t* n_start, uint32_t* n_end,
                           uint32_t* n_parent, int32_t* n_depth, uint32_t* n_next, V* n_lo,
                           V* n_hi, T* n_diam) {
#pragma clang fp contract(off)
  T dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
  T diam = sqrt_rn(sqnorm_exact<D>(dx, dy, dz));  // include/tree3d.h:249
  n_start[0] = 0;
  n_end[0] = n;
  n_parent[0] = 0;
  n_depth[0] = 0;
  n_next[0] = kEnd;
  n_lo[0] = lo;
  n_hi[0] = hi;
  n_diam[0] = diam;
}

// ---------------------------------------------------------------------------
// Device exclusive scan (1024 values per block, recursive over block sums).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_scan_blocks(const uint32_t* in, uint32_t* out, uint32_t n,
                                                     uint32_t* sums) {
  __shared__ uint32_t wsum[4];
  uint32_t base = blockIdx.x * 1024 + threadIdx.x * 4;
  uint32_t v[4];
  uint32_t loc = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    v[j] = (base + j < n) ? in[base + j] : 0u;
    loc += v[j];
  }
  uint32_t total;
  uint32_t pre = block_exclusive_scan<256>(loc, wsum, total);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (base + j < n) out[base + j] = pre;
    pre += v[j];
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ void k_scan_add(uint32_t* out, uint32_t n, const uint32_t* sums) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] += sums[i / 1024];
}

__global__ void k_copy_u32(const uint32_t* src, uint32_t* dst) { *dst = *src; }

void scan_rec(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* total_dev,
              uint32_t* scratch, hipStream_t s) {
  uint32_t nb = (n + 1023) / 1024;
  uint32_t* sums = scratch;
  k_scan_blocks<<<nb, 256, 0, s>>>(in, out, n, sums);
  SBH_HIP(hipGetLastError());
  if (nb == 1) {
    if (total_dev) k_copy_u32<<<1, 1, 0, s>>>(sums, total_dev);
    return;
  }
  scan_rec(sums, sums, nb, total_dev, scratch + nb, s);
  k_scan_add<<<(n + 255) / 256, 256, 0, s>>>(out, n, sums);
  SBH_HIP(hipGetLastError());
}

size_t scan_scratch_size(uint32_t n) {
  size_t tot = 0;
  do {
    n = (n + 1023) / 1024;
    tot += n;
  } while (n > 1);
  return tot + 4;
}

template <typename T>
void h2d(T* dst, const T* src, size_t n, hipStream_t s) {
  SBH_HIP(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, s));
}
template <typename T>
void d2h(T* dst, const T* src, size_t n, hipStream_t s) {
  SBH_HIP(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, s));
  SBH_HIP(hipStreamSynchronize(s));
}

inline uint32_t blocks_for(uint64_t n, uint32_t t) { return (uint32_t)((n + t - 1) / t); }

// Byte offsets of the pieces of one scratch allocation, each 256-byte aligned.
struct ArenaPlan {
  size_t end = 0;
  template <typename T>
  size_t add(size_t n) {
    const size_t o = end;
    end += (n * sizeof(T) + 255) & ~(size_t)255;
    return o;
  }
  size_t bytes() const { return end ? end : 256; }
};

// The arrays a build of scalar type T writes: the fp32 set, or the float64
// set of the double build.
template <typename T>
struct BuildArrays;
template <>
struct BuildArrays<float> {
  static DevBuf<float4>& lo(DeviceTree& t) { return t.n_lo; }
  static DevBuf<float4>& hi(DeviceTree& t) { return t.n_hi; }
  static DevBuf<float>& diam(DeviceTree& t) { return t.n_diam; }
  static DevBuf<float4>& pts(DeviceTree& t) { return t.pts; }
  static DevBuf<float4>& nrm(DeviceTree& t) { return t.nrm; }
  static const float* root_lo(const DeviceTree& t) { return t.lo; }
  static const float* root_hi(const DeviceTree& t) { return t.hi; }
};
template <>
struct BuildArrays<double> {
  static DevBuf<dvec4>& lo(DeviceTree& t) { return t.n_lo64; }
  static DevBuf<dvec4>& hi(DeviceTree& t) { return t.n_hi64; }
  static DevBuf<double>& diam(DeviceTree& t) { return t.n_diam64; }
  static DevBuf<dvec4>& pts(DeviceTree& t) { return t.pts64; }
  static DevBuf<dvec4>& nrm(DeviceTree& t) { return t.nrm64; }
  static const double* root_lo(const DeviceTree& t) { return t.lo64; }
  static const double* root_hi(const DeviceTree& t) { return t.hi64; }
};

template <int D, typename T>
void build_impl(DeviceTree& t, const T* points, const T* normals, const T* masses) {
  using V = typename VecOf<T>::type;
  using A = BuildArrays<T>;
  hipStream_t s = t.stream;
  const uint32_t N = t.n;
  const int k = t.split_log;
  const uint32_t B = 1u << (D * k);
  DevBuf<V>& n_lo = A::lo(t);
  DevBuf<V>& n_hi = A::hi(t);
  DevBuf<T>& n_diam = A::diam(t);

  // Inputs as {x, y, z, m} / {nx, ny, nz, 0} rows in input order, packed on
  // the device from the caller's row-major arrays; perm starts as the identity.
  DevBuf<V> src_pts(N), src_nrm(N);
  t.perm.alloc(N);
  {
    DevBuf<T> raw((size_t)N * (2 * D + 1));
    T* r_pts = raw.get();
    T* r_m = r_pts + (size_t)N * D;
    T* r_nrm = r_m + N;
    h2d(r_pts, points, (size_t)N * D, s);
    h2d(r_m, masses, N, s);
    if (normals) h2d(r_nrm, normals, (size_t)N * D, s);
    k_pack_inputs<D><<<blocks_for(N, 256), 256, 0, s>>>(N, r_pts, r_m, normals ? r_nrm : nullptr,
                                                        src_pts.get(), src_nrm.get(),
                                                        t.perm.get());
    SBH_HIP(hipGetLastError());
    SBH_HIP(hipStreamSynchronize(s));  // raw is freed next
  }

  // Node arrays with growable capacity.
  size_t cap = std::max<size_t>(16, (size_t)N * 2 + 2);
  t.n_start.alloc(cap);
  t.n_end.alloc(cap);
  t.n_parent.alloc(cap);
  t.n_first.alloc(cap);
  t.n_nchild.alloc(cap);
  t.n_next.alloc(cap);
  t.n_depth.alloc(cap);
  n_lo.alloc(cap);
  n_hi.alloc(cap);
  n_diam.alloc(cap);

  auto grow_all = [&](size_t need, size_t keep) {
    if (need <= cap) return;
    size_t nc = std::max(need, cap * 2);
    t.n_start.grow(nc, keep, s);
    t.n_end.grow(nc, keep, s);
    t.n_parent.grow(nc, keep, s);
    t.n_first.grow(nc, keep, s);
    t.n_nchild.grow(nc, keep, s);
    t.n_next.grow(nc, keep, s);
    t.n_depth.grow(nc, keep, s);
    n_lo.grow(nc, keep, s);
    n_hi.grow(nc, keep, s);
    n_diam.grow(nc, keep, s);
    cap = nc;
  };

  // Scratch: one allocation for the whole build, carved into the level
  // lists, the partition work arrays and the large nodes' tables at their
  // largest sizes (a level's large nodes hold more than kSmallMax points each
  // and disjoint ranges, so there are at most N / kSmallMax of them, and
  // their chunks number at most N / kChunk plus one per node).
  const size_t max_large = (size_t)N / kSmallMax + 2;
  const size_t max_chunks = (size_t)N / kChunk + max_large;
  ArenaPlan plan;
  const size_t o_small = plan.add<uint32_t>(N + 1), o_tiny = plan.add<uint32_t>(N + 1),
               o_large = plan.add<uint32_t>(max_large),
               o_counters = plan.add<uint32_t>(4), o_temp = plan.add<uint4>(N + 1),
               o_cbase = plan.add<uint32_t>(N + 1), o_total = plan.add<uint32_t>(1),
               o_scan = plan.add<uint32_t>(scan_scratch_size(N + 1)),
               o_digit = plan.add<uint16_t>(N + 1), o_lg = plan.add<uint32_t>(N + 1),
               o_rg = plan.add<uint32_t>(N + 1), o_lrange = plan.add<uint32_t>(3 * max_large),
               o_tables = plan.add<uint32_t>(4 * max_chunks + 4 * max_large),
               o_hist = plan.add<uint32_t>(max_large * B),
               o_cellstart = plan.add<uint32_t>(max_large * B),
               o_m = plan.add<uint32_t>(max_large), o_ccnt = plan.add<uint2>(max_chunks);
  DevBuf<uint8_t> arena(plan.bytes());
  uint8_t* ab = arena.get();
  uint32_t* small_list = (uint32_t*)(ab + o_small);
  uint32_t* tiny_list = (uint32_t*)(ab + o_tiny);
  uint32_t* large_list = (uint32_t*)(ab + o_large);
  uint32_t* counters = (uint32_t*)(ab + o_counters);
  uint4* temp_child = (uint4*)(ab + o_temp);
  uint32_t* cbase = (uint32_t*)(ab + o_cbase);
  uint32_t* total_dev = (uint32_t*)(ab + o_total);
  uint32_t* scan_scratch = (uint32_t*)(ab + o_scan);
  uint16_t* digit = (uint16_t*)(ab + o_digit);
  uint32_t* Lg = (uint32_t*)(ab + o_lg);
  uint32_t* Rg = (uint32_t*)(ab + o_rg);
  uint32_t* d_lrange = (uint32_t*)(ab + o_lrange);
  uint32_t* d_tables = (uint32_t*)(ab + o_tables);

  const T* tlo = A::root_lo(t);
  const T* thi = A::root_hi(t);
  const V rlo = make_vec(tlo[0], tlo[1], D == 3 ? tlo[2] : (T)0, (T)0);
  const V rhi = make_vec(thi[0], thi[1], D == 3 ? thi[2] : (T)0, (T)0);
  k_set_root<D><<<1, 1, 0, s>>>(N, rlo, rhi, t.n_start.get(), t.n_end.get(), t.n_parent.get(),
                                t.n_depth.get(), t.n_next.get(), n_lo.get(), n_hi.get(),
                                n_diam.get());
  SBH_HIP(hipGetLastError());

  t.level_begin.clear();
  uint32_t lb = 0, lcount = 1, num_nodes = 1;
  std::vector<uint32_t> h_large, h_lstart, h_lend, h_tables, h_hist;
  // Once a level has no node of more than kSmallMax points, no later level
  // has one (children hold subsets of their parent's points): from then on
  // the small-node partition runs over an upper bound of blocks that read the
  // level's count on the device, and the classification is not read back.
  bool large_levels = true;
  while (lcount > 0) {
    t.level_begin.push_back(lb);
    k_zero_u32x4<<<1, 4, 0, s>>>(counters);
    k_classify<<<blocks_for(lcount, 256), 256, 0, s>>>(
        lb, lcount, t.n_start.get(), t.n_end.get(), t.n_depth.get(), t.leaf_size, t.max_depth,
        t.n_nchild.get(), t.n_first.get(), small_list, large_list, tiny_list, counters);
    SBH_HIP(hipGetLastError());
    uint32_t nsmall = lcount, nlarge = 0, ntiny = lcount;  // upper bounds when not read back
    if (large_levels) {
      uint32_t hc[4];
      d2h(hc, counters, 4, s);
      nsmall = hc[0];
      nlarge = hc[1];
      ntiny = hc[2];
      large_levels = nlarge > 0;
    }

    if (nlarge > 0) {
      // the large nodes and their point ranges, gathered into one copy
      k_gather_large<<<blocks_for(nlarge, 64), 64, 0, s>>>(large_list, nlarge,
                                                          t.n_start.get(), t.n_end.get(),
                                                          d_lrange);
      SBH_HIP(hipGetLastError());
      std::vector<uint32_t> h_lrange((size_t)nlarge * 3);
      d2h(h_lrange.data(), d_lrange, (size_t)nlarge * 3, s);
      h_large.assign(h_lrange.begin(), h_lrange.begin() + nlarge);
      h_lstart.assign(h_lrange.begin() + nlarge, h_lrange.begin() + 2 * nlarge);
      h_lend.assign(h_lrange.begin() + 2 * nlarge, h_lrange.end());
      // the chunk table {large index, begin, end, node} and the per-node
      // starts, ends, first chunks and chunk counts, in one copy
      uint32_t nchunks = 0;
      for (uint32_t li = 0; li < nlarge; ++li)
        nchunks += (h_lend[li] - h_lstart[li] + kChunk - 1) / kChunk;
      SBH_REQUIRE(nlarge <= max_large && nchunks <= max_chunks, "large-node tables overflow");
      h_tables.assign(4 * (size_t)nchunks + 4 * (size_t)nlarge, 0u);
      uint32_t* hc4 = h_tables.data();
      uint32_t* hls = hc4 + 4 * (size_t)nchunks;
      uint32_t *hle = hls + nlarge, *hc0 = hle + nlarge, *hnc = hc0 + nlarge;
      uint32_t ci = 0;
      for (uint32_t li = 0; li < nlarge; ++li) {
        hls[li] = h_lstart[li];
        hle[li] = h_lend[li];
        hc0[li] = ci;
        for (uint32_t p = h_lstart[li]; p < h_lend[li]; p += kChunk, ++ci) {
          hc4[4 * ci] = li;
          hc4[4 * ci + 1] = p;
          hc4[4 * ci + 2] = std::min(p + (uint32_t)kChunk, h_lend[li]);
          hc4[4 * ci + 3] = h_large[li];
        }
        hnc[li] = ci - hc0[li];
      }
      h2d(d_tables, h_tables.data(), h_tables.size(), s);
      const uint32_t* dt = d_tables;
      LargeCtx cx;
      cx.chunks = reinterpret_cast<const uint4*>(dt);
      cx.l_start = dt + 4 * (size_t)nchunks;
      cx.l_end = cx.l_start + nlarge;
      cx.l_chunk0 = cx.l_end + nlarge;
      cx.l_nchunk = cx.l_chunk0 + nlarge;
      cx.hist = (uint32_t*)(ab + o_hist);
      cx.cellstart = (uint32_t*)(ab + o_cellstart);
      cx.mcount = (uint32_t*)(ab + o_m);
      cx.ccnt = (uint2*)(ab + o_ccnt);
      cx.digit = digit;
      cx.Lg = Lg;
      cx.Rg = Rg;
      k_zero_u32s<<<blocks_for((size_t)nlarge * B, 256), 256, 0, s>>>(cx.hist, nlarge * B);
      k_large_digits<D><<<nchunks, kChunkThreads, 0, s>>>(cx, n_lo.get(), n_hi.get(),
                                                          t.perm.get(), src_pts.get(), k);
      SBH_HIP(hipGetLastError());
      k_large_cells<<<blocks_for(nlarge, 64), 64, 0, s>>>(cx, large_list, nlarge, B,
                                                         t.n_nchild.get(), temp_child);
      SBH_HIP(hipGetLastError());
      // The pass of cell c moves a node's points of cell c to the front of
      // its unsorted remainder.  It changes nothing in a node whose cell c is
      // empty or is its last non-empty cell (the remainder then holds only
      // cell c), so only passes some large node needs are launched.
      h_hist.resize((size_t)nlarge * B);
      d2h(h_hist.data(), cx.hist, (size_t)nlarge * B, s);
      std::vector<uint8_t> need(B, 0);
      for (uint32_t li = 0; li < nlarge; ++li) {
        const uint32_t* h = h_hist.data() + (size_t)li * B;
        int last = (int)B - 1;
        while (last > 0 && h[last] == 0) --last;
        for (int c = 0; c < last; ++c) need[c] |= h[c] != 0;
      }
      for (uint32_t c = 0; c < B; ++c) {
        if (!need[c]) continue;
        k_large_count<<<nchunks, kChunkThreads, 0, s>>>(cx, B, c);
        k_large_rank<<<nchunks, kChunkThreads, 0, s>>>(cx, B, c);
        k_large_swap<<<nchunks, kChunkThreads, 0, s>>>(cx, B, c, t.perm.get());
      }
      SBH_HIP(hipGetLastError());
    }
    if (nsmall > 0) {
      // blocks past the level's small-node count (on the device) exit at once
      k_partition_small<D><<<nsmall, kSmallThreads, 0, s>>>(
          small_list, counters, t.n_start.get(), t.n_end.get(), n_lo.get(), n_hi.get(),
          t.perm.get(), src_pts.get(), k, t.n_nchild.get(), temp_child);
      SBH_HIP(hipGetLastError());
    }
    if (ntiny > 0) {
      k_partition_tiny<D><<<blocks_for(ntiny, 4), 256, 0, s>>>(
          tiny_list, counters, t.n_start.get(), t.n_end.get(), n_lo.get(), n_hi.get(),
          t.perm.get(), src_pts.get(), k, t.n_nchild.get(), temp_child);
      SBH_HIP(hipGetLastError());
    }
    // Breadth-first numbering of the next level.
    scan_rec(t.n_nchild.get() + lb, cbase, lcount, total_dev, scan_scratch, s);
    uint32_t nnext = 0;
    d2h(&nnext, total_dev, 1, s);
    grow_all((size_t)num_nodes + nnext, num_nodes);
    k_emit<D><<<blocks_for(lcount, 128), 128, 0, s>>>(
        lb, lcount, num_nodes, cbase, k, t.n_start.get(), t.n_end.get(), t.n_parent.get(),
        t.n_first.get(), t.n_nchild.get(), t.n_next.get(), t.n_depth.get(), n_lo.get(),
        n_hi.get(), n_diam.get(), temp_child);
    SBH_HIP(hipGetLastError());
    lb = num_nodes;
    lcount = nnext;
    num_nodes += nnext;
  }
  t.level_begin.push_back(num_nodes);
  t.num_nodes = num_nodes;
  t.depth = (int)t.level_begin.size() - 2;

  // Points in leaf order.
  DevBuf<V>& pts = A::pts(t);
  DevBuf<V>& nrm = A::nrm(t);
  pts.alloc(N);
  nrm.alloc(N);
  k_gather_points<<<blocks_for(N, 256), 256, 0, s>>>(N, t.perm.get(), src_pts.get(),
                                                     src_nrm.get(), pts.get(), nrm.get());
  SBH_HIP(hipGetLastError());
  if constexpr (sizeof(T) == 8) {
    // Moments of the double build, deepest level first, and the walk records.
    t.nodeC64.alloc(num_nodes);
    t.nodeN64.alloc(num_nodes);
    t.nodeL.alloc(num_nodes);
    t.n_m64.alloc((size_t)num_nodes * 8);
    t.n_skip.alloc(num_nodes);
    t.n_size.alloc(num_nodes);
    for (int l = t.depth; l >= 0; --l) {
      uint32_t b = t.level_begin[l], cnt = t.level_begin[l + 1] - b;
      k_moments_double<D><<<blocks_for(cnt, 128), 128, 0, s>>>(
          b, cnt, t.n_start.get(), t.n_end.get(), t.n_first.get(), t.n_nchild.get(),
          pts.get(), nrm.get(), t.nodeC64.get(), t.nodeN64.get(), t.n_m64.get(),
          t.n_skip.get(), t.n_size.get());
      SBH_HIP(hipGetLastError());
    }
    k_finalize<<<blocks_for(num_nodes, 256), 256, 0, s>>>(
        num_nodes, t.n_start.get(), t.n_end.get(), t.n_first.get(), t.n_nchild.get(),
        t.n_next.get(), n_diam.get(), t.nodeL.get());
    SBH_HIP(hipGetLastError());
    t.n_dfs.alloc(num_nodes);
    t.dfsR64.alloc(num_nodes);
    t.dfsN64.alloc(num_nodes);
    k_zero_u32<<<1, 1, 0, s>>>(t.n_dfs.get());
    for (int l = 0; l <= t.depth; ++l) {
      uint32_t b = t.level_begin[l], cnt = t.level_begin[l + 1] - b;
      k_dfs_positions<<<blocks_for(cnt, 128), 128, 0, s>>>(b, cnt, t.n_first.get(),
                                                            t.n_nchild.get(), t.n_size.get(),
                                                            t.n_dfs.get());
    }
    k_dfs_scatter_double<<<blocks_for(num_nodes, 256), 256, 0, s>>>(
        num_nodes, t.n_dfs.get(), t.n_size.get(), t.nodeC64.get(), t.nodeN64.get(),
        t.nodeL.get(), n_diam.get(), pts.get(), nrm.get(), t.dfsR64.get(), t.dfsN64.get());
    SBH_HIP(hipGetLastError());
    SBH_HIP(hipStreamSynchronize(s));
    return;
  }

  // Moments, deepest level first.
  t.nodeC.alloc(num_nodes);
  t.nodeN.alloc(num_nodes);
  t.nodeL.alloc(num_nodes);
  t.n_m64.alloc((size_t)num_nodes * 8);
  t.n_skip.alloc(num_nodes);
  t.n_size.alloc(num_nodes);
  for (int l = t.depth; l >= 0; --l) {
    uint32_t b = t.level_begin[l], cnt = t.level_begin[l + 1] - b;
    k_moments<D><<<blocks_for(cnt, 128), 128, 0, s>>>(
        b, cnt, t.n_start.get(), t.n_end.get(), t.n_first.get(), t.n_nchild.get(), t.pts.get(),
        t.nrm.get(), t.nodeC.get(), t.nodeN.get(), t.n_m64.get(), t.n_skip.get(), t.n_size.get(),
        t.moments64);
    SBH_HIP(hipGetLastError());
  }
  k_finalize<<<blocks_for(num_nodes, 256), 256, 0, s>>>(
      num_nodes, t.n_start.get(), t.n_end.get(), t.n_first.get(), t.n_nchild.get(),
      t.n_next.get(), t.n_diam.get(), t.nodeL.get());
  SBH_HIP(hipGetLastError());

  // Depth-first walk records.
  t.n_dfs.alloc(num_nodes);
  t.dfsC.alloc(num_nodes);
  t.dfsN.alloc(num_nodes);
  t.dfsL.alloc(num_nodes);
  k_zero_u32<<<1, 1, 0, s>>>(t.n_dfs.get());
  for (int l = 0; l <= t.depth; ++l) {
    uint32_t b = t.level_begin[l], cnt = t.level_begin[l + 1] - b;
    k_dfs_positions<<<blocks_for(cnt, 128), 128, 0, s>>>(b, cnt, t.n_first.get(),
                                                          t.n_nchild.get(), t.n_size.get(),
                                                          t.n_dfs.get());
  }
  k_dfs_scatter<<<blocks_for(num_nodes, 256), 256, 0, s>>>(
      num_nodes, t.n_dfs.get(), t.n_size.get(), t.nodeC.get(), t.nodeN.get(), t.nodeL.get(),
      t.pts.get(), t.nrm.get(), t.dfsC.get(), t.dfsN.get(), t.dfsL.get());
  SBH_HIP(hipGetLastError());
  derive_child_sources(t);
  SBH_HIP(hipStreamSynchronize(s));
}

}  // namespace

void derive_child_sources(DeviceTree& t) {
  derive_density_grid(t);
  const uint32_t nn = t.num_nodes;
  t.kidC.alloc(nn);
  t.kidN.alloc(nn);
  t.dfsR.alloc(nn);
  t.bfsR.alloc(nn);
  t.stepR.alloc(nn);
  if (nn == 0) return;
  k_child_sources<<<blocks_for(nn, 256), 256, 0, t.stream>>>(
      nn, t.nodeC.get(), t.nodeN.get(), t.nodeL.get(), t.pts.get(), t.nrm.get(), t.dfsC.get(),
      t.dfsL.get(), t.n_start.get(), t.n_end.get(), t.n_skip.get(), t.kidC.get(), t.kidN.get(),
      t.dfsR.get(), t.bfsR.get(), t.stepR.get());
  SBH_HIP(hipGetLastError());
  derive_walk_tasks(t);
  derive_subdomain_tasks(t);
}

void derive_subdomain_tasks(DeviceTree& t) {
  t.n_sub_tasks = 0;
  if (t.level_begin.size() < 3) {
    t.sub_tasks.alloc(1);
    return;
  }
  const uint32_t s0 = t.level_begin[1], nsub = t.level_begin[2] - s0;
  if (nsub < 2 || nsub > 64) {
    t.sub_tasks.alloc(1);
    return;
  }
  std::vector<uint32_t> nch(nsub), st(nsub), en(nsub);
  hipStream_t s = t.stream;
  SBH_HIP(hipMemcpyAsync(nch.data(), t.n_nchild.get() + s0, nsub * 4, hipMemcpyDeviceToHost, s));
  SBH_HIP(hipMemcpyAsync(st.data(), t.n_start.get() + s0, nsub * 4, hipMemcpyDeviceToHost, s));
  SBH_HIP(hipMemcpyAsync(en.data(), t.n_end.get() + s0, nsub * 4, hipMemcpyDeviceToHost, s));
  SBH_HIP(hipStreamSynchronize(s));
  // a range starts at every internal subdomain but the first
  std::vector<uint4> r;
  std::vector<uint32_t> w;
  for (uint32_t i = 0; i < nsub; ++i) {
    if (nch[i] == 0) continue;
    if (r.empty()) {
      r.push_back(make_uint4(0u, 0u, 0u, 0u));
    } else {
      r.back().y = i;
      r.push_back(make_uint4(i, 0u, (uint32_t)r.size(), 0u));
    }
    w.push_back(en[i] - st[i]);
  }
  t.sub_tasks.alloc(1);
  if (r.size() < 2) return;
  r.back().y = nsub;
  std::vector<uint32_t> rank(r.size());
  for (uint32_t i = 0; i < rank.size(); ++i) rank[i] = i;
  std::stable_sort(rank.begin(), rank.end(), [&](uint32_t a, uint32_t b) { return w[a] > w[b]; });
  std::vector<uint4> tasks;
  for (uint32_t i : rank) tasks.push_back(r[i]);
  t.sub_tasks.alloc(tasks.size());
  SBH_HIP(hipMemcpyAsync(t.sub_tasks.get(), tasks.data(), tasks.size() * sizeof(uint4),
                         hipMemcpyHostToDevice, s));
  SBH_HIP(hipStreamSynchronize(s));
  t.n_sub_tasks = (uint32_t)tasks.size();
}

// The cells of the density grid (cell = ((x - lo) * inv) per axis, clamped)
// of `per` consecutive points per thread, counted: points are in the tree's
// leaf order, so a run of them mostly shares a cell and adds it once (one
// atomic per change of cell, not per point: a dense cell takes few).
__global__ void k_density(const float4* __restrict__ pts, uint32_t n, uint32_t per, int dim,
                          float lx, float ly, float lz, float ix, float iy, float iz, uint32_t g,
                          uint32_t* __restrict__ dens) {
  const uint64_t i0 = (uint64_t)(blockIdx.x * blockDim.x + threadIdx.x) * per;
  if (i0 >= n) return;
  const uint32_t i1 = (uint32_t)(i0 + per < n ? i0 + per : n);
  const float lo[3] = {lx, ly, lz}, inv[3] = {ix, iy, iz};
  uint32_t run_cell = 0xFFFFFFFFu, run = 0;
  for (uint32_t i = (uint32_t)i0; i < i1; ++i) {
    const float4 p = pts[i];
    const float q[3] = {p.x, p.y, p.z};
    uint32_t cell = 0;
    for (int a = dim - 1; a >= 0; --a) {
      float c = (q[a] - lo[a]) * inv[a];
      c = fminf(fmaxf(c, 0.0f), (float)(g - 1));
      cell = cell * g + (uint32_t)c;
    }
    if (cell != run_cell) {
      if (run) atomicAdd(&dens[run_cell], run);
      run_cell = cell;
      run = 0;
    }
    ++run;
  }
  if (run) atomicAdd(&dens[run_cell], run);
}

__global__ void k_zero_dens(uint32_t* __restrict__ p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

void derive_density_grid(DeviceTree& t) {
  // 32^3 cells in 3D, 128^2 in 2D (about 2^15 counts)
  const uint32_t g = t.dim == 3 ? 32u : 128u;
  const uint32_t cells = t.dim == 3 ? g * g * g : g * g;
  t.dens.alloc(cells);
  t.dens_g = g;
  float d2 = 0.0f;
  for (int a = 0; a < 3; ++a) {
    const float w = a < t.dim ? t.hi[a] - t.lo[a] : 0.0f;
    t.dens_inv[a] = w > 0.0f ? (float)g / w : 0.0f;
    d2 += (w / (float)g) * (w / (float)g);
  }
  t.dens_diag = d2 > 0.0f ? std::sqrt(d2) : 1.0f;
  k_zero_dens<<<blocks_for(cells, 256), 256, 0, t.stream>>>(t.dens.get(), cells);
  SBH_HIP(hipGetLastError());
  if (t.n > 0 && !t.is_double) {
    // runs of up to 64 points, fewer while that leaves under 2^14 threads
    const uint32_t per = std::min<uint32_t>(64u, std::max<uint32_t>(1u, t.n >> 14));
    k_density<<<blocks_for(blocks_for(t.n, per), 256), 256, 0, t.stream>>>(
        t.pts.get(), t.n, per, t.dim, t.lo[0], t.lo[1], t.lo[2], t.dens_inv[0], t.dens_inv[1],
        t.dens_inv[2], g, t.dens.get());
    SBH_HIP(hipGetLastError());
  }
}

void derive_walk_tasks(DeviceTree& t) {
  // Balanced frontiers: starting from the root's children, the heaviest
  // subtree (most nodes) that has children, lies above depth kMaxWalkAnc and
  // leaves the frontier within kMaxWalkTasks tasks is replaced by its
  // children, until none is left.  The frontier is kept each time its size
  // first reaches a power of two (4, 8, ...), and at the end, smallest
  // first.  A frontier's tasks are stored heaviest first, each with its
  // depth-first rank as the slot of its partial sum.
  t.walk_frontiers.clear();
  const int levels = (int)t.level_begin.size() - 1;
  // nodes of depths 0 .. top on the host (at most kWalkHostNodes of them)
  constexpr uint32_t kWalkHostNodes = 1u << 18;
  int top = 0;
  while (top + 1 < levels && top + 1 <= kMaxWalkAnc && t.level_begin[top + 2] <= kWalkHostNodes)
    ++top;
  if (top == 0) {
    t.walk_tasks.alloc(1);
    return;
  }
  const uint32_t m = t.level_begin[top + 1];
  std::vector<uint32_t> nch(m), first(m), dfs(m), size(m);
  hipStream_t s = t.stream;
  SBH_HIP(hipMemcpyAsync(nch.data(), t.n_nchild.get(), m * 4, hipMemcpyDeviceToHost, s));
  SBH_HIP(hipMemcpyAsync(first.data(), t.n_first.get(), m * 4, hipMemcpyDeviceToHost, s));
  SBH_HIP(hipMemcpyAsync(dfs.data(), t.n_dfs.get(), m * 4, hipMemcpyDeviceToHost, s));
  SBH_HIP(hipMemcpyAsync(size.data(), t.n_size.get(), m * 4, hipMemcpyDeviceToHost, s));
  SBH_HIP(hipStreamSynchronize(s));
  struct Sub {
    uint32_t v;                   // breadth-first node id
    uint32_t depth;               // = number of ancestors
    uint32_t path[kMaxWalkAnc];   // ancestors' breadth-first ids, root first
  };
  std::vector<Sub> front;
  for (uint32_t c = first[0]; c < first[0] + nch[0]; ++c) {
    Sub u{};
    u.v = c;
    u.depth = 1;
    u.path[0] = 0;
    front.push_back(u);
  }
  std::vector<WalkTask> all;
  size_t kept_size = 0;
  auto keep = [&]() {
    if (front.size() < 2 || front.size() == kept_size) return;
    kept_size = front.size();
    std::vector<Sub> f = front;
    std::sort(f.begin(), f.end(), [&](const Sub& a, const Sub& b) { return dfs[a.v] < dfs[b.v]; });
    std::vector<WalkTask> w(f.size());
    for (size_t k = 0; k < f.size(); ++k) {
      WalkTask& x = w[k];
      x = WalkTask{};
      x.v0 = dfs[f[k].v];
      const uint32_t rope = dfs[f[k].v] + size[f[k].v];
      x.vstop = rope < t.num_nodes ? rope : kEnd;
      x.nanc = f[k].depth;
      x.part = (uint32_t)k;
      for (uint32_t d = 0; d < f[k].depth; ++d) {
        x.anc[d] = dfs[f[k].path[d]];
        // tasks below one ancestor are consecutive in depth-first order
        const bool below_prev = k > 0 && f[k - 1].depth > d && f[k - 1].path[d] == f[k].path[d];
        if (!below_prev) x.first |= 1u << d;
      }
    }
    // heaviest first (launch order); ties in depth-first order
    std::stable_sort(w.begin(), w.end(), [&](const WalkTask& a, const WalkTask& b) {
      const uint32_t sa = (a.vstop == kEnd ? t.num_nodes : a.vstop) - a.v0;
      const uint32_t sb = (b.vstop == kEnd ? t.num_nodes : b.vstop) - b.v0;
      return sa > sb;
    });
    t.walk_frontiers.emplace_back((uint32_t)all.size(), (uint32_t)w.size());
    all.insert(all.end(), w.begin(), w.end());
  };
  for (;;) {
    if (front.size() >= 4 && (front.size() & (front.size() - 1)) == 0) keep();
    size_t best = front.size();
    for (size_t k = 0; k < front.size(); ++k) {
      const Sub& u = front[k];
      if (u.depth >= (uint32_t)top || nch[u.v] == 0 ||
          front.size() - 1 + nch[u.v] > kMaxWalkTasks)
        continue;
      if (best == front.size() || size[u.v] > size[front[best].v]) best = k;
    }
    if (best == front.size()) break;
    // a power of two passed over is kept before the frontier grows past it
    const size_t grown = front.size() - 1 + nch[front[best].v];
    for (size_t p = 4; p <= kMaxWalkTasks; p <<= 1)
      if (front.size() < p && grown > p) keep();
    const Sub u = front[best];
    front.erase(front.begin() + best);
    for (uint32_t c = first[u.v]; c < first[u.v] + nch[u.v]; ++c) {
      Sub w = u;
      w.v = c;
      w.path[u.depth] = u.v;
      w.depth = u.depth + 1;
      front.push_back(w);
    }
  }
  keep();
  t.walk_tasks.alloc(all.empty() ? 1 : all.size());
  if (!all.empty()) {
    SBH_HIP(hipMemcpyAsync(t.walk_tasks.get(), all.data(), all.size() * sizeof(WalkTask),
                           hipMemcpyHostToDevice, s));
    SBH_HIP(hipStreamSynchronize(s));
  }
}

size_t exclusive_scan_scratch(uint32_t n) { return scan_scratch_size(n); }

void exclusive_scan(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* total_dev,
                    uint32_t* scratch, hipStream_t s) {
  scan_rec(in, out, n, total_dev, scratch, s);
}

void build_device_tree(DeviceTree& t, const float* points, const float* normals,
                       const float* masses, float* build_ms) {
  auto t0 = std::chrono::steady_clock::now();
  if (t.dim == 2)
    build_impl<2, float>(t, points, normals, masses);
  else
    build_impl<3, float>(t, points, normals, masses);
  auto t1 = std::chrono::steady_clock::now();
  if (build_ms) *build_ms = std::chrono::duration<float, std::milli>(t1 - t0).count();
}

void build_device_tree_double(DeviceTree& t, const double* points, const double* normals,
                              const double* masses, float* build_ms) {
  auto t0 = std::chrono::steady_clock::now();
  if (t.dim == 2)
    build_impl<2, double>(t, points, normals, masses);
  else
    build_impl<3, double>(t, points, normals, masses);
  auto t1 = std::chrono::steady_clock::now();
  if (build_ms) *build_ms = std::chrono::duration<float, std::milli>(t1 - t0).count();
}

}  // namespace sbh
