// Brute force with the squared distances on the matrix cores.
//
// For a wavefront of 64 spatially ordered queries and a tile of 16 sources,
// with every position taken relative to a centre c:
//     d2[i][j] = |q'_i|^2 + sum_k A[i][k] B[k][j]           q' = q - c, p' = p - c
// with |q'|^2 entering through the accumulator input, in one of two operand
// formats:
//   fp32   A[i] = (q'_x, q'_y, q'_z, 1), B[j] = (-2p'_x, -2p'_y, -2p'_z, |p'|^2):
//          one V_MFMA_F32_16X16X4_F32 per 16 x 16 block;
//   bf16   every coordinate split into bf16 hi + lo parts (x = h + l up to
//          2^-18 |x|) and |p'|^2 into three bf16 parts; the 16 K slots carry
//          the products h.h, h.l, l.h, l.l of each axis and the three parts of
//          |p'|^2, so the fp32 accumulation returns -2 q'.p' + |p'|^2 to
//          ~2^-17 |q'||p'|: one V_MFMA_F32_16X16X16_BF16 per 16 x 16 block.
// The workgroup (256 queries) stages 256 sources at a time in LDS together
// with their B operands relative to its own centre, computed once per source
// instead of once per wavefront and tile.  bf16 operands use that centre
// directly; fp32 operands are re-centred on the wavefront (c_w = c + delta:
// -2(p' - delta) and |p' - delta|^2 = |p'|^2 - 2p'.delta + |delta|^2, four
// fused operations), which keeps the near field the wavefront's.  Centring
// keeps the expansion's cancellation small: the error of d2 stays a few ulp
// of (|q'| + |p'|)^2 (fp32) or ~1e-5 |q'||p'| (bf16), small against d2 once
// d >= 2r (fp32) or d >= 4r (bf16), r the radius of the queries around c.
// Pairs closer than that (the near field) get their distance directly in
// fp32; a tile whose sources all lie beyond the near field of every query
// (known from the smallest staged |p'|^2 of the tile) skips that test.  The
// remaining per-pair work (sqrt, add, reciprocal or exp2, fma) runs on the
// VALU in one loop shared by far and near pairs.  Partial sums per lane are
// kept for 4 x 4 query rows per stage of sources, reduced over the 16 lanes
// that share them by a halving butterfly, and accumulated in float64 by one
// lane per row.
//
// Results match the direct formulation to ~1e-6 relative (checked against
// the oracle and float64 in the GPU tests).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "sbh_device.hpp"
#include "sbh_eval.hpp"
#include "sbh_runtime.hpp"

namespace sbh {
namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef short s16x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf16_round(float x) { return (float)(__bf16)x; }

// bf16 bits of four values (each already a bf16 value, so the conversion is exact)
__device__ __forceinline__ s16x4 bf16x4(float a, float b, float c, float d) {
  s16x4 r;
  r[0] = __builtin_bit_cast(short, (__bf16)a);
  r[1] = __builtin_bit_cast(short, (__bf16)b);
  r[2] = __builtin_bit_cast(short, (__bf16)c);
  r[3] = __builtin_bit_cast(short, (__bf16)d);
  return r;
}

// Item t of K-slot group g (lane / 16) of the bf16 operands; slot 4g + t:
//   g0: x.hh x.hl x.lh y.hh   g1: y.hl y.lh z.hh z.hl
//   g2: z.lh s.h  s.m  s.l    g3: x.ll y.ll z.ll 0
// A slots hold the query factor, B slots the source factor (-2 p' parts, or
// the parts of |p'|^2 against A = 1).
__device__ __forceinline__ s16x4 a_slots(uint32_t g, float xh, float xl, float yh, float yl,
                                         float zh, float zl) {
  if (g == 0) return bf16x4(xh, xh, xl, yh);
  if (g == 1) return bf16x4(yh, yl, zh, zh);
  if (g == 2) return bf16x4(zl, 1.0f, 1.0f, 1.0f);
  return bf16x4(xl, yl, zl, 0.0f);
}
__device__ __forceinline__ s16x4 b_slots(uint32_t g, float xh, float xl, float yh, float yl,
                                         float zh, float zl, float sh, float sm, float sl) {
  const float v0 = g == 0 ? xh : g == 1 ? yl : g == 2 ? zh : xl;
  const float v1 = g == 0 ? xl : g == 1 ? yh : g == 2 ? sh : yl;
  const float v2 = g == 0 ? xh : g == 1 ? zh : g == 2 ? sm : zl;
  const float v3 = g == 0 ? yh : g == 1 ? zl : g == 2 ? sl : 0.0f;
  return bf16x4(v0, v1, v2, v3);
}

constexpr int kThreads = 256;   // 4 wavefronts
constexpr int kStage = 256;     // sources staged in LDS per step

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = fminf(v, __shfl_xor(v, off, 64));
  return v;
}

template <int K>
__device__ __forceinline__ float far_term(float d2, float m, float alpha_k) {
  const float dist = __builtin_amdgcn_sqrtf(d2);
  if constexpr (K == GRAVITY)
    return -m * __builtin_amdgcn_rcpf(dist + 1e-1f);
  else
    return __builtin_amdgcn_exp2f(-alpha_k * dist);
}

template <int K, bool BF16>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(BF16 ? 6 : 5, BF16 ? 6 : 5))) void k_brute_mfma(
    const float4* __restrict__ pts, uint32_t n, float alpha_k, const float* __restrict__ qx,
    const float* __restrict__ qy, const float* __restrict__ qz, uint32_t nq,
    const uint32_t* __restrict__ order, uint32_t per_split, double* __restrict__ partial,
    float* __restrict__ out) {
  __shared__ float4 stage[kStage];  // sources {x, y, z, m}
  // B operands of the staged sources, centred on the workgroup: fp32
  // {-2p'_x, -2p'_y, -2p'_z, |p'|^2}, or the four bf16 K-slot groups
  __shared__ float4 stageB[BF16 ? 1 : kStage];
  __shared__ s16x4 stageBb[BF16 ? kStage * 4 : 1];
  __shared__ float tmin[kStage / 16];  // smallest |p'|^2 of each 16-source tile
  __shared__ float4 wq[kThreads];      // each wavefront's 64 query positions
  __shared__ float bred[kThreads / 64][8];
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wave = threadIdx.x >> 6;
  const uint32_t wbase = threadIdx.x & ~63u;
  const uint32_t base = (blockIdx.x * (kThreads / 64) + wave) * 64;
  const uint32_t slot = base + lane;
  const bool live = slot < nq;
  // lanes past the end repeat the workgroup's first query (inside its box)
  const uint32_t qi = order[live ? slot : blockIdx.x * kThreads];
  const float x = qx[qi], y = qy[qi], z = qz[qi];

  // workgroup centre (bounding-box midpoint) and radius
  {
    const float v[6] = {wave_min(x), wave_min(y), wave_min(z),
                        wave_max(x), wave_max(y), wave_max(z)};
    if (lane == 0)
      for (int a = 0; a < 6; ++a) bred[wave][a] = v[a];
  }
  __syncthreads();
  float lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = bred[0][a];
    hi[a] = bred[0][3 + a];
    for (int w = 1; w < kThreads / 64; ++w) {
      lo[a] = fminf(lo[a], bred[w][a]);
      hi[a] = fmaxf(hi[a], bred[w][3 + a]);
    }
  }
  const float cx = 0.5f * (lo[0] + hi[0]), cy = 0.5f * (lo[1] + hi[1]),
              cz = 0.5f * (lo[2] + hi[2]);
  const float ux = x - cx, uy = y - cy, uz = z - cz;
  const float u2 = ux * ux + uy * uy + uz * uz;
  {
    const float wr = wave_max(u2);
    if (lane == 0) bred[wave][6] = wr;
  }
  __syncthreads();
  float rb2 = bred[0][6];
  for (int w = 1; w < kThreads / 64; ++w) rb2 = fmaxf(rb2, bred[w][6]);
  // The expansion's distances are centred on c: the workgroup's centre for
  // bf16 operands; for fp32 operands the wavefront's own centre, whose B
  // operands follow from the staged workgroup-centred ones
  // (p - c_w = p' - delta, delta = c_w - c), so the near field is the
  // smaller wavefront's.
  float wcx = cx, wcy = cy, wcz = cz, r2 = rb2, dl = 0.0f;
  if constexpr (!BF16) {
    wcx = 0.5f * (wave_min(x) + wave_max(x));
    wcy = 0.5f * (wave_min(y) + wave_max(y));
    wcz = 0.5f * (wave_min(z) + wave_max(z));
    const float vx = x - wcx, vy = y - wcy, vz = z - wcz;
    r2 = wave_max(vx * vx + vy * vy + vz * vz);
    const float ex = wcx - cx, ey = wcy - cy, ez = wcz - cz;
    dl = __builtin_sqrtf(ex * ex + ey * ey + ez * ez);
  }
  const float dx0 = wcx - cx, dy0 = wcy - cy, dz0 = wcz - cz;  // delta
  const float dd = dx0 * dx0 + dy0 * dy0 + dz0 * dz0;          // |delta|^2
  // pairs closer than sqrt(near2) are recomputed directly; a tile whose
  // staged sources all lie at least dl + r + sqrt(near2) from the workgroup
  // centre has none
  const float near2 = (BF16 ? 16.0f : 4.5f) * r2;
  const float tile_far = 1.001f * (dl + (BF16 ? 5.0f : 3.1214f) * __builtin_sqrtf(r2));
  const float tile_far2 = tile_far * tile_far;
  const float qwx = x - wcx, qwy = y - wcy, qwz = z - wcz;
  wq[threadIdx.x] = make_float4(x, y, z, qwx * qwx + qwy * qwy + qwz * qwz);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

  // A operands of the four 16-query blocks: lane holds A[i = lane % 16][k = lane / 16]
  const uint32_t i16 = lane & 15, k4 = lane >> 4;
  float a[4];
  s16x4 ab[4];
  f32x4 cinit[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const float4 qq = wq[wbase + 16 * t + i16];
    const float qx = qq.x - wcx, qy = qq.y - wcy, qz = qq.z - wcz;
    if constexpr (BF16) {
      const float xh = bf16_round(qx), yh = bf16_round(qy), zh = bf16_round(qz);
      ab[t] = a_slots(k4, xh, bf16_round(qx - xh), yh, bf16_round(qy - yh), zh,
                      bf16_round(qz - zh));
    } else {
      a[t] = k4 == 0 ? qx : k4 == 1 ? qy : k4 == 2 ? qz : 1.0f;
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) cinit[t][v] = wq[wbase + 16 * t + 4 * k4 + v].w;
  }

  // float64 running sum of query row (16t + 4(lane / 16) + v) for t, v with
  // 4t + v = lane % 16: each stage's fp32 partial sums are reduced over the 16
  // lanes that share a row and added here
  double acc64 = 0.0;

  const uint32_t s0 = blockIdx.y * per_split;
  const uint32_t s1 = min(n, s0 + per_split);
  for (uint32_t sb = s0; sb < s1; sb += kStage) {
    const uint32_t cnt = min((uint32_t)kStage, s1 - sb);
    __syncthreads();
    {
      // zero-mass padding far away: its terms are exactly 0 (gravity: mass 0;
      // smooth distance: exp2 of a huge negative number)
      const float4 p = threadIdx.x < cnt ? pts[sb + threadIdx.x]
                                         : make_float4(3.0e18f, 3.0e18f, 3.0e18f, 0.0f);
      stage[threadIdx.x] = p;
      const float px = p.x - cx, py = p.y - cy, pz = p.z - cz;
      const float s2 = px * px + py * py + pz * pz;
      if constexpr (BF16) {
        const float xh = bf16_round(px), yh = bf16_round(py), zh = bf16_round(pz);
        const float sh = bf16_round(s2), sm = bf16_round(s2 - sh);
        const float xl = bf16_round(px - xh), yl = bf16_round(py - yh),
                    zl = bf16_round(pz - zh), sl = bf16_round(s2 - sh - sm);
#pragma unroll
        for (uint32_t g = 0; g < 4; ++g)
          stageBb[threadIdx.x * 4 + g] =
              b_slots(g, -2.0f * xh, -2.0f * xl, -2.0f * yh, -2.0f * yl, -2.0f * zh,
                      -2.0f * zl, sh, sm, sl);
      } else {
        stageB[threadIdx.x] = make_float4(-2.0f * px, -2.0f * py, -2.0f * pz, s2);
      }
      float m = s2;
#pragma unroll
      for (int off = 8; off >= 1; off >>= 1) m = fminf(m, __shfl_xor(m, off, 16));
      if ((threadIdx.x & 15) == 0) tmin[threadIdx.x >> 4] = m;
    }
    __syncthreads();
    float acc[4][4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[t][v] = 0.0f;
    const uint32_t ntiles = (cnt + 15) / 16;
    for (uint32_t st = 0; st < ntiles; ++st) {
      const uint32_t j = st * 16 + i16;
      f32x4 d[4];
      if constexpr (BF16) {
        const s16x4 bb = stageBb[j * 4 + k4];
#pragma unroll
        for (int t = 0; t < 4; ++t)
          d[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ab[t], bb, cinit[t], 0, 0, 0);
      } else {
        // B of the wavefront-centred expansion: -2(p' - delta) on the axes,
        // |p' - delta|^2 = |p'|^2 + (-2p').delta + |delta|^2 in the last slot
        const float4 sb = stageB[j];
        const float axis = k4 == 0 ? sb.x : k4 == 1 ? sb.y : sb.z;
        const float dax = k4 == 0 ? dx0 : k4 == 1 ? dy0 : dz0;
        const float s2w = __builtin_fmaf(sb.x, dx0, __builtin_fmaf(sb.y, dy0,
                                         __builtin_fmaf(sb.z, dz0, sb.w + dd)));
        const float b = k4 < 3 ? __builtin_fmaf(2.0f, dax, axis) : s2w;
#pragma unroll
        for (int t = 0; t < 4; ++t)
          d[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[t], b, cinit[t], 0, 0, 0);
      }
      if (tmin[st] < tile_far2) {  // uniform: the tile may hold near pairs
        float dmin = d[0][0];
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
          for (int v = 0; v < 4; ++v) dmin = fminf(dmin, d[t][v]);
        if (__ballot(dmin < near2)) {
          // near field: those pairs' distances directly in fp32 (>= 0;
          // the expansion's others are >= near2 > 0)
          const float4 p = stage[j];
#pragma unroll
          for (int t = 0; t < 4; ++t) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              if (d[t][v] < near2) {
                const float4 qq = wq[wbase + 16 * t + 4 * k4 + v];
                const float dx = p.x - qq.x, dy = p.y - qq.y, dz = p.z - qq.z;
                d[t][v] = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
              }
            }
          }
        }
      }
      const float mass = stage[j].w;
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[t][v] += far_term<K>(d[t][v], mass, alpha_k);
    }
    // lane (i16) gets the sum over its 16-lane row group of accumulator i16
    // (index 4t + v): a butterfly that halves the values each step
    float h8[8], h4[4], h2[2];
    {
      const bool up = (i16 & 8) != 0;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const float lo = acc[q >> 2][q & 3], hi = acc[(q + 8) >> 2][(q + 8) & 3];
        h8[q] = (up ? hi : lo) + __shfl_xor(up ? lo : hi, 8, 16);
      }
    }
    {
      const bool up = (i16 & 4) != 0;
#pragma unroll
      for (int q = 0; q < 4; ++q)
        h4[q] = (up ? h8[q + 4] : h8[q]) + __shfl_xor(up ? h8[q] : h8[q + 4], 4, 16);
    }
    {
      const bool up = (i16 & 2) != 0;
#pragma unroll
      for (int q = 0; q < 2; ++q)
        h2[q] = (up ? h4[q + 2] : h4[q]) + __shfl_xor(up ? h4[q] : h4[q + 2], 2, 16);
    }
    {
      const bool up = (i16 & 1) != 0;
      acc64 += (double)((up ? h2[1] : h2[0]) + __shfl_xor(up ? h2[0] : h2[1], 1, 16));
    }
  }

  const uint32_t row = 16 * (i16 >> 2) + 4 * k4 + (i16 & 3);
  const uint32_t srow = base + row;
  if (srow < nq) {
    const uint32_t q = order[srow];
    if (partial)
      partial[(size_t)blockIdx.y * nq + q] = acc64;
    else
      out[q] = (float)acc64;
  }
}

__global__ void k_reduce_splits_mfma(const double* partial, uint32_t nsplit, uint32_t nq,
                                     float* out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  double s = 0.0;
  for (uint32_t k = 0; k < nsplit; ++k) s += partial[(size_t)k * nq + i];
  out[i] = (float)s;
}

inline uint32_t nblk(uint64_t n, uint32_t t) { return (uint32_t)((n + t - 1) / t); }

}  // namespace

uint32_t brute_mfma_split_count(uint32_t n, uint32_t nq) {
  // a dozen or more rounds of the 5-6 resident wavefronts per SIMD: with
  // only a few rounds, the last, partly filled one idles much of the GPU
  const uint32_t qblocks = nblk(nq, kThreads);
  const uint32_t ns = std::max(1u, 16384u / std::max(1u, qblocks));
  return std::min(ns, std::max(1u, n / (8 * kStage)));
}

void run_brute_mfma(const DeviceTree& t, int kernel, bool bf16, const float* q, uint32_t nq,
                    const uint32_t* order, float* out, double* partial, uint32_t nsplit,
                    hipStream_t s) {
  SBH_REQUIRE(t.dim == 3 && (kernel == GRAVITY || kernel == SMOOTH_DIST),
              "the matrix-core brute force covers 3D gravity and smooth distance");
  const float* qx = q;
  const float* qy = q + nq;
  const float* qz = q + 2ull * nq;
  const uint32_t per_split = nblk(nblk(t.n, nsplit), kStage) * kStage;
  dim3 grid(nblk(nq, kThreads), nsplit);
  const float alpha_k = t.alpha * 1.4426950408889634f;
  double* part = nsplit > 1 ? partial : nullptr;
  auto launch = [&](auto kern) {
    kern<<<grid, kThreads, 0, s>>>(t.pts.get(), t.n, alpha_k, qx, qy, qz, nq, order, per_split,
                                   part, out);
  };
  if (kernel == GRAVITY)
    bf16 ? launch(k_brute_mfma<GRAVITY, true>) : launch(k_brute_mfma<GRAVITY, false>);
  else
    bf16 ? launch(k_brute_mfma<SMOOTH_DIST, true>) : launch(k_brute_mfma<SMOOTH_DIST, false>);
  SBH_HIP(hipGetLastError());
  if (nsplit > 1) {
    k_reduce_splits_mfma<<<nblk(nq, 256), 256, 0, s>>>(partial, nsplit, nq, out);
    SBH_HIP(hipGetLastError());
  }
}

}  // namespace sbh
