// Brute force with the squared distances on the matrix cores.
//
// For a wavefront of 64 spatially ordered queries with centre c (and
// radius r = max |q - c|), and a tile of 16 sources:
//     d2[i][j] = |q'_i|^2 + sum_k A[i][k] B[k][j]
//     A[i] = (q'_x, q'_y, q'_z, 1)          q' = q - c
//     B[j] = (-2 p'_x, -2 p'_y, -2 p'_z, |p'|^2)   p' = p - c
// one V_MFMA_F32_16X16X4_F32 per 16x16 block, |q'|^2 entering through the
// accumulator input.  Centring on the wavefront keeps the expansion's
// cancellation small: the fp32 error of d2 is a few ulp of (|q'| + |p'|)^2,
// within a few ulp of d2 itself once d >= 2r.  Pairs closer than that (the
// near field, d2 < 4.5 r^2) are recomputed directly in fp32.  The remaining
// per-pair work (sqrt, add, reciprocal or exp2, fma) runs on the VALU.
// Partial sums per lane are kept for 4 x 4 query rows and reduced over the 16
// lanes that share them at the end.
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

template <int K>
__global__ __launch_bounds__(kThreads) void k_brute_mfma(
    const float4* __restrict__ pts, uint32_t n, float alpha_k, const float* __restrict__ qx,
    const float* __restrict__ qy, const float* __restrict__ qz, uint32_t nq,
    const uint32_t* __restrict__ order, uint32_t per_split, double* __restrict__ partial,
    float* __restrict__ out) {
  __shared__ float4 stage[kStage];
  __shared__ float4 wq[kThreads];  // each wavefront's 64 query positions
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wbase = threadIdx.x & ~63u;
  const uint32_t base = (blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6)) * 64;
  const uint32_t slot = base + lane;
  const bool live = slot < nq;
  const uint32_t qi = order[live ? slot : (base < nq ? base : 0)];
  const float x = qx[qi], y = qy[qi], z = qz[qi];

  // wavefront centre (bounding-box midpoint) and radius
  const float cx = 0.5f * (wave_min(x) + wave_max(x));
  const float cy = 0.5f * (wave_min(y) + wave_max(y));
  const float cz = 0.5f * (wave_min(z) + wave_max(z));
  const float ux = x - cx, uy = y - cy, uz = z - cz;
  const float r2 = wave_max(ux * ux + uy * uy + uz * uz);
  const float near2 = 4.5f * r2;
  wq[threadIdx.x] = make_float4(x, y, z, ux * ux + uy * uy + uz * uz);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

  // A operands of the four 16-query blocks: lane holds A[i = lane % 16][k = lane / 16]
  const uint32_t i16 = lane & 15, k4 = lane >> 4;
  float a[4];
  f32x4 cinit[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const float4 qq = wq[wbase + 16 * t + i16];
    const float comp = k4 == 0 ? qq.x - cx : k4 == 1 ? qq.y - cy : k4 == 2 ? qq.z - cz : 1.0f;
    a[t] = comp;
#pragma unroll
    for (int v = 0; v < 4; ++v) cinit[t][v] = wq[wbase + 16 * t + 4 * k4 + v].w;
  }

  double acc64[4][4];
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int v = 0; v < 4; ++v) acc64[t][v] = 0.0;

  const uint32_t s0 = blockIdx.y * per_split;
  const uint32_t s1 = min(n, s0 + per_split);
  for (uint32_t sb = s0; sb < s1; sb += kStage) {
    const uint32_t cnt = min((uint32_t)kStage, s1 - sb);
    __syncthreads();
    if (threadIdx.x < cnt) {
      stage[threadIdx.x] = pts[sb + threadIdx.x];
    } else {
      // zero-mass far-away padding: contributes nothing to gravity; smooth
      // distance is masked below
      stage[threadIdx.x] = make_float4(3.0e18f, 3.0e18f, 3.0e18f, 0.0f);
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
      const float4 p = stage[j];
      const bool valid = (sb + j) < s1;
      const float px = p.x - cx, py = p.y - cy, pz = p.z - cz;
      const float b = k4 == 0 ? -2.0f * px
                    : k4 == 1 ? -2.0f * py
                    : k4 == 2 ? -2.0f * pz
                              : px * px + py * py + pz * pz;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const f32x4 d = __builtin_amdgcn_mfma_f32_16x16x4f32(a[t], b, cinit[t], 0, 0, 0);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          float d2 = d[v];
          if (d2 < near2) {
            // near field: direct fp32 distance
            const float4 qq = wq[wbase + 16 * t + 4 * k4 + v];
            const float dx = p.x - qq.x, dy = p.y - qq.y, dz = p.z - qq.z;
            d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
          }
          float term = far_term<K>(fmaxf(d2, 0.0f), p.w, alpha_k);
          if (K != GRAVITY && !valid) term = 0.0f;
          acc[t][v] += term;
        }
      }
    }
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc64[t][v] += (double)acc[t][v];
  }

  // reduce over the 16 lanes that hold the same query rows (same lane / 16)
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      double s = acc64[t][v];
#pragma unroll
      for (int off = 8; off >= 1; off >>= 1) s += __shfl_xor(s, off, 16);
      acc64[t][v] = s;
    }
  if (i16 == 0) {
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const uint32_t row = 16 * t + 4 * k4 + v;
        const uint32_t s = base + row;
        if (s < nq) {
          const uint32_t q = order[s];
          if (partial)
            partial[(size_t)blockIdx.y * nq + q] = acc64[t][v];
          else
            out[q] = (float)acc64[t][v];
        }
      }
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
  const uint32_t qblocks = nblk(nq, kThreads);
  const uint32_t ns = std::max(1u, 4096u / std::max(1u, qblocks));
  return std::min(ns, std::max(1u, n / (8 * kStage)));
}

void run_brute_mfma(const DeviceTree& t, int kernel, const float* q, uint32_t nq,
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
  if (kernel == GRAVITY)
    k_brute_mfma<GRAVITY><<<grid, kThreads, 0, s>>>(t.pts.get(), t.n, alpha_k, qx, qy, qz, nq,
                                                    order, per_split, part, out);
  else
    k_brute_mfma<SMOOTH_DIST><<<grid, kThreads, 0, s>>>(t.pts.get(), t.n, alpha_k, qx, qy, qz,
                                                        nq, order, per_split, part, out);
  SBH_HIP(hipGetLastError());
  if (nsplit > 1) {
    k_reduce_splits_mfma<<<nblk(nq, 256), 256, 0, s>>>(partial, nsplit, nq, out);
    SBH_HIP(hipGetLastError());
  }
}

}  // namespace sbh
