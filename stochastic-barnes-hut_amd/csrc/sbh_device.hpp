// Device-side building blocks shared by the tree build and the evaluation
// kernels: source/node record layout, the three summation kernels, Philox and
// the exact-fp32 helpers the parity-critical paths use.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sbh {

constexpr int kWave = 64;
constexpr uint32_t kEnd = 0xFFFFFFFFu;   // rope terminator
constexpr uint32_t kLeafBit = 0x80000000u;

// A walk record: centroid and mass, then {diameter bits, first child or
// first point, rope, child count | leaf flag and point count}.  The
// breadth-first copy the fast sampler reads holds the diameter's hardware
// reciprocal in place of the diameter.
struct alignas(32) WalkRecord {
  float4 c;
  uint4 l;
};

// One task of a Barnes-Hut walk split over a frontier of disjoint subtrees
// that covers the tree: the subtree at depth-first positions [v0, vstop) and
// the depth-first positions of its nanc ancestors, root first.  Bit d of
// `first` is set when this is the first task (in depth-first order) below
// anc[d]: a query that accepts anc[d] takes that moment there.  `part` is
// the task's depth-first rank: its slot among the walk's partial sums, which
// are added in that order (tasks are stored, and launched, heaviest first).
constexpr int kMaxWalkAnc = 8;
constexpr uint32_t kMaxWalkTasks = 256;
struct alignas(64) WalkTask {
  uint32_t v0, vstop, nanc, first, part;
  uint32_t anc[kMaxWalkAnc];
  uint32_t pad[3];
};

// include/kernel.h:24-28 (bound kernels) and :61-83 (unbound: Helmholtz,
// and the gravity gradient as one scalar kernel per axis)
enum Kernel : int {
  GRAVITY = 0,
  WINDING_NUMBER = 1,
  SMOOTH_DIST = 2,
  HELMHOLTZ_X = 3,
  HELMHOLTZ_Y = 4,
  GRAVITY_GRAD_X = 5,
  GRAVITY_GRAD_Y = 6,
  GRAVITY_GRAD_Z = 7,
};
constexpr bool is_helmholtz(int k) { return k == HELMHOLTZ_X || k == HELMHOLTZ_Y; }
constexpr bool is_gradient(int k) { return k >= GRAVITY_GRAD_X && k <= GRAVITY_GRAD_Z; }

// Constants of include/kernel.h:31-59 in the reference's float evaluation:
// 4 * (float)M_PI and 2 * (float)M_PI are exact power-of-two multiples.
constexpr float kPiF = 3.14159265358979323846f;
constexpr float kFourPi = 4.0f * kPiF;
constexpr float kTwoPi = 2.0f * kPiF;

// ---------------------------------------------------------------------------
// Exact fp32 kernels: one rounding per operation, operations in the
// reference's order (Eigen fixed-size reductions sum a0 + (a1 + a2)),
// correctly rounded sqrt and divide.  Used where results must match the
// oracle bit for bit: Barnes-Hut and the stochastic estimator.
// ---------------------------------------------------------------------------
template <int D>
__device__ __forceinline__ float sqnorm_exact(float dx, float dy, float dz) {
#pragma clang fp contract(off)
  if constexpr (D == 3) {
    return dx * dx + (dy * dy + dz * dz);
  } else {
    return dx * dx + dy * dy;
  }
}

// src = {x, y, z, m}, nrm = {nx, ny, nz, *}, q = {qx, qy, qz}
template <int D, int K>
__device__ __forceinline__ float kernel_exact(float4 src, float4 nrm, float alpha, float qx,
                                              float qy, float qz) {
#pragma clang fp contract(off)
  float dx = src.x - qx;
  float dy = src.y - qy;
  float dz = (D == 3) ? (src.z - qz) : 0.0f;
  if constexpr (K == GRAVITY) {
    float dist = __builtin_sqrtf(sqnorm_exact<D>(dx, dy, dz));
    return -src.w / (dist + 1e-1f);
  } else if constexpr (K == WINDING_NUMBER) {
    if constexpr (D == 3) {
      float dist = __builtin_sqrtf(sqnorm_exact<3>(dx, dy, dz));
      float qdn = nrm.x * dx + (nrm.y * dy + nrm.z * dz);
      return src.w * qdn / (kFourPi * (dist * dist * dist + 1e-5f));
    } else {
      float dsq = sqnorm_exact<2>(dx, dy, dz);
      float qdn = nrm.x * dx + nrm.y * dy;
      return src.w * qdn / (kTwoPi * (dsq + 1e-5f));
    }
  } else if constexpr (is_helmholtz(K)) {
    // m cos(freq 2 pi d) / (d + 0.1), freq 2 (X, cos) or 1 (Y, sin): kernel.h:68-83
    float dist = __builtin_sqrtf(sqnorm_exact<D>(dx, dy, dz));
    const float freq = (K == HELMHOLTZ_X) ? 2.0f : 1.0f;
    float arg = freq * 2.0f * kPiF * dist;
    float c = (K == HELMHOLTZ_X) ? cosf(arg) : sinf(arg);
    return src.w * c / (dist + 1e-1f);
  } else if constexpr (is_gradient(K)) {
    // -m / (d + 0.1)^3 * (q - p), component K - GRAVITY_GRAD_X: kernel.h:61-66
    float dist = __builtin_sqrtf(sqnorm_exact<D>(dx, dy, dz));
    float t = dist + 1e-1f;
    float s = -src.w / (t * t * t);
    float comp = (K == GRAVITY_GRAD_X) ? -dx : (K == GRAVITY_GRAD_Y) ? -dy : -dz;
    return s * comp;
  } else {
    float dist = __builtin_sqrtf(sqnorm_exact<D>(dx, dy, dz));
    return expf(-alpha * dist);
  }
}

// Barnes-Hut opening ratio ||q - c|| / diam (include/tree3d.h:134-140).
template <int D>
__device__ __forceinline__ float ratio_exact(float4 c, float diam, float qx, float qy, float qz) {
#pragma clang fp contract(off)
  float dx = qx - c.x;
  float dy = qy - c.y;
  float dz = (D == 3) ? (qz - c.z) : 0.0f;
  return __builtin_sqrtf(sqnorm_exact<D>(dx, dy, dz)) / diam;
}

// Separation q - p of a query from a source or node centroid and its
// squared norm with fused squares.  The fast kernels, the opening ratio and
// the opening test all start from it, so a walk that tests a node and then
// takes its term computes the distance once.
struct Sep {
  float dx, dy, dz, d2;
};
template <int D>
__device__ __forceinline__ Sep separation(float4 p, float qx, float qy, float qz) {
  Sep s;
  s.dx = qx - p.x;
  s.dy = qy - p.y;
  s.dz = (D == 3) ? (qz - p.z) : 0.0f;
  s.d2 = (D == 3) ? __builtin_fmaf(s.dx, s.dx, __builtin_fmaf(s.dy, s.dy, s.dz * s.dz))
                  : __builtin_fmaf(s.dx, s.dx, s.dy * s.dy);
  return s;
}
__device__ __forceinline__ float fast_dist(const Sep& s) { return __builtin_amdgcn_sqrtf(s.d2); }

// ratio >= beta, decided exactly as ratio_exact would decide it, in two
// parts: a fast estimate (fused squares, hardware sqrt and reciprocal; within
// 7e-7 relative of ratio_exact) settles every case farther than 2e-6 from
// beta (1 = far, 0 = near, -1 = undecided); the correctly rounded formula
// decides the rest.
// inv_diam = hardware reciprocal of diam, computed by the caller; s and dist
// = separation<D>(c, q) and fast_dist(s).
__device__ __forceinline__ int opens_far_fast(const Sep& s, float dist, float inv_diam,
                                              float beta) {
  float r = dist * inv_diam;
  const bool ok = s.d2 >= 1e-30f;
  return (ok && r >= beta * 1.000002f) ? 1 : (ok && r <= beta * 0.999998f) ? 0 : -1;
}
template <int D>
__device__ __forceinline__ int opens_far_fast(float4 c, float inv_diam, float beta, float qx,
                                              float qy, float qz) {
  const Sep s = separation<D>(c, qx, qy, qz);
  return opens_far_fast(s, fast_dist(s), inv_diam, beta);
}
template <int D>
__device__ __forceinline__ bool opens_far_exact(float4 c, float diam, float beta, float qx,
                                                float qy, float qz) {
#pragma clang fp contract(off)
  float dx = qx - c.x;
  float dy = qy - c.y;
  float dz = (D == 3) ? (qz - c.z) : 0.0f;
  return __builtin_sqrtf(sqnorm_exact<D>(dx, dy, dz)) / diam >= beta;
}

// ---------------------------------------------------------------------------
// Fast kernels for brute force and the walks: hardware sqrt / reciprocal /
// exp2 (about one ulp each), fused multiply-adds allowed.  Sums stay within
// 1e-5 relative of the exact formulation.  They take the separation q - p:
// where the formula needs p - q (the winding number's normal projection, the
// gradient's direction) the sign is folded into the mass, which is exact.
// ---------------------------------------------------------------------------
// acc + the fast kernel of source src at separation s (dist = fast_dist(s)),
// with the term's final multiply fused into the accumulation.
template <int D, int K>
__device__ __forceinline__ float kernel_fast_acc(float acc, float4 src, float4 nrm, float alpha,
                                                 const Sep& s, float dist) {
  if constexpr (K == GRAVITY) {
    return __builtin_fmaf(-src.w, __builtin_amdgcn_rcpf(dist + 1e-1f), acc);
  } else if constexpr (K == WINDING_NUMBER) {
    // n . (p - q) = -(n . s)
    if constexpr (D == 3) {
      float nds = __builtin_fmaf(nrm.x, s.dx, __builtin_fmaf(nrm.y, s.dy, nrm.z * s.dz));
      float den = kFourPi * __builtin_fmaf(s.d2, dist, 1e-5f);
      return __builtin_fmaf(-src.w * nds, __builtin_amdgcn_rcpf(den), acc);
    } else {
      float nds = __builtin_fmaf(nrm.x, s.dx, nrm.y * s.dy);
      return __builtin_fmaf(-src.w * nds, __builtin_amdgcn_rcpf(kTwoPi * (s.d2 + 1e-5f)), acc);
    }
  } else if constexpr (is_helmholtz(K)) {
    // the hardware sine / cosine take revolutions: cos(2 pi freq d) = cos_rev(freq d)
    float rev = __builtin_amdgcn_fractf((K == HELMHOLTZ_X ? 2.0f : 1.0f) * dist);
    float c = (K == HELMHOLTZ_X) ? __builtin_amdgcn_cosf(rev) : __builtin_amdgcn_sinf(rev);
    return __builtin_fmaf(src.w * c, __builtin_amdgcn_rcpf(dist + 1e-1f), acc);
  } else if constexpr (is_gradient(K)) {
    // component of -m / (d + 0.1)^3 * (q - p)
    float r = __builtin_amdgcn_rcpf(dist + 1e-1f);
    float comp = (K == GRAVITY_GRAD_X) ? s.dx : (K == GRAVITY_GRAD_Y) ? s.dy : s.dz;
    return __builtin_fmaf(-src.w * (r * r * r), comp, acc);
  } else {
    // exp(-a d) = 2^(-a d log2 e); alpha is pre-scaled by log2(e) by the caller.
    return acc + __builtin_amdgcn_exp2f(-alpha * dist);
  }
}
template <int D, int K>
__device__ __forceinline__ float kernel_fast_acc(float acc, float4 src, float4 nrm, float alpha,
                                                 float qx, float qy, float qz) {
  const Sep s = separation<D>(src, qx, qy, qz);
  return kernel_fast_acc<D, K>(acc, src, nrm, alpha, s, fast_dist(s));
}

// The fast kernel's term alone.
template <int D, int K>
__device__ __forceinline__ float kernel_fast(float4 src, float4 nrm, float alpha, const Sep& s,
                                             float dist) {
  if constexpr (K == GRAVITY) {
    return -src.w * __builtin_amdgcn_rcpf(dist + 1e-1f);
  } else if constexpr (K == WINDING_NUMBER) {
    if constexpr (D == 3) {
      float nds = __builtin_fmaf(nrm.x, s.dx, __builtin_fmaf(nrm.y, s.dy, nrm.z * s.dz));
      float den = kFourPi * __builtin_fmaf(s.d2, dist, 1e-5f);
      return -src.w * nds * __builtin_amdgcn_rcpf(den);
    } else {
      float nds = __builtin_fmaf(nrm.x, s.dx, nrm.y * s.dy);
      return -src.w * nds * __builtin_amdgcn_rcpf(kTwoPi * (s.d2 + 1e-5f));
    }
  } else if constexpr (is_helmholtz(K)) {
    float rev = __builtin_amdgcn_fractf((K == HELMHOLTZ_X ? 2.0f : 1.0f) * dist);
    float c = (K == HELMHOLTZ_X) ? __builtin_amdgcn_cosf(rev) : __builtin_amdgcn_sinf(rev);
    return src.w * c * __builtin_amdgcn_rcpf(dist + 1e-1f);
  } else if constexpr (is_gradient(K)) {
    float r = __builtin_amdgcn_rcpf(dist + 1e-1f);
    float comp = (K == GRAVITY_GRAD_X) ? s.dx : (K == GRAVITY_GRAD_Y) ? s.dy : s.dz;
    return -src.w * (r * r * r) * comp;
  } else {
    return __builtin_amdgcn_exp2f(-alpha * dist);
  }
}
template <int D, int K>
__device__ __forceinline__ float kernel_fast(float4 src, float4 nrm, float alpha, float qx,
                                             float qy, float qz) {
  const Sep s = separation<D>(src, qx, qy, qz);
  return kernel_fast<D, K>(src, nrm, alpha, s, fast_dist(s));
}

// Opening ratio ||q - c|| / diam with hardware sqrt and the caller's
// hardware reciprocal of diam (a few ulp from ratio_exact).
template <int D>
__device__ __forceinline__ float ratio_fast(float4 c, float inv_diam, float qx, float qy,
                                            float qz) {
  return fast_dist(separation<D>(c, qx, qy, qz)) * inv_diam;
}

// float64 formulation of the same kernels (error reporting only).
template <int D, int K>
__device__ __forceinline__ double kernel_f64(float4 src, float4 nrm, float alpha, float qx,
                                             float qy, float qz) {
  double dx = (double)src.x - (double)qx;
  double dy = (double)src.y - (double)qy;
  double dz = (D == 3) ? ((double)src.z - (double)qz) : 0.0;
  double dsq = dx * dx + dy * dy + dz * dz;
  double dist = sqrt(dsq);
  if constexpr (K == GRAVITY) {
    return -(double)src.w / (dist + (double)1e-1f);
  } else if constexpr (K == WINDING_NUMBER) {
    double qdn = (double)nrm.x * dx + (double)nrm.y * dy + (D == 3 ? (double)nrm.z * dz : 0.0);
    if constexpr (D == 3)
      return (double)src.w * qdn / ((double)kFourPi * (dist * dsq + (double)1e-5f));
    else
      return (double)src.w * qdn / ((double)kTwoPi * (dsq + (double)1e-5f));
  } else if constexpr (is_helmholtz(K)) {
    const double pi = (double)kPiF;
    const double c = (K == HELMHOLTZ_X) ? cos(4.0 * pi * dist) : sin(2.0 * pi * dist);
    return (double)src.w * c / (dist + (double)1e-1f);
  } else if constexpr (is_gradient(K)) {
    const double t = dist + (double)1e-1f;
    const double comp = (K == GRAVITY_GRAD_X) ? -dx : (K == GRAVITY_GRAD_Y) ? -dy : -dz;
    return -(double)src.w / (t * t * t) * comp;
  } else {
    return exp(-(double)alpha * dist);
  }
}

// ---------------------------------------------------------------------------
// Double build: the reference compiled with USE_DOUBLE (include/utils.h:13-29,
// FLOAT = double, literals FP(x) = x, PI = M_PI).  Points, cell bounds,
// moments, terms, opening ratios, roulette and sums in float64, one rounding
// per operation in the reference's order (no contraction).
// ---------------------------------------------------------------------------
struct alignas(32) dvec4 {
  double x, y, z, w;
};
constexpr double kPiD = 3.14159265358979323846;

// Depth-first walk record of the double build: centroid and mass (a one-point
// leaf: its point), diameter, first child (depth-first position) or first
// point, rope, child count | leaf flag and point count.
struct alignas(64) WalkRecordD {
  dvec4 c;
  double diam;
  uint32_t first, rope, info, pad;
};

template <int D>
__device__ __forceinline__ double sqnorm_exact(double dx, double dy, double dz) {
#pragma clang fp contract(off)
  if constexpr (D == 3) {
    return dx * dx + (dy * dy + dz * dz);
  } else {
    return dx * dx + dy * dy;
  }
}

// include/kernel.h:31-83 in float64; src = {x, y, z, m}, nrm = {nx, ny, nz, *}
template <int D, int K>
__device__ __forceinline__ double kernel_double(const dvec4& src, const dvec4& nrm, double alpha,
                                                double qx, double qy, double qz) {
#pragma clang fp contract(off)
  double dx = src.x - qx;
  double dy = src.y - qy;
  double dz = (D == 3) ? (src.z - qz) : 0.0;
  if constexpr (K == GRAVITY) {
    double dist = __builtin_sqrt(sqnorm_exact<D>(dx, dy, dz));
    return -src.w / (dist + 1e-1);
  } else if constexpr (K == WINDING_NUMBER) {
    if constexpr (D == 3) {
      double dist = __builtin_sqrt(sqnorm_exact<3>(dx, dy, dz));
      double qdn = nrm.x * dx + (nrm.y * dy + nrm.z * dz);
      return src.w * qdn / (4.0 * kPiD * (dist * dist * dist + 1e-5));
    } else {
      double dsq = sqnorm_exact<2>(dx, dy, dz);
      double qdn = nrm.x * dx + nrm.y * dy;
      return src.w * qdn / (2.0 * kPiD * (dsq + 1e-5));
    }
  } else if constexpr (is_helmholtz(K)) {
    double dist = __builtin_sqrt(sqnorm_exact<D>(dx, dy, dz));
    const double freq = (K == HELMHOLTZ_X) ? 2.0 : 1.0;
    double arg = freq * 2.0 * kPiD * dist;
    double c = (K == HELMHOLTZ_X) ? cos(arg) : sin(arg);
    return src.w * c / (dist + 1e-1);
  } else if constexpr (is_gradient(K)) {
    double dist = __builtin_sqrt(sqnorm_exact<D>(dx, dy, dz));
    double t = dist + 1e-1;
    double s = -src.w / (t * t * t);
    double comp = (K == GRAVITY_GRAD_X) ? -dx : (K == GRAVITY_GRAD_Y) ? -dy : -dz;
    return s * comp;
  } else {
    double dist = __builtin_sqrt(sqnorm_exact<D>(dx, dy, dz));
    return exp(-alpha * dist);
  }
}

// Opening ratio ||q - c|| / diam (include/tree3d.h:134-140) in float64.
template <int D>
__device__ __forceinline__ double ratio_double(const dvec4& c, double diam, double qx, double qy,
                                               double qz) {
#pragma clang fp contract(off)
  double dx = qx - c.x;
  double dy = qy - c.y;
  double dz = (D == 3) ? (qz - c.z) : 0.0;
  return __builtin_sqrt(sqnorm_exact<D>(dx, dy, dz)) / diam;
}

// Russian-roulette uniform in (0, 1] of the double build: all 32 bits of the
// word (the float build keeps the top 24).
__device__ __forceinline__ double rr_uniform_double(uint32_t w) {
  return ((double)w + 1.0) * 0x1p-32;
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel Random Numbers: As
// Easy as 1, 2, 3", SC'11).  10 rounds, key bumped by the Weyl constants
// between rounds.
// ---------------------------------------------------------------------------
struct Philox4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ Philox4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2,
                                                 uint32_t c3, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += W0;
      k1 += W1;
    }
    uint32_t lo0 = M0 * c0, hi0 = __umulhi(M0, c0);
    uint32_t lo1 = M1 * c2, hi1 = __umulhi(M1, c2);
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
  return {c0, c1, c2, c3};
}

// Russian-roulette uniform in (0, 1] from one 32-bit word.
__device__ __forceinline__ float rr_uniform(uint32_t w) {
  return (float)((w >> 8) + 1u) * 0x1p-24f;
}

// Uniform index in [0, n) from a 64-bit draw (multiply-high).
__device__ __forceinline__ uint32_t scale_draw(uint32_t lo, uint32_t hi, uint32_t n) {
  uint64_t u = ((uint64_t)hi << 32) | lo;
  return (uint32_t)__umul64hi(u, (uint64_t)n);
}

__device__ __forceinline__ uint32_t uniform_u32(uint32_t v) {
  return __builtin_amdgcn_readfirstlane(v);
}

// A set of lanes of the wavefront, one bit per lane, held in scalar
// registers.  Predicates that combine others (a && b, a && !b) are kept as
// masks and combined with scalar operations: the ballot of a comparison is
// the comparison's own result, while the ballot of a combined per-lane bool
// costs two vector instructions (materialise it, compare it).  lane_in reads
// this lane's bit as a per-lane condition for selects, at no cost.
// Orders this wavefront's LDS writes before its later LDS reads by other
// lanes (a wavefront-scope fence and wave barrier; no workgroup barrier).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

using LaneMask = uint64_t;
__device__ __forceinline__ bool lane_in(LaneMask m) {
  return __builtin_amdgcn_inverse_ballot_w64(m);
}

// Element i of an array no kernel writes while it runs, read through the
// constant address space: with a wave-uniform i this is a scalar load.
template <typename T>
__device__ __forceinline__ T load_invariant(const T* p, uint32_t i) {
  static_assert(sizeof(T) % 4 == 0, "whole 32-bit words");
  using Word = const __attribute__((address_space(4))) uint32_t;
  Word* w = (Word*)(p + i);
  T out;
  uint32_t* o = reinterpret_cast<uint32_t*>(&out);
#pragma unroll
  for (uint32_t k = 0; k < sizeof(T) / 4; ++k) o[k] = w[k];
  return out;
}

}  // namespace sbh
