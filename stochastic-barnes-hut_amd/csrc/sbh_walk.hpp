// Per-wavefront tree walks shared by the one-shot evaluation kernels
// (eval.hip) and the serving kernel (serve.hip): deterministic Barnes-Hut and
// the stochastic control-variate estimator over a device-resident tree.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sbh_device.hpp"

namespace sbh {

// Device pointers of one resident tree.
struct TreeView {
  const float4* nodeC;  // breadth-first records
  const uint4* nodeL;
  const float4* nodeN;
  const float4* dfsC;   // depth-first records (Barnes-Hut walk)
  const uint4* dfsL;
  const float4* dfsN;
  const uint32_t* n_start;
  const uint32_t* n_end;
  const uint32_t* n_skip;
  const float4* pts;
  const float4* nrm;
  const WalkRecord* dfsR;  // depth-first walk records, centroid and record together
  const WalkRecord* bfsR;  // nodeC and nodeL together (breadth-first)
  const uint4* stepR;      // {range start, range end, chain target, 0}
  const float4* kidC;  // term source of each node as a child (one-point leaf: its point)
  const float4* kidN;
  float alpha;    // smooth-distance exponent
  float alpha_k;  // alpha * log2(e), for exp2
  // point counts of a grid of dg^dim cells over the root box: the cell of
  // (x, y, z) is ((x - dlo) * dinv) per axis, clamped; ddiag = a cell's diagonal
  const uint32_t* dens;
  float dlo[3], dinv[3];
  float ddiag;
  uint32_t dg;
};

struct LaneResult {
  float value;
  int32_t points;  // point terms evaluated
  int32_t nodes;   // node terms evaluated
};

__device__ __forceinline__ void load_query(const float* qx, const float* qy, const float* qz,
                                           uint32_t i, int dim, float& x, float& y, float& z) {
  x = qx[i];
  y = qy[i];
  z = dim == 3 ? qz[i] : 0.0f;
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

// acc + term: exact (reference order, no fusion) or fast (the term's final
// multiply fused into the accumulation).
template <int D, int K, bool EXACT>
__device__ __forceinline__ float accum(float acc, float4 src, float4 nrm, float alpha,
                                       float alpha_k, float x, float y, float z) {
#pragma clang fp contract(off)
  if constexpr (EXACT)
    return acc + kernel_exact<D, K>(src, nrm, alpha, x, y, z);
  else
    return kernel_fast_acc<D, K>(acc, src, nrm, alpha_k, x, y, z);
}

// Opening ratio of a sampler path node and the roulette probability
// min(cur_beta / ratio, 1) (include/tree_sampler_gpu.h:83-92): correctly
// rounded in the exact mode; with hardware sqrt / reciprocal in the fast mode,
// as the reference's fast-math GPU build computes them.  Decisions and weights
// use the same value either way, so the estimator stays unbiased.
template <int D, bool EXACT>
__device__ __forceinline__ float path_ratio(float4 c, float diam, float x, float y, float z) {
  if constexpr (EXACT)
    return ratio_exact<D>(c, diam, x, y, z);
  else
    return ratio_fast<D>(c, __builtin_amdgcn_rcpf(diam), x, y, z);
}
template <bool EXACT>
__device__ __forceinline__ float traverse_prob(float cur_beta, float ratio) {
  if constexpr (EXACT)
    return fminf(cur_beta / ratio, 1.0f);
  else  // exactly 1 whenever the correctly rounded quotient is >= 1
    return ratio <= cur_beta ? 1.0f : fminf(cur_beta * __builtin_amdgcn_rcpf(ratio), 1.0f);
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
// A wavefront whose live queries are spread over a region that holds, at the
// density of the grid cells its queries lie in (their mean point count), more
// than kLaneSpread of the tree's points (mean count x (spread / cell
// diagonal)^D; spread = the diagonal of the queries' bounding box) walks with
// every lane on its own path (bh_walk_lanes): the shared walk would visit the
// union of the lanes' walks, which such lanes share little of.
// Split walks' tasks on subtrees of fewer than kLaneMinNodes nodes skip the
// measurement (bh_walk_task): their walks are short either way.  A tree
// without the grid (dens null) always walks shared.
constexpr float kLaneSpread = 512.0f;
constexpr uint32_t kLaneMinNodes = 1024;
template <int D>
__device__ __forceinline__ bool lanes_spread_out(const TreeView& t, bool live, float x, float y,
                                                 float z) {
  if (t.dens == nullptr || t.dg == 0) return false;
  constexpr float kInf = __builtin_huge_valf();
  const float q[3] = {x, y, z};
  float lo[3], hi[3];
  uint32_t cell = 0;
#pragma unroll
  for (int a = D - 1; a >= 0; --a) {
    lo[a] = live ? q[a] : kInf;
    hi[a] = live ? q[a] : -kInf;
    float c = (q[a] - t.dlo[a]) * t.dinv[a];
    c = fminf(fmaxf(c, 0.0f), (float)(t.dg - 1));  // NaN-free: clamps to [0, dg - 1]
    cell = cell * t.dg + (uint32_t)c;
  }
  float count = live ? (float)t.dens[cell] : 0.0f;
  float lanes = live ? 1.0f : 0.0f;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
    for (int a = 0; a < D; ++a) {
      lo[a] = fminf(lo[a], __shfl_xor(lo[a], off, 64));
      hi[a] = fmaxf(hi[a], __shfl_xor(hi[a], off, 64));
    }
    count += __shfl_xor(count, off, 64);
    lanes += __shfl_xor(lanes, off, 64);
  }
  if (!(lanes > 0.0f)) return false;  // no live lane
  float d2 = 0.0f;
#pragma unroll
  for (int a = 0; a < D; ++a) d2 += (hi[a] - lo[a]) * (hi[a] - lo[a]);
  const float r = __builtin_sqrtf(d2) / t.ddiag;
  const float frac = (D == 3) ? r * r * r : r * r;
  return uniform_u32((count / lanes) * frac > kLaneSpread ? 1u : 0u) != 0;
}

// The per-query fast walk with every lane on its own path: the same rope
// walk, the same opening decisions and the same terms in the same
// (depth-first) order as bh_walk<D, K, 1, false>, so the same values; each
// lane loads its own records (vector loads).  A wavefront then takes as many
// steps as its longest walk, where the shared walk takes as many as the union
// of its lanes' walks: far fewer when the lanes' queries are spread out
// (sparse queries over a deep tree).
// v0 is per lane; acc0, cnt0, nodes0: the sums a walk begun in the shared
// mode carries over (it switches to this one part-way, see bh_walk).
template <int D, int K, bool COUNT = true>
__device__ __forceinline__ LaneResult bh_walk_lanes(const TreeView& t, float beta, bool live,
                                                    float x, float y, float z, uint32_t v0 = 0,
                                                    uint32_t vstop = kEnd, float acc0 = 0.0f,
                                                    int32_t cnt0 = 0, int32_t nodes0 = 0) {
  const WalkRecord* __restrict__ rec = t.dfsR;
  const float4* __restrict__ nodeN = t.dfsN;
  const float beta_hi = beta * 1.000002f, beta_lo = beta * 0.999998f;
  float acc = acc0;
  int32_t cnt = cnt0, cnt_nodes = nodes0;
  uint32_t v = live ? v0 : vstop;
  while (v != vstop) {
    const WalkRecord R = rec[v];
    const float4 c = R.c;
    const uint4 l = R.l;
    const bool leaf = (int32_t)l.w < 0;
    const uint32_t np = leaf ? (l.w & ~kLeafBit) : 0u;
    const Sep s = separation<D>(c, x, y, z);
    const float dist = fast_dist(s);
    bool far = true;
    if (!leaf) {
      // the shared walk's opening test (fast ratio, exact where undecided)
      float r = dist * __builtin_amdgcn_rcpf(__uint_as_float(l.x));
      r = s.d2 >= 1e-30f ? r : __builtin_nanf("");
      far = r >= beta_hi;
      if (!far && !(r <= beta_lo)) far = opens_far_exact<D>(c, __uint_as_float(l.x), beta, x, y, z);
    }
    if (l.w > (kLeafBit | 1u)) {  // a leaf of several points
      for (uint32_t j = l.y; j < l.y + np; ++j)
        acc = accum<D, K, false>(acc, t.pts[j], t.nrm[j], t.alpha, t.alpha_k, x, y, z);
    } else if (far) {
      acc = kernel_fast_acc<D, K>(acc, c, nodeN[v], t.alpha_k, s, dist);
    }
    if constexpr (COUNT) {
      cnt += far ? (int32_t)np : 0;
      cnt_nodes += (far && !leaf) ? 1 : 0;
    }
    v = far ? l.z : l.y;
  }
  return {acc, cnt, cnt_nodes};
}

// COUNT: also count the terms (the counting mode).  [v0, vstop): the walk
// starts at depth-first position v0 and ends where the ropes reach vstop (the
// whole tree by default; a subtree's position and its rope walk the subtree).
// HYBRID (per query, fast terms): the walk takes bh_walk_lanes instead when
// lanes_spread_out; same values either way.
// (spread_check false: walk shared without asking, for walks known to be
// short.)
template <int D, int K, int GROUP, bool EXACT, bool COUNT = true, bool HYBRID = false>
__device__ __forceinline__ LaneResult bh_walk(const TreeView& t, float beta, bool live, float x,
                                              float y, float z, uint32_t v0 = 0,
                                              uint32_t vstop = kEnd, bool spread_check = true) {
  static_assert(!HYBRID || (GROUP == 1 && !EXACT), "the lane walk is per query, fast terms");
  // Lane sets (live, active, far, reject, accept) are lane masks in scalar
  // registers, combined with scalar operations and read per lane with
  // lane_in (see LaneMask).
  // records addressed by 32-bit byte offsets (a tree has < 2^27 nodes)
  const char* __restrict__ rec = reinterpret_cast<const char*>(t.dfsR);
  const float4* __restrict__ nodeN = t.dfsN;
  const float beta_hi = beta * 1.000002f, beta_lo = beta * 0.999998f;
  float acc = 0.0f;
  int32_t cnt = 0, cnt_nodes = 0;
  const LaneMask live_m = __ballot(live);
  LaneMask active = live_m;
  uint32_t resume = kEnd;
  uint32_t v = v0;
  if constexpr (HYBRID) {
    if (spread_check && lanes_spread_out<D>(t, live, x, y, z))
      return bh_walk_lanes<D, K, COUNT>(t, beta, live, x, y, z, v0, vstop);
  }
  WalkRecord R = *reinterpret_cast<const WalkRecord*>(rec + (v0 << 5));
  while (v != vstop) {
    v = uniform_u32(v);
    const float4 c = R.c;
    const uint4 l = R.l;
    active |= live_m & __ballot(resume == v);
    const bool leaf = (int32_t)l.w < 0;
    // the record (c, l) of v was loaded at the end of the previous step, as
    // soon as v was known.  The step is branch-free per lane (selects, not
    // exec-mask branches).  Every field of the record is read on every path,
    // so it arrives in one scalar load.
    const float diam = __uint_as_float(l.x);
    const uint32_t np = leaf ? (l.w & ~kLeafBit) : 0u;
    // the separation from the record's centroid (or one-point leaf's point),
    // shared by the opening test and the fast term
    const Sep s = separation<D>(c, x, y, z);
    const float dist = fast_dist(s);
    LaneMask rej = 0;
    if (!leaf) {
      // far: ratio >= beta.  The fast ratio (within 7e-7 of the correctly
      // rounded one) settles every lane farther than 2e-6 from beta; the
      // correctly rounded formula decides the rest, so every lane opens
      // exactly the nodes the reference walk opens.
      // (a lane whose squared distance is below 1e-30 gets a NaN ratio,
      // which neither comparison settles)
      float r = dist * __builtin_amdgcn_rcpf(diam);
      r = s.d2 >= 1e-30f ? r : __builtin_nanf("");
      LaneMask far = __ballot(r >= beta_hi);
      const LaneMask undecided = active & ~(far | __ballot(r <= beta_lo));
      if (undecided) far |= undecided & __ballot(opens_far_exact<D>(c, diam, beta, x, y, z));
      rej = active & ~far;
    }
    // per query, a lane's own test decides; a vote group needs every lane
    LaneMask accept;
    if constexpr (GROUP == 1) {
      accept = active & ~rej;
    } else if constexpr (GROUP == 32) {
      constexpr LaneMask lo = 0x00000000FFFFFFFFull, hi = 0xFFFFFFFF00000000ull;
      accept = active & (((rej & lo) ? 0ull : lo) | ((rej & hi) ? 0ull : hi));
    } else {
      accept = rej ? 0ull : active;
    }
    const bool acc_lane = lane_in(accept);
    if (l.w > (kLeafBit | 1u)) {  // a leaf of several points
      if (acc_lane) {
        for (uint32_t j = l.y; j < l.y + np; ++j)
          acc = accum<D, K, EXACT>(acc, t.pts[j], t.nrm[j], t.alpha, t.alpha_k, x, y, z);
      }
    } else {
      // an internal node's moment, or a one-point leaf's point (its record
      // holds the point); computed for every lane (cheaper than a branch on
      // the scalar unit), kept where the lane accepts v
      float with;
      if constexpr (EXACT)
        with = accum<D, K, true>(acc, c, nodeN[v], t.alpha, t.alpha_k, x, y, z);
      else
        with = kernel_fast_acc<D, K>(acc, c, nodeN[v], t.alpha_k, s, dist);
      acc = acc_lane ? with : acc;
    }
    if constexpr (COUNT) {
      cnt += acc_lane ? (int32_t)np : 0;
      cnt_nodes += (acc_lane && !leaf) ? 1 : 0;
    }
    // descend when some lane rejects v (the accepting lanes resume at v's
    // rope), else follow the rope
    const uint32_t next = uniform_u32(rej ? l.y : l.z);
    resume = lane_in(rej ? accept : 0ull) ? l.z : resume;
    if constexpr (GROUP == 1)
      active = rej ? rej : active;  // the lanes that reject v are the ones left
    else
      active = rej ? (active & ~accept) : active;
    if (next != vstop) R = *reinterpret_cast<const WalkRecord*>(rec + (next << 5));
    v = next;
  }
  return {acc, cnt, cnt_nodes};
}

// One task of the per-query Barnes-Hut walk split over a frontier of
// disjoint subtrees (for launches too small to fill the GPU): task k's
// subtree for this lane's query.  A query whose opening test accepts an
// ancestor of the subtree takes the moment of the topmost ancestor it accepts
// in the first task below that ancestor, and nothing else, as the whole walk
// would.  .value is the task's part of the sum, for slot tasks[k].part; the
// parts are added in slot (depth-first) order (sum_bh_parts).  Fast terms
// only.
template <int D, int K, bool COUNT>
__device__ __forceinline__ LaneResult bh_walk_task(const TreeView& t,
                                                   const WalkTask* __restrict__ tasks,
                                                   uint32_t k, float beta, bool live, float x,
                                                   float y, float z) {
  const WalkTask w = tasks[k];
  float acc = 0.0f;
  bool taken = false;
  int32_t nodes = 0;
#pragma unroll
  for (int d = 0; d < kMaxWalkAnc; ++d) {
    if (d < (int)w.nanc) {
      const uint32_t a = w.anc[d];
      const WalkRecord r = t.dfsR[a];
      // the ancestor's opening test, decided exactly as the whole walk decides it
      const float diam = __uint_as_float(r.l.x);
      int far = opens_far_fast<D>(r.c, __builtin_amdgcn_rcpf(diam), beta, x, y, z);
      if (far < 0) far = (int)opens_far_exact<D>(r.c, diam, beta, x, y, z);
      const bool take = !taken && far != 0;
      if (((w.first >> d) & 1u) && __ballot(take)) {
        const float with = accum<D, K, false>(acc, r.c, t.dfsN[a], t.alpha, t.alpha_k, x, y, z);
        acc = take ? with : acc;
        nodes = take ? 1 : nodes;
      }
      taken = taken || far != 0;
    }
  }
  // a subtree of fewer than kLaneMinNodes nodes is walked shared without
  // measuring the queries' spread (its walks are short either way)
  const LaneResult r = bh_walk<D, K, 1, false, COUNT, true>(
      t, beta, live && !taken, x, y, z, w.v0, w.vstop, w.vstop - w.v0 >= kLaneMinNodes);
  if (taken) return {acc, 0, nodes};
  return r;
}

// The parts of a split walk for launch slot `slot`, added in depth-first
// order: partial[k * nq + slot] is the part of the task whose .part is k.
__device__ __forceinline__ float sum_bh_parts(const float* __restrict__ partial, uint32_t nsplit,
                                              uint32_t nq, uint32_t slot) {
  // eight loads in flight at a time; the additions keep task order
  constexpr uint32_t kBatch = 8;
  float s = 0.0f;
  uint32_t k = 0;
  for (; k + kBatch <= nsplit; k += kBatch) {
    float v[kBatch];
#pragma unroll
    for (uint32_t j = 0; j < kBatch; ++j) v[j] = partial[(size_t)(k + j) * nq + slot];
#pragma unroll
    for (uint32_t j = 0; j < kBatch; ++j) s = (k + j == 0) ? v[j] : s + v[j];
  }
  for (; k < nsplit; ++k) {
    const float v = partial[(size_t)k * nq + slot];
    s = k == 0 ? v : s + v;
  }
  return s;
}

// The parts of a split estimate ({Barnes-Hut part, control-variate part} per
// subdomain), each summed in subdomain order, then added.
__device__ __forceinline__ float sum_stochastic_parts(const float2* __restrict__ partial,
                                                      uint32_t nsplit, uint32_t nq,
                                                      uint32_t slot) {
  // eight loads in flight at a time; the additions keep subdomain order
  constexpr uint32_t kBatch = 8;
  float bh = 0.0f, mlcv = 0.0f;
  uint32_t k = 0;
  for (; k + kBatch <= nsplit; k += kBatch) {
    float2 v[kBatch];
#pragma unroll
    for (uint32_t j = 0; j < kBatch; ++j) v[j] = partial[(size_t)(k + j) * nq + slot];
#pragma unroll
    for (uint32_t j = 0; j < kBatch; ++j) {
      bh += v[j].x;
      mlcv += v[j].y;
    }
  }
  for (; k < nsplit; ++k) {
    const float2 p = partial[(size_t)k * nq + slot];
    bh += p.x;
    mlcv += p.y;
  }
  return bh + mlcv;
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
// Adds node v's terms to acc: its points one by one if v is a leaf (in point
// order, as the reference sums them), else its moment.
template <int D, int K, bool EXACT>
__device__ __forceinline__ void node_accumulate(const TreeView& t, uint32_t v, float x, float y,
                                                float z, float& acc, int32_t& cnt,
                                                int32_t& cnt_nodes) {
  const uint4 l = t.nodeL[v];
  const bool leaf = (l.w & kLeafBit) != 0;
  uint32_t np = 0;
  if (leaf) {
    const uint32_t p0 = l.y, p1 = l.y + (l.w & ~kLeafBit);
    for (uint32_t j = p0; j < p1; ++j)
      acc = accum<D, K, EXACT>(acc, t.pts[j], t.nrm[j], t.alpha, t.alpha_k, x, y, z);
    np = p1 - p0;
  } else {
    acc = accum<D, K, EXACT>(acc, t.nodeC[v], t.nodeN[v], t.alpha, t.alpha_k, x, y, z);
  }
  // value selects, not a selected reference: both counters stay in registers
  cnt += (int32_t)np;
  cnt_nodes += leaf ? 0 : 1;
}

// This wavefront's 64 LDS slots for staging the children of a path node.
struct SamplerLds {
  float4* src;  // term source: node moment, or the point of a one-point leaf
  float4* nrm;  // its normal (winding number only)
  uint4* rec;   // the child's walk record (nodeL)
  uint2* aux;   // {end of the child's single-child chain, the child's point count}
};

// Weight of one level: nS / (P_RR * n_prev) * diff, with the reference's
// operation order; the fast mode uses the hardware reciprocal for the divide.
template <bool EXACT>
__device__ __forceinline__ float level_weight(float nS, float event_prob, float n_prev,
                                              float diff) {
#pragma clang fp contract(off)
  const float den = event_prob * n_prev;
  if constexpr (EXACT)
    return nS / den * diff;
  else
    return nS * __builtin_amdgcn_rcpf(den) * diff;
}

// g: wavefront index within the request (the path key).
template <int D, int K, bool EXACT>
__device__ __forceinline__ LaneResult stochastic_estimate(const TreeView& t, int spl,
                                                          uint32_t key0, uint32_t key1,
                                                          uint32_t g, bool live, float x, float y,
                                                          float z, SamplerLds lds,
                                                          bool count = true) {
#pragma clang fp contract(off)
  constexpr bool kWN = (K == WINDING_NUMBER);
  const uint32_t lane = threadIdx.x & 63;
  const float4* __restrict__ nodeC = t.nodeC;
  const uint4* __restrict__ nodeL = t.nodeL;
  const float4* __restrict__ nodeN = t.nodeN;
  const uint32_t* __restrict__ n_start = t.n_start;
  const uint32_t* __restrict__ n_end = t.n_end;
  const uint32_t* __restrict__ n_skip = t.n_skip;
  const float4* __restrict__ pts = t.pts;
  const float4* __restrict__ nrm = t.nrm;
  const float alpha = t.alpha, alpha_k = t.alpha_k;
  int32_t cnt = 0, cnt_nodes = 0;
  const uint4 rl = nodeL[0];
  float result;
  if (rl.w & kLeafBit) {
    result = 0.0f;
    node_accumulate<D, K, EXACT>(t, 0, x, y, z, result, cnt, cnt_nodes);
  } else {
    float bh = 0.0f, mlcv = 0.0f;
    const uint32_t nsub = rl.w, first = rl.y;
    for (uint32_t si = 0; si < nsub; ++si) {
      const uint32_t S = uniform_u32(first + si);
      node_accumulate<D, K, EXACT>(t, S, x, y, z, bh, cnt, cnt_nodes);
      const uint4 Sl = nodeL[S];
      const uint32_t sstart = n_start[S];
      const uint32_t nS = n_end[S] - sstart;
      float sub_total = 0.0f;
      if (!(Sl.w & kLeafBit)) {
        const float4 SC = nodeC[S];
        const float4 SN = kWN ? nodeN[S] : SC;
        const float cS = path_ratio<D, EXACT>(SC, __uint_as_float(Sl.x), x, y, z);
        // Level 0 is the same for every sample: at the subdomain the ratio
        // equals cur_beta (or is below 1), so the roulette keeps every path
        // (tp = 1, P_RR stays 1) and the level adds nS / (1 * nS) * (sum of
        // S's children - f(S)) whatever point was drawn.  It is evaluated
        // once here; each sample then starts from the child holding its point
        // with that weight as the first term of its sum, as it would have
        // computed it.  (Subdomains with more than 64 children take the
        // general loop.)
        const uint32_t f0 = uniform_u32(Sl.y), nch0 = uniform_u32(Sl.w);
        const bool cached0 = nch0 <= 64;
        float wd0 = 0.0f;
        int32_t cnt0 = 0, cnt0_nodes = 0;
        uint32_t c0_start = 0xFFFFFFFFu;  // lane j: start of S's child j
        if (cached0) {
          uint32_t np_lane = 0;
          if (lane < nch0) {
            const uint32_t c = f0 + lane;
            const uint4 l = nodeL[c];
            c0_start = n_start[c];
            float4 src, nv = make_float4(0.f, 0.f, 0.f, 0.f);
            if (l.w & kLeafBit) {
              np_lane = l.w & ~kLeafBit;
              src = pts[l.y];
              if (kWN) nv = nrm[l.y];
            } else {
              src = nodeC[c];
              if (kWN) nv = nodeN[c];
            }
            lds.src[lane] = src;
            if (kWN) lds.nrm[lane] = nv;
            lds.rec[lane] = l;
          }
          const uint64_t multi = __ballot(np_lane > 1);
          wave_lds_sync();
          float r0sum = 0.0f;
          if (live) {
            for (uint32_t j = 0; j < nch0; ++j) {
              const uint4 l = lds.rec[j];
              const uint32_t np = l.w & ~kLeafBit;
              if (multi != 0 && (l.w & kLeafBit) && np > 1) {
                for (uint32_t pp = l.y; pp < l.y + np; ++pp)
                  r0sum = accum<D, K, EXACT>(r0sum, pts[pp], nrm[pp], alpha, alpha_k, x, y, z);
              } else {
                const float4 src = lds.src[j];
                const float4 nv = kWN ? lds.nrm[j] : src;
                r0sum = accum<D, K, EXACT>(r0sum, src, nv, alpha, alpha_k, x, y, z);
              }
            }
          }
          if (count) {
            uint32_t pt = np_lane, nd = (lane < nch0 && np_lane == 0) ? 1u : 0u;
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) {
              pt += __shfl_xor(pt, off, 64);
              nd += __shfl_xor(nd, off, 64);
            }
            cnt0 = (int32_t)pt;
            cnt0_nodes = (int32_t)nd;
          }
          wave_lds_sync();
          const float prev0 = term<D, K, EXACT>(SC, SN, alpha, alpha_k, x, y, z);
          wd0 = level_weight<EXACT>((float)nS, 1.0f, (float)nS, r0sum - prev0);
        }
        for (int ss = 0; ss < spl; ++ss) {
          Philox4 r0 = philox4x32_10((uint32_t)ss, si, g, 0u, key0, key1);
          const uint32_t idx = uniform_u32(sstart + scale_draw(r0.x, r0.y, nS));
          Philox4 blk = r0;
          uint32_t blk_id = 0;
          float event_prob = 1.0f;
          float wd_all = 0.0f;
          float cur_beta = fmaxf(1.0f, cS);
          bool alive = live;
          uint32_t level = 0;
          // the path node: its walk record, moment, normal and point count
          uint4 curL = Sl;
          float4 curC = SC, curN = SN;
          uint32_t cur_cnt = nS;
          if (cached0) {
            // level 0 from the cache: descend to the child holding the point
            const uint32_t below =
                (uint32_t)__popcll(__ballot(lane < nch0 && c0_start <= idx));
            const uint32_t chosen = uniform_u32(f0 + below - 1);
            const uint32_t sk = n_skip[chosen];
            cur_cnt = n_end[chosen] - n_start[chosen];
            curL = nodeL[sk];
            curC = nodeC[sk];
            curN = kWN ? nodeN[sk] : curC;
            cur_beta = fmaxf(cur_beta, cS);
            if (live) {
              wd_all += wd0;
              cnt += cnt0;
              cnt_nodes += cnt0_nodes;
            }
            level = 1;
          }
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
            const float ffr = path_ratio<D, EXACT>(curC, __uint_as_float(curL.x), x, y, z);
            const float tp = traverse_prob<EXACT>(cur_beta, ffr);
            if (alive) {
              if (p > tp) {
                event_prob *= 1.0f - tp;
                alive = false;
              } else {
                event_prob *= tp;
              }
            }
            if (__ballot(alive) == 0) break;
            const uint32_t f = uniform_u32(curL.y), nch = uniform_u32(curL.w);
            float cur_result = 0.0f;
            // record of the child holding the sample (captured from LDS)
            uint4 nextL = curL;
            float4 nextC = curC, nextN = curN;
            uint2 nextAux = make_uint2(0u, 0u);
            uint32_t chosen = 0;
            for (uint32_t b = 0; b < nch; b += 64) {
              // stage children b .. b+63 of the path node, one per lane
              const uint32_t cj = b + lane;
              const bool valid = cj < nch;
              uint32_t cstart = 0xFFFFFFFFu;
              uint32_t np_lane = 0;
              if (valid) {
                const uint32_t c = f + cj;
                const uint4 l = nodeL[c];
                cstart = n_start[c];
                const uint32_t cend = n_end[c];
                const uint32_t sk = n_skip[c];
                float4 src, nv = make_float4(0.f, 0.f, 0.f, 0.f);
                if (l.w & kLeafBit) {
                  np_lane = l.w & ~kLeafBit;
                  src = pts[l.y];
                  if (kWN) nv = nrm[l.y];
                } else {
                  src = nodeC[c];
                  if (kWN) nv = nodeN[c];
                }
                lds.src[lane] = src;
                if (kWN) lds.nrm[lane] = nv;
                lds.rec[lane] = l;
                lds.aux[lane] = make_uint2(sk, cend - cstart);
              }
              // children are in point order: those starting at or before the
              // sample form a prefix, the last of which holds it
              const uint32_t below = (uint32_t)__popcll(__ballot(valid && cstart <= idx));
              const uint64_t multi = __ballot(np_lane > 1);
              const uint32_t m = min(64u, nch - b);
              wave_lds_sync();
              if (alive) {
                if (multi == 0) {
                  // every child is one source: a node moment or a one-point leaf
#pragma unroll 4
                  for (uint32_t j = 0; j < m; ++j) {
                    const float4 src = lds.src[j];
                    const float4 nv = kWN ? lds.nrm[j] : src;
                    cur_result = accum<D, K, EXACT>(cur_result, src, nv, alpha, alpha_k, x, y, z);
                  }
                } else {
                  for (uint32_t j = 0; j < m; ++j) {
                    const uint4 l = lds.rec[j];
                    const uint32_t np = l.w & ~kLeafBit;
                    if ((l.w & kLeafBit) && np > 1) {
                      for (uint32_t pp = l.y; pp < l.y + np; ++pp)
                        cur_result =
                            accum<D, K, EXACT>(cur_result, pts[pp], nrm[pp], alpha, alpha_k, x, y, z);
                    } else {
                      const float4 src = lds.src[j];
                      const float4 nv = kWN ? lds.nrm[j] : src;
                      cur_result = accum<D, K, EXACT>(cur_result, src, nv, alpha, alpha_k, x, y, z);
                    }
                  }
                }
              }
              if (below > 0) {
                nextL = lds.rec[below - 1];
                nextC = lds.src[below - 1];
                if (kWN) nextN = lds.nrm[below - 1];
                nextAux = lds.aux[below - 1];
                chosen = f + b + below - 1;
              }
              // term counts of this batch, the same for every lane that evaluated it
              if (count) {
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
            // descend to the child holding the sample, then down any
            // single-child chain (the same points, so the same count)
            const float4 prevC = curC, prevN = curN;
            const uint32_t prev_cnt = cur_cnt;
            const uint32_t sk = uniform_u32(nextAux.x);
            cur_cnt = uniform_u32(nextAux.y);
            if (sk == uniform_u32(chosen)) {
              curL = make_uint4(uniform_u32(nextL.x), uniform_u32(nextL.y), nextL.z,
                                uniform_u32(nextL.w));
              curC = nextC;
              curN = nextN;
            } else {
              curL = nodeL[sk];
              curC = nodeC[sk];
              curN = kWN ? nodeN[sk] : curC;
            }
            cur_beta = fmaxf(cur_beta, ffr);
            if (alive) {
              const float prev_result = term<D, K, EXACT>(prevC, prevN, alpha, alpha_k, x, y, z);
              const float diff = cur_result - prev_result;
              wd_all += level_weight<EXACT>((float)nS, event_prob, (float)prev_cnt, diff);
            }
          }
          sub_total += wd_all;
        }
      }
      mlcv += sub_total / (float)spl;
    }
    result = bh + mlcv;
  }
  return {result, cnt, cnt_nodes};
}

// ---------------------------------------------------------------------------
// Fast-mode estimator for trees whose nodes have at most 64 children
// (octrees k <= 2, quadtrees k <= 3): the same samples, the same per-sample
// paths, roulette draws and weights as stochastic_estimate<D, K, false>,
// evaluated with reuse across samples.
//
// A path node v contributes diff(v) = (sum of v's children's terms) - f(v)
// weighted by the path's roulette probability; diff(v) and v's opening ratio
// depend only on v and the query, not on the path that reached v.  The
// samples of a subdomain are therefore drawn a chunk at a time, sorted by
// their sampled point (so consecutive samples share the upper part of their
// paths), and diff of the most recent node at each of the top kCacheLevels
// levels is kept in LDS, keyed by node.  A sample whose path passes a cached
// node reads it instead of evaluating the node's children (its opening
// ratio is recomputed, which keeps the cache small enough for seven levels at
// seven wavefronts per SIMD).  Each cached node also records the step the
// latest sample took from it (the chosen child's point range and the node
// stepped to), so a sample whose point falls in the same child steps
// without reading the tree.  Only the order in which the samples'
// contributions are summed differs from the per-sample loop.
// ---------------------------------------------------------------------------
constexpr uint32_t kChunk = 512;        // samples drawn and sorted together (at most)
// A 512-sample chunk's keys pack {point << 9 | sample, roulette word}, which
// needs point indices below 2^23; in larger trees chunks hold 256 samples and
// the sample index rides in the roulette word's low 8 bits (which the
// roulette does not use).  The winding-number kernels stage normals too;
// half-size chunks keep their LDS within seven workgroups per CU.
constexpr uint32_t kPackedPointLimit = 1u << 23;
template <int K>
constexpr uint32_t chunk_for() {
  return K == WINDING_NUMBER ? kChunk / 2 : kChunk;
}
constexpr uint32_t kCacheLevels = 7;    // path levels 1..7 cached
constexpr uint32_t kNoNode = 0xFFFFFFFFu;
// Lanes of the estimator's cache register per field (kCacheLevels + 1 step
// records): node ids, point and node term counts, step ranges' ends.
constexpr uint32_t kCacheF = kCacheLevels + 1;
constexpr uint32_t kLanePt = kCacheF, kLaneNd = 2 * kCacheF;
constexpr uint32_t kLaneLo = 3 * kCacheF, kLaneHi = 4 * kCacheF;
static_assert(5 * kCacheF <= 64, "the cache register's fields fit in 64 lanes");

// This wavefront's LDS for the fast estimator.
struct SortedLds {
  float4* src;        // 64 staging slots: term sources
  float4* nrm;        // 64 staging slots: normals (winding number only)
  uint2* keys;        // kChunk samples: {sampled point << 9 | sample, level-1 roulette
                      // word}, or {sampled point, roulette word with the sample in its
                      // low 8 bits}
  uint4* nrec;        // (kCacheLevels + 1) x 3: step records (node and walk record,
                      // centroid, normal)
};

// Sorts keys[0, n) by .x, n a power of two <= kChunk (bitonic network).
// Each lane takes up to kChunk / 128 of a stage's n / 2 compare-exchanges,
// addressed by pair: pair p joins i = p with a 0 inserted at bit log2(j) and
// i + j, and both slots are written back (no per-lane branch on the order).
template <uint32_t CHUNK>
__device__ __forceinline__ void wave_sort_keys(uint2* keys, uint32_t n) {
  const uint32_t lane = threadIdx.x & 63;
  for (uint32_t k = 2; k <= n; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
#pragma unroll
      for (uint32_t m = 0; m < CHUNK / 128; ++m) {
        const uint32_t p = lane + 64 * m;
        if (p < n / 2) {
          const uint32_t i = (p & (j - 1)) | ((p & ~(j - 1)) << 1);
          const uint2 a = keys[i], b = keys[i + j];
          const bool swap = (a.x > b.x) == ((i & k) == 0);
          keys[i] = swap ? b : a;
          keys[i + j] = swap ? a : b;
        }
      }
      wave_lds_sync();
    }
  }
}

// COUNT: also count the terms each query's estimate sums (the counting
// mode); without it the counts are not formed at all.
// RANGE: only subdomains [si0, si1) of the root; the Barnes-Hut part of the
// sum goes to *bh_part and the control-variate part is returned (the caller
// adds the parts of all ranges in subdomain order).
// WORK (with COUNT): count instead the work this kernel performs for the
// query after its reuse: .points = kernel terms evaluated (a cached node's
// children and the subdomain's level 0 count once, not once per sample),
// .nodes = sample-path levels stepped (opening ratio, roulette and weight).
template <int D, int K, bool COUNT, bool RANGE = false, bool WORK = false>
__device__ __forceinline__ LaneResult stochastic_estimate_sorted(const TreeView& t, int spl,
                                                                 uint32_t key0, uint32_t key1,
                                                                 uint32_t g, bool live, float x,
                                                                 float y, float z, SortedLds lds,
                                                                 uint32_t si0 = 0,
                                                                 uint32_t si1 = 0,
                                                                 float* bh_part = nullptr) {
#pragma clang fp contract(off)
  constexpr bool EXACT = false;
  constexpr bool kWN = (K == WINDING_NUMBER);
  constexpr uint32_t kCh = chunk_for<K>();
  static_assert(!WORK || COUNT, "work counts are a counting mode");
  constexpr bool REF = COUNT && !WORK;  // the reference's term counts
  int32_t wterms = 0, wlevels = 0;      // work counts
  const uint32_t lane = threadIdx.x & 63;
  const LaneMask live_m = __ballot(live);
  const float4* __restrict__ nodeC = t.nodeC;
  const uint4* __restrict__ nodeL = t.nodeL;
  const float4* __restrict__ nodeN = t.nodeN;
  const uint32_t* __restrict__ n_start = t.n_start;
  const uint32_t* __restrict__ n_end = t.n_end;
  const float4* __restrict__ pts = t.pts;
  const float4* __restrict__ nrm = t.nrm;
  // read-only for the whole launch: wave-uniform reads become scalar loads
  const uint4* __restrict__ stepR = t.stepR;
  const WalkRecord* __restrict__ bfsR = t.bfsR;
  const float4* __restrict__ kidC = t.kidC;
  const float4* __restrict__ kidN = t.kidN;
  const float alpha = t.alpha, alpha_k = t.alpha_k;
  int32_t cnt = 0, cnt_nodes = 0;
  const uint4 rl = nodeL[0];
  if (rl.w & kLeafBit) {
    float v = 0.0f;
    node_accumulate<D, K, EXACT>(t, 0, x, y, z, v, cnt, cnt_nodes);
    if constexpr (RANGE) {
      *bh_part = v;
      return {0.0f, cnt, cnt_nodes};
    }
    return {v, cnt, cnt_nodes};
  }
  // the cached nodes' identities and term counts live in one register:
  // lane l < kCacheLevels holds level l + 1's node, lane kCacheF + l its
  // point terms and lane 2 kCacheF + l its node terms (read back with
  // readlane); the step records' ranges follow (lanes 3 kCacheF + r, 4 kCacheF + r)
  uint32_t cinfo = lane < kCacheLevels ? kNoNode : 0u;
  // per cached path level, per lane: the cached node's opening probability,
  // its weight (the level's term of a sample through it) and the running
  // beta after it; registers, indexed only by compile-time levels
  float c_tp[kCacheLevels], c_wd[kCacheLevels], c_beta[kCacheLevels];

  // Evaluates the children of v (at most 64) for every live lane: returns
  // their sum; *start = this lane's child's start (~0 past the last child);
  // *pt, *nd = the batch's point and node term counts (if counted).
  auto children_sum = [&](uint32_t f, uint32_t nch, uint32_t* start, uint32_t* pt,
                          uint32_t* nd) -> float {
    uint32_t cstart = 0xFFFFFFFFu, np_lane = 0;
    if (lane < nch) {
      // independent loads: the child's record, range start and term source
      // (a one-point leaf's point; a leaf of several points is summed below)
      const uint32_t c = f + lane;
      const uint4 l = nodeL[c];
      cstart = n_start[c];
      lds.src[lane] = kidC[c];
      if (kWN) lds.nrm[lane] = kidN[c];
      if (l.w & kLeafBit) np_lane = l.w & ~kLeafBit;
    }
    const uint64_t multi = __ballot(np_lane > 1);
    wave_lds_sync();
    // every lane sums (a lane past the last query holds query 0's position
    // and its sums are dropped), so the loop needs no lane mask
    float r = 0.0f;
    {
      if (multi == 0) {
#pragma unroll 4
        for (uint32_t j = 0; j < nch; ++j) {
          const float4 src = lds.src[j];
          const float4 nv = kWN ? lds.nrm[j] : src;
          r = accum<D, K, EXACT>(r, src, nv, alpha, alpha_k, x, y, z);
        }
      } else {
        // some child is a leaf of several points (rare: coincident points
        // or leaf_size > 1): its record is read again, as a scalar load
        for (uint32_t j = 0; j < nch; ++j) {
          const uint4 l = load_invariant(nodeL, f + j);
          const uint32_t np = l.w & ~kLeafBit;
          if ((l.w & kLeafBit) && np > 1) {
            for (uint32_t pp = l.y; pp < l.y + np; ++pp)
              r = accum<D, K, EXACT>(r, pts[pp], nrm[pp], alpha, alpha_k, x, y, z);
          } else {
            const float4 src = lds.src[j];
            const float4 nv = kWN ? lds.nrm[j] : src;
            r = accum<D, K, EXACT>(r, src, nv, alpha, alpha_k, x, y, z);
          }
        }
      }
    }
    *pt = 0;
    *nd = 0;
    if constexpr (COUNT) {
      uint32_t a = np_lane, b = (lane < nch && np_lane == 0) ? 1u : 0u;
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
      }
      *pt = a;
      *nd = b;
    }
    wave_lds_sync();
    *start = cstart;
    return r;
  };

  // The step taken from each cached node by the latest sample through it
  // (record r = 1..kCacheLevels for path level r, record 0 for the
  // subdomain's root): the chosen child's point range [lo, hi) in lanes
  // kLaneLo + r and kLaneHi + r of `cinfo`, and in LDS the node stepped to (single-child
  // chains skipped) with its walk record, centroid and normal.  A sorted
  // sample whose point lies in the same child takes the step without a load.
  constexpr uint32_t kNoRecord = 0xFFu;
  // The level-1 node of record 0 (the node stepped to from the subdomain's
  // root), once a sample has reached it: a leaf, or an internal node whose
  // traversal probability (a function of the node and the query) is kept per
  // lane in l1_tp.  A later sample through record 0 then needs only its
  // level-1 roulette draw to know whether any lane goes on.
  constexpr uint32_t kL1Unknown = 0, kL1Internal = 1, kL1Leaf = 2;
  uint32_t l1_state = kL1Unknown;
  float l1_tp = 0.0f;
  auto range_has = [&](uint32_t r, uint32_t idx) {
    return idx >= (uint32_t)__builtin_amdgcn_readlane((int)cinfo, (int)(kLaneLo + r)) &&
           idx < (uint32_t)__builtin_amdgcn_readlane((int)cinfo, (int)(kLaneHi + r));
  };
  auto take = [&](uint32_t r, uint32_t& cur, uint32_t& cur_cnt, uint4& curL, float4& curC,
                  float4& curN) {
    const uint4 a = lds.nrec[3 * r], c = lds.nrec[3 * r + 1];
    cur = uniform_u32(a.x);
    cur_cnt = (uint32_t)__builtin_amdgcn_readlane((int)cinfo, (int)(kLaneHi + r)) -
              (uint32_t)__builtin_amdgcn_readlane((int)cinfo, (int)(kLaneLo + r));
    curL = make_uint4(a.y, uniform_u32(a.z), 0u, uniform_u32(a.w));
    curC = make_float4(__uint_as_float(c.x), __uint_as_float(c.y), __uint_as_float(c.z),
                       __uint_as_float(c.w));
    if (kWN) {
      const uint4 n = lds.nrec[3 * r + 2];
      curN = make_float4(__uint_as_float(n.x), __uint_as_float(n.y), __uint_as_float(n.z),
                         __uint_as_float(n.w));
    } else {
      curN = curC;
    }
  };
  auto step = [&](uint32_t chosen, uint32_t r, uint32_t& cur, uint32_t& cur_cnt, uint4& curL,
                  float4& curC, float4& curN) {
    const uint4 sr = load_invariant(stepR, chosen);
    const uint32_t lo = sr.x, hi = sr.y;
    cur = sr.z;
    cur_cnt = hi - lo;
    const WalkRecord cr = load_invariant(bfsR, cur);
    curL = cr.l;
    curC = cr.c;
    curN = kWN ? nodeN[cur] : curC;
    if (r != kNoRecord) {
      if (r == 0) l1_state = kL1Unknown;  // record 0 now names another node
      cinfo = lane == kLaneLo + r ? lo : lane == kLaneHi + r ? hi : cinfo;
      if (lane == 0) {
        lds.nrec[3 * r] = make_uint4(cur, curL.x, curL.y, curL.w);
        lds.nrec[3 * r + 1] = make_uint4(__float_as_uint(curC.x), __float_as_uint(curC.y),
                                         __float_as_uint(curC.z), __float_as_uint(curC.w));
        if (kWN)
          lds.nrec[3 * r + 2] = make_uint4(__float_as_uint(curN.x), __float_as_uint(curN.y),
                                           __float_as_uint(curN.z), __float_as_uint(curN.w));
      }
      wave_lds_sync();
    }
  };

  float bh = 0.0f, mlcv = 0.0f;
  const uint32_t nsub = rl.w, first = rl.y;
  const uint32_t si_begin = RANGE ? si0 : 0u;
  const uint32_t si_end = RANGE ? (si1 < nsub ? si1 : nsub) : nsub;
  for (uint32_t si = si_begin; si < si_end; ++si) {
    const uint32_t S = uniform_u32(first + si);
    node_accumulate<D, K, EXACT>(t, S, x, y, z, bh, cnt, cnt_nodes);
    const uint4 Sl = nodeL[S];
    if (Sl.w & kLeafBit) continue;  // the subdomain's points were summed exactly
    cinfo = lane == kLaneHi ? 0u : cinfo;  // record 0 belongs to the previous subdomain: empty it
    l1_state = kL1Unknown;
    const uint32_t sstart = n_start[S];
    const uint32_t nS = n_end[S] - sstart;
    const float4 SC = nodeC[S];
    const float4 SN = kWN ? nodeN[S] : SC;
    const float cS = path_ratio<D, EXACT>(SC, __uint_as_float(Sl.x), x, y, z);
    // level 0, the same for every sample (see stochastic_estimate)
    const uint32_t f0 = uniform_u32(Sl.y), nch0 = uniform_u32(Sl.w);
    uint32_t c0_start, pt0, nd0;
    const float r0sum = children_sum(f0, nch0, &c0_start, &pt0, &nd0);
    const LaneMask kids0 = __ballot(lane < nch0);
    const float beta0 = fmaxf(fmaxf(1.0f, cS), cS);  // the roulette's starting beta
    const float wd0 = level_weight<EXACT>((float)nS, 1.0f, (float)nS,
                                          r0sum - term<D, K, EXACT>(SC, SN, alpha, alpha_k, x, y, z));
    if constexpr (WORK) wterms += live ? (int32_t)(pt0 + nd0 + 1) : 0;
    float sub_total = 0.0f;
    // When every child of S steps (past its single-child chain) to a leaf,
    // every sample's path ends at level 1 and adds the level-0 weight alone,
    // whichever point it draws: the samples need not be drawn.  (Small
    // outlying subdomains; the sum is the per-sample loop's, term for term.)
    {
      uint32_t inner = 0;
      if (lane < nch0) inner = (bfsR[stepR[f0 + lane].z].l.w & kLeafBit) ? 0u : 1u;
      if (__ballot(inner) == 0) {
        const float w = live ? 0.0f + wd0 : 0.0f;
        for (int s = 0; s < spl; ++s) sub_total += w;
        if constexpr (REF) {
          cnt += live ? spl * (int32_t)pt0 : 0;
          cnt_nodes += live ? spl * (int32_t)nd0 : 0;
        }
        mlcv += sub_total / (float)spl;
        continue;
      }
    }
    // the chunk size and key format of this subdomain (its point indices
    // must fit 23 bits for packed keys)
    const bool packed = kCh > 256 && sstart + nS <= kPackedPointLimit;
    const uint32_t ch_size = packed ? kCh : (kCh < 256 ? kCh : 256u);
    for (int s0 = 0; s0 < spl; s0 += (int)ch_size) {
      const uint32_t ncur = (uint32_t)min((int)ch_size, spl - s0);
      // draw the chunk's sampled points (Philox block 0 of each sample)
#pragma unroll
      for (uint32_t m = 0; m < kCh / 64; ++m) {
        const uint32_t i = lane + 64 * m;
        uint2 key = make_uint2(0xFFFFFFFFu, 0u);
        if (i < ncur) {
          const uint32_t ss = (uint32_t)s0 + i;
          const Philox4 r = philox4x32_10(ss, si, g, 0u, key0, key1);
          const uint32_t pt = sstart + scale_draw(r.x, r.y, nS);
          key = packed ? make_uint2((pt << 9) | i, r.w) : make_uint2(pt, (r.w & 0xFFFFFF00u) | i);
        }
        if (i < ch_size) lds.keys[i] = key;
      }
      wave_lds_sync();
      // sort the drawn samples (and sentinels up to the next power of two)
      uint32_t nsort = 1;
      while (nsort < ncur) nsort <<= 1;
      wave_sort_keys<kCh>(lds.keys, nsort);
      for (uint32_t i = 0; i < ncur; ++i) {
        const uint2 key = lds.keys[i];
        const uint32_t kx = uniform_u32(key.x);
        const uint32_t w1 = uniform_u32(key.y);
        const uint32_t idx = packed ? kx >> 9 : kx;
        const uint32_t ss = (uint32_t)s0 + (packed ? kx & 511u : w1 & 0xFFu);
        if (l1_state != kL1Unknown && range_has(0, idx)) {
          // the sample's level-1 node is record 0's: when no lane goes on
          // from it, the sample adds the level-0 weight alone, as the full
          // step below would
          const LaneMask on =
              l1_state == kL1Internal ? live_m & __ballot(!(rr_uniform(w1) > l1_tp)) : 0ull;
          if (on == 0) {
            if constexpr (WORK) wlevels += (live && l1_state == kL1Internal) ? 1 : 0;
            if constexpr (REF) {
              cnt += live ? (int32_t)pt0 : 0;
              cnt_nodes += live ? (int32_t)nd0 : 0;
            }
            sub_total += live ? 0.0f + wd0 : 0.0f;
            continue;
          }
        }
        uint32_t cur, cur_cnt;
        uint4 curL;
        float4 curC, curN;
        // level 0 from the subdomain's cache
        if (range_has(0, idx)) {
          take(0, cur, cur_cnt, curL, curC, curN);
        } else {
          const uint32_t below0 =
              (uint32_t)__popcll(kids0 & __ballot(c0_start <= idx));
          step(uniform_u32(f0 + below0 - 1), 0, cur, cur_cnt, curL, curC, curN);
        }
        // epc: the product of the opening probabilities along the path (the
        // event probability of a lane still on it)
        float epc = 1.0f;
        float wd_all = live ? 0.0f + wd0 : 0.0f;
        float cur_beta = beta0;
        LaneMask alive = live_m;
        if constexpr (REF) {
          cnt += live ? (int32_t)pt0 : 0;
          cnt_nodes += live ? (int32_t)nd0 : 0;
        }
        Philox4 blk{0u, 0u, 0u, w1};  // block 0: only its level-1 word is used here
        uint32_t blk_id = 0;
        uint32_t level = 1;
        if (curL.w & kLeafBit) l1_state = kL1Leaf;
        // One path level.  SLOT >= 0: path level SLOT + 1, whose node may be
        // the cached one (its opening probability, weight and running beta
        // are then read back instead of computed: they depend only on the
        // node and the query); SLOT < 0: an uncached deeper level.  Returns
        // false when the sample's path ends.
        auto path_level = [&]<int SLOT>() -> bool {
          if (curL.w & kLeafBit) return false;
          const uint32_t wi = level + 2;
          if ((wi >> 2) != blk_id) {
            blk_id = wi >> 2;
            blk = philox4x32_10(ss, si, g, blk_id, key0, key1);
          }
          const uint32_t wsel = wi & 3;
          const uint32_t word = wsel == 0 ? blk.x : wsel == 1 ? blk.y : wsel == 2 ? blk.z : blk.w;
          const float p = rr_uniform(word);
          bool hit = false;
          if constexpr (SLOT >= 0)
            hit = (uint32_t)__builtin_amdgcn_readlane((int)cinfo, SLOT) == cur;
          ++level;
          const uint32_t f = uniform_u32(curL.y), nch = uniform_u32(curL.w);
          uint32_t below = 0, pt = 0, nd = 0;
          bool known = false;  // the next node is the one recorded for this cached node
          if (hit) {
            if constexpr (SLOT >= 0) {
              const float tp = c_tp[SLOT];
              const LaneMask stop = __ballot(p > tp);
              if constexpr (WORK) wlevels += lane_in(alive) ? 1 : 0;
              alive &= ~stop;
              if (alive == 0) return false;
              epc = epc * tp;
              cur_beta = c_beta[SLOT];
              wd_all = lane_in(alive) ? wd_all + c_wd[SLOT] : wd_all;
              if constexpr (REF) {
                pt = (uint32_t)__builtin_amdgcn_readlane((int)cinfo, (int)(kLanePt + SLOT));
                nd = (uint32_t)__builtin_amdgcn_readlane((int)cinfo, (int)(kLaneNd + SLOT));
              }
              known = range_has(SLOT + 1, idx);
              if (!known) {
                const uint32_t cs = lane < nch ? n_start[f + lane] : 0xFFFFFFFFu;
                below = (uint32_t)__popcll(__ballot(cs <= idx));
              }
            }
          } else {
            // the breadth-first records hold the diameter's hardware reciprocal
            const float ffr = ratio_fast<D>(curC, __uint_as_float(curL.x), x, y, z);
            const float tp = traverse_prob<EXACT>(cur_beta, ffr);
            if (SLOT == 0) {
              l1_state = kL1Internal;
              l1_tp = tp;
            }
            const LaneMask stop = __ballot(p > tp);
            if constexpr (WORK) wlevels += lane_in(alive) ? 1 : 0;
            alive &= ~stop;
            if (alive == 0) return false;
            epc = epc * tp;
            uint32_t cs;
            const float r = children_sum(f, nch, &cs, &pt, &nd);
            const float diff = r - term<D, K, EXACT>(curC, curN, alpha, alpha_k, x, y, z);
            if constexpr (WORK) wterms += lane_in(alive) ? (int32_t)(pt + nd + 1) : 0;
            below = (uint32_t)__popcll(__ballot(cs <= idx));
            const float wd = level_weight<EXACT>((float)nS, epc, (float)cur_cnt, diff);
            wd_all = lane_in(alive) ? wd_all + wd : wd_all;
            cur_beta = ffr > cur_beta ? ffr : cur_beta;  // max, as a compare and select
            if constexpr (SLOT >= 0) {
              c_tp[SLOT] = tp;
              c_wd[SLOT] = wd;
              c_beta[SLOT] = cur_beta;
              if constexpr (REF)
                cinfo = lane == SLOT            ? cur
                        : lane == kLanePt + SLOT ? pt
                        : lane == kLaneNd + SLOT ? nd
                                                 : cinfo;
              else
                cinfo = lane == SLOT ? cur : cinfo;
            }
          }
          if constexpr (REF) {
            cnt += lane_in(alive) ? (int32_t)pt : 0;
            cnt_nodes += lane_in(alive) ? (int32_t)nd : 0;
          }
          if (known)
            take(SLOT + 1, cur, cur_cnt, curL, curC, curN);
          else
            step(uniform_u32(f + below - 1), SLOT >= 0 ? SLOT + 1 : kNoRecord, cur, cur_cnt,
                 curL, curC, curN);
          return true;
        };
        static_assert(kCacheLevels == 7, "the cached path levels are unrolled below");
        if (path_level.template operator()<0>() && path_level.template operator()<1>() &&
            path_level.template operator()<2>() && path_level.template operator()<3>() &&
            path_level.template operator()<4>() && path_level.template operator()<5>() &&
            path_level.template operator()<6>()) {
          while (path_level.template operator()<-1>()) {
          }
        }
        sub_total += wd_all;
      }
      wave_lds_sync();
    }
    mlcv += sub_total / (float)spl;
  }
  if constexpr (WORK) {
    // the subdomains' own terms (node_accumulate) were counted in cnt / cnt_nodes
    cnt = wterms + cnt + cnt_nodes;
    cnt_nodes = wlevels;
  }
  if constexpr (RANGE) {
    *bh_part = bh;
    return {mlcv, cnt, cnt_nodes};
  }
  return {bh + mlcv, cnt, cnt_nodes};
}

// ---------------------------------------------------------------------------
// The same estimator with one sample path per query, as the reference's CPU
// sampler draws them (include/tree_sampler.h:21-113: every query picks its
// own point and runs its own roulette).  The path of sample `ss` in subdomain
// `si` of query `qid` comes from Philox block {ss, si, qid, b}: the stream a
// 1-query group gets in the wavefront-shared estimator, so a query's result
// does not depend on which queries share its launch or in what order.  Lanes
// walk independently, reading node records and children straight from
// memory (neighbouring queries in Hilbert order touch the same records).
// ---------------------------------------------------------------------------
template <int D, int K, bool EXACT>
__device__ __forceinline__ LaneResult stochastic_estimate_query(const TreeView& t, int spl,
                                                                uint32_t key0, uint32_t key1,
                                                                uint32_t qid, float x, float y,
                                                                float z) {
#pragma clang fp contract(off)
  const float4* __restrict__ nodeC = t.nodeC;
  const uint4* __restrict__ nodeL = t.nodeL;
  const float4* __restrict__ nodeN = t.nodeN;
  const uint32_t* __restrict__ n_start = t.n_start;
  const uint32_t* __restrict__ n_end = t.n_end;
  const uint32_t* __restrict__ n_skip = t.n_skip;
  const float alpha = t.alpha, alpha_k = t.alpha_k;
  int32_t cnt = 0, cnt_nodes = 0;
  const uint4 rl = nodeL[0];
  if (rl.w & kLeafBit) {
    float v = 0.0f;
    node_accumulate<D, K, EXACT>(t, 0, x, y, z, v, cnt, cnt_nodes);
    return {v, cnt, cnt_nodes};
  }
  float bh = 0.0f, mlcv = 0.0f;
  const uint32_t nsub = rl.w, first = rl.y;
  for (uint32_t si = 0; si < nsub; ++si) {
    const uint32_t S = first + si;
    node_accumulate<D, K, EXACT>(t, S, x, y, z, bh, cnt, cnt_nodes);
    const uint4 Sl = nodeL[S];
    if (Sl.w & kLeafBit) continue;  // the subdomain's points were summed exactly
    const uint32_t sstart = n_start[S];
    const uint32_t nS = n_end[S] - sstart;
    const float4 SC = nodeC[S];
    const float cS = path_ratio<D, EXACT>(SC, __uint_as_float(Sl.x), x, y, z);
    // level 0, the same for every sample (see stochastic_estimate)
    const uint32_t f0 = Sl.y, nch0 = Sl.w;
    float r0sum = 0.0f;
    int32_t cnt0 = 0, cnt0_nodes = 0;
    for (uint32_t j = 0; j < nch0; ++j)
      node_accumulate<D, K, EXACT>(t, f0 + j, x, y, z, r0sum, cnt0, cnt0_nodes);
    const float wd0 = level_weight<EXACT>(
        (float)nS, 1.0f, (float)nS, r0sum - term<D, K, EXACT>(SC, nodeN[S], alpha, alpha_k, x, y, z));
    float sub_total = 0.0f;
    for (int ss = 0; ss < spl; ++ss) {
      const Philox4 r0 = philox4x32_10((uint32_t)ss, si, qid, 0u, key0, key1);
      const uint32_t idx = sstart + scale_draw(r0.x, r0.y, nS);
      Philox4 blk = r0;
      uint32_t blk_id = 0;
      float event_prob = 1.0f;
      float wd_all = 0.0f;
      float cur_beta = fmaxf(cS, fmaxf(1.0f, cS));
      uint32_t pick = 0;
      for (uint32_t j = 1; j < nch0; ++j) pick = n_start[f0 + j] <= idx ? j : pick;
      uint32_t cur = n_skip[f0 + pick];
      uint4 curL = nodeL[cur];
      float4 curC = nodeC[cur];
      wd_all += wd0;
      cnt += cnt0;
      cnt_nodes += cnt0_nodes;
      uint32_t level = 1;
      while (!(curL.w & kLeafBit)) {
        const uint32_t wi = level + 2;
        if ((wi >> 2) != blk_id) {
          blk_id = wi >> 2;
          blk = philox4x32_10((uint32_t)ss, si, qid, blk_id, key0, key1);
        }
        const uint32_t wsel = wi & 3;
        const uint32_t word = wsel == 0 ? blk.x : wsel == 1 ? blk.y : wsel == 2 ? blk.z : blk.w;
        const float p = rr_uniform(word);
        ++level;
        const float ffr = path_ratio<D, EXACT>(curC, __uint_as_float(curL.x), x, y, z);
        const float tp = traverse_prob<EXACT>(cur_beta, ffr);
        if (p > tp) {
          event_prob *= 1.0f - tp;
          break;
        }
        event_prob *= tp;
        const uint32_t prev = cur;
        const uint32_t f = curL.y, nch = curL.w;
        float cur_result = 0.0f;
        uint32_t pick = 0;
        for (uint32_t j = 0; j < nch; ++j) {
          const uint32_t c = f + j;
          node_accumulate<D, K, EXACT>(t, c, x, y, z, cur_result, cnt, cnt_nodes);
          pick = n_start[c] <= idx ? j : pick;  // children are in point order
        }
        cur = n_skip[f + pick];
        curL = nodeL[cur];
        curC = nodeC[cur];
        cur_beta = fmaxf(cur_beta, ffr);
        const float prev_result =
            term<D, K, EXACT>(nodeC[prev], nodeN[prev], alpha, alpha_k, x, y, z);
        const float diff = cur_result - prev_result;
        wd_all += level_weight<EXACT>((float)nS, event_prob,
                                      (float)(n_end[prev] - n_start[prev]), diff);
      }
      sub_total += wd_all;
    }
    mlcv += sub_total / (float)spl;
  }
  return {bh + mlcv, cnt, cnt_nodes};
}

}  // namespace sbh
