// CPU oracle for the Stochastic Barnes-Hut fast-summation path.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, in plain C++ on the host, the
// reference's tree build, kernel functions, brute force, deterministic
// Barnes-Hut and stochastic control-variate estimator so that the HIP product
// path can be checked against it.  Only tests/, the smoke check and the CPU
// baseline leg of bench.py load it; nothing in the product links or calls it.
//
// Build: see oracle/Makefile (g++ -O2 -ffp-contract=off, no fast-math, so every
// fp32 operation rounds exactly once, in the order written below).
//
// Parity pins (each one is exercised by tests/test_oracle_*.py):
//   * tree build, leaf order  -> dyadic-grid closed form (integer Morton order),
//                                invariants, brute-force recursion on tiny sets
//   * node moments            -> float64 numpy sums (closed form)
//   * kernel functions        -> float64 closed forms
//   * brute force             -> float64 all-pairs sums
//   * Barnes-Hut              -> beta=inf == brute force, tiny beta == root
//                                moment, vote group 1 vs per-query walk
//   * stochastic estimator    -> exact expectation by enumerating every
//                                (sample point, Russian-roulette) outcome on a
//                                tiny tree, and unbiasedness over many seeds
//   * Philox4x32-10           -> published known-answer vectors and PyTorch's
//                                independent PhiloxRNGEngine.h
// Parity unpinned: bitwise agreement with the reference's own fp32 outputs
// (Eigen 3.4 is absent, so the reference cannot be compiled here); expf ulps.

#include <algorithm>
#include <atomic>
#include <cassert>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <deque>
#include <thread>
#include <vector>

namespace {

constexpr float kPi = (float)M_PI;  // include/utils.h:41-45 (PI = M_PI, cast)

// ---------------------------------------------------------------------------
// Source record: include/leaf_data.h:8-16 (position, normal, mass, alpha).
// `orig` is the input row index; it rides along with the record so the final
// permutation (leaf order) can be reported.  It never enters a predicate.
// ---------------------------------------------------------------------------
template <int D>
struct Rec {
  float pos[D];
  float nrm[D];
  float mass;
  float alpha;
  int64_t orig;
};

template <int D>
struct Box {
  float lo[D];
  float hi[D];
};

// x86 cvttss2si semantics for float -> int (the reference's Eigen cast<int>
// at include/aabb.h:34-41 compiles to it): out-of-range and NaN give INT_MIN.
inline int trunc_x86(float v) {
  if (!(v >= -2147483648.0f && v < 2147483648.0f)) return INT_MIN;
  return (int)v;
}

// include/aabb.h:34-41: clamp(int((D*(p-min)) / (max-min)), 0, D-1) per axis.
template <int D>
inline void quantize(const Box<D>& b, const float* p, int dsplit, int* out) {
  for (int a = 0; a < D; ++a) {
    float num = (float)dsplit * (p[a] - b.lo[a]);
    float den = b.hi[a] - b.lo[a];
    int v = trunc_x86(num / den);
    v = std::min(v, dsplit - 1);
    v = std::max(v, 0);
    out[a] = v;
  }
}

// Eigen fixed-size reductions split in halves: size 3 sums a0 + (a1 + a2).
template <int D>
inline float sqnorm(const float* d);
template <>
inline float sqnorm<2>(const float* d) {
  return d[0] * d[0] + d[1] * d[1];
}
template <>
inline float sqnorm<3>(const float* d) {
  return d[0] * d[0] + (d[1] * d[1] + d[2] * d[2]);
}
template <int D>
inline float dot(const float* a, const float* b);
template <>
inline float dot<2>(const float* a, const float* b) {
  return a[0] * b[0] + a[1] * b[1];
}
template <>
inline float dot<3>(const float* a, const float* b) {
  return a[0] * b[0] + (a[1] * b[1] + a[2] * b[2]);
}

// ---------------------------------------------------------------------------
// Kernel functions: include/kernel.h:31-59.  A "source" is anything with
// position / normal / mass / alpha (a point record or a node moment).
// ---------------------------------------------------------------------------
template <int D>
struct Src {
  const float* pos;
  const float* nrm;
  float mass;
  float alpha;
};

template <int D>
using KernelFn = float (*)(const Src<D>&, const float* q);

template <int D>
float k_gravity(const Src<D>& s, const float* q) {  // kernel.h:31-35
  float d[D];
  for (int a = 0; a < D; ++a) d[a] = s.pos[a] - q[a];
  float dist = std::sqrt(sqnorm<D>(d));
  return -s.mass / (dist + 1e-1f);
}

template <int D>
float k_wn(const Src<D>& s, const float* q);
template <>
float k_wn<2>(const Src<2>& s, const float* q) {  // kernel.h:40-45
  float d[2] = {s.pos[0] - q[0], s.pos[1] - q[1]};
  float dist_sq = sqnorm<2>(d);
  float qdn = dot<2>(s.nrm, d);
  return s.mass * qdn / (2 * kPi * (dist_sq + 1e-5f));
}
template <>
float k_wn<3>(const Src<3>& s, const float* q) {  // kernel.h:47-53
  float d[3] = {s.pos[0] - q[0], s.pos[1] - q[1], s.pos[2] - q[2]};
  float dist = std::sqrt(sqnorm<3>(d));
  float qdn = dot<3>(s.nrm, d);
  return s.mass * qdn / (4 * kPi * (dist * dist * dist + 1e-5f));
}

template <int D>
float k_smooth(const Src<D>& s, const float* q) {  // kernel.h:55-59
  float d[D];
  for (int a = 0; a < D; ++a) d[a] = s.pos[a] - q[a];
  float dist = std::sqrt(sqnorm<D>(d));
  return std::exp(-s.alpha * dist);
}

// helmholtzKernelX / helmholtzKernelY (kernel.h:68-83): m cos(freq 2 pi d) /
// (d + 0.1) with freq 2, and m sin(freq 2 pi d) / (d + 0.1) with freq 1; the
// argument is ((freq * 2) * pi) * dist in fp32, cosf / sinf.
template <int D, int Y>
float k_helm(const Src<D>& s, const float* q) {
  float d[D];
  for (int a = 0; a < D; ++a) d[a] = s.pos[a] - q[a];
  float dist = std::sqrt(sqnorm<D>(d));
  const float freq = Y ? 1.0f : 2.0f;
  float arg = freq * 2 * kPi * dist;
  float c = Y ? std::sin(arg) : std::cos(arg);
  return s.mass * c / (dist + 1e-1f);
}

// Component `A` of gravityGrad (kernel.h:61-66): -m / (d + 0.1)^3 * (q - p).
// The reference's template is never instantiated (no plugin binds it, and as
// written it would mix a double pow with a float vector); the fp32 form here
// is t = d + 0.1, s = -m / ((t * t) * t), s * (q_a - p_a).
template <int D, int A>
float k_grad(const Src<D>& s, const float* q) {
  float d[D];
  for (int a = 0; a < D; ++a) d[a] = q[a] - s.pos[a];
  float dist = std::sqrt(sqnorm<D>(d));
  float t = dist + 1e-1f;
  float sc = -s.mass / (t * t * t);
  return sc * d[A];
}

// float64 restatement of the same formulas (constants are the fp32 literals
// promoted, so this is the exact value of the fp32 formula's mathematics).
template <int D>
double k64(int kernel, const Src<D>& s, const float* q) {
  double d[D];
  double dd = 0.0, dn = 0.0;
  for (int a = 0; a < D; ++a) {
    d[a] = (double)s.pos[a] - (double)q[a];
    dd += d[a] * d[a];
    dn += (double)s.nrm[a] * d[a];
  }
  double dist = std::sqrt(dd);
  const double pi = (double)kPi;
  switch (kernel) {
    case 0:
      return -(double)s.mass / (dist + (double)1e-1f);
    case 1:
      if (D == 2) return (double)s.mass * dn / (2.0 * pi * (dd + (double)1e-5f));
      return (double)s.mass * dn / (4.0 * pi * (dist * dd + (double)1e-5f));
    case 3:
      return (double)s.mass * std::cos(4.0 * pi * dist) / (dist + (double)1e-1f);
    case 4:
      return (double)s.mass * std::sin(2.0 * pi * dist) / (dist + (double)1e-1f);
    case 5:
    case 6:
    case 7: {
      const int a = kernel - 5;
      if (a >= D) return 0.0;
      const double t = dist + (double)1e-1f;
      return -(double)s.mass / (t * t * t) * -d[a];  // (q - p) = -(p - q)
    }
    default:
      return std::exp(-(double)s.alpha * dist);
  }
}

template <int D>
KernelFn<D> pick_kernel(int kernel) {
  switch (kernel) {
    case 0:
      return k_gravity<D>;
    case 1:
      return k_wn<D>;
    case 2:
      return k_smooth<D>;
    case 3:
      return k_helm<D, 0>;
    case 4:
      return k_helm<D, 1>;
    case 5:
      return k_grad<D, 0>;
    case 6:
      return k_grad<D, 1>;
    case 7:
      if constexpr (D == 3) return k_grad<D, 2>;
      return nullptr;
  }
  return nullptr;
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
// ---------------------------------------------------------------------------
struct U4 {
  uint32_t v[4];
};

inline U4 philox(U4 c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += W0;
      k1 += W1;
    }
    uint64_t p0 = (uint64_t)M0 * c.v[0];
    uint64_t p1 = (uint64_t)M1 * c.v[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    U4 n;
    n.v[0] = hi1 ^ c.v[1] ^ k0;
    n.v[1] = lo1;
    n.v[2] = hi0 ^ c.v[3] ^ k1;
    n.v[3] = lo0;
    c = n;
  }
  return c;
}

// Random stream of one sample path.  Path = (seed, group, subdomain ordinal,
// sample index).  Word k of the stream is word k%4 of Philox block
// {sample, subdomain, group, k/4}.  Words 0,1 form the 64-bit point draw; word
// 1+l is the Russian-roulette uniform at descent level l (l = 0 is the
// subdomain itself).
struct PathStream {
  uint32_t k0, k1;
  uint32_t c0, c1, c2;
  uint32_t blk = 0xFFFFFFFFu;
  U4 cur;
  uint32_t word(uint32_t k) {
    uint32_t b = k >> 2;
    if (b != blk) {
      U4 c = {{c0, c1, c2, b}};
      cur = philox(c, k0, k1);
      blk = b;
    }
    return cur.v[k & 3];
  }
  uint64_t point_draw() { return (uint64_t)word(0) | ((uint64_t)word(1) << 32); }
  float rr_uniform(uint32_t level) {  // in (0, 1]
    uint32_t w = word(level + 2);
    return (float)((w >> 8) + 1u) * 0x1p-24f;
  }
};

// Uniform index in [0, n) from a 64-bit draw (multiply-high, Lemire).
inline uint64_t scale_draw(uint64_t u, uint64_t n) {
  return (uint64_t)(((unsigned __int128)u * (unsigned __int128)n) >> 64);
}

// ---------------------------------------------------------------------------
// Tree: include/tree3d.h:238-371 and include/tree2d.h:237-365.
// ---------------------------------------------------------------------------
template <int D>
struct Node {
  int64_t start, end, parent;
  int64_t first_child;  // children are consecutive (tree3d.h:400-411)
  int nchild;
  int depth;
  bool leaf;
  Box<D> box;
  float diam;
  // moment (contrib_data): position, normal, mass, alpha
  float c[D];
  float n[D];
  float mass;
  float alpha;
  // float64 aggregation of the same moments (reported alongside)
  double c64[D];
  double n64[D];
  double mass64;
};

template <int D>
struct Tree {
  int dsplit;      // 1 << split_log
  int split_log;
  int leaf_size;
  int max_depth;
  std::vector<Rec<D>> recs;  // permuted in place by the build (leaf order)
  std::vector<Node<D>> nodes;
  int global_max_depth = 0;

  float ratio(const Node<D>& nd, const float* q) const {  // tree3d.h:134-140
    if (nd.leaf) return INFINITY;
    float d[D];
    for (int a = 0; a < D; ++a) d[a] = q[a] - nd.c[a];
    float dist = std::sqrt(sqnorm<D>(d));
    return dist / nd.diam;
  }
  Src<D> node_src(const Node<D>& nd) const { return {nd.c, nd.n, nd.mass, nd.alpha}; }
  Src<D> rec_src(int64_t i) const {
    const Rec<D>& r = recs[i];
    return {r.pos, r.nrm, r.mass, r.alpha};
  }
  // tree3d.h:155-170 (index form): child holding data index `idx`, skipping
  // single-child chains.
  int64_t child_containing(int64_t node, int64_t idx) const {
    const Node<D>& nd = nodes[node];
    for (int c = 0; c < nd.nchild; ++c) {
      int64_t ci = nd.first_child + c;
      const Node<D>& ch = nodes[ci];
      if (ch.start <= idx && idx < ch.end) {
        if (ch.nchild == 1) return child_containing(ci, idx);
        return ci;
      }
    }
    return -1;
  }
};

template <int D>
struct BuildInfo {
  int64_t start, end;
  Box<D> box;
  int depth;
  int64_t idx;
  int64_t parent;
};

template <int D>
void build_tree(Tree<D>& t, const Box<D>& root_box) {
  const int Ds = t.dsplit;
  const int B = (D == 3) ? Ds * Ds * Ds : Ds * Ds;
  std::deque<BuildInfo<D>> queue;
  int64_t next_idx = 0;
  queue.push_back({0, (int64_t)t.recs.size(), root_box, 0, next_idx++, 0});
  t.nodes.clear();
  t.nodes.reserve(t.recs.size() * 2 + 1);
  while (!queue.empty()) {
    BuildInfo<D> bi = queue.front();
    queue.pop_front();
    if ((int64_t)t.nodes.size() <= bi.idx) t.nodes.resize(bi.idx + 1);
    Node<D>& nd = t.nodes[bi.idx];
    nd = Node<D>();
    nd.start = bi.start;
    nd.end = bi.end;
    nd.parent = bi.parent;
    nd.depth = bi.depth;
    nd.box = bi.box;
    float diag[D];
    for (int a = 0; a < D; ++a) diag[a] = bi.box.hi[a] - bi.box.lo[a];
    nd.diam = std::sqrt(sqnorm<D>(diag));  // tree3d.h:249
    nd.nchild = 0;
    nd.first_child = -1;
    nd.leaf = bi.depth >= t.max_depth || (bi.end - bi.start) <= t.leaf_size;
    for (int a = 0; a < D; ++a) nd.c[a] = nd.n[a] = 0.f;
    nd.mass = 0.f;
    nd.alpha = 0.f;
    t.global_max_depth = std::max(t.global_max_depth, bi.depth);
    if (nd.leaf) {  // tree3d.h:255-264: array order, fp32
      for (int64_t i = bi.start; i < bi.end; ++i) {
        const Rec<D>& r = t.recs[i];
        nd.alpha = r.alpha;
        nd.mass += r.mass;
        for (int a = 0; a < D; ++a) nd.c[a] += r.mass * r.pos[a];
        for (int a = 0; a < D; ++a) nd.n[a] += r.mass * r.nrm[a];
      }
      for (int a = 0; a < D; ++a) nd.c[a] /= nd.mass;
      for (int a = 0; a < D; ++a) nd.n[a] /= nd.mass;
      continue;
    }
    // tree3d.h:268-307: one two-ended std::partition per child cell, cells in
    // raster order (x fastest), each against the node's own bounds.
    float inc[D];
    for (int a = 0; a < D; ++a) inc[a] = diag[a] / (float)Ds;
    int64_t s = bi.start;
    const Box<D> nb = bi.box;
    for (int cell = 0; cell < B; ++cell) {
      int coord[3] = {cell & (Ds - 1), (cell >> t.split_log) & (Ds - 1),
                      (cell >> (2 * t.split_log)) & (Ds - 1)};
      Box<D> sub;
      for (int a = 0; a < D; ++a) {
        sub.lo[a] = nb.lo[a] + (float)coord[a] * inc[a];
        sub.hi[a] = sub.lo[a] + inc[a];
      }
      Rec<D>* first = t.recs.data() + s;
      Rec<D>* last = t.recs.data() + bi.end;
      Rec<D>* split = std::partition(first, last, [&](const Rec<D>& r) {
        int qv[D];
        quantize<D>(nb, r.pos, Ds, qv);
        for (int a = 0; a < D; ++a)
          if (qv[a] != coord[a]) return false;
        return true;
      });
      int64_t split_idx = (split - t.recs.data());
      if (split_idx > s) {
        int64_t ci = next_idx++;
        if (nd.nchild == 0) nd.first_child = ci;
        nd.nchild++;
        queue.push_back({s, split_idx, sub, bi.depth + 1, ci, bi.idx});
      }
      s = split_idx;
    }
  }
  // tree3d.h:360-365 / 310-327: reverse sweep, children in bitmask order.
  for (int64_t i = (int64_t)t.nodes.size() - 1; i >= 0; --i) {
    Node<D>& nd = t.nodes[i];
    if (!nd.leaf) {
      for (int c = 0; c < nd.nchild; ++c) {
        const Node<D>& ch = t.nodes[nd.first_child + c];
        nd.alpha = ch.alpha;
        nd.mass += ch.mass;
        for (int a = 0; a < D; ++a) nd.c[a] += ch.mass * ch.c[a];
        for (int a = 0; a < D; ++a) nd.n[a] += ch.mass * ch.n[a];
      }
      for (int a = 0; a < D; ++a) nd.c[a] /= nd.mass;
      for (int a = 0; a < D; ++a) nd.n[a] /= nd.mass;
    }
  }
  // float64 moments: leaves sum their points, internal nodes their children.
  for (int64_t i = (int64_t)t.nodes.size() - 1; i >= 0; --i) {
    Node<D>& nd = t.nodes[i];
    double m = 0, c[D] = {}, n[D] = {};
    if (nd.leaf) {
      for (int64_t j = nd.start; j < nd.end; ++j) {
        const Rec<D>& r = t.recs[j];
        m += r.mass;
        for (int a = 0; a < D; ++a) {
          c[a] += (double)r.mass * r.pos[a];
          n[a] += (double)r.mass * r.nrm[a];
        }
      }
    } else {
      for (int k = 0; k < nd.nchild; ++k) {
        const Node<D>& ch = t.nodes[nd.first_child + k];
        m += ch.mass64;
        for (int a = 0; a < D; ++a) {
          c[a] += ch.mass64 * ch.c64[a];
          n[a] += ch.mass64 * ch.n64[a];
        }
      }
    }
    nd.mass64 = m;
    for (int a = 0; a < D; ++a) {
      nd.c64[a] = c[a] / m;
      nd.n64[a] = n[a] / m;
    }
  }
}

// ---------------------------------------------------------------------------
// Brute force.  mode 0: fp32 Kahan in leaf order, include/brute_force_gpu.h:45-63.
// mode 1: float64 kernels summed in float64.  mode 2: float64 sum of |term|
// (the condition of the sum, used to bound rounding error).
// ---------------------------------------------------------------------------
template <int D>
void brute_force(const Tree<D>& t, int kernel, int mode, const float* Q, int64_t nq,
                 double* out, int nthreads) {
  KernelFn<D> f = pick_kernel<D>(kernel);
  auto work = [&](int tid) {
    for (int64_t qi = tid; qi < nq; qi += nthreads) {
      const float* q = Q + qi * D;
      if (mode == 0) {
        float total = 0.f, c = 0.f;
        for (size_t j = 0; j < t.recs.size(); ++j) {
          float kv = f(t.rec_src(j), q);
          float y = kv - c;
          float tt = total + y;
          c = (tt - total) - y;
          total = tt;
        }
        out[qi] = total;
      } else if (mode == 1) {
        double total = 0.0;
        for (size_t j = 0; j < t.recs.size(); ++j) total += k64<D>(kernel, t.rec_src(j), q);
        out[qi] = total;
      } else {
        double total = 0.0;
        for (size_t j = 0; j < t.recs.size(); ++j)
          total += std::fabs(k64<D>(kernel, t.rec_src(j), q));
        out[qi] = total;
      }
    }
  };
  std::vector<std::thread> th;
  for (int i = 0; i < nthreads; ++i) th.emplace_back(work, i);
  for (auto& x : th) x.join();
}

// ---------------------------------------------------------------------------
// Deterministic Barnes-Hut: include/barnes_hut.h:23-78 (per query) and
// include/barnes_hut_gpu.h:19-72 (vote).  Depth-first, children in bitmask
// order.  A vote group of G consecutive queries descends together unless every
// live member accepts the node; G = 1 is the per-query walk.
// ---------------------------------------------------------------------------
template <int D>
void barnes_hut(const Tree<D>& t, int kernel, float beta, int group, const float* Q,
                int64_t nq, float* out, int32_t* counts, int nthreads) {
  KernelFn<D> f = pick_kernel<D>(kernel);
  if (group < 1) group = 1;
  int64_t ngroups = (nq + group - 1) / group;
  auto work = [&](int tid) {
    std::vector<int64_t> stack;
    std::vector<float> tot(group);
    std::vector<int32_t> cnt(group);
    for (int64_t g = tid; g < ngroups; g += nthreads) {
      int64_t q0 = g * group;
      int live = (int)std::min<int64_t>(group, nq - q0);
      for (int l = 0; l < live; ++l) {
        tot[l] = 0.f;
        cnt[l] = 0;
      }
      stack.clear();
      stack.push_back(0);
      while (!stack.empty()) {
        int64_t v = stack.back();
        stack.pop_back();
        const Node<D>& nd = t.nodes[v];
        bool all = true;
        for (int l = 0; l < live && all; ++l) {
          float r = t.ratio(nd, Q + (q0 + l) * D);
          if (!(r >= beta)) all = false;
        }
        if (all) {
          for (int l = 0; l < live; ++l) {
            const float* q = Q + (q0 + l) * D;
            if (nd.leaf) {
              for (int64_t i = nd.start; i < nd.end; ++i) {
                tot[l] += f(t.rec_src(i), q);
                cnt[l]++;
              }
            } else {
              tot[l] += f(t.node_src(nd), q);
              cnt[l]++;
            }
          }
        } else if (!nd.leaf) {
          for (int c = nd.nchild - 1; c >= 0; --c) stack.push_back(nd.first_child + c);
        }
      }
      for (int l = 0; l < live; ++l) {
        out[q0 + l] = tot[l];
        if (counts) counts[q0 + l] = cnt[l];
      }
    }
  };
  std::vector<std::thread> th;
  for (int i = 0; i < nthreads; ++i) th.emplace_back(work, i);
  for (auto& x : th) x.join();
}

// ---------------------------------------------------------------------------
// Stochastic control-variate estimator: include/tree_sampler.h:21-113 and
// include/tree_sampler_gpu.h:25-140.  Queries qi share their random paths with
// the other members of group qi / group (the reference shares sample points
// per thread block); each query still follows its own Russian roulette.
// If the root is itself a leaf its points are summed exactly.
// ---------------------------------------------------------------------------
template <int D>
float stochastic_one(const Tree<D>& t, KernelFn<D> f, const float* q, int spl, uint64_t seed,
                     uint32_t group_id, int32_t* count_out) {
  const Node<D>& root = t.nodes[0];
  int32_t contribs = 0;
  if (root.leaf) {
    float s = 0.f;
    for (int64_t i = root.start; i < root.end; ++i) {
      s += f(t.rec_src(i), q);
      contribs++;
    }
    if (count_out) *count_out = contribs;
    return s;
  }
  float bh = 0.f, mlcv = 0.f;
  for (int si = 0; si < root.nchild; ++si) {
    int64_t sidx = root.first_child + si;
    const Node<D>& S = t.nodes[sidx];
    if (S.leaf) {
      for (int64_t i = S.start; i < S.end; ++i) {
        bh += f(t.rec_src(i), q);
        contribs++;
      }
    } else {
      bh += f(t.node_src(S), q);
      contribs++;
    }
    uint64_t nS = (uint64_t)(S.end - S.start);
    float sub_total = 0.f;
    for (int ss = 0; ss < spl; ++ss) {
      PathStream ps;
      ps.k0 = (uint32_t)seed;
      ps.k1 = (uint32_t)(seed >> 32);
      ps.c0 = (uint32_t)ss;
      ps.c1 = (uint32_t)si;
      ps.c2 = group_id;
      int64_t idx = S.start + (int64_t)scale_draw(ps.point_draw(), nS);
      float sample_prob_inv = (float)nS;
      int64_t prev = sidx, cur = sidx;
      float event_prob = 1.0f;
      float wd_all = 0.f;
      float cur_beta = std::fmax(1.0f, t.ratio(t.nodes[cur], q));
      uint32_t level = 0;
      while (!t.nodes[cur].leaf) {
        float ffr = t.ratio(t.nodes[cur], q);
        float tp = std::fmin(cur_beta / ffr, 1.0f);
        float p = ps.rr_uniform(level++);
        if (p > tp) {
          event_prob *= 1.0f - tp;
          break;
        }
        event_prob *= tp;
        prev = cur;
        cur = t.child_containing(cur, idx);
        cur_beta = std::fmax(cur_beta, ffr);
        const Node<D>& P = t.nodes[prev];
        float prev_result = f(t.node_src(P), q);
        float cur_result = 0.f;
        for (int c = 0; c < P.nchild; ++c) {
          const Node<D>& ch = t.nodes[P.first_child + c];
          if (ch.leaf) {
            for (int64_t i = ch.start; i < ch.end; ++i) {
              cur_result += f(t.rec_src(i), q);
              contribs++;
            }
          } else {
            cur_result += f(t.node_src(ch), q);
            contribs++;
          }
        }
        float diff = cur_result - prev_result;
        float wd = sample_prob_inv / (event_prob * (float)(uint64_t)(P.end - P.start)) * diff;
        wd_all += wd;
      }
      sub_total += wd_all;
    }
    mlcv += sub_total / (float)spl;
  }
  if (count_out) *count_out = contribs;
  return bh + mlcv;
}

template <int D>
void stochastic(const Tree<D>& t, int kernel, int spl, uint64_t seed, int group, const float* Q,
                int64_t nq, float* out, int32_t* counts, int nthreads) {
  KernelFn<D> f = pick_kernel<D>(kernel);
  if (group < 1) group = 1;
  auto work = [&](int tid) {
    for (int64_t qi = tid; qi < nq; qi += nthreads) {
      out[qi] = stochastic_one<D>(t, f, Q + qi * D, spl, seed, (uint32_t)(qi / group),
                                  counts ? counts + qi : nullptr);
    }
  };
  std::vector<std::thread> th;
  for (int i = 0; i < nthreads; ++i) th.emplace_back(work, i);
  for (auto& x : th) x.join();
}

struct AnyTree {
  int dim;
  Tree<2>* t2 = nullptr;
  Tree<3>* t3 = nullptr;
};

int resolve_threads(int n) {
  if (n > 0) return n;
  int hc = (int)std::thread::hardware_concurrency();
  return hc == 0 ? 8 : hc;  // src/sbh.cpp:32-36
}

template <int D>
Tree<D>* make_tree(int split_log, int leaf_size, int max_depth, int64_t n, const float* pts,
                   const float* nrm, const float* mass, const float* minc, const float* maxc,
                   float alpha) {
  Tree<D>* t = new Tree<D>();
  t->split_log = split_log;
  t->dsplit = 1 << split_log;
  t->leaf_size = leaf_size;
  t->max_depth = max_depth;
  t->recs.resize(n);
  for (int64_t i = 0; i < n; ++i) {
    Rec<D>& r = t->recs[i];
    for (int a = 0; a < D; ++a) {
      r.pos[a] = pts[i * D + a];
      r.nrm[a] = nrm ? nrm[i * D + a] : 0.f;
    }
    r.mass = mass[i];
    r.alpha = alpha;
    r.orig = i;
  }
  Box<D> b;
  for (int a = 0; a < D; ++a) {
    b.lo[a] = minc[a];
    b.hi[a] = maxc[a];
  }
  build_tree<D>(*t, b);
  return t;
}

template <int D>
void export_nodes(const Tree<D>& t, int64_t* start, int64_t* end, int64_t* first_child,
                  int32_t* nchild, int32_t* depth, int64_t* parent, uint8_t* leaf, float* bmin,
                  float* bmax, float* diam, float* centroid, float* mass, float* normal,
                  double* centroid64, double* mass64, double* normal64) {
  for (size_t i = 0; i < t.nodes.size(); ++i) {
    const Node<D>& nd = t.nodes[i];
    if (start) start[i] = nd.start;
    if (end) end[i] = nd.end;
    if (first_child) first_child[i] = nd.first_child;
    if (nchild) nchild[i] = nd.nchild;
    if (depth) depth[i] = nd.depth;
    if (parent) parent[i] = nd.parent;
    if (leaf) leaf[i] = nd.leaf ? 1 : 0;
    if (diam) diam[i] = nd.diam;
    if (mass) mass[i] = nd.mass;
    if (mass64) mass64[i] = nd.mass64;
    for (int a = 0; a < D; ++a) {
      if (bmin) bmin[i * D + a] = nd.box.lo[a];
      if (bmax) bmax[i * D + a] = nd.box.hi[a];
      if (centroid) centroid[i * D + a] = nd.c[a];
      if (normal) normal[i * D + a] = nd.n[a];
      if (centroid64) centroid64[i * D + a] = nd.c64[a];
      if (normal64) normal64[i * D + a] = nd.n64[a];
    }
  }
}

}  // namespace

extern "C" {

void* ot_build(int dim, int split_log, int leaf_size, int max_depth, int64_t n, const float* pts,
               const float* nrm, const float* mass, const float* minc, const float* maxc,
               float alpha) {
  AnyTree* at = new AnyTree();
  at->dim = dim;
  if (dim == 2)
    at->t2 = make_tree<2>(split_log, leaf_size, max_depth, n, pts, nrm, mass, minc, maxc, alpha);
  else
    at->t3 = make_tree<3>(split_log, leaf_size, max_depth, n, pts, nrm, mass, minc, maxc, alpha);
  return at;
}

void ot_free(void* h) {
  AnyTree* at = (AnyTree*)h;
  delete at->t2;
  delete at->t3;
  delete at;
}

int64_t ot_num_nodes(void* h) {
  AnyTree* at = (AnyTree*)h;
  return at->dim == 2 ? (int64_t)at->t2->nodes.size() : (int64_t)at->t3->nodes.size();
}

int ot_max_depth(void* h) {
  AnyTree* at = (AnyTree*)h;
  return at->dim == 2 ? at->t2->global_max_depth : at->t3->global_max_depth;
}

void ot_leaf_order(void* h, int64_t* out) {
  AnyTree* at = (AnyTree*)h;
  if (at->dim == 2)
    for (size_t i = 0; i < at->t2->recs.size(); ++i) out[i] = at->t2->recs[i].orig;
  else
    for (size_t i = 0; i < at->t3->recs.size(); ++i) out[i] = at->t3->recs[i].orig;
}

void ot_nodes(void* h, int64_t* start, int64_t* end, int64_t* first_child, int32_t* nchild,
              int32_t* depth, int64_t* parent, uint8_t* leaf, float* bmin, float* bmax,
              float* diam, float* centroid, float* mass, float* normal, double* centroid64,
              double* mass64, double* normal64) {
  AnyTree* at = (AnyTree*)h;
  if (at->dim == 2)
    export_nodes<2>(*at->t2, start, end, first_child, nchild, depth, parent, leaf, bmin, bmax,
                    diam, centroid, mass, normal, centroid64, mass64, normal64);
  else
    export_nodes<3>(*at->t3, start, end, first_child, nchild, depth, parent, leaf, bmin, bmax,
                    diam, centroid, mass, normal, centroid64, mass64, normal64);
}

void ot_brute_force(void* h, int kernel, int mode, const float* q, int64_t nq, double* out,
                    int nthreads) {
  AnyTree* at = (AnyTree*)h;
  nthreads = resolve_threads(nthreads);
  if (at->dim == 2)
    brute_force<2>(*at->t2, kernel, mode, q, nq, out, nthreads);
  else
    brute_force<3>(*at->t3, kernel, mode, q, nq, out, nthreads);
}

void ot_barnes_hut(void* h, int kernel, float beta, int group, const float* q, int64_t nq,
                   float* out, int32_t* counts, int nthreads) {
  AnyTree* at = (AnyTree*)h;
  nthreads = resolve_threads(nthreads);
  if (at->dim == 2)
    barnes_hut<2>(*at->t2, kernel, beta, group, q, nq, out, counts, nthreads);
  else
    barnes_hut<3>(*at->t3, kernel, beta, group, q, nq, out, counts, nthreads);
}

void ot_stochastic(void* h, int kernel, int spl, uint64_t seed, int group, const float* q,
                   int64_t nq, float* out, int32_t* counts, int nthreads) {
  AnyTree* at = (AnyTree*)h;
  nthreads = resolve_threads(nthreads);
  if (at->dim == 2)
    stochastic<2>(*at->t2, kernel, spl, seed, group, q, nq, out, counts, nthreads);
  else
    stochastic<3>(*at->t3, kernel, spl, seed, group, q, nq, out, counts, nthreads);
}

// Single-pair kernel evaluations (fp32 formula and float64), for unit tests.
void ot_kernel_pairs(int dim, int kernel, int64_t n, const float* pos, const float* nrm,
                     const float* mass, float alpha, const float* q, float* out32,
                     double* out64) {
  for (int64_t i = 0; i < n; ++i) {
    if (dim == 2) {
      Src<2> s = {pos + 2 * i, nrm + 2 * i, mass[i], alpha};
      out32[i] = pick_kernel<2>(kernel)(s, q + 2 * i);
      out64[i] = k64<2>(kernel, s, q + 2 * i);
    } else {
      Src<3> s = {pos + 3 * i, nrm + 3 * i, mass[i], alpha};
      out32[i] = pick_kernel<3>(kernel)(s, q + 3 * i);
      out64[i] = k64<3>(kernel, s, q + 3 * i);
    }
  }
}

void ot_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
  U4 c = {{ctr[0], ctr[1], ctr[2], ctr[3]}};
  U4 r = philox(c, key[0], key[1]);
  for (int i = 0; i < 4; ++i) out[i] = r.v[i];
}

// Path-stream words as the estimator consumes them (for RNG layout tests).
void ot_path_words(uint64_t seed, uint32_t group, uint32_t subdomain, uint32_t sample,
                   uint32_t nwords, uint32_t* out) {
  PathStream ps;
  ps.k0 = (uint32_t)seed;
  ps.k1 = (uint32_t)(seed >> 32);
  ps.c0 = sample;
  ps.c1 = subdomain;
  ps.c2 = group;
  for (uint32_t k = 0; k < nwords; ++k) out[k] = ps.word(k);
}

}  // extern "C"
