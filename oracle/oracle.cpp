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
#include <type_traits>
#include <utility>
#include <vector>

namespace {

// Scalar type R of a build: float (the reference's default) or double (its
// USE_DOUBLE build, include/utils.h:13-29, where FLOAT = double, literals are
// written FP(x) and PI = M_PI).
template <typename R>
struct Lit;
template <>
struct Lit<float> {
  static constexpr float tenth = 1e-1f;
  static constexpr float eps = 1e-5f;
  static constexpr float pi = (float)M_PI;  // include/utils.h:41-45 (PI = M_PI, cast)
};
template <>
struct Lit<double> {
  static constexpr double tenth = 1e-1;
  static constexpr double eps = 1e-5;
  static constexpr double pi = M_PI;
};

// ---------------------------------------------------------------------------
// Source record: include/leaf_data.h:8-16 (position, normal, mass, alpha).
// `orig` is the input row index; it rides along with the record so the final
// permutation (leaf order) can be reported.  It never enters a predicate.
// ---------------------------------------------------------------------------
template <int D, typename R>
struct Rec {
  R pos[D];
  R nrm[D];
  R mass;
  R alpha;
  int64_t orig;
};

template <int D, typename R>
struct Box {
  R lo[D];
  R hi[D];
};

// x86 cvttss2si / cvttsd2si semantics for float / double -> int (the
// reference's Eigen cast<int> at include/aabb.h:34-41 compiles to them):
// out-of-range and NaN give INT_MIN.
template <typename R>
inline int trunc_x86(R v) {
  if (!(v >= (R)-2147483648.0 && v < (R)2147483648.0)) return INT_MIN;
  return (int)v;
}

// include/aabb.h:34-41: clamp(int((D*(p-min)) / (max-min)), 0, D-1) per axis.
template <int D, typename R>
inline void quantize(const Box<D, R>& b, const R* p, int dsplit, int* out) {
  for (int a = 0; a < D; ++a) {
    R num = (R)dsplit * (p[a] - b.lo[a]);
    R den = b.hi[a] - b.lo[a];
    int v = trunc_x86<R>(num / den);
    v = std::min(v, dsplit - 1);
    v = std::max(v, 0);
    out[a] = v;
  }
}

// Eigen fixed-size reductions split in halves: size 3 sums a0 + (a1 + a2).
template <int D, typename R>
inline R sqnorm(const R* d) {
  if constexpr (D == 2)
    return d[0] * d[0] + d[1] * d[1];
  else
    return d[0] * d[0] + (d[1] * d[1] + d[2] * d[2]);
}
template <int D, typename R>
inline R dot(const R* a, const R* b) {
  if constexpr (D == 2)
    return a[0] * b[0] + a[1] * b[1];
  else
    return a[0] * b[0] + (a[1] * b[1] + a[2] * b[2]);
}

// ---------------------------------------------------------------------------
// Kernel functions: include/kernel.h:31-59.  A "source" is anything with
// position / normal / mass / alpha (a point record or a node moment).
// ---------------------------------------------------------------------------
template <int D, typename R>
struct Src {
  const R* pos;
  const R* nrm;
  R mass;
  R alpha;
};

template <int D, typename R>
using KernelFn = R (*)(const Src<D, R>&, const R* q);

template <int D, typename R>
R k_gravity(const Src<D, R>& s, const R* q) {  // kernel.h:31-35
  R d[D];
  for (int a = 0; a < D; ++a) d[a] = s.pos[a] - q[a];
  R dist = std::sqrt(sqnorm<D, R>(d));
  return -s.mass / (dist + Lit<R>::tenth);
}

template <int D, typename R>
R k_wn(const Src<D, R>& s, const R* q) {
  if constexpr (D == 2) {  // kernel.h:40-45
    R d[2] = {s.pos[0] - q[0], s.pos[1] - q[1]};
    R dist_sq = sqnorm<2, R>(d);
    R qdn = dot<2, R>(s.nrm, d);
    return s.mass * qdn / (2 * Lit<R>::pi * (dist_sq + Lit<R>::eps));
  } else {  // kernel.h:47-53
    R d[3] = {s.pos[0] - q[0], s.pos[1] - q[1], s.pos[2] - q[2]};
    R dist = std::sqrt(sqnorm<3, R>(d));
    R qdn = dot<3, R>(s.nrm, d);
    return s.mass * qdn / (4 * Lit<R>::pi * (dist * dist * dist + Lit<R>::eps));
  }
}

template <int D, typename R>
R k_smooth(const Src<D, R>& s, const R* q) {  // kernel.h:55-59
  R d[D];
  for (int a = 0; a < D; ++a) d[a] = s.pos[a] - q[a];
  R dist = std::sqrt(sqnorm<D, R>(d));
  return std::exp(-s.alpha * dist);
}

// helmholtzKernelX / helmholtzKernelY (kernel.h:68-83): m cos(freq 2 pi d) /
// (d + 0.1) with freq 2, and m sin(freq 2 pi d) / (d + 0.1) with freq 1; the
// argument is ((freq * 2) * pi) * dist, cos / sin of the build's type.
template <int D, int Y, typename R>
R k_helm(const Src<D, R>& s, const R* q) {
  R d[D];
  for (int a = 0; a < D; ++a) d[a] = s.pos[a] - q[a];
  R dist = std::sqrt(sqnorm<D, R>(d));
  const R freq = Y ? (R)1 : (R)2;
  R arg = freq * 2 * Lit<R>::pi * dist;
  R c = Y ? std::sin(arg) : std::cos(arg);
  return s.mass * c / (dist + Lit<R>::tenth);
}

// Component `A` of gravityGrad (kernel.h:61-66): -m / (d + 0.1)^3 * (q - p).
// The reference's template is never instantiated (no plugin binds it, and as
// written it would mix a double pow with a float vector); the form here is
// t = d + 0.1, s = -m / ((t * t) * t), s * (q_a - p_a).
template <int D, int A, typename R>
R k_grad(const Src<D, R>& s, const R* q) {
  R d[D];
  for (int a = 0; a < D; ++a) d[a] = q[a] - s.pos[a];
  R dist = std::sqrt(sqnorm<D, R>(d));
  R t = dist + Lit<R>::tenth;
  R sc = -s.mass / (t * t * t);
  return sc * d[A];
}

// float64 restatement of the fp32 formulas (constants are the fp32 literals
// promoted, so this is the exact value of the fp32 formula's mathematics).
template <int D>
double k64(int kernel, const Src<D, float>& s, const float* q) {
  double d[D];
  double dd = 0.0, dn = 0.0;
  for (int a = 0; a < D; ++a) {
    d[a] = (double)s.pos[a] - (double)q[a];
    dd += d[a] * d[a];
    dn += (double)s.nrm[a] * d[a];
  }
  double dist = std::sqrt(dd);
  const double pi = (double)Lit<float>::pi;
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

template <int D, typename R>
KernelFn<D, R> pick_kernel(int kernel) {
  switch (kernel) {
    case 0:
      return k_gravity<D, R>;
    case 1:
      return k_wn<D, R>;
    case 2:
      return k_smooth<D, R>;
    case 3:
      return k_helm<D, 0, R>;
    case 4:
      return k_helm<D, 1, R>;
    case 5:
      return k_grad<D, 0, R>;
    case 6:
      return k_grad<D, 1, R>;
    case 7:
      if constexpr (D == 3) return k_grad<D, 2, R>;
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
  // Russian-roulette uniform in (0, 1]: the word's top 24 bits in a float
  // build, all 32 bits in a double build.
  template <typename R>
  R rr_uniform(uint32_t level) {
    uint32_t w = word(level + 2);
    if constexpr (sizeof(R) == 4)
      return (float)((w >> 8) + 1u) * 0x1p-24f;
    else
      return ((double)w + 1.0) * 0x1p-32;
  }
};

// Uniform index in [0, n) from a 64-bit draw (multiply-high, Lemire).
inline uint64_t scale_draw(uint64_t u, uint64_t n) {
  return (uint64_t)(((unsigned __int128)u * (unsigned __int128)n) >> 64);
}

// ---------------------------------------------------------------------------
// Tree: include/tree3d.h:238-371 and include/tree2d.h:237-365.
// ---------------------------------------------------------------------------
template <int D, typename R>
struct Node {
  int64_t start, end, parent;
  int64_t first_child;  // children are consecutive (tree3d.h:400-411)
  int nchild;
  int depth;
  bool leaf;
  Box<D, R> box;
  R diam;
  // moment (contrib_data): position, normal, mass, alpha
  R c[D];
  R n[D];
  R mass;
  R alpha;
  // float64 aggregation of the same moments (reported alongside)
  double c64[D];
  double n64[D];
  double mass64;
};

template <int D, typename R>
struct Tree {
  static constexpr int kDim = D;
  int dsplit;      // 1 << split_log
  int split_log;
  int leaf_size;
  int max_depth;
  std::vector<Rec<D, R>> recs;  // permuted in place by the build (leaf order)
  std::vector<Node<D, R>> nodes;
  int global_max_depth = 0;

  R ratio(const Node<D, R>& nd, const R* q) const {  // tree3d.h:134-140
    if (nd.leaf) return (R)INFINITY;
    R d[D];
    for (int a = 0; a < D; ++a) d[a] = q[a] - nd.c[a];
    R dist = std::sqrt(sqnorm<D, R>(d));
    return dist / nd.diam;
  }
  Src<D, R> node_src(const Node<D, R>& nd) const { return {nd.c, nd.n, nd.mass, nd.alpha}; }
  Src<D, R> rec_src(int64_t i) const {
    const Rec<D, R>& r = recs[i];
    return {r.pos, r.nrm, r.mass, r.alpha};
  }
  // tree3d.h:155-170 (index form): child holding data index `idx`, skipping
  // single-child chains.
  int64_t child_containing(int64_t node, int64_t idx) const {
    const Node<D, R>& nd = nodes[node];
    for (int c = 0; c < nd.nchild; ++c) {
      int64_t ci = nd.first_child + c;
      const Node<D, R>& ch = nodes[ci];
      if (ch.start <= idx && idx < ch.end) {
        if (ch.nchild == 1) return child_containing(ci, idx);
        return ci;
      }
    }
    return -1;
  }
};

template <int D, typename R>
struct BuildInfo {
  int64_t start, end;
  Box<D, R> box;
  int depth;
  int64_t idx;
  int64_t parent;
};

template <int D, typename R>
void build_tree(Tree<D, R>& t, const Box<D, R>& root_box) {
  const int Ds = t.dsplit;
  const int B = (D == 3) ? Ds * Ds * Ds : Ds * Ds;
  std::deque<BuildInfo<D, R>> queue;
  int64_t next_idx = 0;
  queue.push_back({0, (int64_t)t.recs.size(), root_box, 0, next_idx++, 0});
  t.nodes.clear();
  t.nodes.reserve(t.recs.size() * 2 + 1);
  while (!queue.empty()) {
    BuildInfo<D, R> bi = queue.front();
    queue.pop_front();
    if ((int64_t)t.nodes.size() <= bi.idx) t.nodes.resize(bi.idx + 1);
    Node<D, R>& nd = t.nodes[bi.idx];
    nd = Node<D, R>();
    nd.start = bi.start;
    nd.end = bi.end;
    nd.parent = bi.parent;
    nd.depth = bi.depth;
    nd.box = bi.box;
    R diag[D];
    for (int a = 0; a < D; ++a) diag[a] = bi.box.hi[a] - bi.box.lo[a];
    nd.diam = std::sqrt(sqnorm<D, R>(diag));  // tree3d.h:249
    nd.nchild = 0;
    nd.first_child = -1;
    nd.leaf = bi.depth >= t.max_depth || (bi.end - bi.start) <= t.leaf_size;
    for (int a = 0; a < D; ++a) nd.c[a] = nd.n[a] = (R)0;
    nd.mass = (R)0;
    nd.alpha = (R)0;
    t.global_max_depth = std::max(t.global_max_depth, bi.depth);
    if (nd.leaf) {  // tree3d.h:255-264: array order, in the build's type
      for (int64_t i = bi.start; i < bi.end; ++i) {
        const Rec<D, R>& r = t.recs[i];
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
    R inc[D];
    for (int a = 0; a < D; ++a) inc[a] = diag[a] / (R)Ds;
    int64_t s = bi.start;
    const Box<D, R> nb = bi.box;
    for (int cell = 0; cell < B; ++cell) {
      int coord[3] = {cell & (Ds - 1), (cell >> t.split_log) & (Ds - 1),
                      (cell >> (2 * t.split_log)) & (Ds - 1)};
      Box<D, R> sub;
      for (int a = 0; a < D; ++a) {
        sub.lo[a] = nb.lo[a] + (R)coord[a] * inc[a];
        sub.hi[a] = sub.lo[a] + inc[a];
      }
      Rec<D, R>* first = t.recs.data() + s;
      Rec<D, R>* last = t.recs.data() + bi.end;
      Rec<D, R>* split = std::partition(first, last, [&](const Rec<D, R>& r) {
        int qv[D];
        quantize<D, R>(nb, r.pos, Ds, qv);
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
    Node<D, R>& nd = t.nodes[i];
    if (!nd.leaf) {
      for (int c = 0; c < nd.nchild; ++c) {
        const Node<D, R>& ch = t.nodes[nd.first_child + c];
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
    Node<D, R>& nd = t.nodes[i];
    double m = 0, c[D] = {}, n[D] = {};
    if (nd.leaf) {
      for (int64_t j = nd.start; j < nd.end; ++j) {
        const Rec<D, R>& r = t.recs[j];
        m += r.mass;
        for (int a = 0; a < D; ++a) {
          c[a] += (double)r.mass * r.pos[a];
          n[a] += (double)r.mass * r.nrm[a];
        }
      }
    } else {
      for (int k = 0; k < nd.nchild; ++k) {
        const Node<D, R>& ch = t.nodes[nd.first_child + k];
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
// Brute force.  mode 0: Kahan summation in the build's type, in leaf order,
// include/brute_force_gpu.h:45-63.  mode 1: float64 kernels of the fp32
// formulas summed in float64 (float builds).  mode 2: float64 sum of |term|
// (the condition of the sum, used to bound rounding error; float builds).
// ---------------------------------------------------------------------------
template <int D, typename R>
void brute_force(const Tree<D, R>& t, int kernel, int mode, const R* Q, int64_t nq,
                 double* out, int nthreads) {
  KernelFn<D, R> f = pick_kernel<D, R>(kernel);
  auto work = [&](int tid) {
    for (int64_t qi = tid; qi < nq; qi += nthreads) {
      const R* q = Q + qi * D;
      if (mode == 0) {
        R total = 0, c = 0;
        for (size_t j = 0; j < t.recs.size(); ++j) {
          R kv = f(t.rec_src(j), q);
          R y = kv - c;
          R tt = total + y;
          c = (tt - total) - y;
          total = tt;
        }
        out[qi] = total;
      } else if constexpr (sizeof(R) == 8) {
        out[qi] = NAN;
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
template <int D, typename R>
void barnes_hut(const Tree<D, R>& t, int kernel, R beta, int group, const R* Q, int64_t nq,
                R* out, int32_t* counts, int nthreads) {
  KernelFn<D, R> f = pick_kernel<D, R>(kernel);
  if (group < 1) group = 1;
  int64_t ngroups = (nq + group - 1) / group;
  auto work = [&](int tid) {
    std::vector<int64_t> stack;
    std::vector<R> tot(group);
    std::vector<int32_t> cnt(group);
    for (int64_t g = tid; g < ngroups; g += nthreads) {
      int64_t q0 = g * group;
      int live = (int)std::min<int64_t>(group, nq - q0);
      for (int l = 0; l < live; ++l) {
        tot[l] = 0;
        cnt[l] = 0;
      }
      stack.clear();
      stack.push_back(0);
      while (!stack.empty()) {
        int64_t v = stack.back();
        stack.pop_back();
        const Node<D, R>& nd = t.nodes[v];
        bool all = true;
        for (int l = 0; l < live && all; ++l) {
          R r = t.ratio(nd, Q + (q0 + l) * D);
          if (!(r >= beta)) all = false;
        }
        if (all) {
          for (int l = 0; l < live; ++l) {
            const R* q = Q + (q0 + l) * D;
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
template <int D, typename R>
R stochastic_one(const Tree<D, R>& t, KernelFn<D, R> f, const R* q, int spl, uint64_t seed,
                 uint32_t group_id, int32_t* count_out) {
  const Node<D, R>& root = t.nodes[0];
  int32_t contribs = 0;
  if (root.leaf) {
    R s = 0;
    for (int64_t i = root.start; i < root.end; ++i) {
      s += f(t.rec_src(i), q);
      contribs++;
    }
    if (count_out) *count_out = contribs;
    return s;
  }
  R bh = 0, mlcv = 0;
  for (int si = 0; si < root.nchild; ++si) {
    int64_t sidx = root.first_child + si;
    const Node<D, R>& S = t.nodes[sidx];
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
    R sub_total = 0;
    for (int ss = 0; ss < spl; ++ss) {
      PathStream ps;
      ps.k0 = (uint32_t)seed;
      ps.k1 = (uint32_t)(seed >> 32);
      ps.c0 = (uint32_t)ss;
      ps.c1 = (uint32_t)si;
      ps.c2 = group_id;
      int64_t idx = S.start + (int64_t)scale_draw(ps.point_draw(), nS);
      R sample_prob_inv = (R)nS;
      int64_t prev = sidx, cur = sidx;
      R event_prob = 1;
      R wd_all = 0;
      R cur_beta = std::fmax((R)1, t.ratio(t.nodes[cur], q));
      uint32_t level = 0;
      while (!t.nodes[cur].leaf) {
        R ffr = t.ratio(t.nodes[cur], q);
        R tp = std::fmin(cur_beta / ffr, (R)1);
        R p = ps.rr_uniform<R>(level++);
        if (p > tp) {
          event_prob *= (R)1 - tp;
          break;
        }
        event_prob *= tp;
        prev = cur;
        cur = t.child_containing(cur, idx);
        cur_beta = std::fmax(cur_beta, ffr);
        const Node<D, R>& P = t.nodes[prev];
        R prev_result = f(t.node_src(P), q);
        R cur_result = 0;
        for (int c = 0; c < P.nchild; ++c) {
          const Node<D, R>& ch = t.nodes[P.first_child + c];
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
        R diff = cur_result - prev_result;
        R wd = sample_prob_inv / (event_prob * (R)(uint64_t)(P.end - P.start)) * diff;
        wd_all += wd;
      }
      sub_total += wd_all;
    }
    mlcv += sub_total / (R)spl;
  }
  if (count_out) *count_out = contribs;
  return bh + mlcv;
}

template <int D, typename R>
void stochastic(const Tree<D, R>& t, int kernel, int spl, uint64_t seed, int group, const R* Q,
                int64_t nq, R* out, int32_t* counts, int nthreads) {
  KernelFn<D, R> f = pick_kernel<D, R>(kernel);
  if (group < 1) group = 1;
  auto work = [&](int tid) {
    for (int64_t qi = tid; qi < nq; qi += nthreads) {
      out[qi] = stochastic_one<D, R>(t, f, Q + qi * D, spl, seed, (uint32_t)(qi / group),
                                     counts ? counts + qi : nullptr);
    }
  };
  std::vector<std::thread> th;
  for (int i = 0; i < nthreads; ++i) th.emplace_back(work, i);
  for (auto& x : th) x.join();
}

// One tree of either dimension and either scalar type (float: the
// reference's default build; double: its USE_DOUBLE build).
struct AnyTree {
  int dim;
  bool dbl = false;
  Tree<2, float>* f2 = nullptr;
  Tree<3, float>* f3 = nullptr;
  Tree<2, double>* d2 = nullptr;
  Tree<3, double>* d3 = nullptr;
  template <typename F>
  void visit(F&& fn) {
    if (dbl) {
      if (dim == 2) fn(*d2); else fn(*d3);
    } else {
      if (dim == 2) fn(*f2); else fn(*f3);
    }
  }
};

int resolve_threads(int n) {
  if (n > 0) return n;
  int hc = (int)std::thread::hardware_concurrency();
  return hc == 0 ? 8 : hc;  // src/sbh.cpp:32-36
}

template <int D, typename R>
Tree<D, R>* make_tree(int split_log, int leaf_size, int max_depth, int64_t n, const R* pts,
                      const R* nrm, const R* mass, const R* minc, const R* maxc, R alpha) {
  Tree<D, R>* t = new Tree<D, R>();
  t->split_log = split_log;
  t->dsplit = 1 << split_log;
  t->leaf_size = leaf_size;
  t->max_depth = max_depth;
  t->recs.resize(n);
  for (int64_t i = 0; i < n; ++i) {
    Rec<D, R>& r = t->recs[i];
    for (int a = 0; a < D; ++a) {
      r.pos[a] = pts[i * D + a];
      r.nrm[a] = nrm ? nrm[i * D + a] : (R)0;
    }
    r.mass = mass[i];
    r.alpha = alpha;
    r.orig = i;
  }
  Box<D, R> b;
  for (int a = 0; a < D; ++a) {
    b.lo[a] = minc[a];
    b.hi[a] = maxc[a];
  }
  build_tree<D, R>(*t, b);
  return t;
}

// Node arrays; `R`-typed outputs (bounds, diameter, moments) in the build's type.
template <int D, typename R>
void export_nodes(const Tree<D, R>& t, int64_t* start, int64_t* end, int64_t* first_child,
                  int32_t* nchild, int32_t* depth, int64_t* parent, uint8_t* leaf, R* bmin,
                  R* bmax, R* diam, R* centroid, R* mass, R* normal, double* centroid64,
                  double* mass64, double* normal64) {
  for (size_t i = 0; i < t.nodes.size(); ++i) {
    const Node<D, R>& nd = t.nodes[i];
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

template <typename R>
AnyTree* build_any(int dim, int split_log, int leaf_size, int max_depth, int64_t n, const R* pts,
                   const R* nrm, const R* mass, const R* minc, const R* maxc, R alpha) {
  AnyTree* at = new AnyTree();
  at->dim = dim;
  at->dbl = sizeof(R) == 8;
  if constexpr (sizeof(R) == 8) {
    if (dim == 2)
      at->d2 = make_tree<2, R>(split_log, leaf_size, max_depth, n, pts, nrm, mass, minc, maxc,
                               alpha);
    else
      at->d3 = make_tree<3, R>(split_log, leaf_size, max_depth, n, pts, nrm, mass, minc, maxc,
                               alpha);
  } else {
    if (dim == 2)
      at->f2 = make_tree<2, R>(split_log, leaf_size, max_depth, n, pts, nrm, mass, minc, maxc,
                               alpha);
    else
      at->f3 = make_tree<3, R>(split_log, leaf_size, max_depth, n, pts, nrm, mass, minc, maxc,
                               alpha);
  }
  return at;
}

// Scalar type of a visited tree.
template <typename T>
using RealOf = std::remove_reference_t<decltype(std::declval<T&>().nodes[0].mass)>;

}  // namespace

extern "C" {

// Handles of float builds (ot_build) and double builds (ot_build_double) share
// the functions below; the _double entry points take and return doubles and
// need a double build, the others a float build.
void* ot_build(int dim, int split_log, int leaf_size, int max_depth, int64_t n, const float* pts,
               const float* nrm, const float* mass, const float* minc, const float* maxc,
               float alpha) {
  return build_any<float>(dim, split_log, leaf_size, max_depth, n, pts, nrm, mass, minc, maxc,
                          alpha);
}

void* ot_build_double(int dim, int split_log, int leaf_size, int max_depth, int64_t n,
                      const double* pts, const double* nrm, const double* mass,
                      const double* minc, const double* maxc, double alpha) {
  return build_any<double>(dim, split_log, leaf_size, max_depth, n, pts, nrm, mass, minc, maxc,
                           alpha);
}

void ot_free(void* h) {
  AnyTree* at = (AnyTree*)h;
  delete at->f2;
  delete at->f3;
  delete at->d2;
  delete at->d3;
  delete at;
}

int64_t ot_num_nodes(void* h) {
  int64_t n = 0;
  ((AnyTree*)h)->visit([&](auto& t) { n = (int64_t)t.nodes.size(); });
  return n;
}

int ot_max_depth(void* h) {
  int d = 0;
  ((AnyTree*)h)->visit([&](auto& t) { d = t.global_max_depth; });
  return d;
}

void ot_leaf_order(void* h, int64_t* out) {
  ((AnyTree*)h)->visit([&](auto& t) {
    for (size_t i = 0; i < t.recs.size(); ++i) out[i] = t.recs[i].orig;
  });
}

void ot_nodes(void* h, int64_t* start, int64_t* end, int64_t* first_child, int32_t* nchild,
              int32_t* depth, int64_t* parent, uint8_t* leaf, void* bmin, void* bmax, void* diam,
              void* centroid, void* mass, void* normal, double* centroid64, double* mass64,
              double* normal64) {
  AnyTree* at = (AnyTree*)h;
  at->visit([&](auto& t) {
    using R = RealOf<decltype(t)>;
    constexpr int D = std::remove_reference_t<decltype(t)>::kDim;
    export_nodes<D, R>(t, start, end, first_child, nchild, depth, parent, leaf, (R*)bmin,
                       (R*)bmax, (R*)diam, (R*)centroid, (R*)mass, (R*)normal, centroid64,
                       mass64, normal64);
  });
}

void ot_brute_force(void* h, int kernel, int mode, const void* q, int64_t nq, double* out,
                    int nthreads) {
  nthreads = resolve_threads(nthreads);
  ((AnyTree*)h)->visit([&](auto& t) {
    using R = RealOf<decltype(t)>;
    constexpr int D = std::remove_reference_t<decltype(t)>::kDim;
    brute_force<D, R>(t, kernel, mode, (const R*)q, nq, out, nthreads);
  });
}

void ot_barnes_hut(void* h, int kernel, float beta, int group, const float* q, int64_t nq,
                   float* out, int32_t* counts, int nthreads) {
  AnyTree* at = (AnyTree*)h;
  if (at->dbl) return;
  nthreads = resolve_threads(nthreads);
  if (at->dim == 2)
    barnes_hut<2, float>(*at->f2, kernel, beta, group, q, nq, out, counts, nthreads);
  else
    barnes_hut<3, float>(*at->f3, kernel, beta, group, q, nq, out, counts, nthreads);
}

void ot_barnes_hut_double(void* h, int kernel, double beta, int group, const double* q,
                          int64_t nq, double* out, int32_t* counts, int nthreads) {
  AnyTree* at = (AnyTree*)h;
  if (!at->dbl) return;
  nthreads = resolve_threads(nthreads);
  if (at->dim == 2)
    barnes_hut<2, double>(*at->d2, kernel, beta, group, q, nq, out, counts, nthreads);
  else
    barnes_hut<3, double>(*at->d3, kernel, beta, group, q, nq, out, counts, nthreads);
}

void ot_stochastic(void* h, int kernel, int spl, uint64_t seed, int group, const float* q,
                   int64_t nq, float* out, int32_t* counts, int nthreads) {
  AnyTree* at = (AnyTree*)h;
  if (at->dbl) return;
  nthreads = resolve_threads(nthreads);
  if (at->dim == 2)
    stochastic<2, float>(*at->f2, kernel, spl, seed, group, q, nq, out, counts, nthreads);
  else
    stochastic<3, float>(*at->f3, kernel, spl, seed, group, q, nq, out, counts, nthreads);
}

void ot_stochastic_double(void* h, int kernel, int spl, uint64_t seed, int group,
                          const double* q, int64_t nq, double* out, int32_t* counts,
                          int nthreads) {
  AnyTree* at = (AnyTree*)h;
  if (!at->dbl) return;
  nthreads = resolve_threads(nthreads);
  if (at->dim == 2)
    stochastic<2, double>(*at->d2, kernel, spl, seed, group, q, nq, out, counts, nthreads);
  else
    stochastic<3, double>(*at->d3, kernel, spl, seed, group, q, nq, out, counts, nthreads);
}

// Single-pair kernel evaluations (fp32 formula and float64), for unit tests.
void ot_kernel_pairs(int dim, int kernel, int64_t n, const float* pos, const float* nrm,
                     const float* mass, float alpha, const float* q, float* out32,
                     double* out64) {
  for (int64_t i = 0; i < n; ++i) {
    if (dim == 2) {
      Src<2, float> s = {pos + 2 * i, nrm + 2 * i, mass[i], alpha};
      out32[i] = pick_kernel<2, float>(kernel)(s, q + 2 * i);
      out64[i] = k64<2>(kernel, s, q + 2 * i);
    } else {
      Src<3, float> s = {pos + 3 * i, nrm + 3 * i, mass[i], alpha};
      out32[i] = pick_kernel<3, float>(kernel)(s, q + 3 * i);
      out64[i] = k64<3>(kernel, s, q + 3 * i);
    }
  }
}

// Single-pair kernel evaluations of the double build's formulas.
void ot_kernel_pairs_double(int dim, int kernel, int64_t n, const double* pos,
                            const double* nrm, const double* mass, double alpha,
                            const double* q, double* out) {
  for (int64_t i = 0; i < n; ++i) {
    if (dim == 2) {
      Src<2, double> s = {pos + 2 * i, nrm + 2 * i, mass[i], alpha};
      out[i] = pick_kernel<2, double>(kernel)(s, q + 2 * i);
    } else {
      Src<3, double> s = {pos + 3 * i, nrm + 3 * i, mass[i], alpha};
      out[i] = pick_kernel<3, double>(kernel)(s, q + 3 * i);
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
