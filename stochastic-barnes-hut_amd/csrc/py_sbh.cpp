// pybind11 module `sbh`: the reference's CPU plugin API (src/sbh.cpp:179-271;
// include/py_tree.h:21-75), executed on the GPU.
//
//   Quadtree1/2/3, Octree1/2/3 (points, normals, masses, min, max, alpha)
//                                                    src/sbh.cpp:183-212
//   fp = 32                                          src/sbh.cpp:214
//   eval_barnes_hut_{gravity,wn,smoothdist}(tree, queries, beta) -> float32[q]
//                                                    src/sbh.cpp:216-233
//   multi_level_prefix_control_variate_{...}(tree, queries, samples_per_subdomain)
//       -> (float32[q], int32[q] terms per query)    src/sbh.cpp:235-270
//
// A host tree holds the fp32 inputs (numpy float64 is rounded to nearest, as
// the reference's Eigen conversion does) and is also accepted by
// sbh_gpu.OctreeGPUk / QuadtreeGPUk.  The evaluation functions build the
// tree on the GPU on first use (kept with the host tree) and evaluate through
// libsbh_hip.so with the reference CPU plugin's semantics: Barnes-Hut with a
// per-query opening test (include/barnes_hut.h:23-78), the stochastic
// estimator with one sample path per query (include/tree_sampler.h:21-113),
// and by default every term with correctly rounded fp32 operations in the
// reference's order (exact_terms=True), which reproduces the CPU plugin's
// arithmetic.  The reference's sampler is seeded from std::random_device; here
// `seed` (default 0) keys the Philox streams.  There is no CPU evaluation: with
// no GPU the calls raise RuntimeError.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>

#include <cmath>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sbh_hip.h"

namespace py = pybind11;
using namespace py::literals;

namespace {

using F32Array = py::array_t<float, py::array::c_style | py::array::forcecast>;

struct HostTreeBase {
  int dim = 3;
  int split_log = 1;
  int leaf_size = 1;
  int64_t n = 0;
  float alpha = 1.0f;
  std::vector<float> points, normals, masses, lo, hi;

  HostTreeBase(int d, int k, F32Array pts, F32Array nrm, F32Array m, F32Array mn, F32Array mx,
               float a, int leaf) {
    dim = d;
    split_log = k;
    leaf_size = leaf;
    alpha = a;
    if (pts.ndim() != 2 || pts.shape(1) != d)
      throw py::value_error("points must have shape (n, " + std::to_string(d) + ")");
    n = pts.shape(0);
    if (nrm.ndim() != 2 || nrm.shape(0) != n || nrm.shape(1) != d)
      throw py::value_error("normals must have the shape of points");
    if (m.ndim() != 1 || m.shape(0) != n) throw py::value_error("masses must have shape (n,)");
    if (mn.size() != d || mx.size() != d)
      throw py::value_error("min_corner / max_corner must have " + std::to_string(d) + " values");
    if (leaf < 1) throw py::value_error("leaf_size must be >= 1");
    points.assign(pts.data(), pts.data() + n * d);
    normals.assign(nrm.data(), nrm.data() + n * d);
    masses.assign(m.data(), m.data() + n);
    lo.assign(mn.data(), mn.data() + d);
    hi.assign(mx.data(), mx.data() + d);
  }

  py::array_t<float> arr(const std::vector<float>& v, std::vector<py::ssize_t> shape) const {
    py::array_t<float> a(shape);
    std::copy(v.begin(), v.end(), a.mutable_data());
    return a;
  }

  ~HostTreeBase() {
    if (dev) sbh_tree_destroy(dev);
  }

  // Device tree, built on first use.  Called without the GIL.
  sbh_tree* device() {
    std::lock_guard<std::mutex> lock(mu);
    if (!dev) {
      float ms = 0.0f;
      check(sbh_tree_create(dim, split_log, leaf_size, 100, n, points.data(), normals.data(),
                            masses.data(), lo.data(), hi.data(), alpha, 0, &dev, &ms),
            "tree build");
    }
    return dev;
  }

  static void check(int status, const char* what) {
    if (status == SBH_OK) return;
    std::string msg = std::string(what) + ": " + sbh_last_error();
    if (status == SBH_ERR_INVALID) throw py::value_error(msg);
    if (status == SBH_ERR_NOMEM) throw std::bad_alloc();
    throw std::runtime_error(msg);
  }

  sbh_tree* dev = nullptr;
  std::mutex mu;
};

template <int D, int K>
struct HostTree : HostTreeBase {
  HostTree(F32Array pts, F32Array nrm, F32Array m, F32Array mn, F32Array mx, float a, int leaf)
      : HostTreeBase(D, K, pts, nrm, m, mn, mx, a, leaf) {}
};

template <int D, int K>
void bind_tree(py::module_& m, const char* name) {
  py::class_<HostTree<D, K>, HostTreeBase>(m, name)
      .def(py::init<F32Array, F32Array, F32Array, F32Array, F32Array, float, int>(), "points"_a,
           "normals"_a, "masses"_a, "min_corner"_a, "max_corner"_a, "alpha"_a = 1.0f,
           "leaf_size"_a = 1);
}

// Queries (q, dim) -> SoA columns [x..., y..., (z...)].
std::vector<float> soa(const HostTreeBase& t, const F32Array& queries) {
  if (queries.ndim() != 2 || queries.shape(1) != t.dim)
    throw py::value_error("queries must have shape (q, " + std::to_string(t.dim) + ")");
  const int64_t nq = queries.shape(0);
  std::vector<float> q((size_t)nq * t.dim);
  const float* src = queries.data();
  for (int64_t i = 0; i < nq; ++i)
    for (int d = 0; d < t.dim; ++d) q[(size_t)d * nq + i] = src[i * t.dim + d];
  return q;
}

py::object evaluate(HostTreeBase& t, int method, int kernel, const F32Array& queries, float beta,
                    int spl, uint64_t seed, bool exact, bool with_counts) {
  std::vector<float> q = soa(t, queries);
  const int64_t nq = queries.shape(0);
  py::array_t<float> out(nq);
  py::array_t<int32_t> cnt(with_counts ? nq : 0);
  sbh_eval_params p{};
  p.beta = beta;
  p.samples_per_subdomain = spl;
  p.seed = seed;
  p.vote_group = 1;
  p.flags = SBH_EVAL_SORT_QUERIES | (exact ? SBH_EVAL_EXACT_TERMS : 0) |
            (method == SBH_STOCHASTIC ? SBH_EVAL_PER_QUERY_PATHS : 0);
  float* op = out.mutable_data();
  int32_t* cp = with_counts ? cnt.mutable_data() : nullptr;
  {
    py::gil_scoped_release nogil;
    sbh_tree* h = t.device();
    float ms = 0.0f;
    HostTreeBase::check(sbh_eval(h, method, kernel, &p, q.data(), nq, op, cp, &ms), "evaluation");
  }
  if (with_counts) return py::make_tuple(out, cnt);
  return out;
}

void bind_eval(py::module_& m, const char* suffix, int kernel) {
  const std::string s = suffix;
  m.def(("eval_barnes_hut_" + s).c_str(),
        [kernel](HostTreeBase& t, F32Array q, float beta, bool exact_terms, bool return_counts) {
          return evaluate(t, SBH_BARNES_HUT, kernel, q, beta, 1, 0, exact_terms, return_counts);
        },
        "tree"_a, "queries"_a, "beta"_a, "exact_terms"_a = true, "return_counts"_a = false);
  m.def(("multi_level_prefix_control_variate_" + s).c_str(),
        [kernel](HostTreeBase& t, F32Array q, int spl, uint64_t seed, bool exact_terms) {
          return evaluate(t, SBH_STOCHASTIC, kernel, q, 0.0f, spl, seed, exact_terms, true);
        },
        "tree"_a, "queries"_a, "samples_per_subdomain"_a, "seed"_a = 0, "exact_terms"_a = true);
}

}  // namespace

PYBIND11_MODULE(sbh, m) {
  m.doc() = "Stochastic Barnes-Hut CPU plugin API (host trees; evaluation runs on the GPU "
            "with the CPU plugin's per-query semantics)";
  py::class_<HostTreeBase>(m, "_HostTree")
      .def_readonly("dim", &HostTreeBase::dim)
      .def_readonly("split_log", &HostTreeBase::split_log)
      .def_readonly("leaf_size", &HostTreeBase::leaf_size)
      .def_readonly("num_points", &HostTreeBase::n)
      .def_readonly("alpha", &HostTreeBase::alpha)
      .def_property_readonly("points",
                             [](const HostTreeBase& t) { return t.arr(t.points, {t.n, t.dim}); })
      .def_property_readonly("normals",
                             [](const HostTreeBase& t) { return t.arr(t.normals, {t.n, t.dim}); })
      .def_property_readonly("masses", [](const HostTreeBase& t) { return t.arr(t.masses, {t.n}); })
      .def_property_readonly("min_corner", [](const HostTreeBase& t) { return t.arr(t.lo, {t.dim}); })
      .def_property_readonly("max_corner", [](const HostTreeBase& t) { return t.arr(t.hi, {t.dim}); });
  bind_tree<2, 1>(m, "Quadtree1");
  bind_tree<2, 2>(m, "Quadtree2");
  bind_tree<2, 3>(m, "Quadtree3");
  bind_tree<3, 1>(m, "Octree1");
  bind_tree<3, 2>(m, "Octree2");
  bind_tree<3, 3>(m, "Octree3");
  m.attr("fp") = 32;  // src/sbh.cpp:214

  bind_eval(m, "gravity", SBH_GRAVITY);
  bind_eval(m, "wn", SBH_WINDING_NUMBER);
  bind_eval(m, "smoothdist", SBH_SMOOTH_DIST);
}
