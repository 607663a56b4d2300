// pybind11 module `sbh_gpu64`: the reference's GPU plugin API
// (src/sbh_gpu.cu:566-661) as a USE_DOUBLE build of the reference would
// expose it (include/utils.h:13-29: FLOAT = double), over the double-build
// entry points of libsbh_hip.so (sbh_tree_create_double, sbh_eval_double).
//
//   QuadtreeGPU1/2/3, OctreeGPU1/2/3 (cpu_tree | arrays)  src/sbh_gpu.cu:571-582
//   eval_brute_force_{kernel}                              src/sbh_gpu.cu:584-604
//   eval_barnes_hut_{kernel}                               src/sbh_gpu.cu:606-623
//   multi_level_prefix_control_variate_{kernel}            src/sbh_gpu.cu:625-660
// for the bound kernels (gravity, wn, smoothdist) and the unbound ones of
// include/kernel.h:61-83.  Inputs are taken as float64 and every evaluation
// returns (float64[q], kernel milliseconds).  Keyword extensions as in
// `sbh_gpu`: vote_group (32), seed (0), sort_queries (False), return_counts,
// per_query_paths, smooth_distance.  The matrix-core forms are float-build
// options and are not offered here.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sbh_hip.h"

namespace py = pybind11;
using namespace py::literals;

namespace {

using F64Array = py::array_t<double, py::array::c_style | py::array::forcecast>;

void check(int status, const char* what) {
  if (status == SBH_OK) return;
  std::string msg = std::string(what) + ": " + sbh_last_error();
  switch (status) {
    case SBH_ERR_INVALID:
      throw py::value_error(msg);
    case SBH_ERR_NOMEM:
      throw std::bad_alloc();
    default:
      throw std::runtime_error(msg);
  }
}

struct TreeBase {
  sbh_tree* h = nullptr;
  int dim = 3;
  int split_log = 1;
  float build_ms = 0.0f;

  TreeBase(int d, int k, F64Array pts, F64Array nrm, F64Array m, F64Array mn, F64Array mx,
           double alpha, int leaf_size, int max_depth)
      : dim(d), split_log(k) {
    if (pts.ndim() != 2 || pts.shape(1) != d)
      throw py::value_error("points must have shape (n, " + std::to_string(d) + ")");
    const int64_t n = pts.shape(0);
    const double* nptr = nullptr;
    if (nrm.size() > 0) {
      if (nrm.ndim() != 2 || nrm.shape(0) != n || nrm.shape(1) != d)
        throw py::value_error("normals must have the shape of points");
      nptr = nrm.data();
    }
    if (m.ndim() != 1 || m.shape(0) != n) throw py::value_error("masses must have shape (n,)");
    if (mn.size() != d || mx.size() != d)
      throw py::value_error("min_corner / max_corner must have " + std::to_string(d) + " values");
    int status;
    {
      py::gil_scoped_release nogil;
      status = sbh_tree_create_double(d, k, leaf_size, max_depth, n, pts.data(), nptr, m.data(),
                                      mn.data(), mx.data(), alpha, &h, &build_ms);
    }
    check(status, "tree build");
  }
  ~TreeBase() {
    if (h) sbh_tree_destroy(h);
  }
  TreeBase(const TreeBase&) = delete;
  TreeBase& operator=(const TreeBase&) = delete;

  py::dict info() const {
    int64_t np = 0, nn = 0, bytes = 0;
    int32_t depth = 0;
    check(sbh_tree_info(h, &np, &nn, &depth, &bytes), "tree info");
    py::dict d;
    d["num_points"] = np;
    d["num_nodes"] = nn;
    d["depth"] = depth;
    d["device_bytes"] = bytes;
    d["build_ms"] = build_ms;
    d["dim"] = dim;
    d["split_log"] = split_log;
    return d;
  }

  py::array_t<int64_t> leaf_order() const {
    int64_t np = 0;
    check(sbh_tree_info(h, &np, nullptr, nullptr, nullptr), "tree info");
    py::array_t<int64_t> out(np);
    check(sbh_tree_leaf_order(h, out.mutable_data()), "leaf order");
    return out;
  }

  py::dict nodes() const {
    int64_t nn = 0;
    check(sbh_tree_info(h, nullptr, &nn, nullptr, nullptr), "tree info");
    auto u32 = [&] { return py::array_t<uint32_t>(nn); };
    py::array_t<uint32_t> start = u32(), end = u32(), first = u32(), nch = u32(), parent = u32(),
                          rope = u32();
    py::array_t<int32_t> depth(nn);
    const int64_t d = dim;
    py::array_t<double> bmin({nn, d}), bmax({nn, d}), diam(nn), centroid({nn, d}), mass(nn),
        normal({nn, d});
    check(sbh_tree_nodes(h, start.mutable_data(), end.mutable_data(), first.mutable_data(),
                         nch.mutable_data(), depth.mutable_data(), parent.mutable_data(),
                         rope.mutable_data(), nullptr, nullptr, nullptr, nullptr, nullptr,
                         nullptr, nullptr),
          "tree nodes");
    check(sbh_tree_nodes_double(h, bmin.mutable_data(), bmax.mutable_data(), diam.mutable_data(),
                                centroid.mutable_data(), mass.mutable_data(),
                                normal.mutable_data()),
          "tree nodes");
    py::dict r;
    r["start"] = start;
    r["end"] = end;
    r["first_child"] = first;
    r["nchild"] = nch;
    r["depth"] = depth;
    r["parent"] = parent;
    r["rope"] = rope;
    r["bmin"] = bmin;
    r["bmax"] = bmax;
    r["diam"] = diam;
    r["centroid"] = centroid;
    r["mass"] = mass;
    r["normal"] = normal;
    return r;
  }

  // (q, dim) row-major -> column-major [x..., y..., z...]
  std::vector<double> soa(const F64Array& q) const {
    if (q.ndim() != 2 || q.shape(1) != dim)
      throw py::value_error("queries must have shape (q, " + std::to_string(dim) + ")");
    const int64_t nq = q.shape(0);
    std::vector<double> v((size_t)nq * dim);
    const double* p = q.data();
    for (int64_t i = 0; i < nq; ++i)
      for (int a = 0; a < dim; ++a) v[(size_t)a * nq + i] = p[i * dim + a];
    return v;
  }
};

template <int D, int K>
struct Tree : TreeBase {
  using TreeBase::TreeBase;
};

F64Array attr_array(const py::object& o, const char* name) {
  return F64Array::ensure(o.attr(name));
}

template <int D, int K>
void bind_tree(py::module_& m, const char* name) {
  py::class_<Tree<D, K>, TreeBase>(m, name)
      .def(py::init([](const py::object& cpu) {
             int d = cpu.attr("dim").cast<int>();
             int k = cpu.attr("split_log").cast<int>();
             if (d != D || k != K)
               throw py::type_error("host tree has dim " + std::to_string(d) + " / split " +
                                    std::to_string(k) + ", expected " + std::to_string(D) +
                                    " / " + std::to_string(K));
             int leaf = py::hasattr(cpu, "leaf_size") ? cpu.attr("leaf_size").cast<int>() : 1;
             return std::make_unique<Tree<D, K>>(
                 D, K, attr_array(cpu, "points"), attr_array(cpu, "normals"),
                 attr_array(cpu, "masses"), attr_array(cpu, "min_corner"),
                 attr_array(cpu, "max_corner"), cpu.attr("alpha").cast<double>(), leaf, 100);
           }),
           "cpu_tree"_a)
      .def(py::init([](F64Array pts, F64Array nrm, F64Array mass, F64Array mn, F64Array mx,
                       double alpha, int leaf_size, int max_depth) {
             return std::make_unique<Tree<D, K>>(D, K, pts, nrm, mass, mn, mx, alpha, leaf_size,
                                                 max_depth);
           }),
           "points"_a, "normals"_a, "masses"_a, "min_corner"_a, "max_corner"_a, "alpha"_a = 1.0,
           "leaf_size"_a = 1, "max_depth"_a = 100);
}

py::object run_eval(TreeBase& t, int method, int kernel, const F64Array& queries, double beta,
                    int spl, uint64_t seed, int vote_group, bool counts, int flags) {
  std::vector<double> q = t.soa(queries);
  const int64_t nq = queries.shape(0);
  py::array_t<double> out(nq);
  py::array_t<int32_t> cnt(counts ? nq : 0);
  sbh_eval_params_double p{};
  p.beta = beta;
  p.samples_per_subdomain = spl;
  p.seed = seed;
  p.vote_group = vote_group;
  p.flags = flags;
  float ms = 0.0f;
  double* op = out.mutable_data();
  int32_t* cp = counts ? cnt.mutable_data() : nullptr;
  int status;
  {
    py::gil_scoped_release nogil;
    status = sbh_eval_double(t.h, method, kernel, &p, q.data(), nq, op, cp, &ms);
  }
  check(status, "evaluation");
  if (counts) return py::make_tuple(out, ms, cnt);
  return py::make_tuple(out, ms);
}

void bind_kernel(py::module_& m, const char* suffix, int kernel) {
  std::string s = suffix;
  m.def(("eval_brute_force_" + s).c_str(),
        [kernel](TreeBase& t, F64Array q, bool smooth_distance) {
          return run_eval(t, SBH_BRUTE_FORCE, kernel, q, 0.0, 1, 0, 1, false,
                          smooth_distance ? SBH_EVAL_SMOOTH_DISTANCE : 0);
        },
        "tree"_a, "queries"_a, "smooth_distance"_a = false);
  m.def(("eval_barnes_hut_" + s).c_str(),
        [kernel](TreeBase& t, F64Array q, double beta, int vote_group, bool sort_queries,
                 bool return_counts, bool smooth_distance) {
          return run_eval(t, SBH_BARNES_HUT, kernel, q, beta, 1, 0, vote_group, return_counts,
                          (sort_queries ? SBH_EVAL_SORT_QUERIES : 0) |
                              (smooth_distance ? SBH_EVAL_SMOOTH_DISTANCE : 0));
        },
        "tree"_a, "queries"_a, "beta"_a, "vote_group"_a = 32, "sort_queries"_a = false,
        "return_counts"_a = false, "smooth_distance"_a = false);
  m.def(("multi_level_prefix_control_variate_" + s).c_str(),
        [kernel](TreeBase& t, F64Array q, int spl, uint64_t seed, bool sort_queries,
                 bool return_counts, bool per_query_paths, bool smooth_distance) {
          return run_eval(t, SBH_STOCHASTIC, kernel, q, 0.0, spl, seed, 1, return_counts,
                          (sort_queries ? SBH_EVAL_SORT_QUERIES : 0) |
                              (per_query_paths ? SBH_EVAL_PER_QUERY_PATHS : 0) |
                              (smooth_distance ? SBH_EVAL_SMOOTH_DISTANCE : 0));
        },
        "tree"_a, "queries"_a, "samples_per_subdomain"_a, "seed"_a = 0, "sort_queries"_a = false,
        "return_counts"_a = false, "per_query_paths"_a = false, "smooth_distance"_a = false);
}

}  // namespace

PYBIND11_MODULE(sbh_gpu64, m) {
  m.doc() =
      "Stochastic Barnes-Hut fast summation on AMD Instinct MI355X (gfx950), double build "
      "(float64 trees, terms and results)";
  py::class_<TreeBase>(m, "_GPUTree64")
      .def("info", &TreeBase::info)
      .def("leaf_order", &TreeBase::leaf_order)
      .def("nodes", &TreeBase::nodes)
      .def_readonly("dim", &TreeBase::dim)
      .def_readonly("split_log", &TreeBase::split_log)
      .def_readonly("build_ms", &TreeBase::build_ms);
  bind_tree<2, 1>(m, "QuadtreeGPU1");
  bind_tree<2, 2>(m, "QuadtreeGPU2");
  bind_tree<2, 3>(m, "QuadtreeGPU3");
  bind_tree<3, 1>(m, "OctreeGPU1");
  bind_tree<3, 2>(m, "OctreeGPU2");
  bind_tree<3, 3>(m, "OctreeGPU3");
  bind_kernel(m, "gravity", SBH_GRAVITY);
  bind_kernel(m, "wn", SBH_WINDING_NUMBER);
  bind_kernel(m, "smoothdist", SBH_SMOOTH_DIST);
  bind_kernel(m, "helmholtz_x", SBH_HELMHOLTZ_X);
  bind_kernel(m, "helmholtz_y", SBH_HELMHOLTZ_Y);
  bind_kernel(m, "gravity_grad_x", SBH_GRAVITY_GRAD_X);
  bind_kernel(m, "gravity_grad_y", SBH_GRAVITY_GRAD_Y);
  bind_kernel(m, "gravity_grad_z", SBH_GRAVITY_GRAD_Z);
  m.def("device_count", [] {
    int32_t n = 0;
    check(sbh_device_count(&n), "device count");
    return n;
  });
  m.attr("fp") = 64;
}
