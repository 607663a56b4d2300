// pybind11 module `sbh_gpu`: the reference's GPU plugin API
// (src/sbh_gpu.cu:566-661) over the C ABI of libsbh_hip.so.
//
//   QuadtreeGPU1/2/3, OctreeGPU1/2/3 (cpu_tree)       src/sbh_gpu.cu:571-582
//   eval_brute_force_{gravity,wn,smoothdist}           src/sbh_gpu.cu:584-604
//   eval_barnes_hut_{gravity,wn,smoothdist}            src/sbh_gpu.cu:606-623
//   multi_level_prefix_control_variate_{...}           src/sbh_gpu.cu:625-660
// and the kernels the reference defines without binding them
// (include/kernel.h:61-83): *_helmholtz_x / _y, *_gravity_grad_x / _y / _z,
// and *_gravity_grad returning (q, dim) gradients.
// Each evaluation returns (float32[q], kernel milliseconds), as the reference.
// Keyword extensions keep the reference defaults: vote_group (32, the
// reference's warp vote), seed (0), sort_queries (False), return_counts,
// per_query_paths (False: the 64 queries of a wavefront share sample paths),
// smooth_distance (False: return the sum w; True: -log(w) / alpha),
// exact_terms (False: hardware sqrt / reciprocal / exp2 as the reference's
// fast-math GPU build; True: correctly rounded, bit-reproducible terms).
// The GIL is released around device work.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>

#include <cmath>
#include <limits>
#include <stdexcept>
#include <string>
#include <memory>
#include <unordered_map>
#include <vector>

#include "../../include/sbh_hip.h"

namespace py = pybind11;
using namespace py::literals;

namespace {

using F32Array = py::array_t<float, py::array::c_style | py::array::forcecast>;

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

struct GPUTreeBase {
  sbh_tree* h = nullptr;
  int dim = 3;
  int split_log = 1;
  float build_ms = 0.0f;

  GPUTreeBase(int d, int k, F32Array pts, F32Array nrm, F32Array m, F32Array mn, F32Array mx,
              float alpha, int leaf_size, bool moments64, int max_depth) {
    dim = d;
    split_log = k;
    if (pts.ndim() != 2 || pts.shape(1) != d)
      throw py::value_error("points must have shape (n, " + std::to_string(d) + ")");
    const int64_t n = pts.shape(0);
    const float* nptr = nullptr;
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
      status = sbh_tree_create(d, k, leaf_size, max_depth, n, pts.data(), nptr, m.data(),
                               mn.data(), mx.data(), alpha, moments64 ? SBH_TREE_MOMENTS_F64 : 0,
                               &h, &build_ms);
    }
    check(status, "tree build");
  }
  // Adopts a tree handle (a loaded image).
  GPUTreeBase(sbh_tree* handle, int d, int k) : h(handle), dim(d), split_log(k) {}
  ~GPUTreeBase() {
    if (h) sbh_tree_destroy(h);
  }

  void save(const std::string& path) const {
    int status;
    {
      py::gil_scoped_release nogil;
      status = sbh_tree_save(h, path.c_str());
    }
    check(status, "tree save");
  }

  py::bytes to_bytes() const {
    std::string img;
    int status;
    {
      py::gil_scoped_release nogil;
      int64_t size = 0;
      status = sbh_tree_serialize(h, nullptr, 0, &size);
      if (status == SBH_OK) {
        img.resize((size_t)size);
        status = sbh_tree_serialize(h, img.data(), size, &size);
      }
    }
    check(status, "tree serialize");
    return py::bytes(img);
  }
  GPUTreeBase(const GPUTreeBase&) = delete;
  GPUTreeBase& operator=(const GPUTreeBase&) = delete;

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
    py::array_t<float> bmin({nn, (int64_t)dim}), bmax({nn, (int64_t)dim}), diam(nn),
        centroid({nn, (int64_t)dim}), mass(nn), normal({nn, (int64_t)dim});
    py::array_t<double> m64({nn, (int64_t)8});
    check(sbh_tree_nodes(h, start.mutable_data(), end.mutable_data(), first.mutable_data(),
                         nch.mutable_data(), depth.mutable_data(), parent.mutable_data(),
                         rope.mutable_data(), bmin.mutable_data(), bmax.mutable_data(),
                         diam.mutable_data(), centroid.mutable_data(), mass.mutable_data(),
                         normal.mutable_data(), m64.mutable_data()),
          "tree nodes");
    py::dict d;
    d["start"] = start;
    d["end"] = end;
    d["first_child"] = first;
    d["nchild"] = nch;
    d["depth"] = depth;
    d["parent"] = parent;
    d["rope"] = rope;
    d["bmin"] = bmin;
    d["bmax"] = bmax;
    d["diam"] = diam;
    d["centroid"] = centroid;
    d["mass"] = mass;
    d["normal"] = normal;
    d["moments64"] = m64;
    return d;
  }

  // (q, dim) row-major -> column-major [x..., y..., z...]
  std::vector<float> soa(const F32Array& q) const {
    if (q.ndim() != 2 || q.shape(1) != dim)
      throw py::value_error("queries must have shape (q, " + std::to_string(dim) + ")");
    const int64_t nq = q.shape(0);
    std::vector<float> v((size_t)nq * dim);
    const float* p = q.data();
    for (int64_t i = 0; i < nq; ++i)
      for (int a = 0; a < dim; ++a) v[(size_t)a * nq + i] = p[i * dim + a];
    return v;
  }
};

template <int D, int K>
struct GPUTree : GPUTreeBase {
  using GPUTreeBase::GPUTreeBase;
};

F32Array attr_array(const py::object& o, const char* name) {
  return F32Array::ensure(o.attr(name));
}

// Device tree from a host tree of the sbh module (or any object with the same
// attributes): include/py_tree_gpu.h:39-80.
template <int D, int K>
std::unique_ptr<GPUTree<D, K>> from_host(const py::object& cpu, bool moments64) {
  int d = cpu.attr("dim").cast<int>();
  int k = cpu.attr("split_log").cast<int>();
  if (d != D || k != K)
    throw py::type_error("host tree has dim " + std::to_string(d) + " / split " +
                         std::to_string(k) + ", expected " + std::to_string(D) + " / " +
                         std::to_string(K));
  int leaf = py::hasattr(cpu, "leaf_size") ? cpu.attr("leaf_size").cast<int>() : 1;
  return std::make_unique<GPUTree<D, K>>(
      D, K, attr_array(cpu, "points"), attr_array(cpu, "normals"), attr_array(cpu, "masses"),
      attr_array(cpu, "min_corner"), attr_array(cpu, "max_corner"),
      cpu.attr("alpha").cast<float>(), leaf, moments64, 100);
}

// Wraps a loaded handle in the class of its shape (OctreeGPUk / QuadtreeGPUk).
py::object adopt(sbh_tree* h) {
  int32_t d = 0, k = 0;
  const int status = sbh_tree_layout(h, &d, &k, nullptr, nullptr);
  if (status != SBH_OK) {
    sbh_tree_destroy(h);
    check(status, "tree layout");
  }
  switch (d * 10 + k) {
    case 21: return py::cast(std::make_unique<GPUTree<2, 1>>(h, 2, 1));
    case 22: return py::cast(std::make_unique<GPUTree<2, 2>>(h, 2, 2));
    case 23: return py::cast(std::make_unique<GPUTree<2, 3>>(h, 2, 3));
    case 31: return py::cast(std::make_unique<GPUTree<3, 1>>(h, 3, 1));
    case 32: return py::cast(std::make_unique<GPUTree<3, 2>>(h, 3, 2));
    case 33: return py::cast(std::make_unique<GPUTree<3, 3>>(h, 3, 3));
    default:
      sbh_tree_destroy(h);
      throw std::runtime_error("loaded tree has an unknown shape");
  }
}

template <int D, int K>
void bind_tree(py::module_& m, const char* name) {
  py::class_<GPUTree<D, K>, GPUTreeBase>(m, name)
      .def(py::init([](const py::object& cpu, bool moments64) {
             return from_host<D, K>(cpu, moments64);
           }),
           "cpu_tree"_a, "moments64"_a = false)
      .def(py::init([](F32Array pts, F32Array nrm, F32Array mass, F32Array mn, F32Array mx,
                       float alpha, int leaf_size, bool moments64, int max_depth) {
             return std::make_unique<GPUTree<D, K>>(D, K, pts, nrm, mass, mn, mx, alpha,
                                                    leaf_size, moments64, max_depth);
           }),
           "points"_a, "normals"_a, "masses"_a, "min_corner"_a, "max_corner"_a,
           "alpha"_a = 1.0f, "leaf_size"_a = 1, "moments64"_a = false, "max_depth"_a = 100);
}

py::object run_eval(GPUTreeBase& t, int method, int kernel, F32Array queries, float beta,
                    int spl, uint64_t seed, int vote_group, bool sort, bool counts,
                    bool exact = false, int extra_flags = 0) {
  std::vector<float> q = t.soa(queries);
  const int64_t nq = queries.shape(0);
  py::array_t<float> out(nq);
  py::array_t<int32_t> cnt(counts ? nq : 0);
  sbh_eval_params p{};
  p.beta = beta;
  p.samples_per_subdomain = spl;
  p.seed = seed;
  p.vote_group = vote_group;
  p.flags = (sort ? SBH_EVAL_SORT_QUERIES : 0) | (exact ? SBH_EVAL_EXACT_TERMS : 0) |
            extra_flags;
  float ms = 0.0f;
  float* op = out.mutable_data();
  int32_t* cp = counts ? cnt.mutable_data() : nullptr;
  int status;
  {
    py::gil_scoped_release nogil;
    status = sbh_eval(t.h, method, kernel, &p, q.data(), nq, op, cp, &ms);
  }
  check(status, "evaluation");
  if (counts) return py::make_tuple(out, ms, cnt);
  return py::make_tuple(out, ms);
}

void bind_kernel(py::module_& m, const char* suffix, int kernel) {
  std::string s = suffix;
  m.def(("eval_brute_force_" + s).c_str(),
        [kernel](GPUTreeBase& t, F32Array q, py::object mfma, bool smooth_distance) {
          // mfma: False, True / "f32" (fp32 operands) or "bf16" (split bf16 operands)
          int flags = smooth_distance ? SBH_EVAL_SMOOTH_DISTANCE : 0;
          if (py::isinstance<py::str>(mfma)) {
            const std::string m = mfma.cast<std::string>();
            if (m == "bf16")
              flags |= SBH_EVAL_MFMA_BF16;
            else if (m == "f32")
              flags |= SBH_EVAL_MFMA;
            else
              throw py::value_error("mfma must be False, True, 'f32' or 'bf16'");
          } else if (mfma.cast<bool>()) {
            flags |= SBH_EVAL_MFMA;
          }
          return run_eval(t, SBH_BRUTE_FORCE, kernel, q, 0.f, 1, 0, 1, false, false, false,
                          flags);
        },
        "tree"_a, "queries"_a, "mfma"_a = false, "smooth_distance"_a = false);
  m.def(("eval_barnes_hut_" + s).c_str(),
        [kernel](GPUTreeBase& t, F32Array q, float beta, int vote_group, bool sort_queries,
                 bool return_counts, bool exact_terms, bool smooth_distance, bool no_split) {
          return run_eval(t, SBH_BARNES_HUT, kernel, q, beta, 1, 0, vote_group, sort_queries,
                          return_counts, exact_terms,
                          (smooth_distance ? SBH_EVAL_SMOOTH_DISTANCE : 0) |
                              (no_split ? SBH_EVAL_NO_SPLIT : 0));
        },
        "tree"_a, "queries"_a, "beta"_a, "vote_group"_a = 32, "sort_queries"_a = false,
        "return_counts"_a = false, "exact_terms"_a = false, "smooth_distance"_a = false,
        "no_split"_a = false);
  m.def(("multi_level_prefix_control_variate_" + s).c_str(),
        [kernel](GPUTreeBase& t, F32Array q, int spl, uint64_t seed, bool sort_queries,
                 bool return_counts, bool exact_terms, bool per_query_paths,
                 bool smooth_distance, bool no_split) {
          return run_eval(t, SBH_STOCHASTIC, kernel, q, 0.f, spl, seed, 1, sort_queries,
                          return_counts, exact_terms,
                          (per_query_paths ? SBH_EVAL_PER_QUERY_PATHS : 0) |
                              (smooth_distance ? SBH_EVAL_SMOOTH_DISTANCE : 0) |
                              (no_split ? SBH_EVAL_NO_SPLIT : 0));
        },
        "tree"_a, "queries"_a, "samples_per_subdomain"_a, "seed"_a = 0, "sort_queries"_a = false,
        "return_counts"_a = false, "exact_terms"_a = false, "per_query_paths"_a = false,
        "smooth_distance"_a = false, "no_split"_a = false);
  m.def(("eval_brute_force_f64_" + s).c_str(),
        [kernel](GPUTreeBase& t, F32Array queries) {
          std::vector<float> q = t.soa(queries);
          const int64_t nq = queries.shape(0);
          py::array_t<double> out(nq);
          float ms = 0.0f;
          double* op = out.mutable_data();
          int status;
          {
            py::gil_scoped_release nogil;
            status = sbh_eval_f64(t.h, kernel, q.data(), nq, op, 0, &ms);
          }
          check(status, "float64 evaluation");
          return py::make_tuple(out, ms);
        },
        "tree"_a, "queries"_a);
}

// gravityGrad (include/kernel.h:61-66) as (q, dim) arrays: one evaluation
// per component kernel with the same parameters, so the stochastic components
// share their sample paths.
py::tuple gradient(GPUTreeBase& t, int method, F32Array q, float beta, int spl, uint64_t seed,
                   int vote_group, bool sort, bool exact, int extra) {
  const int64_t nq = q.ndim() == 2 ? q.shape(0) : 0;
  py::array_t<float> out({nq, (int64_t)t.dim});
  float total_ms = 0.0f;
  for (int a = 0; a < t.dim; ++a) {
    py::tuple r = run_eval(t, method, SBH_GRAVITY_GRAD_X + a, q, beta, spl, seed, vote_group,
                           sort, false, exact, extra);
    auto col = r[0].cast<py::array_t<float>>();
    const float* c = col.data();
    float* o = out.mutable_data();
    for (int64_t i = 0; i < nq; ++i) o[i * t.dim + a] = c[i];
    total_ms += r[1].cast<float>();
  }
  return py::make_tuple(out, total_ms);
}

void bind_gradient(py::module_& m) {
  m.def("eval_brute_force_gravity_grad",
        [](GPUTreeBase& t, F32Array q) {
          return gradient(t, SBH_BRUTE_FORCE, q, 0.f, 1, 0, 1, false, false, 0);
        },
        "tree"_a, "queries"_a);
  m.def("eval_barnes_hut_gravity_grad",
        [](GPUTreeBase& t, F32Array q, float beta, int vote_group, bool sort_queries,
           bool exact_terms) {
          return gradient(t, SBH_BARNES_HUT, q, beta, 1, 0, vote_group, sort_queries,
                          exact_terms, 0);
        },
        "tree"_a, "queries"_a, "beta"_a, "vote_group"_a = 32, "sort_queries"_a = false,
        "exact_terms"_a = false);
  m.def("multi_level_prefix_control_variate_gravity_grad",
        [](GPUTreeBase& t, F32Array q, int spl, uint64_t seed, bool sort_queries,
           bool exact_terms, bool per_query_paths) {
          return gradient(t, SBH_STOCHASTIC, q, 0.f, spl, seed, 1, sort_queries, exact_terms,
                          per_query_paths ? SBH_EVAL_PER_QUERY_PATHS : 0);
        },
        "tree"_a, "queries"_a, "samples_per_subdomain"_a, "seed"_a = 0, "sort_queries"_a = false,
        "exact_terms"_a = false, "per_query_paths"_a = false);
}

// Serving (sbh_server_* in include/sbh_hip.h).  Query and result buffers of a
// request are owned here and kept alive until the engine has finished with
// them, whether or not the caller still holds the request object.
struct RequestBufs {
  std::vector<float> queries;  // column-major dim x q
  py::array_t<float> results;
};

struct PyServer;

struct PyRequest {
  std::shared_ptr<RequestBufs> bufs;
  sbh_server* srv = nullptr;
  int64_t id = 0;
  int state() const {
    int32_t st = 0;
    check(sbh_server_poll(srv, id, &st, nullptr), "poll");
    return st;
  }
  int wait(double timeout) const {
    int32_t st = 0;
    double lat = 0;
    {
      py::gil_scoped_release nogil;
      check(sbh_server_wait(srv, id, timeout, &st, &lat), "wait");
    }
    return st;
  }
  py::tuple result(double timeout) const {
    int st = wait(timeout);
    if (st != SBH_REQ_DONE)
      throw std::runtime_error("request " + std::to_string(id) + " not done (state " +
                               std::to_string(st) + ")");
    int32_t s2 = 0;
    double lat = 0;
    check(sbh_server_poll(srv, id, &s2, &lat), "poll");
    return py::make_tuple(bufs->results, lat);
  }
};

struct PyServer {
  sbh_server* h = nullptr;
  std::unordered_map<int64_t, std::shared_ptr<RequestBufs>> live;
  std::vector<py::object> trees;  // resident trees outlive the server

  PyServer(int64_t budget, int max_inflight) {
    check(sbh_server_create(budget, max_inflight, &h), "server");
  }
  ~PyServer() {
    if (h) {
      py::gil_scoped_release nogil;
      sbh_server_destroy(h);
    }
  }
  int add_tree(py::object tree) {
    GPUTreeBase& t = tree.cast<GPUTreeBase&>();
    int32_t id = 0;
    check(sbh_server_add_tree(h, t.h, &id), "add tree");
    trees.push_back(tree);
    return id;
  }
  void sweep() {
    for (auto it = live.begin(); it != live.end();) {
      int32_t st = 0;
      if (it->second.use_count() == 1 &&
          sbh_server_poll(h, it->first, &st, nullptr) == SBH_OK && st >= SBH_REQ_DONE) {
        sbh_server_release(h, it->first);
        it = live.erase(it);
      } else {
        ++it;
      }
    }
  }
  PyRequest submit(int tree_id, const std::string& method, const std::string& kernel,
                   F32Array queries, float beta, int spl, uint64_t seed, bool no_split) {
    sweep();
    if (tree_id < 0 || tree_id >= (int)trees.size()) throw py::value_error("unknown tree id");
    GPUTreeBase& t = trees[tree_id].cast<GPUTreeBase&>();
    auto b = std::make_shared<RequestBufs>();
    b->queries = t.soa(queries);
    const int64_t nq = queries.shape(0);
    b->results = py::array_t<float>(nq);
    sbh_request r{};
    r.tree_id = tree_id;
    r.method = method == "stochastic" ? SBH_STOCHASTIC : SBH_BARNES_HUT;
    if (method != "stochastic" && method != "barnes_hut")
      throw py::value_error("method must be 'barnes_hut' or 'stochastic'");
    r.kernel = kernel == "gravity" ? SBH_GRAVITY : kernel == "wn" ? SBH_WINDING_NUMBER
                                                                   : SBH_SMOOTH_DIST;
    if (kernel != "gravity" && kernel != "wn" && kernel != "smoothdist")
      throw py::value_error("kernel must be 'gravity', 'wn' or 'smoothdist'");
    r.params.beta = beta;
    r.params.samples_per_subdomain = spl;
    r.params.seed = seed;
    r.params.vote_group = 1;
    r.params.flags = no_split ? SBH_EVAL_NO_SPLIT : 0;
    r.queries = b->queries.data();
    r.q = nq;
    r.results = b->results.mutable_data();
    int64_t id = 0;
    check(sbh_server_submit(h, &r, &id), "submit");
    live[id] = b;
    return PyRequest{b, h, id};
  }
  bool cancel(const PyRequest& r) {
    int32_t c = 0;
    check(sbh_server_cancel(h, r.id, &c), "cancel");
    return c != 0;
  }
  py::dict stats() {
    sbh_server_stats st{};
    check(sbh_server_stats_get(h, &st), "stats");
    py::dict d;
    d["completed"] = st.completed;
    d["launches"] = st.launches;
    d["queued"] = st.queued;
    d["bytes_in_use"] = st.bytes_in_use;
    d["peak_bytes"] = st.peak_bytes;
    d["budget_bytes"] = st.budget_bytes;
    d["tree_bytes"] = st.tree_bytes;
    return d;
  }
};

}  // namespace

PYBIND11_MODULE(sbh_gpu, m) {
  m.doc() = "Stochastic Barnes-Hut fast summation on AMD Instinct MI355X (gfx950)";
  py::class_<GPUTreeBase>(m, "_GPUTree")
      .def("info", &GPUTreeBase::info)
      .def("leaf_order", &GPUTreeBase::leaf_order)
      .def("nodes", &GPUTreeBase::nodes)
      .def_readonly("dim", &GPUTreeBase::dim)
      .def_readonly("split_log", &GPUTreeBase::split_log)
      .def_readonly("build_ms", &GPUTreeBase::build_ms)
      .def("save", &GPUTreeBase::save, "path"_a,
           "write the tree's image (format of sbh_tree_save) to a file")
      .def("to_bytes", &GPUTreeBase::to_bytes);
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
  bind_gradient(m);
  py::class_<PyRequest>(m, "Request")
      .def_readonly("id", &PyRequest::id)
      .def("state", &PyRequest::state)
      .def("wait", &PyRequest::wait, "timeout"_a = -1.0)
      .def("result", &PyRequest::result, "timeout"_a = -1.0,
           "(values, latency_ms) once done; latency is submit -> results in place");
  py::class_<PyServer>(m, "Server")
      .def(py::init<int64_t, int>(), "budget_bytes"_a, "max_inflight"_a = 4)
      .def("add_tree", &PyServer::add_tree, "tree"_a)
      // a Request keeps its Server alive: wait / poll never run on a destroyed engine
      .def("submit", &PyServer::submit, "tree_id"_a, "method"_a, "kernel"_a, "queries"_a,
           "beta"_a = 2.0f, "samples_per_subdomain"_a = 1, "seed"_a = 0, "no_split"_a = false,
           py::keep_alive<0, 1>())
      .def("cancel", &PyServer::cancel, "request"_a)
      .def("stats", &PyServer::stats);
  m.attr("REQ_QUEUED") = SBH_REQ_QUEUED;
  m.attr("REQ_RUNNING") = SBH_REQ_RUNNING;
  m.attr("REQ_DONE") = SBH_REQ_DONE;
  m.attr("REQ_FAILED") = SBH_REQ_FAILED;
  m.attr("REQ_CANCELLED") = SBH_REQ_CANCELLED;
  m.def(
      "load_tree",
      [](const std::string& path) {
        sbh_tree* h = nullptr;
        int status;
        {
          py::gil_scoped_release nogil;
          status = sbh_tree_load(path.c_str(), &h);
        }
        check(status, "tree load");
        return adopt(h);
      },
      "path"_a, "a tree saved with tree.save(path), resident on the current GPU");
  m.def(
      "tree_from_bytes",
      [](const py::bytes& b) {
        std::string img = b;
        sbh_tree* h = nullptr;
        int status;
        {
          py::gil_scoped_release nogil;
          status = sbh_tree_deserialize(img.data(), (int64_t)img.size(), &h);
        }
        check(status, "tree load");
        return adopt(h);
      },
      "image"_a);
  m.def("device_count", [] {
    int32_t n = 0;
    check(sbh_device_count(&n), "device count");
    return n;
  });
  m.attr("fp") = 32;
}
