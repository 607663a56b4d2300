// pybind11 module `sbh`: the host-side tree classes of the reference's CPU
// plugin (src/sbh.cpp:179-214; include/py_tree.h:21-75), kept so that
// `sbh_gpu.OctreeGPUk(sbh.Octreek(...))` works unchanged.  A host tree holds
// the fp32 inputs (numpy float64 is rounded to nearest, as the reference's
// Eigen conversion does); the tree itself is built on the GPU by sbh_gpu.
//
// The CPU evaluation entry points of the reference plugin are not part of
// this library: they raise NotImplementedError and name the sbh_gpu call.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>

#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

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

[[noreturn]] void cpu_eval_unavailable(const char* fn) {
  throw py::type_error(std::string("sbh.") + fn +
                       ": CPU evaluation is not part of this library; build the device tree "
                       "with sbh_gpu.OctreeGPUk(tree) / sbh_gpu.QuadtreeGPUk(tree) and call "
                       "sbh_gpu." + fn);
}

}  // namespace

PYBIND11_MODULE(sbh, m) {
  m.doc() = "Host-side tree classes of the Stochastic Barnes-Hut plugin (trees are built on "
            "the GPU by sbh_gpu)";
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

  for (const char* fn : {"eval_barnes_hut_gravity", "eval_barnes_hut_wn",
                         "eval_barnes_hut_smoothdist", "multi_level_prefix_control_variate_gravity",
                         "multi_level_prefix_control_variate_wn",
                         "multi_level_prefix_control_variate_smoothdist"}) {
    std::string name = fn;
    m.def(fn, [name](py::args, py::kwargs) { cpu_eval_unavailable(name.c_str()); });
  }
}
