// Host-side runtime pieces shared by the library's translation units: error
// handling, device buffers, the device tree object.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "sbh_device.hpp"

namespace sbh {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// Status codes shared with include/sbh_hip.h.
constexpr int kOk = 0;
constexpr int kErrInvalid = 1;
constexpr int kErrHip = 2;
constexpr int kErrNoMemory = 3;
constexpr int kErrNoDevice = 4;
constexpr int kErrBudget = 5;
constexpr int kErrInternal = 6;

#define SBH_HIP(expr)                                                                  \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess) {                                                            \
      throw ::sbh::Error(_e == hipErrorOutOfMemory ? ::sbh::kErrNoMemory               \
                                                   : ::sbh::kErrHip,                   \
                         std::string(#expr) + ": " + hipGetErrorString(_e));           \
    }                                                                                  \
  } while (0)

#define SBH_REQUIRE(cond, msg)                                     \
  do {                                                             \
    if (!(cond)) throw ::sbh::Error(::sbh::kErrInvalid, (msg));    \
  } while (0)

// Owning device allocation.
template <typename T>
class DevBuf {
 public:
  using value_type = T;
  DevBuf() = default;
  explicit DevBuf(size_t n) { alloc(n); }
  ~DevBuf() { release(); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_) {
    o.p_ = nullptr;
    o.n_ = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p_ = o.p_;
      n_ = o.n_;
      o.p_ = nullptr;
      o.n_ = 0;
    }
    return *this;
  }
  void alloc(size_t n) {
    release();
    if (n == 0) n = 1;
    SBH_HIP(hipMalloc((void**)&p_, n * sizeof(T)));
    n_ = n;
  }
  // Grow to at least n elements, keeping the first `keep` elements.
  void grow(size_t n, size_t keep, hipStream_t s) {
    if (n <= n_) return;
    T* np = nullptr;
    SBH_HIP(hipMalloc((void**)&np, n * sizeof(T)));
    if (p_ && keep) SBH_HIP(hipMemcpyAsync(np, p_, keep * sizeof(T), hipMemcpyDeviceToDevice, s));
    SBH_HIP(hipStreamSynchronize(s));
    release();
    p_ = np;
    n_ = n;
  }
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T* get() const { return p_; }
  size_t size() const { return n_; }
  size_t bytes() const { return n_ * sizeof(T); }

 private:
  T* p_ = nullptr;
  size_t n_ = 0;
};

// Device-resident tree.  All node arrays are indexed by the reference's
// breadth-first node number (include/tree3d.h:338-371); points are stored in
// the tree's leaf order.
struct DeviceTree {
  int dim = 3;
  int split_log = 1;
  int leaf_size = 1;
  int max_depth = 100;
  int moments64 = 0;  // node moments aggregated in float64 instead of reference fp32 order
  uint32_t n = 0;
  uint32_t num_nodes = 0;
  int depth = 0;  // deepest level
  float alpha = 1.0f;
  float lo[3] = {0, 0, 0};
  float hi[3] = {0, 0, 0};
  std::vector<uint32_t> level_begin;  // level l holds nodes [level_begin[l], level_begin[l+1])
  hipStream_t stream = nullptr;

  // points, leaf order
  DevBuf<float4> pts;   // {x, y, z, mass}
  DevBuf<float4> nrm;   // {nx, ny, nz, 0}
  DevBuf<uint32_t> perm;  // input row of each sorted point

  // nodes (breadth-first numbering)
  DevBuf<uint32_t> n_start, n_end, n_parent, n_first, n_nchild, n_next;
  DevBuf<int32_t> n_depth;
  DevBuf<float4> n_lo, n_hi;  // cell bounds (w unused)
  DevBuf<float> n_diam;
  DevBuf<float4> nodeC;  // {cx, cy, cz, mass}
  DevBuf<float4> nodeN;  // {nx, ny, nz, 0}
  DevBuf<uint4> nodeL;   // {diam bits, first child | first point, rope, info}
  DevBuf<double> n_m64;  // float64 moments {cx, cy, cz, mass, nx, ny, nz, 0} per node
  // stochastic descent helpers
  DevBuf<uint32_t> n_skip;  // node reached by following single-child chains
  // Depth-first copy of the walk records for Barnes-Hut: node v's record sits
  // at dfs position n_dfs[v]; its first child is the next record and its rope
  // is the position after its subtree, so a walk reads forward in memory.
  DevBuf<uint32_t> n_size;  // nodes in the subtree
  DevBuf<uint32_t> n_dfs;
  DevBuf<float4> dfsC, dfsN;
  DevBuf<uint4> dfsL;
  // Term source of each node as a child on a sampled path: a one-point
  // leaf's point and normal, any other node's moment (derived from the
  // arrays above; not part of a tree image).
  DevBuf<float4> kidC, kidN;
  // The depth-first walk records (dfsC, dfsL) interleaved, 32 bytes a node,
  // so that a walk step is one scalar load (derived, not in a tree image).
  DevBuf<WalkRecord> dfsR;
  // Breadth-first records for the sampler's steps: {nodeC, nodeL} together,
  // and {range start, range end, single-child chain target, 0} (derived).
  DevBuf<WalkRecord> bfsR;
  DevBuf<uint4> stepR;
  // Split frontiers of the Barnes-Hut walk (derived), shallowest first: the
  // frontier at depth L holds the nodes at depth L and the leaves above them
  // (2 to kMaxWalkTasks subtrees), as the tasks
  // walk_tasks[walk_frontiers[f].first, + .second).
  DevBuf<WalkTask> walk_tasks;
  std::vector<std::pair<uint32_t, uint32_t>> walk_frontiers;
  // Tasks of the split sorted estimator (derived): the root's subdomains in
  // ranges {first, end, part, 0} of at most one internal subdomain each (leaf
  // and empty subdomains join the range before them, or the first range),
  // listed heaviest first (points of the internal subdomain), the order a
  // split launch dispatches them in; `part` is the range's position in
  // subdomain order, the slot of its partial sums.  n_sub_tasks is 0 unless
  // the root has 2 to 64 subdomains, two or more of them internal.
  DevBuf<uint4> sub_tasks;
  uint32_t n_sub_tasks = 0;
  // Point counts of a grid of dens_g^dim cells over the root box (derived):
  // the walks' local density estimate (lanes_spread_out).
  DevBuf<uint32_t> dens;
  uint32_t dens_g = 0;
  float dens_inv[3] = {0, 0, 0};
  float dens_diag = 0.0f;

  // Double build (the reference's USE_DOUBLE configuration): points, cell
  // bounds, diameters and moments in float64 instead of the float arrays
  // above; the index arrays (ranges, children, ropes, n_skip, perm) and the
  // float64 moments n_m64 are shared.  A double tree has no fp32 point, node,
  // depth-first or derived records of its own.
  bool is_double = false;
  double alpha64 = 1.0;
  double lo64[3] = {0, 0, 0};
  double hi64[3] = {0, 0, 0};
  DevBuf<dvec4> pts64, nrm64;      // {x, y, z, mass}, {nx, ny, nz, 0}
  DevBuf<dvec4> n_lo64, n_hi64;    // cell bounds
  DevBuf<double> n_diam64;
  DevBuf<dvec4> nodeC64, nodeN64;  // {cx, cy, cz, mass}, {nx, ny, nz, 0}
  // Depth-first walk records of the double build (Barnes-Hut), and the
  // normal of each record's term source.
  DevBuf<WalkRecordD> dfsR64;
  DevBuf<dvec4> dfsN64;

  size_t device_bytes() const {
    return pts.bytes() + nrm.bytes() + perm.bytes() + n_start.bytes() + n_end.bytes() +
           n_parent.bytes() + n_first.bytes() + n_nchild.bytes() + n_next.bytes() +
           n_depth.bytes() + n_lo.bytes() + n_hi.bytes() + n_diam.bytes() + nodeC.bytes() +
           nodeN.bytes() + nodeL.bytes() + n_m64.bytes() + n_skip.bytes() + n_size.bytes() +
           n_dfs.bytes() + dfsC.bytes() + dfsN.bytes() + dfsL.bytes() + kidC.bytes() +
           kidN.bytes() + dfsR.bytes() + bfsR.bytes() + stepR.bytes() + walk_tasks.bytes() +
           sub_tasks.bytes() + dens.bytes() +
           pts64.bytes() + nrm64.bytes() + n_lo64.bytes() + n_hi64.bytes() +
           n_diam64.bytes() + nodeC64.bytes() + nodeN64.bytes() + dfsR64.bytes() +
           dfsN64.bytes();
  }
};

// Builds the tree on the device from input rows (row-major n x dim).
// Host pointers; blocks until the build is complete.
void build_device_tree(DeviceTree& t, const float* points, const float* normals,
                       const float* masses, float* build_ms);
// The double build of the same tree (t.is_double set by the caller).
void build_device_tree_double(DeviceTree& t, const double* points, const double* normals,
                              const double* masses, float* build_ms);

// Byte image of a built tree (tree_io.hip), and the tree back from one: the
// image is checked (checksum, sizes, and every index the walks follow) before
// any device memory is written; the loader creates t.stream if it is unset.
void serialize_tree(const DeviceTree& t, std::vector<uint8_t>& out);
void deserialize_tree(DeviceTree& t, const uint8_t* data, size_t size);
// Fills kidC / kidN, dfsR, bfsR and stepR from the node, point and
// depth-first arrays (after a build or a load), then the walk frontiers.
void derive_child_sources(DeviceTree& t);
// The split frontiers of the Barnes-Hut walk (walk_tasks, walk_frontiers),
// from host copies of the top levels; blocks on t.stream.
void derive_walk_tasks(DeviceTree& t);
// The split sorted estimator's tasks (sub_tasks), from host copies of the
// root's subdomains; blocks on t.stream.
void derive_subdomain_tasks(DeviceTree& t);
// The point-count grid over the root box (dens and its geometry).
void derive_density_grid(DeviceTree& t);

}  // namespace sbh
