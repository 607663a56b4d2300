// C ABI of libsbh_hip.so (declared in include/sbh_hip.h).  Validates
// arguments, owns per-tree device scratch, times the evaluation kernels with
// hipEvents on the tree's stream, and turns every C++ exception into a status
// code plus a thread-local message.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/sbh_hip.h"
#include "sbh_eval.hpp"
#include "sbh_runtime.hpp"
#include "sbh_serve.hpp"

struct sbh_server {
  std::unique_ptr<sbh::Server> s;
};

struct sbh_tree {
  sbh::DeviceTree t;
  // evaluation scratch, grown on demand
  sbh::DevBuf<float> q;
  sbh::DevBuf<float> out;
  sbh::DevBuf<int32_t> counts;
  sbh::DevBuf<double> partial;
  sbh::DevBuf<double> out64;
  sbh::DevBuf<double> qd;  // double build: queries
  sbh::DevBuf<double> memo64;  // double build: the estimator's level-1 memo
  sbh::DevBuf<int32_t> memo_cnt;
  sbh::DevBuf<uint32_t> order;
  sbh::DevBuf<uint32_t> sort_scratch;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int device = 0;  // the GPU the tree lives on
};

namespace {

thread_local std::string g_last_error;

template <typename F>
int guarded(F&& f) {
  try {
    f();
    g_last_error.clear();
    return SBH_OK;
  } catch (const sbh::Error& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return SBH_ERR_NOMEM;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return SBH_ERR_INTERNAL;
  } catch (...) {
    g_last_error = "unknown error";
    return SBH_ERR_INTERNAL;
  }
}

void require_device() {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0)
    throw sbh::Error(SBH_ERR_NODEVICE, "no GPU visible to the HIP runtime");
}

template <typename T>
void ensure(sbh::DevBuf<T>& b, size_t n) {
  if (b.size() < n) b.alloc(n);
}

template <typename T>
bool all_finite(const T* p, size_t n) {
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(p[i])) return false;
  return true;
}

void require_float_build(const sbh::DeviceTree& t, const char* what) {
  SBH_REQUIRE(!t.is_double, std::string(what) +
                                " takes a float-build tree; double-build trees are evaluated by "
                                "sbh_eval_double");
}

// Shape and input checks shared by the float and the double build.
template <typename T>
void check_tree_inputs(int32_t dim, int32_t split_log, int32_t leaf_size, int32_t max_depth,
                       int64_t n, const T* points, const T* normals, const T* masses,
                       const T* min_corner, const T* max_corner, T alpha) {
  SBH_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3");
  SBH_REQUIRE(split_log >= 1 && split_log <= 3, "split_log must be 1, 2 or 3");
  SBH_REQUIRE(leaf_size >= 1, "leaf_size must be >= 1");
  SBH_REQUIRE(max_depth >= 0 && max_depth <= 1000, "max_depth must be in [0, 1000]");
  SBH_REQUIRE(n >= 1, "a tree needs at least one point");
  SBH_REQUIRE(n < (int64_t)0x7FFFFFFF, "too many points (limit 2^31 - 1)");
  SBH_REQUIRE(points && masses && min_corner && max_corner, "NULL input array");
  SBH_REQUIRE(all_finite(points, (size_t)n * dim), "points must be finite");
  SBH_REQUIRE(!normals || all_finite(normals, (size_t)n * dim), "normals must be finite");
  for (int64_t i = 0; i < n; ++i)
    SBH_REQUIRE(std::isfinite(masses[i]) && masses[i] > (T)0,
                "masses must be finite and > 0 (node centroids divide by the mass sum)");
  for (int a = 0; a < dim; ++a)
    SBH_REQUIRE(std::isfinite(min_corner[a]) && std::isfinite(max_corner[a]) &&
                    min_corner[a] < max_corner[a],
                "min_corner must be finite and below max_corner on every axis");
  SBH_REQUIRE(std::isfinite(alpha), "alpha must be finite");
}

// Scratch of sbh_eval_noalloc: one array per GPU (each device holds its own
// copy of a __device__ variable), kept until the process ends.  Host code
// never uses the name itself (on the host it is only a stand-in for the
// symbol): the device address comes from hipGetSymbolAddress, once per GPU.
static __device__ __align__(256) unsigned char scratch[SBH_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The address of `scratch` on `device`, looked up on the first call for that
// device with the device current.
unsigned char* scratch_on(int device) {
  SBH_REQUIRE(device >= 0 && device < kMaxDevices, "device index out of range");
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int cur = 0;
    SBH_HIP(hipGetDevice(&cur));
    if (cur != device) SBH_HIP(hipSetDevice(device));
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (cur != device) SBH_HIP(hipSetDevice(cur));
    SBH_HIP(e);
    SBH_REQUIRE(ptr != nullptr, "no device address for the evaluation scratch");
    g_scratch_ptr[device] = ptr;
  }
  return static_cast<unsigned char*>(g_scratch_ptr[device]);
}

// Device scratch of one float-build evaluation: the launch order of a sorted
// launch, the sort's work space, the split launches' partial sums.  Either
// grown on demand in the tree's own buffers, or carved from a fixed region
// (sbh_eval_noalloc), which is never grown: a request beyond it fails.
class EvalScratch {
 public:
  explicit EvalScratch(sbh_tree* tree) : tree_(tree) {}
  EvalScratch(unsigned char* base, size_t bytes) : base_(base), cap_(bytes) {}
  bool fixed() const { return base_ != nullptr; }
  uint32_t* order(size_t n) {
    if (!fixed()) {
      ensure(tree_->order, n);
      return tree_->order.get();
    }
    return static_cast<uint32_t*>(carve(n * 4, "launch order"));
  }
  uint32_t* sort(size_t n) {
    if (!fixed()) {
      ensure(tree_->sort_scratch, n);
      return tree_->sort_scratch.get();
    }
    return static_cast<uint32_t*>(carve(n * 4, "query sort"));
  }
  double* partial(size_t n) {
    if (!fixed()) {
      ensure(tree_->partial, n);
      return tree_->partial.get();
    }
    return static_cast<double*>(carve(n * 8, "partial sums"));
  }

 private:
  void* carve(size_t bytes, const char* what) {
    const size_t off = (used_ + 255) & ~(size_t)255;
    if (off + bytes > cap_)
      throw sbh::Error(SBH_ERR_INVALID,
                       std::string("sbh_eval_noalloc: this evaluation needs more than the ") +
                           std::to_string(cap_) + " bytes of its scratch (at the " + what +
                           ": " + std::to_string(off + bytes) +
                           " bytes); fewer queries, SBH_EVAL_NO_SPLIT or sbh_eval fit");
    used_ = off + bytes;
    return base_ + off;
  }
  sbh_tree* tree_ = nullptr;
  unsigned char* base_ = nullptr;
  size_t cap_ = 0, used_ = 0;
};

// sbh_eval's body; with fixed scratch (sbh_eval_noalloc) it allocates no
// device memory: device pointers only, every scratch region carved.
void eval_float(sbh_tree* tree, int32_t method, int32_t kernel, const sbh_eval_params* params,
                const float* queries, int64_t q, float* results, int32_t* counts,
                float* kernel_ms, EvalScratch& sc) {
  SBH_REQUIRE(tree && params, "NULL argument");
  require_float_build(tree->t, sc.fixed() ? "sbh_eval_noalloc" : "sbh_eval");
  SBH_REQUIRE(method == SBH_BRUTE_FORCE || method == SBH_BARNES_HUT || method == SBH_STOCHASTIC,
              "unknown method");
  SBH_REQUIRE(kernel >= SBH_GRAVITY && kernel <= SBH_GRAVITY_GRAD_Z, "unknown kernel");
  SBH_REQUIRE(kernel != SBH_GRAVITY_GRAD_Z || tree->t.dim == 3,
              "the gradient's z component needs a 3D tree");
  SBH_REQUIRE(q >= 0 && q < (int64_t)0x7FFFFFFF, "query count out of range");
  if (kernel_ms) *kernel_ms = 0.0f;
  if (q == 0) return;
  SBH_REQUIRE(queries && results, "NULL queries or results");
  const sbh_eval_params& p = *params;
  SBH_REQUIRE((p.flags & ~(SBH_EVAL_DEVICE_POINTERS | SBH_EVAL_SORT_QUERIES |
                           SBH_EVAL_SPLIT_COUNTS | SBH_EVAL_EXACT_TERMS | SBH_EVAL_MFMA |
                           SBH_EVAL_PER_QUERY_PATHS | SBH_EVAL_MFMA_BF16 |
                           SBH_EVAL_SMOOTH_DISTANCE | SBH_EVAL_NO_SPLIT |
                           SBH_EVAL_WORK_COUNTS)) == 0,
              "unknown eval flags");
  const bool work = (p.flags & SBH_EVAL_WORK_COUNTS) != 0;
  if (work)
    SBH_REQUIRE(method == SBH_STOCHASTIC && counts &&
                    !(p.flags & (SBH_EVAL_EXACT_TERMS | SBH_EVAL_PER_QUERY_PATHS)),
                "SBH_EVAL_WORK_COUNTS applies to the fast shared-path estimator, with counts");
  SBH_REQUIRE(!(p.flags & SBH_EVAL_SMOOTH_DISTANCE) || kernel == SBH_SMOOTH_DIST,
              "SBH_EVAL_SMOOTH_DISTANCE applies to the smooth-distance kernel");
  if (method == SBH_BARNES_HUT) {
    SBH_REQUIRE(!std::isnan(p.beta), "beta is NaN");
    SBH_REQUIRE(p.vote_group == 1 || p.vote_group == 32 || p.vote_group == 64,
                "vote_group must be 1, 32 or 64");
  }
  if (method == SBH_STOCHASTIC)
    SBH_REQUIRE(p.samples_per_subdomain >= 1, "samples_per_subdomain must be >= 1");
  sbh::DeviceTree& t = tree->t;
  hipStream_t s = t.stream;
  const bool dev = p.flags & SBH_EVAL_DEVICE_POINTERS;
  const int split = (p.flags & (SBH_EVAL_SPLIT_COUNTS | SBH_EVAL_WORK_COUNTS)) ? 1 : 0;
  const bool exact = (p.flags & SBH_EVAL_EXACT_TERMS) != 0;
  const uint32_t nq = (uint32_t)q;
  const size_t ncnt = split ? 2ull * nq : nq;
  const size_t qfloats = (size_t)nq * t.dim;
  const float* dq = queries;
  float* dout = results;
  int32_t* dcnt = counts;
  if (!dev) {
    SBH_REQUIRE(!sc.fixed(), "sbh_eval_noalloc takes device pointers (SBH_EVAL_DEVICE_POINTERS)");
    SBH_REQUIRE(all_finite(queries, qfloats), "queries must be finite");
    ensure(tree->q, qfloats);
    ensure(tree->out, nq);
    SBH_HIP(hipMemcpyAsync(tree->q.get(), queries, qfloats * 4, hipMemcpyHostToDevice, s));
    dq = tree->q.get();
    dout = tree->out.get();
    if (counts) {
      ensure(tree->counts, ncnt);
      dcnt = tree->counts.get();
    }
  }
  // Counting only exists for the tree walks.
  if (method == SBH_BRUTE_FORCE) {
    const bool mfma = (p.flags & (SBH_EVAL_MFMA | SBH_EVAL_MFMA_BF16)) != 0;
    if (mfma)
      SBH_REQUIRE(t.dim == 3 && (kernel == SBH_GRAVITY || kernel == SBH_SMOOTH_DIST),
                  "SBH_EVAL_MFMA covers 3D gravity and smooth distance");
    uint32_t ns = mfma ? sbh::brute_mfma_split_count(t.n, nq) : sbh::brute_split_count(t.n, nq);
    double* partial = ns > 1 ? sc.partial((size_t)ns * nq) : nullptr;
    uint32_t* ord = mfma ? sc.order(nq) : nullptr;
    uint32_t* sort_ws = mfma ? sc.sort(sbh::sort_scratch_u32(nq)) : nullptr;
    SBH_HIP(hipEventRecord(tree->ev0, s));
    if (mfma) {
      sbh::query_order(t, dq, nq, ord, sort_ws, s);
      sbh::run_brute_mfma(t, kernel, (p.flags & SBH_EVAL_MFMA_BF16) != 0, dq, nq, ord, dout,
                          partial, ns, s);
    } else {
      sbh::run_brute(t, kernel, dq, nq, dout, partial, ns, s);
    }
    SBH_HIP(hipEventRecord(tree->ev1, s));
    if (dcnt) {
      // every query evaluates every point; split counts have no node terms
      if (split) sbh::fill_i32(dcnt, nq, 0, s);
      sbh::fill_i32(dcnt + (split ? nq : 0), nq, (int32_t)t.n, s);
    }
  } else {
    const bool sort = (p.flags & SBH_EVAL_SORT_QUERIES) != 0;
    uint32_t* ord = sort ? sc.order(nq) : nullptr;
    uint32_t* sort_ws = sort ? sc.sort(sbh::sort_scratch_u32(nq)) : nullptr;
    float* bh_partial = nullptr;
    const bool no_split = (p.flags & (SBH_EVAL_NO_SPLIT | SBH_EVAL_WORK_COUNTS)) != 0;
    if (no_split) {
      // whole-walk kernels at any size: no partial sums
    } else if (method == SBH_BARNES_HUT) {
      const uint32_t ns = sbh::bh_split_count(t, nq, p.vote_group, exact, dcnt != nullptr);
      if (ns > 1) {
        // ns * nq floats (rounded up to an even count), then ns * nq int2 counts
        bh_partial = reinterpret_cast<float*>(
            sc.partial(((size_t)ns * nq + 1) / 2 + (dcnt ? (size_t)ns * nq : 0)));
      }
    } else {
      const uint32_t ns = sbh::stochastic_split_count(
          t, nq, (p.flags & SBH_EVAL_PER_QUERY_PATHS) != 0, exact);
      if (ns > 1) {
        // ns * nq float2 parts, then ns * nq int2 counts
        bh_partial = reinterpret_cast<float*>(sc.partial((size_t)ns * nq * (dcnt ? 2 : 1)));
      }
    }
    SBH_HIP(hipEventRecord(tree->ev0, s));
    if (sort) sbh::query_order(t, dq, nq, ord, sort_ws, s);
    if (method == SBH_BARNES_HUT)
      sbh::run_bh(t, kernel, p.beta, p.vote_group, dq, nq, ord, dout, dcnt, split, exact,
                  bh_partial, s);
    else
      sbh::run_stochastic(t, kernel, p.samples_per_subdomain, p.seed,
                          (p.flags & SBH_EVAL_PER_QUERY_PATHS) != 0, dq, nq, ord, dout, dcnt,
                          split, exact, bh_partial, s, work);
    SBH_HIP(hipEventRecord(tree->ev1, s));
  }
  if (p.flags & SBH_EVAL_SMOOTH_DISTANCE) sbh::run_smooth_distance(dout, nq, t.alpha, s);
  if (!dev) {
    SBH_HIP(hipMemcpyAsync(results, dout, nq * 4, hipMemcpyDeviceToHost, s));
    if (counts) SBH_HIP(hipMemcpyAsync(counts, dcnt, ncnt * 4, hipMemcpyDeviceToHost, s));
  }
  SBH_HIP(hipEventSynchronize(tree->ev1));
  if (kernel_ms) SBH_HIP(hipEventElapsedTime(kernel_ms, tree->ev0, tree->ev1));
  SBH_HIP(hipStreamSynchronize(s));
}

}  // namespace

extern "C" {

const char* sbh_last_error(void) { return g_last_error.c_str(); }

int32_t sbh_version(void) { return 100; }

int sbh_device_count(int32_t* count) {
  return guarded([&] {
    SBH_REQUIRE(count, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    *count = (e == hipSuccess) ? n : 0;
  });
}

int sbh_tree_create(int32_t dim, int32_t split_log, int32_t leaf_size, int32_t max_depth,
                    int64_t n, const float* points, const float* normals, const float* masses,
                    const float* min_corner, const float* max_corner, float alpha, int32_t flags,
                    sbh_tree** out, float* build_ms) {
  return guarded([&] {
    SBH_REQUIRE(out, "out is NULL");
    *out = nullptr;
    check_tree_inputs<float>(dim, split_log, leaf_size, max_depth, n, points, normals, masses,
                             min_corner, max_corner, alpha);
    require_device();
    sbh_tree* h = new sbh_tree();
    try {
      sbh::DeviceTree& t = h->t;
      t.dim = dim;
      t.split_log = split_log;
      t.leaf_size = leaf_size;
      t.max_depth = max_depth;
      t.moments64 = (flags & SBH_TREE_MOMENTS_F64) ? 1 : 0;
      t.n = (uint32_t)n;
      t.alpha = alpha;
      for (int a = 0; a < dim; ++a) {
        t.lo[a] = min_corner[a];
        t.hi[a] = max_corner[a];
      }
      SBH_HIP(hipStreamCreateWithFlags(&t.stream, hipStreamNonBlocking));
      SBH_HIP(hipEventCreate(&h->ev0));
      SBH_HIP(hipEventCreate(&h->ev1));
      SBH_HIP(hipGetDevice(&h->device));
      sbh::build_device_tree(t, points, normals, masses, build_ms);
    } catch (...) {
      sbh_tree_destroy(h);
      throw;
    }
    *out = h;
  });
}

int sbh_tree_create_double(int32_t dim, int32_t split_log, int32_t leaf_size,
                           int32_t max_depth, int64_t n, const double* points,
                           const double* normals, const double* masses,
                           const double* min_corner, const double* max_corner, double alpha,
                           sbh_tree** out, float* build_ms) {
  return guarded([&] {
    SBH_REQUIRE(out, "out is NULL");
    *out = nullptr;
    check_tree_inputs<double>(dim, split_log, leaf_size, max_depth, n, points, normals, masses,
                              min_corner, max_corner, alpha);
    require_device();
    sbh_tree* h = new sbh_tree();
    try {
      sbh::DeviceTree& t = h->t;
      t.is_double = true;
      t.dim = dim;
      t.split_log = split_log;
      t.leaf_size = leaf_size;
      t.max_depth = max_depth;
      t.n = (uint32_t)n;
      t.alpha64 = alpha;
      t.alpha = (float)alpha;
      for (int a = 0; a < dim; ++a) {
        t.lo64[a] = min_corner[a];
        t.hi64[a] = max_corner[a];
        t.lo[a] = (float)min_corner[a];
        t.hi[a] = (float)max_corner[a];
      }
      SBH_HIP(hipStreamCreateWithFlags(&t.stream, hipStreamNonBlocking));
      SBH_HIP(hipEventCreate(&h->ev0));
      SBH_HIP(hipEventCreate(&h->ev1));
      SBH_HIP(hipGetDevice(&h->device));
      sbh::build_device_tree_double(t, points, normals, masses, build_ms);
    } catch (...) {
      sbh_tree_destroy(h);
      throw;
    }
    *out = h;
  });
}

int sbh_tree_is_double(const sbh_tree* tree, int32_t* is_double) {
  return guarded([&] {
    SBH_REQUIRE(tree && is_double, "NULL argument");
    *is_double = tree->t.is_double ? 1 : 0;
  });
}

int sbh_tree_destroy(sbh_tree* tree) {
  return guarded([&] {
    if (!tree) return;
    if (tree->t.stream) (void)hipStreamSynchronize(tree->t.stream);
    if (tree->ev0) (void)hipEventDestroy(tree->ev0);
    if (tree->ev1) (void)hipEventDestroy(tree->ev1);
    hipStream_t s = tree->t.stream;
    delete tree;
    if (s) (void)hipStreamDestroy(s);
  });
}

int sbh_tree_info(const sbh_tree* tree, int64_t* num_points, int64_t* num_nodes, int32_t* depth,
                  int64_t* device_bytes) {
  return guarded([&] {
    SBH_REQUIRE(tree, "tree is NULL");
    if (num_points) *num_points = tree->t.n;
    if (num_nodes) *num_nodes = tree->t.num_nodes;
    if (depth) *depth = tree->t.depth;
    if (device_bytes) *device_bytes = (int64_t)tree->t.device_bytes();
  });
}

int sbh_tree_stream(sbh_tree* tree, void** stream) {
  return guarded([&] {
    SBH_REQUIRE(tree && stream, "NULL argument");
    *stream = (void*)tree->t.stream;
  });
}

int sbh_tree_serialize(const sbh_tree* tree, void* buffer, int64_t capacity, int64_t* size) {
  return guarded([&] {
    SBH_REQUIRE(tree && size, "NULL argument");
    SBH_REQUIRE(!tree->t.is_double, "tree images hold float-build trees");
    std::vector<uint8_t> img;
    sbh::serialize_tree(tree->t, img);
    *size = (int64_t)img.size();
    if (!buffer) return;
    SBH_REQUIRE(capacity >= (int64_t)img.size(), "buffer smaller than the tree image");
    std::memcpy(buffer, img.data(), img.size());
  });
}

namespace {
sbh_tree* tree_from_image(const uint8_t* data, size_t size) {
  sbh_tree* h = new sbh_tree();
  try {
    sbh::deserialize_tree(h->t, data, size);  // validates, then creates the stream
    SBH_HIP(hipEventCreate(&h->ev0));
    SBH_HIP(hipEventCreate(&h->ev1));
    SBH_HIP(hipGetDevice(&h->device));
  } catch (...) {
    sbh_tree_destroy(h);
    throw;
  }
  return h;
}
}  // namespace

int sbh_tree_layout(const sbh_tree* tree, int32_t* dim, int32_t* split_log, int32_t* leaf_size,
                    float* alpha) {
  return guarded([&] {
    SBH_REQUIRE(tree, "tree is NULL");
    if (dim) *dim = tree->t.dim;
    if (split_log) *split_log = tree->t.split_log;
    if (leaf_size) *leaf_size = tree->t.leaf_size;
    if (alpha) *alpha = tree->t.alpha;
  });
}

int sbh_tree_deserialize(const void* buffer, int64_t size, sbh_tree** out) {
  return guarded([&] {
    SBH_REQUIRE(out, "out is NULL");
    *out = nullptr;
    SBH_REQUIRE(buffer && size > 0, "empty tree image");
    *out = tree_from_image(static_cast<const uint8_t*>(buffer), (size_t)size);
  });
}

int sbh_tree_save(const sbh_tree* tree, const char* path) {
  return guarded([&] {
    SBH_REQUIRE(tree && path, "NULL argument");
    SBH_REQUIRE(!tree->t.is_double, "tree images hold float-build trees");
    std::vector<uint8_t> img;
    sbh::serialize_tree(tree->t, img);
    FILE* f = std::fopen(path, "wb");
    SBH_REQUIRE(f, std::string("cannot open ") + path + " for writing");
    const size_t w = std::fwrite(img.data(), 1, img.size(), f);
    const bool ok = (w == img.size()) && std::fflush(f) == 0;
    std::fclose(f);
    SBH_REQUIRE(ok, std::string("short write to ") + path);
  });
}

int sbh_tree_load(const char* path, sbh_tree** out) {
  return guarded([&] {
    SBH_REQUIRE(out && path, "NULL argument");
    *out = nullptr;
    FILE* f = std::fopen(path, "rb");
    SBH_REQUIRE(f, std::string("cannot open ") + path);
    std::vector<uint8_t> img;
    uint8_t buf[1 << 16];
    size_t r;
    while ((r = std::fread(buf, 1, sizeof(buf), f)) > 0) img.insert(img.end(), buf, buf + r);
    const bool err = std::ferror(f) != 0;
    std::fclose(f);
    SBH_REQUIRE(!err, std::string("read error on ") + path);
    SBH_REQUIRE(!img.empty(), std::string(path) + " is empty");
    *out = tree_from_image(img.data(), img.size());
  });
}

int sbh_tree_leaf_order(const sbh_tree* tree, int64_t* out) {
  return guarded([&] {
    SBH_REQUIRE(tree && out, "NULL argument");
    const sbh::DeviceTree& t = tree->t;
    std::vector<uint32_t> p(t.n);
    SBH_HIP(hipMemcpyAsync(p.data(), t.perm.get(), t.n * sizeof(uint32_t),
                           hipMemcpyDeviceToHost, t.stream));
    SBH_HIP(hipStreamSynchronize(t.stream));
    for (uint32_t i = 0; i < t.n; ++i) out[i] = p[i];
  });
}

int sbh_tree_nodes(const sbh_tree* tree, uint32_t* start, uint32_t* end, uint32_t* first_child,
                   uint32_t* num_children, int32_t* depth, uint32_t* parent, uint32_t* rope,
                   float* min_corner, float* max_corner, float* diameter, float* centroid,
                   float* mass, float* normal, double* moments64) {
  return guarded([&] {
    SBH_REQUIRE(tree, "tree is NULL");
    const sbh::DeviceTree& t = tree->t;
    SBH_REQUIRE(!t.is_double || !(min_corner || max_corner || diameter || centroid || mass ||
                                  normal),
                "a double-build tree's bounds, diameters and moments are float64: "
                "sbh_tree_nodes_double");
    const size_t nn = t.num_nodes;
    const int d = t.dim;
    hipStream_t s = t.stream;
    auto cp = [&](void* dst, const void* src, size_t bytes) {
      if (dst) SBH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    };
    cp(start, t.n_start.get(), nn * 4);
    cp(end, t.n_end.get(), nn * 4);
    cp(first_child, t.n_first.get(), nn * 4);
    cp(num_children, t.n_nchild.get(), nn * 4);
    cp(depth, t.n_depth.get(), nn * 4);
    cp(parent, t.n_parent.get(), nn * 4);
    cp(rope, t.n_next.get(), nn * 4);
    cp(diameter, t.n_diam.get(), nn * 4);
    cp(moments64, t.n_m64.get(), nn * 8 * sizeof(double));
    std::vector<float4> lo, hi, c, nr;
    if (min_corner) {
      lo.resize(nn);
      cp(lo.data(), t.n_lo.get(), nn * sizeof(float4));
    }
    if (max_corner) {
      hi.resize(nn);
      cp(hi.data(), t.n_hi.get(), nn * sizeof(float4));
    }
    if (centroid || mass) {
      c.resize(nn);
      cp(c.data(), t.nodeC.get(), nn * sizeof(float4));
    }
    if (normal) {
      nr.resize(nn);
      cp(nr.data(), t.nodeN.get(), nn * sizeof(float4));
    }
    SBH_HIP(hipStreamSynchronize(s));
    auto put = [&](float* dst, const std::vector<float4>& v) {
      if (!dst) return;
      for (size_t i = 0; i < nn; ++i) {
        dst[i * d] = v[i].x;
        dst[i * d + 1] = v[i].y;
        if (d == 3) dst[i * d + 2] = v[i].z;
      }
    };
    put(min_corner, lo);
    put(max_corner, hi);
    put(centroid, c);
    put(normal, nr);
    if (mass)
      for (size_t i = 0; i < nn; ++i) mass[i] = c[i].w;
  });
}

int sbh_tree_nodes_double(const sbh_tree* tree, double* min_corner, double* max_corner,
                          double* diameter, double* centroid, double* mass, double* normal) {
  return guarded([&] {
    SBH_REQUIRE(tree, "tree is NULL");
    const sbh::DeviceTree& t = tree->t;
    SBH_REQUIRE(t.is_double, "sbh_tree_nodes_double takes a double-build tree");
    const size_t nn = t.num_nodes;
    const int d = t.dim;
    hipStream_t s = t.stream;
    std::vector<sbh::dvec4> lo, hi, c, nr;
    auto fetch = [&](std::vector<sbh::dvec4>& v, const sbh::dvec4* src) {
      v.resize(nn);
      SBH_HIP(hipMemcpyAsync(v.data(), src, nn * sizeof(sbh::dvec4), hipMemcpyDeviceToHost, s));
    };
    if (min_corner) fetch(lo, t.n_lo64.get());
    if (max_corner) fetch(hi, t.n_hi64.get());
    if (centroid || mass) fetch(c, t.nodeC64.get());
    if (normal) fetch(nr, t.nodeN64.get());
    if (diameter)
      SBH_HIP(hipMemcpyAsync(diameter, t.n_diam64.get(), nn * sizeof(double),
                             hipMemcpyDeviceToHost, s));
    SBH_HIP(hipStreamSynchronize(s));
    auto put = [&](double* dst, const std::vector<sbh::dvec4>& v) {
      if (!dst) return;
      for (size_t i = 0; i < nn; ++i) {
        dst[i * d] = v[i].x;
        dst[i * d + 1] = v[i].y;
        if (d == 3) dst[i * d + 2] = v[i].z;
      }
    };
    put(min_corner, lo);
    put(max_corner, hi);
    put(centroid, c);
    put(normal, nr);
    if (mass)
      for (size_t i = 0; i < nn; ++i) mass[i] = c[i].w;
  });
}

int sbh_query_order(sbh_tree* tree, const float* queries, int64_t q, uint32_t* order,
                    int32_t flags) {
  return guarded([&] {
    SBH_REQUIRE(tree && queries && order, "NULL argument");
    require_float_build(tree->t, "sbh_query_order");
    SBH_REQUIRE(q >= 0 && q < (int64_t)0x7FFFFFFF, "query count out of range");
    if (q == 0) return;
    sbh::DeviceTree& t = tree->t;
    hipStream_t s = t.stream;
    const bool dev = flags & SBH_EVAL_DEVICE_POINTERS;
    const size_t nq = (size_t)q;
    const float* dq = queries;
    if (!dev) {
      ensure(tree->q, nq * t.dim);
      SBH_HIP(hipMemcpyAsync(tree->q.get(), queries, nq * t.dim * 4, hipMemcpyHostToDevice, s));
      dq = tree->q.get();
    }
    ensure(tree->sort_scratch, sbh::sort_scratch_u32((uint32_t)nq));
    uint32_t* dorder = order;
    if (!dev) {
      ensure(tree->order, nq);
      dorder = tree->order.get();
    }
    sbh::query_order(t, dq, (uint32_t)nq, dorder, tree->sort_scratch.get(), s);
    if (!dev) SBH_HIP(hipMemcpyAsync(order, dorder, nq * 4, hipMemcpyDeviceToHost, s));
    SBH_HIP(hipStreamSynchronize(s));
  });
}

int sbh_eval(sbh_tree* tree, int32_t method, int32_t kernel, const sbh_eval_params* params,
             const float* queries, int64_t q, float* results, int32_t* counts, float* kernel_ms) {
  return guarded([&] {
    SBH_REQUIRE(tree, "NULL argument");
    EvalScratch sc(tree);
    eval_float(tree, method, kernel, params, queries, q, results, counts, kernel_ms, sc);
  });
}

int sbh_eval_noalloc(sbh_tree* tree, int32_t method, int32_t kernel,
                     const sbh_eval_params* params, const float* queries, int64_t q,
                     float* results, int32_t* counts, float* kernel_ms) {
  return guarded([&] {
    SBH_REQUIRE(tree && params, "NULL argument");
    SBH_REQUIRE(params->flags & SBH_EVAL_DEVICE_POINTERS,
                "sbh_eval_noalloc takes device pointers (SBH_EVAL_DEVICE_POINTERS)");
    EvalScratch sc(scratch_on(tree->device), SBH_NOALLOC_SCRATCH_BYTES);
    eval_float(tree, method, kernel, params, queries, q, results, counts, kernel_ms, sc);
  });
}

int sbh_eval_f64(sbh_tree* tree, int32_t kernel, const float* queries, int64_t q,
                 double* results, int32_t flags, float* kernel_ms) {
  return guarded([&] {
    SBH_REQUIRE(tree, "tree is NULL");
    require_float_build(tree->t, "sbh_eval_f64");
    SBH_REQUIRE(kernel >= SBH_GRAVITY && kernel <= SBH_GRAVITY_GRAD_Z, "unknown kernel");
    SBH_REQUIRE(kernel != SBH_GRAVITY_GRAD_Z || tree->t.dim == 3,
                "the gradient's z component needs a 3D tree");
    SBH_REQUIRE(q >= 0 && q < (int64_t)0x7FFFFFFF, "query count out of range");
    if (kernel_ms) *kernel_ms = 0.0f;
    if (q == 0) return;
    SBH_REQUIRE(queries && results, "NULL queries or results");
    sbh::DeviceTree& t = tree->t;
    hipStream_t s = t.stream;
    const bool dev = flags & SBH_EVAL_DEVICE_POINTERS;
    const uint32_t nq = (uint32_t)q;
    const size_t qfloats = (size_t)nq * t.dim;
    const float* dq = queries;
    double* dout = results;
    if (!dev) {
      SBH_REQUIRE(all_finite(queries, qfloats), "queries must be finite");
      ensure(tree->q, qfloats);
      ensure(tree->out64, nq);
      SBH_HIP(hipMemcpyAsync(tree->q.get(), queries, qfloats * 4, hipMemcpyHostToDevice, s));
      dq = tree->q.get();
      dout = tree->out64.get();
    }
    uint32_t ns = std::max(1u, sbh::brute_split_count(t.n, nq));
    ensure(tree->partial, (size_t)ns * nq);
    SBH_HIP(hipEventRecord(tree->ev0, s));
    sbh::run_brute_f64(t, kernel, dq, nq, dout, tree->partial.get(), ns, s);
    SBH_HIP(hipEventRecord(tree->ev1, s));
    if (!dev) SBH_HIP(hipMemcpyAsync(results, dout, nq * 8, hipMemcpyDeviceToHost, s));
    SBH_HIP(hipEventSynchronize(tree->ev1));
    if (kernel_ms) SBH_HIP(hipEventElapsedTime(kernel_ms, tree->ev0, tree->ev1));
    SBH_HIP(hipStreamSynchronize(s));
  });
}

int sbh_eval_double(sbh_tree* tree, int32_t method, int32_t kernel,
                    const sbh_eval_params_double* params, const double* queries, int64_t q,
                    double* results, int32_t* counts, float* kernel_ms) {
  return guarded([&] {
    SBH_REQUIRE(tree && params, "NULL argument");
    sbh::DeviceTree& t = tree->t;
    SBH_REQUIRE(t.is_double, "sbh_eval_double takes a double-build tree (sbh_tree_create_double)");
    SBH_REQUIRE(method == SBH_BRUTE_FORCE || method == SBH_BARNES_HUT || method == SBH_STOCHASTIC,
                "unknown method");
    SBH_REQUIRE(kernel >= SBH_GRAVITY && kernel <= SBH_GRAVITY_GRAD_Z, "unknown kernel");
    SBH_REQUIRE(kernel != SBH_GRAVITY_GRAD_Z || t.dim == 3,
                "the gradient's z component needs a 3D tree");
    SBH_REQUIRE(q >= 0 && q < (int64_t)0x7FFFFFFF, "query count out of range");
    if (kernel_ms) *kernel_ms = 0.0f;
    if (q == 0) return;
    SBH_REQUIRE(queries && results, "NULL queries or results");
    const sbh_eval_params_double& p = *params;
    SBH_REQUIRE((p.flags & ~(SBH_EVAL_DEVICE_POINTERS | SBH_EVAL_SORT_QUERIES |
                             SBH_EVAL_SPLIT_COUNTS | SBH_EVAL_EXACT_TERMS |
                             SBH_EVAL_PER_QUERY_PATHS | SBH_EVAL_SMOOTH_DISTANCE)) == 0,
                "unsupported eval flags for a double-build tree (the matrix-core forms are "
                "float-build options)");
    SBH_REQUIRE(!(p.flags & SBH_EVAL_SMOOTH_DISTANCE) || kernel == SBH_SMOOTH_DIST,
                "SBH_EVAL_SMOOTH_DISTANCE applies to the smooth-distance kernel");
    if (method == SBH_BARNES_HUT) {
      SBH_REQUIRE(!std::isnan(p.beta), "beta is NaN");
      SBH_REQUIRE(p.vote_group == 1 || p.vote_group == 32 || p.vote_group == 64,
                  "vote_group must be 1, 32 or 64");
    }
    if (method == SBH_STOCHASTIC)
      SBH_REQUIRE(p.samples_per_subdomain >= 1, "samples_per_subdomain must be >= 1");
    hipStream_t s = t.stream;
    const bool dev = p.flags & SBH_EVAL_DEVICE_POINTERS;
    const int split = (p.flags & SBH_EVAL_SPLIT_COUNTS) ? 1 : 0;
    const uint32_t nq = (uint32_t)q;
    const size_t ncnt = split ? 2ull * nq : nq;
    const size_t qvals = (size_t)nq * t.dim;
    const double* dq = queries;
    double* dout = results;
    int32_t* dcnt = counts;
    if (!dev) {
      SBH_REQUIRE(all_finite(queries, qvals), "queries must be finite");
      ensure(tree->qd, qvals);
      ensure(tree->out64, nq);
      SBH_HIP(hipMemcpyAsync(tree->qd.get(), queries, qvals * sizeof(double),
                             hipMemcpyHostToDevice, s));
      dq = tree->qd.get();
      dout = tree->out64.get();
      if (counts) {
        ensure(tree->counts, ncnt);
        dcnt = tree->counts.get();
      }
    }
    // the estimator's level-1 memo: 64 doubles (and with counts 128 int32) a
    // query, kept by the tree and grown on demand; skipped beyond 4 GiB
    double* memo = nullptr;
    int32_t* memo_cnt = nullptr;
    if (method == SBH_STOCHASTIC && (size_t)nq * 64 * sizeof(double) <= ((size_t)4 << 30)) {
      ensure(tree->memo64, (size_t)nq * 64);
      memo = tree->memo64.get();
      if (dcnt) {
        ensure(tree->memo_cnt, (size_t)nq * 128);
        memo_cnt = tree->memo_cnt.get();
      }
    }
    const bool sort = method != SBH_BRUTE_FORCE && (p.flags & SBH_EVAL_SORT_QUERIES) != 0;
    uint32_t* ord = nullptr;
    if (sort) {
      ensure(tree->q, qvals);
      ensure(tree->order, nq);
      ensure(tree->sort_scratch, sbh::sort_scratch_u32(nq));
      ord = tree->order.get();
    }
    SBH_HIP(hipEventRecord(tree->ev0, s));
    if (sort) sbh::query_order_double(t, dq, nq, tree->q.get(), ord, tree->sort_scratch.get(), s);
    if (method == SBH_BRUTE_FORCE)
      sbh::run_brute_double(t, kernel, dq, nq, dout, s);
    else if (method == SBH_BARNES_HUT)
      sbh::run_bh_double(t, kernel, p.beta, p.vote_group, dq, nq, ord, dout, dcnt, split, s);
    else
      sbh::run_stochastic_double(t, kernel, p.samples_per_subdomain, p.seed,
                                 (p.flags & SBH_EVAL_PER_QUERY_PATHS) != 0, dq, nq, ord, dout,
                                 dcnt, split, memo, memo_cnt, s);
    SBH_HIP(hipEventRecord(tree->ev1, s));
    if (method == SBH_BRUTE_FORCE && dcnt) {
      // every query evaluates every point; split counts have no node terms
      std::vector<int32_t> c(ncnt, (int32_t)t.n);
      if (split) std::fill(c.begin(), c.begin() + nq, 0);
      SBH_HIP(hipMemcpyAsync(dcnt, c.data(), ncnt * 4, hipMemcpyHostToDevice, s));
      SBH_HIP(hipStreamSynchronize(s));
    }
    if (p.flags & SBH_EVAL_SMOOTH_DISTANCE) sbh::run_smooth_distance_double(dout, nq, t.alpha64, s);
    if (!dev) {
      SBH_HIP(hipMemcpyAsync(results, dout, nq * sizeof(double), hipMemcpyDeviceToHost, s));
      if (counts) SBH_HIP(hipMemcpyAsync(counts, dcnt, ncnt * 4, hipMemcpyDeviceToHost, s));
    }
    SBH_HIP(hipEventSynchronize(tree->ev1));
    if (kernel_ms) SBH_HIP(hipEventElapsedTime(kernel_ms, tree->ev0, tree->ev1));
    SBH_HIP(hipStreamSynchronize(s));
  });
}

int sbh_server_create(int64_t budget_bytes, int32_t max_inflight, sbh_server** out) {
  return guarded([&] {
    SBH_REQUIRE(out, "out is NULL");
    *out = nullptr;
    SBH_REQUIRE(budget_bytes > 0, "budget must be positive");
    require_device();
    auto h = std::make_unique<sbh_server>();
    h->s = std::make_unique<sbh::Server>((size_t)budget_bytes, max_inflight);
    *out = h.release();
  });
}

int sbh_server_destroy(sbh_server* server) {
  return guarded([&] { delete server; });
}

int sbh_server_add_tree(sbh_server* server, sbh_tree* tree, int32_t* tree_id) {
  return guarded([&] {
    SBH_REQUIRE(server && tree && tree_id, "NULL argument");
    SBH_REQUIRE(!tree->t.is_double, "the serving engine holds float-build trees");
    *tree_id = server->s->add_tree(&tree->t);
  });
}

int sbh_server_submit(sbh_server* server, const sbh_request* request, int64_t* request_id) {
  return guarded([&] {
    SBH_REQUIRE(server && request && request_id, "NULL argument");
    const sbh_request& r = *request;
    *request_id = server->s->submit(r.tree_id, r.method, r.kernel, r.params, r.queries, r.q,
                                    r.results);
  });
}

int sbh_server_poll(sbh_server* server, int64_t request_id, int32_t* state, double* latency_ms) {
  return guarded([&] {
    SBH_REQUIRE(server && state, "NULL argument");
    *state = server->s->state(request_id, latency_ms);
  });
}

int sbh_server_wait(sbh_server* server, int64_t request_id, double timeout_s, int32_t* state,
                    double* latency_ms) {
  return guarded([&] {
    SBH_REQUIRE(server && state, "NULL argument");
    *state = server->s->wait(request_id, timeout_s, latency_ms);
  });
}

int sbh_server_cancel(sbh_server* server, int64_t request_id, int32_t* cancelled) {
  return guarded([&] {
    SBH_REQUIRE(server && cancelled, "NULL argument");
    *cancelled = server->s->cancel(request_id) ? 1 : 0;
  });
}

int sbh_server_release(sbh_server* server, int64_t request_id) {
  return guarded([&] {
    SBH_REQUIRE(server, "NULL argument");
    server->s->release(request_id);
  });
}

int sbh_server_stats_get(sbh_server* server, sbh_server_stats* stats) {
  return guarded([&] {
    SBH_REQUIRE(server && stats, "NULL argument");
    sbh::ServerStats st = server->s->stats();
    stats->completed = st.completed;
    stats->launches = st.launches;
    stats->queued = st.queued;
    stats->bytes_in_use = st.bytes_in_use;
    stats->peak_bytes = st.peak_bytes;
    stats->budget_bytes = st.budget_bytes;
    stats->tree_bytes = st.tree_bytes;
  });
}

}  // extern "C"
