// Serving: resident trees, continuous batching, admission by HBM footprint.
//
// Requests (a query set against one resident tree, Barnes-Hut or stochastic,
// any kernel) are queued by sbh_server_submit.  A scheduler thread admits
// queued requests in arrival order while their working sets (queries,
// results, the queries' spatial order and its sort scratch) fit in the
// server's device arena, puts each request's queries in Hilbert-curve order
// (so a wavefront's 64 queries are neighbours and its walks coherent), and
// launches every admitted batch as ONE work-queue kernel: each wavefront
// repeatedly takes the next (request, 64-query block of the sorted order)
// task and runs the matching walk, whatever the tree, dimension, kernel or
// method.  Up to `max_inflight` launches run at once
// on separate streams, so requests that arrive while others are on the GPU
// are admitted as soon as memory frees and their work overlaps on the device.
//
// A request's values depend only on its own queries and parameters: the walks
// are per query (Barnes-Hut without votes) or keyed on (seed, block of the
// request's sorted order) (stochastic), so a result is identical alone and in
// any mix, and equal to a one-shot sorted evaluation of the request.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/sbh_hip.h"
#include "sbh_eval.hpp"
#include "sbh_runtime.hpp"
#include "sbh_serve.hpp"
#include "sbh_walk.hpp"

namespace sbh {
namespace {

constexpr int kServeThreads = 256;

// Device descriptor of one admitted request.
struct ReqDesc {
  TreeView tree;
  const float* qx;
  const float* qy;
  const float* qz;
  const uint32_t* order;  // query evaluated in each slot (Hilbert order)
  float* out;
  uint32_t nq;
  int32_t variant;  // ((dim == 3) * 3 + kernel) * 2 + (method == stochastic)
  float beta;
  int32_t spl;
  uint32_t key0, key1;
  int32_t sorted;  // nodes have at most 64 children: the sorted-sample estimator
};

struct Task {
  uint32_t req;    // index into the launch's descriptor array
  uint32_t block;  // 64-query block within the request
};

template <int D, int K>
__device__ __forceinline__ float serve_one(const ReqDesc& r, bool stochastic, uint32_t block,
                                           bool live, float x, float y, float z,
                                           SamplerLds lds, SortedLds sorted_lds) {
  // the same estimator as a one-shot sorted evaluation of the request picks,
  // so a served result equals that one-shot result
  if (stochastic && uniform_u32((uint32_t)r.sorted))
    return stochastic_estimate_sorted<D, K, false>(r.tree, r.spl, r.key0, r.key1, block, live, x,
                                                   y, z, sorted_lds)
        .value;
  if (stochastic)
    return stochastic_estimate<D, K, false>(r.tree, r.spl, r.key0, r.key1, block, live, x, y, z,
                                            lds, false)
        .value;
  return bh_walk<D, K, 1, false, false>(r.tree, r.beta, live, x, y, z).value;
}

__global__ __launch_bounds__(kServeThreads) __attribute__((amdgpu_waves_per_eu(5, 5))) void k_serve(const ReqDesc* __restrict__ reqs,
                                                          const Task* __restrict__ tasks,
                                                          uint32_t ntasks,
                                                          uint32_t* __restrict__ next_task) {
  __shared__ float4 sSrc[kServeThreads];
  __shared__ float4 sNrm[kServeThreads];
  __shared__ uint4 sRec[kServeThreads];
  __shared__ uint2 sAux[kServeThreads];
  constexpr uint32_t kWaves = kServeThreads / 64;
  __shared__ uint2 sKeys[kWaves * kChunk];
  __shared__ uint2 sRec2[kServeThreads];
  __shared__ float sDiff[kWaves * kCacheLevels * 64];
  __shared__ uint4 sNext[kWaves * (kCacheLevels + 1) * 3];
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wbase = threadIdx.x & ~63u;
  while (true) {
    uint32_t t = 0;
    if (lane == 0) t = atomicAdd(next_task, 1u);
    t = uniform_u32(__shfl(t, 0, 64));
    if (t >= ntasks) break;  // every wavefront reaches this once the list is drained
    const Task tk = tasks[t];
    const ReqDesc& r = reqs[tk.req];
    const uint32_t slot = tk.block * 64 + lane;
    const bool live = slot < r.nq;
    const uint32_t qi = live ? r.order[slot] : 0u;
    const int variant = (int)uniform_u32((uint32_t)r.variant);
    const int dim3 = variant >= 6;
    float x, y, z;
    load_query(r.qx, r.qy, r.qz, qi, dim3 ? 3 : 2, x, y, z);
    const bool stoch = variant & 1;
    const SamplerLds c{sSrc + wbase, sNrm + wbase, sRec + wbase, sAux + wbase};
    const uint32_t wave = threadIdx.x >> 6;
    const SortedLds sc{sSrc + wbase, sNrm + wbase, sRec2 + wbase, sKeys + wave * kChunk,
                       sDiff + wave * kCacheLevels * 64, sNext + wave * (kCacheLevels + 1) * 3};
    float v;
    switch (variant >> 1) {
      case 0:
        v = serve_one<2, GRAVITY>(r, stoch, tk.block, live, x, y, z, c, sc);
        break;
      case 1:
        v = serve_one<2, WINDING_NUMBER>(r, stoch, tk.block, live, x, y, z, c, sc);
        break;
      case 2:
        v = serve_one<2, SMOOTH_DIST>(r, stoch, tk.block, live, x, y, z, c, sc);
        break;
      case 3:
        v = serve_one<3, GRAVITY>(r, stoch, tk.block, live, x, y, z, c, sc);
        break;
      case 4:
        v = serve_one<3, WINDING_NUMBER>(r, stoch, tk.block, live, x, y, z, c, sc);
        break;
      default:
        v = serve_one<3, SMOOTH_DIST>(r, stoch, tk.block, live, x, y, z, c, sc);
        break;
    }
    if (live) r.out[qi] = v;
  }
}

__global__ void k_reset_counter(uint32_t* p) { *p = 0; }

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Device bytes of a request in the arena: queries and results (floats), the
// launch order (u32), and the scratch of the sort that makes it.
size_t request_bytes(uint32_t nq, int dim) {
  return align256((size_t)nq * (dim + 2) * 4) + align256(sort_scratch_u32(nq) * 4);
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

}  // namespace

// First-fit allocator over one device allocation.
class Arena {
 public:
  void init(size_t bytes) {
    buf_.alloc(bytes);
    size_ = bytes;
    free_.clear();
    free_[0] = bytes;
  }
  size_t size() const { return size_; }
  char* base() const { return reinterpret_cast<char*>(buf_.get()); }
  bool alloc(size_t bytes, size_t* off) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        *off = it->first;
        size_t rest = it->second - bytes;
        size_t noff = it->first + bytes;
        free_.erase(it);
        if (rest) free_[noff] = rest;
        used_ += bytes;
        peak_ = std::max(peak_, used_);
        return true;
      }
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t used() const { return used_; }
  size_t peak() const { return peak_; }

 private:
  DevBuf<uint8_t> buf_;
  size_t size_ = 0, used_ = 0, peak_ = 0;
  std::map<size_t, size_t> free_;
};

struct Server::Request {
  int64_t id = 0;
  int32_t tree_id = 0, method = 0, kernel = 0;
  sbh_eval_params params{};
  const float* queries = nullptr;  // host, column-major dim x q
  float* results = nullptr;        // host
  uint32_t nq = 0;
  int dim = 3;
  size_t bytes = 0, off = 0;
  int state = SBH_REQ_QUEUED;
  double t_submit = 0, t_done = 0;
};

struct Server::Launch {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;
  std::vector<Request*> reqs;
  std::vector<size_t> res_off;  // offset of each request's results in the staging buffer
  float* stage = nullptr;       // pinned host staging: queries then results
  size_t stage_floats = 0;
  DevBuf<ReqDesc> descs;
  DevBuf<Task> tasks;
  DevBuf<uint32_t> counter;
  std::vector<ReqDesc> h_descs;
  std::vector<Task> h_tasks;
};

Server::Server(size_t budget_bytes, int max_inflight) {
  SBH_REQUIRE(budget_bytes >= 4096, "budget must be at least 4 KiB");
  SBH_REQUIRE(max_inflight >= 1 && max_inflight <= 8, "max_inflight must be in [1, 8]");
  arena_ = std::make_unique<Arena>();
  arena_->init(budget_bytes);
  for (int i = 0; i < max_inflight; ++i) {
    auto L = std::make_unique<Launch>();
    SBH_HIP(hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking));
    SBH_HIP(hipEventCreateWithFlags(&L->done, hipEventDisableTiming));
    L->counter.alloc(1);
    launches_.push_back(std::move(L));
  }
  worker_ = std::thread([this] { loop(); });
}

Server::~Server() {
  {
    std::unique_lock<std::mutex> g(mu_);
    stop_ = true;
    // requests that never started are cancelled; threads blocked in wait()
    // wake up, see the state, and leave before the members go away
    for (Request* r : queue_) r->state = SBH_REQ_CANCELLED;
    queue_.clear();
    cv_.notify_all();
    cv_.wait(g, [&] { return waiters_ == 0; });
  }
  cv_.notify_all();
  if (worker_.joinable()) worker_.join();
  for (auto& L : launches_) {
    if (L->stream) (void)hipStreamSynchronize(L->stream);
    if (L->done) (void)hipEventDestroy(L->done);
    if (L->stream) (void)hipStreamDestroy(L->stream);
    if (L->stage) (void)hipHostFree(L->stage);
  }
}

int32_t Server::add_tree(DeviceTree* t) {
  std::lock_guard<std::mutex> g(mu_);
  trees_.push_back(t);
  tree_bytes_ += t->device_bytes();
  return (int32_t)trees_.size() - 1;
}

int64_t Server::submit(int32_t tree_id, int32_t method, int32_t kernel, const sbh_eval_params& p,
                       const float* queries, int64_t q, float* results) {
  std::unique_lock<std::mutex> g(mu_);
  SBH_REQUIRE(tree_id >= 0 && tree_id < (int32_t)trees_.size(), "unknown tree id");
  SBH_REQUIRE(method == SBH_BARNES_HUT || method == SBH_STOCHASTIC,
              "serving runs Barnes-Hut or stochastic requests");
  SBH_REQUIRE(kernel >= 0 && kernel <= 2,
              "serving evaluates the gravity, winding-number and smooth-distance kernels");
  SBH_REQUIRE(q >= 1 && q < (int64_t)0x7FFFFFFF, "query count out of range");
  SBH_REQUIRE(queries && results, "NULL queries or results");
  SBH_REQUIRE(!stop_, "the server is shutting down");
  // the engine always evaluates in Hilbert order with the whole-walk kernels
  // and fast terms; any other mode is refused rather than ignored
  SBH_REQUIRE((p.flags & ~(SBH_EVAL_SORT_QUERIES | SBH_EVAL_NO_SPLIT)) == 0,
              "serving supports no eval flags besides SBH_EVAL_SORT_QUERIES and "
              "SBH_EVAL_NO_SPLIT (what it always does)");
  if (method == SBH_BARNES_HUT) SBH_REQUIRE(p.beta == p.beta, "beta is NaN");
  if (method == SBH_STOCHASTIC)
    SBH_REQUIRE(p.samples_per_subdomain >= 1, "samples_per_subdomain must be >= 1");
  auto r = std::make_unique<Request>();
  r->id = next_id_++;
  r->tree_id = tree_id;
  r->method = method;
  r->kernel = kernel;
  r->params = p;
  r->queries = queries;
  r->results = results;
  r->nq = (uint32_t)q;
  r->dim = trees_[tree_id]->dim;
  r->bytes = request_bytes(r->nq, r->dim);
  if (r->bytes > arena_->size())
    throw Error(kErrBudget, "request working set (" + std::to_string(r->bytes) +
                                " bytes) exceeds the server's HBM budget");
  r->t_submit = now_s();
  Request* raw = r.get();
  requests_[raw->id] = std::move(r);
  queue_.push_back(raw);
  g.unlock();
  cv_.notify_all();
  return raw->id;
}

int Server::state(int64_t id, double* latency_ms) {
  std::lock_guard<std::mutex> g(mu_);
  auto it = requests_.find(id);
  SBH_REQUIRE(it != requests_.end(), "unknown request id");
  if (latency_ms)
    *latency_ms = it->second->state == SBH_REQ_DONE
                      ? (it->second->t_done - it->second->t_submit) * 1e3
                      : 0.0;
  return it->second->state;
}

int Server::wait(int64_t id, double timeout_s, double* latency_ms) {
  std::unique_lock<std::mutex> g(mu_);
  auto it = requests_.find(id);
  SBH_REQUIRE(it != requests_.end(), "unknown request id");
  Request* r = it->second.get();
  auto pred = [&] { return r->state >= SBH_REQ_DONE || !error_.empty() || stop_; };
  ++waiters_;
  if (timeout_s < 0)
    cv_.wait(g, pred);
  else
    cv_.wait_for(g, std::chrono::duration<double>(timeout_s), pred);
  if (--waiters_ == 0 && stop_) cv_.notify_all();
  if (!error_.empty() && r->state < SBH_REQ_DONE) throw Error(kErrHip, error_);
  if (latency_ms) *latency_ms = r->state == SBH_REQ_DONE ? (r->t_done - r->t_submit) * 1e3 : 0.0;
  return r->state;
}

bool Server::cancel(int64_t id) {
  std::lock_guard<std::mutex> g(mu_);
  auto it = std::find_if(queue_.begin(), queue_.end(), [&](Request* r) { return r->id == id; });
  if (it == queue_.end()) return false;
  (*it)->state = SBH_REQ_CANCELLED;
  queue_.erase(it);
  cv_.notify_all();
  return true;
}

void Server::release(int64_t id) {
  std::lock_guard<std::mutex> g(mu_);
  auto it = requests_.find(id);
  SBH_REQUIRE(it != requests_.end(), "unknown request id");
  SBH_REQUIRE(it->second->state >= SBH_REQ_DONE, "request still queued or running");
  requests_.erase(it);
}

ServerStats Server::stats() {
  std::lock_guard<std::mutex> g(mu_);
  ServerStats s;
  s.completed = completed_;
  s.launches = nlaunches_;
  s.queued = (int64_t)queue_.size();
  s.bytes_in_use = (int64_t)arena_->used();
  s.peak_bytes = (int64_t)arena_->peak();
  s.budget_bytes = (int64_t)arena_->size();
  s.tree_bytes = (int64_t)tree_bytes_;
  return s;
}

// Scheduler thread: retire finished launches, admit and launch, sleep.
void Server::loop() {
  std::unique_lock<std::mutex> g(mu_);
  while (true) {
    bool progressed = false;
    for (auto& L : launches_) {
      if (!L->busy) continue;
      hipError_t e = hipEventQuery(L->done);
      if (e == hipErrorNotReady) continue;
      if (e != hipSuccess) error_ = std::string("serving launch failed: ") + hipGetErrorString(e);
      retire(*L, e == hipSuccess);
      progressed = true;
      cv_.notify_all();
    }
    if (stop_) {
      bool any = false;
      for (auto& L : launches_) any |= L->busy;
      if (!any) break;
    }
    if (!stop_ && error_.empty() && !queue_.empty()) {
      Launch* free_slot = nullptr;
      for (auto& L : launches_)
        if (!L->busy) {
          free_slot = L.get();
          break;
        }
      if (free_slot) {
        std::vector<Request*> batch;
        while (!queue_.empty()) {
          Request* r = queue_.front();
          size_t off;
          if (!arena_->alloc(r->bytes, &off)) break;  // head of line waits for memory
          r->off = off;
          r->state = SBH_REQ_RUNNING;
          queue_.pop_front();
          batch.push_back(r);
        }
        if (!batch.empty()) {
          try {
            launch(*free_slot, batch);
            ++nlaunches_;
          } catch (const std::exception& ex) {
            error_ = ex.what();
            for (Request* r : batch) {
              r->state = SBH_REQ_FAILED;
              arena_->release(r->off, r->bytes);
            }
          }
          progressed = true;
          cv_.notify_all();
        }
      }
    }
    if (!progressed) {
      bool inflight = false;
      for (auto& L : launches_) inflight |= L->busy;
      // poll finished launches every 50 us while work is in flight
      cv_.wait_for(g, std::chrono::microseconds(inflight ? 50 : 2000));
    }
  }
}

// Results leave the pinned staging buffer for the caller's array, then the
// request's arena space is returned.  Called with mu_ held.
void Server::retire(Launch& L, bool ok) {
  double t = now_s();
  for (size_t i = 0; i < L.reqs.size(); ++i) {
    Request* r = L.reqs[i];
    if (ok) {
      std::memcpy(r->results, L.stage + L.res_off[i], (size_t)r->nq * sizeof(float));
      r->state = SBH_REQ_DONE;
      r->t_done = now_s();
      ++completed_;
    } else {
      r->state = SBH_REQ_FAILED;
      r->t_done = t;
    }
    arena_->release(r->off, r->bytes);
  }
  L.reqs.clear();
  L.busy = false;
}

// Copies the batch's queries (through the launch's pinned staging buffer) into
// the arena, builds descriptors and the task list, launches the work-queue
// kernel and the result copies, all asynchronously on L.stream.  Called with
// mu_ held.
void Server::launch(Launch& L, std::vector<Request*>& batch) {
  hipStream_t s = L.stream;
  L.h_descs.clear();
  L.h_tasks.clear();
  L.res_off.clear();
  size_t need = 0;
  for (Request* r : batch) need += (size_t)r->nq * (r->dim + 1);
  if (need > L.stage_floats) {
    if (L.stage) SBH_HIP(hipHostFree(L.stage));
    L.stage = nullptr;
    L.stage_floats = 0;
    size_t cap = std::max(need, (size_t)1 << 20);
    SBH_HIP(hipHostMalloc((void**)&L.stage, cap * sizeof(float), hipHostMallocDefault));
    L.stage_floats = cap;
  }
  char* base = arena_->base();
  size_t so = 0;
  for (size_t i = 0; i < batch.size(); ++i) {
    Request* r = batch[i];
    const DeviceTree& t = *trees_[r->tree_id];
    float* dq = reinterpret_cast<float*>(base + r->off);
    float* dout = dq + (size_t)r->nq * r->dim;
    uint32_t* dorder = reinterpret_cast<uint32_t*>(dout + r->nq);
    uint32_t* dscratch =
        reinterpret_cast<uint32_t*>(base + r->off + align256((size_t)r->nq * (r->dim + 2) * 4));
    const size_t qf = (size_t)r->nq * r->dim;
    std::memcpy(L.stage + so, r->queries, qf * sizeof(float));
    SBH_HIP(hipMemcpyAsync(dq, L.stage + so, qf * sizeof(float), hipMemcpyHostToDevice, s));
    query_order(t, dq, r->nq, dorder, dscratch, s);
    so += qf;
    L.res_off.push_back(so);
    so += r->nq;
    ReqDesc d;
    d.tree = view_of(t);
    d.qx = dq;
    d.qy = dq + r->nq;
    d.qz = r->dim == 3 ? dq + 2ull * r->nq : nullptr;
    d.order = dorder;
    d.out = dout;
    d.nq = r->nq;
    d.variant = (((r->dim == 3) ? 3 : 0) + r->kernel) * 2 + (r->method == SBH_STOCHASTIC ? 1 : 0);
    d.beta = r->params.beta;
    d.spl = r->params.samples_per_subdomain;
    d.key0 = (uint32_t)r->params.seed;
    d.key1 = (uint32_t)(r->params.seed >> 32);
    d.sorted = t.dim * t.split_log <= 6 ? 1 : 0;  // at most 2^6 children per node
    L.h_descs.push_back(d);
    const uint32_t nb = (r->nq + 63) / 64;
    for (uint32_t b = 0; b < nb; ++b) L.h_tasks.push_back(Task{(uint32_t)i, b});
  }
  // largest requests first: their blocks start early and the tail is short
  std::stable_sort(L.h_tasks.begin(), L.h_tasks.end(), [&](const Task& a, const Task& b) {
    return L.h_descs[a.req].nq > L.h_descs[b.req].nq;
  });
  if (L.descs.size() < L.h_descs.size()) L.descs.alloc(L.h_descs.size() * 2);
  if (L.tasks.size() < L.h_tasks.size()) L.tasks.alloc(L.h_tasks.size() * 2);
  SBH_HIP(hipMemcpyAsync(L.descs.get(), L.h_descs.data(), L.h_descs.size() * sizeof(ReqDesc),
                         hipMemcpyHostToDevice, s));
  SBH_HIP(hipMemcpyAsync(L.tasks.get(), L.h_tasks.data(), L.h_tasks.size() * sizeof(Task),
                         hipMemcpyHostToDevice, s));
  k_reset_counter<<<1, 1, 0, s>>>(L.counter.get());
  const uint32_t ntasks = (uint32_t)L.h_tasks.size();
  const uint32_t waves_per_block = kServeThreads / 64;
  const uint32_t grid = std::min<uint32_t>((ntasks + waves_per_block - 1) / waves_per_block, 2048);
  k_serve<<<grid, kServeThreads, 0, s>>>(L.descs.get(), L.tasks.get(), ntasks, L.counter.get());
  SBH_HIP(hipGetLastError());
  for (size_t i = 0; i < batch.size(); ++i) {
    Request* r = batch[i];
    SBH_HIP(hipMemcpyAsync(L.stage + L.res_off[i], L.h_descs[i].out,
                           (size_t)r->nq * sizeof(float), hipMemcpyDeviceToHost, s));
  }
  SBH_HIP(hipEventRecord(L.done, s));
  L.reqs = batch;
  L.busy = true;
}

}  // namespace sbh
