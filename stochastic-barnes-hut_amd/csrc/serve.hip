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
// any mix, and equal to a one-shot sorted evaluation of the request.  Like a
// one-shot launch, a request has its walks split where that launch splits
// them (one task per 64-query block and subtree of the Barnes-Hut walk's
// split frontier when the request is too small to fill the GPU; per block and
// range of root subdomains for the sorted estimator), the parts added in
// subtree or subdomain order by a reduce pass after the work-queue kernel: a
// small request on a large tree then spreads over the whole GPU instead of a
// few wavefronts walking serially.
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
// Tasks (64-query blocks) a wavefront of the work-queue kernel takes at most.
constexpr uint32_t kTasksPerWave = 1;
// A request of at least this many queries is large: it runs alone in its
// launch, and large launches leave one launch slot to the other requests.
// (Classing by measured time per query of earlier launches on the same tree
// and method instead admitted fewer requests per launch: lower throughput at
// the same latency.)
constexpr uint32_t kLargeQueries = 65536;

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
  int32_t sorted;   // nodes have at most 64 children: the sorted-sample estimator
  uint32_t nsplit;  // > 1: the walks are split into nsplit tasks
  void* partial;    // split requests: nsplit x nq parts (float, or float2 for the estimator)
  const WalkTask* wtasks;  // split Barnes-Hut requests: the frontier's task table
  const uint4* stasks;     // split estimator requests: the tree's subdomain ranges
};

constexpr uint32_t kWhole = 0xFFFFFFFFu;

struct Task {
  uint32_t req;    // index into the launch's descriptor array
  uint32_t block;  // 64-query block within the request
  uint32_t sub;    // task of a split request (a walk task, or the estimator's
                   // range of root subdomains), or kWhole
};

// One task for one lane: the whole walk (result to out[qi]) or one task's
// part of a split walk (to the request's partial sums), with the
// kernels and the split rule a one-shot sorted evaluation of the request
// uses, so a served result equals that evaluation's bit for bit.
template <int D, int K>
__device__ __forceinline__ void serve_one(const ReqDesc& r, bool stochastic, uint32_t block,
                                          uint32_t sub, bool live, uint32_t slot, uint32_t qi,
                                          float x, float y, float z, SamplerLds lds,
                                          SortedLds sorted_lds) {
  if (sub != kWhole) {
    if (stochastic) {
      float bh = 0.0f;
      const uint4 tk = load_invariant(r.stasks, sub);
      const LaneResult p = stochastic_estimate_sorted<D, K, false, true>(
          r.tree, r.spl, r.key0, r.key1, block, live, x, y, z, sorted_lds, tk.x, tk.y, &bh);
      if (live)
        static_cast<float2*>(r.partial)[(size_t)tk.z * r.nq + slot] = make_float2(bh, p.value);
    } else {
      const LaneResult p =
          bh_walk_task<D, K, false>(r.tree, r.wtasks, sub, r.beta, live, x, y, z);
      if (live)
        static_cast<float*>(r.partial)[(size_t)load_invariant(&r.wtasks[sub].part, 0) * r.nq +
                                       slot] = p.value;
    }
    return;
  }
  float v;
  if (stochastic && uniform_u32((uint32_t)r.sorted))
    v = stochastic_estimate_sorted<D, K, false>(r.tree, r.spl, r.key0, r.key1, block, live, x, y,
                                                z, sorted_lds)
            .value;
  else if (stochastic)
    v = stochastic_estimate<D, K, false>(r.tree, r.spl, r.key0, r.key1, block, live, x, y, z,
                                         lds, false)
            .value;
  else
    v = bh_walk<D, K, 1, false, false, true>(r.tree, r.beta, live, x, y, z).value;
  if (live) r.out[qi] = v;
}

__global__ __launch_bounds__(kServeThreads) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_serve(const ReqDesc* __restrict__ reqs,
                                                          const Task* __restrict__ tasks,
                                                          uint32_t ntasks,
                                                          uint32_t* __restrict__ next_task,
                                                          uint32_t per_wave) {
  __shared__ float4 sSrc[kServeThreads];
  __shared__ float4 sNrm[kServeThreads];
  __shared__ uint4 sRec[kServeThreads];
  __shared__ uint2 sAux[kServeThreads];
  constexpr uint32_t kWaves = kServeThreads / 64;
  __shared__ uint2 sKeys[kWaves * kChunk];
  __shared__ uint4 sNext[kWaves * (kCacheLevels + 1) * 3];
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wbase = threadIdx.x & ~63u;
  // at most per_wave tasks a wavefront: blocks retire as they finish, so a
  // launch on another stream gets CUs without waiting for this one to drain
  for (uint32_t taken = 0; taken < per_wave; ++taken) {
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
    const SortedLds sc{sSrc + wbase, sNrm + wbase, sKeys + wave * kChunk,
                       sNext + wave * (kCacheLevels + 1) * 3};
    const uint32_t sub = uniform_u32(tk.sub);
    switch (variant >> 1) {
      case 0:
        serve_one<2, GRAVITY>(r, stoch, tk.block, sub, live, slot, qi, x, y, z, c, sc);
        break;
      case 1:
        serve_one<2, WINDING_NUMBER>(r, stoch, tk.block, sub, live, slot, qi, x, y, z, c, sc);
        break;
      case 2:
        serve_one<2, SMOOTH_DIST>(r, stoch, tk.block, sub, live, slot, qi, x, y, z, c, sc);
        break;
      case 3:
        serve_one<3, GRAVITY>(r, stoch, tk.block, sub, live, slot, qi, x, y, z, c, sc);
        break;
      case 4:
        serve_one<3, WINDING_NUMBER>(r, stoch, tk.block, sub, live, slot, qi, x, y, z, c, sc);
        break;
      default:
        serve_one<3, SMOOTH_DIST>(r, stoch, tk.block, sub, live, slot, qi, x, y, z, c, sc);
        break;
    }
  }
}

// The split requests of a launch (blockIdx.y = index into split_reqs): their
// parts added in child order, as a one-shot split launch adds them.
__global__ void k_serve_reduce(const ReqDesc* __restrict__ reqs,
                               const uint32_t* __restrict__ split_reqs) {
  const ReqDesc& r = reqs[split_reqs[blockIdx.y]];
  const bool stoch = r.variant & 1;
  for (uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x; slot < r.nq;
       slot += gridDim.x * blockDim.x) {
    const float v =
        stoch ? sum_stochastic_parts(static_cast<const float2*>(r.partial), r.nsplit, r.nq, slot)
              : sum_bh_parts(static_cast<const float*>(r.partial), r.nsplit, r.nq, slot);
    r.out[r.order[slot]] = v;
  }
}

__global__ void k_reset_counter(uint32_t* p) { *p = 0; }

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Device bytes of a request in the arena: queries and results (floats), the
// launch order (u32), the scratch of the sort that makes it, and a split
// request's parts (nsplit x nq floats, float2 for the estimator).
size_t request_bytes(uint32_t nq, int dim, uint32_t nsplit, bool stochastic) {
  const size_t parts = nsplit > 1 ? (size_t)nsplit * nq * (stochastic ? 8 : 4) : 0;
  return align256((size_t)nq * (dim + 2) * 4) + align256(sort_scratch_u32(nq) * 4) +
         align256(parts);
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
  uint32_t nsplit = 1;  // > 1: walks split into nsplit tasks
  size_t bytes = 0, off = 0;
  int state = SBH_REQ_QUEUED;
  double t_submit = 0, t_done = 0;
};

struct Server::Launch {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;
  bool large = false;  // runs one large request alone
  std::vector<Request*> reqs;
  std::vector<size_t> res_off;  // offset of each request's results in the staging buffer
  float* stage = nullptr;       // pinned host staging: queries then results
  size_t stage_floats = 0;
  DevBuf<ReqDesc> descs;
  DevBuf<Task> tasks;
  DevBuf<uint32_t> counter;
  DevBuf<uint32_t> split_reqs;
  std::vector<ReqDesc> h_descs;
  std::vector<Task> h_tasks;
  std::vector<uint32_t> h_split;
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
  // the engine always evaluates in Hilbert order with fast terms; it honours
  // SBH_EVAL_NO_SPLIT, and refuses any other mode rather than ignoring it
  SBH_REQUIRE((p.flags & ~(SBH_EVAL_SORT_QUERIES | SBH_EVAL_NO_SPLIT)) == 0,
              "serving supports no eval flags besides SBH_EVAL_SORT_QUERIES (what it always "
              "does) and SBH_EVAL_NO_SPLIT");
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
  const DeviceTree& t = *trees_[tree_id];
  r->dim = t.dim;
  // the split rule of a one-shot launch of the same size (fast terms, per
  // query, shared paths)
  r->nsplit = (p.flags & SBH_EVAL_NO_SPLIT) ? 1u
              : method == SBH_STOCHASTIC   ? stochastic_split_count(t, r->nq, false, false)
                                           : bh_split_count(t, r->nq, 1, false, false);
  r->bytes = request_bytes(r->nq, r->dim, r->nsplit, method == SBH_STOCHASTIC);
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
        // In arrival order: small requests join the batch while memory
        // allows; a large request runs alone, and large launches take at most
        // all but one slot, so small requests never queue behind a launch
        // that holds the GPU for long.
        int large_busy = 0;
        for (auto& L : launches_) large_busy += (L->busy && L->large) ? 1 : 0;
        const int large_cap = std::max(1, (int)launches_.size() - 1);
        std::vector<Request*> batch;
        bool large = false;
        for (auto it = queue_.begin(); it != queue_.end();) {
          Request* r = *it;
          const bool is_large = r->nq >= kLargeQueries;
          if (is_large && (!batch.empty() || large_busy >= large_cap)) {
            ++it;  // runs later, alone
            continue;
          }
          size_t off;
          if (!arena_->alloc(r->bytes, &off)) break;  // head of line waits for memory
          r->off = off;
          r->state = SBH_REQ_RUNNING;
          it = queue_.erase(it);
          batch.push_back(r);
          if (is_large) {
            large = true;
            break;
          }
        }
        if (!batch.empty()) {
          free_slot->large = large;
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
  L.h_split.clear();
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
    void* dparts = base + r->off + align256((size_t)r->nq * (r->dim + 2) * 4) +
                   align256(sort_scratch_u32(r->nq) * 4);
    const size_t qf = (size_t)r->nq * r->dim;
    std::memcpy(L.stage + so, r->queries, qf * sizeof(float));
    SBH_HIP(hipMemcpyAsync(dq, L.stage + so, qf * sizeof(float), hipMemcpyHostToDevice, s));
    query_order(t, dq, r->nq, dorder, dscratch, s, /*compact=*/true);
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
    d.nsplit = r->nsplit;
    d.partial = r->nsplit > 1 ? dparts : nullptr;
    d.wtasks = nullptr;
    d.stasks = nullptr;
    if (r->nsplit > 1 && r->method != SBH_STOCHASTIC) {
      const WalkSplit ws = bh_split(t, r->nq, 1, false);
      SBH_REQUIRE(ws.n == r->nsplit, "serving: split plan changed");
      d.wtasks = ws.tasks;
    } else if (r->nsplit > 1) {
      SBH_REQUIRE(t.n_sub_tasks == r->nsplit, "serving: split plan changed");
      d.stasks = t.sub_tasks.get();
    }
    L.h_descs.push_back(d);
    const uint32_t nb = (r->nq + 63) / 64;
    if (r->nsplit > 1) {
      L.h_split.push_back((uint32_t)i);
      for (uint32_t b = 0; b < nb; ++b)
        for (uint32_t c = 0; c < r->nsplit; ++c) L.h_tasks.push_back(Task{(uint32_t)i, b, c});
    } else {
      for (uint32_t b = 0; b < nb; ++b) L.h_tasks.push_back(Task{(uint32_t)i, b, kWhole});
    }
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
  const uint32_t per_block = waves_per_block * kTasksPerWave;
  const uint32_t grid = (ntasks + per_block - 1) / per_block;
  k_serve<<<grid, kServeThreads, 0, s>>>(L.descs.get(), L.tasks.get(), ntasks, L.counter.get(),
                                         kTasksPerWave);
  SBH_HIP(hipGetLastError());
  if (!L.h_split.empty()) {
    if (L.split_reqs.size() < L.h_split.size()) L.split_reqs.alloc(L.h_split.size() * 2);
    SBH_HIP(hipMemcpyAsync(L.split_reqs.get(), L.h_split.data(),
                           L.h_split.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    uint32_t max_nq = 0;
    for (uint32_t i : L.h_split) max_nq = std::max(max_nq, L.h_descs[i].nq);
    const dim3 rgrid(std::min<uint32_t>((max_nq + 255) / 256, 1024),
                     (uint32_t)L.h_split.size());
    k_serve_reduce<<<rgrid, 256, 0, s>>>(L.descs.get(), L.split_reqs.get());
    SBH_HIP(hipGetLastError());
  }
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
