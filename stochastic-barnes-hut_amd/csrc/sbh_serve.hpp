// Serving engine behind the sbh_server_* C ABI (serve.hip).
#pragma once

#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/sbh_hip.h"
#include "sbh_runtime.hpp"

namespace sbh {

class Arena;

struct ServerStats {
  int64_t completed = 0, launches = 0, queued = 0;
  int64_t bytes_in_use = 0, peak_bytes = 0, budget_bytes = 0, tree_bytes = 0;
};

class Server {
 public:
  Server(size_t budget_bytes, int max_inflight);
  ~Server();
  int32_t add_tree(DeviceTree* t);
  int64_t submit(int32_t tree_id, int32_t method, int32_t kernel, const sbh_eval_params& p,
                 const float* queries, int64_t q, float* results);
  int state(int64_t id, double* latency_ms);
  int wait(int64_t id, double timeout_s, double* latency_ms);
  bool cancel(int64_t id);
  void release(int64_t id);
  ServerStats stats();

 private:
  struct Request;
  struct Launch;
  void loop();
  void launch(Launch& L, std::vector<Request*>& batch);
  void retire(Launch& L, bool ok);

  std::mutex mu_;
  std::condition_variable cv_;
  std::thread worker_;
  bool stop_ = false;
  int waiters_ = 0;  // threads inside wait(); the destructor lets them leave first
  std::string error_;
  std::unique_ptr<Arena> arena_;
  std::vector<DeviceTree*> trees_;
  size_t tree_bytes_ = 0;
  std::deque<Request*> queue_;
  std::unordered_map<int64_t, std::unique_ptr<Request>> requests_;
  std::vector<std::unique_ptr<Launch>> launches_;
  int64_t next_id_ = 1, completed_ = 0, nlaunches_ = 0;
};

}  // namespace sbh
