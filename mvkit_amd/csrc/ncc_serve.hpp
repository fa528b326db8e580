// A request loop for NCC finds on one device: requests (one host frame and
// the models to find in it) are queued, admitted while the estimated device
// working sets of the requests in flight stay within an HBM budget, and run
// by worker threads in micro-batches (several requests in one find_requests
// call, all their queries' coarse searches in one persistent launch); the
// workers' device work overlaps on the GPU.  Models stay resident on the
// device (their device copies are cached on first use).  The reference's
// contract is a synchronous find that may run concurrently on shared
// immutable models (docs/runtime.md:31, matcher.cpp:305); each request's
// result is the one a find_models call returns for it alone.
#pragma once

#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "ncc_device.hpp"

namespace mvamd {

struct ServeConfig {
  int workers = 2;                          // request executors, each with its own streams and workspaces
  size_t max_batch = 8;                     // requests per micro-batch
  size_t hbm_budget = size_t(16) << 30;     // admission budget of the requests in flight (estimated bytes)
  size_t queue_limit = 4096;                // requests waiting for admission
  bool coarse_queue = true;                 // persistent coarse launch per micro-batch
};

struct ServeTimes {
  double admitted_us = 0, started_us = 0, finished_us = 0;  // after submission
  double bytes = 0;                                         // the estimated working set admitted
  int batch_size = 0, worker = -1;
};

class NccServer {
 public:
  explicit NccServer(const ServeConfig& config);
  ~NccServer();
  NccServer(const NccServer&) = delete;
  NccServer& operator=(const NccServer&) = delete;

  // Queues one request: a copy of the host frame (u8 or u16 rows `pitch`
  // bytes apart) and the models with find_models' options and selection.
  uint64_t submit(const uint8_t* pixels, size_t width, size_t height, size_t pitch, int pixel_bytes,
                  std::vector<std::shared_ptr<NccModel>> models, std::vector<FindOptions> options,
                  const Selection& selection);
  // Waits for a request and hands its result over (rethrowing its error).
  std::vector<ModelMatch> wait(uint64_t id, ServeTimes* times);
  // The working set a request is admitted with: its frame, search pyramid,
  // and per model the top-level prefixes, score planes, candidate lists and
  // hash tables.
  static size_t estimate_bytes(size_t width, size_t height, int pixel_bytes,
                               const std::vector<std::shared_ptr<NccModel>>& models,
                               const std::vector<FindOptions>& options);

 private:
  struct Request {
    uint64_t id = 0;
    std::vector<uint8_t> pixels;  // rows packed
    size_t width = 0, height = 0;
    int pixel_bytes = 1;
    std::vector<std::shared_ptr<NccModel>> models;
    std::vector<FindOptions> options;
    Selection selection;
    size_t bytes = 0;
    double t_submit = 0;
    ServeTimes times;
    std::vector<ModelMatch> result;
    bool failed = false;
    Error error{kInternalError, "", "", ""};
    bool done = false;
  };
  void worker_loop(int index);
  bool admissible(const Request& r) const { return in_flight_ == 0 || in_flight_ + r.bytes <= config_.hbm_budget; }

  ServeConfig config_;
  int device_ = 0;
  std::mutex mu_;
  std::condition_variable work_cv_, done_cv_;
  std::deque<Request*> waiting_;
  std::map<uint64_t, std::unique_ptr<Request>> requests_;
  size_t in_flight_ = 0;
  uint64_t next_id_ = 1;
  bool stop_ = false;
  std::vector<std::thread> workers_;
};

}  // namespace mvamd
