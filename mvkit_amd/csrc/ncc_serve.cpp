// The NCC request loop (ncc_serve.hpp).
#include "ncc_serve.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

namespace mvamd {

namespace {
double now_us() {
  return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
const char* const kServeOp = "ncc_server";
}  // namespace

NccServer::NccServer(const ServeConfig& config) : config_(config) {
  if (config_.workers < 1 || config_.workers > 16 || !config_.max_batch || config_.max_batch > NCC_MAX_BATCH ||
      !config_.hbm_budget || !config_.queue_limit)
    throw Error(kInvalidArgument, kServeOp, "config", "workers 1..16, batch 1..256, nonzero budget and queue limit");
  check(hipGetDevice(&device_), "hipGetDevice");
  for (int i = 0; i < config_.workers; ++i) workers_.emplace_back([this, i] { worker_loop(i); });
}

NccServer::~NccServer() {
  {
    std::lock_guard<std::mutex> lock(mu_);
    stop_ = true;
  }
  work_cv_.notify_all();
  for (auto& t : workers_) t.join();
}

size_t NccServer::estimate_bytes(size_t width, size_t height, int pixel_bytes,
                                 const std::vector<std::shared_ptr<NccModel>>& models,
                                 const std::vector<FindOptions>& options) {
  const double px = double(width) * double(height);
  double bytes = px * pixel_bytes + px * (pixel_bytes == 2 ? 4.0 : 2.0) / 3.0;  // frame + block sums of every level
  for (size_t i = 0; i < models.size(); ++i) {
    const NccModel& m = *models[i];
    const FindOptions& o = options[i < options.size() ? i : 0];
    const size_t levels = o.highest ? std::min(m.levels.size(), size_t(o.highest)) : m.levels.size();
    const double top = std::ldexp(1.0, -2 * int(levels - 1));
    const double cap = double(o.max_candidates + 1);
    double table = 2;
    while (table < 2 * cap) table *= 2;
    bytes += px * top * (double(m.n_angles) * 8 + 20)      // score planes (at most every angle), row prefixes
             + cap * 32 * double(levels + 1) + table * 4 * double(levels);  // candidate lists, hash tables
  }
  return size_t(bytes);
}

uint64_t NccServer::submit(const uint8_t* pixels, size_t width, size_t height, size_t pitch, int pixel_bytes,
                           std::vector<std::shared_ptr<NccModel>> models, std::vector<FindOptions> options,
                           const Selection& selection) {
  if (!pixels || !width || !height || (pixel_bytes != 1 && pixel_bytes != 2) || pitch < width * size_t(pixel_bytes))
    throw Error(kInvalidArgument, kServeOp, "Image", "one nonempty byte or uint2 image required");
  if (models.empty() || (options.size() != 1 && options.size() != models.size()))
    throw Error(kInvalidArgument, kServeOp, "options", "one option set or one per model required");
  if (options.size() == 1 && models.size() > 1) options.resize(models.size(), options.front());
  auto r = std::make_unique<Request>();
  const size_t row = width * size_t(pixel_bytes);
  r->pixels.resize(row * height);
  for (size_t y = 0; y < height; ++y) std::memcpy(r->pixels.data() + y * row, pixels + y * pitch, row);
  r->width = width; r->height = height; r->pixel_bytes = pixel_bytes;
  r->bytes = estimate_bytes(width, height, pixel_bytes, models, options);
  r->models = std::move(models);
  r->options = std::move(options);
  r->selection = selection;
  r->t_submit = now_us();
  std::lock_guard<std::mutex> lock(mu_);
  if (waiting_.size() >= config_.queue_limit) throw Error(kLimitExceeded, kServeOp, "queue", "request queue is full");
  r->id = next_id_++;
  const uint64_t id = r->id;
  waiting_.push_back(r.get());
  requests_[id] = std::move(r);
  work_cv_.notify_one();
  return id;
}

std::vector<ModelMatch> NccServer::wait(uint64_t id, ServeTimes* times) {
  std::unique_lock<std::mutex> lock(mu_);
  auto it = requests_.find(id);
  if (it == requests_.end()) throw Error(kInvalidState, kServeOp, "request", "no such request");
  Request* r = it->second.get();
  done_cv_.wait(lock, [&] { return r->done; });
  std::unique_ptr<Request> owned = std::move(it->second);
  requests_.erase(it);
  lock.unlock();
  if (times) *times = owned->times;
  if (owned->failed) throw owned->error;
  return std::move(owned->result);
}

void NccServer::worker_loop(int index) {
  if (hipSetDevice(device_) != hipSuccess) return;
  std::unique_ptr<FindContext> ctx;
  std::vector<DeviceBuffer> images;
  for (;;) {
    std::vector<Request*> batch;
    size_t batch_bytes = 0;
    {
      std::unique_lock<std::mutex> lock(mu_);
      // admission: the oldest waiting request when its working set fits the
      // budget beside the requests in flight (always when none is)
      work_cv_.wait(lock, [&] { return stop_ || (!waiting_.empty() && admissible(*waiting_.front())); });
      if (stop_ && waiting_.empty()) break;
      if (waiting_.empty()) continue;
      const double t = now_us();
      while (!waiting_.empty() && batch.size() < config_.max_batch && admissible(*waiting_.front())) {
        Request* r = waiting_.front();
        waiting_.pop_front();
        in_flight_ += r->bytes;
        batch_bytes += r->bytes;
        r->times.admitted_us = t - r->t_submit;
        batch.push_back(r);
      }
    }
    const double t_start = now_us();
    std::vector<std::vector<ModelMatch>> results;
    bool failed = false;
    Error error{kInternalError, kServeOp, "", ""};
    try {
      if (!ctx) ctx = std::make_unique<FindContext>();
      if (images.size() < batch.size()) images.resize(batch.size());
      std::vector<FrameRequest> reqs(batch.size());
      for (size_t k = 0; k < batch.size(); ++k) {
        Request& r = *batch[k];
        if (images[k].bytes < r.pixels.size()) images[k] = DeviceBuffer(std::max<size_t>(r.pixels.size(), 1 << 20));
        check(hipMemcpyAsync(images[k].ptr, r.pixels.data(), r.pixels.size(), hipMemcpyHostToDevice, ctx->stream()),
              "image upload");
        FrameRequest& q = reqs[k];
        q.frame = FrameInput{images[k].as<uint8_t>(), r.width, r.height, r.width * size_t(r.pixel_bytes)};
        q.frame.pixel_bytes = r.pixel_bytes;
        for (const auto& m : r.models) q.models.push_back(m.get());
        q.options = r.options;
        q.selection = r.selection;
      }
      results = ctx->find_requests(reqs.data(), reqs.size(), config_.coarse_queue);
    } catch (const Error& e) {
      failed = true;
      error = e;
    } catch (const std::exception& e) {
      failed = true;
      error = Error(kInternalError, kServeOp, "", e.what());
    }
    {
      std::lock_guard<std::mutex> lock(mu_);
      const double t_end = now_us();
      for (size_t k = 0; k < batch.size(); ++k) {
        Request& r = *batch[k];
        r.times.started_us = t_start - r.t_submit;
        r.times.finished_us = t_end - r.t_submit;
        r.times.bytes = double(r.bytes);
        r.times.batch_size = int(batch.size());
        r.times.worker = index;
        r.failed = failed;
        if (failed) r.error = error;
        else r.result = std::move(results[k]);
        r.done = true;
      }
      in_flight_ -= batch_bytes;
    }
    done_cv_.notify_all();
    work_cv_.notify_all();
  }
  // the worker's context and buffers are released here, on a live thread of a
  // live process (the server is destroyed by an explicit call)
}

}  // namespace mvamd
