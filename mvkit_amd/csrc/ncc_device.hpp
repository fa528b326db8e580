// Device side of the NCC find path: device copies of models and the per-caller
// find pipeline (search pyramid, coarse correlation, peaks, tracking,
// speculative refinement) driving the HIP kernels through the ncchip C ABI.
#pragma once

#include <hip/hip_runtime_api.h>

#include <memory>
#include <string>
#include <vector>

#include "../../include/ncchip.h"
#include "ncc_model.hpp"
#include "ncc_region.hpp"

namespace mvamd {

// RAII device allocation.
struct DeviceBuffer {
  void* ptr = nullptr;
  size_t bytes = 0;
  DeviceBuffer() = default;
  explicit DeviceBuffer(size_t n);
  ~DeviceBuffer();
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& o) noexcept : ptr(o.ptr), bytes(o.bytes) { o.ptr = nullptr; o.bytes = 0; }
  DeviceBuffer& operator=(DeviceBuffer&& o) noexcept;
  template <class T> T* as() const { return static_cast<T*>(ptr); }
};

// Pinned host staging buffer.
struct PinnedBuffer {
  void* ptr = nullptr;
  size_t bytes = 0;
  PinnedBuffer() = default;
  ~PinnedBuffer();
  void reserve(size_t n);
  PinnedBuffer(const PinnedBuffer&) = delete;
  PinnedBuffer& operator=(const PinnedBuffer&) = delete;
};

void check(hipError_t e, const char* what);

// A raster pool resident on one device.
struct DevicePool {
  DeviceBuffer storage;
  NccPatchPool view{};            // device pointers into `storage`
  size_t n_rasters = 0;
  // Uploads a host pool (synchronously on `stream`).
  void upload(const RasterPool& pool, hipStream_t stream);
};

class DeviceModel {
 public:
  int device = 0;
  DevicePool pool;
};

// Returns the model's copy on the current device, uploading it on first use.
std::shared_ptr<DeviceModel> device_model(const NccModel& m, hipStream_t stream);

constexpr int kFormAuto = -1;

struct FindOptions {
  double start = 0, extent = 0, min_score = .8, max_overlap = .5;
  size_t num_matches = 1, max_candidates = 10000;
  bool subpixel = true;
  int highest = 0, lowest = 1;
  // coarse correlation form (NCCHIP_FORM_*); kFormAuto: split-BF16 on byte
  // images, FP32 on u16 images.  A split-BF16 or FP8-screened search whose
  // top level or union bounding box does not fit the matrix-core kernel runs
  // the FP32 form (FindTrace::coarse_form says which form ran).
  int form = kFormAuto;
  double screen = -1.0 / 0.0;      // FP8 screen threshold (form NCCHIP_FORM_FP8_SCREEN; -inf: 0.55)
  int profile = 0;                 // bit s: time stage s (KernelStage) with HIP events on the context's stream
  bool trace_stages = false;       // fill FindTrace's stage lists (coarse, tracked, storage)
  // find_models' query cache (matcher.cpp:281-301): a candidate below the
  // minimum score is recovered before its fit
  bool query_cache = false;
  // find_models over several models: the whole level 0 is the storage region
  // and the recovery windows use its row prefix statistics (matcher.cpp:596-611, 991-1003)
  bool whole_level0 = false;
  // tracking of levels 1..4 on matrix cores with the split-BF16 and
  // FP8-screened forms (exact results; off: FP64 VALU kernel only).  The FP32
  // and FP64 forms track on the FP64 VALU kernel.
  bool track_mfma = true;
  // replay the device part as a captured HIP graph when the batch's key
  // (model, options, frame geometry, workspace) repeats
  bool graph = true;
  // tile shape of the split-BF16 coarse kernel: 0 the model's autotuned shape
  // for this kind of search (tuned on the first batch that needs it), -1 the
  // default shape, else a shape packed as pack_tile() does
  int coarse_tile = 0;
};

// NcchipCoarseTile packed into an int: rows | col_waves << 4 | row_waves << 8 |
// tiles << 12 | b << 16 with b = 0, 1, 2, 3 for bound 0, 8, 16, 32.
inline int pack_tile(const NcchipCoarseTile& t) {
  const int b = t.bound == 8 ? 1 : t.bound == 16 ? 2 : t.bound == 32 ? 3 : 0;
  return t.rows | (t.col_waves << 4) | (t.row_waves << 8) | (t.tiles << 12) | (b << 16);
}
inline NcchipCoarseTile unpack_tile(int v) {
  const int b = (v >> 16) & 3;
  return NcchipCoarseTile{v & 15, (v >> 4) & 15, (v >> 8) & 15, (v >> 12) & 15, b ? 4 << b : 0};
}

// Quota and overlap shared by all models of find_models (multi.hpp:14);
// negative fields mean "per model".
struct Selection { int64_t num_matches = -1; double max_overlap = -1; };
// A find_models result: the pose and the index of its model in the call.
struct ModelMatch { Pose pose; size_t model = 0; };

// Totals of a find_models_frames / find_requests call.
struct BatchStats {
  double coarse_queue_us = 0;   // device time of the persistent coarse launches (HIP events on their stream)
  double coarse_flops = 0;      // algorithmic FLOPs of every query's coarse search
  double total_us = 0;          // host time of the call
  double tune_us = 0;           // host time spent tuning coarse tile shapes
  int queries = 0, queue_queries = 0, candidates = 0;
  int tuned = 0;                // queries whose tile shape the call tuned
  int refined = 0, refine_rounds = 0;  // candidates refined, in this many speculative rounds
};

// Kernel stages timed in profiling mode.
// kStageTrack: tracking levels on the FP64 vector kernel; kStageTrackMfma:
// tracking levels on matrix cores.
enum KernelStage { kStagePyramid = 0, kStagePrefix, kStageCoarse, kStagePeaks, kStageTrack, kStageFull, kStageRefine,
                   kStageTrackMfma, kStageCount };

// Intermediate results of one find, for stage-by-stage checks.
struct FindTrace {
  std::vector<Pose> coarse;                 // peaks, sorted in the reference's emission order
  std::vector<std::vector<Pose>> tracked;   // after each tracking level (merged, filtered)
  std::vector<Pose> prepared;               // full-scored, sorted
  std::vector<Pose> refined;                // per prepared candidate
  std::vector<int> verdict;                 // per prepared candidate
  int top = 0, finest = 0, angle_stride = 1;
  int storage[4] = {0, 0, 0, 0};
  double t_coarse_us = 0, t_track_us = 0, t_refine_us = 0, t_total_us = 0;
  double kernel_us[kStageCount] = {};  // device time per stage (profiling mode)
  double coarse_flops = 0;             // 2 * sum over evaluated coarse windows of their sample counts
  double track_flops = 0;              // 2 * samples of the tracked 5x5 windows (counted when tracking is profiled)
  double track_mfma_flops = 0;         // the part of track_flops tracked on matrix cores
  int coarse_form = -1;                // correlation form the coarse search ran (NCCHIP_FORM_*)
  int coarse_tile = 0;                 // tile shape of the split-BF16 coarse kernel (pack_tile; 0: it did not run)
};

// A u8 or u16 search image in device memory (`pitch` bytes per row), with
// its domain: null for the full image, else a region inside the image (the
// reference's reduced domain, image.hpp:108-113; it may be empty).
struct FrameInput {
  const uint8_t* image;
  size_t width, height, pitch;
  int pixel_bytes = 1;  // 1: u8 ("byte"), 2: u16 ("uint2")
  const Region* domain = nullptr;
};

// One request of find_requests: a frame and the models to find in it, with
// find_models' per-model options and selection.
struct FrameRequest {
  FrameInput frame{};
  std::vector<const NccModel*> models;
  std::vector<FindOptions> options;  // one per model
  Selection selection;
  bool per_model_quota_before_overlap = true;
};

// One find pipeline: owns a stream and grow-only device/pinned workspaces.
// Not thread-safe; use one context per host thread.
class FindContext {
 public:
  FindContext();
  ~FindContext();
  FindContext(const FindContext&) = delete;
  FindContext& operator=(const FindContext&) = delete;

  // `image` is a device pointer to a u8 image with `pitch` bytes per row.
  std::vector<Pose> find(const NccModel& model, const uint8_t* image, size_t width, size_t height, size_t pitch,
                         const FindOptions& options, FindTrace* trace = nullptr);
  // Finds the model in several frames with the same options.  Every stage
  // runs for all frames before the next, and the per-candidate kernels take
  // the candidates of every frame in one launch; each frame's result is the
  // one find() returns for it.  traces (optional) has n_frames entries; the
  // kernel timings of the batch are reported in every trace.
  std::vector<std::vector<Pose>> find_batch(const NccModel& model, const FrameInput* frames, size_t n_frames,
                                            const FindOptions& options, FindTrace* traces = nullptr);
  // The two halves of find_batch, for overlapping consecutive batches: submit
  // enqueues the device-resident part (pyramid to full score, at most
  // NCC_MAX_BATCH frames) without waiting and returns a ticket; complete
  // waits for it and runs the refinement rounds (on a second stream, so they
  // overlap the device work of a batch submitted after this one).  At most
  // two batches may be pending; the model, the frames and the traces must
  // stay alive until complete returns.
  uint64_t submit_batch(const NccModel& model, const FrameInput* frames, size_t n_frames, const FindOptions& options,
                        FindTrace* traces = nullptr);
  std::vector<std::vector<Pose>> complete_batch(uint64_t ticket);
  // Same as find, from host memory (copied to the device first).
  std::vector<Pose> find_host(const NccModel& model, const uint8_t* image, size_t width, size_t height,
                              size_t pitch, const FindOptions& options, FindTrace* trace = nullptr,
                              int pixel_bytes = 1, const Region* domain = nullptr);
  // find_models (matcher.cpp:1077-1124) over one u8 frame in device memory:
  // one query per distinct (model, options), the candidates of all queries
  // selected together in (score, model, row, col, angle) order with the
  // selection's quota/overlap rules, refined on the device.  Each distinct
  // query keeps a workspace of its own in the context.
  std::vector<ModelMatch> find_models(const NccModel* const* models, const FindOptions* options, size_t n_models,
                                      const FrameInput& frame, const Selection& selection,
                                      bool per_model_quota_before_overlap = true);
  // find_models over several frames at once, each frame's result the one
  // find_models returns for it: one query per distinct (model, options) over
  // all frames (one launch per stage for the frames of a query), the
  // split-BF16 coarse searches of all queries in one persistent launch per
  // kernel instantiation fed by a work queue of (image, angle tiles, block)
  // items when coarse_queue is set, and the frames' greedy selections sharing
  // their refinement rounds.
  std::vector<std::vector<ModelMatch>> find_models_frames(const NccModel* const* models, const FindOptions* options,
                                                          size_t n_models, const FrameInput* frames, size_t n_frames,
                                                          const Selection& selection,
                                                          bool per_model_quota_before_overlap = true,
                                                          bool coarse_queue = true, BatchStats* stats = nullptr);
  // find_models for several requests at once, each with its own frame,
  // models, options and selection; each request's result is the one
  // find_models returns for it alone.  One query per distinct (model,
  // options, several models in the request) covers the frames of the
  // requests that ask for it.
  std::vector<std::vector<ModelMatch>> find_requests(const FrameRequest* requests, size_t n_requests,
                                                     bool coarse_queue = true, BatchStats* stats = nullptr);
  // Same, from host memory.
  std::vector<ModelMatch> find_models_host(const NccModel* const* models, const FindOptions* options,
                                           size_t n_models, const uint8_t* image, size_t width, size_t height,
                                           size_t pitch, const Selection& selection,
                                           bool per_model_quota_before_overlap = true, int pixel_bytes = 1,
                                           const Region* domain = nullptr);

  // Waits for all work queued by this context and drops its pending batches
  // (their results are lost): the state of a context handed from a thread
  // that ended to another thread.
  void reset();

  hipStream_t stream() const { return stream_; }
  int device() const { return device_; }

  // rebuild (model.cpp:173-242) with the rotated rasters built on this
  // context's device; the model is bit-identical to rebuild_model's.
  std::unique_ptr<NccModel> rebuild(LevelGeom first, const CreateParams& params, double origin_row,
                                    double origin_col);

 private:
  struct Slot;  // state of a submitted batch
  struct Plan;  // what the back half of a batch's enqueue needs
  struct RefineRequest {
    Slot* slot;      // the query that refines it
    uint32_t frame;  // frame of the slot's batch
    Pose in, out;
    int verdict;
  };
  void enqueue(Slot& S, const NccModel& model, const FrameInput* frames, size_t n_frames, const FindOptions& options,
               FindTrace* traces);
  bool enqueue_front(Slot& S, const NccModel& model, const FrameInput* frames, size_t n_frames,
                     const FindOptions& options, FindTrace* traces);
  void front_launches(Slot& S);
  std::string graph_key(Slot& S);
  void launch_graph(Slot& S);
  void enqueue_back(Slot& S);
  NcchipCoarseJob coarse_job(Slot& S);
  void choose_tile(Slot& S);
  void tune_coarse(Slot& S);
  void collect(Slot& S);
  void refine_round(std::vector<RefineRequest>& reqs);
  void upload_image(const uint8_t* image, size_t row, size_t height, size_t pitch);
  void time_begin(int stage, hipStream_t stream);
  void time_end(hipStream_t stream);
  void time_collect(FindTrace* trace);
  std::unique_ptr<Slot> slots_[2];
  std::vector<std::unique_ptr<Slot>> multi_slots_;  // one per query of find_models
  Slot* cur_ = nullptr;  // the slot being submitted or completed
  uint64_t next_ticket_ = 0;
  hipStream_t stream_ = nullptr;   // device-resident part
  hipStream_t stream2_ = nullptr;  // refinement rounds
  hipStream_t slot_streams_[2] = {nullptr, nullptr};  // device-resident part of each batch slot
  hipEvent_t uploaded_ = nullptr;  // orders the context stream's uploads before a slot's stream
  int device_ = 0;
  DeviceBuffer image_;
  PinnedBuffer upload_stage_;  // strided host images gathered for one transfer
  bool dry_ = false;            // the back half runs its host bookkeeping only (graph replay)
  bool graphs_enabled_ = true;
  bool queued_ = false;         // searches being planned run in the shared persistent coarse launch
  DeviceBuffer create_;  // model construction workspace (grow-only)
  DeviceBuffer queue_dev_;     // the persistent coarse search's work queue
  PinnedBuffer queue_stage_;   // its host-side staging
  hipEvent_t queue_events_[2] = {nullptr, nullptr};  // around the persistent coarse launches
  hipEvent_t tune_events_[2] = {nullptr, nullptr};   // around the autotuner's sample launches
};

// Polygon overlap of two poses for a template geometry (matcher.cpp:252-277).
double pose_overlap(const Pose& a, const Pose& b, const LevelGeom& level);
// find_models' overlap of two template rectangles: intersection over the
// smaller area (multi.hpp:64-81).
double box_overlap(const Pose& a, const LevelGeom& ga, const Pose& b, const LevelGeom& gb);

}  // namespace mvamd
