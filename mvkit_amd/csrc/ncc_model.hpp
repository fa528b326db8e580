// NCC model: host FP64 construction of the rotated, zero-mean, unit-energy
// template rasters of every pyramid level, stored as flat structure-of-arrays
// pools that upload to the device as-is.
//
// Semantics follow mvkit's create/rebuild (src/ncc_matching/model.cpp:120-242)
// operation for operation, so every weight is bit-identical to the reference.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "device_types.h"
#include "ncc_region.hpp"

namespace mvamd {

// Status codes of the public C ABI (mv_status.h of the reference).
enum Status : int32_t {
  kOk = 0, kInvalidArgument = -1, kOutOfRange = -2, kTypeMismatch = -3, kNotFound = -4, kOutOfMemory = -5,
  kUnsupported = -6, kInvalidState = -7, kIoError = -100, kDecodeError = -101, kEncodeError = -102,
  kLimitExceeded = -103, kInternalError = -1000
};
// Error codes of the public C ABI (mv_api.h of the reference).
enum ErrorCode : int32_t {
  kErrUnspecified = 0, kErrNullArgument = 1, kErrInvalidCardinality = 2, kErrInvalidType = 3, kErrInvalidValue = 4,
  kErrIndexOutOfRange = 5, kErrMissingItem = 6, kErrUnsupportedValue = 7, kErrInvalidObjectState = 8,
  kErrIoFailure = 9, kErrInternalFailure = 10
};

ErrorCode default_code(Status s);

struct Error {
  Status status;
  ErrorCode code;
  std::string operation, parameter, message;
  int64_t index = -1;
  Error(Status s, std::string op, std::string param, std::string msg, int64_t idx = -1, ErrorCode c = kErrUnspecified)
      : status(s), code(c == kErrUnspecified ? default_code(s) : c), operation(std::move(op)),
        parameter(std::move(param)), message(std::move(msg)), index(idx) {}
};

struct Pose { double row = 0, col = 0, angle = 0, score = 0; };

// One template raster (a rotated patch of one level), indices into a pool.
struct Raster {
  double angle = 0, orow = 0, ocol = 0;
  double wsum = 0;    // sum of the FP64 weights in sample order
  double wsum_f = 0;  // sum of the FP32-rounded weights (exact in FP64)
  int32_t rmin = 0, rmax = 0, cmin = 0, cmax = 0;
  uint32_t first = 0, count = 0;          // samples
  uint32_t run_first = 0, run_count = 0;  // row runs of consecutive samples
  uint32_t wrun_first = 0, wrun_count = 0;// runs of equal consecutive weights
};

// Flat sample storage shared by many rasters.
struct RasterPool {
  std::vector<Raster> rasters;
  std::vector<int32_t> rc;      // (row << 16) | (col & 0xffff)
  std::vector<double> w;
  std::vector<float> wf;
  std::vector<int32_t> runs;    // 4 ints per run: row, begin, end, 0
  std::vector<int32_t> wruns;   // 4 ints per weight run
  std::vector<double> wrun_w;
  size_t bytes() const;
};

// Dense geometry of one model level (model.hpp:44-53).
struct LevelGeom {
  size_t w = 0, h = 0;
  double orow = 0, ocol = 0;
  std::vector<double> gray;     // mean-free template values
  std::vector<uint8_t> mask;
  uint32_t first_raster = 0;    // index of angle 0 in the pool
};

struct CreateParams {
  int levels = 0;               // 0: automatic
  bool absolute = false;        // "ignore_global_polarity"
  double start = 0, extent = 0, step = 0;  // step 0: automatic
};

class DeviceModel;  // device mirror (ncc_device.cpp)

// A coarse tile shape chosen by the autotuner for one kind of search (the
// fields of NcchipCoarseTile), with the measured times that chose it.
struct CoarseTune {
  int rows = 4, col_waves = 4, row_waves = 1, tiles = 1, bound = 0;
  double us = 0, default_us = 0;  // time of the sample launch with this shape and with the default shape
};

struct NccModel {
  CreateParams params;          // step holds the resolved angle step
  size_t sample_count = 0;
  double orow = 0, ocol = 0;    // centroid in template coordinates
  std::vector<LevelGeom> levels;
  size_t n_angles = 0;
  RasterPool pool;              // level-major rasters, then the original patch
  uint32_t original = 0;        // unrotated level-0 patch, divided by its norm

  // Lazily created device copies, one per device.
  mutable std::mutex device_mutex;
  mutable std::vector<std::shared_ptr<DeviceModel>> device;

  // Coarse tile shapes chosen by the autotuner, keyed by the search they
  // were measured on (level, pixel kind, coarse angles, union box, screen,
  // frame count and size class); shared by every find pipeline.
  mutable std::mutex tune_mutex;
  mutable std::map<std::string, CoarseTune> coarse_tuning;

  uint32_t raster(size_t level, size_t angle) const { return levels[level].first_raster + uint32_t(angle); }
  ~NccModel();
};

// create(): model.cpp:120-171 + rebuild (model.cpp:173-242) for a
// single-channel template of u8 or u16 samples over its domain (null: the
// whole image; else a region inside the image).
std::unique_ptr<NccModel> create_model(const void* pixels, int pixel_bytes, size_t width, size_t height,
                                       size_t stride_bytes, const CreateParams& params,
                                       const Region* domain = nullptr);

// The part of create() before rebuild (model.cpp:120-166): the centred level 0
// over the domain and the template centroid in image coordinates.
struct FirstLevel {
  LevelGeom level;
  double orow = 0, ocol = 0;
};
FirstLevel create_first_level(const void* pixels, int pixel_bytes, size_t width, size_t height, size_t stride_bytes,
                              const CreateParams& params, const Region* domain = nullptr);

// rebuild(): model.cpp:173-242 from a centred level 0 with any mask and the
// template centroid in image coordinates (persisted models).
std::unique_ptr<NccModel> rebuild_model(LevelGeom first, const CreateParams& params, double origin_row,
                                        double origin_col);

// rebuild() in parts, for builders of the rotated rasters elsewhere (the
// device): plan_model does everything up to the rasters (pyramid, angle grid,
// the original patch's samples), the caller appends level-major, angle-minor
// rasters to model->pool (level.first_raster set) exactly as build_raster
// would, and finish_model appends the original patch.
struct ModelPlan {
  std::unique_ptr<NccModel> model;
  double first_mean = 0, first_energy = 0;  // level 0's (unrotated raster's) mean and energy
  std::vector<int32_t> orig_rc;
  std::vector<double> orig_w;
};
ModelPlan plan_model(LevelGeom first, const CreateParams& params, double origin_row, double origin_col);
std::unique_ptr<NccModel> finish_model(ModelPlan plan);
// Angle i of the model's grid (model.cpp:232-238).
double raster_angle(const NccModel& m, size_t i);
// The rebuild's raster errors, and its sample budget.
[[noreturn]] void raise_degenerate_raster();
[[noreturn]] void raise_sample_budget();
size_t max_model_samples();

// Sampling grid of one raster (model.cpp:66-98): the level grid for an
// unrotated angle, else the rotated bounding box (raises LIMIT_EXCEEDED past
// the rotated-box budget) with the angle's cosine and sine.
struct RasterGrid {
  bool unrotated = false;
  double ca = 1, sa = 0;
  int r_lo = 0, r_hi = -1, c_lo = 0, c_hi = -1;
};
RasterGrid raster_grid(const LevelGeom& level, double angle);

// MVKIT_NCC 1 text model (persist.cpp:14-84).
std::string serialize_model(const NccModel& model);
std::unique_ptr<NccModel> deserialize_model(const std::string& bytes);

// Builds the raster of one level at `angle` into `pool` (rotate + normalize,
// model.cpp:39-98).  Returns the raster index, or -1 when degenerate.
int64_t build_raster(const LevelGeom& level, double angle, RasterPool& pool, const double* prepared_mean = nullptr,
                     const double* prepared_energy = nullptr);

}  // namespace mvamd
