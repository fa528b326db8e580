// C ABI of the NCC find path: control tuples, image objects, model handles,
// thread-local error state and the NCC operators, with the reference's
// argument rules (src/c_api/ncc_matching_api.cpp, tuple_input.hpp,
// error_state.hpp).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <optional>
#include <string>
#include <variant>
#include <vector>

#include "../../include/mv_ncc.h"
#include "../../include/mvkit_amd.h"
#include "ncc_device.hpp"
#include "ncc_model.hpp"
#include "ncc_serve.hpp"

using mvamd::Error;

// ------------------------------------------------------------ object model
// A handle's object: an NCC model, or a serialized item (the bytes of a
// model, mv_handle.h:63-97).  Both are null once the handle is cleared.
struct ModelObject {
  std::mutex mu;
  std::shared_ptr<mvamd::NccModel> model;
  std::shared_ptr<const std::string> item;
  bool is_item = false;
};

struct MvHandleImpl {
  std::shared_ptr<ModelObject> object;
};

struct Value {
  MvValueKind kind = MV_VALUE_EMPTY;
  int64_t i = 0;
  double d = 0;
  std::string s;
  std::unique_ptr<MvHandleImpl> h;
  Value() = default;
  Value(const Value& o) : kind(o.kind), i(o.i), d(o.d), s(o.s) {
    if (o.h) h = std::make_unique<MvHandleImpl>(*o.h);
  }
  Value& operator=(const Value& o) {
    if (this != &o) {
      kind = o.kind; i = o.i; d = o.d; s = o.s;
      h = o.h ? std::make_unique<MvHandleImpl>(*o.h) : nullptr;
    }
    return *this;
  }
};

struct MvTupleImpl {
  std::vector<Value> values;
};

// Pixel storage shared by an image and the images reduced from it.
struct PixelStore {
  std::vector<uint8_t> owned;
  const uint8_t* pixels = nullptr;
  void (*clear)(void*) = nullptr;  // caller's release procedure (gen_image1_extern)
  ~PixelStore() {
    if (clear && pixels) clear(const_cast<uint8_t*>(pixels));
  }
};

// An iconic object: an empty object array, one image (with an optional
// reduced domain, image.hpp:108-113) or an array of regions.
struct MvObjectImpl {
  enum Kind { kNone, kImage, kRegion } kind = kNone;
  bool empty = true;  // no object
  // image
  size_t width = 0, height = 0;
  int pixel_bytes = 1;  // 1: byte, 2: uint2
  std::shared_ptr<PixelStore> store;
  const uint8_t* pixels = nullptr;
  std::shared_ptr<const mvamd::Region> domain;  // null: the full image
  // regions
  std::vector<std::shared_ptr<const mvamd::Region>> regions;
};

// ------------------------------------------------------------- error state
namespace {
struct LastError {
  MvStatus status = MV_STATUS_OK;
  MvErrorCode code = MV_ERROR_UNSPECIFIED;
  std::string operation, parameter, message, text;
  int64_t index = MV_ERROR_PARAMETER_INDEX_NONE;
};
thread_local LastError g_error;

void set_error(MvStatus st, MvErrorCode code, std::string op, std::string param, std::string msg, int64_t index) {
  g_error.status = st; g_error.code = code; g_error.operation = std::move(op); g_error.parameter = std::move(param);
  g_error.message = std::move(msg); g_error.index = index;
  g_error.text = g_error.operation + (g_error.parameter.empty() ? "" : "(" + g_error.parameter + ")") + ": " +
                 g_error.message;
}

// Exceptions never cross the ABI (error_state.hpp:28-57).
template <class F>
MvStatus translate(F&& f) noexcept {
  try {
    f();
    g_error = LastError{};
    return MV_STATUS_OK;
  } catch (const Error& e) {
    set_error(e.status, e.code, e.operation, e.parameter, e.message, e.index);
    return e.status;
  } catch (const std::bad_alloc&) {
    set_error(MV_STATUS_OUT_OF_MEMORY, MV_ERROR_INTERNAL_FAILURE, "", "", "out of memory", -1);
    return MV_STATUS_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    set_error(MV_STATUS_INTERNAL_ERROR, MV_ERROR_INTERNAL_FAILURE, "", "", e.what(), -1);
    return MV_STATUS_INTERNAL_ERROR;
  } catch (...) {
    set_error(MV_STATUS_INTERNAL_ERROR, MV_ERROR_INTERNAL_FAILURE, "", "", "unknown internal error", -1);
    return MV_STATUS_INTERNAL_ERROR;
  }
}

[[noreturn]] void fail(mvamd::Status st, mvamd::ErrorCode code, const char* op, const char* param, const char* msg,
                       int64_t index = -1) {
  throw Error(st, op, param, msg, index, code);
}
void required(const void* p, const char* op, const char* param) {
  if (!p) fail(mvamd::kInvalidArgument, mvamd::kErrNullArgument, op, param, "must not be null");
}

// Read-only view of an input tuple with the reference's coercions
// (tuple_input.hpp:79-133).  A null tuple is an empty input.
class Input {
 public:
  Input(MvTuple t, const char* op, const char* param) : t_(t), op_(op), param_(param) {}
  size_t size() const { return t_ ? t_->values.size() : 0; }
  const Value& at(size_t i) const {
    if (i >= size()) fail(mvamd::kInvalidArgument, mvamd::kErrIndexOutOfRange, op_, param_, "control tuple index out of range", int64_t(i));
    return t_->values[i];
  }
  [[noreturn]] void bad(mvamd::ErrorCode code, int64_t index, const char* msg) const {
    fail(code == mvamd::kErrInvalidType ? mvamd::kTypeMismatch : mvamd::kInvalidArgument, code, op_, param_, msg, index);
  }
  double number(size_t i) const {
    const Value& v = at(i);
    if (v.kind == MV_VALUE_INT64) return double(v.i);
    if (v.kind == MV_VALUE_REAL64) return v.d;
    bad(mvamd::kErrInvalidType, int64_t(i), "numeric control value required");
  }
  double scalar_number() const {
    if (size() != 1) bad(mvamd::kErrInvalidCardinality, -1, "exactly one control value required");
    const double v = number(0);
    if (!std::isfinite(v)) bad(mvamd::kErrInvalidValue, 0, "finite control value required");
    return v;
  }
  int64_t integer(size_t i) const {
    const Value& v = at(i);
    if (v.kind == MV_VALUE_INT64) return v.i;
    bad(mvamd::kErrInvalidType, int64_t(i), "integer control value required");
  }
  int64_t scalar_integer() const {
    if (size() != 1) bad(mvamd::kErrInvalidCardinality, -1, "exactly one control value required");
    return integer(0);
  }
  const std::string& string(size_t i) const {
    const Value& v = at(i);
    if (v.kind != MV_VALUE_STRING) bad(mvamd::kErrInvalidType, int64_t(i), "string control value required");
    return v.s;
  }
  MvValueKind kind(size_t i) const { return at(i).kind; }

 private:
  MvTuple t_;
  const char* op_;
  const char* param_;
};

// All outputs of one call are committed together (control_operator_outputs.hpp).
template <size_t N>
class Outputs {
 public:
  Outputs(const char* op, std::initializer_list<MvTuple*> targets) {
    size_t k = 0;
    for (MvTuple* t : targets) {
      required(t, op, "output");
      targets_[k++] = t;
    }
    for (auto& v : values_) v = std::make_unique<MvTupleImpl>();
  }
  MvTupleImpl& operator[](size_t i) { return *values_[i]; }
  void commit() {
    for (size_t i = 0; i < N; ++i) *targets_[i] = values_[i].release();
  }

 private:
  MvTuple* targets_[N];
  std::unique_ptr<MvTupleImpl> values_[N];
};

void push(MvTupleImpl& t, double v) { Value x; x.kind = MV_VALUE_REAL64; x.d = v; t.values.push_back(std::move(x)); }
void push(MvTupleImpl& t, int64_t v) { Value x; x.kind = MV_VALUE_INT64; x.i = v; t.values.push_back(std::move(x)); }
void push(MvTupleImpl& t, const char* v) { Value x; x.kind = MV_VALUE_STRING; x.s = v; t.values.push_back(std::move(x)); }

// IconicTuple::require_kind (iconic_tuple.cpp:198-210): every object of the
// array is of the kind (an empty array is of any kind).
void require_kind(MvObject o, MvObjectImpl::Kind kind, const char* op) {
  if (!o->empty && o->kind != kind)
    fail(mvamd::kTypeMismatch, mvamd::kErrInvalidType, op, "Objects", "iconic tuple contains an incompatible object kind", 0);
}

const MvObjectImpl& single_image(MvObject o, const char* op, const char* param) {
  required(o, op, param);
  require_kind(o, MvObjectImpl::kImage, op);
  if (o->empty) fail(mvamd::kInvalidArgument, mvamd::kErrInvalidCardinality, op, param, "one image required");
  return *o;
}

std::shared_ptr<mvamd::NccModel> model_of(MvTuple t, const char* op) {
  const Input view(t, op, "ModelID");
  if (view.size() != 1) view.bad(mvamd::kErrInvalidCardinality, -1, "exactly one NCC model required");
  if (view.kind(0) != MV_VALUE_HANDLE) view.bad(mvamd::kErrInvalidType, 0, "NCC model handle required");
  const Value& v = view.at(0);
  if (!v.h || !v.h->object) view.bad(mvamd::kErrInvalidType, 0, "control handle is not an NCC model");
  std::lock_guard<std::mutex> lock(v.h->object->mu);
  if (v.h->object->is_item) view.bad(mvamd::kErrInvalidType, 0, "control handle is not an NCC model");
  if (!v.h->object->model)
    fail(mvamd::kInvalidState, mvamd::kErrInvalidObjectState, op, "ModelID", "NCC model handle was cleared", 0);
  return v.h->object->model;
}

// Find pipelines.  A context owns HIP streams, events and device memory, so
// it must never be destroyed from thread-local storage: at thread or process
// exit the HIP runtime and profilers may already have torn down their own
// per-thread state.  Contexts belong to a process-wide pool instead; a thread
// borrows one per device on its first call and hands it back, without any
// HIP call, when it ends.  The next thread to need a context on that device
// takes the idle one and drains it first.  The pool is never destroyed: its
// contexts' device memory goes back to the driver with the process.
class ContextPool {
 public:
  mvamd::FindContext* take(int device) {
    {
      std::lock_guard<std::mutex> lock(mu_);
      auto& idle = idle_[device];
      if (!idle.empty()) {
        mvamd::FindContext* c = idle.back();
        idle.pop_back();
        c->reset();
        return c;
      }
    }
    return new mvamd::FindContext();  // on the calling thread's current device
  }
  void give_back(int device, mvamd::FindContext* c) {
    std::lock_guard<std::mutex> lock(mu_);
    idle_[device].push_back(c);
  }

 private:
  std::mutex mu_;
  std::map<int, std::vector<mvamd::FindContext*>> idle_;
};
ContextPool& context_pool() {
  static ContextPool* pool = new ContextPool;  // deliberately never destroyed
  return *pool;
}
struct BorrowedContexts {
  std::map<int, mvamd::FindContext*> by_device;
  ~BorrowedContexts() {
    for (auto& [device, c] : by_device) context_pool().give_back(device, c);
  }
};
mvamd::FindContext& context() {
  thread_local BorrowedContexts mine;
  int dev = 0;
  mvamd::check(hipGetDevice(&dev), "hipGetDevice");
  mvamd::FindContext*& c = mine.by_device[dev];
  if (!c) c = context_pool().take(dev);
  return *c;
}

// ---------------------------------------------------- find control schema
// The find operators (ncc_matching_api.cpp:126-161 gives the rules) take six
// value controls and NumLevels.  Each value control holds one value shared
// by every model of the call or one value per model; NumLevels holds the
// highest level or a (highest, lowest) pair, shared or per model.  The table
// below says how one value becomes a FindOptions field.  Validation runs in
// two passes so that the first fault reported is the same as mvkit's: the
// cardinality of every control in argument order, then per model the values
// by form (ranged reals, then the count, then the keyword), then the levels.
enum class ValueForm { kReal, kCount, kKeyword };
struct FindControl {
  const char* param;
  ValueForm form;
  double lo, hi;  // kReal: closed range of accepted values
  void (*store)(mvamd::FindOptions&, double);
};
constexpr double kLargest = std::numeric_limits<double>::max();
constexpr double kFullTurn = 2 * 3.14159265358979323846;
constexpr FindControl kFindControls[] = {
    {"AngleStart", ValueForm::kReal, -kLargest, kLargest, [](mvamd::FindOptions& q, double v) { q.start = v; }},
    {"AngleExtent", ValueForm::kReal, 0, kFullTurn, [](mvamd::FindOptions& q, double v) { q.extent = v; }},
    {"MinScore", ValueForm::kReal, 0, 1, [](mvamd::FindOptions& q, double v) { q.min_score = v; }},
    {"NumMatches", ValueForm::kCount, 0, 0, [](mvamd::FindOptions& q, double v) { q.num_matches = size_t(v); }},
    {"MaxOverlap", ValueForm::kReal, 0, 1, [](mvamd::FindOptions& q, double v) { q.max_overlap = v; }},
    {"SubPixel", ValueForm::kKeyword, 0, 0, [](mvamd::FindOptions& q, double v) { q.subpixel = v != 0; }},
};
constexpr size_t kValueControls = sizeof(kFindControls) / sizeof(kFindControls[0]);

// Index of model i's value in a control holding one shared value or one per model.
size_t broadcast(const Input& in, size_t i) { return in.size() == 1 ? 0 : i; }

double control_value(const Input& in, const FindControl& c, size_t i) {
  const size_t k = broadcast(in, i);
  switch (c.form) {
    case ValueForm::kReal: {
      const double x = in.number(k);
      if (!(x >= c.lo && x <= c.hi)) in.bad(mvamd::kErrInvalidValue, int64_t(k), "value lies outside the control's range");
      return x;
    }
    case ValueForm::kCount: {
      const int64_t x = in.integer(k);
      if (x < 0) in.bad(mvamd::kErrInvalidValue, int64_t(k), "a count cannot be negative");
      return double(x);
    }
    case ValueForm::kKeyword: {
      const std::string& word = in.string(k);
      if (word == "true") return 1;
      if (word != "false") in.bad(mvamd::kErrUnsupportedValue, int64_t(k), "keyword must be 'true' or 'false'");
      return 0;
    }
  }
  return 0;
}

// Where model i's highest and (optional) lowest level sit in NumLevels.
struct LevelSlots { size_t highest = 0, lowest = 0; bool paired = false; };
LevelSlots level_slots(size_t n_values, size_t models, size_t i) {
  if (n_values == models && models > 1) return {i, 0, false};     // one highest level per model
  if (n_values == 2 * models) return {2 * i, 2 * i + 1, true};    // one pair per model
  if (n_values == 2) return {0, 1, true};                         // one shared pair
  return {0, 0, false};                                           // one shared highest level
}

struct Controls {
  std::vector<mvamd::FindOptions> options;
  mvamd::Selection selection;
};
Controls find_controls(const char* op, size_t models, const MvTuple (&tuples)[kValueControls], MvTuple num_levels) {
  std::vector<Input> in;
  in.reserve(kValueControls);
  for (size_t c = 0; c < kValueControls; ++c) in.emplace_back(tuples[c], op, kFindControls[c].param);
  const Input levels(num_levels, op, "NumLevels");
  for (const Input& x : in)
    if (x.size() != 1 && x.size() != models)
      x.bad(mvamd::kErrInvalidCardinality, -1, "control needs one value or one value per model");
  const size_t nl = levels.size();
  if (nl != 1 && nl != 2 && nl != models && nl != 2 * models)
    levels.bad(mvamd::kErrInvalidCardinality, -1, "levels need one highest level or a pair, shared or per model");
  Controls out;
  out.options.resize(models);
  for (size_t i = 0; i < models; ++i) {
    mvamd::FindOptions& q = out.options[i];
    for (const ValueForm form : {ValueForm::kReal, ValueForm::kCount, ValueForm::kKeyword})
      for (size_t c = 0; c < kValueControls; ++c)
        if (kFindControls[c].form == form) kFindControls[c].store(q, control_value(in[c], kFindControls[c], i));
    const LevelSlots at = level_slots(nl, models, i);
    const int64_t highest = levels.integer(at.highest);
    const int64_t lowest = at.paired ? levels.integer(at.lowest) : 1;
    constexpr int64_t kIntMax = std::numeric_limits<int>::max();
    if (highest < 0 || highest > kIntMax || lowest < 0 || lowest > kIntMax)
      levels.bad(mvamd::kErrInvalidValue, int64_t(at.highest), "pyramid levels out of range");
    q.highest = int(highest);
    q.lowest = int(std::max<int64_t>(1, lowest));
  }
  // one-value NumMatches / MaxOverlap apply to the whole selection (mv_ncc_matching.h:96-100)
  out.selection.num_matches = in[3].size() == 1 ? int64_t(out.options.front().num_matches) : -1;
  out.selection.max_overlap = in[4].size() == 1 ? out.options.front().max_overlap : -1;
  return out;
}

// ModelIDs of find_ncc_models (ncc_matching_api.cpp:111-124): 1..4096 handles.
std::vector<std::shared_ptr<mvamd::NccModel>> models_of(MvTuple t, const char* op) {
  const Input view(t, op, "ModelIDs");
  if (!view.size() || view.size() > 4096)
    view.bad(mvamd::kErrInvalidCardinality, -1, "one model or a nonempty model array (maximum 4096) required");
  std::vector<std::shared_ptr<mvamd::NccModel>> out;
  for (size_t i = 0; i < view.size(); ++i) {
    if (view.kind(i) != MV_VALUE_HANDLE) view.bad(mvamd::kErrInvalidType, int64_t(i), "NCC model handle required");
    const Value& v = view.at(i);
    if (!v.h || !v.h->object) view.bad(mvamd::kErrInvalidType, int64_t(i), "control handle is not an NCC model");
    std::lock_guard<std::mutex> lock(v.h->object->mu);
    if (v.h->object->is_item) view.bad(mvamd::kErrInvalidType, int64_t(i), "control handle is not an NCC model");
    if (!v.h->object->model)
      fail(mvamd::kInvalidState, mvamd::kErrInvalidObjectState, op, "ModelIDs", "NCC model handle was cleared", int64_t(i));
    out.push_back(v.h->object->model);
  }
  return out;
}

mvamd::FindOptions from_ext(const MvAmdFindOptions* o) {
  required(o, "find_ncc_model", "options");
  mvamd::FindOptions q;
  q.start = o->angle_start; q.extent = o->angle_extent; q.min_score = o->min_score; q.max_overlap = o->max_overlap;
  if (o->num_matches < 0) fail(mvamd::kInvalidArgument, mvamd::kErrInvalidValue, "find_ncc_model", "NumMatches", "match count must be nonnegative");
  q.num_matches = size_t(o->num_matches);
  q.subpixel = o->subpixel != 0;
  q.highest = o->highest_level; q.lowest = std::max(1, o->lowest_level);
  q.form = o->form; q.screen = o->screen; q.profile = o->profile;
  q.track_mfma = o->vector_tracking == 0;
  if (o->max_candidates < 0 || o->max_candidates > (int64_t(1) << 24))
    fail(mvamd::kInvalidArgument, mvamd::kErrInvalidValue, "find_ncc_model", "max_candidates",
         "candidate budget must be 0 (the default) or 1..2^24");
  if (o->max_candidates > 0) q.max_candidates = size_t(o->max_candidates);
  q.graph = o->no_graph == 0;
  if (o->coarse_tile < MVAMD_COARSE_TILE_DEFAULT || o->coarse_tile >= (1 << 18))
    fail(mvamd::kInvalidArgument, mvamd::kErrInvalidValue, "find_ncc_model", "coarse_tile", "invalid coarse tile shape");
  q.coarse_tile = o->coarse_tile;
  if (q.form < MVAMD_FORM_AUTO || q.form > MVAMD_FORM_FP8_SCREEN)
    fail(mvamd::kInvalidArgument, mvamd::kErrUnsupportedValue, "find_ncc_model", "form", "unknown correlation form");
  if (o->pixel_type != MVAMD_PIXEL_BYTE && o->pixel_type != MVAMD_PIXEL_UINT2)
    fail(mvamd::kUnsupported, mvamd::kErrUnsupportedValue, "find_ncc_model", "pixel_type", "unsupported pixel type");
  return q;
}

// A device-resident search image of the pixel type the options name.
mvamd::FrameInput device_frame(const char* op, const uint8_t* image, int64_t width, int64_t height, int64_t pitch,
                               const MvAmdFindOptions* o, int64_t index = -1) {
  const int bytes = o->pixel_type == MVAMD_PIXEL_UINT2 ? 2 : 1;
  if (!image || width <= 0 || height <= 0 || pitch < width * bytes || pitch % bytes)
    fail(mvamd::kInvalidArgument, mvamd::kErrInvalidValue, op, "Image", "invalid image geometry", index);
  mvamd::FrameInput in{image, size_t(width), size_t(height), size_t(pitch)};
  in.pixel_bytes = bytes;
  return in;
}

// Rethrows algorithm errors under the C-ABI operation name (ncc_matching_api.cpp:21-26).
template <class F>
auto algorithm(const char* op, F&& f) -> decltype(f()) {
  try {
    return f();
  } catch (Error& e) {
    e.operation = op;
    throw;
  }
}

void put_handle(MvTupleImpl& t, std::shared_ptr<mvamd::NccModel> m) {
  Value v;
  v.kind = MV_VALUE_HANDLE;
  v.h = std::make_unique<MvHandleImpl>();
  v.h->object = std::make_shared<ModelObject>();
  v.h->object->model = std::move(m);
  t.values.push_back(std::move(v));
}
void put_item(MvTupleImpl& t, std::string bytes) {
  Value v;
  v.kind = MV_VALUE_HANDLE;
  v.h = std::make_unique<MvHandleImpl>();
  v.h->object = std::make_shared<ModelObject>();
  v.h->object->is_item = true;
  v.h->object->item = std::make_shared<const std::string>(std::move(bytes));
  t.values.push_back(std::move(v));
}

// One live handle of a control tuple (serialization_api.cpp:22-31).
std::shared_ptr<ModelObject> single_handle(MvTuple t, const char* op, const char* param) {
  const Input view(t, op, param);
  if (view.size() != 1) view.bad(mvamd::kErrInvalidCardinality, -1, "exactly one handle required");
  if (view.kind(0) != MV_VALUE_HANDLE) view.bad(mvamd::kErrInvalidType, 0, "handle required");
  const Value& v = view.at(0);
  std::shared_ptr<ModelObject> obj = v.h ? v.h->object : nullptr;
  bool live = false;
  if (obj) {
    std::lock_guard<std::mutex> lock(obj->mu);
    live = obj->model || obj->item;
  }
  if (!live) fail(mvamd::kInvalidState, mvamd::kErrInvalidObjectState, op, param, "handle is empty or cleared", 0);
  return obj;
}

// FileName control of read/write (ncc_matching_api.cpp:77-83).
std::string filename(MvTuple t, const char* op) {
  const Input view(t, op, "FileName");
  if (view.size() != 1) view.bad(mvamd::kErrInvalidCardinality, -1, "exactly one file name required");
  const std::string& name = view.string(0);
  if (name.empty() || name.find('\0') != std::string::npos)
    view.bad(mvamd::kErrInvalidValue, 0, "nonempty file name without NUL required");
  return name;
}

// Rethrows a codec error under the operation and parameter of a
// serialization entry point (serialization_api.cpp:14-20).
template <class F>
auto codec(const char* op, const char* param, F&& f) -> decltype(f()) {
  try {
    return f();
  } catch (Error& e) {
    e.operation = op;
    e.parameter = param;
    e.index = 0;
    throw;
  }
}
}  // namespace

// ================================================================= errors
extern "C" MvStatus mv_get_last_error(MvStringView* error) {
  if (!error) return MV_STATUS_INVALID_ARGUMENT;
  *error = {g_error.text.c_str(), g_error.text.size()};
  return MV_STATUS_OK;
}
extern "C" MvStatus mv_get_last_error_info(MvErrorInfo* e) {
  if (!e) return MV_STATUS_INVALID_ARGUMENT;
  e->status = g_error.status;
  e->code = g_error.code;
  e->operation = {g_error.operation.c_str(), g_error.operation.size()};
  e->parameter = {g_error.parameter.c_str(), g_error.parameter.size()};
  e->parameter_index = g_error.index;
  e->message = {g_error.message.c_str(), g_error.message.size()};
  return MV_STATUS_OK;
}

// ================================================================= tuples
extern "C" MvStatus mv_create_tuple(MvTuple* tuple, size_t length) {
  return translate([&] {
    required(tuple, "create_tuple", "tuple");
    auto t = std::make_unique<MvTupleImpl>();
    t->values.resize(length);
    *tuple = t.release();
  });
}
extern "C" MvStatus mv_copy_tuple(MvTuple tuple, MvTuple* clone) {
  return translate([&] {
    required(tuple, "copy_tuple", "tuple");
    required(clone, "copy_tuple", "clone");
    *clone = new MvTupleImpl(*tuple);
  });
}
extern "C" MvStatus mv_destroy_tuple(MvTuple tuple) {
  delete tuple;
  return MV_STATUS_OK;
}
extern "C" MvStatus mv_length_tuple(MvTuple tuple, size_t* size) {
  return translate([&] {
    required(tuple, "length_tuple", "tuple");
    required(size, "length_tuple", "size");
    *size = tuple->values.size();
  });
}
extern "C" MvStatus mv_resize_tuple(MvTuple tuple, size_t length) {
  return translate([&] {
    required(tuple, "resize_tuple", "tuple");
    tuple->values.resize(length);
  });
}
namespace {
Value& slot(MvTuple t, size_t index, const char* op) {
  required(t, op, "tuple");
  if (index >= t->values.size())
    fail(mvamd::kOutOfRange, mvamd::kErrIndexOutOfRange, op, "index", "tuple index out of range", int64_t(index));
  return t->values[index];
}
}  // namespace
extern "C" MvStatus mv_get_type(MvTuple tuple, size_t index, MvValueKind* kind) {
  return translate([&] { required(kind, "get_type", "kind"); *kind = slot(tuple, index, "get_type").kind; });
}
extern "C" MvStatus mv_get_i(MvTuple tuple, size_t index, int64_t* value) {
  return translate([&] {
    required(value, "get_i", "value");
    const Value& v = slot(tuple, index, "get_i");
    if (v.kind != MV_VALUE_INT64) fail(mvamd::kTypeMismatch, mvamd::kErrInvalidType, "get_i", "tuple", "integer element required", int64_t(index));
    *value = v.i;
  });
}
extern "C" MvStatus mv_get_d(MvTuple tuple, size_t index, double* value) {
  return translate([&] {
    required(value, "get_d", "value");
    const Value& v = slot(tuple, index, "get_d");
    if (v.kind == MV_VALUE_REAL64) *value = v.d;
    else if (v.kind == MV_VALUE_INT64) *value = double(v.i);
    else fail(mvamd::kTypeMismatch, mvamd::kErrInvalidType, "get_d", "tuple", "numeric element required", int64_t(index));
  });
}
extern "C" MvStatus mv_get_s(MvTuple tuple, size_t index, MvStringView* value) {
  return translate([&] {
    required(value, "get_s", "value");
    const Value& v = slot(tuple, index, "get_s");
    if (v.kind != MV_VALUE_STRING) fail(mvamd::kTypeMismatch, mvamd::kErrInvalidType, "get_s", "tuple", "string element required", int64_t(index));
    *value = {v.s.c_str(), v.s.size()};
  });
}
extern "C" MvStatus mv_get_h(MvTuple tuple, size_t index, MvHandle* handle) {
  return translate([&] {
    required(handle, "get_h", "handle");
    const Value& v = slot(tuple, index, "get_h");
    if (v.kind != MV_VALUE_HANDLE) fail(mvamd::kTypeMismatch, mvamd::kErrInvalidType, "get_h", "tuple", "handle element required", int64_t(index));
    *handle = v.h.get();
  });
}
extern "C" MvStatus mv_set_i(MvTuple tuple, int64_t value, size_t index) {
  return translate([&] { Value& v = slot(tuple, index, "set_i"); v = Value{}; v.kind = MV_VALUE_INT64; v.i = value; });
}
extern "C" MvStatus mv_set_d(MvTuple tuple, double value, size_t index) {
  return translate([&] { Value& v = slot(tuple, index, "set_d"); v = Value{}; v.kind = MV_VALUE_REAL64; v.d = value; });
}
extern "C" MvStatus mv_set_s(MvTuple tuple, MvStringView value, size_t index) {
  return translate([&] {
    if (!value.data && value.size) fail(mvamd::kInvalidArgument, mvamd::kErrNullArgument, "set_s", "value", "string data must not be null");
    Value& v = slot(tuple, index, "set_s");
    v = Value{};
    v.kind = MV_VALUE_STRING;
    v.s.assign(value.data ? value.data : "", value.size);
  });
}
extern "C" MvStatus mv_set_h(MvTuple tuple, MvHandle value, size_t index) {
  return translate([&] {
    required(value, "set_h", "value");
    Value& v = slot(tuple, index, "set_h");
    Value x;
    x.kind = MV_VALUE_HANDLE;
    x.h = std::make_unique<MvHandleImpl>(*value);
    v = std::move(x);
  });
}

// ================================================================ objects
extern "C" MvStatus mv_gen_empty_obj(MvObject* objects) {
  return translate([&] { required(objects, "gen_empty_obj", "objects"); *objects = new MvObjectImpl(); });
}
extern "C" MvStatus mv_clear_obj(MvObject objects) {
  delete objects;
  return MV_STATUS_OK;
}
namespace {
void image_controls(const char* op, MvTuple type, MvTuple width, MvTuple height, MvTuple pointer, int& bytes,
                    size_t& w, size_t& h, uintptr_t& address) {
  const Input t(type, op, "Type"), wi(width, op, "Width"), hi(height, op, "Height"), pi(pointer, op, "PixelPointer");
  if (t.size() != 1) t.bad(mvamd::kErrInvalidCardinality, -1, "exactly one pixel type required");
  const std::string& name = t.string(0);
  if (name == "byte") bytes = 1;
  else if (name == "uint2") bytes = 2;
  else fail(mvamd::kUnsupported, mvamd::kErrUnsupportedValue, op, "Type", "unsupported pixel type", 0);
  const int64_t ww = wi.scalar_integer(), hh = hi.scalar_integer();
  if (ww <= 0 || hh <= 0) fail(mvamd::kInvalidArgument, mvamd::kErrInvalidValue, op, "Width", "positive image size required");
  w = size_t(ww); h = size_t(hh);
  const int64_t addr = pi.scalar_integer();
  if (!addr) fail(mvamd::kInvalidArgument, mvamd::kErrNullArgument, op, "PixelPointer", "must not be null");
  address = uintptr_t(addr);
}
}  // namespace
extern "C" MvStatus mv_gen_image1(MvObject* Image, MvTuple Type, MvTuple Width, MvTuple Height, MvTuple PixelPointer) {
  return translate([&] {
    constexpr auto op = "gen_image1";
    required(Image, op, "Image");
    int bytes; size_t w, h; uintptr_t addr;
    image_controls(op, Type, Width, Height, PixelPointer, bytes, w, h, addr);
    auto img = std::make_unique<MvObjectImpl>();
    img->kind = MvObjectImpl::kImage;
    img->empty = false; img->width = w; img->height = h; img->pixel_bytes = bytes;
    img->store = std::make_shared<PixelStore>();
    img->store->owned.assign(reinterpret_cast<const uint8_t*>(addr),
                             reinterpret_cast<const uint8_t*>(addr) + w * h * bytes);
    img->store->pixels = img->store->owned.data();
    img->pixels = img->store->pixels;
    *Image = img.release();
  });
}
extern "C" MvStatus mv_gen_image1_extern(MvObject* Image, MvTuple Type, MvTuple Width, MvTuple Height,
                                         MvTuple PixelPointer, MvTuple ClearProc) {
  return translate([&] {
    constexpr auto op = "gen_image1_extern";
    required(Image, op, "Image");
    int bytes; size_t w, h; uintptr_t addr;
    image_controls(op, Type, Width, Height, PixelPointer, bytes, w, h, addr);
    const Input cp(ClearProc, op, "ClearProc");
    const int64_t proc = cp.size() ? cp.scalar_integer() : 0;
    auto img = std::make_unique<MvObjectImpl>();
    img->kind = MvObjectImpl::kImage;
    img->empty = false; img->width = w; img->height = h; img->pixel_bytes = bytes;
    img->store = std::make_shared<PixelStore>();
    img->store->pixels = reinterpret_cast<const uint8_t*>(addr);
    img->store->clear = reinterpret_cast<void (*)(void*)>(uintptr_t(proc));
    img->pixels = img->store->pixels;
    *Image = img.release();
  });
}

// ================================================================ regions
namespace {
std::unique_ptr<MvObjectImpl> region_object(std::vector<mvamd::Region> regions) {
  auto o = std::make_unique<MvObjectImpl>();
  o->kind = MvObjectImpl::kRegion;
  o->empty = regions.empty();
  for (auto& g : regions) o->regions.push_back(std::make_shared<const mvamd::Region>(std::move(g)));
  return o;
}
// Run coordinate of a control value: integers as they are, reals rounded by
// +0.5 and truncation (region_api.cpp:39-46).
int32_t run_coordinate(const Input& in, size_t i) {
  double v;
  if (in.kind(i) == MV_VALUE_INT64) v = double(in.integer(i));
  else v = std::trunc(in.number(i) + .5);
  if (!std::isfinite(v) || v < -2147483648.0 || v > 2147483647.0)
    in.bad(mvamd::kErrIndexOutOfRange, int64_t(i), "run coordinate exceeds int32 range");
  return int32_t(v);
}
}  // namespace

extern "C" MvStatus mv_gen_region_runs(MvObject* Region, MvTuple Row, MvTuple ColumnBegin, MvTuple ColumnEnd) {
  return translate([&] {
    constexpr auto op = "gen_region_runs";
    required(Region, op, "Region");
    const Input rows(Row, op, "Row"), begins(ColumnBegin, op, "ColumnBegin"), ends(ColumnEnd, op, "ColumnEnd");
    const size_t n = rows.size();
    if (begins.size() != n) begins.bad(mvamd::kErrInvalidCardinality, -1, "equal run coordinate lengths required");
    if (ends.size() != n) ends.bad(mvamd::kErrInvalidCardinality, -1, "equal run coordinate lengths required");
    std::vector<mvamd::RegionRun> runs;
    runs.reserve(n);
    for (size_t i = 0; i < n; ++i) {
      const int32_t b = run_coordinate(begins, i), e = run_coordinate(ends, i);
      if (b > e) ends.bad(mvamd::kErrInvalidValue, int64_t(i), "run end must not precede its beginning");
      runs.push_back({run_coordinate(rows, i), b, e});
    }
    std::vector<mvamd::Region> one;
    one.push_back(algorithm(op, [&] { return mvamd::Region::from_runs(std::move(runs)); }));
    *Region = region_object(std::move(one)).release();
  });
}

extern "C" MvStatus mv_get_region_runs(MvObject Region, MvTuple* Row, MvTuple* ColumnBegin, MvTuple* ColumnEnd) {
  return translate([&] {
    constexpr auto op = "get_region_runs";
    Outputs<3> out(op, {Row, ColumnBegin, ColumnEnd});
    required(Region, op, "Region");
    require_kind(Region, MvObjectImpl::kRegion, op);
    if (Region->regions.size() > 1)
      fail(mvamd::kInvalidArgument, mvamd::kErrInvalidCardinality, op, "Region", "one region or a zero-object array required");
    if (!Region->regions.empty())
      for (const auto& r : Region->regions[0]->runs) {
        push(out[0], int64_t(r.row));
        push(out[1], int64_t(r.cb));
        push(out[2], int64_t(r.ce));
      }
    out.commit();
  });
}

extern "C" MvStatus mv_gen_rectangle1(MvObject* Rectangle, MvTuple Row1, MvTuple Column1, MvTuple Row2,
                                      MvTuple Column2) {
  return translate([&] {
    constexpr auto op = "mv_gen_rectangle1";
    required(Rectangle, op, "Rectangle");
    const Input in[4] = {Input(Row1, op, "Row1"), Input(Column1, op, "Column1"), Input(Row2, op, "Row2"),
                         Input(Column2, op, "Column2")};
    // generation_count (region_control_api.cpp:20-27)
    size_t count = 0;
    bool none = false;
    for (const auto& x : in) { count = std::max(count, x.size()); none = none || x.size() == 0; }
    if (none) count = 0;
    for (const auto& x : in)
      if (x.size() != 1 && x.size() != count)
        x.bad(mvamd::kErrInvalidCardinality, -1, "generation controls must have matching lengths or be scalar");
    std::vector<mvamd::Region> out;
    for (size_t i = 0; i < count; ++i) {
      int64_t v[4];
      for (int k = 0; k < 4; ++k) {
        const size_t j = in[k].size() == 1 ? 0 : i;
        const double x = in[k].number(j);
        if (!std::isfinite(x)) in[k].bad(mvamd::kErrInvalidValue, int64_t(j), "finite control value required");
        if (x < -2147483648.0 || x > 2147483647.0) in[k].bad(mvamd::kErrIndexOutOfRange, int64_t(j), "coordinate exceeds int32 range");
        v[k] = int64_t(std::round(x));
      }
      out.push_back(mvamd::Region::rectangle(v[0], v[1], v[2], v[3]));
    }
    *Rectangle = region_object(std::move(out)).release();
  });
}

// ================================================================ domains
extern "C" MvStatus mv_reduce_domain(MvObject Image, MvObject Region, MvObject* ImageReduced) {
  return translate([&] {
    constexpr auto op = "reduce_domain";
    required(Image, op, "Image");
    required(Region, op, "Region");
    required(ImageReduced, op, "ImageReduced");
    require_kind(Region, MvObjectImpl::kRegion, op);
    require_kind(Image, MvObjectImpl::kImage, op);
    auto o = std::make_unique<MvObjectImpl>();
    if (!Image->empty && !Region->empty) {
      // the first region (image_processing_api.cpp:247-248), clipped to the
      // image and intersected with the current domain (image.cpp:267-320)
      *o = *Image;
      auto d = Region->regions[0]->clipped(0, 0, int64_t(Image->height) - 1, int64_t(Image->width) - 1);
      if (Image->domain) d = Image->domain->intersect(d);
      o->domain = std::make_shared<const mvamd::Region>(std::move(d));
    }
    *ImageReduced = o.release();
  });
}

extern "C" MvStatus mv_full_domain(MvObject Image, MvObject* ImageFull) {
  return translate([&] {
    constexpr auto op = "full_domain";
    required(Image, op, "Image");
    required(ImageFull, op, "ImageFull");
    require_kind(Image, MvObjectImpl::kImage, op);
    auto o = std::make_unique<MvObjectImpl>(*Image);
    o->domain.reset();
    *ImageFull = o.release();
  });
}

extern "C" MvStatus mv_get_domain(MvObject Image, MvObject* Domain) {
  return translate([&] {
    constexpr auto op = "get_domain";
    required(Image, op, "Image");
    required(Domain, op, "Domain");
    require_kind(Image, MvObjectImpl::kImage, op);
    std::vector<mvamd::Region> out;
    if (!Image->empty)
      out.push_back(Image->domain ? *Image->domain
                                  : mvamd::Region::rectangle(0, 0, int64_t(Image->height) - 1, int64_t(Image->width) - 1));
    *Domain = region_object(std::move(out)).release();
  });
}

// ================================================================ handles
extern "C" MvStatus mv_clear_handle(MvTuple Handle) {
  return translate([&] {
    const Input view(Handle, "clear_handle", "Handle");
    for (size_t i = 0; i < view.size(); ++i) {
      const Value& v = view.at(i);
      if (v.kind != MV_VALUE_HANDLE || !v.h || !v.h->object) view.bad(mvamd::kErrInvalidType, int64_t(i), "handle required");
      std::lock_guard<std::mutex> lock(v.h->object->mu);
      v.h->object->model.reset();
      v.h->object->item.reset();
    }
  });
}
// The semantic type of a control tuple (handle_api.cpp:67-100): integer, real
// or string when every element has that kind; a tuple of live handles of one
// kind names that kind (ncc_model, serialized_item); handles that are cleared
// or of mixed kinds are "handle"; an empty or mixed tuple is "any".
extern "C" MvStatus mv_tuple_sem_type(MvTuple T, MvTuple* SemType) {
  return translate([&] {
    constexpr auto op = "tuple_sem_type";
    Outputs<1> out(op, {SemType});
    const Input view(T, op, "T");
    const char* name = "any";
    if (view.size()) {
      const MvValueKind kind = view.at(0).kind;
      bool same = true;
      for (size_t i = 1; i < view.size() && same; ++i) same = view.at(i).kind == kind;
      if (same && kind == MV_VALUE_INT64) name = "integer";
      else if (same && kind == MV_VALUE_REAL64) name = "real";
      else if (same && kind == MV_VALUE_STRING) name = "string";
      else if (same && kind == MV_VALUE_HANDLE) {
        int type = -1;  // 0: model, 1: serialized item, 2: cleared or mixed
        for (size_t i = 0; i < view.size() && type != 2; ++i) {
          const MvHandleImpl* h = view.at(i).h.get();
          int t = 2;
          if (h && h->object) {
            std::lock_guard<std::mutex> lock(h->object->mu);
            if (h->object->is_item ? bool(h->object->item) : bool(h->object->model)) t = h->object->is_item ? 1 : 0;
          }
          type = (type < 0 || type == t) ? t : 2;
        }
        name = type == 0 ? "ncc_model" : type == 1 ? "serialized_item" : "handle";
      }
    }
    push(out[0], name);
    out.commit();
  });
}

// ============================================================ NCC operators
namespace {
bool have_device() {
  static const bool present = [] {
    int n = 0;
    return hipGetDeviceCount(&n) == hipSuccess && n > 0;
  }();
  return present;
}

// create_ncc_model's controls (the rules of ncc_matching_api.cpp:29-66).
// Every control holds exactly one value; NumLevels and AngleStep may hold the
// keyword "auto" instead of a number, Metric one of the metric names.  A
// control is read in table order, so the first fault reported is mvkit's.
struct CreateControl {
  const char* param;
  bool may_be_auto;
  bool integral;       // the value must be an integer element
  double lo, hi;       // closed range of accepted values
};
constexpr CreateControl kLevelsControl{"NumLevels", true, true, 0, 16};
constexpr CreateControl kStartControl{"AngleStart", false, false, -kLargest, kLargest};
constexpr CreateControl kExtentControl{"AngleExtent", false, false, -kLargest, kLargest};
constexpr CreateControl kStepControl{"AngleStep", true, false, 0, 3.14159265358979323846 / 16};
// Metric names in the order of CreateParams::absolute (false, true).
constexpr const char* kMetrics[] = {"use_polarity", "ignore_global_polarity"};

// The control's number, or nothing for "auto".
std::optional<double> create_value(MvTuple t, const char* op, const CreateControl& c) {
  const Input in(t, op, c.param);
  if (in.size() != 1) in.bad(mvamd::kErrInvalidCardinality, -1, "control takes a single value");
  if (c.may_be_auto && in.kind(0) == MV_VALUE_STRING) {
    if (in.string(0) != "auto") in.bad(mvamd::kErrUnsupportedValue, 0, "the only keyword accepted here is 'auto'");
    return std::nullopt;
  }
  const double x = c.integral ? double(in.integer(0)) : in.number(0);
  if (!std::isfinite(x)) in.bad(mvamd::kErrInvalidValue, 0, "value is not a finite number");
  if (x < c.lo || x > c.hi) in.bad(mvamd::kErrInvalidValue, 0, "value lies outside the control's range");
  return x;
}

mvamd::CreateParams create_controls(const char* op, MvTuple levels, MvTuple start, MvTuple extent, MvTuple step,
                                    MvTuple metric) {
  mvamd::CreateParams p;
  if (const auto n = create_value(levels, op, kLevelsControl)) p.levels = int(*n);
  p.start = *create_value(start, op, kStartControl);
  p.extent = *create_value(extent, op, kExtentControl);
  if (const auto s = create_value(step, op, kStepControl)) p.step = *s;
  const Input m(metric, op, "Metric");
  if (m.size() != 1) m.bad(mvamd::kErrInvalidCardinality, -1, "control takes a single value");
  const std::string& name = m.string(0);
  const auto* hit = std::find_if(std::begin(kMetrics), std::end(kMetrics), [&](const char* k) { return name == k; });
  if (hit == std::end(kMetrics)) m.bad(mvamd::kErrUnsupportedValue, 0, "not an NCC metric name");
  p.absolute = hit != std::begin(kMetrics);
  return p;
}

// Where the rotated rasters of a new model are built (MVAMD_CREATE_*).
bool build_on_device(int32_t builder, const char* op) {
  switch (builder) {
    case MVAMD_CREATE_HOST: return false;
    case MVAMD_CREATE_AUTO: return have_device();
    case MVAMD_CREATE_DEVICE:
      if (!have_device()) fail(mvamd::kUnsupported, mvamd::kErrUnsupportedValue, op, "builder", "no HIP device present");
      return true;
    default: fail(mvamd::kInvalidArgument, mvamd::kErrUnsupportedValue, op, "builder", "unknown raster builder");
  }
}

MvStatus create_model_handle(MvObject Template, MvTuple NumLevels, MvTuple AngleStart, MvTuple AngleExtent,
                             MvTuple AngleStep, MvTuple Metric, int32_t builder, MvTuple* ModelID) {
  return translate([&] {
    constexpr auto op = "create_ncc_model";
    Outputs<1> out(op, {ModelID});
    const bool on_device = build_on_device(builder, op);
    const MvObjectImpl& tmpl = single_image(Template, op, "Template");
    const mvamd::CreateParams p = create_controls(op, NumLevels, AngleStart, AngleExtent, AngleStep, Metric);
    const size_t pitch = tmpl.width * size_t(tmpl.pixel_bytes);
    std::shared_ptr<mvamd::NccModel> model = algorithm(op, [&]() -> std::unique_ptr<mvamd::NccModel> {
      if (!on_device)
        return mvamd::create_model(tmpl.pixels, tmpl.pixel_bytes, tmpl.width, tmpl.height, pitch, p, tmpl.domain.get());
      // the O(area) first level on the host, the rotated rasters on the device
      mvamd::FirstLevel first = mvamd::create_first_level(tmpl.pixels, tmpl.pixel_bytes, tmpl.width, tmpl.height,
                                                          pitch, p, tmpl.domain.get());
      return context().rebuild(std::move(first.level), p, first.orow, first.ocol);
    });
    put_handle(out[0], std::move(model));
    out.commit();
  });
}
}  // namespace

extern "C" MvStatus mv_create_ncc_model(MvObject Template, MvTuple NumLevels, MvTuple AngleStart, MvTuple AngleExtent,
                                        MvTuple AngleStep, MvTuple Metric, MvTuple* ModelID) {
  return create_model_handle(Template, NumLevels, AngleStart, AngleExtent, AngleStep, Metric, MVAMD_CREATE_AUTO,
                             ModelID);
}

extern "C" MvStatus mvamd_create_ncc_model(MvObject Template, MvTuple NumLevels, MvTuple AngleStart,
                                           MvTuple AngleExtent, MvTuple AngleStep, MvTuple Metric, int32_t builder,
                                           MvTuple* ModelID) {
  return create_model_handle(Template, NumLevels, AngleStart, AngleExtent, AngleStep, Metric, builder, ModelID);
}

extern "C" MvStatus mv_get_ncc_model_params(MvTuple ModelID, MvTuple* NumLevels, MvTuple* AngleStart,
                                            MvTuple* AngleExtent, MvTuple* AngleStep, MvTuple* Metric) {
  return translate([&] {
    constexpr auto op = "get_ncc_model_params";
    Outputs<5> out(op, {NumLevels, AngleStart, AngleExtent, AngleStep, Metric});
    const auto m = model_of(ModelID, op);
    push(out[0], int64_t(m->levels.size()));
    push(out[1], m->params.start);
    push(out[2], m->params.extent);
    push(out[3], m->params.extent == 0 ? 0. : m->params.step);
    push(out[4], m->params.absolute ? "ignore_global_polarity" : "use_polarity");
    out.commit();
  });
}

extern "C" MvStatus mv_find_ncc_model(MvObject Image, MvTuple ModelID, MvTuple AngleStart, MvTuple AngleExtent,
                                      MvTuple MinScore, MvTuple NumMatches, MvTuple MaxOverlap, MvTuple SubPixel,
                                      MvTuple NumLevels, MvTuple* Row, MvTuple* Column, MvTuple* Angle, MvTuple* Score) {
  return translate([&] {
    constexpr auto op = "find_ncc_model";
    Outputs<4> out(op, {Row, Column, Angle, Score});
    required(Image, op, "Image");
    require_kind(Image, MvObjectImpl::kImage, op);
    const auto m = model_of(ModelID, op);
    const mvamd::FindOptions q =
        find_controls(op, 1, {AngleStart, AngleExtent, MinScore, NumMatches, MaxOverlap, SubPixel}, NumLevels)
            .options.front();
    if (Image->empty) { out.commit(); return; }
    const size_t pitch = Image->width * size_t(Image->pixel_bytes);
    const auto found = algorithm(op, [&] {
      return context().find_host(*m, Image->pixels, Image->width, Image->height, pitch, q, nullptr, Image->pixel_bytes,
                                 Image->domain.get());
    });
    // HALCON pixel-edge convention of the translation (ncc_matching_api.cpp:171-176)
    for (const auto& p : found) {
      const double c = std::cos(p.angle), s = std::sin(p.angle);
      push(out[0], p.row + .5 * (1 - c + s));
      push(out[1], p.col + .5 * (1 - c - s));
      push(out[2], p.angle);
      push(out[3], p.score);
    }
    out.commit();
  });
}

extern "C" MvStatus mv_find_ncc_models(MvObject Image, MvTuple ModelIDs, MvTuple AngleStart, MvTuple AngleExtent,
                                       MvTuple MinScore, MvTuple NumMatches, MvTuple MaxOverlap, MvTuple SubPixel,
                                       MvTuple NumLevels, MvTuple* Row, MvTuple* Column, MvTuple* Angle, MvTuple* Score,
                                       MvTuple* Model) {
  return translate([&] {
    constexpr auto op = "find_ncc_models";
    Outputs<5> out(op, {Row, Column, Angle, Score, Model});
    required(Image, op, "Image");
    require_kind(Image, MvObjectImpl::kImage, op);
    const auto ms = models_of(ModelIDs, op);
    const Controls c =
        find_controls(op, ms.size(), {AngleStart, AngleExtent, MinScore, NumMatches, MaxOverlap, SubPixel}, NumLevels);
    if (Image->empty) { out.commit(); return; }
    std::vector<const mvamd::NccModel*> ptrs;
    for (const auto& m : ms) ptrs.push_back(m.get());
    const size_t pitch = Image->width * size_t(Image->pixel_bytes);
    const auto found = algorithm(op, [&] {
      return context().find_models_host(ptrs.data(), c.options.data(), ptrs.size(), Image->pixels, Image->width,
                                        Image->height, pitch, c.selection, true, Image->pixel_bytes,
                                        Image->domain.get());
    });
    for (const auto& m : found) {
      const double cs = std::cos(m.pose.angle), sn = std::sin(m.pose.angle);
      push(out[0], m.pose.row + .5 * (1 - cs + sn));
      push(out[1], m.pose.col + .5 * (1 - cs - sn));
      push(out[2], m.pose.angle);
      push(out[3], m.pose.score);
      push(out[4], int64_t(m.model));
    }
    out.commit();
  });
}

extern "C" MvStatus mv_write_ncc_model(MvTuple ModelID, MvTuple FileName) {
  return translate([&] {
    constexpr auto op = "write_ncc_model";
    const auto m = model_of(ModelID, op);
    const std::string name = filename(FileName, op);
    const std::string bytes = algorithm(op, [&] { return mvamd::serialize_model(*m); });
    // core/model_io.hpp:14-20
    std::FILE* f = std::fopen(name.c_str(), "wb");
    if (!f) fail(mvamd::kIoError, mvamd::kErrIoFailure, op, "", "cannot open model for writing");
    const bool ok = std::fwrite(bytes.data(), 1, bytes.size(), f) == bytes.size() && std::fflush(f) == 0;
    const bool closed = std::fclose(f) == 0;
    if (!ok || !closed) fail(mvamd::kIoError, mvamd::kErrIoFailure, op, "", "cannot write model file");
  });
}

extern "C" MvStatus mv_read_ncc_model(MvTuple FileName, MvTuple* ModelID) {
  return translate([&] {
    constexpr auto op = "read_ncc_model";
    Outputs<1> out(op, {ModelID});
    const std::string name = filename(FileName, op);
    // core/model_io.cpp:13-49: the whole file, bounded
    constexpr long kLimit = 64L * 1024 * 1024 + 128;
    std::FILE* f = std::fopen(name.c_str(), "rb");
    if (!f) fail(mvamd::kIoError, mvamd::kErrIoFailure, op, "", "cannot open model for reading");
    struct Closer {
      std::FILE* f;
      ~Closer() { std::fclose(f); }
    } closer{f};
    long size = -1;
    if (std::fseek(f, 0, SEEK_END) == 0) size = std::ftell(f);
    if (size < 0 || size > kLimit) fail(mvamd::kLimitExceeded, mvamd::kErrInvalidValue, op, "", "model file exceeds size limit");
    std::string bytes(size_t(size), '\0');
    if (std::fseek(f, 0, SEEK_SET) != 0 || (size && std::fread(bytes.data(), 1, bytes.size(), f) != bytes.size()))
      fail(mvamd::kIoError, mvamd::kErrIoFailure, op, "", "cannot read model file");
    std::shared_ptr<mvamd::NccModel> m = algorithm(op, [&] { return mvamd::deserialize_model(bytes); });
    put_handle(out[0], std::move(m));
    out.commit();
  });
}

// ======================================================= serialized items
extern "C" MvStatus mv_serialize_handle(MvTuple Handle, MvTuple* SerializedItem) {
  return translate([&] {
    constexpr auto op = "serialize_handle";
    Outputs<1> out(op, {SerializedItem});
    const auto obj = single_handle(Handle, op, "Handle");
    std::shared_ptr<mvamd::NccModel> m;
    {
      std::lock_guard<std::mutex> lock(obj->mu);
      if (!obj->is_item) m = obj->model;
    }
    if (!m) fail(mvamd::kUnsupported, mvamd::kErrUnsupportedValue, op, "Handle", "handle type does not support serialization", 0);
    put_item(out[0], codec(op, "Handle", [&] { return mvamd::serialize_model(*m); }));
    out.commit();
  });
}

extern "C" MvStatus mv_deserialize_handle(MvTuple SerializedItem, MvTuple* Handle) {
  return translate([&] {
    constexpr auto op = "deserialize_handle";
    Outputs<1> out(op, {Handle});
    const auto obj = single_handle(SerializedItem, op, "SerializedItem");
    std::shared_ptr<const std::string> bytes;
    {
      std::lock_guard<std::mutex> lock(obj->mu);
      if (obj->is_item) bytes = obj->item;
    }
    if (!bytes) fail(mvamd::kTypeMismatch, mvamd::kErrInvalidType, op, "SerializedItem", "serialized item required", 0);
    if (bytes->compare(0, 10, "MVKIT_NCC ") != 0)
      fail(mvamd::kDecodeError, mvamd::kErrIoFailure, op, "SerializedItem", "unknown serialized model signature", 0);
    std::shared_ptr<mvamd::NccModel> m = codec(op, "SerializedItem", [&] { return mvamd::deserialize_model(*bytes); });
    put_handle(out[0], std::move(m));
    out.commit();
  });
}

extern "C" MvStatus mv_create_serialized_item_ptr(MvTuple Pointer, MvTuple Size, MvTuple Copy,
                                                  MvTuple* SerializedItemHandle) {
  return translate([&] {
    constexpr auto op = "create_serialized_item_ptr";
    Outputs<1> out(op, {SerializedItemHandle});
    const Input pointer(Pointer, op, "Pointer"), size(Size, op, "Size"), copy(Copy, op, "Copy");
    const auto address = static_cast<uintptr_t>(pointer.scalar_integer());
    const int64_t length = size.scalar_integer();
    if (copy.size() != 1) copy.bad(mvamd::kErrInvalidCardinality, -1, "one copy mode required");
    if (copy.string(0) != "true")
      fail(mvamd::kUnsupported, mvamd::kErrUnsupportedValue, op, "Copy", "only owned copy mode true is supported", 0);
    if (length <= 0) size.bad(mvamd::kErrInvalidValue, 0, "positive byte count required");
    constexpr uint64_t kMaxBytes = 256ull * 1024 * 1024 + 128;
    if (uint64_t(length) > kMaxBytes)
      fail(mvamd::kLimitExceeded, mvamd::kErrInvalidValue, op, "Size", "serialized item exceeds size limit", 0);
    if (!address || address > std::numeric_limits<uintptr_t>::max() - size_t(length))
      pointer.bad(mvamd::kErrInvalidValue, 0, "non-null address without range overflow required");
    put_item(out[0], std::string(reinterpret_cast<const char*>(address), size_t(length)));
    out.commit();
  });
}

extern "C" MvStatus mv_get_serialized_item_ptr(MvTuple SerializedItemHandle, MvTuple* Pointer, MvTuple* Size) {
  return translate([&] {
    constexpr auto op = "get_serialized_item_ptr";
    Outputs<2> out(op, {Pointer, Size});
    const auto obj = single_handle(SerializedItemHandle, op, "SerializedItemHandle");
    std::shared_ptr<const std::string> bytes;
    {
      std::lock_guard<std::mutex> lock(obj->mu);
      if (obj->is_item) bytes = obj->item;
    }
    if (!bytes)
      fail(mvamd::kTypeMismatch, mvamd::kErrInvalidType, op, "SerializedItemHandle", "serialized item required", 0);
    // the bytes live as long as the item (its handle object)
    push(out[0], int64_t(reinterpret_cast<intptr_t>(bytes->data())));
    push(out[1], int64_t(bytes->size()));
    out.commit();
  });
}

// ============================================================ extensions
namespace {
void fill_stats(const mvamd::FindTrace& t, MvAmdFindStats* s) {
  if (!s) return;
  s->n_coarse = int32_t(t.coarse.size());
  s->n_prepared = int32_t(t.prepared.size());
  s->top = t.top; s->finest = t.finest; s->angle_stride = t.angle_stride;
  std::copy(t.storage, t.storage + 4, s->storage);
  s->t_coarse_us = t.t_coarse_us; s->t_track_us = t.t_track_us; s->t_refine_us = t.t_refine_us;
  s->t_total_us = t.t_total_us;
  for (int k = 0; k < MVAMD_STAGE_COUNT; ++k) s->kernel_us[k] = t.kernel_us[k];
  s->coarse_flops = t.coarse_flops;
  s->track_flops = t.track_flops;
  s->track_mfma_flops = t.track_mfma_flops;
  s->coarse_form = t.coarse_form;
  s->coarse_tile = t.coarse_tile;
}
void put_poses(const std::vector<mvamd::Pose>& v, double* dst, size_t cap) {
  for (size_t i = 0; i < v.size() && i < cap; ++i) {
    dst[4 * i] = v[i].row; dst[4 * i + 1] = v[i].col; dst[4 * i + 2] = v[i].angle; dst[4 * i + 3] = v[i].score;
  }
}
}  // namespace

extern "C" MvStatus mvamd_find_ncc_model_device(MvTuple ModelID, const uint8_t* d_image, int64_t width, int64_t height,
                                                int64_t pitch, const MvAmdFindOptions* options, double* results,
                                                size_t cap, size_t* count, MvAmdFindStats* stats) {
  return translate([&] {
    constexpr auto op = "find_ncc_model";
    required(d_image, op, "Image");
    required(count, op, "count");
    const auto m = model_of(ModelID, op);
    const mvamd::FindOptions q = from_ext(options);
    const mvamd::FrameInput in = device_frame(op, d_image, width, height, pitch, options);
    mvamd::FindTrace trace;
    const auto found = algorithm(op, [&] {
      return std::move(context().find_batch(*m, &in, 1, q, stats ? &trace : nullptr).front());
    });
    if (results) put_poses(found, results, cap);
    *count = found.size();
    fill_stats(trace, stats);
  });
}

extern "C" MvStatus mvamd_find_ncc_model_image(MvObject Image, MvTuple ModelID, const MvAmdFindOptions* options,
                                               double* results, size_t cap, size_t* count, MvAmdFindStats* stats) {
  return translate([&] {
    constexpr auto op = "find_ncc_model";
    required(Image, op, "Image");
    require_kind(Image, MvObjectImpl::kImage, op);
    required(count, op, "count");
    const auto m = model_of(ModelID, op);
    const mvamd::FindOptions q = from_ext(options);
    mvamd::FindTrace trace;
    std::vector<mvamd::Pose> found;
    if (!Image->empty) {
      const size_t pitch = Image->width * size_t(Image->pixel_bytes);
      found = algorithm(op, [&] {
        return context().find_host(*m, Image->pixels, Image->width, Image->height, pitch, q, stats ? &trace : nullptr,
                                   Image->pixel_bytes, Image->domain.get());
      });
    }
    if (results) put_poses(found, results, cap);
    *count = found.size();
    fill_stats(trace, stats);
  });
}

extern "C" MvStatus mvamd_find_ncc_models_device(MvTuple ModelIDs, const uint8_t* d_image, int64_t width,
                                                 int64_t height, int64_t pitch, const MvAmdFindOptions* options,
                                                 size_t n_options, int64_t selection_matches,
                                                 double selection_overlap, double* results, size_t cap,
                                                 size_t* count) {
  return translate([&] {
    constexpr auto op = "find_ncc_models";
    required(d_image, op, "Image");
    required(options, op, "options");
    required(count, op, "count");
    const auto ms = models_of(ModelIDs, op);
    if (n_options != 1 && n_options != ms.size())
      fail(mvamd::kInvalidArgument, mvamd::kErrInvalidCardinality, op, "options", "one value or one value per model required");
    std::vector<mvamd::FindOptions> qs;
    std::vector<const mvamd::NccModel*> ptrs;
    for (size_t i = 0; i < ms.size(); ++i) {
      qs.push_back(from_ext(options + (n_options == 1 ? 0 : i)));
      qs.back().profile = 0;
      ptrs.push_back(ms[i].get());
    }
    mvamd::Selection sel;
    sel.num_matches = selection_matches;
    sel.max_overlap = selection_overlap;
    for (size_t i = 1; i < n_options; ++i)
      if (options[i].pixel_type != options[0].pixel_type)
        fail(mvamd::kInvalidArgument, mvamd::kErrInvalidValue, op, "options", "one pixel type for the image", int64_t(i));
    const mvamd::FrameInput in = device_frame(op, d_image, width, height, pitch, options);
    const auto found = algorithm(op, [&] {
      return context().find_models(ptrs.data(), qs.data(), ptrs.size(), in, sel, true);
    });
    for (size_t i = 0; results && i < found.size() && i < cap; ++i) {
      double* d = results + 5 * i;
      d[0] = found[i].pose.row; d[1] = found[i].pose.col; d[2] = found[i].pose.angle; d[3] = found[i].pose.score;
      d[4] = double(found[i].model);
    }
    *count = found.size();
  });
}

extern "C" MvStatus mvamd_find_ncc_models_batch(MvTuple ModelIDs, const uint8_t* const* d_images,
                                                const int64_t* widths, const int64_t* heights, const int64_t* pitches,
                                                size_t n_images, const MvAmdFindOptions* options, size_t n_options,
                                                int64_t selection_matches, double selection_overlap,
                                                int32_t coarse_queue, double* results, size_t cap, size_t* counts,
                                                MvAmdBatchStats* stats) {
  return translate([&] {
    constexpr auto op = "find_ncc_models";
    required(d_images, op, "Images");
    required(widths, op, "widths");
    required(heights, op, "heights");
    required(pitches, op, "pitches");
    required(options, op, "options");
    required(counts, op, "counts");
    if (n_images > NCC_MAX_BATCH)
      fail(mvamd::kLimitExceeded, mvamd::kErrInvalidValue, op, "Images", "too many images for one batch");
    const auto ms = models_of(ModelIDs, op);
    if (n_options != 1 && n_options != ms.size())
      fail(mvamd::kInvalidArgument, mvamd::kErrInvalidCardinality, op, "options", "one value or one value per model required");
    std::vector<mvamd::FindOptions> qs;
    std::vector<const mvamd::NccModel*> ptrs;
    for (size_t i = 0; i < ms.size(); ++i) {
      qs.push_back(from_ext(options + (n_options == 1 ? 0 : i)));
      qs.back().profile = 0;
      ptrs.push_back(ms[i].get());
    }
    for (size_t i = 1; i < n_options; ++i)
      if (options[i].pixel_type != options[0].pixel_type)
        fail(mvamd::kInvalidArgument, mvamd::kErrInvalidValue, op, "options", "one pixel type for the images", int64_t(i));
    std::vector<mvamd::FrameInput> frames;
    for (size_t k = 0; k < n_images; ++k)
      frames.push_back(device_frame(op, d_images[k], widths[k], heights[k], pitches[k], options, int64_t(k)));
    mvamd::Selection sel;
    sel.num_matches = selection_matches;
    sel.max_overlap = selection_overlap;
    mvamd::BatchStats bs;
    const auto found = algorithm(op, [&] {
      return context().find_models_frames(ptrs.data(), qs.data(), ptrs.size(), frames.data(), n_images, sel, true,
                                          coarse_queue != 0, stats ? &bs : nullptr);
    });
    if (stats) {
      stats->coarse_queue_us = bs.coarse_queue_us;
      stats->coarse_flops = bs.coarse_flops;
      stats->total_us = bs.total_us;
      stats->tune_us = bs.tune_us;
      stats->queries = bs.queries;
      stats->queue_queries = bs.queue_queries;
      stats->candidates = bs.candidates;
      stats->tuned = bs.tuned;
      stats->refined = bs.refined;
      stats->refine_rounds = bs.refine_rounds;
    }
    for (size_t k = 0; k < n_images; ++k) {
      for (size_t i = 0; results && i < found[k].size() && i < cap; ++i) {
        double* d = results + 5 * (cap * k + i);
        const auto& m = found[k][i];
        d[0] = m.pose.row; d[1] = m.pose.col; d[2] = m.pose.angle; d[3] = m.pose.score; d[4] = double(m.model);
      }
      counts[k] = found[k].size();
    }
  });
}

// ============================================================ request loop
extern "C" MvStatus mvamd_server_create(const MvAmdServerConfig* config, void** server) {
  return translate([&] {
    constexpr auto op = "server_create";
    required(config, op, "config");
    required(server, op, "server");
    mvamd::ServeConfig c;
    if (config->workers) c.workers = config->workers;
    if (config->max_batch) c.max_batch = size_t(std::max(config->max_batch, 0));
    if (config->hbm_budget > 0) c.hbm_budget = size_t(config->hbm_budget);
    if (config->queue_limit > 0) c.queue_limit = size_t(config->queue_limit);
    c.coarse_queue = config->no_coarse_queue == 0;
    if (config->workers < 0 || config->max_batch < 0)
      fail(mvamd::kInvalidArgument, mvamd::kErrInvalidValue, op, "config", "negative worker or batch count");
    *server = algorithm(op, [&] { return new mvamd::NccServer(c); });
  });
}

extern "C" MvStatus mvamd_server_destroy(void* server) {
  return translate([&] { delete static_cast<mvamd::NccServer*>(server); });
}

extern "C" MvStatus mvamd_server_submit(void* server, const uint8_t* image, int64_t width, int64_t height,
                                        int64_t pitch, MvTuple ModelIDs, const MvAmdFindOptions* options,
                                        size_t n_options, int64_t selection_matches, double selection_overlap,
                                        uint64_t* request) {
  return translate([&] {
    constexpr auto op = "find_ncc_models";
    required(server, op, "server");
    required(image, op, "Image");
    required(options, op, "options");
    required(request, op, "request");
    const auto ms = models_of(ModelIDs, op);
    if (n_options != 1 && n_options != ms.size())
      fail(mvamd::kInvalidArgument, mvamd::kErrInvalidCardinality, op, "options", "one value or one value per model required");
    std::vector<mvamd::FindOptions> qs;
    for (size_t i = 0; i < n_options; ++i) {
      qs.push_back(from_ext(options + i));
      qs.back().profile = 0;
      if (options[i].pixel_type != options[0].pixel_type)
        fail(mvamd::kInvalidArgument, mvamd::kErrInvalidValue, op, "options", "one pixel type for the image", int64_t(i));
    }
    const int bytes = options->pixel_type == MVAMD_PIXEL_UINT2 ? 2 : 1;
    if (width <= 0 || height <= 0 || pitch < width * bytes)
      fail(mvamd::kInvalidArgument, mvamd::kErrInvalidValue, op, "Image", "invalid image geometry");
    mvamd::Selection sel;
    sel.num_matches = selection_matches;
    sel.max_overlap = selection_overlap;
    *request = algorithm(op, [&] {
      return static_cast<mvamd::NccServer*>(server)->submit(image, size_t(width), size_t(height), size_t(pitch), bytes,
                                                            ms, std::move(qs), sel);
    });
  });
}

extern "C" MvStatus mvamd_server_wait(void* server, uint64_t request, double* results, size_t cap, size_t* count,
                                      MvAmdRequestTimes* times) {
  return translate([&] {
    constexpr auto op = "find_ncc_models";
    required(server, op, "server");
    required(count, op, "count");
    mvamd::ServeTimes t;
    const auto found = algorithm(op, [&] { return static_cast<mvamd::NccServer*>(server)->wait(request, &t); });
    for (size_t i = 0; results && i < found.size() && i < cap; ++i) {
      double* d = results + 5 * i;
      d[0] = found[i].pose.row; d[1] = found[i].pose.col; d[2] = found[i].pose.angle; d[3] = found[i].pose.score;
      d[4] = double(found[i].model);
    }
    *count = found.size();
    if (times) {
      times->admitted_us = t.admitted_us; times->started_us = t.started_us; times->finished_us = t.finished_us;
      times->bytes = t.bytes; times->batch_size = t.batch_size; times->worker = t.worker;
    }
  });
}

namespace {
// Submitted batches of this thread, by ticket (device << 56 | context ticket).
struct PendingBatch {
  std::shared_ptr<mvamd::NccModel> model;  // kept alive until the wait
  std::vector<mvamd::FindTrace> traces;
  size_t n = 0;
  int device = 0;
  uint64_t local = 0;
};
std::map<uint64_t, PendingBatch>& pending_batches() {
  thread_local std::map<uint64_t, PendingBatch> pending;
  return pending;
}
void fill_batch_stats(const std::vector<mvamd::FindTrace>& traces, MvAmdFindStats* stats) {
  fill_stats(traces[0], stats);
  double flops = 0, tflops = 0, mflops = 0;
  for (const auto& t : traces) { flops += t.coarse_flops; tflops += t.track_flops; mflops += t.track_mfma_flops; }
  stats->coarse_flops = flops;
  stats->track_flops = tflops;
  stats->track_mfma_flops = mflops;
}
}  // namespace

extern "C" MvStatus mvamd_find_ncc_model_batch(MvTuple ModelID, const uint8_t* const* d_images, const int64_t* widths,
                                               const int64_t* heights, const int64_t* pitches, size_t n,
                                               const MvAmdFindOptions* options, double* results, size_t cap,
                                               size_t* counts, MvAmdFindStats* stats) {
  return translate([&] {
    constexpr auto op = "find_ncc_model";
    required(d_images, op, "Images");
    required(widths, op, "widths");
    required(heights, op, "heights");
    required(pitches, op, "pitches");
    required(counts, op, "counts");
    required(options, op, "options");
    std::vector<mvamd::FrameInput> frames;
    for (size_t k = 0; k < n; ++k)
      frames.push_back(device_frame(op, d_images[k], widths[k], heights[k], pitches[k], options, int64_t(k)));
    const auto m = model_of(ModelID, op);
    const mvamd::FindOptions q = from_ext(options);
    std::vector<mvamd::FindTrace> traces(stats ? n : 0);
    const auto found = algorithm(op, [&] {
      return context().find_batch(*m, frames.data(), n, q, stats ? traces.data() : nullptr);
    });
    for (size_t k = 0; k < n; ++k) {
      if (results) put_poses(found[k], results + 4 * cap * k, cap);
      counts[k] = found[k].size();
    }
    if (stats && n) fill_batch_stats(traces, stats);
  });
}


extern "C" MvStatus mvamd_find_ncc_model_batch_submit(MvTuple ModelID, const uint8_t* const* d_images,
                                                      const int64_t* widths, const int64_t* heights,
                                                      const int64_t* pitches, size_t n, const MvAmdFindOptions* options,
                                                      int32_t with_stats, uint64_t* ticket) {
  return translate([&] {
    constexpr auto op = "find_ncc_model";
    required(d_images, op, "Images");
    required(widths, op, "widths");
    required(heights, op, "heights");
    required(pitches, op, "pitches");
    required(ticket, op, "ticket");
    if (n > NCC_MAX_BATCH)
      fail(mvamd::kLimitExceeded, mvamd::kErrInvalidValue, op, "Images", "too many frames for one submitted batch");
    required(options, op, "options");
    std::vector<mvamd::FrameInput> frames;
    for (size_t k = 0; k < n; ++k)
      frames.push_back(device_frame(op, d_images[k], widths[k], heights[k], pitches[k], options, int64_t(k)));
    PendingBatch b;
    b.model = model_of(ModelID, op);
    b.n = n;
    mvamd::check(hipGetDevice(&b.device), "hipGetDevice");
    b.traces.resize(with_stats ? n : 0);
    const mvamd::FindOptions q = from_ext(options);
    auto& pend = pending_batches();
    // the context keeps pointers into the traces' buffer, which moves with
    // the vector into the map
    auto node = std::make_unique<PendingBatch>(std::move(b));
    const uint64_t local = algorithm(op, [&] {
      return context().submit_batch(*node->model, frames.data(), n, q, with_stats && n ? node->traces.data() : nullptr);
    });
    node->local = local;
    const uint64_t key = (uint64_t(node->device) << 56) | local;
    pend[key] = std::move(*node);
    *ticket = key;
  });
}

extern "C" MvStatus mvamd_find_ncc_model_batch_wait(uint64_t ticket, double* results, size_t cap, size_t* counts,
                                                    MvAmdFindStats* stats) {
  return translate([&] {
    constexpr auto op = "find_ncc_model";
    required(counts, op, "counts");
    auto& pend = pending_batches();
    const auto it = pend.find(ticket);
    if (it == pend.end()) fail(mvamd::kInvalidState, mvamd::kErrInvalidValue, op, "ticket", "no such pending batch");
    int dev = 0;
    mvamd::check(hipGetDevice(&dev), "hipGetDevice");
    if (dev != it->second.device) {
      fail(mvamd::kInvalidState, mvamd::kErrInvalidValue, op, "ticket", "waited on another device than submitted");
    }
    PendingBatch b = std::move(it->second);
    pend.erase(it);
    const auto found = algorithm(op, [&] { return context().complete_batch(b.local); });
    for (size_t k = 0; k < b.n; ++k) {
      if (results) put_poses(found[k], results + 4 * cap * k, cap);
      counts[k] = found[k].size();
    }
    if (stats && !b.traces.empty()) fill_batch_stats(b.traces, stats);
  });
}

extern "C" MvStatus mvamd_find_ncc_model_trace(MvTuple ModelID, const uint8_t* image, int64_t width, int64_t height,
                                               int64_t pitch, const MvAmdFindOptions* options, double* results,
                                               size_t cap, size_t* count, double* coarse, size_t cap_coarse,
                                               size_t* n_coarse, double* prepared, double* refined, size_t cap_prepared,
                                               size_t* n_prepared, MvAmdFindStats* stats) {
  return translate([&] {
    constexpr auto op = "find_ncc_model";
    required(image, op, "Image");
    required(count, op, "count");
    const auto m = model_of(ModelID, op);
    mvamd::FindOptions q = from_ext(options);
    q.trace_stages = true;
    const int bytes = options->pixel_type == MVAMD_PIXEL_UINT2 ? 2 : 1;
    if (width <= 0 || height <= 0 || pitch < width * bytes || pitch % bytes)
      fail(mvamd::kInvalidArgument, mvamd::kErrInvalidValue, op, "Image", "invalid image geometry");
    mvamd::FindTrace trace;
    const auto found = algorithm(op, [&] {
      return context().find_host(*m, image, size_t(width), size_t(height), size_t(pitch), q, &trace, bytes);
    });
    if (results) put_poses(found, results, cap);
    *count = found.size();
    if (coarse) put_poses(trace.coarse, coarse, cap_coarse);
    if (n_coarse) *n_coarse = trace.coarse.size();
    if (prepared) put_poses(trace.prepared, prepared, cap_prepared);
    if (refined)
      for (size_t i = 0; i < trace.refined.size() && i < cap_prepared; ++i) {
        refined[5 * i] = trace.refined[i].row; refined[5 * i + 1] = trace.refined[i].col;
        refined[5 * i + 2] = trace.refined[i].angle; refined[5 * i + 3] = trace.refined[i].score;
        refined[5 * i + 4] = trace.verdict[i];
      }
    if (n_prepared) *n_prepared = trace.prepared.size();
    fill_stats(trace, stats);
  });
}

extern "C" MvStatus mvamd_get_ncc_model_tuning(MvTuple ModelID, char* text, size_t cap, size_t* length) {
  return translate([&] {
    constexpr auto op = "get_ncc_model_tuning";
    required(length, op, "length");
    if (cap) required(text, op, "text");
    const auto m = model_of(ModelID, op);
    std::string out;
    {
      std::lock_guard<std::mutex> lock(m->tune_mutex);
      char line[320];
      for (const auto& [key, t] : m->coarse_tuning) {
        std::snprintf(line, sizeof line,
                      "%s : rows=%d col_waves=%d row_waves=%d tiles=%d bound=%d us=%.17g default_us=%.17g\n",
                      key.c_str(), t.rows, t.col_waves, t.row_waves, t.tiles, t.bound, t.us, t.default_us);
        out += line;
      }
    }
    *length = out.size();
    if (cap) {
      const size_t n = std::min(cap - 1, out.size());
      std::memcpy(text, out.data(), n);
      text[n] = 0;
    }
  });
}

extern "C" MvStatus mvamd_set_ncc_model_tuning(MvTuple ModelID, const char* text) {
  return translate([&] {
    constexpr auto op = "set_ncc_model_tuning";
    required(text, op, "text");
    const auto m = model_of(ModelID, op);
    std::map<std::string, mvamd::CoarseTune> table;
    const std::string all(text);
    size_t pos = 0;
    int64_t index = 0;
    while (pos < all.size()) {
      size_t end = all.find('\n', pos);
      if (end == std::string::npos) end = all.size();
      const std::string line = all.substr(pos, end - pos);
      pos = end + 1;
      if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
      const size_t sep = line.find(" : ");
      mvamd::CoarseTune t;
      // "bound=" may be absent (bound 0)
      const char* f = sep != std::string::npos ? line.c_str() + sep + 3 : "";
      const bool ok = sep != std::string::npos && sep > 0 &&
                      (std::sscanf(f, "rows=%d col_waves=%d row_waves=%d tiles=%d bound=%d us=%lf default_us=%lf",
                                   &t.rows, &t.col_waves, &t.row_waves, &t.tiles, &t.bound, &t.us, &t.default_us) == 7 ||
                       (t.bound = 0, std::sscanf(f, "rows=%d col_waves=%d row_waves=%d tiles=%d us=%lf default_us=%lf",
                                                 &t.rows, &t.col_waves, &t.row_waves, &t.tiles, &t.us,
                                                 &t.default_us) == 6));
      const NcchipCoarseTile shape{t.rows, t.col_waves, t.row_waves, t.tiles, t.bound};
      // the shape alone (no level or box): valid when it fits a one-row,
      // one-chunk box at level 1
      if (!ok || !ncchip_coarse_mfma_fits(1, NCC_U16, 1, 1, 0, &shape))
        fail(mvamd::kInvalidArgument, mvamd::kErrInvalidValue, op, "text", "invalid tuning line", index);
      table[line.substr(0, sep)] = t;
      ++index;
    }
    std::lock_guard<std::mutex> lock(m->tune_mutex);
    m->coarse_tuning = std::move(table);
  });
}

extern "C" MvStatus mvamd_model_info(MvTuple ModelID, double* info) {
  return translate([&] {
    required(info, "model_info", "info");
    const auto m = model_of(ModelID, "model_info");
    info[0] = double(m->levels.size()); info[1] = double(m->n_angles); info[2] = m->params.step;
    info[3] = m->orow; info[4] = m->ocol; info[5] = double(m->sample_count);
  });
}

extern "C" MvStatus mvamd_model_raster(MvTuple ModelID, int32_t level, int32_t angle, double* header, int32_t* rows,
                                       int32_t* cols, double* weights, size_t cap) {
  return translate([&] {
    constexpr auto op = "model_raster";
    required(header, op, "header");
    const auto m = model_of(ModelID, op);
    uint32_t id = m->original;
    if (level >= 0) {
      if (size_t(level) >= m->levels.size() || angle < 0 || size_t(angle) >= m->n_angles)
        fail(mvamd::kOutOfRange, mvamd::kErrIndexOutOfRange, op, "level", "raster index out of range");
      id = m->raster(size_t(level), size_t(angle));
    }
    const mvamd::Raster& r = m->pool.rasters[id];
    const double h[11] = {double(r.count), r.angle, r.orow, r.ocol, double(r.rmin), double(r.rmax), double(r.cmin),
                          double(r.cmax), r.wsum, double(r.run_count), double(r.wrun_count)};
    std::copy(h, h + 11, header);
    for (size_t i = 0; i < r.count && i < cap; ++i) {
      const int32_t rc = m->pool.rc[r.first + i];
      if (rows) rows[i] = rc >> 16;
      if (cols) cols[i] = int32_t(int16_t(rc & 0xffff));
      if (weights) weights[i] = m->pool.w[r.first + i];
    }
  });
}
