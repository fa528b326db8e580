// Host driver of the NCC find path on one HIP device.  The discrete logic of
// prepare()/find() (level selection, angle clipping, coarse decimation,
// candidate bookkeeping, merging, greedy overlap selection) runs here in FP64
// with the reference's expressions; all per-pixel and per-sample work runs in
// the HIP kernels behind include/ncchip.h.
#include "ncc_device.hpp"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>

namespace mvamd {

namespace {
constexpr double kMissing = -std::numeric_limits<double>::infinity();
const char* const kFindOp = "ncc_model_find";
[[noreturn]] void raise(Status s, const char* param, const char* msg) { throw Error(s, kFindOp, param, msg); }

size_t align_up(size_t v, size_t a = 256) { return (v + a - 1) / a * a; }

double now_us() {
  return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Byte image of a raster pool: contiguous, 256-byte aligned sections.
struct PoolLayout {
  size_t patches = 0, rc = 0, w = 0, wf = 0, runs = 0, wruns = 0, wrun_w = 0, total = 0;
  explicit PoolLayout(const RasterPool& p) {
    size_t o = 0;
    patches = o; o = align_up(o + p.rasters.size() * sizeof(NccPatch));
    rc = o; o = align_up(o + p.rc.size() * 4);
    w = o; o = align_up(o + p.w.size() * 8);
    wf = o; o = align_up(o + p.wf.size() * 4);
    runs = o; o = align_up(o + p.runs.size() * 4);
    wruns = o; o = align_up(o + p.wruns.size() * 4);
    wrun_w = o; o = align_up(o + p.wrun_w.size() * 8);
    total = std::max<size_t>(o, 256);
  }
};

NccPatch describe(const Raster& r) {
  NccPatch P{};
  P.angle = r.angle; P.orow = r.orow; P.ocol = r.ocol; P.wsum = r.wsum; P.wsum_f = r.wsum_f;
  P.n = int32_t(r.count); P.rmin = r.rmin; P.rmax = r.rmax; P.cmin = r.cmin; P.cmax = r.cmax;
  P.sample_off = int32_t(r.first); P.run_off = int32_t(r.run_first); P.n_runs = int32_t(r.run_count);
  P.wrun_off = int32_t(r.wrun_first); P.n_wruns = int32_t(r.wrun_count);
  // weight-run dot products when 2*#runs < #samples (kernels_avx2.cpp:150, 175, 211)
  P.use_wruns = r.wrun_count > 0 && size_t(r.wrun_count) * 2 < size_t(r.count);
  P.dense_off = -1;
  return P;
}

void serialize(const RasterPool& p, const PoolLayout& L, uint8_t* dst) {
  auto* patches = reinterpret_cast<NccPatch*>(dst + L.patches);
  for (size_t i = 0; i < p.rasters.size(); ++i) patches[i] = describe(p.rasters[i]);
  std::memcpy(dst + L.rc, p.rc.data(), p.rc.size() * 4);
  std::memcpy(dst + L.w, p.w.data(), p.w.size() * 8);
  std::memcpy(dst + L.wf, p.wf.data(), p.wf.size() * 4);
  std::memcpy(dst + L.runs, p.runs.data(), p.runs.size() * 4);
  std::memcpy(dst + L.wruns, p.wruns.data(), p.wruns.size() * 4);
  std::memcpy(dst + L.wrun_w, p.wrun_w.data(), p.wrun_w.size() * 8);
}

NccPatchPool view_at(uint8_t* base, const PoolLayout& L) {
  NccPatchPool v{};
  v.patches = reinterpret_cast<const NccPatch*>(base + L.patches);
  v.rc = reinterpret_cast<const int32_t*>(base + L.rc);
  v.w = reinterpret_cast<const double*>(base + L.w);
  v.wf = reinterpret_cast<const float*>(base + L.wf);
  v.runs = reinterpret_cast<const int32_t*>(base + L.runs);
  v.wruns = reinterpret_cast<const int32_t*>(base + L.wruns);
  v.wrun_w = reinterpret_cast<const double*>(base + L.wrun_w);
  return v;
}

int32_t ref(int pool, uint32_t index) { return int32_t((uint32_t(pool) << 24) | index); }

bool pose_less(const Pose& a, const Pose& b) {  // merge_identical order (matcher.cpp:242-244)
  if (a.row != b.row) return a.row < b.row;
  if (a.col != b.col) return a.col < b.col;
  return a.angle < b.angle;
}
bool score_order(const Pose& a, const Pose& b) {  // matcher.cpp:824-827, 1072-1075
  if (a.score != b.score) return a.score > b.score;
  return pose_less(a, b);
}
void merge_identical(std::vector<Pose>& c) {
  std::sort(c.begin(), c.end(), pose_less);
  c.erase(std::unique(c.begin(), c.end(),
                      [](const Pose& a, const Pose& b) { return a.row == b.row && a.col == b.col && a.angle == b.angle; }),
          c.end());
}
}  // namespace

void check(hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  if (e == hipErrorOutOfMemory) throw Error(kOutOfMemory, kFindOp, "device", what);
  throw Error(kInternalError, kFindOp, "device", std::string(what) + ": " + hipGetErrorString(e));
}

DeviceBuffer::DeviceBuffer(size_t n) : bytes(n) {
  if (n) check(hipMalloc(&ptr, n), "hipMalloc");
}
DeviceBuffer::~DeviceBuffer() {
  if (ptr) (void)hipFree(ptr);
}
DeviceBuffer& DeviceBuffer::operator=(DeviceBuffer&& o) noexcept {
  if (this != &o) {
    if (ptr) (void)hipFree(ptr);
    ptr = o.ptr; bytes = o.bytes;
    o.ptr = nullptr; o.bytes = 0;
  }
  return *this;
}
PinnedBuffer::~PinnedBuffer() {
  if (ptr) (void)hipHostFree(ptr);
}
void PinnedBuffer::reserve(size_t n) {
  if (n <= bytes) return;
  if (ptr) (void)hipHostFree(ptr);
  ptr = nullptr; bytes = 0;
  n = std::max<size_t>(align_up(n, 1 << 16), 1 << 16);
  check(hipHostMalloc(&ptr, n, 0), "hipHostMalloc");
  bytes = n;
}

void DevicePool::upload(const RasterPool& pool, hipStream_t stream) {
  const PoolLayout L(pool);
  std::vector<uint8_t> host(L.total);
  serialize(pool, L, host.data());
  storage = DeviceBuffer(L.total);
  check(hipMemcpyAsync(storage.ptr, host.data(), L.total, hipMemcpyHostToDevice, stream), "model upload");
  check(hipStreamSynchronize(stream), "model upload");
  view = view_at(storage.as<uint8_t>(), L);
  n_rasters = pool.rasters.size();
}

std::shared_ptr<DeviceModel> device_model(const NccModel& m, hipStream_t stream) {
  int dev = 0;
  check(hipGetDevice(&dev), "hipGetDevice");
  std::lock_guard<std::mutex> lock(m.device_mutex);
  if (m.device.size() <= size_t(dev)) m.device.resize(size_t(dev) + 1);
  if (!m.device[size_t(dev)]) {
    auto d = std::make_shared<DeviceModel>();
    d->device = dev;
    d->pool.upload(m.pool, stream);
    m.device[size_t(dev)] = d;
  }
  return m.device[size_t(dev)];
}

double pose_overlap(const Pose& a, const Pose& b, const LevelGeom& lv) {
  struct P { double x, y; };
  // corners of the template box [-o-1/2, size-o-1/2] under each pose, relative to a
  auto corners = [&](const Pose& m, P* out) {
    const double ca = std::cos(m.angle), sa = std::sin(m.angle);
    const double left = -lv.ocol - .5, right = lv.w - lv.ocol - .5, top = -lv.orow - .5, bottom = lv.h - lv.orow - .5;
    const P in[4] = {{left, top}, {right, top}, {right, bottom}, {left, bottom}};
    for (int i = 0; i < 4; ++i)
      out[i] = {m.col - a.col + ca * in[i].x + sa * in[i].y, m.row - a.row - sa * in[i].x + ca * in[i].y};
  };
  P clip[4], poly[16], next[16];
  corners(b, clip);
  corners(a, poly);
  size_t count = 4;
  auto cross = [](P u, P v, P w) { return (v.x - u.x) * (w.y - u.y) - (v.y - u.y) * (w.x - u.x); };
  for (int e = 0; e < 4 && count; ++e) {
    const P p = clip[e], q = clip[(e + 1) % 4];
    size_t out = 0;
    for (size_t i = 0; i < count; ++i) {
      const P u = poly[i], v = poly[(i + 1) % count];
      const double du = cross(p, q, u), dv = cross(p, q, v);
      if (du >= 0) next[out++] = u;
      if ((du >= 0) != (dv >= 0)) {
        const double t = du / (du - dv);
        next[out++] = {u.x + t * (v.x - u.x), u.y + t * (v.y - u.y)};
      }
    }
    std::copy(next, next + 16, poly);
    count = out;
  }
  double area = 0;
  for (size_t i = 0; i < count; ++i)
    area += poly[i].x * poly[(i + 1) % count].y - poly[(i + 1) % count].x * poly[i].y;
  return std::abs(area) * .5 / (double(lv.w) * lv.h);
}

FindContext::FindContext() {
  check(hipGetDevice(&device_), "hipGetDevice");
  check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
}
FindContext::~FindContext() {
  if (stream_) {
    (void)hipStreamSynchronize(stream_);
    for (auto& t : timers_) { (void)hipEventDestroy(t.begin); (void)hipEventDestroy(t.end); }
    (void)hipStreamDestroy(stream_);
  }
}

void FindContext::time_begin(int stage) {
  if (!profiling_) return;
  if (timers_used_ == timers_.size()) {
    Timer t;
    check(hipEventCreate(&t.begin), "hipEventCreate");
    check(hipEventCreate(&t.end), "hipEventCreate");
    timers_.push_back(t);
  }
  timers_[timers_used_].stage = stage;
  check(hipEventRecord(timers_[timers_used_].begin, stream_), "hipEventRecord");
}
void FindContext::time_end() {
  if (!profiling_) return;
  check(hipEventRecord(timers_[timers_used_].end, stream_), "hipEventRecord");
  ++timers_used_;
}
void FindContext::time_collect(FindTrace* trace) {
  if (profiling_ && trace) {
    check(hipStreamSynchronize(stream_), "sync");
    for (size_t i = 0; i < timers_used_; ++i) {
      float ms = 0;
      check(hipEventElapsedTime(&ms, timers_[i].begin, timers_[i].end), "hipEventElapsedTime");
      trace->kernel_us[timers_[i].stage] += 1000.0 * ms;
    }
  }
  timers_used_ = 0;
}

std::vector<Pose> FindContext::find_host(const NccModel& model, const uint8_t* image, size_t width, size_t height,
                                         size_t pitch, const FindOptions& options, FindTrace* trace) {
  if (!width || !height) raise(kInvalidArgument, "image", "one nonempty single-channel image required");
  const size_t bytes = width * height;
  if (image_.bytes < bytes) image_ = DeviceBuffer(align_up(bytes, 1 << 20));
  check(hipMemcpy2DAsync(image_.ptr, width, image, pitch, width, height, hipMemcpyHostToDevice, stream_), "image upload");
  return find(model, image_.as<uint8_t>(), width, height, width, options, trace);
}

std::vector<Pose> FindContext::find(const NccModel& model, const uint8_t* image, size_t W, size_t H, size_t pitch,
                                    const FindOptions& req, FindTrace* trace) {
  const double t0 = now_us();
  profiling_ = req.profile && trace;
  timers_used_ = 0;
  if (trace) *trace = FindTrace{};
  // ---- options and levels (matcher.cpp:345-374)
  FindOptions o = req;
  if (!W || !H) raise(kInvalidArgument, "image", "one nonempty single-channel image required");
  if (model.levels.empty()) raise(kInvalidState, "model", "model is empty");
  if (o.highest < 0 || o.lowest < 1) raise(kInvalidArgument, "NumLevels", "invalid pyramid levels");
  const size_t level_count = o.highest ? std::min(model.levels.size(), size_t(o.highest)) : model.levels.size();
  o.lowest = int(std::min(size_t(o.lowest), level_count));
  const size_t finest = size_t(o.lowest - 1);
  RasterPool qpool;  // per-find rasters: endpoint angles and the coarse original patch
  int64_t coarse_original = -1;
  if (finest) {
    coarse_original = build_raster(model.levels[finest], 0, qpool);
    if (coarse_original < 0) raise(kInvalidArgument, "model", "degenerate finest level");
  }
  if (!std::isfinite(o.start) || !std::isfinite(o.extent) || o.extent < 0 || !std::isfinite(o.min_score) ||
      o.min_score < 0 || o.min_score > 1 || !std::isfinite(o.max_overlap) || o.max_overlap < 0 || o.max_overlap > 1 ||
      !o.max_candidates)
    raise(kInvalidArgument, "options", "invalid query range, score, overlap or candidate budget");
  const double angle_end = o.start + o.extent;
  if (!std::isfinite(angle_end)) raise(kInvalidArgument, "angle_extent", "query angle range must be finite");
  const double cs = std::max(o.start, model.params.start);
  const double ce = std::min(angle_end, model.params.start + model.params.extent);
  if (cs > ce + 1e-12) return {};
  o.start = cs;
  o.extent = std::max(0., ce - cs);
  if (W > size_t(std::numeric_limits<int>::max()) || H > size_t(std::numeric_limits<int>::max()) || H > 32000000 / W)
    raise(kLimitExceeded, "image", "search image exceeds pixel budget");
  auto dmodel = device_model(model, stream_);

  // ---- crop (matcher.cpp:383-393; full-image domain)
  const LevelGeom& base = model.levels.front();
  const double radius = std::hypot(std::max(base.orow, double(base.h - 1) - base.orow),
                                   std::max(base.ocol, double(base.w - 1) - base.ocol));
  const int64_t lattice = int64_t(size_t(1) << (level_count - 1));
  const int64_t halo = int64_t(std::ceil(radius)) + 4 * lattice;
  const int64_t crop_r = std::max<int64_t>(0, 0 - halo) / lattice * lattice;
  const int64_t crop_c = std::max<int64_t>(0, 0 - halo) / lattice * lattice;
  const int64_t crop_b = std::min<int64_t>(int64_t(H), ((int64_t(H) - 1 + halo) / lattice + 1) * lattice);
  const int64_t crop_e = std::min<int64_t>(int64_t(W), ((int64_t(W) - 1 + halo) / lattice + 1) * lattice);
  const uint8_t* crop = image + crop_r * int64_t(pitch) + crop_c;

  // ---- level geometry of the search pyramid (matcher.cpp:407, 437)
  struct Lv { int w, h, kind, pitch; size_t off, bytes; };
  std::vector<Lv> lv(level_count);
  lv[0] = {int(crop_e - crop_c), int(crop_b - crop_r), NCC_U8, int(pitch), 0, 0};
  for (size_t l = 1; l < level_count; ++l) {
    lv[l].w = std::max(1, lv[l - 1].w / 2);
    lv[l].h = std::max(1, lv[l - 1].h / 2);
    lv[l].kind = l <= 4 ? NCC_U16 : NCC_U32;
    lv[l].pitch = lv[l].w;
    lv[l].bytes = size_t(lv[l].w) * lv[l].h * (l <= 4 ? 2 : 4);
  }

  // ---- query patches per level (matcher.cpp:463-482)
  std::vector<std::vector<int32_t>> refs(level_count);
  std::vector<std::vector<double>> ref_angles(level_count);
  for (size_t li = 0; li < level_count; ++li) {
    const LevelGeom& g = model.levels[li];
    for (size_t a = 0; a < model.n_angles; ++a) {
      const Raster& r = model.pool.rasters[model.raster(li, a)];
      if (r.angle >= o.start - 1e-12 && r.angle <= angle_end + 1e-12) {
        refs[li].push_back(ref(0, model.raster(li, a)));
        ref_angles[li].push_back(r.angle);
      }
    }
    for (const double a : {o.start, angle_end}) {
      bool found = false;
      for (double x : ref_angles[li]) if (std::abs(x - a) < 1e-12) found = true;
      if (found) continue;
      const int64_t id = build_raster(g, a, qpool);
      if (id >= 0) { refs[li].push_back(ref(1, uint32_t(id))); ref_angles[li].push_back(a); }
    }
  }
  const size_t top = level_count - 1;
  const double coarse_radius = .5 * std::hypot(double(model.levels[top].w), double(model.levels[top].h));
  const size_t stride = std::max<size_t>(
      1, size_t(std::min(double(refs[top].size()), std::ceil(.5 / std::max(1e-20, coarse_radius * model.params.step)))));
  {
    std::vector<size_t> order(refs[top].size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return ref_angles[top][a] < ref_angles[top][b]; });
    std::vector<int32_t> kept;
    std::vector<double> kept_angles;
    for (size_t i = 0; i < order.size(); ++i)
      if (i % stride == 0 || i + 1 == order.size()) { kept.push_back(refs[top][order[i]]); kept_angles.push_back(ref_angles[top][order[i]]); }
    refs[top] = kept;
    ref_angles[top] = kept_angles;
  }
  auto raster_of = [&](int32_t r) -> const Raster& {
    return (r >> 24) ? qpool.rasters[size_t(r & 0xffffff)] : model.pool.rasters[size_t(r & 0xffffff)];
  };

  // ---- coarse anchor ranges (matcher.cpp:549-553; full-image domain)
  const double scale = std::ldexp(1., int(top)), phase = (scale - 1) * .5;
  const size_t A = refs[top].size();
  std::vector<int32_t> ranges(4 * A);
  int max_bh = 0, max_bw = 0;
  for (size_t i = 0; i < A; ++i) {
    const Raster& p = raster_of(refs[top][i]);
    const double margin = scale == 1 ? .5 : 2 * scale;
    const int64_t rmin_box = 0, rmax_box = int64_t(H) - 1, cmin_box = 0, cmax_box = int64_t(W) - 1;
    ranges[4 * i + 0] = std::max({0, -p.rmin, int(std::ceil((rmin_box - margin - crop_r - phase) / scale - p.orow))});
    ranges[4 * i + 1] = std::min(lv[top].h - p.rmax, int(std::floor((rmax_box + margin - crop_r - phase) / scale - p.orow)) + 1);
    ranges[4 * i + 2] = std::max({0, -p.cmin, int(std::ceil((cmin_box - margin - crop_c - phase) / scale - p.ocol))});
    ranges[4 * i + 3] = std::min(lv[top].w - p.cmax, int(std::floor((cmax_box + margin - crop_c - phase) / scale - p.ocol)) + 1);
    max_bh = std::max(max_bh, p.rmax - p.rmin);
    max_bw = std::max(max_bw, p.cmax - p.cmin);
    if (trace)
      trace->coarse_flops += 2.0 * double(p.count) * std::max(0, ranges[4 * i + 1] - ranges[4 * i]) *
                             std::max(0, ranges[4 * i + 3] - ranges[4 * i + 2]);
  }

  // ---- workspace layout
  const size_t cap = o.max_candidates + 1;
  const int Wt = lv[top].w, Ht = lv[top].h;
  const PoolLayout QL(qpool);
  size_t off = 0;
  for (size_t l = 1; l < level_count; ++l) { lv[l].off = off; off = align_up(off + lv[l].bytes); }
  const size_t off_p1 = off; off = align_up(off + size_t(Wt + 1) * Ht * 4);
  const size_t off_p2 = off; off = align_up(off + size_t(Wt + 1) * Ht * 8);
  const size_t off_nz = off; off = align_up(off + size_t(Wt + 1) * Ht * 4);
  const size_t off_planes = off; off = align_up(off + A * size_t(Ht) * Wt * 8);
  const size_t off_peaks = off; off = align_up(off + cap * sizeof(NccPeak));
  const size_t off_count = off; off = align_up(off + 16);
  const size_t off_cand = off; off = align_up(off + cap * sizeof(NccMatch));
  const size_t off_out = off; off = align_up(off + cap * sizeof(NccMatch));
  const size_t off_refined = off; off = align_up(off + cap * sizeof(NccMatch));
  const size_t off_verdict = off; off = align_up(off + cap * 4);
  // small per-find uploads: refs of every level, coarse ranges, query pool
  std::vector<size_t> off_refs(level_count);
  const size_t off_small = off;
  size_t small = 0;
  for (size_t l = 0; l < level_count; ++l) { off_refs[l] = off_small + small; small = align_up(small + refs[l].size() * 4 + 4); }
  const size_t off_ranges = off_small + small; small = align_up(small + ranges.size() * 4);
  const size_t off_qpool = off_small + small; small = align_up(small + QL.total);
  off = off_small + small;
  if (work_.bytes < off) {
    check(hipStreamSynchronize(stream_), "sync");
    work_ = DeviceBuffer(align_up(off + off / 4, 1 << 20));
  }
  uint8_t* dev = work_.as<uint8_t>();
  stage_.reserve(std::max<size_t>(small, cap * sizeof(NccMatch)));
  readback_.reserve(cap * (2 * sizeof(NccMatch) + 4) + cap * sizeof(NccPeak) + 64);
  {
    uint8_t* s = static_cast<uint8_t*>(stage_.ptr);
    for (size_t l = 0; l < level_count; ++l)
      std::memcpy(s + (off_refs[l] - off_small), refs[l].data(), refs[l].size() * 4);
    std::memcpy(s + (off_ranges - off_small), ranges.data(), ranges.size() * 4);
    serialize(qpool, QL, s + (off_qpool - off_small));
    check(hipMemcpyAsync(dev + off_small, s, small, hipMemcpyHostToDevice, stream_), "upload");
  }
  NccPatchPool pools[2] = {dmodel->pool.view, view_at(dev + off_qpool, QL)};
  auto level_desc = [&](size_t l) {
    NccLevel d{};
    d.data = l ? static_cast<const void*>(dev + lv[l].off) : static_cast<const void*>(crop);
    d.kind = lv[l].kind; d.pitch = lv[l].pitch; d.width = lv[l].w; d.height = lv[l].h; d.shift = int32_t(2 * l);
    return d;
  };

  // ---- search pyramid, top-level statistics, coarse search, peaks
  time_begin(kStagePyramid);
  for (size_t l = 1; l < level_count; ++l) {
    const NccLevel src = level_desc(l - 1);
    if (ncchip_pyramid_level(src.data, src.kind, src.pitch, src.width, src.height, dev + lv[l].off, lv[l].kind,
                             lv[l].pitch, lv[l].w, lv[l].h, stream_))
      raise(kInternalError, "device", "pyramid launch failed");
  }
  time_end();
  const NccLevel topd = level_desc(top);
  auto* p1 = reinterpret_cast<uint32_t*>(dev + off_p1);
  auto* p2 = reinterpret_cast<uint64_t*>(dev + off_p2);
  time_begin(kStagePrefix);
  if (ncchip_row_prefix(topd.data, topd.kind, topd.pitch, Wt, Ht, crop, int(2 * top), p1, p2,
                        reinterpret_cast<uint32_t*>(dev + off_nz), nullptr, stream_))
    raise(kInternalError, "device", "prefix launch failed");
  time_end();
  auto* planes = reinterpret_cast<double*>(dev + off_planes);
  const double C = std::ldexp(128.0, int(2 * top));
  const int exact_top = top <= 2;
  int rc = 0;
  time_begin(kStageCoarse);
  if (o.form == NCCHIP_FORM_BF16X3 || o.form == NCCHIP_FORM_FP8_SCREEN)
    rc = ncchip_coarse_mfma(&topd, crop, C, p1, p2, &pools[0], reinterpret_cast<const int32_t*>(dev + off_refs[top]),
                            reinterpret_cast<const int32_t*>(dev + off_ranges), int(A), max_bh, max_bw, exact_top,
                            model.params.absolute, o.form == NCCHIP_FORM_FP8_SCREEN ? o.screen : kMissing, planes, stream_);
  else
    rc = ncchip_coarse(&topd, crop, C, p1, p2, &pools[0], reinterpret_cast<const int32_t*>(dev + off_refs[top]),
                       reinterpret_cast<const int32_t*>(dev + off_ranges), int(A), max_bh, max_bw, o.form, exact_top,
                       model.params.absolute, planes, stream_);
  if (rc) raise(kInternalError, "device", "coarse launch failed");
  time_end();
  auto* count = reinterpret_cast<int32_t*>(dev + off_count);
  check(hipMemsetAsync(count, 0, 4, stream_), "memset");
  time_begin(kStagePeaks);
  if (ncchip_peaks(planes, int(A), Ht, Wt, o.min_score * .4, reinterpret_cast<NccPeak*>(dev + off_peaks), int(cap),
                   count, stream_))
    raise(kInternalError, "device", "peak launch failed");
  time_end();
  auto* rb = static_cast<uint8_t*>(readback_.ptr);
  check(hipMemcpyAsync(rb, count, 4, hipMemcpyDeviceToHost, stream_), "readback");
  check(hipStreamSynchronize(stream_), "coarse");
  int32_t n_peaks = *reinterpret_cast<int32_t*>(rb);
  if (size_t(n_peaks) > o.max_candidates) raise(kLimitExceeded, "max_candidates", "coarse candidate budget exceeded");
  std::vector<Pose> cand;
  if (n_peaks > 0) {
    check(hipMemcpyAsync(rb, dev + off_peaks, size_t(n_peaks) * sizeof(NccPeak), hipMemcpyDeviceToHost, stream_),
          "readback");
    check(hipStreamSynchronize(stream_), "coarse");
    std::vector<NccPeak> pk(static_cast<size_t>(n_peaks));
    std::memcpy(pk.data(), rb, size_t(n_peaks) * sizeof(NccPeak));
    // the reference emits peaks in (angle, row, col) order
    std::sort(pk.begin(), pk.end(), [](const NccPeak& a, const NccPeak& b) {
      if (a.angle != b.angle) return a.angle < b.angle;
      if (a.row != b.row) return a.row < b.row;
      return a.col < b.col;
    });
    cand.reserve(pk.size());
    for (const auto& k : pk) {
      const Raster& p = raster_of(refs[top][size_t(k.angle)]);
      cand.push_back({double(crop_r) + scale * (k.row + p.orow) + phase, double(crop_c) + scale * (k.col + p.ocol) + phase,
                      p.angle, k.score});
    }
  }
  if (trace) {
    trace->coarse = cand; trace->top = int(top); trace->finest = int(finest); trace->angle_stride = int(stride);
    trace->tracked.clear();
  }
  const double t1 = now_us();
  if (cand.empty()) {
    time_collect(trace);
    if (trace) { trace->t_coarse_us = t1 - t0; trace->t_total_us = t1 - t0; }
    return {};
  }

  // ---- level-0 storage of the candidate region (matcher.cpp:612-629)
  int32_t storage0[4] = {0, 0, lv[0].h, lv[0].w};
  if (!finest) {
    double rmin = lv[0].h, cmin = lv[0].w, rmax = 0, cmax = 0;
    for (const auto& p : cand) {
      rmin = std::min(rmin, p.row - crop_r); rmax = std::max(rmax, p.row - crop_r);
      cmin = std::min(cmin, p.col - crop_c); cmax = std::max(cmax, p.col - crop_c);
    }
    const double margin = radius + 4 * scale + 8;
    storage0[0] = int32_t(size_t(std::max(0., std::floor(rmin - margin))));
    storage0[1] = int32_t(size_t(std::max(0., std::floor(cmin - margin))));
    storage0[2] = int32_t(size_t(std::min(double(lv[0].h), std::ceil(rmax + margin) + 1)) - size_t(storage0[0]));
    storage0[3] = int32_t(size_t(std::min(double(lv[0].w), std::ceil(cmax + margin) + 1)) - size_t(storage0[1]));
  }
  if (trace) std::copy(storage0, storage0 + 4, trace->storage);
  auto* dcand = reinterpret_cast<NccMatch*>(dev + off_cand);
  auto* dout = reinterpret_cast<NccMatch*>(dev + off_out);
  auto upload_cand = [&](const std::vector<Pose>& c) {
    auto* s = static_cast<NccMatch*>(stage_.ptr);
    for (size_t i = 0; i < c.size(); ++i) s[i] = {c[i].row, c[i].col, c[i].angle, c[i].score};
    check(hipMemcpyAsync(dcand, s, c.size() * sizeof(NccMatch), hipMemcpyHostToDevice, stream_), "upload");
  };

  // ---- coarse-to-fine tracking (matcher.cpp:630-761)
  for (size_t l = top; l > finest; --l) {
    if (l == 1 && o.subpixel) {
      for (auto& p : cand) {
        const double r = std::floor(p.row + .5), c = std::floor(p.col + .5);
        if (!(r >= 0 && c >= 0 && r < double(H) && c < double(W))) p.score = kMissing;
      }
      merge_identical(cand);
      break;
    }
    const NccLevel tl = level_desc(l - 1);
    const double step = std::ldexp(1., int(l - 1)), offset = (step - 1) * .5;
    const double tol = model.params.step * (l == top ? double(stride) : 1.) * 1.01;
    int32_t storage[4] = {0, 0, tl.height, tl.width};
    if (l - 1 == 0) std::copy(storage0, storage0 + 4, storage);
    upload_cand(cand);
    time_begin(kStageTrack);
    if (ncchip_track(&tl, crop, pools, reinterpret_cast<const int32_t*>(dev + off_refs[l - 1]), int(refs[l - 1].size()),
                     dcand, dout, int(cand.size()), tol, double(crop_r), double(crop_c), step, offset, storage,
                     l - 1 <= 2, model.params.absolute, stream_))
      raise(kInternalError, "device", "tracking launch failed");
    time_end();
    check(hipMemcpyAsync(rb, dout, cand.size() * sizeof(NccMatch), hipMemcpyDeviceToHost, stream_), "readback");
    check(hipStreamSynchronize(stream_), "tracking");
    const auto* res = reinterpret_cast<const NccMatch*>(rb);
    for (size_t i = 0; i < cand.size(); ++i) cand[i] = {res[i].row, res[i].col, res[i].angle, res[i].score};
    merge_identical(cand);
    cand.erase(std::remove_if(cand.begin(), cand.end(),
                              [&](const Pose& p) { return !std::isfinite(p.score) || p.score < o.min_score * .4; }),
               cand.end());
    if (trace) trace->tracked.push_back(cand);
  }
  const double t2 = now_us();

  // ---- full score at the finest level, then candidate order (matcher.cpp:813-827)
  NcchipRefineParams rp{};
  rp.level = level_desc(finest);
  rp.dc_src = crop;
  if (finest) { rp.storage[0] = 0; rp.storage[1] = 0; rp.storage[2] = lv[finest].h; rp.storage[3] = lv[finest].w; }
  else std::copy(storage0, storage0 + 4, rp.storage);
  rp.crop_r = double(crop_r); rp.crop_c = double(crop_c);
  rp.scale = std::ldexp(1., int(finest)); rp.phase = (rp.scale - 1) * .5;
  rp.finest = int32_t(finest); rp.top = int32_t(top);
  rp.orig_ref = finest ? ref(1, uint32_t(coarse_original)) : ref(0, model.original);
  rp.level_orow = model.levels[finest].orow; rp.level_ocol = model.levels[finest].ocol;
  rp.refs0 = reinterpret_cast<const int32_t*>(dev + off_refs[0]);
  rp.n_refs0 = int32_t(refs[0].size());
  rp.step = model.params.step; rp.angle_start = o.start; rp.angle_end = o.start + o.extent;
  rp.min_score = o.min_score; rp.subpixel = o.subpixel; rp.rotate = o.extent > 0; rp.absolute = model.params.absolute;
  rp.image_h = int32_t(H); rp.image_w = int32_t(W);
  std::vector<Pose> full(cand.size());
  if (!cand.empty()) {
    upload_cand(cand);
    time_begin(kStageFull);
    if (ncchip_full_score(&rp, pools, dcand, int(cand.size()), dout, stream_))
      raise(kInternalError, "device", "full score launch failed");
    time_end();
    check(hipMemcpyAsync(rb, dout, cand.size() * sizeof(NccMatch), hipMemcpyDeviceToHost, stream_), "readback");
    check(hipStreamSynchronize(stream_), "full score");
    const auto* f = reinterpret_cast<const NccMatch*>(rb);
    for (size_t i = 0; i < cand.size(); ++i) full[i] = {f[i].row, f[i].col, f[i].angle, f[i].score};
  }
  std::sort(full.begin(), full.end(), score_order);

  // ---- greedy selection in score order (matcher.cpp:1063-1076).  Each
  // candidate's refinement is independent of the others, so candidates are
  // refined on the device in batches, in the order the greedy walk needs them.
  const size_t n = full.size();
  std::vector<Pose> refined(n);
  std::vector<int> verdict(n, -1);  // -1: not refined
  std::vector<Pose> result;
  auto suppressed = [&](const Pose& c) {
    for (const auto& p : result) if (pose_overlap(c, p, base) > o.max_overlap + 1e-12) return true;
    return false;
  };
  auto* drefined = reinterpret_cast<NccMatch*>(dev + off_refined);
  auto* dverdict = reinterpret_cast<int32_t*>(dev + off_verdict);
  size_t pos = 0, batch_size = o.num_matches ? std::max<size_t>(2 * o.num_matches, 4) : n;
  bool done = false;
  while (!done && pos < n) {
    // walk as far as the refined results allow
    while (pos < n) {
      const Pose& c = full[pos];
      if (!std::isfinite(c.score) || suppressed(c)) { ++pos; continue; }
      if (verdict[pos] < 0) break;
      if (verdict[pos] && !suppressed(refined[pos])) {
        result.push_back(refined[pos]);
        if (o.num_matches && result.size() == o.num_matches) { done = true; break; }
      }
      ++pos;
    }
    if (done || pos >= n) break;
    // refine the next candidates the walk may need
    std::vector<size_t> batch;
    std::vector<Pose> poses;
    for (size_t i = pos; i < n && batch.size() < batch_size; ++i)
      if (std::isfinite(full[i].score) && verdict[i] < 0 && !suppressed(full[i])) { batch.push_back(i); poses.push_back(full[i]); }
    upload_cand(poses);
    time_begin(kStageRefine);
    if (ncchip_refine(&rp, pools, dcand, int(poses.size()), drefined, dverdict, stream_))
      raise(kInternalError, "device", "refine launch failed");
    time_end();
    const size_t m = poses.size();
    check(hipMemcpyAsync(rb, drefined, m * sizeof(NccMatch), hipMemcpyDeviceToHost, stream_), "readback");
    check(hipMemcpyAsync(rb + m * sizeof(NccMatch), dverdict, m * 4, hipMemcpyDeviceToHost, stream_), "readback");
    check(hipStreamSynchronize(stream_), "refine");
    const auto* g = reinterpret_cast<const NccMatch*>(rb);
    const auto* v = reinterpret_cast<const int32_t*>(rb + m * sizeof(NccMatch));
    for (size_t k = 0; k < m; ++k) {
      refined[batch[k]] = {g[k].row, g[k].col, g[k].angle, g[k].score};
      verdict[batch[k]] = v[k];
    }
    batch_size *= 2;
  }
  const double t3 = now_us();
  if (trace) {
    trace->prepared = full;
    trace->refined = refined;
    trace->verdict = verdict;
  }
  std::sort(result.begin(), result.end(), score_order);
  time_collect(trace);
  if (trace) {
    trace->t_coarse_us = t1 - t0; trace->t_track_us = t2 - t1; trace->t_refine_us = t3 - t2;
    trace->t_total_us = now_us() - t0;
  }
  return result;
}

}  // namespace mvamd
