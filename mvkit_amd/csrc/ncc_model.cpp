// Host construction of NCC models.  Compiled with -ffp-contract=off: every
// FP64 expression below is evaluated in the reference's order, so rotated
// template weights match mvkit bit for bit (model.cpp:39-242).
#include "ncc_model.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

namespace mvamd {

namespace {
constexpr double kTau = 6.283185307179586476925286766559;
constexpr size_t kMaxPixels = 1048576;   // model.cpp:15
constexpr size_t kMaxAngles = 4096;
constexpr size_t kMaxSamples = 16000000;
const char* const kCreateOp = "ncc_model_create";

[[noreturn]] void raise(Status s, const char* param, const char* msg) { throw Error(s, kCreateOp, param, msg); }

int32_t pack_rc(int32_t r, int32_t c) { return int32_t((uint32_t(r) << 16) | (uint32_t(c) & 0xffffu)); }
int32_t unpack_row(int32_t rc) { return rc >> 16; }
int32_t unpack_col(int32_t rc) { return int32_t(int16_t(uint16_t(uint32_t(rc) & 0xffffu))); }
}  // namespace

ErrorCode default_code(Status s) {
  switch (s) {
    case kInvalidArgument: return kErrInvalidValue;
    case kOutOfRange: return kErrIndexOutOfRange;
    case kTypeMismatch: return kErrInvalidType;
    case kNotFound: return kErrMissingItem;
    case kUnsupported: return kErrUnsupportedValue;
    case kInvalidState: return kErrInvalidObjectState;
    case kIoError: case kDecodeError: case kEncodeError: return kErrIoFailure;
    case kInternalError: case kOutOfMemory: return kErrInternalFailure;
    case kLimitExceeded: return kErrInvalidValue;
    default: return kErrUnspecified;
  }
}

size_t RasterPool::bytes() const {
  return rasters.size() * sizeof(Raster) + rc.size() * 4 + w.size() * 8 + wf.size() * 4 + runs.size() * 4 +
         wruns.size() * 4 + wrun_w.size() * 8;
}

// Samples are appended in raster order as (row, col, value); `finish` centres
// and scales them and derives bounds and runs (normalize, model.cpp:39-65).
namespace {
struct Staging {
  std::vector<int32_t> r, c;
  std::vector<double> v;
};

bool finish(const Staging& s, double angle, double orow, double ocol, RasterPool& pool, const double* pmean,
            const double* penergy, Raster& out) {
  const size_t n = s.v.size();
  if (n < 4) return false;
  std::vector<double> w(s.v);
  double mean = 0, energy = 0;
  if (pmean) {
    mean = *pmean;
    energy = *penergy;
    for (auto& x : w) x -= mean;
  } else {
    for (size_t i = 0; i < n; ++i) mean += (w[i] - mean) / double(i + 1);
    for (auto& x : w) { x -= mean; energy += x * x; }
  }
  if (!(energy > 0) || !std::isfinite(energy)) return false;
  const double scale = 1 / std::sqrt(energy);
  out = Raster{};
  out.angle = angle; out.orow = orow; out.ocol = ocol;
  out.first = uint32_t(pool.w.size());
  out.count = uint32_t(n);
  out.run_first = uint32_t(pool.runs.size() / 4);
  out.wrun_first = uint32_t(pool.wrun_w.size());
  out.rmin = out.rmax = s.r[0];
  out.cmin = out.cmax = s.c[0];
  int32_t run_row = 0, run_end = 0, wrun_row = 0, wrun_end = 0;
  double wrun_value = 0;
  bool have_run = false, have_wrun = false;
  for (size_t i = 0; i < n; ++i) {
    const double wi = w[i] * scale;
    const int32_t r = s.r[i], c = s.c[i];
    pool.rc.push_back(pack_rc(r, c));
    pool.w.push_back(wi);
    pool.wf.push_back(float(wi));
    out.wsum += wi;
    out.wsum_f += double(float(wi));
    out.rmin = std::min(out.rmin, r); out.rmax = std::max(out.rmax, r);
    out.cmin = std::min(out.cmin, c); out.cmax = std::max(out.cmax, c);
    if (have_run && run_row == r && run_end + 1 == c) {
      run_end = c;
      pool.runs[pool.runs.size() - 2] = c;
    } else {
      pool.runs.insert(pool.runs.end(), {r, c, c, 0});
      run_row = r; run_end = c; have_run = true;
    }
    if (have_wrun && wrun_row == r && wrun_end + 1 == c && wrun_value == wi) {
      wrun_end = c;
      pool.wruns[pool.wruns.size() - 2] = c;
    } else {
      pool.wruns.insert(pool.wruns.end(), {r, c, c, 0});
      pool.wrun_w.push_back(wi);
      wrun_row = r; wrun_end = c; wrun_value = wi; have_wrun = true;
    }
  }
  out.run_count = uint32_t(pool.runs.size() / 4) - out.run_first;
  out.wrun_count = uint32_t(pool.wrun_w.size()) - out.wrun_first;
  return true;
}
}  // namespace

RasterGrid raster_grid(const LevelGeom& lv, double angle) {
  RasterGrid g;
  g.unrotated = std::abs(std::remainder(angle, kTau)) < 1e-12;
  if (g.unrotated) {
    g.r_lo = 0; g.r_hi = int(lv.h) - 1; g.c_lo = 0; g.c_hi = int(lv.w) - 1;
    return g;
  }
  // inverse bilinear rotation about the centroid; raster coordinates are
  // offsets from the centroid (model.cpp:77-96)
  g.ca = std::cos(angle); g.sa = std::sin(angle);
  const double ca = g.ca, sa = g.sa;
  double lo_r = 0, hi_r = 0, lo_c = 0, hi_c = 0;
  const double corner_r[2] = {0., double(lv.h - 1)}, corner_c[2] = {0., double(lv.w - 1)};
  for (double cr : corner_r)
    for (double cc : corner_c) {
      const double dr = cr - lv.orow, dc = cc - lv.ocol;
      const double y = ca * dr - sa * dc, x = sa * dr + ca * dc;
      lo_r = std::min(lo_r, y); hi_r = std::max(hi_r, y);
      lo_c = std::min(lo_c, x); hi_c = std::max(hi_c, x);
    }
  const double rows = std::ceil(hi_r) - std::floor(lo_r) + 1, cols = std::ceil(hi_c) - std::floor(lo_c) + 1;
  if (rows > 4 * kMaxPixels / cols) raise(kLimitExceeded, "domain", "rotated bounding box exceeds budget");
  g.r_lo = int(std::floor(lo_r)); g.r_hi = int(std::ceil(hi_r));
  g.c_lo = int(std::floor(lo_c)); g.c_hi = int(std::ceil(hi_c));
  return g;
}

int64_t build_raster(const LevelGeom& lv, double angle, RasterPool& pool, const double* pmean, const double* penergy) {
  Staging s;
  double orow = 0, ocol = 0;
  const RasterGrid g = raster_grid(lv, angle);
  if (g.unrotated) {
    // unrotated raster on the level grid, origin at the level centroid
    orow = lv.orow; ocol = lv.ocol;
    for (size_t r = 0; r < lv.h; ++r)
      for (size_t c = 0; c < lv.w; ++c)
        if (lv.mask[r * lv.w + c]) { s.r.push_back(int32_t(r)); s.c.push_back(int32_t(c)); s.v.push_back(lv.gray[r * lv.w + c]); }
  } else {
    const double ca = g.ca, sa = g.sa;
    const double ymax = double(lv.h - 1), xmax = double(lv.w - 1);
    for (int r = g.r_lo; r <= g.r_hi; ++r)
      for (int c = g.c_lo; c <= g.c_hi; ++c) {
        const double ty = ca * r + sa * c + lv.orow, tx = -sa * r + ca * c + lv.ocol;
        if (ty < 0 || tx < 0 || ty > ymax || tx > xmax) continue;
        const size_t y0 = size_t(ty), x0 = size_t(tx);
        const size_t y1 = std::min(y0 + 1, lv.h - 1), x1 = std::min(x0 + 1, lv.w - 1);
        const size_t i00 = y0 * lv.w + x0, i01 = y0 * lv.w + x1, i10 = y1 * lv.w + x0, i11 = y1 * lv.w + x1;
        if (!lv.mask[i00] || !lv.mask[i01] || !lv.mask[i10] || !lv.mask[i11]) continue;
        const double fy = ty - y0, fx = tx - x0;
        const double top = (1 - fx) * lv.gray[i00] + fx * lv.gray[i01];
        const double bottom = (1 - fx) * lv.gray[i10] + fx * lv.gray[i11];
        s.r.push_back(r); s.c.push_back(c); s.v.push_back((1 - fy) * top + fy * bottom);
      }
  }
  // the shared first-level mean/energy only replaces the reduction of the
  // unrotated raster (model.cpp:154-158); rotated rasters normalise themselves
  const bool unrotated = g.unrotated;
  Raster out;
  if (!finish(s, angle, orow, ocol, pool, unrotated ? pmean : nullptr, unrotated ? penergy : nullptr, out)) return -1;
  pool.rasters.push_back(out);
  return int64_t(pool.rasters.size() - 1);
}

namespace {
// validate_zero (model.cpp:22-36): sample count and finite positive energy.
bool usable(const LevelGeom& lv, size_t& count) {
  double mean = 0;
  size_t n = 0;
  for (size_t i = 0; i < lv.gray.size(); ++i)
    if (lv.mask[i]) mean += (lv.gray[i] - mean) / double(++n);
  count = n;
  if (n < 4) return false;
  double energy = 0;
  for (size_t i = 0; i < lv.gray.size(); ++i)
    if (lv.mask[i]) { const double v = lv.gray[i] - mean; energy += v * v; }
  return energy > 0 && std::isfinite(energy);
}

// 2x2 mean level (model.cpp:206-220 with gray_pyramid.hpp:8-13).
LevelGeom halve(const LevelGeom& pv) {
  LevelGeom nx;
  nx.w = pv.w / 2; nx.h = pv.h / 2;
  nx.orow = (pv.orow - .5) * .5; nx.ocol = (pv.ocol - .5) * .5;
  nx.gray.resize(nx.w * nx.h);
  nx.mask.resize(nx.w * nx.h);
  for (size_t r = 0; r < nx.h; ++r)
    for (size_t c = 0; c < nx.w; ++c) {
      const size_t top = 2 * r * pv.w, bottom = (2 * r + 1) * pv.w;
      const size_t c0 = 2 * c, c1 = c0 + 1 < pv.w ? c0 + 1 : c0;
      nx.mask[r * nx.w + c] = pv.mask[top + c0] && pv.mask[top + c0 + 1] && pv.mask[bottom + c0] && pv.mask[bottom + c0 + 1];
      nx.gray[r * nx.w + c] = .25 * (pv.gray[top + c0] + pv.gray[top + c1] + pv.gray[bottom + c0] + pv.gray[bottom + c1]);
    }
  return nx;
}
}  // namespace

NccModel::~NccModel() = default;

std::unique_ptr<NccModel> create_model(const void* pixels, int pixel_bytes, size_t width, size_t height,
                                       size_t stride_bytes, const CreateParams& p, const Region* domain) {
  FirstLevel f = create_first_level(pixels, pixel_bytes, width, height, stride_bytes, p, domain);
  return rebuild_model(std::move(f.level), p, f.orow, f.ocol);
}

FirstLevel create_first_level(const void* pixels, int pixel_bytes, size_t width, size_t height, size_t stride_bytes,
                              const CreateParams& p, const Region* domain) {
  if (!width || !height) raise(kInvalidArgument, "image", "one nonempty single-channel image required");
  if (!std::isfinite(p.start) || !std::isfinite(p.extent) || !std::isfinite(p.step) || p.extent < 0 || p.extent > kTau ||
      !std::isfinite(p.start + p.extent) || p.step < 0 || p.levels < 0 || p.levels > 16)
    raise(kInvalidArgument, "options", "invalid levels or angle range");
  // the domain's runs and bounding box (the whole image without a reduced domain)
  const Region full = domain ? Region{} : Region::rectangle(0, 0, int64_t(height) - 1, int64_t(width) - 1);
  const Region& D = domain ? *domain : full;
  if (D.empty() || D.area < 4) raise(kInvalidArgument, "domain", "template domain requires at least four pixels");
  if (D.row_min < 0 || D.col_min < 0 || D.row_max >= int64_t(height) || D.col_max >= int64_t(width))
    raise(kInvalidArgument, "domain", "template domain must be inside the image");
  const size_t bw = size_t(D.col_max - D.col_min + 1), bh = size_t(D.row_max - D.row_min + 1);
  if (bh > kMaxPixels / bw) raise(kLimitExceeded, "domain", "template bounding box exceeds one million pixels");

  FirstLevel out;
  LevelGeom& first = out.level;
  first.w = bw; first.h = bh;
  first.gray.assign(bw * bh, 0.0);
  first.mask.assign(bw * bh, 0);
  // centroid of the domain, accumulated per row run in long double
  // (model.cpp:138-161); integer samples summed exactly in int64
  long double sum_row = 0, sum_col = 0;
  int64_t total = 0;
  const auto* base = static_cast<const uint8_t*>(pixels);
  for (const RegionRun& run : D.runs) {
    const int64_t count = int64_t(run.ce) - run.cb + 1;
    sum_row += static_cast<long double>(run.row) * count;
    sum_col += (static_cast<long double>(run.cb) + run.ce) * count / 2;
    const uint8_t* row = base + size_t(run.row) * stride_bytes;
    const size_t off = size_t(run.row - D.row_min) * bw;
    for (int64_t c = run.cb; c <= run.ce; ++c) {
      uint32_t v;
      if (pixel_bytes == 1) v = row[c];
      else { uint16_t t; std::memcpy(&t, row + 2 * c, 2); v = t; }
      const size_t i = off + size_t(c - D.col_min);
      first.gray[i] = double(v);
      first.mask[i] = 1;
      total += v;
    }
  }
  const long double area = static_cast<long double>(D.area);
  out.orow = double(sum_row / area);
  out.ocol = double(sum_col / area);
  first.orow = double((sum_row - static_cast<long double>(D.row_min) * area) / area);
  first.ocol = double((sum_col - static_cast<long double>(D.col_min) * area) / area);
  const double mean = double(total) / double(D.area);
  for (size_t i = 0; i < first.gray.size(); ++i)
    if (first.mask[i]) first.gray[i] -= mean;
  return out;
}

ModelPlan plan_model(LevelGeom first, const CreateParams& p, double origin_row, double origin_col) {
  // rasters store sample coordinates as 16-bit pairs
  if (first.w > 16383 || first.h > 16383)
    raise(kLimitExceeded, "domain", "template side exceeds 16383 pixels");
  ModelPlan plan;
  plan.model = std::make_unique<NccModel>();
  NccModel* m = plan.model.get();
  m->params = p;
  m->orow = origin_row;
  m->ocol = origin_col;
  // original patch (division by the norm), pyramid, angle grid
  m->sample_count = size_t(std::count(first.mask.begin(), first.mask.end(), uint8_t(1)));
  m->levels.push_back(std::move(first));
  // original patch: centred by the incremental mean, divided by the norm
  // (model.cpp:182-204); appended after the angle grid
  double& first_mean = plan.first_mean;
  double& first_energy = plan.first_energy;
  std::vector<int32_t>& orig_rc = plan.orig_rc;
  std::vector<double>& orig_w = plan.orig_w;
  {
    const LevelGeom& lv = m->levels.front();
    size_t cnt = 0;
    for (size_t i = 0; i < lv.gray.size(); ++i)
      if (lv.mask[i]) first_mean += (lv.gray[i] - first_mean) / double(++cnt);
    for (size_t r = 0; r < lv.h; ++r)
      for (size_t c = 0; c < lv.w; ++c)
        if (lv.mask[r * lv.w + c]) {
          const double v = lv.gray[r * lv.w + c] - first_mean;
          first_energy += v * v;
          orig_w.push_back(v);
          orig_rc.push_back(pack_rc(int32_t(r), int32_t(c)));
        }
    if (cnt < 4 || !(first_energy > 0) || !std::isfinite(first_energy))
      raise(kInvalidArgument, "image", "constant template cannot define an NCC model");
    const double norm = std::sqrt(first_energy);
    for (auto& v : orig_w) v /= norm;
  }
  // pyramid of the template (model.cpp:205-221)
  const size_t wanted = p.levels ? size_t(p.levels) : 16;
  while (m->levels.size() < wanted) {
    const LevelGeom& pv = m->levels.back();
    if (pv.w < 8 || pv.h < 8) break;
    LevelGeom nx = halve(pv);
    size_t count = 0;
    const bool ok = usable(nx, count);
    if (!ok || count < 8 || (m->levels.size() >= 2 && count < 64)) break;
    m->levels.push_back(std::move(nx));
  }
  if (p.levels && m->levels.size() != size_t(p.levels))
    raise(kInvalidArgument, "levels", "requested pyramid level is degenerate or too small");
  // angle grid (model.cpp:222-238)
  const double radius = std::hypot(double(m->levels.front().w), double(m->levels.front().h)) * .5;
  m->params.step = p.step > 0 ? p.step : std::max(.01, std::min(.15, 1 / std::max(1., radius)));
  const double intervals = std::ceil(p.extent / m->params.step);
  if (!std::isfinite(intervals) || intervals >= double(kMaxAngles))
    raise(kLimitExceeded, "angle_step", "angle grid exceeds limit");
  m->n_angles = size_t(intervals) + 1;
  return plan;
}

double raster_angle(const NccModel& m, size_t i) { return m.params.start + std::min(m.params.extent, i * m.params.step); }
void raise_degenerate_raster() { raise(kInvalidArgument, "domain", "rotation produces a degenerate template"); }
void raise_sample_budget() { raise(kLimitExceeded, "angle_extent", "model raster budget exceeded"); }
size_t max_model_samples() { return kMaxSamples; }

std::unique_ptr<NccModel> rebuild_model(LevelGeom first, const CreateParams& p, double origin_row, double origin_col) {
  ModelPlan plan = plan_model(std::move(first), p, origin_row, origin_col);
  NccModel* m = plan.model.get();
  size_t total_samples = 0;
  for (size_t li = 0; li < m->levels.size(); ++li) {
    LevelGeom& lv = m->levels[li];
    lv.first_raster = uint32_t(m->pool.rasters.size());
    for (size_t i = 0; i < m->n_angles; ++i) {
      const double angle = raster_angle(*m, i);
      const int64_t id = li == 0 ? build_raster(lv, angle, m->pool, &plan.first_mean, &plan.first_energy)
                                 : build_raster(lv, angle, m->pool);
      if (id < 0) raise_degenerate_raster();
      const size_t n = m->pool.rasters[size_t(id)].count;
      if (n > kMaxSamples - total_samples) raise_sample_budget();
      total_samples += n;
    }
  }
  return finish_model(std::move(plan));
}

std::unique_ptr<NccModel> finish_model(ModelPlan plan) {
  NccModel* m = plan.model.get();
  const std::vector<int32_t>& orig_rc = plan.orig_rc;
  const std::vector<double>& orig_w = plan.orig_w;
  Raster original;
  const LevelGeom& l0 = m->levels.front();
  original.orow = l0.orow;
  original.ocol = l0.ocol;
  // bounds of the masked samples (model.cpp:195-199)
  original.rmin = original.rmax = unpack_row(orig_rc.front());
  original.cmin = original.cmax = unpack_col(orig_rc.front());
  for (const int32_t rc : orig_rc) {
    original.rmin = std::min(original.rmin, unpack_row(rc)); original.rmax = std::max(original.rmax, unpack_row(rc));
    original.cmin = std::min(original.cmin, unpack_col(rc)); original.cmax = std::max(original.cmax, unpack_col(rc));
  }
  original.first = uint32_t(m->pool.w.size());
  original.count = uint32_t(orig_w.size());
  original.run_first = uint32_t(m->pool.runs.size() / 4);
  original.wrun_first = uint32_t(m->pool.wrun_w.size());
  for (size_t i = 0; i < orig_w.size(); ++i) {
    m->pool.rc.push_back(orig_rc[i]);
    m->pool.w.push_back(orig_w[i]);
    m->pool.wf.push_back(float(orig_w[i]));
  }
  // like the reference, the original patch carries no weight sum: the
  // continuous score centres the image samples instead (matcher.cpp:146-156)
  m->pool.rasters.push_back(original);
  m->original = uint32_t(m->pool.rasters.size() - 1);
  return std::move(plan.model);
}

}  // namespace mvamd
