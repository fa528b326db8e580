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
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <utility>

namespace mvamd {

namespace {
constexpr double kMissing = -std::numeric_limits<double>::infinity();
// FP8-screened form: the bound threshold a wave of coarse windows must reach
// to be scored, when the options give none (instances score well above it at
// the coarse level; peaks below it may be dropped)
constexpr double kDefaultScreen = 0.55;
const char* const kFindOp = "ncc_model_find";
[[noreturn]] void raise(Status s, const char* param, const char* msg) { throw Error(s, kFindOp, param, msg); }

size_t align_up(size_t v, size_t a = 256) { return (v + a - 1) / a * a; }

double now_us() {
  return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Dense split-BF16 rasters are built for rasters whose padded bounding box
// holds at most this many elements (every tracking level of templates up to
// several hundred pixels; the level-0 rasters of large templates are left out).
constexpr size_t kDenseMaxElems = size_t(1) << 16;

struct DenseGeom { int32_t h = 0, w = 0; size_t elems = 0; };
DenseGeom dense_geom(const Raster& r) {
  DenseGeom g;
  g.h = r.rmax - r.rmin + 1;
  g.w = ((r.cmax - r.cmin + 1) + 16 + 7) & ~7;  // rows of whole 16-byte groups
  const size_t e = size_t(g.h) * size_t(g.w);
  if (r.count == 0 || e > kDenseMaxElems) return DenseGeom{};
  g.elems = 2 * e;  // two copies
  return g;
}

// Byte image of a raster pool: contiguous, 256-byte aligned sections.
struct PoolLayout {
  size_t patches = 0, rc = 0, w = 0, wf = 0, runs = 0, wruns = 0, wrun_w = 0, dense_hi = 0, dense_lo = 0, total = 0;
  std::vector<int64_t> dense_off;  // per raster, elements (-1: none)
  size_t dense_elems = 0;
  explicit PoolLayout(const RasterPool& p) {
    dense_off.assign(p.rasters.size(), -1);
    for (size_t i = 0; i < p.rasters.size(); ++i) {
      const DenseGeom g = dense_geom(p.rasters[i]);
      if (!g.elems || dense_elems + g.elems > size_t(INT32_MAX)) continue;
      dense_off[i] = int64_t(dense_elems);
      dense_elems += g.elems;
    }
    size_t o = 0;
    patches = o; o = align_up(o + p.rasters.size() * sizeof(NccPatch));
    rc = o; o = align_up(o + p.rc.size() * 4);
    w = o; o = align_up(o + p.w.size() * 8);
    wf = o; o = align_up(o + p.wf.size() * 4);
    runs = o; o = align_up(o + p.runs.size() * 4);
    wruns = o; o = align_up(o + p.wruns.size() * 4);
    wrun_w = o; o = align_up(o + p.wrun_w.size() * 8);
    dense_hi = o; o = align_up(o + dense_elems * 2);
    dense_lo = o; o = align_up(o + dense_elems * 2);
    total = std::max<size_t>(o, 256);
  }
};

uint16_t bf16_bits(double v);
double bf16_value(uint16_t h);

// Fills raster r's dense copies (hi and lo parts) and its split-BF16 error
// coefficients.
void dense_raster(const RasterPool& p, const Raster& r, uint16_t* hi, uint16_t* lo, NccPatch& P) {
  const DenseGeom g = dense_geom(r);
  const size_t plane = size_t(g.h) * size_t(g.w);
  std::fill(hi, hi + plane, uint16_t(0));
  std::fill(lo, lo + plane, uint16_t(0));
  double wh1 = 0, wl1 = 0, res1 = 0;
  for (uint32_t i = 0; i < r.count; ++i) {
    const int32_t rc = p.rc[r.first + i];
    const int32_t row = rc >> 16, col = int32_t(int16_t(uint16_t(uint32_t(rc) & 0xffffu)));
    const double w = p.w[r.first + i];
    const uint16_t h = bf16_bits(w);
    const uint16_t l = bf16_bits(w - bf16_value(h));
    const size_t e = size_t(row - r.rmin) * size_t(g.w) + size_t(col - r.cmin + 8);
    hi[e] = h;
    lo[e] = l;
    wh1 += std::abs(bf16_value(h));
    wl1 += std::abs(bf16_value(l));
    res1 += std::abs(w - bf16_value(h) - bf16_value(l));
  }
  for (size_t e = 0; e < plane; ++e) {  // copy 1 = copy 0 shifted left by one element
    hi[plane + e] = e + 1 < plane ? hi[e + 1] : uint16_t(0);
    lo[plane + e] = e + 1 < plane ? lo[e + 1] : uint16_t(0);
  }
  const double k = 3.0 * r.count + 8, u = std::ldexp(1.0, -23);
  const double gam = k * u / (1 - k * u);
  // the sums above round; a relative slack covers them
  P.err_lo = (wl1 + gam * wh1) * (1 + 1e-9);
  P.err_x = (res1 + gam * (wh1 + wl1)) * (1 + 1e-9) + 1e-300;
  P.dense_h = g.h;
  P.dense_w = g.w;
}

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
  auto* dhi = reinterpret_cast<uint16_t*>(dst + L.dense_hi);
  auto* dlo = reinterpret_cast<uint16_t*>(dst + L.dense_lo);
  for (size_t i = 0; i < p.rasters.size(); ++i) {
    patches[i] = describe(p.rasters[i]);
    if (L.dense_off[i] >= 0) {
      patches[i].dense_off = int32_t(L.dense_off[i]);
      dense_raster(p, p.rasters[i], dhi + L.dense_off[i], dlo + L.dense_off[i], patches[i]);
    }
  }
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
  v.dense_hi = reinterpret_cast<const uint16_t*>(base + L.dense_hi);
  v.dense_lo = reinterpret_cast<const uint16_t*>(base + L.dense_lo);
  return v;
}

int32_t ref(int pool, uint32_t index) { return int32_t((uint32_t(pool) << 24) | index); }

// Round-to-nearest-even BF16 bits of a double (via FP32).
uint16_t bf16_bits(double v) {
  const float f = float(v);
  uint32_t b;
  std::memcpy(&b, &f, 4);
  b += 0x7fffu + ((b >> 16) & 1u);
  return uint16_t(b >> 16);
}
double bf16_value(uint16_t h) {
  const uint32_t b = uint32_t(h) << 16;
  float f;
  std::memcpy(&f, &b, 4);
  return f;
}

// Split-BF16 dense weights of the coarse patches, laid out as MFMA B
// fragments by K step (see NcchipDenseWeights): per angle tile, the rows of
// the union of its patches' supports cut into 8-column segments, two per step.
struct DenseWeights {
  std::vector<uint16_t> wh, wl;      // [K step][64 lanes][8]
  std::vector<int32_t> step0;        // first step of each tile, tiles + 1 entries
  std::vector<int32_t> segs;         // 2 per step: (y << 16 | x) of half-wave 0 and 1
  std::vector<double> wsum_eff;      // per angle
  std::vector<int32_t> common_runs;  // 4 ints per run of the common support
  int n_common = 0;
  // bound groups (see NcchipDenseWeights): one array of group first runs,
  // sample counts and runs, sizes 8, 16, 32 in turn
  std::vector<int32_t> groups;
  int32_t first_base[3] = {0, 0, 0}, n_base[3] = {0, 0, 0}, n_off = 0, runs_off = 0;
  int ymin = 0, xmin = 0, yspan = 0, xchunks = 0, max_run = 0;
  double w8_scale = 1;  // power of two with max |w_hi| * w8_scale <= 128
};
DenseWeights dense_weights(const std::vector<const Raster*>& patches, const std::vector<const RasterPool*>& pools) {
  DenseWeights d;
  int ymax = 0, xmax = 0;
  d.ymin = 1 << 30; d.xmin = 1 << 30; ymax = -(1 << 30); xmax = -(1 << 30);
  for (const Raster* p : patches) {
    d.ymin = std::min(d.ymin, p->rmin); ymax = std::max(ymax, p->rmax);
    d.xmin = std::min(d.xmin, p->cmin); xmax = std::max(xmax, p->cmax);
  }
  d.yspan = ymax - d.ymin + 1;
  d.xchunks = (xmax - d.xmin + 16) / 16;
  const size_t tiles = (patches.size() + 31) / 32;
  // segments of every tile: row by row over the union of its supports
  std::vector<std::vector<int>> first_seg(tiles, std::vector<int>(size_t(d.yspan), -1));
  std::vector<std::vector<int>> row_lo(tiles, std::vector<int>(size_t(d.yspan), 1 << 30));
  d.step0.assign(tiles + 1, 0);
  for (size_t t = 0; t < tiles; ++t) {
    std::vector<int> hi(size_t(d.yspan), -(1 << 30));
    for (size_t a = 32 * t; a < std::min(patches.size(), 32 * t + 32); ++a) {
      const Raster& p = *patches[a];
      const RasterPool& pool = *pools[a];
      for (uint32_t i = 0; i < p.count; ++i) {
        const int32_t rc = pool.rc[p.first + i];
        const int yi = (rc >> 16) - d.ymin, xi = int(int16_t(rc & 0xffff)) - d.xmin;
        row_lo[t][size_t(yi)] = std::min(row_lo[t][size_t(yi)], xi);
        hi[size_t(yi)] = std::max(hi[size_t(yi)], xi);
      }
    }
    std::vector<int32_t> seg;
    for (int y = 0; y < d.yspan; ++y) {
      if (row_lo[t][size_t(y)] > hi[size_t(y)]) continue;
      first_seg[t][size_t(y)] = int(seg.size());
      for (int x = row_lo[t][size_t(y)]; x <= hi[size_t(y)]; x += 8) seg.push_back((y << 16) | x);
    }
    if (seg.size() & 1) seg.push_back(0);  // a step's second half over zero weights
    d.step0[t + 1] = d.step0[t] + int(seg.size() / 2);
    d.segs.insert(d.segs.end(), seg.begin(), seg.end());
  }
  d.wh.assign(size_t(d.step0[tiles]) * 64 * 8, 0);
  d.wl.assign(d.wh.size(), 0);
  d.wsum_eff.assign(patches.size(), 0.0);
  double wmax = 0;
  for (size_t a = 0; a < patches.size(); ++a) {
    const Raster& p = *patches[a];
    const RasterPool& pool = *pools[a];
    const size_t t = a / 32, slot = a % 32;
    double ws = 0;
    for (uint32_t i = 0; i < p.count; ++i) {
      const int32_t rc = pool.rc[p.first + i];
      const int y = rc >> 16, x = int(int16_t(rc & 0xffff));
      const double w = pool.w[p.first + i];
      const uint16_t h = bf16_bits(w), l = bf16_bits(w - bf16_value(h));
      ws += bf16_value(h) + bf16_value(l);
      const int yi = y - d.ymin, xi = x - d.xmin;
      const int k = first_seg[t][size_t(yi)] + (xi - row_lo[t][size_t(yi)]) / 8;
      const int j = (xi - row_lo[t][size_t(yi)]) % 8;
      const size_t step = size_t(d.step0[t]) + size_t(k / 2), lane = size_t(k % 2) * 32 + slot;
      const size_t e = (step * 64 + lane) * 8 + size_t(j);
      d.wh[e] = h;
      d.wl[e] = l;
    }
    d.wsum_eff[a] = ws;
    for (uint32_t i = 0; i < p.count; ++i) wmax = std::max(wmax, std::abs(bf16_value(bf16_bits(pool.w[p.first + i]))));
    for (uint32_t k = 0; k < p.run_count; ++k) {
      const int32_t* run = &pool.runs[4 * size_t(p.run_first + k)];
      d.max_run = std::max(d.max_run, run[2] - run[1] + 1);
    }
  }
  if (wmax > 0) d.w8_scale = std::ldexp(1.0, int(std::floor(std::log2(128.0 / wmax))));
  // support common to every patch, as row runs
  const int xspan = xmax - d.xmin + 1;
  std::vector<int> hits(size_t(d.yspan) * xspan, 0);
  for (size_t a = 0; a < patches.size(); ++a) {
    const Raster& p = *patches[a];
    const RasterPool& pool = *pools[a];
    for (uint32_t i = 0; i < p.count; ++i) {
      const int32_t rc = pool.rc[p.first + i];
      ++hits[size_t((rc >> 16) - d.ymin) * xspan + (int(int16_t(rc & 0xffff)) - d.xmin)];
    }
  }
  for (int y = 0; y < d.yspan; ++y)
    for (int x = 0; x < xspan;) {
      if (hits[size_t(y) * xspan + x] != int(patches.size())) { ++x; continue; }
      const int b = x;
      while (x < xspan && hits[size_t(y) * xspan + x] == int(patches.size())) ++x;
      d.common_runs.insert(d.common_runs.end(), {y + d.ymin, b + d.xmin, x - 1 + d.xmin, 0});
      d.n_common += x - b;
    }
  // supports common to the patches of each bound group of 8, 16 and 32
  // consecutive angles (groups never straddle an angle tile), within the
  // rows and columns the group's patches span
  std::vector<int32_t> first, count, runs;
  for (int i = 0; i < 3; ++i) {
    const size_t gs = size_t(8) << i;
    d.first_base[i] = int32_t(first.size());
    d.n_base[i] = int32_t(count.size());
    for (size_t a0 = 0; a0 < patches.size(); a0 += gs) {
      const size_t a1 = std::min(patches.size(), a0 + gs);
      int gy0 = d.yspan, gy1 = -1, gx0 = xspan, gx1 = -1;
      for (size_t a = a0; a < a1; ++a) {
        gy0 = std::min(gy0, patches[a]->rmin - d.ymin); gy1 = std::max(gy1, patches[a]->rmax - d.ymin);
        gx0 = std::min(gx0, patches[a]->cmin - d.xmin); gx1 = std::max(gx1, patches[a]->cmax - d.xmin);
      }
      for (int y = gy0; y <= gy1; ++y) std::fill(&hits[size_t(y) * xspan + gx0], &hits[size_t(y) * xspan + gx1] + 1, 0);
      for (size_t a = a0; a < a1; ++a) {
        const Raster& p = *patches[a];
        const RasterPool& pool = *pools[a];
        for (uint32_t k = 0; k < p.count; ++k) {
          const int32_t rc = pool.rc[p.first + k];
          ++hits[size_t((rc >> 16) - d.ymin) * xspan + (int(int16_t(rc & 0xffff)) - d.xmin)];
        }
      }
      first.push_back(int32_t(runs.size() / 4));
      int n = 0;
      for (int y = gy0; y <= gy1; ++y)
        for (int x = gx0; x <= gx1;) {
          if (hits[size_t(y) * xspan + x] != int(a1 - a0)) { ++x; continue; }
          const int b = x;
          while (x <= gx1 && hits[size_t(y) * xspan + x] == int(a1 - a0)) ++x;
          runs.insert(runs.end(), {y + d.ymin, b + d.xmin, x - 1 + d.xmin, 0});
          n += x - b;
        }
      count.push_back(n);
    }
    first.push_back(int32_t(runs.size() / 4));
  }
  d.n_off = int32_t(first.size());
  d.runs_off = d.n_off + int32_t(count.size());
  d.runs_off = (d.runs_off + 3) & ~3;  // 16-byte aligned runs
  d.groups.assign(size_t(d.runs_off) + runs.size(), 0);
  std::copy(first.begin(), first.end(), d.groups.begin());
  std::copy(count.begin(), count.end(), d.groups.begin() + d.n_off);
  std::copy(runs.begin(), runs.end(), d.groups.begin() + d.runs_off);
  return d;
}

// Speculative refinement rounds of walks with no match quota: when the
// walks of one call hold at most kOneRound candidates together, each refines
// its whole list in one round; otherwise each starts with
// kUnboundedFirstRound of its next candidates, each later round twice as
// many.  Candidates whose unrefined pose an accepted match suppresses are
// never refined (select_refined checks them first), so bounded rounds let
// the first accepted matches suppress part of a long list before it is
// refined, which pays when the refinement is throughput-bound; a round costs
// at least one candidate's whole Gauss-Newton latency, which dominates
// smaller calls.  Which candidates the walks accept does not depend on the
// round sizes.
constexpr size_t kOneRound = 65536;
constexpr size_t kUnboundedFirstRound = 1024;

bool pose_less(const Pose& a, const Pose& b) {  // merge_identical order (matcher.cpp:242-244)
  if (a.row != b.row) return a.row < b.row;
  if (a.col != b.col) return a.col < b.col;
  return a.angle < b.angle;
}
bool score_order(const Pose& a, const Pose& b) {  // matcher.cpp:824-827, 1072-1075
  if (a.score != b.score) return a.score > b.score;
  return pose_less(a, b);
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
  check(hipHostMalloc(&ptr, n, hipHostMallocCoherent | hipHostMallocMapped), "hipHostMalloc");
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

namespace {
// Host state of one frame of a batch.
struct Job {
  FrameInput in;
  FindTrace* trace = nullptr;
  int64_t crop_r = 0, crop_c = 0, crop_b = 0, crop_e = 0;
  struct Lv { int w, h, kind, pitch; size_t off, bytes; };
  std::vector<Lv> lv;
  std::vector<int32_t> ranges;
  size_t off_p1 = 0, off_p2 = 0, off_nz = 0, off_planes = 0, off_ranges = 0;
  int64_t box[4] = {0, 0, -1, -1};  // search box {row_min, col_min, row_max, col_max} (matcher.cpp:376-383)
  // reduced domain: summed-area table, runs (int4) and per-crop-row run index
  size_t off_sat = 0, off_runs = 0, off_row = 0;
  size_t off_bits = 0;  // split-BF16 form: per top-level position, a bit per coarse angle at or above the threshold
  std::vector<int32_t> dom_runs, dom_row;
  std::vector<Pose> cand, full, refined, results;
  std::vector<int> verdict;
  int32_t storage0[4] = {0, 0, 0, 0};
  size_t pos = 0, batch_size = 0;
  bool done = false;
  double coarse_flops = 0, track_flops = 0, track_mfma_flops = 0;
  // find_models' background guard for this frame (matcher.cpp:839-846, 912-927; < 0: off)
  double guard_first = -1, guard = -1;
};
struct Timer { hipEvent_t begin = nullptr, end = nullptr; int stage = 0; };
}  // namespace

// What the back half of a batch needs from its front half.
struct FindContext::Plan {
  hipStream_t qs = nullptr;
  const NccModel* model = nullptr;
  std::shared_ptr<DeviceModel> dmodel;
  FindOptions o;
  FindTrace* traces = nullptr;
  size_t n_frames = 0, level_count = 0, top = 0, finest = 0, stride = 1, n_track = 0, A = 0;
  size_t cap = 0, cap_all = 0, table_size = 0, n_count_words = 0;
  std::vector<size_t> track_levels;
  std::vector<std::vector<int32_t>> refs;
  RasterPool qpool;
  int64_t coarse_original = -1;
  std::vector<Job> jobs;
  DenseWeights dense;
  bool use_mfma = false, track_on_mfma = false, wide = false, reduced = false;
  bool coarse_deferred = false;  // the coarse search runs in a shared persistent launch
  int max_bh = 0, max_bw = 0, F = 0, Wt_max = 0, Ht_max = 0;
  double scale = 1, phase = 0, radius = 0, t0 = 0;
  uint8_t* dev = nullptr;
  NccFrame* dframes = nullptr;
  NccMatch* lists = nullptr;
  int32_t* counts = nullptr;
  unsigned long long* extent = nullptr;
  unsigned long long* work = nullptr;
  NccPatchPool pools[2] = {};
  size_t off_key = 0, off_raw = 0, off_table = 0, off_wh = 0, off_wl = 0, off_wsum = 0, off_common = 0, off_groups = 0;
  size_t off_step0 = 0, off_segs = 0;
  size_t rb_full = 0, rb_counts = 0;
  std::vector<size_t> off_refs;
  uint8_t* rb = nullptr;
  std::vector<int32_t> frame_dims;  // top-level (w, h) per frame, for the coarse queue
  size_t off_small = 0, small = 0, off_bits = 0, n_bit_words = 0;
  std::vector<int> track_rc;        // per tracking level: ncchip_track_mfma's return (-6: vector kernel)
  NcchipCoarseTile tile{};          // split-BF16 coarse tile shape (rows == 0: the default shape)
  std::string tune_key;             // the model's autotuning key of this search
  bool tune_pending = false;        // the shape is tuned on this batch before its coarse launch
};

// A captured device part of a batch, replayed while its key repeats.
struct SlotGraph {
  std::string key;
  hipGraphExec_t exec = nullptr;
  std::shared_ptr<DeviceModel> dmodel;  // its device raster pool stays at the captured address
  std::vector<int> track_rc;
};

// A submitted batch: its buffers (reused by later batches in the same slot)
// and what the completion needs.
struct FindContext::Slot {
  DeviceBuffer work;
  PinnedBuffer stage, readback;
  hipEvent_t done = nullptr;  // the device-resident part has finished
  hipStream_t stream = nullptr;  // the device-resident part's stream (batches of both slots overlap)
  std::vector<Timer> timers;
  size_t timers_used = 0;
  int profiling = 0;          // stage mask
  bool timing = false;        // the open stage is timed
  bool busy = false;
  uint64_t ticket = 0;
  // the batch
  const NccModel* model = nullptr;
  std::shared_ptr<DeviceModel> dmodel;
  FindOptions o;
  FindTrace* traces = nullptr;
  size_t n_frames = 0;
  bool empty = false;  // the query range misses the model's angles
  std::vector<Job> jobs;
  RasterPool qpool;
  size_t cap = 0, cap_all = 0, n_track = 0, top = 0, finest = 0, stride = 1, off_key = 0;
  uint8_t* dev = nullptr;
  NccFrame* dframes = nullptr;
  NccMatch* lists = nullptr;
  NccMatch* hfull = nullptr;
  int32_t* hcounts = nullptr;
  uint8_t* rb = nullptr;
  NccPatchPool pools[2] = {};
  NcchipRefineParams rp{};
  std::vector<bool> track_mfma;  // per tracking level: launched on matrix cores
  double t0 = 0, t1 = 0, t2 = 0;
  std::unique_ptr<Plan> plan;  // between the two halves of enqueue
  std::vector<int> track_rc;   // the last batch's tracking kernel choices
  int coarse_tile = 0;         // the split-BF16 coarse tile shape that ran (pack_tile; 0: none)
  std::vector<SlotGraph> graphs;  // captured device parts, most recent first
};

FindContext::FindContext() {
  check(hipGetDevice(&device_), "hipGetDevice");
  check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
  check(hipStreamCreateWithFlags(&stream2_, hipStreamNonBlocking), "hipStreamCreate");
  check(hipEventCreateWithFlags(&uploaded_, hipEventDisableTiming), "hipEventCreate");
  for (size_t i = 0; i < 2; ++i) {
    auto& sl = slots_[i];
    sl = std::make_unique<Slot>();
    check(hipStreamCreateWithFlags(&slot_streams_[i], hipStreamNonBlocking), "hipStreamCreate");
    sl->stream = slot_streams_[i];
    check(hipEventCreateWithFlags(&sl->done, hipEventDisableTiming), "hipEventCreate");
  }
}
FindContext::~FindContext() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  if (stream2_) (void)hipStreamSynchronize(stream2_);
  for (hipStream_t q : slot_streams_)
    if (q) (void)hipStreamSynchronize(q);
  for (auto& sl : slots_) {
    if (!sl) continue;
    for (auto& t : sl->timers) { (void)hipEventDestroy(t.begin); (void)hipEventDestroy(t.end); }
    if (sl->done) (void)hipEventDestroy(sl->done);
  }
  if (uploaded_) (void)hipEventDestroy(uploaded_);
  for (hipEvent_t e : queue_events_)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : tune_events_)
    if (e) (void)hipEventDestroy(e);
  for (hipStream_t q : slot_streams_)
    if (q) (void)hipStreamDestroy(q);
  if (stream2_) (void)hipStreamDestroy(stream2_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

void FindContext::reset() {
  check(hipStreamSynchronize(stream_), "sync");
  check(hipStreamSynchronize(stream2_), "sync");
  for (hipStream_t q : slot_streams_) check(hipStreamSynchronize(q), "sync");
  for (auto& sl : slots_) {
    sl->busy = false;
    sl->traces = nullptr;
    for (auto& J : sl->jobs) J.trace = nullptr;
  }
  cur_ = nullptr;
}

void FindContext::time_begin(int stage, hipStream_t stream) {
  Slot& S = *cur_;
  S.timing = S.profiling & (1 << stage);
  if (!S.timing) return;
  if (S.timers_used == S.timers.size()) {
    Timer t;
    check(hipEventCreate(&t.begin), "hipEventCreate");
    check(hipEventCreate(&t.end), "hipEventCreate");
    S.timers.push_back(t);
  }
  S.timers[S.timers_used].stage = stage;
  check(hipEventRecord(S.timers[S.timers_used].begin, stream), "hipEventRecord");
}
void FindContext::time_end(hipStream_t stream) {
  Slot& S = *cur_;
  if (!S.timing) return;
  check(hipEventRecord(S.timers[S.timers_used].end, stream), "hipEventRecord");
  ++S.timers_used;
}
// Every timed event of the slot has completed when this runs (the slot's
// device part and its refinement rounds were waited for).
void FindContext::time_collect(FindTrace* trace) {
  Slot& S = *cur_;
  if (S.profiling && trace) {
    for (size_t i = 0; i < S.timers_used; ++i) {
      float ms = 0;
      check(hipEventElapsedTime(&ms, S.timers[i].begin, S.timers[i].end), "hipEventElapsedTime");
      trace->kernel_us[S.timers[i].stage] += 1000.0 * ms;
    }
  }
  S.timers_used = 0;
}

// Copies a host image (rows of `row` bytes, `pitch` apart) to image_ on the
// context stream as one transfer: directly when the rows are contiguous,
// else gathered into pinned staging first (a strided copy from pageable
// memory moves row by row).
void FindContext::upload_image(const uint8_t* image, size_t row, size_t height, size_t pitch) {
  const size_t bytes = row * height;
  if (pitch == row) {
    check(hipMemcpyAsync(image_.ptr, image, bytes, hipMemcpyHostToDevice, stream_), "image upload");
    return;
  }
  check(hipStreamSynchronize(stream_), "sync");  // the staging of an earlier upload is free
  upload_stage_.reserve(bytes);
  auto* dst = static_cast<uint8_t*>(upload_stage_.ptr);
  for (size_t r = 0; r < height; ++r) std::memcpy(dst + r * row, image + r * pitch, row);
  check(hipMemcpyAsync(image_.ptr, dst, bytes, hipMemcpyHostToDevice, stream_), "image upload");
}

std::vector<Pose> FindContext::find_host(const NccModel& model, const uint8_t* image, size_t width, size_t height,
                                         size_t pitch, const FindOptions& options, FindTrace* trace, int pixel_bytes,
                                         const Region* domain) {
  if (!width || !height) raise(kInvalidArgument, "image", "one nonempty single-channel image required");
  const size_t row = width * size_t(pixel_bytes), bytes = row * height;
  if (image_.bytes < bytes) image_ = DeviceBuffer(align_up(bytes, 1 << 20));
  upload_image(image, row, height, pitch);
  FrameInput in{image_.as<uint8_t>(), width, height, row};
  in.pixel_bytes = pixel_bytes;
  in.domain = domain;
  auto out = find_batch(model, &in, 1, options, trace);
  return std::move(out.front());
}

std::vector<ModelMatch> FindContext::find_models_host(const NccModel* const* models, const FindOptions* options,
                                                      size_t n_models, const uint8_t* image, size_t width,
                                                      size_t height, size_t pitch, const Selection& selection,
                                                      bool per_model_quota_before_overlap, int pixel_bytes,
                                                      const Region* domain) {
  if (!width || !height) raise(kInvalidArgument, "image", "one nonempty single-channel image required");
  const size_t row = width * size_t(pixel_bytes), bytes = row * height;
  check(hipStreamSynchronize(stream2_), "sync");
  if (image_.bytes < bytes) {
    check(hipStreamSynchronize(stream_), "sync");
    image_ = DeviceBuffer(align_up(bytes, 1 << 20));
  }
  upload_image(image, row, height, pitch);
  FrameInput in{image_.as<uint8_t>(), width, height, row};
  in.pixel_bytes = pixel_bytes;
  in.domain = domain;
  return find_models(models, options, n_models, in, selection, per_model_quota_before_overlap);
}

std::vector<Pose> FindContext::find(const NccModel& model, const uint8_t* image, size_t W, size_t H, size_t pitch,
                                    const FindOptions& req, FindTrace* trace) {
  FrameInput in{image, W, H, pitch};
  auto out = find_batch(model, &in, 1, req, trace);
  return std::move(out.front());
}


std::vector<std::vector<Pose>> FindContext::find_batch(const NccModel& model, const FrameInput* inputs, size_t n_frames,
                                                       const FindOptions& req, FindTrace* traces) {
  if (n_frames > NCC_MAX_BATCH) {  // larger batches run as consecutive batches
    std::vector<std::vector<Pose>> all;
    for (size_t b = 0; b < n_frames; b += NCC_MAX_BATCH) {
      const size_t n = std::min<size_t>(NCC_MAX_BATCH, n_frames - b);
      for (auto& part : find_batch(model, inputs + b, n, req, traces ? traces + b : nullptr)) all.push_back(std::move(part));
    }
    return all;
  }
  return complete_batch(submit_batch(model, inputs, n_frames, req, traces));
}

uint64_t FindContext::submit_batch(const NccModel& model, const FrameInput* inputs, size_t n_frames,
                                   const FindOptions& req, FindTrace* traces) {
  if (n_frames > NCC_MAX_BATCH) raise(kLimitExceeded, "Images", "too many frames for one submitted batch");
  const uint64_t ticket = next_ticket_;
  Slot& S = *slots_[ticket % 2];
  if (S.busy) raise(kInvalidState, "batch", "two batches are pending; complete the oldest first");
  cur_ = &S;
  enqueue(S, model, inputs, n_frames, req, traces);
  S.ticket = ticket;
  S.busy = true;
  ++next_ticket_;
  return ticket;
}

void FindContext::enqueue(Slot& S, const NccModel& model, const FrameInput* inputs, size_t n_frames,
                          const FindOptions& req, FindTrace* traces) {
  if (!enqueue_front(S, model, inputs, n_frames, req, traces)) return;
  hipStream_t qs = S.plan->qs;
  if (S.plan->tune_pending) {  // the first batch of its kind: tuned, and launched kernel by kernel
    front_launches(S);
    tune_coarse(S);
    enqueue_back(S);
  } else if (graphs_enabled_ && S.plan->o.graph && !S.profiling && !S.plan->reduced) {
    launch_graph(S);
  } else {
    front_launches(S);
    enqueue_back(S);
  }
  check(hipEventRecord(S.done, qs), "hipEventRecord");
}

// The key of a planned batch's device work: everything its kernel arguments
// depend on (the frames' pixels and addresses reach the kernels through the
// uploaded descriptors, not through arguments).
std::string FindContext::graph_key(Slot& S) {
  const Plan& P = *S.plan;
  std::string k;
  auto put = [&](const auto& v) { k.append(reinterpret_cast<const char*>(&v), sizeof(v)); };
  put(P.dmodel.get()); put(S.work.ptr); put(S.stage.ptr); put(S.readback.ptr); put(P.qs);
  const FindOptions& o = P.o;
  put(o.start); put(o.extent); put(o.min_score); put(o.max_candidates); put(o.subpixel); put(o.highest); put(o.lowest);
  put(o.form); put(o.screen); put(o.track_mfma); put(o.query_cache); put(o.whole_level0);
  put(P.tile);
  put(P.n_frames);
  for (const Job& J : P.jobs) { put(J.in.width); put(J.in.height); put(J.in.pitch); put(J.in.pixel_bytes); }
  return k;
}

// The device work of a planned batch as a HIP graph: captured the first time
// a key is seen (launching it kernel by kernel inside the capture), replayed
// afterwards with the host bookkeeping of the back half run without launches.
void FindContext::launch_graph(Slot& S) {
  Plan& P = *S.plan;
  hipStream_t qs = P.qs;
  const std::string key = graph_key(S);
  for (size_t i = 0; i < S.graphs.size(); ++i) {
    if (S.graphs[i].key != key) continue;
    std::rotate(S.graphs.begin(), S.graphs.begin() + long(i), S.graphs.begin() + long(i) + 1);  // most recent first
    SlotGraph& g = S.graphs.front();
    P.track_rc = g.track_rc;
    dry_ = true;
    try {
      enqueue_back(S);
    } catch (...) {
      dry_ = false;
      throw;
    }
    dry_ = false;
    check(hipGraphLaunch(g.exec, qs), "hipGraphLaunch");
    return;
  }
  SlotGraph entry;
  entry.key = key;
  entry.dmodel = P.dmodel;
  check(hipStreamBeginCapture(qs, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
  hipGraph_t graph = nullptr;
  try {
    front_launches(S);
    enqueue_back(S);
  } catch (...) {
    (void)hipStreamEndCapture(qs, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    throw;
  }
  check(hipStreamEndCapture(qs, &graph), "hipStreamEndCapture");
  const hipError_t e = hipGraphInstantiate(&entry.exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  check(e, "hipGraphInstantiate");
  entry.track_rc = S.track_rc;
  check(hipGraphLaunch(entry.exec, qs), "hipGraphLaunch");
  constexpr size_t kGraphsPerSlot = 8;
  if (S.graphs.size() == kGraphsPerSlot) {  // the least recent is idle: its slot's last batch completed
    (void)hipGraphExecDestroy(S.graphs.back().exec);
    S.graphs.pop_back();
  }
  S.graphs.insert(S.graphs.begin(), std::move(entry));
}

// The front half of a batch: host planning, the small uploads, the search
// pyramid and the top-level statistics.  Returns false when there is nothing
// to search (the slot is then complete); else the plan for the back half is
// left in S.plan.
bool FindContext::enqueue_front(Slot& S, const NccModel& model, const FrameInput* inputs, size_t n_frames,
                                const FindOptions& req, FindTrace* traces) {
  const double t0 = now_us();
  // the slot's stream; work queued on the context stream before (image
  // uploads) completes first
  hipStream_t qs = S.stream ? S.stream : stream_;
  if (qs != stream_) {
    check(hipEventRecord(uploaded_, stream_), "hipEventRecord");
    check(hipStreamWaitEvent(qs, uploaded_, 0), "hipStreamWaitEvent");
  }
  S.profiling = traces ? req.profile : 0;
  S.timers_used = 0;
  S.jobs.clear();
  S.empty = false;
  for (size_t f = 0; f < n_frames && traces; ++f) traces[f] = FindTrace{};
  S.plan.reset();
  auto finish_empty = [&]() {
    S.model = &model; S.o = req; S.traces = traces; S.n_frames = n_frames; S.empty = true;
  };
  if (!n_frames) return finish_empty(), false;
  // ---- options and levels, shared by the batch (matcher.cpp:345-374)
  FindOptions o = req;
  for (size_t f = 0; f < n_frames; ++f)
    if (!inputs[f].width || !inputs[f].height) raise(kInvalidArgument, "image", "one nonempty single-channel image required");
  // u16 ("uint2") frames: level 0 in u16, block sums in u32 from level 1
  const bool wide = inputs[0].pixel_bytes == 2;
  for (size_t f = 0; f < n_frames; ++f)
    if (inputs[f].pixel_bytes != inputs[0].pixel_bytes || (inputs[f].pixel_bytes != 1 && inputs[f].pixel_bytes != 2) ||
        (wide && inputs[f].pitch % 2))
      raise(kInvalidArgument, "image", "frames of one batch share one pixel type (byte or uint2, even pitch)");
  if (model.levels.empty()) raise(kInvalidState, "model", "model is empty");
  if (o.highest < 0 || o.lowest < 1) raise(kInvalidArgument, "NumLevels", "invalid pyramid levels");
  const size_t level_count = o.highest ? std::min(model.levels.size(), size_t(o.highest)) : model.levels.size();
  if (inputs[0].pixel_bytes == 2 && level_count > 9)
    raise(kUnsupported, "NumLevels", "uint2 search images support at most 9 pyramid levels");
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
  if (cs > ce + 1e-12) return finish_empty(), false;
  o.start = cs;
  o.extent = std::max(0., ce - cs);
  for (size_t f = 0; f < n_frames; ++f) {
    const size_t W = inputs[f].width, H = inputs[f].height;
    if (W > size_t(std::numeric_limits<int>::max()) || H > size_t(std::numeric_limits<int>::max()) || H > 32000000 / W)
      raise(kLimitExceeded, "image", "search image exceeds pixel budget");
  }
  // reduced domains: one frame per find; an empty domain finds nothing (matcher.cpp:381)
  const bool reduced = inputs[0].domain != nullptr;
  for (size_t f = 0; f < n_frames; ++f)
    if ((inputs[f].domain != nullptr) != reduced || (reduced && n_frames != 1))
      raise(kInvalidArgument, "image", "reduced search domains are supported for single-frame finds");
  if (reduced && inputs[0].domain->empty()) return finish_empty(), false;
  auto dmodel = device_model(model, qs);

  // ---- query patches per level (matcher.cpp:463-482), shared by the batch
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
  const LevelGeom& base = model.levels.front();
  const double radius = std::hypot(std::max(base.orow, double(base.h - 1) - base.orow),
                                   std::max(base.ocol, double(base.w - 1) - base.ocol));
  const double scale = std::ldexp(1., int(top)), phase = (scale - 1) * .5;
  const size_t A = refs[top].size();
  int max_bh = 0, max_bw = 0;
  for (size_t i = 0; i < A; ++i) {
    const Raster& p = raster_of(refs[top][i]);
    max_bh = std::max(max_bh, p.rmax - p.rmin);
    max_bw = std::max(max_bw, p.cmax - p.cmin);
  }

  if (o.form == kFormAuto) o.form = wide ? NCCHIP_FORM_FP32 : NCCHIP_FORM_BF16X3;
  if (o.form < NCCHIP_FORM_FP32 || o.form > NCCHIP_FORM_FP8_SCREEN) raise(kInvalidArgument, "form", "unknown correlation form");
  if (wide && (o.form == NCCHIP_FORM_BF16X3 || o.form == NCCHIP_FORM_FP8_SCREEN))
    raise(kUnsupported, "form", "the split-BF16 and FP8-screened forms need byte search images");
  if (wide) {
    // The window moments are exact 64-bit integers: sum((S - D)^2) over a
    // window must stay below 2^61 at every level searched
    for (size_t l = finest; l <= top; ++l) {
      double n_max = 0;
      for (const int32_t r : refs[l]) n_max = std::max(n_max, double(raster_of(r).count));
      const double x = 65535. * std::ldexp(1., 2 * int(l));
      if (n_max * x * x > 0x1p61)
        raise(kUnsupported, "NumLevels", "template too large for exact uint2 window moments at this pyramid level");
    }
    // the FP32 form holds S - C exactly in FP32 up to level 4
    if (o.form == NCCHIP_FORM_FP32 && top > 4) o.form = NCCHIP_FORM_FP64;
  }
  // the matrix-core tracking kernel goes with the split-BF16 and FP8-screened
  // forms; the plain FP32 and FP64 forms keep the whole search off matrix cores
  const bool track_on_mfma = o.track_mfma && (o.form == NCCHIP_FORM_BF16X3 || o.form == NCCHIP_FORM_FP8_SCREEN);
  bool use_mfma = o.form == NCCHIP_FORM_BF16X3 || o.form == NCCHIP_FORM_FP8_SCREEN;
  DenseWeights dense;
  if (use_mfma) {
    // the union bounding box of the coarse patches must fit the kernel's LDS
    // tile at this level; otherwise the coarse search runs the FP32 form
    int ymin = 1 << 30, ymax = -(1 << 30), xmin = 1 << 30, xmax = -(1 << 30);
    for (size_t i = 0; i < A; ++i) {
      const Raster& p = raster_of(refs[top][i]);
      ymin = std::min(ymin, p.rmin); ymax = std::max(ymax, p.rmax);
      xmin = std::min(xmin, p.cmin); xmax = std::max(xmax, p.cmax);
    }
    const int top_kind = wide || top > 4 ? NCC_U32 : (top ? NCC_U16 : NCC_U8);
    if (!ncchip_coarse_mfma_fits(int(top), top_kind, ymax - ymin + 1, (xmax - xmin + 16) / 16,
                                 o.form == NCCHIP_FORM_FP8_SCREEN, nullptr)) {
      o.form = NCCHIP_FORM_FP32;
      use_mfma = false;
    }
  }
  if (use_mfma) {
    std::vector<const Raster*> ps;
    std::vector<const RasterPool*> pp;
    for (size_t i = 0; i < A; ++i) {
      ps.push_back(&raster_of(refs[top][i]));
      pp.push_back((refs[top][i] >> 24) ? &qpool : &model.pool);
    }
    dense = dense_weights(ps, pp);
  }

  // ---- per-frame geometry: crop (matcher.cpp:383-393), pyramid levels, coarse anchor ranges (:549-553)
  std::vector<Job> jobs(n_frames);
  const int64_t lattice = int64_t(size_t(1) << (level_count - 1));
  const int64_t halo = int64_t(std::ceil(radius)) + 4 * lattice;
  for (size_t f = 0; f < n_frames; ++f) {
    Job& J = jobs[f];
    J.in = inputs[f];
    J.trace = traces ? &traces[f] : nullptr;
    const int64_t W = int64_t(J.in.width), H = int64_t(J.in.height);
    if (const Region* D = J.in.domain) {
      J.box[0] = D->row_min; J.box[1] = D->col_min; J.box[2] = D->row_max; J.box[3] = D->col_max;
    } else {
      J.box[0] = 0; J.box[1] = 0; J.box[2] = H - 1; J.box[3] = W - 1;
    }
    J.crop_r = std::max<int64_t>(0, J.box[0] - halo) / lattice * lattice;
    J.crop_c = std::max<int64_t>(0, J.box[1] - halo) / lattice * lattice;
    J.crop_b = std::min<int64_t>(H, ((J.box[2] + halo) / lattice + 1) * lattice);
    J.crop_e = std::min<int64_t>(W, ((J.box[3] + halo) / lattice + 1) * lattice);
    J.lv.resize(level_count);
    J.lv[0] = {int(J.crop_e - J.crop_c), int(J.crop_b - J.crop_r), wide ? NCC_U16 : NCC_U8,
               int(J.in.pitch / size_t(J.in.pixel_bytes)), 0, 0};
    for (size_t l = 1; l < level_count; ++l) {
      const bool u16 = !wide && l <= 4;  // block sums of u8 pixels fit 16 bits up to level 4
      J.lv[l].w = std::max(1, J.lv[l - 1].w / 2);
      J.lv[l].h = std::max(1, J.lv[l - 1].h / 2);
      J.lv[l].kind = u16 ? NCC_U16 : NCC_U32;
      J.lv[l].pitch = J.lv[l].w;
      J.lv[l].bytes = size_t(J.lv[l].w) * J.lv[l].h * (u16 ? 2 : 4);
    }
    J.ranges.resize(4 * A);
    for (size_t i = 0; i < A; ++i) {
      const Raster& p = raster_of(refs[top][i]);
      const double margin = scale == 1 ? .5 : 2 * scale;
      int32_t* rg = &J.ranges[4 * i];
      rg[0] = std::max({0, -p.rmin, int(std::ceil((J.box[0] - margin - J.crop_r - phase) / scale - p.orow))});
      rg[1] = std::min(J.lv[top].h - p.rmax, int(std::floor((J.box[2] + margin - J.crop_r - phase) / scale - p.orow)) + 1);
      rg[2] = std::max({0, -p.cmin, int(std::ceil((J.box[1] - margin - J.crop_c - phase) / scale - p.ocol))});
      rg[3] = std::min(J.lv[top].w - p.cmax, int(std::floor((J.box[3] + margin - J.crop_c - phase) / scale - p.ocol)) + 1);
      J.coarse_flops += 2.0 * double(p.count) * std::max(0, rg[1] - rg[0]) * std::max(0, rg[3] - rg[2]);
    }
    if (const Region* D = J.in.domain) {
      // the domain's runs (all inside the crop: the crop covers the box) and
      // the index of each crop row's first run
      J.dom_row.assign(size_t(J.lv[0].h) + 1, 0);
      size_t k = 0;
      for (int r = 0; r <= J.lv[0].h; ++r) {
        while (k < D->runs.size() && D->runs[k].row < J.crop_r + r) ++k;
        J.dom_row[size_t(r)] = int32_t(k);
      }
      for (const RegionRun& run : D->runs) {
        J.dom_runs.push_back(run.row); J.dom_runs.push_back(run.cb); J.dom_runs.push_back(run.ce); J.dom_runs.push_back(0);
      }
    }
  }

  // ---- tracking levels (matcher.cpp:630-645): l = top .. finest+1, ending at
  // l == 1 with subpixel (the fit takes over from level 1)
  std::vector<size_t> track_levels;
  for (size_t l = top; l > finest; --l) {
    if (l == 1 && o.subpixel) break;
    track_levels.push_back(l);
  }
  const size_t n_track = track_levels.size();

  // ---- workspace layout.  Candidate lists: list 0 = coarse peaks, list k =
  // the kept poses after tracking level k; frame f's entries at f * cap.
  const size_t cap = o.max_candidates + 1;  // one past the budget detects overflow
  const size_t cap_all = cap * n_frames;
  size_t table_size = 2;
  while (table_size < 2 * cap) table_size *= 2;
  const PoolLayout QL(qpool);
  size_t off = 0;
  for (auto& J : jobs) {
    const int Wt = J.lv[top].w, Ht = J.lv[top].h;
    for (size_t l = 1; l < level_count; ++l) { J.lv[l].off = off; off = align_up(off + J.lv[l].bytes); }
    J.off_p1 = off; off = align_up(off + size_t(Wt + 1) * Ht * 8);
    J.off_p2 = off; off = align_up(off + size_t(Wt + 1) * Ht * 8);
    J.off_nz = off; off = align_up(off + size_t(Wt + 1) * Ht * 4);
    J.off_planes = off; off = align_up(off + A * size_t(Ht) * Wt * 8);
    if (J.in.domain) { J.off_sat = off; off = align_up(off + (size_t(J.lv[0].h) + 1) * (size_t(J.lv[0].w) + 1) * 4); }
  }
  // the peak bits of every frame in one region (one fill clears them)
  const size_t bit_words = (A + 31) / 32;
  const size_t off_bits = off;
  for (auto& J : jobs) {
    J.off_bits = off;
    if (use_mfma) off += size_t(J.lv[top].h) * size_t(J.lv[top].w) * bit_words * 4;
  }
  const size_t n_bit_words = (off - off_bits) / 4;
  off = align_up(off);
  const size_t off_lists = off; off = align_up(off + (n_track + 1) * cap_all * sizeof(NccMatch));
  const size_t off_raw = off; off = align_up(off + cap_all * sizeof(NccMatch));
  const size_t off_key = off; off = align_up(off + cap_all * 4);
  // stage counts, then one u64 work counter per tracking level (copied together)
  const size_t n_count_words = align_up((n_track + 1) * n_frames * 4, 8) / 4 + 2 * std::max<size_t>(n_track, 1);
  const size_t off_counts = off; off = align_up(off + n_count_words * 4);
  const size_t off_table = off; off = align_up(off + std::max<size_t>(n_track, 1) * n_frames * table_size * 4);
  // small uploads: refs of every level, per-frame coarse ranges, query pool,
  // candidate extents, frame descriptors
  std::vector<size_t> off_refs(level_count);
  const size_t off_small = off;
  size_t small = 0;
  for (size_t l = 0; l < level_count; ++l) { off_refs[l] = off_small + small; small = align_up(small + refs[l].size() * 4 + 4); }
  for (auto& J : jobs) { J.off_ranges = off_small + small; small = align_up(small + J.ranges.size() * 4); }
  for (auto& J : jobs) {
    if (!J.in.domain) continue;
    J.off_runs = off_small + small; small = align_up(small + J.dom_runs.size() * 4);
    J.off_row = off_small + small; small = align_up(small + J.dom_row.size() * 4);
  }
  const size_t off_qpool = off_small + small; small = align_up(small + QL.total);
  const size_t off_extent = off_small + small; small = align_up(small + n_frames * 32);
  const size_t off_wh = off_small + small; small = align_up(small + dense.wh.size() * 2);
  const size_t off_wl = off_small + small; small = align_up(small + dense.wl.size() * 2);
  const size_t off_step0 = off_small + small; small = align_up(small + dense.step0.size() * 4);
  const size_t off_segs = off_small + small; small = align_up(small + dense.segs.size() * 4);
  const size_t off_wsum = off_small + small; small = align_up(small + dense.wsum_eff.size() * 8);
  const size_t off_common = off_small + small; small = align_up(small + dense.common_runs.size() * 4);
  const size_t off_groups = off_small + small; small = align_up(small + dense.groups.size() * 4);
  const size_t off_frames = off_small + small; small = align_up(small + n_frames * sizeof(NccFrame));
  off = off_small + small;
  if (S.work.bytes < off) {
    check(hipStreamSynchronize(qs), "sync");
    S.work = DeviceBuffer(align_up(off + off / 4, 1 << 20));
  }
  uint8_t* dev = S.work.as<uint8_t>();
  // Host-visible (pinned, coherent) staging: the refinement rounds' candidate
  // lists are read by the kernels in place and the full scores and refined
  // poses are written back in place.  A staging area is rewritten only after
  // the synchronisation that ends the launch reading it.
  S.stage.reserve(std::max(small, cap_all * (sizeof(NccMatch) + 4)) + 256);
  const size_t rb_full = align_up(cap_all * (sizeof(NccMatch) + 4));
  const size_t rb_counts = rb_full + align_up(cap_all * sizeof(NccMatch));
  S.readback.reserve(rb_counts + n_count_words * 4 + 64);
  auto* rb = static_cast<uint8_t*>(S.readback.ptr);
  const NccPatchPool pools[2] = {dmodel->pool.view, view_at(dev + off_qpool, QL)};
  auto* dframes = reinterpret_cast<NccFrame*>(dev + off_frames);
  auto* lists = reinterpret_cast<NccMatch*>(dev + off_lists);
  auto* counts = reinterpret_cast<int32_t*>(dev + off_counts);
  auto* work = reinterpret_cast<unsigned long long*>(dev + off_counts + align_up((n_track + 1) * n_frames * 4, 8));
  auto* extent = reinterpret_cast<unsigned long long*>(dev + off_extent);
  {
    uint8_t* st = static_cast<uint8_t*>(S.stage.ptr);
    for (size_t l = 0; l < level_count; ++l)
      std::memcpy(st + (off_refs[l] - off_small), refs[l].data(), refs[l].size() * 4);
    for (auto& J : jobs) std::memcpy(st + (J.off_ranges - off_small), J.ranges.data(), J.ranges.size() * 4);
    for (auto& J : jobs) {
      if (!J.in.domain) continue;
      std::memcpy(st + (J.off_runs - off_small), J.dom_runs.data(), J.dom_runs.size() * 4);
      std::memcpy(st + (J.off_row - off_small), J.dom_row.data(), J.dom_row.size() * 4);
    }
    serialize(qpool, QL, st + (off_qpool - off_small));
    if (use_mfma) {
      std::memcpy(st + (off_wh - off_small), dense.wh.data(), dense.wh.size() * 2);
      std::memcpy(st + (off_wl - off_small), dense.wl.data(), dense.wl.size() * 2);
      std::memcpy(st + (off_step0 - off_small), dense.step0.data(), dense.step0.size() * 4);
      std::memcpy(st + (off_segs - off_small), dense.segs.data(), dense.segs.size() * 4);
      std::memcpy(st + (off_wsum - off_small), dense.wsum_eff.data(), dense.wsum_eff.size() * 8);
      std::memcpy(st + (off_common - off_small), dense.common_runs.data(), dense.common_runs.size() * 4);
      std::memcpy(st + (off_groups - off_small), dense.groups.data(), dense.groups.size() * 4);
    }
    auto* ext = reinterpret_cast<double*>(st + (off_extent - off_small));
    auto* fr = reinterpret_cast<NccFrame*>(st + (off_frames - off_small));
    for (size_t f = 0; f < n_frames; ++f) {
      const Job& J = jobs[f];
      // the host's running min/max start at (h, 0, w, 0) (matcher.cpp:612-618)
      ext[4 * f + 0] = J.lv[0].h; ext[4 * f + 1] = 0; ext[4 * f + 2] = J.lv[0].w; ext[4 * f + 3] = 0;
      NccFrame d{};
      for (size_t l = 0; l < level_count; ++l) {
        NccLevel& L = d.levels[l];
        L.data = l ? static_cast<const void*>(dev + J.lv[l].off)
                   : static_cast<const void*>(J.in.image + J.crop_r * int64_t(J.in.pitch) +
                                              J.crop_c * int64_t(J.in.pixel_bytes));
        L.kind = J.lv[l].kind; L.pitch = J.lv[l].pitch; L.width = J.lv[l].w; L.height = J.lv[l].h;
        L.shift = int32_t(2 * l);
      }
      d.dc_src = J.in.image + J.crop_r * int64_t(J.in.pitch) + J.crop_c * int64_t(J.in.pixel_bytes);
      d.dc_wide = wide;
      d.storage0[0] = 0; d.storage0[1] = 0; d.storage0[2] = J.lv[0].h; d.storage0[3] = J.lv[0].w;
      d.crop_r = double(J.crop_r); d.crop_c = double(J.crop_c);
      d.image_h = int32_t(J.in.height); d.image_w = int32_t(J.in.width);
      d.p1 = reinterpret_cast<uint64_t*>(dev + J.off_p1);
      d.p2 = reinterpret_cast<uint64_t*>(dev + J.off_p2);
      d.nz = reinterpret_cast<uint32_t*>(dev + J.off_nz);
      d.planes = reinterpret_cast<double*>(dev + J.off_planes);
      d.peak_bits = use_mfma ? reinterpret_cast<uint32_t*>(dev + J.off_bits) : nullptr;
      d.ranges = reinterpret_cast<const int32_t*>(dev + J.off_ranges);
      if (const Region* D = J.in.domain) {
        d.dom_sat = reinterpret_cast<uint32_t*>(dev + J.off_sat);
        d.dom_runs = reinterpret_cast<const int32_t*>(dev + J.off_runs);
        d.dom_row = reinterpret_cast<const int32_t*>(dev + J.off_row);
        for (int i = 0; i < 4; ++i) d.dom_box[i] = int32_t(J.box[i]);
        d.dom_sparse = double(D->area) < .7 * double(D->row_max - D->row_min + 1) * double(D->col_max - D->col_min + 1);
      }
      fr[f] = d;
    }
  }
  auto max_dims = [&](size_t l) {
    int w = 0, h = 0;
    for (const auto& J : jobs) { w = std::max(w, J.lv[l].w); h = std::max(h, J.lv[l].h); }
    return std::pair<int, int>(w, h);
  };
  const int F = int(n_frames);

  const auto [Wt_max, Ht_max] = max_dims(top);
  // ---- the plan of the back half
  S.plan = std::make_unique<Plan>();
  Plan& P = *S.plan;
  P.qs = qs; P.model = &model; P.dmodel = dmodel; P.o = o; P.traces = traces; P.n_frames = n_frames;
  P.level_count = level_count; P.top = top; P.finest = finest; P.stride = stride; P.n_track = n_track; P.A = A;
  P.cap = cap; P.cap_all = cap_all; P.table_size = table_size; P.n_count_words = n_count_words;
  P.track_levels = std::move(track_levels); P.refs = std::move(refs); P.qpool = std::move(qpool);
  P.coarse_original = coarse_original; P.jobs = std::move(jobs); P.dense = std::move(dense);
  P.use_mfma = use_mfma; P.track_on_mfma = track_on_mfma; P.wide = wide; P.reduced = reduced;
  P.max_bh = max_bh; P.max_bw = max_bw; P.F = F; P.Wt_max = Wt_max; P.Ht_max = Ht_max;
  P.scale = scale; P.phase = phase; P.radius = radius; P.t0 = t0;
  P.dev = dev; P.dframes = dframes; P.lists = lists; P.counts = counts; P.extent = extent; P.work = work;
  P.pools[0] = pools[0]; P.pools[1] = pools[1];
  P.off_key = off_key; P.off_raw = off_raw; P.off_table = off_table; P.off_wh = off_wh; P.off_wl = off_wl;
  P.off_wsum = off_wsum; P.off_common = off_common; P.off_groups = off_groups; P.off_step0 = off_step0; P.off_segs = off_segs; P.rb_full = rb_full; P.rb_counts = rb_counts;
  P.off_refs = std::move(off_refs); P.rb = rb;
  P.off_small = off_small; P.small = small; P.off_bits = off_bits; P.n_bit_words = n_bit_words;
  if (use_mfma) choose_tile(S);
  return true;
}

// The front half's device work: the small uploads, the clearing of counters,
// hash tables and peak bits, the reduced-domain table, the search pyramid
// and the top-level row prefixes.
void FindContext::front_launches(Slot& S) {
  Plan& P = *S.plan;
  cur_ = &S;
  hipStream_t qs = P.qs;
  uint8_t* const dev = P.dev;
  const size_t n_frames = P.n_frames, n_track = P.n_track, table_size = P.table_size, level_count = P.level_count;
  const size_t top = P.top, n_count_words = P.n_count_words, n_bit_words = P.n_bit_words, off_bits = P.off_bits;
  const size_t off_table = P.off_table;
  const std::vector<Job>& jobs = P.jobs;
  const int F = P.F, Ht_max = P.Ht_max;
  const bool wide = P.wide, reduced = P.reduced;
  NccFrame* const dframes = P.dframes;
  int32_t* const counts = P.counts;
  auto max_dims = [&](size_t l) {
    int w = 0, h = 0;
    for (const auto& J : jobs) { w = std::max(w, J.lv[l].w); h = std::max(h, J.lv[l].h); }
    return std::pair<int, int>(w, h);
  };
  check(hipMemcpyAsync(dev + P.off_small, S.stage.ptr, P.small, hipMemcpyHostToDevice, qs), "upload");
  // ---- the device-resident part: pyramid, top-level statistics, coarse
  // search, peaks, tracking, full score; one synchronisation at the end
  if (ncchip_fill_u32(reinterpret_cast<uint32_t*>(counts), n_count_words, 0u, qs) ||
      ncchip_fill_u32(reinterpret_cast<uint32_t*>(dev + off_bits), n_bit_words, 0u, qs) ||
      ncchip_fill_u32(reinterpret_cast<uint32_t*>(dev + off_table), std::max<size_t>(n_track, 1) * n_frames * table_size,
                      0xffffffffu, qs))
    raise(kInternalError, "device", "fill launch failed");
  if (reduced && ncchip_domain_sat(dframes, F, max_dims(0).second, max_dims(0).first, qs))
    raise(kInternalError, "device", "domain launch failed");
  time_begin(kStagePyramid, qs);
  // levels 1-3 of byte frames in one pass, the others level by level
  size_t fused = 0;
  if (!wide && level_count > 1) {
    fused = std::min<size_t>(3, level_count - 1);
    const auto [w0, h0] = max_dims(0);
    if (ncchip_pyramid3_u8(dframes, F, int(fused), w0, h0, qs)) raise(kInternalError, "device", "pyramid launch failed");
  }
  for (size_t l = 1 + fused; l < level_count; ++l) {
    const auto [w, h] = max_dims(l);
    if (ncchip_pyramid_batch(dframes, F, int(l), jobs[0].lv[l - 1].kind, jobs[0].lv[l].kind, w, h, qs))
      raise(kInternalError, "device", "pyramid launch failed");
  }
  time_end(qs);
  // the top level's row prefixes feed the FP32 and FP64 coarse kernels only:
  // the split-BF16 kernel takes its exact window moments from the image tile
  // it stages
  time_begin(kStagePrefix, qs);
  if (!P.use_mfma && ncchip_row_prefix_batch(dframes, F, int(top), jobs[0].lv[top].kind, Ht_max, qs))
    raise(kInternalError, "device", "prefix launch failed");
  time_end(qs);
}

// The split-BF16 coarse search of a planned batch as one ncchip_coarse_mfma
// job (the arguments the back half passes to it).
NcchipCoarseJob FindContext::coarse_job(Slot& S) {
  Plan& P = *S.plan;
  const NccModel& model = *P.model;
  const FindOptions& o = P.o;
  const DenseWeights& dense = P.dense;
  uint8_t* const dev = P.dev;
  NcchipCoarseJob j{};
  j.frames = P.dframes; j.n_frames = P.F; j.level = int(P.top); j.level_kind = P.jobs[0].lv[P.top].kind;
  j.max_w = P.Wt_max; j.max_h = P.Ht_max;
  P.frame_dims.clear();
  for (const Job& J : P.jobs) { P.frame_dims.push_back(J.lv[P.top].w); P.frame_dims.push_back(J.lv[P.top].h); }
  j.frame_dims = P.frame_dims.data();
  j.C = std::ldexp(P.wide ? 32768.0 : 128.0, int(2 * P.top));
  j.pools[0] = P.pools[0]; j.pools[1] = P.pools[1];
  j.patch_ids = reinterpret_cast<const int32_t*>(dev + P.off_refs[P.top]);
  j.n_angles = int(P.A);
  NcchipDenseWeights& dw = j.weights;
  dw.wh = dev + P.off_wh; dw.wl = dev + P.off_wl; dw.wsum_eff = reinterpret_cast<const double*>(dev + P.off_wsum);
  dw.step0 = reinterpret_cast<const int32_t*>(dev + P.off_step0);
  dw.segs = reinterpret_cast<const int32_t*>(dev + P.off_segs);
  dw.ymin = dense.ymin; dw.xmin = dense.xmin; dw.yspan = dense.yspan; dw.xchunks = dense.xchunks;
  dw.max_run = dense.max_run;
  dw.common_runs = reinterpret_cast<const int32_t*>(dev + P.off_common);
  dw.n_common_runs = int32_t(dense.common_runs.size() / 4);
  dw.n_common = dense.n_common;
  const int32_t* groups = reinterpret_cast<const int32_t*>(dev + P.off_groups);
  dw.group_first = groups;
  dw.group_n = groups + dense.n_off;
  dw.group_runs = groups + dense.runs_off;
  for (int i = 0; i < 3; ++i) { dw.first_base[i] = dense.first_base[i]; dw.n_base[i] = dense.n_base[i]; }
  dw.prune_below = 0.9 * (o.min_score * .4);
  dw.peak_thr = o.min_score * .4;  // sparse planes: the peaks kernel reads the flagged scores only
  dw.w8_scale = dense.w8_scale;
  j.exact_lattice = !P.wide && P.top <= 2;  // exact byte lattice (matcher.cpp:347, 561)
  j.absolute = model.params.absolute;
  j.screen = o.form != NCCHIP_FORM_FP8_SCREEN ? kMissing : o.screen > kMissing ? o.screen : kDefaultScreen;
  j.tile = P.tile;
  return j;
}

namespace {
constexpr NcchipCoarseTile kDefaultCoarseTile = {4, 4, 1, 1, 0};

// Smallest power of two >= v (v >= 1).
int pow2_ceil(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}
}  // namespace

// The coarse tile shape of a planned split-BF16 search: forced by the
// options, the default, or the model's tuned shape for this kind of search
// (level, pixel kind, coarse angles, union box, screen, frame count and
// top-level size classes; a search run in the shared persistent launch is
// keyed without the frame count and size, which set only how many items it
// adds to the queue); a search whose kind has no tuned shape yet is
// tuned on this batch (tune_pending) unless its top level is level 0 (the
// caller's own image, which a later tuning pass could not rely on).
void FindContext::choose_tile(Slot& S) {
  Plan& P = *S.plan;
  const FindOptions& o = P.o;
  const bool screened = o.form == NCCHIP_FORM_FP8_SCREEN;
  const int kind = P.jobs[0].lv[P.top].kind;
  P.tile = kDefaultCoarseTile;
  P.tune_pending = false;
  if (o.coarse_tile > 0) {
    const NcchipCoarseTile t = unpack_tile(o.coarse_tile);
    if (!ncchip_coarse_mfma_fits(int(P.top), kind, P.dense.yspan, P.dense.xchunks, screened, &t))
      raise(kInvalidArgument, "coarse_tile", "tile shape is invalid or does not fit this search");
    P.tile = t;
    return;
  }
  if (o.coarse_tile < 0) return;
  char key[160];
  if (queued_)
    std::snprintf(key, sizeof key, "L%zu K%d A%zu Y%d X%d S%d queue", P.top, kind, P.A, P.dense.yspan,
                  P.dense.xchunks, int(screened));
  else
    std::snprintf(key, sizeof key, "L%zu K%d A%zu Y%d X%d S%d F%d W%d H%d", P.top, kind, P.A, P.dense.yspan,
                  P.dense.xchunks, int(screened), pow2_ceil(std::min(P.F, 256)), pow2_ceil(std::max(P.Wt_max, 1)),
                  pow2_ceil(std::max(P.Ht_max, 1)));
  P.tune_key = key;
  const NccModel& m = *P.model;
  std::lock_guard<std::mutex> lock(m.tune_mutex);
  const auto it = m.coarse_tuning.find(P.tune_key);
  if (it != m.coarse_tuning.end()) {
    const NcchipCoarseTile t{it->second.rows, it->second.col_waves, it->second.row_waves, it->second.tiles,
                             it->second.bound};
    if (ncchip_coarse_mfma_fits(int(P.top), kind, P.dense.yspan, P.dense.xchunks, screened, &t)) P.tile = t;
    return;
  }
  P.tune_pending = P.top > 0;
}

// The autotuner: runs this batch's coarse search with candidate tile shapes
// on the batch's own top level, after the front half's launches, and keeps
// the fastest: the bound groups (32, 16 or 8 angles) on the default shape,
// then every shape (rows per wave 4 or 8; waves across x down 4x1, 2x2, 1x4,
// 2x1, 1x2; 1, 2 or 4 angle tiles per staged image tile) with the best
// bound, then the other bounds on the best shape.  Each shape is
// timed over one launch with HIP events on the batch's stream, after one
// untimed launch for the first shape of each kernel instantiation (its code
// load).  Every shape writes the same sparse planes and peak bits (the
// accumulation order per output does not depend on the shape, and a bound
// only drops windows below the peak threshold), so the repeated launches
// leave the batch's state as one launch would.  A search whose
// default-shape launch is estimated (from a probe on its first frame's
// first rows) to take over 1 ms is timed on its first frames and a band of
// their top-level rows (at least 2048 default-shape blocks), so tuning costs
// about a millisecond per shape.
// Not run inside a graph capture.
void FindContext::tune_coarse(Slot& S) {
  Plan& P = *S.plan;
  hipStream_t qs = P.qs;
  NcchipCoarseJob j = coarse_job(S);
  const bool screened = j.screen > kMissing;
  const int n_tiles = (j.n_angles + 31) / 32;
  if (!tune_events_[0]) {
    check(hipEventCreate(&tune_events_[0]), "hipEventCreate");
    check(hipEventCreate(&tune_events_[1]), "hipEventCreate");
  }
  auto run = [&](const NcchipCoarseTile& t, int reps) {
    check(hipEventRecord(tune_events_[0], qs), "hipEventRecord");
    for (int r = 0; r < reps; ++r)
      if (ncchip_coarse_mfma(j.frames, j.n_frames, j.level, j.level_kind, j.max_w, j.max_h, j.C, j.pools, j.patch_ids,
                             j.n_angles, &j.weights, j.exact_lattice, j.absolute, j.screen, &t, qs))
        raise(kInternalError, "device", "coarse launch failed");
    check(hipEventRecord(tune_events_[1], qs), "hipEventRecord");
    check(hipEventSynchronize(tune_events_[1]), "hipEventSynchronize");
    float ms = 0;
    check(hipEventElapsedTime(&ms, tune_events_[0], tune_events_[1]), "hipEventElapsedTime");
    return double(ms) * 1e3 / reps;
  };
  // The sample: the default shape is timed on a probe (the first frame's
  // first rows, at least kSampleBlocks default-shape blocks) after one
  // warm-up launch there, and its time extrapolated to the batch; a batch
  // estimated over kSampleUs is timed on its first frames, then a band of
  // their top-level rows, keeping at least kSampleBlocks default blocks (so
  // that the sample fills the GPU as the batch does) and about kSampleUs per
  // launch.  No launch of the whole batch is needed to choose it.
  constexpr double kSampleUs = 1000.0;
  constexpr int64_t kSampleBlocks = 2048;
  auto blocks = [&](int f, int rows) {  // default-shape blocks (one per angle tile) of frame f within its first rows
    const int w = P.jobs[size_t(f)].lv[P.top].w, h = std::min(rows, P.jobs[size_t(f)].lv[P.top].h);
    return int64_t((w + 127) / 128) * ((h + 3) / 4) * n_tiles;
  };
  int64_t all = 0;
  for (int f = 0; f < j.n_frames; ++f) all += blocks(f, j.max_h);
  const NcchipCoarseJob whole = j;
  int probe_rows = 16;
  while (probe_rows < j.max_h && blocks(0, probe_rows) < kSampleBlocks) probe_rows *= 2;
  probe_rows = std::min(probe_rows, j.max_h);
  j.n_frames = 1;
  j.max_h = probe_rows;
  run(kDefaultCoarseTile, 1);
  const double probe_us = run(kDefaultCoarseTile, 1);
  const double est = probe_us * double(all) / double(std::max<int64_t>(blocks(0, probe_rows), 1));
  j = whole;
  if (est > kSampleUs) {
    int frames = 0;
    int64_t kept = 0;
    while (frames < j.n_frames && (kept < kSampleBlocks || est * double(kept) / double(all) < kSampleUs / 2)) {
      kept += blocks(frames, j.max_h);
      ++frames;
    }
    double part = est * double(kept) / double(std::max<int64_t>(all, 1));
    int rows = j.max_h;
    while (rows > 16 && part > kSampleUs) {
      const int r2 = std::max(16, rows / 2 / 16 * 16);
      int64_t k2 = 0;
      for (int f = 0; f < frames; ++f) k2 += blocks(f, r2);
      if (k2 < kSampleBlocks) break;
      part *= double(k2) / double(std::max<int64_t>(kept, 1));
      kept = k2;
      rows = r2;
    }
    j.n_frames = frames;
    j.max_h = rows;
  }
  const double def_us = run(kDefaultCoarseTile, 1);
  CoarseTune best;
  best.us = def_us;
  best.default_us = def_us;
  NcchipCoarseTile best_t = kDefaultCoarseTile;
  auto same = [](const NcchipCoarseTile& a, const NcchipCoarseTile& b) {
    return a.rows == b.rows && a.col_waves == b.col_waves && a.row_waves == b.row_waves && a.tiles == b.tiles &&
           a.bound == b.bound;
  };
  bool warm[2][2] = {{true, false}, {false, false}};  // [rows == 8][tiles > 1]: instantiation already run
  auto consider = [&](const NcchipCoarseTile& t) {
    if (same(t, best_t) || same(t, kDefaultCoarseTile)) return;
    if (!ncchip_coarse_mfma_fits(j.level, j.level_kind, j.weights.yspan, j.weights.xchunks, screened, &t)) return;
    bool& w = warm[t.rows == 8][std::min(t.tiles, n_tiles) > 1];
    if (!w) {
      run(t, 1);
      w = true;
    }
    const double us = run(t, 1);
    if (us < best.us) {
      best_t = t;
      best.us = us;
    }
  };
  static const int kBounds[] = {32, 16, 8};
  static const int kShapes[][2] = {{4, 1}, {2, 2}, {1, 4}, {2, 1}, {1, 2}};
  // the bound group on the default shape, then the shapes with the best
  // bound, then the other bounds on the best shape
  if (!screened)
    for (const int b : kBounds) consider(NcchipCoarseTile{4, 4, 1, 1, b});
  const int bound = best_t.bound;
  for (const int rows : {4, 8})
    for (const auto& sh : kShapes)
      for (const int tiles : {1, 2, 4}) {
        if (tiles > 1 && tiles / 2 >= n_tiles) continue;  // the same shape as a smaller tile count
        consider(NcchipCoarseTile{rows, sh[0], sh[1], tiles, bound});
      }
  if (!screened && (best_t.rows != 4 || best_t.col_waves != 4 || best_t.row_waves != 1 || best_t.tiles != 1))
    for (const int b : {0, 32, 16, 8})
      if (b != bound) consider(NcchipCoarseTile{best_t.rows, best_t.col_waves, best_t.row_waves, best_t.tiles, b});
  best.rows = best_t.rows; best.col_waves = best_t.col_waves; best.row_waves = best_t.row_waves;
  best.tiles = best_t.tiles; best.bound = best_t.bound;
  P.tile = best_t;
  P.tune_pending = false;
  const NccModel& m = *P.model;
  std::lock_guard<std::mutex> lock(m.tune_mutex);
  m.coarse_tuning[P.tune_key] = best;
}

// The back half: coarse search (unless the caller ran it in a shared
// persistent launch), peaks, tracking, full score and the read-back copy.
void FindContext::enqueue_back(Slot& S) {
  Plan& P = *S.plan;
  cur_ = &S;
  hipStream_t qs = P.qs;
  const NccModel& model = *P.model;
  const std::shared_ptr<DeviceModel> dmodel = P.dmodel;
  const FindOptions o = P.o;
  FindTrace* const traces = P.traces;
  const size_t n_frames = P.n_frames, top = P.top, finest = P.finest, stride = P.stride, n_track = P.n_track;
  const size_t A = P.A, cap = P.cap, cap_all = P.cap_all, table_size = P.table_size, n_count_words = P.n_count_words;
  const std::vector<size_t>& track_levels = P.track_levels;
  const std::vector<std::vector<int32_t>>& refs = P.refs;
  RasterPool& qpool = P.qpool;
  const int64_t coarse_original = P.coarse_original;
  std::vector<Job>& jobs = P.jobs;
  const bool use_mfma = P.use_mfma, track_on_mfma = P.track_on_mfma, wide = P.wide, reduced = P.reduced;
  const int max_bh = P.max_bh, max_bw = P.max_bw, F = P.F, Wt_max = P.Wt_max, Ht_max = P.Ht_max;
  const double scale = P.scale, phase = P.phase, radius = P.radius, t0 = P.t0;
  uint8_t* const dev = P.dev;
  NccFrame* const dframes = P.dframes;
  NccMatch* const lists = P.lists;
  int32_t* const counts = P.counts;
  unsigned long long* const extent = P.extent;
  unsigned long long* const work = P.work;
  const NccPatchPool pools[2] = {P.pools[0], P.pools[1]};
  const size_t off_key = P.off_key, off_raw = P.off_raw, off_table = P.off_table;
  const std::vector<size_t>& off_refs = P.off_refs;
  uint8_t* const rb = P.rb;
  const size_t rb_full = P.rb_full, rb_counts = P.rb_counts;
  auto raster_of = [&](int32_t r) -> const Raster& {
    return (r >> 24) ? qpool.rasters[size_t(r & 0xffffff)] : model.pool.rasters[size_t(r & 0xffffff)];
  };
  // centring of the coarse forms: mid-range of the pixel type at the top level
  const double C = std::ldexp(wide ? 32768.0 : 128.0, int(2 * top));
  const int exact_top = !wide && top <= 2;  // exact byte lattice (matcher.cpp:347, 561)
  const auto* coarse_refs = reinterpret_cast<const int32_t*>(dev + off_refs[top]);
  time_begin(kStageCoarse, qs);
  int rc = 0;
  if (dry_) {
  } else if (use_mfma && !P.coarse_deferred) {
    const NcchipCoarseJob j = coarse_job(S);
    rc = ncchip_coarse_mfma(j.frames, j.n_frames, j.level, j.level_kind, j.max_w, j.max_h, j.C, j.pools, j.patch_ids,
                            j.n_angles, &j.weights, j.exact_lattice, j.absolute, j.screen, &j.tile, qs);
  } else if (!use_mfma) {
    rc = ncchip_coarse(dframes, F, int(top), jobs[0].lv[top].kind, Wt_max, Ht_max, C, pools, coarse_refs, int(A),
                       max_bh, max_bw, o.form, exact_top, model.params.absolute, qs);
  }
  if (rc == -6) raise(kUnsupported, "form", "correlation form not available in this build");
  if (rc) raise(kInternalError, "device", "coarse launch failed");
  time_end(qs);
  time_begin(kStagePeaks, qs);
  // reduced domain: anchors whose pose may not reach it score missing (matcher.cpp:556-578)
  if (reduced && !dry_ && ncchip_domain_planes(dframes, F, int(top), Wt_max, Ht_max, pools, coarse_refs, int(A), scale, phase,
                                      use_mfma ? 2 : 0, qs))
    raise(kInternalError, "device", "domain launch failed");
  if (!dry_ && ncchip_peaks(dframes, F, int(top), Wt_max, Ht_max, pools, coarse_refs, int(A), o.min_score * .4, scale, phase,
                   lists, reinterpret_cast<int32_t*>(dev + off_key), int(cap), counts, extent, use_mfma ? 2 : 0,
                   qs))
    raise(kInternalError, "device", "peak launch failed");
  // find_models over several models converts the whole level 0 (matcher.cpp:596-611)
  // u16 frames keep the whole level 0 (matcher.cpp:612: only byte images are cut down)
  if (!finest && !o.whole_level0 && !wide && !dry_ && ncchip_storage0(dframes, F, counts, extent, radius + 4 * scale + 8, qs))
    raise(kInternalError, "device", "storage launch failed");
  time_end(qs);
  // ---- coarse-to-fine tracking (matcher.cpp:630-761), all frames per launch
  constexpr int kListGrid = 2048;
  S.track_mfma.assign(n_track, false);
  if (!dry_) P.track_rc.assign(n_track, -6);
  for (size_t k = 0; k < n_track; ++k) {
    const size_t l = track_levels[k];
    const double step = std::ldexp(1., int(l - 1)), offset = (step - 1) * .5;
    const double tol = model.params.step * (l == top ? double(stride) : 1.) * 1.01;
    // bound of a candidate's neighbourhood (the union of the 5x5 windows of
    // every query patch) and of the FP64-exact moment sums
    double lo_r = 1e300, hi_r = -1e300, lo_c = 1e300, hi_c = -1e300, n_max = 0;
    for (const int32_t r : refs[l - 1]) {
      const Raster& p = raster_of(r);
      lo_r = std::min(lo_r, p.rmin - p.orow); hi_r = std::max(hi_r, p.rmax - p.orow);
      lo_c = std::min(lo_c, p.cmin - p.ocol); hi_c = std::max(hi_c, p.cmax - p.ocol);
      n_max = std::max(n_max, double(p.count));
    }
    const double tile_px = (std::ceil(hi_r - lo_r) + 7) * (std::ceil(hi_c - lo_c) + 7);
    // the matrix-core kernel's tile: a candidate's eligible patches lie within
    // tol of its angle, hence within 2 tol of each other; bound the union of
    // every such group (candidates beyond the tile take the kernel's exact
    // global-memory path)
    int mt_h = 0, mt_w = 0;
    for (const int32_t ri : refs[l - 1]) {
      const Raster& pi = raster_of(ri);
      double gr0 = 1e300, gr1 = -1e300, gc0 = 1e300, gc1 = -1e300;
      for (const int32_t rj : refs[l - 1]) {
        const Raster& pj = raster_of(rj);
        if (std::abs(pj.angle - pi.angle) > 2 * tol) continue;
        gr0 = std::min(gr0, pj.rmin - pj.orow); gr1 = std::max(gr1, pj.rmax - pj.orow);
        gc0 = std::min(gc0, pj.cmin - pj.ocol); gc1 = std::max(gc1, pj.cmax - pj.ocol);
      }
      mt_h = std::max(mt_h, int(std::ceil(gr1 - gr0)) + 7);
      mt_w = std::max(mt_w, int(std::ceil(gc1 - gc0)) + 7);
    }
    const double xmax = (wide ? 65535. : 255.) * std::ldexp(1., 2 * int(l - 1));
    const int exact_f64 = n_max * xmax * xmax < 0x1p53;
    // the most query patches within any candidate's angle window: within tol
    // of the candidate's angle, so within 2 tol of the smallest of them
    // (widened by a relative 1e-9 against rounding of the eligibility test)
    int max_eligible = 0;
    {
      std::vector<double> ang;
      for (const int32_t r : refs[l - 1]) ang.push_back(raster_of(r).angle);
      std::sort(ang.begin(), ang.end());
      for (size_t i = 0, j = 0; i < ang.size(); ++i) {
        while (j < ang.size() && ang[j] <= ang[i] + 2 * tol * (1 + 1e-9)) ++j;
        max_eligible = std::max(max_eligible, int(j - i));
      }
    }
    // matrix cores where every query patch has a dense raster and at most 32
    // patches fall in any candidate's angle window
    bool mfma = track_on_mfma && l - 1 >= 1 && l - 1 <= 4 && jobs[0].lv[l - 1].kind == NCC_U16 && max_eligible <= 32;
    if (mfma)
      for (const int32_t r : refs[l - 1]) mfma &= dense_geom(raster_of(r)).elems > 0;
    int rc_track = -6;
    if (mfma) {
      time_begin(kStageTrackMfma, qs);
      rc_track = dry_ ? P.track_rc[k] : ncchip_track_mfma(dframes, F, int(l - 1), jobs[0].lv[l - 1].kind, pools,
                                   reinterpret_cast<const int32_t*>(dev + off_refs[l - 1]), int(refs[l - 1].size()),
                                   lists + k * cap_all, counts + k * n_frames, int(cap),
                                   reinterpret_cast<NccMatch*>(dev + off_raw), lists + (k + 1) * cap_all,
                                   counts + (k + 1) * n_frames,
                                   reinterpret_cast<int32_t*>(dev + off_table) + k * n_frames * table_size,
                                   int(table_size), o.min_score * .4, tol, step, offset, !wide && l - 1 <= 2,
                                   model.params.absolute, mt_h, mt_w, xmax, work + k, kListGrid, qs);
      time_end(qs);
      S.track_mfma[k] = rc_track == 0;
      P.track_rc[k] = rc_track;
    }
    if (rc_track == -6) time_begin(kStageTrack, qs);
    if (rc_track == -6 && !dry_ &&
        ncchip_track(dframes, F, int(l - 1), jobs[0].lv[l - 1].kind, pools,
                     reinterpret_cast<const int32_t*>(dev + off_refs[l - 1]), int(refs[l - 1].size()),
                     lists + k * cap_all, counts + k * n_frames, int(cap), reinterpret_cast<NccMatch*>(dev + off_raw),
                     lists + (k + 1) * cap_all, counts + (k + 1) * n_frames,
                     reinterpret_cast<int32_t*>(dev + off_table) + k * n_frames * table_size, int(table_size),
                     o.min_score * .4, tol, step, offset, !wide && l - 1 <= 2, model.params.absolute,
                     tile_px < 1e9 ? int(tile_px) : 0, exact_f64, max_eligible, xmax, work + k, kListGrid, qs))
      raise(kInternalError, "device", "tracking launch failed");
    if (rc_track != -6 && rc_track != 0) raise(kInternalError, "device", "tracking launch failed");
    if (rc_track == -6) time_end(qs);
  }
  // ---- full score at the finest level (matcher.cpp:813-822)
  NcchipRefineParams rp{};
  rp.frames = dframes;
  rp.level_kind = jobs[0].lv[finest].kind;
  rp.scale = std::ldexp(1., int(finest)); rp.phase = (rp.scale - 1) * .5;
  rp.finest = int32_t(finest); rp.top = int32_t(top);
  rp.orig_ref = finest ? ref(1, uint32_t(coarse_original)) : ref(0, model.original);
  rp.level_orow = model.levels[finest].orow; rp.level_ocol = model.levels[finest].ocol;
  rp.refs0 = reinterpret_cast<const int32_t*>(dev + off_refs[0]);
  rp.n_refs0 = int32_t(refs[0].size());
  rp.step = model.params.step; rp.angle_start = o.start; rp.angle_end = o.start + o.extent;
  rp.min_score = o.min_score; rp.subpixel = o.subpixel; rp.rotate = o.extent > 0; rp.absolute = model.params.absolute;
  rp.recover_first = o.query_cache;
  rp.prefix_recovery = o.whole_level0 && !wide;  // byte images only (matcher.cpp:596-611, 991)
  rp.guard_first = -1;  // set by find_models once the candidate count is known
  rp.guard = -1;
  auto* hfull = reinterpret_cast<NccMatch*>(rb + rb_full);
  auto* hcounts = reinterpret_cast<int32_t*>(rb + rb_counts);
  // reduced domain: the level-1 hand-over to the fit snaps the candidates to
  // the domain (matcher.cpp:631-645); duplicates it makes are merged after the
  // full score, which only depends on the pose
  if (reduced && o.subpixel && finest == 0 && top >= 1 &&
      !dry_ && ncchip_domain_snap(dframes, F, lists + n_track * cap_all, counts + n_track * n_frames, int(cap), qs))
    raise(kInternalError, "device", "domain launch failed");
  time_begin(kStageFull, qs);
  if (!dry_ && ncchip_full_score(&rp, pools, F, lists + n_track * cap_all, counts + n_track * n_frames, int(cap), hfull, 1024,
                        qs))
    raise(kInternalError, "device", "full score launch failed");
  time_end(qs);
  if (!dry_ && ncchip_copy_u32(reinterpret_cast<const uint32_t*>(counts), reinterpret_cast<uint32_t*>(hcounts), n_count_words,
                      qs))
    raise(kInternalError, "device", "copy launch failed");
  S.model = &model; S.dmodel = dmodel; S.o = o; S.traces = traces; S.n_frames = n_frames; S.empty = false;
  S.jobs = std::move(jobs); S.qpool = std::move(qpool);
  S.cap = cap; S.cap_all = cap_all; S.n_track = n_track; S.top = top; S.finest = finest; S.stride = stride;
  S.off_key = off_key; S.dev = dev; S.dframes = dframes; S.lists = lists; S.hfull = hfull; S.hcounts = hcounts;
  S.rb = rb; S.pools[0] = pools[0]; S.pools[1] = pools[1]; S.rp = rp; S.t0 = t0;
  S.track_rc = P.track_rc;
  S.coarse_tile = use_mfma ? pack_tile(P.tile) : 0;
  S.plan.reset();
}

// Waits for a slot's device-resident part and turns its full-scored lists into
// each frame's candidate order (matcher.cpp:824-827); fills the stage traces.
void FindContext::collect(Slot& S) {
  cur_ = &S;
  const FindOptions& o = S.o;
  FindTrace* const traces = S.traces;
  const size_t n_frames = S.n_frames;
  std::vector<Job>& jobs = S.jobs;
  const size_t cap = S.cap, cap_all = S.cap_all, n_track = S.n_track;
  int32_t* const hcounts = S.hcounts;
  check(hipEventSynchronize(S.done), "find");
  S.t1 = now_us();
  for (size_t f = 0; f < n_frames; ++f)
    if (size_t(hcounts[f]) > o.max_candidates) raise(kLimitExceeded, "max_candidates", "coarse candidate budget exceeded");
  auto list_count = [&](size_t k, size_t f) { return std::min<size_t>(size_t(std::max(hcounts[k * n_frames + f], 0)), cap); };

  // tracking work of the batch, reported with the first frame's trace
  if (traces) {
    const auto* hwork = reinterpret_cast<const unsigned long long*>(
        reinterpret_cast<const uint8_t*>(hcounts) + align_up((n_track + 1) * n_frames * 4, 8));
    for (size_t k = 0; k < n_track; ++k) {
      jobs[0].track_flops += 2.0 * double(hwork[k]);
      if (S.track_mfma[k]) jobs[0].track_mfma_flops += 2.0 * double(hwork[k]);
    }
  }
  // ---- stage dumps for traces (not on the benchmarked path)
  if (traces && o.trace_stages) {
    std::vector<NccMatch> all((n_track + 1) * cap_all);
    std::vector<int32_t> keys(cap_all);
    std::vector<NccFrame> frs(n_frames);
    check(hipMemcpy(all.data(), S.lists, all.size() * sizeof(NccMatch), hipMemcpyDeviceToHost), "trace");
    check(hipMemcpy(keys.data(), S.dev + S.off_key, keys.size() * 4, hipMemcpyDeviceToHost), "trace");
    check(hipMemcpy(frs.data(), S.dframes, n_frames * sizeof(NccFrame), hipMemcpyDeviceToHost), "trace");
    for (size_t f = 0; f < n_frames; ++f) {
      Job& J = jobs[f];
      if (!J.trace) continue;
      std::copy(frs[f].storage0, frs[f].storage0 + 4, J.storage0);
      // the reference emits peaks in (angle, row, col) order (matcher.cpp:580-594)
      std::vector<size_t> order(list_count(0, f));
      for (size_t i = 0; i < order.size(); ++i) order[i] = i;
      const NccMatch* c0 = &all[f * cap];
      const int32_t* k0 = &keys[f * cap];
      std::sort(order.begin(), order.end(), [&](size_t a, size_t b) {
        if (k0[a] != k0[b]) return k0[a] < k0[b];
        if (c0[a].row != c0[b].row) return c0[a].row < c0[b].row;
        return c0[a].col < c0[b].col;
      });
      for (size_t i : order) J.trace->coarse.push_back({c0[i].row, c0[i].col, c0[i].angle, c0[i].score});
      for (size_t k = 1; k <= n_track; ++k) {
        std::vector<Pose> v;
        for (size_t i = 0; i < list_count(k, f); ++i) {
          const NccMatch& m = all[k * cap_all + f * cap + i];
          v.push_back({m.row, m.col, m.angle, m.score});
        }
        std::sort(v.begin(), v.end(), pose_less);  // merge_identical's order
        J.trace->tracked.push_back(std::move(v));
      }
    }
  }
  S.t2 = now_us();
  // ---- candidate order (matcher.cpp:824-827)
  for (size_t f = 0; f < n_frames; ++f) {
    Job& J = jobs[f];
    const size_t n = list_count(n_track, f);
    J.full.reserve(n);
    for (size_t i = 0; i < n; ++i) {
      const NccMatch& m = S.hfull[f * cap + i];
      J.full.push_back({m.row, m.col, m.angle, m.score});
    }
    std::sort(J.full.begin(), J.full.end(), score_order);
    if (J.in.domain) {
      // merge_identical after the snap (matcher.cpp:644): equal poses have equal
      // full scores, so they are adjacent here
      J.full.erase(std::unique(J.full.begin(), J.full.end(), [](const Pose& a, const Pose& b) {
        return a.row == b.row && a.col == b.col && a.angle == b.angle;
      }), J.full.end());
    }
    J.refined.assign(J.full.size(), Pose{});
    J.verdict.assign(J.full.size(), -1);  // -1: not refined
    J.batch_size = o.num_matches ? std::max<size_t>(2 * o.num_matches, 4) : J.full.size();
  }
  size_t total = 0;
  for (const Job& J : jobs) total += J.full.size();
  if (!o.num_matches && total > kOneRound)
    for (Job& J : jobs) J.batch_size = std::min(J.batch_size, kUnboundedFirstRound);
}

// One refinement round: every request is refined by its slot's query in one
// launch per slot (on the refinement stream, one synchronisation for all).
// A candidate's refinement does not depend on the other candidates, so the
// greedy walks ask for the candidates they may need next and take the
// results in their own order.  Candidate lists and results live in the
// slots' host-visible staging buffers, free once the slot was collected.
void FindContext::refine_round(std::vector<RefineRequest>& reqs) {
  if (reqs.empty()) return;
  // the background guard of find_models is per frame: a frame with the guard
  // on and one without are refined by separate launches of the same slot
  auto guard_of = [](const RefineRequest& r) { return r.slot->jobs[r.frame].guard_first >= 0; };
  std::vector<size_t> order(reqs.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
    if (reqs[a].slot != reqs[b].slot) return reqs[a].slot < reqs[b].slot;
    return guard_of(reqs[a]) < guard_of(reqs[b]);
  });
  struct Launch { Slot* slot; size_t begin, end, offset; };
  std::vector<Launch> launches;
  for (size_t i = 0; i < order.size();) {
    size_t j = i;
    while (j < order.size() && reqs[order[j]].slot == reqs[order[i]].slot &&
           guard_of(reqs[order[j]]) == guard_of(reqs[order[i]]))
      ++j;
    // the launches of one slot use consecutive parts of its staging areas
    size_t offset = 0;
    if (!launches.empty() && launches.back().slot == reqs[order[i]].slot)
      offset = launches.back().offset + (launches.back().end - launches.back().begin);
    launches.push_back({reqs[order[i]].slot, i, j, offset});
    i = j;
  }
  std::vector<size_t> slot_total(launches.size());
  for (size_t k = 0; k < launches.size(); ++k) {
    size_t n = 0;
    for (const Launch& L : launches) if (L.slot == launches[k].slot) n += L.end - L.begin;
    slot_total[k] = n;
  }
  for (size_t li = 0; li < launches.size(); ++li) {
    const Launch& L = launches[li];
    Slot& S = *L.slot;
    cur_ = &S;
    const size_t n = L.end - L.begin, total = slot_total[li];
    auto* dcand = static_cast<NccMatch*>(S.stage.ptr) + L.offset;
    auto* dcand_frame = reinterpret_cast<int32_t*>(static_cast<NccMatch*>(S.stage.ptr) + total) + L.offset;
    auto* dout = reinterpret_cast<NccMatch*>(S.rb) + L.offset;
    auto* dverdict = reinterpret_cast<int32_t*>(reinterpret_cast<NccMatch*>(S.rb) + total) + L.offset;
    for (size_t k = 0; k < n; ++k) {
      const RefineRequest& r = reqs[order[L.begin + k]];
      dcand[k] = {r.in.row, r.in.col, r.in.angle, r.in.score};
      dcand_frame[k] = int32_t(r.frame);
    }
    NcchipRefineParams rp = S.rp;
    const Job& J = S.jobs[reqs[order[L.begin]].frame];
    rp.guard_first = J.guard_first;
    rp.guard = J.guard;
    time_begin(kStageRefine, stream2_);
    if (ncchip_refine(&rp, S.pools, dcand, dcand_frame, int(n), dout, dverdict, stream2_))
      raise(kInternalError, "device", "refine launch failed");
    time_end(stream2_);
  }
  check(hipStreamSynchronize(stream2_), "refine");
  for (size_t li = 0; li < launches.size(); ++li) {
    const Launch& L = launches[li];
    const size_t n = L.end - L.begin, total = slot_total[li];
    const auto* dout = reinterpret_cast<const NccMatch*>(L.slot->rb) + L.offset;
    const auto* dverdict = reinterpret_cast<const int32_t*>(reinterpret_cast<const NccMatch*>(L.slot->rb) + total) + L.offset;
    for (size_t k = 0; k < n; ++k) {
      RefineRequest& r = reqs[order[L.begin + k]];
      const NccMatch& m = dout[k];
      r.out = {m.row, m.col, m.angle, m.score};
      r.verdict = dverdict[k];
    }
  }
}

std::vector<std::vector<Pose>> FindContext::complete_batch(uint64_t ticket) {
  Slot& S = *slots_[ticket % 2];
  if (!S.busy || S.ticket != ticket) raise(kInvalidState, "batch", "no such pending batch");
  cur_ = &S;
  // On every exit, including an error between refinement launches, the
  // slot's staging buffers are free again only once the refinement stream
  // has drained (a launch queued before the error may still read them).
  struct Release {
    Slot& s;
    hipStream_t refine;
    ~Release() {
      (void)hipStreamSynchronize(refine);
      s.busy = false;
    }
  } release{S, stream2_};
  const NccModel& model = *S.model;
  const FindOptions& o = S.o;
  const size_t n_frames = S.n_frames;
  std::vector<std::vector<Pose>> results(n_frames);
  if (S.empty) return results;
  collect(S);
  std::vector<Job>& jobs = S.jobs;
  const LevelGeom& base = model.levels.front();

  // ---- greedy selection in score order (matcher.cpp:1063-1076); every
  // frame's next requests are refined together in one launch.
  auto suppressed = [&](const Job& J, const Pose& c) {
    for (const auto& p : J.results) if (pose_overlap(c, p, base) > o.max_overlap + 1e-12) return true;
    return false;
  };
  for (;;) {
    std::vector<RefineRequest> reqs;
    std::vector<size_t> index;  // candidate index within its frame
    for (size_t f = 0; f < n_frames; ++f) {
      Job& J = jobs[f];
      const size_t n = J.full.size();
      while (!J.done && J.pos < n) {
        const Pose& c = J.full[J.pos];
        if (!std::isfinite(c.score) || suppressed(J, c)) { ++J.pos; continue; }
        if (J.verdict[J.pos] < 0) break;
        if (J.verdict[J.pos] && !suppressed(J, J.refined[J.pos])) {
          J.results.push_back(J.refined[J.pos]);
          if (o.num_matches && J.results.size() == o.num_matches) { J.done = true; break; }
        }
        ++J.pos;
      }
      if (J.done || J.pos >= n) { J.done = true; continue; }
      size_t taken = 0;
      for (size_t i = J.pos; i < n && taken < J.batch_size; ++i)
        if (std::isfinite(J.full[i].score) && J.verdict[i] < 0 && !suppressed(J, J.full[i])) {
          reqs.push_back({&S, uint32_t(f), J.full[i], {}, 0});
          index.push_back(i);
          ++taken;
        }
      J.batch_size *= 2;
    }
    if (reqs.empty()) break;
    refine_round(reqs);
    for (size_t k = 0; k < reqs.size(); ++k) {
      Job& J = jobs[reqs[k].frame];
      J.refined[index[k]] = reqs[k].out;
      J.verdict[index[k]] = reqs[k].verdict;
    }
  }
  cur_ = &S;
  const double t3 = now_us();
  FindTrace* const traces = S.traces;
  time_collect(traces);  // event timings of the whole batch, into the first trace
  for (size_t f = 0; f < n_frames; ++f) {
    Job& J = jobs[f];
    std::sort(J.results.begin(), J.results.end(), score_order);
    results[f] = J.results;
    if (J.trace) {
      FindTrace& tr = *J.trace;
      tr.top = int(S.top); tr.finest = int(S.finest); tr.angle_stride = int(S.stride);
      std::copy(J.storage0, J.storage0 + 4, tr.storage);
      tr.prepared = J.full; tr.refined = J.refined; tr.verdict = J.verdict;
      tr.coarse_flops = J.coarse_flops;
      tr.track_flops = J.track_flops;
      tr.track_mfma_flops = J.track_mfma_flops;
      tr.coarse_form = o.form;
      tr.coarse_tile = S.coarse_tile;
      tr.t_coarse_us = S.t1 - S.t0; tr.t_track_us = S.t2 - S.t1; tr.t_refine_us = t3 - S.t2; tr.t_total_us = now_us() - S.t0;
      if (f > 0 && traces) std::copy(traces[0].kernel_us, traces[0].kernel_us + kStageCount, tr.kernel_us);
    }
  }
  return results;
}

// ------------------------------------------------------------- find_models
namespace {
struct Box { double x[4], y[4]; };
struct Bounds { double top, bottom, left, right; };

// Template rectangle [-o-1/2, w-o-1/2] at a pose (multi.hpp:54-60, scale 1).
Box pose_box(const LevelGeom& lv, const Pose& p) {
  const double top = -lv.orow - .5, bottom = double(lv.h) - lv.orow - .5;
  const double left = -lv.ocol - .5, right = double(lv.w) - lv.ocol - .5;
  const double ca = std::cos(p.angle), sa = std::sin(p.angle);
  const double xs[4] = {left, right, right, left}, ys[4] = {top, top, bottom, bottom};
  Box b;
  for (int i = 0; i < 4; ++i) {
    b.x[i] = p.col + ca * xs[i] + sa * ys[i];
    b.y[i] = p.row - sa * xs[i] + ca * ys[i];
  }
  return b;
}
Bounds enclosure(const Box& b) {  // multi.hpp:19-23
  Bounds o{b.y[0], b.y[0], b.x[0], b.x[0]};
  for (int i = 0; i < 4; ++i) {
    o.top = std::min(o.top, b.y[i]); o.bottom = std::max(o.bottom, b.y[i]);
    o.left = std::min(o.left, b.x[i]); o.right = std::max(o.right, b.x[i]);
  }
  return o;
}
bool disjoint(const Bounds& a, const Bounds& b) {  // multi.hpp:24-26
  return a.right < b.left || b.right < a.left || a.bottom < b.top || b.bottom < a.top;
}
double polygon_area(const double* x, const double* y, size_t n) {  // multi.hpp:61-63
  double a = 0;
  for (size_t i = 0; i < n; ++i) a += x[i] * y[(i + 1) % n] - x[(i + 1) % n] * y[i];
  return std::abs(a) * .5;
}
}  // namespace

// Intersection over the smaller rectangle, both translated by a's first
// corner (multi.hpp:64-81).
double box_overlap(const Pose& pa, const LevelGeom& ga, const Pose& pb, const LevelGeom& gb) {
  Box a = pose_box(ga, pa), b = pose_box(gb, pb);
  const double ox = a.x[0], oy = a.y[0];
  for (int i = 0; i < 4; ++i) { a.x[i] -= ox; a.y[i] -= oy; b.x[i] -= ox; b.y[i] -= oy; }
  const double den = std::min(polygon_area(a.x, a.y, 4), polygon_area(b.x, b.y, 4));
  if (den <= 1e-12) return 0;
  double px[16], py[16], nx[16], ny[16];
  std::copy(a.x, a.x + 4, px); std::copy(a.y, a.y + 4, py);
  size_t count = 4;
  for (int e = 0; e < 4 && count; ++e) {
    const double ux = b.x[e], uy = b.y[e], vx = b.x[(e + 1) % 4], vy = b.y[(e + 1) % 4];
    auto cross = [&](double x, double y) { return (vx - ux) * (y - uy) - (vy - uy) * (x - ux); };
    size_t m = 0;
    for (size_t i = 0; i < count; ++i) {
      const size_t j = (i + 1) % count;
      const double dp = cross(px[i], py[i]), dq = cross(px[j], py[j]);
      if (dp >= 0) { nx[m] = px[i]; ny[m] = py[i]; ++m; }
      if ((dp >= 0) != (dq >= 0)) {
        const double t = dp / (dp - dq);
        nx[m] = px[i] + t * (px[j] - px[i]); ny[m] = py[i] + t * (py[j] - py[i]); ++m;
      }
    }
    std::copy(nx, nx + 16, px); std::copy(ny, ny + 16, py);
    count = m;
  }
  return std::clamp(polygon_area(px, py, count) / den, 0., 1.);
}

namespace {
bool same_request(const FindOptions& a, const FindOptions& b) {  // matcher.cpp:1083-1086
  return a.start == b.start && a.extent == b.extent && a.min_score == b.min_score && a.max_overlap == b.max_overlap &&
         a.num_matches == b.num_matches && a.max_candidates == b.max_candidates && a.subpixel == b.subpixel &&
         a.highest == b.highest && a.lowest == b.lowest;
}
bool multi_less(const ModelMatch& a, const ModelMatch& b) {  // multi.hpp:87-93
  if (a.pose.score != b.pose.score) return a.pose.score > b.pose.score;
  if (a.model != b.model) return a.model < b.model;
  return pose_less(a.pose, b.pose);
}
}  // namespace

std::vector<ModelMatch> FindContext::find_models(const NccModel* const* models, const FindOptions* opts, size_t n_models,
                                                 const FrameInput& frame, const Selection& sel,
                                                 bool per_model_quota_before_overlap) {
  return std::move(find_models_frames(models, opts, n_models, &frame, 1, sel, per_model_quota_before_overlap).front());
}

namespace {
// One frame's greedy selection (select_refined, multi.hpp:83-119) as a walk
// that stops where it needs refinements it does not have yet, so that the
// walks of all frames of a batch share their refinement rounds.
struct SelectionWalk {
  struct Ref { size_t slot, frame, index; };  // query, frame within the query's batch, candidate
  const NccModel* const* models;
  const FindOptions* opts;
  size_t n_models;
  const std::vector<ModelMatch>* pool;
  const std::vector<Ref>* where;       // per pool entry (with_refine)
  std::vector<size_t> ord;
  Selection s;
  bool with_refine;
  std::vector<ModelMatch> result;
  std::vector<Box> boxes;
  std::vector<Bounds> encl;
  std::vector<size_t> counts;
  size_t batch = 0, pos = 0;
  bool stop = false;

  bool unbounded = false;  // no match quota: the round sizes follow the call's candidate count
  void init() {
    counts.assign(n_models, 0);
    batch = s.num_matches > 0 ? std::max<size_t>(2 * size_t(s.num_matches), 4) : 0;
    if (!batch) {
      size_t total = 0;
      bool bounded = true;
      for (size_t i = 0; i < n_models; ++i) { if (!opts[i].num_matches) bounded = false; total += opts[i].num_matches; }
      unbounded = !bounded;
      batch = bounded ? std::max<size_t>(2 * total, 4) : ord.size();
    }
  }
  const LevelGeom& geom(size_t model) const { return models[model]->levels.front(); }
  bool quota_full(size_t model) const {
    const int64_t limit = s.num_matches >= 0 ? s.num_matches : int64_t(opts[model].num_matches);
    return s.num_matches < 0 && limit && counts[model] >= uint64_t(limit);
  }
  bool suppressed(const ModelMatch& c) const {
    const double threshold = s.max_overlap >= 0 ? s.max_overlap : opts[c.model].max_overlap;
    const Bounds aabb = enclosure(pose_box(geom(c.model), c.pose));
    for (size_t i = 0; i < result.size(); ++i) {
      if (s.max_overlap < 0 && result[i].model != c.model) continue;
      if (disjoint(aabb, encl[i])) continue;
      if (box_overlap(c.pose, geom(c.model), result[i].pose, geom(result[i].model)) > threshold + 1e-12) return true;
    }
    return false;
  }
  bool done() const { return stop || pos >= ord.size(); }
  // Walks while every needed refinement is known (verdict: -1 unknown, -2
  // requested, else the verdict).
  template <class Verdict, class Refined>
  void advance(Verdict verdict, Refined refined) {
    while (pos < ord.size()) {
      const ModelMatch& c = (*pool)[ord[pos]];
      if (quota_full(c.model) || suppressed(c)) { ++pos; continue; }
      ModelMatch r = c;
      bool ok = true;
      if (with_refine) {
        const Ref& w = (*where)[ord[pos]];
        if (verdict(w) < 0) return;
        ok = verdict(w) != 0;
        r.pose = refined(w);
      }
      ++pos;
      if (!ok || suppressed(r)) continue;
      const Box rect = pose_box(geom(r.model), r.pose);
      result.push_back(r); boxes.push_back(rect); encl.push_back(enclosure(rect)); ++counts[r.model];
      if (s.num_matches > 0 && result.size() >= uint64_t(s.num_matches)) { stop = true; return; }
    }
  }
  // The next candidates the walk may need, as far as the current results tell.
  template <class Verdict, class Mark>
  void requests(std::vector<Ref>& targets, Verdict verdict, Mark mark) {
    size_t taken = 0;
    for (size_t i = pos; i < ord.size() && taken < batch; ++i) {
      const ModelMatch& c = (*pool)[ord[i]];
      const Ref& w = (*where)[ord[i]];
      if (verdict(w) != -1 || quota_full(c.model) || suppressed(c)) continue;
      mark(w);  // requested in this round (aliases share the request)
      targets.push_back(w);
      ++taken;
    }
    batch *= 2;
  }
  std::vector<ModelMatch> finish() {
    std::sort(result.begin(), result.end(), multi_less);
    return result;
  }
};
}  // namespace

std::vector<std::vector<ModelMatch>> FindContext::find_models_frames(const NccModel* const* models,
                                                                     const FindOptions* opts, size_t n_models,
                                                                     const FrameInput* frames, size_t n_frames,
                                                                     const Selection& sel,
                                                                     bool per_model_quota_before_overlap,
                                                                     bool coarse_queue, BatchStats* stats) {
  std::vector<FrameRequest> reqs(n_frames);
  for (size_t f = 0; f < n_frames; ++f) {
    FrameRequest& r = reqs[f];
    r.frame = frames[f];
    r.models.assign(models, models + n_models);
    r.options.assign(opts, opts + n_models);
    r.selection = sel;
    r.per_model_quota_before_overlap = per_model_quota_before_overlap;
  }
  return find_requests(reqs.data(), n_frames, coarse_queue, stats);
}

std::vector<std::vector<ModelMatch>> FindContext::find_requests(const FrameRequest* requests, size_t n_requests,
                                                                bool coarse_queue, BatchStats* stats) {
  const double t_call = now_us();
  if (stats) *stats = BatchStats{};
  std::vector<std::vector<ModelMatch>> out(n_requests);
  if (!n_requests) return out;
  // every slot of this call is reused: let earlier work on both streams end
  check(hipStreamSynchronize(stream_), "sync");
  check(hipStreamSynchronize(stream2_), "sync");
  // ---- queries: one per distinct (model, options, several models in the
  // request), each over the frames of the requests that ask for it; within
  // a request, aliases (same model, same options) share a query's
  // candidates (matcher.cpp:1087-1098)
  struct Query {
    const NccModel* model;
    FindOptions o;
    bool several;
    std::vector<size_t> requests;  // frame j of the query's batch is request requests[j]
  };
  std::vector<Query> queries;
  // per request and model position: (query, frame within the query's batch)
  std::vector<std::vector<std::pair<size_t, size_t>>> at(n_requests);
  for (size_t r = 0; r < n_requests; ++r) {
    const FrameRequest& R = requests[r];
    const bool several = R.models.size() > 1;
    at[r].resize(R.models.size());
    for (size_t i = 0; i < R.models.size(); ++i) {
      size_t q = 0;
      while (q < queries.size() && !(queries[q].model == R.models[i] && queries[q].several == several &&
                                     same_request(queries[q].o, R.options[i])))
        ++q;
      if (q == queries.size()) queries.push_back({R.models[i], R.options[i], several, {}});
      std::vector<size_t>& rs = queries[q].requests;
      if (rs.empty() || rs.back() != r) rs.push_back(r);
      at[r][i] = {q, rs.size() - 1};
    }
  }
  for (const Query& q : queries)
    if (q.requests.size() > NCC_MAX_BATCH) raise(kLimitExceeded, "Images", "too many images for one batch");
  while (multi_slots_.size() < queries.size()) {
    multi_slots_.push_back(std::make_unique<Slot>());
    multi_slots_.back()->stream = stream_;
    check(hipEventCreateWithFlags(&multi_slots_.back()->done, hipEventDisableTiming), "hipEventCreate");
  }
  // ---- the device-resident parts of every query, on one stream: all front
  // halves, the split-BF16 coarse searches of all queries in one persistent
  // launch (per kernel instantiation), then all back halves
  std::vector<Slot*> planned;
  for (size_t q = 0; q < queries.size(); ++q) {
    FindOptions o = queries[q].o;
    o.query_cache = true;
    o.whole_level0 = queries[q].several;
    o.profile = 0;
    o.trace_stages = false;
    std::vector<FrameInput> fr;
    for (size_t r : queries[q].requests) fr.push_back(requests[r].frame);
    Slot& S = *multi_slots_[q];
    cur_ = &S;
    queued_ = coarse_queue;
    bool planned_q = false;
    try {
      planned_q = enqueue_front(S, *queries[q].model, fr.data(), fr.size(), o, nullptr);
    } catch (...) {
      queued_ = false;
      throw;
    }
    queued_ = false;
    if (planned_q) {
      front_launches(S);
      planned.push_back(&S);
    }
  }
  for (Slot* S : planned)
    if (S->plan->tune_pending) {
      cur_ = S;
      const double t_tune = now_us();
      tune_coarse(*S);
      if (stats) {
        stats->tune_us += now_us() - t_tune;
        ++stats->tuned;
      }
    }
  if (coarse_queue) {
    std::vector<NcchipCoarseJob> jobs;
    std::vector<Slot*> deferred;
    size_t n_units = 0;
    for (Slot* S : planned)
      if (S->plan->use_mfma) {
        jobs.push_back(coarse_job(*S));
        deferred.push_back(S);
        n_units += S->plan->n_frames;
      }
    if (!jobs.empty()) {
      const size_t bytes = ncchip_coarse_mfma_queue_bytes(int(jobs.size()), int(n_units));
      if (queue_dev_.bytes < bytes) queue_dev_ = DeviceBuffer(align_up(bytes, 1 << 16));
      queue_stage_.reserve(bytes);
      if (stats && !queue_events_[0]) {
        check(hipEventCreate(&queue_events_[0]), "hipEventCreate");
        check(hipEventCreate(&queue_events_[1]), "hipEventCreate");
      }
      if (stats) check(hipEventRecord(queue_events_[0], stream_), "hipEventRecord");
      const int rc = ncchip_coarse_mfma_queue(jobs.data(), int(jobs.size()), queue_dev_.ptr, queue_dev_.bytes,
                                              queue_stage_.ptr, queue_stage_.bytes, stream_);
      if (rc) raise(kInternalError, "device", "coarse queue launch failed");
      if (stats) check(hipEventRecord(queue_events_[1], stream_), "hipEventRecord");
      for (Slot* S : deferred) S->plan->coarse_deferred = true;
      if (stats) stats->queue_queries = int(deferred.size());
    }
  }
  for (Slot* S : planned) {
    enqueue_back(*S);
    check(hipEventRecord(S->done, stream_), "hipEventRecord");
  }
  for (size_t q = 0; q < queries.size(); ++q)
    if (!multi_slots_[q]->empty) collect(*multi_slots_[q]);
  // ---- per request: the background guard of each of its queries' frames
  // (matcher.cpp:839-846, for an AVX2 host rounding to nearest and byte
  // images; the whole level 0 is converted when the query kept candidates,
  // which it did when there are 32 of them), and its candidates in
  // (score, model, row, col, angle) order
  std::vector<std::vector<ModelMatch>> cand(n_requests);
  std::vector<std::vector<SelectionWalk::Ref>> where(n_requests);
  for (size_t r = 0; r < n_requests; ++r) {
    const FrameRequest& R = requests[r];
    const bool several = R.models.size() > 1;
    bool distinct = false;
    for (const NccModel* m : R.models) distinct |= m != R.models[0];
    std::vector<bool> seen(R.models.size(), false);
    for (size_t i = 0; i < R.models.size(); ++i) {
      const auto [q, j] = at[r][i];
      Slot& S = *multi_slots_[q];
      if (S.empty) continue;
      bool owner = true;  // the first model position of this request on query q
      for (size_t k = 0; k < i; ++k) owner &= at[r][k].first != q;
      if (owner) {
        const NccModel& m = *R.models[i];
        const FindOptions& o = S.o;  // the angle range clipped to the model's
        Job& J = S.jobs[j];
        if (!S.finest && distinct && R.selection.num_matches == 0 && J.full.size() >= 32 && o.subpixel &&
            !o.num_matches && o.min_score >= .75 && o.min_score <= .85 && S.top + 1 == 3 && !m.params.absolute &&
            m.params.step <= .02 && o.extent <= .4 && m.pool.rasters[m.original].count <= 32768 && several &&
            R.frame.pixel_bytes == 1 && !R.frame.domain) {
          J.guard_first = o.min_score >= .8 ? .1 : .2;
          J.guard = .1;
        }
      }
      const std::vector<Pose>& full = S.jobs[j].full;
      for (size_t k = 0; k < full.size(); ++k)
        if (std::isfinite(full[k].score)) { cand[r].push_back({full[k], i}); where[r].push_back({q, j, k}); }
    }
  }
  // refined pose and verdict per (query, frame, candidate): aliases share them
  std::vector<std::vector<std::vector<Pose>>> refined(queries.size());
  std::vector<std::vector<std::vector<int>>> verdict(queries.size());
  for (size_t q = 0; q < queries.size(); ++q) {
    const Slot& S = *multi_slots_[q];
    refined[q].resize(queries[q].requests.size());
    verdict[q].resize(queries[q].requests.size());
    if (S.empty) continue;
    for (size_t j = 0; j < queries[q].requests.size(); ++j) {
      refined[q][j].assign(S.jobs[j].full.size(), Pose{});
      verdict[q][j].assign(S.jobs[j].full.size(), -1);
    }
  }
  std::vector<SelectionWalk> walks(n_requests);
  std::vector<bool> independent(n_requests);
  for (size_t r = 0; r < n_requests; ++r) {
    const FrameRequest& R = requests[r];
    bool any_quota = false;
    for (const FindOptions& o : R.options) any_quota |= o.num_matches > 0;
    // HALCON's two-stage quota: per-model quotas and overlaps first, then the
    // cross-model overlap without refilling quotas (matcher.cpp:1108-1112)
    independent[r] = R.per_model_quota_before_overlap && R.selection.num_matches < 0 && R.selection.max_overlap >= 0 &&
                     any_quota;
    SelectionWalk& w = walks[r];
    w.models = R.models.data(); w.opts = R.options.data(); w.n_models = R.models.size();
    w.pool = &cand[r]; w.where = &where[r];
    w.s = independent[r] ? Selection{-1, -1} : R.selection;
    w.with_refine = true;
    w.ord.resize(cand[r].size());
    for (size_t i = 0; i < w.ord.size(); ++i) w.ord[i] = i;
    std::sort(w.ord.begin(), w.ord.end(), [&](size_t a, size_t b) { return multi_less(cand[r][a], cand[r][b]); });
    w.init();
  }
  {
    size_t total = 0;
    for (size_t r = 0; r < n_requests; ++r) total += walks[r].ord.size();
    if (total > kOneRound)
      for (size_t r = 0; r < n_requests; ++r)
        if (walks[r].unbounded) walks[r].batch = std::min(walks[r].batch, kUnboundedFirstRound);
  }
  // ---- the walks of all requests, with shared speculative refinement rounds
  auto v = [&](const SelectionWalk::Ref& x) { return verdict[x.slot][x.frame][x.index]; };
  auto pose = [&](const SelectionWalk::Ref& x) { return refined[x.slot][x.frame][x.index]; };
  for (;;) {
    std::vector<RefineRequest> reqs;
    std::vector<SelectionWalk::Ref> targets;
    for (size_t r = 0; r < n_requests; ++r) {
      SelectionWalk& w = walks[r];
      w.advance(v, pose);
      if (w.done()) continue;
      std::vector<SelectionWalk::Ref> t;
      w.requests(t, v, [&](const SelectionWalk::Ref& x) { verdict[x.slot][x.frame][x.index] = -2; });
      for (const auto& x : t) {
        reqs.push_back({multi_slots_[x.slot].get(), uint32_t(x.frame), multi_slots_[x.slot]->jobs[x.frame].full[x.index],
                        {}, 0});
        targets.push_back(x);
      }
    }
    if (reqs.empty()) break;
    refine_round(reqs);
    if (stats) {
      stats->refined += int(reqs.size());
      ++stats->refine_rounds;
    }
    for (size_t k = 0; k < reqs.size(); ++k) {
      const SelectionWalk::Ref& x = targets[k];
      refined[x.slot][x.frame][x.index] = reqs[k].out;
      verdict[x.slot][x.frame][x.index] = reqs[k].verdict;
    }
  }
  for (size_t r = 0; r < n_requests; ++r) {
    const FrameRequest& R = requests[r];
    std::vector<ModelMatch> result = walks[r].finish();
    if (independent[r]) {
      SelectionWalk w2;
      w2.models = R.models.data(); w2.opts = R.options.data(); w2.n_models = R.models.size();
      w2.pool = &result; w2.where = nullptr;
      w2.s = R.selection;
      w2.with_refine = false;
      w2.ord.resize(result.size());
      for (size_t i = 0; i < w2.ord.size(); ++i) w2.ord[i] = i;
      w2.init();
      auto never = [](const SelectionWalk::Ref&) { return 0; };
      w2.advance(never, [](const SelectionWalk::Ref&) { return Pose{}; });
      result = w2.finish();
    }
    out[r] = std::move(result);
  }
  if (stats) {
    stats->queries = int(queries.size());
    for (size_t q = 0; q < queries.size(); ++q) {
      const Slot& S = *multi_slots_[q];
      if (S.empty) continue;
      for (const Job& J : S.jobs) { stats->coarse_flops += J.coarse_flops; stats->candidates += int(J.full.size()); }
    }
    if (stats->queue_queries) {
      float ms = 0;
      check(hipEventElapsedTime(&ms, queue_events_[0], queue_events_[1]), "hipEventElapsedTime");
      stats->coarse_queue_us = 1000.0 * ms;
    }
    stats->total_us = now_us() - t_call;
  }
  return out;
}

}  // namespace mvamd
