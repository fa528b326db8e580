// Test infrastructure only: a scalar FP64 CPU restatement of mvkit's NCC
// template-matching create/find path, used as the checker for the HIP build.
// Nothing in the product library links, loads or calls this file; only the
// tests, the smoke check and the benchmark's CPU-baseline leg do.
//
// Every function cites the reference file:line it restates (paths relative to
// the mvkit source tree).  The reference's x86 build runs its AVX2 kernels by
// default (src/runtime/cpu/cpu_features.cpp:78), so where the AVX2 and scalar
// formulations differ (window moments from row prefixes, the low-variance
// rule, 4-lane accumulation order) this file follows the AVX2 choice.  The
// order in which the 16/8/4-lane kernels sum a dot product lane-wise is kept
// where it is cheap to keep; the CPU-only multi-model caches are not restated
// (they are designed to be result-neutral, matcher.cpp:1077-1124).
//
// Scope: single-channel u8 ("byte") and u16 ("uint2") search images, u8
// templates with a full rectangular domain or persisted ones with any mask,
// find (matcher.cpp:1063-1076) and find_models (matcher.cpp:1077-1124).
//
// Build: g++ -O2 -ffp-contract=off (the reference compiles its NCC kernels
// with -ffp-contract=off, cmake/Targets.cmake:177); see oracle/Makefile.

#include <algorithm>
#include <array>
#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc {

constexpr double kMissing = -std::numeric_limits<double>::infinity();
constexpr double kTau = 6.283185307179586476925286766559;
constexpr double kEps = std::numeric_limits<double>::epsilon();

// Status codes of include/mvkit/c/mv_status.h:10-24.
enum Status : int {
  OK = 0, INVALID_ARGUMENT = -1, INVALID_STATE = -7, LIMIT_EXCEEDED = -103, UNSUPPORTED = -6, DECODE_ERROR = -101
};

struct Failure : std::runtime_error {
  int status;
  Failure(int s, const char* m) : std::runtime_error(m), status(s) {}
};
[[noreturn]] static void fail(int s, const char* m) { throw Failure(s, m); }

// ---------------------------------------------------------------- model types
// src/ncc_matching/model.hpp:26-63
struct Sample { int32_t r, c; double w; };
struct Run { int32_t r, b, e; };
struct WRun { int32_t r, b, e; double w; };
struct Patch {
  double angle = 0, orow = 0, ocol = 0;
  int32_t rmin = 0, rmax = 0, cmin = 0, cmax = 0;
  bool bounds = false;
  std::vector<Sample> s;
  std::vector<double> relr, relc;
  std::vector<Run> runs;
  std::vector<WRun> wruns;
  double wsum = 0;
};
struct Level {
  size_t w = 0, h = 0;
  double orow = 0, ocol = 0;
  std::vector<double> gray;
  std::vector<uint8_t> mask;
  std::vector<Patch> patches;
};
struct CreateOpts { int levels = 0; bool absolute = false; double start = 0, extent = 0, step = 0; };
struct Model {
  CreateOpts opt;
  size_t sample_count = 0;
  double orow = 0, ocol = 0;
  Patch original;
  std::vector<Level> levels;
};
struct Match { double row = 0, col = 0, angle = 0, score = 0; };
struct FindOpts {
  double start = 0, extent = 0, min_score = .8, max_overlap = .5;
  size_t num_matches = 1, max_candidates = 10000;
  bool subpixel = true;
  int highest = 0, lowest = 1;
};

// ------------------------------------------------------------------- create
// normalize: model.cpp:39-65 (file lines 122-148)
struct Prepared { double mean, energy; };
static bool normalize(Patch& p, const Prepared* pre) {
  if (p.s.size() < 4) return false;
  double mean = pre ? pre->mean : 0;
  if (!pre)
    for (size_t i = 0; i < p.s.size(); ++i) mean += (p.s[i].w - mean) / double(i + 1);
  double energy = pre ? pre->energy : 0;
  if (pre) {
    for (auto& x : p.s) x.w -= mean;
  } else {
    for (auto& x : p.s) { x.w -= mean; energy += x.w * x.w; }
  }
  if (!(energy > 0) || !std::isfinite(energy)) return false;
  const double scale = 1 / std::sqrt(energy);
  p.rmin = p.rmax = p.s.front().r;
  p.cmin = p.cmax = p.s.front().c;
  for (auto& x : p.s) {
    x.w *= scale;
    p.wsum += x.w;
    p.rmin = std::min(p.rmin, x.r); p.rmax = std::max(p.rmax, x.r);
    p.cmin = std::min(p.cmin, x.c); p.cmax = std::max(p.cmax, x.c);
    if (!p.runs.empty() && p.runs.back().r == x.r && p.runs.back().e + 1 == x.c) p.runs.back().e = x.c;
    else p.runs.push_back({x.r, x.c, x.c});
    if (!p.wruns.empty() && p.wruns.back().r == x.r && p.wruns.back().e + 1 == x.c && p.wruns.back().w == x.w)
      p.wruns.back().e = x.c;
    else p.wruns.push_back({x.r, x.c, x.c, x.w});
  }
  p.bounds = true;
  return true;
}

// rotate: model.cpp:66-98 (file lines 149-181)
static Patch rotate(const Level& lv, double angle, const Prepared* pre = nullptr) {
  Patch p; p.angle = angle;
  if (std::abs(std::remainder(angle, kTau)) < 1e-12) {
    p.orow = lv.orow; p.ocol = lv.ocol;
    for (size_t r = 0; r < lv.h; ++r)
      for (size_t c = 0; c < lv.w; ++c)
        if (lv.mask[r * lv.w + c]) p.s.push_back({int(r), int(c), lv.gray[r * lv.w + c]});
    if (!normalize(p, pre)) p.s.clear();
    return p;
  }
  const double ca = std::cos(angle), sa = std::sin(angle);
  double rmin = 0, rmax = 0, cmin = 0, cmax = 0;
  for (double r : {0., double(lv.h - 1)})
    for (double c : {0., double(lv.w - 1)}) {
      const double dr = r - lv.orow, dc = c - lv.ocol;
      const double rr = ca * dr - sa * dc, cc = sa * dr + ca * dc;
      rmin = std::min(rmin, rr); rmax = std::max(rmax, rr);
      cmin = std::min(cmin, cc); cmax = std::max(cmax, cc);
    }
  const double rw = std::ceil(rmax) - std::floor(rmin) + 1, cw = std::ceil(cmax) - std::floor(cmin) + 1;
  if (rw > 4 * 1048576.0 / cw) fail(LIMIT_EXCEEDED, "rotated bounding box exceeds budget");
  for (int r = int(std::floor(rmin)); r <= int(std::ceil(rmax)); ++r)
    for (int c = int(std::floor(cmin)); c <= int(std::ceil(cmax)); ++c) {
      const double rr = ca * r + sa * c + lv.orow, cc = -sa * r + ca * c + lv.ocol;
      if (rr < 0 || cc < 0 || rr > lv.h - 1 || cc > lv.w - 1) continue;
      const size_t r0 = size_t(rr), c0 = size_t(cc);
      const size_t r1 = std::min(r0 + 1, lv.h - 1), c1 = std::min(c0 + 1, lv.w - 1);
      const size_t a = r0 * lv.w + c0, b = r0 * lv.w + c1, d = r1 * lv.w + c0, e = r1 * lv.w + c1;
      if (!lv.mask[a] || !lv.mask[b] || !lv.mask[d] || !lv.mask[e]) continue;
      const double fr = rr - r0, fc = cc - c0;
      p.s.push_back({r, c, (1 - fr) * ((1 - fc) * lv.gray[a] + fc * lv.gray[b]) +
                               fr * ((1 - fc) * lv.gray[d] + fc * lv.gray[e])});
    }
  if (!normalize(p, nullptr)) p.s.clear();
  return p;
}

// validate_zero: model.cpp:22-36 (file lines 105-119)
static bool level_usable(const Level& lv, size_t& count) {
  const size_t n_px = lv.w * lv.h;
  size_t n = 0; double mean = 0;
  for (size_t i = 0; i < n_px; ++i) if (lv.mask[i]) mean += (lv.gray[i] - mean) / double(++n);
  count = n;
  if (n < 4) return false;
  double energy = 0;
  for (size_t i = 0; i < n_px; ++i) if (lv.mask[i]) { const double v = lv.gray[i] - mean; energy += v * v; }
  return energy > 0 && std::isfinite(energy);
}

static Model rebuild(Level first, const CreateOpts& o, double orow, double ocol);

// create: model.cpp:120-171 (file lines 203-254), u8 template over its
// domain given as canonical runs (row, first column, last column) inside the
// image; an empty list means the whole image.
static Model create(const uint8_t* px, size_t w, size_t h, size_t stride, const CreateOpts& o,
                    std::vector<std::array<int64_t, 3>> runs = {}) {
  if (!w || !h) fail(INVALID_ARGUMENT, "one nonempty single-channel image required");
  if (!std::isfinite(o.start) || !std::isfinite(o.extent) || !std::isfinite(o.step) || o.extent < 0 ||
      o.extent > kTau || !std::isfinite(o.start + o.extent) || o.step < 0 || o.levels < 0 || o.levels > 16)
    fail(INVALID_ARGUMENT, "invalid levels or angle range");
  if (runs.empty())
    for (size_t r = 0; r < h; ++r) runs.push_back({int64_t(r), 0, int64_t(w) - 1});
  int64_t area_i = 0, r0 = runs.front()[0], r1 = runs.back()[0], c0 = runs.front()[1], c1 = runs.front()[2];
  for (const auto& x : runs) { area_i += x[2] - x[1] + 1; c0 = std::min(c0, x[1]); c1 = std::max(c1, x[2]); }
  if (area_i < 4) fail(INVALID_ARGUMENT, "template domain requires at least four pixels");
  if (r0 < 0 || c0 < 0 || r1 >= int64_t(h) || c1 >= int64_t(w)) fail(INVALID_ARGUMENT, "template domain must be inside the image");
  const size_t bw = size_t(c1 - c0 + 1), bh = size_t(r1 - r0 + 1);
  if (bh > 1048576 / bw) fail(LIMIT_EXCEEDED, "template bounding box exceeds one million pixels");
  Level first; first.w = bw; first.h = bh;
  first.gray.assign(bw * bh, 0.0); first.mask.assign(bw * bh, 0);
  long double sum_row = 0, sum_col = 0; int64_t isum = 0;
  for (const auto& x : runs) {
    const int64_t count = x[2] - x[1] + 1;
    sum_row += static_cast<long double>(x[0]) * count;
    sum_col += (static_cast<long double>(x[1]) + x[2]) * count / 2;
    for (int64_t c = x[1]; c <= x[2]; ++c) {
      const size_t i = size_t(x[0] - r0) * bw + size_t(c - c0);
      first.gray[i] = double(px[size_t(x[0]) * stride + size_t(c)]); first.mask[i] = 1;
      isum += px[size_t(x[0]) * stride + size_t(c)];
    }
  }
  const long double area = (long double)area_i;
  Model m0; m0.orow = double(sum_row / area); m0.ocol = double(sum_col / area);
  first.orow = double((sum_row - static_cast<long double>(r0) * area) / area);
  first.ocol = double((sum_col - static_cast<long double>(c0) * area) / area);
  const double mean = double(isum) / double(area_i);
  for (size_t i = 0; i < first.gray.size(); ++i) if (first.mask[i]) first.gray[i] -= mean;
  return rebuild(std::move(first), o, m0.orow, m0.ocol);
}

// rebuild: model.cpp:173-242 (file lines 256-325), from a centred level 0
// with any mask (the reference's own templates may have reduced domains).
static Model rebuild(Level first, const CreateOpts& o, double orow, double ocol) {
  Model m; m.opt = o; m.orow = orow; m.ocol = ocol;
  m.sample_count = size_t(std::count(first.mask.begin(), first.mask.end(), uint8_t(1)));
  m.levels.push_back(std::move(first));
  Prepared first_norm{};
  {
    const Level& lv = m.levels.front(); Patch& p = m.original;
    p.orow = lv.orow; p.ocol = lv.ocol;
    double mu = 0; size_t cnt = 0;
    for (size_t i = 0; i < lv.gray.size(); ++i) if (lv.mask[i]) mu += (lv.gray[i] - mu) / double(++cnt);
    double energy = 0;
    for (size_t r = 0; r < lv.h; ++r)
      for (size_t c = 0; c < lv.w; ++c)
        if (lv.mask[r * lv.w + c]) {
          const double v = lv.gray[r * lv.w + c] - mu; energy += v * v; p.s.push_back({int(r), int(c), v});
        }
    if (cnt < 4 || !(energy > 0) || !std::isfinite(energy)) fail(INVALID_ARGUMENT, "constant template cannot define an NCC model");
    first_norm = {mu, energy};
    const double norm = std::sqrt(energy);
    p.rmin = p.rmax = p.s.front().r; p.cmin = p.cmax = p.s.front().c;
    for (auto& x : p.s) {
      x.w /= norm;
      p.rmin = std::min(p.rmin, x.r); p.rmax = std::max(p.rmax, x.r);
      p.cmin = std::min(p.cmin, x.c); p.cmax = std::max(p.cmax, x.c);
      p.relr.push_back(x.r - lv.orow); p.relc.push_back(x.c - lv.ocol);
    }
    p.bounds = true;
  }
  const size_t wanted = o.levels ? size_t(o.levels) : 16;
  while (m.levels.size() < wanted) {
    const Level& pv = m.levels.back();
    if (pv.w < 8 || pv.h < 8) break;
    Level nx; nx.w = pv.w / 2; nx.h = pv.h / 2;
    nx.orow = (pv.orow - .5) * .5; nx.ocol = (pv.ocol - .5) * .5;
    nx.gray.resize(nx.w * nx.h); nx.mask.resize(nx.w * nx.h);
    for (size_t r = 0; r < nx.h; ++r)
      for (size_t c = 0; c < nx.w; ++c) {
        const size_t i = 2 * r * pv.w + 2 * c;
        nx.mask[r * nx.w + c] = pv.mask[i] && pv.mask[i + 1] && pv.mask[i + pv.w] && pv.mask[i + pv.w + 1];
        // box2_gray: src/imageproc/gray_pyramid.hpp:8-13
        const double* row0 = pv.gray.data() + 2 * r * pv.w;
        const double* row1 = pv.gray.data() + (2 * r + 1) * pv.w;
        const size_t c0 = 2 * c, c1 = c0 + 1 < pv.w ? c0 + 1 : c0;
        nx.gray[r * nx.w + c] = .25 * (row0[c0] + row0[c1] + row1[c0] + row1[c1]);
      }
    size_t count = 0;
    const bool ok = level_usable(nx, count);
    if (!ok || count < 8 || (m.levels.size() >= 2 && count < 64)) break;
    m.levels.push_back(std::move(nx));
  }
  if (o.levels && m.levels.size() != size_t(o.levels))
    fail(INVALID_ARGUMENT, "requested pyramid level is degenerate or too small");
  const double radius = std::hypot(double(m.levels.front().w), double(m.levels.front().h)) * .5;
  m.opt.step = o.step > 0 ? o.step : std::max(.01, std::min(.15, 1 / std::max(1., radius)));
  const double intervals = std::ceil(o.extent / m.opt.step);
  if (!std::isfinite(intervals) || intervals >= 4096) fail(LIMIT_EXCEEDED, "angle grid exceeds limit");
  const size_t n_angles = size_t(intervals) + 1;
  size_t total = 0;
  for (auto& lv : m.levels) {
    for (size_t i = 0; i < n_angles; ++i) {
      const double angle = o.start + std::min(o.extent, i * m.opt.step);
      Patch p = rotate(lv, angle, &lv == &m.levels.front() ? &first_norm : nullptr);
      if (p.s.empty()) fail(INVALID_ARGUMENT, "rotation produces a degenerate template");
      if (p.s.size() > 16000000 - total) fail(LIMIT_EXCEEDED, "model raster budget exceeded");
      total += p.s.size();
      lv.patches.push_back(std::move(p));
    }
  }
  return m;
}

// ------------------------------------------------------------- persistence
// "MVKIT_NCC 1" text model: persist.cpp:14-84, number formatting of
// core/model_text.hpp:18-62 (classic locale, integers within 2^53 as
// integers, other reals with 17 significant digits), FNV-1a-64 checksum of
// the body (persist.cpp:18-21).
static uint64_t fnv1a64(const std::string& t) {
  uint64_t h = 14695981039346656037ULL;
  for (unsigned char c : t) { h ^= c; h *= 1099511628211ULL; }
  return h;
}
static void put_int(std::string& s, long long v) { s += std::to_string(v); }
static void put_real(std::string& s, double v) {
  if (v == 0.) { if (std::signbit(v)) s += '-'; s += '0'; return; }
  if (v == 1.) { s += '1'; return; }
  if (v >= -9007199254740992. && v <= 9007199254740992.) {
    const long long i = static_cast<long long>(v);
    if (static_cast<double>(i) == v) { put_int(s, i); return; }
  }
  char buf[64];
  std::snprintf(buf, sizeof buf, "%.17g", v);
  s += buf;
}
static std::string serialize(const Model& m) {
  if (m.levels.empty()) fail(INVALID_STATE, "cannot save empty model");
  const Level& lv = m.levels.front();
  std::string t;
  put_int(t, m.opt.levels); t += ' '; put_int(t, m.opt.absolute ? 1 : 0); t += ' ';
  put_real(t, m.opt.start); t += ' '; put_real(t, m.opt.extent); t += ' '; put_real(t, m.opt.step); t += '\n';
  put_real(t, m.orow); t += ' '; put_real(t, m.ocol); t += ' ';
  put_int(t, (long long)lv.w); t += ' '; put_int(t, (long long)lv.h); t += '\n';
  put_real(t, lv.orow); t += ' '; put_real(t, lv.ocol); t += '\n';
  for (size_t i = 0; i < lv.gray.size(); ++i) { put_int(t, lv.mask[i]); t += ' '; put_real(t, lv.gray[i]); t += '\n'; }
  if (t.size() > 64u * 1024u * 1024u) fail(LIMIT_EXCEEDED, "serialized model exceeds file budget");
  return "MVKIT_NCC 1\n" + std::to_string(fnv1a64(t)) + "\n" + t;
}

// Fields of the body: ASCII whitespace separated, an optional leading '+',
// then the longest number (core/model_text.hpp:113-189).
struct FieldReader {
  const char* p; const char* e;
  void ws() { while (p < e && (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t' || *p == '\v' || *p == '\f')) ++p; }
  bool sign() {
    ws();
    if (p == e) return false;
    if (*p == '+') { ++p; if (p == e || *p == '+' || *p == '-') return false; }
    return true;
  }
  long long integer() {
    long long v = 0;
    if (!sign()) fail(DECODE_ERROR, "invalid or truncated model field");
    auto r = std::from_chars(p, e, v);
    if (r.ec != std::errc{}) fail(DECODE_ERROR, "invalid or truncated model field");
    p = r.ptr;
    return v;
  }
  double real() {
    double v = 0;
    if (!sign()) fail(DECODE_ERROR, "invalid or truncated model field");
    auto r = std::from_chars(p, e, v);
    if (r.ec != std::errc{}) fail(DECODE_ERROR, "invalid or truncated model field");
    p = r.ptr;
    if (!std::isfinite(v)) fail(DECODE_ERROR, "non-finite model field");
    return v;
  }
  size_t count(size_t limit) {
    const long long v = integer();
    if (v < 0 || (unsigned long long)v > limit) fail(LIMIT_EXCEEDED, "model count exceeds budget");
    return size_t(v);
  }
};

static Model deserialize(const std::string& bytes) {
  if (bytes.size() > 64u * 1024u * 1024u + 128) fail(LIMIT_EXCEEDED, "model file exceeds budget");
  const size_t nl = bytes.find('\n');
  const std::string magic = bytes.substr(0, nl);
  if (magic != "MVKIT_NCC 1")
    fail(magic.rfind("MVKIT_NCC ", 0) == 0 ? UNSUPPORTED : DECODE_ERROR, "invalid model signature or version");
  size_t pos = nl == std::string::npos ? bytes.size() : nl + 1;
  while (pos < bytes.size() && std::isspace((unsigned char)bytes[pos])) ++pos;
  uint64_t expected = 0;
  auto r = std::from_chars(bytes.data() + pos, bytes.data() + bytes.size(), expected);
  if (r.ec != std::errc{}) fail(DECODE_ERROR, "invalid or truncated model field");
  pos = size_t(r.ptr - bytes.data());
  if (pos >= bytes.size() || bytes[pos] != '\n') fail(DECODE_ERROR, "invalid checksum header");
  const std::string text = bytes.substr(pos + 1);
  if (fnv1a64(text) != expected) fail(DECODE_ERROR, "NCC model checksum mismatch");
  FieldReader in{text.data(), text.data() + text.size()};
  CreateOpts o;
  o.levels = int(in.count(16)); o.absolute = in.count(1) != 0;
  o.start = in.real(); o.extent = in.real(); o.step = in.real();
  if (o.extent < 0 || o.extent > 6.283185307179586 || o.step <= 0 || !std::isfinite(o.start + o.extent))
    fail(DECODE_ERROR, "invalid trained angle interval");
  const double orow = in.real(), ocol = in.real();
  Level first;
  first.w = in.count(1048576); first.h = in.count(1048576);
  if (!first.w || !first.h || first.h > 1048576 / first.w) fail(LIMIT_EXCEEDED, "invalid model dimensions");
  first.orow = in.real(); first.ocol = in.real();
  first.gray.resize(first.w * first.h); first.mask.resize(first.gray.size());
  long double rs = 0, cs = 0; size_t n = 0;
  for (size_t i = 0; i < first.gray.size(); ++i) {
    first.mask[i] = uint8_t(in.count(1)); first.gray[i] = in.real();
    if (first.mask[i]) { rs += i / first.w; cs += i % first.w; ++n; }
  }
  if (n < 4 || std::abs(first.orow - double(rs / n)) > 1e-10 || std::abs(first.ocol - double(cs / n)) > 1e-10 ||
      orow < first.orow || ocol < first.ocol || std::abs((orow - first.orow) - std::round(orow - first.orow)) > 1e-8 ||
      std::abs((ocol - first.ocol) - std::round(ocol - first.ocol)) > 1e-8)
    fail(DECODE_ERROR, "invalid template domain centroid");
  in.ws();
  if (in.p != in.e) fail(DECODE_ERROR, "trailing model data");
  try {
    return rebuild(std::move(first), o, orow, ocol);
  } catch (const Failure& f) {
    fail(f.status == LIMIT_EXCEEDED ? LIMIT_EXCEEDED : DECODE_ERROR, "invalid prepared NCC model");
  }
}

// ------------------------------------------------------------- search image
// Gray: matcher.cpp:28-40.  Level 0 borrows bytes until the candidate region is
// converted to doubles (matcher.cpp:612-629); levels >= 1 are dense doubles.
struct Gray {
  size_t w = 0, h = 0;
  const uint8_t* bytes = nullptr; size_t stride = 0; double dc = 0;
  std::vector<double> v;
  size_t sw = 0, sh = 0, sr = 0, sc = 0;  // storage rectangle
  double at(size_t r, size_t c) const {
    if (bytes) return double(bytes[r * stride + c]) - dc;
    return v[(r - sr) * sw + c - sc];
  }
  const double* data() const { return v.data(); }
};
struct Prefix { std::vector<double> sum, sq; };

static Prefix build_prefix(const Gray& g) {  // matcher.cpp:487-495
  Prefix p; p.sum.resize((g.w + 1) * g.h); p.sq.resize(p.sum.size());
  for (size_t r = 0; r < g.h; ++r) {
    const size_t b = r * (g.w + 1); double s = 0, q = 0; p.sum[b] = 0; p.sq[b] = 0;
    for (size_t c = 0; c < g.w; ++c) { const double x = g.v[r * g.w + c]; s += x; q += x * x; p.sum[b + c + 1] = s; p.sq[b + c + 1] = q; }
  }
  return p;
}

// ------------------------------------------------------------ window scores
static inline double clamp1(double x) { return std::clamp(x, -1., 1.); }
static inline double polar(double x, bool absolute) { return absolute ? std::abs(x) : x; }

// Scalar discrete: matcher.cpp:87-106 (non-vectorized branch).
static double discrete_scalar(const Gray& g, const Patch& p, int row, int col, bool absolute) {
  const auto& f = p.s.front();
  const double off = g.at(row + f.r, col + f.c);
  double sum = 0, sq = 0, dot = 0;
  for (const auto& x : p.s) { const double v = g.at(row + x.r, col + x.c) - off; sum += v; sq += v * v; dot += x.w * v; }
  const double var = sq - sum * sum / double(p.s.size());
  if (!(var > 0)) return kMissing;
  return polar(clamp1(dot / std::sqrt(var)), absolute);
}

// correlation_avx2: kernels_avx2.cpp:239-258 (4-lane accumulators over groups
// of four samples, lanes combined as (a0+a1)+(a2+a3), then the scalar tail).
static double correlation_gather(const double* img, size_t w, const Patch& p, int row, int col, bool absolute) {
  const auto& f = p.s.front();
  const double dc = img[(row + f.r) * w + col + f.c];
  double s4[4] = {0, 0, 0, 0}, q4[4] = {0, 0, 0, 0}, d4[4] = {0, 0, 0, 0};
  size_t i = 0; const size_t n = p.s.size();
  for (; i + 4 <= n; i += 4)
    for (int l = 0; l < 4; ++l) {
      const auto& x = p.s[i + l];
      const double q = img[(row + x.r) * w + col + x.c] - dc;
      s4[l] += q; q4[l] += q * q; d4[l] += q * x.w;
    }
  double s = (s4[0] + s4[1]) + (s4[2] + s4[3]), sq = (q4[0] + q4[1]) + (q4[2] + q4[3]), d = (d4[0] + d4[1]) + (d4[2] + d4[3]);
  for (; i < n; ++i) { const auto& x = p.s[i]; const double q = img[(row + x.r) * w + col + x.c] - dc; s += q; sq += q * q; d += q * x.w; }
  const double var = sq - s * s / double(n);
  if (!(var > 0)) return kMissing;
  return polar(clamp1(d / std::sqrt(var)), absolute);
}

// discrete with vectorize: matcher.cpp:87-94.
static double discrete(const Gray& g, const Patch& p, int row, int col, bool absolute, bool vectorize) {
  if (row + p.rmin < 0 || col + p.cmin < 0 || row + p.rmax >= int64_t(g.h) || col + p.cmax >= int64_t(g.w)) return kMissing;
  if (row + p.rmin < int64_t(g.sr) || col + p.cmin < int64_t(g.sc) || row + p.rmax >= int64_t(g.sr + g.sh) ||
      col + p.cmax >= int64_t(g.sc + g.sw))
    return kMissing;
  if (vectorize && !g.bytes) return correlation_gather(g.data(), g.sw, p, row - int(g.sr), col - int(g.sc), absolute);
  return discrete_scalar(g, p, row, col, absolute);
}

// Row-prefix window moments in run order: kernels_avx2.cpp:18-34, 123.
static void prefix_moments(const Prefix& P, size_t w, const Patch& p, int row, int col, double& sum, double& sq) {
  sum = 0; sq = 0; const size_t pitch = w + 1;
  for (const auto& r : p.runs) {
    const size_t a = (row + r.r) * pitch + col + r.b, b = (row + r.r) * pitch + col + r.e + 1;
    sum += P.sum[b] - P.sum[a]; sq += P.sq[b] - P.sq[a];
  }
}

enum class DotOrder { Single, Four };
// One window of the 4/8/16-lane prefix kernels (kernels_avx2.cpp:145-238):
// the dot product is taken either from weight runs over the prefix sums (when
// 2*#weight_runs < #samples) or sample by sample with the lane's accumulator
// layout; the low-variance rule returns -inf on the exact byte lattice and a
// direct recompute otherwise.
static double prefix_window(const double* img, const Prefix& P, size_t w, const Patch& p, int row, int col,
                            bool absolute, bool exact_lattice, DotOrder order) {
  double sum, sq; prefix_moments(P, w, p, row, col, sum, sq);
  const size_t n = p.s.size();
  double dot = 0;
  if (!p.wruns.empty() && p.wruns.size() * 2 < n) {
    const size_t pitch = w + 1;
    for (const auto& r : p.wruns) {
      const size_t a = (row + r.r) * pitch + col + r.b, b = (row + r.r) * pitch + col + r.e + 1;
      dot += (P.sum[b] - P.sum[a]) * r.w;
    }
  } else if (order == DotOrder::Single) {
    for (const auto& x : p.s) dot += img[(row + x.r) * w + col + x.c] * x.w;
  } else {
    double d4[4] = {0, 0, 0, 0}; size_t i = 0;
    for (; i + 4 <= n; i += 4)
      for (int l = 0; l < 4; ++l) { const auto& x = p.s[i + l]; d4[l] += img[(row + x.r) * w + col + x.c] * x.w; }
    dot = (d4[0] + d4[1]) + (d4[2] + d4[3]);
    for (; i < n; ++i) { const auto& x = p.s[i]; dot += img[(row + x.r) * w + col + x.c] * x.w; }
  }
  const double cnt = double(n);
  const double var = sq - sum * sum / cnt;
  dot = dot - sum / cnt * p.wsum;
  if (!(var > 128 * kEps * sq)) {
    if (exact_lattice && w <= 32000000 && n <= 1048576 && !p.runs.empty()) return kMissing;
    return correlation_gather(img, w, p, row, col, absolute);
  }
  return polar(clamp1(dot / std::sqrt(var)), absolute);
}

// correlation_four_avx2 (kernels_avx2.cpp:105-117): direct 4 adjacent windows,
// per-window offset, single accumulator per lane.
static void direct_four(const double* img, size_t w, const Patch& p, int row, int col, bool absolute, double* out) {
  const auto& f = p.s.front();
  for (int l = 0; l < 4; ++l) {
    const double off = img[(row + f.r) * w + col + f.c + l];
    double s = 0, sq = 0, d = 0;
    for (const auto& x : p.s) { const double q = img[(row + x.r) * w + col + x.c + l] - off; s += q; sq += q * q; d += q * x.w; }
    const double var = sq - s * s / double(p.s.size());
    const double sc = var > 0 ? clamp1(d / std::sqrt(var)) : kMissing;
    out[l] = absolute && std::isfinite(sc) ? std::abs(sc) : sc;
  }
}

// correlation_one_prefix_avx2 (kernels_avx2.cpp:118-144) on an exact byte
// lattice: moments from row prefix sums over the runs, the dot product from
// the weight runs when 8 * #weight_runs < n, else four lanes per run (lane =
// position within the run mod 4) plus a scalar tail, combined (l0+l1)+(l2+l3).
static double one_prefix(const double* img, const Prefix& P, size_t w, const Patch& p, int row, int col, bool absolute) {
  const size_t pitch = w + 1;
  double sum = 0, sq = 0;
  for (const auto& run : p.runs) {
    const size_t a = size_t(row + run.r) * pitch + size_t(col + run.b), b = size_t(row + run.r) * pitch + size_t(col + run.e + 1);
    sum += P.sum[b] - P.sum[a]; sq += P.sq[b] - P.sq[a];
  }
  const double count = double(p.s.size()), var = sq - sum * sum / count;
  if (var <= 128 * kEps * sq)
    return w <= 32000000 && p.s.size() <= 1048576 && !p.runs.empty() ? kMissing
                                                                     : correlation_gather(img, w, p, row, col, absolute);
  if (!p.wruns.empty() && p.wruns.size() * 8 < p.s.size()) {
    double dot = 0;
    for (const auto& run : p.wruns) {
      const size_t a = size_t(row + run.r) * pitch + size_t(col + run.b), b = size_t(row + run.r) * pitch + size_t(col + run.e + 1);
      dot += (P.sum[b] - P.sum[a]) * run.w;
    }
    return polar(clamp1((dot - sum / count * p.wsum) / std::sqrt(var)), absolute);
  }
  double lane[4] = {0, 0, 0, 0}, tail = 0;
  size_t index = 0;
  for (const auto& run : p.runs) {
    const double* v = img + size_t(row + run.r) * w + size_t(col + run.b);
    const size_t len = size_t(run.e - run.b + 1);
    size_t i = 0;
    for (; i + 4 <= len; i += 4)
      for (int l = 0; l < 4; ++l) lane[l] += v[i + l] * p.s[index + i + l].w;
    for (; i < len; ++i) tail += v[i] * p.s[index + i].w;
    index += len;
  }
  const double num = (lane[0] + lane[1]) + (lane[2] + lane[3]) + tail - sum / count * p.wsum;
  return polar(clamp1(num / std::sqrt(var)), absolute);
}

// correlation_four_prefix_avx2 (kernels_avx2.cpp:145-171): four neighbouring
// windows (col .. col+3).  Dot products from the weight runs when
// 2 * #weight_runs < n, else four accumulators by sample index mod 4 combined
// (d0+d1)+(d2+d3) before the tail; low-variance windows are -inf on a byte lattice.
static void four_prefix(const double* img, const Prefix& P, size_t w, const Patch& p, int row, int col, bool absolute,
                        double* out) {
  const size_t pitch = w + 1, n = p.s.size();
  double sum[4] = {}, sq[4] = {}, d0[4] = {}, d1[4] = {}, d2[4] = {}, d3[4] = {};
  for (const auto& run : p.runs) {
    const size_t a = size_t(row + run.r) * pitch + size_t(col + run.b), b = size_t(row + run.r) * pitch + size_t(col + run.e + 1);
    for (int l = 0; l < 4; ++l) { sum[l] += P.sum[b + l] - P.sum[a + l]; sq[l] += P.sq[b + l] - P.sq[a + l]; }
  }
  size_t i = 0;
  if (!p.wruns.empty() && p.wruns.size() * 2 < n) {
    for (const auto& run : p.wruns) {
      const size_t a = size_t(row + run.r) * pitch + size_t(col + run.b), b = size_t(row + run.r) * pitch + size_t(col + run.e + 1);
      for (int l = 0; l < 4; ++l) d0[l] += (P.sum[b + l] - P.sum[a + l]) * run.w;
    }
    i = n;
  }
  auto px = [&](size_t k, int l) { const auto& x = p.s[k]; return img[size_t(row + x.r) * w + size_t(col + x.c + l)]; };
  for (; i + 4 <= n; i += 4)
    for (int l = 0; l < 4; ++l) {
      d0[l] += px(i, l) * p.s[i].w; d1[l] += px(i + 1, l) * p.s[i + 1].w;
      d2[l] += px(i + 2, l) * p.s[i + 2].w; d3[l] += px(i + 3, l) * p.s[i + 3].w;
    }
  double dot[4], var[4];
  for (int l = 0; l < 4; ++l) dot[l] = (d0[l] + d1[l]) + (d2[l] + d3[l]);
  for (; i < n; ++i)
    for (int l = 0; l < 4; ++l) dot[l] += px(i, l) * p.s[i].w;
  const double count = double(n);
  for (int l = 0; l < 4; ++l) {
    var[l] = sq[l] - sum[l] * sum[l] / count;
    dot[l] = dot[l] - sum[l] / count * p.wsum;
  }
  for (int l = 0; l < 4; ++l) {
    if (var[l] <= 128 * kEps * sq[l]) {
      out[l] = w <= 32000000 && n <= 1048576 && !p.runs.empty() ? kMissing
                                                                 : correlation_gather(img, w, p, row, col + l, absolute);
      continue;
    }
    out[l] = polar(clamp1(dot[l] / std::sqrt(var[l])), absolute);
  }
}

// ---------------------------------------------------- continuous (bilinear)
// strictly_inside: kernels_avx2.cpp:261-277
static bool strictly_inside(size_t w, size_t h, const Level& lv, const Patch& p, const Match& pose, double ca, double sa) {
  if (!p.bounds || (p.rmin == p.rmax && p.cmin == p.cmax)) return false;
  const double dr[2] = {p.rmin - lv.orow, p.rmax - lv.orow};
  const double dc[2] = {p.cmin - lv.ocol, p.cmax - lv.ocol};
  double ymin = std::numeric_limits<double>::infinity(), xmin = ymin, ymax = -ymin, xmax = -ymin;
  for (double r : dr)
    for (double c : dc) {
      const double y = pose.row + ca * r - sa * c, x = pose.col + sa * r + ca * c;
      ymin = std::min(ymin, y); ymax = std::max(ymax, y); xmin = std::min(xmin, x); xmax = std::max(xmax, x);
    }
  const double margin = 64 * kEps * (1 + std::abs(pose.row) + std::abs(pose.col) + std::abs(dr[0]) + std::abs(dr[1]) +
                                     std::abs(dc[0]) + std::abs(dc[1]) + double(w) + double(h));
  return ymin > margin && xmin > margin && ymax < double(h - 1) - margin && xmax < double(w - 1) - margin;
}

struct Moments { double sum = 0, sq = 0, dot = 0, gs[3] = {}, qg[3] = {}, wg[3] = {}, gg[3][3] = {}; };

// bilinear<Derivative,...> non-staged (kernels_avx2.cpp:327-477) on a dense
// storage image: 4-lane accumulation over groups of four, lanes combined as
// (a0+a1)+(a2+a3), then scalar tail.  Returns false where a sample is outside.
static bool bilinear(const double* img, size_t w, size_t h, const Level& lv, const Patch& p, const Match& pose,
                     bool deriv, bool rot, Moments& out) {
  const double ca = std::cos(pose.angle), sa = std::sin(pose.angle);
  const bool interior = strictly_inside(w, h, lv, p, pose, ca, sa);
  const bool aligned = pose.angle == 0;
  auto coord = [&](size_t i, double& y, double& x) {
    const double dr = p.relr[i], dc = p.relc[i];
    if (aligned) { y = pose.row + dr; x = pose.col + dc; }
    else { y = (pose.row + ca * dr) - sa * dc; x = (pose.col + sa * dr) + ca * dc; }
  };
  // scalar lambda (kernels_avx2.cpp:330-339): always the rotated form
  auto scalar = [&](size_t i, double& q, double* g) {
    const double dr = p.s[i].r - lv.orow, dc = p.s[i].c - lv.ocol;
    const double y = pose.row + ca * dr - sa * dc, x = pose.col + sa * dr + ca * dc;
    if (y < 0 || x < 0 || y > h - 1 || x > w - 1 || (deriv && (y >= h - 1 || x >= w - 1))) return false;
    const size_t r = size_t(y), c = size_t(x), r1 = std::min(r + 1, h - 1), c1 = std::min(c + 1, w - 1);
    const double fy = y - r, fx = x - c;
    const double a = img[r * w + c], b = img[r * w + c1], d = img[r1 * w + c], e = img[r1 * w + c1];
    q = (1 - fy) * ((1 - fx) * a + fx * b) + fy * ((1 - fx) * d + fx * e);
    if (deriv) {
      g[0] = (1 - fx) * (d - a) + fx * (e - b);
      g[1] = (1 - fy) * (b - a) + fy * (e - d);
      g[2] = g[0] * (-sa * dr - ca * dc) + g[1] * (ca * dr - sa * dc);
    }
    return true;
  };
  double off = 0, unused[3];
  if (!scalar(0, off, unused)) return false;
  const int axes = rot ? 3 : 2;
  double S[4] = {}, Q[4] = {}, D[4] = {};
  double GS[3][4] = {}, QG[3][4] = {}, WG[3][4] = {}, GG[3][3][4] = {};
  const size_t n = p.s.size(); size_t i = 0;
  for (; i + 4 <= n; i += 4) {
    double y[4], x[4];
    for (int l = 0; l < 4; ++l) coord(i + l, y[l], x[l]);
    if (!interior)
      for (int l = 0; l < 4; ++l) {
        const bool ok = y[l] >= 0 && x[l] >= 0 && (deriv ? y[l] < double(h - 1) : y[l] <= double(h - 1)) &&
                        (deriv ? x[l] < double(w - 1) : x[l] <= double(w - 1));
        if (!ok) return false;
      }
    for (int l = 0; l < 4; ++l) {
      const double yr = std::floor(y[l]), xc = std::floor(x[l]), fy = y[l] - yr, fx = x[l] - xc;
      const long r = long(yr), c = long(xc);
      const long r1 = interior ? r + 1 : std::min(r + 1, long(h - 1)), c1 = interior ? c + 1 : std::min(c + 1, long(w - 1));
      const double a = img[r * w + c], b = img[r * w + c1], d = img[r1 * w + c], e = img[r1 * w + c1];
      const double ux = 1 - fx, uy = 1 - fy;
      const double q = (uy * (ux * a + fx * b) + fy * (ux * d + fx * e)) - off;
      const double wt = p.s[i + l].w;
      S[l] += q; Q[l] += q * q; D[l] += wt * q;
      if (deriv) {
        double g[3];
        g[0] = ux * (d - a) + fx * (e - b);
        g[1] = uy * (b - a) + fy * (e - d);
        const double dr = p.relr[i + l], dc = p.relc[i + l];
        g[2] = rot ? g[0] * ((0 - sa * dr) - ca * dc) + g[1] * (ca * dr - sa * dc) : 0;
        for (int j = 0; j < axes; ++j) {
          GS[j][l] += g[j]; QG[j][l] += q * g[j]; WG[j][l] += wt * g[j];
          for (int k = 0; k <= j; ++k) GG[j][k][l] += g[j] * g[k];
        }
      }
    }
  }
  auto tot = [](const double* a) { return (a[0] + a[1]) + (a[2] + a[3]); };
  out.sum = tot(S); out.sq = tot(Q); out.dot = tot(D);
  if (deriv)
    for (int j = 0; j < axes; ++j) {
      out.gs[j] = tot(GS[j]); out.qg[j] = tot(QG[j]); out.wg[j] = tot(WG[j]);
      for (int k = 0; k <= j; ++k) out.gg[j][k] = tot(GG[j][k]);
    }
  for (; i < n; ++i) {
    double q, g[3];
    if (!scalar(i, q, g)) return false;
    q -= off;
    const double wt = p.s[i].w;
    out.sum += q; out.sq += q * q; out.dot += q * wt;
    if (deriv)
      for (int j = 0; j < axes; ++j) {
        out.gs[j] += g[j]; out.qg[j] += q * g[j]; out.wg[j] += wt * g[j];
        for (int k = 0; k <= j; ++k) out.gg[j][k] += g[j] * g[k];
      }
  }
  return true;
}

// continuous: matcher.cpp:121-157 (vectorized branch on a dense storage image).
static double continuous(const Gray& g, const Level& lv, const Patch& orig, const Match& pose, bool absolute) {
  Match local = pose; local.row -= g.sr; local.col -= g.sc;
  Moments m;
  if (!bilinear(g.data(), g.sw, g.sh, lv, orig, local, false, false, m)) return kMissing;
  const double var = m.sq - m.sum * m.sum / double(orig.s.size());
  if (!(var > 0)) return kMissing;
  return polar(clamp1(m.dot / std::sqrt(var)), absolute);
}

// refinement_step: matcher.cpp:170-240 (vectorized, non-cached branch).
static bool refinement_step(const Gray& g, const Level& lv, const Patch& orig, const Match& pose, bool absolute,
                            bool rot, double* delta, double* out_score, const double* min_score) {
  Match local = pose; local.row -= g.sr; local.col -= g.sc;
  Moments m;
  if (!bilinear(g.data(), g.sw, g.sh, lv, orig, local, true, rot, m)) return false;
  double sum = m.sum, energy = m.sq, dot = m.dot, gs[3], qg[3], wg[3], gg[3][3];
  for (int i = 0; i < 3; ++i) { gs[i] = m.gs[i]; qg[i] = m.qg[i]; wg[i] = m.wg[i]; for (int j = 0; j < 3; ++j) gg[i][j] = m.gg[i][j]; }
  const double n = double(orig.s.size());
  energy -= sum * sum / n;
  if (!(energy > 0)) return false;
  const double sign = absolute && dot < 0 ? -1 : 1, score = sign * dot / std::sqrt(energy);
  if (out_score) *out_score = clamp1(score);
  if (min_score && clamp1(score) <= *min_score) return false;
  const int axes = rot ? 3 : 2; double M[3][4] = {};
  for (int i = 0; i < axes; ++i) qg[i] -= sum * gs[i] / n;
  for (int i = 0; i < axes; ++i) {
    for (int j = 0; j < axes; ++j)
      M[i][j] = (gg[std::max(i, j)][std::min(i, j)] - gs[i] * gs[j] / n - qg[i] * qg[j] / energy) / energy;
    M[i][axes] = sign * wg[i] / std::sqrt(energy) - score * qg[i] / energy;
  }
  for (int i = 0; i < axes; ++i) {
    int pv = i;
    for (int j = i + 1; j < axes; ++j) if (std::abs(M[j][i]) > std::abs(M[pv][i])) pv = j;
    if (std::abs(M[pv][i]) < 1e-14) return false;
    for (int k = i; k <= axes; ++k) std::swap(M[i][k], M[pv][k]);
    const double dv = M[i][i];
    for (int k = i; k <= axes; ++k) M[i][k] /= dv;
    for (int j = 0; j < axes; ++j)
      if (j != i) { const double f = M[j][i]; for (int k = i; k <= axes; ++k) M[j][k] -= f * M[i][k]; }
  }
  for (int i = 0; i < axes; ++i) delta[i] = M[i][axes];
  if (!rot) delta[2] = 0;
  return std::isfinite(delta[0]) && std::isfinite(delta[1]) && std::isfinite(delta[2]);
}

// merge_identical: matcher.cpp:241-248
static void merge_identical(std::vector<Match>& c) {
  std::sort(c.begin(), c.end(), [](const Match& a, const Match& b) {
    if (a.row != b.row) return a.row < b.row;
    if (a.col != b.col) return a.col < b.col;
    return a.angle < b.angle;
  });
  c.erase(std::unique(c.begin(), c.end(), [](const Match& a, const Match& b) {
            return a.row == b.row && a.col == b.col && a.angle == b.angle;
          }), c.end());
}

// overlap: matcher.cpp:252-277
static double overlap(const Match& a, const Match& b, const Level& lv) {
  struct P { double x, y; };
  auto corners = [&](const Match& m) {
    std::array<P, 4> out; const double ca = std::cos(m.angle), sa = std::sin(m.angle);
    const double left = -lv.ocol - .5, right = lv.w - lv.ocol - .5, top = -lv.orow - .5, bottom = lv.h - lv.orow - .5;
    const std::array<P, 4> in = {P{left, top}, P{right, top}, P{right, bottom}, P{left, bottom}};
    for (int i = 0; i < 4; ++i) out[i] = {m.col - a.col + ca * in[i].x + sa * in[i].y, m.row - a.row - sa * in[i].x + ca * in[i].y};
    return out;
  };
  const auto clip = corners(b);
  std::array<P, 16> poly{}, next{};
  const auto src = corners(a);
  std::copy(src.begin(), src.end(), poly.begin());
  size_t count = 4;
  auto cross = [](P a, P b, P c) { return (b.x - a.x) * (c.y - a.y) - (b.y - a.y) * (c.x - a.x); };
  for (int edge = 0; edge < 4 && count; ++edge) {
    const P p = clip[edge], q = clip[(edge + 1) % 4]; size_t out = 0;
    for (size_t i = 0; i < count; ++i) {
      const P u = poly[i], v = poly[(i + 1) % count];
      const double du = cross(p, q, u), dv = cross(p, q, v);
      if (du >= 0) next[out++] = u;
      if ((du >= 0) != (dv >= 0)) { const double t = du / (du - dv); next[out++] = {u.x + t * (v.x - u.x), u.y + t * (v.y - u.y)}; }
    }
    poly = next; count = out;
  }
  double area = 0;
  for (size_t i = 0; i < count; ++i) area += poly[i].x * poly[(i + 1) % count].y - poly[(i + 1) % count].x * poly[i].y;
  return std::abs(area) * .5 / (double(lv.w) * lv.h);
}

// ---------------------------------------------------------------- find
struct Trace {
  std::vector<Match> coarse;                   // after coarse NMS (emission order)
  std::vector<std::vector<Match>> tracked;     // after each tracking level (merged, filtered)
  std::vector<Match> prepared;                 // full-scored and sorted
  std::vector<Match> refined;                  // per prepared candidate, after refine
  std::vector<int> accepted_flag;              // refine verdict per prepared candidate (-1: not refined)
  double coarse_windows = 0, coarse_samples = 0, track_windows = 0, track_samples = 0;
  int top = 0, finest = 0, angle_stride = 1;
  long crop_r = 0, crop_c = 0, store_r = 0, store_c = 0, store_h = 0, store_w = 0;
};

// ------------------------------------------------------------------ domain
// Region: canonical run-length encoding (region.hpp:32-117, region.cpp:64-117):
// runs sorted by (row, col_begin), disjoint and non-adjacent within a row.
struct DRun { int64_t row, cb, ce; };
struct Domain {
  std::vector<DRun> runs;
  int64_t area = 0, rmin = 0, cmin = 0, rmax = -1, cmax = -1;
  bool empty() const { return runs.empty(); }
  // Region::contains (region.cpp:937-945): lexicographic (row, col_end) search
  bool contains(int64_t r, int64_t c) const {
    auto it = std::lower_bound(runs.begin(), runs.end(), std::make_pair(r, c), [](const DRun& x, std::pair<int64_t, int64_t> t) {
      return x.row < t.first || (x.row == t.first && x.ce < t.second);
    });
    return it != runs.end() && it->row == r && it->cb <= c;
  }
};

// Region::from_runs (region.cpp:64-117): sort, merge overlapping or adjacent
// runs of a row, area and bounding box.
static Domain make_domain(std::vector<DRun> in) {
  for (const auto& x : in) if (x.cb > x.ce) fail(INVALID_ARGUMENT, "run begin column must not exceed end column");
  std::sort(in.begin(), in.end(), [](const DRun& a, const DRun& b) {
    if (a.row != b.row) return a.row < b.row;
    if (a.cb != b.cb) return a.cb < b.cb;
    return a.ce < b.ce;
  });
  Domain d;
  if (in.empty()) return d;
  DRun cur = in.front();
  auto emit = [&](const DRun& x) { d.runs.push_back(x); d.area += x.ce - x.cb + 1; };
  for (size_t i = 1; i < in.size(); ++i) {
    const DRun& x = in[i];
    if (cur.row == x.row && (x.cb <= cur.ce || x.cb == cur.ce + 1)) cur.ce = std::max(cur.ce, x.ce);
    else { emit(cur); cur = x; }
  }
  emit(cur);
  d.rmin = d.runs.front().row; d.rmax = d.runs.back().row;
  d.cmin = d.runs.front().cb; d.cmax = d.runs.front().ce;
  for (const auto& x : d.runs) { d.cmin = std::min(d.cmin, x.cb); d.cmax = std::max(d.cmax, x.ce); }
  return d;
}

// Image::reduce_domain on a full-domain image (image.cpp:267-320): the
// region clipped to the image rectangle.
static Domain clip_domain(const Domain& d, int64_t h, int64_t w) {
  std::vector<DRun> out;
  for (const auto& x : d.runs)
    if (x.row >= 0 && x.row < h && x.ce >= 0 && x.cb < w) out.push_back({x.row, std::max<int64_t>(x.cb, 0), std::min<int64_t>(x.ce, w - 1)});
  return make_domain(std::move(out));
}

struct Query {
  const Model* model = nullptr;
  FindOpts opt;
  size_t level_count = 0, finest = 0;
  Patch coarse_original;
  std::vector<Gray> pyr;
  std::vector<std::vector<const Patch*>> qpatches;
  std::vector<std::vector<Patch>> endpoints;
  std::vector<Match> cand;
  long crop_r = 0, crop_c = 0;
  size_t img_w = 0, img_h = 0;
  bool empty = false;
  // the image's domain (image.hpp:108-113): null for the full image
  const Domain* dom = nullptr;
  bool sparse = false;                      // area < 0.7 bbox area (matcher.cpp:505)
  int64_t box_rmin = 0, box_cmin = 0, box_rmax = -1, box_cmax = -1;  // search box (domain bounding box)
  // find_models' query cache (matcher.cpp:281-301, 1078-1081): `cache` is set
  // for every query of find_models; `cache_samples` when it holds more than one
  // model, which converts the whole level 0 to doubles (matcher.cpp:596-611).
  bool cache = false, cache_samples = false;
  bool u8 = true;                           // byte image: exact_byte_statistics (matcher.cpp:347)
  bool shared_full = false;                 // level 0 is the whole crop (shared_values)
  double guard_first = -1, guard = -1;      // background guard thresholds (matcher.cpp:912-927)
  Prefix fine_prefix;                       // row prefix statistics of the whole level 0 (built on first use)
};

static bool domain_contains(const Query& q, const Match& m) {  // matcher.cpp:249-251
  const long r = long(std::floor(m.row + .5)), c = long(std::floor(m.col + .5));
  if (q.dom) return q.dom->contains(r, c);
  return r >= 0 && c >= 0 && r < long(q.img_h) && c < long(q.img_w);
}

// may_contain: matcher.cpp:506-534 (the run scan; the cached interval lists
// of a query cache hold the same runs)
static bool may_contain(const Query& q, const Match& pose, double spacing, double span = 0) {
  if (spacing == 1 && span == 0) return domain_contains(q, pose);
  const double margin = spacing == 1 ? .5 : 2 * spacing;
  if (pose.row < q.box_rmin - margin || pose.row > q.box_rmax + margin || pose.col + span < q.box_cmin - margin ||
      pose.col > q.box_cmax + margin)
    return false;
  if (!q.sparse) return true;
  const auto& runs = q.dom->runs;
  auto it = std::lower_bound(runs.begin(), runs.end(), pose.row - margin, [](const DRun& x, double row) { return double(x.row) < row; });
  for (; it != runs.end() && double(it->row) <= pose.row + margin; ++it)
    if (double(it->cb) <= pose.col + span + margin && double(it->ce) >= pose.col - margin) return true;
  return false;
}

// Pixel (r, c) of a u8 or u16 image with `stride` bytes per row.
static double pixel(const uint8_t* img, size_t stride, int bytes, size_t r, size_t c) {
  if (bytes == 1) return double(img[r * stride + c]);
  uint16_t v;
  std::memcpy(&v, img + r * stride + 2 * c, 2);
  return double(v);
}

static void prepare(Query& Q, const uint8_t* img, size_t W, size_t H, size_t stride, const Model& model, const FindOpts& req, Trace* tr,
                    int pixel_bytes = 1, const Domain* dom = nullptr) {
  FindOpts o = req;
  Q.model = &model; Q.opt = o; Q.img_w = W; Q.img_h = H;
  Q.u8 = pixel_bytes == 1;
  Q.dom = dom;
  if (!W || !H) fail(INVALID_ARGUMENT, "one nonempty single-channel image required");
  if (model.levels.empty()) fail(INVALID_STATE, "model is empty");
  if (o.highest < 0 || o.lowest < 1) fail(INVALID_ARGUMENT, "invalid pyramid levels");
  const size_t level_count = o.highest ? std::min(model.levels.size(), size_t(o.highest)) : model.levels.size();
  o.lowest = int(std::min(size_t(o.lowest), level_count));
  const size_t finest = size_t(o.lowest - 1);
  Q.level_count = level_count; Q.finest = finest;
  if (finest) {  // matcher.cpp:356-361
    Q.coarse_original = rotate(model.levels[finest], 0);
    for (const auto& s : Q.coarse_original.s) {
      Q.coarse_original.relr.push_back(s.r - model.levels[finest].orow);
      Q.coarse_original.relc.push_back(s.c - model.levels[finest].ocol);
    }
  }
  if (!std::isfinite(o.start) || !std::isfinite(o.extent) || o.extent < 0 || !std::isfinite(o.min_score) ||
      o.min_score < 0 || o.min_score > 1 || !std::isfinite(o.max_overlap) || o.max_overlap < 0 || o.max_overlap > 1 ||
      !o.max_candidates)
    fail(INVALID_ARGUMENT, "invalid query range, score, overlap or candidate budget");
  const double angle_end = o.start + o.extent;
  if (!std::isfinite(angle_end)) fail(INVALID_ARGUMENT, "query angle range must be finite");
  const double cs = std::max(o.start, model.opt.start);
  const double ce = std::min(angle_end, model.opt.start + model.opt.extent);
  if (cs > ce + 1e-12) { Q.empty = true; Q.opt = o; return; }
  o.start = cs; o.extent = std::max(0., ce - cs);
  Q.opt = o;
  if (W > size_t(std::numeric_limits<int>::max()) || H > size_t(std::numeric_limits<int>::max()) || H > 32000000 / W)
    fail(LIMIT_EXCEEDED, "search image exceeds pixel budget");
  // search box: the domain's bounding box (matcher.cpp:376-383)
  if (dom) {
    if (dom->empty()) { Q.empty = true; return; }
    Q.box_rmin = dom->rmin; Q.box_cmin = dom->cmin; Q.box_rmax = dom->rmax; Q.box_cmax = dom->cmax;
    Q.sparse = double(dom->area) < .7 * double(dom->rmax - dom->rmin + 1) * double(dom->cmax - dom->cmin + 1);
  } else {
    Q.box_rmin = 0; Q.box_cmin = 0; Q.box_rmax = int64_t(H) - 1; Q.box_cmax = int64_t(W) - 1;
  }
  // crop + halo: matcher.cpp:383-393
  const Level& base = model.levels.front();
  const double radius = std::hypot(std::max(base.orow, double(base.h - 1) - base.orow),
                                   std::max(base.ocol, double(base.w - 1) - base.ocol));
  const long lattice = long(size_t(1) << (level_count - 1));
  const long halo = long(std::ceil(radius)) + 4 * lattice;
  const long crop_r = std::max<long>(0, long(Q.box_rmin) - halo) / lattice * lattice;
  const long crop_c = std::max<long>(0, long(Q.box_cmin) - halo) / lattice * lattice;
  const long crop_b = std::min<long>(long(H), ((long(Q.box_rmax) + halo) / lattice + 1) * lattice);
  const long crop_rt = std::min<long>(long(W), ((long(Q.box_cmax) + halo) / lattice + 1) * lattice);
  Q.crop_r = crop_r; Q.crop_c = crop_c;
  // search pyramid: matcher.cpp:401-455
  auto& pyr = Q.pyr;
  {
    Gray g0; g0.w = size_t(crop_rt - crop_c); g0.h = size_t(crop_b - crop_r);
    g0.dc = pixel(img, stride, pixel_bytes, size_t(crop_r), size_t(crop_c));
    g0.sw = g0.w; g0.sh = g0.h;
    if (Q.u8) { g0.bytes = img + crop_r * stride + crop_c; g0.stride = stride; }
    else {  // u16: the whole crop as doubles, pixel - dc (matcher.cpp:415-433)
      g0.v.resize(g0.w * g0.h);
      for (size_t r = 0; r < g0.h; ++r)
        for (size_t c = 0; c < g0.w; ++c)
          g0.v[r * g0.w + c] = pixel(img, stride, 2, size_t(crop_r) + r, size_t(crop_c) + c) - g0.dc;
    }
    pyr.push_back(std::move(g0));
  }
  for (size_t l = 1; l < level_count; ++l) {
    const Gray& pv = pyr.back();
    Gray nx; nx.w = std::max<size_t>(1, pv.w / 2); nx.h = std::max<size_t>(1, pv.h / 2);
    nx.sw = nx.w; nx.sh = nx.h; nx.v.resize(nx.w * nx.h);
    for (size_t r = 0; r < nx.h; ++r)
      for (size_t c = 0; c < nx.w; ++c) {
        const size_t r0 = 2 * r, r1 = std::min(r0 + 1, pv.h - 1), c0 = 2 * c, c1 = std::min(c0 + 1, pv.w - 1);
        if (pv.bytes) nx.v[r * nx.w + c] = .25 * (pv.at(r0, c0) + pv.at(r0, c1) + pv.at(r1, c0) + pv.at(r1, c1));
        else {
          const double* row0 = pv.v.data() + r0 * pv.w; const double* row1 = pv.v.data() + r1 * pv.w;
          nx.v[r * nx.w + c] = .25 * (row0[c0] + row0[c1] + row1[c0] + row1[c1]);
        }
      }
    pyr.push_back(std::move(nx));
  }
  // query patches: matcher.cpp:463-475
  Q.qpatches.resize(level_count); Q.endpoints.resize(level_count);
  for (size_t li = 0; li < level_count; ++li) {
    const Level& lv = model.levels[li];
    auto& qp = Q.qpatches[li]; auto& ep = Q.endpoints[li]; ep.reserve(2);
    for (const auto& p : lv.patches) if (p.angle >= o.start - 1e-12 && p.angle <= angle_end + 1e-12) qp.push_back(&p);
    for (const double a : {o.start, angle_end}) {
      bool found = false;
      for (const auto* p : qp) if (std::abs(p->angle - a) < 1e-12) found = true;
      if (found) continue;
      ep.push_back(rotate(lv, a));
      if (!ep.back().s.empty()) qp.push_back(&ep.back());
    }
  }
  const size_t top = level_count - 1;
  const Gray& coarse = pyr[top];
  const double coarse_radius = .5 * std::hypot(double(model.levels[top].w), double(model.levels[top].h));
  const size_t stride_a = std::max<size_t>(1, size_t(std::min(double(Q.qpatches[top].size()),
                                                               std::ceil(.5 / std::max(1e-20, coarse_radius * model.opt.step)))));
  auto& ca = Q.qpatches[top];
  std::sort(ca.begin(), ca.end(), [](const Patch* a, const Patch* b) { return a->angle < b->angle; });
  { size_t kept = 0; for (size_t i = 0; i < ca.size(); ++i) if (i % stride_a == 0 || i + 1 == ca.size()) ca[kept++] = ca[i]; ca.resize(kept); }
  if (tr) { tr->top = int(top); tr->finest = int(finest); tr->angle_stride = int(stride_a); tr->crop_r = crop_r; tr->crop_c = crop_c; }
  // row prefixes of the top level: byte images only (matcher.cpp:489)
  Prefix top_prefix; const bool have_top_prefix = Q.u8 && !coarse.bytes;
  if (have_top_prefix) top_prefix = build_prefix(coarse);
  // support integral: matcher.cpp:498-503
  const size_t pitch = coarse.w + 1;
  std::vector<uint32_t> support(pitch * (coarse.h + 1), 0);
  for (size_t r = 0; r < coarse.h; ++r) {
    uint32_t rs = 0;
    for (size_t c = 0; c < coarse.w; ++c) { rs += coarse.at(r, c) != 0; support[(r + 1) * pitch + c + 1] = support[r * pitch + c + 1] + rs; }
  }
  auto has_support = [&](const Patch& p, int r, int c, int extra) {
    const size_t r0 = r + p.rmin, r1 = r + p.rmax + 1, c0 = c + p.cmin, c1 = c + p.cmax + 1 + extra;
    return support[r1 * pitch + c1] - support[r1 * pitch + c0] - support[r0 * pitch + c1] + support[r0 * pitch + c0] != 0;
  };
  const double scale = std::ldexp(1., int(top)), phase = (scale - 1) * .5;
  const bool absolute = model.opt.absolute;
  const bool exact_top = Q.u8 && top <= 2;  // exact_byte_statistics && top<=2
  auto& cand = Q.cand;
  std::vector<double> scores(coarse.w * coarse.h);
  // coarse exhaustive NCC: matcher.cpp:542-594
  for (const Patch* pp : Q.qpatches[top]) {
    const Patch& p = *pp;
    std::fill(scores.begin(), scores.end(), kMissing);
    const double margin = scale == 1 ? .5 : 2 * scale;
    const long row_min_box = long(Q.box_rmin), row_max_box = long(Q.box_rmax), col_min_box = long(Q.box_cmin),
               col_max_box = long(Q.box_cmax);
    auto pose_of = [&](int r, int c) {
      return Match{crop_r + scale * (r + p.orow) + phase, crop_c + scale * (c + p.ocol) + phase, p.angle, 0};
    };
    const int rb = std::max({0, -p.rmin, int(std::ceil((row_min_box - margin - crop_r - phase) / scale - p.orow))});
    const int re = std::min(int(coarse.h) - p.rmax, int(std::floor((row_max_box + margin - crop_r - phase) / scale - p.orow)) + 1);
    const int cb = std::max({0, -p.cmin, int(std::ceil((col_min_box - margin - crop_c - phase) / scale - p.ocol))});
    const int ce2 = std::min(int(coarse.w) - p.cmax, int(std::floor((col_max_box + margin - crop_c - phase) / scale - p.ocol)) + 1);
    for (int r = rb; r < re; ++r) {
      int c = cb;
      if (have_top_prefix) {
        for (; c + 15 < ce2; c += 16) {
          if (!may_contain(Q, pose_of(r, c), scale, 15 * scale) || !has_support(p, r, c, 15)) continue;
          for (int l = 0; l < 16; ++l)
            scores[r * coarse.w + c + l] = prefix_window(coarse.data(), top_prefix, coarse.w, p, r, c + l, absolute, exact_top, DotOrder::Single);
          for (int l = 0; l < 16; ++l) if (!may_contain(Q, pose_of(r, c + l), scale)) scores[r * coarse.w + c + l] = kMissing;
          if (tr) { tr->coarse_windows += 16; tr->coarse_samples += 16.0 * p.s.size(); }
        }
      }
      if (have_top_prefix || !coarse.bytes) {
        // four windows: prefix form, or without row prefixes (u16 image) the
        // direct form (matcher.cpp:565-571)
        for (; c + 3 < ce2; c += 4) {
          if (!may_contain(Q, pose_of(r, c), scale, 3 * scale) || !has_support(p, r, c, 3)) continue;
          if (have_top_prefix)
            for (int l = 0; l < 4; ++l)
              scores[r * coarse.w + c + l] = prefix_window(coarse.data(), top_prefix, coarse.w, p, r, c + l, absolute, exact_top, DotOrder::Four);
          else direct_four(coarse.data(), coarse.w, p, r, c, absolute, scores.data() + r * coarse.w + c);
          for (int l = 0; l < 4; ++l) if (!may_contain(Q, pose_of(r, c + l), scale)) scores[r * coarse.w + c + l] = kMissing;
          if (tr) { tr->coarse_windows += 4; tr->coarse_samples += 4.0 * p.s.size(); }
        }
      }
      for (; c < ce2; ++c)
        if (may_contain(Q, pose_of(r, c), scale) && has_support(p, r, c, 0)) {
          scores[r * coarse.w + c] = discrete(coarse, p, r, c, absolute, false);
          if (tr) { tr->coarse_windows += 1; tr->coarse_samples += double(p.s.size()); }
        }
    }
    for (int r = rb; r < re; ++r)
      for (int c = cb; c < ce2; ++c) {
        const double v = scores[r * coarse.w + c];
        if (v < o.min_score * .4 || !std::isfinite(v)) continue;
        bool mx = true;
        for (int dr = -1; dr <= 1; ++dr)
          for (int dc = -1; dc <= 1; ++dc) {
            const long rr = long(r) + dr, cc = long(c) + dc;
            if (rr < 0 || cc < 0 || rr >= long(coarse.h) || cc >= long(coarse.w) || (!dr && !dc)) continue;
            const double other = scores[rr * coarse.w + cc];
            if (other > v || (other == v && (dr < 0 || (!dr && dc < 0)))) mx = false;
          }
        if (!mx) continue;
        if (cand.size() == o.max_candidates) fail(LIMIT_EXCEEDED, "coarse candidate budget exceeded");
        cand.push_back({crop_r + scale * (r + p.orow) + phase, crop_c + scale * (c + p.ocol) + phase, p.angle, v});
      }
  }
  if (tr) tr->coarse = cand;
  // find_models over several models: the whole level 0 as doubles (matcher.cpp:596-611)
  if (!finest && Q.cache_samples && pyr.front().bytes && !cand.empty()) {
    Gray& f = pyr.front();
    f.sr = 0; f.sc = 0; f.sh = f.h; f.sw = f.w;
    f.v.resize(f.sw * f.sh);
    for (size_t r = 0; r < f.sh; ++r)
      for (size_t c = 0; c < f.sw; ++c) f.v[r * f.sw + c] = double(f.bytes[r * f.stride + c]) - f.dc;
    f.bytes = nullptr;
    Q.shared_full = true;
    if (tr) { tr->store_r = 0; tr->store_c = 0; tr->store_h = long(f.sh); tr->store_w = long(f.sw); }
  }
  // level-0 candidate region to doubles: matcher.cpp:612-629
  if (!finest && pyr.front().bytes && !cand.empty()) {
    Gray& f = pyr.front();
    double rmin = double(f.h), cmin = double(f.w), rmax = 0, cmax = 0;
    for (const auto& p : cand) {
      rmin = std::min(rmin, p.row - crop_r); rmax = std::max(rmax, p.row - crop_r);
      cmin = std::min(cmin, p.col - crop_c); cmax = std::max(cmax, p.col - crop_c);
    }
    const double mg = radius + 4 * scale + 8;
    f.sr = size_t(std::max(0., std::floor(rmin - mg))); f.sc = size_t(std::max(0., std::floor(cmin - mg)));
    f.sh = size_t(std::min(double(f.h), std::ceil(rmax + mg) + 1)) - f.sr;
    f.sw = size_t(std::min(double(f.w), std::ceil(cmax + mg) + 1)) - f.sc;
    f.v.resize(f.sw * f.sh);
    for (size_t r = 0; r < f.sh; ++r)
      for (size_t c = 0; c < f.sw; ++c) f.v[r * f.sw + c] = double(f.bytes[(r + f.sr) * f.stride + f.sc + c]) - f.dc;
    f.bytes = nullptr;
    if (tr) { tr->store_r = long(f.sr); tr->store_c = long(f.sc); tr->store_h = long(f.sh); tr->store_w = long(f.sw); }
  }
  // coarse-to-fine tracking: matcher.cpp:630-761 (no query cache)
  for (size_t l = top; l > finest; --l) {
    if (l == 1 && o.subpixel) {
      // candidates outside the domain snap to the nearest domain pixel within
      // 4 rows and 4 columns (first run wins ties), else drop (matcher.cpp:631-645)
      for (auto& p : cand) {
        if (domain_contains(Q, p)) continue;
        if (!Q.dom) { p.score = kMissing; continue; }
        double best = std::numeric_limits<double>::infinity();
        Match nearest = p;
        const auto& runs = Q.dom->runs;
        auto it = std::lower_bound(runs.begin(), runs.end(), p.row - 4, [](const DRun& x, double row) { return double(x.row) < row; });
        for (; it != runs.end() && double(it->row) <= p.row + 4; ++it) {
          const double c = std::clamp(std::round(p.col), double(it->cb), double(it->ce));
          if (std::abs(c - p.col) > 4) continue;
          const double d = std::hypot(double(it->row) - p.row, c - p.col);
          if (d < best) { best = d; nearest.row = double(it->row); nearest.col = c; }
        }
        if (std::isfinite(best)) p = nearest; else p.score = kMissing;
      }
      merge_identical(cand);
      break;
    }
    const Gray& tg = pyr[l - 1];
    Prefix pf; bool have_pf = false;
    if (Q.u8 && !tg.bytes && tg.sw == tg.w && tg.sh == tg.h && cand.size() >= 4) { pf = build_prefix(tg); have_pf = true; }
    const bool exact_l = Q.u8 && l <= 3;
    const double step = std::ldexp(1., int(l - 1)), offset = (step - 1) * .5;
    for (auto& cd : cand) {
      Match best = cd; best.score = kMissing;
      for (const Patch* pp : Q.qpatches[l - 1]) {
        const Patch& p = *pp;
        if (std::abs(p.angle - cd.angle) > model.opt.step * (l == top ? double(stride_a) : 1.) * 1.01) continue;
        const int row = int(std::round((cd.row - crop_r - offset) / step - p.orow));
        const int col = int(std::round((cd.col - crop_c - offset) / step - p.ocol));
        auto pose_at = [&](int rr, int cc, double s) {
          return Match{crop_r + step * (rr + p.orow) + offset, crop_c + step * (cc + p.ocol) + offset, p.angle, s};
        };
        for (int dr = -2; dr <= 2; ++dr) {
          int dc = -2;
          if (have_pf && row + dr + p.rmin >= 0 && row + dr + p.rmax < int(tg.h) && col - 2 + p.cmin >= 0 &&
              col + 5 + p.cmax < int(tg.w)) {
            double vals[5];
            for (int ln = 0; ln < 5; ++ln)
              vals[ln] = prefix_window(tg.data(), pf, tg.w, p, row + dr, col - 2 + ln, absolute, exact_l, DotOrder::Single);
            for (int ln = 0; ln < 5; ++ln) {
              Match ps = pose_at(row + dr, col - 2 + ln, vals[ln]);
              if (may_contain(Q, ps, step) && ps.score > best.score) best = ps;
            }
            if (tr) { tr->track_windows += 5; tr->track_samples += 5.0 * p.s.size(); }
            dc = 3;
          } else if (!tg.bytes && row + dr + p.rmin >= int(tg.sr) && row + dr + p.rmax < int(tg.sr + tg.sh) &&
                     col - 2 + p.cmin >= int(tg.sc) && col + 1 + p.cmax < int(tg.sc + tg.sw)) {
            double vals[4];
            if (have_pf)
              for (int ln = 0; ln < 4; ++ln)
                vals[ln] = prefix_window(tg.data(), pf, tg.w, p, row + dr, col - 2 + ln, absolute, exact_l, DotOrder::Four);
            else
              direct_four(tg.data(), tg.sw, p, row + dr - int(tg.sr), col - 2 - int(tg.sc), absolute, vals);
            for (int ln = 0; ln < 4; ++ln) {
              Match ps = pose_at(row + dr, col - 2 + ln, vals[ln]);
              if (may_contain(Q, ps, step) && ps.score > best.score) best = ps;
            }
            if (tr) { tr->track_windows += 4; tr->track_samples += 4.0 * p.s.size(); }
            dc = 2;
          }
          for (; dc <= 2; ++dc) {
            Match ps = pose_at(row + dr, col + dc, 0);
            if (!may_contain(Q, ps, step)) continue;
            ps.score = discrete(tg, p, row + dr, col + dc, absolute, true);
            if (tr) { tr->track_windows += 1; tr->track_samples += double(p.s.size()); }
            if (ps.score > best.score) best = ps;
          }
        }
      }
      cd = best;
    }
    merge_identical(cand);
    cand.erase(std::remove_if(cand.begin(), cand.end(), [&](const Match& p) { return !std::isfinite(p.score) || p.score < o.min_score * .4; }), cand.end());
    if (tr) tr->tracked.push_back(cand);
  }
  // full score at the finest level: matcher.cpp:813-822
  const Patch& orig = finest ? Q.coarse_original : model.original;
  {
    const double sc = std::ldexp(1., int(finest)), ph = (sc - 1) * .5;
    for (auto& cd : cand)
      if (std::isfinite(cd.score)) {
        Match ps = cd; ps.row = (ps.row - crop_r - ph) / sc; ps.col = (ps.col - crop_c - ph) / sc;
        if (finest == 0 && pyr.front().bytes) {
          // no candidate ever reached level 0 storage: unreachable when candidates exist
          cd.score = kMissing;
        } else cd.score = continuous(pyr[finest], model.levels[finest], orig, ps, absolute);
      }
  }
  std::sort(cand.begin(), cand.end(), [](const Match& a, const Match& b) {
    if (a.score != b.score) return a.score > b.score;
    if (a.row != b.row) return a.row < b.row;
    if (a.col != b.col) return a.col < b.col;
    return a.angle < b.angle;
  });
  if (tr) tr->prepared = cand;
}

// refine_coarse_candidate: matcher.cpp:853-877
static bool refine_coarse(Query& Q, Match& cd) {
  const Model& model = *Q.model; const FindOpts& o = Q.opt;
  const Level& lv = model.levels[Q.finest]; const Patch& orig = Q.coarse_original;
  const double scale = std::ldexp(1., int(Q.finest)), phase = (scale - 1) * .5;
  const Gray& fine = Q.pyr[Q.finest];
  auto local = [&](Match p) { p.row = (p.row - Q.crop_r - phase) / scale; p.col = (p.col - Q.crop_c - phase) / scale; return p; };
  auto score = [&](const Match& p) { return continuous(fine, lv, orig, local(p), model.opt.absolute); };
  if (o.subpixel)
    for (int it = 0; it < 6; ++it) {
      double delta[3] = {};
      if (!refinement_step(fine, lv, orig, local(cd), model.opt.absolute, o.extent > 0, delta, nullptr, nullptr)) break;
      const double factor = std::min({1., .75 / std::max(1e-20, std::abs(delta[0])), .75 / std::max(1e-20, std::abs(delta[1])),
                                      o.extent > 0 ? model.opt.step / std::max(1e-20, std::abs(delta[2])) : 1.});
      bool accepted = false;
      for (int bt = 0; bt < 4; ++bt) {
        const double st = std::ldexp(factor, -bt); Match nx = cd;
        nx.row += scale * st * delta[0]; nx.col += scale * st * delta[1];
        nx.angle = std::clamp(nx.angle + st * delta[2], o.start, o.start + o.extent);
        if (!domain_contains(Q, nx)) continue;
        nx.score = score(nx);
        if (nx.score > cd.score) { cd = nx; accepted = true; break; }
      }
      if (!accepted) break;
    }
  return cd.score >= o.min_score && domain_contains(Q, cd);
}

// refine_candidate: matcher.cpp:878-1061.  Without the query cache (find):
// fit, then the recovery search when the fit stays below min_score.  With it
// (find_models): a candidate already below min_score is recovered first
// (matcher.cpp:984-985, 1052-1055), the fit may stop early under the
// background guard (matcher.cpp:912-927, 958-960, 977-978), and over several
// models the recovery windows use row prefix statistics of the whole level 0
// (matcher.cpp:991-1003, 1018-1045).  The memo tables of the cache
// (fitted poses, score planes, derivatives) only reuse values and are not
// restated.
static bool refine(Query& Q, Match& cd) {
  if (Q.finest) return refine_coarse(Q, cd);
  const Model& model = *Q.model; const FindOpts& o = Q.opt;
  const long crop_r = Q.crop_r, crop_c = Q.crop_c; const size_t top = Q.level_count - 1;
  const double angle_end = o.start + o.extent;
  const Patch& orig = model.original; const Level& lv0 = model.levels.front();
  const Gray& fine = Q.pyr.front();
  const bool absolute = model.opt.absolute;
  auto full_score = [&](Match p) { p.row -= crop_r; p.col -= crop_c; return continuous(fine, lv0, orig, p, absolute); };
  auto fit = [&](Match& c) {
    bool cached = false; double next_delta[3] = {};
    if (o.subpixel)
      for (int it = 0; it < 6; ++it) {
        Match local = c; local.row -= crop_r; local.col -= crop_c; double delta[3] = {};
        if (cached) std::copy(next_delta, next_delta + 3, delta);
        else if (!refinement_step(fine, lv0, orig, local, absolute, o.extent > 0, delta, nullptr, nullptr)) break;
        cached = false;
        const double factor = std::min({1., .75 / std::max(1e-20, std::abs(delta[0])), .75 / std::max(1e-20, std::abs(delta[1])),
                                        o.extent > 0 ? model.opt.step / std::max(1e-20, std::abs(delta[2])) : 1.});
        bool accepted = false;
        for (int bt = 0; bt < 4; ++bt) {
          const double st = std::ldexp(factor, -bt); Match im = c;
          im.row += st * delta[0]; im.col += st * delta[1];
          im.angle = std::clamp(im.angle + st * delta[2], o.start, angle_end);
          if (!domain_contains(Q, im)) continue;
          bool next_valid = false;
          const bool guard_score_only = (it == 0 && Q.guard_first >= 0 && c.score < o.min_score - Q.guard_first) ||
                                        (it == 1 && Q.guard >= 0 && c.score < o.min_score - Q.guard);
          if (bt == 0 && it + 1 < 6 && !guard_score_only) {
            Match nl = im; nl.row -= crop_r; nl.col -= crop_c; im.score = kMissing;
            next_valid = refinement_step(fine, lv0, orig, nl, absolute, o.extent > 0, next_delta, &im.score, &c.score);
            if (!std::isfinite(im.score)) im.score = full_score(im);
          } else im.score = full_score(im);
          if (im.score > c.score) { c = im; accepted = true; cached = next_valid; break; }
        }
        if (it == 1 && Q.guard >= 0 && c.score < o.min_score - Q.guard) break;
        if (it == 0 && Q.guard_first >= 0 && c.score < o.min_score - Q.guard_first) break;
        if (!accepted || (std::hypot(delta[0], delta[1]) < 1e-4 && std::abs(delta[2]) < 1e-5)) break;
      }
  };
  const Match seed = cd;
  const bool recover_first = Q.cache && o.subpixel && top > 0 && cd.score < o.min_score;
  if (!recover_first) fit(cd);
  if (o.subpixel && top > 0 && cd.score < o.min_score) {
    // recovery 5x5 level-0 search: matcher.cpp:986-1059
    const bool stats = Q.cache && !fine.bytes && Q.shared_full && fine.sw == fine.w && fine.sh == fine.h;
    if (stats && Q.fine_prefix.sum.empty()) Q.fine_prefix = build_prefix(fine);
    const Prefix& pf = Q.fine_prefix;
    Match best = cd;
    for (const Patch* pp : Q.qpatches.front()) {
      const Patch& p = *pp;
      if (std::abs(p.angle - seed.angle) > model.opt.step * 1.01) continue;
      const int r = int(std::round(seed.row - crop_r - p.orow)), c = int(std::round(seed.col - crop_c - p.ocol));
      for (int dr = -2; dr <= 2; ++dr) {
        int dc = -2;
        if (r + dr + p.rmin >= int(fine.sr) && r + dr + p.rmax < int(fine.sr + fine.sh) && c - 2 + p.cmin >= int(fine.sc) &&
            c + 1 + p.cmax < int(fine.sc + fine.sw)) {
          double vals[4];
          if (stats) four_prefix(fine.data(), pf, fine.sw, p, r + dr - int(fine.sr), c - 2 - int(fine.sc), absolute, vals);
          else direct_four(fine.data(), fine.sw, p, r + dr - int(fine.sr), c - 2 - int(fine.sc), absolute, vals);
          for (int ln = 0; ln < 4; ++ln) {
            Match ps{crop_r + r + dr + p.orow, crop_c + c - 2 + ln + p.ocol, p.angle, vals[ln]};
            if (domain_contains(Q, ps) && ps.score > best.score) best = ps;
          }
          dc = 2;
        }
        for (; dc <= 2; ++dc) {
          Match ps{crop_r + r + dr + p.orow, crop_c + c + dc + p.ocol, p.angle, 0};
          if (!domain_contains(Q, ps)) continue;
          if (stats && r + dr + p.rmin >= int(fine.sr) && r + dr + p.rmax < int(fine.sr + fine.sh) &&
              c + dc + p.cmin >= int(fine.sc) && c + dc + p.cmax < int(fine.sc + fine.sw))
            ps.score = one_prefix(fine.data(), pf, fine.sw, p, r + dr - int(fine.sr), c + dc - int(fine.sc), absolute);
          else
            ps.score = discrete(fine, p, r + dr, c + dc, absolute, true);
          if (ps.score > best.score) best = ps;
        }
      }
    }
    if (best.row != cd.row || best.col != cd.col || best.angle != cd.angle) {
      best.score = full_score(best); fit(best);
      if (best.score > cd.score) cd = best;
    } else if (recover_first) fit(cd);
  }
  return cd.score >= o.min_score && domain_contains(Q, cd);
}

// find: matcher.cpp:1063-1076
static std::vector<Match> find(const uint8_t* img, size_t W, size_t H, size_t stride, const Model& model, const FindOpts& o, Trace* tr,
                               int pixel_bytes = 1, const Domain* dom = nullptr) {
  Query Q; prepare(Q, img, W, H, stride, model, o, tr, pixel_bytes, dom);
  std::vector<Match> result;
  if (tr) { tr->refined.assign(Q.cand.size(), Match{}); tr->accepted_flag.assign(Q.cand.size(), -1); }
  const Level& lv0 = model.levels.front();
  for (size_t k = 0; k < Q.cand.size(); ++k) {
    Match cd = Q.cand[k];
    if (!std::isfinite(cd.score)) continue;
    bool dup = false;
    for (const auto& p : result) if (overlap(cd, p, lv0) > Q.opt.max_overlap + 1e-12) { dup = true; break; }
    if (dup) continue;
    const bool ok = refine(Q, cd);
    if (tr) { tr->refined[k] = cd; tr->accepted_flag[k] = ok; }
    if (!ok) continue;
    dup = false;
    for (const auto& p : result) if (overlap(cd, p, lv0) > Q.opt.max_overlap + 1e-12) { dup = true; break; }
    if (dup) continue;
    result.push_back(cd);
    if (Q.opt.num_matches && result.size() == Q.opt.num_matches) break;
  }
  std::sort(result.begin(), result.end(), [](const Match& a, const Match& b) {
    if (a.score != b.score) return a.score > b.score;
    if (a.row != b.row) return a.row < b.row;
    if (a.col != b.col) return a.col < b.col;
    return a.angle < b.angle;
  });
  return result;
}

// Refines one given candidate exactly as find() would (used to compare the
// device's speculative refinement with the oracle, candidate by candidate).
static bool refine_one(const uint8_t* img, size_t W, size_t H, size_t stride, const Model& model, const FindOpts& o,
                       Match& cd, int pixel_bytes = 1) {
  Query Q; prepare(Q, img, W, H, stride, model, o, nullptr, pixel_bytes);
  return refine(Q, cd);
}

// ------------------------------------------------------------ find_models
// src/matching/multi.hpp:14-119 and matcher.cpp:1077-1124.
struct Selection { int64_t num_matches = -1; double max_overlap = -1; };
struct MultiMatch { Match pose; size_t model = 0; };
struct MPoint { double x, y; };
using MBox = std::array<MPoint, 4>;
struct MBounds { double top, bottom, left, right; };

static MBounds enclosure(const MBox& b) {  // multi.hpp:19-23
  MBounds out{b[0].y, b[0].y, b[0].x, b[0].x};
  for (const auto& p : b) {
    out.top = std::min(out.top, p.y); out.bottom = std::max(out.bottom, p.y);
    out.left = std::min(out.left, p.x); out.right = std::max(out.right, p.x);
  }
  return out;
}
static bool disjoint(const MBounds& a, const MBounds& b) {  // multi.hpp:24-26
  return a.right < b.left || b.right < a.left || a.bottom < b.top || b.bottom < a.top;
}
// Template rectangle [-o-1/2, w-o-1/2] at a pose: multi.hpp:54-60 (scale 1).
static MBox pose_box(const Level& lv, const Match& p) {
  const MBounds b{-lv.orow - .5, double(lv.h) - lv.orow - .5, -lv.ocol - .5, double(lv.w) - lv.ocol - .5};
  const double ca = std::cos(p.angle), sa = std::sin(p.angle);
  MBox q{{{b.left, b.top}, {b.right, b.top}, {b.right, b.bottom}, {b.left, b.bottom}}};
  for (auto& v : q) { const double x = v.x, y = v.y; v = {p.col + ca * x + sa * y, p.row - sa * x + ca * y}; }
  return q;
}
static double poly_area(const MPoint* p, size_t n) {  // multi.hpp:61-63
  double a = 0;
  for (size_t i = 0; i < n; ++i) { const auto& u = p[i]; const auto& v = p[(i + 1) % n]; a += u.x * v.y - v.x * u.y; }
  return std::abs(a) * .5;
}
// Intersection over the smaller area, after translating both boxes by a's
// first corner: multi.hpp:64-81.
static double box_overlap(MBox a, MBox b) {
  const MPoint origin = a[0];
  for (auto& p : a) { p.x -= origin.x; p.y -= origin.y; }
  for (auto& p : b) { p.x -= origin.x; p.y -= origin.y; }
  const double den = std::min(poly_area(a.data(), 4), poly_area(b.data(), 4));
  if (den <= 1e-12) return 0;
  std::array<MPoint, 16> poly{}, next{};
  std::copy(a.begin(), a.end(), poly.begin());
  size_t count = 4;
  auto cross = [](MPoint u, MPoint v, MPoint p) { return (v.x - u.x) * (p.y - u.y) - (v.y - u.y) * (p.x - u.x); };
  for (int e = 0; e < 4 && count; ++e) {
    const MPoint u = b[size_t(e)], v = b[size_t((e + 1) % 4)];
    size_t m = 0;
    for (size_t i = 0; i < count; ++i) {
      const MPoint p = poly[i], q = poly[(i + 1) % count];
      const double dp = cross(u, v, p), dq = cross(u, v, q);
      if (dp >= 0) next[m++] = p;
      if ((dp >= 0) != (dq >= 0)) { const double t = dp / (dp - dq); next[m++] = {p.x + t * (q.x - p.x), p.y + t * (q.y - p.y)}; }
    }
    poly = next; count = m;
  }
  return std::clamp(poly_area(poly.data(), count) / den, 0., 1.);
}
static bool multi_less(const MultiMatch& a, const MultiMatch& b) {  // multi.hpp:87-93
  if (a.pose.score != b.pose.score) return a.pose.score > b.pose.score;
  if (a.model != b.model) return a.model < b.model;
  if (a.pose.row != b.pose.row) return a.pose.row < b.pose.row;
  if (a.pose.col != b.pose.col) return a.pose.col < b.pose.col;
  return a.pose.angle < b.pose.angle;
}
// select_refined: multi.hpp:83-119.  Overlap is tested before and after the
// refinement; a suppressed candidate does not use up its model's quota.
template <class Refine>
static std::vector<MultiMatch> select_refined(std::vector<MultiMatch> cand, const std::vector<FindOpts>& opts,
                                              const std::vector<const Model*>& models, Selection sel, Refine refine) {
  std::sort(cand.begin(), cand.end(), multi_less);
  std::vector<MultiMatch> result; std::vector<MBox> boxes; std::vector<MBounds> encl;
  std::vector<size_t> counts(opts.size());
  auto suppressed = [&](size_t index, const MBox& rect, const MBounds& aabb, double threshold) {
    for (size_t i = 0; i < result.size(); ++i) {
      if (sel.max_overlap < 0 && result[i].model != index) continue;
      if (disjoint(aabb, encl[i])) continue;
      if (box_overlap(rect, boxes[i]) > threshold + 1e-12) return true;
    }
    return false;
  };
  for (auto c : cand) {
    const size_t index = c.model;
    const int64_t limit = sel.num_matches >= 0 ? sel.num_matches : int64_t(opts[index].num_matches);
    if (sel.num_matches < 0 && limit && counts[index] >= uint64_t(limit)) continue;
    const double threshold = sel.max_overlap >= 0 ? sel.max_overlap : opts[index].max_overlap;
    const Level& lv = models[index]->levels.front();
    MBox rect = pose_box(lv, c.pose); MBounds aabb = enclosure(rect);
    if (suppressed(index, rect, aabb, threshold)) continue;
    if (!refine(c)) continue;
    rect = pose_box(lv, c.pose); aabb = enclosure(rect);
    if (suppressed(index, rect, aabb, threshold)) continue;
    result.push_back(c); boxes.push_back(rect); encl.push_back(aabb); ++counts[index];
    if (sel.num_matches > 0 && result.size() >= uint64_t(sel.num_matches)) break;
  }
  std::sort(result.begin(), result.end(), multi_less);
  return result;
}

static bool same_options(const FindOpts& a, const FindOpts& b) {  // matcher.cpp:1083-1086
  return a.start == b.start && a.extent == b.extent && a.min_score == b.min_score && a.max_overlap == b.max_overlap &&
         a.num_matches == b.num_matches && a.max_candidates == b.max_candidates && a.subpixel == b.subpixel &&
         a.highest == b.highest && a.lowest == b.lowest;
}

// find_models: matcher.cpp:1077-1124 with the C ABI's two-stage quota
// (per_model_quota_before_overlap = true, ncc_matching_api.cpp:199).
static std::vector<MultiMatch> find_models(const uint8_t* img, size_t W, size_t H, size_t stride,
                                           const std::vector<const Model*>& models, const std::vector<FindOpts>& opts,
                                           Selection sel, bool per_model_quota_before_overlap,
                                           std::vector<std::vector<Match>>* prepared = nullptr, int pixel_bytes = 1,
                                           const Domain* dom = nullptr) {
  const bool cache_samples = models.size() > 1, unlimited = sel.num_matches == 0;
  bool cache_derivatives = false;
  for (const Model* m : models) if (m != models.front()) cache_derivatives = true;
  std::vector<std::unique_ptr<Query>> queries;
  std::vector<size_t> owners;
  std::vector<MultiMatch> cand;
  if (prepared) prepared->assign(models.size(), {});
  for (size_t i = 0; i < models.size(); ++i) {
    size_t owner = i;
    for (size_t j = 0; j < i; ++j)
      if (models[i] == models[j] && same_options(opts[i], opts[j])) { owner = owners[j]; break; }
    owners.push_back(owner);
    if (owner != i) {
      for (const auto& p : queries[owner]->cand) if (std::isfinite(p.score)) cand.push_back({p, i});
      if (prepared) (*prepared)[i] = queries[owner]->cand;
      queries.push_back(nullptr);
      continue;
    }
    auto Q = std::make_unique<Query>();
    Q->cache = true; Q->cache_samples = cache_samples;
    prepare(*Q, img, W, H, stride, *models[i], opts[i], nullptr, pixel_bytes, dom);
    // default background guard: matcher.cpp:839-846 (AVX2 host, round to
    // nearest, byte image, no reduced domain)
    const Model& m = *models[i]; const FindOpts& o = Q->opt;
    if (!Q->empty && !Q->finest && cache_derivatives && unlimited && Q->cand.size() >= 32 && o.subpixel && !o.num_matches &&
        o.min_score >= .75 && o.min_score <= .85 && Q->level_count == 3 && !m.opt.absolute && m.opt.step <= .02 &&
        o.extent <= .4 && m.original.s.size() <= 32768 && Q->shared_full && !dom) {
      Q->guard_first = o.min_score >= .8 ? .1 : .2;
      Q->guard = .1;
    }
    for (const auto& p : Q->cand) if (std::isfinite(p.score)) cand.push_back({p, i});
    if (prepared) (*prepared)[i] = Q->cand;
    queries.push_back(std::move(Q));
  }
  auto refine_c = [&](MultiMatch& c) { return refine(*queries[owners[c.model]], c.pose); };
  bool any_quota = false;
  for (const auto& o : opts) if (o.num_matches > 0) any_quota = true;
  const bool independent = per_model_quota_before_overlap && sel.num_matches < 0 && sel.max_overlap >= 0 && any_quota;
  auto result = select_refined(std::move(cand), opts, models, independent ? Selection{-1, -1} : sel, refine_c);
  if (independent) result = select_refined(std::move(result), opts, models, sel, [](MultiMatch&) { return true; });
  return result;
}

}  // namespace orc

// ====================================================================== C ABI
// Flat entry points for the Python test harness (ctypes).
using namespace orc;

static thread_local std::string g_last_error;

template <class F> static int guard(F&& f) {
  try { f(); return 0; }
  catch (const Failure& e) { g_last_error = e.what(); return e.status; }
  catch (const std::exception& e) { g_last_error = e.what(); return -1000; }
}

extern "C" {

const char* orc_last_error() { return g_last_error.c_str(); }

// Creates a model from a u8 template; dom_runs (n_dom_runs x 3, canonical)
// is its reduced domain when reduced != 0.
int orc_create(const uint8_t* px, int w, int h, int stride, int levels, double start, double extent, double step,
               int absolute, const int64_t* dom_runs, int n_dom_runs, int reduced, void** out) {
  return guard([&] {
    CreateOpts o; o.levels = levels; o.absolute = absolute != 0; o.start = start; o.extent = extent; o.step = step;
    std::vector<std::array<int64_t, 3>> runs;
    for (int i = 0; i < n_dom_runs; ++i) runs.push_back({dom_runs[3 * i], dom_runs[3 * i + 1], dom_runs[3 * i + 2]});
    if (reduced && runs.empty()) fail(INVALID_ARGUMENT, "template domain requires at least four pixels");
    *out = new Model(create(px, size_t(w), size_t(h), size_t(stride), o, std::move(runs)));
  });
}
void orc_free(void* m) { delete static_cast<Model*>(m); }

// info[0]=levels, [1]=angles per level, [2]=step, [3]=origin_row, [4]=origin_col, [5]=sample_count
void orc_model_info(void* mp, double* info) {
  const Model& m = *static_cast<Model*>(mp);
  info[0] = double(m.levels.size()); info[1] = double(m.levels.front().patches.size()); info[2] = m.opt.step;
  info[3] = m.orow; info[4] = m.ocol; info[5] = double(m.sample_count);
}
// level geometry: [w, h, orow, ocol]
void orc_level_info(void* mp, int level, double* out) {
  const Level& lv = static_cast<Model*>(mp)->levels[size_t(level)];
  out[0] = double(lv.w); out[1] = double(lv.h); out[2] = lv.orow; out[3] = lv.ocol;
}
// patch header: [n, angle, orow, ocol, rmin, rmax, cmin, cmax, wsum, n_runs, n_wruns]; level -1: original
void orc_patch_info(void* mp, int level, int angle, double* out) {
  const Model& m = *static_cast<Model*>(mp);
  const Patch& p = level < 0 ? m.original : m.levels[size_t(level)].patches[size_t(angle)];
  out[0] = double(p.s.size()); out[1] = p.angle; out[2] = p.orow; out[3] = p.ocol;
  out[4] = p.rmin; out[5] = p.rmax; out[6] = p.cmin; out[7] = p.cmax; out[8] = p.wsum;
  out[9] = double(p.runs.size()); out[10] = double(p.wruns.size());
}
void orc_patch_samples(void* mp, int level, int angle, int32_t* rows, int32_t* cols, double* w) {
  const Model& m = *static_cast<Model*>(mp);
  const Patch& p = level < 0 ? m.original : m.levels[size_t(level)].patches[size_t(angle)];
  for (size_t i = 0; i < p.s.size(); ++i) { rows[i] = p.s[i].r; cols[i] = p.s[i].c; w[i] = p.s[i].w; }
}
void orc_level_gray(void* mp, int level, double* gray) {
  const Level& lv = static_cast<Model*>(mp)->levels[size_t(level)];
  std::copy(lv.gray.begin(), lv.gray.end(), gray);
}

// MVKIT_NCC 1 text of a model (length in *len; bytes copied when cap allows).
int orc_serialize(void* mp, char* buf, size_t cap, size_t* len) {
  return guard([&] {
    const std::string t = serialize(*static_cast<Model*>(mp));
    *len = t.size();
    if (buf && cap >= t.size()) std::memcpy(buf, t.data(), t.size());
  });
}
uint64_t orc_checksum(const char* bytes, size_t len) { return fnv1a64(std::string(bytes, len)); }
int orc_deserialize(const char* bytes, size_t len, void** out) {
  return guard([&] { *out = new Model(deserialize(std::string(bytes, len))); });
}

// Search pyramid level values (S_l/4^l - dc, doubles) for a full-domain u8 image.
int orc_pyramid(const uint8_t* img, int w, int h, int stride, int levels, double* out) {
  return guard([&] {
    Gray g0; g0.w = size_t(w); g0.h = size_t(h); g0.bytes = img; g0.stride = size_t(stride); g0.dc = double(img[0]);
    std::vector<Gray> pyr; pyr.push_back(g0);
    size_t off = 0;
    for (int l = 1; l < levels; ++l) {
      const Gray& pv = pyr.back();
      Gray nx; nx.w = std::max<size_t>(1, pv.w / 2); nx.h = std::max<size_t>(1, pv.h / 2); nx.v.resize(nx.w * nx.h);
      for (size_t r = 0; r < nx.h; ++r)
        for (size_t c = 0; c < nx.w; ++c) {
          const size_t r0 = 2 * r, r1 = std::min(r0 + 1, pv.h - 1), c0 = 2 * c, c1 = std::min(c0 + 1, pv.w - 1);
          nx.v[r * nx.w + c] = pv.bytes ? .25 * (pv.at(r0, c0) + pv.at(r0, c1) + pv.at(r1, c0) + pv.at(r1, c1))
                                        : .25 * (pv.v[r0 * pv.w + c0] + pv.v[r0 * pv.w + c1] + pv.v[r1 * pv.w + c0] + pv.v[r1 * pv.w + c1]);
        }
      std::copy(nx.v.begin(), nx.v.end(), out + off); off += nx.v.size();
      pyr.push_back(std::move(nx));
    }
  });
}

// Row prefix sums of a double level (layout (w+1) x h) -- matcher.cpp:487-495.
void orc_row_prefix(const double* v, int w, int h, double* psum, double* psq) {
  Gray g; g.w = size_t(w); g.h = size_t(h); g.v.assign(v, v + size_t(w) * h);
  Prefix p = build_prefix(g);
  std::copy(p.sum.begin(), p.sum.end(), psum); std::copy(p.sq.begin(), p.sq.end(), psq);
}

struct OrcFindOut {
  int n_result, n_coarse, n_prepared, top, finest, angle_stride;
  double coarse_windows, coarse_samples, track_windows, track_samples;
  long crop_r, crop_c, store_r, store_c, store_h, store_w;
};

// The domain of an image reduced to the given runs (n x 3: row, first
// column, last column; any order, overlaps allowed), clipped to the image.
static Domain image_domain(const int64_t* runs, int n, int h, int w) {
  std::vector<DRun> v;
  for (int i = 0; i < n; ++i) v.push_back({runs[3 * i], runs[3 * i + 1], runs[3 * i + 2]});
  return clip_domain(make_domain(std::move(v)), h, w);
}

// The canonical runs of image_domain (out: cap x 3); returns their number.
int orc_domain_runs(const int64_t* runs, int n, int h, int w, int64_t* out, int cap) {
  int count = -1;
  const int st = guard([&] {
    const Domain d = image_domain(runs, n, h, w);
    for (size_t i = 0; i < d.runs.size() && int(i) < cap; ++i) {
      out[3 * i] = d.runs[i].row; out[3 * i + 1] = d.runs[i].cb; out[3 * i + 2] = d.runs[i].ce;
    }
    count = int(d.runs.size());
  });
  return st ? st : count;
}

// Runs find on a u8 or u16 image (pixel_bytes 1 or 2, stride in bytes),
// optionally with its domain reduced to dom_runs (reduced != 0).
// results: cap x 4 (row, col, angle, score) in the reference's
// internal convention (centroid coordinates, no output shift).  Optional
// traces: coarse (cap_c x 4), prepared (cap_p x 4), refined (cap_p x 5 with
// the verdict in column 4).
int orc_find(void* mp, const uint8_t* img, int w, int h, int stride, int pixel_bytes, const int64_t* dom_runs,
             int n_dom_runs, int reduced, double start, double extent, double min_score, int num_matches,
             double max_overlap, int subpixel, int highest, int lowest, double* results, int cap, double* coarse,
             int cap_c, double* prepared, int cap_p, double* refined, OrcFindOut* info) {
  return guard([&] {
    FindOpts o; o.start = start; o.extent = extent; o.min_score = min_score; o.num_matches = size_t(num_matches);
    o.max_overlap = max_overlap; o.subpixel = subpixel != 0; o.highest = highest; o.lowest = lowest;
    Trace tr;
    if (pixel_bytes != 1 && pixel_bytes != 2) fail(INVALID_ARGUMENT, "pixel type must be byte or uint2");
    const Domain dom = reduced ? image_domain(dom_runs, n_dom_runs, h, w) : Domain{};
    auto res = find(img, size_t(w), size_t(h), size_t(stride), *static_cast<Model*>(mp), o, &tr, pixel_bytes,
                    reduced ? &dom : nullptr);
    auto put = [](const Match& m, double* d) { d[0] = m.row; d[1] = m.col; d[2] = m.angle; d[3] = m.score; };
    for (size_t i = 0; i < res.size() && int(i) < cap; ++i) put(res[i], results + 4 * i);
    if (coarse) for (size_t i = 0; i < tr.coarse.size() && int(i) < cap_c; ++i) put(tr.coarse[i], coarse + 4 * i);
    if (prepared) for (size_t i = 0; i < tr.prepared.size() && int(i) < cap_p; ++i) put(tr.prepared[i], prepared + 4 * i);
    if (refined)
      for (size_t i = 0; i < tr.refined.size() && int(i) < cap_p; ++i) { put(tr.refined[i], refined + 5 * i); refined[5 * i + 4] = tr.accepted_flag[i]; }
    if (info) {
      info->n_result = int(res.size()); info->n_coarse = int(tr.coarse.size()); info->n_prepared = int(tr.prepared.size());
      info->top = tr.top; info->finest = tr.finest; info->angle_stride = tr.angle_stride;
      info->coarse_windows = tr.coarse_windows; info->coarse_samples = tr.coarse_samples;
      info->track_windows = tr.track_windows; info->track_samples = tr.track_samples;
      info->crop_r = tr.crop_r; info->crop_c = tr.crop_c; info->store_r = tr.store_r; info->store_c = tr.store_c;
      info->store_h = tr.store_h; info->store_w = tr.store_w;
    }
  });
}

// The tracked candidate lists of a find on a u8 image (full domain), one per
// tracking level in tracking order, each in merge_identical's order
// (row, col, angle), concatenated into out (cap x 4); counts[k] is the size
// of list k (at most max_lists lists; *n_lists receives their number).
int orc_find_tracked(void* mp, const uint8_t* img, int w, int h, int stride, double start, double extent,
                     double min_score, int num_matches, double max_overlap, int subpixel, int highest, int lowest,
                     double* out, int cap, int* counts, int max_lists, int* n_lists) {
  return guard([&] {
    FindOpts o; o.start = start; o.extent = extent; o.min_score = min_score; o.num_matches = size_t(num_matches);
    o.max_overlap = max_overlap; o.subpixel = subpixel != 0; o.highest = highest; o.lowest = lowest;
    Trace tr;
    find(img, size_t(w), size_t(h), size_t(stride), *static_cast<Model*>(mp), o, &tr, 1, nullptr);
    int k = 0;
    for (size_t l = 0; l < tr.tracked.size() && int(l) < max_lists; ++l) {
      counts[l] = int(tr.tracked[l].size());
      for (const Match& m : tr.tracked[l]) {
        if (k < cap) { out[4 * k] = m.row; out[4 * k + 1] = m.col; out[4 * k + 2] = m.angle; out[4 * k + 3] = m.score; }
        ++k;
      }
    }
    *n_lists = int(tr.tracked.size());
  });
}

// The candidate list after merge_identical (matcher.cpp:241-248): in (n x 4),
// out receives the merged list; returns its length.
int orc_merge_identical(const double* in, int n, double* out) {
  std::vector<Match> c(size_t(std::max(n, 0)));
  for (int i = 0; i < n; ++i) c[size_t(i)] = Match{in[4 * i], in[4 * i + 1], in[4 * i + 2], in[4 * i + 3]};
  merge_identical(c);
  for (size_t i = 0; i < c.size(); ++i) {
    out[4 * i] = c[i].row; out[4 * i + 1] = c[i].col; out[4 * i + 2] = c[i].angle; out[4 * i + 3] = c[i].score;
  }
  return int(c.size());
}

// Textbook ZNCC on two sample arrays (normalized_correlation, model.cpp:105-118).
double orc_zncc(const double* a, const double* b, int n, int absolute) {
  double ma = 0, mb = 0;
  for (int i = 0; i < n; ++i) { ma += (a[i] - ma) / double(i + 1); mb += (b[i] - mb) / double(i + 1); }
  double aa = 0, bb = 0, ab = 0;
  for (int i = 0; i < n; ++i) { const double x = a[i] - ma, y = b[i] - mb; aa += x * x; bb += y * y; ab += x * y; }
  if (!(aa > 0 && bb > 0)) return -1;
  const double s = std::clamp(ab / std::sqrt(aa) / std::sqrt(bb), -1., 1.);
  return absolute ? std::abs(s) : s;
}

// Overlap of two poses under a level-0 template geometry (matcher.cpp:252-277).
double orc_overlap(double ar, double ac, double aa, double br, double bc, double ba, int w, int h, double orow, double ocol) {
  Level lv; lv.w = size_t(w); lv.h = size_t(h); lv.orow = orow; lv.ocol = ocol;
  return overlap(Match{ar, ac, aa, 0}, Match{br, bc, ba, 0}, lv);
}

// Continuous bilinear NCC score of the model's original patch at a pose on a
// full u8 image (level 0, dense storage) -- used for first-order checks.
double orc_continuous(void* mp, const uint8_t* img, int w, int h, int stride, double row, double col, double angle) {
  const Model& m = *static_cast<Model*>(mp);
  Gray g; g.w = size_t(w); g.h = size_t(h); g.sw = g.w; g.sh = g.h; g.v.resize(g.w * g.h);
  const double dc = double(img[0]);
  for (size_t r = 0; r < g.h; ++r) for (size_t c = 0; c < g.w; ++c) g.v[r * g.w + c] = double(img[r * size_t(stride) + c]) - dc;
  return continuous(g, m.levels.front(), m.original, Match{row, col, angle, 0}, m.opt.absolute);
}

// find_models (domain as for orc_find).  opts: n_models x 9 doubles (start, extent, min_score,
// max_overlap, num_matches, subpixel, highest, lowest, max_candidates with 0 the default 10 000); selection as the C
// ABI derives it (num_matches -1 / max_overlap -1 mean per model).  results:
// cap x 5 (row, col, angle, score, model) in the internal convention.
// prepared (optional): n_models x cap_p x 4, with prepared_n the list sizes.
int orc_find_models(void* const* mps, int n_models, const uint8_t* img, int w, int h, int stride, int pixel_bytes,
                    const int64_t* dom_runs, int n_dom_runs, int reduced, const double* opts,
                    int64_t sel_num, double sel_overlap, int per_model_quota, double* results, int cap, int* n_result,
                    double* prepared, int cap_p, int* prepared_n) {
  return guard([&] {
    std::vector<const Model*> models;
    std::vector<FindOpts> o(static_cast<size_t>(n_models));
    for (int i = 0; i < n_models; ++i) {
      models.push_back(static_cast<const Model*>(mps[i]));
      const double* q = opts + 9 * i;
      FindOpts& f = o[size_t(i)];
      f.start = q[0]; f.extent = q[1]; f.min_score = q[2]; f.max_overlap = q[3]; f.num_matches = size_t(q[4]);
      f.subpixel = q[5] != 0; f.highest = int(q[6]); f.lowest = int(q[7]);
      if (q[8] > 0) f.max_candidates = size_t(q[8]);  // the reference's FindOptions::max_candidates (model.hpp:21)
    }
    std::vector<std::vector<Match>> prep;
    const Domain dom = reduced ? image_domain(dom_runs, n_dom_runs, h, w) : Domain{};
    auto res = find_models(img, size_t(w), size_t(h), size_t(stride), models, o, Selection{sel_num, sel_overlap},
                           per_model_quota != 0, prepared ? &prep : nullptr, pixel_bytes, reduced ? &dom : nullptr);
    *n_result = int(res.size());
    for (size_t i = 0; i < res.size() && int(i) < cap; ++i) {
      double* d = results + 5 * i;
      d[0] = res[i].pose.row; d[1] = res[i].pose.col; d[2] = res[i].pose.angle; d[3] = res[i].pose.score;
      d[4] = double(res[i].model);
    }
    if (prepared)
      for (int m = 0; m < n_models; ++m) {
        prepared_n[m] = int(prep[size_t(m)].size());
        for (size_t i = 0; i < prep[size_t(m)].size() && int(i) < cap_p; ++i) {
          double* d = prepared + (size_t(m) * size_t(cap_p) + i) * 4;
          const Match& x = prep[size_t(m)][i];
          d[0] = x.row; d[1] = x.col; d[2] = x.angle; d[3] = x.score;
        }
      }
  });
}

// Overlap of two template rectangles under multi-model rules (multi.hpp:54-81):
// geometry per pose [w, h, orow, ocol].
double orc_box_overlap(double ar, double ac, double aa, const double* ga, double br, double bc, double ba, const double* gb) {
  Level la; la.w = size_t(ga[0]); la.h = size_t(ga[1]); la.orow = ga[2]; la.ocol = ga[3];
  Level lb; lb.w = size_t(gb[0]); lb.h = size_t(gb[1]); lb.orow = gb[2]; lb.ocol = gb[3];
  return box_overlap(pose_box(la, Match{ar, ac, aa, 0}), pose_box(lb, Match{br, bc, ba, 0}));
}

// Refine a single candidate (row, col, angle, score in/out); returns verdict.
int orc_refine_one(void* mp, const uint8_t* img, int w, int h, int stride, double start, double extent,
                   double min_score, int num_matches, double max_overlap, int subpixel, double* pose) {
  int verdict = 0;
  const int st = guard([&] {
    FindOpts o; o.start = start; o.extent = extent; o.min_score = min_score; o.num_matches = size_t(num_matches);
    o.max_overlap = max_overlap; o.subpixel = subpixel != 0;
    Match cd{pose[0], pose[1], pose[2], pose[3]};
    verdict = refine_one(img, size_t(w), size_t(h), size_t(stride), *static_cast<Model*>(mp), o, cd) ? 1 : 0;
    pose[0] = cd.row; pose[1] = cd.col; pose[2] = cd.angle; pose[3] = cd.score;
  });
  return st ? st : verdict;
}

}  // extern "C"
