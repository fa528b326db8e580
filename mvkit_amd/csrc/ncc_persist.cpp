// Model persistence: the "MVKIT_NCC 1" text format of persist.cpp:14-84.
// Body: levels, polarity flag, angle start / extent / resolved step; the
// template centroid in image coordinates and level-0 size; the level-0
// centroid; then one "mask gray" line per level-0 pixel.  Numbers use the
// classic locale: integers (and reals that are integers within 2^53) as
// decimal integers, other reals with 17 significant digits
// (core/model_text.hpp:18-62).  The header carries the FNV-1a-64 checksum of
// the body.  Loading validates every field and the centroid, then rebuilds
// the pyramid and the rotated rasters exactly as create does.
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <sstream>
#include <string>
#include <string_view>

#include "ncc_model.hpp"

namespace mvamd {

namespace {
constexpr size_t kMaxFile = 64u * 1024u * 1024u;
constexpr size_t kMaxPixels = 1048576;
const char* const kPersistOp = "ncc_model_persistence";

[[noreturn]] void fail(Status s, const char* msg) { throw Error(s, kPersistOp, "path", msg); }

uint64_t fnv1a64(std::string_view text) {
  uint64_t h = 14695981039346656037ULL;
  for (unsigned char c : text) {
    h ^= c;
    h *= 1099511628211ULL;
  }
  return h;
}

void put(std::string& out, int64_t v) {
  char buf[32];
  const auto r = std::to_chars(buf, buf + sizeof buf, v);
  out.append(buf, r.ptr);
}
void put(std::string& out, double v) {
  if (v == 0.) {
    if (std::signbit(v)) out.push_back('-');
    out.push_back('0');
    return;
  }
  if (v >= -9007199254740992. && v <= 9007199254740992.) {
    const auto i = static_cast<int64_t>(v);
    if (static_cast<double>(i) == v) return put(out, i);
  }
  char buf[64];
  const auto r = std::to_chars(buf, buf + sizeof buf, v, std::chars_format::general, 17);
  out.append(buf, r.ptr);
}

bool space(char c) { return c == ' ' || c == '\n' || c == '\r' || c == '\t' || c == '\v' || c == '\f'; }

// Field lexer of core/model_text.hpp:113-189: skip ASCII whitespace, allow
// one leading '+', then the longest number; unusual real spellings
// (exponent without digits, inf/nan words, hex) go through a classic-locale
// stream.
class Fields {
 public:
  explicit Fields(std::string_view t) : t_(t) {}
  template <class T> T read() {
    skip();
    T v{};
    const std::string_view original = t_;
    if (t_.empty()) bad();
    const bool plus = t_.front() == '+';
    if (plus) t_.remove_prefix(1);
    if (t_.empty() || (plus && (t_.front() == '+' || t_.front() == '-'))) bad();
    const auto r = std::from_chars(t_.data(), t_.data() + t_.size(), v);
    bool fallback = r.ec != std::errc{};
    if constexpr (std::is_floating_point_v<T>) {
      fallback = fallback || (r.ptr != t_.data() + t_.size() &&
                              (*r.ptr == 'e' || *r.ptr == 'E' || *r.ptr == 'x' || *r.ptr == 'X'));
      fallback = fallback || t_.front() == 'i' || t_.front() == 'I' || t_.front() == 'n' || t_.front() == 'N';
    }
    if (!fallback) {
      t_.remove_prefix(size_t(r.ptr - t_.data()));
      return v;
    }
    std::istringstream legacy{std::string(original)};
    legacy.imbue(std::locale::classic());
    if (!(legacy >> v)) bad();
    const auto off = legacy.tellg();
    t_ = off < 0 ? std::string_view{} : original.substr(size_t(off));
    return v;
  }
  double real() {
    const double v = read<double>();
    if (!std::isfinite(v)) fail(kDecodeError, "non-finite model field");
    return v;
  }
  size_t count(size_t limit) {
    const int64_t v = read<int64_t>();
    if (v < 0 || uint64_t(v) > limit) fail(kLimitExceeded, "model count exceeds budget");
    return size_t(v);
  }
  void skip() {
    while (!t_.empty() && space(t_.front())) t_.remove_prefix(1);
  }
  bool done() const { return t_.empty(); }

 private:
  [[noreturn]] static void bad() { fail(kDecodeError, "invalid or truncated model field"); }
  std::string_view t_;
};
}  // namespace

std::string serialize_model(const NccModel& m) {
  if (m.levels.empty()) fail(kInvalidState, "cannot save empty model");
  const LevelGeom& lv = m.levels.front();
  std::string body;
  body.reserve(lv.gray.size() * 24 + 128);
  put(body, int64_t(m.params.levels)); body.push_back(' ');
  put(body, int64_t(m.params.absolute ? 1 : 0)); body.push_back(' ');
  put(body, m.params.start); body.push_back(' ');
  put(body, m.params.extent); body.push_back(' ');
  put(body, m.params.step); body.push_back('\n');
  put(body, m.orow); body.push_back(' ');
  put(body, m.ocol); body.push_back(' ');
  put(body, int64_t(lv.w)); body.push_back(' ');
  put(body, int64_t(lv.h)); body.push_back('\n');
  put(body, lv.orow); body.push_back(' ');
  put(body, lv.ocol); body.push_back('\n');
  for (size_t i = 0; i < lv.gray.size(); ++i) {
    put(body, int64_t(lv.mask[i]));
    body.push_back(' ');
    put(body, lv.gray[i]);
    body.push_back('\n');
  }
  if (body.size() > kMaxFile) fail(kLimitExceeded, "serialized model exceeds file budget");
  std::string out = "MVKIT_NCC 1\n";
  char digits[32];
  const auto r = std::to_chars(digits, digits + sizeof digits, fnv1a64(body));
  out.append(digits, r.ptr);
  out.push_back('\n');
  out += body;
  return out;
}

std::unique_ptr<NccModel> deserialize_model(const std::string& bytes) {
  if (bytes.size() > kMaxFile + 128) fail(kLimitExceeded, "model file exceeds budget");
  const size_t nl = bytes.find('\n');
  const std::string_view magic(bytes.data(), nl == std::string::npos ? bytes.size() : nl);
  if (magic != "MVKIT_NCC 1")
    fail(magic.substr(0, 10) == "MVKIT_NCC " ? kUnsupported : kDecodeError, "invalid model signature or version");
  // checksum: stream extraction of an unsigned integer, then exactly one newline
  size_t pos = nl == std::string::npos ? bytes.size() : nl + 1;
  while (pos < bytes.size() && space(bytes[pos])) ++pos;
  if (pos < bytes.size() && bytes[pos] == '+') ++pos;
  uint64_t expected = 0;
  const auto r = std::from_chars(bytes.data() + pos, bytes.data() + bytes.size(), expected);
  if (r.ec != std::errc{}) fail(kDecodeError, "invalid or truncated model field");
  pos = size_t(r.ptr - bytes.data());
  if (pos >= bytes.size() || bytes[pos] != '\n') fail(kDecodeError, "invalid checksum header");
  const std::string_view text(bytes.data() + pos + 1, bytes.size() - pos - 1);
  if (fnv1a64(text) != expected) fail(kDecodeError, "NCC model checksum mismatch");
  Fields in(text);
  CreateParams p;
  p.levels = int(in.count(16));
  p.absolute = in.count(1) != 0;
  p.start = in.real();
  p.extent = in.real();
  p.step = in.real();
  if (p.extent < 0 || p.extent > 6.283185307179586 || p.step <= 0 || !std::isfinite(p.start + p.extent))
    fail(kDecodeError, "invalid trained angle interval");
  const double orow = in.real(), ocol = in.real();
  LevelGeom first;
  first.w = in.count(kMaxPixels);
  first.h = in.count(kMaxPixels);
  if (!first.w || !first.h || first.h > kMaxPixels / first.w) fail(kLimitExceeded, "invalid model dimensions");
  first.orow = in.real();
  first.ocol = in.real();
  first.gray.resize(first.w * first.h);
  first.mask.resize(first.gray.size());
  long double row_sum = 0, col_sum = 0;
  size_t samples = 0;
  for (size_t i = 0; i < first.gray.size(); ++i) {
    first.mask[i] = uint8_t(in.count(1));
    first.gray[i] = in.real();
    if (first.mask[i]) {
      row_sum += i / first.w;
      col_sum += i % first.w;
      ++samples;
    }
  }
  if (samples < 4 || std::abs(first.orow - double(row_sum / samples)) > 1e-10 ||
      std::abs(first.ocol - double(col_sum / samples)) > 1e-10 || orow < first.orow || ocol < first.ocol ||
      std::abs((orow - first.orow) - std::round(orow - first.orow)) > 1e-8 ||
      std::abs((ocol - first.ocol) - std::round(ocol - first.ocol)) > 1e-8)
    fail(kDecodeError, "invalid template domain centroid");
  in.skip();
  if (!in.done()) fail(kDecodeError, "trailing model data");
  try {
    return rebuild_model(std::move(first), p, orow, ocol);
  } catch (const Error& e) {
    fail(e.status == kLimitExceeded ? kLimitExceeded : kDecodeError, "invalid prepared NCC model");
  }
}

}  // namespace mvamd
