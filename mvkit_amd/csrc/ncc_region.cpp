#include "ncc_region.hpp"

#include <algorithm>

#include "ncc_model.hpp"

namespace mvamd {

namespace {
void raise(Status s, const char* param, const char* msg) { throw Error(s, "Region::from_runs", param, msg); }

// Builds a region from runs that are already canonical.
Region from_canonical(std::vector<RegionRun> runs) {
  Region g;
  g.runs = std::move(runs);
  if (g.runs.empty()) return g;
  g.row_min = g.runs.front().row;
  g.row_max = g.runs.back().row;
  g.col_min = g.runs.front().cb;
  g.col_max = g.runs.front().ce;
  for (const auto& r : g.runs) {
    g.area += int64_t(r.ce) - r.cb + 1;
    g.col_min = std::min<int64_t>(g.col_min, r.cb);
    g.col_max = std::max<int64_t>(g.col_max, r.ce);
  }
  return g;
}
}  // namespace

Region Region::from_runs(std::vector<RegionRun> in) {
  for (const auto& r : in)
    if (r.cb > r.ce) raise(kInvalidArgument, "Region", "run begin column must not exceed end column");
  std::sort(in.begin(), in.end(), [](const RegionRun& a, const RegionRun& b) {
    if (a.row != b.row) return a.row < b.row;
    if (a.cb != b.cb) return a.cb < b.cb;
    return a.ce < b.ce;
  });
  std::vector<RegionRun> out;
  for (const auto& r : in) {
    if (!out.empty() && out.back().row == r.row && int64_t(r.cb) <= int64_t(out.back().ce) + 1)
      out.back().ce = std::max(out.back().ce, r.ce);
    else
      out.push_back(r);
  }
  return from_canonical(std::move(out));
}

Region Region::rectangle(int64_t r0, int64_t c0, int64_t r1, int64_t c1) {
  std::vector<RegionRun> runs;
  if (r0 <= r1 && c0 <= c1)
    for (int64_t r = r0; r <= r1; ++r) runs.push_back({int32_t(r), int32_t(c0), int32_t(c1)});
  return from_canonical(std::move(runs));
}

Region Region::clipped(int64_t r0, int64_t c0, int64_t r1, int64_t c1) const {
  std::vector<RegionRun> out;
  for (const auto& r : runs)
    if (r.row >= r0 && r.row <= r1 && r.ce >= c0 && r.cb <= c1)
      out.push_back({r.row, int32_t(std::max<int64_t>(r.cb, c0)), int32_t(std::min<int64_t>(r.ce, c1))});
  return from_canonical(std::move(out));
}

Region Region::intersect(const Region& o) const {
  std::vector<RegionRun> out;
  size_t i = 0, j = 0;
  while (i < runs.size() && j < o.runs.size()) {
    const RegionRun& a = runs[i];
    const RegionRun& b = o.runs[j];
    if (a.row != b.row) {
      if (a.row < b.row) ++i; else ++j;
      continue;
    }
    const int32_t lo = std::max(a.cb, b.cb), hi = std::min(a.ce, b.ce);
    if (lo <= hi) out.push_back({a.row, lo, hi});
    if (a.ce < b.ce) ++i; else ++j;
  }
  return from_canonical(std::move(out));
}

bool Region::contains(int64_t row, int64_t col) const {
  const auto it = std::lower_bound(runs.begin(), runs.end(), std::make_pair(row, col),
                                   [](const RegionRun& r, const std::pair<int64_t, int64_t>& t) {
                                     return r.row < t.first || (r.row == t.first && r.ce < t.second);
                                   });
  return it != runs.end() && it->row == row && it->cb <= col;
}

}  // namespace mvamd
