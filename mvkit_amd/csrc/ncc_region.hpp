// Search domains: a region as canonical run-length encoding, the form the
// reference's image domains take (src/core/region.hpp:32-117).  Runs are
// sorted by (row, first column), disjoint and not adjacent within a row.
#pragma once

#include <cstdint>
#include <vector>

namespace mvamd {

struct RegionRun { int32_t row, cb, ce; };  // columns cb..ce inclusive

struct Region {
  std::vector<RegionRun> runs;
  int64_t area = 0;
  int64_t row_min = 0, col_min = 0, row_max = -1, col_max = -1;  // bounding box (valid when not empty)

  bool empty() const { return runs.empty(); }
  // Region::from_runs (region.cpp:64-117): sorts, merges overlapping or
  // adjacent runs of a row; rejects a run whose end precedes its beginning.
  static Region from_runs(std::vector<RegionRun> runs);
  // Region::rectangle (region.cpp:543-560): rows r0..r1 x columns c0..c1, empty
  // when inverted.
  static Region rectangle(int64_t r0, int64_t c0, int64_t r1, int64_t c1);
  // Pixels of this region inside rows r0..r1 x columns c0..c1.
  Region clipped(int64_t r0, int64_t c0, int64_t r1, int64_t c1) const;
  Region intersect(const Region& other) const;
  // Region::contains (region.cpp:937-945).
  bool contains(int64_t row, int64_t col) const;
};

}  // namespace mvamd
