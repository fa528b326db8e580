// Host side of the split-BF16 MFMA coarse correlation: kernel arguments,
// tile shapes and the launches (one grid over a batch's frames, or one
// persistent launch over a work queue of many queries).  The kernels are in
// ncc_mfma.hip.h, instantiated per rows-per-wave in ncc_mfma_r4.hip and
// ncc_mfma_r8.hip.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "ncc_mfma.hip.h"
#include "../../include/ncchip.h"

using namespace ncc;

namespace {
constexpr NcchipCoarseTile kDefaultTile = {4, 4, 1, 1, 0};

bool valid_tile(const NcchipCoarseTile& t) {
  auto pow2_4 = [](int v) { return v == 1 || v == 2 || v == 4; };
  return (t.rows == 4 || t.rows == 8) && pow2_4(t.col_waves) && pow2_4(t.row_waves) &&
         t.col_waves * t.row_waves <= kMaxWaves && t.tiles >= 1 && t.tiles <= 8 &&
         (t.bound == 0 || t.bound == 8 || t.bound == 16 || t.bound == 32);
}

size_t static_lds_of(int rows) { return rows == 8 ? static_lds<8>() : static_lds<4>(); }

// LDS bytes of one coarse block with tile shape t, beside its static tables;
// 0 when the level or the shape cannot run in the split-BF16 form.
size_t coarse_mfma_lds(int level, int level_kind, int yspan, int xchunks, bool screened, const NcchipCoarseTile& t,
                       int* pitch, int* pitch8) {
  if (level_kind == NCC_U32 || level > 4) return 0;  // T_lo exact in BF16 only up to level 4
  if (!valid_tile(t)) return 0;
  if (screened && t.bound) return 0;  // the screen decides on the bound of every patch
  const int cols = kWaveCols * t.col_waves + 16 * xchunks + 9;
  const int rows = t.rows * t.row_waves + yspan - 1;
  // the staging's float-reciprocal index needs rows * cols < 2^16; the run
  // codes a column below 2^8 and a row below 2^15
  if (size_t(rows) * size_t(cols) >= (size_t(1) << 16)) return 0;
  *pitch = (cols + 1) & ~1;
  *pitch8 = (cols + 11) & ~3;  // an 8-byte A fragment read as three dwords from any column
  size_t lds = size_t(5) * rows * *pitch * sizeof(uint16_t);
  if (screened) lds += size_t(rows) * *pitch8;
  // bound groups: each group's deviations per wave, row and anchor
  if (t.bound) lds += size_t(kWaveCols / t.bound) * t.col_waves * t.row_waves * t.rows * kWaveCols * sizeof(float);
  const size_t kLdsMax = 160 * 1024 - static_lds_of(t.rows);
  return lds > kLdsMax ? 0 : lds;
}

const NcchipCoarseTile& tile_or_default(const NcchipCoarseTile* t) {
  return t && t->rows ? *t : kDefaultTile;
}
}  // namespace

extern "C" int ncchip_coarse_mfma_fits(int level, int level_kind, int yspan, int xchunks, int screened,
                                       const NcchipCoarseTile* tile) {
  int pitch = 0, pitch8 = 0;
  return coarse_mfma_lds(level, level_kind, yspan, xchunks, screened != 0, tile_or_default(tile), &pitch, &pitch8) > 0;
}

namespace {
// Kernel arguments of one split-BF16 coarse search; 0 on success, -6 when the
// level or the tile does not fit the kernel, -1 on bad arguments.
int make_args(const NccFrame* frames, int n_frames, int level, int level_kind, double C, const NccPatchPool* pools,
              const int32_t* patch_ids, int n_angles, const NcchipDenseWeights* dw, int exact_lattice, int absolute,
              double screen, const NcchipCoarseTile& t, MfmaArgs& a, size_t& lds, bool& narrow) {
  const bool screened = screen > -__builtin_inf();
  if (screened && !(dw->w8_scale > 0)) return -1;
  a = MfmaArgs{};
  a.frames = frames; a.level = level; a.n_angles = n_angles; a.n_tiles = (n_angles + 31) / 32; a.C = C;
  a.pools.p0 = pools[0]; a.pools.p1 = pools[1]; a.patch_ids = patch_ids;
  a.wh = reinterpret_cast<const uint4*>(dw->wh); a.wl = reinterpret_cast<const uint4*>(dw->wl);
  a.step0 = dw->step0; a.segs = reinterpret_cast<const int2*>(dw->segs);
  a.wsum_eff = dw->wsum_eff; a.ymin = dw->ymin; a.xmin = dw->xmin; a.yspan = dw->yspan; a.xc = dw->xchunks;
  a.exact_lattice = exact_lattice; a.absolute = absolute;
  a.common_runs = reinterpret_cast<const int4*>(dw->common_runs); a.n_common_runs = dw->n_common_runs;
  a.n_common = dw->n_common; a.prune_below = dw->prune_below; a.peak_thr = dw->peak_thr;
  if (t.bound) {
    if (!dw->group_first || !dw->group_n || !dw->group_runs) return -1;
    const int i = t.bound == 8 ? 0 : t.bound == 16 ? 1 : 2;
    a.group_slots = t.bound;
    a.group_first = dw->group_first + dw->first_base[i];
    a.group_n = dw->group_n + dw->n_base[i];
    a.group_runs = reinterpret_cast<const int4*>(dw->group_runs);
  }
  lds = coarse_mfma_lds(level, level_kind, a.yspan, a.xc, screened, t, &a.pitch, &a.pitch8);
  if (!lds) return -6;
  a.cw = t.col_waves; a.rw = t.row_waves;
  a.tiles_per_block = std::min(t.tiles, a.n_tiles);
  a.n_groups = (a.n_tiles + a.tiles_per_block - 1) / a.tiles_per_block;
  if (size_t(n_frames) * a.n_groups > 65535) return -1;
  if (screened) {
    a.screen = screen;
    a.img8_scale = float(std::ldexp(1.0, -2 * level));
    a.w8_scale = float(dw->w8_scale);
    a.dot8_unscale = 1.0 / (double(a.img8_scale) * double(a.w8_scale));
  }
  // per-run sums of t^2 fit 32 bits when (longest run) * (128 * 4^level)^2 < 2^32
  const double tmax = 128.0 * double(1 << (2 * level));
  narrow = double(dw->max_run) * tmax * tmax < 4294967296.0;
  return 0;
}

// Instantiation of a coarse launch: pixel type, run accumulator width, screen.
int variant(int level_kind, bool narrow, bool screened) {
  return (level_kind == NCC_U8 ? 0 : 4) | (narrow ? 0 : 2) | (screened ? 1 : 0);
}

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

// persistent launches of one queue call (kernel instantiation x block size x
// occupancy class), each with its own item counter
constexpr int kMaxLaunches = 256;
}  // namespace

extern "C" int ncchip_coarse_mfma(const NccFrame* frames, int n_frames, int level, int level_kind, int max_w,
                                  int max_h, double C, const NccPatchPool* pools, const int32_t* patch_ids,
                                  int n_angles, const NcchipDenseWeights* dw, int exact_lattice, int absolute,
                                  double screen, const NcchipCoarseTile* tile, void* stream) {
  if (n_frames <= 0 || n_angles <= 0) return 0;
  const NcchipCoarseTile& t = tile_or_default(tile);
  MfmaArgs a;
  size_t lds = 0;
  bool narrow = false;
  const int rc = make_args(frames, n_frames, level, level_kind, C, pools, patch_ids, n_angles, dw, exact_lattice,
                           absolute, screen, t, a, lds, narrow);
  if (rc) return rc;
  const int bw = kWaveCols * t.col_waves, bh = t.rows * t.row_waves;
  const dim3 grid((max_w + bw - 1) / bw, (max_h + bh - 1) / bh, unsigned(n_frames * a.n_groups));
  const int threads = 64 * t.col_waves * t.row_waves;
  const int v = variant(level_kind, narrow, screen > -__builtin_inf()) | (a.tiles_per_block > 1 ? 8 : 0);
  hipStream_t st = static_cast<hipStream_t>(stream);
  return t.rows == 8 ? coarse_launch_r8(v, grid, threads, lds, st, a) : coarse_launch_r4(v, grid, threads, lds, st, a);
}

extern "C" size_t ncchip_coarse_mfma_queue_bytes(int n_jobs, int n_units) {
  return align256(size_t(n_jobs) * sizeof(MfmaArgs)) + align256(size_t(n_units) * sizeof(QueueUnit)) +
         align256((size_t(n_units) + kMaxLaunches) * sizeof(int)) + kMaxLaunches * sizeof(unsigned);
}

extern "C" int ncchip_coarse_mfma_queue(const NcchipCoarseJob* jobs, int n_jobs, void* d_work, size_t work_bytes,
                                        void* h_stage, size_t stage_bytes, void* stream) {
  if (n_jobs <= 0) return 0;
  if (!jobs || !d_work || !h_stage) return -1;
  std::vector<MfmaArgs> args(static_cast<size_t>(n_jobs));
  std::vector<size_t> lds(static_cast<size_t>(n_jobs));
  // the launch a job joins: kernel instantiation (variant, rows per wave),
  // the block's thread count and the occupancy its LDS allows
  std::vector<int> key(static_cast<size_t>(n_jobs), -1);
  int n_units = 0;
  for (int j = 0; j < n_jobs; ++j) {
    const NcchipCoarseJob& J = jobs[j];
    if (J.n_frames <= 0 || J.n_angles <= 0) continue;
    bool narrow = false;
    const NcchipCoarseTile& t = tile_or_default(&J.tile);
    const int rc = make_args(J.frames, J.n_frames, J.level, J.level_kind, J.C, J.pools, J.patch_ids, J.n_angles,
                             &J.weights, J.exact_lattice, J.absolute, J.screen, t, args[size_t(j)], lds[size_t(j)],
                             narrow);
    if (rc) return rc;
    // the workgroups per CU its LDS allows: a job sharing a launch with one
    // of larger LDS would run at that job's lower occupancy
    const int occ = std::min(8, int((160 * 1024) / (lds[size_t(j)] + static_lds_of(t.rows))));
    key[size_t(j)] = variant(J.level_kind, narrow, J.screen > -__builtin_inf()) |
                     (args[size_t(j)].tiles_per_block > 1 ? 8 : 0) | (t.rows == 8 ? 16 : 0) |
                     ((t.col_waves * t.row_waves) << 5) | (occ << 8);
    n_units += J.n_frames;
  }
  const size_t need = ncchip_coarse_mfma_queue_bytes(n_jobs, n_units);
  if (work_bytes < need || stage_bytes < need) return -103;
  auto* base = static_cast<uint8_t*>(h_stage);
  const size_t off_units = align256(size_t(n_jobs) * sizeof(MfmaArgs));
  const size_t off_first = off_units + align256(size_t(n_units) * sizeof(QueueUnit));
  const size_t off_counters = off_first + align256((size_t(n_units) + kMaxLaunches) * sizeof(int));
  std::memcpy(base, args.data(), size_t(n_jobs) * sizeof(MfmaArgs));
  auto* units = reinterpret_cast<QueueUnit*>(base + off_units);
  auto* first = reinterpret_cast<int*>(base + off_first);
  std::memset(base + off_counters, 0, kMaxLaunches * sizeof(unsigned));
  // units grouped by launch key; each group's first[] ends with its item count
  struct Group { int key, unit0, n_units, first0; size_t lds; int64_t items; };
  std::vector<Group> groups;
  std::vector<int> keys;
  for (int j = 0; j < n_jobs; ++j)
    if (key[size_t(j)] >= 0 && std::find(keys.begin(), keys.end(), key[size_t(j)]) == keys.end()) keys.push_back(key[size_t(j)]);
  if (keys.size() > size_t(kMaxLaunches)) return -103;  // one counter per launch
  // within a launch, the jobs whose items cost most come first (items are
  // taken in order: long items left for the end would idle the other CUs);
  // an item's cost ~ its angle tiles x block area x the union box
  std::vector<double> item_cost(static_cast<size_t>(n_jobs), 0.0);
  for (int j = 0; j < n_jobs; ++j) {
    if (key[size_t(j)] < 0) continue;
    const MfmaArgs& a = args[size_t(j)];
    const NcchipCoarseTile& t = tile_or_default(&jobs[j].tile);
    item_cost[size_t(j)] = double(a.tiles_per_block) * (t.rows * a.rw) * (kWaveCols * a.cw) * double(a.yspan) * a.xc;
  }
  std::vector<int> order(static_cast<size_t>(n_jobs));
  for (int j = 0; j < n_jobs; ++j) order[size_t(j)] = j;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return item_cost[size_t(x)] > item_cost[size_t(y)]; });
  int u = 0, fpos = 0;
  for (const int k : keys) {
    Group g{k, u, 0, fpos, 0, 0};
    for (const int j : order) {
      const NcchipCoarseJob& J = jobs[j];
      if (key[size_t(j)] != k) continue;
      const MfmaArgs& a = args[size_t(j)];
      g.lds = std::max(g.lds, lds[size_t(j)]);
      const int bw = kWaveCols * a.cw, bh = tile_or_default(&J.tile).rows * a.rw;
      for (int f = 0; f < J.n_frames; ++f) {
        const int w = J.frame_dims ? J.frame_dims[2 * f] : J.max_w, h = J.frame_dims ? J.frame_dims[2 * f + 1] : J.max_h;
        const int gx = (w + bw - 1) / bw, gy = (h + bh - 1) / bh;
        units[u] = QueueUnit{j, f, std::max(gx, 1), std::max(gy, 1)};
        first[fpos++] = int(g.items);
        g.items += int64_t(std::max(gx, 1)) * std::max(gy, 1) * a.n_groups;
        ++u;
        ++g.n_units;
      }
    }
    if (g.items >= (int64_t(1) << 31)) return -103;
    first[fpos++] = int(g.items);
    groups.push_back(g);
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (hipMemcpyAsync(d_work, h_stage, need, hipMemcpyHostToDevice, st) != hipSuccess) return -1000;
  int dev = 0, n_cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&n_cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n_cus <= 0)
    n_cus = 256;
  auto* dbase = static_cast<uint8_t*>(d_work);
  for (size_t k = 0; k < groups.size(); ++k) {
    const Group& g = groups[k];
    QueueArgs q{};
    q.args = reinterpret_cast<const MfmaArgs*>(dbase);
    q.units = reinterpret_cast<const QueueUnit*>(dbase + off_units) + g.unit0;
    q.first = reinterpret_cast<const int*>(dbase + off_first) + g.first0;
    q.n_units = g.n_units;
    q.n_items = int(g.items);
    q.counter = reinterpret_cast<unsigned*>(dbase + off_counters) + k;
    const int v = g.key & 15, threads = 64 * ((g.key >> 5) & 7);
    const int rc = (g.key & 16) ? coarse_queue_r8(v, threads, g.lds, st, q, n_cus)
                                : coarse_queue_r4(v, threads, g.lds, st, q, n_cus);
    if (rc) return rc;
  }
  return 0;
}
