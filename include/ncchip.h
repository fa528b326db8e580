/* ncchip: the thin C ABI between the C++ host driver and the HIP kernels of
 * the NCC find path.  Plain pointers and sizes only; every pointer argument
 * named d_* or documented as device memory is a HIP device pointer, `stream`
 * is a hipStream_t passed as void*.  Each entry point names the reference
 * routine (mvkit src/ncc_matching, file:line) whose work it replaces.
 * Return value: 0 on success, negative MvStatus on failure. */
#ifndef MVKIT_AMD_NCCHIP_H
#define MVKIT_AMD_NCCHIP_H

#include <stddef.h>
#include <stdint.h>
#include "../mvkit_amd/csrc/device_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Correlation forms of the coarse search. */
#define NCCHIP_FORM_FP32 0      /* FP32 FMA dot products (plain FP32 form) */
#define NCCHIP_FORM_BF16X3 1    /* split-BF16 MFMA: hi/lo template x hi/lo image, FP32 accumulate */
#define NCCHIP_FORM_FP64 2      /* FP64, reference sample order (checker / error reports) */
#define NCCHIP_FORM_FP8_SCREEN 3 /* FP8-E4M3 screen + split-BF16 re-score */

/* One pyramid level from the previous one: exact integer 2x2 block sums with
 * the reference's edge clamp.  Replaces the level loop of prepare()
 * (matcher.cpp:436-454; kernels_avx2.cpp:43-61; gray_pyramid.hpp:8-13).
 * src/dst are device pointers; kinds are NCC_U8/NCC_U16/NCC_U32. */
int ncchip_pyramid_level(const void* src, int src_kind, int src_pitch, int src_w, int src_h, void* dst, int dst_kind,
                         int dst_pitch, int dst_w, int dst_h, void* stream);

/* Row prefix sums of S and S^2 (u64), (w+1) per row with a leading
 * zero, plus the support integral of (S != D) if `support` is non-null
 * ((w+1)*(h+1) u32).  `nz` is (w+1)*h u32 scratch.  Replaces the top-level
 * prefix statistics and support integral (matcher.cpp:487-503). */
int ncchip_row_prefix(const void* src, int kind, int pitch, int w, int h, const uint8_t* dc_src, int shift,
                      uint64_t* p1, uint64_t* p2, uint32_t* nz, uint32_t* support, void* stream);

/* ncchip_row_prefix without the `nz` workspace argument and without any
 * allocation: no counterpart in the reference; for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The support indicator's row prefixes go
 * to a static device array of 2^22 u32 words (16 MiB) per GPU, held until
 * the process ends, so (w + 1) * h must not exceed 4194304 (else -103,
 * LIMIT_EXCEEDED).  All calls on one GPU share that array: calls that
 * overlap on one GPU (other streams, other host threads) race.  Outputs are
 * those of ncchip_row_prefix. */
int ncchip_row_prefix_noalloc(const void* src, int kind, int pitch, int w, int h, const uint8_t* dc_src, int shift,
                              uint64_t* p1, uint64_t* p2, uint32_t* support, void* stream);

/* Batched front end.  `frames` (device) holds one NccFrame per frame of a
 * batch (at most NCC_MAX_BATCH); each batched kernel covers every frame in
 * one launch, sized by the largest frame (max_w x max_h of the level
 * concerned); blocks past a smaller frame's edge exit. */

/* Pyramid level `level` from level - 1 for every frame (the levels' data
 * pointers are in the descriptors).  Replaces matcher.cpp:436-454. */
int ncchip_pyramid_batch(const NccFrame* frames, int n_frames, int level, int src_kind, int dst_kind, int max_w,
                         int max_h, void* stream);

/* Levels 1..n_levels (n_levels <= 3) of every frame from a u8 level 0 in
 * one launch (u16 block sums, as ncchip_pyramid_batch level by level). */
int ncchip_pyramid3_u8(const NccFrame* frames, int n_frames, int n_levels, int max_w, int max_h, void* stream);

/* Row prefixes of `level` (p1, p2, nz of each descriptor).
 * Replaces matcher.cpp:487-495. */
int ncchip_row_prefix_batch(const NccFrame* frames, int n_frames, int level, int kind, int max_h, void* stream);

/* Fills n u32 words of device memory with `value` (a kernel, not a memset). */
int ncchip_fill_u32(uint32_t* p, size_t n, uint32_t value, void* stream);

/* Copies n u32 words from device memory to device or host-visible memory. */
int ncchip_copy_u32(const uint32_t* src, uint32_t* dst, size_t n, void* stream);

/* Coarse exhaustive NCC at the top level `level` of every frame for each
 * coarse angle, writing each frame's FP64 score planes [n_angles][h][w]
 * (-inf outside each angle's valid anchor range, descriptor `ranges`).
 * `pools` points to two raster pools (the model's and the per-find rasters);
 * `patch_ids` are (pool << 24 | index).  C = 128 * 4^level centres S.
 * Replaces matcher.cpp:542-579 with kernels_avx2.cpp:105-238. */
int ncchip_coarse(const NccFrame* frames, int n_frames, int level, int level_kind, int max_w, int max_h, double C,
                  const NccPatchPool* pools, const int32_t* patch_ids, int n_angles, int max_bbox_h, int max_bbox_w,
                  int form, int exact_lattice, int absolute, void* stream);

/* Dense split-BF16 weights of the coarse patches within their union bounding
 * box (rows ymin .. ymin + yspan - 1, columns xmin .. xmin + 16 * xchunks - 1,
 * which sizes the kernel's image tile), laid out as MFMA B fragments by K
 * step: the steps of angle tile t (32 angles) are step0[t] .. step0[t+1] - 1;
 * step s covers two 8-column segments, segs[2 s] (half-wave 0) and
 * segs[2 s + 1] (half-wave 1), each packed as (y << 16 | x) relative to
 * (ymin, xmin); a tile's segments cover, row by row, the union of its
 * patches' supports.  wh / wl hold per step 64 lanes x 8 bf16 (lane L: angle
 * 32 t + L % 32, the 8 columns of segment L / 32).  wsum_eff[a] is the FP64
 * sum of w_hi + w_lo over angle a's samples; max_run the longest row run. */
typedef struct NcchipDenseWeights {
  const void* wh;          /* device, 16-byte aligned */
  const void* wl;
  const int32_t* step0;    /* device, tiles + 1 entries */
  const int32_t* segs;     /* device, 2 per step */
  const double* wsum_eff;  /* device */
  int32_t ymin, xmin, yspan, xchunks, max_run;
  /* Runs (4 ints: row, first col, last col, 0) of the support common to
   * every coarse patch, and its sample count.  Its variance bounds every
   * patch's window variance from below, so |score| <= |num| / sqrt(var_common);
   * windows whose bound is below prune_below are written -inf without exact
   * moments (they can neither be peaks nor suppress one when prune_below is
   * under the peak threshold).  n_common = 0 or prune_below <= 0 disables it. */
  const int32_t* common_runs;
  int32_t n_common_runs, n_common;
  double prune_below;
  /* Bound groups (NcchipCoarseTile.bound): for group sizes 8, 16 and 32
   * (index i = 0, 1, 2) the angles a / size form groups, each with the runs
   * of the support common to its own patches: group g of size index i has
   * runs group_first[first_base[i] + g] .. group_first[first_base[i] + g + 1]
   * - 1 of group_runs (4 ints each, as common_runs) and group_n[n_base[i] + g]
   * samples.  Null group_first: no bound groups. */
  const int32_t* group_first;
  const int32_t* group_n;
  const int32_t* group_runs;
  int32_t first_base[3], n_base[3];
  /* > -inf: sparse planes -- only scores >= peak_thr are written, each with
   * its bit in the frame's peak_bits (zeroed by the caller); the peaks
   * kernel then runs with hwa = 2.  -inf: every plane entry is written. */
  double peak_thr;
  /* FP8 screen: power of two with max |w_hi| * w8_scale <= 128 (E4M3 range
   * for either encoding of the format). */
  double w8_scale;
} NcchipDenseWeights;

/* Split-BF16 MFMA coarse search (the planes of ncchip_coarse, stored in
 * [row][col][angle] order so that the angles of one anchor are adjacent): the dot
 * products on v_mfma_f32_32x32x16_bf16 (image split T_hi + T_lo, weights
 * w_hi + w_lo, three products in FP32), exact integer window moments and the
 * shared FP64 normalisation.  Top levels of kind NCC_U8 / NCC_U16 up to level 4.
 * `screen` > -inf adds the FP8-E4M3 screen: an E4M3 MFMA pass over each
 * wave's 32 anchors x 4 rows x 32 angles, after which a wave none of whose
 * windows reaches (|dot8| + C |wsum|) / sd_common >= screen skips its
 * split-BF16 products (needs weights->w8_scale; the surviving waves' scores
 * are the split-BF16 form's). */
/* Tile shape of the split-BF16 coarse kernel: each wave computes `rows`
 * output rows (4 or 8) x 32 anchors; a block has col_waves waves across and
 * row_waves waves down (col_waves * row_waves <= 4; each 1, 2 or 4) and runs
 * `tiles` angle tiles of 32 rotations (1..8) over one staged image tile.
 * `bound`: 0 bounds every window's score with the support common to every
 * coarse patch; 8, 16 or 32 with the support common to the patches of its
 * group of that many consecutive angles (a tighter bound, computed per group:
 * fewer windows need exact moments; not with the FP8 screen).  The planes are
 * bit-identical for every shape; only the time differs.  The default shape is
 * {4, 4, 1, 1, 0}. */
typedef struct NcchipCoarseTile {
  int32_t rows, col_waves, row_waves, tiles, bound;
} NcchipCoarseTile;

/* 1 when ncchip_coarse_mfma can run the top level `level` of pixel kind
 * level_kind over a union bounding box of yspan rows and xchunks 16-column
 * chunks with tile shape `tile` (null: the default shape): the shape is
 * valid and its image tile fits in LDS (levels up to 4 of u8/u16 block
 * sums), else 0.  The host driver then runs the FP32 form for that search
 * (default shape) or another shape. */
int ncchip_coarse_mfma_fits(int level, int level_kind, int yspan, int xchunks, int screened,
                            const NcchipCoarseTile* tile);

int ncchip_coarse_mfma(const NccFrame* frames, int n_frames, int level, int level_kind, int max_w, int max_h,
                       double C, const NccPatchPool* pools, const int32_t* patch_ids, int n_angles,
                       const NcchipDenseWeights* weights, int exact_lattice, int absolute, double screen,
                       const NcchipCoarseTile* tile, void* stream);

/* The split-BF16 coarse searches of several queries (models of uneven
 * template size and angle count, over frames of uneven size) in one
 * persistent launch per kernel instantiation, fed by an (image, angle tile,
 * block row, block column) work queue: each job holds the arguments of one
 * ncchip_coarse_mfma call, and the planes written are the ones that call
 * writes.  frame_dims (host, optional) gives each frame's width and height at
 * the searched level, so that the queue holds the frame's own blocks only.
 * d_work (device) and h_stage (pinned host, read by an asynchronous copy
 * that must complete before h_stage is reused) hold
 * ncchip_coarse_mfma_queue_bytes(n_jobs, total frames) bytes.  Returns -6
 * when a job does not fit the kernel (see ncchip_coarse_mfma_fits), -103
 * when the buffers are too small.  Replaces, for a batch of queries, the
 * coarse loop of prepare() (matcher.cpp:542-579). */
typedef struct NcchipCoarseJob {
  const NccFrame* frames;     /* device */
  int32_t n_frames, level, level_kind, max_w, max_h;
  const int32_t* frame_dims;  /* host: (w, h) per frame, or null */
  double C;
  NccPatchPool pools[2];
  const int32_t* patch_ids;   /* device */
  int32_t n_angles;
  NcchipDenseWeights weights;
  int32_t exact_lattice, absolute;
  double screen;
  NcchipCoarseTile tile;      /* tile shape; rows == 0: the default shape */
} NcchipCoarseJob;
size_t ncchip_coarse_mfma_queue_bytes(int n_jobs, int n_frames_total);
/* The queue of one launch lists the items of the jobs that cost most first
 * (long items taken last would leave the other CUs idle); jobs share a
 * launch when they share the kernel instantiation, the block size and the
 * occupancy their LDS allows. */
int ncchip_coarse_mfma_queue(const NcchipCoarseJob* jobs, int n_jobs, void* d_work, size_t work_bytes, void* h_stage,
                             size_t stage_bytes, void* stream);

/* Per-angle 3x3 peaks of every frame's planes with the reference's tie rule
 * (matcher.cpp:580-594), emitted as candidate poses (matcher.cpp:590-592:
 * row = crop_r + scale * (r + origin_r) + phase) into frame f's list at
 * cand + f * cap (unordered), with the angle index in key; count[f] (zeroed
 * by the caller) may exceed cap.  extent[4f..4f+3] (min row, max row,
 * min col, max col of the candidates' offsets from the crop origin, as FP64
 * bit patterns; the caller initialises them to h, 0, w, 0) feeds
 * ncchip_storage0.  hwa != 0: the planes are in [row][col][angle] order
 * (as ncchip_coarse_mfma writes them), else [angle][row][col]. */
/* hwa = 2: the sparse planes of the split-BF16 form (peak_bits set). */
int ncchip_peaks(const NccFrame* frames, int n_frames, int level, int max_w, int max_h, const NccPatchPool* pools,
                 const int32_t* patch_ids, int n_angles, double threshold, double scale, double phase, NccMatch* cand,
                 int32_t* key, int cap, int32_t* count, unsigned long long* extent, int hwa, void* stream);

/* Level-0 storage rectangle of each frame (matcher.cpp:612-629) from its
 * candidate extent and margin = radius + 4 * 2^top + 8, written into the
 * descriptors' storage0 (frames without candidates: the whole level). */
int ncchip_storage0(NccFrame* frames, int n_frames, const int32_t* count, const unsigned long long* extent,
                    double margin, void* stream);

/* Reduced search domains (matcher.cpp:376-393, 505-534, 631-645).  Frames
 * whose dom_sat is null are skipped by all three.
 * ncchip_domain_sat: builds each frame's summed-area table of its domain mask
 * over the crop from dom_runs / dom_row (grid: max crop height, max crop
 * width).
 * ncchip_domain_planes: sets the coarse score of every anchor whose pose may
 * not reach the domain (may_contain at spacing `scale`) to missing; between
 * the coarse correlation and ncchip_peaks, same plane order flag `hwa`.
 * ncchip_domain_snap: moves the candidates of a list (frame f's at
 * cand + f * cap, count[f] of them) that are not on a domain pixel to the
 * nearest domain pixel within 4 rows and columns, else scores them missing. */
int ncchip_domain_sat(const NccFrame* frames, int n_frames, int max_h, int max_w, void* stream);
int ncchip_domain_planes(const NccFrame* frames, int n_frames, int level, int max_w, int max_h,
                         const NccPatchPool* pools, const int32_t* patch_ids, int n_angles, double scale, double phase,
                         int hwa, void* stream);
int ncchip_domain_snap(const NccFrame* frames, int n_frames, NccMatch* cand, const int32_t* count, int cap,
                       void* stream);

/* Model construction on the device (model.cpp:39-98, 173-242), one
 * workgroup per raster job.  Stages, in order:
 *   NCCHIP_CREATE_COUNT: stats[j].count = samples of job j;
 *   NCCHIP_CREATE_SAMPLES: with jobs[j].sample_off set, the samples (rc, v)
 *     in grid order, then the weights (w, wf) and the rest of stats[j]
 *     (status, bounds, weight sums, run counts);
 *   NCCHIP_CREATE_RUNS: with jobs[j].run_off / wrun_off set, the row runs
 *     (4 ints: row, first column, last column, 0), the weight runs and their
 *     weights. */
#define NCCHIP_CREATE_COUNT 0
#define NCCHIP_CREATE_SAMPLES 1
#define NCCHIP_CREATE_RUNS 2
typedef struct NcchipCreateBuffers {
  const NccTemplateLevel* levels;
  const NccRasterJob* jobs;
  NccRasterStats* stats;
  int32_t* rc;
  double* v;
  double* w;
  float* wf;
  int32_t* runs;
  int32_t* wruns;
  double* wrun_w;
  int32_t* run_start;   /* scratch: one int per row run */
  int32_t* wrun_start;  /* scratch: one int per weight run */
} NcchipCreateBuffers;
int ncchip_create_rasters(int stage, const NcchipCreateBuffers* buffers, int n_jobs, void* stream);

/* Every kernel taking `dc_src` reads the reference pixel dc (the first
 * pixel of the search crop, matcher.cpp:410-413) from that device address;
 * integer levels use D = dc << shift. */

/* Coarse-to-fine tracking of every candidate of a batch at one level: the
 * first strict maximum over (query patch within `tol` of the candidate's
 * angle, 5x5 anchors), then the reference's merge and filter.  Frame f's
 * candidates are in + f*cap, in_count[f] of them (clamped to cap); the
 * tracked pose of each goes to the same slot of `raw`; those with a finite
 * score >= keep (0.4 * MinScore) and no identical (row, col, angle) kept
 * before are appended to out + f*cap (unordered), counted in out_count[f]
 * (zeroed by the caller).  `table` holds table_size int32 per frame (a power
 * of two >= 2*cap, filled with -1 by the caller).  `level` is the tracked
 * level and `level_kind` its pixel kind; `refs` lists the level's query
 * patches as (pool << 24 | index).  `tile_pixels` bounds the pixels of a
 * candidate's neighbourhood (the union of its windows), staged in LDS when
 * it fits in 64 KiB (0: read the level from HBM); exact_f64 != 0 asserts
 * n * (max |S - D|)^2 < 2^53 for every query patch, which lets the moments
 * accumulate exactly in FP64.  max_eligible bounds the eligible patches of
 * any candidate (0: unknown); when it is at most 32, the neighbourhood fits
 * in LDS and the level is u8 or u16, each sample is loaded once for all 25
 * windows of its patch and the moments come from sliding integer sums along
 * the support runs (same results); x_max bounds |S - D| at the level (below
 * 2^15 the neighbourhood is staged as x = S - D in 16 bits).  `work` (optional, device) accumulates the
 * samples of every evaluated window (algorithmic work = 2 FLOP per sample).
 * `grid` workgroups loop over the candidates.
 * Replaces matcher.cpp:670-760 (kernels_avx2.cpp:118-238) and
 * merge_identical (matcher.cpp:241-248). */
int ncchip_track(const NccFrame* frames, int n_frames, int level, int level_kind, const NccPatchPool* pools,
                 const int32_t* refs, int n_refs, const NccMatch* in, const int32_t* in_count, int cap, NccMatch* raw,
                 NccMatch* out, int32_t* out_count, int32_t* table, int table_size, double keep, double tol,
                 double step, double offset, int exact_lattice, int absolute, int tile_pixels, int exact_f64,
                 int max_eligible, double x_max, unsigned long long* work, int grid, void* stream);

/* ncchip_track on matrix cores, for u16 levels 1..4: per candidate, the 5x5
 * windows of all eligible patches as one split-BF16 GEMM
 * (v_mfma_f32_32x32x16_bf16) with a rigorous error bound per window; the
 * windows that can hold the maximum are re-scored exactly in FP64, so the
 * tracked pose is the one ncchip_track returns.  tile_h / tile_w bound the
 * rows / columns of any candidate's neighbourhood (the union of its
 * windows); x_max bounds |S - dc 4^l| at the level.  Candidates with more
 * than 32 eligible patches or a larger neighbourhood are scored exactly.
 * Returns -6 when the level, kind or LDS budget does not allow it (then use
 * ncchip_track).  Same inputs, outputs and reference lines as ncchip_track. */
int ncchip_track_mfma(const NccFrame* frames, int n_frames, int level, int level_kind, const NccPatchPool* pools,
                      const int32_t* refs, int n_refs, const NccMatch* in, const int32_t* in_count, int cap,
                      NccMatch* raw, NccMatch* out, int32_t* out_count, int32_t* table, int table_size, double keep,
                      double tol, double step, double offset, int exact_lattice, int absolute, int tile_h,
                      int tile_w, double x_max, unsigned long long* work, int grid, void* stream);

typedef struct NcchipRefineParams {
  const NccFrame* frames;  /* per-frame descriptors (device) */
  int32_t level_kind;      /* pixel kind of the finest level */
  double scale, phase;     /* 2^finest and (scale - 1) / 2 */
  int32_t finest, top;
  int32_t orig_ref;        /* unrotated patch at the finest level */
  double level_orow, level_ocol;
  const int32_t* refs0;    /* level-0 query patches (device) for the recovery search */
  int32_t n_refs0;
  double step, angle_start, angle_end, min_score;
  int32_t subpixel, rotate, absolute;
  /* find_models' query-cache rules (matcher.cpp:912-927, 984-1055); all zero
   * or negative for find: */
  int32_t recover_first;   /* a candidate below min_score is recovered before its fit */
  int32_t prefix_recovery; /* recovery windows in the row-prefix form on the whole level 0 */
  double guard_first;      /* background guard after fit iteration 0 (< 0: off) */
  double guard;            /* background guard after fit iteration 1 (< 0: off) */
} NcchipRefineParams;

/* Continuous bilinear score at the finest level of every candidate of a
 * batch (list form as ncchip_track: frame f's at in + f*cap, in_count[f]
 * of them), written to the same slot of out_full; non-finite scores are
 * kept.  `grid` workgroups loop over the candidates.  Replaces
 * matcher.cpp:813-822 (kernels_avx2.cpp:327-477, 517-541). */
int ncchip_full_score(const NcchipRefineParams* p, const NccPatchPool* pools, int n_frames, const NccMatch* in,
                      const int32_t* in_count, int cap, NccMatch* out_full, int grid, void* stream);

/* refine_candidate of every given full-scored candidate (Gauss-Newton fit
 * with backtracking, and the level-0 recovery search), one workgroup per
 * candidate; outputs the refined pose and the verdict (score >= MinScore
 * inside the domain).  Replaces matcher.cpp:853-1061, with find_models'
 * recover-first order, prefix-form recovery windows and background guard
 * when the params ask for them. */
int ncchip_refine(const NcchipRefineParams* p, const NccPatchPool* pools, const NccMatch* cand,
                  const int32_t* cand_frame, int n_cand, NccMatch* out_refined, int32_t* verdict, void* stream);

#ifdef __cplusplus
}
#endif
#endif
