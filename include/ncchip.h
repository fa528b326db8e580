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

/* Row prefix sums of S (u32) and S^2 (u64), (w+1) per row with a leading
 * zero, plus the support integral of (S != D) if `support` is non-null
 * ((w+1)*(h+1) u32).  `nz` is (w+1)*h u32 scratch.  Replaces the top-level
 * prefix statistics and support integral (matcher.cpp:487-503). */
int ncchip_row_prefix(const void* src, int kind, int pitch, int w, int h, const uint8_t* dc_src, int shift,
                      uint32_t* p1, uint64_t* p2, uint32_t* nz, uint32_t* support, void* stream);

/* Coarse exhaustive NCC over every anchor of the top level for each coarse
 * angle, writing FP64 score planes [n_angles][h][w] (-inf outside each
 * angle's valid anchor range).  `pools` points to two raster pools (the
 * model's and the per-find rasters); `patch_ids` are (pool << 24 | index).
 * `ranges` holds 4 ints per angle: row begin, row end, col begin, col end.
 * Replaces matcher.cpp:542-579 with kernels_avx2.cpp:105-238. */
int ncchip_coarse(const NccLevel* lv, const uint8_t* dc_src, double C, const uint32_t* p1, const uint64_t* p2,
                  const NccPatchPool* pools, const int32_t* patch_ids, const int32_t* ranges, int n_angles,
                  int max_bbox_h, int max_bbox_w, int form, int exact_lattice, int absolute, double* planes,
                  void* stream);

/* Split-BF16 / FP8-screen MFMA coarse search (same planes as ncchip_coarse).
 * `screen` > -inf enables the FP8-E4M3 screen: anchors whose screened score
 * is below it are written -inf, survivors are re-scored in split-BF16. */
int ncchip_coarse_mfma(const NccLevel* lv, const uint8_t* dc_src, double C, const uint32_t* p1, const uint64_t* p2,
                       const NccPatchPool* pools, const int32_t* patch_ids, const int32_t* ranges, int n_angles,
                       int bbox_h, int bbox_w, int exact_lattice, int absolute, double screen, double* planes,
                       void* stream);

/* Every kernel taking `dc_src` reads the reference pixel dc (the first
 * pixel of the search crop, matcher.cpp:410-413) from that device address;
 * integer levels use D = dc << shift. */

/* Per-angle 3x3 peak extraction with the reference's tie rule and atomic
 * compaction into `out` (unordered); `count` (zeroed by the caller) receives
 * the number of peaks found, which may exceed cap.
 * Replaces matcher.cpp:580-594. */
int ncchip_peaks(const double* planes, int n_angles, int h, int w, double threshold, NccPeak* out, int cap,
                 int32_t* count, void* stream);

/* Copies the peak counts of n_frames frames and each frame's first
 * min(count, cap) peaks (frame f's at peaks + f * cap) into host-visible
 * memory (pinned, coherent host allocations), laid out the same way. */
int ncchip_export_peaks(const int32_t* counts, const NccPeak* peaks, int n_frames, int cap, int32_t* host_counts,
                        NccPeak* host_peaks, void* stream);

/* Coarse-to-fine tracking of every candidate at one level: the first strict
 * maximum over (query patch within `tol` of the candidate's angle, 5x5
 * anchors).  Candidates may come from several frames: `frames` (device) holds
 * one NccFrame per frame and `cand_frame` (device) each candidate's frame;
 * `level` is the tracked level and `level_kind` its pixel kind.  `refs` lists
 * the level's query patches as (pool << 24 | index).  `tile_pixels` bounds
 * the pixels of a candidate's neighbourhood (the union of its windows),
 * staged in LDS when it fits in 64 KiB (0: read the level from HBM);
 * exact_f64 != 0 asserts n * (max |S - D|)^2 < 2^53 for every query patch,
 * which lets the moments accumulate exactly in FP64.
 * Replaces matcher.cpp:670-756 (kernels_avx2.cpp:118-238). */
int ncchip_track(const NccFrame* frames, const int32_t* cand_frame, int level, int level_kind,
                 const NccPatchPool* pools, const int32_t* refs, int n_refs, const NccMatch* cand, NccMatch* out,
                 int n_cand, double tol, double step, double offset, int exact_lattice, int absolute,
                 int tile_pixels, int exact_f64, void* stream);

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
} NcchipRefineParams;

/* Continuous bilinear score of every candidate at the finest level of its
 * frame, one workgroup per candidate.  Replaces matcher.cpp:813-822
 * (kernels_avx2.cpp:327-477, 517-541). */
int ncchip_full_score(const NcchipRefineParams* p, const NccPatchPool* pools, const NccMatch* cand,
                      const int32_t* cand_frame, int n_cand, NccMatch* out_full, void* stream);

/* refine_candidate of every given full-scored candidate (Gauss-Newton fit
 * with backtracking, and the level-0 recovery search), one workgroup per
 * candidate; outputs the refined pose and the verdict (score >= MinScore
 * inside the domain).  Replaces matcher.cpp:853-1061. */
int ncchip_refine(const NcchipRefineParams* p, const NccPatchPool* pools, const NccMatch* cand,
                  const int32_t* cand_frame, int n_cand, NccMatch* out_refined, int32_t* verdict, void* stream);

#ifdef __cplusplus
}
#endif
#endif
