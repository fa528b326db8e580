// Plain-C descriptors shared by the host driver and the HIP kernels.
// Every field is a plain integer, floating-point value or raw device pointer.
#ifndef MVKIT_AMD_DEVICE_TYPES_H
#define MVKIT_AMD_DEVICE_TYPES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

// Pixel storage of one search-pyramid level.  Level 0 is the caller's u8
// image; level l >= 1 stores the exact integer block sum S_l of the 4^l
// source pixels (u16 while it fits, i.e. l <= 4, u32 beyond).  The value the
// reference computes in FP64 (matcher.cpp:436-454) is S_l * 4^-l - dc.
enum { NCC_U8 = 0, NCC_U16 = 1, NCC_U32 = 2 };

typedef struct NccLevel {
  const void* data;   // device pointer
  int32_t kind;       // NCC_U8 / NCC_U16 / NCC_U32
  int32_t pitch;      // elements per row
  int32_t width, height;
  int32_t shift;      // 2*l: value = S * 2^-shift - dc
  int32_t pad;
} NccLevel;

// One rotated template raster on the device (model.hpp:30-43).  Sample
// coordinates are packed as (row << 16) | (col & 0xffff) in `rc`.
typedef struct NccPatch {
  double angle, orow, ocol, wsum;
  double wsum_f;             // sum of the fp32 weights (exact in FP64)
  int32_t n, rmin, rmax, cmin, cmax;
  int32_t sample_off;        // into rc / w / wf
  int32_t run_off, n_runs;   // into runs (int4: row, begin, end, 0)
  int32_t wrun_off, n_wruns; // into wruns (int4) / wrun_w (double)
  int32_t use_wruns;         // 2*#weight_runs < #samples (kernels_avx2.cpp:150, 175, 211)
  // Split-BF16 dense raster for the MFMA tracking kernel (element offset
  // into dense_hi / dense_lo, -1: none): rows rmin..rmax, columns
  // cmin-8 .. cmax+8 (zero padded, dense_w a multiple of 8), stored twice,
  // the second copy shifted left by one element, so any 8-element run is
  // one 16-byte load at a 4-byte aligned address of copy (start mod 2).
  int32_t dense_off;
  int32_t dense_h, dense_w;
  int32_t pad;
  // |sum w x - split-BF16 FP32 dot| <= (4^l - 1) err_lo + max|x| err_x for
  // x = S - dc 4^l split as in the tracking kernel (hi part multiple of 4^l,
  // lo part in [0, 4^l)): err_lo = sum|w_lo| + g sum|w_hi|,
  // err_x = sum|w - w_hi - w_lo| + g sum(|w_hi| + |w_lo|), g = k u / (1 - k u),
  // k = 3n + 8 FP32 additions, u = 2^-23.
  double err_lo, err_x;
} NccPatch;

// Device-resident pool of patches (all pointers are device pointers).
typedef struct NccPatchPool {
  const NccPatch* patches;
  const int32_t* rc;
  const double* w;
  const float* wf;
  const int32_t* runs;      // 4 ints per run
  const int32_t* wruns;     // 4 ints per weight run
  const double* wrun_w;
  const uint16_t* dense_hi; // dense rasters (NccPatch::dense_off), bf16 bits of w
  const uint16_t* dense_lo; // bf16 bits of w - bf16(w)
} NccPatchPool;

// A pose in the reference's internal (centroid) convention (model.hpp:26).
typedef struct NccMatch { double row, col, angle, score; } NccMatch;

// Everything a per-candidate kernel needs to know about the candidate's
// frame: its search pyramid, reference pixel, level-0 storage rectangle,
// crop origin and search domain.
// The front-end fields point at the frame's top-level statistics and score
// planes, written by the batched pyramid / prefix / coarse kernels.
#define NCC_MAX_LEVELS 16
#define NCC_MAX_BATCH 256
typedef struct NccFrame {
  NccLevel levels[NCC_MAX_LEVELS];  // level 0: the caller's u8 / u16 image (crop origin); above: workspace
  const void* dc_src;        // reference pixel dc (first crop pixel), device memory
  int32_t dc_wide;           // dc_src points to a u16 pixel (else u8)
  int32_t storage0[4];       // level-0 valid rectangle {row, col, height, width}
  double crop_r, crop_c;
  int32_t image_h, image_w;  // search domain (full image)
  uint64_t* p1;              // top-level row prefix of S, (w+1) per row
  uint64_t* p2;              // top-level row prefix of S^2
  uint32_t* nz;              // top-level row prefix of [S != D]
  double* planes;            // coarse score planes [angle][h][w] (split-BF16 form: [h][w][angle])
  // split-BF16 form: per top-level position, ceil(angles / 32) words with a
  // bit per angle scoring at least the peak threshold; only those scores are
  // written to the planes (null: every plane entry is written)
  uint32_t* peak_bits;
  const int32_t* ranges;     // per coarse angle: valid anchor range {rb, re, cb, ce}
  // Reduced search domain (image.hpp:108-113); dom_sat null: the full image.
  // dom_sat: summed-area table of the domain mask over the crop, (level-0
  // height + 1) x (width + 1) u32, entry (r, c) = domain pixels in crop rows
  // < r and columns < c.  dom_runs: the domain's runs as int4 {row, first
  // column, last column, 0} in image coordinates, dom_row[r] the index of
  // crop row r's first run (level-0 height + 1 entries).
  uint32_t* dom_sat;
  const int32_t* dom_runs;
  const int32_t* dom_row;
  int32_t dom_box[4];        // bounding box {row_min, col_min, row_max, col_max}
  int32_t dom_sparse;        // area < 0.7 x bounding box area (matcher.cpp:505)
  int32_t dom_pad;
} NccFrame;

// Model construction on the device (create / rebuild, model.cpp:39-98,
// 173-242).  One template pyramid level: centred values and mask, h x w.
typedef struct NccTemplateLevel {
  const double* gray;
  const uint8_t* mask;
  int32_t w, h;
  double orow, ocol;         // level centroid
} NccTemplateLevel;

// One raster to build: level, sampling grid (the level grid when unrotated,
// else the rotated bounding box of offsets from the centroid) and the
// angle's cosine and sine from the host's libm; prepared: the level-0
// unrotated raster, centred by the template's mean and energy.
typedef struct NccRasterJob {
  double ca, sa;
  double mean, energy;
  int32_t level, unrotated, prepared, pad;
  int32_t r_lo, r_hi, c_lo, c_hi;
  int64_t sample_off;        // first sample in the output arrays
  int64_t run_off, wrun_off; // first row run / weight run
} NccRasterJob;

// What the host needs back of one built raster.
typedef struct NccRasterStats {
  int64_t count;             // samples
  int32_t status;            // 0: usable, 1: degenerate (fewer than 4 samples or no energy)
  int32_t rmin, rmax, cmin, cmax;
  int32_t n_runs, n_wruns;
  int32_t pad;
  double wsum, wsum_f;
} NccRasterStats;

#ifdef __cplusplus
}
#endif
#endif
