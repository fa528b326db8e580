// Device helpers shared by the NCC kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "device_types.h"

#define NCC_MISSING (-__builtin_inf())

namespace ncc {

constexpr int kWave = 64;

// Integer pixel value (u8 at level 0, block sum S_l above) at (r, c).
__device__ __forceinline__ uint32_t load_s(const NccLevel& L, int r, int c) {
  const size_t i = size_t(r) * size_t(L.pitch) + size_t(c);
  if (L.kind == NCC_U8) return static_cast<const uint8_t*>(L.data)[i];
  if (L.kind == NCC_U16) return static_cast<const uint16_t*>(L.data)[i];
  return static_cast<const uint32_t*>(L.data)[i];
}

// The reference's FP64 level value S * 4^-l - dc (exact in FP64).
__device__ __forceinline__ double load_q(const NccLevel& L, int r, int c, double dc) {
  return ldexp(double(load_s(L, r, c)), -L.shift) - dc;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int m = kWave / 2; m > 0; m >>= 1) v += __shfl_xor(v, m, kWave);
  return v;
}
__device__ __forceinline__ int64_t wave_sum_i64(int64_t v) {
#pragma unroll
  for (int m = kWave / 2; m > 0; m >>= 1) v += __shfl_xor(v, m, kWave);
  return v;
}

// Block-wide sum of N doubles (blockDim.x a multiple of 64, at most 1024);
// every thread gets the totals, summed over waves in a fixed order.
// `scratch` holds at least 16*N doubles of LDS.
template <int N>
__device__ __forceinline__ void block_sum(double (&v)[N], double* scratch) {
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave, n_waves = blockDim.x / kWave;
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] = wave_sum(v[i]);
  if (lane == 0)
#pragma unroll
    for (int i = 0; i < N; ++i) scratch[wave * N + i] = v[i];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < N; ++i) {
    double t = 0;
    for (int w = 0; w < n_waves; ++w) t += scratch[w * N + i];
    v[i] = t;
  }
  __syncthreads();
}

__device__ __forceinline__ int block_any(int flag, int* scratch) {
  if (threadIdx.x == 0) *scratch = 0;
  __syncthreads();
  if (flag) atomicOr(scratch, 1);
  __syncthreads();
  const int r = *scratch;
  __syncthreads();
  return r;
}

__device__ __forceinline__ double clamp1(double x) { return x < -1.0 ? -1.0 : (x > 1.0 ? 1.0 : x); }

// Raster references are (pool << 24) | index with pool 0 the model's rasters
// and pool 1 the per-find rasters; the pools are two separate values so that
// selecting one never indexes a private array.
struct Pools {
  NccPatchPool p0, p1;
  __device__ __forceinline__ const NccPatchPool& of(int32_t ref) const { return (ref >> 24) ? p1 : p0; }
  __device__ __forceinline__ const NccPatch& patch(int32_t ref) const { return of(ref).patches[ref & 0xffffff]; }
};

// Candidate lists of a batch: frame f's entries at f * cap, counts[f] of
// them (clamped to cap).  A block's prologue builds the exclusive offsets of
// the per-frame counts in LDS so that a flat index g in [0, off[n]) maps to
// (frame, entry).  Every thread of the block calls it.
__device__ __forceinline__ void list_offsets(const int32_t* counts, int n_frames, int cap, int* off) {
  // the counts in one round of loads, then a serial scan in LDS
  for (int f = threadIdx.x; f < n_frames; f += blockDim.x) off[f + 1] = min(max(counts[f], 0), cap);
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int f = 0; f < n_frames; ++f) { const int c = off[f + 1]; off[f] = t; t += c; }
    off[n_frames] = t;
  }
  __syncthreads();
}
__device__ __forceinline__ int list_frame(const int* off, int n_frames, int g) {
  int lo = 0, hi = n_frames - 1;  // last f with off[f] <= g
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= g) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ int sample_row(int32_t rc) { return rc >> 16; }
__device__ __forceinline__ int sample_col(int32_t rc) { return int(int16_t(rc & 0xffff)); }

// The frame's reference pixel dc (u8 or u16 search image).
__device__ __forceinline__ uint32_t frame_dc(const NccFrame& F) {
  return F.dc_wide ? uint32_t(*static_cast<const uint16_t*>(F.dc_src)) : uint32_t(*static_cast<const uint8_t*>(F.dc_src));
}

// Domain pixels of a frame in image rows r0..r1 x columns c0..c1 (inclusive),
// from its summed-area table over the crop.
__device__ __forceinline__ uint32_t dom_count(const NccFrame& F, int64_t r0, int64_t c0, int64_t r1, int64_t c1) {
  const int64_t h = F.levels[0].height, w = F.levels[0].width;
  const int64_t cr = int64_t(F.crop_r), cc = int64_t(F.crop_c);
  r0 = max(r0 - cr, int64_t(0)); r1 = min(r1 - cr, h - 1);
  c0 = max(c0 - cc, int64_t(0)); c1 = min(c1 - cc, w - 1);
  if (r0 > r1 || c0 > c1) return 0;
  const uint32_t* S = F.dom_sat;
  const size_t p = size_t(w) + 1;
  return S[size_t(r1 + 1) * p + size_t(c1 + 1)] - S[size_t(r0) * p + size_t(c1 + 1)] -
         S[size_t(r1 + 1) * p + size_t(c0)] + S[size_t(r0) * p + size_t(c0)];
}

// domain_contains (matcher.cpp:249-251): the pose's nearest pixel is a domain
// pixel.
__device__ __forceinline__ bool in_domain(const NccFrame& F, double row, double col) {
  const double r = floor(row + .5), c = floor(col + .5);
  if (!F.dom_sat) return r >= 0 && c >= 0 && r < double(F.image_h) && c < double(F.image_w);
  if (r < -1e9 || r > 1e9 || c < -1e9 || c > 1e9) return false;
  return dom_count(F, int64_t(r), int64_t(c), int64_t(r), int64_t(c)) != 0;
}

// may_contain (matcher.cpp:506-534) for a reduced domain: at spacing 1 the
// pose's pixel; otherwise the pose within 2*spacing of the bounding box and,
// for a sparse domain, a domain pixel in the rows and columns within
// 2*spacing of the pose (a run overlapping [col - m, col + m] holds such a
// pixel because the run ends are integers and 2m >= 1).
__device__ __forceinline__ bool may_contain(const NccFrame& F, double row, double col, double spacing) {
  if (spacing == 1) return in_domain(F, row, col);
  const double m = 2 * spacing;
  if (row < F.dom_box[0] - m || row > F.dom_box[2] + m || col < F.dom_box[1] - m || col > F.dom_box[3] + m)
    return false;
  if (!F.dom_sparse) return true;
  return dom_count(F, int64_t(ceil(row - m)), int64_t(ceil(col - m)), int64_t(floor(row + m)),
                   int64_t(floor(col + m))) != 0;
}

// Centred integer moments t1 = sum(S - D), t2 = sum((S - D)^2) from the
// plain sums of S and S^2.  Computed modulo 2^64, so they are exact whenever
// t2 < 2^63, also when sum(S^2) itself wraps (u16 images at high levels).
__device__ __forceinline__ void centred_moments(int64_t sum_s, int64_t sum_s2, int n, int64_t D, int64_t& t1,
                                                int64_t& t2) {
  const uint64_t s = uint64_t(sum_s), d = uint64_t(D), nn = uint64_t(n);
  t1 = int64_t(s - nn * d);
  t2 = int64_t(uint64_t(sum_s2) - 2 * d * s + nn * d * d);
}

// Exact window moments from integer sums of S over n samples: returns the
// reference's sum(q) and sum(q^2) with q = S*4^-l - dc (exact in FP64 while
// |t1|, t2 < 2^53).
__device__ __forceinline__ void exact_moments(int64_t sum_s, int64_t sum_s2, int n, int64_t D, int shift,
                                              double& sum_q, double& sq_q) {
  int64_t t1, t2;
  centred_moments(sum_s, sum_s2, n, D, t1, t2);
  sum_q = ldexp(double(t1), -shift);
  sq_q = ldexp(double(t2), -2 * shift);
}

// FP64 normalisation of one window from its dot product with the centred
// pixels.  dot_c = sum_i wf_i * (S_i - C) (any precision), wsum_f = sum_i wf_i.
// The reference's numerator is sum_i w_i q_i - (sum q / n) * wsum with
// q = S 4^-l - dc; the window moments are exact.
__device__ __forceinline__ double normalise(double dot_c, double C, double wsum_f, double wsum, int n, int64_t sum_s,
                                            int64_t sum_s2, int64_t D, int shift, double dc, bool exact_lattice,
                                            bool absolute) {
  double sum, sq;
  exact_moments(sum_s, sum_s2, n, D, shift, sum, sq);
  const double cnt = double(n);
  const double var = sq - sum * sum / cnt;
  const double dot_q = ldexp(dot_c + C * wsum_f, -shift) - dc * wsum;  // sum_i w_i q_i
  const double num = dot_q - sum / cnt * wsum;
  if (!(var > 128 * __DBL_EPSILON__ * sq)) {
    // Low-variance rule (kernels_avx2.cpp:128-129, 166-167): -inf on the exact
    // byte lattice; elsewhere the reference recomputes directly, which only
    // differs from -inf when the window is not constant.
    int64_t t1, t2;
    centred_moments(sum_s, sum_s2, n, D, t1, t2);
    const bool constant = (__int128)t2 * n == (__int128)t1 * t1;
    if (exact_lattice || constant) return NCC_MISSING;
    const double v = ldexp(double((__int128)t2 * n - (__int128)t1 * t1), -2 * shift) / cnt;
    const double s = clamp1(num / sqrt(v));
    return absolute ? fabs(s) : s;
  }
  const double s = clamp1(num / sqrt(var));
  return absolute ? fabs(s) : s;
}

}  // namespace ncc
