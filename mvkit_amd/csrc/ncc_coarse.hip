// Coarse exhaustive NCC at the top pyramid level (matcher.cpp:542-594) and
// its per-angle 3x3 peak extraction with candidate compaction.
//
// Correlation forms:
//  * FP32: the dot product sum_i w_i * (S_i - C) in FP32 FMA from an LDS
//    image tile (the "plain FP32" GPU form);
//  * split-BF16 MFMA lives in ncc_mfma.hip and writes the same score planes.
// Both feed the same FP64 normalisation, where the window moments come from
// exact integer row prefixes (u32 sum of S, u64 sum of S^2), so the variance
// is the reference's value bit for bit (kernels_avx2.cpp:18-34, 145-171).
#include "ncc_common.hip.h"
#include "../../include/ncchip.h"

namespace ncc {

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
    const int64_t t1 = sum_s - int64_t(n) * D;
    const int64_t t2 = sum_s2 - 2 * D * sum_s + int64_t(n) * D * D;
    const bool constant = (__int128)t2 * n == (__int128)t1 * t1;
    if (exact_lattice || constant) return NCC_MISSING;
    const double v = ldexp(double((__int128)t2 * n - (__int128)t1 * t1), -2 * shift) / cnt;
    const double s = clamp1(num / sqrt(v));
    return absolute ? fabs(s) : s;
  }
  const double s = clamp1(num / sqrt(var));
  return absolute ? fabs(s) : s;
}

// Exact window moments over a patch's runs from the row prefixes.
__device__ __forceinline__ void window_moments(const uint32_t* __restrict__ p1, const uint64_t* __restrict__ p2,
                                               int pitch, const int32_t* __restrict__ runs, int n_runs, int r, int c,
                                               int64_t& s1, int64_t& s2) {
  uint32_t a1 = 0;
  uint64_t a2 = 0;
  for (int k = 0; k < n_runs; ++k) {
    const int4 run = reinterpret_cast<const int4*>(runs)[k];
    const size_t row = size_t(r + run.x) * pitch;
    a1 += p1[row + c + run.z + 1] - p1[row + c + run.y];
    a2 += p2[row + c + run.z + 1] - p2[row + c + run.y];
  }
  s1 = int64_t(a1);
  s2 = int64_t(a2);
}

struct CoarseArgs {
  NccLevel lv;
  const uint8_t* dc_src;   // reference pixel (first crop pixel), device memory
  double C;                // centring constant subtracted from S before the dot product
  const uint32_t* p1;      // row prefix of S, (w+1) per row
  const uint64_t* p2;      // row prefix of S^2
  Pools pools;
  const int32_t* patch_ids;  // coarse patches as (pool << 24) | index
  const int4* ranges;        // per coarse patch: valid anchor range [rb, re) x [cb, ce)
  double* planes;            // [A][h][w] scores, -inf outside the valid range
  int exact_lattice, absolute;
  int lds_w;               // LDS row length (floats)
};

constexpr int kTileR = 32, kTileC = 64, kColsPerThread = 8;

// FP32 form: one block per (64x32 anchor tile, coarse angle); each thread
// scores 8 consecutive anchors of one row.  The image tile holds the centred
// integers T = S - C.  Per template run, a sliding register window feeds the
// FP32 dot products of the 8 anchors, and the exact integer window moments
// sum(T) and sum(T^2) are accumulated for the first anchor and slid to the
// other seven with the run's boundary pixels.
template <class T>
__global__ void __launch_bounds__(256) k_coarse_fp32(const CoarseArgs a) {
  extern __shared__ int32_t tile[];
  const double dc = double(*a.dc_src);
  const int64_t D = int64_t(*a.dc_src) << a.lv.shift;
  const int ai = blockIdx.z;
  const NccPatchPool& pool = a.pools.of(a.patch_ids[ai]);
  const NccPatch P = a.pools.patch(a.patch_ids[ai]);
  const int4 rg = a.ranges[ai];
  const int W = a.lv.width, H = a.lv.height;
  const int r0 = blockIdx.y * kTileR, c0 = blockIdx.x * kTileC;
  const int LH = kTileR + P.rmax - P.rmin, LW = a.lds_w;
  const int Ci = int(a.C);
  constexpr int U = 8;  // staging loads issued together per thread
  for (int base = threadIdx.x; base < LH * LW; base += 256 * U) {
    int32_t v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = base + 256 * u;
      const int gr = r0 + P.rmin + i / LW, gc = c0 + P.cmin + i % LW;
      v[u] = (i < LH * LW && gr >= 0 && gr < H && gc >= 0 && gc < W)
                 ? int32_t(static_cast<const T*>(a.lv.data)[size_t(gr) * a.lv.pitch + gc]) - Ci
                 : 0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (base + 256 * u < LH * LW) tile[base + 256 * u] = v[u];
  }
  __syncthreads();
  const int ty = threadIdx.x / (kTileC / kColsPerThread), tx = (threadIdx.x % (kTileC / kColsPerThread)) * kColsPerThread;
  float acc[kColsPerThread];
  int64_t m1[kColsPerThread], m2[kColsPerThread];
#pragma unroll
  for (int j = 0; j < kColsPerThread; ++j) { acc[j] = 0.f; m1[j] = 0; m2[j] = 0; }
  const int32_t* runs = pool.runs + 4 * P.run_off;
  const float* wf = pool.wf + P.sample_off;
  int k = 0;
  for (int ri = 0; ri < P.n_runs; ++ri) {
    const int4 run = reinterpret_cast<const int4*>(runs)[ri];
    const int32_t* src = tile + (ty + run.x - P.rmin) * LW + (tx + run.y - P.cmin);
    const int len = run.z - run.y + 1;
    float win[16];
#pragma unroll
    for (int j = 0; j < 8; ++j) win[j] = float(src[j]);
    int64_t s1 = 0, s2 = 0;  // moments of the run for the first anchor
    int s = 0;
    for (; s + 8 <= len; s += 8) {
#pragma unroll
      for (int j = 0; j < 8; ++j) win[8 + j] = float(src[s + 8 + j]);
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const float w = wf[k + s + u];
        const int32_t t = int32_t(win[u]);
        s1 += t;
        s2 += int64_t(t) * t;
#pragma unroll
        for (int j = 0; j < kColsPerThread; ++j) acc[j] = fmaf(w, win[u + j], acc[j]);
      }
#pragma unroll
      for (int j = 0; j < 8; ++j) win[j] = win[8 + j];
    }
    for (; s < len; ++s) {
      const float w = wf[k + s];
      const int32_t t = int32_t(win[0]);
      s1 += t;
      s2 += int64_t(t) * t;
#pragma unroll
      for (int j = 0; j < kColsPerThread; ++j) acc[j] = fmaf(w, win[j], acc[j]);
#pragma unroll
      for (int j = 0; j < 7; ++j) win[j] = win[j + 1];
      win[7] = float(src[s + 8]);
    }
    // slide the run's moments to anchors 1..7: drop src[j-1], add src[len+j-1]
    m1[0] += s1;
    m2[0] += s2;
#pragma unroll
    for (int j = 1; j < kColsPerThread; ++j) {
      const int32_t out = src[j - 1], in = src[len + j - 1];
      s1 += in - out;
      s2 += int64_t(in) * in - int64_t(out) * out;
      m1[j] += s1;
      m2[j] += s2;
    }
    k += len;
  }
  const int r = r0 + ty;
  if (r >= H) return;
  double* out = a.planes + (size_t(ai) * H + r) * W;
  const int64_t Cl = Ci;
#pragma unroll
  for (int j = 0; j < kColsPerThread; ++j) {
    const int c = c0 + tx + j;
    if (c >= W) break;
    double score = NCC_MISSING;
    if (r >= rg.x && r < rg.y && c >= rg.z && c < rg.w) {
      // sums of S from sums of T = S - C
      const int64_t n = P.n;
      const int64_t sum_s = m1[j] + n * Cl;
      const int64_t sum_s2 = m2[j] + 2 * Cl * m1[j] + n * Cl * Cl;
      score = normalise(double(acc[j]), a.C, P.wsum_f, P.wsum, P.n, sum_s, sum_s2, D, a.lv.shift, dc, a.exact_lattice,
                        a.absolute);
    }
    out[c] = score;
  }
}

// Reference-order FP64 scores of every anchor (one thread per anchor); used
// when the LDS tile would not fit, and as the FP64 form for error reports.
template <class T>
__global__ void __launch_bounds__(256) k_coarse_fp64(const CoarseArgs a) {
  const double dc = double(*a.dc_src);
  const int64_t D = int64_t(*a.dc_src) << a.lv.shift;
  const int ai = blockIdx.z;
  const NccPatchPool& pool = a.pools.of(a.patch_ids[ai]);
  const NccPatch P = a.pools.patch(a.patch_ids[ai]);
  const int4 rg = a.ranges[ai];
  const int W = a.lv.width, H = a.lv.height;
  const int c = blockIdx.x * 256 + threadIdx.x, r = blockIdx.y;
  if (c >= W || r >= H) return;
  double score = NCC_MISSING;
  if (r >= rg.x && r < rg.y && c >= rg.z && c < rg.w) {
    const int32_t* rc = pool.rc + P.sample_off;
    const double* w = pool.w + P.sample_off;
    double dot = 0;
    const T* px = static_cast<const T*>(a.lv.data);
    const double unit = ldexp(1.0, -a.lv.shift);
    for (int i = 0; i < P.n; ++i)
      dot += (double(px[size_t(r + sample_row(rc[i])) * a.lv.pitch + c + sample_col(rc[i])]) * unit - dc) * w[i];
    int64_t s1, s2;
    window_moments(a.p1, a.p2, W + 1, pool.runs + 4 * P.run_off, P.n_runs, r, c, s1, s2);
    double sum, sq;
    exact_moments(s1, s2, P.n, D, a.lv.shift, sum, sq);
    const double cnt = double(P.n), var = sq - sum * sum / cnt;
    const double num = dot - sum / cnt * P.wsum;
    if (!(var > 128 * __DBL_EPSILON__ * sq)) score = NCC_MISSING;
    else { score = clamp1(num / sqrt(var)); if (a.absolute) score = fabs(score); }
  }
  a.planes[(size_t(ai) * H + r) * W + c] = score;
}

// Per-angle 3x3 peak extraction (matcher.cpp:580-594): keep finite v >= thr
// with no neighbour above it and no equal neighbour above or to the left.
__global__ void __launch_bounds__(256) k_peaks(const double* __restrict__ planes, int A, int H, int W, double thr,
                                               NccPeak* __restrict__ out, int cap, int* __restrict__ count) {
  const int c = blockIdx.x * 256 + threadIdx.x, r = blockIdx.y, ai = blockIdx.z;
  if (c >= W || r >= H || ai >= A) return;
  const double* pl = planes + size_t(ai) * H * W;
  const double v = pl[size_t(r) * W + c];
  if (v < thr || !isfinite(v)) return;
  bool mx = true;
  for (int dr = -1; dr <= 1 && mx; ++dr)
    for (int dc = -1; dc <= 1; ++dc) {
      const int rr = r + dr, cc = c + dc;
      if (rr < 0 || cc < 0 || rr >= H || cc >= W || (!dr && !dc)) continue;
      const double o = pl[size_t(rr) * W + cc];
      if (o > v || (o == v && (dr < 0 || (!dr && dc < 0)))) { mx = false; break; }
    }
  if (!mx) return;
  const int k = atomicAdd(count, 1);
  if (k < cap) {
    NccPeak p;
    p.angle = ai; p.row = r; p.col = c; p.pad = 0; p.score = v;
    out[k] = p;
  }
}

// Copies every frame's peak count and its first min(count, cap) peaks to
// host-visible memory (frame f's peaks at f * cap), so one stream
// synchronisation returns all frames' peaks.
__global__ void __launch_bounds__(256) k_export_peaks(const int32_t* __restrict__ counts,
                                                      const NccPeak* __restrict__ peaks, int cap,
                                                      int32_t* __restrict__ out_counts, NccPeak* __restrict__ out) {
  const int f = blockIdx.y;
  const int n = min(counts[f], cap);
  if (blockIdx.x == 0 && threadIdx.x == 0) out_counts[f] = counts[f];
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
    out[size_t(f) * cap + i] = peaks[size_t(f) * cap + i];
}

}  // namespace ncc

using namespace ncc;

extern "C" int ncchip_export_peaks(const int32_t* counts, const NccPeak* peaks, int n_frames, int cap,
                                   int32_t* host_counts, NccPeak* host_peaks, void* stream) {
  if (n_frames <= 0) return 0;
  hipLaunchKernelGGL(k_export_peaks, dim3(8, n_frames), dim3(256), 0, static_cast<hipStream_t>(stream), counts, peaks,
                     cap, host_counts, host_peaks);
  return hipGetLastError() == hipSuccess ? 0 : -1000;
}

extern "C" int ncchip_coarse(const NccLevel* lv, const uint8_t* dc_src, double C, const uint32_t* p1,
                             const uint64_t* p2, const NccPatchPool* pools, const int32_t* patch_ids,
                             const int32_t* ranges, int n_angles, int max_bbox_h, int max_bbox_w, int form,
                             int exact_lattice, int absolute, double* planes, void* stream) {
  CoarseArgs a;
  a.lv = *lv; a.dc_src = dc_src; a.C = C; a.p1 = p1; a.p2 = p2; a.pools.p0 = pools[0]; a.pools.p1 = pools[1]; a.patch_ids = patch_ids;
  a.ranges = reinterpret_cast<const int4*>(ranges); a.planes = planes; a.exact_lattice = exact_lattice;
  a.absolute = absolute;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int LH = kTileR + max_bbox_h;
  a.lds_w = ((kTileC + max_bbox_w + 16) + 3) & ~3;
  const size_t lds = size_t(LH) * a.lds_w * sizeof(int32_t);
  if (form == NCCHIP_FORM_FP32 && lds <= 64 * 1024) {
    const dim3 grid((lv->width + kTileC - 1) / kTileC, (lv->height + kTileR - 1) / kTileR, n_angles);
    if (lv->kind == NCC_U8) hipLaunchKernelGGL(k_coarse_fp32<uint8_t>, grid, dim3(256), lds, st, a);
    else if (lv->kind == NCC_U16) hipLaunchKernelGGL(k_coarse_fp32<uint16_t>, grid, dim3(256), lds, st, a);
    else hipLaunchKernelGGL(k_coarse_fp32<uint32_t>, grid, dim3(256), lds, st, a);
  } else {
    const dim3 grid((lv->width + 255) / 256, lv->height, n_angles);
    if (lv->kind == NCC_U8) hipLaunchKernelGGL(k_coarse_fp64<uint8_t>, grid, dim3(256), 0, st, a);
    else if (lv->kind == NCC_U16) hipLaunchKernelGGL(k_coarse_fp64<uint16_t>, grid, dim3(256), 0, st, a);
    else hipLaunchKernelGGL(k_coarse_fp64<uint32_t>, grid, dim3(256), 0, st, a);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1000;
}

extern "C" int ncchip_peaks(const double* planes, int n_angles, int h, int w, double threshold, NccPeak* out,
                            int cap, int32_t* count, void* stream) {
  hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 grid((w + 255) / 256, h, n_angles);
  hipLaunchKernelGGL(k_peaks, grid, dim3(256), 0, st, planes, n_angles, h, w, threshold, out, cap, count);
  return hipGetLastError() == hipSuccess ? 0 : -1000;
}
