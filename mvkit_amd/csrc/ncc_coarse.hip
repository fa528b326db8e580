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

// Exact window moments over a patch's runs from the row prefixes.
__device__ __forceinline__ void window_moments(const uint64_t* __restrict__ p1, const uint64_t* __restrict__ p2,
                                               int pitch, const int32_t* __restrict__ runs, int n_runs, int r, int c,
                                               int64_t& s1, int64_t& s2) {
  uint64_t a1 = 0;
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
  const NccFrame* frames;  // per-frame descriptors (device)
  int level;               // the top level
  int n_angles;            // coarse angles; blockIdx.z = frame * n_angles + angle
  double C;                // centring constant subtracted from S before the dot product
  Pools pools;
  const int32_t* patch_ids;  // coarse patches as (pool << 24) | index
  int exact_lattice, absolute;
  int lds_w;               // LDS row length (floats)
};

// The frame-dependent part of a coarse block's work.
struct CoarseFrame {
  NccLevel lv;
  uint32_t dc;
  const uint64_t* p1;
  const uint64_t* p2;
  int4 rg;
  double* planes;
};
__device__ __forceinline__ CoarseFrame coarse_frame(const CoarseArgs& a, int f, int ai) {
  const NccFrame& F = a.frames[f];
  CoarseFrame c;
  c.lv = F.levels[a.level];
  c.dc = frame_dc(F);
  c.p1 = F.p1;
  c.p2 = F.p2;
  c.rg = reinterpret_cast<const int4*>(F.ranges)[ai];
  c.planes = F.planes;
  return c;
}

constexpr int kTileR = 32, kTileC = 64, kColsPerThread = 8;

// FP32 form: one block per (64x32 anchor tile, coarse angle); each thread
// scores 8 consecutive anchors of one row.  The image tile holds the centred
// integers T = S - C.  Per template run, a sliding register window feeds the
// FP32 dot products of the 8 anchors, and the exact integer window moments
// sum(T) and sum(T^2) are accumulated for the first anchor and slid to the
// other seven with the run's boundary pixels.
template <class T>
__global__ void __launch_bounds__(256) k_coarse_fp32(const CoarseArgs args) {
  extern __shared__ int32_t tile[];
  const int f = blockIdx.z / args.n_angles, ai = blockIdx.z % args.n_angles;
  const CoarseFrame a = coarse_frame(args, f, ai);
  const int W = a.lv.width, H = a.lv.height;
  const int r0 = blockIdx.y * kTileR, c0 = blockIdx.x * kTileC;
  if (r0 >= H || c0 >= W) return;
  const double dc = double(a.dc);
  const int64_t D = int64_t(a.dc) << a.lv.shift;
  const NccPatchPool& pool = args.pools.of(args.patch_ids[ai]);
  const NccPatch P = args.pools.patch(args.patch_ids[ai]);
  const int4 rg = a.rg;
  const int LH = kTileR + P.rmax - P.rmin, LW = args.lds_w;
  const int Ci = int(args.C);
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
      // sums of S from sums of T = S - C (modulo 2^64, see centred_moments)
      const uint64_t n = uint64_t(P.n), cu = uint64_t(Cl), t1 = uint64_t(m1[j]);
      const int64_t sum_s = int64_t(t1 + n * cu);
      const int64_t sum_s2 = int64_t(uint64_t(m2[j]) + 2 * cu * t1 + n * cu * cu);
      score = normalise(double(acc[j]), args.C, P.wsum_f, P.wsum, P.n, sum_s, sum_s2, D, a.lv.shift, dc,
                        args.exact_lattice, args.absolute);
    }
    out[c] = score;
  }
}

// Reference-order FP64 scores of every anchor (one thread per anchor); used
// when the LDS tile would not fit, and as the FP64 form for error reports.
template <class T>
__global__ void __launch_bounds__(256) k_coarse_fp64(const CoarseArgs args) {
  const int f = blockIdx.z / args.n_angles, ai = blockIdx.z % args.n_angles;
  const CoarseFrame a = coarse_frame(args, f, ai);
  const int W = a.lv.width, H = a.lv.height;
  const int c = blockIdx.x * 256 + threadIdx.x, r = blockIdx.y;
  if (c >= W || r >= H) return;
  const double dc = double(a.dc);
  const int64_t D = int64_t(a.dc) << a.lv.shift;
  const NccPatchPool& pool = args.pools.of(args.patch_ids[ai]);
  const NccPatch P = args.pools.patch(args.patch_ids[ai]);
  const int4 rg = a.rg;
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
    else { score = clamp1(num / sqrt(var)); if (args.absolute) score = fabs(score); }
  }
  a.planes[(size_t(ai) * H + r) * W + c] = score;
}

// Per-angle 3x3 peak extraction (matcher.cpp:580-594): keep finite v >= thr
// with no neighbour above it and no equal neighbour above or to the left.
// Batched over frames (blockIdx.z = frame * A + angle).  A peak becomes a
// candidate pose in the reference's internal convention (matcher.cpp:590-592)
// appended to its frame's list (unordered; `key` keeps the angle index so the
// reference's (angle, row, col) emission order can be restored), and widens
// the frame's candidate extent (row and column offsets from the crop origin,
// non-negative doubles kept as ordered bit patterns) used for the level-0
// storage region.
struct PeakArgs {
  const NccFrame* frames;
  int level, n_angles;
  Pools pools;
  const int32_t* patch_ids;
  double thr, scale, phase;
  NccMatch* cand;        // frame f's candidates at f * cap
  int32_t* key;          // angle index of each candidate
  int cap;
  int32_t* count;        // per frame (may exceed cap)
  unsigned long long* extent;  // per frame: min row, max row, min col, max col (FP64 bits)
  int hwa;               // planes in [row][col][angle] order (else [angle][row][col])
};

constexpr int kPeakRows = 8;  // plane rows per block

__global__ void __launch_bounds__(256) k_peaks(const PeakArgs a) {
  // [angle][row][col] planes: blockIdx.z = frame * A + angle, threads over
  // columns; [row][col][angle] planes: blockIdx.z = frame, threads over the
  // (column, angle) pairs of a row, angles innermost (coalesced either way)
  const int f = a.hwa ? blockIdx.z : blockIdx.z / a.n_angles;
  const NccFrame& F = a.frames[f];
  const int H = F.levels[a.level].height, W = F.levels[a.level].width;
  const int e = blockIdx.x * 256 + threadIdx.x;
  const int c = a.hwa ? e / a.n_angles : e;
  const int ai = a.hwa ? e % a.n_angles : blockIdx.z % a.n_angles;
  if (int(blockIdx.x) * 256 >= (a.hwa ? W * a.n_angles : W)) return;  // uniform per block
  // element (r, c) of this angle's plane
  const size_t sr = a.hwa ? size_t(W) * a.n_angles : size_t(W), sc = a.hwa ? size_t(a.n_angles) : 1;
  const double* pl = F.planes + (a.hwa ? size_t(ai) : size_t(ai) * H * W);
  for (int r = blockIdx.y * kPeakRows; r < min(H, int(blockIdx.y + 1) * kPeakRows); ++r) {
  bool peak = false;
  double v = 0;
  if (c < W) {
    v = pl[r * sr + c * sc];
    if (v >= a.thr && isfinite(v)) {
      peak = true;
      for (int dr = -1; dr <= 1 && peak; ++dr)
        for (int dc = -1; dc <= 1; ++dc) {
          const int rr = r + dr, cc = c + dc;
          if (rr < 0 || cc < 0 || rr >= H || cc >= W || (!dr && !dc)) continue;
          const double o = pl[rr * sr + cc * sc];
          if (o > v || (o == v && (dr < 0 || (!dr && dc < 0)))) { peak = false; break; }
        }
    }
  }
  // one atomic per wave for the slots and the extent
  const uint64_t m = __ballot(peak);
  if (!m) continue;
  const int lane = threadIdx.x & 63, leader = __builtin_ctzll(m);
  int base = 0;
  if (lane == leader) base = atomicAdd(a.count + f, __popcll(m));
  base = __shfl(base, leader, 64);
  double orow = __builtin_inf(), ocol = __builtin_inf(), xrow = 0, xcol = 0;
  if (peak) {
    const int k = base + __popcll(m & ((uint64_t(1) << lane) - 1));
    const NccPatch& P = a.pools.patch(a.patch_ids[ai]);
    NccMatch q;
    const double dr = a.scale * (r + P.orow) + a.phase, dcol = a.scale * (c + P.ocol) + a.phase;
    q.row = F.crop_r + dr;
    q.col = F.crop_c + dcol;
    q.angle = P.angle;
    q.score = v;
    if (k < a.cap) {
      a.cand[size_t(f) * a.cap + k] = q;
      a.key[size_t(f) * a.cap + k] = ai;
      // offsets from the crop origin, as the host computes them (row - crop_r)
      orow = xrow = q.row - F.crop_r;
      ocol = xcol = q.col - F.crop_c;
    }
  }
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) {
    orow = fmin(orow, __shfl_xor(orow, s, 64));
    ocol = fmin(ocol, __shfl_xor(ocol, s, 64));
    xrow = fmax(xrow, __shfl_xor(xrow, s, 64));
    xcol = fmax(xcol, __shfl_xor(xcol, s, 64));
  }
  if (lane == leader && isfinite(orow)) {
    unsigned long long* e = a.extent + 4 * size_t(f);
    atomicMin(e + 0, __double_as_longlong(orow));
    atomicMax(e + 1, __double_as_longlong(xrow));
    atomicMin(e + 2, __double_as_longlong(ocol));
    atomicMax(e + 3, __double_as_longlong(xcol));
  }
  }
}

// Peaks of the sparse planes of the split-BF16 form: only scores at or above
// the threshold are written, each with its bit in the frame's peak_bits
// (per position, one word per 32 angles).  Windows below the threshold can
// neither be peaks nor suppress one, so the 3x3 rule only needs the flagged
// neighbours.  One thread per column over the kPeakBitRows rows of its block;
// a lane may emit several angles of its positions.  A wave claims its slots
// in the frame's list with one atomic per word of angles over all its rows,
// and folds the extent of its candidates into the frame's with one set of
// atomics at the end (all rows of a frame meet on those few addresses).
constexpr int kPeakBitRows = 8;
__global__ void __launch_bounds__(256) k_peaks_bits(const PeakArgs a) {
  const int f = blockIdx.z;
  const NccFrame& F = a.frames[f];
  const int H = F.levels[a.level].height, W = F.levels[a.level].width;
  if (int(blockIdx.x) * 256 >= W) return;  // uniform per block
  const int c = blockIdx.x * 256 + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const int nw = (a.n_angles + 31) / 32;
  const size_t A = size_t(a.n_angles);
  const int r0 = blockIdx.y * kPeakBitRows;
  const int r_end = min(H, r0 + kPeakBitRows);
  double orow = __builtin_inf(), ocol = __builtin_inf(), xrow = 0, xcol = 0;
  for (int w = 0; w < nw; ++w) {
    uint32_t words[kPeakBitRows], peaks[kPeakBitRows];
#pragma unroll
    for (int i = 0; i < kPeakBitRows; ++i)
      words[i] = (c < W && r0 + i < r_end) ? F.peak_bits[(size_t(r0 + i) * W + c) * nw + w] : 0u;
    int cnt = 0;
#pragma unroll
    for (int i = 0; i < kPeakBitRows; ++i) {
      const int r = r0 + i;
      peaks[i] = 0;
      const uint32_t bits = words[i];
      if (bits) {
        const size_t pos = size_t(r) * W + c;
        uint32_t nb[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) {
          const int rr = r + k / 3 - 1, cc = c + k % 3 - 1;
          nb[k] = (k != 4 && rr >= 0 && cc >= 0 && rr < H && cc < W) ? F.peak_bits[(size_t(rr) * W + cc) * nw + w] : 0u;
        }
        uint32_t pend = bits;
        while (pend) {
          const int s = __builtin_ctz(pend);
          pend &= pend - 1;
          const double v = F.planes[pos * A + size_t(w) * 32 + s];
          bool peak = true;
          for (int k = 0; k < 9 && peak; ++k) {
            if (!((nb[k] >> s) & 1u)) continue;
            const int dr = k / 3 - 1, dc = k % 3 - 1;
            const double o = F.planes[(size_t(r + dr) * W + (c + dc)) * A + size_t(w) * 32 + s];
            if (o > v || (o == v && (dr < 0 || (!dr && dc < 0)))) peak = false;
          }
          if (peak) peaks[i] |= 1u << s;
        }
      }
      cnt += __popc(peaks[i]);
    }
    // emission: a wave-wide exclusive scan of the lanes' peak counts
    int incl = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int t = __shfl_up(incl, d, 64);
      if (lane >= d) incl += t;
    }
    const int total = __shfl(incl, 63, 64);
    if (!total) continue;  // wave-uniform
    int base = 0;
    if (lane == 63) base = atomicAdd(a.count + f, total);
    base = __shfl(base, 63, 64);
    int k = base + incl - cnt;
#pragma unroll
    for (int i = 0; i < kPeakBitRows; ++i) {
      const int r = r0 + i;
      const size_t pos = size_t(r) * W + c;
      uint32_t pm = peaks[i];
      while (pm) {
        const int s = __builtin_ctz(pm);
        pm &= pm - 1;
        const int ai = w * 32 + s;
        const NccPatch& P = a.pools.patch(a.patch_ids[ai]);
        NccMatch q;
        const double dr = a.scale * (r + P.orow) + a.phase, dcol = a.scale * (c + P.ocol) + a.phase;
        q.row = F.crop_r + dr;
        q.col = F.crop_c + dcol;
        q.angle = P.angle;
        q.score = F.planes[pos * A + size_t(ai)];
        if (k < a.cap) {
          a.cand[size_t(f) * a.cap + k] = q;
          a.key[size_t(f) * a.cap + k] = ai;
          orow = fmin(orow, q.row - F.crop_r); xrow = fmax(xrow, q.row - F.crop_r);
          ocol = fmin(ocol, q.col - F.crop_c); xcol = fmax(xcol, q.col - F.crop_c);
        }
        ++k;
      }
    }
  }
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) {
    orow = fmin(orow, __shfl_xor(orow, s, 64));
    ocol = fmin(ocol, __shfl_xor(ocol, s, 64));
    xrow = fmax(xrow, __shfl_xor(xrow, s, 64));
    xcol = fmax(xcol, __shfl_xor(xcol, s, 64));
  }
  if (lane == 0 && isfinite(orow)) {
    unsigned long long* e = a.extent + 4 * size_t(f);
    atomicMin(e + 0, __double_as_longlong(orow));
    atomicMax(e + 1, __double_as_longlong(xrow));
    atomicMin(e + 2, __double_as_longlong(ocol));
    atomicMax(e + 3, __double_as_longlong(xcol));
  }
}

// Level-0 storage rectangle of every frame from its candidate extent
// (matcher.cpp:612-629), written into the frame descriptors.  Frames without
// candidates keep the full level.
__global__ void k_storage0(NccFrame* __restrict__ frames, int n_frames, const int32_t* __restrict__ count,
                           const unsigned long long* __restrict__ extent, double margin) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n_frames) return;
  NccFrame& F = frames[f];
  const int h = F.levels[0].height, w = F.levels[0].width;
  int32_t s[4] = {0, 0, h, w};
  if (count[f] > 0) {
    const double rmin = __longlong_as_double(extent[4 * f + 0]), rmax = __longlong_as_double(extent[4 * f + 1]);
    const double cmin = __longlong_as_double(extent[4 * f + 2]), cmax = __longlong_as_double(extent[4 * f + 3]);
    s[0] = int32_t(size_t(fmax(0., floor(rmin - margin))));
    s[1] = int32_t(size_t(fmax(0., floor(cmin - margin))));
    s[2] = int32_t(size_t(fmin(double(h), ceil(rmax + margin) + 1)) - size_t(s[0]));
    s[3] = int32_t(size_t(fmin(double(w), ceil(cmax + margin) + 1)) - size_t(s[1]));
  }
  for (int i = 0; i < 4; ++i) F.storage0[i] = s[i];
}

// Reduced search domain on the coarse planes (matcher.cpp:535-579): an anchor
// whose pose may not reach the domain (may_contain at the top-level spacing)
// scores missing, as the reference leaves it.  Runs between the coarse
// correlation and the peaks, for frames with a reduced domain only.
struct DomainPlaneArgs {
  const NccFrame* frames;
  int level, n_angles;
  Pools pools;
  const int32_t* patch_ids;
  double scale, phase;
  int hwa;  // planes in [row][col][angle] order (else [angle][row][col])
};

constexpr int kDomainPlaneRows = 8;

__global__ void __launch_bounds__(256) k_domain_planes(const DomainPlaneArgs a) {
  const int f = blockIdx.z / a.n_angles, ai = blockIdx.z % a.n_angles;
  const NccFrame& F = a.frames[f];
  if (!F.dom_sat) return;
  const int H = F.levels[a.level].height, W = F.levels[a.level].width;
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= W) return;
  const NccPatch& P = a.pools.patch(a.patch_ids[ai]);
  const int4 rg = reinterpret_cast<const int4*>(F.ranges)[ai];
  const int r_end = min(H, int(blockIdx.y + 1) * kDomainPlaneRows);
  for (int r = blockIdx.y * kDomainPlaneRows; r < r_end; ++r) {
    if (r < rg.x || r >= rg.y || c < rg.z || c >= rg.w) continue;  // scored missing already
    const double row = F.crop_r + a.scale * (r + P.orow) + a.phase;
    const double col = F.crop_c + a.scale * (c + P.ocol) + a.phase;
    if (may_contain(F, row, col, a.scale)) continue;
    if (a.hwa == 2) {
      // sparse planes: drop the angle's peak bit
      uint32_t* word = F.peak_bits + (size_t(r) * W + c) * size_t((a.n_angles + 31) / 32) + ai / 32;
      const uint32_t bit = 1u << (ai % 32);
      if (*word & bit) atomicAnd(word, ~bit);
      continue;
    }
    const size_t idx = a.hwa ? (size_t(r) * W + c) * a.n_angles + ai : (size_t(ai) * H + r) * W + c;
    F.planes[idx] = NCC_MISSING;
  }
}

}  // namespace ncc

using namespace ncc;

extern "C" int ncchip_domain_planes(const NccFrame* frames, int n_frames, int level, int max_w, int max_h,
                                    const NccPatchPool* pools, const int32_t* patch_ids, int n_angles, double scale,
                                    double phase, int hwa, void* stream) {
  if (n_frames <= 0 || n_angles <= 0) return 0;
  if (size_t(n_frames) * n_angles > 65535) return -1;
  DomainPlaneArgs a;
  a.frames = frames; a.level = level; a.n_angles = n_angles; a.pools.p0 = pools[0]; a.pools.p1 = pools[1];
  a.patch_ids = patch_ids; a.scale = scale; a.phase = phase; a.hwa = hwa;
  const dim3 grid((max_w + 255) / 256, (max_h + kDomainPlaneRows - 1) / kDomainPlaneRows, unsigned(n_frames * n_angles));
  hipLaunchKernelGGL(k_domain_planes, grid, dim3(256), 0, static_cast<hipStream_t>(stream), a);
  return hipGetLastError() == hipSuccess ? 0 : -1000;
}

extern "C" int ncchip_coarse(const NccFrame* frames, int n_frames, int level, int level_kind, int max_w, int max_h,
                             double C, const NccPatchPool* pools, const int32_t* patch_ids, int n_angles,
                             int max_bbox_h, int max_bbox_w, int form, int exact_lattice, int absolute,
                             void* stream) {
  if (n_frames <= 0 || n_angles <= 0) return 0;
  if (size_t(n_frames) * n_angles > 65535) return -1;
  CoarseArgs a;
  a.frames = frames; a.level = level; a.n_angles = n_angles; a.C = C; a.pools.p0 = pools[0]; a.pools.p1 = pools[1];
  a.patch_ids = patch_ids; a.exact_lattice = exact_lattice; a.absolute = absolute;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int LH = kTileR + max_bbox_h;
  a.lds_w = ((kTileC + max_bbox_w + 16) + 3) & ~3;
  const size_t lds = size_t(LH) * a.lds_w * sizeof(int32_t);
  const unsigned z = unsigned(n_frames * n_angles);
  if (form == NCCHIP_FORM_FP32 && lds <= 64 * 1024) {
    const dim3 grid((max_w + kTileC - 1) / kTileC, (max_h + kTileR - 1) / kTileR, z);
    if (level_kind == NCC_U8) hipLaunchKernelGGL(k_coarse_fp32<uint8_t>, grid, dim3(256), lds, st, a);
    else if (level_kind == NCC_U16) hipLaunchKernelGGL(k_coarse_fp32<uint16_t>, grid, dim3(256), lds, st, a);
    else hipLaunchKernelGGL(k_coarse_fp32<uint32_t>, grid, dim3(256), lds, st, a);
  } else {
    if (max_h > 65535) return -1;
    const dim3 grid((max_w + 255) / 256, max_h, z);
    if (level_kind == NCC_U8) hipLaunchKernelGGL(k_coarse_fp64<uint8_t>, grid, dim3(256), 0, st, a);
    else if (level_kind == NCC_U16) hipLaunchKernelGGL(k_coarse_fp64<uint16_t>, grid, dim3(256), 0, st, a);
    else hipLaunchKernelGGL(k_coarse_fp64<uint32_t>, grid, dim3(256), 0, st, a);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1000;
}

extern "C" int ncchip_peaks(const NccFrame* frames, int n_frames, int level, int max_w, int max_h,
                            const NccPatchPool* pools, const int32_t* patch_ids, int n_angles, double threshold,
                            double scale, double phase, NccMatch* cand, int32_t* key, int cap, int32_t* count,
                            unsigned long long* extent, int hwa, void* stream) {
  if (n_frames <= 0 || n_angles <= 0) return 0;
  if (size_t(n_frames) * n_angles > 65535 || max_h > 65535) return -1;
  PeakArgs a;
  a.frames = frames; a.level = level; a.n_angles = n_angles; a.pools.p0 = pools[0]; a.pools.p1 = pools[1];
  a.patch_ids = patch_ids; a.thr = threshold; a.scale = scale; a.phase = phase; a.cand = cand; a.key = key;
  a.cap = cap; a.count = count; a.extent = extent; a.hwa = hwa;
  if (hwa == 2) {
    const dim3 grid((max_w + 255) / 256, (max_h + kPeakBitRows - 1) / kPeakBitRows, unsigned(n_frames));
    hipLaunchKernelGGL(k_peaks_bits, grid, dim3(256), 0, static_cast<hipStream_t>(stream), a);
    return hipGetLastError() == hipSuccess ? 0 : -1000;
  }
  const dim3 grid(hwa ? (max_w * n_angles + 255) / 256 : (max_w + 255) / 256, (max_h + kPeakRows - 1) / kPeakRows,
                  unsigned(hwa ? n_frames : n_frames * n_angles));
  hipLaunchKernelGGL(k_peaks, grid, dim3(256), 0, static_cast<hipStream_t>(stream), a);
  return hipGetLastError() == hipSuccess ? 0 : -1000;
}

extern "C" int ncchip_storage0(NccFrame* frames, int n_frames, const int32_t* count, const unsigned long long* extent,
                               double margin, void* stream) {
  if (n_frames <= 0) return 0;
  hipLaunchKernelGGL(k_storage0, dim3((n_frames + 63) / 64), dim3(64), 0, static_cast<hipStream_t>(stream), frames,
                     n_frames, count, extent, margin);
  return hipGetLastError() == hipSuccess ? 0 : -1000;
}
