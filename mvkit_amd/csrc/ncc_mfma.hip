// Split-BF16 MFMA coarse correlation (matcher.cpp:542-579 with the window
// statistics of kernels_avx2.cpp:145-171) on gfx950 matrix cores.
//
// The coarse score planes are an implicit GEMM per output row:
//   D[anchor c][angle a] = sum over taps (y, x) of W_a[y][x] * T[r + y][c + x]
// with T = S - C the centred top-level block sums and W_a the dense weights of
// coarse patch a over the union bounding box of all coarse patches (zero
// where the patch has no sample).  v_mfma_f32_32x32x16_bf16 takes 32 anchors
// (M), 32 angle slots (N) and 16 taps (K) as two runs of 8 consecutive
// template columns, one per half-wave, each from any row: the K steps of an
// angle tile walk 8-column segments covering, row by row, the union of the
// tile's supports (not the whole bounding box).  Precision: T = T_hi + T_lo with T_hi = 4^l * (floor(S / 4^l) - 128)
// and T_lo = S mod 4^l, both exact in BF16 for levels <= 4; each weight is
// split into BF16 w_hi + w_lo (w_lo = bf16(w - w_hi)); the products
// T_hi*w_hi, T_hi*w_lo and T_lo*w_hi are exact in FP32 and accumulate in FP32
// (T_lo*w_lo, below 2^-8 of the smallest kept term, is dropped).
//
// FP8 screen (kScreen): before the split-BF16 products, one pass of
// v_mfma_f32_32x32x16_fp8_fp8 over the same tile (image T scaled by 4^-l,
// weights w_hi scaled by a power of two, both converted by the hardware to
// E4M3) gives approximate dot products; a wave whose windows all have
// (|dot8| + C |wsum|) / sd_common below the screen threshold skips its
// split-BF16 products and scoring (its windows are written nowhere in
// sparse mode, -inf in dense mode).  The survivors' scores are the
// split-BF16 form's, bit for bit; the screen only decides which waves run.
//
// The window moments sum(S) and sum(S^2) over each patch's exact support are
// exact integers, slid along each row run in 32-bit (64-bit where a run could
// overflow) arithmetic on the vector ALUs, and the FP64 normalisation is the
// one every correlation form shares, so only the dot product carries the
// split-BF16 error.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "ncc_common.hip.h"
#include "../../include/ncchip.h"

namespace ncc {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float floatx16;

constexpr int kThreads = 256;              // 4 waves, one per column group
constexpr int kWaveRows = 4;               // output rows per wave
constexpr int kMRows = kWaveRows;          // output rows per block
constexpr int kWaveCols = 32;              // anchors per wave (MFMA M)
constexpr int kBlockCols = 4 * kWaveCols;
constexpr int kRunCache = 512;             // support runs of a tile's 32 patches kept in LDS
constexpr int kCommonCache = 64;           // runs of the support common to every patch kept in LDS
// static LDS of coarse_tile beside the image planes: slot descriptors, run
// caches, per-wave deviations and score queues
constexpr size_t kStaticLds = 32 * (4 + 4 + 4 + 8 + 8) + 33 * 4 + kRunCache * 8 + kCommonCache * 16 +
                              (kThreads / 64) * (kWaveRows * kWaveCols * 4 + 64 * 8) + 64;

struct MfmaArgs {
  const NccFrame* frames;
  int level;
  int n_angles;              // coarse angles; slots beyond are zero weights
  int n_tiles;               // angle tiles of 32; blockIdx.z = frame * n_tiles + tile
  double C;                  // 128 * 4^level
  Pools pools;
  const int32_t* patch_ids;  // coarse patches as (pool << 24) | index
  const uint4* wh;           // [K step][64 lanes] of 8 bf16 (w_hi); the steps of tile t from step0[t]
  const uint4* wl;           // same layout (w_lo)
  const int* step0;          // first K step of each angle tile, n_tiles + 1 entries
  const int2* segs;          // per K step: (y << 16 | x) of the segment of half-wave 0 and of half-wave 1
  const double* wsum_eff;    // per angle: sum of (w_hi + w_lo) in FP64
  int ymin, xmin, yspan, xc; // union bounding box (rows, first column) and 16-column chunks (LDS tile width)
  int pitch;                 // LDS row pitch (elements, even)
  int exact_lattice, absolute;
  const int4* common_runs;   // support shared by every coarse patch
  int n_common_runs, n_common;
  double prune_below;        // score bound under which windows are written -inf
  double peak_thr;           // sparse planes: scores below are not written (-inf: dense planes)
  // FP8 screen
  double screen;             // bound threshold under which a wave skips its split-BF16 work
  float img8_scale;          // 4^-l: T * img8_scale is at most 128 in magnitude
  float w8_scale;            // power of two: w_hi * w8_scale is at most 128 in magnitude
  double dot8_unscale;       // 1 / (img8_scale * w8_scale)
  int pitch8;                // LDS row pitch of the FP8 image plane (bytes, multiple of 4)
};

// Two E4M3 bytes of a and b, by the hardware's conversion, into bytes
// (2 kHi, 2 kHi + 1) of old.
template <bool kHi>
__device__ __forceinline__ int fp8_pack2(float a, float b, int old) {
  return __builtin_amdgcn_cvt_pk_fp8_f32(a, b, old, kHi);
}

// Calls fn(integral_constant<int, i>) for i = 0 .. N-1 (compile-time indices).
template <int N, class Fn>
__device__ __forceinline__ void static_for(Fn&& fn) {
  if constexpr (N > 0) {
    static_for<N - 1>(fn);
    fn(std::integral_constant<int, N - 1>{});
  }
}

// Exact BF16 bits of an integer |v| <= 2^15 with at most 8 significant bits.
__device__ __forceinline__ uint16_t bf16_bits_exact(int v) {
  return uint16_t(__float_as_uint(float(v)) >> 16);
}

// One block's work: output rows 4 by .. 4 by + 3, anchors 128 bx .. + 127 of
// frame f, angle tile `tile` (32 coarse angles).  Returns uniformly when the
// block lies past the frame; after the screen decision a wave may return on
// its own (no block barrier follows it).
template <class Tpix, class RunAcc, bool kScreen>
__device__ __forceinline__ void coarse_tile(const MfmaArgs& args, const int f, const int tile, const int bx,
                                            const int by) {
  extern __shared__ uint16_t lds[];
  const NccFrame& F = args.frames[f];
  const NccLevel lv = F.levels[args.level];
  const int H = lv.height, W = lv.width;
  const int r0 = by * kMRows, c0 = bx * kBlockCols;
  if (r0 >= H || c0 >= W) return;  // uniform per block
  const int P = args.pitch;
  const int rows = kMRows + args.yspan - 1;
  const int cols = kBlockCols + 16 * args.xc + 9;  // a segment may start at the last column of the box
  uint16_t* th0 = lds;                // T_hi bits
  uint16_t* th1 = th0 + rows * P;     // T_hi shifted left by one column
  uint16_t* tl0 = th1 + rows * P;     // T_lo bits
  uint16_t* tl1 = tl0 + rows * P;
  uint16_t* su = tl1 + rows * P;      // S
  uint8_t* x8 = reinterpret_cast<uint8_t*>(su + rows * P);  // E4M3 of T * img8_scale (screen only)
  // the angle tile's split weights (MFMA B fragments), read from L2 one
  // iteration ahead of their use
  const int step0 = args.step0[tile];
  const int n_steps = args.step0[tile + 1] - step0;
  const uint4* gwh = args.wh + size_t(step0) * 64;
  const uint4* gwl = args.wl + size_t(step0) * 64;
  const int2* segs = args.segs + step0;
  const int Ci = int(args.C);
  const int shift = lv.shift;
  const int lo_mask = (1 << shift) - 1;
  // The kept windows' exact moments need each slot's patch descriptor and
  // support runs, the common-support deviations the common runs: they are
  // loaded with the tile (threads 0-31: the slots' descriptors, 64..: the
  // common runs; then every thread at most two runs) and kept in LDS.
  __shared__ int s_slot_n[32], s_slot_roff[32], s_run0[33];
  __shared__ int32_t s_slot_ref[32];
  __shared__ double s_slot_wsum[32], s_slot_weff[32];
  __shared__ int2 s_runs[kRunCache];      // (row, begin | end << 16)
  __shared__ int4 s_common[kCommonCache];
  int slot_n = 0, slot_nruns = 0, slot_roff = 0;
  int32_t slot_ref = 0;
  double slot_wsum = 0, slot_weff = 0;
  if (threadIdx.x < 32) {
    const int aa = tile * 32 + int(threadIdx.x);
    if (aa < args.n_angles) {
      slot_ref = args.patch_ids[aa];
      const NccPatch& Pq = args.pools.patch(slot_ref);
      slot_n = Pq.n; slot_nruns = Pq.n_runs; slot_roff = Pq.run_off; slot_wsum = Pq.wsum;
      slot_weff = args.wsum_eff[aa];
    }
  }
  const bool common_cached = args.n_common_runs <= kCommonCache;
  const int cj = int(threadIdx.x) - 64;
  int4 crun = make_int4(0, 0, 0, 0);
  if (common_cached && cj >= 0 && cj < args.n_common_runs) crun = args.common_runs[cj];
  // ---- stage the tile: image rows r0 + ymin .., columns c0 + xmin ..
  // (U loads in flight per thread before the first use).  Element i is
  // (i / cols, i % cols); the quotient comes from a float reciprocal, exact
  // here (i < 2^16, cols < 2^10: (i + 0.5) / cols stays more than 2^-11 away
  // from an integer).
  {
    constexpr int U = 8;
    const int tr0 = r0 + args.ymin, tc0 = c0 + args.xmin;
    const Tpix* img = static_cast<const Tpix*>(lv.data);
    const int total = rows * cols;
    const float inv_cols = 1.f / float(cols);
    for (int base = threadIdx.x; base < total; base += kThreads * U) {
      int Sv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int i = base + kThreads * u;
        const int tr = int((float(i) + .5f) * inv_cols), tc = i - tr * cols;
        const int gr = tr0 + tr, gc = tc0 + tc;
        Sv[u] = (i < total && gr >= 0 && gr < H && gc >= 0 && gc < W) ? int(img[size_t(gr) * lv.pitch + gc]) : Ci;
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int i = base + kThreads * u;
        if (i >= total) break;
        const int tr = int((float(i) + .5f) * inv_cols), tc = i - tr * cols;
        const int S = Sv[u];
        const int lo = S & lo_mask;
        const uint16_t hb = bf16_bits_exact(S - Ci - lo), lb = bf16_bits_exact(lo);
        th0[tr * P + tc] = hb;
        tl0[tr * P + tc] = lb;
        if (tc > 0) { th1[tr * P + tc - 1] = hb; tl1[tr * P + tc - 1] = lb; }
        su[tr * P + tc] = uint16_t(S);
        if constexpr (kScreen) x8[tr * args.pitch8 + tc] = uint8_t(fp8_pack2<false>(float(S - Ci) * args.img8_scale, 0.f, 0));
      }
    }
  }
  if (threadIdx.x < 32) {  // first run of each slot: exclusive scan over lanes 0-31 of wave 0
    int incl = slot_nruns;
#pragma unroll
    for (int d = 1; d < 32; d <<= 1) {
      const int o = __shfl_up(incl, d, 64);
      if (int(threadIdx.x) >= d) incl += o;
    }
    s_run0[threadIdx.x] = incl - slot_nruns;
    if (threadIdx.x == 31) s_run0[32] = incl;
    s_slot_n[threadIdx.x] = slot_n;
    s_slot_roff[threadIdx.x] = slot_roff;
    s_slot_ref[threadIdx.x] = slot_ref;
    s_slot_wsum[threadIdx.x] = slot_wsum;
    s_slot_weff[threadIdx.x] = slot_weff;
  }
  if (common_cached && cj >= 0 && cj < args.n_common_runs) s_common[cj] = crun;
  __syncthreads();
  // the slots' runs: loads issued here, stored before the epilogue's barrier
  const int n_runs_all = s_run0[32];
  const bool runs_cached = n_runs_all <= kRunCache;
  int4 myrun[2] = {make_int4(0, 0, 0, 0), make_int4(0, 0, 0, 0)};
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int j = int(threadIdx.x) + kThreads * u;
    if (runs_cached && j < n_runs_all) {
      int lo = 0, hi = 31;  // last slot whose first run is <= j
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (s_run0[mid] <= j) lo = mid;
        else hi = mid - 1;
      }
      myrun[u] = reinterpret_cast<const int4*>(args.pools.of(s_slot_ref[lo]).runs)[s_slot_roff[lo] + j - s_run0[lo]];
    }
  }
  auto store_runs = [&]() __attribute__((always_inline)) {
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int j = int(threadIdx.x) + kThreads * u;
      if (runs_cached && j < n_runs_all)
        s_runs[j] = make_int2(myrun[u].x, (myrun[u].y & 0xffff) | (myrun[u].z << 16));
    }
  };
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wc = wave & 3, wr0 = (wave >> 2) * kWaveRows;  // column group, first block row of this wave
  const int m = lane & 31, g = lane >> 5;
  // ---- epilogue, part 1: for each of the wave's 8 x 32 windows, the inverse
  // standard deviation of the support shared by every patch.  Its variance is
  // at most any patch's window variance (a subset's squared deviations about
  // its own mean), so by Cauchy-Schwarz (unit-energy weights)
  //   |score| <= (|dot| + 128 * 4^l * |wsum|) * inv_sd_common.
  // Lane l: wave row l >> 4, anchors 2 (l & 15) .. + 1, slid along the shared runs.
  __shared__ float s_inv_sd_f[kThreads / 64][kWaveRows][kWaveCols];
  const bool prune = args.prune_below > 0 && args.n_common > 0;
  auto common_sd = [&]() __attribute__((always_inline)) {
    if (prune) {
      const int rr = lane >> 4, ca = wc * kWaveCols + 2 * (lane & 15);
      int64_t m1[2] = {0, 0}, m2[2] = {0, 0};
      for (int k = 0; k < args.n_common_runs; ++k) {
        const int4 run = common_cached ? s_common[k] : args.common_runs[k];
        const uint16_t* src = su + (wr0 + rr + run.x - args.ymin) * P + (ca - args.xmin);
        int32_t s1 = 0;
        RunAcc s2 = 0;
        int x = run.y;
        for (; x + 7 <= run.z; x += 8) {  // eight reads in flight
          int32_t t[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) t[u] = int32_t(src[x + u]) - Ci;
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            s1 += t[u];
            s2 += RunAcc(uint32_t(t[u] * t[u]));
          }
        }
        for (; x <= run.z; ++x) {
          const int32_t t = int32_t(src[x]) - Ci;
          s1 += t;
          s2 += RunAcc(uint32_t(t * t));
        }
        m1[0] += s1;
        m2[0] += int64_t(s2);
        {
          const int32_t out = int32_t(src[run.y]) - Ci, in = int32_t(src[run.z + 1]) - Ci;
          s1 += in - out;
          s2 += RunAcc(uint32_t(in * in)) - RunAcc(uint32_t(out * out));
          m1[1] += s1;
          m2[1] += int64_t(s2);
        }
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        // n * sum(t^2) - sum(t)^2 = n^2 * variance, exact in 128 bits
        const __int128 vn = (__int128)args.n_common * m2[j] - (__int128)m1[j] * m1[j];
        s_inv_sd_f[wave][rr][2 * (lane & 15) + j] = vn > 0 ? float(sqrt(double(args.n_common) / double(vn))) : -1.f;
      }
    }
  };
  // ---- MFMA: this wave's 32 anchors x 32 angle slots x kMRows rows
  const int col_base = wc * kWaveCols + m;  // tile column of this lane's first A element for segment column 0
  bool run_bf16 = true;  // wave-uniform
  if constexpr (kScreen) {
    // the screen needs the common-support deviations first
    store_runs();
    common_sd();
    __syncthreads();
    floatx16 acc8[kWaveRows];
#pragma unroll
    for (int r = 0; r < kWaveRows; ++r)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc8[r][v] = 0.f;
    for (int it = 0; it < n_steps; ++it) {
      const int2 sg = segs[it];
      const int sv = g ? sg.y : sg.x;
      const int yi = sv >> 16, xi = sv & 0xffff;
      const uint4 bq = gwh[it * 64 + lane];
      int b0 = fp8_pack2<false>(__uint_as_float(bq.x << 16) * args.w8_scale, __uint_as_float(bq.x & 0xffff0000u) * args.w8_scale, 0);
      b0 = fp8_pack2<true>(__uint_as_float(bq.y << 16) * args.w8_scale, __uint_as_float(bq.y & 0xffff0000u) * args.w8_scale, b0);
      int b1 = fp8_pack2<false>(__uint_as_float(bq.z << 16) * args.w8_scale, __uint_as_float(bq.z & 0xffff0000u) * args.w8_scale, 0);
      b1 = fp8_pack2<true>(__uint_as_float(bq.w << 16) * args.w8_scale, __uint_as_float(bq.w & 0xffff0000u) * args.w8_scale, b1);
      const long B8 = long(uint32_t(b0)) | (long(uint32_t(b1)) << 32);
      const int cs = col_base + xi;
      const int sh = cs & 3;
#pragma unroll
      for (int r = 0; r < kWaveRows; ++r) {
        const uint32_t* w = reinterpret_cast<const uint32_t*>(x8 + (wr0 + r + yi) * args.pitch8 + (cs & ~3));
        const uint32_t d0 = w[0], d1 = w[1], d2 = w[2];
        const long A8 = long(__builtin_amdgcn_alignbyte(d1, d0, sh)) | (long(__builtin_amdgcn_alignbyte(d2, d1, sh)) << 32);
        acc8[r] = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(A8, B8, acc8[r], 0, 0, 0);
      }
    }
    // does any of the wave's windows pass the screen?
    const int a8 = tile * 32 + m;
    const bool act8 = a8 < args.n_angles;
    const int4 rg8 = reinterpret_cast<const int4*>(F.ranges)[act8 ? a8 : 0];
    const double slack8 = double(Ci) * fabs(args.wsum_eff[act8 ? a8 : 0]);
    const int cw8 = 16 * g, cb8 = c0 + wc * kWaveCols + cw8;
    bool pass = false;
#pragma unroll
    for (int r = 0; r < kWaveRows; ++r) {
      const int row = r0 + wr0 + r;
      const bool row_in = act8 && row < H && row >= rg8.x && row < rg8.y;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(acc8[r][i]), __float_as_uint(acc8[r][8 + i]),
                                                         false, false);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int j = 8 * (i / 4) + 4 * h + (i % 4);
          const int c = cb8 + j;
          const float isd = s_inv_sd_f[wave][r][cw8 + j];
          const double bnd = (fabs(double(__uint_as_float(sw[h])) * args.dot8_unscale) + slack8) * double(isd);
          pass |= row_in && c < W && c >= rg8.z && c < rg8.w && (!prune || isd < 0.f || bnd >= args.screen);
        }
      }
    }
    run_bf16 = __ballot(pass) != 0;
  }
  floatx16 acc[kWaveRows];
#pragma unroll
  for (int r = 0; r < kWaveRows; ++r)
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[r][v] = 0.f;
  const int n_iter = n_steps;
  uint4 bh_next = make_uint4(0, 0, 0, 0), bl_next = bh_next;
  if (n_iter > 0) {
    bh_next = gwh[lane];
    bl_next = gwl[lane];
  }
  for (int it = 0; run_bf16 && it < n_iter; ++it) {
    const int2 sg = segs[it];
    const int sv = g ? sg.y : sg.x;
    const int yi = sv >> 16, xi = sv & 0xffff;
    {
      const uint4 bh4 = bh_next, bl4 = bl_next;
      if (it + 1 < n_iter) {
        bh_next = gwh[(it + 1) * 64 + lane];
        bl_next = gwl[(it + 1) * 64 + lane];
      }
      const bf16x8 bh = __builtin_bit_cast(bf16x8, bh4), bl = __builtin_bit_cast(bf16x8, bl4);
      const int cs = col_base + xi;
      const bool odd = cs & 1;
      const int ce = cs - (odd ? 1 : 0);
      const uint16_t* ph = odd ? th1 : th0;
      const uint16_t* pl = odd ? tl1 : tl0;
      bf16x8 ah[kWaveRows], al[kWaveRows];
#pragma unroll
      for (int r = 0; r < kWaveRows; ++r) {
        // 4-byte aligned dword reads (ce is even)
        const uint32_t* h = reinterpret_cast<const uint32_t*>(ph + (wr0 + r + yi) * P + ce);
        const uint32_t* l = reinterpret_cast<const uint32_t*>(pl + (wr0 + r + yi) * P + ce);
        ah[r] = __builtin_bit_cast(bf16x8, make_uint4(h[0], h[1], h[2], h[3]));
        al[r] = __builtin_bit_cast(bf16x8, make_uint4(l[0], l[1], l[2], l[3]));
      }
#pragma unroll
      for (int r = 0; r < kWaveRows; ++r) acc[r] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[r], bh, acc[r], 0, 0, 0);
#pragma unroll
      for (int r = 0; r < kWaveRows; ++r) acc[r] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[r], bl, acc[r], 0, 0, 0);
#pragma unroll
      for (int r = 0; r < kWaveRows; ++r) acc[r] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[r], bh, acc[r], 0, 0, 0);
    }
  }
  if constexpr (!kScreen) {
    store_runs();
    common_sd();
    __syncthreads();
  }
  // part 2.  Lane (slot m = lane & 31, half g = lane >> 5) holds the anchors
  // 8 * (v / 4) + 4 * g + (v % 4) of its wave's 32; one exchange with the
  // partner half gives it the 16 consecutive anchors 16 g .. 16 g + 15.
  // Windows outside the valid range or under the bound are written -inf; the
  // others are queued per wave (at most 64) and scored lane-parallel: exact
  // moments over the patch's runs and the FP64 normalisation.
  __shared__ int s_code[kThreads / 64][64];
  __shared__ float s_dot[kThreads / 64][64];
  int queued = 0;  // wave-uniform
  const bool sparse = args.peak_thr > -__builtin_inf();
  const double dc = double(frame_dc(F));
  const int64_t D = int64_t(frame_dc(F)) << shift;
  const int64_t Cl = Ci;
  auto flush = [&]() __attribute__((always_inline)) {
    __builtin_amdgcn_wave_barrier();
    if (lane < queued) {
      const int code = s_code[wave][lane];
      const float dv = s_dot[wave][lane];
      const int r = code >> 16, col = (code >> 8) & 0xff, slot = code & 0xff;
      const int aa = tile * 32 + slot;
      const int run0 = s_run0[slot], n_runs = s_run0[slot + 1] - run0;
      const int4* rq = reinterpret_cast<const int4*>(args.pools.of(s_slot_ref[slot]).runs) + s_slot_roff[slot];
      int64_t m1 = 0, m2 = 0;
      for (int k = 0; k < n_runs; ++k) {
        int4 run;
        if (runs_cached) {
          const int2 c = s_runs[run0 + k];
          run = make_int4(c.x, int(int16_t(c.y & 0xffff)), c.y >> 16, 0);
        } else {
          run = rq[k];
        }
        const uint16_t* src = su + (r + run.x - args.ymin) * P + (wc * kWaveCols + col - args.xmin);
        int32_t s1 = 0;
        RunAcc s2 = 0;
        int x = run.y;
        for (; x + 7 <= run.z; x += 8) {  // eight reads in flight
          int32_t t[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) t[u] = int32_t(src[x + u]) - Ci;
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            s1 += t[u];
            s2 += RunAcc(uint32_t(t[u] * t[u]));
          }
        }
        for (; x <= run.z; ++x) {
          const int32_t t = int32_t(src[x]) - Ci;
          s1 += t;
          s2 += RunAcc(uint32_t(t * t));
        }
        m1 += s1;
        m2 += int64_t(s2);
      }
      const int64_t nq = s_slot_n[slot];
      const int64_t sum_s = m1 + nq * Cl;
      const int64_t sum_s2 = m2 + 2 * Cl * m1 + nq * Cl * Cl;
      const double score = normalise(double(dv), args.C, s_slot_weff[slot], s_slot_wsum[slot], int(nq), sum_s, sum_s2, D, shift,
                                     dc, args.exact_lattice, args.absolute);
      const size_t pos = size_t(r0 + r) * W + c0 + wc * kWaveCols + col;
      if (!sparse) {
        F.planes[pos * args.n_angles + aa] = score;
      } else if (score >= args.peak_thr) {
        F.planes[pos * args.n_angles + aa] = score;
        atomicOr(F.peak_bits + pos * args.n_tiles + tile, 1u << slot);
      }
    }
    __builtin_amdgcn_wave_barrier();
    queued = 0;
  };
  if (!run_bf16) {  // screened out: nothing to score (no block barrier follows)
    const int aa = tile * 32 + m;
    if (!sparse && aa < args.n_angles)
      for (int r = 0; r < kWaveRows; ++r) {
        const int row = r0 + wr0 + r;
        for (int j = 0; j < 16; ++j) {
          const int c = c0 + wc * kWaveCols + 16 * g + j;
          if (row < H && c < W) F.planes[(size_t(row) * W + c) * args.n_angles + aa] = NCC_MISSING;
        }
      }
    return;
  }
  const int a = tile * 32 + m;
  const bool active = a < args.n_angles;
  const int4 rg = reinterpret_cast<const int4*>(F.ranges)[active ? a : 0];
  const double wsum_eff = args.wsum_eff[active ? a : 0];
  const double num_slack = double(Ci) * fabs(wsum_eff);  // |(mean(S) - C) * wsum| <= C * |wsum|
  // planes in [row][col][angle] order: the 32 slots of one anchor are adjacent
  double* plane = F.planes + (active ? a : 0);
  const int cw = 16 * g;  // first of this lane's 16 anchors, relative to the wave's 32
  const uint64_t below = (uint64_t(1) << lane) - 1;
  // The bound test runs in FP32 (it only selects the windows scored exactly;
  // the bound sits a tenth under the peak threshold, far beyond FP32's
  // rounding), one bit per window (wave row r, anchor j: bit 16 r + j); the
  // kept windows are then queued one per lane per pass, so the scoring code
  // appears once.
  const float prune_f = float(args.prune_below), slack_f = float(num_slack);
  float dot[kWaveRows][16];
  uint64_t keep = 0;
  // this lane's 16 anchors c_j = cb + j: bits of those inside the frame
  // (in_width) and inside the angle's valid column range (in_range)
  const int cb = c0 + wc * kWaveCols + cw;
  auto bits_below = [](int k) { return k <= 0 ? 0u : k >= 16 ? 0xffffu : (1u << k) - 1u; };
  const uint32_t in_width = bits_below(W - cb);
  const uint32_t in_range = bits_below(min(W, rg.w) - cb) & ~bits_below(rg.z - cb);
  static_for<kWaveRows>([&](auto rc) __attribute__((always_inline)) {
    constexpr int r = decltype(rc)::value;  // wave row; block row wr0 + r
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(acc[r][i]), __float_as_uint(acc[r][8 + i]),
                                                       false, false);
      // first result: anchors 16 g + 8 (i / 4) + (i % 4); second: the same + 4
      dot[r][8 * (i / 4) + (i % 4)] = __uint_as_float(sw[0]);
      dot[r][8 * (i / 4) + 4 + (i % 4)] = __uint_as_float(sw[1]);
    }
    const int row = r0 + wr0 + r;
    const bool row_in = active && row < H && row >= rg.x && row < rg.y;
    uint32_t pass = 0xffffu;
    if (prune) {
      pass = 0;
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const float isd = s_inv_sd_f[wave][r][cw + j];
        pass |= uint32_t((isd < 0.f) | ((fabsf(dot[r][j]) + slack_f) * isd >= prune_f)) << j;
      }
    }
    const uint32_t kr = (row_in ? in_range : 0u) & pass;
    keep |= uint64_t(kr) << (16 * r);
    if (!sparse) {
      uint32_t miss = (active && row < H ? in_width : 0u) & ~kr;
      while (miss) {
        const int j = __builtin_ctz(miss);
        plane[(size_t(row) * W + c0 + wc * kWaveCols + cw + j) * args.n_angles] = NCC_MISSING;
        miss &= miss - 1;
      }
    }
  });
  while (__ballot(keep != 0)) {
    const bool has = keep != 0;
    const int b = has ? __builtin_ctzll(keep) : 0;
    const uint64_t mask = __ballot(has);
    const int cnt = __popcll(mask);
    if (queued + cnt > 64) flush();
    if (has) {
      float dv = dot[0][0];
#pragma unroll
      for (int q = 1; q < 16 * kWaveRows; ++q) dv = b == q ? dot[q / 16][q % 16] : dv;
      const int pos = queued + __popcll(mask & below);
      s_code[wave][pos] = ((wr0 + (b >> 4)) << 16) | ((cw + (b & 15)) << 8) | m;
      s_dot[wave][pos] = dv;
      keep &= keep - 1;
    }
    queued += cnt;
  }
  if (queued) flush();
}

template <class Tpix, class RunAcc, bool kScreen>
__global__ void __launch_bounds__(kThreads, 3) k_coarse_mfma(const MfmaArgs args) {
  coarse_tile<Tpix, RunAcc, kScreen>(args, int(blockIdx.z) / args.n_tiles, int(blockIdx.z) % args.n_tiles,
                                     int(blockIdx.x), int(blockIdx.y));
}

// The coarse work queue: every (job, frame) unit contributes gx x gy blocks
// per angle tile; its items are numbered from first[unit], angle tile
// outermost, then block rows, then block columns (neighbouring items share
// image rows in L2).
struct QueueUnit { int job, frame, gx, gy; };
struct QueueArgs {
  const MfmaArgs* args;      // per job (device)
  const QueueUnit* units;    // this launch's units (device)
  const int* first;          // first item of each unit, n_units + 1 entries (device)
  int n_units, n_items;
  unsigned* counter;         // next item to take; zero when the launch starts
};

// Persistent coarse search: each block takes items from the queue until it
// is drained (every block exits once the counter passes n_items).
template <class Tpix, class RunAcc, bool kScreen>
__global__ void __launch_bounds__(kThreads, 3) k_coarse_queue(const QueueArgs q) {
  __shared__ int s_item;
  for (;;) {
    __syncthreads();  // the previous item is done with LDS and s_item
    if (threadIdx.x == 0) s_item = int(atomicAdd(q.counter, 1u));
    __syncthreads();
    const int item = __builtin_amdgcn_readfirstlane(s_item);
    if (item >= q.n_items) return;
    int lo = 0, hi = q.n_units - 1;  // last unit whose first item is <= item
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (q.first[mid] <= item) lo = mid;
      else hi = mid - 1;
    }
    const QueueUnit u = q.units[lo];
    const int local = item - q.first[lo];
    const int per_tile = u.gx * u.gy;
    const int tile = local / per_tile, rem = local - tile * per_tile;
    coarse_tile<Tpix, RunAcc, kScreen>(q.args[u.job], u.frame, tile, rem % u.gx, rem / u.gx);
  }
}

}  // namespace ncc

using namespace ncc;

namespace {
// LDS bytes of one k_coarse_mfma block, beside its static tables; 0 when the
// level cannot run in the split-BF16 form.
size_t coarse_mfma_lds(int level, int level_kind, int yspan, int xchunks, bool screened, int* pitch, int* pitch8) {
  if (level_kind == NCC_U32 || level > 4) return 0;  // T_lo exact in BF16 only up to level 4
  const int cols = kBlockCols + 16 * xchunks + 9;
  const int rows = kMRows + yspan - 1;
  *pitch = (cols + 1) & ~1;
  *pitch8 = (cols + 11) & ~3;  // an 8-byte A fragment read as three dwords from any column
  size_t lds = size_t(5) * rows * *pitch * sizeof(uint16_t);
  if (screened) lds += size_t(rows) * *pitch8;
  constexpr size_t kLdsMax = 160 * 1024 - kStaticLds;
  return lds > kLdsMax ? 0 : lds;
}
}  // namespace

extern "C" int ncchip_coarse_mfma_fits(int level, int level_kind, int yspan, int xchunks, int screened) {
  int pitch = 0, pitch8 = 0;
  return coarse_mfma_lds(level, level_kind, yspan, xchunks, screened != 0, &pitch, &pitch8) > 0;
}

namespace {
// Kernel arguments of one split-BF16 coarse search; 0 on success, -6 when the
// level or the tile does not fit the kernel, -1 on bad arguments.
int make_args(const NccFrame* frames, int n_frames, int level, int level_kind, double C, const NccPatchPool* pools,
              const int32_t* patch_ids, int n_angles, const NcchipDenseWeights* dw, int exact_lattice, int absolute,
              double screen, MfmaArgs& a, size_t& lds, bool& narrow) {
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
  if (size_t(n_frames) * a.n_tiles > 65535) return -1;
  lds = coarse_mfma_lds(level, level_kind, a.yspan, a.xc, screened, &a.pitch, &a.pitch8);
  if (!lds) return -6;
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

template <class TP, class RA, bool kScreen>
void launch_grid(const dim3& grid, size_t lds, hipStream_t st, const MfmaArgs& a) {
  hipLaunchKernelGGL((k_coarse_mfma<TP, RA, kScreen>), grid, dim3(kThreads), lds, st, a);
}
void launch_grid_variant(int v, const dim3& grid, size_t lds, hipStream_t st, const MfmaArgs& a) {
  switch (v) {
    case 0: launch_grid<uint8_t, uint32_t, false>(grid, lds, st, a); break;
    case 1: launch_grid<uint8_t, uint32_t, true>(grid, lds, st, a); break;
    case 2: launch_grid<uint8_t, uint64_t, false>(grid, lds, st, a); break;
    case 3: launch_grid<uint8_t, uint64_t, true>(grid, lds, st, a); break;
    case 4: launch_grid<uint16_t, uint32_t, false>(grid, lds, st, a); break;
    case 5: launch_grid<uint16_t, uint32_t, true>(grid, lds, st, a); break;
    case 6: launch_grid<uint16_t, uint64_t, false>(grid, lds, st, a); break;
    default: launch_grid<uint16_t, uint64_t, true>(grid, lds, st, a); break;
  }
}

template <class TP, class RA, bool kScreen>
int launch_queue(size_t lds, hipStream_t st, const QueueArgs& q, int n_cus) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(&k_coarse_queue<TP, RA, kScreen>),
                                                   kThreads, lds) != hipSuccess || per_cu <= 0)
    per_cu = 1;
  const int blocks = std::max(1, std::min(q.n_items, n_cus * per_cu));
  hipLaunchKernelGGL((k_coarse_queue<TP, RA, kScreen>), dim3(blocks), dim3(kThreads), lds, st, q);
  return 0;
}
int launch_queue_variant(int v, size_t lds, hipStream_t st, const QueueArgs& q, int n_cus) {
  switch (v) {
    case 0: return launch_queue<uint8_t, uint32_t, false>(lds, st, q, n_cus);
    case 1: return launch_queue<uint8_t, uint32_t, true>(lds, st, q, n_cus);
    case 2: return launch_queue<uint8_t, uint64_t, false>(lds, st, q, n_cus);
    case 3: return launch_queue<uint8_t, uint64_t, true>(lds, st, q, n_cus);
    case 4: return launch_queue<uint16_t, uint32_t, false>(lds, st, q, n_cus);
    case 5: return launch_queue<uint16_t, uint32_t, true>(lds, st, q, n_cus);
    case 6: return launch_queue<uint16_t, uint64_t, false>(lds, st, q, n_cus);
    default: return launch_queue<uint16_t, uint64_t, true>(lds, st, q, n_cus);
  }
}

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }
}  // namespace

extern "C" int ncchip_coarse_mfma(const NccFrame* frames, int n_frames, int level, int level_kind, int max_w,
                                  int max_h, double C, const NccPatchPool* pools, const int32_t* patch_ids,
                                  int n_angles, const NcchipDenseWeights* dw, int exact_lattice, int absolute,
                                  double screen, void* stream) {
  if (n_frames <= 0 || n_angles <= 0) return 0;
  MfmaArgs a;
  size_t lds = 0;
  bool narrow = false;
  const int rc = make_args(frames, n_frames, level, level_kind, C, pools, patch_ids, n_angles, dw, exact_lattice,
                           absolute, screen, a, lds, narrow);
  if (rc) return rc;
  const dim3 grid((max_w + kBlockCols - 1) / kBlockCols, (max_h + kMRows - 1) / kMRows, unsigned(n_frames * a.n_tiles));
  launch_grid_variant(variant(level_kind, narrow, screen > -__builtin_inf()), grid, lds, static_cast<hipStream_t>(stream), a);
  return hipGetLastError() == hipSuccess ? 0 : -1000;
}

extern "C" size_t ncchip_coarse_mfma_queue_bytes(int n_jobs, int n_units) {
  return align256(size_t(n_jobs) * sizeof(MfmaArgs)) + align256(size_t(n_units) * sizeof(QueueUnit)) +
         align256((size_t(n_units) + 8) * sizeof(int)) + 256;
}

extern "C" int ncchip_coarse_mfma_queue(const NcchipCoarseJob* jobs, int n_jobs, void* d_work, size_t work_bytes,
                                        void* h_stage, size_t stage_bytes, void* stream) {
  if (n_jobs <= 0) return 0;
  if (!jobs || !d_work || !h_stage) return -1;
  std::vector<MfmaArgs> args(static_cast<size_t>(n_jobs));
  std::vector<size_t> lds(static_cast<size_t>(n_jobs));
  std::vector<int> var(static_cast<size_t>(n_jobs));
  int n_units = 0;
  for (int j = 0; j < n_jobs; ++j) {
    const NcchipCoarseJob& J = jobs[j];
    if (J.n_frames <= 0 || J.n_angles <= 0) continue;
    bool narrow = false;
    const int rc = make_args(J.frames, J.n_frames, J.level, J.level_kind, J.C, J.pools, J.patch_ids, J.n_angles,
                             &J.weights, J.exact_lattice, J.absolute, J.screen, args[size_t(j)], lds[size_t(j)], narrow);
    if (rc) return rc;
    var[size_t(j)] = variant(J.level_kind, narrow, J.screen > -__builtin_inf());
    n_units += J.n_frames;
  }
  const size_t need = ncchip_coarse_mfma_queue_bytes(n_jobs, n_units);
  if (work_bytes < need || stage_bytes < need) return -103;
  auto* base = static_cast<uint8_t*>(h_stage);
  const size_t off_units = align256(size_t(n_jobs) * sizeof(MfmaArgs));
  const size_t off_first = off_units + align256(size_t(n_units) * sizeof(QueueUnit));
  const size_t off_counters = off_first + align256((size_t(n_units) + 8) * sizeof(int));
  std::memcpy(base, args.data(), size_t(n_jobs) * sizeof(MfmaArgs));
  auto* units = reinterpret_cast<QueueUnit*>(base + off_units);
  auto* first = reinterpret_cast<int*>(base + off_first);
  std::memset(base + off_counters, 0, 256);
  // units grouped by instantiation; each group's first[] ends with its item count
  struct Group { int variant, unit0, n_units, first0; size_t lds; int64_t items; };
  std::vector<Group> groups;
  int u = 0, fpos = 0;
  for (int v = 0; v < 8; ++v) {
    Group g{v, u, 0, fpos, 0, 0};
    for (int j = 0; j < n_jobs; ++j) {
      const NcchipCoarseJob& J = jobs[j];
      if (J.n_frames <= 0 || J.n_angles <= 0 || var[size_t(j)] != v) continue;
      g.lds = std::max(g.lds, lds[size_t(j)]);
      for (int f = 0; f < J.n_frames; ++f) {
        const int w = J.frame_dims ? J.frame_dims[2 * f] : J.max_w, h = J.frame_dims ? J.frame_dims[2 * f + 1] : J.max_h;
        const int gx = (w + kBlockCols - 1) / kBlockCols, gy = (h + kMRows - 1) / kMRows;
        units[u] = QueueUnit{j, f, std::max(gx, 1), std::max(gy, 1)};
        first[fpos++] = int(g.items);
        g.items += int64_t(std::max(gx, 1)) * std::max(gy, 1) * args[size_t(j)].n_tiles;
        ++u;
        ++g.n_units;
      }
    }
    if (!g.n_units) continue;
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
    launch_queue_variant(g.variant, g.lds, st, q, n_cus);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1000;
}
