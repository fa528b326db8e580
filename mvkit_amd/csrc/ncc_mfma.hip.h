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
// tile's supports (not the whole bounding box).  Precision: T = T_hi + T_lo
// with T_hi = 4^l * (floor(S / 4^l) - 128) and T_lo = S mod 4^l, both exact
// in BF16 for levels <= 4; each weight is split into BF16 w_hi + w_lo
// (w_lo = bf16(w - w_hi)); the products T_hi*w_hi, T_hi*w_lo and T_lo*w_hi
// are exact in FP32 and accumulate in FP32 (T_lo*w_lo, below 2^-8 of the
// smallest kept term, is dropped).
//
// Tile shape (chosen per search by the autotuner, NcchipCoarseTile): a wave
// computes kR output rows x 32 anchors; a block has `cw` waves across and
// `rw` waves down (32 cw anchors x kR rw rows) and runs `tiles_per_block`
// angle tiles of 32 rotations over one staged image tile.  Every output's
// FP32 accumulation runs over the same K steps in the same order whatever
// the shape, so the planes are bit-identical across shapes.
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
#pragma once
#include <type_traits>

#include "ncc_common.hip.h"

namespace ncc {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float floatx16;

constexpr int kMaxThreads = 256;           // at most 4 waves per block
constexpr int kMaxWaves = kMaxThreads / 64;
constexpr int kWaveCols = 32;              // anchors per wave (MFMA M)
constexpr int kRunCache = 512;             // support runs of a tile's 32 patches kept in LDS
constexpr int kCommonCache = 64;           // runs of the support common to every patch kept in LDS
// static LDS of coarse_tile beside the image planes: slot descriptors, run
// caches, per-wave deviations and score queues
template <int kR>
constexpr size_t static_lds() {
  return 32 * (4 + 4 + 4 + 8 + 8) + 33 * 4 + kRunCache * 8 + kCommonCache * 16 +
         kMaxWaves * (kR * kWaveCols * 4 + 64 * 8) + 64;
}
// workgroups per CU the register allocation is sized for
template <int kR>
constexpr int min_blocks() { return kR >= 8 ? 2 : 3; }

struct MfmaArgs {
  const NccFrame* frames;
  int level;
  int n_angles;              // coarse angles; slots beyond are zero weights
  int n_tiles;               // angle tiles of 32
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
  // bound groups (group_slots > 0): consecutive angles a / group_slots share
  // the support common to their own patches, runs group_first[g] ..
  // group_first[g + 1] - 1 of group_runs with group_n[g] samples (g counted
  // from this size's first group); 0: one group over every angle
  int group_slots;
  const int* group_first;
  const int* group_n;
  const int4* group_runs;
  double prune_below;        // score bound under which windows are written -inf
  double peak_thr;           // sparse planes: scores below are not written (-inf: dense planes)
  // FP8 screen
  double screen;             // bound threshold under which a wave skips its split-BF16 work
  float img8_scale;          // 4^-l: T * img8_scale is at most 128 in magnitude
  float w8_scale;            // power of two: w_hi * w8_scale is at most 128 in magnitude
  double dot8_unscale;       // 1 / (img8_scale * w8_scale)
  int pitch8;                // LDS row pitch of the FP8 image plane (bytes, multiple of 4)
  // tile shape (the rows per wave are the kernel's kR)
  int cw, rw;                // waves across and down a block
  int tiles_per_block;       // angle tiles run over one staged image tile
  int n_groups;              // ceil(n_tiles / tiles_per_block); blockIdx.z = frame * n_groups + group
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

// One block's work: output rows kR rw by .. + kR rw - 1, anchors 32 cw bx ..
// + 32 cw - 1 of frame f, angle tiles group * tiles_per_block .. (32 coarse
// angles each) over one staged image tile.  Returns uniformly when the block
// lies past the frame.
template <class Tpix, class RunAcc, bool kScreen, int kR, bool kMulti>
__device__ __forceinline__ void coarse_tile(const MfmaArgs& args, const int f, const int group, const int bx,
                                            const int by) {
  extern __shared__ uint16_t lds[];
  const NccFrame& F = args.frames[f];
  const NccLevel lv = F.levels[args.level];
  const int H = lv.height, W = lv.width;
  const int block_rows = kR * args.rw, block_cols = kWaveCols * args.cw;
  const int r0 = by * block_rows, c0 = bx * block_cols;
  if (r0 >= H || c0 >= W) return;  // uniform per block
  const int tid = int(threadIdx.x), nthreads = 64 * args.cw * args.rw;
  const int P = args.pitch;
  const int rows = block_rows + args.yspan - 1;
  const int cols = block_cols + 16 * args.xc + 9;  // a segment may start at the last column of the box
  uint16_t* th0 = lds;                // T_hi bits
  uint16_t* th1 = th0 + rows * P;     // T_hi shifted left by one column
  uint16_t* tl0 = th1 + rows * P;     // T_lo bits
  uint16_t* tl1 = tl0 + rows * P;
  uint16_t* su = tl1 + rows * P;      // S
  uint8_t* x8 = reinterpret_cast<uint8_t*>(su + rows * P);  // E4M3 of T * img8_scale (screen only)
  const int Ci = int(args.C);
  const int shift = lv.shift;
  const int lo_mask = (1 << shift) - 1;
  const int tile_lo = group * args.tiles_per_block;
  // one angle tile per block unless kMulti (the loop over several tiles keeps
  // more values live: a separate instantiation)
  const int tile_hi = kMulti ? min(args.n_tiles, tile_lo + args.tiles_per_block) : tile_lo + 1;
  // Each angle tile's kept windows need its slots' patch descriptors and
  // support runs, the common-support deviations the common runs: loaded with
  // the tile (threads 0-31: the slots' descriptors; threads 0..: the common
  // runs, once; then every thread at most two runs) and kept in LDS.
  __shared__ int s_slot_n[32], s_slot_roff[32], s_run0[33];
  __shared__ int32_t s_slot_ref[32];
  __shared__ double s_slot_wsum[32], s_slot_weff[32];
  __shared__ int2 s_runs[kRunCache];      // (row, begin | end << 16)
  __shared__ int4 s_common[kCommonCache];
  // slot descriptors of an angle tile (threads 0-31), loaded into registers
  // and stored in LDS after the loads of other work were issued
  int slot_n = 0, slot_nruns = 0, slot_roff = 0;
  int32_t slot_ref = 0;
  double slot_wsum = 0, slot_weff = 0;
  auto load_slot = [&](int tile) __attribute__((always_inline)) {
    slot_n = 0; slot_nruns = 0; slot_roff = 0; slot_ref = 0; slot_wsum = 0; slot_weff = 0;
    const int aa = tile * 32 + tid;
    if (tid < 32 && aa < args.n_angles) {
      slot_ref = args.patch_ids[aa];
      const NccPatch& Pq = args.pools.patch(slot_ref);
      slot_n = Pq.n; slot_nruns = Pq.n_runs; slot_roff = Pq.run_off; slot_wsum = Pq.wsum;
      slot_weff = args.wsum_eff[aa];
    }
  };
  auto store_slot = [&]() __attribute__((always_inline)) {
    if (tid < 32) {  // first run of each slot: exclusive scan over lanes 0-31 of wave 0
      int incl = slot_nruns;
#pragma unroll
      for (int d = 1; d < 32; d <<= 1) {
        const int o = __shfl_up(incl, d, 64);
        if (tid >= d) incl += o;
      }
      s_run0[tid] = incl - slot_nruns;
      if (tid == 31) s_run0[32] = incl;
      s_slot_n[tid] = slot_n;
      s_slot_roff[tid] = slot_roff;
      s_slot_ref[tid] = slot_ref;
      s_slot_wsum[tid] = slot_wsum;
      s_slot_weff[tid] = slot_weff;
    }
  };
  load_slot(tile_lo);
  const bool common_cached = args.n_common_runs <= kCommonCache;
  int4 crun = make_int4(0, 0, 0, 0);
  if (common_cached && tid < args.n_common_runs) crun = args.common_runs[tid];
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
    for (int base = tid; base < total; base += nthreads * U) {
      int Sv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int i = base + nthreads * u;
        const int tr = int((float(i) + .5f) * inv_cols), tc = i - tr * cols;
        const int gr = tr0 + tr, gc = tc0 + tc;
        Sv[u] = (i < total && gr >= 0 && gr < H && gc >= 0 && gc < W) ? int(img[size_t(gr) * lv.pitch + gc]) : Ci;
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int i = base + nthreads * u;
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
  if (common_cached && tid < args.n_common_runs) s_common[tid] = crun;
  const int lane = tid & 63, wave = tid >> 6;
  const int wc = wave % args.cw, wr0 = (wave / args.cw) * kR;  // column group, first block row of this wave
  const int m = lane & 31, g = lane >> 5;
  // ---- epilogue, part 1: for each of the wave's kR x 32 windows, the
  // inverse standard deviation of the support shared by every patch.  Its
  // variance is at most any patch's window variance (a subset's squared
  // deviations about its own mean), so by Cauchy-Schwarz (unit-energy weights)
  //   |score| <= (|dot| + 128 * 4^l * |wsum|) * inv_sd_common.
  // Lane l: wave row (l >> 4) + 4 j, anchors 2 (l & 15) .. + 1, slid along
  // the shared runs.  Computed once per block (it depends on no angle tile).
  // With bound groups the support is the one common to the patches of the
  // group (larger than the support common to every patch, so a tighter
  // bound), and each angle tile computes every group's deviations into an
  // area of dynamic LDS after the image planes: [group][wave][row][anchor].
  // The deviations are private to the wave: written and read by its own
  // lanes, whose LDS operations complete in order.
  __shared__ float s_inv_sd_f[kMaxWaves][kR][kWaveCols];
  const bool grouped = args.group_slots > 0;
  const bool prune = args.prune_below > 0 && (grouped || args.n_common > 0);
  float* const gsd = reinterpret_cast<float*>(su + rows * P);  // no FP8 plane with bound groups
  const int n_waves = args.cw * args.rw;
  auto isd_row = [&](const int sub, const int r) __attribute__((always_inline)) -> float* {
    return grouped ? gsd + ((sub * n_waves + wave) * kR + r) * kWaveCols : &s_inv_sd_f[wave][r][0];
  };
  auto common_sd = [&](const int k0, const int k1, const int n_common, const int4* runs, const int sub)
                       __attribute__((always_inline)) {
    if (prune) {
      const int ca = wc * kWaveCols + 2 * (lane & 15);
#pragma unroll
      for (int rq = 0; rq < kR / 4; ++rq) {
        const int rr = (lane >> 4) + 4 * rq;
        int64_t m1[2] = {0, 0}, m2[2] = {0, 0};
        for (int k = k0; k < k1; ++k) {
          const int4 run = runs ? runs[k] : s_common[k];
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
          const __int128 vn = (__int128)n_common * m2[j] - (__int128)m1[j] * m1[j];
          isd_row(sub, rr)[2 * (lane & 15) + j] = vn > 0 ? float(sqrt(double(n_common) / double(vn))) : -1.f;
        }
      }
    }
  };
  const int col_base = wc * kWaveCols + m;  // tile column of this lane's first A element for segment column 0
  const bool sparse = args.peak_thr > -__builtin_inf();
  const double dc = double(frame_dc(F));
  const int64_t D = int64_t(frame_dc(F)) << shift;
  const int64_t Cl = Ci;
  const int cw16 = 16 * g;  // first of this lane's 16 anchors, relative to the wave's 32
  auto bits_below = [](int k) { return k <= 0 ? 0u : k >= 16 ? 0xffffu : (1u << k) - 1u; };
  __shared__ int s_code[kMaxWaves][64];
  __shared__ float s_dot[kMaxWaves][64];
  for (int tile = tile_lo; tile < tile_hi; ++tile) {
    if (tile > tile_lo) __syncthreads();  // the previous tile is done with the slot tables and run cache
    // the angle tile's split weights (MFMA B fragments), read from L2 one
    // iteration ahead of their use
    const int step0 = args.step0[tile];
    const int n_steps = args.step0[tile + 1] - step0;
    const uint4* gwh = args.wh + size_t(step0) * 64;
    const uint4* gwl = args.wl + size_t(step0) * 64;
    const int2* segs = args.segs + step0;
    if (tile > tile_lo) load_slot(tile);
    store_slot();
    __syncthreads();  // the slot tables (and, for the first tile, the staged image)
    // the slots' runs: loads issued here, stored before the epilogue's barrier
    const int n_runs_all = s_run0[32];
    const bool runs_cached = n_runs_all <= kRunCache && n_runs_all <= 2 * nthreads;
    int4 myrun[2] = {make_int4(0, 0, 0, 0), make_int4(0, 0, 0, 0)};
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int j = tid + nthreads * u;
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
        const int j = tid + nthreads * u;
        if (runs_cached && j < n_runs_all)
          s_runs[j] = make_int2(myrun[u].x, (myrun[u].y & 0xffff) | (myrun[u].z << 16));
      }
    };
    // ---- MFMA: this wave's 32 anchors x 32 angle slots x kR rows
    bool run_bf16 = true;  // wave-uniform
    if constexpr (kScreen) {
      // the screen needs the common-support deviations first
      store_runs();
      if (tile == tile_lo) common_sd(0, args.n_common_runs, args.n_common, common_cached ? nullptr : args.common_runs, 0);
      __syncthreads();
      floatx16 acc8[kR];
#pragma unroll
      for (int r = 0; r < kR; ++r)
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
        for (int r = 0; r < kR; ++r) {
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
      const int cb8 = c0 + wc * kWaveCols + cw16;
      bool pass = false;
#pragma unroll
      for (int r = 0; r < kR; ++r) {
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
            const float isd = s_inv_sd_f[wave][r][cw16 + j];
            const double bnd = (fabs(double(__uint_as_float(sw[h])) * args.dot8_unscale) + slack8) * double(isd);
            pass |= row_in && c < W && c >= rg8.z && c < rg8.w && (!prune || isd < 0.f || bnd >= args.screen);
          }
        }
      }
      run_bf16 = __ballot(pass) != 0;
    }
    floatx16 acc[kR];
#pragma unroll
    for (int r = 0; r < kR; ++r)
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
      bf16x8 ah[kR], al[kR];
#pragma unroll
      for (int r = 0; r < kR; ++r) {
        // 4-byte aligned dword reads (ce is even)
        const uint32_t* h = reinterpret_cast<const uint32_t*>(ph + (wr0 + r + yi) * P + ce);
        const uint32_t* l = reinterpret_cast<const uint32_t*>(pl + (wr0 + r + yi) * P + ce);
        ah[r] = __builtin_bit_cast(bf16x8, make_uint4(h[0], h[1], h[2], h[3]));
        al[r] = __builtin_bit_cast(bf16x8, make_uint4(l[0], l[1], l[2], l[3]));
      }
#pragma unroll
      for (int r = 0; r < kR; ++r) acc[r] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[r], bh, acc[r], 0, 0, 0);
#pragma unroll
      for (int r = 0; r < kR; ++r) acc[r] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[r], bl, acc[r], 0, 0, 0);
#pragma unroll
      for (int r = 0; r < kR; ++r) acc[r] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[r], bh, acc[r], 0, 0, 0);
    }
    if constexpr (!kScreen) {
      store_runs();
      if (!grouped && tile == tile_lo)
        common_sd(0, args.n_common_runs, args.n_common, common_cached ? nullptr : args.common_runs, 0);
      __syncthreads();
    }
    if (grouped) {  // every bound group's deviations, before the bound pass reads them
      const int n_sub = kWaveCols / args.group_slots;
      for (int gi = 0; gi < n_sub && tile * kWaveCols + gi * args.group_slots < args.n_angles; ++gi) {
        const int grp = (tile * kWaveCols) / args.group_slots + gi;
        common_sd(args.group_first[grp], args.group_first[grp + 1], args.group_n[grp], args.group_runs, gi);
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    // part 2.  Lane (slot m = lane & 31, half g = lane >> 5) holds the anchors
    // 8 * (v / 4) + 4 * g + (v % 4) of its wave's 32; one exchange with the
    // partner half gives it the 16 consecutive anchors 16 g .. 16 g + 15.
    // Windows outside the valid range or under the bound are written -inf; the
    // others are queued per wave (at most 64) and scored lane-parallel: exact
    // moments over the patch's runs and the FP64 normalisation.  A flush of q
    // windows gives each the largest power of two of lanes with q * lanes <=
    // 64; a window's lanes take its runs in turn and their integer sums are
    // added across them, so sparse survivors do not leave one lane walking
    // every run of its patch while the others wait.
    int queued = 0;  // wave-uniform
    auto flush = [&]() __attribute__((always_inline)) {
      __builtin_amdgcn_wave_barrier();
      const int lps_log = 31 - __builtin_clz(64 / queued);  // 1 <= queued <= 64
      const int lps = 1 << lps_log;
      const int qi = lane >> lps_log, sub = lane & (lps - 1);
      const bool act = qi < queued;
      int code = 0;
      float dv = 0.f;
      int64_t m1 = 0, m2 = 0;
      if (act) {
        code = s_code[wave][qi];
        dv = s_dot[wave][qi];
        const int r = code >> 16, col = (code >> 8) & 0xff, slot = code & 0xff;
        const int run0 = s_run0[slot], n_runs = s_run0[slot + 1] - run0;
        const int4* rq = reinterpret_cast<const int4*>(args.pools.of(s_slot_ref[slot]).runs) + s_slot_roff[slot];
        for (int k = sub; k < n_runs; k += lps) {
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
      }
      for (int o = 1; o < lps; o <<= 1) {  // wave-uniform: every lane takes part
        m1 += __shfl_xor(m1, o, 64);
        m2 += __shfl_xor(m2, o, 64);
      }
      if (act && sub == 0) {
        const int r = code >> 16, col = (code >> 8) & 0xff, slot = code & 0xff;
        const int aa = tile * 32 + slot;
        const int64_t nq = s_slot_n[slot];
        const int64_t sum_s = m1 + nq * Cl;
        const int64_t sum_s2 = m2 + 2 * Cl * m1 + nq * Cl * Cl;
        const double score = normalise(double(dv), args.C, s_slot_weff[slot], s_slot_wsum[slot], int(nq), sum_s, sum_s2,
                                       D, shift, dc, args.exact_lattice, args.absolute);
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
    const int a = tile * 32 + m;
    const bool active = a < args.n_angles;
    const uint64_t below = (uint64_t(1) << lane) - 1;
    const int my_sub = grouped ? m / args.group_slots : 0;  // this lane's slot's bound group in the tile
    const int cb = c0 + wc * kWaveCols + cw16;
    const uint32_t in_width = bits_below(W - cb);
    if (!run_bf16) {  // screened out: nothing to score
      if (!sparse && active)
        for (int r = 0; r < kR; ++r) {
          const int row = r0 + wr0 + r;
          for (int j = 0; j < 16; ++j) {
            const int c = cb + j;
            if (row < H && c < W) F.planes[(size_t(row) * W + c) * args.n_angles + a] = NCC_MISSING;
          }
        }
      continue;  // the next tile starts with a block barrier every wave reaches
    }
    const int4 rg = reinterpret_cast<const int4*>(F.ranges)[active ? a : 0];
    const double wsum_eff = args.wsum_eff[active ? a : 0];
    const double num_slack = double(Ci) * fabs(wsum_eff);  // |(mean(S) - C) * wsum| <= C * |wsum|
    // planes in [row][col][angle] order: the 32 slots of one anchor are adjacent
    double* plane = F.planes + (active ? a : 0);
    // The bound test runs in FP32 (it only selects the windows scored exactly;
    // the bound sits a tenth under the peak threshold, far beyond FP32's
    // rounding), one bit per window (wave row r, anchor j: bit 16 r + j); the
    // kept windows are then queued one per lane per pass, so the scoring code
    // appears once (64 bits of kept windows per four rows).
    const float prune_f = float(args.prune_below), slack_f = float(num_slack);
    // this lane's 16 anchors c_j = cb + j: bits of those inside the frame
    // (in_width) and inside the angle's valid column range (in_range)
    const uint32_t in_range = bits_below(min(W, rg.w) - cb) & ~bits_below(rg.z - cb);
    static_for<kR / 4>([&](auto qc) __attribute__((always_inline)) {
      constexpr int q = decltype(qc)::value;  // rows 4 q .. 4 q + 3 of the wave
      float dot[4][16];
      uint64_t keep = 0;
      static_for<4>([&](auto rc) __attribute__((always_inline)) {
        constexpr int rl = decltype(rc)::value;
        constexpr int r = 4 * q + rl;  // wave row; block row wr0 + r
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(acc[r][i]), __float_as_uint(acc[r][8 + i]),
                                                           false, false);
          // first result: anchors 16 g + 8 (i / 4) + (i % 4); second: the same + 4
          dot[rl][8 * (i / 4) + (i % 4)] = __uint_as_float(sw[0]);
          dot[rl][8 * (i / 4) + 4 + (i % 4)] = __uint_as_float(sw[1]);
        }
        const int row = r0 + wr0 + r;
        const bool row_in = active && row < H && row >= rg.x && row < rg.y;
        uint32_t pass = 0xffffu;
        if (prune) {
          pass = 0;
          const float* isd = isd_row(my_sub, r) + cw16;
#pragma unroll
          for (int j = 0; j < 16; ++j)
            pass |= uint32_t((isd[j] < 0.f) | ((fabsf(dot[rl][j]) + slack_f) * isd[j] >= prune_f)) << j;
        }
        const uint32_t kr = (row_in ? in_range : 0u) & pass;
        keep |= uint64_t(kr) << (16 * rl);
        if (!sparse) {
          uint32_t miss = (active && row < H ? in_width : 0u) & ~kr;
          while (miss) {
            const int j = __builtin_ctz(miss);
            plane[(size_t(row) * W + cb + j) * args.n_angles] = NCC_MISSING;
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
          for (int qq = 1; qq < 64; ++qq) dv = b == qq ? dot[qq / 16][qq % 16] : dv;
          const int pos = queued + __popcll(mask & below);
          s_code[wave][pos] = ((wr0 + 4 * q + (b >> 4)) << 16) | ((cw16 + (b & 15)) << 8) | m;
          s_dot[wave][pos] = dv;
          keep &= keep - 1;
        }
        queued += cnt;
      }
    });
    if (queued) flush();
  }
}

template <class Tpix, class RunAcc, bool kScreen, int kR, bool kMulti>
__global__ void __launch_bounds__(kMaxThreads, min_blocks<kR>()) k_coarse_mfma(const MfmaArgs args) {
  coarse_tile<Tpix, RunAcc, kScreen, kR, kMulti>(args, int(blockIdx.z) / args.n_groups,
                                                 int(blockIdx.z) % args.n_groups, int(blockIdx.x), int(blockIdx.y));
}

// The coarse work queue: every (job, frame) unit contributes gx x gy blocks
// per group of angle tiles; its items are numbered from first[unit], tile
// group outermost, then block rows, then block columns (neighbouring items
// share image rows in L2).
struct QueueUnit { int job, frame, gx, gy; };
struct QueueArgs {
  const MfmaArgs* args;      // per job (device)
  const QueueUnit* units;    // this launch's units (device)
  const int* first;          // first item of each unit, n_units + 1 entries (device)
  int n_units, n_items;
  unsigned* counter;         // next item to take; zero when the launch starts
};

// Runs one queue item: finds its unit (the last whose first item is <= item)
// and its (angle-tile group, block row, block column).  The unit and job are
// wave-uniform: kept in scalar registers, so the job's arguments are read
// with scalar loads.
template <class Tpix, class RunAcc, bool kScreen, int kR, bool kMulti>
__device__ __forceinline__ void queue_item(const QueueArgs& q, const int item) {
  int lo = 0, hi = q.n_units - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (__builtin_amdgcn_readfirstlane(q.first[mid]) <= item) lo = mid;
    else hi = mid - 1;
  }
  const int job = __builtin_amdgcn_readfirstlane(q.units[lo].job);
  const int frame = __builtin_amdgcn_readfirstlane(q.units[lo].frame);
  const int gx = __builtin_amdgcn_readfirstlane(q.units[lo].gx);
  const int gy = __builtin_amdgcn_readfirstlane(q.units[lo].gy);
  const int local = item - __builtin_amdgcn_readfirstlane(q.first[lo]);
  const int per_group = gx * gy;
  const int group = local / per_group, rem = local - group * per_group;
  coarse_tile<Tpix, RunAcc, kScreen, kR, kMulti>(q.args[job], frame, group, rem % gx, rem / gx);
}

// Persistent coarse search: each block takes items from the queue until it
// is drained (every block exits once the counter passes n_items).  The jobs
// of one launch share kR and the block's thread count.
template <class Tpix, class RunAcc, bool kScreen, int kR, bool kMulti>
__global__ void __launch_bounds__(kMaxThreads, min_blocks<kR>()) k_coarse_queue(const QueueArgs q) {
  __shared__ int s_item;
  for (;;) {
    __syncthreads();  // the previous item is done with LDS and s_item
    if (threadIdx.x == 0) s_item = int(atomicAdd(q.counter, 1u));
    __syncthreads();
    const int item = __builtin_amdgcn_readfirstlane(s_item);
    if (item >= q.n_items) return;
    queue_item<Tpix, RunAcc, kScreen, kR, kMulti>(q, item);
  }
}

// Launchers of one rows-per-wave instantiation (ncc_mfma_r<kR>.hip):
// variant bit 0 the FP8 screen, bit 1 64-bit run sums, bit 2 u16 pixels,
// bit 3 several angle tiles per block.
int coarse_launch_r4(int variant, dim3 grid, int threads, size_t lds, hipStream_t st, const MfmaArgs& a);
int coarse_launch_r8(int variant, dim3 grid, int threads, size_t lds, hipStream_t st, const MfmaArgs& a);
int coarse_queue_r4(int variant, int threads, size_t lds, hipStream_t st, const QueueArgs& q, int n_cus);
int coarse_queue_r8(int variant, int threads, size_t lds, hipStream_t st, const QueueArgs& q, int n_cus);

// Instantiates the launchers of rows-per-wave kR.
template <int kR>
struct CoarseLaunchers {
  // calls fn.template operator()<Tpix, RunAcc, kScreen, kMulti>() for variant v
  template <class Fn>
  static void dispatch(int v, Fn&& fn) {
    auto multi = [&](auto tp, auto ra, auto scr) {
      using TP = typename decltype(tp)::type;
      using RA = typename decltype(ra)::type;
      constexpr bool kS = decltype(scr)::value;
      if (v & 8) fn.template operator()<TP, RA, kS, true>();
      else fn.template operator()<TP, RA, kS, false>();
    };
    using u8 = std::type_identity<uint8_t>;
    using u16 = std::type_identity<uint16_t>;
    using a32 = std::type_identity<uint32_t>;
    using a64 = std::type_identity<uint64_t>;
    using T = std::true_type;
    using F = std::false_type;
    switch (v & 7) {
      case 0: multi(u8{}, a32{}, F{}); break;
      case 1: multi(u8{}, a32{}, T{}); break;
      case 2: multi(u8{}, a64{}, F{}); break;
      case 3: multi(u8{}, a64{}, T{}); break;
      case 4: multi(u16{}, a32{}, F{}); break;
      case 5: multi(u16{}, a32{}, T{}); break;
      case 6: multi(u16{}, a64{}, F{}); break;
      default: multi(u16{}, a64{}, T{}); break;
    }
  }
  static int grid_launch(int v, dim3 grid, int threads, size_t lds, hipStream_t st, const MfmaArgs& a) {
    dispatch(v, [&]<class TP, class RA, bool kScreen, bool kMulti>() {
      hipLaunchKernelGGL((k_coarse_mfma<TP, RA, kScreen, kR, kMulti>), grid, dim3(threads), lds, st, a);
    });
    return hipGetLastError() == hipSuccess ? 0 : -1000;
  }
  static int queue_launch(int v, int threads, size_t lds, hipStream_t st, const QueueArgs& q, int n_cus) {
    dispatch(v, [&]<class TP, class RA, bool kScreen, bool kMulti>() {
      int per_cu = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
              &per_cu, reinterpret_cast<const void*>(&k_coarse_queue<TP, RA, kScreen, kR, kMulti>), threads, lds) !=
              hipSuccess || per_cu <= 0)
        per_cu = 1;
      const int blocks = q.n_items < n_cus * per_cu ? q.n_items : n_cus * per_cu;
      hipLaunchKernelGGL((k_coarse_queue<TP, RA, kScreen, kR, kMulti>), dim3(blocks > 0 ? blocks : 1), dim3(threads),
                         lds, st, q);
    });
    return hipGetLastError() == hipSuccess ? 0 : -1000;
  }
};

}  // namespace ncc
