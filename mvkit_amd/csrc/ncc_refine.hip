// Coarse-to-fine tracking (matcher.cpp:630-761), continuous bilinear scoring
// (matcher.cpp:121-157, kernels_avx2.cpp:327-477) and Gauss-Newton subpixel /
// angle refinement with its level-0 recovery search (matcher.cpp:170-240,
// 853-1061), all in FP64 and compiled with -ffp-contract=off so each sample
// coordinate, bilinear value and window moment is the reference's value.
// Sums over samples are block tree reductions (the reference sums in 4 lanes),
// so reduced totals may differ from the reference in the last bits.
#include <algorithm>
#include <cmath>
#include <type_traits>

#include "ncc_common.hip.h"
#include "../../include/ncchip.h"

namespace ncc {

struct PatchRef {
  NccPatch P;
  const int32_t* rc;
  const double* w;
  const int32_t* wruns;
  const double* wrun_w;
};

__device__ __forceinline__ PatchRef patch_ref(const Pools& pools, int32_t ref) {
  const NccPatchPool& pool = pools.of(ref);
  PatchRef out;
  out.P = pool.patches[ref & 0xffffff];
  out.rc = pool.rc + out.P.sample_off;
  out.w = pool.w + out.P.sample_off;
  out.wruns = pool.wruns + 4 * out.P.wrun_off;
  out.wrun_w = pool.wrun_w + out.P.wrun_off;
  return out;
}
__device__ __forceinline__ double ref_angle(const Pools& pools, int32_t ref) { return pools.patch(ref).angle; }

// A device struct in registers, its 8-byte words loaded in one round (a
// field-by-field copy whose fields are used as they arrive waits for each
// load before issuing the next).
template <class T>
__device__ __forceinline__ T load_words(const T& src) {
  static_assert(sizeof(T) % 8 == 0 && alignof(T) >= 8, "a struct of whole 8-byte words");
  constexpr int kWords = sizeof(T) / 8;
  const uint2* p = reinterpret_cast<const uint2*>(&src);
  uint2 w[kWords];
#pragma unroll
  for (int k = 0; k < kWords; ++k) w[k] = p[k];
  T v;
  __builtin_memcpy(&v, w, sizeof v);
  return v;
}
__device__ __forceinline__ NccPatch load_patch(const Pools& pools, int32_t ref) {
  return load_words(pools.patch(ref));
}

// ------------------------------------------------------------------ tracking
struct TrackArgs {
  const NccFrame* frames;     // per-frame descriptors (device)
  int n_frames;
  int level;                  // tracked level l-1
  Pools pools;
  const int32_t* refs;
  int n_refs;
  const NccMatch* in;         // frame f's candidates at f * cap
  const int32_t* in_count;
  int cap;
  NccMatch* raw;              // tracked pose of every input candidate (same slot)
  NccMatch* out;              // kept, distinct tracked poses (unordered)
  int32_t* out_count;
  int32_t* table;             // per frame: open-addressing set of kept raw slots (-1: empty)
  int table_size;             // power of two
  double keep;                // 0.4 * MinScore
  unsigned long long* work;   // algorithmic work counter: samples over the evaluated 5x5 windows
  double tol, step, offset;
  int exact_lattice, absolute;
  int tile_cap;               // LDS tile capacity in pixels
  // MFMA tracking (k_track_mfma): LDS tile bounds and carve-up
  int mt_h, mt_w, mt_pitch;   // tile rows / columns, plane pitch (elements)
  int mt_s, mt_p1, mt_p2, mt_pl;  // byte offsets of S, row prefixes, planes
  double x_max;               // bound of |S - dc 4^l| at this level
  // the candidate's frame, filled in-kernel
  NccLevel lv;
  double dc, crop_r, crop_c;
  int64_t D;
  int sr, sc, sh, sw;
  const NccFrame* dom;        // the frame when it has a reduced domain, else null
};

// Loads frame f into the per-block copy of the arguments: the level's
// descriptor and the frame's fields in one round of loads, then the
// reference pixel through its pointer.
__device__ __forceinline__ void bind_frame(TrackArgs& a, int f) {
  const NccFrame& F = a.frames[f];
  const NccLevel lv = load_words(F.levels[a.level]);
  const void* const dc_src = F.dc_src;
  const int32_t dc_wide = F.dc_wide;
  const int4 st = make_int4(F.storage0[0], F.storage0[1], F.storage0[2], F.storage0[3]);
  const double crop_r = F.crop_r, crop_c = F.crop_c;
  const bool has_dom = F.dom_sat != nullptr;
  const uint32_t dc = dc_wide ? uint32_t(*static_cast<const uint16_t*>(dc_src)) : uint32_t(*static_cast<const uint8_t*>(dc_src));
  a.lv = lv;
  a.dc = double(dc);
  a.D = int64_t(dc) << lv.shift;
  a.crop_r = crop_r;
  a.crop_c = crop_c;
  if (a.level == 0) { a.sr = st.x; a.sc = st.y; a.sh = st.z; a.sw = st.w; }
  else { a.sr = 0; a.sc = 0; a.sh = lv.height; a.sw = lv.width; }
  a.dom = has_dom ? &F : nullptr;
}

// Pixel source of a tracking window: an LDS tile of the candidate's
// neighbourhood, or the level in global memory when the tile is too large.
template <class T>
struct TileReader {
  const T* tile;
  int r0, c0, tw;
  __device__ __forceinline__ uint32_t at(int r, int c) const { return tile[(r - r0) * tw + (c - c0)]; }
};
template <class T>
struct LevelReader {
  const T* data;
  int pitch;
  __device__ __forceinline__ uint32_t at(int r, int c) const { return data[size_t(r) * pitch + c]; }
};

// Score of one window from its dot product and the exact centred moments
// t1 = sum(S - D), t2 = sum((S - D)^2), with the reference's prefix-kernel
// normalisation and low-variance rule (kernels_avx2.cpp:209-238, 128-129).
__device__ __forceinline__ double window_score(const TrackArgs& a, const NccPatch& P, double dot, int64_t t1,
                                               int64_t t2) {
  const double sum = ldexp(double(t1), -a.lv.shift), sq = ldexp(double(t2), -2 * a.lv.shift);
  const double cnt = double(P.n);
  const double var = sq - sum * sum / cnt;
  dot = dot - sum / cnt * P.wsum;
  double den = var;
  if (!(var > 128 * __DBL_EPSILON__ * sq)) {
    if (a.exact_lattice || (__int128)t2 * P.n == (__int128)t1 * t1) return NCC_MISSING;
    den = ldexp(double((__int128)t2 * P.n - (__int128)t1 * t1), -2 * a.lv.shift) / cnt;
  }
  const double sc = clamp1(dot / sqrt(den));
  return a.absolute ? fabs(sc) : sc;
}

__device__ __forceinline__ bool window_inside(const TrackArgs& a, const NccPatch& P, int row, int col) {
  return row + P.rmin >= 0 && col + P.cmin >= 0 && row + P.rmax < a.lv.height && col + P.cmax < a.lv.width &&
         row + P.rmin >= a.sr && col + P.cmin >= a.sc && row + P.rmax < a.sr + a.sh && col + P.cmax < a.sc + a.sw;
}

// A tracking window that is scored: inside the level's valid region and, for
// a reduced domain, at a pose that may reach the domain (matcher.cpp:673-751,
// may_contain at this level's spacing).
__device__ __forceinline__ bool window_ok(const TrackArgs& a, const NccPatch& P, int row, int col) {
  if (!window_inside(a, P, row, col)) return false;
  if (!a.dom) return true;
  return may_contain(*a.dom, a.crop_r + a.step * (row + P.orow) + a.offset,
                     a.crop_c + a.step * (col + P.ocol) + a.offset, a.step);
}

// Exchanges of doubles for the transposed butterfly, as pairs of dwords:
// permlane32_swap gives x' = (x lanes 0-31, y lanes 0-31) and
// y' = (x lanes 32-63, y lanes 32-63); permlane16_swap does the same per
// pair of 16-lane rows.
__device__ __forceinline__ double sum_swap32(double x, double y) {
  const uint64_t a = uint64_t(__double_as_longlong(x)), b = uint64_t(__double_as_longlong(y));
  const auto lo = __builtin_amdgcn_permlane32_swap(uint32_t(a), uint32_t(b), false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap(uint32_t(a >> 32), uint32_t(b >> 32), false, false);
  const double x2 = __longlong_as_double(int64_t((uint64_t(hi[0]) << 32) | lo[0]));
  const double y2 = __longlong_as_double(int64_t((uint64_t(hi[1]) << 32) | lo[1]));
  return x2 + y2;
}
__device__ __forceinline__ double sum_swap16(double x, double y) {
  const uint64_t a = uint64_t(__double_as_longlong(x)), b = uint64_t(__double_as_longlong(y));
  const auto lo = __builtin_amdgcn_permlane16_swap(uint32_t(a), uint32_t(b), false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap(uint32_t(a >> 32), uint32_t(b >> 32), false, false);
  const double x2 = __longlong_as_double(int64_t((uint64_t(hi[0]) << 32) | lo[0]));
  const double y2 = __longlong_as_double(int64_t((uint64_t(hi[1]) << 32) | lo[1]));
  return x2 + y2;
}

// Sums 16 values over the wave with one transposed butterfly: at each of the
// first four halvings a lane keeps half of its values and adds its partner's
// copies of them, so the totals cost 17 exchanges instead of 96.  On return
// lane L holds the total of value (L >> 2) & 15.
__device__ __forceinline__ double wave_transpose_sum16(double (&v)[16]) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int j = 0; j < 8; ++j) v[j] = sum_swap32(v[j], v[j + 8]);  // lanes 32-63 keep j + 8
#pragma unroll
  for (int j = 0; j < 4; ++j) v[j] = sum_swap16(v[j], v[j + 4]);  // odd rows keep j + 4
#pragma unroll
  for (int half = 2, m = 8; half >= 1; half >>= 1, m >>= 1) {
    const bool hi = (lane & m) != 0;
#pragma unroll
    for (int j = 0; j < half; ++j) {
      // both values loaded, then selected (a select of addresses would keep v in memory)
      const double lo_v = v[j], hi_v = v[j + half];
      const double send = hi ? lo_v : hi_v;
      const double keep = hi ? hi_v : lo_v;
      v[j] = keep + __shfl_xor(send, m, 64);
    }
  }
  double t = v[0];
  t += __shfl_xor(t, 2, 64);
  t += __shfl_xor(t, 1, 64);
  return t;
}

// The five windows (row, col-2 .. col+2) of one template row offset, evaluated
// by one wave with lanes striding over the samples, so each sample's position
// and weight are loaded once for the five windows.  Per sample and window the
// centred value x = S - D (an integer) feeds the dot product and both moments.
// Acc = double keeps the moments exact while every partial sum stays below
// 2^53 (the host checks n * max|x|^2 for the level); Acc = int64_t otherwise.
// The dot product is sum(x * w) scaled by 4^-l at the end, the reference's
// per-sample products up to a power of two; the sums run in a different order
// than the reference's single accumulator, so scores agree to ~1e-15.  Lane L
// reports window k = (L >> 2) & 15 when k < 5 (score, window = k), and
// window = -1 otherwise; the window scores are computed lane-parallel.
template <class Acc, class Reader>
__device__ __forceinline__ void track_row(const TrackArgs& a, const PatchRef& p, int row, int col, const Reader& rd,
                                          double& score, int& window) {
  const NccPatch& P = p.P;
  bool in[5];
  bool any = false;
#pragma unroll
  for (int k = 0; k < 5; ++k) { in[k] = window_inside(a, P, row, col - 2 + k); any |= in[k]; }
  const int lane = threadIdx.x & 63;
  const int q = (lane >> 2) & 15;  // the value (and window) this lane reports
  window = q < 5 ? q : -1;
  score = NCC_MISSING;
  if (!any) return;
  // windows partly outside read clamped columns and are discarded below
  Acc t1[5] = {0, 0, 0, 0, 0}, t2[5] = {0, 0, 0, 0, 0};
  double dot[5] = {0, 0, 0, 0, 0};
  int kmin = 0, kmax = 4;
  while (!in[kmin]) ++kmin;
  while (!in[kmax]) --kmax;
  const int32_t Di = int32_t(a.D);
  constexpr int U = 8;
  // Branch-free over the U samples of a chunk: samples past the end re-read
  // the chunk's first sample with x forced to 0, so all 5U loads issue before
  // the first wait.  Weights are 0 when the dot product uses weight runs.
  const bool dense_w = !P.use_wruns;
  for (int base = lane; base < P.n; base += 64 * U) {
    int32_t rcv[U];
    double wv[U];
    bool ok[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = base + 64 * u;
      ok[u] = i < P.n;
      const int j = ok[u] ? i : base;
      rcv[u] = p.rc[j];
      const double w = p.w[j];
      wv[u] = ok[u] && dense_w ? w : 0.0;
    }
    uint32_t v[U][5];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int r = row + sample_row(rcv[u]), c = col + sample_col(rcv[u]);
#pragma unroll
      for (int k = 0; k < 5; ++k) v[u][k] = rd.at(r, c - 2 + min(max(k, kmin), kmax));
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        if constexpr (std::is_same<Acc, double>::value) {
          const double x = ok[u] ? double(int32_t(v[u][k]) - Di) : 0.0;
          t1[k] += x;
          t2[k] = fma(x, x, t2[k]);
          dot[k] = fma(x, wv[u], dot[k]);
        } else {
          const int64_t x = ok[u] ? int64_t(v[u][k]) - a.D : 0;
          t1[k] += x;
          t2[k] += x * x;
          dot[k] = fma(double(x), wv[u], dot[k]);
        }
      }
    }
  }
  if (P.use_wruns) {
    for (int j = lane; j < P.n_wruns; j += 64) {
      const int4 run = reinterpret_cast<const int4*>(p.wruns)[j];
      const double w = p.wrun_w[j];
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        const int cc = col - 2 + min(max(k, kmin), kmax);
        int64_t rs = 0;
        for (int c = run.y; c <= run.z; ++c) rs += rd.at(row + run.x, cc + c);
        dot[k] += double(rs - int64_t(run.z - run.y + 1) * a.D) * w;
      }
    }
  }
  const double unit = ldexp(1.0, -a.lv.shift);
  double d = 0;
  int64_t T1 = 0, T2 = 0;
  if constexpr (std::is_same<Acc, double>::value) {
    // one transposed reduction of the 15 sums; lane L ends with value
    // (L >> 2) & 15: dot products 0-4, sum(x) 5-9, sum(x^2) 10-14
    double v[16];
#pragma unroll
    for (int k = 0; k < 5; ++k) { v[k] = dot[k]; v[5 + k] = t1[k]; v[10 + k] = t2[k]; }
    v[15] = 0;
    const double tot = wave_transpose_sum16(v);
    d = tot;
    T1 = int64_t(__shfl(tot, ((q + 5) & 15) * 4 + (lane & 3), 64));
    T2 = int64_t(__shfl(tot, ((q + 10) & 15) * 4 + (lane & 3), 64));
  } else {
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const int64_t s1 = wave_sum_i64(t1[k]), s2 = wave_sum_i64(t2[k]);
      const double dk = wave_sum(dot[k]);
      if (k == q) { d = dk; T1 = s1; T2 = s2; }
    }
  }
  if (q < 5 && window_ok(a, P, row, col - 2 + q)) score = window_score(a, P, d * unit, T1, T2);
}

constexpr int kMaxEligible = 1024;

// Ordered compaction of the patch indices whose angle is within `tol` of
// `angle` (query order kept: it defines the reference's tie-breaking).
// Works for any block size that is a multiple of the wave size.
__device__ __forceinline__ int eligible_patches(const Pools& pools, const int32_t* refs, int n_refs, double angle, double tol,
                                int* elig, int* wave_cnt) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, n_waves = blockDim.x >> 6;
  int n_e = 0;
  for (int base = 0; base < n_refs; base += blockDim.x) {
    const int k = base + threadIdx.x;
    const bool f = k < n_refs && fabs(ref_angle(pools, refs[k]) - angle) <= tol;
    const uint64_t m = __ballot(f);
    const int pre = __popcll(m & ((uint64_t(1) << lane) - 1));
    if (lane == 0) wave_cnt[wave] = __popcll(m);
    __syncthreads();
    int off = n_e;
    for (int w = 0; w < wave; ++w) off += wave_cnt[w];
    if (f && off + pre < kMaxEligible) elig[off + pre] = k;
    for (int w = 0; w < n_waves; ++w) n_e += wave_cnt[w];
    __syncthreads();
  }
  return min(n_e, kMaxEligible);
}

// eligible_patches over angles cached in LDS (ang[k] = angle of refs[k]).
__device__ __forceinline__ int eligible_cached(const double* ang, int n_refs, double angle, double tol, int* elig,
                                               int* wave_cnt) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, n_waves = blockDim.x >> 6;
  int n_e = 0;
  for (int base = 0; base < n_refs; base += blockDim.x) {
    const int k = base + threadIdx.x;
    const bool f = k < n_refs && fabs(ang[k] - angle) <= tol;
    const uint64_t m = __ballot(f);
    const int pre = __popcll(m & ((uint64_t(1) << lane) - 1));
    if (lane == 0) wave_cnt[wave] = __popcll(m);
    __syncthreads();
    int off = n_e;
    for (int w = 0; w < wave; ++w) off += wave_cnt[w];
    if (f && off + pre < kMaxEligible) elig[off + pre] = k;
    for (int w = 0; w < n_waves; ++w) n_e += wave_cnt[w];
    __syncthreads();
  }
  return min(n_e, kMaxEligible);
}

// Post-tracking merge and filter of one frame (matcher.cpp:241-248, 758-760):
// a tracked pose with a finite score >= keep is appended to the frame's
// output list unless an identical (row, col, angle) is already there.
// Identical tracked poses come from the same window, so they carry the same
// score, and the kept set equals the reference's merged list (its order is
// restored by the host where it matters).  Called by one thread.
__device__ __forceinline__ uint32_t pose_hash(double row, double col, double angle) {
  uint64_t h = uint64_t(__double_as_longlong(row + 0.0)) * 0x9E3779B97F4A7C15ull;
  h ^= uint64_t(__double_as_longlong(col + 0.0)) + 0x632BE59BD9B4E019ull + (h << 6) + (h >> 2);
  h ^= uint64_t(__double_as_longlong(angle + 0.0)) * 0xC2B2AE3D27D4EB4Full;
  h ^= h >> 29;
  h *= 0xBF58476D1CE4E5B9ull;
  h ^= h >> 32;
  return uint32_t(h);
}
__device__ __forceinline__ double load_shared_double(const double* p) {
  return __longlong_as_double(
      __hip_atomic_load(reinterpret_cast<const long long*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}
__device__ __forceinline__ void keep_distinct(const TrackArgs& a, int f, int ci, const NccMatch& res) {
  const size_t base = size_t(f) * a.cap;
  a.raw[base + ci] = res;
  if (!isfinite(res.score) || res.score < a.keep) return;
  __threadfence();  // the raw pose is visible before its slot is published
  int32_t* T = a.table + size_t(f) * a.table_size;
  const uint32_t mask = uint32_t(a.table_size - 1);
  uint32_t h = pose_hash(res.row, res.col, res.angle) & mask;
  for (int probe = 0; probe < a.table_size; ++probe, h = (h + 1) & mask) {
    const int prev = atomicCAS(T + h, -1, ci);
    if (prev == -1) {
      const int k = atomicAdd(a.out_count + f, 1);
      if (k < a.cap) a.out[base + k] = res;
      return;
    }
    __threadfence();
    const NccMatch* o = a.raw + base + prev;
    if (load_shared_double(&o->row) == res.row && load_shared_double(&o->col) == res.col &&
        load_shared_double(&o->angle) == res.angle)
      return;  // duplicate of a kept pose
  }
}

constexpr int kTrackThreads = 256;

__device__ __forceinline__ void anchor(const TrackArgs& a, const NccMatch& cd, const NccPatch& P, int& row, int& col) {
  row = int(round((cd.row - a.crop_r - a.offset) / a.step - P.orow));
  col = int(round((cd.col - a.crop_c - a.offset) / a.step - P.ocol));
}

// Block-wide (score, index) maximum, ties to the lower index; every thread
// gets the result.
__device__ __forceinline__ void block_argmax(double& s, int& i, double* ws, int* wi) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) {
    const double o = __shfl_xor(s, m, 64);
    const int oi = __shfl_xor(i, m, 64);
    if (o > s || (o == s && oi < i)) { s = o; i = oi; }
  }
  if (lane == 0) { ws[wave] = s; wi[wave] = i; }
  __syncthreads();
  s = ws[0];
  i = wi[0];
  for (int w = 1; w < kTrackThreads / 64; ++w)
    if (ws[w] > s || (ws[w] == s && wi[w] < i)) { s = ws[w]; i = wi[w]; }
  __syncthreads();
}

// One block per candidate: score the 5x5 anchors around the candidate for every
// query patch within the angle window and keep the first strict maximum in the
// reference's (patch, dr, dc) order (matcher.cpp:670-756).  Windows are
// spread over the waves; the candidate's neighbourhood is staged in LDS.
template <class T, class Acc>
__global__ void __launch_bounds__(kTrackThreads, 3) k_track(const TrackArgs args) {
  extern __shared__ uint32_t tile_words[];
  T* tile = reinterpret_cast<T*>(tile_words);
  __shared__ int elig[kMaxEligible];
  __shared__ int wave_cnt[kTrackThreads / 64];
  __shared__ int box[4];
  __shared__ double red_s[kTrackThreads / 64];
  __shared__ int red_i[kTrackThreads / 64];
  __shared__ int off[NCC_MAX_BATCH + 1];
  __shared__ unsigned long long s_work;  // 25 * samples of the block's candidates
  if (threadIdx.x == 0) s_work = 0;
  list_offsets(args.in_count, args.n_frames, args.cap, off);
  const int total = off[args.n_frames];
  for (int g = blockIdx.x; g < total; g += gridDim.x) {
  TrackArgs a = args;
  // (frame, entry) are block-uniform: keep them, and what is loaded through
  // them, in scalar registers
  const int f = __builtin_amdgcn_readfirstlane(list_frame(off, a.n_frames, g));
  const int ci = __builtin_amdgcn_readfirstlane(g - off[f]);
  bind_frame(a, f);
  const NccMatch cd = a.in[size_t(f) * a.cap + ci];
  const int n_e = eligible_patches(a.pools, a.refs, a.n_refs, cd.angle, a.tol, elig, wave_cnt);
  // union of the windows' sample rectangles, clipped to the valid region
  const int lo_r = max(0, a.sr), lo_c = max(0, a.sc);
  const int hi_r = min(a.lv.height, a.sr + a.sh) - 1, hi_c = min(a.lv.width, a.sc + a.sw) - 1;
  if (threadIdx.x == 0) { box[0] = INT_MAX; box[1] = INT_MIN; box[2] = INT_MAX; box[3] = INT_MIN; }
  __syncthreads();
  unsigned long long samples = 0;  // the work counter, in the same round of descriptor loads
  for (int e = threadIdx.x; e < n_e; e += kTrackThreads) {
    const NccPatch& P = a.pools.patch(a.refs[elig[e]]);
    samples += 25ull * uint32_t(P.n);
    int row, col;
    anchor(a, cd, P, row, col);
    atomicMin(&box[0], max(lo_r, row - 2 + P.rmin));
    atomicMax(&box[1], min(hi_r, row + 2 + P.rmax));
    atomicMin(&box[2], max(lo_c, col - 2 + P.cmin));
    atomicMax(&box[3], min(hi_c, col + 2 + P.cmax));
  }
  if (args.work && samples) atomicAdd(&s_work, samples);
  __syncthreads();
  const int r0 = box[0], c0 = box[2];
  const int th = box[1] - box[0] + 1, tw = box[3] - box[2] + 1;
  const T* lvp = static_cast<const T*>(a.lv.data);
  const bool use_tile = n_e > 0 && th > 0 && tw > 0 && th * tw <= a.tile_cap;
  if (use_tile) {
    constexpr int U = 8;
    for (int base = threadIdx.x; base < th * tw; base += kTrackThreads * U) {
      T v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int i = base + kTrackThreads * u;
        v[u] = i < th * tw ? lvp[size_t(r0 + i / tw) * a.lv.pitch + c0 + i % tw] : T(0);
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (base + kTrackThreads * u < th * tw) tile[base + kTrackThreads * u] = v[u];
    }
  }
  __syncthreads();
  double best = NCC_MISSING;
  int best_i = 0x7fffffff;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  for (int task = wave; task < n_e * 5; task += kTrackThreads / 64) {
    // the task is wave-uniform: keep the patch descriptor in scalar registers
    const int32_t ref = __builtin_amdgcn_readfirstlane(a.refs[__builtin_amdgcn_readfirstlane(elig[task / 5])]);
    const PatchRef p = patch_ref(a.pools, ref);
    int row, col;
    anchor(a, cd, p.P, row, col);
    row = __builtin_amdgcn_readfirstlane(row);
    col = __builtin_amdgcn_readfirstlane(col);
    const int dr = task % 5 - 2;
    double sc;
    int k;
    if (use_tile) track_row<Acc>(a, p, row + dr, col, TileReader<T>{tile, r0, c0, tw}, sc, k);
    else track_row<Acc>(a, p, row + dr, col, LevelReader<T>{lvp, a.lv.pitch}, sc, k);
    if (k >= 0 && sc > best) { best = sc; best_i = (task / 5) * 25 + (dr + 2) * 5 + k; }
  }
  block_argmax(best, best_i, red_s, red_i);  // first strict maximum in (patch, dr, dc) order
  if (threadIdx.x == 0) {
    NccMatch res = cd;
    res.score = NCC_MISSING;
    if (best > NCC_MISSING) {
      const int wi = best_i;
      const NccPatch& P = a.pools.patch(a.refs[elig[wi / 25]]);
      int row, col;
      anchor(a, cd, P, row, col);
      res.row = a.crop_r + a.step * (row + (wi % 25) / 5 - 2 + P.orow) + a.offset;
      res.col = a.crop_c + a.step * (col + wi % 5 - 2 + P.ocol) + a.offset;
      res.angle = P.angle;
      res.score = best;
    }
    keep_distinct(a, f, ci, res);
  }
  __syncthreads();  // the LDS of this candidate is reused by the next
  }
  // the work counter: the block's total, added once (one contended global
  // atomic per candidate would wait at the next barrier)
  if (args.work && threadIdx.x == 0 && s_work) atomicAdd(args.work, s_work);
}

// ------------------------------------------------- tracking, window-major
// k_track with every sample loaded once for all 25 windows of its patch: a
// wave takes a slice of one eligible patch's samples and accumulates the 25
// dot products sum(x * w) (x = S - D from the LDS tile, FP64), which one
// transposed butterfly reduces over the wave; the slices of a patch are added
// in a fixed order.  The exact moments sum(S) and sum(S^2) of the 25 windows
// come from sliding sums along the patch's support runs (lanes over runs,
// integers), and a patch whose dot product uses weight runs slides along those
// the same way.  Per sample this reads the pixel of each window once (as
// k_track does) but the sample's position and weight once instead of five
// times.  Windows outside the valid region read the nearest inside window's
// pixels and are discarded.  The tracked pose is the first strict maximum in
// the reference's (patch, dr, dc) order, as in k_track.
constexpr int kWinPatches = 32;  // eligible patches per candidate (the host guarantees the bound)
constexpr int kWinSlices = 4;    // sample slices per patch when at most 8 patches are eligible
constexpr int kWinRunCache = 512;  // support runs of a candidate's eligible patches cached in LDS

// Sums 32 values over the wave with one transposed butterfly (32 exchanges
// instead of 192); lane L ends with the total of value (L >> 1) & 31.
__device__ __forceinline__ double wave_transpose_sum32(double (&v)[32]) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int j = 0; j < 16; ++j) v[j] = sum_swap32(v[j], v[j + 16]);  // lanes 32-63 keep j + 16
#pragma unroll
  for (int j = 0; j < 8; ++j) v[j] = sum_swap16(v[j], v[j + 8]);    // odd 16-lane rows keep j + 8
#pragma unroll
  for (int half = 4, m = 8; half >= 1; half >>= 1, m >>= 1) {
    const bool hi = (lane & m) != 0;
#pragma unroll
    for (int j = 0; j < half; ++j) {
      // both values loaded, then selected (a select of addresses would keep v in memory)
      const double lo_v = v[j], hi_v = v[j + half];
      const double send = hi ? lo_v : hi_v;
      const double keep = hi ? hi_v : lo_v;
      v[j] = keep + __shfl_xor(send, m, 64);
    }
  }
  return v[0] + __shfl_xor(v[0], 1, 64);
}

// kX16: the tile holds x = S - D as int16 (every |S - D| of the level is
// below 2^15), so the dot loop reads x directly; else it holds S.
template <class T, bool kX16>
__global__ void __launch_bounds__(kTrackThreads, 3) k_track_win(const TrackArgs args) {
  using Tile = typename std::conditional<kX16, int16_t, T>::type;
  extern __shared__ uint32_t tile_words[];
  Tile* tile = reinterpret_cast<Tile*>(tile_words);
  __shared__ int elig[kMaxEligible];
  __shared__ int wave_cnt[kTrackThreads / 64];
  __shared__ int box[4];
  __shared__ double red_s[kTrackThreads / 64];
  __shared__ int red_i[kTrackThreads / 64];
  __shared__ int off[NCC_MAX_BATCH + 1];
  __shared__ unsigned long long s_work;  // 25 * samples of the block's candidates
  if (threadIdx.x == 0) s_work = 0;
  __shared__ NccPatch s_pat[kWinPatches];
  __shared__ int32_t s_ref[kWinPatches];
  __shared__ int s_row[kWinPatches], s_col[kWinPatches];
  __shared__ int s_lim[kWinPatches][4];          // inside windows: dr_lo, dr_hi, dc_lo, dc_hi (dr_lo > dr_hi: none)
  __shared__ double s_dot[kWinPatches][25];      // per (patch, slice) task
  __shared__ long long s_m1[kWinPatches][25], s_m2[kWinPatches][25];
  __shared__ int2 s_runs[kWinRunCache];          // support runs of the eligible patches: (row, begin | end << 16)
  __shared__ int s_run0[kWinPatches + 1];        // first run of each patch in s_runs
  list_offsets(args.in_count, args.n_frames, args.cap, off);
  const int total = off[args.n_frames];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  for (int g = blockIdx.x; g < total; g += gridDim.x) {
    TrackArgs a = args;
    const int f = __builtin_amdgcn_readfirstlane(list_frame(off, a.n_frames, g));
    const int ci = __builtin_amdgcn_readfirstlane(g - off[f]);
    bind_frame(a, f);
    const NccMatch cd = a.in[size_t(f) * a.cap + ci];
    const int n_e = min(eligible_patches(a.pools, a.refs, a.n_refs, cd.angle, a.tol, elig, wave_cnt), kWinPatches);
    const int lo_r = max(0, a.sr), lo_c = max(0, a.sc);
    const int hi_r = min(a.lv.height, a.sr + a.sh) - 1, hi_c = min(a.lv.width, a.sc + a.sw) - 1;
    if (threadIdx.x == 0) { box[0] = INT_MAX; box[1] = INT_MIN; box[2] = INT_MAX; box[3] = INT_MIN; }
    __syncthreads();
    unsigned long long samples = 0;
    for (int e = threadIdx.x; e < n_e; e += kTrackThreads) {
      const int32_t ref = a.refs[elig[e]];
      const NccPatch P = load_patch(a.pools, ref);
      samples += 25ull * uint32_t(P.n);
      int row, col;
      anchor(a, cd, P, row, col);
      s_pat[e] = P; s_ref[e] = ref; s_row[e] = row; s_col[e] = col;
      // the windows (row + dr, col + dc) inside the valid region
      const int dr_lo = max(-2, lo_r - row - P.rmin), dr_hi = min(2, hi_r - row - P.rmax);
      const int dc_lo = max(-2, lo_c - col - P.cmin), dc_hi = min(2, hi_c - col - P.cmax);
      const bool any = dr_lo <= dr_hi && dc_lo <= dc_hi;
      s_lim[e][0] = any ? dr_lo : 1; s_lim[e][1] = any ? dr_hi : 0; s_lim[e][2] = dc_lo; s_lim[e][3] = dc_hi;
      if (any) {
        atomicMin(&box[0], row + dr_lo + P.rmin);
        atomicMax(&box[1], row + dr_hi + P.rmax);
        atomicMin(&box[2], col + dc_lo + P.cmin);
        atomicMax(&box[3], col + dc_hi + P.cmax);
      }
    }
    if (a.work && samples) atomicAdd(&s_work, samples);
    __syncthreads();
    if (threadIdx.x == 0) {
      int t = 0;
      for (int e = 0; e < n_e; ++e) { s_run0[e] = t; t += s_pat[e].n_runs; }
      s_run0[n_e] = t;
    }
    // no eligible patch or no inside window leaves the box empty (INT_MAX,
    // INT_MIN): compare the bounds before forming sizes.  th * tw <= tile_cap
    // holds by the host's bound on the neighbourhood; it is checked anyway.
    const bool any_window = box[0] <= box[1] && box[2] <= box[3] &&
                            int64_t(box[1] - box[0] + 1) * (box[3] - box[2] + 1) <= a.tile_cap;  // block-uniform
    const int r0 = any_window ? box[0] : 0, c0 = any_window ? box[2] : 0;
    const int th = any_window ? box[1] - box[0] + 1 : 0, tw = any_window ? box[3] - box[2] + 1 : 0;
    const T* lvp = static_cast<const T*>(a.lv.data);
    const int32_t Di = int32_t(a.D);
    __syncthreads();  // s_run0
    const int n_runs_all = s_run0[n_e];
    const bool runs_cached = n_runs_all <= kWinRunCache;
    if (any_window) {
      // the runs (at most two per thread) and the tile in one round of loads
      int4 myrun[2];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int j = threadIdx.x + kTrackThreads * u;
        if (runs_cached && j < n_runs_all) {
          int e = 0;
          while (s_run0[e + 1] <= j) ++e;
          myrun[u] = reinterpret_cast<const int4*>(a.pools.of(s_ref[e]).runs)[s_pat[e].run_off + j - s_run0[e]];
        }
      }
      constexpr int U = 16;
      for (int base = threadIdx.x; base < th * tw; base += kTrackThreads * U) {
        T v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int i = base + kTrackThreads * u;
          v[u] = i < th * tw ? lvp[size_t(r0 + i / tw) * a.lv.pitch + c0 + i % tw] : T(0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (base + kTrackThreads * u < th * tw) {
            if constexpr (kX16) tile[base + kTrackThreads * u] = int16_t(int32_t(v[u]) - Di);
            else tile[base + kTrackThreads * u] = v[u];
          }
      }
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int j = threadIdx.x + kTrackThreads * u;
        if (runs_cached && j < n_runs_all)
          s_runs[j] = make_int2(myrun[u].x, (myrun[u].y & 0xffff) | (myrun[u].z << 16));
      }
    }
    __syncthreads();
    const int Q = n_e <= 8 ? kWinSlices : 1;
    // ---- dot products: task (patch e, slice q)
    for (int task = wave; any_window && task < n_e * Q; task += kTrackThreads / 64) {
      const int e = task / Q, q = task - e * Q;
      const NccPatch& P = s_pat[e];
      if (P.use_wruns) continue;  // wave-uniform: its dot product comes from the weight runs below
      const int dr_lo = s_lim[e][0], dr_hi = s_lim[e][1], dc_lo = s_lim[e][2], dc_hi = s_lim[e][3];
      double dot[32];
#pragma unroll
      for (int k = 0; k < 32; ++k) dot[k] = 0;
      if (dr_lo <= dr_hi) {
        const NccPatchPool& pool = a.pools.of(s_ref[e]);
        const int32_t* rc = pool.rc + P.sample_off;
        const double* w = pool.w + P.sample_off;
        // tile offsets of window row dr relative to the sample's pixel in
        // window (0, 0); columns dc - 2 = 0..4 as immediate offsets when no
        // column is clamped, else through coff
        int roff[5], coff[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
          roff[k] = min(max(k - 2, dr_lo), dr_hi) * tw;
          coff[k] = min(max(k - 2, dc_lo), dc_hi) + 2;
        }
        const bool full_cols = dc_lo == -2 && dc_hi == 2;  // wave-uniform
        const int base0 = (s_row[e] - r0) * tw + (s_col[e] - c0) - 2;
        // slice [j0, j1) (j0 <= P.n - 1); lane takes j0 + lane + 64 t.  Samples
        // are loaded kRing steps ahead of their use; past the slice a lane
        // re-reads sample j0 with weight 0.
        const int j0 = (P.n * q) / Q, j1 = (P.n * (q + 1)) / Q;
        const int n_it = (j1 - j0 + 63) / 64;  // wave-uniform
        constexpr int kRing = 4;
        int32_t ring_rc[kRing];
        double ring_w[kRing];
        // the ring holds the loads as loaded (a select on a loaded value
        // waits for that load at once); past the slice the weight is zeroed
        // where the sample is used
        auto fetch = [&](int t, int32_t& r, double& wt) __attribute__((always_inline)) {
          const int j = j0 + lane + 64 * t;
          const bool ok = j < j1;
          r = rc[ok ? j : j0];
          wt = w[ok ? j : j0];
        };
        auto accumulate = [&](const int32_t (&pix)[25], double wt) __attribute__((always_inline)) {
#pragma unroll
          for (int k = 0; k < 25; ++k) dot[k] = fma(double(kX16 ? pix[k] : pix[k] - Di), wt, dot[k]);
        };
#pragma unroll
        for (int u = 0; u < kRing; ++u) fetch(u, ring_rc[u], ring_w[u]);
        for (int t0 = 0; t0 < n_it; t0 += kRing) {
#pragma unroll
          for (int u = 0; u < kRing; ++u) {
            if (t0 + u >= n_it) break;
            const int32_t r = ring_rc[u];
            const double wt = j0 + lane + 64 * (t0 + u) < j1 ? ring_w[u] : 0.0;
            fetch(t0 + u + kRing, ring_rc[u], ring_w[u]);
            const int p0 = base0 + sample_row(r) * tw + sample_col(r);
            // the 25 pixels first (all reads in flight), then the products
            int32_t pix[25];
            if (full_cols) {
#pragma unroll
              for (int dr = 0; dr < 5; ++dr) {
                const Tile* rowp = tile + (p0 + roff[dr]);
#pragma unroll
                for (int dc = 0; dc < 5; ++dc) pix[5 * dr + dc] = int32_t(rowp[dc]);
              }
            } else {
#pragma unroll
              for (int dr = 0; dr < 5; ++dr)
#pragma unroll
                for (int dc = 0; dc < 5; ++dc) pix[5 * dr + dc] = int32_t(tile[p0 + roff[dr] + coff[dc]]);
            }
            accumulate(pix, wt);
          }
        }
      }
      const double tot = wave_transpose_sum32(dot);
      const int k = (lane >> 1) & 31;
      if (!(lane & 1) && k < 25) s_dot[task][k] = tot;
    }
    // ---- exact moments (and weight-run dot products): task (patch e, window
    // row dr).  Sums of the tile value v (x or S) and v^2 over each support
    // run for the first inside column, then slid one column at a time; the 15
    // totals (5 sums, 5 sums of squares, 5 weight-run dots) are reduced in one
    // transposed butterfly (the integer sums stay below 2^53, exact in FP64).
    for (int task = wave; any_window && task < n_e * 5; task += kTrackThreads / 64) {
      const int e = task / 5, dri = task - e * 5;
      const NccPatch& P = s_pat[e];
      const int dr_lo = s_lim[e][0], dr_hi = s_lim[e][1], dc_lo = s_lim[e][2], dc_hi = s_lim[e][3];
      const int dr = min(max(dri - 2, dr_lo), dr_hi);
      double v[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) v[k] = 0;
      if (dr_lo <= dr_hi) {
        const NccPatchPool& pool = a.pools.of(s_ref[e]);
        const int base0 = (s_row[e] + dr - r0) * tw + (s_col[e] - c0);
        const int4* runs = reinterpret_cast<const int4*>(pool.runs) + P.run_off;
        int64_t m1[5] = {0, 0, 0, 0, 0}, m2[5] = {0, 0, 0, 0, 0};
        for (int j = lane; j < P.n_runs; j += 64) {
          int4 run;
          if (runs_cached) {
            const int2 c = s_runs[s_run0[e] + j];
            run = make_int4(c.x, int(int16_t(c.y & 0xffff)), c.y >> 16, 0);
          } else {
            run = runs[j];
          }
          const Tile* src = tile + base0 + run.x * tw;
          int64_t s1 = 0, s2 = 0;
          int c = run.y + dc_lo;
          const int ce = run.z + dc_lo;
          for (; c + 7 <= ce; c += 8) {  // eight reads in flight
            int64_t t[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) t[u] = int64_t(src[c + u]);
#pragma unroll
            for (int u = 0; u < 8; ++u) {
              s1 += t[u];
              s2 += t[u] * t[u];
            }
          }
          for (; c <= ce; ++c) {
            const int64_t t = int64_t(src[c]);
            s1 += t;
            s2 += t * t;
          }
#pragma unroll
          for (int k = 0; k < 5; ++k) {
            const int dc = k - 2;
            if (dc > dc_lo && dc <= dc_hi) {  // slide from dc - 1 to dc
              const int64_t in = int64_t(src[run.z + dc]), out = int64_t(src[run.y + dc - 1]);
              s1 += in - out;
              s2 += in * in - out * out;
            }
            m1[k] += s1;
            m2[k] += s2;
          }
        }
#pragma unroll
        for (int k = 0; k < 5; ++k) { v[k] = double(m1[k]); v[5 + k] = double(m2[k]); }
        if (P.use_wruns) {
          const int4* wr = reinterpret_cast<const int4*>(pool.wruns) + P.wrun_off;
          const double* ww = pool.wrun_w + P.wrun_off;
          for (int j = lane; j < P.n_wruns; j += 64) {
            const int4 run = wr[j];
            const Tile* src = tile + base0 + run.x * tw;
            int64_t rs = 0;  // sum of x over the run
            for (int c = run.y + dc_lo; c <= run.z + dc_lo; ++c) rs += int64_t(src[c]) - (kX16 ? 0 : a.D);
            const double wj = ww[j];
#pragma unroll
            for (int k = 0; k < 5; ++k) {
              const int dc = k - 2;
              if (dc > dc_lo && dc <= dc_hi) rs += int64_t(src[run.z + dc]) - int64_t(src[run.y + dc - 1]);
              v[10 + k] += double(rs) * wj;
            }
          }
        }
      }
      const double tot = wave_transpose_sum16(v);
      const int k = (lane >> 2) & 15;
      if ((lane & 3) == 0) {
        if (k < 5) s_m1[e][5 * dri + k] = (long long)tot;
        else if (k < 10) s_m2[e][5 * dri + k - 5] = (long long)tot;
        else if (k < 15 && P.use_wruns) s_dot[e * Q][5 * dri + k - 10] = tot;  // slice 0 only (the dot pass skips such patches)
      }
    }
    __syncthreads();
    // ---- scores, and the first strict maximum in (patch, dr, dc) order
    double best = NCC_MISSING;
    int best_i = 0x7fffffff;
    const double unit = ldexp(1.0, -a.lv.shift);
    for (int k = threadIdx.x; any_window && k < n_e * 25; k += kTrackThreads) {
      const int e = k / 25, wi = k - e * 25;
      const int dr = wi / 5 - 2, dc = wi % 5 - 2;
      if (dr < s_lim[e][0] || dr > s_lim[e][1] || dc < s_lim[e][2] || dc > s_lim[e][3]) continue;
      const NccPatch& P = s_pat[e];
      const int row = s_row[e] + dr, col = s_col[e] + dc;
      if (!window_ok(a, P, row, col)) continue;
      double d = s_dot[e * Q][wi];
      if (!P.use_wruns)
        for (int q = 1; q < Q; ++q) d += s_dot[e * Q + q][wi];
      const int64_t n = P.n, m1 = s_m1[e][wi], m2 = s_m2[e][wi];
      // centred moments t1 = sum(S - D), t2 = sum((S - D)^2)
      const int64_t t1 = kX16 ? m1 : m1 - n * a.D;
      const int64_t t2 = kX16 ? m2 : m2 - 2 * a.D * m1 + n * a.D * a.D;
      const double sc = window_score(a, P, d * unit, t1, t2);
      if (sc > best) { best = sc; best_i = k; }
    }
    block_argmax(best, best_i, red_s, red_i);
    if (threadIdx.x == 0) {
      NccMatch res = cd;
      res.score = NCC_MISSING;
      if (best > NCC_MISSING) {
        const int e = best_i / 25, wi = best_i % 25;
        const NccPatch& P = s_pat[e];
        res.row = a.crop_r + a.step * (s_row[e] + wi / 5 - 2 + P.orow) + a.offset;
        res.col = a.crop_c + a.step * (s_col[e] + wi % 5 - 2 + P.ocol) + a.offset;
        res.angle = P.angle;
        res.score = best;
      }
      keep_distinct(a, f, ci, res);
    }
    __syncthreads();  // the LDS of this candidate is reused by the next
  }
  // the work counter: the block's total, added once (one contended global
  // atomic per candidate would wait at the next barrier)
  if (args.work && threadIdx.x == 0 && s_work) atomicAdd(args.work, s_work);
}

// ----------------------------------------------------------- MFMA tracking
// The 25 windows x E eligible patches of one candidate as one GEMM:
//   D[w][a] = sum over (Y, X) of x(U + Y + dr_w, V + X + dc_w) * W_a(Y, X)
// where (U, V) is the kernel-space origin shared by all patches and W_a is
// patch a's dense weight raster placed at its anchor.  M = 32 window slots
// (25 used), N = 32 patch slots, K = 16 consecutive columns of one row;
// v_mfma_f32_32x32x16_bf16 with x = x_hi + x_lo exact in BF16 (x_hi =
// S - (S mod 4^l) - D, x_lo = S mod 4^l) and w = w_hi + w_lo, three products.
// The four waves split K and their partial tiles are added in a fixed order.
// Each window's approximate score comes with a bound on its error (the
// patch's split and FP32 accumulation terms over the window's exact standard
// deviation); the windows whose upper bound reaches the best lower bound are
// re-scored exactly in FP64, and the first strict maximum among them in the
// reference's (patch, dr, dc) order is the candidate's tracked pose, as in
// k_track.
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8_t;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float floatx16_t;

constexpr int kTrkSlots = 32;
constexpr int kTrkAngCache = 128;
constexpr int kTrkRunCache = 512;

__device__ __forceinline__ uint16_t bf16_exact_bits(int v) { return uint16_t(__float_as_uint(float(v)) >> 16); }
// 16 bytes from a 4-byte aligned global address (one dwordx4 load)
typedef uint32_t u32x4_a4 __attribute__((ext_vector_type(4), aligned(4)));
__device__ __forceinline__ uint4 load_dw4(const uint32_t* p) {
  const u32x4_a4 v = *reinterpret_cast<const u32x4_a4*>(p);
  return make_uint4(v.x, v.y, v.z, v.w);
}

// Exact FP64 score of one window (one wave, samples lane-strided).
template <class Reader>
__device__ __forceinline__ double exact_window_wave(const TrackArgs& a, const PatchRef& p, int row, int col,
                                                    const Reader& rd) {
  const NccPatch& P = p.P;
  const int lane = threadIdx.x & 63;
  int64_t t1 = 0, t2 = 0;
  double dot = 0;
  const bool dense_w = !P.use_wruns;
  constexpr int U = 8;
  for (int base = lane; base < P.n; base += 64 * U) {
    int32_t rcv[U];
    double wv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = base + 64 * u;
      const int j = i < P.n ? i : base;
      rcv[u] = p.rc[j];
      wv[u] = i < P.n && dense_w ? p.w[j] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t x = base + 64 * u < P.n
                            ? int64_t(rd.at(row + sample_row(rcv[u]), col + sample_col(rcv[u]))) - a.D
                            : 0;
      t1 += x;
      t2 += x * x;
      dot = fma(double(x), wv[u], dot);
    }
  }
  if (!dense_w) {
    for (int j = lane; j < P.n_wruns; j += 64) {
      const int4 run = reinterpret_cast<const int4*>(p.wruns)[j];
      int64_t rs = 0;
      for (int c = run.y; c <= run.z; ++c) rs += rd.at(row + run.x, col + c);
      dot += double(rs - int64_t(run.z - run.y + 1) * a.D) * p.wrun_w[j];
    }
  }
  t1 = wave_sum_i64(t1);
  t2 = wave_sum_i64(t2);
  dot = wave_sum(dot);
  return window_score(a, P, ldexp(dot, -a.lv.shift), t1, t2);
}

template <class T, class P2>
__global__ void __launch_bounds__(kTrackThreads, 3) k_track_mfma(const TrackArgs args) {
  extern __shared__ __attribute__((aligned(16))) uint8_t dyn[];
  __shared__ int elig[kMaxEligible];
  __shared__ int wave_cnt[kTrackThreads / 64];
  __shared__ int box[4];
  __shared__ int s_ref[kTrkSlots], s_row[kTrkSlots], s_col[kTrkSlots];
  __shared__ NccPatch s_pat[kTrkSlots];        // the candidate's eligible patches
  __shared__ double s_ang[kTrkAngCache];       // angle of every query patch of the level
  __shared__ int32_t s_refs[kTrkAngCache];     // and its (pool << 24 | index)
  __shared__ int2 s_runs[kTrkRunCache];        // runs of the eligible patches, slot after slot:
                                               // (row, begin | end << 16)
  __shared__ int s_run0[kTrkSlots + 1];        // first run of each slot in s_runs
  __shared__ int s_dense;
  __shared__ double red_s[kTrackThreads / 64];
  __shared__ double red_u[kTrackThreads / 64];
  __shared__ int red_i[kTrackThreads / 64];
  __shared__ int off[NCC_MAX_BATCH + 1];
  __shared__ unsigned long long s_work;  // 25 * samples of the block's candidates
  if (threadIdx.x == 0) s_work = 0;
  uint16_t* sS = reinterpret_cast<uint16_t*>(dyn + args.mt_s);
  uint32_t* P1 = reinterpret_cast<uint32_t*>(dyn + args.mt_p1);
  P2* Q2 = reinterpret_cast<P2*>(dyn + args.mt_p2);
  uint16_t* th0 = reinterpret_cast<uint16_t*>(dyn + args.mt_pl);
  const int PP = args.mt_pitch;
  uint16_t* tl0 = th0 + args.mt_h * PP;
  // after the MFMA loop the planes' space holds two partial tiles [2][32][32],
  // then the summed tile [32][32] and the windows' upper bounds; the list of
  // windows to re-score replaces the summed tile once the bounds are known
  float* part = reinterpret_cast<float*>(dyn + args.mt_pl);
  float* s_up = part + 1024;
  int* s_list = reinterpret_cast<int*>(part);
  list_offsets(args.in_count, args.n_frames, args.cap, off);
  const bool ang_cached = args.n_refs <= kTrkAngCache;
  if (ang_cached)
    for (int k = threadIdx.x; k < args.n_refs; k += kTrackThreads) {
      const int32_t r = args.refs[k];
      s_refs[k] = r;
      s_ang[k] = ref_angle(args.pools, r);
    }
  __syncthreads();
  const int total = off[args.n_frames];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  for (int g = blockIdx.x; g < total; g += gridDim.x) {
    TrackArgs a = args;
    const int f = __builtin_amdgcn_readfirstlane(list_frame(off, a.n_frames, g));
    const int ci = __builtin_amdgcn_readfirstlane(g - off[f]);
    bind_frame(a, f);
    const NccMatch cd = a.in[size_t(f) * a.cap + ci];
    const int n_e = ang_cached ? eligible_cached(s_ang, a.n_refs, cd.angle, a.tol, elig, wave_cnt)
                               : eligible_patches(a.pools, a.refs, a.n_refs, cd.angle, a.tol, elig, wave_cnt);
    if (threadIdx.x == 0) { box[0] = INT_MAX; box[1] = INT_MIN; box[2] = INT_MAX; box[3] = INT_MIN; s_dense = 1; }
    __syncthreads();
    // one round of descriptor loads: each thread takes one eligible patch
    unsigned long long samples = 0;
    for (int e = threadIdx.x; e < n_e; e += kTrackThreads) {
      const int32_t ref = ang_cached ? s_refs[elig[e]] : a.refs[elig[e]];
      const NccPatch P = load_patch(a.pools, ref);
      samples += 25ull * uint32_t(P.n);
      int row, col;
      anchor(a, cd, P, row, col);
      if (e < kTrkSlots) { s_ref[e] = ref; s_row[e] = row; s_col[e] = col; s_pat[e] = P; }
      if (P.dense_off < 0) s_dense = 0;
      atomicMin(&box[0], row - 2 + P.rmin);
      atomicMax(&box[1], row + 2 + P.rmax);
      atomicMin(&box[2], col - 2 + P.cmin);
      atomicMax(&box[3], col + 2 + P.cmax);
    }
    if (a.work && samples) atomicAdd(&s_work, samples);
    __syncthreads();
    if (threadIdx.x == 0) {
      int t = 0;
      for (int e = 0; e < min(n_e, kTrkSlots); ++e) { s_run0[e] = t; t += s_pat[e].n_runs; }
      s_run0[min(n_e, kTrkSlots)] = t;
    }
    const int U0 = box[0], V0 = box[2];
    const int th = box[1] - U0 + 1, tw = box[3] - V0 + 1;
    const int KH = th - 4, XC = (tw - 4 + 15) / 16;
    const bool fast = n_e > 0 && n_e <= kTrkSlots && s_dense && th <= a.mt_h && tw <= a.mt_w &&
                      16 * XC + 5 <= PP && KH > 0;
    double best = NCC_MISSING;
    int best_i = 0x7fffffff;
    bool n_e_skip = false;  // block-uniform
    const T* lvp = static_cast<const T*>(a.lv.data);
    if (!fast) {
      // every window exactly, from the level in global memory
      const LevelReader<T> rd{lvp, a.lv.pitch};
      for (int k = wave; k < n_e * 25; k += kTrackThreads / 64) {
        const int e = k / 25, wi = k % 25;
        const int32_t ref = __builtin_amdgcn_readfirstlane(a.refs[elig[e]]);
        const PatchRef p = patch_ref(a.pools, ref);
        int row, col;
        anchor(a, cd, p.P, row, col);
        row += wi / 5 - 2;
        col += wi % 5 - 2;
        if (!window_ok(a, p.P, row, col)) continue;
        const double sc = exact_window_wave(a, p, row, col, rd);
        if (sc > best) { best = sc; best_i = k; }
      }
    } else {
      // ---- stage S, the split planes (zero past the tile) and the row prefixes
      const int shift = a.lv.shift, lo_mask = (1 << shift) - 1;
      const int Di = int(a.D);
      const int vr0 = max(0, a.sr), vc0 = max(0, a.sc);
      const int vr1 = min(a.lv.height, a.sr + a.sh), vc1 = min(a.lv.width, a.sc + a.sw);
      __syncthreads();  // s_run0
      const bool runs_cached = s_run0[n_e] <= kTrkRunCache;
      if (runs_cached)  // every eligible patch's runs in one round of loads
        for (int j = threadIdx.x; j < s_run0[n_e]; j += kTrackThreads) {
          int e = 0;
          while (s_run0[e + 1] <= j) ++e;
          const int4 r = reinterpret_cast<const int4*>(a.pools.of(s_ref[e]).runs)[s_pat[e].run_off + j - s_run0[e]];
          s_runs[j] = make_int2(r.x, (r.y & 0xffff) | (r.z << 16));
        }
      {
        // element i of the tile is (i / PP, i % PP); the quotient comes from
        // a float reciprocal (exact here: i < 2^16 and PP < 2^10, so
        // (i + 0.5) / PP stays more than 2^-11 away from an integer)
        constexpr int U = 16;  // loads in flight per thread
        const float inv_pp = 1.f / float(PP);
        for (int base = threadIdx.x; base < th * PP; base += kTrackThreads * U) {
          int Sv[U];
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const int i = base + kTrackThreads * u;
            const int tr = int((float(i) + .5f) * inv_pp), tc = i - tr * PP;
            const int r = U0 + tr, c = V0 + tc;
            Sv[u] = i < th * PP && tc < tw && r >= vr0 && r < vr1 && c >= vc0 && c < vc1
                        ? int(lvp[size_t(r) * a.lv.pitch + c])
                        : Di;
          }
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const int i = base + kTrackThreads * u;
            if (i >= th * PP) break;
            const int tr = int((float(i) + .5f) * inv_pp), tc = i - tr * PP;
            const int S = Sv[u];
            if (tc < tw) sS[tr * tw + tc] = uint16_t(S);
            const int lo = S & lo_mask;
            th0[tr * PP + tc] = bf16_exact_bits(S - lo - Di);
            tl0[tr * PP + tc] = bf16_exact_bits(lo);
          }
        }
      }
      __syncthreads();
      // row prefixes of S and S^2: one thread per row, loads in groups of 8
      for (int tr = threadIdx.x; tr < th; tr += kTrackThreads) {
        uint32_t s1 = 0;
        P2 s2 = 0;
        uint32_t* p1 = P1 + tr * (tw + 1);
        P2* p2 = Q2 + tr * (tw + 1);
        const uint16_t* src = sS + tr * tw;
        p1[0] = 0;
        p2[0] = 0;
        for (int c0 = 0; c0 < tw; c0 += 8) {
          uint32_t v[8];
#pragma unroll
          for (int j = 0; j < 8; ++j) v[j] = c0 + j < tw ? src[c0 + j] : 0u;
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            if (c0 + j >= tw) break;
            s1 += v[j];
            s2 += P2(v[j]) * P2(v[j]);
            p1[c0 + j + 1] = s1;
            p2[c0 + j + 1] = s2;
          }
        }
      }
      // ---- MFMA: lane (m, gg) = window slot m / patch slot m, K half gg
      const int m = lane & 31, gg = lane >> 5;
      const int wsl = min(m, 24);
      const int arow = wsl / 5, acol = 8 * gg + wsl % 5;
      const bool has = m < n_e;
      floatx16_t acc;
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[v] = 0.f;
      // this wave's band of K rows; iteration t covers (Y, cx) = (y_lo + t / XC, t % XC)
      const int y_lo = (KH * wave) / 4, y_hi = (KH * (wave + 1)) / 4;
      const int n_my = (y_hi - y_lo) * XC;
      constexpr int kRing = 4;  // weight fragments in flight per lane
      // B fragment of step (Y, cx): elements e .. e + 7 of the patch's dense
      // raster, e = (Y + ybase) dw + 16 cx + xbase.  dw and 16 cx are even,
      // so the copy (e mod 2) is the lane's constant and the 4-byte aligned
      // offset moves by dw per row and 16 per step: the dword index is
      // qrow + 8 cx, qrow advanced by dw / 2 per row.  The fragment is
      // inside the raster for Y in [ylo, yhi) and cx in [cxl, cxh] (lane
      // constants; empty for a lane without a patch).  Loads are issued
      // unconditionally and used as loaded (a mask applied to a loaded value
      // waits for that load at once): a fragment outside the raster, and every
      // fragment of a lane without a patch (slot 0's raster), is read at index
      // 0, row 0's eight zero padding columns.
      const int ms = has ? m : 0;
      const NccPatch& Pm = s_pat[ms];
      const NccPatchPool& pool_m = a.pools.of(s_ref[ms]);
      const uint32_t* bh_q = reinterpret_cast<const uint32_t*>(pool_m.dense_hi + Pm.dense_off);
      const uint32_t* bl_q = reinterpret_cast<const uint32_t*>(pool_m.dense_lo + Pm.dense_off);
      const int dh = Pm.dense_h, dw = Pm.dense_w;
      const int ybase = U0 + 2 - s_row[ms] - Pm.rmin;
      const int xbase = V0 + 2 - s_col[ms] - Pm.cmin + 8 + 8 * gg;
      const int ylo = has ? -ybase : 1, yhi = has ? dh - ybase : 0;
      // 16 cx + xbase >= 0 and 16 cx + xbase + 8 <= dw: cx >= ceil(-xbase / 16)
      // = -floor(xbase / 16) and cx <= floor((dw - 8 - xbase) / 16) (floor
      // division of possibly negative numerators: an arithmetic shift)
      const int cxl = -(xbase >> 4), cxh = (dw - 8 - xbase) >> 4;
      const int qcopy = ((xbase & 1) * dh * dw + (xbase & ~1) + ybase * dw) >> 1;  // qrow at Y = 0
      const int qdw = dw >> 1;
      auto load_b = [&](bool ok, int Y, int cx, uint4& bh4, uint4& bl4) __attribute__((always_inline)) {
        const bool in = ok && Y >= ylo && Y < yhi && cx >= cxl && cx <= cxh;
        const int q = in ? qcopy + Y * qdw + 8 * cx : 0;
        bh4 = load_dw4(bh_q + q);
        bl4 = load_dw4(bl_q + q);
      };
      // A fragment of step (Y, cx): 8 elements from tile row min(Y, KH - 1) +
      // arow, column 16 cx + acol: five aligned dwords, shifted by one element
      // (alignbyte) for the lanes whose column is odd (acol odd: a lane
      // constant).  A prefetch past the band reads a valid row, unused.
      const int sh = (acol & 1) * 2;
      const uint32_t* ah0 = reinterpret_cast<const uint32_t*>(th0 + arow * PP + (acol & ~1));
      const uint32_t* al0 = reinterpret_cast<const uint32_t*>(tl0 + arow * PP + (acol & ~1));
      auto load_a = [&](int Y, int cx, uint4& ah4, uint4& al4) __attribute__((always_inline)) {
        const int off = (min(Y, KH - 1) * PP + 16 * cx) >> 1;  // dwords (PP and 16 cx even), wave-uniform
        const uint32_t* ah = ah0 + off;
        const uint32_t* al = al0 + off;
        uint32_t h[5], l[5];
#pragma unroll
        for (int j = 0; j < 5; ++j) { h[j] = ah[j]; l[j] = al[j]; }
        ah4 = make_uint4(__builtin_amdgcn_alignbyte(h[1], h[0], sh), __builtin_amdgcn_alignbyte(h[2], h[1], sh),
                         __builtin_amdgcn_alignbyte(h[3], h[2], sh), __builtin_amdgcn_alignbyte(h[4], h[3], sh));
        al4 = make_uint4(__builtin_amdgcn_alignbyte(l[1], l[0], sh), __builtin_amdgcn_alignbyte(l[2], l[1], sh),
                         __builtin_amdgcn_alignbyte(l[3], l[2], sh), __builtin_amdgcn_alignbyte(l[4], l[3], sh));
      };
      auto step = [&](int& Y, int& cx) __attribute__((always_inline)) {
        if (++cx == XC) { cx = 0; ++Y; }
      };
      uint4 rh[kRing], rl[kRing];
      int pY = y_lo, pcx = 0;  // prefetch cursor
#pragma unroll
      for (int j = 0; j < kRing; ++j) {
        load_b(j < n_my, pY, pcx, rh[j], rl[j]);
        step(pY, pcx);
      }
      int Y = y_lo, cx = 0;
      uint4 a_h, a_l;
      load_a(Y, cx, a_h, a_l);
      // whole groups of kRing steps (no exit inside the unrolled body, so the
      // loads stay in flight across it); the steps past n_my multiply zero
      // weight fragments and add nothing
      const int n_pad = (n_my + kRing - 1) / kRing * kRing;
      for (int t0 = 0; t0 < n_pad; t0 += kRing) {
#pragma unroll
        for (int j = 0; j < kRing; ++j) {
          const int t = t0 + j;
          int Yn = Y, cxn = cx;
          step(Yn, cxn);
          uint4 n_h, n_l;
          load_a(Yn, cxn, n_h, n_l);
          const bf16x8_t A_h = __builtin_bit_cast(bf16x8_t, a_h), A_l = __builtin_bit_cast(bf16x8_t, a_l);
          const bf16x8_t B_h = __builtin_bit_cast(bf16x8_t, rh[j]), B_l = __builtin_bit_cast(bf16x8_t, rl[j]);
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A_h, B_h, acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A_h, B_l, acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A_l, B_h, acc, 0, 0, 0);
          // the ring slot's next use, loaded once its fragment is consumed, so
          // that it keeps its registers across the loop (no copies at the
          // back-edge, which would wait for every load in flight)
          load_b(t + kRing < n_my, pY, pcx, rh[j], rl[j]);
          step(pY, pcx);
          a_h = n_h;
          a_l = n_l;
          Y = Yn;
          cx = cxn;
        }
      }
      __syncthreads();  // the planes are dead: their space takes the partial tiles
      // C layout: lane (column m, half gg), register v: row 8 (v / 4) + 4 gg + v % 4.
      // The four partial tiles add as (w0 + w1) + (w2 + w3).
      auto cidx = [&](int v) __attribute__((always_inline)) { return (8 * (v / 4) + 4 * gg + (v % 4)) * 32 + m; };
      if (wave & 1) {
#pragma unroll
        for (int v = 0; v < 16; ++v) part[(wave >> 1) * 1024 + cidx(v)] = acc[v];
      }
      __syncthreads();
      if (!(wave & 1)) {
#pragma unroll
        for (int v = 0; v < 16; ++v) acc[v] += part[(wave >> 1) * 1024 + cidx(v)];
      }
      __syncthreads();
      if (wave == 2) {
#pragma unroll
        for (int v = 0; v < 16; ++v) part[cidx(v)] = acc[v];
      }
      __syncthreads();
      if (wave == 0) {
#pragma unroll
        for (int v = 0; v < 16; ++v) part[cidx(v)] = acc[v] + part[cidx(v)];
      }
      __syncthreads();
      // ---- approximate scores with error bounds; the best lower bound
      const double unit = ldexp(1.0, -a.lv.shift);
      const double err_scale = unit;
      double lower_max = NCC_MISSING;
      for (int k = threadIdx.x; k < n_e * 25; k += kTrackThreads) {
        const int e = k / 25, wi = k - e * 25;
        const NccPatch& P = s_pat[e];
        const int row = s_row[e] + wi / 5 - 2, col = s_col[e] + wi % 5 - 2;
        float up = -__builtin_inff();
        if (window_ok(a, P, row, col)) {
          const int4* runs = reinterpret_cast<const int4*>(a.pools.of(s_ref[e]).runs) + P.run_off;
          const int2* cruns = s_runs + s_run0[e];
          int64_t m1 = 0, m2 = 0;
#pragma unroll 8
          for (int j = 0; j < P.n_runs; ++j) {
            int4 run;
            if (runs_cached) {
              const int2 c = cruns[j];
              run = make_int4(c.x, int(int16_t(c.y & 0xffff)), c.y >> 16, 0);
            } else {
              run = runs[j];
            }
            const int tr = row + run.x - U0;
            const int c0 = col + run.y - V0, c1 = col + run.z + 1 - V0;
            m1 += int64_t(P1[tr * (tw + 1) + c1]) - int64_t(P1[tr * (tw + 1) + c0]);
            m2 += int64_t(Q2[tr * (tw + 1) + c1]) - int64_t(Q2[tr * (tw + 1) + c0]);
          }
          const int64_t n = P.n;
          const int64_t t1 = m1 - n * a.D;
          const int64_t t2 = m2 - 2 * a.D * m1 + n * a.D * a.D;
          const double sum = ldexp(double(t1), -a.lv.shift), sq = ldexp(double(t2), -2 * a.lv.shift);
          const double var = sq - sum * sum / double(P.n);
          if (!(var > 128 * __DBL_EPSILON__ * sq)) {
            // low variance: -inf on an exact lattice, else only the exact form decides
            up = a.exact_lattice ? -__builtin_inff() : __builtin_inff();
          } else {
            const double dot = double(part[wi * 32 + e]) * unit;
            const double sc = window_score(a, P, dot, t1, t2);
            const double errd = (double(lo_mask) * P.err_lo + a.x_max * P.err_x) * err_scale;
            const double beta = errd / sqrt(var) * (1 + 1e-6) + 1e-12;
            lower_max = fmax(lower_max, sc - beta);
            up = __double2float_ru(sc + beta);
          }
        }
        s_up[k] = up;
      }
      // one block reduction for both maxima: the best lower bound L and the
      // best upper bound (a candidate whose windows all stay below the keep
      // threshold is dropped whichever window is the maximum: no exact
      // scoring needed)
      double um = NCC_MISSING;
      for (int k = threadIdx.x; k < n_e * 25; k += kTrackThreads) um = fmax(um, double(s_up[k]));
#pragma unroll
      for (int d = 32; d > 0; d >>= 1) {
        lower_max = fmax(lower_max, __shfl_xor(lower_max, d, 64));
        um = fmax(um, __shfl_xor(um, d, 64));
      }
      if (lane == 0) { red_s[wave] = lower_max; red_u[wave] = um; }
      __syncthreads();
      lower_max = red_s[0];
      um = red_u[0];
      for (int w = 1; w < kTrackThreads / 64; ++w) { lower_max = fmax(lower_max, red_s[w]); um = fmax(um, red_u[w]); }
      __syncthreads();
      const double L = lower_max;
      if (um < a.keep) n_e_skip = true;
      // ---- the windows that may hold the maximum, in (patch, dr, dc) order
      int n_list = 0;
      for (int base = 0; !n_e_skip && base < n_e * 25; base += kTrackThreads) {
        const int k = base + threadIdx.x;
        const bool take = k < n_e * 25 && s_up[k] > -__builtin_inff() && double(s_up[k]) >= L;
        const uint64_t mk = __ballot(take);
        const int pre = __popcll(mk & ((uint64_t(1) << lane) - 1));
        if (lane == 0) wave_cnt[wave] = __popcll(mk);
        __syncthreads();
        int o = n_list;
        for (int w = 0; w < wave; ++w) o += wave_cnt[w];
        if (take) s_list[o + pre] = k;
        for (int w = 0; w < kTrackThreads / 64; ++w) n_list += wave_cnt[w];
        __syncthreads();
      }
      const TileReader<uint16_t> rd{sS, U0, V0, tw};
      for (int j = wave; j < n_list; j += kTrackThreads / 64) {
        const int k = __builtin_amdgcn_readfirstlane(s_list[j]);
        const int e = k / 25, wi = k - e * 25;
        const int32_t ref = __builtin_amdgcn_readfirstlane(s_ref[e]);
        const NccPatchPool& pool = a.pools.of(ref);
        PatchRef p;
        p.P = s_pat[e];
        p.rc = pool.rc + p.P.sample_off;
        p.w = pool.w + p.P.sample_off;
        p.wruns = pool.wruns + 4 * p.P.wrun_off;
        p.wrun_w = pool.wrun_w + p.P.wrun_off;
        const int row = s_row[e] + wi / 5 - 2, col = s_col[e] + wi % 5 - 2;
        const double sc = exact_window_wave(a, p, row, col, rd);
        if (sc > best) { best = sc; best_i = k; }
      }
    }
    block_argmax(best, best_i, red_s, red_i);
    if (threadIdx.x == 0) {
      NccMatch res = cd;
      res.score = NCC_MISSING;
      if (best > NCC_MISSING) {
        const int e = best_i / 25;
        const NccPatch P = e < kTrkSlots ? s_pat[e] : a.pools.patch(a.refs[elig[e]]);
        int row, col;
        anchor(a, cd, P, row, col);
        res.row = a.crop_r + a.step * (row + (best_i % 25) / 5 - 2 + P.orow) + a.offset;
        res.col = a.crop_c + a.step * (col + best_i % 5 - 2 + P.ocol) + a.offset;
        res.angle = P.angle;
        res.score = best;
      }
      keep_distinct(a, f, ci, res);
    }
    __syncthreads();  // LDS is reused by the next candidate
  }
  // the work counter: the block's total, added once (one contended global
  // atomic per candidate would wait at the next barrier)
  if (args.work && threadIdx.x == 0 && s_work) atomicAdd(args.work, s_work);
}

// ------------------------------------------------------- continuous / refine
struct RefineArgs {
  const NccFrame* frames;     // per-frame descriptors (device)
  const int32_t* cand_frame;  // frame index of every candidate
  double scale, phase;        // 2^finest, (scale - 1) / 2
  int finest, top;
  Pools pools;
  int32_t orig_ref;       // unrotated patch of the finest level (model.original at level 0)
  double lv_orow, lv_ocol;
  const int32_t* refs0;   // level-0 query patches (recovery search)
  int n_refs0;
  double step, a_start, a_end, min_score;
  int subpixel, rot, absolute;
  int recover_first, prefix_recovery;
  double guard_first, guard;
  const NccMatch* cand;
  int n_cand;
  // list form (full score): frame f's candidates at cand + f * cap
  int n_frames, cap;
  const int32_t* in_count;
  NccMatch* out_full;
  NccMatch* out_refined;
  int32_t* verdict;
  // the candidate's frame, filled in-kernel
  NccLevel lv;            // finest search level
  double dc, crop_r, crop_c;
  int sr, sc, sh, sw;     // storage rectangle at the finest level
  int img_h, img_w;       // full image
  const NccFrame* dom;    // the frame when it has a reduced domain, else null
  int H0, W0;             // level-0 dims (crop)
  // LDS copy of the candidate's neighbourhood (storage-local rectangle)
  const void* tile;
  int tr0, tc0, th, tw;
  // LDS room after the tile for a sampling pass's rotation tables, in
  // 16-byte entries (0: none)
  int rtab_off, rtab_cap;
};

__device__ __forceinline__ void bind_frame(RefineArgs& a, int f) {
  const NccFrame& F = a.frames[f];
  a.lv = F.levels[a.finest];
  a.dc = double(frame_dc(F));
  a.crop_r = F.crop_r;
  a.crop_c = F.crop_c;
  if (a.finest == 0) { a.sr = F.storage0[0]; a.sc = F.storage0[1]; a.sh = F.storage0[2]; a.sw = F.storage0[3]; }
  else { a.sr = 0; a.sc = 0; a.sh = a.lv.height; a.sw = a.lv.width; }
  a.img_h = F.image_h;
  a.img_w = F.image_w;
  a.dom = F.dom_sat ? &F : nullptr;
  a.H0 = F.levels[0].height;
  a.W0 = F.levels[0].width;
}

constexpr int kRefineThreads = 512;
constexpr int kRefineTileBytes = 64 * 1024;

struct Shared {
  double red[(kRefineThreads / 64) * 18];
  double tot[18];
  int flag;
  double bcast[8];
  double win_s[kRefineThreads / 64];
  int win_i[kRefineThreads / 64];
};

__device__ __forceinline__ bool domain_contains(const RefineArgs& a, double row, double col) {
  if (a.dom) return in_domain(*a.dom, row, col);
  const double r = floor(row + .5), c = floor(col + .5);
  return r >= 0 && c >= 0 && r < double(a.img_h) && c < double(a.img_w);
}

// Integer pixel at storage-local coordinates, from the LDS tile when the
// point lies inside it.
template <class T>
__device__ __forceinline__ uint32_t pix_s(const RefineArgs& a, int r, int c) {
  const int y = r - a.tr0, x = c - a.tc0;
  if (unsigned(y) < unsigned(a.th) && unsigned(x) < unsigned(a.tw)) return static_cast<const T*>(a.tile)[y * a.tw + x];
  return static_cast<const T*>(a.lv.data)[size_t(r + a.sr) * a.lv.pitch + (c + a.sc)];
}
// Value at storage-local integer coordinates.
template <class T>
__device__ __forceinline__ double pix(const RefineArgs& a, int r, int c) {
  return ldexp(double(pix_s<T>(a, r, c)), -a.lv.shift) - a.dc;
}

// Stages the storage-local neighbourhood of radius `reach` around (pr, pc)
// into LDS (u8 or u16 levels); records an empty tile when it does not fit.
template <class T>
__device__ __forceinline__ void stage_tile(RefineArgs& a, void* lds, double pr, double pc, double reach) {
  a.tile = lds;
  a.th = 0; a.tw = 0; a.tr0 = 0; a.tc0 = 0;
  a.rtab_off = 0; a.rtab_cap = kRefineTileBytes / 16;  // no tile: the whole area holds the tables
  if (sizeof(T) > 2 || !isfinite(pr) || !isfinite(pc)) return;
  const int r0 = max(0, int(floor(pr - reach))), r1 = min(a.sh - 1, int(ceil(pr + reach)));
  const int c0 = max(0, int(floor(pc - reach))), c1 = min(a.sw - 1, int(ceil(pc + reach)));
  const int h = r1 - r0 + 1, w = c1 - c0 + 1;
  if (h <= 0 || w <= 0 || h * w * int(sizeof(T)) > kRefineTileBytes) return;
  a.rtab_off = (h * w * int(sizeof(T)) + 15) & ~15;
  a.rtab_cap = (kRefineTileBytes - a.rtab_off) / 16;
  const T* src = static_cast<const T*>(a.lv.data);
  constexpr int U = 8;
  for (int base = threadIdx.x; base < h * w; base += blockDim.x * U) {
    uint32_t v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = base + int(blockDim.x) * u;
      v[u] = i < h * w ? uint32_t(src[size_t(r0 + i / w + a.sr) * a.lv.pitch + (c0 + i % w + a.sc)]) : 0u;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = base + int(blockDim.x) * u;
      if (i < h * w) static_cast<T*>(lds)[i] = T(v[u]);
    }
  }
  __syncthreads();
  a.tr0 = r0; a.tc0 = c0; a.th = h; a.tw = w;
}

// strictly_inside: kernels_avx2.cpp:261-277
__device__ __forceinline__ bool strictly_inside(const RefineArgs& a, const NccPatch& P, double pr, double pc, double ca, double sa) {
  if (P.rmin == P.rmax && P.cmin == P.cmax) return false;
  const double dr[2] = {P.rmin - a.lv_orow, P.rmax - a.lv_orow};
  const double dc[2] = {P.cmin - a.lv_ocol, P.cmax - a.lv_ocol};
  double ymin = __builtin_inf(), xmin = ymin, ymax = -ymin, xmax = -ymin;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) {
      const double y = pr + ca * dr[i] - sa * dc[j], x = pc + sa * dr[i] + ca * dc[j];
      ymin = fmin(ymin, y); ymax = fmax(ymax, y); xmin = fmin(xmin, x); xmax = fmax(xmax, x);
    }
  const double margin = 64 * __DBL_EPSILON__ *
      (1 + fabs(pr) + fabs(pc) + fabs(dr[0]) + fabs(dr[1]) + fabs(dc[0]) + fabs(dc[1]) + double(a.sw) + double(a.sh));
  return ymin > margin && xmin > margin && ymax < double(a.sh - 1) - margin && xmax < double(a.sw - 1) - margin;
}

// Block-wide sums of the 18 refinement moments, for thread 0 only (the
// Gauss-Newton solve runs there): the first 16 with one transposed butterfly
// per wave (17 exchanges instead of 96), the last two with plain wave sums;
// then thread t < 18 adds value t over the waves (in wave order) and thread
// 0 reads the 18 totals.  Other threads' v is left unspecified.
__device__ __forceinline__ void block_sum18(double (&v)[18], double* scratch, double* tot) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, n_waves = blockDim.x >> 6;
  double head[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) head[i] = v[i];
  const double t = wave_transpose_sum16(head);  // lane L: total of value L >> 2
  const double t16 = wave_sum(v[16]), t17 = wave_sum(v[17]);
  if ((lane & 3) == 0) scratch[wave * 18 + (lane >> 2)] = t;
  if (lane == 0) { scratch[wave * 18 + 16] = t16; scratch[wave * 18 + 17] = t17; }
  __syncthreads();
  if (threadIdx.x < 18) {
    double a = 0;
    for (int w = 0; w < n_waves; ++w) a += scratch[w * 18 + threadIdx.x];
    tot[threadIdx.x] = a;
  }
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int i = 0; i < 18; ++i) v[i] = tot[i];
}

// Bilinear sample (and gradients) at storage-local (y, x) known to be inside.
template <class T>
__device__ __forceinline__ void bilinear_at(const RefineArgs& a, double y, double x, bool interior, double& q,
                                            double& gy, double& gx) {
  const double yr = floor(y), xc = floor(x), fy = y - yr, fx = x - xc;
  const int r = int(yr), c = int(xc);
  const int r1 = interior ? r + 1 : min(r + 1, a.sh - 1), c1 = interior ? c + 1 : min(c + 1, a.sw - 1);
  const double va = pix<T>(a, r, c), vb = pix<T>(a, r, c1), vd = pix<T>(a, r1, c), ve = pix<T>(a, r1, c1);
  const double ux = 1 - fx, uy = 1 - fy;
  q = uy * (ux * va + fx * vb) + fy * (ux * vd + fx * ve);
  gy = ux * (vd - va) + fx * (ve - vb);
  gx = uy * (vb - va) + fy * (ve - vd);
}

__device__ __forceinline__ bool inside(const RefineArgs& a, double y, double x, bool deriv) {
  if (!(y >= 0 && x >= 0)) return false;
  return deriv ? (y < double(a.sh - 1) && x < double(a.sw - 1)) : (y <= double(a.sh - 1) && x <= double(a.sw - 1));
}

// Block-wide bilinear moments of the unrotated patch at a storage-local pose.
// out: [0] sum q, [1] sum q^2, [2] sum w q, then gs[3], qg[3], wg[3], and
// gg00, gg10, gg11, gg20, gg21, gg22 (with deriv, on thread 0 only).  Returns
// false if any sample is outside.
template <class T>
__device__ __forceinline__ bool bilinear_pass(const RefineArgs& a, const PatchRef& O, double pr, double pc, double angle, bool deriv,
                              double* out, Shared& sm) {
  const double ca = cos(angle), sa = sin(angle);
  const NccPatch& P = O.P;
  const bool interior = strictly_inside(a, P, pr, pc, ca, sa);
  const bool aligned = angle == 0;
  double off;
  {
    const int32_t rc = O.rc[0];
    const double dr = sample_row(rc) - a.lv_orow, dc = sample_col(rc) - a.lv_ocol;
    const double y = pr + ca * dr - sa * dc, x = pc + sa * dr + ca * dc;
    if (!inside(a, y, x, deriv)) return false;
    double gy, gx;
    bilinear_at<T>(a, y, x, false, off, gy, gx);
  }
  double v[18];
#pragma unroll
  for (int k = 0; k < 18; ++k) v[k] = 0;
  int bad = 0;
  constexpr int U = 8;  // samples per thread whose template loads are issued together
  // Fast form of the loop below, for a pass at level 0 whose samples are all
  // strictly inside: the products ca*dr, sa*dr (per template row) and sa*dc,
  // ca*dc (per column) come from tables built in LDS for the pass, and the
  // four bilinear taps are read without bounds checks, from the LDS tile when
  // the pass's footprint lies in it, else from the level.  Every FP64
  // operation is the loop's own, in the same order (S - dc is an exact
  // integer), so the moments are bit-identical to the loop's.
  const int rows = P.rmax - P.rmin + 1, cols = P.cmax - P.cmin + 1;
  int mode = 0;  // 0: the loop below; 1: fast, taps from the LDS tile; 2: fast, taps from the level
  if (interior && a.lv.shift == 0 && rows + cols <= a.rtab_cap) {
    const double dr0 = P.rmin - a.lv_orow, dr1 = P.rmax - a.lv_orow;
    const double dc0 = P.cmin - a.lv_ocol, dc1 = P.cmax - a.lv_ocol;
    double ylo = __builtin_inf(), xlo = ylo, yhi = -ylo, xhi = -ylo;
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) {
        const double dr = i ? dr1 : dr0, dc = j ? dc1 : dc0;
        const double y = pr + ca * dr - sa * dc, x = pc + sa * dr + ca * dc;
        ylo = fmin(ylo, y); yhi = fmax(yhi, y); xlo = fmin(xlo, x); xhi = fmax(xhi, x);
      }
    // a pixel of margin each side for the rounding of single samples
    const bool in_tile = a.th > 0 && int(floor(ylo)) - 1 >= a.tr0 && int(floor(yhi)) + 2 <= a.tr0 + a.th - 1 &&
                         int(floor(xlo)) - 1 >= a.tc0 && int(floor(xhi)) + 2 <= a.tc0 + a.tw - 1;
    mode = in_tile ? 1 : 2;
  }
  if (mode) {
    extern __shared__ uint8_t lds_tile[];
    double2* const rt = reinterpret_cast<double2*>(lds_tile + a.rtab_off);  // [row - rmin]: (ca dr, sa dr)
    double2* const ct = rt + rows;                                          // [col - cmin]: (sa dc, ca dc)
    for (int i = threadIdx.x; i < rows + cols; i += kRefineThreads) {
      if (i < rows) {
        const double dr = (P.rmin + i) - a.lv_orow;
        rt[i] = make_double2(ca * dr, sa * dr);
      } else {
        const double dc = (P.cmin + i - rows) - a.lv_ocol;
        rt[i] = make_double2(sa * dc, ca * dc);
      }
    }
    __syncthreads();
    const int idc = int(a.dc);
    auto fast = [&]<bool kLds>() __attribute__((always_inline)) {
      const T* lt = reinterpret_cast<const T*>(lds_tile);
      const T* gl = static_cast<const T*>(a.lv.data) + (size_t(a.sr) * a.lv.pitch + a.sc);
      const int stride = kLds ? a.tw : int(a.lv.pitch);
      const int org = kLds ? -a.tr0 * a.tw - a.tc0 : 0;
      for (int base = threadIdx.x; base < P.n; base += kRefineThreads * U) {
        int32_t rcv[U];
        double wv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int i = base + kRefineThreads * u;
          rcv[u] = i < P.n ? O.rc[i] : 0;
          wv[u] = i < P.n ? O.w[i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          if (base + kRefineThreads * u >= P.n) continue;
          const double2 tr = rt[sample_row(rcv[u]) - P.rmin], tc = ct[sample_col(rcv[u]) - P.cmin];
          const double y = (pr + tr.x) - tc.x, x = (pc + tr.y) + tc.y;
          const double yr = floor(y), xc = floor(x), fy = y - yr, fx = x - xc;
          const int r = int(yr), c = int(xc);
          double va, vb, vd, ve;
          if constexpr (kLds) {
            const int i0 = r * stride + c + org;
            va = double(int(lt[i0]) - idc); vb = double(int(lt[i0 + 1]) - idc);
            vd = double(int(lt[i0 + stride]) - idc); ve = double(int(lt[i0 + stride + 1]) - idc);
          } else {
            const T* g = gl + (size_t(r) * size_t(stride) + c);
            va = double(int(g[0]) - idc); vb = double(int(g[1]) - idc);
            vd = double(int(g[stride]) - idc); ve = double(int(g[stride + 1]) - idc);
          }
          const double ux = 1 - fx, uy = 1 - fy;
          double q = uy * (ux * va + fx * vb) + fy * (ux * vd + fx * ve);
          const double gy = ux * (vd - va) + fx * (ve - vb);
          const double gx = uy * (vb - va) + fy * (ve - vd);
          q = q - off;
          const double w = wv[u];
          v[0] += q; v[1] += q * q; v[2] += w * q;
          if (deriv) {
            const double g[3] = {gy, gx, gy * ((0 - tr.y) - tc.y) + gx * (tr.x - tc.x)};
#pragma unroll
            for (int j = 0; j < 3; ++j) { v[3 + j] += g[j]; v[6 + j] += q * g[j]; v[9 + j] += w * g[j]; }
            v[12] += g[0] * g[0]; v[13] += g[1] * g[0]; v[14] += g[1] * g[1];
            v[15] += g[2] * g[0]; v[16] += g[2] * g[1]; v[17] += g[2] * g[2];
          }
        }
      }
    };
    if (mode == 1) fast.template operator()<true>();
    else fast.template operator()<false>();
  } else
  for (int base = threadIdx.x; base < P.n; base += kRefineThreads * U) {
    int32_t rcv[U];
    double wv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = base + kRefineThreads * u;
      rcv[u] = i < P.n ? O.rc[i] : 0;
      wv[u] = i < P.n ? O.w[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (base + kRefineThreads * u >= P.n) continue;
      const double dr = sample_row(rcv[u]) - a.lv_orow, dc = sample_col(rcv[u]) - a.lv_ocol;
      double y, x;
      if (aligned) { y = pr + dr; x = pc + dc; }
      else { y = (pr + ca * dr) - sa * dc; x = (pc + sa * dr) + ca * dc; }
      if (!interior && !inside(a, y, x, deriv)) { bad = 1; continue; }
      double q, gy, gx;
      bilinear_at<T>(a, y, x, interior, q, gy, gx);
      q = q - off;
      const double w = wv[u];
      v[0] += q; v[1] += q * q; v[2] += w * q;
      if (deriv) {
        const double g[3] = {gy, gx, gy * ((0 - sa * dr) - ca * dc) + gx * (ca * dr - sa * dc)};
#pragma unroll
        for (int j = 0; j < 3; ++j) { v[3 + j] += g[j]; v[6 + j] += q * g[j]; v[9 + j] += w * g[j]; }
        v[12] += g[0] * g[0]; v[13] += g[1] * g[0]; v[14] += g[1] * g[1];
        v[15] += g[2] * g[0]; v[16] += g[2] * g[1]; v[17] += g[2] * g[2];
      }
    }
  }
  if (block_any(bad, &sm.flag)) return false;
  if (deriv) block_sum18(v, sm.red, sm.tot);  // the derivative pass feeds thread 0's solve only
  else {
    double t[3] = {v[0], v[1], v[2]};
    block_sum<3>(t, sm.red);
    v[0] = t[0]; v[1] = t[1]; v[2] = t[2];
  }
#pragma unroll
  for (int k = 0; k < 18; ++k) out[k] = v[k];
  return true;
}

// Storage-local pose of a search-image pose (matcher.cpp:813-816, 896, 858).
__device__ __forceinline__ void to_local(const RefineArgs& a, double row, double col, double& pr, double& pc) {
  pr = (row - a.crop_r - a.phase) / a.scale;
  pc = (col - a.crop_c - a.phase) / a.scale;
  pr -= a.sr;
  pc -= a.sc;
}

// continuous(): matcher.cpp:121-157
template <class T>
__device__ __forceinline__ double continuous(const RefineArgs& a, const PatchRef& O, double row, double col, double angle, Shared& sm) {
  double pr, pc, m[18];
  to_local(a, row, col, pr, pc);
  if (!bilinear_pass<T>(a, O, pr, pc, angle, false, m, sm)) return NCC_MISSING;
  const double var = m[1] - m[0] * m[0] / double(O.P.n);
  if (!(var > 0)) return NCC_MISSING;
  const double s = clamp1(m[2] / sqrt(var));
  return a.absolute ? fabs(s) : s;
}

// Gauss-Jordan elimination with partial pivoting on an AXES x (AXES+1)
// system, in the reference's operation order (matcher.cpp:231-238); row
// swaps use selects so the matrix stays in registers.
template <int AXES>
__device__ __forceinline__ bool solve(double (&M)[3][4], double* d) {
#pragma unroll
  for (int i = 0; i < AXES; ++i) {
    int pv = i;
#pragma unroll
    for (int j = i + 1; j < AXES; ++j) if (fabs(M[j][i]) > fabs(M[pv][i])) pv = j;
    double piv = M[i][i];
#pragma unroll
    for (int j = i + 1; j < AXES; ++j) if (j == pv) piv = M[j][i];
    if (fabs(piv) < 1e-14) return false;
#pragma unroll
    for (int j = i + 1; j < AXES; ++j)
      if (j == pv)
#pragma unroll
        for (int k = 0; k <= AXES; ++k) { const double t = M[i][k]; M[i][k] = M[j][k]; M[j][k] = t; }
    const double dv = M[i][i];
#pragma unroll
    for (int k = 0; k <= AXES; ++k) if (k >= i) M[i][k] /= dv;
#pragma unroll
    for (int j = 0; j < AXES; ++j)
      if (j != i) {
        const double f = M[j][i];
#pragma unroll
        for (int k = 0; k <= AXES; ++k) if (k >= i) M[j][k] -= f * M[i][k];
      }
  }
#pragma unroll
  for (int i = 0; i < AXES; ++i) d[i] = M[i][AXES];
  return true;
}

// refinement_step(): matcher.cpp:170-240.  Returns the verdict; writes delta
// and (when out_score) the clamped score.
template <class T>
__device__ __forceinline__ bool refinement_step(const RefineArgs& a, const PatchRef& O, double row, double col,
                                                double angle, double* delta, double* out_score,
                                                const double* min_score, Shared& sm) {
  double pr, pc, m[18];
  to_local(a, row, col, pr, pc);
  if (!bilinear_pass<T>(a, O, pr, pc, angle, true, m, sm)) return false;
  if (threadIdx.x == 0) {
    const double n = double(O.P.n);
    const double sum = m[0], dot = m[2];
    double energy = m[1];
    const double gs[3] = {m[3], m[4], m[5]}, wg[3] = {m[9], m[10], m[11]};
    double qg[3] = {m[6], m[7], m[8]};
    const double gg[3][3] = {{m[12], 0, 0}, {m[13], m[14], 0}, {m[15], m[16], m[17]}};
    double sc_out = NCC_MISSING;
    energy -= sum * sum / n;
    int st = 0;
    double d[3] = {0, 0, 0};
    if (energy > 0) {
      const double sign = a.absolute && dot < 0 ? -1 : 1, score = sign * dot / sqrt(energy);
      sc_out = clamp1(score);
      st = !(min_score && clamp1(score) <= *min_score);
      if (st) {
        const int axes = a.rot ? 3 : 2;
        double M[3][4];
#pragma unroll
        for (int i = 0; i < 3; ++i) qg[i] -= sum * gs[i] / n;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
#pragma unroll
          for (int j = 0; j < 3; ++j) {
            const int hi = i > j ? i : j, lo = i > j ? j : i;
            M[i][j] = (gg[hi][lo] - gs[i] * gs[j] / n - qg[i] * qg[j] / energy) / energy;
          }
          M[i][3] = sign * wg[i] / sqrt(energy) - score * qg[i] / energy;
        }
        if (axes == 3) {
          st = solve<3>(M, d);
        } else {
          // the translation-only system: the right-hand side sits in column 2
#pragma unroll
          for (int i = 0; i < 2; ++i) M[i][2] = M[i][3];
          st = solve<2>(M, d);
          d[2] = 0;
        }
        if (st) st = isfinite(d[0]) && isfinite(d[1]) && isfinite(d[2]);
      }
    }
    sm.bcast[0] = st; sm.bcast[1] = d[0]; sm.bcast[2] = d[1]; sm.bcast[3] = d[2]; sm.bcast[4] = sc_out;
    sm.bcast[5] = energy > 0 ? 1 : 0;
  }
  __syncthreads();
  const bool ok = sm.bcast[0] != 0;
  if (sm.bcast[5] != 0 && out_score) *out_score = sm.bcast[4];
  delta[0] = sm.bcast[1]; delta[1] = sm.bcast[2]; delta[2] = sm.bcast[3];
  __syncthreads();
  return ok;
}

__device__ __forceinline__ double clampd(double v, double lo, double hi) { return v < lo ? lo : (hi < v ? hi : v); }

// fit(): matcher.cpp:903-982 (the multi-model memo tables only reuse values).
// With the background guard a fit whose score stays far below min_score after
// iteration 0 or 1 stops; those iterations then score their proposals
// without the next derivatives (matcher.cpp:955-978).
// coarse: refine_coarse_candidate() (matcher.cpp:853-877, the lowest level
// above 0) -- no look-ahead derivatives, no guard, no small-step stop.
// Each kind of evaluation has one call site, so the sampling pass is inlined
// twice (with and without derivatives) rather than once per use.
template <class T>
__device__ __forceinline__ void fit(const RefineArgs& a, const PatchRef& O, NccMatch& c, bool coarse, Shared& sm) {
  if (!a.subpixel) return;
  bool cached = false;
  double next_delta[3] = {0, 0, 0};
  for (int it = 0; it < 6; ++it) {
    double delta[3] = {0, 0, 0};
    if (cached) { delta[0] = next_delta[0]; delta[1] = next_delta[1]; delta[2] = next_delta[2]; }
    else if (!refinement_step<T>(a, O, c.row, c.col, c.angle, delta, nullptr, nullptr, sm)) break;
    cached = false;
    const double factor = fmin(fmin(1., .75 / fmax(1e-20, fabs(delta[0]))),
                               fmin(.75 / fmax(1e-20, fabs(delta[1])), a.rot ? a.step / fmax(1e-20, fabs(delta[2])) : 1.));
    const bool guard0 = !coarse && it == 0 && a.guard_first >= 0 && c.score < a.min_score - a.guard_first;
    const bool guard1 = !coarse && it == 1 && a.guard >= 0 && c.score < a.min_score - a.guard;
    bool accepted = false;
    for (int bt = 0; bt < 4; ++bt) {
      const double st = ldexp(factor, -bt);
      NccMatch im = c;
      im.row += a.scale * st * delta[0];
      im.col += a.scale * st * delta[1];
      im.angle = clampd(im.angle + st * delta[2], a.a_start, a.a_end);
      if (!domain_contains(a, im.row, im.col)) continue;
      bool next_valid = false, need_score = true;
      if (!coarse && bt == 0 && it + 1 < 6 && !guard0 && !guard1) {
        im.score = NCC_MISSING;
        const double floor_score = c.score;
        next_valid = refinement_step<T>(a, O, im.row, im.col, im.angle, next_delta, &im.score, &floor_score, sm);
        need_score = !isfinite(im.score);
      }
      if (need_score) im.score = continuous<T>(a, O, im.row, im.col, im.angle, sm);
      if (im.score > c.score) { c = im; accepted = true; cached = next_valid; break; }
    }
    // the guard's stop tests the score after this iteration's step
    if (!coarse && it == 1 && a.guard >= 0 && c.score < a.min_score - a.guard) break;
    if (!coarse && it == 0 && a.guard_first >= 0 && c.score < a.min_score - a.guard_first) break;
    if (!accepted || (!coarse && hypot(delta[0], delta[1]) < 1e-4 && fabs(delta[2]) < 1e-5)) break;
  }
}

// Direct window NCC at level 0 (correlation_four_avx2 / correlation_avx2,
// kernels_avx2.cpp:105-117, 239-258): per-window offset at the first sample.
// Evaluated by one wave; every lane of the wave gets the score.
template <class T>
__device__ __forceinline__ double direct_window_wave(const RefineArgs& a, const PatchRef& p, int row, int col) {
  const NccPatch& P = p.P;
  if (row + P.rmin < 0 || col + P.cmin < 0 || row + P.rmax >= a.H0 || col + P.cmax >= a.W0) return NCC_MISSING;
  if (row + P.rmin < a.sr || col + P.cmin < a.sc || row + P.rmax >= a.sr + a.sh || col + P.cmax >= a.sc + a.sw)
    return NCC_MISSING;
  const int lr = row - a.sr, lc = col - a.sc;
  const double off = pix<T>(a, lr + sample_row(p.rc[0]), lc + sample_col(p.rc[0]));
  double s = 0, sq = 0, d = 0;
  constexpr int U = 8;
  for (int base = threadIdx.x & 63; base < P.n; base += 64 * U) {
    int32_t rcv[U];
    double wv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = base + 64 * u;
      rcv[u] = i < P.n ? p.rc[i] : 0;
      wv[u] = i < P.n ? p.w[i] : 0.0;
    }
    double qv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) qv[u] = base + 64 * u < P.n ? pix<T>(a, lr + sample_row(rcv[u]), lc + sample_col(rcv[u])) - off : 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (base + 64 * u < P.n) { s += qv[u]; sq += qv[u] * qv[u]; d += qv[u] * wv[u]; }
  }
  s = wave_sum(s); sq = wave_sum(sq); d = wave_sum(d);
  const double var = sq - s * s / double(P.n);
  if (!(var > 0)) return NCC_MISSING;
  const double sc = clamp1(d / sqrt(var));
  return a.absolute ? fabs(sc) : sc;
}

// Window NCC in the row-prefix form on an exact byte lattice
// (correlation_four_prefix_avx2 / correlation_one_prefix_avx2,
// kernels_avx2.cpp:118-171): moments of the dc-relative pixels without a
// per-window offset (exact integers in FP64, so their order does not
// matter), var = sq - sum^2/n, -inf when var <= 128 eps sq, numerator
// D - (sum/n) * weight_sum.  Used by find_models' recovery over the whole level 0.
template <class T>
__device__ __forceinline__ double prefix_window_wave(const RefineArgs& a, const PatchRef& p, int row, int col) {
  const NccPatch& P = p.P;
  if (row + P.rmin < 0 || col + P.cmin < 0 || row + P.rmax >= a.H0 || col + P.cmax >= a.W0) return NCC_MISSING;
  if (row + P.rmin < a.sr || col + P.cmin < a.sc || row + P.rmax >= a.sr + a.sh || col + P.cmax >= a.sc + a.sw)
    return NCC_MISSING;
  const int lr = row - a.sr, lc = col - a.sc;
  double s = 0, sq = 0, d = 0;
  constexpr int U = 8;
  for (int base = threadIdx.x & 63; base < P.n; base += 64 * U) {
    int32_t rcv[U];
    double wv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = base + 64 * u;
      rcv[u] = i < P.n ? p.rc[i] : 0;
      wv[u] = i < P.n ? p.w[i] : 0.0;
    }
    double qv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) qv[u] = base + 64 * u < P.n ? pix<T>(a, lr + sample_row(rcv[u]), lc + sample_col(rcv[u])) : 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (base + 64 * u < P.n) { s += qv[u]; sq += qv[u] * qv[u]; d += qv[u] * wv[u]; }
  }
  s = wave_sum(s); sq = wave_sum(sq); d = wave_sum(d);
  const double n = double(P.n), var = sq - s * s / n;
  if (var <= 128 * 0x1p-52 * sq) return NCC_MISSING;
  const double sc = clamp1((d - s / n * P.wsum) / sqrt(var));
  return a.absolute ? fabs(sc) : sc;
}

// The windows (row, col0 + j), j0 <= j <= j1 (within 0..4), of one level-0
// patch, scored by one wave as direct_window_wave / prefix_window_wave score
// them, every sample's position and weight loaded once for all of them: per
// window the same per-lane order of the sums and the same wave sums (the
// moments are exact integers either way), so the same scores.  The windows
// must lie in the level's storage rectangle (and, with kLds, their pixels in
// the LDS tile).
template <class T, bool kLds>
__device__ __forceinline__ void window_row_wave(const RefineArgs& a, const PatchRef& p, int row, int col0, int j0, int j1,
                                                double (&sc)[5]) {
  extern __shared__ uint8_t lds_tile[];
  const NccPatch& P = p.P;
  const T* lt = reinterpret_cast<const T*>(lds_tile);
  const T* gl = static_cast<const T*>(a.lv.data) + (size_t(a.sr) * a.lv.pitch + a.sc);
  const int stride = kLds ? a.tw : int(a.lv.pitch);
  const int org = kLds ? -a.tr0 * a.tw - a.tc0 : 0;
  const int lr = row - a.sr, lc = col0 - a.sc;
  auto at = [&](int r, int c) __attribute__((always_inline)) -> int {
    if constexpr (kLds) return int(lt[r * stride + c + org]);
    else return int(gl[size_t(r) * size_t(stride) + c]);
  };
  // per window the value the moments are taken from: the first sample's
  // pixel (direct form) or dc (prefix form), S - base an exact integer
  int base[5];
  const int32_t rc0 = p.rc[0];
#pragma unroll
  for (int j = 0; j < 5; ++j)
    base[j] = a.prefix_recovery || j < j0 || j > j1 ? int(a.dc) : at(lr + sample_row(rc0), lc + j + sample_col(rc0));
  int64_t s[5] = {0, 0, 0, 0, 0}, sq[5] = {0, 0, 0, 0, 0};
  double d[5] = {0, 0, 0, 0, 0};
  constexpr int U = 4;
  for (int b = threadIdx.x & 63; b < P.n; b += 64 * U) {
    int32_t rcv[U];
    double wv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = b + 64 * u;
      rcv[u] = i < P.n ? p.rc[i] : 0;
      wv[u] = i < P.n ? p.w[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (b + 64 * u >= P.n) continue;
      const int r = lr + sample_row(rcv[u]), c = lc + sample_col(rcv[u]);
#pragma unroll
      for (int j = 0; j < 5; ++j) {
        if (j < j0 || j > j1) continue;  // uniform
        const int q = at(r, c + j) - base[j];
        const double qd = double(q);
        s[j] += q;
        sq[j] += int64_t(q) * q;
        d[j] += qd * wv[u];
      }
    }
  }
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    sc[j] = NCC_MISSING;
    if (j < j0 || j > j1) continue;
    int64_t ss = s[j], qq = sq[j];
    for (int o = 32; o > 0; o >>= 1) { ss += __shfl_xor(ss, o, 64); qq += __shfl_xor(qq, o, 64); }
    const double sum = double(ss), sumsq = double(qq), dot = wave_sum(d[j]);
    const double n = double(P.n), var = sumsq - sum * sum / n;
    double v;
    if (a.prefix_recovery) {
      if (var <= 128 * 0x1p-52 * sumsq) continue;
      v = clamp1((dot - sum / n * P.wsum) / sqrt(var));
    } else {
      if (!(var > 0)) continue;
      v = clamp1(dot / sqrt(var));
    }
    sc[j] = a.absolute ? fabs(v) : v;
  }
}

// Recovery search (matcher.cpp:986-1059): the 5x5 level-0 anchors around the
// seed for every level-0 patch within one angle step.  A wave takes one
// patch's row of five windows at a time (window_row_wave); the first strict
// maximum above `floor_score` in the reference's (patch, dr, dc) order wins
// (each wave keeps its best with the lowest index, then the block).  Returns
// false when no window beats it.
template <class T>
__device__ __forceinline__ bool recovery_search(const RefineArgs& a, const NccMatch& seed, double floor_score, NccMatch& best,
                                int* elig, int* wave_cnt, Shared& sm) {
  const int n_e = eligible_patches(a.pools, a.refs0, a.n_refs0, seed.angle, a.step * 1.01, elig, wave_cnt);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, n_waves = kRefineThreads / 64;
  double wb = NCC_MISSING;
  int wi = 0x7fffffff;
  for (int k = wave; k < n_e * 5; k += n_waves) {
    const PatchRef p = patch_ref(a.pools, a.refs0[elig[k / 5]]);
    const NccPatch& P = p.P;
    const int r = int(round(seed.row - a.crop_r - P.orow)) + k % 5 - 2;
    const int c0 = int(round(seed.col - a.crop_c - P.ocol)) - 2;
    // the windows inside level 0 and the storage rectangle: columns c0 + j0 .. c0 + j1
    if (r + P.rmin < 0 || r + P.rmax >= a.H0 || r + P.rmin < a.sr || r + P.rmax >= a.sr + a.sh) continue;
    const int j0 = max(0, max(-P.cmin, a.sc - P.cmin) - c0);
    const int j1 = min(4, min(a.W0 - 1 - P.cmax, a.sc + a.sw - 1 - P.cmax) - c0);
    if (j0 > j1) continue;
    const int lr = r - a.sr, lc = c0 - a.sc;
    const bool in_tile = a.th > 0 && lr + P.rmin >= a.tr0 && lr + P.rmax <= a.tr0 + a.th - 1 &&
                         lc + j0 + P.cmin >= a.tc0 && lc + j1 + P.cmax <= a.tc0 + a.tw - 1;
    double sc[5];
    if (in_tile) window_row_wave<T, true>(a, p, r, c0, j0, j1, sc);
    else window_row_wave<T, false>(a, p, r, c0, j0, j1, sc);
    for (int j = j0; j <= j1; ++j) {
      if (!domain_contains(a, a.crop_r + r + P.orow, a.crop_c + c0 + j + P.ocol)) continue;
      if (sc[j] > wb) { wb = sc[j]; wi = (k / 5) * 25 + (k % 5) * 5 + j; }
    }
  }
  if (lane == 0) { sm.win_s[wave] = wb; sm.win_i[wave] = wi; }
  __syncthreads();
  double bs = NCC_MISSING;
  int bi = 0x7fffffff;
  for (int w = 0; w < n_waves; ++w)
    if (sm.win_s[w] > bs || (sm.win_s[w] == bs && sm.win_i[w] < bi)) { bs = sm.win_s[w]; bi = sm.win_i[w]; }
  __syncthreads();
  if (!(bs > floor_score)) return false;
  const PatchRef p = patch_ref(a.pools, a.refs0[elig[bi / 25]]);
  const int r = int(round(seed.row - a.crop_r - p.P.orow)) + (bi % 25) / 5 - 2;
  const int c = int(round(seed.col - a.crop_c - p.P.ocol)) + bi % 5 - 2;
  best.row = a.crop_r + r + p.P.orow;
  best.col = a.crop_c + c + p.P.ocol;
  best.angle = p.P.angle;
  best.score = bs;
  return true;
}

// Largest distance of the unrotated patch's bounding box from its origin.
__device__ __forceinline__ double template_reach(const RefineArgs& a, const NccPatch& P) {
  const double dy = fmax(a.lv_orow - P.rmin, P.rmax - a.lv_orow), dx = fmax(a.lv_ocol - P.cmin, P.cmax - a.lv_ocol);
  return sqrt(dy * dy + dx * dx);
}

// One block per candidate: the continuous score at the finest level
// (matcher.cpp:813-822).
// Candidates in list form (frame f's at f * cap); results in the same slots.
template <class T>
__global__ void __launch_bounds__(kRefineThreads) k_full_score(const RefineArgs args) {
  __shared__ Shared sm;
  __shared__ int off[NCC_MAX_BATCH + 1];
  extern __shared__ uint8_t lds_tile[];
  list_offsets(args.in_count, args.n_frames, args.cap, off);
  const int total = off[args.n_frames];
  for (int g = blockIdx.x; g < total; g += gridDim.x) {
    RefineArgs a = args;
    const int f = list_frame(off, a.n_frames, g);
    const size_t slot = size_t(f) * a.cap + (g - off[f]);
    bind_frame(a, f);
    const PatchRef O = patch_ref(a.pools, a.orig_ref);
    NccMatch cd = a.cand[slot];
    {
      double pr, pc;
      to_local(a, cd.row, cd.col, pr, pc);
      stage_tile<T>(a, lds_tile, pr, pc, template_reach(a, O.P) + 2);
      a.rtab_cap = 0;  // one pass per candidate: building the rotation tables does not pay
    }
    if (isfinite(cd.score)) cd.score = continuous<T>(a, O, cd.row, cd.col, cd.angle, sm);
    if (threadIdx.x == 0) a.out_full[slot] = cd;
    __syncthreads();  // the LDS tile is restaged for the next candidate
  }
}

// One block per candidate: refine_candidate (matcher.cpp:853-1061) from a
// full-scored pose -- Gauss-Newton fit, then for a level-0 fit that stays
// below the minimum score the recovery search and a second fit.
template <class T>
__global__ void __launch_bounds__(kRefineThreads) k_refine(const RefineArgs args) {
  RefineArgs a = args;
  __shared__ Shared sm;
  __shared__ int elig[kMaxEligible];
  __shared__ int wave_cnt[kRefineThreads / 64];
  extern __shared__ uint8_t lds_tile[];
  const int ci = blockIdx.x;
  if (ci >= a.n_cand) return;
  bind_frame(a, a.cand_frame[ci]);
  const PatchRef O = patch_ref(a.pools, a.orig_ref);
  NccMatch cd = a.cand[ci];
  if (!isfinite(cd.score)) {
    if (threadIdx.x == 0) { a.out_refined[ci] = cd; a.verdict[ci] = 0; }
    return;
  }
  {
    // the fit moves at most a few pixels; farther samples fall back to global memory
    double pr, pc;
    to_local(a, cd.row, cd.col, pr, pc);
    stage_tile<T>(a, lds_tile, pr, pc, template_reach(a, O.P) + 12);
  }
  // The sequence below -- fit the pose; at level 0, if it stays under
  // min_score, the recovery search, then fit the recovered pose and keep it
  // if it scores higher -- runs as a loop with one fit call (at most two
  // fits), so the fit's code is inlined once.  find_models recovers a
  // candidate already below min_score before fitting it (matcher.cpp:984-985);
  // when that recovery finds nothing, the candidate itself is fitted.
  {
    const bool coarse = a.finest != 0;
    const bool may_recover = !coarse && a.subpixel && a.top > 0;
    const NccMatch seed = cd;
    const bool recover_first = a.recover_first && may_recover && cd.score < a.min_score;
    NccMatch work = cd;
    bool do_fit = !recover_first, fitting_best = false, recovered = false;
    for (int step = 0; step < 3; ++step) {
      if (do_fit) {
        fit<T>(a, O, work, coarse, sm);
        do_fit = false;
        if (fitting_best) {
          if (work.score > cd.score) cd = work;
          break;
        }
        cd = work;
        if (recovered) break;
      }
      if (recovered || !(may_recover && cd.score < a.min_score)) break;
      recovered = true;
      NccMatch best = cd;
      if (recovery_search<T>(a, seed, cd.score, best, elig, wave_cnt, sm) &&
          (best.row != cd.row || best.col != cd.col || best.angle != cd.angle)) {
        best.score = continuous<T>(a, O, best.row, best.col, best.angle, sm);
        work = best;
        fitting_best = true;
        do_fit = true;
      } else if (recover_first) {
        work = cd;
        do_fit = true;
      } else {
        break;
      }
    }
  }
  if (threadIdx.x == 0) {
    a.out_refined[ci] = cd;
    a.verdict[ci] = (cd.score >= a.min_score && domain_contains(a, cd.row, cd.col)) ? 1 : 0;
  }
}

// Level-1 hand-over to the fit with a reduced domain (matcher.cpp:631-645): a
// candidate whose pose is not on a domain pixel moves to the nearest domain
// pixel of the runs within 4 rows (column clamped to the run and at most 4
// from the pose; the first run in (row, column) order wins ties), else its
// score becomes missing.  One thread per candidate of frame blockIdx.y's list.
__global__ void __launch_bounds__(256) k_domain_snap(const NccFrame* __restrict__ frames, NccMatch* cand,
                                                     const int32_t* __restrict__ count, int cap) {
  const int f = blockIdx.y;
  const NccFrame& F = frames[f];
  if (!F.dom_sat) return;
  const int n = min(max(count[f], 0), cap);
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  NccMatch p = cand[size_t(f) * cap + i];
  if (in_domain(F, p.row, p.col)) return;
  const int64_t h = F.levels[0].height, cr = int64_t(F.crop_r);
  const int64_t r0 = max(int64_t(ceil(p.row - 4)) - cr, int64_t(0));
  const int64_t r1 = min(int64_t(floor(p.row + 4)) - cr, h - 1);
  const int4* runs = reinterpret_cast<const int4*>(F.dom_runs);
  double best = __builtin_inf(), nr = p.row, nc = p.col;
  for (int64_t r = r0; r <= r1; ++r)
    for (int k = F.dom_row[r]; k < F.dom_row[r + 1]; ++k) {
      const int4 run = runs[k];
      double c = round(p.col);
      if (c < double(run.y)) c = double(run.y);
      else if (double(run.z) < c) c = double(run.z);
      if (fabs(c - p.col) > 4) continue;
      const double d = hypot(double(run.x) - p.row, c - p.col);
      if (d < best) { best = d; nr = double(run.x); nc = c; }
    }
  if (isfinite(best)) { p.row = nr; p.col = nc; }
  else p.score = NCC_MISSING;
  cand[size_t(f) * cap + i] = p;
}

}  // namespace ncc

using namespace ncc;

extern "C" int ncchip_track(const NccFrame* frames, int n_frames, int level, int level_kind, const NccPatchPool* pools,
                            const int32_t* refs, int n_refs, const NccMatch* in, const int32_t* in_count, int cap,
                            NccMatch* raw, NccMatch* out, int32_t* out_count, int32_t* table, int table_size,
                            double keep, double tol, double step, double offset, int exact_lattice, int absolute,
                            int tile_pixels, int exact_f64, int max_eligible, double x_max,
                            unsigned long long* work, int grid, void* stream) {
  if (n_frames <= 0 || grid <= 0) return 0;
  if (n_frames > NCC_MAX_BATCH || table_size < 2 || (table_size & (table_size - 1))) return -1;
  TrackArgs a{};
  a.frames = frames; a.n_frames = n_frames; a.level = level; a.pools.p0 = pools[0]; a.pools.p1 = pools[1];
  a.refs = refs; a.n_refs = n_refs; a.in = in; a.in_count = in_count; a.cap = cap; a.raw = raw; a.out = out;
  a.out_count = out_count; a.table = table; a.table_size = table_size; a.keep = keep; a.work = work;
  a.tol = tol; a.step = step; a.offset = offset; a.exact_lattice = exact_lattice; a.absolute = absolute;
  const size_t px = level_kind == NCC_U8 ? 1 : level_kind == NCC_U16 ? 2 : 4;
  constexpr size_t kMaxTileBytes = 64 * 1024;
  a.tile_cap = tile_pixels > 0 && size_t(tile_pixels) * px <= kMaxTileBytes ? tile_pixels : 0;
  const size_t lds = (size_t(a.tile_cap) * px + 15) & ~size_t(15);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 g(grid), b(kTrackThreads);
  if (level_kind != NCC_U32 && a.tile_cap > 0 && size_t(a.tile_cap) * 2 <= kMaxTileBytes && max_eligible > 0 &&
      max_eligible <= kWinPatches) {
    // every candidate's neighbourhood in LDS and at most 32 eligible patches:
    // the window-major kernel (moments exact in integers, whatever exact_f64)
    const bool x16 = x_max < 32768;  // x = S - D as int16 (the tile has the pixel size of 16-bit levels)
    const size_t lds16 = (size_t(a.tile_cap) * 2 + 15) & ~size_t(15);
    if (level_kind == NCC_U8) {
      if (x16) hipLaunchKernelGGL((k_track_win<uint8_t, true>), g, b, lds16, st, a);
      else hipLaunchKernelGGL((k_track_win<uint8_t, false>), g, b, lds, st, a);
    } else {
      if (x16) hipLaunchKernelGGL((k_track_win<uint16_t, true>), g, b, lds16, st, a);
      else hipLaunchKernelGGL((k_track_win<uint16_t, false>), g, b, lds, st, a);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1000;
  }
  if (level_kind == NCC_U8) {
    if (exact_f64) hipLaunchKernelGGL((k_track<uint8_t, double>), g, b, lds, st, a);
    else hipLaunchKernelGGL((k_track<uint8_t, int64_t>), g, b, lds, st, a);
  } else if (level_kind == NCC_U16) {
    if (exact_f64) hipLaunchKernelGGL((k_track<uint16_t, double>), g, b, lds, st, a);
    else hipLaunchKernelGGL((k_track<uint16_t, int64_t>), g, b, lds, st, a);
  } else {
    if (exact_f64) hipLaunchKernelGGL((k_track<uint32_t, double>), g, b, lds, st, a);
    else hipLaunchKernelGGL((k_track<uint32_t, int64_t>), g, b, lds, st, a);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1000;
}

extern "C" int ncchip_track_mfma(const NccFrame* frames, int n_frames, int level, int level_kind,
                                 const NccPatchPool* pools, const int32_t* refs, int n_refs, const NccMatch* in,
                                 const int32_t* in_count, int cap, NccMatch* raw, NccMatch* out, int32_t* out_count,
                                 int32_t* table, int table_size, double keep, double tol, double step, double offset,
                                 int exact_lattice, int absolute, int tile_h, int tile_w, double x_max,
                                 unsigned long long* work, int grid, void* stream) {
  if (n_frames <= 0 || grid <= 0) return 0;
  if (n_frames > NCC_MAX_BATCH || table_size < 2 || (table_size & (table_size - 1))) return -1;
  // the exact BF16 split of S - dc 4^l holds for 16-bit block sums up to level 4
  if (level_kind != NCC_U16 || level < 1 || level > 4 || tile_h < 5 || tile_w < 5) return -6;
  TrackArgs a{};
  a.frames = frames; a.n_frames = n_frames; a.level = level; a.pools.p0 = pools[0]; a.pools.p1 = pools[1];
  a.refs = refs; a.n_refs = n_refs; a.in = in; a.in_count = in_count; a.cap = cap; a.raw = raw; a.out = out;
  a.out_count = out_count; a.table = table; a.table_size = table_size; a.keep = keep; a.work = work;
  a.tol = tol; a.step = step; a.offset = offset; a.exact_lattice = exact_lattice; a.absolute = absolute;
  a.x_max = x_max;
  const int xc = (tile_w - 4 + 15) / 16;
  a.mt_h = tile_h; a.mt_w = tile_w; a.mt_pitch = (16 * xc + 6) & ~1;
  auto al16 = [](size_t v) { return (v + 15) & ~size_t(15); };
  // sum of S^2 over a tile row in 32 bits when it cannot overflow
  const bool p2_32 = 255.0 * 255.0 * std::ldexp(1.0, 4 * level) * double(tile_w + 1) < 4294967295.0;
  size_t o = 0;
  a.mt_s = int(o); o = al16(o + size_t(tile_h) * tile_w * 2);
  a.mt_p1 = int(o); o = al16(o + size_t(tile_h) * (tile_w + 1) * 4);
  a.mt_p2 = int(o); o = al16(o + size_t(tile_h) * (tile_w + 1) * (p2_32 ? 4 : 8));
  a.mt_pl = int(o); o = al16(o + std::max<size_t>(size_t(2) * tile_h * a.mt_pitch * 2, 2 * 32 * 32 * 4));
  // at least two workgroups per CU: a larger neighbourhood (one workgroup
  // per CU) tracks slower than the FP64 vector kernel
  if (o > 64 * 1024) return -6;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (p2_32) hipLaunchKernelGGL((k_track_mfma<uint16_t, uint32_t>), dim3(grid), dim3(kTrackThreads), o, st, a);
  else hipLaunchKernelGGL((k_track_mfma<uint16_t, uint64_t>), dim3(grid), dim3(kTrackThreads), o, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -1000;
}

namespace {
RefineArgs refine_args(const NcchipRefineParams* p, const NccPatchPool* pools) {
  RefineArgs a{};
  a.frames = p->frames; a.scale = p->scale; a.phase = p->phase;
  a.finest = p->finest; a.top = p->top; a.pools.p0 = pools[0]; a.pools.p1 = pools[1]; a.orig_ref = p->orig_ref;
  a.lv_orow = p->level_orow; a.lv_ocol = p->level_ocol; a.refs0 = p->refs0; a.n_refs0 = p->n_refs0;
  a.step = p->step; a.a_start = p->angle_start; a.a_end = p->angle_end; a.min_score = p->min_score;
  a.subpixel = p->subpixel; a.rot = p->rotate; a.absolute = p->absolute;
  a.recover_first = p->recover_first; a.prefix_recovery = p->prefix_recovery;
  a.guard_first = p->guard_first; a.guard = p->guard;
  return a;
}
}  // namespace

extern "C" int ncchip_full_score(const NcchipRefineParams* p, const NccPatchPool* pools, int n_frames,
                                 const NccMatch* in, const int32_t* in_count, int cap, NccMatch* out_full, int grid,
                                 void* stream) {
  if (n_frames <= 0 || grid <= 0) return 0;
  if (n_frames > NCC_MAX_BATCH) return -1;
  RefineArgs a = refine_args(p, pools);
  a.cand = in; a.n_frames = n_frames; a.in_count = in_count; a.cap = cap; a.out_full = out_full;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 g(grid), b(kRefineThreads);
  if (p->level_kind == NCC_U8) hipLaunchKernelGGL(k_full_score<uint8_t>, g, b, kRefineTileBytes, st, a);
  else if (p->level_kind == NCC_U16) hipLaunchKernelGGL(k_full_score<uint16_t>, g, b, kRefineTileBytes, st, a);
  else hipLaunchKernelGGL(k_full_score<uint32_t>, g, b, kRefineTileBytes, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -1000;
}

extern "C" int ncchip_refine(const NcchipRefineParams* p, const NccPatchPool* pools, const NccMatch* cand,
                             const int32_t* cand_frame, int n_cand, NccMatch* out_refined, int32_t* verdict,
                             void* stream) {
  if (n_cand <= 0) return 0;
  RefineArgs a = refine_args(p, pools);
  a.cand = cand; a.cand_frame = cand_frame; a.n_cand = n_cand; a.out_refined = out_refined; a.verdict = verdict;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 g(n_cand), b(kRefineThreads);
  if (p->level_kind == NCC_U8) hipLaunchKernelGGL(k_refine<uint8_t>, g, b, kRefineTileBytes, st, a);
  else if (p->level_kind == NCC_U16) hipLaunchKernelGGL(k_refine<uint16_t>, g, b, kRefineTileBytes, st, a);
  else hipLaunchKernelGGL(k_refine<uint32_t>, g, b, kRefineTileBytes, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -1000;
}

extern "C" int ncchip_domain_snap(const NccFrame* frames, int n_frames, NccMatch* cand, const int32_t* count, int cap,
                                  void* stream) {
  if (n_frames <= 0 || cap <= 0) return 0;
  if (n_frames > 65535) return -1;
  hipLaunchKernelGGL(k_domain_snap, dim3((cap + 255) / 256, n_frames), dim3(256), 0, static_cast<hipStream_t>(stream),
                     frames, cand, count, cap);
  return hipGetLastError() == hipSuccess ? 0 : -1000;
}
