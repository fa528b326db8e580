// Integer passes of the NCC find path: search pyramid (exact block sums),
// top-level row prefix sums of S and S^2, and the 2D support integral.
// All results are bit-exact images of the reference's FP64 values
// (matcher.cpp:401-455, 487-495, 498-503).
#include <algorithm>
#include <mutex>

#include "ncc_common.hip.h"
#include "../../include/ncchip.h"

namespace ncc {

template <class Tin>
__device__ __forceinline__ uint32_t ld(const Tin* p, size_t i) { return uint32_t(p[i]); }

// One output level from the previous one: S_{l+1}(r,c) = sum of the 2x2 block
// at (2r, 2c) with the reference's edge clamp (matcher.cpp:442, 448-450).
// Each thread writes 4 consecutive outputs of row r starting at column c0.
template <class Tin, class Tout>
__device__ __forceinline__ void pyramid_row4(const Tin* __restrict__ src, int sp, int sw, int sh,
                                             Tout* __restrict__ dst, int dp, int dw, int dh, int r, int c0) {
  if (r >= dh || c0 >= dw) return;
  const int r0 = 2 * r, r1 = min(r0 + 1, sh - 1);
  const Tin* a = src + size_t(r0) * sp;
  const Tin* b = src + size_t(r1) * sp;
  Tout out[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int c = c0 + j;
    if (c < dw) {
      const int x0 = 2 * c, x1 = min(x0 + 1, sw - 1);
      out[j] = Tout(ld(a, x0) + ld(a, x1) + ld(b, x0) + ld(b, x1));
    }
  }
  Tout* o = dst + size_t(r) * dp + c0;
  if (c0 + 3 < dw) {
#pragma unroll
    for (int j = 0; j < 4; ++j) o[j] = out[j];
  } else {
    for (int j = 0; j < 4 && c0 + j < dw; ++j) o[j] = out[j];
  }
}

template <class Tin, class Tout>
__global__ void __launch_bounds__(256) k_pyramid(const Tin* __restrict__ src, int sp, int sw, int sh,
                                                 Tout* __restrict__ dst, int dp, int dw, int dh) {
  for (int r = blockIdx.y; r < dh; r += gridDim.y)
    pyramid_row4(src, sp, sw, sh, dst, dp, dw, dh, r, (blockIdx.x * 256 + threadIdx.x) * 4);
}

// Batched form: frame blockIdx.z builds `level` from level - 1.
template <class Tin, class Tout>
__global__ void __launch_bounds__(256) k_pyramid_batch(const NccFrame* __restrict__ frames, int level) {
  const NccLevel& s = frames[blockIdx.z].levels[level - 1];
  const NccLevel& d = frames[blockIdx.z].levels[level];
  for (int r = blockIdx.y; r < d.height; r += gridDim.y)
    pyramid_row4(static_cast<const Tin*>(s.data), s.pitch, s.width, s.height,
                 static_cast<Tout*>(const_cast<void*>(d.data)), d.pitch, d.width, d.height, r,
                 (blockIdx.x * 256 + threadIdx.x) * 4);
}

// Row prefix of S (u32) and S^2 (u64) over one level row, plus the row
// prefix of the support indicator (S != D, i.e. q != 0), one block per row.
// Layout: (w+1) entries per row with a leading zero (matcher.cpp:490-494).
template <class T>
__device__ __forceinline__ void row_prefix(const T* __restrict__ src, int sp, int w, int h, uint32_t D, int r,
                                           uint64_t* __restrict__ p1, uint64_t* __restrict__ p2,
                                           uint32_t* __restrict__ nz) {
  __shared__ uint64_t s1[4];
  __shared__ uint64_t s2[4];
  __shared__ uint32_t s3[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const T* row = src + size_t(r) * sp;
  uint64_t* o1 = p1 + size_t(r) * (w + 1);
  uint64_t* o2 = p2 + size_t(r) * (w + 1);
  uint32_t* o3 = nz + size_t(r) * (w + 1);
  if (threadIdx.x == 0) { o1[0] = 0; o2[0] = 0; o3[0] = 0; }
  uint32_t carry3 = 0;
  uint64_t carry1 = 0, carry2 = 0;
  for (int base = 0; base < w; base += 256) {
    const int c = base + threadIdx.x;
    const uint32_t s = c < w ? uint32_t(row[c]) : 0u;
    uint64_t v1 = s;
    uint32_t v3 = (c < w && s != D) ? 1u : 0u;
    uint64_t v2 = uint64_t(s) * s;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint64_t t1 = __shfl_up(v1, d, 64);
      const uint64_t t2 = __shfl_up(v2, d, 64);
      const uint32_t t3 = __shfl_up(v3, d, 64);
      if (lane >= d) { v1 += t1; v2 += t2; v3 += t3; }
    }
    if (lane == 63) { s1[wave] = v1; s2[wave] = v2; s3[wave] = v3; }
    __syncthreads();
    uint32_t add3 = carry3;
    uint64_t add1 = carry1, add2 = carry2;
    for (int k = 0; k < wave; ++k) { add1 += s1[k]; add2 += s2[k]; add3 += s3[k]; }
    if (c < w) { o1[c + 1] = v1 + add1; o2[c + 1] = v2 + add2; o3[c + 1] = v3 + add3; }
    carry1 += s1[0] + s1[1] + s1[2] + s1[3];
    carry2 += s2[0] + s2[1] + s2[2] + s2[3];
    carry3 += s3[0] + s3[1] + s3[2] + s3[3];
    __syncthreads();
  }
}

template <class T>
__global__ void __launch_bounds__(256) k_row_prefix(const T* __restrict__ src, int sp, int w, int h,
                                                    const uint8_t* __restrict__ dc_src, int shift,
                                                    uint64_t* __restrict__ p1, uint64_t* __restrict__ p2,
                                                    uint32_t* __restrict__ nz) {
  if (int(blockIdx.x) >= h) return;
  row_prefix(src, sp, w, h, uint32_t(*dc_src) << shift, blockIdx.x, p1, p2, nz);
}

// Batched form: row blockIdx.x of frame blockIdx.y's `level`.
template <class T>
__global__ void __launch_bounds__(256) k_row_prefix_batch(const NccFrame* __restrict__ frames, int level) {
  const NccFrame& F = frames[blockIdx.y];
  const NccLevel& L = F.levels[level];
  if (int(blockIdx.x) >= L.height) return;
  row_prefix(static_cast<const T*>(L.data), L.pitch, L.width, L.height, frame_dc(F) << L.shift, blockIdx.x,
             F.p1, F.p2, F.nz);
}

// Copies n 32-bit words (device to device or to host-visible memory).
__global__ void __launch_bounds__(256) k_copy_u32(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst,
                                                  size_t n) {
  for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += size_t(gridDim.x) * 256) dst[i] = src[i];
}

// Fills n 32-bit words (a kernel rather than a memset, so that the sequence
// can be captured in a HIP graph).
__global__ void __launch_bounds__(256) k_fill_u32(uint32_t* __restrict__ p, size_t n, uint32_t v) {
  for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += size_t(gridDim.x) * 256) p[i] = v;
}

// Column pass turning row prefixes of the support indicator into the 2D
// integral support[(r+1)*(w+1) + c] (first row zero), matcher.cpp:498-499.
__global__ void __launch_bounds__(256) k_support_cols(const uint32_t* __restrict__ nz, int w, int h,
                                                      uint32_t* __restrict__ sup) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c > w) return;
  uint32_t acc = 0;
  sup[c] = 0;
  for (int r = 0; r < h; ++r) {
    acc += nz[size_t(r) * (w + 1) + c];
    sup[size_t(r + 1) * (w + 1) + c] = acc;
  }
}

// Levels 1..n (n <= 3) of a u8 level 0 in one pass.  The general tile: a
// block stages 8 rows x 1024 columns of level 0 in LDS and derives the
// 4 x 512, 2 x 256 and 1 x 128 block sums below it with the per-level edge
// clamp (a clamped source row or column stays inside the block's tile, as the
// tiles are aligned to the 2x2 lattice of every level), writing each level
// once.
constexpr int kFuseRows = 8, kFuseCols = 1024;
__device__ __forceinline__ void pyramid3_tile(const NccFrame& F, int n_levels, const int R0, const int C0) {
  __shared__ uint16_t t0[kFuseRows][kFuseCols];       // level 0 (u8 values)
  __shared__ uint16_t t1[kFuseRows / 2][kFuseCols / 2];
  __shared__ uint16_t t2[kFuseRows / 4][kFuseCols / 4];
  const NccLevel& L0 = F.levels[0];
  if (R0 >= L0.height || C0 >= L0.width) return;  // uniform per block
  const uint8_t* src = static_cast<const uint8_t*>(L0.data);
  // level 0: 4 consecutive bytes per thread and row, one dword when the
  // rows are dword-aligned (all rows' loads issue before the LDS writes)
  const int c = 4 * threadIdx.x, gc = C0 + c;
  const bool words = ((reinterpret_cast<uintptr_t>(src) | uintptr_t(L0.pitch)) & 3) == 0 && gc + 3 < L0.width;
  uint32_t v[kFuseRows];
#pragma unroll
  for (int r = 0; r < kFuseRows; ++r) {
    const int gr = R0 + r;
    v[r] = 0;
    if (gr < L0.height) {
      const uint8_t* p = src + size_t(gr) * L0.pitch + gc;
      if (words) v[r] = *reinterpret_cast<const uint32_t*>(p);
      else
        for (int j = 0; j < 4; ++j) v[r] |= uint32_t(gc + j < L0.width ? p[j] : 0) << (8 * j);
    }
  }
#pragma unroll
  for (int r = 0; r < kFuseRows; ++r)
#pragma unroll
    for (int j = 0; j < 4; ++j) t0[r][c + j] = uint16_t((v[r] >> (8 * j)) & 0xff);
  __syncthreads();
  // level l + 1 from level l: out (r, c) = sum of rows 2r, min(2r + 1, h - 1)
  // and columns 2c, min(2c + 1, w - 1) of level l (gray_pyramid's clamp);
  // the result also goes to the LDS tile of the next level when there is one
  auto step = [&](const uint16_t* in, int in_pitch, int in_r0, int in_c0, int in_h, int in_w, uint16_t* out,
                  int out_pitch, const NccLevel& D, int rows, int cols) {
    uint16_t* dst = static_cast<uint16_t*>(const_cast<void*>(D.data));
    for (int i = threadIdx.x; i < rows * cols; i += 256) {
      const int r = i / cols, c = i % cols;
      const int gr = in_r0 / 2 + r, gc = in_c0 / 2 + c;
      if (gr >= D.height || gc >= D.width) continue;
      const int y0 = 2 * gr - in_r0, y1 = min(2 * gr + 1, in_h - 1) - in_r0;
      const int x0 = 2 * gc - in_c0, x1 = min(2 * gc + 1, in_w - 1) - in_c0;
      const uint32_t v = uint32_t(in[y0 * in_pitch + x0]) + in[y0 * in_pitch + x1] + in[y1 * in_pitch + x0] +
                         in[y1 * in_pitch + x1];
      if (out) out[r * out_pitch + c] = uint16_t(v);
      dst[size_t(gr) * D.pitch + gc] = uint16_t(v);
    }
  };
  step(&t0[0][0], kFuseCols, R0, C0, L0.height, L0.width, n_levels > 1 ? &t1[0][0] : nullptr, kFuseCols / 2,
       F.levels[1], kFuseRows / 2, kFuseCols / 2);
  if (n_levels < 2) return;
  __syncthreads();
  step(&t1[0][0], kFuseCols / 2, R0 / 2, C0 / 2, F.levels[1].height, F.levels[1].width,
       n_levels > 2 ? &t2[0][0] : nullptr, kFuseCols / 4, F.levels[2], kFuseRows / 4, kFuseCols / 4);
  if (n_levels < 3) return;
  __syncthreads();
  step(&t2[0][0], kFuseCols / 4, R0 / 4, C0 / 4, F.levels[2].height, F.levels[2].width, nullptr, 0, F.levels[3],
       kFuseRows / 8, kFuseCols / 8);
}

// Sums of the byte pairs (0,1) and (2,3) of x, as two 16-bit halves.
__device__ __forceinline__ uint32_t byte_pairs(uint32_t x) { return (x & 0x00ff00ffu) + ((x >> 8) & 0x00ff00ffu); }
// Sum of the two 16-bit halves of x.
__device__ __forceinline__ uint32_t half_sum(uint32_t x) { return (x & 0xffffu) + (x >> 16); }

// A block covers 16 rows x 1024 columns of level 0.  A tile inside the frame
// whose rows and level rows are 16-byte aligned runs in registers: wave w
// loads rows 4w .. 4w + 3 and lane l columns 16l .. 16l + 15 (one 16-byte
// load per row), forms its 2 x 8 level-1 sums as packed byte-pair sums
// (stored as one 16-byte vector per row) and its 4 level-2 sums from them;
// level 3 joins the level-2 rows of two waves through LDS.  No clamp applies
// inside the frame (its last rows and columns lie in edge tiles), and every
// value is an integer sum of the same four inputs, so the levels equal the
// general tile's.  Edge and unaligned tiles run the general tile on each
// half.
__global__ void __launch_bounds__(256) k_pyramid3_u8(const NccFrame* __restrict__ frames, int n_levels) {
  __shared__ uint32_t l2p[4][64][2];  // per wave and lane: its level-2 sums added in column pairs
  const NccFrame& F = frames[blockIdx.z];
  const NccLevel& L0 = F.levels[0];
  const int R0 = blockIdx.y * 2 * kFuseRows, C0 = blockIdx.x * kFuseCols;
  if (R0 >= L0.height || C0 >= L0.width) return;  // uniform per block
  const uint8_t* src = static_cast<const uint8_t*>(L0.data);
  uintptr_t align = reinterpret_cast<uintptr_t>(src) | uintptr_t(L0.pitch);
  for (int l = 1; l <= n_levels; ++l)
    align |= reinterpret_cast<uintptr_t>(F.levels[l].data) | (uintptr_t(F.levels[l].pitch) * 2);
  if (R0 + 2 * kFuseRows > L0.height || C0 + kFuseCols > L0.width || (align & 15)) {
    pyramid3_tile(F, n_levels, R0, C0);
    __syncthreads();  // the tile's LDS is reused by the second half
    pyramid3_tile(F, n_levels, R0 + kFuseRows, C0);
    return;
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint4 v[4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
    v[r] = *reinterpret_cast<const uint4*>(src + size_t(R0 + 4 * wave + r) * L0.pitch + C0 + 16 * lane);
  // level 1: rows R0/2 + 2 wave + {0, 1}, columns C0/2 + 8 lane .. + 7 (two per dword)
  uint4 q[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    q[h].x = byte_pairs(v[2 * h].x) + byte_pairs(v[2 * h + 1].x);
    q[h].y = byte_pairs(v[2 * h].y) + byte_pairs(v[2 * h + 1].y);
    q[h].z = byte_pairs(v[2 * h].z) + byte_pairs(v[2 * h + 1].z);
    q[h].w = byte_pairs(v[2 * h].w) + byte_pairs(v[2 * h + 1].w);
    const NccLevel& D = F.levels[1];
    *reinterpret_cast<uint4*>(static_cast<uint16_t*>(const_cast<void*>(D.data)) +
                              size_t(R0 / 2 + 2 * wave + h) * D.pitch + C0 / 2 + 8 * lane) = q[h];
  }
  if (n_levels < 2) return;
  // level 2: row R0/4 + wave, columns C0/4 + 4 lane .. + 3
  const uint32_t s0 = half_sum(q[0].x) + half_sum(q[1].x), s1 = half_sum(q[0].y) + half_sum(q[1].y);
  const uint32_t s2 = half_sum(q[0].z) + half_sum(q[1].z), s3 = half_sum(q[0].w) + half_sum(q[1].w);
  {
    const NccLevel& D = F.levels[2];
    *reinterpret_cast<uint2*>(static_cast<uint16_t*>(const_cast<void*>(D.data)) + size_t(R0 / 4 + wave) * D.pitch +
                              C0 / 4 + 4 * lane) = make_uint2(s0 | (s1 << 16), s2 | (s3 << 16));
  }
  if (n_levels < 3) return;
  // level 3: rows R0/8 + {0, 1} from level-2 rows of waves (0, 1) and (2, 3),
  // columns C0/8 + 2 lane + {0, 1}
  l2p[wave][lane][0] = s0 + s1;
  l2p[wave][lane][1] = s2 + s3;
  __syncthreads();
  if (wave == 0 || wave == 2) {
    const uint32_t a = l2p[wave][lane][0] + l2p[wave + 1][lane][0], b = l2p[wave][lane][1] + l2p[wave + 1][lane][1];
    const NccLevel& D = F.levels[3];
    *reinterpret_cast<uint32_t*>(static_cast<uint16_t*>(const_cast<void*>(D.data)) +
                                 size_t(R0 / 8 + wave / 2) * D.pitch + C0 / 8 + 2 * lane) = a | (b << 16);
  }
}

// Row pass of a frame's domain summed-area table: entry (r + 1, c + 1) = the
// domain pixels of crop row r in columns 0..c (one block per crop row,
// blockIdx.y = frame; frames without a reduced domain return at once).
__global__ void __launch_bounds__(256) k_domain_rows(const NccFrame* __restrict__ frames) {
  const NccFrame& F = frames[blockIdx.y];
  if (!F.dom_sat) return;
  const int h = F.levels[0].height, w = F.levels[0].width;
  const int r = blockIdx.x;
  if (r >= h) return;
  __shared__ uint32_t part[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t row = int64_t(F.crop_r) + r, c0 = int64_t(F.crop_c);
  const int first = F.dom_row[r], last = F.dom_row[r + 1];
  uint32_t* out = F.dom_sat + size_t(r + 1) * (size_t(w) + 1);
  if (threadIdx.x == 0) out[0] = 0;
  uint32_t carry = 0;
  for (int base = 0; base < w; base += 256) {
    const int c = base + threadIdx.x;
    uint32_t v = 0;
    if (c < w && first < last) {
      // the row's last run starting at or before the column (runs of a row
      // are sorted by first column)
      const int4* runs = reinterpret_cast<const int4*>(F.dom_runs);
      const int64_t x = c0 + c;
      int lo = first, hi = last - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (runs[mid].y <= x) lo = mid; else hi = mid - 1;
      }
      const int4 run = runs[lo];
      v = (run.x == row && run.y <= x && x <= run.z) ? 1u : 0u;
    }
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t t = __shfl_up(v, d, 64);
      if (lane >= d) v += t;
    }
    if (lane == 63) part[wave] = v;
    __syncthreads();
    uint32_t add = carry;
    for (int k = 0; k < wave; ++k) add += part[k];
    if (c < w) out[c + 1] = v + add;
    carry += part[0] + part[1] + part[2] + part[3];
    __syncthreads();
  }
}

// Column pass: accumulates the row prefixes down each column in place and
// writes the zero first row (one thread per column, blockIdx.y = frame).
__global__ void __launch_bounds__(256) k_domain_cols(const NccFrame* __restrict__ frames) {
  const NccFrame& F = frames[blockIdx.y];
  if (!F.dom_sat) return;
  const int h = F.levels[0].height, w = F.levels[0].width;
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c > w) return;
  const size_t p = size_t(w) + 1;
  uint32_t* S = F.dom_sat;
  S[c] = 0;
  uint32_t acc = 0;
  for (int r = 1; r <= h; ++r) {
    acc += S[size_t(r) * p + c];
    S[size_t(r) * p + c] = acc;
  }
}

}  // namespace ncc

using namespace ncc;

extern "C" int ncchip_pyramid3_u8(const NccFrame* frames, int n_frames, int n_levels, int max_w, int max_h,
                                  void* stream) {
  if (n_frames <= 0 || n_levels <= 0) return 0;
  if (n_frames > 65535 || n_levels > 3) return -1;
  const dim3 grid((max_w + kFuseCols - 1) / kFuseCols, (max_h + 2 * kFuseRows - 1) / (2 * kFuseRows),
                  unsigned(n_frames));
  hipLaunchKernelGGL(k_pyramid3_u8, grid, dim3(256), 0, static_cast<hipStream_t>(stream), frames, n_levels);
  return hipGetLastError() == hipSuccess ? 0 : -1000;
}

extern "C" int ncchip_domain_sat(const NccFrame* frames, int n_frames, int max_h, int max_w, void* stream) {
  if (n_frames <= 0 || n_frames > 65535 || max_h <= 0 || max_w <= 0) return n_frames ? -1 : 0;
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_domain_rows, dim3(max_h, n_frames), dim3(256), 0, st, frames);
  hipLaunchKernelGGL(k_domain_cols, dim3((max_w + 1 + 255) / 256, n_frames), dim3(256), 0, st, frames);
  return hipGetLastError() == hipSuccess ? 0 : -1000;
}

static inline hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

extern "C" int ncchip_pyramid_level(const void* src, int src_kind, int src_pitch, int src_w, int src_h, void* dst,
                                    int dst_kind, int dst_pitch, int dst_w, int dst_h, void* stream) {
  const dim3 grid((dst_w + 1023) / 1024, std::min(dst_h, 65535));
  hipStream_t st = as_stream(stream);
#define NCC_PYR(TI, TO) \
  hipLaunchKernelGGL((k_pyramid<TI, TO>), grid, dim3(256), 0, st, static_cast<const TI*>(src), src_pitch, src_w, \
                     src_h, static_cast<TO*>(dst), dst_pitch, dst_w, dst_h)
  if (src_kind == NCC_U8 && dst_kind == NCC_U16) NCC_PYR(uint8_t, uint16_t);
  else if (src_kind == NCC_U16 && dst_kind == NCC_U16) NCC_PYR(uint16_t, uint16_t);
  else if (src_kind == NCC_U16 && dst_kind == NCC_U32) NCC_PYR(uint16_t, uint32_t);
  else if (src_kind == NCC_U32 && dst_kind == NCC_U32) NCC_PYR(uint32_t, uint32_t);
  else if (src_kind == NCC_U8 && dst_kind == NCC_U32) NCC_PYR(uint8_t, uint32_t);
  else return -1;
#undef NCC_PYR
  return hipGetLastError() == hipSuccess ? 0 : -1000;
}

extern "C" int ncchip_pyramid_batch(const NccFrame* frames, int n_frames, int level, int src_kind, int dst_kind,
                                    int max_w, int max_h, void* stream) {
  if (n_frames <= 0 || n_frames > 65535) return n_frames ? -1 : 0;
  const dim3 grid((max_w + 1023) / 1024, std::min(max_h, 65535), n_frames);
  hipStream_t st = as_stream(stream);
#define NCC_PYR(TI, TO) hipLaunchKernelGGL((k_pyramid_batch<TI, TO>), grid, dim3(256), 0, st, frames, level)
  if (src_kind == NCC_U8 && dst_kind == NCC_U16) NCC_PYR(uint8_t, uint16_t);
  else if (src_kind == NCC_U16 && dst_kind == NCC_U16) NCC_PYR(uint16_t, uint16_t);
  else if (src_kind == NCC_U16 && dst_kind == NCC_U32) NCC_PYR(uint16_t, uint32_t);
  else if (src_kind == NCC_U32 && dst_kind == NCC_U32) NCC_PYR(uint32_t, uint32_t);
  else if (src_kind == NCC_U8 && dst_kind == NCC_U32) NCC_PYR(uint8_t, uint32_t);
  else return -1;
#undef NCC_PYR
  return hipGetLastError() == hipSuccess ? 0 : -1000;
}

extern "C" int ncchip_row_prefix_batch(const NccFrame* frames, int n_frames, int level, int kind, int max_h,
                                       void* stream) {
  if (n_frames <= 0 || n_frames > 65535) return n_frames ? -1 : 0;
  const dim3 grid(max_h, n_frames);
  hipStream_t st = as_stream(stream);
  if (kind == NCC_U8) hipLaunchKernelGGL((k_row_prefix_batch<uint8_t>), grid, dim3(256), 0, st, frames, level);
  else if (kind == NCC_U16) hipLaunchKernelGGL((k_row_prefix_batch<uint16_t>), grid, dim3(256), 0, st, frames, level);
  else hipLaunchKernelGGL((k_row_prefix_batch<uint32_t>), grid, dim3(256), 0, st, frames, level);
  return hipGetLastError() == hipSuccess ? 0 : -1000;
}

extern "C" int ncchip_fill_u32(uint32_t* p, size_t n, uint32_t value, void* stream) {
  if (!n) return 0;
  const size_t blocks = std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_u32, dim3(unsigned(blocks)), dim3(256), 0, as_stream(stream), p, n, value);
  return hipGetLastError() == hipSuccess ? 0 : -1000;
}

extern "C" int ncchip_copy_u32(const uint32_t* src, uint32_t* dst, size_t n, void* stream) {
  if (!n) return 0;
  const size_t blocks = std::min<size_t>((n + 255) / 256, 1024);
  hipLaunchKernelGGL(k_copy_u32, dim3(unsigned(blocks)), dim3(256), 0, as_stream(stream), src, dst, n);
  return hipGetLastError() == hipSuccess ? 0 : -1000;
}

extern "C" int ncchip_row_prefix(const void* src, int kind, int pitch, int w, int h, const uint8_t* dc_src, int shift,
                                 uint64_t* p1, uint64_t* p2, uint32_t* nz, uint32_t* support, void* stream) {
  hipStream_t st = as_stream(stream);
  if (kind == NCC_U8) hipLaunchKernelGGL((k_row_prefix<uint8_t>), dim3(h), dim3(256), 0, st, static_cast<const uint8_t*>(src), pitch, w, h, dc_src, shift, p1, p2, nz);
  else if (kind == NCC_U16) hipLaunchKernelGGL((k_row_prefix<uint16_t>), dim3(h), dim3(256), 0, st, static_cast<const uint16_t*>(src), pitch, w, h, dc_src, shift, p1, p2, nz);
  else hipLaunchKernelGGL((k_row_prefix<uint32_t>), dim3(h), dim3(256), 0, st, static_cast<const uint32_t*>(src), pitch, w, h, dc_src, shift, p1, p2, nz);
  if (support) hipLaunchKernelGGL(k_support_cols, dim3((w + 1 + 255) / 256), dim3(256), 0, st, nz, w, h, support);
  return hipGetLastError() == hipSuccess ? 0 : -1000;
}

// ------------------------------------------------------------------------
// Row prefixes without allocation.  The support indicator's row prefixes
// (the `nz` workspace of ncchip_row_prefix) live in a file-scope device
// array: one copy per GPU, held until the process ends.  Host code never
// uses the array's name as an address; the address of each device's copy
// comes from hipGetSymbolAddress, taken once per device with that device
// current.
namespace {
constexpr size_t kNoAllocWords = size_t(1) << 22;  // 16 MiB: (w + 1) * h <= 4194304
constexpr int kMaxDevices = 64;
__device__ __align__(256) uint32_t scratch[kNoAllocWords];

std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[kMaxDevices] = {};

// The current device's copy of `scratch` (null when the device cannot be
// queried or the symbol has no address on it).
uint32_t* device_scratch() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return nullptr;
    g_scratch_ptr[dev] = static_cast<uint32_t*>(ptr);
  }
  return g_scratch_ptr[dev];
}
}  // namespace

extern "C" int ncchip_row_prefix_noalloc(const void* src, int kind, int pitch, int w, int h, const uint8_t* dc_src,
                                         int shift, uint64_t* p1, uint64_t* p2, uint32_t* support, void* stream) {
  if (w <= 0 || h <= 0 || !src || !dc_src || !p1 || !p2) return -1;
  if (size_t(w + 1) * size_t(h) > kNoAllocWords) return -103;  // beyond the static workspace
  uint32_t* nz = device_scratch();
  if (!nz) return -1000;
  // every (w + 1) * h word of nz is written by the row pass before the
  // column pass reads it, so nothing of an earlier call needs clearing
  return ncchip_row_prefix(src, kind, pitch, w, h, dc_src, shift, p1, p2, nz, support, stream);
}
