// Model construction on the device: the rotated, zero-mean, unit-energy
// rasters of every (level, angle) of a model (normalize / rotate,
// model.cpp:39-98, built level-major by rebuild, model.cpp:173-242).  One
// workgroup per raster.  Every FP64 operation is the host builder's, in its
// order (compiled without contraction), so the rasters are bit-identical to
// the host's:
//  * sampling: row-major over the raster's grid, inverse bilinear rotation
//    about the level centroid with the four-neighbour mask rule; samples are
//    compacted in grid order;
//  * normalisation: the incremental mean and the energy are sequential sums
//    in sample order (one lane), the scaled weights are independent;
//  * row runs and runs of equal weights, in sample order.
// The host counts first (to place each raster's samples), then samples and
// normalises, then places the runs.
#include "ncc_common.hip.h"
#include "../../include/ncchip.h"

namespace ncc {

constexpr int kCreateThreads = 256;

__device__ __forceinline__ int32_t pack_rc(int32_t r, int32_t c) {
  return int32_t((uint32_t(r) << 16) | (uint32_t(c) & 0xffffu));
}

// Grid pixel k of a raster: whether it is a sample, its packed coordinates
// and its value.
__device__ __forceinline__ bool grid_sample(const NccTemplateLevel& L, const NccRasterJob& J, int64_t k,
                                            int32_t& rc, double& val) {
  const int64_t cols = int64_t(J.c_hi) - J.c_lo + 1;
  const int r = J.r_lo + int(k / cols), c = J.c_lo + int(k % cols);
  if (J.unrotated) {
    const size_t i = size_t(r) * size_t(L.w) + size_t(c);
    if (!L.mask[i]) return false;
    rc = pack_rc(r, c);
    val = L.gray[i];
    return true;
  }
  const double ca = J.ca, sa = J.sa;
  const double ty = ca * r + sa * c + L.orow, tx = -sa * r + ca * c + L.ocol;
  const double ymax = double(L.h - 1), xmax = double(L.w - 1);
  if (ty < 0 || tx < 0 || ty > ymax || tx > xmax) return false;
  const size_t y0 = size_t(ty), x0 = size_t(tx);
  const size_t y1 = min(y0 + 1, size_t(L.h - 1)), x1 = min(x0 + 1, size_t(L.w - 1));
  const size_t w = size_t(L.w);
  const size_t i00 = y0 * w + x0, i01 = y0 * w + x1, i10 = y1 * w + x0, i11 = y1 * w + x1;
  if (!L.mask[i00] || !L.mask[i01] || !L.mask[i10] || !L.mask[i11]) return false;
  const double fy = ty - double(y0), fx = tx - double(x0);
  const double top = (1 - fx) * L.gray[i00] + fx * L.gray[i01];
  const double bottom = (1 - fx) * L.gray[i10] + fx * L.gray[i11];
  val = (1 - fy) * top + fy * bottom;
  rc = pack_rc(r, c);
  return true;
}

__device__ __forceinline__ int64_t grid_size(const NccRasterJob& J) {
  return (int64_t(J.r_hi) - J.r_lo + 1) * (int64_t(J.c_hi) - J.c_lo + 1);
}

// Block sum of one value per thread (every thread gets the total).
template <class T>
__device__ __forceinline__ T block_sum(T v, T* red) {
  for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
  const int wave = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[wave] = v;
  __syncthreads();
  T t = 0;
  for (int k = 0; k < kCreateThreads / 64; ++k) t += red[k];
  return t;
}
__device__ __forceinline__ int block_min(int v, int* red) {
  for (int d = 32; d > 0; d >>= 1) v = min(v, __shfl_xor(v, d, 64));
  const int wave = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[wave] = v;
  __syncthreads();
  int t = red[0];
  for (int k = 1; k < kCreateThreads / 64; ++k) t = min(t, red[k]);
  return t;
}

__device__ __forceinline__ int block_max(int v, int* red) {
  for (int d = 32; d > 0; d >>= 1) v = max(v, __shfl_xor(v, d, 64));
  const int wave = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[wave] = v;
  __syncthreads();
  int t = red[0];
  for (int k = 1; k < kCreateThreads / 64; ++k) t = max(t, red[k]);
  return t;
}

struct CreateArgs {
  const NccTemplateLevel* levels;
  const NccRasterJob* jobs;
  NccRasterStats* stats;
  int32_t* rc;
  double* v;        // sampled values
  double* w;        // weights
  float* wf;
  int32_t* runs;    // 4 ints per run
  int32_t* wruns;
  double* wrun_w;
  int32_t* run_start;   // sample index of each run's first sample
  int32_t* wrun_start;
};

__global__ void __launch_bounds__(kCreateThreads) k_raster_count(const CreateArgs a) {
  __shared__ int64_t red[kCreateThreads / 64];
  const NccRasterJob J = a.jobs[blockIdx.x];
  const NccTemplateLevel L = a.levels[J.level];
  const int64_t n = grid_size(J);
  int64_t cnt = 0;
  for (int64_t k = threadIdx.x; k < n; k += kCreateThreads) {
    int32_t rc;
    double val;
    cnt += grid_sample(L, J, k, rc, val) ? 1 : 0;
  }
  cnt = block_sum(cnt, red);
  if (threadIdx.x == 0) {
    NccRasterStats st{};
    st.count = cnt;
    a.stats[blockIdx.x] = st;
  }
}

// Ordered compaction of the samples to jobs[b].sample_off.
__global__ void __launch_bounds__(kCreateThreads) k_raster_samples(const CreateArgs a) {
  __shared__ int wave_tot[kCreateThreads / 64];
  const NccRasterJob J = a.jobs[blockIdx.x];
  const NccTemplateLevel L = a.levels[J.level];
  const int64_t n = grid_size(J);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int64_t base = J.sample_off;
  for (int64_t k0 = 0; k0 < n; k0 += kCreateThreads) {
    const int64_t k = k0 + threadIdx.x;
    int32_t rc = 0;
    double val = 0;
    const bool ok = k < n && grid_sample(L, J, k, rc, val);
    const uint64_t m = __ballot(ok);
    const int pre = __popcll(m & ((uint64_t(1) << lane) - 1));
    if (lane == 0) wave_tot[wave] = __popcll(m);
    __syncthreads();
    int off = 0, tot = 0;
    for (int q = 0; q < kCreateThreads / 64; ++q) {
      if (q < wave) off += wave_tot[q];
      tot += wave_tot[q];
    }
    if (ok) {
      a.rc[base + off + pre] = rc;
      a.v[base + off + pre] = val;
    }
    base += tot;
    __syncthreads();
  }
}

__device__ __forceinline__ int rc_row(int32_t rc) { return rc >> 16; }
__device__ __forceinline__ int rc_col(int32_t rc) { return int(int16_t(rc & 0xffff)); }

// normalize (model.cpp:39-65): centre by the mean, scale to unit energy;
// bounds, weight sums and the numbers of row runs and weight runs.
__global__ void __launch_bounds__(kCreateThreads) k_raster_normalize(const CreateArgs a) {
  __shared__ double s_mean, s_energy;
  __shared__ int64_t red64[kCreateThreads / 64];
  __shared__ int red32[kCreateThreads / 64];
  const NccRasterJob J = a.jobs[blockIdx.x];
  NccRasterStats st = a.stats[blockIdx.x];
  const int64_t n = st.count;
  const double* v = a.v + J.sample_off;
  const int32_t* rc = a.rc + J.sample_off;
  if (n < 4) {
    if (threadIdx.x == 0) { st.status = 1; a.stats[blockIdx.x] = st; }
    return;
  }
  if (threadIdx.x == 0) {
    double mean = 0, energy = 0;
    if (J.prepared) {
      mean = J.mean;
      energy = J.energy;
    } else {
      // sequential sums in sample order; the loads of a chunk issue together
      constexpr int U = 8;
      int64_t i = 0;
      for (; i + U <= n; i += U) {
        double x[U];
#pragma unroll
        for (int u = 0; u < U; ++u) x[u] = v[i + u];
#pragma unroll
        for (int u = 0; u < U; ++u) mean += (x[u] - mean) / double(i + u + 1);
      }
      for (; i < n; ++i) mean += (v[i] - mean) / double(i + 1);
      for (i = 0; i + U <= n; i += U) {
        double x[U];
#pragma unroll
        for (int u = 0; u < U; ++u) x[u] = v[i + u];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const double d = x[u] - mean;
          energy += d * d;
        }
      }
      for (; i < n; ++i) {
        const double d = v[i] - mean;
        energy += d * d;
      }
    }
    s_mean = mean;
    s_energy = energy;
  }
  __syncthreads();
  const double mean = s_mean, energy = s_energy;
  if (!(energy > 0) || !isfinite(energy)) {
    if (threadIdx.x == 0) { st.status = 1; a.stats[blockIdx.x] = st; }
    return;
  }
  const double scale = 1 / sqrt(energy);
  int rmin = INT_MAX, rmax = INT_MIN, cmin = INT_MAX, cmax = INT_MIN;
  int64_t nr = 0, nw = 0;
  for (int64_t i = threadIdx.x; i < n; i += kCreateThreads) {
    const double wi = (v[i] - mean) * scale;
    a.w[J.sample_off + i] = wi;
    a.wf[J.sample_off + i] = float(wi);
    const int r = rc_row(rc[i]), c = rc_col(rc[i]);
    rmin = min(rmin, r); rmax = max(rmax, r); cmin = min(cmin, c); cmax = max(cmax, c);
    bool run_on = false, wrun_on = false;
    if (i > 0) {
      const int rp = rc_row(rc[i - 1]), cp = rc_col(rc[i - 1]);
      run_on = rp == r && cp + 1 == c;
      wrun_on = run_on && (v[i - 1] - mean) * scale == wi;
    }
    nr += run_on ? 0 : 1;
    nw += wrun_on ? 0 : 1;
  }
  rmin = block_min(rmin, red32);
  cmin = block_min(cmin, red32);
  rmax = block_max(rmax, red32);
  cmax = block_max(cmax, red32);
  nr = block_sum(nr, red64);
  nw = block_sum(nw, red64);
  if (threadIdx.x == 0) {
    double wsum = 0, wsum_f = 0;
    constexpr int U = 8;
    int64_t i = 0;
    for (; i + U <= n; i += U) {
      double x[U];
#pragma unroll
      for (int u = 0; u < U; ++u) x[u] = v[i + u];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const double wi = (x[u] - mean) * scale;
        wsum += wi;
        wsum_f += double(float(wi));
      }
    }
    for (; i < n; ++i) {
      const double wi = (v[i] - mean) * scale;
      wsum += wi;
      wsum_f += double(float(wi));
    }
    st.status = 0;
    st.rmin = rmin; st.rmax = rmax; st.cmin = cmin; st.cmax = cmax;
    st.n_runs = int32_t(nr);
    st.n_wruns = int32_t(nw);
    st.wsum = wsum;
    st.wsum_f = wsum_f;
    a.stats[blockIdx.x] = st;
  }
}

// Row runs and weight runs of one raster to jobs[b].run_off / wrun_off:
// ordered compaction of the run starts, then each run's end from the next
// run's start.
__global__ void __launch_bounds__(kCreateThreads) k_raster_runs(const CreateArgs a) {
  __shared__ int wave_tot[2][kCreateThreads / 64];
  const NccRasterJob J = a.jobs[blockIdx.x];
  const NccRasterStats st = a.stats[blockIdx.x];
  const int64_t n = st.count;
  const int32_t* rc = a.rc + J.sample_off;
  const double* w = a.w + J.sample_off;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int64_t base_r = 0, base_w = 0;
  for (int64_t i0 = 0; i0 < n; i0 += kCreateThreads) {
    const int64_t i = i0 + threadIdx.x;
    bool start_r = false, start_w = false;
    if (i < n) {
      start_r = start_w = true;
      if (i > 0) {
        const bool run_on = rc_row(rc[i - 1]) == rc_row(rc[i]) && rc_col(rc[i - 1]) + 1 == rc_col(rc[i]);
        start_r = !run_on;
        start_w = !(run_on && w[i - 1] == w[i]);
      }
    }
    const uint64_t mr = __ballot(start_r), mw = __ballot(start_w);
    const uint64_t lt = (uint64_t(1) << lane) - 1;
    if (lane == 0) { wave_tot[0][wave] = __popcll(mr); wave_tot[1][wave] = __popcll(mw); }
    __syncthreads();
    int off_r = 0, off_w = 0, tot_r = 0, tot_w = 0;
    for (int q = 0; q < kCreateThreads / 64; ++q) {
      if (q < wave) { off_r += wave_tot[0][q]; off_w += wave_tot[1][q]; }
      tot_r += wave_tot[0][q];
      tot_w += wave_tot[1][q];
    }
    if (start_r) a.run_start[J.run_off + base_r + off_r + __popcll(mr & lt)] = int32_t(i);
    if (start_w) a.wrun_start[J.wrun_off + base_w + off_w + __popcll(mw & lt)] = int32_t(i);
    base_r += tot_r;
    base_w += tot_w;
    __syncthreads();
  }
  for (int64_t k = threadIdx.x; k < st.n_runs; k += kCreateThreads) {
    const int64_t s = a.run_start[J.run_off + k];
    const int64_t e = (k + 1 < st.n_runs ? int64_t(a.run_start[J.run_off + k + 1]) : n) - 1;
    int32_t* out = a.runs + 4 * (J.run_off + k);
    out[0] = rc_row(rc[s]); out[1] = rc_col(rc[s]); out[2] = rc_col(rc[e]); out[3] = 0;
  }
  for (int64_t k = threadIdx.x; k < st.n_wruns; k += kCreateThreads) {
    const int64_t s = a.wrun_start[J.wrun_off + k];
    const int64_t e = (k + 1 < st.n_wruns ? int64_t(a.wrun_start[J.wrun_off + k + 1]) : n) - 1;
    int32_t* out = a.wruns + 4 * (J.wrun_off + k);
    out[0] = rc_row(rc[s]); out[1] = rc_col(rc[s]); out[2] = rc_col(rc[e]); out[3] = 0;
    a.wrun_w[J.wrun_off + k] = w[s];
  }
}

}  // namespace ncc

using namespace ncc;

extern "C" int ncchip_create_rasters(int stage, const NcchipCreateBuffers* b, int n_jobs, void* stream) {
  if (n_jobs <= 0) return 0;
  CreateArgs a;
  a.levels = b->levels; a.jobs = b->jobs; a.stats = b->stats; a.rc = b->rc; a.v = b->v; a.w = b->w; a.wf = b->wf;
  a.runs = b->runs; a.wruns = b->wruns; a.wrun_w = b->wrun_w; a.run_start = b->run_start;
  a.wrun_start = b->wrun_start;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 grid{unsigned(n_jobs), 1, 1}, block{unsigned(kCreateThreads), 1, 1};
  switch (stage) {
    case NCCHIP_CREATE_COUNT: hipLaunchKernelGGL(k_raster_count, grid, block, 0, st, a); break;
    case NCCHIP_CREATE_SAMPLES:
      hipLaunchKernelGGL(k_raster_samples, grid, block, 0, st, a);
      hipLaunchKernelGGL(k_raster_normalize, grid, block, 0, st, a);
      break;
    case NCCHIP_CREATE_RUNS: hipLaunchKernelGGL(k_raster_runs, grid, block, 0, st, a); break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -1000;
}
