// Coarse-to-fine tracking (matcher.cpp:630-761), continuous bilinear scoring
// (matcher.cpp:121-157, kernels_avx2.cpp:327-477) and Gauss-Newton subpixel /
// angle refinement with its level-0 recovery search (matcher.cpp:170-240,
// 853-1061), all in FP64 and compiled with -ffp-contract=off so each sample
// coordinate, bilinear value and window moment is the reference's value.
// Sums over samples are block tree reductions (the reference sums in 4 lanes),
// so reduced totals may differ from the reference in the last bits.
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

__device__ __forceinline__ PatchRef patch_ref(const NccPatchPool* pools, int32_t ref) {
  const NccPatchPool& pool = pools[ref >> 24];
  PatchRef out;
  out.P = pool.patches[ref & 0xffffff];
  out.rc = pool.rc + out.P.sample_off;
  out.w = pool.w + out.P.sample_off;
  out.wruns = pool.wruns + 4 * out.P.wrun_off;
  out.wrun_w = pool.wrun_w + out.P.wrun_off;
  return out;
}
__device__ __forceinline__ double ref_angle(const NccPatchPool* pools, int32_t ref) {
  return pools[ref >> 24].patches[ref & 0xffffff].angle;
}

// ------------------------------------------------------------------ tracking
struct TrackArgs {
  NccLevel lv;
  double dc;              // filled in-kernel from dc_src
  int64_t D;
  const uint8_t* dc_src;
  NccPatchPool pools[2];
  const int32_t* refs;
  int n_refs;
  const NccMatch* cand;
  NccMatch* out;
  int n_cand;
  double tol, crop_r, crop_c, step, offset;
  int sr, sc, sh, sw;
  int exact_lattice, absolute;
};

// One window of the reference's 8/4-lane prefix kernels (kernels_avx2.cpp:
// 145-238): exact moments, dot in sample order or by weight runs.
__device__ double track_window(const TrackArgs& a, const PatchRef& p, int row, int col) {
  const NccPatch& P = p.P;
  const int H = a.lv.height, W = a.lv.width;
  if (row + P.rmin < 0 || col + P.cmin < 0 || row + P.rmax >= H || col + P.cmax >= W) return NCC_MISSING;
  if (row + P.rmin < a.sr || col + P.cmin < a.sc || row + P.rmax >= a.sr + a.sh || col + P.cmax >= a.sc + a.sw)
    return NCC_MISSING;
  int64_t s1 = 0, s2 = 0;
  double dot = 0;
  for (int i = 0; i < P.n; ++i) {
    const int32_t rc = p.rc[i];
    const uint32_t s = load_s(a.lv, row + sample_row(rc), col + sample_col(rc));
    s1 += s;
    s2 += int64_t(s) * s;
    if (!P.use_wruns) dot += (ldexp(double(s), -a.lv.shift) - a.dc) * p.w[i];
  }
  if (P.use_wruns) {
    for (int k = 0; k < P.n_wruns; ++k) {
      const int4 run = reinterpret_cast<const int4*>(p.wruns)[k];
      int64_t rs = 0;
      for (int c = run.y; c <= run.z; ++c) rs += load_s(a.lv, row + run.x, col + c);
      const double run_q = ldexp(double(rs - int64_t(run.z - run.y + 1) * a.D), -a.lv.shift);
      dot += run_q * p.wrun_w[k];
    }
  }
  double sum, sq;
  exact_moments(s1, s2, P.n, a.D, a.lv.shift, sum, sq);
  const double cnt = double(P.n);
  const double var = sq - sum * sum / cnt;
  dot = dot - sum / cnt * P.wsum;
  if (!(var > 128 * __DBL_EPSILON__ * sq)) {
    const int64_t t1 = s1 - int64_t(P.n) * a.D;
    const int64_t t2 = s2 - 2 * a.D * s1 + int64_t(P.n) * a.D * a.D;
    if (a.exact_lattice || (__int128)t2 * P.n == (__int128)t1 * t1) return NCC_MISSING;
    const double v = ldexp(double((__int128)t2 * P.n - (__int128)t1 * t1), -2 * a.lv.shift) / cnt;
    const double s = clamp1(dot / sqrt(v));
    return a.absolute ? fabs(s) : s;
  }
  const double s = clamp1(dot / sqrt(var));
  return a.absolute ? fabs(s) : s;
}

constexpr int kMaxEligible = 1024;

// Ordered compaction of the patch indices whose angle is within `tol` of
// `angle` (query order kept: it defines the reference's tie-breaking).
__device__ int eligible_patches(const NccPatchPool* pools, const int32_t* refs, int n_refs, double angle, double tol,
                                int* elig, int* wave_cnt) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int n_e = 0;
  for (int base = 0; base < n_refs; base += 256) {
    const int k = base + threadIdx.x;
    const bool f = k < n_refs && fabs(ref_angle(pools, refs[k]) - angle) <= tol;
    const uint64_t m = __ballot(f);
    const int pre = __popcll(m & ((uint64_t(1) << lane) - 1));
    if (lane == 0) wave_cnt[wave] = __popcll(m);
    __syncthreads();
    int off = n_e;
    for (int w = 0; w < wave; ++w) off += wave_cnt[w];
    if (f && off + pre < kMaxEligible) elig[off + pre] = k;
    n_e += wave_cnt[0] + wave_cnt[1] + wave_cnt[2] + wave_cnt[3];
    __syncthreads();
  }
  return min(n_e, kMaxEligible);
}

// One block per candidate: score the 5x5 anchors around the candidate for every
// query patch within the angle window and keep the first strict maximum in the
// reference's (patch, dr, dc) order (matcher.cpp:670-756).
__global__ void __launch_bounds__(256) k_track(TrackArgs a) {
  a.dc = double(*a.dc_src);
  a.D = int64_t(*a.dc_src) << a.lv.shift;
  __shared__ int elig[kMaxEligible];
  __shared__ int wave_cnt[4];
  __shared__ double red_s[256];
  __shared__ int red_i[256];
  const int ci = blockIdx.x;
  if (ci >= a.n_cand) return;
  const NccMatch cd = a.cand[ci];
  const int n_e = eligible_patches(a.pools, a.refs, a.n_refs, cd.angle, a.tol, elig, wave_cnt);
  const int n_win = n_e * 25;
  double best = NCC_MISSING;
  int best_i = 0x7fffffff;
  for (int wi = threadIdx.x; wi < n_win; wi += 256) {
    const PatchRef p = patch_ref(a.pools, a.refs[elig[wi / 25]]);
    const int dr = (wi % 25) / 5 - 2, dc = wi % 5 - 2;
    const int row = int(round((cd.row - a.crop_r - a.offset) / a.step - p.P.orow));
    const int col = int(round((cd.col - a.crop_c - a.offset) / a.step - p.P.ocol));
    const double s = track_window(a, p, row + dr, col + dc);
    if (s > best) { best = s; best_i = wi; }
  }
  red_s[threadIdx.x] = best;
  red_i[threadIdx.x] = best_i;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) {
      const double o = red_s[threadIdx.x + h];
      const int oi = red_i[threadIdx.x + h];
      if (o > red_s[threadIdx.x] || (o == red_s[threadIdx.x] && oi < red_i[threadIdx.x])) {
        red_s[threadIdx.x] = o;
        red_i[threadIdx.x] = oi;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    NccMatch res = cd;
    res.score = NCC_MISSING;
    if (red_s[0] > NCC_MISSING) {
      const int wi = red_i[0];
      const PatchRef p = patch_ref(a.pools, a.refs[elig[wi / 25]]);
      const int dr = (wi % 25) / 5 - 2, dc = wi % 5 - 2;
      const int row = int(round((cd.row - a.crop_r - a.offset) / a.step - p.P.orow));
      const int col = int(round((cd.col - a.crop_c - a.offset) / a.step - p.P.ocol));
      res.row = a.crop_r + a.step * (row + dr + p.P.orow) + a.offset;
      res.col = a.crop_c + a.step * (col + dc + p.P.ocol) + a.offset;
      res.angle = p.P.angle;
      res.score = red_s[0];
    }
    a.out[ci] = res;
  }
}

// ------------------------------------------------------- continuous / refine
struct RefineArgs {
  NccLevel lv;            // finest search level
  double dc;              // filled in-kernel from dc_src
  const uint8_t* dc_src;
  int sr, sc, sh, sw;     // storage rectangle at the finest level
  double crop_r, crop_c, scale, phase;
  int finest, top;
  NccPatchPool pools[2];
  int32_t orig_ref;       // unrotated patch of the finest level (model.original at level 0)
  double lv_orow, lv_ocol;
  const int32_t* refs0;   // level-0 query patches (recovery search)
  int n_refs0;
  double step, a_start, a_end, min_score;
  int subpixel, rot, absolute;
  int img_h, img_w;       // search domain (full image)
  int H0, W0;             // level-0 dims (crop)
  const NccMatch* cand;
  int n_cand;
  NccMatch* out_full;
  NccMatch* out_refined;
  int32_t* verdict;
};

struct Shared {
  double red[4 * 18];
  int flag;
  double bcast[8];
};

__device__ __forceinline__ bool domain_contains(const RefineArgs& a, double row, double col) {
  const double r = floor(row + .5), c = floor(col + .5);
  return r >= 0 && c >= 0 && r < double(a.img_h) && c < double(a.img_w);
}

// Value at storage-local integer coordinates.
__device__ __forceinline__ double pix(const RefineArgs& a, int r, int c) {
  return load_q(a.lv, r + a.sr, c + a.sc, a.dc);
}

// strictly_inside: kernels_avx2.cpp:261-277
__device__ bool strictly_inside(const RefineArgs& a, const NccPatch& P, double pr, double pc, double ca, double sa) {
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

// Bilinear sample (and gradients) at storage-local (y, x) known to be inside.
__device__ __forceinline__ void bilinear_at(const RefineArgs& a, double y, double x, bool interior, double& q,
                                            double& gy, double& gx) {
  const double yr = floor(y), xc = floor(x), fy = y - yr, fx = x - xc;
  const int r = int(yr), c = int(xc);
  const int r1 = interior ? r + 1 : min(r + 1, a.sh - 1), c1 = interior ? c + 1 : min(c + 1, a.sw - 1);
  const double va = pix(a, r, c), vb = pix(a, r, c1), vd = pix(a, r1, c), ve = pix(a, r1, c1);
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
// gg00, gg10, gg11, gg20, gg21, gg22.  Returns false if any sample is outside.
__device__ bool bilinear_pass(const RefineArgs& a, const PatchRef& O, double pr, double pc, double angle, bool deriv,
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
    bilinear_at(a, y, x, false, off, gy, gx);
  }
  double v[18];
#pragma unroll
  for (int k = 0; k < 18; ++k) v[k] = 0;
  int bad = 0;
  for (int i = threadIdx.x; i < P.n; i += 256) {
    const int32_t rc = O.rc[i];
    const double dr = sample_row(rc) - a.lv_orow, dc = sample_col(rc) - a.lv_ocol;
    double y, x;
    if (aligned) { y = pr + dr; x = pc + dc; }
    else { y = (pr + ca * dr) - sa * dc; x = (pc + sa * dr) + ca * dc; }
    if (!interior && !inside(a, y, x, deriv)) { bad = 1; continue; }
    double q, gy, gx;
    bilinear_at(a, y, x, interior, q, gy, gx);
    q = q - off;
    const double w = O.w[i];
    v[0] += q; v[1] += q * q; v[2] += w * q;
    if (deriv) {
      const double g[3] = {gy, gx, gy * ((0 - sa * dr) - ca * dc) + gx * (ca * dr - sa * dc)};
#pragma unroll
      for (int j = 0; j < 3; ++j) { v[3 + j] += g[j]; v[6 + j] += q * g[j]; v[9 + j] += w * g[j]; }
      v[12] += g[0] * g[0]; v[13] += g[1] * g[0]; v[14] += g[1] * g[1];
      v[15] += g[2] * g[0]; v[16] += g[2] * g[1]; v[17] += g[2] * g[2];
    }
  }
  if (block_any(bad, &sm.flag)) return false;
  if (deriv) block_sum<18>(v, sm.red);
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
__device__ double continuous(const RefineArgs& a, const PatchRef& O, double row, double col, double angle, Shared& sm) {
  double pr, pc, m[18];
  to_local(a, row, col, pr, pc);
  if (!bilinear_pass(a, O, pr, pc, angle, false, m, sm)) return NCC_MISSING;
  const double var = m[1] - m[0] * m[0] / double(O.P.n);
  if (!(var > 0)) return NCC_MISSING;
  const double s = clamp1(m[2] / sqrt(var));
  return a.absolute ? fabs(s) : s;
}

// refinement_step(): matcher.cpp:170-240.  Returns the verdict; writes delta
// and (when out_score) the clamped score.
__device__ bool refinement_step(const RefineArgs& a, const PatchRef& O, double row, double col, double angle,
                                double* delta, double* out_score, const double* min_score, Shared& sm) {
  double pr, pc, m[18];
  to_local(a, row, col, pr, pc);
  if (!bilinear_pass(a, O, pr, pc, angle, true, m, sm)) return false;
  bool ok = false;
  if (threadIdx.x == 0) {
    const double n = double(O.P.n);
    const double sum = m[0], dot = m[2];
    double energy = m[1];
    double gs[3] = {m[3], m[4], m[5]}, qg[3] = {m[6], m[7], m[8]}, wg[3] = {m[9], m[10], m[11]};
    double gg[3][3] = {{m[12], 0, 0}, {m[13], m[14], 0}, {m[15], m[16], m[17]}};
    double sc_out = NCC_MISSING;
    energy -= sum * sum / n;
    int st = 0;
    double d[3] = {0, 0, 0};
    if (energy > 0) {
      const double sign = a.absolute && dot < 0 ? -1 : 1, score = sign * dot / sqrt(energy);
      sc_out = clamp1(score);
      st = 1;
      if (min_score && clamp1(score) <= *min_score) st = 0;
      if (st) {
        const int axes = a.rot ? 3 : 2;
        double M[3][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
        for (int i = 0; i < axes; ++i) qg[i] -= sum * gs[i] / n;
        for (int i = 0; i < axes; ++i) {
          for (int j = 0; j < axes; ++j) {
            const int hi = i > j ? i : j, lo = i > j ? j : i;
            M[i][j] = (gg[hi][lo] - gs[i] * gs[j] / n - qg[i] * qg[j] / energy) / energy;
          }
          M[i][axes] = sign * wg[i] / sqrt(energy) - score * qg[i] / energy;
        }
        for (int i = 0; i < axes && st; ++i) {
          int pv = i;
          for (int j = i + 1; j < axes; ++j) if (fabs(M[j][i]) > fabs(M[pv][i])) pv = j;
          if (fabs(M[pv][i]) < 1e-14) { st = 0; break; }
          for (int k = i; k <= axes; ++k) { const double t = M[i][k]; M[i][k] = M[pv][k]; M[pv][k] = t; }
          const double dv = M[i][i];
          for (int k = i; k <= axes; ++k) M[i][k] /= dv;
          for (int j = 0; j < axes; ++j)
            if (j != i) { const double f = M[j][i]; for (int k = i; k <= axes; ++k) M[j][k] -= f * M[i][k]; }
        }
        if (st) {
          for (int i = 0; i < axes; ++i) d[i] = M[i][axes];
          if (!a.rot) d[2] = 0;
          st = isfinite(d[0]) && isfinite(d[1]) && isfinite(d[2]);
        }
      }
    }
    sm.bcast[0] = st; sm.bcast[1] = d[0]; sm.bcast[2] = d[1]; sm.bcast[3] = d[2]; sm.bcast[4] = sc_out;
    sm.bcast[5] = energy > 0 ? 1 : 0;
  }
  __syncthreads();
  ok = sm.bcast[0] != 0;
  if (sm.bcast[5] != 0 && out_score) *out_score = sm.bcast[4];
  delta[0] = sm.bcast[1]; delta[1] = sm.bcast[2]; delta[2] = sm.bcast[3];
  __syncthreads();
  return ok;
}

__device__ __forceinline__ double clampd(double v, double lo, double hi) { return v < lo ? lo : (hi < v ? hi : v); }

// fit(): matcher.cpp:903-982 without the multi-model memo or guard.
__device__ void fit(const RefineArgs& a, const PatchRef& O, NccMatch& c, Shared& sm) {
  if (!a.subpixel) return;
  bool cached = false;
  double next_delta[3] = {0, 0, 0};
  for (int it = 0; it < 6; ++it) {
    double delta[3] = {0, 0, 0};
    if (cached) { delta[0] = next_delta[0]; delta[1] = next_delta[1]; delta[2] = next_delta[2]; }
    else if (!refinement_step(a, O, c.row, c.col, c.angle, delta, nullptr, nullptr, sm)) break;
    cached = false;
    const double factor = fmin(fmin(1., .75 / fmax(1e-20, fabs(delta[0]))),
                               fmin(.75 / fmax(1e-20, fabs(delta[1])), a.rot ? a.step / fmax(1e-20, fabs(delta[2])) : 1.));
    bool accepted = false;
    for (int bt = 0; bt < 4; ++bt) {
      const double st = ldexp(factor, -bt);
      NccMatch im = c;
      im.row += a.scale * st * delta[0];
      im.col += a.scale * st * delta[1];
      im.angle = clampd(im.angle + st * delta[2], a.a_start, a.a_end);
      if (!domain_contains(a, im.row, im.col)) continue;
      bool next_valid = false;
      if (bt == 0 && it + 1 < 6) {
        im.score = NCC_MISSING;
        const double floor_score = c.score;
        next_valid = refinement_step(a, O, im.row, im.col, im.angle, next_delta, &im.score, &floor_score, sm);
        if (!isfinite(im.score)) im.score = continuous(a, O, im.row, im.col, im.angle, sm);
      } else {
        im.score = continuous(a, O, im.row, im.col, im.angle, sm);
      }
      if (im.score > c.score) { c = im; accepted = true; cached = next_valid; break; }
    }
    if (!accepted || (hypot(delta[0], delta[1]) < 1e-4 && fabs(delta[2]) < 1e-5)) break;
  }
}

// refine_coarse_candidate(): matcher.cpp:853-877 (lowest level above 0).
__device__ void fit_coarse(const RefineArgs& a, const PatchRef& O, NccMatch& c, Shared& sm) {
  if (!a.subpixel) return;
  for (int it = 0; it < 6; ++it) {
    double delta[3] = {0, 0, 0};
    if (!refinement_step(a, O, c.row, c.col, c.angle, delta, nullptr, nullptr, sm)) break;
    const double factor = fmin(fmin(1., .75 / fmax(1e-20, fabs(delta[0]))),
                               fmin(.75 / fmax(1e-20, fabs(delta[1])), a.rot ? a.step / fmax(1e-20, fabs(delta[2])) : 1.));
    bool accepted = false;
    for (int bt = 0; bt < 4; ++bt) {
      const double st = ldexp(factor, -bt);
      NccMatch nx = c;
      nx.row += a.scale * st * delta[0];
      nx.col += a.scale * st * delta[1];
      nx.angle = clampd(nx.angle + st * delta[2], a.a_start, a.a_end);
      if (!domain_contains(a, nx.row, nx.col)) continue;
      nx.score = continuous(a, O, nx.row, nx.col, nx.angle, sm);
      if (nx.score > c.score) { c = nx; accepted = true; break; }
    }
    if (!accepted) break;
  }
}

// Direct window NCC at level 0 (correlation_four_avx2 / correlation_avx2):
// per-window offset at the first sample, block-wide sums.
__device__ double direct_window(const RefineArgs& a, const PatchRef& p, int row, int col, Shared& sm) {
  const NccPatch& P = p.P;
  if (row + P.rmin < 0 || col + P.cmin < 0 || row + P.rmax >= a.H0 || col + P.cmax >= a.W0) return NCC_MISSING;
  if (row + P.rmin < a.sr || col + P.cmin < a.sc || row + P.rmax >= a.sr + a.sh || col + P.cmax >= a.sc + a.sw)
    return NCC_MISSING;
  const double off = load_q(a.lv, row + sample_row(p.rc[0]), col + sample_col(p.rc[0]), a.dc);
  double v[3] = {0, 0, 0};
  for (int i = threadIdx.x; i < P.n; i += 256) {
    const int32_t rc = p.rc[i];
    const double q = load_q(a.lv, row + sample_row(rc), col + sample_col(rc), a.dc) - off;
    v[0] += q; v[1] += q * q; v[2] += q * p.w[i];
  }
  block_sum<3>(v, sm.red);
  const double var = v[1] - v[0] * v[0] / double(P.n);
  if (!(var > 0)) return NCC_MISSING;
  const double s = clamp1(v[2] / sqrt(var));
  return a.absolute ? fabs(s) : s;
}

// One block per candidate: full continuous score (matcher.cpp:822), then the
// speculative refine_candidate (matcher.cpp:878-1061) from that score.
__global__ void __launch_bounds__(256) k_refine(RefineArgs a) {
  a.dc = double(*a.dc_src);
  __shared__ Shared sm;
  __shared__ int elig[kMaxEligible];
  __shared__ int wave_cnt[4];
  const int ci = blockIdx.x;
  if (ci >= a.n_cand) return;
  const PatchRef O = patch_ref(a.pools, a.orig_ref);
  NccMatch cd = a.cand[ci];
  if (isfinite(cd.score)) cd.score = continuous(a, O, cd.row, cd.col, cd.angle, sm);
  if (threadIdx.x == 0) a.out_full[ci] = cd;
  if (!isfinite(cd.score)) {
    if (threadIdx.x == 0) { a.out_refined[ci] = cd; a.verdict[ci] = 0; }
    return;
  }
  if (a.finest) {
    fit_coarse(a, O, cd, sm);
  } else {
    const NccMatch seed = cd;
    fit(a, O, cd, sm);
    if (a.subpixel && a.top > 0 && cd.score < a.min_score) {
      NccMatch best = cd;
      const int n_e = eligible_patches(a.pools, a.refs0, a.n_refs0, seed.angle, a.step * 1.01, elig, wave_cnt);
      for (int e = 0; e < n_e; ++e) {
        const PatchRef p = patch_ref(a.pools, a.refs0[elig[e]]);
        const int r = int(round(seed.row - a.crop_r - p.P.orow)), c = int(round(seed.col - a.crop_c - p.P.ocol));
        for (int dr = -2; dr <= 2; ++dr)
          for (int dc = -2; dc <= 2; ++dc) {
            NccMatch ps;
            ps.row = a.crop_r + r + dr + p.P.orow;
            ps.col = a.crop_c + c + dc + p.P.ocol;
            ps.angle = p.P.angle;
            if (!domain_contains(a, ps.row, ps.col)) continue;
            ps.score = direct_window(a, p, r + dr, c + dc, sm);
            if (ps.score > best.score) best = ps;
          }
      }
      if (best.row != cd.row || best.col != cd.col || best.angle != cd.angle) {
        best.score = continuous(a, O, best.row, best.col, best.angle, sm);
        fit(a, O, best, sm);
        if (best.score > cd.score) cd = best;
      }
    }
  }
  if (threadIdx.x == 0) {
    a.out_refined[ci] = cd;
    a.verdict[ci] = (cd.score >= a.min_score && domain_contains(a, cd.row, cd.col)) ? 1 : 0;
  }
}

}  // namespace ncc

using namespace ncc;

extern "C" int ncchip_track(const NccLevel* lv, const uint8_t* dc_src, const NccPatchPool* pools, const int32_t* refs,
                            int n_refs, const NccMatch* cand, NccMatch* out, int n_cand, double tol, double crop_r,
                            double crop_c, double step, double offset, const int32_t* storage, int exact_lattice,
                            int absolute, void* stream) {
  if (n_cand <= 0) return 0;
  TrackArgs a;
  a.lv = *lv; a.dc_src = dc_src; a.dc = 0; a.D = 0; a.pools[0] = pools[0]; a.pools[1] = pools[1]; a.refs = refs; a.n_refs = n_refs;
  a.cand = cand; a.out = out; a.n_cand = n_cand; a.tol = tol; a.crop_r = crop_r; a.crop_c = crop_c; a.step = step;
  a.offset = offset; a.sr = storage[0]; a.sc = storage[1]; a.sh = storage[2]; a.sw = storage[3];
  a.exact_lattice = exact_lattice; a.absolute = absolute;
  hipLaunchKernelGGL(k_track, dim3(n_cand), dim3(256), 0, static_cast<hipStream_t>(stream), a);
  return hipGetLastError() == hipSuccess ? 0 : -1000;
}

extern "C" int ncchip_refine(const NcchipRefineParams* p, const NccPatchPool* pools, const NccMatch* cand, int n_cand,
                             NccMatch* out_full, NccMatch* out_refined, int32_t* verdict, void* stream) {
  if (n_cand <= 0) return 0;
  RefineArgs a;
  a.lv = p->level; a.dc_src = p->dc_src; a.dc = 0; a.sr = p->storage[0]; a.sc = p->storage[1]; a.sh = p->storage[2];
  a.sw = p->storage[3]; a.crop_r = p->crop_r; a.crop_c = p->crop_c; a.scale = p->scale; a.phase = p->phase;
  a.finest = p->finest; a.top = p->top; a.pools[0] = pools[0]; a.pools[1] = pools[1]; a.orig_ref = p->orig_ref;
  a.lv_orow = p->level_orow; a.lv_ocol = p->level_ocol; a.refs0 = p->refs0; a.n_refs0 = p->n_refs0;
  a.step = p->step; a.a_start = p->angle_start; a.a_end = p->angle_end; a.min_score = p->min_score;
  a.subpixel = p->subpixel; a.rot = p->rotate; a.absolute = p->absolute; a.img_h = p->image_h; a.img_w = p->image_w;
  a.H0 = p->level.height; a.W0 = p->level.width;
  a.cand = cand; a.n_cand = n_cand; a.out_full = out_full; a.out_refined = out_refined; a.verdict = verdict;
  hipLaunchKernelGGL(k_refine, dim3(n_cand), dim3(256), 0, static_cast<hipStream_t>(stream), a);
  return hipGetLastError() == hipSuccess ? 0 : -1000;
}
