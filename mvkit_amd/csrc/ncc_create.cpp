// Host driver of model construction on the device (rebuild, model.cpp:173-242):
// the pyramid, angle grid and original patch come from plan_model; the
// rotated rasters of every (level, angle) are built by ncchip_create_rasters
// in three launches (count, sample + normalise, runs) and copied into the
// model's host pool in the layout build_raster produces, so the model is the
// host builder's bit for bit.  The raster errors are raised in the host
// builder's order: per raster, the rotated-box budget, then a degenerate
// raster, then the sample budget.
#include <algorithm>
#include <cstring>
#include <optional>

#include "ncc_device.hpp"

namespace mvamd {

namespace {
size_t align_up(size_t v, size_t a = 256) { return (v + a - 1) / a * a; }

template <class T>
void download(std::vector<T>& dst, const void* src, size_t n, hipStream_t stream) {
  dst.resize(n);
  if (n) check(hipMemcpyAsync(dst.data(), src, n * sizeof(T), hipMemcpyDeviceToHost, stream), "model download");
}
}  // namespace

std::unique_ptr<NccModel> FindContext::rebuild(LevelGeom first, const CreateParams& p, double origin_row,
                                               double origin_col) {
  ModelPlan plan = plan_model(std::move(first), p, origin_row, origin_col);
  NccModel& m = *plan.model;
  // ---- jobs in the host builder's order, up to the first rotated-box error
  std::vector<NccRasterJob> jobs;
  std::optional<Error> grid_error;
  for (size_t li = 0; li < m.levels.size() && !grid_error; ++li) {
    for (size_t i = 0; i < m.n_angles; ++i) {
      const double angle = raster_angle(m, i);
      RasterGrid g;
      try {
        g = raster_grid(m.levels[li], angle);
      } catch (Error& e) {
        grid_error = e;
        break;
      }
      NccRasterJob J{};
      J.ca = g.ca; J.sa = g.sa;
      J.level = int32_t(li); J.unrotated = g.unrotated;
      J.prepared = li == 0 && g.unrotated;
      J.mean = plan.first_mean; J.energy = plan.first_energy;
      J.r_lo = g.r_lo; J.r_hi = g.r_hi; J.c_lo = g.c_lo; J.c_hi = g.c_hi;
      jobs.push_back(J);
    }
  }
  const size_t n_jobs = jobs.size();
  // ---- workspace: levels, jobs, stats, then the sample and run arrays
  size_t level_bytes = 0;
  for (const auto& lv : m.levels) level_bytes += align_up(lv.gray.size() * 8) + align_up(lv.mask.size());
  const size_t off_levels = 0;
  const size_t off_desc = align_up(off_levels + level_bytes);
  const size_t off_jobs = align_up(off_desc + m.levels.size() * sizeof(NccTemplateLevel));
  const size_t off_stats = align_up(off_jobs + n_jobs * sizeof(NccRasterJob));
  const size_t head = align_up(off_stats + n_jobs * sizeof(NccRasterStats));
  auto grow = [&](size_t bytes) {
    if (create_.bytes < bytes) {
      check(hipStreamSynchronize(stream_), "sync");
      create_ = DeviceBuffer(align_up(bytes + bytes / 4, 1 << 20));
    }
  };
  grow(head);
  uint8_t* dev = create_.as<uint8_t>();
  std::vector<NccTemplateLevel> desc(m.levels.size());
  {
    size_t off = off_levels;
    for (size_t li = 0; li < m.levels.size(); ++li) {
      const LevelGeom& lv = m.levels[li];
      check(hipMemcpyAsync(dev + off, lv.gray.data(), lv.gray.size() * 8, hipMemcpyHostToDevice, stream_), "upload");
      desc[li].gray = reinterpret_cast<const double*>(dev + off);
      off += align_up(lv.gray.size() * 8);
      check(hipMemcpyAsync(dev + off, lv.mask.data(), lv.mask.size(), hipMemcpyHostToDevice, stream_), "upload");
      desc[li].mask = dev + off;
      off += align_up(lv.mask.size());
      desc[li].w = int32_t(lv.w); desc[li].h = int32_t(lv.h);
      desc[li].orow = lv.orow; desc[li].ocol = lv.ocol;
    }
  }
  check(hipMemcpyAsync(dev + off_desc, desc.data(), desc.size() * sizeof(NccTemplateLevel), hipMemcpyHostToDevice,
                       stream_), "upload");
  NcchipCreateBuffers B{};
  B.levels = reinterpret_cast<const NccTemplateLevel*>(dev + off_desc);
  B.jobs = reinterpret_cast<const NccRasterJob*>(dev + off_jobs);
  B.stats = reinterpret_cast<NccRasterStats*>(dev + off_stats);
  std::vector<NccRasterStats> stats(n_jobs);
  if (n_jobs) {
    // ---- 1: sample counts
    check(hipMemcpyAsync(dev + off_jobs, jobs.data(), n_jobs * sizeof(NccRasterJob), hipMemcpyHostToDevice, stream_),
          "upload");
    if (ncchip_create_rasters(NCCHIP_CREATE_COUNT, &B, int(n_jobs), stream_))
      throw Error(kInternalError, "ncc_model_create", "device", "raster count launch failed");
    check(hipMemcpyAsync(stats.data(), B.stats, n_jobs * sizeof(NccRasterStats), hipMemcpyDeviceToHost, stream_),
          "download");
    check(hipStreamSynchronize(stream_), "model create");
  }
  // jobs to build: up to the first one past the sample budget (its
  // degenerate check comes before the budget check)
  size_t total = 0, n_build = n_jobs;
  std::optional<size_t> over_budget;
  for (size_t j = 0; j < n_jobs; ++j) {
    jobs[j].sample_off = int64_t(total);
    const size_t n = size_t(stats[j].count);
    if (n >= 4 && n > max_model_samples() - std::min(total, max_model_samples())) {
      over_budget = j;
      n_build = j + 1;
      total += n;
      break;
    }
    total += n;
  }
  const size_t off_rc = head;
  const size_t off_v = align_up(off_rc + total * 4);
  const size_t off_w = align_up(off_v + total * 8);
  const size_t off_wf = align_up(off_w + total * 8);
  const size_t after_samples = align_up(off_wf + total * 4);
  grow(after_samples);
  dev = create_.as<uint8_t>();
  B.levels = reinterpret_cast<const NccTemplateLevel*>(dev + off_desc);
  B.jobs = reinterpret_cast<const NccRasterJob*>(dev + off_jobs);
  B.stats = reinterpret_cast<NccRasterStats*>(dev + off_stats);
  B.rc = reinterpret_cast<int32_t*>(dev + off_rc);
  B.v = reinterpret_cast<double*>(dev + off_v);
  B.w = reinterpret_cast<double*>(dev + off_w);
  B.wf = reinterpret_cast<float*>(dev + off_wf);
  if (n_build) {
    // ---- 2: samples and normalisation (a grown workspace lost the earlier
    // uploads: send the level data again)
    {
      size_t off = off_levels;
      for (size_t li = 0; li < m.levels.size(); ++li) {
        const LevelGeom& lv = m.levels[li];
        check(hipMemcpyAsync(dev + off, lv.gray.data(), lv.gray.size() * 8, hipMemcpyHostToDevice, stream_), "upload");
        desc[li].gray = reinterpret_cast<const double*>(dev + off);
        off += align_up(lv.gray.size() * 8);
        check(hipMemcpyAsync(dev + off, lv.mask.data(), lv.mask.size(), hipMemcpyHostToDevice, stream_), "upload");
        desc[li].mask = dev + off;
        off += align_up(lv.mask.size());
      }
      check(hipMemcpyAsync(dev + off_desc, desc.data(), desc.size() * sizeof(NccTemplateLevel),
                           hipMemcpyHostToDevice, stream_), "upload");
    }
    check(hipMemcpyAsync(dev + off_jobs, jobs.data(), n_build * sizeof(NccRasterJob), hipMemcpyHostToDevice, stream_),
          "upload");
    check(hipMemcpyAsync(B.stats, stats.data(), n_build * sizeof(NccRasterStats), hipMemcpyHostToDevice, stream_),
          "upload");
    if (ncchip_create_rasters(NCCHIP_CREATE_SAMPLES, &B, int(n_build), stream_))
      throw Error(kInternalError, "ncc_model_create", "device", "raster launch failed");
    check(hipMemcpyAsync(stats.data(), B.stats, n_build * sizeof(NccRasterStats), hipMemcpyDeviceToHost, stream_),
          "download");
    check(hipStreamSynchronize(stream_), "model create");
  }
  // ---- errors in the host builder's order
  for (size_t j = 0; j < n_build; ++j) {
    if (stats[j].status) raise_degenerate_raster();
    if (over_budget && *over_budget == j) raise_sample_budget();
  }
  if (grid_error) throw *grid_error;
  // ---- 3: runs
  size_t n_runs = 0, n_wruns = 0;
  for (size_t j = 0; j < n_jobs; ++j) {
    jobs[j].run_off = int64_t(n_runs);
    jobs[j].wrun_off = int64_t(n_wruns);
    n_runs += size_t(stats[j].n_runs);
    n_wruns += size_t(stats[j].n_wruns);
  }
  const size_t off_runs = after_samples;
  const size_t off_wruns = align_up(off_runs + n_runs * 16);
  const size_t off_wrun_w = align_up(off_wruns + n_wruns * 16);
  const size_t off_rs = align_up(off_wrun_w + n_wruns * 8);
  const size_t off_ws = align_up(off_rs + n_runs * 4);
  const size_t end = align_up(off_ws + n_wruns * 4);
  if (create_.bytes < end) {
    // keep the samples: move them to a larger workspace
    DeviceBuffer bigger(align_up(end + end / 4, 1 << 20));
    check(hipMemcpyAsync(bigger.ptr, create_.ptr, after_samples, hipMemcpyDeviceToDevice, stream_), "copy");
    check(hipStreamSynchronize(stream_), "sync");
    create_ = std::move(bigger);
    dev = create_.as<uint8_t>();
  }
  B.jobs = reinterpret_cast<const NccRasterJob*>(dev + off_jobs);
  B.stats = reinterpret_cast<NccRasterStats*>(dev + off_stats);
  B.rc = reinterpret_cast<int32_t*>(dev + off_rc);
  B.v = reinterpret_cast<double*>(dev + off_v);
  B.w = reinterpret_cast<double*>(dev + off_w);
  B.wf = reinterpret_cast<float*>(dev + off_wf);
  B.runs = reinterpret_cast<int32_t*>(dev + off_runs);
  B.wruns = reinterpret_cast<int32_t*>(dev + off_wruns);
  B.wrun_w = reinterpret_cast<double*>(dev + off_wrun_w);
  B.run_start = reinterpret_cast<int32_t*>(dev + off_rs);
  B.wrun_start = reinterpret_cast<int32_t*>(dev + off_ws);
  RasterPool& pool = m.pool;
  if (n_jobs) {
    check(hipMemcpyAsync(dev + off_jobs, jobs.data(), n_jobs * sizeof(NccRasterJob), hipMemcpyHostToDevice, stream_),
          "upload");
    if (ncchip_create_rasters(NCCHIP_CREATE_RUNS, &B, int(n_jobs), stream_))
      throw Error(kInternalError, "ncc_model_create", "device", "raster run launch failed");
    download(pool.rc, B.rc, total, stream_);
    download(pool.w, B.w, total, stream_);
    download(pool.wf, B.wf, total, stream_);
    download(pool.runs, B.runs, 4 * n_runs, stream_);
    download(pool.wruns, B.wruns, 4 * n_wruns, stream_);
    download(pool.wrun_w, B.wrun_w, n_wruns, stream_);
    check(hipStreamSynchronize(stream_), "model create");
  }
  // ---- raster descriptors, level-major as build_raster appends them
  for (size_t j = 0; j < n_jobs; ++j) {
    const NccRasterJob& J = jobs[j];
    const NccRasterStats& st = stats[j];
    LevelGeom& lv = m.levels[size_t(J.level)];
    if (j % m.n_angles == 0) lv.first_raster = uint32_t(pool.rasters.size());
    Raster r;
    r.angle = raster_angle(m, j % m.n_angles);
    r.orow = J.unrotated ? lv.orow : 0;
    r.ocol = J.unrotated ? lv.ocol : 0;
    r.wsum = st.wsum;
    r.wsum_f = st.wsum_f;
    r.rmin = st.rmin; r.rmax = st.rmax; r.cmin = st.cmin; r.cmax = st.cmax;
    r.first = uint32_t(J.sample_off);
    r.count = uint32_t(st.count);
    r.run_first = uint32_t(J.run_off);
    r.run_count = uint32_t(st.n_runs);
    r.wrun_first = uint32_t(J.wrun_off);
    r.wrun_count = uint32_t(st.n_wruns);
    pool.rasters.push_back(r);
  }
  return finish_model(std::move(plan));
}

}  // namespace mvamd
