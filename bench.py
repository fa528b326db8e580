"""NCC find benchmark on one MI355X: the 2048x2048, +-30 degree rotation
search workload (128x128 template pasted 5 times with bilinear rotation in
[-30, 30] degrees, N(0,8) noise and a 0.8-1.2 gain ramp; model with 4 levels
and the automatic angle step; MinScore 0.7, NumMatches 5, MaxOverlap 0.5,
subpixel refinement).

A step is one batched find over `--inflight` frames (default 16) already
resident in HBM: every stage runs for all frames before the next, and the
per-candidate kernels (tracking, full score, refinement) take the candidates
of every frame in one launch.  Consecutive steps are overlapped the way a
serving loop would run them: step k+1 is submitted before step k is waited
for, so step k's refinement and host work overlap step k+1's device work.  The throughput `value` is search megapixels
per second over the whole timed region; the reported latency is that of one
batch, from its submission to the end of its wait.  The coarse correlation runs in the split-BF16 MFMA form by default;
the plain FP32 form is timed on the same frames for the comparison, and the
coarse scores of both are compared with the FP64 form on frame 0.

Prints one JSON line (rank 0).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FORMS = {"fp32": 0, "bf16x3": 1, "fp64": 2, "fp8": 3}
WORKLOAD = {"workload": "rot2048_pm30", "image": [2048, 2048], "template": [128, 128], "instances": 5,
            "angle_range_deg": [-30, 30], "levels": 4, "angle_step": "auto", "min_score": 0.7, "num_matches": 5,
            "max_overlap": 0.5, "subpixel": True}
START, EXTENT = -math.pi / 6, math.pi / 3


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# the HIP kernel behind each reported stage (one launch per batch)
HIP_KERNELS = {("track_mfma", f): "ncc::k_track_mfma" for f in range(4)}
HIP_KERNELS.update({("coarse", 0): "ncc::k_coarse_fp32", ("coarse", 1): "ncc::k_coarse_mfma",
                    ("coarse", 2): "ncc::k_coarse_fp64"})
HIP_KERNELS.update({("track", f): "ncc::k_track_win" for f in range(4)})
HIP_KERNELS.update({("refine", f): "ncc::k_refine" for f in range(4)})
TRAFFIC = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                       "ncc_find_rot2048_pm30_b16_hbm_traffic.md")


def hbm_bytes_per_frame(hip_kernel):
    """Calibrated FETCH_SIZE + WRITE_SIZE bytes per frame of a kernel on this
    workload, from the committed rocprofv3 counter summary (None if absent)."""
    if not os.path.exists(TRAFFIC):
        return None
    for line in open(TRAFFIC):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 5 and cells[0].strip("`").split("<")[0] == hip_kernel:
            return float(cells[4])
    return None


def physical_cores() -> int:
    """Physical cores of the host available to this process."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    cores = set()
    try:
        phys = core = None
        for line in open("/proc/cpuinfo"):
            if line.startswith("physical id"):
                phys = line.split(":")[1].strip()
            elif line.startswith("core id"):
                core = line.split(":")[1].strip()
            elif not line.strip() and core is not None:
                cores.add((phys, core))
                phys = core = None
    except OSError:
        pass
    return max(1, min(avail, len(cores) if cores else avail))


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(scene, args):
    """The CPU restatement (AVX2 build where the host has AVX2, one find per
    thread) on frame 0 of the workload: the median latency of one core, and
    the throughput of one find per physical core (at most 16, the GPU box's
    share of the host) running concurrently."""
    import threading
    from oracle import pyoracle
    avx2 = pyoracle.host_has_avx2()
    o = pyoracle.Model(scene.template, levels=4, start=START, extent=EXTENT, avx2=avx2)
    o.find(scene.image, START, EXTENT, 0.7, 5, 0.5)
    lat = []
    for _ in range(args.cpu_finds):
        t = time.perf_counter()
        o.find(scene.image, START, EXTENT, 0.7, 5, 0.5)
        lat.append(time.perf_counter() - t)
    one = float(np.median(lat))
    n = min(16, physical_cores())
    log(f"CPU baseline: 1-core median {one * 1e3:.1f} ms; {n} threads x {args.cpu_finds} finds")
    models = [pyoracle.Model(scene.template, levels=4, start=START, extent=EXTENT, avx2=avx2) for _ in range(n)]

    def worker(m):
        for _ in range(args.cpu_finds):
            m.find(scene.image, START, EXTENT, 0.7, 5, 0.5)
    threads = [threading.Thread(target=worker, args=(m,)) for m in models]
    t = time.perf_counter()
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    wall = time.perf_counter() - t
    mpx = 2048 * 2048 * n * args.cpu_finds / wall / 1e6
    return {"value": round(mpx, 3), "unit": "Mpx/s", "cores": n, "kind": "port",
            "sample": f"{n} threads x {args.cpu_finds} finds of frame 0, one find per physical core at a time "
                      f"(FP64 restatement, {'AVX2' if avx2 else 'scalar'} build)",
            "latency_ms_1core_median": round(one * 1e3, 2), "mpx_s_1core": round(2048 * 2048 / one / 1e6, 3),
            "nproc": os.cpu_count(), "cpu_model": cpu_model(), "avx2_build": avx2}


def batched_models(capi, torch, args):
    """The batched heterogeneous workload: 32 models (templates 32-256 px per
    side, +-10 deg / +-45 deg / full turn) x 16 images (640x480 .. 4096x3072,
    0-6 instances each, N(0,6)) in one mvamd_find_ncc_models_batch call,
    split-BF16, MinScore 0.8, the coarse candidate budget raised to 100 000
    (small templates exceed the reference's 10 000 on the large images); the
    persistent coarse launch timed by HIP events, and against one coarse
    launch per query."""
    from mvkit_amd import synth
    log("batched models: generating 32 models x 16 images")
    w = synth.batch_workload()
    t = time.perf_counter()
    models = [capi.NccModel(m.template, m.levels, m.start, m.extent, m.step if m.step else "auto") for m in w.models]
    create_s = time.perf_counter() - t
    dev = [torch.from_numpy(im).cuda() for im in w.images]
    shapes = [im.shape for im in w.images]
    mpx = sum(a * b for a, b in shapes) / 1e6
    opts = [capi.NccModel.options(m.start, m.extent, 0.8, 0, 0.5, form=capi.FORM_BF16X3, max_candidates=100000)
            for m in w.models]
    ptrs, ws, hs = [d.data_ptr() for d in dev], [s[1] for s in shapes], [s[0] for s in shapes]

    def call(queue):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res, st = capi.find_models_batch(models, ptrs, ws, hs, ws, opts, 0, 0.5, coarse_queue=queue, stats=True)
        return res, st, time.perf_counter() - t0
    _, st0, first_s = call(True)  # the first call tunes each query's coarse tile shape
    runs = [call(True) for _ in range(3)]
    res, st, _ = runs[-1]
    dt = float(np.median([r[2] for r in runs]))
    call(False)  # tunes the per-query launches' own kinds of search
    per_query = float(np.median([call(False)[2] for _ in range(2)]))
    hit = sum(any(int(x[4]) == k and math.hypot(x[0] - cy, x[1] - cx) < 1.0 for x in r)
              for r, tr in zip(res, w.truth) for k, cy, cx, _ in tr)
    tflops = st.coarse_flops / st.coarse_queue_us / 1e6 if st.coarse_queue_us else 0.0
    out = {"workload": "32 models x 16 images (batch_workload, seed 0x6D766B6974+4)", "images": len(dev),
           "models": len(models), "mpx": round(mpx, 1), "ms_per_batch": round(dt * 1e3, 1),
           "images_per_s": round(len(dev) / dt, 2), "mpx_s": round(mpx / dt, 1),
           "per_query_launches_ms": round(per_query * 1e3, 1), "speedup_persistent": round(per_query / dt, 3),
           "instances": sum(len(tr) for tr in w.truth), "instances_found": hit,
           "matches": sum(len(r) for r in res), "candidates": st.candidates, "model_create_s": round(create_s, 2),
           "first_call_s": round(first_s, 2), "first_call_tune_s": round(st0.tune_us * 1e-6, 2),
           "first_call_tuned": st0.tuned, "refined": st.refined, "refine_rounds": st.refine_rounds,
           "tuned_shapes": sorted({tuple(int(v.split("=")[1]) for v in line.split(" : ")[1].split()[:5])
                                   for m in models for line in m.tuning.splitlines()}),
           "coarse_queue": {"kernel": "ncc::k_coarse_queue", "us": round(st.coarse_queue_us, 1),
                            "algorithmic_gflop": round(st.coarse_flops / 1e9, 1), "achieved": round(tflops, 2),
                            "peak": 2500.0, "unit": "TFLOP/s", "frac": round(tflops / 2500.0, 5)},
           "min_score": 0.8, "max_candidates": 100000}
    log(f"batched models: {out}")
    return out


def serving(capi, torch, args):
    """The serving loop under open-loop Poisson arrivals: each request is one
    frame (1280x1024, 2448x2048 or 5472x3648) with 1-4 models drawn from a
    64-model pool (templates 48^2-320^2, +-15 deg to the full turn), submitted
    to the request loop at its arrival time; latency is from arrival to the
    end of the request's device and host work.  Split-BF16 with micro-batches
    and two workers, against plain FP32 per-frame launches (one worker, one
    request at a time, no persistent coarse launch)."""
    import threading
    from mvkit_amd import synth
    log("serving: generating the 64-model pool and frames")
    w = synth.serving_workload()
    t = time.perf_counter()
    models = [capi.NccModel(m.template, m.levels, m.start, m.extent, m.step if m.step else "auto") for m in w.models]
    create_s = time.perf_counter() - t
    rng = np.random.default_rng(0x6D766B6974 + 5)
    n_max = int(max(args.serve_rates) * args.serve_seconds) + 1
    trace = [(int(rng.integers(len(w.images))), list(rng.integers(len(models), size=int(rng.integers(1, 5)))))
             for _ in range(n_max)]

    def opts_for(subset, form):
        return [capi.NccModel.options(w.models[i].start, w.models[i].extent, 0.8, 0, 0.5, form=form,
                                      max_candidates=300000, vector_tracking=form == capi.FORM_FP32) for i in subset]

    def run(rate, form, **server_kw):
        n = max(1, int(rate * args.serve_seconds))
        gaps = rng.exponential(1.0 / rate, size=n)
        arrivals = np.cumsum(gaps) - gaps[0]
        prepared = [(w.images[f], [models[i] for i in sub], opts_for(sub, form)) for f, sub in trace[:n]]
        server = capi.Server(**server_kw)
        subs = [None] * n
        ready = threading.Semaphore(0)

        def submitter():
            t0 = time.perf_counter()
            for k in range(n):
                delay = t0 + arrivals[k] - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                img, ms, os_ = prepared[k]
                rid = server.submit(img, ms, os_, 0, 0.5)
                subs[k] = (rid, t0 + arrivals[k], time.perf_counter())
                ready.release()
        th = threading.Thread(target=submitter)
        th.start()
        lat, ends, batches, failed = [], [], [], 0
        for k in range(n):
            ready.acquire()
            rid, arrival, submitted = subs[k]
            try:
                _, tm = server.wait(rid)
            except capi.MvError:  # e.g. a request over the candidate budget (the reference raises the same)
                failed += 1
                continue
            end = submitted + tm["finished_us"] * 1e-6
            lat.append(end - arrival)
            ends.append(end)
            batches.append(tm["batch_size"])
        th.join()
        server.close()
        first = subs[0][1]
        lat = np.array(lat) * 1e3
        return {"offered_req_s": rate, "requests": n, "failed": failed,
                "achieved_req_s": round(float((n - failed) / (max(ends) - first)), 1),
                "p50_ms": round(float(np.percentile(lat, 50)), 2), "p99_ms": round(float(np.percentile(lat, 99)), 2),
                "max_ms": round(float(lat.max()), 2), "mean_batch": round(float(np.mean(batches)), 2)}
    # warm: every frame and model once on a server of each kind
    for form, kw in ((capi.FORM_BF16X3, {"workers": 2, "max_batch": 8}),
                     (capi.FORM_FP32, {"workers": 1, "max_batch": 1, "coarse_queue": False})):
        srv = capi.Server(**kw)
        ids = [srv.submit(w.images[i % len(w.images)], [models[i]], opts_for([i], form), 0, 0.5)
               for i in range(len(models))]
        for i in ids:
            try:
                srv.wait(i)
            except capi.MvError:
                pass
        srv.close()
    out = {"workload": "serving_workload (seed 0x6D766B6974+5): 64 models, 2 frames of each size, 1-4 models per "
                       "request, MinScore 0.8, candidate budget 300 000", "model_create_s": round(create_s, 2),
           "seconds_per_rate": args.serve_seconds, "split_bf16_batched": [], "fp32_per_frame": []}
    for rate in args.serve_rates:
        r = run(rate, capi.FORM_BF16X3, workers=2, max_batch=8)
        log(f"serving split-BF16, micro-batches: {r}")
        out["split_bf16_batched"].append(r)
    for rate in args.serve_rates[:2]:
        r = run(rate, capi.FORM_FP32, workers=1, max_batch=1, coarse_queue=False)
        log(f"serving FP32 per frame: {r}")
        out["fp32_per_frame"].append(r)
    return out


def precision_sweep(capi, args):
    """Each correlation form's coarse score error against the FP64 oracle on
    the full-turn workload (2048^2, 256^2 template, 5 levels, AngleStep 0.035),
    and whether its matches are the oracle's within the parity tolerances."""
    from mvkit_amd import synth
    from oracle import pyoracle
    log("precision sweep: full-turn workload")
    s = synth.workload_full_2048()
    o = pyoracle.Model(s.template, levels=5, start=0.0, extent=2 * math.pi, step=0.035)
    want, tr = o.find(s.image, 0.0, 2 * math.pi, 0.7, 5, 0.5, True, trace=True)
    ref = {(r, c, a): sc for r, c, a, sc in tr["coarse"]}
    m = capi.NccModel(s.template, 5, 0.0, 2 * math.pi, 0.035)
    out = {"workload": "full2048: 2048^2, 256^2 template x5, 0..2pi, 5 levels, AngleStep 0.035", "forms": {}}
    for name, form in (("fp32", 0), ("bf16x3", 1), ("fp8_screen", 3)):
        g = m.find_trace(s.image, capi.NccModel.options(0.0, 2 * math.pi, 0.7, 5, 0.5, form=form))
        errs = np.array([abs(sc - ref[(r, c, a)]) for r, c, a, sc in g["coarse"] if (r, c, a) in ref])
        res = g["results"]
        same = res.shape == want.shape and (len(want) == 0 or (
            np.abs(res[:, :2] - want[:, :2]).max() <= 0.02 and np.abs(res[:, 2] - want[:, 2]).max() <= math.radians(0.01)
            and np.abs(res[:, 3] - want[:, 3]).max() <= 1e-4))
        out["forms"][name] = {"coarse_err_max": float(errs.max()), "coarse_err_rms": float(np.sqrt((errs ** 2).mean())),
                              "peaks": int(len(errs)), "oracle_peaks": len(ref), "matches_oracle": bool(same)}
    log(f"precision sweep: {out}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--form", default="bf16x3", choices=sorted(FORMS))
    ap.add_argument("--inflight", type=int, default=16)
    ap.add_argument("--compare-steps", type=int, default=20,
                    help="timed steps of the plain FP32 form for the comparison (0: skip)")
    ap.add_argument("--cpu-finds", type=int, default=6,
                    help="finds per thread in the CPU baseline sample: the 1-core latency sample and each core's "
                         "share of the all-core sample (0: skip)")
    ap.add_argument("--latency-calls", type=int, default=200,
                    help="single-frame mv_find_ncc_model calls timed for p50/p99 (0: skip)")
    ap.add_argument("--batch-models", type=int, default=1,
                    help="time the batched heterogeneous workload (32 models x 16 images, 0: skip)")
    ap.add_argument("--serve-seconds", type=float, default=1.5,
                    help="seconds of Poisson arrivals per rate in the serving measurement (0: skip)")
    ap.add_argument("--serve-rates", type=float, nargs="+", default=[50.0, 100.0, 200.0, 400.0],
                    help="offered request rates (requests/s) of the serving measurement")
    ap.add_argument("--precision-sweep", type=int, default=1,
                    help="report each correlation form's coarse score error on the full-turn workload (0: skip)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run three plain 64 MiB device-to-device copies (the reference kernel that scales "
                         "the FETCH_SIZE / WRITE_SIZE counters to bytes) and exit")
    ap.add_argument("--part", default="all", choices=["all", "batched", "serving"],
                    help="batched / serving: run only that measurement (for profiling it) and print its JSON")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the NCC find path shards by frame, and only the single-GPU "
              "configuration is benchmarked here (use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import torch
    from mvkit_amd import capi, synth

    assert torch.cuda.is_available(), "no GPU"
    torch.cuda.set_device(0)
    if args.copy_calibration:
        n = 64 << 20
        a = torch.empty(n, dtype=torch.uint8, device="cuda").random_(0, 255)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"copied {n} bytes three times")
        return
    if args.part == "batched":
        print(json.dumps({"batched_models": batched_models(capi, torch, args)}), flush=True)
        return
    if args.part == "serving":
        print(json.dumps({"serving": serving(capi, torch, args)}), flush=True)
        return
    F = max(1, args.inflight)
    log(f"generating {F} frames")
    scenes = [synth.workload_rot_2048(frame=k) for k in range(F)]
    model = capi.NccModel(scenes[0].template, 4, START, EXTENT)
    frames = [torch.from_numpy(s.image).cuda() for s in scenes]
    torch.cuda.synchronize()
    form = FORMS[args.form]
    opts = capi.NccModel.options(START, EXTENT, 0.7, 5, 0.5, True, form=form)

    # parity guard: the benchmarked path must reproduce the truth of frame 0
    res, st = model.find_device(frames[0].data_ptr(), 2048, 2048, 2048, opts)
    assert len(res) == 5, res
    for r, c, a, s in res:
        t = min(scenes[0].truth, key=lambda t: math.hypot(r - t[0], c - t[1]))
        assert math.hypot(r - t[0], c - t[1]) < 0.2 and abs(a - t[2]) < 2e-3, (res, scenes[0].truth)

    ptrs = [f.data_ptr() for f in frames]
    latencies = []

    def step(o):
        t = time.perf_counter()
        out, st = model.find_batch(ptrs, 2048, 2048, 2048, o)
        latencies.append(time.perf_counter() - t)
        return out, st

    def pipelined(o, steps, on_stats=None):
        """`steps` batches with the next batch submitted before the previous
        one is waited for (its refinement overlaps the next device part)."""
        pending = None
        for i in range(steps + 1):
            nxt = (model.submit_batch(ptrs, 2048, 2048, 2048, o), time.perf_counter()) if i < steps else None
            if pending is not None:
                _, st = model.wait_batch(pending[0])
                latencies.append(time.perf_counter() - pending[1])
                if on_stats:
                    on_stats(st)
            pending = nxt

    # the batch must return, frame by frame, what a single find returns
    batch0, _ = step(opts)
    assert len(batch0) == F and np.array_equal(batch0[0], res), (batch0[0], res)

    log(f"warmup {args.warmup} steps x {F} frames")
    pipelined(opts, args.warmup)

    # the coarse tile shape the autotuner chose on the first batch (the
    # sample launch's time with it and with the default shape)
    def tuned_shape(m):
        out = []
        for line in m.tuning.splitlines():
            key, rest = line.split(" : ")
            f = dict(kv.split("=") for kv in rest.split())
            out.append({"key": key, "shape": [int(f[k]) for k in ("rows", "col_waves", "row_waves", "tiles", "bound")],
                        "us": round(float(f["us"]), 2), "default_us": round(float(f["default_us"]), 2)})
        return out
    coarse_tiles = tuned_shape(model)
    log(f"autotuned coarse tiles: {coarse_tiles}")

    # per-stage device time from a short pass with every stage timed
    # (an event pair around each stage's launches on the library's stream)
    kernel = {k: 0.0 for k in capi.STAGES}
    prof_steps = max(1, min(args.steps, 5))
    for _ in range(prof_steps):
        _, st = step(capi.NccModel.options(START, EXTENT, 0.7, 5, 0.5, True, form=form, profile=True))
        for i, name in enumerate(capi.STAGES):
            kernel[name] += st.kernel_us[i] / (prof_steps * F)
    dominant = max(kernel, key=kernel.get)
    dom_bit = 1 << capi.STAGES.index(dominant)

    # timed region: only the dominant stage carries events, for its roofline
    latencies.clear()
    torch.cuda.synchronize()
    log(f"timed {args.steps} steps x {F} frames, form {args.form}, events on {dominant}")
    timed = capi.NccModel.options(START, EXTENT, 0.7, 5, 0.5, True, form=form, profile=dom_bit)
    acc = {"us": 0.0, "flops": 0.0}

    def take(st):
        acc["us"] += st.kernel_us[capi.STAGES.index(dominant)]
        acc["flops"] += {"track": st.track_flops - st.track_mfma_flops,
                         "track_mfma": st.track_mfma_flops}.get(dominant, st.coarse_flops)
    t0 = time.perf_counter()
    pipelined(timed, args.steps, take)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    dom_us, flops = acc["us"], acc["flops"]
    frames_done = args.steps * F
    mpx = 2048 * 2048 * frames_done / elapsed / 1e6
    lat = np.array(latencies) * 1e3

    # roofline of the dominant stage: the coarse correlation (FP32 VALU or
    # BF16/FP8 MFMA) or tracking (split-BF16 MFMA on the levels it applies to,
    # FP64 VALU on the others); per frame, device time from the timed
    # region's events
    per_frame_us = dom_us / frames_done
    per_frame_flop = flops / frames_done
    achieved = per_frame_flop / (per_frame_us * 1e-6) / 1e12 if per_frame_us > 0 else 0.0
    if dominant == "track":
        bound, peak = "valu-fp64", 78.6
    elif dominant == "track_mfma":
        bound, peak = "mfma", 2500.0
    elif dominant == "coarse":
        bound, peak = ("mfma" if form in (1, 3) else "valu"), {0: 157.3, 1: 2500.0, 2: 78.6, 3: 5000.0}[form]
    else:
        bound, peak = "latency", None
    total_kernel = sum(kernel.values())
    # HBM bytes of one launch of the dominant kernel (the batch's frames)
    per_frame_bytes = hbm_bytes_per_frame(HIP_KERNELS.get((dominant, form), ""))
    traffic = round(per_frame_bytes * F) if per_frame_bytes is not None else None

    # the plain FP32 form on the same frames, for the split-BF16 comparison
    fp32 = None
    if args.compare_steps > 0 and form != FORMS["fp32"]:
        o32 = capi.NccModel.options(START, EXTENT, 0.7, 5, 0.5, True, form=FORMS["fp32"])
        pipelined(o32, 2)
        torch.cuda.synchronize()
        t = time.perf_counter()
        pipelined(o32, args.compare_steps)
        torch.cuda.synchronize()
        e32 = time.perf_counter() - t
        fp32 = {"value": round(2048 * 2048 * args.compare_steps * F / e32 / 1e6, 2), "unit": "Mpx/s",
                "ms_per_frame": round(e32 / (args.compare_steps * F) * 1e3, 4), "steps": args.compare_steps}

    # the FP8-screened form on the same frames: throughput, and its matches
    # against the split-BF16 form's (same instances; poses and scores within
    # the parity tolerances)
    fp8 = None
    if args.compare_steps > 0 and form == FORMS["bf16x3"]:
        o8 = capi.NccModel.options(START, EXTENT, 0.7, 5, 0.5, True, form=FORMS["fp8"])
        got8, _ = model.find_batch(ptrs, 2048, 2048, 2048, o8)
        want16, _ = step(opts)
        same = sum(int(np.array_equal(a, b)) for a, b in zip(got8, want16))
        within = sum(int(a.shape == b.shape and (len(a) == 0 or (np.abs(a[:, :2] - b[:, :2]).max() <= 0.02 and
                                                              np.abs(a[:, 2] - b[:, 2]).max() <= math.radians(0.01) and
                                                              np.abs(a[:, 3] - b[:, 3]).max() <= 1e-4)))
                     for a, b in zip(got8, want16))
        pipelined(o8, 2)
        torch.cuda.synchronize()
        t = time.perf_counter()
        pipelined(o8, args.compare_steps)
        torch.cuda.synchronize()
        e8 = time.perf_counter() - t
        v8 = 2048 * 2048 * args.compare_steps * F / e8 / 1e6
        fp8 = {"value": round(v8, 2), "unit": "Mpx/s", "ms_per_frame": round(e8 / (args.compare_steps * F) * 1e3, 4),
               "steps": args.compare_steps, "speedup_vs_bf16x3": round(v8 / mpx, 3), "screen": 0.55,
               "frames_identical_to_bf16x3": same, "frames_within_tolerance": within, "frames": F}
        log(f"fp8-screened form: {fp8}")

    # coarse score error of each form against this build's FP64 form (the
    # reference's sample order in FP64) on frame 0, matched by (row, col, angle)
    def coarse(f):
        o = capi.NccModel.options(START, EXTENT, 0.7, 5, 0.5, True, form=f)
        return {(r, c, a): sc for r, c, a, sc in model.find_trace(scenes[0].image, o)["coarse"]}
    ref = coarse(FORMS["fp64"])
    score_err = {}
    for name in ("fp32", "bf16x3", "fp8"):
        got = coarse(FORMS[name])
        errs = [abs(got[k] - v) for k, v in ref.items() if k in got]
        score_err[name] = {"max": float(max(errs)), "rms": float(np.sqrt(np.mean(np.square(errs)))),
                           "peaks": len(errs), "missing": len(ref) - len(errs)}

    # model construction, host and device builders (identical models)
    create_ms = {}
    capi.NccModel(scenes[0].template, 4, START, EXTENT, builder=capi.CREATE_DEVICE)
    for name, b in (("host", capi.CREATE_HOST), ("device", capi.CREATE_DEVICE)):
        t = time.perf_counter()
        for _ in range(3):
            capi.NccModel(scenes[0].template, 4, START, EXTENT, builder=b)
        create_ms[name] = round((time.perf_counter() - t) / 3 * 1e3, 2)
    log(f"model create ms: {create_ms}")

    # single-frame latency of the drop-in entry: mv_find_ncc_model on a host
    # image object (upload, find, refinement, tuples with the HALCON shift)
    latency = None
    if args.latency_calls > 0:
        log(f"single-frame latency: {args.latency_calls} mv_find_ncc_model calls")
        mc = capi.NccModel(scenes[0].template, 4, START, EXTENT)
        img = capi.Image(scenes[0].image)
        for _ in range(10):
            mc.find(img, START, EXTENT, 0.7, 5, 0.5, "true", 0)
        lat1 = []
        for _ in range(args.latency_calls):
            t = time.perf_counter()
            got = mc.find(img, START, EXTENT, 0.7, 5, 0.5, "true", 0)
            lat1.append(time.perf_counter() - t)
        assert len(got) == 5
        lat1 = np.array(lat1) * 1e3
        latency = {"entry": "mv_find_ncc_model (host u8 image object, C ABI)", "calls": args.latency_calls,
                   "p50_ms": round(float(np.percentile(lat1, 50)), 4), "p99_ms": round(float(np.percentile(lat1, 99)), 4),
                   "mpx_s_at_p50": round(float(2048 * 2048 / (np.percentile(lat1, 50) * 1e-3) / 1e6), 1)}
        log(f"single-frame latency: {latency}")

    batched = batched_models(capi, torch, args) if args.batch_models else None
    sweep = precision_sweep(capi, args) if args.precision_sweep else None
    serve = serving(capi, torch, args) if args.serve_seconds > 0 else None
    cpu = cpu_baseline(scenes[0], args) if args.cpu_finds > 0 else None

    line = {
        "metric": "search Mpx/s & ms/frame, 2048\u00b2 \u00b130\u00b0: split-BF16 vs FP32 at \u22641e-4 score err",
        "value": round(mpx, 2), "unit": "Mpx/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "ms_per_frame": round(elapsed / frames_done * 1e3, 4),
        "batch_latency_ms_p50": round(float(np.percentile(lat, 50)), 4),
        "batch_latency_ms_p99": round(float(np.percentile(lat, 99)), 4),
        "frames_in_flight": F, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {0: "fp32", 1: "bf16x3", 2: "f64", 3: "fp8+bf16x3"}[form] + " correlation / f64 refinement",
        "form": args.form, "data": "synthetic (deterministic SplitMix64 generator)",
        "config": WORKLOAD,
        "roofline": {"kernel": dominant, "hip_kernel": HIP_KERNELS.get((dominant, form), dominant),
                     "us_per_launch": round(per_frame_us * F, 2),  # one launch of the stage per batch
                     "bound": bound, "achieved": round(achieved, 3), "peak": peak,
                     "unit": "TFLOP/s", "frac": round(achieved / peak, 5) if peak else None,
                     "traffic": traffic,
                     "us_per_frame": round(per_frame_us, 2), "algorithmic_gflop_per_frame": round(per_frame_flop / 1e9, 4)},
        "kernel_us_per_frame": {k: round(v, 2) for k, v in kernel.items()},
        "coarse_tile_autotuned": coarse_tiles,
        "dominant_kernel": {"stage": dominant, "share": round(kernel[dominant] / total_kernel, 3) if total_kernel else 0},
        "fp32_form": fp32,
        "speedup_vs_fp32": round(mpx / fp32["value"], 3) if fp32 else None,
        "fp8_screened_form": fp8,
        "coarse_score_error_vs_fp64": score_err,
        "model_create_ms": create_ms,
        "single_frame_latency": latency,
        "batched_models": batched,
        "precision_sweep": sweep,
        "serving": serve,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
