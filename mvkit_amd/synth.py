"""Deterministic synthetic scenes for the NCC find workloads.

Generator: SplitMix64 -> u64 -> double as (x >> 11) * 2**-53; normals by
Box-Muller in FP64.  Buffers are row-major u8 with stride == width.

* blurred-noise base: uniform [0,1) per pixel, separable Gaussian sigma=2
  (13 taps, radius 6, normalised, reflect-101 border) in FP64, linearly
  rescaled min->40 / max->215 and rounded half-to-even to u8;
* template: the same base with its own seed, then 6 filled shapes (ellipse or
  rectangle with p=1/2; centre uniform in the template; semi-axes uniform in
  [w/10, w/3]; rotation uniform in [0, pi); integer level in [30, 220]);
* pastes: backward bilinear warp of the template about its centroid
  ((h-1)/2, (w-1)/2), so an instance pasted at centre (cy, cx) with angle t is
  found by the reference's find at pose (cy, cx, t) in its internal
  (centroid) convention.

Each workload seed is 0x6D766B6974 + k (k = workload index).
"""
from __future__ import annotations

import hashlib
import math
from dataclasses import dataclass, field

import numpy as np

SEED_BASE = 0x6D766B6974
_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


class SplitMix64:
    """Vectorised SplitMix64 stream."""

    def __init__(self, seed: int):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)

    def u64(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            k = np.arange(1, n + 1, dtype=np.uint64)
            z = self.state + k * _GAMMA
            self.state = self.state + np.uint64(n) * _GAMMA
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            z = z ^ (z >> np.uint64(31))
        return z

    def uniform(self, n: int) -> np.ndarray:
        return (self.u64(n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)

    def uniform1(self) -> float:
        return float(self.uniform(1)[0])

    def normal(self, n: int) -> np.ndarray:
        m = (n + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log1p(-u1))  # 1-u1 in (0,1]
        z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
        return z[:n]


def _gauss_kernel(sigma: float = 2.0, radius: int = 6) -> np.ndarray:
    x = np.arange(-radius, radius + 1, dtype=np.float64)
    k = np.exp(-0.5 * (x / sigma) ** 2)
    return k / k.sum()


def _blur_axis(a: np.ndarray, k: np.ndarray, axis: int) -> np.ndarray:
    r = (len(k) - 1) // 2
    pad = [(0, 0), (0, 0)]
    pad[axis] = (r, r)
    p = np.pad(a, pad, mode="reflect")  # numpy 'reflect' == reflect-101
    out = np.zeros_like(a)
    n = a.shape[axis]
    for i, w in enumerate(k):
        sl = [slice(None), slice(None)]
        sl[axis] = slice(i, i + n)
        out += w * p[tuple(sl)]
    return out


def blurred_noise(rng: SplitMix64, h: int, w: int) -> np.ndarray:
    """Blurred-noise base image as u8 in [40, 215]."""
    u = rng.uniform(h * w).reshape(h, w)
    k = _gauss_kernel()
    b = _blur_axis(_blur_axis(u, k, 1), k, 0)
    lo, hi = b.min(), b.max()
    s = 40.0 + (b - lo) * (175.0 / (hi - lo))
    return np.rint(s).clip(0, 255).astype(np.uint8)


def make_template(rng: SplitMix64, h: int, w: int, shapes: int = 6) -> np.ndarray:
    t = blurred_noise(rng, h, w).astype(np.float64)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    for _ in range(shapes):
        ellipse = rng.uniform1() < 0.5
        cy = rng.uniform1() * (h - 1)
        cx = rng.uniform1() * (w - 1)
        ay = w / 10 + rng.uniform1() * (w / 3 - w / 10)
        ax = w / 10 + rng.uniform1() * (w / 3 - w / 10)
        th = rng.uniform1() * math.pi
        level = 30 + int(rng.uniform1() * 191)  # integer uniform in [30, 220]
        c, s = math.cos(th), math.sin(th)
        u = c * (xx - cx) + s * (yy - cy)
        v = -s * (xx - cx) + c * (yy - cy)
        if ellipse:
            inside = (u / ax) ** 2 + (v / ay) ** 2 <= 1.0
        else:
            inside = (np.abs(u) <= ax) & (np.abs(v) <= ay)
        t[inside] = level
    return t.astype(np.uint8)


def paste(image: np.ndarray, tmpl: np.ndarray, cy: float, cx: float, angle: float) -> None:
    """Backward bilinear warp of ``tmpl`` rotated by ``angle`` about its centroid
    to centre (cy, cx); pixels whose template coordinate is outside are kept."""
    th, tw = tmpl.shape
    oy, ox = (th - 1) / 2.0, (tw - 1) / 2.0
    rad = math.hypot(th, tw) / 2 + 2
    r0, r1 = max(0, int(math.floor(cy - rad))), min(image.shape[0], int(math.ceil(cy + rad)) + 1)
    c0, c1 = max(0, int(math.floor(cx - rad))), min(image.shape[1], int(math.ceil(cx + rad)) + 1)
    yy, xx = np.mgrid[r0:r1, c0:c1].astype(np.float64)
    ca, sa = math.cos(angle), math.sin(angle)
    dy, dx = yy - cy, xx - cx
    ty = ca * dy + sa * dx + oy
    tx = -sa * dy + ca * dx + ox
    ok = (ty >= 0) & (tx >= 0) & (ty <= th - 1) & (tx <= tw - 1)
    y0 = np.floor(ty).astype(np.int64).clip(0, th - 1)
    x0 = np.floor(tx).astype(np.int64).clip(0, tw - 1)
    y1 = np.minimum(y0 + 1, th - 1)
    x1 = np.minimum(x0 + 1, tw - 1)
    fy, fx = ty - y0, tx - x0
    t = tmpl.astype(np.float64)
    v = (1 - fy) * ((1 - fx) * t[y0, x0] + fx * t[y0, x1]) + fy * ((1 - fx) * t[y1, x0] + fx * t[y1, x1])
    region = image[r0:r1, c0:c1]
    region[ok] = v[ok]


@dataclass
class Scene:
    image: np.ndarray                  # u8 (or u16) search image
    template: np.ndarray               # u8 template
    truth: list = field(default_factory=list)  # [(row, col, angle)] centroid convention
    seed: int = 0

    def sha256(self) -> dict:
        return {"image": hashlib.sha256(self.image.tobytes()).hexdigest(),
                "template": hashlib.sha256(self.template.tobytes()).hexdigest()}


def _finish(img: np.ndarray, rng: SplitMix64, sigma: float, gain: bool, bits: int = 8) -> np.ndarray:
    """Gain ramp, N(0, sigma) noise (in 8-bit units) and quantisation to u8,
    or for bits=16 to u16 with the 8-bit range scaled by 257 (so the noise
    fills the low byte)."""
    h, w = img.shape
    if gain:
        img = img * (0.8 + 0.4 * np.arange(w, dtype=np.float64) / (w - 1))[None, :]
    img = img + sigma * rng.normal(h * w).reshape(h, w)
    if bits == 16:
        return np.rint(np.clip(img * 257.0, 0, 65535)).astype(np.uint16)
    return np.rint(np.clip(img, 0, 255)).astype(np.uint8)


def scene(width: int, height: int, tw: int, th: int, count: int, angle_lo: float, angle_hi: float,
          noise: float, gain: bool, seed: int, integer_positions: bool = False,
          min_separation: float | None = None, template_seed: int | None = None, bits: int = 8) -> Scene:
    """Generic scene: ``count`` rotated pastes of one template (drawn from
    ``template_seed``, by default the scene seed); a u16 image for bits=16."""
    rng = SplitMix64(seed)
    base = blurred_noise(rng, height, width).astype(np.float64)
    tseed = seed if template_seed is None else template_seed
    tmpl = make_template(SplitMix64(tseed ^ 0x5EED5EED), th, tw)
    diag = math.hypot(th, tw)
    sep = min_separation if min_separation is not None else 1.5 * diag
    margin = diag / 2 + 2
    truth = []
    tries = 0
    while len(truth) < count and tries < 10000:
        tries += 1
        if integer_positions:
            y0 = int(rng.uniform1() * (height - th + 1))
            x0 = int(rng.uniform1() * (width - tw + 1))
            cy, cx = y0 + (th - 1) / 2.0, x0 + (tw - 1) / 2.0
            ang = 0.0
        else:
            cy = margin + rng.uniform1() * (height - 1 - 2 * margin)
            cx = margin + rng.uniform1() * (width - 1 - 2 * margin)
            ang = angle_lo + rng.uniform1() * (angle_hi - angle_lo)
        if any(math.hypot(cy - r, cx - c) < sep for r, c, _ in truth):
            continue
        truth.append((cy, cx, ang))
    for cy, cx, ang in truth:
        if integer_positions:
            y0, x0 = int(round(cy - (th - 1) / 2.0)), int(round(cx - (tw - 1) / 2.0))
            base[y0:y0 + th, x0:x0 + tw] = tmpl
        else:
            paste(base, tmpl, cy, cx, ang)
    img = _finish(base, rng, noise, gain, bits)
    return Scene(img, tmpl, truth, seed)


# ----------------------------------------------------------------- workloads
def workload_plain_640(seed_k: int = 0) -> Scene:
    """640x480 image, 64x64 template pasted 3x at integer positions, N(0,5)."""
    return scene(640, 480, 64, 64, 3, 0.0, 0.0, 5.0, False, SEED_BASE + seed_k, integer_positions=True)


def workload_rot_2048(seed_k: int = 1, frame: int = 0) -> Scene:
    """2048x2048 image, 128x128 template pasted 5x at subpixel centres with
    rotation in [-30, 30] degrees, gain ramp 0.8-1.2 and N(0,8).  Frames
    k > 0 are further scenes of the same template (other placements, noise)."""
    return scene(2048, 2048, 128, 128, 5, -math.pi / 6, math.pi / 6, 8.0, True, SEED_BASE + seed_k + 1000 * frame,
                 min_separation=1.2 * math.hypot(128, 128), template_seed=SEED_BASE + seed_k)


def workload_full_2048(seed_k: int = 2) -> Scene:
    """2048x2048 image, 256x256 template pasted 5x with rotation in [0, 2pi), N(0,8)."""
    return scene(2048, 2048, 256, 256, 5, 0.0, 2 * math.pi, 8.0, False, SEED_BASE + seed_k,
                 min_separation=1.2 * math.hypot(256, 256))


def small_rot(width=256, height=192, t=32, count=2, extent_deg=20.0, seed_k=10, noise=4.0, frame: int = 0) -> Scene:
    """Small rotated scene for oracle-speed parity tests (frames k > 0: same
    template, other placements and noise)."""
    e = math.radians(extent_deg)
    return scene(width, height, t, t, count, -e, e, noise, True, SEED_BASE + seed_k + 1000 * frame,
                 min_separation=1.3 * math.hypot(t, t), template_seed=SEED_BASE + seed_k)


def small_rot16(width=256, height=192, t=32, count=2, extent_deg=20.0, seed_k=11, noise=4.0) -> Scene:
    """Small rotated scene with a u16 ("uint2") search image."""
    e = math.radians(extent_deg)
    return scene(width, height, t, t, count, -e, e, noise, True, SEED_BASE + seed_k,
                 min_separation=1.3 * math.hypot(t, t), template_seed=SEED_BASE + seed_k, bits=16)


def plain16(width=160, height=128, t=32, count=2, seed_k=12, noise=3.0) -> Scene:
    """u16 scene with pastes at integer positions and angle 0."""
    return scene(width, height, t, t, count, 0.0, 0.0, noise, False, SEED_BASE + seed_k, integer_positions=True,
                 bits=16)


def workload_rot_2048_u16(seed_k: int = 1) -> Scene:
    """The 2048x2048 rotated workload with a u16 search image."""
    return scene(2048, 2048, 128, 128, 5, -math.pi / 6, math.pi / 6, 8.0, True, SEED_BASE + seed_k,
                 min_separation=1.2 * math.hypot(128, 128), template_seed=SEED_BASE + seed_k, bits=16)


@dataclass
class MultiScene:
    image: np.ndarray                                # u8 search image
    templates: list = field(default_factory=list)    # u8 templates
    truth: list = field(default_factory=list)        # [(template index, row, col, angle)]
    seed: int = 0


def multi_scene(width: int, height: int, sizes: list, counts: list, angle_lo: float, angle_hi: float, noise: float,
                seed: int, gain: bool = False) -> MultiScene:
    """Several templates (sizes[k] = (h, w), template seed derived from the
    scene seed and k) pasted counts[k] times each at subpixel centres with
    rotation in [angle_lo, angle_hi]; placements keep 1.2 x the larger
    diagonal between centres."""
    rng = SplitMix64(seed)
    base = blurred_noise(rng, height, width).astype(np.float64)
    tmpls = [make_template(SplitMix64((seed + 7919 * (k + 1)) ^ 0x5EED5EED), h, w) for k, (h, w) in enumerate(sizes)]
    truth = []
    for k, (th, tw) in enumerate(sizes):
        diag = math.hypot(th, tw)
        margin = diag / 2 + 2
        placed, tries = 0, 0
        while placed < counts[k] and tries < 10000:
            tries += 1
            cy = margin + rng.uniform1() * (height - 1 - 2 * margin)
            cx = margin + rng.uniform1() * (width - 1 - 2 * margin)
            ang = angle_lo + rng.uniform1() * (angle_hi - angle_lo)
            if any(math.hypot(cy - r, cx - c) < 0.6 * (diag + math.hypot(*sizes[j]))
                   for j, r, c, _ in truth):
                continue
            truth.append((k, cy, cx, ang))
            placed += 1
    for k, cy, cx, ang in truth:
        paste(base, tmpls[k], cy, cx, ang)
    img = _finish(base, rng, noise, gain)
    return MultiScene(img, tmpls, truth, seed)


def small_multi(seed_k: int = 20) -> MultiScene:
    """Two templates (40x40, 32x48) pasted twice each into 320x256, +-15
    degrees, N(0,4): oracle-speed find_models cases."""
    e = math.radians(15)
    return multi_scene(320, 256, [(40, 40), (32, 48)], [2, 2], -e, e, 4.0, SEED_BASE + seed_k)


@dataclass
class BatchModel:
    template: np.ndarray
    start: float
    extent: float
    step: float       # 0: the automatic step
    levels: int = 0   # 0: automatic


@dataclass
class BatchWorkload:
    models: list                                     # BatchModel
    images: list = field(default_factory=list)       # u8 search images of uneven size
    truth: list = field(default_factory=list)        # per image: [(model, row, col, angle)]


RANGES = ((-math.radians(10), math.radians(20)), (-math.radians(45), math.radians(90)), (0.0, 2 * math.pi))
RASTER_BUDGET = 16_000_000  # samples over all levels and angles (the reference's create budget)


def budget_step(h: int, w: int, extent: float) -> float:
    """0 (the automatic step) when the model fits the raster budget with it,
    else the smallest step of 1/1000 rad granularity whose angle grid keeps
    the rasters (about 4/3 w h samples per angle over the pyramid) within
    0.9 of the budget."""
    auto = max(.01, min(.15, 1 / max(1.0, math.hypot(w, h) / 2)))
    per_angle = 4.0 / 3.0 * w * h
    if (math.ceil(extent / auto) + 1) * per_angle <= 0.9 * RASTER_BUDGET:
        return 0.0
    n = int(0.9 * RASTER_BUDGET / per_angle) - 1
    return math.ceil(extent / max(n, 1) * 1000) / 1000


def batch_levels(h: int, w: int, min_top: int = 256) -> int:
    """The most pyramid levels that keep at least min_top template pixels at
    the top level (a coarser top level makes the coarse search of small
    templates exceed the reference's 10 000-candidate budget)."""
    levels = 1
    while (h >> levels) * (w >> levels) >= min_top and min(h >> levels, w >> levels) >= 8:
        levels += 1
    return levels


def batch_workload(n_models: int = 32, n_images: int = 16, side: tuple = (32, 256),
                   image_sizes: tuple = ((640, 480), (4096, 3072)), max_instances: int = 6, noise: float = 6.0,
                   seed_k: int = 4) -> BatchWorkload:
    """Heterogeneous batch: n_models templates with sides uniform in `side`
    (width and height drawn apart) and angle ranges drawn from +-10 deg,
    +-45 deg and the full turn (an explicit step where the automatic one
    would exceed the raster budget); n_images blurred-noise images with
    width and height uniform between image_sizes[0] and image_sizes[1], each
    holding 0..max_instances rotated pastes of models drawn at random, with
    N(0, noise).  Each model's pyramid levels: batch_levels."""
    rng = SplitMix64(SEED_BASE + seed_k)
    models = []
    for k in range(n_models):
        h = side[0] + int(rng.uniform1() * (side[1] - side[0] + 1))
        w = side[0] + int(rng.uniform1() * (side[1] - side[0] + 1))
        start, extent = RANGES[min(2, int(rng.uniform1() * 3))]
        tmpl = make_template(SplitMix64((SEED_BASE + seed_k + 104729 * (k + 1)) ^ 0x5EED5EED), h, w)
        models.append(BatchModel(tmpl, start, extent, budget_step(h, w, extent), batch_levels(h, w)))
    work = BatchWorkload(models)
    (w0, h0), (w1, h1) = image_sizes
    for j in range(n_images):
        W = w0 + int(rng.uniform1() * (w1 - w0 + 1))
        H = h0 + int(rng.uniform1() * (h1 - h0 + 1))
        base = blurred_noise(rng, H, W).astype(np.float64)
        count = int(rng.uniform1() * (max_instances + 1))
        truth, tries = [], 0
        while len(truth) < count and tries < 1000:
            tries += 1
            k = min(n_models - 1, int(rng.uniform1() * n_models))
            th, tw = models[k].template.shape
            diag = math.hypot(th, tw)
            if diag + 4 >= min(H, W):
                continue
            cy = diag / 2 + 2 + rng.uniform1() * (H - 1 - diag - 4)
            cx = diag / 2 + 2 + rng.uniform1() * (W - 1 - diag - 4)
            ang = models[k].start + rng.uniform1() * models[k].extent
            if any(math.hypot(cy - r, cx - c) < 0.6 * (diag + math.hypot(*models[m].template.shape))
                   for m, r, c, _ in truth):
                continue
            truth.append((k, cy, cx, ang))
        for k, cy, cx, ang in truth:
            paste(base, models[k].template, cy, cx, ang)
        work.images.append(_finish(base, rng, noise, False))
        work.truth.append(truth)
    return work


SERVE_RANGES = (math.radians(15), math.radians(30), math.radians(60), math.radians(90), math.pi)
SERVE_FRAMES = ((1280, 1024), (2448, 2048), (5472, 3648))


def serving_workload(n_models: int = 64, side: tuple = (48, 320), frames_per_size: int = 2,
                     instances: int = 4, noise: float = 6.0, seed_k: int = 5) -> BatchWorkload:
    """The serving pool: n_models square templates with sides uniform in
    `side` and half-ranges drawn from +-15, +-30, +-60, +-90 and +-180 deg
    (the full turn; an explicit step where the automatic one would exceed
    the raster budget, levels as batch_levels), and frames_per_size frames
    of each of 1280x1024, 2448x2048 and 5472x3648, each holding `instances`
    rotated pastes of pool models, N(0, noise)."""
    rng = SplitMix64(SEED_BASE + seed_k)
    models = []
    for k in range(n_models):
        t = side[0] + int(rng.uniform1() * (side[1] - side[0] + 1))
        half = SERVE_RANGES[min(len(SERVE_RANGES) - 1, int(rng.uniform1() * len(SERVE_RANGES)))]
        start, extent = (0.0, 2 * math.pi) if half == math.pi else (-half, 2 * half)
        tmpl = make_template(SplitMix64((SEED_BASE + seed_k + 7907 * (k + 1)) ^ 0x5EED5EED), t, t)
        models.append(BatchModel(tmpl, start, extent, budget_step(t, t, extent), batch_levels(t, t)))
    work = BatchWorkload(models)
    for (W, H) in SERVE_FRAMES:
        for _ in range(frames_per_size):
            base = blurred_noise(rng, H, W).astype(np.float64)
            truth, tries = [], 0
            while len(truth) < instances and tries < 1000:
                tries += 1
                k = min(n_models - 1, int(rng.uniform1() * n_models))
                diag = math.hypot(*models[k].template.shape)
                cy = diag / 2 + 2 + rng.uniform1() * (H - 1 - diag - 4)
                cx = diag / 2 + 2 + rng.uniform1() * (W - 1 - diag - 4)
                ang = models[k].start + rng.uniform1() * models[k].extent
                if any(math.hypot(cy - r, cx - c) < 0.6 * (diag + math.hypot(*models[m].template.shape))
                       for m, r, c, _ in truth):
                    continue
                truth.append((k, cy, cx, ang))
            for k, cy, cx, ang in truth:
                paste(base, models[k].template, cy, cx, ang)
            work.images.append(_finish(base, rng, noise, False))
            work.truth.append(truth)
    return work


# ------------------------------------------------------------- search domains
def mask_runs(mask: np.ndarray, row0: int = 0, col0: int = 0) -> np.ndarray:
    """Canonical runs (n x 3: row, first column, last column) of a boolean
    mask placed at (row0, col0)."""
    out = []
    for r in range(mask.shape[0]):
        m = np.concatenate([[False], mask[r].astype(bool), [False]])
        d = np.diff(m.astype(np.int8))
        starts, ends = np.nonzero(d == 1)[0], np.nonzero(d == -1)[0] - 1
        out += [(row0 + r, col0 + a, col0 + b) for a, b in zip(starts, ends)]
    return np.array(out, dtype=np.int64).reshape(-1, 3)


def rect_runs(r0: int, c0: int, r1: int, c1: int) -> np.ndarray:
    return np.array([(r, c0, c1) for r in range(r0, r1 + 1)], dtype=np.int64).reshape(-1, 3)


def disk_runs(cy: float, cx: float, radius: float) -> np.ndarray:
    """Pixels within `radius` of (cy, cx)."""
    r0, r1 = int(math.floor(cy - radius)), int(math.ceil(cy + radius))
    c0, c1 = int(math.floor(cx - radius)), int(math.ceil(cx + radius))
    yy, xx = np.mgrid[r0:r1 + 1, c0:c1 + 1]
    return mask_runs((yy - cy) ** 2 + (xx - cx) ** 2 <= radius * radius, r0, c0)


def random_domain(rng: SplitMix64, h: int, w: int, n: int, max_len: int = 40) -> np.ndarray:
    """n random runs (unsorted, possibly overlapping) inside an h x w image."""
    rows = (rng.uniform(n) * h).astype(np.int64)
    starts = (rng.uniform(n) * w).astype(np.int64)
    lens = (rng.uniform(n) * max_len).astype(np.int64)
    return np.stack([rows, starts, np.minimum(starts + lens, w - 1)], 1)
