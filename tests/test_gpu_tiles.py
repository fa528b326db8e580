"""Tile shapes of the split-BF16 coarse kernel and the autotuner that picks
them.  Every shape (rows per wave, waves across and down a block, angle
tiles per staged image tile, bound groups) accumulates each window over the
same K steps in the same order, and a tighter bound only drops windows that
can be neither peaks nor suppressors, so the coarse peaks, the candidates
and the matches are those of the default shape exactly, for the split-BF16
and the FP8-screened forms and for the batched persistent launch.  The autotuner times the shapes
on the first find of a kind of search, keeps the fastest in the model, and
later finds use it (or a shape restored from text)."""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

ONE_TILE = [(4, 4, 1, 1, 0), (8, 4, 1, 1, 0), (4, 2, 2, 1, 0), (8, 2, 2, 1, 0), (4, 1, 4, 1, 0), (8, 1, 4, 1, 0),
            (4, 2, 1, 1, 0), (8, 2, 1, 1, 0), (4, 1, 2, 1, 0), (8, 1, 2, 1, 0), (4, 1, 1, 1, 0),
            (4, 4, 1, 1, 8), (4, 4, 1, 1, 16), (4, 4, 1, 1, 32), (8, 2, 2, 1, 16), (4, 1, 4, 1, 8)]
MULTI_TILE = [(4, 4, 1, 2, 0), (4, 4, 1, 4, 0), (8, 2, 2, 2, 0), (8, 1, 4, 4, 0), (4, 2, 1, 8, 0), (8, 4, 1, 3, 0)]
# bound groups on the full turn (179 angles: the last angle tile holds 19,
# so its last group of 8 is partial and a fourth is empty)
MULTI_BOUND = [(4, 4, 1, 1, 8), (4, 4, 1, 2, 8), (8, 2, 2, 4, 16), (4, 1, 4, 8, 32), (4, 2, 2, 1, 32),
               (8, 1, 2, 1, 16)]
START, EXTENT = -math.pi / 6, math.pi / 3


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def scenes(capi):
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    torch.cuda.set_device(0)
    from mvkit_amd import synth
    rot = synth.workload_rot_2048(frame=0)
    full = synth.small_rot(width=320, height=256, t=40, count=2, extent_deg=180.0, seed_k=30)
    return rot, full


def trace(capi, model, image, start, extent, form, tile, min_score=0.7, matches=5):
    o = capi.NccModel.options(start, extent, min_score, matches, 0.5, True, form=form, tile=tile)
    t = model.find_trace(image, o)
    return t


def fields(capi, v):
    return capi.tile_fields(v) + (capi.tile_bound(v),)


def same_trace(a, b):
    assert np.array_equal(a["coarse"], b["coarse"])
    assert np.array_equal(a["prepared"], b["prepared"])
    assert np.array_equal(a["results"], b["results"])


@pytest.mark.parametrize("shape", ONE_TILE)
def test_shape_reproduces_the_default_planes(capi, scenes, shape):
    """2048^2 +-30 deg workload (25 coarse angles, one angle tile)."""
    rot, _ = scenes
    model = capi.NccModel(rot.template, 4, START, EXTENT)
    say(f"rot2048 shape {shape}")
    want = trace(capi, model, rot.image, START, EXTENT, 1, capi.COARSE_TILE_DEFAULT)
    got = trace(capi, model, rot.image, START, EXTENT, 1, capi.coarse_tile(*shape))
    assert len(want["results"]) == 5
    assert fields(capi, got["stats"].coarse_tile) == shape
    same_trace(got, want)


@pytest.mark.parametrize("form", [1, 3])
def test_multi_tile_shapes_reproduce_the_default_planes(capi, scenes, form):
    """Full-turn 40 px template, 2 levels: 179 coarse angles in 6 angle
    tiles, run 1..8 at a time over one staged image tile; split-BF16 and
    FP8-screened forms (the screen's skipped waves join every later tile's
    barriers); bound groups with the split-BF16 form."""
    _, full = scenes
    model = capi.NccModel(full.template, 2, 0.0, 2 * math.pi)
    want = trace(capi, model, full.image, 0.0, 2 * math.pi, form, capi.COARSE_TILE_DEFAULT, matches=2)
    assert len(want["results"]) == 2
    shapes = MULTI_TILE + [(8, 2, 2, 1, 0), (4, 1, 2, 1, 0)] + (MULTI_BOUND if form == 1 else [])
    for shape in shapes:
        say(f"full-turn form {form} shape {shape}")
        got = trace(capi, model, full.image, 0.0, 2 * math.pi, form, capi.coarse_tile(*shape), matches=2)
        assert fields(capi, got["stats"].coarse_tile) == shape
        same_trace(got, want)


def test_autotuner_picks_keeps_and_restores_shapes(capi, scenes):
    rot, _ = scenes
    model = capi.NccModel(rot.template, 4, START, EXTENT)
    d = torch.from_numpy(rot.image).cuda()
    torch.cuda.synchronize()
    opt = lambda tile: capi.NccModel.options(START, EXTENT, 0.7, 5, 0.5, True, form=1, tile=tile)
    want, st = model.find_device(d.data_ptr(), 2048, 2048, 2048, opt(capi.COARSE_TILE_DEFAULT))
    assert st.coarse_tile == capi.coarse_tile(4, 4, 1, 1)
    assert model.tuning == ""
    say("autotune: first find")
    got, st = model.find_device(d.data_ptr(), 2048, 2048, 2048, opt(capi.COARSE_TILE_AUTO))
    assert np.array_equal(got, want)
    lines = model.tuning.splitlines()
    assert len(lines) == 1, lines
    key, rest = lines[0].split(" : ")
    f = dict(kv.split("=") for kv in rest.split())
    shape = tuple(int(f[k]) for k in ("rows", "col_waves", "row_waves", "tiles", "bound"))
    assert fields(capi, st.coarse_tile) == shape
    assert float(f["us"]) <= float(f["default_us"])
    say(f"autotune chose {shape} for {key}: {f['us']} us against {f['default_us']} us")
    for _ in range(2):  # later finds (graph replays) keep the shape
        got, st = model.find_device(d.data_ptr(), 2048, 2048, 2048, opt(capi.COARSE_TILE_AUTO))
        assert np.array_equal(got, want) and fields(capi, st.coarse_tile) == shape
    assert model.tuning.splitlines() == lines
    # a restored table decides the shape without tuning
    other = (8, 1, 4, 1, 16) if shape != (8, 1, 4, 1, 16) else (4, 2, 2, 1, 8)
    model.tuning = (f"{key} : rows={other[0]} col_waves={other[1]} row_waves={other[2]} tiles={other[3]} "
                    f"bound={other[4]} us=1 default_us=2\n")
    got, st = model.find_device(d.data_ptr(), 2048, 2048, 2048, opt(capi.COARSE_TILE_AUTO))
    assert np.array_equal(got, want) and fields(capi, st.coarse_tile) == other
    assert model.tuning.splitlines()[0].startswith(key)
    # another kind of search (one frame of a different size) is tuned apart
    small = np.ascontiguousarray(rot.image[:1024, :1536])
    ds = torch.from_numpy(small).cuda()
    torch.cuda.synchronize()
    model.find_device(ds.data_ptr(), 1536, 1024, 1536, opt(capi.COARSE_TILE_AUTO))
    assert len(model.tuning.splitlines()) == 2


def test_batched_queue_with_tuned_shapes(capi, scenes):
    """The persistent coarse launch over queries whose models carry different
    tuned shapes (different thread counts: separate launches; the same count
    with other block shapes: one launch) returns the default shape's matches."""
    rot, full = scenes
    from mvkit_amd import synth
    s2 = synth.small_rot(width=320, height=256, t=48, count=2, extent_deg=30.0, seed_k=31)
    models = [capi.NccModel(rot.template, 4, START, EXTENT), capi.NccModel(full.template, 3, 0.0, 2 * math.pi),
              capi.NccModel(s2.template, 0, -math.radians(30), math.radians(60))]
    imgs = [np.ascontiguousarray(rot.image[:512, :640]), full.image, s2.image]
    dev = [torch.from_numpy(np.ascontiguousarray(i)).cuda() for i in imgs]
    torch.cuda.synchronize()
    ptrs = [x.data_ptr() for x in dev]
    ws = [i.shape[1] for i in imgs]
    hs = [i.shape[0] for i in imgs]
    kw = dict(form=1, max_candidates=100000)
    opts = lambda tile: [capi.NccModel.options(START, EXTENT, 0.7, 0, 0.5, True, tile=tile, **kw),
                         capi.NccModel.options(0.0, 2 * math.pi, 0.7, 0, 0.5, True, tile=tile, **kw),
                         capi.NccModel.options(-math.radians(30), math.radians(60), 0.7, 0, 0.5, True, tile=tile, **kw)]
    say("batched: default shapes")
    want = capi.find_models_batch(models, ptrs, ws, hs, ws, opts(capi.COARSE_TILE_DEFAULT))
    assert sum(len(w) for w in want) > 0
    say("batched: tuned shapes")
    got = capi.find_models_batch(models, ptrs, ws, hs, ws, opts(capi.COARSE_TILE_AUTO))
    for g, w in zip(got, want):
        assert np.array_equal(g, w)
    assert all(m.tuning for m in models)
    for shapes in ([(8, 2, 2, 1, 0), (4, 1, 2, 4, 8), (4, 4, 1, 2, 0)], [(4, 2, 2, 1, 32), (8, 4, 1, 1, 16),
                                                                        (8, 1, 4, 2, 8)]):
        for m, sh in zip(models, shapes):
            key = m.tuning.splitlines()[0].split(" : ")[0]
            m.tuning = (f"{key} : rows={sh[0]} col_waves={sh[1]} row_waves={sh[2]} tiles={sh[3]} bound={sh[4]} "
                        "us=1 default_us=1\n")
        say(f"batched: shapes {shapes}")
        got = capi.find_models_batch(models, ptrs, ws, hs, ws, opts(capi.COARSE_TILE_AUTO))
        for g, w in zip(got, want):
            assert np.array_equal(g, w)
