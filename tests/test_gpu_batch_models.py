"""Batched heterogeneous find on the GPU: several models of uneven template
size and angle range (+-10 deg, +-45 deg, full turn) over several images of
uneven size in one call (mvamd_find_ncc_models_batch): every query's
front half, one persistent split-BF16 coarse launch over the (image, angle
tile, block) work queue of all queries, then every query's back half and
the images' greedy selections with shared refinement rounds.

Each image's matches are compared with the oracle's find_models on that
image (instance count and order, model index, row/col within 0.02 px,
angle within 0.01 deg, score within 1e-4; matcher.cpp:1077-1124,
multi.hpp:79-119); the persistent launch gives exactly what one launch per
query gives, and the batch exactly what the single-image entry gives."""
import math
import sys

import numpy as np
import pytest

from conftest import ANGLE_TOL, POS_TOL, SCORE_TOL

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

MIN_SCORE = 0.8


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def batch(capi, oracle):
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    torch.cuda.set_device(0)
    from mvkit_amd import synth
    w = synth.batch_workload(n_models=8, n_images=4, side=(48, 112), image_sizes=((480, 360), (960, 720)), seed_k=41)
    say("batched models: " + ", ".join(f"{m.template.shape} {math.degrees(m.extent):.0f}deg L{m.levels}"
                                       for m in w.models))
    models = [capi.NccModel(m.template, m.levels, m.start, m.extent, m.step if m.step else "auto") for m in w.models]
    om = [oracle.Model(m.template, levels=m.levels, start=m.start, extent=m.extent, step=m.step) for m in w.models]
    opts = [capi.NccModel.options(m.start, m.extent, MIN_SCORE, 0, 0.5, form=capi.FORM_BF16X3) for m in w.models]
    dev = [torch.from_numpy(im).cuda() for im in w.images]
    torch.cuda.synchronize()
    return w, models, om, opts, dev


def check(got, want, label):
    assert got.shape == want.shape, f"{label}: {len(got)} matches, oracle {len(want)}\n{got}\n{want}"
    if not len(want):
        return
    assert np.array_equal(got[:, 4], want[:, 4]), f"{label}: model order\n{got}\n{want}"
    assert np.abs(got[:, :2] - want[:, :2]).max() <= POS_TOL, label
    assert np.abs(got[:, 2] - want[:, 2]).max() <= ANGLE_TOL, label
    assert np.abs(got[:, 3] - want[:, 3]).max() <= SCORE_TOL, label


def run(capi, w, models, opts, dev, queue=True):
    shapes = [im.shape for im in w.images]
    return capi.find_models_batch(models, [d.data_ptr() for d in dev], [s[1] for s in shapes],
                                  [s[0] for s in shapes], [s[1] for s in shapes], opts, 0, 0.5, coarse_queue=queue)


def test_batched_models_match_oracle(capi, oracle, batch):
    w, models, om, opts, dev = batch
    say("batched models: one persistent coarse launch over 8 models x 4 images")
    got = run(capi, w, models, opts, dev)
    oo = [dict(start=m.start, extent=m.extent, min_score=MIN_SCORE, max_overlap=0.5, num_matches=0) for m in w.models]
    found = 0
    for k, im in enumerate(w.images):
        say(f"  image {k} {im.shape}: oracle find_models")
        want = oracle.find_models(om, im, oo, selection_matches=0, selection_overlap=0.5)
        check(got[k], want, f"image {k}")
        found += len(want)
    assert found == sum(len(t) for t in w.truth)  # every pasted instance, and nothing else


def test_persistent_launch_equals_per_query_launches(capi, batch):
    w, models, om, opts, dev = batch
    say("batched models: persistent coarse launch vs one launch per query")
    a = run(capi, w, models, opts, dev, queue=True)
    b = run(capi, w, models, opts, dev, queue=False)
    for k in range(len(a)):
        assert np.array_equal(a[k], b[k]), k


def test_batch_equals_single_image_calls(capi, batch):
    w, models, om, opts, dev = batch
    say("batched models: batch vs single-image calls")
    a = run(capi, w, models, opts, dev)
    for k, (im, d) in enumerate(zip(w.images, dev)):
        h, wd = im.shape
        single = capi.find_models_device(models, d.data_ptr(), wd, h, wd, opts, 0, 0.5)
        assert np.array_equal(a[k], single), k
