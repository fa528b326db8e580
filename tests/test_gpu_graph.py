"""HIP graph replay of the device part of a find (pyramid through full
score): the first batch with a given model, options, frame geometry and
workspace is captured while it runs, later ones replay the graph with the
new frames' descriptors uploaded by its copy node.  Every replay, with other
finds (other options, other geometry, the other batch slot) run eagerly
between the replays, returns exactly what the kernel-by-kernel launch
returns for the same frames."""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

START, EXTENT = -math.pi / 6, math.pi / 3


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def frames(capi):
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    torch.cuda.set_device(0)
    from mvkit_amd import synth
    scenes = [synth.workload_rot_2048(frame=k) for k in range(3)]
    model = capi.NccModel(scenes[0].template, 4, START, EXTENT)
    dev = [torch.from_numpy(s.image).cuda() for s in scenes]
    small = synth.small_rot()
    sm = capi.NccModel(small.template, "auto", -math.radians(20), math.radians(40))
    sdev = torch.from_numpy(small.image).cuda()
    torch.cuda.synchronize()
    return scenes, model, dev, small, sm, sdev


def opts(capi, graph, form=1, min_score=0.7):
    return capi.NccModel.options(START, EXTENT, min_score, 5, 0.5, True, form=form, graph=graph)


def test_graph_replays_equal_eager_launches(capi, frames):
    scenes, model, dev, small, sm, sdev = frames
    want = [model.find_device(d.data_ptr(), 2048, 2048, 2048, opts(capi, False))[0] for d in dev]
    for w in want:
        assert len(w) == 5
    o_small = capi.NccModel.options(-math.radians(20), math.radians(40), 0.7, 2, 0.5, True, form=1)
    want_small = sm.find_device(sdev.data_ptr(), small.image.shape[1], small.image.shape[0], small.image.shape[1],
                                capi.NccModel.options(-math.radians(20), math.radians(40), 0.7, 2, 0.5, True, form=1,
                                                      graph=False))[0]
    for rnd in range(3):
        for k, d in enumerate(dev):
            say(f"graph round {rnd} frame {k}")
            got = model.find_device(d.data_ptr(), 2048, 2048, 2048, opts(capi, True))[0]
            assert np.array_equal(got, want[k]), (rnd, k)
            # eager work between replays: another model and geometry, another form
            g2 = sm.find_device(sdev.data_ptr(), small.image.shape[1], small.image.shape[0], small.image.shape[1],
                                o_small)[0]
            assert np.array_equal(g2, want_small)
            g3 = model.find_device(d.data_ptr(), 2048, 2048, 2048, opts(capi, False, form=0))[0]
            assert len(g3) == 5
        torch.cuda.synchronize()


def test_graph_replays_in_overlapped_batches(capi, frames):
    """Both batch slots capture and replay their own graphs; each batch's
    results equal the eager launch's."""
    scenes, model, dev, small, sm, sdev = frames
    ptrs = [d.data_ptr() for d in dev]
    want, _ = model.find_batch(ptrs, 2048, 2048, 2048, opts(capi, False))
    pending = None
    for step in range(6):
        order = ptrs[step % 3:] + ptrs[:step % 3]  # new descriptors, same geometry
        nxt = model.submit_batch(order, 2048, 2048, 2048, opts(capi, True))
        if pending is not None:
            got, _ = model.wait_batch(pending[0])
            for g, p in zip(got, pending[1]):
                assert np.array_equal(g, want[ptrs.index(p)]), step
        pending = (nxt, order)
    got, _ = model.wait_batch(pending[0])
    for g, p in zip(got, pending[1]):
        assert np.array_equal(g, want[ptrs.index(p)])


def test_host_entry_with_graphs_matches_eager(capi, frames):
    """mv_find_ncc_model (graphs on by default) repeated on host images of
    the same geometry returns what the eager device launch returns."""
    scenes, model, dev, small, sm, sdev = frames
    for rnd in range(2):
        for k, s in enumerate(scenes):
            say(f"host entry round {rnd} frame {k}")
            got = model.find(s.image, START, EXTENT, 0.7, 5, 0.5, "true", 0)
            want = capi.halcon_shift(model.find_device(dev[k].data_ptr(), 2048, 2048, 2048, opts(capi, False))[0])
            # the same poses; the output shift's cos/sin are numpy's here
            assert got.shape == want.shape and np.allclose(got, want, rtol=0, atol=1e-12), (rnd, k)
