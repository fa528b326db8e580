"""Concurrent finds from several host threads on one shared model handle.

The reference's find is reentrant with per-call scratch (matcher.cpp:305)
and a created model is shareable across threads for concurrent finds
(docs/runtime.md:31).  Here every host thread gets its own find context
(streams, workspaces) per device, so finds from several threads -- through
the host-image entry (mv_find_ncc_model), the device-pointer entry and
batches -- must each return exactly what the same request returns alone.
"""
import math
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from mvkit_amd import synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

E = math.radians(20)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    torch.cuda.set_device(0)


def test_threads_match_serial_requests(capi):
    scenes = [synth.small_rot(frame=k) for k in range(6)]
    m = capi.NccModel(scenes[0].template, "auto", -E, 2 * E)
    h, w = scenes[0].image.shape
    dev = [torch.from_numpy(s.image).cuda() for s in scenes]
    torch.cuda.synchronize()
    opts = capi.NccModel.options(-E, 2 * E, 0.6, 0, 0.5)

    def host(k):
        return m.find(scenes[k].image, -E, 2 * E, 0.6, 0, 0.5, "true", 0)

    def device_ptr(k):
        return m.find_device(dev[k].data_ptr(), w, h, w, opts)[0]

    def batch(k):
        ks = [k, (k + 1) % len(scenes)]
        return m.find_batch([dev[j].data_ptr() for j in ks], w, h, w, opts)[0]

    say("serial requests")
    want = {("host", k): host(k) for k in range(len(scenes))}
    want.update({("dev", k): device_ptr(k) for k in range(len(scenes))})
    want.update({("batch", k): batch(k) for k in range(len(scenes))})
    for k in range(len(scenes)):
        # the host entry reports the C ABI's output convention (half-pixel shift)
        assert len(want[("host", k)]) == 2 and len(want[("dev", k)]) == 2
        assert np.array_equal(want[("dev", k)], want[("batch", k)][0])

    # every request three times, interleaved over four threads
    rng = np.random.default_rng(7)
    jobs = [key for key in want for _ in range(3)]
    order = rng.permutation(len(jobs))
    run = {"host": host, "dev": device_ptr, "batch": batch}
    say(f"{len(jobs)} requests on 4 threads")
    with ThreadPoolExecutor(max_workers=4) as ex:
        futures = [(jobs[i], ex.submit(run[jobs[i][0]], jobs[i][1])) for i in order]
        results = [(key, f.result(timeout=120)) for key, f in futures]
    for key, got in results:
        exp = want[key]
        if key[0] == "batch":
            assert len(got) == len(exp)
            for g, x in zip(got, exp):
                assert np.array_equal(g, x), key
        else:
            assert np.array_equal(got, exp), key
