"""The request loop (mvamd_server_*): requests, each a frame with 1-4 models
drawn from a pool, are submitted together; they are admitted by the HBM
budget, micro-batched and run by workers whose device work overlaps.  Each
request's result in the mix equals, exactly, the result of the same request
run alone (mvamd_find_ncc_models_device on that frame and those models),
and a budget that admits one request at a time gives the same results
(docs/runtime.md:31: concurrent finds on shared immutable models)."""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def pool(capi):
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    torch.cuda.set_device(0)
    from mvkit_amd import synth
    w = synth.batch_workload(n_models=6, n_images=6, side=(48, 96), image_sizes=((320, 240), (720, 540)), seed_k=43)
    models = [capi.NccModel(m.template, m.levels, m.start, m.extent, m.step if m.step else "auto") for m in w.models]
    rng = synth.SplitMix64(99)
    requests = []
    for k in range(12):  # every image twice, with different model subsets
        img = w.images[k % len(w.images)]
        n = 1 + int(rng.uniform1() * 4)
        subset = sorted({int(rng.uniform1() * len(models)) for _ in range(n)})
        requests.append((img, subset))
    return w, models, requests


def options(capi, w, subset):
    return [capi.NccModel.options(w.models[i].start, w.models[i].extent, 0.8, 0, 0.5, form=capi.FORM_BF16X3,
                                  max_candidates=100000) for i in subset]


def alone(capi, w, models, img, subset):
    d = torch.from_numpy(img).cuda()
    h, wd = img.shape
    return capi.find_models_device([models[i] for i in subset], d.data_ptr(), wd, h, wd, options(capi, w, subset),
                                   0, 0.5)


@pytest.mark.parametrize("budget,workers,batch", [(0, 2, 4), (1, 2, 4), (0, 3, 1)])
def test_results_same_alone_and_in_the_mix(capi, pool, budget, workers, batch):
    w, models, requests = pool
    say(f"serving: {len(requests)} requests, budget {budget or 'default'}, {workers} workers, batch {batch}")
    want = [alone(capi, w, models, img, subset) for img, subset in requests]
    server = capi.Server(workers=workers, max_batch=batch, hbm_budget=budget)
    try:
        ids = [server.submit(img, [models[i] for i in subset], options(capi, w, subset), 0, 0.5)
               for img, subset in requests]
        got = [server.wait(i) for i in ids]
    finally:
        server.close()
    sizes = [t["batch_size"] for _, t in got]
    say(f"  batch sizes {sizes}, workers {sorted({t['worker'] for _, t in got})}")
    for k, ((res, t), exp) in enumerate(zip(got, want)):
        assert np.array_equal(res, exp), (k, res, exp)
        assert 0 <= t["admitted_us"] <= t["started_us"] <= t["finished_us"]
    if budget == 1:
        assert max(sizes) == 1  # one request's working set exceeds the budget: one in flight at a time
    assert sum(len(r) for r, _ in got) > 0
