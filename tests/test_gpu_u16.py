"""GPU parity on u16 ("uint2") search images (matcher.cpp:347, 401-433, 489,
565-577, 612, 647, 736; c_api/ncc_matching_api.cpp:167-168): the HIP path
through mv_find_ncc_model / mv_find_ncc_models and the device entries with
MVAMD_PIXEL_UINT2, against the oracle's u16 path.  The integer pyramid of
u16 pixels (u32 block sums up to level 8) is bit-exact; finds meet the same
tolerances as byte images."""
import ctypes as C
import math
import sys

import numpy as np
import pytest

from conftest import ANGLE_TOL, POS_TOL, SCORE_TOL

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
    return torch.device("cuda:0")


def check(got, want, label=""):
    assert got.shape == want.shape, f"{label}: count {len(got)} != {len(want)}\n{got}\n{want}"
    if len(want) == 0:
        return
    assert np.abs(got[:, 0] - want[:, 0]).max() <= POS_TOL, label
    assert np.abs(got[:, 1] - want[:, 1]).max() <= POS_TOL, label
    assert np.abs(got[:, 2] - want[:, 2]).max() <= ANGLE_TOL, label
    assert np.abs(got[:, 3] - want[:, 3]).max() <= SCORE_TOL, label


def on_device(img16):
    """A u16 image in device memory (torch has no unsigned 16-bit arithmetic:
    the bits travel as int16)."""
    return torch.from_numpy(np.ascontiguousarray(img16).view(np.int16)).cuda()


@pytest.fixture(scope="module")
def s16():
    from mvkit_amd import synth
    return synth.small_rot16()


def test_pyramid_u16_bit_exact(capi):
    say("pyramid u16")
    rng = np.random.default_rng(21)
    img = rng.integers(0, 65536, size=(1000, 1300), dtype=np.uint16)
    lib = capi.lib()
    lib.ncchip_pyramid_level.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int,
                                         C.c_int, C.c_int, C.c_void_p]
    prev, prev_kind, cur = on_device(img), 1, img.astype(np.int64)
    for level in range(1, 9):
        h, w = cur.shape
        nh, nw = max(1, h // 2), max(1, w // 2)
        r0 = 2 * np.arange(nh); r1 = np.minimum(r0 + 1, h - 1)
        c0 = 2 * np.arange(nw); c1 = np.minimum(c0 + 1, w - 1)
        cur = cur[r0][:, c0] + cur[r0][:, c1] + cur[r1][:, c0] + cur[r1][:, c1]
        dst = torch.zeros((nh, nw), dtype=torch.int32, device="cuda")
        assert lib.ncchip_pyramid_level(prev.data_ptr(), prev_kind, prev.shape[1], prev.shape[1], prev.shape[0],
                                        dst.data_ptr(), 2, nw, nw, nh, None) == 0
        torch.cuda.synchronize()
        got = dst.cpu().numpy().view(np.uint32).astype(np.int64)
        assert np.array_equal(got, cur), f"level {level}"
        prev, prev_kind = dst, 2


@pytest.mark.parametrize("nm,min_score", [(2, 0.7), (0, 0.5)])
def test_find_u16_matches_oracle(capi, oracle, s16, nm, min_score):
    say(f"find u16 small: nm={nm} min_score={min_score}")
    m = capi.NccModel(s16.template, "auto", -E, 2 * E)
    o = oracle.Model(s16.template, levels=0, start=-E, extent=2 * E)
    want = capi.halcon_shift(o.find(s16.image, -E, 2 * E, min_score, nm, 0.5))
    got = m.find(s16.image, -E, 2 * E, min_score, nm, 0.5, "true", 0)
    check(got, want, "u16 small")
    if nm:
        assert len(got) == 2


@pytest.mark.parametrize("form", [0, 2], ids=["fp32", "fp64"])
@pytest.mark.parametrize("vector", [False, True], ids=["matrix", "vector"])
def test_device_entry_u16(capi, oracle, s16, form, vector):
    say(f"device find u16: form={form} vector_tracking={vector}")
    m = capi.NccModel(s16.template, "auto", -E, 2 * E)
    o = oracle.Model(s16.template, levels=0, start=-E, extent=2 * E)
    want = o.find(s16.image, -E, 2 * E, 0.7, 2, 0.5)
    img = on_device(s16.image)
    h, w = s16.image.shape
    opts = capi.NccModel.options(-E, 2 * E, 0.7, 2, 0.5, form=form, vector_tracking=vector,
                                 pixel_type=capi.PIXEL_UINT2)
    got, _ = m.find_device(img.data_ptr(), w, h, 2 * w, opts)
    check(got, want, "device u16")


def test_affine_u16_equals_byte_find(capi, scene_small):
    """257 * I as u16 finds what I finds as u8."""
    say("affine u16 vs u8")
    m = capi.NccModel(scene_small.template, "auto", -E, 2 * E)
    a = m.find(scene_small.image, -E, 2 * E, 0.7, 0, 0.5, "true", 0)
    b = m.find(scene_small.image.astype(np.uint16) * 257, -E, 2 * E, 0.7, 0, 0.5, "true", 0)
    assert a.shape == b.shape and len(a) == 2
    assert np.abs(a - b).max() < 1e-9


def test_rot2048_u16_matches_oracle(capi, oracle):
    from mvkit_amd import synth
    s = synth.workload_rot_2048_u16()
    a, x = -math.pi / 6, math.pi / 3
    say("find u16 rot2048")
    m = capi.NccModel(s.template, 4, a, x)
    o = oracle.Model(s.template, levels=4, start=a, extent=x)
    want = capi.halcon_shift(o.find(s.image, a, x, 0.7, 5, 0.5))
    got = m.find(s.image, a, x, 0.7, 5, 0.5, "true", 0)
    check(got, want, "u16 rot2048")
    assert len(got) == 5


def test_six_levels_u16(capi, oracle):
    """Top level 5: the coarse FP32 form cannot hold S - C of u16 block sums
    exactly in FP32 there, so the find scores the top level in FP64."""
    from mvkit_amd import synth
    s = synth.scene(1024, 1024, 256, 256, 2, -0.15, 0.15, 6.0, False, 777, min_separation=400, bits=16)
    say("find u16 six levels")
    m = capi.NccModel(s.template, 6, -0.15, 0.3)
    o = oracle.Model(s.template, levels=6, start=-0.15, extent=0.3)
    want = capi.halcon_shift(o.find(s.image, -0.15, 0.3, 0.7, 0, 0.5))
    got = m.find(s.image, -0.15, 0.3, 0.7, 0, 0.5, "true", 0)
    check(got, want, "u16 six levels")
    assert len(got) == 2


def test_find_models_u16_matches_oracle(capi, oracle):
    from mvkit_amd import synth
    s = synth.small_multi()
    rng = np.random.default_rng(5)
    img = (s.image.astype(np.uint16) * 256 + rng.integers(0, 256, size=s.image.shape, dtype=np.uint16))
    e = math.radians(16)
    c = [capi.NccModel(t, 3, -e, 2 * e) for t in s.templates]
    o = [oracle.Model(t, levels=3, start=-e, extent=2 * e) for t in s.templates]
    for nm, ov, sn, so in [(0, 0.5, 0, 0.5), ([2, 2], [0.5, 0.5], -1, -1.0)]:
        say(f"find_models u16 nm={nm}")
        got = capi.find_models(c, img, -e, 2 * e, 0.7, nm, ov, "true", 0)
        opts = [dict(start=-e, extent=2 * e, min_score=0.7, max_overlap=0.5, num_matches=nm if sn == 0 else 2)] * 2
        want = oracle.find_models(o, img, opts, sn, so)
        want[:, :4] = capi.halcon_shift(want[:, :4])
        assert got.shape == want.shape and len(got) == 4
        assert np.array_equal(got[:, 4], want[:, 4])
        check(got[:, :4], want[:, :4], f"find_models u16 {nm}")


def test_batch_u16_equals_single(capi, s16):
    from mvkit_amd import synth
    frames = [s16.image, synth.small_rot16(seed_k=13).image]
    m = capi.NccModel(s16.template, "auto", -E, 2 * E)
    h, w = s16.image.shape
    opts = capi.NccModel.options(-E, 2 * E, 0.7, 2, 0.5, pixel_type=capi.PIXEL_UINT2)
    dev = [on_device(f) for f in frames]
    say("batch u16")
    batch, _ = m.find_batch([d.data_ptr() for d in dev], w, h, 2 * w, opts)
    for d, b in zip(dev, batch):
        single, _ = m.find_device(d.data_ptr(), w, h, 2 * w, opts)
        assert np.array_equal(single, b)


def test_u16_rejections(capi, s16):
    m = capi.NccModel(s16.template, "auto", -E, 2 * E)
    img = on_device(s16.image)
    h, w = s16.image.shape
    say("u16 rejections")
    opts = capi.NccModel.options(-E, 2 * E, 0.7, 2, 0.5, form=capi.FORM_BF16X3, pixel_type=capi.PIXEL_UINT2)
    with pytest.raises(capi.MvError) as e:
        m.find_device(img.data_ptr(), w, h, 2 * w, opts)
    assert e.value.status == -6
    opts = capi.NccModel.options(-E, 2 * E, 0.7, 2, 0.5, pixel_type=capi.PIXEL_UINT2)
    with pytest.raises(capi.MvError) as e:  # pitch shorter than a u16 row
        m.find_device(img.data_ptr(), w, h, w, opts)
    assert e.value.status == -1
