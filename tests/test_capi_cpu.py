"""Host-side checks of the product library that need no GPU: exported symbols,
model construction (bit-exact against the oracle), tuple API, argument
validation and status codes of the C ABI (ncc_matching_api.cpp:29-203)."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in ("mv_ncc.h", "mvkit_amd.h", "ncchip.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        names |= set(re.findall(r"^(?:MV_API\s+)?(?:MvStatus|int|size_t)\s+((?:mv|mvamd|ncchip)_\w+)\s*\(", text, re.M))
    return sorted(names)


def test_library_exports_every_declared_symbol(capi):
    names = declared_symbols()
    assert len(names) >= 30
    lib = capi.lib()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(capi.PUBLIC_SYMBOLS + capi.NCCHIP_SYMBOLS) == set(names)


def compare_models(capi, oracle, tmpl, **kw):
    levels = kw.get("levels", 0)
    start, extent, step = kw.get("start", 0.0), kw.get("extent", 0.0), kw.get("step", 0.0)
    metric = "ignore_global_polarity" if kw.get("absolute") else "use_polarity"
    m = capi.NccModel(tmpl, levels if levels else "auto", start, extent, step if step else "auto", metric)
    o = oracle.Model(tmpl, levels=levels, start=start, extent=extent, step=step, absolute=kw.get("absolute", False))
    assert (m.n_levels, m.n_angles, m.step, m.origin_row, m.origin_col) == (o.n_levels, o.n_angles, o.step, o.origin_row, o.origin_col)
    for l in range(-1, m.n_levels):
        for a in ([0] if l < 0 else range(m.n_angles)):
            A, B = m.raster(l, a), o.patch(l, a)
            for k in ("n", "angle", "orow", "ocol", "rmin", "rmax", "cmin", "cmax", "wsum", "n_runs", "n_wruns"):
                assert A[k] == B[k], (l, a, k, A[k], B[k])
            assert np.array_equal(A["rows"], B["rows"]) and np.array_equal(A["cols"], B["cols"])
            assert np.array_equal(A["w"], B["w"]), (l, a)  # bit-exact weights
    return m


def test_model_bit_exact_unrotated(capi, oracle, scene_plain):
    compare_models(capi, oracle, scene_plain.template, levels=3)


def test_model_bit_exact_rotated(capi, oracle, scene_rot):
    compare_models(capi, oracle, scene_rot.template, levels=4, start=-math.pi / 6, extent=math.pi / 3)


@pytest.mark.parametrize("kw", [
    dict(levels=0, start=-0.3, extent=0.6, absolute=True),
    dict(levels=2, start=1.0, extent=0.5, step=0.05),
    dict(levels=0, start=0.0, extent=2 * math.pi, step=0.1),
])
def test_model_bit_exact_variants(capi, oracle, kw):
    from mvkit_amd import synth
    t = synth.make_template(synth.SplitMix64(42), 37, 45)
    compare_models(capi, oracle, t, **kw)


def test_model_params(capi, scene_plain):
    m = capi.NccModel(scene_plain.template, 3, 0.0, 0.0)
    assert m.params() == [3, 0.0, 0.0, 0.0, "use_polarity"]
    m = capi.NccModel(scene_plain.template, "auto", -0.5, 1.0, 0.02, "ignore_global_polarity")
    lv, a0, ae, st, metric = m.params()
    assert (a0, ae, st, metric) == (-0.5, 1.0, 0.02, "ignore_global_polarity") and lv >= 2


def test_uint16_template(capi, oracle, scene_plain):
    t16 = scene_plain.template.astype(np.uint16) * 3
    m = capi.NccModel(t16, 2, 0.0, 0.0)
    o = oracle.Model(scene_plain.template, levels=2)
    # scaling a template leaves its normalised weights unchanged up to rounding
    assert np.allclose(m.raster(0, 0)["w"], o.patch(0, 0)["w"], atol=1e-14)


def expect_error(capi, fn, status, code=None, operation=None, parameter=None):
    with pytest.raises(capi.MvError) as e:
        fn()
    assert e.value.status == status, e.value
    if code is not None:
        assert e.value.code == code, e.value
    if operation is not None:
        assert e.value.operation == operation, e.value
    if parameter is not None:
        assert e.value.parameter == parameter, e.value
    return e.value


def test_create_errors(capi, scene_plain):
    t = scene_plain.template
    expect_error(capi, lambda: capi.NccModel(np.full((32, 32), 9, np.uint8)), -1, operation="create_ncc_model")
    expect_error(capi, lambda: capi.NccModel(t, 17), -1, code=4, parameter="NumLevels")
    expect_error(capi, lambda: capi.NccModel(t, "many"), -1, code=7, parameter="NumLevels")
    expect_error(capi, lambda: capi.NccModel(t, 3, 0.0, 0.0, "auto", "sad"), -1, code=7, parameter="Metric")
    expect_error(capi, lambda: capi.NccModel(t, 3, 0.0, 0.0, 0.5), -1, code=4, parameter="AngleStep")
    expect_error(capi, lambda: capi.NccModel(t, 3, 0.0, -1.0), -1, operation="create_ncc_model")
    expect_error(capi, lambda: capi.NccModel(t, 9), -1, parameter="levels")
    from mvkit_amd import synth
    big = synth.make_template(synth.SplitMix64(1), 256, 256)
    expect_error(capi, lambda: capi.NccModel(big, 5, 0.0, 2 * math.pi), -103)


def test_find_control_errors_before_device(capi, scene_plain):
    m = capi.NccModel(scene_plain.template, 3, 0.0, 0.0)
    img = capi.Image(scene_plain.image)
    expect_error(capi, lambda: m.find(img, 0.0, 0.0, 1.5, 1), -1, code=4, operation="find_ncc_model", parameter="MinScore")
    expect_error(capi, lambda: m.find(img, 0.0, 0.0, 0.8, -1), -1, parameter="NumMatches")
    expect_error(capi, lambda: m.find(img, 0.0, 0.0, 0.8, 1, 0.5, "maybe"), -1, code=7, parameter="SubPixel")
    expect_error(capi, lambda: m.find(img, 0.0, 7.0, 0.8, 1), -1, parameter="AngleExtent")
    expect_error(capi, lambda: m.find(img, 0.0, 0.0, 0.8, 1, 0.5, "true", [1, 2, 3]), -1, code=2, parameter="NumLevels")
    expect_error(capi, lambda: m.find(img, 0.0, 0.0, "x", 1), -3, code=3, parameter="MinScore")


def test_empty_image_gives_empty_outputs(capi, scene_plain):
    m = capi.NccModel(scene_plain.template, 3, 0.0, 0.0)
    out = m.find(capi.Image(None), 0.0, 0.0, 0.8, 1)
    assert out.shape == (0, 4)


def test_cleared_handle(capi, scene_plain):
    m = capi.NccModel(scene_plain.template, 3, 0.0, 0.0)
    m.clear()
    expect_error(capi, lambda: m.params(), -7, code=8)
    expect_error(capi, lambda: m.find(capi.Image(None), 0.0, 0.0, 0.8, 1), -7)


def test_tuple_api(capi):
    t = capi.Tuple([1, 2.5, "abc"])
    assert t.values() == [1, 2.5, "abc"]
    lib = capi.lib()
    v = C.c_int64()
    st = lib.mv_get_i(t.h, 1, C.byref(v))
    assert st == -3
    st = lib.mv_get_i(t.h, 5, C.byref(v))
    assert st == -2
    clone = C.c_void_p()
    assert lib.mv_copy_tuple(t.h, C.byref(clone)) == 0
    c = capi.Tuple(handle=clone)
    assert c.values() == [1, 2.5, "abc"]
    assert lib.mv_resize_tuple(c.h, 1) == 0 and c.values() == [1]
    d = C.c_double()
    assert lib.mv_get_d(t.h, 0, C.byref(d)) == 0 and d.value == 1.0
    info = capi.ErrorInfo()
    lib.mv_get_i(t.h, 9, C.byref(v))
    lib.mv_get_last_error_info(C.byref(info))
    assert info.status == -2 and info.operation.text() == "get_i"


def test_gen_image_rejects_bad_type(capi):
    a = np.zeros((4, 4), np.uint8)
    args = [capi.Tuple("real"), capi.Tuple(4), capi.Tuple(4), capi.Tuple(int(a.ctypes.data))]
    h = C.c_void_p()
    assert capi.lib().mv_gen_image1(C.byref(h), *[x.h for x in args]) == -6


def test_find_control_fault_order(capi, scene_plain):
    """When several controls are wrong the first fault reported is mvkit's:
    cardinalities in argument order, then per model the ranged reals, the
    count, the keyword and the levels (ncc_matching_api.cpp:126-161)."""
    m = capi.NccModel(scene_plain.template, 3, 0.0, 0.0)
    img = capi.Image(scene_plain.image)
    # NumMatches precedes MaxOverlap among the cardinalities ...
    expect_error(capi, lambda: m.find(img, 0.0, 0.0, 0.8, [1, 2], [0.5, 0.5]), -1, code=2, parameter="NumMatches")
    # ... while MaxOverlap's value is checked before NumMatches' value
    expect_error(capi, lambda: m.find(img, 0.0, 0.0, 0.8, -1, 2.0), -1, code=4, parameter="MaxOverlap")
    # the keyword after the count, the levels last
    expect_error(capi, lambda: m.find(img, 0.0, 0.0, 0.8, -1, 0.5, "maybe"), -1, code=4, parameter="NumMatches")
    expect_error(capi, lambda: m.find(img, 0.0, 0.0, 0.8, 1, 0.5, "maybe", -1), -1, code=7, parameter="SubPixel")
    e = expect_error(capi, lambda: m.find(img, 0.0, 0.0, 0.8, 1, 0.5, "true", [2, -1]), -1, code=4,
                     parameter="NumLevels")
    assert e.parameter_index == 0
    # an integer control given as a real is a type mismatch at its index
    e = expect_error(capi, lambda: m.find(img, 0.0, 0.0, 0.8, 1.0), -3, code=3, parameter="NumMatches")
    assert e.parameter_index == 0


def test_find_models_control_indices(capi, scene_plain):
    m = capi.NccModel(scene_plain.template, 3, 0.0, 0.0)
    img = capi.Image(scene_plain.image)
    e = expect_error(capi, lambda: capi.find_models([m, m], img, 0.0, [0.0, 7.0], 0.8, 1), -1, code=4,
                     parameter="AngleExtent")
    assert e.parameter_index == 1
    e = expect_error(capi, lambda: capi.find_models([m, m], img, 0.0, 0.0, 0.8, 1, 0.5, "true", [1, 2, 3, -4]),
                     -1, code=4, parameter="NumLevels")
    assert e.parameter_index == 2
    expect_error(capi, lambda: capi.find_models([m, m], img, 0.0, 0.0, 0.8, 1, 0.5, "true", [1, 2, 3]), -1, code=2,
                 parameter="NumLevels")


def test_create_control_fault_order(capi, scene_plain):
    t = scene_plain.template
    # NumLevels is read before the angles, AngleStep before Metric
    expect_error(capi, lambda: capi.NccModel(t, 17, math.inf), -1, code=4, parameter="NumLevels")
    expect_error(capi, lambda: capi.NccModel(t, 3, math.inf), -1, code=4, parameter="AngleStart")
    expect_error(capi, lambda: capi.NccModel(t, 3, 0.0, 0.0, "sometimes", "sad"), -1, code=7, parameter="AngleStep")
    expect_error(capi, lambda: capi.NccModel(t, 3.0), -3, code=3, parameter="NumLevels")
    expect_error(capi, lambda: capi.NccModel(t, [3, 4]), -1, code=2, parameter="NumLevels")
    expect_error(capi, lambda: capi.NccModel(t, 3, 0.0, 0.0, "auto", 5), -3, code=3, parameter="Metric")
    expect_error(capi, lambda: capi.NccModel(t, 3, 0.0, 0.0, "auto", "sad", builder=7), -1, code=7,
                 parameter="builder")


def test_extension_pixel_type_checks(capi, scene_plain):
    """Only MVAMD_PIXEL_BYTE and MVAMD_PIXEL_UINT2 are pixel types; the trace
    entry checks the pitch against the pixel size (both before any device
    work)."""
    m = capi.NccModel(scene_plain.template, 3, 0.0, 0.0)
    opts = capi.NccModel.options(0.0, 0.0, 0.8, 1, pixel_type=1)
    expect_error(capi, lambda: m.find_device(0x1000, 64, 64, 64, opts), -6, code=7, parameter="pixel_type")
    opts = capi.NccModel.options(0.0, 0.0, 0.8, 1, form=-2)
    expect_error(capi, lambda: m.find_device(0x1000, 64, 64, 64, opts), -1, parameter="form")
    img = np.zeros((16, 33), np.uint8)
    res = np.zeros((4, 4)); n = C.c_size_t()
    st = capi.FindStats()
    o16 = capi.NccModel.options(0.0, 0.0, 0.8, 1, pixel_type=capi.PIXEL_UINT2)
    lib = capi.lib()
    for w, pitch in ((16, 31), (17, 33)):  # odd pitch, pitch short of a u16 row
        status = lib.mvamd_find_ncc_model_trace(m.tuple.h, img.ctypes.data_as(C.c_void_p), w, 16, pitch,
                                                C.byref(o16), res.ctypes.data_as(C.c_void_p), 4, C.byref(n), None,
                                                0, None, None, None, 0, None, C.byref(st))
        assert status == -1, (w, pitch, status)


def test_coarse_tile_shapes_fit_rule(capi):
    """ncchip_coarse_mfma_fits: rows per wave 4 or 8, waves across and down
    1, 2 or 4 with at most four waves, 1-8 angle tiles per image tile, bound
    groups of 0, 8, 16 or 32 angles (0 only with the FP8 screen), levels up
    to 4 of u8/u16 block sums, and the staged tile within LDS."""
    lib = capi.lib()

    class Tile(C.Structure):
        _fields_ = [("rows", C.c_int32), ("col_waves", C.c_int32), ("row_waves", C.c_int32), ("tiles", C.c_int32),
                    ("bound", C.c_int32)]

    fits = lib.ncchip_coarse_mfma_fits
    fits.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p]
    u16, u32 = 1, 2

    def ok(t, level=3, kind=u16, yspan=19, xc=2, screened=0):
        return fits(level, kind, yspan, xc, screened, C.byref(Tile(*t)) if t else None) == 1

    assert ok(None) and ok((4, 4, 1, 1))
    for t in [(8, 4, 1, 1), (4, 2, 2, 1), (8, 1, 4, 4), (4, 2, 1, 8), (8, 1, 2, 2), (4, 1, 1, 1)]:
        assert ok(t), t
    for t in [(2, 4, 1, 1), (16, 4, 1, 1), (4, 3, 1, 1), (4, 4, 2, 1), (4, 2, 4, 1), (4, 4, 1, 0), (4, 4, 1, 9),
              (4, 0, 1, 1)]:
        assert not ok(t), t
    assert not ok((4, 4, 1, 1), level=5) and not ok((4, 4, 1, 1), kind=u32)
    # the staged tile: 10 bytes per element (11 screened) within LDS
    assert ok((4, 4, 1, 1), yspan=40, xc=4) and not ok((4, 4, 1, 1), yspan=200, xc=20)
    assert ok((4, 1, 1, 1), yspan=120, xc=4) and not ok((8, 1, 4, 1), yspan=120, xc=4)
    for b in (8, 16, 32):
        assert ok((4, 4, 1, 1, b)) and ok((8, 1, 4, 2, b)) and not ok((4, 4, 1, 1, b), screened=1)
    for b in (4, 12, 64, -8):
        assert not ok((4, 4, 1, 1, b)), b
    assert ok((4, 4, 1, 1, 0), screened=1)


def test_model_tuning_text_round_trip(capi, scene_plain):
    """mvamd_set/get_ncc_model_tuning: the text written is read back line by
    line (keys sorted), an empty text clears the table, and a line with an
    invalid shape or no shape is rejected without changing the table."""
    m = capi.NccModel(scene_plain.template, 3, 0.0, 0.0)
    assert m.tuning == ""
    lines = ["L2 K1 A1 Y17 X2 S0 F1 W64 H64 : rows=8 col_waves=2 row_waves=2 tiles=1 bound=16 us=12.5 default_us=20.25",
             "L1 K1 A3 Y33 X3 S1 F16 W128 H128 : rows=4 col_waves=1 row_waves=4 tiles=2 bound=0 us=100 default_us=100"]
    m.tuning = "\n".join(lines) + "\n"
    got = m.tuning.splitlines()
    assert got == sorted(lines), got
    # a line without "bound=" reads as bound 0
    m.tuning = "L3 queue : rows=4 col_waves=2 row_waves=2 tiles=1 us=3 default_us=4\n"
    assert m.tuning == "L3 queue : rows=4 col_waves=2 row_waves=2 tiles=1 bound=0 us=3 default_us=4\n"
    m.tuning = "\n".join(lines) + "\n"
    for bad in ["L2 K1 : rows=2 col_waves=4 row_waves=1 tiles=1 us=1 default_us=1",
                "L2 K1 : rows=4 col_waves=4 row_waves=2 tiles=1 us=1 default_us=1",
                "L2 K1 : rows=4 col_waves=4 row_waves=1 tiles=1 bound=12 us=1 default_us=1",
                "L2 K1 rows=4 col_waves=4 row_waves=1 tiles=1 us=1 default_us=1",
                "L2 K1 : rows=4 col_waves=4"]:
        expect_error(capi, lambda: setattr(m, "tuning", bad), -1, code=4, parameter="text")
        assert m.tuning.splitlines() == sorted(lines)
    m.tuning = ""
    assert m.tuning == ""


def test_coarse_tile_option_checks(capi, scene_plain):
    """coarse_tile in the extension options: -1 (default), 0 (autotuned) or a
    packed shape; anything else is rejected before any device work."""
    m = capi.NccModel(scene_plain.template, 3, 0.0, 0.0)
    for bad in (-2, 1 << 18):
        opts = capi.NccModel.options(0.0, 0.0, 0.8, 1, tile=bad)
        expect_error(capi, lambda: m.find_device(0x1000, 64, 64, 64, opts), -1, parameter="coarse_tile")
    assert capi.coarse_tile(8, 2, 2, 1) == 0x1228 and capi.tile_fields(0x1228) == (8, 2, 2, 1)
    assert capi.coarse_tile(8, 2, 2, 1, 32) == 0x31228 and capi.tile_bound(0x31228) == 32
    assert [capi.tile_bound(capi.coarse_tile(4, 4, 1, 1, b)) for b in (0, 8, 16, 32)] == [0, 8, 16, 32]
