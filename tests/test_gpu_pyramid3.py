"""ncchip_pyramid3_u8 (levels 1-3 of byte frames in one launch) against the
exact 2x2 block sums with the edge clamp (matcher.cpp:436-454;
gray_pyramid.hpp:8-13), for frames whose tiles take the register path (inside
the frame, 16-byte aligned rows) and the general LDS path (edge tiles,
unaligned pitches), several frames of different sizes in one launch."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


class Level(C.Structure):
    _fields_ = [("data", C.c_void_p), ("kind", C.c_int32), ("pitch", C.c_int32), ("width", C.c_int32),
                ("height", C.c_int32), ("shift", C.c_int32), ("pad", C.c_int32)]


class Frame(C.Structure):  # NccFrame (device_types.h)
    _fields_ = [("levels", Level * 16), ("dc_src", C.c_void_p), ("dc_wide", C.c_int32), ("storage0", C.c_int32 * 4),
                ("crop_r", C.c_double), ("crop_c", C.c_double), ("image_h", C.c_int32), ("image_w", C.c_int32),
                ("p1", C.c_void_p), ("p2", C.c_void_p), ("nz", C.c_void_p), ("planes", C.c_void_p),
                ("peak_bits", C.c_void_p), ("ranges", C.c_void_p), ("dom_sat", C.c_void_p), ("dom_runs", C.c_void_p),
                ("dom_row", C.c_void_p), ("dom_box", C.c_int32 * 4), ("dom_sparse", C.c_int32), ("dom_pad", C.c_int32)]


def block_sums(img, levels):
    out, cur = [], img.astype(np.int64)
    for _ in range(levels):
        h, w = cur.shape
        nh, nw = max(1, h // 2), max(1, w // 2)
        r0 = 2 * np.arange(nh); r1 = np.minimum(r0 + 1, h - 1)
        c0 = 2 * np.arange(nw); c1 = np.minimum(c0 + 1, w - 1)
        cur = cur[r0][:, c0] + cur[r0][:, c1] + cur[r1][:, c0] + cur[r1][:, c1]
        out.append(cur)
    return out


CASES = [
    # (frames as (height, width, pitch, level rows 16-byte aligned), levels)
    ([(2048, 2048, 2048, True)], 3),
    ([(2048, 2048, 2048, True), (1080, 1920, 1920, True), (480, 640, 640, True)], 3),
    ([(1024, 4096, 4096, True), (37, 53, 64, True), (100, 2050, 2056, True)], 3),
    ([(512, 1024, 1040, False)], 3),   # level-0 rows aligned, level rows not: the general path
    ([(512, 1024, 1025, True)], 2),    # unaligned level-0 pitch: the general path
    ([(2048, 2048, 2048, True), (64, 1024, 1024, True)], 1),
]


@pytest.mark.parametrize("frames,levels", CASES)
def test_pyramid3_u8_bit_exact(capi, frames, levels):
    print(f"pyramid3 {frames} levels {levels}", flush=True)
    lib = capi.lib()
    lib.ncchip_pyramid3_u8.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p]
    rng = np.random.default_rng(17)
    keep, want, descs = [], [], (Frame * len(frames))()
    for f, (h, w, pitch, aligned) in enumerate(frames):
        img = rng.integers(0, 256, size=(h, w), dtype=np.uint8)
        buf = np.zeros((h, pitch), dtype=np.uint8)
        buf[:, :w] = img
        src = torch.from_numpy(buf).cuda()
        keep.append(src)
        want.append(block_sums(img, levels))
        d = descs[f]
        d.levels[0] = Level(src.data_ptr(), 0, pitch, w, h, 0, 0)
        for l in range(1, levels + 1):
            lh, lw = want[-1][l - 1].shape
            lp = (lw + 7) // 8 * 8 + 8 if aligned else lw + 3
            dst = torch.zeros((lh, lp), dtype=torch.int16, device="cuda")
            keep.append(dst)
            d.levels[l] = Level(dst.data_ptr(), 1, lp, lw, lh, 2 * l, 0)
    dev = torch.from_numpy(np.frombuffer(bytes(descs), dtype=np.uint8).copy()).cuda()
    torch.cuda.synchronize()
    max_w, max_h = max(f[1] for f in frames), max(f[0] for f in frames)
    assert lib.ncchip_pyramid3_u8(dev.data_ptr(), len(frames), levels, max_w, max_h, None) == 0
    torch.cuda.synchronize()
    for f in range(len(frames)):
        for l in range(1, levels + 1):
            lv = descs[f].levels[l]
            got = next(t for t in keep if t.data_ptr() == lv.data).cpu().numpy().view(np.uint16)[:, :lv.width]
            assert np.array_equal(got.astype(np.int64), want[f][l - 1]), (frames[f], l)
