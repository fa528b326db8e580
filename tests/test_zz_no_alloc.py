"""The allocation-free row prefix entry point (ncchip_row_prefix_noalloc,
whose workspace is a static device array) gives what ncchip_row_prefix with
a caller-supplied workspace gives, bit for bit, on two calls in a row (the
second smaller, so stale contents of the first would show), for u8 and u16
levels; sizes beyond the static workspace are refused before any launch."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

ARGS_WITH = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
             C.c_void_p, C.c_void_p]
ARGS_WITHOUT = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                C.c_void_p, C.c_void_p]
MAX_WORDS = 1 << 22


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def prefixes(lib, img, noalloc):
    """Row prefixes and support integral of a level given as a numpy array
    (u8 or u16), through one of the two entry points."""
    h, w = img.shape
    kind = 0 if img.dtype == np.uint8 else 1
    src = torch.from_numpy(img.view(np.int16) if kind else img).cuda()
    p1 = torch.full((h, w + 1), -1, dtype=torch.int64, device="cuda")
    p2 = torch.full((h, w + 1), -1, dtype=torch.int64, device="cuda")
    sup = torch.full((h + 1, w + 1), -1, dtype=torch.int32, device="cuda")
    if noalloc:
        st = lib.ncchip_row_prefix_noalloc(src.data_ptr(), kind, w, w, h, src.data_ptr(), 2, p1.data_ptr(),
                                           p2.data_ptr(), sup.data_ptr(), None)
    else:
        nz = torch.zeros((h, w + 1), dtype=torch.int32, device="cuda")
        st = lib.ncchip_row_prefix(src.data_ptr(), kind, w, w, h, src.data_ptr(), 2, p1.data_ptr(), p2.data_ptr(),
                                   nz.data_ptr(), sup.data_ptr(), None)
    assert st == 0, st
    torch.cuda.synchronize()
    return p1.cpu().numpy(), p2.cpu().numpy(), sup.cpu().numpy()


def test_no_alloc_row_prefix_matches(capi):
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    torch.cuda.set_device(0)
    lib = capi.lib()
    lib.ncchip_row_prefix.argtypes = ARGS_WITH
    lib.ncchip_row_prefix_noalloc.argtypes = ARGS_WITHOUT
    rng = np.random.default_rng(77)
    cases = [rng.integers(0, 256, size=(1024, 1023), dtype=np.uint8),   # (w + 1) * h = 2^20
             rng.integers(0, 256, size=(157, 301), dtype=np.uint8),
             rng.integers(0, 65536, size=(96, 513), dtype=np.uint16),
             rng.integers(0, 256, size=(2048, 2047), dtype=np.uint8)]   # exactly the maximum
    for img in cases:
        img[5:9, 3:40] = img.flat[0]  # a flat run: zero support there
    for img in cases:
        say(f"no-alloc row prefix {img.shape} {img.dtype}")
        want = prefixes(lib, img, False)
        got = prefixes(lib, img, True)
        for g, w in zip(got, want):
            assert np.array_equal(g, w)
    # beyond the static workspace: refused, nothing launched
    h, w = 2048, 2048
    dummy = torch.zeros(16, dtype=torch.int64, device="cuda")
    assert (w + 1) * h > MAX_WORDS
    assert lib.ncchip_row_prefix_noalloc(dummy.data_ptr(), 0, w, w, h, dummy.data_ptr(), 0, dummy.data_ptr(),
                                         dummy.data_ptr(), None, None) == -103
