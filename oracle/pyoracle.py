"""ctypes binding of the CPU oracle (test infrastructure only).

Imported by tests/, the smoke check and the benchmark's CPU-baseline leg --
never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libncc_oracle.so")
LIB_AVX2 = os.path.join(HERE, "libncc_oracle_avx2.so")  # the same source built for AVX2 hosts (CPU baseline)


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(os.path.join(HERE, "ncc_oracle.cpp")):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "all"], check=True)
    return LIB


class FindOut(C.Structure):
    _fields_ = [("n_result", C.c_int), ("n_coarse", C.c_int), ("n_prepared", C.c_int), ("top", C.c_int),
                ("finest", C.c_int), ("angle_stride", C.c_int), ("coarse_windows", C.c_double),
                ("coarse_samples", C.c_double), ("track_windows", C.c_double), ("track_samples", C.c_double),
                ("crop_r", C.c_long), ("crop_c", C.c_long), ("store_r", C.c_long), ("store_c", C.c_long),
                ("store_h", C.c_long), ("store_w", C.c_long)]


_libs = {}


def host_has_avx2() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            return any(line.startswith("flags") and " avx2" in line for line in f)
    except OSError:
        return False


def lib(avx2: bool = False):
    key = "avx2" if avx2 else "scalar"
    if key not in _libs:
        path = LIB_AVX2 if avx2 else LIB
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        P = C.c_void_p
        D = C.c_double
        I = C.c_int
        L.orc_last_error.restype = C.c_char_p
        L.orc_create.argtypes = [P, I, I, I, I, D, D, D, I, P, I, I, C.POINTER(P)]
        L.orc_free.argtypes = [P]
        L.orc_model_info.argtypes = [P, P]
        L.orc_level_info.argtypes = [P, I, P]
        L.orc_patch_info.argtypes = [P, I, I, P]
        L.orc_patch_samples.argtypes = [P, I, I, P, P, P]
        L.orc_level_gray.argtypes = [P, I, P]
        L.orc_pyramid.argtypes = [P, I, I, I, I, P]
        L.orc_row_prefix.argtypes = [P, I, I, P, P]
        L.orc_find.argtypes = [P, P, I, I, I, I, P, I, I, D, D, D, I, D, I, I, I, P, I, P, I, P, I, P, C.POINTER(FindOut)]
        L.orc_zncc.argtypes = [P, P, I, I]
        L.orc_zncc.restype = D
        L.orc_overlap.argtypes = [D, D, D, D, D, D, I, I, D, D]
        L.orc_overlap.restype = D
        L.orc_find_tracked.argtypes = [P, P, I, I, I, D, D, D, I, D, I, I, I, P, I, P, I, C.POINTER(I)]
        L.orc_merge_identical.argtypes = [P, I, P]
        L.orc_continuous.argtypes = [P, P, I, I, I, D, D, D]
        L.orc_continuous.restype = D
        L.orc_refine_one.argtypes = [P, P, I, I, I, D, D, D, I, D, I, P]
        L.orc_find_models.argtypes = [P, I, P, I, I, I, I, P, I, I, P, C.c_int64, D, I, P, I, P, P, I, P]
        L.orc_box_overlap.argtypes = [D, D, D, P, D, D, D, P]
        L.orc_box_overlap.restype = D
        L.orc_serialize.argtypes = [P, P, C.c_size_t, C.POINTER(C.c_size_t)]
        L.orc_deserialize.argtypes = [P, C.c_size_t, C.POINTER(P)]
        L.orc_domain_runs.argtypes = [P, I, I, I, P, I]
        L.orc_checksum.argtypes = [P, C.c_size_t]
        L.orc_checksum.restype = C.c_uint64
        _libs[key] = L
    return _libs[key]


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class OracleError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"status {status}: {message}")
        self.status = status


class Model:
    def __init__(self, template: np.ndarray, levels: int = 0, start: float = 0.0, extent: float = 0.0,
                 step: float = 0.0, absolute: bool = False, domain=None, avx2: bool = False):
        """domain: None (the whole template) or canonical runs (n x 3: row,
        first column, last column) of the template's reduced domain.  avx2:
        use the AVX2 build of the oracle (same results)."""
        self._L = lib(avx2)
        t = np.ascontiguousarray(template, dtype=np.uint8)
        runs, n_runs, reduced = _domain(domain)
        h = C.c_void_p()
        st = self._L.orc_create(_ptr(t), t.shape[1], t.shape[0], t.shape[1], levels, start, extent, step,
                              int(absolute), _ptr(runs), n_runs, reduced, C.byref(h))
        if st:
            raise OracleError(st, self._L.orc_last_error().decode())
        self.h = h
        info = np.zeros(6)
        self._L.orc_model_info(self.h, _ptr(info))
        self.n_levels, self.n_angles = int(info[0]), int(info[1])
        self.step, self.origin_row, self.origin_col, self.sample_count = info[2], info[3], info[4], int(info[5])

    @classmethod
    def deserialize(cls, data: bytes) -> "Model":
        h = C.c_void_p()
        st = lib().orc_deserialize(data, len(data), C.byref(h))
        if st:
            raise OracleError(st, lib().orc_last_error().decode())
        m = cls.__new__(cls)
        m._L = lib()
        m.h = h
        info = np.zeros(6)
        m._L.orc_model_info(m.h, _ptr(info))
        m.n_levels, m.n_angles = int(info[0]), int(info[1])
        m.step, m.origin_row, m.origin_col, m.sample_count = info[2], info[3], info[4], int(info[5])
        return m

    def serialize(self) -> bytes:
        n = C.c_size_t()
        st = self._L.orc_serialize(self.h, None, 0, C.byref(n))
        if st:
            raise OracleError(st, self._L.orc_last_error().decode())
        buf = C.create_string_buffer(n.value)
        self._L.orc_serialize(self.h, buf, n.value, C.byref(n))
        return buf.raw[:n.value]

    def __del__(self):
        if getattr(self, "h", None):
            self._L.orc_free(self.h)
            self.h = None

    def level(self, l: int) -> dict:
        out = np.zeros(4)
        self._L.orc_level_info(self.h, l, _ptr(out))
        w, h = int(out[0]), int(out[1])
        g = np.zeros(w * h)
        self._L.orc_level_gray(self.h, l, _ptr(g))
        return {"w": w, "h": h, "orow": out[2], "ocol": out[3], "gray": g.reshape(h, w)}

    def patch(self, level: int, angle: int) -> dict:
        hdr = np.zeros(11)
        self._L.orc_patch_info(self.h, level, angle, _ptr(hdr))
        n = int(hdr[0])
        rows = np.zeros(n, np.int32)
        cols = np.zeros(n, np.int32)
        w = np.zeros(n)
        self._L.orc_patch_samples(self.h, level, angle, _ptr(rows), _ptr(cols), _ptr(w))
        return {"n": n, "angle": hdr[1], "orow": hdr[2], "ocol": hdr[3], "rmin": int(hdr[4]), "rmax": int(hdr[5]),
                "cmin": int(hdr[6]), "cmax": int(hdr[7]), "wsum": hdr[8], "n_runs": int(hdr[9]),
                "n_wruns": int(hdr[10]), "rows": rows, "cols": cols, "w": w}

    def find(self, image: np.ndarray, start: float, extent: float, min_score: float, num_matches: int,
             max_overlap: float, subpixel: bool = True, highest: int = 0, lowest: int = 1, trace: bool = False,
             domain=None):
        """find on a u8 or u16 image (the dtype selects the pixel type);
        domain: None (full image) or runs (n x 3: row, first column, last
        column) the image's domain is reduced to."""
        img = search_image(image)
        runs, n_runs, reduced = _domain(domain)
        cap = 4096
        res = np.zeros((cap, 4))
        coarse = np.zeros((20000, 4)) if trace else None
        prep = np.zeros((20000, 4)) if trace else None
        refd = np.zeros((20000, 5)) if trace else None
        info = FindOut()
        st = self._L.orc_find(self.h, _ptr(img), img.shape[1], img.shape[0], img.strides[0], img.itemsize,
                            _ptr(runs), n_runs, reduced, start, extent, min_score,
                            num_matches, max_overlap, int(subpixel), highest, lowest, _ptr(res), cap,
                            _ptr(coarse) if trace else None, 20000, _ptr(prep) if trace else None, 20000,
                            _ptr(refd) if trace else None, C.byref(info))
        if st:
            raise OracleError(st, self._L.orc_last_error().decode())
        out = res[:info.n_result].copy()
        if not trace:
            return out
        return out, {"coarse": coarse[:info.n_coarse].copy(), "prepared": prep[:info.n_prepared].copy(),
                     "refined": refd[:info.n_prepared].copy(), "info": info}

    def tracked(self, image: np.ndarray, start: float, extent: float, min_score: float, num_matches: int,
                max_overlap: float, subpixel: bool = True, highest: int = 0, lowest: int = 1) -> list:
        """The candidate list after each tracking level of a find on a u8
        image (each in (row, col, angle) order)."""
        img = np.ascontiguousarray(image, dtype=np.uint8)
        cap, lists = 100000, 16
        out = np.zeros((cap, 4))
        counts = np.zeros(lists, np.int32)
        n = C.c_int()
        st = self._L.orc_find_tracked(self.h, _ptr(img), img.shape[1], img.shape[0], img.shape[1], start, extent,
                                    min_score, num_matches, max_overlap, int(subpixel), highest, lowest, _ptr(out),
                                    cap, _ptr(counts), lists, C.byref(n))
        if st:
            raise OracleError(st, self._L.orc_last_error().decode())
        res, k = [], 0
        for i in range(n.value):
            res.append(out[k:k + counts[i]].copy())
            k += counts[i]
        return res

    def continuous(self, image: np.ndarray, row: float, col: float, angle: float) -> float:
        img = np.ascontiguousarray(image, dtype=np.uint8)
        return self._L.orc_continuous(self.h, _ptr(img), img.shape[1], img.shape[0], img.shape[1], row, col, angle)


def pyramid(image: np.ndarray, levels: int) -> list:
    img = np.ascontiguousarray(image, dtype=np.uint8)
    h, w = img.shape
    dims = []
    for _ in range(1, levels):
        w, h = max(1, w // 2), max(1, h // 2)
        dims.append((h, w))
    out = np.zeros(sum(a * b for a, b in dims))
    st = lib().orc_pyramid(_ptr(img), img.shape[1], img.shape[0], img.shape[1], levels, _ptr(out))
    if st:
        raise OracleError(st, lib().orc_last_error().decode())
    res, off = [], 0
    for a, b in dims:
        res.append(out[off:off + a * b].reshape(a, b))
        off += a * b
    return res


def row_prefix(level: np.ndarray):
    v = np.ascontiguousarray(level, dtype=np.float64)
    h, w = v.shape
    s = np.zeros((h, w + 1))
    q = np.zeros((h, w + 1))
    lib().orc_row_prefix(_ptr(v), w, h, _ptr(s), _ptr(q))
    return s, q


def zncc(a: np.ndarray, b: np.ndarray, absolute: bool = False) -> float:
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return lib().orc_zncc(_ptr(a), _ptr(b), len(a), int(absolute))


def overlap(a, b, w, h, orow, ocol) -> float:
    return lib().orc_overlap(a[0], a[1], a[2], b[0], b[1], b[2], w, h, orow, ocol)


def _domain(domain):
    if domain is None:
        return np.zeros((1, 3), np.int64), 0, 0
    runs = np.ascontiguousarray(np.asarray(domain, dtype=np.int64).reshape(-1, 3))
    return (runs if len(runs) else np.zeros((1, 3), np.int64)), len(runs), 1


def domain_runs(runs, h: int, w: int) -> np.ndarray:
    """The oracle's canonical form of a domain given by runs, clipped to an
    h x w image (n x 3: row, first column, last column)."""
    r, n, _ = _domain(runs)
    cap = max(1, 2 * n + 16)
    out = np.zeros((cap, 3), np.int64)
    k = lib().orc_domain_runs(_ptr(r), n, h, w, _ptr(out), cap)
    if k < 0:
        raise OracleError(k, lib().orc_last_error().decode())
    return out[:k].copy()


def search_image(image: np.ndarray) -> np.ndarray:
    """A contiguous u16 array for u16 input, else u8."""
    return np.ascontiguousarray(image, dtype=np.uint16 if image.dtype == np.uint16 else np.uint8)


def find_models(models: list, image: np.ndarray, opts: list, selection_matches: int = -1,
                selection_overlap: float = -1.0, per_model_quota: bool = True, prepared: bool = False, domain=None):
    """find_models over oracle models.  opts: one dict per model with keys
    start, extent, min_score, max_overlap, num_matches, subpixel, highest,
    lowest and optionally max_candidates (the coarse candidate budget, 10 000
    by default).  Returns (n, 5): row, col, angle, score (internal convention),
    model index; with prepared=True also each model's prepared list.  The
    image is u8 or u16 (by dtype)."""
    img = search_image(image)
    runs, n_runs, reduced = _domain(domain)
    n = len(models)
    handles = (C.c_void_p * n)(*[m.h.value for m in models])
    o = np.zeros((n, 9))
    for i, q in enumerate(opts):
        o[i] = [q["start"], q["extent"], q["min_score"], q["max_overlap"], q["num_matches"], int(q.get("subpixel", True)),
                q.get("highest", 0), q.get("lowest", 1), q.get("max_candidates", 0)]
    cap, cap_p = 4096, 20000
    res = np.zeros((cap, 5))
    nres = C.c_int()
    prep = np.zeros((n, cap_p, 4)) if prepared else None
    prep_n = np.zeros(n, np.int32) if prepared else None
    st = lib().orc_find_models(handles, n, _ptr(img), img.shape[1], img.shape[0], img.strides[0], img.itemsize,
                               _ptr(runs), n_runs, reduced, _ptr(o),
                               int(selection_matches), float(selection_overlap), int(per_model_quota), _ptr(res), cap,
                               C.byref(nres), _ptr(prep) if prepared else None, cap_p,
                               _ptr(prep_n) if prepared else None)
    if st:
        raise OracleError(st, lib().orc_last_error().decode())
    out = res[:nres.value].copy()
    if not prepared:
        return out
    return out, [prep[i, :prep_n[i]].copy() for i in range(n)]


def merge_identical(cand: np.ndarray) -> np.ndarray:
    """The oracle's merge_identical of an (n, 4) candidate list."""
    c = np.ascontiguousarray(cand, dtype=np.float64).reshape(-1, 4)
    out = np.zeros((max(len(c), 1), 4))
    k = lib().orc_merge_identical(_ptr(c), len(c), _ptr(out))
    return out[:k].copy()


def box_overlap(a, ga, b, gb) -> float:
    """Multi-model overlap of poses a, b (row, col, angle) with template
    geometries ga, gb = (w, h, orow, ocol)."""
    ga = np.asarray(ga, dtype=np.float64)
    gb = np.asarray(gb, dtype=np.float64)
    return lib().orc_box_overlap(a[0], a[1], a[2], _ptr(ga), b[0], b[1], b[2], _ptr(gb))


def checksum(data: bytes) -> int:
    """FNV-1a-64 of the model body (persist.cpp:18-21)."""
    return lib().orc_checksum(data, len(data))
