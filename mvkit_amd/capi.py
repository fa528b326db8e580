"""ctypes binding of libmvkit_amd.so: the reference-compatible C ABI
(include/mv_ncc.h) and the device/trace extensions (include/mvkit_amd.h).

The NCC operators run on the GPU; importing this module only loads the
library and fails loudly when it has not been built.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmvkit_amd.so")

MV_VALUE_EMPTY, MV_VALUE_INT64, MV_VALUE_REAL64, MV_VALUE_STRING, MV_VALUE_HANDLE = 0, 1, 2, 4, 16
FORM_AUTO, FORM_FP32, FORM_BF16X3, FORM_FP64, FORM_FP8_SCREEN = -1, 0, 1, 2, 3
PIXEL_BYTE, PIXEL_UINT2 = 0, 2  # device image pixel types (MVAMD_PIXEL_*)
CREATE_AUTO, CREATE_HOST, CREATE_DEVICE = 0, 1, 2  # raster builders (MVAMD_CREATE_*)

STATUS_NAMES = {0: "OK", -1: "INVALID_ARGUMENT", -2: "OUT_OF_RANGE", -3: "TYPE_MISMATCH", -4: "NOT_FOUND",
                -5: "OUT_OF_MEMORY", -6: "UNSUPPORTED", -7: "INVALID_STATE", -100: "IO_ERROR",
                -101: "DECODE_ERROR", -102: "ENCODE_ERROR", -103: "LIMIT_EXCEEDED", -1000: "INTERNAL_ERROR"}

# Symbols declared by include/mv_ncc.h and include/mvkit_amd.h.
PUBLIC_SYMBOLS = [
    "mv_get_last_error", "mv_get_last_error_info", "mv_create_tuple", "mv_copy_tuple", "mv_destroy_tuple",
    "mv_length_tuple", "mv_get_type", "mv_get_i", "mv_get_d", "mv_get_s", "mv_get_h", "mv_resize_tuple", "mv_set_i",
    "mv_set_d", "mv_set_s", "mv_set_h", "mv_gen_empty_obj", "mv_clear_obj", "mv_gen_image1", "mv_gen_image1_extern",
    "mv_gen_region_runs", "mv_get_region_runs", "mv_gen_rectangle1", "mv_reduce_domain", "mv_full_domain",
    "mv_get_domain",
    "mv_clear_handle", "mv_tuple_sem_type", "mv_serialize_handle", "mv_deserialize_handle", "mv_create_serialized_item_ptr",
    "mv_get_serialized_item_ptr", "mv_write_ncc_model", "mv_read_ncc_model", "mv_create_ncc_model", "mv_get_ncc_model_params", "mv_find_ncc_model", "mv_find_ncc_models",
    "mvamd_create_ncc_model", "mvamd_find_ncc_model_device", "mvamd_find_ncc_model_image",
    "mvamd_find_ncc_models_device", "mvamd_find_ncc_models_batch", "mvamd_find_ncc_model_batch",
    "mvamd_server_create", "mvamd_server_destroy", "mvamd_server_submit", "mvamd_server_wait", "mvamd_find_ncc_model_batch_submit",
                  "mvamd_find_ncc_model_batch_wait", "mvamd_find_ncc_model_trace", "mvamd_model_info", "mvamd_model_raster",
    "mvamd_get_ncc_model_tuning", "mvamd_set_ncc_model_tuning",
]
# Symbols declared by include/ncchip.h.
NCCHIP_SYMBOLS = ["ncchip_pyramid_level", "ncchip_row_prefix", "ncchip_pyramid_batch", "ncchip_row_prefix_batch",
                  "ncchip_fill_u32", "ncchip_copy_u32", "ncchip_coarse", "ncchip_coarse_mfma", "ncchip_coarse_mfma_fits",
                  "ncchip_coarse_mfma_queue", "ncchip_coarse_mfma_queue_bytes",
                  "ncchip_peaks", "ncchip_storage0", "ncchip_track", "ncchip_track_mfma", "ncchip_full_score", "ncchip_refine",
                  "ncchip_domain_sat", "ncchip_domain_planes", "ncchip_domain_snap", "ncchip_create_rasters",
                  "ncchip_pyramid3_u8", "ncchip_row_prefix_noalloc"]


class StringView(C.Structure):
    _fields_ = [("data", C.c_char_p), ("size", C.c_size_t)]

    def text(self) -> str:
        return C.string_at(self.data, self.size).decode() if self.data else ""


class ErrorInfo(C.Structure):
    _fields_ = [("status", C.c_int32), ("code", C.c_int32), ("operation", StringView), ("parameter", StringView),
                ("parameter_index", C.c_int64), ("message", StringView)]


class FindOptions(C.Structure):
    _fields_ = [("angle_start", C.c_double), ("angle_extent", C.c_double), ("min_score", C.c_double),
                ("max_overlap", C.c_double), ("num_matches", C.c_int64), ("subpixel", C.c_int32),
                ("highest_level", C.c_int32), ("lowest_level", C.c_int32), ("form", C.c_int32),
                ("screen", C.c_double), ("profile", C.c_int32), ("vector_tracking", C.c_int32),
                ("pixel_type", C.c_int32), ("max_candidates", C.c_int64), ("no_graph", C.c_int32),
                ("coarse_tile", C.c_int32)]


class FindStats(C.Structure):
    _fields_ = [("n_coarse", C.c_int32), ("n_prepared", C.c_int32), ("top", C.c_int32), ("finest", C.c_int32),
                ("angle_stride", C.c_int32), ("storage", C.c_int32 * 4), ("t_coarse_us", C.c_double),
                ("t_track_us", C.c_double), ("t_refine_us", C.c_double), ("t_total_us", C.c_double),
                ("kernel_us", C.c_double * 8), ("coarse_flops", C.c_double),
                ("track_flops", C.c_double), ("track_mfma_flops", C.c_double), ("coarse_form", C.c_int32),
                ("coarse_tile", C.c_int32)]

class ServerConfig(C.Structure):
    _fields_ = [("workers", C.c_int32), ("max_batch", C.c_int32), ("hbm_budget", C.c_int64),
                ("queue_limit", C.c_int64), ("no_coarse_queue", C.c_int32), ("reserved", C.c_int32)]


class RequestTimes(C.Structure):
    _fields_ = [("admitted_us", C.c_double), ("started_us", C.c_double), ("finished_us", C.c_double),
                ("bytes", C.c_double), ("batch_size", C.c_int32), ("worker", C.c_int32)]


class BatchStats(C.Structure):
    _fields_ = [("coarse_queue_us", C.c_double), ("coarse_flops", C.c_double), ("total_us", C.c_double),
                ("tune_us", C.c_double), ("queries", C.c_int32), ("queue_queries", C.c_int32),
                ("candidates", C.c_int32), ("tuned", C.c_int32), ("refined", C.c_int32),
                ("refine_rounds", C.c_int32)]


PROFILE_ALL = 0xff
COARSE_TILE_AUTO = 0
COARSE_TILE_DEFAULT = -1


def coarse_tile(rows: int, col_waves: int, row_waves: int, tiles: int, bound: int = 0) -> int:
    """MVAMD_COARSE_TILE_BOUND: a split-BF16 coarse tile shape packed for the
    options (bound: 0, 8, 16 or 32 angles per bound group)."""
    b = {0: 0, 8: 1, 16: 2, 32: 3}[bound]
    return rows | (col_waves << 4) | (row_waves << 8) | (tiles << 12) | (b << 16)


def tile_fields(v: int) -> tuple:
    """(rows, col_waves, row_waves, tiles) of a packed shape (0: none)."""
    return (v & 15, (v >> 4) & 15, (v >> 8) & 15, (v >> 12) & 15)


def tile_bound(v: int) -> int:
    """Angles per bound group of a packed shape (0: the common support of every angle)."""
    b = (v >> 16) & 3
    return 4 << b if b else 0
STAGES = ["pyramid", "prefix", "coarse", "peaks", "track", "full_score", "refine", "track_mfma"]


class MvError(RuntimeError):
    def __init__(self, status: int, info: ErrorInfo | None):
        self.status = status
        self.code = info.code if info else 0
        self.operation = info.operation.text() if info else ""
        self.parameter = info.parameter.text() if info else ""
        self.parameter_index = info.parameter_index if info else -1
        self.message = info.message.text() if info else ""
        super().__init__(f"{STATUS_NAMES.get(status, status)} in {self.operation}({self.parameter}): {self.message}")


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run mvkit_amd/build.py (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        P, S, D, I64, I32 = C.c_void_p, C.c_size_t, C.c_double, C.c_int64, C.c_int32
        for name in PUBLIC_SYMBOLS + NCCHIP_SYMBOLS:
            getattr(L, name).restype = C.c_int32
        L.mv_get_last_error_info.argtypes = [C.POINTER(ErrorInfo)]
        L.mv_create_tuple.argtypes = [C.POINTER(P), S]
        L.mv_copy_tuple.argtypes = [P, C.POINTER(P)]
        L.mv_destroy_tuple.argtypes = [P]
        L.mv_length_tuple.argtypes = [P, C.POINTER(S)]
        L.mv_get_type.argtypes = [P, S, C.POINTER(I32)]
        L.mv_get_i.argtypes = [P, S, C.POINTER(I64)]
        L.mv_get_d.argtypes = [P, S, C.POINTER(D)]
        L.mv_get_s.argtypes = [P, S, C.POINTER(StringView)]
        L.mv_get_h.argtypes = [P, S, C.POINTER(P)]
        L.mv_set_i.argtypes = [P, I64, S]
        L.mv_set_d.argtypes = [P, D, S]
        L.mv_set_s.argtypes = [P, StringView, S]
        L.mv_set_h.argtypes = [P, P, S]
        L.mv_gen_empty_obj.argtypes = [C.POINTER(P)]
        L.mv_clear_obj.argtypes = [P]
        L.mv_gen_image1.argtypes = [C.POINTER(P), P, P, P, P]
        L.mv_gen_image1_extern.argtypes = [C.POINTER(P), P, P, P, P, P]
        L.mv_gen_region_runs.argtypes = [C.POINTER(P), P, P, P]
        L.mv_get_region_runs.argtypes = [P] + [C.POINTER(P)] * 3
        L.mv_gen_rectangle1.argtypes = [C.POINTER(P), P, P, P, P]
        L.mv_reduce_domain.argtypes = [P, P, C.POINTER(P)]
        L.mv_full_domain.argtypes = [P, C.POINTER(P)]
        L.mv_get_domain.argtypes = [P, C.POINTER(P)]
        L.mv_clear_handle.argtypes = [P]
        L.mv_tuple_sem_type.argtypes = [P, C.POINTER(P)]
        L.mv_create_ncc_model.argtypes = [P, P, P, P, P, P, C.POINTER(P)]
        L.mvamd_create_ncc_model.argtypes = [P, P, P, P, P, P, C.c_int32, C.POINTER(P)]
        L.mv_get_ncc_model_params.argtypes = [P] + [C.POINTER(P)] * 5
        L.mv_find_ncc_model.argtypes = [P] * 9 + [C.POINTER(P)] * 4
        L.mv_find_ncc_models.argtypes = [P] * 9 + [C.POINTER(P)] * 5
        L.mv_write_ncc_model.argtypes = [P, P]
        L.mv_read_ncc_model.argtypes = [P, C.POINTER(P)]
        L.mv_serialize_handle.argtypes = [P, C.POINTER(P)]
        L.mv_deserialize_handle.argtypes = [P, C.POINTER(P)]
        L.mv_create_serialized_item_ptr.argtypes = [P, P, P, C.POINTER(P)]
        L.mv_get_serialized_item_ptr.argtypes = [P, C.POINTER(P), C.POINTER(P)]
        L.mvamd_find_ncc_models_device.argtypes = [P, P, I64, I64, I64, C.POINTER(FindOptions), S, I64, D, P, S,
                                                   C.POINTER(S)]
        L.mvamd_find_ncc_models_batch.argtypes = [P, P, P, P, P, S, C.POINTER(FindOptions), S, I64, D, I32, P, S, P,
                                                  C.POINTER(BatchStats)]
        L.ncchip_coarse_mfma_queue_bytes.restype = S
        L.mvamd_server_create.argtypes = [C.POINTER(ServerConfig), C.POINTER(P)]
        L.mvamd_server_destroy.argtypes = [P]
        L.mvamd_server_submit.argtypes = [P, P, I64, I64, I64, P, C.POINTER(FindOptions), S, I64, D,
                                          C.POINTER(C.c_uint64)]
        L.mvamd_server_wait.argtypes = [P, C.c_uint64, P, S, C.POINTER(S), C.POINTER(RequestTimes)]
        L.mvamd_find_ncc_model_image.argtypes = [P, P, C.POINTER(FindOptions), P, S, C.POINTER(S),
                                                 C.POINTER(FindStats)]
        L.mvamd_find_ncc_model_device.argtypes = [P, P, I64, I64, I64, C.POINTER(FindOptions), P, S, C.POINTER(S),
                                                  C.POINTER(FindStats)]
        L.mvamd_find_ncc_model_batch.argtypes = [P, P, P, P, P, S, C.POINTER(FindOptions), P, S, P,
                                                 C.POINTER(FindStats)]
        L.mvamd_find_ncc_model_batch_submit.argtypes = [P, P, P, P, P, S, C.POINTER(FindOptions), C.c_int32,
                                                        C.POINTER(C.c_uint64)]
        L.mvamd_find_ncc_model_batch_wait.argtypes = [C.c_uint64, P, S, P, C.POINTER(FindStats)]
        L.mvamd_find_ncc_model_trace.argtypes = [P, P, I64, I64, I64, C.POINTER(FindOptions), P, S, C.POINTER(S), P, S,
                                                 C.POINTER(S), P, P, S, C.POINTER(S), C.POINTER(FindStats)]
        L.mvamd_model_info.argtypes = [P, P]
        L.mvamd_get_ncc_model_tuning.argtypes = [P, C.c_char_p, S, C.POINTER(S)]
        L.mvamd_set_ncc_model_tuning.argtypes = [P, C.c_char_p]
        L.mvamd_model_raster.argtypes = [P, I32, I32, P, P, P, P, S]
        _lib = L
    return _lib


def _check(status: int):
    if status != 0:
        info = ErrorInfo()
        lib().mv_get_last_error_info(C.byref(info))
        raise MvError(status, info)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class Tuple:
    """Owned control tuple."""

    def __init__(self, values=None, handle=None):
        if handle is not None:
            self.h = handle
            return
        vals = [] if values is None else (list(values) if isinstance(values, (list, tuple)) else [values])
        h = C.c_void_p()
        _check(lib().mv_create_tuple(C.byref(h), len(vals)))
        self.h = h
        self._keep = []
        for i, v in enumerate(vals):
            if isinstance(v, bool):
                raise TypeError("booleans are not control values")
            if isinstance(v, (int, np.integer)):
                _check(lib().mv_set_i(self.h, int(v), i))
            elif isinstance(v, (float, np.floating)):
                _check(lib().mv_set_d(self.h, float(v), i))
            elif isinstance(v, str):
                b = v.encode()
                self._keep.append(b)
                _check(lib().mv_set_s(self.h, StringView(b, len(b)), i))
            else:
                raise TypeError(f"unsupported control value {v!r}")

    def __len__(self):
        n = C.c_size_t()
        _check(lib().mv_length_tuple(self.h, C.byref(n)))
        return n.value

    def values(self) -> list:
        out = []
        for i in range(len(self)):
            k = C.c_int32()
            _check(lib().mv_get_type(self.h, i, C.byref(k)))
            if k.value == MV_VALUE_INT64:
                v = C.c_int64(); _check(lib().mv_get_i(self.h, i, C.byref(v))); out.append(v.value)
            elif k.value == MV_VALUE_REAL64:
                v = C.c_double(); _check(lib().mv_get_d(self.h, i, C.byref(v))); out.append(v.value)
            elif k.value == MV_VALUE_STRING:
                v = StringView(); _check(lib().mv_get_s(self.h, i, C.byref(v))); out.append(v.text())
            elif k.value == MV_VALUE_HANDLE:
                out.append("<handle>")
            else:
                out.append(None)
        return out

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            try:
                _lib.mv_destroy_tuple(self.h)
            except Exception:  # interpreter shutdown
                pass
            self.h = None


class Region:
    """Region object: from runs (n x 3: row, first column, last column) via
    mv_gen_region_runs, or rectangles via mv_gen_rectangle1."""

    def __init__(self, runs=None, handle=None):
        if handle is not None:
            self.h = handle
            return
        r = np.asarray(runs if runs is not None else np.zeros((0, 3)), dtype=np.int64).reshape(-1, 3)
        h = C.c_void_p()
        t = [Tuple([int(x) for x in r[:, k]]) for k in range(3)]
        _check(lib().mv_gen_region_runs(C.byref(h), *[x.h for x in t]))
        self.h = h

    @classmethod
    def rectangle1(cls, row1, col1, row2, col2) -> "Region":
        h = C.c_void_p()
        t = [Tuple(v) for v in (row1, col1, row2, col2)]
        _check(lib().mv_gen_rectangle1(C.byref(h), *[x.h for x in t]))
        return cls(handle=h)

    def runs(self) -> np.ndarray:
        outs = [C.c_void_p() for _ in range(3)]
        _check(lib().mv_get_region_runs(self.h, *[C.byref(o) for o in outs]))
        cols = [Tuple(handle=o).values() for o in outs]
        return np.array(cols, dtype=np.int64).T.reshape(-1, 3)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            try:
                _lib.mv_clear_obj(self.h)
            except Exception:  # interpreter shutdown
                pass
            self.h = None


class Image:
    """Image object holding a copy of a u8 (or u16) plane, optionally with a
    reduced domain (runs n x 3: row, first column, last column, or a Region)."""

    def __init__(self, pixels: np.ndarray | None, domain=None, handle=None):
        h = C.c_void_p()
        if handle is not None:
            h = handle
        elif pixels is None:
            _check(lib().mv_gen_empty_obj(C.byref(h)))
        else:
            a = np.ascontiguousarray(pixels)
            kind = {np.dtype(np.uint8): "byte", np.dtype(np.uint16): "uint2"}[a.dtype]
            args = [Tuple(kind), Tuple(int(a.shape[1])), Tuple(int(a.shape[0])), Tuple(int(a.ctypes.data))]
            _check(lib().mv_gen_image1(C.byref(h), *[t.h for t in args]))
        self.h = h
        if domain is not None:
            reduced = self.reduce_domain(domain)
            self.h, reduced.h = reduced.h, self.h

    def reduce_domain(self, region) -> "Image":
        """mv_reduce_domain with a Region or runs."""
        g = region if isinstance(region, Region) else Region(region)
        out = C.c_void_p()
        _check(lib().mv_reduce_domain(self.h, g.h, C.byref(out)))
        return Image(None, handle=out)

    def full_domain(self) -> "Image":
        out = C.c_void_p()
        _check(lib().mv_full_domain(self.h, C.byref(out)))
        return Image(None, handle=out)

    def domain(self) -> "Region":
        out = C.c_void_p()
        _check(lib().mv_get_domain(self.h, C.byref(out)))
        return Region(None, handle=out)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            try:
                _lib.mv_clear_obj(self.h)
            except Exception:  # interpreter shutdown
                pass
            self.h = None


def _to_tuple(v):
    return v if isinstance(v, Tuple) else Tuple(v)


def _real(v):
    """Numbers become reals; anything else is passed through unchanged."""
    return float(v) if isinstance(v, (int, float, np.integer, np.floating)) and not isinstance(v, bool) else v


class NccModel:
    """NCC model handle (mirrors mvkit's MvNCCModel, mv_ncc_matching.hpp)."""

    def __init__(self, template, num_levels="auto", angle_start=0.0, angle_extent=0.0,
                 angle_step="auto", metric="use_polarity", builder=None):
        """template: an array or an Image (which may have a reduced domain).
        builder: None (mv_create_ncc_model) or CREATE_HOST / CREATE_DEVICE /
        CREATE_AUTO (mvamd_create_ncc_model)."""
        img = template if isinstance(template, Image) else Image(template)
        out = C.c_void_p()
        args = [_to_tuple(num_levels), _to_tuple(float(angle_start)), _to_tuple(float(angle_extent)),
                _to_tuple(angle_step), _to_tuple(metric)]
        if builder is None:
            _check(lib().mv_create_ncc_model(img.h, *[t.h for t in args], C.byref(out)))
        else:
            _check(lib().mvamd_create_ncc_model(img.h, *[t.h for t in args], int(builder), C.byref(out)))
        self._bind(Tuple(handle=out))

    @classmethod
    def _from_tuple(cls, t: "Tuple") -> "NccModel":
        m = cls.__new__(cls)
        m._bind(t)
        return m

    @classmethod
    def read(cls, path: str) -> "NccModel":
        """mv_read_ncc_model."""
        out = C.c_void_p()
        name = Tuple(str(path))
        _check(lib().mv_read_ncc_model(name.h, C.byref(out)))
        return cls._from_tuple(Tuple(handle=out))

    @classmethod
    def deserialize(cls, data: bytes) -> "NccModel":
        """mv_create_serialized_item_ptr + mv_deserialize_handle."""
        buf = C.create_string_buffer(bytes(data), len(data))
        item = C.c_void_p()
        args = [Tuple(C.addressof(buf)), Tuple(len(data)), Tuple("true")]  # alive across the call
        _check(lib().mv_create_serialized_item_ptr(*[t.h for t in args], C.byref(item)))
        it = Tuple(handle=item)
        out = C.c_void_p()
        _check(lib().mv_deserialize_handle(it.h, C.byref(out)))
        return cls._from_tuple(Tuple(handle=out))

    def write(self, path: str) -> None:
        """mv_write_ncc_model."""
        name = Tuple(str(path))
        _check(lib().mv_write_ncc_model(self.tuple.h, name.h))

    def serialize(self) -> bytes:
        """mv_serialize_handle + mv_get_serialized_item_ptr: the MVKIT_NCC 1 bytes."""
        item = C.c_void_p()
        _check(lib().mv_serialize_handle(self.tuple.h, C.byref(item)))
        it = Tuple(handle=item)
        ptr, size = C.c_void_p(), C.c_void_p()
        _check(lib().mv_get_serialized_item_ptr(it.h, C.byref(ptr), C.byref(size)))
        addr, n = Tuple(handle=ptr).values()[0], Tuple(handle=size).values()[0]
        return C.string_at(addr, n)

    def _bind(self, t: "Tuple") -> None:
        self.tuple = t
        info = np.zeros(6)
        _check(lib().mvamd_model_info(self.tuple.h, _ptr(info)))
        self.n_levels, self.n_angles = int(info[0]), int(info[1])
        self.step, self.origin_row, self.origin_col, self.sample_count = info[2], info[3], info[4], int(info[5])

    def params(self):
        outs = [C.c_void_p() for _ in range(5)]
        _check(lib().mv_get_ncc_model_params(self.tuple.h, *[C.byref(o) for o in outs]))
        return [Tuple(handle=o).values()[0] for o in outs]

    def raster(self, level: int, angle: int) -> dict:
        hdr = np.zeros(11)
        _check(lib().mvamd_model_raster(self.tuple.h, level, angle, _ptr(hdr), None, None, None, 0))
        n = int(hdr[0])
        rows, cols, w = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(n)
        _check(lib().mvamd_model_raster(self.tuple.h, level, angle, _ptr(hdr), _ptr(rows), _ptr(cols), _ptr(w), n))
        return {"n": n, "angle": hdr[1], "orow": hdr[2], "ocol": hdr[3], "rmin": int(hdr[4]), "rmax": int(hdr[5]),
                "cmin": int(hdr[6]), "cmax": int(hdr[7]), "wsum": hdr[8], "n_runs": int(hdr[9]),
                "n_wruns": int(hdr[10]), "rows": rows, "cols": cols, "w": w}

    def clear(self):
        _check(lib().mv_clear_handle(self.tuple.h))

    @property
    def tuning(self) -> str:
        """The autotuned coarse tile shapes (mvamd_get_ncc_model_tuning)."""
        n = C.c_size_t()
        _check(lib().mvamd_get_ncc_model_tuning(self.tuple.h, None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value + 1)
        _check(lib().mvamd_get_ncc_model_tuning(self.tuple.h, buf, n.value + 1, C.byref(n)))
        return buf.value.decode()

    @tuning.setter
    def tuning(self, text: str) -> None:
        _check(lib().mvamd_set_ncc_model_tuning(self.tuple.h, text.encode()))

    def find(self, image, angle_start=0.0, angle_extent=0.0, min_score=0.8, num_matches=1, max_overlap=0.5,
             subpixel="true", num_levels=0):
        """mv_find_ncc_model; returns (n, 4) array of Row, Column, Angle, Score
        in the C ABI's HALCON convention."""
        img = image if isinstance(image, Image) else Image(image)
        outs = [C.c_void_p() for _ in range(4)]
        args = [_to_tuple(_real(angle_start)), _to_tuple(_real(angle_extent)), _to_tuple(_real(min_score)),
                _to_tuple(num_matches), _to_tuple(_real(max_overlap)), _to_tuple(subpixel), _to_tuple(num_levels)]
        _check(lib().mv_find_ncc_model(img.h, self.tuple.h, *[t.h for t in args], *[C.byref(o) for o in outs]))
        cols = [Tuple(handle=o).values() for o in outs]
        return np.array(cols, dtype=np.float64).T.reshape(-1, 4)

    @staticmethod
    def options(angle_start=0.0, angle_extent=0.0, min_score=0.8, num_matches=1, max_overlap=0.5, subpixel=True,
                highest=0, lowest=1, form=FORM_FP32, screen=-math.inf, profile=0,
                vector_tracking=False, pixel_type=PIXEL_BYTE, max_candidates=0, graph=True,
                tile=COARSE_TILE_AUTO) -> FindOptions:
        """profile: True times every stage, an int is a mask of stage bits
        (bit i = STAGES[i]).  vector_tracking: track every level with the FP64
        vector kernel instead of matrix cores (same results).  pixel_type:
        PIXEL_BYTE or PIXEL_UINT2 for the device images (pitches in bytes).
        graph: replay the device part as a captured HIP graph when it repeats.
        tile: the split-BF16 coarse tile shape (COARSE_TILE_AUTO: the model's
        autotuned shape, COARSE_TILE_DEFAULT, or coarse_tile(...))."""
        mask = PROFILE_ALL if profile is True else int(profile)
        return FindOptions(angle_start, angle_extent, min_score, max_overlap, num_matches, int(subpixel), highest,
                           lowest, form, screen, mask, int(vector_tracking), int(pixel_type), int(max_candidates),
                           int(not graph), int(tile))

    def find_image(self, image, opts: FindOptions, cap: int = 4096):
        """mvamd_find_ncc_model_image: find on an Image (or array) with the
        extension options; poses in the internal convention."""
        img = image if isinstance(image, Image) else Image(image)
        res = np.zeros((cap, 4))
        n = C.c_size_t()
        st = FindStats()
        _check(lib().mvamd_find_ncc_model_image(img.h, self.tuple.h, C.byref(opts), _ptr(res), cap, C.byref(n),
                                                C.byref(st)))
        return res[:min(n.value, cap)].copy(), st

    def find_device(self, d_image: int, width: int, height: int, pitch: int, opts: FindOptions, cap: int = 4096):
        """Find on a device-resident u8 or u16 image (integer device address;
        opts.pixel_type names the type, pitch is in bytes)."""
        res = np.zeros((cap, 4))
        n = C.c_size_t()
        st = FindStats()
        _check(lib().mvamd_find_ncc_model_device(self.tuple.h, C.c_void_p(d_image), width, height, pitch,
                                                 C.byref(opts), _ptr(res), cap, C.byref(n), C.byref(st)))
        return res[:n.value].copy(), st

    def find_batch(self, d_images: list, width, height, pitch, opts: FindOptions, cap: int = 256):
        """Find in several device-resident u8 images in one batch; width,
        height and pitch are one value for all frames or one per frame."""
        n = len(d_images)

        def per_frame(v):
            return list(v) if isinstance(v, (list, tuple)) else [v] * n
        ptrs = (C.c_void_p * n)(*d_images)
        ws = (C.c_int64 * n)(*per_frame(width))
        hs = (C.c_int64 * n)(*per_frame(height))
        ps = (C.c_int64 * n)(*per_frame(pitch))
        res = np.zeros((n, cap, 4))
        counts = np.zeros(n, np.uint64)
        st = FindStats()
        _check(lib().mvamd_find_ncc_model_batch(self.tuple.h, ptrs, ws, hs, ps, n, C.byref(opts), _ptr(res), cap,
                                                _ptr(counts), C.byref(st)))
        return [res[k, :int(counts[k])].copy() for k in range(n)], st

    def submit_batch(self, d_images: list, width, height, pitch, opts: FindOptions, stats: bool = True) -> tuple:
        """Submits a batch (see find_batch) without waiting; returns a handle
        for wait_batch.  At most two batches may be pending."""
        n = len(d_images)

        def per_frame(v):
            return list(v) if isinstance(v, (list, tuple)) else [v] * n
        ptrs = (C.c_void_p * n)(*d_images)
        ws = (C.c_int64 * n)(*per_frame(width))
        hs = (C.c_int64 * n)(*per_frame(height))
        ps = (C.c_int64 * n)(*per_frame(pitch))
        ticket = C.c_uint64()
        _check(lib().mvamd_find_ncc_model_batch_submit(self.tuple.h, ptrs, ws, hs, ps, n, C.byref(opts), int(stats),
                                                       C.byref(ticket)))
        return ticket.value, n

    @staticmethod
    def wait_batch(handle: tuple, cap: int = 256):
        ticket, n = handle
        res = np.zeros((max(n, 1), cap, 4))
        counts = np.zeros(max(n, 1), np.uint64)
        st = FindStats()
        _check(lib().mvamd_find_ncc_model_batch_wait(ticket, _ptr(res), cap, _ptr(counts), C.byref(st)))
        return [res[k, :int(counts[k])].copy() for k in range(n)], st

    def find_trace(self, image: np.ndarray, opts: FindOptions, cap: int = 20000):
        img = np.ascontiguousarray(image, dtype=np.uint8)
        res = np.zeros((4096, 4))
        coarse = np.zeros((cap, 4))
        prep = np.zeros((cap, 4))
        refd = np.zeros((cap, 5))
        n, nc, nprep = C.c_size_t(), C.c_size_t(), C.c_size_t()
        st = FindStats()
        _check(lib().mvamd_find_ncc_model_trace(self.tuple.h, _ptr(img), img.shape[1], img.shape[0], img.shape[1],
                                                C.byref(opts), _ptr(res), 4096, C.byref(n), _ptr(coarse), cap,
                                                C.byref(nc), _ptr(prep), _ptr(refd), cap, C.byref(nprep),
                                                C.byref(st)))
        return {"results": res[:n.value].copy(), "coarse": coarse[:min(nc.value, cap)].copy(),
                "prepared": prep[:min(nprep.value, cap)].copy(), "refined": refd[:min(nprep.value, cap)].copy(),
                "stats": st}


def _reals(v):
    return [_real(x) for x in v] if isinstance(v, (list, tuple)) else _real(v)


def _model_ids(models) -> Tuple:
    """One tuple holding the handles of several models."""
    h = C.c_void_p()
    _check(lib().mv_create_tuple(C.byref(h), len(models)))
    t = Tuple(handle=h)
    for i, m in enumerate(models):
        hv = C.c_void_p()
        _check(lib().mv_get_h(m.tuple.h, 0, C.byref(hv)))
        _check(lib().mv_set_h(t.h, hv, i))
    return t


def find_models(models, image, angle_start=0.0, angle_extent=0.0, min_score=0.8, num_matches=1, max_overlap=0.5,
                subpixel="true", num_levels=0):
    """mv_find_ncc_models (mirrors MvNCCModel::FindNccModels); every control
    is one value or a list with one per model.  Returns an (n, 5) array of
    Row, Column, Angle, Score in the C ABI's HALCON convention and the model
    index."""
    img = image if isinstance(image, Image) else Image(image)
    ids = _model_ids(models)
    outs = [C.c_void_p() for _ in range(5)]
    args = [_to_tuple(_reals(angle_start)), _to_tuple(_reals(angle_extent)), _to_tuple(_reals(min_score)),
            _to_tuple(num_matches), _to_tuple(_reals(max_overlap)), _to_tuple(subpixel), _to_tuple(num_levels)]
    _check(lib().mv_find_ncc_models(img.h, ids.h, *[t.h for t in args], *[C.byref(o) for o in outs]))
    cols = [Tuple(handle=o).values() for o in outs]
    return np.array(cols, dtype=np.float64).T.reshape(-1, 5)


def find_models_device(models, d_image: int, width: int, height: int, pitch: int, opts, selection_matches=-1,
                       selection_overlap=-1.0, cap: int = 4096):
    """mvamd_find_ncc_models_device: opts is one FindOptions or one per model.
    Returns (n, 5): row, col, angle, score (internal convention), model index."""
    ids = _model_ids(models)
    ol = list(opts) if isinstance(opts, (list, tuple)) else [opts]
    arr = (FindOptions * len(ol))(*ol)
    res = np.zeros((cap, 5))
    n = C.c_size_t()
    _check(lib().mvamd_find_ncc_models_device(ids.h, C.c_void_p(d_image), width, height, pitch, arr, len(ol),
                                              int(selection_matches), float(selection_overlap), _ptr(res), cap,
                                              C.byref(n)))
    return res[:min(n.value, cap)].copy()


def find_models_batch(models, d_images: list, widths, heights, pitches, opts, selection_matches=-1,
                      selection_overlap=-1.0, coarse_queue=True, cap: int = 1024, stats: bool = False):
    """mvamd_find_ncc_models_batch: find_models over several device-resident
    images (widths/heights/pitches: one value or one per image).  Returns one
    (n, 5) array per image: row, col, angle, score (internal convention),
    model index; with stats=True also the call's BatchStats."""
    n = len(d_images)

    def per_image(v):
        return list(v) if isinstance(v, (list, tuple)) else [v] * n
    ids = _model_ids(models)
    ol = list(opts) if isinstance(opts, (list, tuple)) else [opts]
    arr = (FindOptions * len(ol))(*ol)
    ptrs = (C.c_void_p * n)(*d_images)
    ws = (C.c_int64 * n)(*per_image(widths))
    hs = (C.c_int64 * n)(*per_image(heights))
    ps = (C.c_int64 * n)(*per_image(pitches))
    res = np.zeros((max(n, 1), cap, 5))
    counts = np.zeros(max(n, 1), np.uint64)
    st = BatchStats()
    _check(lib().mvamd_find_ncc_models_batch(ids.h, ptrs, ws, hs, ps, n, arr, len(ol), int(selection_matches),
                                             float(selection_overlap), int(bool(coarse_queue)), _ptr(res), cap,
                                             _ptr(counts), C.byref(st) if stats else None))
    out = [res[k, :min(int(counts[k]), cap)].copy() for k in range(n)]
    return (out, st) if stats else out


class Server:
    """The request loop (mvamd_server_*): submit() queues a host image and
    its models, wait() returns the matches ((n, 5): row, col, angle, score,
    model index) and the request's times."""

    def __init__(self, workers=0, max_batch=0, hbm_budget=0, queue_limit=0, coarse_queue=True):
        cfg = ServerConfig(workers, max_batch, hbm_budget, queue_limit, int(not coarse_queue), 0)
        h = C.c_void_p()
        _check(lib().mvamd_server_create(C.byref(cfg), C.byref(h)))
        self.h = h

    def submit(self, image: np.ndarray, models, opts, selection_matches=-1, selection_overlap=-1.0) -> int:
        img = np.ascontiguousarray(image)
        ids = _model_ids(models)
        ol = list(opts) if isinstance(opts, (list, tuple)) else [opts]
        arr = (FindOptions * len(ol))(*ol)
        rid = C.c_uint64()
        _check(lib().mvamd_server_submit(self.h, _ptr(img), img.shape[1], img.shape[0], img.strides[0], ids.h, arr,
                                         len(ol), int(selection_matches), float(selection_overlap), C.byref(rid)))
        return rid.value

    def wait(self, request: int, cap: int = 1024):
        res = np.zeros((cap, 5))
        n = C.c_size_t()
        t = RequestTimes()
        _check(lib().mvamd_server_wait(self.h, request, _ptr(res), cap, C.byref(n), C.byref(t)))
        return res[:min(n.value, cap)].copy(), {"admitted_us": t.admitted_us, "started_us": t.started_us,
                                                "finished_us": t.finished_us, "bytes": t.bytes,
                                                "batch_size": t.batch_size, "worker": t.worker}

    def close(self):
        if getattr(self, "h", None):
            _check(lib().mvamd_server_destroy(self.h))
            self.h = None


def halcon_shift(poses: np.ndarray) -> np.ndarray:
    """Internal (centroid) poses -> the C ABI's Row/Column convention
    (ncc_matching_api.cpp:171-176)."""
    out = poses.copy()
    c, s = np.cos(poses[:, 2]), np.sin(poses[:, 2])
    out[:, 0] = poses[:, 0] + 0.5 * (1 - c + s)
    out[:, 1] = poses[:, 1] + 0.5 * (1 - c - s)
    return out
