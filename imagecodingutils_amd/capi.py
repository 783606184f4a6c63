"""ctypes binding of libicu_hip.so (the C ABI in include/icu_hip.h).

There is no CPU fallback: importing this module loads the in-tree HIP
library and raises if it is missing.  Device buffers are managed through the
library's own small runtime (icu_malloc / icu_memcpy_*), so no framework is
needed to drive it.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libicu_hip.so"
HEADER = Path(__file__).resolve().parents[1] / "include" / "icu_hip.h"

ICU_OK, ICU_E_ARG, ICU_E_RANGE, ICU_E_DEVICE, ICU_E_STATE = 0, -1, -2, -3, -4


class IcuError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"icu error {code}: {msg}")
        self.code = code


class MeConfig(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "width", "height", "range", "metric", "window", "filter", "subpel",
        "predictor", "best_vector", "tile_rows", "mbs_per_task", "grid_mode")]


class ServerConfig(C.Structure):
    _fields_ = [("lanes", C.c_int), ("max_inflight", C.c_int), ("mem_budget", C.c_longlong),
                ("qp", C.c_int)]


class ServerResult(C.Structure):
    _fields_ = [("tag", C.c_longlong), ("stream", C.c_int), ("frame", C.c_int), ("nmb", C.c_int),
                ("submit_s", C.c_double), ("done_s", C.c_double),
                ("levels_sum", C.c_ulonglong), ("recon_sum", C.c_ulonglong),
                ("mv", C.POINTER(C.c_int16))]


class VbsConfig(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("width", "height", "range", "nrefs", "lambda_")]


_vp = C.c_void_p
_SIGS = {
    "icu_me_config_reference": (C.c_int, [C.POINTER(MeConfig), C.c_int, C.c_int, C.c_int]),
    "icu_me_create": (C.c_int, [C.POINTER(MeConfig), C.POINTER(_vp)]),
    "icu_me_destroy": (C.c_int, [_vp]),
    "icu_me_estimate": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, C.POINTER(C.c_long)]),
    "icu_me_estimate_device": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "icu_me_estimate_timed": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, C.c_int, C.POINTER(C.c_float)]),
    "icu_me_estimate_device_noalloc": (C.c_int, [C.POINTER(MeConfig), _vp, _vp, _vp, _vp, _vp, _vp]),
    "icu_me_check": (C.c_int, [_vp]),
    "icu_me_tie_count": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "icu_me_debug_set_wave_error": (C.c_int, [_vp]),
    "icu_me_get_tiling": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "icu_me_autotune": (C.c_int, [_vp, _vp, _vp, C.c_int]),
    "icu_code_frame_device": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp, _vp]),
    "icu_code_frame_qps_device": (C.c_int, [_vp, _vp, _vp, C.c_int, C.POINTER(C.c_int), C.c_int, C.c_int,
                                            C.POINTER(_vp), C.POINTER(_vp), _vp, _vp]),
    "icu_fwd4x4": (C.c_int, [_vp, _vp, C.c_long, C.c_int, C.c_int, C.c_int, C.c_int]),
    "icu_inv4x4": (C.c_int, [_vp, _vp, C.c_long, C.c_int, C.c_int, C.c_int, C.c_int, _vp]),
    "icu_fwd4x4_device": (C.c_int, [_vp, _vp, C.c_long, C.c_int, C.c_int, C.c_int, C.c_int, _vp]),
    "icu_inv4x4_device": (C.c_int, [_vp, _vp, C.c_long, C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp]),
    "icu_me_batch_create": (C.c_int, [C.POINTER(MeConfig), C.c_int, C.POINTER(_vp)]),
    "icu_me_batch_destroy": (C.c_int, [_vp]),
    "icu_me_batch_num_mbs": (C.c_int, [_vp]),
    "icu_me_batch_mb_base": (C.c_int, [_vp, C.c_int]),
    "icu_me_batch_estimate_device": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_vp), _vp, _vp, _vp]),
    "icu_me_batch_capture": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_vp), _vp, _vp]),
    "icu_me_batch_graph_launch": (C.c_int, [_vp, _vp]),
    "icu_me_batch_autotune": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_vp), _vp, _vp, C.c_int,
                                        C.POINTER(C.c_float)]),
    "icu_me_batch_get_plan": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "icu_me_batch_set_plan": (C.c_int, [_vp, C.c_int, C.c_int]),
    "icu_server_create": (C.c_int, [C.POINTER(ServerConfig), C.POINTER(_vp)]),
    "icu_server_destroy": (C.c_int, [_vp]),
    "icu_server_open_stream": (C.c_int, [_vp, C.POINTER(MeConfig), C.POINTER(C.c_int)]),
    "icu_server_submit": (C.c_int, [_vp, C.c_int, _vp, C.c_longlong]),
    "icu_server_poll": (C.c_int, [_vp, C.POINTER(ServerResult), C.c_int, C.c_int]),
    "icu_server_stream_estimator": (C.c_int, [_vp, C.c_int, C.POINTER(_vp)]),
    "icu_server_stats": (C.c_int, [_vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                                   C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "icu_server_clock": (C.c_double, []),
    "icu_vbs_create": (C.c_int, [C.POINTER(VbsConfig), C.POINTER(_vp)]),
    "icu_vbs_destroy": (C.c_int, [_vp]),
    "icu_vbs_estimate_device": (C.c_int, [_vp, _vp, C.POINTER(_vp), _vp, _vp, _vp, _vp, _vp]),
    "icu_vbs_code_device": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.POINTER(C.c_int), C.c_int, C.c_int,
                                      C.POINTER(_vp), C.POINTER(_vp), _vp, _vp]),
    "icu_vbs_get_plan": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "icu_mc_create": (C.c_int, [C.c_int, C.c_int, C.POINTER(_vp)]),
    "icu_mc_destroy": (C.c_int, [_vp]),
    "icu_mc_compensate": (C.c_int, [_vp, _vp, _vp]),
    "icu_mc_compensate_device": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "icu_yuv_file_open": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(_vp)]),
    "icu_yuv_file_next_unit": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(C.c_int)]),
    "icu_yuv_file_put_unit": (C.c_int, [_vp, _vp, C.c_int]),
    "icu_yuv_file_close": (C.c_int, [_vp]),
    "icu_yuv_unpack_device": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp, _vp]),
    "icu_mv_dump": (C.c_int, [C.c_char_p, C.c_char_p, _vp, C.c_int]),
    "icu_mv_load": (C.c_int, [C.c_char_p, _vp, C.c_int, C.POINTER(C.c_int)]),
    "icu_cavlc_encode": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _vp, C.c_long, C.POINTER(C.c_long),
                                   C.POINTER(C.c_int)]),
    "icu_cavlc_decode": (C.c_int, [_vp, C.c_long, C.POINTER(C.c_long), C.c_int, C.c_int, C.c_int, _vp,
                                   C.POINTER(C.c_int)]),
    "icu_cavlc_create": (C.c_int, [C.POINTER(_vp)]),
    "icu_cavlc_destroy": (C.c_int, [_vp]),
    "icu_cavlc_encode_device": (C.c_int, [_vp, _vp, C.c_long, C.c_int, _vp, C.c_int, C.c_int, _vp, C.c_long, _vp,
                                          C.POINTER(C.c_longlong), _vp]),
    "icu_cavlc_decode_device": (C.c_int, [_vp, _vp, C.c_long, _vp, C.c_long, C.c_int, _vp, C.c_int, C.c_int, _vp,
                                          _vp, _vp]),
    "icu_umhs_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(_vp)]),
    "icu_umhs_destroy": (C.c_int, [_vp]),
    "icu_umhs_estimate": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.POINTER(C.c_long)]),
    "icu_umhs_estimate_device": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "icu_umhs_check": (C.c_int, [_vp]),
    "icu_cross_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(_vp)]),
    "icu_cross_destroy": (C.c_int, [_vp]),
    "icu_cross_estimate": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, C.POINTER(C.c_long)]),
    "icu_cross_estimate_device": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "icu_cross_check": (C.c_int, [_vp]),
    "icu_fhs_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(_vp)]),
    "icu_fhs_destroy": (C.c_int, [_vp]),
    "icu_fhs_estimate": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, C.POINTER(C.c_long)]),
    "icu_fhs_estimate_device": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "icu_fhs_check": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "icu_last_error": (C.c_char_p, []),
    "icu_device_info": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "icu_malloc": (_vp, [C.c_size_t]),
    "icu_free": (C.c_int, [_vp]),
    "icu_memcpy_htod": (C.c_int, [_vp, _vp, C.c_size_t, _vp]),
    "icu_memcpy_dtoh": (C.c_int, [_vp, _vp, C.c_size_t, _vp]),
    "icu_memset": (C.c_int, [_vp, C.c_int, C.c_size_t, _vp]),
    "icu_memcpy_dtod": (C.c_int, [_vp, _vp, C.c_size_t, _vp]),
    "icu_stream_create": (_vp, []),
    "icu_stream_destroy": (C.c_int, [_vp]),
    "icu_stream_sync": (C.c_int, [_vp]),
    "icu_event_create": (_vp, []),
    "icu_event_destroy": (C.c_int, [_vp]),
    "icu_event_record": (C.c_int, [_vp, _vp]),
    "icu_event_elapsed_ms": (C.c_int, [_vp, _vp, C.POINTER(C.c_float)]),
    "icu_sqrt_check_device": (C.c_int, [_vp, _vp, C.c_int, _vp]),
}


def header_symbols() -> list[str]:
    """Every function the public header declares."""
    import re
    text = HEADER.read_text()
    return sorted(set(re.findall(r"\b(icu_[a-z0-9_]+)\s*\(", text)))


def _load():
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is not built: run python -m imagecodingutils_amd.build")
    L = C.CDLL(str(LIB_PATH))
    for name, (res, args) in _SIGS.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


lib = _load()


def _check(rc: int) -> None:
    if rc != ICU_OK:
        raise IcuError(rc, lib.icu_last_error().decode())


# ------------------------------------------------------------ device memory
class DeviceBuffer:
    """A device allocation owned by the library's runtime."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = lib.icu_malloc(max(1, self.nbytes))
        if not self.ptr:
            raise IcuError(ICU_E_DEVICE, lib.icu_last_error().decode())

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray, stream=None) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        _check(lib.icu_memcpy_htod(self.ptr, a.ctypes.data, a.nbytes, stream))

    def download(self, dtype, shape, stream=None) -> np.ndarray:
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        _check(lib.icu_memcpy_dtoh(out.ctypes.data, self.ptr, out.nbytes, stream))
        return out

    def zero(self) -> None:
        _check(lib.icu_memset(self.ptr, 0, self.nbytes, None))
        self.sync()

    @staticmethod
    def sync(stream=None) -> None:
        _check(lib.icu_stream_sync(stream))

    def __del__(self):
        if getattr(self, "ptr", None):
            lib.icu_free(self.ptr)
            self.ptr = None


class Stream:
    def __init__(self):
        self.handle = lib.icu_stream_create()
        if not self.handle:
            raise IcuError(ICU_E_DEVICE, lib.icu_last_error().decode())

    def sync(self) -> None:
        _check(lib.icu_stream_sync(self.handle))

    def __del__(self):
        if getattr(self, "handle", None):
            lib.icu_stream_destroy(self.handle)
            self.handle = None


class Event:
    def __init__(self):
        self.handle = lib.icu_event_create()

    def record(self, stream: Stream | None) -> None:
        _check(lib.icu_event_record(self.handle, stream.handle if stream else None))

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float(0)
        _check(lib.icu_event_elapsed_ms(self.handle, end.handle, C.byref(ms)))
        return float(ms.value)

    def __del__(self):
        if getattr(self, "handle", None):
            lib.icu_event_destroy(self.handle)
            self.handle = None


def device_info() -> dict:
    name = C.create_string_buffer(256)
    cus, clk = C.c_int(0), C.c_int(0)
    _check(lib.icu_device_info(name, 256, C.byref(cus), C.byref(clk)))
    return {"name": name.value.decode(), "compute_units": cus.value, "clock_khz": clk.value}


# --------------------------------------------------------------- estimator
def reference_config(width: int, height: int, motion_range_qpel: int) -> MeConfig:
    cfg = MeConfig()
    _check(lib.icu_me_config_reference(C.byref(cfg), width, height, motion_range_qpel))
    return cfg


def make_config(width, height, R, *, metric=0, window=0, filt=1, subpel=2, predictor=1,
                best_vector=1, tile_rows=0, mbs_per_task=0, grid_mode=0) -> MeConfig:
    return MeConfig(width, height, R, metric, window, filt, subpel, predictor, best_vector,
                    tile_rows, mbs_per_task, grid_mode)


class MotionEstimator:
    """One estimator handle (icu_me)."""

    def __init__(self, cfg: MeConfig):
        self.cfg = cfg
        h = _vp()
        _check(lib.icu_me_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self.nmb = (cfg.width // 16) * (cfg.height // 16)

    def estimate(self, cur: np.ndarray, ref: np.ndarray, include=None):
        """Host int16 planes -> (mv (nMB,2) int16, dist (nMB,) int32, avg)."""
        cur = np.ascontiguousarray(cur, dtype=np.int16)
        ref = np.ascontiguousarray(ref, dtype=np.int16)
        mv = np.zeros((self.nmb, 2), np.int16)
        dist = np.zeros(self.nmb, np.int32)
        avg = C.c_long(0)
        inc = None
        if include is not None:
            include = np.ascontiguousarray(include, dtype=np.uint8)
            inc = include.ctypes.data
        _check(lib.icu_me_estimate(self.h, cur.ctypes.data, ref.ctypes.data, inc, mv.ctypes.data,
                                   dist.ctypes.data, C.byref(avg)))
        return mv, dist, avg.value

    def estimate_device(self, d_cur, d_ref, d_mv, d_dist, d_int=None, stream: Stream | None = None,
                        check: bool = False):
        """Asynchronous on `stream`.  check=True waits for the device and
        raises if a median-predictor wavefront timed out (icu_me_check)."""
        _check(lib.icu_me_estimate_device(self.h, _p(d_cur), _p(d_ref), _p(d_mv), _p(d_dist),
                                          _p(d_int), stream.handle if stream else None))
        if check:
            self.check()

    def check(self):
        """Waits for the device; IcuError(ICU_E_STATE) if a median-predictor
        wavefront of this handle timed out since the last check."""
        _check(lib.icu_me_check(self.h))

    def tie_count(self) -> int:
        """MBs of the last median-predictor estimate decided again in the
        wavefront because their integer scan met an equal cost."""
        t = C.c_int(0)
        _check(lib.icu_me_tie_count(self.h, C.byref(t)))
        return t.value

    def estimate_timed(self, d_cur, d_ref, d_mv, d_dist, stream: Stream | None = None, search_reps: int = 1):
        """Per-stage device times in ms: (pad, search, decide); the search time
        is the mean of search_reps back-to-back launches."""
        ms = (C.c_float * 3)()
        _check(lib.icu_me_estimate_timed(self.h, _p(d_cur), _p(d_ref), _p(d_mv), _p(d_dist),
                                         stream.handle if stream else None, search_reps, ms))
        return tuple(float(x) for x in ms)

    def code_frame_device(self, d_cur, d_mv, qp, intra, cols_first, d_levels, d_recon, d_pred=None,
                          stream: Stream | None = None):
        _check(lib.icu_code_frame_device(self.h, _p(d_cur), _p(d_mv), qp, intra, cols_first,
                                         _p(d_levels), _p(d_recon), _p(d_pred),
                                         stream.handle if stream else None))

    def code_frame_qps_device(self, d_cur, d_mv, qps, intra, cols_first, d_levels, d_recons, d_pred=None,
                              stream: Stream | None = None):
        """One prediction, coded at every QP in qps (d_levels / d_recons: one buffer per QP)."""
        n = len(qps)
        _check(lib.icu_code_frame_qps_device(
            self.h, _p(d_cur), _p(d_mv), n, (C.c_int * n)(*qps), intra, cols_first,
            (_vp * n)(*[_p(x) for x in d_levels]), (_vp * n)(*[_p(x) for x in d_recons]), _p(d_pred),
            stream.handle if stream else None))

    def tiling(self):
        """(candidate rows per lane, MBs per task, search grid mode 0/1/2)."""
        v, g, p = C.c_int(0), C.c_int(0), C.c_int(0)
        _check(lib.icu_me_get_tiling(self.h, C.byref(v), C.byref(g), C.byref(p)))
        return v.value, g.value, p.value

    def autotune(self, d_cur, d_ref, reps=3):
        _check(lib.icu_me_autotune(self.h, _p(d_cur), _p(d_ref), reps))

    def close(self):
        if getattr(self, "h", None):
            lib.icu_me_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def _p(x):
    if x is None:
        return None
    return x.ptr if isinstance(x, DeviceBuffer) else x


class MotionEstimatorBatch:
    """Many streams of mixed size / range estimated by one launch sequence."""

    def __init__(self, cfgs: list[MeConfig]):
        arr = (MeConfig * len(cfgs))(*cfgs)
        h = _vp()
        _check(lib.icu_me_batch_create(arr, len(cfgs), C.byref(h)))
        self.h = h
        self.n = len(cfgs)
        self.total_mbs = lib.icu_me_batch_num_mbs(h)
        self.bases = [lib.icu_me_batch_mb_base(h, i) for i in range(self.n)]

    def _ptrs(self, bufs):
        return (_vp * self.n)(*[_p(b) for b in bufs])

    def estimate_device(self, curs, refs, d_mv, d_dist, stream: Stream | None = None):
        _check(lib.icu_me_batch_estimate_device(self.h, self._ptrs(curs), self._ptrs(refs), _p(d_mv),
                                                _p(d_dist), stream.handle if stream else None))

    def capture(self, curs, refs, d_mv, d_dist):
        _check(lib.icu_me_batch_capture(self.h, self._ptrs(curs), self._ptrs(refs), _p(d_mv), _p(d_dist)))

    def graph_launch(self, stream: Stream | None = None):
        _check(lib.icu_me_batch_graph_launch(self.h, stream.handle if stream else None))

    def autotune(self, curs, refs, d_mv, d_dist, reps: int = 10) -> float:
        """Times the launch plans and keeps the fastest; returns its ms per
        estimate.  Capture again afterwards."""
        ms = C.c_float()
        _check(lib.icu_me_batch_autotune(self.h, self._ptrs(curs), self._ptrs(refs), _p(d_mv), _p(d_dist), reps,
                                         C.byref(ms)))
        return ms.value

    def set_plan(self, grid_mode: int, queue_waves: int = 6):
        _check(lib.icu_me_batch_set_plan(self.h, grid_mode, queue_waves))

    def plan(self) -> tuple[int, int]:
        """(grid mode, queue waves per SIMD)."""
        g, w = C.c_int(), C.c_int()
        _check(lib.icu_me_batch_get_plan(self.h, C.byref(g), C.byref(w)))
        return g.value, w.value

    def close(self):
        if getattr(self, "h", None):
            lib.icu_me_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


# -------------------------------------------------------------------- CAVLC
CAVLC_2X2, CAVLC_4X4, CAVLC_8X8 = 0, 1, 2
CAVLC_SIZE = {CAVLC_2X2: 4, CAVLC_4X4: 16, CAVLC_8X8: 64}


def cavlc_encode(blocks: np.ndarray, mode: int, ncs, dc_skip: int = 0):
    """Host encode of blocks (n x 4, 16 or 64 int16) one after another into
    one MSB-first stream.  Returns (stream bytes, bits per block)."""
    b = np.ascontiguousarray(blocks, np.int16)
    n = b.shape[0]
    ncs = np.broadcast_to(np.asarray(ncs, np.int64), (n,))
    cap = 64 + 1200 * n
    out = np.zeros((cap + 7) // 8, np.uint8)
    pos = C.c_long(0)
    bits = np.zeros(n, np.int32)
    nb = C.c_int()
    for i in range(n):
        _check(lib.icu_cavlc_encode(b[i].ctypes.data, mode, int(ncs[i]), dc_skip, out.ctypes.data, cap,
                                    C.byref(pos), C.byref(nb)))
        bits[i] = nb.value
    return out[:(pos.value + 7) // 8], bits


def cavlc_decode(stream: np.ndarray, nblocks: int, mode: int, ncs, dc_skip: int = 0):
    s = np.ascontiguousarray(stream, np.uint8)
    ncs = np.broadcast_to(np.asarray(ncs, np.int64), (nblocks,))
    size = CAVLC_SIZE[mode]
    out = np.zeros((nblocks, size), np.int16)
    bits = np.zeros(nblocks, np.int32)
    pos = C.c_long(0)
    nb = C.c_int()
    for i in range(nblocks):
        _check(lib.icu_cavlc_decode(s.ctypes.data, s.size * 8, C.byref(pos), mode, int(ncs[i]), dc_skip,
                                    out[i].ctypes.data, C.byref(nb)))
        bits[i] = nb.value
    return out, bits


class CavlcDevice:
    """Batched device CAVLC (icu_cavlc_*_device)."""

    def __init__(self):
        h = _vp()
        _check(lib.icu_cavlc_create(C.byref(h)))
        self.h = h

    def encode(self, d_coef, nblocks, mode, d_nc, nc, dc_skip, d_stream, stream_bytes, d_offsets,
               stream: Stream | None = None) -> int:
        tot = C.c_longlong()
        _check(lib.icu_cavlc_encode_device(self.h, _p(d_coef), nblocks, mode, _p(d_nc), nc, dc_skip, _p(d_stream),
                                           stream_bytes, _p(d_offsets), C.byref(tot),
                                           stream.handle if stream else None))
        return tot.value

    def decode(self, d_stream, stream_bytes, d_offsets, nblocks, mode, d_nc, nc, dc_skip, d_coef, d_status=None,
               stream: Stream | None = None):
        _check(lib.icu_cavlc_decode_device(self.h, _p(d_stream), stream_bytes, _p(d_offsets), nblocks, mode,
                                           _p(d_nc), nc, dc_skip, _p(d_coef), _p(d_status),
                                           stream.handle if stream else None))

    def close(self):
        if getattr(self, "h", None):
            lib.icu_cavlc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


# ------------------------------------------------------ raw YUV / MV files
YUV420_8P, YUV420_NV12, YUV420_16P = 0, 1, 16


class YuvFile:
    """YuvRawFileHandler over icu_yuv_file_* (read: next_unit(); write:
    put_unit(), written on close)."""

    def __init__(self, path, width, height, in_type=YUV420_8P, out_type=YUV420_16P, write=False):
        h = _vp()
        _check(lib.icu_yuv_file_open(str(path).encode(), int(write), width, height, in_type, out_type,
                                     C.byref(h)))
        self.h, self.W, self.H, self.out_type = h, width, height, out_type

    def next_unit(self):
        """(samples of the whole unit buffer as a copy, unit length)."""
        p, n = _vp(), C.c_int()
        _check(lib.icu_yuv_file_next_unit(self.h, C.byref(p), C.byref(n)))
        total = self.W * self.H * 3 // 2
        ct = C.c_int16 if self.out_type == YUV420_16P else C.c_uint8
        buf = np.ctypeslib.as_array(C.cast(p, C.POINTER(ct)), shape=(total,)).copy()
        return buf, n.value

    def put_unit(self, data: bytes | np.ndarray):
        b = np.ascontiguousarray(np.frombuffer(bytes(data), np.uint8) if isinstance(data, bytes) else data)
        _check(lib.icu_yuv_file_put_unit(self.h, b.ctypes.data, b.nbytes))

    def close(self):
        if getattr(self, "h", None):
            h, self.h = self.h, None
            _check(lib.icu_yuv_file_close(h))

    def __del__(self):
        if getattr(self, "h", None):
            lib.icu_yuv_file_close(self.h)
            self.h = None


def yuv_unpack_device(d_raw, in_type, width, height, d_y, d_u, d_v, stream: Stream | None = None):
    _check(lib.icu_yuv_unpack_device(_p(d_raw), in_type, width, height, _p(d_y), _p(d_u), _p(d_v),
                                     stream.handle if stream else None))


def mv_dump(path, mv: np.ndarray, title: str = ""):
    m = np.ascontiguousarray(mv, np.int16)
    _check(lib.icu_mv_dump(str(path).encode(), title.encode(), m.ctypes.data, m.size // 2))


def mv_load(path, max_n: int) -> np.ndarray:
    out = np.zeros((max_n, 2), np.int16)
    n = C.c_int()
    _check(lib.icu_mv_load(str(path).encode(), out.ctypes.data, max_n, C.byref(n)))
    return out[:n.value]


# ------------------------------------------------- YUV motion compensation
class MotionCompensator:
    """Whole-frame YUV 4:2:0 compensation (icu_mc_*)."""

    def __init__(self, width: int, height: int):
        h = _vp()
        _check(lib.icu_mc_create(width, height, C.byref(h)))
        self.h, self.W, self.H = h, width, height

    def compensate(self, y, u, v, mv):
        """Host planes in, compensated host planes out (uint8)."""
        yuv = np.concatenate([np.asarray(p, np.int16).ravel() for p in (y, u, v)])
        _check(lib.icu_mc_compensate(self.h, yuv.ctypes.data,
                                     np.ascontiguousarray(mv, np.int16).ctypes.data))
        ny, nc = self.W * self.H, (self.W // 2) * (self.H // 2)
        return (yuv[:ny].reshape(self.H, self.W).astype(np.uint8),
                yuv[ny:ny + nc].reshape(self.H // 2, self.W // 2).astype(np.uint8),
                yuv[ny + nc:].reshape(self.H // 2, self.W // 2).astype(np.uint8))

    def compensate_device(self, d_ref, d_mv, d_out, stream: Stream | None = None):
        """d_ref, d_out: (Y, U, V) device planes; d_out may be d_ref."""
        _check(lib.icu_mc_compensate_device(self.h, *[_p(x) for x in d_ref], _p(d_mv),
                                            *[_p(x) for x in d_out], stream.handle if stream else None))

    def close(self):
        if getattr(self, "h", None):
            lib.icu_mc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


# ------------------------------------------------------------ UMHS estimator
class UmhsEstimator:
    """MotionEstimatorH264ImplUMHS on the device (icu_umhs_*).  motion_range
    in quarter pels as the reference's constructor; mode 0 quarter, 1 half,
    2 full pel."""

    def __init__(self, width, height, motion_range, mode=0):
        h = _vp()
        _check(lib.icu_umhs_create(width, height, motion_range, mode, C.byref(h)))
        self.h, self.W, self.H = h, width, height
        self.nmb = (width // 16) * (height // 16)

    def estimate(self, cur, ref, prev_mv=None, prev_intra=None, include=None):
        """Host planes (int16, 0..255).  Returns (mv (nMB, 2) int16, dist
        (nMB,) int32, avg over include)."""
        c = np.ascontiguousarray(cur, np.int16)
        r = np.ascontiguousarray(ref, np.int16)
        mv = np.zeros((self.nmb, 2), np.int16)
        dist = np.zeros(self.nmb, np.int32)
        avg = C.c_long(0)
        pm = pi = inc = None
        if prev_mv is not None:
            pm = np.ascontiguousarray(prev_mv, np.int16)
            pi = np.ascontiguousarray(prev_intra if prev_intra is not None else np.zeros(self.nmb), np.uint8)
        if include is not None:
            inc = np.ascontiguousarray(include, np.uint8)
        _check(lib.icu_umhs_estimate(self.h, c.ctypes.data, r.ctypes.data, None if pm is None else pm.ctypes.data,
                                     None if pi is None else pi.ctypes.data,
                                     None if inc is None else inc.ctypes.data, mv.ctypes.data, dist.ctypes.data,
                                     C.byref(avg)))
        return mv, dist, avg.value

    def estimate_device(self, d_cur, d_ref, d_mv, d_dist, d_prev_mv=None, d_prev_intra=None,
                        stream: Stream | None = None):
        _check(lib.icu_umhs_estimate_device(self.h, _p(d_cur), _p(d_ref), _p(d_prev_mv), _p(d_prev_intra),
                                            _p(d_mv), _p(d_dist), stream.handle if stream else None))

    def check(self):
        _check(lib.icu_umhs_check(self.h))

    def close(self):
        if getattr(self, "h", None):
            lib.icu_umhs_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class CrossEstimator:
    """MotionEstimatorH264ImplCross on the device (icu_cross_*).  predictor
    0 zero, 1 H.264 median."""

    def __init__(self, width, height, motion_range, mode=0, predictor=0):
        h = _vp()
        _check(lib.icu_cross_create(width, height, motion_range, mode, predictor, C.byref(h)))
        self.h, self.W, self.H = h, width, height
        self.nmb = (width // 16) * (height // 16)

    def estimate(self, cur, ref, include=None):
        c = np.ascontiguousarray(cur, np.int16)
        r = np.ascontiguousarray(ref, np.int16)
        mv = np.zeros((self.nmb, 2), np.int16)
        dist = np.zeros(self.nmb, np.int32)
        avg = C.c_long(0)
        inc = None if include is None else np.ascontiguousarray(include, np.uint8)
        _check(lib.icu_cross_estimate(self.h, c.ctypes.data, r.ctypes.data, None if inc is None else inc.ctypes.data,
                                      mv.ctypes.data, dist.ctypes.data, C.byref(avg)))
        return mv, dist, avg.value

    def estimate_device(self, d_cur, d_ref, d_mv, d_dist, stream: Stream | None = None):
        _check(lib.icu_cross_estimate_device(self.h, _p(d_cur), _p(d_ref), _p(d_mv), _p(d_dist),
                                             stream.handle if stream else None))

    def check(self):
        _check(lib.icu_cross_check(self.h))

    def close(self):
        if getattr(self, "h", None):
            lib.icu_cross_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class FhsEstimator:
    """MotionEstimatorH264ImplFHS on the device (icu_fhs_*); one handle per
    stream, frames in order (the half-sample cache carries over)."""

    def __init__(self, width, height, motion_range, mode=0):
        h = _vp()
        _check(lib.icu_fhs_create(width, height, motion_range, mode, C.byref(h)))
        self.h, self.W, self.H = h, width, height
        self.nmb = (width // 16) * (height // 16)

    def estimate(self, cur, ref, include=None):
        c = np.ascontiguousarray(cur, np.int16)
        r = np.ascontiguousarray(ref, np.int16)
        mv = np.zeros((self.nmb, 2), np.int16)
        dist = np.zeros(self.nmb, np.int32)
        avg = C.c_long(0)
        inc = None if include is None else np.ascontiguousarray(include, np.uint8)
        _check(lib.icu_fhs_estimate(self.h, c.ctypes.data, r.ctypes.data, None if inc is None else inc.ctypes.data,
                                    mv.ctypes.data, dist.ctypes.data, C.byref(avg)))
        return mv, dist, avg.value

    def estimate_device(self, d_cur, d_ref, d_mv, d_dist, stream: Stream | None = None):
        _check(lib.icu_fhs_estimate_device(self.h, _p(d_cur), _p(d_ref), _p(d_mv), _p(d_dist),
                                           stream.handle if stream else None))

    def check(self) -> int:
        u = C.c_int(0)
        _check(lib.icu_fhs_check(self.h, C.byref(u)))
        return u.value

    def close(self):
        if getattr(self, "h", None):
            lib.icu_fhs_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


# ---------------------------------------- multi-reference, partitioned search
class VbsEstimator:
    """Multi-reference variable-block-size search and its residual coder
    (icu_vbs_*).  Device buffers in, device buffers out."""

    def __init__(self, width: int, height: int, R: int, nrefs: int, lam: int = 4):
        cfg = VbsConfig(width, height, R, nrefs, lam)
        h = _vp()
        _check(lib.icu_vbs_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self.W, self.H, self.R, self.nrefs, self.lam = width, height, R, nrefs, lam
        self.nmb = (width // 16) * (height // 16)
        self.nb8 = (width // 8) * (height // 8)

    def plan(self) -> tuple[int, int]:
        v, lds = C.c_int(), C.c_int()
        _check(lib.icu_vbs_get_plan(self.h, C.byref(v), C.byref(lds)))
        return v.value, lds.value

    def estimate_device(self, d_cur, d_refs, d_mode, d_mv8, d_ref8, d_cost, stream: Stream | None = None):
        refs = (_vp * len(d_refs))(*[_p(r) for r in d_refs])
        _check(lib.icu_vbs_estimate_device(self.h, _p(d_cur), refs, _p(d_mode), _p(d_mv8), _p(d_ref8),
                                           _p(d_cost), stream.handle if stream else None))

    def code_device(self, d_cur, d_mv8, d_ref8, qps, intra, cols_first, d_levels, d_recons, d_pred=None,
                    stream: Stream | None = None):
        n = len(qps)
        _check(lib.icu_vbs_code_device(self.h, _p(d_cur), _p(d_mv8), _p(d_ref8), n, (C.c_int * n)(*qps),
                                       intra, cols_first, (_vp * n)(*[_p(x) for x in d_levels]),
                                       (_vp * n)(*[_p(x) for x in d_recons]), _p(d_pred),
                                       stream.handle if stream else None))

    def close(self):
        if getattr(self, "h", None):
            lib.icu_vbs_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


# ------------------------------------------------------------------ serving
def server_checksum(a: np.ndarray) -> int:
    """The checksum icu_server_result carries for an array's bytes: the sum
    over bytes b[i] of (b[i] + 1) * ((i * 0x9E3779B1 mod 2^32) | 1), mod 2^64."""
    b = np.ascontiguousarray(a).view(np.uint8).ravel().astype(np.uint64)
    i = np.arange(b.size, dtype=np.uint64)
    w = ((i * np.uint64(0x9E3779B1)) & np.uint64(0xFFFFFFFF)) | np.uint64(1)
    return int(np.sum((b + np.uint64(1)) * w, dtype=np.uint64))


class Server:
    """Request loop with continuous batching over many streams (icu_server_*).

    submit() queues a frame and starts what can start; poll() returns the
    finished frames as dicts (tag, stream, frame, submit_s, done_s,
    levels_sum, recon_sum, mv as an (nmb, 2) int16 array)."""

    def __init__(self, lanes: int = 4, max_inflight: int = 16, mem_budget: int = 64 << 30, qp: int = 28):
        cfg = ServerConfig(lanes, max_inflight, mem_budget, qp)
        h = _vp()
        _check(lib.icu_server_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self._res = (ServerResult * 256)()

    def open_stream(self, cfg: MeConfig) -> int:
        sid = C.c_int()
        _check(lib.icu_server_open_stream(self.h, C.byref(cfg), C.byref(sid)))
        return sid.value

    def submit(self, stream: int, frame: np.ndarray, tag: int = 0) -> None:
        f = np.ascontiguousarray(frame, dtype=np.uint8)
        _check(lib.icu_server_submit(self.h, stream, f.ctypes.data, tag))

    def poll(self, wait: bool = False) -> list[dict]:
        n = lib.icu_server_poll(self.h, self._res, len(self._res), int(wait))
        if n < 0:
            _check(n)
        out = []
        for r in self._res[:n]:
            mv = np.ctypeslib.as_array(r.mv, shape=(2 * r.nmb,)).reshape(r.nmb, 2).copy()
            out.append(dict(tag=r.tag, stream=r.stream, frame=r.frame, submit_s=r.submit_s,
                            done_s=r.done_s, levels_sum=r.levels_sum, recon_sum=r.recon_sum, mv=mv))
        return out

    def debug_set_wave_error(self, stream: int) -> None:
        """Test hook: sets the wavefront error flag of a median-predictor
        stream's estimator (icu_me_debug_set_wave_error)."""
        h = _vp()
        _check(lib.icu_server_stream_estimator(self.h, stream, C.byref(h)))
        _check(lib.icu_me_debug_set_wave_error(h))

    def stats(self) -> dict:
        b, pk, q, f = C.c_longlong(), C.c_longlong(), C.c_int(), C.c_int()
        _check(lib.icu_server_stats(self.h, C.byref(b), C.byref(pk), C.byref(q), C.byref(f)))
        return dict(bytes_in_use=b.value, peak_bytes=pk.value, queued=q.value, inflight=f.value)

    @staticmethod
    def clock() -> float:
        return lib.icu_server_clock()

    def close(self):
        if getattr(self, "h", None):
            lib.icu_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


# --------------------------------------------------------------- transforms
def fwd4x4(blocks: np.ndarray, qp: int, intra: int, mode: int, layout: int = 0) -> np.ndarray:
    b = np.ascontiguousarray(blocks, dtype=np.int16)
    out = np.empty_like(b)
    _check(lib.icu_fwd4x4(b.ctypes.data, out.ctypes.data, b.size // 16, qp, intra, mode, layout))
    return out


def inv4x4(blocks: np.ndarray, qp: int, mode: int, layout: int = 0, cols_first: int = 0,
           weights=None) -> np.ndarray:
    b = np.ascontiguousarray(blocks, dtype=np.int16)
    out = np.empty_like(b)
    w = None
    if weights is not None:
        weights = np.ascontiguousarray(weights, dtype=np.int32)
        w = weights.ctypes.data
    _check(lib.icu_inv4x4(b.ctypes.data, out.ctypes.data, b.size // 16, qp, mode, layout, cols_first, w))
    return out
