"""ctypes front of the C oracle (liboracle.so) and, where it has been built,
the reference shim (_ref/libvppref.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, the smoke entry point and the
CPU-baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
# ICU_ORACLE_LIB: another build of the same sources (the sanitizer build)
LIB = Path(os.environ.get("ICU_ORACLE_LIB", HERE / "liboracle.so")).resolve()
REF_LIB = HERE / "_ref" / "libvppref.so"

_i16p = np.ctypeslib.ndpointer(np.int16, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


class MeCfg(C.Structure):
    _fields_ = [(n, C.c_int) for n in
                ("width", "height", "range", "metric", "window", "filter", "subpel", "pred", "bestvec")]


AS_SHIPPED = dict(metric=0, window=0, filter=1, subpel=2, pred=1, bestvec=1)


def build(force: bool = False) -> None:
    if "ICU_ORACLE_LIB" in os.environ:
        return
    if force or not LIB.exists():
        subprocess.run(["make", "-s", "-C", str(HERE), "all"], check=True)


def build_ref() -> bool:
    """Compile the reference sources into _ref/ (only where they are present)."""
    if not Path(os.environ.get("ICU_REFERENCE", "/root/reference")).is_dir():
        return False
    subprocess.run(["make", "-s", "-C", str(HERE), "ref",
                    "REF=" + os.environ.get("ICU_REFERENCE", "/root/reference")], check=True)
    return REF_LIB.exists()


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(str(LIB))
        L.orc_me_estimate.argtypes = [C.POINTER(MeCfg), _i16p, _i16p, C.c_void_p, _i16p,
                                      C.c_void_p, C.c_void_p, C.POINTER(C.c_long)]
        L.orc_me_estimate.restype = C.c_int
        L.orc_me_estimate_subset.argtypes = [C.POINTER(MeCfg), _i16p, _i16p, C.c_void_p, C.c_int,
                                             _i16p, C.c_void_p, C.c_void_p, C.POINTER(C.c_long)]
        L.orc_me_estimate_phase.argtypes = [C.POINTER(MeCfg), _i16p, _i16p, C.c_void_p, C.c_int, C.c_int,
                                            _i16p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_me_estimate_phase.restype = C.c_int
        L.orc_quarter_read.argtypes = [_i16p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.c_int, _i32p]
        L.orc_fwd4x4_n.argtypes = [_i16p, C.c_long, C.c_int, C.c_int, C.c_int]
        L.orc_inv4x4_n.argtypes = [_i16p, C.c_long, C.c_int, C.c_int, C.c_int]
        L.orc_fwd_mb_n.argtypes = [_i16p, C.c_long, C.c_int, C.c_int, C.c_int]
        L.orc_inv_mb_n.argtypes = [_i16p, C.c_long, C.c_int, C.c_int, C.c_int]
        L.orc_inv4x4.argtypes = [_i16p, C.c_int, C.c_int, C.c_void_p]
        L.orc_mc_luma.argtypes = [_i16p, C.c_int, C.c_int, _i16p, _i16p]
        L.orc_block_dist.argtypes = [_i16p, _i16p, C.c_int]
        L.orc_code_residual.argtypes = [_i16p, _i16p, C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.c_int, _i16p, _i16p]
        L.orc_vbs_estimate.argtypes = [_u8p, _u8p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                       _u8p, _i16p, _u8p, _i32p]
        L.orc_vbs_estimate.restype = C.c_int
        L.orc_vbs_pred.argtypes = [_u8p, C.c_int, C.c_int, _i16p, _u8p, _i16p]
        L.orc_umhs_stats.argtypes = [C.c_void_p]
        L.orc_cross_stats.argtypes = [C.c_void_p]
        L.orc_fhs_stats.argtypes = [C.c_void_p]
        L.orc_fhs_estimate.argtypes = [_i16p, _i16p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, _i32p, _u8p,
                                       _i16p, C.c_void_p, C.POINTER(C.c_long)]
        L.orc_cross_estimate.argtypes = [_i16p, _i16p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                         _i16p, C.c_void_p, C.POINTER(C.c_long)]
        L.orc_umhs_estimate.argtypes = [_i16p, _i16p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                        C.c_void_p, _i16p, C.c_void_p, C.POINTER(C.c_long)]
        L.orc_mc_chroma.argtypes = [_u8p, C.c_int, C.c_int, _i16p, _u8p]
        _lib = L
    return _lib


def ref_lib():
    """The compiled reference shim, or None where it was not built."""
    global _ref
    if _ref is None and REF_LIB.exists():
        L = C.CDLL(str(REF_LIB))
        L.ref_me_estimate.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _i16p, _i16p,
                                      C.c_void_p, _i16p, C.POINTER(C.c_long)]
        L.ref_me_time.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _i16p, _i16p, C.c_int]
        L.ref_me_time.restype = C.c_double
        L.ref_fwd.argtypes = [C.c_int, _i16p, C.c_long, C.c_int, C.c_int, C.c_int]
        L.ref_inv.argtypes = [C.c_int, _i16p, C.c_long, C.c_int, C.c_int, C.c_void_p]
        L.ref_quarter_read.argtypes = [_i16p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.c_int, _i16p]
        L.ref_block_dist.argtypes = [_i16p, _i16p, C.c_int]
        L.ref_block_dist_lt.argtypes = [_i16p, _i16p, C.c_int, C.c_int]
        L.ref_mc_luma.argtypes = [_i16p, C.c_int, C.c_int, C.c_int, _i16p, _i16p]
        L.ref_mc_yuv.argtypes = [_i16p, C.c_int, C.c_int, C.c_int, _i16p]
        L.ref_yuv_read.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _i32p, _i32p]
        L.ref_mv_dump.argtypes = [C.c_char_p, C.c_char_p, _i16p, C.c_int]
        L.ref_cavlc.argtypes = [_i16p, C.c_int, C.c_int, _i32p, C.c_int, _u8p, C.c_int, _i32p, _i16p, _i32p]
        L.ref_cavlc_each.argtypes = [_i16p, C.c_int, C.c_int, _i32p, C.c_int, _u8p, C.c_int, _i32p, _i32p,
                                     _i16p, _i32p]
        L.ref_cross_estimate.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _i16p, _i16p, C.c_void_p,
                                         C.c_void_p, C.POINTER(C.c_long), C.c_int, C.POINTER(C.c_double)]
        L.ref_fhs_chain.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _i16p, _i16p, C.c_void_p,
                                    _i16p, _i32p, C.POINTER(C.c_long), C.POINTER(C.c_double)]
        L.ref_umhs_estimate.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _i16p, _i16p, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_long), C.c_int,
                                        C.POINTER(C.c_double)]
        _ref = L
    return _ref


def me_estimate(cur, ref, R, include=None, step=1, **modes):
    """Run the oracle estimator. Returns (mv (nMB,2) int16, dist (nMB,) int32,
    integer stage (nMB,3) int32, avg).  step > 1 decides only every step-th
    MB (zero predictor only); the rest stay zero."""
    cur = np.ascontiguousarray(cur, dtype=np.int16)
    ref = np.ascontiguousarray(ref, dtype=np.int16)
    H, W = cur.shape
    m = dict(AS_SHIPPED)
    m.update(modes)
    cfg = MeCfg(W, H, R, m["metric"], m["window"], m["filter"], m["subpel"], m["pred"], m["bestvec"])
    n = (W // 16) * (H // 16)
    mv = np.zeros(2 * n, np.int16)
    dist = np.zeros(n, np.int32)
    integ = np.zeros(3 * n, np.int32)
    avg = C.c_long(0)
    inc = None
    if include is not None:
        include = np.ascontiguousarray(include, dtype=np.uint8)
        inc = include.ctypes.data
    rc = lib().orc_me_estimate_subset(C.byref(cfg), cur, ref, inc, step, mv, dist.ctypes.data,
                                      integ.ctypes.data, C.byref(avg))
    if rc != 0:
        raise ValueError(f"oracle me_estimate failed ({rc})")
    return mv.reshape(n, 2), dist, integ.reshape(n, 3), avg.value


def me_estimate_par(cur, ref, R, step=1, threads=None, **modes):
    """me_estimate over every step-th MB (zero predictor), the MBs split
    over threads (the C calls release the GIL and write disjoint MBs).
    Returns (mv, dist, integer stage) with the undecided MBs zero."""
    import threading

    cur = np.ascontiguousarray(cur, dtype=np.int16)
    ref = np.ascontiguousarray(ref, dtype=np.int16)
    H, W = cur.shape
    m = dict(AS_SHIPPED)
    m.update(modes)
    if m["pred"]:
        raise ValueError("parallel oracle needs the zero predictor")
    cfg = MeCfg(W, H, R, m["metric"], m["window"], m["filter"], m["subpel"], m["pred"], m["bestvec"])
    n = (W // 16) * (H // 16)
    mv = np.zeros(2 * n, np.int16)
    dist = np.zeros(n, np.int32)
    integ = np.zeros(3 * n, np.int32)
    T = threads or max(1, min(16, len(os.sched_getaffinity(0))))
    T = max(1, min(T, (n + step - 1) // step))
    rcs = [0] * T

    def work(t):
        rcs[t] = lib().orc_me_estimate_phase(C.byref(cfg), cur, ref, None, step * T, step * t, mv,
                                             dist.ctypes.data, integ.ctypes.data, None)

    lib()
    ths = [threading.Thread(target=work, args=(t,)) for t in range(T)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    if any(rcs):
        raise ValueError(f"oracle me_estimate failed ({rcs})")
    return mv.reshape(n, 2), dist, integ.reshape(n, 3)


def fwd4x4(blocks, qp, intra, mode):
    b = np.ascontiguousarray(blocks, dtype=np.int16).copy()
    lib().orc_fwd4x4_n(b, b.size // 16, qp, intra, mode)
    return b


def inv4x4(blocks, qp, mode, cols_first=0):
    b = np.ascontiguousarray(blocks, dtype=np.int16).copy()
    lib().orc_inv4x4_n(b, b.size // 16, qp, mode, cols_first)
    return b


def fwd_mb(mbs, qp, intra, mode):
    b = np.ascontiguousarray(mbs, dtype=np.int16).copy()
    lib().orc_fwd_mb_n(b, b.size // 256, qp, intra, mode)
    return b


def inv_mb(mbs, qp, mode, cols_first=1):
    b = np.ascontiguousarray(mbs, dtype=np.int16).copy()
    lib().orc_inv_mb_n(b, b.size // 256, qp, mode, cols_first)
    return b


def quarter_read(ref, x, y, qx, qy):
    ref = np.ascontiguousarray(ref, dtype=np.int16)
    out = np.zeros(256, np.int32)
    lib().orc_quarter_read(ref, ref.shape[1], ref.shape[0], x, y, qx, qy, out)
    return out.reshape(16, 16)


def mc_luma(ref, mv):
    ref = np.ascontiguousarray(ref, dtype=np.int16)
    mv = np.ascontiguousarray(mv, dtype=np.int16).reshape(-1)
    out = np.zeros_like(ref)
    lib().orc_mc_luma(ref, ref.shape[1], ref.shape[0], mv, out)
    return out


def code_residual(cur, pred, qp, intra=0, cols_first=0):
    cur = np.ascontiguousarray(cur, dtype=np.int16)
    pred = np.ascontiguousarray(pred, dtype=np.int16)
    lv = np.zeros_like(cur)
    rc = np.zeros_like(cur)
    lib().orc_code_residual(cur, pred, cur.shape[1], cur.shape[0], qp, intra, cols_first, lv, rc)
    return lv, rc


def block_dist(a, b, metric):
    return lib().orc_block_dist(np.ascontiguousarray(a, dtype=np.int16).reshape(-1),
                                np.ascontiguousarray(b, dtype=np.int16).reshape(-1), metric)


def vbs_estimate(cur, refs, R, lam=4):
    """Multi-reference variable-block-size search.  refs: list of (H, W)
    uint8 planes.  Returns (mode (mbh, mbw) uint8, mv8 (H/8, W/8, 2) int16
    quarter-pel, ref8 (H/8, W/8) uint8, cost (mbh, mbw) int32)."""
    cur = np.ascontiguousarray(cur, dtype=np.uint8)
    H, W = cur.shape
    rs = np.ascontiguousarray(np.stack([np.asarray(r, np.uint8) for r in refs]))
    mode = np.zeros((H // 16) * (W // 16), np.uint8)
    cost = np.zeros_like(mode, dtype=np.int32)
    mv8 = np.zeros(2 * (H // 8) * (W // 8), np.int16)
    ref8 = np.zeros((H // 8) * (W // 8), np.uint8)
    rc = lib().orc_vbs_estimate(cur, rs.reshape(-1), len(refs), W, H, R, lam, mode, mv8, ref8, cost)
    if rc != 0:
        raise ValueError(f"oracle vbs_estimate failed ({rc})")
    return (mode.reshape(H // 16, W // 16), mv8.reshape(H // 8, W // 8, 2), ref8.reshape(H // 8, W // 8),
            cost.reshape(H // 16, W // 16))


def umhs_estimate(cur, ref, qrange, mode=0, prev_mv=None, prev_intra=None, include=None):
    """MotionEstimatorH264ImplUMHS through the oracle.  qrange: the
    reference's motionRange (quarter pels).  Returns (mv (nMB, 2) int16,
    dist (nMB,) int32, avg)."""
    cur = np.ascontiguousarray(cur, dtype=np.int16)
    ref = np.ascontiguousarray(ref, dtype=np.int16)
    H, W = cur.shape
    n = (W // 16) * (H // 16)
    mv = np.zeros(2 * n, np.int16)
    dist = np.zeros(n, np.int32)
    avg = C.c_long(0)
    pm = pi = inc = None
    if prev_mv is not None:
        prev_mv = np.ascontiguousarray(prev_mv, np.int16).reshape(-1)
        prev_intra = np.ascontiguousarray(prev_intra if prev_intra is not None else np.zeros(n), np.uint8)
        pm, pi = prev_mv.ctypes.data, prev_intra.ctypes.data
    if include is not None:
        include = np.ascontiguousarray(include, np.uint8)
        inc = include.ctypes.data
    rc = lib().orc_umhs_estimate(cur, ref, W, H, qrange, mode, C.c_void_p(pm), C.c_void_p(pi), C.c_void_p(inc),
                                 mv, dist.ctypes.data, C.byref(avg))
    if rc != 0:
        raise ValueError(f"oracle umhs_estimate failed ({rc})")
    return mv.reshape(n, 2), dist, avg.value


def cross_estimate(cur, ref, qrange, mode=0, pred=0, include=None):
    """MotionEstimatorH264ImplCross through the oracle.  Returns (mv (nMB, 2)
    int16, dist (nMB,) int32, avg)."""
    cur = np.ascontiguousarray(cur, dtype=np.int16)
    ref = np.ascontiguousarray(ref, dtype=np.int16)
    H, W = cur.shape
    n = (W // 16) * (H // 16)
    mv = np.zeros(2 * n, np.int16)
    dist = np.zeros(n, np.int32)
    avg = C.c_long(0)
    inc = None
    if include is not None:
        include = np.ascontiguousarray(include, np.uint8)
        inc = include.ctypes.data
    rc = lib().orc_cross_estimate(cur, ref, W, H, qrange, mode, pred, C.c_void_p(inc), mv, dist.ctypes.data,
                                  C.byref(avg))
    if rc != 0:
        raise ValueError(f"oracle cross_estimate failed ({rc})")
    return mv.reshape(n, 2), dist, avg.value


class FhsCache:
    """The FHS estimator's half-sample cache, carried from call to call."""

    def __init__(self):
        self.c = np.zeros(3 * 18 * 18, np.int32)
        self.w = np.zeros(3 * 18 * 18, np.uint8)


def fhs_estimate(cur, ref, qrange, mode=0, include=None, cache: FhsCache | None = None):
    """MotionEstimatorH264ImplFHS through the oracle (one call; pass the same
    cache for a chain of frames).  Returns (mv, dist, avg)."""
    cur = np.ascontiguousarray(cur, dtype=np.int16)
    ref = np.ascontiguousarray(ref, dtype=np.int16)
    H, W = cur.shape
    n = (W // 16) * (H // 16)
    cache = cache or FhsCache()
    mv = np.zeros(2 * n, np.int16)
    dist = np.zeros(n, np.int32)
    avg = C.c_long(0)
    inc = None
    if include is not None:
        include = np.ascontiguousarray(include, np.uint8)
        inc = include.ctypes.data
    rc = lib().orc_fhs_estimate(cur, ref, W, H, qrange, mode, C.c_void_p(inc), cache.c, cache.w, mv,
                                dist.ctypes.data, C.byref(avg))
    if rc != 0:
        raise ValueError(f"oracle fhs_estimate failed ({rc})")
    return mv.reshape(n, 2), dist, avg.value


def fhs_stats():
    """Branch counts of the FHS oracle since the previous call (8 counters,
    named in me_oracle.c; 0 = reads of never-written cache entries)."""
    st = np.zeros(8, np.int64)
    lib().orc_fhs_stats(st.ctypes.data)
    return st


def ref_fhs_chain(R_lib, curs, refs, qrange, mode=0, include=None):
    """Frames through one compiled reference instance.  Returns (mv (F, nMB,
    2), dist (F, nMB), avg (F,), seconds)."""
    curs = np.ascontiguousarray(np.stack(curs), dtype=np.int16)
    refs = np.ascontiguousarray(np.stack(refs), dtype=np.int16)
    F, H, W = curs.shape
    n = (W // 16) * (H // 16)
    mv = np.zeros(F * 2 * n, np.int16)
    dist = np.zeros(F * n, np.int32)
    avg = (C.c_long * F)()
    sec = C.c_double(0)
    inc = None
    if include is not None:
        include = np.ascontiguousarray(include, np.uint8)
        inc = include.ctypes.data
    rc = R_lib.ref_fhs_chain(W, H, qrange, mode, F, curs.reshape(-1), refs.reshape(-1), inc, mv, dist, avg,
                             C.byref(sec))
    if rc != 0:
        raise ValueError(f"reference fhs failed ({rc})")
    return mv.reshape(F, n, 2), dist.reshape(F, n), np.array(avg[:], np.int64), sec.value


def cross_stats():
    """Branch counts of the Cross oracle since the previous call (6
    counters, named in me_oracle.c)."""
    st = np.zeros(6, np.int64)
    lib().orc_cross_stats(st.ctypes.data)
    return st


def ref_cross_estimate(R_lib, cur, ref, qrange, mode=0, pred=0, include=None, reps=1):
    """The compiled reference class (test infrastructure).  Returns (mv, avg,
    seconds)."""
    cur = np.ascontiguousarray(cur, dtype=np.int16)
    ref = np.ascontiguousarray(ref, dtype=np.int16)
    H, W = cur.shape
    n = (W // 16) * (H // 16)
    mv = np.zeros(2 * n, np.int16)
    avg = C.c_long(0)
    sec = C.c_double(0)
    inc = None
    if include is not None:
        include = np.ascontiguousarray(include, np.uint8)
        inc = include.ctypes.data
    rc = R_lib.ref_cross_estimate(W, H, qrange, mode, pred, cur, ref, inc, mv.ctypes.data, C.byref(avg), reps,
                                  C.byref(sec))
    if rc != 0:
        raise ValueError(f"reference cross failed ({rc})")
    return mv.reshape(n, 2), avg.value, sec.value


def umhs_stats():
    """Branch counts of the UMHS oracle since the previous call (13 counters,
    named in me_oracle.c)."""
    st = np.zeros(13, np.int64)
    lib().orc_umhs_stats(st.ctypes.data)
    return st


def ref_umhs_estimate(R_lib, cur, ref, qrange, mode=0, prev_mv=None, prev_intra=None, include=None, reps=1):
    """The compiled reference class (test infrastructure).  Returns (mv, dist,
    avg, seconds)."""
    cur = np.ascontiguousarray(cur, dtype=np.int16)
    ref = np.ascontiguousarray(ref, dtype=np.int16)
    H, W = cur.shape
    n = (W // 16) * (H // 16)
    mv = np.zeros(2 * n, np.int16)
    dist = np.zeros(n, np.int32)
    avg = C.c_long(0)
    sec = C.c_double(0)
    pm = pi = inc = None
    if prev_mv is not None:
        prev_mv = np.ascontiguousarray(prev_mv, np.int16).reshape(-1)
        prev_intra = np.ascontiguousarray(prev_intra if prev_intra is not None else np.zeros(n), np.uint8)
        pm, pi = prev_mv.ctypes.data, prev_intra.ctypes.data
    if include is not None:
        include = np.ascontiguousarray(include, np.uint8)
        inc = include.ctypes.data
    rc = R_lib.ref_umhs_estimate(W, H, qrange, mode, cur, ref, pm, pi, inc, mv.ctypes.data, dist.ctypes.data,
                                 C.byref(avg), reps, C.byref(sec))
    if rc != 0:
        raise ValueError(f"reference umhs failed ({rc})")
    return mv.reshape(n, 2), dist, avg.value, sec.value


def vbs_pred(refs, mv8, ref8):
    rs = np.ascontiguousarray(np.stack([np.asarray(r, np.uint8) for r in refs]))
    H, W = rs.shape[1:]
    out = np.zeros((H, W), np.int16)
    lib().orc_vbs_pred(rs.reshape(-1), W, H, np.ascontiguousarray(mv8, np.int16).reshape(-1),
                       np.ascontiguousarray(ref8, np.uint8).reshape(-1), out)
    return out


def mc_chroma(ref_c, mv):
    """One 4:2:0 chroma plane compensated with the per-MB quarter-pel luma
    vectors (mv: (nMB, 2))."""
    ref_c = np.ascontiguousarray(ref_c, dtype=np.uint8)
    out = np.zeros_like(ref_c)
    lib().orc_mc_chroma(ref_c, ref_c.shape[1], ref_c.shape[0],
                        np.ascontiguousarray(mv, dtype=np.int16).reshape(-1), out)
    return out
