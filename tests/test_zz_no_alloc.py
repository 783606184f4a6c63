"""icu_me_estimate_device_noalloc: the estimate that allocates nothing (its
device scratch is a static array) equals icu_me_estimate_device's output,
on two calls in a row with the second picture smaller, and refuses what it
cannot do without allocating.  Integer work: exact comparisons."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

capi = pytest.importorskip("imagecodingutils_amd.capi")
from imagecodingutils_amd import synth  # noqa: E402

BENCH = dict(metric=1, window=1, filt=0, subpel=1, predictor=0, best_vector=0)


def _noalloc(cfg, d_cur, d_ref, d_mv, d_dist, d_int, s):
    capi._check(capi.lib.icu_me_estimate_device_noalloc(C.byref(cfg), d_cur.ptr, d_ref.ptr, d_mv.ptr, d_dist.ptr,
                                                        d_int.ptr if d_int is not None else None, s.handle))


def _both(w, h, R, seed, modes, s):
    """(no-allocation outputs, handle outputs) for one frame pair."""
    print(f"no-alloc: {w}x{h} R {R} {modes}", flush=True)
    fp = synth.rect_motion_pair(w, h, seed, n_rects=12, max_mv=min(R, 24))
    cfg = capi.make_config(w, h, R, **modes)
    n = (w // 16) * (h // 16)
    d_cur, d_ref = capi.DeviceBuffer.from_array(fp.cur), capi.DeviceBuffer.from_array(fp.ref)
    out = []
    for use_handle in (False, True):
        d_mv, d_dist, d_int = capi.DeviceBuffer(4 * n), capi.DeviceBuffer(4 * n), capi.DeviceBuffer(12 * n)
        if use_handle:
            est = capi.MotionEstimator(cfg)
            est.estimate_device(d_cur, d_ref, d_mv, d_dist, d_int, stream=s)
        else:
            _noalloc(cfg, d_cur, d_ref, d_mv, d_dist, d_int, s)
        s.sync()
        out.append((d_mv.download(np.int16, (n, 2)), d_dist.download(np.int32, n), d_int.download(np.int32, (n, 3))))
        if use_handle:
            est.close()
    return out


def _same(a, b, what):
    for x, y, name in zip(a, b, ("vectors", "distortions", "integer stage")):
        np.testing.assert_array_equal(x, y, err_msg=f"{what}: {name}")


def test_noalloc_equals_handle_path_on_two_calls_the_second_smaller():
    s = capi.Stream()
    big = _both(1920, 1088, 32, 1080, BENCH, s)
    _same(big[0], big[1], "1080p")
    assert (big[0][0] != 0).any()
    small = _both(352, 288, 16, 77, BENCH, s)  # the scratch keeps the 1080p call's contents
    _same(small[0], small[1], "CIF after 1080p")


def test_noalloc_other_modes_and_ranges():
    s = capi.Stream()
    for (w, h, R, modes) in [
        (1280, 720, 64, BENCH),
        (176, 144, 8, dict(metric=0, window=0, filt=0, subpel=2, predictor=0, best_vector=1)),
        (352, 288, 24, dict(metric=1, window=0, filt=0, subpel=2, predictor=0, best_vector=1)),
        (3840, 2160, 16, BENCH),
    ]:
        a, b = _both(w, h, R, 5 + R, modes, s)
        _same(a, b, f"{w}x{h} R {R}")


def test_noalloc_refusals():
    d = capi.DeviceBuffer(16)
    for cfg, code in [
        (capi.make_config(352, 288, 16, metric=1, window=1, filt=1, subpel=1, predictor=0), capi.ICU_E_ARG),
        (capi.make_config(352, 288, 16, metric=1, window=1, filt=0, subpel=1, predictor=1), capi.ICU_E_ARG),
        (capi.make_config(8192, 4096, 16, **BENCH), capi.ICU_E_RANGE),
        (capi.make_config(100, 288, 16, **BENCH), capi.ICU_E_ARG),
    ]:
        rc = capi.lib.icu_me_estimate_device_noalloc(C.byref(cfg), d.ptr, d.ptr, d.ptr, d.ptr, None, None)
        assert rc == code, capi.lib.icu_last_error().decode()
