"""CPU-side checks of the C ABI: the library loads, exports every declared
symbol, and rejects bad arguments before touching a device."""
import ctypes as C
import subprocess

import numpy as np
import pytest

from imagecodingutils_amd import capi


def test_library_exports_every_header_symbol():
    syms = capi.header_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(capi.lib, s)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", str(capi.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert set(syms) <= exported


def test_binding_covers_header():
    assert set(capi.header_symbols()) == set(capi._SIGS)


@pytest.mark.parametrize("bad", [
    dict(width=30), dict(height=0), dict(R=0), dict(R=65), dict(metric=2), dict(subpel=3),
    dict(predictor=2), dict(tile_rows=5), dict(mbs_per_task=17), dict(grid_mode=8), dict(grid_mode=-1)])
def test_create_rejects_bad_config(bad):
    kw = dict(width=64, height=48, R=8)
    kw.update(bad)
    cfg = capi.make_config(kw.pop("width"), kw.pop("height"), kw.pop("R"), **kw)
    h = C.c_void_p()
    rc = capi.lib.icu_me_create(C.byref(cfg), C.byref(h))
    assert rc == capi.ICU_E_ARG
    assert capi.lib.icu_last_error()


def test_reference_config_matches_shipped_class():
    cfg = capi.reference_config(352, 288, 64)
    assert (cfg.range, cfg.metric, cfg.window, cfg.filter, cfg.subpel, cfg.predictor,
            cfg.best_vector) == (16, 0, 0, 1, 2, 1, 1)


def test_transform_argument_errors():
    x = np.zeros((16, 16), np.int16)
    with pytest.raises(capi.IcuError) as e:
        capi.fwd4x4(x, 52, 0, 0)
    assert e.value.code == capi.ICU_E_ARG
    with pytest.raises(capi.IcuError):
        capi.fwd4x4(x, 28, 0, 3)
    with pytest.raises(capi.IcuError):
        capi.inv4x4(x[:3], 28, 0, layout=1)  # 3 blocks are not whole MBs


def test_batch_rejects_predictor_and_mixed_metrics():
    a = capi.make_config(64, 48, 8, metric=1, filt=0, predictor=1)
    with pytest.raises(capi.IcuError):
        capi.MotionEstimatorBatch([a])
    b = capi.make_config(64, 48, 8, metric=1, filt=0, predictor=0)
    c = capi.make_config(64, 48, 8, metric=0, filt=0, predictor=0)
    with pytest.raises(capi.IcuError):
        capi.MotionEstimatorBatch([b, c])


@pytest.mark.parametrize("bad", [dict(lanes=0), dict(lanes=17), dict(max_inflight=0), dict(mem_budget=0),
                                 dict(qp=-1), dict(qp=52)])
def test_server_rejects_bad_config(bad):
    kw = dict(lanes=4, max_inflight=8, mem_budget=1 << 30, qp=28)
    kw.update(bad)
    with pytest.raises(capi.IcuError) as e:
        capi.Server(**kw)
    assert e.value.code == capi.ICU_E_ARG


def test_server_checksum_definition():
    # sum over bytes of (b + 1) * ((i * 0x9E3779B1 mod 2^32) | 1), mod 2^64
    a = np.array([0, 255, 7], np.uint8)
    w = [1, 0x9E3779B1, (2 * 0x9E3779B1) % 2**32 | 1]
    assert capi.server_checksum(a) == (1 * w[0] + 256 * w[1] + 8 * w[2]) % 2**64
    b = np.array([-1, 2], np.int16)  # bytes ff ff 02 00
    wb = [(i * 0x9E3779B1) % 2**32 | 1 for i in range(4)]
    assert capi.server_checksum(b) == (256 * wb[0] + 256 * wb[1] + 3 * wb[2] + 1 * wb[3]) % 2**64
    big = np.full(1 << 16, 255, np.uint8)
    ref = sum(256 * ((i * 0x9E3779B1) % 2**32 | 1) for i in range(1 << 16)) % 2**64
    assert capi.server_checksum(big) == ref


def test_vbs_rejects_bad_configs():
    # checked before any device call
    for args in ((30, 32, 8, 1, 4), (32, 32, 0, 1, 4), (32, 32, 33, 1, 4), (32, 32, 8, 0, 4),
                 (32, 32, 8, 5, 4), (32, 32, 32, 1, 1100), (32, 32, 8, 1, -1)):
        with pytest.raises(capi.IcuError) as e:
            capi.VbsEstimator(*args)
        assert e.value.code == capi.ICU_E_ARG


def test_compensator_rejects_bad_sizes():
    for w, h in ((30, 32), (32, 0), (0, 16), (48, 40)):
        with pytest.raises(capi.IcuError) as e:
            capi.MotionCompensator(w, h)
        assert e.value.code == capi.ICU_E_ARG
