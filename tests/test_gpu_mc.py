"""GPU parity of the whole-frame YUV 4:2:0 compensator (icu_mc_*) against the
compiled reference's fixture and the oracle.  Integer work: exact."""
import numpy as np
import pytest

from conftest import GOLDEN
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

capi = pytest.importorskip("imagecodingutils_amd.capi")


def _device(mc, y, u, v, mv, in_place):
    d_ref = [capi.DeviceBuffer.from_array(np.ascontiguousarray(p)) for p in (y, u, v)]
    d_mv = capi.DeviceBuffer.from_array(np.ascontiguousarray(mv, np.int16))
    d_out = d_ref if in_place else [capi.DeviceBuffer(p.size) for p in (y, u, v)]
    mc.compensate_device(d_ref, d_mv, d_out)
    capi.lib.icu_stream_sync(None)
    return [d.download(np.uint8, p.shape) for d, p in zip(d_out, (y, u, v))]


def test_mc_yuv_matches_reference_fixture():
    fx = np.load(GOLDEN / "mc_yuv_cif.npz")
    H, W = fx["y"].shape
    mc = capi.MotionCompensator(W, H)
    want = (fx["out_y"], fx["out_u"], fx["out_v"])
    for got in (mc.compensate(fx["y"], fx["u"], fx["v"], fx["mv"]),
                _device(mc, fx["y"], fx["u"], fx["v"], fx["mv"], in_place=True),
                _device(mc, fx["y"], fx["u"], fx["v"], fx["mv"], in_place=False)):
        for g, w in zip(got, want):
            np.testing.assert_array_equal(g, w)


@pytest.mark.parametrize("W,H,span,seed", [(176, 144, 48, 1), (64, 48, 2000, 2), (1920, 1088, 60, 3)])
def test_mc_yuv_matches_oracle(W, H, span, seed):
    """Random planes and vectors, including vectors far past the picture
    (samples there repeat the edge).  The luma oracle reads fractional
    samples only within the reference's 19-pel border, so far vectors are
    checked on luma with whole-pel vectors and on chroma with any."""
    rng = np.random.default_rng(seed)
    y = rng.integers(0, 256, (H, W)).astype(np.uint8)
    u = rng.integers(0, 256, (H // 2, W // 2)).astype(np.uint8)
    v = rng.integers(0, 256, (H // 2, W // 2)).astype(np.uint8)
    n = (W // 16) * (H // 16)
    mv = rng.integers(-span, span + 1, (n, 2)).astype(np.int16)
    mc = capi.MotionCompensator(W, H)
    gy, gu, gv = _device(mc, y, u, v, mv, in_place=False)
    np.testing.assert_array_equal(gu, orc.mc_chroma(u, mv))
    np.testing.assert_array_equal(gv, orc.mc_chroma(v, mv))
    if span <= 60:
        np.testing.assert_array_equal(gy, orc.mc_luma(y, mv).astype(np.uint8))
    else:
        mvi = (mv // 4 * 4).astype(np.int16)
        gy, gu, gv = _device(mc, y, u, v, mvi, in_place=True)
        np.testing.assert_array_equal(gy, orc.mc_luma(y, mvi).astype(np.uint8))
        np.testing.assert_array_equal(gu, orc.mc_chroma(u, mvi))


def test_mc_zero_vectors_are_identity_and_bad_input_rejected():
    W, H = 96, 64
    rng = np.random.default_rng(5)
    y = rng.integers(0, 256, (H, W)).astype(np.uint8)
    u = rng.integers(0, 256, (H // 2, W // 2)).astype(np.uint8)
    v = rng.integers(0, 256, (H // 2, W // 2)).astype(np.uint8)
    mc = capi.MotionCompensator(W, H)
    got = mc.compensate(y, u, v, np.zeros(((W // 16) * (H // 16), 2), np.int16))
    for g, w in zip(got, (y, u, v)):
        np.testing.assert_array_equal(g, w)
    bad = y.astype(np.int16)
    bad[3, 4] = 256
    with pytest.raises(capi.IcuError) as e:
        mc.compensate(bad, u, v, np.zeros(((W // 16) * (H // 16), 2), np.int16))
    assert e.value.code == capi.ICU_E_RANGE
