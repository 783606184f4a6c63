"""GPU parity of the Cross estimator (icu_cross_*) against the compiled
reference's fixture (16 cases) and against the oracle at 720p / 1080p with
both predictors.  Exact."""
import numpy as np
import pytest

from conftest import GOLDEN
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

capi = pytest.importorskip("imagecodingutils_amd.capi")


def test_cross_matches_reference_fixture():
    fx = np.load(GOLDEN / "cross.npz")
    sc = np.load(GOLDEN / "umhs.npz")
    for k in range(int(fx["ncases"])):
        si, qr, mode, pred = (int(v) for v in fx[f"c{k}_case"])
        cur, ref = sc[f"s{si}_cur"], sc[f"s{si}_ref"]
        H, W = cur.shape
        print(f"case {k}: {W}x{H} range {qr} mode {mode} predictor {pred}", flush=True)
        est = capi.CrossEstimator(W, H, qr, mode, pred)
        mv, dist, avg = est.estimate(cur, ref, fx[f"c{k}_include"])
        np.testing.assert_array_equal(mv, fx[f"c{k}_mv"], err_msg=f"case {k}")
        assert avg == int(fx[f"c{k}_avg"]), k
        # the per-MB distortions the oracle reports for the same choice
        _, od, _ = orc.cross_estimate(cur, ref, qr, mode, pred)
        np.testing.assert_array_equal(dist, od)
        est.close()


@pytest.mark.parametrize("W,H,qr,mode,pred,seed", [(1280, 720, 128, 0, 1, 21), (1920, 1088, 128, 1, 0, 22),
                                                    (1920, 1088, 256, 0, 1, 23)])
def test_cross_device_matches_oracle(W, H, qr, mode, pred, seed):
    from imagecodingutils_amd import synth
    fp = synth.rect_motion_pair(W, H, seed, n_rects=20, max_mv=min(32, qr // 4), sigma_cur=3.0)
    print(f"{W}x{H} range {qr} mode {mode} predictor {pred}", flush=True)
    est = capi.CrossEstimator(W, H, qr, mode, pred)
    n = est.nmb
    d_cur, d_ref = capi.DeviceBuffer.from_array(fp.cur), capi.DeviceBuffer.from_array(fp.ref)
    d_mv, d_dist = capi.DeviceBuffer(4 * n), capi.DeviceBuffer(4 * n)
    s = capi.Stream()
    est.estimate_device(d_cur, d_ref, d_mv, d_dist, stream=s)
    s.sync()
    est.check()
    want, wd, _ = orc.cross_estimate(fp.cur, fp.ref, qr, mode, pred)
    np.testing.assert_array_equal(d_mv.download(np.int16, (n, 2)), want)
    np.testing.assert_array_equal(d_dist.download(np.int32, (n,)), wd)


def test_cross_bad_arguments():
    for args in ((100, 64, 64, 0, 0), (64, 64, 4, 0, 0), (64, 64, 64, 3, 0), (64, 64, 64, 0, 2)):
        with pytest.raises(capi.IcuError):
            capi.CrossEstimator(*args)
