"""GPU parity of the FHS estimator (icu_fhs_*) against the compiled
reference's fixture (17 cases of three frames through one handle: the
half-sample cache the reference keeps between MBs and calls, stale reads
included) and against the oracle on chains at 720p / 1080p.  Exact."""
import numpy as np
import pytest

from conftest import GOLDEN
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

capi = pytest.importorskip("imagecodingutils_amd.capi")


def scenes():
    d = {}
    for name in ("umhs.npz", "fhs.npz"):
        f = np.load(GOLDEN / name)
        d.update({k: f[k] for k in f.files if k.startswith("s")})
    return d


def chain(cur, ref):
    return ([cur, np.roll(cur, 3, axis=1), np.roll(cur, -2, axis=0)], [ref, ref, np.roll(ref, 1, axis=1)])


def test_fhs_matches_reference_fixture():
    fx = np.load(GOLDEN / "fhs.npz")
    sc = scenes()
    for k in range(int(fx["ncases"])):
        si, qr, mode = (int(v) for v in fx[f"c{k}_case"])
        curs, refs = chain(sc[f"s{si}_cur"], sc[f"s{si}_ref"])
        H, W = curs[0].shape
        print(f"case {k}: {W}x{H} range {qr} mode {mode}", flush=True)
        est = capi.FhsEstimator(W, H, qr, mode)
        for f in range(3):
            mv, dist, avg = est.estimate(curs[f], refs[f], fx[f"c{k}_include"])
            np.testing.assert_array_equal(mv, fx[f"c{k}_mv"][f], err_msg=f"case {k} frame {f}")
            np.testing.assert_array_equal(dist, fx[f"c{k}_dist"][f], err_msg=f"case {k} frame {f}")
            assert avg == int(fx[f"c{k}_avg"][f]), (k, f)
        est.close()


@pytest.mark.parametrize("W,H,qr,mode,seed", [(1280, 720, 128, 0, 31), (1920, 1088, 128, 0, 32),
                                              (1920, 1088, 256, 1, 33)])
def test_fhs_device_chain_matches_oracle(W, H, qr, mode, seed):
    from imagecodingutils_amd import synth
    src = synth.StreamSource(W, H, seed, n_rects=16, max_mv=20, sigma=1.0)
    frames = [src.frame(t) for t in range(4)]
    est = capi.FhsEstimator(W, H, qr, mode)
    n = est.nmb
    d_cur, d_ref = capi.DeviceBuffer(W * H), capi.DeviceBuffer(W * H)
    d_mv, d_dist = capi.DeviceBuffer(4 * n), capi.DeviceBuffer(4 * n)
    s = capi.Stream()
    cache = orc.FhsCache()
    orc.fhs_stats()
    for t in range(1, 4):
        print(f"{W}x{H} frame {t}", flush=True)
        d_cur.upload(frames[t])
        d_ref.upload(frames[t - 1])
        est.estimate_device(d_cur, d_ref, d_mv, d_dist, stream=s)
        s.sync()
        est.check()
        want, wd, _ = orc.fhs_estimate(frames[t], frames[t - 1], qr, mode, None, cache)
        np.testing.assert_array_equal(d_mv.download(np.int16, (n, 2)), want)
        np.testing.assert_array_equal(d_dist.download(np.int32, (n,)), wd)
    st = orc.fhs_stats()
    print("oracle branch counts", st.tolist(), flush=True)


def test_fhs_bad_arguments():
    for args in ((100, 64, 64, 0), (64, 64, 4, 0), (64, 64, 64, 3)):
        with pytest.raises(capi.IcuError):
            capi.FhsEstimator(*args)
