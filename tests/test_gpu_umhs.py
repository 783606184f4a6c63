"""GPU parity of the UMHS estimator (icu_umhs_*) against the compiled
reference's fixture (19 cases, every branch of the search) and against the
oracle at 720p and 1080p (a chain of frames, each using the previous
frame's vectors).  Integer work and doubles rounded as the reference:
exact."""
import numpy as np
import pytest

from conftest import GOLDEN
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

capi = pytest.importorskip("imagecodingutils_amd.capi")


def case_inputs(fx, k):
    si, qr, mode, prev = (int(v) for v in fx[f"c{k}_case"])
    cur, ref = fx[f"s{si}_cur"], fx[f"s{si}_ref"]
    if prev >= 0:
        cur = np.roll(cur, 3, axis=1)
    pm = fx[f"c{k}_prev_mv"] if prev >= 0 else None
    pi = fx[f"c{k}_prev_intra"] if prev >= 0 else None
    return cur, ref, qr, mode, pm, pi, fx[f"c{k}_include"]


def test_umhs_matches_reference_fixture():
    fx = np.load(GOLDEN / "umhs.npz")
    for k in range(int(fx["ncases"])):
        cur, ref, qr, mode, pm, pi, inc = case_inputs(fx, k)
        H, W = cur.shape
        print(f"case {k}: {W}x{H} range {qr} mode {mode} prev {pm is not None}", flush=True)
        est = capi.UmhsEstimator(W, H, qr, mode)
        mv, dist, avg = est.estimate(cur, ref, pm, pi, inc)
        np.testing.assert_array_equal(mv, fx[f"c{k}_mv"], err_msg=f"case {k}")
        np.testing.assert_array_equal(dist, fx[f"c{k}_dist"], err_msg=f"case {k}")
        assert avg == int(fx[f"c{k}_avg"]), k
        est.close()


@pytest.mark.parametrize("W,H,qr,mode,seed", [(1280, 720, 128, 0, 11), (1920, 1088, 128, 1, 12)])
def test_umhs_device_chain_matches_oracle(W, H, qr, mode, seed):
    from imagecodingutils_amd import synth
    src = synth.StreamSource(W, H, seed, n_rects=16, max_mv=20)
    est = capi.UmhsEstimator(W, H, qr, mode)
    n = est.nmb
    d_cur, d_ref = capi.DeviceBuffer(W * H), capi.DeviceBuffer(W * H)
    d_mv = [capi.DeviceBuffer(4 * n) for _ in range(2)]
    d_dist = capi.DeviceBuffer(4 * n)
    d_intra = capi.DeviceBuffer.from_array(np.zeros(n, np.uint8))
    s = capi.Stream()
    prev = None
    frames = [src.frame(t) for t in range(3)]
    for t in range(1, 3):
        print(f"{W}x{H} frame {t}", flush=True)
        d_cur.upload(frames[t])
        d_ref.upload(frames[t - 1])
        out, pin = d_mv[t % 2], d_mv[(t + 1) % 2]
        est.estimate_device(d_cur, d_ref, out, d_dist, pin if prev is not None else None,
                            d_intra if prev is not None else None, stream=s)
        s.sync()
        est.check()
        got = out.download(np.int16, (n, 2))
        gd = d_dist.download(np.int32, (n,))
        want, wd, _ = orc.umhs_estimate(frames[t], frames[t - 1], qr, mode, prev,
                                        np.zeros(n, np.uint8) if prev is not None else None)
        np.testing.assert_array_equal(got, want)
        np.testing.assert_array_equal(gd, wd)
        prev = want


def test_umhs_bad_arguments():
    with pytest.raises(capi.IcuError):
        capi.UmhsEstimator(100, 64, 64, 0)
    with pytest.raises(capi.IcuError):
        capi.UmhsEstimator(64, 64, 4, 0)
    with pytest.raises(capi.IcuError):
        capi.UmhsEstimator(64, 64, 64, 3)
    est = capi.UmhsEstimator(32, 32, 64, 0)
    bad = np.full((32, 32), 300, np.int16)
    with pytest.raises(capi.IcuError) as e:
        est.estimate(bad, bad)
    assert e.value.code == capi.ICU_E_RANGE
