"""The FHS oracle (oracle/me_oracle.c, orc_fhs_estimate) against the
compiled MotionEstimatorH264ImplFHS: 17 cases of three frames through one
instance (the estimator's half-sample cache persists from MB to MB and from
call to call), vectors, per-MB distortion and averages exactly.  The
fixture reaches quarter-pel candidates that read cache entries an earlier
MB wrote, every early exit (predictor, zero vector, after the h and b
passes of the half-pel stage, in the quarter-pel stage) and the hexagon
stage."""
import numpy as np
import pytest

from conftest import GOLDEN
from oracle import oracle as orc


def scenes():
    d = {}
    for name in ("umhs.npz", "fhs.npz"):
        f = np.load(GOLDEN / name)
        d.update({k: f[k] for k in f.files if k.startswith("s")})
    return d


def chain(cur, ref):
    return ([cur, np.roll(cur, 3, axis=1), np.roll(cur, -2, axis=0)], [ref, ref, np.roll(ref, 1, axis=1)])


@pytest.fixture(scope="module")
def fx():
    return np.load(GOLDEN / "fhs.npz")


def test_oracle_matches_reference_and_reaches_every_branch(fx):
    sc = scenes()
    orc.fhs_stats()
    total = np.zeros(8, np.int64)
    for k in range(int(fx["ncases"])):
        si, qr, mode = (int(v) for v in fx[f"c{k}_case"])
        curs, refs = chain(sc[f"s{si}_cur"], sc[f"s{si}_ref"])
        cache = orc.FhsCache()
        for f in range(3):
            mv, dist, avg = orc.fhs_estimate(curs[f], refs[f], qr, mode, fx[f"c{k}_include"], cache)
            np.testing.assert_array_equal(mv, fx[f"c{k}_mv"][f], err_msg=f"case {k} frame {f}")
            np.testing.assert_array_equal(dist, fx[f"c{k}_dist"][f], err_msg=f"case {k} frame {f}")
            assert avg == int(fx[f"c{k}_avg"][f]), (k, f)
            total += orc.fhs_stats()
    names = ["uninitialised cache read", "cache read from an earlier MB", "predictor exit", "zero-vector exit",
             "hexagon stage", "half exit after h", "half exit after b", "quarter exit"]
    assert total[0] == 0
    for name, v in list(zip(names, total))[1:]:
        assert v > 0, f"fixture never reaches: {name}"
