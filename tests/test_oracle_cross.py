"""The Cross oracle (oracle/me_oracle.c, orc_cross_estimate) against the
compiled MotionEstimatorH264ImplCross: 16 cases (zero and median
predictors, all three sub-pel modes, ranges +-8..+-64, pans past the
picture edge), vectors and the average distortion exactly.  The fixture
reaches the batch extrapolation's early rejections, both tie outcomes,
the predictor and the zero vector chosen by GetBestVector and arms outside
the window."""
import numpy as np
import pytest

from conftest import GOLDEN
from oracle import oracle as orc


@pytest.fixture(scope="module")
def fx():
    """The cases, with the scene planes of the UMHS fixture."""
    d = dict(np.load(GOLDEN / "cross.npz"))
    sc = np.load(GOLDEN / "umhs.npz")
    d.update({k: sc[k] for k in sc.files if k.startswith("s")})
    return d


def test_oracle_matches_reference_and_reaches_every_branch(fx):
    orc.cross_stats()
    for k in range(int(fx["ncases"])):
        si, qr, mode, pred = (int(v) for v in fx[f"c{k}_case"])
        mv, dist, avg = orc.cross_estimate(fx[f"s{si}_cur"], fx[f"s{si}_ref"], qr, mode, pred, fx[f"c{k}_include"])
        np.testing.assert_array_equal(mv, fx[f"c{k}_mv"], err_msg=f"case {k}")
        assert avg == int(fx[f"c{k}_avg"]), k
    st = orc.cross_stats()
    names = ["extrapolated rejection", "tie taken", "tie kept", "predictor chosen", "zero chosen",
             "arm outside window"]
    for name, v in zip(names, st):
        assert v > 0, f"fixture never reaches: {name}"


def test_path_is_64_distinct_points_in_batches():
    """The distortion samples 64 distinct pels of the 16x16 block; a
    constant error e on every pel gives 64 e^2, so the zero vector's
    distortion on a picture offset by a constant from its reference is
    exactly that."""
    rng = np.random.default_rng(3)
    ref = rng.integers(0, 200, (48, 64)).astype(np.int16)
    cur = ref + 7
    mv, dist, _ = orc.cross_estimate(cur, ref, 32, 2, 0)
    assert (dist <= 64 * 49).all()
    assert (dist[(mv == 0).all(1)] == 64 * 49).all()
