"""The UMHS oracle (oracle/me_oracle.c, orc_umhs_estimate) against the
compiled MotionEstimatorH264ImplUMHS: 18 cases over 9 scenes (moving
rectangles from QCIF to CIF, ranges +-8..+-64, all three sub-pel modes,
previous-frame vectors with intra MBs, include maps, pans that push the
predictor past the picture edge), vectors, per-MB distortion and average
exactly.  The fixture reaches every branch of the search."""
import numpy as np
import pytest

from conftest import GOLDEN
from oracle import oracle as orc


@pytest.fixture(scope="module")
def fx():
    return np.load(GOLDEN / "umhs.npz")


def case_inputs(fx, k):
    si, qr, mode, prev = (int(v) for v in fx[f"c{k}_case"])
    cur, ref = fx[f"s{si}_cur"], fx[f"s{si}_ref"]
    if prev >= 0:
        cur = np.roll(cur, 3, axis=1)
    pm = fx[f"c{k}_prev_mv"] if prev >= 0 else None
    pi = fx[f"c{k}_prev_intra"] if prev >= 0 else None
    return cur, ref, qr, mode, pm, pi, fx[f"c{k}_include"]


def test_oracle_matches_reference_and_reaches_every_branch(fx):
    orc.umhs_stats()
    for k in range(int(fx["ncases"])):
        cur, ref, qr, mode, pm, pi, inc = case_inputs(fx, k)
        mv, dist, avg = orc.umhs_estimate(cur, ref, qr, mode, pm, pi, inc)
        np.testing.assert_array_equal(mv, fx[f"c{k}_mv"], err_msg=f"case {k}")
        np.testing.assert_array_equal(dist, fx[f"c{k}_dist"], err_msg=f"case {k}")
        assert avg == int(fx[f"c{k}_avg"]), k
    st = orc.umhs_stats()
    names = ["cross", "5x5", "hexagon", "hexagon break", "ext-hexagon moves", "diamond moves",
             "previous candidate", "predictor truncated", "predictor chosen", "exit to diamond",
             "exit to hexagon first", "exit after cross", "exit after 5x5"]
    for name, v in zip(names, st):
        assert v > 0, f"fixture never reaches: {name}"


def test_identical_frames_give_zero_vectors():
    """Same picture as reference: the predictor (zero) costs 0 and nothing
    is strictly cheaper, so every vector and distortion is 0."""
    rng = np.random.default_rng(1)
    f = rng.integers(0, 256, (64, 80)).astype(np.int16)
    for mode in (0, 1, 2):
        mv, dist, _ = orc.umhs_estimate(f, f, 64, mode)
        assert not mv.any() and not dist.any()


def test_full_pel_mode_without_quarter_predictor_is_whole_pels(fx):
    """Mode 2 only searches whole pels; a fractional result can only be
    the predictor's own vector (the final test against it)."""
    for k in range(int(fx["ncases"])):
        cur, ref, qr, mode, pm, pi, inc = case_inputs(fx, k)
        if mode != 2:
            continue
        mv = fx[f"c{k}_mv"]
        frac = (mv % 4 != 0).any(1)
        # fractional vectors must repeat a neighbour's vector (the median)
        H, W = cur.shape
        mbw = W // 16
        for i in np.nonzero(frac)[0]:
            r, c = divmod(i, mbw)
            nb = [mv[j] for j in (i - 1 if c else -1, i - mbw if r else -1,
                                  i - mbw + 1 if r and c + 1 < mbw else -1, i - mbw - 1 if r and c else -1) if j >= 0]
            assert any((mv[i] == v).all() for v in nb) or not nb, (k, i)
