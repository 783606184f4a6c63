"""GPU parity: the HIP path through the C ABI against the oracle and the
reference's golden fixtures.  Integer work, so every comparison is exact."""
import numpy as np
import pytest

from conftest import GOLDEN
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

capi = pytest.importorskip("imagecodingutils_amd.capi")
from imagecodingutils_amd import synth  # noqa: E402


def _cfg(w, h, R, **m):
    base = dict(metric=0, window=0, filt=1, subpel=2, predictor=1, best_vector=1)
    base.update(m)
    return capi.make_config(w, h, R, **base)


def _oracle_modes(m):
    return dict(metric=m.get("metric", 0), window=m.get("window", 0), filter=m.get("filt", 1),
                subpel=m.get("subpel", 2), pred=m.get("predictor", 1), bestvec=m.get("best_vector", 1))


# ------------------------------------------------------------- transforms
def test_forward_transform_matches_reference_fixtures():
    fx = np.load(GOLDEN / "transforms.npz")
    for qp in range(52):
        for intra in (0, 1):
            for mode in (0, 1, 2):
                got = capi.fwd4x4(fx["fwd_in"], qp, intra, mode, layout=0)
                np.testing.assert_array_equal(got, fx["fwd"][qp, intra, mode], err_msg=f"qp{qp} i{intra} m{mode}")
                got = capi.fwd4x4(fx["fwd_mb_in"], qp, intra, mode, layout=1)
                np.testing.assert_array_equal(got, fx["fwd_mb"][qp, intra, mode])


def test_inverse_transform_matches_reference_fixtures():
    fx = np.load(GOLDEN / "transforms.npz")
    for qp in range(52):
        x = fx["inv_in"][qp]
        for mode in (0, 1, 2):
            np.testing.assert_array_equal(capi.inv4x4(x, qp, mode, 0, 0), fx["inv"][qp, mode])
            np.testing.assert_array_equal(capi.inv4x4(x.reshape(4, 256), qp, mode, 1, 1), fx["inv_mb"][qp, mode])
        np.testing.assert_array_equal(capi.inv4x4(x, qp, 0, 0, 0, weights=fx["weights"]),
                                      fx["inv_weighted"][qp])


def test_device_transform_kernels_match_reference_fixtures():
    """The device kernels themselves (icu_*4x4_device; host calls of a few
    blocks run on the CPU), both layouts, both inverse orders, SetScale
    weights, against the compiled reference's outputs."""
    fx = np.load(GOLDEN / "transforms.npz")
    s = capi.Stream()

    def run(fn, x, n, *args, dw=None):
        x = np.ascontiguousarray(x, np.int16)
        d_in, d_out = capi.DeviceBuffer.from_array(x), capi.DeviceBuffer(x.nbytes)
        if fn == "fwd":
            capi._check(capi.lib.icu_fwd4x4_device(d_in.ptr, d_out.ptr, n, *args, s.handle))
        else:
            capi._check(capi.lib.icu_inv4x4_device(d_in.ptr, d_out.ptr, n, *args,
                                                    dw.ptr if dw is not None else None, s.handle))
        s.sync()
        return d_out.download(np.int16, x.shape)

    for qp in range(0, 52, 3):
        for intra in (0, 1):
            for mode in (0, 1, 2):
                np.testing.assert_array_equal(run("fwd", fx["fwd_in"], len(fx["fwd_in"]), qp, intra, mode, 0),
                                              fx["fwd"][qp, intra, mode])
                np.testing.assert_array_equal(
                    run("fwd", fx["fwd_mb_in"], 16 * len(fx["fwd_mb_in"]), qp, intra, mode, 1),
                    fx["fwd_mb"][qp, intra, mode])
    dw = capi.DeviceBuffer.from_array(np.ascontiguousarray(fx["weights"], np.int32))
    for qp in range(52):
        x = fx["inv_in"][qp]
        for mode in (0, 1, 2):
            np.testing.assert_array_equal(run("inv", x, len(x), qp, mode, 0, 0), fx["inv"][qp, mode])
            np.testing.assert_array_equal(run("inv", x.reshape(4, 256), len(x), qp, mode, 1, 1),
                                          fx["inv_mb"][qp, mode])
        np.testing.assert_array_equal(run("inv", x, len(x), qp, 0, 0, 0, dw=dw), fx["inv_weighted"][qp])


def test_transforms_match_oracle_on_large_batch():
    rng = np.random.default_rng(11)
    x = rng.integers(-255, 256, (20000, 16)).astype(np.int16)
    for qp in (0, 17, 22, 28, 34, 40, 51):
        lv = capi.fwd4x4(x, qp, 0, 0)
        np.testing.assert_array_equal(lv, orc.fwd4x4(x, qp, 0, 0))
        for cf in (0, 1):
            np.testing.assert_array_equal(capi.inv4x4(lv, qp, 0, 0, cf), orc.inv4x4(lv, qp, 0, cf))


# ------------------------------------------------------- final-choice sqrt
def test_sqrt_is_correctly_rounded_over_the_vector_domain():
    n = 1 << 21  # |v - pred|^2 stays below 2 * 640^2 < 2^20
    m = np.arange(n, dtype=np.int32)
    dm = capi.DeviceBuffer.from_array(m)
    dout = capi.DeviceBuffer(8 * n)
    capi._check(capi.lib.icu_sqrt_check_device(dm.ptr, dout.ptr, n, None))
    capi.DeviceBuffer.sync()
    got = dout.download(np.float64, n)
    np.testing.assert_array_equal(got, np.sqrt(m.astype(np.float64)))


# ----------------------------------------------------- estimator, shipped
ME_FIXTURES = sorted(GOLDEN.glob("me_*.npz"))


@pytest.mark.parametrize("path", ME_FIXTURES, ids=lambda p: p.stem)
def test_as_shipped_estimator_matches_reference_fixture(path):
    fx = np.load(path)
    H, W = fx["cur"].shape
    R = int(fx["R"])
    est = capi.MotionEstimator(capi.reference_config(W, H, 4 * R) if int(fx["pred"]) else
                               _cfg(W, H, R, predictor=0))
    inc = fx["include"] if "include" in fx.files else None
    mv, dist, avg = est.estimate(fx["cur"].astype(np.int16), fx["ref"].astype(np.int16), include=inc)
    np.testing.assert_array_equal(mv, fx["mv"])
    assert avg == int(fx["avg"])


# ----------------------------------------------------------- mode matrix
MODES = [
    dict(metric=1, window=1, filt=0, subpel=1, predictor=0, best_vector=1),   # bench mode
    dict(metric=1, window=0, filt=0, subpel=0, predictor=0, best_vector=0),
    dict(metric=1, window=1, filt=0, subpel=2, predictor=0, best_vector=0),
    dict(metric=0, window=1, filt=0, subpel=1, predictor=0, best_vector=1),
    dict(metric=1, window=0, filt=1, subpel=2, predictor=0, best_vector=1),
    dict(metric=1, window=1, filt=1, subpel=2, predictor=1, best_vector=1),
    dict(metric=0, window=0, filt=0, subpel=2, predictor=1, best_vector=1),
    dict(metric=1, window=0, filt=0, subpel=1, predictor=1, best_vector=0),
]


@pytest.mark.parametrize("modes", MODES, ids=lambda m: "-".join(f"{k}{v}" for k, v in m.items()))
@pytest.mark.parametrize("size,R,seed", [((176, 144), 8, 1), ((352, 288), 16, 2), ((64, 48), 32, 3)])
def test_estimator_modes_match_oracle(modes, size, R, seed):
    W, H = size
    fp = synth.rect_motion_pair(W, H, seed, n_rects=6, max_mv=min(R, 12))
    est = capi.MotionEstimator(_cfg(W, H, R, **modes))
    d_cur = capi.DeviceBuffer.from_array(fp.cur)
    d_ref = capi.DeviceBuffer.from_array(fp.ref)
    n = est.nmb
    d_mv, d_dist, d_int = capi.DeviceBuffer(4 * n), capi.DeviceBuffer(4 * n), capi.DeviceBuffer(12 * n)
    est.estimate_device(d_cur, d_ref, d_mv, d_dist, d_int)
    capi.DeviceBuffer.sync()
    capi.lib.icu_stream_sync(None)
    mv = d_mv.download(np.int16, (n, 2))
    dist = d_dist.download(np.int32, n)
    integ = d_int.download(np.int32, (n, 3))
    omv, odist, ointeg, _ = orc.me_estimate(fp.cur, fp.ref, R, **_oracle_modes(modes))
    np.testing.assert_array_equal(integ, ointeg)
    np.testing.assert_array_equal(mv, omv)
    np.testing.assert_array_equal(dist, odist)


def test_bench_workload_1080p_whole_frame_matches_oracle():
    """The benchmark's step at its full size: 1080p, +-32 exhaustive SAD,
    half-pel, then MC and residual coding at QP 22/28/34 in one pass.  All
    8160 MBs' vectors, distortions and integer-stage results equal the
    oracle's, the prediction equals the oracle's compensation, and every QP's
    levels and reconstruction equal the oracle's coding of the whole frame."""
    W, H, R = 1920, 1088, 32
    qps = [22, 28, 34]
    fp = synth.rect_motion_pair(W, H, 1080, n_rects=24, max_mv=32, sigma_cur=3.0)
    modes = dict(metric=1, window=1, filt=0, subpel=1, predictor=0, best_vector=0)
    est = capi.MotionEstimator(_cfg(W, H, R, **modes))
    n = est.nmb
    d_cur, d_ref = capi.DeviceBuffer.from_array(fp.cur), capi.DeviceBuffer.from_array(fp.ref)
    d_mv, d_dist, d_int = capi.DeviceBuffer(4 * n), capi.DeviceBuffer(4 * n), capi.DeviceBuffer(12 * n)
    d_lv = [capi.DeviceBuffer(2 * W * H) for _ in qps]
    d_rc = [capi.DeviceBuffer(W * H) for _ in qps]
    d_pr = capi.DeviceBuffer(2 * W * H)
    s = capi.Stream()
    est.estimate_device(d_cur, d_ref, d_mv, d_dist, d_int, stream=s)
    est.code_frame_qps_device(d_cur, d_mv, qps, 0, 0, d_lv, d_rc, d_pr, stream=s)
    s.sync()
    mv = d_mv.download(np.int16, (n, 2))
    omv, odist, ointeg = orc.me_estimate_par(fp.cur, fp.ref, R, **_oracle_modes(modes))
    np.testing.assert_array_equal(d_int.download(np.int32, (n, 3)), ointeg)
    np.testing.assert_array_equal(mv, omv)
    np.testing.assert_array_equal(d_dist.download(np.int32, n), odist)
    truth = (4 * fp.true_mv.reshape(-1, 2))
    assert (np.abs(mv - truth) <= 2).all(axis=1).mean() > 0.8
    pred = orc.mc_luma(fp.ref, omv)
    np.testing.assert_array_equal(d_pr.download(np.int16, (H, W)), pred)
    for i, qp in enumerate(qps):
        lv, rc = orc.code_residual(fp.cur, pred, qp, 0, 0)
        np.testing.assert_array_equal(d_lv[i].download(np.int16, (H, W)), lv, err_msg=f"QP {qp}")
        np.testing.assert_array_equal(d_rc[i].download(np.uint8, (H, W)).astype(np.int16), rc, err_msg=f"QP {qp}")


def test_estimate_is_deterministic_and_rerunnable():
    fp = synth.rect_motion_pair(352, 288, 5, n_rects=8, max_mv=16)
    est = capi.MotionEstimator(_cfg(352, 288, 16))
    a = est.estimate(fp.cur.astype(np.int16), fp.ref.astype(np.int16))
    b = est.estimate(fp.cur.astype(np.int16), fp.ref.astype(np.int16))
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])


@pytest.mark.parametrize("tile_rows,mbs_per_task,R,filt", [(8, 1, 24, 0), (11, 5, 24, 0), (13, 16, 24, 0),
                                                           (16, 3, 16, 1), (8, 2, 64, 0), (13, 4, 16, 1),
                                                           (13, 2, 64, 0), (11, 3, 64, 0), (13, 3, 40, 0),
                                                           (11, 1, 47, 0)])
def test_explicit_tilings_match_oracle(tile_rows, mbs_per_task, R, filt):
    """Argmin tilings (decided inside the search kernel) and cost-table
    tilings (decided by the replay kernel)."""
    W, H = 352, 288
    fp = synth.rect_motion_pair(W, H, 31, n_rects=8, max_mv=min(R - 4, 30))
    modes = dict(metric=1, window=1, filt=filt, subpel=1, predictor=0, best_vector=0)
    est = capi.MotionEstimator(_cfg(W, H, R, tile_rows=tile_rows, mbs_per_task=mbs_per_task, **modes))
    assert est.tiling()[:2] == (tile_rows, mbs_per_task)
    mv, dist, _ = est.estimate(fp.cur.astype(np.int16), fp.ref.astype(np.int16))
    omv, odist, _, _ = orc.me_estimate(fp.cur, fp.ref, R, **_oracle_modes(modes))
    np.testing.assert_array_equal(mv, omv)
    np.testing.assert_array_equal(dist, odist)


def test_argmin_keys_tie_break_on_flat_and_periodic_content():
    """Flat areas and a period-4 texture make many candidates cost the same:
    the lane-local argmin keys must still pick the reference's candidate
    (closest to the predictor, then latest in raster order)."""
    W, H, R = 128, 96, 20
    yy, xx = np.mgrid[0:H, 0:W]
    cur = np.where((xx // 32 + yy // 32) % 2 == 0, 100, ((xx % 4) * 40 + (yy % 4) * 7)).astype(np.uint8)
    ref = np.roll(cur, (3, -5), axis=(0, 1))
    ref[:, :16] = 100
    modes = dict(metric=1, window=1, filt=0, subpel=0, predictor=0, best_vector=0)
    for V in (8, 11, 13):
        est = capi.MotionEstimator(_cfg(W, H, R, tile_rows=V, **modes))
        mv, dist, _ = est.estimate(cur.astype(np.int16), ref.astype(np.int16))
        omv, odist, _, _ = orc.me_estimate(cur, ref, R, **_oracle_modes(modes))
        np.testing.assert_array_equal(mv, omv)
        np.testing.assert_array_equal(dist, odist)


@pytest.mark.parametrize("R", [64, 37])
def test_split_argmin_keys_tie_break_at_large_ranges(R):
    """Tiles taller than 8 rows keep one key per half; far from the window
    centre the halves' distances differ most.  Flat and periodic content with
    wide windows makes the tie-breaks decide."""
    W, H = 352, 288
    yy, xx = np.mgrid[0:H, 0:W]
    cur = np.where((xx // 48 + yy // 48) % 2 == 0, 90, ((xx % 4) * 30 + (yy % 8) * 9)).astype(np.uint8)
    ref = np.roll(cur, (-R + 3, R - 5), axis=(0, 1))
    modes = dict(metric=1, window=1, filt=0, subpel=0, predictor=0, best_vector=0)
    omv, odist, _, _ = orc.me_estimate(cur, ref, R, **_oracle_modes(modes))
    for V in (8, 11, 13):
        est = capi.MotionEstimator(_cfg(W, H, R, tile_rows=V, mbs_per_task=2, **modes))
        mv, dist, _ = est.estimate(cur.astype(np.int16), ref.astype(np.int16))
        np.testing.assert_array_equal(mv, omv, err_msg=f"V {V}")
        np.testing.assert_array_equal(dist, odist)


# One MB per wave (grid mode 3, config grid_mode 4): the default for SAD
# argmin streams at +-7/8, +-15/16, +-31/32 and +-63/64 (the last as four
# quadrant passes).  Pictures small enough that every window clamps (some
# quadrants of the 64 class lie wholly outside and are skipped), the
# reference and exhaustive windows, all sub-pel depths and the final vector
# choice; the queue kernel (config grid_mode 5) and the workgroup kernel
# (grid mode 0) must agree too.
@pytest.mark.parametrize("R", [8, 16, 32, 64])
@pytest.mark.parametrize("size,seed", [((64, 48), 1), ((176, 144), 2), ((352, 288), 3), ((96, 160), 4)])
@pytest.mark.parametrize("window,subpel,bestvec", [(1, 1, 0), (0, 2, 1), (1, 0, 0), (0, 1, 1)])
def test_wave_kernel_matches_oracle(R, size, seed, window, subpel, bestvec):
    W, H = size
    fp = synth.rect_motion_pair(W, H, 60 + seed + R, n_rects=8, max_mv=min(W // 3, H // 3, R))
    modes = dict(metric=1, window=window, filt=0, subpel=subpel, predictor=0, best_vector=bestvec)
    omv, odist, ointeg, _ = orc.me_estimate(fp.cur, fp.ref, R, **_oracle_modes(modes))
    for gm in (4, 5, 1, 0):
        est = capi.MotionEstimator(_cfg(W, H, R, grid_mode=gm, **modes))
        if gm != 1:
            assert est.tiling()[2] == {4: 3, 5: 4, 0: 3}[gm]
        d_cur, d_ref = capi.DeviceBuffer.from_array(fp.cur), capi.DeviceBuffer.from_array(fp.ref)
        n = est.nmb
        d_mv, d_dist, d_int = capi.DeviceBuffer(4 * n), capi.DeviceBuffer(4 * n), capi.DeviceBuffer(12 * n)
        est.estimate_device(d_cur, d_ref, d_mv, d_dist, d_int)
        capi.lib.icu_stream_sync(None)
        np.testing.assert_array_equal(d_int.download(np.int32, (n, 3)), ointeg, err_msg=f"grid {gm}")
        np.testing.assert_array_equal(d_mv.download(np.int16, (n, 2)), omv, err_msg=f"grid {gm}")
        np.testing.assert_array_equal(d_dist.download(np.int32, n), odist, err_msg=f"grid {gm}")


@pytest.mark.parametrize("R", [8, 16, 32, 64])
@pytest.mark.parametrize("window", [0, 1])
def test_wave_kernel_tie_break_on_flat_and_periodic_content(window, R):
    """Many equal costs across the wave's tiles, its leftover column and row
    and the (R, R) corner: the wave's key reduction must pick the
    reference's candidate."""
    W, H = 352, 288
    yy, xx = np.mgrid[0:H, 0:W]
    cur = np.where((xx // 48 + yy // 48) % 2 == 0, 90, ((xx % 4) * 30 + (yy % 8) * 9)).astype(np.uint8)
    ref = np.roll(cur, (-R + 3, R - 5), axis=(0, 1))
    ref[:, 300:] = 90
    modes = dict(metric=1, window=window, filt=0, subpel=0, predictor=0, best_vector=0)
    omv, odist, _, _ = orc.me_estimate(cur, ref, R, **_oracle_modes(modes))
    est = capi.MotionEstimator(_cfg(W, H, R, **modes))
    assert est.tiling()[2] == 3
    mv, dist, _ = est.estimate(cur.astype(np.int16), ref.astype(np.int16))
    np.testing.assert_array_equal(mv, omv)
    np.testing.assert_array_equal(dist, odist)


@pytest.mark.parametrize("tile_rows,R", [(16, 32), (16, 8), (16, 64)])
def test_too_tall_tile_for_exact_keys_is_rejected(tile_rows, R):
    with pytest.raises(capi.IcuError) as e:
        capi.MotionEstimator(_cfg(352, 288, R, tile_rows=tile_rows, metric=1, window=1, filt=0, predictor=0))
    assert e.value.code == capi.ICU_E_ARG


def test_tuned_estimator_matches_oracle():
    """The tuner retiles the handle (rows per lane, MBs per task, grid mode)
    many times; results afterwards must not change."""
    W, H, R = 352, 288, 16
    fp = synth.rect_motion_pair(W, H, 32, n_rects=8, max_mv=14)
    modes = dict(metric=1, window=1, filt=0, subpel=1, predictor=0, best_vector=0)
    est = capi.MotionEstimator(_cfg(W, H, R, **modes))
    d_cur, d_ref = capi.DeviceBuffer.from_array(fp.cur), capi.DeviceBuffer.from_array(fp.ref)
    est.autotune(d_cur, d_ref, reps=1)
    v, g, p = est.tiling()
    assert v in (8, 11, 13, 16) and 1 <= g <= 16 and p in (0, 1, 2, 3, 4)
    mv, dist, _ = est.estimate(fp.cur.astype(np.int16), fp.ref.astype(np.int16))
    omv, odist, _, _ = orc.me_estimate(fp.cur, fp.ref, R, **_oracle_modes(modes))
    np.testing.assert_array_equal(mv, omv)
    np.testing.assert_array_equal(dist, odist)


def test_timed_estimate_stages_and_results():
    """icu_me_estimate_timed: the same vectors as the plain device estimate;
    the decision stage is 0 when the search kernel decides every MB (no cost
    table) and positive when decision kernels run (filter on)."""
    W, H, R = 352, 288, 16
    fp = synth.rect_motion_pair(W, H, 77, n_rects=8, max_mv=14)
    d_cur, d_ref = capi.DeviceBuffer.from_array(fp.cur), capi.DeviceBuffer.from_array(fp.ref)
    n = (W // 16) * (H // 16)
    for filt, decide_runs in ((0, False), (1, True)):
        est = capi.MotionEstimator(_cfg(W, H, R, metric=1, window=1, filt=filt, subpel=1, predictor=0,
                                        best_vector=0))
        d_mv, d_dist = capi.DeviceBuffer(4 * n), capi.DeviceBuffer(4 * n)
        e_mv, e_dist = capi.DeviceBuffer(4 * n), capi.DeviceBuffer(4 * n)
        s = capi.Stream()
        pad_ms, search_ms, decide_ms = est.estimate_timed(d_cur, d_ref, d_mv, d_dist, stream=s, search_reps=3)
        est.estimate_device(d_cur, d_ref, e_mv, e_dist, stream=s)
        s.sync()
        assert pad_ms > 0 and search_ms > 0
        assert (decide_ms > 0) == decide_runs
        np.testing.assert_array_equal(d_mv.download(np.int16, 2 * n), e_mv.download(np.int16, 2 * n))
        np.testing.assert_array_equal(d_dist.download(np.int32, n), e_dist.download(np.int32, n))


def test_out_of_range_input_is_rejected():
    est = capi.MotionEstimator(_cfg(64, 48, 8))
    cur = np.full((48, 64), 300, np.int16)
    with pytest.raises(capi.IcuError) as e:
        est.estimate(cur, cur)
    assert e.value.code == capi.ICU_E_RANGE


# ----------------------------------------------------- MC + residual coding
def test_motion_compensation_matches_reference_fixture():
    fx = np.load(GOLDEN / "mc_cif.npz")
    ref = fx["ref"]
    H, W = ref.shape
    est = capi.MotionEstimator(_cfg(W, H, 16, predictor=0))
    d_ref = capi.DeviceBuffer.from_array(ref)
    n = est.nmb
    d_mv, d_dist = capi.DeviceBuffer(4 * n), capi.DeviceBuffer(4 * n)
    est.estimate_device(d_ref, d_ref, d_mv, d_dist)  # pads the reference
    d_mv.upload(fx["mv"].astype(np.int16))
    d_lv, d_rc, d_pr = capi.DeviceBuffer(2 * W * H), capi.DeviceBuffer(W * H), capi.DeviceBuffer(2 * W * H)
    est.code_frame_device(d_ref, d_mv, 28, 0, 0, d_lv, d_rc, d_pr)
    capi.lib.icu_stream_sync(None)
    capi.DeviceBuffer.sync(None)
    import ctypes
    pred = d_pr.download(np.int16, (H, W))
    np.testing.assert_array_equal(pred, fx["pred"].astype(np.int16))


@pytest.mark.parametrize("qp", [22, 28, 34])
@pytest.mark.parametrize("cols_first", [0, 1])
def test_code_frame_matches_oracle(qp, cols_first):
    W, H, R = 352, 288, 16
    fp = synth.rect_motion_pair(W, H, 40 + qp, n_rects=8, max_mv=12)
    est = capi.MotionEstimator(_cfg(W, H, R, metric=1, window=1, filt=0, subpel=2, predictor=0))
    d_cur, d_ref = capi.DeviceBuffer.from_array(fp.cur), capi.DeviceBuffer.from_array(fp.ref)
    n = est.nmb
    d_mv, d_dist = capi.DeviceBuffer(4 * n), capi.DeviceBuffer(4 * n)
    d_lv, d_rc, d_pr = capi.DeviceBuffer(2 * W * H), capi.DeviceBuffer(W * H), capi.DeviceBuffer(2 * W * H)
    est.estimate_device(d_cur, d_ref, d_mv, d_dist)
    est.code_frame_device(d_cur, d_mv, qp, 0, cols_first, d_lv, d_rc, d_pr)
    capi.lib.icu_stream_sync(None)
    mv = d_mv.download(np.int16, (n, 2))
    pred = orc.mc_luma(fp.ref, mv)
    np.testing.assert_array_equal(d_pr.download(np.int16, (H, W)), pred)
    lv, rc = orc.code_residual(fp.cur, pred, qp, 0, cols_first)
    np.testing.assert_array_equal(d_lv.download(np.int16, (H, W)), lv)
    np.testing.assert_array_equal(d_rc.download(np.uint8, (H, W)).astype(np.int16), rc)


@pytest.mark.parametrize("intra", [0, 1])
def test_code_frame_at_several_qps_matches_oracle(intra):
    """One prediction coded at QP 22..40 in one pass equals the oracle at
    each QP, with quarter-pel vectors (every interpolation case)."""
    W, H, R = 352, 288, 16
    fp = synth.rect_motion_pair(W, H, 77 + intra, n_rects=10, max_mv=12)
    est = capi.MotionEstimator(_cfg(W, H, R, metric=1, window=1, filt=0, subpel=2, predictor=0))
    d_cur, d_ref = capi.DeviceBuffer.from_array(fp.cur), capi.DeviceBuffer.from_array(fp.ref)
    n = est.nmb
    d_mv, d_dist = capi.DeviceBuffer(4 * n), capi.DeviceBuffer(4 * n)
    est.estimate_device(d_cur, d_ref, d_mv, d_dist)  # pads the reference the coder reads
    # random quarter-pel vectors within 12 pels (the reference's padded
    # border bounds what the oracle may read), every fraction pair present
    rng = np.random.default_rng(5 + intra)
    mv = rng.integers(-48, 49, size=(n, 2)).astype(np.int16)
    mv[:16] = np.array([(fx + 4, fy - 8) for fx in range(4) for fy in range(4)], np.int16)
    assert len({(int(x) % 4, int(y) % 4) for x, y in mv}) == 16
    capi.lib.icu_stream_sync(None)
    d_mv = capi.DeviceBuffer.from_array(mv)
    qps = [22, 25, 28, 31, 34, 37, 40, 51]
    d_lv = [capi.DeviceBuffer(2 * W * H) for _ in qps]
    d_rc = [capi.DeviceBuffer(W * H) for _ in qps]
    d_pr = capi.DeviceBuffer(2 * W * H)
    est.code_frame_qps_device(d_cur, d_mv, qps, intra, 0, d_lv, d_rc, d_pr)
    capi.lib.icu_stream_sync(None)
    pred = orc.mc_luma(fp.ref, mv)
    np.testing.assert_array_equal(d_pr.download(np.int16, (H, W)), pred)
    for i, qp in enumerate(qps):
        lv, rc = orc.code_residual(fp.cur, pred, qp, intra, 0)
        np.testing.assert_array_equal(d_lv[i].download(np.int16, (H, W)), lv)
        np.testing.assert_array_equal(d_rc[i].download(np.uint8, (H, W)).astype(np.int16), rc)


# ------------------------------------------------------------- batching
def _mixed_streams(seed=0):
    rng = np.random.default_rng(seed)
    specs = [((176, 144), 8), ((176, 144), 16), ((352, 288), 16), ((352, 288), 32),
             ((1280, 720), 32), ((176, 144), 64)]
    out = []
    for i, ((w, h), R) in enumerate(specs):
        fp = synth.rect_motion_pair(w, h, 500 + i, n_rects=int(rng.integers(8, 25)), max_mv=min(R, 32))
        out.append((w, h, R, fp))
    return out


def test_batch_alone_equals_in_mix_and_graph_replays():
    streams = _mixed_streams()
    modes = dict(metric=1, window=1, filt=0, subpel=1, predictor=0, best_vector=1)
    cfgs = [_cfg(w, h, R, **modes) for (w, h, R, _) in streams]
    batch = capi.MotionEstimatorBatch(cfgs)
    curs = [capi.DeviceBuffer.from_array(fp.cur) for (*_, fp) in streams]
    refs = [capi.DeviceBuffer.from_array(fp.ref) for (*_, fp) in streams]
    T = batch.total_mbs
    d_mv, d_dist = capi.DeviceBuffer(4 * T), capi.DeviceBuffer(4 * T)
    batch.estimate_device(curs, refs, d_mv, d_dist)
    capi.lib.icu_stream_sync(None)
    capi.DeviceBuffer.sync()
    s = capi.Stream()
    # eager result through the batch's own stream
    batch.estimate_device(curs, refs, d_mv, d_dist, stream=s)
    s.sync()
    mv_b = d_mv.download(np.int16, (T, 2))
    dist_b = d_dist.download(np.int32, T)
    for i, (w, h, R, fp) in enumerate(streams):
        est = capi.MotionEstimator(cfgs[i])
        mv, dist, _ = est.estimate(fp.cur.astype(np.int16), fp.ref.astype(np.int16))
        b0 = batch.bases[i]
        np.testing.assert_array_equal(mv_b[b0:b0 + est.nmb], mv, err_msg=f"stream {i}")
        np.testing.assert_array_equal(dist_b[b0:b0 + est.nmb], dist)
        if w <= 352:
            omv, _, _, _ = orc.me_estimate(fp.cur, fp.ref, R, **_oracle_modes(modes))
            np.testing.assert_array_equal(mv, omv)
    # graph: capture once, replay several times with eager work in between
    g_mv, g_dist = capi.DeviceBuffer(4 * T), capi.DeviceBuffer(4 * T)
    batch.capture(curs, refs, g_mv, g_dist)
    for rep in range(3):
        g_mv.zero()
        batch.graph_launch(s)
        s.sync()
        np.testing.assert_array_equal(g_mv.download(np.int16, (T, 2)), mv_b, err_msg=f"replay {rep}")
        np.testing.assert_array_equal(g_dist.download(np.int32, T), dist_b)
        batch.estimate_device(curs, refs, d_mv, d_dist, stream=s)
        s.sync()


def test_batch_plans_agree_and_tuner_keeps_one():
    """Every launch plan of a batch (workgroup tiles, one launch per window
    class on one stream or two, the persistent queue kernel for every class
    or for the large ones) gives the same vectors and
    distortions, over streams of all four window classes under both window
    rules; the queue kernel's counters reset between back-to-back launches;
    the tuner keeps a valid plan and the graph captured after it agrees."""
    streams = _mixed_streams(3)
    modes = dict(metric=1, filt=0, subpel=1, predictor=0, best_vector=1)
    cfgs = [_cfg(w, h, R, window=i % 2, **modes) for i, (w, h, R, _) in enumerate(streams)]
    curs = [capi.DeviceBuffer.from_array(fp.cur) for (*_, fp) in streams]
    refs = [capi.DeviceBuffer.from_array(fp.ref) for (*_, fp) in streams]
    s = capi.Stream()
    want = None
    for gm in (0, 4, 5, 6, 7, 3, 2):
        for c in cfgs:
            c.grid_mode = gm
        batch = capi.MotionEstimatorBatch(cfgs)
        assert batch.plan()[0] == {0: 5, 4: 3, 5: 4, 6: 5, 7: 6, 3: 2, 2: 1}[gm]
        T = batch.total_mbs
        d_mv, d_dist = capi.DeviceBuffer(4 * T), capi.DeviceBuffer(4 * T)
        for rep in range(2):
            d_mv.zero()
            batch.estimate_device(curs, refs, d_mv, d_dist, stream=s)
            s.sync()
            got = (d_mv.download(np.int16, (T, 2)), d_dist.download(np.int32, T))
            if want is None:
                want = got
                for i, (w, h, R, fp) in enumerate(streams):
                    if w > 352:
                        continue
                    b0, n = batch.bases[i], (w // 16) * (h // 16)
                    omv, odist, _, _ = orc.me_estimate(fp.cur, fp.ref, R, **_oracle_modes(dict(modes, window=i % 2)))
                    np.testing.assert_array_equal(want[0][b0:b0 + n], omv, err_msg=f"stream {i}")
                    np.testing.assert_array_equal(want[1][b0:b0 + n], odist, err_msg=f"stream {i}")
            np.testing.assert_array_equal(got[0], want[0], err_msg=f"grid {gm} run {rep}")
            np.testing.assert_array_equal(got[1], want[1], err_msg=f"grid {gm} run {rep}")
        if gm == 0:
            ms = batch.autotune(curs, refs, d_mv, d_dist, reps=2)
            g, w = batch.plan()
            assert ms > 0 and 0 <= g <= 6 and w in (6, 8)
            g_mv, g_dist = capi.DeviceBuffer(4 * T), capi.DeviceBuffer(4 * T)
            batch.capture(curs, refs, g_mv, g_dist)
            for rep in range(2):
                g_mv.zero()
                batch.graph_launch(s)
                s.sync()
                np.testing.assert_array_equal(g_mv.download(np.int16, (T, 2)), want[0], err_msg=f"tuned replay {rep}")
                np.testing.assert_array_equal(g_dist.download(np.int32, T), want[1])
        batch.close()
