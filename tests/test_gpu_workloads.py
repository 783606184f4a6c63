"""The batched and serving workloads at their stated sizes.

Batched: the 64-stream mixed batch the benchmark replays (24 QCIF + 20 CIF
+ 12 720p + 8 1080p, ranges 8..64, seed 2026), captured in a HIP graph and
replayed; every stream's whole-frame result equals its own plain estimate,
and the oracle on every MB of the QCIF / CIF streams and every 8th MB of the
720p / 1080p streams.

Serving: streams drawn CIF:720p:1080p = 4:2:1 (4 + 2 + 1), three frames each
through the server: the same result alone, in the mix and on the plain
per-frame path; the CIF streams' whole chains (vectors, levels,
reconstructions) equal the oracle's, and the 720p / 1080p frames equal it on
every 8th MB's vector and on the whole frame's coding.

Integer work, so every comparison is exact."""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

capi = pytest.importorskip("imagecodingutils_amd.capi")
from imagecodingutils_amd import synth  # noqa: E402

# the benchmark's estimator mode: SAD, exhaustive window, no filter, half-pel,
# zero predictor, the best integer / half-pel vector kept
MODES = dict(metric=1, window=1, filt=0, subpel=1, predictor=0, best_vector=0)
ORC_MODES = dict(metric=1, window=1, filter=0, subpel=1, pred=0, bestvec=0)


def _stride(w):
    return 1 if w <= 352 else 8


def test_batched_64_stream_mix_matches_plain_and_oracle():
    streams = synth.mixed_batch(2026)
    assert len(streams) == 64
    cfgs = [capi.make_config(w, h, R, **MODES) for (w, h, R, _) in streams]
    batch = capi.MotionEstimatorBatch(cfgs)
    curs = [capi.DeviceBuffer.from_array(fp.cur) for (*_, fp) in streams]
    refs = [capi.DeviceBuffer.from_array(fp.ref) for (*_, fp) in streams]
    T = batch.total_mbs
    assert T == 118776
    d_mv, d_dist = capi.DeviceBuffer(4 * T), capi.DeviceBuffer(4 * T)
    s = capi.Stream()
    batch.capture(curs, refs, d_mv, d_dist)
    replays = []
    for rep in range(3):
        d_mv.zero()
        batch.graph_launch(s)
        s.sync()
        replays.append((d_mv.download(np.int16, (T, 2)), d_dist.download(np.int32, T)))
    mv_b, dist_b = replays[-1]
    for rep, (m, d) in enumerate(replays[:-1]):
        np.testing.assert_array_equal(m, mv_b, err_msg=f"replay {rep}")
        np.testing.assert_array_equal(d, dist_b, err_msg=f"replay {rep}")
    for i, (w, h, R, fp) in enumerate(streams):
        print(f"stream {i}: {w}x{h} R {R}", flush=True)
        est = capi.MotionEstimator(cfgs[i])
        n = est.nmb
        d_m, d_d = capi.DeviceBuffer(4 * n), capi.DeviceBuffer(4 * n)
        est.estimate_device(curs[i], refs[i], d_m, d_d, stream=s)
        s.sync()
        mv, dist = d_m.download(np.int16, (n, 2)), d_d.download(np.int32, n)
        est.close()
        b0 = batch.bases[i]
        np.testing.assert_array_equal(mv_b[b0:b0 + n], mv, err_msg=f"stream {i} batch vs plain")
        np.testing.assert_array_equal(dist_b[b0:b0 + n], dist, err_msg=f"stream {i} batch vs plain")
        k = _stride(w)
        omv, odist, _ = orc.me_estimate_par(fp.cur, fp.ref, R, step=k, **ORC_MODES)
        np.testing.assert_array_equal(mv[::k], omv[::k], err_msg=f"stream {i} vs oracle")
        np.testing.assert_array_equal(dist[::k], odist[::k], err_msg=f"stream {i} vs oracle")


# ---------------------------------------------------------------- serving
QP = 28
SERVE_SIZES = ["CIF"] * 4 + ["720p"] * 2 + ["1080p"]
N_FRAMES = 3


def _serve_streams():
    out = []
    for k, name in enumerate(SERVE_SIZES):
        w, h = synth.SIZES[name]
        src = synth.StreamSource(w, h, 7100 + k, n_rects=8 + 2 * k, max_mv=8)
        out.append((w, h, [src.frame(t) for t in range(N_FRAMES)]))
    return out


def _drain(srv, expect, got):
    while len(got) < expect:
        assert srv.stats()["inflight"] + srv.stats()["queued"] > 0, "waiting for frames never submitted"
        for r in srv.poll(wait=True):
            key = (r["stream"], r["frame"])
            assert key not in got
            got[key] = r
    return got


def _serve_alone(streams):
    out = {}
    for i, (w, h, frames) in enumerate(streams):
        srv = capi.Server(lanes=1, max_inflight=1, qp=QP)
        sid = srv.open_stream(capi.make_config(w, h, 32, **MODES))
        for t, f in enumerate(frames):
            srv.submit(sid, f, tag=t)
            (r,) = _drain(srv, 1, {}).values()
            assert r["frame"] == t
            out[(i, t)] = r
        srv.close()
    return out


def _serve_mix(streams):
    srv = capi.Server(lanes=4, max_inflight=32, qp=QP)
    ids = [srv.open_stream(capi.make_config(w, h, 32, **MODES)) for (w, h, _) in streams]
    peak = 0
    for t in range(N_FRAMES):
        for i, sid in enumerate(ids):
            srv.submit(sid, streams[i][2][t], tag=100 * i + t)
            peak = max(peak, srv.stats()["inflight"])
    got = {}
    n = N_FRAMES * len(streams)
    while len(got) < n:
        for r in srv.poll(wait=True):
            got[(ids.index(r["stream"]), r["frame"])] = r
        peak = max(peak, srv.stats()["inflight"])
    srv.close()
    return got, peak


def test_serving_cif_720p_1080p_mix_matches_alone_plain_and_oracle():
    streams = _serve_streams()
    print("serving: alone", flush=True)
    alone = _serve_alone(streams)
    print("serving: mix", flush=True)
    mix, peak = _serve_mix(streams)
    assert peak > 1, "frames of different streams never overlapped"
    assert alone.keys() == mix.keys()
    for k in alone:
        np.testing.assert_array_equal(alone[k]["mv"], mix[k]["mv"], err_msg=f"alone vs mix {k}")
        assert alone[k]["levels_sum"] == mix[k]["levels_sum"], k
        assert alone[k]["recon_sum"] == mix[k]["recon_sum"], k
        assert mix[k]["tag"] == 100 * k[0] + k[1]
    s = capi.Stream()
    for i, (w, h, frames) in enumerate(streams):
        print(f"serving: stream {i} {w}x{h} plain path and oracle", flush=True)
        est = capi.MotionEstimator(capi.make_config(w, h, 32, **MODES))
        n = est.nmb
        d_mv, d_dist = capi.DeviceBuffer(4 * n), capi.DeviceBuffer(4 * n)
        d_lv, d_rc = capi.DeviceBuffer(2 * w * h), capi.DeviceBuffer(w * h)
        d_ref = None
        ref_host = frames[0]  # the first frame is its own reference
        k = _stride(w)
        for t, f in enumerate(frames):
            d_cur = capi.DeviceBuffer.from_array(f)
            est.estimate_device(d_cur, d_cur if d_ref is None else d_ref, d_mv, d_dist, stream=s)
            est.code_frame_device(d_cur, d_mv, QP, 0, 0, d_lv, d_rc, stream=s)
            s.sync()
            mv = d_mv.download(np.int16, (n, 2))
            lv = d_lv.download(np.int16, (h, w))
            rc = d_rc.download(np.uint8, (h, w))
            r = mix[(i, t)]
            np.testing.assert_array_equal(r["mv"], mv, err_msg=f"stream {i} frame {t} server vs plain")
            assert r["levels_sum"] == capi.server_checksum(lv)
            assert r["recon_sum"] == capi.server_checksum(rc)
            omv, _, _ = orc.me_estimate_par(f, ref_host, 32, step=k, **ORC_MODES)
            np.testing.assert_array_equal(mv[::k], omv[::k], err_msg=f"stream {i} frame {t} vs oracle")
            # the whole frame coded from the device's vectors (checked above
            # on every MB of the CIF streams)
            olv, orc_rc = orc.code_residual(f, orc.mc_luma(ref_host, mv), QP, 0, 0)
            np.testing.assert_array_equal(lv, olv, err_msg=f"stream {i} frame {t} levels")
            np.testing.assert_array_equal(rc.astype(np.int16), orc_rc, err_msg=f"stream {i} frame {t} recon")
            ref_host = rc
            d_ref, d_rc = d_rc, capi.DeviceBuffer(w * h)
        est.close()
