"""Benchmark: 1080p macroblocks/s for +-32 full search + half-pel + 4x4
IT/quant, plain per-frame launch vs batched, with p50/p99 step latency.

A step is one pass of the hot path over one frame pair already resident in
HBM: reference padding, integer full search (SAD, +-32 exhaustive window,
zero predictor), decision + half-pel refinement, then
luma motion compensation and residual coding (forward IT+Q, dequant +
inverse IT, clipped reconstruction) at QP 22, 28 and 34.

Prints one JSON line.  The batched leg replays a captured HIP graph over a
mixed batch (24 QCIF + 20 CIF + 12 720p + 8 1080p, ranges 8..64).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

W, H, R = 1920, 1088, 32
QPS = (22, 28, 34)

# HBM traffic of one search launch at this workload, from two rocprofv3
# --pmc runs of `bench.py --core-only --no-cpu --no-batched --steps 20
# --warmup 5 --copy-probe 256` (KiB per dispatch of me_wave_kernel<32, true>),
# each counter scaled by what it reads for the 256 MiB device copy / fill of
# --copy-probe: FETCH_SIZE reads half the copy's 262144 KiB (x2.0),
# WRITE_SIZE reads it exactly (x1.0).  Measured counters, not computed here:
# refresh them when the kernel changes.
SEARCH_FETCH_KIB, FETCH_SCALE = 3995.3, 2.0
SEARCH_WRITE_KIB, WRITE_SCALE = 287.0, 1.0
MODES = dict(metric=1, window=1, filt=0, subpel=1, predictor=0, best_vector=0)


def sad_ops(W, H, L):
    """Candidates inside each MB's clamped window x 256 pels."""
    tot = 0
    for m in range(0, H, 16):
        yu = -L if m - L >= -16 else -(m + 16)
        yd = L if m + L < H else H - m
        for n in range(0, W, 16):
            xl = -L if n - L >= -16 else -(n + 16)
            xr = L if n + L < W else W - n
            tot += (xr - xl + 1) * (yd - yu + 1)
    return tot * 256


def pctl(xs, p):
    return float(np.percentile(np.asarray(xs), p))


def run_plain(capi, synth, steps, warmup, tune=False, settle_ms=100.0):
    fp = synth.rect_motion_pair(W, H, 1080, n_rects=24, max_mv=32, sigma_cur=3.0)
    est = capi.MotionEstimator(capi.make_config(W, H, R, **MODES))
    s = capi.Stream()
    d_cur, d_ref = capi.DeviceBuffer.from_array(fp.cur), capi.DeviceBuffer.from_array(fp.ref)
    if tune:
        est.autotune(d_cur, d_ref, reps=3)  # the tile tuner: rows/lane x MBs/task x grid mode
    n = est.nmb
    d_mv, d_dist = capi.DeviceBuffer(4 * n), capi.DeviceBuffer(4 * n)
    d_lv = [capi.DeviceBuffer(2 * W * H) for _ in QPS]
    d_rc = [capi.DeviceBuffer(W * H) for _ in QPS]

    def step():
        est.estimate_device(d_cur, d_ref, d_mv, d_dist, stream=s)
        # one prediction per frame, coded and reconstructed at each QP
        est.code_frame_qps_device(d_cur, d_mv, QPS, 0, 0, d_lv, d_rc, stream=s)

    # The GPU's clocks ramp up over the first ~20 ms of work after idle
    # (the search measured 89 us falling to a steady 78 us): run the step
    # untimed for settle_ms first, then the warm-up steps, then the timed ones.
    t_end = time.perf_counter() + settle_ms / 1e3
    while time.perf_counter() < t_end:
        for _ in range(20):
            step()
        s.sync()
    for _ in range(warmup):
        step()
    s.sync()
    # the timed steps: one event pair around all of them, nothing between
    a, b = capi.Event(), capi.Event()
    t0 = time.perf_counter()
    a.record(s)
    for k in range(steps):
        step()
    b.record(s)
    s.sync()
    wall = time.perf_counter() - t0
    total_ms = a.elapsed_ms(b)
    # per-step latencies (p50 / p99) in a second pass, an event after every
    # step (recording one holds the next step back by ~2.5 us, so this pass
    # does not give the throughput)
    ev = [capi.Event() for _ in range(steps + 1)]
    ev[0].record(s)
    for k in range(steps):
        step()
        ev[k + 1].record(s)
    s.sync()
    per = [ev[k].elapsed_ms(ev[k + 1]) for k in range(steps)]
    # per-stage device times of the estimate on the same stream; the search
    # stage is the mean of 20 back-to-back launches between one event pair
    stages = np.array([est.estimate_timed(d_cur, d_ref, d_mv, d_dist, stream=s, search_reps=20)
                       for _ in range(5)])
    return dict(nmb=n, total_ms=total_ms, wall_s=wall, per_ms=per, stages_ms=np.median(stages, 0).tolist(),
                tiling=est.tiling(), est=est, frames=fp)


def mixed_batch_configs(synth, capi):
    streams = synth.mixed_batch(2026)
    return [capi.make_config(w, h, R, **MODES) for (w, h, R, _) in streams], [fp for (*_, fp) in streams]


def run_batched(capi, synth, steps, warmup):
    cfgs, frames = mixed_batch_configs(synth, capi)
    batch = capi.MotionEstimatorBatch(cfgs)
    curs = [capi.DeviceBuffer.from_array(fp.cur) for fp in frames]
    refs = [capi.DeviceBuffer.from_array(fp.ref) for fp in frames]
    T = batch.total_mbs
    d_mv, d_dist = capi.DeviceBuffer(4 * T), capi.DeviceBuffer(4 * T)
    s = capi.Stream()
    # the tuner picks the launch plan (workgroup tiles / a launch per window
    # class / the persistent queue kernel and its waves per SIMD), then the
    # graph is captured with it
    default_plan = batch.plan()
    tuned_ms = batch.autotune(curs, refs, d_mv, d_dist, reps=10)
    plan = batch.plan()
    batch.capture(curs, refs, d_mv, d_dist)
    for _ in range(warmup):
        batch.graph_launch(s)
    s.sync()
    a, b = capi.Event(), capi.Event()
    a.record(s)
    for _ in range(steps):
        batch.graph_launch(s)
    b.record(s)
    s.sync()
    ms = a.elapsed_ms(b) / steps
    # the same streams one plain estimate launch sequence per stream
    ests = [capi.MotionEstimator(c) for c in cfgs]
    mvs = [capi.DeviceBuffer(4 * e.nmb) for e in ests]
    dss = [capi.DeviceBuffer(4 * e.nmb) for e in ests]
    for e, c, r, m, d in zip(ests, curs, refs, mvs, dss):
        e.estimate_device(c, r, m, d, stream=s)
    s.sync()
    a.record(s)
    for _ in range(max(1, steps // 2)):
        for e, c, r, m, d in zip(ests, curs, refs, mvs, dss):
            e.estimate_device(c, r, m, d, stream=s)
    b.record(s)
    s.sync()
    plain_ms = a.elapsed_ms(b) / max(1, steps // 2)
    # a few SAD-op counts for the roofline of the batch
    ops = sum(sad_ops(c.width, c.height, c.range) for c in cfgs)
    return dict(mbs=T, ms=ms, mbps=T / (ms / 1e3), plain_ms=plain_ms, plain_mbps=T / (plain_ms / 1e3),
                sad_ops=ops, streams=len(cfgs), plan=plan, default_plan=default_plan, tuned_ms=tuned_ms)


VBS_W, VBS_H, VBS_R, VBS_REFS, VBS_LAMBDA = 3840, 2160, 16, 4, 4
VBS_QPS = tuple(range(22, 41))


def vbs_sad_ops(W, H, R, nrefs):
    """Candidates in each MB's clamped +-R window x references x 256 pels."""
    tot = 0
    for m in range(0, H, 16):
        yu = -R if m - R >= -16 else -(m + 16)
        yd = R if m + R < H else H - m
        for n in range(0, W, 16):
            xl = -R if n - R >= -16 else -(n + 16)
            xr = R if n + R < W else W - n
            tot += (xr - xl + 1) * (yd - yu + 1)
    return tot * nrefs * 256


def run_multiref_4k(capi, synth, steps, warmup):
    """4K, 4 references (shifted, noised copies of one texture), +-16
    exhaustive search with 16x16/16x8/8x16/8x8 partitions from 4x4 SADs and
    the rate term 4 (|vx| + |vy|), then the residual coded at QP 22..40."""
    W, H = VBS_W, VBS_H
    cur, refs, _ = synth.shifted_copies(W, H, 4096, VBS_REFS, max_mv=16)
    est = capi.VbsEstimator(W, H, VBS_R, VBS_REFS, VBS_LAMBDA)
    d_cur = capi.DeviceBuffer.from_array(cur)
    d_refs = [capi.DeviceBuffer.from_array(r) for r in refs]
    d_mode, d_cost = capi.DeviceBuffer(est.nmb), capi.DeviceBuffer(4 * est.nmb)
    d_mv8, d_ref8 = capi.DeviceBuffer(4 * est.nb8), capi.DeviceBuffer(est.nb8)
    d_lv = [capi.DeviceBuffer(2 * W * H) for _ in VBS_QPS]
    d_rc = [capi.DeviceBuffer(W * H) for _ in VBS_QPS]
    s = capi.Stream()

    def step():
        est.estimate_device(d_cur, d_refs, d_mode, d_mv8, d_ref8, d_cost, s)
        est.code_device(d_cur, d_mv8, d_ref8, list(VBS_QPS), 0, 0, d_lv, d_rc, None, s)

    for _ in range(warmup):
        step()
    s.sync()
    ev = [capi.Event() for _ in range(3)]
    ev[0].record(s)
    for _ in range(steps):
        step()
    ev[1].record(s)
    s.sync()
    ms = ev[0].elapsed_ms(ev[1]) / steps
    # the two stages apart: back-to-back launches of each
    ev[0].record(s)
    for _ in range(steps):
        est.estimate_device(d_cur, d_refs, d_mode, d_mv8, d_ref8, d_cost, s)
    ev[1].record(s)
    for _ in range(steps):
        est.code_device(d_cur, d_mv8, d_ref8, list(VBS_QPS), 0, 0, d_lv, d_rc, None, s)
    ev[2].record(s)
    s.sync()
    est_ms = ev[0].elapsed_ms(ev[1]) / steps
    code_ms = ev[1].elapsed_ms(ev[2]) / steps
    mode = d_mode.download(np.uint8, est.nmb)
    ops = vbs_sad_ops(W, H, VBS_R, VBS_REFS)
    # coder traffic: current + prediction source read once, levels (2 B) and
    # reconstruction (1 B) written per pel per QP, vectors and indices
    code_bytes = W * H * (2 + 3 * len(VBS_QPS)) + est.nb8 * 5
    return {"workload": f"4K {W}x{H}, {VBS_REFS} references (shifted copies, sigma 2..6), "
                        f"+-{VBS_R} exhaustive, 16x16/16x8/8x16/8x8 from 4x4 SADs, rate {VBS_LAMBDA}(|vx|+|vy|), "
                        f"residual at QP 22..40 ({len(VBS_QPS)} QPs)",
            "mbs_per_step": est.nmb, "ms_per_step": round(ms, 4), "value": round(est.nmb / (ms / 1e3), 1),
            "unit": "MB/s", "estimate_ms": round(est_ms, 4), "code_ms": round(code_ms, 4),
            "search_T_sad_ops": round(ops / (est_ms / 1e3) / 1e12, 2), "sad_ops_per_frame": ops,
            "code_GBps": round(code_bytes / (code_ms / 1e3) / 1e9, 1), "code_bytes": code_bytes,
            "tile_rows": est.plan()[0],
            "modes": {k: int((mode == i).sum()) for i, k in enumerate(("16x16", "16x8", "8x16", "8x8"))}}


def run_mc_yuv(capi, synth, steps, warmup):
    """Whole-frame YUV 4:2:0 compensation of a 1080p frame (the padding of
    the three reference planes included) at the vectors the 1080p search
    finds, out of place."""
    fp = synth.rect_motion_pair(W, H, 1080, n_rects=24, max_mv=32, sigma_cur=3.0)
    rng = np.random.default_rng(7)
    u = rng.integers(0, 256, (H // 2, W // 2)).astype(np.uint8)
    v = rng.integers(0, 256, (H // 2, W // 2)).astype(np.uint8)
    est = capi.MotionEstimator(capi.make_config(W, H, R, **MODES))
    d_cur, d_ref = capi.DeviceBuffer.from_array(fp.cur), capi.DeviceBuffer.from_array(fp.ref)
    n = (W // 16) * (H // 16)
    d_mv, d_dist = capi.DeviceBuffer(4 * n), capi.DeviceBuffer(4 * n)
    s = capi.Stream()
    est.estimate_device(d_cur, d_ref, d_mv, d_dist, stream=s)
    mc = capi.MotionCompensator(W, H)
    d_in = [d_ref, capi.DeviceBuffer.from_array(u), capi.DeviceBuffer.from_array(v)]
    d_out = [capi.DeviceBuffer(W * H), capi.DeviceBuffer(W * H // 4), capi.DeviceBuffer(W * H // 4)]
    for _ in range(warmup):
        mc.compensate_device(d_in, d_mv, d_out, s)
    s.sync()
    a, b = capi.Event(), capi.Event()
    a.record(s)
    for _ in range(steps):
        mc.compensate_device(d_in, d_mv, d_out, s)
    b.record(s)
    s.sync()
    ms = a.elapsed_ms(b) / steps
    nbytes = 2 * (W * H * 3 // 2)  # three planes read once and written once
    return {"workload": "1080p YUV 4:2:0 compensation (luma quarter-pel, chroma eighth-pel) at the search's "
                        "vectors, reference padding included",
            "ms": round(ms, 4), "mbs_per_s": round(n / (ms / 1e3), 1),
            "GBps_algorithmic": round(nbytes / (ms / 1e3) / 1e9, 1), "bytes_algorithmic": nbytes}


def run_cavlc(capi, synth, steps, warmup):
    """CAVLC of one 1080p frame's luma levels: the 130560 4x4 blocks that
    the search + MC + IT/quant at QP 28 produce, each with its nC from its
    left and upper neighbours' coefficient counts (H.264 9.2.1), encoded into
    one stream (count, scan, one length read back, write), then decoded
    from the offsets."""
    fp = synth.rect_motion_pair(W, H, 1080, n_rects=24, max_mv=32, sigma_cur=3.0)
    est = capi.MotionEstimator(capi.make_config(W, H, R, **MODES))
    d_cur, d_ref = capi.DeviceBuffer.from_array(fp.cur), capi.DeviceBuffer.from_array(fp.ref)
    n_mb = (W // 16) * (H // 16)
    d_mv, d_dist = capi.DeviceBuffer(4 * n_mb), capi.DeviceBuffer(4 * n_mb)
    s = capi.Stream()
    est.estimate_device(d_cur, d_ref, d_mv, d_dist, stream=s)
    d_lv, d_rec = capi.DeviceBuffer(2 * W * H), capi.DeviceBuffer(W * H)
    est.code_frame_device(d_cur, d_mv, 28, 0, 0, d_lv, d_rec, stream=s)
    s.sync()
    # picture layout -> raster 4x4 blocks, block-major (input preparation)
    lv = d_lv.download(np.int16, (H, W))
    blocks = np.ascontiguousarray(lv.reshape(H // 4, 4, W // 4, 4).transpose(0, 2, 1, 3).reshape(-1, 16))
    n = len(blocks)
    tc = np.count_nonzero(blocks, axis=1).reshape(H // 4, W // 4)
    na = np.full_like(tc, -1)
    nb = np.full_like(tc, -1)
    na[:, 1:], nb[1:, :] = tc[:, :-1], tc[:-1, :]
    nc = np.where((na >= 0) & (nb >= 0), (na + nb + 1) >> 1, np.maximum(np.maximum(na, nb), 0))
    d_coef = capi.DeviceBuffer.from_array(blocks)
    d_nc = capi.DeviceBuffer.from_array(nc.reshape(-1).astype(np.int8))
    cap = n * 160
    d_stream, d_off = capi.DeviceBuffer(cap), capi.DeviceBuffer(8 * (n + 1))
    d_dec, d_status = capi.DeviceBuffer(2 * 16 * n), capi.DeviceBuffer(4 * n)
    dev = capi.CavlcDevice()
    for _ in range(warmup):
        bits = dev.encode(d_coef, n, 1, d_nc, 0, 0, d_stream, cap, d_off, stream=s)
        dev.decode(d_stream, cap, d_off, n, 1, d_nc, 0, 0, d_dec, d_status, stream=s)
    s.sync()
    a, b, c = capi.Event(), capi.Event(), capi.Event()
    a.record(s)
    for _ in range(steps):
        dev.encode(d_coef, n, 1, d_nc, 0, 0, d_stream, cap, d_off, stream=s)
    b.record(s)
    for _ in range(steps):
        dev.decode(d_stream, cap, d_off, n, 1, d_nc, 0, 0, d_dec, d_status, stream=s)
    c.record(s)
    s.sync()
    enc_ms, dec_ms = a.elapsed_ms(b) / steps, b.elapsed_ms(c) / steps
    ok = bool((d_dec.download(np.int16, (n, 16)) == blocks).all())
    return {"workload": "CAVLC of a 1080p frame's luma levels at QP 28 (130560 4x4 blocks, nC from neighbours)",
            "encode_ms": round(enc_ms, 4), "decode_ms": round(dec_ms, 4),
            "encode_mbs_per_s": round(n_mb / (enc_ms / 1e3), 1),
            "decode_mbs_per_s": round(n_mb / (dec_ms / 1e3), 1),
            "stream_bits": int(bits), "nonzero_blocks": int(np.count_nonzero(tc)), "round_trip_exact": ok}


def run_as_shipped(capi, synth, steps, warmup):
    """The full-search estimator in the reference class's shipped
    configuration (SSD, running-best filter, +-31 window, H.264 median
    predictor, half + quarter pel, GetBestVector; bit-exact with the class),
    on the 1080p frame pair of the plain leg: device time per estimate, the
    same workload as the CPU baseline's reference-class timing."""
    fp = synth.rect_motion_pair(W, H, 1080, n_rects=24, max_mv=32, sigma_cur=3.0)
    est = capi.MotionEstimator(capi.reference_config(W, H, 4 * R))
    d_cur, d_ref = capi.DeviceBuffer.from_array(fp.cur), capi.DeviceBuffer.from_array(fp.ref)
    n = est.nmb
    d_mv, d_dist = capi.DeviceBuffer(4 * n), capi.DeviceBuffer(4 * n)
    s = capi.Stream()
    for _ in range(warmup):
        est.estimate_device(d_cur, d_ref, d_mv, d_dist, stream=s)
    s.sync()
    a, b = capi.Event(), capi.Event()
    a.record(s)
    for _ in range(steps):
        est.estimate_device(d_cur, d_ref, d_mv, d_dist, stream=s)
    b.record(s)
    s.sync()
    ms = a.elapsed_ms(b) / steps
    return {"workload": "1080p, reference class as shipped (SSD, filter, +-31, median predictor, quarter pel, "
                        "GetBestVector)", "ms_per_frame": round(ms, 4), "mbs_per_s": round(n / (ms / 1e3), 1),
            "tie_mbs": est.tie_count(), "mbs": n}


FAST_RANGE = 128  # +-32 in the reference's quarter-pel motionRange


def run_fast(capi, synth, steps, warmup):
    """The fast estimators (UMHS, Cross with the median and the zero
    predictor, FHS) on 1080p frames of a moving-rectangle stream, quarter-pel
    mode, inputs resident: device time per frame.  UMHS also with 16 streams
    of 1080p in flight on 4 HIP streams (each stream its own handle), for the
    throughput the wavefront leaves room for."""
    src = synth.StreamSource(W, H, 4040, n_rects=24, max_mv=16, sigma=2.0)
    f0, f1 = src.frame(0), src.frame(1)
    n_mb = (W // 16) * (H // 16)
    d_cur, d_ref = capi.DeviceBuffer.from_array(f1), capi.DeviceBuffer.from_array(f0)
    d_mv, d_dist = capi.DeviceBuffer(4 * n_mb), capi.DeviceBuffer(4 * n_mb)
    s = capi.Stream()
    out = {"workload": "1080p 1920x1088, +-32 (motionRange 128), quarter pel, moving rectangles"}
    kinds = [("umhs", lambda: capi.UmhsEstimator(W, H, FAST_RANGE, 0)),
             ("cross_median", lambda: capi.CrossEstimator(W, H, FAST_RANGE, 0, 1)),
             ("cross_zero", lambda: capi.CrossEstimator(W, H, FAST_RANGE, 0, 0)),
             ("fhs", lambda: capi.FhsEstimator(W, H, FAST_RANGE, 0))]
    for name, make in kinds:
        est = make()
        for _ in range(warmup):
            est.estimate_device(d_cur, d_ref, d_mv, d_dist, stream=s)
        s.sync()
        a, b = capi.Event(), capi.Event()
        a.record(s)
        for _ in range(steps):
            est.estimate_device(d_cur, d_ref, d_mv, d_dist, stream=s)
        b.record(s)
        s.sync()
        ms = a.elapsed_ms(b) / steps
        out[name] = {"ms_per_frame": round(ms, 4), "mbs_per_s": round(n_mb / (ms / 1e3), 1)}
        est.close()
    # 16 streams in flight
    nstream, lanes = 16, 4
    ests = [capi.UmhsEstimator(W, H, FAST_RANGE, 0) for _ in range(nstream)]
    mvs = [capi.DeviceBuffer(4 * n_mb) for _ in range(nstream)]
    dss = [capi.DeviceBuffer(4 * n_mb) for _ in range(nstream)]
    sts = [capi.Stream() for _ in range(lanes)]
    for e, mv, ds, k in zip(ests, mvs, dss, range(nstream)):
        e.estimate_device(d_cur, d_ref, mv, ds, stream=sts[k % lanes])
    for st in sts:
        st.sync()
    reps = max(2, steps // 4)
    t0 = time.perf_counter()
    for _ in range(reps):
        for e, mv, ds, k in zip(ests, mvs, dss, range(nstream)):
            e.estimate_device(d_cur, d_ref, mv, ds, stream=sts[k % lanes])
    for st in sts:
        st.sync()
    wall = time.perf_counter() - t0
    out["umhs_16_streams"] = {"frames": nstream * reps, "mbs_per_s": round(nstream * reps * n_mb / wall, 1),
                              "ms_per_frame_aggregate": round(wall * 1e3 / (nstream * reps), 4)}
    return out


SERVE_STREAMS, SERVE_RATE, SERVE_QP, SERVE_BUDGET = 32, 2000.0, 28, 64 << 30
SERVE_RING = 3  # distinct frames generated per stream, submitted in turn


def run_serving(capi, synth, seconds, lanes=4, max_inflight=32):
    """Serving: 32 streams (CIF:720p:1080p drawn 4:2:1), frames arriving as
    a Poisson process at 2000 frames/s in aggregate, each coded through the
    server (+-32 full search + half-pel + MC + IT/quant at QP 28) with its
    stream's reference resident on the device, under a 64 GB budget.
    Latency of a frame: from its scheduled arrival to the poll that returns
    it.  Then a saturated pass (every stream always has a frame queued) for
    the server's throughput ceiling."""
    rng = np.random.default_rng(4242)
    names = rng.choice(["CIF", "720p", "1080p"], size=SERVE_STREAMS, p=[4 / 7, 2 / 7, 1 / 7])
    srv = capi.Server(lanes=lanes, max_inflight=max_inflight, mem_budget=SERVE_BUDGET, qp=SERVE_QP)
    ring, nmb, ids = [], [], []
    for k, name in enumerate(names):
        w, h = synth.SIZES[str(name)]
        src = synth.StreamSource(w, h, 7000 + k, n_rects=int(rng.integers(8, 25)), max_mv=8)
        ring.append([src.frame(t) for t in range(SERVE_RING)])
        nmb.append((w // 16) * (h // 16))
        ids.append(srv.open_stream(capi.make_config(w, h, R, **MODES)))
    sent = [0] * SERVE_STREAMS
    call_max = [0.0]

    def submit(k, tag):
        t = capi.Server.clock()
        srv.submit(ids[k], ring[k][sent[k] % SERVE_RING], tag=tag)
        call_max[0] = max(call_max[0], capi.Server.clock() - t)
        sent[k] += 1

    # warm-up: a few frames per stream (first frames, kernels, buffer pools)
    for _ in range(2):
        for k in range(SERVE_STREAMS):
            submit(k, -1)
    done = 0
    while done < 2 * SERVE_STREAMS:
        done += len(srv.poll(wait=True))
    # Poisson arrivals
    call_max[0] = 0.0
    n = max(1, int(SERVE_RATE * seconds))
    arrive = np.cumsum(rng.exponential(1.0 / SERVE_RATE, n))
    who = rng.integers(0, SERVE_STREAMS, n)
    lat = np.zeros(n)
    mbs, got, i = 0, 0, 0
    t0 = capi.Server.clock()
    last_done = t0
    while got < n:
        now = capi.Server.clock() - t0
        while i < n and arrive[i] <= now:
            submit(int(who[i]), i)
            i += 1
        for r in srv.poll(wait=(i >= n)):
            lat[r["tag"]] = r["done_s"] - (t0 + arrive[r["tag"]])
            mbs += nmb[r["stream"]]
            last_done = max(last_done, r["done_s"])
            got += 1
    span = last_done - t0
    st = srv.stats()
    # saturated: keep two frames queued per stream for about a second
    sat_n = 0
    sat_mbs = 0
    for k in range(SERVE_STREAMS):
        submit(k, -1)
        submit(k, -1)
    t1 = capi.Server.clock()
    while capi.Server.clock() - t1 < min(2.0, max(0.5, seconds / 5)):
        for r in srv.poll(wait=True):
            sat_n += 1
            sat_mbs += nmb[r["stream"]]
            submit(r["stream"], -1)
    sat_s = capi.Server.clock() - t1
    srv.close()  # waits for the frames in flight, drops the queued ones
    mix = {str(x): int((names == x).sum()) for x in ("CIF", "720p", "1080p")}
    return {"workload": f"{SERVE_STREAMS} streams {mix}, Poisson {SERVE_RATE:.0f} frames/s for {seconds:g} s, "
                        f"+-{R} full search + half-pel + MC + IT/quant at QP {SERVE_QP}, "
                        f"{SERVE_BUDGET >> 30} GB budget, {lanes} lanes, up to {max_inflight} frames in flight",
            "frames": n, "mbs_per_s": round(mbs / span, 1), "frames_per_s": round(n / span, 1),
            "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 4),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 4),
            "max_ms": round(float(lat.max()) * 1e3, 4),
            # the longest a single submit call blocked the request loop (the
            # runtime's pinned upload of a frame has stalled for ms at times)
            "max_submit_call_ms": round(call_max[0] * 1e3, 4),
            "frames_over_1ms": int((lat > 1e-3).sum()),
            "device_bytes_peak": st["peak_bytes"],
            "saturated_mbs_per_s": round(sat_mbs / sat_s, 1),
            "saturated_frames_per_s": round(sat_n / sat_s, 1)}


def cpu_baseline(synth, budget_s=12.0):
    """The compiled reference class on the host, one thread: as shipped
    (SSD, running-best filter, +-31, half+quarter pel, median predictor),
    plus its forward/inverse 4x4 transforms over the frame at the three QPs.
    Sample: 1080p frame pairs, as many as fit the time budget."""
    from oracle import oracle as orc
    ref_lib = orc.ref_lib()
    fp = synth.rect_motion_pair(W, H, 1080, n_rects=24, max_mv=32, sigma_cur=3.0)
    cur, ref = fp.cur.astype(np.int16), fp.ref.astype(np.int16)
    nmb = (W // 16) * (H // 16)
    if ref_lib is not None:
        kind = "reference"
        t_me = ref_lib.ref_me_time(W, H, 4 * R, 1, cur, ref, 1)
        reps = max(1, int(budget_s / max(t_me, 1e-3)) - 1)
        if reps > 1:
            t_me = ref_lib.ref_me_time(W, H, 4 * R, 1, cur, ref, reps) / reps
        blocks = (cur - ref).reshape(H // 4, 4, W // 4, 4).transpose(0, 2, 1, 3).reshape(-1, 16).copy()
        t0 = time.perf_counter()
        for qp in QPS:
            b = blocks.copy()
            ref_lib.ref_fwd(0, b.reshape(-1), b.shape[0], qp, 0, 0)
            ref_lib.ref_inv(0, b.reshape(-1), b.shape[0], qp, 0, None)
        t_it = time.perf_counter() - t0
        sample = f"{reps} x 1080p frame pair, reference class as shipped (SSD, filter, +-31, qpel) + 3 QP IT/quant"
    else:
        kind = "port"
        t0 = time.perf_counter()
        orc.me_estimate(cur, ref, R, step=8, metric=1, window=1, filter=0, subpel=1, pred=0, bestvec=1)
        t_me = (time.perf_counter() - t0) * 8
        t_it = 0.0
        sample = "1/8 of the MBs of a 1080p frame pair with the oracle in the benchmark mode"
    per_frame = t_me + t_it
    out = {"value": nmb / per_frame, "unit": "MB/s", "cores": 1, "kind": kind, "sample": sample,
           "s_per_frame": per_frame}
    # the fast estimators' reference classes on the same 1080p stream as
    # the fast-estimator leg, one Estimate each
    if ref_lib is not None:
        src = synth.StreamSource(W, H, 4040, n_rects=24, max_mv=16, sigma=2.0)
        c1, c0 = src.frame(1), src.frame(0)
        _, _, _, t_u = orc.ref_umhs_estimate(ref_lib, c1, c0, 128, 0)
        _, _, t_c = orc.ref_cross_estimate(ref_lib, c1, c0, 128, 0, 1)
        _, _, _, t_f = orc.ref_fhs_chain(ref_lib, [c1], [c0], 128, 0)
        out["fast_estimators_1080p"] = {"kind": "reference", "cores": 1, "unit": "ms per frame",
                                        "umhs": round(t_u * 1e3, 2), "cross_median": round(t_c * 1e3, 2),
                                        "fhs": round(t_f * 1e3, 2)}
    model, nproc, avail = _cpu_info()
    out.update(cpu_model=model, nproc=nproc, cpus_available=avail)
    # the same mode as the GPU (SAD, exhaustive +-32, filter off, zero
    # predictor, half-pel) through the oracle on 1/8 of the MBs: the reference
    # class cannot be built in that mode
    t0 = time.perf_counter()
    orc.me_estimate(cur, ref, R, step=8, metric=1, window=1, filter=0, subpel=1, pred=0, bestvec=0)
    t_same = (time.perf_counter() - t0) * 8
    out["same_mode_as_gpu"] = {"value": round(nmb / t_same, 1), "unit": "MB/s", "cores": 1, "kind": "port",
                               "sample": "1/8 of the MBs of a 1080p frame pair, oracle in the benchmark's "
                                         "search mode (IT/quant excluded)"}
    if ref_lib is not None:
        # one estimator instance per thread over independent frame pairs (the
        # reference has no threading of its own); the calls release the GIL
        import threading
        nthr = max(1, min(16, avail))
        per = 2

        def work():
            ref_lib.ref_me_time(W, H, 4 * R, 1, cur, ref, per)

        ths = [threading.Thread(target=work) for _ in range(nthr)]
        t0 = time.perf_counter()
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        wall = time.perf_counter() - t0
        out["threads"] = {"value": round(nthr * per * nmb / wall, 1), "unit": "MB/s", "cores": nthr,
                          "kind": "reference",
                          "sample": f"{nthr} threads x {per} 1080p frame pairs, reference estimator as shipped "
                                    "(IT/quant excluded)"}
    return out


def _cpu_info():
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count(), len(os.sched_getaffinity(0))


def note(msg):
    """One progress line per finished leg, on stderr (stdout carries only the JSON line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--tune", action="store_true",
                    help="run the tile tuner first (the planner's tiling, 13 rows x 3 MBs per task, one task "
                         "per workgroup, is what it picks for this workload)")
    ap.add_argument("--no-4k", action="store_true", help="skip the 4K multi-reference leg")
    ap.add_argument("--core-only", action="store_true",
                    help="only the 1080p step and the batched leg (no CPU baseline, compensation, CAVLC, "
                         "shipped-configuration, fast-estimator, 4K or serving legs)")
    ap.add_argument("--settle-ms", type=float, default=100.0,
                    help="untimed steps before the warm-up, so the clocks have left their idle state")
    ap.add_argument("--serve-seconds", type=float, default=60.0,
                    help="length of the serving leg's Poisson arrival run (0: skip the leg)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first run one device-to-device copy of MIB MiB: a kernel of known traffic "
                         "that calibrates hardware byte counters read in the same profile")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path shards by "
              "stream, one process per GPU, with no exchange).", file=sys.stderr)
        return 2
    from imagecodingutils_amd import capi, synth

    info = capi.device_info()
    if args.copy_probe > 0:
        nb = args.copy_probe << 20
        a, b = capi.DeviceBuffer(nb), capi.DeviceBuffer(nb)
        capi.lib.icu_memset(a.ptr, 1, nb, None)
        capi.lib.icu_memcpy_dtod(b.ptr, a.ptr, nb, None)
        capi.lib.icu_stream_sync(None)
        del a, b
    plain = run_plain(capi, synth, args.steps, args.warmup, tune=args.tune, settle_ms=args.settle_ms)
    note(f"plain 1080p leg: {plain['total_ms'] / args.steps:.4f} ms per step")
    ms_per_step = plain["total_ms"] / args.steps
    value = plain["nmb"] / (ms_per_step / 1e3)
    ops = sad_ops(W, H, R)
    search_ms = plain["stages_ms"][1]
    achieved = ops / (search_ms / 1e3) / 1e12
    clk_ghz = info["clock_khz"] / 1e6
    cus = info["compute_units"]
    peak = cus * 64 * clk_ghz * 1e9 * 4 / 1e12  # v_sad_u8: 4 SAD-ops per lane-op
    out = {
        "metric": "1080p macroblocks/sec (±32 full search + IT/quant), batched vs plain; p99 ms",
        "value": round(value, 1),
        "unit": "MB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (seeded sinusoid texture, 24 moving rectangles, sigma 3)",
        "config": {"workload": "1080p 1920x1088 single stream, SAD full search +-32 (exhaustive), "
                               "half-pel, MC + 4x4 IT/quant/dequant/inverse at QP 22/28/34",
                   "mbs_per_step": plain["nmb"], "tiling": list(plain["tiling"])},
        "p50_ms": round(pctl(plain["per_ms"], 50), 4),
        "p99_ms": round(pctl(plain["per_ms"], 99), 4),
        "stage_ms": {"pad": plain["stages_ms"][0], "search": search_ms, "decision_kernels": plain["stages_ms"][2]},
        "roofline": {"bound": "valu", "achieved": round(achieved, 2), "peak": round(peak, 2),
                     "unit": "T SAD-op/s", "frac": round(achieved / peak, 4),
                     "traffic": round((SEARCH_FETCH_KIB * FETCH_SCALE + SEARCH_WRITE_KIB * WRITE_SCALE) * 1024),
                     # padded reference + current plane read once, 8 B of results per MB
                     "traffic_algorithmic": (W + 160) * (H + 160) + W * H + W * H // 256 * 8,
                     "kernel": "me_wave_kernel<32, lean> (one MB per wave: integer search + in-kernel half-pel "
                               "decision)",
                     "sad_ops_per_launch": ops, "search_ms": round(search_ms, 5)},
        "device": info,
    }
    if not args.no_batched:
        bt = run_batched(capi, synth, max(4, args.steps // 4), 2)
        out["batched"] = {"workload": "64 streams: 24 QCIF + 20 CIF + 12 720p + 8 1080p, R in {8,16,32,64}, "
                                      "one hipGraph replay",
                          "mbs_per_launch": bt["mbs"], "ms_per_launch": round(bt["ms"], 4),
                          "value": round(bt["mbps"], 1), "plain_same_streams_mbps": round(bt["plain_mbps"], 1),
                          "batched_vs_plain": round(bt["mbps"] / bt["plain_mbps"], 3),
                          "achieved_T_sad_ops": round(bt["sad_ops"] / (bt["ms"] / 1e3) / 1e12, 2),
                          "plan": {"grid_mode": bt["plan"][0], "queue_waves_per_simd": bt["plan"][1],
                                   "tuned_ms": round(bt["tuned_ms"], 4),
                                   "untuned_grid_mode": bt["default_plan"][0],
                                   "modes": "0-2 workgroup tiles, 3 a launch per window class, "
                                            "4 one persistent launch over per-XCD MB queues, 5 as 3 with "
                                            "the +-8/+-16 classes on a second stream, 6 the +-32/+-64 "
                                            "classes as in 4 and the others as in 5"}}
        note("batched leg done")
    if args.core_only:
        print(json.dumps(out), flush=True)
        return 0
    out["mc_yuv_1080p"] = run_mc_yuv(capi, synth, 20, 3)
    out["cavlc_1080p"] = run_cavlc(capi, synth, 20, 3)
    note("YUV compensation and CAVLC legs done")
    out["as_shipped_1080p"] = run_as_shipped(capi, synth, 10, 2)
    note("as-shipped leg done")
    out["fast_estimators_1080p"] = run_fast(capi, synth, 10, 2)
    note("fast estimator legs done")
    if not args.no_4k:
        out["multiref_4k"] = run_multiref_4k(capi, synth, max(5, args.steps // 5), 2)
        note("4K multi-reference leg done")
    if args.serve_seconds > 0:
        out["serving"] = run_serving(capi, synth, args.serve_seconds)
        note("serving leg done")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(synth)
        note("CPU baseline done")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
