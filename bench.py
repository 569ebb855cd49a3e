#!/usr/bin/env python3
"""Headline benchmark of the MI355X matcher.

Default workload (one GPU): 8 probe streams of 384 channels x 1024 templates
(Gaussian footprints, rates log-uniform 0.5-20 Hz, amplitudes N(12, 3),
banded-correlation whitening), 1500-sample chunks framed by 61-sample
margins (W = 1622 columns).  One step = one tick: a chunk of every stream
matched by one graph launch (int16 -> split-BF16 whitening -> proposal network
and gate -> split-BF16 PC conv and projection -> peel loop -> events + PC
features on the host).  The recordings are held in HBM before the timed
region.  Beside the gated split-BF16 group (``value``, one tick at a time) the
same ticks run through the plain FP32 group, BF16 with FP32 rescoring (gated
and not), split-BF16 without the gate and the plain FP32 matcher one stream at
a time, for agreement and speedups; each group mode is also timed with two
ticks in flight on two pipelined lanes (``pipelined``: the next tick's front
stages run beside the current peel).

Prints one JSON line: aggregate realtime factor (seconds of data per wall
second), p50/p99 chunk latency from an int16 hand-off on the host to events
on the host, agreement with the plain FP32 matcher, the peel kernel's
roofline figure and the CPU baseline (the FP32 C restatement of the same
algorithm on the host cores, over a bounded sample of the same chunks).  The
plain FP32 single-stream workload (384 ch x 512 templates, 3000-sample
chunks) follows as secondary keys; ``--workload cfg2|group|bf16`` runs one of
the other workloads alone.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "realtime factor, p99 chunk latency (384ch); BF16/gated vs FP32 @99% spike match"
FS = 30000.0
HBM_PEAK_GBS = 8000.0
FP32_MFMA_PEAK_TFLOPS = 157.3
BF16_MFMA_PEAK_TFLOPS = 2500.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_data(cache: Path, seconds: float):
    from livespikesorter_amd import synth
    cache.mkdir(parents=True, exist_ok=True)
    d = cache / "cfg2_oss_input"
    m = synth.make_model(**synth.CONFIGS["cfg2"])
    if not (d / "ctc.npy").exists() or not (d / "hp_filter.npy").exists():
        synth.write_oss_input(m, d)
    rec_path = cache / f"cfg2_rec_{int(seconds)}s.npy"
    if rec_path.exists():
        x = np.load(rec_path)
    else:
        x, _ = synth.make_recording(m, int(seconds * FS))
        np.save(rec_path, x)
    return m, d, x


def make_group_data(cache: Path, seconds: float):
    """8 probe streams, 384 ch each, template counts 256..1024 (the batching workload)."""
    from livespikesorter_amd import synth
    cache.mkdir(parents=True, exist_ok=True)
    dirs, recs = [], []
    for k, T in enumerate(GROUP_T):
        d = cache / f"grp{k}_T{T}_oss_input"
        m = synth.make_model(C=384, T=T, seed=synth.TEST_SEED + 10 + k, whitening="banded",
                             rate=(0.5, 20.0), amp=(12.0, 3.0))
        if not (d / "ctc.npy").exists():
            synth.write_oss_input(m, d)
        rp = cache / f"grp{k}_T{T}_rec_{int(seconds)}s.npy"
        if rp.exists():
            x = np.load(rp)
        else:
            x, _ = synth.make_recording(m, int(seconds * FS), seed=synth.TEST_SEED + 20 + k)
            np.save(rp, x)
        dirs.append(d)
        recs.append(x)
    return dirs, recs


GROUP_T = (256, 384, 512, 640, 768, 896, 1024, 1024)


def bench_group(args) -> dict:
    """Batching workload: 8 streams in one graph launch per 1500-sample tick."""
    from livespikesorter_amd import OnlineSpikesV2, StreamGroup
    from livespikesorter_amd._hip import DeviceBuffer, synchronize
    N, M = 1500, 61
    W = N + 2 * M
    t0 = time.time()
    dirs, recs = make_group_data(Path(args.cache), args.group_seconds)
    n_ticks = min((len(x) - W) // N + 1 for x in recs)
    log(f"group data ready in {time.time() - t0:.1f}s: {len(dirs)} streams, {n_ticks} ticks")
    bufs = [DeviceBuffer.from_array(x) for x in recs]
    C = recs[0].shape[1]
    synchronize()
    t0 = time.time()
    grp = StreamGroup(dirs, batch=W, peel_split=args.group_split)
    geom = group_geometry(grp.info)
    log(f"group created in {time.time() - t0:.1f}s, peel geometry {geom}")
    rng = np.random.default_rng(20261203)
    jitter = rng.uniform(-5.0, 5.0, size=(args.steps + args.warmup, len(dirs)))

    def tick(j):
        b = j % n_ticks
        return grp.process([bf.ptr + b * N * C * 2 for bf in bufs], [b * N] * len(dirs), device=True,
                           lds=[C] * len(dirs))

    for j in range(args.warmup):
        tick(j)
    wall, lat, spikes = [], [], 0
    synchronize()
    ts = time.perf_counter()
    for j in range(args.warmup, args.warmup + args.steps):
        t1 = time.perf_counter()
        out = tick(j)
        dt = (time.perf_counter() - t1) * 1e3
        wall.append(dt)
        spikes += sum(len(e) for e, _ in out)
        # a tick starts when its last chunk has arrived: each chunk also waits for the slowest
        lat.extend(dt + (jitter[j].max() - jitter[j]))
    synchronize()
    elapsed = time.perf_counter() - ts
    data_s = args.steps * len(dirs) * N / FS
    grp.close()
    # the same streams, one single-stream handle each, run one after another
    singles = [OnlineSpikesV2(d, max_batch=W) for d in dirs]
    for j in range(3):
        for k, sp in enumerate(singles):
            sp.process_batch_device(bufs[k].ptr + (j % n_ticks) * N * C * 2, W, C, first_sample=(j % n_ticks) * N)
    synchronize()
    ts2 = time.perf_counter()
    seq_steps = min(args.steps, 100)
    for j in range(seq_steps):
        b = j % n_ticks
        for k, sp in enumerate(singles):
            sp.process_batch_device(bufs[k].ptr + b * N * C * 2, W, C, first_sample=b * N)
    synchronize()
    seq = time.perf_counter() - ts2
    for sp in singles:
        sp.close()
    grp = StreamGroup(dirs, batch=W, peel_split=args.group_split)
    plat, pbusy, pticks = paced_latency(grp, bufs, n_ticks, N, C, args.group_seconds, seed=20261205,
                                        lanes=args.lanes)
    grp.close()
    for bf in bufs:
        bf.free()
    rtf = data_s / elapsed
    seq_rtf = seq_steps * len(dirs) * N / FS / seq
    log(f"group: {args.steps} ticks in {elapsed:.3f}s: {rtf:.1f}x realtime aggregate, tick p50 "
        f"{pct(wall, 50):.3f} ms p99 {pct(wall, 99):.3f} ms; chunk latency with arrival jitter p50 "
        f"{pct(lat, 50):.3f} p99 {pct(lat, 99):.3f} ms; sequential single handles {seq_rtf:.1f}x; "
        f"paced (launch on arrival) latency p50 {pct(plat, 50):.3f} p99 {pct(plat, 99):.3f} ms")
    return {
        "metric": METRIC + " [batching workload]",
        "value": round(rtf, 3),
        "unit": "x realtime (aggregate over streams)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded generator)",
        "config": {"workload": "8 probe streams x 384 ch, templates 256,384,512,640,768,896,1024,1024, "
                               "1500-sample chunks + 61-sample margins (W=1622), one graph launch per tick, "
                               "arrivals jittered +-5 ms",
                   "streams": len(dirs), "templates": list(GROUP_T), "chunk": N, "batch_columns": W},
        "samples_per_s": round(args.steps * len(dirs) * N / elapsed, 1),
        "tick_p50_ms": round(pct(wall, 50), 4), "tick_p99_ms": round(pct(wall, 99), 4),
        "chunk_latency_p50_ms": round(pct(lat, 50), 4), "chunk_latency_p99_ms": round(pct(lat, 99), 4),
        "paced_latency_p50_ms": round(pct(plat, 50), 4), "paced_latency_p99_ms": round(pct(plat, 99), 4),
        "paced_gpu_busy_fraction": round(pbusy, 4), "paced_launches": pticks,
        "spikes_per_tick": round(spikes / args.steps, 1),
        "sequential_single_handles_x_realtime": round(seq_rtf, 3),
        "speedup_over_sequential": round(rtf / seq_rtf, 3),
        "peel_geometry": geom,
    }


def group_geometry(infos):
    """The group's peel grid: tile rows per thread, subsets per stream, workgroups."""
    return {"rows_per_thread": int(infos[0].peel_rows), "split": [int(x.peel_split) for x in infos],
            "workgroups": int(sum(int(x.peel_groups) for x in infos))}


def pct(a, q):
    return float(np.percentile(np.asarray(a), q)) if len(a) else float("nan")


def paced_latency(grp, bufs, n_ticks, N, C, seconds, seed, lanes=2, host_recs=None):
    """Stream the group's recordings in real time: chunk k of stream s is handed
    over at k * N / FS plus a uniform +-5 ms jitter.  Whenever chunks are
    waiting and a launch lane is free, one launch on that lane matches the
    oldest waiting chunk of every stream that has one and is not in flight
    (the others idle); with two lanes, a chunk arriving during a launch starts
    at once on the other lane.  Latency = events on the host - hand-over.
    With ``host_recs`` (the int16 recordings in host memory) the chunks are
    handed over from the host, so the latency includes their copy to the GPU;
    otherwise they are taken from ``bufs`` in device memory."""
    from concurrent.futures import ThreadPoolExecutor
    S = grp.n
    period = N / FS
    W = grp.batch
    rng = np.random.default_rng(seed)
    n_chunks = min(n_ticks, int(seconds / period))
    arrive = np.arange(n_chunks)[:, None] * period + 0.005 + rng.uniform(-0.005, 0.005, size=(n_chunks, S))
    nxt = [0] * S
    busy_s = [False] * S
    lat, busy, launches = [], 0.0, 0
    free = list(range(lanes))
    inflight = {}
    t0 = time.perf_counter()

    def work(lane, ks, first):
        ts = time.perf_counter()
        if host_recs is not None:
            grp.process([host_recs[s][k * N:k * N + W] if k is not None else None for s, k in enumerate(ks)],
                        first, lane=lane)
        else:
            grp.process([bufs[s].ptr + k * N * C * 2 if k is not None else None for s, k in enumerate(ks)], first,
                        device=True, lds=[C] * S, lane=lane)
        return ts, time.perf_counter()

    with ThreadPoolExecutor(lanes) as pool:
        while min(nxt) < n_chunks or inflight:
            for lane in [l for l, (f, _, _) in inflight.items() if f.done()]:
                f, streams, ks = inflight.pop(lane)
                ts, te = f.result()
                busy += te - ts
                for s, k in zip(streams, ks):
                    lat.append((te - t0 - arrive[k, s]) * 1e3)
                    busy_s[s] = False
                free.append(lane)
            now = time.perf_counter() - t0
            ready = [s for s in range(S) if not busy_s[s] and nxt[s] < n_chunks and arrive[nxt[s], s] <= now]
            if ready and free:
                lane = free.pop(0)
                ks_all = [nxt[s] if s in ready else None for s in range(S)]
                first = [nxt[s] * N for s in range(S)]
                ks = [nxt[s] for s in ready]
                for s in ready:
                    busy_s[s] = True
                    nxt[s] += 1
                inflight[lane] = (pool.submit(work, lane, ks_all, first), ready, ks)
                launches += 1
                continue
            time.sleep(0.0002)
    wall = time.perf_counter() - t0
    return lat, busy / wall, launches


def make_target_data(cache: Path, seconds: float, n_streams: int = 8):
    """n_streams recordings of the 384-channel, 1024-template probe (the model the
    proposal network was trained for), seeds TEST_SEED + 100 + k."""
    from livespikesorter_amd import synth
    cache.mkdir(parents=True, exist_ok=True)
    d = cache / "cfg5_oss_input"
    m = synth.make_model(**synth.CONFIGS["cfg5"])
    if not (d / "ctc.npy").exists():
        synth.write_oss_input(m, d)
    recs = []
    for k in range(n_streams):
        rp = cache / f"cfg5_rec{k}_{int(seconds)}s.npy"
        if rp.exists():
            x = np.load(rp)
        else:
            x, _ = synth.make_recording(m, int(seconds * FS), seed=synth.TEST_SEED + 100 + k,
                                        workers=min(16, os.cpu_count() or 1))
            np.save(rp, x)
        recs.append(x)
    return m, d, recs


def peel_bytes(T, W, M, nC, iters, spikes):
    """Algorithmic HBM bytes of one stream's peel over one batch: the score tile
    read per iteration, the column maxima exchanged per iteration, the ctc rows
    and score updates per spike, the feature reads per spike."""
    return iters * 4.0 * T * W + iters * 8.0 * W + spikes * T * (2 * M + 1) * 12.0 + spikes * nC * M * 4.0


def cpu_baseline_legs(tab, chunks, N, budget_s, threads):
    """The FP32 C restatement of the matcher (oracle/liborc32) timed on host cores
    over chunks of the workload: all threads and one thread, under Kilosort's
    rules (100 iterations) and under the CUDA matcher's (50 iterations, integer
    Th).  Each leg runs until its share of the budget is spent."""
    import oracle
    legs = {}
    plan = [("kilosort", oracle.SEM_KILOSORT, 100, threads), ("kilosort_1thread", oracle.SEM_KILOSORT, 100, 1),
            ("lss", oracle.SEM_LSS, 50, threads), ("lss_1thread", oracle.SEM_LSS, 50, 1)]
    for name, sem, max_iter, th in plan:
        spent, done = 0.0, 0
        Th = float(int(tab["Th"])) if sem == oracle.SEM_LSS else tab["Th"]
        for batch in chunks:
            t1 = time.perf_counter()
            oracle.run_batch(batch, tab["whiteningMat"], tab["wPCA"], tab["Wall3"], tab["ctc"], tab["iCC"], tab["iU"],
                             Th=Th, nC=tab["nC"], max_iter=max_iter, semantics=sem, bits=32, threads=th)
            spent += time.perf_counter() - t1
            done += 1
            if spent >= budget_s / len(plan):
                break
        legs[name] = {"x_realtime": round(done * N / FS / spent, 4), "threads": th, "chunks": done,
                      "seconds": round(spent, 2), "max_iter": max_iter}
        log(f"cpu baseline {name}: {legs[name]}")
    return legs


def host_cpu_info():
    info = {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}
    try:
        txt = Path("/proc/cpuinfo").read_text()
        names = [l.split(":", 1)[1].strip() for l in txt.splitlines() if l.startswith("model name")]
        info["model"] = names[0] if names else None
        info["sockets"] = len({l for l in txt.splitlines() if l.startswith("physical id")}) or None
    except OSError:
        pass
    return info


def copy_calibration(mib: int) -> None:
    """Three device-to-device copies of a known size: the reference kernel for
    scaling rocprofv3's FETCH_SIZE / WRITE_SIZE counters to HBM bytes."""
    from livespikesorter_amd._hip import DeviceBuffer, synchronize
    nbytes = mib << 20
    src = DeviceBuffer(nbytes)
    dst = DeviceBuffer(nbytes)
    for _ in range(3):
        dst.copy_from(src, nbytes)
    synchronize()
    log(f"copy calibration: 3 x {mib} MiB device-to-device copies")
    src.free()
    dst.free()


HEADLINE_DTYPE = {
    "group_bf16x3_gated": "bf16 (whitening and projection as split-BF16 pairs on MFMA, FP32 accumulation; "
                          "proposal net) + f32 (peel)",
    "group_bf16x3": "bf16 (whitening and projection as split-BF16 pairs on MFMA, FP32 accumulation) + f32 (peel)",
    "group_bf16_gated": "bf16 (projection, proposal net) + f32 (rescoring, peel)",
    "group_bf16": "bf16 (projection) + f32 (rescoring, peel)",
    "group_fp32": "f32",
}
HEADLINE_FORM = {
    "group_bf16x3_gated": "split-BF16 whitening and projection (three BF16 MFMA products per FP32 product, FP32 "
                          "accumulation), proposal-network gate",
    "group_bf16x3": "split-BF16 whitening and projection (three BF16 MFMA products per FP32 product, FP32 accumulation)",
    "group_bf16_gated": "BF16 projection with FP32 rescoring, proposal-network gate",
    "group_bf16": "BF16 projection with FP32 rescoring",
    "group_fp32": "plain FP32 matcher",
}


def bench_target(args) -> dict:
    """``--streams`` (default 8) probe streams of 384 channels x 1024 templates,
    1500-sample chunks with 61-sample margins (W = 1622): the plain FP32 matcher
    one stream at a time (the reference's arrangement, one sorter per probe),
    and one graph launch per tick for the plain FP32, BF16 and gated BF16
    (trained proposal network) matchers.  Per mode: aggregate realtime factor
    with the recordings resident in HBM, paced latency with chunks handed over
    from host memory under +-5 ms arrival jitter, the peel kernel's device time
    (HIP events inside the graph) and spike agreement with the plain FP32
    matcher."""
    from livespikesorter_amd import OnlineSpikesV2, StreamGroup
    from livespikesorter_amd._hip import DeviceBuffer, synchronize
    from livespikesorter_amd.matcher import PREC_BF16_RESCORE, PREC_BF16X3, PREC_FP32
    from livespikesorter_amd.proposal import ProposalNet
    from tests.parity import agreement
    N, M = 1500, 61
    W = N + 2 * M
    t0 = time.time()
    m, d, recs = make_target_data(Path(args.cache), args.group_seconds, n_streams=args.streams)
    S = len(recs)
    C = recs[0].shape[1]
    n_ticks = min((len(x) - W) // N + 1 for x in recs)
    log(f"target data ready in {time.time() - t0:.1f}s: {S} streams x {n_ticks} chunks")
    bufs = [DeviceBuffer.from_array(x) for x in recs]
    synchronize()
    ptr = lambda s, j: bufs[s].ptr + (j % n_ticks) * N * C * 2
    steps = args.steps
    # plain FP32, one handle per stream, streams one after another: the agreement reference
    singles = [OnlineSpikesV2(d, chunk=N, max_batch=W) for _ in range(S)]
    ref = [[None] * n_ticks for _ in range(S)]
    for j in range(3):
        for s in range(S):
            singles[s].process_batch_device(ptr(s, j), W, C, first_sample=j * N)
    synchronize()
    ts = time.perf_counter()
    for j in range(n_ticks):
        for s in range(S):
            ref[s][j] = singles[s].process_batch_device(ptr(s, j), W, C, first_sample=j * N)[0]
    seq_s = time.perf_counter() - ts
    for sp in singles:
        sp.close()
    seq_rtf = n_ticks * S * N / FS / seq_s
    log(f"plain FP32, single handles in sequence: {seq_rtf:.1f}x realtime aggregate")
    weights = ROOT / "livespikesorter_amd" / "data" / "proposal_cfg5.npy"
    net = ProposalNet(C, W, seed=1)
    net.load(weights)
    modes = {}
    all_modes = (("group_fp32", PREC_FP32, False), ("group_bf16", PREC_BF16_RESCORE, False),
                 ("group_bf16_gated", PREC_BF16_RESCORE, True), ("group_bf16x3", PREC_BF16X3, False),
                 ("group_bf16x3_gated", PREC_BF16X3, True))
    headline = "group_bf16x3_gated" if "group_bf16x3_gated" in args.modes.split(",") else "group_bf16_gated"
    outs_by_mode = {}
    peel_by_mode = {}
    for name, prec, gated in [mm for mm in all_modes if mm[0] in args.modes.split(",")]:
        def make(profile):
            # profiled groups time the front stages and the peel with events (level 2:
            # without the peel's phase clock, so the peel runs as in production)
            gr = StreamGroup([d] * S, batch=W, precision=prec, peel_split=args.group_split,
                             profile=2 if profile else 0)
            if gated:
                for s in range(S):
                    gr.set_proposal(s, net, args.gate_lo, args.gate_hi, args.gate_reach)
            return gr
        # timed: one graph launch per tick, no stage events (the production configuration);
        # the results stay in the lane's host buffers (no Python-side copies in the loop)
        grp = make(False)
        for j in range(args.warmup):
            grp.process([ptr(s, j) for s in range(S)], [(j % n_ticks) * N] * S, device=True, lds=[C] * S,
                        copy=False)
        synchronize()
        wall = []
        ts = time.perf_counter()
        for j in range(steps):
            t1 = time.perf_counter()
            grp.process([ptr(s, j) for s in range(S)], [(j % n_ticks) * N] * S, device=True, lds=[C] * S,
                        copy=False)
            wall.append((time.perf_counter() - t1) * 1e3)
        el = time.perf_counter() - ts
        # pipelined: the same ticks on two lanes from two host threads (consecutive ticks in
        # flight at once: the next tick's front stages beside the current peel)
        el_pipe = None
        if args.pipeline:
            grp.set_pipelined(True)

            def lane_run(lane, lo, hi):
                for jj in range(lo + lane, hi, 2):
                    grp.process([ptr(s, jj) for s in range(S)], [(jj % n_ticks) * N] * S, device=True,
                                lds=[C] * S, lane=lane, copy=False)

            for lo_, hi_ in ((0, args.warmup), (0, steps)):
                synchronize()
                tp = time.perf_counter()
                th = [threading.Thread(target=lane_run, args=(lane, lo_, hi_)) for lane in (0, 1)]
                for t_ in th:
                    t_.start()
                for t_ in th:
                    t_.join()
                el_pipe = time.perf_counter() - tp
            grp.set_pipelined(False)
        # paced run, chunks handed over from host memory (int16 hand-off -> events on the host)
        lat, busy, nt = paced_latency(grp, bufs, n_ticks, N, C, args.group_seconds, seed=20261205,
                                      lanes=args.lanes, host_recs=recs)
        grp.close()
        # untimed pass over the recording on a profiled group: the device split of a tick
        # (HIP events around the front stages' graph and the peel's), the outputs for the
        # agreement, and the peel's iterations and spikes per stream and tick (the
        # algorithmic work)
        grp = make(True)
        for j in range(3):
            grp.process([ptr(s, j) for s in range(S)], [j * N] * S, device=True, lds=[C] * S)
        grp.stage_times(0, reset=True)
        outs, work_b = [], []
        for j in range(n_ticks):
            outs.append(grp.process([ptr(s, j) for s in range(S)], [j * N] * S, device=True, lds=[C] * S))
            qs = [grp.query(s) for s in range(S)]
            work_b.append(sum(peel_bytes(q.T, W, M, q.nC, q.last_iters, q.last_spikes) for q in qs))
        front_ms, peel_ms, launches = grp.stage_times(0, reset=True)
        tot, ident_n, within_n = 0, 0.0, 0.0
        for j, out in enumerate(outs):
            for s in range(S):
                e, r = out[s][0], ref[s][j]
                i, w_, _ = agreement(r["sample"], r["template"], e["sample"], e["template"])
                nn = max(len(r), len(e))
                ident_n += i * nn
                within_n += w_ * nn
                tot += nn
        q = [grp.query(s) for s in range(S)]
        grp.close()
        rtf = steps * S * N / FS / el
        peel_avg = peel_ms / max(launches, 1)
        res = {"x_realtime": round(rtf, 2), "ms_per_tick": round(el / steps * 1e3, 4),
               "pipelined": (None if el_pipe is None else
                             {"x_realtime": round(steps * S * N / FS / el_pipe, 2),
                              "ms_per_tick": round(el_pipe / steps * 1e3, 4)}),
               "tick_p50_ms": round(pct(wall, 50), 3), "tick_p99_ms": round(pct(wall, 99), 3),
               "paced_latency_p50_ms": round(pct(lat, 50), 3), "paced_latency_p99_ms": round(pct(lat, 99), 3),
               "paced_input": "host int16", "paced_gpu_busy_fraction": round(busy, 4), "paced_launches": nt,
               "device_ms_per_tick": {"front_stages": round(front_ms / max(launches, 1), 4),
                                      "peel": round(peel_avg, 4)},
               "agreement_vs_plain_fp32": {"identical": round(ident_n / max(tot, 1), 6),
                                           "within_1_sample": round(within_n / max(tot, 1), 6), "events": tot},
               "speedup_over_plain_fp32_sequential": round(rtf / seq_rtf, 3),
               "peel_geometry": group_geometry(q),
               "rescored_per_tick": round(sum(x.rescored for x in q) / max(1, q[0].batches), 1)}
        if gated:
            sc = sum(x.gate_tiles_scored for x in q)
            tt = sum(x.gate_tiles for x in q)
            res["tiles_scored_fraction"] = round(sc / max(tt, 1), 4)
        modes[name] = res
        outs_by_mode[name] = outs
        peel_by_mode[name] = (peel_avg, float(np.mean(work_b)))
        log(f"{name}: {res}")
    net.close()
    for b in bufs:
        b.free()
    if "group_fp32" in modes:
        for name, res in modes.items():
            res["speedup_over_plain_fp32_group"] = round(res["x_realtime"] / modes["group_fp32"]["x_realtime"], 3)
    if headline not in modes:
        headline = next(iter(modes))
    g = modes[headline]
    # roofline of the dominant kernel, the group's peel: all streams' algorithmic
    # bytes per launch over its HIP-event-timed device time per launch
    peel_avg, peel_b = peel_by_mode[headline]
    achieved = peel_b / (peel_avg / 1e3) / 1e9
    traffic = None
    tf = ROOT / "profiles" / "target_traffic.json"
    if tf.exists():
        tr = json.loads(tf.read_text())
        rs = "true" if headline in ("group_bf16", "group_bf16_gated") else "false"   # the headline's peel instance
        bpl = tr.get("bytes_per_launch", {})
        rows = g['peel_geometry']['rows_per_thread']
        # the production instance: no phase clock, Kilosort rules
        traffic = bpl.get(f"lss::k_peel<{rows}, {rs}, false, false>")
        traffic = round(traffic) if traffic else None
    roofline = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "kernel": f"k_peel<{g['peel_geometry']['rows_per_thread']}> ({S} streams in one grid)",
                "ms_per_launch": round(peel_avg, 5), "algorithmic_bytes": round(peel_b, 1)}
    # spike-list differences of every mode against the FP64 oracle on a sample, each
    # classified by the oracle's decision margins (near tie / cascade / unexplained; the gated
    # mode's also against its ungated run: gate miss)
    explain = None
    import oracle
    tab = oracle.load_tables(d)
    threads = min(len(os.sched_getaffinity(0)), int(os.environ.get("OMP_NUM_THREADS", "16") or 16))
    if not args.no_cpu and args.explain_ticks > 0:
        from tests.parity import explain_mismatches
        explain = {}
        refs64 = {}
        for j in range(min(args.explain_ticks, n_ticks)):
            for s in range(S):
                refs64[(j, s)] = oracle.run_model_batch(tab, recs[s][j * N:j * N + W], margins=True, threads=threads,
                                                        features=False)
        for name, outs in outs_by_mode.items():
            # the gated mode's differences are also checked against the same matcher without the gate
            ung = outs_by_mode.get(name[:-len("_gated")]) if name.endswith("_gated") else None
            agg = {"events_fp64": 0, "identical": 0, "differing": 0, "near_tie": 0, "cascade": 0, "unexplained": 0}
            if ung is not None:
                agg["gate_miss"] = 0
            for (j, s), r in refs64.items():
                ev = outs[j][s][0]
                i, _, _ = agreement(r.w + j * N - M // 2 + m.nt0min, r.t, ev["sample"], ev["template"])
                ex = explain_mismatches(r, ev, first_sample=j * N, M=M, nt0min=m.nt0min,
                                        ungated=None if ung is None else ung[j][s][0])
                agg["events_fp64"] += len(r.w)
                agg["identical"] += int(round(i * max(len(r.w), len(ev))))
                agg["differing"] += ex["events"]
                for k in ("near_tie", "cascade", "unexplained") + (("gate_miss",) if ung is not None else ()):
                    agg[k] += ex[k]
            agg["chunks"] = len(refs64)
            explain[name] = agg
            log(f"vs FP64 oracle, {name}: {agg}")
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
    # the CPU baseline on the same workload: chunks of every stream, tick by tick
    cpu = None
    if not args.no_cpu:
        chunks = [recs[s][j * N:j * N + W] for j in range(n_ticks) for s in range(S)]
        legs = cpu_baseline_legs(tab, chunks, N, args.cpu_seconds, threads)
        k = legs["kilosort"]
        cpu = {"value": k["x_realtime"], "unit": "x realtime (one stream)", "cores": threads, "kind": "port",
               "sample": f"{k['chunks']} chunks of the {S} streams' recordings (384 ch x 1024 templates, W={W}), "
                         f"FP32 C restatement (oracle/liborc32, Kilosort rules, 100 iterations, features), "
                         f"OpenMP {threads} threads, {k['seconds']} s",
               "legs": legs, "host": host_cpu_info()}
    return {
        "metric": METRIC,
        "value": g["x_realtime"],
        "unit": f"x realtime (aggregate over {S} streams)",
        "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": g["ms_per_tick"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": HEADLINE_DTYPE.get(headline, "f32"),
        "data": f"synthetic (seeded generator: {S} recordings of one 384-ch, 1024-template probe)",
        "config": {"workload": f"{S} streams x 384 ch x 1024 templates, 1500-sample chunks + 61-sample margins "
                               f"(W={W}), one graph launch per tick, {HEADLINE_FORM.get(headline, headline)}",
                   "streams": S, "channels": C, "templates": 1024, "chunk": N, "batch_columns": W,
                   "mode": headline},
        "p50_ms": g["paced_latency_p50_ms"],
        "p99_ms": g["paced_latency_p99_ms"],
        "latency": "paced: chunks handed over as int16 in host memory at 30 kHz with +-5 ms jitter, "
                   "to events and features on the host",
        "agreement_vs_plain_fp32": g["agreement_vs_plain_fp32"],
        "speedup_over_plain_fp32_group": g.get("speedup_over_plain_fp32_group"),
        "speedup_over_plain_fp32_sequential": g["speedup_over_plain_fp32_sequential"],
        "plain_fp32_sequential_x_realtime": round(seq_rtf, 2),
        # the same ticks with two in flight (two pipelined lanes): the capacity with a
        # backlog, where every mode is bound by the peel alone
        "pipelined": g.get("pipelined"),
        "modes": modes,
        "vs_fp64_oracle": explain,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }


def stage_work(info, C, W, M, iters, spikes, bf16=False):
    """Algorithmic work per launch of each device stage: (bound, amount, unit, peak).

    whiten / project: MFMA flops over the listed (nonzero) 64x16 operand blocks;
    pcconv: int16-converted input read once plus the PC projection written once
    (f32; BF16 mode also writes the BF16 copy); peel: the score tile read per
    iteration, Cfmax exchange, ctc rows and score updates per spike, and the
    feature reads per spike."""
    T, K = info.T, info.K
    Wc = (W + 63) // 64 * 64
    I, S = float(np.mean(iters)), float(np.mean(spikes))
    pk = BF16_MFMA_PEAK_TFLOPS if bf16 else FP32_MFMA_PEAK_TFLOPS
    return {
        "whiten": ("mfma", 2.0 * info.white_kblocks * 64 * 16 * Wc / 1e12, "TFLOP/s", FP32_MFMA_PEAK_TFLOPS),
        "pcconv": ("hbm", (4.0 * C * W + (6.0 if bf16 else 4.0) * K * C * W) / 1e9, "GB/s", HBM_PEAK_GBS),
        "project": ("mfma", 2.0 * info.proj_kblocks * 64 * 16 * Wc / 1e12, "TFLOP/s", pk),
        "peel": ("hbm", (I * 4.0 * T * W + I * 8.0 * W + S * T * (2 * M + 1) * 12.0
                         + S * 10 * M * 4.0) / 1e9, "GB/s", HBM_PEAK_GBS),
    }


def make_bf16_data(cache: Path, seconds: float):
    """384 ch, 1024 templates, bursty gamma(0.5) ISIs (the BF16 precision workload)."""
    from livespikesorter_amd import synth
    cache.mkdir(parents=True, exist_ok=True)
    d = cache / "cfg3_oss_input"
    m = synth.make_model(**synth.CONFIGS["cfg3"])
    if not (d / "ctc.npy").exists():
        synth.write_oss_input(m, d)
    rp = cache / f"cfg3_rec_{int(seconds)}s.npy"
    if rp.exists():
        x = np.load(rp)
    else:
        x, _ = synth.make_recording(m, int(seconds * FS))
        np.save(rp, x)
    return m, d, x


def bench_bf16(args) -> dict:
    """BF16 projection + FP32 rescoring on 60000-sample batches, against the FP32 GPU
    matcher on the same batches and the FP64 oracle on a bounded sample."""
    import oracle
    from livespikesorter_amd import OnlineSpikesV2
    from livespikesorter_amd._hip import DeviceBuffer, synchronize
    from livespikesorter_amd.matcher import PREC_BF16_RESCORE, PREC_BF16X3, PREC_FP32
    from tests.parity import agreement
    N, M = 60000, 61
    W = N + 2 * M
    t0 = time.time()
    m, d, x = make_bf16_data(Path(args.cache), args.bf16_seconds)
    n_batches = (len(x) - W) // N + 1
    log(f"bf16 data ready in {time.time() - t0:.1f}s: {len(x)} samples, {n_batches} batches of {W}")
    dx = DeviceBuffer.from_array(x)
    C = x.shape[1]
    synchronize()
    steps = min(args.steps, n_batches)

    def run(prec):
        sp = OnlineSpikesV2(d, chunk=N, max_batch=W, precision=prec, profile=True)
        for j in range(min(args.warmup, 2)):
            sp.process_batch_device(dx.ptr + (j % n_batches) * N * C * 2, W, C, first_sample=j * N)
        sp.stage_times(reset=True)
        lat, evs, iters, spikes = [], [], [], []
        synchronize()
        ts = time.perf_counter()
        for j in range(steps):
            t1 = time.perf_counter()
            ev, _ = sp.process_batch_device(dx.ptr + j * N * C * 2, W, C, first_sample=j * N)
            lat.append((time.perf_counter() - t1) * 1e3)
            q = sp.query()
            iters.append(q.last_iters)
            spikes.append(q.last_spikes)
            evs.append(ev)
        synchronize()
        el = time.perf_counter() - ts
        st, nb = sp.stage_times(reset=True)
        return sp, el, lat, evs, iters, spikes, {k: v / max(nb, 1) for k, v in st.items()}

    sb, el_b, lat_b, ev_b, it_b, sk_b, st_b = run(PREC_BF16_RESCORE)
    sf, el_f, lat_f, ev_f, it_f, sk_f, st_f = run(PREC_FP32)
    sx, el_x, lat_x, ev_x, it_x, sk_x, st_x = run(PREC_BF16X3)
    sx.close()
    rtf_b = steps * N / FS / el_b
    rtf_f = steps * N / FS / el_f
    rtf_x = steps * N / FS / el_x
    log(f"split-bf16: {rtf_x:.1f}x realtime, p99 {pct(lat_x, 99):.2f} ms; stages ms: "
        + ", ".join(f"{k} {v:.3f}" for k, v in st_x.items()))
    log(f"bf16: {rtf_b:.1f}x realtime, p99 {pct(lat_b, 99):.2f} ms; fp32: {rtf_f:.1f}x, p99 {pct(lat_f, 99):.2f} ms")
    log("bf16 stages ms: " + ", ".join(f"{k} {v:.3f}" for k, v in st_b.items()))
    log("fp32 stages ms: " + ", ".join(f"{k} {v:.3f}" for k, v in st_f.items()))
    # spike agreement with the plain FP32 GPU matcher over the compared batches
    n_cmp = min(steps, args.bf16_compare)

    def vs_fp32_of(evs):
        tot = ident_n = within_n = 0
        for j in range(n_cmp):
            a, b = ev_f[j], evs[j]
            i, w, _ = agreement(a["sample"], a["template"], b["sample"], b["template"])
            n = max(len(a), len(b))
            ident_n += i * n
            within_n += w * n
            tot += n
        return {"batches": n_cmp, "events": tot, "identical": round(ident_n / max(tot, 1), 6),
                "within_1_sample": round(within_n / max(tot, 1), 6)}
    vs_fp32 = vs_fp32_of(ev_b)
    vs_fp32_x = vs_fp32_of(ev_x)
    log(f"bf16 vs fp32 GPU: {vs_fp32}; split-bf16 vs fp32 GPU: {vs_fp32_x}")
    info = sb.query()
    # the FP64 oracle (and the FP32 CPU restatement, the CPU baseline) on the first batch
    tab = oracle.load_tables(d)
    threads = min(len(os.sched_getaffinity(0)), int(os.environ.get("OMP_NUM_THREADS", "16") or 16))
    batch = x[:W]
    vs_fp64 = cpu = None
    if not args.no_cpu:
        t1 = time.perf_counter()
        ref = oracle.run_model_batch(tab, batch, bits=64, threads=threads, features=False)
        t64 = time.perf_counter() - t1
        t1 = time.perf_counter()
        oracle.run_model_batch(tab, batch, bits=32, threads=threads, features=False)
        t32 = time.perf_counter() - t1
        rs = ref.w - M // 2 + m.nt0min
        res = {}
        for name, ev in (("bf16", ev_b[0]), ("fp32", ev_f[0]), ("split_bf16", ev_x[0])):
            i, w, pairs = agreement(rs, ref.t, ev["sample"], ev["template"])
            ri = np.array([p[0] for p in pairs])
            gi = np.array([p[1] for p in pairs])
            ex = rs[ri] == ev["sample"][gi]
            err = np.abs(ev["amplitude"][gi[ex]] - ref.amp[ri[ex]]) / np.abs(ref.amp[ri[ex]])
            res[name] = {"identical": round(i, 6), "within_1_sample": round(w, 6),
                         "amp_rel_err_max": float(err.max()) if len(err) else None,
                         "amp_rel_err_median": float(np.median(err)) if len(err) else None}
        vs_fp64 = {"batches": 1, "events_fp64": int(len(rs)), **res, "oracle_seconds": round(t64, 2)}
        log(f"vs FP64 oracle: {vs_fp64}")
        cpu = {"value": round(N / FS / t32, 4), "unit": "x realtime", "cores": threads, "kind": "port",
               "sample": f"first 60000-sample batch (W={W}), FP32 C restatement without features, "
                         f"OpenMP {threads} threads, {t32:.1f}s"}
    sb.close()
    sf.close()
    dx.free()
    work = stage_work(info, C, W, M, it_b, sk_b, bf16=True)
    dom = max(work, key=lambda k: st_b[k])
    bound, amount, unit, peak = work[dom]
    achieved = amount / (st_b[dom] / 1e3)
    proj = work["project"]
    return {
        "metric": METRIC + " [BF16 workload]",
        "value": round(rtf_b, 3),
        "unit": "x realtime",
        "n_gpus": 1, "steps": steps, "warmup": min(args.warmup, 2),
        "ms_per_step": round(el_b / steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16 (projection), f32",
        "data": "synthetic (seeded generator: 384 ch, 1024 templates, gamma(0.5) ISIs, banded whitening)",
        "config": {"workload": "BF16 MFMA projection with FP32 accumulation, local maxima with Cf in "
                               "[0.9, 1.1] Th^2 rescored in FP32; 384 channels, 1024 templates, "
                               f"60000-sample batches + 61-sample margins (W={W}), "
                               f"{args.bf16_seconds:.0f} s recording resident in HBM",
                   "channels": C, "templates": info.T, "batch_columns": W},
        "p50_ms": round(pct(lat_b, 50), 3), "p99_ms": round(pct(lat_b, 99), 3),
        "fp32_x_realtime": round(rtf_f, 3), "fp32_p99_ms": round(pct(lat_f, 99), 3),
        "speedup_over_fp32": round(rtf_b / rtf_f, 4),
        "split_bf16": {"x_realtime": round(rtf_x, 3), "p99_ms": round(pct(lat_x, 99), 3),
                       "speedup_over_fp32": round(rtf_x / rtf_f, 4), "vs_fp32_gpu": vs_fp32_x,
                       "stages_ms": {k: round(v, 4) for k, v in st_x.items()}},
        "project_ms": {"bf16": round(st_b["project"], 4), "fp32": round(st_f["project"], 4)},
        "project_tflops": {"bf16": round(proj[1] / (st_b["project"] / 1e3), 2),
                           "fp32": round(proj[1] / (st_f["project"] / 1e3), 2)},
        "stages_ms": {k: round(v, 4) for k, v in st_b.items()},
        "mean_iterations": round(float(np.mean(it_b)), 2), "mean_spikes_per_batch": round(float(np.mean(sk_b)), 1),
        "rescored": int(info.rescored), "rescore_flips": int(info.rescore_flips),
        "vs_fp32_gpu": vs_fp32, "vs_fp64_oracle": vs_fp64,
        "roofline": {"bound": bound, "achieved": round(achieved, 3), "peak": peak, "unit": unit,
                     "frac": round(achieved / peak, 5), "traffic": None, "kernel": dom,
                     "ms_per_launch": round(st_b[dom], 5)},
        "cpu_baseline": cpu,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--seconds", type=float, default=60.0, help="length of the synthetic recording")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--host-steps", type=int, default=100, help="chunks timed with host-resident input")
    ap.add_argument("--pre-steps", type=int, default=200,
                    help="chunks timed again with preprocessing (means, CAR, high-pass) switched on")
    ap.add_argument("--cache", default=os.environ.get("LSS_BENCH_CACHE", "/tmp/lss_bench_cache"))
    ap.add_argument("--no-autotune", action="store_true",
                    help="cfg2: keep the default peel geometry instead of lss_autotune's choice")
    ap.add_argument("--workload", choices=("cfg2", "group", "bf16", "target"), default="target",
                    help="target (default): 8 streams x 384 ch x 1024 templates per graph launch, gated BF16 as "
                         "the headline beside plain FP32 and BF16; cfg2: the plain FP32 matcher on one 512-template "
                         "stream; group: 8 streams of 256-1024 templates; bf16: BF16 projection with FP32 "
                         "rescoring, 1024 templates, 60000-sample batches")
    ap.add_argument("--no-secondary", action="store_true", help="target: skip the cfg2 secondary keys")
    ap.add_argument("--explain-ticks", type=int, default=1,
                    help="target: ticks whose spike lists are compared with the FP64 oracle (mismatch margins)")
    ap.add_argument("--gate-lo", type=float, default=0.1, help="target: proposal gate fallback band, low edge")
    ap.add_argument("--gate-hi", type=float, default=0.5, help="target: proposal gate acceptance threshold")
    ap.add_argument("--gate-reach", type=int, default=12, help="target: proposal gate reach in columns")
    ap.add_argument("--bf16-seconds", type=float, default=120.0, help="recording length (bf16)")
    ap.add_argument("--bf16-compare", type=int, default=12, help="batches compared with FP32 (bf16)")
    ap.add_argument("--group-seconds", type=float, default=6.0, help="recording length per stream (group)")
    ap.add_argument("--inflight", type=int, default=2,
                    help="cfg2: also time this many handles on alternating chunks from as many threads "
                         "(reported as `inflight`, beside `value`; 1 = skip)")
    ap.add_argument("--streams", type=int, default=8, help="target workload: probe streams (1-8)")
    ap.add_argument("--group-split", type=int, default=0,
                    help="group and target workloads: peel template split (0 automatic, 1 none, n ways)")
    ap.add_argument("--lanes", type=int, default=4, help="concurrent launch lanes of the paced group runs (1-4)")
    ap.add_argument("--pipeline", type=int, default=1,
                    help="target: also time the ticks on two pipelined lanes from two host threads (0 = skip)")
    ap.add_argument("--modes", default="group_fp32,group_bf16,group_bf16_gated,group_bf16x3,group_bf16x3_gated",
                    help="target workload: which group modes to run")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="after the run, copy a MIB-sized buffer three times (a known-size kernel "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the matcher runs one independent instance per probe "
              "stream per GPU and this build has no multi-GPU path; run with --gpus 1", file=sys.stderr)
        return 2

    if args.workload == "group":
        print(json.dumps(bench_group(args)), flush=True)
        return 0
    if args.workload == "bf16":
        print(json.dumps(bench_bf16(args)), flush=True)
        return 0
    if args.workload == "cfg2":
        print(json.dumps(bench_cfg2(args)), flush=True)
        return 0
    out = bench_target(args)
    if not args.no_secondary:
        # the plain FP32 matcher on one 384-channel, 512-template stream in 3000-sample
        # chunks, as secondary keys of the same line
        a2 = argparse.Namespace(**vars(args))
        a2.no_cpu, a2.pre_steps, a2.inflight, a2.host_steps, a2.copy_calibration = True, 0, 1, 50, 0
        c2 = bench_cfg2(a2)
        out["secondary"] = {"cfg2_plain_fp32_single_stream": {
            k: c2[k] for k in ("value", "unit", "config", "ms_per_step", "p50_ms", "p99_ms", "p99_ms_host_input",
                               "mean_iterations", "mean_spikes_per_chunk", "peel_geometry", "roofline",
                               "peel_phases_us", "stages")}}
    print(json.dumps(out), flush=True)
    return 0


def bench_cfg2(args) -> dict:
    """Plain FP32 matcher, one 384-channel, 512-template stream, 3000-sample
    chunks with 61-sample margins (W = 3122), 60 s resident in HBM."""
    from livespikesorter_amd import OnlineSpikesV2
    from livespikesorter_amd._hip import DeviceBuffer, synchronize
    from livespikesorter_amd.matcher import NPHASE, PHASE_NAMES, STAGE_PHASES

    N, M = 3000, 61
    W = N + 2 * M
    t0 = time.time()
    m, d, x = make_data(Path(args.cache), args.seconds)
    n_batches = (len(x) - W) // N + 1
    log(f"data ready in {time.time() - t0:.1f}s: {len(x)} samples x {x.shape[1]} ch, {n_batches} chunks")

    dx = DeviceBuffer.from_array(x)            # the whole recording resident in HBM
    synchronize()
    base = dx.ptr
    C = x.shape[1]

    # HIP events around the peel only (profile level 2): the per-stage events of
    # level 1 leave the GPU idle between the kernels (about 30 us per chunk)
    sp = OnlineSpikesV2(d, chunk=N, profile=2)
    info = sp.info
    # the peel geometry for this batch width, timed on a chunk (lss_autotune), as a
    # caller would pick it before streaming
    tuned = [] if args.no_autotune else sp.autotune(base + 10 * N * C * 2, W, C, reps=20)
    if tuned:
        log("peel autotune (split, rows, ms, chosen): " + str(tuned))

    def step(j):
        b = j % n_batches
        return sp.process_batch_device(base + b * N * C * 2, W, C, first_sample=b * N)

    for j in range(args.warmup):
        step(j)
    sp.stage_times(reset=True)
    lat, iters, spikes = [], [], []
    synchronize()
    t_start = time.perf_counter()
    for j in range(args.warmup, args.warmup + args.steps):
        t1 = time.perf_counter()
        ev, _ = step(j)
        lat.append((time.perf_counter() - t1) * 1e3)
        q = sp.query()                       # host-side counters of the last batch
        iters.append(q.last_iters)
        spikes.append(q.last_spikes)
    synchronize()
    elapsed = time.perf_counter() - t_start
    q_last = sp.query()
    stage_ms, nb = sp.stage_times(reset=True)
    # Untimed diagnostics on a second handle with every stage timed (profile level 1)
    # and the same peel split, over the first chunks of the timed range: the other
    # stages' device times and the peel's phase clock (read back per chunk with a
    # synchronous copy).
    phases = np.zeros(2 * NPHASE + 2)
    n_ph = min(args.steps, 100)
    with OnlineSpikesV2(d, chunk=N, profile=1, peel_split=int(q_last.peel_split)) as spd:
        for j in range(args.warmup, args.warmup + n_ph):
            b = j % n_batches
            spd.process_batch_device(base + b * N * C * 2, W, C, first_sample=b * N)
            phases += spd.debug_stage(STAGE_PHASES)
        dst, dnb = spd.stage_times(reset=True)
    rtf = args.steps * N / FS / elapsed
    log(f"timed {args.steps} chunks in {elapsed:.3f}s: {rtf:.1f}x realtime, p50 {pct(lat, 50):.3f} ms, "
        f"p99 {pct(lat, 99):.3f} ms; mean iterations {np.mean(iters):.1f}, spikes {np.mean(spikes):.1f}")
    # the peel from the timed region; the other stages from the diagnostic pass
    avg = {k: v / max(dnb, 1) for k, v in dst.items()}
    avg["peel"] = stage_ms["peel"] / max(nb, 1)
    log("device ms per chunk by stage (peel: timed region; others: diagnostic pass): " +
        ", ".join(f"{k} {v:.4f}" for k, v in avg.items()))
    # the peel kernel's own clock (100 MHz ticks), workgroup 0, per chunk
    nwg = max(phases[2 * NPHASE + 1], 1.0)
    peel_phases = {k: {"mean_wg": round(phases[i] / nwg / 100.0, 3),
                       "max_wg": round(phases[NPHASE + i] / n_ph / 100.0, 3)}
                   for i, k in enumerate(PHASE_NAMES)}
    peel_phases["scan_passes_per_wg"] = round(phases[2 * NPHASE] / nwg, 2)
    log(f"peel phases, us per chunk (mean over workgroups / largest): {peel_phases}")

    # host-resident input: chunk latency from int16 on the host to events on the host
    host_lat = []
    for j in range(args.host_steps):
        b = j % n_batches
        t1 = time.perf_counter()
        sp.process_batch(x[b * N:b * N + W], first_sample=b * N)
        host_lat.append((time.perf_counter() - t1) * 1e3)
    sp.stage_times(reset=True)

    # the same chunks with preprocessing (channel means, CAR, high-pass) in front of whitening
    pre = None
    if args.pre_steps > 0:
        from livespikesorter_amd import PRE_ALL
        with OnlineSpikesV2(d, chunk=N, profile=1, preprocess=PRE_ALL) as spp:
            def pstep(j):
                b = j % n_batches
                return spp.process_batch_device(base + b * N * C * 2, W, C, first_sample=b * N)
            for j in range(5):
                pstep(j)
            spp.stage_times(reset=True)
            plat = []
            synchronize()
            ts = time.perf_counter()
            for j in range(args.pre_steps):
                t1 = time.perf_counter()
                pstep(j)
                plat.append((time.perf_counter() - t1) * 1e3)
            synchronize()
            pel = time.perf_counter() - ts
            pst, pnb = spp.stage_times(reset=True)
        pavg = {k: round(v / max(pnb, 1), 5) for k, v in pst.items() if k not in ("h2d_copy", "proposal")}
        pre = {"steps": "mean removal + CAR lower median + high-pass (hp_filter.npy, Kilosort 300 Hz) per chunk",
               "chunks": args.pre_steps, "value": round(args.pre_steps * N / FS / pel, 3), "unit": "x realtime",
               "p50_ms": round(pct(plat, 50), 4), "p99_ms": round(pct(plat, 99), 4), "stage_ms": pavg}
        log(f"with preprocessing: {pre['value']}x realtime, p99 {pre['p99_ms']} ms, stages {pavg}")

    # Optional: N handles of the same probe on alternating chunks from N host
    # threads, as a caller with a backlog could run them: one chunk's whitening,
    # PC conv and projection can run on the CUs that another chunk's peel leaves
    # idle.  Reported beside `value` (one chunk in flight), never as it.
    inflight = None
    if args.inflight > 1:
        from concurrent.futures import ThreadPoolExecutor
        n_h = args.inflight
        hs = [OnlineSpikesV2(d, chunk=N, peel_split=int(q_last.peel_split)) for _ in range(n_h)]
        n_cmp = min(50, args.steps)
        outs = {}

        def worker(k, record):
            for j in range(k, args.steps, n_h):
                b = j % n_batches
                r = hs[k].process_batch_device(base + b * N * C * 2, W, C, first_sample=b * N)
                if record and j < n_cmp:
                    outs[j] = r

        with ThreadPoolExecutor(n_h) as pool:
            list(pool.map(lambda k: worker(k, True), range(n_h)))      # warm-up pass, outputs kept
            synchronize()
            ts = time.perf_counter()
            list(pool.map(lambda k: worker(k, False), range(n_h)))
            synchronize()
            iel = time.perf_counter() - ts
        for h in hs:
            h.close()
        same = all(np.array_equal(outs[j][0], ref[0]) and np.array_equal(outs[j][1], ref[1])
                   for j in range(n_cmp)
                   for ref in [sp.process_batch_device(base + (j % n_batches) * N * C * 2, W, C,
                                                       first_sample=(j % n_batches) * N)])
        inflight = {"handles": n_h, "chunks": args.steps, "x_realtime": round(args.steps * N / FS / iel, 3),
                    "identical_to_one_handle": bool(same), "chunks_compared": n_cmp}
        log(f"{n_h} handles on alternating chunks: {inflight}")

    # roofline of the dominant device stage (algorithmic work per launch / measured time)
    T = info.T
    I = float(np.mean(iters))
    S = float(np.mean(spikes))
    work = stage_work(info, C, W, M, iters, spikes)
    dom = max(work, key=lambda k: avg[k])
    bound, amount, unit, peak = work[dom]
    achieved = amount / (avg[dom] / 1e3)
    # measured HBM bytes per launch of that kernel, from the committed rocprofv3
    # FETCH_SIZE / WRITE_SIZE passes over this same workload (scaled by a copy kernel)
    traffic = None
    tf = ROOT / "profiles" / "cfg2_plain_fp32_traffic.json"
    if tf.exists():
        # the peel's entry is the kernel of the timed chunks (its tile height)
        kname = {"whiten": "lss::k_whiten", "pcconv": "lss::k_pcconv", "project": "lss::k_project",
                 "peel": f"lss::k_peel<{int(q_last.peel_rows)}"}[dom]
        for k, v in json.loads(tf.read_text())["bytes_per_launch"].items():
            if k.startswith(kname) and (dom != "peel" or k[len(kname)] in ">,"):
                traffic = round(v)
    roofline = {"bound": bound, "achieved": round(achieved, 3), "peak": peak, "unit": unit,
                "frac": round(achieved / peak, 5), "traffic": traffic, "kernel": dom,
                "ms_per_launch": round(avg[dom], 5),
                "algorithmic_bytes" if bound == "hbm" else "algorithmic_tflop": round(amount * (1e9 if bound == "hbm" else 1.0), 3)}
    # stage times (the proposal stage is empty without a network: it only measures the event gap)
    stages = {k: {"ms": round(avg[k], 5), "bound": work[k][0] if k in work else None,
                  "achieved": round(work[k][1] / (avg[k] / 1e3), 3) if k in work and avg[k] > 0 else None,
                  "unit": work[k][2] if k in work else None} for k in avg if k not in ("h2d_copy", "proposal", "preprocess")}

    cpu = None
    parity = None
    if not args.no_cpu:
        import oracle
        from tests.parity import agreement
        tab = oracle.load_tables(d)
        # the GPU box's CPU share is 16 threads per GPU (OMP_NUM_THREADS is set to it there)
        threads = min(len(os.sched_getaffinity(0)), int(os.environ.get("OMP_NUM_THREADS", "16") or 16))
        done, cpu_t, ident_n, within_n, tot_n, amp_err = 0, 0.0, 0, 0, 0, 0.0
        j = 0
        while cpu_t < args.cpu_seconds and j < n_batches:
            batch = x[j * N:j * N + W]
            t1 = time.perf_counter()
            ref = oracle.run_model_batch(tab, batch, bits=32, threads=threads)
            cpu_t += time.perf_counter() - t1
            ev, _ = sp.process_batch(batch, first_sample=j * N)
            rs = ref.w + j * N - M // 2 + 20
            ident, within, pairs = agreement(rs, ref.t, ev["sample"], ev["template"])
            n = max(len(rs), len(ev))
            ident_n += ident * n
            within_n += within * n
            tot_n += n
            if pairs:
                ri = np.array([p[0] for p in pairs])
                gi = np.array([p[1] for p in pairs])
                amp_err = max(amp_err, float(np.max(np.abs(ev["amplitude"][gi] - ref.amp[ri]) / np.abs(ref.amp[ri]))))
            done += 1
            j += 1
        cpu = {"value": round(done * N / FS / cpu_t, 4), "unit": "x realtime", "cores": threads,
               "kind": "port",
               "sample": f"first {done} chunks of the same recording (W={W}), FP32 C restatement "
                         f"(Kilosort rules, 100 iterations), OpenMP {threads} threads, {cpu_t:.1f}s"}
        parity = {"vs": "CPU FP32 restatement", "chunks": done, "events": tot_n,
                  "identical": round(ident_n / max(tot_n, 1), 6),
                  "within_1_sample": round(within_n / max(tot_n, 1), 6), "amp_rel_err_max": amp_err}
        log(f"cpu baseline {cpu['value']}x realtime over {done} chunks; parity {parity}")
    sp.close()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
    dx.free()

    out = {
        "metric": METRIC,
        "value": round(rtf, 3),
        "unit": "x realtime",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator: 384 ch, 512 Gaussian-footprint templates, banded whitening)",
        "config": {"workload": "plain FP32 matcher, 384 channels, 512 templates, 3000-sample chunks "
                               "+ 61-sample margins (W=3122), 60 s recording resident in HBM",
                   "channels": C, "templates": T, "chunk": N, "batch_columns": W, "semantics": "kilosort",
                   "max_iter": info.max_iter},
        "p50_ms": round(pct(lat, 50), 4),
        "p99_ms": round(pct(lat, 99), 4),
        "p99_ms_host_input": round(pct(host_lat, 99), 4),
        "p50_ms_host_input": round(pct(host_lat, 50), 4),
        "mean_iterations": round(I, 2),
        "mean_spikes_per_chunk": round(S, 2),
        "stages": stages,
        "with_preprocessing": pre,
        **({"inflight": inflight} if inflight else {}),
        "peel_phases_us": peel_phases,
        "peel_geometry": {"split": int(q_last.peel_split), "rows_per_thread": int(q_last.peel_rows),
                          "workgroups": int(q_last.peel_groups), "autotuned": bool(tuned)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "parity": parity,
    }
    return out


if __name__ == "__main__":
    sys.exit(main())
