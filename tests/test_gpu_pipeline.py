"""Pipelined group lanes (lss_group_set_pipelined): consecutive ticks of the same
8 streams in flight on two lanes at once, driven from two host threads, give
exactly the events and features of one lane taking the ticks in turn.

Model: the 384-channel, 1024-template probe the proposal network was trained
for, 8 streams of 1500-sample chunks (W = 1622), 12 ticks; the plain FP32
group and the gated split-BF16 group.  Lane 1 then works in buffers of its own
while lane 0 uses the handles', and every peel waits for the peel enqueued
before it, so the device never holds two peel grids; the front stages of one
tick run beside the other tick's peel.  Also: two lanes of a plain group
driven at once with disjoint halves of the streams active.
"""
import threading
from pathlib import Path

import numpy as np
import pytest

from livespikesorter_amd import StreamGroup, synth
from livespikesorter_amd._hip import DeviceBuffer, synchronize
from livespikesorter_amd.matcher import PREC_BF16X3, PREC_FP32
from livespikesorter_amd.proposal import ProposalNet

pytestmark = pytest.mark.gpu

N, W, C, S, TICKS = 1500, 1622, 384, 8, 12
WEIGHTS = Path(__file__).resolve().parent.parent / "livespikesorter_amd" / "data" / "proposal_cfg5.npy"


@pytest.fixture(scope="module")
def streams(tmp_path_factory):
    m = synth.make_model(**synth.CONFIGS["cfg5"])
    d = tmp_path_factory.mktemp("pipe") / "cfg5"
    synth.write_oss_input(m, d)
    recs = [synth.make_recording(m, TICKS * N + W - N, seed=synth.TEST_SEED + 400 + k, workers=8)[0]
            for k in range(S)]
    bufs = [DeviceBuffer.from_array(x) for x in recs]
    synchronize()
    yield d, bufs
    for b in bufs:
        b.free()


def _ptrs(bufs, j):
    return [bufs[k].ptr + j * N * C * 2 for k in range(S)]


@pytest.mark.parametrize("mode", ["fp32", "x3_gated"])
def test_pipelined_lanes_match_one_lane(streams, mode):
    d, bufs = streams
    prec = PREC_FP32 if mode == "fp32" else PREC_BF16X3
    net = None
    if mode == "x3_gated":
        net = ProposalNet(C, W, seed=1)
        net.load(WEIGHTS)
    try:
        with StreamGroup([d] * S, batch=W, precision=prec) as g:
            if net is not None:
                for k in range(S):
                    g.set_proposal(k, net, 0.1, 0.5, 12)
            print(f"[pipeline {mode}] one lane, {TICKS} ticks", flush=True)
            ref = [g.process(_ptrs(bufs, j), [j * N] * S, device=True, lds=[C] * S) for j in range(TICKS)]
            g.set_pipelined(True)
            got = [None] * TICKS
            errs = []

            def run(lane):
                try:
                    for j in range(lane, TICKS, 2):
                        got[j] = g.process(_ptrs(bufs, j), [j * N] * S, device=True, lds=[C] * S, lane=lane)
                except Exception as e:          # reported by the main thread
                    errs.append(e)

            print(f"[pipeline {mode}] two lanes, two threads", flush=True)
            th = [threading.Thread(target=run, args=(lane,)) for lane in (0, 1)]
            for t in th:
                t.start()
            for t in th:
                t.join()
            assert not errs, errs
            # host input is refused in a pipelined group
            with pytest.raises(Exception):
                g.process([np.zeros((W, C), np.int16)] * S, [0] * S)
            g.set_pipelined(False)
            again = g.process(_ptrs(bufs, 3), [3 * N] * S, device=True, lds=[C] * S)
    finally:
        if net is not None:
            net.close()
    n_ev = 0
    for j in range(TICKS):
        for k in range(S):
            assert np.array_equal(got[j][k][0], ref[j][k][0]), (mode, j, k)
            assert np.array_equal(got[j][k][1], ref[j][k][1]), (mode, j, k)
            n_ev += len(ref[j][k][0])
    for k in range(S):
        assert np.array_equal(again[k][0], ref[3][k][0])
    print(f"[pipeline {mode}] {n_ev} events identical over {TICKS} ticks x {S} streams")
    assert n_ev > 1000


def test_concurrent_lanes_disjoint_streams(streams):
    """Two lanes of a plain (not pipelined) group driven at once from two threads,
    each with half of the streams active and the others idle: the two launches'
    peel grids run side by side on the device, and each stream's events and
    features equal those of one lane taking every stream in turn (with every
    stream active, the launch that places each stream's workgroups on XCDs of
    its own)."""
    d, bufs = streams
    with StreamGroup([d] * S, batch=W, precision=PREC_FP32) as g:
        print(f"[lanes] one lane, every stream active, {TICKS} ticks", flush=True)
        ref = [g.process(_ptrs(bufs, j), [j * N] * S, device=True, lds=[C] * S) for j in range(TICKS)]
        got = [[None] * S for _ in range(TICKS)]
        errs = []

        def run(lane):
            mine = range(lane * S // 2, (lane + 1) * S // 2)
            try:
                for j in range(TICKS):
                    p = [q if k in mine else None for k, q in enumerate(_ptrs(bufs, j))]
                    out = g.process(p, [j * N] * S, device=True, lds=[C] * S, lane=lane)
                    for k in mine:
                        got[j][k] = out[k]
            except Exception as e:              # reported by the main thread
                errs.append(e)

        print("[lanes] two lanes, two threads, disjoint halves of the streams", flush=True)
        th = [threading.Thread(target=run, args=(lane,)) for lane in (0, 1)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        assert not errs, errs
    n_ev = 0
    for j in range(TICKS):
        for k in range(S):
            assert np.array_equal(got[j][k][0], ref[j][k][0]), (j, k)
            assert np.array_equal(got[j][k][1], ref[j][k][1]), (j, k)
            n_ev += len(ref[j][k][0])
    print(f"[lanes] {n_ev} events identical over {TICKS} ticks x {S} streams")
    assert n_ev > 1000
