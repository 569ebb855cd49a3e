"""Proposal network on the GPU (lss_proposal_*) against its CPU oracle
(oracle/proposal.py), and the gate's effect on the matcher.

Forward: GPU P(spike) within 2e-3 of the oracle run with the same BF16
storage points (the remaining difference is FP32 accumulation order and rare
BF16 rounding flips); against the unrounded FP64 oracle the logits differ by
the BF16 storage error, stated in the output.
Gradients: relative L2 error per parameter block <= 3e-2 and cosine >= 0.999
against the oracle with BF16 storage (the GPU also stores the data gradients
of layers 1-3 in BF16).
"""
import numpy as np
import pytest

from oracle import proposal as P
from livespikesorter_amd import synth
from livespikesorter_amd.matcher import OnlineSpikesV2
from livespikesorter_amd.proposal import ProposalNet, n_params

pytestmark = pytest.mark.gpu


def _inputs(C=40, W=300, seed=3):
    rng = np.random.default_rng(seed)
    xw = rng.normal(size=(C, W)).astype(np.float32)
    xw[:, 100:110] += 4.0 * rng.normal(size=(C, 1)).astype(np.float32)   # a few large transients
    return xw


def test_layout_and_count():
    assert n_params() == P.N_PARAMS


def test_forward_matches_oracle():
    xw = _inputs()
    with ProposalNet(40, 512, seed=7) as net:
        flat = net.params
        p_gpu = net.forward(xw)
    params = P.unpack(flat)
    x0 = P.group_input(xw.astype(np.float64))
    z_bf, _ = P.forward(params, x0, bf16_storage=True)
    z64, _ = P.forward(params, x0)
    p_bf = 1 / (1 + np.exp(-z_bf))
    p64 = 1 / (1 + np.exp(-z64))
    err = np.abs(p_gpu - p_bf).max()
    print(f"P(spike): max |gpu - oracle(bf16 storage)| {err:.2e}, max |gpu - oracle(fp64)| "
          f"{np.abs(p_gpu - p64).max():.2e}, logit range {z64.min():.2f}..{z64.max():.2f}")
    assert p_gpu.shape == (3, 300)
    assert err <= 2e-3


def test_gradient_matches_oracle():
    xw = _inputs(seed=4)
    rng = np.random.default_rng(5)
    y = (rng.random((3, 300)) < 0.05).astype(np.uint8)
    with ProposalNet(40, 512, seed=8) as net:
        flat = net.params
        loss_gpu, g_gpu = net.grad(xw, y, pos_weight=8.0)
    params = P.unpack(flat)
    x0 = P.group_input(xw.astype(np.float64))
    loss_ref, grads = P.loss_and_grad(params, x0, y.astype(np.float64), 8.0, bf16_storage=True)
    g_ref = P.pack(grads)
    print(f"loss gpu {loss_gpu:.6f} oracle {loss_ref:.6f}")
    assert abs(loss_gpu - loss_ref) <= 1e-3 * abs(loss_ref)
    off = 0
    for name, shape in P.LAYOUT:
        n = int(np.prod(shape))
        a, b = g_gpu[off:off + n].astype(np.float64), g_ref[off:off + n]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        cos = float(a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30))
        print(f"  {name}: rel L2 err {rel:.2e}, cosine {cos:.6f}")
        assert rel <= 3e-2 and cos >= 0.999, name
        off += n


def test_training_reduces_loss(model_dirs):
    """A few Adam steps on matcher-whitened data with fixed labels lower the loss."""
    from livespikesorter_amd._hip import DeviceBuffer, synchronize
    m, d = model_dirs("cfg1")
    x, truth = synth.make_recording(m, 1622, seed=synth.TEST_SEED)
    dx = DeviceBuffer.from_array(x)
    synchronize()
    G = 4
    y = np.zeros((G, 1622), np.uint8)
    # truth centres as labels (group of the main channel)
    iu = np.load(d / "iU.npy").ravel()
    for c, t in zip(truth.centre, truth.template):
        if 1 <= c < 1621:
            y[iu[t] // 16, c - 1:c + 2] = 1
    with OnlineSpikesV2(d, max_batch=1622) as sp, ProposalNet(64, 1622, seed=9) as net:
        losses = [net.train_step(sp, dx.ptr, 1622, 64, y, 8.0, 3e-3) for _ in range(60)]
    dx.free()
    print(f"loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    assert losses[-1] < 0.7 * losses[0]


def _const_net(C, W, b4):
    net = ProposalNet(C, W, seed=1)
    p = net.params
    p[P.N_PARAMS - 1 - P.CH * P.KW:P.N_PARAMS - 1] = 0.0     # W4 = 0
    p[P.N_PARAMS - 1] = b4                                    # z = b4 everywhere
    net.params = p
    return net


def test_gate_extremes(model_dirs):
    """All tiles proposed, or the fallback everywhere: identical to plain
    scoring; nothing proposed: no spikes and no tile computed."""
    m, d = model_dirs("cfg2")
    x, _ = synth.make_recording(m, 3122, seed=synth.TEST_SEED)
    with OnlineSpikesV2(d) as plain:
        e0, f0 = plain.process_batch(x)
    for b4, expect in ((20.0, "all"), (0.0, "fallback"), (-20.0, "none")):
        with OnlineSpikesV2(d) as sp:
            net = _const_net(384, 3122, b4)
            sp.set_proposal(net)
            e, f = sp.process_batch(x)
            q = sp.query()
            print(f"b4 {b4}: tiles {q.gate_tiles_scored}/{q.gate_tiles}, fallback blocks "
                  f"{q.gate_fallback_blocks}/{q.gate_blocks}, spikes {len(e)}")
            if expect == "none":
                assert len(e) == 0 and q.gate_tiles_scored == 0
            else:
                assert np.array_equal(e, e0) and np.array_equal(f, f0)
                assert q.gate_tiles_scored == q.gate_tiles
                assert (q.gate_fallback_blocks == q.gate_blocks) == (expect == "fallback")
            sp.set_proposal(None)
            net.close()


def test_trained_gate_gain(tmp_path):
    """What the trained network gains over the method without it, on a test-seed
    recording of the 1024-template probe it was trained for (training used
    TRAIN_SEED recordings): the fraction of score tiles left uncomputed at >= 99%
    spike agreement with the plain FP32 matcher, against an untrained network
    (same gate) and no network (every tile).  Then the time gain of the gated
    split-BF16 matcher (the network gate with split-BF16 whitening, PC conv and
    projection) over the plain FP32 matcher, at the shape they are built for:
    8 streams of this probe in one group launch per tick (1500-sample chunks,
    W = 1622).  The device time of the stages before the peel (the peel is the
    same kernel in both) must be at least 5% shorter, with >= 99.9% of the
    spikes identical.  At this spike density the gate leaves few tiles
    uncomputed, so the gain comes from the split-BF16 stages; the network's own
    cost is inside the measured time."""
    import time
    from pathlib import Path
    from livespikesorter_amd._hip import DeviceBuffer, synchronize
    from livespikesorter_amd.matcher import PREC_BF16_RESCORE
    from livespikesorter_amd.train_proposal import _agreement
    weights = Path(__file__).resolve().parent.parent / "livespikesorter_amd" / "data" / "proposal_cfg5.npy"
    m = synth.make_model(**synth.CONFIGS["cfg5"])
    d = tmp_path / "cfg5"
    synth.write_oss_input(m, d)
    x, _ = synth.make_recording(m, 8 * 30000, seed=synth.TEST_SEED, workers=8)
    N, W, C = 1500, 1622, 384
    nb = (len(x) - W) // N + 1
    dx = DeviceBuffer.from_array(x)
    synchronize()

    def run(net, prec):
        with OnlineSpikesV2(d, chunk=N, max_batch=W, precision=prec) as sp:
            if net is not None:
                sp.set_proposal(net, 0.25, 0.75, 4)
            evs = []
            t0 = time.perf_counter()
            for j in range(nb):
                evs.append(sp.process_batch_device(dx.ptr + j * N * C * 2, W, C, first_sample=j * N)[0])
            el = time.perf_counter() - t0
            q = sp.query()
            sp.set_proposal(None)
        frac = q.gate_tiles_scored / q.gate_tiles if q.gate_tiles else 1.0
        return np.concatenate(evs), frac, nb * N / 30000 / el

    ref, _, rt_plain = run(None, 0)
    res = {}
    with ProposalNet(C, W, seed=1) as untrained, ProposalNet(C, W, seed=1) as trained:
        trained.load(weights)
        for name, net in (("none", None), ("untrained", untrained), ("trained", trained)):
            e, frac, rt = run(net, PREC_BF16_RESCORE)
            ident, within = _agreement(ref, e)
            res[name] = (ident, within, frac, rt)
            print(f"{name:9s}: agreement identical {ident:.5f} within1 {within:.5f}, tiles scored {frac:.4f}, "
                  f"{rt:.1f}x realtime (plain FP32 {rt_plain:.1f}x)")
    assert res["trained"][1] >= 0.99
    assert res["untrained"][2] == 1.0 and res["none"][2] == 1.0
    assert res["trained"][2] < 0.98

    from livespikesorter_amd import StreamGroup
    from livespikesorter_amd.matcher import PREC_BF16X3, PREC_FP32
    S, ticks, warm = 8, 12, 2
    off = [k * 27000 for k in range(S)]            # eight different windows of the recording
    assert off[-1] + ticks * N + W - N <= len(x)

    def group(prec, net):
        g = StreamGroup([d] * S, batch=W, precision=prec, profile=2)
        if net is not None:
            for k in range(S):
                g.set_proposal(k, net, 0.1, 0.5, 12)
        evs = [[] for _ in range(S)]
        for j in range(ticks):
            out = g.process([dx.ptr + (off[k] + j * N) * C * 2 for k in range(S)], [j * N] * S, device=True,
                            lds=[C] * S)
            if j == warm - 1:
                g.stage_times(0, reset=True)
            for k in range(S):
                evs[k].append(out[k][0])
        front, peel, n = g.stage_times(0)
        g.close()
        return [np.concatenate(e) for e in evs], front / n * 1e3, peel / n * 1e3

    e32, f32, p32 = group(PREC_FP32, None)
    with ProposalNet(C, W, seed=1) as trained:
        trained.load(weights)
        egx, fgx, pgx = group(PREC_BF16X3, trained)
    dx.free()
    ident = np.mean([_agreement(a, b)[0] for a, b in zip(e32, egx)])
    print(f"8-stream group: front stages plain FP32 {f32:.1f} us, gated split-BF16 {fgx:.1f} us; "
          f"peel {p32:.1f} / {pgx:.1f} us; spikes identical {ident:.5f} over {sum(len(e) for e in e32)}")
    assert ident >= 0.999
    assert fgx <= 0.95 * f32


def test_gate_at_low_spike_density(tmp_path):
    """The trained gate on recordings of the same probe at 3% of its firing rates
    (about 160 spikes/s per stream), 8 streams in one split-BF16 group launch per
    tick: most score tiles are left uncomputed (fewer than half scored) with >= 99%
    of the spikes identical to the same group without the gate.  The stages'
    device times with and without the gate are printed: the network costs about
    as much as the skipped tiles save, even here."""
    import dataclasses
    from pathlib import Path
    from livespikesorter_amd import StreamGroup
    from livespikesorter_amd._hip import DeviceBuffer, synchronize
    from livespikesorter_amd.matcher import PREC_BF16X3
    from livespikesorter_amd.train_proposal import _agreement
    weights = Path(__file__).resolve().parent.parent / "livespikesorter_amd" / "data" / "proposal_cfg5.npy"
    m = synth.make_model(**synth.CONFIGS["cfg5"])
    d = tmp_path / "cfg5"
    synth.write_oss_input(m, d)
    ms = dataclasses.replace(m, rates=m.rates * 0.03)
    N, W, C, S, ticks = 1500, 1622, 384, 8, 10
    recs = [synth.make_recording(ms, ticks * N + W - N, seed=synth.TEST_SEED + 300 + k, workers=8)[0]
            for k in range(S)]
    bufs = [DeviceBuffer.from_array(x) for x in recs]
    synchronize()

    def group(net):
        g = StreamGroup([d] * S, batch=W, precision=PREC_BF16X3, profile=2)
        if net is not None:
            for k in range(S):
                g.set_proposal(k, net, 0.1, 0.5, 12)
        evs = [[] for _ in range(S)]
        for j in range(ticks):
            out = g.process([bufs[k].ptr + j * N * C * 2 for k in range(S)], [j * N] * S, device=True, lds=[C] * S)
            if j == 1:
                g.stage_times(0, reset=True)
            for k in range(S):
                evs[k].append(out[k][0])
        front, peel, n = g.stage_times(0)
        qs = [g.query(k) for k in range(S)]
        g.close()
        frac = sum(q.gate_tiles_scored for q in qs) / max(1, sum(q.gate_tiles for q in qs))
        return [np.concatenate(e) for e in evs], front / n * 1e3, peel / n * 1e3, frac

    e0, f0, p0, _ = group(None)
    with ProposalNet(C, W, seed=1) as net:
        net.load(weights)
        e1, f1, p1, frac = group(net)
    for b in bufs:
        b.free()
    ident = np.mean([_agreement(a, b)[0] for a, b in zip(e0, e1)])
    print(f"3% firing rates: {sum(len(e) for e in e0)} spikes, tiles scored {frac:.3f}, identical {ident:.5f}; "
          f"front stages {f0:.1f} us ungated, {f1:.1f} us gated; peel {p0:.1f} / {p1:.1f} us")
    assert frac < 0.5
    assert ident >= 0.99
