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
    (same gate) and no network (every tile).  Timings are printed, not asserted."""
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
    dx.free()
    assert res["trained"][1] >= 0.99
    assert res["untrained"][2] == 1.0 and res["none"][2] == 1.0
    assert res["trained"][2] < 0.98
