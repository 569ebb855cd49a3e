"""Train the spike-proposal network on the GPU.

    python -m livespikesorter_amd.train_proposal --out livespikesorter_amd/data/proposal_cfg5.npy

Training data: a recording of the 384-channel, 1024-template probe model
(synth.CONFIGS["cfg5"], the model drawn from the test seed) whose spike
trains and noise are drawn from synth.TRAIN_SEED, 600 s by default, held in
HBM.  The plain FP32 matcher runs over it in framed 1500-sample chunks
(columns [j*1500, j*1500 + 1622)); each chunk's spike list, mapped to
(16-channel group of the template's main channel, batch column +-1), is that
chunk's label map.  The network then takes Adam steps (the library's own
forward/backward kernels) on the chunks in shuffled order; the matcher's own
whitening kernel produces its inputs.  The last 5% of the chunks are held out
and scored at the end: spike agreement of the gated BF16 matcher with the
plain FP32 matcher, and the fraction of score tiles the gate computed.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

from . import synth
from ._hip import DeviceBuffer, synchronize
from .matcher import PREC_BF16_RESCORE, OnlineSpikesV2
from .proposal import GROUP, ProposalNet, spike_labels

N, M = 1500, 61
W = N + 2 * M


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _agreement(ref, got, tol=1):
    """(identical, within tol) fractions over the longer list, per template greedy."""
    from collections import defaultdict
    by_t = defaultdict(list)
    for s in got:
        by_t[int(s["template"])].append(int(s["sample"]))
    used = {t: np.zeros(len(v), bool) for t, v in by_t.items()}
    arr = {t: np.array(v) for t, v in by_t.items()}
    exact = near = 0
    for pass_tol in (0, tol):
        for s in ref:
            t = int(s["template"])
            if t not in arr:
                continue
            d = np.abs(arr[t] - int(s["sample"]))
            d = np.where(used[t], 1 << 40, d)
            j = int(np.argmin(d))
            if d[j] <= pass_tol and not used[t][j]:
                used[t][j] = True
                if pass_tol == 0:
                    exact += 1
                else:
                    near += 1
    n = max(len(ref), len(got), 1)
    return exact / n, (exact + near) / n


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", required=True, help="weights file (.npy); a .json summary is written beside it")
    ap.add_argument("--seconds", type=float, default=600.0)
    ap.add_argument("--epochs", type=int, default=3)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--pos-weight", type=float, default=20.0)
    ap.add_argument("--workers", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--seed", type=int, default=1, help="weight initialisation and chunk order")
    ap.add_argument("--lo", type=float, default=0.25)
    ap.add_argument("--hi", type=float, default=0.75)
    ap.add_argument("--reach", type=int, default=4)
    args = ap.parse_args(argv)

    t_all = time.time()
    model = synth.make_model(**synth.CONFIGS["cfg5"])
    tmp = tempfile.TemporaryDirectory()
    d = Path(tmp.name) / "cfg5"
    synth.write_oss_input(model, d)
    t0 = time.time()
    x, _ = synth.make_recording(model, int(args.seconds * model.fs), seed=synth.TRAIN_SEED, workers=args.workers)
    C = x.shape[1]
    n_chunks = (len(x) - W) // N + 1
    log(f"training recording: {len(x)} samples x {C} ch from seed {synth.TRAIN_SEED} in {time.time() - t0:.1f}s, "
        f"{n_chunks} chunks")
    dx = DeviceBuffer.from_array(x)
    synchronize()
    G = (C + GROUP - 1) // GROUP
    iu = np.load(d / "iU.npy").ravel().astype(np.int64)
    tgroup = iu // GROUP

    # labels from the plain FP32 matcher
    t0 = time.time()
    labels = np.zeros((n_chunks, G, W), np.uint8)
    plain_events = []
    n_spikes = 0
    with OnlineSpikesV2(d, chunk=N, max_batch=W) as sp:
        for j in range(n_chunks):
            ev, _ = sp.process_batch_device(dx.ptr + j * N * C * 2, W, C, first_sample=j * N)
            labels[j] = spike_labels(ev, tgroup, j * N, W, G, M // 2, model.nt0min)
            plain_events.append(ev)
            n_spikes += len(ev)
    log(f"plain matcher labels: {n_spikes} spikes in {time.time() - t0:.1f}s, positive fraction "
        f"{labels.mean():.4f}")

    n_val = max(1, n_chunks // 20)
    train_ids = np.arange(n_chunks - n_val)
    val_ids = np.arange(n_chunks - n_val, n_chunks)
    rng = np.random.default_rng(args.seed)
    history = []
    t0 = time.time()
    with OnlineSpikesV2(d, chunk=N, max_batch=W) as sp, ProposalNet(C, W, seed=args.seed) as net:
        for ep in range(args.epochs):
            lr = args.lr * (0.3 ** ep)
            order = rng.permutation(train_ids)
            losses = [net.train_step(sp, dx.ptr + int(j) * N * C * 2, W, C, labels[j], args.pos_weight, lr)
                      for j in order]
            history.append({"epoch": ep, "lr": lr, "mean_loss": float(np.mean(losses)),
                            "last_loss": float(np.mean(losses[-200:]))})
            log(f"epoch {ep}: lr {lr:.2e} mean loss {history[-1]['mean_loss']:.5f} "
                f"(last 200 {history[-1]['last_loss']:.5f})")
        train_s = time.time() - t0
        params = net.params
        out = Path(args.out)
        out.parent.mkdir(parents=True, exist_ok=True)
        np.save(out, params)
        # held-out chunks: gated BF16 matcher against the plain FP32 matcher
        with OnlineSpikesV2(d, chunk=N, max_batch=W, precision=PREC_BF16_RESCORE) as gm:
            gm.set_proposal(net, args.lo, args.hi, args.reach)
            ref, got = [], []
            for j in val_ids:
                ev, _ = gm.process_batch_device(dx.ptr + int(j) * N * C * 2, W, C, first_sample=int(j) * N)
                got.append(ev)
                ref.append(plain_events[j])
            q = gm.query()
            gm.set_proposal(None)
    ref = np.concatenate(ref)
    got = np.concatenate(got)
    ident, within = _agreement(ref, got)
    summary = {
        "model": "synth cfg5: 384 channels, 1024 templates (test-seed probe model)",
        "training_recording": {"seed": synth.TRAIN_SEED, "seconds": args.seconds, "chunks": int(len(train_ids)),
                               "chunk": N, "batch_columns": W, "spikes": int(n_spikes)},
        "epochs": args.epochs, "lr": args.lr, "pos_weight": args.pos_weight, "history": history,
        "train_seconds": round(train_s, 1),
        "held_out": {"chunks": int(n_val), "plain_spikes": int(len(ref)), "gated_spikes": int(len(got)),
                     "identical": round(ident, 6), "within_1_sample": round(within, 6),
                     "tiles_scored_fraction": round(q.gate_tiles_scored / max(q.gate_tiles, 1), 4),
                     "fallback_block_fraction": round(q.gate_fallback_blocks / max(q.gate_blocks, 1), 4),
                     "gate": {"lo": args.lo, "hi": args.hi, "reach": args.reach}},
        "total_seconds": round(time.time() - t_all, 1),
    }
    out.with_suffix(".json").write_text(json.dumps(summary, indent=1))
    log(json.dumps(summary["held_out"]))
    dx.free()
    tmp.cleanup()
    print(json.dumps(summary), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
