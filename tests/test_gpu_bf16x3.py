"""Split-BF16 whitening, PC conv and projection (precision LSS_PREC_BF16X3).

Each FP32 operand is carried as a BF16 pair x = hi + lo + e with
hi = bf16(x), lo = bf16(x - hi), |e| <= 2^-16 |x|, and each product as the
three BF16 MFMA products lo*hi + hi*lo + hi*hi with FP32 accumulation.  The
dropped lo*lo term and the two residuals e bound the product error by
3 * 2^-16 |U||Bpc| (plus 2^-24-order terms), so a score is within
3 * 2^-16 * sum|U||Bpc| of the exact dot product, plus FP32 accumulation
error; the peel then runs on those scores as on FP32 ones.  The whitening
GEMM runs the same way (an int16 sample is exactly a BF16 hi + lo pair, so its
bound is 2 * 2^-16 * sum|Wr||x|), and so does the PC conv, as a GEMM of the
whitened rows with a banded matrix of the PCA basis (3 * 2^-16 * sum|wPCA||Xw|).

Model: 384 channels, 1024 templates, bursty gamma(0.5) ISIs, banded whitening
(cfg3).  Bars: the whitening and projection stages within those bounds of the
FP64 oracle's, the PC conv within its bound of the FP64 conv of the same
whitened rows; events against the FP64 oracle at the plain matcher's bar
(tests/parity.py check_events: identical >= 99.5%, the rest +-1 sample and each
a near tie of the oracle's decision margins, amplitudes 1e-3, features 1e-3),
on 12000-sample batches (register-tile peel) and on one 60000-sample batch
with its margins (W = 60122, the global-tile peel); agreement with the plain
FP32 GPU matcher >= 99.9%.
"""
import numpy as np
import pytest

import oracle
from livespikesorter_amd import StreamGroup, synth
from livespikesorter_amd.matcher import PREC_BF16X3, PREC_FP32, STAGE_B, STAGE_BPC, STAGE_XW, OnlineSpikesV2
from tests.parity import agreement, check_events

pytestmark = pytest.mark.gpu

W = 12122          # 12000-sample batches with 61-sample margins
# 3 * 2^-16 from the split, with the FP32 accumulation of up to 3 * 2304 products
# (at most 2^-24 per addition, relative to the running sum of magnitudes) on top
X3_BOUND = 3 * 2.0 ** -16 + 3 * 2304 * 2.0 ** -24


@pytest.fixture(scope="module")
def cfg3(model_dirs):
    m, d = model_dirs("cfg3")
    x, _ = synth.make_recording(m, 2 * 12000 + 122, seed=synth.TEST_SEED)
    return m, d, oracle.load_tables(d), x


def test_x3_whitening_and_projection_error_bounds(cfg3):
    m, d, tab, x = cfg3
    batch = x[:3122]
    print("[x3] projection stage vs FP64, W=3122", flush=True)
    ref = oracle.run_model_batch(tab, batch, stages=True, features=False)
    out = {}
    for prec in (PREC_BF16X3, PREC_FP32):
        with OnlineSpikesV2(d, precision=prec, keep_stages=True, max_batch=3122) as sp:
            sp.process_batch(batch)
            out[prec] = (sp.debug_stage(STAGE_B).astype(np.float64), sp.debug_stage(STAGE_BPC).astype(np.float64),
                         sp.debug_stage(STAGE_XW).astype(np.float64))
    B, Bpc, Xw = out[PREC_BF16X3]
    # whitening: Wr split into hi + lo, each int16 sample exactly hi + lo, so a whitened
    # value is within 2 * 2^-16 * sum |Wr||x| of the exact one, plus FP32 accumulation
    Wr = tab["whiteningMat"].astype(np.float64)
    absX = np.abs(Wr) @ np.abs(batch.astype(np.float64).T)
    errw = np.abs(Xw - ref.Xw)
    boundw = (2 * 2.0 ** -16 + 3 * 384 * 2.0 ** -24) * absX + 1e-7 * np.abs(ref.Xw).max()
    print(f"split-BF16 whitening: max |err|/sum|Wr||x| {(errw / np.maximum(absX, 1e-30)).max():.2e}, "
          f"max |err|/bound {(errw / boundw).max():.3f}")
    assert (errw <= boundw).all()
    # PC conv: split-BF16 products of wPCA and the GPU's own whitened rows with FP32
    # accumulation of 3 * 80 products, so within (3 * 2^-16 + 240 * 2^-24) * sum |wPCA||Xw|
    # of the FP64 conv of the same rows (zero outside the batch)
    P = tab["wPCA"].astype(np.float64)                           # [K][M]
    Kp, Mp = P.shape
    hp = Mp // 2
    Wb = Xw.shape[1]
    Xpad = np.zeros((Xw.shape[0], Wb + 2 * hp))
    Xpad[:, hp:hp + Wb] = Xw
    conv = np.stack([sum(P[k, m] * Xpad[:, m:m + Wb] for m in range(Mp)) for k in range(Kp)])
    aconv = np.stack([sum(abs(P[k, m]) * np.abs(Xpad[:, m:m + Wb]) for m in range(Mp)) for k in range(Kp)])
    errc = np.abs(Bpc - conv)
    boundc = (3 * 2.0 ** -16 + 240 * 2.0 ** -24) * aconv + 1e-30
    print(f"split-BF16 PC conv: max |err|/sum|wPCA||Xw| {(errc / np.maximum(aconv, 1e-30)).max():.2e}, "
          f"max |err|/bound {(errc / boundc).max():.3f}")
    assert (errc <= boundc).all()
    U = tab["Wall3"].astype(np.float64)                          # [T][K][C]
    T, K, C = U.shape
    absB = np.abs(U).reshape(T, K * C) @ np.abs(Bpc).reshape(K * C, -1)
    err = np.abs(B - ref.B)
    err32 = np.abs(out[PREC_FP32][0] - ref.B)
    bound = X3_BOUND * absB + 1e-6 * np.abs(ref.B).max()
    scale = np.maximum(absB, 1e-30)
    print(f"split-BF16: max |err|/sum|U||Bpc| {(err / scale).max():.2e} (bound {X3_BOUND:.2e}), "
          f"max |err|/bound {(err / bound).max():.3f}; FP32: max |err|/sum|U||Bpc| {(err32 / scale).max():.2e}; "
          f"rms rel err split {np.sqrt(np.mean((err / scale) ** 2)):.2e} vs FP32 {np.sqrt(np.mean((err32 / scale) ** 2)):.2e}")
    assert (err <= bound).all()


def test_x3_events_against_oracle(cfg3):
    m, d, tab, x = cfg3
    M = m.M
    with OnlineSpikesV2(d, precision=PREC_BF16X3, max_batch=W) as sx, \
            OnlineSpikesV2(d, precision=PREC_FP32, max_batch=W) as sf:
        n_x = n_id = 0
        for j in range(2):
            batch = x[j * 12000:j * 12000 + W]
            print(f"[x3] batch {j}: FP64 oracle and GPU, W={W}", flush=True)
            ref = oracle.run_model_batch(tab, batch, margins=True)
            ev, feat = sx.process_batch(batch)
            ef, _ = sf.process_batch(batch)
            check_events(ref, ev, feat, M=M, nt0min=m.nt0min, label=f"split-BF16 W={W} batch {j}")
            ident, _, _ = agreement(ef["sample"], ef["template"], ev["sample"], ev["template"])
            n = max(len(ev), len(ef))
            n_x += n
            n_id += ident * n
        info = sx.query()
    print(f"split-BF16 vs FP32 GPU: identical {n_id / n_x:.5f} over {n_x} events")
    assert info.precision == PREC_BF16X3
    assert info.peel_rows > 0
    assert n_id / n_x >= 0.999


def test_x3_wide_batch_global_tile(model_dirs):
    m, d = model_dirs("cfg3")
    x, _ = synth.make_recording(m, 60122, seed=synth.TEST_SEED + 7)
    tab = oracle.load_tables(d)
    print("[x3 wide] FP64 oracle, W=60122", flush=True)
    ref = oracle.run_model_batch(tab, x, margins=True)
    print("[x3 wide] GPU split-BF16", flush=True)
    with OnlineSpikesV2(d, max_batch=len(x), precision=PREC_BF16X3, max_events=1 << 17) as sp:
        ev, feat = sp.process_batch(x, cap=1 << 17)
        info = sp.query()
    assert info.peel_rows == 0, "a 60122-column batch runs the global-tile peel"
    assert len(ref.w) > 5000
    check_events(ref, ev, feat, M=m.M, nt0min=m.nt0min, label="split-BF16 W=60122")


def test_x3_deterministic_and_group(cfg3):
    """Bitwise repeatable, and a group of split-BF16 streams gives each stream
    what a single split-BF16 handle gives."""
    m, d, tab, x = cfg3
    Wg = 1622
    with OnlineSpikesV2(d, precision=PREC_BF16X3, max_batch=Wg) as sp:
        singles = [sp.process_batch(x[k * 1500:k * 1500 + Wg], first_sample=k * 1500) for k in range(3)]
        again = sp.process_batch(x[:Wg])
    assert np.array_equal(singles[0][0], again[0]) and np.array_equal(singles[0][1], again[1])
    with StreamGroup([d] * 3, batch=Wg, precision=PREC_BF16X3) as g:
        out = g.process([x[k * 1500:k * 1500 + Wg] for k in range(3)], [k * 1500 for k in range(3)])
    for k in range(3):
        assert np.array_equal(out[k][0], singles[k][0]) and np.array_equal(out[k][1], singles[k][1])
