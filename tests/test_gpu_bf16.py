"""BF16 projection with FP32 rescoring (precision LSS_PREC_BF16_RESCORE).

Model: 384 channels, 1024 templates, bursty gamma(0.5) ISIs, banded whitening.
The projection runs in BF16 on MFMA with FP32 accumulation; every column whose
best score reaches 0.9 Th^2 has that score recomputed in FP32 before pooling, so
near-threshold decisions, the order of overlapping spikes and the amplitudes
follow FP32 scores.

Bars: spike agreement (+-1 sample, per template, greedy as the reference's
match_train, src/Python/template_tradeoff.py:88-113) >= 99% against the plain
FP32 GPU matcher and against the FP64 oracle; amplitudes of identically placed
spikes within 1e-3 relative of FP64 (the plain matcher's bar).  The projection stage itself is checked
against the BF16 error bound: |B_bf16 - B64| <= 2^-7 * sum |U| |Bpc| (two BF16
operand roundings of 2^-9 each, FP32 accumulation far below).
"""
import numpy as np
import pytest

import oracle
from livespikesorter_amd import synth
from livespikesorter_amd.matcher import PREC_BF16_RESCORE, PREC_FP32, STAGE_B, STAGE_BPC, OnlineSpikesV2
from tests.parity import agreement

pytestmark = pytest.mark.gpu

W = 12122          # 12000-sample batches with 61-sample margins


@pytest.fixture(scope="module")
def cfg3(model_dirs):
    m, d = model_dirs("cfg3")
    x, _ = synth.make_recording(m, 2 * 12000 + 122, seed=synth.TEST_SEED)
    return m, d, oracle.load_tables(d), x


def _agree(ref_s, ref_t, ev):
    ident, within, pairs = agreement(ref_s, ref_t, ev["sample"], ev["template"], tol=1)
    return ident, within, pairs


def _amp_err(ref_s, ref_t, ref_a, ev):
    """Relative amplitude error over spikes placed identically (sample and template)."""
    _, _, pairs = agreement(ref_s, ref_t, ev["sample"], ev["template"], tol=1)
    ri = np.array([p[0] for p in pairs])
    gi = np.array([p[1] for p in pairs])
    ex = ref_s[ri] == ev["sample"][gi]
    ri, gi = ri[ex], gi[ex]
    return np.abs(ev["amplitude"][gi] - ref_a[ri]) / np.abs(ref_a[ri]), ri, gi


def test_bf16_projection_error_bound(cfg3):
    m, d, tab, x = cfg3
    batch = x[:3122]
    ref = oracle.run_model_batch(tab, batch, stages=True, features=False)
    with OnlineSpikesV2(d, precision=PREC_BF16_RESCORE, keep_stages=True, max_batch=3122) as sp:
        sp.process_batch(batch)
        B = sp.debug_stage(STAGE_B).astype(np.float64)
        Bpc = sp.debug_stage(STAGE_BPC).astype(np.float64)      # [K][C][W]
    U = tab["Wall3"].astype(np.float64)                          # [T][K][C]
    T, K, C = U.shape
    absB = np.abs(U).reshape(T, K * C) @ np.abs(Bpc).reshape(K * C, -1)
    err = np.abs(B - ref.B)
    bound = 2.0 ** -7 * absB + 1e-6 * np.abs(ref.B).max()
    print(f"BF16 projection: max |err| {err.max():.3e}, max |err|/bound {(err / bound).max():.3f}, "
          f"median rel err {np.median(err / np.maximum(np.abs(ref.B), 1e-3)):.2e}")
    assert (err <= bound).all()


def test_bf16_rescore_agreement(cfg3):
    m, d, tab, x = cfg3
    M = m.M
    ref_s, ref_t, ref_a = [], [], []
    bf, fp = [], []
    with OnlineSpikesV2(d, precision=PREC_BF16_RESCORE, max_batch=W) as sb, \
            OnlineSpikesV2(d, precision=PREC_FP32, max_batch=W) as sf:
        for j in range(2):
            batch = x[j * 12000:j * 12000 + W]
            r = oracle.run_model_batch(tab, batch, features=False)
            ref_s.append(r.w + j * 12000 - M // 2 + m.nt0min)
            ref_t.append(r.t)
            ref_a.append(r.amp)
            bf.append(sb.process_batch(batch, first_sample=j * 12000)[0])
            fp.append(sf.process_batch(batch, first_sample=j * 12000)[0])
        info = sb.query()
    ref_s, ref_t, ref_a = map(np.concatenate, (ref_s, ref_t, ref_a))
    bf, fp = np.concatenate(bf), np.concatenate(fp)
    i_fp, w_fp, _ = _agree(fp["sample"], fp["template"], bf)
    i_64, w_64, pairs = _agree(ref_s, ref_t, bf)
    i_f64, w_f64, _ = _agree(ref_s, ref_t, fp)
    amp_b, ri, gi = _amp_err(ref_s, ref_t, ref_a, bf)
    amp_f, rf, gf = _amp_err(ref_s, ref_t, ref_a, fp)
    print(f"spikes: FP64 {len(ref_s)}, FP32 GPU {len(fp)}, BF16 GPU {len(bf)}")
    print(f"BF16 vs FP32 GPU: identical {i_fp:.5f} within1 {w_fp:.5f}")
    print(f"BF16 vs FP64:     identical {i_64:.5f} within1 {w_64:.5f}")
    print(f"FP32 vs FP64:     identical {i_f64:.5f} within1 {w_f64:.5f}")
    for name, e in (("BF16", amp_b), ("FP32", amp_f)):
        print(f"{name} amplitude rel err vs FP64: median {np.median(e):.2e} p99 {np.percentile(e, 99):.2e} "
              f"max {e.max():.2e}, within 1e-3 {np.mean(e <= 1e-3):.5f}")
    print(f"rescored {info.rescored}, decisions flipped {info.rescore_flips}")
    assert info.precision == PREC_BF16_RESCORE
    assert info.rescored >= len(bf)          # every accepted spike passed the rescore
    assert w_fp >= 0.99 and w_64 >= 0.99
    assert amp_f.max() <= 1e-3 and amp_b.max() <= 1e-3


def test_bf16_deterministic(cfg3):
    m, d, tab, x = cfg3
    batch = x[:W]
    with OnlineSpikesV2(d, precision=PREC_BF16_RESCORE, max_batch=W) as sp:
        e1, f1 = sp.process_batch(batch)
        e2, f2 = sp.process_batch(batch)
    assert np.array_equal(e1, e2) and np.array_equal(f1, f2)


def test_bf16_requires_kilosort_semantics(model_dirs):
    from livespikesorter_amd.matcher import SEM_LSS, LssError
    m, d = model_dirs("cfg1")
    with pytest.raises(LssError):
        OnlineSpikesV2(d, precision=PREC_BF16_RESCORE, semantics=SEM_LSS)
