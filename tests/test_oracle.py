"""Pins the CPU oracle (CPU only).

Each check is chosen so a plausible slip in the oracle (wrong pad side,
flipped correlation, transposed operand, dropped term, wrong tie rule) fails
one of them: hand-computed known answers on the reference's own literal
inputs, closed forms of the generative model, algebraic invariants, and
brute-force loops on tiny inputs.
"""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from livespikesorter_amd import synth

GOLD = Path(__file__).parent / "golden"


def literal():
    return json.loads((GOLD / "conv1d_literal.json").read_text())


# ---------------------------------------------------------------------------
# conv1d_test.py literal (src/Python/conv1d_test.py:14-26)
# ---------------------------------------------------------------------------

def test_literal_known_answers():
    """Values worked by hand from the literal: 'same' pad 1, cross-correlation."""
    L = literal()
    x = np.array(L["batch"], np.int16).T                   # [W][C]
    wPCA = np.array(L["wPCA"], np.float32)
    U = np.array(L["Wall3"], np.float32)                   # [T][K][C]
    Xw, Bpc, B = oracle.stage_conv_project(x, np.eye(2, dtype=np.float32), wPCA, U)
    assert np.array_equal(Xw, np.array(L["batch"], np.float64))
    # Bpc[c=0][k=0]: w=0 -> 0*1 + 1*2 + 2*3 = 8; w=1 -> 1 + 4 + 12 = 17; w=10 -> 56 + 2*67 = 190
    assert Bpc[0, 0, 0] == 8 and Bpc[0, 0, 1] == 17 and Bpc[0, 0, 10] == 190
    # Bpc[c=1][k=0][0] = 321*2 + 12*3 = 678 ; rows k=1,2 of wPCA are 2x and 3x row 0
    assert Bpc[1, 0, 0] == 678
    assert np.array_equal(Bpc[:, 1], 2 * Bpc[:, 0]) and np.array_equal(Bpc[:, 2], 3 * Bpc[:, 0])
    # B[0][0] = 1*8 + 2*678 + 2*(3*8 + 4*678) + 3*(5*8 + 6*678) = 19160
    # B[1][1] = (7*17 + 8*447) + 2*(9*17 + 10*447) + 3*(11*17 + 12*447) = 29594
    assert B[0, 0] == 19160 and B[1, 1] == 29594


def test_literal_brute_force_stages():
    L = literal()
    X = L["batch"]
    x = np.array(X, np.int16).T
    wPCA = np.array(L["wPCA"], np.float32)
    U = np.array(L["Wall3"], np.float32)
    _, Bpc, B = oracle.stage_conv_project(x, np.eye(2, dtype=np.float32), wPCA, U)
    ref_bpc = oracle.lit_conv1d_same(X, L["wPCA"])
    assert np.array_equal(Bpc, np.array(ref_bpc))
    ref_B = oracle.lit_project(L["Wall3"], ref_bpc)
    assert np.array_equal(B, np.array(ref_B))
    # scores with the literal's norms (42, 45), M = 3
    nm = np.array(L["nm"], np.float64)
    M = L["M"]
    Cf, Cfmax, imax, Cmax = oracle.lit_scores(ref_B, list(nm), M)
    # the oracle peel with an unreachable threshold only computes the first Cfmax/imax
    K3 = np.zeros((3, 3), np.float32)
    K3[:, 1] = 1.0                                          # any M=3 PCs; unused without spikes
    ctc = np.zeros((2, 2, 2 * M + 1), np.float32)
    _, _, _, _, n_iter, cf, im, _ = oracle.peel(B, nm, ctc, U, K3, Th=1e9)
    assert n_iter == 0
    assert np.allclose(cf, Cfmax, rtol=1e-15, atol=0)
    assert list(im) == imax
    # edge columns are zero and the template max of column 3 is template 1 (B grows with t)
    assert all(cf[w] == 0 for w in (0, 1, 2, 8, 9, 10))
    assert Cmax[5] == max(Cfmax)


# ---------------------------------------------------------------------------
# template-derived tables
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def small_model():
    return synth.make_model(C=12, T=5, seed=7, whitening="identity", sigma=1.5)


def test_pcs_orthonormal(small_model):
    W = small_model.wPCA.astype(np.float64)
    assert np.allclose(W @ W.T, np.eye(W.shape[0]), atol=1e-6)


def test_wtw_brute_force(small_model):
    """WtW is flip(conv1d(W, W, padding=M)) (KS/template_matching.py:113-114)."""
    W = small_model.wPCA.astype(np.float64)
    K, M = W.shape
    ref = np.zeros((K, K, 2 * M + 1))
    for k1 in range(K):
        for k2 in range(K):
            pad = np.concatenate([np.zeros(M), W[k1], np.zeros(M)])
            conv = np.array([np.dot(pad[l:l + M], W[k2]) for l in range(2 * M + 1)])
            ref[k1, k2] = conv[::-1]
    assert np.allclose(synth.wtw(W), ref, atol=1e-12)


def test_ctc_closed_form_and_invariants(small_model):
    m = small_model
    U = m.U.astype(np.float64)
    M = m.M
    ctc = synth.ctc_table(m.U, m.wPCA).astype(np.float64)
    WtW = synth.wtw(m.wPCA)
    utu = np.einsum("ikl,jml->ijkm", U, U)
    ref = np.einsum("ijkm,kml->ijl", utu, WtW)
    assert np.allclose(ctc, ref, rtol=1e-5, atol=1e-6)
    nm = (U ** 2).sum((1, 2))
    # orthonormal PCs: zero-lag self overlap is the template norm
    assert np.allclose(ctc[np.arange(m.T), np.arange(m.T), M], nm, rtol=1e-5)
    # symmetry ctc[t][s][l] = ctc[s][t][2M - l]
    assert np.allclose(ctc, ctc.transpose(1, 0, 2)[:, :, ::-1], rtol=1e-5, atol=1e-6)
    # ctc is what subtracting a spike does to B: B of a noiseless waveform of s placed at w0
    s, a, W, w0 = 2, 3.0, 4 * M, 2 * M
    Xw = np.zeros((m.C, W))
    Xw[:, w0 - M // 2:w0 + M // 2 + 1] = a * m.waves[s].astype(np.float64).T
    Wp = m.wPCA.astype(np.float64)
    Bpc = np.stack([[np.convolve(Xw[c], Wp[k][::-1], mode="same") for k in range(m.K)]
                    for c in range(m.C)])
    B = np.einsum("tkc,ckw->tw", U, Bpc)
    assert np.allclose(B[:, w0 - M:w0 + M + 1], a * ctc[:, s, :], rtol=1e-4, atol=1e-4)


def test_extract_tables_definition(small_model):
    m = small_model
    xc = np.zeros(m.C)
    yc = 20.0 * np.arange(m.C)
    iCC, iU, Ucc = synth.extract_tables(m.U, xc, yc, 4)
    for c in range(m.C):
        d = np.abs(np.arange(m.C) - c)
        order = sorted(range(m.C), key=lambda r: (d[r], r))
        assert list(iCC[:, c]) == order[:4]
    energy = (m.U.astype(np.float64) ** 2).sum(1)
    assert np.array_equal(iU, energy.argmax(1))
    for t in range(m.T):
        for j in range(4):
            assert np.array_equal(Ucc[j, t], m.U[t, :, iCC[j, iU[t]]])


# ---------------------------------------------------------------------------
# the peel loop
# ---------------------------------------------------------------------------

def test_peel_recovers_isolated_noiseless_spikes(small_model):
    """B built from ctc (the exact effect of noiseless spikes) is peeled back to zero."""
    m = small_model
    M, T = m.M, m.T
    ctc = synth.ctc_table(m.U, m.wPCA)
    nm = (m.U.astype(np.float64) ** 2).sum((1, 2))
    W = 12 * M
    spikes = [(2 * M + 3, 1, 9.0), (5 * M, 3, 14.0), (8 * M + 7, 0, 11.0)]
    B = np.zeros((T, W))
    for w0, s, a in spikes:
        B[:, w0 - M:w0 + M + 1] += a * ctc[:, s, :].astype(np.float64)
    w, t, amp, it, n_iter, _, _, B_after = oracle.peel(B, nm, ctc, m.U, m.wPCA, Th=2.0)
    assert [(int(a), int(b)) for a, b in zip(w, t)] == [(s[0], s[1]) for s in spikes]
    assert np.allclose(amp, [s[2] for s in spikes], rtol=1e-6)
    assert n_iter == 1
    assert np.abs(B_after).max() < 1e-3


def test_peel_two_overlapping_spikes_resolve_over_iterations(small_model):
    m = small_model
    M, T = m.M, m.T
    ctc = synth.ctc_table(m.U, m.wPCA)
    nm = (m.U.astype(np.float64) ** 2).sum((1, 2))
    W = 8 * M
    big, small = (3 * M, 2, 20.0), (3 * M + 20, 4, 6.0)
    B = np.zeros((T, W))
    for w0, s, a in (big, small):
        B[:, w0 - M:w0 + M + 1] += a * ctc[:, s, :].astype(np.float64)
    w, t, amp, it, n_iter, _, _, _ = oracle.peel(B, nm, ctc, m.U, m.wPCA, Th=1.0)
    got = sorted(zip(w.tolist(), t.tolist()))
    assert (big[0], big[1]) in got and (small[0], small[1]) in got
    assert n_iter >= 2


def test_peel_zero_input_zero_events(small_model):
    m = small_model
    ctc = synth.ctc_table(m.U, m.wPCA)
    nm = (m.U.astype(np.float64) ** 2).sum((1, 2))
    w, *_ = oracle.peel(np.zeros((m.T, 500)), nm, ctc, m.U, m.wPCA)
    assert len(w) == 0


def brute_run_matching(Xw, U, wPCA, ctc, Th, max_iter=100):
    """run_matching as a direct numpy loop (KS/template_matching.py:141-198)."""
    U = U.astype(np.float64)
    W_ = wPCA.astype(np.float64)
    K, M = W_.shape
    nm = (U ** 2).sum(-1).sum(-1)
    Bpc = np.stack([[np.convolve(Xw[c], W_[k][::-1], mode="same") for k in range(K)]
                    for c in range(Xw.shape[0])])
    B = np.einsum("ijk,kjl->il", U, Bpc)
    Xres = Xw.copy()
    wave = np.einsum("tkc,km->tcm", U, W_)
    out = []
    for it in range(max_iter):
        Cf = np.maximum(B, 0) ** 2 / nm[:, None]
        Cf[:, :M] = 0
        Cf[:, -M:] = 0
        Cfmax = Cf.max(0)
        imax = Cf.argmax(0)
        Wd = B.shape[1]
        Cmax = np.array([Cfmax[max(0, w - M):w + M + 1].max() for w in range(Wd)])
        xs = np.nonzero((Cmax > Th ** 2) & (np.abs(Cmax - Cfmax) < 1e-9))[0]
        if len(xs) == 0:
            break
        iY = imax[xs]
        amps = B[iY, xs] / nm[iY]
        out += [(int(a), int(b), float(c), it) for a, b, c in zip(xs, iY, amps)]
        for j in range(2):
            for w0, s, a in zip(xs[j::2], iY[j::2], amps[j::2]):
                Xres[:, w0 - M // 2:w0 + M // 2 + 1] -= a * wave[s]
                B[:, w0 - M:w0 + M + 1] -= a * ctc[:, s, :]
    return out, Xres


def test_run_batch_matches_brute_force_numpy():
    m = synth.make_model(C=10, T=6, seed=11, whitening="identity", sigma=1.5, rate=60.0, amp=(9, 2))
    x, _ = synth.make_recording(m, 900, seed=12)
    ctc = synth.ctc_table(m.U, m.wPCA)
    xc, yc = np.zeros(m.C), 20.0 * np.arange(m.C)
    iCC, iU, Ucc = synth.extract_tables(m.U, xc, yc, 4)
    r = oracle.run_batch(x, m.whiten, m.wPCA, m.U, ctc, iCC, iU, Th=6.0, nC=4)
    Xw = m.whiten.astype(np.float64) @ x.T.astype(np.float64)
    ref, Xres = brute_run_matching(Xw, m.U, m.wPCA, ctc.astype(np.float64), 6.0)
    assert len(ref) > 5
    ref_sorted = sorted(ref, key=lambda e: (e[0], e[3]))
    assert [(e[0], e[1]) for e in ref_sorted] == list(zip(r.w.tolist(), r.t.tolist()))
    assert np.allclose(r.amp, [e[2] for e in ref_sorted], rtol=1e-9)
    # features on the final residual (KS/template_matching.py:50-51)
    half = m.M // 2
    for i, (w0, t) in enumerate(zip(r.w, r.t)):
        ch = iCC[:, iU[t]]
        seg = Xres[ch][:, w0 - half:w0 + half + 1]
        f = seg @ m.wPCA.astype(np.float64).T + r.amp[i] * Ucc[:, t].astype(np.float64)
        assert np.allclose(r.feat[i], f, rtol=1e-9, atol=1e-9)


def test_fp32_build_tracks_fp64():
    m = synth.make_model(C=16, T=8, seed=3, whitening="identity", rate=20.0)
    x, _ = synth.make_recording(m, 3122, seed=4)
    ctc = synth.ctc_table(m.U, m.wPCA)
    xc, yc = np.zeros(m.C), 20.0 * np.arange(m.C)
    iCC, iU, _ = synth.extract_tables(m.U, xc, yc, 10)
    r64 = oracle.run_batch(x, m.whiten, m.wPCA, m.U, ctc, iCC, iU)
    r32 = oracle.run_batch(x, m.whiten, m.wPCA, m.U, ctc, iCC, iU, bits=32)
    assert np.array_equal(r64.w, r32.w) and np.array_equal(r64.t, r32.t)
    assert np.max(np.abs(r32.amp - r64.amp) / np.abs(r64.amp)) < 1e-4


def test_ground_truth_recall_cfg1_sized():
    """cfg1 model (64 ch, 32 templates, 1 s): every planted spike is found at its sample."""
    m = synth.make_model(**synth.CONFIGS["cfg1"])
    x, truth = synth.make_recording(m, 30000 + 122)
    ctc = synth.ctc_table(m.U, m.wPCA)
    xc, yc = np.zeros(m.C), 20.0 * np.arange(m.C)
    iCC, iU, _ = synth.extract_tables(m.U, xc, yc, 10)
    r = oracle.run_batch(x, m.whiten, m.wPCA, m.U, ctc, iCC, iU, features=False)
    found = set(zip(r.w.tolist(), r.t.tolist()))
    inside = (truth.centre >= m.M) & (truth.centre < len(x) - m.M)
    hits = sum((int(c), int(t)) in found for c, t in zip(truth.centre[inside], truth.template[inside]))
    assert hits >= 0.99 * inside.sum()


def test_lss_semantics_switch():
    """LSS rules accept >= Th^2 and per-template near ties; Kilosort rules do not."""
    m = synth.make_model(C=12, T=5, seed=7, whitening="identity", sigma=1.5)
    M, T = m.M, m.T
    ctc = synth.ctc_table(m.U, m.wPCA)
    nm = (m.U.astype(np.float64) ** 2).sum((1, 2))
    W = 6 * M
    B = np.zeros((T, W))
    w0 = 3 * M
    # two templates with identical score at w0: Cf = Th^2 exactly for both
    for t in (1, 3):
        B[t, w0] = np.sqrt(64.0 * nm[t])
    w, t, *_ = oracle.peel(B, nm, ctc, m.U, m.wPCA, Th=8.0, max_iter=1)
    assert len(w) == 0                                   # Kilosort: Cmax > Th^2 is strict
    w, t, *_ = oracle.peel(B, nm, ctc, m.U, m.wPCA, Th=8.0, max_iter=1, semantics=oracle.SEM_LSS)
    assert sorted(t.tolist()) == [1, 3] and set(w.tolist()) == {w0}


def test_spike_positions_closed_forms():
    """Kilosort's spike position: mass on one channel gives that channel's
    coordinates; equal masses give the midpoint; no mass gives the origin
    (compute_spike_positions_kernel sets weights to 0 when the total is 0);
    a random case equals a loop over the definition."""
    iCC = np.array([[0, 1, 2], [1, 2, 1], [2, 0, 0]])        # [nC=3][C=3]
    iU = np.array([0, 1, 2])
    xc = np.array([0.0, 10.0, 30.0])
    yc = np.array([5.0, 25.0, 45.0])
    f = np.zeros((4, 3, 2))
    f[0, 0, :] = [3.0, 4.0]                                     # template 0: all mass on channel 0
    f[1, 0, 1] = 2.0; f[1, 1, 0] = 2.0                          # template 1: channels 1 and 2, equal
    x, y = oracle.spike_positions(f, np.array([0, 1, 2, 0]), iCC, iU, xc, yc)
    assert (x[0], y[0]) == (0.0, 5.0)
    assert (x[1], y[1]) == (20.0, 35.0)
    assert (x[2], y[2]) == (0.0, 0.0) and (x[3], y[3]) == (0.0, 0.0)
    rng = np.random.default_rng(3)
    f = rng.normal(size=(5, 3, 2))
    t = rng.integers(0, 3, 5)
    x, y = oracle.spike_positions(f, t, iCC, iU, xc, yc)
    for i in range(5):
        m = [(f[i, j] ** 2).sum() for j in range(3)]
        s = sum(m)
        ch = [iCC[j, iU[t[i]]] for j in range(3)]
        assert np.isclose(x[i], sum(xc[ch[j]] * m[j] / s for j in range(3)))
        assert np.isclose(y[i], sum(yc[ch[j]] * m[j] / s for j in range(3)))


def test_closest_cluster_brute_force_and_ties():
    cent = np.array([[0.0, 0.0], [10.0, 0.0], [0.0, 10.0], [10.0, 0.0]])
    x = np.array([1.0, 9.0, 5.0, 0.0])
    y = np.array([1.0, 1.0, 0.0, 9.0])
    got = oracle.closest_cluster(x, y, cent)
    assert got.tolist() == [0, 1, 0, 2]                          # (5, 0): equidistant 0 and 1 -> first
    rng = np.random.default_rng(4)
    cent = rng.normal(size=(40, 2)) * 50
    x, y = rng.normal(size=20) * 50, rng.normal(size=20) * 50
    got = oracle.closest_cluster(x, y, cent)
    for i in range(20):
        d = [((x[i] - cx) ** 2 + (y[i] - cy) ** 2) ** 0.5 for cx, cy in cent]
        assert got[i] == min(range(40), key=lambda k: (d[k], k))


def test_peel_decision_margins_known_ties():
    """The oracle's decision margins on a hand-built score matrix (no
    subtraction: zero ctc, one iteration).  Column 150 sits exactly on the
    threshold (Cf = 64 = Th^2), column 400 has two templates with equal scores,
    columns 700 and 710 are equal peaks of one pooling window, column 900 is a
    clean spike (Cf 144, nothing else near)."""
    T, K, C, M, W = 2, 6, 4, 61, 1000
    B = np.zeros((T, W))
    B[0, 150] = 8.0
    B[0, 400] = B[1, 400] = 10.0
    B[0, 700] = B[0, 710] = 9.0
    B[1, 900] = 12.0
    B[1, 260] = 5.0
    nm = np.ones(T)
    ctc = np.zeros((T, T, 2 * M + 1), np.float32)
    U = np.zeros((T, K, C), np.float32)
    wPCA = np.zeros((K, M), np.float32)
    w, t, amp, it, n_iter, cf, im, _, mg = oracle.peel(B, nm, ctc, U, wPCA, max_iter=1, margins=True)
    assert mg.shape == (3, W)
    assert mg[0, 150] == 0.0                              # threshold: |64 - 64| / 64
    assert mg[1, 400] == 0.0                              # template: (100 - 100) / 100
    assert mg[2, 700] == 0.0 and mg[2, 710] == 0.0        # pooling: equal peaks
    assert mg[0, 900] == pytest.approx((144 - 64) / 64)
    assert mg[1, 900] == 1.0 and mg[2, 900] == 1.0        # runner-up template and column at 0
    # a window far below threshold (Cf 25 at 260) counts no template or pooling tie,
    # and its column is no threshold decision either once Cf < Cmax
    assert np.isinf(mg[1, 260]) and np.isinf(mg[2, 260])
    assert np.isinf(mg[0, 240]) and mg[0, 260] == pytest.approx((64 - 25) / 64)
    # Kilosort rules: 150 is not accepted (> Th^2 needed); the tie at 400 goes to template 0
    got = sorted(zip(w.tolist(), t.tolist()))
    assert (150, 0) not in got and (400, 0) in got and (900, 1) in got
    assert (700, 0) in got and (710, 0) in got            # equal pooled maxima: both accepted


def test_explain_mismatches_gate_miss():
    """A spike the gated run lacks while its ungated run has it is a gate miss;
    a difference near it is a cascade; one far from both is unexplained."""
    from types import SimpleNamespace
    from tests.parity import explain_mismatches
    from livespikesorter_amd._capi import EVENT_DTYPE
    W, M, nt0min = 1000, 61, 20
    shift = -(M // 2) + nt0min
    ref = SimpleNamespace(w=np.array([100, 300, 800]), t=np.array([1, 2, 3]), margins=np.ones((3, W)))

    def evs(pairs):
        e = np.zeros(len(pairs), EVENT_DTYPE)
        e["sample"] = [c + shift for c, _ in pairs]
        e["template"] = [t for _, t in pairs]
        return e
    gated = evs([(100, 1), (350, 4)])                     # misses 300 and 800, extra at 350
    ungated = evs([(100, 1), (300, 2), (350, 4)])         # has 300 (the gate's miss), not 800
    ex = explain_mismatches(ref, gated, M=M, nt0min=nt0min, ungated=ungated)
    assert ex["events"] == 3 and ex["gate_miss"] == 1 and ex["cascade"] == 1 and ex["unexplained"] == 1
    assert ex["unexplained_events"][0]["column"] == 800
    plain = explain_mismatches(ref, gated, M=M, nt0min=nt0min)
    assert "gate_miss" not in plain and plain["unexplained"] == 3
