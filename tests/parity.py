"""Spike-list comparison helpers shared by the parity tests.

``match_events`` pairs two spike lists the way the reference's
``match_train`` compares runs (src/Python/template_tradeoff.py:88-113): per
template, greedily in sample order, nearest unused spike within a tolerance.
Exact pairs are taken first; a pair one sample apart counts as near.
"""
from __future__ import annotations

import numpy as np


def match_events(ref_s, ref_t, got_s, got_t, tol: int = 1):
    """Return (pairs [(i_ref, i_got)], exact_count, near_count)."""
    ref_s = np.asarray(ref_s, np.int64)
    got_s = np.asarray(got_s, np.int64)
    ref_t = np.asarray(ref_t)
    got_t = np.asarray(got_t)
    pairs = []
    exact = near = 0
    big = np.iinfo(np.int64).max
    for t in np.intersect1d(np.unique(ref_t), np.unique(got_t)):
        ri = np.nonzero(ref_t == t)[0]
        gi = np.nonzero(got_t == t)[0]
        ri = ri[np.argsort(ref_s[ri], kind="stable")]
        gs = got_s[gi]
        used = np.zeros(len(gi), bool)
        taken = np.zeros(len(ri), bool)
        for pass_tol in (0, tol):
            for n, r in enumerate(ri):
                if taken[n]:
                    continue
                d = np.where(used, big, np.abs(gs - ref_s[r]))
                j = int(np.argmin(d))
                if d[j] <= pass_tol:
                    used[j] = True
                    taken[n] = True
                    pairs.append((int(r), int(gi[j])))
                    if d[j] == 0:
                        exact += 1
                    else:
                        near += 1
    return pairs, exact, near


def agreement(ref_s, ref_t, got_s, got_t, tol: int = 1):
    """(identical fraction, within-tol fraction, pairs); fractions over the longer list."""
    pairs, exact, near = match_events(ref_s, ref_t, got_s, got_t, tol)
    n = max(len(ref_s), len(got_s), 1)
    return exact / n, (exact + near) / n, pairs


AMP_RTOL = 1e-3
FEAT_RTOL = 1e-3
# Relative score error of an FP32 run against the FP64 oracle that a decision
# margin must lie under for a spike-list difference to count as a near tie:
# a score is an FP32 dot product of n = K*C <= 2304 terms (worst case about
# n * 2^-24 = 1.4e-4 of sum |U||Bpc|, itself a few times |B| at a peak) followed
# by the peel's FP32 subtractions; Cf = relu(B)^2 / nm doubles the relative error.
TIE_RTOL = 1e-3


def explain_mismatches(ref, ev, *, first_sample: int = 0, M: int = 61, nt0min: int = 20,
                       tol: float = TIE_RTOL, tol_samples: int = 1, ungated=None):
    """Classify every event that is not identical in (sample, template) between
    an oracle result ``ref`` (run with ``margins=True``) and GPU events ``ev``.

    An event at batch column c is a *near tie* when one of the oracle's decision
    margins (threshold, template, pooling; ``oracle.c`` peel_m) over columns
    c - 1 .. c + 1 is below ``tol``; with ``ungated`` (the events of the same
    matcher on the same batch without the proposal gate) a *gate miss* when the
    ungated run agrees with the oracle there (it has the oracle's event, or not
    the gated run's extra one): the gate left that spike's score tile
    uncomputed; a *cascade* when it lies within 2M columns of a near tie or a
    gate miss (a decision taken the other way changes the subtractions, and so
    the later decisions, within 2M); otherwise *unexplained*.  Returns a dict
    with the counts, the events' smallest margins and a list of the
    unexplained ones."""
    mg = ref.margins
    if mg is None:
        raise ValueError("the oracle result carries no margins (run it with margins=True)")
    W = mg.shape[1]
    shift = first_sample - M // 2 + nt0min
    ref_s = ref.w + shift
    pairs, _, _ = match_events(ref_s, ref.t, ev["sample"], ev["template"], tol_samples)
    exact_r = {i for i, j in pairs if ref_s[i] == ev["sample"][j]}
    exact_g = {j for i, j in pairs if ref_s[i] == ev["sample"][j]}
    cols = [("oracle", int(ref.w[i]), int(ref.t[i])) for i in range(len(ref_s)) if i not in exact_r]
    cols += [("gpu", int(ev["sample"][j] - shift), int(ev["template"][j]))
             for j in range(len(ev)) if j not in exact_g]
    best = mg.min(axis=0)                       # smallest of the three margins per column
    kind = mg.argmin(axis=0)

    def local(c):
        lo, hi = max(0, c - 1), min(W, c + 2)
        k = lo + int(np.argmin(best[lo:hi]))
        return float(best[k]), ("threshold", "template", "pooling")[int(kind[k])]

    ties, out = [], {"near_tie": 0, "cascade": 0, "unexplained": 0, "events": len(cols), "kinds": {},
                     "margins": [], "unexplained_events": []}
    ug = None
    if ungated is not None:
        out["gate_miss"] = 0
        ug = {(int(a), int(b)) for a, b in zip(ungated["sample"] - shift, ungated["template"])}

    def gate_caused(side, c, t):
        if ug is None:
            return False
        return ((c, t) in ug) if side == "oracle" else ((c, t) not in ug)

    info = []
    for side, c, t in cols:
        m, k = local(c) if 0 <= c < W else (float("inf"), "outside")
        g = gate_caused(side, c, t)
        info.append((side, c, t, m, k, g))
        if m <= tol or g:
            ties.append(c)
    ties = np.asarray(sorted(ties), np.int64)
    for side, c, t, m, k, g in info:
        out["margins"].append(m)
        if m <= tol:
            out["near_tie"] += 1
            out["kinds"][k] = out["kinds"].get(k, 0) + 1
        elif g:
            out["gate_miss"] += 1
        elif len(ties) and np.abs(ties - c).min() <= 2 * M:
            out["cascade"] += 1
        else:
            out["unexplained"] += 1
            out["unexplained_events"].append({"side": side, "column": c, "template": t, "margin": m, "kind": k})
    return out


def check_events(ref, ev, feat, *, M=61, nt0min=20, first_sample=0, label="", explain=True,
                 tol=TIE_RTOL):
    """The parity bar against an oracle result (columns -> Kilosort's sample
    convention): (sample, template) identical for >= 99.5% of events, the rest
    within +-1 sample, each difference explained by a near tie when the oracle
    ran with margins; amplitudes of identical events within 1e-3 relative;
    features within 1e-3 of the spike's feature-vector magnitude."""
    ref_s = ref.w + first_sample - M // 2 + nt0min
    ident, within, pairs = agreement(ref_s, ref.t, ev["sample"], ev["template"], tol=1)
    print(f"[{label}] oracle {len(ref_s)} gpu {len(ev)} identical {ident:.5f} within1 {within:.5f}")
    assert ident >= 0.995, f"{label}: identical fraction {ident}"
    assert within >= 0.999, f"{label}: within-1 fraction {within}"
    if explain and ref.margins is not None:
        ex = explain_mismatches(ref, ev, first_sample=first_sample, M=M, nt0min=nt0min, tol=tol)
        print(f"[{label}] differing events {ex['events']}: near ties {ex['near_tie']} {ex['kinds']}, "
              f"cascades {ex['cascade']}, unexplained {ex['unexplained']}")
        assert ex["unexplained"] == 0, f"{label}: unexplained differences {ex['unexplained_events'][:10]}"
    ri = np.array([p[0] for p in pairs], np.int64)
    gi = np.array([p[1] for p in pairs], np.int64)
    exact = ref_s[ri] == ev["sample"][gi] if len(ri) else np.zeros(0, bool)
    ri, gi = ri[exact], gi[exact]
    if len(ri):
        amp_err = np.abs(ev["amplitude"][gi] - ref.amp[ri]) / np.abs(ref.amp[ri])
        print(f"[{label}] amp rel err max {amp_err.max():.3e}")
        assert amp_err.max() <= AMP_RTOL
    if feat is not None and len(ri):
        fr = ref.feat[ri]
        fg = feat[gi].astype(np.float64)
        scale = np.abs(fr).reshape(len(ri), -1).max(1)
        ferr = np.abs(fg - fr).reshape(len(ri), -1).max(1) / scale
        print(f"[{label}] feature err max {ferr.max():.3e}")
        assert ferr.max() <= FEAT_RTOL
    return ident, within
