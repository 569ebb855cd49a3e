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
