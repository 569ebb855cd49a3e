"""CPU restatement of the Decoder's z-score baseline -- TEST INFRASTRUCTURE ONLY
(imported by tests/ as the checker; the product never imports it).

ZScoreSdmProcessor (src/Decoder/ZScoreSdmProcessor.cpp): onSpikes bins each
spike by time / binSamples (:85-94); computeBinValue folds a completed bin's
count into a running mean / M2 by Welford's recurrence while the baseline is
open (:96-114) and scores z = (count - mean) / sd, direction +-1 beyond the
trigger, once frozen (:116-140).  Series 0 counts every spike, series 1 + t
the spikes of template t.  Plain Python floats are IEEE FP64 with each
operation rounded as written, which is the recurrence the device runs.

Pinned in tests/test_zscore_oracle.py by closed forms: after n bins the
running mean and M2 equal the mean and the summed squared deviations of the
counts (numpy, pairwise sums) to rounding; a hand-worked three-bin case; and
the bin index as C++ integer division (truncation toward zero).
"""
from __future__ import annotations

import math

import numpy as np


def bin_samples(bin_ms: float, rate: float) -> int:
    """max(1, llround(bin_ms / 1000 * rate)) (:27-28; llround rounds halves away from zero)."""
    v = (bin_ms / 1000.0) * float(np.float32(rate))
    r = math.floor(abs(v) + 0.5)
    return max(1, int(math.copysign(r, v)))


def cpp_div(a: int, b: int) -> int:
    """C++ integer division (truncation toward zero)."""
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b > 0) else -q


class ZScore:
    def __init__(self, n_templates: int, bin_samples: int):
        self.S = 1 + n_templates
        self.bin = bin_samples
        self.counts: dict[int, np.ndarray] = {}
        self.first_open = 0
        self.seen = [0] * self.S
        self.mean = [0.0] * self.S
        self.m2 = [0.0] * self.S
        self.frozen = False
        self.base_mean = [0.0] * self.S
        self.base_sd = [0.0] * self.S
        self.trigger = 1.0
        self.late = 0

    def add(self, samples, templates) -> None:
        for s, t in zip(samples, templates):
            b = cpp_div(int(s), self.bin)
            if b < self.first_open:
                self.late += 1
                continue
            row = self.counts.setdefault(b, np.zeros(self.S, np.int64))
            row[0] += 1
            if 0 <= int(t) < self.S - 1:
                row[1 + int(t)] += 1

    def complete(self, end_sample: int):
        last = math.floor(end_sample / self.bin)
        zs, dirs = [], []
        for b in range(self.first_open, last):
            row = self.counts.pop(b, np.zeros(self.S, np.int64))
            z = np.zeros(self.S, np.float32)
            d = np.zeros(self.S, np.int8)
            for s in range(self.S):
                x = float(row[s])
                if not self.frozen:
                    self.seen[s] += 1
                    delta = x - self.mean[s]
                    self.mean[s] += delta / float(self.seen[s])
                    delta2 = x - self.mean[s]
                    self.m2[s] += delta * delta2
                elif self.base_sd[s] > 0.0:
                    zv = (x - self.base_mean[s]) / self.base_sd[s]
                    d[s] = 1 if zv > self.trigger else (-1 if zv < -self.trigger else 0)
                    z[s] = np.float32(zv)
            zs.append(z)
            dirs.append(d)
        self.first_open = max(self.first_open, last)
        return np.array(zs, np.float32).reshape(-1, self.S), np.array(dirs, np.int8).reshape(-1, self.S)

    def freeze(self, mean=None, sd=None, trigger: float = 1.0) -> None:
        if mean is None:
            self.base_mean = list(self.mean)
            self.base_sd = [math.sqrt(m2 / (n - 1)) if n > 1 else 0.0 for m2, n in zip(self.m2, self.seen)]
        else:
            self.base_mean = [float(v) for v in mean]
            self.base_sd = [float(v) for v in sd]
        self.trigger = float(trigger)
        self.frozen = True
