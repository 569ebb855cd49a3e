"""Pinning the z-score baseline oracle (oracle/zscore.py): Welford's running
mean and M2 against the closed forms, a hand-worked case, C++ integer
division for the bin index and llround for the bin size."""
import numpy as np

from oracle import zscore as Z


def test_hand_worked_three_bins():
    z = Z.ZScore(0, 10)
    z.add([1, 5, 20, 21, 22, 29], [0] * 6)       # bins: 0 -> 2 spikes, 1 -> 0, 2 -> 4
    z.complete(30)
    assert z.seen[0] == 3 and z.mean[0] == 2.0 and z.m2[0] == 8.0


def test_welford_matches_closed_forms():
    r = np.random.default_rng(0)
    counts = r.poisson(7.3, size=5000)
    z = Z.ZScore(0, 1)
    z.add(np.repeat(np.arange(len(counts)), counts), np.zeros(counts.sum(), int))
    z.complete(len(counts))
    x = counts.astype(np.float64)
    assert z.seen[0] == len(x)
    assert abs(z.mean[0] - x.mean()) <= 1e-12 * x.mean()
    assert abs(z.m2[0] - ((x - x.mean()) ** 2).sum()) <= 1e-9 * ((x - x.mean()) ** 2).sum()


def test_bin_index_truncates_and_late_spikes_drop():
    assert Z.cpp_div(-7, 3) == -2 and Z.cpp_div(7, 3) == 2 and Z.cpp_div(-6, 3) == -2
    z = Z.ZScore(2, 100)
    z.add([-50, 10, 150], [0, 1, 1])              # -50 / 100 = 0 in C++: bin 0
    z.complete(100)
    assert z.m2 == [0.0, 0.0, 0.0] and z.mean == [2.0, 1.0, 1.0]
    z.add([99, 250], [0, 0])                      # bin 0 already closed: late
    assert z.late == 1


def test_bin_samples_llround():
    assert Z.bin_samples(50, 30000.0) == 1500
    assert Z.bin_samples(0.01, 30000.0) == 1      # llround(0.3) = 0 -> at least 1
    assert Z.bin_samples(0.05, 30000.0) == 2      # 1.5 rounds away from zero
    assert Z.bin_samples(1.0, 29999.5) == 30      # 29.9995


def test_z_scores_and_directions():
    z = Z.ZScore(1, 10)
    z.freeze([4.0, 1.0], [2.0, 0.5], trigger=1.0)
    z.add([0, 1, 2, 3, 4, 5, 6, 7, 12, 33], [0, 0, 0, 0, 0, 0, 0, 0, 0, 0])
    zs, ds = z.complete(40)
    # bins 0..3 hold 8, 1, 0, 1 spikes (template 0 all of them)
    assert np.allclose(zs[:, 0], [2.0, -1.5, -2.0, -1.5]) and list(ds[:, 0]) == [1, -1, -1, -1]
    assert np.allclose(zs[:, 1], [14.0, 0.0, -2.0, 0.0]) and list(ds[:, 1]) == [1, 0, -1, 0]
