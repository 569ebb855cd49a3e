"""The long-run z-score baseline on the GPU (FP64, lss_zscore_*) against the
oracle (oracle/zscore.py): running means and M2 bit-identical (the same FP64
recurrence, rounded as written), z values and directions identical, late
events counted alike, and a ring overflow reported as an error."""
import numpy as np
import pytest

from livespikesorter_amd import EVENT_DTYPE, LssError
from livespikesorter_amd.stream import ZScoreBaseline
from oracle import zscore as Z

pytestmark = pytest.mark.gpu


def events(r, lo, hi, n, T):
    ev = np.zeros(n, EVENT_DTYPE)
    ev["sample"] = r.integers(lo, hi, n)
    ev["template"] = r.integers(-1, T + 2, n)        # ids outside 0..T-1 count in the total only
    ev["amplitude"] = r.random(n)
    return ev


def test_zscore_matches_oracle():
    r = np.random.default_rng(11)
    T, B = 37, 1500
    ref = Z.ZScore(T, B)
    print("[zscore] running baseline", flush=True)
    with ZScoreBaseline(T, bin_ms=50.0, ring_bins=64) as zb:
        assert zb.bin_samples == Z.bin_samples(50.0, 30000.0) == B
        end = 0
        for step in range(40):
            new_end = end + int(r.integers(1000, 9000))
            ev = events(r, max(0, end - 3000), new_end, int(r.integers(0, 400)), T)   # some late ones
            zb.add(ev)
            ref.add(ev["sample"], ev["template"])
            z, d = zb.complete(new_end)
            rz, rd = ref.complete(new_end)
            assert z.shape == rz.shape and not z.any() and not d.any()
            end = new_end
        st = zb.stats()
        assert np.array_equal(st["bins_seen"], ref.seen) and st["late"] == ref.late > 0
        assert np.array_equal(st["mean"], np.array(ref.mean)) and np.array_equal(st["m2"], np.array(ref.m2))
        # frozen at the running baseline, then at a given one
        zb.freeze(trigger_z=1.5)
        ref.freeze(trigger=1.5)
        for step in range(10):
            new_end = end + int(r.integers(1000, 9000))
            ev = events(r, end, new_end, int(r.integers(0, 600)), T)
            zb.add(ev)
            ref.add(ev["sample"], ev["template"])
            z, d = zb.complete(new_end)
            rz, rd = ref.complete(new_end)
            assert np.array_equal(z, rz) and np.array_equal(d, rd)
            end = new_end
        mean = r.random(T + 1) * 5
        sd = r.random(T + 1) + 0.1
        sd[3] = 0.0                                   # non-positive sd scores 0
        zb.freeze(mean, sd, trigger_z=0.5)
        ref.freeze(mean, sd, trigger=0.5)
        ev = events(r, end, end + 30000, 900, T)
        zb.add(ev)
        ref.add(ev["sample"], ev["template"])
        z, d = zb.complete(end + 30000)
        rz, rd = ref.complete(end + 30000)
        assert len(z) == 20 and np.array_equal(z, rz) and np.array_equal(d, rd) and d.any()
        assert not z[:, 3].any()
        # an event beyond the ring of pending bins
        far = np.zeros(1, EVENT_DTYPE)
        far["sample"] = end + 30000 + 100 * B
        zb.add(far)
        with pytest.raises(LssError, match="ring"):
            zb.complete(end + 31000)
