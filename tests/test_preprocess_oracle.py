"""Pinning the preprocessing oracle (oracle/preprocess.py) against independent
forms: the direct circular-correlation sum, torch.fft executing Kilosort's
formula (src/Python/OnlineKilosort.py:124-128, KS/preprocessing.py:139-154),
the Butterworth design's frequency response (scipy.signal.freqz), a sort,
torch.median and exact rational arithmetic."""
from fractions import Fraction

import numpy as np
import pytest

from oracle import preprocess as pp


def rng(seed=0):
    return np.random.default_rng(seed)


@pytest.mark.parametrize("W", [16, 17, 40, 41])
@pytest.mark.parametrize("F", [101, 9, 40])
@pytest.mark.parametrize("sem", [pp.SEM_KILOSORT, pp.SEM_LSS])
def test_fft_form_equals_direct_sum(W, F, sem):
    r = rng(W * 1000 + F)
    X = r.standard_normal((3, W))
    hp = r.standard_normal(F).astype(np.float32)
    a = pp.highpass(X, hp, sem)
    b = pp.highpass_direct(X, hp, sem)
    assert np.allclose(a, b, rtol=0, atol=1e-10)


def _torch_kilosort(X, hp):
    """Kilosort's own lines, in torch FP64 on the CPU."""
    import torch
    from torch.fft import fft, fftshift, ifft
    Xt = torch.from_numpy(np.asarray(X, np.float64))
    h = torch.from_numpy(hp.astype(np.float64))
    NT = Xt.shape[1]
    ft = h.shape[0]
    if ft < NT:
        pad = (NT - ft) // 2
        fhp = fft(torch.cat((torch.zeros(pad, dtype=h.dtype), h, torch.zeros(pad + (NT - pad * 2 - ft), dtype=h.dtype))))
    elif ft > NT:
        crop = (ft - NT) // 2
        fhp = fft(h[crop:crop + NT])
    else:
        fhp = fft(h)
    Y = torch.real(ifft(fft(Xt) * torch.conj(fhp)))
    return fftshift(Y, dim=-1).numpy()


@pytest.mark.parametrize("W", [64, 65, 3000, 3001])
@pytest.mark.parametrize("F", [30122, 51, 64])
def test_kilosort_convention_matches_torch(W, F):
    r = rng(W + F)
    X = r.standard_normal((4, W))
    hp = pp.highpass_filter() if F == 30122 else r.standard_normal(F).astype(np.float32)
    assert np.allclose(pp.highpass(X, hp, pp.SEM_KILOSORT), _torch_kilosort(X, hp), rtol=0, atol=1e-9)


@pytest.mark.parametrize("W", [200, 201, 3122, 3123])
def test_lss_shift_rule(W):
    """shift_kernel reads y[(k + W//2) mod W] (myGPUhelpers.cu:72-83): equal to
    Kilosort's fftshift for even W, one sample later for odd W."""
    r = rng(W)
    X = r.standard_normal((2, W))
    hp = pp.highpass_filter()
    ks = pp.highpass(X, hp, pp.SEM_KILOSORT)
    lss = pp.highpass(X, hp, pp.SEM_LSS)
    if W % 2 == 0:
        assert np.array_equal(ks, lss)
    else:
        assert np.allclose(lss, np.roll(ks, 1, axis=1), rtol=0, atol=1e-12)
        assert not np.allclose(lss, ks, atol=1e-3)


@pytest.mark.parametrize("W", [3000, 3001, 6000])
def test_sinusoid_gain_is_the_filtfilt_response(W):
    """A sinusoid periodic in W comes out scaled by |H(f)|^2 of butter(3, 300 Hz)
    (filtfilt applies the filter forward and backward), with zero phase for
    even W (and for LSS at odd W), advanced one sample for Kilosort at odd W."""
    from scipy.signal import butter, freqz
    b, a = butter(3, 300, fs=30000, btype="high")
    hp = pp.highpass_filter()
    n = np.arange(W)
    for cyc in (2, 7, 20, 31, 60, 150):
        f = cyc / W
        _, H = freqz(b, a, worN=[2 * np.pi * f])
        g = abs(H[0]) ** 2
        x = np.cos(2 * np.pi * f * n + 0.3)
        for sem in (pp.SEM_KILOSORT, pp.SEM_LSS):
            adv = 1 if (sem == pp.SEM_KILOSORT and W % 2) else 0
            want = g * np.cos(2 * np.pi * f * (n + adv) + 0.3)
            got = pp.highpass(x[None, :], hp, sem)[0]
            assert np.allclose(got, want, rtol=0, atol=2e-6), (cyc, sem, np.abs(got - want).max())


def test_filter_fixture_is_symmetric_and_highpass():
    hp = pp.highpass_filter().astype(np.float64)
    c = len(hp) // 2
    assert np.allclose(hp[c - 2000:c + 2001], hp[c - 2000:c + 2001][::-1], atol=1e-9)
    assert abs(hp.sum()) < 1e-6          # no DC gain
    assert 0.9 < hp[c] < 1.0


@pytest.mark.parametrize("C", [1, 2, 7, 8, 383, 384])
def test_lower_median(C):
    import torch
    r = rng(C)
    x = r.integers(-50, 50, size=(64, C)).astype(np.float32)   # many ties
    med = pp.lower_median(x)
    assert np.array_equal(med, np.sort(x, axis=1)[:, (C - 1) // 2])
    assert np.array_equal(med, torch.median(torch.from_numpy(x), dim=1).values.numpy())
    if C % 2:
        assert np.array_equal(med, np.median(x, axis=1).astype(np.float32))


def _f32_nearest(q: Fraction) -> np.float32:
    """Correctly rounded FP32 of a rational (ties to even)."""
    f = np.float32(float(q))
    cands = [np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))]
    errs = [abs(Fraction(float(c)) - q) for c in cands]
    best = min(errs)
    picks = [c for c, e in zip(cands, errs) if e == best]
    if len(picks) > 1:
        picks = [c for c in picks if (np.frombuffer(np.float32(c).tobytes(), np.uint32)[0] & 1) == 0]
    return np.float32(picks[0])


def test_channel_means_are_correctly_rounded():
    r = rng(5)
    for W in (3, 7, 3122):
        x = r.integers(-2000, 2000, size=(W, 5)).astype(np.int16)
        m = pp.channel_means(x)
        for c in range(5):
            assert m[c] == _f32_nearest(Fraction(int(x[:, c].astype(np.int64).sum()), W))
        # equal to the FP32 running sum divided by W while partial sums stay exact
        s32 = np.zeros(5, np.float32)
        for w in range(W):
            s32 = (s32 + x[w].astype(np.float32)).astype(np.float32)
        assert np.array_equal(m, (s32 / np.float32(W)).astype(np.float32))


def test_preprocess_stages_and_flags():
    r = rng(9)
    W, C = 500, 12
    x = (r.integers(-300, 300, size=(W, C)) + 1000).astype(np.int16)
    hp = pp.highpass_filter()
    out = pp.preprocess(x, pp.PRE_MEAN | pp.PRE_CAR | pp.PRE_HP, hp)
    assert out["p2p"] == float(x.max()) - float(x.min())
    assert np.allclose(out["x1"].mean(1), 0, atol=200)           # means removed (then CAR)
    assert np.all(np.sort(out["x1"], axis=0)[(C - 1) // 2] == 0)  # each sample's lower median is 0
    assert np.allclose(out["out"], pp.highpass(out["x1"], hp), atol=0)
    only_mean = pp.preprocess(x, pp.PRE_MEAN)
    assert np.all(only_mean["med"] == 0)
    assert np.array_equal(only_mean["x1"], (x.astype(np.float32) - pp.channel_means(x)).T)
