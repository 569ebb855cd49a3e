"""Pins the CPU oracle of the proposal network (oracle/proposal.py).

The oracle's hand-derived backward pass is checked against central finite
differences of its own loss, and its convolution against a brute-force loop
over the definition, so a transposed weight, a flipped tap or a dropped ReLU
mask fails here before any GPU gradient is compared with it.
"""
import numpy as np

from oracle import proposal as P


def _params(rng, scale=0.3):
    return P.unpack(rng.normal(0, scale, P.N_PARAMS))


def test_conv_matches_definition():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(2, 3, 11))
    Wt = rng.normal(size=(4, 3, P.KW))
    b = rng.normal(size=4)
    out = P.conv(x, Wt, b)
    ref = np.zeros_like(out)
    for g in range(2):
        for co in range(4):
            for w in range(11):
                s = b[co]
                for ci in range(3):
                    for k in range(P.KW):
                        j = w + k - P.PAD
                        if 0 <= j < 11:
                            s += Wt[co, ci, k] * x[g, ci, j]
                ref[g, co, w] = s
    assert np.allclose(out, ref, atol=1e-12)


def test_group_input_halo():
    xw = np.arange(40 * 5, dtype=np.float64).reshape(40, 5)
    x0 = P.group_input(xw)
    assert x0.shape == (3, P.CIN, 5)
    assert np.all(x0[0, :4] == 0)                  # channels -4..-1 are off the probe
    assert np.array_equal(x0[0, 4], xw[0])
    assert np.array_equal(x0[1, 0], xw[12])        # group 1 starts at channel 16 - 4
    assert np.array_equal(x0[2, 4 + 7], xw[39])
    assert np.all(x0[2, 4 + 8:] == 0)              # channels 40.. are off the probe


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(2)
    params = _params(rng)
    x0 = rng.normal(size=(2, P.CIN, 13))
    y = (rng.random((2, 13)) < 0.2).astype(np.float64)
    loss, grads = P.loss_and_grad(params, x0, y, pos_weight=5.0)
    flat = P.pack(params)
    g_ana = P.pack(grads)
    # a fixed sample of coordinates from every parameter block
    idx = np.unique(np.concatenate([rng.choice(P.N_PARAMS, 60, replace=False),
                                    [0, P.N_PARAMS - 1, P.N_PARAMS - 2]]))
    eps = 1e-6
    for i in idx:
        fp = flat.copy(); fp[i] += eps
        fm = flat.copy(); fm[i] -= eps
        lp, _ = P.loss_and_grad(P.unpack(fp), x0, y, 5.0)
        lm, _ = P.loss_and_grad(P.unpack(fm), x0, y, 5.0)
        num = (lp - lm) / (2 * eps)
        assert abs(num - g_ana[i]) <= 1e-6 + 1e-5 * abs(num), (i, num, g_ana[i])


def test_labels():
    y = P.labels(np.array([0, 5]), np.array([1, 0]), 2, 8, tol=1)
    assert y[1].tolist() == [1, 1, 0, 0, 0, 0, 0, 0]
    assert y[0].tolist() == [0, 0, 0, 0, 1, 1, 1, 0]


def test_bf16_rounding_known_answers():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -7, -2.5, 3.0e-3])
    r = P.bf16(v)
    assert r[0] == 1.0
    assert r[1] == 1.0                       # halfway: ties to the even neighbour
    assert r[2] == 1.0 + 2 ** -6             # halfway between 1+2^-7 and 1+2^-6: even is 1+2^-6
    assert r[3] == 1.0 + 2 ** -7             # representable
    assert r[4] == -2.5
    assert abs(r[5] - 3.0e-3) <= 3.0e-3 * 2 ** -8      # half a unit in the 8th significant bit
