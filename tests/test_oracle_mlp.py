"""Pins the numpy network restatement (oracle/mlp.py) before it checks the GPU:
its analytic gradients agree with central finite differences (float64, no
bf16 rounding), its bf16 rounding is round-to-nearest-even, and GELU matches
the closed form."""
import math

import numpy as np
import pytest

from oracle import mlp as R


@pytest.fixture
def exact():
    R.EXACT = True
    yield
    R.EXACT = False


def _loss(p, X, T, norm):
    y, _ = R.forward(p, X, norm)
    tn = (T - norm[32:35]) / norm[35:38]
    return float(np.sum((y - tn) ** 2) / len(X))


def test_gradients_match_finite_differences(exact):
    rng = np.random.default_rng(0)
    p = rng.normal(size=R.NPARAM) * 0.2
    X = rng.normal(size=(7, 16))
    T = rng.normal(size=(7, 3))
    norm = np.concatenate([rng.normal(size=16) * 0.1, rng.uniform(0.5, 2, 16), rng.normal(size=3) * 0.1,
                           rng.uniform(0.5, 2, 3)])
    g, loss = R.gradients(p, X, T, norm)
    assert loss == pytest.approx(_loss(p, X, T, norm), rel=1e-12)
    # check a spread of coordinates from every parameter block
    off = 0
    for name, shape in R.SIZES:
        k = int(np.prod(shape))
        for idx in rng.choice(k, size=min(6, k), replace=False):
            i = off + idx
            e = 1e-6
            pp, pm = p.copy(), p.copy()
            pp[i] += e
            pm[i] -= e
            fd = (_loss(pp, X, T, norm) - _loss(pm, X, T, norm)) / (2 * e)
            assert g[i] == pytest.approx(fd, rel=1e-5, abs=1e-8), (name, idx)
        off += k


def test_bf16_round_to_nearest_even():
    vals = np.array([1.0, 1.00390625, 1.005859375, 1.0078125, -3.14159, 65504.0, 1e-3], np.float32)
    got = R.bf16(vals)
    # bf16 keeps 8 significant bits: 1 + 2^-8 is a tie between 1 and 1 + 2^-7 -> even (1.0)
    assert got[0] == 1.0
    assert got[1] == 1.0
    assert got[2] == np.float32(1.0078125)
    assert got[3] == np.float32(1.0078125)
    assert abs(got[4] + 3.140625) < 1e-7
    for v, r in zip(vals, got):
        assert abs(float(r) - float(v)) <= abs(float(v)) * 2 ** -8


def test_gelu_closed_form(exact):
    x = np.linspace(-4, 4, 41)
    y, d = R.gelu(x)
    for xi, yi, di in zip(x, y, d):
        c = math.sqrt(2 / math.pi)
        u = c * (xi + 0.044715 * xi ** 3)
        assert yi == pytest.approx(0.5 * xi * (1 + math.tanh(u)), abs=1e-12)
        e = 1e-6
        fd = (0.5 * (xi + e) * (1 + math.tanh(c * ((xi + e) + 0.044715 * (xi + e) ** 3))) -
              0.5 * (xi - e) * (1 + math.tanh(c * ((xi - e) + 0.044715 * (xi - e) ** 3)))) / (2 * e)
        assert di == pytest.approx(fd, abs=1e-6)
