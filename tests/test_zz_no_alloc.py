"""odxMlpPredictNoAlloc (no allocation, fixed per-device scratch) against
odxMlpPredict: two calls in a row, the second smaller, give the same
impulses bit for bit; a batch beyond the fixed maximum is refused."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX = 16384


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def features(rng, n):
    X = rng.normal(size=(n, 16)).astype(np.float32)
    X[:, 15] = rng.uniform(5.0, 50.0, size=n).astype(np.float32)  # force scale s > 0
    return X


def test_predict_without_allocation_matches_predict():
    from ode_amd import odx
    L = odx.lib()
    odx._setup_mlp_api(L)
    L.odxMlpPredictNoAlloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    rng = np.random.default_rng(0)
    m = odx.Mlp(3)
    X0 = features(rng, 2048)
    norm = np.concatenate([X0.mean(0), X0.std(0) + 1e-3, np.zeros(3), np.ones(3)]).astype(np.float32)
    m.set_normalization(norm)
    for n in (9000, 1234):
        say("case: odxMlpPredictNoAlloc n =", n)
        X = features(rng, n)
        want = m.predict(X)
        got = np.zeros((n, 3), np.float32)
        assert L.odxMlpPredictNoAlloc(m.h, X.ctypes.data, n, got.ctypes.data) == n
        assert np.array_equal(got, want)
    say("case: odxMlpPredictNoAlloc beyond the maximum")
    X = features(rng, MAX + 1)
    Y = np.zeros((MAX + 1, 3), np.float32)
    odx.last_messages()
    assert L.odxMlpPredictNoAlloc(m.h, X.ctypes.data, MAX + 1, Y.ctypes.data) == 0
    assert any("exceed" in msg for _, msg in odx.last_messages())
    m.close()
