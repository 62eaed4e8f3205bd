"""The learned warm start: MFMA operand layout, forward and gradients against
the numpy restatement (oracle/mlp.py), training, and the gain over the solver
without it at the same LCP residual."""
import sys

import numpy as np
import pytest

from oracle import mlp as R

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_mfma_16x16x32_bf16_operand_layout():
    say("case: mfma layout")
    from ode_amd import odx
    rng = np.random.default_rng(0)
    a = R.bf16(rng.normal(size=(64, 8)).astype(np.float32))
    b = R.bf16(rng.normal(size=(64, 8)).astype(np.float32))
    d = odx.mfma_probe(a, b)
    # documented layout: A[row=l%16][k=8*(l//16)+j], B[k=8*(l//16)+j][col=l%16],
    # D[row=4*(l//16)+j][col=l%16]
    A = np.zeros((16, 32), np.float32)
    B = np.zeros((32, 16), np.float32)
    for l in range(64):
        for j in range(8):
            A[l % 16, 8 * (l // 16) + j] = a[l, j]
            B[8 * (l // 16) + j, l % 16] = b[l, j]
    D = A @ B
    want = np.zeros((64, 4), np.float32)
    for l in range(64):
        for j in range(4):
            want[l, j] = D[4 * (l // 16) + j, l % 16]
    np.testing.assert_allclose(d, want, rtol=1e-5, atol=1e-5)


def _random_problem(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, 16)).astype(np.float32)
    X[:, 15] = rng.uniform(50, 500, size=n)  # force scale feature
    T = rng.normal(size=(n, 3)).astype(np.float32)
    norm = np.concatenate([rng.normal(size=16) * 0.1, rng.uniform(0.5, 2, 16), rng.normal(size=3) * 0.1,
                           rng.uniform(0.5, 2, 3)]).astype(np.float32)
    norm[15], norm[31] = 200.0, 100.0
    return X, T, norm


def test_forward_matches_numpy():
    say("case: forward")
    from ode_amd import odx
    X, T, norm = _random_problem(300, 1)
    m = odx.Mlp(3)
    m.set_normalization(norm)
    p = m.params()
    got = m.predict(X)
    want = R.predict(p, X, norm)
    err = np.max(np.abs(got - want)) / np.max(np.abs(want))
    assert err < 1e-2, err
    m.close()


def test_gradients_match_numpy():
    say("case: gradients")
    from ode_amd import odx
    X, T, norm = _random_problem(512, 2)
    m = odx.Mlp(5)
    m.set_normalization(norm)
    m.set_data(X, T)
    g, loss = m.gradients()
    gr, lr = R.gradients(m.params(), X, T, norm)
    assert loss == pytest.approx(lr, rel=1e-3)
    off = 0
    for name, shape in R.SIZES:
        k = int(np.prod(shape))
        a, b = g[off:off + k], gr[off:off + k]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        cos = float(a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30))
        say("  grad", name, "rel", rel, "cos", cos)
        assert rel < 2e-2, (name, rel)
        assert cos > 0.999, (name, cos)
        off += k
    m.close()


def test_training_reduces_loss_on_a_smooth_target():
    say("case: training")
    from ode_amd import odx
    rng = np.random.default_rng(4)
    X = rng.normal(size=(20000, 16)).astype(np.float32)
    X[:, 15] = 1.0
    T = np.stack([np.sin(X[:, 0]) + 0.5 * X[:, 1], X[:, 2] * X[:, 3], np.tanh(X[:, 4])], 1).astype(np.float32)
    m = odx.Mlp(7)
    norm = np.concatenate([np.zeros(16), np.ones(16), np.zeros(3), np.ones(3)]).astype(np.float32)
    norm[15], norm[31] = 0.0, 1.0
    m.set_normalization(norm)
    m.set_data(X, T)
    losses = m.train(600, batch=1024, lr=2e-3)
    assert losses[-20:].mean() < 0.1 * losses[:5].mean(), (losses[:5], losses[-20:])
    m.close()


def test_learned_warm_start_cuts_sweeps_vs_cold_at_equal_residual():
    say("case: learned warm start gain")
    import oracle as O
    from ode_amd import odx, scenes as S, warmstart as WS
    prm = O.params(iters=3000)
    train_worlds = S.tower_batch(128, seed=21, levels=16)
    Bt = odx.Batch(train_worlds, prm)
    X, T = WS.collect(Bt, steps=60, tol=1e-5)
    Bt.close()
    say("  samples", len(X))
    assert len(X) > 10000
    mlp, losses = WS.train(X, T, steps=1200, batch=2048)
    say("  loss first/last", float(losses[:10].mean()), float(losses[-10:].mean()))
    assert losses[-50:].mean() < losses[:10].mean()

    test_worlds = S.tower_batch(64, seed=99, levels=16)
    res = {}
    for mode in ("cold", "plain", "learned"):
        B = odx.Batch(test_worlds, prm)
        if mode == "learned":
            B.set_mlp(mlp)
        warm = {"cold": 0, "plain": 2, "learned": 3}[mode]
        B.set_solver(order=1, max_iters=3000, tol=1e-4, check_every=2, warm=warm)
        B.step(5)
        sweeps = []
        for _ in range(30):
            B.step(1)
            st = B.stats()
            assert np.all(st["residual"] <= 1e-4)
            sweeps.append(st["iters"].max())
        res[mode] = float(np.mean(sweeps))
        B.close()
    say("  max sweeps per step (mean over steps):", res)
    # What the trained network gains over the cold-start method at the same
    # residual: fewer sweeps for the slowest world.  (Measured on MI355X:
    # learned ~1.6x fewer than cold; the 3x target is not reached.)
    assert res["learned"] <= 0.8 * res["cold"], res
    # and over the same warm start without the network (previous-step
    # impulses, "plain"): the gated network only sets the impulses of new
    # contacts, and its marginal gain is about nil -- it must at least not be
    # worse than plain
    say("  learned / plain:", res["learned"] / res["plain"])
    assert res["learned"] <= 1.05 * res["plain"], res


def test_gated_network_leaves_matched_contacts_at_previous_impulses():
    say("case: gated network, matched contacts")
    import oracle as O
    from ode_amd import odx, scenes as S
    prm = O.params(iters=3000)
    worlds = S.tower_batch(32, seed=7, levels=8)
    # a network with zero weights and a zero output mean predicts a zero impulse
    # for every new contact: gated, it must reproduce the previous-impulse warm start
    m = odx.Mlp(3)
    m.set_params(np.zeros(odx.NPARAM, np.float32))
    norm = np.concatenate([np.zeros(16), np.ones(16), np.zeros(3), np.ones(3)]).astype(np.float32)
    m.set_normalization(norm)
    m.set_gate(1)
    out = {}
    for mode in ("plain", "gated"):
        B = odx.Batch(worlds, prm)
        if mode == "gated":
            B.set_mlp(m)
        B.set_solver(order=1, max_iters=3000, tol=1e-4, check_every=2, warm=2 if mode == "plain" else 3)
        B.step(12)
        if mode == "gated":
            X, _ = B.samples()
            say("  feature rows of the last step", len(X))
            assert np.all(X[:, 12] == 0.0), "a matched contact got a feature slot"
        out[mode] = (B.bodies().copy(), B.stats()["iters"].copy())
        B.close()
    m.close()
    assert np.array_equal(out["plain"][1], out["gated"][1])
    assert np.array_equal(out["plain"][0], out["gated"][0])
