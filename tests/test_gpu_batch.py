"""Parity of the HIP batched step (libodx) with the CPU oracle.

Deterministic mode (reference row order) is compared step by step against the
oracle's sequential QuickStep: contact counts and geom pairs exactly,
contact geometry, impulses and body state within the stated relative
tolerance.  The coloured mode is a different Gauss-Seidel order and is judged
by its LCP residual and the LCP's invariants.
"""
import sys

import numpy as np
import pytest

import oracle as O
from ode_amd import scenes as S

pytestmark = pytest.mark.gpu

# Tolerances (relative to the largest magnitude of the compared quantity).
TOL_STATE = 1e-4     # positions, quaternions, velocities after one step
TOL_LAMBDA = 1e-4    # constraint impulses after one step (float32 vs the float64 oracle)
TOL_CONTACT = 1e-5   # contact positions / normals / depths


def _odx():
    from ode_amd import odx
    return odx


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_joints(contacts, geoms):
    """Oracle contacts -> JOINT layout with the dJointAttach reversal applied."""
    j = O.contacts_to_joints(contacts, geoms)
    for r in j:
        if r[1] < 0:
            r[1], r[2] = r[2], -1
            r[6:9] = -r[6:9]
    return j


def relerr(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if a.size == 0:
        return 0.0
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-12))


def run_one_step(worlds, prm, order=0):
    odx = _odx()
    B = odx.Batch(worlds, prm)
    B.set_solver(order=order, max_iters=int(prm[7]), tol=0.0)
    B.step(1)
    return B


@pytest.mark.parametrize("damped", [False, True])
@pytest.mark.parametrize("seed", [0, 1])
def test_single_step_matches_oracle_reference_order(seed, damped):
    say("case: uneven worlds seed", seed, "damping and angular speed cap", damped)
    rng = np.random.default_rng(100 + seed)
    worlds = [S.uneven_world(rng, 8, 24) for _ in range(12)]
    # sink a few bodies into the ground so every world has contacts
    for w in worlds:
        w.bodies[:, 2] = np.minimum(w.bodies[:, 2], 0.3)
        if damped:
            # record flags bits 2-4 with slots 3/11/15/19/23 (dxStepBody damping and cap)
            nb = len(w.bodies)
            w.bodies[:, 12:15] = rng.normal(0, 3.0, size=(nb, 3))
            sel = rng.random(nb) < 0.6
            bits = rng.integers(1, 8, size=nb) * sel
            w.bodies[:, 31] = w.bodies[:, 31].astype(int) | (bits << 2)
            w.bodies[:, 3] = rng.uniform(0.01, 0.3, nb)
            w.bodies[:, 11] = rng.uniform(0.01, 0.3, nb)
            w.bodies[:, 15] = rng.uniform(0.0, 0.5, nb)
            w.bodies[:, 19] = rng.uniform(0.0, 2.0, nb)
            w.bodies[:, 23] = rng.uniform(0.5, 2.0, nb)
    prm = O.params()
    odx = _odx()
    B = run_one_step(worlds, prm)
    gb = B.bodies()
    if damped:
        # the damping fields travel back with the records
        allb = np.concatenate([w.bodies for w in worlds])
        for k in (3, 11, 15, 19, 23):
            sel = (allb[:, 31].astype(int) >> 2) != 0
            np.testing.assert_allclose(gb[sel, k], allb[sel, k], rtol=1e-7)
        np.testing.assert_array_equal(gb[:, 31], allb[:, 31].astype(np.float32))
    st = B.stats()
    assert B.overflow().sum() == 0
    for wi, w in enumerate(worlds):
        ref_b, ref_c, ref_l = O.world_step(w.bodies, w.geoms, prm, dtype=np.float64)
        gj = B.joints(wi)
        assert st["contacts"][wi] == len(ref_c), "contact count differs in world %d" % wi
        rj = oracle_joints(ref_c, w.geoms)
        # same geom pairs in the same (canonical) order
        np.testing.assert_array_equal(gj[:, 14:16].astype(int), ref_c[:, 7:9].astype(int))
        np.testing.assert_array_equal(gj[:, 1:3].astype(int), rj[:, 1:3].astype(int))
        assert relerr(gj[:, 3:6], rj[:, 3:6]) < TOL_CONTACT
        assert np.max(np.abs(gj[:, 6:9] - rj[:, 6:9])) < TOL_CONTACT
        assert np.max(np.abs(gj[:, 9] - rj[:, 9])) < 1e-5
        lam = B.lam(wi)
        assert len(lam) == len(ref_l)
        assert relerr(lam, ref_l) < TOL_LAMBDA, "lambda world %d: %g" % (wi, relerr(lam, ref_l))
        wb = B.world_bodies(wi, gb)
        assert relerr(wb[:, 0:3], ref_b[:, 0:3]) < TOL_STATE
        assert relerr(wb[:, 4:8], ref_b[:, 4:8]) < TOL_STATE
        vs = np.hstack([wb[:, 8:11], wb[:, 12:15]])
        vr = np.hstack([ref_b[:, 8:11], ref_b[:, 12:15]])
        assert np.max(np.abs(vs - vr)) / max(np.max(np.abs(vr)), 1.0) < TOL_STATE
    B.close()


def test_cylinders_on_a_plane_match_oracle():
    # cylinders standing on a cap, lying on their side and tilted, sunk a
    # little into the plane, among spheres: per world the float32 oracle's
    # contacts exactly (count and pairs) and its geometry, the float64
    # oracle's impulses and states after one step
    say("case: cylinders on a plane")
    rng = np.random.default_rng(21)
    worlds = []
    for w in range(8):
        bodies, geoms = [], []
        for k in range(6):
            r, L = rng.uniform(0.1, 0.3), rng.uniform(0.3, 1.0)
            m, I = S.cylinder_mass(1000.0, r, L)
            kind = k % 3
            if kind == 0:
                q, z = (1, 0, 0, 0), L / 2 - 0.01
            elif kind == 1:
                q, z = (np.cos(np.pi / 4), 0, np.sin(np.pi / 4), 0), r - 0.01
            else:
                q = rng.normal(size=4)
                z = 0.5 * L
            x, y = 1.5 * (k - 2.5), rng.uniform(-0.5, 0.5)
            bodies.append(S._body((x, y, z), q, m, I))
            geoms.append(S.geom(S.CYLINDER, k, (r, L)))
        m, I = S.sphere_mass(1000.0, 0.2)
        bodies.append(S._body((0.0, 2.0, 0.19), (1, 0, 0, 0), m, I))
        geoms.append(S.geom(S.SPHERE, 6, (0.2,)))
        geoms.append(S.plane(0, 0, 1, 0))
        worlds.append(S.World(bodies, geoms))
    prm = O.params()
    B = run_one_step(worlds, prm)
    gb = B.bodies()
    st = B.stats()
    assert B.overflow().sum() == 0
    total = 0
    for wi, w in enumerate(worlds):
        _, ref32, _ = O.world_step(w.bodies, w.geoms, prm, dtype=np.float32)
        ref_b, ref_c, ref_l = O.world_step(w.bodies, w.geoms, prm, dtype=np.float64)
        gj = B.joints(wi)
        assert st["contacts"][wi] == len(ref32) == len(ref_c), wi
        np.testing.assert_array_equal(gj[:, 14:16].astype(int), ref32[:, 7:9].astype(int))
        rj = oracle_joints(ref32, w.geoms)
        assert np.max(np.abs(gj[:, 3:10] - rj[:, 3:10])) < 2e-6
        lam = B.lam(wi)
        assert relerr(lam, ref_l) < TOL_LAMBDA, "lambda world %d: %g" % (wi, relerr(lam, ref_l))
        wb = B.world_bodies(wi, gb)
        assert relerr(wb[:, 0:3], ref_b[:, 0:3]) < TOL_STATE
        vs = np.hstack([wb[:, 8:11], wb[:, 12:15]])
        vr = np.hstack([ref_b[:, 8:11], ref_b[:, 12:15]])
        assert np.max(np.abs(vs - vr)) / max(np.max(np.abs(vr)), 1.0) < TOL_STATE
        total += len(ref_c)
    say("  contacts", total)
    assert total > 8 * 10
    B.close()


def test_towers_single_step_and_rows():
    say("case: towers single step")
    rng = np.random.default_rng(7)
    worlds = [S.tower_world(rng, levels=12) for _ in range(8)]
    prm = O.params()
    B = run_one_step(worlds, prm)
    gb = B.bodies()
    for wi, w in enumerate(worlds):
        ref_b, ref_c, ref_l = O.world_step(w.bodies, w.geoms, prm, dtype=np.float64)
        assert B.stats()["contacts"][wi] == len(ref_c)
        lam = B.lam(wi)
        assert relerr(lam, ref_l) < TOL_LAMBDA
        wb = B.world_bodies(wi, gb)
        assert relerr(wb[:, 0:3], ref_b[:, 0:3]) < TOL_STATE
        rows = B.rows(wi)
        # row layout: normal rows have lo = 0, hi = inf; friction rows carry mu and findex
        m = len(rows)
        assert m == 3 * len(ref_c)
        normals = rows[0::3]
        assert np.all(normals[:, 14] == 0) and np.all(np.isinf(normals[:, 15]))
        fr = rows[1::3]
        np.testing.assert_allclose(fr[:, 15], 0.5)
        fidx = fr[:, 30].view(np.int32)
        np.testing.assert_array_equal(fidx, np.arange(0, m, 3))
    B.close()


def test_multi_step_trajectory_tracks_float32_oracle():
    say("case: 10-step trajectory")
    rng = np.random.default_rng(21)
    worlds = [S.tower_world(rng, levels=6) for _ in range(4)]
    prm = O.params()
    odx = _odx()
    B = odx.Batch(worlds, prm)
    B.set_solver(order=0, max_iters=20)
    ref = [w.bodies.copy() for w in worlds]
    for s in range(10):
        B.step(1)
        for wi, w in enumerate(worlds):
            ref[wi], _, _ = O.world_step(ref[wi], w.geoms, prm, dtype=np.float32)
    gb = B.bodies()
    for wi in range(len(worlds)):
        wb = B.world_bodies(wi, gb)
        assert relerr(wb[:, 0:3], ref[wi][:, 0:3]) < 1e-3
    B.close()


def test_capsule_chain_with_ball_joints():
    say("case: capsule chains")
    rng = np.random.default_rng(3)
    worlds = [S.chain_world(rng) for _ in range(4)]
    prm = O.params(iters=40, skip_connected=1)
    B = run_one_step(worlds, prm)
    gb = B.bodies()
    for wi, w in enumerate(worlds):
        ref_b, ref_c, ref_l = O.world_step(w.bodies, w.geoms, prm, balls=w.balls, dtype=np.float64)
        assert B.stats()["contacts"][wi] == len(ref_c)
        lam = B.lam(wi)
        assert len(lam) == len(ref_l)
        assert relerr(lam, ref_l) < TOL_LAMBDA
        wb = B.world_bodies(wi, gb)
        assert relerr(wb[:, 0:3], ref_b[:, 0:3]) < TOL_STATE
        vs = np.hstack([wb[:, 8:11], wb[:, 12:15]])
        vr = np.hstack([ref_b[:, 8:11], ref_b[:, 12:15]])
        assert np.max(np.abs(vs - vr)) / max(np.max(np.abs(vr)), 1.0) < TOL_STATE
    B.close()


def hinge_slider_fixed_world(rng):
    """Five spheres (r 0.3 m) hung in a chain of hinges along x from the world
    (a motor on one hinge, a limit active on another, a motor pushing against
    its limit on a third), a sphere on a powered slider to the world, a fixed
    pair, and the z = 0 plane under the chain's lowest link so that it has
    contacts.  Joint frames are set with every body at identity orientation,
    as tests/test_gpu_joints.py does through the drop-in API; velocities are
    random."""
    INF = float("inf")
    m, I = S.sphere_mass(1000.0 / 9.0, 0.3)
    bodies, geoms, recs, exts = [], [], [], []

    def body(pos):
        bodies.append(S._body(pos, (1, 0, 0, 0), m, I, lvel=rng.normal(0, 0.5, 3), avel=rng.normal(0, 1.0, 3)))
        geoms.append(S.geom(S.SPHERE, len(bodies) - 1, (0.3,)))
        return len(bodies) - 1

    def ext(**kw):
        x = np.zeros(32)
        x[6] = 1.0
        x[16], x[17], x[18], x[19], x[20] = kw.get("lostop", -INF), kw.get("histop", INF), kw.get("vel", 0.0), \
            kw.get("fmax", 0.0), 1.0
        x[21], x[22], x[23], x[26], x[27] = 1e-5, 0.2, 1e-5, 0.2, 1e-5
        return x

    chain = [body((float(i), 0.0, 0.29 + 0.05 * i)) for i in range(5)]
    limits = {2: dict(fmax=3.0, vel=-1.0), 3: dict(lostop=0.1), 4: dict(lostop=0.2, fmax=2.0, vel=1.0)}
    for i in range(5):
        b1, b2 = (chain[0], -1) if i == 0 else (chain[i - 1], chain[i])
        anchor = np.array([i - 0.5, 0.0, 0.29 + 0.05 * i - 0.025])
        rec = np.zeros(16)
        rec[0], rec[1], rec[2] = 2, b1, b2
        rec[3:6] = anchor - bodies[b1][0:3]
        rec[6:9] = anchor - bodies[b2][0:3] if b2 >= 0 else anchor
        rec[9], rec[10] = 0.2, 1e-5
        x = ext(**limits.get(i, {}))
        x[0:3] = x[3:6] = (0.0, 1.0, 0.0)
        recs.append(rec)
        exts.append(x)
    s = body((0.0, 4.0, 1.0))
    rec = np.zeros(16)
    rec[0], rec[1], rec[2] = 3, s, -1
    x = ext(fmax=50.0, vel=0.5, lostop=-0.5, histop=0.5)
    x[0:3] = np.array([1.0, 0.0, 1.0]) / np.sqrt(2.0)
    x[10:13] = bodies[s][0:3]
    recs.append(rec)
    exts.append(x)
    fa, fb = body((3.0, 4.0, 1.0)), body((3.5, 4.5, 1.5))
    rec = np.zeros(16)
    rec[0], rec[1], rec[2] = 4, fa, fb
    x = ext()
    x[10:13] = bodies[fa][0:3] - bodies[fb][0:3]
    recs.append(rec)
    exts.append(x)
    geoms.append(S.plane(0, 0, 1, 0))
    return S.World(bodies, geoms, recs, exts)


def test_hinge_slider_fixed_joints_match_oracle():
    # batched worlds with hinge, slider and fixed joints (limits, motors and a
    # motor pushing against its limit evaluated on the device every step) and
    # contacts, in the reference row order: one step against the float64
    # oracle, then five more steps tracking it
    say("case: batched hinge / slider / fixed joints")
    rng = np.random.default_rng(14)
    worlds = [hinge_slider_fixed_world(rng) for _ in range(6)]
    prm = O.params(iters=30, skip_connected=1)
    odx = _odx()
    B = odx.Batch(worlds, prm)
    B.set_solver(order=0, max_iters=30, tol=0.0)
    B.step(1)
    gb = B.bodies()
    assert B.overflow().sum() == 0
    st = B.stats()
    for wi, w in enumerate(worlds):
        ref_b, ref_c, ref_l = O.world_step(w.bodies, w.geoms, prm, balls=w.balls, jext=w.jext, dtype=np.float64)
        assert len(ref_c) > 0 and st["contacts"][wi] == len(ref_c)
        lam = B.lam(wi)
        assert len(lam) == len(ref_l), (len(lam), len(ref_l))
        assert relerr(lam, ref_l) < TOL_LAMBDA, relerr(lam, ref_l)
        wb = B.world_bodies(wi, gb)
        assert relerr(wb[:, 0:3], ref_b[:, 0:3]) < TOL_STATE
        assert np.max(np.abs(wb[:, 4:8] - ref_b[:, 4:8])) < TOL_STATE
        vs = np.hstack([wb[:, 8:11], wb[:, 12:15]])
        vr = np.hstack([ref_b[:, 8:11], ref_b[:, 12:15]])
        assert np.max(np.abs(vs - vr)) / max(np.max(np.abs(vr)), 1.0) < TOL_STATE
    # rows per joint: 6 for the hinges with a motor or an active limit, the
    # slider and the fixed joint, 5 for the others
    m = len(B.lam(0)) - 3 * st["contacts"][0]
    say("  joint rows", m, "contacts", st["contacts"][:6])
    assert m == 5 + 5 + 6 + 6 + 6 + 6 + 6
    B.step(5)
    gb = B.bodies()
    for wi, w in enumerate(worlds):
        b = w.bodies.copy()
        for _ in range(6):
            b, _, _ = O.world_step(b, w.geoms, prm, balls=w.balls, jext=w.jext, dtype=np.float64)
        wb = B.world_bodies(wi, gb)
        err = max(relerr(wb[:, 0:3], b[:, 0:3]), np.max(np.abs(wb[:, 4:8] - b[:, 4:8])))
        assert err < 5 * TOL_STATE, (wi, err)
    B.close()


def test_hash_space_semantics_same_contacts():
    say("case: hash space")
    rng = np.random.default_rng(5)
    worlds = [S.uneven_world(rng, 8, 32) for _ in range(6)]
    for w in worlds:
        w.bodies[:, 2] = np.minimum(w.bodies[:, 2], 0.3)
    prm = O.params(space=1)
    B = run_one_step(worlds, prm)
    for wi, w in enumerate(worlds):
        ref_b, ref_c, ref_l = O.world_step(w.bodies, w.geoms, prm, dtype=np.float64)
        assert B.stats()["contacts"][wi] == len(ref_c)
        assert relerr(B.lam(wi), ref_l) < TOL_LAMBDA
    B.close()


def test_coloured_order_converges_and_respects_friction_cone():
    say("case: coloured order residual")
    rng = np.random.default_rng(9)
    worlds = [S.tower_world(rng, levels=16) for _ in range(16)]
    prm = O.params()
    odx = _odx()
    B = odx.Batch(worlds, prm)
    B.set_solver(order=1, max_iters=2000, tol=1e-4, check_every=5)
    B.step(1)
    st = B.stats()
    assert np.all(st["residual"] <= 1e-4), st["residual"]
    assert np.all(st["iters"] < 2000)
    for wi in range(len(worlds)):
        lam = B.lam(wi).reshape(-1, 3)
        assert np.all(lam[:, 0] >= 0)
        assert np.all(np.abs(lam[:, 1:]) <= 0.5 * lam[:, 0:1] * (1 + 1e-6) + 1e-6)
    # the GPU residual agrees with the float64 re-evaluation on the CPU
    for wi, w in enumerate(worlds[:4]):
        j = B.joints(wi).astype(np.float64)
        r64 = O.lcp_residual(w.bodies, j, prm, B.lam(wi))
        assert r64 < 2e-4, r64
    B.close()


def test_tolerance_mode_stops_early_and_reference_order_residual_matches_oracle():
    say("case: tolerance mode")
    rng = np.random.default_rng(11)
    worlds = [S.tower_world(rng, levels=8) for _ in range(4)]
    prm = O.params()
    odx = _odx()
    B = odx.Batch(worlds, prm)
    B.set_solver(order=0, max_iters=5000, tol=1e-4, check_every=1)
    B.step(1)
    st = B.stats()
    for wi, w in enumerate(worlds):
        ref_b, ref_c, _ = O.world_step(w.bodies, w.geoms, prm, dtype=np.float64)
        j = O.contacts_to_joints(ref_c, w.geoms)
        it, res, _ = O.solve_to_residual(w.bodies, j, prm, 1e-4, 5000)
        assert st["residual"][wi] <= 1e-4
        # same Gauss-Seidel order: the sweep count to the target matches within
        # float32 rounding effects
        assert abs(int(st["iters"][wi]) - it) <= max(2, it // 10), (st["iters"][wi], it)
    B.close()


def test_warm_start_from_previous_lambda_cuts_sweeps():
    say("case: warm start")
    rng = np.random.default_rng(13)
    worlds = [S.tower_world(rng, levels=12) for _ in range(8)]
    prm = O.params()
    odx = _odx()
    B = odx.Batch(worlds, prm)
    B.set_solver(order=1, max_iters=3000, tol=1e-4, check_every=2, warm=0)
    B.step(1)
    snap = B.bodies()
    B.step(1)
    cold = B.stats()["iters"].copy()
    lam_prev = np.zeros(B.row_start[-1], np.float32)
    for wi in range(len(worlds)):
        l = B.lam(wi)
        lam_prev[B.row_start[wi]:B.row_start[wi] + len(l)] = l
    # re-run the same step warm-started from its own converged impulses
    B.set_bodies(snap)
    B.set_lambda0(lam_prev)
    B.set_solver(order=1, max_iters=3000, tol=1e-4, check_every=2, warm=1)
    B.step(1)
    warm = B.stats()["iters"]
    assert np.all(warm <= 2), warm
    assert np.mean(cold) > 5 * np.mean(warm)
    B.close()


@pytest.mark.parametrize("order", [0, 1])
def test_lds_and_global_row_paths_agree(order):
    """Rows swept from LDS and rows streamed from global memory (LDS budget 0)
    give the same step within float32 rounding; a budget of 96 rows splits the
    batch between the two paths."""
    say("case: LDS vs global rows, order", order)
    rng = np.random.default_rng(31 + order)
    worlds = [S.uneven_world(rng, 8, 40) for _ in range(16)]
    for w in worlds:
        w.bodies[:, 2] = np.minimum(w.bodies[:, 2], 0.3)
    prm = O.params()
    odx = _odx()
    runs = {}
    for budget in (0, 96, 400):
        B = odx.Batch(worlds, prm)
        B.set_solver(order=order, max_iters=20, tol=0.0)
        B.set_solve_options(lds_rows=budget)
        B.step(1)
        runs[budget] = (B.bodies(), [B.lam(wi) for wi in range(len(worlds))], B.stats()["rows"])
        B.close()
    rows = runs[0][2]
    assert (rows > 96).any() and (rows <= 96).any(), rows
    ref_b, ref_l, _ = runs[0]
    for budget in (96, 400):
        b, l, r = runs[budget]
        np.testing.assert_array_equal(r, rows)
        assert relerr(b[:, 0:3], ref_b[:, 0:3]) < 1e-6
        v, vr = np.hstack([b[:, 8:11], b[:, 12:15]]), np.hstack([ref_b[:, 8:11], ref_b[:, 12:15]])
        assert np.max(np.abs(v - vr)) / max(np.max(np.abs(vr)), 1.0) < 1e-5
        for wi in range(len(worlds)):
            assert relerr(l[wi], ref_l[wi]) < 1e-4, (budget, wi, relerr(l[wi], ref_l[wi]))


@pytest.mark.parametrize("lds_rows", [0, 400])
def test_unconverged_warm_start_restarts_cold(lds_rows):
    """A warm start that misses the residual target by sweep 6 restarts from
    lambda = 0 and then repeats the cold solve exactly: 6 more sweeps and the
    same impulses."""
    say("case: warm-start restart, lds_rows", lds_rows)
    rng = np.random.default_rng(17)
    worlds = [S.tower_world(rng, levels=8) for _ in range(6)]
    prm = O.params()
    odx = _odx()
    out = {}
    for mode in ("cold", "restart"):
        B = odx.Batch(worlds, prm)
        B.set_solve_options(restart_after=6 if mode == "restart" else 0, lds_rows=lds_rows)
        if mode == "restart":
            B.set_lambda0(np.random.default_rng(3).uniform(-50, 50, size=B.row_start[-1]).astype(np.float32))
        B.set_solver(order=1, max_iters=3000, tol=1e-4, check_every=2, warm=1 if mode == "restart" else 0)
        B.step(1)
        st = B.stats()
        assert np.all(st["residual"] <= 1e-4)
        out[mode] = (st["iters"].copy(), [B.lam(wi) for wi in range(len(worlds))])
        B.close()
    np.testing.assert_array_equal(out["restart"][0], out["cold"][0] + 6)
    for a, b in zip(out["restart"][1], out["cold"][1]):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("warm", [0, 2, 3])
def test_split_streams_fused_kernel_and_multi_step_graphs_change_nothing(warm):
    """World ranges stepped on separate streams inside multi-step graphs, and
    the persistent fused kernel (several whole steps per launch, network run
    inside it), give bit-identical bodies, impulses and sweep counts to one
    range stepped one graph launch per step (worlds are independent)."""
    say("case: splits / fused, warm", warm)
    rng = np.random.default_rng(23)
    worlds = [S.uneven_world(rng, 8, 40) for _ in range(37)]
    prm = O.params()
    odx = _odx()
    mlp = None
    if warm == 3:
        mlp = odx.Mlp(11)
        X = np.random.default_rng(1).normal(size=(64, 16)).astype(np.float32)
        norm = np.concatenate([np.zeros(16), np.ones(16), np.zeros(3), np.full(3, 0.05)]).astype(np.float32)
        mlp.set_normalization(norm)
    out = []
    for nsplit, gsteps, fused in ((1, 1, 0), (4, 3, 0), (1, 1, 3), (2, 4, 2)):
        B = odx.Batch(worlds, prm)
        B.set_splits(nsplit, gsteps)
        B.set_fused(fused)
        if mlp is not None:
            B.set_mlp(mlp)
        B.set_solver(order=1, max_iters=3000, tol=1e-4, check_every=2, warm=warm)
        B.step(7)
        st = B.stats()
        out.append((B.bodies(), [B.lam(wi) for wi in range(len(worlds))], st["iters"].copy(),
                    B.samples()[0] if warm == 3 else None))
        B.close()
    for k in range(1, len(out)):
        np.testing.assert_array_equal(out[0][0], out[k][0])
        np.testing.assert_array_equal(out[0][2], out[k][2])
        for a, b in zip(out[0][1], out[k][1]):
            np.testing.assert_array_equal(a, b)
        if warm == 3:
            # the same feature vectors, gathered in a different slot order
            A, Bf = out[0][3], out[k][3]
            assert A.shape == Bf.shape
            np.testing.assert_array_equal(np.sort(A, axis=0), np.sort(Bf, axis=0))


def test_surface_table_and_category_bits_match_oracle():
    # the near-callback policy on the device: each geom's material picks the
    # pair's surface (friction, bounce, soft CFM, mode, contact limit) and its
    # category / collide bits decide whether the pair collides at all
    say("case: material-pair surface table with category bits")
    rng = np.random.default_rng(77)
    worlds = [S.uneven_world(rng, 8, 20) for _ in range(10)]
    for w in worlds:
        w.bodies[:, 2] = np.minimum(w.bodies[:, 2], 0.35)
        w.bodies[:, 8:11] = rng.normal(0, 0.5, size=(len(w.bodies), 3))
        body = w.geoms[:, 1] >= 0
        w.geoms[body, 13] = rng.integers(1, 3, size=int(body.sum()))
        w.geoms[~body, 13] = 0
        grp = rng.integers(0, 2, size=len(w.geoms))
        w.geoms[:, 14] = np.where(grp == 0, 1, 2)        # category: group bit
        w.geoms[:, 15] = np.where(grp == 0, 1 | 4, 2 | 4)  # collide: own group and the ground's bit
        w.geoms[~body, 14] = 4
        w.geoms[~body, 15] = 1 | 2
    approx = 0x1000 | 0x2000
    t = O.surface_table(3, {
        (0, 1): {"mode": approx, "mu": 0.2, "max_contacts": 2},
        (0, 2): {"mode": approx | 0x004 | 0x010, "mu": 0.8, "bounce": 0.5, "bounce_vel": 0.1, "soft_cfm": 1e-3},
        (1, 1): {"mode": approx, "mu": 0.0, "max_contacts": 1},
        (1, 2): {"mode": approx | 0x008, "mu": 0.5, "soft_erp": 0.5, "max_contacts": 3},
        (2, 2): {"mode": 0x001 | approx, "mu": 0.4, "mu2": 0.1},
    })
    prm = O.params()
    odx = _odx()
    B = odx.Batch(worlds, prm)
    B.set_surface_table(t, category_filter=True)
    B.set_solver(order=0, max_iters=int(prm[7]), tol=0.0)
    B.step(1)
    gb = B.bodies()
    st = B.stats()
    assert B.overflow().sum() == 0
    dropped = 0
    for wi, w in enumerate(worlds):
        ref_b, ref_c, ref_l = O.world_step(w.bodies, w.geoms, prm, dtype=np.float64, table=t, category_filter=True)
        _, all_c, _ = O.world_step(w.bodies, w.geoms, prm, dtype=np.float64)
        dropped += len(all_c) - len(ref_c)
        gj = B.joints(wi)
        assert st["contacts"][wi] == len(ref_c), "contact count differs in world %d" % wi
        np.testing.assert_array_equal(gj[:, 14:16].astype(int), ref_c[:, 7:9].astype(int))
        # the surface index of each contact is the material pair's entry
        mat = w.geoms[:, 13].astype(int)
        np.testing.assert_array_equal(gj[:, 13].astype(int), mat[ref_c[:, 7].astype(int)] * 3 + mat[ref_c[:, 8].astype(int)])
        lam = B.lam(wi)
        assert len(lam) == len(ref_l)
        assert relerr(lam, ref_l) < TOL_LAMBDA, "lambda world %d: %g" % (wi, relerr(lam, ref_l))
        wb = B.world_bodies(wi, gb)
        assert relerr(wb[:, 0:3], ref_b[:, 0:3]) < TOL_STATE
        vs = np.hstack([wb[:, 8:11], wb[:, 12:15]])
        vr = np.hstack([ref_b[:, 8:11], ref_b[:, 12:15]])
        assert np.max(np.abs(vs - vr)) / max(np.max(np.abs(vr)), 1.0) < TOL_STATE
    say("  contacts dropped by the policy (bits and limits):", dropped)
    assert dropped > 0
    B.close()


def test_snapshot_round_trip_continues_the_same_steps(tmp_path):
    # odxBatchSaveSnapshot / odxBatchLoadSnapshot: the loaded batch holds the
    # state, the records, the solver settings and the surface table, and
    # continues as the original does (cold start)
    say("case: batch snapshot round trip")
    odx = _odx()
    rng = np.random.default_rng(31)
    worlds = [S.uneven_world(rng, 8, 20) for _ in range(16)]
    for w in worlds:
        w.bodies[:, 2] = np.minimum(w.bodies[:, 2], 0.6)
        w.geoms[:, 13] = np.where(w.geoms[:, 1] >= 0, 1, 0)
    t = O.surface_table(2, {(0, 1): {"mode": 0x1000 | 0x2000, "mu": 0.3, "max_contacts": 3},
                            (1, 1): {"mode": 0x1000 | 0x2000, "mu": 0.7}})
    B = odx.Batch(worlds, O.params())
    B.set_surface_table(t)
    B.set_solver(order=1, max_iters=30, tol=0.0)
    B.step(5)
    path = tmp_path / "batch.snap"
    B.save_snapshot(path)
    C = odx.Batch.load_snapshot(path)
    assert C.nworlds == B.nworlds and C.nbodies == B.nbodies
    np.testing.assert_array_equal(C.body_start, B.body_start)
    b0, c0 = B.bodies(), C.bodies()
    np.testing.assert_allclose(c0, b0, rtol=1e-6, atol=1e-7)  # quaternions renormalised on upload
    B.step(3)
    C.step(3)
    sb, sc = B.stats(), C.stats()
    np.testing.assert_array_equal(sb["contacts"], sc["contacts"])
    np.testing.assert_array_equal(sb["iters"], sc["iters"])
    gb, gc = B.bodies(), C.bodies()
    assert relerr(gc[:, 0:3], gb[:, 0:3]) < TOL_STATE
    vel = np.r_[8:11, 12:15]
    assert np.max(np.abs(gc[:, vel] - gb[:, vel])) / max(np.max(np.abs(gb[:, vel])), 1.0) < TOL_STATE
    B.close()
    C.close()


@pytest.mark.parametrize("warm", [0, 2])
def test_autotuned_schedule_keeps_the_results(warm):
    # the autotuner only picks a launch schedule and puts the state back;
    # stepping afterwards gives the bodies that the default schedule gives
    say("case: autotune, warm", warm)
    from ode_amd import odx, scenes as S
    prm = O.params(iters=3000)
    worlds = S.tower_batch(64, seed=17, levels=8)
    out = []
    for tune in (False, True):
        B = odx.Batch(worlds, prm)
        B.set_solver(order=1, max_iters=3000, tol=1e-4, check_every=2, warm=warm)
        B.step(3)
        if tune:
            B.set_splits(1, 4)
            before = B.bodies().copy()
            sch, ms = B.autotune(steps=6, reps=1)
            say("  schedule", sch, "ms/step", ms)
            assert sch["splits"] in (1, 2, 4, 8) and sch["fused"] in (0, 4) and ms > 0
            assert np.array_equal(B.bodies(), before)
        B.step(5)
        out.append(B.bodies().copy())
        B.close()
    assert np.array_equal(out[0], out[1])
