"""Parity of the HIP large-world step (odxLarge*) with the CPU oracle.

The hash-space pair set and the contacts are compared with the oracle's
(dxHashSpace semantics, dCollide).  The solve sweeps contacts colour by
colour.  In row order 0 the normal rows of all colours come first and the
friction rows after; giving the oracle's QuickStep the contacts in the GPU's
colour order makes its ReorderPrep order (non-findex rows, then findex rows,
each in joint order) the same Gauss-Seidel order.  In row order 2 each
contact's rows are swept together, which is the oracle's joint order over the
same colour-ordered contacts.  Either way impulses and body states after the
step must match the float64 oracle within the stated tolerance.
"""
import sys

import numpy as np
import pytest

import oracle as O
from ode_amd import scenes as S

pytestmark = pytest.mark.gpu

TOL_STATE = 1e-4     # positions, quaternions, velocities after one step (relative)
TOL_LAMBDA = 1e-4    # impulses after one step, float32 GPU vs float64 oracle (relative to max |lambda|)
TOL_CONTACT_F64 = 1e-4  # contact positions / normals / depths, float32 GPU vs the float64 oracle
TOL_CONTACT_F32 = 2e-5  # the same against the float32 oracle (same arithmetic; cullPoints' atan2 may differ)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def relerr(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if a.size == 0:
        return 0.0
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-12))


def make(world, prm, **kw):
    from ode_amd import odx
    return odx.Large(world, prm, **kw)


def colour_order(colours):
    return np.argsort(colours, kind="stable")


@pytest.mark.parametrize("seed,row_order", [(0, 0), (1, 0), (0, 2), (1, 2)])
def test_pairs_contacts_and_coloured_step_match_oracle(seed, row_order):
    say("case: large mixed pile seed", seed, "row order", row_order)
    w = S.mixed_pile(1500, seed=60 + seed)
    prm = O.params(space=1, w=1.0, row_order=row_order)
    L = make(w, prm)
    L.step(1)
    L.sync()
    cnt = L.counts()
    assert cnt["pair_overflow"] == 0 and cnt["contact_overflow"] == 0 and cnt["uncoloured"] == 0
    assert cnt["barrier_error"] == 0

    # broadphase: the hash-space pair set, exactly (pairs of two static geoms
    # make no rows and are left out); AABBs and poses are computed without
    # fused multiply-add on both sides, so they agree bit for bit
    body = w.geoms[:, 1].astype(np.int32)
    ab = O.aabbs(w.geoms, w.bodies, dtype=np.float32)
    ref = O.pairs(ab, body, space="hash")
    ref = {(int(a), int(b)) for a, b in ref if body[a] >= 0 or body[b] >= 0}
    got = [tuple(map(int, p)) for p in L.pairs()]
    assert len(got) == len(set(got)), "duplicate pairs"
    got = set(got)
    for a, b in got:
        assert a < b
    assert got == ref, sorted(got ^ ref)[:5]

    # narrowphase: per pair the float32 oracle's contact count, exactly, and
    # its geometry; the float64 oracle's geometry within the stated tolerance
    _, ref32, _ = O.world_step(w.bodies, w.geoms, prm, dtype=np.float32)
    ref_b, ref_c, _ = O.world_step(w.bodies, w.geoms, prm, dtype=np.float64)
    gj = L.joints()
    assert len(gj) == cnt["contacts"] == len(ref32)
    key = lambda g1, g2: (int(g1), int(g2))
    gpu_pair = {}
    for r in gj:
        gpu_pair.setdefault(key(r[14], r[15]), []).append(r)
    for tol, refc in ((TOL_CONTACT_F32, ref32), (TOL_CONTACT_F64, ref_c)):
        by_pair = {}
        for r in refc:
            by_pair.setdefault(key(r[7], r[8]), []).append(r)
        if refc is ref32:
            assert set(by_pair) == set(gpu_pair)
        err = np.zeros(3)
        for k, rs in by_pair.items():
            gs = gpu_pair.get(k, [])
            if refc is ref32:
                assert len(gs) == len(rs), k
            elif len(gs) != len(rs):
                continue  # a float64 borderline count; the float32 counts are exact above
            for g, r in zip(gs, rs):
                sg = -1.0 if body[k[0]] < 0 else 1.0   # dJointAttach reversal negates the normal
                err = np.maximum(err, [np.max(np.abs(g[3:6] - r[0:3])) / max(1.0, np.max(np.abs(r[0:3]))),
                                       np.max(np.abs(g[6:9] - sg * r[3:6])), abs(g[9] - r[6])])
        say("contact errors (pos, normal, depth) vs", "float32" if refc is ref32 else "float64", err)
        assert np.all(err < tol), err

    # colouring: every contact coloured, no dynamic body twice in one colour
    col = L.colours()
    assert len(col) == len(gj) and col.min() >= 0 and col.max() < cnt["colours"]
    b1 = gj[:, 1].astype(int)
    b2 = gj[:, 2].astype(int)
    for c in range(cnt["colours"]):
        sel = col == c
        bs = np.concatenate([b1[sel], b2[sel][b2[sel] >= 0]])
        assert len(bs) == len(np.unique(bs)), "colour %d shares a body" % c
    say("contacts", len(gj), "colours", cnt["colours"], "rounds", cnt["colour_rounds"])

    # solve + integration: the oracle's QuickStep with the contacts in colour order
    order = colour_order(col)
    joints = gj[order].astype(np.float64)
    joints[:, 0] = 0
    ob, olam = O.quickstep(w.bodies, joints, prm, dtype=np.float64)
    lam = L.lam()
    rc = np.zeros(len(gj) + 1, int)
    rc[1:] = np.cumsum(np.full(len(gj), 3))
    assert len(lam) == rc[-1] == len(olam)
    glam = np.concatenate([lam[rc[k]:rc[k + 1]] for k in order])
    assert relerr(glam, olam) < TOL_LAMBDA, relerr(glam, olam)
    gb = L.bodies()
    assert relerr(gb[:, 0:3], ob[:, 0:3]) < TOL_STATE
    assert relerr(gb[:, 4:8], ob[:, 4:8]) < TOL_STATE
    vs = np.hstack([gb[:, 8:11], gb[:, 12:15]])
    vr = np.hstack([ob[:, 8:11], ob[:, 12:15]])
    assert np.max(np.abs(vs - vr)) / max(np.max(np.abs(vr)), 1.0) < TOL_STATE
    L.close()


def test_geom_with_more_partners_than_the_counting_pass_keeps():
    # a box carrying 7 x 7 small spheres: its pair list (49 partners of a lower
    # index) is longer than the 16 the counting pass keeps, so the writing pass
    # searches again; the pair set still equals the hash space's
    say("case: large, box with 49 partners")
    bodies, geoms = [], []
    m, I = S.box_mass(1000.0, 1.0, 1.0, 1.0)
    bodies.append(S._body((0.0, 0.0, 0.5), (1, 0, 0, 0), m, I))
    geoms.append(S.geom(S.BOX, 0, (1.0, 1.0, 1.0)))
    ms, Is = S.sphere_mass(1000.0, 0.05)
    for a in range(7):
        for b in range(7):
            k = len(bodies)
            bodies.append(S._body((-0.42 + 0.14 * a, -0.42 + 0.14 * b, 1.045), (1, 0, 0, 0), ms, Is))
            geoms.append(S.geom(S.SPHERE, k, (0.05,)))
    geoms.append(S.plane(0, 0, 1, 0))
    w = S.World(bodies, geoms)
    prm = O.params(space=1, w=1.0)
    L = make(w, prm)
    L.step(1)
    L.sync()
    body = w.geoms[:, 1].astype(np.int32)
    ab = O.aabbs(w.geoms, w.bodies, dtype=np.float32)
    ref = {(int(a), int(b)) for a, b in O.pairs(ab, body, space="hash") if body[a] >= 0 or body[b] >= 0}
    got = [tuple(map(int, p)) for p in L.pairs()]
    assert len(got) == len(set(got))
    assert sum(1 for a, b in got if a == 0) >= 49
    assert set(got) == ref
    L.close()


def contact_rows(surface):
    """Rows of one contact joint for a surface-table entry (dxJointContact::getInfo1,
    ode/src/joints/contact.cpp:49-122): the normal row, then one friction row
    per positive mu (mu and mu2 with dContactMu2, else two rows for mu)."""
    mode, mu, mu2 = int(surface[0]), surface[1], surface[2]
    if mode & 0x001:
        return 1 + int(mu > 0) + int(mu2 > 0)
    return 1 + 2 * int(mu > 0)


@pytest.mark.parametrize("row_order", [0, 2])
def test_surface_table_and_category_bits_match_oracle(row_order):
    # the near-callback policy on the large world: the category / collide bit
    # test removes hash-space pairs, each geom's material picks the pair's
    # surface and contact limit; pairs, contacts, surface indices and the
    # coloured step match the oracle's (contacts of 1, 2 and 3 rows)
    say("case: large mixed pile with a material-pair surface table, row order", row_order)
    rng = np.random.default_rng(91)
    w = S.mixed_pile(1500, seed=62)
    body = w.geoms[:, 1].astype(np.int32)
    dyn = body >= 0
    w.geoms[dyn, 13] = rng.integers(1, 3, size=int(dyn.sum()))
    w.geoms[~dyn, 13] = 0
    grp = rng.integers(0, 2, size=len(w.geoms))
    w.geoms[:, 14] = np.where(grp == 0, 1, 2)
    w.geoms[:, 15] = np.where(grp == 0, 1 | 4, 2 | 4)
    w.geoms[~dyn, 14] = 4
    w.geoms[~dyn, 15] = 1 | 2
    approx = 0x1000 | 0x2000
    t = O.surface_table(3, {
        (0, 1): {"mode": approx, "mu": 0.2, "max_contacts": 2},
        (0, 2): {"mode": approx | 0x004 | 0x010, "mu": 0.8, "bounce": 0.5, "bounce_vel": 0.1, "soft_cfm": 1e-3},
        (1, 1): {"mode": approx, "mu": 0.0, "max_contacts": 1},
        (1, 2): {"mode": approx | 0x008, "mu": 0.5, "soft_erp": 0.5, "max_contacts": 3},
        (2, 2): {"mode": 0x001 | approx, "mu": 0.4, "mu2": 0.1},
    })
    prm = O.params(space=1, w=1.0, row_order=row_order)
    L = make(w, prm)
    L.set_surface_table(t, category_filter=True)
    L.step(1)
    L.sync()
    cnt = L.counts()
    assert cnt["pair_overflow"] == 0 and cnt["contact_overflow"] == 0 and cnt["uncoloured"] == 0

    cat = w.geoms[:, 14].astype(np.uint32)
    col = w.geoms[:, 15].astype(np.uint32)
    ab = O.aabbs(w.geoms, w.bodies, dtype=np.float32)
    every = {(int(a), int(b)) for a, b in O.pairs(ab, body, space="hash") if body[a] >= 0 or body[b] >= 0}
    ref = {(int(a), int(b)) for a, b in O.pairs(ab, body, space="hash", cat=cat, col=col)
           if body[a] >= 0 or body[b] >= 0}
    got = {tuple(map(int, p)) for p in L.pairs()}
    assert got == ref, sorted(got ^ ref)[:5]
    say("  pairs", len(got), "of", len(every), "pass the bit test")
    assert len(ref) < len(every)

    _, ref32, _ = O.world_step(w.bodies, w.geoms, prm, dtype=np.float32, table=t, category_filter=True)
    gj = L.joints()
    assert len(gj) == len(ref32)
    count = lambda c: {k: n for k, n in zip(*np.unique(c[:, 0] * 4096 + c[:, 1], return_counts=True))}
    assert count(gj[:, 14:16].astype(np.int64)) == count(ref32[:, 7:9].astype(np.int64))
    mat = w.geoms[:, 13].astype(int)
    np.testing.assert_array_equal(gj[:, 13].astype(int), mat[gj[:, 14].astype(int)] * 3 + mat[gj[:, 15].astype(int)])

    col_ = L.colours()
    order = colour_order(col_)
    joints = gj[order].astype(np.float64)
    joints[:, 0] = 0
    ob, olam = O.quickstep(w.bodies, joints, prm, dtype=np.float64, table=t)
    tab = t.reshape(-1, 16)
    m = np.array([contact_rows(tab[int(s)]) for s in gj[:, 13]], int)
    rc = np.concatenate([[0], np.cumsum(m)])
    lam = L.lam()
    assert len(lam) == rc[-1] == len(olam)
    glam = np.concatenate([lam[rc[k]:rc[k + 1]] for k in order])
    assert relerr(glam, olam) < TOL_LAMBDA, relerr(glam, olam)
    gb = L.bodies()
    assert relerr(gb[:, 0:3], ob[:, 0:3]) < TOL_STATE
    vs = np.hstack([gb[:, 8:11], gb[:, 12:15]])
    vr = np.hstack([ob[:, 8:11], ob[:, 12:15]])
    assert np.max(np.abs(vs - vr)) / max(np.max(np.abs(vr)), 1.0) < TOL_STATE
    say("  contacts", len(gj), "rows", rc[-1], "by surface", np.bincount(gj[:, 13].astype(int), minlength=9))
    L.close()


def test_rows_match_oracle_contact_rows():
    say("case: large rows")
    w = S.mixed_pile(300, seed=70)
    prm = O.params(space=1)
    L = make(w, prm)
    L.step(1)
    L.sync()
    gj = L.joints()
    rows = L.rows()
    assert len(rows) == 3 * len(gj)
    # normal rows: lo 0, hi inf; friction rows: mu bound, findex -> the contact's normal row
    np.testing.assert_array_equal(rows[0::3, 14], 0)
    assert np.all(np.isinf(rows[0::3, 15]))
    np.testing.assert_allclose(rows[1::3, 15], 0.5)
    np.testing.assert_array_equal(rows[1::3, 30].view(np.int32), np.arange(0, len(rows), 3))
    np.testing.assert_array_equal(rows[2::3, 30].view(np.int32), np.arange(0, len(rows), 3))
    np.testing.assert_array_equal(rows[0::3, 28].view(np.int32), gj[:, 1].astype(np.int32))
    np.testing.assert_array_equal(rows[0::3, 29].view(np.int32), gj[:, 2].astype(np.int32))
    # Jacobians (stored times Ad, Ad in the last slot) against the oracle's getInfo2 rows
    surf = np.zeros(16)
    surf[0], surf[1] = prm[O.P_MODE], prm[O.P_MU]
    bodies = w.bodies.astype(np.float64)
    for k in range(0, len(gj), max(1, len(gj) // 40)):
        g = gj[k].astype(np.float64)
        b1, b2 = int(g[1]), int(g[2])
        contact = np.concatenate([g[3:6], g[6:9], [g[9]]])
        J, fi = O.contact_rows(contact, surf, bodies[b1], bodies[b2] if b2 >= 0 else None, fps=1.0 / prm[O.P_H],
                               erp=prm[O.P_ERP])
        assert len(J) == 3
        for q in range(3):
            row = rows[3 * k + q].astype(np.float64)
            ad = row[31]
            np.testing.assert_allclose(row[0:6] / ad, J[q][0:6], rtol=1e-4, atol=1e-5)
            if b2 >= 0:
                np.testing.assert_allclose(row[8:14] / ad, J[q][8:14], rtol=1e-4, atol=1e-5)
    L.close()


@pytest.mark.parametrize("row_order", [0, 2])
def test_step_is_deterministic(row_order):
    say("case: large determinism, row order", row_order)
    w = S.mixed_pile(800, seed=71)
    prm = O.params(space=1, iters=10, row_order=row_order)
    out = []
    for _ in range(2):
        L = make(w, prm)
        L.step(3)
        L.sync()
        out.append((L.bodies(), L.colours(), L.lam()))
        L.close()
    for a, b in zip(out[0], out[1]):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("row_order", [0, 2])
def test_more_sweeps_lower_the_lcp_residual(row_order):
    say("case: large residual, row order", row_order)
    w = S.mixed_pile(600, seed=72)
    res = []
    for iters in (5, 60):
        prm = O.params(space=1, iters=iters, w=1.0, row_order=row_order)
        L = make(w, prm)
        L.step(1)
        L.sync()
        gj = L.joints().astype(np.float64)
        col = L.colours()
        lam = L.lam().astype(np.float64)
        # natural row order of the joint list as given
        r = O.lcp_residual(w.bodies, gj, prm, lam)
        res.append(r)
        lamr = lam.reshape(-1, 3)
        assert np.all(lamr[:, 0] >= 0)
        assert np.all(np.abs(lamr[:, 1:]) <= 0.5 * lamr[:, 0:1] * (1 + 1e-5) + 1e-6)
        assert len(col) == len(gj)
        L.close()
    say("residual 5 / 60 sweeps", res)
    assert res[1] < 0.5 * res[0]


def test_capacity_overflow_fails_the_sync():
    say("case: large overflow")
    from ode_amd import odx
    w = S.mixed_pile(300, seed=73)
    prm = O.params(space=1)
    L = make(w, prm, contact_cap=16)
    L.step(1)
    with pytest.raises(odx.OdxError, match="contacts"):
        L.sync()
    assert L.counts()["contact_overflow"] > 0
    L.close()


def test_barrier_timeout_fails_the_step_and_the_world_recovers():
    # the timeout path of the grid barrier, forced from the host: with a spin
    # limit of 0 every wait fails at once, every workgroup still leaves both
    # cooperative kernels, and the next sync reports the failed step
    say("case: large barrier timeout (spin limit 0)")
    from ode_amd import odx
    w = S.sphere_pour(2000, seed=6)
    prm = O.params(space=1, mu=0.3, iters=10, w=1.0)
    L = make(w, prm)
    L.step(3)
    L.sync()
    before = L.bodies()
    L.set_spin_limit(0)
    L.step(1)
    with pytest.raises(odx.OdxError, match="barrier 1"):
        L.sync()
    # the failed step changed no body record (positions, velocities and the
    # force accumulators)
    assert np.array_equal(L.bodies(), before)
    L.set_spin_limit()
    L.step(2)
    L.sync()
    cnt = L.counts()
    assert cnt["barrier_error"] == 0 and cnt["uncoloured"] == 0
    assert np.all(np.isfinite(L.bodies()[:, :15]))
    L.close()


def test_sphere_pour_steps_and_settles():
    say("case: sphere pour 20000")
    w = S.sphere_pour(20000, seed=4)
    prm = O.params(space=1, mu=0.3, iters=30, w=1.0)
    L = make(w, prm)
    L.step(40)
    L.sync()
    cnt = L.counts()
    b = L.bodies()
    assert np.all(np.isfinite(b[:, :15]))
    assert cnt["contacts"] > 500 and cnt["uncoloured"] == 0
    # nothing leaves the open box through the walls or the floor
    assert np.all(b[:, 2] > -0.05)
    assert np.all(np.abs(b[:, 0:2]) < 5.0 + 0.05)
    say("counts", cnt)
    L.close()


def test_long_pour_is_deterministic():
    # 200k spheres poured for 900 steps, twice: the settled pile has colours of
    # more than one contact per thread of the solve's grid, and a race in the
    # solve would show as a different state after this many steps
    say("case: 200k sphere pour, 900 steps, twice")
    w = S.sphere_pour(200000, seed=5)
    prm = O.params(space=1, mu=0.3, iters=30, w=1.3, row_order=2)
    out = []
    for _ in range(2):
        L = make(w, prm)
        for _ in range(9):
            L.step(100)
            L.sync()
        cnt = L.counts()
        out.append((L.bodies(), cnt))
        say("  counts", cnt, "largest colour", int(np.bincount(L.colours()).max()))
        L.close()
    assert out[0][1] == out[1][1]
    np.testing.assert_array_equal(out[0][0], out[1][0])
