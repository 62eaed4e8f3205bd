"""The ODE-compatible C API (include/ode.h) driven through a near-callback
frame loop, checked against the oracle: candidate pairs, dCollide contacts,
and dWorldQuickStep in the reference's island row order."""
import ctypes
import sys

import numpy as np
import pytest

import oracle as O
from oracle.islands import island_order
from ode_amd import scenes as S

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class Frame:
    """Builds a world through the C API and records what the callback did."""

    def __init__(self, world_scene, space="sap", mu=0.5, iters=20, dynamic=False, levels=None):
        from ode_amd import ode as A
        self.A = A
        L = A.api()
        self.L = L
        self.w = L.dWorldCreate()
        L.dWorldSetGravity(self.w, 0, 0, -9.81)
        L.dWorldSetQuickStepNumIterations(self.w, iters)
        if not dynamic:
            # a fixed iteration count (the dynamic adjustment off: all three zero)
            z = A.dReal(0.0)
            L.dWorldSetQuickStepDynamicIterationParameters(self.w, ctypes.byref(z), ctypes.byref(z), ctypes.byref(z))
        self.space = {"sap": lambda: L.dSweepAndPruneSpaceCreate(None, A.dSAP_AXES_XYZ),
                      "hash": lambda: L.dHashSpaceCreate(None),
                      "simple": lambda: L.dSimpleSpaceCreate(None)}[space]()
        if levels is not None:
            L.dHashSpaceSetLevels(self.space, *levels)
        # the oracle's model of the space's lists, kept in step with the library
        self.model = O.SpaceModel(space, levels=levels or (-3, 10))
        self.space_kind = space
        self.group = L.dJointGroupCreate(0)
        self.mu = mu
        self.scene = world_scene
        self.bodies, self.geoms, self.geom_index = [], [], {}
        for rec in world_scene.bodies:
            b = L.dBodyCreate(self.w)
            L.dBodySetPosition(b, *[float(x) for x in rec[0:3]])
            q = (A.dReal * 4)(*[float(x) for x in rec[4:8]])
            L.dBodySetQuaternion(b, q)
            L.dBodySetLinearVel(b, *[float(x) for x in rec[8:11]])
            L.dBodySetAngularVel(b, *[float(x) for x in rec[12:15]])
            m = A.dMass()
            L.dMassSetParameters(ctypes.byref(m), float(rec[24]), 0, 0, 0, float(rec[25]), float(rec[26]), float(rec[27]),
                                 float(rec[28]), float(rec[29]), float(rec[30]))
            L.dBodySetMass(b, ctypes.byref(m))
            self.bodies.append(b)
        for gi, g in enumerate(world_scene.geoms):
            cls = int(g[0])
            if cls == S.SPHERE:
                h = L.dCreateSphere(self.space, float(g[2]))
            elif cls == S.BOX:
                h = L.dCreateBox(self.space, float(g[2]), float(g[3]), float(g[4]))
            elif cls == S.CAPSULE:
                h = L.dCreateCapsule(self.space, float(g[2]), float(g[3]))
            else:
                h = L.dCreatePlane(self.space, float(g[2]), float(g[3]), float(g[4]), float(g[5]))
            self.model.add(gi)
            if int(g[1]) >= 0:
                L.dGeomSetBody(h, self.bodies[int(g[1])])
            self.geoms.append(h)
            self.geom_index[h] = gi
        self.body_index = {b: i for i, b in enumerate(self.bodies)}
        # each body's geoms, most recently attached first (the body's geom list)
        self.body_geoms = {i: [] for i in range(len(self.bodies))}
        for gi, g in enumerate(world_scene.geoms):
            if int(g[1]) >= 0:
                self.body_geoms[int(g[1])].insert(0, gi)

    def collide(self):
        A, L = self.A, self.L
        self.pairs, self.contacts = [], []
        buf = (A.dContact * 4)()

        def near(data, o1, o2):
            self.pairs.append((self.geom_index[o1], self.geom_index[o2]))
            n = L.dCollide(o1, o2, 4, ctypes.cast(ctypes.byref(buf[0], A.dContact.geom.offset),
                                                   ctypes.POINTER(A.dContactGeom)), ctypes.sizeof(A.dContact))
            b1, b2 = L.dGeomGetBody(o1), L.dGeomGetBody(o2)
            for i in range(n):
                c = buf[i]
                c.surface.mode = A.dContactApprox1
                c.surface.mu = self.mu
                j = L.dJointCreateContact(self.w, self.group, ctypes.byref(c))
                L.dJointAttach(j, b1, b2)
                i1 = self.body_index.get(b1, -1) if b1 else -1
                i2 = self.body_index.get(b2, -1) if b2 else -1
                self.contacts.append((i1, i2, list(c.geom.pos)[:3], list(c.geom.normal)[:3], c.geom.depth,
                                      self.geom_index[o1], self.geom_index[o2]))

        self._cb = A.NearCallback(near)
        L.dSpaceCollide(self.space, None, self._cb)

    def expected_pairs(self, enabled=None, cat=None, col=None):
        """The oracle's callback sequence for the library's current state."""
        ab = O.aabbs(self.scene.geoms, self.state(), np.float32)
        body = self.scene.geoms[:, 1].astype(np.int32)
        return self.model.collide(ab, body, cat=cat, col=col, enabled=enabled)

    def step(self, h=0.01):
        """dWorldQuickStep; the oracle's model sees the geoms of the stepped
        bodies move in stepping order (dxStepBody's dGeomMoved)."""
        norm = [(c[0], c[1]) if c[0] >= 0 else (c[1], -1) for c in self.contacts]
        bodies, _ = island_order(len(self.bodies), norm)
        rc = self.L.dWorldQuickStep(self.w, h)
        for b in bodies:
            for gi in self.body_geoms[b]:
                self.model.moved(gi)
        return rc

    def state(self):
        A, L = self.A, self.L
        out = np.zeros((len(self.bodies), 32))
        for i, b in enumerate(self.bodies):
            out[i, 0:3] = A.vec(L.dBodyGetPosition(b))
            out[i, 4:8] = A.vec(L.dBodyGetQuaternion(b), 4)
            out[i, 8:11] = A.vec(L.dBodyGetLinearVel(b))
            out[i, 12:15] = A.vec(L.dBodyGetAngularVel(b))
            out[i, 24:32] = self.scene.bodies[i, 24:32]
        return out

    def close(self):
        L = self.L
        L.dJointGroupDestroy(self.group)
        L.dSpaceDestroy(self.space)
        L.dWorldDestroy(self.w)


def settle_scene(seed, n=14):
    rng = np.random.default_rng(seed)
    w = S.uneven_world(rng, n, n)
    w.bodies[:, 2] = np.minimum(w.bodies[:, 2], 0.35) + rng.uniform(0, 0.05, size=len(w.bodies))
    return w


def check_contacts(F, exact_counts=True):
    """Every callback's dCollide against the float32 oracle's dCollide of the
    same ordered pair in the same state: the same count, the same geometry."""
    st = F.state()
    by_pair = {}
    for c in F.contacts:
        by_pair.setdefault((c[5], c[6]), []).append(c)
    worst = 0.0
    for g1, g2 in F.pairs:
        bi, bj = int(F.scene.geoms[g1, 1]), int(F.scene.geoms[g2, 1])
        ref = O.collide(F.scene.geoms[g1], F.scene.geoms[g2], st[bi] if bi >= 0 else None,
                        st[bj] if bj >= 0 else None, maxc=4, dtype=np.float32)
        got_c = by_pair.get((g1, g2), [])
        assert len(got_c) == len(ref), (g1, g2, len(got_c), len(ref))
        for gc, rc in zip(got_c, ref):
            worst = max(worst, float(np.max(np.abs(np.array(gc[2] + gc[3] + [gc[4]]) - rc[0:7]))))
    assert worst < 2e-5, worst
    return worst


@pytest.mark.parametrize("space", ["sap", "hash", "simple"])
def test_space_collide_callback_order_matches_oracle_over_a_frame_loop(space):
    say("case: dSpaceCollide callback order", space)
    scene = settle_scene(3, 20)
    F = Frame(scene, space=space)
    for step in range(40):
        F.collide()
        want = F.expected_pairs()
        assert F.pairs == want, (step, F.pairs[:8], want[:8])
        if step % 10 == 0:
            check_contacts(F)
        assert F.step() == 1
        F.L.dJointGroupEmpty(F.group)
    F.close()


def test_space_collide_hash_levels_and_big_boxes():
    say("case: dHashSpace levels (-1, 0)")
    # with maxlevel 0 the larger boxes go to the big-box list
    scene = settle_scene(5, 24)
    F = Frame(scene, space="hash", levels=(-1, 0))
    for step in range(10):
        F.collide()
        assert F.pairs == F.expected_pairs(), step
        assert F.step() == 1
        F.L.dJointGroupEmpty(F.group)
    F.close()


def test_space_collide_enable_remove_bits_and_moves():
    say("case: dSpaceCollide with disabled, removed and filtered geoms")
    scene = settle_scene(6, 16)
    F = Frame(scene, space="sap")
    L, A = F.L, F.A
    ng = len(scene.geoms)
    enabled = np.ones(ng, np.uint8)
    cat = np.full(ng, np.iinfo(np.uint64).max, np.uint64)
    col = np.full(ng, np.iinfo(np.uint64).max, np.uint64)
    F.collide()
    assert F.pairs == F.expected_pairs()
    # a disabled geom stays in the lists and is skipped
    L.dGeomDisable(F.geoms[2])
    enabled[2] = 0
    # 64-bit category bits: geom 3 only meets geoms with collide bit 40
    cat[3] = np.uint64(1) << np.uint64(40)
    col[3] = np.uint64(1) << np.uint64(40)
    L.dGeomSetCategoryBits(F.geoms[3], int(cat[3]))
    L.dGeomSetCollideBits(F.geoms[3], int(col[3]))
    col[5] = np.uint64(1) << np.uint64(40) | np.uint64(1)
    L.dGeomSetCollideBits(F.geoms[5], int(col[5]))
    # bodies moved by the API, in this order
    for b in (4, 1, 4):
        p = A.vec(L.dBodyGetPosition(F.bodies[b]))
        L.dBodySetPosition(F.bodies[b], p[0] + 0.01, p[1], p[2])
        for gi in F.body_geoms[b]:
            F.model.moved(gi)
    F.collide()
    assert F.pairs == F.expected_pairs(enabled=enabled, cat=cat, col=col)
    # removed and re-added: appended to the dirty list again
    L.dSpaceRemove(F.space, F.geoms[0])
    F.model.remove(0)
    L.dSpaceAdd(F.space, F.geoms[0])
    F.model.add(0)
    F.collide()
    assert F.pairs == F.expected_pairs(enabled=enabled, cat=cat, col=col)
    assert F.step() == 1
    F.L.dJointGroupEmpty(F.group)
    F.collide()
    assert F.pairs == F.expected_pairs(enabled=enabled, cat=cat, col=col)
    F.close()


def quickstep_parity(F, dynamic=False):
    """One dWorldQuickStep against the float64 oracle's QuickStep of the same
    contacts in the reference's island order; returns (err_x, err_v, sweeps,
    oracle float32 sweeps)."""
    before = F.state()
    recs = []
    for (i1, i2, pos, nrm, depth, g1, g2) in F.contacts:
        r = np.zeros(16)
        r[1], r[2] = i1, i2
        r[3:6], r[6:9], r[9] = pos, nrm, depth
        recs.append(r)
    # dJointAttach with a world geom first: body1 NULL -> reversed joint
    norm_pairs = [(c[0], c[1]) if c[0] >= 0 else (c[1], -1) for c in F.contacts]
    _, order = island_order(len(F.bodies), norm_pairs)
    assert F.step() == 1
    got = F.state()
    sweeps, res, rows = ctypes.c_int(0), ctypes.c_float(0), ctypes.c_int(0)
    F.L.odxWorldGetSolveStats(F.w, ctypes.byref(sweeps), ctypes.byref(res), ctypes.byref(rows))
    # dWorldCreate's dynamic-iteration defaults (include/ode/objects.h:473-475)
    prm = O.params(**(dict(dyn_premature=1e-8, dyn_extra_factor=1.0, dyn_extra_delta=1e-2) if dynamic else {}))
    jr = np.array([recs[k] for k in order]).reshape(-1, 16)
    ref_b, ref_l = O.quickstep(before, jr, prm, dtype=np.float64)
    err_x = np.max(np.abs(got[:, 0:3] - ref_b[:, 0:3])) / np.max(np.abs(ref_b[:, 0:3]))
    err_v = np.max(np.abs(got[:, 8:15] - ref_b[:, 8:15])) / max(np.max(np.abs(ref_b[:, 8:15])), 1.0)
    # the sweep count (largest over islands) against the float32 oracle: float32
    # adjustments stall near rounding level, as in a dSINGLE build
    O.quickstep(before, jr, prm, dtype=np.float32)
    return err_x, err_v, sweeps.value, O.last_sweeps(), len(recs)


@pytest.mark.parametrize("dynamic", [False, True])
@pytest.mark.parametrize("seed", [1, 2])
def test_quickstep_matches_oracle_in_island_order(seed, dynamic):
    say("case: dWorldQuickStep seed", seed, "dynamic iterations", dynamic)
    scene = settle_scene(seed)
    F = Frame(scene, dynamic=dynamic)
    F.collide()
    assert F.pairs == F.expected_pairs()
    err_x, err_v, sweeps, want, nc = quickstep_parity(F, dynamic)
    say("  contacts", nc, "err_x", err_x, "err_v", err_v, "sweeps", sweeps, "oracle f32", want)
    assert err_x < 1e-4 and err_v < 1e-4
    if dynamic:
        assert abs(sweeps - want) <= 1, (sweeps, want)
    else:
        assert sweeps == 20
    F.L.dJointGroupEmpty(F.group)
    F.close()


def test_boxes_1000_frame_loop_matches_oracle():
    """The 1,000-box configuration through the drop-in API: sweep-and-prune
    space, near callback with dCollide (4 contacts, mu 0.5, Approx1) and
    dJointCreateContact, dWorldQuickStep with 20 iterations (fixed count),
    dJointGroupEmpty, 300 steps.  Every step's callback sequence equals the
    oracle's; every 30th step each callback's contacts equal the float32
    oracle's; the last step matches the float64 oracle's QuickStep."""
    say("case: 1000 boxes, 300 steps")
    scene = S.boxes_on_plane(1000, seed=1)
    F = Frame(scene, space="sap", iters=20)
    ncb = []
    for step in range(300):
        F.collide()
        want = F.expected_pairs()
        assert F.pairs == want, (step, len(F.pairs), len(want))
        ncb.append(len(F.pairs))
        if step % 30 == 29:
            worst = check_contacts(F)
            say("  step", step, "callbacks", len(F.pairs), "contacts", len(F.contacts), "worst", worst)
        if step == 299:
            err_x, err_v, sweeps, _, nc = quickstep_parity(F)
            say("  last step: contacts", nc, "err_x", err_x, "err_v", err_v, "sweeps", sweeps)
            assert nc > 1000
            assert err_x < 1e-4 and err_v < 1e-4
            assert sweeps == 20
        else:
            assert F.step() == 1
        F.L.dJointGroupEmpty(F.group)
    F.close()


def test_frame_loop_runs_and_stays_above_ground():
    say("case: frame loop")
    scene = settle_scene(4)
    F = Frame(scene)
    for _ in range(100):
        F.collide()
        assert F.L.dWorldQuickStep(F.w, 0.01) == 1
        F.L.dJointGroupEmpty(F.group)
    st = F.state()
    assert np.all(np.isfinite(st[:, 0:15]))
    assert np.min(st[:, 2]) > -0.05
    F.close()


def test_ball_joint_feedback_carries_the_weight():
    say("case: joint feedback")
    from ode_amd import ode as A
    L = A.api()
    w = L.dWorldCreate()
    L.dWorldSetGravity(w, 0, 0, -9.81)
    L.dWorldSetQuickStepNumIterations(w, 200)
    b = L.dBodyCreate(w)
    L.dBodySetPosition(b, 0, 0, 1)
    m = A.dMass()
    L.dMassSetSphere(ctypes.byref(m), 1000.0, 0.1)
    L.dBodySetMass(b, ctypes.byref(m))
    j = L.dJointCreateBall(w, None)
    L.dJointAttach(j, b, None)
    L.dJointSetBallAnchor(j, 0, 0, 1.2)
    fb = A.dJointFeedback()
    L.dJointSetFeedback(j, ctypes.byref(fb))
    for _ in range(50):
        assert L.dWorldQuickStep(w, 0.01) == 1
    # the joint holds the hanging body: force on body 1 balances gravity
    assert fb.f1[2] == pytest.approx(m.mass * 9.81, rel=0.02)
    pos = A.vec(L.dBodyGetPosition(b))
    assert abs(pos[2] - 1.0) < 0.01
    L.dWorldDestroy(w)


def test_dcollide_alone_and_drand():
    say("case: dCollide single pair, dRand")
    from ode_amd import ode as A
    L = A.api()
    s1 = L.dCreateSphere(None, 0.5)
    pl = L.dCreatePlane(None, 0, 0, 1, 0)
    L.dGeomSetPosition(s1, 1, 2, 0.4)
    c = (A.dContactGeom * 4)()
    n = L.dCollide(s1, pl, 4, c, ctypes.sizeof(A.dContactGeom))
    assert n == 1
    assert c[0].depth == pytest.approx(0.1, abs=1e-6)
    assert list(c[0].normal)[:3] == pytest.approx([0, 0, 1])
    n = L.dCollide(pl, s1, 4, c, ctypes.sizeof(A.dContactGeom))
    assert list(c[0].normal)[:3] == pytest.approx([0, 0, -1])
    L.dRandSetSeed(0)
    assert [L.dRand() for _ in range(5)] == O.drand_seq(0, 5)
    L.dGeomDestroy(s1)
    L.dGeomDestroy(pl)


def _dyn_frame(bodies, geoms, premature=1e-4):
    F = Frame(S.World(bodies, geoms), dynamic=True)
    A = F.A
    p, fac, xd = A.dReal(premature), A.dReal(1.0), A.dReal(1e-2)
    F.L.dWorldSetQuickStepDynamicIterationParameters(F.w, ctypes.byref(p), ctypes.byref(fac), ctypes.byref(xd))
    F.collide()
    assert F.L.dWorldQuickStep(F.w, 0.01) == 1
    sweeps, res, rows = ctypes.c_int(0), ctypes.c_float(0), ctypes.c_int(0)
    F.L.odxWorldGetSolveStats(F.w, ctypes.byref(sweeps), ctypes.byref(res), ctypes.byref(rows))
    st = F.state()
    F.L.dJointGroupEmpty(F.group)
    F.close()
    return sweeps.value, st


def test_dynamic_iterations_exit_per_island_at_the_closed_form_sweep():
    # one contact per sphere: the SOR iterate is scalar, lambda_k = L (1 - (1 - w)^k),
    # and the sweep-k adjustment is L w |1 - w|^(k-1) / m (w = 1.3); with the
    # premature delta 1e-4 a resting sphere stops after 11 sweeps and one falling
    # at 1 m/s (L 11x larger) after 13; together each island stops on its own
    say("case: dynamic iteration count per island")
    m, r = 10.0, 0.5
    I = 0.4 * m * r * r

    def sphere(x, vz):
        b = np.zeros(32)
        b[0:3] = (x, 0, r)
        b[4] = 1.0
        b[10] = vz
        b[24], b[25:28], b[31] = m, (I, I, I), 2
        return b

    plane = S.plane(0, 0, 1, 0)
    n_rest, st_rest = _dyn_frame([sphere(-3, 0)], [S.geom(S.SPHERE, 0, (r,)), plane])
    n_fall, st_fall = _dyn_frame([sphere(3, -1.0)], [S.geom(S.SPHERE, 0, (r,)), plane])
    n_both, st_both = _dyn_frame([sphere(-3, 0), sphere(3, -1.0)],
                                 [S.geom(S.SPHERE, 0, (r,)), S.geom(S.SPHERE, 1, (r,)), plane])
    say("  sweeps", n_rest, n_fall, n_both)
    assert (n_rest, n_fall, n_both) == (11, 13, 13)
    np.testing.assert_allclose(st_both[0, :15], st_rest[0, :15], rtol=0, atol=1e-6)
    np.testing.assert_allclose(st_both[1, :15], st_fall[0, :15], rtol=0, atol=1e-6)


def test_dynamic_iteration_statistics_sink_counts_exits_per_island():
    # two resting spheres, two islands.  With w = 1.3 and the premature delta
    # 1e-4 each stops after 11 of 20 sweeps: two premature exits per step.
    # With both deltas 0 no sweep meets the requirement (adjustment < 0) and
    # each uses all 20 extra sweeps.  With w = 1.9 the sweep-k adjustment of a
    # resting sphere, g w |1 - w|^(k-1), decays slowly enough to stay exact in
    # float32; a requirement between its values at sweeps 24 and 25 stops both
    # inside the extra sweeps (objects.h:538-576, quickstep.cpp:1829-1853)
    say("case: dynamic iteration statistics sink")
    from ode_amd import ode as A
    m, r = 10.0, 0.5
    I = 0.4 * m * r * r

    def sphere(x):
        b = np.zeros(32)
        b[0:3] = (x, 0, r)
        b[4] = 1.0
        b[24], b[25:28], b[31] = m, (I, I, I), 2
        return b

    class Stats(ctypes.Structure):
        _fields_ = [("struct_size", ctypes.c_uint), ("iteration_count", ctypes.c_uint),
                    ("premature_exits", ctypes.c_uint), ("prolonged_execs", ctypes.c_uint),
                    ("full_extra_execs", ctypes.c_uint)]

    adj = lambda k, w: 9.81 * w * abs(1 - w) ** (k - 1)
    # (the closed form holds for the first step from rest: the third case
    # runs that step only)
    cases = [((1e-4, 1.0, 1e-2), 1.3, (2, 0, 0), 3),
             ((0.0, 1.0, 0.0), 1.3, (0, 0, 2), 3),
             ((0.0, 1.0, 0.5 * (adj(24, 1.9) + adj(25, 1.9))), 1.9, (0, 2, 0), 1)]
    for (pre, fac, xd), w, want, steps in cases:
        F = Frame(S.World([sphere(-3), sphere(3)], [S.geom(S.SPHERE, 0, (r,)), S.geom(S.SPHERE, 1, (r,)),
                                                   S.plane(0, 0, 1, 0)]), dynamic=True)
        L = F.L
        L.dWorldAttachQuickStepDynamicIterationStatisticsSink.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        st = Stats()
        st.struct_size = ctypes.sizeof(Stats) - 4   # too small: refused
        assert L.dWorldAttachQuickStepDynamicIterationStatisticsSink(F.w, ctypes.byref(st)) == 0
        st.struct_size = ctypes.sizeof(Stats)
        assert L.dWorldAttachQuickStepDynamicIterationStatisticsSink(F.w, ctypes.byref(st)) == 1
        L.dWorldSetQuickStepW(F.w, w)
        p, f, x = A.dReal(pre), A.dReal(fac), A.dReal(xd)
        L.dWorldSetQuickStepDynamicIterationParameters(F.w, ctypes.byref(p), ctypes.byref(f), ctypes.byref(x))
        for step in range(steps):
            F.collide()
            assert L.dWorldQuickStep(F.w, 0.01) == 1
            L.dJointGroupEmpty(F.group)
        got = (st.premature_exits, st.prolonged_execs, st.full_extra_execs)
        say("  parameters", (pre, fac, xd), "statistics", got)
        assert got == tuple(steps * k for k in want)
        assert st.iteration_count == 0
        # detached: the world counts into its own statistics again
        assert L.dWorldAttachQuickStepDynamicIterationStatisticsSink(F.w, None) == 1
        F.collide()
        assert L.dWorldQuickStep(F.w, 0.01) == 1
        L.dJointGroupEmpty(F.group)
        assert (st.premature_exits, st.prolonged_execs, st.full_extra_execs) == got
        F.close()


@pytest.mark.parametrize("dynamic", [False, True])
def test_random_reordering_matches_oracle_and_draws_as_the_reference(dynamic):
    # row order 2: the reference's default build; the same dRand seed in the
    # library and the oracle gives the same shuffles, iterate and final seed
    say("case: random reordering, dynamic", dynamic)
    scene = settle_scene(3, 16)
    F = Frame(scene, dynamic=dynamic)
    F.L.odxWorldSetRowOrder(F.w, 2)
    F.collide()
    before = scene.bodies.copy()
    F.L.dRandSetSeed(4242)
    assert F.L.dWorldQuickStep(F.w, 0.01) == 1
    seed_gpu = int(F.L.dRandGetSeed())
    got = F.state()
    sweeps, res, rows = ctypes.c_int(0), ctypes.c_float(0), ctypes.c_int(0)
    F.L.odxWorldGetSolveStats(F.w, ctypes.byref(sweeps), ctypes.byref(res), ctypes.byref(rows))
    recs = []
    for (i1, i2, pos, nrm, depth, g1, g2) in F.contacts:
        r = np.zeros(16)
        r[1], r[2] = i1, i2
        r[3:6], r[6:9], r[9] = pos, nrm, depth
        recs.append(r)
    norm_pairs = [(a, b) if a >= 0 else (b, -1) for a, b in [(c[0], c[1]) for c in F.contacts]]
    _, order = island_order(len(scene.bodies), norm_pairs)
    jr = np.array([recs[k] for k in order]).reshape(-1, 16)
    prm = O.params(random_reorder=1, **(dict(dyn_premature=1e-8, dyn_extra_factor=1.0, dyn_extra_delta=1e-2)
                                        if dynamic else {}))
    O.rand_seed(4242)
    ref_b, ref_l = O.quickstep(before, jr, prm, dtype=np.float32)
    seed_ref = O.rand_seed()
    n_ref = O.last_sweeps()
    O.rand_seed(4242)
    ref64, _ = O.quickstep(before, jr, prm, dtype=np.float64)
    err_x = np.max(np.abs(got[:, 0:3] - ref64[:, 0:3])) / np.max(np.abs(ref64[:, 0:3]))
    err_v = np.max(np.abs(got[:, 8:15] - ref64[:, 8:15])) / max(np.max(np.abs(ref64[:, 8:15])), 1.0)
    say("  contacts", len(recs), "sweeps", sweeps.value, n_ref, "seeds", seed_gpu, seed_ref, "err", err_x, err_v)
    assert sweeps.value == n_ref
    assert seed_gpu == seed_ref
    assert seed_ref != 4242 or sweeps.value <= 8
    assert err_x < 1e-4 and err_v < 1e-4
    F.L.dJointGroupEmpty(F.group)
    F.close()


def _oracle_step_of_frame(F, before, prm, dtype=np.float64):
    """The oracle's QuickStep over the frame's contacts in the island order."""
    recs, pairs = [], []
    for (i1, i2, pos, nrm, depth, g1, g2) in F.contacts:
        r = np.zeros(16)
        r[1], r[2] = i1, i2
        r[3:6], r[6:9], r[9] = pos, nrm, depth
        recs.append(r)
        pairs.append((i1, i2) if i1 >= 0 else (i2, -1))
    _, order = island_order(len(before), pairs)
    jr = np.array([recs[k] for k in order]).reshape(-1, 16)
    return O.quickstep(before, jr, prm, dtype=dtype)[0]


def test_damping_and_angular_speed_cap_match_oracle():
    # dxStepBody's cap before the move and damping after it (util.cpp:583-699),
    # per body through the API, against the oracle's record flags and slots
    say("case: damping and angular speed cap")
    scene = settle_scene(2)
    rng = np.random.default_rng(9)
    scene.bodies[:, 12:15] = rng.normal(0, 3.0, size=(len(scene.bodies), 3))
    F = Frame(scene)
    L = F.L
    before = scene.bodies.astype(np.float64).copy()
    for i, b in enumerate(F.bodies):
        L.dBodySetAngularVel(b, *[float(x) for x in scene.bodies[i, 12:15]])
        kind = i % 4
        if kind == 0:
            L.dBodySetDamping(b, 0.1, 0.2)
            L.dBodySetLinearDampingThreshold(b, 0.0)
            L.dBodySetAngularDampingThreshold(b, 0.5)
            before[i, 31] = int(before[i, 31]) | 4 | 8
            before[i, 3], before[i, 11], before[i, 15], before[i, 19] = 0.1, 0.2, 0.0, 0.5
        elif kind == 1:
            L.dBodySetMaxAngularSpeed(b, 1.5)
            before[i, 31] = int(before[i, 31]) | 16
            before[i, 23] = 1.5
        elif kind == 2:
            L.dBodySetMaxAngularSpeed(b, 1.0)
            L.dBodySetAngularDamping(b, 0.3)
            L.dBodySetAngularDampingThreshold(b, 0.0)
            before[i, 31] = int(before[i, 31]) | 8 | 16
            before[i, 11], before[i, 19], before[i, 23] = 0.3, 0.0, 1.0
    # getters return what was set (thresholds through their squares)
    assert L.dBodyGetLinearDamping(F.bodies[0]) == pytest.approx(0.1)
    assert L.dBodyGetAngularDampingThreshold(F.bodies[0]) == pytest.approx(0.5)
    assert L.dBodyGetMaxAngularSpeed(F.bodies[1]) == pytest.approx(1.5)
    F.collide()
    assert L.dWorldQuickStep(F.w, 0.01) == 1
    got = F.state()
    ref = _oracle_step_of_frame(F, before, O.params())
    err_x = np.max(np.abs(got[:, 0:3] - ref[:, 0:3])) / np.max(np.abs(ref[:, 0:3]))
    vel = np.r_[8:11, 12:15]  # slot 11 holds the angular damping scale, not a velocity
    err_v = np.max(np.abs(got[:, vel] - ref[:, vel])) / max(np.max(np.abs(ref[:, vel])), 1.0)
    say("  contacts", len(F.contacts), "err_x", err_x, "err_v", err_v)
    assert err_x < 1e-4 and err_v < 1e-4
    # capped bodies end at or under their cap (damping after it only lowers it)
    for i in range(len(F.bodies)):
        if i % 4 in (1, 2):
            assert np.linalg.norm(got[i, 12:15]) <= (1.5 if i % 4 == 1 else 1.0) * (1 + 1e-6)
    F.L.dJointGroupEmpty(F.group)
    F.close()


def test_world_damping_and_auto_disable_defaults_reach_new_bodies():
    say("case: world defaults")
    from ode_amd import ode as A
    L = A.api()
    w = L.dWorldCreate()
    b0 = L.dBodyCreate(w)
    # dWorldCreate defaults (objects.cpp:54-69, 99-117)
    assert L.dBodyGetAutoDisableFlag(b0) == 0 and L.dBodyGetAutoDisableSteps(b0) == 10
    assert L.dBodyGetAutoDisableAverageSamplesCount(b0) == 1
    assert L.dBodyGetAutoDisableLinearThreshold(b0) == pytest.approx(0.01)
    assert L.dBodyGetLinearDamping(b0) == 0 and L.dBodyGetMaxAngularSpeed(b0) == float("inf")
    L.dWorldSetAutoDisableFlag(w, 1)
    L.dWorldSetAutoDisableSteps(w, 7)
    L.dWorldSetAutoDisableTime(w, 0.25)
    L.dWorldSetAutoDisableAverageSamplesCount(w, 4)
    L.dWorldSetDamping(w, 0.05, 0.125)
    L.dWorldSetLinearDampingThreshold(w, 0.5)
    L.dWorldSetMaxAngularSpeed(w, 3.0)
    b1 = L.dBodyCreate(w)
    assert L.dBodyGetAutoDisableFlag(b1) == 1 and L.dBodyGetAutoDisableSteps(b1) == 7
    assert L.dBodyGetAutoDisableTime(b1) == pytest.approx(0.25)
    assert L.dBodyGetAutoDisableAverageSamplesCount(b1) == 4
    assert L.dBodyGetLinearDamping(b1) == pytest.approx(0.05)
    assert L.dBodyGetAngularDamping(b1) == pytest.approx(0.125)
    assert L.dBodyGetLinearDampingThreshold(b1) == pytest.approx(0.5)
    assert L.dBodyGetMaxAngularSpeed(b1) == pytest.approx(3.0)
    # bodies made earlier keep their own settings
    assert L.dBodyGetAutoDisableFlag(b0) == 0 and L.dBodyGetLinearDamping(b0) == 0
    # clearing a body's flag takes the world's counts again and enables it
    L.dBodySetAutoDisableSteps(b1, 2)
    L.dBodyDisable(b1)
    L.dBodySetAutoDisableFlag(b1, 0)
    assert L.dBodyIsEnabled(b1) == 1 and L.dBodyGetAutoDisableSteps(b1) == 7
    L.dWorldDestroy(w)


def test_auto_disable_matches_oracle_through_sleep_and_wake():
    # four spheres resting on the plane fall asleep after 5 idle steps; a fifth
    # rolls into the first, whose contact with the awake one wakes it; the
    # library's enabled flags follow the oracle's auto-disable pass fed with
    # the library's own velocities, step by step
    from oracle import autodisable as AD
    say("case: auto-disable sleep and wake")
    m, r = 10.0, 0.25
    I = 0.4 * m * r * r

    def sphere(x, vx=0.0):
        b = np.zeros(32)
        b[0:3] = (x, 0.0, r)
        b[4] = 1.0
        b[8] = vx
        b[24], b[25:28], b[31] = m, (I, I, I), 2
        return b

    bodies = [sphere(0.0), sphere(2.0), sphere(4.0), sphere(6.0), sphere(-0.6, 1.5)]
    geoms = [S.geom(S.SPHERE, i, (r,)) for i in range(5)] + [S.plane(0, 0, 1, 0)]
    F = Frame(S.World(np.array(bodies), np.array(geoms)))
    L = F.L
    # the step budget set after creation takes effect at the next reset
    # (dBodyEnable restarts the countdown), as in the reference
    ref = [AD.Body(idle_steps=10) for _ in F.bodies]
    for b, rb in zip(F.bodies, ref):
        L.dBodySetAutoDisableFlag(b, 1)
        L.dBodySetAutoDisableSteps(b, 5)
        L.dBodyEnable(b)
        rb.set_idle_steps(5)
        rb.enable()
    n = len(F.bodies)
    history = []
    for step in range(1, 41):
        F.collide()
        lv = np.array([F.A.vec(L.dBodyGetLinearVel(b)) for b in F.bodies], np.float32)
        av = np.array([F.A.vec(L.dBodyGetAngularVel(b)) for b in F.bodies], np.float32)
        joints = [(c[0], c[1]) for c in F.contacts]
        has = [any(i in j for j in joints) for i in range(n)]
        AD.handle(ref, lv, av, has, 0.01)
        AD.enabled_after_islands(ref, joints)
        pos_before = F.state()[:, 0:3]
        assert L.dWorldQuickStep(F.w, 0.01) == 1
        got = [L.dBodyIsEnabled(b) for b in F.bodies]
        want = [int(b.enabled) for b in ref]
        assert got == want, (step, got, want)
        st = F.state()
        for i in range(n):
            if not got[i]:
                # asleep: zero velocities, pose untouched
                assert np.all(st[i, 8:15] == 0)
                np.testing.assert_array_equal(st[i, 0:3], pos_before[i])
        history.append(got)
        L.dJointGroupEmpty(F.group)
    h = np.array(history)
    say("  sphere 0 enabled per step:", "".join(map(str, h[:, 0])))
    # the resting spheres fall asleep on the 5th step, exactly
    assert np.all(h[:4, 1:4] == 1) and np.all(h[4:, 1:4] == 0)
    # sphere 0 sleeps from step 5 until the rolling sphere reaches it
    assert h[4, 0] == 0
    woke = np.nonzero(h[5:, 0])[0]
    assert len(woke) > 0, "the rolling sphere never woke sphere 0"
    F.close()


@pytest.mark.parametrize("dynamic", [False, True])
def test_sweep_modes_give_identical_steps(dynamic):
    # the workgroup-wide, the one-wave and the island-group level sweeps run
    # the same Gauss-Seidel order (a level's items touch disjoint bodies, and
    # islands share none): bit-equal states over a 30-step frame loop (the
    # island groups need a fixed sweep count; with the dynamic count mode 3
    # falls back to the automatic choice)
    say("case: sweep modes, dynamic", dynamic)
    out = []
    for mode in (1, 2, 3):
        scene = settle_scene(6, 30)
        F = Frame(scene, dynamic=dynamic)
        F.L.odxWorldSetSweepMode(F.w, mode)
        for _ in range(30):
            F.collide()
            assert F.L.dWorldQuickStep(F.w, 0.01) == 1
            F.L.dJointGroupEmpty(F.group)
        out.append(F.state())
        F.close()
    np.testing.assert_array_equal(out[0], out[1])
    np.testing.assert_array_equal(out[0], out[2])


def test_quickstep_matches_oracle_with_rows_beyond_lds():
    # ~2,000 rows (over 200 KB) do not fit in LDS next to the rest: the step
    # streams rows from global memory; small worlds keep them in LDS
    say("case: dWorldQuickStep, rows streamed from global memory")
    rng = np.random.default_rng(12)
    scene = S.uneven_world(rng, 300, 300)
    scene.bodies[:, 0:2] *= 4.0
    scene.bodies[:, 2] = np.minimum(scene.bodies[:, 2], 0.35) + rng.uniform(0, 0.05, size=len(scene.bodies))
    F = Frame(scene)
    F.collide()
    before = scene.bodies.copy()
    assert F.L.dWorldQuickStep(F.w, 0.01) == 1
    got = F.state()
    sweeps, res, rows = ctypes.c_int(0), ctypes.c_float(0), ctypes.c_int(0)
    F.L.odxWorldGetSolveStats(F.w, ctypes.byref(sweeps), ctypes.byref(res), ctypes.byref(rows))
    say("  rows", rows.value, "contacts", len(F.contacts))
    assert rows.value * 128 > 160 * 1024
    ref = _oracle_step_of_frame(F, before, O.params())
    err_x = np.max(np.abs(got[:, 0:3] - ref[:, 0:3])) / np.max(np.abs(ref[:, 0:3]))
    vel = np.r_[8:11, 12:15]
    err_v = np.max(np.abs(got[:, vel] - ref[:, vel])) / max(np.max(np.abs(ref[:, vel])), 1.0)
    say("  err_x", err_x, "err_v", err_v)
    assert err_x < 1e-4 and err_v < 1e-4
    F.L.dJointGroupEmpty(F.group)
    F.close()


def test_export_dif_matches_oracle_text(tmp_path):
    # dWorldExportDIF of a world with flags, several geoms per body, a ball
    # joint and contact joints (one reversed by a NULL body1), against the
    # oracle's restatement of export-dif.cpp
    from oracle import dif
    from ode_amd import ode as A
    say("case: dWorldExportDIF")
    L = A.api()
    L.dWorldExportDIF.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_char_p]
    L.dGeomSetCategoryBits.argtypes = [ctypes.c_void_p, ctypes.c_ulong]
    L.dGeomSetCollideBits.argtypes = [ctypes.c_void_p, ctypes.c_ulong]
    L.dGeomDisable.argtypes = [ctypes.c_void_p]
    L.dBodyDisable.argtypes = [ctypes.c_void_p]
    libc = ctypes.CDLL(None)
    libc.fopen.restype = ctypes.c_void_p
    libc.fopen.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    libc.fclose.argtypes = [ctypes.c_void_p]
    w = L.dWorldCreate()
    L.dWorldSetGravity(w, 0, 0, -9.81)
    space = L.dSimpleSpaceCreate(None)
    group = L.dJointGroupCreate(0)
    F32 = np.float32
    specs = [((0.5, -1.25, 2.0), (1, 0, 0, 0), 2.5), ((0.1, 0.2, 0.7), (0.8, 0.6, 0, 0), 1.5), ((3, 3, 3), (1, 0, 0, 0), 4.0)]
    bodies = []
    for pos, q, mass in specs:
        b = L.dBodyCreate(w)
        L.dBodySetPosition(b, *pos)
        L.dBodySetQuaternion(b, (A.dReal * 4)(*q))
        L.dBodySetLinearVel(b, 0.25, -0.5, 1.0)
        m = A.dMass()
        L.dMassSetBox(ctypes.byref(m), 1000.0 * mass, 0.5, 0.25, 0.75)
        L.dBodySetMass(b, ctypes.byref(m))
        bodies.append((b, m))
    L.dBodySetGravityMode(bodies[1][0], 0)
    L.dBodySetAutoDisableFlag(bodies[2][0], 1)
    L.dBodySetAutoDisableSteps(bodies[2][0], 4)
    L.dBodyDisable(bodies[2][0])
    L.dBodyAddTorque(bodies[0][0], 1.0, 2.0, 3.0)
    g0 = L.dCreateSphere(space, 0.3)
    L.dGeomSetBody(g0, bodies[0][0])
    g1 = L.dCreateBox(space, 0.5, 0.25, 0.75)
    L.dGeomSetBody(g1, bodies[0][0])
    L.dGeomSetCategoryBits(g1, 5)
    g2 = L.dCreateCapsule(space, 0.1, 0.6)
    L.dGeomSetBody(g2, bodies[1][0])
    L.dGeomDisable(g2)
    L.dCreatePlane(space, 0, 0, 1, 0)
    jb = L.dJointCreateBall(w, None)
    L.dJointAttach(jb, bodies[0][0], bodies[1][0])
    L.dJointSetBallAnchor(jb, 0.3, -0.5, 1.5)
    cs = []
    for k, (b1, b2, mode) in enumerate([(bodies[0][0], None, 0x1000 | 0x004 | 0x001),
                                        (None, bodies[1][0], 0x1000 | 0x2000 | 0x010 | 0x002)]):
        c = A.dContact()
        c.surface.mode = mode
        c.surface.mu = 0.5 + k
        c.surface.mu2 = 0.25
        c.surface.bounce, c.surface.bounce_vel, c.surface.soft_cfm = 0.3, 0.1, 1e-3
        c.geom.pos[0], c.geom.pos[1], c.geom.pos[2] = 1.0, 2.0, 0.0
        c.geom.normal[2] = 1.0
        c.geom.depth = 0.015
        c.fdir1[0] = 1.0
        j = L.dJointCreateContact(w, group, ctypes.byref(c))
        L.dJointAttach(j, b1, b2)
        cs.append(c)
    path = str(tmp_path / "world.dif").encode()
    f = libc.fopen(path, b"w")
    L.dWorldExportDIF(w, f, b"X")
    libc.fclose(f)
    got = open(path).read()

    # the oracle's description of the same world: body list newest first
    adis0 = {"lin_thr2": F32(0.01) * F32(0.01), "ang_thr2": F32(0.01) * F32(0.01), "samples": 1, "idle_time": 0.0,
             "idle_steps": 10}
    world = {"gravity": (0.0, 0.0, F32(-9.81)), "erp": 0.2, "cfm": 1e-5, "adis": adis0}
    order = [2, 1, 0]
    geoms = {0: [{"cls": 1, "p": (0.5, 0.25, 0.75, 0), "cat": 5, "col": dif.ALL_BITS, "enabled": True},
                 {"cls": 0, "p": (0.3, 0, 0, 0), "cat": dif.ALL_BITS, "col": dif.ALL_BITS, "enabled": True}],
             1: [{"cls": 2, "p": (0.1, 0.6, 0, 0), "cat": dif.ALL_BITS, "col": dif.ALL_BITS, "enabled": False}],
             2: []}
    desc = []
    for i in order:
        b, m = bodies[i]
        d = {"pos": A.vec(L.dBodyGetPosition(b)), "q": A.vec(L.dBodyGetQuaternion(b), 4),
             "lvel": A.vec(L.dBodyGetLinearVel(b)), "avel": A.vec(L.dBodyGetAngularVel(b)), "mass": m.mass,
             "I": list(m.I), "c": list(m.c)[:3], "enabled": i != 2, "gravity": i != 1,
             "adis": None, "facc": (0.0, 0.0, 0.0), "tacc": (1.0, 2.0, 3.0) if i == 0 else (0.0, 0.0, 0.0),
             "geoms": geoms[i]}
        if i == 2:
            d["adis"] = dict(adis0, idle_steps=4, time_left=0.0, steps_left=10)
        desc.append(d)
    tag = {i: k for k, i in enumerate(order)}
    joints = []
    # joints newest first: the second contact (attached NULL, body 1 -> reversed to (1)), the first, the ball
    c1, c0 = cs[1], cs[0]
    joints.append({"type": "contact", "bodies": (tag[1], None), "pos": (1.0, 2.0, 0.0), "normal": (0.0, 0.0, 1.0),
                   "depth": 0.015, "mode": c1.surface.mode, "mu": 1.5, "mu2": 0.25, "bounce": 0.3, "bounce_vel": 0.1,
                   "soft_cfm": 1e-3, "fdir1": (1.0, 0.0, 0.0)})
    joints.append({"type": "contact", "bodies": (tag[0], None), "pos": (1.0, 2.0, 0.0), "normal": (0.0, 0.0, 1.0),
                   "depth": 0.015, "mode": c0.surface.mode, "mu": 0.5, "mu2": 0.25, "bounce": 0.3, "bounce_vel": 0.1,
                   "soft_cfm": 1e-3, "fdir1": (1.0, 0.0, 0.0)})
    ball_a1 = np.zeros(3)
    ball_a2 = np.zeros(3)
    # anchors in each body's frame: the library computed them from the world point
    for line in got.split("\n"):
        if line.startswith("\tanchor1 = "):
            ball_a1 = [float(x) for x in line.split("{")[1].split("}")[0].split(",")]
        if line.startswith("\tanchor2 = "):
            ball_a2 = [float(x) for x in line.split("{")[1].split("}")[0].split(",")]
    joints.append({"type": "ball", "bodies": (tag[0], tag[1]), "anchor1": ball_a1, "anchor2": ball_a2})
    want = dif.export(world, desc, joints, prefix="X")
    if got != want:
        import difflib
        say("".join(list(difflib.unified_diff(want.splitlines(True), got.splitlines(True)))[:60]))
    assert got == want
    L.dJointGroupDestroy(group)
    L.dSpaceDestroy(space)
    L.dWorldDestroy(w)
