"""Hinge, slider and fixed joints through the drop-in API: the closed forms
the oracle is pinned by (tests/test_oracle_joints.py), the API's own getters,
and one step of a mixed scene against the float64 oracle in island order."""
import ctypes
import math
import sys

import numpy as np
import pytest

import oracle as O
from oracle.islands import island_order

pytestmark = pytest.mark.gpu
INF = float("inf")
H = 0.01


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class World:
    def __init__(self, erp=0.0, cfm=0.0, iters=200, gravity=(0.0, 0.0, 0.0)):
        from ode_amd import ode as A
        self.A, self.L = A, A.api()
        L = self.L
        for name, res, args in (("dJointCreateHinge", ctypes.c_void_p, [ctypes.c_void_p, ctypes.c_void_p]),
                                ("dJointCreateSlider", ctypes.c_void_p, [ctypes.c_void_p, ctypes.c_void_p]),
                                ("dJointCreateFixed", ctypes.c_void_p, [ctypes.c_void_p, ctypes.c_void_p]),
                                ("dJointSetFixed", None, [ctypes.c_void_p])):
            getattr(L, name).restype = res
            getattr(L, name).argtypes = args
        f = A.dReal
        for name in ("dJointSetHingeAnchor", "dJointSetHingeAxis", "dJointSetSliderAxis"):
            getattr(L, name).argtypes = [ctypes.c_void_p, f, f, f]
        for name in ("dJointSetHingeParam", "dJointSetSliderParam", "dJointSetFixedParam"):
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_int, f]
        for name in ("dJointGetHingeAngle", "dJointGetHingeAngleRate", "dJointGetSliderPosition",
                     "dJointGetSliderPositionRate"):
            getattr(L, name).restype = f
            getattr(L, name).argtypes = [ctypes.c_void_p]
        for name in ("dJointGetHingeParam", "dJointGetSliderParam"):
            getattr(L, name).restype = f
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_int]
        for name in ("dJointGetHingeAnchor", "dJointGetHingeAxis", "dJointGetSliderAxis"):
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.POINTER(f)]
        self.w = L.dWorldCreate()
        L.dWorldSetGravity(self.w, *gravity)
        L.dWorldSetERP(self.w, erp)
        L.dWorldSetCFM(self.w, cfm)
        L.dWorldSetQuickStepNumIterations(self.w, iters)
        L.dWorldSetQuickStepW(self.w, 1.0)
        z = A.dReal(0.0)
        L.dWorldSetQuickStepDynamicIterationParameters(self.w, ctypes.byref(z), ctypes.byref(z), ctypes.byref(z))
        self.bodies = []

    def sphere(self, pos, m=2.0, r=0.5):
        L, A = self.L, self.A
        b = L.dBodyCreate(self.w)
        L.dBodySetPosition(b, *pos)
        ms = A.dMass()
        ii = 0.4 * m * r * r
        L.dMassSetParameters(ctypes.byref(ms), m, 0, 0, 0, ii, ii, ii, 0, 0, 0)
        L.dBodySetMass(b, ctypes.byref(ms))
        self.bodies.append(b)
        return b

    def state(self, b):
        A, L = self.A, self.L
        return np.array(A.vec(L.dBodyGetLinearVel(b)) + A.vec(L.dBodyGetAngularVel(b)))

    def step(self):
        assert self.L.dWorldQuickStep(self.w, H) == 1

    def close(self):
        self.L.dWorldDestroy(self.w)


def test_hinge_closed_forms_through_the_api():
    say("case: hinge closed forms")
    W = World()
    L = W.L
    b = W.sphere((0.0, 0.0, 0.0))
    j = L.dJointCreateHinge(W.w, None)
    L.dJointAttach(j, b, None)
    L.dJointSetHingeAnchor(j, 0.0, 0.0, 0.0)
    L.dJointSetHingeAxis(j, 0.0, 0.0, 1.0)
    assert abs(L.dJointGetHingeAngle(j)) < 1e-7
    ax = (W.A.dReal * 4)()
    L.dJointGetHingeAxis(j, ax)
    assert list(ax)[:3] == pytest.approx([0.0, 0.0, 1.0])
    L.dBodySetLinearVel(b, 0.5, -1.0, 0.25)
    L.dBodySetAngularVel(b, 1.0, 2.0, 3.0)
    W.step()
    np.testing.assert_allclose(W.state(b), [0, 0, 0, 0, 0, 3.0], atol=2e-6)
    assert L.dJointGetHingeAngleRate(j) == pytest.approx(3.0, rel=1e-6)
    assert L.dJointGetHingeAngle(j) == pytest.approx(3.0 * H, rel=1e-4)
    # the motor reaches its velocity in one step
    L.dJointSetHingeParam(j, 5, 1e9)  # dParamFMax
    L.dJointSetHingeParam(j, 2, 2.0)  # dParamVel
    assert L.dJointGetHingeParam(j, 2) == pytest.approx(2.0)
    W.step()
    assert W.state(b)[5] == pytest.approx(2.0, abs=2e-6)
    W.close()


@pytest.mark.parametrize("side", [1, -1])
def test_hinge_limit_through_the_api(side):
    say("case: hinge limit", side)
    W = World()
    L = W.L
    b = W.sphere((0.0, 0.0, 0.0))
    j = L.dJointCreateHinge(W.w, None)
    L.dJointAttach(j, b, None)
    L.dJointSetHingeAnchor(j, 0.0, 0.0, 0.0)
    L.dJointSetHingeAxis(j, 0.0, 0.0, 1.0)
    a = 0.5 * side
    L.dBodySetQuaternion(b, (W.A.dReal * 4)(math.cos(a / 2), 0.0, 0.0, math.sin(a / 2)))
    assert L.dJointGetHingeAngle(j) == pytest.approx(a, rel=1e-5)
    L.dJointSetHingeParam(j, 1 if side > 0 else 0, 0.3 * side)  # dParamHiStop / dParamLoStop
    L.dJointSetHingeParam(j, 9, 0.2)   # dParamStopERP
    L.dJointSetHingeParam(j, 10, 0.0)  # dParamStopCFM
    W.step()
    assert W.state(b)[5] == pytest.approx(-side * 0.2 / H * 0.2, rel=1e-4)
    W.close()


def test_slider_and_fixed_closed_forms_through_the_api():
    say("case: slider and fixed closed forms")
    W = World()
    L = W.L
    s = W.sphere((0.0, 0.0, 0.0))
    js = L.dJointCreateSlider(W.w, None)
    L.dJointAttach(js, s, None)
    L.dJointSetSliderAxis(js, 1.0, 0.0, 0.0)
    L.dBodySetLinearVel(s, 1.0, 2.0, 3.0)
    L.dBodySetAngularVel(s, 0.5, -0.5, 0.5)
    f = W.sphere((5.0, 0.0, 0.0))
    jf = L.dJointCreateFixed(W.w, None)
    L.dJointAttach(jf, f, None)
    L.dJointSetFixed(jf)
    L.dBodySetLinearVel(f, 1.0, 2.0, 3.0)
    L.dBodySetAngularVel(f, 3.0, -2.0, 1.0)
    W.step()
    np.testing.assert_allclose(W.state(s), [1.0, 0, 0, 0, 0, 0], atol=2e-6)
    np.testing.assert_allclose(W.state(f), 0.0, atol=2e-6)
    assert L.dJointGetSliderPosition(js) == pytest.approx(1.0 * H, rel=1e-4)
    assert L.dJointGetSliderPositionRate(js) == pytest.approx(1.0, rel=1e-6)
    # the slider's high stop: 0.01 past 0.0 pushes back at stop_erp / h * err
    L.dJointSetSliderParam(js, 1, 0.0)
    L.dJointSetSliderParam(js, 9, 0.2)
    L.dJointSetSliderParam(js, 10, 0.0)
    L.dBodySetLinearVel(s, 0.0, 0.0, 0.0)
    W.step()
    assert W.state(s)[0] == pytest.approx(-0.2 / H * 0.01, rel=1e-3)
    W.close()


def test_fixed_pair_moves_as_one_body():
    say("case: fixed pair")
    W = World(iters=2000)
    L = W.L
    a = W.sphere((0.0, 0.0, 0.0))
    b = W.sphere((1.0, 0.0, 0.0))
    j = L.dJointCreateFixed(W.w, None)
    L.dJointAttach(j, a, b)
    L.dJointSetFixed(j)
    L.dBodySetLinearVel(a, 0.0, 2.0, 0.0)
    W.step()
    m, I = 2.0, 0.4 * 2.0 * 0.25
    wz = -(m * 2.0 * 0.5) / (2 * I + 2 * m * 0.25)
    sa, sb = W.state(a), W.state(b)
    assert sa[5] == pytest.approx(wz, rel=1e-4) and sb[5] == pytest.approx(wz, rel=1e-4)
    assert sa[1] == pytest.approx(1.0 - 0.5 * wz, rel=1e-4) and sb[1] == pytest.approx(1.0 + 0.5 * wz, rel=1e-4)
    W.close()


def test_mixed_joint_scene_matches_oracle():
    # a chain of five hinged spheres hanging from the world (a motor on one
    # hinge, a limit active on another), a slider with its motor, a fixed
    # pair, under gravity with the world's ERP/CFM; one step against the
    # float64 oracle in the island order.  Joints are set up with every body
    # at identity, so the joint frames are the world's.
    say("case: mixed joint scene")
    rng = np.random.default_rng(4)
    W = World(erp=0.2, cfm=1e-5, iters=30, gravity=(0.0, 0.0, -9.81))
    L = W.L
    bodies, joints, specs = [], [], []
    for i in range(5):
        bodies.append(W.sphere((float(i), 0.0, 2.0), m=1.0 + 0.2 * i, r=0.3))
    for i in range(5):
        j = L.dJointCreateHinge(W.w, None)
        anchor = (i - 0.5, 0.0, 2.0)
        if i == 0:
            L.dJointAttach(j, bodies[0], None)
        else:
            L.dJointAttach(j, bodies[i - 1], bodies[i])
        L.dJointSetHingeAnchor(j, *anchor)
        L.dJointSetHingeAxis(j, 0.0, 1.0, 0.0)
        lim = {}
        if i == 2:
            L.dJointSetHingeParam(j, 5, 3.0)
            L.dJointSetHingeParam(j, 2, -1.0)
            lim = dict(fmax=3.0, vel=-1.0)
        if i == 3:
            L.dJointSetHingeParam(j, 0, 0.1)   # lostop above the angle 0: the limit row is active
            lim = dict(lostop=0.1)
        joints.append(j)
        specs.append(("hinge", i - 1 if i else 0, i if i else -1, anchor, lim))
    s = W.sphere((0.0, 4.0, 1.0), m=2.0, r=0.4)
    bodies.append(s)
    jsl = L.dJointCreateSlider(W.w, None)
    L.dJointAttach(jsl, s, None)
    L.dJointSetSliderAxis(jsl, 1.0, 0.0, 1.0)
    L.dJointSetSliderParam(jsl, 5, 50.0)
    L.dJointSetSliderParam(jsl, 2, 0.5)
    joints.append(jsl)
    specs.append(("slider", 5, -1, None, dict(fmax=50.0, vel=0.5)))
    fa, fb = W.sphere((3.0, 4.0, 1.0), m=1.5, r=0.3), W.sphere((3.5, 4.5, 1.5), m=1.0, r=0.3)
    bodies += [fa, fb]
    jfx = L.dJointCreateFixed(W.w, None)
    L.dJointAttach(jfx, fa, fb)
    L.dJointSetFixed(jfx)
    joints.append(jfx)
    specs.append(("fixed", 6, 7, None, {}))
    # perturb the velocities (joint frames stay those of the setup)
    for b in bodies:
        v = rng.normal(0, 0.5, 3)
        w = rng.normal(0, 1.0, 3)
        L.dBodySetLinearVel(b, *v)
        L.dBodySetAngularVel(b, *w)
    before = []
    for b in bodies:
        rec = np.zeros(32)
        rec[0:3] = W.A.vec(L.dBodyGetPosition(b))
        rec[4:8] = W.A.vec(L.dBodyGetQuaternion(b), 4)
        rec[8:11] = W.A.vec(L.dBodyGetLinearVel(b))
        rec[12:15] = W.A.vec(L.dBodyGetAngularVel(b))
        m = W.A.dMass()
        L.dBodyGetMass(b, ctypes.byref(m))
        rec[24], rec[25:28], rec[31] = m.mass, (m.I[0], m.I[5], m.I[10]), 2
        before.append(rec)
    before = np.array(before)
    W.step()
    got = np.array([np.r_[W.A.vec(L.dBodyGetPosition(b)), 0, W.A.vec(L.dBodyGetQuaternion(b), 4), W.state(b)[0:3], 0,
                          W.state(b)[3:6]] for b in bodies])
    # the oracle's records: joint frames at identity orientation
    recs, exts = [], []
    for kind, b1, b2, anchor, lim in specs:
        rec = np.zeros(16)
        x = np.zeros(32)
        x[6] = 1.0
        x[16], x[17], x[20] = -INF, INF, 1.0
        x[21], x[22], x[23] = 1e-5, 0.2, 1e-5
        x[26], x[27] = 0.2, 1e-5
        p1 = before[b1, 0:3]
        if kind == "hinge":
            rec[0] = 2
            rec[3:6] = np.array(anchor) - p1
            rec[6:9] = np.array(anchor) - before[b2, 0:3] if b2 >= 0 else anchor
            rec[9], rec[10] = 0.2, 1e-5
            x[0:3] = x[3:6] = (0.0, 1.0, 0.0)
        elif kind == "slider":
            rec[0] = 3
            x[0:3] = np.array([1.0, 0.0, 1.0]) / math.sqrt(2.0)
            x[10:13] = p1
        else:
            rec[0] = 4
            x[10:13] = p1 - before[b2, 0:3]
        rec[1], rec[2] = b1, b2
        x[16] = lim.get("lostop", -INF)
        x[18], x[19] = lim.get("vel", 0.0), lim.get("fmax", 0.0)
        recs.append(rec)
        exts.append(x)
    pairs = [(int(r[1]), int(r[2])) for r in recs]
    _, order = island_order(len(bodies), pairs)
    ref, _ = O.quickstep(before, np.array([recs[k] for k in order]), O.params(iters=30, w=1.0), dtype=np.float64,
                         jext=np.array([exts[k] for k in order]))
    vel = np.r_[8:11, 12:15]
    err_x = np.max(np.abs(got[:, 0:3] - ref[:, 0:3])) / np.max(np.abs(ref[:, 0:3]))
    err_q = np.max(np.abs(got[:, 4:8] - ref[:, 4:8]))
    err_v = np.max(np.abs(got[:, vel] - ref[:, vel])) / max(np.max(np.abs(ref[:, vel])), 1.0)
    say("  err_x", err_x, "err_q", err_q, "err_v", err_v)
    assert err_x < 1e-4 and err_q < 1e-4 and err_v < 1e-4
    W.close()


def test_export_dif_of_hinge_slider_fixed_matches_oracle_text(tmp_path):
    # printHinge / printSlider / printFixed with printLimot (export-dif.cpp:
    # 142-338); the joint frames are computed as the library's setters do, in
    # float32, from bodies at identity orientation
    from oracle import dif
    say("case: DIF of hinge, slider, fixed")
    F = np.float32
    W = World(erp=0.2, cfm=1e-5)
    L = W.L
    L.dWorldExportDIF.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_char_p]
    libc = ctypes.CDLL(None)
    libc.fopen.restype = ctypes.c_void_p
    libc.fopen.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    libc.fclose.argtypes = [ctypes.c_void_p]
    x0, x1 = (0.5, -1.0, 2.0), (1.5, 0.25, 2.0)
    b0, b1 = W.sphere(x0), W.sphere(x1)
    jh = L.dJointCreateHinge(W.w, None)
    L.dJointAttach(jh, b0, b1)
    L.dJointSetHingeAnchor(jh, 1.0, -0.5, 2.0)
    L.dJointSetHingeAxis(jh, 0.0, 3.0, 4.0)
    L.dJointSetHingeParam(jh, 0, -0.5)
    L.dJointSetHingeParam(jh, 1, 0.5)
    L.dJointSetHingeParam(jh, 5, 2.0)
    L.dJointSetHingeParam(jh, 2, 1.0)
    L.dJointSetHingeParam(jh, 7, 0.25)
    js = L.dJointCreateSlider(W.w, None)
    L.dJointAttach(js, b1, None)
    L.dJointSetSliderAxis(js, 1.0, 0.0, 0.0)
    jf = L.dJointCreateFixed(W.w, None)
    L.dJointAttach(jf, b0, b1)
    L.dJointSetFixed(jf)
    path = str(tmp_path / "joints.dif").encode()
    f = libc.fopen(path, b"w")
    L.dWorldExportDIF(W.w, f, b"")
    libc.fclose(f)
    got = open(path).read()

    def qmul(qb, qc, cb, cc):
        qb, qc = [F(v) for v in qb], [F(v) for v in qc]
        b1_, b2_, b3_ = (-qb[1], -qb[2], -qb[3]) if cb else (qb[1], qb[2], qb[3])
        c1, c2, c3 = (-qc[1], -qc[2], -qc[3]) if cc else (qc[1], qc[2], qc[3])
        return [qb[0] * qc[0] - b1_ * c1 - b2_ * c2 - b3_ * c3, qb[0] * c1 + b1_ * qc[0] + b2_ * c3 - b3_ * c2,
                qb[0] * c2 + b2_ * qc[0] + b3_ * c1 - b1_ * c3, qb[0] * c3 + b3_ * qc[0] + b1_ * c2 - b2_ * c1]

    one = (1.0, 0.0, 0.0, 0.0)
    ax = np.array([0.0, 3.0, 4.0], F) / F(5.0)
    p = np.array([1.0, -0.5, 2.0], F)
    X0, X1 = np.array(x0, F), np.array(x1, F)
    lim0 = dict(lostop=-INF, histop=INF, bounce=0.0, stop_erp=0.2, stop_cfm=1e-5, vel=0.0, fmax=0.0,
                fudge=1.0, normal_cfm=1e-5)
    joints = [
        {"type": "fixed", "bodies": (0, 1), "qrel": qmul(one, one, True, False), "offset": X0 - X1},
        {"type": "slider", "bodies": (1, None), "axis1": (1.0, 0.0, 0.0), "qrel": (1.0, -0.0, -0.0, -0.0),
         "offset": X1, "limot": lim0},
        {"type": "hinge", "bodies": (0, 1), "anchor1": p - X0, "anchor2": p - X1, "axis1": ax, "axis2": ax,
         "qrel": qmul(one, one, True, False),
         "limot": dict(lim0, lostop=-0.5, histop=0.5, fmax=2.0, vel=1.0, bounce=0.25)},
    ]
    adis0 = {"lin_thr2": F(0.01) * F(0.01), "ang_thr2": F(0.01) * F(0.01), "samples": 1, "idle_time": 0.0,
             "idle_steps": 10}
    I = np.zeros(12)
    I[0] = I[5] = I[10] = 0.4 * 2.0 * 0.25
    bodies = [{"pos": x, "q": one, "lvel": (0, 0, 0), "avel": (0, 0, 0), "mass": 2.0, "I": I, "c": (0, 0, 0),
               "enabled": True, "gravity": True, "adis": None, "facc": (0, 0, 0), "tacc": (0, 0, 0), "geoms": []}
              for x in (x1, x0)]
    # body list newest first: b1 is body[0], b0 is body[1]
    for j in joints:
        j["bodies"] = tuple(None if b is None else 1 - b for b in j["bodies"])
    want = dif.export({"gravity": (0.0, 0.0, 0.0), "erp": 0.2, "cfm": 1e-5, "adis": adis0}, bodies, joints)
    if got != want:
        import difflib
        say("".join(list(difflib.unified_diff(want.splitlines(True), got.splitlines(True)))[:60]))
    assert got == want
    W.close()
