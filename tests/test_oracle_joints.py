"""Hinge, slider and fixed joints of the oracle pinned by closed forms of one
step (joints/hinge.cpp, slider.cpp, fixed.cpp, joint.cpp setBall /
setFixedOrientation / addLimot).  Rigid rows (cfm 0), no position correction
(erp 0) unless a limit is tested, no gravity, and enough sweeps that the
Gauss-Seidel iterate is the exact constrained velocity."""
import math

import numpy as np
import pytest

import oracle as O

INF = float("inf")
H = 0.01


def sphere(pos=(0.0, 0.0, 0.0), v=(0.0, 0.0, 0.0), w=(0.0, 0.0, 0.0), q=(1.0, 0.0, 0.0, 0.0), m=2.0, r=0.5):
    b = np.zeros(32)
    b[0:3] = pos
    b[4:8] = q
    b[8:11] = v
    b[12:15] = w
    ii = 0.4 * m * r * r
    b[24], b[25:28] = m, (ii, ii, ii)
    return b


def joint(jtype, b1, b2, anchor1=(0, 0, 0), anchor2=(0, 0, 0), axis1=(0, 0, 1), axis2=(0, 0, 1), qrel=(1, 0, 0, 0),
          offset=(0, 0, 0), **limot):
    rec = np.zeros(16)
    rec[0], rec[1], rec[2] = jtype, b1, b2
    rec[3:6], rec[6:9] = anchor1, anchor2
    rec[9], rec[10] = 0.0, 0.0   # erp, cfm of the ball rows
    x = np.zeros(32)
    x[0:3], x[3:6], x[6:10], x[10:13] = axis1, axis2, qrel, offset
    # dxJointLimitMotor::init (joint.cpp:495-508) with cfm 0
    lm = dict(lostop=-INF, histop=INF, vel=0.0, fmax=0.0, fudge=1.0, normal_cfm=0.0, stop_erp=0.2, stop_cfm=0.0,
              bounce=0.0, jerp=0.0, jcfm=0.0)
    lm.update(limot)
    x[16], x[17], x[18], x[19], x[20] = lm["lostop"], lm["histop"], lm["vel"], lm["fmax"], lm["fudge"]
    x[21], x[22], x[23], x[24] = lm["normal_cfm"], lm["stop_erp"], lm["stop_cfm"], lm["bounce"]
    x[26], x[27] = lm["jerp"], lm["jcfm"]
    return rec, x


def step(bodies, joints, iters=200):
    prm = O.params(h=H, gravity=(0.0, 0.0, 0.0), erp=0.0, cfm=0.0, iters=iters, w=1.0)
    recs = np.array([j[0] for j in joints])
    ext = np.array([j[1] for j in joints])
    out, _ = O.quickstep(np.array(bodies), recs, prm, dtype=np.float64, jext=ext)
    return out


def test_hinge_to_the_world_keeps_only_the_spin_about_its_axis():
    out = step([sphere(v=(0.5, -1.0, 0.25), w=(1.0, 2.0, 3.0))], [joint(2, 0, -1)])
    np.testing.assert_allclose(out[0, 8:11], 0.0, atol=1e-12)
    np.testing.assert_allclose(out[0, 12:15], [0.0, 0.0, 3.0], atol=1e-12)


def test_hinge_motor_reaches_its_velocity_in_one_step():
    out = step([sphere(w=(0.0, 0.0, -1.0))], [joint(2, 0, -1, vel=2.0, fmax=1e9)])
    assert out[0, 14] == pytest.approx(2.0, abs=1e-12)
    # a weak motor moves the spin by at most fmax * h / I
    I = 0.4 * 2.0 * 0.25
    out = step([sphere(w=(0.0, 0.0, -1.0))], [joint(2, 0, -1, vel=2.0, fmax=1.0)])
    assert out[0, 14] == pytest.approx(-1.0 + 1.0 * H / I, abs=1e-12)


@pytest.mark.parametrize("side", [1, -1])
def test_hinge_limit_pushes_back_by_stop_erp(side):
    # the body sits 0.5 rad about +z from its zero angle, the stop is at 0.3:
    # the limit row asks for -stop_erp / h * (0.5 - 0.3) rad/s (getHingeAngle's
    # sign: a body rotated +0.5 about the axis reads +0.5)
    a = 0.5 * side
    q = (math.cos(a / 2), 0.0, 0.0, math.sin(a / 2))
    lim = dict(histop=0.3) if side > 0 else dict(lostop=-0.3)
    out = step([sphere(q=q)], [joint(2, 0, -1, **lim)])
    assert out[0, 14] == pytest.approx(-side * 0.2 / H * 0.2, abs=1e-9)
    # inside the stops nothing happens
    out = step([sphere(q=q)], [joint(2, 0, -1, lostop=-1.0, histop=1.0)])
    assert abs(out[0, 14]) < 1e-12


def test_slider_to_the_world_keeps_only_the_motion_along_its_axis():
    out = step([sphere(v=(1.0, 2.0, 3.0), w=(0.5, -0.5, 0.5))], [joint(3, 0, -1, axis1=(1, 0, 0))])
    np.testing.assert_allclose(out[0, 8:11], [1.0, 0.0, 0.0], atol=1e-12)
    np.testing.assert_allclose(out[0, 12:15], 0.0, atol=1e-12)
    out = step([sphere(v=(1.0, 0.0, 0.0))], [joint(3, 0, -1, axis1=(1, 0, 0), vel=3.0, fmax=1e9)])
    assert out[0, 8] == pytest.approx(3.0, abs=1e-12)


def test_slider_limit_on_its_position():
    # the body is 0.5 along the axis from the offset point; the high stop at 0.3
    out = step([sphere(pos=(0.5, 0.0, 0.0))], [joint(3, 0, -1, axis1=(1, 0, 0), histop=0.3)])
    assert out[0, 8] == pytest.approx(-0.2 / H * 0.2, abs=1e-9)


def test_fixed_to_the_world_stops_the_body():
    out = step([sphere(pos=(1.0, 2.0, 3.0), v=(1.0, 2.0, 3.0), w=(3.0, -2.0, 1.0))],
               [joint(4, 0, -1, offset=(1.0, 2.0, 3.0))])
    np.testing.assert_allclose(out[0, 8:15], 0.0, atol=1e-12)


def test_fixed_pair_moves_as_one_body_with_its_momentum():
    # two equal spheres 1 m apart along x, one pushed along the line: both
    # move at half the speed, with no spin (offset: body 1's centre from
    # body 2's, in body 1's frame)
    out = step([sphere(pos=(0.0, 0.0, 0.0), v=(2.0, 0.0, 0.0)), sphere(pos=(1.0, 0.0, 0.0))],
               [joint(4, 0, 1, offset=(-1.0, 0.0, 0.0))], iters=400)
    np.testing.assert_allclose(out[:, 8:11], [[1.0, 0.0, 0.0], [1.0, 0.0, 0.0]], atol=1e-9)
    np.testing.assert_allclose(out[:, 12:15], 0.0, atol=1e-9)
    # pushed across the line, the pair also turns: momentum 2 kg x 2 m/s at
    # r = -0.5 m from the centre of mass gives L_z = -2; each body then moves
    # with v_cm + w x r
    out = step([sphere(pos=(0.0, 0.0, 0.0), v=(0.0, 2.0, 0.0)), sphere(pos=(1.0, 0.0, 0.0))],
               [joint(4, 0, 1, offset=(-1.0, 0.0, 0.0))], iters=2000)
    m, I = 2.0, 0.4 * 2.0 * 0.25
    wz = -(m * 2.0 * 0.5) / (2 * I + 2 * m * 0.25)
    np.testing.assert_allclose(out[:, 14], [wz, wz], atol=1e-8)
    np.testing.assert_allclose(out[:, 9], [1.0 - wz * 0.5, 1.0 + wz * 0.5], atol=1e-8)
