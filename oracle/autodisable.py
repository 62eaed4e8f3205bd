"""Auto-disable of dWorldQuickStep (TEST INFRASTRUCTURE ONLY).

Restates dInternalHandleAutoDisabling (ode/src/util.cpp:427-565) with the
body state and parameters it reads (ode/src/objects.h:87-96, 237-243; world
defaults ode/src/objects.cpp:54-61; setters ode/src/ode.cpp:891-1021), in
float32 arithmetic as a dSINGLE build computes it.  The island pass that
follows (util.cpp:724-855) re-enables a disabled body reached through a joint
from an enabled one; `enabled_after_islands` restates that with
oracle.islands.island_order.
"""
import numpy as np

from .islands import island_order

F = np.float32


class Body:
    """Auto-disable state of one body, as dBodyCreate leaves it when the
    world's auto-disable flag is set."""

    def __init__(self, idle_steps=10, idle_time=0.0, average_samples=1, linear_threshold=0.01,
                 angular_threshold=0.01, auto_disable=True):
        self.enabled = True
        self.auto_disable = auto_disable
        self.idle_steps = int(idle_steps)
        self.idle_time = F(idle_time)
        self.samples = int(average_samples)
        # thresholds are kept squared (ode.cpp:898-916)
        self.lin_thr2 = F(linear_threshold) * F(linear_threshold)
        self.ang_thr2 = F(angular_threshold) * F(angular_threshold)
        self.stepsleft = self.idle_steps
        self.timeleft = self.idle_time
        self.buf_l = np.zeros((max(self.samples, 1), 3), F)
        self.buf_a = np.zeros((max(self.samples, 1), 3), F)
        self.counter = 0
        self.ready = False

    def set_idle_steps(self, n):
        """dBodySetAutoDisableSteps (ode.cpp:964-968): the budget only; the
        running countdown keeps its value until the next reset."""
        self.idle_steps = int(n)

    def enable(self):
        """dBodyEnable (ode.cpp:850-857): enabled, countdowns reset."""
        self.enabled = True
        self.stepsleft = self.idle_steps
        self.timeleft = self.idle_time


def _dot(v):
    v = np.asarray(v, F)
    return F(F(v[0] * v[0]) + F(v[1] * v[1])) + F(v[2] * v[2])


def handle(bodies, lvel, avel, has_joints, h):
    """One dInternalHandleAutoDisabling pass.  bodies: list of Body; lvel,
    avel: [n, 3] float32 velocities, zeroed in place for a body disabled here;
    has_joints[i]: body i has at least one joint attached.  Returns the
    indices disabled by this pass."""
    h = F(h)
    out = []
    for i, b in enumerate(bodies):
        if not has_joints[i]:
            continue                      # never freeze a body in mid-air
        if not (b.auto_disable and b.enabled):
            continue
        n = b.samples
        if n == 0:
            continue
        if b.counter >= n:
            b.ready, b.counter = False, 0
        b.buf_l[b.counter] = np.asarray(lvel[i], F)
        b.buf_a[b.counter] = np.asarray(avel[i], F)
        b.counter += 1
        if b.counter >= n:
            b.counter, b.ready = 0, True
        idle = False
        if b.ready:
            idle = True
            al = b.buf_l[0].copy()
            aa = b.buf_a[0].copy()
            if n > 1:
                for k in range(1, n):
                    al = (al + b.buf_l[k]).astype(F)
                    aa = (aa + b.buf_a[k]).astype(F)
                r1 = F(F(1.0) / F(n))
                al = (al * r1).astype(F)
                aa = (aa * r1).astype(F)
            if _dot(al) > b.lin_thr2:
                idle = False
            elif _dot(aa) > b.ang_thr2:
                idle = False
        if idle:
            b.stepsleft -= 1
            b.timeleft = F(b.timeleft - h)
        else:
            b.stepsleft = b.idle_steps
            b.timeleft = b.idle_time
        if b.stepsleft <= 0 and b.timeleft <= 0:
            b.enabled = False
            lvel[i] = 0
            avel[i] = 0
            out.append(i)
    return out


def enabled_after_islands(bodies, joints):
    """Bodies the island pass steps (and so leaves enabled): every enabled
    body, and every body reached from one through joints.  joints: (b1, b2)
    pairs in attach order, b2 = -1 for the world."""
    n = len(bodies)
    norm = [(a, c) if a >= 0 else (c, -1) for a, c in joints]
    stepped, _ = island_order(n, norm, [b.enabled for b in bodies])
    for i in stepped:
        bodies[i].enabled = True
    return sorted(stepped)
