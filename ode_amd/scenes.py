"""Seeded synthetic scenes for the step path (the benchmark workloads and test
cases).  Each builder returns numpy record arrays in the BODY/GEOM layouts of
include/odx.h:

  BODY [32] float: 0-2 pos, 4-7 q (w,x,y,z), 8-10 lvel, 12-14 avel,
                   16-18 facc, 20-22 tacc, 24 mass, 25-30 I (xx,yy,zz,xy,xz,yz),
                   31 flags (bit0 no gravity, bit1 gyroscopic)
  GEOM [16] float: 0 class (0 sphere, 1 box, 2 capsule, 3 cylinder, 4 plane),
                   1 body (-1 static), 2-5 params, 6-8 static pos, 9-12 static q

Mass properties follow dMassSetBox / dMassSetSphere / dMassSetCapsule
(ode/src/mass.cpp:118-215) with the body-frame centre of mass at the origin.
"""
import math

import numpy as np

BODY = 32
GEOM = 16
SPHERE, BOX, CAPSULE, CYLINDER, PLANE = 0, 1, 2, 3, 4
FLAG_NO_GRAVITY = 1
FLAG_GYRO = 2


def _body(pos, q, mass, inertia_diag, lvel=(0, 0, 0), avel=(0, 0, 0)):
    b = np.zeros(BODY, np.float64)
    b[0:3] = pos
    q = np.asarray(q, np.float64)
    b[4:8] = q / np.linalg.norm(q)
    b[8:11] = lvel
    b[12:15] = avel
    b[24] = mass
    b[25:28] = inertia_diag
    b[31] = FLAG_GYRO
    return b


def box_mass(density, lx, ly, lz):
    m = lx * ly * lz * density
    return m, (m / 12.0 * (ly * ly + lz * lz), m / 12.0 * (lx * lx + lz * lz), m / 12.0 * (lx * lx + ly * ly))


def sphere_mass(density, r):
    m = (4.0 / 3.0) * math.pi * r * r * r * density
    ii = 0.4 * m * r * r
    return m, (ii, ii, ii)


def capsule_mass(density, r, length, direction=3):
    m1 = math.pi * r * r * length * density
    m2 = (4.0 / 3.0) * math.pi * r * r * r * density
    ia = m1 * (0.25 * r * r + (1.0 / 12.0) * length * length) + m2 * (0.4 * r * r + 0.375 * r * length + 0.25 * length * length)
    ib = (m1 * 0.5 + m2 * 0.4) * r * r
    d = [ia, ia, ia]
    d[direction - 1] = ib
    return m1 + m2, tuple(d)


def cylinder_mass(density, r, length, direction=3):
    """dMassSetCylinder (ode/src/mass.cpp:173-198)."""
    m = math.pi * r * r * length * density
    ia = m * (0.25 * r * r + (1.0 / 12.0) * length * length)
    d = [ia, ia, ia]
    d[direction - 1] = m * 0.5 * r * r
    return m, tuple(d)


def geom(cls, body, params, pos=(0, 0, 0), q=(1, 0, 0, 0)):
    g = np.zeros(GEOM, np.float64)
    g[0] = cls
    g[1] = body
    g[2:2 + len(params)] = params
    g[6:9] = pos
    g[9:13] = q
    return g


def plane(a, b, c, d):
    n = math.sqrt(a * a + b * b + c * c)
    return geom(PLANE, -1, (a / n, b / n, c / n, d / n))


class World:
    """One world: body records, geom records, joint records (ball, or hinge /
    slider / fixed with their extended records in jext, 32 floats each)."""

    def __init__(self, bodies, geoms, balls=None, jext=None):
        self.bodies = np.asarray(bodies, np.float64).reshape(-1, BODY)
        self.geoms = np.asarray(geoms, np.float64).reshape(-1, GEOM)
        self.balls = np.zeros((0, 16)) if balls is None else np.asarray(balls, np.float64).reshape(-1, 16)
        self.jext = None if jext is None else np.asarray(jext, np.float64).reshape(-1, 32)


def _randq(rng):
    q = rng.normal(size=4)
    return q / np.linalg.norm(q)


def boxes_on_plane(n=1000, side=0.5, density=1000.0, extent=20.0, seed=1):
    """1,000 boxes (side 0.5 m, density 1000) with centres uniform in a
    20x20x20 m volume above the z=0 plane, random unit-quaternion orientations,
    at rest."""
    rng = np.random.default_rng(seed)
    m, I = box_mass(density, side, side, side)
    bodies, geoms = [], []
    for i in range(n):
        pos = (rng.uniform(-extent / 2, extent / 2), rng.uniform(-extent / 2, extent / 2), rng.uniform(0.5, 0.5 + extent))
        bodies.append(_body(pos, _randq(rng), m, I))
        geoms.append(geom(BOX, i, (side, side, side)))
    geoms.append(plane(0, 0, 1, 0))
    return World(bodies, geoms)


def uneven_world(rng, nmin=8, nmax=64, density=1000.0):
    """Bodies uniform in [nmin, nmax]; half spheres (radius log-normal, median
    0.25 m, sigma 0.3), half boxes (sides uniform 0.2-0.8 m), dropped from 0-5 m
    onto a ground plane, x/y uniform in [-2, 2] m."""
    nb = int(rng.integers(nmin, nmax + 1))
    bodies, geoms = [], []
    for i in range(nb):
        xy = rng.uniform(-2, 2, size=2)
        drop = rng.uniform(0, 5)
        if i % 2 == 0:
            r = float(np.exp(rng.normal(math.log(0.25), 0.3)))
            m, I = sphere_mass(density, r)
            bodies.append(_body((xy[0], xy[1], drop + r), (1, 0, 0, 0), m, I))
            geoms.append(geom(SPHERE, i, (r,)))
        else:
            s = rng.uniform(0.2, 0.8, size=3)
            m, I = box_mass(density, *s)
            bodies.append(_body((xy[0], xy[1], drop + 0.5 * float(np.max(s))), _randq(rng), m, I))
            geoms.append(geom(BOX, i, tuple(s)))
    geoms.append(plane(0, 0, 1, 0))
    return World(bodies, geoms)


def uneven_batch(nworlds=4096, seed=2, nmin=8, nmax=64):
    rng = np.random.default_rng(seed)
    return [uneven_world(rng, nmin, nmax) for _ in range(nworlds)]


def tower_world(rng, levels=32, side=0.4, jitter=0.05, density=1000.0, sphere_r=0.2):
    """A 32-box tower (side 0.4 m, +-5 cm seeded x/y jitter per level) on a
    plane with one projectile sphere (r = 0.2 m, speed 5-15 m/s) aimed at a
    random height of the tower from a random bearing."""
    m, I = box_mass(density, side, side, side)
    bodies, geoms = [], []
    for k in range(levels):
        j = rng.uniform(-jitter, jitter, size=2)
        bodies.append(_body((j[0], j[1], side * 0.5 + side * k), (1, 0, 0, 0), m, I))
        geoms.append(geom(BOX, k, (side, side, side)))
    th = rng.uniform(0, 2 * math.pi)
    z0 = rng.uniform(0.5, levels * side)
    start = np.array([3.0 * math.cos(th), 3.0 * math.sin(th), z0])
    target = np.array([0.0, 0.0, rng.uniform(0.2, levels * side)])
    d = target - start + rng.normal(scale=0.3, size=3)
    d /= np.linalg.norm(d)
    speed = rng.uniform(5, 15)
    ms, Is = sphere_mass(density, sphere_r)
    bodies.append(_body(tuple(start), (1, 0, 0, 0), ms, Is, lvel=tuple(d * speed)))
    geoms.append(geom(SPHERE, levels, (sphere_r,)))
    geoms.append(plane(0, 0, 1, 0))
    return World(bodies, geoms)


def tower_batch(nworlds=4096, seed=5, levels=32):
    rng = np.random.default_rng(seed)
    return [tower_world(rng, levels) for _ in range(nworlds)]


def chain_world(rng, links=16, length=0.3, radius=0.05, nboxes=8, density=1000.0):
    """A 16-link capsule chain (link length 0.3 m, radius 0.05 m) joined by ball
    joints and anchored to the world at one end, with initial link angular
    velocities N(0, 1 rad/s), plus 8 free boxes, on a plane.  Capsule axes lie
    along body z (dCreateCapsule convention), chain laid out along +x at 1 m."""
    bodies, geoms, balls = [], [], []
    m, I = capsule_mass(density, radius, length, 3)
    # rotate body z onto world x: quaternion for +90 deg about y
    qy = (math.cos(math.pi / 4), 0.0, math.sin(math.pi / 4), 0.0)
    seg = length + 2 * radius
    for k in range(links):
        pos = (seg * (k + 0.5), 0.0, 1.0)
        bodies.append(_body(pos, qy, m, I, avel=tuple(rng.normal(size=3))))
        geoms.append(geom(CAPSULE, k, (radius, length)))
    # joints: world anchor at x=0, then between consecutive links
    for k in range(links):
        anchor = np.array([seg * k, 0.0, 1.0])
        rec = np.zeros(16)
        rec[0] = 1
        rec[1] = k
        rec[2] = k - 1 if k > 0 else -1
        rec[9] = 0.2
        rec[10] = 1e-5
        rec[3:6] = _to_body_frame(anchor, bodies[k])
        rec[6:9] = _to_body_frame(anchor, bodies[k - 1]) if k > 0 else anchor
        balls.append(rec)
    for i in range(nboxes):
        s = rng.uniform(0.2, 0.5, size=3)
        mb, Ib = box_mass(density, *s)
        pos = (rng.uniform(0.2, seg * links), rng.uniform(-0.6, 0.6), rng.uniform(0.3, 2.0))
        bodies.append(_body(pos, _randq(rng), mb, Ib))
        geoms.append(geom(BOX, links + i, tuple(s)))
    geoms.append(plane(0, 0, 1, 0))
    return World(bodies, geoms, balls)


def _rot(q):
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
    ])


def _to_body_frame(p, body):
    R = _rot(body[4:8])
    return R.T @ (np.asarray(p) - body[0:3])


def chain_batch(nworlds=1024, seed=3):
    rng = np.random.default_rng(seed)
    return [chain_world(rng) for _ in range(nworlds)]


def sphere_pour(n=200000, rmin=0.05, rmax=0.15, half=5.0, density=1000.0, seed=4, layer_gap=0.0):
    """Spheres (radius uniform 0.05-0.15 m) placed on a jittered lattice inside
    a 10x10 m open box of 5 static planes (ground plus four walls)."""
    rng = np.random.default_rng(seed)
    cell = 2 * rmax + layer_gap
    per = int((2 * half - 2 * rmax) // cell)
    bodies, geoms = [], []
    for i in range(n):
        lx, ly, lz = i % per, (i // per) % per, i // (per * per)
        r = rng.uniform(rmin, rmax)
        pos = (-half + rmax + cell * lx + rng.uniform(-0.01, 0.01), -half + rmax + cell * ly + rng.uniform(-0.01, 0.01),
               rmax + cell * lz + 0.01)
        m, I = sphere_mass(density, r)
        bodies.append(_body(pos, (1, 0, 0, 0), m, I))
        geoms.append(geom(SPHERE, i, (r,)))
    geoms.append(plane(0, 0, 1, 0))
    geoms.append(plane(1, 0, 0, -half))
    geoms.append(plane(-1, 0, 0, -half))
    geoms.append(plane(0, 1, 0, -half))
    geoms.append(plane(0, -1, 0, -half))
    return World(bodies, geoms)


def mixed_pile(n=1500, spacing=0.2, seed=6, density=1000.0, walls=True):
    """A packed pile of spheres (r 0.05-0.15 m), boxes (sides 0.1-0.2 m) and
    capsules (r 0.05 m, length 0.15 m) on a jittered cubic lattice of the given
    spacing, so that neighbours touch or overlap and the bottom layer rests on
    the z=0 plane; four walls bound the lattice's columns."""
    rng = np.random.default_rng(seed)
    per = max(2, int(round(n ** (1.0 / 3.0))))
    half = 0.5 * per * spacing
    bodies, geoms = [], []
    for i in range(n):
        lx, ly, lz = i % per, (i // per) % per, i // (per * per)
        pos = (-half + spacing * (lx + 0.5) + rng.uniform(-0.01, 0.01),
               -half + spacing * (ly + 0.5) + rng.uniform(-0.01, 0.01),
               0.09 + spacing * lz + rng.uniform(-0.01, 0.01))
        kind = rng.integers(0, 3)
        if kind == 0:
            r = rng.uniform(0.05, 0.15)
            m, I = sphere_mass(density, r)
            bodies.append(_body(pos, (1, 0, 0, 0), m, I))
            geoms.append(geom(SPHERE, i, (r,)))
        elif kind == 1:
            s = rng.uniform(0.1, 0.2, size=3)
            m, I = box_mass(density, *s)
            bodies.append(_body(pos, _randq(rng), m, I))
            geoms.append(geom(BOX, i, tuple(s)))
        else:
            m, I = capsule_mass(density, 0.05, 0.15)
            bodies.append(_body(pos, _randq(rng), m, I))
            geoms.append(geom(CAPSULE, i, (0.05, 0.15)))
    geoms.append(plane(0, 0, 1, 0))
    if walls:
        geoms.append(plane(1, 0, 0, -half))
        geoms.append(plane(-1, 0, 0, -half))
        geoms.append(plane(0, 1, 0, -half))
        geoms.append(plane(0, -1, 0, -half))
    return World(bodies, geoms)
