"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP implementation.  Only tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg import this module.
Record layouts are documented in oracle/oracle.h.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

BODY = 32
GEOM = 16
CONTACT = 12
JOINT = 16

# params vector indices (oracle.h ORC_P_*)
P_H, P_GX, P_GY, P_GZ, P_ERP, P_CFM, P_SOR_W, P_ITERS, P_MAX_VEL, P_MIN_DEPTH = range(10)
P_MODE, P_MU, P_MU2, P_BOUNCE, P_BOUNCE_VEL, P_SOFT_ERP, P_SOFT_CFM, P_MAXC, P_SPACE = range(10, 19)
P_MOTION1, P_MOTION2, P_MOTIONN, P_SLIP1, P_SLIP2, P_SKIP_CONNECTED, P_LAMBDA_FLOOR, P_TOL, P_CHECK_EVERY = range(19, 28)
P_DYN_PREMATURE, P_DYN_EXTRA_FACTOR, P_DYN_EXTRA_DELTA, P_RANDOM_REORDER = range(28, 32)


def build():
    subprocess.check_call(["sh", os.path.join(_HERE, "build.sh")])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = ctypes.CDLL(_LIB)
        _lib.orc_lcp_residual.restype = ctypes.c_double
        _lib.orc_drand.restype = ctypes.c_uint
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _real(dtype):
    return "f32" if np.dtype(dtype) == np.float32 else "f64"


def params(**kw):
    """Default step parameters (dWorldCreate defaults of objects.cpp:37-99 for
    dSINGLE, deterministic QuickStep, contact surface mu=0.5 with Approx1)."""
    p = np.zeros(32, np.float64)
    p[P_H] = 0.01
    p[P_GZ] = -9.81
    p[P_ERP] = 0.2
    p[P_CFM] = 1e-5
    p[P_SOR_W] = 1.3
    p[P_ITERS] = 20
    p[P_MAX_VEL] = np.inf
    p[P_MIN_DEPTH] = 0.0
    p[P_MODE] = 0x7000
    p[P_MU] = 0.5
    p[P_MAXC] = 4
    p[P_SPACE] = 0
    p[P_LAMBDA_FLOOR] = 1e-3
    names = {
        "h": P_H, "erp": P_ERP, "cfm": P_CFM, "w": P_SOR_W, "iters": P_ITERS, "max_vel": P_MAX_VEL,
        "min_depth": P_MIN_DEPTH, "mode": P_MODE, "mu": P_MU, "mu2": P_MU2, "bounce": P_BOUNCE,
        "bounce_vel": P_BOUNCE_VEL, "soft_erp": P_SOFT_ERP, "soft_cfm": P_SOFT_CFM, "maxc": P_MAXC,
        "space": P_SPACE, "skip_connected": P_SKIP_CONNECTED, "lambda_floor": P_LAMBDA_FLOOR,
        "tol": P_TOL, "check_every": P_CHECK_EVERY,
        "dyn_premature": P_DYN_PREMATURE, "dyn_extra_factor": P_DYN_EXTRA_FACTOR, "dyn_extra_delta": P_DYN_EXTRA_DELTA,
        "random_reorder": P_RANDOM_REORDER,
        "row_order": P_RANDOM_REORDER,   # 0 ReorderPrep, 1 random reordering, 2 joint order
    }
    for k, v in kw.items():
        if k == "gravity":
            p[P_GX:P_GZ + 1] = v
        else:
            p[names[k]] = v
    return p


def last_sweeps():
    return lib().orc_last_sweeps()


def rand_seed(seed=None):
    """The oracle's dRand state used by the random reordering: sets it when
    seed is given; returns the current value."""
    if seed is not None:
        lib().orc_rand_set_seed(ctypes.c_uint(seed))
    lib().orc_rand_get_seed.restype = ctypes.c_uint
    return int(lib().orc_rand_get_seed())


def drand_seq(seed, n):
    s = ctypes.c_uint(seed)
    return [lib().orc_drand(ctypes.byref(s)) for _ in range(n)]


def drand_int_seq(seed, n, k):
    s = ctypes.c_uint(seed)
    return [lib().orc_drand_int(ctypes.byref(s), k) for _ in range(n)]


def collide(g1, g2, b1=None, b2=None, maxc=4, dtype=np.float64):
    g1 = np.ascontiguousarray(g1, dtype)
    g2 = np.ascontiguousarray(g2, dtype)
    b1 = np.ascontiguousarray(b1 if b1 is not None else np.zeros(BODY), dtype)
    b2 = np.ascontiguousarray(b2 if b2 is not None else np.zeros(BODY), dtype)
    out = np.zeros((16, CONTACT), dtype)
    fn = getattr(lib(), "orc_collide_" + _real(dtype))
    n = fn(_p(g1), _p(g2), _p(b1), _p(b2), ctypes.c_int(maxc), _p(out))
    return out[:n]


def aabbs(geoms, bodies, dtype=np.float32):
    geoms = np.ascontiguousarray(geoms, dtype)
    bodies = np.ascontiguousarray(bodies, dtype)
    out = np.zeros((len(geoms), 6), dtype)
    getattr(lib(), "orc_aabbs_" + _real(dtype))(ctypes.c_int(len(geoms)), _p(geoms), _p(bodies), _p(out))
    return out


def pairs(aabb, body, space="sap", cat=None, col=None, levels=(-3, 10)):
    aabb = np.ascontiguousarray(aabb)
    dtype = aabb.dtype
    body = np.ascontiguousarray(body, np.int32)
    ng = len(body)
    cap = max(16, ng * ng)
    out = np.zeros(2 * cap, np.int32)
    catp = _p(np.ascontiguousarray(cat, np.uint32)) if cat is not None else None
    colp = _p(np.ascontiguousarray(col, np.uint32)) if col is not None else None
    if space == "sap":
        n = getattr(lib(), "orc_pairs_sap_" + _real(dtype))(ctypes.c_int(ng), _p(aabb), _p(body), catp, colp, _p(out), ctypes.c_int(cap))
    else:
        n = getattr(lib(), "orc_pairs_hash_" + _real(dtype))(ctypes.c_int(ng), _p(aabb), _p(body), catp, colp,
                                                           ctypes.c_int(levels[0]), ctypes.c_int(levels[1]), _p(out), ctypes.c_int(cap))
    return out[: 2 * n].reshape(n, 2)


def contact_rows(contact, surface, body1, body2=None, reversed_=False, fps=100.0, erp=0.0, min_depth=0.0, max_vel=np.inf):
    contact = np.ascontiguousarray(contact, np.float64)
    surface = np.ascontiguousarray(surface, np.float64)
    body1 = np.ascontiguousarray(body1, np.float64)
    b2 = np.ascontiguousarray(body2, np.float64) if body2 is not None else None
    J = np.zeros((6, 16), np.float64)
    fi = np.zeros(6, np.int32)
    m = lib().orc_contact_rows_f64(_p(contact), _p(surface), _p(body1), _p(b2), ctypes.c_int(int(reversed_)),
                                   ctypes.c_double(fps), ctypes.c_double(erp), ctypes.c_double(min_depth),
                                   ctypes.c_double(max_vel), _p(J), _p(fi))
    return J[:m], fi[:m]


def quickstep(bodies, joints, prm, dtype=np.float64, jext=None, table=None):
    """One QuickStep of the given joints (JOINT records in stepping order).
    jext: [nj, 32] extended records of hinge / slider / fixed joints (types
    2 / 3 / 4; layout in oracle.cpp's ORC_JEXT_STRIDE).  table: a
    surface_table; each contact record's slot 13 then names its entry."""
    bodies = np.ascontiguousarray(bodies, dtype).copy()
    joints = np.ascontiguousarray(joints, dtype).reshape(-1, JOINT)
    lam = np.zeros(max(1, 6 * len(joints)), dtype)
    if table is not None:
        t = np.ascontiguousarray(table, np.float64)
        m = getattr(lib(), "orc_quickstep_table_" + _real(dtype))(
            ctypes.c_int(len(bodies)), _p(bodies), ctypes.c_int(len(joints)), _p(joints),
            _p(np.ascontiguousarray(prm, np.float64)), ctypes.c_int(t.shape[0]), _p(t), _p(lam), ctypes.c_int(len(lam)))
        if m < 0:
            raise ValueError("a contact record names a surface outside the table")
        return bodies, lam[:m]
    if jext is not None:
        x = np.ascontiguousarray(jext, dtype).reshape(-1, 32)
        m = getattr(lib(), "orc_quickstep_ext_" + _real(dtype))(
            ctypes.c_int(len(bodies)), _p(bodies), ctypes.c_int(len(joints)), _p(joints), _p(x),
            _p(np.ascontiguousarray(prm, np.float64)), _p(lam), ctypes.c_int(len(lam)))
        return bodies, lam[:m]
    m = getattr(lib(), "orc_quickstep_" + _real(dtype))(ctypes.c_int(len(bodies)), _p(bodies), ctypes.c_int(len(joints)),
                                                       _p(joints), _p(np.ascontiguousarray(prm, np.float64)), _p(lam),
                                                       ctypes.c_int(len(lam)))
    return bodies, lam[:m]


def surface_table(nmat, entries, default=None):
    """A symmetric material-pair surface table (double[nmat * nmat][16]):
    entries maps (m1, m2) to a dict of surface fields (mode, mu, mu2, bounce,
    bounce_vel, soft_erp, soft_cfm, motion1, motion2, motionN, slip1, slip2,
    max_contacts); unspecified pairs take `default` (the parameters' surface
    when None)."""
    fields = ["mode", "mu", "mu2", "bounce", "bounce_vel", "soft_erp", "soft_cfm", "motion1", "motion2",
              "motionN", "slip1", "slip2", "max_contacts"]
    base = {"mode": 0x7000, "mu": 0.5}
    base.update(default or {})
    t = np.zeros((nmat, nmat, 16), np.float64)
    for a in range(nmat):
        for b in range(nmat):
            e = dict(base)
            e.update(entries.get((a, b), entries.get((b, a), {})))
            for k, v in e.items():
                t[a, b, fields.index(k)] = v
    return t


def world_step(bodies, geoms, prm, balls=None, dtype=np.float32, ccap=None, table=None, category_filter=False,
               jext=None):
    """One canonical world step; returns (new_bodies, contacts[nc,12], lambda[m]).
    ccap bounds the contacts (default: 4 per geom pair).  table: a
    surface_table; geoms then carry their material in slot 13 (and category /
    collide bits in slots 14 / 15, tested when category_filter).  jext:
    [len(balls), 32] extended records of the hinge, slider and fixed joints
    among `balls` (types 2 / 3 / 4; rows of ball joints are ignored)."""
    bodies = np.ascontiguousarray(bodies, dtype).copy()
    geoms = np.ascontiguousarray(geoms, dtype)
    balls = np.zeros((0, JOINT), dtype) if balls is None else np.ascontiguousarray(balls, dtype)
    ng = len(geoms)
    if ccap is None:
        ccap = 4 * ng * ng + 64
    contacts = np.zeros((ccap, CONTACT), dtype)
    lam = np.zeros(3 * ccap + 6 * len(balls) + 8, dtype)
    m = ctypes.c_int(0)
    if jext is not None:
        if table is not None:
            raise ValueError("world_step: jext and table together are not supported")
        x = np.ascontiguousarray(jext, dtype).reshape(-1, 32)
        nc = getattr(lib(), "orc_world_step_ext_" + _real(dtype))(
            ctypes.c_int(len(bodies)), _p(bodies), ctypes.c_int(ng), _p(geoms), ctypes.c_int(len(balls)), _p(balls),
            _p(x), _p(np.ascontiguousarray(prm, np.float64)), _p(contacts), ctypes.c_int(ccap), _p(lam),
            ctypes.c_int(len(lam)), ctypes.byref(m))
        return bodies, contacts[:nc], lam[: m.value]
    if table is not None:
        t = np.ascontiguousarray(table, np.float64)
        nc = getattr(lib(), "orc_world_step_table_" + _real(dtype))(
            ctypes.c_int(len(bodies)), _p(bodies), ctypes.c_int(ng), _p(geoms), ctypes.c_int(len(balls)), _p(balls),
            _p(np.ascontiguousarray(prm, np.float64)), ctypes.c_int(t.shape[0]), _p(t), ctypes.c_int(int(category_filter)),
            _p(contacts), ctypes.c_int(ccap), _p(lam), ctypes.c_int(len(lam)), ctypes.byref(m))
        return bodies, contacts[:nc], lam[: m.value]
    nc = getattr(lib(), "orc_world_step_" + _real(dtype))(
        ctypes.c_int(len(bodies)), _p(bodies), ctypes.c_int(ng), _p(geoms), ctypes.c_int(len(balls)), _p(balls),
        _p(np.ascontiguousarray(prm, np.float64)), _p(contacts), ctypes.c_int(ccap), _p(lam), ctypes.c_int(len(lam)),
        ctypes.byref(m))
    return bodies, contacts[:nc], lam[: m.value]


def contacts_to_joints(contacts, geoms):
    """Contact records -> JOINT records (type 0, body(g1), body(g2), pos, normal, depth)."""
    nc = len(contacts)
    j = np.zeros((nc, JOINT), np.float64)
    if nc:
        g1 = contacts[:, 7].astype(int)
        g2 = contacts[:, 8].astype(int)
        j[:, 1] = geoms[g1, 1]
        j[:, 2] = geoms[g2, 1]
        j[:, 3:6] = contacts[:, 0:3]
        j[:, 6:9] = contacts[:, 3:6]
        j[:, 9] = contacts[:, 6]
    return j


def lcp_residual(bodies, joints, prm, lam, balls=None):
    bodies = np.ascontiguousarray(bodies, np.float64)
    joints = np.ascontiguousarray(joints, np.float64).reshape(-1, JOINT)
    balls = np.zeros((0, JOINT)) if balls is None else np.ascontiguousarray(balls, np.float64)
    lam = np.ascontiguousarray(lam, np.float64)
    return lib().orc_lcp_residual(ctypes.c_int(len(bodies)), _p(bodies), ctypes.c_int(len(balls)), _p(balls),
                                  ctypes.c_int(len(joints)), _p(joints), _p(np.ascontiguousarray(prm, np.float64)), _p(lam))


def solve_to_residual(bodies, joints, prm, target, maxit, balls=None):
    bodies = np.ascontiguousarray(bodies, np.float64)
    joints = np.ascontiguousarray(joints, np.float64).reshape(-1, JOINT)
    balls = np.zeros((0, JOINT)) if balls is None else np.ascontiguousarray(balls, np.float64)
    lam = np.zeros(3 * len(joints) + 3 * len(balls) + 1)
    res = ctypes.c_double(0)
    it = lib().orc_solve_to_residual(ctypes.c_int(len(bodies)), _p(bodies), ctypes.c_int(len(balls)), _p(balls),
                                     ctypes.c_int(len(joints)), _p(joints), _p(np.ascontiguousarray(prm, np.float64)),
                                     ctypes.c_double(target), ctypes.c_int(maxit), _p(lam), ctypes.byref(res))
    return it, res.value, lam


class SpaceModel:
    """The callback order of one ODE space across calls (oracle/spaces.cpp):
    kind "sap" (axisorder as dSweepAndPruneSpaceCreate, dSAP_AXES_XYZ by
    default), "hash" (levels) or
    "simple".  Geoms are integer ids; add/remove/moved mirror dSpaceAdd,
    dSpaceRemove and dGeomMoved."""

    KINDS = {"sap": 0, "hash": 1, "simple": 2}

    def __init__(self, kind, axisorder=(0) | (1 << 2) | (2 << 4), levels=(-3, 10)):
        self._lib = lib()
        self._lib.orc_space_create.restype = ctypes.c_void_p
        self.h = ctypes.c_void_p(self._lib.orc_space_create(ctypes.c_int(self.KINDS[kind]), ctypes.c_int(axisorder),
                                                            ctypes.c_int(levels[0]), ctypes.c_int(levels[1])))

    def add(self, gid):
        self._lib.orc_space_add(self.h, ctypes.c_int(gid))

    def remove(self, gid):
        self._lib.orc_space_remove(self.h, ctypes.c_int(gid))

    def moved(self, gid):
        self._lib.orc_space_moved(self.h, ctypes.c_int(gid))

    def order(self, cap=1 << 16):
        out = np.zeros(cap, np.int32)
        n = self._lib.orc_space_order(self.h, _p(out), ctypes.c_int(cap))
        return [int(x) for x in out[:n]]

    def collide(self, aabb, body, cat=None, col=None, enabled=None):
        """The callback sequence [(o1, o2), ...] of one dSpaceCollide.  Per-id
        arrays: aabb [n, 6] float32, body [n] (-1 none), cat/col [n] uint64
        (default all bits), enabled [n] (default all)."""
        aabb = np.ascontiguousarray(aabb, np.float32)
        n = len(aabb)
        body = np.ascontiguousarray(body, np.int32)
        ones = np.full(n, np.iinfo(np.uint64).max, np.uint64)
        cat = np.ascontiguousarray(ones if cat is None else cat, np.uint64)
        col = np.ascontiguousarray(ones if col is None else col, np.uint64)
        en = np.ascontiguousarray(np.ones(n, np.uint8) if enabled is None else enabled, np.uint8)
        cap = max(64, n * n)
        out = np.zeros(2 * cap, np.int32)
        k = self._lib.orc_space_collide(self.h, _p(aabb), _p(body), _p(cat), _p(col), _p(en), _p(out), ctypes.c_int(cap))
        return [(int(out[2 * i]), int(out[2 * i + 1])) for i in range(min(k, cap))]

    def __del__(self):
        try:
            self._lib.orc_space_destroy(self.h)
        except Exception:
            pass


class RadixModel:
    """The SAP space's radix sort alone, ranks kept between calls."""

    def __init__(self):
        self._lib = lib()
        self._lib.orc_radix_create.restype = ctypes.c_void_p
        self.h = ctypes.c_void_p(self._lib.orc_radix_create())

    def sort(self, keys):
        keys = np.ascontiguousarray(keys, np.float32)
        out = np.zeros(max(1, len(keys)), np.uint32)
        self._lib.orc_radix_sort(self.h, _p(keys), ctypes.c_int(len(keys)), _p(out))
        return [int(x) for x in out[: len(keys)]]

    def valid(self):
        return bool(self._lib.orc_radix_valid(self.h))

    def __del__(self):
        try:
            self._lib.orc_radix_destroy(self.h)
        except Exception:
            pass
