"""Dynamics Interchange Format text of a world (TEST INFRASTRUCTURE ONLY).

Restates dWorldExportDIF (ode/src/export-dif.cpp:45-140 printing helpers,
:185-262 joints, :379-489 geoms, :491-620 the world) for what a
sphere/box/capsule/plane world with ball and contact joints holds, in a
dSINGLE build: reals with 7 significant digits ("%.7g"), "inf"/"-inf",
thresholds printed as stored (squared).  The world is a plain description:

  world  = {"gravity": (3,), "erp", "cfm", "adis": {"lin_thr2", "ang_thr2",
            "samples", "idle_time", "idle_steps"}}
  bodies = [{"pos": (3,), "q": (4,), "lvel", "avel", "mass", "I": (12,),
             "c": (3,), "enabled", "gravity", "adis": None or {... "time_left",
             "steps_left"}, "facc", "tacc", "geoms": [geom, ...]}]  (list order;
            geoms most recently attached first)
  geom   = {"cls": 0 sphere | 1 box | 2 capsule | 4 plane, "p": (4,),
            "cat": int, "col": int, "enabled": bool}
  joints = [{"type": "ball" | "contact" | "hinge" | "slider" | "fixed",
             "bodies": (i, j or None), ...}]
            (newest first; ball: anchor1, anchor2; contact: pos, normal, depth,
            mode, mu, mu2, bounce, bounce_vel, soft_erp, soft_cfm, motion1,
            motion2, slip1, slip2, fdir1; hinge: anchor1, anchor2, axis1, axis2,
            qrel, limot; slider: axis1, qrel, offset, limot; fixed: qrel,
            offset; limot: lostop, histop, bounce, stop_erp, stop_cfm, vel,
            fmax, fudge, normal_cfm)

Parity note: the reference holds no DIF fixtures and its exporter needs the
whole library to build, so this restatement is checked against the library
and by round trips, not against reference output.
"""
import numpy as np

ALL_BITS = (1 << 64) - 1


def _real(x):
    x = float(np.float32(x))
    if x == float("inf"):
        return "inf"
    if x == float("-inf"):
        return "-inf"
    return "%.7g" % x


class _W:
    def __init__(self):
        self.out = []
        self.indent = 1

    def raw(self, s):
        self.out.append(s)

    def tabs(self):
        self.out.append("\t" * self.indent)

    def line(self, name):
        self.tabs()
        self.raw(name + "\n")

    def num(self, name, x, integer=False):
        self.tabs()
        self.raw("%s = %s,\n" % (name, ("%d" % x) if integer else _real(x)))

    def vec(self, name, x, n=3):
        self.tabs()
        self.raw("%s = {%s},\n" % (name, ",".join(_real(v) for v in list(x)[:n])))

    def str_(self, name, x):
        self.tabs()
        self.raw('%s = "%s",\n' % (name, x))

    def vec_nonzero(self, name, x):
        if x[0] != 0 and x[1] != 0 and x[2] != 0:
            self.vec(name, x)


def _geom(c, g):
    if g.get("cat", ALL_BITS) != ALL_BITS:
        c.tabs()
        c.raw("category_bits = %d\n" % g["cat"])
    if g.get("col", ALL_BITS) != ALL_BITS:
        c.tabs()
        c.raw("collide_bits = %d\n" % g["col"])
    if not g.get("enabled", True):
        c.num("disabled", 1, integer=True)
    p = g["p"]
    if g["cls"] == 0:
        c.str_("type", "sphere")
        c.num("radius", p[0])
    elif g["cls"] == 1:
        c.str_("type", "box")
        c.vec("sides", p)
    elif g["cls"] == 2:
        c.str_("type", "capsule")
        c.num("radius", p[0])
        c.num("length", p[1])
    elif g["cls"] == 4:
        c.str_("type", "plane")
        c.vec("normal", p)
        c.num("d", p[3])


def _limot(c, lm):
    """printLimot export-dif.cpp:142-183; lm: lostop, histop, bounce,
    stop_erp, stop_cfm, vel, fmax, fudge, normal_cfm."""
    c.line("limit = {")
    c.indent += 1
    c.num("low_stop", lm["lostop"])
    c.num("high_stop", lm["histop"])
    if lm["bounce"] != 0:
        c.num("bounce", lm["bounce"])
    c.line("ODE = {")
    c.indent += 1
    if lm["stop_erp"] != 0:
        c.num("stop_erp", lm["stop_erp"])
    if lm["stop_cfm"] != 0:
        c.num("stop_cfm", lm["stop_cfm"])
    c.indent -= 1
    c.line("},")
    c.indent -= 1
    c.line("},")
    c.line("motor = {")
    c.indent += 1
    if lm["vel"] != 0:
        c.num("vel", lm["vel"])
    if lm["fmax"] != 0:
        c.num("fmax", lm["fmax"])
    c.line("ODE = {")
    c.indent += 1
    if lm["fudge"] != 0:
        c.num("fudge_factor", lm["fudge"])
    if lm["normal_cfm"] != 0:
        c.num("normal_cfm", lm["normal_cfm"])
    c.indent -= 1
    c.line("},")
    c.indent -= 1
    c.line("},")


def export(world, bodies, joints, prefix=""):
    c = _W()
    c.raw("-- Dynamics Interchange Format v0.1\n\n%sworld = dynamics.world {\n" % prefix)
    c.vec("gravity", world["gravity"])
    c.line("ODE = {")
    c.indent += 1
    c.num("ERP", world["erp"])
    c.num("CFM", world["cfm"])
    c.line("auto_disable = {")
    c.indent += 1
    a = world["adis"]
    c.num("linear_threshold", a["lin_thr2"])
    c.num("angular_threshold", a["ang_thr2"])
    c.num("average_samples", a["samples"], integer=True)
    c.num("idle_time", a["idle_time"])
    c.num("idle_steps", a["idle_steps"], integer=True)
    c.raw("\t\t},\n\t},\n}\n")
    c.indent -= 3
    c.raw("%sbody = {}\n" % prefix)
    for num, b in enumerate(bodies):
        c.raw("%sbody[%d] = dynamics.body {\n\tworld = %sworld,\n" % (prefix, num, prefix))
        c.indent += 1
        c.vec("pos", b["pos"])
        c.vec("q", b["q"], 4)
        c.vec("lvel", b["lvel"])
        c.vec("avel", b["avel"])
        c.num("mass", b["mass"])
        I = b["I"]
        c.raw("\tI = {" + ",".join("{" + ",".join(_real(I[i * 4 + k]) for k in range(3)) + "}" for i in range(3))
              + "},\n")
        c.vec_nonzero("com", b["c"])
        c.line("ODE = {")
        c.indent += 1
        if not b["enabled"]:
            c.num("disabled", 1, integer=True)
        if not b["gravity"]:
            c.num("no_gravity", 1, integer=True)
        if b["adis"] is not None:
            d = b["adis"]
            c.line("auto_disable = {")
            c.indent += 1
            c.num("linear_threshold", d["lin_thr2"])
            c.num("angular_threshold", d["ang_thr2"])
            c.num("average_samples", d["samples"], integer=True)
            c.num("idle_time", d["idle_time"])
            c.num("idle_steps", d["idle_steps"], integer=True)
            c.num("time_left", d["time_left"])
            c.num("steps_left", d["steps_left"], integer=True)
            c.indent -= 1
            c.line("},")
        c.vec_nonzero("facc", b["facc"])
        c.vec_nonzero("tacc", b["tacc"])
        c.indent -= 1
        c.line("},")
        if b["geoms"]:
            c.line("geometry = {")
            c.indent += 1
            for g in b["geoms"]:
                c.line("{")
                c.indent += 1
                _geom(c, g)
                c.indent -= 1
                c.line("},")
            c.indent -= 1
            c.line("},")
        c.indent -= 1
        c.line("}")
    c.raw("%sjoint = {}\n" % prefix)
    for num, j in enumerate(joints):
        c.indent += 1
        c.raw("%sjoint[%d] = dynamics.%s_joint {\n\tworld = %sworld,\n\tbody = {" % (prefix, num, j["type"], prefix))
        b1, b2 = j["bodies"]
        if b1 is not None:
            c.raw("%sbody[%d]" % (prefix, b1))
        if b2 is not None:
            c.raw(",%sbody[%d]" % (prefix, b2))
        c.raw("}\n")
        if j["type"] == "ball":
            c.vec("anchor1", j["anchor1"])
            c.vec("anchor2", j["anchor2"])
        elif j["type"] == "hinge":
            c.vec("anchor1", j["anchor1"])
            c.vec("anchor2", j["anchor2"])
            c.vec("axis1", j["axis1"])
            c.vec("axis2", j["axis2"])
            c.vec("qrel", j["qrel"], 4)
            _limot(c, j["limot"])
        elif j["type"] == "slider":
            c.vec("axis1", j["axis1"])
            c.vec("qrel", j["qrel"], 4)
            c.vec("offset", j["offset"])
            _limot(c, j["limot"])
        elif j["type"] == "fixed":
            c.vec("qrel", j["qrel"])      # three components, as printFixed prints it
            c.vec("offset", j["offset"])
        else:
            mode = j["mode"]
            c.vec("pos", j["pos"])
            c.vec("normal", j["normal"])
            c.num("depth", j["depth"])
            c.num("mu", j["mu"])
            for bit, name, key in ((0x001, "mu2", "mu2"), (0x004, "bounce", "bounce"), (0x004, "bounce_vel", "bounce_vel"),
                                   (0x008, "soft_ERP", "soft_erp"), (0x010, "soft_CFM", "soft_cfm"),
                                   (0x020, "motion1", "motion1"), (0x040, "motion2", "motion2"),
                                   (0x100, "slip1", "slip1"), (0x200, "slip2", "slip2")):
                if mode & bit:
                    c.num(name, j[key])
            fa = (1 if mode & 0x1000 else 0) | (2 if mode & 0x2000 else 0)
            if fa:
                c.num("friction_approximation", fa, integer=True)
            if mode & 0x002:
                c.vec("fdir1", j["fdir1"])
        c.indent -= 1
        c.line("}")
    return "".join(c.out)
