"""The oracle's Dynamics Interchange Format text (oracle/dif.py): the literal
lines of ode/src/export-dif.cpp, number formatting of a dSINGLE build, and a
round trip of every printed value."""
import re

import numpy as np

from oracle import dif

F = np.float32


def small_world():
    world = {"gravity": (0.0, 0.0, -9.81), "erp": 0.2, "cfm": 1e-5,
             "adis": {"lin_thr2": F(0.01) * F(0.01), "ang_thr2": F(0.01) * F(0.01), "samples": 1, "idle_time": 0.0,
                      "idle_steps": 10}}
    I = np.zeros(12)
    I[0], I[5], I[10] = 0.1, 0.2, 0.3
    bodies = [
        {"pos": (1.0, 2.5, -3.25), "q": (1.0, 0.0, 0.0, 0.0), "lvel": (0.0, 0.0, 0.0), "avel": (0.125, 0.0, 0.0),
         "mass": 2.0, "I": I, "c": (0.0, 0.0, 0.0), "enabled": True, "gravity": True, "adis": None,
         "facc": (0.0, 0.0, 0.0), "tacc": (1.0, 2.0, 3.0),
         "geoms": [{"cls": 1, "p": (0.5, 0.25, 1.0, 0.0), "cat": dif.ALL_BITS, "col": dif.ALL_BITS, "enabled": True},
                   {"cls": 0, "p": (0.75, 0, 0, 0), "cat": 6, "col": dif.ALL_BITS, "enabled": False}]},
        {"pos": (0.0, 0.0, 1e-9), "q": (0.5, 0.5, 0.5, 0.5), "lvel": (float("inf"), 0.0, 0.0), "avel": (0, 0, 0),
         "mass": 1.0 / 3.0, "I": I, "c": (0.1, 0.2, 0.3), "enabled": False, "gravity": False,
         "adis": {"lin_thr2": 1e-4, "ang_thr2": 4e-4, "samples": 3, "idle_time": 0.5, "idle_steps": 7, "time_left": 0.25,
                  "steps_left": 2},
         "facc": (0.0, 1.0, 1.0), "tacc": (0.0, 0.0, 0.0), "geoms": []},
    ]
    joints = [
        {"type": "contact", "bodies": (1, None), "pos": (0.0, 0.0, 0.0), "normal": (0.0, 0.0, 1.0), "depth": 0.01,
         "mode": 0x001 | 0x004 | 0x1000 | 0x2000 | 0x002, "mu": float("inf"), "mu2": 0.5, "bounce": 0.25,
         "bounce_vel": 0.1, "fdir1": (1.0, 0.0, 0.0)},
        {"type": "ball", "bodies": (0, 1), "anchor1": (0.5, 0.0, 0.0), "anchor2": (-0.5, 0.0, 0.0)},
    ]
    return world, bodies, joints


def test_literal_lines_of_the_reference_format():
    text = dif.export(*small_world(), prefix="X")
    lines = text.split("\n")
    # export-dif.cpp:505-520
    assert lines[0] == "-- Dynamics Interchange Format v0.1"
    assert lines[1] == ""
    assert lines[2] == "Xworld = dynamics.world {"
    assert lines[3] == "\tgravity = {0,0,-9.81},"
    assert lines[4] == "\tODE = {"
    assert lines[5] == "\t\tERP = 0.2,"
    assert lines[6] == "\t\tCFM = 1e-05,"
    assert lines[7] == "\t\tauto_disable = {"
    assert lines[8] == "\t\t\tlinear_threshold = 0.0001,"
    assert lines[10] == "\t\t\taverage_samples = 1,"
    assert lines[13:16] == ["\t\t},", "\t},", "}"]
    # bodies (:523-595) and joints (:598-619)
    assert "Xbody = {}" in lines and "Xjoint = {}" in lines
    assert "Xbody[0] = dynamics.body {" in lines and "\tworld = Xworld," in lines
    assert "\tI = {{0.1,0,0},{0,0.2,0},{0,0,0.3}}," in lines
    assert "Xjoint[0] = dynamics.contact_joint {" in lines
    assert "\tbody = {Xbody[1]}" in lines and "\tbody = {Xbody[0],Xbody[1]}" in lines
    assert "\tmu = inf," in lines and "\tlvel = {inf,0,0}," in lines
    assert "\tfriction_approximation = 3," in lines and "\tfdir1 = {1,0,0}," in lines
    # category bits without a trailing comma (:462-470); a disabled geom
    assert "\t\t\tcategory_bits = 6" in lines and "\t\t\tdisabled = 1," in lines
    # all-nonzero vectors only (:133-136): com (0,0,0) and facc (0,1,1) are left out
    assert "\tcom = {0.1,0.2,0.3}," in lines
    assert not any(l.startswith("\t\tfacc") for l in lines)
    assert "\t\ttacc = {1,2,3}," in lines
    # the body auto-disable block with its countdowns
    assert "\t\t\tsteps_left = 2," in lines and "\t\t\ttime_left = 0.25," in lines
    assert text.endswith("}\n")


def test_reals_have_seven_significant_digits_of_float32():
    world, bodies, joints = small_world()
    text = dif.export(world, bodies, joints)
    assert "\tmass = 0.3333333," in text.split("\n")
    # 1e-9 as float32 is 9.99999972e-10: 7 digits give 1e-09
    assert "\tpos = {0,0,1e-09}," in text.split("\n")


def test_every_printed_real_round_trips_to_its_float32():
    rng = np.random.default_rng(3)
    world, bodies, joints = small_world()
    for _ in range(50):
        v = rng.normal(0, 10.0 ** rng.integers(-6, 6), size=3).astype(np.float32)
        bodies[0]["pos"] = tuple(float(x) for x in v)
        text = dif.export(world, bodies, joints)
        m = re.search(r"\n\tpos = \{([^}]*)\},\n", text)
        got = np.array([float(x) for x in m.group(1).split(",")], np.float32)
        np.testing.assert_allclose(got, v, rtol=1e-6)
