"""Python front end of the ODE-compatible C API in libodx.so (include/ode.h).

A thin ctypes layer with the call shapes of the reference's Cython binding
(bindings/python/ode.pyx: World.quickStep, Body, Space.collide, collide,
ContactJoint, JointGroup) over the same C entry points an ODE program links.
"""
import ctypes

from . import odx

dReal = ctypes.c_float
dVector3 = dReal * 4
dMatrix3 = dReal * 12


class dSurfaceParameters(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int), ("mu", dReal), ("mu2", dReal), ("rho", dReal), ("rho2", dReal),
                ("rhoN", dReal), ("bounce", dReal), ("bounce_vel", dReal), ("soft_erp", dReal), ("soft_cfm", dReal),
                ("motion1", dReal), ("motion2", dReal), ("motionN", dReal), ("slip1", dReal), ("slip2", dReal)]


class dContactGeom(ctypes.Structure):
    _fields_ = [("pos", dVector3), ("normal", dVector3), ("depth", dReal), ("g1", ctypes.c_void_p),
                ("g2", ctypes.c_void_p), ("side1", ctypes.c_int), ("side2", ctypes.c_int)]


class dContact(ctypes.Structure):
    _fields_ = [("surface", dSurfaceParameters), ("geom", dContactGeom), ("fdir1", dVector3)]


class dMass(ctypes.Structure):
    _fields_ = [("mass", dReal), ("c", dVector3), ("I", dMatrix3)]


class dJointFeedback(ctypes.Structure):
    _fields_ = [("f1", dVector3), ("t1", dVector3), ("f2", dVector3), ("t2", dVector3)]


NearCallback = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p)

dContactMu2, dContactFDir1, dContactBounce, dContactSoftERP, dContactSoftCFM = 0x1, 0x2, 0x4, 0x8, 0x10
dContactApprox1 = 0x7000
dSphereClass, dBoxClass, dCapsuleClass, dPlaneClass = 0, 1, 2, 4
dJointTypeBall, dJointTypeContact = 1, 4
dSAP_AXES_XYZ = (0) | (1 << 2) | (2 << 4)
dSAP_AXES_ZYX = (2) | (1 << 2) | (0 << 4)

_api = None


def api():
    """libodx.so with the ODE entry points' signatures declared."""
    global _api
    if _api is not None:
        return _api
    L = odx.lib()
    vp, f, i = ctypes.c_void_p, dReal, ctypes.c_int
    P = ctypes.POINTER
    sig = {
        "dInitODE2": (i, [ctypes.c_uint]), "dCloseODE": (None, []),
        "dWorldCreate": (vp, []), "dWorldDestroy": (None, [vp]),
        "dWorldSetGravity": (None, [vp, f, f, f]), "dWorldSetERP": (None, [vp, f]), "dWorldSetCFM": (None, [vp, f]),
        "dWorldGetERP": (f, [vp]), "dWorldGetCFM": (f, [vp]),
        "dWorldQuickStep": (i, [vp, f]), "dWorldSetQuickStepNumIterations": (None, [vp, i]),
        "dWorldGetQuickStepNumIterations": (i, [vp]), "dWorldSetQuickStepW": (None, [vp, f]),
        "dWorldSetContactMaxCorrectingVel": (None, [vp, f]), "dWorldSetContactSurfaceLayer": (None, [vp, f]),
        "odxWorldSetRowOrder": (None, [vp, i]), "odxWorldSetSweepMode": (None, [vp, i]), "odxWorldGetSolveStats": (None, [vp, P(i), P(ctypes.c_float), P(i)]),
        "dWorldSetQuickStepDynamicIterationParameters": (None, [vp, P(dReal), P(dReal), P(dReal)]),
        "dWorldGetQuickStepDynamicIterationParameters": (None, [vp, P(dReal), P(dReal), P(dReal)]),
        "dBodyCreate": (vp, [vp]), "dBodyDestroy": (None, [vp]),
        "dBodySetPosition": (None, [vp, f, f, f]), "dBodySetQuaternion": (None, [vp, P(dReal)]),
        "dBodySetRotation": (None, [vp, P(dReal)]),
        "dBodySetLinearVel": (None, [vp, f, f, f]), "dBodySetAngularVel": (None, [vp, f, f, f]),
        "dBodyGetPosition": (P(dReal), [vp]), "dBodyGetQuaternion": (P(dReal), [vp]),
        "dBodyGetRotation": (P(dReal), [vp]), "dBodyGetLinearVel": (P(dReal), [vp]),
        "dBodyGetAngularVel": (P(dReal), [vp]), "dBodySetMass": (None, [vp, P(dMass)]),
        "dBodyGetMass": (None, [vp, P(dMass)]),
        "dBodyAddForce": (None, [vp, f, f, f]), "dBodyAddTorque": (None, [vp, f, f, f]),
        "dBodySetGravityMode": (None, [vp, i]), "dBodySetGyroscopicMode": (None, [vp, i]),
        "dMassSetBox": (None, [P(dMass), f, f, f, f]), "dMassSetSphere": (None, [P(dMass), f, f]),
        "dMassSetCapsule": (None, [P(dMass), f, i, f, f]),
        "dMassSetParameters": (None, [P(dMass), f, f, f, f, f, f, f, f, f, f]),
        "dCreateSphere": (vp, [vp, f]), "dCreateBox": (vp, [vp, f, f, f]), "dCreatePlane": (vp, [vp, f, f, f, f]),
        "dCreateCapsule": (vp, [vp, f, f]), "dGeomSetBody": (None, [vp, vp]), "dGeomGetBody": (vp, [vp]),
        "dGeomSetPosition": (None, [vp, f, f, f]), "dGeomGetClass": (i, [vp]), "dGeomDestroy": (None, [vp]),
        "dSimpleSpaceCreate": (vp, [vp]), "dHashSpaceCreate": (vp, [vp]),
        "dSweepAndPruneSpaceCreate": (vp, [vp, i]), "dSpaceDestroy": (None, [vp]),
        "dHashSpaceSetLevels": (None, [vp, i, i]), "dSpaceCollide": (None, [vp, vp, NearCallback]),
        "dSpaceAdd": (None, [vp, vp]), "dSpaceRemove": (None, [vp, vp]), "dSpaceGetNumGeoms": (i, [vp]),
        "dGeomGetAABB": (None, [vp, P(dReal)]), "dGeomEnable": (None, [vp]), "dGeomDisable": (None, [vp]),
        "dGeomSetCategoryBits": (None, [vp, ctypes.c_ulong]), "dGeomSetCollideBits": (None, [vp, ctypes.c_ulong]),
        "dGeomSetQuaternion": (None, [vp, P(dReal)]),
        "dCollide": (i, [vp, vp, i, P(dContactGeom), i]),
        "dJointGroupCreate": (vp, [i]), "dJointGroupDestroy": (None, [vp]), "dJointGroupEmpty": (None, [vp]),
        "dJointCreateContact": (vp, [vp, vp, P(dContact)]), "dJointCreateBall": (vp, [vp, vp]),
        "dJointAttach": (None, [vp, vp, vp]), "dJointSetBallAnchor": (None, [vp, f, f, f]),
        "dJointSetFeedback": (None, [vp, P(dJointFeedback)]), "dJointGetType": (i, [vp]),
        "dAreConnectedExcluding": (i, [vp, vp, i]),
        "dRand": (ctypes.c_ulong, []), "dRandSetSeed": (None, [ctypes.c_ulong]), "dRandInt": (i, [i]),
        "dRandGetSeed": (ctypes.c_ulong, []),
        "dBodyEnable": (None, [vp]), "dBodyDisable": (None, [vp]), "dBodyIsEnabled": (i, [vp]),
    }
    u = ctypes.c_uint
    # auto-disable, damping and angular speed cap: world defaults and per body
    for who in ("World", "Body"):
        for name in ("AutoDisableLinearThreshold", "AutoDisableAngularThreshold", "AutoDisableTime",
                     "LinearDampingThreshold", "AngularDampingThreshold", "LinearDamping", "AngularDamping",
                     "MaxAngularSpeed"):
            sig["d%sGet%s" % (who, name)] = (f, [vp])
            sig["d%sSet%s" % (who, name)] = (None, [vp, f])
        for name in ("AutoDisableSteps", "AutoDisableFlag"):
            sig["d%sGet%s" % (who, name)] = (i, [vp])
            sig["d%sSet%s" % (who, name)] = (None, [vp, i])
        sig["d%sGetAutoDisableAverageSamplesCount" % who] = (i, [vp])
        sig["d%sSetAutoDisableAverageSamplesCount" % who] = (None, [vp, u])
        sig["d%sSetDamping" % who] = (None, [vp, f, f])
    sig["dBodySetAutoDisableDefaults"] = (None, [vp])
    sig["dBodySetDampingDefaults"] = (None, [vp])
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    L.dInitODE2(0)
    _api = L
    return L


def vec(ptr, n=3):
    return [ptr[k] for k in range(n)]
