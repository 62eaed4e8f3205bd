/*
 * ode.h -- the ODE public API subset on the step hot path, implemented by
 * libodx on one MI355X (dReal = float).
 *
 * A program written against the reference's include/ode/ *.h for this path
 * compiles against this header unchanged and links libodx.so instead of
 * libode.  Each declaration names the reference declaration it replaces
 * (include/ode/<file>:<line> of thomasmarsh/ODE).  Semantics kept: normal into
 * body 1, reversed attach negates the normal, world ERP/CFM defaults 0.2/1e-5,
 * SOR w = 1.3, gyroscopic torque on, QuickStep row order with findex rows last,
 * the dynamic iteration count per island (premature exit 1e-8, up to
 * iterations x 1.0 extra sweeps while an adjustment is >= 1e-2; all three set
 * to 0 give a fixed count).
 * Differences: the solve is deterministic unless odxWorldSetRowOrder(w, 2)
 * selects the reference's random reordering; bodies may carry
 * sphere/box/capsule/plane geoms; joints are contact and ball.
 */
#ifndef ODX_ODE_H
#define ODX_ODE_H

#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef float dReal;                         /* include/ode/precision.h (dSINGLE) */
typedef dReal dVector3[4];                   /* include/ode/common.h:270 */
typedef dReal dVector4[4];                   /* include/ode/common.h:271 */
typedef dReal dMatrix3[4 * 3];               /* include/ode/common.h:272 */
typedef dReal dMatrix4[4 * 4];
typedef dReal dQuaternion[4];                /* include/ode/common.h:275 */

typedef struct dxWorld *dWorldID;            /* include/ode/common.h:364 */
typedef struct dxSpace *dSpaceID;
typedef struct dxBody *dBodyID;
typedef struct dxGeom *dGeomID;
typedef struct dxJoint *dJointID;
typedef struct dxJointGroup *dJointGroupID;

#define dInfinity (__builtin_huge_valf())

typedef enum {                               /* include/ode/common.h:382-400 */
    dJointTypeNone = 0,
    dJointTypeBall,
    dJointTypeHinge,
    dJointTypeSlider,
    dJointTypeContact,
    dJointTypeUniversal,
    dJointTypeHinge2,
    dJointTypeFixed
} dJointType;

enum {                                       /* include/ode/common.h:439-456 (group 1) */
    dParamLoStop = 0, dParamHiStop, dParamVel, dParamLoVel, dParamHiVel, dParamFMax, dParamFudgeFactor,
    dParamBounce, dParamCFM, dParamStopERP, dParamStopCFM, dParamSuspensionERP, dParamSuspensionCFM, dParamERP
};

typedef struct dJointFeedback {              /* include/ode/common.h:516-521 */
    dVector3 f1, t1, f2, t2;
} dJointFeedback;

typedef struct dMass {                       /* include/ode/mass.h:88-91 */
    dReal mass;
    dVector3 c;
    dMatrix3 I;
} dMass;

enum {                                       /* include/ode/contact.h:33-52 */
    dContactMu2 = 0x001, dContactAxisDep = 0x001, dContactFDir1 = 0x002, dContactBounce = 0x004,
    dContactSoftERP = 0x008, dContactSoftCFM = 0x010, dContactMotion1 = 0x020, dContactMotion2 = 0x040,
    dContactMotionN = 0x080, dContactSlip1 = 0x100, dContactSlip2 = 0x200, dContactRolling = 0x400,
    dContactApprox0 = 0x0000, dContactApprox1_1 = 0x1000, dContactApprox1_2 = 0x2000,
    dContactApprox1_N = 0x4000, dContactApprox1 = 0x7000
};

typedef struct dSurfaceParameters {          /* include/ode/contact.h:55-72 */
    int mode;
    dReal mu, mu2, rho, rho2, rhoN, bounce, bounce_vel, soft_erp, soft_cfm, motion1, motion2, motionN, slip1, slip2;
} dSurfaceParameters;

typedef struct dContactGeom {                /* include/ode/contact.h:86-92 */
    dVector3 pos;
    dVector3 normal;
    dReal depth;
    dGeomID g1, g2;
    int side1, side2;
} dContactGeom;

typedef struct dContact {                    /* include/ode/contact.h:97-101 */
    dSurfaceParameters surface;
    dContactGeom geom;
    dVector3 fdir1;
} dContact;

enum { dSphereClass = 0, dBoxClass, dCapsuleClass, dCylinderClass, dPlaneClass };  /* include/ode/collision.h:877-882 */
#define dSAP_AXES_XYZ ((0) | (1 << 2) | (2 << 4))  /* include/ode/collision_space.h:59 */

typedef void dNearCallback(void *data, dGeomID o1, dGeomID o2);  /* include/ode/collision_space.h:49 */
typedef void dMessageFunction(int errnum, const char *msg, ...);

/* ---------------------------------------------------------------- init */
void dInitODE(void);                         /* include/ode/odeinit.h:95 */
int dInitODE2(unsigned int uiInitFlags);     /* include/ode/odeinit.h:119 */
void dCloseODE(void);                        /* include/ode/odeinit.h:227 */

/* --------------------------------------------------------------- world */
dWorldID dWorldCreate(void);                                   /* objects.h:52 */
void dWorldDestroy(dWorldID w);                                /* objects.h:64 */
void dWorldSetGravity(dWorldID w, dReal x, dReal y, dReal z);  /* objects.h:93 */
void dWorldGetGravity(dWorldID w, dVector3 gravity);           /* objects.h:100 */
void dWorldSetERP(dWorldID w, dReal erp);                      /* objects.h:110 */
dReal dWorldGetERP(dWorldID w);                                /* objects.h:117 */
void dWorldSetCFM(dWorldID w, dReal cfm);                      /* objects.h:127 */
dReal dWorldGetCFM(dWorldID w);                                /* objects.h:134 */
int dWorldQuickStep(dWorldID w, dReal stepsize);               /* objects.h:427 */
void dWorldSetQuickStepNumIterations(dWorldID w, int num);     /* objects.h:462 */
int dWorldGetQuickStepNumIterations(dWorldID w);               /* objects.h:470 */
void dWorldSetQuickStepDynamicIterationParameters(dWorldID w, const dReal *premature_exit_delta,
                                                  const dReal *max_num_extra_factor,
                                                  const dReal *extra_iteration_requirement_delta); /* objects.h:509 */
void dWorldGetQuickStepDynamicIterationParameters(dWorldID w, dReal *premature_exit_delta,
                                                  dReal *max_num_extra_factor,
                                                  dReal *extra_iteration_requirement_delta); /* objects.h:528 */
/* Dynamic-iteration statistics (include/ode/objects.h:538-556): per island
 * step, premature_exits counts solves that stopped before the iteration
 * count, prolonged_execs solves that stopped inside the extra iterations,
 * full_extra_execs solves that used every extra iteration; iteration_count
 * is not written (as in the reference's QuickStep). */
typedef struct {
    unsigned struct_size;
    unsigned iteration_count;
    unsigned premature_exits;
    unsigned prolonged_execs;
    unsigned full_extra_execs;
} dWorldQuickStepIterationCount_DynamicAdjustmentStatistics;
static inline void dWorldInitializeQuickStepIterationCount_DynamicAdjustmentStatistics(
    dWorldQuickStepIterationCount_DynamicAdjustmentStatistics *ptr_stat) {
    ptr_stat->struct_size = sizeof(*ptr_stat);
    ptr_stat->iteration_count = 0;
    ptr_stat->premature_exits = 0;
    ptr_stat->prolonged_execs = 0;
    ptr_stat->full_extra_execs = 0;
}
/* Attaches var_stats (NULL: the world's own) to the world's QuickStep; 0 when
 * var_stats->struct_size is too small or not a multiple of 4. */
int dWorldAttachQuickStepDynamicIterationStatisticsSink(
    dWorldID w, dWorldQuickStepIterationCount_DynamicAdjustmentStatistics *var_stats); /* objects.h:576 */
void dWorldSetQuickStepW(dWorldID w, dReal over_relaxation);   /* objects.h:584 */
dReal dWorldGetQuickStepW(dWorldID w);                         /* objects.h:591 */
void dWorldSetContactMaxCorrectingVel(dWorldID w, dReal vel);  /* objects.h:603 */
void dWorldSetContactSurfaceLayer(dWorldID w, dReal depth);    /* objects.h:623 */
/* Defaults that bodies take at dBodyCreate: auto-disable (objects.h:670-748),
 * damping and the angular speed cap (objects.h:790-865).  Thresholds are
 * speeds; the world keeps them squared, as the reference does. */
dReal dWorldGetAutoDisableLinearThreshold(dWorldID w);                        /* objects.h:670 */
void dWorldSetAutoDisableLinearThreshold(dWorldID w, dReal threshold);        /* objects.h:677 */
dReal dWorldGetAutoDisableAngularThreshold(dWorldID w);                       /* objects.h:684 */
void dWorldSetAutoDisableAngularThreshold(dWorldID w, dReal threshold);       /* objects.h:691 */
int dWorldGetAutoDisableAverageSamplesCount(dWorldID w);                      /* objects.h:698 */
void dWorldSetAutoDisableAverageSamplesCount(dWorldID w, unsigned int count); /* objects.h:706 */
int dWorldGetAutoDisableSteps(dWorldID w);                                    /* objects.h:713 */
void dWorldSetAutoDisableSteps(dWorldID w, int steps);                        /* objects.h:720 */
dReal dWorldGetAutoDisableTime(dWorldID w);                                   /* objects.h:727 */
void dWorldSetAutoDisableTime(dWorldID w, dReal time);                        /* objects.h:734 */
int dWorldGetAutoDisableFlag(dWorldID w);                                     /* objects.h:741 */
void dWorldSetAutoDisableFlag(dWorldID w, int do_auto_disable);               /* objects.h:748 */
dReal dWorldGetLinearDampingThreshold(dWorldID w);                            /* objects.h:790 */
void dWorldSetLinearDampingThreshold(dWorldID w, dReal threshold);            /* objects.h:798 */
dReal dWorldGetAngularDampingThreshold(dWorldID w);                           /* objects.h:804 */
void dWorldSetAngularDampingThreshold(dWorldID w, dReal threshold);           /* objects.h:812 */
dReal dWorldGetLinearDamping(dWorldID w);                                     /* objects.h:818 */
void dWorldSetLinearDamping(dWorldID w, dReal scale);                         /* objects.h:826 */
dReal dWorldGetAngularDamping(dWorldID w);                                    /* objects.h:832 */
void dWorldSetAngularDamping(dWorldID w, dReal scale);                        /* objects.h:840 */
void dWorldSetDamping(dWorldID w, dReal linear_scale, dReal angular_scale);   /* objects.h:848 */
dReal dWorldGetMaxAngularSpeed(dWorldID w);                                   /* objects.h:857 */
void dWorldSetMaxAngularSpeed(dWorldID w, dReal max_speed);                   /* objects.h:865 */

/* ---------------------------------------------------------------- body */
dBodyID dBodyCreate(dWorldID w);                               /* objects.h:1011 */
void dBodyDestroy(dBodyID b);                                  /* objects.h:1021 */
void dBodySetData(dBodyID b, void *data);                      /* objects.h:1028 */
void *dBodyGetData(dBodyID b);                                 /* objects.h:1035 */
void dBodySetPosition(dBodyID b, dReal x, dReal y, dReal z);   /* objects.h:1045 */
void dBodySetRotation(dBodyID b, const dMatrix3 R);            /* objects.h:1055 */
void dBodySetQuaternion(dBodyID b, const dQuaternion q);       /* objects.h:1065 */
void dBodySetLinearVel(dBodyID b, dReal x, dReal y, dReal z);  /* objects.h:1071 */
void dBodySetAngularVel(dBodyID b, dReal x, dReal y, dReal z); /* objects.h:1077 */
const dReal *dBodyGetPosition(dBodyID b);                      /* objects.h:1088 */
const dReal *dBodyGetRotation(dBodyID b);                      /* objects.h:1106 */
const dReal *dBodyGetQuaternion(dBodyID b);                    /* objects.h:1124 */
const dReal *dBodyGetLinearVel(dBodyID b);                     /* objects.h:1141 */
const dReal *dBodyGetAngularVel(dBodyID b);                    /* objects.h:1147 */
void dBodySetMass(dBodyID b, const dMass *mass);               /* objects.h:1153 */
void dBodyGetMass(dBodyID b, dMass *mass);                     /* objects.h:1159 */
void dBodyAddForce(dBodyID b, dReal fx, dReal fy, dReal fz);   /* objects.h:1165 */
void dBodyAddTorque(dBodyID b, dReal fx, dReal fy, dReal fz);  /* objects.h:1171 */
void dBodyEnable(dBodyID b);                                   /* objects.h:1424 */
void dBodyDisable(dBodyID b);                                  /* objects.h:1433 */
int dBodyIsEnabled(dBodyID b);                                 /* objects.h:1440 */
void dBodySetGravityMode(dBodyID b, int mode);                 /* objects.h:1449 */
void dBodySetGyroscopicMode(dBodyID b, int enabled);           /* objects.h:1613 */
/* Auto-disable (objects.h:908-994): dWorldQuickStep first samples the
 * velocities of each enabled auto-disabling body that has joints; a body idle
 * for the step and time budgets is disabled with zero velocities and left out
 * of the islands until a joint to an enabled body re-enables it. */
dReal dBodyGetAutoDisableLinearThreshold(dBodyID b);                          /* objects.h:908 */
void dBodySetAutoDisableLinearThreshold(dBodyID b, dReal threshold);          /* objects.h:915 */
dReal dBodyGetAutoDisableAngularThreshold(dBodyID b);                         /* objects.h:922 */
void dBodySetAutoDisableAngularThreshold(dBodyID b, dReal threshold);         /* objects.h:929 */
int dBodyGetAutoDisableAverageSamplesCount(dBodyID b);                        /* objects.h:936 */
void dBodySetAutoDisableAverageSamplesCount(dBodyID b, unsigned int count);   /* objects.h:943 */
int dBodyGetAutoDisableSteps(dBodyID b);                                      /* objects.h:951 */
void dBodySetAutoDisableSteps(dBodyID b, int steps);                          /* objects.h:958 */
dReal dBodyGetAutoDisableTime(dBodyID b);                                     /* objects.h:965 */
void dBodySetAutoDisableTime(dBodyID b, dReal time);                          /* objects.h:972 */
int dBodyGetAutoDisableFlag(dBodyID b);                                       /* objects.h:979 */
void dBodySetAutoDisableFlag(dBodyID b, int do_auto_disable);                 /* objects.h:986 */
void dBodySetAutoDisableDefaults(dBodyID b);                                  /* objects.h:994 */
/* Damping and the angular speed cap, applied by the kernel's integration
 * (dxStepBody, ode/src/util.cpp:583-699): the cap before the move, damping
 * (velocity times 1 - scale when the squared speed exceeds the squared
 * threshold) after it. */
void dBodySetDampingDefaults(dBodyID b);                                      /* objects.h:1500 */
dReal dBodyGetLinearDamping(dBodyID b);                                       /* objects.h:1506 */
void dBodySetLinearDamping(dBodyID b, dReal scale);                           /* objects.h:1516 */
dReal dBodyGetAngularDamping(dBodyID b);                                      /* objects.h:1524 */
void dBodySetAngularDamping(dBodyID b, dReal scale);                          /* objects.h:1534 */
void dBodySetDamping(dBodyID b, dReal linear_scale, dReal angular_scale);     /* objects.h:1543 */
dReal dBodyGetLinearDampingThreshold(dBodyID b);                              /* objects.h:1549 */
void dBodySetLinearDampingThreshold(dBodyID b, dReal threshold);              /* objects.h:1557 */
dReal dBodyGetAngularDampingThreshold(dBodyID b);                             /* objects.h:1563 */
void dBodySetAngularDampingThreshold(dBodyID b, dReal threshold);             /* objects.h:1571 */
dReal dBodyGetMaxAngularSpeed(dBodyID b);                                     /* objects.h:1578 */
void dBodySetMaxAngularSpeed(dBodyID b, dReal max_speed);                     /* objects.h:1588 */

/* --------------------------------------------------------------- joints */
dJointID dJointCreateBall(dWorldID w, dJointGroupID g);                    /* objects.h:1693 */
dJointID dJointCreateContact(dWorldID w, dJointGroupID g, const dContact *c); /* objects.h:1717 */
void dJointDestroy(dJointID j);                                            /* objects.h:1827 */
dJointGroupID dJointGroupCreate(int max_size);                             /* objects.h:1835 */
void dJointGroupDestroy(dJointGroupID g);                                  /* objects.h:1843 */
void dJointGroupEmpty(dJointGroupID g);                                    /* objects.h:1852 */
void dJointAttach(dJointID j, dBodyID body1, dBodyID body2);               /* objects.h:1873 */
dJointType dJointGetType(dJointID j);                                      /* objects.h:1929 */
dBodyID dJointGetBody(dJointID j, int index);                              /* objects.h:1941 */
void dJointSetFeedback(dJointID j, dJointFeedback *fb);                    /* objects.h:1950 */
void dJointSetBallAnchor(dJointID j, dReal x, dReal y, dReal z);           /* objects.h:1965 */
void dJointSetBallAnchor2(dJointID j, dReal x, dReal y, dReal z);          /* objects.h:1971 */
/* Hinge, slider and fixed joints with their limits and motors.  The rows are
 * built in the step kernel; limit states are found at the start of
 * dWorldQuickStep as getInfo1 does. */
dJointID dJointCreateHinge(dWorldID w, dJointGroupID g);                   /* objects.h:1701 */
dJointID dJointCreateSlider(dWorldID w, dJointGroupID g);                  /* objects.h:1709 */
dJointID dJointCreateFixed(dWorldID w, dJointGroupID g);                   /* objects.h:1766 */
void dJointSetHingeAnchor(dJointID j, dReal x, dReal y, dReal z);          /* objects.h:1983 */
void dJointSetHingeAxis(dJointID j, dReal x, dReal y, dReal z);            /* objects.h:1991 */
void dJointSetHingeParam(dJointID j, int parameter, dReal value);          /* objects.h:2027 */
void dJointSetSliderAxis(dJointID j, dReal x, dReal y, dReal z);           /* objects.h:2043 */
void dJointSetSliderParam(dJointID j, int parameter, dReal value);         /* objects.h:2054 */
void dJointSetFixed(dJointID j);                                           /* objects.h:2451 */
void dJointSetFixedParam(dJointID j, int parameter, dReal value);          /* objects.h:2458 */
void dJointGetHingeAnchor(dJointID j, dVector3 result);                    /* objects.h:2582 */
void dJointGetHingeAxis(dJointID j, dVector3 result);                      /* objects.h:2598 */
dReal dJointGetHingeParam(dJointID j, int parameter);                      /* objects.h:2604 */
dReal dJointGetHingeAngle(dJointID j);                                     /* objects.h:2618 */
dReal dJointGetHingeAngleRate(dJointID j);                                 /* objects.h:2624 */
dReal dJointGetSliderPosition(dJointID j);                                 /* objects.h:2637 */
dReal dJointGetSliderPositionRate(dJointID j);                             /* objects.h:2643 */
void dJointGetSliderAxis(dJointID j, dVector3 result);                     /* objects.h:2649 */
dReal dJointGetSliderParam(dJointID j, int parameter);                     /* objects.h:2655 */
dReal dJointGetFixedParam(dJointID j, int parameter);                      /* objects.h:3144 */
int dAreConnected(dBodyID a, dBodyID b);                                   /* objects.h:3484 */
int dAreConnectedExcluding(dBodyID a, dBodyID b, int joint_type);          /* objects.h:3499 */

/* ----------------------------------------------------------------- mass */
void dMassSetZero(dMass *m);                                               /* mass.h:45 */
void dMassSetParameters(dMass *m, dReal themass, dReal cgx, dReal cgy, dReal cgz, dReal I11, dReal I22,
                        dReal I33, dReal I12, dReal I13, dReal I23);       /* mass.h:47 */
void dMassSetSphere(dMass *m, dReal density, dReal radius);                /* mass.h:52 */
void dMassSetSphereTotal(dMass *m, dReal total_mass, dReal radius);        /* mass.h:53 */
void dMassSetCapsule(dMass *m, dReal density, int direction, dReal radius, dReal length);       /* mass.h:55 */
void dMassSetCapsuleTotal(dMass *m, dReal total_mass, int direction, dReal radius, dReal length); /* mass.h:57 */
void dMassSetCylinder(dMass *m, dReal density, int direction, dReal radius, dReal length);      /* mass.h:60 */
void dMassSetCylinderTotal(dMass *m, dReal total_mass, int direction, dReal radius, dReal length); /* mass.h:62 */
void dMassSetBox(dMass *m, dReal density, dReal lx, dReal ly, dReal lz);   /* mass.h:65 */
void dMassSetBoxTotal(dMass *m, dReal total_mass, dReal lx, dReal ly, dReal lz); /* mass.h:67 */

/* ------------------------------------------------------------ collision */
void dGeomDestroy(dGeomID g);                                /* collision.h:65 */
void dGeomSetData(dGeomID g, void *data);                    /* collision.h:75 */
void *dGeomGetData(dGeomID g);                               /* collision.h:84 */
void dGeomSetBody(dGeomID g, dBodyID b);                     /* collision.h:105 */
dBodyID dGeomGetBody(dGeomID g);                             /* collision.h:114 */
void dGeomSetPosition(dGeomID g, dReal x, dReal y, dReal z); /* collision.h:131 */
void dGeomSetRotation(dGeomID g, const dMatrix3 R);          /* collision.h:146 */
void dGeomSetQuaternion(dGeomID g, const dQuaternion q);     /* collision.h:162 */
const dReal *dGeomGetPosition(dGeomID g);                    /* collision.h:181 */
void dGeomGetAABB(dGeomID g, dReal aabb[6]);                 /* collision.h:261 */
dSpaceID dGeomGetSpace(dGeomID g);                           /* collision.h:280 */
int dGeomGetClass(dGeomID g);                                /* collision.h:307 */
void dGeomSetCategoryBits(dGeomID g, unsigned long bits);    /* collision.h:322 */
void dGeomSetCollideBits(dGeomID g, unsigned long bits);     /* collision.h:337 */
void dGeomEnable(dGeomID g);                                 /* collision.h:374 */
void dGeomDisable(dGeomID g);                                /* collision.h:389 */
int dCollide(dGeomID o1, dGeomID o2, int flags, dContactGeom *contact, int skip); /* collision.h:792 */
void dSpaceCollide(dSpaceID space, void *data, dNearCallback *callback);          /* collision.h:822 */
dGeomID dCreateSphere(dSpaceID space, dReal radius);         /* collision.h:921 */
dReal dGeomSphereGetRadius(dGeomID sphere);                  /* collision.h:944 */
dGeomID dCreateBox(dSpaceID space, dReal lx, dReal ly, dReal lz); /* collision.h:1001 */
void dGeomBoxGetLengths(dGeomID box, dVector3 result);       /* collision.h:1027 */
dGeomID dCreatePlane(dSpaceID space, dReal a, dReal b, dReal c, dReal d); /* collision.h:1045 */
void dGeomPlaneGetParams(dGeomID plane, dVector4 result);    /* collision.h:1047 */
dGeomID dCreateCapsule(dSpaceID space, dReal radius, dReal length); /* collision.h:1050 */
void dGeomCapsuleGetParams(dGeomID capsule, dReal *radius, dReal *length); /* collision.h:1052 */
/* Cylinders collide with planes only in this build (dCollideCylinderPlane,
 * ode/src/collision_cylinder_plane.cpp); dCollide of a cylinder with any
 * other class returns 0. */
dGeomID dCreateCylinder(dSpaceID space, dReal radius, dReal length); /* collision.h:1062 */
void dGeomCylinderSetParams(dGeomID cylinder, dReal radius, dReal length); /* collision.h:1063 */
void dGeomCylinderGetParams(dGeomID cylinder, dReal *radius, dReal *length); /* collision.h:1064 */

dSpaceID dSimpleSpaceCreate(dSpaceID space);                 /* collision_space.h:52 */
dSpaceID dHashSpaceCreate(dSpaceID space);                   /* collision_space.h:53 */
dSpaceID dSweepAndPruneSpaceCreate(dSpaceID space, int axisorder); /* collision_space.h:66 */
void dSpaceDestroy(dSpaceID space);                          /* collision_space.h:70 */
void dHashSpaceSetLevels(dSpaceID space, int minlevel, int maxlevel); /* collision_space.h:72 */
void dSpaceSetCleanup(dSpaceID space, int mode);             /* collision_space.h:75 */
void dSpaceAdd(dSpaceID space, dGeomID g);                   /* collision_space.h:149 */
void dSpaceRemove(dSpaceID space, dGeomID g);                /* collision_space.h:150 */
int dSpaceGetNumGeoms(dSpaceID space);                       /* collision_space.h:153 */

/* ----------------------------------------------------------------- misc */
unsigned long dRand(void);                                   /* misc.h:42 */
unsigned long dRandGetSeed(void);                            /* misc.h:45 */
void dRandSetSeed(unsigned long s);                          /* misc.h:46 */
int dRandInt(int n);                                         /* misc.h:51 */

/* ------------------------------------------- libodx controls (no ODE twin) */
/* Row order of dWorldQuickStep: 0 = the reference's order (findex rows last,
 * joints in island order, no random reordering; default), 1 = coloured contact
 * blocks, 2 = the reference's default build (REORDERING_METHOD__RANDOMLY,
 * quickstep.cpp:54-85): the order of 0, shuffled per island with dRandInt
 * before iterations 8, 16, ... (quickstep.cpp:2589-2611), drawing from the
 * dRand state of dRandSetSeed island by island.  Applies to subsequent steps
 * of the world. */
void odxWorldSetRowOrder(dWorldID w, int order);
/* How the step kernel sweeps the dependency levels of rows (no counterpart in
 * the reference, which sweeps on one CPU thread): 0 = automatic (island
 * groups when the reference row order runs a fixed sweep count over at least
 * 32 islands with rows; else one wave when levels hold at most 64 items on
 * average; default), 1 = the whole workgroup per level, 2 = one wave, 3 =
 * island groups: runs of consecutive islands on workgroups of their own (one
 * wave each, on separate CUs), or among the waves of one workgroup when the
 * runs do not fit, each sweeping its own islands' levels (reference row order
 * with a fixed sweep count; otherwise as 0).  The Gauss-Seidel order and the
 * results are the same in every mode. */
void odxWorldSetSweepMode(dWorldID w, int mode);

/* Writes the world in the Dynamics Interchange Format (export-dif.h:33,
 * ode/src/export-dif.cpp:491-620): parameters, bodies in list order with the
 * geoms attached to each, joints newest first (ball and contact here), reals
 * with 7 significant digits as a dSINGLE build prints them. */
void dWorldExportDIF(dWorldID w, FILE *file, const char *world_name);
/* Sweeps / LCP residual of the last dWorldQuickStep (residual < 0 when no rows). */
void odxWorldGetSolveStats(dWorldID w, int *sweeps, float *residual, int *rows);

#ifdef __cplusplus
}
#endif
#endif
