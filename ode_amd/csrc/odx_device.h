// odx_device.h -- shared device/host layouts of the batched step path.
#pragma once
#include <stdint.h>

namespace odx {

// Record strides (floats) shared with include/odx.h and the test harness.
constexpr int BODY_STRIDE = 32;
constexpr int GEOM_STRIDE = 16;
constexpr int JOINT_STRIDE = 16;
constexpr int ROW_STRIDE = 32;    // one assembled constraint row, 128 B
constexpr int BSCR_STRIDE = 32;   // per-body scratch: R(12) invI_w(12) rhs_tmp(6) invMass
constexpr int FEAT_DIM = 16;      // warm-start feature vector per contact

// Per-world limits of the one-wave-per-world kernels.
constexpr int SW_MAXB = 64;       // bodies
constexpr int SW_MAXG = 96;       // geoms
constexpr int SW_MAXPAIRS = 1024; // candidate pairs after broadphase
constexpr int SW_MAXROWS = 2048;  // constraint rows
// Schedule keys: colours are < 64 + joints <= 64 + SW_MAXROWS (a joint whose
// bodies have used all 64 mask colours gets a phase of its own).
constexpr int SW_MAXPH = SW_MAXROWS + 65;

// Limit of the single-world (dWorldQuickStep) kernel: bodies per island set.
constexpr int SL_MAXB = 2048;

// Contacts a device collider returns per geom pair at most.
#ifndef ODX_MAXC
#define ODX_MAXC 4
#endif

// Row record layout (floats): J1[0..5] rhs[6] cfm[7] J2[8..13] lo[14] hi[15]
// iMJ1[16..21] iMJ2[22..27] b1[28] b2[29] findex[30] (int bits) Ad[31].
enum { R_J1 = 0, R_RHS = 6, R_CFM = 7, R_J2 = 8, R_LO = 14, R_HI = 15, R_IMJ1 = 16, R_IMJ2 = 22,
       R_B1 = 28, R_B2 = 29, R_FINDEX = 30, R_AD = 31 };

// JOINT record (floats): 0 type (0 contact, 1 ball), 1 b1, 2 b2 (world-local,
// -1 none), contact: 3-5 pos 6-8 normal 9 depth 10-12 fdir1 13 surface 14 g1 15 g2;
// ball: 3-5 anchor1 6-8 anchor2 9 erp 10 cfm.

struct Surface {
    int mode;
    float mu, mu2, bounce, bounce_vel, soft_erp, soft_cfm, motion1, motion2, motionN, slip1, slip2;
    int max_contacts;  // batched collide: contacts per geom pair of this surface (0: StepParams::maxc)
};

// Material-pair surface tables of the batched path hold at most this many materials.
constexpr int ODX_MAX_MATERIALS = 16;

struct StepParams {
    float h, gx, gy, gz, erp, cfm, sor_w;
    int max_iters;       // sweep cap (the fixed count when tol <= 0)
    float tol;           // LCP residual target; <= 0 runs exactly max_iters sweeps
    int check_every;     // residual check period in sweeps
    float max_vel, min_depth;
    int maxc;            // contacts per geom pair
    int space;           // 0 SAP semantics, 1 hash semantics
    int skip_connected;  // drop contacts between bodies joined by a ball joint
    int order;           // 0 reference row order (dependency levels), 1 coloured contact blocks
    int warm;            // 0 cold start, 1 host lambda0 per row, 2 matched previous-step impulses,
                         // 3 learned (the MLP fills lambda0 from the contact features)
    float lambda_floor;
    int final_residual;  // 1: compute the residual after the last sweep
    int collide;         // 1: generate contacts; 0: joints were uploaded
    int features;        // 1: write warm-start features per contact
    int feedback;        // 1: write joint forces J^T lambda per joint (dJointFeedback)
    int restart_after;   // > 0: a warm-started world not at tol after this many sweeps restarts from lambda = 0
    int lds_rows;        // batched solve: worlds with at most this many rows sweep from LDS (0: never)
    int lds_bodies;      // LDS body records per world (max bodies per world + 1 zero body)
    int wbase, wcount;   // batched launch covers worlds [wbase, wbase + wcount)
    int split;           // feature counter pair of this launch: feat_count[2 * split], [2 * split + 1]
    int feat_base;       // first feature slot of this launch (joint_start[wbase])
    int feat_local;      // 1: features go to slots joint_start[w] + i, in joint order (fused kernel)
    int gate;            // warm 3 with a gated network: matched contacts take their previous impulses
                         // directly and only unmatched contacts get feature slots
    // dynamic iteration count per island (drop-in world, quickstep.cpp:1820-1853)
    int dyn;             // 1: on (then max_iters is the base count)
    int dyn_max_extra;   // extra sweeps allowed after max_iters
    float dyn_premature, dyn_extra_delta;
    int nislands;        // islands of the drop-in world (body_island values)
    // near-callback policy of the batched collide: with nmat > 0 a contact of
    // geoms g1 < g2 takes surface material(g1) * nmat + material(g2); with
    // cat_filter a pair needs (cat1 & col2) | (cat2 & col1) != 0
    int nmat, cat_filter;
    int wave_sweep;      // drop-in world: 0 auto, 1 block-wide levels, 2 one wave (the kernel sees 0 or 1)
    int wave_groups;     // drop-in world: > 0 = islands split into this many wave groups (phase table header)
};

// Independent world ranges a batch may be split into (one stream each).
constexpr int ODX_MAX_SPLITS = 8;

// Largest dynamic LDS of the batched solve kernel (bytes per workgroup).
constexpr int SOLVE_LDS_MAX = 65536;
// waves of the staged single-world step kernel (512 threads): the island
// groups of sweep mode 3 when they share one workgroup
constexpr int SL_STAGED_WAVES = 8;
// island groups of sweep mode 3 on one workgroup each: at most this many
// groups, and the LDS one may take (fc of its bodies, lambda of its rows, its
// schedule)
constexpr int SL_GROUP_BLOCKS = 128;
constexpr int SL_GROUP_LDS = 65536;
// ... and with the run's rows in LDS as well (capped by the device's limit)
constexpr int SL_GROUP_LDS_ROWS = 160 * 1024;
// LDS budgets of the batched solve besides the rows: work items per world
// (three quarters of the row budget, plus 16) and sweep phases per world.  A
// world over any budget sweeps from global memory instead.
__host__ __device__ inline int lds_items_of(int lds_rows) { return lds_rows * 3 / 4 + 16; }
constexpr int LDS_PHASES = 128;
// dynamic LDS of the batched solve, in floats: bodies [lds_bodies][16], rows
// [lds_rows][16], lambda [lds_rows], items [lds_items_of(lds_rows)], phase
// table [LDS_PHASES + 1] (16-bit, padded to whole floats), 64 trash words
__host__ __device__ inline int solve_lds_floats(int lds_bodies, int lds_rows) {
    return lds_bodies * 16 + lds_rows * 17 + lds_items_of(lds_rows) + (LDS_PHASES + 2) / 2 + 64;
}

// Split of a joint's rows for the coloured schedule, which keeps the
// reference's two passes (ReorderPrep, quickstep.cpp:2329-2355: rows without
// findex first, findex rows after): rfl[r] = 1 for a findex row.  A joint
// whose rows are n_a non-findex rows followed by n_f findex rows gives
// (n_a, n_f); any other pattern stays one item in the first pass (n_f = 0).
template <class Flag>
__host__ __device__ inline void split_joint_rows(const Flag *rfl, int mj, int &n_a, int &n_f) {
    int a = 0;
    while (a < mj && rfl[a] == 0) a++;
    int f = a;
    while (f < mj && rfl[f] == 1) f++;
    if (a > 0 && f == mj) { n_a = a; n_f = mj - a; }
    else { n_a = mj; n_f = 0; }
}

struct DevWorlds {
    int nworlds;
    const int *body_start, *geom_start, *ball_start, *joint_start, *row_start;  // [W+1]
    float4 *pos, *quat, *lvel, *avel, *facc, *tacc;   // [NB]
    const float4 *mass;                                // [NB] mass, invMass, flags, 0
    const float *invIb;                                // [NB*12] body-frame inverse inertia
    const float *Ib;                                   // [NB*12] body-frame inertia
    // [NB*2] per-body damping and angular speed cap (util.cpp:583-594, 675-698),
    // null when no body uses them: (linear scale, angular scale, linear
    // threshold^2, angular threshold^2), (max angular speed, DAMP_* bits,
    // linear threshold, angular threshold)
    const float4 *damp;
    const float4 *geom0, *geom1, *gpos, *gquat;        // [NG] cls/body/p0/p1, p2/p3/material/category bits,
                                                       // static pos + collide bits, static q
    const float *balls;                                // [NBALL*16]
    const Surface *surf;
    float *joints;          // [joint cap * 16]
    int *njoints, *ncontacts, *nrows, *nphases, *niters;  // [W]
    float *residual;        // [W]
    float *rows;            // [row cap * 32]
    float *lambda;          // [row cap]
    float *lambda0;         // [row cap] warm-start values (host- or device-written)
    int *sched;             // [row cap] work items
    int *phase_start;       // [row cap + W] per-world phase offsets (row_start[w] + w based)
    float *bscr;            // [NB * 32]
    int *joff;              // [joint cap] world-local first row of each joint
    float *feat;            // [joint cap * 16] compacted warm-start features (learned mode)
    int *feat_row;          // [joint cap] global first row of each feature entry
    int *feat_count;        // [2 * (ODX_MAX_SPLITS + 1)] entries written per launch range: this step, last step
    int *nfeat;             // [W] entries of each world (world-local slots, fused kernel)
    float *prev;            // [joint cap * 16] previous-step contacts: g1 g2 plocal(3) n(3) lambda(3)
    int *nprev;             // [W]
    int *overflow;          // [4] pairs / contacts / rows / phases overflow counters
    float *fb;              // [joint cap * 12] f1 t1 f2 t2 per joint (feedback mode)
    int *perm, *inv;        // [row cap] sweep position of each row and its inverse (per world, row_start based)
    int *order;             // [W] solve dispatch order (longest previous solve first); null = index order
    long long *wtime;       // [W][2] solve start / end of each world (device wall clock), may be null
    long long *ptime;       // [W][PTIME_SLOTS] prepare stage boundaries (device wall clock), may be null
    const int *body_island; // [NB] island of each body (drop-in world with the dynamic iteration count), may be null
    // drop-in world with random reordering: sweeps 8s .. 8s+7 use segment s of
    // the schedule (seg_table[3s] = first item, [3s+1] = first phase start,
    // [3s+2] = phases; the last segment serves any later sweep)
    // drop-in world: one extended record per joint for hinge, slider and fixed
    // joints (JEXT_* layout), may be null
    float *jext;            // extended records of hinge / slider / fixed joints: drop-in world [joint
                            // slot * JEXT_STRIDE]; batched worlds [ball-joint index * JEXT_STRIDE]
    const int *seg_table;   // [3 * nseg], may be null
    int nseg;
    int *island_sweeps;     // [nislands] sweeps each island ran, may be null
    // drop-in world, island groups on separate workgroups: the first body and
    // joint of each group's run of islands ([G + 1] each), may be null
    const int *grp_body_start, *grp_joint_start;
};
constexpr int PTIME_SLOTS = 6;

// JOINT record types: 0 contact, 1 ball, 2 hinge, 3 slider, 4 fixed.  Hinge,
// slider and fixed joints carry an extended record of JEXT_STRIDE floats:
// axis1 (body-1 frame), axis2 (body-2 frame or world), the initial relative
// rotation qrel, the offset (slider / fixed), whether the limit / motor row is
// present, the limit state and error from getInfo1, the limit / motor
// parameters (dxJointLimitMotor, joints/joint.h) and the fixed joint's own
// erp / cfm.
constexpr int JEXT_STRIDE = 32;
enum { JT_HINGE = 2, JT_SLIDER = 3, JT_FIXED = 4 };
enum { JX_AXIS1 = 0, JX_AXIS2 = 3, JX_QREL = 6, JX_OFFSET = 10, JX_LIMOT_ROW = 13, JX_LIMIT = 14, JX_LIMIT_ERR = 15,
       JX_LOSTOP = 16, JX_HISTOP = 17, JX_VEL = 18, JX_FMAX = 19, JX_FUDGE = 20, JX_NORMAL_CFM = 21,
       JX_STOP_ERP = 22, JX_STOP_CFM = 23, JX_BOUNCE = 24, JX_REVERSE = 25, JX_ERP = 26, JX_CFM = 27 };

// DAMP_* bits of a body's damping record (BODY record flags bits 2-4 >> 2)
enum { DAMP_LINEAR = 1, DAMP_ANGULAR = 2, DAMP_MAX_ANGULAR_SPEED = 4 };

// Geom view for the collision entry points of the single-world API:
// cls, body identity (-1 none), params[4], pos[3], R[12] (3x4).
constexpr int GVIEW_STRIDE = 24;

// One dSpaceCollide on the device (k_space_collide, odx_space.hip).
struct SpaceCall {
    const float *gviews;             // [n][GVIEW_STRIDE] in the space's order (SAP: the sorted set first)
    const unsigned long long *bits;  // [n][2] category, collide
    int n;
    int nsort;                       // SAP: finite geoms (views 0..nsort-1); hash, simple: n
    int kind;                        // 0 sweep and prune, 1 hash, 2 simple
    int ax0, ax1, ax2;               // SAP axes as AABB indices (min = idx, max = idx + 1)
    int minlevel, maxlevel;          // hash levels
    const unsigned *ranks_in;        // SAP: ranks kept from the previous sort
    int ranks_valid;
    unsigned *ranks_out;
    float *aabb;                     // [n][6]
    int *cells;                      // hash: [n][8] level (INT_MAX: big box), discretised bounds [6]
    int *pairs;                      // [cap][2] view indices (o1, o2)
    int cap;
    int *info;                       // [0] pairs; [1] SAP: ranks valid after the call
};

// Previous-contact record (floats) kept by the solve kernel for matching.
enum { PV_G1 = 0, PV_G2 = 1, PV_LOCAL = 2, PV_N = 5, PV_LAM = 8 };
constexpr float MATCH_DIST = 0.05f;   // max body-local distance (m) of a matched contact point
constexpr float GRAVITY_SCALE = 9.81f;  // force scale of the learned warm start: s = g / (im1 + im2)

// Feature vector of a contact (learned warm start), raw units:
// 0-2 normal, 3 depth, 4 v_n, 5 v_t1, 6 v_t2 (relative velocity of the contact
// point), 7-9 matched previous impulses / s, 10 im1, 11 im2, 12 matched flag,
// 13 contact height over body 1, 14 two-body flag, 15 s (force scale).

}  // namespace odx
