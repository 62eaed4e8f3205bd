/*
 * odx.h -- C ABI of libodx: the ODE step hot path on AMD Instinct MI355X.
 *
 * Two groups of entry points:
 *
 *  1. The ODE-compatible subset an existing ODE program calls on its step path
 *     (see ode.h in this directory): dWorldQuickStep, dSpaceCollide with its
 *     near-callback, dCollide, dJointCreateContact and the world/body/geom/
 *     joint-group calls they need.  Same names, argument meaning and return
 *     codes as the reference's include/ode/ *.h; dReal is float.
 *
 *  2. A batched-worlds extension (odxBatch*): thousands of independent worlds
 *     stepped by one launch sequence.  No ODE counterpart; it replaces the
 *     caller's frame loop (dSpaceCollide + near-callback + dWorldQuickStep +
 *     dJointGroupEmpty, e.g. ode/demo/demo_space_stress.cpp:363-369) for many
 *     worlds at once, with the near-callback policy given as a surface table.
 *
 * Plain pointers and sizes only; no exceptions cross this boundary.  Functions
 * returning int return 1 (or a count) on success and 0 on failure, with a
 * message through the handler installed by odxSetMessageHandler.
 *
 * Record layouts (float arrays):
 *   BODY  [32]: 0-2 pos, 4-7 q (w,x,y,z), 8-10 lvel, 12-14 avel, 16-18 facc,
 *               20-22 tacc, 24 mass, 25-30 I (xx,yy,zz,xy,xz,yz),
 *               31 flags (bit0 no gravity, bit1 gyroscopic, bit2 linear
 *               damping, bit3 angular damping, bit4 angular speed cap);
 *               3 linear damping scale, 11 angular damping scale, 15 linear
 *               damping threshold, 19 angular damping threshold, 23 maximum
 *               angular speed (dBodySetDamping*, dBodySetMaxAngularSpeed,
 *               include/ode/objects.h:1500-1588; applied by dxStepBody,
 *               ode/src/util.cpp:583-699)
 *   GEOM  [16]: 0 class (0 sphere, 1 box, 2 capsule, 4 plane), 1 body (world-
 *               local index, -1 static), 2-5 params (sphere r | box lx,ly,lz |
 *               capsule r,len | plane a,b,c,d), 6-8 static pos, 9-12 static q,
 *               13 material, 14 category bits, 15 collide bits (bits 0-23;
 *               read by odxBatchSetSurfaceTable's policy)
 *   JOINT [16]: 0 type (0 contact, 1 ball), 1 body1, 2 body2 (-1 none);
 *               contact: 3-5 pos 6-8 normal 9 depth 10-12 fdir1 13 surface 14 g1 15 g2
 *               ball:    3-5 anchor1 (body1 frame) 6-8 anchor2 (body2 frame or world) 9 erp 10 cfm
 *   PARAMS double[32]: 0 h, 1-3 gravity, 4 erp, 5 cfm, 6 sor w, 7 iterations,
 *               8 contact max correcting vel, 9 contact surface layer,
 *               10 surface mode, 11 mu, 12 mu2, 13 bounce, 14 bounce_vel,
 *               15 soft_erp, 16 soft_cfm, 17 max contacts per pair, 18 space
 *               (0 sweep-and-prune, 1 hash), 19 motion1, 20 motion2, 21 motionN,
 *               22 slip1, 23 slip2, 24 skip pairs joined by a ball joint,
 *               25 residual floor, 31 row order of odxLarge's sweeps (0
 *               ReorderPrep's, 2 each contact's rows together)
 */
#ifndef ODX_H
#define ODX_H

#ifdef __cplusplus
extern "C" {
#endif

#define ODX_BODY_STRIDE 32
#define ODX_GEOM_STRIDE 16
#define ODX_JOINT_STRIDE 16
#define ODX_FEAT_DIM 16

typedef void odxMessageFunction(int errnum, const char *msg);

/* Library version string and a message sink (NULL restores stderr). */
const char *odxVersion(void);
void odxSetMessageHandler(odxMessageFunction *fn);
/* Number of visible GPUs (0 when none). */
int odxDeviceCount(void);

/* ------------------------------------------------------------ batched API */
typedef struct odxBatch odxBatch;

/* Creates nworlds worlds on the current device.  body_start/geom_start/
 * ball_start are prefix offsets [nworlds+1] into the record arrays; bodies are
 * world-local in GEOM and JOINT records.  Worlds may hold at most 64 bodies and
 * 96 geoms.  contacts_per_body sizes each world's contact capacity
 * (capacity = contacts_per_body * bodies + 16).  Returns NULL on failure. */
odxBatch *odxBatchCreate(int nworlds, const int *body_start, const float *bodies, const int *geom_start,
                         const float *geoms, const int *ball_start, const float *balls, const double *params,
                         int contacts_per_body);
/* odxBatchCreate with hinge, slider and fixed joints among the joints
 * (dJointCreateHinge / Slider / Fixed, ode/src/joints/hinge.cpp, slider.cpp,
 * fixed.cpp; rows of getInfo2 and the limit / motor of joint.cpp
 * dxJointLimitMotor).  joints are JOINT records of type 1 (ball), 2 (hinge),
 * 3 (slider) or 4 (fixed): 3-5 anchor1 (body1 frame), 6-8 anchor2 (body2
 * frame, or world when body2 is -1), 9 erp and 10 cfm of the hinge's ball rows
 * (the world's).  jext holds one extended record of 32 floats per joint (those
 * of ball joints are ignored): 0-2 axis1 (body1 frame), 3-5 axis2 (body2 frame,
 * or world), 6-9 qrel (dJointSetFixed / setFixedOrientation), 10-12 offset
 * (slider: body1's position from body2 in body2's frame; fixed: body1's
 * position from body2 in body1's frame; either: body1's world position when
 * body2 is -1), 14 limit state
 * (0 to start), 15 limit error, 16 lostop, 17 histop, 18 vel, 19 fmax, 20 fudge,
 * 21 normal cfm, 22 stop erp, 23 stop cfm, 24 bounce, 25 slider reverse, 26 erp
 * and 27 cfm of the fixed joint.  The limit state, the limit / motor row and
 * the force of a motor at its limit (getInfo1, addLimot) are evaluated on the
 * device every step.  jext may be NULL when every joint is a ball joint. */
odxBatch *odxBatchCreateJoints(int nworlds, const int *body_start, const float *bodies, const int *geom_start,
                               const float *geoms, const int *joint_start, const float *joints, const float *jext,
                               const double *params, int contacts_per_body);
void odxBatchDestroy(odxBatch *b);

/* Solver controls.  order: 0 = reference row order (deterministic; results
 * match a sequential QuickStep with row reordering off), 1 = coloured contact
 * blocks.  tol <= 0 runs exactly max_iters sweeps; otherwise sweeps stop per
 * world once the LCP residual is <= tol (checked every check_every sweeps),
 * capped at max_iters.  warm: 0 cold start, 1 lambda0 from odxBatchSetLambda0. */
int odxBatchSetSolver(odxBatch *b, int order, int max_iters, float tol, int check_every, int warm);

/* restart_after > 0: a warm-started world that has not reached tol after that
 * many sweeps restarts from lambda = 0 (0 = never).  lds_rows: worlds with at
 * most this many rows sweep with rows and bodies held in LDS (64 B per row);
 * larger worlds stream rows from global memory; 0 disables the LDS path.
 * Fails when lds_rows does not fit the kernel's 64 KB LDS budget. */
int odxBatchSetSolveOptions(odxBatch *b, int restart_after, int lds_rows);

/* The near-callback policy on the device, in place of a per-pair callback
 * that sets dContact::surface before dJointCreateContact (as in
 * ode/demo/demo_boxstack.cpp:131-160).  table: float[nmat][nmat][16], rows
 * 0 mode, 1 mu, 2 mu2, 3 bounce, 4 bounce_vel, 5 soft_erp, 6 soft_cfm,
 * 7 motion1, 8 motion2, 9 motionN, 10 slip1, 11 slip2 (dSurfaceParameters,
 * include/ode/contact.h:55-71), 12 contacts per pair (1..4; 0 = the batch's
 * count).  A contact of geoms g1 < g2 takes entry (material(g1),
 * material(g2)); the table must be symmetric and cover every geom's
 * material (GEOM slot 13).  category_filter != 0 drops pairs whose category
 * and collide bits (GEOM slots 14, 15) fail the test of dSpaceCollide's
 * collideAABBs, ode/src/collision_space_internal.h:44-79: (cat1 & col2) ||
 * (cat2 & col1).  nmat = 0 returns to the one surface of the parameters. */
int odxBatchSetSurfaceTable(odxBatch *b, int nmat, const float *table, int category_filter);

/* World snapshot I/O (no counterpart in the reference, whose nearest is the
 * text export dWorldExportDIF; for fixtures and reproducible benchmarks).
 * odxBatchSaveSnapshot writes the batch to a binary file: the bodies as they
 * are now, the geoms, balls and parameters it was created with, its surface
 * table and its solver settings.  odxBatchLoadSnapshot creates a batch from
 * such a file.  A cold-start step of the loaded batch equals the original's
 * next step up to the quaternion renormalisation of the upload; warm-start
 * history (previous contacts) is not saved, so warm modes start afresh.
 * Both return 0 / NULL with a message on failure. */
int odxBatchSaveSnapshot(odxBatch *b, const char *path);
odxBatch *odxBatchLoadSnapshot(const char *path);
/* Worlds, bodies and the world offsets (nworlds + 1 ints; may be NULL). */
int odxBatchGetLayout(odxBatch *b, int *nworlds, int *nbodies, int *body_start);

/* Splits the worlds into nsplit (1..8) equal index ranges that step on streams
 * of their own inside one hipGraph of graph_steps steps: the worlds are
 * independent, so while one range finishes its slowest solve the others keep
 * the GPU busy with their next steps.  Results are those of nsplit = 1. */
int odxBatchSetSplits(odxBatch *b, int nsplit, int graph_steps);

/* steps_per_launch > 0: every world goes through that many complete steps
 * (collide, rows, network, solve, integrate) inside one launch of a
 * persistent fused kernel, one wave per world, so a world's slow solve delays
 * only that world.  0 (default): one launch per stage per step.  Results are
 * those of the staged launches. */
int odxBatchSetFused(odxBatch *b, int steps_per_launch);

/* Autotuner of the launch schedule (no ODE counterpart).  Times every
 * candidate -- nsplit 1, 2, 4, 8 (odxBatchSetSplits, graph_steps as set) by
 * staged or fused (odxBatchSetFused(b, graph_steps)) launches -- over `steps`
 * steps of this batch, `reps` times each, and keeps the fastest; the schedule
 * in effect at the call is kept unless a candidate beats it by more than 5 %.
 * Every run starts from the state at the call (bodies, previous-step contacts
 * and impulses, sweep counts), and that state is put back at the end.
 * schedule (may be NULL): [nsplit, graph_steps, steps per fused launch or 0];
 * ms_per_step (may be NULL): the winner's mean time per step. */
int odxBatchAutotune(odxBatch *b, int steps, int reps, int *schedule, float *ms_per_step);

/* Runs nsteps steps on the batch's stream (captured once into a hipGraph of
 * graph_steps steps and replayed).  Returns 1 on success. */
int odxBatchStep(odxBatch *b, int nsteps);
/* Blocks until queued work is done. */
int odxBatchSync(odxBatch *b);

/* Times nsteps graph-replayed steps with hipEvents on the batch stream
 * (total milliseconds). */
int odxBatchTime(odxBatch *b, int nsteps, float *total_ms);
/* Runs nsteps directly launched steps with hipEvents around each kernel and
 * returns the average durations [prepare, warm-start inference, solve] in
 * milliseconds (kernel_ms[3]; inference is 0 when not used). */
int odxBatchKernelTimes(odxBatch *b, int nsteps, float *kernel_ms);
/* Runs nsteps steps launched on the batch's current schedule (its world
 * ranges on their own streams; no counterpart in the reference) with
 * hipEvents on each range's stream around each kernel.  Writes, per range r,
 * kernel_ms[3r..3r+2] = mean [prepare, inference, solve] milliseconds per
 * launch (a fused schedule reports its one kernel as solve) and
 * ranges[2r..2r+1] = its first and one-past-last world; returns the number
 * of ranges (at most 8), 0 on failure. */
int odxBatchScheduleKernelTimes(odxBatch *b, int nsteps, float *kernel_ms, int *ranges);

/* Readback.  bodies: [NB][32].  stats per world: contacts, rows, sweeps,
 * residual (NULL pointers are skipped).  Returns the total count written. */
int odxBatchGetBodies(odxBatch *b, float *bodies);
int odxBatchSetBodies(odxBatch *b, const float *bodies);
int odxBatchGetStats(odxBatch *b, int *ncontacts, int *nrows, int *niters, float *residual);
/* Joint records of world w after the last step (balls first, then contacts). */
int odxBatchGetJoints(odxBatch *b, int w, float *joints, int cap);
/* Lambda of world w in natural row order. */
int odxBatchGetLambda(odxBatch *b, int w, float *lambda, int cap);
/* Assembled rows of world w ([m][32]: J1 rhs cfm J2 lo hi iMJ1 iMJ2 b1 b2 findex Ad). */
int odxBatchGetRows(odxBatch *b, int w, float *rows, int cap);
/* Overflow counters: [pairs, contacts, rows, phases] dropped since creation. */
int odxBatchGetOverflow(odxBatch *b, int *counts);
/* Start and end (microseconds from the earliest start) of each world's solve
 * in the last step, [nworlds][2], from the device wall clock. */
int odxBatchGetSolveTimes(odxBatch *b, double *start_end_us);
/* Stage boundaries of each world's prepare stage in the last step,
 * [nworlds][6] microseconds from the earliest stamp: start, after bodies,
 * after broadphase, after narrowphase, after rows, after the schedule. */
int odxBatchGetPrepareTimes(odxBatch *b, double *stamps_us);
/* Row-capacity offsets [nworlds+1] (the lambda0 layout). */
int odxBatchRowStarts(odxBatch *b, int *row_start);
/* Warm-start values per row in the row-capacity layout (host array). */
int odxBatchSetLambda0(odxBatch *b, const float *lambda0, int n);

/* Kernel time of the device step inside dWorldQuickStep (include/ode.h):
 * odxApiSetKernelTiming(1) clears the totals and brackets every following
 * step's kernel with hipEvents; odxApiGetKernelTiming fills stats[3] = steps,
 * total kernel milliseconds, total row-sweeps (rows x sweeps).  No ODE
 * counterpart (measurement only). */
void odxApiSetKernelTiming(int on);
int odxApiGetKernelTiming(double *stats);

/* ------------------------------------------------------- one large world */
/* One world of up to millions of bodies stepped across the whole GPU:
 * dSpaceCollide over a dHashSpace (ode/src/collision_space.cpp:421-614; the
 * pair set is every pair of geoms on different bodies whose closed AABBs
 * overlap, planes tested against all), dCollide per pair
 * (collision_kernel.cpp:292-329), one contact joint per contact
 * (dJointCreateContact ode.cpp:1192-1200 with dJointAttach's reversal), and
 * dWorldQuickStep (quickstep.cpp) with the rows swept colour by colour: a
 * colour is a set of contacts with disjoint bodies, so a phase runs in
 * parallel and a sweep is a Gauss-Seidel pass in colour order.  params[31]
 * picks the rows' order within a sweep: 0, as ReorderPrep
 * (quickstep.cpp:2329-2355), every contact's normal row colour by colour,
 * then its friction rows colour by colour (two passes over the colours); 2,
 * each contact's rows together in joint order (one pass: half the colour
 * phases, a different Gauss-Seidel order of the same LCP).  Random reordering
 * (1) is not supported.  Contacts only (no ball joints).  bodies [nb][32], geoms [ng][16], params as
 * odxBatchCreate (params[7] = sweeps per step).  pair_cap / contact_cap bound
 * the candidate pairs and contacts of one step (<= 0: 12 and 8 per geom).
 * Returns NULL on failure. */
typedef struct odxLarge odxLarge;
odxLarge *odxLargeCreate(int nb, const float *bodies, int ng, const float *geoms, const double *params, int pair_cap,
                         int contact_cap);
void odxLargeDestroy(odxLarge *w);
/* Sweeps per step (dWorldSetQuickStepNumIterations, objects.h:462). */
int odxLargeSetIterations(odxLarge *w, int iters);
/* Workgroups per CU of the cooperative solve (1..8, capped by occupancy;
 * default 1).  Returns the resulting grid size, 0 on failure.  No ODE
 * counterpart (launch shape only; results do not depend on it). */
int odxLargeSetSolveBlocks(odxLarge *w, int per_cu);
/* Polls one grid-barrier wait may make before the step fails (at most and by
 * default 2^21, about a second).  No ODE counterpart: it exists so that a test
 * can force the barrier-timeout path from the host (0: every wait fails at
 * once); every workgroup still leaves the kernel normally, and the next
 * odxLargeSync reports the failed step. */
int odxLargeSetSpinLimit(odxLarge *w, int polls);
/* The near-callback policy of odxBatchSetSurfaceTable for the large world:
 * table [nmat][nmat][16] as there, geoms' materials in GEOM slot 13; with
 * category_filter the hash-space pairs whose category / collide bits fail
 * collideAABBs' test (collision_space_internal.h:50-57) are not formed.
 * Waits for queued steps.  nmat = 0 returns to the parameters' surface. */
int odxLargeSetSurfaceTable(odxLarge *w, int nmat, const float *table, int category_filter);
/* Queues nsteps steps (dSpaceCollide + contacts + dWorldQuickStep +
 * dJointGroupEmpty each) on the world's stream.  Returns 1 when queued. */
int odxLargeStep(odxLarge *w, int nsteps);
/* Waits for the queued steps.  Returns 0 (and reports why) when a step since
 * the last sync dropped pairs or contacts past the capacities, left contacts
 * uncoloured or failed a grid barrier. */
int odxLargeSync(odxLarge *w);
/* Times nsteps steps with hipEvents (total milliseconds). */
int odxLargeTime(odxLarge *w, int nsteps, float *total_ms);
/* Average [prepare (bodies, broadphase, narrowphase, rows), colouring, solve]
 * milliseconds over nsteps steps (ms[3]). */
int odxLargeKernelTimes(odxLarge *w, int nsteps, float *ms);
int odxLargeGetBodies(odxLarge *w, float *bodies);
int odxLargeSetBodies(odxLarge *w, const float *bodies);
/* Counters of the last step [16]: pairs, contacts, rows, colours, dropped
 * pairs, dropped contacts, uncoloured contacts, barrier failure, colouring
 * rounds. */
int odxLargeGetCounts(odxLarge *w, int *counts);
/* Last step's candidate pairs [n][2] (lower geom first), contact JOINT
 * records [n][16] (14, 15 = geoms), rows [m][32] (contact k's rows are
 * consecutive, in contact order), lambda [m] and contact colours [n]. */
int odxLargeGetPairs(odxLarge *w, int *pairs, int cap);
int odxLargeGetJoints(odxLarge *w, float *joints, int cap);
int odxLargeGetRows(odxLarge *w, float *rows, int cap);
int odxLargeGetLambda(odxLarge *w, float *lambda, int cap);
int odxLargeGetColours(odxLarge *w, int *colour, int cap);

/* ------------------------------------------------ learned warm start (MLP) */
/* A 16 -> 128 -> 128 -> 3 GELU network on bf16 MFMA with fp32 accumulation
 * that predicts a contact's initial (normal, friction1, friction2) impulses
 * from its features (layout in ode_amd/csrc/odx_device.h).  No ODE
 * counterpart: the reference's warm start is compiled out
 * (ode/src/quickstep.cpp:46-51).  Parameters are a flat float vector
 * W1[128][16] b1[128] W2[128][128] b2[128] W3[3][128] b3[3] (19,075 floats). */
typedef struct odxMlp odxMlp;
odxMlp *odxMlpCreate(unsigned seed);
void odxMlpDestroy(odxMlp *m);
/* norm38: input mean[16], input scale[16], output mean[3], output scale[3]. */
int odxMlpSetNormalization(odxMlp *m, const float *norm38);
/* gate 1: contacts matched to a previous-step contact keep its impulses and
 * the network predicts only new contacts; gate 0 (default): the network's
 * predicted change is added for every contact.  Takes effect for batches
 * attached (odxBatchSetMlp) afterwards: such a batch writes features (and
 * runs the network) for its unmatched contacts only, so odxBatchGetSamples
 * returns those. */
int odxMlpSetGate(odxMlp *m, int gate);
int odxMlpGetParams(odxMlp *m, float *params);
int odxMlpSetParams(odxMlp *m, const float *params);
/* Training set: features X[N][16] (raw) and targets T[N][3] (impulse / force scale). */
int odxMlpSetData(odxMlp *m, const float *X, const float *T, int N);
/* Loss gradient over the whole uploaded set (no update); grads[19075]. */
int odxMlpGradients(odxMlp *m, float *grads, float *loss);
/* nsteps Adam steps on random mini-batches; per-step losses (may be NULL). */
int odxMlpTrain(odxMlp *m, int nsteps, int batch, float lr, unsigned seed, float *losses);
/* Predicted impulses (force units) for features X[N][16] -> Y[N][3]. */
int odxMlpPredict(odxMlp *m, const float *X, int N, float *Y);
/* odxMlpPredict without any allocation (no counterpart in the reference):
 * for callers that cannot allocate device memory, such as tight serving
 * loops and harnesses that forbid hipMalloc.  It stages the features and
 * impulses in a fixed array of 1.25 MiB on each device and accepts at most
 * 16384 contacts (more is an error, reported through the message handler).
 * Calls that overlap on one device share that array and race. */
int odxMlpPredictNoAlloc(odxMlp *m, const float *X, int N, float *Y);
/* MFMA operand-layout probe: per-lane inputs a,b [64][8] -> d [64][4]. */
int odxMfmaProbe(const float *a, const float *b, float *d);

/* Use the network for warm mode 3 (NULL detaches it; mode 3 then warm-starts
 * from matched previous-step impulses and only records features). */
int odxBatchSetMlp(odxBatch *b, const odxMlp *mlp);
/* Features and converged targets (impulse / force scale) of the last step's
 * contacts recorded in warm mode 3; returns the count. */
int odxBatchGetSamples(odxBatch *b, float *X, float *T, int cap);

#ifdef __cplusplus
}
#endif
#endif
