// odx_step.hip -- batched world-step kernels for gfx950 (one wavefront per world).
//
// Kernel A (k_prepare_small): body stage (dxQuickStepIsland Stage0_Bodies and
//   Stage2b, quickstep.cpp:1176-1306, 1644-1690), geom AABBs, broadphase with
//   dxSAPSpace / dxHashSpace pair semantics, narrowphase (dCollide), contact and
//   ball rows (getInfo2 + Stage2a/2c + compute_invM_JT + Ad,
//   quickstep.cpp:1486-1747, 859-897, 2251-2316) and the sweep schedule.
// Kernel C (k_solve_small): SOR-LCP sweeps (IterationStep quickstep.cpp:2917-3033)
//   over the schedule with body accumulators fc in LDS and rows streamed from
//   HBM, an LCP residual check, then Stage4b/6a/6b and dxStepBody
//   (quickstep.cpp:3082-3102, 3299-3337; util.cpp:583-699).
//
// Schedules.  Gauss-Seidel is sequential; order 0 reproduces the reference's
// order exactly: rows are processed in "dependency levels" where a row's level
// is one more than the last level that touched either of its bodies, taken in
// the reference's ReorderPrep order (findex == -1 rows first, :2329-2355).
// Rows on one level touch disjoint bodies, so executing a level in parallel
// gives the same iterate as the sequential loop.  Order 1 colours whole joints
// (a contact's normal row then its friction rows in one lane) greedily so that
// no two joints of a colour share a body: a different, equally valid
// Gauss-Seidel order with far fewer phases.
#include <hip/hip_runtime.h>

#include "odx_collide.h"
#include "odx_device.h"
#include "odx_math.h"
#include "odx_mlp_dev.h"

namespace odx {

__device__ __forceinline__ int lane_id() { return (int)threadIdx.x & 63; }
__device__ __forceinline__ int as_int(float f) { return __float_as_int(f); }
__device__ __forceinline__ float as_float(int i) { return __int_as_float(i); }

// Sweep work item of the batched solve: first row in sweep order (11 bits),
// row count (4 bits), body 1 and body 2 (7 bits each; body 2 = nb when the
// rows have no second body).  Its rows are consecutive in sweep order.
constexpr int ITEM_MAXROWS = 15;
__device__ __forceinline__ int pack_item(int start, int nrow, int b1, int b2) {
    return start | (nrow << 11) | (b1 << 15) | (b2 << 22);
}
__device__ __forceinline__ int item_start(int it) { return it & 2047; }
__device__ __forceinline__ int item_rows(int it) { return (it >> 11) & 15; }
__device__ __forceinline__ int item_b1(int it) { return (it >> 15) & 127; }
__device__ __forceinline__ int item_b2(int it) { return (it >> 22) & 127; }

__device__ __forceinline__ int wave_incl_scan(int v) {
    const int l = lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(v, o, 64);
        if (l >= o) v += y;
    }
    return v;
}
__device__ __forceinline__ int wave_bcast_last(int v) { return __shfl(v, 63, 64); }
// Orders the LDS accesses of a one-wave workgroup: a wave's LDS operations
// execute in issue order, so a wavefront-scope fence (no wait) and a compiler
// barrier suffice where a workgroup barrier would also wait for every
// outstanding LDS operation.  Only for kernels launched with 64 threads.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// ------------------------------------------------------------ body stage --
// The total force and torque (accumulators, gravity, gyroscopic torque) go to
// fout / tout: the batched and single-world paths pass the accumulators
// themselves, the large world arrays of its own, so that a step that fails
// after this stage leaves every body record as it was.
ODX_DEV void body_prep(const DevWorlds &D, const StepParams &P, int b, float4 *fout, float4 *tout) {
    float4 p4 = D.pos[b], q4 = D.quat[b], lv = D.lvel[b], av = D.avel[b], fa = D.facc[b], ta = D.tacc[b];
    float4 ms = D.mass[b];
    const float mass = ms.x, im = ms.y;
    const int flags = (int)ms.z;
    float q[4] = {q4.x, q4.y, q4.z, q4.w};
    float R[12];
    RfromQ(R, q);
    float f[3] = {fa.x, fa.y, fa.z}, tq[3] = {ta.x, ta.y, ta.z};
    if (!(flags & 1)) {
        if (P.gx != 0.f) f[0] += mass * P.gx;
        if (P.gy != 0.f) f[1] += mass * P.gy;
        if (P.gz != 0.f) f[2] += mass * P.gz;
    }
    float ib[12], tmp[12], iw[12];
#pragma unroll
    for (int k = 0; k < 12; k++) ib[k] = D.invIb[(size_t)b * 12 + k];
    mul2_333(tmp, ib, R);
    mul0_333(iw, R, tmp);
    if ((flags & 2) && im > 0.f) {
        // implicit gyroscopic torque (Lacoursiere 2006), quickstep.cpp:1239-1303
        float Ibd[12], Iw[12], L[3], w[3] = {av.x, av.y, av.z};
#pragma unroll
        for (int k = 0; k < 12; k++) Ibd[k] = D.Ib[(size_t)b * 12 + k];
        mul2_333(tmp, Ibd, R);
        mul0_333(Iw, R, tmp);
        mul0_331(L, Iw, w);
        float It[12] = {0.f, L[2], -L[1], 0.f, -L[2], 0.f, L[0], 0.f, L[1], -L[0], 0.f, 0.f};
#pragma unroll
        for (int k = 0; k < 12; k++) It[k] = It[k] * P.h + Iw[k];
        const float hr = 1.f / P.h;
        L[0] *= hr; L[1] *= hr; L[2] *= hr;
        float itInv[12];
        if (invert3(itInv, It) != 0.f) {
            mul0_333(It, Iw, itInv);
            It[0] -= 1.f; It[5] -= 1.f; It[10] -= 1.f;
            float t0[3];
            mul0_331(t0, It, L);
            tq[0] += t0[0]; tq[1] += t0[1]; tq[2] += t0[2];
        }
    }
    fout[b] = make_float4(f[0], f[1], f[2], 0.f);
    tout[b] = make_float4(tq[0], tq[1], tq[2], 0.f);
    // Stage2b: rhs_tmp = -(v/h + M^-1 f)
    const float hr = 1.f / P.h;
    float ra[3];
    mul0_331(ra, iw, tq);
    float *o = D.bscr + (size_t)b * BSCR_STRIDE;
#pragma unroll
    for (int k = 0; k < 12; k++) { o[k] = R[k]; o[12 + k] = iw[k]; }
    o[24] = -(f[0] * im + lv.x * hr);
    o[25] = -(f[1] * im + lv.y * hr);
    o[26] = -(f[2] * im + lv.z * hr);
    o[27] = -(av.x * hr) - ra[0];
    o[28] = -(av.y * hr) - ra[1];
    o[29] = -(av.z * hr) - ra[2];
    o[30] = im;
    o[31] = 0.f;
    (void)p4;
}

// ------------------------------------------------------------- geoms ------
ODX_DEV void load_geom(const DevWorlds &D, int g, int b0, DGeom &G) {
    float4 a = D.geom0[g], c = D.geom1[g];
    G.cls = (int)a.x;
    G.body = (int)a.y;
    G.p[0] = a.z; G.p[1] = a.w; G.p[2] = c.x; G.p[3] = c.y;
    if (G.body >= 0) {
        float4 p = D.pos[b0 + G.body];
        G.pos[0] = p.x; G.pos[1] = p.y; G.pos[2] = p.z;
        const float *s = D.bscr + (size_t)(b0 + G.body) * BSCR_STRIDE;
#pragma unroll
        for (int k = 0; k < 12; k++) G.R[k] = s[k];
    } else {
        float4 p = D.gpos[g], q4 = D.gquat[g];
        G.pos[0] = p.x; G.pos[1] = p.y; G.pos[2] = p.z;
        float q[4] = {q4.x, q4.y, q4.z, q4.w};
        RfromQ(G.R, q);
    }
}

// ------------------------------------------------------------- rows -------
__device__ __forceinline__ int contact_rows_m(const Surface &s) {
    // getInfo1 contact.cpp:48-123 (negative mu counts as zero)
    if (s.mode & 0x001) return 1 + (s.mu > 0.f ? 1 : 0) + (s.mu2 > 0.f ? 1 : 0);
    return 1 + (s.mu > 0.f ? 2 : 0);
}

struct RowBuf {
    float J1[6], rhs, cfm, J2[6], lo, hi;
    int findex;
};

// getInfo2 of a contact joint, contact.cpp:125-345 (rolling friction not on
// this path).  b1: body row index >= 0, b2 may be -1.
ODX_DEV int contact_info2(const DevWorlds &D, const StepParams &P, const float *rec, int b0, RowBuf *rows) {
    const Surface s = D.surf[(int)rec[13]];
    const int m = contact_rows_m(s);
    const int b1 = (int)rec[1], b2 = (int)rec[2];
    const int mode = s.mode;
    const float fps = 1.f / P.h;
    float erp = (mode & 0x008) ? s.soft_erp : P.erp;
    float k = fps * erp;
    float depth = rec[9] - P.min_depth;
    if (depth < 0.f) depth = 0.f;
    float motionN = (mode & 0x080) ? s.motionN : 0.f;
    const float pushout = k * depth + motionN;
    const bool bounce = (mode & 0x004) && s.bounce_vel >= 0.f;
    float c = pushout > P.max_vel ? P.max_vel : pushout;
    float n[3] = {rec[6], rec[7], rec[8]};
    float pos[3] = {rec[3], rec[4], rec[5]};
    float c1[3], c2[3] = {0.f, 0.f, 0.f};
    float4 p1 = D.pos[b0 + b1];
    c1[0] = pos[0] - p1.x; c1[1] = pos[1] - p1.y; c1[2] = pos[2] - p1.z;
    float outgoing = 0.f;
    if (b2 >= 0) {
        float4 p2 = D.pos[b0 + b2];
        c2[0] = pos[0] - p2.x; c2[1] = pos[1] - p2.y; c2[2] = pos[2] - p2.z;
        rows[0].J2[0] = -n[0]; rows[0].J2[1] = -n[1]; rows[0].J2[2] = -n[2];
        cross3(rows[0].J2 + 3, n, c2);
        if (bounce) {
            float4 av = D.avel[b0 + b2], lv = D.lvel[b0 + b2];
            float a[3] = {av.x, av.y, av.z}, l[3] = {lv.x, lv.y, lv.z};
            outgoing = dot3(rows[0].J2 + 3, a) - dot3(n, l);
        }
    }
    rows[0].J1[0] = n[0]; rows[0].J1[1] = n[1]; rows[0].J1[2] = n[2];
    cross3(rows[0].J1 + 3, c1, n);
    if (bounce) {
        float4 av = D.avel[b0 + b1], lv = D.lvel[b0 + b1];
        float a[3] = {av.x, av.y, av.z}, l[3] = {lv.x, lv.y, lv.z};
        outgoing += dot3(rows[0].J1 + 3, a) + dot3(n, l);
        float neg = motionN - outgoing;
        if (neg > s.bounce_vel) {
            float newc = s.bounce * neg + motionN;
            if (newc > c) c = newc;
        }
    }
    rows[0].rhs = c;
    if (mode & 0x010) rows[0].cfm = s.soft_cfm;
    rows[0].lo = 0.f;
    rows[0].hi = kInf;
    if (m > 1) {
        float t1[3], t2[3];
        if (mode & 0x002) {
            t1[0] = rec[10]; t1[1] = rec[11]; t1[2] = rec[12];
            cross3(t2, n, t1);
        } else {
            planeSpace(n, t1, t2);
        }
        // friction rows: t1 in row 1 when mu > 0, then t2 in the next row
        // (static row indices keep the rows in registers)
        auto friction = [&](RowBuf &w, const float *t, float mu_, int mrhs, float rhs_, int mcfm, float cfm_, int mf) {
            w.J1[0] = t[0]; w.J1[1] = t[1]; w.J1[2] = t[2];
            cross3(w.J1 + 3, c1, t);
            if (b2 >= 0) {
                w.J2[0] = -t[0]; w.J2[1] = -t[1]; w.J2[2] = -t[2];
                cross3(w.J2 + 3, t, c2);
            }
            if (mode & mrhs) w.rhs = rhs_;
            if (mode & mcfm) w.cfm = cfm_;
            w.lo = -mu_; w.hi = mu_;
            if (mode & mf) w.findex = 0;
        };
        const bool f1 = s.mu > 0.f;
        if (f1) friction(rows[1], t1, s.mu, 0x020, s.motion1, 0x100, s.slip1, 0x1000);
        const float mu2 = (mode & 0x001) ? s.mu2 : s.mu;
        if (mu2 > 0.f) {
            if (f1) { if (m > 2) friction(rows[2], t2, mu2, 0x040, s.motion2, 0x200, s.slip2, 0x2000); }
            else friction(rows[1], t2, mu2, 0x040, s.motion2, 0x200, s.slip2, 0x2000);
        }
    }
    return m;
}

// getInfo2 of a ball joint, ball.cpp:58-67 -> setBall joint.cpp:111-162
ODX_DEV int ball_info2(const DevWorlds &D, const StepParams &P, const float *rec, int b0, RowBuf *rows) {
    const int b1 = (int)rec[1], b2 = (int)rec[2];
    const float cfm = rec[10], erp = rec[9];
    for (int r = 0; r < 3; r++) rows[r].cfm = cfm;
    const float *R1 = D.bscr + (size_t)(b0 + b1) * BSCR_STRIDE;
    float an1[3] = {rec[3], rec[4], rec[5]}, an2[3] = {rec[6], rec[7], rec[8]}, a1[3], a2[3] = {0.f, 0.f, 0.f};
    rows[0].J1[0] = 1.f; rows[1].J1[1] = 1.f; rows[2].J1[2] = 1.f;
    mul0_331(a1, R1, an1);
    rows[0].J1[4] = a1[2]; rows[0].J1[5] = -a1[1];
    rows[1].J1[3] = -a1[2]; rows[1].J1[5] = a1[0];
    rows[2].J1[3] = a1[1]; rows[2].J1[4] = -a1[0];
    float4 p1 = D.pos[b0 + b1];
    float x1[3] = {p1.x, p1.y, p1.z};
    const float k = (1.f / P.h) * erp;
    if (b2 >= 0) {
        const float *R2 = D.bscr + (size_t)(b0 + b2) * BSCR_STRIDE;
        rows[0].J2[0] = -1.f; rows[1].J2[1] = -1.f; rows[2].J2[2] = -1.f;
        mul0_331(a2, R2, an2);
        rows[0].J2[4] = -a2[2]; rows[0].J2[5] = a2[1];
        rows[1].J2[3] = a2[2]; rows[1].J2[5] = -a2[0];
        rows[2].J2[3] = -a2[1]; rows[2].J2[4] = a2[0];
        float4 p2 = D.pos[b0 + b2];
        float x2[3] = {p2.x, p2.y, p2.z};
        for (int r = 0; r < 3; r++) rows[r].rhs = k * (a2[r] + x2[r] - a1[r] - x1[r]);
    } else {
        for (int r = 0; r < 3; r++) rows[r].rhs = k * (an2[r] - a1[r] - x1[r]);
    }
    return 3;
}

// qa = qb * qc with qb and / or qc conjugated (dQMultiply0..3, rotation.cpp:41-107)
__device__ __forceinline__ void qmul(float *qa, const float *qb, const float *qc, bool cb, bool cc) {
    const float b1 = cb ? -qb[1] : qb[1], b2 = cb ? -qb[2] : qb[2], b3 = cb ? -qb[3] : qb[3];
    const float c1 = cc ? -qc[1] : qc[1], c2 = cc ? -qc[2] : qc[2], c3 = cc ? -qc[3] : qc[3];
    qa[0] = qb[0] * qc[0] - b1 * c1 - b2 * c2 - b3 * c3;
    qa[1] = qb[0] * c1 + b1 * qc[0] + b2 * c3 - b3 * c2;
    qa[2] = qb[0] * c2 + b2 * qc[0] + b3 * c1 - b1 * c3;
    qa[3] = qb[0] * c3 + b3 * qc[0] + b1 * c2 - b2 * c1;
}

// setFixedOrientation joints/joint.cpp:228-285: angular rows making the
// relative rotation equal qrel; rows[0..2]
ODX_DEV void fixed_orientation_rows(const DevWorlds &D, int b0, int b1, int b2, const float *x, float fps, float erp,
                                       RowBuf *rows) {
    rows[0].J1[3] = 1.f; rows[1].J1[4] = 1.f; rows[2].J1[5] = 1.f;
    if (b2 >= 0) { rows[0].J2[3] = -1.f; rows[1].J2[4] = -1.f; rows[2].J2[5] = -1.f; }
    const float4 q1v = D.quat[b0 + b1];
    const float q1[4] = {q1v.x, q1v.y, q1v.z, q1v.w};
    float qerr[4];
    if (b2 >= 0) {
        const float4 q2v = D.quat[b0 + b2];
        const float q2[4] = {q2v.x, q2v.y, q2v.z, q2v.w};
        float qq[4];
        qmul(qq, q1, q2, true, false);
        qmul(qerr, qq, x + JX_QREL, false, true);
    } else {
        qmul(qerr, q1, x + JX_QREL, true, true);
    }
    if (qerr[0] < 0.f) { qerr[1] = -qerr[1]; qerr[2] = -qerr[2]; qerr[3] = -qerr[3]; }
    float e[3];
    mul0_331(e, D.bscr + (size_t)(b0 + b1) * BSCR_STRIDE, qerr + 1);
    const float k2 = fps * erp * 2.f;
    for (int r = 0; r < 3; r++) rows[r].rhs = k2 * e[r];
}

// dxJointLimitMotor::addLimot joints/joint.cpp:598-786 for the row along ax1;
// the force of a motor powered at its limit is added to the body accumulators
// on the host before the step
ODX_DEV void limot_row(const DevWorlds &D, int b0, int b1, int b2, const float *x, float fps, const float *ax1,
                          bool rotational, RowBuf &row) {
    const int limit = (int)x[JX_LIMIT];
    bool powered = x[JX_FMAX] > 0.f;
    float *J1u = rotational ? row.J1 + 3 : row.J1, *J2u = rotational ? row.J2 + 3 : row.J2;
    for (int k = 0; k < 3; k++) J1u[k] = ax1[k];
    if (b2 >= 0) for (int k = 0; k < 3; k++) J2u[k] = -ax1[k];
    if (!rotational && b2 >= 0) {  // linear limot torque decoupling
        const float4 p1 = D.pos[b0 + b1], p2 = D.pos[b0 + b2];
        const float c[3] = {0.5f * (p2.x - p1.x), 0.5f * (p2.y - p1.y), 0.5f * (p2.z - p1.z)};
        float ltd[3];
        cross3(ltd, c, ax1);
        for (int k = 0; k < 3; k++) { row.J1[3 + k] = ltd[k]; row.J2[3 + k] = ltd[k]; }
    }
    const float lo = x[JX_LOSTOP], hi = x[JX_HISTOP];
    if (limit && lo == hi) powered = false;
    if (powered) {
        row.cfm = x[JX_NORMAL_CFM];
        if (!limit) {
            row.rhs = x[JX_VEL];
            row.lo = -x[JX_FMAX];
            row.hi = x[JX_FMAX];
        }
    }
    if (limit) {
        row.rhs = -fps * x[JX_STOP_ERP] * x[JX_LIMIT_ERR];
        row.cfm = x[JX_STOP_CFM];
        if (lo == hi) {
            row.lo = -kInf;
            row.hi = kInf;
        } else {
            if (limit == 1) { row.lo = 0.f; row.hi = kInf; }
            else { row.lo = -kInf; row.hi = 0.f; }
            const float bounce = x[JX_BOUNCE];
            if (bounce > 0.f) {
                float v;
                if (rotational) {
                    const float4 w1 = D.avel[b0 + b1];
                    v = w1.x * ax1[0] + w1.y * ax1[1] + w1.z * ax1[2];
                    if (b2 >= 0) { const float4 w2 = D.avel[b0 + b2]; v -= w2.x * ax1[0] + w2.y * ax1[1] + w2.z * ax1[2]; }
                } else {
                    const float4 v1 = D.lvel[b0 + b1];
                    v = v1.x * ax1[0] + v1.y * ax1[1] + v1.z * ax1[2];
                    if (b2 >= 0) { const float4 v2 = D.lvel[b0 + b2]; v -= v2.x * ax1[0] + v2.y * ax1[1] + v2.z * ax1[2]; }
                }
                if (limit == 1) {
                    if (v < 0.f) { const float nc = -bounce * v; if (nc > row.rhs) row.rhs = nc; }
                } else {
                    if (v > 0.f) { const float nc = -bounce * v; if (nc < row.rhs) row.rhs = nc; }
                }
            }
        }
    }
}

// getInfo2 of hinge (hinge.cpp:80-153), slider (slider.cpp:146-250) and fixed
// (fixed.cpp:68-105) joints
ODX_DEV int hinge_slider_fixed_info2(const DevWorlds &D, const StepParams &P, const float *rec, const float *x,
                                        int b0, RowBuf *rows) {
    const int type = (int)rec[0];
    const int b1 = (int)rec[1], b2 = (int)rec[2];
    const float fps = 1.f / P.h;
    const float *R1 = D.bscr + (size_t)(b0 + b1) * BSCR_STRIDE;
    const float4 p1v = D.pos[b0 + b1];
    const float x1[3] = {p1v.x, p1v.y, p1v.z};
    float ax1[3];
    mul0_331(ax1, R1, x + JX_AXIS1);
    if (type == JT_FIXED) {
        fixed_orientation_rows(D, b0, b1, b2, x, fps, P.erp, rows + 3);
        rows[0].J1[0] = 1.f; rows[1].J1[1] = 1.f; rows[2].J1[2] = 1.f;
        const float k = fps * x[JX_ERP];
        float ofs[3];
        mul0_331(ofs, R1, x + JX_OFFSET);
        if (b2 >= 0) {
            rows[0].J1[4] = -ofs[2]; rows[0].J1[5] = ofs[1];
            rows[1].J1[3] = ofs[2]; rows[1].J1[5] = -ofs[0];
            rows[2].J1[3] = -ofs[1]; rows[2].J1[4] = ofs[0];
            rows[0].J2[0] = -1.f; rows[1].J2[1] = -1.f; rows[2].J2[2] = -1.f;
            const float4 p2v = D.pos[b0 + b2];
            const float x2[3] = {p2v.x, p2v.y, p2v.z};
            for (int r = 0; r < 3; r++) rows[r].rhs = k * (x2[r] - x1[r] + ofs[r]);
        } else {
            for (int r = 0; r < 3; r++) rows[r].rhs = k * (x[JX_OFFSET + r] - x1[r]);
        }
        for (int r = 0; r < 3; r++) rows[r].cfm = x[JX_CFM];
        return 6;
    }
    float p[3], q[3];
    planeSpace(ax1, p, q);
    const float k = fps * P.erp;
    if (type == JT_HINGE) {
        // the ball rows with the world's erp, then the two axis rows
        float brec[16];
        for (int i = 0; i < 16; i++) brec[i] = rec[i];
        brec[9] = P.erp;
        ball_info2(D, P, brec, b0, rows);
        for (int c = 0; c < 3; c++) { rows[3].J1[3 + c] = p[c]; rows[4].J1[3 + c] = q[c]; }
        if (b2 >= 0) for (int c = 0; c < 3; c++) { rows[3].J2[3 + c] = -p[c]; rows[4].J2[3 + c] = -q[c]; }
        float bv[3];
        if (b2 >= 0) {
            float ax2[3];
            mul0_331(ax2, D.bscr + (size_t)(b0 + b2) * BSCR_STRIDE, x + JX_AXIS2);
            cross3(bv, ax1, ax2);
        } else {
            cross3(bv, ax1, x + JX_AXIS2);
        }
        rows[3].rhs = k * dot3(bv, p);
        rows[4].rhs = k * dot3(bv, q);
        if (x[JX_LIMOT_ROW] != 0.f) { limot_row(D, b0, b1, b2, x, fps, ax1, true, rows[5]); return 6; }
        return 5;
    }
    // slider
    fixed_orientation_rows(D, b0, b1, b2, x, fps, P.erp, rows);
    float c[3] = {0.f, 0.f, 0.f};
    for (int i = 0; i < 3; i++) { rows[3].J1[i] = p[i]; rows[4].J1[i] = q[i]; }
    if (b2 >= 0) {
        const float4 p2v = D.pos[b0 + b2];
        c[0] = p2v.x - x1[0]; c[1] = p2v.y - x1[1]; c[2] = p2v.z - x1[2];
        float t[3];
        for (int i = 0; i < 3; i++) { rows[3].J2[i] = -p[i]; rows[4].J2[i] = -q[i]; }
        cross3(t, c, p);
        for (int i = 0; i < 3; i++) { rows[3].J1[3 + i] = 0.5f * t[i]; rows[3].J2[3 + i] = rows[3].J1[3 + i]; }
        cross3(t, c, q);
        for (int i = 0; i < 3; i++) { rows[4].J1[3 + i] = 0.5f * t[i]; rows[4].J2[3 + i] = rows[4].J1[3 + i]; }
        float ofs[3];
        mul0_331(ofs, D.bscr + (size_t)(b0 + b2) * BSCR_STRIDE, x + JX_OFFSET);
        for (int i = 0; i < 3; i++) c[i] += ofs[i];
        rows[3].rhs = k * dot3(p, c);
        rows[4].rhs = k * dot3(q, c);
    } else {
        float ofs[3];
        for (int i = 0; i < 3; i++) ofs[i] = x[JX_OFFSET + i] - x1[i];
        rows[3].rhs = k * dot3(p, ofs);
        rows[4].rhs = k * dot3(q, ofs);
        if (x[JX_REVERSE] != 0.f) for (int i = 0; i < 3; i++) ax1[i] = -ax1[i];
    }
    if (x[JX_LIMOT_ROW] != 0.f) { limot_row(D, b0, b1, b2, x, fps, ax1, false, rows[5]); return 6; }
    return 5;
}

// Builds the rows of one joint (Stage2a scaling, Stage2c rhs, iMJ, Ad) and
// writes them at world-local row index ri.
// getInfo1 of a hinge (hinge.cpp:58-77) or slider (slider.cpp:112-143) of a
// batched world, from the bodies' poses at the start of the step: the limit
// state (testRotationalLimit joint.cpp:574-595 for the hinge angle of
// joint.cpp:421-490; the slider position of slider.cpp:45-82), whether the
// limit / motor row is present (x[JX_LIMOT_ROW]) and the force of a motor
// powered at its limit, added to the bodies' accumulators as addLimot does
// (joint.cpp:640-705).  A hinge whose stops are not finite keeps the limit
// state it has.  Accumulator adds are atomic: several joints may share a body.
ODX_DEV void joint_info1(const DevWorlds &D, int b0, const float *rec, float *x) {
    const int type = (int)rec[0];
    if (type != JT_HINGE && type != JT_SLIDER) return;
    const float pi = 3.14159265358979f;
    const int b1 = (int)rec[1], b2 = (int)rec[2];
    const float4 q1v = D.quat[b0 + b1];
    const float q1[4] = {q1v.x, q1v.y, q1v.z, q1v.w};
    float R1[12];
    RfromQ(R1, q1);
    const float lo = x[JX_LOSTOP], hi = x[JX_HISTOP];
    int limit = (int)x[JX_LIMIT];
    float err = x[JX_LIMIT_ERR];
    if (type == JT_HINGE) {
        if ((lo >= -pi || hi <= pi) && lo <= hi) {
            float qr[4];
            if (b2 >= 0) {
                const float4 q2v = D.quat[b0 + b2];
                const float q2[4] = {q2v.x, q2v.y, q2v.z, q2v.w};
                float qq[4];
                qmul(qq, q1, q2, true, false);
                qmul(qr, qq, x + JX_QREL, false, true);
            } else {
                qmul(qr, q1, x + JX_QREL, true, true);
            }
            const float cost2 = qr[0], sint2 = sqrtf(qr[1] * qr[1] + qr[2] * qr[2] + qr[3] * qr[3]);
            const float d = qr[1] * x[JX_AXIS1] + qr[2] * x[JX_AXIS1 + 1] + qr[3] * x[JX_AXIS1 + 2];
            float th = d >= 0.f ? 2.f * atan2f(sint2, cost2) : 2.f * atan2f(sint2, -cost2);
            if (th > pi) th -= 2.f * pi;
            const float a = -th;
            if (a <= lo) { limit = 1; err = a - lo; }
            else if (a >= hi) { limit = 2; err = a - hi; }
            else limit = 0;
        }
    } else {
        limit = 0;
        if ((lo > -kInf || hi < kInf) && lo <= hi) {
            float ax1[3], q[3];
            mul0_331(ax1, R1, x + JX_AXIS1);
            const float4 p1 = D.pos[b0 + b1];
            if (b2 >= 0) {
                const float4 q2v = D.quat[b0 + b2], p2 = D.pos[b0 + b2];
                const float q2[4] = {q2v.x, q2v.y, q2v.z, q2v.w};
                float R2[12];
                RfromQ(R2, q2);
                mul0_331(q, R2, x + JX_OFFSET);
                q[0] = p1.x - q[0] - p2.x; q[1] = p1.y - q[1] - p2.y; q[2] = p1.z - q[2] - p2.z;
            } else {
                q[0] = p1.x - x[JX_OFFSET]; q[1] = p1.y - x[JX_OFFSET + 1]; q[2] = p1.z - x[JX_OFFSET + 2];
                if (x[JX_REVERSE] != 0.f) { ax1[0] = -ax1[0]; ax1[1] = -ax1[1]; ax1[2] = -ax1[2]; }
            }
            const float pos = dot3(ax1, q);
            if (pos <= lo) { limit = 1; err = pos - lo; }
            else if (pos >= hi) { limit = 2; err = pos - hi; }
        }
    }
    const float fmax = x[JX_FMAX];
    x[JX_LIMIT] = (float)limit;
    x[JX_LIMIT_ERR] = err;
    x[JX_LIMOT_ROW] = (limit || fmax > 0.f) ? 1.f : 0.f;
    if (fmax > 0.f && limit && lo != hi) {
        float fm = fmax;
        const float vel = x[JX_VEL];
        if (vel > 0.f || (vel == 0.f && limit == 2)) fm = -fm;
        if ((limit == 1 && vel > 0.f) || (limit == 2 && vel < 0.f)) fm *= x[JX_FUDGE];
        float ax1[3];
        mul0_331(ax1, R1, x + JX_AXIS1);
        if (type == JT_SLIDER && b2 < 0 && x[JX_REVERSE] != 0.f) { ax1[0] = -ax1[0]; ax1[1] = -ax1[1]; ax1[2] = -ax1[2]; }
        const float f[3] = {fm * ax1[0], fm * ax1[1], fm * ax1[2]};
        float4 *T1 = D.tacc + b0 + b1, *F1 = D.facc + b0 + b1;
        if (type == JT_HINGE) {
            if (b2 >= 0) {
                float4 *T2 = D.tacc + b0 + b2;
                atomicAdd(&T2->x, f[0]); atomicAdd(&T2->y, f[1]); atomicAdd(&T2->z, f[2]);
            }
            atomicAdd(&T1->x, -f[0]); atomicAdd(&T1->y, -f[1]); atomicAdd(&T1->z, -f[2]);
        } else {
            if (b2 >= 0) {
                const float4 p1 = D.pos[b0 + b1], p2 = D.pos[b0 + b2];
                const float c[3] = {0.5f * (p2.x - p1.x), 0.5f * (p2.y - p1.y), 0.5f * (p2.z - p1.z)};
                float ltd[3];
                cross3(ltd, c, ax1);
                float4 *T2 = D.tacc + b0 + b2, *F2 = D.facc + b0 + b2;
                atomicAdd(&T1->x, -fm * ltd[0]); atomicAdd(&T1->y, -fm * ltd[1]); atomicAdd(&T1->z, -fm * ltd[2]);
                atomicAdd(&T2->x, -fm * ltd[0]); atomicAdd(&T2->y, -fm * ltd[1]); atomicAdd(&T2->z, -fm * ltd[2]);
                atomicAdd(&F2->x, f[0]); atomicAdd(&F2->y, f[1]); atomicAdd(&F2->z, f[2]);
            }
            atomicAdd(&F1->x, -f[0]); atomicAdd(&F1->y, -f[1]); atomicAdd(&F1->z, -f[2]);
        }
    }
}

// x: the joint's extended record (hinge / slider / fixed), else unused.
ODX_DEV int build_joint_rows(const DevWorlds &D, const StepParams &P, const float *rec, const float *x, int b0,
                                int r0, int ri) {
    RowBuf rows[6];
#pragma unroll
    for (int r = 0; r < 6; r++) {
#pragma unroll
        for (int k = 0; k < 6; k++) { rows[r].J1[k] = 0.f; rows[r].J2[k] = 0.f; }
        rows[r].rhs = 0.f;
        rows[r].cfm = P.cfm;
        rows[r].lo = -kInf;
        rows[r].hi = kInf;
        rows[r].findex = -1;
    }
    const int type = (int)rec[0];
    int m;
    if (type == 0) m = contact_info2(D, P, rec, b0, rows);
    else if (type == 1) m = ball_info2(D, P, rec, b0, rows);
    else m = hinge_slider_fixed_info2(D, P, rec, x, b0, rows);
    const int b1 = (int)rec[1], b2 = (int)rec[2];
    const float fps = 1.f / P.h;
    const float *s1 = D.bscr + (size_t)(b0 + b1) * BSCR_STRIDE;
    const float *s2 = b2 >= 0 ? D.bscr + (size_t)(b0 + b2) * BSCR_STRIDE : nullptr;
#pragma unroll
    for (int r = 0; r < 6; r++) {
        if (r >= m) break;
        RowBuf &w = rows[r];
        w.rhs *= fps;
        w.cfm *= fps;
        float sum = 0.f;
#pragma unroll
        for (int k = 0; k < 6; k++) sum += w.J1[k] * s1[24 + k];
        if (s2) {
#pragma unroll
            for (int k = 0; k < 6; k++) sum += w.J2[k] * s2[24 + k];
        }
        w.rhs += sum;
        float imj1[6], imj2[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        const float k1 = s1[30];
        imj1[0] = k1 * w.J1[0]; imj1[1] = k1 * w.J1[1]; imj1[2] = k1 * w.J1[2];
        mul0_331(imj1 + 3, s1 + 12, w.J1 + 3);
        if (s2) {
            const float k2 = s2[30];
            imj2[0] = k2 * w.J2[0]; imj2[1] = k2 * w.J2[1]; imj2[2] = k2 * w.J2[2];
            mul0_331(imj2 + 3, s2 + 12, w.J2 + 3);
        }
        float ssum = 0.f;
#pragma unroll
        for (int k = 0; k < 6; k++) ssum += imj1[k] * w.J1[k];
        if (s2) {
#pragma unroll
            for (int k = 0; k < 6; k++) ssum += imj2[k] * w.J2[k];
        }
        const float Ad = P.sor_w / (ssum + w.cfm);
        float *o = D.rows + (size_t)(r0 + ri + r) * ROW_STRIDE;
        float4 *o4 = reinterpret_cast<float4 *>(o);
        float J1s[6], J2s[6];
#pragma unroll
        for (int k = 0; k < 6; k++) { J1s[k] = w.J1[k] * Ad; J2s[k] = s2 ? w.J2[k] * Ad : w.J2[k]; }
        o4[0] = make_float4(J1s[0], J1s[1], J1s[2], J1s[3]);
        o4[1] = make_float4(J1s[4], J1s[5], w.rhs * Ad, w.cfm * Ad);
        o4[2] = make_float4(J2s[0], J2s[1], J2s[2], J2s[3]);
        o4[3] = make_float4(J2s[4], J2s[5], w.lo, w.hi);
        o4[4] = make_float4(imj1[0], imj1[1], imj1[2], imj1[3]);
        o4[5] = make_float4(imj1[4], imj1[5], imj2[0], imj2[1]);
        o4[6] = make_float4(imj2[2], imj2[3], imj2[4], imj2[5]);
        const int fidx = w.findex >= 0 ? ri + w.findex : -1;
        o4[7] = make_float4(as_float(b1), as_float(b2), as_float(fidx), Ad);
    }
    return m;
}

// Warm-start feature vector of a contact (its first row at world-local ri).
// Nearest previous-step contact of the same geom pair, compared in body-1
// coordinates (contacts are rebuilt every step; the reference keeps no contact
// identity, quickstep.cpp:46-51).  The previous friction impulse is carried
// over as a tangential force vector and re-expressed in the new tangent basis.
// Returns 1 and lam[3] = (normal, t1, t2) when a match is found.
ODX_DEV int match_previous(const DevWorlds &D, int j0, int nprev, const float *rec, int b0, float *lam) {
    lam[0] = lam[1] = lam[2] = 0.f;
    if (nprev <= 0) return 0;
    const int g1 = (int)rec[14], g2 = (int)rec[15];
    const unsigned key = ((unsigned)g1 << 16) | (unsigned)g2;
    const float *pv = D.prev + (size_t)j0 * JOINT_STRIDE;
    int lo = 0, hi = nprev;
    while (lo < hi) {  // first record with pair key >= key (records are in pair order)
        const int mid = (lo + hi) >> 1;
        const float *r = pv + (size_t)mid * JOINT_STRIDE;
        const unsigned km = ((unsigned)(int)r[PV_G1] << 16) | (unsigned)(int)r[PV_G2];
        if (km < key) lo = mid + 1; else hi = mid;
    }
    const int b1 = (int)rec[1];
    const float *s1 = D.bscr + (size_t)(b0 + b1) * BSCR_STRIDE;
    float4 x1 = D.pos[b0 + b1];
    float d[3] = {rec[3] - x1.x, rec[4] - x1.y, rec[5] - x1.z}, pl[3];
    mul1_331(pl, s1, d);
    int best = -1;
    float bestd = MATCH_DIST * MATCH_DIST;
    for (int i = lo; i < nprev && i < lo + 8; i++) {
        const float *r = pv + (size_t)i * JOINT_STRIDE;
        if ((int)r[PV_G1] != g1 || (int)r[PV_G2] != g2) break;
        const float dx = r[PV_LOCAL] - pl[0], dy = r[PV_LOCAL + 1] - pl[1], dz = r[PV_LOCAL + 2] - pl[2];
        const float dd = dx * dx + dy * dy + dz * dz;
        if (dd <= bestd) { bestd = dd; best = i; }
    }
    if (best < 0) return 0;
    const float *r = pv + (size_t)best * JOINT_STRIDE;
    float no[3] = {r[PV_N], r[PV_N + 1], r[PV_N + 2]}, to1[3], to2[3];
    planeSpace(no, to1, to2);
    float f[3];
    for (int k = 0; k < 3; k++) f[k] = r[PV_LAM + 1] * to1[k] + r[PV_LAM + 2] * to2[k];
    float n[3] = {rec[6], rec[7], rec[8]}, t1[3], t2[3];
    planeSpace(n, t1, t2);
    lam[0] = r[PV_LAM];
    lam[1] = dot3(f, t1);
    lam[2] = dot3(f, t2);
    return 1;
}

ODX_DEV void contact_features(const DevWorlds &D, const float *rec, int b0, const float *lam_prev, int matched,
                                 float *out) {
    const int b1 = (int)rec[1], b2 = (int)rec[2];
    float n[3] = {rec[6], rec[7], rec[8]}, t1[3], t2[3];
    planeSpace(n, t1, t2);
    float4 p1 = D.pos[b0 + b1], v1 = D.lvel[b0 + b1], w1 = D.avel[b0 + b1];
    float c1[3] = {rec[3] - p1.x, rec[4] - p1.y, rec[5] - p1.z};
    float w1a[3] = {w1.x, w1.y, w1.z}, u[3];
    cross3(u, w1a, c1);
    float vr[3] = {v1.x + u[0], v1.y + u[1], v1.z + u[2]};
    float im2 = 0.f;
    if (b2 >= 0) {
        float4 p2 = D.pos[b0 + b2], v2 = D.lvel[b0 + b2], w2 = D.avel[b0 + b2];
        float c2[3] = {rec[3] - p2.x, rec[4] - p2.y, rec[5] - p2.z};
        float w2a[3] = {w2.x, w2.y, w2.z};
        cross3(u, w2a, c2);
        vr[0] -= v2.x + u[0]; vr[1] -= v2.y + u[1]; vr[2] -= v2.z + u[2];
        im2 = D.bscr[(size_t)(b0 + b2) * BSCR_STRIDE + 30];
    }
    const float im1 = D.bscr[(size_t)(b0 + b1) * BSCR_STRIDE + 30];
    const float s = GRAVITY_SCALE / (im1 + im2);
    const float si = 1.f / s;
    out[0] = n[0]; out[1] = n[1]; out[2] = n[2];
    out[3] = rec[9];
    out[4] = dot3(vr, n); out[5] = dot3(vr, t1); out[6] = dot3(vr, t2);
    out[7] = lam_prev[0] * si; out[8] = lam_prev[1] * si; out[9] = lam_prev[2] * si;
    out[10] = im1; out[11] = im2;
    out[12] = (float)matched;
    out[13] = rec[5] - p1.z;
    out[14] = (float)(b2 >= 0);
    out[15] = s;
}

// Warm-start value of each row of a contact from (normal, t1, t2) impulses,
// following the row order of contact_info2.
ODX_DEV void write_contact_lambda0(const DevWorlds &D, const float *rec, int r0, int ri, const float *lam) {
    const Surface s = D.surf[(int)rec[13]];
    const int m = contact_rows_m(s);
    float *o = D.lambda0 + r0 + ri;
    o[0] = lam[0];
    int row = 1;
    if (m > 1) {
        if (s.mu > 0.f) o[row++] = lam[1];
        const float mu2 = (s.mode & 0x001) ? s.mu2 : s.mu;
        if (mu2 > 0.f && row < m) o[row++] = lam[2];
    }
}

// ============================================================== kernel A ===
// LDS of the prepare stage of one world (one wave).  Laid out by
// prep_lds_carve over one buffer so that a fused kernel can reuse the space
// for the solve stage.
struct PrepLds {
    unsigned long long *mask;  // [SW_MAXB]
    float (*ab)[6];            // [SW_MAXG]
    uint32_t *pairs;           // [SW_MAXPAIRS]
    int *cnt;                  // [SW_MAXPH + 1]
    unsigned short *lvl, *joff;  // [SW_MAXROWS]
    unsigned char *rb1, *rb2, *rfl, *jm;  // [SW_MAXROWS]
    signed char *gb;           // [SW_MAXG]
    int *L, *nblk;             // scalars
    float *clip;               // [CLIP_WS * 64]: the narrowphase's box-box clipper, one column per lane;
                               // aliases cnt..jm, which the row and schedule stages fill afterwards
};
constexpr size_t PREP_LDS_BYTES = SW_MAXB * 8 + SW_MAXG * 6 * 4 + SW_MAXPAIRS * 4 + (SW_MAXPH + 1) * 4 +
                                  2 * SW_MAXROWS * 2 + 4 * SW_MAXROWS + SW_MAXG + 2 * 4 + 16;
__device__ inline PrepLds prep_lds_carve(void *base) {
    char *p = reinterpret_cast<char *>(base);  // 16-byte aligned
    PrepLds S;
    S.mask = reinterpret_cast<unsigned long long *>(p); p += SW_MAXB * 8;
    S.ab = reinterpret_cast<float (*)[6]>(p); p += SW_MAXG * 6 * 4;
    S.pairs = reinterpret_cast<uint32_t *>(p); p += SW_MAXPAIRS * 4;
    S.cnt = reinterpret_cast<int *>(p); p += (SW_MAXPH + 1) * 4;
    S.L = reinterpret_cast<int *>(p); p += 4;
    S.nblk = reinterpret_cast<int *>(p); p += 4;
    S.lvl = reinterpret_cast<unsigned short *>(p); p += SW_MAXROWS * 2;
    S.joff = reinterpret_cast<unsigned short *>(p); p += SW_MAXROWS * 2;
    S.rb1 = reinterpret_cast<unsigned char *>(p); p += SW_MAXROWS;
    S.rb2 = reinterpret_cast<unsigned char *>(p); p += SW_MAXROWS;
    S.rfl = reinterpret_cast<unsigned char *>(p); p += SW_MAXROWS;
    S.jm = reinterpret_cast<unsigned char *>(p); p += SW_MAXROWS;
    S.gb = reinterpret_cast<signed char *>(p);
    S.clip = reinterpret_cast<float *>(S.cnt);
    static_assert((SW_MAXPH + 1) * 4 + 2 * 4 + 2 * SW_MAXROWS * 2 + 4 * SW_MAXROWS >= CLIP_WS * 64 * 4,
                  "the clipper columns fit in cnt..jm");
    return S;
}

// Prepare stage of world w: bodies, AABBs, pairs, contacts, rows, schedule.
// Returns the number of warm-start feature entries it wrote when features go
// to world-local slots (P.feat_local), else 0.
__device__ __forceinline__ int prepare_world(const DevWorlds &D, const StepParams &P, int w, const PrepLds &S) {
    const int t = threadIdx.x;
    float (*s_ab)[6] = S.ab;
    signed char *s_gb = S.gb;
    uint32_t *s_pairs = S.pairs;
    unsigned char *s_rb1 = S.rb1, *s_rb2 = S.rb2, *s_rfl = S.rfl, *s_jm = S.jm;
    unsigned short *s_lvl = S.lvl, *s_joff = S.joff;
    int *s_cnt = S.cnt;
    unsigned long long *s_mask = S.mask;
    int &s_L = *S.L, &s_nblk = *S.nblk;
    static_assert(SW_MAXPAIRS * sizeof(uint32_t) >= SW_MAXROWS * sizeof(unsigned short), "block starts alias s_pairs");
    static_assert(SW_MAXG * 6 * sizeof(float) >= SW_MAXROWS, "block lengths alias s_ab");
    int nfeat = 0;

    const int b0 = D.body_start[w], nb = D.body_start[w + 1] - b0;
    const int g0 = D.geom_start[w], ng = D.geom_start[w + 1] - g0;
    const int l0 = D.ball_start[w], nball = D.ball_start[w + 1] - l0;
    const int j0 = D.joint_start[w], jcap = D.joint_start[w + 1] - j0;
    const int r0 = D.row_start[w], rcap = min(D.row_start[w + 1] - r0, SW_MAXROWS);

    // stage timestamps (device wall clock): start, S1, S3, S4, S5, S6
    long long *pt = D.ptime ? D.ptime + (size_t)w * PTIME_SLOTS : nullptr;
    if (pt && t == 0) pt[0] = (long long)wall_clock64();

    // getInfo1 of the world's hinges and sliders before the body stage reads
    // the accumulators a motor at its limit pushes through
    if (D.jext && nball > 0) {
        for (int k = t; k < nball; k += 64) {
            const float *rec = D.balls + (size_t)(l0 + k) * JOINT_STRIDE;
            if ((int)rec[0] >= JT_HINGE) joint_info1(D, b0, rec, D.jext + (size_t)(l0 + k) * JEXT_STRIDE);
        }
        __syncthreads();
    }
    // S1: bodies
    for (int i = t; i < nb; i += 64) body_prep(D, P, b0 + i, D.facc, D.tacc);
    __syncthreads();
    if (pt && t == 0) pt[1] = (long long)wall_clock64();

    int njoint, ncont;
    if (P.collide) {
        for (int k = t; k < nball * JOINT_STRIDE; k += 64) D.joints[(size_t)j0 * JOINT_STRIDE + k] = D.balls[(size_t)l0 * JOINT_STRIDE + k];
        // S2: AABBs
        for (int i = t; i < ng; i += 64) {
            DGeom G;
            load_geom(D, g0 + i, b0, G);
            geom_aabb(G, s_ab[i]);
            s_gb[i] = (signed char)G.body;
        }
        __syncthreads();
        // S3: all-pairs broadphase in geom-index order (lower index first)
        const int T = ng * (ng - 1) / 2;
        int npairs = 0;
        for (int base = 0; base < T; base += 64) {
            const int k = base + t;
            bool ok = false;
            int gi = 0, gj = 0;
            if (k < T) {
                // row gi of the strict upper triangle holds ng-1-gi entries
                float fn = (float)(2 * ng - 1);
                gi = (int)floorf((fn - sqrtf(fn * fn - 8.f * (float)k)) * 0.5f);
                if (gi < 0) gi = 0;
                while (gi > 0 && gi * (2 * ng - gi - 1) / 2 > k) gi--;
                while ((gi + 1) * (2 * ng - gi - 2) / 2 <= k) gi++;
                gj = k - gi * (2 * ng - gi - 1) / 2 + gi + 1;
                const int bi = s_gb[gi], bj = s_gb[gj];
                bool bits = true;
                if (P.cat_filter) {  // collideAABBs' category test (collision_space_internal.h)
                    const unsigned ci = (unsigned)D.geom1[g0 + gi].w, li = (unsigned)D.gpos[g0 + gi].w;
                    const unsigned cj = (unsigned)D.geom1[g0 + gj].w, lj = (unsigned)D.gpos[g0 + gj].w;
                    bits = ((ci & lj) | (cj & li)) != 0u;
                }
                if ((bi >= 0 || bj >= 0) && !(bi == bj) && bits) {
                    const float *x = s_ab[gi], *y = s_ab[gj];
                    if (P.space == 0 && (x[1] == kInf || y[1] == kInf)) ok = true;
                    else ok = !(x[0] > y[1] || x[1] < y[0] || x[2] > y[3] || x[3] < y[2] || x[4] > y[5] || x[5] < y[4]);
                    if (ok && P.skip_connected && bi >= 0 && bj >= 0) {
                        for (int q = 0; q < nball; q++) {
                            const float *jr = D.balls + (size_t)(l0 + q) * JOINT_STRIDE;
                            int x1 = (int)jr[1], x2 = (int)jr[2];
                            if ((x1 == bi && x2 == bj) || (x1 == bj && x2 == bi)) ok = false;
                        }
                    }
                }
            }
            const unsigned long long bal = __ballot(ok);
            const int off = __popcll(bal & ((1ull << t) - 1ull));
            if (ok && npairs + off < SW_MAXPAIRS) s_pairs[npairs + off] = ((uint32_t)gi << 16) | (uint32_t)gj;
            npairs += __popcll(bal);
        }
        if (npairs > SW_MAXPAIRS) {
            if (t == 0) atomicAdd(&D.overflow[0], npairs - SW_MAXPAIRS);
            npairs = SW_MAXPAIRS;
        }
        __syncthreads();
        if (pt && t == 0) pt[2] = (long long)wall_clock64();
        // S4: narrowphase, contacts appended after the ball joints
        const int ccap = jcap - nball;
        ncont = 0;
        for (int base = 0; base < npairs; base += 64) {
            const int k = base + t;
            int n = 0, gi = 0, gj = 0, si = 0;
            DContact c[ODX_MAXC];
            ClipLds ws{S.clip + t};
            if (k < npairs) {
                gi = (int)(s_pairs[k] >> 16);
                gj = (int)(s_pairs[k] & 0xffff);
                DGeom A, B;
                load_geom(D, g0 + gi, b0, A);
                load_geom(D, g0 + gj, b0, B);
                int maxc = P.maxc;
                if (P.nmat > 0) {  // the material pair's surface and contact limit
                    si = (int)D.geom1[g0 + gi].z * P.nmat + (int)D.geom1[g0 + gj].z;
                    if (D.surf[si].max_contacts > 0) maxc = D.surf[si].max_contacts;
                }
                n = dcollide(A, B, min(maxc, ODX_MAXC), c, ws);
            }
            const int incl = wave_incl_scan(n);
            const int off = incl - n;
#pragma unroll
            for (int q = 0; q < ODX_MAXC; q++) {
                if (q >= n) break;
                const int slot = ncont + off + q;
                if (slot < ccap) {
                    int bb1 = s_gb[gi], bb2 = s_gb[gj];
                    float sg = 1.f;
                    if (bb1 < 0) { bb1 = bb2; bb2 = -1; sg = -1.f; }  // dJointAttach: reversed joint
                    float *o = D.joints + (size_t)(j0 + nball + slot) * JOINT_STRIDE;
                    float4 *o4 = reinterpret_cast<float4 *>(o);
                    o4[0] = make_float4(0.f, (float)bb1, (float)bb2, c[q].pos[0]);
                    o4[1] = make_float4(c[q].pos[1], c[q].pos[2], sg * c[q].normal[0], sg * c[q].normal[1]);
                    o4[2] = make_float4(sg * c[q].normal[2], c[q].depth, 0.f, 0.f);
                    o4[3] = make_float4(0.f, (float)si, (float)gi, (float)gj);
                }
            }
            ncont += wave_bcast_last(incl);
        }
        if (ncont > ccap) {
            if (t == 0) atomicAdd(&D.overflow[1], ncont - ccap);
            ncont = ccap;
        }
        njoint = nball + ncont;
        if (t == 0) { D.njoints[w] = njoint; D.ncontacts[w] = ncont; }
    } else {
        njoint = D.njoints[w];
        ncont = D.ncontacts[w];
    }
    __syncthreads();
    if (pt && t == 0) pt[3] = (long long)wall_clock64();

    // S5: rows, in joint order
    int mrow = 0;
    for (int base = 0; base < njoint; base += 64) {
        const int k = base + t;
        int mj = 0;
        const float *rec = D.joints + (size_t)(j0 + k) * JOINT_STRIDE;
        // ball joints and the extended types come first (k < nball); their
        // extended records are indexed by ball-joint index
        const int ty = k < njoint ? (int)rec[0] : 0;
        const float *xr = (ty >= JT_HINGE && D.jext) ? D.jext + (size_t)(l0 + k) * JEXT_STRIDE : nullptr;
        if (k < njoint)
            mj = ty == 0 ? contact_rows_m(D.surf[(int)rec[13]])
                         : ty == 1 ? 3 : ty == JT_FIXED ? 6 : 5 + (xr && xr[JX_LIMOT_ROW] != 0.f ? 1 : 0);
        const int incl = wave_incl_scan(mj);
        const int ri = mrow + incl - mj;
        if (k < njoint) {
            if (ri + mj <= rcap && k < SW_MAXROWS) {
                build_joint_rows(D, P, rec, xr, b0, r0, ri);
                const int bb1 = (int)rec[1], bb2 = (int)rec[2];
                s_joff[k] = (unsigned short)ri;
                s_jm[k] = (unsigned char)mj;
                for (int r = 0; r < mj; r++) {
                    s_rb1[ri + r] = (unsigned char)bb1;
                    s_rb2[ri + r] = (unsigned char)(bb2 < 0 ? 255 : bb2);
                }
                // findex rows of a contact are its friction rows (Approx1)
                if ((int)rec[0] == 0) {
                    const int mode = D.surf[(int)rec[13]].mode;
                    s_rfl[ri] = 0;
                    int rr = 1;
                    const Surface sf = D.surf[(int)rec[13]];
                    if (mj > 1) {
                        if (sf.mu > 0.f) { s_rfl[ri + rr] = (mode & 0x1000) ? 1 : 0; rr++; }
                        if (rr < mj) s_rfl[ri + rr] = (mode & 0x2000) ? 1 : 0;
                    }
                } else {
                    for (int r = 0; r < mj; r++) s_rfl[ri + r] = 0;
                }
                D.joff[j0 + k] = ri;
            } else {
                if (k < SW_MAXROWS) s_jm[k] = 0;
                D.joff[j0 + k] = 1 << 30;  // no rows (capacity exceeded)
            }
        }
        // warm start: matched previous impulses (mode 2) or features for the
        // learned predictor (mode 3); ball joints keep last step's impulses
        if (P.warm >= 2) {
            const bool fits = k < njoint && ri + mj <= rcap && k < SW_MAXROWS;
            const bool is_contact = fits && (int)rec[0] == 0;
            float lp[3] = {0.f, 0.f, 0.f};
            int matched = 0;
            if (is_contact && P.collide) matched = match_previous(D, j0, D.nprev[w], rec, b0, lp);
            if (fits && !is_contact)
                for (int r = 0; r < mj; r++) D.lambda0[r0 + ri + r] = D.lambda[r0 + ri + r];
            if (is_contact && P.warm == 2) write_contact_lambda0(D, rec, r0, ri, lp);
            if (P.warm == 3) {
                // the network predicts (normal, t1, t2): contacts with exactly those three rows;
                // a gated network leaves matched contacts at their previous impulses
                const bool learn = is_contact && mj == 3 && !(P.gate && matched);
                if (is_contact && !learn) write_contact_lambda0(D, rec, r0, ri, lp);
                const unsigned long long bal = __ballot(learn);
                int base = 0;
                int fbase;
                if (P.feat_local) {
                    // world-local slots from the world's first joint, in joint order
                    fbase = j0 + nfeat;
                    nfeat += __popcll(bal);
                } else {
                    if (t == 0 && bal) base = atomicAdd(D.feat_count + 2 * P.split, __popcll(bal));
                    fbase = P.feat_base + __shfl(base, 0, 64);
                }
                if (learn) {
                    const int slot = fbase + __popcll(bal & ((1ull << t) - 1ull));
                    contact_features(D, rec, b0, lp, matched, D.feat + (size_t)slot * FEAT_DIM);
                    D.feat_row[slot] = r0 + ri;
                    // rows beyond the predicted three (none for this surface) start at zero
                    write_contact_lambda0(D, rec, r0, ri, lp);
                }
            }
        }
        mrow += wave_bcast_last(incl);
    }
    int nj_used = njoint;
    if (mrow > rcap) {
        if (t == 0) atomicAdd(&D.overflow[2], mrow - rcap);
        // keep only joints whose rows fit
        mrow = 0;
        nj_used = 0;
        for (int k = 0; k < njoint && k < SW_MAXROWS; k++) {
            if (s_jm[k] == 0) break;
            mrow += s_jm[k];
            nj_used++;
        }
    }
    __syncthreads();
    if (pt && t == 0) pt[4] = (long long)wall_clock64();

    // S6: schedule.  Work items are packed (first row, row count, body 1,
    // body 2; pack_item) and address rows in "sweep order": row r of the world
    // sits at sweep position perm[r] (inverse: inv), the order in which the
    // solve kernel stages rows into LDS, so that every item's rows are
    // contiguous.  Order 0 keeps rows in place (perm = identity).
    int *ps = D.phase_start + r0 + w;
    int *sched = D.sched + r0;
    int *perm = D.perm + r0, *inv = D.inv + r0;
    if (P.order == 0) {
        // dependency levels over the ReorderPrep order (lane 0, LDS-resident)
        if (t == 0) {
            unsigned short *last = reinterpret_cast<unsigned short *>(s_mask);
            for (int b = 0; b < nb; b++) last[b] = 0;
            int L = 0;
            for (int pass = 0; pass < 2; pass++) {
                for (int r = 0; r < mrow; r++) {
                    if (s_rfl[r] != pass) continue;
                    const int a = s_rb1[r], bb = s_rb2[r];
                    int lv = last[a];
                    if (bb != 255 && last[bb] > lv) lv = last[bb];
                    lv += 1;
                    last[a] = (unsigned short)lv;
                    if (bb != 255) last[bb] = (unsigned short)lv;
                    s_lvl[r] = (unsigned short)(lv - 1);
                    if (lv > L) L = lv;
                }
            }
            s_L = L;
        }
        __syncthreads();
        const int L = s_L;
        for (int p = t; p <= L; p += 64) s_cnt[p] = 0;
        __syncthreads();
        for (int r = t; r < mrow; r += 64) atomicAdd(&s_cnt[s_lvl[r] + 1], 1);
        __syncthreads();
        if (t == 0) {
            for (int p = 1; p <= L; p++) s_cnt[p] += s_cnt[p - 1];
            for (int p = 0; p <= L; p++) ps[p] = s_cnt[p];
        }
        __syncthreads();
        for (int r = t; r < mrow; r += 64) {
            const int pos = atomicAdd(&s_cnt[s_lvl[r]], 1);
            sched[pos] = pack_item(r, 1, s_rb1[r], s_rb2[r] == 255 ? nb : s_rb2[r]);
            perm[r] = r;
            inv[r] = r;
        }
        if (t == 0) D.nphases[w] = L;
    } else {
        // Contact blocks: consecutive joints on the same body pair (the up to
        // four contacts of one geom pair) form one work item per pass, so a
        // sweep has a few long items instead of many short phases.  Two passes
        // like ReorderPrep (quickstep.cpp:2329-2355): each block's non-findex
        // rows (normals, ball rows) in colour phases first, then its findex
        // (friction-cone) rows in a second set of colour phases; interleaving
        // a contact's friction rows with its normal row instead can settle into
        // a limit cycle short of the residual target.
        unsigned short *s_bstart = reinterpret_cast<unsigned short *>(s_pairs);  // free after S4
        unsigned char *s_blen = reinterpret_cast<unsigned char *>(&s_ab[0][0]);
        // per joint: its rows of the first pass (na) and of the findex pass
        // (nf), packed na | nf << 4 in place of its row count
        for (int k = t; k < nj_used; k += 64) {
            int na, nf;
            split_joint_rows(s_rfl + s_joff[k], (int)s_jm[k], na, nf);
            s_jm[k] = (unsigned char)(na | (nf << 4));
        }
        __syncthreads();
        // Blocks: one walk over the joints in order, the same on every lane
        // (scalar registers); each chunk of 64 joints is read into a register
        // first and walked with readlane, so no LDS read sits on the walk.
        int nblk = 0, pa = -2, pb = -2, acc_a = 0, acc_f = 0, cur_len = 0;
        for (int base = 0; base < nj_used; base += 64) {
            const int k = base + t;
            int pk = 0;
            if (k < nj_used) {
                const int r = s_joff[k];
                pk = (int)s_rb1[r] | ((int)s_rb2[r] << 8) | ((int)s_jm[k] << 16);
            }
            const int n = min(64, nj_used - base);
            for (int j = 0; j < n; j++) {
                const int v = __builtin_amdgcn_readlane(pk, j);
                const int a = v & 255, bb = (v >> 8) & 255, na = (v >> 16) & 15, nf = (v >> 20) & 15;
                if (nblk > 0 && a == pa && bb == pb && acc_a + na <= ITEM_MAXROWS && acc_f + nf <= ITEM_MAXROWS) {
                    cur_len++;
                    acc_a += na;
                    acc_f += nf;
                } else {
                    if (t == 0) {
                        if (nblk > 0) s_blen[nblk - 1] = (unsigned char)cur_len;
                        s_bstart[nblk] = (unsigned short)(base + j);
                    }
                    nblk++;
                    cur_len = 1;
                    pa = a; pb = bb; acc_a = na; acc_f = nf;
                }
            }
        }
        if (t == 0 && nblk > 0) s_blen[nblk - 1] = (unsigned char)cur_len;
        __syncthreads();
        // Greedy block colouring: block q takes the lowest colour neither of
        // its bodies holds.  Each body's 64-bit colour mask lives in lane b
        // (nb <= 64); the walk reads it with readlane and updates it by lane select.
        unsigned mlo = 0u, mhi = 0u;
        int L = 0, extra = 0;
        for (int base = 0; base < nblk; base += 64) {
            const int q = base + t;
            int ab = 0;
            if (q < nblk) {
                const int r = s_joff[s_bstart[q]];
                ab = (int)s_rb1[r] | ((int)s_rb2[r] << 8);
            }
            int lvl = 0;
            const int n = min(64, nblk - base);
            for (int j = 0; j < n; j++) {
                const int v = __builtin_amdgcn_readlane(ab, j);
                const int a = v & 255, bb = (v >> 8) & 255;
                const unsigned alo = __builtin_amdgcn_readlane(mlo, a), ahi = __builtin_amdgcn_readlane(mhi, a);
                unsigned ulo = alo, uhi = ahi, blo = 0u, bhi = 0u;
                if (bb != 255) {
                    blo = __builtin_amdgcn_readlane(mlo, bb);
                    bhi = __builtin_amdgcn_readlane(mhi, bb);
                    ulo |= blo;
                    uhi |= bhi;
                }
                int c;
                if (~ulo) {
                    c = __ffs((int)~ulo) - 1;
                    const unsigned bit = 1u << c;
                    mlo = t == a ? (alo | bit) : mlo;
                    mlo = t == bb ? (blo | bit) : mlo;
                } else if (~uhi) {
                    c = 32 + __ffs((int)~uhi) - 1;
                    const unsigned bit = 1u << (c - 32);
                    mhi = t == a ? (ahi | bit) : mhi;
                    mhi = t == bb ? (bhi | bit) : mhi;
                } else {
                    c = 64 + extra++;  // no free colour: a phase of its own
                }
                lvl = t == j ? c : lvl;
                if (c + 1 > L) L = c + 1;
            }
            if (q < nblk) s_lvl[q] = (unsigned short)lvl;
        }
        __syncthreads();
        // Emit the items colour by colour (pass A then pass B); lanes take
        // blocks in order, a wave scan gives each item its slot and rows.
        int nitem = 0, npos = 0, nph = 0;
        for (int pass = 0; pass < 2; pass++) {
            for (int c = 0; c < L; c++) {
                int phase_items = 0;
                for (int base = 0; base < nblk; base += 64) {
                    const int q = base + t;
                    int nrow = 0;
                    if (q < nblk && s_lvl[q] == c) {
                        const int k0 = s_bstart[q], k1 = k0 + s_blen[q];
                        for (int k = k0; k < k1; k++) nrow += pass == 0 ? (s_jm[k] & 15) : (s_jm[k] >> 4);
                    }
                    const unsigned long long bal = __ballot(nrow > 0);
                    const int incl = wave_incl_scan(nrow);
                    if (nrow > 0) {
                        const int item = nitem + __popcll(bal & ((1ull << t) - 1ull));
                        const int start = npos + incl - nrow;
                        const int r = s_joff[s_bstart[q]];
                        sched[item] = pack_item(start, nrow, s_rb1[r], s_rb2[r] == 255 ? nb : s_rb2[r]);
                        int p = start;
                        const int k0 = s_bstart[q], k1 = k0 + s_blen[q];
                        for (int k = k0; k < k1; k++) {
                            const int na = s_jm[k] & 15, nf = s_jm[k] >> 4;
                            const int rr = s_joff[k] + (pass == 0 ? 0 : na), n = pass == 0 ? na : nf;
                            for (int i = 0; i < n; i++, p++) { perm[rr + i] = p; inv[p] = rr + i; }
                        }
                    }
                    phase_items += __popcll(bal);
                    nitem += __popcll(bal);
                    npos += wave_bcast_last(incl);
                }
                if (phase_items > 0 && t == 0) ps[nph] = nitem - phase_items;
                if (phase_items > 0) nph++;
            }
        }
        if (t == 0) {
            ps[nph] = nitem;
            D.nphases[w] = nph;
        }
    }
    if (t == 0) D.nrows[w] = mrow;
    if (t == 0 && P.feat_local) D.nfeat[w] = nfeat;
    if (pt && t == 0) pt[5] = (long long)wall_clock64();
    return nfeat;
}

__global__ __launch_bounds__(64) void k_prepare_small(DevWorlds D, StepParams P) {
    __shared__ __align__(16) char s_prep[PREP_LDS_BYTES];
    prepare_world(D, P, P.wbase + (int)blockIdx.x, prep_lds_carve(s_prep));
}

// ============================================================== kernel C ===
__device__ __forceinline__ void process_item(const DevWorlds &D, int r0, int item, float *s_fc) {
#pragma clang fp contract(fast)
    const int rs = item >> 3, nr = item & 7;
    const float4 *row4 = reinterpret_cast<const float4 *>(D.rows + (size_t)(r0 + rs) * ROW_STRIDE);
    const float4 ids = row4[7];
    const int b1 = as_int(ids.x), b2 = as_int(ids.y);
    float f1[6], f2[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < 6; k++) f1[k] = s_fc[b1 * 6 + k];
    if (b2 >= 0) {
#pragma unroll
        for (int k = 0; k < 6; k++) f2[k] = s_fc[b2 * 6 + k];
    }
    float *lam = D.lambda + r0;
    float lam_first = 0.f;
    for (int q = 0; q < nr; q++) {
        const float4 *rr = row4 + q * 8;
        const float4 a0 = rr[0], a1 = rr[1], a2 = rr[2], a3 = rr[3], a4 = rr[4], a5 = rr[5], a6 = rr[6], a7 = rr[7];
        const float old = lam[rs + q];
        float delta = a1.z - old * a1.w;
        delta -= f1[0] * a0.x + f1[1] * a0.y + f1[2] * a0.z + f1[3] * a0.w + f1[4] * a1.x + f1[5] * a1.y;
        if (b2 >= 0) delta -= f2[0] * a2.x + f2[1] * a2.y + f2[2] * a2.z + f2[3] * a2.w + f2[4] * a3.x + f2[5] * a3.y;
        float lo = a3.z, hi = a3.w;
        const int fidx = as_int(a7.z);
        if (fidx >= 0) {
            const float lf = (fidx == rs) ? lam_first : lam[fidx];
            hi = fabsf(hi * lf);
            lo = -hi;
        }
        float nl = old + delta;
        if (nl < lo) { delta = lo - old; nl = lo; }
        else if (nl > hi) { delta = hi - old; nl = hi; }
        lam[rs + q] = nl;
        if (q == 0) lam_first = nl;
        f1[0] += delta * a4.x; f1[1] += delta * a4.y; f1[2] += delta * a4.z;
        f1[3] += delta * a4.w; f1[4] += delta * a5.x; f1[5] += delta * a5.y;
        if (b2 >= 0) {
            f2[0] += delta * a5.z; f2[1] += delta * a5.w; f2[2] += delta * a6.x;
            f2[3] += delta * a6.y; f2[4] += delta * a6.z; f2[5] += delta * a6.w;
        }
    }
#pragma unroll
    for (int k = 0; k < 6; k++) s_fc[b1 * 6 + k] = f1[k];
    if (b2 >= 0) {
#pragma unroll
        for (int k = 0; k < 6; k++) s_fc[b2 * 6 + k] = f2[k];
    }
}

// Batched solve, rows streamed from global memory: one packed item (its rows
// in sweep order, mapped back to rows through inv).
__device__ __forceinline__ void process_item_packed(const DevWorlds &D, int r0, const int *inv, int nb, int item,
                                                    float *s_fc) {
#pragma clang fp contract(fast)
    const int rs = item_start(item), nr = item_rows(item);
    const int b1 = item_b1(item), b2 = item_b2(item) == nb ? -1 : item_b2(item);
    float f1[6], f2[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < 6; k++) f1[k] = s_fc[b1 * 6 + k];
    if (b2 >= 0) {
#pragma unroll
        for (int k = 0; k < 6; k++) f2[k] = s_fc[b2 * 6 + k];
    }
    float *lam = D.lambda + r0;
    for (int q = 0; q < nr; q++) {
        const int g = inv[rs + q];
        const float4 *rr = reinterpret_cast<const float4 *>(D.rows + (size_t)(r0 + g) * ROW_STRIDE);
        const float4 a0 = rr[0], a1 = rr[1], a2 = rr[2], a3 = rr[3], a4 = rr[4], a5 = rr[5], a6 = rr[6], a7 = rr[7];
        const float old = lam[g];
        float delta = a1.z - old * a1.w;
        delta -= f1[0] * a0.x + f1[1] * a0.y + f1[2] * a0.z + f1[3] * a0.w + f1[4] * a1.x + f1[5] * a1.y;
        if (b2 >= 0) delta -= f2[0] * a2.x + f2[1] * a2.y + f2[2] * a2.z + f2[3] * a2.w + f2[4] * a3.x + f2[5] * a3.y;
        float lo = a3.z, hi = a3.w;
        const int fidx = as_int(a7.z);
        if (fidx >= 0) {
            hi = fabsf(hi * lam[fidx]);
            lo = -hi;
        }
        float nl = old + delta;
        if (nl < lo) { delta = lo - old; nl = lo; }
        else if (nl > hi) { delta = hi - old; nl = hi; }
        lam[g] = nl;
        f1[0] += delta * a4.x; f1[1] += delta * a4.y; f1[2] += delta * a4.z;
        f1[3] += delta * a4.w; f1[4] += delta * a5.x; f1[5] += delta * a5.y;
        if (b2 >= 0) {
            f2[0] += delta * a5.z; f2[1] += delta * a5.w; f2[2] += delta * a6.x;
            f2[3] += delta * a6.y; f2[4] += delta * a6.z; f2[5] += delta * a6.w;
        }
    }
#pragma unroll
    for (int k = 0; k < 6; k++) s_fc[b1 * 6 + k] = f1[k];
    if (b2 >= 0) {
#pragma unroll
        for (int k = 0; k < 6; k++) s_fc[b2 * 6 + k] = f2[k];
    }
}

// Residual of the current iterate: the un-relaxed projected correction of
// each row relative to max |lambda| (SOR's Ad already carries the factor w).
__device__ float world_residual(const DevWorlds &D, const StepParams &P, int r0, int m, const float *s_fc) {
    float mx = 0.f, lmax = 0.f;
    const float *lam = D.lambda + r0;
    const float iw = 1.f / P.sor_w;
    for (int r = lane_id(); r < m; r += 64) {
        const float4 *rr = reinterpret_cast<const float4 *>(D.rows + (size_t)(r0 + r) * ROW_STRIDE);
        const float4 a0 = rr[0], a1 = rr[1], a2 = rr[2], a3 = rr[3], a7 = rr[7];
        const int b1 = as_int(a7.x), b2 = as_int(a7.y), fidx = as_int(a7.z);
        const float *f1 = s_fc + b1 * 6;
        const float l = lam[r];
        float d = a1.z - l * a1.w;
        d -= f1[0] * a0.x + f1[1] * a0.y + f1[2] * a0.z + f1[3] * a0.w + f1[4] * a1.x + f1[5] * a1.y;
        if (b2 >= 0) {
            const float *f2 = s_fc + b2 * 6;
            d -= f2[0] * a2.x + f2[1] * a2.y + f2[2] * a2.z + f2[3] * a2.w + f2[4] * a3.x + f2[5] * a3.y;
        }
        d *= iw;
        float lo = a3.z, hi = a3.w;
        if (fidx >= 0) { hi = fabsf(hi * lam[fidx]); lo = -hi; }
        float nl = fminf(fmaxf(l + d, lo), hi);
        mx = fmaxf(mx, fabsf(nl - l));
        lmax = fmaxf(lmax, fabsf(l));
    }
    mx = wave_max(mx);
    lmax = wave_max(lmax);
    return mx / fmaxf(lmax, P.lambda_floor);
}

// ---- LDS-resident sweeps -------------------------------------------------
// A world whose rows fit the launch's LDS budget is solved from LDS: each
// sweep then costs LDS latency instead of a round trip to HBM per row.
// Row (16 floats): 0-2 J1 linear, 3-5 J1 angular, 6-8 J2 angular, 9 rhs,
// 10 cfm (J, rhs and cfm Ad-scaled as in the global row), 11 1/Ad,
// 12 lo, 13 hi, 14 ids = b1 | b2 << 8 | (findex + 1) << 16, 15 unused.
// J2 linear is -J1 linear for every contact and ball row (checked when the
// rows are staged).  The body-side terms of iMJ = M^-1 J^T are formed from
// the body record, iMJ = (1/Ad) [invMass J_lin, invI_w J_ang] with J
// Ad-scaled, so a row needs 64 B instead of 128 B.
// Body (16 floats): 0-5 fc, 6 invMass, 7-15 invI_w (3x3 row-major).  A row
// without a second body points at the zero body at index nb.
constexpr int LDS_ROW = 16, LDS_BODY = 16;
// Offset (floats) of the trash words, one per lane for the masked row
// updates, at the end of the solve's dynamic LDS (solve_lds_floats).
__host__ __device__ __forceinline__ int solve_lds_trash_offset(const StepParams &P) {
    return solve_lds_floats(P.lds_bodies, P.lds_rows) - 64;
}

// 64-byte records start on one of only two 16-bank boundaries of LDS, so the
// lanes of a 128-bit read that touch different records would pile onto the
// same banks.  The four 16-byte slots of record r are therefore stored in the
// order j ^ ((r >> 1) & 3): records r .. r + 7 then read slot j from eight
// different bank groups.
__device__ __forceinline__ int sw4(int r, int j) { return r * 4 + (j ^ ((r >> 1) & 3)); }
__device__ __forceinline__ float4 ld4(const float *base, int r, int j) {
    return reinterpret_cast<const float4 *>(base)[sw4(r, j)];
}
__device__ __forceinline__ void st4(float *base, int r, int j, float4 v) {
    reinterpret_cast<float4 *>(base)[sw4(r, j)] = v;
}
// element k (0..15) of record r
__device__ __forceinline__ float *rec_elem(float *base, int r, int k) { return base + 4 * sw4(r, k >> 2) + (k & 3); }

// Body record in LDS (16 floats): slot 0 = fc linear xyz, 1/m; slot 1 = fc
// angular xyz, invI[0][0]; slots 2-3 = invI[0][1] .. invI[2][2] (world frame,
// row-major).
__device__ __forceinline__ void lds_fc(const float *s_body, int b, float *fc) {
    const float4 a = ld4(s_body, b, 0), c = ld4(s_body, b, 1);
    fc[0] = a.x; fc[1] = a.y; fc[2] = a.z; fc[3] = c.x; fc[4] = c.y; fc[5] = c.z;
}

__device__ __forceinline__ void lds_body(const float *s_body, int b, float *fc, float &im, float *I) {
    const float4 a = ld4(s_body, b, 0), c = ld4(s_body, b, 1), d = ld4(s_body, b, 2), e = ld4(s_body, b, 3);
    fc[0] = a.x; fc[1] = a.y; fc[2] = a.z; fc[3] = c.x; fc[4] = c.y; fc[5] = c.z;
    im = a.w;
    I[0] = c.w; I[1] = d.x; I[2] = d.y; I[3] = d.z; I[4] = d.w; I[5] = e.x; I[6] = e.y; I[7] = e.z; I[8] = e.w;
}

// fc component k (0..5) of body b
__device__ __forceinline__ float *fc_elem(float *s_body, int b, int k) { return rec_elem(s_body, b, k < 3 ? k : k + 1); }

// x, y, z into the first three floats of slot `slot` of body b
__device__ __forceinline__ void lds_store3(float *s_body, int b, int slot, float x, float y, float z) {
    float *o = s_body + 4 * sw4(b, slot);
    o[0] = x; o[1] = y; o[2] = z;
}

__device__ __forceinline__ void lds_store_fc(float *s_body, int b, const float *fc) {
    lds_store3(s_body, b, 0, fc[0], fc[1], fc[2]);
    lds_store3(s_body, b, 1, fc[3], fc[4], fc[5]);
}

// Row record in LDS (16 floats): slot 0 = J1 linear xyz, rhs; slot 1 = J1
// angular xyz, cfm; slot 2 = J2 angular xyz, 0; slot 3 = lo, hi, ids (body 1
// bits 0-7, body 2 bits 8-15, sweep position of the findex row + 1 from bit
// 16), 1/Ad.  J2 linear is -J1 linear.  rhs, cfm and J are scaled by Ad as the
// reference's SOR loop takes them, so a row's step is rhs - cfm lambda - J.fc.

// Quad-cooperative sweeps.  The four lanes of a quad share one item (one body
// pair) and split each row update of quickstep.cpp:2917-3033 (IterationStep)
// by the blocks of J: lane q of the quad holds one 3-vector x of the pair's
// fc, the matching 3x3 block of M^-1 and reads the J block that multiplies it:
//   q = 0  body 1 linear   (slot 0, J1 linear; M^-1 block 1/m1)
//   q = 1  body 1 angular  (slot 1, J1 angular; invI1)
//   q = 2  body 2 angular  (slot 2, J2 angular; invI2)
//   q = 3  body 2 linear, held negated (slot 0: J1 linear = -J2 linear; 1/m2)
// The row's step rhs - cfm lambda - J.fc is the sum of the four lanes' parts
// (lane 0 adds rhs, lane 1 subtracts cfm lambda), formed with two DPP adds
// inside the quad; every lane of the quad then takes the same clamp and the
// same delta, and adds delta/Ad (M^-1 J)_q to its own 3-vector.
struct QuadRow {
    float4 j;   // this lane's J block and its scalar (rhs, cfm, 0 or rhs)
    float4 t;   // lo, hi, ids, 1/Ad
    float old, lf;
};

__device__ __forceinline__ float quad_sum(float v) {
    v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));  // lanes 1,0,3,2
    v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, false));  // lanes 2,3,0,1
    return v;
}

__device__ __forceinline__ void quad_load_row(const float *s_rows, int r, int js, QuadRow &x) {
    x.j = ld4(s_rows, r, js);
    x.t = ld4(s_rows, r, 3);
}
__device__ __forceinline__ void quad_load_lam(const float *s_lam, int r, QuadRow &x) {
    const int fi = (as_int(x.t.z) >> 16) - 1;
    x.old = s_lam[r];
    x.lf = s_lam[fi >= 0 ? fi : 0];
}

// One row step of the quad (all four lanes active).  k1 / k2 select the
// lane's share of rhs - cfm lambda.  Returns the new lambda (the same on the
// quad's four lanes); live = false leaves x as it is.
__device__ __forceinline__ float quad_row_step(const QuadRow &r, float k1, float k2, float *x, const float *M,
                                               bool live) {
#pragma clang fp contract(fast)
    const float d = r.j.x * x[0] + r.j.y * x[1] + r.j.z * x[2];
    float delta = quad_sum(r.j.w * (k2 * r.old + k1) - d);
    const int fidx = (as_int(r.t.z) >> 16) - 1;
    const float hf = fabsf(r.t.y * r.lf);
    const float hi = fidx >= 0 ? hf : r.t.y, lo = fidx >= 0 ? -hf : r.t.x;
    const float raw = r.old + delta;
    const float nl = fminf(fmaxf(raw, lo), hi);
    if (nl != raw) delta = nl - r.old;  // clamped: delta to the bound, as the reference
    const float e = live ? delta * r.t.w : 0.f;
    x[0] += e * (M[0] * r.j.x + M[1] * r.j.y + M[2] * r.j.z);
    x[1] += e * (M[3] * r.j.x + M[4] * r.j.y + M[5] * r.j.z);
    x[2] += e * (M[6] * r.j.x + M[7] * r.j.y + M[8] * r.j.z);
    return nl;
}

// One stage of the item pipeline: updates row i (in cur), reads the record of
// row i + 2 into far and the lambdas of row i + 1 (whose record is in nxt).
// Indices past the item's end re-read its last row.  A row's lambda is
// written by this quad only and a friction row's normal lambda was settled in
// an earlier phase (normals and friction rows sweep in separate passes), so
// none of these reads depends on the item's own writes.
__device__ __forceinline__ void quad_stage(const float *s_rows, float *s_lam, int rs, int nr, int last, int i, int js,
                                           float k1, float k2, QuadRow &cur, QuadRow &nxt, QuadRow &far, float *x,
                                           const float *M, float *trash) {
    quad_load_row(s_rows, min(rs + i + 2, last), js, far);
    quad_load_lam(s_lam, min(rs + i + 1, last), nxt);
    const bool live = i < nr;
    const float nl = quad_row_step(cur, k1, k2, x, M, live);
    *(live ? s_lam + rs + i : trash) = nl;
}

// The rows of one item (one body pair, nr <= ITEM_MAXROWS rows in sweep
// order; item < 0: none) on a quad, for a wave whose longest item has nmax
// rows (wave-uniform).  Row i + 2's record and row i + 1's lambdas are read
// while row i updates, in three register sets used in turn.  Every lane
// issues the same LDS operations in the same order (lanes past their item's
// end update nothing and write their lambda to the lane's trash word).
__device__ __forceinline__ void process_item_quad(const float *s_rows, float *s_lam, float *s_body, int nb, int item,
                                                  int nmax, float *trash) {
#pragma clang fp contract(fast)
    const int q = lane_id() & 3;
    const int js = q == 3 ? 0 : q;
    const float k1 = q == 0 ? 1.f : 0.f, k2 = q == 1 ? -1.f : 0.f;
    const bool lin = q == 0 || q == 3;
    const float sg = q == 3 ? -1.f : 1.f;
    const bool has = item >= 0;
    const int rs = has ? item_start(item) : 0, nr = has ? item_rows(item) : 0;
    const int b = has ? (q < 2 ? item_b1(item) : item_b2(item)) : nb;
    const int last = rs + max(nr - 1, 0);
    QuadRow A, B, C;
    quad_load_row(s_rows, rs, js, A);
    quad_load_row(s_rows, min(rs + 1, last), js, B);
    float x[3], M[9];
    {
        const float4 c0 = ld4(s_body, b, 0), c1 = ld4(s_body, b, 1), c2 = ld4(s_body, b, 2), c3 = ld4(s_body, b, 3);
        x[0] = sg * (lin ? c0.x : c1.x);
        x[1] = sg * (lin ? c0.y : c1.y);
        x[2] = sg * (lin ? c0.z : c1.z);
        M[0] = lin ? c0.w : c1.w; M[1] = lin ? 0.f : c2.x; M[2] = lin ? 0.f : c2.y;
        M[3] = lin ? 0.f : c2.z; M[4] = lin ? c0.w : c2.w; M[5] = lin ? 0.f : c3.x;
        M[6] = lin ? 0.f : c3.y; M[7] = lin ? 0.f : c3.z; M[8] = lin ? c0.w : c3.w;
    }
    quad_load_lam(s_lam, rs, A);
    for (int i = 0; i < nmax; i += 3) {
        quad_stage(s_rows, s_lam, rs, nr, last, i, js, k1, k2, A, B, C, x, M, trash);
        if (i + 1 >= nmax) break;
        quad_stage(s_rows, s_lam, rs, nr, last, i + 1, js, k1, k2, B, C, A, x, M, trash);
        if (i + 2 >= nmax) break;
        quad_stage(s_rows, s_lam, rs, nr, last, i + 2, js, k1, k2, C, A, B, x, M, trash);
    }
    if (has && b != nb) lds_store3(s_body, b, lin ? 0 : 1, sg * x[0], sg * x[1], sg * x[2]);
}

__device__ float world_residual_lds(const float *s_rows, const float *s_lam, const float *s_body, int m,
                                    const StepParams &P) {
    float mx = 0.f, lmax = 0.f;
    const float iw = 1.f / P.sor_w;
    for (int r = lane_id(); r < m; r += 64) {
        const float4 a0 = ld4(s_rows, r, 0), a1 = ld4(s_rows, r, 1), a2 = ld4(s_rows, r, 2), a3 = ld4(s_rows, r, 3);
        const int ids = as_int(a3.z);
        float f1[6], f2[6];
        lds_fc(s_body, ids & 255, f1);
        lds_fc(s_body, (ids >> 8) & 255, f2);
        const int fidx = (ids >> 16) - 1;
        const float l = s_lam[r];
        float d = a0.w - l * a1.w;
        d -= a0.x * (f1[0] - f2[0]) + a0.y * (f1[1] - f2[1]) + a0.z * (f1[2] - f2[2]);
        d -= a1.x * f1[3] + a1.y * f1[4] + a1.z * f1[5];
        d -= a2.x * f2[3] + a2.y * f2[4] + a2.z * f2[5];
        d *= iw;
        float lo = a3.x, hi = a3.y;
        if (fidx >= 0) { hi = fabsf(hi * s_lam[fidx]); lo = -hi; }
        const float nl = fminf(fmaxf(l + d, lo), hi);
        mx = fmaxf(mx, fabsf(nl - l));
        lmax = fmaxf(lmax, fabsf(l));
    }
    mx = wave_max(mx);
    lmax = wave_max(lmax);
    return mx / fmaxf(lmax, P.lambda_floor);
}

// fc += lambda iMJ for the rows' current lambda (warm start), LDS atomics.
__device__ void lds_accumulate_fc(const float *s_rows, const float *s_lam, float *s_body, int nb, int m) {
#pragma clang fp contract(fast)
    for (int r = lane_id(); r < m; r += 64) {
        const float l = s_lam[r];
        if (l == 0.f) continue;
        const float4 a0 = ld4(s_rows, r, 0), a1 = ld4(s_rows, r, 1), a2 = ld4(s_rows, r, 2), a3 = ld4(s_rows, r, 3);
        const int ids = as_int(a3.z);
        const int b1 = ids & 255, b2 = (ids >> 8) & 255;
        float f[6], I[9], im;
        const float e = l * a3.w;
        lds_body(s_body, b1, f, im, I);
        atomicAdd(fc_elem(s_body, b1, 0), e * im * a0.x);
        atomicAdd(fc_elem(s_body, b1, 1), e * im * a0.y);
        atomicAdd(fc_elem(s_body, b1, 2), e * im * a0.z);
        atomicAdd(fc_elem(s_body, b1, 3), e * (I[0] * a1.x + I[1] * a1.y + I[2] * a1.z));
        atomicAdd(fc_elem(s_body, b1, 4), e * (I[3] * a1.x + I[4] * a1.y + I[5] * a1.z));
        atomicAdd(fc_elem(s_body, b1, 5), e * (I[6] * a1.x + I[7] * a1.y + I[8] * a1.z));
        if (b2 != nb) {
            lds_body(s_body, b2, f, im, I);
            atomicAdd(fc_elem(s_body, b2, 0), -e * im * a0.x);
            atomicAdd(fc_elem(s_body, b2, 1), -e * im * a0.y);
            atomicAdd(fc_elem(s_body, b2, 2), -e * im * a0.z);
            atomicAdd(fc_elem(s_body, b2, 3), e * (I[0] * a2.x + I[1] * a2.y + I[2] * a2.z));
            atomicAdd(fc_elem(s_body, b2, 4), e * (I[3] * a2.x + I[4] * a2.y + I[5] * a2.z));
            atomicAdd(fc_elem(s_body, b2, 5), e * (I[6] * a2.x + I[7] * a2.y + I[8] * a2.z));
        }
    }
}

// Stages the world's bodies, rows (each at its sweep position perm[r], with
// findex mapped the same way) and sweep schedule into LDS.  Returns false
// (wave-uniform) when a row breaks the J2 linear = -J1 linear form; the
// caller then solves from global memory.
__device__ bool lds_stage(const DevWorlds &D, int b0, int nb, int r0, int m, int L, const int *ps,
                          const int *sched, const int *perm, float *s_body, float *s_rows, int *s_sched,
                          unsigned short *s_ps) {
    const int t = lane_id();
    // phase starts (bits 0-10) with the phase's longest item in rows (bits 11-14)
    for (int p = t; p <= L; p += 64) {
        int nmax = 0;
        if (p < L)
            for (int k = ps[p]; k < ps[p + 1]; k++) nmax = max(nmax, item_rows(sched[k]));
        s_ps[p] = (unsigned short)(ps[p] | (nmax << 11));
    }
    const int nitems = ps[L];
    for (int k = t; k < nitems; k += 64) s_sched[k] = sched[k];
    for (int i = t; i <= nb; i += 64) {
        const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
        if (i < nb) {
            const float *s = D.bscr + (size_t)(b0 + i) * BSCR_STRIDE;
            st4(s_body, i, 0, make_float4(0.f, 0.f, 0.f, s[30]));
            st4(s_body, i, 1, make_float4(0.f, 0.f, 0.f, s[12]));
            st4(s_body, i, 2, make_float4(s[13], s[14], s[16], s[17]));
            st4(s_body, i, 3, make_float4(s[18], s[20], s[21], s[22]));
        } else {
            for (int j = 0; j < 4; j++) st4(s_body, i, j, z);
        }
    }
    int bad = 0;
    for (int r = t; r < m; r += 64) {
        const float4 *g = reinterpret_cast<const float4 *>(D.rows + (size_t)(r0 + r) * ROW_STRIDE);
        const float4 a0 = g[0], a1 = g[1], a2 = g[2], a3 = g[3], a7 = g[7];
        const int b1 = as_int(a7.x), b2 = as_int(a7.y), fidx = as_int(a7.z);
        if (b2 >= 0 && (a2.x != -a0.x || a2.y != -a0.y || a2.z != -a0.z)) bad = 1;
        const int ids = b1 | ((b2 >= 0 ? b2 : nb) << 8) | (((fidx >= 0 ? perm[fidx] : -1) + 1) << 16);
        const int p = perm[r];
        st4(s_rows, p, 0, make_float4(a0.x, a0.y, a0.z, a1.z));
        st4(s_rows, p, 1, make_float4(a0.w, a1.x, a1.y, a1.w));
        st4(s_rows, p, 2, make_float4(a2.w, a3.x, a3.y, 0.f));
        st4(s_rows, p, 3, make_float4(a3.z, a3.w, as_float(ids), 1.f / a7.w));
    }
    return !__any(bad);
}

// Dispatch order of the solve: worlds by descending sweep count of their
// previous solve (a counting sort over 1024 buckets of 4 sweeps), so the
// longest solves start first and short ones fill in behind them.  Only the
// assignment of worlds to workgroups changes, never a world's arithmetic.
constexpr int ORDER_BUCKETS = 1024;
__global__ __launch_bounds__(1024) void k_order_worlds(const int *niters, int wbase, int nworlds, int *order) {
    __shared__ int s_h[ORDER_BUCKETS + 1];
    const int t = threadIdx.x;
    for (int i = t; i <= ORDER_BUCKETS; i += blockDim.x) s_h[i] = 0;
    __syncthreads();
    for (int w = t; w < nworlds; w += blockDim.x) {
        const int key = ORDER_BUCKETS - 1 - min(max(niters[wbase + w], 0) >> 2, ORDER_BUCKETS - 1);
        atomicAdd(&s_h[key + 1], 1);
    }
    __syncthreads();
    if (t == 0)
        for (int i = 1; i <= ORDER_BUCKETS; i++) s_h[i] += s_h[i - 1];
    __syncthreads();
    for (int w = t; w < nworlds; w += blockDim.x) {
        const int key = ORDER_BUCKETS - 1 - min(max(niters[wbase + w], 0) >> 2, ORDER_BUCKETS - 1);
        order[wbase + atomicAdd(&s_h[key], 1)] = wbase + w;
    }
}

// Solve stage of world w (sweeps, residual stop, warm-start records,
// integration); s_dyn: the launch's dynamic LDS (solve_small_lds_bytes).
// Stage4b (v += h fc), Stage6a (v += h M^-1 f), Stage6b + dxStepBody
// (angular speed cap, x += h v, q += h dq(w, q), normalised, damping;
// f = t = 0) of body b, quickstep.cpp:3082-3102, 3299-3337,
// util.cpp:583-699 (infinitesimal rotation).
// fin / tin: the total force and torque body_prep wrote.
__device__ __forceinline__ void integrate_body(const DevWorlds &D, const StepParams &P, int b, const float *fc,
                                               bool rows, const float4 *fin, const float4 *tin) {
    const float h = P.h;
    const float *s = D.bscr + (size_t)b * BSCR_STRIDE;
    float4 p4 = D.pos[b], q4 = D.quat[b], lv = D.lvel[b], av = D.avel[b], fa = fin[b], ta = tin[b];
    if (rows) {
        lv.x += h * fc[0]; lv.y += h * fc[1]; lv.z += h * fc[2];
        av.x += h * fc[3]; av.y += h * fc[4]; av.z += h * fc[5];
    }
    const float ims = h * s[30];
    lv.x += ims * fa.x; lv.y += ims * fa.y; lv.z += ims * fa.z;
    float tq[3] = {ta.x * h, ta.y * h, ta.z * h}, dw[3];
    mul0_331(dw, s + 12, tq);
    av.x += dw[0]; av.y += dw[1]; av.z += dw[2];
    float4 d0 = make_float4(0.f, 0.f, 0.f, 0.f), d1 = d0;
    int dfl = 0;
    if (D.damp) {
        d0 = D.damp[2 * b];
        d1 = D.damp[2 * b + 1];
        dfl = (int)d1.y;
    }
    if (dfl & DAMP_MAX_ANGULAR_SPEED) {  // util.cpp:585-594
        const float a2 = av.x * av.x + av.y * av.y + av.z * av.z;
        if (a2 > d1.x * d1.x) {
            const float c = d1.x / sqrtf(a2);
            av.x *= c; av.y *= c; av.z *= c;
        }
    }
    p4.x += h * lv.x; p4.y += h * lv.y; p4.z += h * lv.z;
    float dq0 = 0.5f * (-av.x * q4.y - av.y * q4.z - av.z * q4.w);
    float dq1 = 0.5f * (av.x * q4.x + av.y * q4.w - av.z * q4.z);
    float dq2 = 0.5f * (-av.x * q4.w + av.y * q4.x + av.z * q4.y);
    float dq3 = 0.5f * (av.x * q4.z - av.y * q4.y + av.z * q4.x);
    q4.x += h * dq0; q4.y += h * dq1; q4.z += h * dq2; q4.w += h * dq3;
    float l2 = q4.x * q4.x + q4.y * q4.y + q4.z * q4.z + q4.w * q4.w;
    if (l2 > 0.f) {
        const float l = 1.f / sqrtf(l2);
        q4.x *= l; q4.y *= l; q4.z *= l; q4.w *= l;
    } else {
        q4 = make_float4(1.f, 0.f, 0.f, 0.f);
    }
    // damping after the move (util.cpp:675-698): scale a speed above its threshold
    if (dfl & DAMP_LINEAR) {
        if (lv.x * lv.x + lv.y * lv.y + lv.z * lv.z > d0.z) {
            const float k = 1.f - d0.x;
            lv.x *= k; lv.y *= k; lv.z *= k;
        }
    }
    if (dfl & DAMP_ANGULAR) {
        if (av.x * av.x + av.y * av.y + av.z * av.z > d0.w) {
            const float k = 1.f - d0.y;
            av.x *= k; av.y *= k; av.z *= k;
        }
    }
    D.pos[b] = p4; D.quat[b] = q4; D.lvel[b] = lv; D.avel[b] = av;
    D.facc[b] = make_float4(0.f, 0.f, 0.f, 0.f);
    D.tacc[b] = make_float4(0.f, 0.f, 0.f, 0.f);
}

__device__ __forceinline__ void solve_world(const DevWorlds &D, const StepParams &P, int w, float *s_dyn) {
    const int t = threadIdx.x;
    const int b0 = D.body_start[w], nb = D.body_start[w + 1] - b0;
    const int r0 = D.row_start[w];
    const int m = D.nrows[w];
    const int L = D.nphases[w];
    const int *ps = D.phase_start + r0 + w;
    const int *sched = D.sched + r0;
    const int *perm = D.perm + r0, *inv = D.inv + r0;
    float *lam = D.lambda + r0;
    const bool warm = P.warm >= 1 && D.lambda0;
    if (t == 0 && D.wtime) D.wtime[2 * w] = (long long)wall_clock64();
    // LDS path: bodies [lds_bodies][16], rows [lds_rows][16], lambda [lds_rows],
    // items [lds_items_of(lds_rows)], phase starts [LDS_PHASES + 1] (16-bit)
    float *s_body = s_dyn;
    float *s_rows = s_dyn + P.lds_bodies * LDS_BODY;
    float *s_lam = s_rows + P.lds_rows * LDS_ROW;
    int *s_sched = reinterpret_cast<int *>(s_lam + P.lds_rows);
    unsigned short *s_ps = reinterpret_cast<unsigned short *>(s_sched + lds_items_of(P.lds_rows));
    float *s_trash = s_dyn + solve_lds_trash_offset(P);
    bool lds = m > 0 && m <= P.lds_rows && nb < P.lds_bodies && L <= LDS_PHASES && ps[L] <= lds_items_of(P.lds_rows);
    if (lds) {
        lds = lds_stage(D, b0, nb, r0, m, L, ps, sched, perm, s_body, s_rows, s_sched, s_ps);
        __syncthreads();
    }
    int it = 0;
    float res = -1.f;
    if (lds) {
        if (warm) {
            // lambda0 clamped to the row bounds (LDS holds rows in sweep order)
            for (int pass = 0; pass < 2; pass++) {
                for (int p = t; p < m; p += 64) {
                    const float4 a3 = ld4(s_rows, p, 3);
                    const int fidx = (as_int(a3.z) >> 16) - 1;
                    if ((fidx >= 0) != (pass == 1)) continue;
                    float lo = a3.x, hi = a3.y;
                    if (fidx >= 0) { hi = fabsf(hi * s_lam[fidx]); lo = -hi; }
                    s_lam[p] = fminf(fmaxf(D.lambda0[r0 + inv[p]], lo), hi);
                }
                __syncthreads();
            }
            lds_accumulate_fc(s_rows, s_lam, s_body, nb, m);
        } else {
            for (int r = t; r < m; r += 64) s_lam[r] = 0.f;
        }
        __syncthreads();
        bool restarted = false;
        const int nitem = s_ps[L] & 2047;
        float *trash = s_trash + t;
        const int quad = t >> 2;
        while (it < P.max_iters) {
            // 16 quads take a phase's items; each quad's first item of the
            // next phase is read while the current phase runs
            int s = 0, pe = s_ps[1], e = pe & 2047;
            int nmax = __builtin_amdgcn_readfirstlane(s_ps[0] >> 11);
            int cur = quad < e ? s_sched[quad] : -1;
            for (int p = 0; p < L; p++) {
                const int pe2 = p + 1 < L ? s_ps[p + 2] : pe;
                const int e2 = pe2 & 2047;
                const int cand = e + quad < nitem ? s_sched[e + quad] : -1;
                process_item_quad(s_rows, s_lam, s_body, nb, cur, nmax, trash);
                for (int k0 = s + 16; k0 < e; k0 += 16) {
                    const int k = k0 + quad;
                    process_item_quad(s_rows, s_lam, s_body, nb, k < e ? s_sched[k] : -1, nmax, trash);
                }
                wave_sync();
                cur = e + quad < e2 ? cand : -1;
                nmax = __builtin_amdgcn_readfirstlane(pe >> 11);
                s = e;
                e = e2;
                pe = pe2;
            }
            it++;
            if (P.tol > 0.f && (it % P.check_every == 0 || it == P.max_iters)) {
                res = world_residual_lds(s_rows, s_lam, s_body, m, P);
                if (res <= P.tol) break;
                if (warm && !restarted && P.restart_after > 0 && it >= P.restart_after) {
                    // a warm start that has not converged restarts from lambda = 0
                    restarted = true;
                    for (int r = t; r < m; r += 64) s_lam[r] = 0.f;
                    for (int i = t; i < nb; i += 64) {
                        const float z[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
                        lds_store_fc(s_body, i, z);
                    }
                    __syncthreads();
                }
            }
        }
        if (P.final_residual && res < 0.f) res = world_residual_lds(s_rows, s_lam, s_body, m, P);
        for (int r = t; r < m; r += 64) lam[r] = s_lam[perm[r]];
        __syncthreads();
    }
    // global-memory path: rows streamed from HBM, fc in LDS
    float *s_fc = s_dyn;
    if (!lds) {
        for (int k = t; k < nb * 6; k += 64) s_fc[k] = 0.f;
        if (warm) {
            // lambda0 clamped to the row bounds; friction bounds use the clamped normal
            for (int pass = 0; pass < 2; pass++) {
                for (int r = t; r < m; r += 64) {
                    const float4 *rr = reinterpret_cast<const float4 *>(D.rows + (size_t)(r0 + r) * ROW_STRIDE);
                    const float4 a3 = rr[3], a7 = rr[7];
                    const int fidx = as_int(a7.z);
                    if ((fidx >= 0) != (pass == 1)) continue;
                    float lo = a3.z, hi = a3.w;
                    if (fidx >= 0) { hi = fabsf(hi * lam[fidx]); lo = -hi; }
                    lam[r] = fminf(fmaxf(D.lambda0[r0 + r], lo), hi);
                }
                __syncthreads();
            }
            for (int r = t; r < m; r += 64) {
                const float4 *rr = reinterpret_cast<const float4 *>(D.rows + (size_t)(r0 + r) * ROW_STRIDE);
                const float4 a4 = rr[4], a5 = rr[5], a6 = rr[6], a7 = rr[7];
                const int b1 = as_int(a7.x), b2 = as_int(a7.y);
                const float l = lam[r];
                if (l != 0.f) {
                    atomicAdd(&s_fc[b1 * 6 + 0], l * a4.x); atomicAdd(&s_fc[b1 * 6 + 1], l * a4.y);
                    atomicAdd(&s_fc[b1 * 6 + 2], l * a4.z); atomicAdd(&s_fc[b1 * 6 + 3], l * a4.w);
                    atomicAdd(&s_fc[b1 * 6 + 4], l * a5.x); atomicAdd(&s_fc[b1 * 6 + 5], l * a5.y);
                    if (b2 >= 0) {
                        atomicAdd(&s_fc[b2 * 6 + 0], l * a5.z); atomicAdd(&s_fc[b2 * 6 + 1], l * a5.w);
                        atomicAdd(&s_fc[b2 * 6 + 2], l * a6.x); atomicAdd(&s_fc[b2 * 6 + 3], l * a6.y);
                        atomicAdd(&s_fc[b2 * 6 + 4], l * a6.z); atomicAdd(&s_fc[b2 * 6 + 5], l * a6.w);
                    }
                }
            }
        } else {
            for (int r = t; r < m; r += 64) lam[r] = 0.f;
        }
        __syncthreads();

        if (m > 0) {
            bool restarted = false;
            while (it < P.max_iters) {
                for (int p = 0; p < L; p++) {
                    const int s = ps[p], e = ps[p + 1];
                    for (int k = s + t; k < e; k += 64) process_item_packed(D, r0, inv, nb, sched[k], s_fc);
                    __syncthreads();
                }
                it++;
                if (P.tol > 0.f && (it % P.check_every == 0 || it == P.max_iters)) {
                    res = world_residual(D, P, r0, m, s_fc);
                    if (res <= P.tol) break;
                    if (warm && !restarted && P.restart_after > 0 && it >= P.restart_after) {
                        // a warm start that has not converged restarts from lambda = 0
                        restarted = true;
                        for (int r = t; r < m; r += 64) lam[r] = 0.f;
                        for (int k = t; k < nb * 6; k += 64) s_fc[k] = 0.f;
                        __syncthreads();
                    }
                }
            }
            if (P.final_residual && res < 0.f) res = world_residual(D, P, r0, m, s_fc);
        } else {
            res = 0.f;
        }
    }
    if (t == 0) { D.niters[w] = it; D.residual[w] = res; }

    // keep this step's contacts and impulses for next step's matching
    // (body-1 frame at the pose the contacts were generated for)
    if (P.warm >= 2 && P.collide) {
        const int j0 = D.joint_start[w];
        const int nball = D.ball_start[w + 1] - D.ball_start[w];
        const int nj = D.njoints[w];
        int nkeep = 0;
        for (int k = nball + t; k < nj; k += 64) {
            const float *rec = D.joints + (size_t)(j0 + k) * JOINT_STRIDE;
            const int ri = D.joff[j0 + k];
            float *o = D.prev + (size_t)(j0 + k - nball) * JOINT_STRIDE;
            o[PV_G1] = rec[14]; o[PV_G2] = rec[15];
            if (ri >= m) {  // a contact that got no rows: keeps the pair order, never matches
                o[PV_LOCAL] = o[PV_LOCAL + 1] = o[PV_LOCAL + 2] = 1e30f;
                o[PV_LAM] = o[PV_LAM + 1] = o[PV_LAM + 2] = 0.f;
                continue;
            }
            const Surface s = D.surf[(int)rec[13]];
            const int mj = contact_rows_m(s);
            float lmb[3] = {lam[ri], 0.f, 0.f};
            int row = 1;
            if (mj > 1) {
                if (s.mu > 0.f) lmb[1] = lam[ri + row++];
                const float mu2 = (s.mode & 0x001) ? s.mu2 : s.mu;
                if (mu2 > 0.f && row < mj) lmb[2] = lam[ri + row++];
            }
            const int b1 = (int)rec[1];
            const float *s1 = D.bscr + (size_t)(b0 + b1) * BSCR_STRIDE;
            float4 x1 = D.pos[b0 + b1];
            float d[3] = {rec[3] - x1.x, rec[4] - x1.y, rec[5] - x1.z}, pl[3];
            mul1_331(pl, s1, d);
            o[PV_LOCAL] = pl[0]; o[PV_LOCAL + 1] = pl[1]; o[PV_LOCAL + 2] = pl[2];
            o[PV_N] = rec[6]; o[PV_N + 1] = rec[7]; o[PV_N + 2] = rec[8];
            o[PV_LAM] = lmb[0]; o[PV_LAM + 1] = lmb[1]; o[PV_LAM + 2] = lmb[2];
        }
        (void)nkeep;
        if (t == 0) D.nprev[w] = nj - nball;
        __syncthreads();
    }

    // Stage4b, Stage6a, Stage6b + dxStepBody (infinitesimal rotation)
    for (int i = t; i < nb; i += 64) {
        float fc[6];
        if (lds) lds_fc(s_body, i, fc);
        else for (int k = 0; k < 6; k++) fc[k] = s_fc[i * 6 + k];
        integrate_body(D, P, b0 + i, fc, m > 0, D.facc, D.tacc);
    }
    if (t == 0 && D.wtime) D.wtime[2 * w + 1] = (long long)wall_clock64();
}

__global__ __launch_bounds__(64) void k_solve_small(DevWorlds D, StepParams P) {
    extern __shared__ float4 s_dyn4[];
    const int w = D.order ? D.order[P.wbase + (int)blockIdx.x] : P.wbase + (int)blockIdx.x;
    if (blockIdx.x == 0 && threadIdx.x == 0 && D.feat_count) {  // features were consumed; keep the count
        D.feat_count[2 * P.split + 1] = D.feat_count[2 * P.split];
        D.feat_count[2 * P.split] = 0;
    }
    solve_world(D, P, w, reinterpret_cast<float *>(s_dyn4));
}

// Warm-start network over world w's nfeat feature entries (world-local slots
// from joint_start[w]): tiles of 16 contacts on MFMA, weights read from
// global memory (L2-resident), activations in the wave's LDS.  Writes
// lambda0 = s (previous + predicted change) like k_mlp_infer.
__device__ __forceinline__ void infer_world(const DevWorlds &D, const MlpDev &M, int w, int nfeat, InferBuf &tb) {
    const int j0 = D.joint_start[w];
    const int l = threadIdx.x & 63, o = l & 15, g = l >> 4;
    const float *nrm = M.norm;
    for (int s0 = 0; s0 < nfeat; s0 += 16) {
        const int nv = min(16, nfeat - s0);
        const float *xr = D.feat + (size_t)(j0 + s0) * FEAT_DIM;
        const f32x4 y = forward_tile<false>(M.W1h, M.W2h, M.W3h, M.b1, M.b2, M.b3, nrm, tb, xr, nv);
        if (o < NOUT) {
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int smp = 4 * g + j;
                if (smp < nv) {
                    const float *x = xr + (size_t)smp * FEAT_DIM;
                    const float use = (M.gate && x[12] > 0.5f) ? 0.f : 1.f;
                    D.lambda0[D.feat_row[j0 + s0 + smp] + o] =
                        x[15] * (x[7 + o] + use * (y[j] * nrm[2 * NIN + NOUT + o] + nrm[2 * NIN + o]));
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// Persistent fused step: each workgroup (one wave) takes one world through
// `nsteps` complete steps -- prepare (collide, rows, schedule), the network
// when one is attached, solve and integrate -- with no kernel boundary in
// between, so a world that needs many sweeps in one step delays only itself.
// LDS holds one stage at a time (prepare arrays, then the network's
// activations, then the solve's rows).
__global__ __launch_bounds__(64) void k_step_fused(DevWorlds D, StepParams P, MlpDev M, int use_mlp, int nsteps) {
    extern __shared__ float4 s_dyn4[];
    const int w = D.order ? D.order[P.wbase + (int)blockIdx.x] : P.wbase + (int)blockIdx.x;
    for (int k = 0; k < nsteps; k++) {
        const int nfeat = prepare_world(D, P, w, prep_lds_carve(s_dyn4));
        __syncthreads();
        if (use_mlp) {
            infer_world(D, M, w, nfeat, *reinterpret_cast<InferBuf *>(s_dyn4));
            __syncthreads();
        }
        solve_world(D, P, w, reinterpret_cast<float *>(s_dyn4));
        __syncthreads();
    }
}

}  // namespace odx

namespace odx {

// ===================================================== single large world ===
// dWorldQuickStep of one world of up to SL_MAXB bodies: one workgroup of 1024
// threads.  The host builds the joint list in island order (the reference's
// DFS, util.cpp:724-855), each joint's first row (joff) and the sweep schedule;
// the kernel runs Stage0 bodies, the rows of every joint, the SOR sweeps with
// fc in LDS, the optional residual, joint feedback (Stage4b,
// quickstep.cpp:3105-3150) and the integration.
constexpr int SL_BLOCK = 1024;
static_assert(SL_MAXB * 6 * sizeof(float) <= 65536, "fc of the single-world kernel must fit 64 KB of LDS");

__device__ float block_max(float v, float *s_red) {
    v = wave_max(v);
    const int wv = threadIdx.x >> 6;
    __syncthreads();
    if (lane_id() == 0) s_red[wv] = v;
    __syncthreads();
    float r = 0.f;
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) r = fmaxf(r, s_red[i]);
    return r;
}

// One work item of the single-world sweeps (first row << 3 | rows) with lambda
// at `lam` (LDS or global); `pre`, when given, holds the item's first row.
// With the dynamic iteration count (fa non-null): items of islands that have
// stopped are skipped, and every non-zero delta adds delta * max|iMJ side| to
// the body's positive or negative force adjustment (quickstep.cpp:3001-3031;
// iMJ is stored unscaled, as compute_invM_JT's maxabs reads it).
// The single-world row update's arithmetic with every fused multiply-add
// written out, so that each sweep variant rounds alike: J . f over one body's
// six entries (j0.xyzw, j4, j5), the product of entry 0 first.
__device__ __forceinline__ float row_dot6(const float *f, float4 j0, float j4, float j5) {
    float s = f[0] * j0.x;
    s = fmaf(f[1], j0.y, s);
    s = fmaf(f[2], j0.z, s);
    s = fmaf(f[3], j0.w, s);
    s = fmaf(f[4], j4, s);
    return fmaf(f[5], j5, s);
}
// f += delta * iMJ for one body's six entries
__device__ __forceinline__ void row_axpy6(float *f, float delta, float4 m0, float m4, float m5) {
    f[0] = fmaf(delta, m0.x, f[0]);
    f[1] = fmaf(delta, m0.y, f[1]);
    f[2] = fmaf(delta, m0.z, f[2]);
    f[3] = fmaf(delta, m0.w, f[3]);
    f[4] = fmaf(delta, m4, f[4]);
    f[5] = fmaf(delta, m5, f[5]);
}

// bbase / rbase: the first body and row whose fc and lambda the arrays hold,
// and row_base the first row of rows4 (a workgroup holding one island group's
// range); row records and items keep global indices.
__device__ __forceinline__ void process_item_l(const float4 *rows4, const float4 *pre, int item, float *lam,
                                               float *s_fc, float *fa = nullptr, const int *isl_b = nullptr,
                                               const int *ist = nullptr, int bbase = 0, int rbase = 0,
                                               int row_base = 0) {
#pragma clang fp contract(fast)
    const int rs = item >> 3, nr = item & 7;
    const float4 *row4 = rows4 + (size_t)(rs - row_base) * (ROW_STRIDE / 4);
    float4 a[8];
    if (pre) {
#pragma unroll
        for (int k = 0; k < 8; k++) a[k] = pre[k];
    } else {
#pragma unroll
        for (int k = 0; k < 8; k++) a[k] = row4[k];
    }
    const int b1 = as_int(a[7].x), b2 = as_int(a[7].y);
    if (fa && !ist[isl_b[b1] * 4 + 2]) return;
    const int l1 = b1 - bbase, l2 = b2 - bbase, rl = rs - rbase;
    float f1[6], f2[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < 6; k++) f1[k] = s_fc[l1 * 6 + k];
    if (b2 >= 0) {
#pragma unroll
        for (int k = 0; k < 6; k++) f2[k] = s_fc[l2 * 6 + k];
    }
    float lam_first = 0.f;
    for (int q = 0; q < nr; q++) {
        if (q > 0) {
#pragma unroll
            for (int k = 0; k < 8; k++) a[k] = row4[q * 8 + k];
        }
        const float old = lam[rl + q];
        float delta = fmaf(-old, a[1].w, a[1].z);
        delta -= row_dot6(f1, a[0], a[1].x, a[1].y);
        if (b2 >= 0) delta -= row_dot6(f2, a[2], a[3].x, a[3].y);
        float lo = a[3].z, hi = a[3].w;
        const int fidx = as_int(a[7].z);
        if (fidx >= 0) {
            const float lf = (fidx == rs) ? lam_first : lam[fidx - rbase];
            hi = fabsf(hi * lf);
            lo = -hi;
        }
        float nl = old + delta;
        if (nl < lo) { delta = lo - old; nl = lo; }
        else if (nl > hi) { delta = hi - old; nl = hi; }
        lam[rl + q] = nl;
        if (q == 0) lam_first = nl;
        row_axpy6(f1, delta, a[4], a[5].x, a[5].y);
        if (b2 >= 0) row_axpy6(f2, delta, make_float4(a[5].z, a[5].w, a[6].x, a[6].y), a[6].z, a[6].w);
        if (fa && delta != 0.f) {
            const int sg = delta > 0.f ? 1 : 0;
            const float m1 = fmaxf(fmaxf(fmaxf(fabsf(a[4].x), fabsf(a[4].y)), fmaxf(fabsf(a[4].z), fabsf(a[4].w))),
                                   fmaxf(fabsf(a[5].x), fabsf(a[5].y)));
            fa[b1 * 2 + sg] += delta * m1;
            if (b2 >= 0) {
                const float m2 = fmaxf(fmaxf(fmaxf(fabsf(a[5].z), fabsf(a[5].w)), fmaxf(fabsf(a[6].x), fabsf(a[6].y))),
                                       fmaxf(fabsf(a[6].z), fabsf(a[6].w)));
                fa[b2 * 2 + sg] += delta * m2;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 6; k++) s_fc[l1 * 6 + k] = f1[k];
    if (b2 >= 0) {
#pragma unroll
        for (int k = 0; k < 6; k++) s_fc[l2 * 6 + k] = f2[k];
    }
}

// Levels ph[0 .. L] of island groups swept without the dynamic iteration
// count: every item is one row (the reference order's levels), so a level is
// one row update per lane with no branch on its operands.  The rows come from
// global memory (L2) three levels ahead: at the top of level p a lane holds
// its items of levels p .. p + 3, the rows of p .. p + 2 in three of four
// register buffers that rotate (the fourth receives p + 3's row during level
// p), and the level ends up to p + 3; level p also requests level p + 4's end
// and item.  Reads past a level's end are clamped to the arrays and
// discarded.  A row without body 2 reads body 1's fc in its place, uses zeros,
// and stores body 2's slot first so that body 1's values win.  The arithmetic
// is process_item_l's, term for term.  bbase / rbase / row_base as for
// process_item_l.
template <class Sync>
__device__ __forceinline__ void lean_levels_fn(const float4 *rows4, const int *sched, const int nit, float *lam,
                                               float *s_fc, const int ln, const int stride, const int *ph,
                                               const int L, Sync sync, const int bbase = 0, const int rbase = 0,
                                               const int row_base = 0) {
#pragma clang fp contract(fast)
    const int last_item = nit - 1;
    // level q spans [ph[q], ph[q + 1]); levels past L - 1 are empty
    int s = ph[0], e = ph[1];
    int e1 = L > 1 ? ph[2] : e, e2 = L > 2 ? ph[3] : e1, e3 = L > 3 ? ph[4] : e2;
    int cur = s + ln < e ? sched[s + ln] : -1;
    int i1 = e + ln < e1 ? sched[e + ln] : -1;
    int i2 = e1 + ln < e2 ? sched[e1 + ln] : -1;
    int i3 = e2 + ln < e3 ? sched[e2 + ln] : -1;
    auto fetch = [&](float4 *X, const int item) {
        const float4 *rr = rows4 + (size_t)(item >= 0 ? (item >> 3) - row_base : 0) * (ROW_STRIDE / 4);
#pragma unroll
        for (int k = 0; k < 8; k++) X[k] = rr[k];
    };
    float4 A[8], B[8], C[8], Dd[8];
    fetch(A, cur);
    fetch(B, i1);
    fetch(C, i2);
    // a: the row of level p; Y: the buffer for level p + 3's row
    auto level = [&](const int p, const float4 *a, float4 *Y) {
#pragma clang fp contract(fast)
        // the row's last word (Ad) is not needed here; keeping it live until
        // its level stops the compiler from reusing its register while the
        // load that fills it is in flight (which would wait for that load)
        asm volatile("" ::"v"(a[7].w));
        // a lane without an item reads slot 0 (its row buffer may hold a row
        // of another range)
        const bool on = cur >= 0;
        const int rs = on ? (cur >> 3) - rbase : 0;
        const int b2 = as_int(a[7].y), fidx = as_int(a[7].z);
        const int b1 = on ? as_int(a[7].x) - bbase : 0;
        const int b2c = on && b2 >= 0 ? b2 - bbase : b1;
        float f1[6], f2[6];
#pragma unroll
        for (int k = 0; k < 6; k++) f1[k] = s_fc[b1 * 6 + k];
#pragma unroll
        for (int k = 0; k < 6; k++) f2[k] = s_fc[b2c * 6 + k];
        const float old = lam[rs];
        const float lf = lam[on && fidx >= 0 ? fidx - rbase : rs];
        // level p + 3's row, level p + 4's end and item
        fetch(Y, i3);
        const int e4 = ph[min(p + 5, L)];
        const int n4 = sched[min(e3 + ln, last_item)];
        if (on) {
#pragma unroll
            for (int k = 0; k < 6; k++) f2[k] = b2 >= 0 ? f2[k] : 0.f;
            float delta = fmaf(-old, a[1].w, a[1].z);
            delta -= row_dot6(f1, a[0], a[1].x, a[1].y);
            if (b2 >= 0) delta -= row_dot6(f2, a[2], a[3].x, a[3].y);
            float lo = a[3].z, hi = a[3].w;
            if (fidx >= 0) {
                hi = fabsf(hi * lf);
                lo = -hi;
            }
            const float nl0 = old + delta;
            const bool bl = nl0 < lo, bh = !bl && nl0 > hi;
            const float nl = bl ? lo : (bh ? hi : nl0);
            delta = bl ? lo - old : (bh ? hi - old : delta);
            lam[rs] = nl;
            row_axpy6(f1, delta, a[4], a[5].x, a[5].y);
            row_axpy6(f2, delta, make_float4(a[5].z, a[5].w, a[6].x, a[6].y), a[6].z, a[6].w);
#pragma unroll
            for (int k = 0; k < 6; k++) s_fc[b2c * 6 + k] = f2[k];
#pragma unroll
            for (int k = 0; k < 6; k++) s_fc[b1 * 6 + k] = f1[k];
        }
        for (int k = s + ln + stride; k < e; k += stride)
            process_item_l(rows4, nullptr, sched[k], lam, s_fc, nullptr, nullptr, nullptr, bbase, rbase, row_base);
        sync();
        const int e4v = p + 4 < L ? e4 : e3;
        s = e;
        e = e1;
        e1 = e2;
        e2 = e3;
        cur = i1;
        i1 = i2;
        i2 = i3;
        i3 = e3 + ln < e4v ? n4 : -1;
        e3 = e4v;
    };
    for (int p = 0; p < L; p += 4) {
        level(p, A, Dd);
        if (p + 1 < L) level(p + 1, B, A);
        if (p + 2 < L) level(p + 2, C, B);
        if (p + 3 < L) level(p + 3, Dd, C);
    }
}

// The single-world step.  STAGED = false: fc in LDS, lambda, the schedule and
// the rows in global memory.  STAGED = true: lambda, the work items and the
// phase starts are copied into LDS next to fc, and every thread loads its item
// and first row of the next phase before the phase barrier, so a phase costs
// LDS latency instead of dependent global loads.  ROWS_LDS (with STAGED): the
// assembled rows are copied into LDS too (128 B each, first in the block), so
// a level of one row -- a stack's long dependency chains -- costs LDS latency
// rather than a global round trip.
template <bool STAGED, int NT, bool ROWS_LDS = false>
__global__ __launch_bounds__(NT) void k_single_step_t(DevWorlds D, StepParams P) {
    extern __shared__ float4 s_dyn4[];
    __shared__ float s_red[32];
    const int t = threadIdx.x;
    const int nb = D.body_start[1];
    const int nj = D.njoints[0];
    const int m = D.nrows[0];
    const int L = D.nphases[0];
    const int nit = STAGED ? D.phase_start[L] : 0;
    float *s_dyn = reinterpret_cast<float *>(s_dyn4) + (ROWS_LDS ? (size_t)m * ROW_STRIDE : 0);
    const float4 *rows4 = ROWS_LDS ? s_dyn4 : reinterpret_cast<const float4 *>(D.rows);
    float *s_fc = s_dyn;
    float *lam = STAGED ? s_dyn + nb * 6 : D.lambda;
    const int *sched = STAGED ? reinterpret_cast<const int *>(s_dyn + nb * 6 + m) : D.sched;
    const int *ph = STAGED ? reinterpret_cast<const int *>(s_dyn + nb * 6 + m + nit) : D.phase_start;
    // dynamic iteration count: force adjustments [nb][2], island of each body,
    // per island (sweeps, extra sweeps, active, not converged), limit per island
    const int nisl = P.dyn ? P.nislands : 0;
    float *fa = P.dyn ? s_dyn + nb * 6 + (STAGED ? m + nit + L + 1 : 0) : nullptr;
    int *isl_b = P.dyn ? reinterpret_cast<int *>(fa + nb * 2) : nullptr;
    int *ist = P.dyn ? isl_b + nb : nullptr;
    float *ilim = P.dyn ? reinterpret_cast<float *>(ist + nisl * 4) : nullptr;
    for (int i = t; i < nb; i += NT) body_prep(D, P, i, D.facc, D.tacc);
    for (int k = t; k < nb * 6; k += NT) s_fc[k] = 0.f;
    if (STAGED) {
        int *ss = reinterpret_cast<int *>(s_dyn + nb * 6 + m);
        for (int k = t; k < nit; k += NT) ss[k] = D.sched[k];
        for (int k = t; k <= L; k += NT) ss[nit + k] = D.phase_start[k];
    }
    __syncthreads();
    for (int j = t; j < nj; j += NT)
        build_joint_rows(D, P, D.joints + (size_t)j * JOINT_STRIDE, D.jext ? D.jext + (size_t)j * JEXT_STRIDE : nullptr,
                         0, 0, D.joff[j]);
    for (int r = t; r < m; r += NT) lam[r] = 0.f;
    __syncthreads();
    if (ROWS_LDS) {
        // the block's own rows, written above: visible after the barrier
        const float4 *g4 = reinterpret_cast<const float4 *>(D.rows);
        for (int k = t; k < m * (ROW_STRIDE / 4); k += NT) s_dyn4[k] = g4[k];
        __syncthreads();
    }
    int it = 0;
    float res = -1.f;
    // one sweep over the phases (items of stopped islands skipped when fa is
    // set); with a segmented schedule (random reordering) sweep nsw uses
    // segment nsw / 8
    int nsw = 0;
    auto sweep = [&]() {
            if (!STAGED && D.seg_table) {
                const int sg = min(nsw / 8, D.nseg - 1);
                const int *sch = D.sched + D.seg_table[3 * sg], *phs = D.phase_start + D.seg_table[3 * sg + 1];
                const int Ls = D.seg_table[3 * sg + 2];
                for (int p = 0; p < Ls; p++) {
                    const int s = phs[p], e = phs[p + 1];
                    for (int k = s + t; k < e; k += NT) process_item_l(rows4, nullptr, sch[k], lam, s_fc, fa, isl_b, ist);
                    __syncthreads();
                }
                nsw++;
                return;
            }
            nsw++;
            // island groups (sweep mode 3): the phase table opens with
            // wave_groups + 1 offsets of each group's level list; wave g sweeps
            // group g's levels alone, ordered by wavefront fences, and the
            // block meets once per sweep
            auto group_levels = [&](auto &&levels_of) {
                const int G = P.wave_groups, wv = t >> 6;
                if (wv < G) {
                    const int *gph = ph + (G + 1) + ph[wv];
                    levels_of(t & 63, 64, gph, ph[wv + 1] - ph[wv] - 1, [] {
                        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                        __builtin_amdgcn_wave_barrier();
                        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    });
                }
                __syncthreads();
            };
            if (!STAGED && P.wave_groups > 0) {
                group_levels([&](const int ln, const int stride, const int *gp, const int Lg, auto sync) {
                    for (int p = 0; p < Lg; p++) {
                        const int s = gp[p], e = gp[p + 1];
                        for (int k = s + ln; k < e; k += stride)
                            process_item_l(rows4, nullptr, sched[k], lam, s_fc, fa, isl_b, ist);
                        sync();
                    }
                });
                return;
            }
            if (STAGED) {
                // levels from LDS with the next level's rows prefetched; lane
                // `ln` of `stride` takes items ln, ln + stride, ...; `sync` ends a
                // level; levels `ph_[0 .. L_]`
                auto levels = [&](const int ln, const int stride, const int *ph_, const int L_, auto sync) {
                    const int *ph = ph_;
                    const int L = L_;
                    int s = ph[0], e = ph[1];
                    int cur = s + ln < e ? sched[s + ln] : -1;
                    float4 pre[8], npre[8];
                    if (cur >= 0) {
                        const float4 *rr = rows4 + (size_t)(cur >> 3) * (ROW_STRIDE / 4);
#pragma unroll
                        for (int k = 0; k < 8; k++) pre[k] = rr[k];
                    }
                    for (int p = 0; p < L; p++) {
                        int nxt = -1, s2 = 0, e2 = 0;
                        if (p + 1 < L) {
                            s2 = ph[p + 1];
                            e2 = ph[p + 2];
                            if (s2 + ln < e2) {
                                nxt = sched[s2 + ln];
                                const float4 *rr = rows4 + (size_t)(nxt >> 3) * (ROW_STRIDE / 4);
#pragma unroll
                                for (int k = 0; k < 8; k++) npre[k] = rr[k];
                            }
                        }
                        if (cur >= 0) process_item_l(rows4, pre, cur, lam, s_fc, fa, isl_b, ist);
                        for (int k = s + ln + stride; k < e; k += stride)
                            process_item_l(rows4, nullptr, sched[k], lam, s_fc, fa, isl_b, ist);
                        sync();
                        s = s2;
                        e = e2;
                        cur = nxt;
                        if (nxt >= 0) {
#pragma unroll
                            for (int k = 0; k < 8; k++) pre[k] = npre[k];
                        }
                    }
                };
                auto lean_levels = [&](const int ln, const int stride, const int *ph, const int L, auto sync) {
                    lean_levels_fn(rows4, sched, nit, lam, s_fc, ln, stride, ph, L, sync);
                };
                if (P.wave_groups > 0 && !fa) {
                    group_levels([&](const int ln, const int stride, const int *gp, const int Lg, auto sync) {
                        lean_levels(ln, stride, gp, Lg, sync);
                    });
                } else if (P.wave_groups > 0) {
                    group_levels([&](const int ln, const int stride, const int *gp, const int Lg, auto sync) {
                        levels(ln, stride, gp, Lg, sync);
                    });
                } else if (P.wave_sweep) {
                    // narrow levels: the first wave sweeps alone.  A wave's LDS
                    // accesses complete in order, so a level needs no block
                    // barrier, only a fence that keeps the compiler from moving
                    // accesses across it; the block meets once per sweep.
                    if (t < 64)
                        levels(t, 64, ph, L, [] {
                            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                            __builtin_amdgcn_wave_barrier();
                            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                        });
                    __syncthreads();
                } else {
                    levels(t, NT, ph, L, [] { __syncthreads(); });
                }
            } else {
                for (int p = 0; p < L; p++) {
                    const int s = ph[p], e = ph[p + 1];
                    for (int k = s + t; k < e; k += NT) process_item_l(rows4, nullptr, sched[k], lam, s_fc, fa, isl_b, ist);
                    __syncthreads();
                }
            }
    };
    if (m > 0 && P.dyn) {
        // the iteration loop of dxQuickStepIsland (quickstep.cpp:1820-1853) per
        // island: after each sweep an island that has run max_iters (+ its extra
        // sweeps) stops; at max_iters it may switch to up to dyn_max_extra extra
        // sweeps with the extra-iteration delta as its limit; it stops early once
        // every body's positive and negative adjustment is below its limit
        // (CheckForMaximumToBeLessThanLimitAndResetMaxAdjustments :3253-3285)
        for (int i = t; i < nisl; i += NT) {
            ist[i * 4 + 0] = 0; ist[i * 4 + 1] = 0; ist[i * 4 + 2] = 0; ist[i * 4 + 3] = 0;
            ilim[i] = P.dyn_premature;
        }
        for (int b = t; b < nb; b += NT) { fa[b * 2] = 0.f; fa[b * 2 + 1] = 0.f; isl_b[b] = D.body_island[b]; }
        __syncthreads();
        for (int r = t; r < m; r += NT) ist[isl_b[as_int(D.rows[(size_t)r * ROW_STRIDE + R_B1])] * 4 + 2] = 1;
        __syncthreads();
        int any = 1;
        while (any) {
            sweep();
            for (int i = t; i < nisl; i += NT) {
                int *S = ist + i * 4;
                if (!S[2]) continue;
                S[0]++;
                if (S[0] - S[1] == P.max_iters) {
                    if (S[1] != 0 || P.dyn_max_extra == 0) { S[2] = 0; continue; }
                    S[1] = P.dyn_max_extra;
                    ilim[i] = P.dyn_extra_delta;
                }
                S[3] = ilim[i] == 0.f ? 1 : 0;
            }
            __syncthreads();
            for (int b = t; b < nb; b += NT) {
                const int i = isl_b[b];
                if (i >= 0 && ist[i * 4 + 2] && (!(fa[b * 2 + 1] < ilim[i]) || !(-fa[b * 2] < ilim[i]))) ist[i * 4 + 3] = 1;
                fa[b * 2] = 0.f;
                fa[b * 2 + 1] = 0.f;
            }
            __syncthreads();
            int mine = 0;
            for (int i = t; i < nisl; i += NT) {
                int *S = ist + i * 4;
                if (S[2] && !S[3]) S[2] = 0;
                mine |= S[2];
            }
            any = __syncthreads_or(mine);
        }
        int mx = 0;
        for (int i = t; i < nisl; i += NT) {
            mx = max(mx, ist[i * 4]);
            if (D.island_sweeps) D.island_sweeps[i] = ist[i * 4];
        }
        it = (int)block_max((float)mx, s_red);
    } else if (m > 0) {
        while (it < P.max_iters) {
            sweep();
            it++;
            if (P.tol > 0.f && (it % P.check_every == 0 || it == P.max_iters)) {
                // residual over all rows (block-wide)
                float mx = 0.f, lmax = 0.f;
                const float iw = 1.f / P.sor_w;
                for (int r = t; r < m; r += NT) {
                    const float4 *rr = reinterpret_cast<const float4 *>(D.rows + (size_t)r * ROW_STRIDE);
                    const float4 a0 = rr[0], a1 = rr[1], a2 = rr[2], a3 = rr[3], a7 = rr[7];
                    const int b1 = as_int(a7.x), b2 = as_int(a7.y), fidx = as_int(a7.z);
                    const float *f1 = s_fc + b1 * 6;
                    const float l = lam[r];
                    float d = a1.z - l * a1.w;
                    d -= f1[0] * a0.x + f1[1] * a0.y + f1[2] * a0.z + f1[3] * a0.w + f1[4] * a1.x + f1[5] * a1.y;
                    if (b2 >= 0) {
                        const float *f2 = s_fc + b2 * 6;
                        d -= f2[0] * a2.x + f2[1] * a2.y + f2[2] * a2.z + f2[3] * a2.w + f2[4] * a3.x + f2[5] * a3.y;
                    }
                    d *= iw;
                    float lo = a3.z, hi = a3.w;
                    if (fidx >= 0) { hi = fabsf(hi * lam[fidx]); lo = -hi; }
                    const float nl = fminf(fmaxf(l + d, lo), hi);
                    mx = fmaxf(mx, fabsf(nl - l));
                    lmax = fmaxf(lmax, fabsf(l));
                }
                mx = block_max(mx, s_red);
                lmax = block_max(lmax, s_red);
                res = mx / fmaxf(lmax, P.lambda_floor);
                __syncthreads();
                if (res <= P.tol) break;
            }
        }
    } else {
        res = -1.f;
    }
    if (t == 0) { D.niters[0] = it; D.residual[0] = res; }
    if (STAGED) {
        for (int r = t; r < m; r += NT) D.lambda[r] = lam[r];
        __syncthreads();
    }
    // joint feedback: f = J^T lambda with the unscaled J (J_row / Ad)
    if (P.feedback) {
        for (int j = t; j < nj; j += NT) {
            const float *rec = D.joints + (size_t)j * JOINT_STRIDE;
            const int ri = D.joff[j];
            const int ty = (int)rec[0];
            const int mj = ty == 0 ? contact_rows_m(D.surf[(int)rec[13]])
                         : ty == 1 ? 3
                         : ty == JT_FIXED ? 6
                                          : 5 + (D.jext[(size_t)j * JEXT_STRIDE + JX_LIMOT_ROW] != 0.f ? 1 : 0);
            float f[12] = {0.f};
            for (int q = 0; q < mj; q++) {
                const float *rw = D.rows + (size_t)(ri + q) * ROW_STRIDE;
                const float sc = D.lambda[ri + q] / rw[R_AD];
                for (int k = 0; k < 6; k++) { f[k] += rw[R_J1 + k] * sc; f[6 + k] += rw[R_J2 + k] * sc; }
            }
            for (int k = 0; k < 12; k++) D.fb[(size_t)j * 12 + k] = f[k];
        }
    }
    __syncthreads();
    for (int i = t; i < nb; i += NT) integrate_body(D, P, i, s_fc + i * 6, m > 0, D.facc, D.tacc);
}

// The single-world step of island groups on one workgroup (one wave) each
// (sweep mode 3 with a fixed sweep count): islands share no body, so each
// group's bodies, rows, sweeps and integration need nothing from another
// group, and the groups run on separate CUs with no grid barrier.  A group is
// a run of consecutive islands, so it owns the consecutive bodies
// [grp_body_start[g], grp_body_start[g + 1]), joints [grp_joint_start[g], ..)
// and their rows; the phase table opens with G + 1 offsets of the groups'
// level lists, whose items are consecutive in the schedule.  LDS: (ROWS: the
// run's row records,) fc of the group's bodies and lambda of its rows, then
// its items and level starts.
template <bool ROWS>
__global__ __launch_bounds__(64) void k_single_groups(DevWorlds D, StepParams P) {
    extern __shared__ float4 s_dyn4[];
    const int g = blockIdx.x, t = threadIdx.x, G = P.wave_groups;
    const int nj = D.njoints[0], m = D.nrows[0];
    const int *gph = D.phase_start + (G + 1) + D.phase_start[g];
    const int Lg = D.phase_start[g + 1] - D.phase_start[g] - 1;
    const int gs = gph[0], ni = gph[Lg] - gs;
    const int b0 = D.grp_body_start[g], b1 = D.grp_body_start[g + 1];
    const int j0 = D.grp_joint_start[g], j1 = D.grp_joint_start[g + 1];
    const int r0 = j0 < nj ? D.joff[j0] : m, r1 = j1 < nj ? D.joff[j1] : m;
    float4 *s_rows = s_dyn4;
    float *s_fc = reinterpret_cast<float *>(s_dyn4 + (ROWS ? (size_t)(r1 - r0) * (ROW_STRIDE / 4) : 0));
    float *lam = s_fc + (size_t)(b1 - b0) * 6;
    int *s_sched = reinterpret_cast<int *>(lam + (r1 - r0));
    int *s_ph = s_sched + ni;
    for (int i = b0 + t; i < b1; i += 64) {
        body_prep(D, P, i, D.facc, D.tacc);
#pragma unroll
        for (int c = 0; c < 6; c++) s_fc[(i - b0) * 6 + c] = 0.f;
    }
    for (int k = t; k < ni; k += 64) s_sched[k] = D.sched[gs + k];
    for (int k = t; k <= Lg; k += 64) s_ph[k] = gph[k] - gs;
    for (int r = r0 + t; r < r1; r += 64) lam[r - r0] = 0.f;
    __syncthreads();
    for (int j = j0 + t; j < j1; j += 64)
        build_joint_rows(D, P, D.joints + (size_t)j * JOINT_STRIDE, D.jext ? D.jext + (size_t)j * JEXT_STRIDE : nullptr,
                         0, 0, D.joff[j]);
    // the rows this workgroup wrote are read back by it below
    __syncthreads();
    if (ROWS) {
        const float4 *g4 = reinterpret_cast<const float4 *>(D.rows) + (size_t)r0 * (ROW_STRIDE / 4);
        for (int k = t; k < (r1 - r0) * (ROW_STRIDE / 4); k += 64) s_rows[k] = g4[k];
        __syncthreads();
    }
    if (Lg > 0 && ni > 0) {
        const float4 *rows4 = ROWS ? s_rows : reinterpret_cast<const float4 *>(D.rows);
        for (int it = 0; it < P.max_iters; it++)
            lean_levels_fn(
                rows4, s_sched, ni, lam, s_fc, t, 64, s_ph, Lg,
                [] {
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                },
                b0, r0, ROWS ? r0 : 0);
    }
    __syncthreads();
    for (int r = r0 + t; r < r1; r += 64) D.lambda[r] = lam[r - r0];
    // joint feedback: f = J^T lambda with the unscaled J (J_row / Ad)
    if (P.feedback) {
        for (int j = j0 + t; j < j1; j += 64) {
            const int ri = D.joff[j], re = j + 1 < nj ? D.joff[j + 1] : m;
            float f[12] = {0.f};
            for (int r = ri; r < re; r++) {
                const float *rw = D.rows + (size_t)r * ROW_STRIDE;
                const float sc = lam[r - r0] / rw[R_AD];
                for (int c = 0; c < 6; c++) { f[c] += rw[R_J1 + c] * sc; f[6 + c] += rw[R_J2 + c] * sc; }
            }
            for (int c = 0; c < 12; c++) D.fb[(size_t)j * 12 + c] = f[c];
        }
    }
    for (int i = b0 + t; i < b1; i += 64) integrate_body(D, P, i, s_fc + (i - b0) * 6, m > 0, D.facc, D.tacc);
    if (g == 0 && t == 0) { D.niters[0] = m > 0 ? P.max_iters : 0; D.residual[0] = -1.f; }
}

// -------------------------------------------- collision entry points -------

// AABBs of geom views (computeAABB of each class).
__global__ void k_aabbs(const float *gviews, int n, float *aabb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    DGeom G;
    load_gview(gviews + (size_t)i * GVIEW_STRIDE, G);
    geom_aabb(G, aabb + (size_t)i * 6);
}

// dCollide for a list of geom-view pairs: up to maxc contacts per pair
// (pos[3] normal[3] depth) into out[pair][maxc][7], counts[pair].
__global__ __launch_bounds__(128) void k_collide_pairs(const float *gviews, const int *pairs, int np, int maxc,
                                                        float *out, int *counts) {
    // the box-box clipper's arrays, one 64-column block per wave
    __shared__ float s_clip[2 * 64 * CLIP_WS];
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    DGeom A, B;
    load_gview(gviews + (size_t)pairs[2 * p] * GVIEW_STRIDE, A);
    load_gview(gviews + (size_t)pairs[2 * p + 1] * GVIEW_STRIDE, B);
    DContact c[ODX_MAXC];
    ClipLds ws{s_clip + (threadIdx.x >> 6) * 64 * CLIP_WS + (threadIdx.x & 63)};
    const int n = dcollide(A, B, min(maxc, ODX_MAXC), c, ws);
    counts[p] = n;
#pragma unroll
    for (int q = 0; q < ODX_MAXC; q++) {
        if (q >= n) break;
        float *o = out + ((size_t)p * ODX_MAXC + q) * 7;
        o[0] = c[q].pos[0]; o[1] = c[q].pos[1]; o[2] = c[q].pos[2];
        o[3] = c[q].normal[0]; o[4] = c[q].normal[1]; o[5] = c[q].normal[2];
        o[6] = c[q].depth;
    }
}

}  // namespace odx

// Host-side launchers (the kernels' only callers).
namespace odx {
void launch_prepare_small(const DevWorlds &D, const StepParams &P, hipStream_t s) {
    if (P.wcount <= 0) return;
    hipLaunchKernelGGL(k_prepare_small, dim3(P.wcount), dim3(64), 0, s, D, P);
}
size_t solve_small_lds_bytes(const StepParams &P) {
    const size_t lds = (size_t)solve_lds_floats(P.lds_bodies, P.lds_rows) * sizeof(float);
    const size_t fc = (size_t)SW_MAXB * 6 * sizeof(float);
    return lds > fc ? lds : fc;
}
void launch_solve_small(const DevWorlds &D, const StepParams &P, hipStream_t s) {
    if (P.wcount <= 0) return;
    if (D.order) hipLaunchKernelGGL(k_order_worlds, dim3(1), dim3(1024), 0, s, D.niters, P.wbase, P.wcount, D.order);
    hipLaunchKernelGGL(k_solve_small, dim3(P.wcount), dim3(64), solve_small_lds_bytes(P), s, D, P);
}
size_t step_fused_lds_bytes(const StepParams &P) {
    size_t b = solve_small_lds_bytes(P);
    if (PREP_LDS_BYTES > b) b = PREP_LDS_BYTES;
    if (sizeof(InferBuf) > b) b = sizeof(InferBuf);
    return b;
}
// P.feat_local must be 1 (features in world-local slots); M may be null.
void launch_step_fused(const DevWorlds &D, const StepParams &P, const MlpDev *M, int nsteps, hipStream_t s) {
    if (P.wcount <= 0 || nsteps <= 0) return;
    if (D.order) hipLaunchKernelGGL(k_order_worlds, dim3(1), dim3(1024), 0, s, D.niters, P.wbase, P.wcount, D.order);
    const MlpDev none{};
    hipLaunchKernelGGL(k_step_fused, dim3(P.wcount), dim3(64), step_fused_lds_bytes(P), s, D, P, M ? *M : none,
                       M ? 1 : 0, nsteps);
}
}  // namespace odx

namespace odx {
// The staged kernel when fc, lambda, the items and the phase starts fit the
// device's LDS per workgroup; otherwise the global-memory kernel.
// Largest dynamic LDS of one workgroup on the current device (bytes).
int single_lds_max() {
    static int lds_max = -1;
    if (lds_max < 0) {
        int dev = 0, v = 0;
        lds_max = (hipGetDevice(&dev) == hipSuccess &&
                   hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess)
                      ? v
                      : 65536;
    }
    return lds_max;
}

// k_single_groups: P.wave_groups workgroups; lds = bytes of the largest
// group's (rows,) fc, lambda, items and level starts.  False if the LDS does
// not fit.
bool launch_single_groups(const DevWorlds &D, const StepParams &P, size_t lds, bool rows_lds, hipStream_t s) {
    if (lds > (size_t)single_lds_max()) return false;
    const void *fn = rows_lds ? (const void *)k_single_groups<true> : (const void *)k_single_groups<false>;
    if (lds > 65536 && hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    if (rows_lds)
        hipLaunchKernelGGL(k_single_groups<true>, dim3(P.wave_groups), dim3(64), lds, s, D, P);
    else
        hipLaunchKernelGGL(k_single_groups<false>, dim3(P.wave_groups), dim3(64), lds, s, D, P);
    return true;
}

void launch_single_step(const DevWorlds &D, const StepParams &P0, int nb, int m, int nitems, int nphases,
                        hipStream_t s) {
    // sweep mode: 1 the whole block per level, 2 the first wave alone, 0 the
    // wave when levels hold at most 64 items on average
    StepParams P = P0;
    P.wave_sweep = P0.wave_sweep == 2 || (P0.wave_sweep == 0 && nitems <= 64 * (nphases > 0 ? nphases : 1)) ? 1 : 0;
    const int lds_max = single_lds_max();
    const size_t dynw = P.dyn ? (size_t)nb * 3 + (size_t)P.nislands * 5 : 0;  // fa, island of body, island state
    const size_t staged = ((size_t)nb * 6 + m + nitems + nphases + 1 + dynw) * sizeof(float);
    // rows in LDS as well when they fit next to the staged arrays
    const size_t with_rows = staged + (size_t)m * ROW_STRIDE * sizeof(float);
    if (!D.seg_table && with_rows + 256 <= (size_t)lds_max &&
        hipFuncSetAttribute((const void *)k_single_step_t<true, 512, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)with_rows) == hipSuccess) {
        hipLaunchKernelGGL((k_single_step_t<true, 512, true>), dim3(1), dim3(512), with_rows, s, D, P);
        return;
    }
    (void)hipGetLastError();
    if (!D.seg_table && staged + 256 <= (size_t)lds_max &&
        hipFuncSetAttribute((const void *)k_single_step_t<true, 512>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)staged) == hipSuccess) {
        hipLaunchKernelGGL((k_single_step_t<true, 512>), dim3(1), dim3(512), staged, s, D, P);
        return;
    }
    (void)hipGetLastError();
    const size_t plain = ((size_t)nb * 6 + dynw) * sizeof(float);
    if (plain > 65536)
        (void)hipFuncSetAttribute((const void *)k_single_step_t<false, SL_BLOCK>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)plain);
    hipLaunchKernelGGL((k_single_step_t<false, SL_BLOCK>), dim3(1), dim3(SL_BLOCK), plain, s, D, P);
}
void launch_aabbs(const float *gviews, int n, float *aabb, hipStream_t s) {
    hipLaunchKernelGGL(k_aabbs, dim3((n + 255) / 256), dim3(256), 0, s, gviews, n, aabb);
}
void launch_collide_pairs(const float *gviews, const int *pairs, int np, int maxc, float *out, int *counts,
                          hipStream_t s) {
    if (np <= 0) return;
    hipLaunchKernelGGL(k_collide_pairs, dim3((np + 127) / 128), dim3(128), 0, s, gviews, pairs, np, maxc, out, counts);
}
}  // namespace odx

// ========================================================= one large world ===
// A single world of up to millions of bodies (hash-space broadphase).  The
// per-body, per-geom, per-pair and per-contact device functions of the batched
// path run one thread per item over the whole world; colouring and the solve
// are cooperative persistent kernels whose phases are separated by grid
// barriers.  The solve sweeps colour phases (contacts whose bodies are
// disjoint), rows streamed from HBM, body accumulators fc in global memory.
#include "odx_large.h"

namespace odx {

__device__ __forceinline__ int lg_hash(int x, int y, int z, int mask) {
    unsigned h = (unsigned)x * 73856093u ^ (unsigned)y * 19349663u ^ (unsigned)z * 83492791u;
    h ^= h >> 15; h *= 0x2c1b3c6du; h ^= h >> 12;
    return (int)(h & (unsigned)mask);
}

__global__ void k_lg_fill(int *p, int n, int v) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// Exclusive scan of one 1024-thread block's values; returns the exclusive
// prefix of v, *total gets the block sum.
__device__ int lg_block_scan(int v, int *s_tmp, int *total) {
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6, nw = blockDim.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_tmp[wv] = x;
    __syncthreads();
    if (t < 64) {
        int w = t < nw ? s_tmp[t] : 0;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(w, o, 64);
            if (lane >= o) w += y;
        }
        if (t < nw) s_tmp[t] = w;
    }
    __syncthreads();
    const int r = (wv > 0 ? s_tmp[wv - 1] : 0) + x - v;
    *total = s_tmp[nw - 1];
    __syncthreads();
    return r;
}

// In-place exclusive scan of data[0, n), n = min(*dn + 1, cap) when dn is
// given (a count the device wrote; entry *dn is zero), else cap.  Blocks of
// 2048 entries; k_lg_scan_sums scans the block sums in chunks of 2048.
__device__ __forceinline__ int lg_scan_n(const int *dn, int cap) { return dn ? min(*dn + 1, cap) : cap; }

__global__ __launch_bounds__(1024) void k_lg_scan_blocks(int *data, int cap, const int *dn, int *sums) {
    __shared__ int s_tmp[64];
    const int n = lg_scan_n(dn, cap);
    const int i0 = blockIdx.x * 2048 + 2 * threadIdx.x;
    const int a = i0 < n ? data[i0] : 0, b = i0 + 1 < n ? data[i0 + 1] : 0;
    int total;
    const int ex = lg_block_scan(a + b, s_tmp, &total);
    if (i0 < n) data[i0] = ex;
    if (i0 + 1 < n) data[i0 + 1] = ex + a;
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(1024) void k_lg_scan_sums(int *sums, int nblk) {
    __shared__ int s_tmp[64];
    int carry = 0;
    for (int base = 0; base < nblk; base += 2048) {
        const int i0 = base + 2 * threadIdx.x;
        const int a = i0 < nblk ? sums[i0] : 0, b = i0 + 1 < nblk ? sums[i0 + 1] : 0;
        int total;
        const int ex = lg_block_scan(a + b, s_tmp, &total);
        if (i0 < nblk) sums[i0] = carry + ex;
        if (i0 + 1 < nblk) sums[i0 + 1] = carry + ex + a;
        carry += total;
    }
}

__global__ __launch_bounds__(1024) void k_lg_scan_add(int *data, int cap, const int *dn, const int *sums) {
    const int n = lg_scan_n(dn, cap);
    const int add = sums[blockIdx.x];
    const int i0 = blockIdx.x * 2048 + 2 * threadIdx.x;
    if (i0 < n) data[i0] += add;
    if (i0 + 1 < n) data[i0 + 1] += add;
}

// scan of at most cap entries (the grid covers cap; blocks past n do nothing)
void lg_scan(int *data, int cap, const int *dn, int *sums, hipStream_t s) {
    const int nblk = (cap + 2047) / 2048;
    hipLaunchKernelGGL(k_lg_scan_blocks, dim3(nblk), dim3(1024), 0, s, data, cap, dn, sums);
    hipLaunchKernelGGL(k_lg_scan_sums, dim3(1), dim3(1024), 0, s, sums, nblk);
    hipLaunchKernelGGL(k_lg_scan_add, dim3(nblk), dim3(1024), 0, s, data, cap, dn, sums);
}

__global__ void k_lg_bodies(DevWorlds D, StepParams P, int nb, float4 *ftot) {
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x)
        body_prep(D, P, b, ftot, ftot + nb);
}

// AABBs; each finite geom gets the cell of its AABB centre and counts into
// that cell's hash bucket
__global__ void k_lg_aabb(DevWorlds D, LargeDev L) {
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < L.ng; g += gridDim.x * blockDim.x) {
        DGeom G;
        load_geom(D, g, 0, G);
        float ab[6];
        geom_aabb(G, ab);
        for (int k = 0; k < 6; k++) L.aabb[(size_t)g * 6 + k] = ab[k];
        if (L.isbig[g]) continue;
        const int cx = (int)floorf(0.5f * (ab[0] + ab[1]) * L.inv_cell);
        const int cy = (int)floorf(0.5f * (ab[2] + ab[3]) * L.inv_cell);
        const int cz = (int)floorf(0.5f * (ab[4] + ab[5]) * L.inv_cell);
        const int h = lg_hash(cx, cy, cz, L.hmask);
        L.gcell[g] = make_int4(cx, cy, cz, h);
        atomicAdd(&L.bucket[h], 1);
    }
}

// after the exclusive scan bucket[h] is h's start; the scatter advances it to
// h's end, so bucket h spans [h ? bucket[h - 1] : 0, bucket[h])
__global__ void k_lg_scatter(LargeDev L) {
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < L.ng; g += gridDim.x * blockDim.x) {
        if (L.isbig[g]) continue;
        L.bucket_geom[atomicAdd(&L.bucket[L.gcell[g].w], 1)] = g;
    }
}

// geoms of a bucket in index order (deterministic pair order), then each
// slot's copy of its geom's cell, AABB and body, so that the pair search
// reads a candidate with two loads that do not wait for its geom index
__global__ void k_lg_sort_buckets(DevWorlds D, LargeDev L) {
    for (int h = blockIdx.x * blockDim.x + threadIdx.x; h <= L.hmask; h += gridDim.x * blockDim.x) {
        const int s = h ? L.bucket[h - 1] : 0, e = L.bucket[h];
        for (int i = s + 1; i < e; i++) {
            const int v = L.bucket_geom[i];
            int j = i - 1;
            while (j >= s && L.bucket_geom[j] > v) { L.bucket_geom[j + 1] = L.bucket_geom[j]; j--; }
            L.bucket_geom[j + 1] = v;
        }
        for (int k = s; k < e; k++) {
            const int g = L.bucket_geom[k];
            const int4 c = L.gcell[g];
            const float *a = L.aabb + (size_t)g * 6;
            L.slot_cell[k] = make_int4(c.x, c.y, c.z, g);
            L.slot_box[2 * (size_t)k] = make_float4(a[0], a[1], a[2], a[3]);
            L.slot_box[2 * (size_t)k + 1] = make_float4(a[4], a[5], D.geom0[g].y, 0.f);
        }
    }
}

__device__ __forceinline__ bool lg_overlap(const float *x, const float *y) {
    return !(x[0] > y[1] || x[1] < y[0] || x[2] > y[3] || x[3] < y[2] || x[4] > y[5] || x[5] < y[4]);
}

// pair filter of collideAABBs (collision_space_internal.h:44-78): not the same
// body; pairs of two static geoms make no rows (their joint is in no island)
// and are left out, as in the batched path
__device__ __forceinline__ bool lg_bodies_ok(int bi, int bj) { return (bi >= 0 || bj >= 0) && bi != bj; }

// Candidate pairs of the hash space (collision_space.cpp:421-614: every pair
// whose closed AABBs overlap).  From each finite geom i: geoms j > i of the 27
// cells around it (the cell edge bounds every finite AABB, so the centres of
// overlapping AABBs lie in adjacent cells; a bucket's geoms count only when
// their cell matches exactly), then every big geom.  A big geom i pairs with
// the big geoms after it.  write = 0 counts into pair_cnt[i]; write = 1 stores
// at the scanned offsets.
// (cat_filter: the category / collide bit test of collideAABBs, GEOM slots
// 14 and 15, collision_space_internal.h:50-57)
__device__ __forceinline__ bool lg_bits_ok(const DevWorlds &D, int i, int j) {
    const unsigned ci = (unsigned)D.geom1[i].w, li = (unsigned)D.gpos[i].w;
    const unsigned cj = (unsigned)D.geom1[j].w, lj = (unsigned)D.gpos[j].w;
    return ((ci & lj) | (cj & li)) != 0u;
}

__global__ void k_lg_pairs(DevWorlds D, LargeDev L, int write, int cat_filter) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < L.ng; i += gridDim.x * blockDim.x) {
        const float *ai = L.aabb + (size_t)i * 6;
        const int bi = (int)D.geom0[i].y;
        const bool ibig = L.isbig[i] != 0;
        const int off = write ? L.pair_cnt[i] : 0;
        int *keep = L.pair_keep + (size_t)i * LG_PAIR_KEEP;
        if (write) {
            // the counting pass kept the partners, in the order found, when
            // they fit: copy them instead of searching again
            const int cnt = L.pair_cnt[i + 1] - off;
            if (cnt <= LG_PAIR_KEEP) {
                for (int q = 0; q < cnt && off + q < L.pair_cap; q++) {
                    const int j = keep[q];
                    L.pairs[off + q] = make_int2(min(i, j), max(i, j));
                }
                continue;
            }
        }
        int n = 0;
        if (!ibig) {
            const int4 c = L.gcell[i];
            // the 27 cells' bucket ranges first (z outer, x inner), loaded together
            int bs[27], be[27];
#pragma unroll
            for (int t = 0; t < 27; t++) {
                const int h = lg_hash(c.x + t % 3 - 1, c.y + (t / 3) % 3 - 1, c.z + t / 9 - 1, L.hmask);
                bs[t] = h ? L.bucket[h - 1] : 0;
                be[t] = L.bucket[h];
            }
            float a[6];
            for (int q = 0; q < 6; q++) a[q] = ai[q];
#pragma unroll
            for (int t = 0; t < 27; t++) {
                const int x = c.x + t % 3 - 1, y = c.y + (t / 3) % 3 - 1, z = c.z + t / 9 - 1;
                for (int k = bs[t]; k < be[t]; k++) {
                    const int4 sc = L.slot_cell[k];
                    const float4 b0 = L.slot_box[2 * (size_t)k], b1 = L.slot_box[2 * (size_t)k + 1];
                    const int j = sc.w;
                    if (j <= i) continue;
                    if (sc.x != x || sc.y != y || sc.z != z) continue;
                    if (!lg_bodies_ok(bi, (int)b1.z)) continue;
                    if (cat_filter && !lg_bits_ok(D, i, j)) continue;
                    const float bj[6] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y};
                    if (!lg_overlap(a, bj)) continue;
                    if (write && off + n < L.pair_cap) L.pairs[off + n] = make_int2(i, j);
                    if (!write && n < LG_PAIR_KEEP) keep[n] = j;
                    n++;
                }
            }
        }
        for (int q = 0; q < L.nbig; q++) {
            const int j = L.big[q];
            if (ibig && j <= i) continue;
            if (!lg_bodies_ok(bi, (int)D.geom0[j].y)) continue;
            if (cat_filter && !lg_bits_ok(D, i, j)) continue;
            if (!lg_overlap(ai, L.aabb + (size_t)j * 6)) continue;
            if (write && off + n < L.pair_cap) L.pairs[off + n] = make_int2(min(i, j), max(i, j));
            if (!write && n < LG_PAIR_KEEP) keep[n] = j;
            n++;
        }
        if (!write) L.pair_cnt[i] = n;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (!write) {
            L.pair_cnt[L.ng] = 0;
        } else {
            const int np = L.pair_cnt[L.ng];
            L.counters[LG_PAIRS] = min(np, L.pair_cap);
            if (np > L.pair_cap) L.counters[LG_PAIR_OVF] = np - L.pair_cap;
        }
    }
}

// dCollide of every pair (collision_kernel.cpp:292-329); contacts become JOINT
// records, the two bodies swapped and the normal negated when body 1 is
// static (dJointAttach ode.cpp:1383-1439, contact.cpp:160-167).
__global__ void k_lg_contacts(DevWorlds D, StepParams P, LargeDev L, int write) {
    const int np = L.counters[LG_PAIRS];
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p <= np; p += gridDim.x * blockDim.x) {
        if (p == np) {
            if (!write) L.contact_cnt[np] = 0;
            continue;
        }
        const int2 pr = L.pairs[p];
        DGeom A, B;
        load_geom(D, pr.x, 0, A);
        load_geom(D, pr.y, 0, B);
        DContact c[ODX_MAXC];
        int maxc = P.maxc, si = 0;
        if (P.nmat > 0) {  // the material pair's surface and contact limit
            si = (int)D.geom1[pr.x].z * P.nmat + (int)D.geom1[pr.y].z;
            if (D.surf[si].max_contacts > 0) maxc = D.surf[si].max_contacts;
        }
        ClipPriv ws;
        const int n = (A.body < 0 && B.body < 0) ? 0 : dcollide(A, B, min(maxc, ODX_MAXC), c, ws);
        if (!write) { L.contact_cnt[p] = n; continue; }
        const int off = L.contact_cnt[p];
        int b1 = A.body, b2 = B.body;
        float sg = 1.f;
        if (b1 < 0) { b1 = b2; b2 = -1; sg = -1.f; }
#pragma unroll
        for (int q = 0; q < ODX_MAXC; q++) {
            if (q >= n || off + q >= L.contact_cap) break;
            float4 *o4 = reinterpret_cast<float4 *>(D.joints + (size_t)(off + q) * JOINT_STRIDE);
            o4[0] = make_float4(0.f, (float)b1, (float)b2, c[q].pos[0]);
            o4[1] = make_float4(c[q].pos[1], c[q].pos[2], sg * c[q].normal[0], sg * c[q].normal[1]);
            o4[2] = make_float4(sg * c[q].normal[2], c[q].depth, 0.f, 0.f);
            o4[3] = make_float4(0.f, (float)si, (float)pr.x, (float)pr.y);
            L.cbody[off + q] = make_int2(b1, b2);
        }
    }
    if (write && blockIdx.x == 0 && threadIdx.x == 0) {
        const int nc = L.contact_cnt[np];
        L.counters[LG_CONTACTS] = min(nc, L.contact_cap);
        if (nc > L.contact_cap) L.counters[LG_CONTACT_OVF] = nc - L.contact_cap;
    }
}

__global__ void k_lg_rowcount(DevWorlds D, LargeDev L) {
    const int nc = L.counters[LG_CONTACTS];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k <= nc; k += gridDim.x * blockDim.x)
        L.row_cnt[k] = k < nc ? contact_rows_m(D.surf[(int)D.joints[(size_t)k * JOINT_STRIDE + 13]]) : 0;
}

// getInfo2, Stage2a/2c, compute_invM_JT and Ad of every contact, at the
// contact's scanned first row (rows of contact k are consecutive; a friction
// row's findex points at its normal row)
__global__ void k_lg_rows(DevWorlds D, StepParams P, LargeDev L) {
    const int nc = L.counters[LG_CONTACTS];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nc; k += gridDim.x * blockDim.x)
        build_joint_rows(D, P, D.joints + (size_t)k * JOINT_STRIDE, nullptr, 0, 0, L.row_cnt[k]);
    if (blockIdx.x == 0 && threadIdx.x == 0) L.counters[LG_ROWS] = L.row_cnt[nc];
}

// Values shared between workgroups inside one cooperative kernel are stored
// and loaded with agent-scope atomics (global_store/load ... sc1: coherent at
// the device level across the XCDs' L2 caches), so a barrier needs no L2
// writeback or invalidation, only the completion of each thread's stores.
__device__ __forceinline__ float ld_dev(const float *p) {
    return __hip_atomic_load(const_cast<float *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int ld_dev(const int *p) {
    return __hip_atomic_load(const_cast<int *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned ld_dev(const unsigned *p) {
    return __hip_atomic_load(const_cast<unsigned *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long ld_dev(const unsigned long long *p) {
    return __hip_atomic_load(const_cast<unsigned long long *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T> __device__ __forceinline__ void st_dev(T *p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned add_dev(unsigned *p, unsigned v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// all of this thread's memory operations have completed (stores acknowledged
// at their coherence level); also a compiler barrier
__device__ __forceinline__ void mem_done() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_s_waitcnt(0);
}

// Grid barrier of the cooperative kernels.  Workgroups count per residue
// class blockIdx % 8 (one XCD each under round-robin dispatch; correctness
// does not depend on it); the last of a class adds one to the generation word,
// and every workgroup waits until that word has counted all classes.  The
// counters are zeroed before each launch and only grow inside it: at barrier g
// a class is complete when its counter reaches per * (g + 1), and the barrier
// when the generation word reaches classes * (g + 1), so no reset store (and
// no wait for one) sits on the barrier's path.  Counters sit 64 B apart.  A
// wait that runs out sets LG_BARRIER_ERR and bit 31 of the generation word;
// every block sees the bit at its next barrier, leaves with false, and the
// host reports the step as failed.  `read`, when given, is loaded by one
// thread after the barrier and returned in *val.
constexpr unsigned LG_GEN_ERR = 0x80000000u;
__device__ bool lg_grid_sync(const LargeDev &L, unsigned &gen, const int *read = nullptr, int *val = nullptr) {
    __shared__ int s_ok, s_val;
    mem_done();
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned x = blockIdx.x & 7u;
        const unsigned per = (gridDim.x + 7u - x) >> 3;
        const unsigned groups = gridDim.x < 8u ? gridDim.x : 8u;
        const unsigned target = groups * (gen + 1u);
        unsigned seen = 0;
        if (add_dev(&L.bar[LG_BAR_GROUP + 16 * x], 1u) + 1u == per * (gen + 1u))
            seen = add_dev(&L.bar[LG_BAR_GEN], 1u) + 1u;
        long long spins = 0;
        while (seen < target) {   // an error bit makes seen >= target
            if (++spins > L.spin_limit) {
                atomicExch(&L.counters[LG_BARRIER_ERR], 1);
                __hip_atomic_fetch_or(&L.bar[LG_BAR_GEN], LG_GEN_ERR, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                seen = LG_GEN_ERR;
                break;
            }
            if (spins > 1) __builtin_amdgcn_s_sleep(1);
            seen = ld_dev(&L.bar[LG_BAR_GEN]);
        }
        gen++;
        mem_done();
        if (read) s_val = ld_dev(read);
        s_ok = (seen & LG_GEN_ERR) ? 0 : 1;
    }
    __syncthreads();
    if (read) *val = s_val;
    return s_ok != 0;
}

// sum over the block (every thread gets it)
__device__ int lg_block_sum(int v, int *s_red) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o; o >>= 1) v += __shfl_xor(v, o, 64);
    __syncthreads();
    if (lane == 0) s_red[wv] = v;
    __syncthreads();
    int s = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) s += s_red[i];
    return s;
}

// exclusive prefix of v over the block in thread order, and the block total
__device__ void lg_block_scan(int v, int *s_red, int &pre, int &tot) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(x, o, 64);
        if (lane >= o) x += t;
    }
    __syncthreads();
    if (lane == 63) s_red[wv] = x;
    __syncthreads();
    int base = 0, all = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) {
        if (i < wv) base += s_red[i];
        all += s_red[i];
    }
    pre = base + x - v;
    tot = all;
}

// Rows of a contact in the first pass of ReorderPrep (quickstep.cpp:2329-2355:
// rows without findex first, findex rows after), by the batched path's rule:
// n_a leading rows without findex followed only by findex rows split there;
// any other pattern keeps all n rows in the first pass.
__device__ __forceinline__ int lg_pass_a(const DevWorlds &D, int r, int n) {
    unsigned char fl[3] = {0, 0, 0};
    for (int q = 0; q < n && q < 3; q++) fl[q] = as_int(D.rows[(size_t)(r + q) * ROW_STRIDE + R_FINDEX]) >= 0;
    int na, nf;
    split_joint_rows(fl, n, na, nf);
    return nf > 0 ? na : n;
}

// Claim key of contact k in claim round `stamp`: rounds increase, so a claim
// always beats those left by earlier rounds and owner[] needs no reset.
__device__ __forceinline__ unsigned long long lg_key(int k, int stamp) {
    unsigned h = (unsigned)k * 2654435761u ^ (unsigned)(stamp + 1) * 0x9e3779b9u;
    h ^= h >> 16; h *= 0x7feb352du; h ^= h >> 15;
    return ((unsigned long long)stamp << 50) | ((unsigned long long)(h & 0x3ffffffu) << 24) |
           (unsigned long long)(k & 0xffffff);
}

constexpr int LG_CCACHE = 8;   // contacts per thread the colouring keeps in registers

// Colouring (cooperative): colour c is a maximal set of contacts with
// disjoint dynamic bodies, built by claim/commit rounds (Luby): every
// uncoloured contact whose bodies no contact of colour c holds claims both
// with its key (atomicMax); one holding both claims joins colour c.  Then the
// contacts are grouped by colour: corder[col_start[c] ..) are colour c's.
// Keys hash the contact index (< 2^24) and the round, so the result is
// deterministic.
__global__ __launch_bounds__(LG_BLOCK) void k_lg_colour(DevWorlds D, LargeDev L) {
    __shared__ int s_red[LG_BLOCK / 64];
    __shared__ int s_hist[LG_MAXCOL], s_base[LG_MAXCOL], s_start[LG_MAXCOL + 1];
    unsigned gen = 0;
    const int nc = L.counters[LG_CONTACTS];
    const int gt = blockIdx.x * blockDim.x + threadIdx.x, gs = gridDim.x * blockDim.x;
    for (int k = gt; k < nc; k += gs) L.colour[k] = -1;
    for (int b = gt; b < L.nb; b += gs) { st_dev(&L.taken[b], -1); st_dev(&L.owner[b], 0ull); }
    bool ok = lg_grid_sync(L, gen);
    // the thread's first LG_CCACHE contacts (k = gt + i * gs) stay in
    // registers: their bodies, and a bit per contact still uncoloured
    int2 cb[LG_CCACHE];
    unsigned live = 0;
#pragma unroll
    for (int i = 0; i < LG_CCACHE; i++) {
        const int k = gt + i * gs;
        cb[i] = make_int2(0, -1);
        if (k < nc) {
            cb[i] = L.cbody[k];
            live |= 1u << i;
        }
    }
    const int kfar = gt + LG_CCACHE * gs;   // contacts past the cache take the loads
    // a colour takes at most L.colour_cap contacts (one per thread of the
    // solve's grid): each workgroup admits at most `quota` of its winners per
    // colour, in thread order, so the cap is deterministic
    const int quota = L.colour_cap > 0 ? max(1, L.colour_cap / (int)gridDim.x) : 0x7fffffff;
    int ncol = 0, rounds = 0, left = nc;
    for (int c = 0; ok && left > 0 && c < LG_MAXCOL; c++) {
        int joined = 0;   // contacts of colour c after its previous rounds
        int used = 0;     // this workgroup's contacts in colour c
        for (int sub = 0; ok && sub < LG_MAXSUB; sub++) {
            const int stamp = c * LG_MAXSUB + sub;
            int e = 0;
            // a workgroup whose quota is full makes no claims for colour c
            // (its winners could not join, and their claims would block others)
            const unsigned claim = used < quota ? live : 0u;
#pragma unroll
            for (int i = 0; i < LG_CCACHE; i++) {
                if (!((claim >> i) & 1u)) continue;
                const int k = gt + i * gs;
                const int2 bb = cb[i];
                if (ld_dev(&L.taken[bb.x]) == c || (bb.y >= 0 && ld_dev(&L.taken[bb.y]) == c)) continue;
                const unsigned long long key = lg_key(k, stamp);
                atomicMax(&L.owner[bb.x], key);
                if (bb.y >= 0) atomicMax(&L.owner[bb.y], key);
                e++;
            }
            for (int k = kfar; k < nc; k += gs) {
                if (L.colour[k] >= 0) continue;
                const int2 bb = L.cbody[k];
                if (ld_dev(&L.taken[bb.x]) == c || (bb.y >= 0 && ld_dev(&L.taken[bb.y]) == c)) continue;
                const unsigned long long key = lg_key(k, stamp);
                atomicMax(&L.owner[bb.x], key);
                if (bb.y >= 0) atomicMax(&L.owner[bb.y], key);
                e++;
            }
            e = lg_block_sum(e, s_red);
            if (threadIdx.x == 0 && e) atomicAdd(&L.elig[stamp], e);
            int elig = 0;
            ok = lg_grid_sync(L, gen, &L.elig[stamp], &elig);
            if (!ok) break;
            rounds++;
            if (elig == 0) break;
            int won = 0;
            unsigned winm = 0;   // cached contacts holding both claims
#pragma unroll
            for (int i = 0; i < LG_CCACHE; i++) {
                if (!((live >> i) & 1u)) continue;
                const int k = gt + i * gs;
                const int2 bb = cb[i];
                const unsigned long long key = lg_key(k, stamp);
                if (ld_dev(&L.owner[bb.x]) != key || (bb.y >= 0 && ld_dev(&L.owner[bb.y]) != key)) continue;
                winm |= 1u << i;
            }
            int pre, btot;
            lg_block_scan(__popc(winm), s_red, pre, btot);
            const int room = quota - used;
#pragma unroll
            for (int i = 0; i < LG_CCACHE; i++) {
                if (!((winm >> i) & 1u)) continue;
                if (pre + __popc(winm & ((1u << i) - 1u)) >= room) continue;   // the workgroup's quota is full
                const int k = gt + i * gs;
                const int2 bb = cb[i];
                L.colour[k] = c;
                live &= ~(1u << i);
                st_dev(&L.taken[bb.x], c);
                if (bb.y >= 0) st_dev(&L.taken[bb.y], c);
                won++;
            }
            used += min(btot, max(room, 0));
            for (int k = kfar; k < nc; k += gs) {
                if (L.colour[k] >= 0) continue;
                const int2 bb = L.cbody[k];
                const unsigned long long key = lg_key(k, stamp);
                if (ld_dev(&L.owner[bb.x]) != key || (bb.y >= 0 && ld_dev(&L.owner[bb.y]) != key)) continue;
                L.colour[k] = c;   // contact k belongs to this thread in every round
                st_dev(&L.taken[bb.x], c);
                if (bb.y >= 0) st_dev(&L.taken[bb.y], c);
                won++;
            }
            won = lg_block_sum(won, s_red);
            if (threadIdx.x == 0 && won) atomicAdd(&L.col_start[c], won);
            int total = 0;
            ok = lg_grid_sync(L, gen, &L.col_start[c], &total);
            // every eligible contact joined: each contact still uncoloured
            // touches a body of colour c, so another round would find none;
            // a round that added none means the quotas are full
            if (total - joined == elig || total == joined) break;
            joined = total;
        }
        if (!ok) break;
        ncol = c + 1;
        int nmem = 0;   // contacts of colour c (complete: its last round ended with a barrier)
        __syncthreads();
        if (threadIdx.x == 0) s_hist[0] = ld_dev(&L.col_start[c]);
        __syncthreads();
        nmem = s_hist[0];
        left -= nmem;
    }
    if (!ok) return;
    // group by colour: per-block histogram, block base per colour, scan
    for (int c = threadIdx.x; c < LG_MAXCOL; c += blockDim.x) s_hist[c] = 0;
    __syncthreads();
    for (int k = gt; k < nc; k += gs) {
        const int c = L.colour[k];
        if (c >= 0) atomicAdd(&s_hist[c], 1);
    }
    __syncthreads();
    for (int c = threadIdx.x; c < ncol; c += blockDim.x)
        s_base[c] = s_hist[c] ? (int)add_dev(&L.bar[LG_BAR_FILL + c], (unsigned)s_hist[c]) : 0;
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int c = 0; c < ncol; c++) {
            s_start[c] = acc;
            acc += ld_dev(&L.col_start[c]);
        }
        s_start[ncol] = acc;
    }
    __syncthreads();
    for (int c = threadIdx.x; c < LG_MAXCOL; c += blockDim.x) s_hist[c] = 0;
    __syncthreads();
    for (int k = gt; k < nc; k += gs) {
        const int c = L.colour[k];
        if (c < 0) continue;
        const int pos = s_start[c] + s_base[c] + atomicAdd(&s_hist[c], 1);
        const int r = L.row_cnt[k], n = L.row_cnt[k + 1] - r;
        const int2 bb = L.cbody[k];
        L.corder[pos] = k;
        L.items[pos] = make_int4(r, (L.one_pass ? n : lg_pass_a(D, r, n)) | (n << 8), bb.x, bb.y);
    }
    ok = lg_grid_sync(L, gen);   // every block has read col_start[] counts
    if (!ok) return;
    if (blockIdx.x == 0) {
        for (int c = threadIdx.x; c <= ncol; c += blockDim.x) L.col_start[c] = s_start[c];
        if (threadIdx.x == 0) {
            L.counters[LG_NCOL] = ncol;
            L.counters[LG_SUBROUNDS] = rounds;
            if (left > 0) L.counters[LG_COLOUR_OVF] = left;
        }
    }
}

// Body accumulators fc of the solve: 8 floats per body (32 B), moved as a
// 16 B + 8 B pair of buffer accesses with the device-coherent cache policy
// (sc1), since bodies are shared between workgroups inside the kernel.  The
// buffer descriptor bounds every access to the array.
typedef __attribute__((ext_vector_type(4))) unsigned lg_u4;
typedef __attribute__((ext_vector_type(2))) unsigned lg_u2;
constexpr int LG_CPOL_DEVICE = 16;  // cache policy bits: sc1
constexpr int LG_FC_STRIDE = 8;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t lg_fc_rsrc(float *fc, int nb) {
    return __builtin_amdgcn_make_buffer_rsrc(fc, 0, nb * LG_FC_STRIDE * 4, 0x00020000);
}
__device__ __forceinline__ void lg_fc_load(__amdgpu_buffer_rsrc_t rs, int b, float *f) {
    const float4 a = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rs, b * 32, 0, LG_CPOL_DEVICE));
    const float2 c = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(rs, b * 32 + 16, 0, LG_CPOL_DEVICE));
    f[0] = a.x; f[1] = a.y; f[2] = a.z; f[3] = a.w; f[4] = c.x; f[5] = c.y;
}
__device__ __forceinline__ void lg_fc_store(__amdgpu_buffer_rsrc_t rs, int b, const float *f) {
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(lg_u4, make_float4(f[0], f[1], f[2], f[3])), rs, b * 32, 0,
                                           LG_CPOL_DEVICE);
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(lg_u2, make_float2(f[4], f[5])), rs, b * 32 + 16, 0,
                                          LG_CPOL_DEVICE);
}

// One row update (IterationStep quickstep.cpp:2917-3033) on a row held in
// registers; lf = lambda of the row's findex row (unused without findex).
__device__ __forceinline__ float lg_row_update(const float4 *a, float old, float lf, float *f1, float *f2, bool two) {
#pragma clang fp contract(fast)
    float delta = a[1].z - old * a[1].w;
    delta -= f1[0] * a[0].x + f1[1] * a[0].y + f1[2] * a[0].z + f1[3] * a[0].w + f1[4] * a[1].x + f1[5] * a[1].y;
    if (two) delta -= f2[0] * a[2].x + f2[1] * a[2].y + f2[2] * a[2].z + f2[3] * a[2].w + f2[4] * a[3].x + f2[5] * a[3].y;
    float lo = a[3].z, hi = a[3].w;
    if (as_int(a[7].z) >= 0) { hi = fabsf(hi * lf); lo = -hi; }
    float nl = old + delta;
    if (nl < lo) { delta = lo - old; nl = lo; }
    else if (nl > hi) { delta = hi - old; nl = hi; }
    f1[0] += delta * a[4].x; f1[1] += delta * a[4].y; f1[2] += delta * a[4].z;
    f1[3] += delta * a[4].w; f1[4] += delta * a[5].x; f1[5] += delta * a[5].y;
    if (two) {
        f2[0] += delta * a[5].z; f2[1] += delta * a[5].w; f2[2] += delta * a[6].x;
        f2[3] += delta * a[6].y; f2[4] += delta * a[6].z; f2[5] += delta * a[6].w;
    }
    return nl;
}

// A slot's work in one phase: rows [r, r + n) of one contact between bodies
// b1, b2 (-1: static).  Up to LG_PRE rows come preloaded in `pre`.  A row's
// findex points at a row of the same contact: inside [r, r + n) the value
// updated here is used, else the contact's first-pass row (fixed in this
// pass).  lambda of a contact is touched only by the thread owning its slot.
template <int PRE> struct LgSlot {
    int r, n, b1, b2;
    float4 pre[PRE][8];
};

template <int LG_PRE>
__device__ __forceinline__ void lg_slot_load(const DevWorlds &D, const int4 it, int pass, LgSlot<LG_PRE> &S) {
    const int na = it.y & 255, n = it.y >> 8;
    S.r = pass == 0 ? it.x : it.x + na;
    S.n = pass == 0 ? na : n - na;
    S.b1 = it.z;
    S.b2 = it.w;
    if (it.x < 0) S.n = 0;
#pragma unroll
    for (int q = 0; q < LG_PRE; q++) {
        if (q < S.n) {
            const float4 *rr = reinterpret_cast<const float4 *>(D.rows + (size_t)(S.r + q) * ROW_STRIDE);
#pragma unroll
            for (int k = 0; k < 8; k++) S.pre[q][k] = rr[k];
        }
    }
}

template <int LG_PRE>
__device__ __forceinline__ void lg_slot_run(const DevWorlds &D, __amdgpu_buffer_rsrc_t rs, const LgSlot<LG_PRE> &S) {
    if (S.n <= 0) return;
    const bool two = S.b2 >= 0;
    float f1[6], f2[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    lg_fc_load(rs, S.b1, f1);
    if (two) lg_fc_load(rs, S.b2, f2);
    float lam[LG_PRE], lf[LG_PRE];
#pragma unroll
    for (int q = 0; q < LG_PRE; q++) {
        lam[q] = 0.f;
        lf[q] = 0.f;
        if (q < S.n) {
            lam[q] = D.lambda[S.r + q];
            const int fidx = as_int(S.pre[q][7].z);
            if (fidx >= 0 && (fidx < S.r || fidx >= S.r + S.n)) lf[q] = D.lambda[fidx];
        }
    }
#pragma unroll
    for (int q = 0; q < LG_PRE; q++) {
        if (q < S.n) {
            const int fidx = as_int(S.pre[q][7].z);
            float l = lf[q];
#pragma unroll
            for (int u = 0; u < q; u++)
                if (fidx == S.r + u) l = lam[u];
            lam[q] = lg_row_update(S.pre[q], lam[q], l, f1, f2, two);
            D.lambda[S.r + q] = lam[q];
        }
    }
    for (int q = LG_PRE; q < S.n; q++) {   // rows beyond the preloaded ones
        const float4 *rr = reinterpret_cast<const float4 *>(D.rows + (size_t)(S.r + q) * ROW_STRIDE);
        float4 a[8];
        for (int k = 0; k < 8; k++) a[k] = rr[k];
        const float old = D.lambda[S.r + q];
        const float l = D.lambda[max(as_int(a[7].z), 0)];
        D.lambda[S.r + q] = lg_row_update(a, old, l, f1, f2, two);
    }
    lg_fc_store(rs, S.b1, f1);
    if (two) lg_fc_store(rs, S.b2, f2);
}

// Colours of at most this many contacts at the end of the colour order run on
// the first workgroup alone, one after another with workgroup barriers, in a
// single grid phase: the greedy colouring ends in a few near-empty colours
// whose grid barriers would cost as much as a full colour's.
constexpr int LG_TAIL = 256;

// The solve (cooperative): lambda = 0, `iters` sweeps, then Stage4b/6 and
// dxStepBody of every body.  A sweep is the contacts' first-pass rows colour
// by colour, then their findex rows colour by colour (two passes); with
// L.one_pass every contact's rows are one pass, colour by colour, each
// contact's rows in joint order.  The thread at a slot of the colour order
// handles that contact in every phase (its lambda is touched by no other
// thread); each thread loads its slot of the next phase (item and rows:
// read-only here) before the barrier.  Tail colours (LG_TAIL) map their slots
// to the first workgroup's threads, in the sweeps and at the start.  PRE rows
// per slot are preloaded (the rows of one pass of a contact with friction).
template <int PRE>
__global__ __launch_bounds__(LG_BLOCK) void k_lg_solve(DevWorlds D, StepParams P, LargeDev L, int iters) {
    __shared__ int s_cs[LG_MAXCOL + 1];
    // a colouring that gave up on a barrier leaves no schedule: the step
    // changes no body (the next sync reports it failed)
    if (L.counters[LG_BARRIER_ERR]) return;
    unsigned gen = 0;
    const int nc = L.counters[LG_CONTACTS];
    const int ncol = L.counters[LG_NCOL];
    const int gt = blockIdx.x * blockDim.x + threadIdx.x, gs = gridDim.x * blockDim.x;
    const __amdgpu_buffer_rsrc_t rs = lg_fc_rsrc(L.fc, L.nb);
    const float zero[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int c = threadIdx.x; c <= ncol; c += blockDim.x) s_cs[c] = L.col_start[c];
    __syncthreads();
    int ct = ncol;  // colours ct .. ncol-1 form the tail
    while (ct > 0 && s_cs[ct] - s_cs[ct - 1] <= LG_TAIL) ct--;
    for (int b = gt; b < L.nb; b += gs) lg_fc_store(rs, b, zero);
    for (int c = 0; c < ncol; c++) {   // slots of colour c map to threads as in the sweeps
        const int e = s_cs[c + 1];
        const int first = c < ct ? s_cs[c] + gt : (blockIdx.x == 0 ? s_cs[c] + (int)threadIdx.x : e);
        const int step = c < ct ? gs : (int)blockDim.x;
        for (int idx = first; idx < e; idx += step) {
            const int4 it = L.items[idx];
            for (int r = it.x; r < it.x + (it.y >> 8); r++) D.lambda[r] = 0.f;
        }
    }
    bool ok = lg_grid_sync(L, gen);
    const int nph = ct + (ct < ncol ? 1 : 0);  // grid phases per pass
    const int npass = L.one_pass ? 1 : 2;
    const int T = (nc > 0 && ncol > 0) ? iters * npass * nph : 0;
    int c = 0, pass = 0;
    int idx = s_cs[0] + gt, e = ct > 0 ? s_cs[1] : 0;
    LgSlot<PRE> cur;
    lg_slot_load(D, idx < e && T > 0 ? L.items[idx] : make_int4(-1, 0, 0, 0), 0, cur);
    for (int p = 0; ok && p < T; p++) {
        const int c2 = (p + 1) % nph, pass2 = ((p + 1) / nph) % npass;
        const bool grid2 = p + 1 < T && c2 < ct;
        const int idx2 = grid2 ? s_cs[c2] + gt : 0, e2 = grid2 ? s_cs[c2 + 1] : 0;
        LgSlot<PRE> nxt;
        lg_slot_load(D, idx2 < e2 ? L.items[idx2] : make_int4(-1, 0, 0, 0), pass2, nxt);
        if (c < ct) {
            lg_slot_run(D, rs, cur);
            for (int j = idx + gs; j < e; j += gs) {   // slots beyond one per thread
                LgSlot<PRE> x;
                lg_slot_load(D, L.items[j], pass, x);
                lg_slot_run(D, rs, x);
            }
        } else if (blockIdx.x == 0) {
            for (int tc = ct; tc < ncol; tc++) {
                for (int j = s_cs[tc] + (int)threadIdx.x; j < s_cs[tc + 1]; j += blockDim.x) {
                    LgSlot<PRE> x;
                    lg_slot_load(D, L.items[j], pass, x);
                    lg_slot_run(D, rs, x);
                }
                mem_done();
                __syncthreads();
            }
        }
        ok = lg_grid_sync(L, gen);
        c = c2; pass = pass2; idx = idx2; e = e2; cur = nxt;
    }
    if (!ok) return;
    for (int b = gt; b < L.nb; b += gs) {
        float f[6];
        lg_fc_load(rs, b, f);
        integrate_body(D, P, b, f, nc > 0, L.ftot, L.ftot + L.nb);
    }
}

// largest value each counter reached since the host last cleared sticky
__global__ void k_lg_fold(const int *counters, int *sticky) {
    const int i = threadIdx.x;
    if (i < LG_NCOUNTERS) sticky[i] = max(sticky[i], counters[i]);
}

// ---- launchers ----
void lg_enqueue_fold(const LargeDev &L, int *sticky, hipStream_t s) {
    hipLaunchKernelGGL(k_lg_fold, dim3(1), dim3(64), 0, s, L.counters, sticky);
}

// workgroups of a cooperative launch: every one resident at once
int lg_coop_blocks(const void *kernel, int per_cu) {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, LG_BLOCK, 0) != hipSuccess) return 0;
    return cus * (per < per_cu ? per : per_cu);
}
int lg_colour_blocks() { return lg_coop_blocks((const void *)k_lg_colour, 2); }
static const void *lg_solve_kernel(int one_pass) {
    return one_pass ? (const void *)k_lg_solve<3> : (const void *)k_lg_solve<2>;
}
int lg_solve_blocks(int per_cu, int one_pass) { return lg_coop_blocks(lg_solve_kernel(one_pass), per_cu); }

// The prepare stage of a step: bodies, broadphase, narrowphase, rows.
// Capacities: ng <= pair scan cap etc. are checked by the host.
void lg_enqueue_prepare(const DevWorlds &D, const StepParams &P, const LargeDev &L, hipStream_t s) {
    const int G = 1024, T = 256;
    hipLaunchKernelGGL(k_lg_fill, dim3(G), dim3(T), 0, s, L.bucket, L.hmask + 2, 0);
    hipLaunchKernelGGL(k_lg_fill, dim3(1), dim3(T), 0, s, L.counters, LG_NCOUNTERS, 0);
    hipLaunchKernelGGL(k_lg_bodies, dim3(G), dim3(T), 0, s, D, P, L.nb, L.ftot);
    hipLaunchKernelGGL(k_lg_aabb, dim3(G), dim3(T), 0, s, D, L);
    lg_scan(L.bucket, L.hmask + 2, nullptr, L.scan_tmp, s);
    hipLaunchKernelGGL(k_lg_scatter, dim3(G), dim3(T), 0, s, L);
    hipLaunchKernelGGL(k_lg_sort_buckets, dim3(G), dim3(T), 0, s, D, L);
    hipLaunchKernelGGL(k_lg_pairs, dim3(G), dim3(T), 0, s, D, L, 0, P.cat_filter);
    lg_scan(L.pair_cnt, L.ng + 1, nullptr, L.scan_tmp, s);
    hipLaunchKernelGGL(k_lg_pairs, dim3(G), dim3(T), 0, s, D, L, 1, P.cat_filter);
    hipLaunchKernelGGL(k_lg_contacts, dim3(G), dim3(T), 0, s, D, P, L, 0);
    lg_scan(L.contact_cnt, L.pair_cap + 1, L.counters + LG_PAIRS, L.scan_tmp, s);
    hipLaunchKernelGGL(k_lg_contacts, dim3(G), dim3(T), 0, s, D, P, L, 1);
    hipLaunchKernelGGL(k_lg_rowcount, dim3(G), dim3(T), 0, s, D, L);
    lg_scan(L.row_cnt, L.contact_cap + 1, L.counters + LG_CONTACTS, L.scan_tmp, s);
    hipLaunchKernelGGL(k_lg_rows, dim3(G), dim3(T), 0, s, D, P, L);
}

bool lg_enqueue_colour(const DevWorlds &D, const LargeDev &L, int blocks, hipStream_t s) {
    hipLaunchKernelGGL(k_lg_fill, dim3(16), dim3(256), 0, s, L.elig, LG_MAXCOL * LG_MAXSUB, 0);
    hipLaunchKernelGGL(k_lg_fill, dim3(1), dim3(256), 0, s, L.col_start, LG_MAXCOL + 1, 0);
    hipLaunchKernelGGL(k_lg_fill, dim3(1), dim3(256), 0, s, reinterpret_cast<int *>(L.bar), LG_BAR_WORDS, 0);
    if (hipGetLastError() != hipSuccess) return false;
    DevWorlds d = D;
    LargeDev l = L;
    void *args[] = {(void *)&d, (void *)&l};
    return hipLaunchCooperativeKernel((const void *)k_lg_colour, dim3(blocks), dim3(LG_BLOCK), args, 0, s) ==
           hipSuccess;
}

bool lg_enqueue_solve(const DevWorlds &D, const StepParams &P, const LargeDev &L, int iters, int blocks,
                      hipStream_t s) {
    hipLaunchKernelGGL(k_lg_fill, dim3(1), dim3(256), 0, s, reinterpret_cast<int *>(L.bar), LG_BAR_FILL, 0);
    if (hipGetLastError() != hipSuccess) return false;
    DevWorlds d = D;
    StepParams p = P;
    LargeDev l = L;
    int n = iters;
    void *args[] = {(void *)&d, (void *)&p, (void *)&l, (void *)&n};
    return hipLaunchCooperativeKernel(lg_solve_kernel(L.one_pass), dim3(blocks), dim3(LG_BLOCK), args, 0, s) ==
           hipSuccess;
}
}  // namespace odx
