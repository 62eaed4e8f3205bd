// odx_collide.h -- device narrowphase for sphere, box, capsule and plane.
//
// One lane evaluates one geom pair; contacts are produced in registers (up to
// ODX_MAXC) and compacted by the caller.  Contact semantics follow dCollide and
// the primitive colliders of the reference (ode/src/sphere.cpp, box.cpp,
// capsule.cpp, collision_util.cpp); each function cites the routine it
// implements.  The contact normal points into geom 1 and the reversed-table
// entries negate it, as dCollide does (collision_kernel.cpp:292-329).
#pragma once
#include "odx_math.h"

#define ODX_MAXC 4

namespace odx {

enum { CLS_SPHERE = 0, CLS_BOX = 1, CLS_CAPSULE = 2, CLS_CYLINDER = 3, CLS_PLANE = 4 };

struct DGeom {
    int cls;
    int body;       // world-local body index, -1 = static
    float p[4];     // sphere r | box lx,ly,lz | capsule r,len | plane a,b,c,d
    float pos[3];
    float R[12];
};

struct DContact {
    float pos[3];
    float normal[3];
    float depth;
};

// Working arrays of the box-box clipper, indexed by a count that differs per
// lane: two clip buffers of 8 points (CW_A, CW_B), the clipped points in 3D
// (CW_PT) with their depths (CW_DEP), and the centroid angles of cullPoints
// (CW_ANG).  ClipPriv keeps them in a private array; ClipLds in one lane's
// column of an LDS block (element i at i * 64, so that a wave's 64 lanes read
// 64 consecutive words).
enum { CW_A = 0, CW_B = 16, CW_PT = 32, CW_DEP = 56, CW_ANG = 64, CLIP_WS = 72 };
struct ClipPriv {
    float v[CLIP_WS];
    ODX_DEV float &operator[](int i) { return v[i]; }
};
struct ClipLds {
    float *p;  // this lane's first element
    ODX_DEV float &operator[](int i) const { return p[i << 6]; }
};

// v[k] for k in 0..2 by selects, so that v can stay in registers.
ODX_DEV float sel3(const float *v, int k) { return k == 0 ? v[0] : (k == 1 ? v[1] : v[2]); }
// Column k (0..2) of a 3x4 row-major matrix.
ODX_DEV void col3(float *o, const float *R, int k) {
    o[0] = sel3(R, k); o[1] = sel3(R + 4, k); o[2] = sel3(R + 8, k);
}

// computeAABB: sphere.cpp:59-67, box.cpp:60-76, capsule.cpp:60-74, plane.cpp:81-105
ODX_DEV void geom_aabb(const DGeom &G, float *ab) {
    const float *p = G.pos, *R = G.R;
    switch (G.cls) {
    case CLS_SPHERE: {
        float r = G.p[0];
        ab[0] = p[0] - r; ab[1] = p[0] + r; ab[2] = p[1] - r; ab[3] = p[1] + r; ab[4] = p[2] - r; ab[5] = p[2] + r;
        break;
    }
    case CLS_BOX: {
        const float *s = G.p;
        float xr = 0.5f * (fabsf(R[0] * s[0]) + fabsf(R[1] * s[1]) + fabsf(R[2] * s[2]));
        float yr = 0.5f * (fabsf(R[4] * s[0]) + fabsf(R[5] * s[1]) + fabsf(R[6] * s[2]));
        float zr = 0.5f * (fabsf(R[8] * s[0]) + fabsf(R[9] * s[1]) + fabsf(R[10] * s[2]));
        ab[0] = p[0] - xr; ab[1] = p[0] + xr; ab[2] = p[1] - yr; ab[3] = p[1] + yr; ab[4] = p[2] - zr; ab[5] = p[2] + zr;
        break;
    }
    case CLS_CAPSULE: {
        float r = G.p[0], lz = G.p[1];
        float xr = fabsf(R[2] * lz) * 0.5f + r, yr = fabsf(R[6] * lz) * 0.5f + r, zr = fabsf(R[10] * lz) * 0.5f + r;
        ab[0] = p[0] - xr; ab[1] = p[0] + xr; ab[2] = p[1] - yr; ab[3] = p[1] + yr; ab[4] = p[2] - zr; ab[5] = p[2] + zr;
        break;
    }
    case CLS_CYLINDER: {  // cylinder.cpp:61-78: axis half-length plus the disc's extent
        float r = G.p[0], lz = G.p[1];
        float o2 = 1.f - R[2] * R[2], o6 = 1.f - R[6] * R[6], o10 = 1.f - R[10] * R[10];
        float xr = fabsf(R[2] * lz * 0.5f) + r * sqrtf(0.f > o2 ? 0.f : o2);
        float yr = fabsf(R[6] * lz * 0.5f) + r * sqrtf(0.f > o6 ? 0.f : o6);
        float zr = fabsf(R[10] * lz * 0.5f) + r * sqrtf(0.f > o10 ? 0.f : o10);
        ab[0] = p[0] - xr; ab[1] = p[0] + xr; ab[2] = p[1] - yr; ab[3] = p[1] + yr; ab[4] = p[2] - zr; ab[5] = p[2] + zr;
        break;
    }
    case CLS_PLANE: {
        const float *n = G.p;
        ab[0] = -kInf; ab[1] = kInf; ab[2] = -kInf; ab[3] = kInf; ab[4] = -kInf; ab[5] = kInf;
        if (n[1] == 0.f && n[2] == 0.f) {
            ab[0] = (n[0] > 0.f) ? -kInf : -n[3];
            ab[1] = (n[0] > 0.f) ? n[3] : kInf;
        } else if (n[0] == 0.f && n[2] == 0.f) {
            ab[2] = (n[1] > 0.f) ? -kInf : -n[3];
            ab[3] = (n[1] > 0.f) ? n[3] : kInf;
        } else if (n[0] == 0.f && n[1] == 0.f) {
            ab[4] = (n[2] > 0.f) ? -kInf : -n[3];
            ab[5] = (n[2] > 0.f) ? n[3] : kInf;
        }
        break;
    }
    default:
        ab[0] = -kInf; ab[1] = kInf; ab[2] = -kInf; ab[3] = kInf; ab[4] = -kInf; ab[5] = kInf;
    }
}


// Geom view (cls, body, params[4], pos[3], R[12]) into a DGeom.
ODX_DEV void load_gview(const float *gv, DGeom &G) {
    G.cls = (int)gv[0];
    G.body = (int)gv[1];
    for (int k = 0; k < 4; k++) G.p[k] = gv[2 + k];
    for (int k = 0; k < 3; k++) G.pos[k] = gv[6 + k];
    for (int k = 0; k < 12; k++) G.R[k] = gv[9 + k];
}

// dCollideSpheres collision_util.cpp:38-67
ODX_DEV int collide_spheres(const float *p1, float r1, const float *p2, float r2, DContact &c) {
    float dx = p1[0] - p2[0], dy = p1[1] - p2[1], dz = p1[2] - p2[2];
    float d = sqrtf(dx * dx + dy * dy + dz * dz);
    if (d > (r1 + r2)) return 0;
    if (d <= 0.f) {
        c.pos[0] = p1[0]; c.pos[1] = p1[1]; c.pos[2] = p1[2];
        c.normal[0] = 1.f; c.normal[1] = 0.f; c.normal[2] = 0.f;
        c.depth = r1 + r2;
    } else {
        float d1 = 1.f / d;
        c.normal[0] = dx * d1; c.normal[1] = dy * d1; c.normal[2] = dz * d1;
        float k = 0.5f * (r2 - r1 - d);
        c.pos[0] = p1[0] + c.normal[0] * k;
        c.pos[1] = p1[1] + c.normal[1] * k;
        c.pos[2] = p1[2] + c.normal[2] * k;
        c.depth = r1 + r2 - d;
    }
    return 1;
}

// dCollideSphereBox sphere.cpp:131-219
ODX_DEV int collide_sphere_box(const DGeom &s, const DGeom &b, DContact &c) {
    float p[3], t[3], l[3];
    bool onborder = false;
#pragma unroll
    for (int i = 0; i < 3; i++) p[i] = s.pos[i] - b.pos[i];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        l[i] = b.p[i] * 0.5f;
        t[i] = dot14(p, b.R + i);
        if (t[i] < -l[i]) { t[i] = -l[i]; onborder = true; }
        if (t[i] > l[i]) { t[i] = l[i]; onborder = true; }
    }
    if (!onborder) {
        float mind = l[0] - fabsf(t[0]);
        int mini = 0;
        float f1 = l[1] - fabsf(t[1]);
        if (f1 < mind) { mind = f1; mini = 1; }
        float f2 = l[2] - fabsf(t[2]);
        if (f2 < mind) { mind = f2; mini = 2; }
        c.pos[0] = s.pos[0]; c.pos[1] = s.pos[1]; c.pos[2] = s.pos[2];
        float tsel = mini == 0 ? t[0] : (mini == 1 ? t[1] : t[2]);
        float sg = tsel > 0.f ? 1.f : -1.f;
        // R * (sg * e_mini)
        c.normal[0] = sg * sel3(b.R + 0, mini);
        c.normal[1] = sg * sel3(b.R + 4, mini);
        c.normal[2] = sg * sel3(b.R + 8, mini);
        c.depth = mind + s.p[0];
        return 1;
    }
    float q[3], r[3];
    mul0_331(q, b.R, t);
#pragma unroll
    for (int i = 0; i < 3; i++) r[i] = p[i] - q[i];
    float depth = s.p[0] - sqrtf(dot3(r, r));
    if (depth < 0.f) return 0;
#pragma unroll
    for (int i = 0; i < 3; i++) { c.pos[i] = q[i] + b.pos[i]; c.normal[i] = r[i]; }
    normalize3(c.normal);
    c.depth = depth;
    return 1;
}

// dCollideSpherePlane sphere.cpp:222-251
ODX_DEV int collide_sphere_plane(const DGeom &s, const DGeom &pl, DContact &c) {
    float depth = pl.p[3] - dot3(s.pos, pl.p) + s.p[0];
    if (depth < 0.f) return 0;
#pragma unroll
    for (int i = 0; i < 3; i++) { c.normal[i] = pl.p[i]; c.pos[i] = s.pos[i] - pl.p[i] * s.p[0]; }
    c.depth = depth;
    return 1;
}

// Clip a polygon (n <= 8 points, at ws[q..]) against sign * x[dir] < h into
// ws[r..].  One pass of intersectRectQuad box.cpp:212-263; returns the new
// count, or a count with bit 8 set when the output filled up (the reference
// stops clipping then).
template <class W>
ODX_DEV int clip_pass(W &ws, int q, int nq, int r, int dir, float sign, float h) {
    int nr = 0;
    for (int i = 0; i < nq; i++) {
        const int pq = q + 2 * i;
        const int nx = (i + 1 < nq) ? pq + 2 : q;
        const float p0 = ws[pq], p1 = ws[pq + 1], n0 = ws[nx], n1 = ws[nx + 1];
        const float pd = dir ? p1 : p0, po = dir ? p0 : p1, nd = dir ? n1 : n0, no = dir ? n0 : n1;
        bool in0 = sign * pd < h;
        if (in0) {
            ws[r + 2 * nr] = p0; ws[r + 2 * nr + 1] = p1;
            nr++;
            if (nr & 8) return nr | 0x100;
        }
        bool in1 = sign * nd < h;
        if (in0 ^ in1) {
            float t = (no - po) / (nd - pd) * (sign * h - pd);
            ws[r + 2 * nr + (1 - dir)] = po + t;
            ws[r + 2 * nr + dir] = sign * h;
            nr++;
            if (nr & 8) return nr | 0x100;
        }
    }
    return nr;
}

// intersectRectQuad box.cpp:212-263: clip the quad at ws[CW_A..CW_A+7]
// against the rectangle |x| < h0, |y| < h1.  Returns the point count (<= 8);
// *at = the buffer (CW_A or CW_B) that holds the points.
template <class W>
ODX_DEV int intersect_rect_quad(float h0, float h1, W &ws, int *at) {
    int n = 4;
    // pass order: (dir 0, sign -1), (0, +1), (1, -1), (1, +1); ping-pong A/B
    n = clip_pass(ws, CW_A, n, CW_B, 0, -1.f, h0);
    if (n & 0x100) { *at = CW_B; return n & 0xff; }
    n = clip_pass(ws, CW_B, n, CW_A, 0, 1.f, h0);
    if (n & 0x100) { *at = CW_A; return n & 0xff; }
    n = clip_pass(ws, CW_A, n, CW_B, 1, -1.f, h1);
    if (n & 0x100) { *at = CW_B; return n & 0xff; }
    n = clip_pass(ws, CW_B, n, CW_A, 1, 1.f, h1);
    *at = CW_A;
    return n & 0xff;
}

// cullPoints box.cpp:274-336: choose m (<= ODX_MAXC) of the n points at
// ws[p..] spread by angle around the centroid, starting with i0.
template <class W>
ODX_DEV void cull_points(int n, W &ws, int p, int m, int i0, int *iret) {
    float a, cx, cy, q;
    if (n == 1) { cx = ws[p]; cy = ws[p + 1]; }
    else if (n == 2) { cx = 0.5f * (ws[p] + ws[p + 2]); cy = 0.5f * (ws[p + 1] + ws[p + 3]); }
    else {
        a = 0.f; cx = 0.f; cy = 0.f;
        for (int i = 0; i < n - 1; i++) {
            const float x0 = ws[p + i * 2], y0 = ws[p + i * 2 + 1], x1 = ws[p + i * 2 + 2], y1 = ws[p + i * 2 + 3];
            q = x0 * y1 - x1 * y0;
            a += q;
            cx += q * (x0 + x1);
            cy += q * (y0 + y1);
        }
        const float xl = ws[p + n * 2 - 2], yl = ws[p + n * 2 - 1], x0 = ws[p], y0 = ws[p + 1];
        q = xl * y0 - x0 * yl;
        a = 1.f / (3.f * (a + q));
        cx = a * (cx + q * (xl + x0));
        cy = a * (cy + q * (yl + y0));
    }
    for (int i = 0; i < n; i++) ws[CW_ANG + i] = atan2f(ws[p + i * 2 + 1] - cy, ws[p + i * 2] - cx);
    unsigned avail = (1u << n) - 1u;
    avail &= ~(1u << i0);
    iret[0] = i0;
    const double PI = 3.14159265358979323846;
    const float ai0 = ws[CW_ANG + i0];
#pragma unroll
    for (int j = 1; j < ODX_MAXC; j++) {
        if (j >= m) break;
        a = (float)(double(j) * (2.0 * PI / m) + (double)ai0);
        if ((double)a > PI) a -= (float)(2.0 * PI);
        float maxdiff = 1e9f;
        int best = i0;
        for (int i = 0; i < n; i++) {
            if (avail & (1u << i)) {
                float diff = fabsf(ws[CW_ANG + i] - a);
                if ((double)diff > PI) diff = (float)(2.0 * PI - (double)diff);
                if (diff < maxdiff) { maxdiff = diff; best = i; }
            }
        }
        iret[j] = best;
        avail &= ~(1u << best);
    }
}

// dLineClosestApproach collision_util.cpp:70-95
ODX_DEV void line_closest_approach(const float *pa, const float *ua, const float *pb, const float *ub, float &alpha,
                                   float &beta) {
    float p[3] = {pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2]};
    float uaub = dot3(ua, ub), q1 = dot3(ua, p), q2 = -dot3(ub, p);
    float d = 1.f - uaub * uaub;
    if (d <= 0.0001f) { alpha = 0.f; beta = 0.f; }
    else {
        d = 1.f / d;
        alpha = (q1 + uaub * q2) * d;
        beta = (uaub * q1 + q2) * d;
    }
}

// dBoxBox box.cpp:356-737.  Writes pos/depth of the contacts, the box-1-to-box-2
// separating normal and returns the count (0 = separated).
// The geoms' matrices, positions and half sides are read with static indices
// and selects only (no pointer chosen between two geoms), so that they stay in
// registers; the arithmetic is the reference's, term for term.
template <class W>
ODX_DEV int box_box(const float *p1, const float *R1, const float *side1, const float *p2, const float *R2,
                    const float *side2, float *normal, int maxc, DContact *con, W &ws) {
    const float fudge = 1.05f;
    float p[3], pp[3], nC[3] = {0.f, 0.f, 0.f};
    int nRcol = -1, nRbox = 0;  // normal = column nRcol of R1 (nRbox 0) or R2 (1); -1 => nC
    float A[3], B[3];
#pragma unroll
    for (int i = 0; i < 3; i++) { p[i] = p2[i] - p1[i]; A[i] = side1[i] * 0.5f; B[i] = side2[i] * 0.5f; }
    mul1_331(pp, R1, p);
    float R11 = dot44(R1 + 0, R2 + 0), R12 = dot44(R1 + 0, R2 + 1), R13 = dot44(R1 + 0, R2 + 2);
    float R21 = dot44(R1 + 1, R2 + 0), R22 = dot44(R1 + 1, R2 + 1), R23 = dot44(R1 + 1, R2 + 2);
    float R31 = dot44(R1 + 2, R2 + 0), R32 = dot44(R1 + 2, R2 + 1), R33 = dot44(R1 + 2, R2 + 2);
    float Q11 = fabsf(R11), Q12 = fabsf(R12), Q13 = fabsf(R13);
    float Q21 = fabsf(R21), Q22 = fabsf(R22), Q23 = fabsf(R23);
    float Q31 = fabsf(R31), Q32 = fabsf(R32), Q33 = fabsf(R33);
    float s = -kInf, s2, e, l;
    int invert = 0, code = 0;
#define ODX_TSTF(expr1, expr2, box, col, cc)                                        \
    e = (expr1);                                                                     \
    s2 = fabsf(e) - (expr2);                                                         \
    if (s2 > 0.f) return 0;                                                          \
    if (s2 > s) { s = s2; nRbox = box; nRcol = col; invert = (e < 0.f); code = cc; }
    ODX_TSTF(pp[0], (A[0] + B[0] * Q11 + B[1] * Q12 + B[2] * Q13), 0, 0, 1);
    ODX_TSTF(pp[1], (A[1] + B[0] * Q21 + B[1] * Q22 + B[2] * Q23), 0, 1, 2);
    ODX_TSTF(pp[2], (A[2] + B[0] * Q31 + B[1] * Q32 + B[2] * Q33), 0, 2, 3);
    ODX_TSTF(dot41(R2 + 0, p), (A[0] * Q11 + A[1] * Q21 + A[2] * Q31 + B[0]), 1, 0, 4);
    ODX_TSTF(dot41(R2 + 1, p), (A[0] * Q12 + A[1] * Q22 + A[2] * Q32 + B[1]), 1, 1, 5);
    ODX_TSTF(dot41(R2 + 2, p), (A[0] * Q13 + A[1] * Q23 + A[2] * Q33 + B[2]), 1, 2, 6);
#undef ODX_TSTF
#define ODX_TSTE(expr1, expr2, n1, n2, n3, cc)                              \
    e = (expr1);                                                             \
    s2 = fabsf(e) - (expr2);                                                 \
    if (s2 > 0.f) return 0;                                                  \
    l = sqrtf((n1) * (n1) + (n2) * (n2) + (n3) * (n3));                      \
    if (l > 0.f) {                                                           \
        s2 /= l;                                                             \
        if (s2 * fudge > s) {                                                \
            s = s2; nRcol = -1;                                              \
            nC[0] = (n1) / l; nC[1] = (n2) / l; nC[2] = (n3) / l;            \
            invert = (e < 0.f); code = cc;                                   \
        }                                                                    \
    }
    ODX_TSTE(pp[2] * R21 - pp[1] * R31, (A[1] * Q31 + A[2] * Q21 + B[1] * Q13 + B[2] * Q12), 0.f, -R31, R21, 7);
    ODX_TSTE(pp[2] * R22 - pp[1] * R32, (A[1] * Q32 + A[2] * Q22 + B[0] * Q13 + B[2] * Q11), 0.f, -R32, R22, 8);
    ODX_TSTE(pp[2] * R23 - pp[1] * R33, (A[1] * Q33 + A[2] * Q23 + B[0] * Q12 + B[1] * Q11), 0.f, -R33, R23, 9);
    ODX_TSTE(pp[0] * R31 - pp[2] * R11, (A[0] * Q31 + A[2] * Q11 + B[1] * Q23 + B[2] * Q22), R31, 0.f, -R11, 10);
    ODX_TSTE(pp[0] * R32 - pp[2] * R12, (A[0] * Q32 + A[2] * Q12 + B[0] * Q23 + B[2] * Q21), R32, 0.f, -R12, 11);
    ODX_TSTE(pp[0] * R33 - pp[2] * R13, (A[0] * Q33 + A[2] * Q13 + B[0] * Q22 + B[1] * Q21), R33, 0.f, -R13, 12);
    ODX_TSTE(pp[1] * R11 - pp[0] * R21, (A[0] * Q21 + A[1] * Q11 + B[1] * Q33 + B[2] * Q32), -R21, R11, 0.f, 13);
    ODX_TSTE(pp[1] * R12 - pp[0] * R22, (A[0] * Q22 + A[1] * Q12 + B[0] * Q33 + B[2] * Q31), -R22, R12, 0.f, 14);
    ODX_TSTE(pp[1] * R13 - pp[0] * R23, (A[0] * Q23 + A[1] * Q13 + B[0] * Q32 + B[1] * Q31), -R23, R13, 0.f, 15);
#undef ODX_TSTE
    if (!code) return 0;
    if (nRcol >= 0) {
        float n1[3], n2[3];
        col3(n1, R1, nRcol);
        col3(n2, R2, nRcol);
#pragma unroll
        for (int i = 0; i < 3; i++) normal[i] = nRbox ? n2[i] : n1[i];
    } else {
        mul0_331(normal, R1, nC);
    }
    if (invert) { normal[0] = -normal[0]; normal[1] = -normal[1]; normal[2] = -normal[2]; }
    const float depth = -s;

    if (code > 6) {
        float pa[3], pb[3], ua[3], ub[3], alpha, beta;
#pragma unroll
        for (int i = 0; i < 3; i++) { pa[i] = p1[i]; pb[i] = p2[i]; }
#pragma unroll
        for (int j = 0; j < 3; j++) {
            float sg = (dot14(normal, R1 + j) > 0.f) ? 1.f : -1.f;
#pragma unroll
            for (int i = 0; i < 3; i++) pa[i] += sg * A[j] * R1[i * 4 + j];
        }
#pragma unroll
        for (int j = 0; j < 3; j++) {
            float sg = (dot14(normal, R2 + j) > 0.f) ? -1.f : 1.f;
#pragma unroll
            for (int i = 0; i < 3; i++) pb[i] += sg * B[j] * R2[i * 4 + j];
        }
        int ca = (code - 7) / 3, cb = (code - 7) % 3;
        col3(ua, R1, ca);
        col3(ub, R2, cb);
        line_closest_approach(pa, ua, pb, ub, alpha, beta);
#pragma unroll
        for (int i = 0; i < 3; i++) {
            pa[i] += ua[i] * alpha;
            pb[i] += ub[i] * beta;
            con[0].pos[i] = 0.5f * (pa[i] + pb[i]);
        }
        con[0].depth = depth;
        return 1;
    }

    const bool one = code <= 3;
    float Ra[12], Rb[12], pa[3], pb[3], Sa[3], Sb[3];
#pragma unroll
    for (int i = 0; i < 12; i++) { Ra[i] = one ? R1[i] : R2[i]; Rb[i] = one ? R2[i] : R1[i]; }
#pragma unroll
    for (int i = 0; i < 3; i++) {
        pa[i] = one ? p1[i] : p2[i]; pb[i] = one ? p2[i] : p1[i];
        Sa[i] = one ? A[i] : B[i]; Sb[i] = one ? B[i] : A[i];
    }
    float normal2[3], nr[3], anr[3];
#pragma unroll
    for (int i = 0; i < 3; i++) normal2[i] = one ? normal[i] : -normal[i];
    mul1_331(nr, Rb, normal2);
#pragma unroll
    for (int i = 0; i < 3; i++) anr[i] = fabsf(nr[i]);
    int lanr, a1, a2;
    if (anr[1] > anr[0]) {
        if (anr[1] > anr[2]) { a1 = 0; lanr = 1; a2 = 2; }
        else { a1 = 0; a2 = 1; lanr = 2; }
    } else {
        if (anr[0] > anr[2]) { lanr = 0; a1 = 1; a2 = 2; }
        else { a1 = 0; a2 = 1; lanr = 2; }
    }
    float center[3], rbl[3];
    const float nlanr = sel3(nr, lanr);
    const float sbl = sel3(Sb, lanr);
    col3(rbl, Rb, lanr);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        float rb = rbl[i];
        center[i] = (nlanr < 0.f) ? (pb[i] - pa[i] + sbl * rb) : (pb[i] - pa[i] - sbl * rb);
    }
    int codeN = one ? code - 1 : code - 4;
    int code1 = codeN == 0 ? 1 : 0;
    int code2 = codeN == 2 ? 1 : 2;
    // columns code1, code2 of Ra and a1, a2 of Rb
    float ra1[3], ra2[3], rb1[3], rb2[3];
    col3(ra1, Ra, code1); col3(ra2, Ra, code2);
    col3(rb1, Rb, a1); col3(rb2, Rb, a2);
    float c1 = dot3(center, ra1), c2 = dot3(center, ra2);
    float m11 = dot3(ra1, rb1), m12 = dot3(ra1, rb2);
    float m21 = dot3(ra2, rb1), m22 = dot3(ra2, rb2);
    {
        float sba1 = sel3(Sb, a1);
        float sba2 = sel3(Sb, a2);
        float k1 = m11 * sba1, k2 = m21 * sba1, k3 = m12 * sba2, k4 = m22 * sba2;
        ws[CW_A + 0] = c1 - k1 - k3; ws[CW_A + 1] = c2 - k2 - k4;
        ws[CW_A + 2] = c1 - k1 + k3; ws[CW_A + 3] = c2 - k2 + k4;
        ws[CW_A + 4] = c1 + k1 + k3; ws[CW_A + 5] = c2 + k2 + k4;
        ws[CW_A + 6] = c1 + k1 - k3; ws[CW_A + 7] = c2 + k2 - k4;
    }
    int rt;  // buffer holding the clipped 2D points
    const int n = intersect_rect_quad(code1 == 0 ? Sa[0] : Sa[1], code2 == 1 ? Sa[1] : Sa[2], ws, &rt);
    if (n < 1) return 0;
    float det1 = 1.f / (m11 * m22 - m12 * m21);
    m11 *= det1; m12 *= det1; m21 *= det1; m22 *= det1;
    const float saN = sel3(Sa, codeN);
    int cnum = 0;
    for (int j = 0; j < n; j++) {
        const float rx = ws[rt + j * 2], ry = ws[rt + j * 2 + 1];
        float k1 = m22 * (rx - c1) - m12 * (ry - c2);
        float k2 = -m21 * (rx - c1) + m11 * (ry - c2);
        float pt[3];
#pragma unroll
        for (int i = 0; i < 3; i++) pt[i] = center[i] + k1 * rb1[i] + k2 * rb2[i];
        float dd = saN - dot3(normal2, pt);
        if (dd >= 0.f) {
            ws[CW_PT + cnum * 3 + 0] = pt[0]; ws[CW_PT + cnum * 3 + 1] = pt[1]; ws[CW_PT + cnum * 3 + 2] = pt[2];
            ws[CW_DEP + cnum] = dd;
            ws[rt + cnum * 2] = rx;
            ws[rt + cnum * 2 + 1] = ry;
            cnum++;
        }
    }
    if (cnum < 1) return 0;
    if (maxc > cnum) maxc = cnum;
    if (maxc < 1) maxc = 1;
    // all points when they fit, else maxc of them culled by angle from the
    // deepest one
    int iret[ODX_MAXC] = {0, 1, 2, 3};
    const bool cull = cnum > maxc;
    if (cull) {
        int i1 = 0;
        float maxdepth = ws[CW_DEP];
        for (int i = 1; i < cnum; i++) {
            const float d = ws[CW_DEP + i];
            if (d > maxdepth) { maxdepth = d; i1 = i; }
        }
        cull_points(cnum, ws, rt, maxc, i1, iret);
    }
    const int nout = cull ? maxc : cnum;
#pragma unroll
    for (int j = 0; j < ODX_MAXC; j++) {
        if (j >= nout) break;
        const int k = iret[j];
#pragma unroll
        for (int i = 0; i < 3; i++) con[j].pos[i] = ws[CW_PT + k * 3 + i] + pa[i];
        con[j].depth = ws[CW_DEP + k];
    }
    return nout;
}

// Sets the normal of contacts 0..n-1 (unrolled over the ODX_MAXC slots, so
// that the contacts stay in registers).
ODX_DEV void set_normals(DContact *c, int n, float n0, float n1, float n2) {
#pragma unroll
    for (int i = 0; i < ODX_MAXC; i++)
        if (i < n) { c[i].normal[0] = n0; c[i].normal[1] = n1; c[i].normal[2] = n2; }
}

// dCollideBoxBox box.cpp:741-767
template <class W>
ODX_DEV int collide_box_box(const DGeom &b1, const DGeom &b2, int maxc, DContact *c, W &ws) {
    float normal[3];
    int n = box_box(b1.pos, b1.R, b1.p, b2.pos, b2.R, b2.p, normal, maxc, c, ws);
    set_normals(c, n, -normal[0], -normal[1], -normal[2]);
    return n;
}

// dCollideBoxPlane box.cpp:770-903
ODX_DEV int collide_box_plane(const DGeom &bx, const DGeom &pl, int maxc, DContact *c) {
    const float *R = bx.R, *n = pl.p, *side = bx.p;
    float Aq[3], Bq[3];
#pragma unroll
    for (int i = 0; i < 3; i++) { Aq[i] = side[i] * dot14(n, R + i); Bq[i] = fabsf(Aq[i]); }
    float depth = pl.p[3] + 0.5f * (Bq[0] + Bq[1] + Bq[2]) - dot3(n, bx.pos);
    if (depth < 0.f) return 0;
    if (maxc > 4) maxc = 4;
    float p[3] = {bx.pos[0], bx.pos[1], bx.pos[2]};
#pragma unroll
    for (int i = 0; i < 3; i++) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            if (Aq[i] > 0.f) p[k] -= 0.5f * side[i] * R[k * 4 + i];
            else p[k] += 0.5f * side[i] * R[k * 4 + i];
        }
    }
    c[0].pos[0] = p[0]; c[0].pos[1] = p[1]; c[0].pos[2] = p[2];
    c[0].depth = depth;
    int ret = 1;
    // the two sides of smallest projected length, in the order of the
    // reference's branch tree (box.cpp:836-872)
    int o0, o1;
    const float B1 = Bq[0], B2 = Bq[1], B3 = Bq[2];
    if (B1 < B2) {
        if (B3 < B1) { o0 = 2; o1 = (B1 < B2) ? 0 : 1; }
        else { o0 = 0; o1 = (B2 < B3) ? 1 : 2; }
    } else {
        if (B3 < B2) { o0 = 2; o1 = (B1 < B2) ? 0 : 1; }
        else { o0 = 1; o1 = (B1 < B3) ? 0 : 2; }
    }
    if (maxc > 1) {
#pragma unroll
        for (int t = 0; t < 2; t++) {
            int sd = t == 0 ? o0 : o1;
            float bsd = sd == 0 ? B1 : (sd == 1 ? B2 : B3);
            float asd = sd == 0 ? Aq[0] : (sd == 1 ? Aq[1] : Aq[2]);
            float ssd = sd == 0 ? side[0] : (sd == 1 ? side[1] : side[2]);
            if (depth - bsd < 0.f) break;
            // ret == 1 + t here: the loop stops at its first miss
#pragma unroll
            for (int k = 0; k < 3; k++) {
                float v = ssd * sel3(R + k * 4, sd);
                c[1 + t].pos[k] = asd > 0.f ? p[k] + v : p[k] - v;
            }
            c[1 + t].depth = depth - bsd;
            ret++;
            if (maxc == 2) break;
        }
    }
    if (maxc == 4 && ret == 3) {
        float d4 = c[1].depth + c[2].depth - depth;
        if (d4 > 0.f) {
#pragma unroll
            for (int k = 0; k < 3; k++) c[3].pos[k] = c[1].pos[k] + c[2].pos[k] - p[k];
            c[3].depth = d4;
            ret++;
        }
    }
    set_normals(c, ret, n[0], n[1], n[2]);
    return ret;
}

// dCollideCapsuleSphere capsule.cpp:130-161
ODX_DEV int collide_capsule_sphere(const DGeom &cp, const DGeom &sp, DContact &c) {
    const float *R = cp.R;
    float alpha = R[2] * (sp.pos[0] - cp.pos[0]) + R[6] * (sp.pos[1] - cp.pos[1]) + R[10] * (sp.pos[2] - cp.pos[2]);
    float lz2 = cp.p[1] * 0.5f;
    alpha = fminf(fmaxf(alpha, -lz2), lz2);
    float p[3] = {cp.pos[0] + alpha * R[2], cp.pos[1] + alpha * R[6], cp.pos[2] + alpha * R[10]};
    return collide_spheres(p, cp.p[0], sp.pos, sp.p[0], c);
}

// dClosestLineSegmentPoints collision_util.cpp:109-223
ODX_DEV void closest_segment_points(const float *a1, const float *a2, const float *b1, const float *b2, float *cp1,
                                    float *cp2) {
    float a1a2[3], b1b2[3], a1b1[3], a1b2[3], a2b1[3], a2b2[3], n[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        a1a2[i] = a2[i] - a1[i]; b1b2[i] = b2[i] - b1[i]; a1b1[i] = b1[i] - a1[i];
        a1b2[i] = b2[i] - a1[i]; a2b1[i] = b1[i] - a2[i]; a2b2[i] = b2[i] - a2[i];
    }
    float da1 = dot3(a1a2, a1b1), db1 = dot3(b1b2, a1b1);
    if (da1 <= 0.f && db1 >= 0.f) { for (int i = 0; i < 3; i++) { cp1[i] = a1[i]; cp2[i] = b1[i]; } return; }
    float da2 = dot3(a1a2, a1b2), db2 = dot3(b1b2, a1b2);
    if (da2 <= 0.f && db2 <= 0.f) { for (int i = 0; i < 3; i++) { cp1[i] = a1[i]; cp2[i] = b2[i]; } return; }
    float da3 = dot3(a1a2, a2b1), db3 = dot3(b1b2, a2b1);
    if (da3 >= 0.f && db3 >= 0.f) { for (int i = 0; i < 3; i++) { cp1[i] = a2[i]; cp2[i] = b1[i]; } return; }
    float da4 = dot3(a1a2, a2b2), db4 = dot3(b1b2, a2b2);
    if (da4 >= 0.f && db4 <= 0.f) { for (int i = 0; i < 3; i++) { cp1[i] = a2[i]; cp2[i] = b2[i]; } return; }
    float la = dot3(a1a2, a1a2), k;
    if (da1 >= 0.f && da3 <= 0.f) {
        k = da1 / la;
        for (int i = 0; i < 3; i++) n[i] = a1b1[i] - k * a1a2[i];
        if (dot3(b1b2, n) >= 0.f) { for (int i = 0; i < 3; i++) { cp1[i] = a1[i] + k * a1a2[i]; cp2[i] = b1[i]; } return; }
    }
    if (da2 >= 0.f && da4 <= 0.f) {
        k = da2 / la;
        for (int i = 0; i < 3; i++) n[i] = a1b2[i] - k * a1a2[i];
        if (dot3(b1b2, n) <= 0.f) { for (int i = 0; i < 3; i++) { cp1[i] = a1[i] + k * a1a2[i]; cp2[i] = b2[i]; } return; }
    }
    float lb = dot3(b1b2, b1b2);
    if (db1 <= 0.f && db2 >= 0.f) {
        k = -db1 / lb;
        for (int i = 0; i < 3; i++) n[i] = -a1b1[i] - k * b1b2[i];
        if (dot3(a1a2, n) >= 0.f) { for (int i = 0; i < 3; i++) { cp1[i] = a1[i]; cp2[i] = b1[i] + k * b1b2[i]; } return; }
    }
    if (db3 <= 0.f && db4 >= 0.f) {
        k = -db3 / lb;
        for (int i = 0; i < 3; i++) n[i] = -a2b1[i] - k * b1b2[i];
        if (dot3(a1a2, n) <= 0.f) { for (int i = 0; i < 3; i++) { cp1[i] = a2[i]; cp2[i] = b1[i] + k * b1b2[i]; } return; }
    }
    k = dot3(a1a2, b1b2);
    float det = la * lb - k * k;
    if (det <= 0.f) { for (int i = 0; i < 3; i++) { cp1[i] = a1[i]; cp2[i] = b1[i]; } return; }
    det = 1.f / det;
    float alpha = (lb * da1 - k * db1) * det, beta = (k * da1 - la * db1) * det;
    for (int i = 0; i < 3; i++) { cp1[i] = a1[i] + alpha * a1a2[i]; cp2[i] = b1[i] + beta * b1b2[i]; }
}

// dClosestLineBoxPoints collision_util.cpp:247-384
ODX_DEV void closest_line_box_points(const float *p1, const float *p2, const float *c, const float *R,
                                     const float *side, float *lret, float *bret) {
    float tmp[3], s[3], v[3], sign[3], v2[3], h[3];
    for (int i = 0; i < 3; i++) tmp[i] = p1[i] - c[i];
    mul1_331(s, R, tmp);
    for (int i = 0; i < 3; i++) tmp[i] = p2[i] - p1[i];
    mul1_331(v, R, tmp);
    for (int i = 0; i < 3; i++) {
        if (v[i] < 0.f) { s[i] = -s[i]; v[i] = -v[i]; sign[i] = -1.f; }
        else sign[i] = 1.f;
        v2[i] = v[i] * v[i];
        h[i] = 0.5f * side[i];
    }
    int region[3];
    float tanchor[3];
    for (int i = 0; i < 3; i++) {
        if (v[i] > 1e-19f) {
            if (s[i] < -h[i]) { region[i] = -1; tanchor[i] = (-h[i] - s[i]) / v[i]; }
            else { region[i] = (s[i] > h[i]); tanchor[i] = (h[i] - s[i]) / v[i]; }
        } else { region[i] = 0; tanchor[i] = 2.f; }
    }
    float t = 0.f, dd2dt = 0.f;
    for (int i = 0; i < 3; i++) dd2dt -= (region[i] ? v2[i] : 0.f) * tanchor[i];
    if (dd2dt < 0.f) {
        bool done = false;
        for (int guard = 0; guard < 8 && !done; guard++) {
            float next_t = 1.f;
            for (int i = 0; i < 3; i++) if (tanchor[i] > t && tanchor[i] < 1.f && tanchor[i] < next_t) next_t = tanchor[i];
            float next_dd2dt = 0.f;
            for (int i = 0; i < 3; i++) next_dd2dt += (region[i] ? v2[i] : 0.f) * (next_t - tanchor[i]);
            if (next_dd2dt >= 0.f) {
                float m = (next_dd2dt - dd2dt) / (next_t - t);
                t -= dd2dt / m;
                done = true;
                break;
            }
            for (int i = 0; i < 3; i++) if (tanchor[i] == next_t) { tanchor[i] = (h[i] - s[i]) / v[i]; region[i]++; }
            t = next_t;
            dd2dt = next_dd2dt;
            if (!(t < 1.f)) break;
        }
        if (!done) t = 1.f;
    }
    for (int i = 0; i < 3; i++) lret[i] = p1[i] + t * tmp[i];
    for (int i = 0; i < 3; i++) {
        float x = sign[i] * (s[i] + t * v[i]);
        tmp[i] = x < -h[i] ? -h[i] : (x > h[i] ? h[i] : x);
    }
    mul0_331(s, R, tmp);
    for (int i = 0; i < 3; i++) bret[i] = s[i] + c[i];
}

// dCollideCapsuleBox capsule.cpp:164-235
template <class W>
ODX_DEV int collide_capsule_box(const DGeom &cp, const DGeom &bx, int maxc, DContact *c, W &ws) {
    const float *R = cp.R;
    float clen = cp.p[1] * 0.5f, radius = cp.p[0];
    float p1[3] = {cp.pos[0] + clen * R[2], cp.pos[1] + clen * R[6], cp.pos[2] + clen * R[10]};
    float p2[3] = {cp.pos[0] - clen * R[2], cp.pos[1] - clen * R[6], cp.pos[2] - clen * R[10]};
    float pl[3], pb[3];
    closest_line_box_points(p1, p2, bx.pos, bx.R, bx.p, pl, pb);
    float dx = pl[0] - pb[0], dy = pl[1] - pb[1], dz = pl[2] - pb[2];
    if (sqrtf(dx * dx + dy * dy + dz * dz) < 1e-6f) {
        float normal[3];
        float r2 = radius * 2.f;
        float capside[3] = {r2, r2, cp.p[1] + r2};
        int n = box_box(bx.pos, bx.R, bx.p, cp.pos, cp.R, capside, normal, maxc, c, ws);
        set_normals(c, n, normal[0], normal[1], normal[2]);
        return n;
    }
    return collide_spheres(pl, radius, pb, 0.f, c[0]);
}

// dCollideCapsuleCapsule capsule.cpp:238-353
ODX_DEV int collide_capsule_capsule(const DGeom &g1, const DGeom &g2, int maxc, DContact *c) {
    float lz1 = g1.p[1] * 0.5f, lz2 = g2.p[1] * 0.5f;
    float ax1[3] = {g1.R[2], g1.R[6], g1.R[10]}, ax2[3] = {g2.R[2], g2.R[6], g2.R[10]};
    float s1[3], s2[3];
    float a1a2 = dot3(ax1, ax2);
    float det = 1.f - a1a2 * a1a2;
    if (det < 1e-5f) {
        if (a1a2 < 0.f) { ax2[0] = -ax2[0]; ax2[1] = -ax2[1]; ax2[2] = -ax2[2]; }
        float q[3] = {g1.pos[0] - g2.pos[0], g1.pos[1] - g2.pos[1], g1.pos[2] - g2.pos[2]};
        float k = dot3(ax1, q);
        float lo = fmaxf(-lz1, -lz2 - k), hi = fminf(lz1, lz2 - k);
        if (lo <= hi) {
            if (maxc >= 2 && lo < hi) {
                for (int i = 0; i < 3; i++) { s1[i] = g1.pos[i] + lo * ax1[i]; s2[i] = g2.pos[i] + (lo + k) * ax2[i]; }
                if (collide_spheres(s1, g1.p[0], s2, g2.p[0], c[0])) {
                    for (int i = 0; i < 3; i++) { s1[i] = g1.pos[i] + hi * ax1[i]; s2[i] = g2.pos[i] + (hi + k) * ax2[i]; }
                    if (collide_spheres(s1, g1.p[0], s2, g2.p[0], c[1])) return 2;
                }
            }
            float al1 = (lo + hi) * 0.5f, al2 = al1 + k;
            for (int i = 0; i < 3; i++) { s1[i] = g1.pos[i] + al1 * ax1[i]; s2[i] = g2.pos[i] + al2 * ax2[i]; }
            return collide_spheres(s1, g1.p[0], s2, g2.p[0], c[0]);
        }
    }
    float a1[3], a2[3], b1[3], b2[3];
    for (int i = 0; i < 3; i++) {
        a1[i] = g1.pos[i] + ax1[i] * lz1; a2[i] = g1.pos[i] - ax1[i] * lz1;
        b1[i] = g2.pos[i] + ax2[i] * lz2; b2[i] = g2.pos[i] - ax2[i] * lz2;
    }
    closest_segment_points(a1, a2, b1, b2, s1, s2);
    return collide_spheres(s1, g1.p[0], s2, g2.p[0], c[0]);
}

// dCollideCapsulePlane capsule.cpp:356-416
ODX_DEV int collide_capsule_plane(const DGeom &cp, const DGeom &pl, int maxc, DContact *c) {
    const float *R = cp.R, *n = pl.p;
    float lz = cp.p[1], r = cp.p[0];
    float sign = (dot14(n, R + 2) > 0.f) ? -1.f : 1.f;
    float p[3];
    for (int i = 0; i < 3; i++) p[i] = cp.pos[i] + R[i * 4 + 2] * lz * 0.5f * sign;
    float depth = n[3] - dot3(p, n) + r;
    if (depth < 0.f) return 0;
    for (int i = 0; i < 3; i++) { c[0].normal[i] = n[i]; c[0].pos[i] = p[i] - n[i] * r; }
    c[0].depth = depth;
    int nc = 1;
    if (maxc >= 2) {
        for (int i = 0; i < 3; i++) p[i] = cp.pos[i] - R[i * 4 + 2] * lz * 0.5f * sign;
        depth = n[3] - dot3(p, n) + r;
        if (depth >= 0.f) {
            for (int i = 0; i < 3; i++) { c[1].normal[i] = n[i]; c[1].pos[i] = p[i] - n[i] * r; }
            c[1].depth = depth;
            nc = 2;
        }
    }
    return nc;
}

// Forward collider table entry (cls1, cls2) or -1 when the reference has no
// forward entry (dInitColliders collision_kernel.cpp:166-268).
// dCollideCylinderPlane collision_cylinder_plane.cpp:42-250.  The disc ends
// e1 (-axis) and e2 (+axis).  When a disc is parallel to the plane (|a.n| = 1
// within 1e-4) the nearer disc gives up to four rim points, two orthogonal
// radius vectors V1, V2 from either side of its centre, each kept while
// depth > 0; otherwise each disc gives its rim point deepest along -n, kept
// while depth >= 0.  The expressions keep the reference's operation order
// (the second tilted depth is summed term by term, as there).
// Writes contact slot n (0 <= n < ODX_MAXC): every slot is rewritten with a
// select, so that the slots keep static indices and stay in registers.
ODX_DEV void put_contact(DContact *c, int n, const float *q, const float *nrm, float depth) {
#pragma unroll
    for (int s = 0; s < ODX_MAXC; s++) {
        const bool h = s == n;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            c[s].pos[i] = h ? q[i] : c[s].pos[i];
            c[s].normal[i] = h ? nrm[i] : c[s].normal[i];
        }
        c[s].depth = h ? depth : c[s].depth;
    }
}

ODX_DEV int collide_cylinder_plane(const DGeom &cy, const DGeom &pl, int maxc, DContact *c) {
    const float tol = 0.0001f;
    const float radius = cy.p[0], s0 = cy.p[1] * 0.5f;
    const float *cp = cy.pos, *pv = pl.p;
    const float dir[3] = {cy.R[2], cy.R[6], cy.R[10]};
    float e1[3], e2[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        e2[i] = dir[i] * s0 + cp[i];
        e1[i] = dir[i] * -s0 + cp[i];
    }
    int n = 0;
    float s = dir[0] * pv[0] + dir[1] * pv[1] + dir[2] * pv[2];
    s = s < 0.f ? s + 1.f : s - 1.f;
    if (s < tol && s > -tol) {
        const float d1 = pv[3] - dot3(pv, e1), d2 = pv[3] - dot3(pv, e2);
        const bool first = d1 >= d2;
        if (!((first ? d1 : d2) >= 0.f)) return 0;
        const float *P = first ? e1 : e2;
        float V1[3] = {dir[0], dir[1], dir[2]}, V2[3];
        if (dir[0] < tol && dir[0] > -tol) V1[0] += 1.f;
        else V1[1] += 1.f;
        cross3(V2, V1, dir);
        const float t = radius / sqrtf(V2[0] * V2[0] + V2[1] * V2[1] + V2[2] * V2[2]);
        V2[0] *= t; V2[1] *= t; V2[2] *= t;
        cross3(V1, V2, dir);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (n >= maxc) break;
            const float *off = k < 2 ? V1 : V2;
            float q[3];
#pragma unroll
            for (int i = 0; i < 3; i++) q[i] = (k & 1) ? P[i] - off[i] : P[i] + off[i];
            const float depth = pv[3] - dot3(pv, q);
            if (depth > 0.f) {
                put_contact(c, n, q, pv, depth);
                n++;
            }
        }
        return n;
    }
    const float t = dot3(pv, dir);
    float C[3] = {dir[0] * t - pv[0], dir[1] * t - pv[1], dir[2] * t - pv[2]};
    const float l = radius / sqrtf(C[0] * C[0] + C[1] * C[1] + C[2] * C[2]);
    C[0] *= l; C[1] *= l; C[2] *= l;
    float q[3] = {C[0] + e1[0], C[1] + e1[1], C[2] + e1[2]};
    float depth = pv[3] - dot3(pv, q);
    if (depth >= 0.f) {
#pragma unroll
        for (int i = 0; i < 3; i++) { c[0].pos[i] = q[i]; c[0].normal[i] = pv[i]; }
        c[0].depth = depth;
        if (++n >= maxc) return n;
    }
    q[0] = C[0] + e2[0]; q[1] = C[1] + e2[1]; q[2] = C[2] + e2[2];
    depth = pv[3] - pv[0] * q[0] - pv[1] * q[1] - pv[2] * q[2];
    if (depth >= 0.f) {
        put_contact(c, n, q, pv, depth);
        n++;
    }
    return n;
}

template <class W>
ODX_DEV int collide_fwd(const DGeom &a, const DGeom &b, int maxc, DContact *c, W &ws) {
    switch (a.cls * 8 + b.cls) {
    case CLS_SPHERE * 8 + CLS_SPHERE: return collide_spheres(a.pos, a.p[0], b.pos, b.p[0], c[0]);
    case CLS_SPHERE * 8 + CLS_BOX: return collide_sphere_box(a, b, c[0]);
    case CLS_SPHERE * 8 + CLS_PLANE: return collide_sphere_plane(a, b, c[0]);
    case CLS_BOX * 8 + CLS_BOX: return collide_box_box(a, b, maxc, c, ws);
    case CLS_BOX * 8 + CLS_PLANE: return collide_box_plane(a, b, maxc, c);
    case CLS_CAPSULE * 8 + CLS_SPHERE: return collide_capsule_sphere(a, b, c[0]);
    case CLS_CAPSULE * 8 + CLS_BOX: return collide_capsule_box(a, b, maxc, c, ws);
    case CLS_CAPSULE * 8 + CLS_CAPSULE: return collide_capsule_capsule(a, b, maxc, c);
    case CLS_CAPSULE * 8 + CLS_PLANE: return collide_capsule_plane(a, b, maxc, c);
    case CLS_CYLINDER * 8 + CLS_PLANE: return collide_cylinder_plane(a, b, maxc, c);
    default: return -1;
    }
}

// dCollide: forward entry, else the reversed entry with negated normals.
// maxc must be in 1..ODX_MAXC (flags & 0xffff of the reference, capped).
// ws: the box-box clipper's working arrays (ClipPriv or ClipLds).
template <class W>
ODX_DEV int dcollide(const DGeom &g1, const DGeom &g2, int maxc, DContact *c, W &ws) {
    if (g1.body == g2.body && g1.body >= 0) return 0;
    int n = collide_fwd(g1, g2, maxc, c, ws);
    if (n >= 0) return n;
    n = collide_fwd(g2, g1, maxc, c, ws);
    if (n < 0) return 0;
#pragma unroll
    for (int i = 0; i < ODX_MAXC; i++)
        if (i < n) { c[i].normal[0] = -c[i].normal[0]; c[i].normal[1] = -c[i].normal[1]; c[i].normal[2] = -c[i].normal[2]; }
    return n;
}

}  // namespace odx
