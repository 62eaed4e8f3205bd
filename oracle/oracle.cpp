/*
 * oracle.cpp -- sequential CPU restatement of ODE's step hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker for the HIP implementation; never
 * linked into the product).  Each function names the reference file:line it
 * restates; paths are relative to the root of thomasmarsh/ODE.
 *
 * Build: oracle/build.sh -> oracle/liboracle.so (g++ -O2, no fast-math, no FMA
 * contraction so float and double runs follow the reference's operation order).
 */
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <functional>
#include <vector>

namespace {

/* ------------------------------------------------------------------ math -- */
/* 3x4 row-major matrices and 3-vectors, as include/ode/common.h:270-275. */
template <class R> struct M3 { R m[12]; };

template <class R> inline R dot3(const R *a, const R *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
/* dCalcVectorDot3_14: a . (column of a 3x4 matrix) */
template <class R> inline R dot14(const R *a, const R *b) { return a[0] * b[0] + a[1] * b[4] + a[2] * b[8]; }
template <class R> inline R dot41(const R *a, const R *b) { return a[0] * b[0] + a[4] * b[1] + a[8] * b[2]; }
template <class R> inline R dot44(const R *a, const R *b) { return a[0] * b[0] + a[4] * b[4] + a[8] * b[8]; }
template <class R> inline void cross3(R *r, const R *a, const R *b) {
    R r0 = a[1] * b[2] - a[2] * b[1];
    R r1 = a[2] * b[0] - a[0] * b[2];
    R r2 = a[0] * b[1] - a[1] * b[0];
    r[0] = r0; r[1] = r1; r[2] = r2;
}
/* dMultiply0_331: r = A b */
template <class R> inline void mul0_331(R *r, const R *A, const R *b) {
    R x = A[0] * b[0] + A[1] * b[1] + A[2] * b[2];
    R y = A[4] * b[0] + A[5] * b[1] + A[6] * b[2];
    R z = A[8] * b[0] + A[9] * b[1] + A[10] * b[2];
    r[0] = x; r[1] = y; r[2] = z;
}
/* dMultiply1_331: r = A^T b */
template <class R> inline void mul1_331(R *r, const R *A, const R *b) {
    R x = A[0] * b[0] + A[4] * b[1] + A[8] * b[2];
    R y = A[1] * b[0] + A[5] * b[1] + A[9] * b[2];
    R z = A[2] * b[0] + A[6] * b[1] + A[10] * b[2];
    r[0] = x; r[1] = y; r[2] = z;
}
/* dMultiply0_333: A = B C ; dMultiply2_333: A = B C^T (3x4 storage) */
template <class R> inline void mul0_333(R *A, const R *B, const R *C) {
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++)
            A[i * 4 + j] = B[i * 4 + 0] * C[0 * 4 + j] + B[i * 4 + 1] * C[1 * 4 + j] + B[i * 4 + 2] * C[2 * 4 + j];
        A[i * 4 + 3] = 0;
    }
}
template <class R> inline void mul2_333(R *A, const R *B, const R *C) {
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++)
            A[i * 4 + j] = B[i * 4 + 0] * C[j * 4 + 0] + B[i * 4 + 1] * C[j * 4 + 1] + B[i * 4 + 2] * C[j * 4 + 2];
        A[i * 4 + 3] = 0;
    }
}
template <class R> inline R recip(R x) { return R(1) / x; }
template <class R> inline R rsqrt(R x) { return R(1) / std::sqrt(x); }

/* dRfromQ rotation.cpp:236-255 */
template <class R> void RfromQ(R *Rm, const R *q) {
    R qq1 = 2 * q[1] * q[1], qq2 = 2 * q[2] * q[2], qq3 = 2 * q[3] * q[3];
    Rm[0] = 1 - qq2 - qq3; Rm[1] = 2 * (q[1] * q[2] - q[0] * q[3]); Rm[2] = 2 * (q[1] * q[3] + q[0] * q[2]); Rm[3] = 0;
    Rm[4] = 2 * (q[1] * q[2] + q[0] * q[3]); Rm[5] = 1 - qq1 - qq3; Rm[6] = 2 * (q[2] * q[3] - q[0] * q[1]); Rm[7] = 0;
    Rm[8] = 2 * (q[1] * q[3] - q[0] * q[2]); Rm[9] = 2 * (q[2] * q[3] + q[0] * q[1]); Rm[10] = 1 - qq1 - qq2; Rm[11] = 0;
}
/* dDQfromW rotation.cpp:310-317 */
template <class R> void DQfromW(R *dq, const R *w, const R *q) {
    dq[0] = R(0.5) * (-w[0] * q[1] - w[1] * q[2] - w[2] * q[3]);
    dq[1] = R(0.5) * (w[0] * q[0] + w[1] * q[3] - w[2] * q[2]);
    dq[2] = R(0.5) * (-w[0] * q[3] + w[1] * q[0] + w[2] * q[1]);
    dq[3] = R(0.5) * (w[0] * q[2] - w[1] * q[1] + w[2] * q[0]);
}
/* dxSafeNormalize4 odemath.cpp:203-221 with the dxNormalize4 fallback (odemath.h:48-57) */
template <class R> void normalize4(R *a) {
    R l = a[0] * a[0] + a[1] * a[1] + a[2] * a[2] + a[3] * a[3];
    if (l > 0) {
        l = rsqrt(l);
        a[0] *= l; a[1] *= l; a[2] *= l; a[3] *= l;
    } else {
        a[0] = 1; a[1] = a[2] = a[3] = 0;
    }
}
/* dxSafeNormalize3 odemath.cpp:94-160 with the dxNormalize3 fallback (odemath.h:36-45) */
template <class R> void normalize3(R *a) {
    R a0 = std::fabs(a[0]), a1 = std::fabs(a[1]), a2 = std::fabs(a[2]);
    int idx;
    if (a1 > a0) idx = (a2 > a1) ? 2 : 1;
    else if (a2 > a0) idx = 2;
    else {
        if (!(a0 > R(0))) { a[0] = 1; a[1] = a[2] = 0; return; }
        idx = 0;
    }
    if (idx == 0) {
        R rr = recip(a0), y = a[1] * rr, z = a[2] * rr, l = rsqrt(R(1) + y * y + z * z);
        a[1] = y * l; a[2] = z * l; a[0] = std::copysign(l, a[0]);
    } else if (idx == 1) {
        R rr = recip(a1), x = a[0] * rr, z = a[2] * rr, l = rsqrt(R(1) + x * x + z * z);
        a[0] = x * l; a[2] = z * l; a[1] = std::copysign(l, a[1]);
    } else {
        R rr = recip(a2), x = a[0] * rr, y = a[1] * rr, l = rsqrt(R(1) + x * x + y * y);
        a[0] = x * l; a[1] = y * l; a[2] = std::copysign(l, a[2]);
    }
}
/* dxPlaneSpace odemath.cpp:224-251 */
template <class R> void planeSpace(const R *n, R *p, R *q) {
    if (std::fabs(n[2]) > R(M_SQRT1_2)) {
        R a = n[1] * n[1] + n[2] * n[2], k = rsqrt(a);
        p[0] = 0; p[1] = -n[2] * k; p[2] = n[1] * k;
        q[0] = a * k; q[1] = -n[0] * p[2]; q[2] = n[0] * p[1];
    } else {
        R a = n[0] * n[0] + n[1] * n[1], k = rsqrt(a);
        p[0] = -n[1] * k; p[1] = n[0] * k; p[2] = 0;
        q[0] = -n[2] * p[1]; q[1] = n[2] * p[0]; q[2] = a * k;
    }
}
/* dCalcMatrix3Det / dInvertMatrix3 include/ode/odemath.h:444-490 */
template <class R> R invert3(R *dst, const R *ma) {
    R det = ma[0] * (ma[5] * ma[10] - ma[9] * ma[6]) - ma[1] * (ma[4] * ma[10] - ma[8] * ma[6]) +
            ma[2] * (ma[4] * ma[9] - ma[8] * ma[5]);
    if (det == 0) return 0;
    R d = recip(det);
    dst[0] = (ma[5] * ma[10] - ma[6] * ma[9]) * d;
    dst[1] = (ma[9] * ma[2] - ma[1] * ma[10]) * d;
    dst[2] = (ma[1] * ma[6] - ma[5] * ma[2]) * d;
    dst[3] = 0;
    dst[4] = (ma[6] * ma[8] - ma[4] * ma[10]) * d;
    dst[5] = (ma[0] * ma[10] - ma[8] * ma[2]) * d;
    dst[6] = (ma[4] * ma[2] - ma[0] * ma[6]) * d;
    dst[7] = 0;
    dst[8] = (ma[4] * ma[9] - ma[8] * ma[5]) * d;
    dst[9] = (ma[8] * ma[1] - ma[0] * ma[9]) * d;
    dst[10] = (ma[0] * ma[5] - ma[1] * ma[4]) * d;
    dst[11] = 0;
    return det;
}
/* dInvertPDMatrix (matrix.cpp:563) for the 3x3 body inertia, via Cholesky
 * L L^T = A and A^-1 = L^-T L^-1 (used by dBodySetMass, ode.cpp:488-504). */
template <class R> bool invertPD3(R *out /*3x4*/, const R *I /*3x4*/) {
    R a[3][3];
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) a[i][j] = I[i * 4 + j];
    R L[3][3] = {{0}};
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j <= i; j++) {
            R s = a[i][j];
            for (int k = 0; k < j; k++) s -= L[i][k] * L[j][k];
            if (i == j) {
                if (s <= 0) return false;
                L[i][i] = std::sqrt(s);
            } else {
                L[i][j] = s / L[j][j];
            }
        }
    }
    /* Li = L^-1 (lower triangular) */
    R Li[3][3] = {{0}};
    for (int i = 0; i < 3; i++) {
        Li[i][i] = R(1) / L[i][i];
        for (int j = 0; j < i; j++) {
            R s = 0;
            for (int k = j; k < i; k++) s += L[i][k] * Li[k][j];
            Li[i][j] = -s / L[i][i];
        }
    }
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) {
            R s = 0;
            for (int k = 0; k < 3; k++) s += Li[k][i] * Li[k][j];
            out[i * 4 + j] = s;
        }
        out[i * 4 + 3] = 0;
    }
    return true;
}

const double kInf = std::numeric_limits<double>::infinity();
thread_local int g_last_sweeps = 0;  /* per calling thread (CPU baselines step worlds on several threads) */
unsigned int g_rand_seed = 0;  /* dRand's global state (misc.cpp:35) */
extern "C" int orc_drand_int(unsigned int *seed, int n);

/* ------------------------------------------------------------- records -- */
enum { CLS_SPHERE = 0, CLS_BOX = 1, CLS_CAPSULE = 2, CLS_CYLINDER = 3, CLS_PLANE = 4 };

template <class R> struct Pose { R pos[4]; R Rm[12]; };

template <class R> struct Contact {
    R pos[3], normal[3], depth;
    int g1, g2, side1, side2;
};

template <class R> struct GeomView {
    int cls, body;
    R p[4];
    Pose<R> pose;
};

template <class R> GeomView<R> geom_view(const R *g, const R *bodies) {
    GeomView<R> v;
    v.cls = (int)g[0];
    v.body = (int)g[1];
    for (int i = 0; i < 4; i++) v.p[i] = g[2 + i];
    if (v.body >= 0) {
        const R *b = bodies + (size_t)v.body * ORC_BODY_STRIDE;
        for (int i = 0; i < 3; i++) v.pose.pos[i] = b[i];
        RfromQ(v.pose.Rm, b + 4);
    } else {
        for (int i = 0; i < 3; i++) v.pose.pos[i] = g[6 + i];
        R q[4] = {g[9], g[10], g[11], g[12]};
        RfromQ(v.pose.Rm, q);
    }
    v.pose.pos[3] = 0;
    return v;
}

/* computeAABB: sphere.cpp:59-67, box.cpp:60-76, capsule.cpp:60-74, plane.cpp:81-105 */
template <class R> void geom_aabb(const GeomView<R> &g, R *aabb) {
    const R *p = g.pose.pos, *Rm = g.pose.Rm;
    R inf = (R)kInf;
    switch (g.cls) {
    case CLS_SPHERE: {
        R r = g.p[0];
        aabb[0] = p[0] - r; aabb[1] = p[0] + r; aabb[2] = p[1] - r; aabb[3] = p[1] + r; aabb[4] = p[2] - r; aabb[5] = p[2] + r;
        break;
    }
    case CLS_BOX: {
        const R *s = g.p;
        R xr = R(0.5) * (std::fabs(Rm[0] * s[0]) + std::fabs(Rm[1] * s[1]) + std::fabs(Rm[2] * s[2]));
        R yr = R(0.5) * (std::fabs(Rm[4] * s[0]) + std::fabs(Rm[5] * s[1]) + std::fabs(Rm[6] * s[2]));
        R zr = R(0.5) * (std::fabs(Rm[8] * s[0]) + std::fabs(Rm[9] * s[1]) + std::fabs(Rm[10] * s[2]));
        aabb[0] = p[0] - xr; aabb[1] = p[0] + xr; aabb[2] = p[1] - yr; aabb[3] = p[1] + yr; aabb[4] = p[2] - zr; aabb[5] = p[2] + zr;
        break;
    }
    case CLS_CAPSULE: {
        R r = g.p[0], lz = g.p[1];
        R xr = std::fabs(Rm[2] * lz) * R(0.5) + r;
        R yr = std::fabs(Rm[6] * lz) * R(0.5) + r;
        R zr = std::fabs(Rm[10] * lz) * R(0.5) + r;
        aabb[0] = p[0] - xr; aabb[1] = p[0] + xr; aabb[2] = p[1] - yr; aabb[3] = p[1] + yr; aabb[4] = p[2] - zr; aabb[5] = p[2] + zr;
        break;
    }
    case CLS_CYLINDER: {  /* cylinder.cpp:61-78 */
        R r = g.p[0], lz = g.p[1];
        R o2 = R(1) - Rm[2] * Rm[2], o6 = R(1) - Rm[6] * Rm[6], o10 = R(1) - Rm[10] * Rm[10];
        R xr = std::fabs(Rm[2] * lz * R(0.5)) + r * std::sqrt(R(0) > o2 ? R(0) : o2);
        R yr = std::fabs(Rm[6] * lz * R(0.5)) + r * std::sqrt(R(0) > o6 ? R(0) : o6);
        R zr = std::fabs(Rm[10] * lz * R(0.5)) + r * std::sqrt(R(0) > o10 ? R(0) : o10);
        aabb[0] = p[0] - xr; aabb[1] = p[0] + xr; aabb[2] = p[1] - yr; aabb[3] = p[1] + yr; aabb[4] = p[2] - zr; aabb[5] = p[2] + zr;
        break;
    }
    case CLS_PLANE: {
        const R *pl = g.p;
        aabb[0] = -inf; aabb[1] = inf; aabb[2] = -inf; aabb[3] = inf; aabb[4] = -inf; aabb[5] = inf;
        if (pl[1] == 0 && pl[2] == 0) {
            aabb[0] = (pl[0] > 0) ? -inf : -pl[3];
            aabb[1] = (pl[0] > 0) ? pl[3] : inf;
        } else if (pl[0] == 0 && pl[2] == 0) {
            aabb[2] = (pl[1] > 0) ? -inf : -pl[3];
            aabb[3] = (pl[1] > 0) ? pl[3] : inf;
        } else if (pl[0] == 0 && pl[1] == 0) {
            aabb[4] = (pl[2] > 0) ? -inf : -pl[3];
            aabb[5] = (pl[2] > 0) ? pl[3] : inf;
        }
        break;
    }
    default:
        for (int i = 0; i < 6; i++) aabb[i] = (i & 1) ? inf : -inf;
    }
}

/* ----------------------------------------------------------- colliders -- */
/* dCollideSpheres collision_util.cpp:38-67 */
template <class R> int collide_spheres(const R *p1, R r1, const R *p2, R r2, Contact<R> *c) {
    R dx = p1[0] - p2[0], dy = p1[1] - p2[1], dz = p1[2] - p2[2];
    R d = std::sqrt(dx * dx + dy * dy + dz * dz);
    if (d > (r1 + r2)) return 0;
    if (d <= 0) {
        c->pos[0] = p1[0]; c->pos[1] = p1[1]; c->pos[2] = p1[2];
        c->normal[0] = 1; c->normal[1] = 0; c->normal[2] = 0;
        c->depth = r1 + r2;
    } else {
        R d1 = recip(d);
        c->normal[0] = (p1[0] - p2[0]) * d1;
        c->normal[1] = (p1[1] - p2[1]) * d1;
        c->normal[2] = (p1[2] - p2[2]) * d1;
        R k = R(0.5) * (r2 - r1 - d);
        c->pos[0] = p1[0] + c->normal[0] * k;
        c->pos[1] = p1[1] + c->normal[1] * k;
        c->pos[2] = p1[2] + c->normal[2] * k;
        c->depth = r1 + r2 - d;
    }
    return 1;
}

/* dCollideSphereBox sphere.cpp:131-219 */
template <class R> int collide_sphere_box(const GeomView<R> &s, const GeomView<R> &b, Contact<R> *c) {
    const R *Rb = b.pose.Rm;
    R p[3], t[3], l[3];
    int onborder = 0;
    for (int i = 0; i < 3; i++) p[i] = s.pose.pos[i] - b.pose.pos[i];
    for (int i = 0; i < 3; i++) {
        l[i] = b.p[i] * R(0.5);
        t[i] = dot14(p, Rb + i);
        if (t[i] < -l[i]) { t[i] = -l[i]; onborder = 1; }
        if (t[i] > l[i]) { t[i] = l[i]; onborder = 1; }
    }
    if (!onborder) {
        R mind = l[0] - std::fabs(t[0]);
        int mini = 0;
        for (int i = 1; i < 3; i++) {
            R fd = l[i] - std::fabs(t[i]);
            if (fd < mind) { mind = fd; mini = i; }
        }
        for (int i = 0; i < 3; i++) c->pos[i] = s.pose.pos[i];
        R tmp[3] = {0, 0, 0};
        tmp[mini] = (t[mini] > 0) ? R(1) : R(-1);
        mul0_331(c->normal, Rb, tmp);
        c->depth = mind + s.p[0];
        return 1;
    }
    R q[3], r[3];
    mul0_331(q, Rb, t);
    for (int i = 0; i < 3; i++) r[i] = p[i] - q[i];
    R depth = s.p[0] - std::sqrt(dot3(r, r));
    if (depth < 0) return 0;
    for (int i = 0; i < 3; i++) c->pos[i] = q[i] + b.pose.pos[i];
    for (int i = 0; i < 3; i++) c->normal[i] = r[i];
    normalize3(c->normal);
    c->depth = depth;
    return 1;
}

/* dCollideSpherePlane sphere.cpp:222-251 */
template <class R> int collide_sphere_plane(const GeomView<R> &s, const GeomView<R> &pl, Contact<R> *c) {
    R k = dot3(s.pose.pos, pl.p);
    R depth = pl.p[3] - k + s.p[0];
    if (depth < 0) return 0;
    for (int i = 0; i < 3; i++) {
        c->normal[i] = pl.p[i];
        c->pos[i] = s.pose.pos[i] - pl.p[i] * s.p[0];
    }
    c->depth = depth;
    return 1;
}

/* intersectRectQuad box.cpp:212-263 */
template <class R> int intersectRectQuad(const R h[2], R p[8], R ret[16]) {
    int nq = 4, nr = 0;
    R buffer[16];
    R *q = p, *r = ret;
    for (int dir = 0; dir <= 1; dir++) {
        for (int sign = -1; sign <= 1; sign += 2) {
            R *pq = q, *pr = r;
            nr = 0;
            for (int i = nq; i > 0; i--) {
                if (sign * pq[dir] < h[dir]) {
                    pr[0] = pq[0]; pr[1] = pq[1];
                    pr += 2; nr++;
                    if (nr & 8) { q = r; goto done; }
                }
                R *nextq = (i > 1) ? pq + 2 : q;
                if ((sign * pq[dir] < h[dir]) ^ (sign * nextq[dir] < h[dir])) {
                    pr[1 - dir] = pq[1 - dir] + (nextq[1 - dir] - pq[1 - dir]) / (nextq[dir] - pq[dir]) * (sign * h[dir] - pq[dir]);
                    pr[dir] = sign * h[dir];
                    pr += 2; nr++;
                    if (nr & 8) { q = r; goto done; }
                }
                pq += 2;
            }
            q = r;
            r = (q == ret) ? buffer : ret;
            nq = nr;
        }
    }
done:
    if (q != ret) std::memcpy(ret, q, nr * 2 * sizeof(R));
    return nr;
}

/* cullPoints box.cpp:274-336 */
template <class R> void cullPoints(int n, const R p[], int m, int i0, int iret[]) {
    R a, cx, cy, q;
    if (n == 1) { cx = p[0]; cy = p[1]; }
    else if (n == 2) { cx = R(0.5) * (p[0] + p[2]); cy = R(0.5) * (p[1] + p[3]); }
    else {
        a = 0; cx = 0; cy = 0;
        for (int i = 0; i < n - 1; i++) {
            q = p[i * 2] * p[i * 2 + 3] - p[i * 2 + 2] * p[i * 2 + 1];
            a += q;
            cx += q * (p[i * 2] + p[i * 2 + 2]);
            cy += q * (p[i * 2 + 1] + p[i * 2 + 3]);
        }
        q = p[n * 2 - 2] * p[1] - p[0] * p[n * 2 - 1];
        a = recip(R(3) * (a + q));
        cx = a * (cx + q * (p[n * 2 - 2] + p[0]));
        cy = a * (cy + q * (p[n * 2 - 1] + p[1]));
    }
    R A[8];
    for (int i = 0; i < n; i++) A[i] = std::atan2(p[i * 2 + 1] - cy, p[i * 2] - cx);
    int avail[8];
    for (int i = 0; i < n; i++) avail[i] = 1;
    avail[i0] = 0;
    iret[0] = i0;
    int *ir = iret + 1;
    for (int j = 1; j < m; j++) {
        a = (R)(R(j) * (2 * M_PI / m) + A[i0]);
        if (a > M_PI) a -= (R)(2 * M_PI);
        R maxdiff = (R)1e9, diff;
        *ir = i0;
        for (int i = 0; i < n; i++) {
            if (avail[i]) {
                diff = std::fabs(A[i] - a);
                if (diff > M_PI) diff = (R)(2 * M_PI - diff);
                if (diff < maxdiff) { maxdiff = diff; *ir = i; }
            }
        }
        avail[*ir] = 0;
        ir++;
    }
}

/* dLineClosestApproach collision_util.cpp:70-95 */
template <class R> void lineClosestApproach(const R *pa, const R *ua, const R *pb, const R *ub, R *alpha, R *beta) {
    R p[3] = {pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2]};
    R uaub = dot3(ua, ub), q1 = dot3(ua, p), q2 = -dot3(ub, p);
    R d = 1 - uaub * uaub;
    if (d <= R(0.0001)) { *alpha = 0; *beta = 0; }
    else {
        d = recip(d);
        *alpha = (q1 + uaub * q2) * d;
        *beta = (uaub * q1 + q2) * d;
    }
}

/* dBoxBox box.cpp:356-737 (CONTACTS_UNIMPORTANT not used by this path).
 * Fills pos/depth of up to maxc contacts, the normal and the return code. */
template <class R> int boxBox(const R *p1, const R *R1, const R *side1, const R *p2, const R *R2, const R *side2,
                              R *normal, R *depth, int *return_code, int maxc_flags, Contact<R> *contact) {
    const R fudge_factor = R(1.05);
    R p[3], pp[3], normalC[3] = {0, 0, 0};
    const R *normalR = 0;
    R A[3], B[3], R11, R12, R13, R21, R22, R23, R31, R32, R33, Q11, Q12, Q13, Q21, Q22, Q23, Q31, Q32, Q33, s, s2, l, e;
    int invert_normal, code;
    for (int i = 0; i < 3; i++) p[i] = p2[i] - p1[i];
    mul1_331(pp, R1, p);
    for (int i = 0; i < 3; i++) { A[i] = side1[i] * R(0.5); B[i] = side2[i] * R(0.5); }
    R11 = dot44(R1 + 0, R2 + 0); R12 = dot44(R1 + 0, R2 + 1); R13 = dot44(R1 + 0, R2 + 2);
    R21 = dot44(R1 + 1, R2 + 0); R22 = dot44(R1 + 1, R2 + 1); R23 = dot44(R1 + 1, R2 + 2);
    R31 = dot44(R1 + 2, R2 + 0); R32 = dot44(R1 + 2, R2 + 1); R33 = dot44(R1 + 2, R2 + 2);
    Q11 = std::fabs(R11); Q12 = std::fabs(R12); Q13 = std::fabs(R13);
    Q21 = std::fabs(R21); Q22 = std::fabs(R22); Q23 = std::fabs(R23);
    Q31 = std::fabs(R31); Q32 = std::fabs(R32); Q33 = std::fabs(R33);

    s = (R)-kInf;
    invert_normal = 0;
    code = 0;
#define ORC_TSTF(expr1, expr2, norm, cc)                  \
    e = (expr1);                                          \
    s2 = std::fabs(e) - (expr2);                          \
    if (s2 > 0) return 0;                                 \
    if (s2 > s) { s = s2; normalR = norm; invert_normal = (e < 0); code = (cc); }
    ORC_TSTF(pp[0], (A[0] + B[0] * Q11 + B[1] * Q12 + B[2] * Q13), R1 + 0, 1);
    ORC_TSTF(pp[1], (A[1] + B[0] * Q21 + B[1] * Q22 + B[2] * Q23), R1 + 1, 2);
    ORC_TSTF(pp[2], (A[2] + B[0] * Q31 + B[1] * Q32 + B[2] * Q33), R1 + 2, 3);
    ORC_TSTF(dot41(R2 + 0, p), (A[0] * Q11 + A[1] * Q21 + A[2] * Q31 + B[0]), R2 + 0, 4);
    ORC_TSTF(dot41(R2 + 1, p), (A[0] * Q12 + A[1] * Q22 + A[2] * Q32 + B[1]), R2 + 1, 5);
    ORC_TSTF(dot41(R2 + 2, p), (A[0] * Q13 + A[1] * Q23 + A[2] * Q33 + B[2]), R2 + 2, 6);
#undef ORC_TSTF
#define ORC_TSTE(expr1, expr2, n1, n2, n3, cc)                                           \
    e = (expr1);                                                                         \
    s2 = std::fabs(e) - (expr2);                                                         \
    if (s2 > 0) return 0;                                                                \
    l = std::sqrt((n1) * (n1) + (n2) * (n2) + (n3) * (n3));                              \
    if (l > 0) {                                                                         \
        s2 /= l;                                                                         \
        if (s2 * fudge_factor > s) {                                                     \
            s = s2; normalR = 0;                                                         \
            normalC[0] = (n1) / l; normalC[1] = (n2) / l; normalC[2] = (n3) / l;         \
            invert_normal = (e < 0); code = (cc);                                        \
        }                                                                                \
    }
    ORC_TSTE(pp[2] * R21 - pp[1] * R31, (A[1] * Q31 + A[2] * Q21 + B[1] * Q13 + B[2] * Q12), 0, -R31, R21, 7);
    ORC_TSTE(pp[2] * R22 - pp[1] * R32, (A[1] * Q32 + A[2] * Q22 + B[0] * Q13 + B[2] * Q11), 0, -R32, R22, 8);
    ORC_TSTE(pp[2] * R23 - pp[1] * R33, (A[1] * Q33 + A[2] * Q23 + B[0] * Q12 + B[1] * Q11), 0, -R33, R23, 9);
    ORC_TSTE(pp[0] * R31 - pp[2] * R11, (A[0] * Q31 + A[2] * Q11 + B[1] * Q23 + B[2] * Q22), R31, 0, -R11, 10);
    ORC_TSTE(pp[0] * R32 - pp[2] * R12, (A[0] * Q32 + A[2] * Q12 + B[0] * Q23 + B[2] * Q21), R32, 0, -R12, 11);
    ORC_TSTE(pp[0] * R33 - pp[2] * R13, (A[0] * Q33 + A[2] * Q13 + B[0] * Q22 + B[1] * Q21), R33, 0, -R13, 12);
    ORC_TSTE(pp[1] * R11 - pp[0] * R21, (A[0] * Q21 + A[1] * Q11 + B[1] * Q33 + B[2] * Q32), -R21, R11, 0, 13);
    ORC_TSTE(pp[1] * R12 - pp[0] * R22, (A[0] * Q22 + A[1] * Q12 + B[0] * Q33 + B[2] * Q31), -R22, R12, 0, 14);
    ORC_TSTE(pp[1] * R13 - pp[0] * R23, (A[0] * Q23 + A[1] * Q13 + B[0] * Q32 + B[1] * Q31), -R23, R13, 0, 15);
#undef ORC_TSTE
    if (!code) return 0;

    if (normalR) { normal[0] = normalR[0]; normal[1] = normalR[4]; normal[2] = normalR[8]; }
    else mul0_331(normal, R1, normalC);
    if (invert_normal) { normal[0] = -normal[0]; normal[1] = -normal[1]; normal[2] = -normal[2]; }
    *depth = -s;

    if (code > 6) {
        R pa[3], pb[3], sign, alpha, beta, ua[3], ub[3];
        for (int i = 0; i < 3; i++) pa[i] = p1[i];
        for (int j = 0; j < 3; j++) {
            sign = (dot14(normal, R1 + j) > 0) ? R(1) : R(-1);
            for (int i = 0; i < 3; i++) pa[i] += sign * A[j] * R1[i * 4 + j];
        }
        for (int i = 0; i < 3; i++) pb[i] = p2[i];
        for (int j = 0; j < 3; j++) {
            sign = (dot14(normal, R2 + j) > 0) ? R(-1) : R(1);
            for (int i = 0; i < 3; i++) pb[i] += sign * B[j] * R2[i * 4 + j];
        }
        for (int i = 0; i < 3; i++) ua[i] = R1[((code) - 7) / 3 + i * 4];
        for (int i = 0; i < 3; i++) ub[i] = R2[((code) - 7) % 3 + i * 4];
        lineClosestApproach(pa, ua, pb, ub, &alpha, &beta);
        for (int i = 0; i < 3; i++) pa[i] += ua[i] * alpha;
        for (int i = 0; i < 3; i++) pb[i] += ub[i] * beta;
        for (int i = 0; i < 3; i++) contact[0].pos[i] = R(0.5) * (pa[i] + pb[i]);
        contact[0].depth = *depth;
        *return_code = code;
        return 1;
    }

    const R *Ra, *Rb, *pa, *pb, *Sa, *Sb;
    if (code <= 3) { Ra = R1; Rb = R2; pa = p1; pb = p2; Sa = A; Sb = B; }
    else { Ra = R2; Rb = R1; pa = p2; pb = p1; Sa = B; Sb = A; }
    R normal2[3], nr[3], anr[3];
    if (code <= 3) { normal2[0] = normal[0]; normal2[1] = normal[1]; normal2[2] = normal[2]; }
    else { normal2[0] = -normal[0]; normal2[1] = -normal[1]; normal2[2] = -normal[2]; }
    mul1_331(nr, Rb, normal2);
    anr[0] = std::fabs(nr[0]); anr[1] = std::fabs(nr[1]); anr[2] = std::fabs(nr[2]);
    int lanr, a1, a2;
    if (anr[1] > anr[0]) {
        if (anr[1] > anr[2]) { a1 = 0; lanr = 1; a2 = 2; }
        else { a1 = 0; a2 = 1; lanr = 2; }
    } else {
        if (anr[0] > anr[2]) { lanr = 0; a1 = 1; a2 = 2; }
        else { a1 = 0; a2 = 1; lanr = 2; }
    }
    R center[3];
    if (nr[lanr] < 0) { for (int i = 0; i < 3; i++) center[i] = pb[i] - pa[i] + Sb[lanr] * Rb[i * 4 + lanr]; }
    else { for (int i = 0; i < 3; i++) center[i] = pb[i] - pa[i] - Sb[lanr] * Rb[i * 4 + lanr]; }
    int codeN, code1, code2;
    if (code <= 3) codeN = code - 1; else codeN = code - 4;
    if (codeN == 0) { code1 = 1; code2 = 2; }
    else if (codeN == 1) { code1 = 0; code2 = 2; }
    else { code1 = 0; code2 = 1; }
    R quad[8], c1, c2, m11, m12, m21, m22;
    c1 = dot14(center, Ra + code1);
    c2 = dot14(center, Ra + code2);
    m11 = dot44(Ra + code1, Rb + a1);
    m12 = dot44(Ra + code1, Rb + a2);
    m21 = dot44(Ra + code2, Rb + a1);
    m22 = dot44(Ra + code2, Rb + a2);
    {
        R k1 = m11 * Sb[a1], k2 = m21 * Sb[a1], k3 = m12 * Sb[a2], k4 = m22 * Sb[a2];
        quad[0] = c1 - k1 - k3; quad[1] = c2 - k2 - k4;
        quad[2] = c1 - k1 + k3; quad[3] = c2 - k2 + k4;
        quad[4] = c1 + k1 + k3; quad[5] = c2 + k2 + k4;
        quad[6] = c1 + k1 - k3; quad[7] = c2 + k2 - k4;
    }
    R rect[2] = {Sa[code1], Sa[code2]};
    R ret[16];
    int n = intersectRectQuad(rect, quad, ret);
    if (n < 1) return 0;
    R point[3 * 8], dep[8];
    R det1 = recip(m11 * m22 - m12 * m21);
    m11 *= det1; m12 *= det1; m21 *= det1; m22 *= det1;
    int cnum = 0;
    for (int j = 0; j < n; j++) {
        R k1 = m22 * (ret[j * 2] - c1) - m12 * (ret[j * 2 + 1] - c2);
        R k2 = -m21 * (ret[j * 2] - c1) + m11 * (ret[j * 2 + 1] - c2);
        for (int i = 0; i < 3; i++) point[cnum * 3 + i] = center[i] + k1 * Rb[i * 4 + a1] + k2 * Rb[i * 4 + a2];
        dep[cnum] = Sa[codeN] - dot3(normal2, point + cnum * 3);
        if (dep[cnum] >= 0) {
            ret[cnum * 2] = ret[j * 2];
            ret[cnum * 2 + 1] = ret[j * 2 + 1];
            cnum++;
        }
    }
    if (cnum < 1) return 0;
    int maxc = maxc_flags & 0xffff;
    if (maxc > cnum) maxc = cnum;
    if (maxc < 1) maxc = 1;
    if (cnum <= maxc) {
        for (int j = 0; j < cnum; j++) {
            for (int i = 0; i < 3; i++) contact[j].pos[i] = point[j * 3 + i] + pa[i];
            contact[j].depth = dep[j];
        }
    } else {
        int i1 = 0;
        R maxdepth = dep[0];
        for (int i = 1; i < cnum; i++) if (dep[i] > maxdepth) { maxdepth = dep[i]; i1 = i; }
        int iret[8];
        cullPoints(cnum, ret, maxc, i1, iret);
        for (int j = 0; j < maxc; j++) {
            for (int i = 0; i < 3; i++) contact[j].pos[i] = point[iret[j] * 3 + i] + pa[i];
            contact[j].depth = dep[iret[j]];
        }
        cnum = maxc;
    }
    *return_code = code;
    return cnum;
}

/* dCollideBoxBox box.cpp:741-767 */
template <class R> int collide_box_box(const GeomView<R> &b1, const GeomView<R> &b2, int flags, Contact<R> *c) {
    R normal[3], depth;
    int code;
    int num = boxBox(b1.pose.pos, b1.pose.Rm, b1.p, b2.pose.pos, b2.pose.Rm, b2.p, normal, &depth, &code, flags, c);
    for (int i = 0; i < num; i++) for (int k = 0; k < 3; k++) c[i].normal[k] = -normal[k];
    return num;
}

/* dCollideBoxPlane box.cpp:770-903 */
template <class R> int collide_box_plane(const GeomView<R> &bx, const GeomView<R> &pl, int flags, Contact<R> *c) {
    const R *Rm = bx.pose.Rm, *n = pl.p, *side = bx.p;
    R Q1 = dot14(n, Rm + 0), Q2 = dot14(n, Rm + 1), Q3 = dot14(n, Rm + 2);
    R A1 = side[0] * Q1, A2 = side[1] * Q2, A3 = side[2] * Q3;
    R B1 = std::fabs(A1), B2 = std::fabs(A2), B3 = std::fabs(A3);
    R depth = pl.p[3] + R(0.5) * (B1 + B2 + B3) - dot3(n, bx.pose.pos);
    if (depth < 0) return 0;
    int maxc = flags & 0xffff;
    if (maxc > 4) maxc = 4;
    R p[3] = {bx.pose.pos[0], bx.pose.pos[1], bx.pose.pos[2]};
    const R Aarr[3] = {A1, A2, A3}, Barr[3] = {B1, B2, B3};
    for (int i = 0; i < 3; i++) {
        R sgn = (Aarr[i] > 0) ? R(-1) : R(1);
        for (int k = 0; k < 3; k++) {
            if (sgn < 0) p[k] -= R(0.5) * side[i] * Rm[k * 4 + i];
            else p[k] += R(0.5) * side[i] * Rm[k * 4 + i];
        }
    }
    for (int k = 0; k < 3; k++) c[0].pos[k] = p[k];
    c[0].depth = depth;
    int ret = 1;
    /* second and third contact points go along the two sides with the smallest
     * projected length; the goto tree of box.cpp:836-872 reduces to this
     * side order. */
    int order[2];
    if (B1 < B2) {
        if (B3 < B1) { order[0] = 2; order[1] = (B1 < B2) ? 0 : 1; }
        else { order[0] = 0; order[1] = (B2 < B3) ? 1 : 2; }
    } else {
        if (B3 < B2) { order[0] = 2; order[1] = (B1 < B2) ? 0 : 1; }
        else { order[0] = 1; order[1] = (B1 < B3) ? 0 : 2; }
    }
    if (maxc > 1) {
        for (int t = 0; t < 2; t++) {
            int sd = order[t];
            if (depth - Barr[sd] < 0) break;
            for (int k = 0; k < 3; k++) {
                if (Aarr[sd] > 0) c[ret].pos[k] = p[k] + side[sd] * Rm[k * 4 + sd];
                else c[ret].pos[k] = p[k] - side[sd] * Rm[k * 4 + sd];
            }
            c[ret].depth = depth - Barr[sd];
            ret++;
            if (maxc == 2) break;
        }
    }
    if (maxc == 4 && ret == 3) {
        R d4 = c[1].depth + c[2].depth - depth;
        if (d4 > 0) {
            for (int k = 0; k < 3; k++) c[3].pos[k] = c[1].pos[k] + c[2].pos[k] - p[k];
            c[3].depth = d4;
            ret++;
        }
    }
    for (int i = 0; i < ret; i++) for (int k = 0; k < 3; k++) c[i].normal[k] = n[k];
    return ret;
}

/* dCollideCapsuleSphere capsule.cpp:130-161 */
template <class R> int collide_capsule_sphere(const GeomView<R> &cp, const GeomView<R> &sp, Contact<R> *c) {
    const R *Rm = cp.pose.Rm, *p1 = cp.pose.pos, *p2 = sp.pose.pos;
    R alpha = Rm[2] * (p2[0] - p1[0]) + Rm[6] * (p2[1] - p1[1]) + Rm[10] * (p2[2] - p1[2]);
    R lz2 = cp.p[1] * R(0.5);
    if (alpha > lz2) alpha = lz2;
    if (alpha < -lz2) alpha = -lz2;
    R p[3] = {p1[0] + alpha * Rm[2], p1[1] + alpha * Rm[6], p1[2] + alpha * Rm[10]};
    return collide_spheres(p, cp.p[0], p2, sp.p[0], c);
}

/* dClosestLineSegmentPoints collision_util.cpp:109-223 */
template <class R> void closestLineSegmentPoints(const R *a1, const R *a2, const R *b1, const R *b2, R *cp1, R *cp2) {
    R a1a2[3], b1b2[3], a1b1[3], a1b2[3], a2b1[3], a2b2[3], n[3];
    R la, lb, k, da1, da2, da3, da4, db1, db2, db3, db4, det;
    auto set2 = [](R *a, const R *b) { a[0] = b[0]; a[1] = b[1]; a[2] = b[2]; };
    auto sub = [](R *a, const R *b, const R *c) { a[0] = b[0] - c[0]; a[1] = b[1] - c[1]; a[2] = b[2] - c[2]; };
    sub(a1a2, a2, a1); sub(b1b2, b2, b1); sub(a1b1, b1, a1);
    da1 = dot3(a1a2, a1b1); db1 = dot3(b1b2, a1b1);
    if (da1 <= 0 && db1 >= 0) { set2(cp1, a1); set2(cp2, b1); return; }
    sub(a1b2, b2, a1);
    da2 = dot3(a1a2, a1b2); db2 = dot3(b1b2, a1b2);
    if (da2 <= 0 && db2 <= 0) { set2(cp1, a1); set2(cp2, b2); return; }
    sub(a2b1, b1, a2);
    da3 = dot3(a1a2, a2b1); db3 = dot3(b1b2, a2b1);
    if (da3 >= 0 && db3 >= 0) { set2(cp1, a2); set2(cp2, b1); return; }
    sub(a2b2, b2, a2);
    da4 = dot3(a1a2, a2b2); db4 = dot3(b1b2, a2b2);
    if (da4 >= 0 && db4 <= 0) { set2(cp1, a2); set2(cp2, b2); return; }
    la = dot3(a1a2, a1a2);
    if (da1 >= 0 && da3 <= 0) {
        k = da1 / la;
        for (int i = 0; i < 3; i++) n[i] = a1b1[i] - k * a1a2[i];
        if (dot3(b1b2, n) >= 0) { for (int i = 0; i < 3; i++) cp1[i] = a1[i] + k * a1a2[i]; set2(cp2, b1); return; }
    }
    if (da2 >= 0 && da4 <= 0) {
        k = da2 / la;
        for (int i = 0; i < 3; i++) n[i] = a1b2[i] - k * a1a2[i];
        if (dot3(b1b2, n) <= 0) { for (int i = 0; i < 3; i++) cp1[i] = a1[i] + k * a1a2[i]; set2(cp2, b2); return; }
    }
    lb = dot3(b1b2, b1b2);
    if (db1 <= 0 && db2 >= 0) {
        k = -db1 / lb;
        for (int i = 0; i < 3; i++) n[i] = -a1b1[i] - k * b1b2[i];
        if (dot3(a1a2, n) >= 0) { set2(cp1, a1); for (int i = 0; i < 3; i++) cp2[i] = b1[i] + k * b1b2[i]; return; }
    }
    if (db3 <= 0 && db4 >= 0) {
        k = -db3 / lb;
        for (int i = 0; i < 3; i++) n[i] = -a2b1[i] - k * b1b2[i];
        if (dot3(a1a2, n) <= 0) { set2(cp1, a2); for (int i = 0; i < 3; i++) cp2[i] = b1[i] + k * b1b2[i]; return; }
    }
    k = dot3(a1a2, b1b2);
    det = la * lb - k * k;
    if (det <= 0) { set2(cp1, a1); set2(cp2, b1); return; }
    det = recip(det);
    R alpha = (lb * da1 - k * db1) * det;
    R beta = (k * da1 - la * db1) * det;
    for (int i = 0; i < 3; i++) { cp1[i] = a1[i] + alpha * a1a2[i]; cp2[i] = b1[i] + beta * b1b2[i]; }
}

/* dClosestLineBoxPoints collision_util.cpp:247-384 */
template <class R> void closestLineBoxPoints(const R *p1, const R *p2, const R *c, const R *Rm, const R *side, R *lret, R *bret) {
    R tmp[3], s[3], v[3], sign[3], v2[3], h[3];
    for (int i = 0; i < 3; i++) tmp[i] = p1[i] - c[i];
    mul1_331(s, Rm, tmp);
    for (int i = 0; i < 3; i++) tmp[i] = p2[i] - p1[i];
    mul1_331(v, Rm, tmp);
    for (int i = 0; i < 3; i++) {
        if (v[i] < 0) { s[i] = -s[i]; v[i] = -v[i]; sign[i] = -1; }
        else sign[i] = 1;
    }
    for (int i = 0; i < 3; i++) { v2[i] = v[i] * v[i]; h[i] = R(0.5) * side[i]; }
    int region[3];
    R tanchor[3];
    const R tanchor_eps = sizeof(R) == 4 ? R(1e-19) : R(1e-307);
    for (int i = 0; i < 3; i++) {
        if (v[i] > tanchor_eps) {
            if (s[i] < -h[i]) { region[i] = -1; tanchor[i] = (-h[i] - s[i]) / v[i]; }
            else { region[i] = (s[i] > h[i]); tanchor[i] = (h[i] - s[i]) / v[i]; }
        } else { region[i] = 0; tanchor[i] = 2; }
    }
    R t = 0, dd2dt = 0;
    for (int i = 0; i < 3; i++) dd2dt -= (region[i] ? v2[i] : 0) * tanchor[i];
    if (dd2dt >= 0) goto got_answer;
    do {
        R next_t = 1;
        for (int i = 0; i < 3; i++) if (tanchor[i] > t && tanchor[i] < 1 && tanchor[i] < next_t) next_t = tanchor[i];
        R next_dd2dt = 0;
        for (int i = 0; i < 3; i++) next_dd2dt += (region[i] ? v2[i] : 0) * (next_t - tanchor[i]);
        if (next_dd2dt >= 0) {
            R m = (next_dd2dt - dd2dt) / (next_t - t);
            t -= dd2dt / m;
            goto got_answer;
        }
        for (int i = 0; i < 3; i++) if (tanchor[i] == next_t) { tanchor[i] = (h[i] - s[i]) / v[i]; region[i]++; }
        t = next_t;
        dd2dt = next_dd2dt;
    } while (t < 1);
    t = 1;
got_answer:
    for (int i = 0; i < 3; i++) lret[i] = p1[i] + t * tmp[i];
    for (int i = 0; i < 3; i++) {
        tmp[i] = sign[i] * (s[i] + t * v[i]);
        if (tmp[i] < -h[i]) tmp[i] = -h[i];
        else if (tmp[i] > h[i]) tmp[i] = h[i];
    }
    mul0_331(s, Rm, tmp);
    for (int i = 0; i < 3; i++) bret[i] = s[i] + c[i];
}

/* dCollideCapsuleBox capsule.cpp:164-235 */
template <class R> int collide_capsule_box(const GeomView<R> &cp, const GeomView<R> &bx, int flags, Contact<R> *c) {
    const R *Rc = cp.pose.Rm, *pc = cp.pose.pos;
    R clen = cp.p[1] * R(0.5), radius = cp.p[0];
    R p1[3] = {pc[0] + clen * Rc[2], pc[1] + clen * Rc[6], pc[2] + clen * Rc[10]};
    R p2[3] = {pc[0] - clen * Rc[2], pc[1] - clen * Rc[6], pc[2] - clen * Rc[10]};
    R pl[3], pb[3];
    closestLineBoxPoints(p1, p2, bx.pose.pos, bx.pose.Rm, bx.p, pl, pb);
    R mindist = sizeof(R) == 4 ? R(1e-6) : R(1e-15);
    R dx = pl[0] - pb[0], dy = pl[1] - pb[1], dz = pl[2] - pb[2];
    if (std::sqrt(dx * dx + dy * dy + dz * dz) < mindist) {
        R normal[3], depth;
        int code;
        R r2 = radius * R(2);
        R capboxside[3] = {r2, r2, cp.p[1] + r2};
        int num = boxBox(bx.pose.pos, bx.pose.Rm, bx.p, pc, Rc, capboxside, normal, &depth, &code, flags, c);
        for (int i = 0; i < num; i++) for (int k = 0; k < 3; k++) c[i].normal[k] = normal[k];
        return num;
    }
    return collide_spheres(pl, radius, pb, R(0), c);
}

/* dCollideCapsuleCapsule capsule.cpp:238-353 */
template <class R> int collide_capsule_capsule(const GeomView<R> &c1g, const GeomView<R> &c2g, int flags, Contact<R> *c) {
    const R tolerance = R(1e-5);
    R lz1 = c1g.p[1] * R(0.5), lz2 = c2g.p[1] * R(0.5);
    const R *pos1 = c1g.pose.pos, *pos2 = c2g.pose.pos;
    R axis1[3] = {c1g.pose.Rm[2], c1g.pose.Rm[6], c1g.pose.Rm[10]};
    R axis2[3] = {c2g.pose.Rm[2], c2g.pose.Rm[6], c2g.pose.Rm[10]};
    R sphere1[3], sphere2[3];
    R a1a2 = dot3(axis1, axis2);
    R det = R(1) - a1a2 * a1a2;
    if (det < tolerance) {
        if (a1a2 < 0) { axis2[0] = -axis2[0]; axis2[1] = -axis2[1]; axis2[2] = -axis2[2]; }
        R q[3];
        for (int i = 0; i < 3; i++) q[i] = pos1[i] - pos2[i];
        R k = dot3(axis1, q);
        R a1lo = -lz1, a1hi = lz1, a2lo = -lz2 - k, a2hi = lz2 - k;
        R lo = (a1lo > a2lo) ? a1lo : a2lo;
        R hi = (a1hi < a2hi) ? a1hi : a2hi;
        if (lo <= hi) {
            int num_contacts = flags & 0xffff;
            if (num_contacts >= 2 && lo < hi) {
                for (int i = 0; i < 3; i++) sphere1[i] = pos1[i] + lo * axis1[i];
                for (int i = 0; i < 3; i++) sphere2[i] = pos2[i] + (lo + k) * axis2[i];
                int n1 = collide_spheres(sphere1, c1g.p[0], sphere2, c2g.p[0], c);
                if (n1) {
                    for (int i = 0; i < 3; i++) sphere1[i] = pos1[i] + hi * axis1[i];
                    for (int i = 0; i < 3; i++) sphere2[i] = pos2[i] + (hi + k) * axis2[i];
                    int n2 = collide_spheres(sphere1, c1g.p[0], sphere2, c2g.p[0], c + 1);
                    if (n2) return 2;
                }
            }
            R alpha1 = (lo + hi) * R(0.5);
            R alpha2 = alpha1 + k;
            for (int i = 0; i < 3; i++) sphere1[i] = pos1[i] + alpha1 * axis1[i];
            for (int i = 0; i < 3; i++) sphere2[i] = pos2[i] + alpha2 * axis2[i];
            return collide_spheres(sphere1, c1g.p[0], sphere2, c2g.p[0], c);
        }
    }
    R a1[3], a2[3], b1[3], b2[3];
    for (int i = 0; i < 3; i++) {
        a1[i] = pos1[i] + axis1[i] * lz1; a2[i] = pos1[i] - axis1[i] * lz1;
        b1[i] = pos2[i] + axis2[i] * lz2; b2[i] = pos2[i] - axis2[i] * lz2;
    }
    closestLineSegmentPoints(a1, a2, b1, b2, sphere1, sphere2);
    return collide_spheres(sphere1, c1g.p[0], sphere2, c2g.p[0], c);
}

/* dCollideCapsulePlane capsule.cpp:356-416 */
template <class R> int collide_capsule_plane(const GeomView<R> &cp, const GeomView<R> &pl, int flags, Contact<R> *c) {
    const R *Rm = cp.pose.Rm, *pos = cp.pose.pos, *n = pl.p;
    R lz = cp.p[1], r = cp.p[0];
    R sign = (dot14(n, Rm + 2) > 0) ? R(-1) : R(1);
    R p[3];
    for (int i = 0; i < 3; i++) p[i] = pos[i] + Rm[i * 4 + 2] * lz * R(0.5) * sign;
    R k = dot3(p, n);
    R depth = n[3] - k + r;
    if (depth < 0) return 0;
    for (int i = 0; i < 3; i++) { c[0].normal[i] = n[i]; c[0].pos[i] = p[i] - n[i] * r; }
    c[0].depth = depth;
    int nc = 1;
    if ((flags & 0xffff) >= 2) {
        for (int i = 0; i < 3; i++) p[i] = pos[i] - Rm[i * 4 + 2] * lz * R(0.5) * sign;
        k = dot3(p, n);
        depth = n[3] - k + r;
        if (depth >= 0) {
            for (int i = 0; i < 3; i++) { c[1].normal[i] = n[i]; c[1].pos[i] = p[i] - n[i] * r; }
            c[1].depth = depth;
            nc = 2;
        }
    }
    return nc;
}

/* dCollideCylinderPlane collision_cylinder_plane.cpp:42-250.  A disc parallel
 * to the plane (axis within the tolerance of the normal): the nearer disc's
 * centre +-V1, +-V2 (two radius vectors built by the reference's cross
 * products), each kept while depth > 0.  Otherwise the rim point of each disc
 * deepest along -n, kept while depth >= 0. */
template <class R> int collide_cylinder_plane(const GeomView<R> &cy, const GeomView<R> &pl, int flags, Contact<R> *c) {
    const R tol = sizeof(R) == 4 ? R(0.0001) : R(0.0000001);
    const int maxc = flags & 0xffff;
    const R radius = cy.p[0], length = cy.p[1];
    const R *cp = cy.pose.pos, *Rm = cy.pose.Rm, *pv = pl.p;
    const R dir[3] = {Rm[2], Rm[6], Rm[10]};
    R s = length * R(0.5);
    R e1[3], e2[3];
    for (int i = 0; i < 3; i++) {
        e2[i] = dir[i] * s + cp[i];
        e1[i] = dir[i] * -s + cp[i];
    }
    auto dot = [](const R *a, const R *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
    auto cross = [](R *r, const R *a, const R *b) {
        const R r0 = a[1] * b[2] - a[2] * b[1], r1 = a[2] * b[0] - a[0] * b[2], r2 = a[0] * b[1] - a[1] * b[0];
        r[0] = r0; r[1] = r1; r[2] = r2;
    };
    int n = 0;
    s = dir[0] * pv[0] + dir[1] * pv[1] + dir[2] * pv[2];
    if (s < 0) s += R(1); else s -= R(1);
    if (s < tol && s > -tol) {
        const R s1 = pv[3] - dot(pv, e1), t1 = pv[3] - dot(pv, e2);
        const R *P;
        if (s1 >= t1) {
            if (!(s1 >= 0)) return 0;
            P = e1;
        } else {
            if (!(t1 >= 0)) return 0;
            P = e2;
        }
        R V1[3], V2[3];
        if (dir[0] < tol && dir[0] > -tol) { V1[0] = dir[0] + R(1); V1[1] = dir[1]; V1[2] = dir[2]; }
        else { V1[0] = dir[0]; V1[1] = dir[1] + R(1); V1[2] = dir[2]; }
        cross(V2, V1, dir);
        R t = std::sqrt(V2[0] * V2[0] + V2[1] * V2[1] + V2[2] * V2[2]);
        t = radius / t;
        for (int i = 0; i < 3; i++) V2[i] *= t;
        cross(V1, V2, dir);
        const R *off[4] = {V1, V1, V2, V2};
        const R sgn[4] = {R(1), R(-1), R(1), R(-1)};
        for (int k = 0; k < 4; k++) {
            R q[3];
            for (int i = 0; i < 3; i++) q[i] = sgn[k] > 0 ? P[i] + off[k][i] : P[i] - off[k][i];
            const R depth = pv[3] - dot(pv, q);
            if (depth > 0) {
                for (int i = 0; i < 3; i++) { c[n].pos[i] = q[i]; c[n].normal[i] = pv[i]; }
                c[n].depth = depth;
                if (++n >= maxc) return n;
            }
        }
    } else {
        const R t = dot(pv, dir);
        R C[3] = {dir[0] * t - pv[0], dir[1] * t - pv[1], dir[2] * t - pv[2]};
        R l = std::sqrt(C[0] * C[0] + C[1] * C[1] + C[2] * C[2]);
        l = radius / l;
        for (int i = 0; i < 3; i++) C[i] *= l;
        R q[3] = {C[0] + e1[0], C[1] + e1[1], C[2] + e1[2]};
        R depth = pv[3] - dot(pv, q);
        if (depth >= 0) {
            for (int i = 0; i < 3; i++) { c[n].pos[i] = q[i]; c[n].normal[i] = pv[i]; }
            c[n].depth = depth;
            if (++n >= maxc) return n;
        }
        for (int i = 0; i < 3; i++) q[i] = C[i] + e2[i];
        depth = pv[3] - pv[0] * q[0] - pv[1] * q[1] - pv[2] * q[2];
        if (depth >= 0) {
            for (int i = 0; i < 3; i++) { c[n].pos[i] = q[i]; c[n].normal[i] = pv[i]; }
            c[n].depth = depth;
            ++n;
        }
    }
    return n;
}

/* Collider for (cls1, cls2) without reversal; returns -1 when no function exists. */
template <class R> int collide_fwd(const GeomView<R> &a, const GeomView<R> &b, int flags, Contact<R> *c) {
    if (a.cls == CLS_SPHERE && b.cls == CLS_SPHERE) return collide_spheres(a.pose.pos, a.p[0], b.pose.pos, b.p[0], c);
    if (a.cls == CLS_SPHERE && b.cls == CLS_BOX) return collide_sphere_box(a, b, c);
    if (a.cls == CLS_SPHERE && b.cls == CLS_PLANE) return collide_sphere_plane(a, b, c);
    if (a.cls == CLS_BOX && b.cls == CLS_BOX) return collide_box_box(a, b, flags, c);
    if (a.cls == CLS_BOX && b.cls == CLS_PLANE) return collide_box_plane(a, b, flags, c);
    if (a.cls == CLS_CAPSULE && b.cls == CLS_SPHERE) return collide_capsule_sphere(a, b, c);
    if (a.cls == CLS_CAPSULE && b.cls == CLS_BOX) return collide_capsule_box(a, b, flags, c);
    if (a.cls == CLS_CAPSULE && b.cls == CLS_CAPSULE) return collide_capsule_capsule(a, b, flags, c);
    if (a.cls == CLS_CAPSULE && b.cls == CLS_PLANE) return collide_capsule_plane(a, b, flags, c);
    if (a.cls == CLS_CYLINDER && b.cls == CLS_PLANE) return collide_cylinder_plane(a, b, flags, c);
    return -1;
}

/* dCollide collision_kernel.cpp:292-329 with the colliders table of :166-268:
 * the reverse entry swaps the geoms and negates the normals. */
template <class R> int dcollide(const GeomView<R> &g1, int i1, const GeomView<R> &g2, int i2, int flags, Contact<R> *c) {
    if ((flags & 0xffff) == 0) return 0;
    if (i1 == i2) return 0;
    if (g1.body == g2.body && g1.body >= 0) return 0;
    int n = collide_fwd(g1, g2, flags, c);
    if (n >= 0) {
        for (int i = 0; i < n; i++) { c[i].g1 = i1; c[i].g2 = i2; c[i].side1 = -1; c[i].side2 = -1; }
        return n;
    }
    n = collide_fwd(g2, g1, flags, c);
    if (n < 0) return 0;
    for (int i = 0; i < n; i++) {
        for (int k = 0; k < 3; k++) c[i].normal[k] = -c[i].normal[k];
        c[i].g1 = i1; c[i].g2 = i2; c[i].side1 = -1; c[i].side2 = -1;
    }
    return n;
}

/* ---------------------------------------------------------- broadphase -- */
inline bool cat_ok(const unsigned *cat, const unsigned *col, int a, int b) {
    unsigned ca = cat ? cat[a] : ~0u, cb = cat ? cat[b] : ~0u;
    unsigned la = col ? col[a] : ~0u, lb = col ? col[b] : ~0u;
    return ((ca & lb) || (cb & la));
}

/* dxSAPSpace::collide collision_sapspace.cpp:467-535 and BoxPruning :560-621
 * (axis order XYZ).  The radix sort of :654-894 is a stable ascending sort of
 * the float keys with -0.0 placed before +0.0 (sign-magnitude order of the
 * IEEE bits); std::stable_sort on that key order reproduces it. */
template <class R> int pairs_sap(int ng, const R *aabb, const int *body, const unsigned *cat, const unsigned *col, int *pairs, int cap) {
    std::vector<int> fin, inf;
    for (int i = 0; i < ng; i++) {
        if (aabb[i * 6 + 1] == (R)kInf) inf.push_back(i); else fin.push_back(i);
    }
    int np = 0;
    auto emit = [&](int a, int b) {
        /* collideGeomsNoAABBs collision_sapspace.cpp:273-297 */
        if (body[a] == body[b] && body[a] >= 0) return;
        if (!cat_ok(cat, col, a, b)) return;
        if (np < cap) { pairs[2 * np] = a; pairs[2 * np + 1] = b; }
        np++;
    };
    int n = (int)fin.size();
    std::vector<float> pos(n);
    for (int i = 0; i < n; i++) pos[i] = (float)aabb[fin[i] * 6 + 0];
    std::vector<int> sorted(n);
    for (int i = 0; i < n; i++) sorted[i] = i;
    auto key = [](float f) {
        uint32_t u; std::memcpy(&u, &f, 4);
        /* map IEEE bits to an order-preserving unsigned key: negatives reversed */
        return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
    };
    std::stable_sort(sorted.begin(), sorted.end(), [&](int a, int b) { return key(pos[a]) < key(pos[b]); });
    for (int si = 0; si + 1 < n; si++) {
        int id0 = sorted[si];
        const R *a0 = aabb + fin[id0] * 6;
        float ax0max = (float)a0[1];
        for (int sj = si + 1; sj < n; sj++) {
            int id1 = sorted[sj];
            if (!(pos[id1] <= ax0max)) break;
            const R *a1 = aabb + fin[id1] * 6;
            if (a0[3] >= a1[2] && a1[3] >= a0[2] && a0[5] >= a1[4] && a1[5] >= a0[4]) emit(fin[id0], fin[id1]);
        }
    }
    int ni = (int)inf.size();
    for (int m = 0; m < ni; m++) {
        for (int k = m + 1; k < ni; k++) emit(inf[m], inf[k]);
        for (int k = 0; k < n; k++) emit(inf[m], fin[k]);
    }
    return np;
}

/* dxHashSpace::collide collision_space.cpp:421-614.  Every pair it reports
 * passes collideAABBs (collision_space_internal.h:44-78); the multi-level cell
 * walk only prunes, because a closed AABB overlap implies an overlap of the
 * floor-discretised cell ranges at the coarser of the two levels.  The set is
 * therefore the closed-AABB-overlap set, restated here as an all-pairs loop
 * (canonical order: lower geom index first). */
template <class R> int pairs_hash(int ng, const R *aabb, const int *body, const unsigned *cat, const unsigned *col, int, int, int *pairs, int cap) {
    int np = 0;
    for (int a = 0; a < ng; a++) {
        for (int b = a + 1; b < ng; b++) {
            if (body[a] == body[b] && body[a] >= 0) continue;
            if (!cat_ok(cat, col, a, b)) continue;
            const R *x = aabb + a * 6, *y = aabb + b * 6;
            if (x[0] > y[1] || x[1] < y[0] || x[2] > y[3] || x[3] < y[2] || x[4] > y[5] || x[5] < y[4]) continue;
            if (np < cap) { pairs[2 * np] = a; pairs[2 * np + 1] = b; }
            np++;
        }
    }
    return np;
}

/* ------------------------------------------------------------- joints --- */
enum { JT_CONTACT = 0, JT_BALL = 1, JT_HINGE = 2, JT_SLIDER = 3, JT_FIXED = 4 };

/* Extended joint record (32 reals) of hinge, slider and fixed joints: 0-2
 * axis1 (body1 frame), 3-5 axis2 (body2 frame, or world), 6-9 qrel, 10-12
 * offset, 14 the limit state the joint carries from its last getInfo1 (the
 * hinge keeps it when its stops are not finite), 16 lostop, 17 histop, 18 vel,
 * 19 fmax, 20 fudge factor, 21 normal cfm, 22 stop erp, 23 stop cfm,
 * 24 bounce, 26 joint erp, 27 joint cfm (fixed joint).  dxJointLimitMotor
 * joints/joint.h, dxJointHinge/Slider/Fixed joints/hinge.h, slider.h, fixed.h. */
constexpr int ORC_JEXT_STRIDE = 32;

template <class R> struct Surface {
    int mode;
    R mu, mu2, bounce, bounce_vel, soft_erp, soft_cfm, motion1, motion2, motionN, slip1, slip2;
};

template <class R> struct Joint {
    int type, b1, b2;   /* node[0].body, node[1].body after dJointAttach (ode.cpp:1383-1439) */
    bool reverse;
    R pos[3], normal[3], depth, fdir1[3];
    Surface<R> s;
    R anchor1[3], anchor2[3], erp, cfm;
    /* hinge / slider / fixed */
    R axis1[3], axis2[3], qrel[4], offset[3];
    int limit;
    R limit_err, lostop, histop, vel, fmax, fudge, normal_cfm, stop_erp, stop_cfm, bounce, jerp, jcfm;
};

struct BodyState;

/* dxJointContact::getInfo1 contact.cpp:48-123 (rolling friction not on this path) */
template <class R> int contact_m(Surface<R> &s) {
    int m = 1;
    if (s.mode & 0x001) {
        if (s.mu < 0) s.mu = 0; else if (s.mu > 0) m++;
        if (s.mu2 < 0) s.mu2 = 0; else if (s.mu2 > 0) m++;
    } else {
        if (s.mu < 0) s.mu = 0; else if (s.mu > 0) m += 2;
    }
    return m;
}

template <class R> struct Row { R J1[6], rhs, cfm, J2[6], lo, hi; };

/* dxJointContact::getInfo2 contact.cpp:125-345.  b1v/b2v: BODY records of
 * node[0]/node[1] (b2v may be NULL).  Rows are pre-initialised by the caller. */
template <class R> void contact_info2(const Joint<R> &j, int the_m, const R *b0v, const R *b1v, R fps, R worldERP,
                                      R min_depth, R max_vel, Row<R> *rows, int *findex) {
    const int mode = j.s.mode;
    R erp = (mode & 0x008) ? j.s.soft_erp : worldERP;
    R k = fps * erp;
    R depth = j.depth - min_depth;
    if (depth < 0) depth = 0;
    R motionN = (mode & 0x080) ? j.s.motionN : R(0);
    const R pushout = k * depth + motionN;
    bool apply_bounce = (mode & 0x004) && j.s.bounce_vel >= 0;
    R outgoing = 0;
    R c = pushout > max_vel ? max_vel : pushout;
    R c1[3], c2[3] = {0, 0, 0}, normal[3];
    for (int i = 0; i < 3; i++) normal[i] = j.reverse ? -j.normal[i] : j.normal[i];
    if (b1v) {
        for (int i = 0; i < 3; i++) c2[i] = j.pos[i] - b1v[i];
        for (int i = 0; i < 3; i++) rows[0].J2[i] = -normal[i];
        cross3(rows[0].J2 + 3, normal, c2);
        if (apply_bounce) outgoing = dot3(rows[0].J2 + 3, b1v + 12) - dot3(normal, b1v + 8);
    }
    for (int i = 0; i < 3; i++) c1[i] = j.pos[i] - b0v[i];
    for (int i = 0; i < 3; i++) rows[0].J1[i] = normal[i];
    cross3(rows[0].J1 + 3, c1, normal);
    if (apply_bounce) {
        outgoing += dot3(rows[0].J1 + 3, b0v + 12) + dot3(normal, b0v + 8);
        R neg = motionN - outgoing;
        if (neg > j.s.bounce_vel) {
            const R newc = j.s.bounce * neg + motionN;
            if (newc > c) c = newc;
        }
    }
    rows[0].rhs = c;
    if (mode & 0x010) rows[0].cfm = j.s.soft_cfm;
    rows[0].lo = 0;
    rows[0].hi = (R)kInf;
    if (the_m > 1) {
        R t1[3], t2[3];
        if (mode & 0x002) { for (int i = 0; i < 3; i++) t1[i] = j.fdir1[i]; cross3(t2, normal, t1); }
        else planeSpace(normal, t1, t2);
        int row = 1;
        const R mu = j.s.mu;
        if (mu > 0) {
            for (int i = 0; i < 3; i++) rows[row].J1[i] = t1[i];
            cross3(rows[row].J1 + 3, c1, t1);
            if (b1v) { for (int i = 0; i < 3; i++) rows[row].J2[i] = -t1[i]; cross3(rows[row].J2 + 3, t1, c2); }
            if (mode & 0x020) rows[row].rhs = j.s.motion1;
            if (mode & 0x100) rows[row].cfm = j.s.slip1;
            rows[row].lo = -mu; rows[row].hi = mu;
            if (mode & 0x1000) findex[row] = 0;
            row++;
        }
        const R mu2 = (mode & 0x001) ? j.s.mu2 : mu;
        if (mu2 > 0) {
            for (int i = 0; i < 3; i++) rows[row].J1[i] = t2[i];
            cross3(rows[row].J1 + 3, c1, t2);
            if (b1v) { for (int i = 0; i < 3; i++) rows[row].J2[i] = -t2[i]; cross3(rows[row].J2 + 3, t2, c2); }
            if (mode & 0x040) rows[row].rhs = j.s.motion2;
            if (mode & 0x200) rows[row].cfm = j.s.slip2;
            rows[row].lo = -mu2; rows[row].hi = mu2;
            if (mode & 0x2000) findex[row] = 0;
            row++;
        }
    }
}

/* dxJointBall::getInfo2 ball.cpp:58-67 -> setBall joints/joint.cpp:111-162 */
template <class R> void ball_info2(const Joint<R> &j, const R *b0v, const R *b1v, R fps, Row<R> *rows) {
    R R0[12], a1[3], a2[3];
    for (int r = 0; r < 3; r++) rows[r].cfm = j.cfm;
    RfromQ(R0, b0v + 4);
    rows[0].J1[0] = 1; rows[1].J1[1] = 1; rows[2].J1[2] = 1;
    mul0_331(a1, R0, j.anchor1);
    /* dSetCrossMatrixMinus(J1 + angular, a1): rows of -[a1]x */
    rows[0].J1[4] = a1[2]; rows[0].J1[5] = -a1[1];
    rows[1].J1[3] = -a1[2]; rows[1].J1[5] = a1[0];
    rows[2].J1[3] = a1[1]; rows[2].J1[4] = -a1[0];
    if (b1v) {
        R R1[12];
        RfromQ(R1, b1v + 4);
        rows[0].J2[0] = -1; rows[1].J2[1] = -1; rows[2].J2[2] = -1;
        mul0_331(a2, R1, j.anchor2);
        /* dSetCrossMatrixPlus(J2 + angular, a2) */
        rows[0].J2[4] = -a2[2]; rows[0].J2[5] = a2[1];
        rows[1].J2[3] = a2[2]; rows[1].J2[5] = -a2[0];
        rows[2].J2[3] = -a2[1]; rows[2].J2[4] = a2[0];
    }
    R k = fps * j.erp;
    for (int r = 0; r < 3; r++) {
        if (b1v) rows[r].rhs = k * (a2[r] + b1v[r] - a1[r] - b0v[r]);
        else rows[r].rhs = k * (j.anchor2[r] - a1[r] - b0v[r]);
    }
}

/* Quaternion products of rotation.cpp:41-107: qa = qb * qc with qb and / or
 * qc conjugated (dQMultiply1: qb', dQMultiply2: qc', dQMultiply3: both). */
template <class R> void qmul(R *qa, const R *qb, const R *qc, bool cb, bool cc) {
    const R b[4] = {qb[0], cb ? -qb[1] : qb[1], cb ? -qb[2] : qb[2], cb ? -qb[3] : qb[3]};
    const R c[4] = {qc[0], cc ? -qc[1] : qc[1], cc ? -qc[2] : qc[2], cc ? -qc[3] : qc[3]};
    qa[0] = b[0] * c[0] - b[1] * c[1] - b[2] * c[2] - b[3] * c[3];
    qa[1] = b[0] * c[1] + b[1] * c[0] + b[2] * c[3] - b[3] * c[2];
    qa[2] = b[0] * c[2] + b[2] * c[0] + b[3] * c[1] - b[1] * c[3];
    qa[3] = b[0] * c[3] + b[3] * c[0] + b[1] * c[2] - b[2] * c[1];
}

/* getHingeAngle joints/joint.cpp:421-490: the relative rotation angle about
 * `axis` (body1 frame), in -pi..pi, zero at the orientation of qrel. */
template <class R> R hinge_angle(const R *q1, const R *q2, const R *axis, const R *qrel) {
    R qr[4];
    if (q2) {
        R qq[4];
        qmul(qq, q1, q2, true, false);
        qmul(qr, qq, qrel, false, true);
    } else {
        qmul(qr, q1, qrel, true, true);
    }
    const R cost2 = qr[0];
    const R sint2 = std::sqrt(qr[1] * qr[1] + qr[2] * qr[2] + qr[3] * qr[3]);
    R theta = dot3(qr + 1, axis) >= 0 ? 2 * std::atan2(sint2, cost2) : 2 * std::atan2(sint2, -cost2);
    if (theta > (R)M_PI) theta -= (R)(2 * M_PI);
    return -theta;
}

/* dJointGetSliderPosition joints/slider.cpp:45-82 */
template <class R> R slider_position(const Joint<R> &j, const R *b0v, const R *b1v) {
    R R0[12], ax1[3], q[3];
    RfromQ(R0, b0v + 4);
    mul0_331(ax1, R0, j.axis1);
    if (b1v) {
        R R1[12];
        RfromQ(R1, b1v + 4);
        mul0_331(q, R1, j.offset);
        for (int i = 0; i < 3; i++) q[i] = b0v[i] - q[i] - b1v[i];
    } else {
        for (int i = 0; i < 3; i++) q[i] = b0v[i] - j.offset[i];
        if (j.reverse) for (int i = 0; i < 3; i++) ax1[i] = -ax1[i];
    }
    return dot3(ax1, q);
}

/* getInfo1 of hinge (hinge.cpp:58-77), slider (slider.cpp:112-143), fixed
 * (fixed.cpp:61-65): rows, and the limit state of the limit / motor. */
template <class R> int joint_m(Joint<R> &j, const R *b0v, const R *b1v) {
    if (j.type == JT_FIXED) return 6;
    if (j.type == JT_HINGE) {
        int m = j.fmax > 0 ? 6 : 5;
        if ((j.lostop >= -(R)M_PI || j.histop <= (R)M_PI) && j.lostop <= j.histop) {
            const R a = hinge_angle(b0v + 4, b1v ? b1v + 4 : (const R *)nullptr, j.axis1, j.qrel);
            /* testRotationalLimit joint.cpp:574-595 */
            if (a <= j.lostop) { j.limit = 1; j.limit_err = a - j.lostop; m = 6; }
            else if (a >= j.histop) { j.limit = 2; j.limit_err = a - j.histop; m = 6; }
            else j.limit = 0;
        }
        if (j.limit) m = 6;
        return m;
    }
    if (j.type == JT_SLIDER) {
        int m = j.fmax > 0 ? 6 : 5;
        j.limit = 0;
        if ((j.lostop > -kInf || j.histop < kInf) && j.lostop <= j.histop) {
            const R pos = slider_position(j, b0v, b1v);
            if (pos <= j.lostop) { j.limit = 1; j.limit_err = pos - j.lostop; m = 6; }
            else if (pos >= j.histop) { j.limit = 2; j.limit_err = pos - j.histop; m = 6; }
        }
        return m;
    }
    return 3;
}

/* setFixedOrientation joints/joint.cpp:228-285: three rows making the
 * relative rotation equal qrel. */
template <class R> void fixed_orientation(const Joint<R> &j, const R *b0v, const R *b1v, R fps, R erp, Row<R> *rows) {
    rows[0].J1[3] = 1; rows[1].J1[4] = 1; rows[2].J1[5] = 1;
    if (b1v) { rows[0].J2[3] = -1; rows[1].J2[4] = -1; rows[2].J2[5] = -1; }
    R qerr[4];
    if (b1v) {
        R qq[4];
        qmul(qq, b0v + 4, b1v + 4, true, false);
        qmul(qerr, qq, j.qrel, false, true);
    } else {
        qmul(qerr, b0v + 4, j.qrel, true, true);
    }
    if (qerr[0] < 0) { qerr[1] = -qerr[1]; qerr[2] = -qerr[2]; qerr[3] = -qerr[3]; }
    R R0[12], e[3];
    RfromQ(R0, b0v + 4);
    mul0_331(e, R0, qerr + 1);
    const R k2 = fps * erp * 2;
    for (int r = 0; r < 3; r++) rows[r].rhs = k2 * e[r];
}

/* dxJointLimitMotor::addLimot joints/joint.cpp:598-786: the limit / motor row
 * along ax1 (rotational or linear), and the force of a motor powered at its
 * limit, added to the bodies' accumulators as the reference does. */
template <class R> void limot_row(const Joint<R> &j, R *b0v, R *b1v, R fps, const R *ax1, bool rotational, Row<R> &row) {
    int powered = j.fmax > 0;
    if (!powered && !j.limit) return;
    R *J1u = rotational ? row.J1 + 3 : row.J1, *J2u = rotational ? row.J2 + 3 : row.J2;
    for (int k = 0; k < 3; k++) J1u[k] = ax1[k];
    if (b1v) for (int k = 0; k < 3; k++) J2u[k] = -ax1[k];
    R ltd[3] = {0, 0, 0};
    if (!rotational && b1v) {
        R c[3];
        for (int k = 0; k < 3; k++) c[k] = (R)0.5 * (b1v[k] - b0v[k]);
        cross3(ltd, c, ax1);
        for (int k = 0; k < 3; k++) { row.J1[3 + k] = ltd[k]; row.J2[3 + k] = ltd[k]; }
    }
    if (j.limit && j.lostop == j.histop) powered = 0;
    if (powered) {
        row.cfm = j.normal_cfm;
        if (!j.limit) {
            row.rhs = j.vel;
            row.lo = -j.fmax;
            row.hi = j.fmax;
        } else {
            R fm = j.fmax;
            if (j.vel > 0 || (j.vel == 0 && j.limit == 2)) fm = -fm;
            if ((j.limit == 1 && j.vel > 0) || (j.limit == 2 && j.vel < 0)) fm *= j.fudge;
            const R f[3] = {fm * ax1[0], fm * ax1[1], fm * ax1[2]};
            if (rotational) {
                if (b1v) for (int k = 0; k < 3; k++) b1v[20 + k] += f[k];
                for (int k = 0; k < 3; k++) b0v[20 + k] -= f[k];
            } else {
                if (b1v) {
                    for (int k = 0; k < 3; k++) { b0v[20 + k] += -fm * ltd[k]; b1v[20 + k] += -fm * ltd[k]; }
                    for (int k = 0; k < 3; k++) b1v[16 + k] += f[k];
                }
                for (int k = 0; k < 3; k++) b0v[16 + k] -= f[k];
            }
        }
    }
    if (j.limit) {
        const R k = fps * j.stop_erp;
        row.rhs = -k * j.limit_err;
        row.cfm = j.stop_cfm;
        if (j.lostop == j.histop) {
            row.lo = (R)-kInf;
            row.hi = (R)kInf;
        } else {
            if (j.limit == 1) { row.lo = 0; row.hi = (R)kInf; }
            else { row.lo = (R)-kInf; row.hi = 0; }
            if (j.bounce > 0) {
                R v;
                if (rotational) {
                    v = dot3(b0v + 12, ax1);
                    if (b1v) v -= dot3(b1v + 12, ax1);
                } else {
                    v = dot3(b0v + 8, ax1);
                    if (b1v) v -= dot3(b1v + 8, ax1);
                }
                if (j.limit == 1) {
                    if (v < 0) { const R nc = -j.bounce * v; if (nc > row.rhs) row.rhs = nc; }
                } else {
                    if (v > 0) { const R nc = -j.bounce * v; if (nc < row.rhs) row.rhs = nc; }
                }
            }
        }
    }
}

/* getInfo2 of a hinge joint, hinge.cpp:80-153 */
template <class R> void hinge_info2(const Joint<R> &j, R *b0v, R *b1v, R fps, R erp, Row<R> *rows) {
    Joint<R> jb = j;
    jb.erp = erp;
    ball_info2(jb, b0v, b1v, fps, rows);
    for (int r = 0; r < 3; r++) rows[r].cfm = j.cfm;
    R R0[12], ax1[3], p[3], q[3], b[3];
    RfromQ(R0, b0v + 4);
    mul0_331(ax1, R0, j.axis1);
    planeSpace(ax1, p, q);
    for (int k = 0; k < 3; k++) { rows[3].J1[3 + k] = p[k]; rows[4].J1[3 + k] = q[k]; }
    if (b1v) for (int k = 0; k < 3; k++) { rows[3].J2[3 + k] = -p[k]; rows[4].J2[3 + k] = -q[k]; }
    if (b1v) {
        R R1[12], ax2[3];
        RfromQ(R1, b1v + 4);
        mul0_331(ax2, R1, j.axis2);
        cross3(b, ax1, ax2);
    } else {
        cross3(b, ax1, j.axis2);
    }
    const R k = fps * erp;
    rows[3].rhs = k * dot3(b, p);
    rows[4].rhs = k * dot3(b, q);
    limot_row(j, b0v, b1v, fps, ax1, true, rows[5]);
}

/* getInfo2 of a slider joint, slider.cpp:146-250 */
template <class R> void slider_info2(const Joint<R> &j, R *b0v, R *b1v, R fps, R erp, Row<R> *rows) {
    fixed_orientation(j, b0v, b1v, fps, erp, rows);
    R R0[12], ax1[3], p[3], q[3], c[3] = {0, 0, 0};
    RfromQ(R0, b0v + 4);
    mul0_331(ax1, R0, j.axis1);
    planeSpace(ax1, p, q);
    if (b1v) for (int k = 0; k < 3; k++) c[k] = b1v[k] - b0v[k];
    for (int k = 0; k < 3; k++) { rows[3].J1[k] = p[k]; rows[4].J1[k] = q[k]; }
    if (b1v) {
        R t[3];
        for (int k = 0; k < 3; k++) { rows[3].J2[k] = -p[k]; rows[4].J2[k] = -q[k]; }
        cross3(t, c, p);
        for (int k = 0; k < 3; k++) { rows[3].J1[3 + k] = (R)0.5 * t[k]; rows[3].J2[3 + k] = rows[3].J1[3 + k]; }
        cross3(t, c, q);
        for (int k = 0; k < 3; k++) { rows[4].J1[3 + k] = (R)0.5 * t[k]; rows[4].J2[3 + k] = rows[4].J1[3 + k]; }
    }
    const R k = fps * erp;
    if (b1v) {
        R R1[12], ofs[3];
        RfromQ(R1, b1v + 4);
        mul0_331(ofs, R1, j.offset);
        for (int i = 0; i < 3; i++) c[i] += ofs[i];
        rows[3].rhs = k * dot3(p, c);
        rows[4].rhs = k * dot3(q, c);
    } else {
        R ofs[3];
        for (int i = 0; i < 3; i++) ofs[i] = j.offset[i] - b0v[i];
        rows[3].rhs = k * dot3(p, ofs);
        rows[4].rhs = k * dot3(q, ofs);
        if (j.reverse) for (int i = 0; i < 3; i++) ax1[i] = -ax1[i];
    }
    limot_row(j, b0v, b1v, fps, ax1, false, rows[5]);
}

/* getInfo2 of a fixed joint, fixed.cpp:68-105: the joint's own erp / cfm on
 * the three position rows, the world's erp on the orientation rows. */
template <class R> void fixed_info2(const Joint<R> &j, R *b0v, R *b1v, R fps, R erp, Row<R> *rows) {
    fixed_orientation(j, b0v, b1v, fps, erp, rows + 3);
    rows[0].J1[0] = 1; rows[1].J1[1] = 1; rows[2].J1[2] = 1;
    const R k = fps * j.jerp;
    R R0[12], ofs[3];
    RfromQ(R0, b0v + 4);
    mul0_331(ofs, R0, j.offset);
    if (b1v) {
        /* dSetCrossMatrixPlus(J1 + angular, ofs) */
        rows[0].J1[4] = -ofs[2]; rows[0].J1[5] = ofs[1];
        rows[1].J1[3] = ofs[2]; rows[1].J1[5] = -ofs[0];
        rows[2].J1[3] = -ofs[1]; rows[2].J1[4] = ofs[0];
        rows[0].J2[0] = -1; rows[1].J2[1] = -1; rows[2].J2[2] = -1;
        for (int r = 0; r < 3; r++) rows[r].rhs = k * (b1v[r] - b0v[r] + ofs[r]);
    } else {
        for (int r = 0; r < 3; r++) rows[r].rhs = k * (j.offset[r] - b0v[r]);
    }
    for (int r = 0; r < 3; r++) rows[r].cfm = j.jcfm;
}

/* ---------------------------------------------------------- quickstep --- */
template <class R> struct Params {
    R h, g[3], erp, cfm, w;
    int iters;
    R max_vel, min_depth;
    Surface<R> s;
    int maxc, space, skip_connected;
    double lambda_floor;
    double tol;
    int check_every;
    /* dynamic iteration count adjustment (objects.h:121-203, quickstep.cpp:1809-1860) */
    bool dyn;
    unsigned max_extra;
    R premature_delta, extra_delta;
    bool random_reorder;  /* REORDERING_METHOD__RANDOMLY, frequency 8 (quickstep.cpp:54-85) */
    bool joint_order;     /* rows swept in joint order, each joint's rows together (no ReorderPrep split) */
};

/* The near-callback policy as a table (what a program's callback decides
 * per pair before dJointCreateContact, e.g. ode/demo/demo_boxstack.cpp:
 * 131-160): each geom carries a material (GEOM slot 13) and category / collide
 * bits (slots 14, 15); a pair passes the bit test of collideAABBs
 * (collision_space_internal.h) when the filter is on, and its contacts take
 * the surface of the material pair, at most that entry's contact count.
 * Table rows (double[16]): 0 mode, 1 mu, 2 mu2, 3 bounce, 4 bounce_vel,
 * 5 soft_erp, 6 soft_cfm, 7 motion1, 8 motion2, 9 motionN, 10 slip1, 11 slip2,
 * 12 max contacts (0: the parameters' count). */
template <class R> struct SurfaceTable {
    int nmat = 0;
    bool filter = false;
    std::vector<Surface<R>> s;
    std::vector<int> maxc;
};

template <class R> SurfaceTable<R> read_table(int nmat, const double *t, int filter) {
    SurfaceTable<R> T;
    T.nmat = nmat;
    T.filter = filter != 0;
    for (int k = 0; k < nmat * nmat; k++) {
        const double *e = t + (size_t)k * 16;
        Surface<R> x;
        x.mode = (int)e[0];
        x.mu = (R)e[1]; x.mu2 = (R)e[2]; x.bounce = (R)e[3]; x.bounce_vel = (R)e[4];
        x.soft_erp = (R)e[5]; x.soft_cfm = (R)e[6];
        x.motion1 = (R)e[7]; x.motion2 = (R)e[8]; x.motionN = (R)e[9];
        x.slip1 = (R)e[10]; x.slip2 = (R)e[11];
        T.s.push_back(x);
        T.maxc.push_back((int)e[12]);
    }
    return T;
}

template <class R> inline int geom_material(const R *geoms, int g) { return (int)geoms[(size_t)g * ORC_GEOM_STRIDE + 13]; }

template <class R> Params<R> read_params(const double *P) {
    Params<R> p;
    p.h = (R)P[ORC_P_H];
    for (int i = 0; i < 3; i++) p.g[i] = (R)P[ORC_P_GX + i];
    p.erp = (R)P[ORC_P_ERP]; p.cfm = (R)P[ORC_P_CFM]; p.w = (R)P[ORC_P_SOR_W];
    p.iters = (int)P[ORC_P_ITERS];
    p.max_vel = (R)P[ORC_P_MAX_VEL]; p.min_depth = (R)P[ORC_P_MIN_DEPTH];
    p.s.mode = (int)P[ORC_P_MODE]; p.s.mu = (R)P[ORC_P_MU]; p.s.mu2 = (R)P[ORC_P_MU2];
    p.s.bounce = (R)P[ORC_P_BOUNCE]; p.s.bounce_vel = (R)P[ORC_P_BOUNCE_VEL];
    p.s.soft_erp = (R)P[ORC_P_SOFT_ERP]; p.s.soft_cfm = (R)P[ORC_P_SOFT_CFM];
    p.s.motion1 = (R)P[ORC_P_MOTION1]; p.s.motion2 = (R)P[ORC_P_MOTION2]; p.s.motionN = (R)P[ORC_P_MOTIONN];
    p.s.slip1 = (R)P[ORC_P_SLIP1]; p.s.slip2 = (R)P[ORC_P_SLIP2];
    p.maxc = (int)P[ORC_P_MAXC]; p.space = (int)P[ORC_P_SPACE];
    p.skip_connected = (int)P[ORC_P_SKIP_CONNECTED];
    p.lambda_floor = P[ORC_P_LAMBDA_FLOOR] > 0 ? P[ORC_P_LAMBDA_FLOOR] : 1e-3;
    p.tol = P[ORC_P_TOL];
    p.check_every = P[ORC_P_CHECK_EVERY] >= 1 ? (int)P[ORC_P_CHECK_EVERY] : 1;
    /* DeriveExtraIterationCount objects.h:177-184 and the enabled rule of
     * UpdateDynamicIterationCountAdjustmentEnabledState objects.h:186-189 */
    const double fac = P[ORC_P_DYN_EXTRA_FACTOR] > 0 ? P[ORC_P_DYN_EXTRA_FACTOR] : 0.0;
    const double xc = (double)p.iters * (double)(R)fac;
    p.max_extra = xc < 4294967295.0 ? (unsigned)xc : 4294967295u;
    p.premature_delta = (R)(P[ORC_P_DYN_PREMATURE] > 0 ? P[ORC_P_DYN_PREMATURE] : 0.0);
    p.extra_delta = (R)(P[ORC_P_DYN_EXTRA_DELTA] > 0 ? P[ORC_P_DYN_EXTRA_DELTA] : 0.0);
    p.dyn = p.max_extra != 0 || p.premature_delta != 0;
    p.random_reorder = P[ORC_P_RANDOM_REORDER] == 1;
    p.joint_order = P[ORC_P_RANDOM_REORDER] == 2;
    return p;
}

/* Assembled system of one QuickStep (Stages 0-2 + iMJ + Ad). */
template <class R> struct System {
    int nb, m;
    std::vector<R> invI;            /* nb x 12 world-frame inverse inertia */
    std::vector<Row<R>> J;          /* m rows (Ad-scaled after scale_ad) */
    std::vector<Row<R>> Jraw;       /* unscaled copy */
    std::vector<int> findex, jb1, jb2;
    std::vector<R> iMJ;             /* m x 12 */
    std::vector<R> Ad;
    std::vector<int> order;
};

template <class R> R body_invmass(const R *b) { return recip(b[24]); }
template <class R> void body_I(const R *b, R *I) {
    I[0] = b[25]; I[5] = b[26]; I[10] = b[27];
    I[1] = I[4] = b[28]; I[2] = I[8] = b[29]; I[6] = I[9] = b[30];
    I[3] = I[7] = I[11] = 0;
}

/* Stage0_Bodies quickstep.cpp:1176-1306 (gravity into facc, world-frame invI,
 * implicit gyroscopic torque into tacc).  Modifies facc/tacc in place. */
template <class R> void stage0_bodies(int nb, R *bodies, const Params<R> &p, std::vector<R> &invI) {
    invI.assign((size_t)nb * 12, 0);
    for (int i = 0; i < nb; i++) {
        R *b = bodies + (size_t)i * ORC_BODY_STRIDE;
        int flags = (int)b[31];
        if (!(flags & 1)) {
            for (int k = 0; k < 3; k++) if (p.g[k] != 0) b[16 + k] += b[24] * p.g[k];
        }
    }
    for (int i = 0; i < nb; i++) {
        R *b = bodies + (size_t)i * ORC_BODY_STRIDE;
        int flags = (int)b[31];
        R Rm[12], Ib[12], invIb[12], tmp[12];
        RfromQ(Rm, b + 4);
        body_I(b, Ib);
        if (!invertPD3(invIb, Ib)) { std::memset(invIb, 0, sizeof invIb); invIb[0] = invIb[5] = invIb[10] = 1; }
        mul2_333(tmp, invIb, Rm);
        mul0_333(&invI[(size_t)i * 12], Rm, tmp);
        R invMass = body_invmass(b);
        if ((flags & 2) && invMass > 0) {
            R Iw[12], L[3];
            mul2_333(tmp, Ib, Rm);
            mul0_333(Iw, Rm, tmp);
            mul0_331(L, Iw, b + 12);
            R It[12] = {0};
            It[1] = L[2]; It[2] = -L[1]; It[4] = -L[2]; It[6] = L[0]; It[8] = L[1]; It[9] = -L[0];
            for (int k = 0; k < 12; k++) It[k] = It[k] * p.h + Iw[k];
            R hr = recip(p.h);
            for (int k = 0; k < 3; k++) L[k] *= hr;
            R itInv[12];
            if (invert3(itInv, It) != 0) {
                mul0_333(It, Iw, itInv);
                It[0] -= 1; It[5] -= 1; It[10] -= 1;
                R tau0[3];
                mul0_331(tau0, It, L);
                for (int k = 0; k < 3; k++) b[20 + k] += tau0[k];
            }
        }
    }
}

/* Stages 0_Joints..4LCP_ReorderPrep (quickstep.cpp:1319-2355) over the joint
 * list in the given order. */
template <class R> void assemble(int nb, R *bodies, std::vector<Joint<R>> &joints, const Params<R> &p, System<R> &S,
                                 const std::vector<R> &invI) {
    S.nb = nb;
    S.invI = invI;
    std::vector<int> jm(joints.size());
    int m = 0;
    for (size_t i = 0; i < joints.size(); i++) {
        Joint<R> &j = joints[i];
        const R *b0v = bodies + (size_t)j.b1 * ORC_BODY_STRIDE;
        const R *b1v = j.b2 >= 0 ? bodies + (size_t)j.b2 * ORC_BODY_STRIDE : nullptr;
        jm[i] = j.type == JT_CONTACT ? contact_m(j.s) : joint_m(j, b0v, b1v);
        m += jm[i];
    }
    S.m = m;
    S.J.assign(m, Row<R>());
    S.findex.assign(m, -1);
    S.jb1.assign(m, -1);
    S.jb2.assign(m, -1);
    const R fps = recip(p.h);
    int ofs = 0;
    for (size_t ji = 0; ji < joints.size(); ji++) {
        Joint<R> &j = joints[ji];
        int im = jm[ji];
        for (int r = 0; r < im; r++) {
            Row<R> &row = S.J[ofs + r];
            std::memset(&row, 0, sizeof row);
            row.cfm = p.cfm;
            row.lo = (R)-kInf;
            row.hi = (R)kInf;
        }
        R *b0v = bodies + (size_t)j.b1 * ORC_BODY_STRIDE;
        R *b1v = j.b2 >= 0 ? bodies + (size_t)j.b2 * ORC_BODY_STRIDE : nullptr;
        int fidx[6] = {-1, -1, -1, -1, -1, -1};
        if (j.type == JT_CONTACT) contact_info2(j, im, b0v, b1v, fps, p.erp, p.min_depth, p.max_vel, &S.J[ofs], fidx);
        else if (j.type == JT_HINGE) hinge_info2(j, b0v, b1v, fps, p.erp, &S.J[ofs]);
        else if (j.type == JT_SLIDER) slider_info2(j, b0v, b1v, fps, p.erp, &S.J[ofs]);
        else if (j.type == JT_FIXED) fixed_info2(j, b0v, b1v, fps, p.erp, &S.J[ofs]);
        else ball_info2(j, b0v, b1v, fps, &S.J[ofs]);
        for (int r = 0; r < im; r++) {
            if (fidx[r] != -1) S.findex[ofs + r] = fidx[r] + ofs;
            S.J[ofs + r].rhs *= fps;
            S.J[ofs + r].cfm *= fps;
            S.jb1[ofs + r] = j.b1;
            S.jb2[ofs + r] = j.b2;
        }
        ofs += im;
    }
    /* Stage2b :1644-1690 and Stage2c :1727-1747 */
    std::vector<R> rhs_tmp((size_t)nb * 6);
    for (int i = 0; i < nb; i++) {
        const R *b = bodies + (size_t)i * ORC_BODY_STRIDE;
        R im = body_invmass(b);
        R *rt = &rhs_tmp[(size_t)i * 6];
        for (int k = 0; k < 3; k++) rt[k] = -(b[16 + k] * im + b[8 + k] * fps);
        mul0_331(rt + 3, &invI[(size_t)i * 12], b + 20);
        for (int k = 0; k < 3; k++) rt[3 + k] = -(b[12 + k] * fps) - rt[3 + k];
    }
    for (int i = 0; i < m; i++) {
        Row<R> &row = S.J[i];
        R sum = 0;
        const R *in = &rhs_tmp[(size_t)S.jb1[i] * 6];
        for (int k = 0; k < 6; k++) sum += row.J1[k] * in[k];
        if (S.jb2[i] != -1) {
            in = &rhs_tmp[(size_t)S.jb2[i] * 6];
            for (int k = 0; k < 6; k++) sum += row.J2[k] * in[k];
        }
        row.rhs += sum;
    }
    S.Jraw = S.J;
    /* compute_invM_JT :859-897 */
    S.iMJ.assign((size_t)m * 12, 0);
    for (int i = 0; i < m; i++) {
        R *o = &S.iMJ[(size_t)i * 12];
        const Row<R> &row = S.J[i];
        int b1 = S.jb1[i], b2 = S.jb2[i];
        R k1 = body_invmass(bodies + (size_t)b1 * ORC_BODY_STRIDE);
        for (int k = 0; k < 3; k++) o[k] = k1 * row.J1[k];
        mul0_331(o + 3, &invI[(size_t)b1 * 12], row.J1 + 3);
        if (b2 != -1) {
            R k2 = body_invmass(bodies + (size_t)b2 * ORC_BODY_STRIDE);
            for (int k = 0; k < 3; k++) o[6 + k] = k2 * row.J2[k];
            mul0_331(o + 9, &invI[(size_t)b2 * 12], row.J2 + 3);
        }
    }
    /* Stage4LCP_AdComputation :2251-2316 */
    S.Ad.assign(m, 0);
    for (int i = 0; i < m; i++) {
        Row<R> &row = S.J[i];
        const R *o = &S.iMJ[(size_t)i * 12];
        R sum = 0;
        for (int k = 0; k < 6; k++) sum += o[k] * row.J1[k];
        if (S.jb2[i] != -1) for (int k = 0; k < 6; k++) sum += o[6 + k] * row.J2[k];
        R Ad = p.w / (sum + row.cfm);
        S.Ad[i] = Ad;
        row.cfm = row.cfm * Ad;
        row.rhs *= Ad;
        for (int k = 0; k < 6; k++) row.J1[k] *= Ad;
        if (S.jb2[i] != -1) for (int k = 0; k < 6; k++) row.J2[k] *= Ad;
    }
    /* Stage4LCP_ReorderPrep :2329-2355 */
    S.order.clear();
    if (p.joint_order) {
        for (int i = 0; i < m; i++) S.order.push_back(i);
    } else {
        for (int i = 0; i < m; i++) if (S.findex[i] == -1) S.order.push_back(i);
        for (int i = 0; i < m; i++) if (S.findex[i] != -1) S.order.push_back(i);
    }
}

/* dxQuickStepIsland_Stage4LCP_IterationStep :2917-3033.  fa, when given,
 * receives the force adjustments of the dynamic iteration count: per body,
 * delta * max|iMJ side| summed separately for positive and negative delta
 * (:3001-3031; the maxabs of compute_invM_JT :859-897). */
template <class R> void sor_row(const System<R> &S, int index, std::vector<R> &lambda, std::vector<R> &fc,
                                std::vector<R> *fa = nullptr) {
    const Row<R> &row = S.J[index];
    R old = lambda[index];
    R delta = row.rhs - old * row.cfm;
    R *f1 = &fc[(size_t)S.jb1[index] * 6];
    R *f2 = S.jb2[index] != -1 ? &fc[(size_t)S.jb2[index] * 6] : nullptr;
    delta -= f1[0] * row.J1[0] + f1[1] * row.J1[1] + f1[2] * row.J1[2] + f1[3] * row.J1[3] + f1[4] * row.J1[4] + f1[5] * row.J1[5];
    if (f2) delta -= f2[0] * row.J2[0] + f2[1] * row.J2[1] + f2[2] * row.J2[2] + f2[3] * row.J2[3] + f2[4] * row.J2[4] + f2[5] * row.J2[5];
    R hi_act, lo_act;
    if (S.findex[index] != -1) { hi_act = std::fabs(row.hi * lambda[S.findex[index]]); lo_act = -hi_act; }
    else { hi_act = row.hi; lo_act = row.lo; }
    R nl = old + delta;
    if (nl < lo_act) { delta = lo_act - old; lambda[index] = lo_act; }
    else if (nl > hi_act) { delta = hi_act - old; lambda[index] = hi_act; }
    else lambda[index] = nl;
    if (delta != 0) {
        const R *o = &S.iMJ[(size_t)index * 12];
        for (int k = 0; k < 6; k++) f1[k] += delta * o[k];
        if (f2) for (int k = 0; k < 6; k++) f2[k] += delta * o[6 + k];
        if (fa) {
            const int sgn = delta > 0 ? 1 : 0;  /* [negative, positive] */
            R m1 = std::fabs(o[0]);
            for (int k = 1; k < 6; k++) m1 = std::max(m1, (R)std::fabs(o[k]));
            (*fa)[(size_t)S.jb1[index] * 2 + sgn] += delta * m1;
            if (f2) {
                R m2 = std::fabs(o[6]);
                for (int k = 7; k < 12; k++) m2 = std::max(m2, (R)std::fabs(o[k]));
                (*fa)[(size_t)S.jb2[index] * 2 + sgn] += delta * m2;
            }
        }
    }
}

/* CheckForMaximumToBeLessThanLimitAndResetMaxAdjustments :3253-3285 over the
 * bodies of one island: true when every body's positive and negative
 * adjustment is strictly below limit (never for limit 0); all reset to 0. */
template <class R> bool adjustments_below(std::vector<R> &fa, const std::vector<int> &bodies, R limit) {
    bool ok = limit != 0;
    for (int b : bodies) {
        if (ok && (!(fa[(size_t)b * 2 + 1] < limit) || !(-fa[(size_t)b * 2] < limit))) ok = false;
        fa[(size_t)b * 2] = 0;
        fa[(size_t)b * 2 + 1] = 0;
    }
    return ok;
}

/* Islands of the assembled rows: connected components of the bodies through
 * the rows, numbered by first appearance in row order (the island builder,
 * util.cpp:724-855, lists each island's joints contiguously).  Returns the
 * island of every row; bodies[i] lists island i's bodies. */
template <class R> std::vector<int> row_islands(const System<R> &S, std::vector<std::vector<int>> &bodies) {
    std::vector<int> parent(S.nb);
    for (int i = 0; i < S.nb; i++) parent[i] = i;
    std::function<int(int)> find = [&](int x) { return parent[x] == x ? x : parent[x] = find(parent[x]); };
    for (int r = 0; r < S.m; r++)
        if (S.jb2[r] >= 0) parent[find(S.jb1[r])] = find(S.jb2[r]);
    std::vector<int> id(S.nb, -1), isl(S.m);
    bodies.clear();
    for (int r = 0; r < S.m; r++) {
        const int root = find(S.jb1[r]);
        if (id[root] < 0) { id[root] = (int)bodies.size(); bodies.emplace_back(); }
        isl[r] = id[root];
    }
    for (int b = 0; b < S.nb; b++) {
        const int root = find(b);
        if (id[root] >= 0) bodies[id[root]].push_back(b);
    }
    return isl;
}

/* LCP residual of the assembled system at (lambda, fc = M^-1 J^T lambda):
 * the un-relaxed projected correction per row, relative to max|lambda|. */
template <class R> double residual(const System<R> &S, const std::vector<R> &lambda, const std::vector<R> &fc, R w, double floor_) {
    double mx = 0, lmax = 0;
    for (int i = 0; i < S.m; i++) {
        const Row<R> &row = S.J[i];
        const R *f1 = &fc[(size_t)S.jb1[i] * 6];
        const R *f2 = S.jb2[i] != -1 ? &fc[(size_t)S.jb2[i] * 6] : nullptr;
        double d = (double)row.rhs - (double)lambda[i] * row.cfm;
        for (int k = 0; k < 6; k++) d -= (double)f1[k] * row.J1[k];
        if (f2) for (int k = 0; k < 6; k++) d -= (double)f2[k] * row.J2[k];
        d /= (double)w;
        double lo, hi;
        if (S.findex[i] != -1) { hi = std::fabs((double)row.hi * lambda[S.findex[i]]); lo = -hi; }
        else { hi = row.hi; lo = row.lo; }
        double nl = lambda[i] + d;
        if (nl < lo) nl = lo;
        if (nl > hi) nl = hi;
        mx = std::max(mx, std::fabs(nl - (double)lambda[i]));
        lmax = std::max(lmax, std::fabs((double)lambda[i]));
    }
    return mx / std::max(lmax, floor_);
}

/* Stages 4b, 6a, 6b: quickstep.cpp:3082-3102, 3299-3337, 3409-3439 and
 * dxStepBody util.cpp:583-699 (infinitesimal rotation).  Body flags bit 4
 * caps the angular speed at slot 23 before the move (util.cpp:585-594); bits
 * 2 and 3 scale the linear / angular velocity by 1 - slot 3 / slot 11 after
 * the move when its squared speed exceeds the square of slot 15 / slot 19
 * (util.cpp:675-698). */
template <class R> void integrate(int nb, R *bodies, const std::vector<R> &invI, const std::vector<R> *fc, R h) {
    for (int i = 0; i < nb; i++) {
        R *b = bodies + (size_t)i * ORC_BODY_STRIDE;
        if (fc) {
            const R *f = &(*fc)[(size_t)i * 6];
            for (int k = 0; k < 3; k++) { b[8 + k] += h * f[k]; b[12 + k] += h * f[3 + k]; }
        }
    }
    for (int i = 0; i < nb; i++) {
        R *b = bodies + (size_t)i * ORC_BODY_STRIDE;
        R ims = h * body_invmass(b);
        for (int k = 0; k < 3; k++) { b[8 + k] += ims * b[16 + k]; b[20 + k] *= h; }
        const R *iI = &invI[(size_t)i * 12];
        R t[3];
        mul0_331(t, iI, b + 20);
        for (int k = 0; k < 3; k++) b[12 + k] += t[k];
    }
    for (int i = 0; i < nb; i++) {
        R *b = bodies + (size_t)i * ORC_BODY_STRIDE;
        const int flags = (int)b[31];
        if (flags & 16) {
            const R cap = b[23], a2 = b[12] * b[12] + b[13] * b[13] + b[14] * b[14];
            if (a2 > cap * cap) {
                const R c = cap / std::sqrt(a2);
                for (int k = 0; k < 3; k++) b[12 + k] *= c;
            }
        }
        for (int k = 0; k < 3; k++) b[k] += h * b[8 + k];
        R dq[4];
        DQfromW(dq, b + 12, b + 4);
        for (int k = 0; k < 4; k++) b[4 + k] += h * dq[k];
        normalize4(b + 4);
        if ((flags & 4) && b[8] * b[8] + b[9] * b[9] + b[10] * b[10] > b[15] * b[15])
            for (int k = 0; k < 3; k++) b[8 + k] *= 1 - b[3];
        if ((flags & 8) && b[12] * b[12] + b[13] * b[13] + b[14] * b[14] > b[19] * b[19])
            for (int k = 0; k < 3; k++) b[12 + k] *= 1 - b[11];
        for (int k = 0; k < 3; k++) { b[16 + k] = 0; b[20 + k] = 0; }
    }
}

template <class R> void read_joint(const R *rec, const Params<R> &p, Joint<R> &j) {
    j.type = (int)rec[0];
    int b1 = (int)rec[1], b2 = (int)rec[2];
    /* dJointAttach: a missing body1 moves body2 into slot 1 and sets dJOINT_REVERSE */
    j.reverse = false;
    if (b1 < 0) { b1 = b2; b2 = -1; j.reverse = true; }
    j.b1 = b1; j.b2 = b2;
    if (j.type == JT_CONTACT) {
        for (int i = 0; i < 3; i++) { j.pos[i] = rec[3 + i]; j.normal[i] = rec[6 + i]; j.fdir1[i] = rec[10 + i]; }
        j.depth = rec[9];
        j.s = p.s;
    } else {
        for (int i = 0; i < 3; i++) { j.anchor1[i] = rec[3 + i]; j.anchor2[i] = rec[6 + i]; }
        j.erp = rec[9];
        j.cfm = rec[10];
    }
}

template <class R> void read_joint_ext(const R *x, Joint<R> &j) {
    for (int i = 0; i < 3; i++) { j.axis1[i] = x[i]; j.axis2[i] = x[3 + i]; j.offset[i] = x[10 + i]; }
    for (int i = 0; i < 4; i++) j.qrel[i] = x[6 + i];
    j.limit = (int)x[14];
    j.limit_err = 0;
    j.lostop = x[16]; j.histop = x[17]; j.vel = x[18]; j.fmax = x[19]; j.fudge = x[20];
    j.normal_cfm = x[21]; j.stop_erp = x[22]; j.stop_cfm = x[23]; j.bounce = x[24];
    j.jerp = x[26]; j.jcfm = x[27];
}

/* dxQuickStepIsland quickstep.cpp:1112 (single-threaded path, deterministic
 * order: REORDERING_METHOD__DONT_REORDER, dynamic iteration count off). */
template <class R> int quickstep(int nb, R *bodies, std::vector<Joint<R>> &joints, const Params<R> &p, R *lambda_out,
                                 int lambda_cap, System<R> *Sout = nullptr) {
    std::vector<R> invI;
    stage0_bodies(nb, bodies, p, invI);
    System<R> S;
    assemble(nb, bodies, joints, p, S, invI);
    std::vector<R> lambda(S.m, 0), fc((size_t)nb * 6, 0);
    int sweeps = 0;
    if (S.m > 0 && (p.dyn || p.random_reorder) && !(p.tol > 0)) {
        /* the iteration loop of dxQuickStepIsland (quickstep.cpp:1820-1853),
         * island by island (dxProcessIslands util.cpp:886-938) with its own
         * count: random reordering of the island's whole order at iterations
         * 8, 16, ... (IsSORConstraintsReorderRequiredForIteration :1080-1109,
         * ConstraintsShuffling :2589-2611: for index 1..m-1 swap with
         * dRandInt(index + 1)); with the dynamic count, a premature exit once
         * every body's adjustment is below the premature delta, and after
         * `iters` up to max_extra more while some adjustment is at least
         * extra_delta */
        std::vector<std::vector<int>> ib;
        const std::vector<int> isl = row_islands(S, ib);
        std::vector<R> fa((size_t)nb * 2, 0);
        for (size_t i = 0; i < ib.size(); i++) {
            std::vector<int> rows;
            for (int k = 0; k < S.m; k++)
                if (isl[S.order[k]] == (int)i) rows.push_back(S.order[k]);
            R limit = p.premature_delta;
            unsigned iteration = 0, extra = 0;
            for (;;) {
                if (p.random_reorder && iteration >= 8 && iteration % 8 == 0) {
                    for (size_t index = 1; index < rows.size(); ++index) {
                        const int swap_index = orc_drand_int(&g_rand_seed, (int)index + 1);
                        std::swap(rows[index], rows[swap_index]);
                    }
                }
                for (int r : rows) sor_row(S, r, lambda, fc, p.dyn ? &fa : nullptr);
                ++iteration;
                if (!p.dyn) {
                    if (iteration == (unsigned)p.iters) break;
                    continue;
                }
                if (iteration - extra == (unsigned)p.iters) {
                    if (extra != 0 || p.max_extra == 0) break;
                    extra = p.max_extra;
                    limit = p.extra_delta;
                }
                if (adjustments_below(fa, ib[i], limit)) break;
            }
            sweeps = std::max(sweeps, (int)iteration);
        }
    } else if (S.m > 0) {
        for (int it = 0; it < p.iters; it++) {
            for (int k = 0; k < S.m; k++) sor_row(S, S.order[k], lambda, fc);
            sweeps = it + 1;
            /* optional stopping rule: LCP residual <= tol, checked every check_every sweeps */
            if (p.tol > 0 && (sweeps % p.check_every == 0 || sweeps == p.iters) &&
                residual(S, lambda, fc, p.w, p.lambda_floor) <= p.tol)
                break;
        }
    }
    g_last_sweeps = sweeps;
    integrate(nb, bodies, invI, S.m > 0 ? &fc : nullptr, p.h);
    if (lambda_out) for (int i = 0; i < S.m && i < lambda_cap; i++) lambda_out[i] = lambda[i];
    if (Sout) *Sout = S;
    return S.m;
}

/* tab: contact joints take the table's surface at their record's slot 13
 * (the material pair's entry, as the near-callback stored it). */
template <class R> int quickstep_api(int nb, R *bodies, int nj, const R *joints, const double *params, R *lambda, int cap,
                                     const R *jext = nullptr, const SurfaceTable<R> *tab = nullptr) {
    Params<R> p = read_params<R>(params);
    std::vector<Joint<R>> js(nj);
    for (int i = 0; i < nj; i++) {
        const R *rec = joints + (size_t)i * ORC_JOINT_STRIDE;
        read_joint(rec, p, js[i]);
        if (jext && js[i].type >= JT_HINGE) read_joint_ext(jext + (size_t)i * ORC_JEXT_STRIDE, js[i]);
        if (tab && tab->nmat && js[i].type == JT_CONTACT) {
            const int e = (int)rec[13];
            if (e < 0 || e >= tab->nmat * tab->nmat) return -1;
            js[i].s = tab->s[e];
        }
    }
    return quickstep(nb, bodies, js, p, lambda, cap);
}

/* Near-callback of the canonical frame: every candidate pair in geom-index
 * order, dCollide with max contacts, one contact joint per contact attached to
 * (body(g1), body(g2)) -- the loop of e.g. ode/demo/demo_space_stress.cpp:76-107. */
template <class R> int gen_contacts(int nb, const R *bodies, int ng, const R *geoms, int nball, const R *balls,
                                    const Params<R> &p, std::vector<Contact<R>> &out,
                                    const SurfaceTable<R> *tab = nullptr) {
    std::vector<GeomView<R>> gv(ng);
    std::vector<R> aabb((size_t)ng * 6);
    std::vector<int> body(ng);
    for (int i = 0; i < ng; i++) {
        gv[i] = geom_view(geoms + (size_t)i * ORC_GEOM_STRIDE, bodies);
        geom_aabb(gv[i], &aabb[(size_t)i * 6]);
        body[i] = gv[i].body;
    }
    int cap = ng * ng + 16;
    std::vector<int> pairs((size_t)cap * 2);
    std::vector<unsigned> cat, col;
    if (tab && tab->filter) {
        for (int i = 0; i < ng; i++) {
            cat.push_back((unsigned)geoms[(size_t)i * ORC_GEOM_STRIDE + 14]);
            col.push_back((unsigned)geoms[(size_t)i * ORC_GEOM_STRIDE + 15]);
        }
    }
    const unsigned *pc = cat.empty() ? nullptr : cat.data(), *pl = col.empty() ? nullptr : col.data();
    int np = (p.space == 1) ? pairs_hash(ng, aabb.data(), body.data(), pc, pl, -3, 10, pairs.data(), cap)
                            : pairs_sap(ng, aabb.data(), body.data(), pc, pl, pairs.data(), cap);
    std::vector<std::pair<int, int>> cp(np);
    for (int k = 0; k < np; k++) {
        int a = pairs[2 * k], b = pairs[2 * k + 1];
        cp[k] = {std::min(a, b), std::max(a, b)};
    }
    std::sort(cp.begin(), cp.end());
    out.clear();
    (void)nb;
    for (auto &pr : cp) {
        const GeomView<R> &a = gv[pr.first], &b = gv[pr.second];
        if (a.body < 0 && b.body < 0) continue;
        if (p.skip_connected && a.body >= 0 && b.body >= 0) {
            bool conn = false;
            for (int k = 0; k < nball; k++) {
                int x = (int)balls[(size_t)k * ORC_JOINT_STRIDE + 1], y = (int)balls[(size_t)k * ORC_JOINT_STRIDE + 2];
                if ((x == a.body && y == b.body) || (x == b.body && y == a.body)) conn = true;
            }
            if (conn) continue;
        }
        int maxc = p.maxc;
        if (tab && tab->nmat) {
            const int e = tab->maxc[geom_material(geoms, pr.first) * tab->nmat + geom_material(geoms, pr.second)];
            if (e > 0) maxc = e;
        }
        Contact<R> c[16];
        int n = dcollide(a, pr.first, b, pr.second, maxc, c);
        for (int i = 0; i < n; i++) out.push_back(c[i]);
    }
    return (int)out.size();
}

/* bext: extended records (ORC_JEXT_STRIDE each) of the hinge, slider and
 * fixed joints among `balls` (null when every one is a ball joint). */
template <class R> void build_joints(const R *geoms, int nball, const R *balls, const std::vector<Contact<R>> &cs,
                                     const Params<R> &p, std::vector<Joint<R>> &js,
                                     const SurfaceTable<R> *tab = nullptr, const R *bext = nullptr) {
    js.clear();
    for (int k = 0; k < nball; k++) {
        Joint<R> j;
        read_joint(balls + (size_t)k * ORC_JOINT_STRIDE, p, j);
        if (bext && j.type >= JT_HINGE) read_joint_ext(bext + (size_t)k * ORC_JEXT_STRIDE, j);
        js.push_back(j);
    }
    for (const Contact<R> &c : cs) {
        R rec[ORC_JOINT_STRIDE] = {0};
        rec[0] = 0;
        rec[1] = geoms[(size_t)c.g1 * ORC_GEOM_STRIDE + 1];
        rec[2] = geoms[(size_t)c.g2 * ORC_GEOM_STRIDE + 1];
        for (int i = 0; i < 3; i++) { rec[3 + i] = c.pos[i]; rec[6 + i] = c.normal[i]; }
        rec[9] = c.depth;
        Joint<R> j;
        read_joint(rec, p, j);
        if (tab && tab->nmat) j.s = tab->s[geom_material(geoms, c.g1) * tab->nmat + geom_material(geoms, c.g2)];
        js.push_back(j);
    }
}

template <class R> int world_step(int nb, R *bodies, int ng, const R *geoms, int nball, const R *balls, const double *P,
                                  R *contacts, int ccap, R *lambda, int lcap, int *m_out,
                                  const SurfaceTable<R> *tab = nullptr, const R *bext = nullptr) {
    Params<R> p = read_params<R>(P);
    std::vector<Contact<R>> cs;
    int nc = gen_contacts(nb, bodies, ng, geoms, nball, balls, p, cs, tab);
    for (int i = 0; i < nc && i < ccap; i++) {
        R *o = contacts + (size_t)i * ORC_CONTACT_STRIDE;
        for (int k = 0; k < 3; k++) { o[k] = cs[i].pos[k]; o[3 + k] = cs[i].normal[k]; }
        o[6] = cs[i].depth; o[7] = (R)cs[i].g1; o[8] = (R)cs[i].g2; o[9] = -1; o[10] = -1; o[11] = 0;
    }
    std::vector<Joint<R>> js;
    build_joints(geoms, nball, balls, cs, p, js, tab, bext);
    int m = quickstep(nb, bodies, js, p, lambda, lcap);
    if (m_out) *m_out = m;
    return nc;
}

/* Double-precision system for (bodies at step start, contacts). */
void build_system_f64(int nb, const double *bodies_in, int nball, const double *balls, int nc, const double *contacts,
                      const double *P, std::vector<double> &bodies, System<double> &S) {
    Params<double> p = read_params<double>(P);
    bodies.assign(bodies_in, bodies_in + (size_t)nb * ORC_BODY_STRIDE);
    std::vector<double> invI;
    stage0_bodies(nb, bodies.data(), p, invI);
    std::vector<Joint<double>> js;
    for (int k = 0; k < nball; k++) {
        Joint<double> j;
        read_joint(balls + (size_t)k * ORC_JOINT_STRIDE, p, j);
        js.push_back(j);
    }
    for (int i = 0; i < nc; i++) {
        const double *c = contacts + (size_t)i * ORC_JOINT_STRIDE;
        Joint<double> j;
        read_joint(c, p, j);
        js.push_back(j);
    }
    assemble(nb, bodies.data(), js, p, S, invI);
}

}  // namespace

/* =========================================================== exports ==== */
extern "C" {

unsigned int orc_drand(unsigned int *seed) {
    *seed = (unsigned int)((1664525u * (uint64_t)*seed + 1013904223u) & 0xffffffffu);
    return *seed;
}

void orc_rand_set_seed(unsigned int seed) { g_rand_seed = seed; }
unsigned int orc_rand_get_seed(void) { return g_rand_seed; }

int orc_drand_int(unsigned int *seed, int n) {
    uint32_t r = orc_drand(seed), un = (uint32_t)n;
    int result;
    if (un <= 0x10u) {
        r ^= (r >> 16); r ^= (r >> 8); r ^= (r >> 4);
        if (un <= 0x2u) { r ^= (r >> 2); r ^= (r >> 1); result = (int)(r & (un >> 1)); }
        else if (un <= 0x4u) { r ^= (r >> 2); result = (int)(((r & 0x03u) * un) >> 2); }
        else result = (int)(((r & 0x0Fu) * un) >> 4);
    } else if (un <= 0x100u) {
        r ^= (r >> 16); r ^= (r >> 8); result = (int)(((r & 0xFFu) * un) >> 8);
    } else if (un <= 0x10000u) {
        r ^= (r >> 16); result = (int)(((r & 0xFFFFu) * un) >> 16);
    } else {
        result = (int)(((uint64_t)r * un) >> 32);
    }
    return result;
}

#define ORC_COLLIDE(SUF, R)                                                                         \
    int orc_collide_##SUF(const R *g1, const R *g2, const R *b1, const R *b2, int flags, R *out) { \
        /* b1/b2: BODY records of the geoms' bodies (ignored for static geoms) */                   \
        R tmpb[2 * ORC_BODY_STRIDE];                                                                \
        R gg1[ORC_GEOM_STRIDE], gg2[ORC_GEOM_STRIDE];                                               \
        std::memcpy(gg1, g1, sizeof gg1); std::memcpy(gg2, g2, sizeof gg2);                         \
        if (gg1[1] >= 0) { std::memcpy(tmpb, b1, ORC_BODY_STRIDE * sizeof(R)); gg1[1] = 0; }        \
        if (gg2[1] >= 0) { std::memcpy(tmpb + ORC_BODY_STRIDE, b2, ORC_BODY_STRIDE * sizeof(R)); gg2[1] = 1; } \
        GeomView<R> a = geom_view(gg1, tmpb), b = geom_view(gg2, tmpb);                             \
        Contact<R> c[16];                                                                           \
        int n = dcollide(a, 0, b, 1, flags, c);                                                     \
        for (int i = 0; i < n; i++) {                                                               \
            R *o = out + (size_t)i * ORC_CONTACT_STRIDE;                                            \
            for (int k = 0; k < 3; k++) { o[k] = c[i].pos[k]; o[3 + k] = c[i].normal[k]; }        \
            o[6] = c[i].depth; o[7] = 0; o[8] = 1; o[9] = -1; o[10] = -1; o[11] = 0;               \
        }                                                                                           \
        return n;                                                                                   \
    }
ORC_COLLIDE(f32, float)
ORC_COLLIDE(f64, double)

#define ORC_AABB(SUF, R)                                                          \
    void orc_aabbs_##SUF(int ng, const R *geoms, const R *bodies, R *aabb) {      \
        for (int i = 0; i < ng; i++) {                                            \
            GeomView<R> v = geom_view(geoms + (size_t)i * ORC_GEOM_STRIDE, bodies); \
            geom_aabb(v, aabb + (size_t)i * 6);                                   \
        }                                                                         \
    }
ORC_AABB(f32, float)
ORC_AABB(f64, double)

int orc_pairs_sap_f32(int ng, const float *aabb, const int *body, const unsigned *cat, const unsigned *col, int *pairs, int cap) {
    return pairs_sap(ng, aabb, body, cat, col, pairs, cap);
}
int orc_pairs_sap_f64(int ng, const double *aabb, const int *body, const unsigned *cat, const unsigned *col, int *pairs, int cap) {
    return pairs_sap(ng, aabb, body, cat, col, pairs, cap);
}
int orc_pairs_hash_f32(int ng, const float *aabb, const int *body, const unsigned *cat, const unsigned *col, int lo, int hi, int *pairs, int cap) {
    return pairs_hash(ng, aabb, body, cat, col, lo, hi, pairs, cap);
}
int orc_pairs_hash_f64(int ng, const double *aabb, const int *body, const unsigned *cat, const unsigned *col, int lo, int hi, int *pairs, int cap) {
    return pairs_hash(ng, aabb, body, cat, col, lo, hi, pairs, cap);
}

int orc_contact_rows_f64(const double *contact, const double *surface, const double *body1, const double *body2,
                         int reversed, double fps, double erp, double min_depth, double max_vel, double *J, int *findex) {
    Joint<double> j;
    j.type = JT_CONTACT;
    j.reverse = reversed != 0;
    for (int i = 0; i < 3; i++) { j.pos[i] = contact[i]; j.normal[i] = contact[3 + i]; j.fdir1[i] = surface[12 + i]; }
    j.depth = contact[6];
    j.s.mode = (int)surface[0]; j.s.mu = surface[1]; j.s.mu2 = surface[2]; j.s.bounce = surface[3];
    j.s.bounce_vel = surface[4]; j.s.soft_erp = surface[5]; j.s.soft_cfm = surface[6]; j.s.motion1 = surface[7];
    j.s.motion2 = surface[8]; j.s.motionN = surface[9]; j.s.slip1 = surface[10]; j.s.slip2 = surface[11];
    int m = contact_m(j.s);
    Row<double> rows[6];
    std::memset(rows, 0, sizeof rows);
    int fidx[6] = {-1, -1, -1, -1, -1, -1};
    contact_info2(j, m, body1, body2, fps, erp, min_depth, max_vel, rows, fidx);
    for (int r = 0; r < m; r++) {
        double *o = J + r * 16;
        for (int k = 0; k < 6; k++) { o[k] = rows[r].J1[k]; o[8 + k] = rows[r].J2[k]; }
        o[6] = rows[r].rhs; o[7] = rows[r].cfm; o[14] = rows[r].lo; o[15] = rows[r].hi;
        findex[r] = fidx[r];
    }
    return m;
}

int orc_quickstep_f32(int nb, float *bodies, int nj, const float *joints, const double *params, float *lambda, int cap) {
    return quickstep_api(nb, bodies, nj, joints, params, lambda, cap);
}
int orc_quickstep_f64(int nb, double *bodies, int nj, const double *joints, const double *params, double *lambda, int cap) {
    return quickstep_api(nb, bodies, nj, joints, params, lambda, cap);
}
int orc_quickstep_ext_f32(int nb, float *bodies, int nj, const float *joints, const float *jext, const double *params,
                          float *lambda, int cap) {
    return quickstep_api(nb, bodies, nj, joints, params, lambda, cap, jext);
}
int orc_quickstep_ext_f64(int nb, double *bodies, int nj, const double *joints, const double *jext,
                          const double *params, double *lambda, int cap) {
    return quickstep_api(nb, bodies, nj, joints, params, lambda, cap, jext);
}

#define ORC_QUICKSTEP_TABLE(SUF, R)                                                                          \
    int orc_quickstep_table_##SUF(int nb, R *bodies, int nj, const R *joints, const double *params, int nmat,  \
                                  const double *table, R *lambda, int cap) {                                  \
        const SurfaceTable<R> T = read_table<R>(nmat, table, 0);                                            \
        return quickstep_api(nb, bodies, nj, joints, params, lambda, cap, (const R *)nullptr, &T);           \
    }
ORC_QUICKSTEP_TABLE(f32, float)
ORC_QUICKSTEP_TABLE(f64, double)

int orc_world_step_f32(int nb, float *bodies, int ng, const float *geoms, int nball, const float *balls, const double *params,
                       float *contacts, int ccap, float *lambda, int lcap, int *m_out) {
    return world_step(nb, bodies, ng, geoms, nball, balls, params, contacts, ccap, lambda, lcap, m_out);
}
int orc_world_step_f64(int nb, double *bodies, int ng, const double *geoms, int nball, const double *balls, const double *params,
                       double *contacts, int ccap, double *lambda, int lcap, int *m_out) {
    return world_step(nb, bodies, ng, geoms, nball, balls, params, contacts, ccap, lambda, lcap, m_out);
}

#define ORC_WORLD_STEP_EXT(SUF, R)                                                                             \
    int orc_world_step_ext_##SUF(int nb, R *bodies, int ng, const R *geoms, int nball, const R *balls,          \
                                 const R *bext, const double *params, R *contacts, int ccap, R *lambda,        \
                                 int lcap, int *m_out) {                                                       \
        return world_step(nb, bodies, ng, geoms, nball, balls, params, contacts, ccap, lambda, lcap, m_out,   \
                          (const SurfaceTable<R> *)nullptr, bext);                                             \
    }
ORC_WORLD_STEP_EXT(f32, float)
ORC_WORLD_STEP_EXT(f64, double)

#define ORC_WORLD_STEP_TABLE(SUF, R)                                                                           \
    int orc_world_step_table_##SUF(int nb, R *bodies, int ng, const R *geoms, int nball, const R *balls,        \
                                   const double *params, int nmat, const double *table, int filter,            \
                                   R *contacts, int ccap, R *lambda, int lcap, int *m_out) {                   \
        const SurfaceTable<R> T = read_table<R>(nmat, table, filter);                                          \
        return world_step(nb, bodies, ng, geoms, nball, balls, params, contacts, ccap, lambda, lcap, m_out, &T); \
    }
ORC_WORLD_STEP_TABLE(f32, float)
ORC_WORLD_STEP_TABLE(f64, double)

/* contacts here are JOINT records [16] (type 0, b1, b2, pos, normal, depth). */
double orc_lcp_residual(int nb, const double *bodies, int nball, const double *balls, int nc, const double *contacts,
                        const double *params, const double *lambda) {
    std::vector<double> bs;
    System<double> S;
    build_system_f64(nb, bodies, nball, balls, nc, contacts, params, bs, S);
    std::vector<double> lam(lambda, lambda + S.m), fc((size_t)nb * 6, 0);
    for (int i = 0; i < S.m; i++) {
        const double *o = &S.iMJ[(size_t)i * 12];
        double *f1 = &fc[(size_t)S.jb1[i] * 6];
        for (int k = 0; k < 6; k++) f1[k] += lam[i] * o[k];
        if (S.jb2[i] != -1) {
            double *f2 = &fc[(size_t)S.jb2[i] * 6];
            for (int k = 0; k < 6; k++) f2[k] += lam[i] * o[6 + k];
        }
    }
    Params<double> p = read_params<double>(params);
    return residual(S, lam, fc, p.w, p.lambda_floor);
}

int orc_solve_to_residual(int nb, const double *bodies, int nball, const double *balls, int nc, const double *contacts,
                          const double *params, double target, int maxit, double *lambda, double *res_out) {
    std::vector<double> bs;
    System<double> S;
    build_system_f64(nb, bodies, nball, balls, nc, contacts, params, bs, S);
    Params<double> p = read_params<double>(params);
    std::vector<double> lam(S.m, 0), fc((size_t)nb * 6, 0);
    int it = 0;
    double res = S.m ? 1.0 : 0.0;
    while (S.m && it < maxit) {
        for (int k = 0; k < S.m; k++) sor_row(S, S.order[k], lam, fc);
        it++;
        res = residual(S, lam, fc, p.w, p.lambda_floor);
        if (res <= target) break;
    }
    for (int i = 0; i < S.m; i++) lambda[i] = lam[i];
    if (res_out) *res_out = res;
    return it;
}

void orc_normalize3_f32(float *a) { normalize3(a); }
void orc_normalize3_f64(double *a) { normalize3(a); }
void orc_plane_space_f64(const double *n, double *p, double *q) { planeSpace(n, p, q); }

int orc_last_sweeps(void) { return g_last_sweeps; }

}  // extern "C"
