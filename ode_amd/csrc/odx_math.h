// odx_math.h -- device-side vector and rotation helpers for the step path.
//
// 3x4 row-major rotation storage and the helper semantics follow the
// reference's include/ode/odemath.h and ode/src/rotation.cpp (cited per
// function).  All device code computes in float (dReal = float).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#define ODX_DEV __device__ __forceinline__

namespace odx {

constexpr float kInf = __builtin_huge_valf();

ODX_DEV float dot3(const float *a, const float *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
// a . column j of a 3x4 matrix (dCalcVectorDot3_14)
ODX_DEV float dot14(const float *a, const float *M) { return a[0] * M[0] + a[1] * M[4] + a[2] * M[8]; }
ODX_DEV float dot44(const float *a, const float *b) { return a[0] * b[0] + a[4] * b[4] + a[8] * b[8]; }
ODX_DEV float dot41(const float *M, const float *b) { return M[0] * b[0] + M[4] * b[1] + M[8] * b[2]; }

ODX_DEV void cross3(float *r, const float *a, const float *b) {
    float x = a[1] * b[2] - a[2] * b[1];
    float y = a[2] * b[0] - a[0] * b[2];
    float z = a[0] * b[1] - a[1] * b[0];
    r[0] = x; r[1] = y; r[2] = z;
}
// r = A b (dMultiply0_331)
ODX_DEV void mul0_331(float *r, const float *A, const float *b) {
    float x = A[0] * b[0] + A[1] * b[1] + A[2] * b[2];
    float y = A[4] * b[0] + A[5] * b[1] + A[6] * b[2];
    float z = A[8] * b[0] + A[9] * b[1] + A[10] * b[2];
    r[0] = x; r[1] = y; r[2] = z;
}
// r = A^T b (dMultiply1_331)
ODX_DEV void mul1_331(float *r, const float *A, const float *b) {
    float x = A[0] * b[0] + A[4] * b[1] + A[8] * b[2];
    float y = A[1] * b[0] + A[5] * b[1] + A[9] * b[2];
    float z = A[2] * b[0] + A[6] * b[1] + A[10] * b[2];
    r[0] = x; r[1] = y; r[2] = z;
}
// A = B C (dMultiply0_333), A = B C^T (dMultiply2_333), 3x4 storage
ODX_DEV void mul0_333(float *A, const float *B, const float *C) {
#pragma unroll
    for (int i = 0; i < 3; i++) {
#pragma unroll
        for (int j = 0; j < 3; j++) A[i * 4 + j] = B[i * 4] * C[j] + B[i * 4 + 1] * C[4 + j] + B[i * 4 + 2] * C[8 + j];
        A[i * 4 + 3] = 0.f;
    }
}
ODX_DEV void mul2_333(float *A, const float *B, const float *C) {
#pragma unroll
    for (int i = 0; i < 3; i++) {
#pragma unroll
        for (int j = 0; j < 3; j++)
            A[i * 4 + j] = B[i * 4] * C[j * 4] + B[i * 4 + 1] * C[j * 4 + 1] + B[i * 4 + 2] * C[j * 4 + 2];
        A[i * 4 + 3] = 0.f;
    }
}

// Rotation matrix of a unit quaternion (w,x,y,z): rotation.cpp:236-255.
ODX_DEV void RfromQ(float *R, const float *q) {
    float qq1 = 2.f * q[1] * q[1], qq2 = 2.f * q[2] * q[2], qq3 = 2.f * q[3] * q[3];
    R[0] = 1.f - qq2 - qq3; R[1] = 2.f * (q[1] * q[2] - q[0] * q[3]); R[2] = 2.f * (q[1] * q[3] + q[0] * q[2]); R[3] = 0.f;
    R[4] = 2.f * (q[1] * q[2] + q[0] * q[3]); R[5] = 1.f - qq1 - qq3; R[6] = 2.f * (q[2] * q[3] - q[0] * q[1]); R[7] = 0.f;
    R[8] = 2.f * (q[1] * q[3] - q[0] * q[2]); R[9] = 2.f * (q[2] * q[3] + q[0] * q[1]); R[10] = 1.f - qq1 - qq2; R[11] = 0.f;
}

// Safe normalisation with the (1,0,0) fallback: odemath.cpp:94-160, odemath.h:36-45.
ODX_DEV void normalize3(float *a) {
    float a0 = fabsf(a[0]), a1 = fabsf(a[1]), a2 = fabsf(a[2]);
    int idx;
    if (a1 > a0) idx = (a2 > a1) ? 2 : 1;
    else if (a2 > a0) idx = 2;
    else {
        if (!(a0 > 0.f)) { a[0] = 1.f; a[1] = 0.f; a[2] = 0.f; return; }
        idx = 0;
    }
    if (idx == 0) {
        float rr = 1.f / a0, y = a[1] * rr, z = a[2] * rr, l = 1.f / sqrtf(1.f + y * y + z * z);
        a[1] = y * l; a[2] = z * l; a[0] = copysignf(l, a[0]);
    } else if (idx == 1) {
        float rr = 1.f / a1, x = a[0] * rr, z = a[2] * rr, l = 1.f / sqrtf(1.f + x * x + z * z);
        a[0] = x * l; a[2] = z * l; a[1] = copysignf(l, a[1]);
    } else {
        float rr = 1.f / a2, x = a[0] * rr, y = a[1] * rr, l = 1.f / sqrtf(1.f + x * x + y * y);
        a[0] = x * l; a[1] = y * l; a[2] = copysignf(l, a[2]);
    }
}

// Orthonormal tangent pair of a unit normal: odemath.cpp:224-251.
ODX_DEV void planeSpace(const float *n, float *p, float *q) {
    if (fabsf(n[2]) > 0.70710678118654752440f) {
        float a = n[1] * n[1] + n[2] * n[2], k = 1.f / sqrtf(a);
        p[0] = 0.f; p[1] = -n[2] * k; p[2] = n[1] * k;
        q[0] = a * k; q[1] = -n[0] * p[2]; q[2] = n[0] * p[1];
    } else {
        float a = n[0] * n[0] + n[1] * n[1], k = 1.f / sqrtf(a);
        p[0] = -n[1] * k; p[1] = n[0] * k; p[2] = 0.f;
        q[0] = -n[2] * p[1]; q[1] = n[2] * p[0]; q[2] = a * k;
    }
}

// Closed-form 3x3 inverse returning the determinant (0 = singular):
// include/ode/odemath.h:444-490.
ODX_DEV float invert3(float *dst, const float *ma) {
    float det = ma[0] * (ma[5] * ma[10] - ma[9] * ma[6]) - ma[1] * (ma[4] * ma[10] - ma[8] * ma[6]) +
                ma[2] * (ma[4] * ma[9] - ma[8] * ma[5]);
    if (det == 0.f) return 0.f;
    float d = 1.f / det;
    dst[0] = (ma[5] * ma[10] - ma[6] * ma[9]) * d;
    dst[1] = (ma[9] * ma[2] - ma[1] * ma[10]) * d;
    dst[2] = (ma[1] * ma[6] - ma[5] * ma[2]) * d;
    dst[3] = 0.f;
    dst[4] = (ma[6] * ma[8] - ma[4] * ma[10]) * d;
    dst[5] = (ma[0] * ma[10] - ma[8] * ma[2]) * d;
    dst[6] = (ma[4] * ma[2] - ma[0] * ma[6]) * d;
    dst[7] = 0.f;
    dst[8] = (ma[4] * ma[9] - ma[8] * ma[5]) * d;
    dst[9] = (ma[8] * ma[1] - ma[0] * ma[9]) * d;
    dst[10] = (ma[0] * ma[5] - ma[1] * ma[4]) * d;
    dst[11] = 0.f;
    return det;
}

}  // namespace odx
