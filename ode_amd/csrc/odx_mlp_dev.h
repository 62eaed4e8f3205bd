// odx_mlp_dev.h -- device side of the warm-start network shared by the
// training/inference kernels (odx_mlp.hip) and the fused step kernel
// (odx_step.hip): operand types, sizes, parameter layout and the 16-sample
// forward tile on v_mfma_f32_16x16x32_bf16.
#pragma once
#include <hip/hip_runtime.h>

namespace odx {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int NIN = 16, NH = 128, NOUT = 3;
constexpr int KIN = 32;     // layer-1 K padded to one MFMA K step
constexpr int NOUTP = 16;   // layer-3 N padded to one MFMA N tile

// Operand layout of v_mfma_f32_16x16x32_bf16 (wave64), lane l, element j:
//   A (16 x 32): row = l % 16, k = 8 * (l / 16) + j      (j < 8)
//   B (32 x 16): col = l % 16, k = 8 * (l / 16) + j      (j < 8)
//   D (16 x 16): col = l % 16, row = 4 * (l / 16) + j    (j < 4)
// (checked on the device by odxMfmaProbe / tests/test_gpu_mlp.py)

struct MlpDev {
    float *W1, *b1, *W2, *b2, *W3, *b3;        // fp32 masters, [out][in]
    __bf16 *W1h, *W2h, *W3h;                  // bf16 copies: [128][32], [128][128], [16][128]
    float *gW1, *gb1, *gW2, *gb2, *gW3, *gb3;  // gradients
    float *mom, *vel;                          // Adam moments over the flat parameter vector
    float *norm;                               // in_mu[16] in_sig[16] out_mu[3] out_sig[3]
    float *loss;                               // [2] sum of squared error, samples
    int gate;                                  // 1: matched contacts keep their previous impulses
};

// parameter vector layout (flat offsets)
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + NH * NIN, OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH,
              OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NOUT * NH, NPARAM = OFF_B3 + NOUT;

__device__ __forceinline__ float gelu(float x, float &dgelu) {
    // tanh form: 0.5 x (1 + tanh(c (x + a x^3))), tanh(u) = 1 - 2 / (1 + e^{2u})
    // (e^{2u} overflows to +inf for large u, which gives 1; the result feeds
    // bf16 activations, so the few-ulp error of the exp/rcp pair is invisible)
    const float c = 0.7978845608028654f, a = 0.044715f;
    const float u = c * (x + a * x * x * x);
    const float th = 1.f - 2.f * __frcp_rn(1.f + __expf(2.f * u));
    dgelu = 0.5f * (1.f + th) + 0.5f * x * (1.f - th * th) * c * (1.f + 3.f * a * x * x);
    return 0.5f * x * (1.f + th);
}

__device__ __forceinline__ f32x4 mfma(const bf16x8 &a, const bf16x8 &b, const f32x4 &c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x8 load8(const __bf16 *p) { return *reinterpret_cast<const bf16x8 *>(p); }

// Per-wave scratch of one 16-sample tile (LDS).
struct TileBuf {
    __bf16 act1[16][NH];
    __bf16 act2[16][NH];
    __bf16 g1[16][NH];   // gelu'(z1), later dZ1
    __bf16 g2[16][NH];   // gelu'(z2), later dZ2
    float x[16][NIN];    // normalised input
    float dz3[16][4];
};

// Inference scratch of one tile: activations only.
struct InferBuf {
    __bf16 act1[16][NH];
    __bf16 act2[16][NH];
};

// Forward pass of one 16-sample tile by one wave.  xr: raw features of the
// tile's samples (invalid samples produce garbage rows that callers ignore).
// Returns the layer-3 accumulator (D layout, cols 0..2 are the outputs).
// keep_grad (training, Buf = TileBuf) also stores gelu' of both layers.
template <bool keep_grad, class Buf>
__device__ f32x4 forward_tile(const __bf16 *W1h, const __bf16 *W2h, const __bf16 *W3h, const float *b1,
                              const float *b2, const float *b3, const float *nrm, Buf &tb, const float *xr,
                              int nvalid) {
    const int l = threadIdx.x & 63;
    const int row = l & 15, g = l >> 4;
    // normalised inputs: lane (row, g < 2) holds features 8g..8g+7 of sample row
    float xn[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const int f = 8 * (g & 1) + j;
        const float v = row < nvalid ? xr[(size_t)row * NIN + f] : 0.f;
        xn[j] = (v - nrm[f]) / nrm[NIN + f];
    }
    if constexpr (keep_grad) {
        if (g < 2) {
#pragma unroll
            for (int j = 0; j < 8; j++) tb.x[row][8 * g + j] = xn[j];
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
    bf16x8 a1;
#pragma unroll
    for (int j = 0; j < 8; j++) a1[j] = g < 2 ? (__bf16)xn[j] : (__bf16)0.f;
    // layer 1
#pragma unroll
    for (int t = 0; t < NH / 16; t++) {
        const int n = 16 * t + row;
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma(a1, load8(W1h + (size_t)n * KIN + 8 * g), acc);
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int r = 4 * g + j;
            float d;
            const float h = gelu(acc[j] + b1[n], d);
            tb.act1[r][n] = (__bf16)h;
            if constexpr (keep_grad) tb.g1[r][n] = (__bf16)d;
        }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    // layer 2
    bf16x8 a2[4];
#pragma unroll
    for (int s = 0; s < 4; s++) a2[s] = load8(&tb.act1[row][32 * s + 8 * g]);
#pragma unroll
    for (int t = 0; t < NH / 16; t++) {
        const int n = 16 * t + row;
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < 4; s++) acc = mfma(a2[s], load8(W2h + (size_t)n * NH + 32 * s + 8 * g), acc);
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int r = 4 * g + j;
            float d;
            const float h = gelu(acc[j] + b2[n], d);
            tb.act2[r][n] = (__bf16)h;
            if constexpr (keep_grad) tb.g2[r][n] = (__bf16)d;
        }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    // layer 3 (outputs padded to 16)
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < 4; s++)
        acc = mfma(load8(&tb.act2[row][32 * s + 8 * g]), load8(W3h + (size_t)row * NH + 32 * s + 8 * g), acc);
    const int n = row;
#pragma unroll
    for (int j = 0; j < 4; j++) acc[j] += n < NOUT ? b3[n] : 0.f;
    return acc;
}

}  // namespace odx
