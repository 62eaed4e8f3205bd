// odx_mlp.hip -- learned warm start: a 16 -> 128 -> 128 -> 3 GELU MLP on bf16
// MFMA (v_mfma_f32_16x16x32_bf16, fp32 accumulate) for gfx950.
//
// It maps a contact's features (normal, depth, relative normal/tangent
// velocity, matched previous-step impulses, inverse masses, ...) to initial
// normal/friction impulses for the SOR-LCP solve.  The reference has no such
// component: its warm start is compiled out (quickstep.cpp:46-51) and the
// solve starts from lambda = 0 every step (:1929-1972).  Cold-start PGS stays
// the fallback and the arbiter: sweeps always continue to the same residual.
//
// Kernels: k_mlp_infer (forward, writes lambda0 rows), k_mlp_train (forward +
// backward over sampled mini-batches, fp32 gradient accumulation), k_mlp_adam
// (Adam on fp32 master weights, refreshes the bf16 copies), k_mfma_probe
// (operand-layout check of the MFMA builtin).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/odx.h"
#include "odx_device.h"
#include "odx_host.h"
#include "odx_mlp_dev.h"

namespace odx {

// ------------------------------------------------------------ inference --
// One wave per workgroup, one 16-contact tile at a time; the weights are read
// from global memory (L2) rather than staged in LDS, so that a workgroup needs
// only its 8 KB activation buffer and finds room on CUs whose LDS is held by
// the concurrent solve kernels.
__global__ __launch_bounds__(64) void k_mlp_infer(MlpDev M, const float *feat, const int *feat_row,
                                                  const int *count, float *lambda0) {
    __shared__ InferBuf tb;
    const int n = *count;
    const int ntiles = (n + 15) / 16;
    // the count is known only on the device: workgroups beyond it leave
    if (n <= 0 || (int)blockIdx.x >= ntiles) return;
    const int l = threadIdx.x & 63;
    const float *nrm = M.norm;
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int s0 = tile * 16;
        const int nv = min(16, n - s0);
        f32x4 y = forward_tile<false>(M.W1h, M.W2h, M.W3h, M.b1, M.b2, M.b3, nrm, tb, feat + (size_t)s0 * NIN, nv);
        const int o = l & 15, g = l >> 4;
        if (o < NOUT) {
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int smp = s0 + 4 * g + j;
                if (smp < n) {
                    // the network predicts the change from the matched previous impulse
                    // (features 7-9, zero when unmatched), in units of the force scale s;
                    // with the gate on, matched contacts (feature 12) keep that impulse
                    const float *x = feat + (size_t)smp * NIN;
                    const float use = (M.gate && x[12] > 0.5f) ? 0.f : 1.f;
                    lambda0[feat_row[smp] + o] =
                        x[15] * (x[7 + o] + use * (y[j] * nrm[2 * NIN + NOUT + o] + nrm[2 * NIN + o]));
                }
            }
        }
        // the next tile reuses tb
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
}

// ------------------------------------------------------------- training --
__device__ __forceinline__ unsigned hash32(unsigned x) {
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}

// Forward + backward over `chunks` 64-sample chunks per block.  Samples are
// X[idx], T[idx] with idx drawn from (seed, position) unless seq != 0 (then
// samples are taken in order from 0).  Gradients of the mean squared error of
// the normalised outputs are accumulated into M.g* with one atomic add per
// weight per block.
__global__ __launch_bounds__(256) void k_mlp_train(MlpDev M, const float *X, const float *T, int N, int batch,
                                                   unsigned seed, int seq) {
    __shared__ __align__(16) __bf16 sW1[NH * KIN];
    __shared__ __align__(16) __bf16 sW2[NH * NH];
    __shared__ __align__(16) __bf16 sW3[NOUTP * NH];
    __shared__ float sb[NH + NH + 4 + 38];
    __shared__ TileBuf tb[4];
    const int tid = threadIdx.x, wv = tid >> 6, l = tid & 63;
    const int row = l & 15, g = l >> 4;
    for (int i = tid; i < NH * KIN / 8; i += 256) reinterpret_cast<bf16x8 *>(sW1)[i] = reinterpret_cast<const bf16x8 *>(M.W1h)[i];
    for (int i = tid; i < NH * NH / 8; i += 256) reinterpret_cast<bf16x8 *>(sW2)[i] = reinterpret_cast<const bf16x8 *>(M.W2h)[i];
    for (int i = tid; i < NOUTP * NH / 8; i += 256) reinterpret_cast<bf16x8 *>(sW3)[i] = reinterpret_cast<const bf16x8 *>(M.W3h)[i];
    for (int i = tid; i < NH; i += 256) { sb[i] = M.b1[i]; sb[NH + i] = M.b2[i]; }
    if (tid < NOUT) sb[2 * NH + tid] = M.b3[tid];
    if (tid < 38) sb[2 * NH + 4 + tid] = M.norm[tid];
    __syncthreads();
    const float *nrm = sb + 2 * NH + 4;
    const float inv_b = 1.f / (float)batch;

    // register accumulators: dW2 tiles owned by this wave (mt 0..7, nt = 2*wv, 2*wv+1)
    f32x4 accW2[8][2];
#pragma unroll
    for (int a = 0; a < 8; a++) for (int b = 0; b < 2; b++) accW2[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    float accW1[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};  // (out = tid/2, in = 8*(tid&1)+j)
    float accW3[2] = {0.f, 0.f};                                 // pairs (o,k) = tid, tid+256 of 384
    float accb1 = 0.f, accb2 = 0.f, accb3 = 0.f, lsum = 0.f;
    __shared__ float sx_raw[64][NIN];
    __shared__ float st[64][NOUT];

    const int nchunks = (batch + 63) / 64;
    for (int c = blockIdx.x; c < nchunks; c += gridDim.x) {
        // gather the chunk's samples
        for (int e = tid; e < 64 * NIN; e += 256) {
            const int s = e / NIN, f = e % NIN;
            const int pos = c * 64 + s;
            float v = 0.f;
            if (pos < batch) {
                const unsigned id = seq ? (unsigned)pos % (unsigned)N : hash32(seed * 0x9e3779b9u + (unsigned)pos) % (unsigned)N;
                v = X[(size_t)id * NIN + f];
                if (f < NOUT) st[s][f] = T[(size_t)id * NOUT + f];
            }
            sx_raw[s][f] = v;
        }
        __syncthreads();
        const int base = wv * 16;
        const int nv = max(0, min(16, batch - (c * 64 + base)));
        f32x4 y = forward_tile<true>(sW1, sW2, sW3, sb, sb + NH, sb + 2 * NH, nrm, tb[wv], &sx_raw[base][0], nv);
        // dZ3 = 2 (y - t_norm) / batch for valid samples
        {
            const int o = row;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int r = 4 * g + j;
                float d = 0.f;
                if (o < NOUT && r < nv) {
                    const float tn = (st[base + r][o] - nrm[2 * NIN + o]) / nrm[2 * NIN + NOUT + o];
                    const float e = y[j] - tn;
                    lsum += e * e;
                    d = 2.f * e * inv_b;
                }
                if (o < 4) tb[wv].dz3[r][o] = d;
            }
        }
        __syncthreads();
        // dW3, db3 (VALU): pairs p = (o, k)
        for (int q = 0; q < 2; q++) {
            const int p = tid + 256 * q;
            if (p < NOUT * NH) {
                const int o = p / NH, k = p % NH;
                float s = 0.f;
                for (int smp = 0; smp < 64; smp++) s += tb[smp >> 4].dz3[smp & 15][o] * (float)tb[smp >> 4].act2[smp & 15][k];
                accW3[q] += s;
            }
        }
        if (tid < NOUT) {
            float s = 0.f;
            for (int smp = 0; smp < 64; smp++) s += tb[smp >> 4].dz3[smp & 15][tid];
            accb3 += s;
        }
        // dZ2 = (dZ3 W3) * gelu'(z2), stored over g2 (bf16)
        for (int e = tid; e < 64 * NH; e += 256) {
            const int smp = e / NH, k = e % NH;
            TileBuf &b = tb[smp >> 4];
            float s = 0.f;
#pragma unroll
            for (int o = 0; o < NOUT; o++) s += b.dz3[smp & 15][o] * M.W3[o * NH + k];
            b.g2[smp & 15][k] = (__bf16)(s * (float)b.g2[smp & 15][k]);
        }
        __syncthreads();
        // dW2 += dZ2^T act1 (MFMA, K = 64 samples in two steps)
#pragma unroll
        for (int mt = 0; mt < 8; mt++) {
#pragma unroll
            for (int q = 0; q < 2; q++) {
                const int nt = 2 * wv + q;
#pragma unroll
                for (int ks = 0; ks < 2; ks++) {
                    bf16x8 a, b;
#pragma unroll
                    for (int j = 0; j < 8; j++) {
                        const int smp = 32 * ks + 8 * g + j;
                        a[j] = tb[smp >> 4].g2[smp & 15][16 * mt + row];
                        b[j] = tb[smp >> 4].act1[smp & 15][16 * nt + row];
                    }
                    accW2[mt][q] = mfma(a, b, accW2[mt][q]);
                }
            }
        }
        if (tid < NH) {
            float s = 0.f;
            for (int smp = 0; smp < 64; smp++) s += (float)tb[smp >> 4].g2[smp & 15][tid];
            accb2 += s;
        }
        // dZ1 = (dZ2 W2) * gelu'(z1) per wave (MFMA), stored over g1
        {
            TileBuf &b = tb[wv];
            bf16x8 a[4];
#pragma unroll
            for (int s = 0; s < 4; s++) a[s] = load8(&b.g2[row][32 * s + 8 * g]);
#pragma unroll
            for (int t = 0; t < NH / 16; t++) {
                const int kin = 16 * t + row;
                f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    bf16x8 w;
#pragma unroll
                    for (int j = 0; j < 8; j++) w[j] = sW2[(size_t)(32 * s + 8 * g + j) * NH + kin];
                    acc = mfma(a[s], w, acc);
                }
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int r = 4 * g + j;
                    b.g1[r][kin] = (__bf16)(acc[j] * (float)b.g1[r][kin]);
                }
            }
        }
        __syncthreads();
        // dW1 += dZ1^T x, db1 (VALU)
        {
            const int o = tid >> 1, i0 = 8 * (tid & 1);
            for (int smp = 0; smp < 64; smp++) {
                const float dz = (float)tb[smp >> 4].g1[smp & 15][o];
#pragma unroll
                for (int j = 0; j < 8; j++) accW1[j] += dz * (float)(__bf16)tb[smp >> 4].x[smp & 15][i0 + j];
            }
        }
        if (tid < NH) {
            float s = 0.f;
            for (int smp = 0; smp < 64; smp++) s += (float)tb[smp >> 4].g1[smp & 15][tid];
            accb1 += s;
        }
        __syncthreads();
    }
    // flush accumulators
#pragma unroll
    for (int mt = 0; mt < 8; mt++)
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int nt = 2 * wv + q;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int out = 16 * mt + 4 * g + j, in = 16 * nt + row;
                atomicAdd(&M.gW2[out * NH + in], accW2[mt][q][j]);
            }
        }
    {
        const int o = tid >> 1, i0 = 8 * (tid & 1);
#pragma unroll
        for (int j = 0; j < 8; j++) atomicAdd(&M.gW1[o * NIN + i0 + j], accW1[j]);
    }
    for (int q = 0; q < 2; q++) {
        const int p = tid + 256 * q;
        if (p < NOUT * NH) atomicAdd(&M.gW3[p], accW3[q]);
    }
    if (tid < NH) { atomicAdd(&M.gb1[tid], accb1); atomicAdd(&M.gb2[tid], accb2); }
    if (tid < NOUT) atomicAdd(&M.gb3[tid], accb3);
    // loss: reduce within the wave then one atomic per wave
    for (int o = 32; o; o >>= 1) lsum += __shfl_xor(lsum, o, 64);
    if (l == 0) atomicAdd(&M.loss[0], lsum);
}

// Adam on the fp32 masters (flat parameter vector), then refresh the bf16
// copies and clear the gradients.
__global__ void k_mlp_adam(MlpDev M, float lr, float b1, float b2, float eps, float bc1, float bc2, int update) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NPARAM) return;
    float *p, *gp;
    int k;
    if (i < OFF_B1) { k = i - OFF_W1; p = M.W1 + k; gp = M.gW1 + k; }
    else if (i < OFF_W2) { k = i - OFF_B1; p = M.b1 + k; gp = M.gb1 + k; }
    else if (i < OFF_B2) { k = i - OFF_W2; p = M.W2 + k; gp = M.gW2 + k; }
    else if (i < OFF_W3) { k = i - OFF_B2; p = M.b2 + k; gp = M.gb2 + k; }
    else if (i < OFF_B3) { k = i - OFF_W3; p = M.W3 + k; gp = M.gW3 + k; }
    else { k = i - OFF_B3; p = M.b3 + k; gp = M.gb3 + k; }
    if (update) {
        const float gr = *gp;
        const float m = b1 * M.mom[i] + (1.f - b1) * gr;
        const float v = b2 * M.vel[i] + (1.f - b2) * gr * gr;
        M.mom[i] = m;
        M.vel[i] = v;
        *p -= lr * (m / bc1) / (sqrtf(v / bc2) + eps);
        *gp = 0.f;
    }
    // bf16 copies
    if (i < OFF_B1) {
        const int o = k / NIN, in = k % NIN;
        M.W1h[o * KIN + in] = (__bf16)*p;
        M.W1h[o * KIN + NIN + in] = (__bf16)0.f;
    } else if (i >= OFF_W2 && i < OFF_B2) {
        M.W2h[k] = (__bf16)*p;
    } else if (i >= OFF_W3 && i < OFF_B3) {
        M.W3h[k] = (__bf16)*p;
    }
}

// ------------------------------------------------------------ MFMA probe --
__global__ void k_mfma_probe(const float *a, const float *b, float *d) {
    const int l = threadIdx.x;
    bf16x8 A, B;
    for (int j = 0; j < 8; j++) { A[j] = (__bf16)a[l * 8 + j]; B[j] = (__bf16)b[l * 8 + j]; }
    f32x4 c = {0.f, 0.f, 0.f, 0.f};
    c = mfma(A, B, c);
    for (int j = 0; j < 4; j++) d[l * 4 + j] = c[j];
}

}  // namespace odx

using namespace odx;

struct odxMlp {
    MlpDev M{};
    std::vector<void *> allocs;
    hipStream_t stream = nullptr;
    float *X = nullptr, *T = nullptr;
    int N = 0, step = 0;
    float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    template <class Tp> bool alloc(Tp **p, size_t n) {
        *p = nullptr;
        if (!odx_check(hipMalloc((void **)p, (n ? n : 1) * sizeof(Tp)), "hipMalloc")) return false;
        allocs.push_back((void *)*p);
        return true;
    }
    ~odxMlp() {
        if (stream) (void)hipStreamDestroy(stream);
        for (void *p : allocs) (void)hipFree(p);
    }
};

namespace odx {
// device view of a network, for the fused step kernel
const MlpDev *mlp_device_view(const odxMlp *m) { return m ? &m->M : nullptr; }
// used by the batch runtime to run inference inside the step graph
// max_entries bounds the device-side count; max_blocks caps the grid (a
// gated network infers only the few unmatched contacts of a step)
void launch_mlp_infer(const odxMlp *m, const float *feat, const int *feat_row, const int *count, float *lambda0,
                      int max_entries, hipStream_t s, int max_blocks) {
    int blocks = (max_entries + 15) / 16;
    if (blocks > max_blocks) blocks = max_blocks;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_mlp_infer, dim3(blocks), dim3(64), 0, s, m->M, feat, feat_row, count, lambda0);
}
int mlp_gate(const odxMlp *m) { return m ? m->M.gate : 0; }
}  // namespace odx

static void mlp_refresh_copies(odxMlp *m) {
    hipLaunchKernelGGL(k_mlp_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, m->stream, m->M, 0.f, 0.f, 0.f, 1.f, 1.f,
                       1.f, 0);
}

extern "C" {

int odxMfmaProbe(const float *a, const float *b, float *d) {
    float *da = nullptr, *db = nullptr, *dd = nullptr;
    bool ok = odx_check(hipMalloc((void **)&da, 512 * sizeof(float)), "hipMalloc") &&
              odx_check(hipMalloc((void **)&db, 512 * sizeof(float)), "hipMalloc") &&
              odx_check(hipMalloc((void **)&dd, 256 * sizeof(float)), "hipMalloc");
    ok = ok && odx_check(hipMemcpy(da, a, 512 * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy") &&
         odx_check(hipMemcpy(db, b, 512 * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
    if (ok) {
        hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, 0, da, db, dd);
        ok = odx_check(hipGetLastError(), "k_mfma_probe") &&
             odx_check(hipMemcpy(d, dd, 256 * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
    }
    if (da) (void)hipFree(da);
    if (db) (void)hipFree(db);
    if (dd) (void)hipFree(dd);
    return ok ? 1 : 0;
}

odxMlp *odxMlpCreate(unsigned seed) {
    auto *m = new odxMlp();
    MlpDev &M = m->M;
    bool ok = m->alloc(&M.W1, NH * NIN) && m->alloc(&M.b1, NH) && m->alloc(&M.W2, NH * NH) && m->alloc(&M.b2, NH) &&
              m->alloc(&M.W3, NOUT * NH) && m->alloc(&M.b3, NOUT) && m->alloc(&M.W1h, NH * KIN) &&
              m->alloc(&M.W2h, NH * NH) && m->alloc(&M.W3h, NOUTP * NH) && m->alloc(&M.gW1, NH * NIN) &&
              m->alloc(&M.gb1, NH) && m->alloc(&M.gW2, NH * NH) && m->alloc(&M.gb2, NH) && m->alloc(&M.gW3, NOUT * NH) &&
              m->alloc(&M.gb3, NOUT) && m->alloc(&M.mom, NPARAM) && m->alloc(&M.vel, NPARAM) && m->alloc(&M.norm, 38) &&
              m->alloc(&M.loss, 2) && odx_check(hipStreamCreate(&m->stream), "hipStreamCreate");
    if (!ok) { delete m; return nullptr; }
    // He-style uniform init from a fixed LCG (host), zero biases
    std::vector<float> P(NPARAM, 0.f);
    unsigned s = seed ? seed : 1u;
    auto rnd = [&]() { s = 1664525u * s + 1013904223u; return ((s >> 8) * (1.0f / 16777216.0f)) * 2.f - 1.f; };
    for (int i = 0; i < NH * NIN; i++) P[OFF_W1 + i] = rnd() * std::sqrt(6.f / NIN);
    for (int i = 0; i < NH * NH; i++) P[OFF_W2 + i] = rnd() * std::sqrt(6.f / NH);
    for (int i = 0; i < NOUT * NH; i++) P[OFF_W3 + i] = rnd() * std::sqrt(1.f / NH);
    std::vector<float> norm(38, 0.f);
    for (int i = 0; i < NIN; i++) norm[NIN + i] = 1.f;
    for (int i = 0; i < NOUT; i++) norm[2 * NIN + NOUT + i] = 1.f;
    ok = odx_check(hipMemcpy(M.W1, &P[OFF_W1], NH * NIN * 4, hipMemcpyHostToDevice), "hipMemcpy") &&
         odx_check(hipMemcpy(M.b1, &P[OFF_B1], NH * 4, hipMemcpyHostToDevice), "hipMemcpy") &&
         odx_check(hipMemcpy(M.W2, &P[OFF_W2], NH * NH * 4, hipMemcpyHostToDevice), "hipMemcpy") &&
         odx_check(hipMemcpy(M.b2, &P[OFF_B2], NH * 4, hipMemcpyHostToDevice), "hipMemcpy") &&
         odx_check(hipMemcpy(M.W3, &P[OFF_W3], NOUT * NH * 4, hipMemcpyHostToDevice), "hipMemcpy") &&
         odx_check(hipMemcpy(M.b3, &P[OFF_B3], NOUT * 4, hipMemcpyHostToDevice), "hipMemcpy") &&
         odx_check(hipMemcpy(M.norm, norm.data(), 38 * 4, hipMemcpyHostToDevice), "hipMemcpy") &&
         odx_check(hipMemset(M.W3h, 0, NOUTP * NH * sizeof(__bf16)), "hipMemset") &&
         odx_check(hipMemset(M.gW1, 0, NH * NIN * 4), "hipMemset") && odx_check(hipMemset(M.gb1, 0, NH * 4), "hipMemset") &&
         odx_check(hipMemset(M.gW2, 0, NH * NH * 4), "hipMemset") && odx_check(hipMemset(M.gb2, 0, NH * 4), "hipMemset") &&
         odx_check(hipMemset(M.gW3, 0, NOUT * NH * 4), "hipMemset") &&
         odx_check(hipMemset(M.gb3, 0, NOUT * 4), "hipMemset") &&
         odx_check(hipMemset(M.mom, 0, NPARAM * 4), "hipMemset") &&
         odx_check(hipMemset(M.vel, 0, NPARAM * 4), "hipMemset") && odx_check(hipMemset(M.loss, 0, 8), "hipMemset");
    if (!ok) { delete m; return nullptr; }
    mlp_refresh_copies(m);
    if (!odx_check(hipStreamSynchronize(m->stream), "hipStreamSynchronize")) { delete m; return nullptr; }
    return m;
}

void odxMlpDestroy(odxMlp *m) { delete m; }

int odxMlpSetGate(odxMlp *m, int gate) {
    if (!m || gate < 0 || gate > 1) return 0;
    m->M.gate = gate;
    return 1;
}

int odxMlpSetNormalization(odxMlp *m, const float *norm38) {
    if (!m || !norm38) return 0;
    for (int i = 0; i < NIN; i++)
        if (!(norm38[NIN + i] > 0.f)) { odx_message(1, "odxMlpSetNormalization: input scale %d must be > 0", i); return 0; }
    for (int i = 0; i < NOUT; i++)
        if (!(norm38[2 * NIN + NOUT + i] > 0.f)) { odx_message(1, "odxMlpSetNormalization: output scale %d must be > 0", i); return 0; }
    return odx_check(hipMemcpy(m->M.norm, norm38, 38 * 4, hipMemcpyHostToDevice), "hipMemcpy") ? 1 : 0;
}

int odxMlpGetParams(odxMlp *m, float *params) {
    if (!m || !params) return 0;
    const MlpDev &M = m->M;
    bool ok = odx_check(hipStreamSynchronize(m->stream), "hipStreamSynchronize") &&
              odx_check(hipMemcpy(params + OFF_W1, M.W1, NH * NIN * 4, hipMemcpyDeviceToHost), "hipMemcpy") &&
              odx_check(hipMemcpy(params + OFF_B1, M.b1, NH * 4, hipMemcpyDeviceToHost), "hipMemcpy") &&
              odx_check(hipMemcpy(params + OFF_W2, M.W2, NH * NH * 4, hipMemcpyDeviceToHost), "hipMemcpy") &&
              odx_check(hipMemcpy(params + OFF_B2, M.b2, NH * 4, hipMemcpyDeviceToHost), "hipMemcpy") &&
              odx_check(hipMemcpy(params + OFF_W3, M.W3, NOUT * NH * 4, hipMemcpyDeviceToHost), "hipMemcpy") &&
              odx_check(hipMemcpy(params + OFF_B3, M.b3, NOUT * 4, hipMemcpyDeviceToHost), "hipMemcpy");
    return ok ? NPARAM : 0;
}

int odxMlpSetParams(odxMlp *m, const float *params) {
    if (!m || !params) return 0;
    MlpDev &M = m->M;
    bool ok = odx_check(hipMemcpy(M.W1, params + OFF_W1, NH * NIN * 4, hipMemcpyHostToDevice), "hipMemcpy") &&
              odx_check(hipMemcpy(M.b1, params + OFF_B1, NH * 4, hipMemcpyHostToDevice), "hipMemcpy") &&
              odx_check(hipMemcpy(M.W2, params + OFF_W2, NH * NH * 4, hipMemcpyHostToDevice), "hipMemcpy") &&
              odx_check(hipMemcpy(M.b2, params + OFF_B2, NH * 4, hipMemcpyHostToDevice), "hipMemcpy") &&
              odx_check(hipMemcpy(M.W3, params + OFF_W3, NOUT * NH * 4, hipMemcpyHostToDevice), "hipMemcpy") &&
              odx_check(hipMemcpy(M.b3, params + OFF_B3, NOUT * 4, hipMemcpyHostToDevice), "hipMemcpy");
    if (!ok) return 0;
    mlp_refresh_copies(m);
    return odx_check(hipStreamSynchronize(m->stream), "hipStreamSynchronize") ? NPARAM : 0;
}

int odxMlpSetData(odxMlp *m, const float *X, const float *T, int N) {
    if (!m || !X || !T || N <= 0) return 0;
    if (N > m->N) {
        if (!m->alloc(&m->X, (size_t)N * NIN) || !m->alloc(&m->T, (size_t)N * NOUT)) return 0;
    }
    m->N = N;
    bool ok = odx_check(hipMemcpy(m->X, X, (size_t)N * NIN * 4, hipMemcpyHostToDevice), "hipMemcpy") &&
              odx_check(hipMemcpy(m->T, T, (size_t)N * NOUT * 4, hipMemcpyHostToDevice), "hipMemcpy");
    return ok ? N : 0;
}

// Gradient of the loss over samples [0, N) of the uploaded data (no update).
// grads: NPARAM floats in the parameter layout; loss: mean squared error.
int odxMlpGradients(odxMlp *m, float *grads, float *loss) {
    if (!m || m->N <= 0 || !grads) return 0;
    MlpDev &M = m->M;
    (void)hipMemsetAsync(M.loss, 0, 8, m->stream);
    int chunks = (m->N + 63) / 64;
    int blocks = chunks < 256 ? chunks : 256;
    hipLaunchKernelGGL(k_mlp_train, dim3(blocks), dim3(256), 0, m->stream, M, m->X, m->T, m->N, m->N, 0u, 1);
    if (!odx_check(hipGetLastError(), "k_mlp_train")) return 0;
    bool ok = odx_check(hipStreamSynchronize(m->stream), "hipStreamSynchronize") &&
              odx_check(hipMemcpy(grads + OFF_W1, M.gW1, NH * NIN * 4, hipMemcpyDeviceToHost), "hipMemcpy") &&
              odx_check(hipMemcpy(grads + OFF_B1, M.gb1, NH * 4, hipMemcpyDeviceToHost), "hipMemcpy") &&
              odx_check(hipMemcpy(grads + OFF_W2, M.gW2, NH * NH * 4, hipMemcpyDeviceToHost), "hipMemcpy") &&
              odx_check(hipMemcpy(grads + OFF_B2, M.gb2, NH * 4, hipMemcpyDeviceToHost), "hipMemcpy") &&
              odx_check(hipMemcpy(grads + OFF_W3, M.gW3, NOUT * NH * 4, hipMemcpyDeviceToHost), "hipMemcpy") &&
              odx_check(hipMemcpy(grads + OFF_B3, M.gb3, NOUT * 4, hipMemcpyDeviceToHost), "hipMemcpy");
    float l[2] = {0.f, 0.f};
    ok = ok && odx_check(hipMemcpy(l, M.loss, 8, hipMemcpyDeviceToHost), "hipMemcpy");
    if (loss) *loss = l[0] / (float)m->N;
    // clear the gradients for later training
    ok = ok && odx_check(hipMemset(M.gW1, 0, NH * NIN * 4), "hipMemset") && odx_check(hipMemset(M.gb1, 0, NH * 4), "hipMemset") &&
         odx_check(hipMemset(M.gW2, 0, NH * NH * 4), "hipMemset") && odx_check(hipMemset(M.gb2, 0, NH * 4), "hipMemset") &&
         odx_check(hipMemset(M.gW3, 0, NOUT * NH * 4), "hipMemset") && odx_check(hipMemset(M.gb3, 0, NOUT * 4), "hipMemset");
    return ok ? NPARAM : 0;
}

// nsteps Adam steps on mini-batches of `batch` samples drawn from the uploaded
// data; losses[i] receives each step's mini-batch mean squared error (may be NULL).
int odxMlpTrain(odxMlp *m, int nsteps, int batch, float lr, unsigned seed, float *losses) {
    if (!m || m->N <= 0 || nsteps <= 0 || batch <= 0) return 0;
    MlpDev &M = m->M;
    int chunks = (batch + 63) / 64;
    int blocks = chunks < 256 ? chunks : 256;
    std::vector<float> l(2);
    for (int s = 0; s < nsteps; s++) {
        if (losses) (void)hipMemsetAsync(M.loss, 0, 8, m->stream);
        hipLaunchKernelGGL(k_mlp_train, dim3(blocks), dim3(256), 0, m->stream, M, m->X, m->T, m->N, batch,
                           seed + (unsigned)m->step * 7919u, 0);
        m->step++;
        const float bc1 = 1.f - std::pow(m->b1, (float)m->step), bc2 = 1.f - std::pow(m->b2, (float)m->step);
        hipLaunchKernelGGL(k_mlp_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, m->stream, M, lr, m->b1, m->b2, m->eps,
                           bc1, bc2, 1);
        if (!odx_check(hipGetLastError(), "mlp train launch")) return 0;
        if (losses) {
            if (!odx_check(hipMemcpyAsync(l.data(), M.loss, 8, hipMemcpyDeviceToHost, m->stream), "hipMemcpyAsync") ||
                !odx_check(hipStreamSynchronize(m->stream), "hipStreamSynchronize"))
                return 0;
            losses[s] = l[0] / (float)batch;
        }
    }
    return odx_check(hipStreamSynchronize(m->stream), "hipStreamSynchronize") ? nsteps : 0;
}

// Forward pass on host features X [N][16] -> lambda [N][3] (force units).
int odxMlpPredict(odxMlp *m, const float *X, int N, float *Y) {
    if (!m || !X || !Y || N <= 0) return 0;
    float *dX = nullptr, *dY = nullptr;
    int *drow = nullptr, *dcnt = nullptr;
    bool ok = odx_check(hipMalloc((void **)&dX, (size_t)N * NIN * 4), "hipMalloc") &&
              odx_check(hipMalloc((void **)&dY, (size_t)N * NOUT * 4), "hipMalloc") &&
              odx_check(hipMalloc((void **)&drow, (size_t)N * 4), "hipMalloc") &&
              odx_check(hipMalloc((void **)&dcnt, 4), "hipMalloc");
    std::vector<int> rows(N);
    for (int i = 0; i < N; i++) rows[i] = NOUT * i;
    ok = ok && odx_check(hipMemcpy(dX, X, (size_t)N * NIN * 4, hipMemcpyHostToDevice), "hipMemcpy") &&
         odx_check(hipMemcpy(drow, rows.data(), (size_t)N * 4, hipMemcpyHostToDevice), "hipMemcpy") &&
         odx_check(hipMemcpy(dcnt, &N, 4, hipMemcpyHostToDevice), "hipMemcpy");
    if (ok) {
        launch_mlp_infer(m, dX, drow, dcnt, dY, N, m->stream, 4096);
        ok = odx_check(hipGetLastError(), "k_mlp_infer") && odx_check(hipStreamSynchronize(m->stream), "sync") &&
             odx_check(hipMemcpy(Y, dY, (size_t)N * NOUT * 4, hipMemcpyDeviceToHost), "hipMemcpy");
    }
    if (dX) (void)hipFree(dX);
    if (dY) (void)hipFree(dY);
    if (drow) (void)hipFree(drow);
    if (dcnt) (void)hipFree(dcnt);
    return ok ? N : 0;
}

}  // extern "C"

// ---- odxMlpPredictNoAlloc: the forward pass of odxMlpPredict on a fixed
// scratch array per device instead of memory allocated per call.
// At most MLP_NOALLOC_MAX contacts: features (1 MiB), impulses (192 KiB), row
// indices (64 KiB) and the count, 1.25 MiB held by each device the library
// runs on until the process ends.  Every call on a device shares its array:
// calls that overlap on one device race.  Each call writes every entry it
// reads (the row indices and the count by k_mlp_rows, the features by the
// copy), so nothing has to start at zero.
constexpr int MLP_NOALLOC_MAX = 16384;
constexpr size_t NA_X = 0;
constexpr size_t NA_Y = NA_X + (size_t)MLP_NOALLOC_MAX * NIN * sizeof(float);
constexpr size_t NA_ROW = NA_Y + (size_t)MLP_NOALLOC_MAX * NOUT * sizeof(float);
constexpr size_t NA_CNT = NA_ROW + (size_t)MLP_NOALLOC_MAX * sizeof(int);
constexpr size_t NA_BYTES = NA_CNT + 256;
static __device__ __align__(256) unsigned char scratch[NA_BYTES];

// identity row map (row i writes impulses 3i..3i+2) and the entry count
__global__ void k_mlp_rows(int *row, int *cnt, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) row[i] = NOUT * i;
    if (i == 0) *cnt = n;
}

namespace {
constexpr int NA_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_dev[NA_MAX_DEVICES] = {};

// The device address of `scratch` on the current device, looked up once per
// device (each device holds its own copy of the array); the array's name is
// a host-side stand-in and is never passed to a kernel or a copy.
unsigned char *scratch_on_current_device() {
    int dev = 0;
    if (!odx_check(hipGetDevice(&dev), "hipGetDevice")) return nullptr;
    if (dev < 0 || dev >= NA_MAX_DEVICES) {
        odx_message(4, "odxMlpPredictNoAlloc: device %d is beyond the %d this build tracks", dev, NA_MAX_DEVICES);
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        if (!odx_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress")) return nullptr;
        g_scratch_dev[dev] = static_cast<unsigned char *>(p);
    }
    return g_scratch_dev[dev];
}
}  // namespace

extern "C" {

int odxMlpPredictNoAlloc(odxMlp *m, const float *X, int N, float *Y) {
    if (!m || !X || !Y || N <= 0) return 0;
    if (N > MLP_NOALLOC_MAX) {
        odx_message(4, "odxMlpPredictNoAlloc: %d contacts exceed the fixed scratch of %d", N, MLP_NOALLOC_MAX);
        return 0;
    }
    unsigned char *base = scratch_on_current_device();
    if (!base) return 0;
    float *dX = reinterpret_cast<float *>(base + NA_X), *dY = reinterpret_cast<float *>(base + NA_Y);
    int *drow = reinterpret_cast<int *>(base + NA_ROW), *dcnt = reinterpret_cast<int *>(base + NA_CNT);
    bool ok = odx_check(hipMemcpyAsync(dX, X, (size_t)N * NIN * sizeof(float), hipMemcpyHostToDevice, m->stream),
                        "hipMemcpyAsync");
    if (ok) {
        hipLaunchKernelGGL(k_mlp_rows, dim3((N + 255) / 256), dim3(256), 0, m->stream, drow, dcnt, N);
        ok = odx_check(hipGetLastError(), "k_mlp_rows");
    }
    if (ok) {
        launch_mlp_infer(m, dX, drow, dcnt, dY, N, m->stream, 4096);
        ok = odx_check(hipGetLastError(), "k_mlp_infer") &&
             odx_check(hipMemcpyAsync(Y, dY, (size_t)N * NOUT * sizeof(float), hipMemcpyDeviceToHost, m->stream),
                       "hipMemcpyAsync") &&
             odx_check(hipStreamSynchronize(m->stream), "hipStreamSynchronize");
    }
    return ok ? N : 0;
}

}  // extern "C"
