// odx_space.hip -- the broadphase behind the drop-in dSpaceCollide.
//
// One workgroup per call computes the AABBs of a space's enabled geoms and
// writes the candidate pairs in the order the reference's space reports them
// to the near callback:
//
//  * sweep and prune (dxSAPSpace::collide, collision_sapspace.cpp:423-535):
//    the finite geoms are sorted on the first axis exactly as the OPCODE radix
//    sort of :654-894 orders them, including the ranks it keeps from the
//    previous call (when the keys read in that order are still non-decreasing
//    the previous order is returned unchanged), then pruned as BoxPruning
//    (:560-621) does, one sorted position per thread;
//  * simple space (collision_space.cpp:231-256): the n^2 loop over the list;
//  * hash space (collision_space.cpp:421-614): each geom's level and
//    discretised cell bounds as the reference computes them; every pair of
//    overlapping AABBs (found by a sweep on x) belongs to the geom that meets
//    it first in the reference's walk -- the lower-level geom, or the earlier
//    one in the list at equal levels -- and is ordered there by the level,
//    then the first shared cell in x, y, z order, then the newest hash-chain
//    node first.
//
// The pairs with infinite (SAP) or big (hash) AABBs are appended by the host.
// One call handles at most SP_MAX enabled geoms; the sort runs in LDS.
#include <hip/hip_runtime.h>
#include <climits>

#include "odx_collide.h"
#include "odx_device.h"

namespace odx {

constexpr int SP_MAX = 8192;
constexpr int SP_THREADS = 1024;


__device__ __forceinline__ unsigned sp_key(float f) {
    // IEEE bits to an order-preserving unsigned key (negatives reversed)
    const unsigned u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ bool sp_overlap(const float *x, const float *y) {
    return !(x[0] > y[1] || x[1] < y[0] || x[2] > y[3] || x[3] < y[2] || x[4] > y[5] || x[5] < y[4]);
}

// same non-null body; category / collide bits (collideAABBs,
// collision_space_internal.h:50-57)
__device__ __forceinline__ bool sp_filters(const SpaceCall &C, int a, int b) {
    const int ba = (int)C.gviews[(size_t)a * GVIEW_STRIDE + 1], bb = (int)C.gviews[(size_t)b * GVIEW_STRIDE + 1];
    if (ba == bb && ba >= 0) return false;
    const unsigned long long ca = C.bits[2 * a], la = C.bits[2 * a + 1], cb = C.bits[2 * b], lb = C.bits[2 * b + 1];
    return ((ca & lb) | (cb & la)) != 0ull;
}

// Bitonic sort of k[0, P) in LDS (P a power of two), ascending.
__device__ void sp_bitonic(unsigned long long *k, int P) {
    for (int size = 2; size <= P; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = threadIdx.x; t < (P >> 1); t += blockDim.x) {
                const int i = 2 * t - (t & (stride - 1)), j = i + stride;
                const bool up = (i & size) == 0;
                const unsigned long long a = k[i], b = k[j];
                if ((a > b) == up) { k[i] = b; k[j] = a; }
            }
            __syncthreads();
        }
}

// Up to this many keys a rank sort beats the bitonic network (n^2 / 1024
// comparisons per thread, no barrier per stage).
constexpr int SP_RANK_MAX = 4096;

// Ranks of the n distinct keys k[0, n): r[p] = #{q : k[q] < k[p]}, the
// position of key p in ascending order.  Every lane of a wave reads the same
// k[q] at once (an LDS broadcast).
__device__ void sp_rank(const unsigned long long *k, int n, int *r) {
    for (int p = threadIdx.x; p < n; p += blockDim.x) {
        const unsigned long long kp = k[p];
        int c = 0;
        int q = 0;
        for (; q + 4 <= n; q += 4)
            c += (k[q] < kp ? 1 : 0) + (k[q + 1] < kp ? 1 : 0) + (k[q + 2] < kp ? 1 : 0) + (k[q + 3] < kp ? 1 : 0);
        for (; q < n; q++) c += k[q] < kp ? 1 : 0;
        r[p] = c;
    }
    __syncthreads();
}

// Exclusive scan of a[0, rows) in place; a[rows] = total, which is returned.
__device__ int sp_scan(int *a, int rows, int *tmp) {
    const int t = threadIdx.x, T = blockDim.x;
    const int per = (rows + T - 1) / T, b = t * per, e = min(rows, b + per);
    int s = 0;
    for (int i = b; i < e; i++) s += a[i];
    tmp[t] = s;
    __syncthreads();
    for (int off = 1; off < T; off <<= 1) {
        const int v = t >= off ? tmp[t - off] : 0;
        __syncthreads();
        tmp[t] += v;
        __syncthreads();
    }
    int run = tmp[t] - s;
    for (int i = b; i < e; i++) { const int c = a[i]; a[i] = run; run += c; }
    const int total = tmp[T - 1];
    __syncthreads();
    if (t == 0) a[rows] = total;
    __syncthreads();
    return total;
}

// BoxPruning for the geom at sorted position k: the later ones while their
// first-axis minimum is <= its maximum, with the two other axes overlapping.
__device__ int sap_row(const SpaceCall &C, const unsigned short *sorted, int k, int *out) {
    const int a = sorted[k];
    const float *A = C.aabb + (size_t)a * 6;
    const float x0max = A[C.ax0 + 1], x1min = A[C.ax1], x1max = A[C.ax1 + 1], x2min = A[C.ax2], x2max = A[C.ax2 + 1];
    int c = 0;
    for (int q = k + 1; q < C.nsort; q++) {
        const int b = sorted[q];
        const float *B = C.aabb + (size_t)b * 6;
        if (!(B[C.ax0] <= x0max)) break;
        if (x1max >= B[C.ax1] && B[C.ax1 + 1] >= x1min && x2max >= B[C.ax2] && B[C.ax2 + 1] >= x2min &&
            sp_filters(C, a, b)) {
            if (out) { out[2 * c] = a; out[2 * c + 1] = b; }
            c++;
        }
    }
    return c;
}

// sap_row over AABBs staged in LDS in sorted order (sab[6 q .. 6 q + 5] is
// the AABB of the geom at sorted position q).
__device__ int sap_row_lds(const SpaceCall &C, const unsigned short *sorted, const float *sab, int k, int *out) {
    const int a = sorted[k];
    const float *A = sab + (size_t)k * 6;
    const float x0max = A[C.ax0 + 1], x1min = A[C.ax1], x1max = A[C.ax1 + 1], x2min = A[C.ax2], x2max = A[C.ax2 + 1];
    int c = 0;
    for (int q = k + 1; q < C.nsort; q++) {
        const float *B = sab + (size_t)q * 6;
        if (!(B[C.ax0] <= x0max)) break;
        if (x1max >= B[C.ax1] && B[C.ax1 + 1] >= x1min && x2max >= B[C.ax2] && B[C.ax2 + 1] >= x2min) {
            const int b = sorted[q];
            if (sp_filters(C, a, b)) {
                if (out) { out[2 * c] = a; out[2 * c + 1] = b; }
                c++;
            }
        }
    }
    return c;
}

// The simple space's row of geom i: the later geoms of the list.
__device__ int simple_row(const SpaceCall &C, int i, int *out) {
    const float *A = C.aabb + (size_t)i * 6;
    int c = 0;
    for (int j = i + 1; j < C.n; j++) {
        if (sp_overlap(A, C.aabb + (size_t)j * 6) && sp_filters(C, i, j)) {
            if (out) { out[2 * c] = i; out[2 * c + 1] = j; }
            c++;
        }
    }
    return c;
}

// Level and cell bounds of a hash-space geom (findLevel :329-350 and the
// discretisation of dxHashSpace::collide :451-471).
__device__ void hash_cells(const SpaceCall &C, const float *ab, int *o) {
    for (int k = 0; k < 8; k++) o[k] = 0;
    for (int k = 0; k < 6; k += 2)
        if (ab[k] <= -kInf || ab[k + 1] >= kInf) { o[0] = INT_MAX; return; }
    float q = ab[1] - ab[0], q2 = ab[3] - ab[2];
    if (q2 > q) q = q2;
    q2 = ab[5] - ab[4];
    if (q2 > q) q = q2;
    int level;
    frexpf(q, &level);
    if (level < C.minlevel) level = C.minlevel;
    if (level > C.maxlevel) { o[0] = INT_MAX; return; }
    o[0] = level;
    const float recip = 1.0f / ldexpf(1.0f, level);
    for (int k = 0; k < 6; k++) o[1 + k] = (int)floorf(ab[k] * recip);
}

__device__ __forceinline__ int sp_shr(int v, int s) { return s > 31 ? (v < 0 ? -1 : 0) : (v >> s); }

// Order of partner p in searcher s's walk: level, first shared cell (x, y, z),
// then the newer node (later in the list) first.  Returns true if p comes
// before p2.
__device__ bool hash_before(const int *cs, const int *cp, int p, const int *cp2, int p2) {
    const int L1 = cp[0], L2 = cp2[0];
    if (L1 != L2) return L1 < L2;
    const int sh = L1 - cs[0];
    for (int ax = 0; ax < 3; ax++) {
        const int sv = sp_shr(cs[1 + 2 * ax], sh);
        const int v1 = max(sv, cp[1 + 2 * ax]), v2 = max(sv, cp2[1 + 2 * ax]);
        if (v1 != v2) return v1 < v2;
    }
    return p > p2;
}

__global__ __launch_bounds__(SP_THREADS) void k_space_collide(SpaceCall C) {
    __shared__ unsigned long long s_key[SP_MAX];
    __shared__ unsigned short s_init[SP_MAX];
    __shared__ unsigned short s_sorted[SP_MAX];
    __shared__ int s_cnt[SP_MAX + 1];
    __shared__ int s_tmp[SP_THREADS];
    __shared__ int s_flag, s_nh;
    const int t = threadIdx.x, T = blockDim.x, n = C.n;

    // AABBs, and the hash levels and cells
    for (int i = t; i < n; i += T) {
        DGeom G;
        load_gview(C.gviews + (size_t)i * GVIEW_STRIDE, G);
        float ab[6];
        geom_aabb(G, ab);
        for (int k = 0; k < 6; k++) C.aabb[(size_t)i * 6 + k] = ab[k];
        if (C.kind == 1) hash_cells(C, ab, C.cells + (size_t)i * 8);
    }
    if (t == 0) { s_flag = 0; s_nh = 0; }
    __syncthreads();

    int rows = n;
    if (C.kind == 0) {
        const int ns = C.nsort;
        rows = ns;
        // (ranks outside [0, ns) would be a host error; they are clamped)
        for (int p = t; p < ns; p += T) s_init[p] = (unsigned short)(C.ranks_valid ? min(C.ranks_in[p], (unsigned)(ns - 1)) : (unsigned)p);
        __syncthreads();
        for (int p = t; p + 1 < ns; p += T)
            if (C.aabb[(size_t)s_init[p + 1] * 6 + C.ax0] < C.aabb[(size_t)s_init[p] * 6 + C.ax0]) s_flag = 1;
        __syncthreads();
        if (s_flag == 0) {
            // already in order: the previous ranks (or the index order) stand
            for (int p = t; p < ns; p += T) { s_sorted[p] = s_init[p]; C.ranks_out[p] = s_init[p]; }
            if (t == 0) C.info[1] = C.ranks_valid;
        } else {
            // a stable sort of the initial order by key, ties of negative keys
            // reversed (the sign pass of the radix sort fills them from the end)
            int P = 1;
            while (P < ns) P <<= 1;
            for (int p = t; p < P; p += T) {
                unsigned long long k = ~0ull;
                if (p < ns) {
                    const float f = C.aabb[(size_t)s_init[p] * 6 + C.ax0];
                    const unsigned tie = (__float_as_uint(f) & 0x80000000u) ? (unsigned)(ns - 1 - p) : (unsigned)p;
                    k = ((unsigned long long)sp_key(f) << 32) | tie;
                }
                s_key[p] = k;
            }
            __syncthreads();
            if (ns <= SP_RANK_MAX) {
                // the keys are distinct (the tie field): key p goes to its rank
                sp_rank(s_key, ns, s_cnt);
                for (int p = t; p < ns; p += T) {
                    const int q = s_cnt[p];
                    s_sorted[q] = s_init[p];
                    C.ranks_out[q] = s_init[p];
                }
                __syncthreads();
            } else {
                sp_bitonic(s_key, P);
                for (int q = t; q < ns; q += T) {
                    const unsigned long long k = s_key[q];
                    const unsigned tie = (unsigned)k;
                    const int p = ((unsigned)(k >> 32) & 0x80000000u) ? (int)tie : ns - 1 - (int)tie;
                    s_sorted[q] = s_init[p];
                    C.ranks_out[q] = s_init[p];
                }
            }
            if (t == 0) C.info[1] = 1;
        }
    } else if (C.kind == 1) {
        int P = 1;
        while (P < n) P <<= 1;
        for (int p = t; p < P; p += T) {
            unsigned long long k = ~0ull;
            if (p < n && C.cells[(size_t)p * 8] != INT_MAX) {
                k = ((unsigned long long)sp_key(C.aabb[(size_t)p * 6]) << 32) | (unsigned)p;
                atomicAdd(&s_nh, 1);
            }
            s_key[p] = k;
        }
        __syncthreads();
        if (n <= SP_RANK_MAX) {
            // hashed geoms' keys are distinct (the geom index); the others are
            // all ones and rank after them
            sp_rank(s_key, n, s_cnt);
            for (int p = t; p < n; p += T)
                if (s_key[p] != ~0ull) s_sorted[s_cnt[p]] = (unsigned short)p;
            __syncthreads();
        } else {
            sp_bitonic(s_key, P);
            for (int q = t; q < s_nh; q += T) s_sorted[q] = (unsigned short)(s_key[q] & 0xffffu);
        }
    }
    // sweep and prune: the AABBs in sorted order into LDS (over the sort keys,
    // no longer needed) when they fit
    float *s_ab = reinterpret_cast<float *>(s_key);
    const bool lds_ab = C.kind == 0 && rows * 6 <= SP_MAX * 2;
    if (lds_ab) {
        __syncthreads();
        for (int q = t; q < rows * 6; q += T) s_ab[q] = C.aabb[(size_t)s_sorted[q / 6] * 6 + q % 6];
    }
    for (int i = t; i <= rows; i += T) s_cnt[i] = 0;
    __syncthreads();

    // count the pairs of every row
    const int nh = s_nh;
    if (C.kind == 0) {
        for (int k = t; k < rows; k += T)
            s_cnt[k] = lds_ab ? sap_row_lds(C, s_sorted, s_ab, k, nullptr) : sap_row(C, s_sorted, k, nullptr);
    } else if (C.kind == 2) {
        for (int i = t; i < rows; i += T) s_cnt[i] = simple_row(C, i, nullptr);
    } else {
        for (int k = t; k < nh; k += T) {
            const int a = s_sorted[k];
            const float *A = C.aabb + (size_t)a * 6;
            for (int q = k + 1; q < nh; q++) {
                const int b = s_sorted[q];
                const float *B = C.aabb + (size_t)b * 6;
                if (B[0] > A[1]) break;
                if (!sp_overlap(A, B) || !sp_filters(C, a, b)) continue;
                const int la = C.cells[(size_t)a * 8], lb = C.cells[(size_t)b * 8];
                const int s = (la < lb || (la == lb && a < b)) ? a : b;
                atomicAdd(&s_cnt[s], 1);
            }
        }
    }
    __syncthreads();
    const int total = sp_scan(s_cnt, rows, s_tmp);
    if (t == 0) C.info[0] = total;
    if (total > C.cap) return;  // the host grows the buffer and calls again

    // write them at their offsets
    if (C.kind == 0) {
        for (int k = t; k < rows; k += T) {
            int *o = C.pairs + 2 * (size_t)s_cnt[k];
            if (lds_ab) sap_row_lds(C, s_sorted, s_ab, k, o);
            else sap_row(C, s_sorted, k, o);
        }
    } else if (C.kind == 2) {
        for (int i = t; i < rows; i += T) simple_row(C, i, C.pairs + 2 * (size_t)s_cnt[i]);
    } else {
        int *fill = reinterpret_cast<int *>(s_key);
        for (int i = t; i < n; i += T) fill[i] = 0;
        __syncthreads();
        for (int k = t; k < nh; k += T) {
            const int a = s_sorted[k];
            const float *A = C.aabb + (size_t)a * 6;
            for (int q = k + 1; q < nh; q++) {
                const int b = s_sorted[q];
                const float *B = C.aabb + (size_t)b * 6;
                if (B[0] > A[1]) break;
                if (!sp_overlap(A, B) || !sp_filters(C, a, b)) continue;
                const int la = C.cells[(size_t)a * 8], lb = C.cells[(size_t)b * 8];
                const bool a_first = la < lb || (la == lb && a < b);
                const int s = a_first ? a : b, p = a_first ? b : a;
                const int slot = s_cnt[s] + atomicAdd(&fill[s], 1);
                C.pairs[2 * (size_t)slot] = s;
                C.pairs[2 * (size_t)slot + 1] = p;
            }
        }
        __syncthreads();
        // each searcher orders its partners (insertion sort: a few per geom)
        for (int s = t; s < n; s += T) {
            const int b0 = s_cnt[s], e = s_cnt[s + 1];
            if (e - b0 < 2) continue;
            const int *cs = C.cells + (size_t)s * 8;
            for (int i = b0 + 1; i < e; i++) {
                const int p = C.pairs[2 * (size_t)i + 1];
                const int *cp = C.cells + (size_t)p * 8;
                int j = i - 1;
                while (j >= b0) {
                    const int pj = C.pairs[2 * (size_t)j + 1];
                    if (!hash_before(cs, cp, p, C.cells + (size_t)pj * 8, pj)) break;
                    C.pairs[2 * (size_t)(j + 1) + 1] = pj;
                    j--;
                }
                C.pairs[2 * (size_t)(j + 1) + 1] = p;
            }
        }
    }
}

int space_max_geoms() { return SP_MAX; }

void launch_space_collide(const SpaceCall &C, hipStream_t s) {
    hipLaunchKernelGGL(k_space_collide, dim3(1), dim3(SP_THREADS), 0, s, C);
}

}  // namespace odx
