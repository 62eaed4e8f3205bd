// odx_large.cpp -- host runtime of one large world (odxLarge* in odx.h).
//
// One world of up to millions of bodies: hash-space broadphase, narrowphase
// and rows one thread per item, then a cooperative colouring kernel and a
// cooperative solve kernel (grid barriers between colour phases).  Device
// counters report overflow; a step whose counters show dropped work or a
// failed barrier makes the next odxLargeSync fail.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/odx.h"
#include "odx_device.h"
#include "odx_host.h"
#include "odx_large.h"

namespace odx {
void lg_enqueue_prepare(const DevWorlds &D, const StepParams &P, const LargeDev &L, hipStream_t s);
bool lg_enqueue_colour(const DevWorlds &D, const LargeDev &L, int blocks, hipStream_t s);
bool lg_enqueue_solve(const DevWorlds &D, const StepParams &P, const LargeDev &L, int iters, int blocks,
                      hipStream_t s);
void lg_enqueue_fold(const LargeDev &L, int *sticky, hipStream_t s);
int lg_colour_blocks();
int lg_solve_blocks(int per_cu, int one_pass);
}  // namespace odx

using namespace odx;

struct odxLarge {
    int nb = 0, ng = 0, iters = 20;
    DevWorlds D{};
    StepParams P{};
    Surface surf{};
    Surface *surf_default = nullptr;  // device copy of the parameters' surface
    Surface *surf_table = nullptr;    // device material-pair table, owned (freed on replacement)
    std::vector<int> geom_mat;        // material of each geom (GEOM slot 13)
    LargeDev L{};
    int *sticky = nullptr;  // [LG_NCOUNTERS] largest value each counter reached since the last sync
    int colour_blocks = 0, solve_blocks = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    std::vector<void *> allocs;

    template <class T> bool alloc(T **p, size_t n) {
        *p = nullptr;
        if (n == 0) n = 1;
        if (!odx_check(hipMalloc((void **)p, n * sizeof(T)), "hipMalloc")) return false;
        allocs.push_back((void *)*p);
        return true;
    }
    ~odxLarge() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto &e : ev) if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
        for (void *p : allocs) (void)hipFree(p);
        if (surf_table) (void)hipFree(surf_table);
    }
    bool enqueue_step() {
        lg_enqueue_prepare(D, P, L, stream);
        if (!lg_enqueue_colour(D, L, colour_blocks, stream)) return odx_check(hipGetLastError(), "colour launch");
        if (!lg_enqueue_solve(D, P, L, iters, solve_blocks, stream))
            return odx_check(hipGetLastError(), "solve launch");
        lg_enqueue_fold(L, sticky, stream);
        return odx_check(hipGetLastError(), "large-world step launch");
    }
};

namespace {
constexpr int SCAN_BLOCK = 2048;  // entries per block of the device scans

float geom_diameter(const float *r) {
    switch ((int)r[0]) {
    case 0: return 2.f * r[2];                                                  // sphere r
    case 1: return std::sqrt(r[2] * r[2] + r[3] * r[3] + r[4] * r[4]);         // box sides
    case 2: return r[3] + 2.f * r[2];                                           // capsule r, length
    default: return INFINITY;                                                   // plane
    }
}
}  // namespace

extern "C" {

odxLarge *odxLargeCreate(int nb, const float *bodies, int ng, const float *geoms, const double *params,
                         int pair_cap, int contact_cap) {
    if (nb <= 0 || ng <= 0 || !bodies || !geoms || !params || nb >= (1 << 24) || ng >= (1 << 24)) {
        odx_message(1, "odxLargeCreate: bad arguments");
        return nullptr;
    }
    // geom records: class and body must be valid before any kernel reads them
    std::vector<int> isbig(ng, 0), big;
    float cell = 0.f;
    for (int g = 0; g < ng; g++) {
        const float *r = geoms + (size_t)g * GEOM_STRIDE;
        const int cls = (int)r[0], gb = (int)r[1];
        if (gb < -1 || gb >= nb || (cls < 0 || cls > 4)) {
            odx_message(2, "odxLargeCreate: geom %d has class %d / body %d", g, cls, gb);
            return nullptr;
        }
        const float d = geom_diameter(r);
        if (!(d < INFINITY)) {
            isbig[g] = 1;
            big.push_back(g);
        } else {
            cell = std::max(cell, d);
        }
    }
    if ((int)big.size() > 64) {
        odx_message(2, "odxLargeCreate: %d infinite geoms (at most 64)", (int)big.size());
        return nullptr;
    }
    if (pair_cap <= 0) pair_cap = 12 * ng + 1024;
    if (contact_cap <= 0) contact_cap = 8 * ng + 1024;
    if (contact_cap >= (1 << 24) || pair_cap >= (1 << 28)) {
        odx_message(2, "odxLargeCreate: capacities %d pairs / %d contacts too large", pair_cap, contact_cap);
        return nullptr;
    }
    const int row_order = (int)params[31];
    if (row_order != 0 && row_order != 2) {
        odx_message(2, "odxLargeCreate: row order %d (0 ReorderPrep's, 2 each contact's rows together)", row_order);
        return nullptr;
    }
    auto *W = new odxLarge();
    W->nb = nb;
    W->ng = ng;
    odx_params_from(params, W->P, W->surf);
    W->iters = std::max(1, W->P.max_iters);
    W->P.lds_rows = 0;
    DevWorlds &D = W->D;
    D.nworlds = 1;
    bool ok = odx_state_upload(D, nb, bodies, ng, geoms, W->allocs);
    Surface *sd = nullptr;
    ok = ok && W->alloc(&sd, 1) && odx_check(hipMemcpy(sd, &W->surf, sizeof(Surface), hipMemcpyHostToDevice), "H2D");
    D.surf = sd;
    W->surf_default = sd;
    W->geom_mat.resize(ng);
    for (int g = 0; g < ng; g++) W->geom_mat[g] = (int)geoms[(size_t)g * GEOM_STRIDE + 13];
    ok = ok && W->alloc(&D.bscr, (size_t)nb * BSCR_STRIDE);
    ok = ok && W->alloc(&D.joints, (size_t)contact_cap * JOINT_STRIDE);
    ok = ok && W->alloc(&D.rows, (size_t)3 * contact_cap * ROW_STRIDE) && W->alloc(&D.lambda, (size_t)3 * contact_cap);

    LargeDev &L = W->L;
    L.nb = nb;
    L.ng = ng;
    L.nbig = (int)big.size();
    int hsize = 1024;
    while (hsize < 2 * ng) hsize <<= 1;
    L.hmask = hsize - 1;
    L.spin_limit = (int)LG_SPIN_LIMIT;
    L.one_pass = row_order == 2;
    cell = cell > 0.f ? cell * 1.001f : 1.f;  // slack keeps floor() of overlapping centres within one cell
    L.inv_cell = 1.f / cell;
    L.pair_cap = pair_cap;
    L.contact_cap = contact_cap;
    int *ib = nullptr, *bg = nullptr;
    ok = ok && W->alloc(&ib, ng) &&
         odx_check(hipMemcpy(ib, isbig.data(), ng * sizeof(int), hipMemcpyHostToDevice), "H2D");
    L.isbig = ib;
    ok = ok && W->alloc(&bg, std::max<size_t>(1, big.size())) &&
         (big.empty() || odx_check(hipMemcpy(bg, big.data(), big.size() * sizeof(int), hipMemcpyHostToDevice), "H2D"));
    L.big = bg;
    ok = ok && W->alloc(&L.aabb, (size_t)ng * 6) && W->alloc(&L.gcell, ng) && W->alloc(&L.bucket, hsize + 1) &&
         W->alloc(&L.bucket_geom, ng) && W->alloc(&L.slot_cell, ng) && W->alloc(&L.slot_box, (size_t)ng * 2) &&
         W->alloc(&L.pair_cnt, (size_t)ng + 1) && W->alloc(&L.pairs, pair_cap) &&
         W->alloc(&L.pair_keep, (size_t)ng * LG_PAIR_KEEP) &&
         W->alloc(&L.contact_cnt, (size_t)pair_cap + 1) && W->alloc(&L.row_cnt, (size_t)contact_cap + 1) &&
         W->alloc(&L.counters, LG_NCOUNTERS) && W->alloc(&L.cbody, contact_cap) && W->alloc(&L.colour, contact_cap) && W->alloc(&L.taken, nb) &&
         W->alloc(&L.owner, nb) && W->alloc(&L.elig, LG_MAXCOL * LG_MAXSUB) && W->alloc(&L.col_start, LG_MAXCOL + 1) &&
         W->alloc(&L.corder, contact_cap) && W->alloc(&L.items, contact_cap) && W->alloc(&L.bar, LG_BAR_WORDS) && W->alloc(&L.fc, (size_t)nb * 8) && W->alloc(&L.ftot, (size_t)nb * 2);
    const size_t scan_cap = std::max({(size_t)hsize + 1, (size_t)ng + 1, (size_t)pair_cap + 1, (size_t)contact_cap + 1});
    ok = ok && W->alloc(&L.scan_tmp, (scan_cap + SCAN_BLOCK - 1) / SCAN_BLOCK);
    ok = ok && W->alloc(&W->sticky, LG_NCOUNTERS);
    if (ok) {
        ok = odx_check(hipMemset(L.counters, 0, LG_NCOUNTERS * sizeof(int)), "hipMemset") &&
             odx_check(hipMemset(W->sticky, 0, LG_NCOUNTERS * sizeof(int)), "hipMemset") &&
             odx_check(hipMemset(L.col_start, 0, (LG_MAXCOL + 1) * sizeof(int)), "hipMemset") &&
             odx_check(hipMemset(L.row_cnt, 0, sizeof(int)), "hipMemset");
    }
    ok = ok && odx_check(hipStreamCreate(&W->stream), "hipStreamCreate");
    for (auto &e : W->ev) ok = ok && odx_check(hipEventCreate(&e), "hipEventCreate");
    if (ok) {
        W->colour_blocks = lg_colour_blocks();
        W->solve_blocks = lg_solve_blocks(1, L.one_pass);
        L.colour_cap = W->solve_blocks * LG_BLOCK;
        if (W->colour_blocks <= 0 || W->solve_blocks <= 0) {
            odx_message(3, "odxLargeCreate: no cooperative launch configuration");
            ok = false;
        }
    }
    if (!ok) {
        delete W;
        return nullptr;
    }
    return W;
}

void odxLargeDestroy(odxLarge *w) { delete w; }

int odxLargeSetIterations(odxLarge *w, int iters) {
    if (!w || iters < 1) return 0;
    w->iters = iters;
    return 1;
}

int odxLargeSetSolveBlocks(odxLarge *w, int per_cu) {
    if (!w || per_cu < 1 || per_cu > 8) return 0;
    const int b = lg_solve_blocks(per_cu, w->L.one_pass);
    if (b <= 0) return 0;
    w->solve_blocks = b;
    w->L.colour_cap = b * LG_BLOCK;
    return b;
}

int odxLargeSetSpinLimit(odxLarge *w, int polls) {
    if (!w || polls < 0) return 0;
    w->L.spin_limit = polls == 0 ? 0 : (polls < (int)LG_SPIN_LIMIT ? polls : (int)LG_SPIN_LIMIT);
    return 1;
}

// The near-callback policy as data, as odxBatchSetSurfaceTable: the surface
// and contact limit of each material pair, and the category / collide bit
// test of collideAABBs (collision_space_internal.h:50-57) when
// category_filter is set.
int odxLargeSetSurfaceTable(odxLarge *w, int nmat, const float *table, int category_filter) {
    if (!w) return 0;
    std::vector<Surface> t;
    if (!odx_surface_table("odxLargeSetSurfaceTable", nmat, table, w->geom_mat, t)) return 0;
    if (!odx_check(hipStreamSynchronize(w->stream), "hipStreamSynchronize")) return 0;
    if (w->surf_table) { (void)hipFree(w->surf_table); w->surf_table = nullptr; }
    if (nmat > 0) {
        if (!odx_check(hipMalloc((void **)&w->surf_table, t.size() * sizeof(Surface)), "hipMalloc") ||
            !odx_check(hipMemcpy(w->surf_table, t.data(), t.size() * sizeof(Surface), hipMemcpyHostToDevice), "H2D"))
            return 0;
        w->D.surf = w->surf_table;
    } else {
        w->D.surf = w->surf_default;
    }
    w->P.nmat = nmat;
    w->P.cat_filter = category_filter ? 1 : 0;
    return 1;
}

int odxLargeStep(odxLarge *w, int nsteps) {
    if (!w || nsteps < 0) return 0;
    for (int i = 0; i < nsteps; i++)
        if (!w->enqueue_step()) return 0;
    return 1;
}

int odxLargeSync(odxLarge *w) {
    if (!w) return 0;
    if (!odx_check(hipStreamSynchronize(w->stream), "hipStreamSynchronize")) return 0;
    int st[LG_NCOUNTERS];
    if (!odx_check(hipMemcpy(st, w->sticky, sizeof st, hipMemcpyDeviceToHost), "D2H")) return 0;
    if (st[LG_BARRIER_ERR] || st[LG_PAIR_OVF] || st[LG_CONTACT_OVF] || st[LG_COLOUR_OVF]) {
        odx_message(4, "odxLarge: a step failed (barrier %d, dropped pairs %d, contacts %d, uncoloured %d)",
                    st[LG_BARRIER_ERR], st[LG_PAIR_OVF], st[LG_CONTACT_OVF], st[LG_COLOUR_OVF]);
        (void)hipMemset(w->sticky, 0, sizeof st);
        return 0;
    }
    return 1;
}

int odxLargeTime(odxLarge *w, int nsteps, float *total_ms) {
    if (!w || nsteps <= 0 || !total_ms) return 0;
    if (!odx_check(hipEventRecord(w->ev[0], w->stream), "hipEventRecord")) return 0;
    if (!odxLargeStep(w, nsteps)) return 0;
    if (!odx_check(hipEventRecord(w->ev[1], w->stream), "hipEventRecord")) return 0;
    if (!odxLargeSync(w)) return 0;
    return odx_check(hipEventElapsedTime(total_ms, w->ev[0], w->ev[1]), "hipEventElapsedTime");
}

int odxLargeKernelTimes(odxLarge *w, int nsteps, float *ms) {
    if (!w || nsteps <= 0 || !ms) return 0;
    double acc[3] = {0, 0, 0};
    for (int i = 0; i < nsteps; i++) {
        bool ok = odx_check(hipEventRecord(w->ev[0], w->stream), "hipEventRecord");
        lg_enqueue_prepare(w->D, w->P, w->L, w->stream);
        ok = ok && odx_check(hipEventRecord(w->ev[1], w->stream), "hipEventRecord");
        ok = ok && lg_enqueue_colour(w->D, w->L, w->colour_blocks, w->stream);
        ok = ok && odx_check(hipEventRecord(w->ev[2], w->stream), "hipEventRecord");
        ok = ok && lg_enqueue_solve(w->D, w->P, w->L, w->iters, w->solve_blocks, w->stream);
        ok = ok && odx_check(hipEventRecord(w->ev[3], w->stream), "hipEventRecord");
        lg_enqueue_fold(w->L, w->sticky, w->stream);
        if (!ok || !odxLargeSync(w)) return 0;
        float t[3];
        for (int k = 0; k < 3; k++)
            if (!odx_check(hipEventElapsedTime(&t[k], w->ev[k], w->ev[k + 1]), "hipEventElapsedTime")) return 0;
        for (int k = 0; k < 3; k++) acc[k] += t[k];
    }
    for (int k = 0; k < 3; k++) ms[k] = (float)(acc[k] / nsteps);
    return 1;
}

int odxLargeGetBodies(odxLarge *w, float *bodies) {
    if (!w || !bodies || !odx_check(hipStreamSynchronize(w->stream), "hipStreamSynchronize")) return 0;
    return odx_bodies_download(w->D, w->nb, bodies) ? w->nb : 0;
}

int odxLargeSetBodies(odxLarge *w, const float *bodies) {
    if (!w || !bodies || !odx_check(hipStreamSynchronize(w->stream), "hipStreamSynchronize")) return 0;
    return odx_bodies_upload(w->D, w->nb, bodies) ? w->nb : 0;
}

int odxLargeGetCounts(odxLarge *w, int *counts) {
    if (!w || !counts || !odx_check(hipStreamSynchronize(w->stream), "hipStreamSynchronize")) return 0;
    return odx_check(hipMemcpy(counts, w->L.counters, LG_NCOUNTERS * sizeof(int), hipMemcpyDeviceToHost), "D2H")
               ? LG_NCOUNTERS
               : 0;
}

static int large_count(odxLarge *w, int which) {
    int c[LG_NCOUNTERS];
    if (odxLargeGetCounts(w, c) != LG_NCOUNTERS) return -1;
    return c[which];
}

int odxLargeGetPairs(odxLarge *w, int *pairs, int cap) {
    const int n = w && pairs ? large_count(w, LG_PAIRS) : -1;
    if (n < 0) return 0;
    const int k = std::min(n, cap);
    if (k > 0 && !odx_check(hipMemcpy(pairs, w->L.pairs, (size_t)k * sizeof(int2), hipMemcpyDeviceToHost), "D2H"))
        return 0;
    return k;
}

int odxLargeGetJoints(odxLarge *w, float *joints, int cap) {
    const int n = w && joints ? large_count(w, LG_CONTACTS) : -1;
    if (n < 0) return 0;
    const int k = std::min(n, cap);
    if (k > 0 && !odx_check(hipMemcpy(joints, w->D.joints, (size_t)k * JOINT_STRIDE * sizeof(float),
                                      hipMemcpyDeviceToHost), "D2H"))
        return 0;
    return k;
}

int odxLargeGetRows(odxLarge *w, float *rows, int cap) {
    const int n = w && rows ? large_count(w, LG_ROWS) : -1;
    if (n < 0) return 0;
    const int k = std::min(n, cap);
    if (k > 0 && !odx_check(hipMemcpy(rows, w->D.rows, (size_t)k * ROW_STRIDE * sizeof(float),
                                      hipMemcpyDeviceToHost), "D2H"))
        return 0;
    return k;
}

int odxLargeGetLambda(odxLarge *w, float *lambda, int cap) {
    const int n = w && lambda ? large_count(w, LG_ROWS) : -1;
    if (n < 0) return 0;
    const int k = std::min(n, cap);
    if (k > 0 && !odx_check(hipMemcpy(lambda, w->D.lambda, (size_t)k * sizeof(float), hipMemcpyDeviceToHost), "D2H"))
        return 0;
    return k;
}

int odxLargeGetColours(odxLarge *w, int *colour, int cap) {
    const int n = w && colour ? large_count(w, LG_CONTACTS) : -1;
    if (n < 0) return 0;
    const int k = std::min(n, cap);
    if (k > 0 && !odx_check(hipMemcpy(colour, w->L.colour, (size_t)k * sizeof(int), hipMemcpyDeviceToHost), "D2H"))
        return 0;
    return k;
}

}  // extern "C"
