// odx_batch.cpp -- host runtime of the batched step path (odxBatch* in odx.h).
//
// Owns the device SoA state of many independent worlds, builds per-world
// capacities, launches the prepare/solve kernels on one stream and replays the
// step as a hipGraph.  Reports errors through the message handler; nothing
// here throws across the C ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/odx.h"
#include "odx_device.h"
#include "odx_host.h"

namespace odx {
struct MlpDev;
void launch_prepare_small(const DevWorlds &D, const StepParams &P, hipStream_t s);
void launch_solve_small(const DevWorlds &D, const StepParams &P, hipStream_t s);
void launch_step_fused(const DevWorlds &D, const StepParams &P, const MlpDev *M, int nsteps, hipStream_t s);
void launch_mlp_infer(const odxMlp *m, const float *feat, const int *feat_row, const int *count, float *lambda0,
                      int max_entries, hipStream_t s, int max_blocks);
const MlpDev *mlp_device_view(const odxMlp *m);
int mlp_gate(const odxMlp *m);
}  // namespace odx

using namespace odx;

namespace {
odxMessageFunction *g_handler = nullptr;
}

void odx_message(int err, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (g_handler) g_handler(err, buf);
    else fprintf(stderr, "odx: %s\n", buf);
}

bool odx_check(hipError_t e, const char *what) {
    if (e == hipSuccess) return true;
    odx_message(int(e), "%s failed: %s", what, hipGetErrorString(e));
    return false;
}

// Body-frame inverse inertia by Cholesky factorisation, as dInvertPDMatrix
// (matrix.cpp:563) does for dBodySetMass (ode.cpp:488-504).  Falls back to the
// identity for a non positive-definite tensor, as dBodySetMass does.
void odx_invert_pd3(const float *I, float *out) {
    float a[3][3] = {{I[0], I[1], I[2]}, {I[4], I[5], I[6]}, {I[8], I[9], I[10]}};
    float L[3][3] = {{0.f}}, Li[3][3] = {{0.f}};
    bool ok = true;
    for (int i = 0; i < 3 && ok; i++) {
        for (int j = 0; j <= i; j++) {
            float s = a[i][j];
            for (int k = 0; k < j; k++) s -= L[i][k] * L[j][k];
            if (i == j) {
                if (s <= 0.f) { ok = false; break; }
                L[i][i] = std::sqrt(s);
            } else {
                L[i][j] = s / L[j][j];
            }
        }
    }
    std::memset(out, 0, 12 * sizeof(float));
    if (!ok) { out[0] = out[5] = out[10] = 1.f; return; }
    for (int i = 0; i < 3; i++) {
        Li[i][i] = 1.f / L[i][i];
        for (int j = 0; j < i; j++) {
            float s = 0.f;
            for (int k = j; k < i; k++) s += L[i][k] * Li[k][j];
            Li[i][j] = -s / L[i][i];
        }
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            float s = 0.f;
            for (int k = 0; k < 3; k++) s += Li[k][i] * Li[k][j];
            out[i * 4 + j] = s;
        }
}

template <class T> static bool dalloc(T **p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    return odx_check(hipMalloc((void **)p, n * sizeof(T)), "hipMalloc");
}
template <class T> static bool upload(T *d, const T *h, size_t n) {
    if (n == 0) return true;
    return odx_check(hipMemcpy(d, h, n * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D");
}
template <class T> static bool download(T *h, const T *d, size_t n) {
    if (n == 0) return true;
    return odx_check(hipMemcpy(h, d, n * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy D2H");
}

// Body and geom records -> device SoA of D (positions, velocities, mass,
// inertia and its inverse, geom parameters and static poses); every
// allocation is appended to allocs.
bool odx_state_upload(DevWorlds &D, int nb, const float *bodies, int ng, const float *geoms,
                      std::vector<void *> &allocs) {
    auto alloc = [&](auto **p, size_t n) {
        if (!dalloc(p, n)) return false;
        allocs.push_back((void *)*p);
        return true;
    };
    std::vector<float4> pos(nb), quat(nb), lvel(nb), avel(nb), facc(nb), tacc(nb), mass(nb);
    std::vector<float> Ib((size_t)nb * 12), invIb((size_t)nb * 12);
    for (int i = 0; i < nb; i++) {
        const float *r = bodies + (size_t)i * BODY_STRIDE;
        pos[i] = make_float4(r[0], r[1], r[2], 0.f);
        float q[4] = {r[4], r[5], r[6], r[7]};
        float l = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
        if (l > 0.f) { l = 1.f / std::sqrt(l); for (float &x : q) x *= l; }
        else { q[0] = 1.f; q[1] = q[2] = q[3] = 0.f; }
        quat[i] = make_float4(q[0], q[1], q[2], q[3]);
        lvel[i] = make_float4(r[8], r[9], r[10], 0.f);
        avel[i] = make_float4(r[12], r[13], r[14], 0.f);
        facc[i] = make_float4(r[16], r[17], r[18], 0.f);
        tacc[i] = make_float4(r[20], r[21], r[22], 0.f);
        mass[i] = make_float4(r[24], 1.f / r[24], r[31], 0.f);
        float *I = &Ib[(size_t)i * 12];
        I[0] = r[25]; I[5] = r[26]; I[10] = r[27];
        I[1] = I[4] = r[28]; I[2] = I[8] = r[29]; I[6] = I[9] = r[30];
        I[3] = I[7] = I[11] = 0.f;
        odx_invert_pd3(I, &invIb[(size_t)i * 12]);
    }
    std::vector<float4> g0(ng), g1(ng), gp(ng), gq(ng);
    for (int i = 0; i < ng; i++) {
        const float *r = geoms + (size_t)i * GEOM_STRIDE;
        g0[i] = make_float4(r[0], r[1], r[2], r[3]);
        g1[i] = make_float4(r[4], r[5], r[13], r[14]);  // material, category bits
        gp[i] = make_float4(r[6], r[7], r[8], r[15]);   // collide bits
        float q[4] = {r[9], r[10], r[11], r[12]};
        float l = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
        if (l > 0.f) { l = 1.f / std::sqrt(l); for (float &x : q) x *= l; }
        else { q[0] = 1.f; q[1] = q[2] = q[3] = 0.f; }
        gq[i] = make_float4(q[0], q[1], q[2], q[3]);
    }
    bool ok = alloc(&D.pos, nb) && upload(D.pos, pos.data(), nb);
    ok = ok && alloc(&D.quat, nb) && upload(D.quat, quat.data(), nb);
    ok = ok && alloc(&D.lvel, nb) && upload(D.lvel, lvel.data(), nb);
    ok = ok && alloc(&D.avel, nb) && upload(D.avel, avel.data(), nb);
    ok = ok && alloc(&D.facc, nb) && upload(D.facc, facc.data(), nb);
    ok = ok && alloc(&D.tacc, nb) && upload(D.tacc, tacc.data(), nb);
    float4 *m4 = nullptr;
    ok = ok && alloc(&m4, nb) && upload(m4, mass.data(), nb);
    D.mass = m4;
    float *f = nullptr;
    ok = ok && alloc(&f, (size_t)nb * 12) && upload(f, invIb.data(), (size_t)nb * 12);
    D.invIb = f;
    ok = ok && alloc(&f, (size_t)nb * 12) && upload(f, Ib.data(), (size_t)nb * 12);
    D.Ib = f;
    float4 *g4 = nullptr;
    ok = ok && alloc(&g4, ng) && upload(g4, g0.data(), ng);
    D.geom0 = g4;
    ok = ok && alloc(&g4, ng) && upload(g4, g1.data(), ng);
    D.geom1 = g4;
    ok = ok && alloc(&g4, ng) && upload(g4, gp.data(), ng);
    D.gpos = g4;
    ok = ok && alloc(&g4, ng) && upload(g4, gq.data(), ng);
    D.gquat = g4;
    // damping / angular speed cap of bodies whose flags ask for them (record
    // slots 3, 11, 15, 19, 23; thresholds squared as dBodySet*DampingThreshold)
    std::vector<float4> damp;
    for (int i = 0; i < nb; i++) {
        const float *r = bodies + (size_t)i * BODY_STRIDE;
        const int bits = ((int)r[31] >> 2) & 7;
        if (!bits) continue;
        if (damp.empty()) damp.assign((size_t)nb * 2, make_float4(0.f, 0.f, 0.f, 0.f));
        damp[2 * i] = make_float4(r[3], r[11], r[15] * r[15], r[19] * r[19]);
        damp[2 * i + 1] = make_float4(r[23], (float)bits, r[15], r[19]);
    }
    D.damp = nullptr;
    if (ok && !damp.empty()) {
        ok = alloc(&g4, damp.size()) && upload(g4, damp.data(), damp.size());
        D.damp = g4;
    }
    return ok;
}

// device body state -> body records [nb][32]
bool odx_bodies_download(const DevWorlds &D, int nb, float *out) {
    std::vector<float4> pos(nb), quat(nb), lvel(nb), avel(nb), facc(nb), tacc(nb), mass(nb);
    std::vector<float> Ib((size_t)nb * 12);
    bool ok = download(pos.data(), D.pos, nb) && download(quat.data(), D.quat, nb) &&
              download(lvel.data(), D.lvel, nb) && download(avel.data(), D.avel, nb) &&
              download(facc.data(), D.facc, nb) && download(tacc.data(), D.tacc, nb) &&
              download(mass.data(), (float4 *)D.mass, nb) && download(Ib.data(), (float *)D.Ib, (size_t)nb * 12);
    if (!ok) return false;
    for (int i = 0; i < nb; i++) {
        float *r = out + (size_t)i * BODY_STRIDE;
        std::memset(r, 0, BODY_STRIDE * sizeof(float));
        r[0] = pos[i].x; r[1] = pos[i].y; r[2] = pos[i].z;
        r[4] = quat[i].x; r[5] = quat[i].y; r[6] = quat[i].z; r[7] = quat[i].w;
        r[8] = lvel[i].x; r[9] = lvel[i].y; r[10] = lvel[i].z;
        r[12] = avel[i].x; r[13] = avel[i].y; r[14] = avel[i].z;
        r[16] = facc[i].x; r[17] = facc[i].y; r[18] = facc[i].z;
        r[20] = tacc[i].x; r[21] = tacc[i].y; r[22] = tacc[i].z;
        r[24] = mass[i].x;
        const float *I = &Ib[(size_t)i * 12];
        r[25] = I[0]; r[26] = I[5]; r[27] = I[10]; r[28] = I[1]; r[29] = I[2]; r[30] = I[6];
        r[31] = mass[i].z;
    }
    if (D.damp) {
        std::vector<float4> damp((size_t)nb * 2);
        if (!download(damp.data(), (float4 *)D.damp, damp.size())) return false;
        for (int i = 0; i < nb; i++) {
            float *r = out + (size_t)i * BODY_STRIDE;
            if (!((int)damp[2 * i + 1].y)) continue;
            r[3] = damp[2 * i].x; r[11] = damp[2 * i].y;
            r[15] = damp[2 * i + 1].z; r[19] = damp[2 * i + 1].w; r[23] = damp[2 * i + 1].x;
        }
    }
    return true;
}

// body records -> device pose, velocities and accumulators (mass unchanged)
bool odx_bodies_upload(const DevWorlds &D, int nb, const float *bodies) {
    std::vector<float4> pos(nb), quat(nb), lvel(nb), avel(nb), facc(nb), tacc(nb);
    for (int i = 0; i < nb; i++) {
        const float *r = bodies + (size_t)i * BODY_STRIDE;
        pos[i] = make_float4(r[0], r[1], r[2], 0.f);
        quat[i] = make_float4(r[4], r[5], r[6], r[7]);
        lvel[i] = make_float4(r[8], r[9], r[10], 0.f);
        avel[i] = make_float4(r[12], r[13], r[14], 0.f);
        facc[i] = make_float4(r[16], r[17], r[18], 0.f);
        tacc[i] = make_float4(r[20], r[21], r[22], 0.f);
    }
    return upload(D.pos, pos.data(), nb) && upload(D.quat, quat.data(), nb) && upload(D.lvel, lvel.data(), nb) &&
           upload(D.avel, avel.data(), nb) && upload(D.facc, facc.data(), nb) && upload(D.tacc, tacc.data(), nb);
}

// The batch's worlds may be split into ranges stepped on streams of their own
// inside one graph: worlds are independent, so while one range finishes its
// slowest solve the others keep the GPU busy with their next steps.
struct odxSplit {
    int w0 = 0, w1 = 0;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
};

struct odxBatch {
    int nw = 0, nb = 0, ng = 0, nball = 0, jcap = 0, rcap = 0;
    std::vector<int> body_start, geom_start, ball_start, joint_start, row_start;
    std::vector<void *> allocs;
    DevWorlds D{};
    StepParams P{};  // solver settings; wbase / wcount / split / feat_base are set per launch
    Surface surf{};
    std::vector<int> geom_mat;  // material of each geom (GEOM slot 13)
    // what a snapshot needs to re-create the batch: the records as created,
    // the parameter vector, the surface table and the contact capacity
    std::vector<float> geoms_h, balls_h, table_h;
    double params_h[32] = {0};
    int contacts_per_body = 8;
    Surface *surf_default = nullptr;  // device copy of the parameters' surface
    Surface *surf_table = nullptr;  // device material-pair table (nmat * nmat), owned
    hipStream_t stream = nullptr;
    hipEvent_t fork = nullptr;
    std::vector<odxSplit> splits;  // empty: the whole batch on `stream`
    int graph_steps = 1;           // steps per graph launch
    hipGraphExec_t graph = nullptr, graph1 = nullptr;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    float *lambda0 = nullptr;
    int samples_split = 0;  // counter pair of the launch mode that ran last (for odxBatchGetSamples)
    bool has_jext = false;  // hinge / slider / fixed joints (extended records on the device)

    template <class T> bool alloc(T **p, size_t n) {
        if (!dalloc(p, n)) return false;
        allocs.push_back((void *)*p);
        return true;
    }
    ~odxBatch() {
        drop_graph();
        drop_splits();
        for (auto &e : ev) if (e) (void)hipEventDestroy(e);
        if (fork) (void)hipEventDestroy(fork);
        if (stream) (void)hipStreamDestroy(stream);
        for (void *p : allocs) (void)hipFree(p);
        if (surf_table) (void)hipFree(surf_table);
    }
    void drop_graph() {
        if (graph) { (void)hipGraphExecDestroy(graph); graph = nullptr; }
        if (graph1) { (void)hipGraphExecDestroy(graph1); graph1 = nullptr; }
    }
    void drop_splits() {
        for (auto &s : splits) {
            if (s.done) (void)hipEventDestroy(s.done);
            if (s.stream) (void)hipStreamDestroy(s.stream);
        }
        splits.clear();
    }
    // launch parameters of worlds [w0, w1) using counter pair `split`
    StepParams launch_params(int w0, int w1, int split) const {
        StepParams q = P;
        q.wbase = w0;
        q.wcount = w1 - w0;
        q.split = split;
        q.feat_base = joint_start[w0];
        return q;
    }
    const odxMlp *mlp = nullptr;
    bool infer_on() const { return P.warm == 3 && mlp; }
    bool enqueue_infer(const StepParams &q, hipStream_t s) {
        if (!infer_on()) return true;
        launch_mlp_infer(mlp, D.feat + (size_t)q.feat_base * FEAT_DIM, D.feat_row + q.feat_base, D.feat_count + 2 * q.split,
                         D.lambda0, joint_start[q.wbase + q.wcount] - q.feat_base, s, P.gate ? 64 : 4096);
        return odx_check(hipGetLastError(), "k_mlp_infer launch");
    }
    bool enqueue_step(const StepParams &q, hipStream_t s) {
        launch_prepare_small(D, q, s);
        if (!odx_check(hipGetLastError(), "k_prepare_small launch")) return false;
        if (!enqueue_infer(q, s)) return false;
        launch_solve_small(D, q, s);
        return odx_check(hipGetLastError(), "k_solve_small launch");
    }
    int fused = 0;  // > 0: steps per launch of the persistent fused kernel
    // nsteps steps of worlds q.wbase.. on stream s
    bool enqueue_range(StepParams q, int nsteps, hipStream_t s) {
        if (fused <= 0) {
            for (int k = 0; k < nsteps; k++)
                if (!enqueue_step(q, s)) return false;
            return true;
        }
        q.feat_local = 1;
        const MlpDev *M = infer_on() ? mlp_device_view(mlp) : nullptr;
        for (int k = 0; k < nsteps; k += fused) {
            launch_step_fused(D, q, M, std::min(fused, nsteps - k), s);
            if (!odx_check(hipGetLastError(), "k_step_fused launch")) return false;
        }
        return true;
    }
    // nsteps steps of every world, captured on `stream` (forked over the splits)
    bool enqueue_steps(int nsteps) {
        if (splits.empty()) {
            if (!enqueue_range(launch_params(0, nw, 0), nsteps, stream)) return false;
            samples_split = fused > 0 ? -2 : 0;
            return true;
        }
        if (!odx_check(hipEventRecord(fork, stream), "hipEventRecord")) return false;
        for (size_t i = 0; i < splits.size(); i++) {
            odxSplit &sp = splits[i];
            if (!odx_check(hipStreamWaitEvent(sp.stream, fork, 0), "hipStreamWaitEvent")) return false;
            if (!enqueue_range(launch_params(sp.w0, sp.w1, (int)i), nsteps, sp.stream)) return false;
            if (!odx_check(hipEventRecord(sp.done, sp.stream), "hipEventRecord") ||
                !odx_check(hipStreamWaitEvent(stream, sp.done, 0), "hipStreamWaitEvent"))
                return false;
        }
        samples_split = fused > 0 ? -2 : -1;
        return true;
    }
    bool capture(int nsteps, hipGraphExec_t *out) {
        hipGraph_t g = nullptr;
        if (!odx_check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture"))
            return false;
        bool ok = enqueue_steps(nsteps);
        hipError_t e = hipStreamEndCapture(stream, &g);
        if (!ok || !odx_check(e, "hipStreamEndCapture")) {
            if (g) (void)hipGraphDestroy(g);
            return false;
        }
        e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (!odx_check(e, "hipGraphInstantiate")) { *out = nullptr; return false; }
        return true;
    }
};

// Rows per world the batched solve keeps in LDS by default: 400 rows (64 B
// each plus lambda, work item and phase start: 74 B) and 65 bodies take
// 33 KB, so four to five worlds share a CU's 160 KB of LDS.
constexpr int ODX_DEFAULT_LDS_ROWS = 400;

// Largest row count <= want whose LDS footprint fits the kernel's limit
// (same accounting as solve_small_lds_bytes).
static int odx_lds_rows_fit(int lds_bodies, int want) {
    int r = std::max(0, std::min(want, SW_MAXROWS));
    while (r > 0 && solve_lds_floats(lds_bodies, r) * (int)sizeof(float) > SOLVE_LDS_MAX) r--;
    return r;
}

void odx_params_from(const double *p, StepParams &P, Surface &S) {
    P.h = (float)p[0];
    P.gx = (float)p[1]; P.gy = (float)p[2]; P.gz = (float)p[3];
    P.erp = (float)p[4]; P.cfm = (float)p[5]; P.sor_w = (float)p[6];
    P.max_iters = (int)p[7];
    P.tol = 0.f;
    P.check_every = 1;
    P.max_vel = (float)p[8]; P.min_depth = (float)p[9];
    P.maxc = (int)p[17] > 0 ? (int)p[17] : 1;
    P.space = (int)p[18];
    P.skip_connected = (int)p[24];
    P.order = 0;
    P.warm = 0;
    P.lambda_floor = p[25] > 0 ? (float)p[25] : 1e-3f;
    P.final_residual = 1;
    P.collide = 1;
    P.features = 0;
    S.mode = (int)p[10];
    S.mu = (float)p[11]; S.mu2 = (float)p[12]; S.bounce = (float)p[13]; S.bounce_vel = (float)p[14];
    S.soft_erp = (float)p[15]; S.soft_cfm = (float)p[16];
    S.motion1 = (float)p[19]; S.motion2 = (float)p[20]; S.motionN = (float)p[21];
    S.slip1 = (float)p[22]; S.slip2 = (float)p[23];
}

extern "C" {

const char *odxVersion(void) { return "odx 0.1 (gfx950)"; }

void odxSetMessageHandler(odxMessageFunction *fn) { g_handler = fn; }

int odxDeviceCount(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

odxBatch *odxBatchCreate(int nworlds, const int *body_start, const float *bodies, const int *geom_start,
                         const float *geoms, const int *ball_start, const float *balls, const double *params,
                         int contacts_per_body) {
    return odxBatchCreateJoints(nworlds, body_start, bodies, geom_start, geoms, ball_start, balls, nullptr, params,
                                contacts_per_body);
}

odxBatch *odxBatchCreateJoints(int nworlds, const int *body_start, const float *bodies, const int *geom_start,
                               const float *geoms, const int *joint_start, const float *joints, const float *jext,
                               const double *params, int contacts_per_body) {
    const int *ball_start = joint_start;
    const float *balls = joints;
    if (nworlds <= 0 || !body_start || !bodies || !geom_start || !geoms || !params) {
        odx_message(1, "odxBatchCreate: bad arguments");
        return nullptr;
    }
    auto *B = new odxBatch();
    B->nw = nworlds;
    B->body_start.assign(body_start, body_start + nworlds + 1);
    B->geom_start.assign(geom_start, geom_start + nworlds + 1);
    if (ball_start) B->ball_start.assign(ball_start, ball_start + nworlds + 1);
    else B->ball_start.assign(nworlds + 1, 0);
    B->nb = B->body_start[nworlds];
    B->ng = B->geom_start[nworlds];
    B->nball = B->ball_start[nworlds];
    B->geom_mat.resize(B->ng);
    for (int g = 0; g < B->ng; g++) B->geom_mat[g] = (int)geoms[(size_t)g * GEOM_STRIDE + 13];
    B->geoms_h.assign(geoms, geoms + (size_t)B->ng * GEOM_STRIDE);
    if (balls) B->balls_h.assign(balls, balls + (size_t)B->nball * JOINT_STRIDE);
    std::memcpy(B->params_h, params, sizeof B->params_h);
    if (contacts_per_body <= 0) contacts_per_body = 8;
    B->contacts_per_body = contacts_per_body;
    B->joint_start.resize(nworlds + 1);
    B->row_start.resize(nworlds + 1);
    B->joint_start[0] = 0;
    B->row_start[0] = 0;
    for (int w = 0; w < nworlds; w++) {
        int nbw = B->body_start[w + 1] - B->body_start[w];
        int ngw = B->geom_start[w + 1] - B->geom_start[w];
        int nlw = B->ball_start[w + 1] - B->ball_start[w];
        if (nbw < 0 || nbw > SW_MAXB || ngw < 0 || ngw > SW_MAXG) {
            odx_message(2, "odxBatchCreate: world %d has %d bodies / %d geoms (limits %d / %d)", w, nbw, ngw, SW_MAXB,
                        SW_MAXG);
            delete B;
            return nullptr;
        }
        // every body index a kernel will dereference must lie inside its world
        for (int g = B->geom_start[w]; g < B->geom_start[w + 1]; g++) {
            const int gb = (int)geoms[(size_t)g * GEOM_STRIDE + 1];
            const int cls = (int)geoms[(size_t)g * GEOM_STRIDE + 0];
            if (gb < -1 || gb >= nbw || (cls < 0 || cls > 4)) {
                odx_message(2, "odxBatchCreate: geom %d of world %d has class %d / body %d", g, w, cls, gb);
                delete B;
                return nullptr;
            }
        }
        for (int l = B->ball_start[w]; l < B->ball_start[w + 1]; l++) {
            const float *r = balls + (size_t)l * JOINT_STRIDE;
            const int ty = (int)r[0], x1 = (int)r[1], x2 = (int)r[2];
            if (ty < 1 || ty > JT_FIXED || x1 < 0 || x1 >= nbw || x2 < -1 || x2 >= nbw || x1 == x2) {
                odx_message(2, "odxBatchCreate: joint %d of world %d has type %d, bodies %d,%d", l, w, ty, x1, x2);
                delete B;
                return nullptr;
            }
            if (ty >= JT_HINGE && !jext) {
                odx_message(2, "odxBatchCreate: joint %d of world %d is a hinge, slider or fixed joint and no "
                               "extended records were given", l, w);
                delete B;
                return nullptr;
            }
            if (ty >= JT_HINGE) B->has_jext = true;
        }
        int cc = contacts_per_body * nbw + 16;
        int jc = nlw + cc;
        int rc = 3 * jc + 3 * nlw;  // hinge / slider / fixed joints take up to six rows
        if (rc > SW_MAXROWS) { rc = SW_MAXROWS; jc = rc / 3; }
        B->joint_start[w + 1] = B->joint_start[w] + jc;
        B->row_start[w + 1] = B->row_start[w] + rc;
    }
    B->jcap = B->joint_start[nworlds];
    B->rcap = B->row_start[nworlds];
    odx_params_from(params, B->P, B->surf);
    int maxnb = 0;
    for (int w = 0; w < nworlds; w++) maxnb = std::max(maxnb, B->body_start[w + 1] - B->body_start[w]);
    B->P.lds_bodies = maxnb + 1;
    B->P.lds_rows = odx_lds_rows_fit(B->P.lds_bodies, ODX_DEFAULT_LDS_ROWS);
    B->P.restart_after = 0;

    DevWorlds &D = B->D;
    D.nworlds = nworlds;
    int *ds = nullptr;
    bool ok = odx_state_upload(D, B->nb, bodies, B->ng, geoms, B->allocs);
    ok = ok && B->alloc(&ds, nworlds + 1) && upload(ds, B->body_start.data(), nworlds + 1);
    D.body_start = ds;
    ok = ok && B->alloc(&ds, nworlds + 1) && upload(ds, B->geom_start.data(), nworlds + 1);
    D.geom_start = ds;
    ok = ok && B->alloc(&ds, nworlds + 1) && upload(ds, B->ball_start.data(), nworlds + 1);
    D.ball_start = ds;
    ok = ok && B->alloc(&ds, nworlds + 1) && upload(ds, B->joint_start.data(), nworlds + 1);
    D.joint_start = ds;
    ok = ok && B->alloc(&ds, nworlds + 1) && upload(ds, B->row_start.data(), nworlds + 1);
    D.row_start = ds;
    float *f = nullptr;
    ok = ok && B->alloc(&f, (size_t)B->nball * JOINT_STRIDE) &&
         upload(f, balls ? balls : (const float *)nullptr, balls ? (size_t)B->nball * JOINT_STRIDE : 0);
    D.balls = f;
    if (B->has_jext) {
        ok = ok && B->alloc(&f, (size_t)B->nball * JEXT_STRIDE) && upload(f, jext, (size_t)B->nball * JEXT_STRIDE);
        D.jext = f;
    }
    Surface *sd = nullptr;
    ok = ok && B->alloc(&sd, 1) && upload(sd, &B->surf, 1);
    D.surf = sd;
    B->surf_default = sd;
    ok = ok && B->alloc(&D.joints, (size_t)B->jcap * JOINT_STRIDE);
    ok = ok && B->alloc(&D.njoints, nworlds) && B->alloc(&D.ncontacts, nworlds) && B->alloc(&D.nrows, nworlds) &&
         B->alloc(&D.nphases, nworlds) && B->alloc(&D.niters, nworlds) && B->alloc(&D.residual, nworlds);
    ok = ok && B->alloc(&D.rows, (size_t)B->rcap * ROW_STRIDE) && B->alloc(&D.lambda, B->rcap);
    ok = ok && B->alloc(&D.sched, B->rcap) && B->alloc(&D.phase_start, (size_t)B->rcap + nworlds);
    ok = ok && B->alloc(&D.perm, B->rcap) && B->alloc(&D.inv, B->rcap);
    ok = ok && B->alloc(&D.bscr, (size_t)B->nb * BSCR_STRIDE);
    ok = ok && B->alloc(&D.feat, (size_t)B->jcap * FEAT_DIM) && B->alloc(&D.feat_row, B->jcap) &&
         B->alloc(&D.feat_count, 2 * (ODX_MAX_SPLITS + 1)) && B->alloc(&D.joff, B->jcap);
    ok = ok && B->alloc(&D.prev, (size_t)B->jcap * JOINT_STRIDE) && B->alloc(&D.nprev, nworlds) &&
         B->alloc(&D.nfeat, nworlds);
    ok = ok && B->alloc(&D.overflow, 4) && B->alloc(&D.order, nworlds) && B->alloc(&D.wtime, 2 * (size_t)nworlds) &&
         B->alloc(&D.ptime, PTIME_SLOTS * (size_t)nworlds);
    ok = ok && B->alloc(&B->lambda0, B->rcap);
    if (ok) {
        ok = odx_check(hipMemset(D.overflow, 0, 4 * sizeof(int)), "hipMemset") &&
             odx_check(hipMemset(D.nrows, 0, nworlds * sizeof(int)), "hipMemset") &&
             odx_check(hipMemset(D.ncontacts, 0, nworlds * sizeof(int)), "hipMemset") &&
             odx_check(hipMemset(D.niters, 0, nworlds * sizeof(int)), "hipMemset") &&
             odx_check(hipMemset(D.nprev, 0, nworlds * sizeof(int)), "hipMemset") &&
             odx_check(hipMemset(D.feat_count, 0, 2 * (ODX_MAX_SPLITS + 1) * sizeof(int)), "hipMemset") &&
             odx_check(hipMemset(B->lambda0, 0, (size_t)B->rcap * sizeof(float)), "hipMemset") &&
             odx_check(hipMemset(D.lambda, 0, (size_t)B->rcap * sizeof(float)), "hipMemset");
    }
    D.lambda0 = B->lambda0;
    ok = ok && odx_check(hipStreamCreate(&B->stream), "hipStreamCreate");
    for (auto &e : B->ev) ok = ok && odx_check(hipEventCreate(&e), "hipEventCreate");
    ok = ok && odx_check(hipEventCreateWithFlags(&B->fork, hipEventDisableTiming), "hipEventCreate");
    if (!ok) {
        delete B;
        return nullptr;
    }
    return B;
}

void odxBatchDestroy(odxBatch *b) { delete b; }

int odxBatchSetSolver(odxBatch *b, int order, int max_iters, float tol, int check_every, int warm) {
    if (!b || max_iters < 0 || check_every <= 0 || order < 0 || order > 1 || warm < 0 || warm > 3) {
        odx_message(1, "odxBatchSetSolver: bad arguments");
        return 0;
    }
    b->P.order = order;
    b->P.max_iters = max_iters;
    b->P.tol = tol;
    b->P.check_every = check_every;
    b->P.warm = warm;
    b->drop_graph();
    return 1;
}

}  // extern "C"

// A material-pair surface table (float[nmat][nmat][16]) checked and
// converted for the device: symmetric, contact limits within 0..ODX_MAXC,
// every geom's material inside the table.  `who` names the caller in messages.
bool odx_surface_table(const char *who, int nmat, const float *table, const std::vector<int> &geom_mat,
                       std::vector<Surface> &t) {
    if (nmat < 0 || nmat > ODX_MAX_MATERIALS || (nmat > 0 && !table)) {
        odx_message(1, "%s: bad arguments (0 <= nmat <= %d)", who, ODX_MAX_MATERIALS);
        return false;
    }
    t.assign((size_t)std::max(nmat * nmat, 1), Surface{});
    for (int a = 0; a < nmat; a++)
        for (int c = 0; c < nmat; c++) {
            const float *e = table + ((size_t)a * nmat + c) * 16, *f = table + ((size_t)c * nmat + a) * 16;
            for (int k = 0; k < 13; k++)
                if (!(e[k] == f[k]) && !(e[k] != e[k] && f[k] != f[k])) {
                    odx_message(1, "%s: entry (%d, %d) differs from (%d, %d)", who, a, c, c, a);
                    return false;
                }
            if (e[12] < 0 || e[12] > ODX_MAXC) {
                odx_message(1, "%s: max contacts of (%d, %d) outside 0..%d", who, a, c, ODX_MAXC);
                return false;
            }
            Surface &s = t[(size_t)a * nmat + c];
            s.mode = (int)e[0];
            s.mu = e[1] < 0 ? 0.f : e[1]; s.mu2 = e[2] < 0 ? 0.f : e[2];
            s.bounce = e[3]; s.bounce_vel = e[4]; s.soft_erp = e[5]; s.soft_cfm = e[6];
            s.motion1 = e[7]; s.motion2 = e[8]; s.motionN = e[9]; s.slip1 = e[10]; s.slip2 = e[11];
            s.max_contacts = (int)e[12];
        }
    for (size_t g = 0; g < geom_mat.size(); g++)
        if (nmat > 0 && (geom_mat[g] < 0 || geom_mat[g] >= nmat)) {
            odx_message(1, "%s: geom %d has material %d, outside the table's %d", who, (int)g, geom_mat[g], nmat);
            return false;
        }
    return true;
}

extern "C" {

int odxBatchSetSurfaceTable(odxBatch *b, int nmat, const float *table, int category_filter) {
    if (!b) return 0;
    std::vector<Surface> t;
    if (!odx_surface_table("odxBatchSetSurfaceTable", nmat, table, b->geom_mat, t)) return 0;
    // no step in flight may still read the table being replaced
    if (!odx_check(hipDeviceSynchronize(), "hipDeviceSynchronize")) return 0;
    if (b->surf_table) { (void)hipFree(b->surf_table); b->surf_table = nullptr; }
    if (nmat > 0) {
        if (!dalloc(&b->surf_table, t.size()) || !upload(b->surf_table, t.data(), t.size())) return 0;
        b->D.surf = b->surf_table;
    } else {
        b->D.surf = b->surf_default;
    }
    b->P.nmat = nmat;
    b->P.cat_filter = category_filter ? 1 : 0;
    b->table_h.assign(table, table + (size_t)nmat * nmat * 16);
    b->drop_graph();
    return 1;
}

int odxBatchSetSolveOptions(odxBatch *b, int restart_after, int lds_rows) {
    if (!b || restart_after < 0 || lds_rows < 0) {
        odx_message(1, "odxBatchSetSolveOptions: bad arguments");
        return 0;
    }
    const int fit = odx_lds_rows_fit(b->P.lds_bodies, lds_rows);
    if (fit < lds_rows) {
        odx_message(1, "odxBatchSetSolveOptions: %d LDS rows exceed the %d-byte budget (at most %d)", lds_rows,
                    SOLVE_LDS_MAX, fit);
        return 0;
    }
    b->P.restart_after = restart_after;
    b->P.lds_rows = lds_rows;
    b->drop_graph();
    return 1;
}

int odxBatchSetSplits(odxBatch *b, int nsplit, int graph_steps) {
    if (!b || nsplit < 1 || nsplit > ODX_MAX_SPLITS || nsplit > b->nw || graph_steps < 1) {
        odx_message(1, "odxBatchSetSplits: bad arguments (1 <= nsplit <= %d and <= worlds, graph_steps >= 1)",
                    ODX_MAX_SPLITS);
        return 0;
    }
    if (!odxBatchSync(b)) return 0;
    b->drop_graph();
    b->drop_splits();
    b->graph_steps = graph_steps;
    if (nsplit == 1) return 1;
    for (int i = 0; i < nsplit; i++) {
        odxSplit s;
        s.w0 = (int)((long long)b->nw * i / nsplit);
        s.w1 = (int)((long long)b->nw * (i + 1) / nsplit);
        if (!odx_check(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking), "hipStreamCreate") ||
            !odx_check(hipEventCreateWithFlags(&s.done, hipEventDisableTiming), "hipEventCreate")) {
            if (s.stream) (void)hipStreamDestroy(s.stream);
            b->drop_splits();
            return 0;
        }
        b->splits.push_back(s);
    }
    return 1;
}

int odxBatchSetFused(odxBatch *b, int steps_per_launch) {
    if (!b || steps_per_launch < 0) {
        odx_message(1, "odxBatchSetFused: bad arguments");
        return 0;
    }
    if (!odxBatchSync(b)) return 0;
    b->fused = steps_per_launch;
    b->drop_graph();
    return 1;
}

// Launch-schedule autotuner: every candidate (world splits 1, 2, 4, 8 times
// staged launches or fused launches of graph_steps steps, with the graph size
// in effect) steps the batch `steps` steps `reps` times from
// the same body state; the fastest mean wins, where the schedule in effect at
// the call counts 5 % faster than measured (a change must pay for itself
// beyond the noise of trials shorter than the run they tune for).  Schedules do not change results, so only time is
// compared.
int odxBatchAutotune(odxBatch *b, int steps, int reps, int *schedule, float *ms_per_step) {
    if (!b || steps < 1 || reps < 1) {
        odx_message(1, "odxBatchAutotune: bad arguments (steps >= 1, reps >= 1)");
        return 0;
    }
    std::vector<float> saved((size_t)b->nb * ODX_BODY_STRIDE);
    if (!odxBatchGetBodies(b, saved.data())) return 0;
    // what the next step reads besides the bodies: the previous step's
    // contacts and impulses (warm starts) and sweep counts (dispatch order)
    DevWorlds &D = b->D;
    const std::pair<void *, size_t> state[] = {{D.prev, (size_t)b->jcap * JOINT_STRIDE * sizeof(float)},
                                               {D.nprev, (size_t)b->nw * sizeof(int)},
                                               {D.lambda, (size_t)b->rcap * sizeof(float)},
                                               {D.niters, (size_t)b->nw * sizeof(int)}};
    std::vector<void *> copy;
    auto release = [&]() {
        for (void *p : copy) (void)hipFree(p);
        copy.clear();
    };
    for (const auto &r : state) {
        void *p = nullptr;
        if (!odx_check(hipMalloc(&p, r.second), "hipMalloc")) { release(); return 0; }
        copy.push_back(p);
        if (!odx_check(hipMemcpy(p, r.first, r.second, hipMemcpyDeviceToDevice), "D2D")) { release(); return 0; }
    }
    auto restore = [&]() {
        if (!odxBatchSetBodies(b, saved.data())) return false;
        for (size_t i = 0; i < copy.size(); i++)
            if (!odx_check(hipMemcpy(state[i].first, copy[i], state[i].second, hipMemcpyDeviceToDevice), "D2D"))
                return false;
        return true;
    };
    const int cur_sp = b->splits.empty() ? 1 : (int)b->splits.size(), cur_fz = b->fused, cur_gs = b->graph_steps;
    const int split_c[] = {1, 2, 4, 8};
    const int fused_c[] = {0, cur_gs};
    float best = 0.f;
    int bs = 1, bf = 0;
    for (int fz : fused_c) {
        for (int sp : split_c) {
            if (sp > b->nw) continue;
            if (!odxBatchSetSplits(b, sp, cur_gs) || !odxBatchSetFused(b, fz)) { release(); return 0; }
            double tot = 0.0;
            for (int r = 0; r < reps; r++) {
                float ms = 0.f;
                if (!restore() || !odxBatchTime(b, steps, &ms)) { release(); return 0; }
                tot += ms;
            }
            float mean = (float)(tot / reps) / (float)steps;
            if (sp == cur_sp && fz == cur_fz) mean *= 0.95f;
            if (best == 0.f || mean < best) {
                best = mean;
                bs = sp;
                bf = fz;
            }
        }
    }
    const bool ok = odxBatchSetSplits(b, bs, cur_gs) && odxBatchSetFused(b, bf) && restore();
    release();
    if (!ok) return 0;
    if (schedule) {
        schedule[0] = bs;
        schedule[1] = cur_gs;
        schedule[2] = bf;
    }
    if (ms_per_step) *ms_per_step = best;
    return 1;
}

int odxBatchStep(odxBatch *b, int nsteps) {
    if (!b || nsteps < 0) return 0;
    if (nsteps == 0) return 1;
    const int K = b->graph_steps;
    if (nsteps >= K && !b->graph && !b->capture(K, &b->graph)) return 0;
    if (nsteps % K && !b->graph1 && !b->capture(1, &b->graph1)) return 0;
    for (int s = 0; s + K <= nsteps; s += K)
        if (!odx_check(hipGraphLaunch(b->graph, b->stream), "hipGraphLaunch")) return 0;
    for (int s = 0; s < nsteps % K; s++)
        if (!odx_check(hipGraphLaunch(b->graph1, b->stream), "hipGraphLaunch")) return 0;
    return 1;
}

int odxBatchSync(odxBatch *b) {
    if (!b) return 0;
    return odx_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize") ? 1 : 0;
}

int odxBatchTime(odxBatch *b, int nsteps, float *total_ms) {
    if (!b || nsteps <= 0) return 0;
    if (!odxBatchSync(b)) return 0;
    // instantiate the graphs outside the timed region
    const int K = b->graph_steps;
    if (nsteps >= K && !b->graph && !b->capture(K, &b->graph)) return 0;
    if (nsteps % K && !b->graph1 && !b->capture(1, &b->graph1)) return 0;
    if (!odx_check(hipEventRecord(b->ev[0], b->stream), "hipEventRecord")) return 0;
    if (!odxBatchStep(b, nsteps)) return 0;
    if (!odx_check(hipEventRecord(b->ev[1], b->stream), "hipEventRecord")) return 0;
    if (!odx_check(hipEventSynchronize(b->ev[1]), "hipEventSynchronize")) return 0;
    float tot = 0.f;
    if (!odx_check(hipEventElapsedTime(&tot, b->ev[0], b->ev[1]), "hipEventElapsedTime")) return 0;
    if (total_ms) *total_ms = tot;
    return 1;
}

int odxBatchKernelTimes(odxBatch *b, int nsteps, float *kernel_ms) {
    if (!b || nsteps <= 0 || !kernel_ms) return 0;
    if (!odxBatchSync(b)) return 0;
    // events: before prepare, before inference, before solve, per step; one final
    std::vector<hipEvent_t> ev(3 * (size_t)nsteps + 1, nullptr);
    bool ok = true;
    for (auto &e : ev) ok = ok && odx_check(hipEventCreate(&e), "hipEventCreate");
    // whole-batch launches (the kernels' full grids) on the batch stream, with
    // a counter pair of their own
    const StepParams q = b->launch_params(0, b->nw, ODX_MAX_SPLITS);
    for (int s = 0; ok && s < nsteps; s++) {
        ok = ok && odx_check(hipEventRecord(ev[3 * s], b->stream), "hipEventRecord");
        launch_prepare_small(b->D, q, b->stream);
        ok = ok && odx_check(hipGetLastError(), "k_prepare_small launch");
        ok = ok && odx_check(hipEventRecord(ev[3 * s + 1], b->stream), "hipEventRecord");
        ok = ok && b->enqueue_infer(q, b->stream);
        ok = ok && odx_check(hipEventRecord(ev[3 * s + 2], b->stream), "hipEventRecord");
        launch_solve_small(b->D, q, b->stream);
        ok = ok && odx_check(hipGetLastError(), "k_solve_small launch");
    }
    b->samples_split = ODX_MAX_SPLITS;
    ok = ok && odx_check(hipEventRecord(ev[3 * (size_t)nsteps], b->stream), "hipEventRecord");
    ok = ok && odx_check(hipEventSynchronize(ev[3 * (size_t)nsteps]), "hipEventSynchronize");
    double ta = 0.0, ti = 0.0, tc = 0.0;
    for (int s = 0; ok && s < nsteps; s++) {
        float a = 0.f, i = 0.f, c = 0.f;
        ok = ok && odx_check(hipEventElapsedTime(&a, ev[3 * s], ev[3 * s + 1]), "hipEventElapsedTime");
        ok = ok && odx_check(hipEventElapsedTime(&i, ev[3 * s + 1], ev[3 * s + 2]), "hipEventElapsedTime");
        ok = ok && odx_check(hipEventElapsedTime(&c, ev[3 * s + 2], ev[3 * s + 3]), "hipEventElapsedTime");
        ta += a;
        ti += i;
        tc += c;
    }
    for (auto &e : ev) if (e) (void)hipEventDestroy(e);
    if (!ok) return 0;
    kernel_ms[0] = (float)(ta / nsteps);
    kernel_ms[1] = (float)(ti / nsteps);
    kernel_ms[2] = (float)(tc / nsteps);
    return 1;
}

int odxBatchScheduleKernelTimes(odxBatch *b, int nsteps, float *kernel_ms, int *ranges) {
    if (!b || nsteps <= 0 || !kernel_ms || !ranges) return 0;
    if (!odxBatchSync(b)) return 0;
    struct Range { int w0, w1; hipStream_t s; };
    std::vector<Range> rs;
    if (b->splits.empty()) rs.push_back({0, b->nw, b->stream});
    else
        for (const odxSplit &sp : b->splits) rs.push_back({sp.w0, sp.w1, sp.stream});
    const int nr = (int)rs.size();
    // per step and range: before prepare, before inference, before solve, after
    std::vector<hipEvent_t> ev((size_t)nsteps * nr * 4, nullptr);
    bool ok = true;
    for (auto &e : ev) ok = ok && odx_check(hipEventCreate(&e), "hipEventCreate");
    auto E = [&](int s, int r, int k) { return ev[((size_t)s * nr + r) * 4 + k]; };
    for (int s = 0; ok && s < nsteps; s++) {
        if (!b->splits.empty()) ok = ok && odx_check(hipEventRecord(b->fork, b->stream), "hipEventRecord");
        for (int r = 0; ok && r < nr; r++) {
            const Range &R = rs[(size_t)r];
            if (!b->splits.empty()) ok = ok && odx_check(hipStreamWaitEvent(R.s, b->fork, 0), "hipStreamWaitEvent");
            StepParams q = b->launch_params(R.w0, R.w1, b->splits.empty() ? 0 : r);
            ok = ok && odx_check(hipEventRecord(E(s, r, 0), R.s), "hipEventRecord");
            if (b->fused > 0) {
                q.feat_local = 1;
                const MlpDev *M = b->infer_on() ? mlp_device_view(b->mlp) : nullptr;
                ok = ok && odx_check(hipEventRecord(E(s, r, 1), R.s), "hipEventRecord") &&
                     odx_check(hipEventRecord(E(s, r, 2), R.s), "hipEventRecord");
                launch_step_fused(b->D, q, M, 1, R.s);
                ok = ok && odx_check(hipGetLastError(), "k_step_fused launch");
            } else {
                launch_prepare_small(b->D, q, R.s);
                ok = ok && odx_check(hipGetLastError(), "k_prepare_small launch");
                ok = ok && odx_check(hipEventRecord(E(s, r, 1), R.s), "hipEventRecord");
                ok = ok && b->enqueue_infer(q, R.s);
                ok = ok && odx_check(hipEventRecord(E(s, r, 2), R.s), "hipEventRecord");
                launch_solve_small(b->D, q, R.s);
                ok = ok && odx_check(hipGetLastError(), "k_solve_small launch");
            }
            ok = ok && odx_check(hipEventRecord(E(s, r, 3), R.s), "hipEventRecord");
            if (!b->splits.empty()) {
                const odxSplit &sp = b->splits[(size_t)r];
                ok = ok && odx_check(hipEventRecord(sp.done, R.s), "hipEventRecord") &&
                     odx_check(hipStreamWaitEvent(b->stream, sp.done, 0), "hipStreamWaitEvent");
            }
        }
    }
    b->samples_split = b->fused > 0 ? -2 : (b->splits.empty() ? 0 : -1);
    ok = ok && odxBatchSync(b);
    for (int r = 0; ok && r < nr; r++) {
        double t[3] = {0.0, 0.0, 0.0};
        for (int s = 0; ok && s < nsteps; s++)
            for (int k = 0; ok && k < 3; k++) {
                float ms = 0.f;
                ok = ok && odx_check(hipEventElapsedTime(&ms, E(s, r, k), E(s, r, k + 1)), "hipEventElapsedTime");
                t[k] += ms;
            }
        for (int k = 0; k < 3; k++) kernel_ms[3 * r + k] = (float)(t[k] / nsteps);
        ranges[2 * r] = rs[(size_t)r].w0;
        ranges[2 * r + 1] = rs[(size_t)r].w1;
    }
    for (auto &e : ev) if (e) (void)hipEventDestroy(e);
    return ok ? nr : 0;
}

int odxBatchGetBodies(odxBatch *b, float *out) {
    if (!b || !out) return 0;
    if (!odxBatchSync(b)) return 0;
    return odx_bodies_download(b->D, b->nb, out) ? b->nb : 0;
}

int odxBatchSetBodies(odxBatch *b, const float *bodies) {
    if (!b || !bodies) return 0;
    if (!odxBatchSync(b)) return 0;
    return odx_bodies_upload(b->D, b->nb, bodies) ? b->nb : 0;
}

// ---- binary snapshot: "ODXSNAP1", then int32 counts (worlds, bodies, geoms,
// balls, contacts per body, materials, category filter, order, max_iters,
// check_every, warm, restart_after, lds_rows), float tol, the int32 world
// offsets, the float records (bodies as they are now), double params[32] and
// the float surface table
namespace {
constexpr char SNAP_MAGIC[8] = {'O', 'D', 'X', 'S', 'N', 'A', 'P', '1'};
constexpr int SNAP_NINT = 13;
}  // namespace

int odxBatchSaveSnapshot(odxBatch *b, const char *path) {
    if (!b || !path) {
        odx_message(1, "odxBatchSaveSnapshot: bad arguments");
        return 0;
    }
    if (b->has_jext) {
        odx_message(1, "odxBatchSaveSnapshot: batches with hinge, slider or fixed joints are not saved");
        return 0;
    }
    std::vector<float> bodies((size_t)b->nb * BODY_STRIDE);
    if (!odxBatchSync(b) || !odx_bodies_download(b->D, b->nb, bodies.data())) return 0;
    FILE *f = std::fopen(path, "wb");
    if (!f) {
        odx_message(2, "odxBatchSaveSnapshot: cannot open %s", path);
        return 0;
    }
    const int hdr[SNAP_NINT] = {b->nw, b->nb, b->ng, b->nball, b->contacts_per_body, b->P.nmat, b->P.cat_filter,
                                b->P.order, b->P.max_iters, b->P.check_every, b->P.warm, b->P.restart_after,
                                b->P.lds_rows};
    const float tol = b->P.tol;
    bool ok = std::fwrite(SNAP_MAGIC, 1, 8, f) == 8 && std::fwrite(hdr, sizeof(int), SNAP_NINT, f) == SNAP_NINT &&
              std::fwrite(&tol, sizeof tol, 1, f) == 1 &&
              std::fwrite(b->body_start.data(), sizeof(int), b->nw + 1, f) == (size_t)b->nw + 1 &&
              std::fwrite(b->geom_start.data(), sizeof(int), b->nw + 1, f) == (size_t)b->nw + 1 &&
              std::fwrite(b->ball_start.data(), sizeof(int), b->nw + 1, f) == (size_t)b->nw + 1 &&
              std::fwrite(bodies.data(), sizeof(float), bodies.size(), f) == bodies.size() &&
              std::fwrite(b->geoms_h.data(), sizeof(float), b->geoms_h.size(), f) == b->geoms_h.size() &&
              std::fwrite(b->balls_h.data(), sizeof(float), b->balls_h.size(), f) == b->balls_h.size() &&
              std::fwrite(b->params_h, sizeof(double), 32, f) == 32 &&
              std::fwrite(b->table_h.data(), sizeof(float), b->table_h.size(), f) == b->table_h.size();
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) odx_message(2, "odxBatchSaveSnapshot: write to %s failed", path);
    return ok ? 1 : 0;
}

int odxBatchGetLayout(odxBatch *b, int *nworlds, int *nbodies, int *body_start) {
    if (!b) return 0;
    if (nworlds) *nworlds = b->nw;
    if (nbodies) *nbodies = b->nb;
    if (body_start) std::memcpy(body_start, b->body_start.data(), (size_t)(b->nw + 1) * sizeof(int));
    return 1;
}

odxBatch *odxBatchLoadSnapshot(const char *path) {
    FILE *f = path ? std::fopen(path, "rb") : nullptr;
    if (!f) {
        odx_message(2, "odxBatchLoadSnapshot: cannot open %s", path ? path : "(null)");
        return nullptr;
    }
    char magic[8];
    int hdr[SNAP_NINT];
    float tol = 0.f;
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, SNAP_MAGIC, 8) == 0 &&
              std::fread(hdr, sizeof(int), SNAP_NINT, f) == SNAP_NINT && std::fread(&tol, sizeof tol, 1, f) == 1;
    const int nw = ok ? hdr[0] : 0, nb = ok ? hdr[1] : 0, ng = ok ? hdr[2] : 0, nball = ok ? hdr[3] : 0;
    const int nmat = ok ? hdr[5] : 0;
    ok = ok && nw > 0 && nb >= 0 && ng >= 0 && nball >= 0 && nmat >= 0 && nmat <= ODX_MAX_MATERIALS;
    std::vector<int> bs(ok ? nw + 1 : 0), gs(bs.size()), ls(bs.size());
    std::vector<float> bodies(ok ? (size_t)nb * BODY_STRIDE : 0), geoms(ok ? (size_t)ng * GEOM_STRIDE : 0),
        balls(ok ? (size_t)nball * JOINT_STRIDE : 0), table(ok ? (size_t)nmat * nmat * 16 : 0);
    double params[32];
    ok = ok && std::fread(bs.data(), sizeof(int), bs.size(), f) == bs.size() &&
         std::fread(gs.data(), sizeof(int), gs.size(), f) == gs.size() &&
         std::fread(ls.data(), sizeof(int), ls.size(), f) == ls.size() &&
         std::fread(bodies.data(), sizeof(float), bodies.size(), f) == bodies.size() &&
         std::fread(geoms.data(), sizeof(float), geoms.size(), f) == geoms.size() &&
         std::fread(balls.data(), sizeof(float), balls.size(), f) == balls.size() &&
         std::fread(params, sizeof(double), 32, f) == 32 &&
         std::fread(table.data(), sizeof(float), table.size(), f) == table.size();
    std::fclose(f);
    ok = ok && bs[0] == 0 && bs[nw] == nb && gs[0] == 0 && gs[nw] == ng && ls[0] == 0 && ls[nw] == nball;
    if (!ok) {
        odx_message(2, "odxBatchLoadSnapshot: %s is not a complete snapshot", path);
        return nullptr;
    }
    odxBatch *b = odxBatchCreate(nw, bs.data(), bodies.data(), gs.data(), geoms.data(), ls.data(),
                                 nball ? balls.data() : nullptr, params, hdr[4]);
    if (!b) return nullptr;
    if (!odxBatchSetSolver(b, hdr[7], hdr[8], tol, hdr[9], hdr[10]) || !odxBatchSetSolveOptions(b, hdr[11], hdr[12]) ||
        (nmat > 0 && !odxBatchSetSurfaceTable(b, nmat, table.data(), hdr[6]))) {
        odxBatchDestroy(b);
        return nullptr;
    }
    return b;
}

int odxBatchGetStats(odxBatch *b, int *ncontacts, int *nrows, int *niters, float *residual) {
    if (!b) return 0;
    if (!odxBatchSync(b)) return 0;
    bool ok = true;
    if (ncontacts) ok = ok && download(ncontacts, b->D.ncontacts, b->nw);
    if (nrows) ok = ok && download(nrows, b->D.nrows, b->nw);
    if (niters) ok = ok && download(niters, b->D.niters, b->nw);
    if (residual) ok = ok && download(residual, b->D.residual, b->nw);
    return ok ? b->nw : 0;
}

int odxBatchGetJoints(odxBatch *b, int w, float *joints, int cap) {
    if (!b || w < 0 || w >= b->nw) return 0;
    if (!odxBatchSync(b)) return 0;
    int nj = 0;
    if (!download(&nj, b->D.njoints + w, 1)) return 0;
    int n = nj < cap ? nj : cap;
    if (!download(joints, b->D.joints + (size_t)b->joint_start[w] * JOINT_STRIDE, (size_t)n * JOINT_STRIDE)) return 0;
    return n;
}

int odxBatchGetLambda(odxBatch *b, int w, float *lambda, int cap) {
    if (!b || w < 0 || w >= b->nw) return 0;
    if (!odxBatchSync(b)) return 0;
    int m = 0;
    if (!download(&m, b->D.nrows + w, 1)) return 0;
    int n = m < cap ? m : cap;
    if (!download(lambda, b->D.lambda + b->row_start[w], n)) return 0;
    return n;
}

int odxBatchGetRows(odxBatch *b, int w, float *rows, int cap) {
    if (!b || w < 0 || w >= b->nw) return 0;
    if (!odxBatchSync(b)) return 0;
    int m = 0;
    if (!download(&m, b->D.nrows + w, 1)) return 0;
    int n = m < cap ? m : cap;
    if (!download(rows, b->D.rows + (size_t)b->row_start[w] * ROW_STRIDE, (size_t)n * ROW_STRIDE)) return 0;
    return n;
}

int odxBatchGetOverflow(odxBatch *b, int *counts) {
    if (!b || !counts) return 0;
    if (!odxBatchSync(b)) return 0;
    return download(counts, b->D.overflow, 4) ? 4 : 0;
}

// Device wall-clock stamps -> microseconds from the earliest stamp.
static int odx_stamps_us(const long long *dev_stamps, size_t n, double *out) {
    std::vector<long long> t(n);
    if (!download(t.data(), dev_stamps, n)) return 0;
    int dev = 0, khz = 0;
    if (!odx_check(hipGetDevice(&dev), "hipGetDevice") ||
        !odx_check(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev), "hipDeviceGetAttribute") || khz <= 0)
        return 0;
    long long t0 = t[0];
    for (long long v : t) t0 = std::min(t0, v);
    for (size_t i = 0; i < n; i++) out[i] = (double)(t[i] - t0) * 1e3 / (double)khz;
    return 1;
}

int odxBatchGetSolveTimes(odxBatch *b, double *start_end_us) {
    if (!b || !start_end_us) return 0;
    if (!odxBatchSync(b)) return 0;
    return odx_stamps_us(b->D.wtime, 2 * (size_t)b->nw, start_end_us) ? b->nw : 0;
}

int odxBatchGetPrepareTimes(odxBatch *b, double *stamps_us) {
    if (!b || !stamps_us) return 0;
    if (!odxBatchSync(b)) return 0;
    return odx_stamps_us(b->D.ptime, PTIME_SLOTS * (size_t)b->nw, stamps_us) ? b->nw : 0;
}

int odxBatchRowStarts(odxBatch *b, int *row_start) {
    if (!b || !row_start) return 0;
    std::memcpy(row_start, b->row_start.data(), (b->nw + 1) * sizeof(int));
    return b->nw + 1;
}

int odxBatchSetMlp(odxBatch *b, const odxMlp *mlp) {
    if (!b) return 0;
    b->mlp = mlp;
    b->P.gate = mlp_gate(mlp);
    b->drop_graph();
    return 1;
}

int odxBatchGetSamples(odxBatch *b, float *X, float *T, int cap) {
    if (!b || !X || !T || cap <= 0) return 0;
    if (!odxBatchSync(b)) return 0;
    std::vector<int> cnt(2 * (ODX_MAX_SPLITS + 1), 0);
    if (!download(cnt.data(), b->D.feat_count, cnt.size())) return 0;
    // feature ranges (first slot, count) of the launches that ran last
    std::vector<std::pair<int, int>> ranges;
    if (b->samples_split >= 0) {
        ranges.push_back({0, cnt[2 * b->samples_split + 1]});
    } else if (b->samples_split == -1) {
        for (size_t i = 0; i < b->splits.size(); i++) ranges.push_back({b->joint_start[b->splits[i].w0], cnt[2 * i + 1]});
    } else {  // fused kernel: world-local slots
        std::vector<int> nf(b->nw);
        if (!download(nf.data(), b->D.nfeat, b->nw)) return 0;
        for (int w = 0; w < b->nw; w++) ranges.push_back({b->joint_start[w], nf[w]});
    }
    std::vector<float> lam(b->rcap);
    if (!download(lam.data(), b->D.lambda, b->rcap)) return 0;
    int n = 0;
    for (const auto &rg : ranges) {
        if (n >= cap) break;
        const int k = std::min(rg.second, cap - n);
        if (k <= 0) continue;
        std::vector<int> rows(k);
        if (!download(X + (size_t)n * FEAT_DIM, b->D.feat + (size_t)rg.first * FEAT_DIM, (size_t)k * FEAT_DIM) ||
            !download(rows.data(), b->D.feat_row + rg.first, k))
            return 0;
        for (int i = 0; i < k; i++) {
            const float s = X[(size_t)(n + i) * FEAT_DIM + 15];
            for (int o = 0; o < 3; o++) T[(size_t)(n + i) * 3 + o] = lam[rows[i] + o] / s;
        }
        n += k;
    }
    return n;
}

int odxBatchSetLambda0(odxBatch *b, const float *lambda0, int n) {
    if (!b || !lambda0 || n != b->rcap) {
        odx_message(1, "odxBatchSetLambda0: expected %d values", b ? b->rcap : 0);
        return 0;
    }
    if (!odxBatchSync(b)) return 0;
    return upload(b->lambda0, lambda0, (size_t)n) ? 1 : 0;
}

}  // extern "C"
