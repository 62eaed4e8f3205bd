// odx_api.cpp -- the ODE-compatible C API (include/ode.h) over the HIP kernels.
//
// Host objects keep ODE's observable structure: bodies and joints are
// head-inserted into LIFO lists (ode.cpp:54-60, 1383-1439) so the island DFS
// of dWorldQuickStep (util.cpp:724-855) visits joints in the reference's order,
// and contact joints live in joint groups emptied LIFO (ode.cpp:1325-1366).
// The numerical work runs on the GPU: dWorldQuickStep -> k_single_step (body
// stage, rows, SOR sweeps, feedback, integration); dSpaceCollide -> k_aabbs +
// k_pairs_all; dCollide -> k_collide_pairs, evaluated for every candidate pair
// of the current dSpaceCollide at the first call and served from that batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/ode.h"
#include "odx_device.h"
#include "odx_host.h"

namespace odx {
void launch_single_step(const DevWorlds &D, const StepParams &P, int nb, int m, int nitems, int nphases,
                        hipStream_t s);
bool launch_single_groups(const DevWorlds &D, const StepParams &P, size_t lds, bool rows_lds, hipStream_t s);
int single_lds_max();
void launch_aabbs(const float *gviews, int n, float *aabb, hipStream_t s);
void launch_space_collide(const SpaceCall &C, hipStream_t s);
int space_max_geoms();
void launch_collide_pairs(const float *gviews, const int *pairs, int np, int maxc, float *out, int *counts,
                          hipStream_t s);
}  // namespace odx

using namespace odx;

// ------------------------------------------------------------ device arena
namespace {

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    bool ensure(size_t bytes) {
        if (bytes <= cap && p) return true;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        size_t want = std::max<size_t>(bytes, 256);
        want = want + want / 2;
        if (!odx_check(hipMalloc(&p, want), "hipMalloc")) return false;
        cap = want;
        return true;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    ~DevBuf() { release(); }
    template <class T> T *as() const { return reinterpret_cast<T *>(p); }
};

hipStream_t g_stream = nullptr;
bool ensure_stream() {
    if (g_stream) return true;
    return odx_check(hipStreamCreate(&g_stream), "hipStreamCreate");
}

template <class T> bool up(DevBuf &b, const T *h, size_t n) {
    if (!b.ensure(n * sizeof(T))) return false;
    if (n == 0) return true;
    return odx_check(hipMemcpyAsync(b.p, h, n * sizeof(T), hipMemcpyHostToDevice, g_stream), "hipMemcpyAsync H2D");
}
template <class T> bool down(T *h, const DevBuf &b, size_t n) {
    if (n == 0) return true;
    return odx_check(hipMemcpyAsync(h, b.p, n * sizeof(T), hipMemcpyDeviceToHost, g_stream), "hipMemcpyAsync D2H");
}

// Page-locked host memory, grown on demand.
struct PinnedBuf {
    void *p = nullptr;
    size_t cap = 0;
    bool ensure(size_t bytes) {
        if (bytes <= cap && p) return true;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
        size_t want = std::max<size_t>(bytes, 256);
        want = want + want / 2;
        if (!odx_check(hipHostMalloc(&p, want, 0), "hipHostMalloc")) return false;
        cap = want;
        return true;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
    ~PinnedBuf() { release(); }
    template <class T> T *as() const { return reinterpret_cast<T *>(p); }
};

// One call's transfers packed into a page-locked host block and its device
// twin, in 256-byte aligned slots: the inputs go up in one copy and the
// outputs, laid out first, come back in one.  Every user synchronises its
// stream before returning, so no copy is in flight when a block is reused.
struct Staging {
    DevBuf dev;
    PinnedBuf host;
    size_t used = 0;
    void reset() { used = 0; }
    size_t slot(size_t bytes) {
        const size_t off = (used + 255) & ~(size_t)255;
        used = off + bytes;
        return off;
    }
    bool alloc() { return host.ensure(used) && dev.ensure(used); }
    template <class T> T *h(size_t off) const { return reinterpret_cast<T *>(host.as<char>() + off); }
    template <class T> T *d(size_t off) const { return reinterpret_cast<T *>(dev.as<char>() + off); }
    template <class T> void put(size_t off, const T *src, size_t n) {
        if (n) std::memcpy(host.as<char>() + off, src, n * sizeof(T));
    }
    bool send(size_t off, size_t bytes) {
        return bytes == 0 || odx_check(hipMemcpyAsync(dev.as<char>() + off, host.as<char>() + off, bytes,
                                                      hipMemcpyHostToDevice, g_stream), "hipMemcpyAsync H2D");
    }
    bool fetch(size_t off, size_t bytes) {
        return bytes == 0 || odx_check(hipMemcpyAsync(host.as<char>() + off, dev.as<char>() + off, bytes,
                                                      hipMemcpyDeviceToHost, g_stream), "hipMemcpyAsync D2H");
    }
};

void rfromq(float *R, const float *q) {
    float qq1 = 2 * q[1] * q[1], qq2 = 2 * q[2] * q[2], qq3 = 2 * q[3] * q[3];
    R[0] = 1 - qq2 - qq3; R[1] = 2 * (q[1] * q[2] - q[0] * q[3]); R[2] = 2 * (q[1] * q[3] + q[0] * q[2]); R[3] = 0;
    R[4] = 2 * (q[1] * q[2] + q[0] * q[3]); R[5] = 1 - qq1 - qq3; R[6] = 2 * (q[2] * q[3] - q[0] * q[1]); R[7] = 0;
    R[8] = 2 * (q[1] * q[3] - q[0] * q[2]); R[9] = 2 * (q[2] * q[3] + q[0] * q[1]); R[10] = 1 - qq1 - qq2; R[11] = 0;
}

// Quaternion of a rotation matrix (the trace / largest-diagonal branches of
// dQfromR, rotation.cpp:258-303).
void qfromr(float *q, const float *R) {
    auto r = [&](int i, int j) { return R[i * 4 + j]; };
    float tr = r(0, 0) + r(1, 1) + r(2, 2), s;
    if (tr >= 0) {
        s = std::sqrt(tr + 1);
        q[0] = 0.5f * s;
        s = 0.5f / s;
        q[1] = (r(2, 1) - r(1, 2)) * s; q[2] = (r(0, 2) - r(2, 0)) * s; q[3] = (r(1, 0) - r(0, 1)) * s;
        return;
    }
    int c = 0;
    if (r(1, 1) > r(0, 0)) c = r(2, 2) > r(1, 1) ? 2 : 1;
    else if (r(2, 2) > r(0, 0)) c = 2;
    if (c == 0) {
        s = std::sqrt((r(0, 0) - (r(1, 1) + r(2, 2))) + 1);
        q[1] = 0.5f * s; s = 0.5f / s;
        q[2] = (r(0, 1) + r(1, 0)) * s; q[3] = (r(2, 0) + r(0, 2)) * s; q[0] = (r(2, 1) - r(1, 2)) * s;
    } else if (c == 1) {
        s = std::sqrt((r(1, 1) - (r(2, 2) + r(0, 0))) + 1);
        q[2] = 0.5f * s; s = 0.5f / s;
        q[3] = (r(1, 2) + r(2, 1)) * s; q[1] = (r(0, 1) + r(1, 0)) * s; q[0] = (r(0, 2) - r(2, 0)) * s;
    } else {
        s = std::sqrt((r(2, 2) - (r(0, 0) + r(1, 1))) + 1);
        q[3] = 0.5f * s; s = 0.5f / s;
        q[1] = (r(2, 0) + r(0, 2)) * s; q[2] = (r(1, 2) + r(2, 1)) * s; q[0] = (r(1, 0) - r(0, 1)) * s;
    }
}

void normq(float *q) {
    float l = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    if (l > 0) { l = 1.f / std::sqrt(l); for (int i = 0; i < 4; i++) q[i] *= l; }
    else { q[0] = 1; q[1] = q[2] = q[3] = 0; }
}

}  // namespace

// ------------------------------------------------------------- objects
struct dxJointNode {
    dxJoint *joint;
    dxBody *body;  // the other body (NULL = world)
    dxJointNode *next;
};

// Auto-disable parameters (dxAutoDisable, ode/src/objects.h:87-96; defaults
// objects.cpp:54-61); thresholds are kept squared, as the reference keeps them.
struct AutoDisableParams {
    float idle_time = 0.f;
    int idle_steps = 10;
    unsigned average_samples = 1;
    float lin_thr2 = 0.01f * 0.01f, ang_thr2 = 0.01f * 0.01f;
};

// Damping parameters (dxDampingParameters, objects.h:100-108; defaults
// objects.cpp:63-69); thresholds squared.
struct DampingParams {
    float lin_scale = 0.f, ang_scale = 0.f;
    float lin_thr2 = 0.01f * 0.01f, ang_thr2 = 0.01f * 0.01f;
};

struct dxBody {
    dxWorld *world = nullptr;
    dxBody *next = nullptr, *prev = nullptr;
    dVector3 pos = {0, 0, 0, 0};
    dQuaternion q = {1, 0, 0, 0};
    dMatrix3 R = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    dVector3 lvel = {0, 0, 0, 0}, avel = {0, 0, 0, 0}, facc = {0, 0, 0, 0}, tacc = {0, 0, 0, 0};
    dMass mass{};
    bool gravity = true, gyro = true, enabled = true;
    // auto-disable state (objects.h:237-243)
    bool auto_disable = false;
    AutoDisableParams adis;
    float adis_timeleft = 0.f;
    int adis_stepsleft = 10;
    std::vector<float> avg_lvel, avg_avel;  // [average_samples][3]
    unsigned avg_counter = 0;
    int avg_ready = 0;
    // damping and angular speed cap (objects.h:246-247)
    bool lin_damping = false, ang_damping = false, max_ang = false;
    DampingParams dampingp;
    float max_angular_speed = dInfinity;
    dxJointNode *firstjoint = nullptr;
    std::vector<dxGeom *> geoms;
    void *data = nullptr;
    int tag = 0;
    int uid = 0;
};

struct dxJoint {
    dxWorld *world = nullptr;
    int type = dJointTypeNone;
    dxJointGroup *group = nullptr;
    dxJointNode node[2] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
    bool reverse = false, enabled = true;
    dContact contact{};
    dVector3 anchor1 = {0, 0, 0, 0}, anchor2 = {0, 0, 0, 0};
    float erp = 0.2f, cfm = 1e-5f;
    dJointFeedback *feedback = nullptr;
    int tag = 0;
    dxJoint *wnext = nullptr, *wprev = nullptr;  // the world's joint list (newest first, as addObjectToList)
    bool listed = false;
    // hinge / slider / fixed (joints/hinge.h, slider.h, fixed.h)
    dVector3 axis1 = {0, 0, 0, 0}, axis2 = {0, 0, 0, 0}, offset = {0, 0, 0, 0};
    dQuaternion qrel = {1, 0, 0, 0};
    // dxJointLimitMotor (joints/joint.h); set up by new_joint as init() does (joint.cpp:495-508)
    float lostop = -dInfinity, histop = dInfinity, vel = 0, fmax = 0, fudge = 1, normal_cfm = 0, stop_erp = 0,
          stop_cfm = 0, bounce = 0, limit_err = 0;
    int limit = 0;
};

struct dxJointGroup {
    std::vector<dxJoint *> joints;
};

// A space keeps its geoms as the reference's space classes do, because the
// order of the near-callback calls depends on it:
//  * sweep and prune (collision_sapspace.cpp:300-421): GeomList and
//    DirtyList, a moved geom swap-removed from the first and appended to the
//    second, the dirty ones appended to GeomList at the next collide;
//  * simple and hash (collision_space.cpp:166-209): one list with new and
//    moved geoms at its front.
// The radix sort's ranks stay on the device between calls.
struct dxSpace {
    int kind = 2;  // 0 SAP, 1 hash, 2 simple
    int ax0 = 0, ax1 = 2, ax2 = 4;  // SAP axes as AABB indices (dSAP_AXES_*)
    int minlevel = -3, maxlevel = 10;
    int cleanup = 1;
    int count = 0;
    dxGeom *first = nullptr;            // simple / hash list
    std::vector<dxGeom *> glist, dlist;  // sweep and prune: GeomList, DirtyList
    DevBuf ranks[2];                    // RaixSortContext ranks (ping-pong)
    int rcur = 0;
    bool ranks_valid = false;
    size_t ranks_used = 0;
    dxSpace *parent = nullptr;
};

struct dxGeom {
    int cls = 0;
    float p[4] = {0, 0, 0, 0};
    dxBody *body = nullptr;
    dxSpace *space = nullptr;
    dVector3 pos = {0, 0, 0, 0};
    dMatrix3 R = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    unsigned long cat = ~0ul, col = ~0ul;
    bool enabled = true;
    bool zero_sized = false;  // GEOM_ZERO_SIZED: never collides
    bool dirty = true;        // GEOM_DIRTY: set at creation, cleared by the space's collide
    int gidx = -1, didx = -1;                   // SAP GeomList / DirtyList slots
    dxGeom *snext = nullptr, *sprev = nullptr;  // simple / hash list links
    void *data = nullptr;
};

struct dxWorld {
    dxBody *firstbody = nullptr;
    dxJoint *firstjoint = nullptr;  // every live joint, newest first
    int nbodies = 0, uid_next = 1;
    std::vector<dxJoint *> free_joints;  // joints not in a group
    dVector3 gravity = {0, 0, 0, 0};
    float erp = 0.2f, cfm = 1e-5f, w = 1.3f;
    int iterations = 20;
    // dynamic iteration count adjustment (objects.h:121-203; dWorldCreate
    // defaults include/ode/objects.h:473-475): premature exit delta, maximal
    // extra iteration factor, extra iteration requirement delta
    float dyn_premature = 1e-8f, dyn_extra_factor = 1.0f, dyn_extra_delta = 1e-2f;
    float max_vel = dInfinity, min_depth = 0.f;
    // defaults new bodies take (dxWorld::adis, body_flags, dampingp,
    // max_angular_speed; objects.cpp:99-117)
    AutoDisableParams adis;
    bool body_auto_disable = false, body_lin_damping = false, body_ang_damping = false, body_max_ang = false;
    DampingParams dampingp;
    float max_angular_speed = dInfinity;
    int order = 0;
    int sweep_mode = 0;  // odxWorldSetSweepMode
    // the dynamic-iteration statistics sink (the world's own unless one is
    // attached; objects.cpp:88, objects.h:168-169)
    dWorldQuickStepIterationCount_DynamicAdjustmentStatistics own_stats{
        sizeof(dWorldQuickStepIterationCount_DynamicAdjustmentStatistics), 0, 0, 0, 0};
    dWorldQuickStepIterationCount_DynamicAdjustmentStatistics *stats = &own_stats;
    int last_sweeps = 0, last_rows = 0;
    float last_res = -1.f;
    // device buffers of the step: rows, impulses and body scratch stay on the
    // device; everything else travels in `stage`
    DevBuf rows, lambda, bscr;
    Staging stage;  // the step's inputs and outputs (one copy each way)
};

// Extra sweeps of the dynamic iteration count are capped here (the reference
// allows up to UINT_MAX); a factor asking for more gets this many.
constexpr int ODX_MAX_EXTRA_SWEEPS = 100000;

namespace {

int g_seed = 0;  // dRand state (misc.cpp:35-48)

// dRand (misc.cpp:35-48) and dRandInt (misc.cpp:78-130) on an explicit state
uint32_t rand_next(uint32_t &seed) {
    seed = (uint32_t)((1664525u * (uint64_t)seed + 1013904223u) & 0xffffffffu);
    return seed;
}
int rand_int(uint32_t &seed, int n) {
    uint32_t r = rand_next(seed), un = (uint32_t)n;
    if (un <= 0x10u) {
        r ^= (r >> 16); r ^= (r >> 8); r ^= (r >> 4);
        if (un <= 0x2u) { r ^= (r >> 2); r ^= (r >> 1); return (int)(r & (un >> 1)); }
        if (un <= 0x4u) { r ^= (r >> 2); return (int)(((r & 0x03u) * un) >> 2); }
        return (int)(((r & 0x0Fu) * un) >> 4);
    }
    if (un <= 0x100u) { r ^= (r >> 16); r ^= (r >> 8); return (int)(((r & 0xFFu) * un) >> 8); }
    if (un <= 0x10000u) { r ^= (r >> 16); return (int)(((r & 0xFFFFu) * un) >> 16); }
    return (int)(((uint64_t)r * un) >> 32);
}

void add_node(dxBody *b, dxJointNode *n) {
    n->next = b->firstjoint;
    b->firstjoint = n;
}
void remove_node(dxBody *b, dxJointNode *n) {
    for (dxJointNode **pp = &b->firstjoint; *pp; pp = &(*pp)->next)
        if (*pp == n) { *pp = n->next; n->next = nullptr; return; }
}
void detach(dxJoint *j) {
    if (j->node[0].body) remove_node(j->node[0].body, &j->node[1]);
    if (j->node[1].body) remove_node(j->node[1].body, &j->node[0]);
    j->node[0].body = j->node[1].body = nullptr;
}

// the world's joint list: head insertion, as addObjectToList (ode.cpp:54-60)
void list_joint(dxWorld *w, dxJoint *j) {
    j->wnext = w->firstjoint;
    j->wprev = nullptr;
    if (w->firstjoint) w->firstjoint->wprev = j;
    w->firstjoint = j;
    j->listed = true;
}
void unlist_joint(dxJoint *j) {
    if (!j->listed) return;
    if (j->wprev) j->wprev->wnext = j->wnext;
    else j->world->firstjoint = j->wnext;
    if (j->wnext) j->wnext->wprev = j->wprev;
    j->wnext = j->wprev = nullptr;
    j->listed = false;
}

dxJoint *new_joint(dxWorld *w, dxJointGroup *g, int type) {
    auto *j = new dxJoint();
    j->world = w;
    j->type = type;
    j->group = g;
    j->node[0].joint = j->node[1].joint = j;
    j->erp = w->erp;
    j->cfm = w->cfm;
    j->normal_cfm = w->cfm;
    j->stop_erp = w->erp;
    j->stop_cfm = w->cfm;
    if (g) g->joints.push_back(j);
    else w->free_joints.push_back(j);
    list_joint(w, j);
    return j;
}

void destroy_joint(dxJoint *j) {
    unlist_joint(j);
    detach(j);
    if (!j->group) {
        auto &v = j->world->free_joints;
        v.erase(std::remove(v.begin(), v.end(), j), v.end());
    }
    delete j;
}

bool joint_enabled(const dxJoint *j) {
    // dxJoint::isEnabled joints/joint.cpp:73-78
    if (!j->enabled || !j->node[0].body) return false;
    return j->node[0].body->mass.mass > 0 || (j->node[1].body && j->node[1].body->mass.mass > 0);
}

void geom_pose(const dxGeom *g, float *pos, float *R) {
    if (g->body) {
        for (int k = 0; k < 3; k++) pos[k] = g->body->pos[k];
        for (int k = 0; k < 12; k++) R[k] = g->body->R[k];
    } else {
        for (int k = 0; k < 3; k++) pos[k] = g->pos[k];
        for (int k = 0; k < 12; k++) R[k] = g->R[k];
    }
}

void geom_view(const dxGeom *g, float *v) {
    v[0] = (float)g->cls;
    v[1] = g->body ? (float)g->body->uid : -1.f;
    for (int k = 0; k < 4; k++) v[2 + k] = g->p[k];
    geom_pose(g, v + 6, v + 9);
    v[21] = v[22] = v[23] = 0.f;
}

// Collision session of the current dSpaceCollide (dCollide serves from it).
struct Session {
    bool active = false;
    std::vector<dxGeom *> geoms;
    std::vector<int> pairs;  // ordered, 2 per pair
    int current = -1;
    int maxc = 0;
    std::vector<float> contacts;  // [pair][4][7]
    std::vector<int> counts;
    Staging space;                // geom views, bits and the counter in; counter and AABBs out
    Staging coll;                 // pairs in; contact counts and contacts out
    PinnedBuf hpairs;             // candidate pairs read back
    DevBuf dpairs;                // candidate pairs (device)
    const float *d_gviews = nullptr;
};
// The sessions live on the heap and are never destroyed at exit: a static
// destructor would free device and pinned memory after the HIP runtime's own
// exit handlers may have run.  dCloseODE releases their memory.
Session &sess() {
    static Session *s = new Session();
    return *s;
}
Session &lone_session() {
    static Session *s = new Session();
    return *s;
}
void release_session(Session &S) {
    S.space.dev.release();
    S.space.host.release();
    S.coll.dev.release();
    S.coll.host.release();
    S.hpairs.release();
    S.dpairs.release();
    S.d_gviews = nullptr;
}

// hipEvent timing of dWorldQuickStep's device step, on request
struct KernelTiming {
    bool on = false;
    hipEvent_t ev[2] = {nullptr, nullptr};
    long long steps = 0;
    double ms = 0.0, row_sweeps = 0.0;
};
KernelTiming g_ktime;

// dMultiply0_331 / dMultiply1_331 with a 3x4 rotation: R v and R^T v
void rmul(float *r, const float *R, const float *v) {
    for (int i = 0; i < 3; i++) r[i] = R[i * 4] * v[0] + R[i * 4 + 1] * v[1] + R[i * 4 + 2] * v[2];
}
void rtmul(float *r, const float *R, const float *v) {
    for (int i = 0; i < 3; i++) r[i] = R[i] * v[0] + R[4 + i] * v[1] + R[8 + i] * v[2];
}
// qa = qb * qc with qb and / or qc conjugated (dQMultiply0..3, rotation.cpp:41-107)
void qmul_h(float *qa, const float *qb, const float *qc, bool cb, bool cc) {
    const float b1 = cb ? -qb[1] : qb[1], b2 = cb ? -qb[2] : qb[2], b3 = cb ? -qb[3] : qb[3];
    const float c1 = cc ? -qc[1] : qc[1], c2 = cc ? -qc[2] : qc[2], c3 = cc ? -qc[3] : qc[3];
    qa[0] = qb[0] * qc[0] - b1 * c1 - b2 * c2 - b3 * c3;
    qa[1] = qb[0] * c1 + b1 * qc[0] + b2 * c3 - b3 * c2;
    qa[2] = qb[0] * c2 + b2 * qc[0] + b3 * c1 - b1 * c3;
    qa[3] = qb[0] * c3 + b3 * qc[0] + b1 * c2 - b2 * c1;
}
// computeInitialRelativeRotation of hinge / slider / fixed (hinge.cpp:376-394)
void joint_qrel(dxJoint *j) {
    dxBody *b0 = j->node[0].body, *b1 = j->node[1].body;
    if (!b0) return;
    if (b1) qmul_h(j->qrel, b0->q, b1->q, true, false);
    else { j->qrel[0] = b0->q[0]; j->qrel[1] = -b0->q[1]; j->qrel[2] = -b0->q[2]; j->qrel[3] = -b0->q[3]; }
}
// getHingeAngle joints/joint.cpp:421-490
float hinge_angle_h(const dxJoint *j) {
    dxBody *b0 = j->node[0].body, *b1 = j->node[1].body;
    float qr[4];
    if (b1) { float qq[4]; qmul_h(qq, b0->q, b1->q, true, false); qmul_h(qr, qq, j->qrel, false, true); }
    else qmul_h(qr, b0->q, j->qrel, true, true);
    const float cost2 = qr[0], sint2 = std::sqrt(qr[1] * qr[1] + qr[2] * qr[2] + qr[3] * qr[3]);
    const float d = qr[1] * j->axis1[0] + qr[2] * j->axis1[1] + qr[3] * j->axis1[2];
    float theta = d >= 0 ? 2 * std::atan2(sint2, cost2) : 2 * std::atan2(sint2, -cost2);
    if (theta > (float)M_PI) theta -= (float)(2 * M_PI);
    return -theta;
}
// dJointGetSliderPosition joints/slider.cpp:45-82 (before the reverse sign of
// the single-body case, which the caller applies)
float slider_position_h(const dxJoint *j) {
    dxBody *b0 = j->node[0].body, *b1 = j->node[1].body;
    float ax1[3], q[3];
    rmul(ax1, b0->R, j->axis1);
    if (b1) {
        rmul(q, b1->R, j->offset);
        for (int i = 0; i < 3; i++) q[i] = b0->pos[i] - q[i] - b1->pos[i];
    } else {
        for (int i = 0; i < 3; i++) q[i] = b0->pos[i] - j->offset[i];
        if (j->reverse) for (int i = 0; i < 3; i++) ax1[i] = -ax1[i];
    }
    return ax1[0] * q[0] + ax1[1] * q[1] + ax1[2] * q[2];
}
// getInfo1 of hinge (hinge.cpp:58-77) and slider (slider.cpp:112-143): rows,
// and the limit state the limit / motor row uses
int hinge_slider_m(dxJoint *j) {
    int m = j->fmax > 0 ? 6 : 5;
    if (j->type == dJointTypeHinge) {
        if ((j->lostop >= -(float)M_PI || j->histop <= (float)M_PI) && j->lostop <= j->histop) {
            const float a = hinge_angle_h(j);
            if (a <= j->lostop) { j->limit = 1; j->limit_err = a - j->lostop; }
            else if (a >= j->histop) { j->limit = 2; j->limit_err = a - j->histop; }
            else j->limit = 0;
        }
    } else {
        j->limit = 0;
        if ((j->lostop > -dInfinity || j->histop < dInfinity) && j->lostop <= j->histop) {
            const float pos = slider_position_h(j);
            if (pos <= j->lostop) { j->limit = 1; j->limit_err = pos - j->lostop; }
            else if (pos >= j->histop) { j->limit = 2; j->limit_err = pos - j->histop; }
        }
    }
    return j->limit ? 6 : m;
}
// dxJointLimitMotor::set joints/joint.cpp:511-545 (limits and motors: the
// group-1 parameters)
void limot_set(dxJoint *j, int num, float v) {
    switch (num) {
    case 0: j->lostop = v; break;
    case 1: j->histop = v; break;
    case 2: j->vel = v; break;
    case 5: if (v >= 0) j->fmax = v; break;
    case 6: if (v >= 0 && v <= 1) j->fudge = v; break;
    case 7: j->bounce = v; break;
    case 8: j->normal_cfm = v; break;
    case 9: j->stop_erp = v; break;
    case 10: j->stop_cfm = v; break;
    default: break;
    }
}
float limot_get(const dxJoint *j, int num) {
    switch (num) {
    case 0: return j->lostop;
    case 1: return j->histop;
    case 2: return j->vel;
    case 5: return j->fmax;
    case 6: return j->fudge;
    case 7: return j->bounce;
    case 8: return j->normal_cfm;
    case 9: return j->stop_erp;
    case 10: return j->stop_cfm;
    default: return 0;
    }
}

bool gpu_collide_pairs(Session &S, int maxc) {
    const int np = (int)S.pairs.size() / 2;
    Staging &C = S.coll;
    C.reset();
    const size_t o_cnt = C.slot((size_t)np * sizeof(int));
    const size_t o_out = C.slot((size_t)np * ODX_MAXC * 7 * sizeof(float));
    const size_t out_end = C.used;
    const size_t o_pairs = C.slot((size_t)np * 2 * sizeof(int));
    if (!C.alloc()) return false;
    if (np > 0) {
        C.put(o_pairs, S.pairs.data(), (size_t)np * 2);
        if (!C.send(o_pairs, (size_t)np * 2 * sizeof(int))) return false;
        launch_collide_pairs(S.d_gviews, C.d<int>(o_pairs), np, maxc, C.d<float>(o_out), C.d<int>(o_cnt), g_stream);
        if (!odx_check(hipGetLastError(), "k_collide_pairs") || !C.fetch(0, out_end)) return false;
    }
    if (!odx_check(hipStreamSynchronize(g_stream), "hipStreamSynchronize")) return false;
    S.counts.assign(C.h<int>(o_cnt), C.h<int>(o_cnt) + np);
    S.contacts.assign(C.h<float>(o_out), C.h<float>(o_out) + (size_t)np * ODX_MAXC * 7);
    S.maxc = maxc;
    return true;
}


// ---------------------------------------------------------- space lists
void list_push_front(dxSpace *s, dxGeom *g) {
    g->sprev = nullptr;
    g->snext = s->first;
    if (s->first) s->first->sprev = g;
    s->first = g;
}
void list_unlink(dxSpace *s, dxGeom *g) {
    if (g->sprev) g->sprev->snext = g->snext; else s->first = g->snext;
    if (g->snext) g->snext->sprev = g->sprev;
    g->snext = g->sprev = nullptr;
}
// swap the last element of a SAP list into slot idx (collision_sapspace.cpp:343-366)
void sap_swap_remove(std::vector<dxGeom *> &list, int idx, bool geomlist) {
    const int last = (int)list.size() - 1;
    if (idx != last) {
        dxGeom *moved = list[(size_t)last];
        list[(size_t)idx] = moved;
        if (geomlist) moved->gidx = idx; else moved->didx = idx;
    }
    list.pop_back();
}
// dxSpace::add / dxSAPSpace::add: the geom's flags are left as they are
void space_add(dxSpace *s, dxGeom *g) {
    g->space = s;
    s->count++;
    if (s->kind == 0) {
        g->gidx = -1;
        g->didx = (int)s->dlist.size();
        s->dlist.push_back(g);
    } else {
        list_push_front(s, g);
    }
}
void space_remove(dxGeom *g) {
    dxSpace *s = g->space;
    if (!s) return;
    if (s->kind == 0) {
        if (g->didx >= 0) sap_swap_remove(s->dlist, g->didx, false);
        else if (g->gidx >= 0) sap_swap_remove(s->glist, g->gidx, true);
    } else {
        list_unlink(s, g);
    }
    g->gidx = g->didx = -1;
    s->count--;
    g->space = nullptr;
}
// dGeomMoved (collision_space.cpp:49-79): a clean geom of a space becomes
// dirty and is moved by its space's dirty() (collision_sapspace.cpp:368-389;
// collision_space.cpp:205-209); a geom in no space is only marked dirty.
void geom_moved(dxGeom *g) {
    dxSpace *s = g->space;
    if (!s) { g->dirty = true; return; }
    if (g->dirty) return;
    g->dirty = true;
    if (s->kind == 0) {
        if (g->didx >= 0) return;
        sap_swap_remove(s->glist, g->gidx, true);
        g->gidx = -1;
        g->didx = (int)s->dlist.size();
        s->dlist.push_back(g);
    } else {
        list_unlink(s, g);
        list_push_front(s, g);
    }
}
// every geom of a body, most recently attached first (the body's geom list,
// bodyAdd collision_kernel.h:202-206)
void body_moved(dxBody *b) {
    for (auto it = b->geoms.rbegin(); it != b->geoms.rend(); ++it) geom_moved(*it);
}
}  // namespace

extern "C" {

// ------------------------------------------------------------------ init
void dInitODE(void) { (void)dInitODE2(0); }
int dInitODE2(unsigned int) { return ensure_stream() ? 1 : 0; }
void dCloseODE(void) {
    if (!g_stream) return;
    (void)hipStreamSynchronize(g_stream);
    release_session(sess());
    release_session(lone_session());
    (void)hipStreamDestroy(g_stream);
    g_stream = nullptr;
}

// ----------------------------------------------------------------- world
dWorldID dWorldCreate(void) {
    if (!ensure_stream()) return nullptr;
    return new dxWorld();
}

void dWorldDestroy(dWorldID w) {
    if (!w) return;
    for (dxJoint *j : std::vector<dxJoint *>(w->free_joints)) destroy_joint(j);
    // group joints outlive the world: detach them from its bodies now
    while (w->firstjoint) {
        dxJoint *j = w->firstjoint;
        unlist_joint(j);
        detach(j);
        j->enabled = false;
    }
    for (dxBody *b = w->firstbody; b;) {
        dxBody *n = b->next;
        // each geom keeps the body's last pose, as dGeomSetBody(g, 0) does
    for (dxGeom *g : b->geoms) {
        geom_pose(g, g->pos, g->R);
        g->body = nullptr;
    }
        delete b;
        b = n;
    }
    delete w;
}

void dWorldSetGravity(dWorldID w, dReal x, dReal y, dReal z) { w->gravity[0] = x; w->gravity[1] = y; w->gravity[2] = z; }
void dWorldGetGravity(dWorldID w, dVector3 g) { for (int k = 0; k < 3; k++) g[k] = w->gravity[k]; }
void dWorldSetERP(dWorldID w, dReal erp) { w->erp = erp; }
dReal dWorldGetERP(dWorldID w) { return w->erp; }
void dWorldSetCFM(dWorldID w, dReal cfm) { w->cfm = cfm; }
dReal dWorldGetCFM(dWorldID w) { return w->cfm; }
void dWorldSetQuickStepNumIterations(dWorldID w, int num) { w->iterations = num; }
int dWorldGetQuickStepNumIterations(dWorldID w) { return w->iterations; }
// ode.cpp:2068-2088 (negative values clamp to 0)
void dWorldSetQuickStepDynamicIterationParameters(dWorldID w, const dReal *premature_exit_delta,
                                                  const dReal *max_num_extra_factor,
                                                  const dReal *extra_iteration_requirement_delta) {
    if (!w) return;
    if (premature_exit_delta) w->dyn_premature = std::max(*premature_exit_delta, 0.f);
    if (extra_iteration_requirement_delta) w->dyn_extra_delta = std::max(*extra_iteration_requirement_delta, 0.f);
    if (max_num_extra_factor) w->dyn_extra_factor = std::max(*max_num_extra_factor, 0.f);
}
// ode.cpp:2091-2108
void dWorldGetQuickStepDynamicIterationParameters(dWorldID w, dReal *premature_exit_delta, dReal *max_num_extra_factor,
                                                  dReal *extra_iteration_requirement_delta) {
    if (!w) return;
    if (premature_exit_delta) *premature_exit_delta = w->dyn_premature;
    if (extra_iteration_requirement_delta) *extra_iteration_requirement_delta = w->dyn_extra_delta;
    if (max_num_extra_factor) *max_num_extra_factor = w->dyn_extra_factor;
}
// dWorldAttachQuickStepDynamicIterationStatisticsSink ode.cpp:2111-2131
int dWorldAttachQuickStepDynamicIterationStatisticsSink(
    dWorldID w, dWorldQuickStepIterationCount_DynamicAdjustmentStatistics *var_stats) {
    if (!w) return 0;
    if (!var_stats) {
        w->stats = &w->own_stats;
        return 1;
    }
    if (var_stats->struct_size < sizeof(*var_stats) || var_stats->struct_size % sizeof(unsigned) != 0) return 0;
    w->stats = var_stats;
    return 1;
}
void dWorldSetQuickStepW(dWorldID w, dReal over_relaxation) { w->w = over_relaxation; }
dReal dWorldGetQuickStepW(dWorldID w) { return w->w; }
void dWorldSetContactMaxCorrectingVel(dWorldID w, dReal vel) { w->max_vel = vel; }
void dWorldSetContactSurfaceLayer(dWorldID w, dReal depth) { w->min_depth = depth; }
void odxWorldSetRowOrder(dWorldID w, int order) { w->order = (order == 1 || order == 2) ? order : 0; }
void odxWorldSetSweepMode(dWorldID w, int mode) { w->sweep_mode = (mode >= 1 && mode <= 3) ? mode : 0; }

}  // extern "C"

namespace {

// Text in the Dynamics Interchange Format (export-dif.cpp): tab indents,
// "name = value," lines, reals with 7 significant digits (a dSINGLE build),
// +-inf spelled out.
struct DifWriter {
    FILE *f;
    int indent = 1;
    void tabs() const { for (int i = 0; i < indent; i++) std::fputc('\t', f); }
    void real(float x) const {
        if (x == dInfinity) std::fputs("inf", f);
        else if (x == -dInfinity) std::fputs("-inf", f);
        else std::fprintf(f, "%.*g", 7, (double)x);
    }
    void line(const char *name) const { tabs(); std::fprintf(f, "%s\n", name); }
    void num(const char *name, int x) const { tabs(); std::fprintf(f, "%s = %d,\n", name, x); }
    void num(const char *name, float x) const {
        tabs();
        std::fprintf(f, "%s = ", name);
        real(x);
        std::fputs(",\n", f);
    }
    void vec(const char *name, const float *x, int n = 3) const {
        tabs();
        std::fprintf(f, "%s = {", name);
        for (int i = 0; i < n; i++) {
            real(x[i]);
            if (i < n - 1) std::fputc(',', f);
        }
        std::fputs("},\n", f);
    }
    void str(const char *name, const char *x) const { tabs(); std::fprintf(f, "%s = \"%s\",\n", name, x); }
    // a vector is written only when all three components are non-zero, as the reference does
    void vec_nonzero(const char *name, const float *x) const {
        if (x[0] != 0 && x[1] != 0 && x[2] != 0) vec(name, x);
    }
};

void dif_geom(DifWriter &c, const dxGeom *g) {
    if (g->cat != ~0ul) { c.tabs(); std::fprintf(c.f, "category_bits = %lu\n", g->cat); }
    if (g->col != ~0ul) { c.tabs(); std::fprintf(c.f, "collide_bits = %lu\n", g->col); }
    if (!g->enabled) c.num("disabled", 1);
    switch (g->cls) {
    case dSphereClass: c.str("type", "sphere"); c.num("radius", g->p[0]); break;
    case dBoxClass: c.str("type", "box"); c.vec("sides", g->p); break;
    case dCapsuleClass: c.str("type", "capsule"); c.num("radius", g->p[0]); c.num("length", g->p[1]); break;
    case dCylinderClass: c.str("type", "cylinder"); c.num("radius", g->p[0]); c.num("length", g->p[1]); break;
    case dPlaneClass: c.str("type", "plane"); c.vec("normal", g->p); c.num("d", g->p[3]); break;
    default: break;
    }
}

// printLimot export-dif.cpp:142-183
void dif_limot(DifWriter &c, const dxJoint *j) {
    c.line("limit = {");
    c.indent++;
    c.num("low_stop", j->lostop);
    c.num("high_stop", j->histop);
    if (j->bounce != 0) c.num("bounce", j->bounce);
    c.line("ODE = {");
    c.indent++;
    if (j->stop_erp != 0) c.num("stop_erp", j->stop_erp);
    if (j->stop_cfm != 0) c.num("stop_cfm", j->stop_cfm);
    c.indent--;
    c.line("},");
    c.indent--;
    c.line("},");
    c.line("motor = {");
    c.indent++;
    if (j->vel != 0) c.num("vel", j->vel);
    if (j->fmax != 0) c.num("fmax", j->fmax);
    c.line("ODE = {");
    c.indent++;
    if (j->fudge != 0) c.num("fudge_factor", j->fudge);
    if (j->normal_cfm != 0) c.num("normal_cfm", j->normal_cfm);
    c.indent--;
    c.line("},");
    c.indent--;
    c.line("},");
}

}  // namespace

extern "C" {

// dWorldExportDIF (export-dif.cpp:491-620): the world's parameters, its
// bodies in list order with the geoms attached to each (most recently
// attached first) and its joints (newest first), for contact and ball joints.
void dWorldExportDIF(dWorldID w, FILE *file, const char *prefix) {
    if (!w || !file) return;
    if (!prefix) prefix = "";
    DifWriter c{file};
    std::fprintf(file, "-- Dynamics Interchange Format v0.1\n\n%sworld = dynamics.world {\n", prefix);
    c.vec("gravity", w->gravity);
    c.line("ODE = {");
    c.indent++;
    c.num("ERP", w->erp);
    c.num("CFM", w->cfm);
    c.line("auto_disable = {");
    c.indent++;
    c.num("linear_threshold", w->adis.lin_thr2);
    c.num("angular_threshold", w->adis.ang_thr2);
    c.num("average_samples", (int)w->adis.average_samples);
    c.num("idle_time", w->adis.idle_time);
    c.num("idle_steps", w->adis.idle_steps);
    std::fputs("\t\t},\n\t},\n}\n", file);
    c.indent -= 3;
    int num = 0;
    std::fprintf(file, "%sbody = {}\n", prefix);
    for (dxBody *b = w->firstbody; b; b = b->next) {
        b->tag = num;
        std::fprintf(file, "%sbody[%d] = dynamics.body {\n\tworld = %sworld,\n", prefix, num, prefix);
        c.indent++;
        c.vec("pos", b->pos);
        c.vec("q", b->q, 4);
        c.vec("lvel", b->lvel);
        c.vec("avel", b->avel);
        c.num("mass", b->mass.mass);
        std::fputs("\tI = {{", file);
        for (int i = 0; i < 3; i++) {
            for (int k = 0; k < 3; k++) {
                c.real(b->mass.I[i * 4 + k]);
                if (k < 2) std::fputc(',', file);
            }
            if (i < 2) std::fputs("},{", file);
        }
        std::fputs("}},\n", file);
        c.vec_nonzero("com", b->mass.c);
        c.line("ODE = {");
        c.indent++;
        if (!b->enabled) c.num("disabled", 1);
        if (!b->gravity) c.num("no_gravity", 1);
        if (b->auto_disable) {
            c.line("auto_disable = {");
            c.indent++;
            c.num("linear_threshold", b->adis.lin_thr2);
            c.num("angular_threshold", b->adis.ang_thr2);
            c.num("average_samples", (int)b->adis.average_samples);
            c.num("idle_time", b->adis.idle_time);
            c.num("idle_steps", b->adis.idle_steps);
            c.num("time_left", b->adis_timeleft);
            c.num("steps_left", b->adis_stepsleft);
            c.indent--;
            c.line("},");
        }
        c.vec_nonzero("facc", b->facc);
        c.vec_nonzero("tacc", b->tacc);
        c.indent--;
        c.line("},");
        if (!b->geoms.empty()) {
            c.line("geometry = {");
            c.indent++;
            for (size_t k = b->geoms.size(); k-- > 0;) {
                c.line("{");
                c.indent++;
                dif_geom(c, b->geoms[k]);
                c.indent--;
                c.line("},");
            }
            c.indent--;
            c.line("},");
        }
        c.indent--;
        c.line("}");
        num++;
    }
    num = 0;
    std::fprintf(file, "%sjoint = {}\n", prefix);
    for (dxJoint *j = w->firstjoint; j; j = j->wnext) {
        c.indent++;
        const char *name = j->type == dJointTypeBall      ? "ball"
                           : j->type == dJointTypeContact ? "contact"
                           : j->type == dJointTypeHinge   ? "hinge"
                           : j->type == dJointTypeSlider  ? "slider"
                           : j->type == dJointTypeFixed   ? "fixed"
                                                          : "unknown";
        std::fprintf(file, "%sjoint[%d] = dynamics.%s_joint {\n\tworld = %sworld,\n\tbody = {", prefix, num, name, prefix);
        if (j->node[0].body) std::fprintf(file, "%sbody[%d]", prefix, j->node[0].body->tag);
        if (j->node[1].body) std::fprintf(file, ",%sbody[%d]", prefix, j->node[1].body->tag);
        std::fputs("}\n", file);
        if (j->type == dJointTypeBall) {
            c.vec("anchor1", j->anchor1);
            c.vec("anchor2", j->anchor2);
        } else if (j->type == dJointTypeContact) {
            const dContact &ct = j->contact;
            const int mode = ct.surface.mode;
            c.vec("pos", ct.geom.pos);
            c.vec("normal", ct.geom.normal);
            c.num("depth", ct.geom.depth);
            c.num("mu", ct.surface.mu);
            if (mode & dContactMu2) c.num("mu2", ct.surface.mu2);
            if (mode & dContactBounce) c.num("bounce", ct.surface.bounce);
            if (mode & dContactBounce) c.num("bounce_vel", ct.surface.bounce_vel);
            if (mode & dContactSoftERP) c.num("soft_ERP", ct.surface.soft_erp);
            if (mode & dContactSoftCFM) c.num("soft_CFM", ct.surface.soft_cfm);
            if (mode & dContactMotion1) c.num("motion1", ct.surface.motion1);
            if (mode & dContactMotion2) c.num("motion2", ct.surface.motion2);
            if (mode & dContactSlip1) c.num("slip1", ct.surface.slip1);
            if (mode & dContactSlip2) c.num("slip2", ct.surface.slip2);
            const int fa = ((mode & dContactApprox1_1) ? 1 : 0) | ((mode & dContactApprox1_2) ? 2 : 0);
            if (fa) c.num("friction_approximation", fa);
            if (mode & dContactFDir1) c.vec("fdir1", ct.fdir1);
        } else if (j->type == dJointTypeHinge) {  // printHinge :214-223
            c.vec("anchor1", j->anchor1);
            c.vec("anchor2", j->anchor2);
            c.vec("axis1", j->axis1);
            c.vec("axis2", j->axis2);
            c.vec("qrel", j->qrel, 4);
            dif_limot(c, j);
        } else if (j->type == dJointTypeSlider) {  // printSlider :226-233
            c.vec("axis1", j->axis1);
            c.vec("qrel", j->qrel, 4);
            c.vec("offset", j->offset);
            dif_limot(c, j);
        } else if (j->type == dJointTypeFixed) {  // printFixed :333-338 (three components of qrel)
            c.vec("qrel", j->qrel);
            c.vec("offset", j->offset);
        } else {
            c.line("unknown joint");
        }
        c.indent--;
        c.line("}");
        num++;
    }
}
void odxWorldGetSolveStats(dWorldID w, int *sweeps, float *residual, int *rows) {
    if (sweeps) *sweeps = w->last_sweeps;
    if (residual) *residual = w->last_res;
    if (rows) *rows = w->last_rows;
}

// ------------------------------------------------------------------ body
dBodyID dBodyCreate(dWorldID w) {
    auto *b = new dxBody();
    b->world = w;
    b->uid = w->uid_next++;
    dMassSetParameters(&b->mass, 1, 0, 0, 0, 1, 1, 1, 0, 0, 0);
    // head insertion, as addObjectToList (ode.cpp:54-60)
    b->next = w->firstbody;
    if (w->firstbody) w->firstbody->prev = b;
    w->firstbody = b;
    w->nbodies++;
    // the world's auto-disable, damping and angular speed defaults (ode.cpp:267-283)
    dBodySetAutoDisableDefaults(b);
    b->adis_stepsleft = b->adis.idle_steps;
    b->adis_timeleft = b->adis.idle_time;
    dBodySetAutoDisableAverageSamplesCount(b, b->adis.average_samples);
    dBodySetDampingDefaults(b);
    b->max_ang = w->body_max_ang;
    b->max_angular_speed = w->max_angular_speed;
    return b;
}

void dBodyDestroy(dBodyID b) {
    if (!b) return;
    dxWorld *w = b->world;
    // each geom keeps the body's last pose, as dGeomSetBody(g, 0) does
    for (dxGeom *g : b->geoms) {
        geom_pose(g, g->pos, g->R);
        g->body = nullptr;
    }
    // detach joints (ODE detaches group joints and destroys free ones' links)
    while (b->firstjoint) detach(b->firstjoint->joint);
    if (b->prev) b->prev->next = b->next;
    else w->firstbody = b->next;
    if (b->next) b->next->prev = b->prev;
    w->nbodies--;
    delete b;
}

void dBodySetData(dBodyID b, void *data) { b->data = data; }
void *dBodyGetData(dBodyID b) { return b->data; }
// the body setters notify the attached geoms (ode.cpp:344-395)
void dBodySetPosition(dBodyID b, dReal x, dReal y, dReal z) {
    b->pos[0] = x; b->pos[1] = y; b->pos[2] = z;
    body_moved(b);
}
void dBodySetRotation(dBodyID b, const dMatrix3 R) {
    std::memcpy(b->R, R, sizeof(dMatrix3));
    qfromr(b->q, b->R);
    normq(b->q);
    rfromq(b->R, b->q);
    body_moved(b);
}
void dBodySetQuaternion(dBodyID b, const dQuaternion q) {
    for (int k = 0; k < 4; k++) b->q[k] = q[k];
    normq(b->q);
    rfromq(b->R, b->q);
    body_moved(b);
}
void dBodySetLinearVel(dBodyID b, dReal x, dReal y, dReal z) { b->lvel[0] = x; b->lvel[1] = y; b->lvel[2] = z; }
void dBodySetAngularVel(dBodyID b, dReal x, dReal y, dReal z) { b->avel[0] = x; b->avel[1] = y; b->avel[2] = z; }
const dReal *dBodyGetPosition(dBodyID b) { return b->pos; }
const dReal *dBodyGetRotation(dBodyID b) { return b->R; }
const dReal *dBodyGetQuaternion(dBodyID b) { return b->q; }
const dReal *dBodyGetLinearVel(dBodyID b) { return b->lvel; }
const dReal *dBodyGetAngularVel(dBodyID b) { return b->avel; }
void dBodySetMass(dBodyID b, const dMass *m) { b->mass = *m; }
void dBodyGetMass(dBodyID b, dMass *m) { *m = b->mass; }
void dBodyAddForce(dBodyID b, dReal fx, dReal fy, dReal fz) { b->facc[0] += fx; b->facc[1] += fy; b->facc[2] += fz; }
void dBodyAddTorque(dBodyID b, dReal fx, dReal fy, dReal fz) { b->tacc[0] += fx; b->tacc[1] += fy; b->tacc[2] += fz; }
void dBodyEnable(dBodyID b) {
    b->enabled = true;
    b->adis_stepsleft = b->adis.idle_steps;
    b->adis_timeleft = b->adis.idle_time;
}
void dBodyDisable(dBodyID b) { b->enabled = false; }
int dBodyIsEnabled(dBodyID b) { return b->enabled ? 1 : 0; }

// ---- auto-disable (ode.cpp:891-1021, 1889-1969): thresholds squared on set,
// square roots on get
dReal dBodyGetAutoDisableLinearThreshold(dBodyID b) { return std::sqrt(b->adis.lin_thr2); }
void dBodySetAutoDisableLinearThreshold(dBodyID b, dReal t) { b->adis.lin_thr2 = t * t; }
dReal dBodyGetAutoDisableAngularThreshold(dBodyID b) { return std::sqrt(b->adis.ang_thr2); }
void dBodySetAutoDisableAngularThreshold(dBodyID b, dReal t) { b->adis.ang_thr2 = t * t; }
int dBodyGetAutoDisableAverageSamplesCount(dBodyID b) { return (int)b->adis.average_samples; }
void dBodySetAutoDisableAverageSamplesCount(dBodyID b, unsigned int n) {
    b->adis.average_samples = n;
    b->avg_lvel.assign((size_t)n * 3, 0.f);
    b->avg_avel.assign((size_t)n * 3, 0.f);
    b->avg_counter = 0;
    b->avg_ready = 0;
}
int dBodyGetAutoDisableSteps(dBodyID b) { return b->adis.idle_steps; }
void dBodySetAutoDisableSteps(dBodyID b, int steps) { b->adis.idle_steps = steps; }
dReal dBodyGetAutoDisableTime(dBodyID b) { return b->adis.idle_time; }
void dBodySetAutoDisableTime(dBodyID b, dReal t) { b->adis.idle_time = t; }
int dBodyGetAutoDisableFlag(dBodyID b) { return b->auto_disable ? 1 : 0; }
void dBodySetAutoDisableFlag(dBodyID b, int on) {
    if (!on) {
        // clearing the flag also enables the body and takes the world's
        // step/time counts and sample count again (ode.cpp:992-1010)
        b->auto_disable = false;
        b->enabled = true;
        b->adis.idle_steps = b->world->adis.idle_steps;
        b->adis.idle_time = b->world->adis.idle_time;
        dBodySetAutoDisableAverageSamplesCount(b, b->world->adis.average_samples);
    } else {
        b->auto_disable = true;
    }
}
void dBodySetAutoDisableDefaults(dBodyID b) {
    b->adis = b->world->adis;
    dBodySetAutoDisableFlag(b, b->world->body_auto_disable ? 1 : 0);
}

dReal dWorldGetAutoDisableLinearThreshold(dWorldID w) { return std::sqrt(w->adis.lin_thr2); }
void dWorldSetAutoDisableLinearThreshold(dWorldID w, dReal t) { w->adis.lin_thr2 = t * t; }
dReal dWorldGetAutoDisableAngularThreshold(dWorldID w) { return std::sqrt(w->adis.ang_thr2); }
void dWorldSetAutoDisableAngularThreshold(dWorldID w, dReal t) { w->adis.ang_thr2 = t * t; }
int dWorldGetAutoDisableAverageSamplesCount(dWorldID w) { return (int)w->adis.average_samples; }
void dWorldSetAutoDisableAverageSamplesCount(dWorldID w, unsigned int n) { w->adis.average_samples = n; }
int dWorldGetAutoDisableSteps(dWorldID w) { return w->adis.idle_steps; }
void dWorldSetAutoDisableSteps(dWorldID w, int steps) { w->adis.idle_steps = steps; }
dReal dWorldGetAutoDisableTime(dWorldID w) { return w->adis.idle_time; }
void dWorldSetAutoDisableTime(dWorldID w, dReal t) { w->adis.idle_time = t; }
int dWorldGetAutoDisableFlag(dWorldID w) { return w->body_auto_disable ? 1 : 0; }
void dWorldSetAutoDisableFlag(dWorldID w, int on) { w->body_auto_disable = on != 0; }

// ---- damping and angular speed cap (ode.cpp:1024-1110, 1971-2050): a
// scale of 0 turns damping off, an infinite cap turns the cap off
dReal dBodyGetLinearDamping(dBodyID b) { return b->dampingp.lin_scale; }
void dBodySetLinearDamping(dBodyID b, dReal scale) { b->lin_damping = scale != 0; b->dampingp.lin_scale = scale; }
dReal dBodyGetAngularDamping(dBodyID b) { return b->dampingp.ang_scale; }
void dBodySetAngularDamping(dBodyID b, dReal scale) { b->ang_damping = scale != 0; b->dampingp.ang_scale = scale; }
void dBodySetDamping(dBodyID b, dReal lin, dReal ang) {
    dBodySetLinearDamping(b, lin);
    dBodySetAngularDamping(b, ang);
}
dReal dBodyGetLinearDampingThreshold(dBodyID b) { return std::sqrt(b->dampingp.lin_thr2); }
void dBodySetLinearDampingThreshold(dBodyID b, dReal t) { b->dampingp.lin_thr2 = t * t; }
dReal dBodyGetAngularDampingThreshold(dBodyID b) { return std::sqrt(b->dampingp.ang_thr2); }
void dBodySetAngularDampingThreshold(dBodyID b, dReal t) { b->dampingp.ang_thr2 = t * t; }
void dBodySetDampingDefaults(dBodyID b) {
    b->dampingp = b->world->dampingp;
    b->lin_damping = b->world->body_lin_damping;
    b->ang_damping = b->world->body_ang_damping;
}
dReal dBodyGetMaxAngularSpeed(dBodyID b) { return b->max_angular_speed; }
void dBodySetMaxAngularSpeed(dBodyID b, dReal s) { b->max_ang = s < dInfinity; b->max_angular_speed = s; }

dReal dWorldGetLinearDampingThreshold(dWorldID w) { return std::sqrt(w->dampingp.lin_thr2); }
void dWorldSetLinearDampingThreshold(dWorldID w, dReal t) { w->dampingp.lin_thr2 = t * t; }
dReal dWorldGetAngularDampingThreshold(dWorldID w) { return std::sqrt(w->dampingp.ang_thr2); }
void dWorldSetAngularDampingThreshold(dWorldID w, dReal t) { w->dampingp.ang_thr2 = t * t; }
dReal dWorldGetLinearDamping(dWorldID w) { return w->dampingp.lin_scale; }
void dWorldSetLinearDamping(dWorldID w, dReal scale) { w->body_lin_damping = scale != 0; w->dampingp.lin_scale = scale; }
dReal dWorldGetAngularDamping(dWorldID w) { return w->dampingp.ang_scale; }
void dWorldSetAngularDamping(dWorldID w, dReal scale) { w->body_ang_damping = scale != 0; w->dampingp.ang_scale = scale; }
void dWorldSetDamping(dWorldID w, dReal lin, dReal ang) {
    dWorldSetLinearDamping(w, lin);
    dWorldSetAngularDamping(w, ang);
}
dReal dWorldGetMaxAngularSpeed(dWorldID w) { return w->max_angular_speed; }
void dWorldSetMaxAngularSpeed(dWorldID w, dReal s) { w->body_max_ang = s < dInfinity; w->max_angular_speed = s; }
void dBodySetGravityMode(dBodyID b, int mode) { b->gravity = mode != 0; }
void dBodySetGyroscopicMode(dBodyID b, int enabled) { b->gyro = enabled != 0; }

// ---------------------------------------------------------------- joints
dJointID dJointCreateBall(dWorldID w, dJointGroupID g) { return new_joint(w, g, dJointTypeBall); }
dJointID dJointCreateContact(dWorldID w, dJointGroupID g, const dContact *c) {
    dxJoint *j = new_joint(w, g, dJointTypeContact);
    j->contact = *c;
    return j;
}
void dJointDestroy(dJointID j) {
    if (!j) return;
    if (j->group) {  // group joints are only detached; the group owns the memory
        unlist_joint(j);
        detach(j);
        j->enabled = false;
        return;
    }
    destroy_joint(j);
}
dJointGroupID dJointGroupCreate(int) { return new dxJointGroup(); }
void dJointGroupDestroy(dJointGroupID g) {
    if (!g) return;
    dJointGroupEmpty(g);
    delete g;
}
void dJointGroupEmpty(dJointGroupID g) {
    // detached from the most recently created (ode.cpp:1325-1366)
    for (size_t i = g->joints.size(); i-- > 0;) {
        unlist_joint(g->joints[i]);
        detach(g->joints[i]);
        delete g->joints[i];
    }
    g->joints.clear();
}
void dJointAttach(dJointID j, dBodyID body1, dBodyID body2) {
    // ode.cpp:1383-1439
    if (j->node[0].body || j->node[1].body) detach(j);
    if (!body1) { body1 = body2; body2 = nullptr; j->reverse = true; }
    else j->reverse = false;
    j->node[0].body = body1;
    j->node[1].body = body2;
    if (body1) add_node(body1, &j->node[1]);
    if (body2) add_node(body2, &j->node[0]);
}
dJointType dJointGetType(dJointID j) { return (dJointType)j->type; }
dBodyID dJointGetBody(dJointID j, int index) {
    // reversed joints report their bodies in the caller's order
    if (index == 0) return j->reverse ? j->node[1].body : j->node[0].body;
    if (index == 1) return j->reverse ? j->node[0].body : j->node[1].body;
    return nullptr;
}
void dJointSetFeedback(dJointID j, dJointFeedback *fb) { j->feedback = fb; }
void dJointSetBallAnchor(dJointID j, dReal x, dReal y, dReal z) {
    // setAnchors joints/joint.cpp:289-319
    dxBody *b0 = j->node[0].body;
    if (!b0) return;
    float q[3] = {x - b0->pos[0], y - b0->pos[1], z - b0->pos[2]};
    for (int i = 0; i < 3; i++) j->anchor1[i] = b0->R[i] * q[0] + b0->R[4 + i] * q[1] + b0->R[8 + i] * q[2];
    dxBody *b1 = j->node[1].body;
    if (b1) {
        float r[3] = {x - b1->pos[0], y - b1->pos[1], z - b1->pos[2]};
        for (int i = 0; i < 3; i++) j->anchor2[i] = b1->R[i] * r[0] + b1->R[4 + i] * r[1] + b1->R[8 + i] * r[2];
    } else {
        j->anchor2[0] = x; j->anchor2[1] = y; j->anchor2[2] = z;
    }
    j->anchor1[3] = j->anchor2[3] = 0;
}
void dJointSetBallAnchor2(dJointID j, dReal x, dReal y, dReal z) {
    j->anchor2[0] = x; j->anchor2[1] = y; j->anchor2[2] = z; j->anchor2[3] = 0;
}

// ---- hinge, slider, fixed (joints/hinge.cpp, slider.cpp, fixed.cpp)
dJointID dJointCreateHinge(dWorldID w, dJointGroupID g) { return new_joint(w, g, dJointTypeHinge); }
dJointID dJointCreateSlider(dWorldID w, dJointGroupID g) { return new_joint(w, g, dJointTypeSlider); }
dJointID dJointCreateFixed(dWorldID w, dJointGroupID g) { return new_joint(w, g, dJointTypeFixed); }

void dJointSetHingeAnchor(dJointID j, dReal x, dReal y, dReal z) {
    dJointSetBallAnchor(j, x, y, z);  // setAnchors
    joint_qrel(j);
}
// setAxes joints/joint.cpp:325-360 (the direction normalised first)
static void set_axes(dxJoint *j, float x, float y, float z, float *axis1, float *axis2) {
    dxBody *b0 = j->node[0].body;
    if (!b0) return;
    float q[3] = {x, y, z};
    const float l = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    if (l > 0) for (int i = 0; i < 3; i++) q[i] /= l;
    if (axis1) { rtmul(axis1, b0->R, q); axis1[3] = 0; }
    if (axis2) {
        if (j->node[1].body) rtmul(axis2, j->node[1].body->R, q);
        else { axis2[0] = x; axis2[1] = y; axis2[2] = z; }
        axis2[3] = 0;
    }
}
void dJointSetHingeAxis(dJointID j, dReal x, dReal y, dReal z) {
    set_axes(j, x, y, z, j->axis1, j->axis2);
    joint_qrel(j);
}
void dJointGetHingeAnchor(dJointID j, dVector3 r) {
    // getAnchor / getAnchor2 (joint.cpp:363-392), swapped for a reversed joint
    dxBody *b0 = j->node[0].body, *b1 = j->node[1].body;
    const bool second = j->reverse;
    if (!second) {
        if (!b0) return;
        rmul(r, b0->R, j->anchor1);
        for (int i = 0; i < 3; i++) r[i] += b0->pos[i];
    } else if (b1) {
        rmul(r, b1->R, j->anchor2);
        for (int i = 0; i < 3; i++) r[i] += b1->pos[i];
    } else {
        for (int i = 0; i < 3; i++) r[i] = j->anchor2[i];
    }
}
void dJointGetHingeAxis(dJointID j, dVector3 r) {
    if (j->node[0].body) rmul(r, j->node[0].body->R, j->axis1);
}
void dJointSetHingeParam(dJointID j, int parameter, dReal value) { limot_set(j, parameter, value); }
dReal dJointGetHingeParam(dJointID j, int parameter) { return limot_get(j, parameter); }
dReal dJointGetHingeAngle(dJointID j) {
    if (!j->node[0].body) return 0;
    const float a = hinge_angle_h(j);
    return j->reverse ? -a : a;
}
dReal dJointGetHingeAngleRate(dJointID j) {
    dxBody *b0 = j->node[0].body, *b1 = j->node[1].body;
    if (!b0) return 0;
    float ax[3];
    rmul(ax, b0->R, j->axis1);
    float rate = ax[0] * b0->avel[0] + ax[1] * b0->avel[1] + ax[2] * b0->avel[2];
    if (b1) rate -= ax[0] * b1->avel[0] + ax[1] * b1->avel[1] + ax[2] * b1->avel[2];
    return j->reverse ? -rate : rate;
}

void dJointSetSliderAxis(dJointID j, dReal x, dReal y, dReal z) {
    set_axes(j, x, y, z, j->axis1, nullptr);
    // computeOffset slider.cpp:406-423
    dxBody *b0 = j->node[0].body, *b1 = j->node[1].body;
    if (b1) {
        float c[3] = {b0->pos[0] - b1->pos[0], b0->pos[1] - b1->pos[1], b0->pos[2] - b1->pos[2]};
        rtmul(j->offset, b1->R, c);
    } else if (b0) {
        for (int i = 0; i < 3; i++) j->offset[i] = b0->pos[i];
    }
    joint_qrel(j);
}
void dJointGetSliderAxis(dJointID j, dVector3 r) {
    if (j->node[0].body) rmul(r, j->node[0].body->R, j->axis1);
}
void dJointSetSliderParam(dJointID j, int parameter, dReal value) { limot_set(j, parameter, value); }
dReal dJointGetSliderParam(dJointID j, int parameter) { return limot_get(j, parameter); }
dReal dJointGetSliderPosition(dJointID j) { return j->node[0].body ? slider_position_h(j) : 0; }
dReal dJointGetSliderPositionRate(dJointID j) {
    dxBody *b0 = j->node[0].body, *b1 = j->node[1].body;
    if (!b0) return 0;
    float ax[3];
    rmul(ax, b0->R, j->axis1);
    float rate = ax[0] * b0->lvel[0] + ax[1] * b0->lvel[1] + ax[2] * b0->lvel[2];
    if (b1) return rate - (ax[0] * b1->lvel[0] + ax[1] * b1->lvel[1] + ax[2] * b1->lvel[2]);
    return j->reverse ? -rate : rate;
}

void dJointSetFixed(dJointID j) {
    // fixed.cpp:108-136: body 1's centre relative to body 2 in body 1's frame
    dxBody *b0 = j->node[0].body, *b1 = j->node[1].body;
    if (b0) {
        if (b1) {
            float ofs[3] = {b0->pos[0] - b1->pos[0], b0->pos[1] - b1->pos[1], b0->pos[2] - b1->pos[2]};
            rtmul(j->offset, b0->R, ofs);
        } else {
            for (int i = 0; i < 3; i++) j->offset[i] = b0->pos[i];
        }
    }
    joint_qrel(j);
}
void dJointSetFixedParam(dJointID j, int parameter, dReal value) {
    // fixed.cpp:138-150: dParamCFM, dParamERP
    if (parameter == 8) j->cfm = value;
    else if (parameter == 13) j->erp = value;
}
dReal dJointGetFixedParam(dJointID j, int parameter) {
    return parameter == 8 ? j->cfm : parameter == 13 ? j->erp : 0;
}
int dAreConnected(dBodyID a, dBodyID b) {
    for (dxJointNode *n = a->firstjoint; n; n = n->next)
        if (n->body == b) return 1;
    return 0;
}
int dAreConnectedExcluding(dBodyID a, dBodyID b, int joint_type) {
    for (dxJointNode *n = a->firstjoint; n; n = n->next)
        if (n->body == b && n->joint->type != joint_type) return 1;
    return 0;
}

// ------------------------------------------------------------------ mass
void dMassSetZero(dMass *m) { std::memset(m, 0, sizeof(dMass)); }
void dMassSetParameters(dMass *m, dReal themass, dReal cgx, dReal cgy, dReal cgz, dReal I11, dReal I22, dReal I33,
                        dReal I12, dReal I13, dReal I23) {
    dMassSetZero(m);
    m->mass = themass;
    m->c[0] = cgx; m->c[1] = cgy; m->c[2] = cgz;
    m->I[0] = I11; m->I[5] = I22; m->I[10] = I33;
    m->I[1] = m->I[4] = I12;
    m->I[2] = m->I[8] = I13;
    m->I[6] = m->I[9] = I23;
}
void dMassSetSphereTotal(dMass *m, dReal total, dReal r) {
    float ii = 0.4f * total * r * r;
    dMassSetParameters(m, total, 0, 0, 0, ii, ii, ii, 0, 0, 0);
}
void dMassSetSphere(dMass *m, dReal density, dReal r) {
    dMassSetSphereTotal(m, (dReal)((4.0 / 3.0) * M_PI * r * r * r * density), r);
}
void dMassSetBoxTotal(dMass *m, dReal total, dReal lx, dReal ly, dReal lz) {
    dMassSetParameters(m, total, 0, 0, 0, total / 12.f * (ly * ly + lz * lz), total / 12.f * (lx * lx + lz * lz),
                       total / 12.f * (lx * lx + ly * ly), 0, 0, 0);
}
void dMassSetBox(dMass *m, dReal density, dReal lx, dReal ly, dReal lz) {
    dMassSetBoxTotal(m, lx * ly * lz * density, lx, ly, lz);
}
void dMassSetCapsule(dMass *m, dReal density, int direction, dReal r, dReal len) {
    // mass.cpp:130-152
    dMassSetZero(m);
    float M1 = (float)(M_PI * r * r * len * density);
    float M2 = (float)((4.0 / 3.0) * M_PI * r * r * r * density);
    m->mass = M1 + M2;
    float Ia = M1 * (0.25f * r * r + (1.f / 12.f) * len * len) + M2 * (0.4f * r * r + 0.375f * r * len + 0.25f * len * len);
    float Ib = (M1 * 0.5f + M2 * 0.4f) * r * r;
    m->I[0] = m->I[5] = m->I[10] = Ia;
    if (direction >= 1 && direction <= 3) m->I[(direction - 1) * 5] = Ib;
}
// dMassSetCylinder / dMassSetCylinderTotal mass.cpp:173-198
void dMassSetCylinderTotal(dMass *m, dReal total, int direction, dReal r, dReal len) {
    dMassSetZero(m);
    const float r2 = r * r;
    m->mass = total;
    const float I = total * (0.25f * r2 + (1.0f / 12.0f) * len * len);
    m->I[0] = m->I[5] = m->I[10] = I;
    if (direction >= 1 && direction <= 3) m->I[(direction - 1) * 5] = total * 0.5f * r2;
}
void dMassSetCylinder(dMass *m, dReal density, int direction, dReal r, dReal len) {
    dMassSetCylinderTotal(m, (float)(M_PI * r * r * len * density), direction, r, len);
}
void dMassSetCapsuleTotal(dMass *m, dReal total, int direction, dReal r, dReal len) {
    dMassSetCapsule(m, 1.f, direction, r, len);
    float s = total / m->mass;
    m->mass = total;
    for (int k = 0; k < 12; k++) m->I[k] *= s;
}

// --------------------------------------------------------------- geoms
static dxGeom *new_geom(dSpaceID space, int cls) {
    auto *g = new dxGeom();
    g->cls = cls;
    if (space) space_add(space, g);
    return g;
}
dGeomID dCreateSphere(dSpaceID space, dReal radius) {
    dxGeom *g = new_geom(space, dSphereClass);
    g->p[0] = radius;
    g->zero_sized = radius == 0;  // sphere.cpp:55
    return g;
}
dGeomID dCreateBox(dSpaceID space, dReal lx, dReal ly, dReal lz) {
    dxGeom *g = new_geom(space, dBoxClass);
    g->p[0] = lx; g->p[1] = ly; g->p[2] = lz;
    g->zero_sized = lx == 0 || ly == 0 || lz == 0;  // box.cpp:56
    return g;
}
// make_sure_plane_normal_has_unit_length (plane.cpp:49-65)
dGeomID dCreatePlane(dSpaceID space, dReal a, dReal b, dReal c, dReal d) {
    dxGeom *g = new_geom(space, dPlaneClass);
    float l = a * a + b * b + c * c;
    if (l > 0) {
        l = 1.0f / std::sqrt(l);
        a *= l; b *= l; c *= l; d *= l;
    } else {
        a = 1; b = 0; c = 0; d = 0;
    }
    g->p[0] = a; g->p[1] = b; g->p[2] = c; g->p[3] = d;
    return g;
}
dGeomID dCreateCapsule(dSpaceID space, dReal radius, dReal length) {
    dxGeom *g = new_geom(space, dCapsuleClass);
    g->p[0] = radius; g->p[1] = length;
    g->zero_sized = radius == 0;  // capsule.cpp:56 (a zero length is not zero-sized)
    return g;
}
dGeomID dCreateCylinder(dSpaceID space, dReal radius, dReal length) {
    dxGeom *g = new_geom(space, dCylinderClass);
    g->p[0] = radius; g->p[1] = length;
    g->zero_sized = radius == 0 || length == 0;  // cylinder.cpp:54
    return g;
}
void dGeomCylinderSetParams(dGeomID g, dReal radius, dReal length) {
    g->p[0] = radius; g->p[1] = length;
    g->zero_sized = radius == 0 || length == 0;
    geom_moved(g);  // cylinder.cpp:84-93
}
void dGeomCylinderGetParams(dGeomID g, dReal *radius, dReal *length) { *radius = g->p[0]; *length = g->p[1]; }
void dGeomDestroy(dGeomID g) {
    if (!g) return;
    space_remove(g);
    if (g->body) {
        auto &v = g->body->geoms;
        v.erase(std::remove(v.begin(), v.end(), g), v.end());
    }
    delete g;
}
void dGeomSetData(dGeomID g, void *data) { g->data = data; }
void *dGeomGetData(dGeomID g) { return g->data; }
// dGeomSetBody collision_kernel.cpp:511-551: attaching moves the geom;
// detaching keeps the body's last pose and does not
void dGeomSetBody(dGeomID g, dBodyID b) {
    if (b) {
        if (g->body != b) {
            if (g->body) {
                auto &v = g->body->geoms;
                v.erase(std::remove(v.begin(), v.end(), g), v.end());
            }
            g->body = b;
            b->geoms.push_back(g);
        }
        geom_moved(g);
    } else if (g->body) {
        geom_pose(g, g->pos, g->R);
        auto &v = g->body->geoms;
        v.erase(std::remove(v.begin(), v.end(), g), v.end());
        g->body = nullptr;
    }
}
dBodyID dGeomGetBody(dGeomID g) { return g->body; }
void dGeomSetPosition(dGeomID g, dReal x, dReal y, dReal z) {
    if (g->body) { dBodySetPosition(g->body, x, y, z); return; }
    g->pos[0] = x; g->pos[1] = y; g->pos[2] = z;
    geom_moved(g);
}
void dGeomSetRotation(dGeomID g, const dMatrix3 R) {
    if (g->body) { dBodySetRotation(g->body, R); return; }
    std::memcpy(g->R, R, sizeof(dMatrix3));
    geom_moved(g);
}
void dGeomSetQuaternion(dGeomID g, const dQuaternion q) {
    if (g->body) { dBodySetQuaternion(g->body, q); return; }
    float qq[4] = {q[0], q[1], q[2], q[3]};
    normq(qq);
    rfromq(g->R, qq);
    geom_moved(g);
}
const dReal *dGeomGetPosition(dGeomID g) { return g->body ? g->body->pos : g->pos; }
dSpaceID dGeomGetSpace(dGeomID g) { return g->space; }
int dGeomGetClass(dGeomID g) { return g->cls; }
void dGeomSetCategoryBits(dGeomID g, unsigned long bits) { g->cat = bits; }
void dGeomSetCollideBits(dGeomID g, unsigned long bits) { g->col = bits; }
void dGeomEnable(dGeomID g) { g->enabled = true; }
void dGeomDisable(dGeomID g) { g->enabled = false; }
dReal dGeomSphereGetRadius(dGeomID g) { return g->p[0]; }
void dGeomBoxGetLengths(dGeomID g, dVector3 r) { r[0] = g->p[0]; r[1] = g->p[1]; r[2] = g->p[2]; }
void dGeomPlaneGetParams(dGeomID g, dVector4 r) { for (int k = 0; k < 4; k++) r[k] = g->p[k]; }
void dGeomCapsuleGetParams(dGeomID g, dReal *radius, dReal *length) { *radius = g->p[0]; *length = g->p[1]; }

void dGeomGetAABB(dGeomID g, dReal aabb[6]) {
    // host restatement of the computeAABB family for the query API
    float pos[3], R[12];
    geom_pose(g, pos, R);
    const float inf = dInfinity;
    switch (g->cls) {
    case dSphereClass:
        for (int k = 0; k < 3; k++) { aabb[2 * k] = pos[k] - g->p[0]; aabb[2 * k + 1] = pos[k] + g->p[0]; }
        break;
    case dBoxClass:
        for (int k = 0; k < 3; k++) {
            float r = 0.5f * (std::fabs(R[4 * k] * g->p[0]) + std::fabs(R[4 * k + 1] * g->p[1]) + std::fabs(R[4 * k + 2] * g->p[2]));
            aabb[2 * k] = pos[k] - r; aabb[2 * k + 1] = pos[k] + r;
        }
        break;
    case dCapsuleClass:
        for (int k = 0; k < 3; k++) {
            float r = std::fabs(R[4 * k + 2] * g->p[1]) * 0.5f + g->p[0];
            aabb[2 * k] = pos[k] - r; aabb[2 * k + 1] = pos[k] + r;
        }
        break;
    case dCylinderClass:
        for (int k = 0; k < 3; k++) {
            const float a = R[4 * k + 2], o = 1.0f - a * a;
            float r = std::fabs(a * g->p[1] * 0.5f) + g->p[0] * std::sqrt(0.f > o ? 0.f : o);
            aabb[2 * k] = pos[k] - r; aabb[2 * k + 1] = pos[k] + r;
        }
        break;
    default:
        for (int k = 0; k < 3; k++) { aabb[2 * k] = -inf; aabb[2 * k + 1] = inf; }
        if (g->p[1] == 0 && g->p[2] == 0) { aabb[0] = g->p[0] > 0 ? -inf : -g->p[3]; aabb[1] = g->p[0] > 0 ? g->p[3] : inf; }
        else if (g->p[0] == 0 && g->p[2] == 0) { aabb[2] = g->p[1] > 0 ? -inf : -g->p[3]; aabb[3] = g->p[1] > 0 ? g->p[3] : inf; }
        else if (g->p[0] == 0 && g->p[1] == 0) { aabb[4] = g->p[2] > 0 ? -inf : -g->p[3]; aabb[5] = g->p[2] > 0 ? g->p[3] : inf; }
    }
}

// ---------------------------------------------------------------- spaces
dSpaceID dSimpleSpaceCreate(dSpaceID parent) { auto *s = new dxSpace(); s->kind = 2; s->parent = parent; return s; }
dSpaceID dHashSpaceCreate(dSpaceID parent) { auto *s = new dxSpace(); s->kind = 1; s->parent = parent; return s; }
// axis order: ax0 = (axisorder & 3), ax1 = (axisorder >> 2) & 3, ax2 =
// (axisorder >> 4) & 3 (collision_sapspace.cpp:291-293)
dSpaceID dSweepAndPruneSpaceCreate(dSpaceID parent, int axisorder) {
    auto *s = new dxSpace();
    s->kind = 0;
    s->parent = parent;
    s->ax0 = (axisorder & 3) << 1;
    s->ax1 = ((axisorder >> 2) & 3) << 1;
    s->ax2 = ((axisorder >> 4) & 3) << 1;
    return s;
}
void dSpaceDestroy(dSpaceID s) {
    if (!s) return;
    std::vector<dxGeom *> gs;
    for (dxGeom *g = s->first; g; g = g->snext) gs.push_back(g);
    gs.insert(gs.end(), s->dlist.begin(), s->dlist.end());
    gs.insert(gs.end(), s->glist.begin(), s->glist.end());
    for (dxGeom *g : gs) {
        if (s->cleanup) dGeomDestroy(g);
        else space_remove(g);
    }
    delete s;
}
void dHashSpaceSetLevels(dSpaceID s, int minlevel, int maxlevel) { s->minlevel = minlevel; s->maxlevel = maxlevel; }
void dSpaceSetCleanup(dSpaceID s, int mode) { s->cleanup = mode; }
void dSpaceAdd(dSpaceID s, dGeomID g) {
    space_remove(g);
    space_add(s, g);
}
void dSpaceRemove(dSpaceID s, dGeomID g) {
    if (g && g->space == s) space_remove(g);
}
int dSpaceGetNumGeoms(dSpaceID s) { return s->count; }

// dSpaceCollide (collision_space.cpp:779-784): the space's geoms in its own
// order go to the device once; k_space_collide computes their AABBs and the
// candidate pairs in the order the space reports them; the pairs with an
// infinite (sweep and prune) or big (hash) AABB follow in the order the
// space checks them; then the callback runs once per pair.
void dSpaceCollide(dSpaceID space, void *data, dNearCallback *callback) {
    if (!space || !callback || !ensure_stream()) return;
    Session &S = sess();
    S.geoms.clear();
    S.pairs.clear();
    S.maxc = 0;
    const int kind = space->kind;
    // cleanGeoms: SAP appends its dirty list (collision_sapspace.cpp:391-421);
    // the lists clear the dirty flags of their leading dirty geoms
    // (collision_space.cpp:235-247, 394-409).  A hash space of fewer than two
    // geoms returns before cleaning (:425).
    if (kind == 0) {
        for (dxGeom *g : space->dlist) {
            g->dirty = false;
            g->didx = -1;
            g->gidx = (int)space->glist.size();
            space->glist.push_back(g);
        }
        space->dlist.clear();
    } else if (kind == 2 || space->count >= 2) {
        for (dxGeom *g = space->first; g && g->dirty; g = g->snext) g->dirty = false;
    }
    std::vector<dxGeom *> order;
    if (kind == 0) order = space->glist;
    else if (kind == 2 || space->count >= 2)
        for (dxGeom *g = space->first; g; g = g->snext) order.push_back(g);
    int nsort = 0;
    if (kind == 0) {
        // TmpGeomList (finite on the first axis) then TmpInfGeomList (:437-449)
        std::vector<dxGeom *> inf;
        for (dxGeom *g : order) {
            if (!g->enabled || g->zero_sized) continue;
            float ab[6];
            dGeomGetAABB(g, ab);
            if (ab[space->ax0 + 1] == dInfinity) inf.push_back(g);
            else S.geoms.push_back(g);
        }
        nsort = (int)S.geoms.size();
        S.geoms.insert(S.geoms.end(), inf.begin(), inf.end());
    } else {
        for (dxGeom *g : order)
            if (g->enabled && !g->zero_sized) S.geoms.push_back(g);
        nsort = (int)S.geoms.size();
    }
    const int n = (int)S.geoms.size();
    if (n > space_max_geoms()) {
        odx_message(4, "dSpaceCollide: %d enabled geoms exceed this build's limit of %d per space", n,
                    space_max_geoms());
        return;
    }
    if (n >= 1) {
        std::vector<float> gv((size_t)n * GVIEW_STRIDE);
        std::vector<unsigned long long> bits((size_t)n * 2);
        for (int i = 0; i < n; i++) {
            geom_view(S.geoms[i], &gv[(size_t)i * GVIEW_STRIDE]);
            bits[2 * i] = (unsigned long long)S.geoms[i]->cat;
            bits[2 * i + 1] = (unsigned long long)S.geoms[i]->col;
        }
        // the sort keeps its ranks while the count of sorted geoms stays the
        // same (RaixSortContext::ReallocateRanksIfNecessary, :111-125); no
        // sort runs without finite geoms
        const bool sorts = kind == 0 && nsort > 0;
        if (sorts) {
            if ((size_t)nsort != space->ranks_used) {
                space->ranks_valid = false;
                space->ranks_used = (size_t)nsort;
            }
            if (!space->ranks[0].ensure((size_t)nsort * sizeof(unsigned)) ||
                !space->ranks[1].ensure((size_t)nsort * sizeof(unsigned)))
                return;
        }
        Staging &G = S.space;
        G.reset();
        const size_t o_info = G.slot(4 * sizeof(int)), o_ab = G.slot((size_t)n * 6 * sizeof(float));
        const size_t out_end = G.used;
        const size_t o_gv = G.slot(gv.size() * sizeof(float));
        const size_t o_bits = G.slot(bits.size() * sizeof(unsigned long long));
        const size_t o_cells = G.slot(kind == 1 ? (size_t)n * 8 * sizeof(int) : 0);
        int cap = std::max(64, 16 * n);
        if (!G.alloc() || !S.dpairs.ensure((size_t)cap * 2 * sizeof(int))) return;
        cap = (int)std::min<size_t>(S.dpairs.cap / (2 * sizeof(int)), (size_t)INT32_MAX / 2);
        G.put(o_gv, gv.data(), gv.size());
        G.put(o_bits, bits.data(), bits.size());
        S.d_gviews = G.d<float>(o_gv);
        if (!G.send(o_gv, o_cells - o_gv)) return;
        SpaceCall C{};
        C.gviews = G.d<float>(o_gv);
        C.bits = G.d<unsigned long long>(o_bits);
        C.n = n;
        C.nsort = nsort;
        C.kind = kind;
        C.ax0 = space->ax0; C.ax1 = space->ax1; C.ax2 = space->ax2;
        C.minlevel = space->minlevel; C.maxlevel = space->maxlevel;
        C.ranks_in = sorts ? space->ranks[space->rcur].as<unsigned>() : nullptr;
        C.ranks_valid = sorts && space->ranks_valid ? 1 : 0;
        C.ranks_out = sorts ? space->ranks[1 - space->rcur].as<unsigned>() : nullptr;
        C.aabb = G.d<float>(o_ab);
        C.cells = kind == 1 ? G.d<int>(o_cells) : nullptr;
        C.info = G.d<int>(o_info);
        int total = 0;
        for (int attempt = 0; attempt < 2; attempt++) {
            C.pairs = S.dpairs.as<int>();
            C.cap = cap;
            launch_space_collide(C, g_stream);
            if (!odx_check(hipGetLastError(), "k_space_collide") || !G.fetch(0, out_end) ||
                !odx_check(hipStreamSynchronize(g_stream), "hipStreamSynchronize"))
                return;
            total = G.h<int>(o_info)[0];
            if (total <= cap) break;
            // more pairs than room: grow to the count and run again from the same ranks
            if (attempt == 1 || !S.dpairs.ensure((size_t)total * 2 * sizeof(int))) {
                odx_message(4, "dSpaceCollide: no room for %d candidate pairs", total);
                return;
            }
            cap = (int)std::min<size_t>(S.dpairs.cap / (2 * sizeof(int)), (size_t)INT32_MAX / 2);
        }
        if (sorts) {
            space->ranks_valid = G.h<int>(o_info)[1] != 0;
            space->rcur = 1 - space->rcur;
        }
        S.pairs.resize((size_t)total * 2);
        if (total > 0) {
            if (!S.hpairs.ensure(S.pairs.size() * sizeof(int)) ||
                !odx_check(hipMemcpyAsync(S.hpairs.p, S.dpairs.p, S.pairs.size() * sizeof(int), hipMemcpyDeviceToHost,
                                          g_stream), "hipMemcpyAsync D2H") ||
                !odx_check(hipStreamSynchronize(g_stream), "hipStreamSynchronize"))
                return;
            std::memcpy(S.pairs.data(), S.hpairs.p, S.pairs.size() * sizeof(int));
        }
        const float *ab = G.h<float>(o_ab);
        auto filters = [&](int a, int b) {
            const dxGeom *x = S.geoms[a], *y = S.geoms[b];
            if (x->body && x->body == y->body) return false;
            return ((x->cat & y->col) | (y->cat & x->col)) != 0ul;
        };
        auto overlap = [&](int a, int b) {
            const float *x = ab + (size_t)a * 6, *y = ab + (size_t)b * 6;
            return !(x[0] > y[1] || x[1] < y[0] || x[2] > y[3] || x[3] < y[2] || x[4] > y[5] || x[5] < y[4]);
        };
        auto add = [&](int a, int b) { S.pairs.push_back(a); S.pairs.push_back(b); };
        if (kind == 0) {
            // infinite x infinite, then infinite x finite, untested (:517-532)
            for (int m = nsort; m < n; m++) {
                for (int k = m + 1; k < n; k++)
                    if (filters(m, k)) add(m, k);
                for (int k = 0; k < nsort; k++)
                    if (filters(m, k)) add(m, k);
            }
        } else if (kind == 1) {
            // the big boxes (infinite, or above the top level): every hashed
            // geom against each of them, then the big boxes among themselves
            // (collision_space.cpp:595-609)
            std::vector<int> hashed, big;
            for (int i = 0; i < n; i++) {
                const float *x = ab + (size_t)i * 6;
                bool inf = false;
                for (int k = 0; k < 6; k += 2) inf = inf || x[k] <= -dInfinity || x[k + 1] >= dInfinity;
                int level = 0;
                if (!inf) {
                    float q = x[1] - x[0], q2 = x[3] - x[2];
                    if (q2 > q) q = q2;
                    q2 = x[5] - x[4];
                    if (q2 > q) q = q2;
                    std::frexp(q, &level);
                    if (level < space->minlevel) level = space->minlevel;
                }
                (inf || level > space->maxlevel ? big : hashed).push_back(i);
            }
            for (int a : hashed)
                for (int b : big)
                    if (filters(a, b) && overlap(a, b)) add(a, b);
            for (size_t a = 0; a < big.size(); a++)
                for (size_t b = a + 1; b < big.size(); b++)
                    if (filters(big[a], big[b]) && overlap(big[a], big[b])) add(big[a], big[b]);
        }
    }
    S.active = true;
    const int np = (int)S.pairs.size() / 2;
    for (int k = 0; k < np; k++) {
        S.current = k;
        callback(data, S.geoms[S.pairs[2 * k]], S.geoms[S.pairs[2 * k + 1]]);
    }
    S.active = false;
    S.current = -1;
}

// dCollide collision_kernel.cpp:292-329.  Inside dSpaceCollide the contacts of
// every candidate pair are computed in one launch at the first call.
int dCollide(dGeomID o1, dGeomID o2, int flags, dContactGeom *contact, int skip) {
    if (!o1 || !o2 || !contact) return 0;
    int maxc = flags & 0xffff;
    if (maxc == 0) return 0;
    if (o1 == o2) return 0;
    if (o1->body && o1->body == o2->body) return 0;
    if (skip < (int)sizeof(dContactGeom)) skip = (int)sizeof(dContactGeom);
    if (maxc > ODX_MAXC) maxc = ODX_MAXC;
    if (!ensure_stream()) return 0;
    Session &S = sess();
    const float *res = nullptr;
    int n = 0;
    std::vector<float> single;
    if (S.active && S.current >= 0 && S.geoms[S.pairs[2 * S.current]] == o1 &&
        S.geoms[S.pairs[2 * S.current + 1]] == o2) {
        if (S.maxc != maxc && !gpu_collide_pairs(S, maxc)) return 0;
        n = S.counts[S.current];
        res = &S.contacts[(size_t)S.current * ODX_MAXC * 7];
    } else {
        // a lone pair: one launch on its own buffers
        Session &one = lone_session();
        one.geoms = {o1, o2};
        std::vector<float> gv(2 * GVIEW_STRIDE);
        geom_view(o1, gv.data());
        geom_view(o2, gv.data() + GVIEW_STRIDE);
        one.pairs = {0, 1};
        Staging &G = one.space;
        G.reset();
        const size_t o_gv = G.slot(gv.size() * sizeof(float));
        if (!G.alloc()) return 0;
        G.put(o_gv, gv.data(), gv.size());
        one.d_gviews = G.d<float>(o_gv);
        if (!G.send(o_gv, gv.size() * sizeof(float)) || !gpu_collide_pairs(one, maxc)) return 0;
        n = one.counts[0];
        single = one.contacts;
        res = single.data();
    }
    for (int i = 0; i < n; i++) {
        dContactGeom *c = (dContactGeom *)((char *)contact + (size_t)i * skip);
        const float *r = res + (size_t)i * 7;
        c->pos[0] = r[0]; c->pos[1] = r[1]; c->pos[2] = r[2]; c->pos[3] = 0;
        c->normal[0] = r[3]; c->normal[1] = r[4]; c->normal[2] = r[5]; c->normal[3] = 0;
        c->depth = r[6];
        c->g1 = o1;
        c->g2 = o2;
        c->side1 = -1;
        c->side2 = -1;
    }
    return n;
}

// ------------------------------------------------------------------ step
// dWorldQuickStep objects.h:427 / ode.cpp:1847: islands by the reference's
// DFS, rows and sweeps on the GPU.  Returns 1, or 0 (state unchanged) when
// the device work fails.
// dInternalHandleAutoDisabling (util.cpp:427-565), run before the islands
// are built: each enabled auto-disabling body with joints samples its
// velocities into a ring of average_samples entries; once the ring is full,
// the squared speeds of the averages are tested against the thresholds, idle
// steps count the step and time budgets down (anything else resets them), and
// a body whose budgets are both spent is disabled with its velocities zeroed.
static void handle_auto_disabling(dxWorld *w, float h) {
    for (dxBody *b = w->firstbody; b; b = b->next) {
        if (!b->firstjoint) continue;  // never freeze a body in mid-air
        if (!b->auto_disable || !b->enabled) continue;
        const unsigned n = b->adis.average_samples;
        if (n == 0) continue;
        if (b->avg_counter >= n) { b->avg_ready = 0; b->avg_counter = 0; }
        for (int k = 0; k < 3; k++) {
            b->avg_lvel[3 * b->avg_counter + k] = b->lvel[k];
            b->avg_avel[3 * b->avg_counter + k] = b->avel[k];
        }
        if (++b->avg_counter >= n) { b->avg_counter = 0; b->avg_ready = 1; }
        bool idle = false;
        if (b->avg_ready) {
            idle = true;
            float al[3], aa[3];
            for (int k = 0; k < 3; k++) { al[k] = b->avg_lvel[k]; aa[k] = b->avg_avel[k]; }
            if (n > 1) {
                for (unsigned i = 1; i < n; i++)
                    for (int k = 0; k < 3; k++) { al[k] += b->avg_lvel[3 * i + k]; aa[k] += b->avg_avel[3 * i + k]; }
                const float r1 = 1.f / (float)n;
                for (int k = 0; k < 3; k++) { al[k] *= r1; aa[k] *= r1; }
            }
            if (al[0] * al[0] + al[1] * al[1] + al[2] * al[2] > b->adis.lin_thr2) idle = false;
            else if (aa[0] * aa[0] + aa[1] * aa[1] + aa[2] * aa[2] > b->adis.ang_thr2) idle = false;
        }
        if (idle) {
            b->adis_stepsleft--;
            b->adis_timeleft -= h;
        } else {
            b->adis_stepsleft = b->adis.idle_steps;
            b->adis_timeleft = b->adis.idle_time;
        }
        if (b->adis_stepsleft <= 0 && b->adis_timeleft <= 0) {
            b->enabled = false;
            for (int k = 0; k < 3; k++) { b->lvel[k] = 0; b->avel[k] = 0; }
        }
    }
}

int dWorldQuickStep(dWorldID w, dReal stepsize) {
    if (!w || !(stepsize > 0) || !ensure_stream()) return 0;
    handle_auto_disabling(w, stepsize);
    // ---- islands (util.cpp:724-855): body list order, LIFO joint lists
    std::vector<dxBody *> bodies;
    std::vector<dxJoint *> joints;
    for (dxBody *b = w->firstbody; b; b = b->next) b->tag = 0;
    std::vector<dxBody *> stack;
    std::vector<char> jtag;
    auto tagged = [&](dxJoint *j) { return j->tag != 0; };
    std::vector<dxJoint *> seen;
    std::vector<int> island;  // island of each listed body
    int nisl = 0;
    for (dxBody *bb = w->firstbody; bb; bb = bb->next) {
        if (bb->tag || !bb->enabled) continue;
        bb->tag = 1;
        bodies.push_back(bb);
        island.push_back(nisl);
        dxBody *b = bb;
        while (true) {
            for (dxJointNode *nd = b->firstjoint; nd; nd = nd->next) {
                dxJoint *j = nd->joint;
                if (tagged(j)) continue;
                if (joint_enabled(j)) {
                    j->tag = 1;
                    seen.push_back(j);
                    joints.push_back(j);
                    dxBody *o = nd->body;
                    if (o && o->tag <= 0) {
                        o->tag = 1;
                        o->enabled = true;
                        stack.push_back(o);
                    }
                } else {
                    j->tag = -1;
                    seen.push_back(j);
                }
            }
            if (stack.empty()) break;
            b = stack.back();
            stack.pop_back();
            bodies.push_back(b);
            island.push_back(nisl);
        }
        nisl++;
    }
    for (dxJoint *j : seen) j->tag = 0;
    const int nb = (int)bodies.size();
    if (nb > SL_MAXB) {
        odx_message(4, "dWorldQuickStep: %d enabled bodies exceed this build's single-world limit (%d)", nb, SL_MAXB);
        return 0;
    }
    for (int i = 0; i < nb; i++) bodies[i]->tag = i;

    // ---- records
    std::vector<float4> pos(nb), quat(nb), lvel(nb), avel(nb), facc(nb), tacc(nb), mass(nb), damp;
    std::vector<float> Ib((size_t)nb * 12), invIb((size_t)nb * 12);
    for (int i = 0; i < nb; i++) {
        dxBody *b = bodies[i];
        pos[i] = make_float4(b->pos[0], b->pos[1], b->pos[2], 0.f);
        quat[i] = make_float4(b->q[0], b->q[1], b->q[2], b->q[3]);
        lvel[i] = make_float4(b->lvel[0], b->lvel[1], b->lvel[2], 0.f);
        avel[i] = make_float4(b->avel[0], b->avel[1], b->avel[2], 0.f);
        facc[i] = make_float4(b->facc[0], b->facc[1], b->facc[2], 0.f);
        tacc[i] = make_float4(b->tacc[0], b->tacc[1], b->tacc[2], 0.f);
        const float flags = (float)((b->gravity ? 0 : 1) | (b->gyro ? 2 : 0));
        mass[i] = make_float4(b->mass.mass, 1.f / b->mass.mass, flags, 0.f);
        const int dbits = (b->lin_damping ? DAMP_LINEAR : 0) | (b->ang_damping ? DAMP_ANGULAR : 0) |
                          (b->max_ang ? DAMP_MAX_ANGULAR_SPEED : 0);
        if (dbits) {
            if (damp.empty()) damp.assign((size_t)nb * 2, make_float4(0.f, 0.f, 0.f, 0.f));
            damp[2 * i] = make_float4(b->dampingp.lin_scale, b->dampingp.ang_scale, b->dampingp.lin_thr2,
                                      b->dampingp.ang_thr2);
            damp[2 * i + 1] = make_float4(b->max_angular_speed, (float)dbits, 0.f, 0.f);
        }
        float *I = &Ib[(size_t)i * 12];
        for (int k = 0; k < 12; k++) I[k] = b->mass.I[k];
        I[3] = I[7] = I[11] = 0.f;
        odx_invert_pd3(I, &invIb[(size_t)i * 12]);
    }
    const int nj = (int)joints.size();
    std::vector<float> jrec((size_t)nj * JOINT_STRIDE, 0.f), jext;
    std::vector<Surface> surf;
    std::vector<int> joff(nj), jm(nj);
    std::vector<int> rb1, rb2, rfl;
    int m = 0;
    for (int k = 0; k < nj; k++) {
        dxJoint *j = joints[k];
        float *r = &jrec[(size_t)k * JOINT_STRIDE];
        r[1] = (float)j->node[0].body->tag;
        r[2] = j->node[1].body ? (float)j->node[1].body->tag : -1.f;
        int mj;
        if (j->type == dJointTypeContact) {
            const dContact &c = j->contact;
            r[0] = 0;
            const float sg = j->reverse ? -1.f : 1.f;
            for (int i = 0; i < 3; i++) { r[3 + i] = c.geom.pos[i]; r[6 + i] = sg * c.geom.normal[i]; r[10 + i] = c.fdir1[i]; }
            r[9] = c.geom.depth;
            r[13] = (float)surf.size();
            Surface s{};
            s.mode = c.surface.mode;
            s.mu = c.surface.mu < 0 ? 0.f : c.surface.mu;
            s.mu2 = c.surface.mu2 < 0 ? 0.f : c.surface.mu2;
            s.bounce = c.surface.bounce; s.bounce_vel = c.surface.bounce_vel;
            s.soft_erp = c.surface.soft_erp; s.soft_cfm = c.surface.soft_cfm;
            s.motion1 = c.surface.motion1; s.motion2 = c.surface.motion2; s.motionN = c.surface.motionN;
            s.slip1 = c.surface.slip1; s.slip2 = c.surface.slip2;
            surf.push_back(s);
            // getInfo1 contact.cpp:48-123
            mj = (s.mode & dContactMu2) ? 1 + (s.mu > 0) + (s.mu2 > 0) : 1 + (s.mu > 0 ? 2 : 0);
            // findex rows (friction with Approx1), in contact_info2's row order
            rfl.push_back(0);
            int row = 1;
            if (mj > 1) {
                if (s.mu > 0) { rfl.push_back((s.mode & dContactApprox1_1) ? 1 : 0); row++; }
                if (row < mj) rfl.push_back((s.mode & dContactApprox1_2) ? 1 : 0);
            }
        } else if (j->type == dJointTypeBall) {
            r[0] = 1;
            for (int i = 0; i < 3; i++) { r[3 + i] = j->anchor1[i]; r[6 + i] = j->anchor2[i]; }
            r[9] = j->erp;
            r[10] = j->cfm;
            mj = 3;
            rfl.push_back(0); rfl.push_back(0); rfl.push_back(0);
        } else {
            // hinge / slider / fixed: getInfo1 here, getInfo2 in the kernel
            r[0] = (float)(j->type == dJointTypeHinge ? JT_HINGE : j->type == dJointTypeSlider ? JT_SLIDER : JT_FIXED);
            for (int i = 0; i < 3; i++) { r[3 + i] = j->anchor1[i]; r[6 + i] = j->anchor2[i]; }
            r[9] = w->erp;
            r[10] = w->cfm;  // the ball rows of a hinge keep the world's cfm
            if (jext.empty()) jext.assign((size_t)nj * JEXT_STRIDE, 0.f);
            float *x = &jext[(size_t)k * JEXT_STRIDE];
            for (int i = 0; i < 3; i++) { x[JX_AXIS1 + i] = j->axis1[i]; x[JX_AXIS2 + i] = j->axis2[i]; x[JX_OFFSET + i] = j->offset[i]; }
            for (int i = 0; i < 4; i++) x[JX_QREL + i] = j->qrel[i];
            mj = j->type == dJointTypeFixed ? 6 : hinge_slider_m(j);
            x[JX_LIMOT_ROW] = (j->type != dJointTypeFixed && mj == 6) ? 1.f : 0.f;
            x[JX_LIMIT] = (float)j->limit;
            x[JX_LIMIT_ERR] = j->limit_err;
            x[JX_LOSTOP] = j->lostop; x[JX_HISTOP] = j->histop; x[JX_VEL] = j->vel; x[JX_FMAX] = j->fmax;
            x[JX_FUDGE] = j->fudge; x[JX_NORMAL_CFM] = j->normal_cfm; x[JX_STOP_ERP] = j->stop_erp;
            x[JX_STOP_CFM] = j->stop_cfm; x[JX_BOUNCE] = j->bounce; x[JX_REVERSE] = j->reverse ? 1.f : 0.f;
            x[JX_ERP] = j->erp; x[JX_CFM] = j->cfm;
            // a motor powered at its limit pushes with (a fraction of) fmax
            // through the body accumulators (addLimot joint.cpp:640-705)
            if (j->type != dJointTypeFixed && j->fmax > 0 && j->limit && j->lostop != j->histop) {
                float fm = j->fmax;
                if (j->vel > 0 || (j->vel == 0 && j->limit == 2)) fm = -fm;
                if ((j->limit == 1 && j->vel > 0) || (j->limit == 2 && j->vel < 0)) fm *= j->fudge;
                dxBody *b0 = j->node[0].body, *b1 = j->node[1].body;
                float ax1[3];
                rmul(ax1, b0->R, j->axis1);
                const bool rot = j->type == dJointTypeHinge;
                if (!rot && !b1 && j->reverse) for (int i = 0; i < 3; i++) ax1[i] = -ax1[i];
                const float f[3] = {fm * ax1[0], fm * ax1[1], fm * ax1[2]};
                float4 &t0 = tacc[b0->tag], &f0 = facc[b0->tag];
                if (rot) {
                    if (b1) { float4 &t1 = tacc[b1->tag]; t1.x += f[0]; t1.y += f[1]; t1.z += f[2]; }
                    t0.x -= f[0]; t0.y -= f[1]; t0.z -= f[2];
                } else {
                    if (b1) {
                        float c[3] = {0.5f * (b1->pos[0] - b0->pos[0]), 0.5f * (b1->pos[1] - b0->pos[1]),
                                      0.5f * (b1->pos[2] - b0->pos[2])}, ltd[3];
                        ltd[0] = c[1] * ax1[2] - c[2] * ax1[1];
                        ltd[1] = c[2] * ax1[0] - c[0] * ax1[2];
                        ltd[2] = c[0] * ax1[1] - c[1] * ax1[0];
                        float4 &t1 = tacc[b1->tag], &f1 = facc[b1->tag];
                        t0.x += -fm * ltd[0]; t0.y += -fm * ltd[1]; t0.z += -fm * ltd[2];
                        t1.x += -fm * ltd[0]; t1.y += -fm * ltd[1]; t1.z += -fm * ltd[2];
                        f1.x += f[0]; f1.y += f[1]; f1.z += f[2];
                    }
                    f0.x -= f[0]; f0.y -= f[1]; f0.z -= f[2];
                }
            }
            for (int q = 0; q < mj; q++) rfl.push_back(0);
        }
        joff[k] = m;
        jm[k] = mj;
        for (int q = 0; q < mj; q++) { rb1.push_back((int)r[1]); rb2.push_back((int)r[2]); }
        m += mj;
    }
    if (surf.empty()) surf.push_back(Surface{});
    // ---- schedule (see odx_step.hip): dependency levels over the ReorderPrep
    // order, or greedy colours of whole joints
    std::vector<int> sched(m), phase;
    int wave_groups = 0;
    // island groups on separate workgroups (k_single_groups): the group of each
    // island and the LDS one such workgroup takes
    bool group_blocks = false, group_rows_lds = false;
    std::vector<int> isl_group;
    size_t group_lds = 0;
    const size_t group_lds_max = std::min((size_t)single_lds_max(), (size_t)SL_GROUP_LDS);
    {
        std::vector<int> level;
        int L = 0;
        if (w->order == 0 || w->order == 2) {
            level.assign(m, 0);
            std::vector<int> last(nb, 0);
            for (int pass = 0; pass < 2; pass++)
                for (int r = 0; r < m; r++) {
                    if (rfl[r] != pass) continue;
                    int lv = last[rb1[r]];
                    if (rb2[r] >= 0) lv = std::max(lv, last[rb2[r]]);
                    lv++;
                    last[rb1[r]] = lv;
                    if (rb2[r] >= 0) last[rb2[r]] = lv;
                    level[r] = lv - 1;
                    L = std::max(L, lv);
                }
            // sweep mode 3 (automatic for the reference order with a fixed
            // sweep count and enough islands): islands never share a body, so
            // each wave of the step kernel can sweep the levels of its own
            // islands without meeting the others; the Gauss-Seidel order
            // inside every island is unchanged
            int isl_rows = 0;
            {
                std::vector<char> has(std::max(nisl, 1), 0);
                for (int r = 0; r < m; r++) has[island[rb1[r]]] = 1;
                for (int i = 0; i < nisl; i++) isl_rows += has[i];
            }
            const float xc = (float)w->iterations * w->dyn_extra_factor;
            const bool dyn_sched = (xc < (float)ODX_MAX_EXTRA_SWEEPS ? (int)xc : ODX_MAX_EXTRA_SWEEPS) != 0 ||
                                   w->dyn_premature != 0.f;
            wave_groups = 0;
            if (w->order == 0 && !dyn_sched &&
                (w->sweep_mode == 3 || (w->sweep_mode == 0 && isl_rows >= 4 * SL_STAGED_WAVES))) {
                wave_groups = SL_STAGED_WAVES;
                group_blocks = true;
            }
            // items per level of each island
            std::vector<std::vector<int>> ilc(std::max(nisl, 1));
            if (wave_groups > 0)
                for (int r = 0; r < m; r++) {
                    std::vector<int> &c = ilc[island[rb1[r]]];
                    if ((int)c.size() <= level[r]) c.resize(level[r] + 1, 0);
                    c[level[r]]++;
                }
            // one workgroup per group: runs of consecutive islands (their
            // bodies, joints and rows are consecutive), cut before an island
            // that would give a level of the run more than 64 items (a second
            // pass of the wave) or take the run's LDS past the budget.  The
            // rows go into LDS too when every run fits with them (the chip's
            // per-workgroup LDS), else the runs are cut for the 64 KB budget
            // without rows.  More than SL_GROUP_BLOCKS runs fall back to
            // groups sharing one workgroup.
            std::vector<int> run_start;  // first island of each run, then nisl
            if (group_blocks) {
                std::vector<long> ibod(std::max(nisl, 1), 0), irow(std::max(nisl, 1), 0);
                for (int i = 0; i < nb; i++) ibod[island[i]]++;
                for (int r = 0; r < m; r++) irow[island[rb1[r]]]++;
                // LDS words: 6 per body, per row 2 (+ 32 with its record), 1 per level, 1
                auto cut_runs = [&](size_t row_words, size_t budget) {
                    std::vector<int> width;  // items per level of the current run
                    size_t words = 1;
                    bool ok = true;
                    run_start.assign(1, 0);
                    for (int i = 0; i < nisl; i++) {
                        const size_t wi = (size_t)ibod[i] * 6 + row_words * (size_t)irow[i];
                        bool fits = (words + wi + std::max(ilc[i].size(), width.size()) - width.size()) *
                                        sizeof(float) <= budget;
                        for (size_t l = 0; fits && l < ilc[i].size(); l++)
                            fits = (l < width.size() ? width[l] : 0) + ilc[i][l] <= 64;
                        if (!fits && i > run_start.back()) {
                            run_start.push_back(i);
                            width.clear();
                            words = 1;
                        }
                        if (width.size() < ilc[i].size()) {
                            words += ilc[i].size() - width.size();
                            width.resize(ilc[i].size(), 0);
                        }
                        for (size_t l = 0; l < ilc[i].size(); l++) width[l] += ilc[i][l];
                        words += wi;
                        if (words * sizeof(float) > budget) ok = false;  // one island alone is over
                    }
                    run_start.push_back(std::max(nisl, 1));
                    return ok && (int)run_start.size() - 1 <= SL_GROUP_BLOCKS;
                };
                const size_t rows_budget = std::min((size_t)single_lds_max(), (size_t)SL_GROUP_LDS_ROWS);
                group_rows_lds = cut_runs(2 + ROW_STRIDE, rows_budget);
                if (!group_rows_lds) cut_runs(2, group_lds_max);
                wave_groups = (int)run_start.size() - 1;
                if (wave_groups > SL_GROUP_BLOCKS) {
                    group_blocks = false;
                    wave_groups = SL_STAGED_WAVES;
                }
            }
            for (int attempt_g = 0; wave_groups > 0 && attempt_g < 2; attempt_g++) {
                const int G = wave_groups;
                // islands (deepest first) to the group whose passes (sum over
                // levels of ceil(items / 64)) grow least, or runs of islands
                std::vector<int> order_isl;
                for (int i = 0; i < nisl; i++)
                    if (!ilc[i].empty()) order_isl.push_back(i);
                std::stable_sort(order_isl.begin(), order_isl.end(),
                                 [&](int a, int b) { return ilc[a].size() > ilc[b].size(); });
                std::vector<std::vector<int>> gc(G);
                std::vector<long> gcost(G, 0);
                std::vector<int> group(std::max(nisl, 1), 0);
                auto passes = [](int n) { return (n + 63) / 64; };
                if (group_blocks) {
                    for (int g = 0; g < G; g++)
                        for (int i = run_start[g]; i < run_start[g + 1] && i < nisl; i++) {
                            group[i] = g;
                            std::vector<int> &c = gc[g];
                            if (c.size() < ilc[i].size()) c.resize(ilc[i].size(), 0);
                            for (size_t l = 0; l < ilc[i].size(); l++) c[l] += ilc[i][l];
                        }
                } else {
                    for (int i : order_isl) {
                        int best = 0;
                        long bestc = -1;
                        for (int g = 0; g < G; g++) {
                            long d = 0;
                            for (size_t l = 0; l < ilc[i].size(); l++) {
                                const int have = l < gc[g].size() ? gc[g][l] : 0;
                                d += passes(have + ilc[i][l]) - passes(have);
                            }
                            if (bestc < 0 || gcost[g] + d < bestc) { bestc = gcost[g] + d; best = g; }
                        }
                        std::vector<int> &c = gc[best];
                        if (c.size() < ilc[i].size()) c.resize(ilc[i].size(), 0);
                        for (size_t l = 0; l < ilc[i].size(); l++) c[l] += ilc[i][l];
                        gcost[best] = bestc;
                        group[i] = best;
                    }
                }
                if (group_blocks) {
                    // LDS of a run: fc of its bodies, lambda of its rows (one
                    // item each), its level starts; over the budget, the
                    // groups share one workgroup instead
                    std::vector<long> ibod(std::max(nisl, 1), 0), irow(std::max(nisl, 1), 0);
                    for (int i = 0; i < nb; i++) ibod[island[i]]++;
                    for (int r = 0; r < m; r++) irow[island[rb1[r]]]++;
                    const size_t row_words = group_rows_lds ? 2 + ROW_STRIDE : 2;
                    size_t most = 0;
                    for (int g = 0; g < G; g++) {
                        size_t fl = gc[g].size() + 1;
                        for (int i = run_start[g]; i < run_start[g + 1] && i < nisl; i++)
                            fl += (size_t)ibod[i] * 6 + row_words * (size_t)irow[i];
                        most = std::max(most, fl);
                    }
                    group_lds = most * sizeof(float);
                    const size_t cap = group_rows_lds
                                           ? std::min((size_t)single_lds_max(), (size_t)SL_GROUP_LDS_ROWS)
                                           : group_lds_max;
                    if (group_lds > cap) {
                        group_blocks = false;
                        wave_groups = SL_STAGED_WAVES;
                        continue;
                    }
                    isl_group = group;
                }
                // phase table: G + 1 offsets of the groups' level lists (from
                // entry G + 1), then per group its level starts and end; items
                // of a level in row order
                std::vector<int> first(G + 1, 0);
                for (int g = 0; g < G; g++) first[g + 1] = first[g] + (int)gc[g].size();
                std::vector<int> start(first[G] + 1, 0);
                for (int g = 0; g < G; g++)
                    for (size_t l = 0; l < gc[g].size(); l++) start[first[g] + l + 1] = gc[g][l];
                for (int k = 1; k <= first[G]; k++) start[k] += start[k - 1];
                phase.assign(G + 1, 0);
                for (int g = 0; g < G; g++) {
                    phase[g] = (int)(phase.size() - (G + 1));
                    for (size_t l = 0; l <= gc[g].size(); l++) phase.push_back(start[first[g] + l]);
                }
                phase[G] = (int)(phase.size() - (G + 1));
                std::vector<int> pos(start.begin(), start.end() - 1);
                for (int r = 0; r < m; r++) sched[pos[first[group[island[rb1[r]]]] + level[r]]++] = (r << 3) | 1;
                break;
            }
            if (wave_groups == 0) {
                std::vector<int> cnt(L + 1, 0);
                for (int r = 0; r < m; r++) cnt[level[r] + 1]++;
                for (int p = 1; p <= L; p++) cnt[p] += cnt[p - 1];
                phase = cnt;
                for (int r = 0; r < m; r++) sched[cnt[level[r]]++] = (r << 3) | 1;
            }
        } else {
            level.assign(nj, 0);
            std::vector<std::vector<char>> used(nb);
            for (int k = 0; k < nj; k++) {
                const int a = rb1[joff[k]], b = rb2[joff[k]];
                int c = 0;
                while (true) {
                    bool ua = c < (int)used[a].size() && used[a][c];
                    bool ub = b >= 0 && c < (int)used[b].size() && used[b][c];
                    if (!ua && !ub) break;
                    c++;
                }
                if ((int)used[a].size() <= c) used[a].resize(c + 1, 0);
                used[a][c] = 1;
                if (b >= 0) { if ((int)used[b].size() <= c) used[b].resize(c + 1, 0); used[b][c] = 1; }
                level[k] = c;
                L = std::max(L, c + 1);
            }
            // two passes like ReorderPrep (quickstep.cpp:2329-2355): each joint's
            // leading non-findex rows in colour phases first, then its findex
            // (friction-cone) rows in a second set of colour phases
            std::vector<int> na(nj), nf(nj);
            for (int k = 0; k < nj; k++) split_joint_rows(&rfl[joff[k]], jm[k], na[k], nf[k]);
            std::vector<int> cnt(2 * L + 1, 0);
            for (int k = 0; k < nj; k++) {
                cnt[level[k] + 1]++;
                if (nf[k]) cnt[L + level[k] + 1]++;
            }
            std::vector<int> start(2 * L, 0);
            phase.clear();
            int run = 0;
            for (int p = 0; p < 2 * L; p++) {
                start[p] = run;
                if (cnt[p + 1]) phase.push_back(run);
                run += cnt[p + 1];
            }
            phase.push_back(run);
            sched.resize(run);
            for (int k = 0; k < nj; k++) {
                sched[start[level[k]]++] = (joff[k] << 3) | na[k];
                if (nf[k]) sched[start[L + level[k]]++] = ((joff[k] + na[k]) << 3) | nf[k];
            }
        }
    }
    int L = (int)phase.size() - 1;

    // ---- random reordering (order 2: REORDERING_METHOD__RANDOMLY, the
    // reference's default build, quickstep.cpp:54-85, 1080-1109, 2589-2611).
    // Each island's order starts as its ReorderPrep order and is shuffled
    // (index 1..m-1 swapped with dRandInt(index + 1)) before sweeps 8, 16, ...;
    // islands draw from dRand one after the other in island order.  Sweeps
    // 8s..8s+7 use schedule segment s (dependency levels of the orders after s
    // shuffles).  An island's draws start where the earlier islands' shuffles
    // ended, so the schedules assume each island's shuffle count; with the
    // dynamic iteration count the counts are known only after the solve, and
    // the step is run again with the counts it reported until they agree.
    const unsigned seed0 = (unsigned)g_seed;
    const int n_iter_max = [&] {
        const float xc = (float)w->iterations * w->dyn_extra_factor;
        const int xm = xc < (float)ODX_MAX_EXTRA_SWEEPS ? (int)xc : ODX_MAX_EXTRA_SWEEPS;
        const bool dyn = xm != 0 || w->dyn_premature != 0.f;
        return w->iterations + (dyn ? xm : 0);
    }();
    const int nseg = w->order == 2 ? std::max(1, (n_iter_max - 1) / 8 + 1) : 0;
    std::vector<std::vector<int>> isl_rows;  // island -> its rows in ReorderPrep order
    if (w->order == 2) {
        isl_rows.assign(nisl, {});
        for (int pass = 0; pass < 2; pass++)
            for (int r = 0; r < m; r++)
                if (rfl[r] == pass) isl_rows[island[rb1[r]]].push_back(r);
    }
    std::vector<int> assumed(nisl, 0), seg_table;
    for (int i = 0; i < (int)isl_rows.size(); i++)
        if (!isl_rows[i].empty()) assumed[i] = std::min(std::max(0, (w->iterations - 1) / 8), nseg - 1);
    std::vector<unsigned> seed_after;  // per island, state after each of its shuffles (nseg entries)
    const auto pos0 = pos, quat0 = quat, lvel0 = lvel, avel0 = avel;
    auto build_segments = [&]() {
        std::vector<std::vector<std::vector<int>>> orders(nseg, std::vector<std::vector<int>>(nisl));
        seed_after.assign((size_t)nisl * nseg, 0u);
        uint32_t seed = seed0;
        for (int i = 0; i < nisl; i++) {
            std::vector<int> o = isl_rows[i];
            orders[0][i] = o;
            uint32_t si = seed;
            seed_after[(size_t)i * nseg] = si;
            for (int sg = 1; sg < nseg; sg++) {
                for (size_t index = 1; index < o.size(); ++index) std::swap(o[index], o[rand_int(si, (int)index + 1)]);
                orders[sg][i] = o;
                seed_after[(size_t)i * nseg + sg] = si;
            }
            seed = seed_after[(size_t)i * nseg + std::min(assumed[i], nseg - 1)];
        }
        sched.clear();
        phase.clear();
        seg_table.clear();
        std::vector<int> lvl(m), last(nb);
        for (int sg = 0; sg < nseg; sg++) {
            std::fill(last.begin(), last.end(), 0);
            int Ls = 0;
            for (int i = 0; i < nisl; i++)
                for (int r : orders[sg][i]) {
                    int lv = last[rb1[r]];
                    if (rb2[r] >= 0) lv = std::max(lv, last[rb2[r]]);
                    lv++;
                    last[rb1[r]] = lv;
                    if (rb2[r] >= 0) last[rb2[r]] = lv;
                    lvl[r] = lv - 1;
                    Ls = std::max(Ls, lv);
                }
            std::vector<int> cnt(Ls + 1, 0);
            for (int r = 0; r < m; r++) cnt[lvl[r] + 1]++;
            for (int q = 1; q <= Ls; q++) cnt[q] += cnt[q - 1];
            seg_table.push_back((int)sched.size());
            seg_table.push_back((int)phase.size());
            seg_table.push_back(Ls);
            const int base = (int)sched.size();
            for (int q = 0; q <= Ls; q++) phase.push_back(cnt[q]);  // relative to the segment's first item
            sched.resize(base + m);
            for (int i = 0; i < nisl; i++)
                for (int r : orders[sg][i]) sched[base + cnt[lvl[r]]++] = (r << 3) | 1;
        }
        L = seg_table[2];
    };
    std::vector<float> fbh;
    int it = 0;
    float res = -1.f;
    bool want_fb = false;
    for (dxJoint *j : joints) want_fb = want_fb || j->feedback;
    std::vector<int> isl_sweeps(std::max(nisl, 1));
    for (int attempt = 0;; attempt++) {
    if (w->order == 2) {
        build_segments();
        pos = pos0; quat = quat0; lvel = lvel0; avel = avel0;
    }

    // ---- every index the kernel will follow stays inside its array: items
    // address rows [r, r + n) < m, phase starts rise within the segment's items
    {
        auto bad = [&](const char *what) {
            odx_message(4, "dWorldQuickStep: malformed schedule (%s)", what);
            return 0;
        };
        std::vector<int> segs = w->order == 2 ? seg_table : std::vector<int>{0, 0, L};
        if (wave_groups > 0) {
            // island groups: a header of wave_groups + 1 offsets (from entry
            // wave_groups + 1) of each group's level list, read as segments
            const int G = wave_groups;
            if ((int)phase.size() < G + 2 || phase[0] != 0) return bad("group table");
            segs.clear();
            for (int g = 0; g < G; g++) {
                const int a = phase[g], b = phase[g + 1];
                if (a < 0 || b <= a || (size_t)(G + 1 + b) > phase.size()) return bad("group table");
                segs.push_back(0);
                segs.push_back(G + 1 + a);
                segs.push_back(b - a - 1);
            }
            if ((size_t)(G + 1 + phase[G]) != phase.size()) return bad("group table");
        }
        for (size_t q = 0; q + 2 < segs.size(); q += 3) {
            const int so = segs[q], po = segs[q + 1], Ls = segs[q + 2];
            if (so < 0 || po < 0 || Ls < 0 || (size_t)po + Ls >= phase.size()) return bad("segment table");
            for (int k = 0; k < Ls; k++)
                if (phase[po + k] > phase[po + k + 1]) return bad("phase order");
            if (phase[po] < 0 || (size_t)so + phase[po + Ls] > sched.size()) return bad("phase range");
            for (int k = phase[po]; k < phase[po + Ls]; k++) {
                const int item = sched[so + k], r = item >> 3, n = item & 7;
                if (r < 0 || n < 1 || r + n > m) return bad("item rows");
            }
        }
    }

    // ---- device: every input in one staged copy; the state, the sweep count,
    // the residual, joint feedback and island sweeps sit first in the block
    // and come back in one copy
    int bstart[2] = {0, nb};
    int cnts[4] = {nj, 0, m, L};  // njoints, ncontacts (unused), nrows, nphases
    const bool dyn_on = [&] {
        const float xc = (float)w->iterations * w->dyn_extra_factor;
        const int xm = xc < (float)ODX_MAX_EXTRA_SWEEPS ? (int)xc : ODX_MAX_EXTRA_SWEEPS;
        return xm != 0 || w->dyn_premature != 0.f;
    }();
    // island groups on separate workgroups: the first body and joint of each
    // group's run of islands (bodies and joints are listed island by island)
    std::vector<int> gbs, gjs;
    if (group_blocks) {
        const int NG = wave_groups;
        gbs.assign(NG + 1, 0);
        gjs.assign(NG + 1, 0);
        for (int i = 0; i < nb; i++) gbs[isl_group[island[i]] + 1]++;
        for (int k = 0; k < nj; k++) gjs[isl_group[island[(int)jrec[(size_t)k * JOINT_STRIDE + 1]]] + 1]++;
        for (int g = 0; g < NG; g++) { gbs[g + 1] += gbs[g]; gjs[g + 1] += gjs[g]; }
        // the runs must cover the lists in order (a body or joint out of its
        // run's range would be stepped by no workgroup)
        for (int i = 0; i < nb; i++)
            if (i < gbs[isl_group[island[i]]] || i >= gbs[isl_group[island[i]] + 1]) group_blocks = false;
        for (int k = 0; k < nj && group_blocks; k++) {
            const int g = isl_group[island[(int)jrec[(size_t)k * JOINT_STRIDE + 1]]];
            if (k < gjs[g] || k >= gjs[g + 1]) group_blocks = false;
        }
        if (!group_blocks) {
            odx_message(4, "dWorldQuickStep: island runs out of list order");
            return 0;
        }
    }
    Staging &G = w->stage;
    G.reset();
    const size_t f4 = (size_t)nb * sizeof(float4);
    const size_t o_pos = G.slot(f4), o_quat = G.slot(f4), o_lvel = G.slot(f4), o_avel = G.slot(f4);
    const size_t o_it = G.slot(2 * sizeof(int));
    const size_t o_fb = G.slot(want_fb ? (size_t)nj * 12 * sizeof(float) : 0);
    const size_t o_isl = G.slot(w->order == 2 || dyn_on ? (size_t)std::max(nisl, 1) * sizeof(int) : 0);
    const size_t out_end = G.used;
    const size_t o_facc = G.slot(f4), o_tacc = G.slot(f4), o_mass = G.slot(f4);
    const size_t o_Ib = G.slot(Ib.size() * sizeof(float)), o_invIb = G.slot(invIb.size() * sizeof(float));
    const size_t o_joints = G.slot(jrec.size() * sizeof(float)), o_surf = G.slot(surf.size() * sizeof(Surface));
    const size_t o_joff = G.slot(joff.size() * sizeof(int)), o_sched = G.slot(sched.size() * sizeof(int));
    const size_t o_phase = G.slot(phase.size() * sizeof(int)), o_bstart = G.slot(sizeof bstart);
    const size_t o_counts = G.slot(sizeof cnts), o_damp = G.slot(damp.size() * sizeof(float4));
    const size_t o_island = G.slot(dyn_on ? island.size() * sizeof(int) : 0);
    const size_t o_seg = G.slot(seg_table.size() * sizeof(int));
    const size_t o_jext = G.slot(jext.size() * sizeof(float));
    const size_t o_gbs = G.slot(gbs.size() * sizeof(int)), o_gjs = G.slot(gjs.size() * sizeof(int));
    bool ok = G.alloc() && w->rows.ensure((size_t)std::max(m, 1) * ROW_STRIDE * sizeof(float)) &&
              w->lambda.ensure((size_t)std::max(m, 1) * sizeof(float)) &&
              w->bscr.ensure((size_t)std::max(nb, 1) * BSCR_STRIDE * sizeof(float));
    if (!ok) return 0;
    G.put(o_pos, pos.data(), nb); G.put(o_quat, quat.data(), nb); G.put(o_lvel, lvel.data(), nb);
    G.put(o_avel, avel.data(), nb); G.put(o_facc, facc.data(), nb); G.put(o_tacc, tacc.data(), nb);
    G.put(o_mass, mass.data(), nb); G.put(o_Ib, Ib.data(), Ib.size()); G.put(o_invIb, invIb.data(), invIb.size());
    G.put(o_joints, jrec.data(), jrec.size()); G.put(o_surf, surf.data(), surf.size());
    G.put(o_joff, joff.data(), joff.size()); G.put(o_sched, sched.data(), sched.size());
    G.put(o_phase, phase.data(), phase.size()); G.put(o_bstart, bstart, 2); G.put(o_counts, cnts, 4);
    G.put(o_damp, damp.data(), damp.size());
    if (dyn_on) G.put(o_island, island.data(), island.size());
    G.put(o_seg, seg_table.data(), seg_table.size());
    G.put(o_jext, jext.data(), jext.size());
    G.put(o_gbs, gbs.data(), gbs.size()); G.put(o_gjs, gjs.data(), gjs.size());
    if (!G.send(0, G.used)) return 0;
    DevWorlds D{};
    D.nworlds = 1;
    D.body_start = G.d<int>(o_bstart);
    D.pos = G.d<float4>(o_pos); D.quat = G.d<float4>(o_quat); D.lvel = G.d<float4>(o_lvel); D.avel = G.d<float4>(o_avel);
    D.facc = G.d<float4>(o_facc); D.tacc = G.d<float4>(o_tacc); D.mass = G.d<float4>(o_mass);
    D.Ib = G.d<float>(o_Ib); D.invIb = G.d<float>(o_invIb);
    D.surf = G.d<Surface>(o_surf);
    D.joints = G.d<float>(o_joints);
    D.njoints = G.d<int>(o_counts);
    D.ncontacts = G.d<int>(o_counts) + 1;
    D.nrows = G.d<int>(o_counts) + 2;
    D.nphases = G.d<int>(o_counts) + 3;
    D.niters = G.d<int>(o_it);
    D.residual = G.d<float>(o_it) + 1;
    D.rows = w->rows.as<float>();
    D.lambda = w->lambda.as<float>();
    D.sched = G.d<int>(o_sched);
    D.phase_start = G.d<int>(o_phase);
    D.bscr = w->bscr.as<float>();
    D.joff = G.d<int>(o_joff);
    D.fb = want_fb ? G.d<float>(o_fb) : nullptr;
    D.damp = damp.empty() ? nullptr : G.d<float4>(o_damp);
    D.jext = jext.empty() ? nullptr : G.d<float>(o_jext);
    if (group_blocks) {
        D.grp_body_start = G.d<int>(o_gbs);
        D.grp_joint_start = G.d<int>(o_gjs);
    }
    StepParams P{};
    P.h = stepsize;
    P.gx = w->gravity[0]; P.gy = w->gravity[1]; P.gz = w->gravity[2];
    P.erp = w->erp; P.cfm = w->cfm; P.sor_w = w->w;
    P.max_iters = w->iterations;
    P.tol = 0.f;
    P.check_every = 1;
    P.max_vel = w->max_vel; P.min_depth = w->min_depth;
    P.maxc = 4;
    P.order = w->order;
    P.wave_sweep = w->sweep_mode == 3 ? 0 : w->sweep_mode;
    P.wave_groups = wave_groups;
    P.lambda_floor = 1e-3f;
    P.collide = 0;
    P.feedback = want_fb ? 1 : 0;
    // dynamic iteration count (objects.h:177-189): extra sweeps = iterations x
    // factor; on when that is non-zero or the premature delta is
    {
        const float xc = (float)w->iterations * w->dyn_extra_factor;
        const int max_extra = xc < (float)ODX_MAX_EXTRA_SWEEPS ? (int)xc : ODX_MAX_EXTRA_SWEEPS;
        P.dyn = (max_extra != 0 || w->dyn_premature != 0.f) ? 1 : 0;
        P.dyn_max_extra = max_extra;
        P.dyn_premature = w->dyn_premature;
        P.dyn_extra_delta = w->dyn_extra_delta;
        P.nislands = nisl;
        D.body_island = P.dyn ? G.d<int>(o_island) : nullptr;
    }
    D.island_sweeps = P.dyn ? G.d<int>(o_isl) : nullptr;
    if (w->order == 2) {
        D.seg_table = G.d<int>(o_seg);
        D.nseg = nseg;
    }
    if (g_ktime.on && !g_ktime.ev[0]) {
        if (!odx_check(hipEventCreate(&g_ktime.ev[0]), "hipEventCreate") ||
            !odx_check(hipEventCreate(&g_ktime.ev[1]), "hipEventCreate"))
            return 0;
    }
    if (g_ktime.on && !odx_check(hipEventRecord(g_ktime.ev[0], g_stream), "hipEventRecord")) return 0;
    if (!(group_blocks && launch_single_groups(D, P, group_lds, group_rows_lds, g_stream))) {
        if (group_blocks) {
            odx_message(4, "dWorldQuickStep: the island groups' workgroups could not be launched");
            return 0;
        }
        launch_single_step(D, P, nb, m, (int)sched.size(), L, g_stream);
    }
    if (!odx_check(hipGetLastError(), "k_single_step")) return 0;
    if (g_ktime.on && !odx_check(hipEventRecord(g_ktime.ev[1], g_stream), "hipEventRecord")) return 0;
    ok = G.fetch(0, out_end) && odx_check(hipStreamSynchronize(g_stream), "hipStreamSynchronize");
    if (!ok) return 0;
    std::memcpy(pos.data(), G.h<float4>(o_pos), f4);
    std::memcpy(quat.data(), G.h<float4>(o_quat), f4);
    std::memcpy(lvel.data(), G.h<float4>(o_lvel), f4);
    std::memcpy(avel.data(), G.h<float4>(o_avel), f4);
    it = G.h<int>(o_it)[0];
    res = G.h<float>(o_it)[1];
    fbh.assign(want_fb ? (size_t)nj * 12 : 0, 0.f);
    if (want_fb) std::memcpy(fbh.data(), G.h<float>(o_fb), fbh.size() * sizeof(float));
    if (P.dyn) std::memcpy(isl_sweeps.data(), G.h<int>(o_isl), (size_t)nisl * sizeof(int));
    if (w->order != 2) break;
    // shuffles each island ran: before sweeps 8, 16, ... that it reached
    bool agree = true;
    for (int i = 0; i < nisl; i++) {
        const int sw = P.dyn ? isl_sweeps[i] : w->iterations;
        const int ran = isl_rows[i].empty() ? 0 : std::min(std::max(0, (sw - 1) / 8), nseg - 1);
        if (ran != assumed[i]) agree = false;
        assumed[i] = ran;
    }
    if (agree) {
        g_seed = (int)(nisl > 0 ? seed_after[(size_t)(nisl - 1) * nseg + assumed[nisl - 1]] : seed0);
        break;
    }
    if (attempt > nisl + 1) {
        odx_message(4, "dWorldQuickStep: the random reordering's shuffle counts did not settle");
        return 0;
    }
    }  // attempts
    for (int i = 0; i < nb; i++) {
        dxBody *b = bodies[i];
        b->pos[0] = pos[i].x; b->pos[1] = pos[i].y; b->pos[2] = pos[i].z;
        b->q[0] = quat[i].x; b->q[1] = quat[i].y; b->q[2] = quat[i].z; b->q[3] = quat[i].w;
        rfromq(b->R, b->q);
        b->lvel[0] = lvel[i].x; b->lvel[1] = lvel[i].y; b->lvel[2] = lvel[i].z;
        b->avel[0] = avel[i].x; b->avel[1] = avel[i].y; b->avel[2] = avel[i].z;
        for (int k = 0; k < 3; k++) { b->facc[k] = 0; b->tacc[k] = 0; }
    }
    // dxStepBody notifies each body's geoms as it integrates it (util.cpp:661-668)
    for (int i = 0; i < nb; i++) body_moved(bodies[i]);
    if (want_fb) {
        for (int k = 0; k < nj; k++) {
            dJointFeedback *fb = joints[k]->feedback;
            if (!fb) continue;
            const float *f = &fbh[(size_t)k * 12];
            for (int i = 0; i < 3; i++) { fb->f1[i] = f[i]; fb->t1[i] = f[3 + i]; }
            if (joints[k]->node[1].body)
                for (int i = 0; i < 3; i++) { fb->f2[i] = f[6 + i]; fb->t2[i] = f[9 + i]; }
        }
    }
    // dynamic-iteration statistics per island with rows (quickstep.cpp:1829-1853):
    // stopped before the iteration count, inside the extra sweeps, or after all
    // of them (the last extra sweep is not followed by the exit test)
    const float xcount = (float)w->iterations * w->dyn_extra_factor;
    const int X = xcount < (float)ODX_MAX_EXTRA_SWEEPS ? (int)xcount : ODX_MAX_EXTRA_SWEEPS;
    if (X != 0 || w->dyn_premature != 0.f) {
        std::vector<int> isl_m((size_t)std::max(nisl, 1), 0);
        for (int k = 0; k < nj; k++) isl_m[(size_t)island[(size_t)joints[k]->node[0].body->tag]] += jm[k];
        const int N = w->iterations;
        for (int i = 0; i < nisl; i++) {
            if (isl_m[(size_t)i] == 0) continue;
            const int sw = isl_sweeps[(size_t)i];
            if (sw < N) w->stats->premature_exits++;
            else if (X > 0 && sw == N + X) w->stats->full_extra_execs++;
            else if (sw > N) w->stats->prolonged_execs++;
        }
    }
    w->last_sweeps = it;
    w->last_res = res;
    w->last_rows = m;
    if (g_ktime.on) {
        float ms = 0.f;
        if (!odx_check(hipEventElapsedTime(&ms, g_ktime.ev[0], g_ktime.ev[1]), "hipEventElapsedTime")) return 0;
        g_ktime.steps += 1;
        g_ktime.ms += ms;
        g_ktime.row_sweeps += (double)m * it;
    }
    return 1;
}

// Kernel time of the device step of dWorldQuickStep (no ODE counterpart).
void odxApiSetKernelTiming(int on) {
    g_ktime.on = on != 0;
    g_ktime.steps = 0;
    g_ktime.ms = 0.0;
    g_ktime.row_sweeps = 0.0;
}

int odxApiGetKernelTiming(double *stats) {
    if (!stats) return 0;
    stats[0] = (double)g_ktime.steps;
    stats[1] = g_ktime.ms;
    stats[2] = g_ktime.row_sweeps;
    return 1;
}

// ------------------------------------------------------------------ misc
unsigned long dRand(void) {
    uint32_t sd = (uint32_t)g_seed;
    rand_next(sd);
    g_seed = (int)sd;
    return (unsigned long)sd;
}
unsigned long dRandGetSeed(void) { return (unsigned long)(uint32_t)g_seed; }
void dRandSetSeed(unsigned long s) { g_seed = (int)(uint32_t)s; }
int dRandInt(int n) {
    uint32_t sd = (uint32_t)g_seed;
    const int r = rand_int(sd, n);
    g_seed = (int)sd;
    return r;
}

}  // extern "C"
