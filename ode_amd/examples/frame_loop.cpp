// frame_loop.cpp -- an ODE frame loop over the drop-in API (include/ode.h),
// written as an existing ODE program would: dSpaceCollide with a near
// callback that calls dCollide and dJointCreateContact, dWorldQuickStep,
// dJointGroupEmpty (the loop of ode/demo/demo_space_stress.cpp:76-107,
// 363-369).  Used by bench.py to time one world through libodx.
//
//   frame_loop SCENE WARMUP STEPS ITERS [MU [ORDER [DYNAMIC [SWEEP_MODE]]]]
//
// ORDER 0 (default) sweeps the rows in the reference's order; 1 in colour
// phases (odxWorldSetRowOrder).  DYNAMIC 0 (default) runs exactly ITERS sweeps
// (dWorldSetQuickStepDynamicIterationParameters with zeros); 1 keeps
// dWorldCreate's dynamic iteration count.  SWEEP_MODE: odxWorldSetSweepMode.
// SCENE is a binary file: int32 nb, int32 ng, float bodies[nb][32],
// float geoms[ng][16] (the record layouts of include/odx.h).  The geoms go
// into a sweep-and-prune space (axes XYZ).  Prints one JSON object.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "ode.h"
#include "odx.h"

namespace {
struct Ctx {
    dWorldID world;
    dJointGroupID group;
    float mu;
    long long contacts = 0;
    long long pairs = 0;
};

void near_callback(void *data, dGeomID o1, dGeomID o2) {
    Ctx *c = static_cast<Ctx *>(data);
    c->pairs++;
    dBodyID b1 = dGeomGetBody(o1), b2 = dGeomGetBody(o2);
    dContact contact[4];
    const int n = dCollide(o1, o2, 4, &contact[0].geom, sizeof(dContact));
    for (int i = 0; i < n; i++) {
        contact[i].surface.mode = dContactApprox1;
        contact[i].surface.mu = c->mu;
        dJointID j = dJointCreateContact(c->world, c->group, &contact[i]);
        dJointAttach(j, b1, b2);
    }
    c->contacts += n;
}
}  // namespace

int main(int argc, char **argv) {
    if (argc < 5) {
        fprintf(stderr, "usage: %s SCENE WARMUP STEPS ITERS [MU [ORDER [DYNAMIC [SWEEP_MODE]]]]\n", argv[0]);
        return 2;
    }
    const int warmup = atoi(argv[2]), steps = atoi(argv[3]), iters = atoi(argv[4]);
    const float mu = argc > 5 ? (float)atof(argv[5]) : 0.5f;
    const int order = argc > 6 ? atoi(argv[6]) : 0;
    const int dynamic = argc > 7 ? atoi(argv[7]) : 0;
    const int sweep_mode = argc > 8 ? atoi(argv[8]) : 0;
    FILE *f = fopen(argv[1], "rb");
    if (!f) { perror("scene"); return 2; }
    int nb = 0, ng = 0;
    if (fread(&nb, 4, 1, f) != 1 || fread(&ng, 4, 1, f) != 1 || nb < 0 || ng < 0) { fprintf(stderr, "bad scene\n"); return 2; }
    std::vector<float> B((size_t)nb * 32), G((size_t)ng * 16);
    if (fread(B.data(), 4, B.size(), f) != B.size() || fread(G.data(), 4, G.size(), f) != G.size()) {
        fprintf(stderr, "short scene\n");
        return 2;
    }
    fclose(f);

    dInitODE2(0);
    Ctx c;
    c.world = dWorldCreate();
    c.group = dJointGroupCreate(0);
    c.mu = mu;
    dWorldSetGravity(c.world, 0, 0, -9.81f);
    dWorldSetQuickStepNumIterations(c.world, iters);
    odxWorldSetRowOrder(c.world, order);
    odxWorldSetSweepMode(c.world, sweep_mode);
    if (!dynamic) {
        const dReal zero = 0;
        dWorldSetQuickStepDynamicIterationParameters(c.world, &zero, &zero, &zero);
    }
    dSpaceID space = dSweepAndPruneSpaceCreate(nullptr, dSAP_AXES_XYZ);
    std::vector<dBodyID> bodies(nb);
    for (int i = 0; i < nb; i++) {
        const float *r = &B[(size_t)i * 32];
        dBodyID b = dBodyCreate(c.world);
        dBodySetPosition(b, r[0], r[1], r[2]);
        dQuaternion q = {r[4], r[5], r[6], r[7]};
        dBodySetQuaternion(b, q);
        dBodySetLinearVel(b, r[8], r[9], r[10]);
        dBodySetAngularVel(b, r[12], r[13], r[14]);
        dMass m;
        dMassSetParameters(&m, r[24], 0, 0, 0, r[25], r[26], r[27], r[28], r[29], r[30]);
        dBodySetMass(b, &m);
        bodies[i] = b;
    }
    for (int i = 0; i < ng; i++) {
        const float *g = &G[(size_t)i * 16];
        dGeomID h = nullptr;
        switch ((int)g[0]) {
        case 0: h = dCreateSphere(space, g[2]); break;
        case 1: h = dCreateBox(space, g[2], g[3], g[4]); break;
        case 2: h = dCreateCapsule(space, g[2], g[3]); break;
        default: h = dCreatePlane(space, g[2], g[3], g[4], g[5]); break;
        }
        if ((int)g[1] >= 0) dGeomSetBody(h, bodies[(int)g[1]]);
    }
    double collide_ms = 0, step_ms = 0;  // host wall time inside each call
    auto frame = [&]() {
        auto a = std::chrono::steady_clock::now();
        dSpaceCollide(space, &c, &near_callback);
        auto b = std::chrono::steady_clock::now();
        const int ok = dWorldQuickStep(c.world, 0.01f);
        auto e = std::chrono::steady_clock::now();
        dJointGroupEmpty(c.group);
        collide_ms += std::chrono::duration<double, std::milli>(b - a).count();
        step_ms += std::chrono::duration<double, std::milli>(e - b).count();
        return ok;
    };
    for (int s = 0; s < warmup; s++)
        if (!frame()) { fprintf(stderr, "dWorldQuickStep failed\n"); return 1; }
    c.contacts = c.pairs = 0;
    collide_ms = step_ms = 0;
    odxApiSetKernelTiming(1);
    auto t0 = std::chrono::steady_clock::now();
    for (int s = 0; s < steps; s++)
        if (!frame()) { fprintf(stderr, "dWorldQuickStep failed\n"); return 1; }
    auto t1 = std::chrono::steady_clock::now();
    const double ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    double kt[3] = {0, 0, 0};
    odxApiGetKernelTiming(kt);
    double zmin = 1e30;
    for (dBodyID b : bodies) zmin = dBodyGetPosition(b)[2] < zmin ? dBodyGetPosition(b)[2] : zmin;
    printf("{\"steps\": %d, \"warmup\": %d, \"total_ms\": %.6f, \"ms_per_step\": %.6f, \"pairs_per_step\": %.2f, "
           "\"contacts_per_step\": %.2f, \"min_z\": %.6f, \"kernel_ms_per_step\": %.6f, \"row_sweeps_per_step\": %.1f, "
           "\"row_order\": %d, \"dynamic_iterations\": %d, \"sweep_mode\": %d, \"collide_ms_per_step\": %.6f, "
           "\"quickstep_ms_per_step\": %.6f}\n",
           steps, warmup, ms, ms / steps, (double)c.pairs / steps, (double)c.contacts / steps, zmin,
           kt[0] > 0 ? kt[1] / kt[0] : 0.0, kt[0] > 0 ? kt[2] / kt[0] : 0.0, order, dynamic, sweep_mode,
           collide_ms / steps, step_ms / steps);
    dWorldDestroy(c.world);
    dCloseODE();
    return 0;
}
