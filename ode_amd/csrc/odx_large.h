// odx_large.h -- device view of one large world (hash-space broadphase,
// colour-parallel SOR-LCP over the whole chip).
#pragma once
#include <hip/hip_runtime.h>

namespace odx {

// Device counters of a large-world step.
enum {
    LG_PAIRS = 0,        // candidate pairs kept
    LG_CONTACTS,         // contacts kept
    LG_ROWS,             // constraint rows
    LG_NCOL,             // colours of the sweep schedule
    LG_PAIR_OVF,         // pairs beyond the pair capacity (dropped)
    LG_CONTACT_OVF,      // contacts beyond the contact capacity (dropped)
    LG_COLOUR_OVF,       // contacts left uncoloured after LG_MAXCOL colours
    LG_BARRIER_ERR,      // a grid barrier wait ran out (the step is invalid)
    LG_SUBROUNDS,        // claim/commit rounds the colouring took
    LG_NCOUNTERS = 16
};

constexpr int LG_MAXCOL = 128;        // colours of the sweep schedule
constexpr int LG_MAXSUB = 32;         // claim/commit rounds per colour
constexpr int LG_PAIR_KEEP = 16;      // partners per geom the counting pass keeps for the writing pass
constexpr long long LG_SPIN_LIMIT = 1ll << 21;  // polls of one barrier wait (~1 s)
constexpr int LG_BLOCK = 256;         // threads per workgroup of the cooperative kernels
// words of the barrier array: generation, 8 group counters (64 B apart), then
// the per-colour fill counters of the colouring
enum { LG_BAR_GEN = 0, LG_BAR_GROUP = 32, LG_BAR_FILL = 160, LG_BAR_WORDS = LG_BAR_FILL + LG_MAXCOL };

struct LargeDev {
    int nb, ng, nbig, hmask;          // bodies, geoms, big geoms, hash buckets - 1 (power of two)
    float inv_cell;                   // 1 / cell edge; the edge bounds every finite AABB
    int pair_cap, contact_cap;
    float *aabb;                      // [ng * 6]
    const int *isbig;                 // [ng] 1: infinite or oversized AABB, tested against every geom
    const int *big;                   // [nbig] geom index of each big geom (ascending)
    int4 *gcell;                      // [ng] cell x, y, z and hash bucket of each finite geom
    int *bucket;                      // [hmask + 2] counts -> starts -> ends
    int *bucket_geom;                 // [ng] finite geoms grouped by bucket, ascending index in a bucket
    int4 *slot_cell;                  // [ng] per bucket slot: the geom's cell x, y, z and the geom
    float4 *slot_box;                 // [2 * ng] per bucket slot: the geom's AABB, then its body (as float)
    int *pair_cnt;                    // [ng + 1] pairs per geom -> first pair of each geom
    int2 *pairs;                      // [pair_cap] (lower geom, higher geom)
    int *pair_keep;                   // [ng * LG_PAIR_KEEP] partners of each geom found by the counting pass
    int *contact_cnt;                 // [pair_cap + 1] contacts per pair -> first contact of each pair
    int *row_cnt;                     // [contact_cap + 1] rows per contact -> first row of each contact
    int *counters;                    // [LG_NCOUNTERS]
    int2 *cbody;                      // [contact_cap] bodies of each contact (body 2 = -1: static)
    int *colour;                      // [contact_cap]
    int *taken;                       // [nb] last colour that holds the body
    unsigned long long *owner;        // [nb] claim of the current round (stamp | priority | contact)
    int *elig;                        // [LG_MAXCOL * LG_MAXSUB] contacts taking part in each round
    int *col_start;                   // [LG_MAXCOL + 1] contacts per colour -> first slot of each colour
    int *corder;                      // [contact_cap] contacts in colour order
    int4 *items;                      // [contact_cap] per slot: first row, first-pass rows | rows << 8, body 1, body 2
    unsigned *bar;                    // [LG_BAR_WORDS] grid barrier counters, colour fill counters
    float *fc;                        // [nb * 8] body accumulators of the solve (6 used)
    float4 *ftot;                     // [2 * nb] total force, then total torque, of the body stage
    int *scan_tmp;                    // block sums of the scans
    int spin_limit;                   // polls of one barrier wait before it fails (LG_SPIN_LIMIT)
    int colour_cap;                   // contacts per colour at most (the solve's threads; 0: no cap)
    int one_pass;                     // 1: a contact's rows together in one colour phase per sweep (joint
                                      // order); 0: ReorderPrep's two passes (rows without findex, then findex rows)
};

}  // namespace odx
