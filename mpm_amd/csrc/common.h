// common.h -- host/device shared definitions of the MI355X MPM solver.
#pragma once

#include <stdint.h>

#include "mpm_hip.h"

#define MPM_BLOCK 4          // grid block edge (nodes); blocks are 4x4x4
#define MPM_BLOCK_NODES 64
#define MPM_MAX_MATERIALS 65536

// Grid field planes (each plane holds one float per padded node, in
// block-major order).  Fields 0..6 are what P2G scatters (mass, momentum
// stored in the reference's "velocity" slot, force); 7..9 hold velocity_star
// for the unfused sub-step path.
enum GridField {
    GF_MASS = 0,
    GF_VEL = 1,   // 1,2,3: momentum after P2G, velocity after updateVelocity
    GF_FORCE = 4, // 4,5,6
    GF_VSTAR = 7, // 7,8,9
    GF_PRIMARY = 7,
    GF_ALL = 10,
    GF_FOLDED = 4  // fused APIC step: mass, then p + dt f in planes 1..3
};

// Material constants shared by groups of particles (hardening xi,
// critical compression/stretch, constitutive model): the per-particle
// fields of MaterialPoint3D that do not change (include/material_point3d.h:19-21).
struct MaterialDev {
    float hardening;
    float crit_c;
    float crit_s;
    int32_t model;
};

// Pointers into device memory inside the descriptors below.  In the device
// pass they are typed as global-address-space pointers, so that accesses
// through a descriptor read from memory (the batch scene table) compile to
// global loads and stores; a generic pointer there would make every access
// a flat access, which also occupies the LDS wait counter.  The host pass
// sees plain pointers of the same size, so the layout is identical.
#if defined(__HIP_DEVICE_COMPILE__)
#define MPM_GLOBAL __attribute__((address_space(1)))
#else
#define MPM_GLOBAL
#endif

// Everything a kernel needs to know about one scene (grid + integrator
// constants).  Passed by value.
struct SceneDev {
    float ox, oy, oz;        // grid origin
    float tx, ty, tz;        // target = origin + h * res (Grid3DSettings)
    float h;                 // stride
    float inv_h;             // 1/h (weight gradients)
    float dt;
    float gravity;           // magnitude along -y
    float friction;          // boundary friction coefficient
    float blend;             // FLIP weight (PIC-FLIP)
    float dinv;              // APIC Dinv diagonal (k h^2)^-1, Eigen cofactor inverse
    float inv_cell_vol;      // 1/(h^3) as the reference computes it (initialize)
    int32_t rx, ry, rz;      // resolution
    int32_t nbx, nby, nbz;   // storage blocks per axis = ceil(res/4)
    int32_t kbx, kby, kbz;   // key blocks per axis = ((res+3)>>2)+1
    uint32_t escape_key;     // kbx*kby*kbz*64
    int32_t particle_collision;
    int64_t plane;           // floats per grid field plane (= nbx*nby*nbz*64)
    float MPM_GLOBAL* grid;        // GF_ALL (or GF_PRIMARY) planes
    uint32_t MPM_GLOBAL* touched;  // one flag per storage block
};

// Particle SoA (planar: component c of a vector field lives at base + c*N).
struct ParticlesDev {
    int64_t n;
    float MPM_GLOBAL* x;      // 3N
    float MPM_GLOBAL* v;      // 3N
    float MPM_GLOBAL* B;      // 9N (APIC) or null
    float MPM_GLOBAL* Fe;     // 9N row-major
    float MPM_GLOBAL* Fp;     // 9N
    float MPM_GLOBAL* Jp;     // N det(Fp), stored with Fp (the stress's hardening needs only it)
    float MPM_GLOBAL* m;      // N
    float MPM_GLOBAL* V0;     // N
    float MPM_GLOBAL* mu0;    // N
    float MPM_GLOBAL* lam0;   // N
    float MPM_GLOBAL* tau;    // 6N cached -V0 * P(Fe) Fe^T (xx, xy, xz, yy, yz, zz) for the next P2G
    uint16_t MPM_GLOBAL* mat; // N
    uint32_t MPM_GLOBAL* pid; // N original particle id
    const MaterialDev MPM_GLOBAL* mats;
};

// One scene as the item kernels see it.  A batch sorts the particles of
// all its scenes together: scene s owns global slots [pstart, pstart + n)
// and global keys key_offset + (its cell key); key_offset is a multiple of
// 64 so that key >> 6 (the block) never mixes scenes.
struct SceneEntry {
    SceneDev S;
    ParticlesDev P;
    int64_t pstart;
    unsigned long long MPM_GLOBAL* active;  // the scene's active-node counter
    uint32_t key_offset;
    uint32_t index;
    // the fused batch step: the second grid buffer, into which the step's
    // G2P scatters the next step's P2G (swapped with S.grid after the
    // step), and whether S.grid already holds the P2G of the current
    // particles (prescat; the step then runs no P2G for this scene)
    float MPM_GLOBAL* grid2;
    uint32_t MPM_GLOBAL* touched2;
    uint32_t prescat;
    uint32_t pad_;
};

// Incremental cell sort (incsort.hip) of a large scene or a batch: the
// transfer appends each particle whose key changed -- new key, global slot,
// old sorted position -- to one of kMoveLists sub-lists (by wave), the
// next step's sort is the incremental update when the plan words allow it,
// else the full radix sort.  Sub-lists, with their counters 256 bytes
// apart: one device-wide counter serialises the ~10^5 appends of a batch
// step (its G2P took 1.31 ms instead of 0.96).
constexpr int kMoveMin = 16384;   // capacity of all sub-lists: max(kMoveMin, particles / 8)
constexpr int kMoveLists = 64;
constexpr int kMoveCountStride = 64;  // words between two sub-list counters
constexpr int kIncTile = 2048;    // snapshot entries per merge workgroup
constexpr int kIncStage = 512;    // removals or insertions a merge workgroup ranks in LDS
enum PlanWord {
    kPlanSnapOk = 1,    // the snapshot is the previous step's order
    kPlanIncSteps = 2,  // steps sorted by the update since creation
    kPlanCap = 3,       // capacity of one sub-list
    kPlanOverflow = 4,  // a sub-list overflowed, or a snapshot tile drew more than kIncStage entries
    kPlanLists = 64,    // the sub-list counters: word kPlanLists + r * kMoveCountStride
    kPlanWords = kPlanLists + kMoveLists * kMoveCountStride
};
// The step's choice, read by every sort kernel: the update when the
// snapshot is the previous step's order and nothing overflowed.  The words
// change only between sorts (the transfer appends and may set the
// overflow; the update's offsets kernel may set it before the radix
// kernels run; the item count of the step resets them).
__host__ __device__ inline bool plan_incremental(const uint32_t* plan) {
    return plan[kPlanSnapOk] != 0u && plan[kPlanOverflow] == 0u;
}
struct MoveList {
    uint32_t *key, *slot, *pos;  // appended by the transfer: sub-list r at [r * cap, r * cap + count_r)
    uint32_t* lb;                // per entry: its place among the snapshot
    uint32_t* tcount;            // [2][tiles + 1]: removed positions, insertion places per snapshot tile (zero between steps)
    uint32_t* tnext;             // [2][tiles + 1]: bucket cursors
    uint32_t *del_b, *ins_b;     // removed positions, entry indices: bucketed by tile (capacity each)
    // a batch (else nranges 0): range_start[0..nranges] the first global slot
    // of each member and the total; a member's particles also fill exactly
    // that range of sorted positions (its keys are one range of the key space)
    const uint32_t* range_start;
    uint32_t nranges;
};

// Work items of the block-tiled kernels: a contiguous run of sorted
// particles that share one key block (and so one scene).
struct ItemsDev {
    const uint32_t* sorted_keys;  // N (global keys)
    const uint32_t* perm;         // N, sorted position -> global slot
    const uint32_t* start;        // item start positions
    const uint32_t* blk;          // item key block (scene-local), so the item needs no key load
    const uint32_t* count;        // device counter: number of items
    const uint32_t* scene;        // per item scene index (batches) or null
    int64_t n;
    // the incremental sort (else null): the transfer's first kernel copies
    // the sorted keys and global slots to snap_k/snap_p, G2P appends moved
    // particles (move_cap: capacity of one sub-list)
    uint32_t* snap_k;
    uint32_t* snap_p;
    uint32_t* plan;
    const MoveList* moves;  // device copy of the list's pointers
    uint32_t move_cap;
};
