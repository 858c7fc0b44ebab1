// solver_impl.h -- host-side state of one scene handle, shared by the
// C-ABI translation units (capi.hip: one scene; batch.hip: many scenes in
// one launch).  Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "launch.h"

namespace mpm {
namespace host {


inline thread_local std::string g_err;

inline int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(call)                                                                                 \
    do {                                                                                              \
        hipError_t e_ = (call);                                                                       \
        if (e_ != hipSuccess) {                                                                       \
            return fail(MPM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_) + " at " +     \
                                         __FILE__ + ":" + std::to_string(__LINE__));                  \
        }                                                                                             \
    } while (0)

// PH_RAW: P2G sums (mass, momentum, force) of a fused step; PH_FOLDED: the
// fused APIC step's mass and p + dt f (momentum and force not separable).
enum GridPhase { PH_CLEAR = 0, PH_MASS = 1, PH_RAW = 2, PH_SUBSTEP = 3, PH_FOLDED = 4 };

struct Pending {
    std::string name;
    hipEvent_t a, b;
};

struct Solver {
    mpm_desc desc{};
    int interp = 0, apic = 0;
    int64_t N = 0;
    int device = 0;
    hipStream_t st = nullptr;
    SceneDev S{};
    ParticlesDev P{};
    ItemsDev I{};
    std::vector<MaterialDev> mats;
    MaterialDev* mats_dev = nullptr;
    int64_t nsb = 0;
    int key_bits = 1;
    uint32_t* kin = nullptr;   // sort input keys (slot order); written by k_keys or the fused G2P
    uint32_t* iota = nullptr;  // sort input values: slot index
    uint32_t *ka = nullptr, *va = nullptr, *kb = nullptr, *vb = nullptr;  // ping-pong / result
    bool keys_valid = false;   // kin matches the current positions
    uint32_t* hist = nullptr;
    uint32_t* scan_temp = nullptr;
    uint32_t* flags = nullptr;
    uint32_t* items = nullptr;
    uint32_t* counters = nullptr;                 // [0] item count
    unsigned long long* acc = nullptr;            // [0] active sum, [1] scratch, [2] escaped
    bool sort_valid = false, grid_dirty = false, vstar_dirty = false, count_on_clear = false;
    bool stress_valid = false;  // P.tau matches the particle state
    // force planes (GF_FORCE..) may be non-zero: a clear must cover them
    bool force_dirty = false;
    // keep momentum and force apart in the fused APIC step (7-plane P2G) so
    // that mpm_read_grid shows them after a step; off = the folded grid
    bool record_grid = false;
    int phase = PH_CLEAR;
    int64_t steps = 0;
    int item_wgs = 0;  // 0: persistent grids sized by occupancy
    double last_step_ms = 0.0;
    std::vector<void*> allocs;
    int64_t device_bytes = 0;
    // profiling
    int prof = 0;  // 0 off, 1 every kernel, 2 the item kernels (p2g, g2p, g2p2g) and the store reorder only
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;
    std::map<std::string, std::pair<int64_t, double>> prof_tot;
    std::vector<std::string> prof_order;
    // graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int64_t graph_captures = 0;
    ItemsDev graph_I{};  // the items the captured step reads (its sort buffers)
    // batch membership: while in a batch the solver works on the batch's
    // stream (own_st keeps its own)
    struct Batch* batch = nullptr;
    hipStream_t own_st = nullptr;
    // incremented whenever particle positions change (steps, advection,
    // state writes): a batch recomputes the member's sort keys on a change
    uint64_t version = 0;
    // density raster (allocated on first use): one mass plane with its own
    // touched flags, block ranks (nsb + 1 words, the last is the count) and
    // scan scratch; valid while dens_version == version
    float* dens = nullptr;
    uint32_t* dens_touched = nullptr;
    uint32_t* dens_rank = nullptr;
    uint32_t* dens_temp = nullptr;
    bool dens_dirty = false;
    uint64_t dens_version = ~0ull;
    // render export (allocated on first use): float4 positions staged on
    // the device, copied to the host on a second stream
    float* pos4 = nullptr;
    hipStream_t copy_st = nullptr;
    // the fused step clears the second grid buffer on aux_st while the sort
    // runs on st (fork/join events)
    hipStream_t aux_st = nullptr;
    hipEvent_t aux_fork = nullptr, aux_join = nullptr;
    hipEvent_t pos4_packed = nullptr, pos4_copied = nullptr;
    bool pos4_pending = false;
    // store reorder: every reorder_interval steps (0 = never) the particle
    // store is permuted into the current cell order; scratch 9N words
    int reorder_interval = 100;
    int64_t last_reorder = 0;
    uint32_t* reorder_scratch = nullptr;
    // incremental cell sort of the fused step (scenes of at least
    // kIncMinParticles): snapshot of the sorted order, plan words, move list;
    // snap_ok: the device snapshot and move list describe the current keys
    bool incr = false;
    bool snap_ok = false;
    uint32_t* snap_k = nullptr;
    uint32_t* snap_p = nullptr;
    uint32_t* plan = nullptr;
    MoveList moves{};
    MoveList* moves_dev = nullptr;  // device copy of `moves` (the kernels' view)
    uint32_t move_cap = 0;          // capacity of one sub-list (also in the plan words)
    // how often the update ran lately: the incremental-sort count read back
    // every few steps through pinned memory (never waited for); mostly
    // updates -> the full sort's launches use fewer passes (few_passes)
    uint32_t* inc_host = nullptr;
    hipEvent_t inc_ev = nullptr;
    bool inc_pending = false;
    int64_t inc_copy_step = 0, inc_prev_step = -1;
    uint32_t inc_prev = 0;
    bool few_passes = false;
    // fused transfer (APIC folded step, linear or quadratic): G2P scatters
    // the next step's P2G into the second grid buffer (same size as the
    // first: the two alternate), which then becomes the grid.  prescat:
    // S.grid holds the P2G of the current particle state and grid2 the grid
    // the last step read; grid2_dirty: grid2 holds values to clear,
    // grid2_count: they are a step's grid (its active nodes are counted).
    // capturing: a graph capture runs the two-kernel path.
    bool fuse_transfer = true;
    bool capturing = false;
    bool prescat = false;
    bool grid2_dirty = false;
    bool grid2_count = false;
    decltype(SceneDev::grid) grid2 = nullptr;
    decltype(SceneDev::touched) touched2 = nullptr;
    decltype(SceneDev::grid) graph_grid = nullptr;  // the grid buffer the captured step uses
    // member of a batch that fuses its transfer: the batch table's grid of
    // this scene holds the P2G of its current particles (the table's grid
    // and grid2 may be swapped relative to S.grid / grid2 here, by the
    // batch's parity)
    bool bprescat = false;
};

// Below this many particles the fused step sorts in full every step: the
// radix sort of a small scene is a few launches and the update would add
// more than it saves.
#ifndef MPM_INC_MIN_PARTICLES
#define MPM_INC_MIN_PARTICLES (int64_t(1) << 20)
#endif
constexpr int64_t kIncMinParticles = MPM_INC_MIN_PARTICLES;


int dalloc(Solver& s, void** p, size_t bytes);
void dfree(Solver& s, void* p);
ProfileHook hook(Solver& s);
void prof_collect(Solver& s);
struct Scope {
    const ProfileHook& h;
    const char* n;
    Scope(const ProfileHook& h_, const char* n_) : h(h_), n(n_) { h.begin(n); }
    ~Scope() { h.end(n); }
};
void run_clear(Solver& s);
void ensure_stress(Solver& s);
void destroy(Solver* s);
int finish(Solver& s);
// batch.hip: removes a scene that is destroyed while a batch member.
void batch_forget(struct Batch* b, Solver* s);
// batch.hip: a member's settings changed; the batch recomputes its layout.
void batch_invalidate(struct Batch* b);
// batch.hip: before a member's own call touches its grid or particles, the
// member leaves the batch's fused state (its grids as a two-kernel step
// leaves them).
void batch_unfuse_member(struct Batch* b, Solver& s);
// The second grid buffer (all planes, cleared) of the fused transfer.
int ensure_grid2_buffers(Solver& s);
// Planes of the grid that P2G writes in the fused step.
inline bool fused_folds(const Solver& s) { return s.apic && !s.record_grid; }
// Before anything but a fused step touches the particles or the grid: the
// grid becomes again the one the last step read (as the two-kernel step
// leaves it) and the next step's pre-scattered sums wait in grid2 to be
// cleared.
inline void unfuse(Solver& s) {
    if (s.batch) batch_unfuse_member(s.batch, s);
    if (!s.prescat) return;
    std::swap(s.S.grid, s.grid2);
    std::swap(s.S.touched, s.touched2);
    s.grid2_dirty = true;
    s.grid2_count = false;  // never a step's grid
    s.prescat = false;
}

}  // namespace host
}  // namespace mpm
