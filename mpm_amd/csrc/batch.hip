// batch.hip -- many independent scenes advanced by one set of launches.
//
// A batch holds scene handles of one interpolation and transfer scheme.
// Each member keeps its own particle store and grid; the batch adds one
// global key space in which scene s owns the key range
//   [key_offset_s, key_offset_s + escape_key_s + 64)   (64-aligned),
// and one global slot space in which it owns [pstart_s, pstart_s + n_s).
// One step is then five launch groups over all scenes together:
//   batched clear of every touched grid block -> one stable radix sort of
//   all keys -> work items (with their scene index) -> P2G -> G2P
// where P2G and G2P are the persistent tiled kernels of the single-scene
// path reading their scene's descriptors from a device table.  Results per
// scene equal the scene's own mpm_step up to the order of the float
// atomics in P2G.  The step can be captured once as a hipGraph and
// replayed.  Launch parameters (particles per work item, G2P workgroup
// size, persistent workgroup count) come from a tuning table keyed by
// (interpolation, transfer, size bucket), filled by mpm_batch_autotune.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <tuple>
#include <vector>

#include "devcache.h"
#include "launch.h"
#include "solver_impl.h"

namespace mpm {
namespace host {

struct Tune {
    int item_max = 256;  // what the autotuner picks for the 256-scene uneven batch
    int g2p_wg = 128;  // as the single-scene G2P
    int wgs = 0;  // 0: one resident round per kernel (occupancy x CUs)
    int fuse = 0;  // the fused transfer (APIC, folded grid, quadratic stencil)
    bool operator==(const Tune& o) const {
        return item_max == o.item_max && g2p_wg == o.g2p_wg && wgs == o.wgs && fuse == o.fuse;
    }
};

struct TuneEntry {
    Tune t;
    double ms = 0.0;
};

std::mutex g_tune_mu;
std::map<std::tuple<int, int, int>, TuneEntry> g_tune;  // (interp, apic, bucket) -> choice

int size_bucket(int64_t n) {
    int b = 0;
    while (b < 62 && (n >> (b + 1)) != 0) ++b;
    return b;
}

struct Batch {
    int interp = 0, apic = 0;
    int device = 0;
    int cus = 1;
    hipStream_t st = nullptr;
    // eager steps clear the grids on aux while the sort runs on st (they
    // share no buffer); P2G waits for the join event
    hipStream_t aux = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    std::vector<Solver*> members;
    std::vector<uint64_t> synced_version;  // member version whose keys are in kin
    // layout
    bool layout_valid = false;
    int64_t ntot = 0, cap = 0;
    int64_t total_blocks = 0;
    int key_bits = 1;
    std::vector<SceneEntry> tab_host;
    std::vector<uint32_t> koff_host;
    std::vector<int64_t> bstart_host;
    SceneEntry* tab = nullptr;
    uint32_t* key_offsets = nullptr;
    int64_t* block_start = nullptr;
    uint32_t* range_start = nullptr;  // tab_cap + 1 words: each member's first global slot, then ntot
    std::vector<uint32_t> range_host;
    int tab_cap = 0;
    uint32_t *kin = nullptr, *iota = nullptr, *ka = nullptr, *va = nullptr, *kb = nullptr, *vb = nullptr;
    uint32_t *flags = nullptr, *items = nullptr, *item_scene = nullptr, *hist = nullptr, *scan_temp = nullptr;
    uint32_t* counters = nullptr;
    ItemsDev I{};
    Tune tune;
    bool tune_fixed = false;  // set explicitly: the tuning table is not consulted
    bool fold = false;        // APIC with no member recording its grid: 4-plane P2G
    // transfer fusion (tune.fuse; APIC, folded grid, quadratic stencil): one
    // kernel runs the step's G2P and the next step's P2G into each member's
    // second grid, then a small kernel swaps every entry's grid and grid2 on
    // the device.  parity: swaps since the members' host pointers last
    // matched the table (odd: each member's S.grid / grid2 are the other way
    // round there).
    int parity = 0;
    // every reorder_interval steps (0 = never) each member's store is
    // permuted into its range of the global cell order, a chunk of members
    // per step (chunks: about ntot / reorder_interval particles each, bounds
    // in global positions, the current chunk index on the device); the
    // autotuner's full reorders use reorder_scratch (9 * cap words)
    int reorder_interval = 16;
    int64_t last_reorder = 0;
    uint32_t* reorder_scratch = nullptr;
    int64_t reorder_cap = 0;
    int nchunks = 0;
    int64_t chunk_max = 0;
    int64_t* chunk_bounds = nullptr;  // nchunks + 1
    int chunk_bounds_cap = 0;
    std::vector<int64_t> chunk_host;  // host copy of the bounds (the source of their upload)
    uint32_t* chunk_rr = nullptr;
    uint32_t* slot_scene = nullptr;  // cap words: each global slot's member
    int64_t slot_scene_cap = 0;
    uint32_t* chunk_scratch = nullptr;
    int64_t chunk_scratch_cap = 0;
    // incremental sort (incsort.hip; on unless switched off): snapshot of the
    // sorted order, plan words, move list; snap_ok: the device snapshot and
    // list describe the current keys (else the next step sorts in full)
    bool incr = true;
    bool snap_ok = false;
    uint32_t *snap_k = nullptr, *snap_p = nullptr, *plan = nullptr;
    MoveList moves{};
    MoveList* moves_dev = nullptr;
    uint32_t move_cap = 0;  // capacity of one sub-list (also in the plan words)
    // per-member health words (3 per member), allocated on first use
    uint32_t* health = nullptr;
    int health_cap = 0;
    // graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    Tune graph_tune;
    // stats and timing
    int64_t steps = 0;
    double last_step_ms = 0.0;
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    std::vector<hipEvent_t> event_pool;
    std::vector<float> step_ms;
    std::vector<void*> allocs;
    int64_t device_bytes = 0;
};

// Device memory from the block cache (devcache.h), as for scenes.
int balloc(Batch& b, void** p, size_t bytes) {
    size_t got = 0;
    hipError_t e = devcache::alloc(p, bytes, &got);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(MPM_ERR_OOM, "hipMalloc(" + std::to_string(bytes) + " B) failed: " + hipGetErrorString(e));
    }
    b.allocs.push_back(*p);
    b.device_bytes += (int64_t)got;
    return MPM_OK;
}

// The batch stream is synchronised first: the block may be handed to
// another stream at once.
void bfree(Batch& b, void*& p) {
    if (!p) return;
    for (size_t i = 0; i < b.allocs.size(); ++i)
        if (b.allocs[i] == p) { b.allocs.erase(b.allocs.begin() + (long)i); break; }
    if (b.st) (void)hipStreamSynchronize(b.st);
    b.device_bytes -= (int64_t)devcache::release(p);
    p = nullptr;
}

void drop_graph(Batch& b) {
    if (b.gexec) { (void)hipGraphExecDestroy(b.gexec); b.gexec = nullptr; }
    if (b.graph) { (void)hipGraphDestroy(b.graph); b.graph = nullptr; }
}

int effective_wgs(const Batch& b) { return b.tune.wgs; }

// Per-particle buffers sized for at least n particles (grown by 1.5x so
// that a stream of joins does not reallocate every time).
int ensure_capacity(Batch& b, int64_t n) {
    if (n <= b.cap && b.kin) return MPM_OK;
    (void)hipStreamSynchronize(b.st);
    int64_t c = std::max<int64_t>(std::max<int64_t>(n, 1), b.cap + b.cap / 2);
    uint32_t** bufs[] = {&b.kin, &b.iota, &b.ka, &b.va, &b.kb, &b.vb, &b.flags, &b.item_scene, &b.snap_k, &b.snap_p};
    uint32_t** lists[] = {&b.moves.key, &b.moves.slot, &b.moves.pos, &b.moves.lb, &b.moves.del_b, &b.moves.ins_b};
    uint32_t** tiles[] = {&b.moves.tcount, &b.moves.tnext};
    for (auto pp : bufs) bfree(b, *(void**)pp);
    for (auto pp : lists) bfree(b, *(void**)pp);
    for (auto pp : tiles) bfree(b, *(void**)pp);
    bfree(b, *(void**)&b.items);
    bfree(b, *(void**)&b.hist);
    bfree(b, *(void**)&b.scan_temp);
    int rc;
    for (auto pp : bufs)
        if ((rc = balloc(b, (void**)pp, (size_t)c * 4))) return rc;
    // the incremental sort's list and tile counts for up to c particles
    b.move_cap = inc_sublist_cap(c);
    for (auto pp : lists)
        if ((rc = balloc(b, (void**)pp, (size_t)kMoveLists * b.move_cap * 4))) return rc;
    for (auto pp : tiles)
        if ((rc = balloc(b, (void**)pp, (size_t)2 * (inc_tiles(c) + 1) * 4))) return rc;
    // the merge keeps the counts zero afterwards
    HIP_TRY(hipMemsetAsync(b.moves.tcount, 0, (size_t)2 * (inc_tiles(c) + 1) * 4, b.st));
    if (!b.moves_dev && (rc = balloc(b, (void**)&b.moves_dev, sizeof(MoveList)))) return rc;
    if (!b.plan && (rc = balloc(b, (void**)&b.plan, kPlanWords * 4))) return rc;
    HIP_TRY(hipMemcpyAsync(b.moves_dev, &b.moves, sizeof(MoveList), hipMemcpyHostToDevice, b.st));
    HIP_TRY(hipMemsetAsync(b.plan, 0, kPlanWords * 4, b.st));  // no snapshot yet
    HIP_TRY(hipMemcpyAsync(b.plan + kPlanCap, &b.move_cap, 4, hipMemcpyHostToDevice, b.st));
    b.snap_ok = false;
    if ((rc = balloc(b, (void**)&b.items, (size_t)c * 8))) return rc;  // starts + blocks
    const int64_t hw = radix_hist_words(c);
    if ((rc = balloc(b, (void**)&b.hist, (size_t)hw * 4))) return rc;
    const int64_t tw = std::max(scan_temp_words(hw), scan_temp_words(c));
    if ((rc = balloc(b, (void**)&b.scan_temp, (size_t)tw * 4))) return rc;
    b.cap = c;
    return MPM_OK;
}

SceneEntry make_entry(const Solver& s, int64_t pstart, uint32_t koff, uint32_t index) {
    SceneEntry e;
    std::memset(&e, 0, sizeof(e));  // padding too: entries are compared bytewise
    std::memcpy(&e.S, &s.S, sizeof(SceneDev));
    std::memcpy(&e.P, &s.P, sizeof(ParticlesDev));
    e.pstart = pstart;
    e.active = (decltype(e.active))s.acc;  // acc[0]: active-node sum over steps
    e.key_offset = koff;
    e.index = index;
    e.grid2 = s.grid2;
    e.touched2 = s.touched2;
    e.prescat = s.bprescat ? 1u : 0u;
    return e;
}

bool can_fuse(const Batch& b) { return b.apic && b.fold && b.interp == MPM_INTERP_QUADRATIC_BSPLINE; }
bool fuses(const Batch& b) { return b.tune.fuse != 0 && can_fuse(b); }

// The members' host pointers and the host table made equal to the device
// table again (after an odd number of fused steps each entry has its grid
// buffers the other way round).
void normalize(Batch& b) {
    if (!b.parity) return;
    for (size_t i = 0; i < b.members.size(); ++i) {
        Solver& s = *b.members[i];
        std::swap(s.S.grid, s.grid2);
        std::swap(s.S.touched, s.touched2);
        if (i < b.tab_host.size()) {
            SceneEntry& e = b.tab_host[i];
            std::swap(e.S.grid, e.grid2);
            std::swap(e.S.touched, e.touched2);
        }
    }
    b.parity = 0;
}

// The reorder chunks of the current layout: consecutive members, closed
// once they hold ntot / reorder_interval particles (a larger member is a
// chunk of its own); their position bounds, each slot's member, the scratch
// of the largest chunk, and the chunk index reset to 0.  No chunks when the
// reorder is off.
int layout_chunks(Batch& b) {
    b.nchunks = 0;
    b.chunk_max = 0;
    const int ns = (int)b.members.size();
    if (b.reorder_interval <= 0 || ns == 0 || b.ntot == 0) return MPM_OK;
    const int64_t target = (b.ntot + b.reorder_interval - 1) / b.reorder_interval;
    std::vector<int64_t>& bounds = b.chunk_host;
    bounds.assign(1, 0);
    int64_t acc = 0;
    for (int i = 0; i < ns; ++i) {
        acc += b.members[(size_t)i]->N;
        if (acc - bounds.back() >= target || i == ns - 1) bounds.push_back(acc);
    }
    const int k = (int)bounds.size() - 1;
    int64_t mx = 0;
    for (int c = 0; c < k; ++c) mx = std::max(mx, bounds[(size_t)c + 1] - bounds[(size_t)c]);
    int rc;
    if (k + 1 > b.chunk_bounds_cap) {
        (void)hipStreamSynchronize(b.st);
        bfree(b, *(void**)&b.chunk_bounds);
        if ((rc = balloc(b, (void**)&b.chunk_bounds, (size_t)(k + 1) * 8))) return rc;
        b.chunk_bounds_cap = k + 1;
    }
    if (!b.chunk_rr && (rc = balloc(b, (void**)&b.chunk_rr, 64))) return rc;
    if (b.slot_scene_cap < b.cap) {
        (void)hipStreamSynchronize(b.st);
        bfree(b, *(void**)&b.slot_scene);
        if ((rc = balloc(b, (void**)&b.slot_scene, (size_t)b.cap * 4))) return rc;
        b.slot_scene_cap = b.cap;
    }
    if (b.chunk_scratch_cap < mx) {
        (void)hipStreamSynchronize(b.st);
        bfree(b, *(void**)&b.chunk_scratch);
        if ((rc = balloc(b, (void**)&b.chunk_scratch, (size_t)mx * kChunkScratchWords * 4))) return rc;
        b.chunk_scratch_cap = mx;
    }
    HIP_TRY(hipMemcpyAsync(b.chunk_bounds, bounds.data(), (size_t)(k + 1) * 8, hipMemcpyHostToDevice, b.st));
    HIP_TRY(hipMemsetAsync(b.chunk_rr, 0, 4, b.st));
    launch_slot_scenes(b.tab, ns, b.ntot, b.slot_scene, b.st);
    b.nchunks = k;
    b.chunk_max = mx;
    return MPM_OK;
}

// Recompute the global layout of the members, (re)upload the tables and
// the sort input (keys with offsets, identity values).
int relayout(Batch& b) {
    drop_graph(b);
    normalize(b);  // the table is rebuilt from the members' host pointers
    const int ns = (int)b.members.size();
    b.fold = b.apic != 0;
    for (const Solver* m : b.members) b.fold = b.fold && !m->record_grid;
    b.tab_host.resize((size_t)ns);
    b.koff_host.resize((size_t)ns);
    b.bstart_host.assign((size_t)ns + 1, 0);
    b.synced_version.assign((size_t)ns, ~0ull);
    int64_t n = 0, blocks = 0;
    uint64_t koff = 0;
    for (int i = 0; i < ns; ++i) {
        const Solver& s = *b.members[(size_t)i];
        const uint64_t span = (uint64_t)s.S.escape_key + 64u;
        if (koff + span > (1ull << 32))
            return fail(MPM_ERR_UNSUPPORTED, "batch key space exhausted (the scenes' grids total more than 2^32 keys)");
        b.tab_host[(size_t)i] = make_entry(s, n, (uint32_t)koff, (uint32_t)i);
        b.koff_host[(size_t)i] = (uint32_t)koff;
        b.bstart_host[(size_t)i] = blocks;
        n += s.N;
        blocks += s.nsb;
        koff += span;
    }
    b.bstart_host[(size_t)ns] = blocks;
    if (n >= (int64_t)0xFFFFFFFFll) return fail(MPM_ERR_UNSUPPORTED, "batch holds too many particles");
    b.ntot = n;
    b.total_blocks = blocks;
    int bits = 1;
    while (bits < 32 && (koff > 0 ? ((koff - 1) >> bits) != 0 : false)) ++bits;
    b.key_bits = bits;
    int rc;
    if ((rc = ensure_capacity(b, n))) return rc;
    if (ns > b.tab_cap || !b.tab) {
        (void)hipStreamSynchronize(b.st);
        bfree(b, *(void**)&b.tab);
        bfree(b, *(void**)&b.key_offsets);
        bfree(b, *(void**)&b.block_start);
        bfree(b, *(void**)&b.range_start);
        const int c = std::max(ns, std::max(16, b.tab_cap * 2));
        if ((rc = balloc(b, (void**)&b.tab, (size_t)c * sizeof(SceneEntry)))) return rc;
        if ((rc = balloc(b, (void**)&b.key_offsets, (size_t)c * 4))) return rc;
        if ((rc = balloc(b, (void**)&b.block_start, (size_t)(c + 1) * 8))) return rc;
        if ((rc = balloc(b, (void**)&b.range_start, (size_t)(c + 1) * 4))) return rc;
        b.tab_cap = c;
    }
    if (ns > 0) {
        HIP_TRY(hipMemcpyAsync(b.tab, b.tab_host.data(), (size_t)ns * sizeof(SceneEntry), hipMemcpyHostToDevice, b.st));
        HIP_TRY(hipMemcpyAsync(b.key_offsets, b.koff_host.data(), (size_t)ns * 4, hipMemcpyHostToDevice, b.st));
    }
    HIP_TRY(hipMemcpyAsync(b.block_start, b.bstart_host.data(), (size_t)(ns + 1) * 8, hipMemcpyHostToDevice, b.st));
    b.range_host.resize((size_t)ns + 1);
    for (int i = 0; i < ns; ++i) b.range_host[(size_t)i] = (uint32_t)b.tab_host[(size_t)i].pstart;
    b.range_host[(size_t)ns] = (uint32_t)n;
    HIP_TRY(hipMemcpyAsync(b.range_start, b.range_host.data(), (size_t)(ns + 1) * 4, hipMemcpyHostToDevice, b.st));
    b.moves.range_start = b.range_start;
    b.moves.nranges = (uint32_t)ns;
    launch_iota(b.iota, n, b.st);
    b.I = ItemsDev{};
    b.I.start = b.items;
    b.I.blk = b.items + n;
    b.I.count = b.counters;
    b.I.scene = b.item_scene;
    b.I.n = n;
    if (b.incr) {
        b.I.snap_k = b.snap_k;
        b.I.snap_p = b.snap_p;
        b.I.plan = b.plan;
        b.I.moves = b.moves_dev;
        b.I.move_cap = b.move_cap;
    }
    b.snap_ok = false;  // slots and keys laid out anew
    if (!b.tune_fixed) {
        std::lock_guard<std::mutex> g(g_tune_mu);
        auto it = g_tune.find({b.interp, b.apic, size_bucket(n)});
        b.tune = it != g_tune.end() ? it->second.t : Tune{};
    }
    if ((rc = layout_chunks(b))) return rc;
    HIP_TRY(hipStreamSynchronize(b.st));  // host tables must outlive the copies
    b.layout_valid = true;
    return MPM_OK;
}

// Bring every member's device state in line with the batch before a step:
// table entries (device pointers change on e.g. a material upload), grids
// left dirty by the member's own sub-steps, the stress cache, and the sort
// keys of members whose particles moved outside the batch.
int sync_members(Batch& b) {
    if (!b.layout_valid) {
        int rc = relayout(b);
        if (rc) return rc;
    }
    bool upload = false;
    const bool fuse = fuses(b);
    for (size_t i = 0; i < b.members.size(); ++i) {
        Solver& s = *b.members[i];
        if (!fuse) {
            batch_unfuse_member(&b, s);  // fusion turned off (tune) or no longer possible
            if (s.grid2) {  // its second grid is not needed while a member (it returns on first use)
                // the scene's own captured step may use either buffer: it is captured again
                if (s.gexec) { (void)hipGraphExecDestroy(s.gexec); s.gexec = nullptr; }
                if (s.graph) { (void)hipGraphDestroy(s.graph); s.graph = nullptr; }
                s.graph_grid = nullptr;
                dfree(s, s.grid2);
                dfree(s, s.touched2);
                s.grid2 = nullptr;
                s.touched2 = nullptr;
                s.grid2_dirty = false;
            }
        }
        if (fuse) {
            int rc = ensure_grid2_buffers(s);
            if (rc) return rc;
        }
        SceneEntry e = make_entry(s, b.tab_host[i].pstart, b.tab_host[i].key_offset, (uint32_t)i);
        if (std::memcmp(&e, &b.tab_host[i], sizeof(SceneEntry)) != 0) {
            b.tab_host[i] = e;
            upload = true;
        }
        if (!s.bprescat) {  // a pre-scattered member's grid holds its next sums; the fused step needs no stress cache
            if (s.grid_dirty && (!s.count_on_clear || s.vstar_dirty || (b.fold && s.force_dirty))) run_clear(s);
            ensure_stress(s);
        }
        if (b.synced_version[i] != s.version) {
            launch_keys(s.S, s.P, b.interp, b.kin + b.tab_host[i].pstart, b.st, b.tab_host[i].key_offset);
            b.synced_version[i] = s.version;
            b.snap_ok = false;  // keys the last step's moves do not describe
        }
    }
    if (upload) {
        drop_graph(b);
        HIP_TRY(hipMemcpyAsync(b.tab, b.tab_host.data(), b.members.size() * sizeof(SceneEntry), hipMemcpyHostToDevice,
                               b.st));
        HIP_TRY(hipStreamSynchronize(b.st));
    }
    return MPM_OK;
}

// fused: the fused transfer; scattered: every member's grid already holds
// this step's P2G (else the step clears and scatters the others first);
// full_sort: the step is known to sort in full (the snapshot was just
// invalidated), else the radix launches use the light plan of a sort that
// mostly returns at once.
// overlap: the grid clears run on the second stream (eager steps; a
// captured step keeps them on the batch stream).
void enqueue_step(Batch& b, bool fused, bool scattered, bool full_sort, bool overlap) {
    const int ns = (int)b.members.size();
    const bool forked = overlap && b.aux && hipEventRecord(b.fork, b.st) == hipSuccess &&
                        hipStreamWaitEvent(b.aux, b.fork, 0) == hipSuccess;
    hipStream_t cst = forked ? b.aux : b.st;
    if (!fused) {
        launch_clear_batch(b.tab, b.block_start, ns, b.total_blocks, b.fold ? GF_FOLDED : GF_PRIMARY, cst);
    } else {
        if (!scattered) launch_clear_batch(b.tab, b.block_start, ns, b.total_blocks, GF_FOLDED, cst, 1);
        launch_clear_batch(b.tab, b.block_start, ns, b.total_blocks, GF_FOLDED, cst, 2);
    }
    bool joined = !forked;
    if (forked && hipEventRecord(b.join, b.aux) != hipSuccess) {
        (void)hipStreamSynchronize(b.aux);
        joined = true;
    }
    // the sort: the update of the last step's order or the full radix sort,
    // as the plan words decide on the device (a captured step holds both)
    uint32_t* const plan = b.incr ? b.plan : nullptr;
    if (plan) launch_inc_prepare(plan, b.moves, b.snap_k, b.snap_p, b.ntot, b.st);
    const int sel = radix_sort_pairs(b.kin, b.iota, b.ka, b.va, b.kb, b.vb, b.ntot, b.key_bits, b.hist, b.scan_temp,
                                     b.st, ProfileHook{}, plan, plan && !full_sort);
    b.I.sorted_keys = sel ? b.kb : b.ka;
    b.I.perm = sel ? b.vb : b.va;
    if (plan) launch_inc_merge(plan, b.moves, b.snap_k, b.snap_p, b.ntot, sel ? b.kb : b.ka, sel ? b.vb : b.va, b.st);
    // one chunk of members into this step's order: their slots become their
    // sorted positions, which keeps the (key, slot) order of every cell (a
    // member's cells hold only its particles), so the snapshot the transfer
    // takes afterwards stays sorted
    if (b.reorder_interval > 0 && b.nchunks > 0)
        launch_reorder_chunk(b.tab, b.slot_scene, sel ? b.vb : b.va, b.chunk_bounds, b.chunk_rr, b.nchunks,
                             b.chunk_max, b.chunk_scratch, b.apic != 0, b.st);
    launch_items(b.I.sorted_keys, b.ntot, b.flags, b.scan_temp, b.items, b.counters, b.st, (uint32_t)b.tune.item_max,
                 b.key_offsets, ns, b.item_scene, plan);
    const SceneDev S0{};
    const ParticlesDev P0{};
    const int wgs = effective_wgs(b);
    if (!joined && hipStreamWaitEvent(b.st, b.join, 0) != hipSuccess) (void)hipStreamSynchronize(b.aux);
    if (fused) {
        // P2G only for members whose grid is not pre-scattered (the kernel
        // skips the others' items), then G2P fused with the next P2G
        if (!scattered) launch_p2g(S0, P0, b.I, b.interp, b.apic, 2, wgs, b.st, b.tab);
        launch_g2p2g_batch(b.tab, b.I, wgs, b.kin, b.st);
        launch_swap_grids(b.tab, ns, b.st);
        return;
    }
    launch_p2g(S0, P0, b.I, b.interp, b.apic, b.fold ? 2 : 0, wgs, b.st, b.tab);
    launch_g2p(S0, P0, b.I, b.interp, b.apic, b.fold ? 3 : 0, 1, 0, wgs, b.st, b.kin, b.tab, b.tune.g2p_wg);
}

// Host-side bookkeeping of the members after n batch steps.
void after_steps(Batch& b, int n, bool fused) {
    if (fused) b.parity ^= (n & 1);
    for (size_t i = 0; i < b.members.size(); ++i) {
        Solver& s = *b.members[i];
        if (fused) {
            s.bprescat = true;
            b.tab_host[i].prescat = 1u;
            s.stress_valid = false;  // the fused kernel uses the new stress at once
        }
        s.steps += n;
        s.version += (uint64_t)n;
        b.synced_version[i] = s.version;  // the batch G2P wrote the keys of the new positions
        s.grid_dirty = true;
        s.count_on_clear = true;
        if (!b.fold) s.force_dirty = true;
        s.phase = b.fold ? PH_FOLDED : PH_RAW;
        s.sort_valid = false;
        s.keys_valid = false;  // the member's own sort input is stale
    }
    b.steps += n;
}

hipEvent_t take_event(Batch& b) {
    if (!b.event_pool.empty()) {
        hipEvent_t e = b.event_pool.back();
        b.event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void collect_times(Batch& b) {
    if (b.pending.empty()) return;
    (void)hipStreamSynchronize(b.st);
    for (auto& pr : b.pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) b.step_ms.push_back(ms);
        b.event_pool.push_back(pr.first);
        b.event_pool.push_back(pr.second);
    }
    b.pending.clear();
}

// Runs n steps eagerly (graph = false) or as replays of a captured step.
// Sort the global keys, permute every member's store into its range of
// the sorted order and make the sorted keys the next sort's input.
int ensure_reorder_scratch(Batch& b) {
    if (b.reorder_cap >= b.ntot) return MPM_OK;
    (void)hipStreamSynchronize(b.st);
    bfree(b, *(void**)&b.reorder_scratch);
    b.reorder_cap = 0;
    int rc = balloc(b, (void**)&b.reorder_scratch, (size_t)b.cap * 9 * 4);
    if (rc) return rc;
    b.reorder_cap = b.cap;
    return MPM_OK;
}

int reorder_members(Batch& b) {
    b.last_reorder = b.steps;
    if (b.ntot <= 0) return MPM_OK;
    int rc0 = ensure_reorder_scratch(b);
    if (rc0) return rc0;
    const int sel = radix_sort_pairs(b.kin, b.iota, b.ka, b.va, b.kb, b.vb, b.ntot, b.key_bits, b.hist, b.scan_temp,
                                     b.st, ProfileHook{});
    const uint32_t* sorted = sel ? b.kb : b.ka;
    const uint32_t* perm = sel ? b.vb : b.va;
    // flags (ntot words, free between steps) holds each slot's scene
    launch_reorder_store(ParticlesDev{}, b.tab, (int)b.members.size(), perm, b.ntot, b.reorder_scratch, b.apic != 0,
                         b.st, b.flags);
    HIP_TRY(hipMemcpyAsync(b.kin, sorted, (size_t)b.ntot * 4, hipMemcpyDeviceToDevice, b.st));
    b.snap_ok = false;  // slots renumbered
    for (Solver* m : b.members) {
        m->keys_valid = false;  // the member's own sort input is in the old slot order
        m->sort_valid = false;
    }
    return MPM_OK;
}

int capture_step(Batch& b, bool fused) {
    drop_graph(b);
    HIP_TRY(hipStreamBeginCapture(b.st, hipStreamCaptureModeThreadLocal));
    enqueue_step(b, fused, fused, false, false);
    hipError_t e = hipStreamEndCapture(b.st, &b.graph);
    if (e != hipSuccess) return fail(MPM_ERR_HIP, std::string("batch graph capture: ") + hipGetErrorString(e));
    HIP_TRY(hipGraphInstantiate(&b.gexec, b.graph, nullptr, nullptr, 0));
    b.graph_tune = b.tune;
    return MPM_OK;
}

int run_steps(Batch& b, int n, bool graph) {
    int rc = sync_members(b);
    if (rc) return rc;
    const bool fused = fuses(b);
    if (b.members.empty() || b.ntot == 0 || n <= 0) {
        if (!b.members.empty() && n > 0) after_steps(b, n, false);
        return MPM_OK;
    }
    // A fused step whose members are not all pre-scattered also scatters
    // the others' P2G first; it runs eagerly, and the steps after it (and a
    // captured graph) start from pre-scattered grids.
    bool scattered = fused;
    for (const Solver* m : b.members) scattered = scattered && m->bprescat;
    auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < n; ++i) {
        const bool first_scatter = fused && !scattered;
        if (graph && !first_scatter && (!b.gexec || !(b.graph_tune == b.tune))) {
            if ((rc = capture_step(b, fused))) return rc;
        }
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (b.timing) {
            e0 = take_event(b);
            e1 = take_event(b);
            if (e0 && e1) HIP_TRY(hipEventRecord(e0, b.st));
        }
        const bool full_sort = !b.incr || !b.snap_ok;
        if (b.incr && !b.snap_ok) {  // the full sort this time (outside a captured step)
            launch_inc_set(b.plan, 0u, b.st);
            b.snap_ok = true;
        }
        if (graph && !first_scatter) HIP_TRY(hipGraphLaunch(b.gexec, b.st));
        else enqueue_step(b, fused, scattered, full_sort, true);
        if (e0 && e1) {
            HIP_TRY(hipEventRecord(e1, b.st));
            b.pending.push_back({e0, e1});
        }
        scattered = fused;
    }
    auto t1 = std::chrono::steady_clock::now();
    b.last_step_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    after_steps(b, n, fused);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(MPM_ERR_HIP, std::string("batch step: ") + hipGetErrorString(e));
    return MPM_OK;
}

int detach(Batch& b, Solver& s) {
    auto it = std::find(b.members.begin(), b.members.end(), &s);
    if (it == b.members.end()) return fail(MPM_ERR_NOT_FOUND, "scene is not a member of this batch");
    HIP_TRY(hipStreamSynchronize(b.st));
    normalize(b);  // while the member list still matches the table
    batch_unfuse_member(&b, s);
    b.members.erase(it);
    s.st = s.own_st;
    s.own_st = nullptr;
    s.batch = nullptr;
    s.keys_valid = false;
    s.sort_valid = false;
    b.layout_valid = false;
    drop_graph(b);
    return MPM_OK;
}

void batch_destroy(Batch* b) {
    if (!b) return;
    if (b->st) (void)hipStreamSynchronize(b->st);
    while (!b->members.empty()) (void)detach(*b, *b->members.back());
    drop_graph(*b);
    for (auto& pr : b->pending) b->event_pool.push_back(pr.first), b->event_pool.push_back(pr.second);
    for (auto e : b->event_pool) (void)hipEventDestroy(e);
    for (void* p : b->allocs) (void)devcache::release(p);  // the stream is synchronised above
    if (b->aux) {
        (void)hipStreamSynchronize(b->aux);
        (void)hipStreamDestroy(b->aux);
    }
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->join) (void)hipEventDestroy(b->join);
    if (b->st) (void)hipStreamDestroy(b->st);
    delete b;
}

// Called by mpm_destroy for a scene that is still a member.
void batch_forget(Batch* b, Solver* s) {
    if (b && s) (void)detach(*b, *s);
}

// A member leaves the fused state: its grid becomes the one its last step
// read and the pre-scattered sums wait in its grid2 to be cleared, as the
// two-kernel step leaves them; the next batch step scatters its P2G again
// (the changed table entry is uploaded before that step).
void batch_unfuse_member(Batch* b, Solver& s) {
    if (!b || !s.bprescat) return;
    normalize(*b);
    std::swap(s.S.grid, s.grid2);
    std::swap(s.S.touched, s.touched2);
    s.grid2_dirty = true;
    s.grid2_count = false;
    s.bprescat = false;
    s.stress_valid = false;
    drop_graph(*b);
}

void batch_invalidate(Batch* b) {
    if (!b) return;
    b->layout_valid = false;
    drop_graph(*b);
}

int check_tune(const mpm_tune* t) {
    if (!t) return fail(MPM_ERR_INVALID, "tune is NULL");
    if (t->item_max < 64 || t->item_max > 65536 || (t->item_max % 64) != 0)
        return fail(MPM_ERR_INVALID, "item_max must be a multiple of 64 in [64, 65536]");
    if (t->g2p_threads != 64 && t->g2p_threads != 128 && t->g2p_threads != 256)
        return fail(MPM_ERR_INVALID, "g2p_threads must be 64, 128 or 256");
    if (t->workgroups < 0 || t->workgroups > (1 << 20)) return fail(MPM_ERR_INVALID, "workgroups out of range");
    if (t->fuse_transfer != 0 && t->fuse_transfer != 1) return fail(MPM_ERR_INVALID, "fuse_transfer must be 0 or 1");
    return MPM_OK;
}

void to_c(const Batch& b, const Tune& t, mpm_tune* out, double ms) {
    out->item_max = t.item_max;
    out->g2p_threads = t.g2p_wg;
    out->workgroups = t.wgs;
    out->bucket = size_bucket(b.ntot);
    out->step_ms = ms;
    out->fuse_transfer = t.fuse;
}

Batch* as_batch(void* h) { return static_cast<Batch*>(h); }

}  // namespace host
}  // namespace mpm

using namespace mpm;
using namespace mpm::host;

#define GET_BATCH(var, h)          \
    Batch* var = as_batch(h);      \
    if (!var) return fail(MPM_ERR_INVALID, "batch handle is NULL")

extern "C" {

MPM_API int mpm_batch_create(int32_t interp, int32_t transfer, void** out_batch) {
    if (!out_batch) return fail(MPM_ERR_INVALID, "out_batch is NULL");
    *out_batch = nullptr;
    if (interp < 0 || interp > 2) return fail(MPM_ERR_INVALID, "unknown interpolation type");
    if (transfer < 0 || transfer > 1) return fail(MPM_ERR_INVALID, "unknown transfer scheme");
    Batch* b = new (std::nothrow) Batch();
    if (!b) return fail(MPM_ERR_OOM, "host allocation failed");
    b->interp = interp;
    b->apic = transfer == MPM_TRANSFER_APIC;
    hipError_t e = hipGetDevice(&b->device);
    hipDeviceProp_t prop;
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, b->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->aux, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->join, hipEventDisableTiming);
    if (e == hipSuccess) b->cus = prop.multiProcessorCount;
    int rc = MPM_OK;
    if (e != hipSuccess) rc = fail(MPM_ERR_HIP, std::string("batch create: ") + hipGetErrorString(e));
    if (!rc) rc = balloc(*b, (void**)&b->counters, 64);
    if (!rc && hipMemsetAsync(b->counters, 0, 64, b->st) != hipSuccess) rc = fail(MPM_ERR_HIP, "batch create: memset");
    if (!rc) rc = relayout(*b);
    if (rc) {
        std::string keep = g_err;
        batch_destroy(b);
        g_err = keep;
        return rc;
    }
    *out_batch = b;
    return MPM_OK;
}

MPM_API void mpm_batch_destroy(void* batch) { batch_destroy(as_batch(batch)); }

MPM_API int mpm_batch_add(void* batch, void* scene) {
    GET_BATCH(b, batch);
    Solver* s = static_cast<Solver*>(scene);
    if (!s) return fail(MPM_ERR_INVALID, "scene handle is NULL");
    if (s->batch) return fail(MPM_ERR_STATE, "scene is already a member of a batch");
    if (s->interp != b->interp || s->apic != b->apic)
        return fail(MPM_ERR_INVALID, "scene interpolation/transfer differs from the batch's");
    uint64_t keys = (uint64_t)s->S.escape_key + 64u;
    for (const Solver* m : b->members) keys += (uint64_t)m->S.escape_key + 64u;
    if (keys > (1ull << 32))
        return fail(MPM_ERR_UNSUPPORTED, "batch key space exhausted: the members' grids would total more than 2^32 "
                                         "cell keys");
    // Finish the scene's own work, then move it onto the batch stream.
    unfuse(*s);
    if (s->grid_dirty && (!s->count_on_clear || s->vstar_dirty)) run_clear(*s);
    ensure_stress(*s);
    HIP_TRY(hipStreamSynchronize(s->st));
    HIP_TRY(hipStreamSynchronize(b->st));
    normalize(*b);  // before the member list changes: the new scene's pointers are its own
    s->own_st = s->st;
    s->st = b->st;
    s->batch = b;
    b->members.push_back(s);
    b->layout_valid = false;
    drop_graph(*b);
    return finish(*s);
}

MPM_API int mpm_batch_remove(void* batch, void* scene) {
    GET_BATCH(b, batch);
    Solver* s = static_cast<Solver*>(scene);
    if (!s) return fail(MPM_ERR_INVALID, "scene handle is NULL");
    return detach(*b, *s);
}

MPM_API int mpm_batch_info(void* batch, mpm_batch_stats* out) {
    GET_BATCH(b, batch);
    if (!out) return fail(MPM_ERR_INVALID, "out is NULL");
    int rc = sync_members(*b);
    if (rc) return rc;
    std::memset(out, 0, sizeof(*out));
    uint32_t items = 0;
    if (b->ntot > 0 && b->steps > 0) {
        HIP_TRY(hipMemcpyAsync(&items, b->counters, 4, hipMemcpyDeviceToHost, b->st));
    }
    HIP_TRY(hipStreamSynchronize(b->st));
    out->scenes = (int64_t)b->members.size();
    out->particles = b->ntot;
    out->steps = b->steps;
    out->storage_blocks = b->total_blocks;
    out->key_bits = b->key_bits;
    out->work_items = items;
    out->device_bytes = b->device_bytes;
    out->last_step_ms = b->last_step_ms;
    to_c(*b, b->tune, &out->tune, 0.0);
    if (b->plan) {
        uint32_t inc = 0;
        HIP_TRY(hipMemcpy(&inc, b->plan + kPlanIncSteps, 4, hipMemcpyDeviceToHost));
        out->incremental_sorts = inc;
    }
    return MPM_OK;
}

MPM_API int mpm_batch_read_sort(void* batch, uint32_t* keys, uint32_t* slots, int64_t count) {
    GET_BATCH(b, batch);
    if (count != b->ntot) return fail(MPM_ERR_INVALID, "count must equal the batch's particle count");
    if (b->steps == 0 || !b->I.sorted_keys) return fail(MPM_ERR_STATE, "the batch has not stepped");
    if (keys) HIP_TRY(hipMemcpyAsync(keys, b->I.sorted_keys, (size_t)count * 4, hipMemcpyDeviceToHost, b->st));
    if (slots) HIP_TRY(hipMemcpyAsync(slots, b->I.perm, (size_t)count * 4, hipMemcpyDeviceToHost, b->st));
    HIP_TRY(hipStreamSynchronize(b->st));
    return MPM_OK;
}

MPM_API int mpm_batch_set_incremental_sort(void* batch, int enable) {
    GET_BATCH(b, batch);
    HIP_TRY(hipStreamSynchronize(b->st));
    b->incr = enable != 0;
    b->layout_valid = false;  // the items' view of the snapshot and list is set by the layout
    drop_graph(*b);
    return MPM_OK;
}

MPM_API int mpm_batch_step(void* batch, int n) {
    GET_BATCH(b, batch);
    if (n < 0) return fail(MPM_ERR_INVALID, "n < 0");
    return run_steps(*b, n, false);
}

MPM_API int mpm_batch_set_reorder_interval(void* batch, int steps) {
    GET_BATCH(b, batch);
    if (steps < 0) return fail(MPM_ERR_INVALID, "steps < 0");
    if (b->reorder_interval != steps) {
        b->reorder_interval = steps;
        b->layout_valid = false;  // new chunks; the captured step changes
        drop_graph(*b);
    }
    return MPM_OK;
}

MPM_API int mpm_batch_step_graph(void* batch, int n) {
    GET_BATCH(b, batch);
    if (n < 0) return fail(MPM_ERR_INVALID, "n < 0");
    return run_steps(*b, n, true);
}

MPM_API int mpm_batch_synchronize(void* batch) {
    GET_BATCH(b, batch);
    HIP_TRY(hipStreamSynchronize(b->st));
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(MPM_ERR_HIP, std::string("batch: ") + hipGetErrorString(e));
    return MPM_OK;
}

MPM_API int mpm_batch_set_tune(void* batch, const mpm_tune* t) {
    GET_BATCH(b, batch);
    if (!t) {  // back to the tuning table
        b->tune_fixed = false;
        b->layout_valid = false;
        return MPM_OK;
    }
    int rc = check_tune(t);
    if (rc) return rc;
    b->tune.item_max = t->item_max;
    b->tune.g2p_wg = t->g2p_threads;
    b->tune.wgs = t->workgroups;
    b->tune.fuse = t->fuse_transfer;
    b->tune_fixed = true;
    return MPM_OK;
}

MPM_API int mpm_batch_autotune(void* batch, int trial_steps, mpm_tune* out) {
    GET_BATCH(b, batch);
    if (trial_steps < 1) return fail(MPM_ERR_INVALID, "trial_steps < 1");
    int rc = sync_members(*b);
    if (rc) return rc;
    if (b->ntot == 0) return fail(MPM_ERR_STATE, "batch is empty");
    // Coordinate search from the defaults: particles per work item, then
    // G2P workgroup size, then persistent workgroups (one resident round,
    // or a fixed 8 or 16 per CU).  Every trial step is a real step of every
    // scene (the parameters change the work split, not the arithmetic).
    const int item_opts[] = {128, 256, 512, 1024};
    const int wg_opts[] = {64, 128, 256};
    const int wgs_opts[] = {0, b->cus * 8, b->cus * 16};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    const bool fixed = b->tune_fixed;
    b->tune_fixed = true;
    // every trial starts from a store in cell order (the order decays within
    // tens of steps, which would favour the first trials), with no reorder
    // inside the timed steps
    const int interval = b->reorder_interval;
    b->reorder_interval = 0;
    auto measure = [&](const Tune& t, double& ms) -> int {
        b->tune = t;
        int r = reorder_members(*b);
        if (r) return r;
        r = run_steps(*b, 1, false);  // warm-up (code objects, caches)
        if (r) return r;
        HIP_TRY(hipEventRecord(e0, b->st));
        r = run_steps(*b, trial_steps, false);
        if (r) return r;
        HIP_TRY(hipEventRecord(e1, b->st));
        HIP_TRY(hipEventSynchronize(e1));
        float t_ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&t_ms, e0, e1));
        ms = (double)t_ms / trial_steps;
        return MPM_OK;
    };
    Tune best;
    double best_ms = 0.0;
    rc = measure(best, best_ms);
    // the last stage: the fused transfer, where the batch allows it (it wins
    // when few particles change cell per step: each one that does scatters
    // its 27 nodes with global atomics)
    for (int stage = 0; stage < 4 && !rc; ++stage) {
        if (stage == 3 && !can_fuse(*b)) break;
        for (int k = 0; k < (stage == 0 ? 4 : stage == 3 ? 2 : 3) && !rc; ++k) {
            Tune t = best;
            if (stage == 0) t.item_max = item_opts[k];
            if (stage == 1) t.g2p_wg = wg_opts[k];
            if (stage == 2) t.wgs = wgs_opts[k];
            if (stage == 3) t.fuse = k;
            if (t == best) continue;
            double ms = 0.0;
            rc = measure(t, ms);
            if (!rc && ms < best_ms) { best = t; best_ms = ms; }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->tune_fixed = fixed;
    b->reorder_interval = interval;
    if (rc) return rc;
    b->tune = best;
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        g_tune[{b->interp, b->apic, size_bucket(b->ntot)}] = TuneEntry{best, best_ms};
    }
    if (out) to_c(*b, best, out, best_ms);
    return MPM_OK;
}

MPM_API int mpm_batch_timing(void* batch, int enable) {
    GET_BATCH(b, batch);
    collect_times(*b);
    b->timing = enable != 0;
    return MPM_OK;
}

MPM_API int mpm_batch_read_times(void* batch, float* ms, int max, int* count) {
    GET_BATCH(b, batch);
    collect_times(*b);
    const int k = (int)b->step_ms.size();
    if (count) *count = k;
    if (!ms) return MPM_OK;  // count only
    for (int i = 0; i < std::min(k, max); ++i) ms[i] = b->step_ms[(size_t)i];
    b->step_ms.clear();
    return MPM_OK;
}

MPM_API int mpm_batch_health(void* batch, mpm_scene_health* out, int n) {
    GET_BATCH(b, batch);
    if (!out && n > 0) return fail(MPM_ERR_INVALID, "out is NULL");
    int rc = sync_members(*b);  // keys of members changed outside the batch
    if (rc) return rc;
    const int ns = (int)b->members.size();
    if (n != ns) return fail(MPM_ERR_INVALID, "n must equal the number of members");
    if (ns == 0) return MPM_OK;
    if (b->health_cap < ns) {
        HIP_TRY(hipStreamSynchronize(b->st));
        bfree(*b, *(void**)&b->health);
        b->health_cap = 0;
        if ((rc = balloc(*b, (void**)&b->health, (size_t)b->tab_cap * 3 * 4))) return rc;
        b->health_cap = b->tab_cap;
    }
    launch_fill_u32(b->health, 0u, (int64_t)ns * 3, b->st);
    launch_batch_health(b->tab, b->key_offsets, ns, b->ntot, b->kin, b->health, b->st);
    std::vector<uint32_t> h((size_t)ns * 3);
    HIP_TRY(hipMemcpyAsync(h.data(), b->health, h.size() * 4, hipMemcpyDeviceToHost, b->st));
    HIP_TRY(hipStreamSynchronize(b->st));
    for (int i = 0; i < ns; ++i) {
        const Solver& s = *b->members[(size_t)i];
        float v2;
        std::memcpy(&v2, &h[(size_t)i * 3], 4);
        out[i].max_speed = std::sqrt((double)v2);
        out[i].cfl = out[i].max_speed * (double)s.S.dt / (double)s.S.h;
        out[i].nonfinite = h[(size_t)i * 3 + 1];
        out[i].escaped = h[(size_t)i * 3 + 2];
    }
    return MPM_OK;
}

MPM_API int mpm_tune_lookup(int32_t interp, int32_t transfer, int64_t particles, mpm_tune* out) {
    if (!out) return fail(MPM_ERR_INVALID, "out is NULL");
    std::lock_guard<std::mutex> g(g_tune_mu);
    auto it = g_tune.find({interp, transfer == MPM_TRANSFER_APIC ? 1 : 0, size_bucket(particles)});
    if (it == g_tune.end()) return fail(MPM_ERR_NOT_FOUND, "no tuning entry for this size bucket");
    out->item_max = it->second.t.item_max;
    out->g2p_threads = it->second.t.g2p_wg;
    out->workgroups = it->second.t.wgs;
    out->bucket = size_bucket(particles);
    out->step_ms = it->second.ms;
    return MPM_OK;
}

}  // extern "C"
