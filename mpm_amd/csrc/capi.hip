// capi.hip -- host side of the C-ABI (include/mpm_hip.h).
//
// A handle owns the SoA particle store, the block-major grid, the sort
// buffers and one HIP stream.  mpm_step runs the fused device path:
//   clear touched grid blocks -> cell keys -> stable radix sort -> work
//   items -> tiled P2G -> tiled G2P with the grid update fused into its
//   tile load (+ F update, particle collision, advection).
// The reference's sub-step methods run unfused kernels that materialise the
// grid exactly as the reference leaves it, so the two paths can be compared.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <tuple>
#include <vector>

#include "devcache.h"
#include "devmath.h"
#include "launch.h"
#include "solver_impl.h"

using namespace mpm;

// Threads per workgroup of the fused G2P of one scene: 128 (two waves per
// item) runs the linear and quadratic APIC G2P at 4 waves per SIMD.
#ifndef MPM_SOLVER_G2P_WG
#define MPM_SOLVER_G2P_WG 128
#endif

namespace mpm {
namespace host {

// Device memory comes from the block cache (devcache.h), so that building
// and destroying scenes never waits for other streams' work.
int dalloc(Solver& s, void** p, size_t bytes) {
    size_t got = 0;
    hipError_t e = devcache::alloc(p, bytes, &got);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(MPM_ERR_OOM, "hipMalloc(" + std::to_string(bytes) + " B) failed: " + hipGetErrorString(e));
    }
    s.allocs.push_back(*p);
    s.device_bytes += (int64_t)got;
    return MPM_OK;
}

// The solver's stream (the batch's while it is a member) is synchronised
// first: the block may be handed to another stream at once.
void dfree(Solver& s, void* p) {
    if (!p) return;
    for (size_t i = 0; i < s.allocs.size(); ++i)
        if (s.allocs[i] == p) { s.allocs.erase(s.allocs.begin() + (long)i); break; }
    if (s.st) (void)hipStreamSynchronize(s.st);
    s.device_bytes -= (int64_t)devcache::release(p);
}

// ---- profiling hook (hipEvent pairs on the solver stream)
hipEvent_t take_event(Solver& s) {
    if (!s.event_pool.empty()) {
        hipEvent_t e = s.event_pool.back();
        s.event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}
// Level 2 times the item kernels and the store reorder only.
bool prof_wants(const Solver& s, const char* name) {
    return s.prof != 2 || std::strcmp(name, "p2g") == 0 || std::strcmp(name, "g2p") == 0 ||
           std::strcmp(name, "g2p2g") == 0 || std::strcmp(name, "reorder") == 0;
}
void prof_begin(void* ctx, const char* name) {
    Solver& s = *static_cast<Solver*>(ctx);
    if (!prof_wants(s, name)) return;
    Pending p;
    p.name = name;
    p.a = take_event(s);
    p.b = take_event(s);
    if (!p.a || !p.b) return;
    (void)hipEventRecord(p.a, s.st);
    s.pending.push_back(p);
}
void prof_end(void* ctx, const char* name) {
    Solver& s = *static_cast<Solver*>(ctx);
    if (!prof_wants(s, name)) return;  // never close an older pair of a skipped name
    for (size_t i = s.pending.size(); i-- > 0;)
        if (s.pending[i].name == name) {
            (void)hipEventRecord(s.pending[i].b, s.st);
            return;
        }
}
void prof_collect(Solver& s) {
    if (s.pending.empty()) return;
    (void)hipStreamSynchronize(s.st);
    for (auto& p : s.pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            auto it = s.prof_tot.find(p.name);
            if (it == s.prof_tot.end()) {
                s.prof_order.push_back(p.name);
                s.prof_tot[p.name] = {1, (double)ms};
            } else {
                it->second.first += 1;
                it->second.second += ms;
            }
        }
        s.event_pool.push_back(p.a);
        s.event_pool.push_back(p.b);
    }
    s.pending.clear();
}
ProfileHook hook(Solver& s) {
    ProfileHook h;
    if (s.prof) {
        h.ctx = &s;
        h.fn_begin = prof_begin;
        h.fn_end = prof_end;
    }
    return h;
}

// ---- validation
int check_desc(const mpm_desc* d) {
    if (!d) return fail(MPM_ERR_INVALID, "desc is NULL");
    if (d->interp < 0 || d->interp > 2) return fail(MPM_ERR_INVALID, "unknown interpolation type");
    if (d->transfer < 0 || d->transfer > 1) return fail(MPM_ERR_INVALID, "unknown transfer scheme");
    if (d->transfer == MPM_TRANSFER_APIC && d->interp == MPM_INTERP_LINEAR)
        return fail(MPM_ERR_UNSUPPORTED,
                    "APIC with linear interpolation has a singular D (Dinv non-finite); rejected");
    for (int i = 0; i < 3; ++i) {
        if (d->resolution[i] <= 0 || d->resolution[i] > 4096) return fail(MPM_ERR_INVALID, "resolution out of range");
        if (!std::isfinite(d->origin[i])) return fail(MPM_ERR_INVALID, "origin not finite");
    }
    if (!(d->stride > 0.0f) || !std::isfinite(d->stride)) return fail(MPM_ERR_INVALID, "stride must be > 0");
    if (!std::isfinite(d->dt) || !std::isfinite(d->gravity) || !std::isfinite(d->friction) ||
        !std::isfinite(d->blend))
        return fail(MPM_ERR_INVALID, "non-finite dt/gravity/friction/blend");
    uint64_t kb = 1;
    for (int i = 0; i < 3; ++i) kb *= (uint64_t)(((d->resolution[i] + 3) >> 2) + 1);
    if (kb * 64ull >= (1ull << 32)) return fail(MPM_ERR_INVALID, "grid too large for 32-bit cell keys");
    for (int i = 0; i < 4; ++i)
        if (d->reserved[i] != 0) return fail(MPM_ERR_INVALID, "reserved fields must be zero");
    return MPM_OK;
}

// Eigen's 3x3 inverse of D = k h^2 I (cofactor / determinant), float.
float apic_dinv(const mpm_desc& d) {
    volatile float k = 0.0f;
    if (d.interp == MPM_INTERP_CUBIC_BSPLINE) k = (float)(1 / 3.0);
    if (d.interp == MPM_INTERP_QUADRATIC_BSPLINE) k = (float)(1 / 4.0);
    volatile float kh = k * d.stride;
    volatile float D = kh * d.stride;
    volatile float c = D * D;
    volatile float det = c * D;
    volatile float id = 1.0f / det;
    volatile float r = c * id;
    return r;
}

void fill_scene(Solver& s) {
    const mpm_desc& d = s.desc;
    SceneDev& S = s.S;
    S.ox = d.origin[0]; S.oy = d.origin[1]; S.oz = d.origin[2];
    volatile float hx = d.stride * (float)d.resolution[0];
    volatile float hy = d.stride * (float)d.resolution[1];
    volatile float hz = d.stride * (float)d.resolution[2];
    S.tx = d.origin[0] + hx; S.ty = d.origin[1] + hy; S.tz = d.origin[2] + hz;
    S.h = d.stride;
    S.inv_h = 1.0f / d.stride;
    S.dt = d.dt;
    S.gravity = d.gravity;
    S.friction = d.friction;
    S.blend = d.blend;
    S.dinv = s.apic ? apic_dinv(d) : 0.0f;
    volatile float h3 = d.stride * d.stride;
    h3 = h3 * d.stride;
    S.inv_cell_vol = (float)(1.0 / (double)h3);
    S.rx = d.resolution[0]; S.ry = d.resolution[1]; S.rz = d.resolution[2];
    S.nbx = (S.rx + 3) / 4; S.nby = (S.ry + 3) / 4; S.nbz = (S.rz + 3) / 4;
    S.kbx = ((S.rx + 3) >> 2) + 1; S.kby = ((S.ry + 3) >> 2) + 1; S.kbz = ((S.rz + 3) >> 2) + 1;
    S.escape_key = (uint32_t)S.kbx * (uint32_t)S.kby * (uint32_t)S.kbz * 64u;
    S.particle_collision = d.particle_collision;
    s.nsb = (int64_t)S.nbx * S.nby * S.nbz;
    S.plane = s.nsb * 64;
    int bits = 1;
    while (bits < 32 && ((uint64_t)S.escape_key >> bits) != 0) ++bits;
    s.key_bits = bits;
}

// Host arrays in the C-ABI layout (original order).
struct HostState {
    std::vector<float> volume0, mass, pos, vel, Fe, Fp, mu0, lam0, B;
    std::vector<uint16_t> mat;
};

int dedupe_materials(Solver& s, int64_t n, const float* xi, const float* thc, const float* ths, const int32_t* model,
                     std::vector<uint16_t>& mat_ids) {
    std::map<std::tuple<uint32_t, uint32_t, uint32_t, int32_t>, uint16_t> table;
    for (size_t i = 0; i < s.mats.size(); ++i) {
        uint32_t a, b, c;
        std::memcpy(&a, &s.mats[i].hardening, 4);
        std::memcpy(&b, &s.mats[i].crit_c, 4);
        std::memcpy(&c, &s.mats[i].crit_s, 4);
        table[{a, b, c, s.mats[i].model}] = (uint16_t)i;
    }
    mat_ids.resize((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        if (model[i] != MPM_MODEL_FIXED_COROTATED && model[i] != MPM_MODEL_NEO_HOOKEAN)
            return fail(MPM_ERR_INVALID, "particle " + std::to_string(i) + ": unknown constitutive model");
        uint32_t a, b, c;
        std::memcpy(&a, &xi[i], 4);
        std::memcpy(&b, &thc[i], 4);
        std::memcpy(&c, &ths[i], 4);
        auto key = std::make_tuple(a, b, c, model[i]);
        auto it = table.find(key);
        if (it == table.end()) {
            if (s.mats.size() >= MPM_MAX_MATERIALS)
                return fail(MPM_ERR_UNSUPPORTED, "more than 65536 distinct (hardening, theta_c, theta_s, model)");
            MaterialDev m{xi[i], thc[i], ths[i], model[i]};
            s.mats.push_back(m);
            uint16_t id = (uint16_t)(s.mats.size() - 1);
            table[key] = id;
            mat_ids[(size_t)i] = id;
        } else {
            mat_ids[(size_t)i] = it->second;
        }
    }
    return MPM_OK;
}

// The captured step holds the particle descriptor (with the material table's
// address) by value: a new table invalidates it.
static void drop_graph(Solver& s) {
    if (s.gexec) { (void)hipGraphExecDestroy(s.gexec); s.gexec = nullptr; }
    if (s.graph) { (void)hipGraphDestroy(s.graph); s.graph = nullptr; }
}

int upload_materials(Solver& s) {
    drop_graph(s);
    unfuse(s);  // the pre-scattered forces used the old materials
    if (s.mats_dev) dfree(s, s.mats_dev);
    s.mats_dev = nullptr;
    size_t bytes = std::max<size_t>(1, s.mats.size()) * sizeof(MaterialDev);
    int rc = dalloc(s, (void**)&s.mats_dev, bytes);
    if (rc) return rc;
    if (!s.mats.empty()) HIP_TRY(hipMemcpyAsync(s.mats_dev, s.mats.data(), s.mats.size() * sizeof(MaterialDev),
                                                hipMemcpyHostToDevice, s.st));
    s.P.mats = (decltype(s.P.mats))s.mats_dev;
    return MPM_OK;
}

int alloc_particles(Solver& s, ParticlesDev& P, int64_t n) {
    P.n = n;
    size_t f = (size_t)std::max<int64_t>(n, 1) * sizeof(float);
    int rc;
    if ((rc = dalloc(s, (void**)&P.x, 3 * f))) return rc;
    if ((rc = dalloc(s, (void**)&P.v, 3 * f))) return rc;
    if (s.apic && (rc = dalloc(s, (void**)&P.B, 9 * f))) return rc;
    if (!s.apic) P.B = nullptr;
    if ((rc = dalloc(s, (void**)&P.Fe, 9 * f))) return rc;
    if ((rc = dalloc(s, (void**)&P.Fp, 9 * f))) return rc;
    if ((rc = dalloc(s, (void**)&P.Jp, f))) return rc;
    if ((rc = dalloc(s, (void**)&P.m, f))) return rc;
    if ((rc = dalloc(s, (void**)&P.V0, f))) return rc;
    if ((rc = dalloc(s, (void**)&P.mu0, f))) return rc;
    if ((rc = dalloc(s, (void**)&P.lam0, f))) return rc;
    if ((rc = dalloc(s, (void**)&P.tau, 6 * f))) return rc;
    if ((rc = dalloc(s, (void**)&P.mat, (size_t)std::max<int64_t>(n, 1) * 2))) return rc;
    if ((rc = dalloc(s, (void**)&P.pid, (size_t)std::max<int64_t>(n, 1) * 4))) return rc;
    P.mats = (decltype(P.mats))s.mats_dev;
    return MPM_OK;
}
void free_particles(Solver& s, ParticlesDev& P) {
    dfree(s, P.x); dfree(s, P.v); dfree(s, P.B); dfree(s, P.Fe); dfree(s, P.Fp); dfree(s, P.Jp); dfree(s, P.m);
    dfree(s, P.V0); dfree(s, P.mu0); dfree(s, P.lam0); dfree(s, P.tau); dfree(s, P.mat); dfree(s, P.pid);
}

// Host layout staging on the device for pack/unpack.
struct Staging {
    HostLayoutDev H{};
    std::vector<void*> ptrs;
};
int alloc_staging(Solver& s, Staging& g, int64_t n) {
    size_t f = (size_t)std::max<int64_t>(n, 1) * sizeof(float);
    float** fields[] = {&g.H.volume0, &g.H.mass, &g.H.mu0, &g.H.lam0};
    for (auto p : fields) {
        int rc = dalloc(s, (void**)p, f);
        if (rc) return rc;
        g.ptrs.push_back(*p);
    }
    int rc;
    if ((rc = dalloc(s, (void**)&g.H.pos, 3 * f))) return rc;
    g.ptrs.push_back(g.H.pos);
    if ((rc = dalloc(s, (void**)&g.H.vel, 3 * f))) return rc;
    g.ptrs.push_back(g.H.vel);
    if ((rc = dalloc(s, (void**)&g.H.Fe, 9 * f))) return rc;
    g.ptrs.push_back(g.H.Fe);
    if ((rc = dalloc(s, (void**)&g.H.Fp, 9 * f))) return rc;
    g.ptrs.push_back(g.H.Fp);
    if ((rc = dalloc(s, (void**)&g.H.B, 9 * f))) return rc;
    g.ptrs.push_back(g.H.B);
    if ((rc = dalloc(s, (void**)&g.H.mat, (size_t)std::max<int64_t>(n, 1) * 2))) return rc;
    g.ptrs.push_back(g.H.mat);
    return MPM_OK;
}
void free_staging(Solver& s, Staging& g) {
    for (void* p : g.ptrs) dfree(s, p);
    g.ptrs.clear();
}

// ---- device phases
void run_sort(Solver& s) {
    ProfileHook h = hook(s);
    s.snap_ok = false;  // a sort outside the fused step: the move list no longer applies
    if (s.N <= 0) { s.sort_valid = true; return; }
    if (!s.keys_valid) {
        Scope sc(h, "keys");
        launch_keys(s.S, s.P, s.interp, s.kin, s.st);
        s.keys_valid = true;
    }
    int sel = radix_sort_pairs(s.kin, s.iota, s.ka, s.va, s.kb, s.vb, s.N, s.key_bits, s.hist, s.scan_temp, s.st, h);
    s.I.sorted_keys = sel ? s.kb : s.ka;
    s.I.perm = sel ? s.vb : s.va;
    s.I.start = s.items;
    s.I.blk = s.items + s.N;
    s.I.count = s.counters;
    s.I.n = s.N;
    {
        Scope sc(h, "items");
        launch_items(s.I.sorted_keys, s.N, s.flags, s.scan_temp, s.items, s.counters, s.st);
    }
    s.sort_valid = true;
}

void run_clear(Solver& s) {
    ProfileHook h = hook(s);
    unfuse(s);
    Scope sc(h, "grid_clear");
    launch_clear(s.S, s.nsb, s.vstar_dirty ? GF_ALL : (s.force_dirty ? GF_PRIMARY : GF_FOLDED),
                 s.count_on_clear ? s.acc : nullptr, s.st);
    s.grid_dirty = false;
    s.vstar_dirty = false;
    s.force_dirty = false;
    s.count_on_clear = false;
    s.phase = PH_CLEAR;
}

void ensure_stress(Solver& s) {
    if (s.stress_valid) return;
    ProfileHook h = hook(s);
    Scope sc(h, "stress");
    launch_stress(s.P, s.st);
    s.stress_valid = true;
}

// Every 8 steps the count of incremental sorts is copied into pinned host
// memory on the stream; once the copy has landed (an event query, no wait)
// the share of updates since the previous reading sets few_passes: mostly
// updates (>= 3/4) -> the full sort's kernels, which then mostly return at
// once, use 3 passes instead of 4.  Only the speed depends on it.
static void sample_inc_history(Solver& s) {
    if (s.capturing || !s.inc_host || !s.inc_ev) return;
    if (s.inc_pending && hipEventQuery(s.inc_ev) == hipSuccess) {
        const uint32_t now = *reinterpret_cast<volatile uint32_t*>(s.inc_host);
        if (s.inc_prev_step >= 0 && s.inc_copy_step > s.inc_prev_step)
            s.few_passes = 4.0 * (double)(now - s.inc_prev) >= 3.0 * (double)(s.inc_copy_step - s.inc_prev_step);
        s.inc_prev = now;
        s.inc_prev_step = s.inc_copy_step;
        s.inc_pending = false;
    }
    if (!s.inc_pending && s.steps - s.inc_copy_step >= 8) {
        if (hipMemcpyAsync(s.inc_host, s.plan + kPlanIncSteps, 4, hipMemcpyDeviceToHost, s.st) != hipSuccess ||
            hipEventRecord(s.inc_ev, s.st) != hipSuccess)
            return;
        s.inc_copy_step = s.steps;
        s.inc_pending = true;
    }
}

// The fused step's sort of a large scene: the plan words choose between the
// incremental update of the last step's order (few moved particles) and the
// full radix sort; both write the same buffers, the other's launches return
// at once.  Bit-identical to run_sort.
void run_sort_incremental(Solver& s) {
    ProfileHook h = hook(s);
    sample_inc_history(s);
    if (!s.keys_valid) {
        Scope sc(h, "keys");
        launch_keys(s.S, s.P, s.interp, s.kin, s.st);
        s.keys_valid = true;
        s.snap_ok = false;
    }
    if (!s.snap_ok) launch_inc_set(s.plan, 0u, s.st);  // the full sort this time
    {
        Scope sc(h, "sort_incr");
        launch_inc_prepare(s.plan, s.moves, s.snap_k, s.snap_p, s.N, s.st);
    }
    const int sel = radix_sort_pairs(s.kin, s.iota, s.ka, s.va, s.kb, s.vb, s.N, s.key_bits, s.hist, s.scan_temp,
                                     s.st, h, s.plan, s.few_passes);
    uint32_t* out_k = sel ? s.kb : s.ka;
    uint32_t* out_p = sel ? s.vb : s.va;
    s.I.sorted_keys = out_k;
    s.I.perm = out_p;
    s.I.start = s.items;
    s.I.blk = s.items + s.N;
    s.I.count = s.counters;
    s.I.n = s.N;
    {
        Scope sc(h, "sort_incr");
        launch_inc_merge(s.plan, s.moves, s.snap_k, s.snap_p, s.N, out_k, out_p, s.st);
    }
    {
        Scope sc(h, "items");  // also resets the plan words for this step's G2P
        launch_items(s.I.sorted_keys, s.N, s.flags, s.scan_temp, s.items, s.counters, s.st, 0, nullptr, 0, nullptr,
                     s.plan);
    }
    s.sort_valid = true;
}

// G2P fused with the next step's P2G (k_g2p2g) for the folded APIC step of
// a single scene with the quadratic stencil (APIC with the linear one is
// rejected at creation; the cubic one keeps the two-kernel step: its 7^3
// tiles leave no room for a second set).
static bool use_fused_transfer(const Solver& s) {
    return s.fuse_transfer && fused_folds(s) && s.interp == MPM_INTERP_QUADRATIC_BSPLINE && !s.capturing &&
           !s.batch && s.N > 0;
}

// The second grid buffer of the fused transfer and the stream that clears
// it (allocated on first use).
int ensure_grid2_buffers(Solver& s) {
    if (s.grid2 && s.touched2) return MPM_OK;
    const size_t gbytes = (size_t)s.S.plane * GF_ALL * sizeof(float);
    int rc;
    if (!s.grid2 && (rc = dalloc(s, (void**)&s.grid2, gbytes))) return rc;
    if (!s.touched2 && (rc = dalloc(s, (void**)&s.touched2, (size_t)s.nsb * 4))) return rc;
    HIP_TRY(hipMemsetAsync(s.grid2, 0, gbytes, s.st));
    HIP_TRY(hipMemsetAsync(s.touched2, 0, (size_t)s.nsb * 4, s.st));
    s.grid2_dirty = false;
    return MPM_OK;
}
static int ensure_grid2(Solver& s) {
    if (!s.aux_st) HIP_TRY(hipStreamCreateWithFlags(&s.aux_st, hipStreamNonBlocking));
    if (!s.aux_fork) HIP_TRY(hipEventCreateWithFlags(&s.aux_fork, hipEventDisableTiming));
    if (!s.aux_join) HIP_TRY(hipEventCreateWithFlags(&s.aux_join, hipEventDisableTiming));
    return ensure_grid2_buffers(s);
}

// Clear of the second grid buffer (the grid the last fused step read, its
// active nodes counted, or pre-scattered sums that were dropped).  Unless
// every launch group is being timed (profile level 1) it runs on aux_st,
// overlapping the sort; the step joins it before the fused kernel writes
// the buffer.
// Returns true when a join is due.
static bool clear_grid2(Solver& s) {
    if (!s.grid2_dirty) return false;
    SceneDev S2 = s.S;
    S2.grid = s.grid2;
    S2.touched = s.touched2;
    unsigned long long* cnt = s.grid2_count ? s.acc : nullptr;
    s.grid2_dirty = false;
    if (s.prof == 1) {
        ProfileHook h = hook(s);
        Scope sc(h, "grid_clear");
        launch_clear(S2, s.nsb, GF_FOLDED, cnt, s.st);
        return false;
    }
    if (hipEventRecord(s.aux_fork, s.st) != hipSuccess || hipStreamWaitEvent(s.aux_st, s.aux_fork, 0) != hipSuccess) {
        launch_clear(S2, s.nsb, GF_FOLDED, cnt, s.st);
        return false;
    }
    launch_clear(S2, s.nsb, GF_FOLDED, cnt, s.aux_st);
    if (hipEventRecord(s.aux_join, s.aux_st) != hipSuccess) {
        (void)hipStreamSynchronize(s.aux_st);
        return false;
    }
    return true;
}

void run_fused_step(Solver& s) {
    ProfileHook h = hook(s);
    const bool fuse = use_fused_transfer(s) && ensure_grid2(s) == MPM_OK;
    const bool scattered = fuse && s.prescat;  // the last step already did this step's P2G
    if (!scattered) {
        if (s.grid_dirty) run_clear(s);
        ensure_stress(s);
    }
    const bool join = fuse && clear_grid2(s);
    // a raster or sub-step since the last change of the positions has
    // already sorted them
    if (!s.sort_valid) {
        if (s.incr) run_sort_incremental(s);
        else run_sort(s);
    } else if (s.incr) {
        // sorted elsewhere (a store reorder, a raster): this step's P2G
        // takes the snapshot of that order and its G2P starts a new list
        launch_inc_set(s.plan, 1u, s.st);
    }
    ItemsDev I = s.I;
    if (s.incr) {
        I.snap_k = s.snap_k;
        I.snap_p = s.snap_p;
        I.plan = s.plan;
        I.moves = s.moves_dev;
        I.move_cap = s.move_cap;
    }
    // APIC's G2P reads v* alone, so by default its P2G sums p + dt f into
    // one set of planes (4 instead of 7 per node)
    const bool fold = fused_folds(s);
    if (fuse) {
        if (!scattered) {
            Scope sc(h, "p2g");
            launch_p2g(s.S, s.P, I, s.interp, s.apic, 2, s.item_wgs, s.st);
        }
        if (join && hipStreamWaitEvent(s.st, s.aux_join, 0) != hipSuccess) (void)hipStreamSynchronize(s.aux_st);
        {
            Scope sc(h, "g2p2g");
            launch_g2p2g(s.S, s.P, I, s.S.particle_collision, s.grid2, s.touched2, s.kin, s.st);
        }
        // the grid just scattered is the next step's
        std::swap(s.S.grid, s.grid2);
        std::swap(s.S.touched, s.touched2);
        s.grid2_dirty = true;
        s.grid2_count = true;
        s.prescat = true;
        s.stress_valid = false;  // the fused kernel uses the new stress at once
    } else {
        {
            Scope sc(h, "p2g");
            launch_p2g(s.S, s.P, I, s.interp, s.apic, fold ? 2 : 0, s.item_wgs, s.st);
        }
        {
            Scope sc(h, "g2p");
            launch_g2p(s.S, s.P, I, s.interp, s.apic, fold ? 3 : 0, 1, s.S.particle_collision, s.item_wgs, s.st,
                       s.kin, nullptr, MPM_SOLVER_G2P_WG);
        }
    }
    s.snap_ok = s.incr;
    s.grid_dirty = true;
    s.count_on_clear = true;
    if (!fold) s.force_dirty = true;
    s.phase = fold ? PH_FOLDED : PH_RAW;
    s.sort_valid = false;
    s.keys_valid = true;  // the G2P wrote the keys of the advected positions
    s.steps += 1;
    s.version += 1;
}

void run_initialize(Solver& s) {
    unfuse(s);
    if (s.grid_dirty) run_clear(s);
    run_sort(s);
    launch_p2g(s.S, s.P, s.I, s.interp, s.apic, 1, s.item_wgs, s.st);
    launch_g2p(s.S, s.P, s.I, s.interp, s.apic, 2, 0, 0, s.item_wgs, s.st);
    s.grid_dirty = true;
    s.count_on_clear = false;
    s.phase = PH_MASS;
    s.stress_valid = false;  // volume0 changed
}

// Materialise the grid as the reference holds it after a step (gravity
// applied to force, velocity normalised, v* after collision).
void materialize(Solver& s) {
    if (s.phase != PH_RAW) return;
    launch_grid_op(s.S, s.nsb, 0, s.S.dt, s.st);
    launch_grid_op(s.S, s.nsb, 1, s.S.dt, s.st);
    launch_grid_op(s.S, s.nsb, 2, s.S.dt, s.st);
    s.vstar_dirty = true;
    s.phase = PH_SUBSTEP;
}

int build(Solver& s, HostState& H, int run_init) {
    int rc;
    HIP_TRY(hipGetDevice(&s.device));
    HIP_TRY(hipStreamCreateWithFlags(&s.st, hipStreamNonBlocking));
    s.item_wgs = 0;  // persistent kernels: one resident round (occupancy x CUs)
    fill_scene(s);
    if ((rc = upload_materials(s))) return rc;
    // grid
    size_t gbytes = (size_t)s.S.plane * GF_ALL * sizeof(float);
    if ((rc = dalloc(s, (void**)&s.S.grid, gbytes))) return rc;
    HIP_TRY(hipMemsetAsync(s.S.grid, 0, gbytes, s.st));
    if ((rc = dalloc(s, (void**)&s.S.touched, (size_t)s.nsb * 4))) return rc;
    HIP_TRY(hipMemsetAsync(s.S.touched, 0, (size_t)s.nsb * 4, s.st));
    // sort buffers
    const int64_t n = std::max<int64_t>(s.N, 1);
    if ((rc = dalloc(s, (void**)&s.kin, (size_t)n * 4))) return rc;
    if ((rc = dalloc(s, (void**)&s.iota, (size_t)n * 4))) return rc;
    if ((rc = dalloc(s, (void**)&s.ka, (size_t)n * 4))) return rc;
    if ((rc = dalloc(s, (void**)&s.va, (size_t)n * 4))) return rc;
    if ((rc = dalloc(s, (void**)&s.kb, (size_t)n * 4))) return rc;
    if ((rc = dalloc(s, (void**)&s.vb, (size_t)n * 4))) return rc;
    launch_iota(s.iota, s.N, s.st);
    if ((rc = dalloc(s, (void**)&s.flags, (size_t)n * 4))) return rc;
    if ((rc = dalloc(s, (void**)&s.items, (size_t)n * 8))) return rc;  // starts + blocks
    int64_t hw = radix_hist_words(n);
    if ((rc = dalloc(s, (void**)&s.hist, (size_t)hw * 4))) return rc;
    int64_t tw = std::max(scan_temp_words(hw), scan_temp_words(n));
    if ((rc = dalloc(s, (void**)&s.scan_temp, (size_t)tw * 4))) return rc;
    if ((rc = dalloc(s, (void**)&s.counters, 64))) return rc;
    HIP_TRY(hipMemsetAsync(s.counters, 0, 64, s.st));
    if ((rc = dalloc(s, (void**)&s.acc, 64))) return rc;
    HIP_TRY(hipMemsetAsync(s.acc, 0, 64, s.st));
    // incremental sort of a large scene
    s.incr = s.N >= kIncMinParticles;
    if (s.incr) {
        if ((rc = dalloc(s, (void**)&s.snap_k, (size_t)n * 4))) return rc;
        if ((rc = dalloc(s, (void**)&s.snap_p, (size_t)n * 4))) return rc;
        if ((rc = dalloc(s, (void**)&s.plan, kPlanWords * 4))) return rc;
        HIP_TRY(hipMemsetAsync(s.plan, 0, kPlanWords * 4, s.st));  // snapshot not valid yet
        s.move_cap = inc_sublist_cap(s.N);
        uint32_t** lists[] = {&s.moves.key, &s.moves.slot, &s.moves.pos, &s.moves.lb, &s.moves.del_b, &s.moves.ins_b};
        for (auto pp : lists)
            if ((rc = dalloc(s, (void**)pp, (size_t)kMoveLists * s.move_cap * 4))) return rc;
        const size_t tw = (size_t)2 * (inc_tiles(s.N) + 1);
        if ((rc = dalloc(s, (void**)&s.moves.tcount, tw * 4))) return rc;
        HIP_TRY(hipMemsetAsync(s.moves.tcount, 0, tw * 4, s.st));  // the merge keeps them zero afterwards
        if ((rc = dalloc(s, (void**)&s.moves.tnext, tw * 4))) return rc;
        if ((rc = dalloc(s, (void**)&s.moves_dev, sizeof(MoveList)))) return rc;
        HIP_TRY(hipMemcpyAsync(s.plan + kPlanCap, &s.move_cap, 4, hipMemcpyHostToDevice, s.st));
        HIP_TRY(hipHostMalloc((void**)&s.inc_host, 64, hipHostMallocDefault));
        HIP_TRY(hipEventCreateWithFlags(&s.inc_ev, hipEventDisableTiming));
        HIP_TRY(hipMemcpyAsync(s.moves_dev, &s.moves, sizeof(MoveList), hipMemcpyHostToDevice, s.st));
        HIP_TRY(hipStreamSynchronize(s.st));  // the host copy is a member: it outlives the copy anyway
    }
    // particles
    if ((rc = alloc_particles(s, s.P, s.N))) return rc;
    if (s.N > 0) {
        std::vector<uint32_t> iota((size_t)s.N);
        for (int64_t i = 0; i < s.N; ++i) iota[(size_t)i] = (uint32_t)i;
        HIP_TRY(hipMemcpyAsync(s.P.pid, iota.data(), (size_t)s.N * 4, hipMemcpyHostToDevice, s.st));
        Staging g;
        if ((rc = alloc_staging(s, g, s.N))) return rc;
        const size_t f = (size_t)s.N * 4;
        HIP_TRY(hipMemcpyAsync(g.H.volume0, H.volume0.data(), f, hipMemcpyHostToDevice, s.st));
        HIP_TRY(hipMemcpyAsync(g.H.mass, H.mass.data(), f, hipMemcpyHostToDevice, s.st));
        HIP_TRY(hipMemcpyAsync(g.H.mu0, H.mu0.data(), f, hipMemcpyHostToDevice, s.st));
        HIP_TRY(hipMemcpyAsync(g.H.lam0, H.lam0.data(), f, hipMemcpyHostToDevice, s.st));
        HIP_TRY(hipMemcpyAsync(g.H.pos, H.pos.data(), 3 * f, hipMemcpyHostToDevice, s.st));
        HIP_TRY(hipMemcpyAsync(g.H.vel, H.vel.data(), 3 * f, hipMemcpyHostToDevice, s.st));
        HIP_TRY(hipMemcpyAsync(g.H.Fe, H.Fe.data(), 9 * f, hipMemcpyHostToDevice, s.st));
        HIP_TRY(hipMemcpyAsync(g.H.Fp, H.Fp.data(), 9 * f, hipMemcpyHostToDevice, s.st));
        HIP_TRY(hipMemcpyAsync(g.H.B, H.B.data(), 9 * f, hipMemcpyHostToDevice, s.st));
        HIP_TRY(hipMemcpyAsync(g.H.mat, H.mat.data(), (size_t)s.N * 2, hipMemcpyHostToDevice, s.st));
        launch_pack(g.H, s.P, s.st);
        HIP_TRY(hipStreamSynchronize(s.st));
        free_staging(s, g);
        // Physical reorder by cell so that per-step accesses stream.
        run_sort(s);
        ParticlesDev Q{};
        if ((rc = alloc_particles(s, Q, s.N))) return rc;
        launch_permute(s.P, Q, s.I.perm, s.st);
        HIP_TRY(hipStreamSynchronize(s.st));
        free_particles(s, s.P);
        s.P = Q;
        s.sort_valid = false;
        s.keys_valid = false;
    }
    HIP_TRY(hipGetLastError());
    if (run_init) run_initialize(s);
    HIP_TRY(hipStreamSynchronize(s.st));
    HIP_TRY(hipGetLastError());
    return MPM_OK;
}

void destroy(Solver* s) {
    if (!s) return;
    if (s->batch) batch_forget(s->batch, s);
    if (s->st) (void)hipStreamSynchronize(s->st);
    if (s->copy_st) {
        (void)hipStreamSynchronize(s->copy_st);
        (void)hipStreamDestroy(s->copy_st);
    }
    if (s->aux_st) {
        (void)hipStreamSynchronize(s->aux_st);
        (void)hipStreamDestroy(s->aux_st);
    }
    if (s->aux_fork) (void)hipEventDestroy(s->aux_fork);
    if (s->inc_ev) (void)hipEventDestroy(s->inc_ev);
    if (s->inc_host) (void)hipHostFree(s->inc_host);
    if (s->aux_join) (void)hipEventDestroy(s->aux_join);
    if (s->pos4_packed) (void)hipEventDestroy(s->pos4_packed);
    if (s->pos4_copied) (void)hipEventDestroy(s->pos4_copied);
    if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    for (auto& p : s->pending) s->event_pool.push_back(p.a), s->event_pool.push_back(p.b);
    for (auto e : s->event_pool) (void)hipEventDestroy(e);
    for (void* p : s->allocs) (void)devcache::release(p);  // every stream of the scene is synchronised above
    if (s->st) (void)hipStreamDestroy(s->st);
    delete s;
}

Solver* as_solver(void* h) { return static_cast<Solver*>(h); }

// A member of a batch leaves the batch's fused state before any call of
// its own (its grid pointers and stress cache become its own again).
#define GET_SOLVER(var, h)                                                         \
    Solver* var = as_solver(h);                                                    \
    if (!var) return fail(MPM_ERR_INVALID, "handle is NULL");                      \
    if (var->batch) batch_unfuse_member(var->batch, *var)

int finish(Solver& s) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(MPM_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return MPM_OK;
}

int create_common(const mpm_desc* desc, HostState& H, int64_t n, const std::vector<MaterialDev>& mats, int run_init,
                  void** out) {
    Solver* s = new (std::nothrow) Solver();
    if (!s) return fail(MPM_ERR_OOM, "host allocation failed");
    s->desc = *desc;
    s->mats = mats;
    s->interp = desc->interp;
    s->apic = desc->transfer == MPM_TRANSFER_APIC;
    s->N = n;
    int rc = build(*s, H, run_init);
    if (rc) {
        std::string keep = g_err;
        destroy(s);
        g_err = keep;
        return rc;
    }
    *out = s;
    return MPM_OK;
}

}  // namespace host
}  // namespace mpm

using namespace mpm::host;

extern "C" {

MPM_API const char* mpm_last_error(void) { return g_err.c_str(); }
MPM_API const char* mpm_version(void) { return "mpm_hip 1.0 (gfx950)"; }

MPM_API int mpm_set_cache_limit(int64_t bytes) {
    if (bytes < 0) return fail(MPM_ERR_INVALID, "cache limit < 0");
    devcache::set_limit((size_t)bytes);
    return MPM_OK;
}

MPM_API int mpm_get_cache_info(mpm_cache_info* out) {
    if (!out) return fail(MPM_ERR_INVALID, "out is NULL");
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= devcache::kMaxDevices) return fail(MPM_ERR_INVALID, "device index beyond the cache's table");
    devcache::State& st = devcache::state();
    std::lock_guard<std::mutex> g(st.mu);
    out->cached_bytes = (int64_t)st.cached[dev];
    out->live_bytes = (int64_t)st.live_bytes[dev];
    out->hits = (int64_t)st.hits;
    out->misses = (int64_t)st.misses;
    return MPM_OK;
}

MPM_API int mpm_release_cache(void) {
    devcache::release_all();
    return MPM_OK;
}

MPM_API int mpm_create(const mpm_desc* desc, const mpm_particle_input* in, void** out_handle) {
    if (!out_handle) return fail(MPM_ERR_INVALID, "out_handle is NULL");
    *out_handle = nullptr;
    int rc = check_desc(desc);
    if (rc) return rc;
    if (!in || in->count < 0) return fail(MPM_ERR_INVALID, "particle input is NULL or count < 0");
    const int64_t n = in->count;
    if (n > 0 && (!in->position || !in->velocity || !in->mass || !in->hardening || !in->young || !in->poisson ||
                  !in->critical_compression || !in->critical_stretch || !in->model))
        return fail(MPM_ERR_INVALID, "particle input has a NULL array");
    if (n >= (int64_t)0xFFFFFFFFll) return fail(MPM_ERR_INVALID, "too many particles");
    HostState H;
    H.volume0.assign((size_t)n, 0.0f);
    H.mass.assign(in->mass, in->mass + n);
    H.pos.assign(in->position, in->position + 3 * n);
    H.vel.assign(in->velocity, in->velocity + 3 * n);
    H.Fe.assign((size_t)(9 * n), 0.0f);
    H.Fp.assign((size_t)(9 * n), 0.0f);
    H.B.assign((size_t)(9 * n), 0.0f);
    H.mu0.resize((size_t)n);
    H.lam0.resize((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) {
            H.Fe[(size_t)(9 * i + 4 * k)] = 1.0f;
            H.Fp[(size_t)(9 * i + 4 * k)] = 1.0f;
        }
        // MaterialPoint3D constructor, src/material_point3d.cu:25-26
        const float E = in->young[i], nu = in->poisson[i];
        H.mu0[(size_t)i] = (float)(E / (2.0 * (1.0 + nu)));
        H.lam0[(size_t)i] = (float)((E * nu) / ((1.0 + nu) * (1.0 - 2.0 * nu)));
    }
    Solver probe;
    rc = dedupe_materials(probe, n, in->hardening, in->critical_compression, in->critical_stretch, in->model, H.mat);
    if (rc) return rc;
    return create_common(desc, H, n, probe.mats, 1, out_handle);
}

MPM_API int mpm_create_state(const mpm_desc* desc, const mpm_particle_state* st, int run_initialize,
                             void** out_handle) {
    if (!out_handle) return fail(MPM_ERR_INVALID, "out_handle is NULL");
    *out_handle = nullptr;
    int rc = check_desc(desc);
    if (rc) return rc;
    if (!st || st->count < 0) return fail(MPM_ERR_INVALID, "state is NULL or count < 0");
    const int64_t n = st->count;
    if (n > 0 && (!st->volume0 || !st->mass || !st->position || !st->velocity || !st->Fe || !st->Fp || !st->mu0 ||
                  !st->lambda0 || !st->hardening || !st->critical_compression || !st->critical_stretch || !st->model))
        return fail(MPM_ERR_INVALID, "state has a NULL array");
    HostState H;
    H.volume0.assign(st->volume0, st->volume0 + n);
    H.mass.assign(st->mass, st->mass + n);
    H.pos.assign(st->position, st->position + 3 * n);
    H.vel.assign(st->velocity, st->velocity + 3 * n);
    H.Fe.assign(st->Fe, st->Fe + 9 * n);
    H.Fp.assign(st->Fp, st->Fp + 9 * n);
    H.mu0.assign(st->mu0, st->mu0 + n);
    H.lam0.assign(st->lambda0, st->lambda0 + n);
    if (st->B) H.B.assign(st->B, st->B + 9 * n);
    else H.B.assign((size_t)(9 * n), 0.0f);
    Solver probe;
    rc = dedupe_materials(probe, n, st->hardening, st->critical_compression, st->critical_stretch, st->model, H.mat);
    if (rc) return rc;
    return create_common(desc, H, n, probe.mats, run_initialize, out_handle);
}

MPM_API void mpm_destroy(void* handle) { destroy(as_solver(handle)); }

// Permute the particle store into the cell order of the current positions
// (the slots each item reads become contiguous again); afterwards the
// sort's permutation is the identity and its input keys are the sorted keys.
static int reorder_store(Solver& s) {
    s.last_reorder = s.steps;
    if (s.N <= 0) return MPM_OK;
    if (!s.reorder_scratch) {
        int rc = dalloc(s, (void**)&s.reorder_scratch, (size_t)s.N * 9 * 4);
        if (rc) return rc;
    }
    ProfileHook h = hook(s);
    if (!s.sort_valid) run_sort(s);
    Scope sc(h, "reorder");
    launch_reorder_store(s.P, nullptr, 0, s.I.perm, s.N, s.reorder_scratch, s.apic != 0, s.st);
    HIP_TRY(hipMemcpyAsync(s.kin, s.I.sorted_keys, (size_t)s.N * 4, hipMemcpyDeviceToDevice, s.st));
    s.I.perm = s.iota;
    // slots renumbered: the last move list refers to the old ones (a
    // replayed step sorts in full; an eager step takes a new snapshot)
    if (s.incr) launch_inc_set(s.plan, 0u, s.st);
    s.snap_ok = false;
    s.keys_valid = true;
    s.sort_valid = true;
    return MPM_OK;
}

static bool reorder_due(const Solver& s) {
    return s.reorder_interval > 0 && s.steps - s.last_reorder >= s.reorder_interval;
}

MPM_API int mpm_set_reorder_interval(void* handle, int steps) {
    GET_SOLVER(s, handle);
    if (steps < 0) return fail(MPM_ERR_INVALID, "steps < 0");
    s->reorder_interval = steps;
    return MPM_OK;
}

MPM_API int mpm_set_transfer_fusion(void* handle, int enable) {
    GET_SOLVER(s, handle);
    if (!enable) unfuse(*s);
    s->fuse_transfer = enable != 0;
    return MPM_OK;
}

MPM_API int mpm_step(void* handle, int n) {
    GET_SOLVER(s, handle);
    if (n < 0) return fail(MPM_ERR_INVALID, "n < 0");
    auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < n; ++i) {
        if (reorder_due(*s)) {
            int rc = reorder_store(*s);
            if (rc) return rc;
        }
        run_fused_step(*s);
    }
    int rc = finish(*s);
    auto t1 = std::chrono::steady_clock::now();
    s->last_step_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    return rc;
}

MPM_API int mpm_synchronize(void* handle) {
    GET_SOLVER(s, handle);
    HIP_TRY(hipStreamSynchronize(s->st));
    return finish(*s);
}

MPM_API int mpm_step_graph(void* handle, int n) {
    GET_SOLVER(s, handle);
    if (n < 0) return fail(MPM_ERR_INVALID, "n < 0");
    if (n == 0) return MPM_OK;
    // A grid left by initialize / sub-steps is cleared outside the graph
    // (without counting); the captured step always starts with a clear.
    unfuse(*s);
    const bool fold = fused_folds(*s);
    if (s->grid_dirty && (!s->count_on_clear || s->vstar_dirty || (fold && s->force_dirty))) run_clear(*s);
    ensure_stress(*s);  // outside the graph: replays rely on G2P keeping it current
    if (!s->keys_valid) {  // likewise the sort input keys
        launch_keys(s->S, s->P, s->interp, s->kin, s->st);
        s->keys_valid = true;
        s->snap_ok = false;
    }
    if (s->gexec && s->S.grid != s->graph_grid) {
        // fused steps since the capture left the grid in the other buffer:
        // both cleared here (the last step's grid counted), then swapped
        if (s->grid_dirty) run_clear(*s);
        if (s->grid2_dirty) {
            SceneDev S2 = s->S;
            S2.grid = s->grid2;
            S2.touched = s->touched2;
            launch_clear(S2, s->nsb, GF_FOLDED, s->grid2_count ? s->acc : nullptr, s->st);
            s->grid2_dirty = false;
        }
        std::swap(s->S.grid, s->grid2);
        std::swap(s->S.touched, s->touched2);
    }
    if (s->incr && !s->snap_ok) {
        // the next replay sorts in full (outside the graph, which always
        // holds the planned sort)
        launch_inc_set(s->plan, 0u, s->st);
        s->snap_ok = true;
    }
    if (!s->gexec) {
        const int prof = s->prof;  // no events inside the capture
        s->prof = 0;
        HIP_TRY(hipStreamBeginCapture(s->st, hipStreamCaptureModeThreadLocal));
        bool dirty = s->grid_dirty;
        // force the clear node (of the planes the step writes, counting the
        // active nodes) and the sort into the graph
        s->grid_dirty = true;
        s->count_on_clear = true;
        s->vstar_dirty = false;
        s->force_dirty = !fold;
        s->sort_valid = false;
        s->capturing = true;  // the captured step is the two-kernel one
        run_fused_step(*s);
        s->capturing = false;
        s->steps -= 1;  // the capture ran nothing
        s->version -= 1;
        hipError_t e = hipStreamEndCapture(s->st, &s->graph);
        s->prof = prof;
        (void)dirty;
        if (e != hipSuccess) return fail(MPM_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        HIP_TRY(hipGraphInstantiate(&s->gexec, s->graph, nullptr, nullptr, 0));
        s->graph_captures += 1;
        s->graph_I = s->I;  // the sort buffers the replays write
        s->graph_grid = s->S.grid;
    }
    auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < n; ++i) {
        if (reorder_due(*s)) {
            // between replays: the eager sort of the keys the last replay's
            // G2P wrote, then the permutation (the next replay sorts again)
            s->sort_valid = false;
            int rc = reorder_store(*s);
            if (rc) return rc;
        }
        HIP_TRY(hipGraphLaunch(s->gexec, s->st));
        s->steps += 1;
    }
    s->version += (uint64_t)n;
    s->grid_dirty = true;
    s->count_on_clear = true;
    s->vstar_dirty = false;
    s->force_dirty = !fold;
    s->phase = fold ? PH_FOLDED : PH_RAW;
    s->sort_valid = false;
    s->snap_ok = s->incr;
    s->I = s->graph_I;  // a reorder between replays pointed the host copy at the identity
    auto t1 = std::chrono::steady_clock::now();
    s->last_step_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    return finish(*s);
}

MPM_API int mpm_initialize(void* handle) {
    GET_SOLVER(s, handle);
    run_initialize(*s);
    return finish(*s);
}

MPM_API int mpm_reset_grid(void* handle) {
    GET_SOLVER(s, handle);
    s->count_on_clear = false;
    run_clear(*s);
    return finish(*s);
}

MPM_API int mpm_particles_to_grid(void* handle) {
    GET_SOLVER(s, handle);
    unfuse(*s);
    ensure_stress(*s);
    if (!s->sort_valid) run_sort(*s);
    launch_p2g(s->S, s->P, s->I, s->interp, s->apic, 0, s->item_wgs, s->st);
    s->grid_dirty = true;
    s->count_on_clear = false;
    s->force_dirty = true;
    s->phase = PH_SUBSTEP;
    return finish(*s);
}

// The folded grid of a fused APIC step holds no momentum or force to work on.
static int require_unfolded(Solver& s, const char* what) {
    if (s.phase != PH_FOLDED) return MPM_OK;
    return fail(MPM_ERR_STATE, std::string(what) +
                                   ": the last fused APIC step summed momentum and dt*force per node; call "
                                   "mpm_set_grid_record(handle, 1) before stepping to keep them apart");
}

MPM_API int mpm_set_grid_record(void* handle, int enable) {
    GET_SOLVER(s, handle);
    if (s->record_grid == (enable != 0)) return MPM_OK;
    s->record_grid = enable != 0;
    // the captured step and a batch's launches depend on the mode
    if (s->gexec) { (void)hipGraphExecDestroy(s->gexec); s->gexec = nullptr; }
    if (s->graph) { (void)hipGraphDestroy(s->graph); s->graph = nullptr; }
    if (s->batch) batch_invalidate(s->batch);
    return MPM_OK;
}

MPM_API int mpm_compute_gravity_forces(void* handle) {
    GET_SOLVER(s, handle);
    if (int rc = require_unfolded(*s, "computeGravityForces")) return rc;
    unfuse(*s);
    s->force_dirty = true;
    launch_grid_op(s->S, s->nsb, 0, s->S.dt, s->st);
    s->phase = PH_SUBSTEP;
    return finish(*s);
}

MPM_API int mpm_compute_external_forces(void* handle) { return mpm_compute_gravity_forces(handle); }

MPM_API int mpm_update_grid_velocities(void* handle, float dt) {
    GET_SOLVER(s, handle);
    if (int rc = require_unfolded(*s, "updateGridVelocities")) return rc;
    unfuse(*s);
    launch_grid_op(s->S, s->nsb, 1, dt, s->st);
    s->vstar_dirty = true;
    s->phase = PH_SUBSTEP;
    return finish(*s);
}

MPM_API int mpm_solve_linear_system(void* handle, float dt) {
    GET_SOLVER(s, handle);
    (void)dt;  // empty in the reference (src/mpm_solver3d.cu:312-314)
    (void)s;
    return MPM_OK;
}

MPM_API int mpm_grid_collision(void* handle, float dt) {
    GET_SOLVER(s, handle);
    if (int rc = require_unfolded(*s, "gridCollision")) return rc;
    unfuse(*s);
    launch_grid_op(s->S, s->nsb, 2, dt, s->st);
    s->vstar_dirty = true;
    s->phase = PH_SUBSTEP;
    return finish(*s);
}

MPM_API int mpm_grid_to_particles(void* handle, float dt, float blend) {
    GET_SOLVER(s, handle);
    if (int rc = require_unfolded(*s, "gridToParticles")) return rc;
    unfuse(*s);
    if (s->phase == PH_RAW) materialize(*s);
    if (!s->sort_valid) run_sort(*s);
    SceneDev S = s->S;
    S.dt = dt;
    S.blend = blend;
    launch_g2p(S, s->P, s->I, s->interp, s->apic, 1, 0, 0, s->item_wgs, s->st);
    return finish(*s);
}

MPM_API int mpm_particles_collision(void* handle, float dt) {
    GET_SOLVER(s, handle);
    unfuse(*s);
    launch_particles_collision(s->S, s->P, dt, s->st);
    return finish(*s);
}

MPM_API int mpm_advect_particles(void* handle, float dt) {
    GET_SOLVER(s, handle);
    unfuse(*s);
    launch_advect(s->P, dt, s->st);
    s->sort_valid = false;
    s->keys_valid = false;
    s->version += 1;
    return finish(*s);
}

MPM_API int mpm_switch_particles_collision(void* handle) {
    GET_SOLVER(s, handle);
    s->S.particle_collision = s->S.particle_collision ? 0 : 1;
    s->desc.particle_collision = s->S.particle_collision;
    if (s->gexec) { (void)hipGraphExecDestroy(s->gexec); s->gexec = nullptr; }
    if (s->graph) { (void)hipGraphDestroy(s->graph); s->graph = nullptr; }
    return MPM_OK;
}

MPM_API int mpm_set_particles_collision(void* handle, int enable) {
    GET_SOLVER(s, handle);
    if ((s->S.particle_collision != 0) != (enable != 0)) return mpm_switch_particles_collision(handle);
    return MPM_OK;
}

MPM_API int mpm_sort(void* handle) {
    GET_SOLVER(s, handle);
    run_sort(*s);
    return finish(*s);
}

// The host-layout staging H (filled by launch_unpack on the solver's stream)
// copied into the caller's arrays; waits for the copies.
static hipError_t copy_out_staging(Solver& s, const HostLayoutDev& H, mpm_particle_state* out,
                                   std::vector<uint16_t>& mat) {
    const int64_t n = s.N;
    const size_t f = (size_t)n * 4;
    mat.resize((size_t)n);
    auto cp = [&](void* dst, const void* src, size_t bytes) -> hipError_t {
        if (!dst) return hipSuccess;
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s.st);
    };
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = cp(out->volume0, H.volume0, f);
    if (e == hipSuccess) e = cp(out->mass, H.mass, f);
    if (e == hipSuccess) e = cp(out->position, H.pos, 3 * f);
    if (e == hipSuccess) e = cp(out->velocity, H.vel, 3 * f);
    if (e == hipSuccess) e = cp(out->Fe, H.Fe, 9 * f);
    if (e == hipSuccess) e = cp(out->Fp, H.Fp, 9 * f);
    if (e == hipSuccess) e = cp(out->mu0, H.mu0, f);
    if (e == hipSuccess) e = cp(out->lambda0, H.lam0, f);
    if (e == hipSuccess && out->B) {
        if (s.apic) e = cp(out->B, H.B, 9 * f);
        else std::memset(out->B, 0, 9 * f);
    }
    if (e == hipSuccess) e = cp(mat.data(), H.mat, (size_t)n * 2);
    if (e == hipSuccess) e = hipStreamSynchronize(s.st);
    return e;
}

// Per-particle material constants from the material indices.
static void fill_materials(const Solver& s, const std::vector<uint16_t>& mat, mpm_particle_state* out) {
    for (int64_t i = 0; i < s.N; ++i) {
        const MaterialDev& m = s.mats[mat[(size_t)i]];
        if (out->hardening) out->hardening[i] = m.hardening;
        if (out->critical_compression) out->critical_compression[i] = m.crit_c;
        if (out->critical_stretch) out->critical_stretch[i] = m.crit_s;
        if (out->model) out->model[i] = m.model;
    }
}

MPM_API int mpm_read_particles(void* handle, mpm_particle_state* out) {
    GET_SOLVER(s, handle);
    if (!out || out->count != s->N) return fail(MPM_ERR_INVALID, "output count must equal the particle count");
    const int64_t n = s->N;
    if (n == 0) return MPM_OK;
    Staging g;
    int rc = alloc_staging(*s, g, n);
    if (rc) { free_staging(*s, g); return rc; }
    launch_unpack(s->P, g.H, s->st);
    std::vector<uint16_t> mat;
    const hipError_t e = copy_out_staging(*s, g.H, out, mat);
    free_staging(*s, g);
    if (e != hipSuccess) return fail(MPM_ERR_HIP, std::string("read particles: ") + hipGetErrorString(e));
    fill_materials(*s, mat, out);
    return finish(*s);
}

// ------------------------------------------- allocation-free readback
// mpm_read_particles_noalloc stages through this file-scope device array
// instead of hipMalloc'd buffers: 37 floats per particle (volume0, mass,
// mu0, lambda0, x3, v3, Fe9, Fp9, B9) for up to kNoAllocParticles
// particles, then their 16-bit material indices.  Each GPU holds its own
// copy (39.3 MB) until the process ends.  Host code never names the array
// except to take its address: `scratch` is a host-side stand-in there.
constexpr int64_t kNoAllocParticles = int64_t(1) << 18;
constexpr int64_t kNoAllocFloats = 37 * kNoAllocParticles + kNoAllocParticles / 2;
static __device__ __align__(256) float scratch[kNoAllocFloats];

constexpr int kNoAllocDevices = 64;
static std::mutex g_scratch_mu;  // guards g_scratch_ptr
static float* g_scratch_ptr[kNoAllocDevices] = {};
static std::mutex g_scratch_use[kNoAllocDevices];  // one readback at a time per GPU

// This device's address of `scratch`, taken once per device with that
// device current (hipGetSymbolAddress; every device has its own copy).
static int scratch_address(int dev, float** out) {
    if (dev < 0 || dev >= kNoAllocDevices) return fail(MPM_ERR_UNSUPPORTED, "device index out of range");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(MPM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<float*>(p);
    }
    *out = g_scratch_ptr[dev];
    return MPM_OK;
}

MPM_API int mpm_read_particles_noalloc(void* handle, mpm_particle_state* out) {
    GET_SOLVER(s, handle);
    if (!out || out->count != s->N) return fail(MPM_ERR_INVALID, "output count must equal the particle count");
    const int64_t n = s->N;
    if (n > kNoAllocParticles)
        return fail(MPM_ERR_UNSUPPORTED, "mpm_read_particles_noalloc reads at most " +
                                             std::to_string(kNoAllocParticles) + " particles (this handle has " +
                                             std::to_string(n) + "); use mpm_read_particles");
    if (n == 0) return MPM_OK;
    // the solver's device is current for the address and the launch
    int prev = 0;
    HIP_TRY(hipGetDevice(&prev));
    if (prev != s->device) HIP_TRY(hipSetDevice(s->device));
    struct Restore {
        int dev, prev;
        ~Restore() { if (dev != prev) (void)hipSetDevice(prev); }
    } restore{s->device, prev};
    float* base = nullptr;
    int rc = scratch_address(s->device, &base);
    if (rc) return rc;
    std::lock_guard<std::mutex> use(g_scratch_use[s->device]);
    // regions sized for kNoAllocParticles: every entry below n is written
    // by the unpack (pid is a permutation), so nothing needs clearing
    const int64_t c = kNoAllocParticles;
    HostLayoutDev H{};
    H.volume0 = base;
    H.mass = base + c;
    H.mu0 = base + 2 * c;
    H.lam0 = base + 3 * c;
    H.pos = base + 4 * c;
    H.vel = base + 7 * c;
    H.Fe = base + 10 * c;
    H.Fp = base + 19 * c;
    H.B = base + 28 * c;
    H.mat = reinterpret_cast<uint16_t*>(base + 37 * c);
    launch_unpack(s->P, H, s->st);
    std::vector<uint16_t> mat;
    const hipError_t e = copy_out_staging(*s, H, out, mat);
    if (e != hipSuccess) return fail(MPM_ERR_HIP, std::string("read particles (no alloc): ") + hipGetErrorString(e));
    fill_materials(*s, mat, out);
    return finish(*s);
}

MPM_API int mpm_write_particles(void* handle, const mpm_particle_state* in) {
    GET_SOLVER(s, handle);
    if (!in || in->count != s->N) return fail(MPM_ERR_INVALID, "state count must equal the particle count");
    const int64_t n = s->N;
    if (n == 0) return MPM_OK;
    if (!in->volume0 || !in->mass || !in->position || !in->velocity || !in->Fe || !in->Fp || !in->mu0 ||
        !in->lambda0 || !in->hardening || !in->critical_compression || !in->critical_stretch || !in->model)
        return fail(MPM_ERR_INVALID, "state has a NULL array");
    unfuse(*s);
    std::vector<uint16_t> mat;
    size_t before = s->mats.size();
    int rc = dedupe_materials(*s, n, in->hardening, in->critical_compression, in->critical_stretch, in->model, mat);
    if (rc) return rc;
    if (s->mats.size() != before) {
        HIP_TRY(hipStreamSynchronize(s->st));
        if ((rc = upload_materials(*s))) return rc;
    }
    Staging g;
    rc = alloc_staging(*s, g, n);
    if (rc) { free_staging(*s, g); return rc; }
    const size_t f = (size_t)n * 4;
    std::vector<float> zeros;
    const float* B = in->B;
    if (!B) { zeros.assign((size_t)(9 * n), 0.0f); B = zeros.data(); }
    hipError_t e = hipSuccess;
    auto cp = [&](void* dst, const void* src, size_t bytes) {
        if (e == hipSuccess) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s->st);
    };
    cp(g.H.volume0, in->volume0, f);
    cp(g.H.mass, in->mass, f);
    cp(g.H.pos, in->position, 3 * f);
    cp(g.H.vel, in->velocity, 3 * f);
    cp(g.H.Fe, in->Fe, 9 * f);
    cp(g.H.Fp, in->Fp, 9 * f);
    cp(g.H.mu0, in->mu0, f);
    cp(g.H.lam0, in->lambda0, f);
    cp(g.H.B, B, 9 * f);
    cp(g.H.mat, mat.data(), (size_t)n * 2);
    if (e == hipSuccess) launch_pack(g.H, s->P, s->st);
    if (e == hipSuccess) e = hipStreamSynchronize(s->st);
    free_staging(*s, g);
    if (e != hipSuccess) return fail(MPM_ERR_HIP, std::string("write particles: ") + hipGetErrorString(e));
    s->sort_valid = false;
    s->keys_valid = false;
    s->version += 1;
    s->stress_valid = false;
    return finish(*s);
}

MPM_API int mpm_read_grid(void* handle, mpm_grid_host* out) {
    GET_SOLVER(s, handle);
    const int64_t total = (int64_t)s->S.rx * s->S.ry * s->S.rz;
    if (!out || out->nodes != total) return fail(MPM_ERR_INVALID, "grid node count mismatch");
    if (int rc = require_unfolded(*s, "read grid")) return rc;
    if (s->phase == PH_RAW) materialize(*s);
    float *m = nullptr, *v = nullptr, *w = nullptr, *f = nullptr;
    int rc;
    if ((rc = dalloc(*s, (void**)&m, (size_t)total * 4))) return rc;
    if ((rc = dalloc(*s, (void**)&v, (size_t)total * 12))) { dfree(*s, m); return rc; }
    if ((rc = dalloc(*s, (void**)&w, (size_t)total * 12))) { dfree(*s, m); dfree(*s, v); return rc; }
    if ((rc = dalloc(*s, (void**)&f, (size_t)total * 12))) { dfree(*s, m); dfree(*s, v); dfree(*s, w); return rc; }
    launch_grid_dense(s->S, m, v, w, f, s->vstar_dirty ? 1 : 0, s->st);
    hipError_t e = hipSuccess;
    if (out->mass) e = hipMemcpyAsync(out->mass, m, (size_t)total * 4, hipMemcpyDeviceToHost, s->st);
    if (e == hipSuccess && out->velocity)
        e = hipMemcpyAsync(out->velocity, v, (size_t)total * 12, hipMemcpyDeviceToHost, s->st);
    if (e == hipSuccess && out->velocity_star)
        e = hipMemcpyAsync(out->velocity_star, w, (size_t)total * 12, hipMemcpyDeviceToHost, s->st);
    if (e == hipSuccess && out->force)
        e = hipMemcpyAsync(out->force, f, (size_t)total * 12, hipMemcpyDeviceToHost, s->st);
    if (e == hipSuccess) e = hipStreamSynchronize(s->st);
    dfree(*s, m); dfree(*s, v); dfree(*s, w); dfree(*s, f);
    if (e != hipSuccess) return fail(MPM_ERR_HIP, std::string("read grid: ") + hipGetErrorString(e));
    return finish(*s);
}

MPM_API int mpm_read_positions4(void* handle, float* out) {
    GET_SOLVER(s, handle);
    if (!out) return fail(MPM_ERR_INVALID, "out is NULL");
    if (s->N == 0) return MPM_OK;
    float* d = nullptr;
    int rc = dalloc(*s, (void**)&d, (size_t)s->N * 16);
    if (rc) return rc;
    launch_positions4(s->P, d, s->st);
    hipError_t e = hipMemcpyAsync(out, d, (size_t)s->N * 16, hipMemcpyDeviceToHost, s->st);
    if (e == hipSuccess) e = hipStreamSynchronize(s->st);
    dfree(*s, d);
    if (e != hipSuccess) return fail(MPM_ERR_HIP, std::string("read positions: ") + hipGetErrorString(e));
    return finish(*s);
}

// ---- render export (updateGLBufferWithCUDA / updateGLBufferByCPU,
// src/mpm_solver3d.cu:457-508)

MPM_API int mpm_pack_positions4(void* handle, float* device_out) {
    GET_SOLVER(s, handle);
    if (!device_out) return fail(MPM_ERR_INVALID, "device_out is NULL");
    launch_positions4(s->P, device_out, s->st);
    return finish(*s);
}

MPM_API int mpm_positions4_async(void* handle, float* host_out) {
    GET_SOLVER(s, handle);
    if (!host_out) return fail(MPM_ERR_INVALID, "host_out is NULL");
    if (s->N == 0) return MPM_OK;
    if (!s->pos4) {
        if (!s->copy_st) HIP_TRY(hipStreamCreateWithFlags(&s->copy_st, hipStreamNonBlocking));
        if (!s->pos4_packed) HIP_TRY(hipEventCreateWithFlags(&s->pos4_packed, hipEventDisableTiming));
        if (!s->pos4_copied) HIP_TRY(hipEventCreateWithFlags(&s->pos4_copied, hipEventDisableTiming));
        int rc = dalloc(*s, (void**)&s->pos4, (size_t)s->N * 16);
        if (rc) return rc;
    }
    // the staging buffer is reused once the previous copy has read it
    if (s->pos4_pending) HIP_TRY(hipStreamWaitEvent(s->st, s->pos4_copied, 0));
    launch_positions4(s->P, s->pos4, s->st);
    HIP_TRY(hipEventRecord(s->pos4_packed, s->st));
    HIP_TRY(hipStreamWaitEvent(s->copy_st, s->pos4_packed, 0));
    HIP_TRY(hipMemcpyAsync(host_out, s->pos4, (size_t)s->N * 16, hipMemcpyDeviceToHost, s->copy_st));
    HIP_TRY(hipEventRecord(s->pos4_copied, s->copy_st));
    s->pos4_pending = true;
    return finish(*s);
}

MPM_API int mpm_positions4_wait(void* handle) {
    GET_SOLVER(s, handle);
    if (s->pos4_pending) {
        HIP_TRY(hipEventSynchronize(s->pos4_copied));
        s->pos4_pending = false;
    }
    return MPM_OK;
}

// ---- density raster (writeToOpenVDB, src/mpm_solver3d.cu:510-572)

// The raster's one-plane grid: the scene's geometry with the density
// buffers in place of the simulation grid -- or, after a fused step, the
// simulation grid itself: its first plane holds the next step's P2G mass,
// sum m w over the current positions, which is the raster's sum.
static SceneDev density_scene(const Solver& s) {
    SceneDev D = s.S;
    if (s.prescat) return D;
    D.grid = decltype(D.grid)(s.dens);
    D.touched = decltype(D.touched)(s.dens_touched);
    return D;
}

// Mass-only P2G of the current positions into the density plane (skipped
// when the particles have not changed since the last raster, and after a
// fused step, whose grid already holds the sums).
static int run_density(Solver& s) {
    if (s.dens && (s.prescat || s.dens_version == s.version)) return MPM_OK;
    ProfileHook h = hook(s);
    Scope sc(h, "density");
    if (!s.dens) {
        const size_t sizes[4] = {(size_t)s.S.plane * 4, (size_t)s.nsb * 4, (size_t)(s.nsb + 1) * 4,
                                 (size_t)scan_temp_words(s.nsb) * 4};
        void* p[4] = {nullptr, nullptr, nullptr, nullptr};
        for (int i = 0; i < 4; ++i) {
            int rc = dalloc(s, &p[i], sizes[i]);
            if (rc) {
                for (int j = 0; j < i; ++j) dfree(s, p[j]);
                return rc;
            }
        }
        HIP_TRY(hipMemsetAsync(p[0], 0, sizes[0], s.st));
        HIP_TRY(hipMemsetAsync(p[1], 0, sizes[1], s.st));
        s.dens = (float*)p[0];
        s.dens_touched = (uint32_t*)p[1];
        s.dens_rank = (uint32_t*)p[2];
        s.dens_temp = (uint32_t*)p[3];
        s.dens_dirty = false;
    }
    if (s.prescat) return MPM_OK;  // the buffers above serve the compaction
    const SceneDev D = density_scene(s);
    if (s.dens_dirty) launch_clear(D, s.nsb, 1, nullptr, s.st);
    if (s.N > 0) {
        if (!s.sort_valid) run_sort(s);
        launch_p2g(D, s.P, s.I, s.interp, s.apic, 1, s.item_wgs, s.st);
    }
    s.dens_dirty = true;
    s.dens_version = s.version;
    return MPM_OK;
}

MPM_API int mpm_density_compute(void* handle) {
    GET_SOLVER(s, handle);
    int rc = run_density(*s);
    if (rc) return rc;
    return finish(*s);
}

MPM_API int mpm_read_density(void* handle, float* out) {
    GET_SOLVER(s, handle);
    if (!out) return fail(MPM_ERR_INVALID, "out is NULL");
    int rc = run_density(*s);
    if (rc) return rc;
    const int64_t total = (int64_t)s->S.rx * s->S.ry * s->S.rz;
    float* d = nullptr;
    if ((rc = dalloc(*s, (void**)&d, (size_t)total * 4))) return rc;
    launch_density_dense(density_scene(*s), d, s->st);
    hipError_t e = hipMemcpyAsync(out, d, (size_t)total * 4, hipMemcpyDeviceToHost, s->st);
    if (e == hipSuccess) e = hipStreamSynchronize(s->st);
    dfree(*s, d);
    if (e != hipSuccess) return fail(MPM_ERR_HIP, std::string("read density: ") + hipGetErrorString(e));
    return finish(*s);
}

// Touched blocks of the raster, compacted in storage-block order.
static int density_blocks(Solver& s, std::vector<int32_t>* origins, std::vector<float>* values, int32_t* h_orig,
                          float* h_vals, int64_t max_blocks, int64_t* count) {
    int rc = run_density(s);
    if (rc) return rc;
    const SceneDev D = density_scene(s);
    launch_touched_flags((const uint32_t*)D.touched, s.nsb, s.dens_rank, s.st);
    exclusive_scan(s.dens_rank, s.nsb, s.dens_temp, s.dens_rank + s.nsb, s.st);
    uint32_t total = 0;
    HIP_TRY(hipMemcpyAsync(&total, s.dens_rank + s.nsb, 4, hipMemcpyDeviceToHost, s.st));
    HIP_TRY(hipStreamSynchronize(s.st));
    if (count) *count = (int64_t)total;
    if (origins) { origins->resize((size_t)total * 3); h_orig = origins->data(); }
    if (values) { values->resize((size_t)total * 64); h_vals = values->data(); max_blocks = total; }
    if (!h_orig && !h_vals) return finish(s);
    if (max_blocks < (int64_t)total)
        return fail(MPM_ERR_INVALID, "max_blocks (" + std::to_string(max_blocks) + ") < touched blocks (" +
                                         std::to_string(total) + ")");
    if (total == 0) return finish(s);
    int32_t* d_orig = nullptr;
    float* d_vals = nullptr;
    if ((rc = dalloc(s, (void**)&d_orig, (size_t)total * 12))) return rc;
    if ((rc = dalloc(s, (void**)&d_vals, (size_t)total * 256))) { dfree(s, d_orig); return rc; }
    launch_density_blocks(D, (const uint32_t*)D.touched, s.dens_rank, s.nsb, d_orig, d_vals, total, s.st);
    hipError_t e = hipSuccess;
    if (h_orig) e = hipMemcpyAsync(h_orig, d_orig, (size_t)total * 12, hipMemcpyDeviceToHost, s.st);
    if (e == hipSuccess && h_vals) e = hipMemcpyAsync(h_vals, d_vals, (size_t)total * 256, hipMemcpyDeviceToHost, s.st);
    if (e == hipSuccess) e = hipStreamSynchronize(s.st);
    dfree(s, d_orig);
    dfree(s, d_vals);
    if (e != hipSuccess) return fail(MPM_ERR_HIP, std::string("read density blocks: ") + hipGetErrorString(e));
    return finish(s);
}

MPM_API int mpm_read_density_blocks(void* handle, int32_t* origins, float* values, int64_t max_blocks,
                                    int64_t* count) {
    GET_SOLVER(s, handle);
    if (!count) return fail(MPM_ERR_INVALID, "count is NULL");
    return density_blocks(*s, nullptr, nullptr, origins, values, max_blocks, count);
}

MPM_API int mpm_write_density(void* handle, const char* path) {
    GET_SOLVER(s, handle);
    if (!path) return fail(MPM_ERR_INVALID, "path is NULL");
    std::vector<int32_t> org;
    std::vector<float> val;
    int64_t nb = 0;
    int rc = density_blocks(*s, &org, &val, nullptr, nullptr, 0, &nb);
    if (rc) return rc;
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(MPM_ERR_IO, std::string("cannot open ") + path + " for writing");
    const int32_t res[3] = {s->S.rx, s->S.ry, s->S.rz};
    const float geo[4] = {s->desc.origin[0], s->desc.origin[1], s->desc.origin[2], s->desc.stride};
    const int32_t edge = MPM_BLOCK;
    bool ok = std::fwrite(MPM_DENSITY_MAGIC, 1, 8, f) == 8 && std::fwrite(res, 4, 3, f) == 3 &&
              std::fwrite(geo, 4, 4, f) == 4 && std::fwrite(&edge, 4, 1, f) == 1 && std::fwrite(&nb, 8, 1, f) == 1;
    for (int64_t b = 0; ok && b < nb; ++b)
        ok = std::fwrite(&org[(size_t)b * 3], 4, 3, f) == 3 && std::fwrite(&val[(size_t)b * 64], 4, 64, f) == 64;
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return fail(MPM_ERR_IO, std::string("write failed: ") + path);
    return MPM_OK;
}

MPM_API int mpm_read_sort(void* handle, uint32_t* keys, uint32_t* ids, uint32_t* seq, int64_t count) {
    GET_SOLVER(s, handle);
    if (count != s->N) return fail(MPM_ERR_INVALID, "count must equal the particle count");
    if (s->N == 0) return MPM_OK;
    if (!s->I.sorted_keys) return fail(MPM_ERR_STATE, "no sort has run");
    uint32_t* tmp = nullptr;
    int rc = dalloc(*s, (void**)&tmp, (size_t)s->N * 4);
    if (rc) return rc;
    // The particle storage order (pid) is fixed after creation, so it is
    // the input sequence of every sort.
    launch_gather_u32(s->P.pid, s->I.perm, tmp, s->N, s->st);
    hipError_t e = hipSuccess;
    if (keys) e = hipMemcpyAsync(keys, s->I.sorted_keys, (size_t)s->N * 4, hipMemcpyDeviceToHost, s->st);
    if (e == hipSuccess && ids) e = hipMemcpyAsync(ids, tmp, (size_t)s->N * 4, hipMemcpyDeviceToHost, s->st);
    if (e == hipSuccess && seq) e = hipMemcpyAsync(seq, s->P.pid, (size_t)s->N * 4, hipMemcpyDeviceToHost, s->st);
    if (e == hipSuccess) e = hipStreamSynchronize(s->st);
    dfree(*s, tmp);
    if (e != hipSuccess) return fail(MPM_ERR_HIP, std::string("read sort: ") + hipGetErrorString(e));
    return finish(*s);
}

MPM_API int mpm_host_keys(void* handle, const float* pos, int64_t n, uint32_t* keys) {
    GET_SOLVER(s, handle);
    if (n < 0 || (n > 0 && (!pos || !keys))) return fail(MPM_ERR_INVALID, "bad arguments");
    const SceneDev& S = s->S;
    const float shift = s->interp == MPM_INTERP_LINEAR ? 0.0f : s->interp == MPM_INTERP_QUADRATIC_BSPLINE ? 0.5f : 1.0f;
    const int width = s->interp == MPM_INTERP_LINEAR ? 2 : s->interp == MPM_INTERP_QUADRATIC_BSPLINE ? 3 : 4;
    const int res[3] = {S.rx, S.ry, S.rz};
    for (int64_t i = 0; i < n; ++i) {
        int b[3];
        bool esc = false;
        for (int a = 0; a < 3; ++a) {
            volatile float u = pos[3 * i + a] / S.h;
            volatile float t = u - shift;
            float fb = std::floor(t);
            if (!(fb >= (float)(-(width - 1)) && fb <= (float)(res[a] - 1))) { esc = true; break; }
            b[a] = (int)fb;
        }
        if (esc) { keys[i] = S.escape_key; continue; }
        uint32_t blk = (uint32_t)((b[0] + 4) >> 2) +
                       (uint32_t)S.kbx * ((uint32_t)((b[1] + 4) >> 2) + (uint32_t)S.kby * (uint32_t)((b[2] + 4) >> 2));
        keys[i] = blk * 64u + ((uint32_t)((b[0] + 4) & 3) | ((uint32_t)((b[1] + 4) & 3) << 2) |
                               ((uint32_t)((b[2] + 4) & 3) << 4));
    }
    return MPM_OK;
}

MPM_API int mpm_get_stats(void* handle, mpm_stats* out) {
    GET_SOLVER(s, handle);
    if (!out) return fail(MPM_ERR_INVALID, "out is NULL");
    std::memset(out, 0, sizeof(*out));
    unsigned long long zero = 0;
    HIP_TRY(hipMemcpyAsync(&s->acc[1], &zero, 8, hipMemcpyHostToDevice, s->st));
    HIP_TRY(hipMemcpyAsync(&s->acc[2], &zero, 8, hipMemcpyHostToDevice, s->st));
    if (s->grid_dirty) launch_count_active(s->S, s->nsb, &s->acc[1], s->st);
    if (s->I.sorted_keys && s->N > 0) launch_count_escaped(s->I.sorted_keys, s->N, s->S.escape_key, &s->acc[2], s->st);
    // acc[3] (low word): max |v|^2 bits
    HIP_TRY(hipMemcpyAsync(&s->acc[3], &zero, 8, hipMemcpyHostToDevice, s->st));
    launch_max_speed2(s->P, reinterpret_cast<unsigned int*>(&s->acc[3]), s->st);
    unsigned long long a[4] = {0, 0, 0, 0};
    uint32_t items = 0, plan[kPlanWords] = {0};
    HIP_TRY(hipMemcpyAsync(a, s->acc, 32, hipMemcpyDeviceToHost, s->st));
    HIP_TRY(hipMemcpyAsync(&items, s->counters, 4, hipMemcpyDeviceToHost, s->st));
    if (s->plan) HIP_TRY(hipMemcpyAsync(plan, s->plan, sizeof(plan), hipMemcpyDeviceToHost, s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
    out->particles = s->N;
    out->steps = s->steps;
    out->active_nodes = (int64_t)a[1];
    out->active_nodes_sum = (int64_t)a[0] + (s->count_on_clear ? (int64_t)a[1] : 0);
    out->work_items = items;
    out->escaped = (int64_t)a[2];
    out->device_bytes = s->device_bytes;
    out->last_step_ms = s->last_step_ms;
    float v2;
    const uint32_t bits = (uint32_t)(a[3] & 0xFFFFFFFFull);
    std::memcpy(&v2, &bits, 4);
    out->max_speed = std::sqrt((double)v2);
    out->cfl = out->max_speed * (double)s->S.dt / (double)s->S.h;
    out->graph_captures = s->graph_captures;
    out->incremental_sorts = plan[kPlanIncSteps];
    return finish(*s);
}

MPM_API int mpm_profile_enable(void* handle, int enable) {
    GET_SOLVER(s, handle);
    if (enable < 0 || enable > 2) return fail(MPM_ERR_INVALID, "profile level must be 0, 1 or 2");
    s->prof = enable;
    return MPM_OK;
}

MPM_API int mpm_profile_reset(void* handle) {
    GET_SOLVER(s, handle);
    prof_collect(*s);
    s->prof_tot.clear();
    s->prof_order.clear();
    return MPM_OK;
}

MPM_API int mpm_profile_read(void* handle, mpm_kernel_time* out, int max, int* count) {
    GET_SOLVER(s, handle);
    prof_collect(*s);
    int k = 0;
    for (auto& name : s->prof_order) {
        if (out && k < max) {
            std::memset(&out[k], 0, sizeof(mpm_kernel_time));
            std::strncpy(out[k].name, name.c_str(), sizeof(out[k].name) - 1);
            out[k].launches = s->prof_tot[name].first;
            out[k].total_ms = s->prof_tot[name].second;
        }
        ++k;
    }
    if (count) *count = k;
    return MPM_OK;
}

MPM_API int mpm_debug_svd3(const float* A, int64_t n, float* U, float* sv, float* V) {
    if (n < 0 || (n > 0 && (!A || !U || !sv || !V))) return fail(MPM_ERR_INVALID, "bad arguments");
    if (n == 0) return MPM_OK;
    float* d[4] = {nullptr, nullptr, nullptr, nullptr};  // A, U, s, V
    const size_t bytes[4] = {(size_t)n * 36, (size_t)n * 36, (size_t)n * 12, (size_t)n * 36};
    hipError_t e = hipSuccess;
    for (int i = 0; i < 4 && e == hipSuccess; ++i) e = hipMalloc((void**)&d[i], bytes[i]);
    if (e == hipSuccess) e = hipMemcpy(d[0], A, bytes[0], hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        launch_debug_svd(d[0], n, d[1], d[2], d[3], nullptr);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(U, d[1], bytes[1], hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(sv, d[2], bytes[2], hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(V, d[3], bytes[3], hipMemcpyDeviceToHost);
    for (float* p : d)
        if (p) (void)hipFree(p);
    if (e != hipSuccess) return fail(MPM_ERR_HIP, std::string("debug svd: ") + hipGetErrorString(e));
    return MPM_OK;
}

}  // extern "C"
