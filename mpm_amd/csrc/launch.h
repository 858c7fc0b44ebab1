// launch.h -- host-side launchers of the device kernels (one TU per kernel
// family; the C-ABI in capi.hip calls these).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.h"

namespace mpm {

// Optional per-kernel timing: begin/end bracket the launches of one named
// kernel on the solver's stream.
struct ProfileHook {
    void* ctx = nullptr;
    void (*fn_begin)(void*, const char*) = nullptr;
    void (*fn_end)(void*, const char*) = nullptr;
    void begin(const char* n) const { if (fn_begin) fn_begin(ctx, n); }
    void end(const char* n) const { if (fn_end) fn_end(ctx, n); }
};

// ---- sort.hip
int64_t scan_temp_words(int64_t n);
// plan (device, may be null): every launch returns at once when the plan
// words choose the incremental update (plan_incremental).
void exclusive_scan(uint32_t* data, int64_t n, uint32_t* temp, uint32_t* total, hipStream_t st,
                    const uint32_t* plan = nullptr, int max_wgs = 0);
int64_t radix_tiles(int64_t n);
int64_t radix_hist_words(int64_t n);
int radix_passes(int key_bits, int64_t n, bool few_passes = false);
int radix_sort_pairs(const uint32_t* kin, const uint32_t* vin, uint32_t* ka, uint32_t* va, uint32_t* kb, uint32_t* vb,
                     int64_t n, int key_bits, uint32_t* hist, uint32_t* scan_temp, hipStream_t st,
                     const ProfileHook& prof, const uint32_t* plan = nullptr, bool few_passes = false);

// ---- incsort.hip
// plan words: reset the move counter and the overflow and set whether the
// snapshot is the previous step's (0 forces the full sort next).
void launch_inc_set(uint32_t* plan, uint32_t snap_ok, hipStream_t st);
// The incremental update (its launches return at once unless the plan
// chose it), in two halves around the radix sort: prepare (places among the
// snapshot, bucket offsets; may still send the step to the full sort), then
// merge (the snapshot, sorted by the previous keys, and the move list ->
// out_k/out_p).  n snapshot entries.
void launch_inc_prepare(uint32_t* plan, const MoveList& L, const uint32_t* snap_k, const uint32_t* snap_p, int64_t n,
                        hipStream_t st);
void launch_inc_merge(const uint32_t* plan, const MoveList& L, const uint32_t* snap_k, const uint32_t* snap_p,
                      int64_t n, uint32_t* out_k, uint32_t* out_p, hipStream_t st);
// Sizes of the update's buffers for n particles: snapshot tiles (the tile
// arrays hold 2 * (tiles + 1) words each) and the capacity of one sub-list
// (the list arrays hold kMoveLists times that).
uint32_t inc_tiles(int64_t n);
inline uint32_t inc_sublist_cap(int64_t n) {
    const int64_t cap = n / 8 > kMoveMin ? n / 8 : kMoveMin;
    return (uint32_t)((cap + kMoveLists - 1) / kMoveLists);
}

// ---- kernels.hip
// Persistent launch width of the item kernels.
constexpr int kItemMax = 512;   // particles per work item (max)

// key_offset is added to every key (the scene's offset in a batch).
void launch_keys(const SceneDev& S, const ParticlesDev& P, int interp, uint32_t* keys, hipStream_t st,
                 uint32_t key_offset = 0);
void launch_iota(uint32_t* p, int64_t n, hipStream_t st);
// Work items over sorted keys: items[0..n) receives the item starts and
// items[n..2n) their key blocks (so items holds 2n words).  item_max = 0
// picks a multiple of 64 up to kItemMax from n (about four items per compute
// unit for small n).  For a batch, key_offsets[nscenes] (ascending) gives
// each item's scene in item_scene.
void launch_items(const uint32_t* sorted_keys, int64_t n, uint32_t* flags, uint32_t* scan_temp, uint32_t* items,
                  uint32_t* item_count, hipStream_t st, uint32_t item_max = 0, const uint32_t* key_offsets = nullptr,
                  int nscenes = 0, uint32_t* item_scene = nullptr, uint32_t* plan = nullptr);

// grid_wgs: persistent workgroups; 0 = one resident round (occupancy of the
// kernel instance x compute units).
// mode: 0 = full P2G (mass, momentum, force), 1 = mass only (initialize,
// density raster), 2 = folded APIC grid (mass, p + dt f).
// With a scene table (tab != null, I.scene set) every item reads its scene
// from the table (modes 0 and 2); S and P are then unused.
void launch_p2g(const SceneDev& S, const ParticlesDev& P, const ItemsDev& I, int interp, int apic, int mode,
                int grid_wgs, hipStream_t st, const SceneEntry* tab = nullptr);
// src: 0 = fused grid update from P2G fields, 1 = materialised v / v*,
//      2 = initialize() volume gather, 3 = fused update from the folded grid
// next_keys (may be null): with advect, the cell key of each particle's new
// position is written at its storage slot (input of the next sort); in a
// batch at the scene's global slot, plus its key offset.
// wg: threads per workgroup of the fused G2P (64, 128 or 256).
void launch_g2p(const SceneDev& S, const ParticlesDev& P, const ItemsDev& I, int interp, int apic, int src,
                int advect, int pcollide, int grid_wgs, hipStream_t st, uint32_t* next_keys = nullptr,
                const SceneEntry* tab = nullptr, int wg = 64);

// gridToParticles of this step fused with particlesToGrid of the next (APIC,
// folded grid, one scene, quadratic stencil): gathers from S.grid and
// scatters into ogrid / otouched (a second grid of the same layout, clear).
void launch_g2p2g(const SceneDev& S, const ParticlesDev& P, const ItemsDev& I, int pcollide, float* ogrid,
                  uint32_t* otouched, uint32_t* next_keys, hipStream_t st);
// The same over a batch's items (scenes from the table, each scattering into
// its entry's grid2); next_keys are the global keys at the global slots.
void launch_g2p2g_batch(const SceneEntry* tab, const ItemsDev& I, int grid_wgs, uint32_t* next_keys,
                        hipStream_t st);
// After a fused batch step: each entry's grid2 becomes its grid and the
// reverse; every entry is marked pre-scattered.
void launch_swap_grids(SceneEntry* tab, int nscenes, hipStream_t st);

void launch_clear(const SceneDev& S, int64_t storage_blocks, int nfields, unsigned long long* active_nodes,
                  hipStream_t st);
// Clear of every scene's touched blocks; block_start[nscenes + 1] is the
// prefix sum of the scenes' storage block counts (on the device).
// mode 0: every grid (counted); 1: the grids of entries not pre-scattered
// (counted); 2: every grid2 (counted where the entry is pre-scattered).
void launch_clear_batch(const SceneEntry* tab, const int64_t* block_start, int nscenes, int64_t total_blocks,
                        int nfields, hipStream_t st, int mode = 0);
void launch_count_active(const SceneDev& S, int64_t storage_blocks, unsigned long long* active_nodes,
                         hipStream_t st);
// op: 0 gravity, 1 update velocities, 2 grid collision (touched blocks only)
void launch_grid_op(const SceneDev& S, int64_t storage_blocks, int op, float dt, hipStream_t st);
void launch_stress(const ParticlesDev& P, hipStream_t st);
void launch_particles_collision(const SceneDev& S, const ParticlesDev& P, float dt, hipStream_t st);
void launch_advect(const ParticlesDev& P, float dt, hipStream_t st);

// Host-layout (interleaved, original order) <-> device SoA.
struct HostLayoutDev {
    float *volume0, *mass, *pos, *vel, *Fe, *Fp, *mu0, *lam0, *B;
    uint16_t* mat;
};
void launch_pack(const HostLayoutDev& H, const ParticlesDev& P, hipStream_t st);
void launch_unpack(const ParticlesDev& P, const HostLayoutDev& H, hipStream_t st);
void launch_positions4(const ParticlesDev& P, float* out, hipStream_t st);
void launch_permute(const ParticlesDev& src, const ParticlesDev& dst, const uint32_t* perm, hipStream_t st);
void launch_gather_u32(const uint32_t* src, const uint32_t* idx, uint32_t* dst, int64_t n, hipStream_t st);
// In-place permutation of the particle store into sorted order (slot i <-
// slot perm[i]) through scratch (9 * ntot words).  With a scene table the
// slots are global (member s owns [pstart_s, pstart_s + n_s)), P is unused
// and slot_sc (ntot words) receives each slot's scene; otherwise P is the
// one store, ntot = P.n and slot_sc is unused.
void launch_reorder_store(const ParticlesDev& P, const SceneEntry* tab, int nscenes, const uint32_t* perm,
                          int64_t ntot, uint32_t* scratch, bool apic, hipStream_t st, uint32_t* slot_sc = nullptr);
// A batch's reorder of one chunk of members per step: the members of chunk
// *rr (global positions [bounds[*rr], bounds[*rr + 1]), at most chunk_max)
// permuted into the step's sorted order (perm, set to the identity there),
// then *rr advanced modulo nchunks.  scratch: kStoreWords * chunk_max words
// plus chunk_max 16-bit words; slot_sc: each global slot's scene.
constexpr int kChunkScratchWords = 46;  // per chunk position (45 words + the 16-bit index, rounded up)
void launch_reorder_chunk(const SceneEntry* tab, const uint32_t* slot_sc, uint32_t* perm, const int64_t* bounds,
                          uint32_t* rr, int nchunks, int64_t chunk_max, uint32_t* scratch, bool apic, hipStream_t st);
void launch_slot_scenes(const SceneEntry* tab, int nscenes, int64_t ntot, uint32_t* slot_sc, hipStream_t st);
void launch_grid_dense(const SceneDev& S, float* mass, float* vel, float* vstar, float* force, int have_vstar,
                       hipStream_t st);
// Density raster readout from a one-plane mass grid D (D.grid = sum m w):
// dense in node order x + rx (y + ry z), or the touched storage blocks
// compacted by rank (exclusive scan of launch_touched_flags' output).
void launch_density_dense(const SceneDev& D, float* out, hipStream_t st);
void launch_touched_flags(const uint32_t* touched, int64_t n, uint32_t* flags, hipStream_t st);
void launch_density_blocks(const SceneDev& D, const uint32_t* touched, const uint32_t* rank, int64_t nsb,
                           int32_t* origin, float* values, int64_t max_blocks, hipStream_t st);
// Largest |v|^2 of the store into *out (float bits; *out must start at 0).
void launch_max_speed2(const ParticlesDev& P, unsigned int* out, hipStream_t st);
// Per-scene health of a batch (3 words per scene, zeroed by the caller):
// largest |v|^2 as float bits, non-finite particles, escaped particles.
void launch_batch_health(const SceneEntry* tab, const uint32_t* key_offsets, int ns, int64_t ntot,
                         const uint32_t* keys, uint32_t* out, hipStream_t st);
void launch_count_escaped(const uint32_t* sorted_keys, int64_t n, uint32_t escape_key, unsigned long long* out,
                          hipStream_t st);
void launch_debug_svd(const float* A, int64_t n, float* U, float* s, float* V, hipStream_t st);
void launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t st);

}  // namespace mpm
