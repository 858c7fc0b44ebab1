/*
 * mpm_hip.h -- C-ABI of the MI355X (gfx950) explicit MPM snow solver.
 *
 * Plain C types only (no torch, no HIP types): every entry point takes an
 * opaque handle, plain pointers and sizes, and returns an int status
 * (MPM_OK == 0, negative on error; the message is in mpm_last_error()).
 * Nothing in the library calls exit().
 *
 * The entry points replace the reference's C++ solver surface one for one:
 *   reference include/mpm_solver3d.h:25-74      (class MPMSolver3D)
 *   reference include/apic_mpm_solver3d.h:24-71 (class APICMPMSolver3D)
 * Both reference classes map onto one handle type; mpm_desc.transfer picks
 * PIC-FLIP or APIC.  The C++ drop-in facade (namespace chains, same class,
 * method and builder names as the reference) forwards to these functions.
 *
 * Threading: a handle owns its device memory and one HIP stream; calls on
 * one handle must not overlap.  Different handles may be used from
 * different host threads.
 */
#ifndef MPM_HIP_H_
#define MPM_HIP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPM_API __attribute__((visibility("default")))

/* ---- status codes (the reference exit()s instead: include/cuda_utils.h:6-20) */
enum {
    MPM_OK = 0,
    MPM_ERR_INVALID = -1,     /* bad argument / shape / enum value          */
    MPM_ERR_HIP = -2,         /* HIP runtime error (message has details)    */
    MPM_ERR_OOM = -3,         /* device allocation failed / budget exceeded */
    MPM_ERR_UNSUPPORTED = -4, /* combination the solver rejects             */
    MPM_ERR_STATE = -5,       /* call not valid in the current state        */
    MPM_ERR_NOT_FOUND = -6,   /* unknown scene id                           */
    MPM_ERR_IO = -7           /* file could not be written                  */
};

/* ---- enums: same values and order as reference include/enums.h:6-15 */
typedef enum mpm_interp {
    MPM_INTERP_LINEAR = 0,
    MPM_INTERP_QUADRATIC_BSPLINE = 1,
    MPM_INTERP_CUBIC_BSPLINE = 2
} mpm_interp;

typedef enum mpm_model {
    MPM_MODEL_FIXED_COROTATED = 0,
    MPM_MODEL_NEO_HOOKEAN = 1
} mpm_model;

/* MPMSolver3D (PIC-FLIP) vs APICMPMSolver3D: the reference has two classes. */
typedef enum mpm_transfer {
    MPM_TRANSFER_PIC_FLIP = 0,
    MPM_TRANSFER_APIC = 1
} mpm_transfer;

/* Solver description: the reference constructor arguments
 * (include/mpm_solver3d.h:25-34, include/apic_mpm_solver3d.h:24-32)
 * plus the gravity magnitude the reference hard-codes to 9.80665 on -y
 * (src/mpm_solver3d.cu:397). */
typedef struct mpm_desc {
    float origin[3];            /* gridOrigin                                  */
    int32_t resolution[3];      /* gridResolution (nodes per axis)             */
    float stride;               /* gridStride h                                */
    float friction;             /* gridBoundaryFrictionCoefficient             */
    float blend;                /* blendCoefficient (FLIP weight), PIC-FLIP    */
    int32_t interp;             /* mpm_interp                                  */
    int32_t transfer;           /* mpm_transfer                                */
    float dt;                   /* deltaTime                                   */
    float gravity;              /* magnitude along -y; 9.80665f = reference    */
    int32_t particle_collision; /* initial _enable_particles_collision (0)     */
    int32_t reserved[4];        /* must be zero                                */
} mpm_desc;

/* Constructor arguments of MaterialPoint3D / APICMaterialPoint3D, one entry
 * per particle (include/material_point3d.h:27-37).  mu0 and lambda0 are
 * derived as the reference constructor does (src/material_point3d.cu:25-26),
 * Fe = Fp = I, B = 0. */
typedef struct mpm_particle_input {
    int64_t count;
    const float* position;             /* [3*count] x,y,z interleaved         */
    const float* velocity;             /* [3*count]                           */
    const float* mass;                 /* [count]                             */
    const float* hardening;            /* [count] xi                          */
    const float* young;                /* [count] E                           */
    const float* poisson;              /* [count] nu                          */
    const float* critical_compression; /* [count] theta_c                     */
    const float* critical_stretch;     /* [count] theta_s                     */
    const int32_t* model;              /* [count] mpm_model                   */
} mpm_particle_input;

/* Full per-particle state: the fields of MaterialPoint3D
 * (include/material_point3d.h:15-21) plus APIC B
 * (include/apic_material_point3d.h:23).  Matrices are row-major 3x3.
 * Arrays are in the caller's (original) particle order.  Any pointer may be
 * NULL on read to skip that field.  On write (mpm_create_state /
 * mpm_write_particles) every pointer except B must be set; B may be NULL
 * (then B = 0). */
typedef struct mpm_particle_state {
    int64_t count;
    float* volume0;              /* [count]   */
    float* mass;                 /* [count]   */
    float* position;             /* [3*count] */
    float* velocity;             /* [3*count] */
    float* Fe;                   /* [9*count] deformation_gradient_elastic */
    float* Fp;                   /* [9*count] deformation_gradient_plastic */
    float* mu0;                  /* [count]   */
    float* lambda0;              /* [count]   */
    float* hardening;            /* [count]   */
    float* critical_compression; /* [count]   */
    float* critical_stretch;     /* [count]   */
    int32_t* model;              /* [count]   */
    float* B;                    /* [9*count] APIC only                    */
} mpm_particle_state;

/* Dense grid in the reference's node order, linear index
 * i = x + rx*(y + ry*z) (src/mpm_solver3d.cu:158); fields of
 * CollocatedGridData3D (include/grid3d.h:28-31).  NULL pointers skipped. */
typedef struct mpm_grid_host {
    int64_t nodes;        /* must equal rx*ry*rz */
    float* mass;          /* [nodes]   */
    float* velocity;      /* [3*nodes] xyz interleaved per node */
    float* velocity_star; /* [3*nodes] */
    float* force;         /* [3*nodes] */
} mpm_grid_host;

typedef struct mpm_stats {
    int64_t particles;          /* N                                         */
    int64_t steps;              /* completed simulateOneStep calls           */
    int64_t active_nodes;       /* nodes with mass > 0 in the current grid   */
    int64_t active_nodes_sum;   /* sum of active_nodes over all steps        */
    int64_t work_items;         /* block work items of the last sort         */
    int64_t escaped;            /* particles whose stencil misses the grid   */
    int64_t device_bytes;       /* device memory owned by the handle         */
    double last_step_ms;        /* host wall time of the last mpm_step call  */
    /* CFL monitor (the reference steps with a fixed dt and has no check,
     * src/mpm_solver3d.cu:101): the largest particle speed |v| now, and
     * max|v| * dt / h.  NaN if a velocity is NaN. */
    double max_speed;
    double cfl;
    /* step graphs captured by mpm_step_graph since creation (a capture is
     * dropped when the particle collision switch, the grid record mode or
     * the material table changes) */
    int64_t graph_captures;
    /* steps whose cell sort updated the previous step's order instead of
     * sorting in full (scenes of at least 2^20 particles: few particles
     * change cell per step; the result is the same permutation) */
    int64_t incremental_sorts;
} mpm_stats;

/* Per-kernel timing (hipEvent pairs on the handle's stream), enabled by
 * mpm_profile_enable.  Durations are sums over launches since reset. */
typedef struct mpm_kernel_time {
    char name[32];
    int64_t launches;
    double total_ms;
} mpm_kernel_time;

/* ---------------------------------------------------------------- lifetime */

/* Build a solver from constructor-style particle input and run
 * initialize() (volume0 from grid density), as the reference constructor
 * does (src/mpm_solver3d.cu:36-91, src/apic_mpm_solver3d.cu:36-90). */
MPM_API int mpm_create(const mpm_desc* desc, const mpm_particle_input* particles,
                       void** out_handle);

/* Build a solver from a full particle state (checkpoint restore).  When
 * run_initialize != 0 volume0 is recomputed as in the reference
 * initialize(); otherwise the given volume0 is kept. */
MPM_API int mpm_create_state(const mpm_desc* desc, const mpm_particle_state* state,
                             int run_initialize, void** out_handle);

/* Reference destructor (src/mpm_solver3d.cu:93-98). */
MPM_API void mpm_destroy(void* handle);

/* Thread-local message for the last failing call on this host thread. */
MPM_API const char* mpm_last_error(void);
MPM_API const char* mpm_version(void);

/* ----------------------------------------------------------- device memory */
/* No counterpart in the reference (one scene per process, its buffers
 * freed by the destructor, src/mpm_solver3d.cu:93-98).  Solvers and
 * batches take device memory from a per-device cache of released blocks,
 * because hipFree waits for every stream of the device: a scene built or
 * destroyed while batches run would otherwise stall behind all their
 * queued steps.  A released block stays cached (up to the limit, per
 * device; 16 GiB by default) for the next allocation of a similar size
 * (a block at most 25% larger than the request's size class).  Cached
 * bytes are not part of any handle's device_bytes.  Thread-safe. */
typedef struct mpm_cache_info {
    int64_t cached_bytes; /* released blocks kept, current device */
    int64_t live_bytes;   /* blocks held by solvers and batches, current device */
    int64_t hits;         /* allocations served from the cache (all devices) */
    int64_t misses;       /* allocations that called hipMalloc (all devices) */
} mpm_cache_info;
MPM_API int mpm_set_cache_limit(int64_t bytes); /* trims every device to the new limit */
MPM_API int mpm_get_cache_info(mpm_cache_info* out);
/* Hands every cached block back to the driver (hipFree: waits for the device). */
MPM_API int mpm_release_cache(void);

/* -------------------------------------------------------------- time steps */

/* simulateOneStep() repeated n times (src/mpm_solver3d.cu:100-125,
 * src/apic_mpm_solver3d.cu:99-124); fused device path.  Asynchronous with
 * respect to the host; mpm_synchronize waits. */
MPM_API int mpm_step(void* handle, int n);
MPM_API int mpm_synchronize(void* handle);

/* Capture one simulateOneStep into a hipGraph and replay it n times
 * (identical results to mpm_step). */
MPM_API int mpm_step_graph(void* handle, int n);

/* Sub-steps, one per reference public method (include/mpm_solver3d.h:40-63).
 * They run the unfused device kernels and leave the grid materialised so
 * that mpm_read_grid shows exactly what the reference grid holds. */
MPM_API int mpm_initialize(void* handle);                         /* :42 */
MPM_API int mpm_reset_grid(void* handle);                         /* :44 */
MPM_API int mpm_particles_to_grid(void* handle);                  /* :46 */
MPM_API int mpm_compute_external_forces(void* handle);            /* :48 */
MPM_API int mpm_compute_gravity_forces(void* handle);             /* :59 */
MPM_API int mpm_update_grid_velocities(void* handle, float dt);   /* :51 */
MPM_API int mpm_solve_linear_system(void* handle, float dt);      /* :53 (no-op) */
MPM_API int mpm_grid_collision(void* handle, float dt);           /* :61 */
MPM_API int mpm_grid_to_particles(void* handle, float dt, float blend); /* :55 */
MPM_API int mpm_particles_collision(void* handle, float dt);      /* :63 */
MPM_API int mpm_advect_particles(void* handle, float dt);         /* :57 */
MPM_API int mpm_switch_particles_collision(void* handle);         /* :40 */
MPM_API int mpm_set_particles_collision(void* handle, int enable); /* field :23 */
/* Grid recording for APIC.  The fused APIC step only needs v*, so by
 * default its P2G sums momentum and dt*force into one set of planes
 * (mass + 3 instead of mass + 6); mpm_read_grid, the grid sub-steps and
 * mpm_grid_to_particles then return MPM_ERR_STATE until the next P2G or
 * reset.  enable = 1 keeps momentum and force apart (the reference's grid
 * after simulateOneStep, src/apic_mpm_solver3d.cu:99-124) at the cost of
 * the wider P2G.  No effect on PIC-FLIP, whose G2P needs v and v*. */
MPM_API int mpm_set_grid_record(void* handle, int enable);
/* Store reorder (no counterpart in the reference, whose particle array is
 * never reordered): every `steps` steps (default 100; 0 = never) the
 * particle store is permuted in place into the cell order of the current
 * positions, so that the per-step reads and writes stay contiguous as
 * particles move (on the 8.4M slab G2P slows from 0.88 to 1.19 ms over
 * 1000 steps without it).  Results do not depend on the interval beyond
 * float-atomic order; reads return the caller's particle order. */
MPM_API int mpm_set_reorder_interval(void* handle, int steps);
/* Transfer fusion (no counterpart in the reference, which runs
 * gridToParticles and the next step's particlesToGrid as separate passes
 * with a fresh grid between them, src/apic_mpm_solver3d.cu:99-124): for an
 * APIC scene with the quadratic stencil stepped alone (not in a
 * batch, not under mpm_step_graph, grid record off), one kernel gathers
 * step n and, with each particle still in registers, scatters step n+1's
 * particlesToGrid into a second grid buffer (one more grid of device
 * memory).  The sums are the same up to float order.  Any other call that
 * touches particles or grid first restores the two-pass state.  1 (the
 * default) enables it, 0 disables it. */
MPM_API int mpm_set_transfer_fusion(void* handle, int enable);

/* Particle-by-cell key generation + stable LSD radix sort (no reference
 * counterpart: the reference scatters in array order,
 * src/mpm_solver3d.cu:292).  Runs at the start of every step; exposed for
 * parity tests. */
MPM_API int mpm_sort(void* handle);

/* ------------------------------------------------------------- state I/O */

/* Copy particle state to the host in the original particle order (the
 * reference never reorders: src/mpm_solver3d_builder.cu:38-65). */
MPM_API int mpm_read_particles(void* handle, mpm_particle_state* out);
/* The same readback without allocating device memory: no counterpart in the
 * reference.  For callers that cannot call hipMalloc (code under graph
 * capture, tight loops and harnesses that forbid allocation).  The device
 * staging is a file-scope array of the
 * library (39.3 MB per GPU, held until the process ends) that every call on
 * that GPU shares: at most 262,144 particles (larger handles get
 * MPM_ERR_UNSUPPORTED; use mpm_read_particles).  Calls from several host
 * threads on one GPU would race on the array, so the library serialises
 * them; the array keeps the last call's contents. */
MPM_API int mpm_read_particles_noalloc(void* handle, mpm_particle_state* out);
/* Overwrite the particle state (same count as at creation). */
MPM_API int mpm_write_particles(void* handle, const mpm_particle_state* in);
/* Copy the dense grid (reference node order) to the host. */
MPM_API int mpm_read_grid(void* handle, mpm_grid_host* out);
/* Device positions packed as float4 (x,y,z,1) in original order into a
 * host buffer of 4*count floats (reference updateGLBufferByCPU,
 * src/mpm_solver3d.cu:484-508, without the GL buffer). */
MPM_API int mpm_read_positions4(void* handle, float* out_xyz1);
/* The same pack into a caller-owned device buffer of 4*count floats (for
 * example a vertex buffer mapped for interop), asynchronously on the
 * solver's stream: updateGLBufferWithCUDA (src/mpm_solver3d.cu:457-482)
 * without the GL registration. */
MPM_API int mpm_pack_positions4(void* handle, float* device_out);
/* Frame export that overlaps later steps: the pack runs on the solver's
 * stream into a device staging buffer and the device-to-host copy into
 * host_out (4*count floats) on a second stream.  host_out must stay valid
 * until mpm_positions4_wait returns; page-locked memory (hipHostMalloc)
 * makes the copy asynchronous.  A second export waits on the device for
 * the first copy before reusing the staging buffer. */
MPM_API int mpm_positions4_async(void* handle, float* host_out);
MPM_API int mpm_positions4_wait(void* handle);

/* ---------------------------------------------------- density raster
 * Replacement of writeToOpenVDB (include/mpm_solver3d.h:73,
 * src/mpm_solver3d.cu:510-572), which rasterises on the host into an
 * OpenVDB FloatGrid named "density".  Here the raster runs on the device:
 * density(node) = sum over particles of mass * w / h^3, over the current
 * positions, the stencil range of P2G (origin ignored, as in the
 * reference) and the nodes inside [0, res); nodes no particle reaches hold
 * 0, the reference grid's background value.  It is recomputed only when
 * the particles changed since the last raster; the reads synchronise. */
MPM_API int mpm_density_compute(void* handle);  /* asynchronous */
/* Dense copy: res.x*res.y*res.z floats, node (x,y,z) at x + rx*(y + ry*z). */
MPM_API int mpm_read_density(void* handle, float* out);
/* Sparse copy: the 4x4x4 node blocks the raster touched, in storage order.
 * *count receives their number; with origins and values both NULL nothing
 * else is written.  Otherwise block b's first node goes to origins[3b..3b+2]
 * and its 64 densities to values[64b..64b+63], node origin + (i,j,k) at
 * i + 4j + 16k (nodes beyond res hold 0); MPM_ERR_INVALID if
 * max_blocks < *count. */
MPM_API int mpm_read_density_blocks(void* handle, int32_t* origins, float* values, int64_t max_blocks,
                                    int64_t* count);
/* Write the sparse blocks to a file (OpenVDB is not used), little-endian:
 *   0  char[8]  "MPMDENS1"
 *   8  int32[3] resolution
 *  20  float[3] grid origin
 *  32  float    stride h
 *  36  int32    block edge (4)
 *  40  int64    block count B
 *  48  B records { int32 first node[3]; float density[64] } as above.
 * Node (x,y,z) sits at origin + h*(x,y,z); MPM_ERR_IO on a write error. */
#define MPM_DENSITY_MAGIC "MPMDENS1"
MPM_API int mpm_write_density(void* handle, const char* path);

/* Last sort: keys[i] and original particle id ids[i] of sorted position i;
 * seq_ids[j] = original id of the particle at input position j of that sort
 * (the sequence the stable sort was applied to).  Any pointer may be NULL. */
MPM_API int mpm_read_sort(void* handle, uint32_t* keys, uint32_t* ids, uint32_t* seq_ids, int64_t count);
/* The key function used by the sort, evaluated on the host for n positions
 * (x,y,z interleaved) under this handle's grid and interpolation. */
MPM_API int mpm_host_keys(void* handle, const float* positions, int64_t n, uint32_t* keys);

MPM_API int mpm_get_stats(void* handle, mpm_stats* out);

/* ------------------------------------------------------------- profiling */
/* enable: 0 off; 1 every kernel group (two events around each launch group:
 * about 30 per step, which add ~0.1 ms of stream gaps per step); 2 only the
 * item kernels p2g and g2p, or the fused g2p2g (2-4 events per step), and
 * the store reorder. */
MPM_API int mpm_profile_enable(void* handle, int enable);
MPM_API int mpm_profile_reset(void* handle);
/* Fills up to max entries; *count gets the number of kernels tracked. */
MPM_API int mpm_profile_read(void* handle, mpm_kernel_time* out, int max, int* count);

/* ---------------------------------------------------------------- batches */
/* No counterpart in the reference, which runs one scene per solver object
 * (src/mpm_solver3d.cu:77-108 simulateOneStep).  A batch advances many scene
 * handles of one interpolation and transfer scheme with one set of launches
 * per step: a batched grid clear, one radix sort over all scenes' keys
 * (scene-offset key ranges), and the tiled P2G/G2P kernels reading each work
 * item's scene from a device table.  Per scene the result equals that
 * scene's own mpm_step up to the summation order of float atomics.
 *
 * While a member, a scene handle runs on the batch's stream: its own calls
 * (reads, sub-steps, mpm_step) stay valid and are ordered with the batch's
 * steps.  mpm_destroy of a member removes it from the batch first.        */

/* Launch parameters of a batch step. */
typedef struct mpm_tune {
    int32_t item_max;     /* particles per work item: multiple of 64 in [64, 65536] */
    int32_t g2p_threads;  /* threads per G2P workgroup: 64, 128 or 256             */
    int32_t workgroups;   /* persistent workgroups of P2G/G2P; 0 = one resident
                             round per kernel (its occupancy x CUs)              */
    int32_t bucket;       /* out: size bucket, floor(log2(particles))              */
    double step_ms;       /* out (autotune): measured ms per step of the choice    */
    int32_t fuse_transfer; /* 1: G2P fused with the next step's P2G (APIC batches
                              with the quadratic stencil whose members do not
                              record their grids; ignored otherwise): one kernel
                              scatters each member's next P2G into a second grid
                              buffer of the member (one more grid of device
                              memory), then the batch table's two grids of every
                              member are swapped.  0 by default: it pays when few
                              particles change cell per step (see the autotuner) */
} mpm_tune;

typedef struct mpm_batch_stats {
    int64_t scenes;
    int64_t particles;
    int64_t steps;           /* batch steps taken                        */
    int64_t storage_blocks;  /* 4x4x4 grid blocks over all scenes        */
    int64_t key_bits;        /* bits of the batch key space (sort width) */
    int64_t work_items;      /* work items of the last step              */
    int64_t device_bytes;    /* batch-owned device memory                */
    double last_step_ms;     /* host time of the last step call (enqueue) */
    mpm_tune tune;           /* parameters in use                        */
    int64_t incremental_sorts; /* steps sorted by the incremental update */
} mpm_batch_stats;

MPM_API int mpm_batch_create(int32_t interp, int32_t transfer, void** out_batch);
/* Removes the members (they stay valid scene handles) and frees the batch. */
MPM_API void mpm_batch_destroy(void* batch);
/* MPM_ERR_INVALID if the scene's interp/transfer differ from the batch's,
 * MPM_ERR_STATE if it is already in a batch, MPM_ERR_UNSUPPORTED if the
 * members' cell keys would exceed the 32-bit batch key space (each scene
 * takes (((rx+3)>>2)+1)*(((ry+3)>>2)+1)*(((rz+3)>>2)+1)*64 + 64 keys). */
MPM_API int mpm_batch_add(void* batch, void* scene);
MPM_API int mpm_batch_remove(void* batch, void* scene);
MPM_API int mpm_batch_info(void* batch, mpm_batch_stats* out);
/* n steps of every member (asynchronous). */
MPM_API int mpm_batch_step(void* batch, int n);
/* Same, replaying one step captured as a hipGraph (recaptured after a
 * membership or tuning change). */
MPM_API int mpm_batch_step_graph(void* batch, int n);
/* The batch's store reorder interval (as mpm_set_reorder_interval, for
 * every member; default 16, 0 = never).  Scenes that move as a whole lose
 * their cell order quickly: the 256-scene uneven batch steps in 2.5 ms
 * right after a reorder and 3.9 ms 100 steps later, and one reorder costs
 * ~3.3 ms, so 16 steps (3.0 ms per step on average) beats 100 (3.5 ms). */
MPM_API int mpm_batch_set_reorder_interval(void* batch, int steps);
/* The batch's cell sort updates the last step's order (the particles that
 * changed cell taken out and put back) instead of sorting in full, when
 * the list of moves fits (up to max(16384, particles / 8)); bit-identical
 * to the full sort.  On by default; 0 sorts in full every step.  The
 * single-scene counterpart is automatic (scenes of 2^20 particles and up).
 * No reference counterpart: the reference sorts nothing, its P2G adds each
 * particle's stencil with global atomics (src/mpm_solver3d.cu:254-285). */
MPM_API int mpm_batch_set_incremental_sort(void* batch, int enable);
/* The last batch step's sorted order: keys[i] (scene key offset included)
 * and the global slot (member stores concatenated in membership order) of
 * sorted position i, after that step's store reorder of a chunk of members
 * renumbered their slots.  count must equal the batch's particle count;
 * either pointer may be NULL.  No reference counterpart (the reference
 * sorts nothing); for tests and debugging. */
MPM_API int mpm_batch_read_sort(void* batch, uint32_t* keys, uint32_t* slots, int64_t count);
MPM_API int mpm_batch_synchronize(void* batch);
/* Fixes the launch parameters; NULL returns to the tuning table. */
MPM_API int mpm_batch_set_tune(void* batch, const mpm_tune* tune);
/* Times candidate parameters on the batch's own scenes (every trial step
 * is a real step of every member: 1 + trial_steps per candidate, at most 9
 * candidates: work-item size, G2P workgroup, persistent workgroups, then
 * the fused transfer where the batch allows it), keeps the fastest for the
 * batch and records it in the
 * process-wide table for this (interp, transfer, size bucket). */
MPM_API int mpm_batch_autotune(void* batch, int trial_steps, mpm_tune* out);
MPM_API int mpm_tune_lookup(int32_t interp, int32_t transfer, int64_t particles, mpm_tune* out);
/* Per-step device time (hipEvents on the batch stream) while enabled. */
MPM_API int mpm_batch_timing(void* batch, int enable);
/* Copies up to max recorded step times (ms) and clears the record;
 * with ms == NULL only *count is set. */
MPM_API int mpm_batch_read_times(void* batch, float* ms, int max, int* count);

/* Health of each member after its steps, in membership order (one pass
 * over all members' particles): the largest particle speed |v| and
 * |v| * dt / h (NaN when a velocity is NaN), the particles with a non-finite
 * position or velocity, and the particles outside their grid (stencil off
 * the grid: the escape key).  n must equal the member count.  No
 * counterpart in the reference, which neither batches scenes nor checks
 * its state (src/mpm_solver3d.cu:100-125); a request server uses it to
 * evict a failing scene and keep the others running. */
typedef struct mpm_scene_health {
    double max_speed;
    double cfl;
    int64_t nonfinite;
    int64_t escaped;
} mpm_scene_health;
MPM_API int mpm_batch_health(void* batch, mpm_scene_health* out, int n);

/* ------------------------------------------------------------- test hooks */
/* Device 3x3 SVD (the solver's own) on n row-major matrices:
 * A = U diag(s) V^T, U,V rotations, s[0] >= s[1] >= |s[2]|. */
MPM_API int mpm_debug_svd3(const float* A, int64_t n, float* U, float* s, float* V);

#ifdef __cplusplus
}
#endif

#endif /* MPM_HIP_H_ */
