/*
 * sph_amd.h -- C ABI of the MI355X (gfx950) adaptive SPH step.
 *
 * Drop-in boundary for the device side of finallyjustice/sca2012: every entry point below
 * replaces one of the C++-linkage device functions declared in the reference's
 * sph_system.cuh (SPH-3D-2/sph_system.cuh:11-90, SPH-3D-4/sph_system.cuh:11-90,
 * SPH-2D/sph_system.cuh:12-71), with the reference's semantics, but with explicit state:
 *   - no process-global __constant__ SysParam (SPH-3D-2/sph_kernel.cu:10): parameters
 *     live in the context, so several scenes can coexist;
 *   - plain pointers and sizes only; every call returns an int status;
 *   - particle records cross the boundary in the reference's AoS layout (Particle, 72 B in
 *     3D / 56 B in 2D); on the device the state is SoA and kept in cell order.
 * Status codes: 0 ok, SPH_EINVAL bad argument / phase order, SPH_EEMPTY no particles
 * (the reference divides by zero in compute_grid_size there, SPH-3D-2/sph_system.cu:72-76),
 * SPH_EDOMAIN a particle outside the cell grid (an out-of-range cell_start write in the
 * reference), SPH_ECAP capacity, SPH_EHIP a HIP runtime error, SPH_ETRI triangle buffer
 * too small.
 */
#ifndef SPH_AMD_H
#define SPH_AMD_H

#include <stddef.h>
#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ---- reference memory layouts (byte-identical; static_assert'ed in the library) ---- */
typedef struct { float x, y, z; } sph_float3;
typedef struct { float x, y; } __attribute__((aligned(8))) sph_float2;
typedef struct { uint32_t x, y; } __attribute__((aligned(8))) sph_uint2;

/* SPH-3D-2/sph_object.h:4-9 */
typedef struct {
  sph_float3 pos;
  float radius;
  float damping;
} SphSphere;

/* SPH-3D-2/sph_param.h:8-87 and SPH-3D-4/sph_param.h:8-88 (296 bytes). The two variants
 * differ only in the trailing bools, so the record is declared three times over the same
 * leading fields: SphSysParam3D2 and SphSysParam3D4 with the reference's names (what the C++
 * drop-in hands to an application as SysParam), and SphSysParam3D with neutral names for
 * code that handles either variant. All three have the same layout; the C ABI takes any of
 * them through a void pointer. */
#define SPH_SYSPARAM3D_FIELDS                              \
  uint32_t num_particle, max_particle;                     \
  float world_width, world_height, world_length;           \
  float large_radius, small_radius;                        \
  float large_kernel, small_kernel;                        \
  float large_mass, small_mass;                            \
  float cell_size;                                         \
  uint32_t row_cell, col_cell, len_cell, total_cell;       \
  sph_float3 gravity;                                      \
  float wall_damping, rest_density, gas_constant, time_step, viscosity, surface_tension; \
  float co_surface, co_energy;                             \
  float split_criteria, merge_criteria;                    \
  float large_poly6, large_spiky, large_visco;             \
  float large_grad_poly6, large_lapc_poly6;                \
  float small_poly6, small_spiky, small_visco;             \
  float small_grad_poly6, small_lapc_poly6;                \
  float large_kernel_2, large_kernel_6;                    \
  float small_kernel_2, small_kernel_6;                    \
  float large_poly6_radius, small_poly6_radius;            \
  uint32_t dens_mul;                                       \
  float den_size;                                          \
  uint32_t row_dens, col_dens, len_dens, tot_dens;         \
  SphSphere sphere1, sphere2, sphere3, sphere4;            \
  float split_energy, merge_energy;

typedef struct {
  SPH_SYSPARAM3D_FIELDS
  bool flag0; /* 3D-2 force_waiting      | 3D-4 force_waiting_left  */
  bool flag1; /* 3D-2 use_cylinder       | 3D-4 force_waiting_right */
  bool flag2; /* 3D-2 (padding)          | 3D-4 use_cylinder        */
} SphSysParam3D;

/* SPH-3D-2/sph_param.h:8-87 */
typedef struct {
  SPH_SYSPARAM3D_FIELDS
  bool force_waiting; /* SPH-3D-2/sph_param.h:85 */
  bool use_cylinder;  /* SPH-3D-2/sph_param.h:86 */
} SphSysParam3D2;

/* SPH-3D-4/sph_param.h:8-88 */
typedef struct {
  SPH_SYSPARAM3D_FIELDS
  bool force_waiting_left;  /* SPH-3D-4/sph_param.h:85 */
  bool force_waiting_right; /* SPH-3D-4/sph_param.h:86 */
  bool use_cylinder;        /* SPH-3D-4/sph_param.h:87 */
} SphSysParam3D4;

/* SPH-3D-2/sph_system.h:8-24 (72 bytes) */
typedef struct {
  sph_float3 pos, vel, ev, acc;
  float dens, press, energy, surface, tension;
  uint32_t level; /* 1 = large, 2 = small */
} SphParticle3D;

/* SPH-2D/sph_param.h:6-72 (200 bytes) */
typedef struct {
  uint32_t num_particle, max_particle;
  float world_width, world_depth;
  float large_mass, small_mass;
  float large_kernel, small_kernel;
  float large_kernel2, small_kernel2;
  float large_kernel6, small_kernel6;
  float large_radius, small_radius;
  float large_poly6, small_poly6;
  float large_spiky, small_spiky;
  float large_visco, small_visco;
  float large_grad_poly6, small_grad_poly6;
  float large_lapc_poly6, small_lapc_poly6;
  sph_float2 gravity;
  float wall_damping, rest_density, gas_constant, viscosity, time_step, surface_normal;
  float cell_size;
  uint32_t row_cell, col_cell, total_cell;
  float co_surface, co_energy;
  float thresh_split, thresh_merge;
  bool force_waiting;
  sph_float2 extern_force, extern_click;
  sph_uint2 extern_pos;
  float extern_len, extern_coe;
} SphSysParam2D;

/* SPH-2D/sph_system.h:7-22 (56 bytes) */
typedef struct {
  sph_float2 pos, vel, ev, acc;
  float dens, pres, energy;
  uint32_t level;
  float surface_normal;
} SphParticle2D;

/* ---- constants ---- */
#define SPH_V2D 0  /* SPH-2D   */
#define SPH_V3D2 1 /* SPH-3D-2 */
#define SPH_V3D4 2 /* SPH-3D-4 */

#define SPH_STEP_SPLIT 1u /* m_use_split (SPH-3D-2/sph_system.cpp:373)   */
#define SPH_STEP_MERGE 2u /* m_use_merge (SPH-3D-2/sph_system.cpp:386)   */
#define SPH_STEP_MC 4u    /* m_disp_mode == 2 (SPH-3D-2/sph_system.cpp:336) */

#define SPH_OK 0
#define SPH_EINVAL (-1)
#define SPH_EEMPTY (-2)
#define SPH_EDOMAIN (-3)
#define SPH_ECAP (-4)
#define SPH_EHIP (-5)
#define SPH_ETRI (-6)

/* debug accessors (sph_get_u32 / sph_get_f32) */
#define SPH_Q_HASH 0      /* sorted hash (N), as dev_hash after sort_particles       */
#define SPH_Q_INDEX 1     /* sorted slot index (N), as dev_index after sort_particles */
#define SPH_Q_START 2     /* cell_start (C)                                          */
#define SPH_Q_END 3       /* cell_end (C)                                            */
#define SPH_Q_SPLITFLAG 4 /* dev_split per slot before the sort, 0 = split (N)       */
#define SPH_Q_STATUS 5    /* dev_status per slot after merge, 1 = removed (N')       */
#define SPH_Q_COUNTERS 6  /* device counters: n m n_pre s_req s_done groups err ntri ... (16) */
#define SPH_Q_DGRID 10    /* density grid (G)                                        */
#define SPH_Q_NORMAL 11   /* normals, 3 floats per voxel (3G)                        */
#define SPH_Q_TRI 12      /* triangles: p0 p1 p2 n0 n1 n2, 18 floats each            */

/* profiling phases (sph_prof_read) */
#define SPH_P_HASH 0
#define SPH_P_SORT 1
#define SPH_P_CELLS 2
#define SPH_P_REORDER 3
#define SPH_P_DENSITY 4
#define SPH_P_FORCE 5
#define SPH_P_INTEGRATE 6
#define SPH_P_ENERGY 7
#define SPH_P_SPLIT 8
#define SPH_P_MERGE 9
#define SPH_P_COMPACT 10
#define SPH_P_DGRID 11
#define SPH_P_NORMALS 12
#define SPH_P_MC 13
#define SPH_P_COUNT 14

typedef struct SphCtx SphCtx;

/* ---- parameters ---- */
/* Fills a SysParam with the constants of SPHSystem::SPHSystem()
 * (SPH-3D-2/sph_system.cpp:6-229, SPH-3D-4/sph_system.cpp:6-230, SPH-2D/sph_system.cpp:5-146). */
int sph_param_defaults(int variant, void *sysparam_out);
/* Recomputes the derived fields (cell grid snapped to whole cells, kernel constants with
 * PI = 3.141592f, kernel powers, density-grid sizes) from the base fields, in the order
 * of the reference constructor (SPH-3D-2/sph_system.cpp:24-32, :63-92). */
int sph_param_derive(int variant, void *sysparam_inout);

/* ---- context ---- */
/* Allocates every buffer at `capacity` particles (SPHSystem ctor, SPH-3D-2/sph_system.cpp:34-47,
 * :94-101 via alloc_array, sph_system.cu:39-42); max_triangles sizes the marching-cubes output. */
int sph_ctx_create(int variant, const void *sysparam, uint32_t capacity, uint32_t max_triangles,
                   SphCtx **out);
/* SPHSystem::~SPHSystem (SPH-3D-2/sph_system.cpp:231-245; free_array, sph_system.cu:44-47) */
int sph_ctx_destroy(SphCtx *ctx);
/* set_parameters (SPH-3D-2/sph_system.cu:34-37): here a per-context copy, no __constant__. */
int sph_set_params(SphCtx *ctx, const void *sysparam);
int sph_get_params(SphCtx *ctx, void *sysparam_out);
/* copy_array(dev_mem, host_mem, n*sizeof(Particle), CUDA_HOST_TO_DEV) (sph_system.cu:49-70,
 * called from add_box_particle, SPH-3D-2/sph_system.cpp:263). */
int sph_upload_particles(SphCtx *ctx, const void *aos, uint32_t n);
/* copy_array(host_mem, dev_mem, n*sizeof(Particle), CUDA_DEV_TO_HOST)
 * (SPH-3D-2/sph_system.cpp:401); writes the live particles in slot order. */
int sph_download_particles(SphCtx *ctx, void *aos, uint32_t n);
int sph_num_particles(SphCtx *ctx, uint32_t *n);

/* ---- phases (each mirrors one reference device function) ---- */
int sph_calc_hash(SphCtx *ctx);   /* calc_hash        SPH-3D-2/sph_system.cu:98-117  */
int sph_sort(SphCtx *ctx);        /* sort_particles   SPH-3D-2/sph_system.cu:119-129 */
int sph_cell_ranges(SphCtx *ctx); /* find_start_end   SPH-3D-2/sph_system.cu:175-202 */
int sph_density(SphCtx *ctx);     /* compute, density half  SPH-3D-2/sph_system.cu:1139-1167 */
int sph_force(SphCtx *ctx);       /* compute, force half    SPH-3D-2/sph_system.cu:1139-1167 */
int sph_integrate(SphCtx *ctx);   /* integrate_velocity SPH-3D-2/sph_system.cu:324-336 */
int sph_energy(SphCtx *ctx);      /* compute_energy   SPH-3D-2/sph_system.cu:1524-1546 */
/* adaptive_particle (SPH-3D-2/sph_system.cu:1617-1659; SPH-3D-4 :1658-1700) / split_particle
 * (SPH-2D/sph_system.cu:1133-1175). Only the first floor((capacity-N)/(n-1)) flagged
 * parents in slot order are split (the reference has no capacity check). */
int sph_split(SphCtx *ctx, uint32_t *n_requested, uint32_t *n_done);
int sph_merge(SphCtx *ctx, uint32_t *groups);  /* merge     SPH-3D-2/sph_system.cu:1791-1815 */
int sph_compact(SphCtx *ctx, uint32_t *n_out); /* rearrange SPH-3D-2/sph_system.cu:1818-1840 */
int sph_density_grid(SphCtx *ctx);  /* integrate_dens   SPH-3D-2/sph_system.cu:2185-2219 */
int sph_grid_normals(SphCtx *ctx);  /* integrate_normal SPH-3D-2/sph_system.cu:2337-2351 */
/* SPHSystem::marching_cube (SPH-3D-2/sph_system.cpp:517-542), on the GPU. */
int sph_marching_cubes(SphCtx *ctx, uint32_t *n_tri);
/* SPHSystem::animation (SPH-3D-2/sph_system.cpp:302-404; SPH-3D-4 :350-453) /
 * animate (SPH-2D/sph_system.cpp:190-264), without the per-step full readback. */
int sph_step(SphCtx *ctx, uint32_t flags, uint32_t *n_out);

/* ---- debug accessors used by the parity tests ---- */
int sph_get_u32(SphCtx *ctx, int which, uint32_t *out, size_t n);
int sph_get_f32(SphCtx *ctx, int which, float *out, size_t n);
int sph_count(SphCtx *ctx, int which, size_t *n);

/* ---- timing: hipEvents around each phase, on the context's stream ----
 * Each event pair costs the stream a few microseconds, so a timed loop that only needs one
 * phase should time only that one: sph_prof_phases takes a mask (bit k = phase k, 0 = off);
 * sph_prof_enable(ctx, on) is the mask of every phase or 0. */
int sph_prof_enable(SphCtx *ctx, int on);
int sph_prof_phases(SphCtx *ctx, uint32_t phase_mask);

/* ---- autotuning of the pair-list lengths (no counterpart in the reference) ----
 * The force and energy kernels queue interacting pairs in per-thread LDS lists; the list
 * length (force 24/32/48, energy 16/24/32 entries) changes only where the lists are
 * drained, never the order in which a particle's terms are summed, so every length gives
 * the same bits. A context tunes itself in rounds of 12 steps (the first round of a tuning
 * run: 6): each step of a round runs both kernels with one candidate length, the
 * candidates taking turns in the order 0 1 2 2 1 0 0 1 2 2 1 0 so that a cost drifting
 * from step to step weighs on all alike, timed with events; after the round each kernel
 * switches to the length with the lowest mean time if it beats the length in use by more
 * than 5% (10% in the later rounds). The first round starts after 8 steps, the next ones
 * every 128 steps (the best length drifts with a scene's neighbour counts).
 * sph_autotune(rounds, after_steps, every_steps) restarts that: `rounds` rounds (0 stops
 * tuning and keeps the current lengths), the first after `after_steps` more steps, then one
 * every `every_steps` steps (at least 12). sph_tuned_lists reports the lengths in use and,
 * in best_ms (6 floats, may be NULL), the mean time per candidate in the last round
 * (force 24/32/48, energy 16/24/32; -1 = not measured). The phase calls (sph_force,
 * sph_energy) use the current lengths. */
int sph_autotune(SphCtx *ctx, uint32_t rounds, uint32_t after_steps, uint32_t every_steps);
int sph_tuned_lists(SphCtx *ctx, uint32_t *force_list, uint32_t *energy_list, float *best_ms);
/* pins the lengths and stops tuning; SPH_EINVAL for a length that is not a candidate */
int sph_set_lists(SphCtx *ctx, uint32_t force_list, uint32_t energy_list);
int sph_prof_reset(SphCtx *ctx);
int sph_prof_read(SphCtx *ctx, int phase, double *ms_total, uint64_t *launches);
int sph_synchronize(SphCtx *ctx);

/* ---- batches of independent scenes ----
 * No counterpart in the reference: its parameter block is one process-wide __constant__
 * (dev_param, SPH-3D-2/sph_kernel.cu:10), so it steps one scene per process. A batch steps
 * up to 256 scenes of one 3D variant (SPH_V3D2 or SPH_V3D4) together: one sort and one
 * launch per phase for all scenes, one host synchronisation per step. Every scene ends
 * each step exactly as its own SphCtx would (sph_step with the same flags); marching
 * cubes is not part of a batch step. */
typedef struct SphBatch SphBatch;
/* sysparams: n_scenes SysParam records; capacity: n_scenes particle capacities (< 2^24 - 1) */
int sph_batch_create(int variant, uint32_t n_scenes, const void *sysparams,
                     const uint32_t *capacity, SphBatch **out);
int sph_batch_destroy(SphBatch *b);
/* The cell grid of a scene is fixed at creation: a record with another total_cell is
 * rejected. Pushes (force_waiting*) last one step, as in sph_step. */
int sph_batch_set_params(SphBatch *b, uint32_t scene, const void *sysparam);
int sph_batch_get_params(SphBatch *b, uint32_t scene, void *sysparam_out);
/* Replaces the particles of one scene (AoS records, as sph_upload_particles). */
int sph_batch_upload(SphBatch *b, uint32_t scene, const void *aos, uint32_t n);
/* Live particles of one scene in slot order (as sph_download_particles). */
int sph_batch_download(SphBatch *b, uint32_t scene, void *aos, uint32_t n);
int sph_batch_num_particles(SphBatch *b, uint32_t *n_per_scene); /* n_scenes entries */
/* Last step, per scene: split parents flagged, parents split, merge groups, status
 * (0 or SPH_EDOMAIN): 4 * n_scenes entries. */
int sph_batch_scene_counts(SphBatch *b, int32_t *out);
/* `steps` steps of every scene with SPH_STEP_SPLIT / SPH_STEP_MERGE; *n_total = live
 * particles of all scenes afterwards. SPH_EEMPTY when no scene has particles; SPH_EDOMAIN
 * when a particle of some scene left its grid (sph_batch_scene_counts names the scenes). */
int sph_batch_step(SphBatch *b, uint32_t flags, uint32_t steps, uint32_t *n_total);
int sph_batch_prof_enable(SphBatch *b, int on);
int sph_batch_prof_phases(SphBatch *b, uint32_t phase_mask);
/* autotuning as sph_autotune / sph_tuned_lists, over the batch's steps */
int sph_batch_autotune(SphBatch *b, uint32_t rounds, uint32_t after_steps, uint32_t every_steps);
int sph_batch_tuned_lists(SphBatch *b, uint32_t *force_list, uint32_t *energy_list, float *best_ms);
int sph_batch_set_lists(SphBatch *b, uint32_t force_list, uint32_t energy_list);
int sph_batch_prof_reset(SphBatch *b);
int sph_batch_prof_read(SphBatch *b, int phase, double *ms_total, uint64_t *launches);
/* How a batch step runs. on = 1 (the default): the whole step is one captured hipGraph of
 * launches that read every count on the device, so a call of `steps` steps launches the
 * graph step after step and synchronises once at the end. Each launch is sized by its
 * buffer's bound (blocks past the live count exit at once; workgroups per CU 0) or is
 * persistent (4, 8 or 16 workgroups per CU, each looping over blocks of 256 particles or
 * cells); the batch's tuner picks the workgroups per CU and the pair-list lengths by timing
 * whole steps (sph_batch_autotune restarts it; phase profiling does not apply inside a
 * graph).
 * on = 0: the kernels are launched one by one each step, sized by the counts read back at
 * the end of the previous step. Both give the same bits for every scene. */
int sph_batch_set_graph(SphBatch *b, int on);
/* pins the workgroups per CU of the graph steps' launches (0 = sized by the bounds, or 4, 8
 * or 16 persistent; SPH_EINVAL otherwise) and stops their tuning */
int sph_batch_set_launch(SphBatch *b, uint32_t groups_per_cu);
/* workgroups per CU in use, the CU count, and in best_ms (4 floats, may be NULL) the mean
 * step time per candidate in the last tuning round (-1 = not measured) */
int sph_batch_tuned_launch(SphBatch *b, uint32_t *groups_per_cu, uint32_t *num_cu, float *best_ms);

/* ---- differentiable rollout and its adjoint ----
 * No counterpart in the reference. The model is the 3D-2 step (SPH-3D-2/sph_system.cpp:
 * 302-404) with marching cubes off, no push and no obstacle spheres, with split and merge
 * as `flags` select (SPH_STEP_SPLIT, SPH_STEP_MERGE; 0 = a fixed particle set), plus a
 * learnable force field: theta holds field_nodes^3 nodes x 3 components (x fastest)
 * spanning the box, sampled trilinearly at each particle and added to its force before the
 * velocity update. The loss after `steps` steps is sum_v (rho_s(x_v) - target_v)^2 over a
 * sample_n^3 grid of points at the cell centres of the box (x fastest), rho_s the
 * particle-radius poly6 splat of the density grid (SPH-3D-2/sph_system.cu:1954-1990). The
 * backward pass returns dL/dtheta (and the adjoint of the initial state) through
 * hand-written adjoints of density, force, integrate, split and merge over each step's
 * topology as the forward took it: which particles split, which merged, the wall clamps
 * and the surface-tension threshold are held, so the gradient is that of the model with
 * the forward's decisions frozen. fp32 throughout, fixed-order reductions: every backward
 * pass of a forward gives the same bits. Capacity is 2n (a 1 -> 2 split of a level-1
 * particle and a 2 -> 1 merge both keep 2 #level1 + #level2, so with split on the initial
 * particles must all be level 1). checkpoint_every = 0 keeps every step on a device tape;
 * k > 0 keeps every k-th state and replays each segment in the backward pass. theta,
 * target, sample and dtheta may be host or device pointers (copied with hipMemcpyDefault);
 * calls synchronise the context's stream before returning. */
typedef struct SphAdj SphAdj;
int sph_adj_create(const void *sysparam3d, uint32_t n, uint32_t field_nodes, uint32_t sample_n,
                   uint32_t checkpoint_every, uint32_t flags, SphAdj **out);
int sph_adj_destroy(SphAdj *a);
int sph_adj_upload(SphAdj *a, const void *aos, uint32_t n); /* initial state, n records */
int sph_adj_set_target(SphAdj *a, const float *target);  /* sample_n^3 floats */
int sph_adj_forward(SphAdj *a, const float *theta, uint32_t steps, float *sample, double *loss);
int sph_adj_backward(SphAdj *a, float *dtheta); /* dL/dtheta of the last forward */
/* live particles after the last forward (the initial count before one) */
int sph_adj_num_particles(SphAdj *a, uint32_t *n);
/* the state after the last forward, n = sph_adj_num_particles records in slot order */
int sph_adj_download(SphAdj *a, void *aos, uint32_t n);
/* per step of the last forward (the first `steps`): live particles before the split,
 * parents split, merge groups (3 uint32 per step) */
int sph_adj_topology(SphAdj *a, uint32_t *out, uint32_t steps);
/* adjoint of the initial positions, velocities and ev (host arrays of 3n floats, or NULL) */
int sph_adj_state_grad(SphAdj *a, float *dpos, float *dvel, float *dev);
int sph_adj_timing(SphAdj *a, double *fwd_ms, double *bwd_ms); /* last forward / backward */

/* ---- measurement helper (no counterpart in the reference) ----
 * Copies a device buffer of `bytes` (rounded down to 16-byte units) `reps` times with a
 * plain vectorised copy kernel, on buffers it allocates and frees. Profiler byte counters
 * of this kernel (16 x bytes/16 read and written per launch) scale those of the other
 * kernels. */
int sph_copy_calibration(size_t bytes, int reps);

/* ---- sort without allocation ----
 * No counterpart in the reference. For callers that cannot allocate device memory (code
 * under HIP stream capture into a graph, serving loops and harnesses that forbid hipMalloc):
 * the sort of sort_particles (SPH-3D-2/sph_system.cu:119-129, thrust::sort_by_key, whose
 * temporary storage thrust allocates) with its working memory in a static device array of
 * the library instead, so the call allocates nothing. Sorts the device arrays hash[0..n)
 * and index[0..n) in place by the low key_bits bits of hash (0 = all 32), ascending and
 * stable (equal keys keep their order), exactly as sph_sort orders a context's keys. The
 * work is enqueued on `stream` (a hipStream_t of the current device; NULL = the null
 * stream) and the call returns without waiting. n at most SPH_SORT_NOALLOC_MAX_KEYS
 * (SPH_ECAP beyond). Each GPU holds its own copy of the working memory (about 10 MB) from
 * the first call until the process ends, shared by every call on that GPU: calls that
 * overlap on one GPU (other streams or threads) race. */
#define SPH_SORT_NOALLOC_MAX_KEYS (1u << 20)
int sph_sort_noalloc(uint32_t *hash, uint32_t *index, uint32_t n, uint32_t key_bits, void *stream);

/* library identification: "sph_amd <version> gfx950" */
const char *sph_version(void);

#ifdef __cplusplus
}
#endif
#endif
