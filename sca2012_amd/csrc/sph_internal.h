// Internal declarations shared by the HIP translation units of libsph_amd.
//
// Device data layout (HBM), per particle, structure-of-arrays and kept in cell order
// (the order of the sorted (hash, slot) list) from the cell-range phase of each step on:
//   pos    float4  x, y, z, w = level (uint bits)        read by every neighbour loop
//   vel    float4  x, y, z, w = 0
//   ev     float4  x, y, z, w = tension (never used by the kernels, kept for AoS I/O)
//   acc    float4  x, y, z, w = surface (|colour-field gradient|)
//   dp     float2  dens, press
//   energy float
//   slot   uint    the particle's index in the reference's Particle array ("slot");
//                  0xffffffff marks a particle removed by merge (dropped at the next sort).
//                  In a batch of scenes the word is scene << 24 | slot within the scene.
// = 80 B per particle; two copies (ping-pong for the reorder).
#ifndef SPH_INTERNAL_H
#define SPH_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/sph_amd.h"
#include "sph_arith.h"

typedef unsigned int uint;

#define SPH_DEAD 0xffffffffu
#define SPH_EMPTY 0xffffffffu

// Parameters as a by-value kernel argument (no __constant__ global, so contexts are
// independent). Index [0] = large particles (level 1), [1] = small (level 2).
struct DevParams {
  int variant, dim, nsplit;
  float world[3];
  float radius[2], kernel[2], mass[2];
  float cell_size;
  uint ncell[3];
  uint total_cell;
  uint cell_base;  // first global cell of this scene (0 for a single scene)
  float gravity[3];
  float wall_damping, rest_density, gas_constant, time_step, viscosity, surface_tension;
  float co_surface, co_energy, split_criteria, merge_criteria;
  float poly6[2], spiky[2], visco[2], grad[2], lapc[2], k2[2], k6[2], poly6_radius[2];
  float radius2[2];
  float visco_k6[2];  // visco / h^6, the per-level factor of the viscosity kernel
  float sqrt_kernel[2];
  uint nd[3];
  uint tot_dens;
  float den_size;
  float sph_pos[3][3], sph_radius[3], sph_damping[3];
  int push_left, push_right, use_cylinder, mouse;
  float extern_force[2], extern_click[2], extern_len, extern_coe;
  float surf_coef, dens_eps;
  uint cap;
  uint dens_mul;
  // RN(1/x) of the pair terms' constant divisors, for sph_div_mk / sph_div_rn (sph_arith.h)
  float rcp_k6[2], rcp_kernel[2], rcp_sqrt_kernel[2];
  // 1 when the force pair term (spiky*(h-r)^2)/h^6 of both levels meets sph_div_mk's
  // preconditions for every pair (checked on the host from the parameters)
  int k6_fast;
  // 1 when h and sqrt(h) of both levels lie within [2^-20, 2^20] (sph_div_rn's divisors)
  int kernel_div_fast;
  // 1 when every particle's energy kernel length h and sqrt(h) are exactly 1 (3D: the large
  // level's, which both levels use; 2D: both levels'), so x/h and x/sqrt(h) equal x
  int energy_unit;
};

// Per-step per-particle factors written by the density kernel for the force kernel.
// Per-particle factors of the force pair terms, written by the density kernel, with the
// particle's ev: what the force kernel reads of a neighbour besides its position, as one
// 48-byte record (3 float4) per particle:
//   vpd.x vpd.y ev.x ev.y | gr_0 lc_0 ev.z - | gr_1 lc_1 ev.z -    (2D: vpd.x vpd.y ev.x ev.y)
// with vpd = (m/dens, press/dens^2), gr_L = ((-grad_L) m)/dens and lc_L = (lapc_L m)/dens
// for L = large (0), small (1): a receiver of level L reads the first part and part 1 + L.
// erec: what the energy kernel reads of a neighbour, (pos.x, pos.y, vel.x, vel.y) after
// integration, written by the integrate kernel (the z components never reach the energy).
// nbl (optional): the force kernel's interacting pairs, written by the density kernel,
// which walks the same candidates in the same order: entries packed as
// (offset index << 27) | j in visiting order, four to a 16-byte group, group g of
// particle i at ((uint4 *)nbl)[g * nbl_stride + i] (group-major, so a wave's loads of a
// group are one contiguous run); nbl_cnt[i] entries. A count above SPH_NBL_K means the
// list overflowed (the entries past SPH_NBL_K are not stored) and the force kernel walks
// the cells itself for that particle's wave.
#define SPH_NBL_K 128u
struct Aux {
  float4 *frec;
  float4 *erec;
  uint *nbl = nullptr;
  uint *nbl_cnt = nullptr;
  uint nbl_stride = 0;
};

struct DevState {
  float4 *pos, *vel, *ev, *acc;
  float2 *dp;
  float *energy;
  uint *slot;
};

// Device-resident counters (the step runs without host round trips).
struct Counters {
  uint n;         // live particles
  uint m;         // physical entries (live + removed, until the next sort)
  uint n_pre;     // live count before split (the cell lists cover [0, n_pre))
  uint s_req;     // split parents flagged
  uint s_done;    // split parents split (capacity policy)
  uint groups;    // merge groups
  uint err;       // bit 0: particle outside the grid
  uint ntri;      // triangles emitted by marching cubes
  uint scan_total;
  uint nmc;  // non-empty marching cubes
  uint pad[6];
};

// Batches of scenes (one step over many independent scenes): scene s owns the global cells
// [cell_base_s, cell_base_s + total_cell_s) and the dense slot range [first_s, first_s + n_s)
// of the sort input. SceneCnt is its device-resident bookkeeping.
#define SPH_SCENE_SHIFT 24
#define SPH_LOCAL_MASK 0xffffffu
#define SPH_MAX_SCENES 256u  // capacities stay below 0xffffff, so no slot word is SPH_DEAD

struct SceneCnt {
  uint n;          // live particles (updated by split and compaction)
  uint n_pre;      // live particles at the start of the step
  uint first;      // dense start of the scene's slots in the sort input (pre-split)
  uint r0;         // split-flag scan at `first`
  uint s_req;      // split parents flagged
  uint s_done;     // split parents split (capacity policy)
  uint done_base;  // split parents of the scenes before this one
  uint first2;     // dense start of the scene's slots after split
  uint a0;         // survivor scan at `first2`
  uint groups;     // merge groups
  uint err;        // bit 0: particle outside the scene's grid
  uint pad[5];
};

struct BatchView {
  const DevParams *Ptab;  // per-scene parameters (cell_base, cap set)
  SceneCnt *scn;
  uint n_scenes;
  uint total_cell;        // cells of all scenes
};

// ---------------------------------------------------------------------------------------
// XCD-aware block order. The dispatcher hands workgroups to the 8 XCDs round-robin
// (block b to XCD b mod 8), and each XCD has its own L2. Particles are stored in cell
// order, so consecutive blocks are spatial neighbours; this bijection gives XCD x one
// contiguous run of the grid's blocks instead, and the neighbour data a block reads is
// then mostly in its own XCD's L2. A pure permutation of the work: results do not change.
// ---------------------------------------------------------------------------------------
#define SPH_NXCD 8u
// Wave priority from a neighbour walk's progress (o of NNB cells): 3 in the first quarter
// down to 0 in the last. The instruction arbiter prefers higher priority, then older waves;
// waves that are ahead step down, so the waves of a SIMD progress together instead of the
// youngest running alone at the end. Only the order of issue changes, never a result.
template <int NNB> __device__ inline void sph_walk_prio(int o) {
  const int p = 3 - (o * 4) / NNB;
  if (o > 0 && p == 3 - ((o - 1) * 4) / NNB) return;
  switch (p) {
    case 3: __builtin_amdgcn_s_setprio(3); break;
    case 2: __builtin_amdgcn_s_setprio(2); break;
    case 1: __builtin_amdgcn_s_setprio(1); break;
    default: __builtin_amdgcn_s_setprio(0); break;
  }
}

__device__ inline uint xcd_block(uint b, uint nb) {
  const uint x = b % SPH_NXCD, k = b / SPH_NXCD;
  const uint q = nb / SPH_NXCD, r = nb % SPH_NXCD;
  return x * q + (x < r ? x : r) + k;
}

// ---------------------------------------------------------------------------------------
// fp32 exp and the reference's wavelet, identical bits on host and device.
// exp: Cody & Waite (1980) argument reduction t = k ln2 + r (ln2 split into a head with
// trailing zero bits and a tail), degree-7 Taylor polynomial of e^r on |r| <= ln2/2 by
// Horner's rule, then scaling by 2^k through the exponent field. Every operation is a
// single IEEE binary32 op in fixed order; the library is compiled without contraction.
// ---------------------------------------------------------------------------------------
__host__ __device__ inline float sph_scale2(float p, int k) {
  if (k > 127) {
    p *= 0x1p127f;
    k -= 127;
  } else if (k < -126) {
    p *= 0x1p-126f;
    k += 126;
    if (k < -126) {
      p *= 0x1p-126f;
      k += 126;
    }
  }
  union { uint u; float f; } s;
  s.u = (uint)(k + 127) << 23;
  return p * s.f;
}

__host__ __device__ inline float sph_expf(float t) {
  if (t != t) return t;
  if (t > 88.75f) return __builtin_huge_valf();
  if (t < -104.0f) return 0.0f;
  const float inv_ln2 = 0x1.715476p+0f;
  const float ln2_hi = 0x1.62e4p-1f;
  const float ln2_lo = 0x1.7f7d1cp-20f;
  float kf = __builtin_rintf(t * inv_ln2);
  float r = (t - kf * ln2_hi) - kf * ln2_lo;
  float p = 0x1.a01a02p-13f;
  p = p * r + 0x1.6c16c2p-10f;
  p = p * r + 0x1.111112p-7f;
  p = p * r + 0x1.555556p-5f;
  p = p * r + 0x1.555556p-3f;
  p = p * r + 0x1p-1f;
  p = p * r + 1.0f;
  p = p * r + 1.0f;
  return sph_scale2(p, (int)kf);
}

// sph_expf for |t| < ln2/2: the reduction above gives k = 0 and r = t exactly and the scale
// is a multiplication by 1, so the polynomial alone returns the same bits.
__host__ __device__ inline float sph_expf_small(float t) {
  float p = 0x1.a01a02p-13f;
  p = p * t + 0x1.6c16c2p-10f;
  p = p * t + 0x1.111112p-7f;
  p = p * t + 0x1.555556p-5f;
  p = p * t + 0x1.555556p-3f;
  p = p * t + 0x1p-1f;
  p = p * t + 1.0f;
  p = p * t + 1.0f;
  return p;
}

// The wavelet for |x| <= 1 (every call site: |x| = |dx|/h with dx a component of a pair
// separation shorter than h), where the exp argument lies in [-0.125, 0].
__host__ __device__ inline float sph_wavelet_unit(float x) {
  const float A = 0x1.3a0156p-1f;
  const float ln_b = 0x1.fffff6p-1f;
  float xx = x * x;
  float b = xx / 4.0f - 1.0f;
  float y = (-xx) / 8.0f;
  return (A * b) * sph_expf_small(y * ln_b);
}

// Entries j .. j+3 of a cell's (or list's) run [.., e), j < e: entries past the end are
// replaced by the run's last entry (in bounds, the same cache line) and the caller skips
// their terms, so a run is walked in groups of four to its end, without a one-by-one tail
// loop.
__device__ inline void sph_group4(const float4 *__restrict__ pos, uint j, uint e, float4 q[4]) {
  const uint l = e - 1;
  q[0] = pos[j];
  q[1] = pos[min(j + 1, l)];
  q[2] = pos[min(j + 2, l)];
  q[3] = pos[min(j + 3, l)];
}

// The same four entries through a buffer resource over the array's first n entries
// (n < 2^28): a load past n returns zeros (the hardware range check) instead of faulting,
// so a run's last group needs no clamping (the caller skips the terms past the run's end,
// and an entry between the run's end and n is a valid particle, merely unused); the four
// loads share one address register with immediate offsets.
__device__ inline __amdgpu_buffer_rsrc_t sph_rsrc(const float4 *a, uint n) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<float4 *>(a), (short)0, (int)(n * 16u),
                                           0x00020000);  // gfx9 raw-buffer word 3
}
__device__ inline void sph_group4(__amdgpu_buffer_rsrc_t r, uint j, float4 q[4]) {
  const uint off = j * 16u;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, off + 16u * k, 0, 0);
    q[k] = make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]),
                       __uint_as_float(v[3]));
  }
}

// Correctly rounded sqrtf for finite x >= 2^-96: the hardware estimate, then whichever
// neighbour one ulp away the two fma residuals select (the correction the compiler's IEEE
// sqrtf applies; its extra input scaling and zero/infinity fix-ups concern only x below
// 2^-96 and non-finite x).
__device__ inline float sph_sqrt_normal(float x) {
  float s = __builtin_amdgcn_sqrtf(x);
  const float sm = __uint_as_float(__float_as_uint(s) - 1u);
  const float sp = __uint_as_float(__float_as_uint(s) + 1u);
  const float rm = __builtin_fmaf(-sm, s, x);
  const float rp = __builtin_fmaf(-sp, s, x);
  if (rm <= 0.0f) s = sm;
  if (rp > 0.0f) s = sp;
  return s;
}

// SPH-3D-2/sph_system.cu:11-16: 2/(PI^.25 (3 sigma)^.5) (x^2/sigma^2 - 1) EXP^(-x^2/(2 sigma^2)),
// sigma = 2, PI = 3.141592f, EXP = 2.718281f; EXP^y = exp(y ln EXP).
__host__ __device__ inline float sph_wavelet(float x) {
  const float A = 0x1.3a0156p-1f;    // 2 / (float(PI^0.25) * sqrtf(6))
  const float ln_b = 0x1.fffff6p-1f; // float(ln 2.718281f)
  float xx = x * x;
  float b = xx / 4.0f - 1.0f;
  float y = (-xx) / 8.0f;
  return (A * b) * sph_expf(y * ln_b);
}

// ---------------------------------------------------------------------------------------
// Launch wrappers (sph_kernels.hip / sph_prims.hip)
// ---------------------------------------------------------------------------------------
struct SortScratch {
  uint *keys_alt, *vals_alt;  // ping-pong partners
  uint *tile_hist;            // RADIX * max_tiles
  uint *scan_tmp;             // block sums for the scan
  size_t max_tiles;
  uint *len_dev;              // one device word: histogram length of a sort by device count
};

size_t sph_scan_tmp_elems(size_t n);
size_t sph_sort_tiles(size_t n);
size_t sph_sort_hist_elems(size_t n);  // tile histogram entries for a sort of n keys

// Exclusive scan of in[0..*n_dev) (host upper bound n_max) into out; total -> *total_dev.
void sph_exclusive_scan(const uint *in, uint *out, const uint *n_dev, uint n_max, uint *total_dev,
                        uint *tmp, hipStream_t st);
// Stable LSD radix sort of keys[0..n) (values = slot index 0..n-1, generated) by the low
// `bits` bits. Result (keys, vals) ends in (*keys_out, *vals_out), which point into
// {keys, vals} or {sc.keys_alt, sc.vals_alt}. n_dev (optional, needs sc.len_dev): the key
// count in device memory, n bounding it; grid (optional): persistent histogram and scatter
// launches of that many workgroups.
void sph_radix_sort(uint *keys, uint *vals, uint n, int bits, SortScratch &sc, uint **keys_out,
                    uint **vals_out, hipStream_t st, const uint *n_dev = nullptr, uint grid = 0);
void sph_fill_u32(uint *p, uint value, size_t n, hipStream_t st);
// cstart[0..n) = SPH_EMPTY and cend[0..n) = 0 in one launch
void sph_fill_cells(uint *cstart, uint *cend, size_t n, hipStream_t st);
// the cells that launch_cell_bounds(skeys, n, ncell, ...) wrote, back to the fill values
void sph_reset_cells(const uint *skeys, uint n, uint ncell, uint *cstart, uint *cend,
                     hipStream_t st);

// `bv` (batches only) selects the kernels that read the scene of each particle from its
// slot word and its parameters from bv->Ptab.
// Launch shapes of the per-item kernels: `n_dev` (optional) is a device word holding the
// item count, and `n` then only bounds it (the buffers' size); `grid` (optional) is a
// persistent grid of that many workgroups, each looping over blocks of 256 items, else the
// grid has one workgroup per block of n. A launch with both is independent of the host's
// view of the counts, as a captured graph needs.
void launch_hash(int variant, const DevParams &P, DevState s, const uint *cnt_m, uint m_max,
                 uint *keys, uint *slot2phys, Counters *cnt, hipStream_t st,
                 const BatchView *bv = nullptr, uint grid = 0);
void launch_cell_bounds(const uint *skeys, uint n, uint ncell, uint *cstart, uint *cend,
                        hipStream_t st, const uint *n_dev = nullptr, uint grid = 0);
// set_m: also cnt->m = cnt->n (removed entries are gone once reordered)
// all_fields == false: acc, dp and energy are not carried (a step recomputes them)
void launch_reorder(DevState src, DevState dst, const uint *svals, const uint *slot2phys, uint n,
                    uint *slot2phys_new, hipStream_t st, Counters *set_m = nullptr,
                    bool all_fields = true, const uint *n_dev = nullptr, uint grid = 0);
// The neighbour kernels (density, force, energy) and merge run one thread per particle (or
// cell, or sorted key) on a grid sized by n, which bounds the device count n_dev if given.
// density with a.nbl: also the force kernel's pair lists; force with use_nbl: reads them
// (they must come from a density pass over the same state, cell lists and parameters).
void launch_density(int variant, const DevParams &P, DevState s, Aux a, const uint *cstart,
                    const uint *cend, uint n, hipStream_t st, const BatchView *bv = nullptr,
                    const uint *n_dev = nullptr);
// Pair-list lengths the force and energy kernels are built with; `cfg` indexes them.
constexpr int kForceLists[3] = {24, 32, 48};
constexpr int kEnergyLists[3] = {16, 24, 32};
constexpr int kForceListDefault = 1, kEnergyListDefault = 0;
// A new context or batch runs a round of 6 tuning steps after its first 8 steps and then
// one every 128 steps (a scene's neighbour counts, and with them the best length, drift).
constexpr uint kTuneRounds = 0xffffffffu, kTuneAfter = 8, kTuneEvery = 128;
constexpr float kTuneMargin = 0.05f;  // a length replaces the one in use when 5% faster
// ... and, in a list-tuning round after the first, when 10% faster: those rounds compare a
// length chosen for the scene against alternatives that are rarely better by more than a
// few percent, so their measurement noise must not move the choice
constexpr float kTuneMarginLater = 0.10f;
void launch_force(int variant, const DevParams &P, DevState s, Aux a, const uint *cstart,
                  const uint *cend, uint n, hipStream_t st, const BatchView *bv = nullptr,
                  int cfg = kForceListDefault, const uint *n_dev = nullptr,
                  bool use_nbl = false);
// erec (optional): also each particle's energy record (struct Aux)
void launch_integrate(int variant, const DevParams &P, DevState s, uint n, hipStream_t st,
                      const BatchView *bv = nullptr, float4 *erec = nullptr,
                      const uint *n_dev = nullptr, uint grid = 0);
// flag_by_slot (optional): also the split decision of each particle (as launch_decide);
// erec: the energy records of the current positions and velocities (written by the
// integrate kernel, or by launch_make_erec), read for the neighbours' pair terms
void launch_energy(int variant, const DevParams &P, DevState s, const uint *cstart,
                   const uint *cend, uint n, hipStream_t st, const BatchView *bv, int cfg,
                   uint *flag_by_slot, const float4 *erec, const uint *n_dev = nullptr);
void launch_make_erec(DevState s, uint n, float4 *erec, hipStream_t st,
                      const uint *n_dev = nullptr);
void launch_decide(int variant, const DevParams &P, DevState s, uint n, uint *flag_by_slot,
                   hipStream_t st, const BatchView *bv = nullptr);
// A single scene's split also applies the capacity policy to the counters (n, m, n_pre,
// s_req, s_done); a batch's counters come from launch_batch_split_plan.
void launch_split(int variant, const DevParams &P, DevState s, uint n, const uint *flag_by_slot,
                  const uint *rank_by_slot, Counters *cnt, uint *status, hipStream_t st,
                  const BatchView *bv = nullptr, const uint *n_dev = nullptr, uint grid = 0);
// 2D merge qualification word per cell (bit 0; the entry + 1 of the cell id's particle in
// bits 1..31 when it is listed in that cell)
void launch_merge_q2d(const DevParams &P, DevState s, const uint *slot2phys, const Counters *cnt,
                      const uint *cstart, const uint *cend, uint *q2d, hipStream_t st);
// skeys/nkeys (optional): the step's sorted cell keys; on a grid with more cells than keys
// the merge then runs one thread per key instead of one per cell (same result).
// qbits (optional, qbits_len words of scratch, at least two per 64 keys): the entries'
// qualification is evaluated first, one thread per entry (k_merge_qual), instead of inside
// each cell's walk (same result)
void launch_merge(int variant, const DevParams &P, DevState s, const uint *cstart,
                  const uint *cend, const uint *q2d, uint *status, Counters *cnt, hipStream_t st,
                  const BatchView *bv = nullptr, const uint *skeys = nullptr, uint nkeys = 0,
                  const uint *nkeys_dev = nullptr, uint *group_out = nullptr,
                  uint *qbits = nullptr, size_t qbits_len = 0);
void launch_alive_by_slot(DevState s, const uint *status, const Counters *cnt, uint m_max,
                          uint *alive_by_slot, hipStream_t st, const BatchView *bv = nullptr,
                          uint grid = 0);
// A single scene's renumbering also sets the live count to the survivors (cnt->scan_total).
void launch_renumber(DevState s, const uint *status, const uint *newslot, Counters *cnt,
                     uint m_max, hipStream_t st, const BatchView *bv = nullptr, uint grid = 0);
void launch_begin_step(Counters *cnt, hipStream_t st);
// batch bookkeeping (one block each): dense slot starts, split plan, survivor counts
void launch_batch_begin(const BatchView &bv, Counters *cnt, hipStream_t st);
void launch_batch_split_plan(int variant, const BatchView &bv, const uint *rank_by_slot, uint m,
                             bool do_split, Counters *cnt, hipStream_t st,
                             const uint *m_dev = nullptr);
void launch_batch_compact_plan(const BatchView &bv, const uint *newslot, Counters *cnt,
                               hipStream_t st);
// entries of one scene: mark removed / slot -> physical index map
void launch_scene_kill(DevState s, uint m, uint scene, hipStream_t st);
void launch_scene_slot2phys(DevState s, uint m, uint scene, uint *slot2phys, hipStream_t st);
// move the live entries of [0, m) to the front of dst, in order; *cnt_live = their number
void launch_pack_live(DevState src, DevState dst, uint m, uint *flags, uint *pos, uint *tmp,
                      uint *total_dev, hipStream_t st);
// occ: 2 * sph_occ_words(total_cell) words of scratch for the cell-occupancy bitmap of the
// general (per-voxel) kernel and its 27-cell dilation; unused by the cell-block kernel
inline size_t sph_occ_words(size_t ncell) { return (ncell + 63) / 64 * 2; }
void launch_density_grid(const DevParams &P, DevState s, const uint *cstart, const uint *cend,
                         float *dgrid, bool cell_blocks, hipStream_t st, uint *occ);
void launch_normals(const DevParams &P, const float *dgrid, float *normal3, hipStream_t st);
size_t mc_segments(const DevParams &P);  // z-segments of marching cubes (count/offset entries)
void launch_mc_count(const DevParams &P, const float *dgrid, uint *seg_tri, uint *seg_cube,
                     hipStream_t st);
void launch_mc_list(const DevParams &P, const float *dgrid, const uint *seg_cube,
                    const uint *seg_toff, const uint *seg_coff, uint cap, uint2 *clist,
                    hipStream_t st);
void launch_mc_emit(const DevParams &P, const float *dgrid, const float *normal3, const uint2 *clist,
                    const Counters *cnt, uint max_tri, float *tris, hipStream_t st);
// AoS records -> SoA entries of s (entry i gets slot word slot_or | i)
void launch_upload_scatter(int variant, const void *aos_dev, uint n, DevState s, hipStream_t st,
                           uint slot_or = 0u);
void launch_download_gather(int variant, DevState s, const uint *slot2phys, uint n, void *aos_dev,
                            hipStream_t st);
void launch_slot2phys(DevState s, uint m, uint *slot2phys, hipStream_t st);

// Online autotuner of the pair-list lengths of the force and energy kernels, shared by
// the single-scene context and the batch. The length never changes the order in which a
// particle's terms are summed, so the results do not depend on tuning. Tuning starts after
// `after` steps (a scene's first steps are not representative: cold caches, and an
// adaptive scene grows over its first few steps). A round is 12 tuning steps (the first
// round of a tuning run 6); each runs each kernel once, with one candidate length,
// bracketed by events; the candidates take turns in the order 0 1 2 2 1 0 0 1 2 2 1 0, so
// a cost that drifts linearly from step to step (a scene's neighbour counts keep changing)
// weighs on every candidate alike. The first round decides between lengths that differ by
// 20% or more at a scene's start; the later ones compare lengths within a few percent, so
// each candidate's mean is over four launches (with two, launch-to-launch noise made a
// later round pick a length 9% slower than the one in use). After a round
// each kernel switches to the candidate with the lowest mean time if it beats the length
// in use by more than kTuneMargin. Rounds repeat every `every` steps, as the best length
// drifts with the neighbour counts (shorter lists win while a scene is sparse).
struct ListTuner {
  static constexpr int kRound = 12;
  static constexpr int kOrder[kRound] = {0, 1, 2, 2, 1, 0, 0, 1, 2, 2, 1, 0};
  int force = kForceListDefault, energy = kEnergyListDefault;
  uint rounds_left = 0, skip = 0, every = 0;
  int steps_left = 0, pos = 0, round_len = 0;
  bool pending = false;         // a tuning step's events are not read yet
  float sum[2][3];              // [force, energy][candidate] total ms over the tuning steps
  int cnt[2][3];
  float best[2][3];             // mean ms per candidate (-1: not measured)
  hipEvent_t ev[2][2] = {};     // [kernel][begin, end]

  bool create() {
    for (auto &k : ev)
      for (auto &e : k)
        if (hipEventCreate(&e) != hipSuccess) return false;
    return true;
  }
  void destroy() {
    for (auto &k : ev)
      for (auto &e : k)
        if (e) (void)hipEventDestroy(e);
  }
  void start(uint rounds, uint after, uint every_steps) {
    rounds_left = rounds;
    round_len = kRound / 2;
    steps_left = rounds ? round_len : 0;
    skip = rounds ? after : 0;
    every = every_steps > (uint)kRound ? every_steps : (uint)kRound;
    pos = 0;
    pending = false;
    clear();
  }
  void clear() {
    for (int k = 0; k < 2; k++)
      for (int c = 0; c < 3; c++) {
        sum[k][c] = 0.0f;
        cnt[k][c] = 0;
        best[k][c] = -1.0f;
      }
  }
  bool tuning() const { return steps_left > 0 && skip == 0; }
  bool active() const { return steps_left > 0; }
  // pin the lengths (stops tuning); false if a length is not a candidate
  bool set(uint force_len, uint energy_len) {
    int f = -1, e = -1;
    for (int c = 0; c < 3; c++) {
      if ((uint)kForceLists[c] == force_len) f = c;
      if ((uint)kEnergyLists[c] == energy_len) e = c;
    }
    if (f < 0 || e < 0) return false;
    force = f;
    energy = e;
    rounds_left = 0;
    steps_left = 0;
    skip = 0;
    pending = false;
    return true;
  }
  // kernel 0 = force, 1 = energy; launch(cfg) enqueues one run with that candidate
  template <class F> void run(int kernel, hipStream_t st, F launch) {
    if (!tuning()) {
      launch(kernel == 0 ? force : energy);
      return;
    }
    (void)hipEventRecord(ev[kernel][0], st);
    launch(kOrder[pos % kRound]);
    (void)hipEventRecord(ev[kernel][1], st);
    pending = true;
  }
  // after the step's synchronisation
  void end_step() {
    if (steps_left == 0) return;
    if (skip > 0) {
      skip--;
      return;
    }
    if (!pending) return;
    const int c = kOrder[pos % kRound];
    for (int k = 0; k < 2; k++) {
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, ev[k][0], ev[k][1]) == hipSuccess) {
        sum[k][c] += ms;
        cnt[k][c] += 1;
      }
    }
    pending = false;
    pos++;
    if (--steps_left == 0) {
      const int cur[2] = {force, energy};
      int pick[2] = {cur[0], cur[1]};
      const float margin = round_len < kRound ? kTuneMargin : kTuneMarginLater;
      for (int k = 0; k < 2; k++) {
        for (int j = 0; j < 3; j++) best[k][j] = cnt[k][j] ? sum[k][j] / (float)cnt[k][j] : -1.0f;
        for (int j = 0; j < 3; j++)
          if (best[k][j] >= 0.0f && best[k][cur[k]] >= 0.0f &&
              best[k][j] < best[k][cur[k]] * (1.0f - margin) && best[k][j] < best[k][pick[k]])
            pick[k] = j;
      }
      force = pick[0];
      energy = pick[1];
      rounds_left--;
      if (rounds_left > 0) {  // the next round, `every` steps after this one started
        skip = every - round_len;
        round_len = kRound;
        steps_left = kRound;
        pos = 0;
        for (int k = 0; k < 2; k++)
          for (int j = 0; j < 3; j++) {
            sum[k][j] = 0.0f;
            cnt[k][j] = 0;
          }
      }
    }
  }
};

// Host-side parameter handling shared by the single-scene context and the batch.
namespace sphhost {
bool valid_variant(int v);
DevParams make_dev_params(int variant, const void *raw, uint cap);
int key_bits(uint total_cell);
size_t sysparam_size(int variant);
size_t particle_size(int variant);
}  // namespace sphhost

#endif
