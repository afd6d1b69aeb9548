// C ABI of libsph_amd (declared in include/sph_amd.h): parameter derivation, the context
// that owns every device buffer, the phase entry points and the whole step.
//
// Per-step device work (one HIP stream, no host round trip until the end of the step):
//   hash -> radix sort -> cell ranges + reorder into cell order -> density -> force ->
//   integrate -> [density grid -> normals -> marching cubes] -> energy ->
//   [decide -> scan -> split] -> [merge -> renumber survivors]
// Counts that change inside the step (after split and merge) stay on the device; kernels
// read them and are launched over host-side upper bounds.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "sph_internal.h"

static_assert(sizeof(SphSysParam3D) == 296, "SysParam 3D layout (SPH-3D-2/sph_param.h)");
static_assert(sizeof(SphParticle3D) == 72, "Particle 3D layout (SPH-3D-2/sph_system.h)");
static_assert(sizeof(SphSysParam2D) == 200, "SysParam 2D layout (SPH-2D/sph_param.h)");
static_assert(sizeof(SphParticle2D) == 56, "Particle 2D layout (SPH-2D/sph_system.h)");
static_assert(offsetof(SphSysParam3D, sphere1) == 204, "SysParam 3D sphere offset");
static_assert(offsetof(SphSysParam3D, flag0) == 292, "SysParam 3D flag offset");
static_assert(offsetof(SphSysParam2D, extern_force) == 168, "SysParam 2D extern_force offset");

#define SPH_VERSION "sph_amd 0.1 gfx950"

namespace {

const float kPI = 3.141592f;  // SPH-3D-2/sph_header.h:22

bool valid_variant(int v) { return v == SPH_V2D || v == SPH_V3D2 || v == SPH_V3D4; }

// Kernel constants of the reference constructor (SPH-3D-2/sph_system.cpp:63-84;
// SPH-2D/sph_system.cpp:19-41). pow(float, int) is evaluated in double, as C++11 does.
void derive_kernel_constants(float h, float &poly6, float &spiky, float &visco, float &grad,
                             float &lapc, float &k2, float &k6) {
  poly6 = (float)(315.0f / (64.0f * kPI * std::pow((double)h, 9)));
  spiky = (float)(-45.0f / (kPI * std::pow((double)h, 6)));
  visco = (float)(45.0f / (kPI * std::pow((double)h, 6)));
  grad = (float)(-945.0f / (32.0f * kPI * std::pow((double)h, 9)));
  lapc = (float)(-945.0f / (8.0f * kPI * std::pow((double)h, 9)));
  k2 = h * h;
  k6 = (float)std::pow((double)h, 6);
}

void derive3(SphSysParam3D &s) {
  s.cell_size = s.large_kernel;
  s.row_cell = (uint)std::ceil(s.world_width / s.cell_size);
  s.col_cell = (uint)std::ceil(s.world_height / s.cell_size);
  s.len_cell = (uint)std::ceil(s.world_length / s.cell_size);
  s.total_cell = s.row_cell * s.col_cell * s.len_cell;
  s.world_width = s.row_cell * s.cell_size;
  s.world_height = s.col_cell * s.cell_size;
  s.world_length = s.len_cell * s.cell_size;
  derive_kernel_constants(s.large_kernel, s.large_poly6, s.large_spiky, s.large_visco,
                          s.large_grad_poly6, s.large_lapc_poly6, s.large_kernel_2, s.large_kernel_6);
  derive_kernel_constants(s.small_kernel, s.small_poly6, s.small_spiky, s.small_visco,
                          s.small_grad_poly6, s.small_lapc_poly6, s.small_kernel_2, s.small_kernel_6);
  s.large_poly6_radius = (float)(315.0f / (64.0f * kPI * std::pow((double)s.large_radius, 9)));
  s.small_poly6_radius = (float)(315.0f / (64.0f * kPI * std::pow((double)s.small_radius, 9)));
  if (s.dens_mul == 0) s.dens_mul = 4;
  s.den_size = s.cell_size / s.dens_mul;
  s.row_dens = s.row_cell * s.dens_mul + 1;
  s.col_dens = s.col_cell * s.dens_mul + 1;
  s.len_dens = s.len_cell * s.dens_mul + 1;
  s.tot_dens = s.row_dens * s.col_dens * s.len_dens;
}

void derive2(SphSysParam2D &s) {
  derive_kernel_constants(s.large_kernel, s.large_poly6, s.large_spiky, s.large_visco,
                          s.large_grad_poly6, s.large_lapc_poly6, s.large_kernel2, s.large_kernel6);
  derive_kernel_constants(s.small_kernel, s.small_poly6, s.small_spiky, s.small_visco,
                          s.small_grad_poly6, s.small_lapc_poly6, s.small_kernel2, s.small_kernel6);
  s.cell_size = s.large_kernel;
  s.row_cell = (uint)std::ceil(s.world_width / s.cell_size);
  s.col_cell = (uint)std::ceil(s.world_depth / s.cell_size);
  s.total_cell = s.row_cell * s.col_cell;
  s.world_width = s.row_cell * s.cell_size;
  s.world_depth = s.col_cell * s.cell_size;
}

// SPH-3D-2/sph_system.cpp:6-229 and SPH-3D-4/sph_system.cpp:6-230
void defaults3(int variant, SphSysParam3D &s) {
  std::memset(&s, 0, sizeof(s));
  const bool v4 = variant == SPH_V3D4;
  s.max_particle = 1000000;
  s.world_width = v4 ? 48.0f : 32.0f;
  s.world_height = v4 ? 24.0f : 16.0f;
  s.world_length = v4 ? 24.0f : 16.0f;
  s.large_radius = v4 ? 1.2f : 1.0f;
  s.small_radius = 0.8f;
  s.large_kernel = v4 ? 1.2f : 1.0f;
  s.large_mass = 1.0f;
  s.small_kernel = v4 ? 1.0f : 0.9f;
  s.small_mass = v4 ? 0.25f : 0.5f;
  s.gravity = {0.0f, v4 ? -4.0f : -5.0f, 0.0f};
  s.wall_damping = -0.6f;
  s.rest_density = v4 ? 0.5f : 1.0f;
  s.gas_constant = 1.0f;
  s.time_step = 0.1f;
  s.viscosity = v4 ? 1.0f : 3.0f;
  s.surface_tension = 0.8f;
  s.co_surface = v4 ? 1.5f : 1.0f;
  s.co_energy = 1.0f;
  s.split_criteria = v4 ? 1.0f : 1.2f;
  s.merge_criteria = v4 ? 0.6f : 0.4f;
  s.dens_mul = 4;
  s.sphere1 = {{11.0f, 3.0f, 4.0f}, 2.0f, 1.0f};
  s.sphere2 = {{11.0f, 3.0f, 12.0f}, 2.0f, 1.0f};
  s.sphere3 = {{18.0f, 3.0f, 8.0f}, 2.0f, 1.0f};
  s.split_energy = 0.8f;
  s.merge_energy = 0.02f;
  derive3(s);
}

// SPH-2D/sph_system.cpp:5-146
void defaults2(SphSysParam2D &s) {
  std::memset(&s, 0, sizeof(s));
  s.max_particle = 100000;
  s.world_width = 32.0f;
  s.world_depth = 32.0f;
  s.large_mass = 1.0f;
  s.small_mass = 0.5f;
  s.large_kernel = 1.0f;
  s.small_kernel = 0.9f;
  s.large_radius = 1.0f;
  s.small_radius = 0.8f;
  s.gravity = {0.0f, -4.5f};
  s.wall_damping = -0.6f;
  s.rest_density = 2.0f;
  s.gas_constant = 1.0f;
  s.viscosity = 2.0f;
  s.time_step = 0.1f;
  s.surface_normal = 0.5f;
  s.co_surface = 0.8f;
  s.co_energy = 1.0f;
  s.thresh_split = 0.7f;
  s.thresh_merge = 0.04f;
  s.extern_len = 2.0f;
  s.extern_coe = 1.0f;
  derive2(s);
}

DevParams make_dev_params(int variant, const void *raw, uint cap) {
  DevParams P;
  std::memset(&P, 0, sizeof(P));
  P.variant = variant;
  P.cap = cap;
  if (variant == SPH_V2D) {
    const SphSysParam2D &s = *(const SphSysParam2D *)raw;
    P.dim = 2;
    P.nsplit = 2;
    P.world[0] = s.world_width; P.world[1] = s.world_depth; P.world[2] = 0.0f;
    P.radius[0] = s.large_radius; P.radius[1] = s.small_radius;
    P.kernel[0] = s.large_kernel; P.kernel[1] = s.small_kernel;
    P.mass[0] = s.large_mass; P.mass[1] = s.small_mass;
    P.cell_size = s.cell_size;
    P.ncell[0] = s.row_cell; P.ncell[1] = s.col_cell; P.ncell[2] = 1;
    P.total_cell = s.total_cell;
    P.gravity[0] = s.gravity.x; P.gravity[1] = s.gravity.y; P.gravity[2] = 0.0f;
    P.wall_damping = s.wall_damping; P.rest_density = s.rest_density;
    P.gas_constant = s.gas_constant; P.time_step = s.time_step; P.viscosity = s.viscosity;
    P.surface_tension = s.surface_normal;
    P.co_surface = s.co_surface; P.co_energy = s.co_energy;
    P.split_criteria = s.thresh_split; P.merge_criteria = s.thresh_merge;
    P.poly6[0] = s.large_poly6; P.poly6[1] = s.small_poly6;
    P.spiky[0] = s.large_spiky; P.spiky[1] = s.small_spiky;
    P.visco[0] = s.large_visco; P.visco[1] = s.small_visco;
    P.grad[0] = s.large_grad_poly6; P.grad[1] = s.small_grad_poly6;
    P.lapc[0] = s.large_lapc_poly6; P.lapc[1] = s.small_lapc_poly6;
    P.k2[0] = s.large_kernel2; P.k2[1] = s.small_kernel2;
    P.k6[0] = s.large_kernel6; P.k6[1] = s.small_kernel6;
    P.mouse = s.force_waiting ? 1 : 0;
    P.extern_force[0] = s.extern_force.x; P.extern_force[1] = s.extern_force.y;
    P.extern_click[0] = s.extern_click.x; P.extern_click[1] = s.extern_click.y;
    P.extern_len = s.extern_len; P.extern_coe = s.extern_coe;
    P.surf_coef = 0.04f;  // SPH-2D/sph_system.cu:770
    P.dens_eps = 1e-12f;  // EXT2, SPH-2D/sph_system.cu:8, :320
  } else {
    const SphSysParam3D &s = *(const SphSysParam3D *)raw;
    P.dim = 3;
    P.nsplit = variant == SPH_V3D4 ? 4 : 2;
    P.world[0] = s.world_width; P.world[1] = s.world_height; P.world[2] = s.world_length;
    P.radius[0] = s.large_radius; P.radius[1] = s.small_radius;
    P.kernel[0] = s.large_kernel; P.kernel[1] = s.small_kernel;
    P.mass[0] = s.large_mass; P.mass[1] = s.small_mass;
    P.cell_size = s.cell_size;
    P.ncell[0] = s.row_cell; P.ncell[1] = s.col_cell; P.ncell[2] = s.len_cell;
    P.total_cell = s.total_cell;
    P.gravity[0] = s.gravity.x; P.gravity[1] = s.gravity.y; P.gravity[2] = s.gravity.z;
    P.wall_damping = s.wall_damping; P.rest_density = s.rest_density;
    P.gas_constant = s.gas_constant; P.time_step = s.time_step; P.viscosity = s.viscosity;
    P.surface_tension = s.surface_tension;
    P.co_surface = s.co_surface; P.co_energy = s.co_energy;
    P.split_criteria = s.split_criteria; P.merge_criteria = s.merge_criteria;
    P.poly6[0] = s.large_poly6; P.poly6[1] = s.small_poly6;
    P.spiky[0] = s.large_spiky; P.spiky[1] = s.small_spiky;
    P.visco[0] = s.large_visco; P.visco[1] = s.small_visco;
    P.grad[0] = s.large_grad_poly6; P.grad[1] = s.small_grad_poly6;
    P.lapc[0] = s.large_lapc_poly6; P.lapc[1] = s.small_lapc_poly6;
    P.k2[0] = s.large_kernel_2; P.k2[1] = s.small_kernel_2;
    P.k6[0] = s.large_kernel_6; P.k6[1] = s.small_kernel_6;
    P.poly6_radius[0] = s.large_poly6_radius; P.poly6_radius[1] = s.small_poly6_radius;
    P.nd[0] = s.row_dens; P.nd[1] = s.col_dens; P.nd[2] = s.len_dens;
    P.tot_dens = s.tot_dens;
    P.den_size = s.den_size;
    P.dens_mul = s.dens_mul;
    const SphSphere *sp[3] = {&s.sphere1, &s.sphere2, &s.sphere3};
    for (int k = 0; k < 3; k++) {
      P.sph_pos[k][0] = sp[k]->pos.x; P.sph_pos[k][1] = sp[k]->pos.y; P.sph_pos[k][2] = sp[k]->pos.z;
      P.sph_radius[k] = sp[k]->radius; P.sph_damping[k] = sp[k]->damping;
    }
    if (variant == SPH_V3D4) {
      P.push_left = s.flag0; P.push_right = s.flag1; P.use_cylinder = s.flag2;
    } else {
      P.push_left = s.flag0; P.push_right = 0; P.use_cylinder = s.flag1;
    }
    P.surf_coef = variant == SPH_V3D4 ? 0.02f : 0.1f;  // 3D-4 :1136, 3D-2 :1129
    P.dens_eps = 0.0001f;
  }
  P.radius2[0] = P.radius[0] * P.radius[0];
  P.radius2[1] = P.radius[1] * P.radius[1];
  P.visco_k6[0] = P.visco[0] / P.k6[0];
  P.visco_k6[1] = P.visco[1] / P.k6[1];
  P.sqrt_kernel[0] = std::sqrt(P.kernel[0]);  // pow(kernel, 0.5f)
  P.sqrt_kernel[1] = std::sqrt(P.kernel[1]);
  P.k6_fast = 1;
  P.kernel_div_fast = 1;
  for (int l = 0; l < 2; l++) {  // after sqrt_kernel
    P.rcp_k6[l] = sph_rcp_rn(P.k6[l]);
    P.rcp_kernel[l] = sph_rcp_rn(P.kernel[l]);
    P.rcp_sqrt_kernel[l] = sph_rcp_rn(P.sqrt_kernel[l]);
    if (!sph_k6_div_fast(P.spiky[l], P.kernel[l], P.k6[l])) P.k6_fast = 0;
    for (float d : {P.kernel[l], P.sqrt_kernel[l]})
      if (!(d >= 0x1p-20f && d <= 0x1p20f)) P.kernel_div_fast = 0;
  }
  P.energy_unit = 1;
  for (int l = 0; l < (P.dim == 3 ? 1 : 2); l++)
    if (!(P.kernel[l] == 1.0f && P.sqrt_kernel[l] == 1.0f)) P.energy_unit = 0;
  return P;
}

// The cell-block density-grid kernel assumes that the voxels dens_mul*c + 1 ... dens_mul*c +
// dens_mul along each axis are exactly those the reference maps to cell c,
// floor((float)v / (float)n_dens * (float)n_cell) (SPH-3D-2/sph_system.cu:1946-1948). This
// checks it for every voxel index, in the same fp32 arithmetic.
bool dgrid_cell_blocks_exact(const DevParams &P) {
  if (P.dim != 3 || P.dens_mul != 4) return false;
  for (int a = 0; a < 3; a++) {
    if (P.nd[a] != P.ncell[a] * P.dens_mul + 1) return false;
    for (uint v = 1; v < P.nd[a]; v++) {
      const uint c = (uint)std::floor((float)v / (float)P.nd[a] * (float)P.ncell[a]);
      if (c != (v - 1) / P.dens_mul) return false;
    }
  }
  return true;
}

int key_bits(uint total_cell) {
  int b = 1;
  while (b < 32 && ((uint64_t)1 << b) < (uint64_t)total_cell) b++;
  return b;
}

}  // namespace

namespace sphhost {
bool valid_variant(int v) { return ::valid_variant(v); }
DevParams make_dev_params(int variant, const void *raw, uint cap) {
  return ::make_dev_params(variant, raw, cap);
}
int key_bits(uint total_cell) { return ::key_bits(total_cell); }
size_t sysparam_size(int variant) {
  return variant == SPH_V2D ? sizeof(SphSysParam2D) : sizeof(SphSysParam3D);
}
size_t particle_size(int variant) {
  return variant == SPH_V2D ? sizeof(SphParticle2D) : sizeof(SphParticle3D);
}
}  // namespace sphhost

struct SphCtx {
  int variant = 0;
  unsigned char raw[296];
  size_t raw_size = 0;
  DevParams P;
  uint cap = 0, max_tri = 0;
  hipStream_t st = nullptr;
  DevState cur{}, alt{};
  Aux aux{};
  std::vector<void *> allocs;
  uint *keys = nullptr, *vals = nullptr;
  SortScratch sc{};
  uint *s2p_hash = nullptr, *s2p_cell = nullptr;
  uint *cstart = nullptr, *cend = nullptr;
  uint *flag_by_slot = nullptr, *rank_by_slot = nullptr, *alive_by_slot = nullptr;
  uint *newslot = nullptr, *status = nullptr, *q2d = nullptr;
  uint *scan_tmp = nullptr;
  Counters *cnt = nullptr;
  Counters *cnt_host = nullptr;
  float *dgrid = nullptr, *normal3 = nullptr, *tris = nullptr;
  uint *cube_count = nullptr, *cube_off = nullptr;  // per z-segment: triangles, their offsets
  uint *seg_cube = nullptr, *seg_coff = nullptr;    // per z-segment: non-empty cubes, offsets
  uint2 *clist = nullptr;                           // non-empty cubes (id, first triangle)
  bool dgrid_cells = false;                         // cell-block density-grid kernel applies
  uint *occ = nullptr;                              // cell occupancy bits and their dilation
  size_t ncube = 0;
  unsigned char *aos_dev = nullptr;
  uint *skeys = nullptr, *svals = nullptr;
  uint n = 0, m = 0, n_pre = 0;
  uint n_sorted = 0, n_flagged = 0, n_status = 0;
  bool hashed = false, sorted = false, lists_valid = false, merged = false, compacted = false;
  bool erec_ok = false;  // aux.erec holds the current positions and velocities
  // aux.nbl holds the force pairs of the current state, cell lists and parameters (the
  // density kernel writes them; anything else that runs invalidates them)
  bool nbl_ok = false;
  // Cell tables: cells_keys/cells_nkeys are the sorted keys they were built from (valid
  // until the next hash or sort overwrites them); cells_clean = every entry holds its fill
  // value. On a sparse grid the next build resets only the cells those keys wrote.
  const uint *cells_keys = nullptr;
  uint cells_nkeys = 0;
  bool cells_clean = false;
  int kbits = 1;
  hipError_t last_err = hipSuccess;
  // profiling
  uint32_t prof_mask = 0;  // phases timed with events (bit k = phase k)
  std::vector<hipEvent_t> ev_pool;
  struct Pending { int phase; hipEvent_t a, b; };
  std::vector<Pending> pending;
  double prof_ms[SPH_P_COUNT] = {0};
  uint64_t prof_n[SPH_P_COUNT] = {0};
  int open_phase = -1;
  hipEvent_t open_ev = nullptr;

  ListTuner tune;  // pair-list lengths of the force and energy kernels

  ~SphCtx() {
    if (st) (void)hipStreamSynchronize(st);
    tune.destroy();
    for (auto &p : pending) {
      (void)hipEventDestroy(p.a);
      (void)hipEventDestroy(p.b);
    }
    for (auto e : ev_pool) (void)hipEventDestroy(e);
    for (void *p : allocs) (void)hipFree(p);
    if (cnt_host) (void)hipHostFree(cnt_host);
    if (st) (void)hipStreamDestroy(st);
  }

  template <class T> bool alloc(T *&p, size_t count) {
    void *q = nullptr;
    if (count == 0) count = 1;
    if (hipMalloc(&q, count * sizeof(T)) != hipSuccess) return false;
    allocs.push_back(q);
    p = (T *)q;
    return true;
  }

  bool alloc_state(DevState &s) {
    return alloc(s.pos, cap) && alloc(s.vel, cap) && alloc(s.ev, cap) && alloc(s.acc, cap) &&
           alloc(s.dp, cap) && alloc(s.energy, cap) && alloc(s.slot, cap);
  }

  hipEvent_t take_event() {
    if (!ev_pool.empty()) {
      hipEvent_t e = ev_pool.back();
      ev_pool.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
  }
  void pbegin(int phase) {
    if (!((prof_mask >> phase) & 1u)) return;
    open_phase = phase;
    open_ev = take_event();
    (void)hipEventRecord(open_ev, st);
  }
  void pend() {
    if (open_phase < 0) return;
    hipEvent_t b = take_event();
    (void)hipEventRecord(b, st);
    pending.push_back({open_phase, open_ev, b});
    open_phase = -1;
  }
  void collect() {
    for (auto &p : pending) {
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
        prof_ms[p.phase] += ms;
        prof_n[p.phase] += 1;
      }
      ev_pool.push_back(p.a);
      ev_pool.push_back(p.b);
    }
    pending.clear();
  }

  // Synchronise, read the counters, collect timings. Returns an SPH status.
  int sync() {
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(cnt_host, cnt, sizeof(Counters), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      last_err = e;
      return SPH_EHIP;
    }
    collect();
    n = cnt_host->n;
    m = cnt_host->m;
    return SPH_OK;
  }

  void set_counts_on_device(uint nn, uint mm) {
    Counters c;
    std::memset(&c, 0, sizeof(c));
    c.n = nn;
    c.m = mm;
    c.n_pre = nn;
    *cnt_host = c;
    (void)hipMemcpyAsync(cnt, cnt_host, sizeof(Counters), hipMemcpyHostToDevice, st);
  }

  // ---- enqueue-only phase bodies ----
  // before the sorted keys behind the cell tables are overwritten
  void retire_cells() {
    if (!cells_keys) return;
    if ((size_t)cells_nkeys * 16 < P.total_cell) {  // sparse: undo the written cells only
      sph_reset_cells(cells_keys, cells_nkeys, P.total_cell, cstart, cend, st);
      cells_clean = true;
    }
    cells_keys = nullptr;
  }
  void q_hash() {
    nbl_ok = false;
    retire_cells();
    launch_begin_step(cnt, st);
    pbegin(SPH_P_HASH);
    launch_hash(variant, P, cur, cnt ? &cnt->m : nullptr, m, keys, s2p_hash, cnt, st);
    pend();
  }
  void q_sort() {
    nbl_ok = false;
    retire_cells();
    pbegin(SPH_P_SORT);
    sph_radix_sort(keys, vals, n, kbits, sc, &skeys, &svals, st);
    n_sorted = n;
    pend();
  }
  // in_step: the step's density, force and energy kernels follow, so the reorder leaves out
  // the fields they overwrite
  void q_cells(bool in_step = false) {
    nbl_ok = false;
    pbegin(SPH_P_CELLS);
    if (!cells_clean) sph_fill_cells(cstart, cend, P.total_cell, st);
    launch_cell_bounds(skeys, n, P.total_cell, cstart, cend, st);
    cells_keys = skeys;
    cells_nkeys = n;
    cells_clean = false;
    pend();
    pbegin(SPH_P_REORDER);
    launch_reorder(cur, alt, svals, s2p_hash, n, s2p_cell, st, cnt, !in_step);  // and m = n
    pend();
    erec_ok = false;
    DevState t = cur;
    cur = alt;
    alt = t;
  }
  void q_density() {
    pbegin(SPH_P_DENSITY);
    launch_density(variant, P, cur, aux, cstart, cend, n, st);
    pend();
    nbl_ok = true;
  }
  // tuned: a step may run the autotuner's candidates (the phase calls use its choice)
  void q_force(bool tuned = false) {
    pbegin(SPH_P_FORCE);
    const bool use = nbl_ok;
    auto go = [&](int c) {
      launch_force(variant, P, cur, aux, cstart, cend, n, st, nullptr, c, nullptr, use);
    };
    if (tuned) tune.run(0, st, go);
    else go(tune.force);
    pend();
  }
  void q_integrate() {
    nbl_ok = false;
    pbegin(SPH_P_INTEGRATE);
    launch_integrate(variant, P, cur, n, st, nullptr, aux.erec);
    pend();
    erec_ok = true;
  }
  // with_flags: the energy kernel also makes the split decisions (q_split(true) follows)
  void q_energy(bool tuned = false, bool with_flags = false) {
    pbegin(SPH_P_ENERGY);
    uint *flags = with_flags ? flag_by_slot : nullptr;
    if (!erec_ok) {  // not straight after integrate: records from the current state
      launch_make_erec(cur, n, aux.erec, st);
      erec_ok = true;
    }
    auto go = [&](int c) {
      launch_energy(variant, P, cur, cstart, cend, n, st, nullptr, c, flags, aux.erec);
    };
    if (tuned) tune.run(1, st, go);
    else go(tune.energy);
    pend();
  }
  void q_split(bool flags_ready = false) {
    nbl_ok = false;
    erec_ok = false;
    pbegin(SPH_P_SPLIT);
    if (!flags_ready) launch_decide(variant, P, cur, n, flag_by_slot, st);
    sph_exclusive_scan(flag_by_slot, rank_by_slot, nullptr, n, &cnt->scan_total, scan_tmp, st);
    launch_split(variant, P, cur, n, flag_by_slot, rank_by_slot, cnt, status, st);
    n_flagged = n;
    pend();
  }
  void q_merge() {
    nbl_ok = false;
    erec_ok = false;
    pbegin(SPH_P_MERGE);
    if (variant == SPH_V2D) launch_merge_q2d(P, cur, s2p_cell, cnt, cstart, cend, q2d, st);
    // flag_by_slot (the split's flags, consumed by now) is the qualification bits' scratch
    launch_merge(variant, P, cur, cstart, cend, q2d, status, cnt, st, nullptr, skeys, n_sorted,
                 nullptr, nullptr, flag_by_slot, cap);
    pend();
  }
  void q_compact() {
    nbl_ok = false;
    pbegin(SPH_P_COMPACT);
    launch_alive_by_slot(cur, status, cnt, cap, alive_by_slot, st);
    sph_exclusive_scan(alive_by_slot, newslot, &cnt->n, cap, &cnt->scan_total, scan_tmp, st);
    launch_renumber(cur, status, newslot, cnt, cap, st);  // and n = survivors
    pend();
  }
  void q_dgrid() {
    pbegin(SPH_P_DGRID);
    launch_density_grid(P, cur, cstart, cend, dgrid, dgrid_cells, st, occ);
    pend();
  }
  void q_normals() {
    pbegin(SPH_P_NORMALS);
    launch_normals(P, dgrid, normal3, st);
    pend();
  }
  void q_mc() {
    pbegin(SPH_P_MC);
    launch_mc_count(P, dgrid, cube_count, seg_cube, st);
    sph_exclusive_scan(cube_count, cube_off, nullptr, (uint)ncube, &cnt->ntri, scan_tmp, st);
    sph_exclusive_scan(seg_cube, seg_coff, nullptr, (uint)ncube, &cnt->nmc, scan_tmp, st);
    launch_mc_list(P, dgrid, seg_cube, cube_off, seg_coff, max_tri, clist, st);
    launch_mc_emit(P, dgrid, normal3, clist, cnt, max_tri, tris, st);
    pend();
  }

  void clear_force_flags() {
    // force_waiting* and the 2D mouse force last one step (SPH-3D-2/sph_system.cpp:403)
    if (variant == SPH_V2D) ((SphSysParam2D *)raw)->force_waiting = false;
    else {
      ((SphSysParam3D *)raw)->flag0 = false;
      if (variant == SPH_V3D4) ((SphSysParam3D *)raw)->flag1 = false;
    }
    P.push_left = P.push_right = 0;
    P.mouse = 0;
  }
};

#define CHECK_CTX(c)            \
  do {                          \
    if (!(c)) return SPH_EINVAL; \
  } while (0)

extern "C" {

const char *sph_version(void) { return SPH_VERSION; }

int sph_param_defaults(int variant, void *out) {
  if (!valid_variant(variant) || !out) return SPH_EINVAL;
  if (variant == SPH_V2D) defaults2(*(SphSysParam2D *)out);
  else defaults3(variant, *(SphSysParam3D *)out);
  return SPH_OK;
}

int sph_param_derive(int variant, void *inout) {
  if (!valid_variant(variant) || !inout) return SPH_EINVAL;
  if (variant == SPH_V2D) {
    SphSysParam2D &s = *(SphSysParam2D *)inout;
    if (!(s.large_kernel > 0.0f) || !(s.small_kernel > 0.0f)) return SPH_EINVAL;
    derive2(s);
  } else {
    SphSysParam3D &s = *(SphSysParam3D *)inout;
    if (!(s.large_kernel > 0.0f) || !(s.small_kernel > 0.0f)) return SPH_EINVAL;
    derive3(s);
  }
  return SPH_OK;
}

int sph_ctx_create(int variant, const void *sysparam, uint32_t capacity, uint32_t max_triangles,
                   SphCtx **out) {
  if (!out) return SPH_EINVAL;
  *out = nullptr;
  // particle indices share a 32-bit word with a 5-bit cell-offset id in the pair lists
  if (!valid_variant(variant) || !sysparam || capacity == 0 || capacity > (1u << 27)) return SPH_EINVAL;
  SphCtx *c = new (std::nothrow) SphCtx();
  if (!c) return SPH_ECAP;
  c->variant = variant;
  c->raw_size = variant == SPH_V2D ? sizeof(SphSysParam2D) : sizeof(SphSysParam3D);
  std::memcpy(c->raw, sysparam, c->raw_size);
  c->cap = capacity;
  c->max_tri = max_triangles;
  c->P = make_dev_params(variant, c->raw, capacity);
  if (c->P.total_cell == 0) {
    delete c;
    return SPH_EINVAL;
  }
  c->kbits = key_bits(c->P.total_cell);
  c->dgrid_cells = dgrid_cell_blocks_exact(c->P);
  if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return SPH_EHIP;
  }
  const size_t cap = capacity;
  const size_t scan_n = std::max<size_t>(cap, sph_sort_hist_elems(cap));
  bool ok = c->alloc_state(c->cur) && c->alloc_state(c->alt) && c->alloc(c->keys, cap) &&
            c->alloc(c->vals, cap) && c->alloc(c->sc.keys_alt, cap) && c->alloc(c->sc.vals_alt, cap) &&
            c->alloc(c->sc.tile_hist, sph_sort_hist_elems(cap)) &&
            c->alloc(c->s2p_hash, cap) && c->alloc(c->s2p_cell, cap) &&
            c->alloc(c->cstart, c->P.total_cell) && c->alloc(c->cend, c->P.total_cell) &&
            c->alloc(c->flag_by_slot, cap) && c->alloc(c->rank_by_slot, cap) &&
            c->alloc(c->alive_by_slot, cap) && c->alloc(c->newslot, cap) &&
            c->alloc(c->status, cap) && c->alloc(c->q2d, std::max<size_t>(c->P.total_cell, 1)) &&
            c->alloc(c->cnt, 1) && c->alloc(c->aos_dev, cap * sizeof(SphParticle3D)) &&
            c->alloc(c->aux.frec, 3 * (size_t)cap) && c->alloc(c->aux.erec, cap) &&
            c->alloc(c->aux.nbl, (size_t)SPH_NBL_K * cap) && c->alloc(c->aux.nbl_cnt, cap);
  c->aux.nbl_stride = cap;
  if (ok && c->P.dim == 3 && c->P.tot_dens > 0) {
    c->ncube = mc_segments(c->P);
    ok = c->alloc(c->dgrid, c->P.tot_dens) && c->alloc(c->normal3, (size_t)c->P.tot_dens * 3) &&
         c->alloc(c->cube_count, c->ncube) && c->alloc(c->cube_off, c->ncube) &&
         c->alloc(c->seg_cube, c->ncube) && c->alloc(c->seg_coff, c->ncube) &&
         c->alloc(c->clist, (size_t)std::max<uint32_t>(max_triangles, 1)) &&
         c->alloc(c->tris, (size_t)std::max<uint32_t>(max_triangles, 1) * 18);
    if (ok) ok = c->alloc(c->occ, 2 * sph_occ_words(c->P.total_cell));
  }
  size_t scan_elems = sph_scan_tmp_elems(std::max(scan_n, c->ncube));
  ok = ok && c->alloc(c->scan_tmp, scan_elems);
  c->sc.scan_tmp = c->scan_tmp;
  c->sc.max_tiles = sph_sort_tiles(cap);
  if (ok && hipHostMalloc((void **)&c->cnt_host, sizeof(Counters), hipHostMallocDefault) != hipSuccess)
    ok = false;
  if (ok && !c->tune.create()) ok = false;
  c->tune.start(kTuneRounds, kTuneAfter, kTuneEvery);
  if (!ok) {
    delete c;
    return SPH_ECAP;
  }
  c->set_counts_on_device(0, 0);
  if (c->sync() != SPH_OK) {
    delete c;
    return SPH_EHIP;
  }
  *out = c;
  return SPH_OK;
}

int sph_ctx_destroy(SphCtx *ctx) {
  CHECK_CTX(ctx);
  delete ctx;
  return SPH_OK;
}

int sph_set_params(SphCtx *ctx, const void *sysparam) {
  CHECK_CTX(ctx);
  if (!sysparam) return SPH_EINVAL;
  DevParams P = make_dev_params(ctx->variant, sysparam, ctx->cap);
  if (P.total_cell != ctx->P.total_cell || P.tot_dens != ctx->P.tot_dens) return SPH_EINVAL;
  std::memcpy(ctx->raw, sysparam, ctx->raw_size);
  ctx->P = P;
  ctx->nbl_ok = false;  // the pair lists were made with the old kernel radii
  return SPH_OK;
}

int sph_get_params(SphCtx *ctx, void *out) {
  CHECK_CTX(ctx);
  if (!out) return SPH_EINVAL;
  std::memcpy(out, ctx->raw, ctx->raw_size);
  return SPH_OK;
}

int sph_upload_particles(SphCtx *ctx, const void *aos, uint32_t n) {
  CHECK_CTX(ctx);
  if (n > ctx->cap) return SPH_ECAP;
  if (n > 0 && !aos) return SPH_EINVAL;
  const size_t rec = ctx->variant == SPH_V2D ? sizeof(SphParticle2D) : sizeof(SphParticle3D);
  if (n > 0) {
    if (hipMemcpyAsync(ctx->aos_dev, aos, rec * n, hipMemcpyHostToDevice, ctx->st) != hipSuccess)
      return SPH_EHIP;
    launch_upload_scatter(ctx->variant, ctx->aos_dev, n, ctx->cur, ctx->st);
  }
  ctx->set_counts_on_device(n, n);
  ctx->hashed = ctx->sorted = ctx->lists_valid = ctx->merged = false;
  ctx->erec_ok = false;
  ctx->nbl_ok = false;
  return ctx->sync();
}

int sph_download_particles(SphCtx *ctx, void *aos, uint32_t n) {
  CHECK_CTX(ctx);
  if (n > ctx->n) return SPH_EINVAL;
  if (n == 0) return SPH_OK;
  if (!aos) return SPH_EINVAL;
  const size_t rec = ctx->variant == SPH_V2D ? sizeof(SphParticle2D) : sizeof(SphParticle3D);
  // slot -> physical map of the current numbering (scratch: newslot)
  launch_slot2phys(ctx->cur, ctx->m, ctx->newslot, ctx->st);
  launch_download_gather(ctx->variant, ctx->cur, ctx->newslot, n, ctx->aos_dev, ctx->st);
  if (hipMemcpyAsync(aos, ctx->aos_dev, rec * n, hipMemcpyDeviceToHost, ctx->st) != hipSuccess)
    return SPH_EHIP;
  return ctx->sync();
}

int sph_num_particles(SphCtx *ctx, uint32_t *n) {
  CHECK_CTX(ctx);
  if (!n) return SPH_EINVAL;
  *n = ctx->n;
  return SPH_OK;
}

int sph_calc_hash(SphCtx *ctx) {
  CHECK_CTX(ctx);
  if (ctx->n == 0) return SPH_OK;
  ctx->q_hash();
  int rc = ctx->sync();
  if (rc) return rc;
  ctx->hashed = true;
  ctx->sorted = ctx->lists_valid = ctx->merged = false;
  return ctx->cnt_host->err ? SPH_EDOMAIN : SPH_OK;
}

int sph_sort(SphCtx *ctx) {
  CHECK_CTX(ctx);
  if (ctx->n == 0) return SPH_OK;
  if (!ctx->hashed) return SPH_EINVAL;
  ctx->q_sort();
  int rc = ctx->sync();
  if (rc) return rc;
  ctx->sorted = true;
  return SPH_OK;
}

int sph_cell_ranges(SphCtx *ctx) {
  CHECK_CTX(ctx);
  if (ctx->n == 0) return SPH_OK;
  if (!ctx->sorted) return SPH_EINVAL;
  ctx->q_cells();
  int rc = ctx->sync();
  if (rc) return rc;
  ctx->lists_valid = true;
  ctx->hashed = false;
  return SPH_OK;
}

#define PHASE_BODY(ctx, need_lists, call)            \
  CHECK_CTX(ctx);                                    \
  if ((ctx)->n == 0) return SPH_EEMPTY;              \
  if ((need_lists) && !(ctx)->lists_valid) return SPH_EINVAL; \
  if ((ctx)->m != (ctx)->n) return SPH_EINVAL;       \
  (ctx)->call;                                       \
  return (ctx)->sync();

int sph_density(SphCtx *ctx) { PHASE_BODY(ctx, true, q_density()) }
int sph_force(SphCtx *ctx) { PHASE_BODY(ctx, true, q_force()) }
int sph_energy(SphCtx *ctx) { PHASE_BODY(ctx, true, q_energy()) }

int sph_integrate(SphCtx *ctx) {
  CHECK_CTX(ctx);
  if (ctx->n == 0) return SPH_OK;
  if (ctx->m != ctx->n) return SPH_EINVAL;
  ctx->q_integrate();
  return ctx->sync();
}

int sph_split(SphCtx *ctx, uint32_t *n_requested, uint32_t *n_done) {
  CHECK_CTX(ctx);
  if (n_requested) *n_requested = 0;
  if (n_done) *n_done = 0;
  if (ctx->n == 0) return SPH_OK;
  if (ctx->m != ctx->n) return SPH_EINVAL;
  launch_begin_step(ctx->cnt, ctx->st);  // resets split/merge counters, n_pre = n
  ctx->q_split();
  int rc = ctx->sync();
  if (rc) return rc;
  if (n_requested) *n_requested = ctx->cnt_host->s_req;
  if (n_done) *n_done = ctx->cnt_host->s_done;
  return SPH_OK;
}

int sph_merge(SphCtx *ctx, uint32_t *groups) {
  CHECK_CTX(ctx);
  if (groups) *groups = 0;
  if (!ctx->lists_valid) return SPH_EINVAL;
  ctx->q_merge();
  int rc = ctx->sync();
  if (rc) return rc;
  ctx->merged = true;
  if (groups) *groups = ctx->cnt_host->groups;
  return SPH_OK;
}

int sph_compact(SphCtx *ctx, uint32_t *n_out) {
  CHECK_CTX(ctx);
  if (!ctx->merged) return SPH_EINVAL;
  ctx->q_compact();
  int rc = ctx->sync();
  if (rc) return rc;
  ctx->merged = false;
  ctx->lists_valid = false;
  ctx->compacted = true;
  ctx->n_status = ctx->m;
  if (n_out) *n_out = ctx->n;
  return SPH_OK;
}

int sph_density_grid(SphCtx *ctx) {
  CHECK_CTX(ctx);
  if (ctx->P.dim != 3 || !ctx->dgrid) return SPH_EINVAL;
  if (!ctx->lists_valid) return SPH_EINVAL;
  ctx->q_dgrid();
  return ctx->sync();
}

int sph_grid_normals(SphCtx *ctx) {
  CHECK_CTX(ctx);
  if (ctx->P.dim != 3 || !ctx->dgrid) return SPH_EINVAL;
  ctx->q_normals();
  return ctx->sync();
}

int sph_marching_cubes(SphCtx *ctx, uint32_t *n_tri) {
  CHECK_CTX(ctx);
  if (ctx->P.dim != 3 || !ctx->dgrid) return SPH_EINVAL;
  ctx->q_mc();
  int rc = ctx->sync();
  if (rc) return rc;
  if (n_tri) *n_tri = ctx->cnt_host->ntri;
  return ctx->cnt_host->ntri > ctx->max_tri ? SPH_ETRI : SPH_OK;
}

int sph_step(SphCtx *ctx, uint32_t flags, uint32_t *n_out) {
  CHECK_CTX(ctx);
  if (ctx->n == 0) return SPH_EEMPTY;
  ctx->q_hash();
  ctx->q_sort();
  ctx->q_cells(true);
  ctx->q_density();
  ctx->q_force(true);
  ctx->q_integrate();
  if ((flags & SPH_STEP_MC) && ctx->P.dim == 3 && ctx->dgrid) {
    ctx->q_dgrid();
    ctx->q_normals();
    ctx->q_mc();
  }
  ctx->q_energy(true, (flags & SPH_STEP_SPLIT) != 0);
  if (flags & SPH_STEP_SPLIT) ctx->q_split(true);
  if (flags & SPH_STEP_MERGE) {
    ctx->q_merge();
    ctx->q_compact();
  }
  // the cell-range phase made physical == live before split appended children
  int rc = ctx->sync();
  if (rc == SPH_OK) ctx->tune.end_step();
  ctx->clear_force_flags();
  ctx->hashed = ctx->sorted = ctx->merged = false;
  ctx->lists_valid = !(flags & SPH_STEP_MERGE);
  if (rc) return rc;
  if (flags & SPH_STEP_MERGE) ctx->n_status = ctx->m;
  if (n_out) *n_out = ctx->n;
  return ctx->cnt_host->err ? SPH_EDOMAIN : SPH_OK;
}

static size_t count_of(SphCtx *ctx, int which) {
  switch (which) {
    case SPH_Q_HASH:
    case SPH_Q_INDEX: return ctx->skeys ? ctx->n_sorted : 0;
    case SPH_Q_START:
    case SPH_Q_END: return ctx->P.total_cell;
    case SPH_Q_SPLITFLAG: return ctx->n_flagged;
    case SPH_Q_STATUS: return ctx->n_status;
    case SPH_Q_COUNTERS: return sizeof(Counters) / 4;
    case SPH_Q_DGRID: return ctx->dgrid ? ctx->P.tot_dens : 0;
    case SPH_Q_NORMAL: return ctx->dgrid ? (size_t)ctx->P.tot_dens * 3 : 0;
    case SPH_Q_TRI: {
      uint nt = ctx->cnt_host->ntri < ctx->max_tri ? ctx->cnt_host->ntri : ctx->max_tri;
      return ctx->tris ? (size_t)nt * 18 : 0;
    }
  }
  return 0;
}

int sph_count(SphCtx *ctx, int which, size_t *n) {
  CHECK_CTX(ctx);
  if (!n) return SPH_EINVAL;
  *n = count_of(ctx, which);
  return SPH_OK;
}

// Reads an integer array; SPLITFLAG and STATUS use the reference's conventions
// (dev_split: 0 = split; dev_status: 1 = removed), indexed by slot.
int sph_get_u32(SphCtx *ctx, int which, uint32_t *out, size_t n) {
  CHECK_CTX(ctx);
  if (n == 0) return SPH_OK;
  if (!out || n > count_of(ctx, which)) return SPH_EINVAL;
  const uint *src = nullptr;
  switch (which) {
    case SPH_Q_HASH: src = ctx->skeys; break;
    case SPH_Q_INDEX: src = ctx->svals; break;
    case SPH_Q_START: src = ctx->cstart; break;
    case SPH_Q_END: src = ctx->cend; break;
    case SPH_Q_SPLITFLAG: src = ctx->flag_by_slot; break;
    case SPH_Q_STATUS: src = ctx->alive_by_slot; break;
    case SPH_Q_COUNTERS: src = (const uint *)ctx->cnt; break;
    default: return SPH_EINVAL;
  }
  if (hipMemcpyAsync(out, src, n * 4, hipMemcpyDeviceToHost, ctx->st) != hipSuccess) return SPH_EHIP;
  int rc = ctx->sync();
  if (rc) return rc;
  if (which == SPH_Q_SPLITFLAG || which == SPH_Q_STATUS)
    for (size_t k = 0; k < n; k++) out[k] = out[k] ? 0u : 1u;
  return SPH_OK;
}

int sph_get_f32(SphCtx *ctx, int which, float *out, size_t n) {
  CHECK_CTX(ctx);
  if (n == 0) return SPH_OK;
  if (!out || n > count_of(ctx, which)) return SPH_EINVAL;
  const float *src = nullptr;
  switch (which) {
    case SPH_Q_DGRID: src = ctx->dgrid; break;
    case SPH_Q_NORMAL: src = ctx->normal3; break;
    case SPH_Q_TRI: src = ctx->tris; break;
    default: return SPH_EINVAL;
  }
  if (hipMemcpyAsync(out, src, n * 4, hipMemcpyDeviceToHost, ctx->st) != hipSuccess) return SPH_EHIP;
  return ctx->sync();
}

int sph_autotune(SphCtx *ctx, uint32_t rounds, uint32_t after_steps, uint32_t every_steps) {
  CHECK_CTX(ctx);
  ctx->tune.start(rounds, after_steps, every_steps);
  return SPH_OK;
}

int sph_set_lists(SphCtx *ctx, uint32_t force_list, uint32_t energy_list) {
  CHECK_CTX(ctx);
  return ctx->tune.set(force_list, energy_list) ? SPH_OK : SPH_EINVAL;
}

int sph_tuned_lists(SphCtx *ctx, uint32_t *force_list, uint32_t *energy_list, float *best_ms) {
  CHECK_CTX(ctx);
  if (force_list) *force_list = (uint32_t)kForceLists[ctx->tune.force];
  if (energy_list) *energy_list = (uint32_t)kEnergyLists[ctx->tune.energy];
  if (best_ms)
    for (int k = 0; k < 2; k++)
      for (int c = 0; c < 3; c++) best_ms[k * 3 + c] = ctx->tune.best[k][c];
  return SPH_OK;
}

int sph_prof_enable(SphCtx *ctx, int on) {
  CHECK_CTX(ctx);
  ctx->prof_mask = on ? (1u << SPH_P_COUNT) - 1u : 0u;
  return SPH_OK;
}

int sph_prof_phases(SphCtx *ctx, uint32_t phase_mask) {
  CHECK_CTX(ctx);
  ctx->prof_mask = phase_mask & ((1u << SPH_P_COUNT) - 1u);
  return SPH_OK;
}

int sph_prof_reset(SphCtx *ctx) {
  CHECK_CTX(ctx);
  int rc = ctx->sync();
  for (int k = 0; k < SPH_P_COUNT; k++) {
    ctx->prof_ms[k] = 0.0;
    ctx->prof_n[k] = 0;
  }
  return rc;
}

int sph_prof_read(SphCtx *ctx, int phase, double *ms_total, uint64_t *launches) {
  CHECK_CTX(ctx);
  if (phase < 0 || phase >= SPH_P_COUNT) return SPH_EINVAL;
  if (ms_total) *ms_total = ctx->prof_ms[phase];
  if (launches) *launches = ctx->prof_n[phase];
  return SPH_OK;
}

int sph_synchronize(SphCtx *ctx) {
  CHECK_CTX(ctx);
  return ctx->sync();
}

}  // extern "C"
