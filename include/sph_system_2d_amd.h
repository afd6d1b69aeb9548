// sph_system_2d_amd.h -- C++ drop-in for the reference's 2D host class SPHSystem
// (SPH-2D/sph_system.h:24-106).
//
// SphSystem2D has the public fields, constructor and methods of the reference's 2D class;
// an application written against the reference includes this header and gets the reference's
// names SPHSystem (= SphSystem2D), Particle (SphParticle2D) and SysParam (SphSysParam2D).
// animate() steps the device state through the C ABI of sph_amd.h and refreshes m_host_mem,
// as the reference's per-step readback does (SPH-2D/sph_system.cpp:260). Define
// SPH_DROPIN_NO_ALIASES to get SphSystem2D only (for code that also includes the 3D
// drop-in, sph_system_amd.h).
//
// Differences, by design (as in the 3D drop-in): the reference's device pointers (m_dev_mem,
// m_dev_hash, ..., m_trans_vel) are null, the device state being owned by a context in a
// structure-of-arrays layout; the split step never exceeds m_max_particle; add_particle
// stops at m_max_particle instead of writing past m_host_mem; errors are reported through
// last_status().
#ifndef SPH_SYSTEM_2D_AMD_H
#define SPH_SYSTEM_2D_AMD_H

#include "sph_amd.h"

typedef unsigned int uint;
typedef sph_float2 SphFloat2;

class SphSystem2D {
 public:
  uint m_num_particle;
  uint m_max_particle;

  float m_world_width, m_world_depth;
  float m_large_mass, m_small_mass;
  float m_large_kernel, m_small_kernel;
  float m_large_kernel2, m_small_kernel2;
  float m_large_kernel6, m_small_kernel6;
  float m_large_radius, m_small_radius;
  float m_large_poly6, m_small_poly6;
  float m_large_spiky, m_small_spiky;
  float m_large_visco, m_small_visco;
  float m_large_grad_poly6, m_small_grad_poly6;
  float m_large_lapc_poly6, m_small_lapc_poly6;

  SphFloat2 m_gravity;
  float m_wall_damping, m_rest_density, m_gas_constant, m_viscosity, m_time_step;
  float m_surface_normal;
  uint m_sys_running;

  float m_cell_size;
  uint m_row_cell, m_col_cell, m_total_cell;

  float m_co_surface, m_co_energy;
  float m_thresh_split, m_thresh_merge;

  SphParticle2D *m_host_mem;
  SphParticle2D *m_dev_mem;                                   // null
  uint *m_dev_hash, *m_dev_index, *m_cell_start, *m_cell_end;  // null
  SphFloat2 *m_trans_vel;                                     // null
  uint *m_dev_status, *m_dev_split, *m_dev_merge, *m_dev_split_index;  // null

  uint m_enable_split;
  uint m_enable_merge;

  SphSysParam2D *m_host_param;

  // SPHSystem::SPHSystem (SPH-2D/sph_system.cpp:5-146); the reference's m_max_particle is
  // 100000
  explicit SphSystem2D(uint max_particle = 100000);
  ~SphSystem2D();
  SphSystem2D(const SphSystem2D &) = delete;
  SphSystem2D &operator=(const SphSystem2D &) = delete;

  void add_init_particle();                                  // SPH-2D/sph_system.cpp:148-162
  void add_particle(SphFloat2 pos, SphFloat2 vel, uint level);  // :164-177
  void animate();                                            // :190-264
  // the mouse force of the next step, pos in [0, 1]^2 of the window (:179-188)
  void add_extern_force(SphFloat2 pos, SphFloat2 force);

  // additions: the last C-ABI status and the device context
  int last_status() const { return m_status; }
  SphCtx *context() { return m_ctx; }

 private:
  int m_status;
  SphCtx *m_ctx;
};

#ifndef SPH_DROPIN_NO_ALIASES
typedef SphSystem2D SPHSystem;
typedef SphParticle2D Particle;
typedef SphSysParam2D SysParam;
#endif

#endif
