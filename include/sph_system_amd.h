// sph_system_amd.h -- C++ drop-in for the reference's 3D host class SPHSystem
// (SPH-3D-2/sph_system.h:26-154, SPH-3D-4/sph_system.h:26-154).
//
// The reference builds each variant as its own program, with its own SysParam (the trailing
// bools differ: SPH-3D-2/sph_param.h:85-86, SPH-3D-4/sph_param.h:85-87). Here both variants
// are one class template, SphSystem3D<SPH_V3D2> and SphSystem3D<SPH_V3D4>, instantiated in
// libsph_amd.so; an application written against the reference includes this header, with
// SPH_DROPIN_3D4 defined for the 3D-4 program, and gets the reference's names:
//   SPHSystem  (SphSystem3D of the variant)      SysParam (SphSysParam3D2 / SphSysParam3D4)
//   Particle   (SphParticle3D)                   Sphere   (SphSphere)
// with the same public fields, the same constructor and methods, and the same sequence of
// host-visible effects: animation() steps the device state and refreshes m_host_mem (and,
// when m_disp_mode == 2, the density grid, its normals and the triangle soup) where the
// reference reads them back. Define SPH_DROPIN_NO_ALIASES to get the class templates only
// (for code that also includes the 2D drop-in, sph_system_2d_amd.h).
//
// Differences, by design:
//   * marching cubes runs on the GPU inside animation() when m_disp_mode == 2;
//     marching_cube() copies the last triangle soup to the host arrays;
//   * the split step never exceeds m_max_particle (the reference does not check);
//   * the device buffers are owned by a context (SphCtx, sph_amd.h) in a structure-of-arrays
//     layout, so the reference's device pointers (m_dev_mem, m_dev_hash, ..., m_trans_vel)
//     have no counterpart and are null; the reference's UI never reads them;
//   * m_num_lines stays 0 as in the reference (its line recording is commented out,
//     SPH-3D-2/sph_system.cpp:503-513), and m_host_line0/1 hold one zero entry each;
//   * like the reference, animation() sends m_host_param (not the m_* mirrors of its
//     fields, which the constructor fills and nothing re-reads);
//   * errors are reported through last_status() instead of being ignored.
#ifndef SPH_SYSTEM_AMD_H
#define SPH_SYSTEM_AMD_H

#include "sph_amd.h"

typedef unsigned int uint;
typedef sph_float3 SphFloat3;

template <int V> struct SphSysParamOf;
template <> struct SphSysParamOf<SPH_V3D2> { typedef SphSysParam3D2 type; };
template <> struct SphSysParamOf<SPH_V3D4> { typedef SphSysParam3D4 type; };

// Everything of SPH-3D-2/sph_system.h:26-154 but the parameter block.
class SphSystem3DBase {
 public:
  SphParticle3D *m_host_mem;
  SphParticle3D *m_dev_mem;  // null (see above)

  uint m_num_particle;
  uint m_max_particle;

  float m_world_width, m_world_height, m_world_length;
  float m_large_radius, m_small_radius;
  float m_large_kernel, m_small_kernel;
  float m_large_mass, m_small_mass;

  float m_cell_size;
  uint m_row_cell, m_col_cell, m_len_cell, m_total_cell;

  uint *m_dev_hash, *m_dev_index, *m_cell_start, *m_cell_end;  // null (see above)
  uint *m_dev_status;                                          // null
  SphFloat3 *m_trans_vel;                                      // null
  uint *m_dev_split, *m_dev_split_index, *m_dev_merge;         // null

  SphFloat3 m_gravity;
  float m_wall_damping, m_rest_density, m_gas_constant, m_time_step, m_viscosity;
  float m_surface_tension;
  float m_co_surface, m_co_energy;
  float m_split_criteria, m_merge_criteria;

  // kernel constants of the constructor (SPH-3D-2/sph_system.cpp:63-92)
  float m_large_poly6, m_large_spiky, m_large_visco;
  float m_large_grad_poly6, m_large_lapc_poly6;
  float m_small_poly6, m_small_spiky, m_small_visco;
  float m_small_grad_poly6, m_small_lapc_poly6;
  float m_large_kernel_2, m_large_kernel_6;
  float m_small_kernel_2, m_small_kernel_6;
  float m_large_poly6_radius, m_small_poly6_radius;

  uint m_dens_mul;
  float m_den_size;
  uint m_row_dens, m_col_dens, m_len_dens, m_tot_dens;
  float *m_host_dens;
  float *m_dev_dens;  // null
  SphFloat3 *m_host_dens_pos;
  SphFloat3 *m_dev_dens_pos;  // null
  SphFloat3 *m_host_dens_normal;
  SphFloat3 *m_dev_dens_normal;  // null

  uint m_num_lines;
  SphFloat3 *m_host_line0, *m_host_line1;

  uint m_num_triangle;
  SphFloat3 *m_host_triangle0, *m_host_triangle1, *m_host_triangle2;
  SphFloat3 *m_host_triangle_normal0, *m_host_triangle_normal1, *m_host_triangle_normal2;

  SphSphere m_sphere1, m_sphere2, m_sphere3, m_sphere4;

  uint m_sys_running;
  uint m_trans_once;
  uint m_use_split;
  uint m_use_merge;
  uint m_disp_mode;  // 1 particles, 2 density grid + marching cubes

  // additions: the last C-ABI status, the variant, the device context
  int last_status() const { return m_status; }
  int variant() const { return m_variant; }
  SphCtx *context() { return m_ctx; }

  SphSystem3DBase(const SphSystem3DBase &) = delete;
  SphSystem3DBase &operator=(const SphSystem3DBase &) = delete;

 protected:
  SphSystem3DBase() = default;
  ~SphSystem3DBase() = default;
  int m_variant = SPH_V3D2;
  int m_status = SPH_OK;
  uint m_max_triangle = 0;
  SphCtx *m_ctx = nullptr;
};

template <int V> class SphSystem3D : public SphSystem3DBase {
 public:
  typedef typename SphSysParamOf<V>::type SysParamT;
  SysParamT *m_host_param;

  // SPHSystem::SPHSystem (SPH-3D-2/sph_system.cpp:6-229; SPH-3D-4 :6-230); max_particle is
  // the reference's m_max_particle (1000000)
  explicit SphSystem3D(uint max_particle = 1000000);
  ~SphSystem3D();

  void add_box_particle();  // SPH-3D-2/sph_system.cpp:247-264 (3D-4 :248-289)
  void add_new_particle();  // SPH-3D-2/sph_system.cpp:266-283 (3D-4 :291-331)
  void animation();         // SPH-3D-2/sph_system.cpp:302-404 (3D-4 :350-453)
  void add_particle(SphFloat3 pos, SphFloat3 vel, uint level);  // :285-300
  void marching_cube();     // SPH-3D-2/sph_system.cpp:517-542

 private:
  void upload();
};

extern template class SphSystem3D<SPH_V3D2>;
extern template class SphSystem3D<SPH_V3D4>;

#ifndef SPH_DROPIN_NO_ALIASES
#ifdef SPH_DROPIN_3D4
typedef SphSystem3D<SPH_V3D4> SPHSystem;
typedef SphSysParam3D4 SysParam;
#else
typedef SphSystem3D<SPH_V3D2> SPHSystem;
typedef SphSysParam3D2 SysParam;
#endif
typedef SphParticle3D Particle;
typedef SphSphere Sphere;
#endif

// ---- host output of the reference application (SPH-3D-2/sph_main.cpp) ----
// Viewer frame: a 30 x 20 x 20 box centred at the origin (sph_main.cpp:280-286) and
// sim_ratio = side / world size (sph_main.cpp:343-345); a point p is drawn at
// p * ratio + origin in float arithmetic.
struct SphViewFrame {
  SphFloat3 origin, ratio;
};
SphViewFrame sph_view_frame(const SphSystem3DBase &s);
// "Povray_Shots/000000.pov" numbering of the frame writers (frame n in the six digits).
void sph_pov_frame_name(int n, char out[24]);
// make_particle (sph_main.cpp:174-220): one sphere per particle of m_host_mem, radius 0.3
// red for level 1 and 0.24 yellow for level 2, then the checker floor plane.
// make_mesh (sph_main.cpp:114-172): one smooth_triangle per triangle of the last
// marching_cube(), vertices in the viewer frame, normals as computed, then the material
// and the floor plane. Both write POV-Ray scene text; return 0, or -1 when the file cannot
// be written.
int sph_write_pov_particles(const SphSystem3DBase &s, const char *path, const SphViewFrame &f);
int sph_write_pov_mesh(const SphSystem3DBase &s, const char *path, const SphViewFrame &f);

#endif
