// sph_system_amd.h -- C++ drop-in for the reference's host class SPHSystem
// (SPH-3D-2/sph_system.h:26-154, SPH-3D-4/sph_system.h:26-154).
//
// Same class name, constructor defaults, public fields and methods, so an application
// written against the reference (GLUT front end, POV-Ray writers) compiles against this
// header and links libsph_amd.so. The step runs on the GPU through the C ABI of
// sph_amd.h; m_host_mem, the density grid and the triangle soup are refreshed after each
// animation() exactly where the reference reads them back.
//
// Differences, by design:
//   * the variant (3D-2: 1->2 split; 3D-4: 1->4 split) is a constructor argument;
//   * marching cubes runs on the GPU inside animation() when m_disp_mode == 2;
//     marching_cube() only copies the last triangle soup to the host arrays;
//   * the split step never exceeds m_max_particle (the reference does not check);
//   * errors are reported through last_status() instead of being ignored.
#ifndef SPH_SYSTEM_AMD_H
#define SPH_SYSTEM_AMD_H

#include "sph_amd.h"

typedef unsigned int uint;
typedef sph_float3 SphFloat3;
typedef SphParticle3D Particle;
typedef SphSysParam3D SysParam;
typedef SphSphere Sphere;

class SPHSystem {
 public:
  Particle *m_host_mem;

  uint m_num_particle;
  uint m_max_particle;

  float m_world_width, m_world_height, m_world_length;
  float m_large_radius, m_small_radius;
  float m_large_kernel, m_small_kernel;
  float m_large_mass, m_small_mass;

  float m_cell_size;
  uint m_row_cell, m_col_cell, m_len_cell, m_total_cell;

  SphFloat3 m_gravity;
  float m_wall_damping, m_rest_density, m_gas_constant, m_time_step, m_viscosity;
  float m_surface_tension;
  float m_co_surface, m_co_energy;
  float m_split_criteria, m_merge_criteria;

  uint m_dens_mul;
  float m_den_size;
  uint m_row_dens, m_col_dens, m_len_dens, m_tot_dens;
  float *m_host_dens;
  SphFloat3 *m_host_dens_pos;
  SphFloat3 *m_host_dens_normal;

  uint m_num_triangle;
  SphFloat3 *m_host_triangle0, *m_host_triangle1, *m_host_triangle2;
  SphFloat3 *m_host_triangle_normal0, *m_host_triangle_normal1, *m_host_triangle_normal2;

  Sphere m_sphere1, m_sphere2, m_sphere3, m_sphere4;

  SysParam *m_host_param;

  uint m_sys_running;
  uint m_trans_once;
  uint m_use_split;
  uint m_use_merge;
  uint m_disp_mode;  // 1 particles, 2 density grid + marching cubes

 public:
  // variant: SPH_V3D2 (default, the constants of SPH-3D-2/sph_system.cpp:6-229) or SPH_V3D4
  explicit SPHSystem(int variant = SPH_V3D2, uint max_particle = 1000000);
  ~SPHSystem();
  SPHSystem(const SPHSystem &) = delete;
  SPHSystem &operator=(const SPHSystem &) = delete;

  void add_box_particle();  // SPH-3D-2/sph_system.cpp:247-264 (3D-4 :248-289)
  void add_new_particle();  // SPH-3D-2/sph_system.cpp:266-283 (3D-4 :291-331)
  void animation();         // SPH-3D-2/sph_system.cpp:302-404
  void add_particle(SphFloat3 pos, SphFloat3 vel, uint level);  // :285-300
  void marching_cube();     // SPH-3D-2/sph_system.cpp:517-542

  // additions: the parameter block changed through the public fields is re-derived and
  // sent on the next animation(); last_status() is the last C-ABI status.
  int last_status() const { return m_status; }
  int variant() const { return m_variant; }
  SphCtx *context() { return m_ctx; }

 private:
  void sync_params_from_fields();
  void upload();
  int m_variant;
  int m_status;
  uint m_max_triangle;
  SphCtx *m_ctx;
};

// ---- host output of the reference application (SPH-3D-2/sph_main.cpp) ----
// Viewer frame: a 30 x 20 x 20 box centred at the origin (sph_main.cpp:280-286) and
// sim_ratio = side / world size (sph_main.cpp:343-345); a point p is drawn at
// p * ratio + origin in float arithmetic.
struct SphViewFrame {
  SphFloat3 origin, ratio;
};
SphViewFrame sph_view_frame(const SPHSystem &s);
// "Povray_Shots/000000.pov" numbering of the frame writers (frame n in the six digits).
void sph_pov_frame_name(int n, char out[24]);
// make_particle (sph_main.cpp:174-220): one sphere per particle of m_host_mem, radius 0.3
// red for level 1 and 0.24 yellow for level 2, then the checker floor plane.
// make_mesh (sph_main.cpp:114-172): one smooth_triangle per triangle of the last
// marching_cube(), vertices in the viewer frame, normals as computed, then the material
// and the floor plane. Both write POV-Ray scene text; return 0, or -1 when the file cannot
// be written.
int sph_write_pov_particles(const SPHSystem &s, const char *path, const SphViewFrame &f);
int sph_write_pov_mesh(const SPHSystem &s, const char *path, const SphViewFrame &f);

#endif
