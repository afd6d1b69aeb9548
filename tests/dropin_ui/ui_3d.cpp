// Compile-and-link check of the 3D drop-in against the way the reference's GLUT front ends
// use SPHSystem (SPH-3D-2/sph_main.cpp, SPH-3D-4/sph_main.cpp): every member they read or
// write, through the reference's names (SPHSystem, SysParam, Particle, Sphere). Built twice
// by tests/test_dropin_boundary.py, without and with SPH_DROPIN_3D4; never run (it needs a
// GPU). Each group cites the reference lines whose accesses it repeats.
#include <cstdio>

#include "../../include/sph_system_amd.h"

static SPHSystem *sph_system;
static SphFloat3 real_world_origin = {-15.0f, -10.0f, -10.0f};
static SphFloat3 sim_ratio = {1.0f, 1.0f, 1.0f};

// init_sph_system (SPH-3D-2/sph_main.cpp:278-323; SPH-3D-4 :304-349)
static void init_sph_system() {
  sph_system = new SPHSystem();
  printf("World Size: %f, %f, %f\n", sph_system->m_world_width, sph_system->m_world_height,
         sph_system->m_world_length);
  printf("%f %f %f\n", sph_system->m_cell_size, sph_system->m_large_kernel, sph_system->m_small_kernel);
  printf("%u %u %u %u\n", sph_system->m_row_cell, sph_system->m_col_cell, sph_system->m_len_cell,
         sph_system->m_total_cell);
  printf("%f %f %f\n", sph_system->m_large_poly6, sph_system->m_large_spiky, sph_system->m_large_visco);
  printf("%f %f %f\n", sph_system->m_small_poly6, sph_system->m_small_spiky, sph_system->m_small_visco);
  printf("%f %f\n", sph_system->m_large_kernel_2, sph_system->m_small_kernel_2);
  printf("%u %u %u %f %u\n", sph_system->m_row_dens, sph_system->m_col_dens, sph_system->m_len_dens,
         sph_system->m_den_size, sph_system->m_tot_dens);
  sph_system->add_box_particle();
  printf("New particle: %u\n", sph_system->m_num_particle);
  sim_ratio.x = 30.0f / sph_system->m_world_width;  // :343-345
  sim_ratio.y = 20.0f / sph_system->m_world_height;
  sim_ratio.z = 20.0f / sph_system->m_world_length;
}

// render_particles (SPH-3D-2/sph_main.cpp:347-400): levels, surface and energy colouring
static float render_particles() {
  float acc = 0.0f;
  for (uint count = 0; count < sph_system->m_num_particle; count++) {
    const Particle &p = sph_system->m_host_mem[count];
    if (p.level == 1 && p.surface > sph_system->m_surface_tension) acc += 1.0f;
    if (p.energy > sph_system->m_host_param->split_energy) acc += 1.0f;
    if (p.energy < sph_system->m_host_param->merge_energy) acc += 1.0f;
    acc += p.pos.x * sim_ratio.x + real_world_origin.x + p.pos.y * sim_ratio.y + p.pos.z * sim_ratio.z;
  }
  return acc;
}

// render_dens / render_dens_pos (SPH-3D-2/sph_main.cpp:419-468)
static float render_dens() {
  float acc = 0.0f;
  for (float count_x = 0; count_x < sph_system->m_row_dens; count_x++)
    for (float count_y = 0; count_y < sph_system->m_col_dens; count_y++)
      for (float count_z = 0; count_z < sph_system->m_len_dens; count_z++) {
        uint index = count_z * sph_system->m_row_dens * sph_system->m_col_dens +
                     count_y * sph_system->m_row_dens + count_x;
        if (sph_system->m_host_dens[index] == 0.0f) continue;
        acc += sph_system->m_host_dens_pos[index].x * sim_ratio.x + real_world_origin.x;
        acc += count_x / sph_system->m_row_dens * sph_system->m_world_width;
        acc += sph_system->m_host_dens[index] * 0.2f;
      }
  return acc;
}

// render_mesh / render_lines (SPH-3D-2/sph_main.cpp:470-520), make_mesh (:114-172)
static float render_mesh() {
  float acc = 0.0f;
  for (uint count = 0; count < sph_system->m_num_triangle; count++) {
    acc += sph_system->m_host_triangle0[count].x + sph_system->m_host_triangle1[count].y +
           sph_system->m_host_triangle2[count].z;
    acc += sph_system->m_host_triangle_normal0[count].x + sph_system->m_host_triangle_normal1[count].y +
           sph_system->m_host_triangle_normal2[count].z;
  }
  for (uint count = 0; count < sph_system->m_num_lines; count++)
    acc += sph_system->m_host_line0[count].x + sph_system->m_host_line1[count].y;
  for (uint count = 0; count < sph_system->m_tot_dens; count++)
    acc += sph_system->m_host_dens_normal[count].x;
  return acc;
}

// display_func (SPH-3D-2/sph_main.cpp:574-720; SPH-3D-4 :605-700): obstacle spheres, the
// step, display modes
static float display_func() {
  const Sphere *s[4] = {&sph_system->m_sphere1, &sph_system->m_sphere2, &sph_system->m_sphere3,
                        &sph_system->m_sphere4};
  float acc = 0.0f;
  for (const Sphere *q : s) acc += q->pos.x * sim_ratio.x + q->radius * sim_ratio.x;
  sph_system->animation();
  if (sph_system->m_disp_mode == 1) acc += render_particles();
  if (sph_system->m_disp_mode == 2) acc += render_dens() + render_mesh();
  if (sph_system->m_host_param->use_cylinder == true) acc += 1.0f;  // :707 (3D-4 :692)
  return acc;
}

// keyboard_func / menu (SPH-3D-2/sph_main.cpp:780-880; SPH-3D-4 :772-888)
static void keyboard_func(unsigned char key) {
  if (key == '1') sph_system->add_new_particle();
  if (key == '2') sph_system->m_trans_once = 1;
  if (key == ' ') sph_system->m_sys_running = 1 - sph_system->m_sys_running;
#ifdef SPH_DROPIN_3D4
  if (key == 'y') sph_system->m_host_param->force_waiting_left = true;   // SPH-3D-4 :797
  if (key == 'u') sph_system->m_host_param->force_waiting_right = true;  // SPH-3D-4 :802
#else
  if (key == 'y') sph_system->m_host_param->force_waiting = true;  // SPH-3D-2 :812
#endif
  if (key == 'c') sph_system->m_host_param->use_cylinder ^= true;  // :876 (3D-4 :866)
  if (key == 's') sph_system->m_use_split = 1 - sph_system->m_use_split;
  if (key == 'm') sph_system->m_use_merge = 1 - sph_system->m_use_merge;
  if (key == '5') sph_system->m_disp_mode = 2;
  if (key == '6') sph_system->m_disp_mode = 1;
  if (key == 'r') {
    delete sph_system;
    init_sph_system();
  }
}

// the parameter block keeps the reference's layout
static_assert(sizeof(SysParam) == 296, "SysParam layout");
static_assert(sizeof(Particle) == 72, "Particle layout");

int main(int argc, char **) {
  init_sph_system();
  float acc = 0.0f;
  for (int k = 0; k < argc; k++) {
    keyboard_func((unsigned char)(" 12fyucsm56"[k % 11]));
    acc += display_func();
  }
  sph_system->marching_cube();
  delete sph_system;
  return acc > 0.0f ? 0 : 1;
}
