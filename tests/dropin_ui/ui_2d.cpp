// Compile-and-link check of the 2D drop-in against the way the reference's 2D GLUT front end
// uses SPHSystem (SPH-2D/sph_main.cpp): every member it reads or writes, through the
// reference's names (SPHSystem, SysParam, Particle). Built by tests/test_dropin_boundary.py;
// never run (it needs a GPU).
#include <cstdio>

#include "../../include/sph_system_2d_amd.h"

static SPHSystem *sph_system;

// init_sph_system (SPH-2D/sph_main.cpp:40-46)
static void init_sph_system() {
  sph_system = new SPHSystem();
  sph_system->add_init_particle();
}

// display_func (SPH-2D/sph_main.cpp:55-135)
static float display_func() {
  float acc = sph_system->m_world_width + sph_system->m_world_depth;  // gluOrtho2D, :59
  for (uint count = 0; count < sph_system->m_num_particle; count++) {
    Particle *p = &(sph_system->m_host_mem[count]);
    if (p->surface_normal > sph_system->m_surface_normal) acc += 1.0f;  // :71
    if (sph_system->m_co_energy * p->energy + sph_system->m_co_surface * p->surface_normal >
        sph_system->m_thresh_split)  // :92
      acc += 1.0f;
    acc += p->pos.x + p->pos.y + (float)p->level;
  }
  SphFloat2 pos = {0.5f, 0.5f}, extern_force = {1.0f, 0.0f};
  sph_system->add_extern_force(pos, extern_force);  // :122
  sph_system->animate();                            // :133
  return acc;
}

// keyboard_func (SPH-2D/sph_main.cpp:175-250)
static void keyboard_func(unsigned char key) {
  if (key == ' ') sph_system->m_sys_running = 1 - sph_system->m_sys_running;     // :181
  if (key == 's') sph_system->m_enable_split = 1 - sph_system->m_enable_split;   // :239
  if (key == 'm') sph_system->m_enable_merge = 1 - sph_system->m_enable_merge;   // :243
  if (key == 'f') sph_system->m_host_param->force_waiting = true;
}

static_assert(sizeof(SysParam) == 200, "SysParam layout");
static_assert(sizeof(Particle) == 56, "Particle layout");

int main(int argc, char **) {
  init_sph_system();
  float acc = 0.0f;
  for (int k = 0; k < argc; k++) {
    keyboard_func((unsigned char)(" smf"[k % 4]));
    acc += display_func();
  }
  delete sph_system;
  return acc > 0.0f ? 0 : 1;
}
