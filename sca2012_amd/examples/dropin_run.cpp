// Drives the C++ drop-in SPHSystem the way the reference's GLUT applications do
// (SPH-3D-2/sph_main.cpp:278-323 init, display_func -> animation() per frame; SPH-2D/
// sph_main.cpp:40-133 with animate()), without a window: prints the live count and a
// checksum per step, and the triangle count of display-mode-2 steps. With an output
// directory (not "-") a 3D run also writes the reference's POV-Ray frames there:
// make_particle for every step (frame = step) and make_mesh for every marching-cubes step
// (frame = step, "mesh_" prefix). new_step >= 0 calls add_new_particle() before that step
// (the reference's '1' key, SPH-3D-2/sph_main.cpp:783-786).
// Usage: sph_dropin_run <variant 0|1|2> <steps> <split 0|1> <merge 0|1> <mc_every>
//                       [pov_dir|-] [new_step]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#define SPH_DROPIN_NO_ALIASES
#include "../../include/sph_system_2d_amd.h"
#include "../../include/sph_system_amd.h"

namespace {

struct Args {
  int steps, split, merge, mc_every, new_step;
  const char *pov_dir;
};

template <int V> int run3d(const Args &a) {
  SphSystem3D<V> sys(1000000);
  if (sys.last_status() != SPH_OK) {
    fprintf(stderr, "context creation failed: %d\n", sys.last_status());
    return 1;
  }
  sys.add_box_particle();
  printf("init %u\n", sys.m_num_particle);
  sys.m_sys_running = 1;
  sys.m_use_split = a.split;
  sys.m_use_merge = a.merge;
  for (int s = 0; s < a.steps; s++) {
    if (s == a.new_step) {
      sys.add_new_particle();
      printf("new %u\n", sys.m_num_particle);
    }
    sys.m_disp_mode = (a.mc_every > 0 && (s + 1) % a.mc_every == 0) ? 2 : 1;
    sys.animation();
    if (sys.last_status() != SPH_OK) {
      fprintf(stderr, "step %d failed: %d\n", s, sys.last_status());
      return 1;
    }
    double cs = 0.0;
    for (uint i = 0; i < sys.m_num_particle; i++)
      cs += sys.m_host_mem[i].pos.x + 2.0 * sys.m_host_mem[i].pos.y + 3.0 * sys.m_host_mem[i].pos.z;
    printf("step %d n %u checksum %.9e", s, sys.m_num_particle, cs);
    if (sys.m_disp_mode == 2) printf(" triangles %u", sys.m_num_triangle);
    printf("\n");
    if (a.pov_dir) {
      const SphViewFrame f = sph_view_frame(sys);
      char name[24];
      sph_pov_frame_name(s, name);
      const std::string base = std::string(a.pov_dir) + "/";
      if (sph_write_pov_particles(sys, (base + name).c_str(), f) != 0) return 1;
      if (sys.m_disp_mode == 2) {
        sys.marching_cube();
        const std::string nm(name);
        const std::string mesh = base + nm.substr(0, 13) + "mesh_" + nm.substr(13);
        if (sph_write_pov_mesh(sys, mesh.c_str(), f) != 0) return 1;
      }
    }
  }
  return 0;
}

int run2d(const Args &a) {
  SphSystem2D sys;
  if (sys.last_status() != SPH_OK) {
    fprintf(stderr, "context creation failed: %d\n", sys.last_status());
    return 1;
  }
  sys.add_init_particle();
  printf("init %u\n", sys.m_num_particle);
  sys.m_sys_running = 1;
  sys.m_enable_split = a.split;
  sys.m_enable_merge = a.merge;
  for (int s = 0; s < a.steps; s++) {
    if (s == a.new_step) {  // a mouse drag at (0.3, 0.6) of the window (SPH-2D/sph_main.cpp:110-123)
      sys.add_extern_force({0.3f, 0.6f}, {4.0f, -2.0f});
      printf("force %u\n", sys.m_num_particle);
    }
    sys.animate();
    if (sys.last_status() != SPH_OK) {
      fprintf(stderr, "step %d failed: %d\n", s, sys.last_status());
      return 1;
    }
    double cs = 0.0;
    for (uint i = 0; i < sys.m_num_particle; i++)
      cs += sys.m_host_mem[i].pos.x + 2.0 * sys.m_host_mem[i].pos.y;
    printf("step %d n %u checksum %.9e\n", s, sys.m_num_particle, cs);
  }
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  const int variant = argc > 1 ? atoi(argv[1]) : SPH_V3D2;
  Args a;
  a.steps = argc > 2 ? atoi(argv[2]) : 10;
  a.split = argc > 3 ? atoi(argv[3]) : 1;
  a.merge = argc > 4 ? atoi(argv[4]) : 1;
  a.mc_every = argc > 5 ? atoi(argv[5]) : 0;
  a.pov_dir = argc > 6 && std::strcmp(argv[6], "-") != 0 ? argv[6] : nullptr;
  a.new_step = argc > 7 ? atoi(argv[7]) : -1;
  if (variant == SPH_V2D) return run2d(a);
  if (variant == SPH_V3D4) return run3d<SPH_V3D4>(a);
  return run3d<SPH_V3D2>(a);
}
