// Drives the C++ drop-in SPHSystem the way the reference's GLUT application does
// (SPH-3D-2/sph_main.cpp:278-323 init, display_func -> animation() per frame), without a
// window: prints the live count per step, and the triangle count of display-mode-2 steps.
// With an output directory it also writes the reference's POV-Ray frames there:
// make_particle for every step (frame = step) and make_mesh for every marching-cubes step
// (frame = step, "mesh_" prefix).
// Usage: sph_dropin_run <variant 1|2> <steps> <split 0|1> <merge 0|1> <mc_every> [pov_dir]
#include <cstdio>
#include <cstdlib>
#include <string>

#include "../../include/sph_system_amd.h"

int main(int argc, char **argv) {
  int variant = argc > 1 ? atoi(argv[1]) : SPH_V3D2;
  int steps = argc > 2 ? atoi(argv[2]) : 10;
  int split = argc > 3 ? atoi(argv[3]) : 1;
  int merge = argc > 4 ? atoi(argv[4]) : 1;
  int mc_every = argc > 5 ? atoi(argv[5]) : 0;
  const char *pov_dir = argc > 6 ? argv[6] : nullptr;
  SPHSystem sys(variant, 1000000);
  if (sys.last_status() != SPH_OK) {
    fprintf(stderr, "context creation failed: %d\n", sys.last_status());
    return 1;
  }
  sys.add_box_particle();
  printf("init %u\n", sys.m_num_particle);
  sys.m_sys_running = 1;
  sys.m_use_split = split;
  sys.m_use_merge = merge;
  for (int s = 0; s < steps; s++) {
    sys.m_disp_mode = (mc_every > 0 && (s + 1) % mc_every == 0) ? 2 : 1;
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
    if (pov_dir) {
      const SphViewFrame f = sph_view_frame(sys);
      char name[24];
      sph_pov_frame_name(s, name);
      const std::string base = std::string(pov_dir) + "/";
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
