// C++ drop-in SPHSystem (include/sph_system_amd.h) over the C ABI of libsph_amd.
// Host seeding loops restate SPH-3D-2/sph_system.cpp:247-300 and SPH-3D-4/sph_system.cpp:
// 248-348 (float-accumulated loops, x outermost); animation() follows
// SPH-3D-2/sph_system.cpp:302-404 with the whole device step in one sph_step call.
#include "../../include/sph_system_amd.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

SPHSystem::SPHSystem(int variant, uint max_particle)
    : m_variant(variant == SPH_V3D4 ? SPH_V3D4 : SPH_V3D2), m_status(SPH_OK), m_ctx(nullptr) {
  m_host_param = new SysParam();
  sph_param_defaults(m_variant, m_host_param);
  m_host_param->max_particle = max_particle;
  SysParam &s = *m_host_param;
  m_num_particle = 0;
  m_max_particle = max_particle;
  m_world_width = s.world_width;
  m_world_height = s.world_height;
  m_world_length = s.world_length;
  m_large_radius = s.large_radius;
  m_small_radius = s.small_radius;
  m_large_kernel = s.large_kernel;
  m_small_kernel = s.small_kernel;
  m_large_mass = s.large_mass;
  m_small_mass = s.small_mass;
  m_cell_size = s.cell_size;
  m_row_cell = s.row_cell;
  m_col_cell = s.col_cell;
  m_len_cell = s.len_cell;
  m_total_cell = s.total_cell;
  m_gravity = s.gravity;
  m_wall_damping = s.wall_damping;
  m_rest_density = s.rest_density;
  m_gas_constant = s.gas_constant;
  m_time_step = s.time_step;
  m_viscosity = s.viscosity;
  m_surface_tension = s.surface_tension;
  m_co_surface = s.co_surface;
  m_co_energy = s.co_energy;
  m_split_criteria = s.split_criteria;
  m_merge_criteria = s.merge_criteria;
  m_dens_mul = s.dens_mul;
  m_den_size = s.den_size;
  m_row_dens = s.row_dens;
  m_col_dens = s.col_dens;
  m_len_dens = s.len_dens;
  m_tot_dens = s.tot_dens;
  m_sphere1 = s.sphere1;
  m_sphere2 = s.sphere2;
  m_sphere3 = s.sphere3;
  m_sphere4 = s.sphere4;

  m_host_mem = (Particle *)std::calloc(max_particle, sizeof(Particle));
  m_host_dens = (float *)std::calloc(m_tot_dens, sizeof(float));
  m_host_dens_pos = (SphFloat3 *)std::calloc(m_tot_dens, sizeof(SphFloat3));
  m_host_dens_normal = (SphFloat3 *)std::calloc(m_tot_dens, sizeof(SphFloat3));
  // triangle capacity as allocated by the reference (3D-2: tot_dens; 3D-4: 3 * tot_dens)
  m_max_triangle = m_tot_dens * (m_variant == SPH_V3D4 ? 3u : 1u);
  m_num_triangle = 0;
  SphFloat3 **tri[6] = {&m_host_triangle0, &m_host_triangle1, &m_host_triangle2,
                        &m_host_triangle_normal0, &m_host_triangle_normal1,
                        &m_host_triangle_normal2};
  for (auto t : tri) *t = (SphFloat3 *)std::calloc(m_max_triangle, sizeof(SphFloat3));
  // voxel positions (SPH-3D-2/sph_system.cpp:214-226)
  for (uint x = 0; x < m_row_dens; x++)
    for (uint y = 0; y < m_col_dens; y++)
      for (uint z = 0; z < m_len_dens; z++) {
        uint v = z * m_row_dens * m_col_dens + y * m_row_dens + x;
        m_host_dens_pos[v].x = (float)x * m_den_size;
        m_host_dens_pos[v].y = (float)y * m_den_size;
        m_host_dens_pos[v].z = (float)z * m_den_size;
      }
  m_sys_running = 0;
  m_trans_once = 0;
  m_use_split = 0;
  m_use_merge = 0;
  m_disp_mode = 1;
  m_status = sph_ctx_create(m_variant, m_host_param, max_particle, m_max_triangle, &m_ctx);
}

SPHSystem::~SPHSystem() {
  if (m_ctx) sph_ctx_destroy(m_ctx);
  std::free(m_host_mem);
  std::free(m_host_dens);
  std::free(m_host_dens_pos);
  std::free(m_host_dens_normal);
  std::free(m_host_triangle0);
  std::free(m_host_triangle1);
  std::free(m_host_triangle2);
  std::free(m_host_triangle_normal0);
  std::free(m_host_triangle_normal1);
  std::free(m_host_triangle_normal2);
  delete m_host_param;
}

void SPHSystem::add_particle(SphFloat3 pos, SphFloat3 vel, uint level) {
  if (m_num_particle >= m_max_particle) return;  // the reference writes past the array
  Particle *p = &m_host_mem[m_num_particle];
  p->pos = pos;
  p->vel = vel;
  p->ev = {0.0f, 0.0f, 0.0f};
  p->acc = {0.0f, 0.0f, 0.0f};
  p->dens = 0.0f;
  p->press = 0.0f;
  p->energy = 0.0f;  // left unset by the reference; computed before its first use
  p->surface = 0.0f;
  p->tension = 0.0f;
  p->level = level;
  m_num_particle++;
}

void SPHSystem::upload() {
  if (m_ctx) m_status = sph_upload_particles(m_ctx, m_host_mem, m_num_particle);
}

void SPHSystem::add_box_particle() {
  SphFloat3 pos;
  const SphFloat3 vel = {0.0f, 0.0f, 0.0f};
  if (m_variant == SPH_V3D2) {
    for (pos.x = m_large_kernel / 2.0f; pos.x <= m_world_width / 2; pos.x += m_large_kernel * 0.8f)
      for (pos.y = m_large_kernel / 2.0f; m_world_height - pos.y >= m_large_kernel / 2.0f;
           pos.y += m_large_kernel * 0.8f)
        for (pos.z = m_large_kernel / 2.0f; m_world_length - pos.z >= m_large_kernel / 2.0f;
             pos.z += m_large_kernel * 0.8f)
          add_particle(pos, vel, 1);
  } else {
    const float space = m_small_kernel;
    for (pos.x = space; m_world_width - pos.x >= space; pos.x += space * 0.5f)
      for (pos.y = space; pos.y < m_world_height / 4; pos.y += space * 0.5f)
        for (pos.z = space; m_world_length - pos.z >= space; pos.z += space * 0.5f)
          add_particle(pos, vel, 2);
  }
  upload();
}

void SPHSystem::add_new_particle() {
  SphFloat3 pos;
  if (m_variant == SPH_V3D2) {
    const SphFloat3 vel = {0.0f, 0.0f, 0.0f};
    for (pos.x = m_large_kernel / 2.0f + 4; m_world_width - 4 - pos.x >= m_large_kernel / 2.0f;
         pos.x += m_small_kernel * 0.4f)
      for (pos.y = m_large_kernel / 2.0f + 4; m_world_height - 4 - pos.y >= m_large_kernel / 2.0f;
           pos.y += m_small_kernel * 0.4f)
        for (pos.z = m_large_kernel / 2.0f + 4; m_world_length - 4 - pos.z >= m_large_kernel / 2.0f;
             pos.z += m_small_kernel * 0.4f)
          add_particle(pos, vel, 2);
  } else {
    // a ball of radius 5 around (24, 18, 12) falling at 10 (SPH-3D-4/sph_system.cpp:291-331);
    // the reference increments the count twice per particle and leaves every other slot
    // uninitialised, which is not reproduced
    const float r = 10.0f;
    const float spacing = (float)(m_small_kernel * 0.5);
    const SphFloat3 vel = {0.0f, -10.0f, 0.0f};
    for (float z = -r; z <= r; z += spacing)
      for (float y = -r; y <= r; y += spacing)
        for (float x = -r; x <= r; x += spacing) {
          float l = std::sqrt(x * x + y * y + z * z);
          if (l <= 5.0f) {
            pos.x = 24.0f + x;
            pos.y = 18.0f + y;
            pos.z = 12.0f + z;
            add_particle(pos, vel, 2);
          }
        }
  }
  upload();
}

void SPHSystem::animation() {
  if (m_sys_running == 0 || !m_ctx) return;
  m_host_param->num_particle = m_num_particle;
  m_status = sph_set_params(m_ctx, m_host_param);
  if (m_status != SPH_OK) return;
  uint32_t flags = 0;
  if (m_use_split == 1) flags |= SPH_STEP_SPLIT;
  if (m_use_merge == 1) flags |= SPH_STEP_MERGE;
  if (m_disp_mode == 2) flags |= SPH_STEP_MC;
  uint32_t n = m_num_particle;
  m_status = sph_step(m_ctx, flags, &n);
  if (m_status == SPH_OK) m_num_particle = n;
  if (m_disp_mode == 2 && m_status == SPH_OK) {
    sph_get_f32(m_ctx, SPH_Q_DGRID, m_host_dens, m_tot_dens);
    sph_get_f32(m_ctx, SPH_Q_NORMAL, (float *)m_host_dens_normal, (size_t)m_tot_dens * 3);
    marching_cube();
  }
  if (m_status == SPH_OK && m_num_particle > 0)
    m_status = sph_download_particles(m_ctx, m_host_mem, m_num_particle);
  m_host_param->flag0 = false;                        // force_waiting(_left)
  if (m_variant == SPH_V3D4) m_host_param->flag1 = false;  // force_waiting_right
}

void SPHSystem::marching_cube() {
  m_num_triangle = 0;
  if (!m_ctx) return;
  size_t nf = 0;
  if (sph_count(m_ctx, SPH_Q_TRI, &nf) != SPH_OK || nf == 0) return;
  std::vector<float> t(nf);
  if (sph_get_f32(m_ctx, SPH_Q_TRI, t.data(), nf) != SPH_OK) return;
  const size_t nt = nf / 18;
  SphFloat3 *dst[6] = {m_host_triangle0, m_host_triangle1, m_host_triangle2,
                       m_host_triangle_normal0, m_host_triangle_normal1, m_host_triangle_normal2};
  for (size_t k = 0; k < nt; k++)
    for (int c = 0; c < 6; c++) {
      const float *s = &t[k * 18 + 3 * c];
      dst[c][k] = {s[0], s[1], s[2]};
    }
  m_num_triangle = (uint)nt;
}

// ---- POV-Ray frame writers (SPH-3D-2/sph_main.cpp:114-220) ----

SphViewFrame sph_view_frame(const SPHSystem &s) {
  SphViewFrame f;
  const float side[3] = {30.0f, 20.0f, 20.0f};
  f.origin.x = 0.0f - side[0] / 2;
  f.origin.y = 0.0f - side[1] / 2;
  f.origin.z = 0.0f - side[2] / 2;
  f.ratio.x = side[0] / s.m_world_width;
  f.ratio.y = side[1] / s.m_world_height;
  f.ratio.z = side[2] / s.m_world_length;
  return f;
}

void sph_pov_frame_name(int n, char out[24]) {
  std::snprintf(out, 24, "Povray_Shots/%06d.pov", n < 0 ? 0 : n % 1000000);
}

namespace {

// the camera and light both writers put first
void pov_header(FILE *fp) {
  std::fprintf(fp, "#include \"colors.inc\"\n");
  std::fprintf(fp, "camera {\n  location <5,10,-35>\n  look_at <0,0,0>\n}\n");
  std::fprintf(fp, "light_source {<13, 13, -20> color White}\n");
}

inline void to_view(const SphViewFrame &f, const SphFloat3 &p, float &x, float &y, float &z) {
  x = p.x * f.ratio.x + f.origin.x;
  y = p.y * f.ratio.y + f.origin.y;
  z = p.z * f.ratio.z + f.origin.z;
}

}  // namespace

int sph_write_pov_particles(const SPHSystem &s, const char *path, const SphViewFrame &f) {
  FILE *fp = std::fopen(path, "w");
  if (!fp) return -1;
  pov_header(fp);
  for (uint k = 0; k < s.m_num_particle; k++) {
    const Particle &q = s.m_host_mem[k];
    float x, y, z;
    to_view(f, q.pos, x, y, z);
    if (q.level == 1)
      std::fprintf(fp, "sphere { <%f, %f, %f>, %f texture { pigment { color Red } } }\n", x, y, z, 0.3f);
    if (q.level == 2)
      std::fprintf(fp, "sphere { <%f, %f, %f>, %f texture { pigment { color Yellow } } }\n", x, y, z,
                   0.24f);
  }
  std::fprintf(fp, "plane { <0,-1,0>, 10 pigment {checker color White, color Black}}\n");
  return std::fclose(fp) == 0 ? 0 : -1;
}

int sph_write_pov_mesh(const SPHSystem &s, const char *path, const SphViewFrame &f) {
  FILE *fp = std::fopen(path, "w");
  if (!fp) return -1;
  pov_header(fp);
  std::fprintf(fp, "mesh {\n");
  for (uint k = 0; k < s.m_num_triangle; k++) {
    const SphFloat3 *v[3] = {&s.m_host_triangle0[k], &s.m_host_triangle1[k], &s.m_host_triangle2[k]};
    const SphFloat3 *nv[3] = {&s.m_host_triangle_normal0[k], &s.m_host_triangle_normal1[k],
                              &s.m_host_triangle_normal2[k]};
    float c[3][3];
    for (int t = 0; t < 3; t++) to_view(f, *v[t], c[t][0], c[t][1], c[t][2]);
    std::fprintf(fp,
                 "smooth_triangle { <%f, %f, %f>, <%f, %f, %f>, <%f, %f, %f>, <%f, %f, %f>, "
                 "<%f, %f, %f>, <%f, %f, %f> }\n",
                 c[0][0], c[0][1], c[0][2], nv[0]->x, nv[0]->y, nv[0]->z, c[1][0], c[1][1], c[1][2],
                 nv[1]->x, nv[1]->y, nv[1]->z, c[2][0], c[2][1], c[2][2], nv[2]->x, nv[2]->y,
                 nv[2]->z);
  }
  std::fprintf(fp,
               "pigment { color red 0.8 green 0.8 blue 0.8 transmit 0.6} finish { ambient 0.2 "
               "diffuse 0.5 } } plane { <0,-1,0>, 10 pigment {checker color White, color Black}}\n");
  return std::fclose(fp) == 0 ? 0 : -1;
}
