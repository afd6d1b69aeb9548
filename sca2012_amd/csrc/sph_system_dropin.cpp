// C++ drop-ins SPHSystem (include/sph_system_amd.h, 3D; include/sph_system_2d_amd.h, 2D)
// over the C ABI of libsph_amd. Host seeding loops restate SPH-3D-2/sph_system.cpp:247-300,
// SPH-3D-4/sph_system.cpp:248-348 and SPH-2D/sph_system.cpp:148-177 (float-accumulated
// loops, x outermost); animation() / animate() follow SPH-3D-2/sph_system.cpp:302-404 and
// SPH-2D/sph_system.cpp:190-264 with the whole device step in one sph_step call.
#define SPH_DROPIN_NO_ALIASES
#include "../../include/sph_system_amd.h"
#include "../../include/sph_system_2d_amd.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

// ---- 3D ----

template <int V>
SphSystem3D<V>::SphSystem3D(uint max_particle) {
  m_variant = V;
  m_status = SPH_OK;
  m_ctx = nullptr;
  m_host_param = new SysParamT();
  sph_param_defaults(V, m_host_param);
  m_host_param->max_particle = max_particle;
  const SysParamT &s = *m_host_param;
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
  m_large_poly6 = s.large_poly6;
  m_large_spiky = s.large_spiky;
  m_large_visco = s.large_visco;
  m_large_grad_poly6 = s.large_grad_poly6;
  m_large_lapc_poly6 = s.large_lapc_poly6;
  m_small_poly6 = s.small_poly6;
  m_small_spiky = s.small_spiky;
  m_small_visco = s.small_visco;
  m_small_grad_poly6 = s.small_grad_poly6;
  m_small_lapc_poly6 = s.small_lapc_poly6;
  m_large_kernel_2 = s.large_kernel_2;
  m_large_kernel_6 = s.large_kernel_6;
  m_small_kernel_2 = s.small_kernel_2;
  m_small_kernel_6 = s.small_kernel_6;
  m_large_poly6_radius = s.large_poly6_radius;
  m_small_poly6_radius = s.small_poly6_radius;
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

  m_dev_mem = nullptr;
  m_dev_hash = m_dev_index = m_cell_start = m_cell_end = nullptr;
  m_dev_status = m_dev_split = m_dev_split_index = m_dev_merge = nullptr;
  m_trans_vel = nullptr;
  m_dev_dens = nullptr;
  m_dev_dens_pos = m_dev_dens_normal = nullptr;

  m_host_mem = (SphParticle3D *)std::calloc(max_particle, sizeof(SphParticle3D));
  m_host_dens = (float *)std::calloc(m_tot_dens, sizeof(float));
  m_host_dens_pos = (SphFloat3 *)std::calloc(m_tot_dens, sizeof(SphFloat3));
  m_host_dens_normal = (SphFloat3 *)std::calloc(m_tot_dens, sizeof(SphFloat3));
  m_num_lines = 0;
  m_host_line0 = (SphFloat3 *)std::calloc(1, sizeof(SphFloat3));
  m_host_line1 = (SphFloat3 *)std::calloc(1, sizeof(SphFloat3));
  // triangle capacity as allocated by the reference (3D-2: tot_dens; 3D-4: 3 * tot_dens)
  m_max_triangle = m_tot_dens * (V == SPH_V3D4 ? 3u : 1u);
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
  m_status = sph_ctx_create(V, m_host_param, max_particle, m_max_triangle, &m_ctx);
}

template <int V>
SphSystem3D<V>::~SphSystem3D() {
  if (m_ctx) sph_ctx_destroy(m_ctx);
  std::free(m_host_mem);
  std::free(m_host_dens);
  std::free(m_host_dens_pos);
  std::free(m_host_dens_normal);
  std::free(m_host_line0);
  std::free(m_host_line1);
  std::free(m_host_triangle0);
  std::free(m_host_triangle1);
  std::free(m_host_triangle2);
  std::free(m_host_triangle_normal0);
  std::free(m_host_triangle_normal1);
  std::free(m_host_triangle_normal2);
  delete m_host_param;
}

template <int V>
void SphSystem3D<V>::add_particle(SphFloat3 pos, SphFloat3 vel, uint level) {
  if (m_num_particle >= m_max_particle) return;  // the reference writes past the array
  SphParticle3D *p = &m_host_mem[m_num_particle];
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

template <int V>
void SphSystem3D<V>::upload() {
  if (m_ctx) m_status = sph_upload_particles(m_ctx, m_host_mem, m_num_particle);
}

template <int V>
void SphSystem3D<V>::add_box_particle() {
  SphFloat3 pos;
  const SphFloat3 vel = {0.0f, 0.0f, 0.0f};
  if (V == SPH_V3D2) {
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

template <int V>
void SphSystem3D<V>::add_new_particle() {
  SphFloat3 pos;
  if (V == SPH_V3D2) {
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

template <int V>
void SphSystem3D<V>::animation() {
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
  // the keyboard pushes last one step (SPH-3D-2/sph_system.cpp:403, SPH-3D-4 :452)
  if constexpr (V == SPH_V3D4) {
    m_host_param->force_waiting_left = false;
    m_host_param->force_waiting_right = false;
  } else {
    m_host_param->force_waiting = false;
  }
}

template <int V>
void SphSystem3D<V>::marching_cube() {
  m_num_triangle = 0;
  m_num_lines = 0;
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

template class SphSystem3D<SPH_V3D2>;
template class SphSystem3D<SPH_V3D4>;

// ---- 2D ----

SphSystem2D::SphSystem2D(uint max_particle) : m_status(SPH_OK), m_ctx(nullptr) {
  m_host_param = new SphSysParam2D();
  sph_param_defaults(SPH_V2D, m_host_param);
  m_host_param->max_particle = max_particle;
  const SphSysParam2D &s = *m_host_param;
  m_num_particle = 0;
  m_max_particle = max_particle;
  m_world_width = s.world_width;
  m_world_depth = s.world_depth;
  m_large_mass = s.large_mass;
  m_small_mass = s.small_mass;
  m_large_kernel = s.large_kernel;
  m_small_kernel = s.small_kernel;
  m_large_kernel2 = s.large_kernel2;
  m_small_kernel2 = s.small_kernel2;
  m_large_kernel6 = s.large_kernel6;
  m_small_kernel6 = s.small_kernel6;
  m_large_radius = s.large_radius;
  m_small_radius = s.small_radius;
  m_large_poly6 = s.large_poly6;
  m_small_poly6 = s.small_poly6;
  m_large_spiky = s.large_spiky;
  m_small_spiky = s.small_spiky;
  m_large_visco = s.large_visco;
  m_small_visco = s.small_visco;
  m_large_grad_poly6 = s.large_grad_poly6;
  m_small_grad_poly6 = s.small_grad_poly6;
  m_large_lapc_poly6 = s.large_lapc_poly6;
  m_small_lapc_poly6 = s.small_lapc_poly6;
  m_gravity = s.gravity;
  m_wall_damping = s.wall_damping;
  m_rest_density = s.rest_density;
  m_gas_constant = s.gas_constant;
  m_viscosity = s.viscosity;
  m_time_step = s.time_step;
  m_surface_normal = s.surface_normal;
  m_sys_running = 0;
  m_cell_size = s.cell_size;
  m_row_cell = s.row_cell;
  m_col_cell = s.col_cell;
  m_total_cell = s.total_cell;
  m_co_surface = s.co_surface;
  m_co_energy = s.co_energy;
  m_thresh_split = s.thresh_split;
  m_thresh_merge = s.thresh_merge;
  m_host_mem = (SphParticle2D *)std::calloc(max_particle, sizeof(SphParticle2D));
  m_dev_mem = nullptr;
  m_dev_hash = m_dev_index = m_cell_start = m_cell_end = nullptr;
  m_trans_vel = nullptr;
  m_dev_status = m_dev_split = m_dev_merge = m_dev_split_index = nullptr;
  m_enable_split = 0;
  m_enable_merge = 0;
  m_status = sph_ctx_create(SPH_V2D, m_host_param, max_particle, 0, &m_ctx);
}

SphSystem2D::~SphSystem2D() {
  if (m_ctx) sph_ctx_destroy(m_ctx);
  std::free(m_host_mem);
  delete m_host_param;
}

void SphSystem2D::add_particle(SphFloat2 pos, SphFloat2 vel, uint level) {
  if (m_num_particle >= m_max_particle) return;  // the reference writes past the array
  SphParticle2D *p = &m_host_mem[m_num_particle];
  p->pos = pos;
  p->vel = vel;
  p->ev = {0.0f, 0.0f};
  p->acc = {0.0f, 0.0f};
  p->dens = 0.0f;
  p->pres = 0.0f;
  p->energy = 0.0f;          // left unset by the reference; computed before its first use
  p->surface_normal = 0.0f;  // likewise
  p->level = level;
  m_num_particle++;
}

void SphSystem2D::add_init_particle() {
  SphFloat2 pos;
  const SphFloat2 vel = {0.0f, 0.0f};
  for (pos.x = m_large_radius / 2.0f; pos.x < m_world_width / 2; pos.x += m_large_kernel * 0.5f)
    for (pos.y = m_large_radius / 2.0f; m_world_depth - pos.y >= m_large_radius / 2.0f;
         pos.y += m_large_kernel * 0.5f)
      add_particle(pos, vel, 1);
  if (m_ctx) m_status = sph_upload_particles(m_ctx, m_host_mem, m_num_particle);
}

void SphSystem2D::add_extern_force(SphFloat2 pos, SphFloat2 force) {
  const uint cell_x = (uint)std::floor(pos.x * m_row_cell);
  const uint cell_y = (uint)std::floor(pos.y * m_col_cell);
  m_host_param->extern_click = {pos.x * m_world_width, pos.y * m_world_depth};
  m_host_param->extern_pos = {cell_x, cell_y};
  m_host_param->extern_force = {force.x, force.y};
  m_host_param->force_waiting = true;
}

void SphSystem2D::animate() {
  if (m_sys_running == 0 || !m_ctx) return;
  m_host_param->num_particle = m_num_particle;
  m_status = sph_set_params(m_ctx, m_host_param);
  if (m_status != SPH_OK) return;
  uint32_t flags = 0;
  if (m_enable_split == 1) flags |= SPH_STEP_SPLIT;
  if (m_enable_merge == 1) flags |= SPH_STEP_MERGE;
  uint32_t n = m_num_particle;
  m_status = sph_step(m_ctx, flags, &n);
  if (m_status == SPH_OK) m_num_particle = n;
  if (m_status == SPH_OK && m_num_particle > 0)
    m_status = sph_download_particles(m_ctx, m_host_mem, m_num_particle);
  m_host_param->force_waiting = false;  // SPH-2D/sph_system.cpp:263
}

// ---- POV-Ray frame writers (SPH-3D-2/sph_main.cpp:114-220) ----

SphViewFrame sph_view_frame(const SphSystem3DBase &s) {
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

int sph_write_pov_particles(const SphSystem3DBase &s, const char *path, const SphViewFrame &f) {
  FILE *fp = std::fopen(path, "w");
  if (!fp) return -1;
  pov_header(fp);
  for (uint k = 0; k < s.m_num_particle; k++) {
    const SphParticle3D &q = s.m_host_mem[k];
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

int sph_write_pov_mesh(const SphSystem3DBase &s, const char *path, const SphViewFrame &f) {
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
