// CPU oracle for the adaptive SPH step -- TEST INFRASTRUCTURE ONLY (see sph_oracle.h).
//
// Restates, phase by phase and in the same floating-point operation order, the kernels of
// SPH-2D/sph_system.cu, SPH-3D-2/sph_system.cu and SPH-3D-4/sph_system.cu and the host
// orchestration of */sph_system.cpp. Built with -O3 -fopenmp -ffp-contract=off (no FMA
// contraction, no fast-math), so the fp32 instantiation is the bit-exact parity checker
// and the fp32 CPU baseline, and the fp64 instantiation is the drift oracle.
//
// Conventions fixed here where the reference leaves the arithmetic to a library:
//   * cutil_math float3/float operators are componentwise IEEE ops, e.g. (v * s) / t is
//     ((v.x * s) / t, ...); normalize(v) = v / sqrt(dot(v, v)).
//   * pow(x, 3) = (x * x) * x and pow(x, 2) = x * x.
//   * pow(EXP, y) = exp(y * ln(EXP)) with the Cody-Waite/Taylor expf below.
//   * pow(kernel, 0.5f) = sqrtf(kernel); pow(PI, 0.25f) = float(double(PI)^0.25).
//   * thrust::sort_by_key is taken as a stable sort (ties in ascending slot).
#include "sph_oracle.h"
#include "dual.h"
#include "mc_tri_table.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <new>
#include <type_traits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#include <parallel/algorithm>
#endif

typedef uint32_t u32;

// ---------------------------------------------------------------------------------------
// Reference memory layouts (SPH-3D-2/sph_system.h:8-24, sph_param.h:8-87, sph_object.h:4-9;
// SPH-3D-4/sph_param.h:8-88; SPH-2D/sph_system.h:7-22, sph_param.h:6-72).
// ---------------------------------------------------------------------------------------
namespace {

struct F3 { float x, y, z; };
struct alignas(8) F2 { float x, y; };
struct alignas(8) U2 { u32 x, y; };
struct SphereR { F3 pos; float radius; float damping; };

struct SysParam3R {
  u32 num_particle, max_particle;
  float world_width, world_height, world_length;
  float large_radius, small_radius;
  float large_kernel, small_kernel;
  float large_mass, small_mass;
  float cell_size;
  u32 row_cell, col_cell, len_cell, total_cell;
  F3 gravity;
  float wall_damping, rest_density, gas_constant, time_step, viscosity, surface_tension;
  float co_surface, co_energy;
  float split_criteria, merge_criteria;
  float large_poly6, large_spiky, large_visco;
  float large_grad_poly6, large_lapc_poly6;
  float small_poly6, small_spiky, small_visco;
  float small_grad_poly6, small_lapc_poly6;
  float large_kernel_2, large_kernel_6;
  float small_kernel_2, small_kernel_6;
  float large_poly6_radius, small_poly6_radius;
  u32 dens_mul;
  float den_size;
  u32 row_dens, col_dens, len_dens, tot_dens;
  SphereR sphere1, sphere2, sphere3, sphere4;
  float split_energy, merge_energy;
  // 3D-2: force_waiting, use_cylinder.  3D-4: force_waiting_left, force_waiting_right,
  // use_cylinder.
  bool b0, b1, b2;
};
static_assert(sizeof(SysParam3R) == 296, "SysParam 3D layout");

struct Particle3R {
  F3 pos, vel, ev, acc;
  float dens, press, energy, surface, tension;
  u32 level;
};
static_assert(sizeof(Particle3R) == 72, "Particle 3D layout");

struct SysParam2R {
  u32 num_particle, max_particle;
  float world_width, world_depth;
  float large_mass, small_mass;
  float large_kernel, small_kernel;
  float large_kernel2, small_kernel2;
  float large_kernel6, small_kernel6;
  float large_radius, small_radius;
  float large_poly6, small_poly6;
  float large_spiky, small_spiky;
  float large_visco, small_visco;
  float large_grad_poly6, small_grad_poly6;
  float large_lapc_poly6, small_lapc_poly6;
  F2 gravity;
  float wall_damping, rest_density, gas_constant, viscosity, time_step, surface_normal;
  float cell_size;
  u32 row_cell, col_cell, total_cell;
  float co_surface, co_energy;
  float thresh_split, thresh_merge;
  bool force_waiting;
  F2 extern_force, extern_click;
  U2 extern_pos;
  float extern_len, extern_coe;
};
static_assert(sizeof(SysParam2R) == 200, "SysParam 2D layout");

struct Particle2R {
  F2 pos, vel, ev, acc;
  float dens, pres, energy;
  u32 level;
  float surface_normal;
};
static_assert(sizeof(Particle2R) == 56, "Particle 2D layout");

// ---------------------------------------------------------------------------------------
// fp32 exp: Cody & Waite (1980) argument reduction t = k ln2 + r with ln2 split into a
// 16-significant-bit head and a tail, then the degree-7 Taylor polynomial of e^r on
// |r| <= ln2/2 by Horner's rule, then scaling by 2^k through the exponent field. Every
// step is one IEEE binary32 operation in a fixed order (no FMA), so any IEEE machine
// without contraction returns the same bits. Max error ~2 ulp (tested against libm).
// ---------------------------------------------------------------------------------------
inline float scale2(float p, int k) {
  // p * 2^k for k in [-252, 254]
  if (k > 127) {
    p *= 0x1p127f;
    k -= 127;
  } else if (k < -126) {
    p *= 0x1p-126f;
    k += 126;
    if (k < -126) {
      p *= 0x1p-126f;
      k += 126;
    }
  }
  u32 bits = (u32)(k + 127) << 23;
  float s;
  std::memcpy(&s, &bits, 4);
  return p * s;
}

inline float expf_cw(float t) {
  if (t != t) return t;
  if (t > 88.75f) return std::numeric_limits<float>::infinity();
  if (t < -104.0f) return 0.0f;
  const float inv_ln2 = 0x1.715476p+0f;
  const float ln2_hi = 0x1.62e4p-1f;
  const float ln2_lo = 0x1.7f7d1cp-20f;
  float kf = std::nearbyint(t * inv_ln2);
  float r = (t - kf * ln2_hi) - kf * ln2_lo;
  float p = 0x1.a01a02p-13f;          // 1/7!
  p = p * r + 0x1.6c16c2p-10f;        // 1/6!
  p = p * r + 0x1.111112p-7f;         // 1/5!
  p = p * r + 0x1.555556p-5f;         // 1/4!
  p = p * r + 0x1.555556p-3f;         // 1/3!
  p = p * r + 0x1p-1f;                // 1/2!
  p = p * r + 1.0f;
  p = p * r + 1.0f;
  return scale2(p, (int)kf);
}

// Mexican-hat "wavelet" of SPH-3D-2/sph_system.cu:11-16 (same in 2D :827-832, 3D-4 :11-16):
//   2/(pow(PI,.25) pow(3 sigma,.5)) * (x^2/sigma^2 - 1) * pow(EXP, -x^2/(2 sigma^2)), sigma=2
// with PI = 3.141592f (sph_header.h:22) and EXP = 2.718281f (sph_system.cu:7).
inline float wavelet_f(float x) {
  const float A = 0x1.3a0156p-1f;   // 2.0f / (float(PI^0.25) * sqrtf(6.0f))
  const float ln_b = 0x1.fffff6p-1f;  // float(ln(2.718281f))
  float xx = x * x;
  float b = xx / 4.0f - 1.0f;
  float y = (-xx) / 8.0f;
  return (A * b) * expf_cw(y * ln_b);
}

inline double wavelet_d(double x) {
  const double PIf = (double)3.141592f;
  const double A = 2.0 / (std::pow(PIf, 0.25) * std::sqrt(6.0));
  double xx = x * x;
  return (A * (xx / 4.0 - 1.0)) * std::pow((double)2.718281f, (-xx) / 8.0);
}

template <class R> inline R wavelet(R x);
template <> inline float wavelet<float>(float x) { return wavelet_f(x); }
template <> inline double wavelet<double>(double x) { return wavelet_d(x); }
// the wavelet of the value type and its derivative A E (x/2 - (x^2/4 - 1) x ln(b) / 4),
// E = b^(-x^2/8), in float64
inline double wavelet_dx(double xv) {
  const double PIf = (double)3.141592f;
  const double A = 2.0 / (std::pow(PIf, 0.25) * std::sqrt(6.0));
  const double lb = std::log((double)2.718281f);
  const double E = std::pow((double)2.718281f, (-(xv * xv)) / 8.0);
  return A * E * (xv / 2.0 - (xv * xv / 4.0 - 1.0) * xv * lb / 4.0);
}
template <> inline Dual wavelet<Dual>(Dual x) {
  return Dual(wavelet_d(x.v), wavelet_dx(x.v) * x.d, 0);
}
template <> inline DualF wavelet<DualF>(DualF x) {
  return DualF(wavelet_f(x.v), wavelet_dx((double)x.v) * x.d, 0);
}

// sqrt and floor of the step's scalar type (float, double or Dual)
template <class R> inline R msqrt(R x) {
  using std::sqrt;
  return sqrt(x);
}
template <class R> inline R mfloor(R x) {
  using std::floor;
  return floor(x);
}
// the float64 value of the step's scalar type
inline double to_d(float x) { return x; }
inline double to_d(double x) { return x; }
template <class V> inline double to_d(const DualT<V> &x) { return (double)x.v; }

// ---------------------------------------------------------------------------------------
// Internal state
// ---------------------------------------------------------------------------------------
template <class R> struct P {
  R pos[3], vel[3], ev[3], acc[3];
  R dens, press, energy, surface, tension;
  u32 level;
};

template <class R> struct Sph {
  R pos[3];
  R radius, damping;
};

template <class R> struct Prm {
  int variant, dim, nsplit;
  R world[3];
  R radius[2], kernel[2], mass[2];  // [0] = large (level 1), [1] = small (level 2)
  R cell_size;
  u32 ncell[3], total_cell;
  R gravity[3];
  R wall_damping, rest_density, gas_constant, time_step, viscosity, surface_tension;
  R co_surface, co_energy, split_criteria, merge_criteria;
  R poly6[2], spiky[2], visco[2], grad[2], lapc[2], k2[2], k6[2], poly6_radius[2];
  R sqrt_kernel[2];
  u32 nd[3], tot_dens;
  R den_size;
  Sph<R> sphere[3];
  bool push_left, push_right, use_cylinder;
  // 2D mouse force (SPH-2D/sph_system.cu:781-789)
  bool mouse;
  R extern_force[2], extern_click[2], extern_len, extern_coe;
  R surf_coef;   // 0.04 (2D), 0.1 (3D-2), 0.02 (3D-4)
  R dens_eps;    // 1e-12f (2D, EXT2) or 1e-4f (3D)
};

inline int lv(u32 level) { return level == 1 ? 0 : 1; }

// Neighbour offsets in the reference's order (SPH-3D-2/sph_system.cu:456-562; the 2D
// order at SPH-2D/sph_system.cu:383-408 is the first nine).
const int OFF[27][3] = {
    {0, 0, 0},   {-1, -1, 0},  {-1, 0, 0},  {0, -1, 0},  {1, 1, 0},   {1, 0, 0},   {0, 1, 0},
    {1, -1, 0},  {-1, 1, 0},   {0, 0, -1},  {-1, -1, -1}, {-1, 0, -1}, {0, -1, -1}, {1, 1, -1},
    {1, 0, -1},  {0, 1, -1},   {1, -1, -1}, {-1, 1, -1}, {0, 0, 1},   {-1, -1, 1}, {-1, 0, 1},
    {0, -1, 1},  {1, 1, 1},    {1, 0, 1},   {0, 1, 1},   {1, -1, 1},  {-1, 1, 1}};

const int EDGE[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                         {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
const int CORNER[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                          {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

}  // namespace

struct OrcSim {
  virtual ~OrcSim() {}
  int variant = 0;
  int fp64 = 0;
};

namespace {

template <class R> struct Sim : OrcSim {
  Prm<R> p;
  SysParam3R raw3;
  SysParam2R raw2;
  u32 cap = 0, n = 0;
  u32 n_flagged = 0, n_status = 0;  // lengths of the split-flag and status arrays
  std::vector<P<R>> mem;
  std::vector<u32> hash, index, cstart, cend, split, sindex, status;
  std::vector<R> dgrid, normal;
  std::vector<float> tris;
  bool lists_valid = false;
  // Learnable force field of the differentiable rollout (no counterpart in the reference):
  // nf^3 nodes x 3 components spanning the box (node a at a * world / (nf - 1)), sampled
  // trilinearly at the particle position and added to its force in integrate. nf = 0: off.
  int nf = 0;
  std::vector<R> field;
  // Topology of a rollout (no counterpart in the reference). With rec_topo each step()
  // keeps the slots of its split parents (numbering before the split) and its merge groups
  // (numbering after it, first member the survivor); with replay_topo a step() takes them
  // from that record instead of deciding them (energy and surface are still computed), so
  // a perturbed run keeps the topology of the recorded one.
  bool rec_topo = false, replay_topo = false;
  u32 topo_step = 0;
  std::vector<std::vector<u32>> topo_split, topo_merge;
  // With the topology, each step's branch decisions per slot (3D): bit 2a = wall clamp at
  // the high end of axis a, bit 2a+1 at the low end (integrate), bit 6 = colour-field force
  // on (|grad c| above the surface tension, force). Replaying them freezes every branch of
  // the step, so finite differences see the smooth model whose derivative the adjoint takes.
  std::vector<std::vector<uint8_t>> topo_branch;
  // this step's decision bytes (recording or replaying), or null; `fresh` (the force
  // phase, the step's first decision) clears them when recording
  uint8_t *branch_rec(bool fresh) {
    if (p.dim != 3) return nullptr;
    if (rec_topo) {
      if (topo_branch.size() <= topo_step) topo_branch.resize(topo_step + 1);
      if (fresh || topo_branch[topo_step].size() < n) topo_branch[topo_step].assign(n, 0);
      return topo_branch[topo_step].data();
    }
    if (replay_topo && topo_step < topo_branch.size() && topo_branch[topo_step].size() >= n)
      return topo_branch[topo_step].data();
    return nullptr;
  }

  // ---- parameters -------------------------------------------------------------------
  void load_params3(const SysParam3R &s) {
    raw3 = s;
    p.dim = 3;
    p.nsplit = variant == ORC_V3D4 ? 4 : 2;
    p.world[0] = s.world_width; p.world[1] = s.world_height; p.world[2] = s.world_length;
    p.radius[0] = s.large_radius; p.radius[1] = s.small_radius;
    p.kernel[0] = s.large_kernel; p.kernel[1] = s.small_kernel;
    p.mass[0] = s.large_mass; p.mass[1] = s.small_mass;
    p.cell_size = s.cell_size;
    p.ncell[0] = s.row_cell; p.ncell[1] = s.col_cell; p.ncell[2] = s.len_cell;
    p.total_cell = s.total_cell;
    p.gravity[0] = s.gravity.x; p.gravity[1] = s.gravity.y; p.gravity[2] = s.gravity.z;
    p.wall_damping = s.wall_damping; p.rest_density = s.rest_density;
    p.gas_constant = s.gas_constant; p.time_step = s.time_step; p.viscosity = s.viscosity;
    p.surface_tension = s.surface_tension;
    p.co_surface = s.co_surface; p.co_energy = s.co_energy;
    p.split_criteria = s.split_criteria; p.merge_criteria = s.merge_criteria;
    p.poly6[0] = s.large_poly6; p.poly6[1] = s.small_poly6;
    p.spiky[0] = s.large_spiky; p.spiky[1] = s.small_spiky;
    p.visco[0] = s.large_visco; p.visco[1] = s.small_visco;
    p.grad[0] = s.large_grad_poly6; p.grad[1] = s.small_grad_poly6;
    p.lapc[0] = s.large_lapc_poly6; p.lapc[1] = s.small_lapc_poly6;
    p.k2[0] = s.large_kernel_2; p.k2[1] = s.small_kernel_2;
    p.k6[0] = s.large_kernel_6; p.k6[1] = s.small_kernel_6;
    p.poly6_radius[0] = s.large_poly6_radius; p.poly6_radius[1] = s.small_poly6_radius;
    p.nd[0] = s.row_dens; p.nd[1] = s.col_dens; p.nd[2] = s.len_dens; p.tot_dens = s.tot_dens;
    p.den_size = s.den_size;
    const SphereR *sp[3] = {&s.sphere1, &s.sphere2, &s.sphere3};
    for (int k = 0; k < 3; k++) {
      p.sphere[k].pos[0] = sp[k]->pos.x; p.sphere[k].pos[1] = sp[k]->pos.y;
      p.sphere[k].pos[2] = sp[k]->pos.z;
      p.sphere[k].radius = sp[k]->radius; p.sphere[k].damping = sp[k]->damping;
    }
    if (variant == ORC_V3D4) {
      p.push_left = s.b0; p.push_right = s.b1; p.use_cylinder = s.b2;
    } else {
      p.push_left = s.b0; p.push_right = false; p.use_cylinder = s.b1;
    }
    p.mouse = false;
    p.surf_coef = variant == ORC_V3D4 ? R(0.02f) : R(0.1f);
    p.dens_eps = R(0.0001f);
    set_sqrt_kernel();
  }

  void load_params2(const SysParam2R &s) {
    raw2 = s;
    p.dim = 2;
    p.nsplit = 2;
    p.world[0] = s.world_width; p.world[1] = s.world_depth; p.world[2] = 0;
    p.radius[0] = s.large_radius; p.radius[1] = s.small_radius;
    p.kernel[0] = s.large_kernel; p.kernel[1] = s.small_kernel;
    p.mass[0] = s.large_mass; p.mass[1] = s.small_mass;
    p.cell_size = s.cell_size;
    p.ncell[0] = s.row_cell; p.ncell[1] = s.col_cell; p.ncell[2] = 1;
    p.total_cell = s.total_cell;
    p.gravity[0] = s.gravity.x; p.gravity[1] = s.gravity.y; p.gravity[2] = 0;
    p.wall_damping = s.wall_damping; p.rest_density = s.rest_density;
    p.gas_constant = s.gas_constant; p.time_step = s.time_step; p.viscosity = s.viscosity;
    p.surface_tension = s.surface_normal;
    p.co_surface = s.co_surface; p.co_energy = s.co_energy;
    p.split_criteria = s.thresh_split; p.merge_criteria = s.thresh_merge;
    p.poly6[0] = s.large_poly6; p.poly6[1] = s.small_poly6;
    p.spiky[0] = s.large_spiky; p.spiky[1] = s.small_spiky;
    p.visco[0] = s.large_visco; p.visco[1] = s.small_visco;
    p.grad[0] = s.large_grad_poly6; p.grad[1] = s.small_grad_poly6;
    p.lapc[0] = s.large_lapc_poly6; p.lapc[1] = s.small_lapc_poly6;
    p.k2[0] = s.large_kernel2; p.k2[1] = s.small_kernel2;
    p.k6[0] = s.large_kernel6; p.k6[1] = s.small_kernel6;
    p.poly6_radius[0] = 0; p.poly6_radius[1] = 0;
    p.nd[0] = p.nd[1] = p.nd[2] = 0; p.tot_dens = 0; p.den_size = 0;
    p.push_left = p.push_right = p.use_cylinder = false;
    p.mouse = s.force_waiting;
    p.extern_force[0] = s.extern_force.x; p.extern_force[1] = s.extern_force.y;
    p.extern_click[0] = s.extern_click.x; p.extern_click[1] = s.extern_click.y;
    p.extern_len = s.extern_len; p.extern_coe = s.extern_coe;
    p.surf_coef = R(0.04f);
    p.dens_eps = R(1e-12f);
    set_sqrt_kernel();
  }

  void set_sqrt_kernel() {
    for (int k = 0; k < 2; k++) p.sqrt_kernel[k] = msqrt(p.kernel[k]);
  }

  int set_params(const void *sp) {
    if (variant == ORC_V2D) load_params2(*(const SysParam2R *)sp);
    else load_params3(*(const SysParam3R *)sp);
    if (p.total_cell == 0) return ORC_EINVAL;
    cstart.assign(p.total_cell, 0xffffffffu);
    cend.assign(p.total_cell, 0u);
    if (p.dim == 3 && p.tot_dens) {
      dgrid.assign(p.tot_dens, R(0));
      normal.assign((size_t)p.tot_dens * 3, R(0));
    }
    lists_valid = false;
    return ORC_OK;
  }

  // ---- particles ----------------------------------------------------------------------
  int set_particles(const void *aos, u32 cnt) {
    if (cnt > cap) return ORC_ECAP;
    n = cnt;
    if (variant == ORC_V2D) {
      const Particle2R *a = (const Particle2R *)aos;
      for (u32 i = 0; i < cnt; i++) {
        P<R> &q = mem[i];
        q.pos[0] = a[i].pos.x; q.pos[1] = a[i].pos.y; q.pos[2] = 0;
        q.vel[0] = a[i].vel.x; q.vel[1] = a[i].vel.y; q.vel[2] = 0;
        q.ev[0] = a[i].ev.x; q.ev[1] = a[i].ev.y; q.ev[2] = 0;
        q.acc[0] = a[i].acc.x; q.acc[1] = a[i].acc.y; q.acc[2] = 0;
        q.dens = a[i].dens; q.press = a[i].pres; q.energy = a[i].energy;
        q.surface = a[i].surface_normal; q.tension = 0; q.level = a[i].level;
      }
    } else {
      const Particle3R *a = (const Particle3R *)aos;
      for (u32 i = 0; i < cnt; i++) {
        P<R> &q = mem[i];
        q.pos[0] = a[i].pos.x; q.pos[1] = a[i].pos.y; q.pos[2] = a[i].pos.z;
        q.vel[0] = a[i].vel.x; q.vel[1] = a[i].vel.y; q.vel[2] = a[i].vel.z;
        q.ev[0] = a[i].ev.x; q.ev[1] = a[i].ev.y; q.ev[2] = a[i].ev.z;
        q.acc[0] = a[i].acc.x; q.acc[1] = a[i].acc.y; q.acc[2] = a[i].acc.z;
        q.dens = a[i].dens; q.press = a[i].press; q.energy = a[i].energy;
        q.surface = a[i].surface; q.tension = a[i].tension; q.level = a[i].level;
      }
    }
    lists_valid = false;
    return ORC_OK;
  }

  int get_particles(void *aos) const {
    if (variant == ORC_V2D) {
      Particle2R *a = (Particle2R *)aos;
      for (u32 i = 0; i < n; i++) {
        const P<R> &q = mem[i];
        a[i].pos.x = (float)q.pos[0]; a[i].pos.y = (float)q.pos[1];
        a[i].vel.x = (float)q.vel[0]; a[i].vel.y = (float)q.vel[1];
        a[i].ev.x = (float)q.ev[0]; a[i].ev.y = (float)q.ev[1];
        a[i].acc.x = (float)q.acc[0]; a[i].acc.y = (float)q.acc[1];
        a[i].dens = (float)q.dens; a[i].pres = (float)q.press; a[i].energy = (float)q.energy;
        a[i].level = q.level; a[i].surface_normal = (float)q.surface;
      }
    } else {
      Particle3R *a = (Particle3R *)aos;
      for (u32 i = 0; i < n; i++) {
        const P<R> &q = mem[i];
        a[i].pos = {(float)q.pos[0], (float)q.pos[1], (float)q.pos[2]};
        a[i].vel = {(float)q.vel[0], (float)q.vel[1], (float)q.vel[2]};
        a[i].ev = {(float)q.ev[0], (float)q.ev[1], (float)q.ev[2]};
        a[i].acc = {(float)q.acc[0], (float)q.acc[1], (float)q.acc[2]};
        a[i].dens = (float)q.dens; a[i].press = (float)q.press;
        a[i].energy = (float)q.energy; a[i].surface = (float)q.surface;
        a[i].tension = (float)q.tension; a[i].level = q.level;
      }
    }
    return ORC_OK;
  }

  // ---- grid ---------------------------------------------------------------------------
  // calc_grid_pos: (uint)floor(p / cell_size) (SPH-3D-2/sph_system.cu:18-27; 2D :11-19).
  // Returns false for a coordinate outside [0, ncell) or NaN (undefined in the reference).
  bool cell_of(const R *pos, u32 c[3]) const {
    for (int a = 0; a < p.dim; a++) {
      R q = mfloor(pos[a] / p.cell_size);
      if (!(q >= R(0)) || !(q < (R)p.ncell[a])) return false;
      c[a] = (u32)q;
    }
    if (p.dim == 2) c[2] = 0;
    return true;
  }
  // calc_grid_hash (SPH-3D-2/sph_system.cu:29-32; 2D :21-26)
  u32 hash_of(const u32 c[3]) const {
    return c[2] * p.ncell[0] * p.ncell[1] + c[1] * p.ncell[0] + c[0];
  }
  // Neighbour cell o of centre c, or false when the reference's neighbor_status disables
  // it (SPH-3D-2/sph_system.cu:564-652; 2D :410-432).
  bool neighbor(int o, const u32 c[3], u32 &h) const {
    if (p.dim == 2 && o >= 9) return false;
    u32 nc[3];
    for (int a = 0; a < 3; a++) {
      int d = OFF[o][a];
      if (d == -1 && c[a] == 0) return false;
      if (d == 1 && c[a] == p.ncell[a] - 1) return false;
      nc[a] = c[a] + (u32)d;
    }
    h = hash_of(nc);
    return true;
  }
  int nnb() const { return p.dim == 2 ? 9 : 27; }

  // ---- A2 calc_hash (SPH-3D-2/sph_system.cu:78-117; 2D :72-111) -----------------------
  int calc_hash() {
    if (n == 0) return ORC_OK;
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (long i = 0; i < (long)n; i++) {
      u32 c[3];
      if (!cell_of(mem[i].pos, c)) { bad |= 1; hash[i] = 0; }
      else hash[i] = hash_of(c);
      index[i] = (u32)i;
    }
    return bad ? ORC_EDOMAIN : ORC_OK;
  }

  // ---- A3 sort_particles: stable sort of (hash, index) by hash (:119-129) -------------
  int sort() {
    if (n == 0) return ORC_OK;
    std::vector<std::pair<u32, u32>> kv(n);
    for (u32 i = 0; i < n; i++) kv[i] = {hash[i], index[i]};
    auto cmp = [](const std::pair<u32, u32> &a, const std::pair<u32, u32> &b) {
      return a.first < b.first;
    };
#ifdef _OPENMP
    __gnu_parallel::stable_sort(kv.begin(), kv.end(), cmp);
#else
    std::stable_sort(kv.begin(), kv.end(), cmp);
#endif
    for (u32 i = 0; i < n; i++) { hash[i] = kv[i].first; index[i] = kv[i].second; }
    return ORC_OK;
  }

  // ---- A4 find_start_end (:131-202) ---------------------------------------------------
  int cell_ranges() {
    if (n == 0) return ORC_OK;
    std::fill(cstart.begin(), cstart.end(), 0xffffffffu);
    std::fill(cend.begin(), cend.end(), 0u);
    for (u32 k = 0; k < n; k++) {
      u32 h = hash[k];
      if (k == 0 || h != hash[k - 1]) {
        cstart[h] = k;
        if (k > 0) cend[hash[k - 1]] = k;
      }
      if (k == n - 1) cend[h] = k + 1;
    }
    lists_valid = true;
    return ORC_OK;
  }

  // ---- A5 density + pressure (3D-2 :338-677; 3D-4 :345-684; 2D :268-460) --------------
  int density() {
    if (n == 0) return ORC_EEMPTY;
#pragma omp parallel for schedule(dynamic, 256)
    for (long il = 0; il < (long)n; il++) {
      u32 i = (u32)il;
      P<R> &pi = mem[i];
      int L = lv(pi.level);
      R k2 = p.k2[L], k6 = p.k6[L], poly6 = p.poly6[L], mi = p.mass[L];
      u32 c[3];
      R total = 0;
      if (cell_of(pi.pos, c)) {
        for (int o = 0; o < nnb(); o++) {
          u32 h;
          if (!neighbor(o, c, h)) continue;
          R cell_sum = 0;
          u32 s = cstart[h];
          if (s != 0xffffffffu) {
            u32 e = cend[h];
            for (u32 k = s; k < e; k++) {
              u32 j = index[k];
              const P<R> &pj = mem[j];
              if (j != i) {
                R dx = pj.pos[0] - pi.pos[0], dy = pj.pos[1] - pi.pos[1];
                R r2;
                if (p.dim == 3) {
                  R dz = pj.pos[2] - pi.pos[2];
                  r2 = dx * dx + dy * dy + dz * dz;
                } else {
                  r2 = dx * dx + dy * dy;
                }
                if (r2 < p.dens_eps) continue;
                if (r2 < k2) {
                  R mj = p.mass[lv(pj.level)];
                  R d = k2 - r2;
                  cell_sum = cell_sum + mj * poly6 * ((d * d) * d);
                }
              }
              if (j == i) cell_sum = cell_sum + mi * poly6 * k6;
            }
          }
          total = total + cell_sum;
        }
      }
      pi.dens = total;
      R q = pi.dens / p.rest_density;
      pi.press = ((q * q) * q - 1) * p.gas_constant;
    }
    return ORC_OK;
  }

  // ---- A6 force (3D-2 :679-1137; 3D-4 :686-1144; 2D :462-790) -------------------------
  int force() {
    if (n == 0) return ORC_EEMPTY;
    const bool d3 = p.dim == 3;
    uint8_t *br = branch_rec(true);
    const bool replay_br = br && replay_topo;
#pragma omp parallel for schedule(dynamic, 256)
    for (long il = 0; il < (long)n; il++) {
      u32 i = (u32)il;
      P<R> &pi = mem[i];
      int Li = lv(pi.level);
      R hi = p.kernel[Li], mi = p.mass[Li], k2i = p.k2[Li], k6i = p.k6[Li];
      R spi = p.spiky[Li], vii = p.visco[Li], gi = p.grad[Li], lci = p.lapc[Li];
      R tf[3] = {0, 0, 0};
      R sf[4] = {0, 0, 0, 0};  // surface gradient xyz (2D: xy) and Laplacian (w; 2D: z)
      u32 c[3];
      if (cell_of(pi.pos, c)) {
        for (int o = 0; o < nnb(); o++) {
          u32 h;
          if (!neighbor(o, c, h)) continue;
          R cf[3] = {0, 0, 0};
          u32 s = cstart[h];
          if (s != 0xffffffffu) {
            u32 e = cend[h];
            for (u32 k = s; k < e; k++) {
              u32 j = index[k];
              const P<R> &pj = mem[j];
              if (j != i) {
                R rel[3] = {pi.pos[0] - pj.pos[0], pi.pos[1] - pj.pos[1],
                            d3 ? pi.pos[2] - pj.pos[2] : R(0)};
                R r2 = d3 ? rel[0] * rel[0] + rel[1] * rel[1] + rel[2] * rel[2]
                          : rel[0] * rel[0] + rel[1] * rel[1];
                int Lj = lv(pj.level);
                R hj = p.kernel[Lj], mj = p.mass[Lj], k2j = p.k2[Lj], k6j = p.k6[Lj];
                R spj = p.spiky[Lj], vij = p.visco[Lj];
                R maxk2 = k2i > k2j ? k2i : k2j;
                if (r2 < maxk2) {
                  if (r2 < R(0.0001f)) continue;
                  R r = msqrt(r2);
                  R iV = mi / pi.dens;
                  R jV = mj / pj.dens;
                  R ikr = hi - r, jkr = hj - r;
                  R ipk, ivk, jpk, jvk;
                  if (ikr > 0) {
                    ipk = spi * ikr * ikr / k6i;
                    ivk = vii / k6i * (ikr);
                    R d = k2i - r2;
                    if (d3) {
                      R t = (-1) * gi * mj / pj.dens * (d * d);
                      sf[0] += t * rel[0];
                      sf[1] += t * rel[1];
                      sf[2] += t * rel[2];
                      sf[3] += lci * mj / pj.dens * d * (r2 - R(0) * d);
                    } else {
                      R t = (-1) * gi * jV * (d * d);
                      sf[0] += t * rel[0];
                      sf[1] += t * rel[1];
                      sf[2] += lci * jV * d * (r2 - R(0) * d);
                    }
                  } else {
                    ipk = 0; ivk = 0;
                  }
                  if (jkr > 0) {
                    jpk = spj * jkr * jkr / k6j;
                    jvk = vij / k6j * (jkr);
                  } else {
                    jpk = 0; jvk = 0;
                  }
                  R tp = mi * mj * (pi.press / (pi.dens * pi.dens) + pj.press / (pj.dens * pj.dens)) *
                         (ipk + jpk) / 2;
                  for (int a = 0; a < p.dim; a++) cf[a] = cf[a] - rel[a] * tp / r;
                  R tv = (iV * jV) * p.viscosity * (ivk + jvk) / 2;
                  for (int a = 0; a < p.dim; a++) cf[a] = cf[a] + (pj.ev[a] - pi.ev[a]) * tv;
                }
              }
              if (j == i && !d3) {
                // SPH-2D/sph_system.cu:640: adds lapc * m / dens * h^2 * (0 - 0*h^2), a signed
                // zero for finite operands; the reference's uninitialised j_mass is taken
                // as the particle's own mass.
                sf[2] += lci * mi / pi.dens * (k2i) * (R(0) - R(0) * (k2i));
              }
            }
          }
          for (int a = 0; a < p.dim; a++) tf[a] = tf[a] + cf[a];
        }
      }
      if (d3) {
        R gx = sf[0], gy = sf[1], gz = sf[2], lap = sf[3];
        R nrm = msqrt(gx * gx + gy * gy + gz * gz);
        pi.surface = nrm;
        R f[3] = {0, 0, 0};
        bool on = nrm > p.surface_tension;
        if (br) {
          if (replay_br) on = (br[i] >> 6) & 1;
          else br[i] |= on ? 64 : 0;
        }
        if (on) {
          R cl = p.surf_coef * lap;
          f[0] = cl * gx / nrm; f[1] = cl * gy / nrm; f[2] = cl * gz / nrm;
        }
        for (int a = 0; a < 3; a++) pi.acc[a] = tf[a] + f[a];
      } else {
        R gx = sf[0], gy = sf[1], lap = sf[2];
        R nrm = msqrt(gx * gx + gy * gy);
        pi.surface = nrm;
        R f[2] = {0, 0};
        if (nrm > p.surface_tension) {
          R cl = p.surf_coef * lap;
          f[0] = cl * gx / nrm; f[1] = cl * gy / nrm;
        }
        pi.acc[0] = tf[0] + f[0];
        pi.acc[1] = tf[1] + f[1];
        if (p.mouse) {
          R xl = pi.pos[0] - p.extern_click[0], yl = pi.pos[1] - p.extern_click[1];
          if ((xl * xl + yl * yl) <= (p.extern_len * p.extern_len)) {
            pi.acc[0] += p.extern_force[0] * p.extern_coe;
            pi.acc[1] += p.extern_force[1] * p.extern_coe;
          }
        }
      }
    }
    return ORC_OK;
  }

  // ---- A8 integrate (3D-2 :204-336; 3D-4 :204-343; 2D :198-266) -----------------------
  int integrate() {
    if (n == 0) return ORC_OK;
    const R dt = p.time_step;
    uint8_t *br = branch_rec(false);
    const bool replay_br = br && replay_topo;
#pragma omp parallel for schedule(static)
    for (long il = 0; il < (long)n; il++) {
      P<R> &q = mem[il];
      R radius = p.radius[lv(q.level)];
      if (p.dim == 3) {
        if (variant == ORC_V3D2) {
          if (p.push_left && q.pos[0] < p.world[0] / 3 && q.pos[1] > p.world[1] / 5)
            q.acc[0] = q.acc[0] + R(20.0f);
        } else {
          if (p.push_left && q.pos[0] < p.world[0] / 3 && q.pos[1] > p.world[1] / 10)
            q.acc[0] = q.acc[0] + R(20.0f);
          if (p.push_right && q.pos[0] > p.world[0] / 3 * 2 && q.pos[1] > p.world[1] / 10)
            q.acc[0] = q.acc[0] - R(20.0f);
        }
        R a3[3] = {q.acc[0], q.acc[1], q.acc[2]};
        if (nf > 0) {
          R f[3];
          field_at(q.pos, f);
          for (int a = 0; a < 3; a++) a3[a] = a3[a] + f[a];
        }
        for (int a = 0; a < 3; a++) q.vel[a] = q.vel[a] + a3[a] * dt / q.dens;
        for (int a = 0; a < 3; a++) q.vel[a] = q.vel[a] + p.gravity[a] * dt / q.dens;
        for (int a = 0; a < 3; a++) q.pos[a] = q.pos[a] + q.vel[a] * dt;
        if (p.use_cylinder) {
          for (int k = 0; k < 3; k++) {
            const Sph<R> &s = p.sphere[k];
            R dv[3] = {q.pos[0] - s.pos[0], q.pos[1] - s.pos[1], q.pos[2] - s.pos[2]};
            R dist = msqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
            if (dist < s.radius + radius) {
              R rr = s.radius + radius;
              R dxyz[3];
              for (int a = 0; a < 3; a++) dxyz[a] = (q.pos[a] - s.pos[a]) / dist * rr;
              R ln = msqrt(dxyz[0] * dxyz[0] + dxyz[1] * dxyz[1] + dxyz[2] * dxyz[2]);
              R nm[3] = {dxyz[0] / ln, dxyz[1] / ln, dxyz[2] / ln};
              for (int a = 0; a < 3; a++) q.pos[a] = s.pos[a] + dxyz[a];
              R dv2 = q.vel[0] * nm[0] + q.vel[1] * nm[1] + q.vel[2] * nm[2];
              R sc = s.damping * (dv2);
              for (int a = 0; a < 3; a++) q.vel[a] = q.vel[a] - sc * nm[a];
            }
          }
        }
        for (int a = 0; a < 3; a++) {
          bool hi = q.pos[a] >= p.world[a] - radius;
          if (replay_br) hi = (br[il] >> (2 * a)) & 1;
          if (hi) {
            q.vel[a] = q.vel[a] * p.wall_damping;
            q.pos[a] = p.world[a] - radius;
          }
          bool lo = q.pos[a] < radius;
          if (replay_br) lo = (br[il] >> (2 * a + 1)) & 1;
          if (lo) {
            q.vel[a] = q.vel[a] * p.wall_damping;
            q.pos[a] = radius;
          }
          if (br && !replay_br) br[il] |= (uint8_t)((hi ? 1 : 0) << (2 * a)) | (uint8_t)((lo ? 1 : 0) << (2 * a + 1));
        }
        for (int a = 0; a < 3; a++) q.ev[a] = (q.ev[a] + q.vel[a]) / 2;
      } else {
        for (int a = 0; a < 2; a++) q.acc[a] = q.acc[a] + p.gravity[a] * dt;
        for (int a = 0; a < 2; a++) q.acc[a] = q.acc[a] / q.dens;
        for (int a = 0; a < 2; a++) q.vel[a] = q.vel[a] + q.acc[a] * dt;
        for (int a = 0; a < 2; a++) q.ev[a] = (q.ev[a] + q.vel[a]) / 2;
        for (int a = 0; a < 2; a++) q.pos[a] = q.pos[a] + q.vel[a] * dt;
        for (int a = 0; a < 2; a++) {
          if (q.pos[a] >= p.world[a] - radius / 2) {
            q.vel[a] = q.vel[a] * p.wall_damping;
            q.pos[a] = p.world[a] - radius / 2;
          }
          if (q.pos[a] < radius / 2) {
            q.vel[a] = q.vel[a] * p.wall_damping;
            q.pos[a] = radius / 2;
          }
        }
      }
    }
    return ORC_OK;
  }

  // ---- learnable force field and density sample (the differentiable rollout) ---------
  // Trilinear sample of the field at pos: g = pos / world * (nf - 1) clamped to
  // [0, nf - 1], cell i = min(floor(g), nf - 2), t = g - i; x, then y, then z blends.
  void field_at(const R pos[3], R f[3]) const {
    const R top = (R)(nf - 1);
    int i0[3];
    R t[3];
    for (int a = 0; a < 3; a++) {
      R g = pos[a] / p.world[a] * top;
      if (!(g > R(0))) g = R(0);
      if (g > top) g = top;
      int i = (int)mfloor(g);
      if (i > nf - 2) i = nf - 2;
      i0[a] = i;
      t[a] = g - (R)i;
    }
    const R ux = R(1) - t[0], uy = R(1) - t[1], uz = R(1) - t[2];
    for (int c = 0; c < 3; c++) {
      auto th = [&](int dx, int dy, int dz) {
        return field[(((size_t)(i0[2] + dz) * nf + (size_t)(i0[1] + dy)) * nf + (size_t)(i0[0] + dx)) * 3 + c];
      };
      const R c00 = th(0, 0, 0) * ux + th(1, 0, 0) * t[0];
      const R c10 = th(0, 1, 0) * ux + th(1, 1, 0) * t[0];
      const R c01 = th(0, 0, 1) * ux + th(1, 0, 1) * t[0];
      const R c11 = th(0, 1, 1) * ux + th(1, 1, 1) * t[0];
      const R c0 = c00 * uy + c10 * t[1];
      const R c1 = c01 * uy + c11 * t[1];
      f[c] = c0 * uz + c1 * t[2];
    }
  }

  template <class T> int set_field(const T *theta, int n_nodes) {
    if (n_nodes == 0) {
      nf = 0;
      field.clear();
      return ORC_OK;
    }
    if (p.dim != 3 || n_nodes < 2 || !theta) return ORC_EINVAL;
    nf = n_nodes;
    field.assign((size_t)nf * nf * nf * 3, R(0));
    for (size_t k = 0; k < field.size(); k++) field[k] = (R)theta[k];
    return ORC_OK;
  }

  // Density of the current particles at the centres of an ns^3 sample grid over the box
  // (x fastest): voxel v at ((float)v + 0.5) * world / ns, the poly6 splat with the
  // particle radius of the density grid (SPH-3D-2/sph_system.cu:1954-1990), the 27 cells
  // around the voxel's cell in offset order, particles in cell order.
  int density_sample(int ns, double *out) {
    if (!out) return ORC_EINVAL;
    std::vector<R> v((size_t)ns * ns * ns);
    int rc = density_sample_r(ns, v.data());
    if (rc) return rc;
    for (size_t k = 0; k < v.size(); k++) out[k] = to_d(v[k]);
    return ORC_OK;
  }
  int density_sample_r(int ns, R *out) {
    if (p.dim != 3 || ns <= 0 || !out) return ORC_EINVAL;
    if (n == 0) return ORC_EEMPTY;
    int rc = calc_hash();
    if (rc) return rc;
    sort();
    cell_ranges();
    const R r2s[2] = {p.radius[0] * p.radius[0], p.radius[1] * p.radius[1]};
    R hs[3];
    for (int a = 0; a < 3; a++) hs[a] = p.world[a] / (R)ns;
    const long tot = (long)ns * ns * ns;
#pragma omp parallel for schedule(dynamic, 1024)
    for (long vl = 0; vl < tot; vl++) {
      const int vv[3] = {(int)(vl % ns), (int)(vl / ns % ns), (int)(vl / ((long)ns * ns))};
      R x[3];
      u32 c[3];
      for (int a = 0; a < 3; a++) {
        x[a] = ((R)vv[a] + R(0.5)) * hs[a];
        R q = mfloor(x[a] / p.cell_size);
        if (q > (R)(p.ncell[a] - 1)) q = (R)(p.ncell[a] - 1);
        c[a] = (u32)q;
      }
      R acc = 0;
      for (int o = 0; o < 27; o++) {
        u32 h;
        if (!neighbor(o, c, h)) continue;
        u32 st = cstart[h];
        if (st == 0xffffffffu) continue;
        u32 e = cend[h];
        for (u32 k = st; k < e; k++) {
          const P<R> &q = mem[index[k]];
          const int L = lv(q.level);
          const R rx = q.pos[0] - x[0], ry = q.pos[1] - x[1], rz = q.pos[2] - x[2];
          const R r2 = rx * rx + ry * ry + rz * rz;
          if (r2 < r2s[L]) {
            const R d = r2s[L] - r2;
            acc += p.mass[L] * p.poly6_radius[L] * ((d * d) * d);
          }
        }
      }
      out[vl] = acc;
    }
    return ORC_OK;
  }

  // ---- A9 energy (3D-2 :1169-1546; 3D-4 :1176-1553; 2D :827-1065) ---------------------
  // Runs after integrate on the start-of-step cell lists with the new positions.
  int energy() {
    if (n == 0) return ORC_EEMPTY;
    const bool d3 = p.dim == 3;
#pragma omp parallel for schedule(dynamic, 256)
    for (long il = 0; il < (long)n; il++) {
      u32 i = (u32)il;
      P<R> &pi = mem[i];
      // 3D uses the large kernel for both levels (3D-2 :1199-1209); 2D the level kernel.
      int Lk = d3 ? 0 : lv(pi.level);
      R ker = p.kernel[Lk], ker2 = p.k2[Lk], sk = p.sqrt_kernel[Lk];
      R tt[3] = {0, 0, 0}, tv[3] = {0, 0, 0};
      u32 c[3];
      if (cell_of(pi.pos, c)) {
        for (int o = 0; o < nnb(); o++) {
          u32 h;
          if (!neighbor(o, c, h)) continue;
          R ct[3] = {0, 0, 0};
          u32 s = cstart[h];
          if (s != 0xffffffffu) {
            u32 e = cend[h];
            for (u32 k = s; k < e; k++) {
              u32 j = index[k];
              if (j == i) continue;
              const P<R> &pj = mem[j];
              R rel[3] = {pi.pos[0] - pj.pos[0], pi.pos[1] - pj.pos[1],
                          d3 ? pi.pos[2] - pj.pos[2] : R(0)};
              R r2 = d3 ? rel[0] * rel[0] + rel[1] * rel[1] + rel[2] * rel[2]
                        : rel[0] * rel[0] + rel[1] * rel[1];
              if (r2 < ker2) {
                // 3D skips r2 < 1e-4; 2D clamps r2 to EXT2 and never uses it again.
                if (d3 && r2 < R(0.0001f)) continue;
                R t[3];
                t[0] = wavelet<R>((pi.pos[0] - pj.pos[0]) / ker);
                t[1] = wavelet<R>((pi.pos[1] - pj.pos[1]) / ker);
                t[2] = d3 ? wavelet<R>((pi.pos[2] - pj.pos[2]) / ker) : R(0);
                for (int a = 0; a < p.dim; a++) ct[a] = ct[a] + t[a];
                for (int a = 0; a < p.dim; a++) tv[a] = tv[a] + (pj.vel[a] * t[a] / sk);
              }
            }
          }
          for (int a = 0; a < p.dim; a++) tt[a] = tt[a] + ct[a];
        }
      }
      for (int a = 0; a < p.dim; a++) tv[a] = tv[a] / tt[a];
      R e2 = tv[0] * tv[0] + tv[1] * tv[1];
      if (variant == ORC_V3D4) pi.energy = R(0.5f) * p.mass[lv(pi.level)] * (e2);
      else pi.energy = R(0.5f) * (e2);
    }
    return ORC_OK;
  }

  // ---- A10-A12 decide + split (3D-2 :1548-1659; 3D-4 :1555-1700; 2D :1067-1175) -------
  // Split parents are taken in ascending slot order. Capacity policy: only the first
  // floor((cap - N)/(n - 1)) parents are split (the reference has no check).
  int do_split(u32 *nreq, u32 *ndone) {
    if (nreq) *nreq = 0;
    if (ndone) *ndone = 0;
    if (n == 0) return ORC_OK;
#pragma omp parallel for schedule(static)
    for (long il = 0; il < (long)n; il++) {
      const P<R> &q = mem[il];
      split[il] = (q.energy * p.co_energy + q.surface * p.co_surface > p.split_criteria &&
                   q.level == 1) ? 0u : 1u;
    }
    if (replay_topo) {
      if (topo_step >= topo_split.size()) return ORC_EINVAL;
      for (u32 i = 0; i < n; i++) split[i] = 1u;
      for (u32 a : topo_split[topo_step]) {
        if (a >= n) return ORC_EINVAL;
        split[a] = 0u;
      }
    }
    u32 S = 0;
    n_flagged = n;
    for (u32 i = 0; i < n; i++)
      if (split[i] == 0) sindex[S++] = i;
    u32 per = (u32)(p.nsplit - 1);
    u32 room = (cap - n) / per;
    u32 Se = S < room ? S : room;
    const R hs = p.kernel[1] / 2;
    if (rec_topo) topo_split.emplace_back(sindex.begin(), sindex.begin() + Se);
    for (u32 k = 0; k < Se; k++) {
      u32 a = sindex[k];
      P<R> &par = mem[a];
      if (p.nsplit == 2) {
        P<R> &ch = mem[n + k];
        ch = par;  // fields the reference leaves unwritten (3D: dens, press, surface,
                   // tension) are defined here as copies of the parent
        ch.level = 2;
        ch.pos[0] = par.pos[0] - hs;
        par.level = 2;
        par.pos[0] = par.pos[0] + hs;
      } else {
        const R e4 = par.energy / 4;
        for (int c = 0; c < 3; c++) {
          P<R> &ch = mem[n + 3 * k + c];
          ch = par;
          ch.level = 2;
          ch.energy = e4;
        }
        mem[n + 3 * k + 0].pos[0] = par.pos[0] - hs;
        mem[n + 3 * k + 1].pos[2] = par.pos[2] - hs;
        mem[n + 3 * k + 2].pos[2] = par.pos[2] + hs;
        par.level = 2;
        par.pos[0] = par.pos[0] + hs;
        par.energy = e4;
      }
      status[a] = 0;
    }
    for (u32 k = 0; k < Se * per; k++) status[n + k] = 0;
    n = n + Se * per;
    if (nreq) *nreq = S;
    if (ndone) *ndone = Se;
    return ORC_OK;
  }

  // ---- A13 merge (3D-2 :1661-1815; 3D-4 :1702-1914; 2D :1177-1296) --------------------
  // One independent walk per cell over the start-of-step lists; statuses of listed
  // particles are reset to 0, then the first floor(q/n)*n qualifying particles form
  // groups of n in list order; the first of each group survives.
  int merge(u32 *groups_out) {
    u32 groups = 0;
    if (groups_out) *groups_out = 0;
    if (!lists_valid) return ORC_EINVAL;
    if (replay_topo) return merge_replay(groups_out);
    const int nm = p.nsplit;
    const bool d3 = p.dim == 3;
    n_status = n;
    // 2D qualifies on the energy/surface of the particle whose SLOT equals the cell id
    // (SPH-2D/sph_system.cu:1207, :1230, :1245), read afresh at every test. When that
    // particle is listed in the cell itself, only this cell's walk writes it (the walk
    // averages its energy when it heads a pair, :1261), so the reference's reads are
    // deterministic and followed exactly: a pair headed by it changes the qualification of
    // the entries after it. When it lives in another cell, that cell's walk may write it
    // concurrently (a race in the reference); the value before any merge is used then. A
    // cell id at or beyond the live count never qualifies.
    std::vector<char> q2d;
    if (!d3) {
      q2d.assign(p.total_cell, 0);
      for (u32 cidx = 0; cidx < p.total_cell && cidx < n; cidx++)
        q2d[cidx] = (mem[cidx].energy * p.co_energy + mem[cidx].surface * p.co_surface <
                     p.merge_criteria) ? 1 : 0;
    }
    std::vector<std::vector<u32>> parts;  // recorded groups, per thread
#pragma omp parallel
    {
      std::vector<u32> mine;
#pragma omp for schedule(dynamic, 1024) reduction(+ : groups)
    for (long cl = 0; cl < (long)p.total_cell; cl++) {
      u32 s = cstart[cl];
      if (s == 0xffffffffu) continue;
      u32 e = cend[cl];
      u32 cnt = 0;
      bool own = false;  // 2D: the particle of slot cl is in this cell's list
      auto cell_q = [&]() -> bool {
        if (!own) return q2d[cl] != 0;
        const P<R> &c = mem[cl];
        return c.energy * p.co_energy + c.surface * p.co_surface < p.merge_criteria;
      };
      auto qual = [&](u32 j) -> bool {
        const P<R> &q = mem[j];
        if (d3) return q.level != 1 && q.energy * p.co_energy + q.surface * p.co_surface < p.merge_criteria;
        return q.level == 2 && cell_q();
      };
      for (u32 k = s; k < e; k++) {
        u32 j = index[k];
        status[j] = 0;
        if (qual(j)) cnt++;  // no write yet: the cell's particle still holds its start value
      }
      if (!d3)
        for (u32 k = s; k < e; k++) own |= index[k] == (u32)cl;
      if (cnt < (u32)nm) continue;
      u32 ng = cnt / nm;
      u32 cur = s;
      // A search that reaches the end of the list leaves its entry as it was: the previous
      // pair's (:1221-1251; possible in 2D only, after the qualification changed)
      u32 pm[4] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu};
      for (u32 g = 0; g < ng; g++) {
        for (int t = 0; t < nm; t++) {
          while (cur < e) {
            u32 j = index[cur++];
            if (qual(j)) { pm[t] = j; break; }
          }
        }
        merge_group(pm);
        if (rec_topo) mine.insert(mine.end(), pm, pm + nm);
        groups++;
      }
    }
#pragma omp critical
      parts.push_back(std::move(mine));
    }  // omp parallel
    if (rec_topo) {
      // groups in ascending order of their survivor (the walks ran in any order)
      std::vector<u32> all;
      for (auto &v : parts) all.insert(all.end(), v.begin(), v.end());
      std::vector<u32> order(all.size() / nm);
      for (u32 g = 0; g < order.size(); g++) order[g] = g;
      std::sort(order.begin(), order.end(), [&](u32 a, u32 b) { return all[a * nm] < all[b * nm]; });
      std::vector<u32> rec;
      for (u32 g : order) rec.insert(rec.end(), all.begin() + g * nm, all.begin() + (g + 1) * nm);
      topo_merge.push_back(std::move(rec));
    }
    if (groups_out) *groups_out = groups;
    return ORC_OK;
  }

  // The survivor pm[0] takes level 1, then the group's volume-weighted average (3D; V =
  // m(level)/dens with the survivor's level already 1, SPH-3D-2/sph_system.cu:1742-1772)
  // or plain average (2D, :1257-1262); the others are removed.
  void merge_group(const u32 *pm) {
    const int nm = p.nsplit;
    P<R> &p1 = mem[pm[0]];
    p1.level = 1;
    if (p.dim == 3) {
      R V[4] = {R(0), R(0), R(0), R(0)};
      for (int t = 0; t < nm; t++) V[t] = p.mass[lv(mem[pm[t]].level)] / mem[pm[t]].dens;
      R vs = V[0];
      for (int t = 1; t < nm; t++) vs = vs + V[t];
      R *f1[4] = {p1.pos, p1.vel, p1.ev, p1.acc};
      for (int fi = 0; fi < 4; fi++) {
        for (int a = 0; a < 3; a++) {
          R acc = f1[fi][a] * V[0];
          for (int t = 1; t < nm; t++) {
            const P<R> &pt = mem[pm[t]];
            const R *ft = fi == 0 ? pt.pos : fi == 1 ? pt.vel : fi == 2 ? pt.ev : pt.acc;
            acc = acc + ft[a] * V[t];
          }
          f1[fi][a] = acc / vs;
        }
      }
    } else {
      const P<R> &p2 = mem[pm[1]];
      for (int a = 0; a < 2; a++) {
        p1.pos[a] = (p1.pos[a] + p2.pos[a]) / 2;
        p1.vel[a] = (p1.vel[a] + p2.vel[a]) / 2;
        p1.ev[a] = (p1.ev[a] + p2.ev[a]) / 2;
        p1.acc[a] = (p1.acc[a] + p2.acc[a]) / 2;
      }
      p1.energy = (p1.energy + p2.energy) / 2;
    }
    status[pm[0]] = 0;
    for (int t = 1; t < nm; t++) status[pm[t]] = 1;
  }

  // merge with the recorded groups of this step (replay_topo)
  int merge_replay(u32 *groups_out) {
    if (topo_step >= topo_merge.size()) return ORC_EINVAL;
    const std::vector<u32> &g = topo_merge[topo_step];
    const u32 nm = (u32)p.nsplit;
    n_status = n;
    for (u32 i = 0; i < n; i++) status[i] = 0;
    for (size_t k = 0; k + nm <= g.size(); k += nm) {
      for (u32 t = 0; t < nm; t++)
        if (g[k + t] >= n) return ORC_EINVAL;
      merge_group(&g[k]);
    }
    if (groups_out) *groups_out = (u32)(g.size() / nm);
    return ORC_OK;
  }

  // ---- A14 rearrange: stable partition, survivors (status 0) first (:1818-1840) -------
  int compact() {
    if (n == 0) return ORC_OK;
    u32 w = 0;
    for (u32 i = 0; i < n; i++) {
      if (status[i] == 0) {
        if (w != i) mem[w] = mem[i];
        w++;
      }
    }
    n = w;
    lists_valid = false;
    return ORC_OK;
  }

  // ---- A15 density grid (3D-2 :1842-2219; 3D-4 :1941-2320) ----------------------------
  int density_grid() {
    if (p.dim != 3 || p.tot_dens == 0) return ORC_EINVAL;
    if (!lists_valid) return ORC_EINVAL;
    const u32 rd = p.nd[0], cd = p.nd[1];
    const R r2s[2] = {p.radius[0] * p.radius[0], p.radius[1] * p.radius[1]};
#pragma omp parallel for schedule(dynamic, 4096)
    for (long vl = 0; vl < (long)p.tot_dens; vl++) {
      u32 v = (u32)vl;
      u32 vx = v % (rd * cd) % rd, vy = v % (rd * cd) / rd, vz = v / (rd * cd);
      R acc = 0;
      if (!(vx == 0 || vy == 0 || vz == 0)) {
        // voxel position (float)v * den_size (SPH-3D-2/sph_system.cpp:214-226)
        R dp[3] = {(R)(float)vx * p.den_size, (R)(float)vy * p.den_size, (R)(float)vz * p.den_size};
        u32 c[3];
        const u32 vv[3] = {vx, vy, vz};
        for (int a = 0; a < 3; a++)
          c[a] = (u32)mfloor((R)(float)vv[a] / (R)(float)p.nd[a] * (R)(float)p.ncell[a]);
        for (int o = 0; o < 27; o++) {
          u32 h;
          if (!neighbor(o, c, h)) continue;
          u32 s = cstart[h];
          if (s == 0xffffffffu) continue;
          u32 e = cend[h];
          for (u32 k = s; k < e; k++) {
            const P<R> &q = mem[index[k]];
            int L = lv(q.level);
            R rx = q.pos[0] - dp[0], ry = q.pos[1] - dp[1], rz = q.pos[2] - dp[2];
            R r2 = rx * rx + ry * ry + rz * rz;
            if (r2 < r2s[L]) {
              R d = r2s[L] - r2;
              acc += p.mass[L] * p.poly6_radius[L] * ((d * d) * d);
            }
          }
        }
      }
      dgrid[v] = acc;
    }
    return ORC_OK;
  }

  // ---- A16 normals (3D-2 :2221-2351) --------------------------------------------------
  int normals() {
    if (p.dim != 3 || p.tot_dens == 0) return ORC_EINVAL;
    const u32 rd = p.nd[0], cd = p.nd[1], ld = p.nd[2];
    const R ds = p.den_size;
#pragma omp parallel for schedule(static)
    for (long vl = 0; vl < (long)p.tot_dens; vl++) {
      u32 v = (u32)vl;
      u32 vx = v % (rd * cd) % rd, vy = v % (rd * cd) / rd, vz = v / (rd * cd);
      const u32 vv[3] = {vx, vy, vz};
      const u32 nn[3] = {rd, cd, ld};
      const u32 st[3] = {1, rd, rd * cd};
      R g[3];
      for (int a = 0; a < 3; a++) {
        R da = vv[a] == 0 ? R(0) : dgrid[v - st[a]];
        R db = vv[a] == nn[a] - 1 ? R(0) : dgrid[v + st[a]];
        g[a] = (da - db) / ds;
      }
      R nr = msqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
      for (int a = 0; a < 3; a++) normal[(size_t)v * 3 + a] = nr == R(0) ? R(0) : g[a] / nr;
    }
    return ORC_OK;
  }

  // ---- A17 marching cubes (SPH-3D-2/sph_system.cpp:417-542) ---------------------------
  // Cubes in x-major order (x outer, z inner); corner inside if value > ISOVALUE (0.5f);
  // edge vertices and normals at edge midpoints. Parallel count / scan / emit gives the
  // same triangle order as the serial loop.
  int marching_cubes(u32 *ntri) {
    if (p.dim != 3 || p.tot_dens == 0) return ORC_EINVAL;
    const u32 rd = p.nd[0], cd = p.nd[1], ld = p.nd[2];
    const u32 nx = rd - 1, ny = cd - 1, nz = ld - 1;
    const size_t ncube = (size_t)nx * ny * nz;
    std::vector<unsigned char> cnt(ncube);
    auto vid = [&](u32 x, u32 y, u32 z) { return (size_t)z * rd * cd + (size_t)y * rd + x; };
    auto cube_case = [&](u32 x, u32 y, u32 z) {
      int f = 0;
      for (int c = 0; c < 8; c++)
        if ((float)dgrid[vid(x + CORNER[c][0], y + CORNER[c][1], z + CORNER[c][2])] > 0.5f)
          f |= 1 << c;
      return f;
    };
#pragma omp parallel for schedule(static)
    for (long xl = 0; xl < (long)nx; xl++)
      for (u32 y = 0; y < ny; y++)
        for (u32 z = 0; z < nz; z++) {
          int f = cube_case((u32)xl, y, z);
          int t = 0;
          while (t < 5 && orc_mc_tri[f][3 * t] >= 0) t++;
          cnt[((size_t)xl * ny + y) * nz + z] = (unsigned char)t;
        }
    std::vector<size_t> off(ncube + 1);
    off[0] = 0;
    for (size_t k = 0; k < ncube; k++) off[k + 1] = off[k] + cnt[k];
    size_t total = off[ncube];
    tris.assign(total * 18, 0.0f);
    const R ds = p.den_size;
#pragma omp parallel for schedule(dynamic, 1)
    for (long xl = 0; xl < (long)nx; xl++)
      for (u32 y = 0; y < ny; y++)
        for (u32 z = 0; z < nz; z++) {
          size_t cid = ((size_t)xl * ny + y) * nz + z;
          if (!cnt[cid]) continue;
          u32 x = (u32)xl;
          int f = cube_case(x, y, z);
          R cp[8][3], cn[8][3];
          for (int c = 0; c < 8; c++) {
            u32 X = x + CORNER[c][0], Y = y + CORNER[c][1], Z = z + CORNER[c][2];
            cp[c][0] = (R)(float)X * ds; cp[c][1] = (R)(float)Y * ds; cp[c][2] = (R)(float)Z * ds;
            size_t v = vid(X, Y, Z);
            for (int a = 0; a < 3; a++) cn[c][a] = normal[v * 3 + a];
          }
          R ev[12][3], en[12][3];
          for (int e = 0; e < 12; e++) {
            int a0 = EDGE[e][0], a1 = EDGE[e][1];
            if (((f >> a0) & 1) == ((f >> a1) & 1)) continue;
            for (int a = 0; a < 3; a++) {
              ev[e][a] = (cp[a0][a] + cp[a1][a]) / 2;
              en[e][a] = (cn[a0][a] + cn[a1][a]) / 2;
            }
          }
          size_t o = off[cid];
          for (int t = 0; t < cnt[cid]; t++) {
            float *dst = &tris[(o + t) * 18];
            for (int vtx = 0; vtx < 3; vtx++) {
              int e = orc_mc_tri[f][3 * t + vtx];
              for (int a = 0; a < 3; a++) {
                dst[vtx * 3 + a] = (float)ev[e][a];
                dst[9 + vtx * 3 + a] = (float)en[e][a];
              }
            }
          }
        }
    if (ntri) *ntri = (u32)total;
    return ORC_OK;
  }

  // ---- A19 one step (SPH-3D-2/sph_system.cpp:302-404; 3D-4 :350-453; 2D :190-264) -----
  int step(u32 flags) {
    if (n == 0) return ORC_EEMPTY;
    int rc;
    if ((rc = calc_hash())) return rc;
    sort();
    cell_ranges();
    if ((rc = density())) return rc;
    if ((rc = force())) return rc;
    integrate();
    if (p.dim == 3 && (flags & ORC_MC)) {
      density_grid();
      normals();
      marching_cubes(nullptr);
    }
    if ((rc = energy())) return rc;
    if (flags & ORC_SPLIT) {
      if ((rc = do_split(nullptr, nullptr))) return rc;
    } else if (rec_topo) {
      topo_split.emplace_back();
    }
    if (flags & ORC_MERGE) {
      if ((rc = merge(nullptr))) return rc;
      compact();
    } else if (rec_topo) {
      topo_merge.emplace_back();
    }
    topo_step++;
    // force_waiting* and the 2D mouse force last one step (3D-2 sph_system.cpp:403).
    p.push_left = p.push_right = false;
    p.mouse = false;
    if (variant == ORC_V2D) raw2.force_waiting = false;
    else { raw3.b0 = false; if (variant == ORC_V3D4) raw3.b1 = false; }
    lists_valid = false;
    return ORC_OK;
  }
};

template <class R> Sim<R> *mk(int variant, u32 cap) {
  Sim<R> *s = new (std::nothrow) Sim<R>();
  if (!s) return nullptr;
  s->variant = variant;
  s->fp64 = std::is_same<R, DualF>::value ? 3 : std::is_same<R, Dual>::value ? 2
            : std::is_same<R, double>::value ? 1 : 0;
  s->cap = cap;
  s->mem.resize(cap);
  s->hash.resize(cap);
  s->index.resize(cap);
  s->split.resize(cap);
  s->sindex.resize(cap);
  s->status.resize(cap);
  return s;
}

}  // namespace

namespace {

template <class R> int state_f64(Sim<R> *s, double *out) {
  for (u32 i = 0; i < s->n; i++) {
    const P<R> &q = s->mem[i];
    double *o = out + (size_t)i * 17;
    for (int a = 0; a < 3; a++) {
      o[a] = to_d(q.pos[a]); o[3 + a] = to_d(q.vel[a]); o[6 + a] = to_d(q.ev[a]);
      o[9 + a] = to_d(q.acc[a]);
    }
    o[12] = to_d(q.dens); o[13] = to_d(q.press); o[14] = to_d(q.energy);
    o[15] = to_d(q.surface); o[16] = q.level;
  }
  return ORC_OK;
}

template <class R> size_t count_of(Sim<R> *s, int which) {
  switch (which) {
    case ORC_HASH: case ORC_INDEX: return s->n;
    case ORC_SPLITFLAG: return s->n_flagged;
    case ORC_STATUS: return s->n_status;
    case ORC_START: case ORC_END: return s->p.total_cell;
    case ORC_DGRID: return s->dgrid.size();
    case ORC_NORMAL: return s->normal.size();
    case ORC_TRI: return s->tris.size();
  }
  return 0;
}

template <class R> int get_u32(Sim<R> *s, int which, uint32_t *out, size_t n) {
  const std::vector<u32> *v = nullptr;
  switch (which) {
    case ORC_HASH: v = &s->hash; break;
    case ORC_INDEX: v = &s->index; break;
    case ORC_START: v = &s->cstart; break;
    case ORC_END: v = &s->cend; break;
    case ORC_SPLITFLAG: v = &s->split; break;
    case ORC_STATUS: v = &s->status; break;
    default: return ORC_EINVAL;
  }
  if (n > v->size()) return ORC_EINVAL;
  std::memcpy(out, v->data(), n * 4);
  return ORC_OK;
}

template <class R> int get_f32(Sim<R> *s, int which, float *out, size_t n) {
  if (which == ORC_TRI) {
    if (n > s->tris.size()) return ORC_EINVAL;
    std::memcpy(out, s->tris.data(), n * 4);
    return ORC_OK;
  }
  const std::vector<R> *v = which == ORC_DGRID ? &s->dgrid : which == ORC_NORMAL ? &s->normal : nullptr;
  if (!v || n > v->size()) return ORC_EINVAL;
  for (size_t k = 0; k < n; k++) out[k] = (float)(*v)[k];
  return ORC_OK;
}
}  // namespace

#define DISPATCH(s, call)                                                              \
  ((s)->fp64 == 3   ? static_cast<Sim<DualF> *>(s)->call                                  \
   : (s)->fp64 == 2 ? static_cast<Sim<Dual> *>(s)->call                                   \
   : (s)->fp64 == 1 ? static_cast<Sim<double> *>(s)->call                                 \
                    : static_cast<Sim<float> *>(s)->call)
// the same for free functions taking the simulation
#define DISPATCH_FN(fn, s, ...)                                                         \
  ((s)->fp64 == 3   ? fn(static_cast<Sim<DualF> *>(s), __VA_ARGS__)                       \
   : (s)->fp64 == 2 ? fn(static_cast<Sim<Dual> *>(s), __VA_ARGS__)                        \
   : (s)->fp64 == 1 ? fn(static_cast<Sim<double> *>(s), __VA_ARGS__)                      \
                    : fn(static_cast<Sim<float> *>(s), __VA_ARGS__))

namespace {
// the derivative builds (DualT<double>, DualT<float>): operations on one of them, given a
// function template taking Sim<D> *; other builds get ORC_EINVAL
#define DUAL_FN(fn, s, ...)                                                              \
  (!(s) ? ORC_EINVAL                                                                      \
   : (s)->fp64 == 3 ? fn(static_cast<Sim<DualF> *>(s), __VA_ARGS__)                       \
   : (s)->fp64 == 2 ? fn(static_cast<Sim<Dual> *>(s), __VA_ARGS__)                        \
                    : ORC_EINVAL)

template <class D> int field_tangent(Sim<D> *d, const double *theta, const double *dtheta, int nf) {
  if (nf < 2 || !theta || !dtheta) return ORC_EINVAL;
  int rc = d->set_field(theta, nf);
  if (rc) return rc;
  for (size_t k = 0; k < d->field.size(); k++) d->field[k].d = dtheta[k];
  return ORC_OK;
}
template <class D> int state_tangent(Sim<D> *d, const double *dpos, const double *dvel, const double *dev) {
  if (d->p.dim != 3) return ORC_EINVAL;
  for (u32 i = 0; i < d->n; i++)
    for (int a = 0; a < 3; a++) {
      d->mem[i].pos[a].d = dpos ? dpos[3 * (size_t)i + a] : 0.0;
      d->mem[i].vel[a].d = dvel ? dvel[3 * (size_t)i + a] : 0.0;
      d->mem[i].ev[a].d = dev ? dev[3 * (size_t)i + a] : 0.0;
    }
  return ORC_OK;
}
template <class D> int sample_tangent(Sim<D> *d, int ns, double *value, double *tangent) {
  if (!value || !tangent || ns <= 0) return ORC_EINVAL;
  std::vector<D> out((size_t)ns * ns * ns);
  int rc = d->density_sample_r(ns, out.data());
  if (rc) return rc;
  for (size_t k = 0; k < out.size(); k++) {
    value[k] = (double)out[k].v;
    tangent[k] = out[k].d;
  }
  return ORC_OK;
}

template <class R> u32 num_of(Sim<R> *s, int) { return s->n; }
template <class R> int topo_mode(Sim<R> *s, int mode) {
  s->rec_topo = mode == 1;
  s->replay_topo = mode == 2;
  s->topo_step = 0;
  if (mode == 1) {
    s->topo_split.clear();
    s->topo_merge.clear();
    s->topo_branch.clear();
  }
  return ORC_OK;
}
struct TopoRef {
  std::vector<std::vector<u32>> *sp, *mg;
  std::vector<std::vector<uint8_t>> *br;
};
template <class R> TopoRef topo_get(Sim<R> *s, int) {
  return TopoRef{&s->topo_split, &s->topo_merge, &s->topo_branch};
}
template <class R> u32 nsplit_of(Sim<R> *s, int) { return (u32)s->p.nsplit; }
}  // namespace

extern "C" {

OrcSim *orc_create(int variant, const void *sysparam, uint32_t capacity, int fp64) {
  if (variant < 0 || variant > 2 || !sysparam || capacity == 0 || fp64 < 0 || fp64 > 3)
    return nullptr;
  OrcSim *s = fp64 == 3   ? (OrcSim *)mk<DualF>(variant, capacity)
              : fp64 == 2 ? (OrcSim *)mk<Dual>(variant, capacity)
              : fp64 == 1 ? (OrcSim *)mk<double>(variant, capacity)
                          : (OrcSim *)mk<float>(variant, capacity);
  if (!s) return nullptr;
  if (DISPATCH(s, set_params(sysparam)) != ORC_OK) {
    delete s;
    return nullptr;
  }
  return s;
}
void orc_destroy(OrcSim *s) { delete s; }
void orc_set_threads(int nthr) {
#ifdef _OPENMP
  if (nthr > 0) omp_set_num_threads(nthr);
#else
  (void)nthr;
#endif
}
int orc_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
int orc_set_params(OrcSim *s, const void *sp) { return DISPATCH(s, set_params(sp)); }
int orc_set_particles(OrcSim *s, const void *aos, uint32_t n) { return DISPATCH(s, set_particles(aos, n)); }
int orc_get_particles(OrcSim *s, void *aos) { return DISPATCH(s, get_particles(aos)); }
uint32_t orc_num_particles(OrcSim *s) { return DISPATCH_FN(num_of, s, 0); }

int orc_get_state_f64(OrcSim *s, double *out) { return DISPATCH_FN(state_f64, s, out); }

int orc_calc_hash(OrcSim *s) { return DISPATCH(s, calc_hash()); }
int orc_sort(OrcSim *s) { return DISPATCH(s, sort()); }
int orc_cell_ranges(OrcSim *s) { return DISPATCH(s, cell_ranges()); }
int orc_density(OrcSim *s) { return DISPATCH(s, density()); }
int orc_force(OrcSim *s) { return DISPATCH(s, force()); }
int orc_integrate(OrcSim *s) { return DISPATCH(s, integrate()); }
int orc_energy(OrcSim *s) { return DISPATCH(s, energy()); }
int orc_split(OrcSim *s, uint32_t *a, uint32_t *b) { return DISPATCH(s, do_split(a, b)); }
int orc_merge(OrcSim *s, uint32_t *g) { return DISPATCH(s, merge(g)); }
int orc_compact(OrcSim *s) { return DISPATCH(s, compact()); }
int orc_density_grid(OrcSim *s) { return DISPATCH(s, density_grid()); }
int orc_normals(OrcSim *s) { return DISPATCH(s, normals()); }
int orc_marching_cubes(OrcSim *s, uint32_t *nt) { return DISPATCH(s, marching_cubes(nt)); }
int orc_step(OrcSim *s, uint32_t flags) { return DISPATCH(s, step(flags)); }

size_t orc_count(OrcSim *s, int which) { return DISPATCH_FN(count_of, s, which); }

int orc_get_u32(OrcSim *s, int which, uint32_t *out, size_t n) {
  return DISPATCH_FN(get_u32, s, which, out, n);
}
int orc_get_f32(OrcSim *s, int which, float *out, size_t n) {
  return DISPATCH_FN(get_f32, s, which, out, n);
}

int orc_set_field(OrcSim *s, const float *theta, int nf) { return DISPATCH(s, set_field(theta, nf)); }
int orc_set_field_f64(OrcSim *s, const double *theta, int nf) { return DISPATCH(s, set_field(theta, nf)); }
int orc_density_sample(OrcSim *s, int ns, double *out) { return DISPATCH(s, density_sample(ns, out)); }

int orc_set_field_tangent(OrcSim *s, const double *theta, const double *dtheta, int nf) {
  return DUAL_FN(field_tangent, s, theta, dtheta, nf);
}

int orc_set_state_tangent(OrcSim *s, const double *dpos, const double *dvel, const double *dev) {
  return DUAL_FN(state_tangent, s, dpos, dvel, dev);
}

int orc_density_sample_tangent(OrcSim *s, int ns, double *value, double *tangent) {
  return DUAL_FN(sample_tangent, s, ns, value, tangent);
}

int orc_topology(OrcSim *s, int mode) {
  if (!s || mode < 0 || mode > 2) return ORC_EINVAL;
  return DISPATCH_FN(topo_mode, s, mode);
}

int orc_topology_copy(OrcSim *dst, const OrcSim *src) {
  if (!dst || !src) return ORC_EINVAL;
  OrcSim *cs = const_cast<OrcSim *>(src);
  const TopoRef a = DISPATCH_FN(topo_get, cs, 0), b = DISPATCH_FN(topo_get, dst, 0);
  *b.sp = *a.sp;
  *b.mg = *a.mg;
  *b.br = *a.br;
  return ORC_OK;
}

int orc_topology_stats(const OrcSim *s, uint32_t out[3]) {
  if (!s || !out) return ORC_EINVAL;
  OrcSim *cs = const_cast<OrcSim *>(s);
  const TopoRef t = DISPATCH_FN(topo_get, cs, 0);
  const u32 nm = DISPATCH_FN(nsplit_of, cs, 0);
  uint32_t ns = 0, ng = 0;
  for (auto &v : *t.sp) ns += (uint32_t)v.size();
  for (auto &v : *t.mg) ng += (uint32_t)(v.size() / nm);
  out[0] = (uint32_t)t.sp->size();
  out[1] = ns;
  out[2] = ng;
  return ORC_OK;
}

float orc_expf(float t) { return expf_cw(t); }
float orc_wavelet_f32(float x) { return wavelet_f(x); }
double orc_wavelet_f64(double x) { return wavelet_d(x); }

}  // extern "C"
