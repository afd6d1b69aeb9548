// Differentiable rollout of the SPH step and its adjoint (declared in include/sph_amd.h,
// "differentiable rollout"). No counterpart in the reference: the model is the reference's
// 3D-2 step with split, merge and marching cubes off (a frozen particle topology), plus a
// learnable force field theta (nf^3 nodes x 3 components over the box, sampled trilinearly
// at the particle and added to its SPH force in integrate); the loss after T steps is
// sum_v (rho_s(x_v) - target_v)^2 over an ns^3 grid of sample points, rho_s the particle-
// radius poly6 splat of the density grid (SPH-3D-2/sph_system.cu:1954-1990).
//
// Forward: the product kernels (hash, sort, cell ranges, reorder, density, force) and a
// field-aware integrate. Each step's pre-integrate state, density, force and cell lists
// are kept on a tape (every step, or the steps of one checkpoint segment at a time, which
// the backward pass replays from the segment's checkpoint).
//
// Backward, per step, in that step's cell order, gathers only (no atomics except the fp64
// field gradient):
//   integrate^T  per particle: adjoints of force, density, field, state (clamps frozen)
//   surface      per particle: the colour-field direction's adjoint (threshold frozen)
//   force^T      per particle over its neighbours: pressure, viscosity and colour-field
//                pair terms, with the particle as receiver and as neighbour
//   density^T    per particle over its neighbours, then the state adjoint by slot
// All particles are level 1 (large), which split and merge being off keeps true.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "sph_internal.h"

namespace {

constexpr int ABLK = 256;
inline dim3 agrid(size_t n) { return dim3((uint)((n + ABLK - 1) / ABLK)); }

// The adjoint neighbour kernels give each particle NBL consecutive lanes: lane g of the
// group walks the neighbour cells o = g, g + NBL, ... and the group adds its partial sums
// with a fixed butterfly, so the result is the same on every run (gradients are checked
// against finite differences, not bit for bit). Small rollouts would otherwise leave most
// of the GPU idle, one thread walking all 27 cells per particle.
constexpr int NBL = 16;
__device__ inline float group_sum(float x) {
#pragma unroll
  for (int m = 1; m < NBL; m <<= 1) x += __shfl_xor(x, m, 64);
  return x;
}

// neighbour offsets in the reference's order (SPH-3D-2/sph_system.cu:456-562)
__device__ constexpr int kOff[27][3] = {
    {0, 0, 0},   {-1, -1, 0},  {-1, 0, 0},  {0, -1, 0},  {1, 1, 0},    {1, 0, 0},   {0, 1, 0},
    {1, -1, 0},  {-1, 1, 0},   {0, 0, -1},  {-1, -1, -1}, {-1, 0, -1}, {0, -1, -1}, {1, 1, -1},
    {1, 0, -1},  {0, 1, -1},   {1, -1, -1}, {-1, 1, -1},  {0, 0, 1},   {-1, -1, 1}, {-1, 0, 1},
    {0, -1, 1},  {1, 1, 1},    {1, 0, 1},   {0, 1, 1},    {1, -1, 1},  {-1, 1, 1}};

__device__ inline bool nb_cell(const DevParams &P, int o, const uint c[3], uint &h) {
  uint q[3];
  for (int a = 0; a < 3; a++) {
    const int d = kOff[o][a];
    if (d == -1 && c[a] == 0) return false;
    if (d == 1 && c[a] == P.ncell[a] - 1) return false;
    q[a] = c[a] + (uint)d;
  }
  h = q[2] * P.ncell[0] * P.ncell[1] + q[1] * P.ncell[0] + q[0];
  return true;
}

__device__ inline bool cell_of3(const DevParams &P, float x, float y, float z, uint c[3]) {
  const float q[3] = {floorf(x / P.cell_size), floorf(y / P.cell_size), floorf(z / P.cell_size)};
  bool ok = true;
  for (int a = 0; a < 3; a++) {
    ok = ok && q[a] >= 0.0f && q[a] < (float)P.ncell[a];
    c[a] = ok ? (uint)q[a] : 0u;
  }
  return ok;
}

// ---- the learnable field: trilinear sample in the oracle's operation order ----
struct FW {
  int i0[3];
  float t[3];
  bool in[3];  // strictly inside the node range (d sample / d position nonzero)
};

__device__ inline FW field_weights(const DevParams &P, int nf, float x, float y, float z) {
  FW w;
  const float top = (float)(nf - 1);
  const float p3[3] = {x, y, z};
  for (int a = 0; a < 3; a++) {
    float g = p3[a] / P.world[a] * top;
    w.in[a] = g > 0.0f && g < top;
    if (!(g > 0.0f)) g = 0.0f;
    if (g > top) g = top;
    int i = (int)floorf(g);
    if (i > nf - 2) i = nf - 2;
    w.i0[a] = i;
    w.t[a] = g - (float)i;
  }
  return w;
}

__device__ inline float th(const float *theta, int nf, const FW &w, int dx, int dy, int dz, int c) {
  return theta[(((size_t)(w.i0[2] + dz) * nf + (size_t)(w.i0[1] + dy)) * nf + (size_t)(w.i0[0] + dx)) * 3 + c];
}

__device__ inline void field_eval(const float *theta, int nf, const FW &w, float f[3]) {
  const float ux = 1.0f - w.t[0], uy = 1.0f - w.t[1], uz = 1.0f - w.t[2];
  for (int c = 0; c < 3; c++) {
    const float c00 = th(theta, nf, w, 0, 0, 0, c) * ux + th(theta, nf, w, 1, 0, 0, c) * w.t[0];
    const float c10 = th(theta, nf, w, 0, 1, 0, c) * ux + th(theta, nf, w, 1, 1, 0, c) * w.t[0];
    const float c01 = th(theta, nf, w, 0, 0, 1, c) * ux + th(theta, nf, w, 1, 0, 1, c) * w.t[0];
    const float c11 = th(theta, nf, w, 0, 1, 1, c) * ux + th(theta, nf, w, 1, 1, 1, c) * w.t[0];
    const float c0 = c00 * uy + c10 * w.t[1];
    const float c1 = c01 * uy + c11 * w.t[1];
    f[c] = c0 * uz + c1 * w.t[2];
  }
}

// ---- forward: integrate with the field (SPH-3D-2/sph_system.cu:204-336, no push, no
// obstacle spheres), from the tape entry T into the next state S ----
__global__ __launch_bounds__(ABLK) void k_adj_integrate_fwd(DevParams P, const float *__restrict__ theta,
                                                            int nf, DevState T, DevState S, uint n) {
  const uint i = blockIdx.x * ABLK + threadIdx.x;
  if (i >= n) return;
  const float4 p = T.pos[i], v = T.vel[i], e = T.ev[i], a = T.acc[i];
  const float2 dp = T.dp[i];
  const float dens = dp.x;
  const float radius = __float_as_uint(p.w) == 1u ? P.radius[0] : P.radius[1];
  const float dt = P.time_step;
  float pp[3] = {p.x, p.y, p.z}, vv[3] = {v.x, v.y, v.z}, ee[3] = {e.x, e.y, e.z};
  float aa[3] = {a.x, a.y, a.z};
  if (nf > 0) {
    float f[3];
    field_eval(theta, nf, field_weights(P, nf, p.x, p.y, p.z), f);
    for (int k = 0; k < 3; k++) aa[k] = aa[k] + f[k];
  }
  for (int k = 0; k < 3; k++) vv[k] = vv[k] + aa[k] * dt / dens;
  for (int k = 0; k < 3; k++) vv[k] = vv[k] + P.gravity[k] * dt / dens;
  for (int k = 0; k < 3; k++) pp[k] = pp[k] + vv[k] * dt;
  for (int k = 0; k < 3; k++) {
    if (pp[k] >= P.world[k] - radius) {
      vv[k] = vv[k] * P.wall_damping;
      pp[k] = P.world[k] - radius;
    }
    if (pp[k] < radius) {
      vv[k] = vv[k] * P.wall_damping;
      pp[k] = radius;
    }
  }
  for (int k = 0; k < 3; k++) ee[k] = (ee[k] + vv[k]) / 2;
  S.pos[i] = make_float4(pp[0], pp[1], pp[2], p.w);
  S.vel[i] = make_float4(vv[0], vv[1], vv[2], v.w);
  S.ev[i] = make_float4(ee[0], ee[1], ee[2], e.w);
  S.acc[i] = a;
  S.dp[i] = dp;
  S.energy[i] = 0.0f;
  S.slot[i] = T.slot[i];
}

// ---- density sample and loss ----
__device__ inline void sample_point(const DevParams &P, int ns, uint v, float x[3]) {
  const uint vv[3] = {v % (uint)ns, v / (uint)ns % (uint)ns, v / ((uint)ns * (uint)ns)};
  for (int a = 0; a < 3; a++) x[a] = ((float)vv[a] + 0.5f) * (P.world[a] / (float)ns);
}

__global__ __launch_bounds__(ABLK) void k_adj_sample(DevParams P, int ns, const float4 *__restrict__ pos,
                                                     const uint *__restrict__ cstart,
                                                     const uint *__restrict__ cend,
                                                     const float *__restrict__ target,
                                                     float *__restrict__ sample,
                                                     float *__restrict__ resid,
                                                     double *__restrict__ part) {
  __shared__ double red[ABLK];
  const uint v = blockIdx.x * ABLK + threadIdx.x;
  const uint nv = (uint)ns * (uint)ns * (uint)ns;
  double sq = 0.0;
  if (v < nv) {
    float x[3];
    sample_point(P, ns, v, x);
    uint c[3];
    for (int a = 0; a < 3; a++) {
      float q = floorf(x[a] / P.cell_size);
      if (q > (float)(P.ncell[a] - 1)) q = (float)(P.ncell[a] - 1);
      c[a] = (uint)q;
    }
    const float r2s[2] = {P.radius[0] * P.radius[0], P.radius[1] * P.radius[1]};
    float acc = 0.0f;
    for (int o = 0; o < 27; o++) {
      uint h;
      if (!nb_cell(P, o, c, h)) continue;
      const uint s = cstart[h];
      if (s == SPH_EMPTY) continue;
      const uint e = cend[h];
      for (uint j = s; j < e; j++) {
        const float4 q = pos[j];
        const int L = __float_as_uint(q.w) == 1u ? 0 : 1;
        const float rx = q.x - x[0], ry = q.y - x[1], rz = q.z - x[2];
        const float r2 = rx * rx + ry * ry + rz * rz;
        if (r2 < r2s[L]) {
          const float d = r2s[L] - r2;
          acc += P.mass[L] * P.poly6_radius[L] * ((d * d) * d);
        }
      }
    }
    sample[v] = acc;
    const double r = (double)acc - (double)target[v];
    resid[v] = (float)r;
    sq = r * r;
  }
  red[threadIdx.x] = sq;
  __syncthreads();
  for (int off = ABLK / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// partial sums in block order: the loss is the same on every run
__global__ __launch_bounds__(ABLK) void k_adj_loss_sum(const double *__restrict__ part, uint nb,
                                                       double *__restrict__ loss) {
  __shared__ double red[ABLK];
  double s = 0.0;
  for (uint k = threadIdx.x; k < nb; k += ABLK) s += part[k];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = ABLK / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) *loss = red[0];
}

// dL/dp_T: p_bar = sum_v 2 r_v m poly6_r (-6) (rad^2 - d^2)^2 (p - x_v) over the sample
// points within the particle radius; velocity and ev adjoints start at zero.
__global__ __launch_bounds__(ABLK) void k_adj_loss_bwd(DevParams P, int ns, DevState F,
                                                       const float *__restrict__ resid,
                                                       float4 *__restrict__ ap, float4 *__restrict__ av,
                                                       float4 *__restrict__ ae, uint n) {
  const uint i = blockIdx.x * ABLK + threadIdx.x;
  if (i >= n) return;
  const float4 p = F.pos[i];
  const int L = __float_as_uint(p.w) == 1u ? 0 : 1;
  const float rad = P.radius[L], r2s = rad * rad;
  const float coef = 2.0f * P.mass[L] * P.poly6_radius[L] * -6.0f;
  const float p3[3] = {p.x, p.y, p.z};
  int lo[3], hi[3];
  for (int a = 0; a < 3; a++) {
    const float hs = P.world[a] / (float)ns;
    lo[a] = max(0, (int)floorf((p3[a] - rad) / hs - 0.5f) - 1);
    hi[a] = min(ns - 1, (int)ceilf((p3[a] + rad) / hs - 0.5f) + 1);
  }
  float gx = 0.0f, gy = 0.0f, gz = 0.0f;
  for (int z = lo[2]; z <= hi[2]; z++)
    for (int y = lo[1]; y <= hi[1]; y++)
      for (int x = lo[0]; x <= hi[0]; x++) {
        const uint v = ((uint)z * (uint)ns + (uint)y) * (uint)ns + (uint)x;
        float xv[3];
        sample_point(P, ns, v, xv);
        const float dx = p.x - xv[0], dy = p.y - xv[1], dz = p.z - xv[2];
        const float d2 = dx * dx + dy * dy + dz * dz;
        if (d2 < r2s) {
          const float q = r2s - d2;
          const float w = coef * resid[v] * q * q;
          gx += w * dx;
          gy += w * dy;
          gz += w * dz;
        }
      }
  const uint s = F.slot[i];
  ap[s] = make_float4(gx, gy, gz, 0.0f);
  av[s] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  ae[s] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// ---- backward of one step ----
struct Bufs {
  float4 *Fbar, *gbar, *pbar, *vbar, *ebar;
  float *rbar;
};

// integrate^T: from the adjoints of the step's output state (slot order) to the adjoints
// of the force, density, field and input state. Wall clamps are recomputed, hence frozen.
// The field gradient of a block (its particles share a few field cells) is summed in LDS
// first, then added to the fp64 global sum once per nonzero node component.
__device__ inline void adj_integrate_particle(const DevParams &P, const float *__restrict__ theta, int nf,
                                              const DevState &T, const float4 *__restrict__ ap,
                                              const float4 *__restrict__ av,
                                              const float4 *__restrict__ ae, const Bufs &B,
                                              double *__restrict__ dtheta, float *th_lds, uint i) {
  const float4 p = T.pos[i], v = T.vel[i], a = T.acc[i];
  const float dens = T.dp[i].x;
  const float radius = __float_as_uint(p.w) == 1u ? P.radius[0] : P.radius[1];
  const float dt = P.time_step;
  const float pp0[3] = {p.x, p.y, p.z};
  float aa[3] = {a.x, a.y, a.z};
  FW w;
  if (nf > 0) {
    float f[3];
    w = field_weights(P, nf, p.x, p.y, p.z);
    field_eval(theta, nf, w, f);
    for (int k = 0; k < 3; k++) aa[k] = aa[k] + f[k];
  }
  // forward to the wall tests, as k_adj_integrate_fwd
  float vv[3] = {v.x, v.y, v.z}, pp[3];
  bool clamped[3];
  for (int k = 0; k < 3; k++) vv[k] = vv[k] + aa[k] * dt / dens;
  for (int k = 0; k < 3; k++) vv[k] = vv[k] + P.gravity[k] * dt / dens;
  for (int k = 0; k < 3; k++) pp[k] = pp0[k] + vv[k] * dt;
  for (int k = 0; k < 3; k++) {
    clamped[k] = false;
    if (pp[k] >= P.world[k] - radius) {
      clamped[k] = true;
      pp[k] = P.world[k] - radius;
    }
    if (pp[k] < radius) clamped[k] = true;
  }
  const uint s = T.slot[i];
  const float4 Pb = ap[s], Vb = av[s], Eb = ae[s];
  const float pb[3] = {Pb.x, Pb.y, Pb.z}, vb[3] = {Vb.x, Vb.y, Vb.z}, eb[3] = {Eb.x, Eb.y, Eb.z};
  float v3b[3], v2b[3], pbar[3], ebar[3];
  for (int k = 0; k < 3; k++) {
    v3b[k] = vb[k] + 0.5f * eb[k];  // ev' = (ev + v')/2
    ebar[k] = 0.5f * eb[k];
    if (clamped[k]) {
      v2b[k] = P.wall_damping * v3b[k];
      pbar[k] = 0.0f;
    } else {
      v2b[k] = v3b[k];
      pbar[k] = pb[k];
    }
    v2b[k] += pbar[k] * dt;  // p' = p + v2 dt
  }
  float rb = 0.0f;
  const float idd = dt / (dens * dens);
  for (int k = 0; k < 3; k++) rb -= v2b[k] * (P.gravity[k] + aa[k]) * idd;
  float ab[3];
  for (int k = 0; k < 3; k++) ab[k] = v2b[k] * dt / dens;
  if (nf > 0) {
    const float ux = 1.0f - w.t[0], uy = 1.0f - w.t[1], uz = 1.0f - w.t[2];
    const float wx[2] = {ux, w.t[0]}, wy[2] = {uy, w.t[1]}, wz[2] = {uz, w.t[2]};
    for (int dz = 0; dz < 2; dz++)
      for (int dy = 0; dy < 2; dy++)
        for (int dx = 0; dx < 2; dx++) {
          const size_t node = ((size_t)(w.i0[2] + dz) * nf + (size_t)(w.i0[1] + dy)) * nf + (size_t)(w.i0[0] + dx);
          const float wt = wx[dx] * wy[dy] * wz[dz];
          for (int c = 0; c < 3; c++) {
            if (th_lds) atomicAdd(&th_lds[node * 3 + c], wt * ab[c]);
            else atomicAdd(&dtheta[node * 3 + c], (double)(wt * ab[c]));
          }
        }
    // d field / d position through t = position / world * (nf - 1)
    for (int ax = 0; ax < 3; ax++) {
      if (!w.in[ax]) continue;
      const float sc = (float)(nf - 1) / P.world[ax];
      for (int c = 0; c < 3; c++) {
        float dfd = 0.0f;
        for (int dz = 0; dz < 2; dz++)
          for (int dy = 0; dy < 2; dy++)
            for (int dx = 0; dx < 2; dx++) {
              const int dd[3] = {dx, dy, dz};
              float wt = 1.0f;
              for (int b = 0; b < 3; b++) {
                const float wb = dd[b] ? w.t[b] : 1.0f - w.t[b];
                wt *= (b == ax) ? (dd[b] ? 1.0f : -1.0f) : wb;
              }
              dfd += wt * th(theta, nf, w, dx, dy, dz, c);
            }
        pbar[ax] += ab[c] * dfd * sc;
      }
    }
  }
  B.Fbar[i] = make_float4(ab[0], ab[1], ab[2], 0.0f);
  B.rbar[i] = rb;
  B.pbar[i] = make_float4(pbar[0], pbar[1], pbar[2], 0.0f);
  B.vbar[i] = make_float4(v2b[0], v2b[1], v2b[2], 0.0f);  // v2 = v + (a + g) dt / dens
  B.ebar[i] = make_float4(ebar[0], ebar[1], ebar[2], 0.0f);
}

constexpr int LDS_THETA = 4096;  // node components summed in LDS (nf <= 11)
// 64-thread blocks: at the rollout's 32k particles, 512 blocks spread over every CU
// (256-thread blocks left half of them idle): backward 34.9 -> 34.2 ms
constexpr int IBLK = 64;
__global__ __launch_bounds__(IBLK) void k_adj_integrate_bwd(DevParams P, const float *__restrict__ theta,
                                                            int nf, DevState T,
                                                            const float4 *__restrict__ ap,
                                                            const float4 *__restrict__ av,
                                                            const float4 *__restrict__ ae, Bufs B,
                                                            double *__restrict__ dtheta, uint n) {
  __shared__ float th_lds[LDS_THETA];
  const int nth = nf * nf * nf * 3;
  const bool use_lds = nf > 0 && nth <= LDS_THETA;
  if (use_lds) {
    for (int k = threadIdx.x; k < nth; k += IBLK) th_lds[k] = 0.0f;
    __syncthreads();
  }
  const uint i = blockIdx.x * IBLK + threadIdx.x;
  if (i < n) adj_integrate_particle(P, theta, nf, T, ap, av, ae, B, dtheta, use_lds ? th_lds : nullptr, i);
  if (use_lds) {
    __syncthreads();
    for (int k = threadIdx.x; k < nth; k += IBLK) {
      const float v = th_lds[k];
      if (v != 0.0f) atomicAdd(&dtheta[k], (double)v);
    }
  }
}


// Reciprocal for the backward kernels: the hardware estimate (1 ulp). The adjoint is checked
// against finite differences of the float64 oracle, not bit for bit, so its pair terms
// multiply by reciprocals instead of dividing.
__device__ inline float adj_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// single-level pair constants (all particles level 1)
struct PairK {
  float h, k2, m, S, V, mu, G, Lc, poly6, cs, sigma;
};
__device__ inline PairK pair_k(const DevParams &P) {
  PairK K;
  K.h = P.kernel[0];
  K.k2 = P.k2[0];
  K.m = P.mass[0];
  K.S = P.spiky[0] / P.k6[0];  // pressure kernel divided by h^6 twice (3D-2 :818)
  K.V = P.visco_k6[0];         // viscosity kernel divided by h^6 twice (3D-2 :819)
  K.mu = P.viscosity;
  K.G = -P.grad[0];
  K.Lc = P.lapc[0];
  K.poly6 = P.poly6[0];
  K.cs = P.surf_coef;
  K.sigma = P.surface_tension;
  return K;
}

// Colour-field direction adjoint: F += cs lap g/|g| when |g| > sigma (the forward's |g|
// is acc.w, so the same particles are above the threshold); g and lap recomputed here.
__global__ __launch_bounds__(ABLK) void k_adj_surface(DevParams P, DevState T,
                                                      const uint *__restrict__ cstart,
                                                      const uint *__restrict__ cend, Bufs B, uint n) {
  const uint i = (xcd_block(blockIdx.x, gridDim.x) * ABLK + threadIdx.x) / NBL;
  const int g0 = (int)(threadIdx.x % NBL);
  if (i >= n) return;  // whole groups
  const PairK K = pair_k(P);
  const float4 pi = T.pos[i];
  if (!(T.acc[i].w > K.sigma)) {  // the same for the whole group
    if (g0 == 0) B.gbar[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return;
  }
  uint c[3];
  float gx = 0.0f, gy = 0.0f, gz = 0.0f, lap = 0.0f;
  if (cell_of3(P, pi.x, pi.y, pi.z, c)) {
    for (int o = g0; o < 27; o += NBL) {
      uint h;
      if (!nb_cell(P, o, c, h)) continue;
      const uint s = cstart[h];
      if (s == SPH_EMPTY) continue;
      const uint e = cend[h];
      for (uint j = s; j < e; j++) {
        const float4 pj = T.pos[j];
        const float ux = pi.x - pj.x, uy = pi.y - pj.y, uz = pi.z - pj.z;
        const float r2 = ux * ux + uy * uy + uz * uz;
        if (!(r2 < K.k2) || r2 < 0.0001f) continue;  // also the particle itself (r2 = 0)
        const float d = K.k2 - r2;
        const float ij = adj_rcp(T.dp[j].x);  // 1/rho_j
        const float kg = K.G * K.m * ij * (d * d);
        gx += kg * ux;
        gy += kg * uy;
        gz += kg * uz;
        lap += K.Lc * K.m * ij * d * r2;
      }
    }
  }
  gx = group_sum(gx);
  gy = group_sum(gy);
  gz = group_sum(gz);
  lap = group_sum(lap);
  if (g0 != 0) return;
  const float nrm = sqrtf(gx * gx + gy * gy + gz * gz);
  const float4 Fb = B.Fbar[i];
  const float nx = gx / nrm, ny = gy / nrm, nz = gz / nrm;
  const float lam = K.cs * (Fb.x * nx + Fb.y * ny + Fb.z * nz);
  const float mx = K.cs * lap * Fb.x, my = K.cs * lap * Fb.y, mz = K.cs * lap * Fb.z;
  const float nm = nx * mx + ny * my + nz * mz;
  B.gbar[i] = make_float4((mx - nm * nx) / nrm, (my - nm * ny) / nrm, (mz - nm * nz) / nrm, lam);
}

// force^T: for particle k, every force pair (r2 < h^2, r2 >= 1e-4) with k as receiver and
// as neighbour; accumulates the position, ev and density adjoints of k.
__global__ __launch_bounds__(ABLK) void k_adj_force_bwd(DevParams P, DevState T,
                                                        const uint *__restrict__ cstart,
                                                        const uint *__restrict__ cend, Bufs B, uint n) {
  const uint k = (xcd_block(blockIdx.x, gridDim.x) * ABLK + threadIdx.x) / NBL;
  const int g0 = (int)(threadIdx.x % NBL);
  if (k >= n) return;  // whole groups
  const PairK K = pair_k(P);
  const float4 pk = T.pos[k], ek = T.ev[k];
  const float2 dk = T.dp[k];
  const float4 Fk = B.Fbar[k], gk = B.gbar[k];
  const float rk = dk.x, Pk = dk.y;
  const float m2 = K.m * K.m;
  const float ik = adj_rcp(rk);  // 1/rho_k
  const float pdk = Pk * ik * ik;
  const float Kk = K.G * K.m * ik, Lk = K.Lc * K.m * ik;
  const float pbk = m2 * ik * ik;                  // d(P_k/rho_k^2 m^2)/dP_k
  const float rfk = m2 * (-2.0f * Pk * ik * ik * ik);  // d(P_k/rho_k^2 m^2)/drho_k
  const float wfk = m2 * K.mu * K.V * ik;
  float px = 0.0f, py = 0.0f, pz = 0.0f, ex = 0.0f, ey = 0.0f, ez = 0.0f, rb = 0.0f, Pb = 0.0f;
  uint c[3];
  if (cell_of3(P, pk.x, pk.y, pk.z, c)) {
    for (int o = g0; o < 27; o += NBL) {
      uint h;
      if (!nb_cell(P, o, c, h)) continue;
      const uint s = cstart[h];
      if (s == SPH_EMPTY) continue;
      const uint e = cend[h];
      for (uint j = s; j < e; j++) {
        const float4 pj = T.pos[j];
        const float ux = pk.x - pj.x, uy = pk.y - pj.y, uz = pk.z - pj.z;
        const float r2 = ux * ux + uy * uy + uz * uz;
        if (!(r2 < K.k2) || r2 < 0.0001f) continue;  // also the particle itself (r2 = 0)
        const float r = sqrtf(r2), hr = K.h - r, d = K.k2 - r2;
        const float ir = adj_rcp(r);
        const float2 dj = T.dp[j];
        const float ij = adj_rcp(dj.x), Pj = dj.y;  // 1/rho_j, P_j
        const float4 Fj = B.Fbar[j], gj = B.gbar[j], ej = T.ev[j];
        const float Dx = Fk.x - Fj.x, Dy = Fk.y - Fj.y, Dz = Fk.z - Fj.z;
        const float Du = Dx * ux + Dy * uy + Dz * uz;
        // pressure: loss term -(D.u) c phi(r), c = m^2 (P_k/rho_k^2 + P_j/rho_j^2)
        const float cc = m2 * (pdk + Pj * ij * ij);
        const float phi = K.S * hr * hr * ir;
        const float dphi = -K.S * hr * (K.h + r) * (ir * ir);
        const float tq = -cc * Du * dphi * ir;
        px += -cc * phi * Dx + tq * ux;
        py += -cc * phi * Dy + tq * uy;
        pz += -cc * phi * Dz + tq * uz;
        const float cb = -Du * phi;
        Pb += cb * pbk;
        rb += cb * rfk;
        // viscosity: loss term (D.(e_j - e_k)) w, w = m^2 mu V (h - r)/(rho_k rho_j)
        const float wf = wfk * ij;
        const float w = wf * hr;
        const float dex = ej.x - ek.x, dey = ej.y - ek.y, dez = ej.z - ek.z;
        const float Dde = Dx * dex + Dy * dey + Dz * dez;
        ex -= Dx * w;
        ey -= Dy * w;
        ez -= Dz * w;
        const float tv = -Dde * wf * ir;
        px += tv * ux;
        py += tv * uy;
        pz += tv * uz;
        rb -= Dde * w * ik;
        // colour field, k as receiver: gbar_k . (K_j d^2 u) + lam_k L_j d r^2
        const float Kj = K.G * K.m * ij, Lj = K.Lc * K.m * ij;
        const float gku = gk.x * ux + gk.y * uy + gk.z * uz;
        const float a1 = Kj * d * d, b1 = -4.0f * Kj * d * gku + 2.0f * gk.w * Lj * (d - r2);
        px += a1 * gk.x + b1 * ux;
        py += a1 * gk.y + b1 * uy;
        pz += a1 * gk.z + b1 * uz;
        // colour field, k as neighbour of receiver j (u_jk = -u)
        const float gju = gj.x * ux + gj.y * uy + gj.z * uz;
        const float a2 = -Kk * d * d, b2 = 4.0f * Kk * d * gju + 2.0f * gj.w * Lk * (d - r2);
        px += a2 * gj.x + b2 * ux;
        py += a2 * gj.y + b2 * uy;
        pz += a2 * gj.z + b2 * uz;
        rb -= (-Kk * d * d * gju + gj.w * Lk * d * r2) * ik;
      }
    }
  }
  px = group_sum(px);
  py = group_sum(py);
  pz = group_sum(pz);
  ex = group_sum(ex);
  ey = group_sum(ey);
  ez = group_sum(ez);
  rb = group_sum(rb);
  Pb = group_sum(Pb);
  if (g0 != 0) return;
  // pressure of k: P = k_gas ((rho/rho0)^3 - 1)
  const float q = rk / P.rest_density;
  rb += Pb * P.gas_constant * 3.0f * q * q / P.rest_density;
  const float4 pb = B.pbar[k], eb = B.ebar[k];
  B.pbar[k] = make_float4(pb.x + px, pb.y + py, pb.z + pz, 0.0f);
  B.ebar[k] = make_float4(eb.x + ex, eb.y + ey, eb.z + ez, 0.0f);
  B.rbar[k] += rb;
}

// density^T, then the adjoint of the step's input state, by slot.
__global__ __launch_bounds__(ABLK) void k_adj_density_bwd(DevParams P, DevState T,
                                                          const uint *__restrict__ cstart,
                                                          const uint *__restrict__ cend, Bufs B,
                                                          float4 *__restrict__ ap, float4 *__restrict__ av,
                                                          float4 *__restrict__ ae, uint n) {
  const uint k = (xcd_block(blockIdx.x, gridDim.x) * ABLK + threadIdx.x) / NBL;
  const int g0 = (int)(threadIdx.x % NBL);
  if (k >= n) return;  // whole groups
  const PairK K = pair_k(P);
  const float4 pk = T.pos[k];
  const float rbk = B.rbar[k];
  const float cf = -6.0f * K.m * K.poly6;
  float px = 0.0f, py = 0.0f, pz = 0.0f;
  uint c[3];
  if (cell_of3(P, pk.x, pk.y, pk.z, c)) {
    for (int o = g0; o < 27; o += NBL) {
      uint h;
      if (!nb_cell(P, o, c, h)) continue;
      const uint s = cstart[h];
      if (s == SPH_EMPTY) continue;
      const uint e = cend[h];
      for (uint j = s; j < e; j++) {
        const float4 pj = T.pos[j];
        const float ux = pk.x - pj.x, uy = pk.y - pj.y, uz = pk.z - pj.z;
        const float r2 = ux * ux + uy * uy + uz * uz;
        if (r2 < P.dens_eps || !(r2 < K.k2)) continue;  // also the particle itself (r2 = 0)
        const float d = K.k2 - r2;
        const float t = (rbk + B.rbar[j]) * cf * d * d;
        px += t * ux;
        py += t * uy;
        pz += t * uz;
      }
    }
  }
  px = group_sum(px);
  py = group_sum(py);
  pz = group_sum(pz);
  if (g0 != 0) return;
  const float4 pb = B.pbar[k];
  const uint s = T.slot[k];
  ap[s] = make_float4(pb.x + px, pb.y + py, pb.z + pz, 0.0f);
  av[s] = B.vbar[k];
  ae[s] = B.ebar[k];
}

__global__ __launch_bounds__(ABLK) void k_fill_f64(double *p, size_t n) {
  const size_t i = (size_t)blockIdx.x * ABLK + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

__global__ __launch_bounds__(ABLK) void k_f64_to_f32(const double *__restrict__ a, float *__restrict__ b,
                                                     size_t n) {
  const size_t i = (size_t)blockIdx.x * ABLK + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

}  // namespace

// ---------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------
struct SphAdj {
  DevParams P;
  unsigned char raw[sizeof(SphSysParam3D)];
  uint n = 0;
  int nf = 0, ns = 0;
  uint ckpt = 0;
  uint cells = 0, nvox = 0, nvb = 0;
  int kbits = 1;
  hipStream_t st = nullptr;
  std::vector<void *> allocs;
  DevState cur{}, scratch{}, state0{};
  Aux aux{};
  uint *keys = nullptr, *vals = nullptr, *skeys = nullptr, *svals = nullptr;
  SortScratch sc{};
  uint *s2p_hash = nullptr, *s2p_cell = nullptr, *cs_scr = nullptr, *ce_scr = nullptr;
  uint *scan_tmp = nullptr;
  Counters *cnt = nullptr;
  Counters *cnt_host = nullptr;
  float *theta = nullptr, *target = nullptr, *sample = nullptr, *resid = nullptr;
  double *dtheta = nullptr, *part = nullptr, *loss_dev = nullptr, *loss_host = nullptr;
  float *dtheta_f = nullptr;
  float4 *ap = nullptr, *av = nullptr, *ae = nullptr;
  Bufs B{};
  unsigned char *aos_dev = nullptr;
  struct Tape {
    DevState T;
    uint *cs, *ce;
  };
  std::vector<Tape> tape;
  std::vector<DevState> ckpts;  // pos, vel, ev, slot at the start of each segment
  uint steps = 0;               // of the last forward
  bool have_state = false, have_target = false, forward_done = false;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float fwd_ms = 0.0f, bwd_ms = 0.0f;

  ~SphAdj() {
    if (st) (void)hipStreamSynchronize(st);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    for (void *p : allocs) (void)hipFree(p);
    if (cnt_host) (void)hipHostFree(cnt_host);
    if (loss_host) (void)hipHostFree(loss_host);
    if (st) (void)hipStreamDestroy(st);
  }
  template <class T> bool alloc(T *&p, size_t count) {
    void *q = nullptr;
    if (count == 0) count = 1;
    if (hipMalloc(&q, count * sizeof(T)) != hipSuccess) return false;
    allocs.push_back(q);
    p = (T *)q;
    return true;
  }
  bool alloc_state(DevState &s) {
    return alloc(s.pos, n) && alloc(s.vel, n) && alloc(s.ev, n) && alloc(s.acc, n) &&
           alloc(s.dp, n) && alloc(s.energy, n) && alloc(s.slot, n);
  }
  bool grow_tape(size_t len) {
    while (tape.size() < len) {
      Tape t{};
      if (!alloc_state(t.T) || !alloc(t.cs, cells) || !alloc(t.ce, cells)) return false;
      tape.push_back(t);
    }
    return true;
  }
  bool grow_ckpts(size_t len) {
    while (ckpts.size() < len) {
      DevState c{};
      if (!alloc(c.pos, n) || !alloc(c.vel, n) || !alloc(c.ev, n) || !alloc(c.slot, n)) return false;
      ckpts.push_back(c);
    }
    return true;
  }
  hipError_t copy_state(DevState dst, DevState src, bool all) {
    hipError_t e = hipSuccess;
    auto cp = [&](void *d, const void *s, size_t b) {
      if (e == hipSuccess && d && s) e = hipMemcpyAsync(d, s, b, hipMemcpyDeviceToDevice, st);
    };
    cp(dst.pos, src.pos, 16 * (size_t)n);
    cp(dst.vel, src.vel, 16 * (size_t)n);
    cp(dst.ev, src.ev, 16 * (size_t)n);
    cp(dst.slot, src.slot, 4 * (size_t)n);
    if (all) {
      cp(dst.acc, src.acc, 16 * (size_t)n);
      cp(dst.dp, src.dp, 8 * (size_t)n);
      cp(dst.energy, src.energy, 4 * (size_t)n);
    }
    return e;
  }

  // one forward step of state S: cell order into T (lists cs/ce), next state back into S
  void fwd_step(DevState S, DevState T, uint *cs, uint *ce) {
    launch_hash(SPH_V3D2, P, S, nullptr, n, keys, s2p_hash, cnt, st);
    sph_radix_sort(keys, vals, n, kbits, sc, &skeys, &svals, st);
    sph_fill_cells(cs, ce, cells, st);
    launch_cell_bounds(skeys, n, cells, cs, ce, st);
    launch_reorder(S, T, svals, s2p_hash, n, s2p_cell, st);
    launch_density(SPH_V3D2, P, T, aux, cs, ce, n, st);
    launch_force(SPH_V3D2, P, T, aux, cs, ce, n, st);
    k_adj_integrate_fwd<<<agrid(n), ABLK, 0, st>>>(P, theta, nf, T, S, n);
  }
  void bwd_step(const Tape &t) {
    k_adj_integrate_bwd<<<dim3((uint)((n + IBLK - 1) / IBLK)), IBLK, 0, st>>>(P, theta, nf, t.T, ap, av, ae, B, dtheta, n);
    k_adj_surface<<<agrid((size_t)NBL * n), ABLK, 0, st>>>(P, t.T, t.cs, t.ce, B, n);
    k_adj_force_bwd<<<agrid((size_t)NBL * n), ABLK, 0, st>>>(P, t.T, t.cs, t.ce, B, n);
    k_adj_density_bwd<<<agrid((size_t)NBL * n), ABLK, 0, st>>>(P, t.T, t.cs, t.ce, B, ap, av, ae, n);
  }
  // final state -> cell order in scratch (lists cs_scr/ce_scr), sample, loss
  void final_sample() {
    launch_hash(SPH_V3D2, P, cur, nullptr, n, keys, s2p_hash, cnt, st);
    sph_radix_sort(keys, vals, n, kbits, sc, &skeys, &svals, st);
    sph_fill_cells(cs_scr, ce_scr, cells, st);
    launch_cell_bounds(skeys, n, cells, cs_scr, ce_scr, st);
    launch_reorder(cur, scratch, svals, s2p_hash, n, s2p_cell, st);
    k_adj_sample<<<dim3(nvb), ABLK, 0, st>>>(P, ns, scratch.pos, cs_scr, ce_scr, target, sample, resid, part);
    k_adj_loss_sum<<<1, ABLK, 0, st>>>(part, nvb, loss_dev);
  }
  int sync() {
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(cnt_host, cnt, sizeof(Counters), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e == hipSuccess ? SPH_OK : SPH_EHIP;
  }
};

#define CHECK_ADJ(a)            \
  do {                          \
    if (!(a)) return SPH_EINVAL; \
  } while (0)

extern "C" {

int sph_adj_create(const void *sysparam, uint32_t n, uint32_t field_nodes, uint32_t sample_n,
                   uint32_t checkpoint_every, SphAdj **out) {
  if (!out) return SPH_EINVAL;
  *out = nullptr;
  if (!sysparam || n == 0 || n > (1u << 27) || field_nodes == 1 || field_nodes > 64 ||
      sample_n == 0 || sample_n > 1024)
    return SPH_EINVAL;
  const SphSysParam3D &sp = *(const SphSysParam3D *)sysparam;
  if (sp.flag0 || sp.flag1) return SPH_EINVAL;  // no push, no obstacle spheres
  SphAdj *a = new (std::nothrow) SphAdj();
  if (!a) return SPH_ECAP;
  std::memcpy(a->raw, sysparam, sizeof(SphSysParam3D));
  a->P = sphhost::make_dev_params(SPH_V3D2, a->raw, n);
  a->n = n;
  a->nf = (int)field_nodes;
  a->ns = (int)sample_n;
  a->ckpt = checkpoint_every;
  a->cells = a->P.total_cell;
  a->nvox = sample_n * sample_n * sample_n;
  a->nvb = (a->nvox + ABLK - 1) / ABLK;
  if (a->cells == 0) {
    delete a;
    return SPH_EINVAL;
  }
  a->kbits = sphhost::key_bits(a->cells);
  if (hipStreamCreateWithFlags(&a->st, hipStreamNonBlocking) != hipSuccess) {
    delete a;
    return SPH_EHIP;
  }
  const size_t nth = (size_t)std::max(a->nf, 2) * std::max(a->nf, 2) * std::max(a->nf, 2) * 3;
  const size_t scan_n = std::max<size_t>(n, sph_sort_hist_elems(n));
  bool ok = a->alloc_state(a->cur) && a->alloc_state(a->scratch) && a->alloc_state(a->state0) &&
            a->alloc(a->keys, n) && a->alloc(a->vals, n) && a->alloc(a->sc.keys_alt, n) &&
            a->alloc(a->sc.vals_alt, n) && a->alloc(a->sc.tile_hist, sph_sort_hist_elems(n)) &&
            a->alloc(a->scan_tmp, sph_scan_tmp_elems(scan_n)) && a->alloc(a->s2p_hash, n) &&
            a->alloc(a->s2p_cell, n) && a->alloc(a->cs_scr, a->cells) && a->alloc(a->ce_scr, a->cells) &&
            a->alloc(a->cnt, 1) && a->alloc(a->aux.frec, 3 * (size_t)n) &&
            a->alloc(a->theta, nth) && a->alloc(a->dtheta, nth) && a->alloc(a->dtheta_f, nth) && a->alloc(a->target, a->nvox) &&
            a->alloc(a->sample, a->nvox) && a->alloc(a->resid, a->nvox) && a->alloc(a->part, a->nvb) &&
            a->alloc(a->loss_dev, 1) && a->alloc(a->ap, n) && a->alloc(a->av, n) && a->alloc(a->ae, n) &&
            a->alloc(a->B.Fbar, n) && a->alloc(a->B.gbar, n) && a->alloc(a->B.pbar, n) &&
            a->alloc(a->B.vbar, n) && a->alloc(a->B.ebar, n) && a->alloc(a->B.rbar, n) &&
            a->alloc(a->aos_dev, (size_t)n * sizeof(SphParticle3D));
  a->sc.scan_tmp = a->scan_tmp;
  a->sc.max_tiles = sph_sort_tiles(n);
  if (ok && hipHostMalloc((void **)&a->cnt_host, sizeof(Counters), hipHostMallocDefault) != hipSuccess) ok = false;
  if (ok && hipHostMalloc((void **)&a->loss_host, sizeof(double), hipHostMallocDefault) != hipSuccess) ok = false;
  if (ok && (hipEventCreate(&a->e0) != hipSuccess || hipEventCreate(&a->e1) != hipSuccess)) ok = false;
  if (!ok) {
    delete a;
    return SPH_ECAP;
  }
  *out = a;
  return SPH_OK;
}

int sph_adj_destroy(SphAdj *a) {
  CHECK_ADJ(a);
  delete a;
  return SPH_OK;
}

int sph_adj_upload(SphAdj *a, const void *aos, uint32_t n) {
  CHECK_ADJ(a);
  if (!aos || n != a->n) return SPH_EINVAL;
  const SphParticle3D *q = (const SphParticle3D *)aos;
  for (uint32_t i = 0; i < n; i++)
    if (q[i].level != 1) return SPH_EINVAL;
  if (hipMemcpyAsync(a->aos_dev, aos, (size_t)n * sizeof(SphParticle3D), hipMemcpyHostToDevice,
                     a->st) != hipSuccess)
    return SPH_EHIP;
  launch_upload_scatter(SPH_V3D2, a->aos_dev, n, a->state0, a->st);
  a->have_state = true;
  a->forward_done = false;
  return a->sync();
}

int sph_adj_set_target(SphAdj *a, const float *target_dev) {
  CHECK_ADJ(a);
  if (!target_dev) return SPH_EINVAL;
  if (hipMemcpyAsync(a->target, target_dev, sizeof(float) * a->nvox, hipMemcpyDefault, a->st) != hipSuccess)
    return SPH_EHIP;
  a->have_target = true;
  return a->sync();
}

int sph_adj_forward(SphAdj *a, const float *theta_dev, uint32_t steps, float *sample_dev, double *loss) {
  CHECK_ADJ(a);
  if (!a->have_state || (a->nf > 0 && !theta_dev)) return SPH_EINVAL;
  hipStream_t st = a->st;
  if (!a->have_target) {
    sph_fill_u32((uint *)a->target, 0u, a->nvox, st);  // 0.0f
    a->have_target = true;
  }
  const size_t nth = (size_t)a->nf * a->nf * a->nf * 3;
  if (a->nf > 0 && hipMemcpyAsync(a->theta, theta_dev, sizeof(float) * nth, hipMemcpyDefault, st) != hipSuccess)
    return SPH_EHIP;
  const bool record = a->ckpt == 0;
  if (record ? !a->grow_tape(steps) : !a->grow_ckpts((steps + a->ckpt - 1) / a->ckpt)) return SPH_ECAP;
  (void)hipEventRecord(a->e0, st);
  if (a->copy_state(a->cur, a->state0, true) != hipSuccess) return SPH_EHIP;
  Counters zero;
  std::memset(&zero, 0, sizeof(zero));
  *a->cnt_host = zero;
  if (hipMemcpyAsync(a->cnt, a->cnt_host, sizeof(Counters), hipMemcpyHostToDevice, st) != hipSuccess)
    return SPH_EHIP;
  for (uint32_t t = 0; t < steps; t++) {
    if (!record && t % a->ckpt == 0 && a->copy_state(a->ckpts[t / a->ckpt], a->cur, false) != hipSuccess)
      return SPH_EHIP;
    if (record) a->fwd_step(a->cur, a->tape[t].T, a->tape[t].cs, a->tape[t].ce);
    else a->fwd_step(a->cur, a->scratch, a->cs_scr, a->ce_scr);
  }
  a->final_sample();
  if (sample_dev &&
      hipMemcpyAsync(sample_dev, a->sample, sizeof(float) * a->nvox, hipMemcpyDefault, st) != hipSuccess)
    return SPH_EHIP;
  if (hipMemcpyAsync(a->loss_host, a->loss_dev, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess)
    return SPH_EHIP;
  (void)hipEventRecord(a->e1, st);
  int rc = a->sync();
  if (rc) return rc;
  (void)hipEventElapsedTime(&a->fwd_ms, a->e0, a->e1);
  a->steps = steps;
  a->forward_done = true;
  if (loss) *loss = *a->loss_host;
  return a->cnt_host->err ? SPH_EDOMAIN : SPH_OK;
}

int sph_adj_backward(SphAdj *a, float *dtheta_dev) {
  CHECK_ADJ(a);
  if (!a->forward_done) return SPH_EINVAL;
  hipStream_t st = a->st;
  const uint n = a->n;
  (void)hipEventRecord(a->e0, st);
  const size_t nth = a->nf > 0 ? (size_t)a->nf * a->nf * a->nf * 3 : 1;
  k_fill_f64<<<agrid(nth), ABLK, 0, st>>>(a->dtheta, nth);
  k_adj_loss_bwd<<<agrid(n), ABLK, 0, st>>>(a->P, a->ns, a->scratch, a->resid, a->ap, a->av, a->ae, n);
  if (a->ckpt == 0) {
    for (uint32_t t = a->steps; t-- > 0;) a->bwd_step(a->tape[t]);
  } else {
    if (!a->grow_tape(std::min(a->ckpt, a->steps))) return SPH_ECAP;
    const uint nseg = (a->steps + a->ckpt - 1) / a->ckpt;
    for (uint seg = nseg; seg-- > 0;) {
      const uint t0 = seg * a->ckpt, len = std::min(a->ckpt, a->steps - t0);
      // replayed in scratch (free once the loss adjoint has read the final state), so the
      // state of the forward stays in cur
      if (a->copy_state(a->scratch, a->ckpts[seg], false) != hipSuccess) return SPH_EHIP;
      for (uint j = 0; j < len; j++) a->fwd_step(a->scratch, a->tape[j].T, a->tape[j].cs, a->tape[j].ce);
      for (uint j = len; j-- > 0;) a->bwd_step(a->tape[j]);
    }
  }
  if (a->nf > 0 && dtheta_dev) {
    k_f64_to_f32<<<agrid(nth), ABLK, 0, st>>>(a->dtheta, a->dtheta_f, nth);
    if (hipMemcpyAsync(dtheta_dev, a->dtheta_f, sizeof(float) * nth, hipMemcpyDefault, st) != hipSuccess)
      return SPH_EHIP;
  }
  (void)hipEventRecord(a->e1, st);
  int rc = a->sync();
  if (rc) return rc;
  (void)hipEventElapsedTime(&a->bwd_ms, a->e0, a->e1);
  return SPH_OK;
}

int sph_adj_download(SphAdj *a, void *aos, uint32_t n) {
  CHECK_ADJ(a);
  if (!aos || n != a->n || !a->forward_done) return SPH_EINVAL;
  launch_slot2phys(a->cur, n, a->s2p_cell, a->st);
  launch_download_gather(SPH_V3D2, a->cur, a->s2p_cell, n, a->aos_dev, a->st);
  if (hipMemcpyAsync(aos, a->aos_dev, (size_t)n * sizeof(SphParticle3D), hipMemcpyDeviceToHost, a->st) != hipSuccess)
    return SPH_EHIP;
  return a->sync();
}

int sph_adj_state_grad(SphAdj *a, float *dpos, float *dvel, float *dev) {
  CHECK_ADJ(a);
  if (!a->forward_done) return SPH_EINVAL;
  std::vector<float> h(4 * (size_t)a->n);
  float *outs[3] = {dpos, dvel, dev};
  float4 *src[3] = {a->ap, a->av, a->ae};
  for (int k = 0; k < 3; k++) {
    if (!outs[k]) continue;
    if (hipMemcpyAsync(h.data(), src[k], 16 * (size_t)a->n, hipMemcpyDeviceToHost, a->st) != hipSuccess ||
        hipStreamSynchronize(a->st) != hipSuccess)
      return SPH_EHIP;
    for (uint i = 0; i < a->n; i++)
      for (int c = 0; c < 3; c++) outs[k][3 * i + c] = h[4 * i + c];
  }
  return SPH_OK;
}

int sph_adj_timing(SphAdj *a, double *fwd_ms, double *bwd_ms) {
  CHECK_ADJ(a);
  if (fwd_ms) *fwd_ms = a->fwd_ms;
  if (bwd_ms) *bwd_ms = a->bwd_ms;
  return SPH_OK;
}

}  // extern "C"
