// Differentiable rollout of the adaptive SPH step and its adjoint (declared in
// include/sph_amd.h, "differentiable rollout"). No counterpart in the reference: the model
// is the reference's 3D-2 step (SPH-3D-2/sph_system.cpp:302-404) -- hash, sort, cell
// ranges, density, force, integrate, and with the flags energy, split (1 -> 2) and merge --
// plus a learnable force field theta (nf^3 nodes x 3 components over the box, sampled
// trilinearly at the particle and added to its SPH force in integrate); the loss after T
// steps is sum_v (rho_s(x_v) - target_v)^2 over an ns^3 grid of sample points, rho_s the
// particle-radius poly6 splat of the density grid (SPH-3D-2/sph_system.cu:1954-1990).
//
// Forward: the product kernels with the counts kept on the device (no host round trip per
// step) and a field-aware integrate. Each step's tape entry holds its pre-integrate state in
// cell order, density, force, cell lists and its topology: the split child of each parent,
// the second member of each merge group (indexed by its survivor) and each post-split
// entry's slot after the step's compaction (SPH_DEAD for a merged-away entry). Every step is
// kept, or the steps of one checkpoint segment at a time, replayed from the segment's
// checkpoint (state and counters) during the backward pass.
//
// Backward, per step, over that step's frozen topology (split and merge decisions, wall
// clamps, the surface-tension threshold as the forward took them), in that step's cell
// order, gathers only:
//   topology^T   the next step's state adjoint (by slot) back to the post-split entries;
//                merge^T spreads a survivor's adjoint over its group by the volume weights
//                V = m/rho and adds the weights' density adjoints; split^T adds a child's
//                adjoint to its parent (the split offsets are constants)
//   integrate^T  per particle: adjoints of force, density, field, state (clamps frozen); the
//                field gradient is summed per block in a fixed order, then over blocks
//   surface      per particle: the colour-field direction's adjoint (threshold frozen)
//   force^T      per particle over its neighbours: pressure, viscosity and colour-field
//                pair terms with each particle's own level constants, as receiver and as
//                neighbour
//   density^T    per particle over its neighbours, then the state adjoint by slot
// All floating point is fp32 and every reduction has a fixed order, so a backward pass
// gives the same bits on every run.
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
// with a fixed butterfly, so the result is the same on every run. Small rollouts would
// otherwise leave most of the GPU idle, one thread walking all 27 cells per particle.
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

__device__ inline int lev_of(float w) { return __float_as_uint(w) == 1u ? 0 : 1; }

// Work-item count from a device counter, clamped to the buffers' size.
__device__ inline uint cnt_of(const uint *n_dev, uint cap) { return min(*n_dev, cap); }

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

// ---- integrate with the field (SPH-3D-2/sph_system.cu:204-336, no push, no obstacle
// spheres): the post-integrate position, velocity and ev of tape entry i, the wall clamps
// taken per axis, the total force and the field weights ----
struct Integ {
  float pp[3], vv[3], ee[3], aa[3];
  bool clamped[3];
  FW w;
};

__device__ inline Integ integrate_one(const DevParams &P, const float *__restrict__ theta, int nf,
                                      const DevState &T, uint i) {
  Integ r;
  const float4 p = T.pos[i], v = T.vel[i], e = T.ev[i], a = T.acc[i];
  const float dens = T.dp[i].x;
  const float radius = P.radius[lev_of(p.w)];
  const float dt = P.time_step;
  float pp[3] = {p.x, p.y, p.z}, vv[3] = {v.x, v.y, v.z}, ee[3] = {e.x, e.y, e.z};
  float aa[3] = {a.x, a.y, a.z};
  if (nf > 0) {
    float f[3];
    r.w = field_weights(P, nf, p.x, p.y, p.z);
    field_eval(theta, nf, r.w, f);
    for (int k = 0; k < 3; k++) aa[k] = aa[k] + f[k];
  }
  for (int k = 0; k < 3; k++) vv[k] = vv[k] + aa[k] * dt / dens;
  for (int k = 0; k < 3; k++) vv[k] = vv[k] + P.gravity[k] * dt / dens;
  for (int k = 0; k < 3; k++) pp[k] = pp[k] + vv[k] * dt;
  for (int k = 0; k < 3; k++) {
    r.clamped[k] = false;
    if (pp[k] >= P.world[k] - radius) {
      vv[k] = vv[k] * P.wall_damping;
      pp[k] = P.world[k] - radius;
      r.clamped[k] = true;
    }
    if (pp[k] < radius) {
      vv[k] = vv[k] * P.wall_damping;
      pp[k] = radius;
      r.clamped[k] = true;
    }
  }
  for (int k = 0; k < 3; k++) ee[k] = (ee[k] + vv[k]) / 2;
  for (int k = 0; k < 3; k++) {
    r.pp[k] = pp[k];
    r.vv[k] = vv[k];
    r.ee[k] = ee[k];
    r.aa[k] = aa[k];
  }
  return r;
}

// forward integrate: tape entry T (cell order) into the next state S, same index
__global__ __launch_bounds__(ABLK) void k_adj_integrate_fwd(DevParams P, const float *__restrict__ theta,
                                                            int nf, DevState T, DevState S,
                                                            const uint *n_dev, uint cap) {
  const uint i = blockIdx.x * ABLK + threadIdx.x;
  if (i >= cnt_of(n_dev, cap)) return;
  const Integ r = integrate_one(P, theta, nf, T, i);
  const float4 p = T.pos[i], v = T.vel[i], e = T.ev[i];
  S.pos[i] = make_float4(r.pp[0], r.pp[1], r.pp[2], p.w);
  S.vel[i] = make_float4(r.vv[0], r.vv[1], r.vv[2], v.w);
  S.ev[i] = make_float4(r.ee[0], r.ee[1], r.ee[2], e.w);
  S.acc[i] = T.acc[i];
  S.dp[i] = T.dp[i];
  S.energy[i] = 0.0f;
  S.slot[i] = T.slot[i];
}

// ---- topology records of a forward step ----
// per step: live entries before the split, entries after it, live entries after merge and
// compaction, parents split, merge groups
struct TapeCnt {
  uint n_pre, m, n_next, s_done, groups, pad[3];
};

__global__ void k_adj_rec_counts(const Counters *cnt, TapeCnt *tc) {
  tc->n_pre = cnt->n_pre;
  tc->m = cnt->m;
  tc->n_next = cnt->n;
  tc->s_done = cnt->s_done;
  tc->groups = cnt->groups;
}

// split child of entry k (cell order, k < n_pre): the parent of split rank r writes its
// child at n_pre + r (the 1 -> 2 split, SPH-3D-2/sph_system.cu:1576-1615)
__global__ __launch_bounds__(ABLK) void k_adj_rec_split(DevState S, const uint *__restrict__ flag_by_slot,
                                                        const uint *__restrict__ rank_by_slot,
                                                        const Counters *cnt, uint *__restrict__ child,
                                                        uint cap) {
  const uint k = blockIdx.x * ABLK + threadIdx.x;
  if (k >= cap) return;
  const uint n_pre = min(cnt->n_pre, cap);
  uint c = SPH_DEAD;
  if (k < n_pre) {
    const uint w = S.slot[k];
    if (flag_by_slot[w] && rank_by_slot[w] < cnt->s_done) c = n_pre + rank_by_slot[w];
  }
  child[k] = c;
}

// ---- density sample and loss (fp32, fixed-order reductions) ----
__device__ inline void sample_point(const DevParams &P, int ns, uint v, float x[3]) {
  const uint vv[3] = {v % (uint)ns, v / (uint)ns % (uint)ns, v / ((uint)ns * (uint)ns)};
  for (int a = 0; a < 3; a++) x[a] = ((float)vv[a] + 0.5f) * (P.world[a] / (float)ns);
}

// sum of one value per thread over the block, in a fixed tree order
__device__ inline float block_tree_sum(float x, float *red) {
  red[threadIdx.x] = x;
  __syncthreads();
  for (int off = ABLK / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  return red[0];
}

__global__ __launch_bounds__(ABLK) void k_adj_sample(DevParams P, int ns, const float4 *__restrict__ pos,
                                                     const uint *__restrict__ cstart,
                                                     const uint *__restrict__ cend,
                                                     const float *__restrict__ target,
                                                     float *__restrict__ sample,
                                                     float *__restrict__ resid,
                                                     float *__restrict__ part) {
  __shared__ float red[ABLK];
  const uint v = blockIdx.x * ABLK + threadIdx.x;
  const uint nv = (uint)ns * (uint)ns * (uint)ns;
  float sq = 0.0f;
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
        const int L = lev_of(q.w);
        const float rx = q.x - x[0], ry = q.y - x[1], rz = q.z - x[2];
        const float r2 = rx * rx + ry * ry + rz * rz;
        if (r2 < r2s[L]) {
          const float d = r2s[L] - r2;
          acc += P.mass[L] * P.poly6_radius[L] * ((d * d) * d);
        }
      }
    }
    sample[v] = acc;
    const float r = acc - target[v];
    resid[v] = r;
    sq = r * r;
  }
  const float tot = block_tree_sum(sq, red);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// the block partials in a fixed order: the loss is the same on every run
__global__ __launch_bounds__(ABLK) void k_adj_loss_sum(const float *__restrict__ part, uint nb,
                                                       float *__restrict__ loss) {
  __shared__ float red[ABLK];
  float s = 0.0f;
  for (uint k = threadIdx.x; k < nb; k += ABLK) s += part[k];
  const float tot = block_tree_sum(s, red);
  if (threadIdx.x == 0) *loss = tot;
}

// dL/dp_T: p_bar = sum_v 2 r_v m poly6_r (-6) (rad^2 - d^2)^2 (p - x_v) over the sample
// points within the particle radius; velocity and ev adjoints start at zero.
__global__ __launch_bounds__(ABLK) void k_adj_loss_bwd(DevParams P, int ns, DevState F,
                                                       const float *__restrict__ resid,
                                                       float4 *__restrict__ ap, float4 *__restrict__ av,
                                                       float4 *__restrict__ ae, const uint *n_dev,
                                                       uint cap) {
  const uint i = blockIdx.x * ABLK + threadIdx.x;
  if (i >= cnt_of(n_dev, cap)) return;
  const float4 p = F.pos[i];
  const int L = lev_of(p.w);
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
// adjoints of the post-split entries of a step (pos, vel, ev), and the density adjoints
// the merge weights add
struct Post {
  float4 *p, *v, *e;
  float *rbar;
};

// topology^T, first part: entry k of the post-split array takes the adjoint of the next
// step's state at its slot after compaction (nothing for a merged-away entry)
__global__ __launch_bounds__(ABLK) void k_adj_topo_gather(const TapeCnt *tc, const uint *__restrict__ slot_after,
                                                          const float4 *__restrict__ ap,
                                                          const float4 *__restrict__ av,
                                                          const float4 *__restrict__ ae, Post Q,
                                                          uint cap) {
  const uint k = blockIdx.x * ABLK + threadIdx.x;
  if (k >= min(tc->m, cap)) return;
  const uint s = slot_after[k];
  const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  Q.p[k] = s != SPH_DEAD ? ap[s] : z;
  Q.v[k] = s != SPH_DEAD ? av[s] : z;
  Q.e[k] = s != SPH_DEAD ? ae[s] : z;
  if (k < tc->n_pre) Q.rbar[k] = 0.0f;
}

// merge^T (SPH-3D-2/sph_system.cu:1742-1772): the survivor k of a group (k, j) holds
// y = (V_k x_k + V_j x_j) / (V_k + V_j) for x = pos, vel, ev after integrate, with
// V_k = m_L / rho_k (its level set to 1 first) and V_j = m(level_j) / rho_j. With the
// survivor's adjoint ybar: xbar_k = V_k/S ybar, xbar_j = V_j/S ybar (S = V_k + V_j), and
// rho_bar_i = -(m_i / rho_i^2) (x_i - y).ybar / S for both.
__global__ __launch_bounds__(ABLK) void k_adj_merge_bwd(DevParams P, const float *__restrict__ theta, int nf,
                                                        DevState T, const TapeCnt *tc,
                                                        const uint *__restrict__ partner, Post Q,
                                                        uint cap) {
  const uint k = blockIdx.x * ABLK + threadIdx.x;
  if (k >= min(tc->n_pre, cap)) return;
  const uint j = partner[k];
  if (j == SPH_DEAD) return;
  const Integ xk = integrate_one(P, theta, nf, T, k);
  const Integ xj = integrate_one(P, theta, nf, T, j);
  const float mk = P.mass[0], mj = P.mass[lev_of(T.pos[j].w)];
  const float rk = T.dp[k].x, rj = T.dp[j].x;
  const float Vk = mk / rk, Vj = mj / rj, iS = 1.0f / (Vk + Vj);
  const float4 yp = Q.p[k], yv = Q.v[k], ye = Q.e[k];
  const float yb[3][3] = {{yp.x, yp.y, yp.z}, {yv.x, yv.y, yv.z}, {ye.x, ye.y, ye.z}};
  const float *xs_k[3] = {xk.pp, xk.vv, xk.ee}, *xs_j[3] = {xj.pp, xj.vv, xj.ee};
  float dk = 0.0f, dj = 0.0f;  // (x_i - y).ybar
  for (int f = 0; f < 3; f++)
    for (int a = 0; a < 3; a++) {
      const float y = (Vk * xs_k[f][a] + Vj * xs_j[f][a]) * iS;
      dk += (xs_k[f][a] - y) * yb[f][a];
      dj += (xs_j[f][a] - y) * yb[f][a];
    }
  const float wk = Vk * iS, wj = Vj * iS;
  Q.p[k] = make_float4(wk * yp.x, wk * yp.y, wk * yp.z, 0.0f);
  Q.v[k] = make_float4(wk * yv.x, wk * yv.y, wk * yv.z, 0.0f);
  Q.e[k] = make_float4(wk * ye.x, wk * ye.y, wk * ye.z, 0.0f);
  Q.p[j] = make_float4(wj * yp.x, wj * yp.y, wj * yp.z, 0.0f);
  Q.v[j] = make_float4(wj * yv.x, wj * yv.y, wj * yv.z, 0.0f);
  Q.e[j] = make_float4(wj * ye.x, wj * ye.y, wj * ye.z, 0.0f);
  Q.rbar[k] = -(mk / (rk * rk)) * dk * iS;
  Q.rbar[j] = -(mj / (rj * rj)) * dj * iS;
}

// integrate^T: from the adjoint of the particle's post-integrate state (its own entry plus
// its split child's: split^T) to the adjoints of the force, density, field and input
// state. Wall clamps are recomputed, hence frozen. The field gradient of a block (its
// particles share a few field cells) is summed in LDS particle by particle in index order,
// then written as the block's partial; k_adj_theta_chunks and k_adj_theta_reduce add the
// partials in a fixed order.
constexpr int IBLK = 64;         // one wave per block
constexpr int LDS_THETA = 4096;  // node components summed in LDS (nf <= 11)
__global__ __launch_bounds__(IBLK) void k_adj_integrate_bwd(DevParams P, const float *__restrict__ theta,
                                                            int nf, DevState T, const TapeCnt *tc,
                                                            const uint *__restrict__ child, Post Q,
                                                            Bufs B, float *__restrict__ part,
                                                            uint cap) {
  __shared__ float th_lds[LDS_THETA];
  __shared__ float cw[IBLK][24];   // each particle's 8 corner weights x 3 force adjoints
  __shared__ uint cn[IBLK][8];     // and the 8 corner nodes
  const int nth = nf * nf * nf * 3;
  const uint n = min(tc->n_pre, cap);
  const uint i = blockIdx.x * IBLK + threadIdx.x;
  if ((size_t)blockIdx.x * IBLK >= n) return;  // uniform: blocks past the live count
  for (int k = threadIdx.x; k < nth; k += IBLK) th_lds[k] = 0.0f;
  bool have = false;
  if (i < n) {
    const Integ r = integrate_one(P, theta, nf, T, i);
    const float dens = T.dp[i].x;
    const float dt = P.time_step;
    float4 Pb = Q.p[i], Vb = Q.v[i], Eb = Q.e[i];
    const uint c = child[i];
    if (c != SPH_DEAD) {  // split^T: the child copied the parent's state (offset constant)
      const float4 a = Q.p[c], b = Q.v[c], e = Q.e[c];
      Pb = make_float4(Pb.x + a.x, Pb.y + a.y, Pb.z + a.z, 0.0f);
      Vb = make_float4(Vb.x + b.x, Vb.y + b.y, Vb.z + b.z, 0.0f);
      Eb = make_float4(Eb.x + e.x, Eb.y + e.y, Eb.z + e.z, 0.0f);
    }
    const float pb[3] = {Pb.x, Pb.y, Pb.z}, vb[3] = {Vb.x, Vb.y, Vb.z}, eb[3] = {Eb.x, Eb.y, Eb.z};
    float v3b[3], v2b[3], pbar[3], ebar[3];
    for (int k = 0; k < 3; k++) {
      v3b[k] = vb[k] + 0.5f * eb[k];  // ev' = (ev + v')/2
      ebar[k] = 0.5f * eb[k];
      if (r.clamped[k]) {
        v2b[k] = P.wall_damping * v3b[k];
        pbar[k] = 0.0f;
      } else {
        v2b[k] = v3b[k];
        pbar[k] = pb[k];
      }
      v2b[k] += pbar[k] * dt;  // p' = p + v2 dt
    }
    float rb = Q.rbar[i];  // from the merge weights
    const float idd = dt / (dens * dens);
    for (int k = 0; k < 3; k++) rb -= v2b[k] * (P.gravity[k] + r.aa[k]) * idd;
    float ab[3];
    for (int k = 0; k < 3; k++) ab[k] = v2b[k] * dt / dens;
    if (nf > 0) {
      const FW &w = r.w;
      const float wx[2] = {1.0f - w.t[0], w.t[0]}, wy[2] = {1.0f - w.t[1], w.t[1]},
                  wz[2] = {1.0f - w.t[2], w.t[2]};
      for (int dz = 0; dz < 2; dz++)
        for (int dy = 0; dy < 2; dy++)
          for (int dx = 0; dx < 2; dx++) {
            const int q = (dz * 2 + dy) * 2 + dx;
            cn[threadIdx.x][q] = (uint)(((size_t)(w.i0[2] + dz) * nf + (size_t)(w.i0[1] + dy)) * nf +
                                        (size_t)(w.i0[0] + dx));
            const float wt = wx[dx] * wy[dy] * wz[dz];
            for (int cc = 0; cc < 3; cc++) cw[threadIdx.x][q * 3 + cc] = wt * ab[cc];
          }
      have = true;
      // d field / d position through t = position / world * (nf - 1)
      for (int ax = 0; ax < 3; ax++) {
        if (!w.in[ax]) continue;
        const float sc = (float)(nf - 1) / P.world[ax];
        for (int cc = 0; cc < 3; cc++) {
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
                dfd += wt * th(theta, nf, w, dx, dy, dz, cc);
              }
          pbar[ax] += ab[cc] * dfd * sc;
        }
      }
    }
    B.Fbar[i] = make_float4(ab[0], ab[1], ab[2], 0.0f);
    B.rbar[i] = rb;
    B.pbar[i] = make_float4(pbar[0], pbar[1], pbar[2], 0.0f);
    B.vbar[i] = make_float4(v2b[0], v2b[1], v2b[2], 0.0f);  // v2 = v + (a + g) dt / dens
    B.ebar[i] = make_float4(ebar[0], ebar[1], ebar[2], 0.0f);
  }
  if (nf <= 0) return;
  // particle by particle in index order, lane l < 24 adding component l of its 24 (within
  // one particle the 8 corners are distinct nodes)
  const unsigned long long hv = __ballot(have);
  __syncthreads();
  for (int q = 0; q < IBLK; q++) {
    if (!((hv >> q) & 1ull)) continue;
    if (threadIdx.x < 24) {
      const uint node = cn[q][threadIdx.x / 3];
      th_lds[node * 3 + threadIdx.x % 3] += cw[q][threadIdx.x];
    }
    __syncthreads();
  }
  float *dst = part + (size_t)blockIdx.x * nth;
  for (int k = threadIdx.x; k < nth; k += IBLK) dst[k] = th_lds[k];
}

// dtheta[k] += the sum of the step's block partials part[b][k] in a fixed order that depends
// on the block count only: chunks of THETA_CHUNK consecutive blocks are summed in block
// order (k_adj_theta_chunks, one thread per component and chunk), then the chunk sums in
// chunk order (k_adj_theta_reduce) -- deterministic, and about THETA_CHUNK times less
// serial latency than one sum over all blocks.
constexpr uint THETA_CHUNK = 16;
__global__ __launch_bounds__(ABLK) void k_adj_theta_chunks(const float *__restrict__ part,
                                                           const TapeCnt *tc, int nth,
                                                           float *__restrict__ csum, uint cap) {
  const int k = blockIdx.x * ABLK + threadIdx.x;
  if (k >= nth) return;
  const uint nb = (min(tc->n_pre, cap) + IBLK - 1) / IBLK;
  const uint b0 = blockIdx.y * THETA_CHUNK;
  if (b0 >= nb) return;
  const uint b1 = min(nb, b0 + THETA_CHUNK);
  float s = 0.0f;
  for (uint b = b0; b < b1; b++) s += part[(size_t)b * nth + k];
  csum[(size_t)blockIdx.y * nth + k] = s;
}
__global__ __launch_bounds__(ABLK) void k_adj_theta_reduce(const float *__restrict__ csum,
                                                           const TapeCnt *tc, int nth,
                                                           float *__restrict__ dtheta, uint cap) {
  const int k = blockIdx.x * ABLK + threadIdx.x;
  if (k >= nth) return;
  const uint nb = (min(tc->n_pre, cap) + IBLK - 1) / IBLK;
  const uint nc = (nb + THETA_CHUNK - 1) / THETA_CHUNK;
  float s = 0.0f;
  for (uint c = 0; c < nc; c++) s += csum[(size_t)c * nth + k];
  dtheta[k] += s;
}

// Per-level pair constants: [0] large (level 1), [1] small (level 2).
struct PairK {
  float h[2], k2[2], m[2], S[2], Vk[2], G[2], Lc[2], poly6[2];
  float mu, cs, sigma;
};
__device__ inline PairK pair_k(const DevParams &P) {
  PairK K;
  for (int l = 0; l < 2; l++) {
    K.h[l] = P.kernel[l];
    K.k2[l] = P.k2[l];
    K.m[l] = P.mass[l];
    K.S[l] = P.spiky[l] / P.k6[l];  // pressure kernel divided by h^6 twice (3D-2 :818)
    K.Vk[l] = P.visco_k6[l];        // viscosity kernel divided by h^6 twice (3D-2 :819)
    K.G[l] = -P.grad[l];
    K.Lc[l] = P.lapc[l];
    K.poly6[l] = P.poly6[l];
  }
  K.mu = P.viscosity;
  K.cs = P.surf_coef;
  K.sigma = P.surface_tension;
  return K;
}

// Reciprocal for the backward kernels: the hardware estimate (1 ulp). The adjoint is checked
// against derivatives of the float64 oracle, not bit for bit, so its pair terms multiply by
// reciprocals instead of dividing.
__device__ inline float adj_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// A force pair of the forward (SPH-3D-2/sph_system.cu:776-800): r2 < max(h_i^2, h_j^2),
// r2 >= 1e-4; r = sqrt(r2) as there.
__device__ inline bool force_pair(const PairK &K, int Li, int Lj, float r2) {
  const float mk2 = K.k2[Li] > K.k2[Lj] ? K.k2[Li] : K.k2[Lj];
  return r2 < mk2 && !(r2 < 0.0001f);
}

// Colour-field direction adjoint: F += cs lap g/|g| when |g| > sigma (the forward's |g|
// is acc.w, so the same particles are above the threshold); g and lap recomputed here with
// the receiver's kernel over the pairs where it reaches (h_i - r > 0, 3D-2 :801-813).
__global__ __launch_bounds__(ABLK) void k_adj_surface(DevParams P, DevState T,
                                                      const uint *__restrict__ cstart,
                                                      const uint *__restrict__ cend, Bufs B,
                                                      const TapeCnt *tc, uint cap) {
  const uint n = min(tc->n_pre, cap);
  const uint nb = (uint)(((size_t)NBL * n + ABLK - 1) / ABLK);
  if (blockIdx.x >= nb) return;
  const uint i = (xcd_block(blockIdx.x, nb) * ABLK + threadIdx.x) / NBL;
  const int g0 = (int)(threadIdx.x % NBL);
  if (i >= n) return;  // whole groups
  const PairK K = pair_k(P);
  const float4 pi = T.pos[i];
  if (!(T.acc[i].w > K.sigma)) {  // the same for the whole group
    if (g0 == 0) B.gbar[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return;
  }
  const int Li = lev_of(pi.w);
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
        const int Lj = lev_of(pj.w);
        const float ux = pi.x - pj.x, uy = pi.y - pj.y, uz = pi.z - pj.z;
        const float r2 = ux * ux + uy * uy + uz * uz;
        if (!force_pair(K, Li, Lj, r2)) continue;  // also the particle itself (r2 = 0)
        const float r = sqrtf(r2);
        if (!(K.h[Li] - r > 0.0f)) continue;
        const float d = K.k2[Li] - r2;
        const float ij = adj_rcp(T.dp[j].x);  // 1/rho_j
        const float kg = K.G[Li] * K.m[Lj] * ij * (d * d);
        gx += kg * ux;
        gy += kg * uy;
        gz += kg * uz;
        lap += K.Lc[Li] * K.m[Lj] * ij * d * r2;
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

// force^T: for particle k, every force pair with k as receiver and as neighbour (each
// particle with its own level's kernel radius, mass and constants); accumulates the
// position, ev and density adjoints of k. With u = p_k - p_j and D = Fbar_k - Fbar_j the
// pair enters the loss as -(D.u) c phi(r) + (D.(e_j - e_k)) w(r), where
// c = m_k m_j (P_k/rho_k^2 + P_j/rho_j^2), phi = (Wp_k + Wp_j)/(2r), Wp_l = S_l (h_l - r)^2,
// w = (m_k/rho_k)(m_j/rho_j) mu (Wv_k + Wv_j)/2, Wv_l = Vk_l (h_l - r), each level term only
// where h_l - r > 0; plus the colour-field terms of both receivers.
__global__ __launch_bounds__(ABLK) void k_adj_force_bwd(DevParams P, DevState T,
                                                        const uint *__restrict__ cstart,
                                                        const uint *__restrict__ cend, Bufs B,
                                                        const TapeCnt *tc, uint cap) {
  const uint n = min(tc->n_pre, cap);
  const uint nb = (uint)(((size_t)NBL * n + ABLK - 1) / ABLK);
  if (blockIdx.x >= nb) return;
  const uint k = (xcd_block(blockIdx.x, nb) * ABLK + threadIdx.x) / NBL;
  const int g0 = (int)(threadIdx.x % NBL);
  if (k >= n) return;  // whole groups
  const PairK K = pair_k(P);
  const float4 pk = T.pos[k], ek = T.ev[k];
  const int Lk = lev_of(pk.w);
  const float2 dk = T.dp[k];
  const float4 Fk = B.Fbar[k], gk = B.gbar[k];
  const float rk = dk.x, Pk = dk.y;
  const float mk = K.m[Lk];
  const float ik = adj_rcp(rk);  // 1/rho_k
  const float pdk = Pk * ik * ik;
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
        const int Lj = lev_of(pj.w);
        const float ux = pk.x - pj.x, uy = pk.y - pj.y, uz = pk.z - pj.z;
        const float r2 = ux * ux + uy * uy + uz * uz;
        if (!force_pair(K, Lk, Lj, r2)) continue;  // also the particle itself (r2 = 0)
        const float r = sqrtf(r2);
        const float ir = adj_rcp(r);
        const float mj = K.m[Lj];
        const float hk = K.h[Lk] - r, hj = K.h[Lj] - r;  // each level's kernel reaches if > 0
        const bool on_k = hk > 0.0f, on_j = hj > 0.0f;
        const float Wp = (on_k ? K.S[Lk] * hk * hk : 0.0f) + (on_j ? K.S[Lj] * hj * hj : 0.0f);
        const float dWp = (on_k ? -2.0f * K.S[Lk] * hk : 0.0f) + (on_j ? -2.0f * K.S[Lj] * hj : 0.0f);
        const float Wv = (on_k ? K.Vk[Lk] * hk : 0.0f) + (on_j ? K.Vk[Lj] * hj : 0.0f);
        const float dWv = (on_k ? -K.Vk[Lk] : 0.0f) + (on_j ? -K.Vk[Lj] : 0.0f);
        const float2 dj = T.dp[j];
        const float ij = adj_rcp(dj.x), Pj = dj.y;  // 1/rho_j, P_j
        const float4 Fj = B.Fbar[j], gj = B.gbar[j], ej = T.ev[j];
        const float Dx = Fk.x - Fj.x, Dy = Fk.y - Fj.y, Dz = Fk.z - Fj.z;
        const float Du = Dx * ux + Dy * uy + Dz * uz;
        const float mm = mk * mj;
        // pressure: -(D.u) c phi, phi = Wp / (2r)
        const float cc = mm * (pdk + Pj * ij * ij);
        const float phi = 0.5f * Wp * ir;
        const float dphi = 0.5f * (dWp * ir - Wp * ir * ir);
        const float tq = -cc * Du * dphi * ir;
        px += -cc * phi * Dx + tq * ux;
        py += -cc * phi * Dy + tq * uy;
        pz += -cc * phi * Dz + tq * uz;
        const float cb = -Du * phi;
        Pb += cb * mm * ik * ik;                    // dc/dP_k
        rb += cb * mm * (-2.0f * Pk * ik * ik * ik);  // dc/drho_k
        // viscosity: (D.(e_j - e_k)) w, w = mm mu Wv / (2 rho_k rho_j)
        const float wf = 0.5f * mm * K.mu * ik * ij;
        const float w = wf * Wv;
        const float dex = ej.x - ek.x, dey = ej.y - ek.y, dez = ej.z - ek.z;
        const float Dde = Dx * dex + Dy * dey + Dz * dez;
        ex -= Dx * w;
        ey -= Dy * w;
        ez -= Dz * w;
        const float tv = Dde * wf * dWv * ir;
        px += tv * ux;
        py += tv * uy;
        pz += tv * uz;
        rb -= Dde * w * ik;
        // colour field, k as receiver (its kernel): gbar_k . (K_j d^2 u) + lam_k L_j d r^2
        if (on_k) {
          const float d = K.k2[Lk] - r2;
          const float Kj = K.G[Lk] * mj * ij, Lj_ = K.Lc[Lk] * mj * ij;
          const float gku = gk.x * ux + gk.y * uy + gk.z * uz;
          const float a1 = Kj * d * d, b1 = -4.0f * Kj * d * gku + 2.0f * gk.w * Lj_ * (d - r2);
          px += a1 * gk.x + b1 * ux;
          py += a1 * gk.y + b1 * uy;
          pz += a1 * gk.z + b1 * uz;
        }
        // colour field, k as neighbour of receiver j (j's kernel; u_jk = -u)
        if (on_j) {
          const float d = K.k2[Lj] - r2;
          const float Kk = K.G[Lj] * mk * ik, Lk_ = K.Lc[Lj] * mk * ik;
          const float gju = gj.x * ux + gj.y * uy + gj.z * uz;
          const float a2 = -Kk * d * d, b2 = 4.0f * Kk * d * gju + 2.0f * gj.w * Lk_ * (d - r2);
          px += a2 * gj.x + b2 * ux;
          py += a2 * gj.y + b2 * uy;
          pz += a2 * gj.z + b2 * uz;
          rb -= (-Kk * d * d * gju + gj.w * Lk_ * d * r2) * ik;
        }
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

// density^T (rho_i = sum_j m_j poly6_i (h_i^2 - r^2)^3 over eps <= r2 < h_i^2, plus the self
// term), then the adjoint of the step's input state, by slot.
__global__ __launch_bounds__(ABLK) void k_adj_density_bwd(DevParams P, DevState T,
                                                          const uint *__restrict__ cstart,
                                                          const uint *__restrict__ cend, Bufs B,
                                                          float4 *__restrict__ ap, float4 *__restrict__ av,
                                                          float4 *__restrict__ ae, const TapeCnt *tc,
                                                          uint cap) {
  const uint n = min(tc->n_pre, cap);
  const uint nb = (uint)(((size_t)NBL * n + ABLK - 1) / ABLK);
  if (blockIdx.x >= nb) return;
  const uint k = (xcd_block(blockIdx.x, nb) * ABLK + threadIdx.x) / NBL;
  const int g0 = (int)(threadIdx.x % NBL);
  if (k >= n) return;  // whole groups
  const PairK K = pair_k(P);
  const float4 pk = T.pos[k];
  const int Lk = lev_of(pk.w);
  const float rbk = B.rbar[k];
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
        const int Lj = lev_of(pj.w);
        const float ux = pk.x - pj.x, uy = pk.y - pj.y, uz = pk.z - pj.z;
        const float r2 = ux * ux + uy * uy + uz * uz;
        if (r2 < P.dens_eps) continue;  // also the particle itself (r2 = 0)
        float t = 0.0f;
        if (r2 < K.k2[Lk]) {  // j in rho_k
          const float d = K.k2[Lk] - r2;
          t += rbk * K.m[Lj] * K.poly6[Lk] * d * d;
        }
        if (r2 < K.k2[Lj]) {  // k in rho_j
          const float d = K.k2[Lj] - r2;
          t += B.rbar[j] * K.m[Lk] * K.poly6[Lj] * d * d;
        }
        t *= -6.0f;
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

__global__ __launch_bounds__(ABLK) void k_fill_f32(float *p, size_t n) {
  const size_t i = (size_t)blockIdx.x * ABLK + threadIdx.x;
  if (i < n) p[i] = 0.0f;
}

}  // namespace

// ---------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------
struct SphAdj {
  DevParams P;
  unsigned char raw[sizeof(SphSysParam3D)];
  uint n0 = 0, cap = 0;  // initial particles; capacity 2 n0 (1 -> 2 splits of level-1
                         // particles and 2 -> 1 merges keep 2 #level1 + #level2 = 2 n0)
  uint flags = 0;
  int nf = 0, ns = 0;
  uint ckpt = 0;
  uint cells = 0, nvox = 0, nvb = 0;
  int kbits = 1;
  uint n_live = 0, m_final = 0;  // live particles and entries of cur after the last forward
  hipStream_t st = nullptr;
  std::vector<void *> allocs;
  DevState cur{}, scratch{}, state0{};
  Aux aux{};
  uint *keys = nullptr, *vals = nullptr, *skeys = nullptr, *svals = nullptr;
  SortScratch sc{};
  uint *s2p_hash = nullptr, *s2p_cell = nullptr, *cs_scr = nullptr, *ce_scr = nullptr;
  uint *flag_by_slot = nullptr, *rank_by_slot = nullptr, *alive_by_slot = nullptr;
  uint *newslot = nullptr, *status = nullptr;
  uint *scan_tmp = nullptr;
  Counters *cnt = nullptr;
  Counters *cnt_host = nullptr;
  float *theta = nullptr, *target = nullptr, *sample = nullptr, *resid = nullptr;
  float *dtheta = nullptr, *part = nullptr, *loss_dev = nullptr, *loss_host = nullptr;
  float *th_part = nullptr;  // per-block field-gradient partials of one step
  float *th_csum = nullptr;  // their sums over chunks of THETA_CHUNK blocks
  float4 *ap = nullptr, *av = nullptr, *ae = nullptr;
  Post post{};
  Bufs B{};
  unsigned char *aos_dev = nullptr;
  struct Tape {
    DevState T;
    uint *cs, *ce;
    uint *child, *partner, *slot_after;
    TapeCnt *tc;
  };
  std::vector<Tape> tape;
  struct Ckpt {
    DevState S;
    Counters *c;
  };
  std::vector<Ckpt> ckpts;  // state and counters at the start of each segment
  std::vector<TapeCnt> topo_host;  // per step of the last forward (read after it)
  TapeCnt *topo_dev = nullptr;     // the same, gathered on the device (capacity: steps)
  uint topo_cap = 0;
  uint steps = 0;  // of the last forward
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
    return alloc(s.pos, cap) && alloc(s.vel, cap) && alloc(s.ev, cap) && alloc(s.acc, cap) &&
           alloc(s.dp, cap) && alloc(s.energy, cap) && alloc(s.slot, cap);
  }
  bool grow_tape(size_t len) {
    while (tape.size() < len) {
      Tape t{};
      if (!alloc_state(t.T) || !alloc(t.cs, cells) || !alloc(t.ce, cells) || !alloc(t.child, cap) ||
          !alloc(t.partner, cap) || !alloc(t.slot_after, cap) || !alloc(t.tc, 1))
        return false;
      tape.push_back(t);
    }
    return true;
  }
  bool grow_ckpts(size_t len) {
    while (ckpts.size() < len) {
      Ckpt c{};
      if (!alloc(c.S.pos, cap) || !alloc(c.S.vel, cap) || !alloc(c.S.ev, cap) ||
          !alloc(c.S.slot, cap) || !alloc(c.c, 1))
        return false;
      ckpts.push_back(c);
    }
    return true;
  }
  bool grow_topo(uint steps_) {
    if (topo_cap >= steps_) return true;
    if (!alloc(topo_dev, steps_)) return false;
    topo_cap = steps_;
    return true;
  }
  hipError_t copy_state(DevState dst, DevState src, bool all) {
    hipError_t e = hipSuccess;
    auto cp = [&](void *d, const void *s, size_t b) {
      if (e == hipSuccess && d && s) e = hipMemcpyAsync(d, s, b, hipMemcpyDeviceToDevice, st);
    };
    cp(dst.pos, src.pos, 16 * (size_t)cap);
    cp(dst.vel, src.vel, 16 * (size_t)cap);
    cp(dst.ev, src.ev, 16 * (size_t)cap);
    cp(dst.slot, src.slot, 4 * (size_t)cap);
    if (all) {
      cp(dst.acc, src.acc, 16 * (size_t)cap);
      cp(dst.dp, src.dp, 8 * (size_t)cap);
      cp(dst.energy, src.energy, 4 * (size_t)cap);
    }
    return e;
  }
  hipError_t copy_cnt(Counters *dst, const Counters *src) {
    return hipMemcpyAsync(dst, src, sizeof(Counters), hipMemcpyDeviceToDevice, st);
  }

  // one forward step of state S (counters in cnt): cell order into the tape entry, next
  // state back into S; no host round trip
  void fwd_step(DevState S, const Tape &t) {
    const uint *nd = &cnt->n;
    launch_begin_step(cnt, st);  // n_pre = n, split and merge counters cleared
    launch_hash(SPH_V3D2, P, S, &cnt->m, cap, keys, s2p_hash, cnt, st);
    sph_radix_sort(keys, vals, cap, kbits, sc, &skeys, &svals, st, nd);
    sph_fill_cells(t.cs, t.ce, cells, st);
    launch_cell_bounds(skeys, cap, cells, t.cs, t.ce, st, nd);
    launch_reorder(S, t.T, svals, s2p_hash, cap, s2p_cell, st, cnt, true, nd);  // and m = n
    launch_density(SPH_V3D2, P, t.T, aux, t.cs, t.ce, cap, st, nullptr, nd);
    launch_force(SPH_V3D2, P, t.T, aux, t.cs, t.ce, cap, st, nullptr, kForceListDefault, nd, true);
    k_adj_integrate_fwd<<<agrid(cap), ABLK, 0, st>>>(P, theta, nf, t.T, S, nd, cap);
    const bool do_split = (flags & SPH_STEP_SPLIT) != 0, do_merge = (flags & SPH_STEP_MERGE) != 0;
    if (do_split || do_merge) {  // the energy decides split and merge (SPH-3D-2/sph_system.cpp:356)
      launch_make_erec(S, cap, aux.erec, st, nd);
      launch_energy(SPH_V3D2, P, S, t.cs, t.ce, cap, st, nullptr, kEnergyListDefault,
                    do_split ? flag_by_slot : nullptr, aux.erec, nd);
    }
    if (do_split) {
      sph_exclusive_scan(flag_by_slot, rank_by_slot, nd, cap, &cnt->scan_total, scan_tmp, st);
      launch_split(SPH_V3D2, P, S, cap, flag_by_slot, rank_by_slot, cnt, status, st, nullptr,
                   &cnt->n_pre);
      k_adj_rec_split<<<agrid(cap), ABLK, 0, st>>>(S, flag_by_slot, rank_by_slot, cnt, t.child, cap);
    } else {
      sph_fill_u32(t.child, SPH_DEAD, cap, st);
    }
    sph_fill_u32(t.partner, SPH_DEAD, cap, st);
    if (do_merge) {
      launch_merge(SPH_V3D2, P, S, t.cs, t.ce, nullptr, status, cnt, st, nullptr, nullptr, 0,
                   nullptr, t.partner);
      launch_alive_by_slot(S, status, cnt, cap, alive_by_slot, st);
      sph_exclusive_scan(alive_by_slot, newslot, &cnt->n, cap, &cnt->scan_total, scan_tmp, st);
      launch_renumber(S, status, newslot, cnt, cap, st);  // and n = survivors
    }
    (void)hipMemcpyAsync(t.slot_after, S.slot, 4 * (size_t)cap, hipMemcpyDeviceToDevice, st);
    k_adj_rec_counts<<<1, 1, 0, st>>>(cnt, t.tc);
  }
  void bwd_step(const Tape &t) {
    const int nth = nf > 0 ? nf * nf * nf * 3 : 0;
    k_adj_topo_gather<<<agrid(cap), ABLK, 0, st>>>(t.tc, t.slot_after, ap, av, ae, post, cap);
    k_adj_merge_bwd<<<agrid(cap), ABLK, 0, st>>>(P, theta, nf, t.T, t.tc, t.partner, post, cap);
    k_adj_integrate_bwd<<<dim3((cap + IBLK - 1) / IBLK), IBLK, 0, st>>>(P, theta, nf, t.T, t.tc,
                                                                        t.child, post, B, th_part, cap);
    if (nth > 0) {
      const uint nchunk = ((cap + IBLK - 1) / IBLK + THETA_CHUNK - 1) / THETA_CHUNK;
      k_adj_theta_chunks<<<dim3(agrid(nth).x, nchunk), ABLK, 0, st>>>(th_part, t.tc, nth, th_csum,
                                                                     cap);
      k_adj_theta_reduce<<<agrid(nth), ABLK, 0, st>>>(th_csum, t.tc, nth, dtheta, cap);
    }
    const dim3 g = agrid((size_t)NBL * cap);
    k_adj_surface<<<g, ABLK, 0, st>>>(P, t.T, t.cs, t.ce, B, t.tc, cap);
    k_adj_force_bwd<<<g, ABLK, 0, st>>>(P, t.T, t.cs, t.ce, B, t.tc, cap);
    k_adj_density_bwd<<<g, ABLK, 0, st>>>(P, t.T, t.cs, t.ce, B, ap, av, ae, t.tc, cap);
  }
  // final state -> cell order in scratch (lists cs_scr/ce_scr), sample, loss
  void final_sample() {
    const uint *nd = &cnt->n;
    launch_hash(SPH_V3D2, P, cur, &cnt->m, cap, keys, s2p_hash, cnt, st);
    sph_radix_sort(keys, vals, cap, kbits, sc, &skeys, &svals, st, nd);
    sph_fill_cells(cs_scr, ce_scr, cells, st);
    launch_cell_bounds(skeys, cap, cells, cs_scr, ce_scr, st, nd);
    // cnt->m keeps the entries of cur (merged-away ones included)
    launch_reorder(cur, scratch, svals, s2p_hash, cap, s2p_cell, st, nullptr, true, nd);
    k_adj_sample<<<dim3(nvb), ABLK, 0, st>>>(P, ns, scratch.pos, cs_scr, ce_scr, target, sample,
                                             resid, part);
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
                   uint32_t checkpoint_every, uint32_t flags, SphAdj **out) {
  if (!out) return SPH_EINVAL;
  *out = nullptr;
  if (!sysparam || n == 0 || n > (1u << 26) || field_nodes == 1 || field_nodes > 11 ||
      sample_n == 0 || sample_n > 1024 || (flags & ~(SPH_STEP_SPLIT | SPH_STEP_MERGE)))
    return SPH_EINVAL;
  const SphSysParam3D &sp = *(const SphSysParam3D *)sysparam;
  if (sp.flag0 || sp.flag1) return SPH_EINVAL;  // no push, no obstacle spheres
  SphAdj *a = new (std::nothrow) SphAdj();
  if (!a) return SPH_ECAP;
  std::memcpy(a->raw, sysparam, sizeof(SphSysParam3D));
  a->n0 = n;
  a->cap = 2 * n;
  a->flags = flags;
  a->P = sphhost::make_dev_params(SPH_V3D2, a->raw, a->cap);
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
  const size_t cap = a->cap;
  const size_t nth = (size_t)std::max(a->nf, 2) * std::max(a->nf, 2) * std::max(a->nf, 2) * 3;
  const size_t scan_n = std::max<size_t>(cap, sph_sort_hist_elems(cap));
  const size_t iblocks = (cap + IBLK - 1) / IBLK;
  bool ok = a->alloc_state(a->cur) && a->alloc_state(a->scratch) && a->alloc_state(a->state0) &&
            a->alloc(a->keys, cap) && a->alloc(a->vals, cap) && a->alloc(a->sc.keys_alt, cap) &&
            a->alloc(a->sc.vals_alt, cap) && a->alloc(a->sc.tile_hist, sph_sort_hist_elems(cap)) &&
            a->alloc(a->sc.len_dev, 1) && a->alloc(a->scan_tmp, sph_scan_tmp_elems(scan_n)) &&
            a->alloc(a->s2p_hash, cap) && a->alloc(a->s2p_cell, cap) &&
            a->alloc(a->cs_scr, a->cells) && a->alloc(a->ce_scr, a->cells) &&
            a->alloc(a->flag_by_slot, cap) && a->alloc(a->rank_by_slot, cap) &&
            a->alloc(a->alive_by_slot, cap) && a->alloc(a->newslot, cap) && a->alloc(a->status, cap) &&
            a->alloc(a->cnt, 1) && a->alloc(a->aux.frec, 3 * cap) && a->alloc(a->aux.erec, cap) &&
            a->alloc(a->aux.nbl, (size_t)SPH_NBL_K * cap) && a->alloc(a->aux.nbl_cnt, cap) &&
            a->alloc(a->theta, nth) && a->alloc(a->dtheta, nth) && a->alloc(a->th_part, iblocks * nth) &&
            a->alloc(a->th_csum, ((iblocks + THETA_CHUNK - 1) / THETA_CHUNK) * nth) &&
            a->alloc(a->target, a->nvox) && a->alloc(a->sample, a->nvox) && a->alloc(a->resid, a->nvox) &&
            a->alloc(a->part, a->nvb) && a->alloc(a->loss_dev, 1) && a->alloc(a->ap, cap) &&
            a->alloc(a->av, cap) && a->alloc(a->ae, cap) && a->alloc(a->post.p, cap) &&
            a->alloc(a->post.v, cap) && a->alloc(a->post.e, cap) && a->alloc(a->post.rbar, cap) &&
            a->alloc(a->B.Fbar, cap) && a->alloc(a->B.gbar, cap) && a->alloc(a->B.pbar, cap) &&
            a->alloc(a->B.vbar, cap) && a->alloc(a->B.ebar, cap) && a->alloc(a->B.rbar, cap) &&
            a->alloc(a->aos_dev, cap * sizeof(SphParticle3D));
  a->aux.nbl_stride = (uint)cap;
  a->sc.scan_tmp = a->scan_tmp;
  a->sc.max_tiles = sph_sort_tiles(cap);
  if (ok && hipHostMalloc((void **)&a->cnt_host, sizeof(Counters), hipHostMallocDefault) != hipSuccess) ok = false;
  if (ok && hipHostMalloc((void **)&a->loss_host, sizeof(float), hipHostMallocDefault) != hipSuccess) ok = false;
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
  if (!aos || n != a->n0) return SPH_EINVAL;
  const SphParticle3D *q = (const SphParticle3D *)aos;
  uint32_t n_small = 0;
  for (uint32_t i = 0; i < n; i++) {
    if (q[i].level != 1 && q[i].level != 2) return SPH_EINVAL;
    n_small += q[i].level == 2 ? 1u : 0u;
  }
  if (n_small > 0 && (a->flags & SPH_STEP_SPLIT)) return SPH_EINVAL;  // the 2 n0 capacity bound
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
  if (record ? !a->grow_tape(std::max<uint32_t>(steps, 1)) : !a->grow_ckpts((steps + a->ckpt - 1) / a->ckpt + 1))
    return SPH_ECAP;
  if (!record && !a->grow_tape(1)) return SPH_ECAP;
  if (!a->grow_topo(std::max<uint32_t>(steps, 1))) return SPH_ECAP;
  (void)hipEventRecord(a->e0, st);
  if (a->copy_state(a->cur, a->state0, true) != hipSuccess) return SPH_EHIP;
  Counters zero;
  std::memset(&zero, 0, sizeof(zero));
  zero.n = zero.m = zero.n_pre = a->n0;
  *a->cnt_host = zero;
  if (hipMemcpyAsync(a->cnt, a->cnt_host, sizeof(Counters), hipMemcpyHostToDevice, st) != hipSuccess)
    return SPH_EHIP;
  for (uint32_t t = 0; t < steps; t++) {
    if (!record && t % a->ckpt == 0) {
      const SphAdj::Ckpt &c = a->ckpts[t / a->ckpt];
      if (a->copy_state(c.S, a->cur, false) != hipSuccess || a->copy_cnt(c.c, a->cnt) != hipSuccess)
        return SPH_EHIP;
    }
    const SphAdj::Tape &tp = record ? a->tape[t] : a->tape[0];
    a->fwd_step(a->cur, tp);
    if (hipMemcpyAsync(a->topo_dev + t, tp.tc, sizeof(TapeCnt), hipMemcpyDeviceToDevice, st) != hipSuccess)
      return SPH_EHIP;
  }
  a->final_sample();
  if (sample_dev &&
      hipMemcpyAsync(sample_dev, a->sample, sizeof(float) * a->nvox, hipMemcpyDefault, st) != hipSuccess)
    return SPH_EHIP;
  if (hipMemcpyAsync(a->loss_host, a->loss_dev, sizeof(float), hipMemcpyDeviceToHost, st) != hipSuccess)
    return SPH_EHIP;
  (void)hipEventRecord(a->e1, st);
  int rc = a->sync();
  if (rc) return rc;
  (void)hipEventElapsedTime(&a->fwd_ms, a->e0, a->e1);
  a->topo_host.resize(steps);
  if (steps > 0 && hipMemcpy(a->topo_host.data(), a->topo_dev, sizeof(TapeCnt) * steps,
                             hipMemcpyDeviceToHost) != hipSuccess)
    return SPH_EHIP;
  a->steps = steps;
  a->n_live = a->cnt_host->n;
  a->m_final = a->cnt_host->m;
  a->forward_done = true;
  if (loss) *loss = (double)*a->loss_host;
  return a->cnt_host->err ? SPH_EDOMAIN : SPH_OK;
}

int sph_adj_backward(SphAdj *a, float *dtheta_dev) {
  CHECK_ADJ(a);
  if (!a->forward_done) return SPH_EINVAL;
  hipStream_t st = a->st;
  (void)hipEventRecord(a->e0, st);
  const size_t nth = a->nf > 0 ? (size_t)a->nf * a->nf * a->nf * 3 : 1;
  k_fill_f32<<<agrid(nth), ABLK, 0, st>>>(a->dtheta, nth);
  // the final state is in cell order in scratch (final_sample), with the live count in cnt
  k_adj_loss_bwd<<<agrid(a->cap), ABLK, 0, st>>>(a->P, a->ns, a->scratch, a->resid, a->ap, a->av,
                                                 a->ae, &a->cnt->n, a->cap);
  if (a->ckpt == 0) {
    for (uint32_t t = a->steps; t-- > 0;) a->bwd_step(a->tape[t]);
  } else {
    if (!a->grow_tape(std::min(a->ckpt, a->steps))) return SPH_ECAP;
    const uint nseg = (a->steps + a->ckpt - 1) / a->ckpt;
    for (uint seg = nseg; seg-- > 0;) {
      const uint t0 = seg * a->ckpt, len = std::min(a->ckpt, a->steps - t0);
      // replayed in scratch (free once the loss adjoint has read the final state), so the
      // state of the forward stays in cur; the counters are the segment's
      const SphAdj::Ckpt &c = a->ckpts[seg];
      if (a->copy_state(a->scratch, c.S, false) != hipSuccess || a->copy_cnt(a->cnt, c.c) != hipSuccess)
        return SPH_EHIP;
      for (uint j = 0; j < len; j++) a->fwd_step(a->scratch, a->tape[j]);
      for (uint j = len; j-- > 0;) a->bwd_step(a->tape[j]);
    }
  }
  if (a->nf > 0 && dtheta_dev &&
      hipMemcpyAsync(dtheta_dev, a->dtheta, sizeof(float) * nth, hipMemcpyDefault, st) != hipSuccess)
    return SPH_EHIP;
  (void)hipEventRecord(a->e1, st);
  int rc = a->sync();
  if (rc) return rc;
  (void)hipEventElapsedTime(&a->bwd_ms, a->e0, a->e1);
  return SPH_OK;
}

int sph_adj_num_particles(SphAdj *a, uint32_t *n) {
  CHECK_ADJ(a);
  if (!n) return SPH_EINVAL;
  *n = a->forward_done ? a->n_live : a->n0;
  return SPH_OK;
}

int sph_adj_topology(SphAdj *a, uint32_t *out, uint32_t steps) {
  CHECK_ADJ(a);
  if (!out || !a->forward_done || steps > a->steps) return SPH_EINVAL;
  for (uint32_t t = 0; t < steps; t++) {
    out[3 * t + 0] = a->topo_host[t].n_pre;
    out[3 * t + 1] = a->topo_host[t].s_done;
    out[3 * t + 2] = a->topo_host[t].groups;
  }
  return SPH_OK;
}

int sph_adj_download(SphAdj *a, void *aos, uint32_t n) {
  CHECK_ADJ(a);
  if (!aos || !a->forward_done || n != a->n_live) return SPH_EINVAL;
  // the forward's final state: cur, m_final entries (merged-away ones marked), by slot
  launch_slot2phys(a->cur, a->m_final, a->s2p_cell, a->st);
  launch_download_gather(SPH_V3D2, a->cur, a->s2p_cell, n, a->aos_dev, a->st);
  if (hipMemcpyAsync(aos, a->aos_dev, (size_t)n * sizeof(SphParticle3D), hipMemcpyDeviceToHost, a->st) != hipSuccess)
    return SPH_EHIP;
  return a->sync();
}

int sph_adj_state_grad(SphAdj *a, float *dpos, float *dvel, float *dev) {
  CHECK_ADJ(a);
  if (!a->forward_done) return SPH_EINVAL;
  std::vector<float> h(4 * (size_t)a->n0);
  float *outs[3] = {dpos, dvel, dev};
  float4 *src[3] = {a->ap, a->av, a->ae};
  for (int k = 0; k < 3; k++) {
    if (!outs[k]) continue;
    if (hipMemcpyAsync(h.data(), src[k], 16 * (size_t)a->n0, hipMemcpyDeviceToHost, a->st) != hipSuccess ||
        hipStreamSynchronize(a->st) != hipSuccess)
      return SPH_EHIP;
    for (uint i = 0; i < a->n0; i++)
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
