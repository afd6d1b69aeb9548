// HIP kernels of the adaptive SPH step for gfx950 (MI355X, CDNA4, wave64).
//
// One thread per particle (or cell / voxel / cube), particles stored SoA in cell order,
// so a neighbour cell is one contiguous run [cell_start, cell_end) of the arrays and the
// neighbour loads of a wave walk a few cache lines. Floating-point operations follow the
// reference's order exactly (the library is compiled with -ffp-contract=off and with
// correctly rounded fp32 division and square root), so integer decisions (cells, split,
// merge) and most floats are bit-identical with the CPU restatement of the reference.
// Kernels are templated on the variant (2D, 3D-2, 3D-4) so the per-variant formula
// differences are resolved at compile time.
#include <algorithm>

#include "sph_internal.h"
#include "mc_tri_table.h"

namespace {

constexpr int BLK = 256;

// Work-item count: the device count when given (clamped to the host bound, which sizes the
// buffers), else the host count.
__device__ inline uint dev_count(const uint *n_dev, uint n_host) {
  return n_dev ? min(*n_dev, n_host) : n_host;
}

// Items [0, n) in blocks of BLK: block vb of nvb = ceil(n / BLK) is handled by workgroup
// vb mod gridDim.x. A plain launch has one workgroup per block and runs one iteration; a
// persistent launch (a fixed grid sized for the GPU, a multiple of the 8 XCDs) loops over
// the blocks. With XCD the blocks map XCD-contiguously (xcd_block): workgroup w runs on XCD
// w mod 8, and so does every block w + k * gridDim.x it takes.
template <bool XCD, class F> __device__ inline void for_each_item(uint n, F &&item) {
  const uint nvb = (n + BLK - 1) / BLK;
  for (uint vb = blockIdx.x; vb < nvb; vb += gridDim.x) {
    const uint i = (XCD ? xcd_block(vb, nvb) : vb) * BLK + threadIdx.x;
    if (i < n) item(i);
  }
}

// One thread per item for the kernels whose blocks cost unevenly (the neighbour loops):
// the launch has one workgroup per block of a bound of n (the host's count, or in a graph
// the buffer's size while n is read on the device), the blocks past n exit at once and the
// hardware dispatcher balances the rest; block vb runs on XCD vb mod 8 and takes block
// xcd_block(vb, nvb). False for a thread with no item.
__device__ inline bool one_item(uint n, uint &i) {
  const uint nvb = (n + BLK - 1) / BLK;
  if (blockIdx.x >= nvb) return false;
  i = xcd_block(blockIdx.x, nvb) * BLK + threadIdx.x;
  return i < n;
}

template <int V> struct VT {
  static constexpr int dim = V == SPH_V2D ? 2 : 3;
  static constexpr int nsplit = V == SPH_V3D4 ? 4 : 2;
  static constexpr int nnb = dim == 2 ? 9 : 27;
};

// Neighbour offsets in the reference's order (SPH-3D-2/sph_system.cu:456-562; 2D :383-408):
// within each z layer (z, z-1, z+1) the in-plane order is
// (0,0) (-1,-1) (-1,0) (0,-1) (1,1) (1,0) (0,1) (1,-1) (-1,1).
__device__ constexpr int kDX[9] = {0, -1, -1, 0, 1, 1, 0, 1, -1};
__device__ constexpr int kDY[9] = {0, -1, 0, -1, 1, 0, 1, -1, 1};
__device__ constexpr int kDZ[3] = {0, -1, 1};

__device__ inline int lvl_index(float w) { return __float_as_uint(w) == 1u ? 0 : 1; }

// Parameters of a particle's scene: the by-value block of a single scene, or the batch
// table entry named by the scene bits of its slot word.
template <bool B>
__device__ inline const DevParams &scene_params(const DevParams &P0,
                                                const DevParams *__restrict__ Ptab, uint w) {
  if constexpr (B) return Ptab[w >> SPH_SCENE_SHIFT];
  else return P0;
}

// calc_grid_pos (SPH-3D-2/sph_system.cu:18-27): (uint)floor(p / cell_size); false when the
// coordinate is outside [0, ncell) or NaN.
template <int V>
__device__ inline bool cell_of(const DevParams &P, float x, float y, float z, uint c[3]) {
  float q0 = floorf(x / P.cell_size);
  float q1 = floorf(y / P.cell_size);
  bool ok = (q0 >= 0.0f) && (q0 < (float)P.ncell[0]) && (q1 >= 0.0f) && (q1 < (float)P.ncell[1]);
  c[0] = ok ? (uint)q0 : 0u;
  c[1] = ok ? (uint)q1 : 0u;
  c[2] = 0u;
  if (VT<V>::dim == 3) {
    float q2 = floorf(z / P.cell_size);
    bool ok2 = (q2 >= 0.0f) && (q2 < (float)P.ncell[2]);
    c[2] = ok2 ? (uint)q2 : 0u;
    ok = ok && ok2;
  }
  return ok;
}

// neighbor_status (SPH-3D-2/sph_system.cu:564-652) and calc_grid_hash (:29-32)
template <int V>
__device__ inline bool neighbor_cell(const DevParams &P, int o, const uint c[3], uint &h) {
  const int g = o % 9, dx = kDX[g], dy = kDY[g], dz = kDZ[o / 9];
  if (dx == -1 && c[0] == 0) return false;
  if (dx == 1 && c[0] == P.ncell[0] - 1) return false;
  if (dy == -1 && c[1] == 0) return false;
  if (dy == 1 && c[1] == P.ncell[1] - 1) return false;
  if (VT<V>::dim == 3) {
    if (dz == -1 && c[2] == 0) return false;
    if (dz == 1 && c[2] == P.ncell[2] - 1) return false;
  }
  h = P.cell_base + (c[2] + (uint)dz) * P.ncell[0] * P.ncell[1] + (c[1] + (uint)dy) * P.ncell[0] +
      (c[0] + (uint)dx);
  return true;
}

// (dx+1) | (dy+1) << 2 of the nine in-plane offsets, 4 bits each, in offset order,
// derived at compile time from the same in-plane order as kDX/kDY.
constexpr unsigned long long make_off_xy() {
  constexpr int dx[9] = {0, -1, -1, 0, 1, 1, 0, 1, -1};
  constexpr int dy[9] = {0, -1, 0, -1, 1, 0, 1, -1, 1};
  unsigned long long v = 0;
  for (int g = 0; g < 9; g++)
    v |= (unsigned long long)((dx[g] + 1) | ((dy[g] + 1) << 2)) << (4 * g);
  return v;
}
constexpr unsigned long long kOffXY = make_off_xy();
static_assert(kOffXY == 0x8296a1405ull, "in-plane offset table");

// Particle range [s, e) of neighbour cell o (o wave-uniform, loop not unrolled); empty
// when the reference's neighbor_status disables the cell or it holds no particle.
template <int V>
__device__ inline void nb_range(const DevParams &P, int o, const uint c[3],
                                const uint *__restrict__ cstart, const uint *__restrict__ cend,
                                uint &s, uint &e) {
  const int g = o % 9, zl = o / 9;
  const int code = (int)((kOffXY >> (4 * g)) & 0xfull);
  const int dx = (code & 3) - 1, dy = (code >> 2) - 1;
  const int dz = zl == 0 ? 0 : (zl == 1 ? -1 : 1);
  s = e = 0;
  if (dx == -1 && c[0] == 0) return;
  if (dx == 1 && c[0] == P.ncell[0] - 1) return;
  if (dy == -1 && c[1] == 0) return;
  if (dy == 1 && c[1] == P.ncell[1] - 1) return;
  if (VT<V>::dim == 3) {
    if (dz == -1 && c[2] == 0) return;
    if (dz == 1 && c[2] == P.ncell[2] - 1) return;
  }
  const uint h = P.cell_base + (c[2] + (uint)dz) * P.ncell[0] * P.ncell[1] +
                 (c[1] + (uint)dy) * P.ncell[0] + (c[0] + (uint)dx);
  const uint st = cstart[h];
  const uint en = cend[h];  // issued with the start load; 0 for an empty cell
  if (st == SPH_EMPTY) return;
  s = st;
  e = en;
}

// Interacting-pair lists: entry = (offset index << 27) | particle index, kept in LDS
// column-major ([k][thread]) so a wave's accesses hit distinct banks.
// The list length only moves the drain points, never the order in which a particle's
// pairs are summed, so every length gives the same bits; the autotuner picks one per
// kernel (sph_internal.h: kForceLists, kEnergyLists).
__device__ inline uint pack_entry(int o, uint j) { return ((uint)o << 27) | j; }

// ---------------------------------------------------------------------------------------
// A2 calcHashD (SPH-3D-2/sph_system.cu:78-96). The key goes to the particle's slot
// position, so the sort input is in slot order and equal hashes stay in ascending slot.
// ---------------------------------------------------------------------------------------
template <int V, bool B>
__global__ __launch_bounds__(BLK) void k_hash(DevParams P0, const DevParams *__restrict__ Ptab,
                                              SceneCnt *scn, const float4 *__restrict__ pos,
                                              const uint *__restrict__ slot, const uint *m_dev,
                                              uint m_host, uint *__restrict__ keys,
                                              uint *__restrict__ slot2phys, Counters *cnt) {
  for_each_item<false>(dev_count(m_dev, m_host), [&](const uint k) {
  const uint w = slot[k];
  if (w == SPH_DEAD) return;
  const DevParams &P = scene_params<B>(P0, Ptab, w);
  // sort input position: the slot (a batch packs its scenes' slots densely, scene-major)
  const uint s = B ? scn[w >> SPH_SCENE_SHIFT].first + (w & SPH_LOCAL_MASK) : w;
  float4 p = pos[k];
  uint c[3];
  uint h = P.cell_base;
  if (cell_of<V>(P, p.x, p.y, p.z, c)) {
    h += c[2] * P.ncell[0] * P.ncell[1] + c[1] * P.ncell[0] + c[0];
  } else {
    atomicOr(&cnt->err, 1u);  // the scene's cell 0 keeps every later access in bounds
    if (B) atomicOr(&scn[w >> SPH_SCENE_SHIFT].err, 1u);
  }
  keys[s] = h;
  slot2phys[s] = k;
  });
}

// A4 find_start_end_kernel (SPH-3D-2/sph_system.cu:131-173): run boundaries of the
// sorted hashes. The cell tables are cleared by a fill kernel beforehand.
__global__ __launch_bounds__(BLK) void k_cell_bounds(const uint *__restrict__ skeys,
                                                     const uint *n_dev, uint n_host, uint ncell,
                                                     uint *__restrict__ cstart,
                                                     uint *__restrict__ cend) {
  const uint n = dev_count(n_dev, n_host);
  for_each_item<false>(n, [&](const uint k) {
    const uint h = skeys[k];
    if (h >= ncell) return;  // keys are cell ids below ncell; never written past the tables
    if (k == 0 || h != skeys[k - 1]) {
      cstart[h] = k;
      if (k > 0 && skeys[k - 1] < ncell) cend[skeys[k - 1]] = k;
    }
    if (k == n - 1) cend[h] = n;
  });
}

// Gather the fields into cell order: destination k takes slot svals[k].
// set_m (optional): thread 0 also sets cnt->m = cnt->n, physical == live from here on.
// all == false leaves out acc, dp and energy: inside a step the force, density and energy
// kernels overwrite them for every particle before anything reads them.
__global__ __launch_bounds__(BLK) void k_reorder(DevState src, DevState dst,
                                                 const uint *__restrict__ svals,
                                                 const uint *__restrict__ slot2phys,
                                                 const uint *n_dev, uint n_host,
                                                 uint *__restrict__ slot2phys_new,
                                                 Counters *set_m, bool all) {
  if (set_m && blockIdx.x == 0 && threadIdx.x == 0) set_m->m = set_m->n;
  for_each_item<false>(dev_count(n_dev, n_host), [&](const uint k) {
  const uint s = svals[k];
  const uint q = slot2phys[s];
  dst.pos[k] = src.pos[q];
  dst.vel[k] = src.vel[q];
  dst.ev[k] = src.ev[q];
  if (all) {
    dst.acc[k] = src.acc[q];
    dst.dp[k] = src.dp[q];
    dst.energy[k] = src.energy[q];
  }
  dst.slot[k] = src.slot[q];  // the slot word (a batch's keeps its scene bits; s is dense)
  slot2phys_new[s] = k;
  });
}

// ---------------------------------------------------------------------------------------
// A5 compute_density_kernel + compute_cell_density (SPH-3D-2/sph_system.cu:338-677;
// 3D-4 :345-684; 2D :268-460). Per-cell partial sums, accumulated in offset order.
// ---------------------------------------------------------------------------------------
template <int V, bool B>
__global__ __launch_bounds__(BLK) void k_density(DevParams P0, const DevParams *__restrict__ Ptab,
                                                 const uint *__restrict__ slot,
                                                 const float4 *__restrict__ pos,
                                                 const float4 *__restrict__ ev,
                                                 float2 *__restrict__ dp,
                                                 float4 *__restrict__ frec,
                                                 const uint *__restrict__ cstart,
                                                 const uint *__restrict__ cend,
                                                 const uint *n_dev, uint n_host,
                                                 uint *__restrict__ nbl,
                                                 uint *__restrict__ nbl_cnt, uint nbl_stride) {
  constexpr bool D3 = VT<V>::dim == 3;
  uint i;
  if (!one_item(dev_count(n_dev, n_host), i)) return;
  const __amdgpu_buffer_rsrc_t prs = sph_rsrc(pos, n_host);  // candidates' positions
  const DevParams &P = scene_params<B>(P0, Ptab, B ? slot[i] : 0u);
  const float4 pi = pos[i];
  const int L = lvl_index(pi.w);
  const float k2 = L ? P.k2[1] : P.k2[0];
  const float k6 = L ? P.k6[1] : P.k6[0];
  const float poly6 = L ? P.poly6[1] : P.poly6[0];
  const float mi = L ? P.mass[1] : P.mass[0];
  const float eps = P.dens_eps;
  // the term's leading factors, each the same single product the term evaluates first:
  // m_j * poly6 for the neighbour's level, and the self term (m_i * poly6) * h^6
  const float mp0 = P.mass[0] * poly6, mp1 = P.mass[1] * poly6;
  const float self_t = mi * poly6 * k6;
  uint c[3];
  float total = 0.0f;
  const bool valid = cell_of<V>(P, pi.x, pi.y, pi.z, c);
  // squared distance to a candidate (from pj - pi: the same float as from pi - pj)
  auto dist2 = [&](const float4 &pj) {
    const float dx = pj.x - pi.x, dy = pj.y - pi.y;
    if (VT<V>::dim == 3) {
      const float dz = pj.z - pi.z;
      return dx * dx + dy * dy + dz * dz;
    }
    return dx * dx + dy * dy;
  };
  // j != i: the term if eps <= r2 < k2; j == i: the self term (j == i has r2 = 0 < eps, or
  // NaN, so testing r2 first picks the same case); otherwise nothing. Evaluated without a
  // branch: adding +0 to the cell sum, which starts at +0 and only gains positive terms,
  // changes no bits.
  auto term = [&](float &cs, uint j, float r2, bool lj) {
    const float d = k2 - r2;
    const float tv = (lj ? mp1 : mp0) * ((d * d) * d);
    cs = cs + ((!(r2 < eps) && r2 < k2) ? tv : (j == i ? self_t : 0.0f));
  };
  if (!nbl) {
    if (valid) {
#pragma unroll
      for (int o = 0; o < VT<V>::nnb; o++) {
        uint h;
        if (!neighbor_cell<V>(P, o, c, h)) continue;
        float cs = 0.0f;
        const uint s = cstart[h];
        if (s != SPH_EMPTY) {
          const uint e = cend[h];
          // four loads in flight, then the terms in list order
          for (uint j = s; j < e; j += 4) {
            float4 q[4];
            sph_group4(prs, j, q);
            term(cs, j, dist2(q[0]), lvl_index(q[0].w));
            if (j + 1 < e) term(cs, j + 1, dist2(q[1]), lvl_index(q[1].w));
            if (j + 2 < e) term(cs, j + 2, dist2(q[2]), lvl_index(q[2].w));
            if (j + 3 < e) term(cs, j + 3, dist2(q[3]), lvl_index(q[3].w));
          }
        }
        total = total + cs;
      }
    }
  } else {
    // The same walk, also collecting the force kernel's pairs (r2 < max(h_i^2, h_j^2),
    // r2 >= 1e-4; 2D also j == i, its self term) in visiting order: each shifted into a
    // group of four registers (oldest first), stored as one 16-byte group when complete;
    // the last, partial group is moved to the front and stored at the end. 3D sets
    // eps = 1e-4 (the force test's bound), so the tests share their comparisons.
    uint4 *const g4 = reinterpret_cast<uint4 *>(nbl) + i;
    const size_t gs = nbl_stride;
    const float feps = D3 ? eps : 0.0001f;
    // r2 < max(h_i^2, h_j^2) for a neighbour of either level: r2 < k2 || r2 < k2j
    const float fk0 = k2 > P.k2[0] ? k2 : P.k2[0], fk1 = k2 > P.k2[1] ? k2 : P.k2[1];
    uint4 grp = make_uint4(0u, 0u, 0u, 0u);
    uint nc = 0;
    if (valid) {
#pragma unroll
      for (int o = 0; o < VT<V>::nnb; o++) {
        sph_walk_prio<VT<V>::nnb>(o);
        uint h;
        if (!neighbor_cell<V>(P, o, c, h)) continue;
        float cs = 0.0f;
        const uint s = cstart[h];
        if (s != SPH_EMPTY) {
          const uint e = cend[h];
          auto cand = [&](uint j, const float4 &pj) {
            const float r2 = dist2(pj);
            const bool lj = lvl_index(pj.w);
            term(cs, j, r2, lj);
            if ((!D3 && j == i) || (!(r2 < feps) && r2 < (lj ? fk1 : fk0))) {
              grp = make_uint4(grp.y, grp.z, grp.w, pack_entry(o, j));
              if ((nc & 3u) == 3u && nc < SPH_NBL_K) g4[(nc >> 2) * gs] = grp;
              nc++;
            }
          };
          for (uint j = s; j < e; j += 4) {
            float4 q[4];
            sph_group4(prs, j, q);
            cand(j, q[0]);
            if (j + 1 < e) cand(j + 1, q[1]);
            if (j + 2 < e) cand(j + 2, q[2]);
            if (j + 3 < e) cand(j + 3, q[3]);
          }
        }
        total = total + cs;
      }
    }
    const uint r = nc & 3u;  // entries of the partial group, in its last r places
    if (r && nc < SPH_NBL_K) {
      const uint4 v = r == 1u ? make_uint4(grp.w, 0u, 0u, 0u)
                    : r == 2u ? make_uint4(grp.z, grp.w, 0u, 0u)
                              : make_uint4(grp.y, grp.z, grp.w, 0u);
      g4[(nc >> 2) * gs] = v;
    }
    nbl_cnt[i] = nc;
  }
  const float q = total / P.rest_density;
  const float press = ((q * q) * q - 1.0f) * P.gas_constant;
  dp[i] = make_float2(total, press);
  // Per-particle factors of the force pair terms, each the same single IEEE operation
  // chain the pair body would evaluate (so hoisting them changes no bits):
  //   m/dens, press/(dens*dens), ((-1*grad_L)*m)/dens and (lapc_L*m)/dens for L = 0, 1,
  // stored with ev as the force kernel's 48-byte neighbour record (layout at struct Aux).
  const float2 v2 = make_float2(mi / total, press / (total * total));
  const float4 evi = ev[i];
  if (VT<V>::dim == 3) {
    const float4 g = make_float4((-1.0f) * P.grad[0] * mi / total, (-1.0f) * P.grad[1] * mi / total,
                                 P.lapc[0] * mi / total, P.lapc[1] * mi / total);
    // a neighbour of level L reads the first part and part 1 + L only
    frec[3 * (size_t)i] = make_float4(v2.x, v2.y, evi.x, evi.y);
    frec[3 * (size_t)i + 1] = make_float4(g.x, g.z, evi.z, 0.0f);
    frec[3 * (size_t)i + 2] = make_float4(g.y, g.w, evi.z, 0.0f);
  } else {
    frec[3 * (size_t)i] = make_float4(v2.x, v2.y, evi.x, evi.y);
  }
}

// ---------------------------------------------------------------------------------------
// A6 compute_force_kernel + compute_cell_force (SPH-3D-2/sph_system.cu:679-1137;
// 3D-4 :686-1144; 2D :462-790). acc.w receives |grad c| ("surface").
// Two phases per round: every lane walks its candidates in the reference order and
// appends the interacting ones (r2 < max(h_i^2, h_j^2), r2 >= 1e-4) to its LDS list; then
// the pair bodies run over the lists with all lanes busy. Per-cell force partials are
// flushed when the cell changes; a cell without interacting pairs would add +0 to a sum
// that is never -0, so leaving it out changes no bits.
// ---------------------------------------------------------------------------------------
// A single scene's force kernel fits 96 VGPRs (5 waves per SIMD, where the fallback list's
// LDS allows it) without spilling; the batch's (scene parameters per particle) does not.
template <int V, bool B, int LIST_F>
__global__ __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu((B || LIST_F > 32) ? 1 : 5)))
void k_force(DevParams P0, const DevParams *__restrict__ Ptab,
                                               const uint *__restrict__ slot,
                                               const float4 *__restrict__ pos,
                                               const float4 *__restrict__ ev,
                                               const float2 *__restrict__ dp,
                                               const float4 *__restrict__ frec,
                                               float4 *__restrict__ acc,
                                               const uint *__restrict__ cstart,
                                               const uint *__restrict__ cend,
                                               const uint *n_dev, uint n_host,
                                               const uint *__restrict__ nbl,
                                               const uint *__restrict__ nbl_cnt,
                                               uint nbl_stride) {
  constexpr bool D3 = VT<V>::dim == 3;
  __shared__ uint list[LIST_F * BLK];  // each thread uses its own column only
  const uint t = threadIdx.x;
  uint i;
  if (!one_item(dev_count(n_dev, n_host), i)) return;
  const __amdgpu_buffer_rsrc_t prs = sph_rsrc(pos, n_host);  // candidates' positions
  const DevParams &P = scene_params<B>(P0, Ptab, B ? slot[i] : 0u);
  const float4 pi = pos[i];
  const float4 evi = ev[i];
  const float2 dpi = dp[i];
  const float4 fri = frec[3 * (size_t)i];
  const float2 vpi = make_float2(fri.x, fri.y);
  const int Li = lvl_index(pi.w);
  const float hi = Li ? P.kernel[1] : P.kernel[0];
  const float mi = Li ? P.mass[1] : P.mass[0];
  const float k2i = Li ? P.k2[1] : P.k2[0];
  const float spi = Li ? P.spiky[1] : P.spiky[0];
  const float k6i = Li ? P.k6[1] : P.k6[0];
  const float rk6i = Li ? P.rcp_k6[1] : P.rcp_k6[0];  // RN(1/k6)
  const bool k6_fast = P.k6_fast != 0;
  const float vki = Li ? P.visco_k6[1] : P.visco_k6[0];  // visco/h^6
  const float gi = Li ? P.grad[1] : P.grad[0];
  const float lci = Li ? P.lapc[1] : P.lapc[0];
  const float iV = vpi.x;   // mi / dens_i
  const float pdi = vpi.y;  // press_i / dens_i^2
  float tfx = 0.0f, tfy = 0.0f, tfz = 0.0f;
  float cfx = 0.0f, cfy = 0.0f, cfz = 0.0f;
  float sx = 0.0f, sy = 0.0f, sz = 0.0f, sw = 0.0f;
  int cur_o = -1;
  uint c[3];
  const bool valid = cell_of<V>(P, pi.x, pi.y, pi.z, c);
  uint cnt = 0;
  // pair bodies over the collected list, in visiting order; the neighbour data of two
  // entries is fetched before either is evaluated
  struct Nb {
    uint j;
    int eo;
    float4 pj, evj;
    float2 vpj, q;  // q: the neighbour's colour-field gradient and laplacian factors for level Li
  };
  auto fetch_ent = [&](uint ent) {
    Nb f;
    f.eo = (int)(ent >> 27);
    f.j = ent & 0x7ffffffu;
    f.pj = pos[f.j];
    const float4 *r = frec + 3 * (size_t)f.j;  // the neighbour's packed record
    const float4 r0 = r[0];
    f.vpj = make_float2(r0.x, r0.y);
    if (D3) {
      const float *rl = reinterpret_cast<const float *>(r + 1 + Li);  // this level's part
      f.q = make_float2(rl[0], rl[1]);
      f.evj = make_float4(r0.z, r0.w, rl[2], 0.0f);
    } else {
      f.evj = make_float4(r0.z, r0.w, 0.0f, 0.0f);
    }
    return f;
  };
  auto fetch = [&](uint k) { return fetch_ent(list[k * BLK + t]); };
  auto body = [&](const Nb &f) {
    const uint j = f.j;
    if (f.eo != cur_o) {
      tfx = tfx + cfx;
      tfy = tfy + cfy;
      if (D3) tfz = tfz + cfz;
      cfx = cfy = cfz = 0.0f;
      cur_o = f.eo;
    }
    if (!D3 && j == i) {
      // SPH-2D/sph_system.cu:640 self term: a signed zero for finite operands
      sz += lci * mi / dpi.x * (k2i) * (0.0f - 0.0f * (k2i));
      return;
    }
    const float4 pj = f.pj;
    const float rx = pi.x - pj.x, ry = pi.y - pj.y;
    const float rz = D3 ? pi.z - pj.z : 0.0f;
    const float r2 = D3 ? rx * rx + ry * ry + rz * rz : rx * rx + ry * ry;
    const int Lj = lvl_index(pj.w);
    const float r = sph_sqrt_normal(r2);  // 1e-4 <= r2 < max(h_i^2, h_j^2) on the list
    const float2 vpj = f.vpj;
    const float4 evj = f.evj;
    const float hj = Lj ? P.kernel[1] : P.kernel[0];
    const float mj = Lj ? P.mass[1] : P.mass[0];
    const float k6j = Lj ? P.k6[1] : P.k6[0];
    const float rk6j = Lj ? P.rcp_k6[1] : P.rcp_k6[0];
    const float spj = Lj ? P.spiky[1] : P.spiky[0];
    const float vkj = Lj ? P.visco_k6[1] : P.visco_k6[0];
    const float jV = vpj.x;
    const float ikr = hi - r, jkr = hj - r;
    float ipk, ivk, jpk, jvk;
    if (ikr > 0) {
      const float a = spi * ikr * ikr;  // / k6i, correctly rounded either way
      ipk = k6_fast ? sph_div_mk(a, k6i, rk6i) : a / k6i;
      ivk = vki * (ikr);
      const float d = k2i - r2;
      if (D3) {
        const float t0 = f.q.x * (d * d);
        sx += t0 * rx;
        sy += t0 * ry;
        sz += t0 * rz;
        sw += f.q.y * d * (r2 - 0.0f * d);
      } else {
        const float t0 = (-1.0f) * gi * jV * (d * d);
        sx += t0 * rx;
        sy += t0 * ry;
        sz += lci * jV * d * (r2 - 0.0f * d);
      }
    } else {
      ipk = 0.0f;
      ivk = 0.0f;
    }
    if (jkr > 0) {
      const float a = spj * jkr * jkr;
      jpk = k6_fast ? sph_div_mk(a, k6j, rk6j) : a / k6j;
      jvk = vkj * (jkr);
    } else {
      jpk = 0.0f;
      jvk = 0.0f;
    }
    const float tp = mi * mj * (pdi + vpj.y) * (ipk + jpk) / 2;
    cfx = cfx - rx * tp / r;
    cfy = cfy - ry * tp / r;
    if (D3) cfz = cfz - rz * tp / r;
    const float tv = (iV * jV) * P.viscosity * (ivk + jvk) / 2;
    cfx = cfx + (evj.x - evi.x) * tv;
    cfy = cfy + (evj.y - evi.y) * tv;
    if (D3) cfz = cfz + (evj.z - evi.z) * tv;
  };
  auto drain = [&]() {
    uint k = 0;
    for (; k + 2 <= cnt; k += 2) {
      const Nb f0 = fetch(k), f1 = fetch(k + 1);
      body(f0);
      body(f1);
    }
    if (k < cnt) body(fetch(k));
    cnt = 0;
  };
  // The density pass's pair lists, when given and no lane's overflowed: the pair bodies
  // over them in their order (the cells' offset order, then list order within a cell),
  // four entries to a 16-byte group, the neighbour data of two pairs fetched before
  // either is evaluated.
  const uint nl = nbl ? nbl_cnt[i] : 0u;
  const bool use_list = nbl && !__any(nl > SPH_NBL_K);
  if (use_list) {
    const uint4 *g4 = reinterpret_cast<const uint4 *>(nbl) + i;
    const size_t gs = nbl_stride;
    uint4 q = nl ? g4[0] : make_uint4(0u, 0u, 0u, 0u);
    for (uint k = 0; k < nl; k += 4) {
      Nb f0 = fetch_ent(q.x), f1, f2, f3;
      if (k + 1 < nl) f1 = fetch_ent(q.y);
      if (k + 2 < nl) f2 = fetch_ent(q.z);
      if (k + 3 < nl) f3 = fetch_ent(q.w);
      if (k + 4 < nl) q = g4[((k >> 2) + 1) * gs];  // the next group, during the bodies
      body(f0);
      if (k + 1 < nl) body(f1);
      if (k + 2 < nl) body(f2);
      if (k + 3 < nl) body(f3);
    }
  }
  // Otherwise: cells in the reference's offset order, the same cell for the whole wave at a
  // time; candidates that interact (r2 < max(h_i^2, h_j^2), r2 >= 1e-4) are queued, and the
  // queue is drained with all lanes busy whenever some lane's queue would overflow
  for (int o = 0; o < (use_list ? 0 : VT<V>::nnb); o++) {
    uint s0 = 0, e0 = 0;
    if (valid) nb_range<V>(P, o, c, cstart, cend, s0, e0);
    while (true) {
      auto cand = [&](uint j, const float4 &pj) {
        // 3D: j == i has r2 = 0, which the r2 >= 1e-4 test rejects, so no separate test
        // (and no branch that would hold back the candidate's load)
        if (!D3 && j == i) {
          list[cnt++ * BLK + t] = pack_entry(o, j);  // 2D self term
          return;
        }
        const float rx = pi.x - pj.x, ry = pi.y - pj.y;
        const float rz = D3 ? pi.z - pj.z : 0.0f;
        const float r2 = D3 ? rx * rx + ry * ry + rz * rz : rx * rx + ry * ry;
        const float k2j = lvl_index(pj.w) ? P.k2[1] : P.k2[0];
        const float maxk2 = k2i > k2j ? k2i : k2j;
        if (r2 < maxk2 && !(r2 < 0.0001f)) list[cnt++ * BLK + t] = pack_entry(o, j);
      };
      // four loads in flight, tests in list order, while the list has room for four more
      for (; s0 < e0 && cnt + 4 <= LIST_F; s0 += 4) {
        float4 q[4];
        sph_group4(prs, s0, q);
        cand(s0, q[0]);
        if (s0 + 1 < e0) cand(s0 + 1, q[1]);
        if (s0 + 2 < e0) cand(s0 + 2, q[2]);
        if (s0 + 3 < e0) cand(s0 + 3, q[3]);
      }
      if (!__any(s0 < e0)) break;
      drain();
    }
  }
  drain();
  tfx = tfx + cfx;
  tfy = tfy + cfy;
  if (D3) tfz = tfz + cfz;
  (void)gi;
  float4 out;
  if (D3) {
    const float nrm = sqrtf(sx * sx + sy * sy + sz * sz);
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    if (nrm > P.surface_tension) {
      const float cl = P.surf_coef * sw;
      fx = cl * sx / nrm;
      fy = cl * sy / nrm;
      fz = cl * sz / nrm;
    }
    out = make_float4(tfx + fx, tfy + fy, tfz + fz, nrm);
  } else {
    const float nrm = sqrtf(sx * sx + sy * sy);
    float fx = 0.0f, fy = 0.0f;
    if (nrm > P.surface_tension) {
      const float cl = P.surf_coef * sz;
      fx = cl * sx / nrm;
      fy = cl * sy / nrm;
    }
    float ax = tfx + fx, ay = tfy + fy;
    if (P.mouse) {  // SPH-2D/sph_system.cu:781-789
      const float xl = pi.x - P.extern_click[0], yl = pi.y - P.extern_click[1];
      if ((xl * xl + yl * yl) <= (P.extern_len * P.extern_len)) {
        ax += P.extern_force[0] * P.extern_coe;
        ay += P.extern_force[1] * P.extern_coe;
      }
    }
    out = make_float4(ax, ay, 0.0f, nrm);
  }
  acc[i] = out;
}

// ---------------------------------------------------------------------------------------
// A8 integrate_velocity_kernel (SPH-3D-2/sph_system.cu:204-322; 3D-4 :204-329;
// 2D :198-252).
// ---------------------------------------------------------------------------------------
template <int V, bool B>
__global__ __launch_bounds__(BLK) void k_integrate(DevParams P0, const DevParams *__restrict__ Ptab,
                                                   DevState S, const uint *n_dev, uint n_host,
                                                   float4 *__restrict__ erec) {
  for_each_item<false>(dev_count(n_dev, n_host), [&](const uint i) {
  const DevParams &P = scene_params<B>(P0, Ptab, B ? S.slot[i] : 0u);
  float4 p = S.pos[i];
  float4 v = S.vel[i];
  float4 e = S.ev[i];
  float4 a = S.acc[i];
  const float dens = S.dp[i].x;
  const float radius = lvl_index(p.w) ? P.radius[1] : P.radius[0];
  const float dt = P.time_step;
  float pp[3] = {p.x, p.y, p.z}, vv[3] = {v.x, v.y, v.z}, ee[3] = {e.x, e.y, e.z};
  float aa[3] = {a.x, a.y, a.z};
  if (VT<V>::dim == 3) {
    bool acc_changed = false;
    if (V == SPH_V3D2) {
      if (P.push_left && pp[0] < P.world[0] / 3 && pp[1] > P.world[1] / 5) {
        aa[0] = aa[0] + 20.0f;
        acc_changed = true;
      }
    } else {
      if (P.push_left && pp[0] < P.world[0] / 3 && pp[1] > P.world[1] / 10) {
        aa[0] = aa[0] + 20.0f;
        acc_changed = true;
      }
      if (P.push_right && pp[0] > P.world[0] / 3 * 2 && pp[1] > P.world[1] / 10) {
        aa[0] = aa[0] - 20.0f;
        acc_changed = true;
      }
    }
#pragma unroll
    for (int k = 0; k < 3; k++) vv[k] = vv[k] + aa[k] * dt / dens;
#pragma unroll
    for (int k = 0; k < 3; k++) vv[k] = vv[k] + P.gravity[k] * dt / dens;
#pragma unroll
    for (int k = 0; k < 3; k++) pp[k] = pp[k] + vv[k] * dt;
    if (P.use_cylinder) {
      for (int q = 0; q < 3; q++) {
        const float d0 = pp[0] - P.sph_pos[q][0], d1 = pp[1] - P.sph_pos[q][1],
                    d2 = pp[2] - P.sph_pos[q][2];
        const float dist = sqrtf(d0 * d0 + d1 * d1 + d2 * d2);
        if (dist < P.sph_radius[q] + radius) {
          const float rr = P.sph_radius[q] + radius;
          float dx[3];
#pragma unroll
          for (int k = 0; k < 3; k++) dx[k] = (pp[k] - P.sph_pos[q][k]) / dist * rr;
          const float ln = sqrtf(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
          const float nm[3] = {dx[0] / ln, dx[1] / ln, dx[2] / ln};
#pragma unroll
          for (int k = 0; k < 3; k++) pp[k] = P.sph_pos[q][k] + dx[k];
          const float dv = vv[0] * nm[0] + vv[1] * nm[1] + vv[2] * nm[2];
          const float sc = P.sph_damping[q] * (dv);
#pragma unroll
          for (int k = 0; k < 3; k++) vv[k] = vv[k] - sc * nm[k];
        }
      }
    }
#pragma unroll
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
#pragma unroll
    for (int k = 0; k < 3; k++) ee[k] = (ee[k] + vv[k]) / 2;
    if (acc_changed) S.acc[i] = make_float4(aa[0], aa[1], aa[2], a.w);
  } else {
#pragma unroll
    for (int k = 0; k < 2; k++) aa[k] = aa[k] + P.gravity[k] * dt;
#pragma unroll
    for (int k = 0; k < 2; k++) aa[k] = aa[k] / dens;
#pragma unroll
    for (int k = 0; k < 2; k++) vv[k] = vv[k] + aa[k] * dt;
#pragma unroll
    for (int k = 0; k < 2; k++) ee[k] = (ee[k] + vv[k]) / 2;
#pragma unroll
    for (int k = 0; k < 2; k++) pp[k] = pp[k] + vv[k] * dt;
#pragma unroll
    for (int k = 0; k < 2; k++) {
      if (pp[k] >= P.world[k] - radius / 2) {
        vv[k] = vv[k] * P.wall_damping;
        pp[k] = P.world[k] - radius / 2;
      }
      if (pp[k] < radius / 2) {
        vv[k] = vv[k] * P.wall_damping;
        pp[k] = radius / 2;
      }
    }
    S.acc[i] = make_float4(aa[0], aa[1], a.z, a.w);
  }
  S.pos[i] = make_float4(pp[0], pp[1], pp[2], p.w);
  S.vel[i] = make_float4(vv[0], vv[1], vv[2], v.w);
  if (erec) erec[i] = make_float4(pp[0], pp[1], vv[0], vv[1]);
  S.ev[i] = make_float4(ee[0], ee[1], ee[2], e.w);
  });
}

// ---------------------------------------------------------------------------------------
// A9 compute_energy_kernel + compute_cell_energy + wavelet (SPH-3D-2/sph_system.cu:11-16,
// :1169-1546; 3D-4 :1176-1553; 2D :827-1065). Start-of-step cell lists, current positions.
// Same two-phase structure as the force kernel; per-cell wavelet sums are flushed when
// the cell changes. When the kernel length is exactly 1, x/h and x/sqrt(h) equal x, so
// those divisions are skipped.
// ---------------------------------------------------------------------------------------
// flag_by_slot (optional): the step's split decision (k_decide) is made here too, from
// the energy just computed, the surface value of the force phase and the level.
template <int V, bool B, int LIST_E, bool U>
__global__ __launch_bounds__(BLK) void k_energy(DevParams P0, const DevParams *__restrict__ Ptab,
                                                const SceneCnt *__restrict__ scn,
                                                const uint *__restrict__ slot,
                                                const float4 *__restrict__ pos,
                                                const float4 *__restrict__ vel,
                                                const float4 *__restrict__ acc,
                                                float *__restrict__ energy,
                                                const uint *__restrict__ cstart,
                                                const uint *__restrict__ cend,
                                                const uint *n_dev, uint n_host,
                                                uint *__restrict__ flag_by_slot,
                                                const float4 *__restrict__ erec) {
  constexpr bool D3 = VT<V>::dim == 3;
  __shared__ uint list[LIST_E * BLK];  // each thread uses its own column only
  const uint t = threadIdx.x;
  uint i;
  if (!one_item(dev_count(n_dev, n_host), i)) return;
  const __amdgpu_buffer_rsrc_t prs = sph_rsrc(pos, n_host);  // candidates' positions
  const DevParams &P = scene_params<B>(P0, Ptab, B ? slot[i] : 0u);
  const float4 pi = pos[i];
  const int Li = lvl_index(pi.w);
  const int Lk = D3 ? 0 : Li;  // 3D: large kernel for both levels (3D-2 :1199-1209)
  const float ker = Lk ? P.kernel[1] : P.kernel[0];
  const float ker2 = Lk ? P.k2[1] : P.k2[0];
  const float sk = Lk ? P.sqrt_kernel[1] : P.sqrt_kernel[0];
  const float rker = Lk ? P.rcp_kernel[1] : P.rcp_kernel[0];  // RN(1/ker), for sph_div_rn
  const float rsk = Lk ? P.rcp_sqrt_kernel[1] : P.rcp_sqrt_kernel[0];
  const bool div_fast = P.kernel_div_fast != 0;
  const bool unit = U || (ker == 1.0f && sk == 1.0f);
  float ttx = 0.0f, tty = 0.0f, tvx = 0.0f, tvy = 0.0f;
  float ctx = 0.0f, cty = 0.0f;
  int cur_o = -1;
  uint c[3];
  const bool valid = cell_of<V>(P, pi.x, pi.y, pi.z, c);
  uint cnt = 0;
  struct Nb {
    int eo;
    float4 pj, vj;
  };
  auto fetch_ent = [&](uint ent) {
    Nb f;
    f.eo = (int)(ent >> 27);
    const uint j = ent & 0x7ffffffu;
    const float4 r = erec[j];  // one 16-byte record: pos.xy, vel.xy
    f.pj = make_float4(r.x, r.y, 0.0f, 0.0f);
    f.vj = make_float4(r.z, r.w, 0.0f, 0.0f);
    return f;
  };
  auto body = [&](const Nb &f) {
    if (f.eo != cur_o) {
      ttx = ttx + ctx;
      tty = tty + cty;
      ctx = cty = 0.0f;
      cur_o = f.eo;
    }
    const float ax = pi.x - f.pj.x, ay = pi.y - f.pj.y;
    // the z component of the transform never reaches the energy (3D-2 :1521)
    if (unit) {
      // |ax|, |ay| <= h = 1 here, so the exp argument is within [-0.125, 0]
      const float tx = sph_wavelet_unit(ax), ty = sph_wavelet_unit(ay);
      ctx = ctx + tx;
      cty = cty + ty;
      tvx = tvx + (f.vj.x * tx);
      tvy = tvy + (f.vj.y * ty);
    } else {
      // ax / ker, ay / ker, (vx * tx) / sk, (vy * ty) / sk, each correctly rounded
      auto dv = [&](float a, float b, float y) { return div_fast ? sph_div_rn(a, b, y) : a / b; };
      const float tx = sph_wavelet_unit(dv(ax, ker, rker)), ty = sph_wavelet_unit(dv(ay, ker, rker));
      ctx = ctx + tx;
      cty = cty + ty;
      tvx = tvx + dv(f.vj.x * tx, sk, rsk);
      tvy = tvy + dv(f.vj.y * ty, sk, rsk);
    }
  };
  // The LDS list as a ring per lane (entries head .. head + cnt - 1, mod LIST_E). The
  // candidate groups advance together over the wave (the same cell, group k of every
  // lane's run); whenever some lane has fewer than four free places, every lane with
  // entries evaluates its oldest two (one, if it has one) -- so a lane's pairs are
  // evaluated in visiting order, mostly while other lanes' lists fill up, and a body row
  // is spent only when some lane needs the room.
  uint head = 0;
  auto ring = [&](uint x) {
    const uint p = head + x;
    if constexpr ((LIST_E & (LIST_E - 1)) == 0) return p & (uint)(LIST_E - 1);
    else return p >= (uint)LIST_E ? p - (uint)LIST_E : p;
  };
  auto pop2 = [&]() {
    const bool h0 = cnt > 0u, h1 = cnt > 1u;
    Nb f0, f1;
    if (h0) f0 = fetch_ent(list[ring(0) * BLK + t]);
    if (h1) f1 = fetch_ent(list[ring(1) * BLK + t]);
    if (h0) body(f0);
    if (h1) body(f1);
    const uint k = h1 ? 2u : (h0 ? 1u : 0u);
    head = ring(k);
    cnt -= k;
  };
  for (int o = 0; o < VT<V>::nnb; o++) {
    sph_walk_prio<VT<V>::nnb>(o);
    uint s0 = 0, e0 = 0;
    if (valid) nb_range<V>(P, o, c, cstart, cend, s0, e0);
    auto cand = [&](uint j, const float4 &pj) {
      if (!D3 && j == i) return;  // 3D: r2 = 0 fails the r2 >= 1e-4 test below
      const float rx = pi.x - pj.x, ry = pi.y - pj.y;
      const float rz = D3 ? pi.z - pj.z : 0.0f;
      const float r2 = D3 ? rx * rx + ry * ry + rz * rz : rx * rx + ry * ry;
      // 3D skips r2 < 1e-4; 2D clamps r2 to EXT2 and never uses it again.
      if (r2 < ker2 && !(D3 && r2 < 0.0001f)) {
        list[ring(cnt) * BLK + t] = pack_entry(o, j);
        cnt++;
      }
    };
    for (;; s0 += 4) {
      // the group's loads are issued first, so they are in flight during the rows below
      float4 q[4];
      if (s0 < e0) sph_group4(prs, s0, q);
      // every lane had room for four before the last group, so two rows restore it
      if (__any(cnt + 4u > (uint)LIST_E)) {
        pop2();
        if (__any(cnt + 4u > (uint)LIST_E)) pop2();
      }
      if (!__any(s0 < e0)) break;
      if (s0 < e0) {
        cand(s0, q[0]);
        if (s0 + 1 < e0) cand(s0 + 1, q[1]);
        if (s0 + 2 < e0) cand(s0 + 2, q[2]);
        if (s0 + 3 < e0) cand(s0 + 3, q[3]);
      }
    }
  }
  while (__any(cnt > 0u)) pop2();
  ttx = ttx + ctx;
  tty = tty + cty;
  const float ux = tvx / ttx, uy = tvy / tty;
  const float e2 = ux * ux + uy * uy;
  float E;
  if (V == SPH_V3D4) E = 0.5f * (Li ? P.mass[1] : P.mass[0]) * (e2);
  else E = 0.5f * (e2);
  energy[i] = E;
  if (flag_by_slot) {  // decide_adaptive_kernel, as k_decide
    const uint w = slot[i];
    const float sf = acc[i].w;
    const bool split = E * P.co_energy + sf * P.co_surface > P.split_criteria &&
                       __float_as_uint(pi.w) == 1u;
    flag_by_slot[B ? scn[w >> SPH_SCENE_SHIFT].first + (w & SPH_LOCAL_MASK) : w] = split ? 1u : 0u;
  }
}

// ---------------------------------------------------------------------------------------
// A10 decide_adaptive_kernel / decide_split_kernel (SPH-3D-2/sph_system.cu:1548-1574;
// 2D :1067-1090): flag 1 = split (the reference stores 0 for split and sorts ascending).
// Flags go to slot positions so the scan ranks parents in slot order.
// ---------------------------------------------------------------------------------------
template <bool B>
__global__ __launch_bounds__(BLK) void k_decide(DevParams P0, const DevParams *__restrict__ Ptab,
                                                const SceneCnt *__restrict__ scn, DevState S,
                                                uint n, uint *__restrict__ flag_by_slot) {
  const uint k = blockIdx.x * BLK + threadIdx.x;
  if (k >= n) return;
  const uint w = S.slot[k];
  const DevParams &P = scene_params<B>(P0, Ptab, w);
  const float E = S.energy[k];
  const float sf = S.acc[k].w;
  const uint lev = __float_as_uint(S.pos[k].w);
  const bool split = E * P.co_energy + sf * P.co_surface > P.split_criteria && lev == 1u;
  flag_by_slot[B ? scn[w >> SPH_SCENE_SHIFT].first + (w & SPH_LOCAL_MASK) : w] = split ? 1u : 0u;
}

// A12 adaptive_particle_kernel (3D-2 :1576-1615; 3D-4 :1583-1656) / split_particle_kernel
// (2D :1092-1131). Parent of rank r (slot order) writes its children at N + r(n-1) + c.
// Capacity policy on the counters after a single scene's split: only the first
// floor((cap - n) / (NS - 1)) flagged parents split.
template <int V> __device__ inline void finish_split_counts(const DevParams &P, Counters *cnt) {
  constexpr int NS = VT<V>::nsplit;
  const uint n = cnt->n;
  const uint total = cnt->scan_total;
  const uint room = (P.cap - n) / (uint)(NS - 1);
  const uint done = total < room ? total : room;
  cnt->n_pre = n;
  cnt->s_req = total;
  cnt->s_done = done;
  cnt->n = n + done * (uint)(NS - 1);
  cnt->m = cnt->n;
}

// A single scene's split: thread 0 also updates the counters (no thread of this kernel
// reads n, m, n_pre, s_req or s_done).
template <int V, bool B>
__global__ __launch_bounds__(BLK) void k_split(DevParams P0, const DevParams *__restrict__ Ptab,
                                               const SceneCnt *__restrict__ scn, DevState S,
                                               const uint *n_dev, uint n_host,
                                               const uint *__restrict__ flag_by_slot,
                                               const uint *__restrict__ rank_by_slot,
                                               Counters *cnt, uint *__restrict__ status) {
  constexpr int NS = VT<V>::nsplit;
  if (!B && blockIdx.x == 0 && threadIdx.x == 0) finish_split_counts<V>(P0, cnt);
  const uint n = dev_count(n_dev, n_host);  // live entries before the split
  for_each_item<false>(n, [&](const uint k) {
  const uint w = S.slot[k];
  const DevParams &P = scene_params<B>(P0, Ptab, w);
  // base: physical index of the first child; sbase: its slot word
  uint base, sbase;
  if (!B) {
    if (!flag_by_slot[w]) return;
    const uint rank = rank_by_slot[w];
    const uint total = cnt->scan_total;
    const uint room = (P.cap - n) / (uint)(NS - 1);
    const uint done = total < room ? total : room;
    if (rank >= done) return;
    base = n + rank * (uint)(NS - 1);
    sbase = base;
  } else {
    // children of a batch go after all live entries, scene by scene in parent rank order;
    // their slots continue the scene's own numbering
    const uint sc = w >> SPH_SCENE_SHIFT;
    const SceneCnt &C = scn[sc];
    const uint d = C.first + (w & SPH_LOCAL_MASK);
    if (!flag_by_slot[d]) return;
    const uint rank = rank_by_slot[d] - C.r0;
    if (rank >= C.s_done) return;
    base = n + (C.done_base + rank) * (uint)(NS - 1);
    sbase = (sc << SPH_SCENE_SHIFT) | (C.n_pre + rank * (uint)(NS - 1));
  }
  const float hs = P.kernel[1] / 2;
  float4 p = S.pos[k];
  const float4 v = S.vel[k], e = S.ev[k], a = S.acc[k];
  const float2 dpk = S.dp[k];
  const float en = S.energy[k];
  const float lev2 = __uint_as_float(2u);
  // children copy every field of the parent (3D leaves dens/press/surface/tension
  // unwritten in the reference; copies make the downloaded state defined)
  if (NS == 2) {
    S.pos[base] = make_float4(p.x - hs, p.y, p.z, lev2);
    S.vel[base] = v;
    S.ev[base] = e;
    S.acc[base] = a;
    S.dp[base] = dpk;
    S.energy[base] = en;
    S.slot[base] = sbase;
    status[base] = 0u;
    S.pos[k] = make_float4(p.x + hs, p.y, p.z, lev2);
  } else {
    const float e4 = en / 4;
    const float4 cp[3] = {make_float4(p.x - hs, p.y, p.z, lev2), make_float4(p.x, p.y, p.z - hs, lev2),
                          make_float4(p.x, p.y, p.z + hs, lev2)};
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const uint q = base + c;
      S.pos[q] = cp[c];
      S.vel[q] = v;
      S.ev[q] = e;
      S.acc[q] = a;
      S.dp[q] = dpk;
      S.energy[q] = e4;
      S.slot[q] = sbase + c;
      status[q] = 0u;
    }
    S.pos[k] = make_float4(p.x + hs, p.y, p.z, lev2);
    S.energy[k] = e4;
  }
  status[k] = 0u;
  });
}

template <int V> __global__ void k_finish_split(DevParams P, Counters *cnt) {
  finish_split_counts<V>(P, cnt);
}

// ---------------------------------------------------------------------------------------
// A13 merge_kernel (3D-2 :1661-1789; 3D-4 :1702-1888; 2D :1177-1270): one thread per cell
// walks its start-of-step list (cell order = ascending slot). The merged-group count is
// summed over the wave before its one atomic add.
// ---------------------------------------------------------------------------------------
__device__ inline uint wave_sum(uint x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

// group_out (optional): entry group_out[p] of a group's survivor p receives its second
// member (1 -> 2 split variants; the differentiable rollout records its merge topology)
template <int V, bool B>
__device__ inline uint merge_cell(const DevParams &P0, const DevParams *__restrict__ Ptab,
                                  SceneCnt *scn, uint cl, DevState &S,
                                  const uint *__restrict__ cstart, const uint *__restrict__ cend,
                                  const uint *__restrict__ q2d, uint *__restrict__ status,
                                  uint *__restrict__ group_out, const uint *__restrict__ qbits) {
  constexpr int NM = VT<V>::nsplit;
  constexpr bool D3 = VT<V>::dim == 3;
  const uint s = cstart[cl];
  if (s == SPH_EMPTY) return 0u;
  const uint e = cend[cl];
  const uint w0 = B ? S.slot[s] : 0u;  // every entry of a cell belongs to one scene
  const DevParams &P = scene_params<B>(P0, Ptab, w0);
  // 2D: qualification of the particle of slot cl (bit 0) and, when that particle is in this
  // cell's list, its entry + 1 (bits 1..31; k_merge_q2d). Only this walk writes it then, and
  // the reference re-reads it at every test, so cell_q follows its merged energy.
  const uint qw = D3 ? 1u : q2d[cl];
  bool cell_q = (qw & 1u) != 0u;
  const uint kown = (qw >> 1) ? (qw >> 1) - 1u : SPH_DEAD;
  // both tests evaluated (no short circuit), so all loads of an entry issue together
  auto qual = [&](uint j) -> bool {
    const uint lev = __float_as_uint(S.pos[j].w);
    if (D3) {
      const bool low = S.energy[j] * P.co_energy + S.acc[j].w * P.co_surface < P.merge_criteria;
      return (lev != 1u) & low;
    }
    return lev == 2u && cell_q;
  };
  // qualification of every entry, eight entries' loads in flight at a time; the first 64
  // results are kept as bits for the group walk (merges below only write entries the walk
  // has passed, so a second evaluation would give the same answer)
  uint q = 0;
  unsigned long long qb = 0ull;
  if (qbits) {
    // the entries' bits from k_merge_qual (which also cleared their status): the count over
    // [s, e) and the first 64 bits relative to s
    for (uint wi = s >> 5; wi <= (e - 1u) >> 5; wi++) {
      const uint b0 = wi << 5;
      const uint lo = s > b0 ? s - b0 : 0u, hi = e - b0 < 32u ? e - b0 : 32u;
      const uint m = (hi == 32u ? 0xffffffffu : (1u << hi) - 1u) & ~((1u << lo) - 1u);
      const uint bits = qbits[wi] & m;
      q += __popc(bits);
      if (b0 >= s) {
        if (b0 - s < 64u) qb |= (unsigned long long)bits << (b0 - s);
      } else {
        qb |= (unsigned long long)(bits >> (s - b0));
      }
    }
    if (!D3 && !cell_q) q = 0u;  // the bits hold level == 2; the cell does not qualify
  }
  for (uint j0 = s; j0 < (qbits ? s : e); j0 += 8) {
    bool qq[8];
#pragma unroll
    for (int t = 0; t < 8; t++) {  // past the run's end: its last entry again, result dropped
      const bool q_t = qual(min(j0 + (uint)t, e - 1u));
      qq[t] = (j0 + t < e) & q_t;
    }
#pragma unroll
    for (int t = 0; t < 8; t++) {
      const uint j = j0 + t;
      if (j < e) {
        status[j] = 0u;
        if (qq[t]) {
          q++;
          if (j - s < 64u) qb |= 1ull << (j - s);
        }
      }
    }
  }
  if (q < (uint)NM) return 0u;
  // the bits hold level == 2 in 2D (cell_q was true when they were set)
  auto is_q = [&](uint j) -> bool {
    const bool b = j - s < 64u ? ((qb >> (j - s)) & 1ull) != 0ull
                 : (qbits ? ((qbits[j >> 5] >> (j & 31u)) & 1u) != 0u : qual(j));
    return D3 ? b : (b && cell_q);
  };
  const uint ng = q / NM;
  uint cur = s;
  // A search that reaches the end of the list keeps its previous entry: the previous
  // pair's (SPH-2D/sph_system.cu:1221-1251). Only in 2D, after the qualification changed;
  // the first pair always exists.
  uint pm[4] = {s, s, s, s};
  for (uint g = 0; g < ng; g++) {
#pragma unroll
    for (int t = 0; t < NM; t++) {
      while (cur < e) {
        const uint j = cur++;
        if (is_q(j)) {
          pm[t] = j;
          break;
        }
      }
    }
    const uint p1 = pm[0];
    float4 P1 = S.pos[p1];
    float4 V1 = S.vel[p1], E1 = S.ev[p1], A1 = S.acc[p1];
    if (D3) {
      float Vw[4];
      Vw[0] = P.mass[0] / S.dp[p1].x;  // level set to 1 before the volume (3D-2 :1742-1755)
#pragma unroll
      for (int t = 1; t < NM; t++) {
        const uint pt = pm[t];
        Vw[t] = (lvl_index(S.pos[pt].w) ? P.mass[1] : P.mass[0]) / S.dp[pt].x;
      }
      float vs = Vw[0];
#pragma unroll
      for (int t = 1; t < NM; t++) vs = vs + Vw[t];
      float px = P1.x * Vw[0], py = P1.y * Vw[0], pz = P1.z * Vw[0];
      float vx = V1.x * Vw[0], vy = V1.y * Vw[0], vz = V1.z * Vw[0];
      float ex = E1.x * Vw[0], ey = E1.y * Vw[0], ez = E1.z * Vw[0];
      float ax = A1.x * Vw[0], ay = A1.y * Vw[0], az = A1.z * Vw[0];
#pragma unroll
      for (int t = 1; t < NM; t++) {
        const uint pt = pm[t];
        const float4 Pt = S.pos[pt], Vt = S.vel[pt], Et = S.ev[pt], At = S.acc[pt];
        px = px + Pt.x * Vw[t]; py = py + Pt.y * Vw[t]; pz = pz + Pt.z * Vw[t];
        vx = vx + Vt.x * Vw[t]; vy = vy + Vt.y * Vw[t]; vz = vz + Vt.z * Vw[t];
        ex = ex + Et.x * Vw[t]; ey = ey + Et.y * Vw[t]; ez = ez + Et.z * Vw[t];
        ax = ax + At.x * Vw[t]; ay = ay + At.y * Vw[t]; az = az + At.z * Vw[t];
      }
      S.pos[p1] = make_float4(px / vs, py / vs, pz / vs, __uint_as_float(1u));
      S.vel[p1] = make_float4(vx / vs, vy / vs, vz / vs, V1.w);
      S.ev[p1] = make_float4(ex / vs, ey / vs, ez / vs, E1.w);
      S.acc[p1] = make_float4(ax / vs, ay / vs, az / vs, A1.w);
    } else {
      const uint p2 = pm[1];
      const float4 P2 = S.pos[p2], V2 = S.vel[p2], E2 = S.ev[p2], A2 = S.acc[p2];
      S.pos[p1] = make_float4((P1.x + P2.x) / 2, (P1.y + P2.y) / 2, P1.z, __uint_as_float(1u));
      S.vel[p1] = make_float4((V1.x + V2.x) / 2, (V1.y + V2.y) / 2, V1.z, V1.w);
      S.ev[p1] = make_float4((E1.x + E2.x) / 2, (E1.y + E2.y) / 2, E1.z, E1.w);
      S.acc[p1] = make_float4((A1.x + A2.x) / 2, (A1.y + A2.y) / 2, A1.z, A1.w);
      const float e1 = (S.energy[p1] + S.energy[p2]) / 2;
      S.energy[p1] = e1;
      if (p1 == kown) cell_q = e1 * P.co_energy + A1.w * P.co_surface < P.merge_criteria;
    }
    status[p1] = 0u;
#pragma unroll
    for (int t = 1; t < NM; t++) status[pm[t]] = 1u;
    if (NM == 2 && group_out) group_out[p1] = pm[1];
  }
  if (B) atomicAdd(&scn[w0 >> SPH_SCENE_SHIFT].groups, ng);
  return ng;
}

// Merge qualification of the sorted entries j < n, one thread each (coalesced loads), as
// merge_cell's walk would evaluate it before any merge of the step (merges write only
// entries of their own cell): bit j of qbits -- 3D: level 2 and
// energy * co_energy + surface * co_surface < merge_criteria; 2D: level 2 (the cell's own
// qualification is applied in the walk) -- and status[j] = 0.
template <int V, bool B>
__global__ __launch_bounds__(BLK) void k_merge_qual(DevParams P0, const DevParams *__restrict__ Ptab,
                                                    DevState S, const uint *n_dev, uint n_host,
                                                    uint *__restrict__ qbits,
                                                    uint *__restrict__ status) {
  const uint j = blockIdx.x * BLK + threadIdx.x;
  const uint n = dev_count(n_dev, n_host);
  bool q = false;
  if (j < n) {
    const DevParams &P = scene_params<B>(P0, Ptab, B ? S.slot[j] : 0u);
    const uint lev = __float_as_uint(S.pos[j].w);
    if (VT<V>::dim == 3)
      q = (lev != 1u) &
          (S.energy[j] * P.co_energy + S.acc[j].w * P.co_surface < P.merge_criteria);
    else
      q = lev == 2u;
    status[j] = 0u;
  }
  const unsigned long long b = __ballot(q);
  const uint j0 = j & ~63u;  // the wave's first entry: its 64 bits are words j0/32, +1
  if ((threadIdx.x & 63u) == 0u && j0 < n) {
    qbits[j0 >> 5] = (uint)b;
    qbits[(j0 >> 5) + 1u] = (uint)(b >> 32);
  }
}

// One thread per cell, or, on a grid with more cells than sorted entries (skeys given),
// one thread per sorted entry, the first entry of each run taking its cell: every occupied
// cell is merged exactly once either way, and empty cells have nothing to merge.
template <int V, bool B>
__global__ __launch_bounds__(BLK) void k_merge(DevParams P0, const DevParams *__restrict__ Ptab,
                                               SceneCnt *scn, uint ncell, DevState S,
                                               const uint *__restrict__ cstart,
                                               const uint *__restrict__ cend,
                                               const uint *__restrict__ q2d,
                                               uint *__restrict__ status, Counters *cnt,
                                               const uint *__restrict__ skeys,
                                               const uint *nkeys_dev, uint nkeys_host,
                                               uint *__restrict__ group_out,
                                               const uint *__restrict__ qbits) {
  uint ng = 0;
  const uint t = blockIdx.x * BLK + threadIdx.x;
  if (skeys) {
    if (t < dev_count(nkeys_dev, nkeys_host)) {
      const uint h = skeys[t];
      if ((t == 0 || skeys[t - 1] != h) && h < ncell)
        ng = merge_cell<V, B>(P0, Ptab, scn, h, S, cstart, cend, q2d, status, group_out, qbits);
    }
  } else if (t < ncell) {
    ng = merge_cell<V, B>(P0, Ptab, scn, t, S, cstart, cend, q2d, status, group_out, qbits);
  }
  ng = wave_sum(ng);  // every lane of the wave reaches this point
  if ((threadIdx.x & 63u) == 0 && ng) atomicAdd(&cnt->groups, ng);
}

// 2D merge qualification reads the particle whose slot equals the cell id
// (SPH-2D/sph_system.cu:1207), before any merge writes: bit 0. When that particle is in the
// list of cell cl itself, bits 1..31 hold its entry + 1, and the walk of cell cl re-reads it
// after merging it; a particle listed in another cell is read here only (its own cell's walk
// may write it, a race in the reference).
__global__ __launch_bounds__(BLK) void k_merge_q2d(DevParams P, DevState S,
                                                   const uint *__restrict__ slot2phys,
                                                   const Counters *cnt,
                                                   const uint *__restrict__ cstart,
                                                   const uint *__restrict__ cend,
                                                   uint *__restrict__ q2d) {
  const uint cl = blockIdx.x * BLK + threadIdx.x;
  if (cl >= P.total_cell) return;
  const uint n_pre = cnt->n_pre, n = cnt->n;
  uint w = 0u;
  if (cl < n) {
    const uint k = cl < n_pre ? slot2phys[cl] : cl;  // children sit at their slot
    w = (S.energy[k] * P.co_energy + S.acc[k].w * P.co_surface < P.merge_criteria) ? 1u : 0u;
    const uint b = cstart[cl];
    if (cl < n_pre && b != SPH_EMPTY && k >= b && k < cend[cl]) w |= (k + 1u) << 1;
  }
  q2d[cl] = w;
}

// A14 rearrange (SPH-3D-2/sph_system.cu:1818-1840) as a renumbering: survivors get the
// rank of their slot among survivors; removed entries are dropped by the next sort.
template <bool B>
__global__ __launch_bounds__(BLK) void k_alive_by_slot(DevState S, const uint *__restrict__ status,
                                                       const SceneCnt *__restrict__ scn,
                                                       const Counters *cnt,
                                                       uint *__restrict__ alive_by_slot) {
  for_each_item<false>(cnt->n, [&](const uint k) {
    const uint w = S.slot[k];
    const uint d = B ? scn[w >> SPH_SCENE_SHIFT].first2 + (w & SPH_LOCAL_MASK) : w;
    alive_by_slot[d] = status[k] == 0u ? 1u : 0u;
  });
}

// Bound = cnt->m, the physical count (a single scene compacts right after merge, where
// m == n; a batch's compaction plan has already set cnt->n). A single scene's thread 0
// also sets the live count to the survivors; no thread here reads cnt->n.
template <bool B>
__global__ __launch_bounds__(BLK) void k_renumber(DevState S, const uint *__restrict__ status,
                                                  const uint *__restrict__ newslot,
                                                  const SceneCnt *__restrict__ scn,
                                                  Counters *cnt) {
  if (!B && blockIdx.x == 0 && threadIdx.x == 0) cnt->n = cnt->scan_total;
  for_each_item<false>(cnt->m, [&](const uint k) {
    const uint w = S.slot[k];
    if (!B) {
      S.slot[k] = status[k] == 0u ? newslot[w] : SPH_DEAD;
    } else {
      const uint sc = w >> SPH_SCENE_SHIFT;
      const SceneCnt &C = scn[sc];
      const uint d = C.first2 + (w & SPH_LOCAL_MASK);
      S.slot[k] = status[k] == 0u ? (sc << SPH_SCENE_SHIFT) | (newslot[d] - C.a0) : SPH_DEAD;
    }
  });
}

__global__ void k_begin_step(Counters *cnt) {
  cnt->n_pre = cnt->n;
  cnt->s_req = 0u;
  cnt->s_done = 0u;
  cnt->groups = 0u;
  cnt->err = 0u;
}

__global__ void k_set_phys_count(Counters *cnt) { cnt->m = cnt->n; }

__global__ void k_finish_compact(Counters *cnt) {
  cnt->m = cnt->n;
  cnt->n = cnt->scan_total;
}

// ---------------------------------------------------------------------------------------
// A15 integrate_dens_kernel + integrate_cell_dens (SPH-3D-2/sph_system.cu:1842-2183):
// voxel v (x fastest) gathers poly6 (particle radius) splats from the 27 cells around
// cell floor(v / n_dens * n_cell); flat accumulation in list order; low faces are zero.
// ---------------------------------------------------------------------------------------
// occ: one bit per cell (set = the cell holds particles); dil: bit h set when any of the
// 27 cells around h may hold particles. A voxel whose cell has no dil bit is 0 after one
// lookup, and the others skip empty neighbour cells before touching the cell tables.
__global__ __launch_bounds__(BLK) void k_density_grid(DevParams P, const float4 *__restrict__ pos,
                                                      const uint *__restrict__ cstart,
                                                      const uint *__restrict__ cend,
                                                      const uint *__restrict__ occ,
                                                      const uint *__restrict__ dil,
                                                      float *__restrict__ dgrid) {
  const uint v = blockIdx.x * BLK + threadIdx.x;
  if (v >= P.tot_dens) return;
  const uint rd = P.nd[0], cd = P.nd[1];
  const uint vx = v % (rd * cd) % rd, vy = v % (rd * cd) / rd, vz = v / (rd * cd);
  float acc = 0.0f;
  uint c[3];
  c[0] = (uint)floorf((float)vx / (float)P.nd[0] * (float)P.ncell[0]);
  c[1] = (uint)floorf((float)vy / (float)P.nd[1] * (float)P.ncell[1]);
  c[2] = (uint)floorf((float)vz / (float)P.nd[2] * (float)P.ncell[2]);
  const uint hc = P.cell_base + (c[2] * P.ncell[1] + c[1]) * P.ncell[0] + c[0];
  if (!(vx == 0 || vy == 0 || vz == 0) && hc < P.total_cell && ((dil[hc >> 5] >> (hc & 31u)) & 1u)) {
    const float ds = P.den_size;
    const float dpx = (float)vx * ds, dpy = (float)vy * ds, dpz = (float)vz * ds;
#pragma unroll
    for (int o = 0; o < 27; o++) {
      uint h;
      if (!neighbor_cell<SPH_V3D2>(P, o, c, h)) continue;
      if (!((occ[h >> 5] >> (h & 31u)) & 1u)) continue;  // empty cell
      const uint s = cstart[h];
      if (s == SPH_EMPTY) continue;
      const uint e = cend[h];
      auto add = [&](const float4 &q) {
        const int L = lvl_index(q.w);
        const float rx = q.x - dpx, ry = q.y - dpy, rz = q.z - dpz;
        const float r2 = rx * rx + ry * ry + rz * rz;
        const float rad2 = L ? P.radius2[1] : P.radius2[0];
        if (r2 < rad2) {
          const float d = rad2 - r2;
          acc += (L ? P.mass[1] : P.mass[0]) * (L ? P.poly6_radius[1] : P.poly6_radius[0]) *
                 ((d * d) * d);
        }
      };
      for (uint j = s; j < e; j += 4) {  // four loads in flight, terms in list order
        float4 q[4];
        sph_group4(pos, j, e, q);
        add(q[0]);
        if (j + 1 < e) add(q[1]);
        if (j + 2 < e) add(q[2]);
        if (j + 3 < e) add(q[3]);
      }
    }
  }
  dgrid[v] = acc;
}

// The same density grid, one wave per cell: when dens_mul = m, the voxels
// (m c + 1 ... m c + m) along each axis are exactly the voxels the reference maps to cell c
// (floor(v / (m n + 1) * n) = c; checked on the host for every voxel before this kernel is
// chosen), so the m^3 voxels of a cell share its 27-cell candidate list. The candidate
// loop is uniform across the wave (one load serves every lane) and each voxel still adds
// its terms in the reference's order. The planes v = 0 are written by k_dgrid_zero_planes.
// Four consecutive cells per block of four waves; mass * poly6 of each level is the one
// product the reference's left-to-right term starts with, so it is formed once.
constexpr int DG_CELLS = BLK / 64;
template <int M>
__global__ __launch_bounds__(BLK) void k_density_grid_cells(DevParams P,
                                                            const float4 *__restrict__ pos,
                                                            const uint *__restrict__ cstart,
                                                            const uint *__restrict__ cend,
                                                            const uint *__restrict__ dil,
                                                            float *__restrict__ dgrid) {
  // the wave's cell, made uniform (scalar) so the candidate loads are scalar loads
  const uint cell = xcd_block(blockIdx.x, gridDim.x) * DG_CELLS +
                    (uint)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  if (cell >= P.total_cell) return;  // whole waves
  const uint lane = threadIdx.x & 63u;
  const uint c[3] = {cell % P.ncell[0], (cell / P.ncell[0]) % P.ncell[1],
                     cell / (P.ncell[0] * P.ncell[1])};
  // dil (optional): bit clear = no particle in any of the cell's 27 neighbours, so every
  // voxel of the cell is +0, the value the sums below would give; no range loads
  const bool empty = dil && !((dil[cell >> 5] >> (cell & 31u)) & 1u);
  const float rad2_0 = P.radius2[0], rad2_1 = P.radius2[1];
  const float mp0 = P.mass[0] * P.poly6_radius[0], mp1 = P.mass[1] * P.poly6_radius[1];
  const bool active = lane < (uint)(M * M * M);
  const uint lx = lane % M, ly = (lane / M) % M, lz = lane / (M * M);
  const uint vx = M * c[0] + 1 + lx, vy = M * c[1] + 1 + ly, vz = M * c[2] + 1 + lz;
  const float ds = P.den_size;
  const float dpx = (float)vx * ds, dpy = (float)vy * ds, dpz = (float)vz * ds;
  float acc = 0.0f;
  auto add = [&](const float4 &q) {
    const int L = lvl_index(q.w);
    const float rx = q.x - dpx, ry = q.y - dpy, rz = q.z - dpz;
    const float r2 = rx * rx + ry * ry + rz * rz;
    const float rad2 = L ? rad2_1 : rad2_0;
    if (r2 < rad2) {
      const float d = rad2 - r2;
      acc += (L ? mp1 : mp0) * ((d * d) * d);
    }
  };
  // the 27 cell ranges are wave-uniform: all are requested up front (scalar loads into
  // SGPRs, one wait), and the fully unrolled cell loop then only waits for particles
  uint rs[27], re[27];
#pragma unroll
  for (int o = 0; o < 27; o++) {
    rs[o] = re[o] = 0u;
    if (!empty) nb_range<SPH_V3D2>(P, o, c, cstart, cend, rs[o], re[o]);
  }
#pragma unroll
  for (int o = 0; o < 27; o++) {
    const uint s0 = rs[o], e0 = re[o];
    uint j = s0;
    for (; j + 4 <= e0; j += 4) {
      const float4 *q = pos + j;  // one 64-byte scalar load for four particles
      const float4 q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
      add(q0);
      add(q1);
      add(q2);
      add(q3);
    }
    for (; j < e0; j++) add(pos[j]);
  }
  if (active) dgrid[(size_t)vz * P.nd[0] * P.nd[1] + (size_t)vy * P.nd[0] + vx] = acc;
}

__global__ __launch_bounds__(BLK) void k_dgrid_zero_planes(DevParams P, float *__restrict__ dgrid) {
  const uint rd = P.nd[0], cd = P.nd[1], ld = P.nd[2];
  const size_t t = (size_t)blockIdx.x * BLK + threadIdx.x;
  const size_t a = (size_t)cd * ld, b = (size_t)rd * ld, c = (size_t)rd * cd;
  if (t < a) {  // x = 0
    const uint y = (uint)(t % cd), z = (uint)(t / cd);
    dgrid[(size_t)z * rd * cd + (size_t)y * rd] = 0.0f;
  } else if (t < a + b) {  // y = 0
    const size_t u = t - a;
    const uint x = (uint)(u % rd), z = (uint)(u / rd);
    dgrid[(size_t)z * rd * cd + x] = 0.0f;
  } else if (t < a + b + c) {  // z = 0
    dgrid[t - a - b] = 0.0f;
  }
}

// A16 integrate_normall (SPH-3D-2/sph_system.cu:2221-2335)
__global__ __launch_bounds__(BLK) void k_normals(DevParams P, const float *__restrict__ dgrid,
                                                 float *__restrict__ normal3) {
  const uint v = blockIdx.x * BLK + threadIdx.x;
  if (v >= P.tot_dens) return;
  const uint rd = P.nd[0], cd = P.nd[1], ld = P.nd[2];
  const uint vx = v % (rd * cd) % rd, vy = v % (rd * cd) / rd, vz = v / (rd * cd);
  const float ds = P.den_size;
  const uint vv[3] = {vx, vy, vz}, nn[3] = {rd, cd, ld}, st[3] = {1u, rd, rd * cd};
  float g[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const float da = vv[a] == 0 ? 0.0f : dgrid[v - st[a]];
    const float db = vv[a] == nn[a] - 1 ? 0.0f : dgrid[v + st[a]];
    g[a] = (da - db) / ds;
  }
  const float nr = sqrtf(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
#pragma unroll
  for (int a = 0; a < 3; a++) normal3[(size_t)v * 3 + a] = nr == 0.0f ? 0.0f : g[a] / nr;
}

// ---------------------------------------------------------------------------------------
// A17 marching cubes (SPH-3D-2/sph_system.cpp:417-542), count / scan / emit. Threads walk
// cubes with x fastest (coalesced voxel reads); counts and triangles are placed in the
// reference's cube order (x outer, z inner).
// ---------------------------------------------------------------------------------------
constexpr int kCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                               {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
constexpr int kEdge[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                              {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
// The same tables as bit fields for the device: corner c is offset along axis a by bit c
// of kCornerBits[a]; edge e joins corners (kEdgeEnd[0] >> 3e) & 7 and (kEdgeEnd[1] >> 3e) & 7.
constexpr uint corner_bits(int a) {
  uint m = 0;
  for (int c = 0; c < 8; c++) m |= (uint)kCorner[c][a] << c;
  return m;
}
constexpr unsigned long long edge_end_bits(int end) {
  unsigned long long v = 0;
  for (int e = 0; e < 12; e++) v |= (unsigned long long)kEdge[e][end] << (3 * e);
  return v;
}
constexpr uint kCornerX = corner_bits(0), kCornerY = corner_bits(1), kCornerZ = corner_bits(2);
constexpr unsigned long long kEdgeEnd0 = edge_end_bits(0), kEdgeEnd1 = edge_end_bits(1);
static_assert(kCornerX == 0x66u && kCornerY == 0xccu && kCornerZ == 0xf0u, "cube corners");

// Cubes are handled in z-segments of MC_SEG cubes: segment (x, y, zc) covers cubes
// (x, y, zc*MC_SEG ...) and segments are numbered in the reference's cube order
// (x outer, then y, then z), so an exclusive scan of per-segment triangle counts gives
// each segment its output offset and triangles keep the reference's order. Threads take
// segments with x fastest, so the voxel reads of a wave are contiguous, and each thread
// slides a 2x2 window of corner values along z (one new plane per cube).
constexpr uint MC_SEG = 16;

struct McGeom {
  uint rd, cd, nx, ny, nz, nzc;
};

__device__ inline McGeom mc_geom(const DevParams &P) {
  McGeom g;
  g.rd = P.nd[0];
  g.cd = P.nd[1];
  g.nx = P.nd[0] - 1;
  g.ny = P.nd[1] - 1;
  g.nz = P.nd[2] - 1;
  g.nzc = (g.nz + MC_SEG - 1) / MC_SEG;
  return g;
}

__device__ inline int mc_tri_count(int f) {
  int tc = 0;
  while (tc < 5 && kMcTri[f][3 * tc] >= 0) tc++;
  return tc;
}

// corner values of plane z at (x, y), (x+1, y), (x+1, y+1), (x, y+1)
__device__ inline void mc_plane(const float *__restrict__ dgrid, const McGeom &g, uint x, uint y,
                                uint z, float v[4]) {
  const size_t b = (size_t)z * g.rd * g.cd + (size_t)y * g.rd + x;
  v[0] = dgrid[b];
  v[1] = dgrid[b + 1];
  v[2] = dgrid[b + 1 + g.rd];
  v[3] = dgrid[b + g.rd];
}

// corner bit c set when the value exceeds ISOVALUE (0.5f, SPH-3D-2/sph_header.h:24);
// corners 0-3 on the lower plane, 4-7 on the upper (kCorner order)
__device__ inline int mc_case_planes(const float lo[4], const float hi[4]) {
  int f = 0;
#pragma unroll
  for (int c = 0; c < 4; c++) {
    if (lo[c] > 0.5f) f |= 1 << c;
    if (hi[c] > 0.5f) f |= 1 << (c + 4);
  }
  return f;
}

__global__ __launch_bounds__(BLK) void k_mc_count(DevParams P, const float *__restrict__ dgrid,
                                                  uint *__restrict__ seg_tri,
                                                  uint *__restrict__ seg_cube) {
  const McGeom g = mc_geom(P);
  const size_t t = (size_t)blockIdx.x * BLK + threadIdx.x;
  if (t >= (size_t)g.nx * g.ny * g.nzc) return;
  const uint x = (uint)(t % g.nx), y = (uint)((t / g.nx) % g.ny), zc = (uint)(t / ((size_t)g.nx * g.ny));
  const uint z0 = zc * MC_SEG, z1 = min(g.nz, z0 + MC_SEG);
  float lo[4], hi[4];
  mc_plane(dgrid, g, x, y, z0, lo);
  uint tri = 0, cubes = 0;
  for (uint z = z0; z < z1; z++) {
    mc_plane(dgrid, g, x, y, z + 1, hi);
    const uint tc = (uint)mc_tri_count(mc_case_planes(lo, hi));
    tri += tc;
    cubes += tc ? 1u : 0u;
#pragma unroll
    for (int c = 0; c < 4; c++) lo[c] = hi[c];
  }
  const size_t sid = ((size_t)x * g.ny + y) * g.nzc + zc;
  seg_tri[sid] = tri;
  seg_cube[sid] = cubes;
}

// Non-empty cubes in the reference's order: (cube id, first triangle index); at most
// `cap` entries (a cube at list index >= cap has its triangles at index >= cap).
__global__ __launch_bounds__(BLK) void k_mc_list(DevParams P, const float *__restrict__ dgrid,
                                                 const uint *__restrict__ seg_cube,
                                                 const uint *__restrict__ seg_toff,
                                                 const uint *__restrict__ seg_coff, uint cap,
                                                 uint2 *__restrict__ clist) {
  const McGeom g = mc_geom(P);
  const size_t t = (size_t)blockIdx.x * BLK + threadIdx.x;
  if (t >= (size_t)g.nx * g.ny * g.nzc) return;
  const uint x = (uint)(t % g.nx), y = (uint)((t / g.nx) % g.ny), zc = (uint)(t / ((size_t)g.nx * g.ny));
  const size_t sid = ((size_t)x * g.ny + y) * g.nzc + zc;
  if (seg_cube[sid] == 0) return;
  uint idx = seg_coff[sid], toff = seg_toff[sid];
  const uint z0 = zc * MC_SEG, z1 = min(g.nz, z0 + MC_SEG);
  float lo[4], hi[4];
  mc_plane(dgrid, g, x, y, z0, lo);
  for (uint z = z0; z < z1 && idx < cap; z++) {
    mc_plane(dgrid, g, x, y, z + 1, hi);
    const uint tc = (uint)mc_tri_count(mc_case_planes(lo, hi));
#pragma unroll
    for (int c = 0; c < 4; c++) lo[c] = hi[c];
    if (tc == 0) continue;
    clist[idx] = make_uint2((uint)(((size_t)x * g.ny + y) * g.nz + z), toff);
    idx++;
    toff += tc;
  }
}

// One thread per non-empty cube: its triangles go straight to their place in the output
// (the cube's first triangle index comes from the scan) as 8-byte stores; the table row
// of the cube's case is read once as 16 bytes. Corner positions and normals are formed
// from the edge's two corners with the same operations as (c0 + c1) / 2 per component.
__device__ inline int mc_entry(const uint4 &row, int k) {
  const uint w = k < 4 ? row.x : (k < 8 ? row.y : (k < 12 ? row.z : row.w));
  return (int)(signed char)((w >> (8 * (k & 3))) & 0xffu);
}

__global__ __launch_bounds__(BLK) void k_mc_emit(DevParams P, const float *__restrict__ dgrid,
                                                 const float *__restrict__ normal3,
                                                 const uint2 *__restrict__ clist,
                                                 const Counters *cnt, uint max_tri,
                                                 float *__restrict__ tris) {
  const McGeom g = mc_geom(P);
  const uint ncube = min(cnt->nmc, max_tri);
  const float ds = P.den_size;
  const size_t plane = (size_t)g.rd * g.cd;
  for (uint k = blockIdx.x * BLK + threadIdx.x; k < ncube; k += gridDim.x * BLK) {
    const uint2 ent = clist[k];
    const uint cid = ent.x, toff = ent.y;
    const uint z = cid % g.nz, y = (cid / g.nz) % g.ny, x = cid / (g.nz * g.ny);
    float lo[4], hi[4];
    mc_plane(dgrid, g, x, y, z, lo);
    mc_plane(dgrid, g, x, y, z + 1, hi);
    const int f = mc_case_planes(lo, hi);
    const uint4 row = *reinterpret_cast<const uint4 *>(kMcTri[f]);
#pragma unroll
    for (int tr = 0; tr < 5; tr++) {
      if (mc_entry(row, 3 * tr) < 0 || toff + (uint)tr >= max_tri) break;
      float v[18];
#pragma unroll
      for (int vtx = 0; vtx < 3; vtx++) {
        const int e = mc_entry(row, 3 * tr + vtx);
        const uint a0 = (uint)(kEdgeEnd0 >> (3 * e)) & 7u, a1 = (uint)(kEdgeEnd1 >> (3 * e)) & 7u;
        const uint X0 = x + ((kCornerX >> a0) & 1u), Y0 = y + ((kCornerY >> a0) & 1u),
                   Z0 = z + ((kCornerZ >> a0) & 1u);
        const uint X1 = x + ((kCornerX >> a1) & 1u), Y1 = y + ((kCornerY >> a1) & 1u),
                   Z1 = z + ((kCornerZ >> a1) & 1u);
        v[vtx * 3 + 0] = ((float)X0 * ds + (float)X1 * ds) / 2;
        v[vtx * 3 + 1] = ((float)Y0 * ds + (float)Y1 * ds) / 2;
        v[vtx * 3 + 2] = ((float)Z0 * ds + (float)Z1 * ds) / 2;
        const size_t b0 = ((size_t)Z0 * plane + (size_t)Y0 * g.rd + X0) * 3;
        const size_t b1 = ((size_t)Z1 * plane + (size_t)Y1 * g.rd + X1) * 3;
#pragma unroll
        for (int a = 0; a < 3; a++) v[9 + vtx * 3 + a] = (normal3[b0 + a] + normal3[b1 + a]) / 2;
      }
      float2 *dst = reinterpret_cast<float2 *>(tris + (size_t)(toff + (uint)tr) * 18);
#pragma unroll
      for (int q = 0; q < 9; q++) dst[q] = make_float2(v[2 * q], v[2 * q + 1]);
    }
  }
}

// ---------------------------------------------------------------------------------------
// AoS <-> SoA (copy_array boundary)
// ---------------------------------------------------------------------------------------
template <int V>
__global__ __launch_bounds__(BLK) void k_upload(const void *__restrict__ aos, uint n, DevState S,
                                                uint slot_or) {
  const uint i = blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  if (V == SPH_V2D) {
    const SphParticle2D q = ((const SphParticle2D *)aos)[i];
    S.pos[i] = make_float4(q.pos.x, q.pos.y, 0.0f, __uint_as_float(q.level));
    S.vel[i] = make_float4(q.vel.x, q.vel.y, 0.0f, 0.0f);
    S.ev[i] = make_float4(q.ev.x, q.ev.y, 0.0f, 0.0f);
    S.acc[i] = make_float4(q.acc.x, q.acc.y, 0.0f, q.surface_normal);
    S.dp[i] = make_float2(q.dens, q.pres);
    S.energy[i] = q.energy;
  } else {
    const SphParticle3D q = ((const SphParticle3D *)aos)[i];
    S.pos[i] = make_float4(q.pos.x, q.pos.y, q.pos.z, __uint_as_float(q.level));
    S.vel[i] = make_float4(q.vel.x, q.vel.y, q.vel.z, 0.0f);
    S.ev[i] = make_float4(q.ev.x, q.ev.y, q.ev.z, q.tension);
    S.acc[i] = make_float4(q.acc.x, q.acc.y, q.acc.z, q.surface);
    S.dp[i] = make_float2(q.dens, q.press);
    S.energy[i] = q.energy;
  }
  S.slot[i] = slot_or | i;
}

template <int V>
__global__ __launch_bounds__(BLK) void k_download(DevState S, const uint *__restrict__ slot2phys,
                                                  uint n, void *__restrict__ aos) {
  const uint s = blockIdx.x * BLK + threadIdx.x;
  if (s >= n) return;
  const uint k = slot2phys[s];
  const float4 p = S.pos[k], v = S.vel[k], e = S.ev[k], a = S.acc[k];
  const float2 d = S.dp[k];
  if (V == SPH_V2D) {
    SphParticle2D q;
    q.pos.x = p.x; q.pos.y = p.y;
    q.vel.x = v.x; q.vel.y = v.y;
    q.ev.x = e.x; q.ev.y = e.y;
    q.acc.x = a.x; q.acc.y = a.y;
    q.dens = d.x; q.pres = d.y; q.energy = S.energy[k];
    q.level = __float_as_uint(p.w);
    q.surface_normal = a.w;
    ((SphParticle2D *)aos)[s] = q;
  } else {
    SphParticle3D q;
    q.pos.x = p.x; q.pos.y = p.y; q.pos.z = p.z;
    q.vel.x = v.x; q.vel.y = v.y; q.vel.z = v.z;
    q.ev.x = e.x; q.ev.y = e.y; q.ev.z = e.z;
    q.acc.x = a.x; q.acc.y = a.y; q.acc.z = a.z;
    q.dens = d.x; q.press = d.y; q.energy = S.energy[k];
    q.surface = a.w; q.tension = e.w;
    q.level = __float_as_uint(p.w);
    ((SphParticle3D *)aos)[s] = q;
  }
}

__global__ __launch_bounds__(BLK) void k_slot2phys(DevState S, uint m, uint *__restrict__ s2p) {
  const uint k = blockIdx.x * BLK + threadIdx.x;
  if (k >= m) return;
  const uint s = S.slot[k];
  if (s != SPH_DEAD) s2p[s] = k;
}


// ---------------------------------------------------------------------------------------
// Batches of scenes. Each scene keeps the reference's semantics on its own: its particles
// see only the cells of their scene (global cell = cell_base + hash), its slots are
// numbered from 0, and split capacity and merge are decided per scene. The sort input
// packs the scenes' slots densely, scene after scene, so one sort orders every scene at
// once and equal keys keep ascending slot within each scene.
// ---------------------------------------------------------------------------------------
constexpr int SCN_BLK = 256;  // one thread per scene (at most SPH_MAX_SCENES)
static_assert(SPH_MAX_SCENES <= SCN_BLK, "one block holds every scene");

// Exclusive scan of one value per thread over the block; total -> *total.
__device__ inline uint block_exscan(uint v, uint *sh, uint &total) {
  const uint t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (uint off = 1; off < SCN_BLK; off <<= 1) {
    const uint x = t >= off ? sh[t - off] : 0u;
    __syncthreads();
    sh[t] += x;
    __syncthreads();
  }
  total = sh[SCN_BLK - 1];
  const uint r = sh[t] - v;
  __syncthreads();
  return r;
}

// Start of a step: dense slot starts of the scenes, per-step counters cleared.
__global__ __launch_bounds__(SCN_BLK) void k_batch_begin(SceneCnt *scn, uint ns, Counters *cnt) {
  __shared__ uint sh[SCN_BLK];
  const uint t = threadIdx.x;
  const uint n = t < ns ? scn[t].n : 0u;
  uint tot;
  const uint first = block_exscan(n, sh, tot);
  if (t < ns) {
    SceneCnt &C = scn[t];
    C.n_pre = n;
    C.first = first;
    C.first2 = first;
    C.r0 = C.s_req = C.s_done = C.done_base = C.a0 = C.groups = C.err = 0u;
  }
  if (t == 0) {
    cnt->n_pre = cnt->n;
    cnt->s_req = cnt->s_done = cnt->groups = cnt->err = 0u;
  }
}

// After the split-flag scan: per scene, parents flagged, parents split under the scene's
// capacity (the first floor((cap - n)/(n_split - 1)) in slot order), where its children
// go, and its dense slot start after split. Without split only the starts are set.
template <int NS>
__global__ __launch_bounds__(SCN_BLK) void k_batch_split_plan(const DevParams *__restrict__ Ptab,
                                                              SceneCnt *scn, uint ns,
                                                              const uint *__restrict__ rank,
                                                              const uint *m_dev, uint m_host,
                                                              int do_split, Counters *cnt) {
  __shared__ uint sh[SCN_BLK];
  const uint t = threadIdx.x;
  const uint m = dev_count(m_dev, m_host);
  uint n = 0u, req = 0u, done = 0u, r0 = 0u;
  if (t < ns) {
    n = scn[t].n_pre;
    if (do_split) {
      const uint total = cnt->scan_total;
      const uint f = scn[t].first, e = f + n;
      r0 = f < m ? rank[f] : total;
      req = (e < m ? rank[e] : total) - r0;
      const uint room = (Ptab[t].cap - n) / (uint)(NS - 1);
      done = req < room ? req : room;
    }
  }
  uint tot_done, tot_n2, tot_req;
  const uint done_base = block_exscan(done, sh, tot_done);
  const uint n2 = n + done * (uint)(NS - 1);
  const uint first2 = block_exscan(n2, sh, tot_n2);
  (void)block_exscan(req, sh, tot_req);
  if (t < ns) {
    SceneCnt &C = scn[t];
    C.r0 = r0;
    C.s_req = req;
    C.s_done = done;
    C.done_base = done_base;
    C.n = n2;
    C.first2 = first2;
  }
  if (t == 0) {
    cnt->s_req = tot_req;
    cnt->s_done = tot_done;
    cnt->n = tot_n2;
    cnt->m = tot_n2;
  }
}

// After the survivor scan: each scene's survivors and the scan value its renumbering
// subtracts. cnt->m (entries, removed ones included) stays for the renumber bound.
__global__ __launch_bounds__(SCN_BLK) void k_batch_compact_plan(SceneCnt *scn, uint ns,
                                                                const uint *__restrict__ newslot,
                                                                Counters *cnt) {
  const uint t = threadIdx.x;
  const uint D = cnt->m;
  const uint total = cnt->scan_total;
  if (t < ns) {
    SceneCnt &C = scn[t];
    const uint f = C.first2, e = f + C.n;
    const uint a0 = f < D ? newslot[f] : total;
    const uint a1 = e < D ? newslot[e] : total;
    C.a0 = a0;
    C.n = a1 - a0;
  }
  __syncthreads();
  if (t == 0) cnt->n = total;
}

__global__ __launch_bounds__(BLK) void k_scene_kill(uint *__restrict__ slot, uint m, uint scene) {
  const uint k = blockIdx.x * BLK + threadIdx.x;
  if (k >= m) return;
  const uint w = slot[k];
  if (w != SPH_DEAD && (w >> SPH_SCENE_SHIFT) == scene) slot[k] = SPH_DEAD;
}

__global__ __launch_bounds__(BLK) void k_scene_slot2phys(const uint *__restrict__ slot, uint m,
                                                         uint scene, uint *__restrict__ s2p) {
  const uint k = blockIdx.x * BLK + threadIdx.x;
  if (k >= m) return;
  const uint w = slot[k];
  if (w != SPH_DEAD && (w >> SPH_SCENE_SHIFT) == scene) s2p[w & SPH_LOCAL_MASK] = k;
}

__global__ __launch_bounds__(BLK) void k_live_flags(const uint *__restrict__ slot, uint m,
                                                    uint *__restrict__ flags) {
  const uint k = blockIdx.x * BLK + threadIdx.x;
  if (k >= m) return;
  flags[k] = slot[k] != SPH_DEAD ? 1u : 0u;
}

__global__ __launch_bounds__(BLK) void k_pack_scatter(DevState src, DevState dst, uint m,
                                                      const uint *__restrict__ flags,
                                                      const uint *__restrict__ pos) {
  const uint k = blockIdx.x * BLK + threadIdx.x;
  if (k >= m || !flags[k]) return;
  const uint q = pos[k];
  dst.pos[q] = src.pos[k];
  dst.vel[q] = src.vel[k];
  dst.ev[q] = src.ev[k];
  dst.acc[q] = src.acc[k];
  dst.dp[q] = src.dp[k];
  dst.energy[q] = src.energy[k];
  dst.slot[q] = src.slot[k];
}

inline dim3 grid_for(size_t n) { return dim3((uint)((n + BLK - 1) / BLK)); }
// a persistent grid when given, else one workgroup per block of n items
inline dim3 grid_or(uint grid, size_t n) { return grid ? dim3(grid) : grid_for(n); }

#define SPH_DISPATCH(variant, KERNEL, ...)                                                   \
  do {                                                                                      \
    if ((variant) == SPH_V2D) hipLaunchKernelGGL(KERNEL<SPH_V2D>, __VA_ARGS__);            \
    else if ((variant) == SPH_V3D2) hipLaunchKernelGGL(KERNEL<SPH_V3D2>, __VA_ARGS__);     \
    else hipLaunchKernelGGL(KERNEL<SPH_V3D4>, __VA_ARGS__);                                 \
  } while (0)

}  // namespace

// KERNEL<variant, batched> for the 3D variants and, single-scene only, 2D
#define SPH_DISPATCH_VB(variant, batched, KERNEL, grid, block, shmem, st, ...)                  \
  do {                                                                                       \
    if (batched) {                                                                           \
      if ((variant) == SPH_V3D2) KERNEL<SPH_V3D2, true><<<grid, block, shmem, st>>>(__VA_ARGS__); \
      else if ((variant) == SPH_V3D4) KERNEL<SPH_V3D4, true><<<grid, block, shmem, st>>>(__VA_ARGS__); \
    } else {                                                                                 \
      if ((variant) == SPH_V2D) KERNEL<SPH_V2D, false><<<grid, block, shmem, st>>>(__VA_ARGS__); \
      else if ((variant) == SPH_V3D2) KERNEL<SPH_V3D2, false><<<grid, block, shmem, st>>>(__VA_ARGS__); \
      else KERNEL<SPH_V3D4, false><<<grid, block, shmem, st>>>(__VA_ARGS__);                \
    }                                                                                        \
  } while (0)

// KERNEL<variant, batched, L> (L: a pair-list length)
#define SPH_DISPATCH_VBL(variant, batched, KERNEL, L, grid, block, shmem, st, ...)             \
  do {                                                                                       \
    if (batched) {                                                                           \
      if ((variant) == SPH_V3D2) KERNEL<SPH_V3D2, true, L><<<grid, block, shmem, st>>>(__VA_ARGS__); \
      else if ((variant) == SPH_V3D4) KERNEL<SPH_V3D4, true, L><<<grid, block, shmem, st>>>(__VA_ARGS__); \
    } else {                                                                                 \
      if ((variant) == SPH_V2D) KERNEL<SPH_V2D, false, L><<<grid, block, shmem, st>>>(__VA_ARGS__); \
      else if ((variant) == SPH_V3D2) KERNEL<SPH_V3D2, false, L><<<grid, block, shmem, st>>>(__VA_ARGS__); \
      else KERNEL<SPH_V3D4, false, L><<<grid, block, shmem, st>>>(__VA_ARGS__);             \
    }                                                                                        \
  } while (0)

namespace {
inline const DevParams *ptab_of(const BatchView *bv) { return bv ? bv->Ptab : nullptr; }
inline SceneCnt *scn_of(const BatchView *bv) { return bv ? bv->scn : nullptr; }
}  // namespace

void launch_hash(int variant, const DevParams &P, DevState s, const uint *cnt_m, uint m_max,
                 uint *keys, uint *slot2phys, Counters *cnt, hipStream_t st, const BatchView *bv,
                 uint grid) {
  if (m_max == 0) return;
  SPH_DISPATCH_VB(variant, bv != nullptr, k_hash, grid_or(grid, m_max), dim3(BLK), 0, st, P, ptab_of(bv),
                  scn_of(bv), s.pos, s.slot, cnt_m, m_max, keys, slot2phys, cnt);
}

void launch_cell_bounds(const uint *skeys, uint n, uint ncell, uint *cstart, uint *cend,
                        hipStream_t st, const uint *n_dev, uint grid) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_cell_bounds, grid_or(grid, n), dim3(BLK), 0, st, skeys, n_dev, n, ncell,
                     cstart, cend);
}

void launch_reorder(DevState src, DevState dst, const uint *svals, const uint *slot2phys, uint n,
                    uint *slot2phys_new, hipStream_t st, Counters *set_m, bool all_fields,
                    const uint *n_dev, uint grid) {
  if (n == 0) {
    if (set_m) hipLaunchKernelGGL(k_set_phys_count, dim3(1), dim3(1), 0, st, set_m);
    return;
  }
  hipLaunchKernelGGL(k_reorder, grid_or(grid, n), dim3(BLK), 0, st, src, dst, svals, slot2phys,
                     n_dev, n, slot2phys_new, set_m, all_fields);
}

void launch_density(int variant, const DevParams &P, DevState s, Aux a, const uint *cstart,
                    const uint *cend, uint n, hipStream_t st, const BatchView *bv,
                    const uint *n_dev) {
  if (n == 0) return;
  SPH_DISPATCH_VB(variant, bv != nullptr, k_density, grid_for(n), dim3(BLK), 0, st, P,
                  ptab_of(bv), s.slot, s.pos, s.ev, s.dp, a.frec, cstart, cend, n_dev, n, a.nbl,
                  a.nbl_cnt, a.nbl_stride);
}

template <int L>
static void launch_force_l(int variant, const DevParams &P, DevState s, Aux a, const uint *cstart,
                           const uint *cend, uint n, hipStream_t st, const BatchView *bv,
                           const uint *n_dev, bool use_nbl) {
  SPH_DISPATCH_VBL(variant, bv != nullptr, k_force, L, grid_for(n), dim3(BLK), 0, st, P,
                   ptab_of(bv), s.slot, s.pos, s.ev, s.dp, a.frec, s.acc, cstart, cend, n_dev, n,
                   use_nbl ? a.nbl : nullptr, a.nbl_cnt, a.nbl_stride);
}

void launch_force(int variant, const DevParams &P, DevState s, Aux a, const uint *cstart,
                  const uint *cend, uint n, hipStream_t st, const BatchView *bv, int cfg,
                  const uint *n_dev, bool use_nbl) {
  if (n == 0) return;
  static_assert(kForceLists[0] == 24 && kForceLists[1] == 32 && kForceLists[2] == 48, "force lists");
  use_nbl = use_nbl && a.nbl != nullptr;
  if (cfg == 0) launch_force_l<24>(variant, P, s, a, cstart, cend, n, st, bv, n_dev, use_nbl);
  else if (cfg == 2) launch_force_l<48>(variant, P, s, a, cstart, cend, n, st, bv, n_dev, use_nbl);
  else launch_force_l<32>(variant, P, s, a, cstart, cend, n, st, bv, n_dev, use_nbl);
}

void launch_integrate(int variant, const DevParams &P, DevState s, uint n, hipStream_t st,
                      const BatchView *bv, float4 *erec, const uint *n_dev, uint grid) {
  if (n == 0) return;
  SPH_DISPATCH_VB(variant, bv != nullptr, k_integrate, grid_or(grid, n), dim3(BLK), 0, st, P,
                  ptab_of(bv), s, n_dev, n, erec);
}

template <int L, bool U>
static void launch_energy_lu(int variant, const DevParams &P, DevState s, const uint *cstart,
                             const uint *cend, uint n, hipStream_t st, const BatchView *bv,
                             uint *flag_by_slot, const float4 *erec, const uint *n_dev) {
#define SPH_EN(VV, BB)                                                                       \
  k_energy<VV, BB, L, U><<<grid_for(n), dim3(BLK), 0, st>>>(                                 \
      P, ptab_of(bv), scn_of(bv), s.slot, s.pos, s.vel, s.acc, s.energy, cstart, cend, n_dev, \
      n, flag_by_slot, erec)
  if (bv) {
    if (variant == SPH_V3D2) SPH_EN(SPH_V3D2, true);
    else if (variant == SPH_V3D4) SPH_EN(SPH_V3D4, true);
  } else {
    if (variant == SPH_V2D) SPH_EN(SPH_V2D, false);
    else if (variant == SPH_V3D2) SPH_EN(SPH_V3D2, false);
    else SPH_EN(SPH_V3D4, false);
  }
#undef SPH_EN
}

template <int L>
static void launch_energy_l(int variant, const DevParams &P, DevState s, const uint *cstart,
                            const uint *cend, uint n, hipStream_t st, const BatchView *bv,
                            uint *flag_by_slot, const float4 *erec, const uint *n_dev) {
  if (P.energy_unit)
    launch_energy_lu<L, true>(variant, P, s, cstart, cend, n, st, bv, flag_by_slot, erec, n_dev);
  else
    launch_energy_lu<L, false>(variant, P, s, cstart, cend, n, st, bv, flag_by_slot, erec, n_dev);
}

// erec (pos.x, pos.y, vel.x, vel.y) from the state, for an energy pass that does not follow
// the integrate kernel (which writes the same records)
__global__ __launch_bounds__(BLK) void k_make_erec(const float4 *__restrict__ pos,
                                                   const float4 *__restrict__ vel,
                                                   const uint *n_dev, uint n_host,
                                                   float4 *__restrict__ erec) {
  for_each_item<false>(dev_count(n_dev, n_host), [&](const uint i) {
    const float4 p = pos[i], v = vel[i];
    erec[i] = make_float4(p.x, p.y, v.x, v.y);
  });
}

void launch_make_erec(DevState s, uint n, float4 *erec, hipStream_t st, const uint *n_dev) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_make_erec, grid_for(n), dim3(BLK), 0, st, s.pos, s.vel, n_dev, n, erec);
}

void launch_energy(int variant, const DevParams &P, DevState s, const uint *cstart,
                   const uint *cend, uint n, hipStream_t st, const BatchView *bv, int cfg,
                   uint *flag_by_slot, const float4 *erec, const uint *n_dev) {
  if (n == 0) return;
  static_assert(kEnergyLists[0] == 16 && kEnergyLists[1] == 24 && kEnergyLists[2] == 32, "energy lists");
  if (cfg == 1)
    launch_energy_l<24>(variant, P, s, cstart, cend, n, st, bv, flag_by_slot, erec, n_dev);
  else if (cfg == 2)
    launch_energy_l<32>(variant, P, s, cstart, cend, n, st, bv, flag_by_slot, erec, n_dev);
  else
    launch_energy_l<16>(variant, P, s, cstart, cend, n, st, bv, flag_by_slot, erec, n_dev);
}

void launch_decide(int variant, const DevParams &P, DevState s, uint n, uint *flag_by_slot,
                   hipStream_t st, const BatchView *bv) {
  (void)variant;
  if (n == 0) return;
  if (bv)
    k_decide<true><<<grid_for(n), dim3(BLK), 0, st>>>(P, bv->Ptab, bv->scn, s, n, flag_by_slot);
  else
    k_decide<false><<<grid_for(n), dim3(BLK), 0, st>>>(P, nullptr, nullptr, s, n, flag_by_slot);
}

void launch_split(int variant, const DevParams &P, DevState s, uint n, const uint *flag_by_slot,
                  const uint *rank_by_slot, Counters *cnt, uint *status, hipStream_t st,
                  const BatchView *bv, const uint *n_dev, uint grid) {
  if (n == 0) {  // nothing to split; a single scene still applies the counter update
    if (!bv) SPH_DISPATCH(variant, k_finish_split, dim3(1), dim3(1), 0, st, P, cnt);
    return;
  }
  SPH_DISPATCH_VB(variant, bv != nullptr, k_split, grid_or(grid, n), dim3(BLK), 0, st, P,
                  ptab_of(bv), scn_of(bv), s, n_dev, n, flag_by_slot, rank_by_slot, cnt, status);
}

void launch_merge_q2d(const DevParams &P, DevState s, const uint *slot2phys, const Counters *cnt,
                      const uint *cstart, const uint *cend, uint *q2d, hipStream_t st) {
  hipLaunchKernelGGL(k_merge_q2d, grid_for(P.total_cell), dim3(BLK), 0, st, P, s, slot2phys, cnt,
                     cstart, cend, q2d);
}

void launch_merge(int variant, const DevParams &P, DevState s, const uint *cstart,
                  const uint *cend, const uint *q2d, uint *status, Counters *cnt, hipStream_t st,
                  const BatchView *bv, const uint *skeys, uint nkeys, const uint *nkeys_dev,
                  uint *group_out, uint *qbits, size_t qbits_len) {
  const uint ncell = bv ? bv->total_cell : P.total_cell;
  // a sparse grid (the 512^3-cell collision: 4M keys, 134M cells) runs by key; on a
  // denser one the per-cell threads keep more lanes of a wave busy
  const bool by_entry = skeys && (size_t)nkeys * 4 < ncell;
  // the qualification pass needs the entry count (the sorted keys) and two words per 64
  // entries of scratch
  if (!(qbits && nkeys && ((size_t)nkeys + 63) / 64 * 2 <= qbits_len)) qbits = nullptr;
  if (qbits)
    SPH_DISPATCH_VB(variant, bv != nullptr, k_merge_qual, grid_for(nkeys), dim3(BLK), 0, st, P,
                    ptab_of(bv), s, nkeys_dev, nkeys, qbits, status);
  SPH_DISPATCH_VB(variant, bv != nullptr, k_merge, grid_for(by_entry ? nkeys : ncell),
                  dim3(BLK), 0, st, P, ptab_of(bv), scn_of(bv), ncell, s, cstart, cend, q2d,
                  status, cnt, by_entry ? skeys : nullptr, nkeys_dev, nkeys, group_out, qbits);
}

void launch_alive_by_slot(DevState s, const uint *status, const Counters *cnt, uint m_max,
                          uint *alive_by_slot, hipStream_t st, const BatchView *bv, uint grid) {
  if (m_max == 0) return;
  if (bv)
    k_alive_by_slot<true><<<grid_or(grid, m_max), dim3(BLK), 0, st>>>(s, status, bv->scn, cnt,
                                                                      alive_by_slot);
  else
    k_alive_by_slot<false><<<grid_or(grid, m_max), dim3(BLK), 0, st>>>(s, status, nullptr, cnt,
                                                                       alive_by_slot);
}

void launch_renumber(DevState s, const uint *status, const uint *newslot, Counters *cnt,
                     uint m_max, hipStream_t st, const BatchView *bv, uint grid) {
  if (m_max == 0) {
    if (!bv) hipLaunchKernelGGL(k_finish_compact, dim3(1), dim3(1), 0, st, cnt);
    return;
  }
  if (bv)
    k_renumber<true><<<grid_or(grid, m_max), dim3(BLK), 0, st>>>(s, status, newslot, bv->scn, cnt);
  else
    k_renumber<false><<<grid_or(grid, m_max), dim3(BLK), 0, st>>>(s, status, newslot, nullptr, cnt);
}

void launch_begin_step(Counters *cnt, hipStream_t st) {
  hipLaunchKernelGGL(k_begin_step, dim3(1), dim3(1), 0, st, cnt);
}

void launch_batch_begin(const BatchView &bv, Counters *cnt, hipStream_t st) {
  k_batch_begin<<<1, SCN_BLK, 0, st>>>(bv.scn, bv.n_scenes, cnt);
}

void launch_batch_split_plan(int variant, const BatchView &bv, const uint *rank_by_slot, uint m,
                             bool do_split, Counters *cnt, hipStream_t st, const uint *m_dev) {
  if (variant == SPH_V3D4)
    k_batch_split_plan<4><<<1, SCN_BLK, 0, st>>>(bv.Ptab, bv.scn, bv.n_scenes, rank_by_slot, m_dev,
                                                 m, do_split ? 1 : 0, cnt);
  else
    k_batch_split_plan<2><<<1, SCN_BLK, 0, st>>>(bv.Ptab, bv.scn, bv.n_scenes, rank_by_slot, m_dev,
                                                 m, do_split ? 1 : 0, cnt);
}

void launch_batch_compact_plan(const BatchView &bv, const uint *newslot, Counters *cnt,
                               hipStream_t st) {
  k_batch_compact_plan<<<1, SCN_BLK, 0, st>>>(bv.scn, bv.n_scenes, newslot, cnt);
}

void launch_scene_kill(DevState s, uint m, uint scene, hipStream_t st) {
  if (m == 0) return;
  k_scene_kill<<<grid_for(m), dim3(BLK), 0, st>>>(s.slot, m, scene);
}

void launch_scene_slot2phys(DevState s, uint m, uint scene, uint *slot2phys, hipStream_t st) {
  if (m == 0) return;
  k_scene_slot2phys<<<grid_for(m), dim3(BLK), 0, st>>>(s.slot, m, scene, slot2phys);
}

void launch_pack_live(DevState src, DevState dst, uint m, uint *flags, uint *pos, uint *tmp,
                      uint *total_dev, hipStream_t st) {
  if (m == 0) return;
  k_live_flags<<<grid_for(m), dim3(BLK), 0, st>>>(src.slot, m, flags);
  sph_exclusive_scan(flags, pos, nullptr, m, total_dev, tmp, st);
  k_pack_scatter<<<grid_for(m), dim3(BLK), 0, st>>>(src, dst, m, flags, pos);
}

// bit h of occ = cell h holds particles: one coalesced pass over cstart, a wave's 64 cells
// gathered by ballot into two words (occ holds 2 * ceil(ncell / 64) words)
__global__ __launch_bounds__(BLK) void k_occ_bits(const uint *__restrict__ cstart, uint ncell,
                                                  uint *__restrict__ occ) {
  const uint t = blockIdx.x * BLK + threadIdx.x;
  const uint base = t & ~63u;
  if (base >= ncell) return;  // whole waves leave together
  const bool full = t < ncell && cstart[t] != SPH_EMPTY;
  const unsigned long long b = __ballot(full);
  const uint lane = threadIdx.x & 63u;
  if (lane == 0) occ[base >> 5] = (uint)b;
  else if (lane == 32) occ[(base >> 5) + 1] = (uint)(b >> 32);
}

// bits b .. b+31 of the bit array occ (nw words), 0 outside it
__device__ inline uint bit_window(const uint *__restrict__ occ, long long b, long long nw) {
  const long long w = b >> 5;  // floor, also for b < 0
  const uint sh = (uint)(b & 31);
  const uint lo = (w >= 0 && w < nw) ? occ[w] : 0u;
  if (sh == 0) return lo;
  const uint hi = (w + 1 >= 0 && w + 1 < nw) ? occ[w + 1] : 0u;
  return (lo >> sh) | (hi << (32u - sh));
}

// dil word w = OR of occ over the 27 offsets dx + dy*nx + dz*nx*ny of every bit of word w.
// Offsets that wrap to the next row or plane only add bits, so a clear bit of dil still
// proves an empty neighbourhood.
__global__ __launch_bounds__(BLK) void k_occ_dilate(const uint *__restrict__ occ, uint nw,
                                                    uint nx, uint nxy, uint *__restrict__ dil) {
  const uint w = blockIdx.x * BLK + threadIdx.x;
  if (w >= nw) return;
  uint r = 0;
  for (int dz = -1; dz <= 1; dz++)
    for (int dy = -1; dy <= 1; dy++) {
      const long long b = (long long)w * 32 + (long long)dz * nxy + (long long)dy * nx;
      r |= bit_window(occ, b - 1, nw) | bit_window(occ, b, nw) | bit_window(occ, b + 1, nw);
    }
  dil[w] = r;
}

void launch_density_grid(const DevParams &P, DevState s, const uint *cstart, const uint *cend,
                         float *dgrid, bool cell_blocks, hipStream_t st, uint *occ) {
  const uint nw = (uint)sph_occ_words(P.total_cell);
  uint *dil = occ ? occ + nw : nullptr;
  if (occ) {  // occupancy bits and their 27-cell dilation
    hipLaunchKernelGGL(k_occ_bits, grid_for(P.total_cell), dim3(BLK), 0, st, cstart, P.total_cell, occ);
    hipLaunchKernelGGL(k_occ_dilate, grid_for(nw), dim3(BLK), 0, st, occ, nw, P.ncell[0],
                       P.ncell[0] * P.ncell[1], dil);
  }
  if (cell_blocks && P.dens_mul == 4) {
    const size_t planes = (size_t)P.nd[1] * P.nd[2] + (size_t)P.nd[0] * P.nd[2] + (size_t)P.nd[0] * P.nd[1];
    hipLaunchKernelGGL(k_dgrid_zero_planes, grid_for(planes), dim3(BLK), 0, st, P, dgrid);
    hipLaunchKernelGGL(k_density_grid_cells<4>, dim3((P.total_cell + DG_CELLS - 1) / DG_CELLS),
                       dim3(BLK), 0, st, P, s.pos, cstart, cend, dil, dgrid);
    return;
  }
  hipLaunchKernelGGL(k_density_grid, grid_for(P.tot_dens), dim3(BLK), 0, st, P, s.pos, cstart, cend,
                     occ, dil, dgrid);
}

void launch_normals(const DevParams &P, const float *dgrid, float *normal3, hipStream_t st) {
  hipLaunchKernelGGL(k_normals, grid_for(P.tot_dens), dim3(BLK), 0, st, P, dgrid, normal3);
}

size_t mc_segments(const DevParams &P) {
  const size_t nz = P.nd[2] - 1;
  return (size_t)(P.nd[0] - 1) * (P.nd[1] - 1) * ((nz + MC_SEG - 1) / MC_SEG);
}

void launch_mc_count(const DevParams &P, const float *dgrid, uint *seg_tri, uint *seg_cube,
                     hipStream_t st) {
  hipLaunchKernelGGL(k_mc_count, grid_for(mc_segments(P)), dim3(BLK), 0, st, P, dgrid, seg_tri,
                     seg_cube);
}

void launch_mc_list(const DevParams &P, const float *dgrid, const uint *seg_cube,
                    const uint *seg_toff, const uint *seg_coff, uint cap, uint2 *clist,
                    hipStream_t st) {
  hipLaunchKernelGGL(k_mc_list, grid_for(mc_segments(P)), dim3(BLK), 0, st, P, dgrid, seg_cube,
                     seg_toff, seg_coff, cap, clist);
}

void launch_mc_emit(const DevParams &P, const float *dgrid, const float *normal3, const uint2 *clist,
                    const Counters *cnt, uint max_tri, float *tris, hipStream_t st) {
  const uint blocks = (uint)std::min<size_t>(std::max<size_t>(grid_for(max_tri).x, 1), 4096);
  hipLaunchKernelGGL(k_mc_emit, dim3(blocks), dim3(BLK), 0, st, P, dgrid, normal3, clist, cnt,
                     max_tri, tris);
}

void launch_upload_scatter(int variant, const void *aos_dev, uint n, DevState s, hipStream_t st,
                           uint slot_or) {
  if (n == 0) return;
  SPH_DISPATCH(variant, k_upload, grid_for(n), dim3(BLK), 0, st, aos_dev, n, s, slot_or);
}

void launch_download_gather(int variant, DevState s, const uint *slot2phys, uint n, void *aos_dev,
                            hipStream_t st) {
  if (n == 0) return;
  SPH_DISPATCH(variant, k_download, grid_for(n), dim3(BLK), 0, st, s, slot2phys, n, aos_dev);
}

void launch_slot2phys(DevState s, uint m, uint *slot2phys, hipStream_t st) {
  if (m == 0) return;
  hipLaunchKernelGGL(k_slot2phys, grid_for(m), dim3(BLK), 0, st, s, m, slot2phys);
}
