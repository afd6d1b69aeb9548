// Device primitives replacing the reference's thrust calls (thrust::sort_by_key and
// thrust::reduce, SPH-3D-2/sph_system.cu:126-128, :1639-1643, :1832-1836):
//   * exclusive scan (deterministic integer reduce-then-scan: per-tile sums, then each tile
//     adds the sums of the tiles before it, which it reads itself when there are at most
//     1024 tiles and otherwise takes from a single-block scan of the sums);
//   * stable LSD radix sort of (cell hash, slot) with digits of up to 9 bits (an 18-bit key,
//     64^3 cells, takes two passes); each pass is a per-tile LDS histogram, a scan of the
//     digit-major tile histograms, and a scatter whose stable in-tile rank comes from
//     64-lane ballot matching plus per-wave digit counts in LDS.
// All integer work, no floating point, results independent of scheduling.
#include <mutex>

#include "sph_internal.h"

namespace {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;  // 2048
constexpr uint SCAN_FUSED_TILES = 4 * SCAN_THREADS;     // two-kernel scan up to 1024 tiles

constexpr int RADIX_BITS = 9;           // widest digit
constexpr int RADIX = 1 << RADIX_BITS;  // 512 bins, two per thread
constexpr int SORT_THREADS = 256;
constexpr int SORT_WAVES = SORT_THREADS / 64;
constexpr int SORT_ITEMS = 4;
constexpr int SORT_TILE = SORT_THREADS * SORT_ITEMS;  // 1024 keys per tile (more tiles fill the GPU)

__device__ inline uint wave_incl_scan(uint x, int lane) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  return x;
}

// Exclusive scan of one value per thread across the block; returns the block total.
__device__ inline uint block_excl_scan(uint x, uint *lds_waves, uint &total) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  uint inc = wave_incl_scan(x, lane);
  if (lane == 63) lds_waves[w] = inc;
  __syncthreads();
  uint before = 0, tot = 0;
  for (int k = 0; k < (int)(blockDim.x >> 6); k++) {
    uint v = lds_waves[k];
    if (k < w) before += v;
    tot += v;
  }
  __syncthreads();
  total = tot;
  return before + inc - x;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(const uint *__restrict__ in,
                                                              const uint *n_dev, uint n_host,
                                                              uint *__restrict__ tmp) {
  __shared__ uint lw[SCAN_THREADS / 64];
  const uint n = n_dev ? min(*n_dev, n_host) : n_host;
  const size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
  uint s = 0;
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; j++)
    if (base + j < n) s += in[base + j];
  uint tot;
  block_excl_scan(s, lw, tot);
  if (threadIdx.x == 0) tmp[blockIdx.x] = tot;
}

__global__ __launch_bounds__(1024) void k_scan_top(uint *__restrict__ tmp, const uint *n_dev,
                                                   uint n_host, uint *total_dev) {
  __shared__ uint lw[16];
  const uint n = n_dev ? min(*n_dev, n_host) : n_host;
  const uint nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  uint carry = 0;
  for (uint base = 0; base < nb; base += blockDim.x) {
    uint i = base + threadIdx.x;
    uint v = i < nb ? tmp[i] : 0u;
    uint tot;
    uint ex = block_excl_scan(v, lw, tot);
    if (i < nb) tmp[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0 && total_dev) *total_dev = carry;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_down(const uint *__restrict__ in,
                                                            uint *__restrict__ out,
                                                            const uint *n_dev, uint n_host,
                                                            const uint *__restrict__ tmp) {
  __shared__ uint lw[SCAN_THREADS / 64];
  const uint n = n_dev ? min(*n_dev, n_host) : n_host;
  const size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
  if ((size_t)blockIdx.x * SCAN_TILE >= n) return;  // uniform per block
  uint v[SCAN_ITEMS];
  uint s = 0;
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; j++) {
    v[j] = base + j < n ? in[base + j] : 0u;
    s += v[j];
  }
  uint tot;
  uint run = block_excl_scan(s, lw, tot) + tmp[blockIdx.x];
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; j++) {
    if (base + j < n) out[base + j] = run;
    run += v[j];
  }
}

// Down-sweep that forms its own carry: the sum of the tile sums before this tile (at most
// SCAN_FUSED_TILES of them, up to four per thread, added in a fixed order). The last tile
// writes the total; with n = 0 tile 0 writes a zero total.
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_down_carry(const uint *__restrict__ in,
                                                                  uint *__restrict__ out,
                                                                  const uint *n_dev, uint n_host,
                                                                  const uint *__restrict__ tmp,
                                                                  uint *total_dev) {
  __shared__ uint lw[SCAN_THREADS / 64];
  const uint n = n_dev ? min(*n_dev, n_host) : n_host;
  if (n == 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0 && total_dev) *total_dev = 0u;
    return;
  }
  if ((size_t)blockIdx.x * SCAN_TILE >= n) return;  // uniform per block
  const uint nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  uint part = 0;
  for (uint k = threadIdx.x; k < blockIdx.x; k += SCAN_THREADS) part += tmp[k];
  uint carry;
  block_excl_scan(part, lw, carry);
  const size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
  uint v[SCAN_ITEMS];
  uint s = 0;
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; j++) {
    v[j] = base + j < n ? in[base + j] : 0u;
    s += v[j];
  }
  uint tot;
  uint run = block_excl_scan(s, lw, tot) + carry;
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; j++) {
    if (base + j < n) out[base + j] = run;
    run += v[j];
  }
  if (blockIdx.x == nb - 1 && threadIdx.x == 0 && total_dev) *total_dev = carry + tot;
}

__global__ __launch_bounds__(256) void k_fill(uint *p, uint value, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = value;
}

__global__ __launch_bounds__(256) void k_fill2(uint *a, uint va, uint *b, uint vb, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    a[i] = va;
    b[i] = vb;
  }
}

// Keys to sort: the device count when given (clamped to the host bound), else the host's.
__device__ inline uint sort_count(const uint *n_dev, uint n_host) {
  return n_dev ? min(*n_dev, n_host) : n_host;
}

// Per-tile digit histogram, stored digit-major: hist[d * ntiles + tile], d < 2^width,
// ntiles = ceil(n / SORT_TILE). Workgroup b takes tiles b, b + gridDim.x, ... (one tile each
// in a plain launch). scan_len (optional): receives the histogram's length, 2^width * ntiles.
__global__ __launch_bounds__(SORT_THREADS) void k_radix_hist(const uint *__restrict__ keys,
                                                             const uint *n_dev, uint n_host,
                                                             int shift, int width,
                                                             uint *__restrict__ hist,
                                                             uint *scan_len) {
  __shared__ uint h[RADIX];
  const int t = threadIdx.x;
  const int nbin = 1 << width;
  const uint mask = (uint)nbin - 1u;
  const uint n = sort_count(n_dev, n_host);
  const uint ntiles = (n + SORT_TILE - 1) / SORT_TILE;
  if (scan_len && blockIdx.x == 0 && t == 0) *scan_len = (uint)nbin * ntiles;
  for (uint tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    for (int d = t; d < nbin; d += SORT_THREADS) h[d] = 0;
    __syncthreads();
    const size_t base = (size_t)tile * SORT_TILE;
#pragma unroll 4
    for (int j = 0; j < SORT_ITEMS; j++) {
      size_t i = base + (size_t)j * SORT_THREADS + t;
      if (i < n) atomicAdd(&h[(keys[i] >> shift) & mask], 1u);
    }
    __syncthreads();
    for (int d = t; d < nbin; d += SORT_THREADS) hist[(size_t)d * ntiles + tile] = h[d];
    __syncthreads();  // h is cleared for the next tile
  }
}

// Stable scatter. Items of a tile are taken in index order (round j, then thread t); the
// rank of an item among equal digits = digits seen in earlier rounds of this tile (run)
// + equal digits in earlier waves of this round (wcnt) + equal digits in lower lanes of
// its own wave (ballot match).
__global__ __launch_bounds__(SORT_THREADS) void k_radix_scatter(
    const uint *__restrict__ keys_in, const uint *__restrict__ vals_in, const uint *n_dev,
    uint n_host, int shift, int width, const uint *__restrict__ offs,
    uint *__restrict__ keys_out, uint *__restrict__ vals_out) {
  __shared__ uint run[RADIX];
  __shared__ uint wcnt[SORT_WAVES][RADIX];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int nbin = 1 << width;
  const uint mask = (uint)nbin - 1u;
  const uint n = sort_count(n_dev, n_host);
  const uint ntiles = (n + SORT_TILE - 1) / SORT_TILE;
  for (uint tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
  for (int d = t; d < nbin; d += SORT_THREADS) {
    run[d] = offs[(size_t)d * ntiles + tile];
#pragma unroll
    for (int k = 0; k < SORT_WAVES; k++) wcnt[k][d] = 0;
  }
  __syncthreads();
  const size_t base = (size_t)tile * SORT_TILE;
  const unsigned long long lt = (1ull << lane) - 1ull;
  for (int j = 0; j < SORT_ITEMS; j++) {
    const size_t i = base + (size_t)j * SORT_THREADS + t;
    const bool valid = i < n;
    uint key = valid ? keys_in[i] : 0u;
    uint val = valid ? (vals_in ? vals_in[i] : (uint)i) : 0u;
    uint d = (key >> shift) & mask;
    unsigned long long peers = __ballot(valid);
    for (int b = 0; b < width; b++) {  // width is uniform: the ballots stay wave-wide
      unsigned long long bb = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? bb : ~bb;
    }
    uint rank = (uint)__popcll(peers & lt);
    if (valid && rank == 0) wcnt[w][d] = (uint)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint pos = run[d] + rank;
      for (int k = 0; k < w; k++) pos += wcnt[k][d];
      keys_out[pos] = key;
      vals_out[pos] = val;
    }
    __syncthreads();
    for (int dd = t; dd < nbin; dd += SORT_THREADS) {
      uint add = 0;
#pragma unroll
      for (int k = 0; k < SORT_WAVES; k++) {
        add += wcnt[k][dd];
        wcnt[k][dd] = 0;
      }
      run[dd] += add;
    }
    __syncthreads();
  }
  }  // tile
}

inline uint blocks_for(size_t n, size_t per) { return (uint)((n + per - 1) / per); }

}  // namespace

size_t sph_scan_tmp_elems(size_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE + 1; }
size_t sph_sort_tiles(size_t n) { return (n + SORT_TILE - 1) / SORT_TILE; }
size_t sph_sort_hist_elems(size_t n) { return (size_t)RADIX * sph_sort_tiles(n); }

void sph_exclusive_scan(const uint *in, uint *out, const uint *n_dev, uint n_max, uint *total_dev,
                        uint *tmp, hipStream_t st) {
  uint nb = blocks_for(n_max, SCAN_TILE);
  if (nb == 0) nb = 1;
  if (nb <= SCAN_FUSED_TILES) {
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(SCAN_THREADS), 0, st, in, n_dev, n_max, tmp);
    hipLaunchKernelGGL(k_scan_down_carry, dim3(nb), dim3(SCAN_THREADS), 0, st, in, out, n_dev,
                       n_max, tmp, total_dev);
    return;
  }
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(SCAN_THREADS), 0, st, in, n_dev, n_max, tmp);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, st, tmp, n_dev, n_max, total_dev);
  hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(SCAN_THREADS), 0, st, in, out, n_dev, n_max, tmp);
}

void sph_fill_u32(uint *p, uint value, size_t n, hipStream_t st) {
  if (n == 0) return;
  size_t nb = (n + 255) / 256;
  if (nb > 65536) nb = 65536;
  hipLaunchKernelGGL(k_fill, dim3((uint)nb), dim3(256), 0, st, p, value, n);
}

// Undo launch_cell_bounds over the same sorted keys: the cells at the run starts get their
// fill values back (the other cells were never written).
__global__ __launch_bounds__(256) void k_reset_cells(const uint *__restrict__ skeys, uint n,
                                                     uint ncell, uint *__restrict__ cstart,
                                                     uint *__restrict__ cend) {
  const uint k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  const uint h = skeys[k];
  if (h < ncell && (k == 0 || skeys[k - 1] != h)) {
    cstart[h] = SPH_EMPTY;
    cend[h] = 0u;
  }
}

void sph_reset_cells(const uint *skeys, uint n, uint ncell, uint *cstart, uint *cend,
                     hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_reset_cells, dim3((n + 255) / 256), dim3(256), 0, st, skeys, n, ncell,
                     cstart, cend);
}

void sph_fill_cells(uint *cstart, uint *cend, size_t n, hipStream_t st) {
  if (n == 0) return;
  size_t nb = (n + 255) / 256;
  if (nb > 65536) nb = 65536;
  hipLaunchKernelGGL(k_fill2, dim3((uint)nb), dim3(256), 0, st, cstart, SPH_EMPTY, cend, 0u, n);
}

namespace {

// LSD passes over keys[0..n): pass 0 reads (k0, v0) -- v0 == nullptr: the values are the
// indices 0..n-1 -- and writes (k1, v1); each later pass reads what the one before wrote
// and writes the other pair, (vbuf0 standing in as v0's storage). The result ends in
// (*keys_out, *vals_out).
// n_dev (optional): a device word holding the key count, n only bounding it; grid
// (optional): a persistent grid for the histogram and scatter kernels; len_dev: a device word
// for the histogram length when n_dev is given.
void radix_passes(uint *k0, const uint *v0, uint *vbuf0, uint *k1, uint *v1, uint n, int bits,
                  uint *tile_hist, uint *scan_tmp, uint **keys_out, uint **vals_out,
                  hipStream_t st, const uint *n_dev = nullptr, uint grid = 0,
                  uint *len_dev = nullptr) {
  const uint *kin = k0, *vin = v0;
  uint *kout = k1, *vout = v1;
  const uint ntiles = (uint)sph_sort_tiles(n);
  const uint tiles_grid = grid ? grid : ntiles;
  if (bits < 1) bits = 1;
  // fewest passes of at most RADIX_BITS bits, the digit width spread evenly over them
  const int passes = (bits + RADIX_BITS - 1) / RADIX_BITS;
  const int width = (bits + passes - 1) / passes;
  for (int p = 0; p < passes; p++) {
    const int shift = p * width;
    hipLaunchKernelGGL(k_radix_hist, dim3(tiles_grid), dim3(SORT_THREADS), 0, st, kin, n_dev, n,
                       shift, width, tile_hist, n_dev ? len_dev : nullptr);
    sph_exclusive_scan(tile_hist, tile_hist, n_dev ? len_dev : nullptr,
                       (uint)((1u << width) * ntiles), nullptr, scan_tmp, st);
    hipLaunchKernelGGL(k_radix_scatter, dim3(tiles_grid), dim3(SORT_THREADS), 0, st, kin, vin,
                       n_dev, n, shift, width, tile_hist, kout, vout);
    // next pass reads what this one wrote
    uint *nk = kout, *nv = vout;
    if (kout == k1) {
      kout = k0;
      vout = vbuf0;
    } else {
      kout = k1;
      vout = v1;
    }
    kin = nk;
    vin = nv;
  }
  *keys_out = const_cast<uint *>(kin);
  *vals_out = const_cast<uint *>(vin);
}

}  // namespace

void sph_radix_sort(uint *keys, uint *vals, uint n, int bits, SortScratch &sc, uint **keys_out,
                    uint **vals_out, hipStream_t st, const uint *n_dev, uint grid) {
  radix_passes(keys, nullptr, vals, sc.keys_alt, sc.vals_alt, n, bits, sc.tile_hist, sc.scan_tmp,
               keys_out, vals_out, st, n_dev, grid, sc.len_dev);
}

// ---------------------------------------------------------------------------------------
// Sort without allocation. The sort's working memory (ping-pong partners, tile histograms,
// scan block sums) is one static device array instead of buffers a context allocates, so
// the call makes no hipMalloc/hipFree: for callers that cannot allocate (code under stream
// capture into a graph, serving loops that forbid allocation). Each GPU holds its own copy
// of the array for the life of the process; every call on one GPU shares it, so calls that
// overlap on one GPU (other streams, other threads) race. Nothing in it needs clearing:
// every pass writes the histogram entries it scans before reading them.
// ---------------------------------------------------------------------------------------
namespace {
constexpr uint kNoAllocMaxKeys = 1u << 20;
constexpr size_t kNoAllocHist = (size_t)RADIX * ((kNoAllocMaxKeys + SORT_TILE - 1) / SORT_TILE);
constexpr size_t kNoAllocScanTmp = (kNoAllocHist + SCAN_TILE - 1) / SCAN_TILE + 1;
// [keys_alt | vals_alt | tile histograms | scan block sums], every part 256-byte aligned
constexpr size_t kNoAllocKeysAlt = 0;
constexpr size_t kNoAllocValsAlt = kNoAllocKeysAlt + kNoAllocMaxKeys;
constexpr size_t kNoAllocHistOff = kNoAllocValsAlt + kNoAllocMaxKeys;
constexpr size_t kNoAllocScanOff = kNoAllocHistOff + kNoAllocHist;
constexpr size_t kNoAllocWords = kNoAllocScanOff + (kNoAllocScanTmp + 63) / 64 * 64;
static_assert(kNoAllocHistOff % 64 == 0 && kNoAllocScanOff % 64 == 0, "scratch part alignment");
static_assert(SPH_SORT_NOALLOC_MAX_KEYS == kNoAllocMaxKeys, "header limit");
}  // namespace

static __device__ __align__(256) uint scratch[kNoAllocWords];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint *g_scratch_dev[kMaxDevices] = {};  // device address of `scratch`, per GPU

// The device address of this GPU's copy of `scratch`, looked up once per device (with that
// device current) on its first call. The array's host-side name is not a device address.
int noalloc_scratch(uint **out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return SPH_EHIP;
  if (dev < 0 || dev >= kMaxDevices) return SPH_EINVAL;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return SPH_EHIP;
    g_scratch_dev[dev] = (uint *)p;
  }
  *out = g_scratch_dev[dev];
  return SPH_OK;
}

__global__ __launch_bounds__(256) void k_copy_pairs(const uint *__restrict__ ka,
                                                    const uint *__restrict__ va, uint *__restrict__ kb,
                                                    uint *__restrict__ vb, uint n) {
  const uint i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) {
    kb[i] = ka[i];
    vb[i] = va[i];
  }
}
}  // namespace

extern "C" int sph_sort_noalloc(uint32_t *hash, uint32_t *index, uint32_t n, uint32_t key_bits,
                                void *stream) {
  if (n == 0) return SPH_OK;
  if (!hash || !index || key_bits > 32) return SPH_EINVAL;
  if (n > kNoAllocMaxKeys) return SPH_ECAP;
  uint *base = nullptr;
  int rc = noalloc_scratch(&base);
  if (rc != SPH_OK) return rc;
  hipStream_t st = (hipStream_t)stream;
  uint *ko = nullptr, *vo = nullptr;
  radix_passes(hash, index, index, base + kNoAllocKeysAlt, base + kNoAllocValsAlt, n,
               key_bits ? (int)key_bits : 32, base + kNoAllocHistOff, base + kNoAllocScanOff, &ko,
               &vo, st);
  if (ko != hash)  // an odd number of passes ends in the scratch pair
    hipLaunchKernelGGL(k_copy_pairs, dim3((n + 255) / 256), dim3(256), 0, st, ko, vo, hash, index, n);
  return hipGetLastError() == hipSuccess ? SPH_OK : SPH_EHIP;
}

// A plain copy of n 16-byte units: one load and one store per unit, grid-stride.
__global__ __launch_bounds__(256) void k_copy_calibration(const float4 *__restrict__ src,
                                                          float4 *__restrict__ dst, size_t n) {
  const size_t stride = (size_t)gridDim.x * 256;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

extern "C" int sph_copy_calibration(size_t bytes, int reps) {
  const size_t n = bytes / 16;
  if (n == 0 || reps <= 0) return SPH_EINVAL;
  float4 *src = nullptr, *dst = nullptr;
  if (hipMalloc(&src, n * 16) != hipSuccess) return SPH_ECAP;
  if (hipMalloc(&dst, n * 16) != hipSuccess) {
    (void)hipFree(src);
    return SPH_ECAP;
  }
  hipError_t e = hipMemset(src, 0, n * 16);
  const size_t blocks = std::min<size_t>((n + 255) / 256, 256 * 32);
  for (int r = 0; r < reps && e == hipSuccess; r++) {
    k_copy_calibration<<<dim3((uint)blocks), dim3(256), 0, 0>>>(src, dst, n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  (void)hipFree(src);
  (void)hipFree(dst);
  return e == hipSuccess ? SPH_OK : SPH_EHIP;
}
