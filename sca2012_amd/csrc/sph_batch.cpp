// Batches of independent scenes (declared in include/sph_amd.h, "batches of independent
// scenes"). The reference keeps its parameters in one process-wide __constant__ block
// (SPH-3D-2/sph_kernel.cu:10), so it steps one scene per process; here the parameters are
// a per-scene table in device memory and every phase of the step is one launch over the
// particles (or cells) of all scenes.
//
// Layout of a batch of S scenes:
//   cells     scene s owns the global cells [cell_base_s, cell_base_s + total_cell_s);
//             a particle's sort key is cell_base_s + its reference hash, so one radix sort
//             orders all scenes, scene after scene, cells ascending within a scene.
//   slots     the slot word of a particle is s << 24 | its slot within scene s. The sort
//             input packs the scenes densely (scene s at [first_s, first_s + n_s)), so ties
//             stay in ascending slot within each scene, as in the single-scene sort.
//   entries   the SoA state arrays hold live entries [0, M) in cell order after the
//             reorder, then the split children of all scenes (scene by scene, in parent
//             rank order); merge victims stay as removed entries until the next sort.
// Split capacity (floor((cap_s - n_s)/(n_split - 1)) parents) and merge groups are decided
// per scene, so each scene evolves exactly as it would in its own SphCtx.
//
// Two ways to run a step:
//   graph     (default) the whole step is one captured hipGraph: every per-item phase is a
//             persistent launch (a fixed grid of workgroups per CU, each looping over blocks
//             of 256 items) that reads its item count from the device counters, so nothing
//             in the step depends on the host's view of the counts and the host launches
//             step after step without waiting. A launch tuner picks the workgroups per CU
//             and the pair-list lengths by timing whole steps.
//   stream    the kernels launched one by one per step, sized by the counts the host read
//             at the end of the previous step (one synchronisation per step).
// Both give the same bits for every scene: the launch shape never changes what an item
// computes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "sph_internal.h"

namespace {

// Workgroups per CU of a graph step's element-wise, sort and scan launches (the candidates
// of the launch tuner): 0 sizes every launch by its buffer's bound, one workgroup per block
// of 256 entries, the blocks past the device count exiting at once; 4, 8, 16 per CU make
// them persistent, each workgroup looping over blocks. The neighbour kernels and merge are
// always sized by the bound: their blocks cost unevenly and the hardware dispatcher
// balances them better than persistent workgroups taking blocks from work queues.
constexpr uint kGridPerCU[4] = {0, 4, 8, 16};
constexpr int kGridDefault = 0;

// Tuner of a graph step's launch shape: the workgroups per CU of the persistent launches
// (4 candidates) and the force and energy pair-list lengths (3 each), every choice giving
// the same bits. Steps are timed whole (events around each graph launch). One dimension is
// tuned per round, in turn: each candidate of that dimension runs two steps, in the order
// 0 1 .. k-1 k-1 .. 1 0 so that a cost drifting from step to step weighs on all alike, the
// others held at their current choice; once the round's times are read back (after the
// call's synchronisation) the dimension takes the candidate with the lowest mean step time
// if it beats the current one by more than kTuneMargin. Rounds follow one another from
// `after` steps on; a cycle over the three dimensions repeats every `every` steps (a
// scene's neighbour counts, and with them the best lengths, drift).
struct LaunchTuner {
  static constexpr int kDims = 3;
  static constexpr int kCands[kDims] = {4, 3, 3};
  static constexpr uint kCycle = 2 * (4 + 3 + 3);  // steps of one cycle of rounds
  int cur[kDims] = {kGridDefault, kForceListDefault, kEnergyListDefault};
  uint rounds_left = 0, skip = 0, every = 0;
  int dim = 0, pos = 0;  // dimension of the round being issued, its steps issued
  uint rid = 0;          // rounds issued so far (tags the timed steps)
  // the round of each dimension whose times are being gathered
  uint acc_rid[kDims] = {0, 0, 0};
  int acc_n[kDims] = {0, 0, 0};
  float sum[kDims][4];
  int cnt[kDims][4];
  float best[kDims][4];  // mean ms per candidate in the last finished round of each dimension
  struct Pending { int dim, cand; uint rid; hipEvent_t a, b; };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;

  ~LaunchTuner() {
    for (auto &p : pending) {
      (void)hipEventDestroy(p.a);
      if (p.b) (void)hipEventDestroy(p.b);
    }
    for (auto e : pool) (void)hipEventDestroy(e);
  }
  void clear_sums(int d) {
    acc_n[d] = 0;
    for (int c = 0; c < 4; c++) {
      sum[d][c] = 0.0f;
      cnt[d][c] = 0;
    }
  }
  void start(uint rounds, uint after, uint every_steps) {
    rounds_left = rounds;
    skip = after;
    every = std::max<uint>(every_steps, kCycle);
    dim = 0;
    pos = 0;
    for (int d = 0; d < kDims; d++) {
      clear_sums(d);
      acc_rid[d] = 0xffffffffu;
      for (int c = 0; c < 4; c++) best[d][c] = -1.0f;
    }
  }
  int round_len(int d) const { return 2 * kCands[d]; }
  bool tuning() const { return rounds_left > 0 && skip == 0; }
  // the shape of the next step: {grid candidate, force list, energy list}
  void next(int shape[kDims]) const {
    for (int d = 0; d < kDims; d++) shape[d] = cur[d];
    if (tuning()) {
      const int k = kCands[dim];
      shape[dim] = pos < k ? pos : 2 * k - 1 - pos;
    }
  }
  hipEvent_t take() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
  }
  // around the graph launch of a step with `shape`
  void before(hipStream_t st, const int shape[kDims], hipEvent_t &a) {
    a = nullptr;
    if (!tuning()) return;
    a = take();
    (void)hipEventRecord(a, st);
    pending.push_back({dim, shape[dim], rid, a, nullptr});
  }
  void after_launch(hipStream_t st, hipEvent_t a) {
    if (a) {
      hipEvent_t b = take();
      (void)hipEventRecord(b, st);
      pending.back().b = b;
    }
    if (rounds_left == 0) return;
    if (skip > 0) {
      skip--;
      return;
    }
    if (++pos < round_len(dim)) return;
    // the round is issued: the next dimension's round follows at once, a new cycle
    // `every` steps after this one began
    rounds_left--;
    rid++;
    dim = (dim + 1) % kDims;
    pos = 0;
    skip = dim == 0 ? every - kCycle : 0;
  }
  void finish(int d) {
    const int k = kCands[d];
    for (int c = 0; c < k; c++) best[d][c] = cnt[d][c] ? sum[d][c] / (float)cnt[d][c] : -1.0f;
    int pick = cur[d];
    for (int c = 0; c < k; c++)
      if (best[d][c] >= 0.0f && best[d][cur[d]] >= 0.0f &&
          best[d][c] < best[d][cur[d]] * (1.0f - kTuneMargin) && best[d][c] < best[d][pick])
        pick = c;
    cur[d] = pick;
    clear_sums(d);
  }
  // after the stream is synchronised: fold the timed steps in, in issue order; a round
  // whose steps are all in decides its dimension
  void collect() {
    for (auto &p : pending) {
      const int d = p.dim;
      if (p.rid != acc_rid[d]) {
        clear_sums(d);
        acc_rid[d] = p.rid;
      }
      float ms = 0.0f;
      if (p.b && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
        sum[d][p.cand] += ms;
        cnt[d][p.cand] += 1;
      }
      if (++acc_n[d] == round_len(d)) finish(d);
      pool.push_back(p.a);
      if (p.b) pool.push_back(p.b);
    }
    pending.clear();
  }
};

}  // namespace

struct SphBatch {
  int variant = 0;
  uint ns = 0;
  size_t rec = 0, raw_size = 0;
  std::vector<unsigned char> raw;  // ns SysParam records
  std::vector<DevParams> P;        // host copy of the table
  std::vector<uint> cap;
  uint total_cell = 0;
  size_t phys_cap = 0;  // entries of the state arrays (sum of capacities)
  uint max_cap = 0;
  hipStream_t st = nullptr;
  std::vector<void *> allocs;
  DevParams *Ptab = nullptr;
  SceneCnt *scn = nullptr;
  SceneCnt *scn_host = nullptr;
  Counters *cnt = nullptr;
  Counters *cnt_host = nullptr;
  DevState cur{}, alt{};
  Aux aux{};
  uint *keys = nullptr, *vals = nullptr;
  SortScratch sc{};
  uint *skeys = nullptr, *svals = nullptr;
  uint *s2p_hash = nullptr, *s2p_cell = nullptr;
  uint *cstart = nullptr, *cend = nullptr;
  uint *flag_by_slot = nullptr, *rank_by_slot = nullptr, *alive_by_slot = nullptr;
  uint *newslot = nullptr, *status = nullptr, *scan_tmp = nullptr;
  unsigned char *aos_dev = nullptr;
  int kbits = 1;
  uint M = 0;      // live particles of all scenes
  uint m = 0;      // entries, removed ones included
  std::vector<uint> n_host;  // live particles per scene
  bool pushes = false;       // some scene has a one-step push pending
  hipError_t last_err = hipSuccess;
  // profiling (hipEvents around each phase on the batch's stream)
  uint32_t prof_mask = 0;  // phases timed with events (bit k = phase k)
  struct Pending { int phase; hipEvent_t a, b; };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> ev_pool;
  double prof_ms[SPH_P_COUNT] = {0};
  uint64_t prof_n[SPH_P_COUNT] = {0};
  int open_phase = -1;
  hipEvent_t open_ev = nullptr;

  ListTuner tune;  // pair-list lengths of the force and energy kernels (stream steps)

  // graph steps
  bool use_graph = true;
  uint num_cu = 256;
  LaunchTuner ltune;
  struct GraphEntry { uint key; hipGraphExec_t exec; };
  std::vector<GraphEntry> graphs;  // one per step shape, flags and state buffer parity
  bool parity = false;             // cur is the second state buffer (the graphs bake pointers)

  ~SphBatch() {
    if (st) (void)hipStreamSynchronize(st);
    for (auto &g : graphs) (void)hipGraphExecDestroy(g.exec);
    tune.destroy();
    for (auto &p : pending) {
      (void)hipEventDestroy(p.a);
      (void)hipEventDestroy(p.b);
    }
    for (auto e : ev_pool) (void)hipEventDestroy(e);
    for (void *p : allocs) (void)hipFree(p);
    if (cnt_host) (void)hipHostFree(cnt_host);
    if (scn_host) (void)hipHostFree(scn_host);
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
    return alloc(s.pos, phys_cap) && alloc(s.vel, phys_cap) && alloc(s.ev, phys_cap) &&
           alloc(s.acc, phys_cap) && alloc(s.dp, phys_cap) && alloc(s.energy, phys_cap) &&
           alloc(s.slot, phys_cap);
  }
  BatchView view() const { return BatchView{Ptab, scn, ns, total_cell}; }

  hipEvent_t take_event() {
    if (!ev_pool.empty()) {
      hipEvent_t e = ev_pool.back();
      ev_pool.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
  }
  void pbegin(int phase) {
    if (!((prof_mask >> phase) & 1u)) return;
    open_phase = phase;
    open_ev = take_event();
    (void)hipEventRecord(open_ev, st);
  }
  void pend() {
    if (open_phase < 0) return;
    hipEvent_t b = take_event();
    (void)hipEventRecord(b, st);
    pending.push_back({open_phase, open_ev, b});
    open_phase = -1;
  }
  void collect() {
    for (auto &p : pending) {
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
        prof_ms[p.phase] += ms;
        prof_n[p.phase] += 1;
      }
      ev_pool.push_back(p.a);
      ev_pool.push_back(p.b);
    }
    pending.clear();
  }

  // host mirrors -> device counters (after uploads)
  int push_counts() {
    Counters c;
    std::memset(&c, 0, sizeof(c));
    c.n = M;
    c.m = m;
    c.n_pre = M;
    *cnt_host = c;
    for (uint s = 0; s < ns; s++) {
      std::memset(&scn_host[s], 0, sizeof(SceneCnt));
      scn_host[s].n = n_host[s];
      scn_host[s].n_pre = n_host[s];
    }
    if (hipMemcpyAsync(cnt, cnt_host, sizeof(Counters), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(scn, scn_host, sizeof(SceneCnt) * ns, hipMemcpyHostToDevice, st) != hipSuccess)
      return SPH_EHIP;
    return SPH_OK;
  }

  // synchronise and read every counter back
  int sync() {
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(cnt_host, cnt, sizeof(Counters), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
      e = hipMemcpyAsync(scn_host, scn, sizeof(SceneCnt) * ns, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      last_err = e;
      return SPH_EHIP;
    }
    collect();
    M = cnt_host->n;
    m = cnt_host->m;
    for (uint s = 0; s < ns; s++) n_host[s] = scn_host[s].n;
    return SPH_OK;
  }

  int upload_params() {
    return hipMemcpyAsync(Ptab, P.data(), sizeof(DevParams) * ns, hipMemcpyHostToDevice, st) ==
                   hipSuccess
               ? SPH_OK
               : SPH_EHIP;
  }

  // the removed entries of [0, m) are dropped: live entries move to the front, in order
  int pack() {
    uint *flags = flag_by_slot, *pos = rank_by_slot;
    launch_pack_live(cur, alt, m, flags, pos, scan_tmp, &cnt->scan_total, st);
    swap_state();
    m = M;
    return SPH_OK;
  }
  void swap_state() {
    std::swap(cur, alt);
    parity = !parity;
  }

  // One step of every scene.
  //   stream step (grid == 0): launches sized by M (live) and m (entries), which the host
  //     read at the end of the previous step; the list tuner picks the pair-list lengths.
  //   graph step (dev): launches that read every count on the device (the host bounds are
  //     the buffers' size), sized by those bounds (grid == 0) or persistent over `grid`
  //     workgroups, with list lengths lf, le; no events, so the enqueued work can be
  //     captured into a graph.
  bool energy_unit_all() const {
    for (const DevParams &p : P)
      if (!p.energy_unit) return false;
    return true;
  }

  void enqueue_step(uint flags, bool dev = false, uint grid = 0, int lf = -1, int le = -1) {
    const BatchView bv = view();
    const DevParams &P0 = P[0];
    const int v = variant;
    const uint Mb = dev ? (uint)phys_cap : M;  // live entries: host count or bound
    const uint mb = dev ? (uint)phys_cap : m;  // entries, removed ones included
    const uint *n_live = dev ? &cnt->n : nullptr;   // live entries at the start of the step
    const uint *n_pre = dev ? &cnt->n_pre : nullptr;  // the same, once split has moved n
    launch_batch_begin(bv, cnt, st);
    pbegin(SPH_P_HASH);
    launch_hash(v, P0, cur, &cnt->m, mb, keys, s2p_hash, cnt, st, &bv, grid);
    pend();
    pbegin(SPH_P_SORT);
    sph_radix_sort(keys, vals, Mb, kbits, sc, &skeys, &svals, st, n_live, grid);
    pend();
    pbegin(SPH_P_CELLS);
    sph_fill_cells(cstart, cend, total_cell, st);
    launch_cell_bounds(skeys, Mb, total_cell, cstart, cend, st, n_live, grid);
    pend();
    pbegin(SPH_P_REORDER);
    // and m = n; acc, dp and energy are recomputed by the step's kernels below
    launch_reorder(cur, alt, svals, s2p_hash, Mb, s2p_cell, st, cnt, false, n_live, grid);
    pend();
    swap_state();
    pbegin(SPH_P_DENSITY);
    launch_density(v, P0, cur, aux, cstart, cend, Mb, st, &bv, n_live);
    pend();
    pbegin(SPH_P_FORCE);
    auto force = [&](int c) {
      launch_force(v, P0, cur, aux, cstart, cend, Mb, st, &bv, c, n_live, true);
    };
    if (dev) force(lf);
    else tune.run(0, st, force);
    pend();
    pbegin(SPH_P_INTEGRATE);
    launch_integrate(v, P0, cur, Mb, st, &bv, aux.erec, n_live, grid);
    pend();
    const bool do_split = (flags & SPH_STEP_SPLIT) != 0;
    pbegin(SPH_P_ENERGY);
    uint *split_flags = do_split ? flag_by_slot : nullptr;  // the energy kernel decides splits
    DevParams Pe = P0;  // the energy kernel's unit-length path only if every scene takes it
    Pe.energy_unit = energy_unit_all() ? 1 : 0;
    auto energy = [&](int c) {
      launch_energy(v, Pe, cur, cstart, cend, Mb, st, &bv, c, split_flags, aux.erec, n_live);
    };
    if (dev) energy(le);
    else tune.run(1, st, energy);
    pend();
    pbegin(SPH_P_SPLIT);
    if (do_split)
      sph_exclusive_scan(flag_by_slot, rank_by_slot, n_live, Mb, &cnt->scan_total, scan_tmp, st);
    launch_batch_split_plan(v, bv, rank_by_slot, Mb, do_split, cnt, st, n_pre);
    if (do_split)
      launch_split(v, P0, cur, Mb, flag_by_slot, rank_by_slot, cnt, status, st, &bv, n_pre, grid);
    pend();
    if (flags & SPH_STEP_MERGE) {
      pbegin(SPH_P_MERGE);
      // flag_by_slot (the split's flags, consumed by now) is the qualification bits' scratch
      launch_merge(v, P0, cur, cstart, cend, nullptr, status, cnt, st, &bv, skeys, Mb, n_pre,
                   nullptr, flag_by_slot, phys_cap);
      pend();
      pbegin(SPH_P_COMPACT);
      const uint bound = (uint)phys_cap;  // entries after split: device count, host bound
      launch_alive_by_slot(cur, status, cnt, bound, alive_by_slot, st, &bv, grid);
      sph_exclusive_scan(alive_by_slot, newslot, &cnt->n, bound, &cnt->scan_total, scan_tmp, st);
      launch_batch_compact_plan(bv, newslot, cnt, st);
      launch_renumber(cur, status, newslot, cnt, bound, st, &bv, grid);
      pend();
    }
  }

  // The graph of one step with this shape, flags and state-buffer parity (captured on first
  // use). Returns nullptr if capture or instantiation failed.
  hipGraphExec_t graph_for(uint flags, const int shape[3]) {
    const uint key = (flags & 3u) | ((uint)shape[0] << 2) | ((uint)shape[1] << 4) |
                     ((uint)shape[2] << 6) | ((parity ? 1u : 0u) << 8) |
                     ((energy_unit_all() ? 1u : 0u) << 9);
    for (auto &g : graphs)
      if (g.key == key) return g.exec;
    const uint grid = kGridPerCU[shape[0]] * num_cu;
    const uint32_t mask = prof_mask;
    prof_mask = 0;  // no events inside the graph
    const DevState c0 = cur, a0 = alt;
    const bool p0 = parity;
    hipGraph_t g = nullptr;
    hipGraphExec_t ex = nullptr;
    hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      enqueue_step(flags, true, grid, shape[1], shape[2]);
      e = hipStreamEndCapture(st, &g);
    }
    // capture only recorded the step: the host's view of the buffers stays as it was
    cur = c0;
    alt = a0;
    parity = p0;
    prof_mask = mask;
    if (e == hipSuccess) e = hipGraphInstantiateWithFlags(&ex, g, 0);
    if (g) (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      last_err = e;
      return nullptr;
    }
    graphs.push_back({key, ex});
    return ex;
  }

  // `steps` graph steps, one synchronisation at the end (after the first step as well when
  // a one-step push is pending).
  int graph_steps(uint flags, uint32_t steps) {
    for (uint32_t k = 0; k < steps; k++) {
      int shape[3];
      ltune.next(shape);
      hipGraphExec_t ex = graph_for(flags, shape);
      if (!ex) return SPH_EHIP;
      hipEvent_t a = nullptr;
      ltune.before(st, shape, a);
      if (hipGraphLaunch(ex, st) != hipSuccess) return SPH_EHIP;
      ltune.after_launch(st, a);
      swap_state();  // the step's reorder moved the state to the other buffer
      if (pushes) {
        const int rc = sync();
        clear_pushes();
        if (rc) return rc;
      }
    }
    const int rc = sync();
    ltune.collect();
    return rc;
  }

  void clear_pushes() {
    if (!pushes) return;
    for (uint s = 0; s < ns; s++) {
      SphSysParam3D *r = (SphSysParam3D *)&raw[s * raw_size];
      r->flag0 = false;  // force_waiting(_left), SPH-3D-2/sph_system.cpp:403
      if (variant == SPH_V3D4) r->flag1 = false;  // force_waiting_right, SPH-3D-4 :451-452
      P[s].push_left = P[s].push_right = 0;
    }
    pushes = false;
    (void)upload_params();
  }
};

#define CHECK_BATCH(b)          \
  do {                          \
    if (!(b)) return SPH_EINVAL; \
  } while (0)

extern "C" {

int sph_batch_create(int variant, uint32_t n_scenes, const void *sysparams,
                     const uint32_t *capacity, SphBatch **out) {
  if (!out) return SPH_EINVAL;
  *out = nullptr;
  if ((variant != SPH_V3D2 && variant != SPH_V3D4) || n_scenes == 0 ||
      n_scenes > SPH_MAX_SCENES || !sysparams || !capacity)
    return SPH_EINVAL;
  SphBatch *b = new (std::nothrow) SphBatch();
  if (!b) return SPH_ECAP;
  b->variant = variant;
  b->ns = n_scenes;
  b->rec = sphhost::particle_size(variant);
  b->raw_size = sphhost::sysparam_size(variant);
  b->raw.assign((const unsigned char *)sysparams,
                (const unsigned char *)sysparams + b->raw_size * n_scenes);
  b->n_host.assign(n_scenes, 0u);
  size_t cells = 0, total = 0;
  for (uint s = 0; s < n_scenes; s++) {
    const uint c = capacity[s];
    if (c == 0 || c >= SPH_LOCAL_MASK) {
      delete b;
      return SPH_EINVAL;
    }
    DevParams p = sphhost::make_dev_params(variant, &b->raw[s * b->raw_size], c);
    if (p.total_cell == 0) {
      delete b;
      return SPH_EINVAL;
    }
    p.cell_base = (uint)cells;
    cells += p.total_cell;
    total += c;
    b->P.push_back(p);
    b->cap.push_back(c);
    b->max_cap = std::max(b->max_cap, c);
  }
  // cell ids are 32-bit keys; entry indices share a word with a 5-bit offset id in the
  // neighbour kernels' pair lists
  if (cells >= 0xffffffffull || total > (1u << 27)) {
    delete b;
    return SPH_EINVAL;
  }
  for (const DevParams &p : b->P) b->pushes = b->pushes || p.push_left || p.push_right;
  b->total_cell = (uint)cells;
  b->phys_cap = total;
  b->kbits = sphhost::key_bits(b->total_cell);
  if (hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return SPH_EHIP;
  }
  const size_t cap = b->phys_cap;
  const size_t scan_n = std::max<size_t>(cap, sph_sort_hist_elems(cap));
  bool ok = b->alloc_state(b->cur) && b->alloc_state(b->alt) && b->alloc(b->keys, cap) &&
            b->alloc(b->vals, cap) && b->alloc(b->sc.keys_alt, cap) && b->alloc(b->sc.vals_alt, cap) &&
            b->alloc(b->sc.tile_hist, sph_sort_hist_elems(cap)) &&
            b->alloc(b->s2p_hash, cap) && b->alloc(b->s2p_cell, cap) &&
            b->alloc(b->cstart, b->total_cell) && b->alloc(b->cend, b->total_cell) &&
            b->alloc(b->flag_by_slot, cap) && b->alloc(b->rank_by_slot, cap) &&
            b->alloc(b->alive_by_slot, cap) && b->alloc(b->newslot, cap) &&
            b->alloc(b->status, cap) && b->alloc(b->cnt, 1) && b->alloc(b->scn, n_scenes) &&
            b->alloc(b->Ptab, n_scenes) && b->alloc(b->aos_dev, (size_t)b->max_cap * b->rec) &&
            b->alloc(b->aux.frec, 3 * cap) && b->alloc(b->aux.erec, cap) &&
            b->alloc(b->aux.nbl, (size_t)SPH_NBL_K * cap) && b->alloc(b->aux.nbl_cnt, cap) &&
            b->alloc(b->scan_tmp, sph_scan_tmp_elems(scan_n)) && b->alloc(b->sc.len_dev, 1);
  b->aux.nbl_stride = (uint)cap;
  b->sc.scan_tmp = b->scan_tmp;
  b->sc.max_tiles = sph_sort_tiles(cap);
  if (ok && hipHostMalloc((void **)&b->cnt_host, sizeof(Counters), hipHostMallocDefault) != hipSuccess)
    ok = false;
  if (ok && hipHostMalloc((void **)&b->scn_host, sizeof(SceneCnt) * n_scenes,
                          hipHostMallocDefault) != hipSuccess)
    ok = false;
  if (!ok) {
    delete b;
    return SPH_ECAP;
  }
  if (!b->tune.create()) {
    delete b;
    return SPH_EHIP;
  }
  b->tune.start(kTuneRounds, kTuneAfter, kTuneEvery);
  b->ltune.start(kTuneRounds, kTuneAfter, kTuneEvery);
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess &&
      prop.multiProcessorCount > 0)
    b->num_cu = (uint)prop.multiProcessorCount;
  if (b->upload_params() != SPH_OK || b->push_counts() != SPH_OK || b->sync() != SPH_OK) {
    delete b;
    return SPH_EHIP;
  }
  *out = b;
  return SPH_OK;
}

int sph_batch_destroy(SphBatch *b) {
  CHECK_BATCH(b);
  delete b;
  return SPH_OK;
}

int sph_batch_set_params(SphBatch *b, uint32_t scene, const void *sysparam) {
  CHECK_BATCH(b);
  if (scene >= b->ns || !sysparam) return SPH_EINVAL;
  DevParams p = sphhost::make_dev_params(b->variant, sysparam, b->cap[scene]);
  if (p.total_cell != b->P[scene].total_cell) return SPH_EINVAL;
  p.cell_base = b->P[scene].cell_base;
  std::memcpy(&b->raw[scene * b->raw_size], sysparam, b->raw_size);
  b->P[scene] = p;
  if (p.push_left || p.push_right) b->pushes = true;
  if (b->upload_params() != SPH_OK) return SPH_EHIP;
  return b->sync();
}

int sph_batch_get_params(SphBatch *b, uint32_t scene, void *out) {
  CHECK_BATCH(b);
  if (scene >= b->ns || !out) return SPH_EINVAL;
  std::memcpy(out, &b->raw[scene * b->raw_size], b->raw_size);
  return SPH_OK;
}

int sph_batch_upload(SphBatch *b, uint32_t scene, const void *aos, uint32_t n) {
  CHECK_BATCH(b);
  if (scene >= b->ns || (n > 0 && !aos)) return SPH_EINVAL;
  if (n > b->cap[scene]) return SPH_ECAP;
  if (b->n_host[scene] > 0) launch_scene_kill(b->cur, b->m, scene, b->st);
  b->M -= b->n_host[scene];
  b->n_host[scene] = 0;
  if ((size_t)b->m + n > b->phys_cap) {  // drop removed entries to make room
    if (b->push_counts() != SPH_OK) return SPH_EHIP;
    b->pack();
  }
  if (n > 0) {
    if (hipMemcpyAsync(b->aos_dev, aos, b->rec * n, hipMemcpyHostToDevice, b->st) != hipSuccess)
      return SPH_EHIP;
    DevState at = b->cur;
    at.pos += b->m; at.vel += b->m; at.ev += b->m; at.acc += b->m;
    at.dp += b->m; at.energy += b->m; at.slot += b->m;
    launch_upload_scatter(b->variant, b->aos_dev, n, at, b->st, scene << SPH_SCENE_SHIFT);
  }
  b->n_host[scene] = n;
  b->M += n;
  b->m += n;
  if (b->push_counts() != SPH_OK) return SPH_EHIP;
  return b->sync();
}

int sph_batch_download(SphBatch *b, uint32_t scene, void *aos, uint32_t n) {
  CHECK_BATCH(b);
  if (scene >= b->ns || n > b->n_host[scene]) return SPH_EINVAL;
  if (n == 0) return SPH_OK;
  if (!aos) return SPH_EINVAL;
  launch_scene_slot2phys(b->cur, b->m, scene, b->newslot, b->st);
  launch_download_gather(b->variant, b->cur, b->newslot, n, b->aos_dev, b->st);
  if (hipMemcpyAsync(aos, b->aos_dev, b->rec * n, hipMemcpyDeviceToHost, b->st) != hipSuccess)
    return SPH_EHIP;
  return b->sync();
}

int sph_batch_num_particles(SphBatch *b, uint32_t *n_per_scene) {
  CHECK_BATCH(b);
  if (!n_per_scene) return SPH_EINVAL;
  for (uint s = 0; s < b->ns; s++) n_per_scene[s] = b->n_host[s];
  return SPH_OK;
}

int sph_batch_scene_counts(SphBatch *b, int32_t *out) {
  CHECK_BATCH(b);
  if (!out) return SPH_EINVAL;
  for (uint s = 0; s < b->ns; s++) {
    const SceneCnt &c = b->scn_host[s];
    out[4 * s + 0] = (int32_t)c.s_req;
    out[4 * s + 1] = (int32_t)c.s_done;
    out[4 * s + 2] = (int32_t)c.groups;
    out[4 * s + 3] = c.err ? SPH_EDOMAIN : SPH_OK;
  }
  return SPH_OK;
}

int sph_batch_step(SphBatch *b, uint32_t flags, uint32_t steps, uint32_t *n_total) {
  CHECK_BATCH(b);
  if (flags & ~(SPH_STEP_SPLIT | SPH_STEP_MERGE)) return SPH_EINVAL;
  int rc = SPH_OK;
  if (b->use_graph && steps > 0) {
    // merges never empty a scene and splits only add, so a batch with particles keeps them
    if (b->M == 0) return SPH_EEMPTY;
    rc = b->graph_steps(flags, steps);
    if (rc) return rc;
    if (n_total) *n_total = b->M;
    return b->cnt_host->err ? SPH_EDOMAIN : SPH_OK;
  }
  for (uint32_t k = 0; k < steps; k++) {
    if (b->M == 0) return SPH_EEMPTY;
    b->enqueue_step(flags);
    rc = b->sync();
    b->clear_pushes();
    if (rc) return rc;
    b->tune.end_step();
    if (b->cnt_host->err) rc = SPH_EDOMAIN;
  }
  if (n_total) *n_total = b->M;
  return rc;
}

int sph_batch_autotune(SphBatch *b, uint32_t rounds, uint32_t after_steps, uint32_t every_steps) {
  CHECK_BATCH(b);
  b->tune.start(rounds, after_steps, every_steps);
  b->ltune.start(rounds, after_steps, every_steps);
  return SPH_OK;
}

int sph_batch_set_lists(SphBatch *b, uint32_t force_list, uint32_t energy_list) {
  CHECK_BATCH(b);
  if (!b->tune.set(force_list, energy_list)) return SPH_EINVAL;
  b->ltune.cur[1] = b->tune.force;
  b->ltune.cur[2] = b->tune.energy;
  b->ltune.rounds_left = 0;
  return SPH_OK;
}

int sph_batch_tuned_lists(SphBatch *b, uint32_t *force_list, uint32_t *energy_list, float *best_ms) {
  CHECK_BATCH(b);
  const bool g = b->use_graph;
  if (force_list) *force_list = (uint32_t)kForceLists[g ? b->ltune.cur[1] : b->tune.force];
  if (energy_list) *energy_list = (uint32_t)kEnergyLists[g ? b->ltune.cur[2] : b->tune.energy];
  if (best_ms)
    for (int k = 0; k < 2; k++)
      for (int c = 0; c < 3; c++) best_ms[k * 3 + c] = g ? b->ltune.best[1 + k][c] : b->tune.best[k][c];
  return SPH_OK;
}

int sph_batch_set_graph(SphBatch *b, int on) {
  CHECK_BATCH(b);
  b->use_graph = on != 0;
  return SPH_OK;
}

int sph_batch_set_launch(SphBatch *b, uint32_t groups_per_cu) {
  CHECK_BATCH(b);
  for (int c = 0; c < 4; c++)
    if (kGridPerCU[c] == groups_per_cu) {
      b->ltune.cur[0] = c;
      b->ltune.rounds_left = 0;
      return SPH_OK;
    }
  return SPH_EINVAL;
}

int sph_batch_tuned_launch(SphBatch *b, uint32_t *groups_per_cu, uint32_t *num_cu, float *best_ms) {
  CHECK_BATCH(b);
  if (groups_per_cu) *groups_per_cu = kGridPerCU[b->ltune.cur[0]];
  if (num_cu) *num_cu = b->num_cu;
  if (best_ms)
    for (int c = 0; c < 4; c++) best_ms[c] = b->ltune.best[0][c];
  return SPH_OK;
}

int sph_batch_prof_enable(SphBatch *b, int on) {
  CHECK_BATCH(b);
  b->prof_mask = on ? (1u << SPH_P_COUNT) - 1u : 0u;
  return SPH_OK;
}

int sph_batch_prof_phases(SphBatch *b, uint32_t phase_mask) {
  CHECK_BATCH(b);
  b->prof_mask = phase_mask & ((1u << SPH_P_COUNT) - 1u);
  return SPH_OK;
}

int sph_batch_prof_reset(SphBatch *b) {
  CHECK_BATCH(b);
  int rc = b->sync();
  for (int k = 0; k < SPH_P_COUNT; k++) {
    b->prof_ms[k] = 0.0;
    b->prof_n[k] = 0;
  }
  return rc;
}

int sph_batch_prof_read(SphBatch *b, int phase, double *ms_total, uint64_t *launches) {
  CHECK_BATCH(b);
  if (phase < 0 || phase >= SPH_P_COUNT) return SPH_EINVAL;
  if (ms_total) *ms_total = b->prof_ms[phase];
  if (launches) *launches = b->prof_n[phase];
  return SPH_OK;
}

}  // extern "C"
