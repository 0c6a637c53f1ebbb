// Batched, continuously admitted IB-LBM cases on one device (compiled as
// part of ilbm_level.hip).
//
// Every admitted case is a resident Level (its own populations, ghost slabs,
// IB band, markers and histories).  A step of the whole batch is the single
// Level step with each kernel launched ONCE for all cases of a lattice /
// storage class: blocks look their case up in a BlockRef map and read the
// case's arguments from a device table.  The step sequence of every class is
// captured in a hipGraph covering `graph_steps` steps, replayed until the
// next admission or retirement changes the batch.
//
// Cases advance in lock step but each stops at its own step count: a call
// advances the batch by at most the smallest remaining count, so no case
// runs past its request.  Results do not depend on the mix: every kernel
// body is the one the single Level runs, with the same per-case block
// decomposition and fixed-order reductions.
#pragma once

#include <map>
#include <memory>

namespace ilbm {

using StepBFn = void (*)(dim3, hipStream_t, const StepArgs*, const BlockRef*);
using GhostBFn = void (*)(dim3, hipStream_t, const GhostArgs*, const BlockRef*);
using FixBFn = void (*)(dim3, hipStream_t, const IbCase*, const BlockRef*);
using GhostRecBFn = void (*)(unsigned, hipStream_t, const GhostArgs*, const BlockRef*, unsigned, const IbCase*,
                             const BlockRef*);
using StepRecBFn = void (*)(unsigned, hipStream_t, const StepArgs*, const BlockRef*, unsigned, const IbCase*,
                            const BlockRef*);

template <class LAT, class ST, int AM>
static void launch_step_b(dim3 g, hipStream_t s, const StepArgs* a, const BlockRef* m) {
  hipLaunchKernelGGL((k_step_b<LAT, ST, AM>), g, dim3(kBlock), 0, s, a, m);
}
template <class LAT, class ST, int AM>
static void launch_step_rec_b(unsigned grid, hipStream_t s, const StepArgs* a, const BlockRef* m, unsigned nstep,
                              const IbCase* ia, const BlockRef* rm) {
  hipLaunchKernelGGL((k_step_record_b<LAT, ST, AM>), dim3(grid), dim3(kBlock), 0, s, a, m, nstep, ia, rm);
}
template <class LAT, class ST>
static void launch_ghost_b(dim3 g, hipStream_t s, const GhostArgs* a, const BlockRef* m) {
  hipLaunchKernelGGL((k_fill_ghosts_b<LAT, ST>), g, dim3(kBlock), 0, s, a, m);
}
template <class LAT, class ST>
static void launch_fix_b(dim3 g, hipStream_t s, const IbCase* a, const BlockRef* m) {
  hipLaunchKernelGGL((k_ib_export_fixup_b<LAT, ST>), g, dim3(256), 0, s, a, m);
}

template <class LAT, class ST>
static void launch_fixw_b(dim3 g, hipStream_t s, const IbCase* a, const BlockRef* m) {
  hipLaunchKernelGGL((k_ib_spread_export_w_b<LAT, ST>), g, dim3(64 * kSpreadWaves), 0, s, a, m);
}

template <class LAT, class ST>
static void launch_ghost_rec_b(unsigned grid, hipStream_t s, const GhostArgs* ga, const BlockRef* gm, unsigned nghost,
                               const IbCase* a, const BlockRef* rm) {
  hipLaunchKernelGGL((k_ghosts_record_b<LAT, ST>), dim3(grid), dim3(kSumsBlock), 0, s, ga, gm, nghost, a, rm);
}

// the fused slab + record launch of a class (lattice, storage)
static GhostRecBFn pick_ghost_rec_b(int lat, int st) {
  GhostRecBFn f = nullptr;
  auto pk = [&](auto latt) {
    using L = decltype(latt);
    f = (st == 0) ? launch_ghost_rec_b<L, StF64> : (st == 1) ? launch_ghost_rec_b<L, StF32> : launch_ghost_rec_b<L, StF16S>;
  };
  if (lat == 0) pk(D2Q9{});
  else if (lat == 1) pk(D3Q19{});
  else pk(D3Q27{});
  return f;
}

template <class LAT, class ST>
static void pick_b(int am, StepBFn& sf, GhostBFn& gf, FixBFn& ff, FixBFn& fw, StepRecBFn* sr) {
  sf = (am == kAddrWrap) ? launch_step_b<LAT, ST, kAddrWrap>
     : (am == kAddrWallX32) ? launch_step_b<LAT, ST, kAddrWallX32>
     : (am == kAddrWall32) ? launch_step_b<LAT, ST, kAddrWall32> : launch_step_b<LAT, ST, kAddrWall64>;
  if (sr)
    *sr = (am == kAddrWrap) ? launch_step_rec_b<LAT, ST, kAddrWrap>
        : (am == kAddrWallX32) ? launch_step_rec_b<LAT, ST, kAddrWallX32>
        : (am == kAddrWall32) ? launch_step_rec_b<LAT, ST, kAddrWall32> : launch_step_rec_b<LAT, ST, kAddrWall64>;
  gf = launch_ghost_b<LAT, ST>;
  ff = launch_fix_b<LAT, ST>;
  fw = launch_fixw_b<LAT, ST>;
}

static void pick_class(int lat, int st, int am, StepBFn& sf, GhostBFn& gf, FixBFn& ff, FixBFn& fw,
                       StepRecBFn* sr = nullptr) {
  auto pk = [&](auto latt) {
    using L = decltype(latt);
    if (st == 0) pick_b<L, StF64>(am, sf, gf, ff, fw, sr);
    else if (st == 1) pick_b<L, StF32>(am, sf, gf, ff, fw, sr);
    else pick_b<L, StF16S>(am, sf, gf, ff, fw, sr);
  };
  if (lat == 0) pk(D2Q9{});
  else if (lat == 1) pk(D3Q19{});
  else pk(D3Q27{});
}

// Tiling of a case's batched stream-collide launch: tile width (0: linear
// tiles of consecutive cells) and cells per thread.
struct Tile {
  int bx = 64;
  int nc = 1;
};

struct BatchCase {
  int id = -1;
  std::unique_ptr<Level> L;
  int target = 0, done = 0;
  int64_t bytes = 0;
  int tile_bx = 64;
  int tile_nc = 1;
  int base_cur = 0;         // L->cur when the batch tables were built
};

// Device tables of one (lattice, storage, layout) class.
struct BatchGroup {
  int lat = 0, st = 0;
  int am = kAddrWrap;  // stream-collide addressing of the class's cases
  int lane = 0;        // stream the group's step chain runs on (0: stream, 1: stream2)
  StepBFn step = nullptr;
  GhostBFn ghost = nullptr;
  GhostRecBFn ghost_rec = nullptr;  // slabs and force records in one launch
  StepRecBFn step_rec = nullptr;    // stream-collide and force records in one launch (no slab launch)
  FixBFn fix = nullptr;
  FixBFn fixw = nullptr;
  std::vector<int> members;  // indices into Batch::cases
  StepArgs* dstep[2] = {nullptr, nullptr};
  GhostArgs* dghost[2] = {nullptr, nullptr};
  IbCase* dib[2] = {nullptr, nullptr};
  CaseTail* dtail = nullptr;
  BlockRef* dmap = nullptr;
  // block maps inside dmap: offset and count of each kernel's grid
  enum { M_STEP, M_GHOST, M_BAND, M_INTERP, M_SPREAD, M_RHS, M_MATVEC, M_IBCASE, M_TAIL, M_FIX, M_FIXW, M_N };
  int64_t moff[M_N] = {}, mcnt[M_N] = {};
  int max_loop = 0;
  bool any_ivc = false, any_iter = false;

  void release() {
    for (int p = 0; p < 2; ++p) { dev_free(dstep[p]); dev_free(dghost[p]); dev_free(dib[p]); }
    dev_free(dtail); dev_free(dmap);
    for (int p = 0; p < 2; ++p) { dstep[p] = nullptr; dghost[p] = nullptr; dib[p] = nullptr; }
    dtail = nullptr; dmap = nullptr;
  }
};

// Device bytes a case holds while resident (what admission charges).
static int64_t case_footprint(const ilbm_case_desc& c0, int nsteps) {
  ilbm_case_desc c = c0;
  if (c.dim == 2) c.n[2] = 1;
  const int Q = c.q;
  const int64_t cells = (int64_t)c.n[0] * c.n[1] * c.n[2];
  const int pb = (c.storage == 0) ? 8 : (c.storage == 1) ? 4 : 2;
  const int64_t sl = row_stride(c.dim, c.n[0]) * c.n[1];
  const int64_t qs = slice_layout(c.dim, cells) ? ((sl * pb + 16383) / 16384 * 16384 + 8192) / pb * c.n[2]
                                                 : plane_stride(sl * c.n[2], pb);
  int64_t slab = 0, faces = 0;
  for (int d = 0; d < c.dim; ++d) {
    if (c.periodic[d]) continue;
    int64_t f = 1;
    for (int e = 0; e < c.dim; ++e)
      if (e != d) f *= (int64_t)c.n[e] + 2;
    slab += 2 * f;
    faces += 2 * f;
  }
  int64_t b = 2 * (int64_t)pb * Q * (qs + slab) + (int64_t)pb * face_stride(Q, pb) * faces;
  b += 8 * (cells / 64 + 64);                          // block partials
  if (c.use_ib) {
    // band: stencil box of the body, markers and their lists
    const double ext = 2.0 * c.R / c.dx[0] + 7.0;
    int64_t band = 1, nm;
    for (int d = 0; d < c.dim; ++d) band *= std::min<int64_t>((int64_t)std::ceil(ext), c.n[d]);
    if (c.dim == 2) nm = (int64_t)std::ceil(2.0 * kPi * c.R / c.dx[0]) + 16;
    else nm = (int64_t)std::ceil(4.0 * kPi * c.R * c.R / (c.dx[0] * c.dx[0])) + 32;
    if (c.n_marker > 0) nm = std::max<int64_t>(nm, c.n_marker);
    b += band * 8 * 16 + band * 4 + nm * 8 * 40;
    b += nm * (c.dim == 3 ? 125 : 25) * 16;              // spreading entries
    if (c.coupling_method == 1) b += nm * nm * 8;
    b += (int64_t)nsteps * 8 * ((c.dim == 3) ? 33 : 23);  // force history
  }
  b += (int64_t)nsteps * 8;                               // mass history
  return b;
}

// Autotuned tile width per (lattice, storage, layout, x-extent bucket),
// shared by every batch of the process so a case gets the same tiling
// whichever batch it joins.
static std::map<int64_t, Tile>& tile_table() {
  static std::map<int64_t, Tile> t;
  return t;
}

// two-lane batches run their lanes in phase (Batch::lanes_together)
constexpr bool kLanesTogether = false;

struct Batch {
  hipStream_t stream = nullptr;
  // A class of two or more cases is split into two lanes of balanced cell
  // counts; each lane's whole step chain (IB kernels, slab fill,
  // stream-collide, tail) runs on its own stream, so one lane's IB kernels
  // overlap the other lane's stream-collide.  The lanes share nothing within
  // a step (cases are independent); advance forks stream2 from stream and
  // joins it back, so every call returns with both ordered behind stream.
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // after a fork, lane 1 starts once lane 0's first IB kernels are done: the
  // lanes then run half a step apart (each lane's IB beside the other's
  // stream-collide) instead of in phase
  hipEvent_t ev_phase = nullptr;
  bool phase_pending = false;
  // lanes in phase (ILBM_BATCH_PHASE=together): both lanes' IB kernels of a
  // step first, then both stream-collide launches once both are done, so no
  // IB kernel waits behind the dispatch of the other lane's stream-collide
  bool lanes_together = kLanesTogether;
  hipEvent_t ev_ib[2] = {nullptr, nullptr};
  // two lanes when the active cases hold at least 2^22 cells: a step is then
  // long enough for eager launches, and a graph replay would run the two
  // lanes' branches one after the other on one queue (measured), so two-lane
  // batches advance eagerly and one-lane batches replay graphs
  static constexpr int64_t kTwoLaneCells = (int64_t)1 << 22;
  int nlanes = 1;
  int64_t budget = 0, used = 0;
  int max_cases = 0;
  int graph_steps = 8;
  int next_id = 1;
  std::vector<BatchCase> cases;
  std::vector<BatchGroup> groups;
  bool dirty = true;
  int phase = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  bool graph_failed = false;
  // statistics
  int64_t graph_launches = 0, captures = 0, eager_steps = 0;
  double lups = 0.0;

  ~Batch() {
    if (stream2) (void)hipStreamSynchronize(stream2);
    if (stream) (void)hipStreamSynchronize(stream);
    drop_graph();
    for (auto& g : groups) g.release();
    cases.clear();
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (ev_phase) (void)hipEventDestroy(ev_phase);
    for (hipEvent_t e : ev_ib)
      if (e) (void)hipEventDestroy(e);
    if (stream2) (void)hipStreamDestroy(stream2);
    if (stream) (void)hipStreamDestroy(stream);
  }

  // stream2 continues after stream's work so far / stream after stream2's
  void fork() {
    HIPCHK(hipEventRecord(ev_fork, stream));
    HIPCHK(hipStreamWaitEvent(stream2, ev_fork, 0));
    phase_pending = true;
  }
  void join() {
    HIPCHK(hipEventRecord(ev_join, stream2));
    HIPCHK(hipStreamWaitEvent(stream, ev_join, 0));
  }

  void drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
  }

  BatchCase* find(int id) {
    for (auto& c : cases)
      if (c.id == id) return &c;
    return nullptr;
  }

  // ---------------- autotuner ----------------
  static int bucket(int n0) {
    int b = 0;
    while ((64 << b) < n0 && b < 6) ++b;
    return b;
  }

  // Tiling of the stream-collide launch for a case of this class and x
  // extent: linear tiles (0) and the widths 64 / 128 / 256, with one and
  // (2D) two cells per thread, are timed on a scratch level of the case's x
  // extent (no IB) and the fastest is kept for the process, so every later
  // case of the bucket uses the same tiling.
  Tile tile_for(const Level& ref) {
    // ILBM_BATCH_TILE=0|64|128|256 and ILBM_BATCH_NC=1|2 fix the tiling
    // (tests of each path); a fixed width alone keeps one cell per thread
    Tile fixed{};
    bool have_fixed = false;
    if (const char* fix = std::getenv("ILBM_BATCH_TILE")) {
      const int t = std::atoi(fix);
      if (t == 0 || t == 64 || t == 128 || t == 256) { fixed.bx = t; have_fixed = true; }
    }
    int fix_nc = 0;
    if (const char* fn = std::getenv("ILBM_BATCH_NC")) {
      const int v = std::atoi(fn);
      if (v == 1 || v == 2) fix_nc = v;
    }
    if (ref.c.dim == 3) fix_nc = 0;   // one cell per thread on 3D lattices
    if (have_fixed) {
      fixed.nc = fix_nc ? fix_nc : 1;
      return fixed;
    }
    const int64_t key = ((((int64_t)ref.lat * 4 + ref.st) * 4 + ref.am) * 16 + bucket(ref.c.n[0])) * 4 + fix_nc;
    auto& tiles = tile_table();
    auto it = tiles.find(key);
    if (it != tiles.end()) return it->second;
    std::vector<Tile> cand;
    for (int nc = 1; nc <= (ref.c.dim == 2 ? 2 : 1); ++nc) {
      if (fix_nc && nc != fix_nc && ref.c.dim == 2) continue;
      cand.push_back(Tile{0, nc});
      for (int bx = 64; bx <= 256; bx *= 2) {
        cand.push_back(Tile{bx, nc});
        if (bx >= ref.c.n[0]) break;
      }
    }
    Tile best = cand[0];
    if (cand.size() > 1) {
      // the batched launch is bound by throughput, not by one case's
      // latency: the scratch level keeps the case's x extent and grows in
      // its outer dimension to at least 2^22 cells
      Level T;
      T.c = ref.c;
      T.c.use_ib = 0;
      T.c.use_prescribed_force = 0;
      {
        const int outer = (T.c.dim == 3) ? 2 : 1;
        const int64_t plane = (int64_t)ref.ncells / std::max(1, T.c.n[outer]);
        const int64_t want = ((int64_t)1 << 22) / std::max<int64_t>(1, plane);
        if (want > T.c.n[outer]) T.c.n[outer] = (int)std::min<int64_t>(want, 1 << 16);
      }
      T.stream = stream;
      T.own_stream = false;
      T.validate();
      T.setup_layout();
      std::vector<double> m0(4 * (size_t)T.ncells, 0.0);
      std::fill(m0.begin(), m0.begin() + T.ncells, 1.0);
      T.init(m0.data());
      int64_t maxblk = 0;
      for (const Tile& t : cand) maxblk = std::max(maxblk, T.tile_blocks(t.bx, t.nc));
      T.ensure_partial(maxblk);
      StepBFn sf; GhostBFn gf; FixBFn ff, fw;
      pick_class(T.lat, T.st, T.am, sf, gf, ff, fw);
      StepArgs* da = dev_alloc<StepArgs>(1);
      BlockRef* dm = dev_alloc<BlockRef>(maxblk);
      hipEvent_t e0, e1;
      HIPCHK(hipEventCreate(&e0));
      HIPCHK(hipEventCreate(&e1));
      double best_ms = 1e30;
      for (const Tile& t : cand) {
        StepArgs a = T.step_args(false);
        a.tile_bx = t.bx;
        a.tile_nc = t.nc;
        const int64_t nb = T.tile_blocks(t.bx, t.nc);
        std::vector<BlockRef> m(nb);
        for (int64_t b = 0; b < nb; ++b) m[b] = step_ref(T.c.n, t.bx, t.nc, 0, b);
        HIPCHK(hipMemcpyAsync(da, &a, sizeof(a), hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(dm, m.data(), sizeof(BlockRef) * nb, hipMemcpyHostToDevice, stream));
        for (int r = 0; r < 3; ++r) sf(dim3((unsigned)nb), stream, da, dm);
        HIPCHK(hipEventRecord(e0, stream));
        const int reps = 10;
        for (int r = 0; r < reps; ++r) sf(dim3((unsigned)nb), stream, da, dm);
        HIPCHK(hipEventRecord(e1, stream));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best_ms) { best_ms = ms; best = t; }
      }
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      dev_free(da);
      dev_free(dm);
    }
    tiles[key] = best;
    return best;
  }

  // A case whose stream-collide kernel writes its own ghost slabs gets them
  // from the boundary layer once, for its first batched step, unless they
  // are current (restored from a checkpoint taken after a step).
  void prime_slabs(Level& L) {
    if (L.wrap || !L.gfuse || L.slabs_fresh) return;
    const int64_t npad = L.slabs.face_off[L.slabs.nface];
    L.ghostfn(dim3((unsigned)((npad + kGhostBlock - 1) / kGhostBlock)), dim3(kGhostBlock), stream, L.ghost_args());
    HIPCHK(hipGetLastError());
    L.slabs_fresh = true;
  }

  // ---------------- admission ----------------
  int add(const ilbm_case_desc& desc, const double* init, int nsteps) {
    if (nsteps <= 0) throw StatusError{ILBM_ERR_INVALID, "nsteps must be positive"};
    if (desc.use_ib && desc.geometry_type == 2)
      throw StatusError{ILBM_ERR_UNSUPPORTED, "moving bodies are not batched"};
    const int64_t bytes = case_footprint(desc, nsteps);
    if ((int)cases.size() >= max_cases || used + bytes > budget)
      throw StatusError{ILBM_ERR_FULL, "batch full (case slots or HBM budget)"};
    BatchCase bc;
    bc.L.reset(new Level());
    Level& L = *bc.L;
    L.c = desc;
    L.stream = stream;
    L.own_stream = false;
    L.validate();
    L.setup_layout();
    L.setup_ib();
    L.init(init);
    prime_slabs(L);
    const Tile t = tile_for(L);
    bc.tile_bx = t.bx;
    bc.tile_nc = t.nc;
    L.ensure_partial(L.tile_blocks(t.bx, t.nc));
    L.grow_histories(nsteps, nsteps);
    const double clk[4] = {L.t_new, (double)L.istep, (double)L.istep, 0.0};
    HIPCHK(hipMemcpyAsync(L.clock_d, clk, sizeof(clk), hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
    bc.id = next_id++;
    bc.target = nsteps;
    bc.bytes = bytes;
    used += bytes;
    cases.push_back(std::move(bc));
    dirty = true;
    return cases.back().id;
  }

  // a case admitted from a checkpoint of the same case description
  // (preempted and resumed); nsteps more steps
  int add_from_checkpoint(const ilbm_case_desc& desc, const char* buf, int64_t size, int nsteps) {
    if (nsteps <= 0) throw StatusError{ILBM_ERR_INVALID, "nsteps must be positive"};
    if (desc.use_ib && desc.geometry_type == 2)
      throw StatusError{ILBM_ERR_UNSUPPORTED, "moving bodies are not batched"};
    const int64_t bytes = case_footprint(desc, nsteps);
    if ((int)cases.size() >= max_cases || used + bytes > budget)
      throw StatusError{ILBM_ERR_FULL, "batch full (case slots or HBM budget)"};
    BatchCase bc;
    bc.L.reset(new Level());
    Level& L = *bc.L;
    L.c = desc;
    L.stream = stream;
    L.own_stream = false;
    L.validate();
    L.setup_layout();
    L.setup_ib();
    std::vector<double> m0(4 * (size_t)L.ncells, 0.0);
    std::fill(m0.begin(), m0.begin() + L.ncells, 1.0);
    L.init(m0.data());
    L.restore(buf, size);
    prime_slabs(L);
    const Tile t = tile_for(L);
    bc.tile_bx = t.bx;
    bc.tile_nc = t.nc;
    L.ensure_partial(L.tile_blocks(t.bx, t.nc));
    L.grow_histories(nsteps, nsteps);
    HIPCHK(hipStreamSynchronize(stream));
    bc.id = next_id++;
    bc.target = nsteps;
    bc.bytes = bytes;
    used += bytes;
    cases.push_back(std::move(bc));
    dirty = true;
    return cases.back().id;
  }

  void remove(int id) {
    for (size_t i = 0; i < cases.size(); ++i) {
      if (cases[i].id != id) continue;
      HIPCHK(hipStreamSynchronize(stream));
      used -= cases[i].bytes;
      cases.erase(cases.begin() + i);
      dirty = true;
      return;
    }
    throw StatusError{ILBM_ERR_INVALID, "unknown case id"};
  }

  // ---------------- tables ----------------
  void rebuild() {
    HIPCHK(hipStreamSynchronize(stream));
    drop_graph();
    graph_failed = false;
    for (auto& g : groups) g.release();
    groups.clear();
    phase = 0;
    // lanes: within a class, cases by decreasing cell count go to the lane
    // with fewer cells so far (a class of one case stays on lane 0)
    std::vector<int> lane_of(cases.size(), 0);
    int64_t active_cells = 0;
    for (const BatchCase& bc : cases)
      if (bc.done < bc.target) active_cells += bc.L->ncells;
    nlanes = (active_cells >= kTwoLaneCells) ? 2 : 1;
    // ILBM_BATCH_LANES=1 keeps every case on one lane (measurement of the
    // lane split, and graph replays of large batches)
    if (const char* nl = std::getenv("ILBM_BATCH_LANES"))
      if (nl[0] == '1') nlanes = 1;
    if (nlanes == 2) {
      std::map<int64_t, std::vector<int>> cls;
      for (size_t i = 0; i < cases.size(); ++i) {
        const BatchCase& bc = cases[i];
        if (bc.done >= bc.target) continue;
        const Level& L = *bc.L;
        cls[((int64_t)L.lat * 4 + L.st) * 4 + L.am].push_back((int)i);
      }
      for (auto& kv : cls) {
        std::vector<int>& v = kv.second;
        std::stable_sort(v.begin(), v.end(), [&](int a, int b) { return cases[a].L->ncells > cases[b].L->ncells; });
        int64_t load[2] = {0, 0};
        for (int i : v) {
          const int ln = (v.size() > 1 && load[1] < load[0]) ? 1 : 0;
          lane_of[i] = ln;
          load[ln] += cases[i].L->ncells;
        }
      }
    }
    for (size_t i = 0; i < cases.size(); ++i) {
      BatchCase& bc = cases[i];
      if (bc.done >= bc.target) continue;  // finished cases stay resident until removed
      Level& L = *bc.L;
      bc.base_cur = L.cur;
      int gi = -1;
      for (size_t k = 0; k < groups.size(); ++k)
        if (groups[k].lat == L.lat && groups[k].st == L.st && groups[k].am == L.am && groups[k].lane == lane_of[i])
          gi = (int)k;
      if (gi < 0) {
        BatchGroup g;
        g.lat = L.lat; g.st = L.st; g.am = L.am; g.lane = lane_of[i];
        pick_class(g.lat, g.st, g.am, g.step, g.ghost, g.fix, g.fixw, &g.step_rec);
        g.ghost_rec = pick_ghost_rec_b(g.lat, g.st);
        groups.push_back(g);
        gi = (int)groups.size() - 1;
      }
      groups[gi].members.push_back((int)i);
    }
    for (auto& g : groups) build_group(g);
    dirty = false;
  }

  void build_group(BatchGroup& G) {
    const size_t n = G.members.size();
    std::vector<StepArgs> sa[2];
    std::vector<GhostArgs> ga[2];
    std::vector<IbCase> ia[2];
    std::vector<CaseTail> tails(n);
    std::vector<BlockRef> maps[BatchGroup::M_N];
    for (int p = 0; p < 2; ++p) { sa[p].resize(n); ga[p].resize(n); ia[p].resize(n); }
    G.max_loop = 0;
    for (size_t ci = 0; ci < n; ++ci) {
      BatchCase& bc = cases[G.members[ci]];
      Level& L = *bc.L;
      const int c = (int)ci;
      for (int p = 0; p < 2; ++p) {
        L.cur = bc.base_cur ^ p;
        sa[p][ci] = L.step_args(false);
        sa[p][ci].tile_bx = bc.tile_bx;
        sa[p][ci].tile_nc = bc.tile_nc;
        ga[p][ci] = L.ghost_args();
        IbCase ic{};
        if (L.ib) {
          ic.ib = L.ib_args();
          ic.g = L.g;
          ic.gbuf = L.gbuf[L.cur];
          ic.tau = L.tau;
          RecArgs& r = ic.rec;
          r.partial = L.fpartial; r.sums = L.sums_d; r.state = L.frstate_d;
          r.hist = L.force_dev; r.row = 0; r.ncol = L.ncol; r.method = L.c.force_eval_method;
          r.dV = (L.c.dim == 3) ? L.c.dx[0] * L.c.dx[1] * L.c.dx[2] : L.c.dx[0] * L.c.dx[1];
          r.U0 = L.c.U0; r.R = L.c.R;
          r.clock = L.clock_d;
          ic.nblk_fix = (int)L.fix_blocks();
        }
        ia[p][ci] = ic;
      }
      L.cur = bc.base_cur;
      const int64_t nb = L.tile_blocks(bc.tile_bx, bc.tile_nc);
      tails[ci] = CaseTail{L.partial, nb, L.mass_dev, L.clock_d, L.dt};
      for (int64_t b = 0; b < nb; ++b) maps[BatchGroup::M_STEP].push_back(step_ref(L.c.n, bc.tile_bx, bc.tile_nc, c, b));
      if (!L.wrap && !L.gfuse) {   // else the stream-collide kernel writes the slabs
        const int64_t npad = L.slabs.face_off[L.slabs.nface];
        for (int64_t b = 0; b < (npad + kBlock - 1) / kBlock; ++b) maps[BatchGroup::M_GHOST].push_back({c, (int)b});
      }
      maps[BatchGroup::M_TAIL].push_back({c, 0});
      if (L.ib) {
        const IbArgs& a = ia[0][ci].ib;
        for (int64_t b = 0; b < (L.band_cells + 255) / 256; ++b) maps[BatchGroup::M_BAND].push_back({c, (int)b});
        for (int b = 0; b < interp_blocks(L.nm, a.tree != 0, L.c.dim); ++b)
          maps[BatchGroup::M_INTERP].push_back({c, b});
        if (a.nloop > 1)
          for (int b = 0; b < (L.n_ne + 255) / 256; ++b) maps[BatchGroup::M_SPREAD].push_back({c, b});
        const int fk = (L.fix_mode() == 3) ? BatchGroup::M_FIXW : BatchGroup::M_FIX;
        for (unsigned b = 0; b < L.fix_blocks(); ++b) maps[fk].push_back({c, (int)b});
        if (a.use_ivc) {
          for (int b = 0; b < (L.nm + 255) / 256; ++b) maps[BatchGroup::M_RHS].push_back({c, b});
          for (int b = 0; b < (L.nm + 3) / 4; ++b) maps[BatchGroup::M_MATVEC].push_back({c, b});
          G.any_ivc = true;
        }
        if (!a.ev_from_macro) G.any_iter = true;
        maps[BatchGroup::M_IBCASE].push_back({c, 0});
        G.max_loop = std::max(G.max_loop, a.nloop);
      }
    }
    int64_t tot = 0;
    for (int m = 0; m < BatchGroup::M_N; ++m) {
      G.moff[m] = tot;
      G.mcnt[m] = (int64_t)maps[m].size();
      tot += G.mcnt[m];
    }
    std::vector<BlockRef> all;
    all.reserve(tot);
    for (int m = 0; m < BatchGroup::M_N; ++m) all.insert(all.end(), maps[m].begin(), maps[m].end());
    for (int64_t m = 0; m < BatchGroup::M_N; ++m)
      if (G.mcnt[m] > 0x7fffffff) throw StatusError{ILBM_ERR_UNSUPPORTED, "batched grid too large"};
    G.dmap = dev_alloc<BlockRef>(std::max<int64_t>(tot, 1));
    HIPCHK(hipMemcpyAsync(G.dmap, all.data(), sizeof(BlockRef) * tot, hipMemcpyHostToDevice, stream));
    for (int p = 0; p < 2; ++p) {
      G.dstep[p] = dev_alloc<StepArgs>(n);
      G.dghost[p] = dev_alloc<GhostArgs>(n);
      G.dib[p] = dev_alloc<IbCase>(n);
      HIPCHK(hipMemcpyAsync(G.dstep[p], sa[p].data(), sizeof(StepArgs) * n, hipMemcpyHostToDevice, stream));
      HIPCHK(hipMemcpyAsync(G.dghost[p], ga[p].data(), sizeof(GhostArgs) * n, hipMemcpyHostToDevice, stream));
      HIPCHK(hipMemcpyAsync(G.dib[p], ia[p].data(), sizeof(IbCase) * n, hipMemcpyHostToDevice, stream));
    }
    G.dtail = dev_alloc<CaseTail>(n);
    HIPCHK(hipMemcpyAsync(G.dtail, tails.data(), sizeof(CaseTail) * n, hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));  // host tables go out of scope
  }

  // ---------------- one batched step ----------------
  // The single Level's order: IB forcing with the export / Guo fix-up, the
  // force records with the ghost slabs (one launch when the group has
  // slabs), stream-collide, Sum(Rho) and the clock.
  // One step of a group on its lane's stream; lane0_ib: set once lane 0's
  // first IB kernels after a fork are queued (and ev_phase recorded)
  void enqueue_group(BatchGroup& G, int p, hipStream_t stream, bool& lane0_ib) {
    enqueue_group_ib(G, p, stream, lane0_ib);
    enqueue_group_lbm(G, p, stream);
  }

  // the IB forcing with the export / Guo fix-up and the force records of a
  // group without slabs
  void enqueue_group_ib(BatchGroup& G, int p, hipStream_t stream, bool& lane0_ib) {
    const BlockRef* m = G.dmap;
    auto grid = [&](int k) { return dim3((unsigned)G.mcnt[k]); };
    if (G.mcnt[BatchGroup::M_IBCASE] > 0) {
      const IbCase* A = G.dib[p];
      if (G.any_iter)
        hipLaunchKernelGGL(k_ib_begin_b, grid(BatchGroup::M_BAND), dim3(256), 0, stream, A,
                           m + G.moff[BatchGroup::M_BAND]);
      for (int it = 0; it < G.max_loop; ++it) {
        hipLaunchKernelGGL(k_ib_interp_b, grid(BatchGroup::M_INTERP), dim3(64 * kInterpWaves), 0, stream, A,
                           m + G.moff[BatchGroup::M_INTERP], it);
        if (G.any_ivc) {
          hipLaunchKernelGGL(k_ib_ivc_rhs_b, grid(BatchGroup::M_RHS), dim3(256), 0, stream, A,
                             m + G.moff[BatchGroup::M_RHS], it);
          hipLaunchKernelGGL(k_ib_ivc_matvec_b, grid(BatchGroup::M_MATVEC), dim3(256), 0, stream, A,
                             m + G.moff[BatchGroup::M_MATVEC], it);
        }
        if (G.mcnt[BatchGroup::M_SPREAD] > 0)
          hipLaunchKernelGGL(k_ib_spread_t_b, grid(BatchGroup::M_SPREAD), dim3(256), 0, stream, A,
                             m + G.moff[BatchGroup::M_SPREAD], it);
        // (a no-op for cases of one iteration, whose sums come with the record)
        if (G.max_loop > 1)
          hipLaunchKernelGGL(k_ib_sums_b, grid(BatchGroup::M_IBCASE), dim3(kSumsBlock), 0, stream, A,
                             m + G.moff[BatchGroup::M_IBCASE], it);
        if (G.any_iter)
          hipLaunchKernelGGL(k_ib_correct_b, grid(BatchGroup::M_BAND), dim3(256), 0, stream, A,
                             m + G.moff[BatchGroup::M_BAND], it);
      }
      if (G.mcnt[BatchGroup::M_FIX] > 0) G.fix(grid(BatchGroup::M_FIX), stream, A, m + G.moff[BatchGroup::M_FIX]);
      if (G.mcnt[BatchGroup::M_FIXW] > 0)
        G.fixw(grid(BatchGroup::M_FIXW), stream, A, m + G.moff[BatchGroup::M_FIXW]);
      // the records go with the slab launch, or without slabs to fill with
      // the stream-collide launch (enqueue_group_lbm)
      if (G.lane == 0 && phase_pending && !lane0_ib) {
        HIPCHK(hipEventRecord(ev_phase, stream));
        lane0_ib = true;
      }
    }
  }

  // the slabs, stream-collide (the force records ride with the slab launch,
  // or with the stream-collide launch when there are no slabs to fill),
  // Sum(Rho) and the clock
  void enqueue_group_lbm(BatchGroup& G, int p, hipStream_t stream) {
    const BlockRef* m = G.dmap;
    auto grid = [&](int k) { return dim3((unsigned)G.mcnt[k]); };
    if (G.mcnt[BatchGroup::M_GHOST] > 0 && G.mcnt[BatchGroup::M_IBCASE] > 0)
      G.ghost_rec((unsigned)(G.mcnt[BatchGroup::M_GHOST] + G.mcnt[BatchGroup::M_IBCASE]), stream, G.dghost[p],
                  m + G.moff[BatchGroup::M_GHOST], (unsigned)G.mcnt[BatchGroup::M_GHOST], G.dib[p],
                  m + G.moff[BatchGroup::M_IBCASE]);
    else if (G.mcnt[BatchGroup::M_GHOST] > 0)
      G.ghost(grid(BatchGroup::M_GHOST), stream, G.dghost[p], m + G.moff[BatchGroup::M_GHOST]);
    if (G.mcnt[BatchGroup::M_GHOST] == 0 && G.mcnt[BatchGroup::M_IBCASE] > 0)
      G.step_rec((unsigned)(G.mcnt[BatchGroup::M_STEP] + G.mcnt[BatchGroup::M_IBCASE]), stream, G.dstep[p],
                 m + G.moff[BatchGroup::M_STEP], (unsigned)G.mcnt[BatchGroup::M_STEP], G.dib[p],
                 m + G.moff[BatchGroup::M_IBCASE]);
    else
      G.step(grid(BatchGroup::M_STEP), stream, G.dstep[p], m + G.moff[BatchGroup::M_STEP]);
    hipLaunchKernelGGL(k_case_tail_b, grid(BatchGroup::M_TAIL), dim3(1024), 0, stream, G.dtail);
  }

  void enqueue_step(int p) {
    bool lane0_ib = false;
    if (lanes_together && nlanes > 1) {
      for (int lane = 0; lane < 2; ++lane)
        for (auto& G : groups)
          if (G.lane == lane) enqueue_group_ib(G, p, lane ? stream2 : stream, lane0_ib);
      HIPCHK(hipEventRecord(ev_ib[0], stream));
      HIPCHK(hipEventRecord(ev_ib[1], stream2));
      HIPCHK(hipStreamWaitEvent(stream, ev_ib[1], 0));
      HIPCHK(hipStreamWaitEvent(stream2, ev_ib[0], 0));
      phase_pending = false;
      for (int lane = 0; lane < 2; ++lane)
        for (auto& G : groups)
          if (G.lane == lane) enqueue_group_lbm(G, p, lane ? stream2 : stream);
      return;
    }
    for (int lane = 0; lane < 2; ++lane) {
      for (auto& G : groups) {
        if (G.lane != lane) continue;
        if (phase_pending && lane == 1) {
          // the first lane-1 work after a fork waits for lane 0's IB kernels
          // (or for lane 0's whole step if lane 0 has none)
          if (!lane0_ib) HIPCHK(hipEventRecord(ev_phase, stream));
          HIPCHK(hipStreamWaitEvent(stream2, ev_phase, 0));
          phase_pending = false;
        }
        enqueue_group(G, p, lane ? stream2 : stream, lane0_ib);
      }
    }
  }

  void capture() {
    drop_graph();
    HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    fork();
    for (int s = 0; s < graph_steps; ++s) enqueue_step(s & 1);
    join();
    hipGraph_t gr = nullptr;
    const hipError_t e = hipStreamEndCapture(stream, &gr);
    if (e != hipSuccess || !gr) {
      graph_failed = true;
      (void)hipGetLastError();
      return;
    }
    graph = gr;
    if (hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0) != hipSuccess) {
      graph_failed = true;
      (void)hipGetLastError();
      drop_graph();
      return;
    }
    ++captures;
  }

  int remaining_min() const {
    int r = 1 << 30;
    for (const auto& c : cases)
      if (c.done < c.target) r = std::min(r, c.target - c.done);
    return r == (1 << 30) ? 0 : r;
  }

  // Advances every unfinished case by k = min(max_steps, smallest remaining)
  // steps; returns k.
  int advance(int max_steps, bool use_graph) {
    if (max_steps <= 0) return 0;
    if (dirty) rebuild();
    const int k = std::min(max_steps, remaining_min());
    if (k <= 0) return 0;
    int left = k;
    bool forked = false;  // eager steps since the last fork run on both streams
    while (left > 0) {
      if (use_graph && nlanes == 1 && phase == 0 && left >= graph_steps && graph_steps >= 2 && !graph_failed) {
        if (forked) { join(); forked = false; }
        if (!gexec) capture();
        if (gexec) {
          HIPCHK(hipGraphLaunch(gexec, stream));
          ++graph_launches;
          left -= graph_steps;
          continue;
        }
      }
      if (!forked) { fork(); forked = true; }
      enqueue_step(phase);
      HIPCHK(hipGetLastError());
      phase ^= 1;
      ++eager_steps;
      --left;
    }
    if (forked) join();
    HIPCHK(hipStreamSynchronize(stream));
    // host bookkeeping and the streamed-back histories of the k steps
    for (auto& bc : cases) {
      if (bc.done >= bc.target) continue;
      Level& L = *bc.L;
      const int row0 = bc.done;
      std::vector<double> mh(k);
      HIPCHK(hipMemcpyAsync(mh.data(), L.mass_dev + row0, sizeof(double) * k, hipMemcpyDeviceToHost, stream));
      std::vector<double> fh;
      if (L.ib) {
        fh.resize((size_t)k * L.ncol);
        HIPCHK(hipMemcpyAsync(fh.data(), L.force_dev + (size_t)row0 * L.ncol, sizeof(double) * k * L.ncol,
                              hipMemcpyDeviceToHost, stream));
      }
      HIPCHK(hipStreamSynchronize(stream));
      L.mass_host.insert(L.mass_host.end(), mh.begin(), mh.end());
      L.force_host.insert(L.force_host.end(), fh.begin(), fh.end());
      for (int s = 0; s < k; ++s) {
        L.t_old = L.t_new;
        L.t_new += L.dt;
      }
      L.istep += k;
      L.cur ^= (k & 1);
      if (k > 0) L.slabs_fresh = L.gfuse;
      L.state_valid = false;
      bc.done += k;
      lups += (double)L.ncells * k;
    }
    // a case that finished leaves the lock-step set: tables change
    for (auto& bc : cases)
      if (bc.done >= bc.target && bc.done - k < bc.target) dirty = true;
    return k;
  }
};

}  // namespace ilbm

// ---------------------------------------------------------------------------
// C ABI of the batch (include/ilbm.h)
// ---------------------------------------------------------------------------
struct ilbm_batch {
  ilbm::Batch B;
};

extern "C" {

int ilbm_batch_create(int64_t hbm_budget_bytes, int max_cases, int graph_steps, ilbm_batch** out) {
  if (!out || hbm_budget_bytes <= 0 || max_cases <= 0 || graph_steps < 0) return ILBM_ERR_INVALID;
  *out = nullptr;
  ilbm_batch* h = new (std::nothrow) ilbm_batch();
  if (!h) return ILBM_ERR_ALLOC;
  const int st = guarded([&] {
    h->B.budget = hbm_budget_bytes;
    h->B.max_cases = max_cases;
    h->B.graph_steps = (graph_steps / 2) * 2;
    HIPCHK(hipStreamCreateWithFlags(&h->B.stream, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&h->B.stream2, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&h->B.ev_fork, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&h->B.ev_join, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&h->B.ev_phase, hipEventDisableTiming));
    for (hipEvent_t& e : h->B.ev_ib) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    if (const char* e = std::getenv("ILBM_BATCH_PHASE")) h->B.lanes_together = std::string(e) == "together";
  });
  if (st != ILBM_OK) {
    delete h;
    return st;
  }
  *out = h;
  return ILBM_OK;
}

int ilbm_batch_destroy(ilbm_batch* h) {
  if (!h) return ILBM_ERR_INVALID;
  delete h;
  return ILBM_OK;
}

int ilbm_batch_footprint(const ilbm_case_desc* desc, int nsteps, int64_t* bytes) {
  if (!desc || !bytes || nsteps < 0) return ILBM_ERR_INVALID;
  *bytes = ilbm::case_footprint(*desc, nsteps);
  return ILBM_OK;
}

int ilbm_batch_add(ilbm_batch* h, const ilbm_case_desc* desc, const double* init_macro, int nsteps, int* case_id) {
  if (!h || !desc || !init_macro || !case_id) return ILBM_ERR_INVALID;
  *case_id = -1;
  return guarded([&] { *case_id = h->B.add(*desc, init_macro, nsteps); });
}

int ilbm_batch_advance(ilbm_batch* h, int max_steps, int use_graph, int* steps_done) {
  if (!h || !steps_done) return ILBM_ERR_INVALID;
  *steps_done = 0;
  return guarded([&] { *steps_done = h->B.advance(max_steps, use_graph != 0); });
}

int ilbm_batch_finished(ilbm_batch* h, int* ids, int max_ids, int* n) {
  if (!h || !n) return ILBM_ERR_INVALID;
  int k = 0;
  for (const auto& c : h->B.cases)
    if (c.done >= c.target) {
      if (ids && k < max_ids) ids[k] = c.id;
      ++k;
    }
  *n = k;
  return ILBM_OK;
}

int ilbm_batch_progress(ilbm_batch* h, int case_id, int* done, int* target) {
  if (!h || !done || !target) return ILBM_ERR_INVALID;
  const ilbm::BatchCase* c = h->B.find(case_id);
  if (!c) return ILBM_ERR_INVALID;
  *done = c->done;
  *target = c->target;
  return ILBM_OK;
}

int ilbm_batch_force_history(ilbm_batch* h, int case_id, double* out, int max_rows, int* n_rows, int* n_cols) {
  if (!h) return ILBM_ERR_INVALID;
  ilbm::BatchCase* c = h->B.find(case_id);
  if (!c) return ILBM_ERR_INVALID;
  return force_history_of(*c->L, out, max_rows, n_rows, n_cols);
}

int ilbm_batch_mass_history(ilbm_batch* h, int case_id, double* out, int max_rows, int* n_out) {
  if (!h) return ILBM_ERR_INVALID;
  ilbm::BatchCase* c = h->B.find(case_id);
  if (!c) return ILBM_ERR_INVALID;
  return mass_history_of(*c->L, out, max_rows, n_out);
}

int ilbm_batch_checkpoint_size(ilbm_batch* h, int case_id, int64_t* bytes) {
  if (!h || !bytes) return ILBM_ERR_INVALID;
  ilbm::BatchCase* c = h->B.find(case_id);
  if (!c) return ILBM_ERR_INVALID;
  *bytes = c->L->checkpoint_size();
  return ILBM_OK;
}

int ilbm_batch_checkpoint(ilbm_batch* h, int case_id, void* host_buf, int64_t bytes) {
  if (!h || !host_buf) return ILBM_ERR_INVALID;
  ilbm::BatchCase* c = h->B.find(case_id);
  if (!c) return ILBM_ERR_INVALID;
  return guarded([&] {
    HIPCHK(hipStreamSynchronize(h->B.stream));
    c->L->checkpoint(static_cast<char*>(host_buf), bytes);
  });
}

int ilbm_batch_add_from_checkpoint(ilbm_batch* h, const ilbm_case_desc* desc, const void* host_buf, int64_t bytes,
                                   int nsteps, int* case_id) {
  if (!h || !desc || !host_buf || !case_id) return ILBM_ERR_INVALID;
  *case_id = -1;
  return guarded([&] {
    *case_id = h->B.add_from_checkpoint(*desc, static_cast<const char*>(host_buf), bytes, nsteps);
  });
}

int ilbm_batch_remove(ilbm_batch* h, int case_id) {
  if (!h) return ILBM_ERR_INVALID;
  return guarded([&] { h->B.remove(case_id); });
}

int ilbm_batch_case_tile(ilbm_batch* h, int case_id, int* tile_bx) {
  if (!h || !tile_bx) return ILBM_ERR_INVALID;
  const ilbm::BatchCase* c = h->B.find(case_id);
  if (!c) return ILBM_ERR_INVALID;
  *tile_bx = c->tile_bx;
  return ILBM_OK;
}

int ilbm_device_memory(int64_t* free_bytes, int64_t* total_bytes) {
  if (!free_bytes || !total_bytes) return ILBM_ERR_INVALID;
  return guarded([&] {
    size_t f = 0, t = 0;
    HIPCHK(hipMemGetInfo(&f, &t));
    *free_bytes = (int64_t)f;
    *total_bytes = (int64_t)t;
  });
}

int ilbm_batch_info(ilbm_batch* h, double out[8]) {
  if (!h || !out) return ILBM_ERR_INVALID;
  const ilbm::Batch& B = h->B;
  int active = 0;
  for (const auto& c : B.cases) active += (c.done < c.target) ? 1 : 0;
  out[0] = (double)B.cases.size();
  out[1] = (double)active;
  out[2] = (double)B.used;
  out[3] = (double)B.budget;
  out[4] = (double)B.graph_launches;
  out[5] = (double)B.captures;
  out[6] = (double)B.eager_steps;
  out[7] = B.lups;
  return ILBM_OK;
}

}  // extern "C"
