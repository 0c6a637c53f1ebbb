// AMR hierarchy on one device (compiled as part of ilbm_level.hip): level 0
// is the case's domain, each finer level one box at refinement ratio 2,
// advanced with subcycling as timeStepWithSubcycling does
// (Source/Core/AmrCoreLBM.cpp:1688-1754), with the regrid of :1692-1720
// when it is switched on (ErrorEst tags on the device, one box per level
// chosen on the host, RemakeLevel on the device).  Every level is a resident Level; a refined level has ghost
// slabs on all faces, filled from its coarser level by k_cf_fill, and the
// immersed body lives on the finest level only (AdvanceLBMAtLevel.cpp:87-99).
//
// Per coarse step:
//   substep(0):  swap window macro_old/new, stream-collide level 0 (its band
//                exports the window macros), derived fields every
//                regrid_int steps, then twice substep(1) and
//                AverageDownTo(0) + derived fields
//   substep(l):  ghost slabs from level l-1 at the substep's time, IB
//                forcing (finest), stream-collide (exports f and macros on
//                the second substep for the average-down), recursion
//   Sum(Rho) of level 0 and the finest level's force record (Evolve,
//                AmrCoreLBM.cpp:366-420), all levels' time set to the
//                coarse time.
//
// Fine levels are created as MakeNewLevelFromCoarse makes them (macros
// interpolated from the coarse level, f = feq); the initial state is
// uniform, where this equals MakeNewLevelFromScratch with a uniform initdata.
#pragma once

#include "ilbm_amr_kernels.h"

namespace ilbm {

using CfFn = void (*)(dim3, hipStream_t, const CfArgs&);
using CfMoveFn = void (*)(unsigned, unsigned, hipStream_t, const CfArgs&, const MoveArgs&, const IbArgs&);
using AvgFn = void (*)(dim3, hipStream_t, const AvgArgs&);

template <class LAT, class ST>
static void launch_cf(dim3 g, hipStream_t s, const CfArgs& a) {
  hipLaunchKernelGGL((k_cf_fill<LAT, ST>), g, dim3(kBlock), 0, s, a);
}
// A refined level's coarse-fine ghost fill and the marker kinematics and
// weights of its moving body in one launch: blocks [0, ncf) fill ghost
// cells, the rest run k_ib_move_weights's blocks.  Neither reads what the
// other writes, and the forcing and stream-collide that follow need both.
template <class LAT, class ST>
__global__ void __launch_bounds__(kBlock) k_cf_fill_move(CfArgs a, unsigned ncf, MoveArgs mv, IbArgs ia) {
  static_assert(kBlock == 256, "ghost and marker blocks share the launch's block size");
  if (blockIdx.x < ncf) cf_fill_body<LAT, ST>(a, (int64_t)blockIdx.x * kBlock + threadIdx.x);
  else move_weights_body(mv, ia, (int)(blockIdx.x - ncf));
}
template <class LAT, class ST>
static void launch_cf_move(unsigned ncf, unsigned nmv, hipStream_t s, const CfArgs& a, const MoveArgs& mv,
                           const IbArgs& ia) {
  hipLaunchKernelGGL((k_cf_fill_move<LAT, ST>), dim3(ncf + nmv), dim3(kBlock), 0, s, a, ncf, mv, ia);
}

template <class LAT, class ST>
static void launch_avg(dim3 g, hipStream_t s, const AvgArgs& a) {
  hipLaunchKernelGGL((k_average_down<LAT, ST>), g, dim3(256), 0, s, a);
}

static void pick_amr(int lat, int st, CfFn& cf, AvgFn& av, CfMoveFn& cm) {
  auto pk = [&](auto latt) {
    using L = decltype(latt);
    if (st == 0) { cf = launch_cf<L, StF64>; av = launch_avg<L, StF64>; cm = launch_cf_move<L, StF64>; }
    else if (st == 1) { cf = launch_cf<L, StF32>; av = launch_avg<L, StF32>; cm = launch_cf_move<L, StF32>; }
    else { cf = launch_cf<L, StF16S>; av = launch_avg<L, StF16S>; cm = launch_cf_move<L, StF16S>; }
  };
  if (lat == 0) pk(D2Q9{});
  else if (lat == 1) pk(D3Q19{});
  else pk(D3Q27{});
}

constexpr int kMaxLev = 3;

struct Amr {
  hipStream_t stream = nullptr;
  int nlev = 1, regrid_int = 1, dim = 2;
  std::vector<std::unique_ptr<Level>> L;
  int blo[kMaxLev][3] = {}, bhi[kMaxLev][3] = {};  // box of level l, level-l global indices
  int wlo[kMaxLev][3] = {}, wn[kMaxLev][3] = {};   // window of level l (l < nlev-1), global
  int clo[kMaxLev][3] = {}, cn[kMaxLev][3] = {};   // covered cells of level l, global
  double* dpartial[kMaxLev] = {nullptr, nullptr, nullptr};
  int64_t dblocks[kMaxLev] = {0, 0, 0};
  bool exported[kMaxLev] = {false, false, false};
  CfFn cffn = nullptr;
  CfMoveFn cfmovefn = nullptr;
  AvgFn avgfn = nullptr;
  int istep = 0;
  double step_end = 0.0;
  double* mass_dev = nullptr;
  int mass_cap = 0;
  std::vector<double> mass_host, force_host;
  int ncol = 0;
  ilbm_case_desc d0{};            // the case as given (refinement-region parameters)
  // regrid: ErrorEst + one box per level + RemakeLevel
  bool regrid_on = false;
  ilbm_regrid_params rg{};
  int last_regrid[kMaxLev] = {0, 0, 0};
  int nregrid = 0;
  double* Dfull[kMaxLev][2] = {};   // windowed levels: [2][cells] vorticity, pressure after step s in [s & 1]
  double* patch[kMaxLev] = {};      // one-ghost macro patch [4][patch cells]
  int64_t patch_cap[kMaxLev] = {};
  double* scratch = nullptr;        // [2][cells] derived fields for tagging
  int64_t scratch_cap = 0;
  int* tag_d = nullptr;             // count, bbox lo[3], hi[3]
  unsigned long long* vmax_d = nullptr;

  ~Amr() {
    if (stream) (void)hipStreamSynchronize(stream);
    L.clear();
    for (int l = 0; l < kMaxLev; ++l) {
      dev_free(dpartial[l]);
      dev_free(Dfull[l][0]);
      dev_free(Dfull[l][1]);
      dev_free(patch[l]);
    }
    dev_free(scratch);
    dev_free(tag_d);
    dev_free(vmax_d);
    dev_free(mass_dev);
    if (stream) (void)hipStreamDestroy(stream);
  }

  // covered cells and windows from the boxes: a fine box covers whole
  // coarse cells and its window (the coarse footprint grown by 3) lies in
  // the coarse box
  void nest_geometry() {
    for (int l = 1; l < nlev; ++l)
      for (int d = 0; d < dim; ++d) {
        if (blo[l][d] % 2 != 0 || bhi[l][d] % 2 != 1 || bhi[l][d] < blo[l][d])
          throw StatusError{ILBM_ERR_INVALID, "fine box must cover whole coarse cells"};
        const int flo = blo[l][d] / 2, fhi = bhi[l][d] / 2;
        clo[l - 1][d] = flo;
        cn[l - 1][d] = fhi - flo + 1;
        wlo[l - 1][d] = flo - 3;
        wn[l - 1][d] = fhi - flo + 7;
        if (flo - 3 < blo[l - 1][d] || fhi + 3 > bhi[l - 1][d])
          throw StatusError{ILBM_ERR_UNSUPPORTED, "fine box too close to the coarse box edge (needs 3 coarse cells)"};
      }
    for (int l = 0; l < nlev - 1; ++l)
      for (int d = dim; d < 3; ++d) { clo[l][d] = 0; cn[l][d] = 1; wlo[l][d] = 0; wn[l][d] = 1; }
  }

  // window storage of level l for the current box of level l+1 (contents
  // left to the caller), and the average-down partials of its footprint
  void alloc_window(int l) {
    Level& V = *L[l];
    V.window = true;
    int64_t cells = 1;
    for (int d = 0; d < 3; ++d) {
      V.band_lo[d] = wlo[l][d] - V.ioff[d];
      V.band_n[d] = wn[l][d];
      cells *= wn[l][d];
    }
    V.band_cells = cells;
    for (int b = 0; b < 2; ++b) {
      dev_free(V.Mw[b]);
      dev_free(V.Dw[b]);
      V.Mw[b] = dev_alloc<double>(4 * cells);
      V.Dw[b] = dev_alloc<double>(2 * cells);
      HIPCHK(hipMemsetAsync(V.Dw[b], 0, sizeof(double) * 2 * cells, stream));
    }
    const int64_t ncov = (int64_t)cn[l][0] * cn[l][1] * cn[l][2];
    dev_free(dpartial[l]);
    dblocks[l] = (ncov + 255) / 256;
    dpartial[l] = dev_alloc<double>(dblocks[l]);
  }

  // a Level for the current box of level l, with its window (l < nlev-1)
  // and its marker set (finest), not yet initialised
  std::unique_ptr<Level> new_level(int l) {
    std::unique_ptr<Level> V(new Level());
    ilbm_case_desc c = d0;
    const int fac = 1 << l;
    for (int d = 0; d < dim; ++d) {
      c.dx[d] = d0.dx[d] / (double)fac;
      c.n[d] = bhi[l][d] - blo[l][d] + 1;
      V->ioff[d] = blo[l][d];
      V->dom_n[d] = d0.n[d] * fac;
    }
    c.use_ib = (l == nlev - 1) ? d0.use_ib : 0;
    V->c = c;
    V->cf = (l > 0);
    V->stream = stream;
    V->own_stream = false;
    V->validate();
    V->setup_layout();
    if ((int)L.size() <= l) L.resize(l + 1);
    std::swap(L[l], V);   // alloc_window addresses L[l]
    if (l < nlev - 1) alloc_window(l);
    L[l]->setup_ib();
    std::swap(L[l], V);
    return V;
  }

  void create(const ilbm_case_desc& dd, int nl, const int* box_lo, const int* box_hi, int rint,
              const double st[4]) {
    if (nl < 1 || nl > kMaxLev) throw StatusError{ILBM_ERR_UNSUPPORTED, "1 to 3 levels"};
    if (rint < 1) throw StatusError{ILBM_ERR_INVALID, "regrid_int must be >= 1"};
    if (dd.nmac != 4 && dd.nmac != 6) throw StatusError{ILBM_ERR_UNSUPPORTED, "nmac must be 4 or 6"};
    d0 = dd;
    nlev = nl;
    regrid_int = rint;
    dim = d0.dim;
    if (dim != 2 && dim != 3) throw StatusError{ILBM_ERR_INVALID, "dim"};
    for (int d = 0; d < 3; ++d) {
      blo[0][d] = 0;
      bhi[0][d] = (d < dim) ? d0.n[d] - 1 : 0;
    }
    for (int l = 1; l < nlev; ++l)
      for (int d = 0; d < 3; ++d) {
        blo[l][d] = (d < dim) ? box_lo[3 * (l - 1) + d] : 0;
        bhi[l][d] = (d < dim) ? box_hi[3 * (l - 1) + d] : 0;
      }
    nest_geometry();
    L.clear();
    L.resize(nlev);
    for (int l = 0; l < nlev; ++l) {
      std::unique_ptr<Level> V = new_level(l);
      std::vector<double> m0(4 * (size_t)V->ncells);
      for (int n = 0; n < 4; ++n) std::fill(m0.begin() + n * V->ncells, m0.begin() + (n + 1) * V->ncells, st[n]);
      V->init(m0.data());
      if (l > 0) V->ensure_exports();
      L[l] = std::move(V);
    }
    // macro_old = macro_new on the windows, derived fields of both
    for (int l = 0; l < nlev - 1; ++l) {
      Level& V = *L[l];
      HIPCHK(hipMemcpyAsync(V.Mw[V.wsel ^ 1], V.Mw[V.wsel], sizeof(double) * 4 * V.band_cells,
                            hipMemcpyDeviceToDevice, stream));
      derived(l, 0);
      derived(l, 1);
    }
    const Level& F = *L[nlev - 1];
    ncol = F.ib ? F.ncol : 0;
    pick_amr(L[0]->lat, L[0]->st, cffn, avgfn, cfmovefn);
    HIPCHK(hipStreamSynchronize(stream));
  }

  // before the first step: the derived fields of the initial state fill
  // both slots (InitData's UpdateDerivedMacroFields, macro_old = macro_new)
  void set_regrid(const ilbm_regrid_params& p) {
    if (p.blocking_factor < 2 || p.blocking_factor % 2 != 0)
      throw StatusError{ILBM_ERR_INVALID, "blocking_factor must be even and >= 2"};
    if (p.n_error_buf < 0) throw StatusError{ILBM_ERR_INVALID, "n_error_buf < 0"};
    if (istep != 0) throw StatusError{ILBM_ERR_STATE, "regridding must be switched on before the first advance"};
    for (int l = 0; l < nlev - 1; ++l) {
      Level& V = *L[l];
      if (!V.mout) throw StatusError{ILBM_ERR_UNSUPPORTED, "level too large for an exported initial state"};
      ensure_dfull(l);
      patch_derived(l, V.t_new, Dfull[l][0]);
      HIPCHK(hipMemcpyAsync(Dfull[l][1], Dfull[l][0], sizeof(double) * 2 * V.ncells, hipMemcpyDeviceToDevice,
                            stream));
    }
    HIPCHK(hipStreamSynchronize(stream));
    rg = p;
    regrid_on = true;
  }

  void derived(int l, int sel) {
    Level& V = *L[l];
    if (!V.window) return;
    DerArgs a{};
    a.M = V.Mw[sel];
    a.D = V.Dw[sel];
    for (int d = 0; d < 3; ++d) a.n[d] = (int)V.band_n[d];
    a.cells = V.band_cells;
    a.dim = dim;
    a.dx = V.c.dx[0]; a.dy = V.c.dx[1]; a.dz = V.c.dx[2];
    const int64_t nint = (int64_t)(a.n[0] - 2) * (a.n[1] - 2) * ((dim == 3) ? a.n[2] - 2 : 1);
    if (nint > 0) hipLaunchKernelGGL(k_derived, dim3((unsigned)((nint + 255) / 256)), dim3(256), 0, stream, a);
  }

  // GetDataMacro of level lc at `time` (AmrCoreLBM.cpp:2239-2269)
  void window_at(int lc, double time, Window& W) const {
    const Level& C = *L[lc];
    W.Mo = C.Mw[C.wsel ^ 1]; W.Mn = C.Mw[C.wsel];
    W.Do = C.Dw[C.wsel ^ 1]; W.Dn = C.Dw[C.wsel];
    for (int d = 0; d < 3; ++d) { W.lo[d] = wlo[lc][d]; W.n[d] = wn[lc][d]; }
    W.cells = C.band_cells;
    W.a = W.b = 0.0;
    const double SENT = 1.0e50;
    if (std::fabs(C.t_old) > SENT || std::fabs(C.t_new) > SENT) { W.mode = 1; return; }
    const double dtloc = std::fabs(C.t_new - C.t_old);
    const double teps = std::max(1.0e-12, dtloc * 1.0e-3);
    if (time >= C.t_new - teps && time <= C.t_new + teps) { W.mode = 1; return; }
    if (time >= C.t_old - teps && time <= C.t_old + teps) { W.mode = 0; return; }
    W.mode = 2;
    W.a = (C.t_new - time) / (C.t_new - C.t_old);
    W.b = (time - C.t_old) / (C.t_new - C.t_old);
  }

  // with_move: the level's moving-body kinematics at `time` go in the same
  // launch (k_cf_fill_move) and its next forcing skips them
  void cf_fill(int l, double time, bool with_move = false) {
    Level& V = *L[l];
    CfArgs a{};
    a.g = V.g;
    a.slabs = V.slabs;
    a.gbuf = V.gbuf[V.cur];
    for (int d = 0; d < 3; ++d) a.ioff[d] = (d < dim) ? V.ioff[d] : 0;
    a.ncomp = V.c.nmac;
    window_at(l - 1, time, a.W);
    const int64_t npad = V.slabs.face_off[V.slabs.nface];
    const unsigned ncf = (unsigned)((npad + kBlock - 1) / kBlock);
    if (with_move) {
      cfmovefn(ncf, (unsigned)((V.nm + kBlock - 1) / kBlock), stream, a, V.move_args(time), V.ib_args());
      V.moved_ext = true;
      return;
    }
    cffn(dim3(ncf), stream, a);
  }

  void average_down(int l) {
    Level& C = *L[l];
    Level& F = *L[l + 1];
    AvgArgs a{};
    a.gc = C.g;
    a.gbuf = C.gbuf[C.cur];
    for (int d = 0; d < 3; ++d) {
      a.clo[d] = clo[l][d] - C.ioff[d];
      a.cn[d] = cn[l][d];
      a.fn[d] = F.c.n[d];
      a.fofs[d] = (d < dim) ? 2 * clo[l][d] - F.ioff[d] : 0;
      a.wofs[d] = clo[l][d] - wlo[l][d];
      a.wn[d] = wn[l][d];
    }
    a.ncov = (int64_t)cn[l][0] * cn[l][1] * cn[l][2];
    a.fout = F.fout;
    a.mout = F.mout;
    a.Mw = C.Mw[C.wsel];
    a.wcells = C.band_cells;
    a.cfout = exported[l] ? C.fout : nullptr;
    a.cmout = exported[l] ? C.mout : nullptr;
    a.tau = C.tau;
    a.inv_tau = 1.0 / C.tau;
    const StepArgs sa = C.step_args(false);
    for (int d = 0; d < 3; ++d) a.Fcol[d] = sa.Fcol[d];
    a.dpartial = dpartial[l];
    avgfn(dim3((unsigned)dblocks[l]), stream, a);
    // the covered cells keep 3 coarse cells inside the coarse box
    // (nest_geometry), so no boundary cell of level 0 is averaged down and
    // the slabs its stream-collide kernel wrote stay current
  }

  // ---------------- regrid ----------------
  // FillPatchMacro(l, time) of the 4 macro components with one ghost layer
  // from the level's exported macros (physical / periodic ghosts on level
  // 0, coarse interpolation on refined levels), then visPara on every
  // valid cell into D[2][cells]
  void patch_derived(int l, double time, double* D) {
    Level& V = *L[l];
    if (!V.mout) throw StatusError{ILBM_ERR_STATE, "level macros not exported"};
    PatchFab P{};
    for (int d = 0; d < 3; ++d) {
      const int g = (d < dim) ? 1 : 0;
      P.lo[d] = -g;
      P.hi[d] = V.c.n[d] - 1 + g;
    }
    P.js = P.hi[0] - P.lo[0] + 1;
    P.ks = P.js * (P.hi[1] - P.lo[1] + 1);
    P.ns = P.ks * (P.hi[2] - P.lo[2] + 1);
    if (4 * P.ns > patch_cap[l]) {
      dev_free(patch[l]);
      patch[l] = dev_alloc<double>(4 * P.ns);
      patch_cap[l] = 4 * P.ns;
    }
    P.p = patch[l];
    Fab F{};
    F.p = P.p;
    for (int d = 0; d < 3; ++d) { F.lo[d] = P.lo[d]; F.hi[d] = P.hi[d]; }
    F.js = P.js; F.ks = P.ks; F.ns = P.ns; F.cells = P.ns; F.nc = 4;
    const int64_t cells = V.ncells;
    hipLaunchKernelGGL(k_fab_from_cells, grid_for(cells), dim3(256), 0, stream, F, (const double*)V.mout, V.c.n[0],
                       V.c.n[1], cells);
    if (l == 0) {
      const BoxGeom G = box_geom(V.c, nullptr, nullptr);
      hipLaunchKernelGGL(k_box_fill_periodic, grid_for(P.ns), dim3(256), 0, stream, F, G);
      with_lattice(V.lat, [&](auto lat) {
        using LAT = decltype(lat);
        for (int pass = 1; pass <= dim; ++pass)
          hipLaunchKernelGGL(k_box_fill_physical<LAT>, grid_for(P.ns), dim3(256), 0, stream, F, G, pass, 0);
      });
    } else {
      PatchCfArgs a{};
      a.P = P;
      for (int d = 0; d < 3; ++d) { a.n[d] = V.c.n[d]; a.ioff[d] = (d < dim) ? V.ioff[d] : 0; }
      a.dim = dim;
      window_at(l - 1, time, a.W);
      hipLaunchKernelGGL(k_patch_cf, grid_for(P.ns), dim3(256), 0, stream, a);
    }
    hipLaunchKernelGGL(k_patch_derived, grid_for(cells), dim3(256), 0, stream, P, D, V.c.n[0], V.c.n[1], cells, dim,
                       V.c.dx[0], V.c.dx[1], V.c.dx[2]);
    HIPCHK(hipGetLastError());
  }

  void ensure_dfull(int l) {
    if (!Dfull[l][0]) {
      Dfull[l][0] = dev_alloc<double>(2 * L[l]->ncells);
      Dfull[l][1] = dev_alloc<double>(2 * L[l]->ncells);
    }
  }

  // immersed_boundary_refinement_region (AmrCoreLBM.cpp:60-104); the user
  // (pitching) section is Examples/Pitching_airfoil/IBMUserDefinedGeometry.H:
  // 25-57 with the vertical reach of the point farthest from the pivot
  bool refine_region(double dx, double lo[3], double hi[3]) const {
    const ilbm_case_desc& c = d0;
    const double pad = std::max(6.0 * dx, 2.0);
    if (c.geometry_type == 1) {
      lo[0] = c.box_lo[0] - pad; hi[0] = c.box_hi[0] + pad;
      lo[1] = (c.box_lid_only > 0) ? (c.box_hi[1] - pad) : (c.box_lo[1] - pad);
      hi[1] = c.box_hi[1] + pad;
      lo[2] = c.box_lo[2] - pad; hi[2] = c.box_hi[2] + pad;
      return true;
    }
    if (c.geometry_type == 2) {
      if (c.dim != 2) return false;
      const double chord = std::max(c.airfoil_chord, 1.0e-12);
      const double ht = 0.5 * c.airfoil_thickness * chord;
      const double arm = std::max(c.pivot_x, chord - c.pivot_x);
      const double yext = arm * std::sin(std::fabs(c.pitch_amplitude)) + ht;
      lo[0] = c.x0 - pad; hi[0] = c.x0 + chord + pad;
      lo[1] = c.y0 - yext - pad; hi[1] = c.y0 + yext + pad;
      lo[2] = c.z0 - pad; hi[2] = c.z0 + pad;
      return true;
    }
    if (c.R > 0.0) {
      const bool ext = rg.refine_upstream > 0.0 && rg.refine_downstream > 0.0 && rg.refine_cross > 0.0;
      const double up = ext ? rg.refine_upstream * c.R : c.R, dn = ext ? rg.refine_downstream * c.R : c.R;
      const double cr = ext ? rg.refine_cross * c.R : c.R;
      lo[0] = c.x0 - up - pad; hi[0] = c.x0 + dn + pad;
      lo[1] = c.y0 - cr - pad; hi[1] = c.y0 + cr + pad;
      lo[2] = c.z0 - cr - pad; hi[2] = c.z0 + cr + pad;
      return true;
    }
    return false;
  }

  // ErrorEst on level l at `time`: tag count and the bounding box of the
  // tags grown by n_error_buf (level-l global indices)
  int64_t tag_bbox(int l, double time, int tlo[3], int thi[3]) {
    Level& V = *L[l];
    const int64_t cells = V.ncells;
    if (2 * cells > scratch_cap) {
      dev_free(scratch);
      scratch = dev_alloc<double>(2 * cells);
      scratch_cap = 2 * cells;
    }
    if (!tag_d) tag_d = dev_alloc<int>(8);
    if (!vmax_d) vmax_d = dev_alloc<unsigned long long>(1);
    patch_derived(l, time, scratch);
    const int init[8] = {0, 1 << 30, 1 << 30, 1 << 30, -(1 << 30), -(1 << 30), -(1 << 30), 0};
    HIPCHK(hipMemcpyAsync(tag_d, init, sizeof(init), hipMemcpyHostToDevice, stream));
    const unsigned long long zero = 0;
    HIPCHK(hipMemcpyAsync(vmax_d, &zero, sizeof(zero), hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(k_vort_max, grid_for(cells), dim3(256), 0, stream, (const double*)scratch, cells, vmax_d);
    TagArgs a{};
    a.vort = scratch;
    a.vmax_key = vmax_d;
    for (int d = 0; d < 3; ++d) {
      a.n[d] = V.c.n[d];
      a.ioff[d] = (d < dim) ? V.ioff[d] : 0;
      a.plo[d] = d0.plo[d];
      a.dx[d] = V.c.dx[d];
    }
    a.dim = dim;
    a.cells = cells;
    a.ratio = rg.threshold_ratio[std::min(l, 3)];
    a.probhi0 = d0.plo[0] + d0.dx[0] * (double)d0.n[0];
    a.nr_lo = rg.no_refine_xlo;
    a.nr_hi = rg.no_refine_xhi;
    a.region = d0.use_ib && refine_region(V.c.dx[0], a.rlo, a.rhi);
    a.lset = d0.use_ib && d0.geometry_type == 0;
    a.x0 = d0.x0; a.y0 = d0.y0; a.z0 = d0.z0; a.R = d0.R;
    a.band = (5.0 + 0.5) * V.c.dx[0];
    a.buf = rg.n_error_buf;
    a.out = tag_d;
    hipLaunchKernelGGL(k_tag_bbox, grid_for(cells), dim3(256), 0, stream, a);
    HIPCHK(hipGetLastError());
    int out[8];
    HIPCHK(hipMemcpyAsync(out, tag_d, sizeof(out), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    for (int d = 0; d < 3; ++d) {
      tlo[d] = (d < dim) ? out[1 + d] : 0;
      thi[d] = (d < dim) ? out[4 + d] : 0;
    }
    return out[0];
  }

  static int fdiv(int a, int b) { return (a >= 0) ? a / b : -((-a + b - 1) / b); }

  // one box for level l+1 (level-(l+1) indices) from a tag bounding box in
  // level-l indices: clipped to the parent box shrunk by 3 cells, aligned
  // outward to the blocking unit, inward where that leaves the clip region
  bool box_from_tags(const int tlo[3], const int thi[3], const int plo[3], const int phi[3], int bu, int olo[3],
                     int ohi[3]) const {
    for (int d = 0; d < 3; ++d) { olo[d] = 0; ohi[d] = 0; }
    for (int d = 0; d < dim; ++d) {
      const int nlo = plo[d] + 3, nhi = phi[d] - 3;
      const int lo = std::max(tlo[d], nlo), hi = std::min(thi[d], nhi);
      if (lo > hi) return false;
      int alo = fdiv(lo, bu) * bu, ahi = (fdiv(hi, bu) + 1) * bu - 1;
      if (alo < nlo) alo = -fdiv(-nlo, bu) * bu;
      if (ahi > nhi) ahi = fdiv(nhi + 1, bu) * bu - 1;
      if (alo > ahi) return false;
      olo[d] = 2 * alo;
      ohi[d] = 2 * ahi + 1;
    }
    return true;
  }

  // crop of the window box of level l out of a [nc][level cells] array
  void crop_window(int l, const double* src, double* dst, int nc) {
    Level& V = *L[l];
    const int64_t wc = V.band_cells;
    hipLaunchKernelGGL(k_crop, grid_for(wc), dim3(256), 0, stream, src, V.c.n[0], V.c.n[1], (int64_t)V.ncells, dst,
                       (int)V.band_n[0], (int)V.band_n[1], wc, (int)V.band_lo[0], (int)V.band_lo[1],
                       (int)V.band_lo[2], nc);
  }

  // window of level l (not re-boxed) moved to the new box of level l+1:
  // macro_new from the level's exported macros, the derived fields of the
  // last two steps (macro_new's, and macro_old's stale ones)
  void move_window(int l) {
    Level& V = *L[l];
    alloc_window(l);
    crop_window(l, V.mout, V.Mw[V.wsel], 4);
    crop_window(l, V.mout, V.Mw[V.wsel ^ 1], 4);
    crop_window(l, Dfull[l][V.istep & 1], V.Dw[V.wsel], 2);
    crop_window(l, Dfull[l][(V.istep + 1) & 1], V.Dw[V.wsel ^ 1], 2);
    HIPCHK(hipGetLastError());
  }

  // RemakeLevel (AmrCoreLBM.cpp:835-914) of level l on its new box
  void remake(int l, double time) {
    Level& O = *L[l];
    const int nmac = d0.nmac;
    std::unique_ptr<Level> N = new_level(l);
    double* M = dev_alloc<double>((size_t)nmac * N->ncells);
    RemakeArgs ra{};
    ra.M = M;
    for (int d = 0; d < 3; ++d) {
      ra.n[d] = N->c.n[d];
      ra.lo[d] = (d < dim) ? blo[l][d] : 0;
      ra.on[d] = O.c.n[d];
      ra.olo[d] = (d < dim) ? O.ioff[d] : 0;
    }
    if (!O.mout) throw StatusError{ILBM_ERR_STATE, "level macros not exported"};
    ra.om = O.mout;
    ra.od = O.window ? Dfull[l][O.istep & 1] : nullptr;
    ra.dim = dim;
    ra.nmac = nmac;
    window_at(l - 1, time, ra.W);
    hipLaunchKernelGGL(k_remake_macro, grid_for(N->ncells), dim3(256), 0, stream, ra);
    HIPCHK(hipGetLastError());
    N->init_dev(M);   // synchronizes: the old level's derived fields are free to go
    for (int b = 0; b < 2; ++b) {
      dev_free(Dfull[l][b]);
      Dfull[l][b] = nullptr;
    }
    N->ensure_exports();
    HIPCHK(hipMemcpyAsync(N->mout, M, sizeof(double) * 4 * N->ncells, hipMemcpyDeviceToDevice, stream));
    N->t_new = O.t_new;
    N->t_old = O.t_old;
    N->istep = O.istep;
    if (N->ib && O.ib) {   // the force record carries on (ComputeIBForce state is the hierarchy's)
      N->grow_histories(1, O.force_cap);
      if (O.force_cap && O.force_dev)
        HIPCHK(hipMemcpyAsync(N->force_dev, O.force_dev, sizeof(double) * (size_t)O.force_cap * O.ncol,
                              hipMemcpyDeviceToDevice, stream));
      HIPCHK(hipMemcpyAsync(N->frstate_d, O.frstate_d, sizeof(double) * 4, hipMemcpyDeviceToDevice, stream));
    }
    L[l] = std::move(N);   // the old level's memory is released here
    Level& V = *L[l];
    if (V.window) {
      ensure_dfull(l);
      crop_window(l, M, V.Mw[V.wsel], 4);
      crop_window(l, M, V.Mw[V.wsel ^ 1], 4);
      derived(l, V.wsel);                 // UpdateDerivedMacroFields of macro_new
      if (nmac > 4) crop_window(l, M + 4 * V.ncells, V.Dw[V.wsel ^ 1], 2);   // macro_old = FillPatch copy
      patch_derived(l, time, Dfull[l][V.istep & 1]);
      if (nmac > 4)
        HIPCHK(hipMemcpyAsync(Dfull[l][(V.istep + 1) & 1], M + 4 * V.ncells, sizeof(double) * 2 * V.ncells,
                              hipMemcpyDeviceToDevice, stream));
    }
    HIPCHK(hipStreamSynchronize(stream));
    dev_free(M);
  }

  // regrid(lbase, time) (AmrCore::regrid with ErrorEst and RemakeLevel):
  // boxes of levels lbase+1.. made finest first from the tags of their
  // parent (old data) and the new child footprint grown by 3, then the
  // changed levels remade coarse to fine
  void regrid(int lbase, double time) {
    const int bu = rg.blocking_factor / 2;
    int tlo[kMaxLev][3], thi[kMaxLev][3];
    int64_t tcnt[kMaxLev] = {0, 0, 0};
    for (int l = lbase; l < nlev - 1; ++l) tcnt[l] = tag_bbox(l, time, tlo[l], thi[l]);
    int nlo[kMaxLev][3], nhi[kMaxLev][3];
    for (int l = 0; l < nlev; ++l)
      for (int d = 0; d < 3; ++d) { nlo[l][d] = blo[l][d]; nhi[l][d] = bhi[l][d]; }
    for (int lf = nlev - 1; lf > lbase; --lf) {
      const int lc = lf - 1;
      int qlo[3], qhi[3];
      bool have = tcnt[lc] > 0;
      for (int d = 0; d < 3; ++d) { qlo[d] = tlo[lc][d]; qhi[d] = thi[lc][d]; }
      if (lf < nlev - 1) {
        for (int d = 0; d < dim; ++d) {
          const int clo_ = fdiv(fdiv(nlo[lf + 1][d], 2) - 3, 2), chi_ = fdiv(fdiv(nhi[lf + 1][d], 2) + 3, 2);
          if (!have || clo_ < qlo[d]) qlo[d] = clo_;
          if (!have || chi_ > qhi[d]) qhi[d] = chi_;
        }
        if (!have)
          for (int d = 0; d < dim; ++d) {
            qlo[d] = std::min(qlo[d], fdiv(nlo[lf][d], 2));
            qhi[d] = std::max(qhi[d], fdiv(nhi[lf][d], 2));
          }
        have = true;
      }
      if (!have) continue;
      int b0[3], b1[3];
      if (box_from_tags(qlo, qhi, nlo[lc], nhi[lc], bu, b0, b1))
        for (int d = 0; d < 3; ++d) { nlo[lf][d] = b0[d]; nhi[lf][d] = b1[d]; }
    }
    for (int lf = lbase + 2; lf < nlev; ++lf) {
      int c0[3], c1[3], b0[3], b1[3];
      for (int d = 0; d < 3; ++d) { c0[d] = fdiv(nlo[lf][d], 2); c1[d] = fdiv(nhi[lf][d], 2); }
      if (box_from_tags(c0, c1, nlo[lf - 1], nhi[lf - 1], bu, b0, b1))
        for (int d = 0; d < 3; ++d) { nlo[lf][d] = b0[d]; nhi[lf][d] = b1[d]; }
    }
    bool changed[kMaxLev] = {false, false, false};
    bool any = false;
    for (int l = lbase + 1; l < nlev; ++l)
      for (int d = 0; d < 3; ++d)
        if (nlo[l][d] != blo[l][d] || nhi[l][d] != bhi[l][d]) changed[l] = true;
    for (int l = lbase + 1; l < nlev; ++l) any = any || changed[l];
    if (!any) return;
    for (int l = lbase + 1; l < nlev; ++l)
      for (int d = 0; d < 3; ++d) { blo[l][d] = nlo[l][d]; bhi[l][d] = nhi[l][d]; }
    nest_geometry();
    for (int l = lbase + 1; l < nlev; ++l) {
      if (!changed[l]) continue;
      if (!changed[l - 1]) move_window(l - 1);
      remake(l, time);
    }
    if (changed[nlev - 1] && L[nlev - 1]->ib) ncol = L[nlev - 1]->ncol;
    ++nregrid;
  }

  void substep(int l, double time, bool last_chain, int row, bool export_state) {
    if (regrid_on && l < nlev - 1 && L[l]->istep > last_regrid[l] && L[l]->istep % regrid_int == 0) {
      regrid(l, time);
      for (int k = l; k < nlev; ++k) last_regrid[k] = L[k]->istep;
    }
    Level& V = *L[l];
    // the two steps before a possible regrid of this level keep their
    // macros and whole-level derived fields (tags, window moves, remakes)
    const int nxt = V.istep + 1;
    const bool pre = regrid_on && l < nlev - 1 && (nxt % regrid_int == 0 || (nxt + 1) % regrid_int == 0);
    if (pre) {
      V.ensure_exports();
      ensure_dfull(l);
    }
    export_state = export_state || pre;
    V.t_old = V.t_new;
    V.t_new += V.dt;
    if (V.window) V.wsel ^= 1;  // swap(macro_new, macro_old)
    if (l > 0) cf_fill(l, time, V.ib && V.dev_kin);
    if (V.ib) {
      V.rec_enable = last_chain;
      V.rec_override = true;
      V.rec_time = step_end;
      V.rec_step = (double)(istep + 1);
      V.rec_to_ghost = true;   // a single-iteration record rides with the stream-collide launch
      V.ib_update_forcing(time, row);
      V.rec_to_ghost = false;
    }
    if (!V.cf && !V.wrap) {
      // a stream-collide kernel that wrote the slabs of this step (level 0
      // walled in x only) leaves nothing to fill
      const int64_t npad = V.slabs.face_off[V.slabs.nface];
      if (npad && !V.slabs_fresh)
        V.ghostfn(dim3((unsigned)((npad + kGhostBlock - 1) / kGhostBlock)), dim3(kGhostBlock), V.stream,
                  V.ghost_args());
    }
    StepArgs sa = V.step_args(export_state);
    if (V.rec_held) {
      V.steprecfn(V.grid, V.block, stream, sa, V.held_a, V.held_r, V.held_nblk);
      V.rec_held = false;
    } else {
      V.stepfn(V.grid, V.block, stream, sa, nullptr, nullptr);
    }
    V.slabs_fresh = V.gfuse;
    exported[l] = export_state;
    V.cur ^= 1;
    const int step_lev = V.istep;
    if ((step_lev + 1) % regrid_int == 0) derived(l, V.wsel);
    ++V.istep;
    if (l < nlev - 1) {
      for (int i = 1; i <= 2; ++i)
        substep(l + 1, time + (double)(i - 1) * L[l + 1]->dt, last_chain && i == 2, row, i == 2);
      average_down(l);
      derived(l, V.wsel);
    }
    if (pre) patch_derived(l, V.t_new, Dfull[l][V.istep & 1]);
    HIPCHK(hipGetLastError());
  }

  void advance(int nsteps, bool want_state) {
    if (nsteps < 0) throw StatusError{ILBM_ERR_INVALID, "nsteps < 0"};
    if (nsteps == 0) return;
    if (nsteps > mass_cap) {
      dev_free(mass_dev);
      mass_cap = std::max(nsteps, 2 * mass_cap);
      mass_dev = dev_alloc<double>(mass_cap);
    }
    L[nlev - 1]->grow_histories(1, nsteps);
    Level& C0 = *L[0];
    if (want_state) C0.ensure_exports();
    for (int s = 0; s < nsteps; ++s) {
      const double time = C0.t_new;
      step_end = time + C0.dt;
      substep(0, time, true, s, want_state && s == nsteps - 1);
      ++istep;
      hipLaunchKernelGGL(k_mass_total, dim3(1), dim3(1024), 0, stream, (const double*)C0.partial, C0.nblocks,
                         (const double*)dpartial[0], (nlev > 1) ? dblocks[0] : (int64_t)0, mass_dev + s);
      const double cur = step_end;
      for (int l = 0; l < nlev; ++l) L[l]->t_new = cur;
      HIPCHK(hipGetLastError());
    }
    std::vector<double> mh(nsteps);
    HIPCHK(hipMemcpyAsync(mh.data(), mass_dev, sizeof(double) * nsteps, hipMemcpyDeviceToHost, stream));
    std::vector<double> fh;
    const Level& F = *L[nlev - 1];
    if (ncol) {
      fh.resize((size_t)nsteps * ncol);
      HIPCHK(hipMemcpyAsync(fh.data(), F.force_dev, sizeof(double) * nsteps * ncol, hipMemcpyDeviceToHost, stream));
    }
    HIPCHK(hipStreamSynchronize(stream));
    mass_host.insert(mass_host.end(), mh.begin(), mh.end());
    force_host.insert(force_host.end(), fh.begin(), fh.end());
    for (int l = 0; l < nlev; ++l) L[l]->state_valid = (l > 0) || want_state;
  }
};

}  // namespace ilbm

// ---------------------------------------------------------------------------
// C ABI of the hierarchy (include/ilbm.h)
// ---------------------------------------------------------------------------
struct ilbm_amr {
  ilbm::Amr A;
};

extern "C" {

int ilbm_amr_create(const ilbm_case_desc* desc, int nlev, const int* box_lo, const int* box_hi, int regrid_int,
                    const double init_state[4], ilbm_amr** out) {
  if (!desc || !init_state || !out || (nlev > 1 && (!box_lo || !box_hi))) return ILBM_ERR_INVALID;
  *out = nullptr;
  ilbm_amr* h = new (std::nothrow) ilbm_amr();
  if (!h) return ILBM_ERR_ALLOC;
  const int st = guarded([&] {
    HIPCHK(hipStreamCreateWithFlags(&h->A.stream, hipStreamNonBlocking));
    h->A.create(*desc, nlev, box_lo, box_hi, regrid_int, init_state);
  });
  if (st != ILBM_OK) {
    delete h;
    return st;
  }
  *out = h;
  return ILBM_OK;
}

int ilbm_amr_destroy(ilbm_amr* h) {
  if (!h) return ILBM_ERR_INVALID;
  delete h;
  return ILBM_OK;
}

int ilbm_amr_advance(ilbm_amr* h, int nsteps, int want_state) {
  if (!h) return ILBM_ERR_INVALID;
  return guarded([&] { h->A.advance(nsteps, want_state != 0); });
}

int ilbm_amr_level_cells(ilbm_amr* h, int lev, int64_t* cells) {
  if (!h || !cells || lev < 0 || lev >= h->A.nlev) return ILBM_ERR_INVALID;
  *cells = h->A.L[lev]->ncells;
  return ILBM_OK;
}

int ilbm_amr_get_f(ilbm_amr* h, int lev, double* out) {
  if (!h || !out || lev < 0 || lev >= h->A.nlev) return ILBM_ERR_INVALID;
  const ilbm::Level& V = *h->A.L[lev];
  if (!V.state_valid || !V.fout) return ILBM_ERR_STATE;
  return guarded([&] {
    HIPCHK(hipMemcpy(out, V.fout, sizeof(double) * V.c.q * V.ncells, hipMemcpyDeviceToHost));
  });
}

int ilbm_amr_get_macro(ilbm_amr* h, int lev, double* out) {
  if (!h || !out || lev < 0 || lev >= h->A.nlev) return ILBM_ERR_INVALID;
  const ilbm::Level& V = *h->A.L[lev];
  if (!V.state_valid || !V.mout) return ILBM_ERR_STATE;
  return guarded([&] {
    HIPCHK(hipMemcpy(out, V.mout, sizeof(double) * 4 * V.ncells, hipMemcpyDeviceToHost));
  });
}

int ilbm_amr_mass_history(ilbm_amr* h, double* out, int max_rows, int* n_out) {
  if (!h || !n_out) return ILBM_ERR_INVALID;
  const auto& v = h->A.mass_host;
  *n_out = (int)v.size();
  if (out) std::copy(v.begin(), v.begin() + std::min<size_t>(v.size(), (size_t)std::max(max_rows, 0)), out);
  return ILBM_OK;
}

int ilbm_amr_force_history(ilbm_amr* h, double* out, int max_rows, int* n_rows, int* n_cols) {
  if (!h || !n_rows || !n_cols) return ILBM_ERR_INVALID;
  const ilbm::Amr& A = h->A;
  *n_cols = A.ncol;
  *n_rows = A.ncol ? (int)(A.force_host.size() / A.ncol) : 0;
  if (out && A.ncol) {
    const size_t n = std::min<size_t>(A.force_host.size(), (size_t)std::max(max_rows, 0) * A.ncol);
    std::copy(A.force_host.begin(), A.force_host.begin() + n, out);
  }
  return ILBM_OK;
}

int ilbm_amr_num_markers(ilbm_amr* h, int* out) {
  if (!h || !out) return ILBM_ERR_INVALID;
  const ilbm::Level& F = *h->A.L[h->A.nlev - 1];
  *out = F.ib ? F.nm : 0;
  return ILBM_OK;
}

int ilbm_amr_get_derived(ilbm_amr* h, int lev, double* out) {
  if (!h || !out || lev < 0 || lev >= h->A.nlev) return ILBM_ERR_INVALID;
  ilbm::Amr& A = h->A;
  ilbm::Level& V = *A.L[lev];
  if (!V.state_valid || !V.mout) return ILBM_ERR_STATE;
  return guarded([&] {
    if (2 * V.ncells > A.scratch_cap) {
      ilbm::dev_free(A.scratch);
      A.scratch = ilbm::dev_alloc<double>(2 * V.ncells);
      A.scratch_cap = 2 * V.ncells;
    }
    A.patch_derived(lev, V.t_new, A.scratch);
    HIPCHK(hipMemcpyAsync(out, A.scratch, sizeof(double) * 2 * V.ncells, hipMemcpyDeviceToHost, A.stream));
    HIPCHK(hipStreamSynchronize(A.stream));
  });
}

int ilbm_amr_set_regrid(ilbm_amr* h, const ilbm_regrid_params* p) {
  if (!h || !p) return ILBM_ERR_INVALID;
  return guarded([&] { h->A.set_regrid(*p); });
}

int ilbm_amr_boxes(ilbm_amr* h, int* lo, int* hi, int* n_regrids) {
  if (!h) return ILBM_ERR_INVALID;
  const ilbm::Amr& A = h->A;
  for (int l = 1; l < A.nlev; ++l)
    for (int d = 0; d < 3; ++d) {
      if (lo) lo[3 * (l - 1) + d] = A.blo[l][d];
      if (hi) hi[3 * (l - 1) + d] = A.bhi[l][d];
    }
  if (n_regrids) *n_regrids = A.nregrid;
  return ILBM_OK;
}

int ilbm_amr_info(ilbm_amr* h, double out[8]) {
  if (!h || !out) return ILBM_ERR_INVALID;
  const ilbm::Amr& A = h->A;
  for (int i = 0; i < 8; ++i) out[i] = 0.0;
  out[0] = A.nlev;
  out[1] = A.istep;
  out[2] = A.L[0]->t_new;
  for (int l = 0; l < A.nlev && l < 3; ++l) out[3 + l] = (double)A.L[l]->ncells;
  out[6] = A.L[A.nlev - 1]->tau;
  out[7] = A.L[0]->tau;
  return ILBM_OK;
}

}  // extern "C"
