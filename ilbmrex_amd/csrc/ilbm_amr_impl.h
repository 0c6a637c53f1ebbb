// AMR hierarchy on one device (compiled as part of ilbm_level.hip): level 0
// is the case's domain, each finer level one box at refinement ratio 2,
// advanced with subcycling as timeStepWithSubcycling does
// (Source/Core/AmrCoreLBM.cpp:1688-1754), regridding left out (the boxes
// are fixed).  Every level is a resident Level; a refined level has ghost
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
using AvgFn = void (*)(dim3, hipStream_t, const AvgArgs&);

template <class LAT, class ST>
static void launch_cf(dim3 g, hipStream_t s, const CfArgs& a) {
  hipLaunchKernelGGL((k_cf_fill<LAT, ST>), g, dim3(kBlock), 0, s, a);
}
template <class LAT, class ST>
static void launch_avg(dim3 g, hipStream_t s, const AvgArgs& a) {
  hipLaunchKernelGGL((k_average_down<LAT, ST>), g, dim3(256), 0, s, a);
}

static void pick_amr(int lat, int st, CfFn& cf, AvgFn& av) {
  auto pk = [&](auto latt) {
    using L = decltype(latt);
    if (st == 0) { cf = launch_cf<L, StF64>; av = launch_avg<L, StF64>; }
    else if (st == 1) { cf = launch_cf<L, StF32>; av = launch_avg<L, StF32>; }
    else { cf = launch_cf<L, StF16S>; av = launch_avg<L, StF16S>; }
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
  AvgFn avgfn = nullptr;
  int istep = 0;
  double step_end = 0.0;
  double* mass_dev = nullptr;
  int mass_cap = 0;
  std::vector<double> mass_host, force_host;
  int ncol = 0;

  ~Amr() {
    if (stream) (void)hipStreamSynchronize(stream);
    L.clear();
    for (int l = 0; l < kMaxLev; ++l) dev_free(dpartial[l]);
    dev_free(mass_dev);
    if (stream) (void)hipStreamDestroy(stream);
  }

  void create(const ilbm_case_desc& d0, int nl, const int* box_lo, const int* box_hi, int rint,
              const double st[4]) {
    if (nl < 1 || nl > kMaxLev) throw StatusError{ILBM_ERR_UNSUPPORTED, "1 to 3 levels"};
    if (rint < 1) throw StatusError{ILBM_ERR_INVALID, "regrid_int must be >= 1"};
    if (d0.nmac != 4 && d0.nmac != 6) throw StatusError{ILBM_ERR_UNSUPPORTED, "nmac must be 4 or 6"};
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
    // nesting: a fine box covers whole coarse cells and its window (the
    // coarse footprint grown by 3) lies in the coarse box
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
    for (int l = 0; l < nlev; ++l) {
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
      if (l < nlev - 1) {
        V->window = true;
        int64_t cells = 1;
        for (int d = 0; d < 3; ++d) {
          V->band_lo[d] = wlo[l][d] - V->ioff[d];
          V->band_n[d] = wn[l][d];
          cells *= wn[l][d];
        }
        V->band_cells = cells;
        for (int b = 0; b < 2; ++b) {
          V->Mw[b] = dev_alloc<double>(4 * cells);
          V->Dw[b] = dev_alloc<double>(2 * cells);
          HIPCHK(hipMemsetAsync(V->Dw[b], 0, sizeof(double) * 2 * cells, stream));
        }
        const int64_t ncov = (int64_t)cn[l][0] * cn[l][1] * cn[l][2];
        dblocks[l] = (ncov + 255) / 256;
        dpartial[l] = dev_alloc<double>(dblocks[l]);
      }
      V->setup_ib();
      std::vector<double> m0(4 * (size_t)V->ncells);
      for (int n = 0; n < 4; ++n) std::fill(m0.begin() + n * V->ncells, m0.begin() + (n + 1) * V->ncells, st[n]);
      V->init(m0.data());
      if (l > 0) V->ensure_exports();
      L.push_back(std::move(V));
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
    pick_amr(L[0]->lat, L[0]->st, cffn, avgfn);
    HIPCHK(hipStreamSynchronize(stream));
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

  void cf_fill(int l, double time) {
    Level& V = *L[l];
    CfArgs a{};
    a.g = V.g;
    a.slabs = V.slabs;
    a.gbuf = V.gbuf[V.cur];
    for (int d = 0; d < 3; ++d) a.ioff[d] = (d < dim) ? V.ioff[d] : 0;
    a.ncomp = V.c.nmac;
    window_at(l - 1, time, a.W);
    const int64_t npad = V.slabs.face_off[V.slabs.nface];
    cffn(dim3((unsigned)((npad + kBlock - 1) / kBlock)), stream, a);
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
  }

  void substep(int l, double time, bool last_chain, int row, bool export_state) {
    Level& V = *L[l];
    V.t_old = V.t_new;
    V.t_new += V.dt;
    if (V.window) V.wsel ^= 1;  // swap(macro_new, macro_old)
    if (l > 0) cf_fill(l, time);
    if (V.ib) {
      V.rec_enable = last_chain;
      V.rec_override = true;
      V.rec_time = step_end;
      V.rec_step = (double)(istep + 1);
      V.ib_update_forcing(time, row);
    }
    if (!V.cf && !V.wrap) {
      GhostArgs ga = V.ghost_args();
      const int64_t npad = V.slabs.face_off[V.slabs.nface];
      if (npad) V.ghostfn(dim3((unsigned)((npad + 255) / 256)), dim3(256), V.stream, ga);
    }
    StepArgs sa = V.step_args(export_state);
    V.stepfn(V.grid, V.block, stream, sa);
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
    Level& F = *L[nlev - 1];
    F.grow_histories(1, nsteps);
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
