// Coarse-fine transfer kernels of the AMR hierarchy (one box per level,
// refinement ratio 2), following Source/Core/AmrCoreLBM.cpp:
//
//   k_cf_fill        fine ghost layer = feq of FillPatchMacro(lev, time): the
//                    coarse macros of a window around the fine box, linearly
//                    interpolated in time (GetDataMacro, :2239-2269) and in
//                    space with AMReX's cell_cons_interp (FillPatchTwoLevels,
//                    :2180-2197), then the equilibrium that
//                    FillPatchMesoscopic writes into every ghost (:1562-1624)
//   k_average_down   AverageDownTo (:1520-1526): covered coarse cells get the
//                    mean of their fine children's f and macros; their
//                    post-collision state is recomputed from those values, as
//                    the next coarse step collides them
//   k_derived        UpdateDerivedMacroFields / visPara (:150-168,
//                    visPara.H:10-54) on the window: vorticity, pressure
//
// The window of a coarse level holds the macro components the fine ghost
// interpolation reads: rho, u, v, w (written by the coarse stream-collide
// kernel as its band) and vorticity, pressure (k_derived).  Two copies, old
// and new, as macro_old / macro_new.
#pragma once

#include "ilbm_lbm_kernels.h"

namespace ilbm {

// an Array4-style fab: comps planes over [lo, hi], x fastest
struct PatchFab {
  double* p;
  int lo[3], hi[3];
  int64_t js, ks, ns;
  __device__ double& at(int i, int j, int k, int n) const {
    return p[(int64_t)(i - lo[0]) + (int64_t)(j - lo[1]) * js + (int64_t)(k - lo[2]) * ks + (int64_t)n * ns];
  }
};

struct Window {
  const double* Mo;   // [4][cells] macro_old
  const double* Mn;   // [4][cells] macro_new
  const double* Do;   // [2][cells] vorticity, pressure of macro_old
  const double* Dn;
  int lo[3];          // window box in coarse global cell indices
  int n[3];
  int64_t cells;
  int mode;           // 0: old, 1: new, 2: a*old + b*new
  double a, b;
};

// MODE is W.mode, fixed at compile time so that no branch separates the
// loads of one interpolation
template <int MODE>
__device__ inline double window_val(const Window& W, int i, int j, int k, int comp) {
  const int64_t w = ((int64_t)(k - W.lo[2]) * W.n[1] + (j - W.lo[1])) * W.n[0] + (i - W.lo[0]);
  const double* o = (comp < 4) ? W.Mo + comp * W.cells : W.Do + (comp - 4) * W.cells;
  const double* nw = (comp < 4) ? W.Mn + comp * W.cells : W.Dn + (comp - 4) * W.cells;
  if constexpr (MODE == 0) return o[w];
  else if constexpr (MODE == 1) return nw[w];
  else return W.a * o[w] + W.b * nw[w];
}

__device__ inline int floordiv2(int i) { return (i >= 0) ? i / 2 : -((1 - i) / 2); }

// cell_cons_interp at fine global cell (I, J, K): slopes limited per
// direction with one factor shared by the NC components; all NC components
// are written to out.  Every window value is loaded before the slopes are
// formed, so the loads of all components and directions are in flight
// together.
template <int DIM, int NC, int MODE>
__device__ inline void cons_interp_m(const Window& W, int I, int J, int K, double* out) {
  const int ic = floordiv2(I), jc = floordiv2(J), kc = (DIM == 3) ? floordiv2(K) : 0;
  double u0[NC], up[DIM][NC], um[DIM][NC];
#pragma unroll
  for (int n = 0; n < NC; ++n) u0[n] = window_val<MODE>(W, ic, jc, kc, n);
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const int pi = ic + (d == 0), pj = jc + (d == 1), pk = kc + (d == 2);
    const int mi = ic - (d == 0), mj = jc - (d == 1), mk = kc - (d == 2);
#pragma unroll
    for (int n = 0; n < NC; ++n) {
      up[d][n] = window_val<MODE>(W, pi, pj, pk, n);
      um[d][n] = window_val<MODE>(W, mi, mj, mk, n);
    }
  }
  double s[DIM][NC];
  double sf[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    sf[d] = 1.0;
#pragma unroll
    for (int n = 0; n < NC; ++n) {
      const double dc = 0.5 * (up[d][n] - um[d][n]);
      const double df = 2.0 * (up[d][n] - u0[n]);
      const double db = 2.0 * (u0[n] - um[d][n]);
      double sx = (df * db >= 0.0) ? fmin(fabs(df), fabs(db)) : 0.0;
      sx = copysign(1.0, dc) * fmin(sx, fabs(dc));
      if (dc != 0.0) sf[d] = fmin(sf[d], sx / dc);
      s[d][n] = dc;
    }
  }
  const double ox = ((double)(I - ic * 2) + 0.5) / 2.0 - 0.5;
  const double oy = ((double)(J - jc * 2) + 0.5) / 2.0 - 0.5;
  const double oz = (DIM == 3) ? ((double)(K - kc * 2) + 0.5) / 2.0 - 0.5 : 0.0;
#pragma unroll
  for (int n = 0; n < NC; ++n) {
    double v = u0[n] + ox * (s[0][n] * sf[0]);
    v = v + oy * (s[1][n] * sf[1]);
    if constexpr (DIM == 3) v = v + oz * (s[2][n] * sf[2]);
    out[n] = v;
  }
}

template <int DIM, int NC>
__device__ inline void cons_interp(const Window& W, int I, int J, int K, double* out) {
  if (W.mode == 0) cons_interp_m<DIM, NC, 0>(W, I, J, K, out);
  else if (W.mode == 1) cons_interp_m<DIM, NC, 1>(W, I, J, K, out);
  else cons_interp_m<DIM, NC, 2>(W, I, J, K, out);
}

// the same with the dimension and component count (4 or 6) known at run time
__device__ inline void cons_interp_rt(const Window& W, int dim, int ncomp, int I, int J, int K, double* out) {
  if (dim == 3) {
    if (ncomp == 6) cons_interp<3, 6>(W, I, J, K, out);
    else cons_interp<3, 4>(W, I, J, K, out);
  } else {
    if (ncomp == 6) cons_interp<2, 6>(W, I, J, K, out);
    else cons_interp<2, 4>(W, I, J, K, out);
  }
}

struct CfArgs {
  GridGeom g;          // fine level
  Slabs slabs;
  void* gbuf;          // fine populations; the ghost slabs follow the Q planes
  int ioff[3];         // global index of the fine level's local cell 0
  int ncomp;           // macro components interpolated together (nmac: 4 or 6)
  Window W;
};

template <class LAT, class ST>
__device__ __forceinline__ void cf_fill_body(const CfArgs& a, int64_t t) {
  using T = typename ST::compute_t;
  using R = typename ST::raw;
  const int64_t npad = a.slabs.face_off[a.slabs.nface];
  if (t >= npad) return;
  int iv[3];
  slab_cell_of(a.slabs, a.g.n, t, iv);
  double m[6];
  if (a.ncomp == 6) cons_interp<LAT::D, 6>(a.W, iv[0] + a.ioff[0], iv[1] + a.ioff[1], iv[2] + a.ioff[2], m);
  else cons_interp<LAT::D, 4>(a.W, iv[0] + a.ioff[0], iv[1] + a.ioff[1], iv[2] + a.ioff[2], m);
  // FillPatchMesoscopic's equilibrium (AmrCoreLBM.cpp:1605-1620)
  const double r = fmax(m[0], 1.0e-12);
  const double u0 = m[1], v0 = m[2], w0 = m[3];
  const double u2 = u0 * u0 + v0 * v0 + w0 * w0;
  R* out = static_cast<R*>(a.gbuf) + a.g.pop_n;
#pragma unroll
  for (int q = 0; q < LAT::Q; ++q) {
    const double cu = double(LAT::cx[q]) * u0 + double(LAT::cy[q]) * v0 + double(LAT::cz[q]) * w0;
    const double fe = LAT::w[q] * r * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * u2);
    ST::store(out, (int64_t)q * npad + t, T(fe), T(LAT::w[q]));
  }
}

template <class LAT, class ST>
__global__ void __launch_bounds__(kBlock) k_cf_fill(CfArgs a) {
  cf_fill_body<LAT, ST>(a, (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

struct AvgArgs {
  GridGeom gc;          // coarse level
  void* gbuf;           // coarse post-collision populations (overwritten on covered cells)
  int clo[3], cn[3];    // covered coarse cells (coarse local indices)
  int64_t ncov;
  const double* fout;   // fine f_new [Q][fine cells]
  const double* mout;   // fine macro_new [4][fine cells]
  int fn[3];            // fine level extents
  int fofs[3];          // fine local index of the first child of coarse cell clo
  double* Mw;           // coarse window macro_new [4][window cells]
  int wofs[3];          // window-local index of coarse cell clo
  int wn[3];
  int64_t wcells;
  double* cfout;        // optional coarse exports [Q][cells], [4][cells]
  double* cmout;
  double tau, inv_tau;
  double Fcol[3];       // uniform force of the coarse collision
  double* dpartial;     // per-block sum of (rho_avg - rho) for Sum(Rho)
};

template <class LAT, class ST>
__global__ void __launch_bounds__(256) k_average_down(AvgArgs a) {
  using T = typename ST::compute_t;
  using R = typename ST::raw;
  __shared__ double lds[4];
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double drho = 0.0;
  if (t < a.ncov) {
    const int ci = (int)(t % a.cn[0]);
    const int cj = (int)((t / a.cn[0]) % a.cn[1]);
    const int ck = (int)(t / ((int64_t)a.cn[0] * a.cn[1]));
    const int fk0 = (LAT::D == 3) ? a.fofs[2] + 2 * ck : 0;
    const int fj0 = a.fofs[1] + 2 * cj, fi0 = a.fofs[0] + 2 * ci;
    const int nkr = (LAT::D == 3) ? 2 : 1;
    const int64_t nf = (int64_t)a.fn[0] * a.fn[1] * a.fn[2];
    const double volfrac = (LAT::D == 3) ? 0.125 : 0.25;
    double fa[LAT::Q], ma[4];
#pragma unroll
    for (int q = 0; q < LAT::Q; ++q) {
      double c = 0.0;
      for (int kr = 0; kr < nkr; ++kr)
        for (int jr = 0; jr < 2; ++jr)
          for (int ir = 0; ir < 2; ++ir)
            c += a.fout[(int64_t)q * nf + ((int64_t)(fk0 + kr) * a.fn[1] + (fj0 + jr)) * a.fn[0] + fi0 + ir];
      fa[q] = volfrac * c;
    }
    for (int n = 0; n < 4; ++n) {
      double c = 0.0;
      for (int kr = 0; kr < nkr; ++kr)
        for (int jr = 0; jr < 2; ++jr)
          for (int ir = 0; ir < 2; ++ir)
            c += a.mout[(int64_t)n * nf + ((int64_t)(fk0 + kr) * a.fn[1] + (fj0 + jr)) * a.fn[0] + fi0 + ir];
      ma[n] = volfrac * c;
    }
    const int64_t w = ((int64_t)(a.wofs[2] + ck) * a.wn[1] + (a.wofs[1] + cj)) * a.wn[0] + a.wofs[0] + ci;
    drho = ma[0] - a.Mw[w];
    for (int n = 0; n < 4; ++n) a.Mw[n * a.wcells + w] = ma[n];
    const int i = a.clo[0] + ci, j = a.clo[1] + cj, k = a.clo[2] + ck;
    const int64_t cell = ((int64_t)k * a.gc.n[1] + j) * a.gc.n[0] + i;
    if (a.cmout) {
      const int64_t nc = (int64_t)a.gc.n[0] * a.gc.n[1] * a.gc.n[2];
      for (int n = 0; n < 4; ++n) a.cmout[n * nc + cell] = ma[n];
#pragma unroll
      for (int q = 0; q < LAT::Q; ++q) a.cfout[(int64_t)q * nc + cell] = fa[q];
    }
    // the coarse collision of the averaged state (AdvanceLBMAtLevel.cpp:196-240)
    const T r = T(ma[0]), u = T(ma[1]), v = T(ma[2]), ww = T(ma[3]);
    const T uv = usq<LAT, T>(u, v, ww);
    const T cfx = T(a.Fcol[0]), cfy = T(a.Fcol[1]), cfz = T(a.Fcol[2]);
    const bool forced = !(cfx == T(0.0) && cfy == T(0.0) && cfz == T(0.0));
    const T fac = T(1.0) - T(0.5) / T(a.tau);
    R* gb = static_cast<R*>(a.gbuf);
    const int64_t self = a.gc.idx(i, j, k);
#pragma unroll
    for (int q = 0; q < LAT::Q; ++q) {
      const T fe = feq_q<LAT, T>(q, r, u, v, ww, uv);
      T gq = bgk<ST>(T(fa[q]), fe, a.tau, a.inv_tau);
      if (forced) gq += guo_q<LAT, T>(q, fac, u, v, ww, cfx, cfy, cfz);
      ST::store(gb, (int64_t)q * a.gc.qs + self, gq, T(LAT::w[q]));
    }
  }
  const double s = wave_sum(drho);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) a.dpartial[blockIdx.x] = ((lds[0] + lds[1]) + lds[2]) + lds[3];
}

// vorticity and pressure on the window shrunk by one cell (visPara.H:10-54)
struct DerArgs {
  const double* M;   // [4][window]
  double* D;         // [2][window]
  int n[3];
  int64_t cells;
  int dim;
  double dx, dy, dz;
};

__global__ void __launch_bounds__(256) k_derived(DerArgs a) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int ni = a.n[0] - 2, nj = a.n[1] - 2, nk = (a.dim == 3) ? a.n[2] - 2 : 1;
  if (t >= (int64_t)ni * nj * nk) return;
  const int i = 1 + (int)(t % ni), j = 1 + (int)((t / ni) % nj);
  const int k = (a.dim == 3) ? 1 + (int)(t / ((int64_t)ni * nj)) : 0;
  auto at = [&](int ii, int jj, int kk, int c) {
    return a.M[c * a.cells + ((int64_t)kk * a.n[1] + jj) * a.n[0] + ii];
  };
  const int64_t w = ((int64_t)k * a.n[1] + j) * a.n[0] + i;
  const double P = (at(i, j, k, 0) - 1.0) / 3.0;
  const double dv_dx = (at(i + 1, j, k, 2) - at(i - 1, j, k, 2)) / (2.0 * a.dx);
  const double du_dy = (at(i, j + 1, k, 1) - at(i, j - 1, k, 1)) / (2.0 * a.dy);
  double vort;
  if (a.dim == 3) {
    const double du_dz = (at(i, j, k + 1, 1) - at(i, j, k - 1, 1)) / (2.0 * a.dz);
    const double dw_dx = (at(i + 1, j, k, 3) - at(i - 1, j, k, 3)) / (2.0 * a.dx);
    const double dv_dz = (at(i, j, k + 1, 2) - at(i, j, k - 1, 2)) / (2.0 * a.dz);
    const double dw_dy = (at(i, j + 1, k, 3) - at(i, j - 1, k, 3)) / (2.0 * a.dy);
    const double ox = dw_dy - dv_dz, oy = du_dz - dw_dx, oz = dv_dx - du_dy;
    vort = sqrt(ox * ox + oy * oy + oz * oz);
  } else {
    vort = dv_dx - du_dy;
  }
  a.D[w] = vort;
  a.D[a.cells + w] = P;
}

// ---------------------------------------------------------------------------
// Regrid (ErrorEst AmrCoreLBM.cpp:986-1180, RemakeLevel :835-914)
// ---------------------------------------------------------------------------

// FillPatchMacro of a refined level with one ghost layer, the ghost cells:
// cell_cons_interp of the 4 macro components from the coarser level's
// window (FillPatchTwoLevels).  P holds the level in local indices.
struct PatchCfArgs {
  PatchFab P;
  int n[3];          // valid extents
  int ioff[3];       // global index of local cell 0
  int dim;
  Window W;
};

__global__ void k_patch_cf(PatchCfArgs a) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.P.ns) return;
  const int i = a.P.lo[0] + (int)(t % a.P.js);
  const int j = a.P.lo[1] + (int)((t / a.P.js) % (a.P.hi[1] - a.P.lo[1] + 1));
  const int k = a.P.lo[2] + (int)(t / a.P.ks);
  if (i >= 0 && i < a.n[0] && j >= 0 && j < a.n[1] && k >= 0 && k < a.n[2]) return;
  double m[4];
  if (a.dim == 3) cons_interp<3, 4>(a.W, i + a.ioff[0], j + a.ioff[1], k + a.ioff[2], m);
  else cons_interp<2, 4>(a.W, i + a.ioff[0], j + a.ioff[1], k + a.ioff[2], m);
  for (int c = 0; c < 4; ++c) a.P.p[c * a.P.ns + t] = m[c];
}

// vorticity and pressure (visPara.H:10-54) of every valid cell of a level
// from its one-ghost macro patch, into D[2][cells]
__global__ void k_patch_derived(PatchFab P, double* D, int nx, int ny, int64_t cells, int dim, double dx, double dy,
                                double dz) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cells) return;
  const int i = (int)(t % nx), j = (int)((t / nx) % ny), k = (int)(t / ((int64_t)nx * ny));
  const double Pr = (P.at(i, j, k, 0) - 1.0) / 3.0;
  const double dv_dx = (P.at(i + 1, j, k, 2) - P.at(i - 1, j, k, 2)) / (2.0 * dx);
  const double du_dy = (P.at(i, j + 1, k, 1) - P.at(i, j - 1, k, 1)) / (2.0 * dy);
  double vort;
  if (dim == 3) {
    const double du_dz = (P.at(i, j, k + 1, 1) - P.at(i, j, k - 1, 1)) / (2.0 * dz);
    const double dw_dx = (P.at(i + 1, j, k, 3) - P.at(i - 1, j, k, 3)) / (2.0 * dx);
    const double dv_dz = (P.at(i, j, k + 1, 2) - P.at(i, j, k - 1, 2)) / (2.0 * dz);
    const double dw_dy = (P.at(i, j + 1, k, 3) - P.at(i, j - 1, k, 3)) / (2.0 * dy);
    const double ox = dw_dy - dv_dz, oy = du_dz - dw_dx, oz = dv_dx - du_dy;
    vort = sqrt(ox * ox + oy * oy + oz * oz);
  } else {
    vort = dv_dx - du_dy;
  }
  D[t] = vort;
  D[cells + t] = Pr;
}

// order-preserving 64-bit key of a double (NaN aside), for atomicMax
__device__ inline unsigned long long order_key(double x) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(x);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ inline double key_value(unsigned long long k) {
  const unsigned long long u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return __longlong_as_double((long long)u);
}

// MultiFab::max of the vorticity: block maxima folded with atomicMax
__global__ void __launch_bounds__(256) k_vort_max(const double* v, int64_t cells, unsigned long long* key) {
  __shared__ double lds[4];
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double m = (t < cells) ? v[t] : -HUGE_VAL;
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double b = fmax(fmax(lds[0], lds[1]), fmax(lds[2], lds[3]));
    atomicMax(key, order_key(b));
  }
}

// ErrorEst's tags on one level (valid cells): |vorticity| > ratio * max,
// the level-set band of a cylinder / sphere (phi = r - R < 5.5 dx), the
// body refinement region, all cleared in the no-refine x bands.  Output:
// tag count and the bounding box of the tags grown by n_error_buf, in
// level-global indices (atomics on ints: order free).
struct TagArgs {
  const double* vort;   // [cells]
  const unsigned long long* vmax_key;
  int n[3], ioff[3], dim;
  int64_t cells;
  double ratio;
  double plo[3], dx[3];
  double probhi0;
  double nr_lo, nr_hi;  // no_refine_xlo / xhi
  int region;           // body refinement box in use
  double rlo[3], rhi[3];
  int lset;             // level-set band of the cylinder / sphere
  double x0, y0, z0, R, band;
  int buf;
  int* out;             // [0] count, [1..3] lo, [4..6] hi
};

__global__ void k_tag_bbox(TagArgs a) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.cells) return;
  const int i = a.ioff[0] + (int)(t % a.n[0]);
  const int j = a.ioff[1] + (int)((t / a.n[0]) % a.n[1]);
  const int k = a.ioff[2] + (int)(t / ((int64_t)a.n[0] * a.n[1]));
  const double x = a.plo[0] + ((double)i + 0.5) * a.dx[0];
  const double y = a.plo[1] + ((double)j + 0.5) * a.dx[1];
  const double z = (a.dim == 3) ? a.plo[2] + ((double)k + 0.5) * a.dx[2] : 0.0;
  if (a.nr_lo > 0.0 || a.nr_hi > 0.0) {
    const double xlo_block = a.plo[0] + a.nr_lo, xhi_block = a.probhi0 - a.nr_hi;
    if (x <= xlo_block || x >= xhi_block) return;
  }
  const double thr = a.ratio * key_value(*a.vmax_key);
  bool tag = fabs(a.vort[t]) > thr;
  if (a.lset) {
    double r2 = (x - a.x0) * (x - a.x0) + (y - a.y0) * (y - a.y0);
    if (a.dim == 3) r2 += (z - a.z0) * (z - a.z0);
    if (sqrt(r2) - a.R < a.band) tag = true;
  }
  if (a.region && x >= a.rlo[0] && x <= a.rhi[0] && y >= a.rlo[1] && y <= a.rhi[1] && z >= a.rlo[2] &&
      z <= a.rhi[2])
    tag = true;
  if (!tag) return;
  atomicAdd(&a.out[0], 1);
  const int iv[3] = {i, j, k};
  for (int d = 0; d < a.dim; ++d) {
    atomicMin(&a.out[1 + d], iv[d] - a.buf);
    atomicMax(&a.out[4 + d], iv[d] + a.buf);
  }
}

// RemakeLevel's FillPatchMacro over the new box (valid cells): the old
// level's macro_new where it was valid (comps 0-3 from its export, 4-5
// from its derived fields), cell_cons_interp of all nmac components from
// the coarser level's window elsewhere.
struct RemakeArgs {
  double* M;            // [nmac][new cells]
  int n[3], lo[3];      // new box extents and global lo
  const double* om;     // old [4][old cells]
  const double* od;     // old [2][old cells] or null
  int on[3], olo[3];
  int dim, nmac;
  Window W;
};

__global__ void k_remake_macro(RemakeArgs a) {
  const int64_t cells = (int64_t)a.n[0] * a.n[1] * a.n[2];
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cells) return;
  const int I = a.lo[0] + (int)(t % a.n[0]);
  const int J = a.lo[1] + (int)((t / a.n[0]) % a.n[1]);
  const int K = a.lo[2] + (int)(t / ((int64_t)a.n[0] * a.n[1]));
  const int oi = I - a.olo[0], oj = J - a.olo[1], ok = K - a.olo[2];
  if (oi >= 0 && oi < a.on[0] && oj >= 0 && oj < a.on[1] && ok >= 0 && ok < a.on[2]) {
    const int64_t ocells = (int64_t)a.on[0] * a.on[1] * a.on[2];
    const int64_t o = ((int64_t)ok * a.on[1] + oj) * a.on[0] + oi;
    for (int c = 0; c < 4; ++c) a.M[c * cells + t] = a.om[c * ocells + o];
    for (int c = 4; c < a.nmac; ++c) a.M[c * cells + t] = a.od ? a.od[(c - 4) * ocells + o] : 0.0;
    return;
  }
  double m[6];
  cons_interp_rt(a.W, a.dim, a.nmac, I, J, K, m);
  for (int c = 0; c < a.nmac; ++c) a.M[c * cells + t] = m[c];
}

// copy of a box of a [nc][src cells] array into a [nc][dst cells] array
// (dst box inside the src box; offsets in src indices)
__global__ void k_crop(const double* src, int sn0, int sn1, int64_t scells, double* dst, int dn0, int dn1,
                       int64_t dcells, int o0, int o1, int o2, int nc) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= dcells) return;
  const int i = (int)(t % dn0), j = (int)((t / dn0) % dn1), k = (int)(t / ((int64_t)dn0 * dn1));
  const int64_t s = ((int64_t)(k + o2) * sn1 + (j + o1)) * sn0 + (i + o0);
  for (int c = 0; c < nc; ++c) dst[c * dcells + t] = src[c * scells + s];
}


}  // namespace ilbm
