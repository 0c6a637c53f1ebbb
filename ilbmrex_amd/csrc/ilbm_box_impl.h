// Per-box entry points on caller-owned device memory in the reference's own
// layout (compiled as part of ilbm_level.hip).  A box's data is an AMReX
// Array4 of FP64: one contiguous array per component over the valid box
// grown by its ghost cells, x fastest (AmrCoreLBM.H:276, nghost = 3).  These
// are the per-MFIter loops of AdvancePhiAtLevel
// (Source/Core/AdvanceLBMAtLevel.cpp:196-386) and the BC functors of
// FillPatchSingleLevel (Source/BC/bc_fill.H:263-347), one kernel each, so a
// caller that keeps AMReX's MultiFabs (and its ParallelCopy / FillBoundary
// between boxes) can route the arithmetic of every box here.  The level-
// resident path (ilbm_level_*) fuses the same steps; both produce the
// reference's FP64 results.
#pragma once

#include "ilbm_amr_kernels.h"

namespace ilbm {

struct Fab {
  double* p;
  int lo[3], hi[3];
  int64_t js, ks, ns;    // element strides
  int64_t cells;         // cells of the box
  int nc;
  __host__ __device__ double& at(int i, int j, int k, int n) const {
    return p[(int64_t)(i - lo[0]) + (int64_t)(j - lo[1]) * js + (int64_t)(k - lo[2]) * ks + (int64_t)n * ns];
  }
  __host__ __device__ bool contains(int i, int j, int k) const {
    return i >= lo[0] && i <= hi[0] && j >= lo[1] && j <= hi[1] && k >= lo[2] && k <= hi[2];
  }
};

// Descriptor -> Fab.  Strides of 0 are the dense Array4 strides; explicit
// strides must keep x fastest and the components apart (no two cells or
// components of the box share an element).
static Fab to_fab(const ilbm_fab& f) {
  if (f.dtype != ILBM_FAB_F64) throw StatusError{ILBM_ERR_UNSUPPORTED, "per-box fabs are FP64 (dtype ILBM_FAB_F64)"};
  Fab F{};
  F.p = static_cast<double*>(f.p);
  for (int d = 0; d < 3; ++d) {
    F.lo[d] = f.lo[d];
    F.hi[d] = f.hi[d];
    if (f.hi[d] < f.lo[d]) throw StatusError{ILBM_ERR_INVALID, "fab box is empty"};
  }
  const int64_t nx = f.hi[0] - f.lo[0] + 1, ny = f.hi[1] - f.lo[1] + 1, nz = f.hi[2] - f.lo[2] + 1;
  F.js = f.jstride ? f.jstride : nx;
  F.ks = f.kstride ? f.kstride : F.js * ny;
  F.ns = f.nstride ? f.nstride : F.ks * nz;
  if (F.js < nx || F.ks < F.js * ny || (f.ncomp > 1 && F.ns < F.ks * nz))
    throw StatusError{ILBM_ERR_INVALID, "fab strides overlap cells or components"};
  F.cells = nx * ny * nz;
  F.nc = f.ncomp;
  return F;
}

struct BoxGeom {
  int dim;
  int n[3];                 // domain [0, n-1]
  int periodic[3];
  int lo_t[3], hi_t[3];     // BC types
  double lo_v[3][4], hi_v[3][4];
  int vlo[3], vhi[3];       // valid box
};

__device__ inline int nonper_out_cnt(const BoxGeom& G, const int iv[3]) {
  int c = 0;
  for (int d = 0; d < G.dim; ++d)
    if (!G.periodic[d] && (iv[d] < 0 || iv[d] > G.n[d] - 1)) ++c;
  return c;
}

// cell of the fab box at a linear index (x fastest)
__device__ inline void fab_cell(const Fab& F, int64_t t, int iv[3]) {
  const int nx = F.hi[0] - F.lo[0] + 1, ny = F.hi[1] - F.lo[1] + 1;
  iv[0] = F.lo[0] + (int)(t % nx);
  iv[1] = F.lo[1] + (int)((t / nx) % ny);
  iv[2] = F.lo[2] + (int)(t / ((int64_t)nx * ny));
}

// InitEquilibrium (AmrCoreLBM.cpp:2146-2161) over a box
template <class LAT>
__global__ void k_box_equilibrium(Fab m, Fab f, int lo0, int lo1, int lo2, int nx, int ny, int64_t cells) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cells) return;
  const int i = lo0 + (int)(t % nx), j = lo1 + (int)((t / nx) % ny), k = lo2 + (int)(t / ((int64_t)nx * ny));
  const double r = fmax(m.at(i, j, k, 0), 1.0e-12);
  const double u0 = m.at(i, j, k, 1), v0 = m.at(i, j, k, 2), w0 = m.at(i, j, k, 3);
  const double u2 = u0 * u0 + v0 * v0 + w0 * w0;
#pragma unroll
  for (int q = 0; q < LAT::Q; ++q) {
    const double cu = double(LAT::cx[q]) * u0 + double(LAT::cy[q]) * v0 + double(LAT::cz[q]) * w0;
    f.at(i, j, k, q) = LAT::w[q] * r * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * u2);
  }
}

// FillBoundary of one box: ghost cells whose periodic image lies in the
// valid box take its value (AMReX FillBoundary with the periodicity)
__global__ void k_box_fill_periodic(Fab F, BoxGeom G) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= F.cells) return;
  int iv[3];
  fab_cell(F, t, iv);
  bool valid = true, perout = false;
  int s[3] = {iv[0], iv[1], iv[2]};
  for (int d = 0; d < 3; ++d) {
    if (iv[d] >= G.vlo[d] && iv[d] <= G.vhi[d]) continue;
    valid = false;
    if (d >= G.dim) return;
    const bool out = iv[d] < 0 || iv[d] > G.n[d] - 1;
    if (out && !G.periodic[d]) return;
    if (out) {
      perout = true;
      int r = iv[d] % G.n[d];
      if (r < 0) r += G.n[d];
      s[d] = r;
    }
  }
  if (valid || !perout) return;
  for (int d = 0; d < 3; ++d)
    if (s[d] < G.vlo[d] || s[d] > G.vhi[d]) return;
  for (int n = 0; n < F.nc; ++n) F.at(iv[0], iv[1], iv[2], n) = F.at(s[0], s[1], s[2], n);
}

// One pass of the physical BC fill (cells outside the domain in `pass`
// non-periodic dimensions): FilccCell's foextrap copies (the last foextrap
// dimension wins), then mesoscopicBcFill (meso) or macroBcFill.
template <class LAT>
__global__ void k_box_fill_physical(Fab F, BoxGeom G, int pass, int meso) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= F.cells) return;
  int iv[3];
  fab_cell(F, t, iv);
  if (nonper_out_cnt(G, iv) != pass) return;
  const int nc = F.nc;
  for (int d = 0; d < G.dim; ++d) {
    int src[3] = {iv[0], iv[1], iv[2]};
    bool hit = false;
    if (iv[d] < 0 && G.lo_t[d] == BC_FOEXTRAP) { src[d] = 0; hit = true; }
    else if (iv[d] > G.n[d] - 1 && G.hi_t[d] == BC_FOEXTRAP) { src[d] = G.n[d] - 1; hit = true; }
    if (hit)
      for (int n = 0; n < nc; ++n) F.at(iv[0], iv[1], iv[2], n) = F.at(src[0], src[1], src[2], n);
  }
  // nearest interior cell, clamped in every dimension (bc_fill.H:15-41)
  int mir[3] = {iv[0], iv[1], iv[2]}, lh[3] = {0, 0, 0}, hh[3] = {0, 0, 0};
  for (int d = 0; d < G.dim; ++d) {
    if (iv[d] < 0) { mir[d] = 0; lh[d] = 1; }
    else if (iv[d] > G.n[d] - 1) { mir[d] = G.n[d] - 1; hh[d] = 1; }
  }
  auto first_face = [&](int type, int* bd, int* lo_side) {
    for (int d = 0; d < G.dim; ++d) {
      if (lh[d] && G.lo_t[d] == type) { *bd = d; *lo_side = 1; return true; }
      if (hh[d] && G.hi_t[d] == type) { *bd = d; *lo_side = 0; return true; }
    }
    return false;
  };
  int bd = -1, los = 1;
  if (!meso) {  // macroBcFill (bc_fill.H:263-301)
    if (first_face(BC_USER_2, &bd, &los)) {
      const double* b = los ? G.lo_v[bd] : G.hi_v[bd];
      for (int n = 0; n < nc; ++n) F.at(iv[0], iv[1], iv[2], n) = F.at(mir[0], mir[1], mir[2], n);
      if (nc > 0) F.at(iv[0], iv[1], iv[2], 0) = fmax(1.0e-20, 2.0 * b[0] - F.at(mir[0], mir[1], mir[2], 0));
      for (int n = 1; n < 4 && n < nc; ++n) F.at(iv[0], iv[1], iv[2], n) = 2.0 * b[n] - F.at(mir[0], mir[1], mir[2], n);
      return;
    }
    if (first_face(BC_USER_1, &bd, &los) || first_face(BC_EXT_DIR, &bd, &los)) {
      const double* b = los ? G.lo_v[bd] : G.hi_v[bd];
      for (int n = 0; n < nc; ++n) F.at(iv[0], iv[1], iv[2], n) = F.at(mir[0], mir[1], mir[2], n);
      for (int n = 0; n < 4 && n < nc; ++n) F.at(iv[0], iv[1], iv[2], n) = b[n];
    }
    return;
  }
  int kind = 0;
  if (first_face(BC_USER_2, &bd, &los)) kind = 2;
  else if (first_face(BC_USER_1, &bd, &los)) kind = 1;
  else if (first_face(BC_EXT_DIR, &bd, &los)) kind = 3;
  if (kind == 0) return;
  if (kind == 1) {  // copyMesoscopicFromMirror (:147-154)
#pragma unroll
    for (int q = 0; q < LAT::Q; ++q) F.at(iv[0], iv[1], iv[2], q) = F.at(mir[0], mir[1], mir[2], q);
    return;
  }
  double fnp[LAT::Q];
  double rhonp = 0.0, unp = 0.0, vnp = 0.0, wnp = 0.0;
#pragma unroll
  for (int q = 0; q < LAT::Q; ++q) {
    fnp[q] = F.at(mir[0], mir[1], mir[2], q);
    rhonp += fnp[q];
    unp += double(LAT::cx[q]) * fnp[q];
    vnp += double(LAT::cy[q]) * fnp[q];
    wnp += double(LAT::cz[q]) * fnp[q];
  }
  rhonp = fmax(rhonp, 1.0e-20);
  unp /= rhonp; vnp /= rhonp; wnp /= rhonp;
  const double* b = los ? G.lo_v[bd] : G.hi_v[bd];
  double rg = b[0], ug = b[1], vg = b[2], wg = b[3];
  if (kind == 2) {  // applyFaceMirroredMesoscopic (:209-261)
    rg = fmax(1.0e-20, 2.0 * b[0] - rhonp);
    ug = 2.0 * b[1] - unp; vg = 2.0 * b[2] - vnp; wg = 2.0 * b[3] - wnp;
  }
#pragma unroll
  for (int q = 0; q < LAT::Q; ++q) {  // applyExtDirMesoscopic (:159-204) / mirrored
    const double cx = LAT::cx[q], cy = LAT::cy[q], cz = LAT::cz[q];
    const double eu_np = cx * unp + cy * vnp + cz * wnp;
    const double uv_np = unp * unp + vnp * vnp + wnp * wnp;
    const double feqnp = LAT::w[q] * rhonp * (1.0 + 3.0 * eu_np + 4.5 * eu_np * eu_np - 1.5 * uv_np);
    const double eu_g = cx * ug + cy * vg + cz * wg;
    const double uv_g = ug * ug + vg * vg + wg * wg;
    const double feqg = LAT::w[q] * rg * (1.0 + 3.0 * eu_g + 4.5 * eu_g * eu_g - 1.5 * uv_g);
    F.at(iv[0], iv[1], iv[2], q) = feqg + (fnp[q] - feqnp);
  }
}

// collide / collide_forced over the valid box (AdvanceLBMAtLevel.cpp:196-240)
template <class LAT>
__global__ void k_box_collide(Fab f, Fab m, Fab F, double tau, BoxGeom G) {
  const int nx = G.vhi[0] - G.vlo[0] + 1, ny = G.vhi[1] - G.vlo[1] + 1;
  const int64_t cells = (int64_t)nx * ny * (G.vhi[2] - G.vlo[2] + 1);
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cells) return;
  const int i = G.vlo[0] + (int)(t % nx), j = G.vlo[1] + (int)((t / nx) % ny);
  const int k = G.vlo[2] + (int)(t / ((int64_t)nx * ny));
  const double rho = m.at(i, j, k, 0), u = m.at(i, j, k, 1), v = m.at(i, j, k, 2), w = m.at(i, j, k, 3);
  const double fx = F.p ? F.at(i, j, k, 0) : 0.0, fy = F.p ? F.at(i, j, k, 1) : 0.0;
  const double fz = F.p ? F.at(i, j, k, 2) : 0.0;
  const double uv = usq<LAT, double>(u, v, w);
  const bool forced = !(fx == 0.0 && fy == 0.0 && fz == 0.0);
  const double fac = 1.0 - 0.5 / tau;
#pragma unroll
  for (int q = 0; q < LAT::Q; ++q) {
    const double fe = feq_q<LAT, double>(q, rho, u, v, w, uv);
    double g = bgk<StF64>(f.at(i, j, k, q), fe, tau, 1.0 / tau);
    if (forced) g += guo_q<LAT, double>(q, fac, u, v, w, fx, fy, fz);
    f.at(i, j, k, q) = g;
  }
}

// halfway bounce-back into user_1 ghost cells (AdvanceLBMAtLevel.cpp:265-373)
template <class LAT>
__global__ void k_box_bounce_back(Fab f, BoxGeom G) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= f.cells) return;
  int iv[3];
  fab_cell(f, t, iv);
  bool valid = true;
  for (int d = 0; d < 3; ++d)
    if (iv[d] < G.vlo[d] || iv[d] > G.vhi[d]) valid = false;
  if (valid) return;
  bool has_bb = false, unsup = false;
  for (int d = 0; d < G.dim; ++d) {
    if (iv[d] < 0) { if (G.lo_t[d] == BC_USER_1) has_bb = true; else unsup = true; }
    else if (iv[d] > G.n[d] - 1) { if (G.hi_t[d] == BC_USER_1) has_bb = true; else unsup = true; }
  }
  if (!has_bb || unsup) return;
#pragma unroll
  for (int q = 0; q < LAT::Q; ++q) {
    const int it = iv[0] + LAT::cx[q], jt = iv[1] + LAT::cy[q], kt = iv[2] + LAT::cz[q];
    if (it < G.vlo[0] || it > G.vhi[0] || jt < G.vlo[1] || jt > G.vhi[1] || kt < G.vlo[2] || kt > G.vhi[2]) continue;
    f.at(iv[0], iv[1], iv[2], q) = f.at(it, jt, kt, opposite<LAT>(q));
  }
}

// stream (stream.H:9-22) and calculateMacroForcing
// (calculateMacroForcing.H:9-32) over the valid box
template <class LAT>
__global__ void k_box_stream_macro(Fab fin, Fab fout, Fab mout, Fab F, BoxGeom G) {
  const int nx = G.vhi[0] - G.vlo[0] + 1, ny = G.vhi[1] - G.vlo[1] + 1;
  const int64_t cells = (int64_t)nx * ny * (G.vhi[2] - G.vlo[2] + 1);
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cells) return;
  const int i = G.vlo[0] + (int)(t % nx), j = G.vlo[1] + (int)((t / nx) % ny);
  const int k = G.vlo[2] + (int)(t / ((int64_t)nx * ny));
  double r = 0.0, m1 = 0.0, m2 = 0.0, m3 = 0.0;
#pragma unroll
  for (int q = 0; q < LAT::Q; ++q) {
    const double fq = fin.at(i - LAT::cx[q], j - LAT::cy[q], k - LAT::cz[q], q);
    fout.at(i, j, k, q) = fq;
    r += fq;
    m1 += double(LAT::cx[q]) * fq;
    m2 += double(LAT::cy[q]) * fq;
    m3 += double(LAT::cz[q]) * fq;
  }
  const double fx = F.p ? F.at(i, j, k, 0) : 0.0, fy = F.p ? F.at(i, j, k, 1) : 0.0;
  const double fz = F.p ? F.at(i, j, k, 2) : 0.0;
  mout.at(i, j, k, 0) = r;
  mout.at(i, j, k, 1) = (m1 + 0.5 * fx) / r;
  mout.at(i, j, k, 2) = (m2 + 0.5 * fy) / r;
  mout.at(i, j, k, 3) = (m3 + 0.5 * fz) / r;
}

// vorticity and pressure (visPara.H:10-54) from a macro patch with one ghost
__global__ void k_box_derived(Fab mp, Fab mout, BoxGeom G, double dx, double dy, double dz) {
  const int nx = G.vhi[0] - G.vlo[0] + 1, ny = G.vhi[1] - G.vlo[1] + 1;
  const int64_t cells = (int64_t)nx * ny * (G.vhi[2] - G.vlo[2] + 1);
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cells) return;
  const int i = G.vlo[0] + (int)(t % nx), j = G.vlo[1] + (int)((t / nx) % ny);
  const int k = G.vlo[2] + (int)(t / ((int64_t)nx * ny));
  const double P = (mp.at(i, j, k, 0) - 1.0) / 3.0;
  const double dv_dx = (mp.at(i + 1, j, k, 2) - mp.at(i - 1, j, k, 2)) / (2.0 * dx);
  const double du_dy = (mp.at(i, j + 1, k, 1) - mp.at(i, j - 1, k, 1)) / (2.0 * dy);
  double vort;
  if (G.dim == 3) {
    const double du_dz = (mp.at(i, j, k + 1, 1) - mp.at(i, j, k - 1, 1)) / (2.0 * dz);
    const double dw_dx = (mp.at(i + 1, j, k, 3) - mp.at(i - 1, j, k, 3)) / (2.0 * dx);
    const double dv_dz = (mp.at(i, j, k + 1, 2) - mp.at(i, j, k - 1, 2)) / (2.0 * dz);
    const double dw_dy = (mp.at(i, j + 1, k, 3) - mp.at(i, j - 1, k, 3)) / (2.0 * dy);
    const double ox = dw_dy - dv_dz, oy = du_dz - dw_dx, oz = dv_dx - du_dy;
    vort = sqrt(ox * ox + oy * oy + oz * oz);
  } else {
    vort = dv_dx - du_dy;
  }
  mout.at(i, j, k, 4) = vort;
  if (mout.nc > 5) mout.at(i, j, k, 5) = P;
}

// per-block partial sums of one component over a box, fixed order
__global__ void __launch_bounds__(256) k_box_sum(Fab F, int comp, BoxGeom G, double* partial) {
  __shared__ double lds[4];
  const int nx = G.vhi[0] - G.vlo[0] + 1, ny = G.vhi[1] - G.vlo[1] + 1;
  const int64_t cells = (int64_t)nx * ny * (G.vhi[2] - G.vlo[2] + 1);
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (t < cells) {
    const int i = G.vlo[0] + (int)(t % nx), j = G.vlo[1] + (int)((t / nx) % ny);
    const int k = G.vlo[2] + (int)(t / ((int64_t)nx * ny));
    v = F.at(i, j, k, comp);
  }
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = ((lds[0] + lds[1]) + lds[2]) + lds[3];
}

static BoxGeom box_geom(const ilbm_case_desc& c, const int* vlo, const int* vhi) {
  BoxGeom G{};
  G.dim = c.dim;
  for (int d = 0; d < 3; ++d) {
    G.n[d] = (d < c.dim) ? c.n[d] : 1;
    G.periodic[d] = (d < c.dim) ? c.periodic[d] : 1;
    G.lo_t[d] = c.bc_lo[d];
    G.hi_t[d] = c.bc_hi[d];
    for (int k = 0; k < 4; ++k) { G.lo_v[d][k] = c.bc_lo_val[d][k]; G.hi_v[d][k] = c.bc_hi_val[d][k]; }
    G.vlo[d] = vlo ? vlo[d] : 0;
    G.vhi[d] = vhi ? vhi[d] : (G.n[d] - 1);
    if (d >= c.dim) { G.vlo[d] = 0; G.vhi[d] = 0; }
  }
  return G;
}

static int box_lattice(const ilbm_case_desc& c) {
  if (lattice_matches<D2Q9>(c)) return 0;
  if (lattice_matches<D3Q19>(c)) return 1;
  if (lattice_matches<D3Q27>(c)) return 2;
  throw StatusError{ILBM_ERR_UNSUPPORTED, "lattice table is not D2Q9/D3Q19/D3Q27 in the reference order"};
}

template <class F>
static void with_lattice(int lat, F&& f) {
  if (lat == 0) f(D2Q9{});
  else if (lat == 1) f(D3Q19{});
  else f(D3Q27{});
}

static dim3 grid_for(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

static int64_t box_cells(const BoxGeom& G) {
  return (int64_t)(G.vhi[0] - G.vlo[0] + 1) * (G.vhi[1] - G.vlo[1] + 1) * (G.vhi[2] - G.vlo[2] + 1);
}

static void check_fab(const ilbm_fab* f, int min_comp, const BoxGeom& G, int ghosts) {
  if (!f || !f->p) throw StatusError{ILBM_ERR_INVALID, "null fab"};
  if (f->ncomp < min_comp) throw StatusError{ILBM_ERR_INVALID, "fab has too few components"};
  for (int d = 0; d < 3; ++d) {
    const int g = (d < G.dim) ? ghosts : 0;
    if (f->lo[d] > G.vlo[d] - g || f->hi[d] < G.vhi[d] + g)
      throw StatusError{ILBM_ERR_INVALID, "fab does not cover the box and the ghost cells the operation reads"};
  }
}

// macro (rho, ux, uy, uz; [4][cells], no ghosts) into the interior of a fab
__global__ void k_fab_from_cells(Fab F, const double* m, int nx, int ny, int64_t cells) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cells) return;
  const int i = (int)(t % nx), j = (int)((t / nx) % ny), k = (int)(t / ((int64_t)nx * ny));
  for (int n = 0; n < 4; ++n) F.at(i, j, k, n) = m[n * cells + t];
}

__global__ void k_cells_from_fab(Fab F, int comp0, int ncomp, double* out, int nx, int ny, int64_t cells) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cells) return;
  const int i = (int)(t % nx), j = (int)((t / nx) % ny), k = (int)(t / ((int64_t)nx * ny));
  for (int n = 0; n < ncomp; ++n) out[n * cells + t] = F.at(i, j, k, comp0 + n);
}

// ParallelCopy of a box region between two fabs (same index space)
__global__ void k_box_copy(Fab S, Fab D, BoxGeom G, int ncomp) {
  const int nx = G.vhi[0] - G.vlo[0] + 1, ny = G.vhi[1] - G.vlo[1] + 1;
  const int64_t cells = (int64_t)nx * ny * (G.vhi[2] - G.vlo[2] + 1);
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cells) return;
  const int i = G.vlo[0] + (int)(t % nx), j = G.vlo[1] + (int)((t / nx) % ny);
  const int k = G.vlo[2] + (int)(t / ((int64_t)nx * ny));
  for (int n = 0; n < ncomp; ++n) D.at(i, j, k, n) = S.at(i, j, k, n);
}

__global__ void k_fab_setval(Fab F, double v) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= F.cells * F.nc) return;
  int iv[3];
  fab_cell(F, t % F.cells, iv);
  F.at(iv[0], iv[1], iv[2], (int)(t / F.cells)) = v;
}

}  // namespace ilbm

extern "C" {

int ilbm_box_copy(const ilbm_fab* src, ilbm_fab* dst, const int lo[3], const int hi[3], int ncomp, void* stream) {
  if (!src || !dst || !src->p || !dst->p || !lo || !hi || ncomp < 1 || ncomp > src->ncomp || ncomp > dst->ncomp)
    return ILBM_ERR_INVALID;
  return guarded([&] {
    using namespace ilbm;
    BoxGeom G{};
    G.dim = 3;
    for (int d = 0; d < 3; ++d) { G.vlo[d] = lo[d]; G.vhi[d] = hi[d]; }
    const Fab S = to_fab(*src), D = to_fab(*dst);
    for (int d = 0; d < 3; ++d)
      if (lo[d] > hi[d]) throw StatusError{ILBM_ERR_INVALID, "empty box"};
    if (!S.contains(lo[0], lo[1], lo[2]) || !S.contains(hi[0], hi[1], hi[2]) || !D.contains(lo[0], lo[1], lo[2]) ||
        !D.contains(hi[0], hi[1], hi[2]))
      throw StatusError{ILBM_ERR_INVALID, "box outside a fab"};
    hipLaunchKernelGGL(k_box_copy, grid_for(box_cells(G)), dim3(256), 0, (hipStream_t)stream, S, D, G, ncomp);
    HIPCHK(hipGetLastError());
  });
}

int ilbm_box_setval(ilbm_fab* f, double value, void* stream) {
  if (!f || !f->p || f->ncomp < 1) return ILBM_ERR_INVALID;
  return guarded([&] {
    using namespace ilbm;
    const Fab F = to_fab(*f);
    hipLaunchKernelGGL(k_fab_setval, grid_for(F.cells * F.nc), dim3(256), 0, (hipStream_t)stream, F, value);
    HIPCHK(hipGetLastError());
  });
}

int ilbm_device_alloc(int64_t bytes, void** out) {
  if (!out || bytes <= 0) return ILBM_ERR_INVALID;
  *out = nullptr;
  return guarded([&] { *out = ilbm::dev_alloc_bytes((size_t)bytes); });
}

int ilbm_device_free(void* p) {
  if (!p) return ILBM_ERR_INVALID;
  return guarded([&] { HIPCHK(hipFree(p)); });
}

int ilbm_copy_to_device(void* dst, const void* src, int64_t bytes) {
  if (!dst || !src || bytes < 0) return ILBM_ERR_INVALID;
  return guarded([&] { HIPCHK(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice)); });
}

int ilbm_copy_to_host(void* dst, const void* src, int64_t bytes) {
  if (!dst || !src || bytes < 0) return ILBM_ERR_INVALID;
  return guarded([&] { HIPCHK(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost)); });
}

int ilbm_device_synchronize(void) {
  return guarded([&] { HIPCHK(hipDeviceSynchronize()); });
}

// UpdateDerivedMacroFields of a resident level (AmrCoreLBM.cpp:150-168):
// FillPatchMacro of the exported macros with one ghost, then visPara.
int ilbm_level_get_derived(ilbm_level* h, double* host_out) {
  if (!h || !host_out) return ILBM_ERR_INVALID;
  ilbm::Level& L = h->L;
  if (!L.state_valid || !L.mout) return ILBM_ERR_STATE;
  if (L.cf) return ILBM_ERR_UNSUPPORTED;
  return guarded([&] {
    using namespace ilbm;
    const ilbm_case_desc& c = L.c;
    const int g[3] = {1, 1, c.dim == 3 ? 1 : 0};
    ilbm_fab pf{}, of{};
    for (int d = 0; d < 3; ++d) {
      pf.lo[d] = -g[d];
      pf.hi[d] = c.n[d] - 1 + g[d];
      of.lo[d] = 0;
      of.hi[d] = c.n[d] - 1;
    }
    pf.ncomp = 4;
    of.ncomp = 6;
    pf.dtype = of.dtype = ILBM_FAB_F64;
    Fab P = to_fab(pf), O = to_fab(of);
    double* buf = dev_alloc<double>(4 * P.ns + 6 * O.ns);
    P.p = buf;
    O.p = buf + 4 * P.ns;
    const BoxGeom G = box_geom(c, nullptr, nullptr);
    const int64_t cells = L.ncells;
    hipStream_t s = L.stream;
    hipLaunchKernelGGL(k_fab_from_cells, grid_for(cells), dim3(256), 0, s, P, (const double*)L.mout, c.n[0], c.n[1],
                       cells);
    hipLaunchKernelGGL(k_box_fill_periodic, grid_for(P.cells), dim3(256), 0, s, P, G);
    with_lattice(L.lat, [&](auto lat) {
      using LAT = decltype(lat);
      for (int pass = 1; pass <= c.dim; ++pass)
        hipLaunchKernelGGL(k_box_fill_physical<LAT>, grid_for(P.cells), dim3(256), 0, s, P, G, pass, 0);
    });
    hipLaunchKernelGGL(k_box_derived, grid_for(cells), dim3(256), 0, s, P, O, G, c.dx[0], c.dx[1], c.dx[2]);
    double* outd = dev_alloc<double>(2 * cells);
    hipLaunchKernelGGL(k_cells_from_fab, grid_for(cells), dim3(256), 0, s, O, 4, 2, outd, c.n[0], c.n[1], cells);
    const hipError_t e1 = hipMemcpyAsync(host_out, outd, sizeof(double) * 2 * cells, hipMemcpyDeviceToHost, s);
    const hipError_t e2 = hipStreamSynchronize(s);
    dev_free(outd);
    dev_free(buf);
    HIPCHK(e1);
    HIPCHK(e2);
  });
}

int ilbm_box_equilibrium(const ilbm_case_desc* d, const ilbm_fab* macro, ilbm_fab* f, const int lo[3],
                         const int hi[3], void* stream) {
  if (!d || !lo || !hi) return ILBM_ERR_INVALID;
  return guarded([&] {
    using namespace ilbm;
    const BoxGeom G = box_geom(*d, lo, hi);
    check_fab(macro, 4, G, 0);
    check_fab(f, d->q, G, 0);
    const int64_t cells = box_cells(G);
    with_lattice(box_lattice(*d), [&](auto L) {
      using LAT = decltype(L);
      hipLaunchKernelGGL(k_box_equilibrium<LAT>, grid_for(cells), dim3(256), 0, (hipStream_t)stream,
                         to_fab(*macro), to_fab(*f), G.vlo[0], G.vlo[1], G.vlo[2], G.vhi[0] - G.vlo[0] + 1,
                         G.vhi[1] - G.vlo[1] + 1, cells);
    });
    HIPCHK(hipGetLastError());
  });
}

int ilbm_box_fill_periodic(const ilbm_case_desc* d, ilbm_fab* fab, const int vlo[3], const int vhi[3],
                           void* stream) {
  if (!d || !fab || !fab->p) return ILBM_ERR_INVALID;
  return guarded([&] {
    using namespace ilbm;
    const BoxGeom G = box_geom(*d, vlo, vhi);
    const Fab F = to_fab(*fab);
    hipLaunchKernelGGL(k_box_fill_periodic, grid_for(F.cells), dim3(256), 0, (hipStream_t)stream, F, G);
    HIPCHK(hipGetLastError());
  });
}

int ilbm_box_fill_physical(const ilbm_case_desc* d, ilbm_fab* fab, int mesoscopic, void* stream) {
  if (!d || !fab || !fab->p) return ILBM_ERR_INVALID;
  return guarded([&] {
    using namespace ilbm;
    const BoxGeom G = box_geom(*d, nullptr, nullptr);
    const Fab F = to_fab(*fab);
    if (mesoscopic && F.nc < d->q) throw StatusError{ILBM_ERR_INVALID, "mesoscopic fab needs q components"};
    with_lattice(box_lattice(*d), [&](auto L) {
      using LAT = decltype(L);
      for (int pass = 1; pass <= d->dim; ++pass)
        hipLaunchKernelGGL(k_box_fill_physical<LAT>, grid_for(F.cells), dim3(256), 0, (hipStream_t)stream, F, G,
                           pass, mesoscopic ? 1 : 0);
    });
    HIPCHK(hipGetLastError());
  });
}

int ilbm_box_collide(const ilbm_case_desc* d, ilbm_fab* f, const ilbm_fab* macro, const ilbm_fab* force,
                     const int vlo[3], const int vhi[3], void* stream) {
  if (!d || !vlo || !vhi) return ILBM_ERR_INVALID;
  return guarded([&] {
    using namespace ilbm;
    const BoxGeom G = box_geom(*d, vlo, vhi);
    check_fab(f, d->q, G, 0);
    check_fab(macro, 4, G, 0);
    Fab Fo{};
    if (force) { check_fab(force, 3, G, 0); Fo = to_fab(*force); }
    const double tau = 3.0 * d->nu / d->dx[0] + 0.5;  // AmrCoreLBM.cpp:225-238
    with_lattice(box_lattice(*d), [&](auto L) {
      using LAT = decltype(L);
      hipLaunchKernelGGL(k_box_collide<LAT>, grid_for(box_cells(G)), dim3(256), 0, (hipStream_t)stream,
                         to_fab(*f), to_fab(*macro), Fo, tau, G);
    });
    HIPCHK(hipGetLastError());
  });
}

int ilbm_box_bounce_back(const ilbm_case_desc* d, ilbm_fab* f, const int vlo[3], const int vhi[3], void* stream) {
  if (!d || !vlo || !vhi) return ILBM_ERR_INVALID;
  return guarded([&] {
    using namespace ilbm;
    const BoxGeom G = box_geom(*d, vlo, vhi);
    check_fab(f, d->q, G, 1);
    const Fab F = to_fab(*f);
    with_lattice(box_lattice(*d), [&](auto L) {
      using LAT = decltype(L);
      hipLaunchKernelGGL(k_box_bounce_back<LAT>, grid_for(F.cells), dim3(256), 0, (hipStream_t)stream, F, G);
    });
    HIPCHK(hipGetLastError());
  });
}

int ilbm_box_stream_macro(const ilbm_case_desc* d, const ilbm_fab* f_in, ilbm_fab* f_out, ilbm_fab* macro_out,
                          const ilbm_fab* force, const int vlo[3], const int vhi[3], void* stream) {
  if (!d || !vlo || !vhi) return ILBM_ERR_INVALID;
  return guarded([&] {
    using namespace ilbm;
    const BoxGeom G = box_geom(*d, vlo, vhi);
    check_fab(f_in, d->q, G, 1);
    check_fab(f_out, d->q, G, 0);
    check_fab(macro_out, 4, G, 0);
    Fab Fo{};
    if (force) { check_fab(force, 3, G, 0); Fo = to_fab(*force); }
    with_lattice(box_lattice(*d), [&](auto L) {
      using LAT = decltype(L);
      hipLaunchKernelGGL(k_box_stream_macro<LAT>, grid_for(box_cells(G)), dim3(256), 0, (hipStream_t)stream,
                         to_fab(*f_in), to_fab(*f_out), to_fab(*macro_out), Fo, G);
    });
    HIPCHK(hipGetLastError());
  });
}

int ilbm_box_derived(const ilbm_case_desc* d, const ilbm_fab* macro_patch, ilbm_fab* macro_out, const int vlo[3],
                     const int vhi[3], void* stream) {
  if (!d || !vlo || !vhi) return ILBM_ERR_INVALID;
  return guarded([&] {
    using namespace ilbm;
    const BoxGeom G = box_geom(*d, vlo, vhi);
    check_fab(macro_patch, 4, G, 1);
    check_fab(macro_out, 5, G, 0);
    hipLaunchKernelGGL(k_box_derived, grid_for(box_cells(G)), dim3(256), 0, (hipStream_t)stream,
                       to_fab(*macro_patch), to_fab(*macro_out), G, d->dx[0], d->dx[1], d->dx[2]);
    HIPCHK(hipGetLastError());
  });
}

int ilbm_box_sum(const ilbm_fab* fab, int comp, const int vlo[3], const int vhi[3], double* host_out,
                 void* stream) {
  if (!fab || !fab->p || !vlo || !vhi || !host_out || comp < 0 || comp >= fab->ncomp) return ILBM_ERR_INVALID;
  return guarded([&] {
    using namespace ilbm;
    BoxGeom G{};
    G.dim = 3;
    for (int dd = 0; dd < 3; ++dd) { G.vlo[dd] = vlo[dd]; G.vhi[dd] = vhi[dd]; }
    const Fab F = to_fab(*fab);
    if (!F.contains(vlo[0], vlo[1], vlo[2]) || !F.contains(vhi[0], vhi[1], vhi[2]))
      throw StatusError{ILBM_ERR_INVALID, "box outside the fab"};
    const int64_t cells = box_cells(G);
    const int64_t nb = (cells + 255) / 256;
    double* part = dev_alloc<double>(nb + 1);
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_box_sum, dim3((unsigned)nb), dim3(256), 0, s, F, comp, G, part);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, s, (const double*)part, nb, part + nb);
    const hipError_t e1 = hipMemcpyAsync(host_out, part + nb, sizeof(double), hipMemcpyDeviceToHost, s);
    const hipError_t e2 = hipStreamSynchronize(s);
    dev_free(part);
    HIPCHK(e1);
    HIPCHK(e2);
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Coarse-fine transfer on caller-owned fabs, refinement ratio 2
// (Source/Core/AmrCoreLBM.cpp): FillPatchMesoscopic / FillPatchMacro of a
// fine box at lev > 0 (:1545-1627, :2180-2197), InterpFromCoarseLevel of
// MakeNewLevelFromCoarse (:2214-2237) and AverageDownTo (:1520-1526).
// ---------------------------------------------------------------------------
namespace ilbm {

// GetDataMacro (:2239-2269): the coarse data at `time` -- new, old, or
// FillPatchSingleLevel's a * old + b * new -- decided on the host
struct TimeMix {
  int mode;   // 0 old, 1 new, 2 a * old + b * new
  double a, b;
};

static TimeMix time_mix(double t_old, double t_new, double time) {
  const double sentinel = 1.0e50;   // a level not yet stepped keeps t_old = -1e200
  if (std::fabs(t_old) > sentinel || std::fabs(t_new) > sentinel) return {1, 0.0, 0.0};
  const double teps = std::fmax(1.0e-12, std::fabs(t_new - t_old) * 1.0e-3);
  if (time >= t_new - teps && time <= t_new + teps) return {1, 0.0, 0.0};
  if (time >= t_old - teps && time <= t_old + teps) return {0, 0.0, 0.0};
  return {2, (t_new - time) / (t_new - t_old), (time - t_old) / (t_new - t_old)};
}

__device__ inline double crse_at(const Fab& Co, const Fab& Cn, const TimeMix& tm, int i, int j, int k, int n) {
  if (tm.mode == 0) return Co.at(i, j, k, n);
  if (tm.mode == 1) return Cn.at(i, j, k, n);
  return tm.a * Co.at(i, j, k, n) + tm.b * Cn.at(i, j, k, n);
}

constexpr int kMaxInterpComp = 8;

// AMReX cell_cons_interp at fine cell (I, J, K): centred slopes limited per
// direction (minmod of the one-sided differences), one factor per direction
// shared by the NC components, offsets of +-1/4 coarse cell.  NC is a
// template parameter so that the per-component arrays stay in registers.
template <int DIM, int NC>
__device__ inline void cons_interp_fab(const Fab& Co, const Fab& Cn, const TimeMix& tm, int I, int J, int K,
                                       double* out) {
  const int ic = floordiv2(I), jc = floordiv2(J), kc = (DIM == 3) ? floordiv2(K) : K;
  double u0[NC], s[DIM][NC];
  double sf[DIM];
#pragma unroll
  for (int n = 0; n < NC; ++n) u0[n] = crse_at(Co, Cn, tm, ic, jc, kc, n);
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    sf[d] = 1.0;
    const int pi = ic + (d == 0), pj = jc + (d == 1), pk = kc + (d == 2);
    const int mi = ic - (d == 0), mj = jc - (d == 1), mk = kc - (d == 2);
#pragma unroll
    for (int n = 0; n < NC; ++n) {
      const double up = crse_at(Co, Cn, tm, pi, pj, pk, n);
      const double um = crse_at(Co, Cn, tm, mi, mj, mk, n);
      const double dc = 0.5 * (up - um);
      const double df = 2.0 * (up - u0[n]);
      const double db = 2.0 * (u0[n] - um);
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

// one cell of k_box_cf_fill with NC interpolated components
template <class LAT, int NC>
__device__ inline void box_cf_cell(const Fab& Co, const Fab& Cn, const TimeMix& tm, const Fab& M, const Fab& Fq,
                                   const int iv[3]) {
  double m[NC < 4 ? 4 : NC] = {};
  cons_interp_fab<LAT::D, NC>(Co, Cn, tm, iv[0], iv[1], iv[2], m);
  if (M.p)
#pragma unroll
    for (int n = 0; n < NC; ++n) M.at(iv[0], iv[1], iv[2], n) = m[n];
  if (!Fq.p) return;
  // FillPatchMesoscopic's equilibrium of the ghost cells (:1605-1620)
  const double r = fmax(m[0], 1.0e-12);
  const double u0 = m[1], v0 = m[2], w0 = m[3];
  const double u2 = u0 * u0 + v0 * v0 + w0 * w0;
#pragma unroll
  for (int q = 0; q < LAT::Q; ++q) {
    const double cu = double(LAT::cx[q]) * u0 + double(LAT::cy[q]) * v0 + double(LAT::cz[q]) * w0;
    Fq.at(iv[0], iv[1], iv[2], q) = LAT::w[q] * r * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * u2);
  }
}

// Interpolated macros on the cells of `dom` (the f fab's box when f is
// given, else the macro fab's) inside (ghosts = 0) or outside (ghosts = 1)
// the valid box; f = feq of them where f is given (ncomp >= 4 then).
template <class LAT>
__global__ void k_box_cf_fill(Fab Co, Fab Cn, TimeMix tm, Fab M, Fab Fq, Fab dom, int ncomp, BoxGeom G, int ghosts) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= dom.cells) return;
  int iv[3];
  fab_cell(dom, t, iv);
  bool in_valid = true;
  for (int d = 0; d < LAT::D; ++d)
    if (iv[d] < G.vlo[d] || iv[d] > G.vhi[d]) in_valid = false;
  if (in_valid == (ghosts != 0)) return;
  switch (ncomp) {
    case 1: box_cf_cell<LAT, 1>(Co, Cn, tm, M, Fq, iv); break;
    case 2: box_cf_cell<LAT, 2>(Co, Cn, tm, M, Fq, iv); break;
    case 3: box_cf_cell<LAT, 3>(Co, Cn, tm, M, Fq, iv); break;
    case 4: box_cf_cell<LAT, 4>(Co, Cn, tm, M, Fq, iv); break;
    case 5: box_cf_cell<LAT, 5>(Co, Cn, tm, M, Fq, iv); break;
    case 6: box_cf_cell<LAT, 6>(Co, Cn, tm, M, Fq, iv); break;
    case 7: box_cf_cell<LAT, 7>(Co, Cn, tm, M, Fq, iv); break;
    default: box_cf_cell<LAT, kMaxInterpComp>(Co, Cn, tm, M, Fq, iv); break;
  }
}

// amrex::average_down of f (Q comps) and macros (nmac comps) over the coarse
// cells [clo, clo + cn): crse = (sum of the children, k then j then i) / 2^d
template <class LAT>
__global__ void k_box_average_down(Fab Ff, Fab Mf, Fab Fc, Fab Mc, int nmac, int clo0, int clo1, int clo2, int cnx,
                                   int cny, int64_t ncov) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncov) return;
  const int i = clo0 + (int)(t % cnx), j = clo1 + (int)((t / cnx) % cny), k = clo2 + (int)(t / ((int64_t)cnx * cny));
  const int ii = 2 * i, jj = 2 * j, kk = (LAT::D == 3) ? 2 * k : k;
  const int nkr = (LAT::D == 3) ? 2 : 1;
  const double volfrac = (LAT::D == 3) ? 0.125 : 0.25;
  for (int q = 0; q < LAT::Q; ++q) {
    double c = 0.0;
    for (int kr = 0; kr < nkr; ++kr)
      for (int jr = 0; jr < 2; ++jr)
        for (int ir = 0; ir < 2; ++ir) c += Ff.at(ii + ir, jj + jr, kk + kr, q);
    Fc.at(i, j, k, q) = volfrac * c;
  }
  for (int n = 0; n < nmac; ++n) {
    double c = 0.0;
    for (int kr = 0; kr < nkr; ++kr)
      for (int jr = 0; jr < 2; ++jr)
        for (int ir = 0; ir < 2; ++ir) c += Mf.at(ii + ir, jj + jr, kk + kr, n);
    Mc.at(i, j, k, n) = volfrac * c;
  }
}

// the coarse cells a fine region's interpolation reads (its parents grown by 1)
static void check_coarse_cover(const ilbm_fab* crse, int ncomp, int dim, const int flo[3], const int fhi[3]) {
  if (!crse || !crse->p) throw StatusError{ILBM_ERR_INVALID, "null coarse fab"};
  if (crse->ncomp < ncomp) throw StatusError{ILBM_ERR_INVALID, "coarse fab has too few components"};
  for (int d = 0; d < dim; ++d) {
    const int lo = (flo[d] >= 0 ? flo[d] / 2 : -((1 - flo[d]) / 2)) - 1;
    const int hi = (fhi[d] >= 0 ? fhi[d] / 2 : -((1 - fhi[d]) / 2)) + 1;
    if (crse->lo[d] > lo || crse->hi[d] < hi)
      throw StatusError{ILBM_ERR_INVALID, "coarse fab does not cover the parents of the fine cells and their neighbours"};
  }
}

}  // namespace ilbm

extern "C" {

int ilbm_box_cf_fill(const ilbm_case_desc* d, const ilbm_fab* crse_old, const ilbm_fab* crse_new, double t_old,
                     double t_new, double time, ilbm_fab* f_fine, ilbm_fab* macro_fine, int ncomp,
                     const int valid_lo[3], const int valid_hi[3], void* stream) {
  if (!d || !valid_lo || !valid_hi || (!f_fine && !macro_fine)) return ILBM_ERR_INVALID;
  if (ncomp < (f_fine ? 4 : 1) || ncomp > ilbm::kMaxInterpComp) return ILBM_ERR_INVALID;
  return guarded([&] {
    using namespace ilbm;
    const BoxGeom G = box_geom(*d, valid_lo, valid_hi);
    const int dim = d->dim;
    const ilbm_fab* domf = f_fine ? f_fine : macro_fine;
    if (f_fine) check_fab(f_fine, d->q, G, 0);
    if (macro_fine) {
      check_fab(macro_fine, ncomp, G, 0);
      for (int e = 0; e < dim; ++e)
        if (macro_fine->lo[e] > domf->lo[e] || macro_fine->hi[e] < domf->hi[e])
          throw StatusError{ILBM_ERR_INVALID, "macro fab does not cover the f fab"};
    }
    // the fine ghost cells are coarse-fine ghosts: inside the level's domain
    for (int e = 0; e < dim; ++e)
      if (domf->lo[e] < 0 || domf->hi[e] > d->n[e] - 1)
        throw StatusError{ILBM_ERR_UNSUPPORTED, "fine fab reaches outside the level's domain (physical ghosts)"};
    check_coarse_cover(crse_old, ncomp, dim, domf->lo, domf->hi);
    check_coarse_cover(crse_new, ncomp, dim, domf->lo, domf->hi);
    const Fab Co = to_fab(*crse_old), Cn = to_fab(*crse_new), D = to_fab(*domf);
    const Fab M = macro_fine ? to_fab(*macro_fine) : Fab{};
    const Fab Fq = f_fine ? to_fab(*f_fine) : Fab{};
    const TimeMix tm = time_mix(t_old, t_new, time);
    with_lattice(box_lattice(*d), [&](auto L) {
      using LAT = decltype(L);
      hipLaunchKernelGGL(k_box_cf_fill<LAT>, grid_for(D.cells), dim3(256), 0, (hipStream_t)stream, Co, Cn, tm, M, Fq,
                         D, ncomp, G, 1);
    });
    HIPCHK(hipGetLastError());
  });
}

int ilbm_box_interp_from_coarse(const ilbm_case_desc* d, const ilbm_fab* crse, ilbm_fab* macro_fine, int ncomp,
                                const int lo[3], const int hi[3], void* stream) {
  if (!d || !crse || !macro_fine || !lo || !hi || ncomp < 1 || ncomp > ilbm::kMaxInterpComp)
    return ILBM_ERR_INVALID;
  return guarded([&] {
    using namespace ilbm;
    const BoxGeom G = box_geom(*d, lo, hi);
    const int dim = d->dim;
    check_fab(macro_fine, ncomp, G, 0);
    for (int e = 0; e < dim; ++e)
      if (lo[e] < 0 || hi[e] > d->n[e] - 1) throw StatusError{ILBM_ERR_INVALID, "box outside the level's domain"};
    check_coarse_cover(crse, ncomp, dim, G.vlo, G.vhi);
    const Fab C = to_fab(*crse), M = to_fab(*macro_fine);
    const TimeMix tm{1, 0.0, 0.0};
    with_lattice(box_lattice(*d), [&](auto L) {
      using LAT = decltype(L);
      hipLaunchKernelGGL(k_box_cf_fill<LAT>, grid_for(M.cells), dim3(256), 0, (hipStream_t)stream, C, C, tm, M, Fab{},
                         M, ncomp, G, 0);
    });
    HIPCHK(hipGetLastError());
  });
}

int ilbm_box_average_down(const ilbm_case_desc* d, const ilbm_fab* f_fine, const ilbm_fab* macro_fine,
                          const int fine_lo[3], const int fine_hi[3], ilbm_fab* f_crse, ilbm_fab* macro_crse, int nmac,
                          void* stream) {
  if (!d || !f_fine || !macro_fine || !f_crse || !macro_crse || !fine_lo || !fine_hi || nmac < 1)
    return ILBM_ERR_INVALID;
  return guarded([&] {
    using namespace ilbm;
    const int dim = d->dim;
    const BoxGeom GF = box_geom(*d, fine_lo, fine_hi);
    int clo[3] = {0, 0, 0}, chi[3] = {0, 0, 0};
    for (int e = 0; e < dim; ++e) {
      if (fine_lo[e] % 2 != 0 || fine_hi[e] % 2 != 1)
        throw StatusError{ILBM_ERR_INVALID, "fine box must cover whole coarse cells"};
      clo[e] = fine_lo[e] / 2;
      chi[e] = (fine_hi[e] - 1) / 2;
    }
    const BoxGeom GC = box_geom(*d, clo, chi);
    check_fab(f_fine, d->q, GF, 0);
    check_fab(macro_fine, nmac, GF, 0);
    check_fab(f_crse, d->q, GC, 0);
    check_fab(macro_crse, nmac, GC, 0);
    const int cnx = chi[0] - clo[0] + 1, cny = chi[1] - clo[1] + 1;
    const int64_t ncov = (int64_t)cnx * cny * (chi[2] - clo[2] + 1);
    with_lattice(box_lattice(*d), [&](auto L) {
      using LAT = decltype(L);
      hipLaunchKernelGGL(k_box_average_down<LAT>, grid_for(ncov), dim3(256), 0, (hipStream_t)stream, to_fab(*f_fine),
                         to_fab(*macro_fine), to_fab(*f_crse), to_fab(*macro_crse), nmac, clo[0], clo[1], clo[2], cnx,
                         cny, ncov);
    });
    HIPCHK(hipGetLastError());
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free box sum.  The block partials and the result live in a
// file-scope device array (one per GPU, held until the process ends), so
// the call allocates nothing; its kernels are ilbm_box_sum's, so the two
// return the same bits.  Calls that overlap on one GPU share the array and
// race; every call rewrites every slot it reads, so nothing needs clearing.
// ---------------------------------------------------------------------------
namespace ilbm {
constexpr int64_t kNoAllocPartials = 131072;                    // 1 MiB of partials
constexpr int64_t kNoAllocMaxCells = (kNoAllocPartials - 1) * 256;
}  // namespace ilbm

static __device__ __align__(256) double scratch[ilbm::kNoAllocPartials];

namespace ilbm {
constexpr int kNoAllocMaxDevices = 64;
static double* g_scratch_ptr[kNoAllocMaxDevices] = {};
static std::mutex g_scratch_mu;

// Device address of `scratch` on the current device, looked up once per
// device with that device current (each GPU holds its own copy).
static double* scratch_on_current_device() {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocMaxDevices) throw StatusError{ILBM_ERR_UNSUPPORTED, "device index"};
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw StatusError{ILBM_ERR_HIP, "hipGetSymbolAddress(scratch) returned null"};
    g_scratch_ptr[dev] = static_cast<double*>(ptr);
  }
  return g_scratch_ptr[dev];
}
}  // namespace ilbm

extern "C" {

int ilbm_box_sum_noalloc(const ilbm_fab* fab, int comp, const int vlo[3], const int vhi[3], double* host_out,
                         void* stream) {
  if (!fab || !fab->p || !vlo || !vhi || !host_out || comp < 0 || comp >= fab->ncomp) return ILBM_ERR_INVALID;
  return guarded([&] {
    using namespace ilbm;
    BoxGeom G{};
    G.dim = 3;
    for (int dd = 0; dd < 3; ++dd) {
      if (vlo[dd] > vhi[dd]) throw StatusError{ILBM_ERR_INVALID, "empty box"};
      G.vlo[dd] = vlo[dd];
      G.vhi[dd] = vhi[dd];
    }
    const Fab F = to_fab(*fab);
    if (!F.contains(vlo[0], vlo[1], vlo[2]) || !F.contains(vhi[0], vhi[1], vhi[2]))
      throw StatusError{ILBM_ERR_INVALID, "box outside the fab"};
    const int64_t cells = box_cells(G);
    if (cells > kNoAllocMaxCells)
      throw StatusError{ILBM_ERR_INVALID, "box larger than the allocation-free sum's limit of " +
                                              std::to_string(kNoAllocMaxCells) + " cells"};
    const int64_t nb = (cells + 255) / 256;
    double* part = scratch_on_current_device();
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_box_sum, dim3((unsigned)nb), dim3(256), 0, s, F, comp, G, part);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, s, (const double*)part, nb, part + nb);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(host_out, part + nb, sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  });
}

int ilbm_box_sum_noalloc_max_cells(int64_t* cells) {
  if (!cells) return ILBM_ERR_INVALID;
  *cells = ilbm::kNoAllocMaxCells;
  return ILBM_OK;
}

}  // extern "C"
