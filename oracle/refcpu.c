/*
 * refcpu -- CPU restatement of iLBMReX's per-timestep IB-LBM path.
 * TEST INFRASTRUCTURE ONLY (the parity oracle); see refcpu.h.
 *
 * Parity status of each part (the pins live in tests/test_oracle_*.py):
 *   feq / collide / Guo / stream / forced macros ... pinned: uniform-force
 *       recurrence and the reference's recorded 5.660377e-08 offset
 *       (Examples/ForceValidation), the square-duct E2/Linf/centre error
 *       table (Examples/SquareDuct_flow/Visualization.ipynb cell 4), mass
 *       conservation, feq fixed point.
 *   user_2 / periodic ghosts ............. pinned by the square duct.
 *   ext_dir / foextrap / user_1 ghosts ... pinned by closed forms on small
 *       cases (feq + non-equilibrium extrapolation identities, halfway
 *       bounce-back = resting wall invariants); corner ghosts across a
 *       periodic dimension follow the reference's clamp-all mirror and are
 *       "parity unpinned" beyond that reading.
 *   delta kernels / overlap_1d ........... pinned: Peskin partition of unity,
 *       zero first moment, sum of squares 3/8 (4-pt) and 1/2 (3-pt).
 *   marker geometry ...................... pinned: counts 101 / 1810 / 256
 *       and dv from the closed-form formulas.
 *   IVC LU ............................... pinned against numpy.linalg.solve.
 *   SumBoundary order across a periodic boundary, AMReX reduction order of
 *       MultiFab::sum ........................ parity unpinned (summation
 *       order only).
 */
#include "refcpu.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define RC_MAXQ 27

/* Planted mistakes for the tests that show each pin fails on a plausible
 * error (rc_set_mutation); 0 = the restatement as written. */
static int g_mut = 0;
enum {
  MUT_NONE = 0,
  MUT_EXTDIR_SIGN = 1,    /* feq - (f_np - feq_np) in applyExtDirMesoscopic */
  MUT_SPREAD_NO_DV = 2,   /* spreading without the marker volume dv */
  MUT_EXPORT_NO_DT = 3,   /* export F = rho * EFT without dt */
  MUT_INTERP_NO_DV = 4,   /* interpolation without the cell volume */
  MUT_CD_NO_HALF = 5,     /* Cd normalised by rho U^2 D */
  MUT_FORCE_SIGN = 6,     /* explicit forcing (U - U_target) / dt */
  MUT_MOMENT_SIGN = 7,    /* moments F x r instead of r x F */
  MUT_GUO_NO_HALF = 8     /* forced macros without the +F/2 */
};
void rc_set_mutation(int id) { g_mut = id; }

/* ------------------------------------------------------------------ */
/* Array4-like fab: one contiguous array per component, x fastest.      */
/* ------------------------------------------------------------------ */
typedef struct {
  double *p;
  int lo[3], hi[3];
  int64_t js, ks, ns;
  int nc;
} fab;

static fab fab_alloc(const int lo[3], const int hi[3], int nc) {
  fab F;
  for (int d = 0; d < 3; ++d) { F.lo[d] = lo[d]; F.hi[d] = hi[d]; }
  F.js = (int64_t)(hi[0] - lo[0] + 1);
  F.ks = F.js * (int64_t)(hi[1] - lo[1] + 1);
  F.ns = F.ks * (int64_t)(hi[2] - lo[2] + 1);
  F.nc = nc;
  F.p = (double *)calloc((size_t)(F.ns * nc), sizeof(double));
  return F;
}
static void fab_free(fab *F) { free(F->p); F->p = NULL; }

#define AT(F, i, j, k, n)                                                     \
  ((F).p[((int64_t)(i) - (F).lo[0]) + ((int64_t)(j) - (F).lo[1]) * (F).js +   \
         ((int64_t)(k) - (F).lo[2]) * (F).ks + (int64_t)(n) * (F).ns])

static void fab_setval(fab *F, double v) {
  for (int64_t t = 0; t < F->ns * F->nc; ++t) F->p[t] = v;
}

typedef struct {
  /* markers, index m = id - 1 */
  int nm;
  double *x, *y, *z;
  double *u, *v, *w;
  double *fx, *fy, *fz;
  double *mx, *my, *mz;
  double dv, wall_tol;
  double *phi, *theta;        /* cylinder / sphere parameterisation, 1-based */
  double *xref, *yref, *zref; /* box / user geometry, 1-based */
  double last_sum[3];
  /* IVC cache */
  int ivc_ready, ivc_n;
  double *ivc_lu;
  int *ivc_piv;
  double *xc, *yc, *zc;
  int have_cache;
} rc_ib;

struct rc_amr;

struct rc_level {
  rc_case c;
  int ng;
  int dom_lo[3], dom_hi[3];   /* the level's domain (geom[lev].Domain()) */
  int vlo[3], vhi[3];         /* the valid box (one box per level) */
  struct rc_amr *amr;         /* hierarchy this level belongs to, or NULL */
  int lev;
  fab f_new, f_old, m_new, m_old, forcing;
  double tau, dt;
  double t_new, t_old;
  int istep;
  rc_ib ib;
  /* ComputeIBForce state */
  int have_prev;
  double prev_body[3];
  double force_rec[40];
  int force_rec_n;
};

/* ------------------------------------------------------------------ */
/* Source/LBM/feq.H:9-29                                                 */
/* ------------------------------------------------------------------ */
double rc_feq(double wi, double dirx, double diry, double dirz, double rho,
              double u, double v, double w, int dim) {
  double eu, uv;
  if (dim == 3) {
    eu = dirx * u + diry * v + dirz * w;
    uv = u * u + v * v + w * w;
  } else {
    eu = dirx * u + diry * v;
    uv = u * u + v * v;
  }
  return wi * rho * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * uv);
}

/* Source/LBM/collide.H:9-25 then Source/LBM/collide_forced.H:10-51 */
void rc_collide_forced(double *f, const double *feq_i, int ndir, double tau,
                       const double *wi, const double *dirx,
                       const double *diry, const double *dirz, double rho,
                       double ux, double uy, double uz, double fx, double fy,
                       double fz, int dim) {
  (void)rho;
  for (int q = 0; q < ndir; ++q) f[q] = 1.0 * f[q] - (f[q] - feq_i[q]) / tau;
  if (fx == 0.0 && fy == 0.0 && fz == 0.0) return;
  const double cs2 = 1.0 / 3.0;
  const double fac = 1.0 - 0.5 / tau;
  for (int q = 0; q < ndir; ++q) {
    double cx = dirx[q], cy = diry[q], cz = dirz[q];
    double eiu = cx * ux + cy * uy + cz * uz;
    double termx = (cx - ux) / cs2 + (eiu / (cs2 * cs2)) * cx;
    double termy = (cy - uy) / cs2 + (eiu / (cs2 * cs2)) * cy;
    double termz = (dim == 3) ? (cz - uz) / cs2 + (eiu / (cs2 * cs2)) * cz : 0.0;
    double Si = wi[q] * fac * (termx * fx + termy * fy + termz * fz);
    f[q] += Si;
  }
}

/* ------------------------------------------------------------------ */
/* Ghost semantics (FillPatchSingleLevel + PhysBCFunct<GpuBndryFuncFab>) */
/* ------------------------------------------------------------------ */
static int out_lo(const rc_level *L, int d, int c) { return c < L->dom_lo[d]; }
static int out_hi(const rc_level *L, int d, int c) { return c > L->dom_hi[d]; }
static int wrapc(const rc_level *L, int d, int c) {
  int n = L->c.n[d];
  int r = (c - L->dom_lo[d]) % n;
  if (r < 0) r += n;
  return r + L->dom_lo[d];
}

/* Source/BC/bc_fill.H:15-41 */
static void nearest_interior(const rc_level *L, const int iv[3], int mir[3],
                             int lo_hit[3], int hi_hit[3]) {
  for (int d = 0; d < 3; ++d) { mir[d] = iv[d]; lo_hit[d] = hi_hit[d] = 0; }
  for (int d = 0; d < L->c.dim; ++d) {
    if (iv[d] < L->dom_lo[d]) { mir[d] = L->dom_lo[d]; lo_hit[d] = 1; }
    else if (iv[d] > L->dom_hi[d]) { mir[d] = L->dom_hi[d]; hi_hit[d] = 1; }
  }
}

/* Source/BC/bc_fill.H:46-67 */
static int first_face_of_type(const rc_level *L, const int lo_hit[3],
                              const int hi_hit[3], int type, int *bc_dim,
                              int *lo_side) {
  *bc_dim = -1; *lo_side = 1;
  for (int d = 0; d < L->c.dim; ++d) {
    if (lo_hit[d] && L->c.bc_lo[d] == type) { *bc_dim = d; *lo_side = 1; return 1; }
    if (hi_hit[d] && L->c.bc_hi[d] == type) { *bc_dim = d; *lo_side = 0; return 1; }
  }
  return 0;
}

static const double *bcv(const rc_level *L, int d, int lo_side) {
  return lo_side ? L->c.bc_lo_val[d] : L->c.bc_hi_val[d];
}

/* mesoscopicBcFill, Source/BC/bc_fill.H:303-347 (with 159-261) */
static void meso_user(const rc_level *L, fab *f, const int iv[3]) {
  const int ndir = L->c.q;
  int mir[3], lh[3], hh[3], bd, lo;
  nearest_interior(L, iv, mir, lh, hh);
  const double *cx = L->c.cx, *cy = L->c.cy, *cz = L->c.cz, *w = L->c.w;
  int kind = 0;
  if (first_face_of_type(L, lh, hh, RC_USER_2, &bd, &lo)) kind = 2;
  else if (first_face_of_type(L, lh, hh, RC_USER_1, &bd, &lo)) kind = 1;
  else if (first_face_of_type(L, lh, hh, RC_EXT_DIR, &bd, &lo)) kind = 3;
  if (kind == 0) return;
  if (kind == 1) { /* copyMesoscopicFromMirror :147-154 */
    for (int n = 0; n < ndir; ++n)
      AT(*f, iv[0], iv[1], iv[2], n) = AT(*f, mir[0], mir[1], mir[2], n);
    return;
  }
  double fnp[RC_MAXQ];
  double rhonp = 0.0, unp = 0.0, vnp = 0.0, wnp = 0.0;
  for (int n = 0; n < ndir; ++n) {
    fnp[n] = AT(*f, mir[0], mir[1], mir[2], n);
    rhonp += fnp[n];
    unp += cx[n] * fnp[n];
    vnp += cy[n] * fnp[n];
    wnp += cz[n] * fnp[n];
  }
  rhonp = fmax(rhonp, 1.0e-20);
  unp /= rhonp; vnp /= rhonp; wnp /= rhonp;
  const double *b = bcv(L, bd, lo);
  const double rho_b = b[0], u_b = b[1], v_b = b[2], w_b = b[3];
  if (kind == 3) { /* applyExtDirMesoscopic :159-204 */
    for (int n = 0; n < ndir; ++n) {
      double eu_np = cx[n] * unp + cy[n] * vnp + cz[n] * wnp;
      double uv_np = unp * unp + vnp * vnp + wnp * wnp;
      double feqnp = w[n] * rhonp * (1.0 + 3.0 * eu_np + 4.5 * eu_np * eu_np - 1.5 * uv_np);
      double eu_b = cx[n] * u_b + cy[n] * v_b + cz[n] * w_b;
      double uv_b = u_b * u_b + v_b * v_b + w_b * w_b;
      double feq = w[n] * rho_b * (1.0 + 3.0 * eu_b + 4.5 * eu_b * eu_b - 1.5 * uv_b);
      AT(*f, iv[0], iv[1], iv[2], n) = (g_mut == MUT_EXTDIR_SIGN) ? feq - (fnp[n] - feqnp) : feq + (fnp[n] - feqnp);
    }
  } else { /* applyFaceMirroredMesoscopic :209-261 */
    const double rho_g = fmax(1.0e-20, 2.0 * rho_b - rhonp);
    const double u_g = 2.0 * u_b - unp;
    const double v_g = 2.0 * v_b - vnp;
    const double w_g = 2.0 * w_b - wnp;
    for (int n = 0; n < ndir; ++n) {
      double eu_np = cx[n] * unp + cy[n] * vnp + cz[n] * wnp;
      double uv_np = unp * unp + vnp * vnp + wnp * wnp;
      double feqnp = w[n] * rhonp * (1.0 + 3.0 * eu_np + 4.5 * eu_np * eu_np - 1.5 * uv_np);
      double eu_g = cx[n] * u_g + cy[n] * v_g + cz[n] * w_g;
      double uv_g = u_g * u_g + v_g * v_g + w_g * w_g;
      double feqg = w[n] * rho_g * (1.0 + 3.0 * eu_g + 4.5 * eu_g * eu_g - 1.5 * uv_g);
      AT(*f, iv[0], iv[1], iv[2], n) = feqg + (fnp[n] - feqnp);
    }
  }
}

/* macroBcFill, Source/BC/bc_fill.H:263-301 */
static void macro_user(const rc_level *L, fab *m, const int iv[3], int numcomp) {
  int mir[3], lh[3], hh[3], bd, lo;
  nearest_interior(L, iv, mir, lh, hh);
  if (first_face_of_type(L, lh, hh, RC_USER_2, &bd, &lo)) {
    for (int n = 0; n < numcomp; ++n)
      AT(*m, iv[0], iv[1], iv[2], n) = AT(*m, mir[0], mir[1], mir[2], n);
    const double *b = bcv(L, bd, lo);
    if (numcomp > 0) AT(*m, iv[0], iv[1], iv[2], 0) = fmax(1.0e-20, 2.0 * b[0] - AT(*m, mir[0], mir[1], mir[2], 0));
    if (numcomp > 1) AT(*m, iv[0], iv[1], iv[2], 1) = 2.0 * b[1] - AT(*m, mir[0], mir[1], mir[2], 1);
    if (numcomp > 2) AT(*m, iv[0], iv[1], iv[2], 2) = 2.0 * b[2] - AT(*m, mir[0], mir[1], mir[2], 2);
    if (numcomp > 3) AT(*m, iv[0], iv[1], iv[2], 3) = 2.0 * b[3] - AT(*m, mir[0], mir[1], mir[2], 3);
    return;
  }
  if (first_face_of_type(L, lh, hh, RC_USER_1, &bd, &lo) ||
      first_face_of_type(L, lh, hh, RC_EXT_DIR, &bd, &lo)) {
    for (int n = 0; n < numcomp; ++n)
      AT(*m, iv[0], iv[1], iv[2], n) = AT(*m, mir[0], mir[1], mir[2], n);
    const double *b = bcv(L, bd, lo);
    for (int n = 0; n < 4 && n < numcomp; ++n) AT(*m, iv[0], iv[1], iv[2], n) = b[n];
  }
}

/* AMReX FilccCell: foextrap copies the nearest cell of that dimension; the
 * dimensions are processed in order, so the last foextrap dimension wins. */
static void filcc_cell(const rc_level *L, fab *F, const int iv[3], int ncomp) {
  for (int d = 0; d < L->c.dim; ++d) {
    int src[3] = {iv[0], iv[1], iv[2]};
    int hit = 0;
    if (out_lo(L, d, iv[d]) && L->c.bc_lo[d] == RC_FOEXTRAP) { src[d] = L->dom_lo[d]; hit = 1; }
    else if (out_hi(L, d, iv[d]) && L->c.bc_hi[d] == RC_FOEXTRAP) { src[d] = L->dom_hi[d]; hit = 1; }
    if (hit)
      for (int n = 0; n < ncomp; ++n)
        AT(*F, iv[0], iv[1], iv[2], n) = AT(*F, src[0], src[1], src[2], n);
  }
}

/* number of non-periodic dimensions in which iv lies outside the domain */
static int nonper_out(const rc_level *L, const int iv[3]) {
  int c = 0;
  for (int d = 0; d < L->c.dim; ++d)
    if (!L->c.periodic[d] && (out_lo(L, d, iv[d]) || out_hi(L, d, iv[d]))) ++c;
  return c;
}

/* FillPatchSingleLevel(dst, time, {src}, ...) with the level's physbc:
 * ParallelCopy (valid + periodic images), then the BC functor on faces,
 * edges and corners (cells outside the periodically grown domain in 1, 2,
 * then 3 dimensions).  is_meso selects mesoscopicBcFill vs macroBcFill. */
static void fill_patch(rc_level *L, fab *dst, const fab *src, int ncomp, int is_meso) {
  const int dim = L->c.dim;
  /* every cell writes only itself, from src (first loop) or from cells of
   * an earlier pass (the later loops): the loops run in parallel */
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = dst->lo[2]; k <= dst->hi[2]; ++k)
    for (int j = dst->lo[1]; j <= dst->hi[1]; ++j)
      for (int i = dst->lo[0]; i <= dst->hi[0]; ++i) {
        int iv[3] = {i, j, k};
        if (nonper_out(L, iv) > 0) continue;
        int s[3] = {i, j, k};
        for (int d = 0; d < dim; ++d) s[d] = wrapc(L, d, iv[d]);
        for (int n = 0; n < ncomp; ++n) AT(*dst, i, j, k, n) = AT(*src, s[0], s[1], s[2], n);
      }
  for (int pass = 1; pass <= dim; ++pass) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = dst->lo[2]; k <= dst->hi[2]; ++k)
      for (int j = dst->lo[1]; j <= dst->hi[1]; ++j)
        for (int i = dst->lo[0]; i <= dst->hi[0]; ++i) {
          int iv[3] = {i, j, k};
          if (nonper_out(L, iv) != pass) continue;
          filcc_cell(L, dst, iv, ncomp);
          if (is_meso) meso_user(L, dst, iv);
          else macro_user(L, dst, iv, ncomp);
        }
  }
}

/* FillBoundary: ghost cells whose periodic image is valid get that value. */
static void fill_boundary(rc_level *L, fab *F, int ncomp) {
  const int dim = L->c.dim;
  /* ghost cells take values of valid cells only: parallel */
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = F->lo[2]; k <= F->hi[2]; ++k)
    for (int j = F->lo[1]; j <= F->hi[1]; ++j)
      for (int i = F->lo[0]; i <= F->hi[0]; ++i) {
        int iv[3] = {i, j, k};
        int inside = 1, perout = 0;
        for (int d = 0; d < dim; ++d) {
          int o = out_lo(L, d, iv[d]) || out_hi(L, d, iv[d]);
          if (o && !L->c.periodic[d]) inside = 0;
          if (o && L->c.periodic[d]) perout = 1;
        }
        if (!inside || !perout) continue;
        int s[3] = {i, j, k};
        for (int d = 0; d < dim; ++d) s[d] = wrapc(L, d, iv[d]);
        for (int n = 0; n < ncomp; ++n) AT(*F, i, j, k, n) = AT(*F, s[0], s[1], s[2], n);
      }
}

/* SumBoundary: ghost contributions whose periodic image is valid are added
 * to that valid cell (ghost traversal order k, j, i: parity unpinned). */
static void sum_boundary(rc_level *L, fab *F, int ncomp) {
  const int dim = L->c.dim;
  for (int k = F->lo[2]; k <= F->hi[2]; ++k)
    for (int j = F->lo[1]; j <= F->hi[1]; ++j)
      for (int i = F->lo[0]; i <= F->hi[0]; ++i) {
        int iv[3] = {i, j, k};
        int inside = 1, perout = 0;
        for (int d = 0; d < dim; ++d) {
          int o = out_lo(L, d, iv[d]) || out_hi(L, d, iv[d]);
          if (o && !L->c.periodic[d]) inside = 0;
          if (o && L->c.periodic[d]) perout = 1;
        }
        if (!inside || !perout) continue;
        int s[3] = {i, j, k};
        for (int d = 0; d < dim; ++d) s[d] = wrapc(L, d, iv[d]);
        for (int n = 0; n < ncomp; ++n) AT(*F, s[0], s[1], s[2], n) += AT(*F, i, j, k, n);
      }
}

/* ------------------------------------------------------------------ */
/* Immersed boundary: Source/IBM/IBMarkerDF.cpp                          */
/* ------------------------------------------------------------------ */
double rc_delta(double xf, double xp, double h, int type) {
  const double rr = fabs(xf - xp) / h;
  double value = 0.0;
  if (type == 0) {
    if (rr < 1.0) value = 0.125 * (3.0 - 2.0 * rr + sqrt(1.0 + 4.0 * rr - 4.0 * rr * rr));
    else if (rr < 2.0) value = 0.125 * (5.0 - 2.0 * rr - sqrt(-7.0 + 12.0 * rr - 4.0 * rr * rr));
    else value = 0.0;
  } else {
    if (rr < 0.5) value = (1.0 + sqrt(1.0 - 3.0 * rr * rr)) / 3.0;
    else if (rr < 1.5) value = (5.0 - 3.0 * rr - sqrt(-2.0 + 6.0 * rr - 3.0 * rr * rr)) / 6.0;
    else value = 0.0;
  }
  return value / h;
}

int rc_delta_type_to_kernel(int t) {
  if (t == 4 || t == 0) return 0;
  if (t == 3 || t == 1) return 1;
  return 0;
}

double rc_overlap_1d(double xa, double xb, double h, int kernel_type,
                     double plo, int ilo, int ihi) {
  const int ia = (int)floor((xa - plo) / h);
  const int ib = (int)floor((xb - plo) / h);
  int lo = (ia - 2 > ib - 2) ? ia - 2 : ib - 2;
  int hi = (ia + 2 < ib + 2) ? ia + 2 : ib + 2;
  if (lo < ilo) lo = ilo;
  if (hi > ihi) hi = ihi;
  if (lo > hi) return 0.0;
  double sum = 0.0;
  for (int i = lo; i <= hi; ++i) {
    const double xc = plo + ((double)i + 0.5) * h;
    const double da = rc_delta(xc, xa, h, kernel_type);
    const double db = rc_delta(xc, xb, h, kernel_type);
    sum += da * db * h;
  }
  return sum;
}

int rc_lu_factorize(double *a, int *pivot, int n, double tiny) {
  for (int k = 0; k < n; ++k) pivot[k] = 0;
  for (int k = 0; k < n; ++k) {
    int p = k;
    double maxv = fabs(a[k * n + k]);
    for (int i = k + 1; i < n; ++i) {
      const double v = fabs(a[i * n + k]);
      if (v > maxv) { maxv = v; p = i; }
    }
    if (maxv <= tiny) return 0;
    pivot[k] = p;
    if (p != k)
      for (int j = 0; j < n; ++j) { double t = a[k * n + j]; a[k * n + j] = a[p * n + j]; a[p * n + j] = t; }
    const double akk = a[k * n + k];
    for (int i = k + 1; i < n; ++i) {
      a[i * n + k] /= akk;
      const double lik = a[i * n + k];
      for (int j = k + 1; j < n; ++j) a[i * n + j] -= lik * a[k * n + j];
    }
  }
  return 1;
}

void rc_lu_solve(const double *a, const int *pivot, const double *rhs,
                 double *x, int n) {
  for (int i = 0; i < n; ++i) x[i] = rhs[i];
  for (int k = 0; k < n; ++k) {
    const int p = pivot[k];
    if (p != k) { double t = x[k]; x[k] = x[p]; x[p] = t; }
  }
  for (int i = 1; i < n; ++i) {
    double sum = x[i];
    for (int j = 0; j < i; ++j) sum -= a[i * n + j] * x[j];
    x[i] = sum;
  }
  for (int i = n - 1; i >= 0; --i) {
    double sum = x[i];
    for (int j = i + 1; j < n; ++j) sum -= a[i * n + j] * x[j];
    x[i] = sum / a[i * n + i];
  }
}

static double rc_pi(void) { return 3.14159265358979323846; }

/* Pitching section outline, Examples/Pitching_airfoil/IBMUserDefinedGeometry.H:59-142 */
static int user_build(const rc_case *c, double h, double **xr, double **yr,
                      double **zr, double *dv) {
  const double chord = fmax(c->airfoil_chord, 1.0e-12);
  const double thick = fmax(c->airfoil_thickness, 1.0e-12);
  int n_half;
  if (c->n_marker > 0) { n_half = c->n_marker / 2; if (n_half < 32) n_half = 32; }
  else { n_half = (int)ceil(1.5 * chord / h); if (n_half < 48) n_half = 48; }
  const int cap = 2 * n_half + 2;
  *xr = (double *)calloc(cap, sizeof(double));
  *yr = (double *)calloc(cap, sizeof(double));
  *zr = (double *)calloc(cap, sizeof(double));
  int m = 1;
  const double pi = rc_pi();
#define THICK(xx)                                                               \
  (5.0 * thick * chord *                                                       \
   (0.2969 * sqrt(fmax(fmin(fmax((xx) / chord, 0.0), 1.0), 0.0)) -             \
    0.1260 * fmin(fmax((xx) / chord, 0.0), 1.0) -                              \
    0.3516 * fmin(fmax((xx) / chord, 0.0), 1.0) * fmin(fmax((xx) / chord, 0.0), 1.0) + \
    0.2843 * fmin(fmax((xx) / chord, 0.0), 1.0) * fmin(fmax((xx) / chord, 0.0), 1.0) * fmin(fmax((xx) / chord, 0.0), 1.0) - \
    0.1036 * fmin(fmax((xx) / chord, 0.0), 1.0) * fmin(fmax((xx) / chord, 0.0), 1.0) * fmin(fmax((xx) / chord, 0.0), 1.0) * fmin(fmax((xx) / chord, 0.0), 1.0)))
  for (int i = 0; i <= n_half; ++i) {
    const double beta = pi * (double)i / (double)n_half;
    const double x = 0.5 * chord * (1.0 - cos(beta));
    (*xr)[m] = x; (*yr)[m] = THICK(x); (*zr)[m] = 0.0; ++m;
  }
  for (int i = n_half - 1; i >= 1; --i) {
    const double beta = pi * (double)i / (double)n_half;
    const double x = 0.5 * chord * (1.0 - cos(beta));
    (*xr)[m] = x; (*yr)[m] = -THICK(x); (*zr)[m] = 0.0; ++m;
  }
#undef THICK
  const int ml = m - 1;
  double perim = 0.0;
  for (int i = 1; i <= ml; ++i) {
    const int inext = (i == ml) ? 1 : (i + 1);
    const double dxs = (*xr)[inext] - (*xr)[i];
    const double dys = (*yr)[inext] - (*yr)[i];
    perim += sqrt(dxs * dxs + dys * dys);
  }
  *dv = (perim / (double)ml) * h;
  return ml;
}

/* marker_state, IBMUserDefinedGeometry.H:144-175 (pivot made a parameter) */
static void user_state(const rc_case *c, double time, double xr, double yr,
                       double zr, double *x, double *y, double *z, double *ux,
                       double *uy, double *uz) {
  const double twopi = 2.0 * rc_pi();
  const double phase = twopi * c->pitch_frequency * time + c->pitch_phase;
  const double angle = c->pitch_mean + c->pitch_amplitude * sin(phase);
  const double omega = twopi * c->pitch_frequency * c->pitch_amplitude * cos(phase);
  const double px = c->x0 + c->pivot_x, py = c->y0;
  const double xrr = xr - c->pivot_x;
  const double cs = cos(angle), sn = sin(angle);
  *x = px + cs * xrr - sn * yr;
  *y = py + sn * xrr + cs * yr;
  *z = c->z0 + zr;
  *ux = -omega * (*y - py);
  *uy = omega * (*x - px);
  *uz = 0.0;
}

/* build_markers, IBMarkerDF.cpp:201-497 */
static void ib_build(rc_level *L) {
  rc_ib *ib = &L->ib;
  const rc_case *c = &L->c;
  const double h = c->dx[0];
  const double pi = rc_pi();
  int Ml = c->n_marker;
  double dv = 0.0;
  ib->phi = ib->theta = ib->xref = ib->yref = ib->zref = NULL;
  ib->wall_tol = 0.0;
  if (c->geometry_type == RC_GEOM_BOX) {
    double xlo = c->box_lo[0], xhi = c->box_hi[0], ylo = c->box_lo[1], yhi = c->box_hi[1];
    const double phx = c->plo[0] + c->dx[0] * c->n[0], phy = c->plo[1] + c->dx[1] * c->n[1];
    if (xhi <= xlo) { xlo = c->plo[0]; xhi = phx; }
    if (yhi <= ylo) { ylo = c->plo[1]; yhi = phy; }
    const double epsx = fmax(1.0e-6 * c->dx[0], 1.0e-12), epsy = fmax(1.0e-6 * c->dx[1], 1.0e-12);
    xlo = fmax(xlo, c->plo[0] + epsx); xhi = fmin(xhi, phx - epsx);
    ylo = fmax(ylo, c->plo[1] + epsy); yhi = fmin(yhi, phy - epsy);
    L->c.box_lo[0] = xlo; L->c.box_hi[0] = xhi; L->c.box_lo[1] = ylo; L->c.box_hi[1] = yhi;
    double ds = (c->box_ds > 0.0) ? c->box_ds : h;
    ds = fmax(ds, 1.0e-12);
    int nx = (int)round((xhi - xlo) / ds) + 1; if (nx < 2) nx = 2;
    int ny = (int)round((yhi - ylo) / ds) + 1; if (ny < 2) ny = 2;
    const double dsx = (xhi - xlo) / (double)(nx - 1), dsy = (yhi - ylo) / (double)(ny - 1);
    const int cap = 2 * nx + 2 * ny + 2;
    ib->xref = calloc(cap, sizeof(double)); ib->yref = calloc(cap, sizeof(double)); ib->zref = calloc(cap, sizeof(double));
    int m = 1;
    if (c->box_lid_only) {
      for (int i = 0; i < nx; ++i) { ib->xref[m] = xlo + (double)i * dsx; ib->yref[m] = yhi; ++m; }
    } else {
      for (int i = 0; i < nx; ++i) { ib->xref[m] = xlo + (double)i * dsx; ib->yref[m] = ylo; ++m; }
      for (int i = 0; i < nx; ++i) { ib->xref[m] = xlo + (double)i * dsx; ib->yref[m] = yhi; ++m; }
      for (int j = 1; j < ny - 1; ++j) {
        const double y = ylo + (double)j * dsy;
        ib->xref[m] = xlo; ib->yref[m] = y; ++m;
        ib->xref[m] = xhi; ib->yref[m] = y; ++m;
      }
    }
    Ml = m - 1;
    const double perim = c->box_lid_only ? (xhi - xlo) : 2.0 * ((xhi - xlo) + (yhi - ylo));
    dv = (perim / (double)Ml) * h;
    ib->wall_tol = (c->box_wall_tol > 0.0) ? c->box_wall_tol : fmax(0.25 * fmin(dsx, dsy), 1.0e-12);
  } else if (c->geometry_type == RC_GEOM_USER) {
    Ml = user_build(c, h, &ib->xref, &ib->yref, &ib->zref, &dv);
  } else if (c->dim == 2) {
    if (Ml <= 0) { Ml = (int)ceil(2.0 * pi * c->R / h); if (Ml < 16) Ml = 16; }
    dv = (2.0 * pi * c->R * h) / (double)Ml;
    ib->phi = calloc(Ml + 1, sizeof(double));
    ib->theta = calloc(Ml + 1, sizeof(double));
    ib->theta[0] = 0.5 * pi;
    for (int i = 1; i <= Ml; ++i) {
      ib->phi[i] = 2.0 * pi * (double)(i - 1) / (double)Ml;
      ib->theta[i] = 0.5 * pi;
    }
  } else {
    const double rd = c->rd;
    const double r1 = c->R - (rd - 0.5) * h, r2 = c->R - (rd + 0.5) * h;
    const double a1 = r1 * r1 * r1, a2 = r2 * r2 * r2;
    if (Ml <= 0) {
      const double denom = 3.0 * h * h * h / 4.0 / pi;
      Ml = (int)((a1 - a2) / denom);
      if (Ml < 32) Ml = 32;
    }
    dv = (a1 - a2) / (3.0 * (double)Ml / 4.0 / pi);
    ib->phi = calloc(Ml + 1, sizeof(double));
    ib->theta = calloc(Ml + 1, sizeof(double));
    double phi = 0.0;
    for (int mi = 0; mi < Ml; ++mi) {
      const double Hk = -1.0 + 2.0 * (double)mi / ((double)Ml - 1.0);
      const double theta = acos(Hk);
      if (mi == 0 || mi == Ml - 1) phi = 0.0;
      else phi = fmod(phi + 3.809 / sqrt((double)Ml) / sqrt(1.0 - Hk * Hk), 2.0 * pi);
      ib->phi[mi + 1] = phi;
      ib->theta[mi + 1] = theta;
    }
  }
  ib->nm = Ml;
  ib->dv = dv;
  const size_t n = (size_t)Ml;
  ib->x = calloc(n, 8); ib->y = calloc(n, 8); ib->z = calloc(n, 8);
  ib->u = calloc(n, 8); ib->v = calloc(n, 8); ib->w = calloc(n, 8);
  ib->fx = calloc(n, 8); ib->fy = calloc(n, 8); ib->fz = calloc(n, 8);
  ib->mx = calloc(n, 8); ib->my = calloc(n, 8); ib->mz = calloc(n, 8);
  ib->ivc_lu = NULL; ib->ivc_piv = NULL; ib->xc = ib->yc = ib->zc = NULL;
  ib->ivc_ready = 0; ib->ivc_n = 0; ib->have_cache = 0;
}

/* update_lagrangian_marker, IBMarkerDF.cpp:499-609 */
static void ib_update_markers(rc_level *L, double time) {
  rc_ib *ib = &L->ib;
  const rc_case *c = &L->c;
  const double z0 = (c->dim == 3) ? c->z0 : 0.0;
  for (int m = 0; m < ib->nm; ++m) {
    const int idx = m + 1;
    if (c->geometry_type == RC_GEOM_BOX) {
      ib->x[m] = ib->xref[idx]; ib->y[m] = ib->yref[idx];
      ib->z[m] = (c->dim == 3) ? ib->zref[idx] : 0.0;
      ib->mx[m] = ib->my[m] = ib->mz[m] = 0.0;
    } else if (c->geometry_type == RC_GEOM_USER) {
      double px, py, pz, ux, uy, uz;
      user_state(c, time, ib->xref[idx], ib->yref[idx], ib->zref[idx], &px, &py, &pz, &ux, &uy, &uz);
      ib->x[m] = px; ib->y[m] = py; ib->z[m] = (c->dim == 3) ? pz : 0.0;
      ib->mx[m] = ux; ib->my[m] = uy; ib->mz[m] = uz;
    } else {
      ib->x[m] = c->x0 + c->R * sin(ib->theta[idx]) * cos(ib->phi[idx]);
      ib->y[m] = c->y0 + c->R * sin(ib->theta[idx]) * sin(ib->phi[idx]);
      ib->z[m] = (c->dim == 3) ? z0 + c->R * cos(ib->theta[idx]) : 0.0;
      ib->mx[m] = ib->my[m] = ib->mz[m] = 0.0;
    }
    ib->u[m] = ib->v[m] = ib->w[m] = 0.0;
    ib->fx[m] = ib->fy[m] = ib->fz[m] = 0.0;
  }
}

static void rigid_target(const rc_level *L, double x, double y, double z,
                         double *ux, double *uy, double *uz) {
  const rc_case *c = &L->c;
  if (c->geometry_type == RC_GEOM_BOX) {
    *ux = *uy = *uz = 0.0;
    if (fabs(y - c->box_hi[1]) <= L->ib.wall_tol) {
      *ux = c->box_lid_u[0]; *uy = c->box_lid_u[1]; *uz = c->box_lid_u[2];
    }
    return;
  }
  const double z0 = (c->dim == 3) ? c->z0 : 0.0;
  const double rx = x - c->x0, ry = y - c->y0, rz = z - z0;
  const double urx = c->omy * rz - c->omz * ry;
  const double ury = c->omz * rx - c->omx * rz;
  const double urz = c->omx * ry - c->omy * rx;
  *ux = c->ubx + urx; *uy = c->uby + ury; *uz = c->ubz + urz;
}

static void marker_base(const rc_level *L, int m, int b[3]) {
  const rc_case *c = &L->c;
  const double pos[3] = {L->ib.x[m], L->ib.y[m], L->ib.z[m]};
  for (int d = 0; d < 3; ++d) b[d] = 0;
  for (int d = 0; d < c->dim; ++d) b[d] = (int)floor((pos[d] - c->plo[d]) / c->dx[d]);
}

/* velocity_interpolation_cic + velocity_interpolation, :611-739 */
static void ib_interp(rc_level *L, fab *E) {
  fill_boundary(L, E, 3);
  rc_ib *ib = &L->ib;
  const rc_case *c = &L->c;
  const int kt = rc_delta_type_to_kernel(c->delta_type);
  const int renorm = (c->renormalize_delta != 0) && (c->coupling_method != RC_COUPLING_IVC);
  const double d = (g_mut == MUT_INTERP_NO_DV) ? 1.0 : (c->dim == 3) ? c->dx[0] * c->dx[1] * c->dx[2] : c->dx[0] * c->dx[1];
  for (int m = 0; m < ib->nm; ++m) {
    int b[3];
    marker_base(L, m, b);
    double u = 0.0, v = 0.0, w = 0.0, part = 0.0;
    for (int ii = -2; ii < 3; ++ii) {
      const double xi = c->plo[0] + ((double)(b[0] + ii) + 0.5) * c->dx[0];
      const double tU = rc_delta(xi, ib->x[m], c->dx[0], kt);
      for (int jj = -2; jj < 3; ++jj) {
        const double yi = c->plo[1] + ((double)(b[1] + jj) + 0.5) * c->dx[1];
        const double tV = rc_delta(yi, ib->y[m], c->dx[1], kt);
        if (c->dim == 3) {
          for (int kk = -2; kk < 3; ++kk) {
            const double zi = c->plo[2] + ((double)(b[2] + kk) + 0.5) * c->dx[2];
            const double tW = rc_delta(zi, ib->z[m], c->dx[2], kt);
            const int ip = b[0] + ii, jp = b[1] + jj, kp = b[2] + kk;
            if (ip < E->lo[0] || ip > E->hi[0] || jp < E->lo[1] || jp > E->hi[1] || kp < E->lo[2] || kp > E->hi[2]) continue;
            const double dval = tU * tV * tW;
            part += dval * d;
            u += dval * AT(*E, ip, jp, kp, 0) * d;
            v += dval * AT(*E, ip, jp, kp, 1) * d;
            w += dval * AT(*E, ip, jp, kp, 2) * d;
          }
        } else {
          const int ip = b[0] + ii, jp = b[1] + jj;
          if (ip < E->lo[0] || ip > E->hi[0] || jp < E->lo[1] || jp > E->hi[1]) continue;
          const double dval = tU * tV;
          part += dval * d;
          u += dval * AT(*E, ip, jp, 0, 0) * d;
          v += dval * AT(*E, ip, jp, 0, 1) * d;
          w += 0.0;
        }
      }
    }
    if (renorm && part > 0.0) { u /= part; v /= part; w /= part; }
    ib->u[m] += u; ib->v[m] += v; ib->w[m] += w;
  }
}

/* compute_lagrangian_force_explicit, :741-873 */
static void ib_force_explicit(rc_level *L, double dt, double time) {
  (void)time;
  rc_ib *ib = &L->ib;
  const rc_case *c = &L->c;
  const double relax = fmax(0.0, fmin(1.0, c->mdf_relax));
  const int diag = (c->coupling_method == RC_COUPLING_EXPLICIT_DIAG);
  const double diag_eps = fmax(c->explicit_diag_eps, 1.0e-30);
  const int kt = rc_delta_type_to_kernel(c->delta_type);
  const double dV = (c->dim == 3) ? c->dx[0] * c->dx[1] * c->dx[2] : c->dx[0] * c->dx[1];
  for (int m = 0; m < ib->nm; ++m) {
    double Ux, Uy, Uz;
    if (c->geometry_type == RC_GEOM_USER) { Ux = ib->mx[m]; Uy = ib->my[m]; Uz = ib->mz[m]; }
    else rigid_target(L, ib->x[m], ib->y[m], ib->z[m], &Ux, &Uy, &Uz);
    double denom = dt;
    if (diag) {
      int b[3];
      marker_base(L, m, b);
      double s2 = 0.0;
      for (int ii = -2; ii < 3; ++ii) {
        const int ip = b[0] + ii;
        if (ip < L->dom_lo[0] || ip > L->dom_hi[0]) continue;
        const double xi = c->plo[0] + ((double)ip + 0.5) * c->dx[0];
        const double tU = rc_delta(xi, ib->x[m], c->dx[0], kt);
        for (int jj = -2; jj < 3; ++jj) {
          const int jp = b[1] + jj;
          if (jp < L->dom_lo[1] || jp > L->dom_hi[1]) continue;
          const double yi = c->plo[1] + ((double)jp + 0.5) * c->dx[1];
          const double tV = rc_delta(yi, ib->y[m], c->dx[1], kt);
          if (c->dim == 3) {
            for (int kk = -2; kk < 3; ++kk) {
              const int kp = b[2] + kk;
              if (kp < L->dom_lo[2] || kp > L->dom_hi[2]) continue;
              const double zi = c->plo[2] + ((double)kp + 0.5) * c->dx[2];
              const double tW = rc_delta(zi, ib->z[m], c->dx[2], kt);
              const double dval = tU * tV * tW;
              s2 += dval * dval * dV;
            }
          } else {
            const double dval = tU * tV;
            s2 += dval * dval * dV;
          }
        }
      }
      const double mkk = ib->dv * s2;
      denom = dt * fmax(mkk, diag_eps);
    }
    const double sg = (g_mut == MUT_FORCE_SIGN) ? -1.0 : 1.0;
    ib->fx[m] += relax * (sg * (Ux - ib->u[m])) / denom;
    ib->fy[m] += relax * (sg * (Uy - ib->v[m])) / denom;
    ib->fz[m] += relax * (sg * (Uz - ib->w[m])) / denom;
  }
}

/* build_ivc_operator :1081-1142 */
static void ib_build_ivc(rc_level *L, const double *xg, const double *yg, const double *zg) {
  rc_ib *ib = &L->ib;
  const rc_case *c = &L->c;
  const int n = ib->nm;
  const int kt = rc_delta_type_to_kernel(c->delta_type);
  ib->ivc_ready = 0;
  ib->ivc_n = n;
  free(ib->ivc_lu); free(ib->ivc_piv);
  ib->ivc_lu = calloc((size_t)n * n, sizeof(double));
  ib->ivc_piv = calloc((size_t)n, sizeof(int));
  for (int i = 0; i < n; ++i)
    for (int j = i; j < n; ++j) {
      const double sx = rc_overlap_1d(xg[i], xg[j], c->dx[0], kt, c->plo[0], L->dom_lo[0], L->dom_hi[0]);
      const double sy = rc_overlap_1d(yg[i], yg[j], c->dx[1], kt, c->plo[1], L->dom_lo[1], L->dom_hi[1]);
      const double sz = (c->dim == 3) ? rc_overlap_1d(zg[i], zg[j], c->dx[2], kt, c->plo[2], L->dom_lo[2], L->dom_hi[2]) : 1.0;
      double val = ib->dv * sx * sy * sz;
      if (i == j) val += c->ivc_diag_reg;
      ib->ivc_lu[(size_t)i * n + j] = val;
      ib->ivc_lu[(size_t)j * n + i] = val;
    }
  const double tiny = fmax(c->ivc_diag_reg * 1.0e-4, 1.0e-30);
  ib->ivc_ready = rc_lu_factorize(ib->ivc_lu, ib->ivc_piv, n, tiny);
  free(ib->xc); free(ib->yc); free(ib->zc);
  ib->xc = malloc((size_t)n * 8); ib->yc = malloc((size_t)n * 8); ib->zc = malloc((size_t)n * 8);
  memcpy(ib->xc, xg, (size_t)n * 8); memcpy(ib->yc, yg, (size_t)n * 8); memcpy(ib->zc, zg, (size_t)n * 8);
  ib->have_cache = 1;
}

/* compute_lagrangian_force_ivc :1144-1328 (one rank: gather = local state) */
static void ib_force_ivc(rc_level *L, double dt, double time) {
  rc_ib *ib = &L->ib;
  const rc_case *c = &L->c;
  const int n = ib->nm;
  if (n <= 0) return;
  /* gather_marker_state recomputes positions from the parameterisation;
   * the values equal the particle positions set by update_lagrangian_marker */
  int changed = !ib->have_cache || ib->ivc_n != n;
  if (!changed)
    for (int i = 0; i < n; ++i)
      if (fabs(ib->x[i] - ib->xc[i]) > 1.0e-12 || fabs(ib->y[i] - ib->yc[i]) > 1.0e-12 ||
          fabs(ib->z[i] - ib->zc[i]) > 1.0e-12) { changed = 1; break; }
  if (!ib->ivc_ready || ib->ivc_n != n || c->ivc_rebuild_matrix > 0 || changed)
    ib_build_ivc(L, ib->x, ib->y, ib->z);
  if (!ib->ivc_ready) { ib_force_explicit(L, dt, time); return; }
  double *bx = calloc(n, 8), *by = calloc(n, 8), *bz = calloc(n, 8);
  double *sx = calloc(n, 8), *sy = calloc(n, 8), *sz = calloc(n, 8);
  for (int i = 0; i < n; ++i) {
    double Ux, Uy, Uz;
    if (c->geometry_type == RC_GEOM_USER) {
      double px, py, pz;
      user_state(c, time, ib->xref[i + 1], ib->yref[i + 1], ib->zref[i + 1], &px, &py, &pz, &Ux, &Uy, &Uz);
    } else {
      rigid_target(L, ib->x[i], ib->y[i], ib->z[i], &Ux, &Uy, &Uz);
    }
    bx[i] = (Ux - ib->u[i]) / dt;
    by[i] = (Uy - ib->v[i]) / dt;
    bz[i] = (Uz - ib->w[i]) / dt;
  }
  rc_lu_solve(ib->ivc_lu, ib->ivc_piv, bx, sx, n);
  rc_lu_solve(ib->ivc_lu, ib->ivc_piv, by, sy, n);
  rc_lu_solve(ib->ivc_lu, ib->ivc_piv, bz, sz, n);
  for (int i = 0; i < n; ++i) { ib->fx[i] += sx[i]; ib->fy[i] += sy[i]; ib->fz[i] += sz[i]; }
  free(bx); free(by); free(bz); free(sx); free(sy); free(sz);
}

/* force_spreading_cic + force_spreading, :1330-1512 */
static void ib_spread(rc_level *L, fab *E) {
  rc_ib *ib = &L->ib;
  const rc_case *c = &L->c;
  const int kt = rc_delta_type_to_kernel(c->delta_type);
  const int renorm = (c->renormalize_delta != 0);
  const double d = (c->dim == 3) ? c->dx[0] * c->dx[1] * c->dx[2] : c->dx[0] * c->dx[1];
  const double z0 = (c->dim == 3) ? c->z0 : 0.0;
  for (int m = 0; m < ib->nm; ++m) {
    const double dvm = (g_mut == MUT_SPREAD_NO_DV) ? 1.0 : ib->dv;
    const double fx = ib->fx[m] * dvm, fy = ib->fy[m] * dvm, fz = ib->fz[m] * dvm;
    const double rx = ib->x[m] - c->x0, ry = ib->y[m] - c->y0;
    const double rz = (c->dim == 3) ? ib->z[m] - z0 : 0.0;
    const double ms = (g_mut == MUT_MOMENT_SIGN) ? -1.0 : 1.0;
    ib->mx[m] = ms * (ry * fz - rz * fy);
    ib->my[m] = ms * (rz * fx - rx * fz);
    ib->mz[m] = ms * (rx * fy - ry * fx);
    int b[3];
    marker_base(L, m, b);
    double tu[5], tv[5], tw[5];
    for (int t = 0; t < 5; ++t) {
      tu[t] = rc_delta(c->plo[0] + ((double)(b[0] + t - 2) + 0.5) * c->dx[0], ib->x[m], c->dx[0], kt);
      tv[t] = rc_delta(c->plo[1] + ((double)(b[1] + t - 2) + 0.5) * c->dx[1], ib->y[m], c->dx[1], kt);
      tw[t] = (c->dim == 3) ? rc_delta(c->plo[2] + ((double)(b[2] + t - 2) + 0.5) * c->dx[2], ib->z[m], c->dx[2], kt) : 1.0;
    }
    double part = 0.0;
    for (int ii = 0; ii < 5; ++ii)
      for (int jj = 0; jj < 5; ++jj) {
        if (c->dim == 3) {
          for (int kk = 0; kk < 5; ++kk) {
            const int ip = b[0] + ii - 2, jp = b[1] + jj - 2, kp = b[2] + kk - 2;
            if (ip < E->lo[0] || ip > E->hi[0] || jp < E->lo[1] || jp > E->hi[1] || kp < E->lo[2] || kp > E->hi[2]) continue;
            part += tu[ii] * tv[jj] * tw[kk] * d;
          }
        } else {
          const int ip = b[0] + ii - 2, jp = b[1] + jj - 2;
          if (ip < E->lo[0] || ip > E->hi[0] || jp < E->lo[1] || jp > E->hi[1]) continue;
          part += tu[ii] * tv[jj] * d;
        }
      }
    double scale = 1.0;
    if (renorm && part > 0.0) scale = 1.0 / part;
    for (int ii = 0; ii < 5; ++ii)
      for (int jj = 0; jj < 5; ++jj) {
        if (c->dim == 3) {
          for (int kk = 0; kk < 5; ++kk) {
            const int ip = b[0] + ii - 2, jp = b[1] + jj - 2, kp = b[2] + kk - 2;
            if (ip < E->lo[0] || ip > E->hi[0] || jp < E->lo[1] || jp > E->hi[1] || kp < E->lo[2] || kp > E->hi[2]) continue;
            const double dval = (tu[ii] * tv[jj] * tw[kk]) * scale;
            AT(*E, ip, jp, kp, 0) += dval * fx;
            AT(*E, ip, jp, kp, 1) += dval * fy;
            AT(*E, ip, jp, kp, 2) += dval * fz;
          }
        } else {
          const int ip = b[0] + ii - 2, jp = b[1] + jj - 2;
          if (ip < E->lo[0] || ip > E->hi[0] || jp < E->lo[1] || jp > E->hi[1]) continue;
          const double dval = (tu[ii] * tv[jj]) * scale;
          AT(*E, ip, jp, 0, 0) += dval * fx;
          AT(*E, ip, jp, 0, 1) += dval * fy;
        }
      }
  }
  sum_boundary(L, E, 3);
}

/* update_forcing, :1522-1917.  smS holds FillPatchMacro(macro_old);
 * F (3 comps, ng ghosts) receives dt * rho * a on the valid region. */
static void ib_update_forcing(rc_level *L, const fab *smS, fab *F, double dt, double time) {
  rc_ib *ib = &L->ib;
  const rc_case *c = &L->c;
  int lo[3], hi[3];
  for (int d = 0; d < 3; ++d) { lo[d] = L->vlo[d]; hi[d] = L->vhi[d]; }
  for (int d = 0; d < c->dim; ++d) { lo[d] -= 2; hi[d] += 2; }
  fab EV = fab_alloc(lo, hi, 3), EFI = fab_alloc(lo, hi, 3), EFT = fab_alloc(lo, hi, 3);
  for (int k = lo[2]; k <= hi[2]; ++k)
    for (int j = lo[1]; j <= hi[1]; ++j)
      for (int i = lo[0]; i <= hi[0]; ++i) {
        AT(EV, i, j, k, 0) = AT(*smS, i, j, k, 1);
        AT(EV, i, j, k, 1) = AT(*smS, i, j, k, 2);
        AT(EV, i, j, k, 2) = (c->dim == 3) ? AT(*smS, i, j, k, 3) : 0.0;
      }
  fill_boundary(L, &EV, 3);
  ib->last_sum[0] = ib->last_sum[1] = ib->last_sum[2] = 0.0;
  const int use_ivc = (c->coupling_method == RC_COUPLING_IVC);
  const int nloop = use_ivc ? 1 : (c->loop_ns > 1 ? c->loop_ns : 1);
  for (int it = 0; it < nloop; ++it) {
    fab_setval(&EFI, 0.0);
    ib_update_markers(L, time);
    ib_interp(L, &EV);
    if (use_ivc) ib_force_ivc(L, dt, time);
    else ib_force_explicit(L, dt, time);
    ib_spread(L, &EFI);
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int m = 0; m < ib->nm; ++m) {
      sx += ib->fx[m] * ib->dv;
      sy += ib->fy[m] * ib->dv;
      sz += ib->fz[m] * ib->dv;
    }
    ib->last_sum[0] += sx; ib->last_sum[1] += sy; ib->last_sum[2] += sz;
    for (int k = L->vlo[2]; k <= L->vhi[2]; ++k)
      for (int j = L->vlo[1]; j <= L->vhi[1]; ++j)
        for (int i = L->vlo[0]; i <= L->vhi[0]; ++i)
          for (int n = 0; n < 3; ++n) {
            AT(EV, i, j, k, n) += dt * AT(EFI, i, j, k, n);
            AT(EFT, i, j, k, n) += AT(EFI, i, j, k, n);
          }
  }
  for (int k = L->vlo[2]; k <= L->vhi[2]; ++k)
    for (int j = L->vlo[1]; j <= L->vhi[1]; ++j)
      for (int i = L->vlo[0]; i <= L->vhi[0]; ++i)
        for (int n = 0; n < 3; ++n)
          AT(*F, i, j, k, n) = ((g_mut == MUT_EXPORT_NO_DT) ? 1.0 : dt) * AT(*smS, i, j, k, 0) * AT(EFT, i, j, k, n);
  fill_boundary(L, F, 3);
  fab_free(&EV); fab_free(&EFI); fab_free(&EFT);
}

/* ------------------------------------------------------------------ */
/* Level lifecycle                                                       */
/* ------------------------------------------------------------------ */
rc_level *rc_create(const rc_case *c, const double *init_macro) {
  rc_level *L = (rc_level *)calloc(1, sizeof(rc_level));
  L->c = *c;
  if (L->c.dim == 2) { L->c.n[2] = 1; L->c.periodic[2] = 1; }
  L->ng = 3;
  int lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    L->dom_lo[d] = 0; L->dom_hi[d] = L->c.n[d] - 1; lo[d] = 0; hi[d] = L->c.n[d] - 1;
    L->vlo[d] = L->dom_lo[d]; L->vhi[d] = L->dom_hi[d];
  }
  for (int d = 0; d < L->c.dim; ++d) { lo[d] -= L->ng; hi[d] += L->ng; }
  L->f_new = fab_alloc(lo, hi, L->c.q);
  L->f_old = fab_alloc(lo, hi, L->c.q);
  L->m_new = fab_alloc(lo, hi, L->c.nmac);
  L->m_old = fab_alloc(lo, hi, L->c.nmac);
  L->forcing = fab_alloc(lo, hi, 3);
  /* AmrCoreLBM.cpp:225-229 */
  L->tau = 3.0 * L->c.nu / L->c.dx[0] + 0.5;
  L->dt = L->c.dx[0];
  L->t_new = 0.0;
  L->t_old = -1.0e100;
  /* initdata on the valid region; ghosts are refilled by FillPatch */
  const int n0 = L->c.n[0], n1 = L->c.n[1], n2 = L->c.n[2];
  const int64_t nv = (int64_t)n0 * n1 * n2;
  for (int k = 0; k < n2; ++k)
    for (int j = 0; j < n1; ++j)
      for (int i = 0; i < n0; ++i) {
        const int64_t t = ((int64_t)k * n1 + j) * n0 + i;
        for (int n = 0; n < 4; ++n) AT(L->m_new, i, j, k, n) = init_macro[n * nv + t];
      }
  memcpy(L->m_old.p, L->m_new.p, sizeof(double) * (size_t)(L->m_new.ns * L->m_new.nc));
  /* InitEquilibrium, AmrCoreLBM.cpp:2146-2161 */
  for (int k = 0; k < n2; ++k)
    for (int j = 0; j < n1; ++j)
      for (int i = 0; i < n0; ++i) {
        const double r = fmax(AT(L->m_new, i, j, k, 0), 1.0e-12);
        const double u0 = AT(L->m_new, i, j, k, 1), v0 = AT(L->m_new, i, j, k, 2), w0 = AT(L->m_new, i, j, k, 3);
        const double u2 = u0 * u0 + v0 * v0 + w0 * w0;
        for (int q = 0; q < L->c.q; ++q) {
          const double cu = L->c.cx[q] * u0 + L->c.cy[q] * v0 + L->c.cz[q] * w0;
          const double feq = L->c.w[q] * r * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * u2);
          AT(L->f_new, i, j, k, q) = feq;
          AT(L->f_old, i, j, k, q) = feq;
        }
      }
  if (L->c.use_ib) {
    ib_build(L);
    ib_update_markers(L, 0.0);
  }
  return L;
}

static void rc_ib_free(rc_ib *ib) {
  free(ib->x); free(ib->y); free(ib->z); free(ib->u); free(ib->v); free(ib->w);
  free(ib->fx); free(ib->fy); free(ib->fz); free(ib->mx); free(ib->my); free(ib->mz);
  free(ib->phi); free(ib->theta); free(ib->xref); free(ib->yref); free(ib->zref);
  free(ib->ivc_lu); free(ib->ivc_piv); free(ib->xc); free(ib->yc); free(ib->zc);
}

void rc_destroy(rc_level *L) {
  if (!L) return;
  fab_free(&L->f_new); fab_free(&L->f_old); fab_free(&L->m_new); fab_free(&L->m_old);
  fab_free(&L->forcing);
  if (L->c.use_ib) rc_ib_free(&L->ib);
  free(L);
}

static void amr_fill_macro(struct rc_amr *A, int l, double time, fab *dst, int ncomp);
static void amr_fill_meso(const rc_level *L, fab *sfS, const fab *smS);

/* AdvancePhiAtLevel, Source/Core/AdvanceLBMAtLevel.cpp:20-387 */
static void advance_level(rc_level *L, double time, double dt_lev) {
  rc_case *c = &L->c;
  const int dim = c->dim, Q = c->q;
  fab t = L->f_new; L->f_new = L->f_old; L->f_old = t;
  t = L->m_new; L->m_new = L->m_old; L->m_old = t;

  fab fx_cc = fab_alloc(L->f_new.lo, L->f_new.hi, 3);
  if (c->use_prescribed_force)
    for (int64_t s = 0; s < fx_cc.ns; ++s) {
      fx_cc.p[s] = c->prescribed_force[0];
      fx_cc.p[s + fx_cc.ns] = c->prescribed_force[1];
      fx_cc.p[s + 2 * fx_cc.ns] = c->prescribed_force[2];
    }
  fab sfS = fab_alloc(L->f_new.lo, L->f_new.hi, Q);
  fab smS = fab_alloc(L->m_new.lo, L->m_new.hi, c->nmac);
  if (L->amr && L->lev > 0) {
    /* FillPatchMacro / FillPatchMesoscopic for lev > 0 (AmrCoreLBM.cpp:1545-1627) */
    amr_fill_macro(L->amr, L->lev, time, &smS, c->nmac);
    amr_fill_meso(L, &sfS, &smS);
  } else {
    fill_patch(L, &sfS, &L->f_old, Q, 1);
    fill_patch(L, &smS, &L->m_old, c->nmac, 0);
  }

  if (c->use_ib) {
    ib_update_forcing(L, &smS, &fx_cc, dt_lev, time);
    for (int k = L->vlo[2]; k <= L->vhi[2]; ++k)
      for (int j = L->vlo[1]; j <= L->vhi[1]; ++j)
        for (int i = L->vlo[0]; i <= L->vhi[0]; ++i)
          for (int n = 0; n < 3; ++n) AT(L->forcing, i, j, k, n) = AT(fx_cc, i, j, k, n);
  }
  fill_boundary(L, &fx_cc, 3);

  const double tau = L->tau;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = L->vlo[2]; k <= L->vhi[2]; ++k)
    for (int j = L->vlo[1]; j <= L->vhi[1]; ++j)
      for (int i = L->vlo[0]; i <= L->vhi[0]; ++i) {
        const double rho = AT(smS, i, j, k, 0), u = AT(smS, i, j, k, 1), v = AT(smS, i, j, k, 2), w = AT(smS, i, j, k, 3);
        double feq[RC_MAXQ], f[RC_MAXQ];
        for (int q = 0; q < Q; ++q) {
          feq[q] = rc_feq(c->w[q], c->cx[q], c->cy[q], c->cz[q], rho, u, v, w, dim);
          f[q] = AT(sfS, i, j, k, q);
        }
        rc_collide_forced(f, feq, Q, tau, c->w, c->cx, c->cy, c->cz, rho, u, v, w,
                          AT(fx_cc, i, j, k, 0), AT(fx_cc, i, j, k, 1), AT(fx_cc, i, j, k, 2), dim);
        for (int q = 0; q < Q; ++q) AT(sfS, i, j, k, q) = f[q];
      }
  fill_boundary(L, &sfS, Q);

  /* bounce-back ghost fill, :265-373 (ghosts from valid cells: parallel) */
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = sfS.lo[2]; k <= sfS.hi[2]; ++k)
    for (int j = sfS.lo[1]; j <= sfS.hi[1]; ++j)
      for (int i = sfS.lo[0]; i <= sfS.hi[0]; ++i) {
        const int iv[3] = {i, j, k};
        int valid = 1;
        for (int d = 0; d < 3; ++d) if (iv[d] < L->vlo[d] || iv[d] > L->vhi[d]) valid = 0;
        if (valid) continue;
        int has_bb = 0, unsup = 0;
        for (int d = 0; d < dim; ++d) {
          if (iv[d] < L->dom_lo[d]) { if (c->bc_lo[d] == RC_USER_1) has_bb = 1; else unsup = 1; }
          else if (iv[d] > L->dom_hi[d]) { if (c->bc_hi[d] == RC_USER_1) has_bb = 1; else unsup = 1; }
        }
        if (!has_bb || unsup) continue;
        for (int q = 0; q < Q; ++q) {
          const int it = i + (int)c->cx[q], jt = j + (int)c->cy[q];
          const int kt = (dim == 3) ? k + (int)c->cz[q] : k;
          if (it < L->vlo[0] || it > L->vhi[0] || jt < L->vlo[1] || jt > L->vhi[1] ||
              kt < L->vlo[2] || kt > L->vhi[2]) continue;
          const int rx = -(int)c->cx[q], ry = -(int)c->cy[q], rz = (dim == 3) ? -(int)c->cz[q] : 0;
          int qr = q;
          for (int p = 0; p < Q; ++p)
            if ((int)c->cx[p] == rx && (int)c->cy[p] == ry && (int)c->cz[p] == rz) { qr = p; break; }
          AT(sfS, i, j, k, q) = AT(sfS, it, jt, kt, qr);
        }
      }

  /* stream (Source/LBM/stream.H:9-22) and forced macros
   * (Source/LBM/calculateMacroForcing.H:9-32) */
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = L->vlo[2]; k <= L->vhi[2]; ++k)
    for (int j = L->vlo[1]; j <= L->vhi[1]; ++j)
      for (int i = L->vlo[0]; i <= L->vhi[0]; ++i) {
        for (int q = 0; q < Q; ++q)
          AT(L->f_new, i, j, k, q) = AT(sfS, i - (int)c->cx[q], j - (int)c->cy[q], k - (int)c->cz[q], q);
        double r = 0, m1 = 0, m2 = 0, m3 = 0;
        for (int q = 0; q < Q; ++q) {
          const double fq = AT(L->f_new, i, j, k, q);
          r += fq;
          m1 += c->cx[q] * fq;
          m2 += c->cy[q] * fq;
          m3 += c->cz[q] * fq;
        }
        const double hf = (g_mut == MUT_GUO_NO_HALF) ? 0.0 : 0.5;
        AT(L->m_new, i, j, k, 0) = r;
        AT(L->m_new, i, j, k, 1) = (m1 + hf * AT(fx_cc, i, j, k, 0)) / r;
        AT(L->m_new, i, j, k, 2) = (m2 + hf * AT(fx_cc, i, j, k, 1)) / r;
        AT(L->m_new, i, j, k, 3) = (m3 + hf * AT(fx_cc, i, j, k, 2)) / r;
      }
  fab_free(&fx_cc); fab_free(&sfS); fab_free(&smS);
}

/* ComputeIBForce, Source/Core/AmrCoreLBM.cpp:456-701 */
static void compute_ib_force(rc_level *L, double time, int step) {
  rc_case *c = &L->c;
  double Fx_eul, Fy_eul, Fz_eul, Fx_m = 0, Fy_m = 0, Fz_m = 0;
  {
    double sx = 0, sy = 0, sz = 0;
    for (int k = L->vlo[2]; k <= L->vhi[2]; ++k)
      for (int j = L->vlo[1]; j <= L->vhi[1]; ++j)
        for (int i = L->vlo[0]; i <= L->vhi[0]; ++i) {
          sx += AT(L->forcing, i, j, k, 0);
          sy += AT(L->forcing, i, j, k, 1);
          sz += AT(L->forcing, i, j, k, 2);
        }
    if (c->dim != 3) sz = 0.0;
    double dV = c->dx[0] * c->dx[1];
    if (c->dim == 3) dV *= c->dx[2];
    Fx_eul = -(sx * dV) / L->dt;
    Fy_eul = -(sy * dV) / L->dt;
    Fz_eul = -(sz * dV) / L->dt;
  }
  Fx_m = -L->ib.last_sum[0]; Fy_m = -L->ib.last_sum[1]; Fz_m = -L->ib.last_sum[2];
  const double Fx_me = Fx_eul, Fy_me = Fy_eul, Fz_me = Fz_eul;
  double Fxb = Fx_eul, Fyb = Fy_eul, Fzb = Fz_eul;
  if (c->force_eval_method == RC_FORCE_EVAL_MARKER) { Fxb = Fx_m; Fyb = Fy_m; Fzb = Fz_m; }
  else if (c->force_eval_method == RC_FORCE_EVAL_ME) { Fxb = Fx_me; Fyb = Fy_me; Fzb = Fz_me; }
  double Fxt = Fxb, Fyt = Fyb, Fzt = Fzb;
  if (L->have_prev) {
    Fxt = 0.5 * (L->prev_body[0] + Fxb);
    Fyt = 0.5 * (L->prev_body[1] + Fyb);
    Fzt = 0.5 * (L->prev_body[2] + Fzb);
  }
  L->prev_body[0] = Fxb; L->prev_body[1] = Fyb; L->prev_body[2] = Fzb;
  L->have_prev = 1;
  const double rho_ref = 1.0, U = c->U0, D = 2.0 * c->R;
  const double half = (g_mut == MUT_CD_NO_HALF) ? 1.0 : 0.5;
  const double denom = (c->dim == 3) ? half * rho_ref * U * U * (rc_pi() * c->R * c->R)
                                     : half * rho_ref * U * U * D;
  double Cd = 0, Cl = 0, Cdr = 0, Clr = 0, Cde = 0, Cle = 0, Cdm = 0, Clm = 0, Cdx = 0, Clx = 0;
  double Cz = 0, Czr = 0, Cze = 0, Czm = 0, Czx = 0;
  if (denom > 0.0) {
    Cdr = Fxb / denom; Clr = Fyb / denom; Cd = Fxt / denom; Cl = Fyt / denom;
    Cde = Fx_eul / denom; Cle = Fy_eul / denom; Cdm = Fx_m / denom; Clm = Fy_m / denom;
    Cdx = Fx_me / denom; Clx = Fy_me / denom;
    Czr = Fzb / denom; Cz = Fzt / denom; Cze = Fz_eul / denom; Czm = Fz_m / denom; Czx = Fz_me / denom;
  }
  const double tconv = (D > 0.0) ? (time * U / D) : 0.0;
  double *r = L->force_rec;
  int n = 0;
  r[n++] = step; r[n++] = time;
  r[n++] = Fxt; r[n++] = Fyt; r[n++] = Cd; r[n++] = Cl;
  r[n++] = Fxb; r[n++] = Fyb; r[n++] = Cdr; r[n++] = Clr;
  r[n++] = Fx_eul; r[n++] = Fy_eul; r[n++] = Cde; r[n++] = Cle;
  r[n++] = Fx_m; r[n++] = Fy_m; r[n++] = Cdm; r[n++] = Clm;
  r[n++] = Fx_me; r[n++] = Fy_me; r[n++] = Cdx; r[n++] = Clx; r[n++] = tconv;
  if (c->dim == 3) {
    r[n++] = Fzt; r[n++] = Cz; r[n++] = Fzb; r[n++] = Czr; r[n++] = Fz_eul; r[n++] = Cze;
    r[n++] = Fz_m; r[n++] = Czm; r[n++] = Fz_me; r[n++] = Czx;
  }
  L->force_rec_n = n;
}

void rc_step(rc_level *L) {
  /* timeStepWithSubcycling, AmrCoreLBM.cpp:1730-1735 */
  const double time = L->t_new;
  L->t_old = L->t_new;
  L->t_new += L->dt;
  advance_level(L, time, L->dt);
  ++L->istep;
  if (L->c.use_ib) compute_ib_force(L, time + L->dt, L->istep);
}

void rc_advance(rc_level *L, int nsteps) {
  for (int s = 0; s < nsteps; ++s) rc_step(L);
}

double rc_mass_sum(const rc_level *L) {
  double s = 0.0;
  for (int k = L->vlo[2]; k <= L->vhi[2]; ++k)
    for (int j = L->vlo[1]; j <= L->vhi[1]; ++j)
      for (int i = L->vlo[0]; i <= L->vhi[0]; ++i) s += AT(L->m_new, i, j, k, 0);
  return s;
}

static int get_valid(const rc_level *L, const fab *F, int ncomp, double *out) {
  const int n0 = L->vhi[0] - L->vlo[0] + 1, n1 = L->vhi[1] - L->vlo[1] + 1, n2 = L->vhi[2] - L->vlo[2] + 1;
  const int64_t nv = (int64_t)n0 * n1 * n2;
  for (int n = 0; n < ncomp; ++n)
    for (int k = 0; k < n2; ++k)
      for (int j = 0; j < n1; ++j)
        for (int i = 0; i < n0; ++i)
          out[n * nv + ((int64_t)k * n1 + j) * n0 + i] = AT(*F, L->vlo[0] + i, L->vlo[1] + j, L->vlo[2] + k, n);
  return 0;
}
int rc_get_f(const rc_level *L, double *out) { return get_valid(L, &L->f_new, L->c.q, out); }
int rc_get_macro(const rc_level *L, double *out) { return get_valid(L, &L->m_new, 4, out); }
int rc_get_forcing(const rc_level *L, double *out) { return get_valid(L, &L->forcing, 3, out); }

int rc_force_record(const rc_level *L, double *out, int maxn) {
  int n = L->force_rec_n < maxn ? L->force_rec_n : maxn;
  for (int i = 0; i < n; ++i) out[i] = L->force_rec[i];
  return L->force_rec_n;
}
int rc_num_markers(const rc_level *L) { return L->c.use_ib ? L->ib.nm : 0; }
void rc_marker_state(const rc_level *L, double *xyz, double *uvw, double *fxyz) {
  const rc_ib *ib = &L->ib;
  for (int m = 0; m < ib->nm; ++m) {
    if (xyz) { xyz[3 * m] = ib->x[m]; xyz[3 * m + 1] = ib->y[m]; xyz[3 * m + 2] = ib->z[m]; }
    if (uvw) { uvw[3 * m] = ib->u[m]; uvw[3 * m + 1] = ib->v[m]; uvw[3 * m + 2] = ib->w[m]; }
    if (fxyz) { fxyz[3 * m] = ib->fx[m]; fxyz[3 * m + 1] = ib->fy[m]; fxyz[3 * m + 2] = ib->fz[m]; }
  }
}
void rc_last_marker_force_sum(const rc_level *L, double out[3]) {
  for (int d = 0; d < 3; ++d) out[d] = L->ib.last_sum[d];
}
double rc_marker_dv(const rc_level *L) { return L->ib.dv; }
double rc_tau(const rc_level *L) { return L->tau; }

int rc_build_markers(const rc_case *c, double time, double *xyz, int maxn, double *dv_out) {
  rc_level tmp;
  memset(&tmp, 0, sizeof(tmp));
  tmp.c = *c;
  if (tmp.c.dim == 2) tmp.c.n[2] = 1;
  for (int d = 0; d < 3; ++d) {
    tmp.dom_lo[d] = 0; tmp.dom_hi[d] = tmp.c.n[d] - 1;
    tmp.vlo[d] = tmp.dom_lo[d]; tmp.vhi[d] = tmp.dom_hi[d];
  }
  ib_build(&tmp);
  ib_update_markers(&tmp, time);
  const int nm = tmp.ib.nm;
  for (int m = 0; m < nm && m < maxn; ++m) {
    xyz[3 * m] = tmp.ib.x[m]; xyz[3 * m + 1] = tmp.ib.y[m]; xyz[3 * m + 2] = tmp.ib.z[m];
  }
  if (dv_out) *dv_out = tmp.ib.dv;
  rc_ib_free(&tmp.ib);
  return nm;
}

/* Stage probes of update_forcing for the unit checks of the IB
 * restatement.  Both place the markers at `time` first
 * (update_lagrangian_marker, IBMarkerDF.cpp:499-609). */
static fab probe_fab(const rc_level *L) {
  int lo[3], hi[3];
  for (int d = 0; d < 3; ++d) { lo[d] = L->vlo[d]; hi[d] = L->vhi[d]; }
  for (int d = 0; d < L->c.dim; ++d) { lo[d] -= 2; hi[d] += 2; }
  return fab_alloc(lo, hi, 3);
}

/* velocity_interpolation (:611-739) of a caller-given EulerVel over the
 * valid box grown by 2 cells ([3][grown cells], x fastest); periodic ghost
 * images are refilled from the valid region as FillBoundary does.
 * U_out [nm][3]; returns the marker count. */
int rc_ib_interp_probe(rc_level *L, double time, const double *ev, double *U_out) {
  if (!L->c.use_ib) return -1;
  fab E = probe_fab(L);
  memcpy(E.p, ev, sizeof(double) * (size_t)(3 * E.ns));
  ib_update_markers(L, time);
  ib_interp(L, &E);
  for (int m = 0; m < L->ib.nm; ++m) {
    U_out[3 * m] = L->ib.u[m]; U_out[3 * m + 1] = L->ib.v[m]; U_out[3 * m + 2] = L->ib.w[m];
  }
  fab_free(&E);
  return L->ib.nm;
}

/* force_spreading (:1330-1512) of caller-given marker forces F [nm][3]:
 * EulerForce on the valid box after SumBoundary ([3][valid cells]) and the
 * marker moments r x (F dv) [nm][3]. */
int rc_ib_spread_probe(rc_level *L, double time, const double *F, double *efi_out, double *mom_out) {
  if (!L->c.use_ib) return -1;
  fab E = probe_fab(L);
  ib_update_markers(L, time);
  for (int m = 0; m < L->ib.nm; ++m) {
    L->ib.fx[m] = F[3 * m]; L->ib.fy[m] = F[3 * m + 1]; L->ib.fz[m] = F[3 * m + 2];
  }
  ib_spread(L, &E);
  get_valid(L, &E, 3, efi_out);
  for (int m = 0; m < L->ib.nm; ++m) {
    mom_out[3 * m] = L->ib.mx[m]; mom_out[3 * m + 1] = L->ib.my[m]; mom_out[3 * m + 2] = L->ib.mz[m];
  }
  fab_free(&E);
  return L->ib.nm;
}

/* ------------------------------------------------------------------ */
/* AMR hierarchy: one box per level, refinement ratio 2, subcycling.     */
/* Source/Core/AmrCoreLBM.cpp: timeStepWithSubcycling :1688-1754,        */
/* FillPatchMesoscopic lev > 0 :1545-1627, FillPatchMacro :2166-2198,    */
/* GetDataMacro :2239-2269, AverageDownTo :1520-1526,                    */
/* MakeNewLevelFromCoarse :751-830, UpdateDerivedMacroFields :150-168.   */
/* ------------------------------------------------------------------ */
#define RC_MAXLEV 4
#define RC_MAXC 8

struct rc_amr {
  int nlev;
  rc_level *L[RC_MAXLEV];
  int regrid_int;
  int istep;
  rc_case c0;                      /* the case as given (refinement-region parameters) */
  int regrid_on;
  rc_regrid rg;
  int last_regrid_step[RC_MAXLEV];
  int nregrid;                     /* regrids that changed at least one box */
};

static void amr_regrid(struct rc_amr *A, int lbase, double time);

static int floordiv2(int i) { return (i >= 0) ? i / 2 : -((1 - i) / 2); }

static int in_box(const rc_level *L, const int iv[3]) {
  for (int d = 0; d < 3; ++d)
    if (iv[d] < L->vlo[d] || iv[d] > L->vhi[d]) return 0;
  return 1;
}

/* GetDataMacro at one cell: the level's macro state at `time` (new, old,
 * or FillPatchSingleLevel's two-time LinComb a*old + b*new with
 * a = (t1 - t)/(t1 - t0), b = (t - t0)/(t1 - t0)).  The form of the
 * LinComb is AMReX's (version unknown): parity unpinned. */
static double macro_at_time(const rc_level *L, double time, int i, int j, int k, int n) {
  const double SENT = 1.0e50;
  if (fabs(L->t_old) > SENT || fabs(L->t_new) > SENT) return AT(L->m_new, i, j, k, n);
  const double dtloc = fabs(L->t_new - L->t_old);
  const double teps = fmax(1.0e-12, dtloc * 1.0e-3);
  if (time >= L->t_new - teps && time <= L->t_new + teps) return AT(L->m_new, i, j, k, n);
  if (time >= L->t_old - teps && time <= L->t_old + teps) return AT(L->m_old, i, j, k, n);
  const double a = (L->t_new - time) / (L->t_new - L->t_old);
  const double b = (time - L->t_old) / (L->t_new - L->t_old);
  return a * AT(L->m_old, i, j, k, n) + b * AT(L->m_new, i, j, k, n);
}

/* AMReX cell_cons_interp (CellConservativeLinear with linear limiting) at
 * fine cell fiv from the coarse level's macro state at `time`: centred
 * slopes dc = (u[+1] - u[-1])/2 per direction, limited slope
 * sign(dc) min(|dc|, min(|df|, |db|)) (0 at an extremum) with
 * df = 2 (u[+1] - u0), db = 2 (u0 - u[-1]); one factor per direction,
 * min over the ncomp components of limited/centred, scales every
 * component; fine = u0 + xoff sx + yoff sy (+ zoff sz), offsets +-1/4.
 * Follows mf_cell_cons_lin_interp_llslope / mf_cell_cons_lin_interp of
 * AMReX; the version the reference builds against is not recoverable, so
 * the limiter details are parity unpinned.  Pins: linear fields are
 * reproduced exactly, the children of a coarse cell average to it. */
static void cons_interp_cell(const rc_level *Lc, double time, int dim, int ncomp, const int fiv[3], double *out) {
  const int ic[3] = {floordiv2(fiv[0]), floordiv2(fiv[1]), dim == 3 ? floordiv2(fiv[2]) : fiv[2]};
  double u0[RC_MAXC], slope[3][RC_MAXC];
  double sf[3] = {1.0, 1.0, 1.0};
  for (int d = 0; d < dim; ++d) {
    int p[3] = {ic[0], ic[1], ic[2]}, m[3] = {ic[0], ic[1], ic[2]};
    p[d] += 1; m[d] -= 1;
    if (!in_box(Lc, p) || !in_box(Lc, m)) abort(); /* nesting margin violated */
  }
  for (int n = 0; n < ncomp; ++n) u0[n] = macro_at_time(Lc, time, ic[0], ic[1], ic[2], n);
  for (int d = 0; d < dim; ++d) {
    int p[3] = {ic[0], ic[1], ic[2]}, m[3] = {ic[0], ic[1], ic[2]};
    p[d] += 1; m[d] -= 1;
    for (int n = 0; n < ncomp; ++n) {
      const double up = macro_at_time(Lc, time, p[0], p[1], p[2], n);
      const double um = macro_at_time(Lc, time, m[0], m[1], m[2], n);
      const double dc = 0.5 * (up - um);
      const double df = 2.0 * (up - u0[n]);
      const double db = 2.0 * (u0[n] - um);
      double sx = (df * db >= 0.0) ? fmin(fabs(df), fabs(db)) : 0.0;
      sx = copysign(1.0, dc) * fmin(sx, fabs(dc));
      if (dc != 0.0) sf[d] = fmin(sf[d], sx / dc);
      slope[d][n] = dc;
    }
  }
  for (int d = 0; d < dim; ++d)
    for (int n = 0; n < ncomp; ++n) slope[d][n] *= sf[d];
  double off[3];
  for (int d = 0; d < dim; ++d) off[d] = ((double)(fiv[d] - ic[d] * 2) + 0.5) / 2.0 - 0.5;
  for (int n = 0; n < ncomp; ++n) {
    double v = u0[n] + off[0] * slope[0][n];
    v = v + off[1] * slope[1][n];
    if (dim == 3) v = v + off[2] * slope[2][n];
    out[n] = v;
  }
}

/* FillPatchMacro(l, time) for l > 0 (FillPatchTwoLevels with
 * cell_cons_interp): valid cells from the level's own data, every other
 * cell of dst by interpolation of the coarse level's data at `time`.
 * Fine boxes sit inside the domain, so no physical ghost occurs. */
static void amr_fill_macro(struct rc_amr *A, int l, double time, fab *dst, int ncomp) {
  const rc_level *L = A->L[l], *Lc = A->L[l - 1];
  const int dim = L->c.dim;
  for (int k = dst->lo[2]; k <= dst->hi[2]; ++k)
    for (int j = dst->lo[1]; j <= dst->hi[1]; ++j)
      for (int i = dst->lo[0]; i <= dst->hi[0]; ++i) {
        const int iv[3] = {i, j, k};
        if (in_box(L, iv)) {
          for (int n = 0; n < ncomp; ++n) AT(*dst, i, j, k, n) = macro_at_time(L, time, i, j, k, n);
          continue;
        }
        double v[RC_MAXC];
        cons_interp_cell(Lc, time, dim, ncomp, iv, v);
        for (int n = 0; n < ncomp; ++n) AT(*dst, i, j, k, n) = v[n];
      }
}

/* FillPatchMesoscopic for l > 0: valid cells from f_old, ghost cells
 * replaced by the equilibrium of the FillPatchMacro values
 * (AmrCoreLBM.cpp:1562-1624; the interpolated populations are discarded). */
static void amr_fill_meso(const rc_level *L, fab *sfS, const fab *smS) {
  const rc_case *c = &L->c;
  for (int k = sfS->lo[2]; k <= sfS->hi[2]; ++k)
    for (int j = sfS->lo[1]; j <= sfS->hi[1]; ++j)
      for (int i = sfS->lo[0]; i <= sfS->hi[0]; ++i) {
        const int iv[3] = {i, j, k};
        if (in_box(L, iv)) {
          for (int q = 0; q < c->q; ++q) AT(*sfS, i, j, k, q) = AT(L->f_old, i, j, k, q);
          continue;
        }
        const double r = fmax(AT(*smS, i, j, k, 0), 1.0e-12);
        const double u0 = AT(*smS, i, j, k, 1), v0 = AT(*smS, i, j, k, 2), w0 = AT(*smS, i, j, k, 3);
        const double u2 = u0 * u0 + v0 * v0 + w0 * w0;
        for (int q = 0; q < c->q; ++q) {
          const double cu = c->cx[q] * u0 + c->cy[q] * v0 + c->cz[q] * w0;
          AT(*sfS, i, j, k, q) = c->w[q] * r * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * u2);
        }
      }
}

/* vorticityDiagnostic / visPara (Source/LBM/visPara.H:10-54) */
static void vis_para(const rc_case *c, const fab *m, int i, int j, int k, double *vort, double *P) {
  const double dx = c->dx[0], dy = c->dx[1], dz = c->dx[2];
  *P = (AT(*m, i, j, k, 0) - 1.0) / 3.0;
  const double dv_dx = (AT(*m, i + 1, j, k, 2) - AT(*m, i - 1, j, k, 2)) / (2.0 * dx);
  const double du_dy = (AT(*m, i, j + 1, k, 1) - AT(*m, i, j - 1, k, 1)) / (2.0 * dy);
  if (c->dim == 3) {
    const double du_dz = (AT(*m, i, j, k + 1, 1) - AT(*m, i, j, k - 1, 1)) / (2.0 * dz);
    const double dw_dx = (AT(*m, i + 1, j, k, 3) - AT(*m, i - 1, j, k, 3)) / (2.0 * dx);
    const double dv_dz = (AT(*m, i, j, k + 1, 2) - AT(*m, i, j, k - 1, 2)) / (2.0 * dz);
    const double dw_dy = (AT(*m, i, j + 1, k, 3) - AT(*m, i, j - 1, k, 3)) / (2.0 * dy);
    const double ox = dw_dy - dv_dz, oy = du_dz - dw_dx, oz = dv_dx - du_dy;
    *vort = sqrt(ox * ox + oy * oy + oz * oz);
  } else {
    *vort = dv_dx - du_dy;
  }
}

/* UpdateDerivedMacroFields (AmrCoreLBM.cpp:150-168): FillPatchMacro of
 * comps 0..3 with one ghost at `time`, then vorticity and pressure into
 * macro_new comps 4 and 5. */
static void update_derived(struct rc_amr *A, rc_level *L, double time) {
  if (L->c.nmac < 5) return;
  int lo[3], hi[3];
  for (int d = 0; d < 3; ++d) { lo[d] = L->vlo[d]; hi[d] = L->vhi[d]; }
  for (int d = 0; d < L->c.dim; ++d) { lo[d] -= 1; hi[d] += 1; }
  fab mp = fab_alloc(lo, hi, 4);
  if (A && L->lev > 0) {
    amr_fill_macro(A, L->lev, time, &mp, 4);
  } else {
    fab src = fab_alloc(L->m_new.lo, L->m_new.hi, 4);
    for (int k = L->vlo[2]; k <= L->vhi[2]; ++k)
      for (int j = L->vlo[1]; j <= L->vhi[1]; ++j)
        for (int i = L->vlo[0]; i <= L->vhi[0]; ++i)
          for (int n = 0; n < 4; ++n) AT(src, i, j, k, n) = macro_at_time(L, time, i, j, k, n);
    fill_patch(L, &mp, &src, 4, 0);
    fab_free(&src);
  }
  for (int k = L->vlo[2]; k <= L->vhi[2]; ++k)
    for (int j = L->vlo[1]; j <= L->vhi[1]; ++j)
      for (int i = L->vlo[0]; i <= L->vhi[0]; ++i) {
        double vort, P;
        vis_para(&L->c, &mp, i, j, k, &vort, &P);
        AT(L->m_new, i, j, k, 4) = vort;
        if (L->c.nmac > 5) AT(L->m_new, i, j, k, 5) = P;
      }
  fab_free(&mp);
}

/* AverageDownTo (AmrCoreLBM.cpp:1520-1526): amrex::average_down of f and
 * of every macro component, c = sum over (kref, jref, iref) of the fine
 * children, crse = c / ratio^dim. */
static void average_down_to(struct rc_amr *A, int lc) {
  rc_level *Lc = A->L[lc], *Lf = A->L[lc + 1];
  const int dim = Lc->c.dim;
  const int fk = (dim == 3) ? 2 : 1;
  const double volfrac = 1.0 / (double)(dim == 3 ? 8 : 4);
  int clo[3], chi[3];
  for (int d = 0; d < 3; ++d) {
    clo[d] = (d < dim) ? floordiv2(Lf->vlo[d]) : Lf->vlo[d];
    chi[d] = (d < dim) ? floordiv2(Lf->vhi[d]) : Lf->vhi[d];
  }
  for (int k = clo[2]; k <= chi[2]; ++k)
    for (int j = clo[1]; j <= chi[1]; ++j)
      for (int i = clo[0]; i <= chi[0]; ++i) {
        const int ii = 2 * i, jj = 2 * j, kk = (dim == 3) ? 2 * k : k;
        for (int n = 0; n < Lc->c.q; ++n) {
          double c = 0.0;
          for (int kr = 0; kr < fk; ++kr)
            for (int jr = 0; jr < 2; ++jr)
              for (int ir = 0; ir < 2; ++ir) c += AT(Lf->f_new, ii + ir, jj + jr, kk + kr, n);
          AT(Lc->f_new, i, j, k, n) = volfrac * c;
        }
        for (int n = 0; n < Lc->c.nmac; ++n) {
          double c = 0.0;
          for (int kr = 0; kr < fk; ++kr)
            for (int jr = 0; jr < 2; ++jr)
              for (int ir = 0; ir < 2; ++ir) c += AT(Lf->m_new, ii + ir, jj + jr, kk + kr, n);
          AT(Lc->m_new, i, j, k, n) = volfrac * c;
        }
      }
}

/* timeStepWithSubcycling (AmrCoreLBM.cpp:1688-1754); the regrid check of
 * :1692-1720 runs when regridding is switched on. */
static void amr_substep(struct rc_amr *A, int l, double time) {
  if (A->regrid_on && l < A->nlev - 1 && A->L[l]->istep > A->last_regrid_step[l] &&
      A->L[l]->istep % A->regrid_int == 0) {
    amr_regrid(A, l, time);
    for (int k = l; k < A->nlev; ++k) A->last_regrid_step[k] = A->L[k]->istep;
  }
  rc_level *L = A->L[l];
  L->t_old = L->t_new;
  L->t_new += L->dt;
  advance_level(L, time, L->dt);
  const int step_lev = L->istep;
  if (A->regrid_int > 0 && (step_lev + 1) % A->regrid_int == 0) update_derived(A, L, L->t_new);
  ++L->istep;
  if (l < A->nlev - 1) {
    for (int i = 1; i <= 2; ++i) amr_substep(A, l + 1, time + (double)(i - 1) * A->L[l + 1]->dt);
    average_down_to(A, l);
    update_derived(A, L, L->t_new);
  }
}

/* a fine level made from the coarse one (MakeNewLevelFromCoarse) */
static rc_level *amr_new_level(struct rc_amr *A, int l, const rc_case *c0, const int blo[3], const int bhi[3]) {
  rc_level *L = (rc_level *)calloc(1, sizeof(rc_level));
  L->c = *c0;
  if (L->c.dim == 2) { L->c.n[2] = 1; L->c.periodic[2] = 1; }
  const int fac = 1 << l;
  for (int d = 0; d < L->c.dim; ++d) { L->c.n[d] = c0->n[d] * fac; L->c.dx[d] = c0->dx[d] / (double)fac; }
  L->c.use_ib = (l == A->nlev - 1) ? c0->use_ib : 0;
  L->ng = 3;
  L->amr = A;
  L->lev = l;
  int lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    L->dom_lo[d] = 0; L->dom_hi[d] = L->c.n[d] - 1;
    L->vlo[d] = (d < L->c.dim) ? blo[d] : 0;
    L->vhi[d] = (d < L->c.dim) ? bhi[d] : 0;
    lo[d] = L->vlo[d]; hi[d] = L->vhi[d];
  }
  for (int d = 0; d < L->c.dim; ++d) { lo[d] -= L->ng; hi[d] += L->ng; }
  L->f_new = fab_alloc(lo, hi, L->c.q);
  L->f_old = fab_alloc(lo, hi, L->c.q);
  L->m_new = fab_alloc(lo, hi, L->c.nmac);
  L->m_old = fab_alloc(lo, hi, L->c.nmac);
  L->forcing = fab_alloc(lo, hi, 3);
  L->tau = 0.5 + 3.0 * L->c.nu / L->c.dx[0];   /* AmrCoreLBM.cpp:232-238 */
  L->dt = L->c.dx[0];
  L->t_new = 0.0;
  L->t_old = -1.0e200;
  /* FillCoarsePatchMacro over the level (InterpFromCoarseLevel), then
   * f = feq(macro) */
  const rc_level *Lc = A->L[l - 1];
  for (int k = L->vlo[2]; k <= L->vhi[2]; ++k)
    for (int j = L->vlo[1]; j <= L->vhi[1]; ++j)
      for (int i = L->vlo[0]; i <= L->vhi[0]; ++i) {
        const int iv[3] = {i, j, k};
        double v[RC_MAXC];
        cons_interp_cell(Lc, 0.0, L->c.dim, L->c.nmac, iv, v);
        for (int n = 0; n < L->c.nmac; ++n) AT(L->m_new, i, j, k, n) = v[n];
        const double r = fmax(v[0], 1.0e-12);
        const double u0 = v[1], v0 = v[2], w0 = v[3];
        const double u2 = u0 * u0 + v0 * v0 + w0 * w0;
        for (int q = 0; q < L->c.q; ++q) {
          const double cu = L->c.cx[q] * u0 + L->c.cy[q] * v0 + L->c.cz[q] * w0;
          const double feq = L->c.w[q] * r * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * u2);
          AT(L->f_new, i, j, k, q) = feq;
          AT(L->f_old, i, j, k, q) = feq;
        }
      }
  memcpy(L->m_old.p, L->m_new.p, sizeof(double) * (size_t)(L->m_new.ns * L->m_new.nc));
  return L;
}


/* ------------------------------------------------------------------ */
/* Regrid: ErrorEst (AmrCoreLBM.cpp:986-1180) with vorticity_tagging /  */
/* levelset_tagging (Source/Tagging.H:12-35) and the body refinement    */
/* region (AmrCoreLBM.cpp:56-104, user geometry                          */
/* Examples/Pitching_airfoil/IBMUserDefinedGeometry.H:25-57), then      */
/* RemakeLevel (AmrCoreLBM.cpp:835-914).  AMReX's clustering of the tags */
/* into boxes (Berger-Rigoutsos, grid_eff, max_grid_size) is replaced by */
/* one box per level: the bounding box of the buffered tags, aligned to */
/* the blocking factor and kept 3 coarse cells inside the coarser box    */
/* (the nesting the coarse-fine interpolation needs).  A level whose     */
/* tags are empty keeps its box.  The box choice is this build's own:   */
/* parity of the boxes with AMReX is unpinned; everything after the box  */
/* choice follows the reference.                                         */
/* ------------------------------------------------------------------ */

void rc_amr_set_regrid(rc_amr *A, const rc_regrid *p) {
  A->rg = *p;
  A->regrid_on = 1;
}

int rc_amr_regrids(const rc_amr *A) { return A->nregrid; }

/* regrid(lbase) at the hierarchy's current time, outside the step loop */
void rc_amr_regrid_now(rc_amr *A, int lbase) { amr_regrid(A, lbase, A->L[lbase]->t_new); }

int rc_amr_boxes(const rc_amr *A, int *lo, int *hi) {
  for (int l = 1; l < A->nlev; ++l)
    for (int d = 0; d < 3; ++d) { lo[3 * (l - 1) + d] = A->L[l]->vlo[d]; hi[3 * (l - 1) + d] = A->L[l]->vhi[d]; }
  return A->nlev - 1;
}

/* immersed_boundary_refinement_region (AmrCoreLBM.cpp:60-104) */
static int ib_refine_region(const rc_case *c, const rc_regrid *rg, double dx, double *lo, double *hi) {
  const double pad = fmax(6.0 * dx, 2.0);   /* ib_refine_padding :56-58 */
  if (c->geometry_type == RC_GEOM_BOX) {
    lo[0] = c->box_lo[0] - pad; hi[0] = c->box_hi[0] + pad;
    lo[1] = (c->box_lid_only > 0) ? (c->box_hi[1] - pad) : (c->box_lo[1] - pad);
    hi[1] = c->box_hi[1] + pad;
    lo[2] = c->box_lo[2] - pad; hi[2] = c->box_hi[2] + pad;
    return 1;
  }
  if (c->geometry_type == RC_GEOM_USER) {
    /* the swept section: leading edge at (x0, y0); the vertical reach is
     * that of the point farthest from the pivot (the trailing edge for the
     * leading-edge pivot of the reference hook) */
    if (c->dim != 2) return 0;
    const double chord = fmax(c->airfoil_chord, 1.0e-12);
    const double ht = 0.5 * c->airfoil_thickness * chord;
    const double arm = fmax(c->pivot_x, chord - c->pivot_x);
    const double yext = arm * sin(fabs(c->pitch_amplitude)) + ht;
    lo[0] = c->x0 - pad; hi[0] = c->x0 + chord + pad;
    lo[1] = c->y0 - yext - pad; hi[1] = c->y0 + yext + pad;
    lo[2] = c->z0 - pad; hi[2] = c->z0 + pad;
    return 1;
  }
  if (c->R > 0.0) {
    const double x0 = c->x0, y0 = c->y0, z0 = c->z0, R = c->R;
    if (rg->refine_upstream > 0.0 && rg->refine_downstream > 0.0 && rg->refine_cross > 0.0) {
      lo[0] = x0 - rg->refine_upstream * R - pad; hi[0] = x0 + rg->refine_downstream * R + pad;
      lo[1] = y0 - rg->refine_cross * R - pad; hi[1] = y0 + rg->refine_cross * R + pad;
      lo[2] = z0 - rg->refine_cross * R - pad; hi[2] = z0 + rg->refine_cross * R + pad;
    } else {
      lo[0] = x0 - R - pad; hi[0] = x0 + R + pad;
      lo[1] = y0 - R - pad; hi[1] = y0 + R + pad;
      lo[2] = z0 - R - pad; hi[2] = z0 + R + pad;
    }
    return 1;
  }
  return 0;
}

/* ErrorEst on level l at `time`: bounding box (level-l indices) of the
 * tagged cells grown by n_error_buf; returns the tag count. */
static int64_t amr_tag_bbox(struct rc_amr *A, int l, double time, int blo[3], int bhi[3]) {
  rc_level *L = A->L[l];
  const rc_case *c = &L->c;
  const rc_case *c0 = &A->c0;
  const int dim = c->dim;
  int lo[3], hi[3];
  for (int d = 0; d < 3; ++d) { lo[d] = L->vlo[d]; hi[d] = L->vhi[d]; }
  for (int d = 0; d < dim; ++d) { lo[d] -= 1; hi[d] += 1; }
  fab mp = fab_alloc(lo, hi, 4);      /* FillPatchMacro(lev, time, macro_tag, 0, 4), one ghost */
  if (l > 0) {
    amr_fill_macro(A, l, time, &mp, 4);
  } else {
    fab src = fab_alloc(L->m_new.lo, L->m_new.hi, 4);
    for (int k = L->vlo[2]; k <= L->vhi[2]; ++k)
      for (int j = L->vlo[1]; j <= L->vhi[1]; ++j)
        for (int i = L->vlo[0]; i <= L->vhi[0]; ++i)
          for (int n = 0; n < 4; ++n) AT(src, i, j, k, n) = macro_at_time(L, time, i, j, k, n);
    fill_patch(L, &mp, &src, 4, 0);
    fab_free(&src);
  }
  fab vort = fab_alloc(L->vlo, L->vhi, 1);
  double vmax = -HUGE_VAL;            /* vort_for_tag.max(0) */
  for (int k = L->vlo[2]; k <= L->vhi[2]; ++k)
    for (int j = L->vlo[1]; j <= L->vhi[1]; ++j)
      for (int i = L->vlo[0]; i <= L->vhi[0]; ++i) {
        double v, P;
        vis_para(c, &mp, i, j, k, &v, &P);
        AT(vort, i, j, k, 0) = v;
        vmax = fmax(vmax, v);
      }
  fab_free(&mp);
  const double thr = A->rg.threshold_ratio[l] * vmax;
  const double dxm = c->dx[0];
  double rlo[3] = {0, 0, 0}, rhi[3] = {0, 0, 0};
  const int region = c0->use_ib && ib_refine_region(c0, &A->rg, dxm, rlo, rhi);
  /* levelset_tagging applies for the cylinder / sphere marker geometry
   * (marker_levelset_ibm, AmrCoreLBM.cpp:1070-1073), phi = r - R */
  const int lset = c0->use_ib && c0->geometry_type == RC_GEOM_CYLINDER;
  const double band = (5.0 + 0.5) * dxm;
  const int buf = A->rg.n_error_buf;
  const double nrlo = A->rg.no_refine_xlo, nrhi = A->rg.no_refine_xhi;
  const double xlo_block = c->plo[0] + nrlo;
  const double xhi_block = c->plo[0] + c->dx[0] * (double)(c->n[0]) - nrhi;   /* probhi[0] - no_refine_xhi */
  int64_t count = 0;
  for (int d = 0; d < 3; ++d) { blo[d] = 1 << 30; bhi[d] = -(1 << 30); }
  for (int k = L->vlo[2]; k <= L->vhi[2]; ++k)
    for (int j = L->vlo[1]; j <= L->vhi[1]; ++j)
      for (int i = L->vlo[0]; i <= L->vhi[0]; ++i) {
        const double x = c->plo[0] + ((double)i + 0.5) * c->dx[0];
        const double y = c->plo[1] + ((double)j + 0.5) * c->dx[1];
        const double z = (dim == 3) ? c->plo[2] + ((double)k + 0.5) * c->dx[2] : 0.0;
        if ((nrlo > 0.0 || nrhi > 0.0) && (x <= xlo_block || x >= xhi_block)) continue;
        int tag = fabs(AT(vort, i, j, k, 0)) > thr;
        if (lset) {
          double r2 = (x - c0->x0) * (x - c0->x0) + (y - c0->y0) * (y - c0->y0);
          if (dim == 3) r2 += (z - c0->z0) * (z - c0->z0);
          if (sqrt(r2) - c0->R < band) tag = 1;
        }
        if (region && x >= rlo[0] && x <= rhi[0] && y >= rlo[1] && y <= rhi[1] && z >= rlo[2] && z <= rhi[2])
          tag = 1;
        if (!tag) continue;
        ++count;
        const int iv[3] = {i, j, k};
        for (int d = 0; d < dim; ++d) {
          if (iv[d] - buf < blo[d]) blo[d] = iv[d] - buf;
          if (iv[d] + buf > bhi[d]) bhi[d] = iv[d] + buf;
        }
      }
  for (int d = dim; d < 3; ++d) { blo[d] = 0; bhi[d] = 0; }
  fab_free(&vort);
  return count;
}

static int fdiv(int a, int b) { return (a >= 0) ? a / b : -((-a + b - 1) / b); }

/* One box for level l+1 (level-(l+1) indices) from a tag bounding box in
 * level-l indices: clipped to the parent box shrunk by 3 cells, aligned
 * outward to the blocking unit (blocking_factor / 2 level-l cells), then
 * inward where the alignment leaves the clip region.  0 if empty. */
static int box_from_tags(int dim, const int tlo[3], const int thi[3], const int plo[3], const int phi[3], int bu,
                         int out_lo[3], int out_hi[3]) {
  for (int d = 0; d < 3; ++d) { out_lo[d] = 0; out_hi[d] = 0; }
  for (int d = 0; d < dim; ++d) {
    const int nlo = plo[d] + 3, nhi = phi[d] - 3;
    const int lo = (tlo[d] > nlo) ? tlo[d] : nlo, hi = (thi[d] < nhi) ? thi[d] : nhi;
    if (lo > hi) return 0;
    int alo = fdiv(lo, bu) * bu, ahi = (fdiv(hi, bu) + 1) * bu - 1;
    if (alo < nlo) alo = -fdiv(-nlo, bu) * bu;
    if (ahi > nhi) ahi = fdiv(nhi + 1, bu) * bu - 1;
    if (alo > ahi) return 0;
    out_lo[d] = 2 * alo;
    out_hi[d] = 2 * ahi + 1;
  }
  return 1;
}

/* RemakeLevel (AmrCoreLBM.cpp:835-914): macros by FillPatchMacro (the old
 * level where it was valid, coarse interpolation elsewhere, all nmac
 * components, ghosts included), populations = feq of those macros,
 * macro_old = macro_new; the finest level gets a new marker backend
 * (:906-909), the force history carries on; then UpdateDerivedMacroFields. */
static void amr_remake(struct rc_amr *A, int l, double time, const int blo[3], const int bhi[3]) {
  rc_level *O = A->L[l];
  rc_level *N = (rc_level *)calloc(1, sizeof(rc_level));
  N->c = O->c;
  N->ng = O->ng;
  N->amr = A;
  N->lev = l;
  int lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    N->dom_lo[d] = O->dom_lo[d]; N->dom_hi[d] = O->dom_hi[d];
    N->vlo[d] = blo[d]; N->vhi[d] = bhi[d];
    lo[d] = blo[d]; hi[d] = bhi[d];
  }
  for (int d = 0; d < N->c.dim; ++d) { lo[d] -= N->ng; hi[d] += N->ng; }
  N->f_new = fab_alloc(lo, hi, N->c.q);
  N->f_old = fab_alloc(lo, hi, N->c.q);
  N->m_new = fab_alloc(lo, hi, N->c.nmac);
  N->m_old = fab_alloc(lo, hi, N->c.nmac);
  N->forcing = fab_alloc(lo, hi, 3);
  N->tau = O->tau; N->dt = O->dt;
  N->t_new = O->t_new; N->t_old = O->t_old;
  N->istep = O->istep;
  N->have_prev = O->have_prev;
  for (int d = 0; d < 3; ++d) N->prev_body[d] = O->prev_body[d];
  amr_fill_macro(A, l, time, &N->m_new, N->c.nmac);   /* A->L[l] is still the old level */
  const rc_case *c = &N->c;
  for (int64_t s = 0; s < N->f_new.ns; ++s) {
    const double r = fmax(N->m_new.p[s], 1.0e-12);
    const double u0 = N->m_new.p[s + N->m_new.ns], v0 = N->m_new.p[s + 2 * N->m_new.ns];
    const double w0 = N->m_new.p[s + 3 * N->m_new.ns];
    const double u2 = u0 * u0 + v0 * v0 + w0 * w0;
    for (int q = 0; q < c->q; ++q) {
      const double cu = c->cx[q] * u0 + c->cy[q] * v0 + c->cz[q] * w0;
      const double feq = c->w[q] * r * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * u2);
      N->f_new.p[s + q * N->f_new.ns] = feq;
      N->f_old.p[s + q * N->f_old.ns] = feq;
    }
  }
  memcpy(N->m_old.p, N->m_new.p, sizeof(double) * (size_t)(N->m_new.ns * N->m_new.nc));
  A->L[l] = N;
  if (l == A->nlev - 1 && N->c.use_ib) {
    ib_build(N);
    ib_update_markers(N, time);
  }
  update_derived(A, N, time);
  rc_destroy(O);
}

/* regrid(lbase, time) with the one-box clustering above: the boxes of
 * levels lbase+1 .. finest are made finest first, each from the tags of
 * its parent level (old data) and, for a parent that is itself re-boxed,
 * from the footprint of the new child box grown by 3 cells; changed
 * levels are remade coarse to fine (AmrCore::regrid). */
static void amr_regrid(struct rc_amr *A, int lbase, double time) {
  const int dim = A->L[0]->c.dim;
  const int bu = (A->rg.blocking_factor / 2 > 0) ? A->rg.blocking_factor / 2 : 1;
  int tlo[RC_MAXLEV][3], thi[RC_MAXLEV][3];
  int64_t tcnt[RC_MAXLEV];
  for (int l = lbase; l < A->nlev - 1; ++l) tcnt[l] = amr_tag_bbox(A, l, time, tlo[l], thi[l]);
  int nlo[RC_MAXLEV][3], nhi[RC_MAXLEV][3];
  for (int l = 1; l < A->nlev; ++l)
    for (int d = 0; d < 3; ++d) { nlo[l][d] = A->L[l]->vlo[d]; nhi[l][d] = A->L[l]->vhi[d]; }
  for (int lf = A->nlev - 1; lf > lbase; --lf) {
    const int lc = lf - 1;
    int qlo[3], qhi[3];
    int have = tcnt[lc] > 0;
    for (int d = 0; d < 3; ++d) { qlo[d] = tlo[lc][d]; qhi[d] = thi[lc][d]; }
    if (lf < A->nlev - 1) {   /* contain the (new) child footprint grown by 3 */
      for (int d = 0; d < dim; ++d) {
        const int flo = fdiv(nlo[lf + 1][d], 2) - 3, fhi = fdiv(nhi[lf + 1][d], 2) + 3;
        const int clo = fdiv(flo, 2), chi = fdiv(fhi, 2);
        if (!have || clo < qlo[d]) qlo[d] = clo;
        if (!have || chi > qhi[d]) qhi[d] = chi;
      }
      if (!have) for (int d = 0; d < dim; ++d) {   /* keep the old box as well */
        const int olo = fdiv(nlo[lf][d], 2), ohi = fdiv(nhi[lf][d], 2);
        if (olo < qlo[d]) qlo[d] = olo;
        if (ohi > qhi[d]) qhi[d] = ohi;
      }
      have = 1;
    }
    if (!have) continue;      /* no tags: the level keeps its box */
    int blo[3], bhi[3];
    const rc_level *P = A->L[lc];
    const int *plo = (lc > lbase) ? nlo[lc] : P->vlo, *phi = (lc > lbase) ? nhi[lc] : P->vhi;
    if (box_from_tags(dim, qlo, qhi, plo, phi, bu, blo, bhi))
      for (int d = 0; d < 3; ++d) { nlo[lf][d] = blo[d]; nhi[lf][d] = bhi[d]; }
  }
  /* a parent box chosen after its child: re-clip the child into it */
  for (int lf = lbase + 2; lf < A->nlev; ++lf) {
    int clo[3], chi[3], blo[3], bhi[3];
    for (int d = 0; d < 3; ++d) { clo[d] = fdiv(nlo[lf][d], 2); chi[d] = fdiv(nhi[lf][d], 2); }
    if (box_from_tags(dim, clo, chi, nlo[lf - 1], nhi[lf - 1], bu, blo, bhi))
      for (int d = 0; d < 3; ++d) { nlo[lf][d] = blo[d]; nhi[lf][d] = bhi[d]; }
  }
  int changed = 0;
  for (int l = lbase + 1; l < A->nlev; ++l) {
    int same = 1;
    for (int d = 0; d < 3; ++d)
      if (nlo[l][d] != A->L[l]->vlo[d] || nhi[l][d] != A->L[l]->vhi[d]) same = 0;
    if (same) continue;
    amr_remake(A, l, time, nlo[l], nhi[l]);
    changed = 1;
  }
  A->nregrid += changed;
}

rc_amr *rc_amr_create(const rc_case *c0, int nlev, const int *box_lo, const int *box_hi, int regrid_int,
                      const double *init_macro) {
  if (nlev < 1 || nlev > RC_MAXLEV || regrid_int < 1 || c0->nmac > RC_MAXC) return NULL;
  rc_amr *A = (rc_amr *)calloc(1, sizeof(rc_amr));
  A->nlev = nlev;
  A->regrid_int = regrid_int;
  A->c0 = *c0;
  rc_case cc = *c0;
  if (nlev > 1) cc.use_ib = 0;
  A->L[0] = rc_create(&cc, init_macro);
  A->L[0]->amr = A;
  A->L[0]->lev = 0;
  /* MakeNewLevelFromScratch: initdata, UpdateDerivedMacroFields, then
   * macro_old = macro_new (AmrCoreLBM.cpp:971-983) */
  update_derived(A, A->L[0], 0.0);
  memcpy(A->L[0]->m_old.p, A->L[0]->m_new.p, sizeof(double) * (size_t)(A->L[0]->m_new.ns * A->L[0]->m_new.nc));
  for (int l = 1; l < nlev; ++l) {
    A->L[l] = amr_new_level(A, l, c0, box_lo + 3 * (l - 1), box_hi + 3 * (l - 1));
    update_derived(A, A->L[l], 0.0);
  }
  rc_level *Lf = A->L[nlev - 1];
  if (nlev > 1 && Lf->c.use_ib) {
    ib_build(Lf);
    ib_update_markers(Lf, 0.0);
  }
  return A;
}

void rc_amr_destroy(rc_amr *A) {
  if (!A) return;
  for (int l = 0; l < A->nlev; ++l) rc_destroy(A->L[l]);
  free(A);
}

/* one coarse step of Evolve (AmrCoreLBM.cpp:366-420) */
void rc_amr_step(rc_amr *A) {
  double cur = A->L[0]->t_new;
  amr_substep(A, 0, cur);
  cur += A->L[0]->dt;
  ++A->istep;
  rc_level *Lf = A->L[A->nlev - 1];
  if (Lf->c.use_ib) compute_ib_force(Lf, cur, A->istep);
  for (int l = 0; l < A->nlev; ++l) A->L[l]->t_new = cur;
}

void rc_amr_advance(rc_amr *A, int nsteps) {
  for (int s = 0; s < nsteps; ++s) rc_amr_step(A);
}

rc_level *rc_amr_level(rc_amr *A, int l) { return (l >= 0 && l < A->nlev) ? A->L[l] : NULL; }

/* macro_new comps 0..nmac-1 on the valid box */
int rc_get_macro_all(const rc_level *L, double *out) { return get_valid(L, &L->m_new, L->c.nmac, out); }

/* public entry points of the interpolation pieces, for the pins */
void rc_cons_interp_test(const double *crse, const int *clo, const int *chi, int dim, int ncomp,
                         const int *fiv, double *out) {
  rc_level tmp;
  memset(&tmp, 0, sizeof(tmp));
  int lo[3], hi[3];
  for (int d = 0; d < 3; ++d) { lo[d] = clo[d]; hi[d] = chi[d]; tmp.vlo[d] = clo[d]; tmp.vhi[d] = chi[d]; }
  tmp.m_new = fab_alloc(lo, hi, ncomp);
  memcpy(tmp.m_new.p, crse, sizeof(double) * (size_t)(tmp.m_new.ns * ncomp));
  tmp.t_old = -1.0e200;
  tmp.t_new = 0.0;
  cons_interp_cell(&tmp, 0.0, dim, ncomp, fiv, out);
  fab_free(&tmp.m_new);
}

