// Immersed-boundary kernels of the resident level (FP64 throughout).
//
// They follow IBMarkerDF::update_forcing (Source/IBM/IBMarkerDF.cpp:1522-1917)
// stage by stage, restricted to a band box around the markers (the
// reference allocates and sweeps the whole level every step,
// IBMarkerDF.cpp:1532-1563, 1878-1912; outside the band every term is zero).
//
//   k_ib_begin     EulerVel = u(macro_old), EulerForceTotal = 0     (:1532-1566)
//   k_ib_weights   marker cell, 5-point delta weights per dimension (:70-99)
//   k_ib_interp    U_m = sum delta * EulerVel * dV                  (:611-739)
//   k_ib_force_*   explicit / explicit-diag / IVC direct forcing    (:741-873, :1144-1328)
//   k_ib_spread    EulerForceIter = sum_m delta * F_m * dv          (:1330-1512)
//   k_ib_sums      sum_m F_m dv (marker order), moments r x F dv    (:1779-1819, :1500-1503)
//   k_ib_correct   EulerVel += dt * Iter, Total += Iter             (:1514-1520, :1878-1879)
//   k_ib_export    F = dt * rho * Total                             (:1889-1912)
//
// Spreading is a deterministic gather: every band cell walks a precomputed
// list of (marker, stencil point) contributions in marker order, so the
// per-cell sums are evaluated in the same order as the reference's serial
// particle loop and there are no atomics.  Contributions that land in a
// periodic ghost image are summed per image and added afterwards, as
// SumBoundary does (:1511).
#pragma once

#include "ilbm_lbm_kernels.h"

namespace ilbm {

struct IbArgs {
  int dim;
  int n[3];
  int periodic[3];
  double plo[3], dx[3];
  BcSpec bc;
  int band_lo[3], band_n[3];
  int64_t band_cells;
  int nm;
  double dt, dv;
  int kernel_type;      // 0: 4-point, 1: 3-point
  int renorm_interp;    // renormalize_delta && coupling != IVC
  int renorm_spread;    // renormalize_delta
  int coupling;         // 0 explicit, 1 IVC, 2 explicit-diag
  double relax, diag_eps;
  int geometry;         // 0 rigid, 1 box, 2 user
  double x0, y0, z0;
  double ub[3], om[3];
  double box_yhi, wall_tol, lid_u[3];
  // marker state
  const double* xm;     // [3][nm] positions
  const double* ut;     // [3][nm] target velocity (user geometry)
  double* U;            // [3][nm]
  double* F;            // [3][nm]
  double* Mm;           // [3][nm] moments
  int* base;            // [3][nm]
  double* wts;          // [3][5][nm]
  double* scale;        // [2][nm] interp / spread renormalisation
  // band fields
  const double* Mband;  // [4][band]
  double* EV;           // [3][band]
  double* EFI;          // [3][band]
  double* EFT;          // [3][band]
  double* Fband;        // [3][band]
  // spreading lists
  const int* rowptr;    // [band+1]
  const int* ent_m;
  const int* ent_code;
  double* ent_w;        // per-entry delta weight
  // IVC
  const double* Ainv;   // [nm][nm]
  double* rhs;          // [3][nm]
  // reductions
  double* sums;         // [0..2] last marker force sum, [3..5] torque
  // single-iteration fast path (loop_ns == 1 or IVC): EulerVel is the
  // macro velocity and EulerForceTotal is the one iteration's force
  int ev_from_macro;
  int fuse_explicit;    // explicit / diag force evaluated by k_ib_interp
  // interpolation and spreading sums in a fixed wave-level reduction tree
  // (levels with reduced storage) instead of the reference's serial order
  int tree;
  int first_iter;
  const int* ne_cells;  // band cells that receive spreading entries
  const int* ne_e;      // [2][n_ne]: each such cell's first and last + 1 entry
  int n_ne;
  int nloop;            // direct-forcing iterations per step
  int use_ivc;          // IVC solve (operator factorized) instead of explicit forcing
  int ioff[3];          // global cell index of local cell 0 (refined levels)
  int dlo[3], dhi[3];   // the level's domain in local cell indices
  int* chunk_box;       // moving body: [6][nchunks] min / max marker cell per dim of each marker chunk
};

constexpr int kMarkerChunk = 32;

__device__ inline double delta_fn(double xf, double xp, double h, int type) {
  const double rr = fabs(xf - xp) / h;
  double value = 0.0;
  if (type == 0) {
    if (rr < 1.0) value = 0.125 * (3.0 - 2.0 * rr + sqrt(1.0 + 4.0 * rr - 4.0 * rr * rr));
    else if (rr < 2.0) value = 0.125 * (5.0 - 2.0 * rr - sqrt(-7.0 + 12.0 * rr - 4.0 * rr * rr));
  } else {
    if (rr < 0.5) value = (1.0 + sqrt(1.0 - 3.0 * rr * rr)) / 3.0;
    else if (rr < 1.5) value = (5.0 - 3.0 * rr - sqrt(-2.0 + 6.0 * rr - 3.0 * rr * rr)) / 6.0;
  }
  return value / h;
}

__device__ inline bool band_index(const IbArgs& a, const int c[3], int64_t* b) {
  const int bi = c[0] - a.band_lo[0], bj = c[1] - a.band_lo[1], bk = c[2] - a.band_lo[2];
  if ((unsigned)bi >= (unsigned)a.band_n[0] || (unsigned)bj >= (unsigned)a.band_n[1] ||
      (unsigned)bk >= (unsigned)a.band_n[2])
    return false;
  *b = ((int64_t)bk * a.band_n[1] + bj) * a.band_n[0] + bi;
  return true;
}

// FillPatchMacro value (macroBcFill, bc_fill.H:263-301, plus FilccCell and
// periodic images) of macro component comp at any cell, from macro_old.
__device__ double macro_fill(const IbArgs& a, const int iv[3], int comp) {
  int c[3] = {iv[0], iv[1], iv[2]};
  for (int iter = 0; iter < 8; ++iter) {
    int lo_hit[3] = {0, 0, 0}, hi_hit[3] = {0, 0, 0};
    bool any_out = false, nonper_out = false;
    for (int d = 0; d < a.dim; ++d) {
      if (c[d] < 0) { lo_hit[d] = 1; any_out = true; if (!a.periodic[d]) nonper_out = true; }
      else if (c[d] > a.n[d] - 1) { hi_hit[d] = 1; any_out = true; if (!a.periodic[d]) nonper_out = true; }
    }
    if (!any_out) {
      int64_t b;
      if (!band_index(a, c, &b)) return 0.0;
      return a.Mband[comp * a.band_cells + b];
    }
    if (!nonper_out) {
      for (int d = 0; d < a.dim; ++d) {
        if (c[d] < 0) c[d] += a.n[d];
        else if (c[d] > a.n[d] - 1) c[d] -= a.n[d];
      }
      continue;
    }
    int bd = -1, lo_side = 1, kind = 0;
    for (int d = 0; d < a.dim && kind == 0; ++d) {
      if (lo_hit[d] && a.bc.lo[d] == BC_USER_2) { kind = 2; bd = d; lo_side = 1; }
      else if (hi_hit[d] && a.bc.hi[d] == BC_USER_2) { kind = 2; bd = d; lo_side = 0; }
    }
    for (int d = 0; d < a.dim && kind == 0; ++d) {
      if (lo_hit[d] && a.bc.lo[d] == BC_USER_1) { kind = 1; bd = d; lo_side = 1; }
      else if (hi_hit[d] && a.bc.hi[d] == BC_USER_1) { kind = 1; bd = d; lo_side = 0; }
    }
    for (int d = 0; d < a.dim && kind == 0; ++d) {
      if (lo_hit[d] && a.bc.lo[d] == BC_EXT_DIR) { kind = 1; bd = d; lo_side = 1; }
      else if (hi_hit[d] && a.bc.hi[d] == BC_EXT_DIR) { kind = 1; bd = d; lo_side = 0; }
    }
    if (kind != 0) {
      const double* bv = lo_side ? a.bc.lo_val[bd] : a.bc.hi_val[bd];
      if (kind == 1) return bv[comp];
      int mir[3] = {c[0], c[1], c[2]};
      for (int d = 0; d < a.dim; ++d) {
        if (mir[d] < 0) mir[d] = 0;
        else if (mir[d] > a.n[d] - 1) mir[d] = a.n[d] - 1;
      }
      int64_t b;
      if (!band_index(a, mir, &b)) return 0.0;
      const double m = a.Mband[comp * a.band_cells + b];
      if (comp == 0) return fmax(1.0e-20, 2.0 * bv[0] - m);
      return 2.0 * bv[comp] - m;
    }
    int dsel = -1;
    for (int d = 0; d < a.dim; ++d)
      if ((lo_hit[d] && a.bc.lo[d] == BC_FOEXTRAP) || (hi_hit[d] && a.bc.hi[d] == BC_FOEXTRAP)) dsel = d;
    if (dsel < 0) return 0.0;
    c[dsel] = lo_hit[dsel] ? 0 : a.n[dsel] - 1;
  }
  return 0.0;
}

// EulerVel at any cell of the 2-ghost fab: corrected values inside the
// domain and at periodic images, FillPatch values at physical ghosts.
__device__ inline double ev_value(const IbArgs& a, const int p[3], int comp) {
  int c[3] = {p[0], p[1], p[2]};
  bool phys = false;
  for (int d = 0; d < a.dim; ++d) {
    if (c[d] < 0 || c[d] > a.n[d] - 1) {
      if (a.periodic[d]) c[d] = (c[d] < 0) ? c[d] + a.n[d] : c[d] - a.n[d];
      else phys = true;
    }
  }
  if (phys) return macro_fill(a, p, comp + 1);
  int64_t b;
  if (!band_index(a, c, &b)) return 0.0;
  if (a.ev_from_macro) return (comp == 2 && a.dim != 3) ? 0.0 : a.Mband[(comp + 1) * a.band_cells + b];
  return a.EV[comp * a.band_cells + b];
}

// The three EulerVel components at stencil point p: read from the band
// directly when p lies inside the domain and the band (the common case, all
// loads issued together), else through ev_value.  Same values either way.
__device__ inline void ev_values3(const IbArgs& a, const int p[3], double v[3]) {
  bool inside = true;
  for (int d = 0; d < a.dim; ++d)
    if (p[d] < 0 || p[d] > a.n[d] - 1) inside = false;
  int64_t b;
  if (inside && band_index(a, p, &b)) {
    if (a.ev_from_macro) {
      v[0] = a.Mband[a.band_cells + b];
      v[1] = a.Mband[2 * a.band_cells + b];
      v[2] = (a.dim == 3) ? a.Mband[3 * a.band_cells + b] : 0.0;
    } else {
      v[0] = a.EV[b];
      v[1] = a.EV[a.band_cells + b];
      v[2] = a.EV[2 * a.band_cells + b];
    }
    return;
  }
  for (int c = 0; c < 3; ++c) v[c] = ev_value(a, p, c);
}

__device__ __forceinline__ void ib_begin_body(const IbArgs& a, int64_t b) {
  if (b >= a.band_cells) return;
  a.EV[b] = a.Mband[a.band_cells + b];
  a.EV[a.band_cells + b] = a.Mband[2 * a.band_cells + b];
  a.EV[2 * a.band_cells + b] = (a.dim == 3) ? a.Mband[3 * a.band_cells + b] : 0.0;
  for (int c = 0; c < 3; ++c) a.EFT[c * a.band_cells + b] = 0.0;
}

__global__ void k_ib_begin(IbArgs a) { ib_begin_body(a, (int64_t)blockIdx.x * blockDim.x + threadIdx.x); }

// cell index and 5-point weights of every marker, and the partition sums
// that the optional renormalisation uses (:640-690, :1358-1403)
__device__ inline void marker_weights(const IbArgs& a, int m, const double pos3[3], int b[3]) {
  const int nm = a.nm;
  double t[3][5];
  b[0] = b[1] = b[2] = 0;
  for (int d = 0; d < 3; ++d) {
    if (d < a.dim) {
      const double pos = pos3[d];
      b[d] = (int)floor((pos - a.plo[d]) / a.dx[d]) - a.ioff[d];
      for (int s = 0; s < 5; ++s) {
        const double xi = a.plo[d] + ((double)(b[d] + a.ioff[d] + s - 2) + 0.5) * a.dx[d];
        t[d][s] = delta_fn(xi, pos, a.dx[d], a.kernel_type);
      }
    } else {
      for (int s = 0; s < 5; ++s) t[d][s] = 1.0;
    }
    a.base[d * nm + m] = b[d];
    for (int s = 0; s < 5; ++s) a.wts[(d * 5 + s) * nm + m] = t[d][s];
  }
  const double dvol = (a.dim == 3) ? a.dx[0] * a.dx[1] * a.dx[2] : a.dx[0] * a.dx[1];
  double part = 0.0;
  for (int ii = 0; ii < 5; ++ii)
    for (int jj = 0; jj < 5; ++jj) {
      if (a.dim == 3) {
        for (int kk = 0; kk < 5; ++kk) {
          const int ip = b[0] + ii - 2, jp = b[1] + jj - 2, kp = b[2] + kk - 2;
          if (ip < -2 || ip > a.n[0] + 1 || jp < -2 || jp > a.n[1] + 1 || kp < -2 || kp > a.n[2] + 1) continue;
          part += t[0][ii] * t[1][jj] * t[2][kk] * dvol;
        }
      } else {
        const int ip = b[0] + ii - 2, jp = b[1] + jj - 2;
        if (ip < -2 || ip > a.n[0] + 1 || jp < -2 || jp > a.n[1] + 1) continue;
        part += t[0][ii] * t[1][jj] * dvol;
      }
    }
  a.scale[m] = part;                                              // interpolation divisor
  a.scale[nm + m] = (a.renorm_spread && part > 0.0) ? 1.0 / part : 1.0;  // spreading factor
}

__global__ void k_ib_weights(IbArgs a) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= a.nm) return;
  double pos[3] = {0.0, 0.0, 0.0};
  for (int d = 0; d < a.dim; ++d) pos[d] = a.xm[d * a.nm + m];
  int b[3];
  marker_weights(a, m, pos, b);
}

__device__ inline void target_velocity(const IbArgs& a, int m, double* ux, double* uy, double* uz) {
  const int nm = a.nm;
  if (a.geometry == 2) {
    *ux = a.ut[m]; *uy = a.ut[nm + m]; *uz = a.ut[2 * nm + m];
    return;
  }
  const double x = a.xm[m], y = a.xm[nm + m], z = (a.dim == 3) ? a.xm[2 * nm + m] : 0.0;
  if (a.geometry == 1) {  // box_target_velocity (:147-161)
    *ux = *uy = *uz = 0.0;
    if (fabs(y - a.box_yhi) <= a.wall_tol) { *ux = a.lid_u[0]; *uy = a.lid_u[1]; *uz = a.lid_u[2]; }
    return;
  }
  const double z0 = (a.dim == 3) ? a.z0 : 0.0;  // rigid_target_velocity (:163-182)
  const double rx = x - a.x0, ry = y - a.y0, rz = z - z0;
  const double cx = a.om[1] * rz - a.om[2] * ry;
  const double cy = a.om[2] * rx - a.om[0] * rz;
  const double cz = a.om[0] * ry - a.om[1] * rx;
  *ux = a.ub[0] + cx; *uy = a.ub[1] + cy; *uz = a.ub[2] + cz;
}

// compute_lagrangian_force_explicit (:741-873): F += relax*(U_t - U)/denom,
// with the target velocity (Ux, Uy, Uz) of marker m given
__device__ inline void explicit_force_tv(const IbArgs& a, int m, double Ux, double Uy, double Uz, double Ux_m,
                                         double Uy_m, double Uz_m, double& F0, double& F1, double& F2) {
  const int nm = a.nm;
  double denom = a.dt;
  if (a.coupling == 2) {
    const double dV = (a.dim == 3) ? a.dx[0] * a.dx[1] * a.dx[2] : a.dx[0] * a.dx[1];
    const int b0 = a.base[m], b1 = a.base[nm + m], b2 = a.base[2 * nm + m];
    double s2 = 0.0;
    for (int ii = 0; ii < 5; ++ii) {
      const int ip = b0 + ii - 2;
      if (ip < a.dlo[0] || ip > a.dhi[0]) continue;
      const double tU = a.wts[(0 * 5 + ii) * nm + m];
      for (int jj = 0; jj < 5; ++jj) {
        const int jp = b1 + jj - 2;
        if (jp < a.dlo[1] || jp > a.dhi[1]) continue;
        const double tV = a.wts[(1 * 5 + jj) * nm + m];
        if (a.dim == 3) {
          for (int kk = 0; kk < 5; ++kk) {
            const int kp = b2 + kk - 2;
            if (kp < a.dlo[2] || kp > a.dhi[2]) continue;
            const double tW = a.wts[(2 * 5 + kk) * nm + m];
            const double dval = tU * tV * tW;
            s2 += dval * dval * dV;
          }
        } else {
          const double dval = tU * tV;
          s2 += dval * dval * dV;
        }
      }
    }
    const double mkk = a.dv * s2;
    denom = a.dt * fmax(mkk, a.diag_eps);
  }
  F0 += a.relax * (Ux - Ux_m) / denom;
  F1 += a.relax * (Uy - Uy_m) / denom;
  F2 += a.relax * (Uz - Uz_m) / denom;
}

__device__ inline void explicit_force(const IbArgs& a, int m, double Ux_m, double Uy_m, double Uz_m,
                                      double& F0, double& F1, double& F2) {
  double Ux, Uy, Uz;
  target_velocity(a, m, &Ux, &Uy, &Uz);
  explicit_force_tv(a, m, Ux, Uy, Uz, Ux_m, Uy_m, Uz_m, F0, F1, F2);
}

// One wavefront per marker: the lanes evaluate the stencil products in
// parallel, then four lanes accumulate them in the reference's (ii, jj, kk)
// order, so the sums are bit-identical to the serial loop (:642-681).
// The loads that do not depend on the marker's cell -- the delta weights of
// each lane's stencil points and lane 0's target velocity -- are issued
// before the cell index arrives, so the chain of dependent memory accesses
// is cell index -> band velocities -> marker U / F.
constexpr int kInterpWaves = 4;
// Interpolation in a tree order (levels with reduced storage): G lanes per
// marker, 256 / G markers per block -- a wavefront per marker in 3D (125
// points, the shortest chain on the step's critical path), 16 lanes in 2D
// (25 points).  Lane s takes the stencil points s, s + G, ..., adds its
// products in that order, and the marker's lanes add their sums in a fixed
// butterfly: deterministic, and equal to the serial sum to rounding.
__host__ __device__ constexpr int interp_group(int dim) { return dim == 3 ? 64 : 16; }
// blocks of a k_ib_interp launch over nm markers
__host__ __device__ inline int interp_blocks(int nm, bool tree, int dim) {
  const int per = tree ? 64 * kInterpWaves / interp_group(dim) : kInterpWaves;
  return (nm + per - 1) / per;
}
template <int DIM>
__device__ __forceinline__ void interp_group_body(const IbArgs& a, int blk) {
  constexpr int kInterpGroup = interp_group(DIM);
  constexpr int kInterpGroupMarkers = 64 * kInterpWaves / kInterpGroup;
  constexpr int npts = (DIM == 3) ? 125 : 25;
  constexpr int kP = (npts + kInterpGroup - 1) / kInterpGroup;   // points per lane
  const int sub = (int)threadIdx.x & (kInterpGroup - 1);
  const int m = blk * kInterpGroupMarkers + (int)threadIdx.x / kInterpGroup;
  const bool act = m < a.nm;
  const int nm = a.nm;
  const int mc = act ? m : 0;
  const double d = (DIM == 3) ? a.dx[0] * a.dx[1] * a.dx[2] : a.dx[0] * a.dx[1];
  const int b0 = a.base[mc], b1 = a.base[nm + mc], b2 = (DIM == 3) ? a.base[2 * nm + mc] : 0;
  double tvx = 0.0, tvy = 0.0, tvz = 0.0;
  if (act && sub == 0 && a.fuse_explicit) target_velocity(a, m, &tvx, &tvy, &tvz);
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  // A marker whose whole stencil lies inside the domain and the band reads
  // its EulerVel values straight from the band, every load of the lane's
  // points issued first; any other marker goes point by point through
  // ev_value.  Same values, same order of the additions.
  bool inner = act;
  for (int e = 0; e < DIM; ++e) {
    const int be = (e == 0) ? b0 : (e == 1) ? b1 : b2;
    inner = inner && be - 2 >= 0 && be + 2 <= a.n[e] - 1 && be - 2 >= a.band_lo[e] &&
            be + 2 <= a.band_lo[e] + a.band_n[e] - 1;
  }
  const double* src = a.ev_from_macro ? a.Mband + a.band_cells : a.EV;
  if (inner) {
    double dval[kP], ev[kP][3];
#pragma unroll
    for (int r = 0; r < kP; ++r) {
      const int t = sub + kInterpGroup * r;
      const int tc = t < npts ? t : 0;
      int ii, jj, kk;
      if (DIM == 3) { ii = tc / 25; jj = (tc / 5) % 5; kk = tc % 5; }
      else { ii = tc / 5; jj = tc % 5; kk = 0; }
      const double tU = a.wts[(0 * 5 + ii) * nm + m];
      const double tV = a.wts[(1 * 5 + jj) * nm + m];
      dval[r] = (DIM == 3) ? tU * tV * a.wts[(2 * 5 + kk) * nm + m] : tU * tV;
      const int64_t bc = ((int64_t)((DIM == 3) ? b2 + kk - 2 - a.band_lo[2] : 0) * a.band_n[1] +
                          (b1 + jj - 2 - a.band_lo[1])) * a.band_n[0] + (b0 + ii - 2 - a.band_lo[0]);
      ev[r][0] = src[bc];
      ev[r][1] = src[a.band_cells + bc];
      ev[r][2] = (DIM == 3) ? src[2 * a.band_cells + bc] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < kP; ++r) {
      if (sub + kInterpGroup * r >= npts) continue;
      s0 += dval[r] * d;
      s1 += dval[r] * ev[r][0] * d;
      s2 += dval[r] * ev[r][1] * d;
      if (DIM == 3) s3 += dval[r] * ev[r][2] * d;
    }
  } else if (act) {
#pragma unroll 1
    for (int t = sub; t < npts; t += kInterpGroup) {
      int ii, jj, kk;
      if (DIM == 3) { ii = t / 25; jj = (t / 5) % 5; kk = t % 5; }
      else { ii = t / 5; jj = t % 5; kk = 0; }
      const int p[3] = {b0 + ii - 2, b1 + jj - 2, (DIM == 3) ? b2 + kk - 2 : 0};
      bool o = !(p[0] < -2 || p[0] > a.n[0] + 1 || p[1] < -2 || p[1] > a.n[1] + 1);
      if (DIM == 3 && (p[2] < -2 || p[2] > a.n[2] + 1)) o = false;
      if (!o) continue;
      const double tU = a.wts[(0 * 5 + ii) * nm + m];
      const double tV = a.wts[(1 * 5 + jj) * nm + m];
      const double dv = (DIM == 3) ? tU * tV * a.wts[(2 * 5 + kk) * nm + m] : tU * tV;
      double ev[3];
      ev_values3(a, p, ev);
      s0 += dv * d;
      s1 += dv * ev[0] * d;
      s2 += dv * ev[1] * d;
      if (DIM == 3) s3 += dv * ev[2] * d;
    }
  }
  // a + b == b + a: every lane of the marker ends with the same sums
#pragma unroll
  for (int off = kInterpGroup / 2; off > 0; off >>= 1) {
    s0 += __shfl_xor(s0, off, kInterpGroup);
    s1 += __shfl_xor(s1, off, kInterpGroup);
    s2 += __shfl_xor(s2, off, kInterpGroup);
    s3 += __shfl_xor(s3, off, kInterpGroup);
  }
  if (!act || sub != 0) return;
  double u = s1, v = s2, w = (DIM == 3) ? s3 : 0.0;
  if (a.renorm_interp && s0 > 0.0) { u /= s0; v /= s0; w /= s0; }
  a.U[m] = u; a.U[nm + m] = v; a.U[2 * nm + m] = w;
  double F0 = 0.0, F1 = 0.0, F2 = 0.0;
  if (a.fuse_explicit) explicit_force_tv(a, m, tvx, tvy, tvz, u, v, w, F0, F1, F2);
  a.F[m] = F0; a.F[nm + m] = F1; a.F[2 * nm + m] = F2;
}

__device__ __forceinline__ void interp_body(const IbArgs& a, int blk) {
  if (a.tree) {
    if (a.dim == 3) interp_group_body<3>(a, blk);
    else interp_group_body<2>(a, blk);
    return;
  }
  __shared__ double prod[kInterpWaves][4][128];
  __shared__ unsigned char valid[kInterpWaves][128];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int m = blk * kInterpWaves + wv;
  const bool active = m < a.nm;
  const int nm = a.nm;
  const double d = (a.dim == 3) ? a.dx[0] * a.dx[1] * a.dx[2] : a.dx[0] * a.dx[1];
  const int npts = (a.dim == 3) ? 125 : 25;
  // delta weights of the lane's stencil points t = lane and lane + 64
  double dval[2] = {0.0, 0.0};
  for (int r = 0; r < 2; ++r) {
    const int t = lane + 64 * r;
    if (!active || t >= npts) continue;
    int ii, jj, kk;
    if (a.dim == 3) { ii = t / 25; jj = (t / 5) % 5; kk = t % 5; }
    else { ii = t / 5; jj = t % 5; kk = 0; }
    const double tU = a.wts[(0 * 5 + ii) * nm + m];
    const double tV = a.wts[(1 * 5 + jj) * nm + m];
    dval[r] = (a.dim == 3) ? tU * tV * a.wts[(2 * 5 + kk) * nm + m] : tU * tV;
  }
  double tvx = 0.0, tvy = 0.0, tvz = 0.0;
  if (active && lane == 0 && a.fuse_explicit) target_velocity(a, m, &tvx, &tvy, &tvz);
  const int b0 = active ? a.base[m] : 0, b1 = active ? a.base[nm + m] : 0;
  const int b2 = active ? a.base[2 * nm + m] : 0;
  for (int r = 0; r < 2; ++r) {
    const int t = lane + 64 * r;
    if (!active || t >= npts) continue;
    int ii, jj, kk;
    if (a.dim == 3) { ii = t / 25; jj = (t / 5) % 5; kk = t % 5; }
    else { ii = t / 5; jj = t % 5; kk = 0; }
    const int p[3] = {b0 + ii - 2, b1 + jj - 2, (a.dim == 3) ? b2 + kk - 2 : 0};
    bool ok = !(p[0] < -2 || p[0] > a.n[0] + 1 || p[1] < -2 || p[1] > a.n[1] + 1);
    if (a.dim == 3 && (p[2] < -2 || p[2] > a.n[2] + 1)) ok = false;
    valid[wv][t] = ok ? 1 : 0;
    if (!ok) continue;
    double ev[3];
    ev_values3(a, p, ev);
    prod[wv][0][t] = dval[r] * d;
    prod[wv][1][t] = dval[r] * ev[0] * d;
    prod[wv][2][t] = dval[r] * ev[1] * d;
    prod[wv][3][t] = (a.dim == 3) ? dval[r] * ev[2] * d : 0.0;
  }
  __syncthreads();
  if (active && lane < 4) {
    double acc = 0.0;
    int t = 0;
    for (; t + 5 <= npts; t += 5) {
      double v[5];
      unsigned char ok[5];
#pragma unroll
      for (int u = 0; u < 5; ++u) { v[u] = prod[wv][lane][t + u]; ok[u] = valid[wv][t + u]; }
#pragma unroll
      for (int u = 0; u < 5; ++u)
        if (ok[u]) acc += v[u];
    }
    prod[wv][lane][127] = acc;
  }
  __syncthreads();
  if (active && lane == 0) {
    const double part = prod[wv][0][127];
    double u = prod[wv][1][127], v = prod[wv][2][127], w = prod[wv][3][127];
    if (a.dim != 3) w = 0.0;
    if (a.renorm_interp && part > 0.0) { u /= part; v /= part; w /= part; }
    // update_lagrangian_marker zeroes U and F before every iteration (:599-604)
    double U0 = 0.0, U1 = 0.0, U2 = 0.0;
    U0 += u; U1 += v; U2 += w;
    a.U[m] = U0; a.U[nm + m] = U1; a.U[2 * nm + m] = U2;
    double F0 = 0.0, F1 = 0.0, F2 = 0.0;
    if (a.fuse_explicit) explicit_force_tv(a, m, tvx, tvy, tvz, U0, U1, U2, F0, F1, F2);
    a.F[m] = F0; a.F[nm + m] = F1; a.F[2 * nm + m] = F2;
  }
}

__global__ void __launch_bounds__(64 * kInterpWaves) k_ib_interp(IbArgs a) { interp_body(a, blockIdx.x); }


__global__ void k_ib_force_explicit(IbArgs a) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= a.nm) return;
  const int nm = a.nm;
  double F0 = a.F[m], F1 = a.F[nm + m], F2 = a.F[2 * nm + m];
  explicit_force(a, m, a.U[m], a.U[nm + m], a.U[2 * nm + m], F0, F1, F2);
  a.F[m] = F0; a.F[nm + m] = F1; a.F[2 * nm + m] = F2;
}

// IVC right-hand side b = (U_target - U) / dt (:1276-1296)
__device__ __forceinline__ void ivc_rhs_body(const IbArgs& a, int m) {
  if (m >= a.nm) return;
  const int nm = a.nm;
  double Ux, Uy, Uz;
  target_velocity(a, m, &Ux, &Uy, &Uz);
  a.rhs[m] = (Ux - a.U[m]) / a.dt;
  a.rhs[nm + m] = (Uy - a.U[nm + m]) / a.dt;
  a.rhs[2 * nm + m] = (Uz - a.U[2 * nm + m]) / a.dt;
}

__global__ void k_ib_ivc_rhs(IbArgs a) { ivc_rhs_body(a, blockIdx.x * blockDim.x + threadIdx.x); }

// F += A^{-1} b, one wavefront per marker row (fixed reduction tree).
__device__ __forceinline__ void ivc_matvec_body(const IbArgs& a, int blk) {
  const int lane = threadIdx.x & 63;
  const int row = blk * 4 + (threadIdx.x >> 6);
  if (row >= a.nm) return;
  const int nm = a.nm;
  const double* Ar = a.Ainv + (int64_t)row * nm;
  double sx = 0.0, sy = 0.0, sz = 0.0;
  for (int j = lane; j < nm; j += 64) {
    const double aij = Ar[j];
    sx += aij * a.rhs[j];
    sy += aij * a.rhs[nm + j];
    sz += aij * a.rhs[2 * nm + j];
  }
  sx = wave_sum(sx); sy = wave_sum(sy); sz = wave_sum(sz);
  if (lane == 0) {
    a.F[row] += sx; a.F[nm + row] += sy; a.F[2 * nm + row] += sz;
  }
}

__global__ void __launch_bounds__(256) k_ib_ivc_matvec(IbArgs a) { ivc_matvec_body(a, blockIdx.x); }

// weight of every spreading entry: ((tU*tV)*tW)*scale as in :1426 / :1435
__global__ void k_ib_entry_weights(IbArgs a, int64_t nent) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nent) return;
  const int nm = a.nm;
  const int code = a.ent_code[e];
  const int m = a.ent_m[e];
  const int ii = code & 7, jj = (code >> 3) & 7, kk = (code >> 6) & 7;
  const double tU = a.wts[(0 * 5 + ii) * nm + m];
  const double tV = a.wts[(1 * 5 + jj) * nm + m];
  const double sc = a.scale[nm + m];
  a.ent_w[e] = (a.dim == 3) ? (tU * tV * a.wts[(2 * 5 + kk) * nm + m]) * sc : (tU * tV) * sc;
}

// The spreading sum of one band cell from its CSR list: entries in list
// order (direct contributions first, then each periodic image group in
// marker order, added as a whole as SumBoundary adds ghost sums) -- the
// serial scatter's order.
__device__ inline void spread_cell_sum(const IbArgs& a, int64_t b, double tot[3]) {
  // kB entries per round: their list and marker-force loads are in flight
  // together; the products are added in list order
  constexpr int kB = 8;
  const int nm = a.nm;
  const int e0 = a.rowptr[b], e1 = a.rowptr[b + 1];
  double gs[3] = {0.0, 0.0, 0.0};
  tot[0] = tot[1] = tot[2] = 0.0;
  int cur = 0;
  for (int e = e0; e < e1; e += kB) {
    int mm[kB], gg[kB];
    double dd[kB];
#pragma unroll
    for (int u = 0; u < kB; ++u) {
      const int ee = (e + u < e1) ? e + u : e0;
      mm[u] = a.ent_m[ee];
      dd[u] = a.ent_w[ee];
      gg[u] = a.ent_code[ee] >> 9;
    }
    double f0[kB], f1[kB], f2[kB];
#pragma unroll
    for (int u = 0; u < kB; ++u) {
      f0[u] = a.F[mm[u]];
      f1[u] = a.F[nm + mm[u]];
      f2[u] = a.F[2 * nm + mm[u]];
    }
#pragma unroll
    for (int u = 0; u < kB; ++u) {
      if (e + u >= e1) break;
      const int grp = gg[u];
      const double p0 = dd[u] * (f0[u] * a.dv);
      const double p1 = dd[u] * (f1[u] * a.dv);
      const double p2 = dd[u] * (f2[u] * a.dv);
      if (grp != cur) {
        if (cur != 0)
          for (int c = 0; c < 3; ++c) tot[c] += gs[c];
        gs[0] = gs[1] = gs[2] = 0.0;
        cur = grp;
      }
      if (grp == 0) {
        tot[0] += p0; tot[1] += p1;
        if (a.dim == 3) tot[2] += p2;
      } else {
        gs[0] += p0; gs[1] += p1;
        if (a.dim == 3) gs[2] += p2;
      }
    }
  }
  if (cur != 0)
    for (int c = 0; c < 3; ++c) tot[c] += gs[c];
}

// multi-iteration forcing: EulerForceIter of the non-empty band cells
__device__ __forceinline__ void spread_thread_body(const IbArgs& a, int ci) {
  if (ci >= a.n_ne) return;
  const int b = a.ne_cells[ci];
  double tot[3];
  spread_cell_sum(a, b, tot);
  a.EFI[b] = tot[0];
  a.EFI[a.band_cells + b] = tot[1];
  a.EFI[2 * a.band_cells + b] = tot[2];
}

__global__ void __launch_bounds__(256) k_ib_spread_t(IbArgs a) {
  spread_thread_body(a, blockIdx.x * blockDim.x + threadIdx.x);
}

// One wavefront per band cell that receives contributions: lanes form the
// products, lane 0 accumulates them in list order (marker order within each
// periodic image group, direct contributions first), bit-identical to the
// serial scatter followed by SumBoundary.
constexpr int kSpreadWaves = 4;
constexpr int kSpreadChunk = 128;
__device__ __forceinline__ void spread_body(const IbArgs& a, int blk) {
  __shared__ double pr[kSpreadWaves][3][kSpreadChunk];
  __shared__ int gp[kSpreadWaves][kSpreadChunk];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ci = blk * kSpreadWaves + wv;
  const bool active = ci < a.n_ne;
  const int nm = a.nm;
  const int b = active ? a.ne_cells[ci] : 0;
  const int e0 = active ? a.rowptr[b] : 0, e1 = active ? a.rowptr[b + 1] : 0;
  double tot[3] = {0.0, 0.0, 0.0}, gs[3] = {0.0, 0.0, 0.0};
  int cur = 0;
  // every wave of the block runs the same number of chunk rounds
  int my_rounds = (e1 - e0 + kSpreadChunk - 1) / kSpreadChunk;
  __shared__ int rounds[kSpreadWaves];
  if (lane == 0) rounds[wv] = my_rounds;
  __syncthreads();
  int nr = 0;
  for (int w2 = 0; w2 < kSpreadWaves; ++w2) nr = rounds[w2] > nr ? rounds[w2] : nr;
  for (int r = 0; r < nr; ++r) {
    const int base = e0 + r * kSpreadChunk;
    for (int t = lane; t < kSpreadChunk; t += 64) {
      const int e = base + t;
      if (e < e1) {
        const int m = a.ent_m[e];
        const double dval = a.ent_w[e];
        pr[wv][0][t] = dval * (a.F[m] * a.dv);
        pr[wv][1][t] = dval * (a.F[nm + m] * a.dv);
        pr[wv][2][t] = dval * (a.F[2 * nm + m] * a.dv);
        gp[wv][t] = a.ent_code[e] >> 9;
      }
    }
    __syncthreads();
    if (active && lane == 0) {
      const int cnt = (e1 - base < kSpreadChunk) ? e1 - base : kSpreadChunk;
      for (int t = 0; t < cnt; ++t) {
        const int grp = gp[wv][t];
        if (grp != cur) {
          if (cur != 0) for (int c = 0; c < 3; ++c) tot[c] += gs[c];
          gs[0] = gs[1] = gs[2] = 0.0;
          cur = grp;
        }
        double* acc = (grp == 0) ? tot : gs;
        acc[0] += pr[wv][0][t];
        acc[1] += pr[wv][1][t];
        if (a.dim == 3) acc[2] += pr[wv][2][t];
      }
    }
    __syncthreads();
  }
  if (active && lane == 0) {
    if (cur != 0) for (int c = 0; c < 3; ++c) tot[c] += gs[c];
    a.EFI[b] = tot[0];
    a.EFI[a.band_cells + b] = tot[1];
    a.EFI[2 * a.band_cells + b] = tot[2];
  }
}

__global__ void __launch_bounds__(64 * kSpreadWaves) k_ib_spread(IbArgs a) { spread_body(a, blockIdx.x); }

// Spreading for a moving body, whose stencils change every step: every
// band cell gathers the contributions of all markers whose 5-point stencil
// covers it, in marker order -- the order of the serial scatter
// (IBMarkerDF.cpp:1405-1512).  The band does not cross a periodic boundary.
__global__ void __launch_bounds__(256) k_ib_spread_all(IbArgs a) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= a.band_cells) return;
  const int nm = a.nm;
  const int i = a.band_lo[0] + (int)(b % a.band_n[0]);
  const int j = a.band_lo[1] + (int)((b / a.band_n[0]) % a.band_n[1]);
  const int k = a.band_lo[2] + (int)(b / ((int64_t)a.band_n[0] * a.band_n[1]));
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int m = 0; m < nm; ++m) {
    const int ii = i - a.base[m] + 2, jj = j - a.base[nm + m] + 2;
    const int kk = (a.dim == 3) ? k - a.base[2 * nm + m] + 2 : 0;
    if ((unsigned)ii > 4u || (unsigned)jj > 4u || (unsigned)kk > 4u) continue;
    const double tU = a.wts[(0 * 5 + ii) * nm + m];
    const double tV = a.wts[(1 * 5 + jj) * nm + m];
    const double sc = a.scale[nm + m];
    const double dval = (a.dim == 3) ? (tU * tV * a.wts[(2 * 5 + kk) * nm + m]) * sc : (tU * tV) * sc;
    s0 += dval * (a.F[m] * a.dv);
    s1 += dval * (a.F[nm + m] * a.dv);
    if (a.dim == 3) s2 += dval * (a.F[2 * nm + m] * a.dv);
  }
  a.EFI[b] = s0;
  a.EFI[a.band_cells + b] = s1;
  a.EFI[2 * a.band_cells + b] = s2;
}

// Positions and target velocities of the pitching section at one time
// (marker_state of Examples/Pitching_airfoil/IBMUserDefinedGeometry.H:144-175,
// with the pivot offset along the chord as a parameter).  The angle terms
// (cos, sin of the pitch angle and the angular rate) come from the host.
struct MoveArgs {
  const double* xref;   // [2][nm] section outline in body coordinates
  double* xm;           // [3][nm]
  double* ut;           // [3][nm]
  int nm;
  double px, py, pivot_x, cs, sn, omega, z;
};

// One pass per step of a moving body: positions and target velocities
// (k_ib_move), stencil cells and weights (k_ib_weights), and the cell bounds
// of each kMarkerChunk-marker chunk (k_ib_chunk_bounds) from a 32-lane
// min / max; blocks hold whole chunks.
__device__ __forceinline__ void move_weights_body(const MoveArgs& mv, const IbArgs& a, int blk) {
  const int m = blk * (int)blockDim.x + (int)threadIdx.x;
  const int nm = a.nm;
  const bool act = m < nm;
  int b[3] = {0, 0, 0};
  if (act) {
    const double xrr = mv.xref[m] - mv.pivot_x, yr = mv.xref[nm + m];
    const double x = mv.px + mv.cs * xrr - mv.sn * yr;
    const double y = mv.py + mv.sn * xrr + mv.cs * yr;
    mv.xm[m] = x;
    mv.xm[nm + m] = y;
    mv.xm[2 * nm + m] = mv.z;
    mv.ut[m] = -mv.omega * (y - mv.py);
    mv.ut[nm + m] = mv.omega * (x - mv.px);
    mv.ut[2 * nm + m] = 0.0;
    const double pos[3] = {x, y, mv.z};
    marker_weights(a, m, pos, b);
  }
  const int nch = (nm + kMarkerChunk - 1) / kMarkerChunk;
  const int ch = m / kMarkerChunk;
  for (int d = 0; d < 3; ++d) {
    int lo = act ? b[d] : (1 << 30), hi = act ? b[d] : -(1 << 30);
    for (int off = kMarkerChunk / 2; off > 0; off >>= 1) {
      const int ol = __shfl_xor(lo, off, kMarkerChunk), oh = __shfl_xor(hi, off, kMarkerChunk);
      lo = ol < lo ? ol : lo;
      hi = oh > hi ? oh : hi;
    }
    if (act && (m % kMarkerChunk) == 0) {
      a.chunk_box[(2 * d) * nch + ch] = lo;
      a.chunk_box[(2 * d + 1) * nch + ch] = hi;
    }
  }
}

__global__ void __launch_bounds__(256) k_ib_move_weights(MoveArgs mv, IbArgs a) { move_weights_body(mv, a, blockIdx.x); }

// Block size of the one-block reductions (marker sums, force record): every
// launch of a level, a batch or a hierarchy uses it, so the trees are the
// same everywhere.  256 threads rather than 1024: a 1024-thread block waits
// long for a CU beside a bandwidth-bound stream-collide (a no-op one took
// 128 us in the two-lane batch).
constexpr int kSumsBlock = 256;

// Marker force integral sum_m F_m dv (:1781-1819, accumulated over MDF
// iterations) and the moments r x (F dv) (:1500-1503), with their sum as
// the body torque.  One block, fixed reduction tree.
// Thread index and count of a block in linear order (x fastest), so that the
// reductions below form the same wavefronts and trees in a block of
// kSumsBlock x 1 threads and in one of the stream-collide kernel's 2D shapes.
__device__ inline int lin_tid() { return (int)(threadIdx.x + threadIdx.y * blockDim.x); }
__device__ inline int lin_nt() { return (int)(blockDim.x * blockDim.y); }

__device__ __forceinline__ void sums_body(const IbArgs& a, int first_iter) {
  __shared__ double lds[6][16];
  const int nm = a.nm;
  const int t = lin_tid(), nt = lin_nt();
  double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const double z0 = (a.dim == 3) ? a.z0 : 0.0;
  for (int m = t; m < nm; m += nt) {
    const double fx = a.F[m] * a.dv, fy = a.F[nm + m] * a.dv, fz = a.F[2 * nm + m] * a.dv;
    const double rx = a.xm[m] - a.x0, ry = a.xm[nm + m] - a.y0;
    const double rz = (a.dim == 3) ? a.xm[2 * nm + m] - z0 : 0.0;
    const double mx = ry * fz - rz * fy, my = rz * fx - rx * fz, mz = rx * fy - ry * fx;
    a.Mm[m] = mx; a.Mm[nm + m] = my; a.Mm[2 * nm + m] = mz;
    acc[0] += fx; acc[1] += fy; acc[2] += fz;
    acc[3] += mx; acc[4] += my; acc[5] += mz;
  }
  for (int c = 0; c < 6; ++c) {
    const double r = wave_sum(acc[c]);
    if ((t & 63) == 0) lds[c][t >> 6] = r;
  }
  __syncthreads();
  if (t < 6) {
    double s = 0.0;
    for (int wv = 0; wv < (nt + 63) / 64; ++wv) s += lds[t][wv];
    if (t < 3) a.sums[t] = first_iter ? s : a.sums[t] + s;
    else a.sums[t] = s;
  }
}

__global__ void __launch_bounds__(kSumsBlock) k_ib_sums(IbArgs a) { sums_body(a, a.first_iter); }

__device__ __forceinline__ void correct_body(const IbArgs& a, int64_t b) {
  if (b >= a.band_cells) return;
  for (int c = 0; c < 3; ++c) {
    const int64_t o = c * a.band_cells + b;
    a.EV[o] += a.dt * a.EFI[o];
    a.EFT[o] += a.EFI[o];
  }
}

__global__ void k_ib_correct(IbArgs a) { correct_body(a, (int64_t)blockIdx.x * blockDim.x + threadIdx.x); }

__global__ void k_ib_export(IbArgs a) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= a.band_cells) return;
  const double rho = a.Mband[b];
  for (int c = 0; c < 3; ++c) a.Fband[c * a.band_cells + b] = a.dt * rho * a.EFT[c * a.band_cells + b];
}

// ComputeIBForce inputs (AmrCoreLBM.cpp:456-631) for the fused record.
struct RecArgs {
  double* partial;      // [3][blocks] band partial sums of the exported force
  const double* sums;   // marker sums [0..2]
  double* state;        // [0] have_prev, [1..3] previous body force
  double* hist;         // history rows of ncol doubles
  int row, ncol, method;
  double time, step, dV, U0, R;
  const double* clock;  // batched runs: [0] t_new, [1] steps taken, [2] history row of step 0
};

__device__ void write_force_record(const IbArgs& a, const RecArgs& r, const double sum[3]);

// Export F = dt * rho * EulerForceTotal (:1889-1912) fused with the Guo
// term of the coming collision on the same cell (collide_forced.H:22-50),
// which the stream-collide kernel left out because F was not yet known.
// Each block also sums its exported force for the step's force record.
// Moving body: the band is covered by tiles of 16 x 16 (i, j) cells per k
// plane, one block each, or of 16 x 4 cells when the level's IB sums may
// use a tree order (reduced storage).  The block stages the marker chunks
// whose stencil bounds reach its tile into LDS (chunk order = marker order);
// every cell sums the markers whose 5-point stencil covers it with
// k_ib_spread_all's arithmetic -- in marker order on a 16 x 16 tile (a
// wavefront per 16 x 4 cells), or on a 16 x 4 tile with the four wavefronts
// taking every fourth batch of the tile's marker list and their partial
// sums added in wavefront order.
constexpr int kGatherTile = 16;
constexpr int kGatherStage = 256 / kMarkerChunk;   // chunks staged per pass
static_assert(kGatherTile * kGatherTile == 256, "one 256-thread block per tile, 16 x 4 cells per wavefront");

// tile height in cells: 16 (serial order) or 4 (tree order)
__host__ __device__ inline int gather_tile_h(bool tree) { return tree ? 64 / kGatherTile : kGatherTile; }

struct GatherTile {
  int bi, bj, bk;     // band-local cell of this thread
  bool active;        // inside the band (tree order: wavefront 0 only)
  int64_t b;          // band cell index (valid when active)
  int i0, i1, j0, j1, k;   // the tile's cells in level-local indices (inclusive)
};

template <bool TREE>
__device__ inline GatherTile gather_tile_of(const IbArgs& a, int blk) {
  constexpr int th = TREE ? 64 / kGatherTile : kGatherTile;
  GatherTile t;
  const int tx = (a.band_n[0] + kGatherTile - 1) / kGatherTile;
  const int ty = (a.band_n[1] + th - 1) / th;
  const int cx = blk % tx, cy = (blk / tx) % ty;
  const int cell = TREE ? (int)(threadIdx.x & 63) : (int)threadIdx.x;
  t.bk = blk / (tx * ty);
  t.bi = cx * kGatherTile + cell % kGatherTile;
  t.bj = cy * th + cell / kGatherTile;
  t.active = t.bi < a.band_n[0] && t.bj < a.band_n[1] && (!TREE || threadIdx.x < 64);
  t.b = ((int64_t)t.bk * a.band_n[1] + t.bj) * a.band_n[0] + t.bi;
  t.i0 = a.band_lo[0] + cx * kGatherTile;
  t.i1 = t.i0 + kGatherTile - 1;
  t.j0 = a.band_lo[1] + cy * th;
  t.j1 = t.j0 + th - 1;
  t.k = a.band_lo[2] + t.bk;
  return t;
}

template <bool D3, bool TREE>
__device__ inline void gather_tile_sum(const IbArgs& a, const GatherTile& t, double tot[3]) {
  constexpr int kRows = D3 ? 15 : 10;
  constexpr int kBatch = 8;       // list entries per round of the cell sum
  __shared__ int s_list[kGatherStage];
  __shared__ int s_n, s_next;
  __shared__ int s_base[3][256];
  __shared__ double s_w[15][256 + 2];   // rows 2 doubles off the bank period
  __shared__ double s_f[4][256];        // F * dv (3 comps), scale
  __shared__ int4 s_wl[4][256];         // per wavefront: {entry, base x, y, z}
  __shared__ double s_part[TREE ? 3 : 1][TREE ? 3 * 64 : 1];   // tree order: wavefronts 1-3's partial sums
  tot[0] = tot[1] = tot[2] = 0.0;
  const int nm = a.nm;
  const int nch = (nm + kMarkerChunk - 1) / kMarkerChunk;
  const int i = a.band_lo[0] + t.bi, j = a.band_lo[1] + t.bj;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (threadIdx.x == 0) s_next = 0;
  __syncthreads();
  while (true) {
    const int next = s_next;
    __syncthreads();   // every thread has read s_next before the scan moves it
    if (next >= nch) break;
    // the first wavefront tests 64 chunks at once; the first kGatherStage
    // hits, in chunk (= marker) order, are staged this pass
    if (wv == 0) {
      const int ch = next + lane;
      bool hit = ch < nch;
      if (hit) {
        hit = !(t.i1 < a.chunk_box[ch] - 2 || t.i0 > a.chunk_box[nch + ch] + 2) &&
              !(t.j1 < a.chunk_box[2 * nch + ch] - 2 || t.j0 > a.chunk_box[3 * nch + ch] + 2) &&
              !(D3 && (t.k < a.chunk_box[4 * nch + ch] - 2 || t.k > a.chunk_box[5 * nch + ch] + 2));
      }
      const unsigned long long mask = __ballot(hit);
      const int total = __popcll(mask);
      const int pos = __popcll(mask & ((1ull << lane) - 1ull));
      if (hit && pos < kGatherStage) s_list[pos] = ch;
      if (lane == 0) {
        s_n = total < kGatherStage ? total : kGatherStage;
        if (total <= kGatherStage) {
          s_next = next + 64;
        } else {
          unsigned long long rest = mask;
          for (int q = 0; q < kGatherStage; ++q) rest &= rest - 1ull;   // drop the staged hits
          s_next = next + __ffsll((long long)rest) - 1;
        }
      }
    }
    __syncthreads();
    const int n = s_n;
    if (n == 0) continue;
    const int slot = threadIdx.x;
    if (slot < n * kMarkerChunk) {
      const int m = s_list[slot / kMarkerChunk] * kMarkerChunk + slot % kMarkerChunk;
      if (m < nm) {
        s_base[0][slot] = a.base[m];
        s_base[1][slot] = a.base[nm + m];
        s_base[2][slot] = D3 ? a.base[2 * nm + m] : 0;
#pragma unroll
        for (int r = 0; r < kRows; ++r) s_w[r][slot] = a.wts[r * nm + m];
        s_f[0][slot] = a.F[m] * a.dv;
        s_f[1][slot] = a.F[nm + m] * a.dv;
        s_f[2][slot] = D3 ? a.F[2 * nm + m] * a.dv : 0.0;
        s_f[3][slot] = a.scale[nm + m];
      } else {
        s_base[0][slot] = -(1 << 29);   // no cell is covered
        s_base[1][slot] = 0;
        s_base[2][slot] = 0;
      }
    }
    __syncthreads();
    // each wavefront lists, in staged (= marker) order, the entries whose
    // stencil reaches its 16 x 4 cells; every entry that covers one of its
    // cells is among them, so each cell adds its markers in the same order.
    // Tree order: all four wavefronts hold the tile's 16 x 4 cells and build
    // the same list.
    const int wj0 = TREE ? t.j0 : t.j0 + wv * (64 / kGatherTile), wj1 = wj0 + 64 / kGatherTile - 1;
    int cnt = 0;   // wave-uniform
    for (int e0 = 0; e0 < n * kMarkerChunk; e0 += 64) {
      const int e = e0 + lane;
      int bx = 0, by = 0, bz = 0;
      bool rel = false;
      if (e < n * kMarkerChunk) {
        bx = s_base[0][e];
        by = s_base[1][e];
        bz = s_base[2][e];
        rel = bx + 2 >= t.i0 && bx - 2 <= t.i1 && by + 2 >= wj0 && by - 2 <= wj1 &&
              (!D3 || (bz + 2 >= t.k && bz - 2 <= t.k));
      }
      const unsigned long long mask = __ballot(rel);
      if (rel) s_wl[wv][cnt + __popcll(mask & ((1ull << lane) - 1ull))] = make_int4(e, bx, by, bz);
      cnt += __popcll(mask);
    }
    __syncthreads();
    // kBatch entries per round: their LDS reads are in flight together, and
    // the additions stay in entry order (a skipped entry keeps the sum)
    // tree order: wavefront w takes batches w, w + 4, ...
    const int x0 = TREE ? wv * kBatch : 0, xs = TREE ? 4 * kBatch : kBatch;
    for (int x = x0; x < cnt; x += xs) {
      double dv[kBatch], f0[kBatch], f1[kBatch], f2[kBatch];
      bool h[kBatch];
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const int4 en = s_wl[wv][x + u];   // x + u < 256: x < cnt <= 256, x a multiple of kBatch
        const bool in = x + u < cnt;
        const int e = in ? en.x : 0;
        int ii = i - en.y + 2, jj = j - en.z + 2, kk = D3 ? t.k - en.w + 2 : 0;
        h[u] = in && (unsigned)ii <= 4u && (unsigned)jj <= 4u && (unsigned)kk <= 4u;
        if (!h[u]) ii = jj = kk = 0;
        const double tU = s_w[ii][e];
        const double tV = s_w[5 + jj][e];
        const double sc = s_f[3][e];
        if constexpr (D3) dv[u] = (tU * tV * s_w[10 + kk][e]) * sc;
        else dv[u] = (tU * tV) * sc;
        f0[u] = s_f[0][e];
        f1[u] = s_f[1][e];
        f2[u] = D3 ? s_f[2][e] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const double n0 = tot[0] + dv[u] * f0[u];
        const double n1 = tot[1] + dv[u] * f1[u];
        tot[0] = h[u] ? n0 : tot[0];
        tot[1] = h[u] ? n1 : tot[1];
        if constexpr (D3) {
          const double n2 = tot[2] + dv[u] * f2[u];
          tot[2] = h[u] ? n2 : tot[2];
        }
      }
    }
    __syncthreads();
  }
  if constexpr (TREE) {
    // partial sums of wavefronts 1-3 to wavefront 0, added in wavefront order
    const int lane = threadIdx.x & 63;
    if (wv > 0)
      for (int c = 0; c < 3; ++c) s_part[wv - 1][c * 64 + lane] = tot[c];
    __syncthreads();
    if (wv == 0)
      for (int w = 0; w < 3; ++w)
        for (int c = 0; c < 3; ++c) tot[c] += s_part[w][c * 64 + lane];
  }
}

// Export of the force of band cell b: MODE 0 reads EulerForceIter /
// EulerForceTotal, MODE 1 sums the cell's CSR spreading list first, MODE 2
// gathers a moving body's markers in marker order and MODE 4 in tree order
// (single-iteration forcing, where the iteration's force is the total).
template <class LAT, class ST, int MODE>
__device__ __forceinline__ void export_fixup_body(const IbArgs& a, const GridGeom& g, void* gbuf, double tau,
                                                  const RecArgs& rec, int blk, int nblk) {
  using T = typename ST::compute_t;
  using R = typename ST::raw;
  __shared__ double lds[3][4];
  int64_t b = (int64_t)blk * 256 + threadIdx.x;
  bool active = b < a.band_cells;
  double fx = 0.0, fy = 0.0, fz = 0.0;
  double cell_tot[3] = {0.0, 0.0, 0.0};
  double mac[4] = {0.0, 0.0, 0.0, 0.0};   // the cell's rho, u, v, w
  // Gathered or CSR-summed forces on a 2D level: the cell's macros and
  // populations are loaded before the sum (they do not depend on it), so
  // they arrive while it runs.  Elsewhere they are loaded after it, and the
  // populations only where the force is not zero.
  constexpr bool kGather = (MODE == 2 || MODE == 4);
  constexpr bool kPre = (MODE != 0) && (LAT::D == 2);
  T gpre[kPre ? LAT::Q : 1];
  int64_t self = 0;
  R* gb = static_cast<R*>(gbuf);
  if constexpr (kGather) {
    const GatherTile t = gather_tile_of<MODE == 4>(a, blk);
    b = t.b;
    active = t.active;
  }
  if (active && (kPre || kGather)) {
    for (int c = 0; c < 4; ++c) mac[c] = a.Mband[c * a.band_cells + b];
    if constexpr (kPre) {
      const int bi = (int)(b % a.band_n[0]);
      const int bj = (int)((b / a.band_n[0]) % a.band_n[1]);
      self = g.idx(a.band_lo[0] + bi, a.band_lo[1] + bj, a.band_lo[2]);
#pragma unroll
      for (int q = 0; q < LAT::Q; ++q) gpre[q] = ST::load(gb, (int64_t)q * g.qs + self, T(LAT::w[q]));
    }
  }
  if constexpr (kGather) gather_tile_sum<LAT::D == 3, MODE == 4>(a, gather_tile_of<MODE == 4>(a, blk), cell_tot);
  if (active) {
    if constexpr (!(kPre || kGather)) mac[0] = a.Mband[b];
    const double rho = mac[0];
    if constexpr (MODE == 0) {
      const double* tot = a.ev_from_macro ? a.EFI : a.EFT;
      cell_tot[0] = tot[b];
      cell_tot[1] = tot[a.band_cells + b];
      cell_tot[2] = tot[2 * a.band_cells + b];
    } else if constexpr (MODE == 1) {
      spread_cell_sum(a, b, cell_tot);
    }
    fx = a.dt * rho * cell_tot[0];
    fy = a.dt * rho * cell_tot[1];
    fz = a.dt * rho * cell_tot[2];
    a.Fband[b] = fx;
    a.Fband[a.band_cells + b] = fy;
    a.Fband[2 * a.band_cells + b] = fz;
    if (!(fx == 0.0 && fy == 0.0 && fz == 0.0)) {
      if constexpr (!kPre) {
        const int bi = (int)(b % a.band_n[0]);
        const int bj = (int)((b / a.band_n[0]) % a.band_n[1]);
        const int bk = (int)(b / ((int64_t)a.band_n[0] * a.band_n[1]));
        self = g.idx(a.band_lo[0] + bi, a.band_lo[1] + bj, a.band_lo[2] + bk);
      }
      if constexpr (!(kPre || kGather))
        for (int c = 1; c < 4; ++c) mac[c] = a.Mband[c * a.band_cells + b];
      const T ux = T(mac[1]), uy = T(mac[2]), uz = T(mac[3]);
      const T fac = T(1.0) - T(0.5) / T(tau);
#pragma unroll
      for (int q = 0; q < LAT::Q; ++q) {
        const int64_t o = (int64_t)q * g.qs + self;
        T gq;
        if constexpr (kPre) gq = gpre[q];
        else gq = ST::load(gb, o, T(LAT::w[q]));
        gq += guo_q<LAT, T>(q, fac, ux, uy, uz, T(fx), T(fy), T(fz));
        ST::store(gb, o, gq, T(LAT::w[q]));
      }
    }
  }
  const double v[3] = {fx, fy, fz};
  for (int c = 0; c < 3; ++c) {
    const double r = wave_sum(v[c]);
    if ((threadIdx.x & 63) == 0) lds[c][threadIdx.x >> 6] = r;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int c = 0; c < 3; ++c) {
      double s = 0.0;
      for (int wv = 0; wv < (int)(blockDim.x + 63) / 64; ++wv) s += lds[c][wv];
      rec.partial[(int64_t)c * nblk + blk] = s;
    }
  }
}

template <class LAT, class ST, int MODE>
__global__ void __launch_bounds__(256) k_ib_export_fixup(IbArgs a, GridGeom g, void* gbuf, double tau, RecArgs rec) {
  export_fixup_body<LAT, ST, MODE>(a, g, gbuf, tau, rec, blockIdx.x, gridDim.x);
}

// 3D single-iteration forcing: one wavefront per non-empty band cell sums
// its CSR list (lanes form the products, lane 0 adds them in list order as
// k_ib_spread does), exports F = dt rho EulerForce and adds the Guo term of
// the coming collision, one population per lane.  Band cells without
// entries keep F = 0.  Block partials of F in wave order for the record.
// The cell's own inputs -- its entry range, macros and populations -- do not
// depend on the spreading sum and are loaded first, so the dependent chain
// is entries -> marker forces -> sum -> store.  Levels whose IB sums may use
// a tree order (reduced storage) take spread_export_group_body instead.
//
// Tree order: kGroupLanes lanes per non-empty band cell, kGroupCells cells
// per block.  Lane s of a cell takes the entries e0 + s, e0 + s + 8, ...
// (four per round, their loads in flight together), the cell's lanes add
// their sums in a fixed butterfly (periodic images in the same sum), and
// each lane adds the Guo term to populations s, s + 8, s + 16.  Block
// partials: the cells' forces in the wave tree, then the waves in order.
constexpr int kGroupLanes = 8;
constexpr int kGroupCells = 256 / kGroupLanes;
template <class LAT, class ST>
__device__ __forceinline__ void spread_export_group_body(const IbArgs& a, const GridGeom& g, void* gbuf, double tau,
                                                         const RecArgs& rec, int blk, int nblk) {
  using T = typename ST::compute_t;
  using R = typename ST::raw;
  constexpr int kQP = (LAT::Q + kGroupLanes - 1) / kGroupLanes;
  constexpr int kE = 4;   // entries per lane per round
  __shared__ double wsum[4][3];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, sub = lane & (kGroupLanes - 1);
  const int ci = blk * kGroupCells + (int)threadIdx.x / kGroupLanes;
  const bool active = ci < a.n_ne;
  const int nm = a.nm;
  const int b = active ? a.ne_cells[ci] : 0;
  const int e0 = active ? a.ne_e[ci] : 0, e1 = active ? a.ne_e[a.n_ne + ci] : 0;
  R* gb = static_cast<R*>(gbuf);
  int64_t self = 0;
  T gq[kQP];
  T ux = T(0), uy = T(0), uz = T(0);
  double rho = 0.0;
  if (active) {
    const int bi = (int)(b % a.band_n[0]);
    const int bj = (int)((b / a.band_n[0]) % a.band_n[1]);
    const int bk = (int)(b / ((int64_t)a.band_n[0] * a.band_n[1]));
    self = g.idx(a.band_lo[0] + bi, a.band_lo[1] + bj, a.band_lo[2] + bk);
#pragma unroll
    for (int p = 0; p < kQP; ++p) {
      const int q = sub + p * kGroupLanes;
      gq[p] = (q < LAT::Q) ? ST::load(gb, (int64_t)q * g.qs + self, T(LAT::w[q])) : T(0);
    }
    rho = a.Mband[b];
    ux = T(a.Mband[a.band_cells + b]);
    uy = T(a.Mband[2 * a.band_cells + b]);
    uz = T(a.Mband[3 * a.band_cells + b]);
  }
  double tot[3] = {0.0, 0.0, 0.0};
  for (int e = e0 + sub; e < e1; e += kE * kGroupLanes) {
    int mm[kE];
    double dd[kE];
#pragma unroll
    for (int u = 0; u < kE; ++u) {
      const int ee = (e + u * kGroupLanes < e1) ? e + u * kGroupLanes : e;
      mm[u] = a.ent_m[ee];
      dd[u] = a.ent_w[ee];
    }
    double f0[kE], f1[kE], f2[kE];
#pragma unroll
    for (int u = 0; u < kE; ++u) {
      f0[u] = a.F[mm[u]];
      f1[u] = a.F[nm + mm[u]];
      f2[u] = a.F[2 * nm + mm[u]];
    }
#pragma unroll
    for (int u = 0; u < kE; ++u) {
      if (e + u * kGroupLanes >= e1) break;
      tot[0] += dd[u] * (f0[u] * a.dv);
      tot[1] += dd[u] * (f1[u] * a.dv);
      if (LAT::D == 3) tot[2] += dd[u] * (f2[u] * a.dv);
    }
  }
  // a + b == b + a: every lane of the cell ends with the same sums
#pragma unroll
  for (int off = kGroupLanes / 2; off > 0; off >>= 1)
    for (int c = 0; c < 3; ++c) tot[c] += __shfl_xor(tot[c], off, kGroupLanes);
  double fx = 0.0, fy = 0.0, fz = 0.0;
  if (active) {
    fx = a.dt * rho * tot[0];
    fy = a.dt * rho * tot[1];
    fz = a.dt * rho * tot[2];
    if (sub == 0) {
      a.Fband[b] = fx;
      a.Fband[a.band_cells + b] = fy;
      a.Fband[2 * a.band_cells + b] = fz;
    }
    if (!(fx == 0.0 && fy == 0.0 && fz == 0.0)) {
      const T fac = T(1.0) - T(0.5) / T(tau);
#pragma unroll
      for (int p = 0; p < kQP; ++p) {
        const int q = sub + p * kGroupLanes;
        if (q < LAT::Q) {
          const T v = gq[p] + guo_q<LAT, T>(q, fac, ux, uy, uz, T(fx), T(fy), T(fz));
          ST::store(gb, (int64_t)q * g.qs + self, v, T(LAT::w[q]));
        }
      }
    }
  }
  const double v[3] = {fx, fy, fz};
  for (int c = 0; c < 3; ++c) {
    const double r = wave_sum(sub == 0 ? v[c] : 0.0);
    if (lane == 0) wsum[wv][c] = r;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int c = 0; c < 3; ++c) {
      double sum = 0.0;
      for (int w2 = 0; w2 < 4; ++w2) sum += wsum[w2][c];
      rec.partial[(int64_t)c * nblk + blk] = sum;
    }
  }
}

template <class LAT, class ST>
__device__ __forceinline__ void spread_export_wave_body(const IbArgs& a, const GridGeom& g, void* gbuf, double tau,
                                                        const RecArgs& rec, int blk, int nblk) {
  if (a.tree) {
    spread_export_group_body<LAT, ST>(a, g, gbuf, tau, rec, blk, nblk);
    return;
  }
  using T = typename ST::compute_t;
  using R = typename ST::raw;
  __shared__ double pr[kSpreadWaves][3][kSpreadChunk];
  __shared__ int gp[kSpreadWaves][kSpreadChunk];
  __shared__ int rounds[kSpreadWaves];
  __shared__ double fsh[kSpreadWaves][3];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ci = blk * kSpreadWaves + wv;
  const bool active = ci < a.n_ne;
  const int nm = a.nm;
  const int b = active ? a.ne_cells[ci] : 0;
  const int e0 = active ? a.ne_e[ci] : 0, e1 = active ? a.ne_e[a.n_ne + ci] : 0;
  // population `lane` of the cell and its macros, for the Guo term
  R* gb = static_cast<R*>(gbuf);
  int64_t o = 0;
  T gq = T(0), ux = T(0), uy = T(0), uz = T(0);
  double rho = 0.0;
  if (active) {
    const int bi = (int)(b % a.band_n[0]);
    const int bj = (int)((b / a.band_n[0]) % a.band_n[1]);
    const int bk = (int)(b / ((int64_t)a.band_n[0] * a.band_n[1]));
    const int i = a.band_lo[0] + bi, j = a.band_lo[1] + bj, k = a.band_lo[2] + bk;
    if (lane < LAT::Q) {
      o = (int64_t)lane * g.qs + g.idx(i, j, k);
      gq = ST::load(gb, o, T(LAT::w[lane]));
      ux = T(a.Mband[a.band_cells + b]);
      uy = T(a.Mband[2 * a.band_cells + b]);
      uz = T(a.Mband[3 * a.band_cells + b]);
    }
    if (lane == 0) rho = a.Mband[b];
  }
  double tot[3] = {0.0, 0.0, 0.0}, gs[3] = {0.0, 0.0, 0.0};
  int cur = 0;
  if (lane == 0) rounds[wv] = (e1 - e0 + kSpreadChunk - 1) / kSpreadChunk;
  __syncthreads();
  int nr = 0;
  for (int w2 = 0; w2 < kSpreadWaves; ++w2) nr = rounds[w2] > nr ? rounds[w2] : nr;
  for (int r = 0; r < nr; ++r) {
    const int base = e0 + r * kSpreadChunk;
    for (int t = lane; t < kSpreadChunk; t += 64) {
      const int e = base + t;
      if (e < e1) {
        const int m = a.ent_m[e];
        const double dval = a.ent_w[e];
        pr[wv][0][t] = dval * (a.F[m] * a.dv);
        pr[wv][1][t] = dval * (a.F[nm + m] * a.dv);
        pr[wv][2][t] = dval * (a.F[2 * nm + m] * a.dv);
        gp[wv][t] = a.ent_code[e] >> 9;
      }
    }
    __syncthreads();
    if (active && lane == 0) {
      const int cnt = (e1 - base < kSpreadChunk) ? e1 - base : kSpreadChunk;
      for (int t = 0; t < cnt; ++t) {
        const int grp = gp[wv][t];
        if (grp != cur) {
          if (cur != 0) for (int c = 0; c < 3; ++c) tot[c] += gs[c];
          gs[0] = gs[1] = gs[2] = 0.0;
          cur = grp;
        }
        double* acc = (grp == 0) ? tot : gs;
        acc[0] += pr[wv][0][t];
        acc[1] += pr[wv][1][t];
        if (a.dim == 3) acc[2] += pr[wv][2][t];
      }
    }
    __syncthreads();
  }
  if (lane == 0) {
    double fx = 0.0, fy = 0.0, fz = 0.0;
    if (active) {
      if (cur != 0) for (int c = 0; c < 3; ++c) tot[c] += gs[c];
      fx = a.dt * rho * tot[0];
      fy = a.dt * rho * tot[1];
      fz = a.dt * rho * tot[2];
      a.Fband[b] = fx;
      a.Fband[a.band_cells + b] = fy;
      a.Fband[2 * a.band_cells + b] = fz;
    }
    fsh[wv][0] = fx; fsh[wv][1] = fy; fsh[wv][2] = fz;
  }
  __syncthreads();
  if (active && lane < LAT::Q) {
    const double fx = fsh[wv][0], fy = fsh[wv][1], fz = fsh[wv][2];
    if (!(fx == 0.0 && fy == 0.0 && fz == 0.0)) {
      const T fac = T(1.0) - T(0.5) / T(tau);
      gq += guo_q<LAT, T>(lane, fac, ux, uy, uz, T(fx), T(fy), T(fz));
      ST::store(gb, o, gq, T(LAT::w[lane]));
    }
  }
  if (threadIdx.x == 0) {
    for (int c = 0; c < 3; ++c) {
      double sum = 0.0;
      for (int w2 = 0; w2 < kSpreadWaves; ++w2) sum += fsh[w2][c];
      rec.partial[(int64_t)c * nblk + blk] = sum;
    }
  }
}

template <class LAT, class ST>
__global__ void __launch_bounds__(64 * kSpreadWaves) k_ib_spread_export_w(IbArgs a, GridGeom g, void* gbuf, double tau,
                                                                          RecArgs rec) {
  spread_export_wave_body<LAT, ST>(a, g, gbuf, tau, rec, blockIdx.x, gridDim.x);
}

// Adds the export kernel's block sums in block order and writes the
// force.dat row of the step.  The first 256 threads reduce, whatever the
// block size, so the tree is the same in every launch.
__device__ __forceinline__ void force_record_body(const IbArgs& a, const RecArgs& rec, int nblk) {
  __shared__ double lds[3][4];
  double acc[3] = {0.0, 0.0, 0.0};
  const int tid = lin_tid();
  if (tid < 256)
    for (int t = tid; t < nblk; t += 256)
      for (int c = 0; c < 3; ++c) acc[c] += rec.partial[(int64_t)c * nblk + t];
  for (int c = 0; c < 3; ++c) {
    const double r = wave_sum(acc[c]);
    if ((tid & 63) == 0 && tid < 256) lds[c][tid >> 6] = r;
  }
  __syncthreads();
  if (tid == 0) {
    double sum[3] = {0.0, 0.0, 0.0};
    for (int c = 0; c < 3; ++c)
      for (int wv = 0; wv < 4; ++wv) sum[c] += lds[c][wv];
    write_force_record(a, rec, sum);
  }
}

__global__ void __launch_bounds__(256) k_force_record(IbArgs a, RecArgs rec, int nblk) { force_record_body(a, rec, nblk); }

// Single-iteration forcing: the marker sums (k_ib_sums, same tree) and the
// record in one launch.
__device__ __forceinline__ void sums_record_body(const IbArgs& a, const RecArgs& rec, int nblk) {
  __shared__ double lds[3][16];
  sums_body(a, 1);
  __syncthreads();
  const int tid = lin_tid(), nt = lin_nt();
  double acc[3] = {0.0, 0.0, 0.0};
  for (int t = tid; t < nblk; t += nt)
    for (int c = 0; c < 3; ++c) acc[c] += rec.partial[(int64_t)c * nblk + t];
  for (int c = 0; c < 3; ++c) {
    const double r = wave_sum(acc[c]);
    if ((tid & 63) == 0) lds[c][tid >> 6] = r;
  }
  __syncthreads();
  if (tid == 0) {
    double sum[3] = {0.0, 0.0, 0.0};
    for (int c = 0; c < 3; ++c)
      for (int wv = 0; wv < (nt + 63) / 64; ++wv) sum[c] += lds[c][wv];
    write_force_record(a, rec, sum);
  }
}

__global__ void __launch_bounds__(kSumsBlock) k_sums_record(IbArgs a, RecArgs rec, int nblk) {
  sums_record_body(a, rec, nblk);
}

// A step's ghost slabs and the force record of its forcing in one launch
// (levels that step on one stream): block 0 writes the record, blocks
// [1, nghost + 1) fill kSumsBlock slab cells each.  Neither reads what the
// other writes, and the stream-collide launch that follows needs both.  The
// record, a chain of dependent reductions, goes first: dispatched last it
// would start once every other block had been placed and end the launch.
template <class LAT, class ST>
__global__ void __launch_bounds__(kSumsBlock) k_ghosts_record(GhostArgs g, unsigned nghost, IbArgs a, RecArgs rec,
                                                               int nblk) {
  (void)nghost;
  if (blockIdx.x == 0) {
    sums_record_body(a, rec, nblk);
    return;
  }
  ghost_body<LAT, ST>(g, (int64_t)(blockIdx.x - 1) * kSumsBlock + threadIdx.x);
}

// A step's stream-collide and the force record of the forcing before it in
// one launch (levels that step on one stream and need no separate ghost
// fill): block 0 writes the record, blocks [1, ntiles + 1) are k_step's
// blocks in linear order (tile t = block - 1 keeps t as its mass partial).
// The record reads the marker forces and the export partials, which the
// stream-collide does not touch; the next forcing, which rewrites them,
// follows the launch.  The record goes first so that it runs beside the
// stream-collide blocks (last, it ended the launch: sphere 256x128x128
// 112.1 against 107.4 us).  The block (k_step's shape) holds kSumsBlock
// threads and the record's reductions number them linearly, so its tree is
// k_sums_record's.
template <class LAT, class ST, int AM>
__global__ void __launch_bounds__(kBlock) k_step_record(StepArgs a, uint32_t tx, uint32_t ty, uint32_t ntiles,
                                                        IbArgs ia, RecArgs rec, int nblk) {
  (void)ntiles;
  if (blockIdx.x == 0) {
    sums_record_body(ia, rec, nblk);
    return;
  }
  const uint32_t t = blockIdx.x - 1;
  __shared__ double lds[kBlock / 64];
  const uint32_t bz = t / (tx * ty), rem = t - bz * tx * ty, by = rem / tx, bx = rem - by * tx;
  const int o0 = (a.sub_mode >= 2) ? a.sub_lo[0] : 0, o1 = (a.sub_mode >= 2) ? a.sub_lo[1] : 0;
  const int o2 = (a.sub_mode >= 2) ? a.sub_lo[2] : 0;
  step_body<LAT, ST, AM>(a, o0 + (int)(bx * blockDim.x + threadIdx.x),
                         wave_row(o1 + (int)(by * blockDim.y + threadIdx.y)), o2 + (int)bz, t, lds);
}

// The ComputeIBForce record (AmrCoreLBM.cpp:456-631) from the band sum of
// the exported force and the marker sums, appended to the history.
__device__ void write_force_record(const IbArgs& a, const RecArgs& rc, const double sum_in[3]) {
  double sum[3] = {sum_in[0], sum_in[1], sum_in[2]};
  if (a.dim != 3) sum[2] = 0.0;
  const double Fxe = -(sum[0] * rc.dV) / a.dt, Fye = -(sum[1] * rc.dV) / a.dt, Fze = -(sum[2] * rc.dV) / a.dt;
  const double Fxm = -rc.sums[0], Fym = -rc.sums[1], Fzm = -rc.sums[2];
  const double Fxx = Fxe, Fyx = Fye, Fzx = Fze;
  double Fxb = Fxe, Fyb = Fye, Fzb = Fze;
  if (rc.method == 1) { Fxb = Fxm; Fyb = Fym; Fzb = Fzm; }
  else if (rc.method == 2) { Fxb = Fxx; Fyb = Fyx; Fzb = Fzx; }
  double Fxt = Fxb, Fyt = Fyb, Fzt = Fzb;
  if (rc.state[0] != 0.0) {
    Fxt = 0.5 * (rc.state[1] + Fxb); Fyt = 0.5 * (rc.state[2] + Fyb); Fzt = 0.5 * (rc.state[3] + Fzb);
  }
  rc.state[0] = 1.0; rc.state[1] = Fxb; rc.state[2] = Fyb; rc.state[3] = Fzb;
  const double D = 2.0 * rc.R;
  const double pi = 3.14159265358979323846;
  const double denom = (a.dim == 3) ? 0.5 * 1.0 * rc.U0 * rc.U0 * (pi * rc.R * rc.R) : 0.5 * 1.0 * rc.U0 * rc.U0 * D;
  double C[15];
  for (int t = 0; t < 15; ++t) C[t] = 0.0;
  if (denom > 0.0) {
    C[0] = Fxt / denom; C[1] = Fyt / denom; C[2] = Fxb / denom; C[3] = Fyb / denom;
    C[4] = Fxe / denom; C[5] = Fye / denom; C[6] = Fxm / denom; C[7] = Fym / denom;
    C[8] = Fxx / denom; C[9] = Fyx / denom;
    C[10] = Fzt / denom; C[11] = Fzb / denom; C[12] = Fze / denom; C[13] = Fzm / denom; C[14] = Fzx / denom;
  }
  double time = rc.time, step = rc.step;
  int row = rc.row;
  if (rc.clock) {  // the step's time and row from the case's device clock
    time = rc.clock[0] + a.dt;
    step = rc.clock[1] + 1.0;
    row = (int)(rc.clock[1] - rc.clock[2]);
  }
  const double tconv = (D > 0.0) ? (time * rc.U0 / D) : 0.0;
  double* r = rc.hist + (int64_t)row * rc.ncol;
  int n = 0;
  r[n++] = step; r[n++] = time;
  r[n++] = Fxt; r[n++] = Fyt; r[n++] = C[0]; r[n++] = C[1];
  r[n++] = Fxb; r[n++] = Fyb; r[n++] = C[2]; r[n++] = C[3];
  r[n++] = Fxe; r[n++] = Fye; r[n++] = C[4]; r[n++] = C[5];
  r[n++] = Fxm; r[n++] = Fym; r[n++] = C[6]; r[n++] = C[7];
  r[n++] = Fxx; r[n++] = Fyx; r[n++] = C[8]; r[n++] = C[9]; r[n++] = tconv;
  if (a.dim == 3) {
    r[n++] = Fzt; r[n++] = C[10]; r[n++] = Fzb; r[n++] = C[11]; r[n++] = Fze; r[n++] = C[12];
    r[n++] = Fzm; r[n++] = C[13]; r[n++] = Fzx; r[n++] = C[14];
  }
}

// ---------------------------------------------------------------------------
// Batched launches: every block reads its case from a BlockRef map; one
// argument table entry per case.  A case whose iteration count is below
// `it` skips the iteration (uniformly for the whole block).
// ---------------------------------------------------------------------------
struct IbCase {
  IbArgs ib;
  GridGeom g;
  void* gbuf;          // populations the export's Guo term goes into
  double tau;
  RecArgs rec;
  int nblk_fix;        // export blocks (256 band cells each)
};

__global__ void k_ib_begin_b(const IbCase* __restrict__ A, const BlockRef* __restrict__ map) {
  const BlockRef r = map[blockIdx.x];
  const IbArgs& a = A[r.c].ib;
  if (a.ev_from_macro) return;
  ib_begin_body(a, (int64_t)r.b * 256 + threadIdx.x);
}

__global__ void __launch_bounds__(64 * kInterpWaves) k_ib_interp_b(const IbCase* __restrict__ A,
                                                                   const BlockRef* __restrict__ map, int it) {
  const BlockRef r = map[blockIdx.x];
  const IbArgs& a = A[r.c].ib;
  if (it >= a.nloop) return;
  interp_body(a, r.b);
}

__global__ void __launch_bounds__(256) k_ib_spread_t_b(const IbCase* __restrict__ A, const BlockRef* __restrict__ map,
                                                       int it) {
  const BlockRef r = map[blockIdx.x];
  const IbArgs& a = A[r.c].ib;
  if (it >= a.nloop || a.nloop == 1) return;
  spread_thread_body(a, r.b * 256 + threadIdx.x);
}

__global__ void k_ib_ivc_rhs_b(const IbCase* __restrict__ A, const BlockRef* __restrict__ map, int it) {
  const BlockRef r = map[blockIdx.x];
  const IbArgs& a = A[r.c].ib;
  if (it >= a.nloop || !a.use_ivc) return;
  ivc_rhs_body(a, r.b * 256 + threadIdx.x);
}

__global__ void __launch_bounds__(256) k_ib_ivc_matvec_b(const IbCase* __restrict__ A,
                                                         const BlockRef* __restrict__ map, int it) {
  const BlockRef r = map[blockIdx.x];
  const IbArgs& a = A[r.c].ib;
  if (it >= a.nloop || !a.use_ivc) return;
  ivc_matvec_body(a, r.b);
}

__global__ void __launch_bounds__(64 * kSpreadWaves) k_ib_spread_b(const IbCase* __restrict__ A,
                                                                   const BlockRef* __restrict__ map, int it) {
  const BlockRef r = map[blockIdx.x];
  const IbArgs& a = A[r.c].ib;
  if (it >= a.nloop) return;
  spread_body(a, r.b);
}

__global__ void __launch_bounds__(kSumsBlock) k_ib_sums_b(const IbCase* __restrict__ A, const BlockRef* __restrict__ map,
                                                    int it) {
  const IbArgs& a = A[map[blockIdx.x].c].ib;
  if (it >= a.nloop || a.nloop == 1) return;
  sums_body(a, it == 0);
}

__global__ void k_ib_correct_b(const IbCase* __restrict__ A, const BlockRef* __restrict__ map, int it) {
  const BlockRef r = map[blockIdx.x];
  const IbArgs& a = A[r.c].ib;
  if (it >= a.nloop || a.ev_from_macro) return;
  correct_body(a, (int64_t)r.b * 256 + threadIdx.x);
}

template <class LAT, class ST>
__global__ void __launch_bounds__(256) k_ib_export_fixup_b(const IbCase* __restrict__ A,
                                                           const BlockRef* __restrict__ map) {
  const BlockRef r = map[blockIdx.x];
  const IbCase& c = A[r.c];
  if (c.ib.nloop == 1) export_fixup_body<LAT, ST, 1>(c.ib, c.g, c.gbuf, c.tau, c.rec, r.b, c.nblk_fix);
  else export_fixup_body<LAT, ST, 0>(c.ib, c.g, c.gbuf, c.tau, c.rec, r.b, c.nblk_fix);
}

template <class LAT, class ST>
__global__ void __launch_bounds__(64 * kSpreadWaves) k_ib_spread_export_w_b(const IbCase* __restrict__ A,
                                                                            const BlockRef* __restrict__ map) {
  const BlockRef r = map[blockIdx.x];
  const IbCase& c = A[r.c];
  spread_export_wave_body<LAT, ST>(c.ib, c.g, c.gbuf, c.tau, c.rec, r.b, c.nblk_fix);
}

// one batched case's force record (a block of kSumsBlock threads)
__device__ __forceinline__ void case_record_body(const IbCase& c) {
  if (c.ib.nloop == 1) sums_record_body(c.ib, c.rec, c.nblk_fix);
  else force_record_body(c.ib, c.rec, c.nblk_fix);
}

// The ghost slabs and the force records of a batched step in one launch:
// the first blocks write one case's record each (gridDim.x - nghost of
// them, first for the reason k_ghosts_record's goes first), the other
// nghost fill the slab cells of their cases (kBlock each).  Neither reads
// what the other writes, and the stream-collide launch that follows needs
// both.
static_assert(kBlock == kSumsBlock, "slab and record blocks share the launch's block size");
template <class LAT, class ST>
__global__ void __launch_bounds__(kSumsBlock) k_ghosts_record_b(const GhostArgs* __restrict__ GA,
                                                               const BlockRef* __restrict__ gmap, unsigned nghost,
                                                               const IbCase* __restrict__ A,
                                                               const BlockRef* __restrict__ rmap) {
  const unsigned nrec = gridDim.x - nghost;
  if (blockIdx.x < nrec) {
    case_record_body(A[rmap[blockIdx.x].c]);
    return;
  }
  const BlockRef r = gmap[blockIdx.x - nrec];
  ghost_body<LAT, ST>(GA[r.c], (int64_t)r.b * kBlock + threadIdx.x);
}

// A batched step's stream-collide and the force records of its IB cases in
// one launch (classes whose stream-collide kernel writes the slabs, so no
// slab launch carries the records): the first gridDim.x - nstep blocks
// write one case's record each (first, as in k_step_record), the other
// nstep are k_step_b's.  Neither reads what the other writes, and the tail
// that advances the cases' clocks follows the launch.
template <class LAT, class ST, int AM>
__global__ void __launch_bounds__(kBlock) k_step_record_b(const StepArgs* __restrict__ A,
                                                          const BlockRef* __restrict__ map, unsigned nstep,
                                                          const IbCase* __restrict__ IA,
                                                          const BlockRef* __restrict__ rmap) {
  const unsigned nrec = gridDim.x - nstep;
  if (blockIdx.x < nrec) {
    case_record_body(IA[rmap[blockIdx.x].c]);
    return;
  }
  __shared__ double lds[kBlock / 64];
  step_b_body<LAT, ST, AM>(A, map, blockIdx.x - nrec, lds);
}

}  // namespace ilbm
