// Resident MI355X IB-LBM level: host driver (the AdvancePhiAtLevel /
// Evolve / IBMarkerDF roles of the reference) and the C ABI of include/ilbm.h.
//
// Per coarse step, all on one stream, no host synchronisation:
//   [IB]   k_ib_*             F^{n+1} on the band from macro^n (IBMarkerDF::update_forcing)
//   [IB]   k_ib_export_fixup  export F, Guo term of step n+1 on band cells
//                             (collide_forced.H:22-50), force.dat row
//                             (AmrCoreLBM.cpp:456-701)
//   [wall] k_fill_ghosts      ghost slabs (FillPatch / bounce-back)
//          k_step             pull + macros + BGK (AdvanceLBMAtLevel.cpp:223-386)
//                             and Sum(Rho) (AmrCoreLBM.cpp:378)
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "ilbm.h"
#include "ilbm_ib_kernels.h"

namespace ilbm {

struct HipError {
  hipError_t e;
};
#define HIPCHK(x)                                   \
  do {                                              \
    hipError_t e_ = (x);                            \
    if (e_ != hipSuccess) throw ::ilbm::HipError{e_}; \
  } while (0)

struct StatusError {
  int status;
  std::string msg;
};

template <class T>
static T* dev_alloc(size_t count) {
  if (count == 0) count = 1;
  void* p = nullptr;
  if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) throw StatusError{ILBM_ERR_ALLOC, "hipMalloc"};
  return static_cast<T*>(p);
}
static void* dev_alloc_bytes(size_t bytes) {
  void* p = nullptr;
  if (bytes == 0) bytes = 1;
  if (hipMalloc(&p, bytes) != hipSuccess) throw StatusError{ILBM_ERR_ALLOC, "hipMalloc"};
  return p;
}
static void dev_free(void* p) {
  if (p) (void)hipFree(p);
}

static constexpr double kPi = 3.14159265358979323846;

// ---------------------------------------------------------------------------
// Marker geometry on the host (IBMarkerDF::build_markers,
// Source/IBM/IBMarkerDF.cpp:201-497, and the user hooks of
// Examples/Pitching_airfoil/IBMUserDefinedGeometry.H:59-175).
// ---------------------------------------------------------------------------
struct MarkerGeom {
  int nm = 0;
  double dv = 0.0, wall_tol = 0.0;
  std::vector<double> phi, theta;     // 1-based
  std::vector<double> xref, yref, zref; // 1-based
};

static double naca_thickness(double x, double chord, double thick) {
  const double xc = std::min(std::max(x / chord, 0.0), 1.0);
  const double sqrt_x = std::sqrt(std::max(xc, 0.0));
  return 5.0 * thick * chord *
         (0.2969 * sqrt_x - 0.1260 * xc - 0.3516 * xc * xc + 0.2843 * xc * xc * xc -
          0.1036 * xc * xc * xc * xc);
}

static MarkerGeom build_marker_geometry(ilbm_case_desc& c) {
  MarkerGeom G;
  const double h = c.dx[0];
  int Ml = c.n_marker;
  if (c.geometry_type == 1) {
    if (c.dim != 2) throw StatusError{ILBM_ERR_UNSUPPORTED, "box marker geometry is 2D only"};
    const double phx = c.plo[0] + c.dx[0] * c.n[0], phy = c.plo[1] + c.dx[1] * c.n[1];
    double xlo = c.box_lo[0], xhi = c.box_hi[0], ylo = c.box_lo[1], yhi = c.box_hi[1];
    if (xhi <= xlo) { xlo = c.plo[0]; xhi = phx; }
    if (yhi <= ylo) { ylo = c.plo[1]; yhi = phy; }
    const double epsx = std::max(1.0e-6 * c.dx[0], 1.0e-12), epsy = std::max(1.0e-6 * c.dx[1], 1.0e-12);
    xlo = std::max(xlo, c.plo[0] + epsx); xhi = std::min(xhi, phx - epsx);
    ylo = std::max(ylo, c.plo[1] + epsy); yhi = std::min(yhi, phy - epsy);
    if (xhi <= xlo || yhi <= ylo) throw StatusError{ILBM_ERR_INVALID, "invalid in-domain box bounds"};
    c.box_lo[0] = xlo; c.box_hi[0] = xhi; c.box_lo[1] = ylo; c.box_hi[1] = yhi;
    double ds = (c.box_ds > 0.0) ? c.box_ds : h;
    ds = std::max(ds, 1.0e-12);
    const int nx = std::max(2, (int)std::round((xhi - xlo) / ds) + 1);
    const int ny = std::max(2, (int)std::round((yhi - ylo) / ds) + 1);
    const double dsx = (xhi - xlo) / (double)(nx - 1), dsy = (yhi - ylo) / (double)(ny - 1);
    G.xref.push_back(0.0); G.yref.push_back(0.0); G.zref.push_back(0.0);
    auto add = [&](double x, double y) { G.xref.push_back(x); G.yref.push_back(y); G.zref.push_back(0.0); };
    if (c.box_lid_only) {
      for (int i = 0; i < nx; ++i) add(xlo + (double)i * dsx, yhi);
    } else {
      for (int i = 0; i < nx; ++i) add(xlo + (double)i * dsx, ylo);
      for (int i = 0; i < nx; ++i) add(xlo + (double)i * dsx, yhi);
      for (int j = 1; j < ny - 1; ++j) {
        const double y = ylo + (double)j * dsy;
        add(xlo, y);
        add(xhi, y);
      }
    }
    Ml = (int)G.xref.size() - 1;
    const double perim = c.box_lid_only ? (xhi - xlo) : 2.0 * ((xhi - xlo) + (yhi - ylo));
    G.dv = (perim / (double)Ml) * h;
    G.wall_tol = (c.box_wall_tol > 0.0) ? c.box_wall_tol : std::max(0.25 * std::min(dsx, dsy), 1.0e-12);
  } else if (c.geometry_type == 2) {
    if (c.dim != 2) throw StatusError{ILBM_ERR_UNSUPPORTED, "pitching section geometry is 2D only"};
    const double chord = std::max(c.airfoil_chord, 1.0e-12);
    const double thick = std::max(c.airfoil_thickness, 1.0e-12);
    const int n_half = (c.n_marker > 0) ? std::max(c.n_marker / 2, 32)
                                        : std::max((int)std::ceil(1.5 * chord / h), 48);
    G.xref.push_back(0.0); G.yref.push_back(0.0); G.zref.push_back(0.0);
    for (int i = 0; i <= n_half; ++i) {
      const double beta = kPi * (double)i / (double)n_half;
      const double x = 0.5 * chord * (1.0 - std::cos(beta));
      G.xref.push_back(x); G.yref.push_back(naca_thickness(x, chord, thick)); G.zref.push_back(0.0);
    }
    for (int i = n_half - 1; i >= 1; --i) {
      const double beta = kPi * (double)i / (double)n_half;
      const double x = 0.5 * chord * (1.0 - std::cos(beta));
      G.xref.push_back(x); G.yref.push_back(-naca_thickness(x, chord, thick)); G.zref.push_back(0.0);
    }
    Ml = (int)G.xref.size() - 1;
    double perim = 0.0;
    for (int i = 1; i <= Ml; ++i) {
      const int inext = (i == Ml) ? 1 : (i + 1);
      const double dxs = G.xref[inext] - G.xref[i], dys = G.yref[inext] - G.yref[i];
      perim += std::sqrt(dxs * dxs + dys * dys);
    }
    G.dv = (perim / (double)Ml) * h;
  } else if (c.dim == 2) {
    if (Ml <= 0) Ml = std::max((int)std::ceil(2.0 * kPi * c.R / h), 16);
    G.dv = (2.0 * kPi * c.R * h) / (double)Ml;
    G.phi.assign(Ml + 1, 0.0);
    G.theta.assign(Ml + 1, 0.5 * kPi);
    for (int i = 1; i <= Ml; ++i) G.phi[i] = 2.0 * kPi * (double)(i - 1) / (double)Ml;
  } else {
    const double r1 = c.R - (c.rd - 0.5) * h, r2 = c.R - (c.rd + 0.5) * h;
    const double a1 = r1 * r1 * r1, a2 = r2 * r2 * r2;
    if (Ml <= 0) {
      const double denom = 3.0 * h * h * h / 4.0 / kPi;
      Ml = std::max((int)((a1 - a2) / denom), 32);
    }
    G.dv = (a1 - a2) / (3.0 * (double)Ml / 4.0 / kPi);
    G.phi.assign(Ml + 1, 0.0);
    G.theta.assign(Ml + 1, 0.0);
    double phi = 0.0;
    for (int mi = 0; mi < Ml; ++mi) {
      const double Hk = -1.0 + 2.0 * (double)mi / ((double)Ml - 1.0);
      const double theta = std::acos(Hk);
      if (mi == 0 || mi == Ml - 1) phi = 0.0;
      else phi = std::fmod(phi + 3.809 / std::sqrt((double)Ml) / std::sqrt(1.0 - Hk * Hk), 2.0 * kPi);
      G.phi[mi + 1] = phi;
      G.theta[mi + 1] = theta;
    }
  }
  if (Ml <= 0) throw StatusError{ILBM_ERR_INVALID, "no markers"};
  G.nm = Ml;
  return G;
}

// Positions (and user target velocities) at `time`, update_lagrangian_marker
// (IBMarkerDF.cpp:544-605) and marker_state (IBMUserDefinedGeometry.H:144-175,
// with the pivot offset along the chord as a parameter).
static void marker_positions(const ilbm_case_desc& c, const MarkerGeom& G, double time,
                             std::vector<double>& xm, std::vector<double>& ut) {
  const int nm = G.nm;
  xm.assign(3 * (size_t)nm, 0.0);
  ut.assign(3 * (size_t)nm, 0.0);
  const double z0 = (c.dim == 3) ? c.z0 : 0.0;
  for (int m = 0; m < nm; ++m) {
    const int idx = m + 1;
    double x, y, z = 0.0;
    if (c.geometry_type == 1) {
      x = G.xref[idx]; y = G.yref[idx]; z = (c.dim == 3) ? G.zref[idx] : 0.0;
    } else if (c.geometry_type == 2) {
      const double twopi = 2.0 * kPi;
      const double phase = twopi * c.pitch_frequency * time + c.pitch_phase;
      const double angle = c.pitch_mean + c.pitch_amplitude * std::sin(phase);
      const double omega = twopi * c.pitch_frequency * c.pitch_amplitude * std::cos(phase);
      const double px = c.x0 + c.pivot_x, py = c.y0;
      const double xrr = G.xref[idx] - c.pivot_x, yr = G.yref[idx];
      const double cs = std::cos(angle), sn = std::sin(angle);
      x = px + cs * xrr - sn * yr;
      y = py + sn * xrr + cs * yr;
      z = (c.dim == 3) ? c.z0 + G.zref[idx] : 0.0;
      ut[m] = -omega * (y - py);
      ut[nm + m] = omega * (x - px);
      ut[2 * nm + m] = 0.0;
    } else {
      x = c.x0 + c.R * std::sin(G.theta[idx]) * std::cos(G.phi[idx]);
      y = c.y0 + c.R * std::sin(G.theta[idx]) * std::sin(G.phi[idx]);
      if (c.dim == 3) z = z0 + c.R * std::cos(G.theta[idx]);
    }
    xm[m] = x; xm[nm + m] = y; xm[2 * (size_t)nm + m] = z;
  }
}

// Peskin kernels and the IVC operator entries (IBMarkerDF.cpp:70-145)
static double delta_host(double xf, double xp, double h, int type) {
  const double rr = std::fabs(xf - xp) / h;
  double value = 0.0;
  if (type == 0) {
    if (rr < 1.0) value = 0.125 * (3.0 - 2.0 * rr + std::sqrt(1.0 + 4.0 * rr - 4.0 * rr * rr));
    else if (rr < 2.0) value = 0.125 * (5.0 - 2.0 * rr - std::sqrt(-7.0 + 12.0 * rr - 4.0 * rr * rr));
  } else {
    if (rr < 0.5) value = (1.0 + std::sqrt(1.0 - 3.0 * rr * rr)) / 3.0;
    else if (rr < 1.5) value = (5.0 - 3.0 * rr - std::sqrt(-2.0 + 6.0 * rr - 3.0 * rr * rr)) / 6.0;
  }
  return value / h;
}

static double overlap_1d(double xa, double xb, double h, int kt, double plo, int ilo, int ihi) {
  const int ia = (int)std::floor((xa - plo) / h), ib = (int)std::floor((xb - plo) / h);
  int lo = std::max(ia - 2, ib - 2), hi = std::min(ia + 2, ib + 2);
  lo = std::max(lo, ilo);
  hi = std::min(hi, ihi);
  if (lo > hi) return 0.0;
  double sum = 0.0;
  for (int i = lo; i <= hi; ++i) {
    const double xc = plo + ((double)i + 0.5) * h;
    sum += delta_host(xc, xa, h, kt) * delta_host(xc, xb, h, kt) * h;
  }
  return sum;
}

// Partial-pivot LU and solve of IBMarkerDF::lu_factorize / lu_solve (:1015-1079)
static bool lu_factorize(std::vector<double>& a, std::vector<int>& piv, int n, double tiny) {
  piv.assign(n, 0);
  for (int k = 0; k < n; ++k) {
    int p = k;
    double maxv = std::fabs(a[(size_t)k * n + k]);
    for (int i = k + 1; i < n; ++i) {
      const double v = std::fabs(a[(size_t)i * n + k]);
      if (v > maxv) { maxv = v; p = i; }
    }
    if (maxv <= tiny) return false;
    piv[k] = p;
    if (p != k)
      for (int j = 0; j < n; ++j) std::swap(a[(size_t)k * n + j], a[(size_t)p * n + j]);
    const double akk = a[(size_t)k * n + k];
    for (int i = k + 1; i < n; ++i) {
      a[(size_t)i * n + k] /= akk;
      const double lik = a[(size_t)i * n + k];
      double* ai = &a[(size_t)i * n];
      const double* ak = &a[(size_t)k * n];
      for (int j = k + 1; j < n; ++j) ai[j] -= lik * ak[j];
    }
  }
  return true;
}

static void lu_solve(const std::vector<double>& a, const std::vector<int>& piv, std::vector<double>& x, int n) {
  for (int k = 0; k < n; ++k)
    if (piv[k] != k) std::swap(x[k], x[piv[k]]);
  for (int i = 1; i < n; ++i) {
    double s = x[i];
    for (int j = 0; j < i; ++j) s -= a[(size_t)i * n + j] * x[j];
    x[i] = s;
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = x[i];
    for (int j = i + 1; j < n; ++j) s -= a[(size_t)i * n + j] * x[j];
    x[i] = s / a[(size_t)i * n + i];
  }
}

// Flags of the events that order one step's kernels across the streams of
// an overlapped level on one device.  Such an event needs no system-scope
// fence: the kernels' own completion makes their writes visible to kernels
// that start after them on the device, and nothing on the host or another
// device reads through these events (the events before host copies keep
// HIP's default).  Without the fence the recorded signal is cheaper and the
// waiting stream starts sooner (sphere 256x128x128 FP32 overlapped step:
// 0.133 -> 0.129 ms; a batch's lane events gained nothing and keep the
// default).  ILBM_XEV_FENCE=system | device restores the default
// system-scope fence or takes a device-scope release instead.
static unsigned step_event_flags() {
  unsigned f = hipEventDisableTiming | hipEventDisableSystemFence;
  if (const char* xf = std::getenv("ILBM_XEV_FENCE")) {
    if (std::string(xf) == "system") f = hipEventDisableTiming;
    else if (std::string(xf) == "device") f = hipEventDisableTiming | hipEventReleaseToDevice;
  }
  return f;
}

// ---------------------------------------------------------------------------
// Level
// ---------------------------------------------------------------------------
// e0 / e1: timing events recorded by the dispatch itself (kernel begin / end), or null
using StepFn = void (*)(dim3, dim3, hipStream_t, const StepArgs&, hipEvent_t, hipEvent_t);
using InitFn = void (*)(dim3, dim3, hipStream_t, const InitArgs&);
using FixFn = void (*)(int, dim3, dim3, hipStream_t, const IbArgs&, const GridGeom&, void*, double, const RecArgs&);
using GhostFn = void (*)(dim3, dim3, hipStream_t, const GhostArgs&);
using GhostRecFn = void (*)(unsigned, hipStream_t, const GhostArgs&, unsigned, const IbArgs&, const RecArgs&, int);
using StepRecFn = void (*)(dim3, dim3, hipStream_t, const StepArgs&, const IbArgs&, const RecArgs&, int);

// the stream-collide kernels take a wavefront's row as uniform (wave_row)
static void check_step_block(dim3 b) {
  if (b.x % 64 != 0 || b.x * b.y > (unsigned)kBlock || b.z != 1)
    throw StatusError{ILBM_ERR_INVALID, "stream-collide block must be bx x by <= 256 threads, bx a multiple of 64"};
}

template <class LAT, class ST, int AM>
static void launch_step(dim3 g, dim3 b, hipStream_t s, const StepArgs& a, hipEvent_t e0, hipEvent_t e1) {
  check_step_block(b);
  const uint32_t nt = g.x * g.y * g.z;
  if (a.persist > 0 && nt > (uint32_t)a.persist) {
    const dim3 pg((unsigned)a.persist);
    if (e0) hipExtLaunchKernelGGL((k_step_p<LAT, ST, AM>), pg, b, 0, s, e0, e1, 0, a, g.x, g.y, nt);
    else hipLaunchKernelGGL((k_step_p<LAT, ST, AM>), pg, b, 0, s, a, g.x, g.y, nt);
    return;
  }
  if (e0) hipExtLaunchKernelGGL((k_step<LAT, ST, AM>), g, b, 0, s, e0, e1, 0, a);
  else hipLaunchKernelGGL((k_step<LAT, ST, AM>), g, b, 0, s, a);
}
template <class LAT, class ST>
static void launch_init(dim3 g, dim3 b, hipStream_t s, const InitArgs& a) {
  hipLaunchKernelGGL((k_init<LAT, ST>), g, b, 0, s, a);
}
// mode 0: export of EulerForceIter / Total; 1: fused CSR spreading (thread
// per band cell); 2: fused gather spreading of a moving body (marker order);
// 3: fused CSR spreading, one wavefront per non-empty band cell; 4: as 2 in
// tree order (16 x 4 tiles, four wavefronts per tile)
template <class LAT, class ST>
static void launch_fix(int mode, dim3 g, dim3 b, hipStream_t s, const IbArgs& a, const GridGeom& gg, void* buf,
                       double tau, const RecArgs& r) {
  if (mode == 3) hipLaunchKernelGGL((k_ib_spread_export_w<LAT, ST>), g, b, 0, s, a, gg, buf, tau, r);
  else if (mode == 1) hipLaunchKernelGGL((k_ib_export_fixup<LAT, ST, 1>), g, b, 0, s, a, gg, buf, tau, r);
  else if (mode == 2) hipLaunchKernelGGL((k_ib_export_fixup<LAT, ST, 2>), g, b, 0, s, a, gg, buf, tau, r);
  else if (mode == 4) hipLaunchKernelGGL((k_ib_export_fixup<LAT, ST, 4>), g, b, 0, s, a, gg, buf, tau, r);
  else hipLaunchKernelGGL((k_ib_export_fixup<LAT, ST, 0>), g, b, 0, s, a, gg, buf, tau, r);
}
template <class LAT, class ST>
static void launch_ghost(dim3 g, dim3 b, hipStream_t s, const GhostArgs& a) {
  hipLaunchKernelGGL((k_fill_ghosts<LAT, ST>), g, b, 0, s, a);
}

template <class LAT, class ST>
static void launch_ghost_rec(unsigned grid, hipStream_t s, const GhostArgs& g, unsigned nghost, const IbArgs& a,
                             const RecArgs& r, int nblk) {
  hipLaunchKernelGGL((k_ghosts_record<LAT, ST>), dim3(grid), dim3(kSumsBlock), 0, s, g, nghost, a, r, nblk);
}

// g, b: k_step's grid and block (kSumsBlock threads); one block more
template <class LAT, class ST, int AM>
static void launch_step_rec(dim3 g, dim3 b, hipStream_t s, const StepArgs& a, const IbArgs& ia, const RecArgs& r,
                            int nblk) {
  check_step_block(b);
  const uint32_t nt = g.x * g.y * g.z;
  hipLaunchKernelGGL((k_step_record<LAT, ST, AM>), dim3(nt + 1), b, 0, s, a, g.x, g.y, nt, ia, r, nblk);
}

template <class LAT, class ST>
static void pick(int am, StepFn& sf, InitFn& inf, FixFn& ff, GhostFn& gf, GhostRecFn& grf, StepRecFn& srf) {
  sf = (am == kAddrWrap) ? launch_step<LAT, ST, kAddrWrap>
     : (am == kAddrWallX32) ? launch_step<LAT, ST, kAddrWallX32>
     : (am == kAddrWall32) ? launch_step<LAT, ST, kAddrWall32> : launch_step<LAT, ST, kAddrWall64>;
  srf = (am == kAddrWrap) ? launch_step_rec<LAT, ST, kAddrWrap>
      : (am == kAddrWallX32) ? launch_step_rec<LAT, ST, kAddrWallX32>
      : (am == kAddrWall32) ? launch_step_rec<LAT, ST, kAddrWall32> : launch_step_rec<LAT, ST, kAddrWall64>;
  inf = launch_init<LAT, ST>;
  ff = launch_fix<LAT, ST>;
  gf = launch_ghost<LAT, ST>;
  grf = launch_ghost_rec<LAT, ST>;
}

template <class LAT>
static bool lattice_matches(const ilbm_case_desc& c) {
  if (c.q != LAT::Q || c.dim != LAT::D) return false;
  for (int q = 0; q < LAT::Q; ++q) {
    if (c.cx[q] != (double)LAT::cx[q] || c.cy[q] != (double)LAT::cy[q] || c.cz[q] != (double)LAT::cz[q])
      return false;
    if (c.w[q] != LAT::w[q]) return false;
  }
  return true;
}

// Elements between population planes: the cell count rounded up to 128.
// Large planes start 8 KiB past a multiple of 16 KiB from each other: the
// 19 loads and 19 stores a cell makes sit at one offset of every plane, and
// plane strides that are multiples of 16 KiB cost up to 10 % of the
// stream-collide bandwidth.  Measured on D3Q19 (TB/s, multiple of 16 KiB ->
// +8 KiB): 256x128x128 FP32 4.62 -> 5.10, FP64 5.30 -> 5.40; 256^3 FP32
// 5.64 -> 5.61, FP64 5.47 -> 5.41, FP16 unchanged.  Odd multiples of 4 KiB
// did as well on the FP32 box but cost 7 % at 256^3 FP64.
// 3D levels of at least kSliceMinCells cells store z-slices of all
// populations (ILBM_SLICE_MIN_CELLS overrides the threshold).  Every 3D
// level does: with padded rows (below) a 98^3 FP32 level went from 2.9 to
// 4.2 TB/s in slices, against 3.8 in planes.
constexpr int64_t kSliceMinCells = 0;
static int64_t slice_min_cells() {
  static const int64_t v = [] {
    const char* e = std::getenv("ILBM_SLICE_MIN_CELLS");
    return e ? (int64_t)std::atoll(e) : kSliceMinCells;
  }();
  return v;
}
static bool slice_layout(int dim, int64_t cells) { return dim == 3 && cells >= slice_min_cells(); }

// Rows of 3D levels padded to a multiple of kRowAlign cells (ILBM_ROW_ALIGN
// overrides; 1: none): every row starts on a 128-byte boundary in FP32, so
// a wavefront's 64 consecutive cells touch two cache lines, not three.
// D3Q19 FP32 cubes, TB/s unpadded -> padded: 98^3 2.9 -> 4.2, 130^3 3.3 ->
// 4.5; the batch of config 5's 36 spheres 19.8 -> 25.1 GLUPS.  2D levels
// keep dense rows: the 2D ensemble lost 5 % with padding (its batched
// linear tiles run across row ends).
constexpr int kRowAlign = 32;
static int row_align() {
  static const int v = [] {
    const char* e = std::getenv("ILBM_ROW_ALIGN");
    return e ? std::max(1, std::atoi(e)) : kRowAlign;
  }();
  return v;
}
static int64_t row_stride(int dim, int n0) {
  if (dim != 3) return n0;   // 2D: the batched linear tiles span dense rows
  const int a = row_align();
  return ((int64_t)n0 + a - 1) / a * a;
}

static int64_t plane_stride(int64_t cells, int64_t pop_bytes) {
  if (cells < ((int64_t)1 << 20)) return (cells + 127) / 128 * 128;
  return ((cells * pop_bytes + 16383) / 16384 * 16384 + 8192) / pop_bytes;
}

// Overlapped step of an IB level: the far cells start at once (0), or once
// this step's interpolation (1) or export (2) is done, as a launch of one
// block per tile (0) or a persistent launch of that many blocks per CU.
// Waiting for the interpolation lets it run on an idle device, but the
// cross-stream wait then sits between the far launches of two steps.
// Measured on the 256x128x128 sphere (GLUPS, FP32 / FP16-shifted / FP64):
// at once with 4 blocks per CU 29.0 / 28.3 / 12.9; after the interpolation,
// full grid 30.7 / 46.1 / 16.3 with the IB and stream-collide kernels of
// that time, 33.3 / 54.4 / 17.0 with the later ones; at once, full grid,
// 34.5 / 56.7 / 17.2.
constexpr int kFarBlocksPerCu = 0;
constexpr int kFarAfter = 0;
constexpr int kRingStream = 0;   // Level::RING_NEAR

struct Level {
  ilbm_case_desc c{};
  int lat = -1, st = -1;
  bool wrap = true;
  int am = kAddrWall64;  // stream-collide addressing (kAddr*)
  GridGeom g{};
  int64_t ncells = 0;
  size_t popbytes = 8;
  void* gbuf[2] = {nullptr, nullptr};
  int cur = 0;
  FaceBuf faces{};
  int nonper[3] = {0, 0, 0};
  double* partial = nullptr;
  int64_t nblocks = 0;
  dim3 grid, block;
  StepFn stepfn = nullptr;
  InitFn initfn = nullptr;
  FixFn fixfn = nullptr;
  GhostFn ghostfn = nullptr;
  GhostRecFn ghostrecfn = nullptr;
  StepRecFn steprecfn = nullptr;
  // the step's marker kinematics were launched with another kernel (an AMR
  // level's coarse-fine ghost fill); the next forcing skips its own launch
  bool moved_ext = false;
  // 3D single-iteration export: one wavefront per non-empty band cell (true)
  // or one thread per band cell
  bool spread_wave = true;
  // one-stream steps (levels and AMR substeps): the force record of the
  // forcing is held back (rec_to_ghost) and launched with the ghost slabs of
  // the same step or, with no slab launch, as the last block of the
  // stream-collide launch
  bool rec_to_ghost = false, rec_held = false;
  // k_step writes the x-face ghost slabs of the next step (ghost_write);
  // slabs_fresh: the current buffer's slabs were written by the last step
  // (or restored with it); a gfuse level's boundary layer is then stale
  bool gfuse = false, slabs_fresh = false;
  // IB sums in a fixed wave tree: levels with reduced storage, unless
  // ILBM_IB_EXACT=1 keeps the reference's serial order (FP64 storage always
  // keeps it)
  bool ib_tree = false;
  IbArgs held_a{};
  RecArgs held_r{};
  int held_nblk = 0;
  double tau = 0, dt = 0, t_new = 0, t_old = -1.0e100;
  int istep = 0;
  hipStream_t stream = nullptr;
  bool own_stream = true;
  double* clock_d = nullptr;     // batched runs: t_new, steps taken, history row base
  int64_t partial_cap = 0;
  // refined levels of an AMR hierarchy
  int ioff[3] = {0, 0, 0};       // global cell index of local cell 0
  int dom_n[3] = {0, 0, 0};      // the level's domain extent (0: the box)
  bool cf = false;               // every face is a coarse-fine interface
  bool window = false;           // keeps macros of a window box for the next finer level
  double* Mw[2] = {nullptr, nullptr};
  double* Dw[2] = {nullptr, nullptr};
  int wsel = 0;                  // Mw[wsel] is macro_new
  bool rec_enable = true;        // the IB pass writes the force record
  bool rec_override = false;
  double rec_time = 0.0, rec_step = 0.0;
  Slabs slabs{};
  double* fpartial = nullptr;    // [3][export blocks]
  // exports
  double* mout = nullptr;
  double* fout = nullptr;
  bool state_valid = false;
  // histories
  double* mass_dev = nullptr;
  int mass_cap = 0;
  std::vector<double> mass_host;
  double* force_dev = nullptr;
  int force_cap = 0, ncol = 0;
  std::vector<double> force_host;
  // IB
  bool ib = false;
  MarkerGeom mg;
  int nm = 0;
  int band_lo[3] = {0, 0, 0}, band_n[3] = {0, 0, 0};
  int64_t band_cells = 0;
  double *Mband = nullptr, *Fband = nullptr, *EV = nullptr, *EFI = nullptr, *EFT = nullptr;
  double *xm_d = nullptr, *ut_d = nullptr, *U_d = nullptr, *F_d = nullptr, *Mm_d = nullptr;
  int* base_d = nullptr;
  double *wts_d = nullptr, *scale_d = nullptr, *rhs_d = nullptr, *sums_d = nullptr;
  double* Ainv_d = nullptr;
  bool ivc_ready = false;
  std::vector<double> ivc_xc;  // positions the operator was built for
  int* rowptr_d = nullptr;
  int* ne_d = nullptr;
  int* nee_d = nullptr;
  int n_ne = 0;
  int* entm_d = nullptr;
  int* entc_d = nullptr;
  double* entw_d = nullptr;
  int64_t nent = 0, ent_cap = 0;
  double* frstate_d = nullptr;
  std::vector<double> xm_h, ut_h;
  bool moving = false;
  bool dev_kin = false;          // moving body advanced on the device (explicit forcing)
  double* xref_d = nullptr;      // [2][nm] section outline in body coordinates
  int* chunk_d = nullptr;        // [6][marker chunks] cell bounds of the moving markers
  // timing
  bool timing = false;
  std::vector<hipEvent_t> ev_start, ev_stop;
  double last_avg_ms = 0.0;
  int last_launches = 0;
  // Overlap of an IB level's step: the cells farther than one cell from the
  // IB band stream-collide on a second stream while the IB forcing runs on
  // the first; the box of the band grown by one cell follows the forcing.
  bool overlap = true;
  int near_lo[3] = {0, 0, 0}, near_n[3] = {0, 0, 0};
  dim3 grid_near;
  int64_t nblocks_near = 0;
  // Sum(Rho) block partials of consecutive steps fill the slots of a ring
  // (near-box partials in a second ring); one k_mass_ring launch reduces the
  // slots filled so far when the ring is full or the advance ends
  double* mring = nullptr;
  double* nring = nullptr;
  int mring_slots = 0;
  int mring_used = 0;   // next slot
  int mring_base = 0;   // first slot not yet reduced
  hipStream_t stream_b = nullptr;
  hipEvent_t ev_a = nullptr, ev_b = nullptr;  // start / end of an overlapped advance
  // Per-step dependencies between the streams: one event per kind of
  // signal, recorded by the producer (xsignal) and waited for by the
  // consumer (xwait), which always waits for the latest record.  (Stream
  // memory operations, hipStreamWrite/WaitValue64, cost less host time but
  // run as runtime kernels of 3-5 us each on the GPU, which lengthens the
  // far stream's chain; the host has slack, the GPU has not.)
  enum { X_GHOST = 0, X_NEAR = 1, X_EXP = 2, X_REC0 = 3, X_FAR0 = X_REC0 + 4, X_RING0 = X_FAR0 + 2,
         X_IB = X_RING0 + 2, X_NEAR0 = X_IB + 1, X_N = X_NEAR0 + 2 };
  hipEvent_t xev[X_N] = {};
  uint64_t xseq[X_N] = {};
  uint64_t rec_seq[4] = {};
  // During an overlapped advance the force record of a fixed body runs on a
  // third stream: step s's marker forces and export block sums go to slot
  // s % kRecRing, which the forcing reuses only after that record is done.
  static constexpr int kRecRing = 4;
  int ring_slot = -1;
  double* F_ring[kRecRing] = {};
  double* fp_ring[kRecRing] = {};
  int64_t fpartial_n = 0;
  hipStream_t stream_c = nullptr;
  hipEvent_t ev_c = nullptr;
  // the near box keeps off the non-periodic faces: the ghost slabs of the
  // next step depend on the far cells only and are filled right after them
  bool near_off_walls = false;
  // Three-region step (the near box grown by one more cell, its "ring
  // box", keeps off the non-periodic faces too): the far cells outside the
  // ring box never read what the near box or the IB forcing of any step
  // writes, so they stream back to back on the second stream, and the ring
  // between the boxes follows the near box on the first
  bool ring_ok = false;
  // resident blocks of the far-cell launch (k_step_p) of an overlapped step:
  // ILBM_FAR_BLOCKS_PER_CU per CU (0: one block per tile)
  int far_persist = 0;
  // the far cells of a three-region step start after this step's
  // interpolation (1) or export (2), 0: at once (ILBM_FAR_AFTER=interp |
  // export | none)
  int far_after = 0;
  int ib_mark = 0;          // ib_update_forcing signals X_IB after that stage
  // the stream of a three-region step's ring (ILBM_RING_STREAM=near | far |
  // own): the near box's, the far cells' after them, or a fourth stream
  // beside both
  enum { RING_NEAR = 0, RING_FAR = 1, RING_OWN = 2 };
  int ring_stream = RING_NEAR;
  bool ring_on_far = false;
  hipStream_t stream_d = nullptr;
  hipEvent_t ev_d = nullptr;
  uint64_t ib_sig = 0;
  int ring_lo[3] = {0, 0, 0}, ring_n[3] = {0, 0, 0};
  dim3 grid_ring;
  int64_t nblocks_ring = 0;
  double* rring = nullptr;
  bool ghost_ahead = false;               // this step's slabs are already queued
  uint64_t ahead_signal = 0;              // and signalled with this value

  ~Level() {
    for (auto e : ev_start) (void)hipEventDestroy(e);
    for (auto e : ev_stop) (void)hipEventDestroy(e);
    if (stream_b) (void)hipStreamSynchronize(stream_b);
    if (stream_c) (void)hipStreamSynchronize(stream_c);
    if (stream_d) (void)hipStreamSynchronize(stream_d);
    for (hipEvent_t e : {ev_a, ev_b, ev_c, ev_d})
      if (e) (void)hipEventDestroy(e);
    for (int k = 1; k < kRecRing; ++k) { dev_free(F_ring[k]); dev_free(fp_ring[k]); }
    for (hipEvent_t e : xev)
      if (e) (void)hipEventDestroy(e);
    if (stream_b) (void)hipStreamDestroy(stream_b);
    if (stream_c) (void)hipStreamDestroy(stream_c);
    if (stream_d) (void)hipStreamDestroy(stream_d);
    dev_free(mring); dev_free(nring); dev_free(rring);
    dev_free(gbuf[0]); dev_free(gbuf[1]);
    for (int f = 0; f < 6; ++f) dev_free(faces.p[f]);
    dev_free(partial); dev_free(mout); dev_free(fout); dev_free(mass_dev); dev_free(force_dev);
    dev_free(Mband); dev_free(Fband); dev_free(EV); dev_free(EFI); dev_free(EFT);
    dev_free(xm_d); dev_free(ut_d); dev_free(U_d); dev_free(F_d); dev_free(Mm_d);
    dev_free(base_d); dev_free(wts_d); dev_free(scale_d); dev_free(rhs_d); dev_free(sums_d);
    dev_free(Ainv_d); dev_free(rowptr_d); dev_free(entm_d); dev_free(entc_d); dev_free(entw_d); dev_free(frstate_d);
    dev_free(ne_d); dev_free(nee_d); dev_free(fpartial); dev_free(clock_d); dev_free(xref_d); dev_free(chunk_d);
    for (int b = 0; b < 2; ++b) { dev_free(Mw[b]); dev_free(Dw[b]); }
    if (stream && own_stream) (void)hipStreamDestroy(stream);
  }

  void validate() {
    if (c.dim != 2 && c.dim != 3) throw StatusError{ILBM_ERR_INVALID, "dim must be 2 or 3"};
    if (c.dim == 2) { c.n[2] = 1; c.periodic[2] = 1; c.bc_lo[2] = c.bc_hi[2] = 0; }
    for (int d = 0; d < 3; ++d)
      if (c.n[d] < 1) throw StatusError{ILBM_ERR_INVALID, "n must be positive"};
    for (int d = 0; d < 3; ++d)
      if (dom_n[d] == 0) dom_n[d] = c.n[d];
    if (cf)
      for (int d = 0; d < c.dim; ++d) { c.periodic[d] = 0; c.bc_lo[d] = c.bc_hi[d] = BC_FOEXTRAP; }
    for (int d = 0; d < c.dim && !cf; ++d) {
      const int ok[5] = {0, 2, 3, 1001, 1002};
      for (int side = 0; side < 2; ++side) {
        const int b = side ? c.bc_hi[d] : c.bc_lo[d];
        if (std::find(ok, ok + 5, b) == ok + 5) throw StatusError{ILBM_ERR_UNSUPPORTED, "BC type"};
        if (c.periodic[d] && b != 0) throw StatusError{ILBM_ERR_INVALID, "periodic dim needs int_dir"};
        if (!c.periodic[d] && b == 0) throw StatusError{ILBM_ERR_INVALID, "int_dir on non-periodic dim"};
        if (!c.periodic[d] && c.n[d] < 2) throw StatusError{ILBM_ERR_UNSUPPORTED, "non-periodic extent < 2"};
      }
    }
    if (lattice_matches<D2Q9>(c)) lat = 0;
    else if (lattice_matches<D3Q19>(c)) lat = 1;
    else if (lattice_matches<D3Q27>(c)) lat = 2;
    else throw StatusError{ILBM_ERR_UNSUPPORTED, "lattice table is not D2Q9/D3Q19/D3Q27 in the reference order"};
    if (c.storage < 0 || c.storage > 2) throw StatusError{ILBM_ERR_INVALID, "storage"};
    st = c.storage;
    if (c.nu <= 0.0) throw StatusError{ILBM_ERR_INVALID, "nu must be positive"};
  }

  void setup_layout() {
    wrap = true;
    for (int d = 0; d < c.dim; ++d)
      if (!c.periodic[d] || cf) { wrap = false; nonper[d] = 1; }
    g.dim = c.dim;
    for (int d = 0; d < 3; ++d) g.n[d] = c.n[d];
    g.pad = 0;
    g.js = row_stride(c.dim, c.n[0]);
    g.off0 = 0;
    {
      const int64_t pb = (st == 0) ? 8 : (st == 1) ? 4 : 2;
      const int64_t slice = g.js * c.n[1], cells = slice * c.n[2];
      if (slice_layout(c.dim, (int64_t)c.n[0] * c.n[1] * c.n[2])) {
        // [z][q][y][x]: a z-slice holds the Q populations of its cells, so a
        // cell's 19 pulls and 19 stores fall in 3 slices instead of 19
        // planes (a trivial 19+19-stream kernel: 5.7 -> 5.95 TB/s)
        // slices of 64 KiB or more: q-chunks 8 KiB past a multiple of 16 KiB
        // apart (256^3 FP32 5.96 -> 6.00 TB/s, 256x128x128 5.68 -> 5.75)
        g.qs = (slice * pb >= 65536) ? ((slice * pb + 16383) / 16384 * 16384 + 8192) / pb
                                     : (slice + 127) / 128 * 128;
        g.ks = (int64_t)c.q * g.qs;
        g.pop_n = g.ks * c.n[2];
      } else {
        g.ks = slice;
        g.qs = plane_stride(cells, pb);
        g.pop_n = (int64_t)c.q * g.qs;
      }
      // the all-periodic kernel addresses the populations with 32-bit byte offsets
      if (g.pop_n * pb >= ((int64_t)1 << 32)) wrap = false;
    }
    // ghost slabs of the non-periodic faces
    for (int d = 0; d < 3; ++d) slabs.slab_of[d][0] = slabs.slab_of[d][1] = -1;
    int nf = 0;
    int64_t off = 0;
    for (int d = 0; d < c.dim; ++d) {
      if (!nonper[d]) continue;
      for (int sd = 0; sd < 2; ++sd) {
        slabs.slab_of[d][sd] = nf;
        slabs.face_d[nf] = d;
        slabs.face_side[nf] = sd;
        int64_t cnt = 1;
        for (int e = 0; e < 3; ++e) {
          if (e == d || e >= c.dim) { slabs.face_lo[nf][e] = 0; slabs.face_n[nf][e] = 1; continue; }
          if (nonper[e] && e < d) { slabs.face_lo[nf][e] = 0; slabs.face_n[nf][e] = c.n[e]; }
          else { slabs.face_lo[nf][e] = -1; slabs.face_n[nf][e] = c.n[e] + 2; }
          cnt *= slabs.face_n[nf][e];
        }
        if (cnt >= ((int64_t)1 << 31)) throw StatusError{ILBM_ERR_INVALID, "face slab of 2^31 cells or more"};
        slabs.face_off[nf] = off;
        off += cnt;
        ++nf;
      }
    }
    slabs.nface = nf;
    for (int f = nf; f < 7; ++f) slabs.face_off[f] = off;
    ncells = (int64_t)c.n[0] * c.n[1] * c.n[2];
    popbytes = (st == 0) ? 8 : (st == 1) ? 4 : 2;
    const int Q = c.q;
    const int64_t npad = slabs.face_off[slabs.nface];
    // walled levels under 4 GiB take 32-bit offsets (a VALU-bound FP16S
    // sphere kernel: 3.78 -> 4.43 TB/s); ILBM_FORCE_ADDR64=1 keeps 64-bit
    // offsets on every walled level (the path of levels of 4 GiB or more)
    const char* f64 = std::getenv("ILBM_FORCE_ADDR64");
    const bool force64 = f64 && f64[0] == '1';
    am = wrap ? kAddrWrap
         : (!force64 && (int64_t)popbytes * (g.pop_n + npad * Q) < ((int64_t)1 << 32)) ? kAddrWall32 : kAddrWall64;
    // non-periodic faces in x only (ILBM_ADDR_GENERIC=1 keeps the select of
    // every dimension)
    const char* gen = std::getenv("ILBM_ADDR_GENERIC");
    if (am == kAddrWall32 && nonper[0] && !nonper[1] && !nonper[2] && !(gen && gen[0] == '1')) am = kAddrWallX32;
    // on such a level with ext_dir / user_2 / foextrap x faces the stream-
    // collide kernel writes the ghost slabs itself (ILBM_GHOST_SEPARATE=1
    // keeps the separate ghost kernel)
    const char* gsep = std::getenv("ILBM_GHOST_SEPARATE");
    auto fusable = [](int b) { return b == BC_FOEXTRAP || b == BC_EXT_DIR || b == BC_USER_2; };
    gfuse = am == kAddrWallX32 && !cf && fusable(c.bc_lo[0]) && fusable(c.bc_hi[0]) && !(gsep && gsep[0] == '1');
    const char* ibx = std::getenv("ILBM_IB_EXACT");
    ib_tree = st != 0 && !(ibx && ibx[0] == '1');
    const char* swv = std::getenv("ILBM_SPREAD_WAVE");
    spread_wave = !(swv && swv[0] == '0');
    for (int b = 0; b < 2; ++b) {
      // Q population planes followed by the Q ghost slabs
      const size_t bytes = popbytes * ((size_t)g.pop_n + (size_t)npad * Q);
      gbuf[b] = dev_alloc_bytes(bytes);
      HIPCHK(hipMemsetAsync(gbuf[b], 0, bytes, stream));
    }
    for (int f = 0; f < 6; ++f) faces.p[f] = nullptr;
    for (int d = 0; d < c.dim; ++d) {
      if (!nonper[d]) continue;
      const int64_t fn = (d == 0) ? (int64_t)c.n[1] * c.n[2] : (d == 1) ? (int64_t)c.n[0] * c.n[2]
                                                                       : (int64_t)c.n[0] * c.n[1];
      for (int side = 0; side < 2; ++side)
        faces.p[2 * d + side] = dev_alloc_bytes(popbytes * (size_t)fn * face_stride(Q, (int)popbytes));
    }
    const int bx = std::min(256, (c.n[0] + 63) / 64 * 64);
    const int by = 256 / bx;
    block = dim3(bx, by, 1);
    grid = dim3((c.n[0] + bx - 1) / bx, (c.n[1] + by - 1) / by, c.n[2]);
    nblocks = (int64_t)grid.x * grid.y * grid.z;
    ensure_partial(nblocks);
    clock_d = dev_alloc<double>(4);

    auto pk = [&](auto latt) {
      using L = decltype(latt);
      if (st == 0) pick<L, StF64>(am, stepfn, initfn, fixfn, ghostfn, ghostrecfn, steprecfn);
      else if (st == 1) pick<L, StF32>(am, stepfn, initfn, fixfn, ghostfn, ghostrecfn, steprecfn);
      else pick<L, StF16S>(am, stepfn, initfn, fixfn, ghostfn, ghostrecfn, steprecfn);
    };
    if (lat == 0) pk(D2Q9{});
    else if (lat == 1) pk(D3Q19{});
    else pk(D3Q27{});
    tau = 3.0 * c.nu / c.dx[0] + 0.5;  // AmrCoreLBM.cpp:225
    dt = c.dx[0];                       // AmrCoreLBM.cpp:228
  }

  // number of blocks of a tile of width bx (height kBlock / bx)
  // blocks of the batched stream-collide launch with tile width bx (0:
  // linear tiles) and nc cells per thread
  int64_t tile_blocks(int bx, int nc) const {
    if (bx == 0) return (ncells + (int64_t)nc * kBlock - 1) / ((int64_t)nc * kBlock);  // linear tiles
    const int by = kBlock / bx;
    return (int64_t)((c.n[0] + bx - 1) / bx) * ((c.n[1] + nc * by - 1) / (nc * by)) * c.n[2];
  }

  void ensure_partial(int64_t n) {
    if (n <= partial_cap) return;
    dev_free(partial);
    partial = dev_alloc<double>(n);
    partial_cap = n;
  }

  // ---------------- immersed boundary set-up ----------------
  void setup_ib() {
    ib = c.use_ib != 0;
    if (!ib) return;
    if (c.coupling_method < 0 || c.coupling_method > 2) throw StatusError{ILBM_ERR_INVALID, "coupling_method"};
    {
      ilbm_case_desc cg = c;  // markers are placed in the level's domain
      for (int d = 0; d < 3; ++d) cg.n[d] = dom_n[d];
      mg = build_marker_geometry(cg);
      c.box_lo[0] = cg.box_lo[0]; c.box_hi[0] = cg.box_hi[0];
      c.box_lo[1] = cg.box_lo[1]; c.box_hi[1] = cg.box_hi[1];
    }
    nm = mg.nm;
    moving = (c.geometry_type == 2);
    dev_kin = moving && c.coupling_method != 1;
    for (int d = 0; d < c.dim; ++d)
      if (c.periodic[d] && c.n[d] < 5) throw StatusError{ILBM_ERR_UNSUPPORTED, "periodic extent < 5 with IB"};
    // band: bounding box of the 5-point stencils (over the motion for a moving body)
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int d = 0; d < 3; ++d) { lo[d] = 1 << 30; hi[d] = -(1 << 30); }
    auto absorb = [&](const std::vector<double>& xm, int margin) {
      for (int m = 0; m < nm; ++m)
        for (int d = 0; d < c.dim; ++d) {
          const int b = (int)std::floor((xm[(size_t)d * nm + m] - c.plo[d]) / c.dx[d]) - ioff[d];
          lo[d] = std::min(lo[d], b - 2 - margin);
          hi[d] = std::max(hi[d], b + 2 + margin);
        }
    };
    std::vector<double> ut;
    if (moving && c.pitch_frequency != 0.0) {
      const double period = 1.0 / std::fabs(c.pitch_frequency);
      for (int s = 0; s < 2048; ++s) {
        marker_positions(c, mg, period * (double)s / 2048.0, xm_h, ut);
        absorb(xm_h, 1);
      }
    }
    marker_positions(c, mg, 0.0, xm_h, ut_h);
    absorb(xm_h, moving ? 1 : 0);
    for (int d = 0; d < 3; ++d) {
      if (d >= c.dim) { band_lo[d] = 0; band_n[d] = 1; continue; }
      if (cf && (lo[d] < 0 || hi[d] > c.n[d] - 1))
        throw StatusError{ILBM_ERR_UNSUPPORTED, "immersed-boundary stencil reaches the coarse-fine boundary"};
      if (c.periodic[d] && (lo[d] < 0 || hi[d] > c.n[d] - 1)) { band_lo[d] = 0; band_n[d] = c.n[d]; continue; }
      const int l = std::max(lo[d], 0), h = std::min(hi[d], c.n[d] - 1);
      if (h < l) throw StatusError{ILBM_ERR_INVALID, "markers outside the domain"};
      band_lo[d] = l;
      band_n[d] = h - l + 1;
    }
    band_cells = (int64_t)band_n[0] * band_n[1] * band_n[2];
    if (dev_kin) {
      for (int d = 0; d < c.dim; ++d)
        if (c.periodic[d] && (lo[d] < 0 || hi[d] > c.n[d] - 1))
          throw StatusError{ILBM_ERR_UNSUPPORTED, "moving body across a periodic boundary"};
      std::vector<double> xr(2 * (size_t)nm);
      for (int m = 0; m < nm; ++m) { xr[m] = mg.xref[m + 1]; xr[nm + m] = mg.yref[m + 1]; }
      xref_d = dev_alloc<double>(2 * nm);
      chunk_d = dev_alloc<int>(6 * ((nm + kMarkerChunk - 1) / kMarkerChunk));
      HIPCHK(hipMemcpyAsync(xref_d, xr.data(), sizeof(double) * 2 * nm, hipMemcpyHostToDevice, stream));
      HIPCHK(hipStreamSynchronize(stream));
    }
    Mband = dev_alloc<double>(4 * band_cells);
    Fband = dev_alloc<double>(3 * band_cells);
    EV = dev_alloc<double>(3 * band_cells);
    EFI = dev_alloc<double>(3 * band_cells);
    EFT = dev_alloc<double>(3 * band_cells);
    HIPCHK(hipMemsetAsync(Fband, 0, sizeof(double) * 3 * band_cells, stream));
    HIPCHK(hipMemsetAsync(EFI, 0, sizeof(double) * 3 * band_cells, stream));
    HIPCHK(hipMemsetAsync(Mband, 0, sizeof(double) * 4 * band_cells, stream));
    xm_d = dev_alloc<double>(3 * nm);
    ut_d = dev_alloc<double>(3 * nm);
    U_d = dev_alloc<double>(3 * nm);
    F_d = dev_alloc<double>(3 * nm);
    Mm_d = dev_alloc<double>(3 * nm);
    base_d = dev_alloc<int>(3 * nm);
    wts_d = dev_alloc<double>(15 * nm);
    scale_d = dev_alloc<double>(2 * nm);
    rhs_d = dev_alloc<double>(3 * nm);
    sums_d = dev_alloc<double>(8);
    frstate_d = dev_alloc<double>(4);
    HIPCHK(hipMemsetAsync(frstate_d, 0, sizeof(double) * 4, stream));
    HIPCHK(hipMemsetAsync(sums_d, 0, sizeof(double) * 8, stream));
    HIPCHK(hipMemsetAsync(F_d, 0, sizeof(double) * 3 * nm, stream));
    HIPCHK(hipMemsetAsync(U_d, 0, sizeof(double) * 3 * nm, stream));
    HIPCHK(hipMemsetAsync(Mm_d, 0, sizeof(double) * 3 * nm, stream));
    rowptr_d = dev_alloc<int>(band_cells + 1);
    const int64_t tiles = (int64_t)((band_n[0] + kGatherTile - 1) / kGatherTile) *
                          ((band_n[1] + gather_tile_h(true) - 1) / gather_tile_h(true)) * band_n[2];
    fpartial_n = 3 * ((band_cells + 255) / 256 + band_cells / kSpreadWaves + tiles + 2);
    fpartial = dev_alloc<double>(fpartial_n);
    upload_markers();
    ncol = (c.dim == 3) ? 33 : 23;
  }

  // positions, spreading lists and (for IVC) the inverse operator
  void upload_markers() {
    HIPCHK(hipMemcpyAsync(xm_d, xm_h.data(), sizeof(double) * 3 * nm, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(ut_d, ut_h.data(), sizeof(double) * 3 * nm, hipMemcpyHostToDevice, stream));
    if (!dev_kin) build_spread_lists();
    if (c.coupling_method == 1) build_ivc();
    // delta weights depend on positions only: once for a fixed body, every
    // step for a moving one
    const IbArgs a = ib_args();
    hipLaunchKernelGGL(k_ib_weights, dim3((nm + 255) / 256), dim3(256), 0, stream, a);
    if (nent)
      hipLaunchKernelGGL(k_ib_entry_weights, dim3((unsigned)((nent + 255) / 256)), dim3(256), 0, stream, a, nent);
    HIPCHK(hipGetLastError());
  }

  void build_spread_lists() {
    struct Ent { int64_t b; int grp; int m; int code; };
    std::vector<Ent> ents;
    ents.reserve((size_t)nm * (c.dim == 3 ? 125 : 25));
    for (int m = 0; m < nm; ++m) {
      int b[3] = {0, 0, 0};
      for (int d = 0; d < c.dim; ++d)
        b[d] = (int)std::floor((xm_h[(size_t)d * nm + m] - c.plo[d]) / c.dx[d]) - ioff[d];
      const int kmax = (c.dim == 3) ? 5 : 1;
      for (int ii = 0; ii < 5; ++ii)
        for (int jj = 0; jj < 5; ++jj)
          for (int kk = 0; kk < kmax; ++kk) {
            int p[3] = {b[0] + ii - 2, b[1] + jj - 2, (c.dim == 3) ? b[2] + kk - 2 : 0};
            bool drop = false;
            int t[3] = {0, 0, 0};
            for (int d = 0; d < c.dim; ++d) {
              if (p[d] < -2 || p[d] > c.n[d] + 1) drop = true;
              if (p[d] < 0 || p[d] > c.n[d] - 1) {
                if (!c.periodic[d]) drop = true;
                else if (p[d] < 0) { t[d] = -1; p[d] += c.n[d]; }
                else { t[d] = 1; p[d] -= c.n[d]; }
              }
            }
            if (drop) continue;
            const int bi = p[0] - band_lo[0], bj = p[1] - band_lo[1], bk = p[2] - band_lo[2];
            if (bi < 0 || bi >= band_n[0] || bj < 0 || bj >= band_n[1] || bk < 0 || bk >= band_n[2])
              throw StatusError{ILBM_ERR_STATE, "marker stencil left the IB band"};
            const int64_t bidx = ((int64_t)bk * band_n[1] + bj) * band_n[0] + bi;
            int grp = 0;
            if (t[0] || t[1] || t[2]) {
              const int code = (t[2] + 1) * 9 + (t[1] + 1) * 3 + (t[0] + 1);
              grp = (code < 13) ? code + 1 : code;
            }
            ents.push_back({bidx, grp, m, ii | (jj << 3) | ((c.dim == 3 ? kk : 0) << 6) | (grp << 9)});
          }
    }
    std::stable_sort(ents.begin(), ents.end(), [](const Ent& x, const Ent& y) {
      if (x.b != y.b) return x.b < y.b;
      if (x.grp != y.grp) return x.grp < y.grp;
      return x.m < y.m;
    });
    std::vector<int> rowptr(band_cells + 1, 0), em(ents.size()), ec(ents.size());
    for (size_t e = 0; e < ents.size(); ++e) {
      rowptr[ents[e].b + 1]++;
      em[e] = ents[e].m;
      ec[e] = ents[e].code;
    }
    std::vector<int> ne;
    for (int64_t b = 0; b < band_cells; ++b)
      if (rowptr[b + 1] > 0) ne.push_back((int)b);
    for (int64_t b = 0; b < band_cells; ++b) rowptr[b + 1] += rowptr[b];
    nent = (int64_t)ents.size();
    dev_free(ne_d);
    dev_free(nee_d);
    n_ne = (int)ne.size();
    // each non-empty cell's entry range beside its index (one load fewer in
    // the dependent chain of the spreading kernels)
    std::vector<int> nee(2 * (size_t)std::max(n_ne, 1), 0);
    for (int ci = 0; ci < n_ne; ++ci) { nee[ci] = rowptr[ne[ci]]; nee[n_ne + ci] = rowptr[ne[ci] + 1]; }
    ne_d = dev_alloc<int>(std::max(n_ne, 1));
    nee_d = dev_alloc<int>(2 * std::max(n_ne, 1));
    if (n_ne) {
      HIPCHK(hipMemcpyAsync(ne_d, ne.data(), sizeof(int) * n_ne, hipMemcpyHostToDevice, stream));
      HIPCHK(hipMemcpyAsync(nee_d, nee.data(), sizeof(int) * 2 * n_ne, hipMemcpyHostToDevice, stream));
    }
    if (nent > ent_cap) {
      dev_free(entm_d); dev_free(entc_d); dev_free(entw_d);
      ent_cap = std::max<int64_t>(nent, 1) * 2;
      entm_d = dev_alloc<int>(ent_cap);
      entc_d = dev_alloc<int>(ent_cap);
      entw_d = dev_alloc<double>(ent_cap);
    }
    HIPCHK(hipMemcpyAsync(rowptr_d, rowptr.data(), sizeof(int) * (band_cells + 1), hipMemcpyHostToDevice, stream));
    if (nent) {
      HIPCHK(hipMemcpyAsync(entm_d, em.data(), sizeof(int) * nent, hipMemcpyHostToDevice, stream));
      HIPCHK(hipMemcpyAsync(entc_d, ec.data(), sizeof(int) * nent, hipMemcpyHostToDevice, stream));
    }
    HIPCHK(hipStreamSynchronize(stream));  // host vectors go out of scope
  }

  // build_ivc_operator + lu_factorize (IBMarkerDF.cpp:1081-1142); the solve
  // is applied on the device as F += A^{-1} b.
  void build_ivc() {
    const int n = nm;
    bool changed = ivc_xc.size() != (size_t)(3 * n);
    if (!changed)
      for (int i = 0; i < 3 * n && !changed; ++i)
        if (std::fabs(xm_h[i] - ivc_xc[i]) > 1.0e-12) changed = true;
    if (!changed && ivc_ready && c.ivc_rebuild_matrix <= 0) return;
    const int kt = (c.delta_type == 3 || c.delta_type == 1) ? 1 : 0;
    std::vector<double> A((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i)
      for (int j = i; j < n; ++j) {
        const double sx = overlap_1d(xm_h[i], xm_h[j], c.dx[0], kt, c.plo[0], 0, dom_n[0] - 1);
        const double sy = overlap_1d(xm_h[n + i], xm_h[n + j], c.dx[1], kt, c.plo[1], 0, dom_n[1] - 1);
        const double sz = (c.dim == 3)
                              ? overlap_1d(xm_h[2 * n + i], xm_h[2 * n + j], c.dx[2], kt, c.plo[2], 0, dom_n[2] - 1)
                              : 1.0;
        double val = mg.dv * sx * sy * sz;
        if (i == j) val += c.ivc_diag_reg;
        A[(size_t)i * n + j] = val;
        A[(size_t)j * n + i] = val;
      }
    const double tiny = std::max(c.ivc_diag_reg * 1.0e-4, 1.0e-30);
    std::vector<int> piv;
    ivc_ready = lu_factorize(A, piv, n, tiny);
    ivc_xc = xm_h;
    if (!ivc_ready) return;  // explicit forcing fallback (:1253-1259)
    std::vector<double> Ainv((size_t)n * n), col(n);
    for (int j = 0; j < n; ++j) {
      std::fill(col.begin(), col.end(), 0.0);
      col[j] = 1.0;
      lu_solve(A, piv, col, n);
      for (int i = 0; i < n; ++i) Ainv[(size_t)i * n + j] = col[i];
    }
    if (!Ainv_d) Ainv_d = dev_alloc<double>((size_t)n * n);
    HIPCHK(hipMemcpyAsync(Ainv_d, Ainv.data(), sizeof(double) * (size_t)n * n, hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
  }

  IbArgs ib_args() const {
    IbArgs a{};
    a.dim = c.dim;
    for (int d = 0; d < 3; ++d) {
      a.n[d] = c.n[d]; a.periodic[d] = c.periodic[d]; a.plo[d] = c.plo[d]; a.dx[d] = c.dx[d];
      a.band_lo[d] = band_lo[d]; a.band_n[d] = band_n[d];
    }
    a.bc = bc_spec();
    a.band_cells = band_cells;
    a.nm = nm;
    a.dt = dt;
    a.dv = mg.dv;
    a.kernel_type = (c.delta_type == 3 || c.delta_type == 1) ? 1 : 0;
    a.renorm_interp = (c.renormalize_delta != 0) && (c.coupling_method != 1);
    a.renorm_spread = (c.renormalize_delta != 0);
    a.coupling = c.coupling_method;
    a.relax = std::max(0.0, std::min(1.0, c.mdf_relax));
    a.diag_eps = std::max(c.explicit_diag_eps, 1.0e-30);
    a.geometry = c.geometry_type;
    a.x0 = c.x0; a.y0 = c.y0; a.z0 = c.z0;
    a.ub[0] = c.ubx; a.ub[1] = c.uby; a.ub[2] = c.ubz;
    a.om[0] = c.omx; a.om[1] = c.omy; a.om[2] = c.omz;
    a.box_yhi = c.box_hi[1]; a.wall_tol = mg.wall_tol;
    for (int d = 0; d < 3; ++d) a.lid_u[d] = c.box_lid_u[d];
    a.xm = xm_d; a.ut = ut_d; a.U = U_d; a.F = (ring_slot >= 0) ? F_ring[ring_slot] : F_d; a.Mm = Mm_d;
    a.base = base_d; a.wts = wts_d; a.scale = scale_d;
    a.Mband = Mband; a.EV = EV; a.EFI = EFI; a.EFT = EFT; a.Fband = Fband;
    a.rowptr = rowptr_d; a.ent_m = entm_d; a.ent_code = entc_d; a.ent_w = entw_d;
    a.ne_cells = ne_d; a.ne_e = nee_d; a.n_ne = n_ne;
    a.chunk_box = chunk_d;
    a.ev_from_macro = (c.coupling_method == 1 || std::max(c.loop_ns, 1) == 1) ? 1 : 0;
    a.use_ivc = (c.coupling_method == 1 && ivc_ready) ? 1 : 0;
    // IVC whose LU failed falls back to explicit forcing (IBMarkerDF.cpp:1253-1259)
    a.fuse_explicit = (c.coupling_method != 1 || !ivc_ready) ? 1 : 0;
    a.tree = ib_tree ? 1 : 0;
    a.nloop = (c.coupling_method == 1) ? 1 : std::max(c.loop_ns, 1);
    a.first_iter = 1;
    for (int d = 0; d < 3; ++d) {
      a.ioff[d] = ioff[d];
      a.dlo[d] = -ioff[d];
      a.dhi[d] = dom_n[d] - 1 - ioff[d];
    }
    a.Ainv = Ainv_d; a.rhs = rhs_d; a.sums = sums_d;
    return a;
  }

  // how the IB force is exported (see launch_fix) and its block count
  int fix_mode() const {
    const int nloop = (c.coupling_method == 1) ? 1 : std::max(c.loop_ns, 1);
    if (nloop > 1) return 0;
    if (dev_kin) return ib_tree ? 4 : 2;
    return (c.dim == 3 && spread_wave) ? 3 : 1;
  }
  unsigned fix_blocks() const {
    if (fix_mode() == 3)
      return (unsigned)(ib_tree ? (n_ne + kGroupCells - 1) / kGroupCells : (n_ne + kSpreadWaves - 1) / kSpreadWaves);
    if (fix_mode() == 2 || fix_mode() == 4) {
      const int th = gather_tile_h(fix_mode() == 4);
      return (unsigned)(((band_n[0] + kGatherTile - 1) / kGatherTile) * ((band_n[1] + th - 1) / th) * band_n[2]);
    }
    return (unsigned)((band_cells + 255) / 256);
  }

  BcSpec bc_spec() const {
    BcSpec b{};
    for (int d = 0; d < 3; ++d) {
      b.periodic[d] = (d < c.dim) ? c.periodic[d] : 1;
      b.lo[d] = c.bc_lo[d]; b.hi[d] = c.bc_hi[d];
      for (int k = 0; k < 4; ++k) { b.lo_val[d][k] = c.bc_lo_val[d][k]; b.hi_val[d][k] = c.bc_hi_val[d][k]; }
    }
    return b;
  }

  // Positions and weights of the pitching section at `time` on the device;
  // the angle terms as marker_positions computes them.
  MoveArgs move_args(double time) const {
    const double twopi = 2.0 * kPi;
    const double phase = twopi * c.pitch_frequency * time + c.pitch_phase;
    const double angle = c.pitch_mean + c.pitch_amplitude * std::sin(phase);
    MoveArgs mv{};
    mv.xref = xref_d; mv.xm = xm_d; mv.ut = ut_d; mv.nm = nm;
    mv.omega = twopi * c.pitch_frequency * c.pitch_amplitude * std::cos(phase);
    mv.px = c.x0 + c.pivot_x; mv.py = c.y0; mv.pivot_x = c.pivot_x;
    mv.cs = std::cos(angle); mv.sn = std::sin(angle);
    mv.z = 0.0;
    return mv;
  }

  // IBMarkerDF::update_forcing (IBMarkerDF.cpp:1522-1917) on the band, ending
  // with the export of F fused with the Guo term of the coming collision and
  // with the ComputeIBForce record of the step (row `row` of the history).
  void ib_update_forcing(double time, int row) {
    // the slot's previous record must be done before its inputs are rewritten
    if (ring_slot >= 0 && rec_seq[ring_slot]) xwait(stream, X_REC0 + ring_slot, rec_seq[ring_slot]);
    if (dev_kin) {
      // the pitching kinematics on the device, unless the caller launched
      // them with another kernel of the step (moved_ext)
      if (!moved_ext) {
        const IbArgs aw = ib_args();
        hipLaunchKernelGGL(k_ib_move_weights, dim3((nm + 255) / 256), dim3(256), 0, stream, move_args(time), aw);
      }
      moved_ext = false;
    } else if (moving) {
      marker_positions(c, mg, time, xm_h, ut_h);
      upload_markers();
    }
    IbArgs a = ib_args();
    const int tb = 256;
    const unsigned gb = (unsigned)((band_cells + tb - 1) / tb);
    const bool use_ivc = a.use_ivc != 0;
    const int nloop = a.nloop;
    if (!a.ev_from_macro) hipLaunchKernelGGL(k_ib_begin, dim3(gb), dim3(tb), 0, stream, a);
    for (int it = 0; it < nloop; ++it) {
      a.first_iter = (it == 0) ? 1 : 0;
      hipLaunchKernelGGL(k_ib_interp, dim3(interp_blocks(nm, a.tree != 0, c.dim)), dim3(64 * kInterpWaves), 0,
                         stream, a);
      if (ib_mark == 1 && nloop == 1 && !use_ivc) ib_sig = xsignal(stream, X_IB);
      if (use_ivc) {
        hipLaunchKernelGGL(k_ib_ivc_rhs, dim3((nm + tb - 1) / tb), dim3(tb), 0, stream, a);
        hipLaunchKernelGGL(k_ib_ivc_matvec, dim3((nm + 3) / 4), dim3(256), 0, stream, a);
      }
      if (nloop == 1) break;  // single iteration: spreading fused into the export
      if (dev_kin)
        hipLaunchKernelGGL(k_ib_spread_all, dim3(gb), dim3(tb), 0, stream, a);
      else if (n_ne)
        hipLaunchKernelGGL(k_ib_spread_t, dim3((n_ne + 255) / 256), dim3(256), 0, stream, a);
      hipLaunchKernelGGL(k_ib_sums, dim3(1), dim3(kSumsBlock), 0, stream, a);
      if (!a.ev_from_macro) hipLaunchKernelGGL(k_ib_correct, dim3(gb), dim3(tb), 0, stream, a);
    }
    RecArgs r{};
    r.partial = (ring_slot >= 0) ? fp_ring[ring_slot] : fpartial; r.sums = sums_d; r.state = frstate_d;
    r.hist = force_dev; r.row = row; r.ncol = ncol; r.method = c.force_eval_method;
    r.time = time + dt; r.step = (double)(istep + 1);
    r.dV = (c.dim == 3) ? c.dx[0] * c.dx[1] * c.dx[2] : c.dx[0] * c.dx[1];
    r.U0 = c.U0; r.R = c.R;
    r.clock = nullptr;
    if (rec_override) { r.time = rec_time; r.step = rec_step; }
    const int mode = fix_mode();
    const unsigned gfix = fix_blocks();
    if (gfix) fixfn(mode, dim3(gfix), dim3(tb), stream, a, g, gbuf[cur], tau, r);
    if (ib_mark == 2 || (ib_mark == 1 && (nloop != 1 || use_ivc))) ib_sig = xsignal(stream, X_IB);
    if (nloop == 1) {
      // the marker sums feed only the force record (a sub-step without a
      // record is followed by one that has it)
      if (rec_enable && ring_slot >= 0) {
        // off the step's critical path: the record reads only its slot
        xwait(stream_c, X_EXP, xsignal(stream, X_EXP));
        hipLaunchKernelGGL(k_sums_record, dim3(1), dim3(kSumsBlock), 0, stream_c, a, r, (int)gfix);
        rec_seq[ring_slot] = xsignal(stream_c, X_REC0 + ring_slot);
      } else if (rec_enable && rec_to_ghost) {
        rec_held = true;
        held_a = a; held_r = r; held_nblk = (int)gfix;
      } else if (rec_enable) {
        hipLaunchKernelGGL(k_sums_record, dim3(1), dim3(kSumsBlock), 0, stream, a, r, (int)gfix);
      }
    } else if (rec_enable) {
      hipLaunchKernelGGL(k_force_record, dim3(1), dim3(256), 0, stream, a, r, (int)gfix);
    }
    HIPCHK(hipGetLastError());
  }

  // The stages of update_forcing without the export and the force record
  // (the per-box path exports into the caller's fabs): kinematics, then per
  // iteration interpolation with the explicit force, the IVC solve,
  // spreading into EulerForceIter, the marker sums and moments, and (more
  // than one iteration) the velocity correction.  The macros of the band
  // (Mband) are the caller's.
  void ib_stages(double time) {
    if (dev_kin) {
      hipLaunchKernelGGL(k_ib_move_weights, dim3((nm + 255) / 256), dim3(256), 0, stream, move_args(time), ib_args());
    } else if (moving) {
      marker_positions(c, mg, time, xm_h, ut_h);
      upload_markers();
    }
    IbArgs a = ib_args();
    const int tb = 256;
    const unsigned gb = (unsigned)((band_cells + tb - 1) / tb);
    if (!a.ev_from_macro) hipLaunchKernelGGL(k_ib_begin, dim3(gb), dim3(tb), 0, stream, a);
    for (int it = 0; it < a.nloop; ++it) {
      a.first_iter = (it == 0) ? 1 : 0;
      hipLaunchKernelGGL(k_ib_interp, dim3(interp_blocks(nm, a.tree != 0, c.dim)), dim3(64 * kInterpWaves), 0,
                         stream, a);
      if (a.use_ivc) {
        hipLaunchKernelGGL(k_ib_ivc_rhs, dim3((nm + tb - 1) / tb), dim3(tb), 0, stream, a);
        hipLaunchKernelGGL(k_ib_ivc_matvec, dim3((nm + 3) / 4), dim3(256), 0, stream, a);
      }
      if (dev_kin)
        hipLaunchKernelGGL(k_ib_spread_all, dim3(gb), dim3(tb), 0, stream, a);
      else if (n_ne)
        hipLaunchKernelGGL(k_ib_spread_t, dim3((n_ne + 255) / 256), dim3(256), 0, stream, a);
      hipLaunchKernelGGL(k_ib_sums, dim3(1), dim3(kSumsBlock), 0, stream, a);
      if (!a.ev_from_macro) hipLaunchKernelGGL(k_ib_correct, dim3(gb), dim3(tb), 0, stream, a);
    }
    HIPCHK(hipGetLastError());
  }

  // ---------------- lattice ----------------
  GhostArgs ghost_args() {
    GhostArgs ga{};
    ga.g = g; ga.gbuf = gbuf[cur]; ga.faces = faces; ga.bc = bc_spec();
    ga.slabs = slabs;
    return ga;
  }

  StepArgs step_args(bool want_state) {
    StepArgs a{};
    a.g = g;
    a.gin = gbuf[cur];
    a.gout = gbuf[cur ^ 1];
    a.tau = tau;
    a.inv_tau = 1.0 / tau;
    for (int d = 0; d < 3; ++d) {
      const double pf = (c.use_prescribed_force && !ib) ? c.prescribed_force[d] : 0.0;
      a.Fmac[d] = pf;
      a.Fcol[d] = pf;
      a.band_lo[d] = ib ? band_lo[d] : 0;
      a.band_n[d] = ib ? band_n[d] : 0;
      a.nonper[d] = nonper[d];
    }
    a.Fband = Fband;
    a.Mband = window ? Mw[wsel] : Mband;
    a.band_force = ib ? 1 : 0;
    a.band_cells = (ib || window) ? band_cells : 0;
    if (window)
      for (int d = 0; d < 3; ++d) { a.band_lo[d] = band_lo[d]; a.band_n[d] = band_n[d]; }
    a.mout = want_state ? mout : nullptr;
    a.fout = want_state ? fout : nullptr;
    a.faces = faces;
    a.partial = partial;
    a.slabs = slabs;
    a.gfuse = gfuse ? 1 : 0;
    a.gkind[0] = c.bc_lo[0];
    a.gkind[1] = c.bc_hi[0];
    for (int k = 0; k < 4; ++k) { a.gval[0][k] = c.bc_lo_val[0][k]; a.gval[1][k] = c.bc_hi_val[0][k]; }
    a.tile_bx = (int)block.x;
    a.tile_nc = 1;
    a.persist = 0;
    a.off32 = ((int64_t)popbytes * g.pop_n < ((int64_t)1 << 32)) ? 1 : 0;
    a.ftau = (float)a.tau;
    a.finv_tau = (float)a.inv_tau;
    a.ffac = 1.0f - 0.5f / a.ftau;
    for (int d = 0; d < 3; ++d) { a.fFmac[d] = (float)a.Fmac[d]; a.fFcol[d] = (float)a.Fcol[d]; }
    return a;
  }

  void ensure_exports() {
    if (!mout) mout = dev_alloc<double>(4 * ncells);
    if (!fout) fout = dev_alloc<double>((size_t)c.q * ncells);
  }

  // ---------------- checkpoint / restart ----------------
  // The dynamic state a step reads: the current populations with their
  // ghost slabs, the pre-collision boundary layer, the IB band macros and
  // exported force, the time-centred force state, and the clock.  Marker
  // geometry, IVC operator and layout are rebuilt from the case.
  struct CkptHeader {
    char magic[8];
    int abi, q, storage, ib;
    int64_t pop_bytes, face_bytes, band_cells;
    double t_new, t_old;
    int istep;
    int slabs_fresh;   // the saved slabs are current (else rebuilt from the boundary layer)
  };

  int64_t face_bytes_total() const {
    int64_t b = 0;
    for (int d = 0; d < c.dim; ++d) {
      if (!nonper[d]) continue;
      const int64_t fn = (d == 0) ? (int64_t)c.n[1] * c.n[2] : (d == 1) ? (int64_t)c.n[0] * c.n[2]
                                                                     : (int64_t)c.n[0] * c.n[1];
      b += 2 * (int64_t)popbytes * fn * face_stride(c.q, (int)popbytes);
    }
    return b;
  }

  int64_t pop_bytes_total() const {
    return (int64_t)popbytes * (g.pop_n + slabs.face_off[slabs.nface] * c.q);
  }

  int64_t checkpoint_size() const {
    int64_t b = (int64_t)sizeof(CkptHeader) + pop_bytes_total() + face_bytes_total();
    if (ib) b += (int64_t)sizeof(double) * (7 * band_cells + 4);
    return b;
  }

  void checkpoint(char* out, int64_t size) {
    if (size < checkpoint_size()) throw StatusError{ILBM_ERR_INVALID, "checkpoint buffer too small"};
    HIPCHK(hipStreamSynchronize(stream));
    CkptHeader h{};
    std::memcpy(h.magic, "ILBMCKP1", 8);
    h.abi = ILBM_ABI_VERSION; h.q = c.q; h.storage = st; h.ib = ib ? 1 : 0;
    h.pop_bytes = pop_bytes_total(); h.face_bytes = face_bytes_total(); h.band_cells = ib ? band_cells : 0;
    h.t_new = t_new; h.t_old = t_old; h.istep = istep;
    h.slabs_fresh = slabs_fresh ? 1 : 0;
    std::memcpy(out, &h, sizeof(h));
    char* p = out + sizeof(h);
    HIPCHK(hipMemcpy(p, gbuf[cur], h.pop_bytes, hipMemcpyDeviceToHost));
    p += h.pop_bytes;
    for (int f = 0; f < 6; ++f) {
      if (!faces.p[f]) continue;
      const int d = f / 2;
      const int64_t fn = (d == 0) ? (int64_t)c.n[1] * c.n[2] : (d == 1) ? (int64_t)c.n[0] * c.n[2]
                                                                     : (int64_t)c.n[0] * c.n[1];
      const int64_t b = (int64_t)popbytes * fn * face_stride(c.q, (int)popbytes);
      HIPCHK(hipMemcpy(p, faces.p[f], b, hipMemcpyDeviceToHost));
      p += b;
    }
    if (ib) {
      HIPCHK(hipMemcpy(p, Mband, sizeof(double) * 4 * band_cells, hipMemcpyDeviceToHost));
      p += sizeof(double) * 4 * band_cells;
      HIPCHK(hipMemcpy(p, Fband, sizeof(double) * 3 * band_cells, hipMemcpyDeviceToHost));
      p += sizeof(double) * 3 * band_cells;
      HIPCHK(hipMemcpy(p, frstate_d, sizeof(double) * 4, hipMemcpyDeviceToHost));
    }
  }

  void restore(const char* in, int64_t size) {
    if (size < (int64_t)sizeof(CkptHeader)) throw StatusError{ILBM_ERR_INVALID, "checkpoint too small"};
    CkptHeader h;
    std::memcpy(&h, in, sizeof(h));
    if (std::memcmp(h.magic, "ILBMCKP1", 8) != 0 || h.abi != ILBM_ABI_VERSION || h.q != c.q || h.storage != st ||
        h.ib != (ib ? 1 : 0) || h.pop_bytes != pop_bytes_total() || h.face_bytes != face_bytes_total() ||
        h.band_cells != (ib ? band_cells : 0) || size < checkpoint_size())
      throw StatusError{ILBM_ERR_INVALID, "checkpoint does not match this case"};
    HIPCHK(hipStreamSynchronize(stream));
    const char* p = in + sizeof(h);
    cur = 0;
    HIPCHK(hipMemcpy(gbuf[cur], p, h.pop_bytes, hipMemcpyHostToDevice));
    p += h.pop_bytes;
    for (int f = 0; f < 6; ++f) {
      if (!faces.p[f]) continue;
      const int d = f / 2;
      const int64_t fn = (d == 0) ? (int64_t)c.n[1] * c.n[2] : (d == 1) ? (int64_t)c.n[0] * c.n[2]
                                                                     : (int64_t)c.n[0] * c.n[1];
      const int64_t b = (int64_t)popbytes * fn * face_stride(c.q, (int)popbytes);
      HIPCHK(hipMemcpy(faces.p[f], p, b, hipMemcpyHostToDevice));
      p += b;
    }
    if (ib) {
      HIPCHK(hipMemcpy(Mband, p, sizeof(double) * 4 * band_cells, hipMemcpyHostToDevice));
      p += sizeof(double) * 4 * band_cells;
      HIPCHK(hipMemcpy(Fband, p, sizeof(double) * 3 * band_cells, hipMemcpyHostToDevice));
      p += sizeof(double) * 3 * band_cells;
      HIPCHK(hipMemcpy(frstate_d, p, sizeof(double) * 4, hipMemcpyHostToDevice));
    }
    t_new = h.t_new; t_old = h.t_old; istep = h.istep;
    slabs_fresh = h.slabs_fresh != 0;
    const double clk[4] = {t_new, (double)istep, (double)istep, 0.0};
    HIPCHK(hipMemcpy(clock_d, clk, sizeof(clk), hipMemcpyHostToDevice));
    state_valid = false;
    mass_host.clear();
    force_host.clear();
  }

  void init(const double* macro_host) {
    double* m0 = dev_alloc<double>(4 * ncells);
    HIPCHK(hipMemcpyAsync(m0, macro_host, sizeof(double) * 4 * ncells, hipMemcpyHostToDevice, stream));
    init_dev(m0);
    dev_free(m0);
  }

  // f = feq of the device macros m0 [4][cells] (InitEquilibrium)
  void init_dev(const double* m0) {
    grow_histories(1, 1);
    const bool small = (int64_t)c.q * ncells <= (int64_t)1 << 27;
    if (small) ensure_exports();
    InitArgs ia{};
    ia.s = step_args(small);
    ia.s.gout = gbuf[cur];
    ia.macro0 = m0;
    initfn(grid, block, stream, ia);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(stream));
    slabs_fresh = false;
    state_valid = small;
    t_new = 0.0;
    t_old = -1.0e100;
    istep = 0;
  }

  void grow_histories(int need_mass, int need_force) {
    if (need_mass > mass_cap) {
      dev_free(mass_dev);
      mass_cap = std::max(need_mass, 2 * mass_cap);
      mass_dev = dev_alloc<double>(mass_cap);
    }
    if (ib && need_force > force_cap) {
      dev_free(force_dev);
      force_cap = std::max(need_force, 2 * force_cap);
      force_dev = dev_alloc<double>((size_t)force_cap * ncol);
    }
  }

  // the near box (IB band grown by one cell; the whole extent of a periodic
  // dimension it would wrap across) and the resources of the overlapped step
  bool split_ready() {
    if (!overlap || timing || !ib || cf || window || band_cells <= 0) return false;
    // below ~1M cells the two cross-stream waits per step cost more than
    // the overlap saves (512x256 D2Q9 FP64: 2.5 -> 1.9 GLUPS with it)
    if (ncells < ((int64_t)1 << 20)) return false;
    if (!stream_b) {
      int per_cu = kFarBlocksPerCu;
      if (const char* e = std::getenv("ILBM_FAR_BLOCKS_PER_CU")) per_cu = std::max(0, std::atoi(e));
      int dev = 0, ncu = 0;
      HIPCHK(hipGetDevice(&dev));
      HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
      far_persist = per_cu * ncu;
      far_after = kFarAfter;
      ring_stream = kRingStream;
      if (const char* e = std::getenv("ILBM_RING_STREAM")) {
        const std::string v(e);
        ring_stream = (v == "far") ? RING_FAR : (v == "own") ? RING_OWN : RING_NEAR;
      }
      ring_on_far = ring_stream == RING_FAR;
      if (const char* e = std::getenv("ILBM_FAR_AFTER")) {
        const std::string v(e);
        far_after = (v == "interp") ? 1 : (v == "export") ? 2 : 0;
      }
      for (int d = 0; d < 3; ++d) {
        int lo = band_lo[d] - 1, hi = band_lo[d] + band_n[d];
        if (d >= c.dim) { lo = 0; hi = 0; }
        else if (c.periodic[d] && (lo < 0 || hi > c.n[d] - 1)) { lo = 0; hi = c.n[d] - 1; }
        else { lo = std::max(lo, 0); hi = std::min(hi, c.n[d] - 1); }
        near_lo[d] = lo;
        near_n[d] = hi - lo + 1;
      }
      grid_near = dim3((near_n[0] + block.x - 1) / block.x, (near_n[1] + block.y - 1) / block.y, near_n[2]);
      near_off_walls = true;
      for (int d = 0; d < c.dim; ++d)
        if (nonper[d] && (near_lo[d] <= 0 || near_lo[d] + near_n[d] >= c.n[d])) near_off_walls = false;
      nblocks_near = (int64_t)grid_near.x * grid_near.y * grid_near.z;
      // the ring box: the near box grown by one cell, the whole extent of a
      // periodic dimension the near box already spans
      ring_ok = near_off_walls;
      for (int d = 0; d < 3; ++d) {
        int lo = near_lo[d] - 1, hi = near_lo[d] + near_n[d];
        if (d >= c.dim) { lo = 0; hi = 0; }
        else if (near_n[d] >= c.n[d]) { lo = 0; hi = c.n[d] - 1; }
        else if (c.periodic[d] && (lo < 0 || hi > c.n[d] - 1)) ring_ok = false;
        else if (nonper[d] && (lo <= 0 || hi >= c.n[d] - 1)) ring_ok = false;
        ring_lo[d] = std::max(lo, 0);
        ring_n[d] = std::min(hi, c.n[d] - 1) - ring_lo[d] + 1;
      }
      grid_ring = dim3((ring_n[0] + block.x - 1) / block.x, (ring_n[1] + block.y - 1) / block.y, ring_n[2]);
      nblocks_ring = (int64_t)grid_ring.x * grid_ring.y * grid_ring.z;
      const char* two = std::getenv("ILBM_OVERLAP_TWO_REGIONS");
      if (two && two[0] == '1') ring_ok = false;
      HIPCHK(hipStreamCreateWithFlags(&stream_b, hipStreamNonBlocking));
      for (hipEvent_t* e : {&ev_a, &ev_b})
        HIPCHK(hipEventCreateWithFlags(e, hipEventDisableTiming));
      if (ring_ok && ring_stream == RING_OWN) {
        HIPCHK(hipStreamCreateWithFlags(&stream_d, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&ev_d, hipEventDisableTiming));
      }
      for (hipEvent_t& e : xev) HIPCHK(hipEventCreateWithFlags(&e, step_event_flags()));
      if (async_record()) {
        HIPCHK(hipStreamCreateWithFlags(&stream_c, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&ev_c, hipEventDisableTiming));
        F_ring[0] = F_d;
        fp_ring[0] = fpartial;
        for (int k = 0; k < kRecRing; ++k) {
          if (k > 0) {
            F_ring[k] = dev_alloc<double>(3 * nm);
            fp_ring[k] = dev_alloc<double>(fpartial_n);
          }
        }
      }
    }
    return true;
  }

  // the record can run apart from the forcing: one iteration, a fixed body
  bool async_record() const {
    const int nloop = (c.coupling_method == 1) ? 1 : std::max(c.loop_ns, 1);
    return nloop == 1 && !dev_kin && !moving;
  }

  // One step of an IB level on two streams (a third takes the force record
  // of a fixed body).  Stream B: ghost slabs, the far cells, then a wait for
  // the near box (the sources of the next far part) and, when the ring of
  // mass partials is full, Sum(Rho) of its steps.  Stream A: IB forcing
  // (reads the band macros of the last near box, adds the Guo term to band
  // cells, which no far cell reads), then the near box after the ghost
  // slabs.  When the near box keeps off the non-periodic faces, stream B
  // fills the next step's ghost slabs right after the far cells (they read
  // only far cells and their boundary layer), before waiting for the near
  // box.
  void split_step(int s, double time, bool want, bool more) {
    StepArgs sf = step_args(want);
    sf.sub_mode = 1;
    sf.persist = far_persist;
    for (int d = 0; d < 3; ++d) { sf.sub_lo[d] = near_lo[d]; sf.sub_n[d] = near_n[d]; }
    sf.partial = mring + (int64_t)mring_used * nblocks;
    StepArgs sn = sf;
    sn.persist = 0;
    sn.sub_mode = 2;
    sn.partial = nring + (int64_t)mring_used * nblocks_near;
    const int64_t npad = wrap ? 0 : slabs.face_off[slabs.nface];
    const dim3 ggrid((unsigned)((npad + kGhostBlock - 1) / kGhostBlock));
    // The near box waits for this signal: the ghost slabs it may read and
    // the far cells of the last step, which read what it overwrites.
    uint64_t ghost_done = ahead_signal;
    if (!ghost_ahead) {
      if (npad && !slabs_fresh) ghostfn(ggrid, dim3(kGhostBlock), stream_b, ghost_args());
      ghost_done = xsignal(stream_b, X_GHOST);
    }
    ghost_ahead = false;
    stepfn(grid, block, stream_b, sf, nullptr, nullptr);
    ring_slot = async_record() ? s % kRecRing : -1;
    ib_update_forcing(time, s);
    ring_slot = -1;
    xwait(stream, X_GHOST, ghost_done);
    stepfn(grid_near, block, stream, sn, nullptr, nullptr);
    const uint64_t near_done = xsignal(stream, X_NEAR);
    if (more && (near_off_walls || !npad)) {
      // the next step's slabs (from the boundary layer of the far cells
      // just written) and its signal to the near box, ahead of the wait
      if (npad && !gfuse) {
        cur ^= 1;
        const GhostArgs ga = ghost_args();
        cur ^= 1;
        ghostfn(ggrid, dim3(kGhostBlock), stream_b, ga);
      }
      ahead_signal = xsignal(stream_b, X_GHOST);
      ghost_ahead = true;
    }
    xwait(stream_b, X_NEAR, near_done);
    ring_step(stream_b, s, more, true);
  }

  // One step in three regions (ring_ok).  Stream B: a wait for the ring of
  // the last step, the far cells (outside the ring box), the next step's
  // ghost slabs.  Stream A: IB forcing (reads the band macros of the last
  // near box, adds the Guo term to band cells, which no far or ring cell
  // reads), the near box, a wait for the far cells of the last step, then
  // the ring (the cells of the ring box outside the near box).  The far
  // cells of step s read, outside the near box, what the far cells and the
  // ring of step s - 1 wrote, and overwrite what those read; the ring of
  // step s reads what the far cells of step s - 1 wrote next to it and
  // overwrites what they read; the near box and the ring read and write
  // only cells of the ring box grown by one, and never the cells of the
  // other buffer that the far cells of the same step read or write.  So the
  // far kernels run back to back while the IB forcing, the near box and the
  // ring of the same step run beside them.
  struct PendingMass { int base = 0, n = 0, s = -1; };
  PendingMass pend{};

  void flush_mass(hipStream_t st) {
    if (pend.n <= 0) return;
    hipLaunchKernelGGL(k_mass_ring, dim3(pend.n), dim3(1024), 0, st,
                       (const double*)(mring + (int64_t)pend.base * nblocks), nblocks, nblocks,
                       (const double*)(nring + (int64_t)pend.base * nblocks_near), nblocks_near, nblocks_near,
                       (const double*)(rring + (int64_t)pend.base * nblocks_ring), nblocks_ring, nblocks_ring,
                       mass_dev + pend.s + 1 - pend.n);
    pend = PendingMass{};
  }

  uint64_t far_sig[2] = {0, 0}, ring_sig[2] = {0, 0}, near_sig[2] = {0, 0};

  void split_step3(int s, double time, bool want, bool more, bool first) {
    StepArgs sf = step_args(want);
    sf.sub_mode = 1;
    sf.persist = far_persist;
    for (int d = 0; d < 3; ++d) { sf.sub_lo[d] = ring_lo[d]; sf.sub_n[d] = ring_n[d]; }
    const int slot = mring_used;
    sf.partial = mring + (int64_t)slot * nblocks;
    StepArgs sn = sf;
    sn.persist = 0;
    sn.sub_mode = 2;
    for (int d = 0; d < 3; ++d) { sn.sub_lo[d] = near_lo[d]; sn.sub_n[d] = near_n[d]; }
    sn.partial = nring + (int64_t)slot * nblocks_near;
    StepArgs sr = sf;
    sr.persist = 0;
    sr.sub_mode = 3;
    for (int d = 0; d < 3; ++d) { sr.in_lo[d] = near_lo[d]; sr.in_n[d] = near_n[d]; }
    sr.partial = rring + (int64_t)slot * nblocks_ring;
    const int64_t npad = wrap ? 0 : slabs.face_off[slabs.nface];
    const dim3 ggrid((unsigned)((npad + kGhostBlock - 1) / kGhostBlock));
    if (stream_d) {
      // Stream A: IB forcing, a wait for the ring of the last step, the near
      // box.  Stream B: a wait for the ring of the last step, the far cells
      // (after the forcing's stage far_after).  Stream D: waits for the far
      // cells and the near box of the last step, the ring, the Sum(Rho) of a
      // finished half of the slot ring, the ring's signal.  The ring of step
      // s reads only cells of step s - 1's output and overwrites only what
      // step s - 1's launches read, so it runs beside step s's far cells and
      // near box; both of those of step s + 1 wait for it.
      ring_slot = async_record() ? s % kRecRing : -1;
      ib_mark = far_after;
      ib_update_forcing(time, s);
      ib_mark = 0;
      ring_slot = -1;
      if (!first) xwait(stream_b, X_RING0 + ((s - 1) & 1), ring_sig[(s - 1) & 1]);
      if (!ghost_ahead && npad && !slabs_fresh) ghostfn(ggrid, dim3(kGhostBlock), stream_b, ghost_args());
      ghost_ahead = false;
      if (far_after) xwait(stream_b, X_IB, ib_sig);
      stepfn(grid, block, stream_b, sf, nullptr, nullptr);
      far_sig[s & 1] = xsignal(stream_b, X_FAR0 + (s & 1));
      if (more && npad && !gfuse) {
        cur ^= 1;
        const GhostArgs ga = ghost_args();
        cur ^= 1;
        ghostfn(ggrid, dim3(kGhostBlock), stream_b, ga);
        ghost_ahead = true;
      }
      if (!first) {
        xwait(stream_d, X_FAR0 + ((s - 1) & 1), far_sig[(s - 1) & 1]);
        xwait(stream_d, X_NEAR0 + ((s - 1) & 1), near_sig[(s - 1) & 1]);
      }
      stepfn(grid_ring, block, stream_d, sr, nullptr, nullptr);
      flush_mass(stream_d);
      ring_sig[s & 1] = xsignal(stream_d, X_RING0 + (s & 1));
      if (!first) xwait(stream, X_RING0 + ((s - 1) & 1), ring_sig[(s - 1) & 1]);
      stepfn(grid_near, block, stream, sn, nullptr, nullptr);
      near_sig[s & 1] = xsignal(stream, X_NEAR0 + (s & 1));
      ++mring_used;
      const int n = mring_used - mring_base;
      if (n == mring_slots / 2 || !more) {
        pend.base = mring_base;
        pend.n = n;
        pend.s = s;
        mring_base = mring_used;
        if (mring_used == mring_slots || !more) mring_used = mring_base = 0;
      }
      return;
    }
    if (ring_on_far) {
      // Stream A: IB forcing, a wait for the ring of the last step, the near
      // box.  Stream B: the far cells (after the forcing's stage far_after),
      // a wait for the near box of the last step, the ring, the Sum(Rho) of
      // a finished half of the slot ring, the ring's signal.  The ring of
      // step s reads what the far cells, the ring and the near box of step
      // s - 1 wrote next to it and overwrites what they read; the near box
      // of step s + 1 reads the ring's cells; the IB forcing reads only the
      // band, inside the near box.
      ring_slot = async_record() ? s % kRecRing : -1;
      ib_mark = far_after;
      ib_update_forcing(time, s);
      ib_mark = 0;
      ring_slot = -1;
      if (!ghost_ahead && npad && !slabs_fresh) ghostfn(ggrid, dim3(kGhostBlock), stream_b, ghost_args());
      ghost_ahead = false;
      if (far_after) xwait(stream_b, X_IB, ib_sig);
      stepfn(grid, block, stream_b, sf, nullptr, nullptr);
      if (more && npad && !gfuse) {
        cur ^= 1;
        const GhostArgs ga = ghost_args();
        cur ^= 1;
        ghostfn(ggrid, dim3(kGhostBlock), stream_b, ga);
        ghost_ahead = true;
      }
      if (!first) xwait(stream_b, X_NEAR0 + ((s - 1) & 1), near_sig[(s - 1) & 1]);
      stepfn(grid_ring, block, stream_b, sr, nullptr, nullptr);
      // the pending half (steps up to s - 1: their near boxes are waited for
      // above) before the signal, so a near box that reuses its slots waits
      flush_mass(stream_b);
      ring_sig[s & 1] = xsignal(stream_b, X_RING0 + (s & 1));
      if (!first) xwait(stream, X_RING0 + ((s - 1) & 1), ring_sig[(s - 1) & 1]);
      stepfn(grid_near, block, stream, sn, nullptr, nullptr);
      near_sig[s & 1] = xsignal(stream, X_NEAR0 + (s & 1));
      ++mring_used;
      const int n = mring_used - mring_base;
      if (n == mring_slots / 2 || !more) {
        pend.base = mring_base;
        pend.n = n;
        pend.s = s;
        mring_base = mring_used;
        if (mring_used == mring_slots || !more) mring_used = mring_base = 0;
      }
      return;
    }
    // with far_after, stream A's IB forcing is enqueued first and signals
    // the stage the far cells wait for (the far cells read and write
    // nothing the forcing touches; the wait only orders the two)
    if (far_after) {
      ring_slot = async_record() ? s % kRecRing : -1;
      ib_mark = far_after;
      ib_update_forcing(time, s);
      ib_mark = 0;
      ring_slot = -1;
    }
    // stream B
    if (!first) xwait(stream_b, X_RING0 + ((s - 1) & 1), ring_sig[(s - 1) & 1]);
    flush_mass(stream_b);
    if (!ghost_ahead && npad && !slabs_fresh) ghostfn(ggrid, dim3(kGhostBlock), stream_b, ghost_args());
    ghost_ahead = false;
    if (far_after) xwait(stream_b, X_IB, ib_sig);
    stepfn(grid, block, stream_b, sf, nullptr, nullptr);
    far_sig[s & 1] = xsignal(stream_b, X_FAR0 + (s & 1));
    if (more && npad && !gfuse) {
      cur ^= 1;
      const GhostArgs ga = ghost_args();
      cur ^= 1;
      ghostfn(ggrid, dim3(kGhostBlock), stream_b, ga);
      ghost_ahead = true;
    }
    // stream A
    if (!far_after) {
      ring_slot = async_record() ? s % kRecRing : -1;
      ib_update_forcing(time, s);
      ring_slot = -1;
    }
    stepfn(grid_near, block, stream, sn, nullptr, nullptr);
    if (!first) xwait(stream, X_FAR0 + ((s - 1) & 1), far_sig[(s - 1) & 1]);
    stepfn(grid_ring, block, stream, sr, nullptr, nullptr);
    ring_sig[s & 1] = xsignal(stream, X_RING0 + (s & 1));
    // Sum(Rho) slots: a finished half is reduced on stream B once the ring
    // of its last step is done (the next step's wait, or the advance's end)
    ++mring_used;
    const int n = mring_used - mring_base;
    if (n == mring_slots / 2 || !more) {
      pend.base = mring_base;
      pend.n = n;
      pend.s = s;
      mring_base = mring_used;
      if (mring_used == mring_slots || !more) mring_used = mring_base = 0;
    }
  }

  // After step s wrote ring slot mring_used: when half of the ring is
  // filled (or the advance ends) its steps' Sum(Rho) are reduced on `st`.
  // The halves alternate, so a near box queued ahead of the reduction
  // writes into the other half.
  void ring_step(hipStream_t st, int s, bool more, bool with_near) {
    ++mring_used;
    const int n = mring_used - mring_base;
    if (n == mring_slots / 2 || !more) {
      const bool with_ring = with_near && rring;
      hipLaunchKernelGGL(k_mass_ring, dim3(n), dim3(1024), 0, st, (const double*)(mring + (int64_t)mring_base * nblocks),
                         nblocks, nblocks, (const double*)(with_near ? nring + (int64_t)mring_base * nblocks_near : nullptr),
                         with_near ? nblocks_near : (int64_t)0, with_near ? nblocks_near : (int64_t)0,
                         (const double*)(with_ring ? rring + (int64_t)mring_base * nblocks_ring : nullptr),
                         with_ring ? nblocks_ring : (int64_t)0, with_ring ? nblocks_ring : (int64_t)0,
                         mass_dev + s + 1 - n);
      mring_base = mring_used;
      if (mring_used == mring_slots || !more) mring_used = mring_base = 0;
    }
  }

  // the record numbers only document which record a wait expects: every
  // wait below expects the latest one of its kind
  uint64_t xsignal(hipStream_t st, int k) {
    HIPCHK(hipEventRecord(xev[k], st));
    return ++xseq[k];
  }
  void xwait(hipStream_t st, int k, uint64_t v) {
    if (v != xseq[k]) throw StatusError{ILBM_ERR_STATE, "stream signal out of order"};
    HIPCHK(hipStreamWaitEvent(st, xev[k], 0));
  }

  // ring slots for the mass partials of this level (and its near box)
  void ensure_mring(bool split) {
    const bool three = split && ring_ok;
    const int64_t per = nblocks + (split ? nblocks_near : 0) + (three ? nblocks_ring : 0);
    if (mring && (!split || nring) && (!three || rring)) return;
    dev_free(mring); dev_free(nring); dev_free(rring);
    mring = nring = rring = nullptr;
    // two halves of at most 32 steps each, at most 64 MiB
    mring_slots = 2 * (int)std::max<int64_t>(1, std::min<int64_t>(32, ((int64_t)32 << 20) / (8 * per)));
    mring = dev_alloc<double>((size_t)nblocks * mring_slots);
    if (split) nring = dev_alloc<double>((size_t)nblocks_near * mring_slots);
    if (three) rring = dev_alloc<double>((size_t)nblocks_ring * mring_slots);
    mring_used = mring_base = 0;
  }

  void advance(int nsteps, bool want_state) {
    if (nsteps < 0) throw StatusError{ILBM_ERR_INVALID, "nsteps < 0"};
    if (nsteps == 0) return;
    if (want_state) ensure_exports();
    grow_histories(nsteps, nsteps);
    const bool split = split_ready();
    ensure_mring(split);
    if (split) {
      HIPCHK(hipEventRecord(ev_a, stream));
      HIPCHK(hipStreamWaitEvent(stream_b, ev_a, 0));
      if (stream_d) HIPCHK(hipStreamWaitEvent(stream_d, ev_a, 0));
    }
    if (timing) {
      while ((int)ev_start.size() < nsteps) {
        hipEvent_t a, b;
        HIPCHK(hipEventCreate(&a));
        HIPCHK(hipEventCreate(&b));
        ev_start.push_back(a);
        ev_stop.push_back(b);
      }
    }
    state_valid = false;
    for (int s = 0; s < nsteps; ++s) {
      const double time = t_new;  // timeStepWithSubcycling, AmrCoreLBM.cpp:1730-1733
      t_old = t_new;
      t_new += dt;
      if (split) {
        if (ring_ok) split_step3(s, time, want_state && s == nsteps - 1, s + 1 < nsteps, s == 0);
        else split_step(s, time, want_state && s == nsteps - 1, s + 1 < nsteps);
        slabs_fresh = gfuse;
        cur ^= 1;
        ++istep;
        HIPCHK(hipGetLastError());
        continue;
      }
      const int64_t npad = wrap ? 0 : slabs.face_off[slabs.nface];
      if (ib) {
        rec_to_ghost = true;
        ib_update_forcing(time, s);
        rec_to_ghost = false;
      }
      if (npad && !slabs_fresh) {
        GhostArgs ga = ghost_args();
        if (rec_held) {
          // one launch fewer per step (512x256 cylinder: 5 -> 4 launches)
          const unsigned ng = (unsigned)((npad + kSumsBlock - 1) / kSumsBlock);
          ghostrecfn(ng + 1, stream, ga, ng, held_a, held_r, held_nblk);
          rec_held = false;
        } else {
          ghostfn(dim3((unsigned)((npad + kGhostBlock - 1) / kGhostBlock)), dim3(kGhostBlock), stream, ga);
        }
      }
      // a record still held goes with the stream-collide launch as its last
      // block, unless the launch is timed (then on its own, before it)
      const bool rec_with_step = rec_held && !timing && block.x * block.y == (unsigned)kSumsBlock;
      if (rec_held && !rec_with_step) {
        hipLaunchKernelGGL(k_sums_record, dim3(1), dim3(kSumsBlock), 0, stream, held_a, held_r, held_nblk);
        rec_held = false;
      }
      const bool last = (s == nsteps - 1);
      StepArgs sa = step_args(want_state && last);
      sa.partial = mring + (int64_t)mring_used * nblocks;
      if (rec_with_step) {
        steprecfn(grid, block, stream, sa, held_a, held_r, held_nblk);
        rec_held = false;
      } else {
        // timing: the dispatch records the kernel's own begin / end
        stepfn(grid, block, stream, sa, timing ? ev_start[s] : nullptr, timing ? ev_stop[s] : nullptr);
      }
      ring_step(stream, s, !last, false);
      slabs_fresh = gfuse;
      cur ^= 1;
      ++istep;
      HIPCHK(hipGetLastError());
    }
    if (split) {
      if (ring_ok) {
        // the last step's Sum(Rho) once its ring (and near box) is done
        const int last = nsteps - 1;
        if (stream_d) {
          xwait(stream_d, X_NEAR0 + (last & 1), near_sig[last & 1]);
          xwait(stream_d, X_FAR0 + (last & 1), far_sig[last & 1]);
          flush_mass(stream_d);
          HIPCHK(hipEventRecord(ev_d, stream_d));
          HIPCHK(hipStreamWaitEvent(stream, ev_d, 0));
        } else {
          if (ring_on_far) xwait(stream_b, X_NEAR0 + (last & 1), near_sig[last & 1]);
          else xwait(stream_b, X_RING0 + (last & 1), ring_sig[last & 1]);
          flush_mass(stream_b);
        }
      }
      HIPCHK(hipEventRecord(ev_b, stream_b));
      HIPCHK(hipStreamWaitEvent(stream, ev_b, 0));
      if (stream_c) {
        HIPCHK(hipEventRecord(ev_c, stream_c));
        HIPCHK(hipStreamWaitEvent(stream, ev_c, 0));
        // the marker forces of the last step where the getters read them
        const int last = (nsteps - 1) % kRecRing;
        if (last != 0)
          HIPCHK(hipMemcpyAsync(F_d, F_ring[last], sizeof(double) * 3 * nm, hipMemcpyDeviceToDevice, stream));
      }
    }
    std::vector<double> mh(nsteps);
    HIPCHK(hipMemcpyAsync(mh.data(), mass_dev, sizeof(double) * nsteps, hipMemcpyDeviceToHost, stream));
    std::vector<double> fh;
    if (ib) {
      fh.resize((size_t)nsteps * ncol);
      HIPCHK(hipMemcpyAsync(fh.data(), force_dev, sizeof(double) * nsteps * ncol, hipMemcpyDeviceToHost, stream));
    }
    HIPCHK(hipStreamSynchronize(stream));
    mass_host.insert(mass_host.end(), mh.begin(), mh.end());
    force_host.insert(force_host.end(), fh.begin(), fh.end());
    if (timing) {
      double tot = 0.0;
      for (int s = 0; s < nsteps; ++s) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, ev_start[s], ev_stop[s]));
        tot += ms;
      }
      last_avg_ms = tot / nsteps;
      last_launches = nsteps;
    }
    state_valid = want_state;
  }
};

}  // namespace ilbm

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
struct ilbm_level {
  ilbm::Level L;
};

static thread_local std::string g_last_error;

template <class F>
static int guarded(F&& f) {
  try {
    f();
    return ILBM_OK;
  } catch (const ilbm::StatusError& e) {
    g_last_error = e.msg;
    return e.status;
  } catch (const ilbm::HipError& e) {
    g_last_error = hipGetErrorString(e.e);
    return ILBM_ERR_HIP;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation";
    return ILBM_ERR_ALLOC;
  }
}

extern "C" {

int ilbm_abi_version(void) { return ILBM_ABI_VERSION; }

int ilbm_case_desc_size(void) { return (int)sizeof(ilbm_case_desc); }

const char* ilbm_status_string(int status) {
  switch (status) {
    case ILBM_OK: return "ok";
    case ILBM_ERR_INVALID: return "invalid argument";
    case ILBM_ERR_UNSUPPORTED: return "unsupported";
    case ILBM_ERR_HIP: return "HIP runtime error";
    case ILBM_ERR_ALLOC: return "allocation failed";
    case ILBM_ERR_STATE: return "invalid state";
    case ILBM_ERR_FULL: return "batch full";
    default: return "unknown";
  }
}

const char* ilbm_last_error(void) { return g_last_error.c_str(); }

int ilbm_level_create(const ilbm_case_desc* desc, const double* init_macro, ilbm_level** out) {
  if (!desc || !init_macro || !out) return ILBM_ERR_INVALID;
  *out = nullptr;
  ilbm_level* h = new (std::nothrow) ilbm_level();
  if (!h) return ILBM_ERR_ALLOC;
  const int st = guarded([&] {
    ilbm::Level& L = h->L;
    L.c = *desc;
    L.validate();
    HIPCHK(hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking));

    L.setup_layout();
    L.setup_ib();
    L.init(init_macro);
  });
  if (st != ILBM_OK) {
    delete h;
    return st;
  }
  *out = h;
  return ILBM_OK;
}

int ilbm_level_destroy(ilbm_level* h) {
  if (!h) return ILBM_ERR_INVALID;
  if (h->L.stream) (void)hipStreamSynchronize(h->L.stream);

  delete h;
  return ILBM_OK;
}

int ilbm_level_advance(ilbm_level* h, int nsteps, int want_state) {
  if (!h) return ILBM_ERR_INVALID;
  return guarded([&] { h->L.advance(nsteps, want_state != 0); });
}

int ilbm_level_synchronize(ilbm_level* h) {
  if (!h) return ILBM_ERR_INVALID;
  return guarded([&] {
    HIPCHK(hipStreamSynchronize(h->L.stream));
  });
}

int ilbm_level_get_f(ilbm_level* h, double* out) {
  if (!h || !out) return ILBM_ERR_INVALID;
  if (!h->L.state_valid) return ILBM_ERR_STATE;
  return guarded([&] {
    HIPCHK(hipMemcpy(out, h->L.fout, sizeof(double) * h->L.c.q * h->L.ncells, hipMemcpyDeviceToHost));
  });
}

int ilbm_level_get_macro(ilbm_level* h, double* out) {
  if (!h || !out) return ILBM_ERR_INVALID;
  if (!h->L.state_valid) return ILBM_ERR_STATE;
  return guarded([&] {
    HIPCHK(hipMemcpy(out, h->L.mout, sizeof(double) * 4 * h->L.ncells, hipMemcpyDeviceToHost));
  });
}

int ilbm_level_get_forcing(ilbm_level* h, double* out) {
  if (!h || !out) return ILBM_ERR_INVALID;
  return guarded([&] {
    ilbm::Level& L = h->L;
    std::fill(out, out + 3 * L.ncells, 0.0);
    if (!L.ib) return;
    std::vector<double> fb(3 * L.band_cells);
    HIPCHK(hipMemcpy(fb.data(), L.Fband, sizeof(double) * 3 * L.band_cells, hipMemcpyDeviceToHost));
    for (int bk = 0; bk < L.band_n[2]; ++bk)
      for (int bj = 0; bj < L.band_n[1]; ++bj)
        for (int bi = 0; bi < L.band_n[0]; ++bi) {
          const int64_t b = ((int64_t)bk * L.band_n[1] + bj) * L.band_n[0] + bi;
          const int i = L.band_lo[0] + bi, j = L.band_lo[1] + bj, k = L.band_lo[2] + bk;
          const int64_t v = ((int64_t)k * L.c.n[1] + j) * L.c.n[0] + i;
          for (int cc = 0; cc < 3; ++cc) out[cc * L.ncells + v] = fb[cc * L.band_cells + b];
        }
  });
}

}  // extern "C"

static int mass_history_of(const ilbm::Level& L, double* out, int max_rows, int* n_out) {
  if (!n_out) return ILBM_ERR_INVALID;
  const auto& v = L.mass_host;
  *n_out = (int)v.size();
  if (out) std::copy(v.begin(), v.begin() + std::min<size_t>(v.size(), (size_t)std::max(max_rows, 0)), out);
  return ILBM_OK;
}

static int force_history_of(const ilbm::Level& L, double* out, int max_rows, int* n_rows, int* n_cols) {
  if (!n_rows || !n_cols) return ILBM_ERR_INVALID;
  *n_cols = L.ncol;
  *n_rows = L.ncol ? (int)(L.force_host.size() / L.ncol) : 0;
  if (out && L.ncol) {
    const size_t n = std::min<size_t>(L.force_host.size(), (size_t)std::max(max_rows, 0) * L.ncol);
    std::copy(L.force_host.begin(), L.force_host.begin() + n, out);
  }
  return ILBM_OK;
}

extern "C" {

int ilbm_level_mass_history(ilbm_level* h, double* out, int max_rows, int* n_out) {
  if (!h) return ILBM_ERR_INVALID;
  return mass_history_of(h->L, out, max_rows, n_out);
}

int ilbm_level_force_history(ilbm_level* h, double* out, int max_rows, int* n_rows, int* n_cols) {
  if (!h) return ILBM_ERR_INVALID;
  return force_history_of(h->L, out, max_rows, n_rows, n_cols);
}

int ilbm_level_num_markers(ilbm_level* h, int* out) {
  if (!h || !out) return ILBM_ERR_INVALID;
  *out = h->L.ib ? h->L.nm : 0;
  return ILBM_OK;
}

int ilbm_level_marker_dv(ilbm_level* h, double* out) {
  if (!h || !out) return ILBM_ERR_INVALID;
  *out = h->L.ib ? h->L.mg.dv : 0.0;
  return ILBM_OK;
}

int ilbm_level_marker_state(ilbm_level* h, double* xyz, double* uvw, double* fxyz) {
  if (!h) return ILBM_ERR_INVALID;
  ilbm::Level& L = h->L;
  if (!L.ib) return ILBM_ERR_STATE;
  return guarded([&] {
    const int nm = L.nm;
    std::vector<double> t(3 * nm);
    auto tr = [&](const double* dev, double* out) {
      if (!out) return;
      HIPCHK(hipMemcpy(t.data(), dev, sizeof(double) * 3 * nm, hipMemcpyDeviceToHost));
      for (int m = 0; m < nm; ++m)
        for (int d = 0; d < 3; ++d) out[3 * m + d] = t[d * nm + m];
    };
    HIPCHK(hipStreamSynchronize(L.stream));
    tr(L.xm_d, xyz);
    tr(L.U_d, uvw);
    tr(L.F_d, fxyz);
  });
}

int ilbm_level_marker_force_sum(ilbm_level* h, double out[3]) {
  if (!h || !out) return ILBM_ERR_INVALID;
  if (!h->L.ib) return ILBM_ERR_STATE;
  return guarded([&] {
    double s[8];
    HIPCHK(hipStreamSynchronize(h->L.stream));
    HIPCHK(hipMemcpy(s, h->L.sums_d, sizeof(double) * 8, hipMemcpyDeviceToHost));
    out[0] = s[0]; out[1] = s[1]; out[2] = s[2];
  });
}

int ilbm_level_marker_torque(ilbm_level* h, double out[3]) {
  if (!h || !out) return ILBM_ERR_INVALID;
  if (!h->L.ib) return ILBM_ERR_STATE;
  return guarded([&] {
    double s[8];
    HIPCHK(hipStreamSynchronize(h->L.stream));
    HIPCHK(hipMemcpy(s, h->L.sums_d, sizeof(double) * 8, hipMemcpyDeviceToHost));
    out[0] = s[3]; out[1] = s[4]; out[2] = s[5];
  });
}

int ilbm_level_checkpoint_size(ilbm_level* h, int64_t* bytes) {
  if (!h || !bytes) return ILBM_ERR_INVALID;
  *bytes = h->L.checkpoint_size();
  return ILBM_OK;
}

int ilbm_level_checkpoint(ilbm_level* h, void* host_buf, int64_t bytes) {
  if (!h || !host_buf) return ILBM_ERR_INVALID;
  return guarded([&] { h->L.checkpoint(static_cast<char*>(host_buf), bytes); });
}

int ilbm_level_restore(ilbm_level* h, const void* host_buf, int64_t bytes) {
  if (!h || !host_buf) return ILBM_ERR_INVALID;
  return guarded([&] { h->L.restore(static_cast<const char*>(host_buf), bytes); });
}

int ilbm_level_info(ilbm_level* h, double out[8]) {
  if (!h || !out) return ILBM_ERR_INVALID;
  const ilbm::Level& L = h->L;
  out[0] = L.tau; out[1] = L.dt; out[2] = L.t_new; out[3] = L.istep;
  out[4] = (double)L.band_cells; out[5] = (double)L.popbytes;
  out[6] = (L.am == ilbm::kAddrWrap) ? 0.0 : (L.am == ilbm::kAddrWall64) ? 1.0 : (L.am == ilbm::kAddrWall32) ? 2.0
           : L.gfuse ? 4.0 : 3.0;
  out[7] = (double)L.nblocks;
  return ILBM_OK;
}

int ilbm_level_kernel_time(ilbm_level* h, double* avg_ms, int* launches) {
  if (!h || !avg_ms || !launches) return ILBM_ERR_INVALID;
  *avg_ms = h->L.last_avg_ms;
  *launches = h->L.last_launches;
  return ILBM_OK;
}

int ilbm_level_set_timing(ilbm_level* h, int enable) {
  if (!h) return ILBM_ERR_INVALID;
  h->L.timing = enable != 0;
  return ILBM_OK;
}

int ilbm_level_set_overlap(ilbm_level* h, int enable) {
  if (!h) return ILBM_ERR_INVALID;
  h->L.overlap = enable != 0;
  return ILBM_OK;
}

// Plain device-to-device copy, used only to measure the achievable HBM
// bandwidth and to calibrate rocprof byte counters against a kernel whose
// traffic is known (bytes read = bytes written = `bytes`).
typedef float v4f __attribute__((ext_vector_type(4)));
// One 16-byte element per thread, non-temporal load and store (the fastest
// of the copy forms tried on MI355X: 6.7 TB/s at 2 GiB).
__global__ void __launch_bounds__(256) k_copy_probe(const v4f* __restrict__ src,
                                                    v4f* __restrict__ dst, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) __builtin_nontemporal_store(__builtin_nontemporal_load(src + t), dst + t);
}

int ilbm_copy_probe(int64_t bytes, int iters, double* avg_ms) {
  if (bytes <= 0 || iters <= 0 || !avg_ms) return ILBM_ERR_INVALID;
  return guarded([&] {
    const int64_t n = bytes / 16;
    void* a = ilbm::dev_alloc_bytes((size_t)n * 16);
    void* b = ilbm::dev_alloc_bytes((size_t)n * 16);
    hipEvent_t e0, e1;
    HIPCHK(hipMemset(a, 0, (size_t)n * 16));
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    const unsigned grid = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(256), 0, nullptr, (const v4f*)a, (v4f*)b, n);
    HIPCHK(hipEventRecord(e0, nullptr));
    for (int it = 0; it < iters; ++it)
      hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(256), 0, nullptr, (const v4f*)a, (v4f*)b, n);
    HIPCHK(hipEventRecord(e1, nullptr));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    *avg_ms = ms / iters;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ilbm::dev_free(a);
    ilbm::dev_free(b);
  });
}

}  // extern "C"

#include "ilbm_batch_impl.h"
#include "ilbm_box_impl.h"
#include "ilbm_box_ib_impl.h"
#include "ilbm_amr_impl.h"
