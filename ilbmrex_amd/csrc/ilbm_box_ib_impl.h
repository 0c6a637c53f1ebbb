// Per-box immersed boundary: IBMarkerDF (Source/IBM/IBMarkerDF.H:68-91) on
// caller-owned fabs, compiled as part of ilbm_level.hip.
//
// An ilbm_ib holds the marker set of one level and the band around its
// stencils (the IB half of a resident level, without populations).  Each
// update_forcing call gathers rho, u, v, w of the band cells from the
// caller's fabs, runs the stages of IBMarkerDF::update_forcing
// (IBMarkerDF.cpp:1522-1917: markers at `time`, interpolation, explicit /
// explicit-diag / IVC forcing, spreading, marker sums and moments, velocity
// correction over loop_ns iterations) and writes F = dt * rho *
// EulerForceTotal into every valid cell of the caller's force fabs (zero
// outside the band), as the reference's export loop does (:1889-1912).
//
// Ghost values: interpolation near a physical boundary reads the FillPatch
// values of the BC functors (macroBcFill of the case's BCs), which is what
// the reference's rho/u/v/w hold there (update_forcing is called on
// FillPatchMacro's result, AdvanceLBMAtLevel.cpp:60-90); periodic images come
// from the valid cells, as FillBoundary gives them.
#pragma once

namespace ilbm {

__global__ void k_ib_gather_box(Fab rho, Fab u, Fab v, Fab w, int has_w, IbArgs a, double* Mband) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= a.band_cells) return;
  const int i = a.band_lo[0] + (int)(b % a.band_n[0]) + a.ioff[0];
  const int j = a.band_lo[1] + (int)((b / a.band_n[0]) % a.band_n[1]) + a.ioff[1];
  const int k = a.band_lo[2] + (int)(b / ((int64_t)a.band_n[0] * a.band_n[1])) + a.ioff[2];
  const int64_t bc = a.band_cells;
  Mband[b] = rho.at(i, j, k, 0);
  Mband[bc + b] = u.at(i, j, k, 0);
  Mband[2 * bc + b] = v.at(i, j, k, 0);
  Mband[3 * bc + b] = has_w ? w.at(i, j, k, 0) : 0.0;
}

// F = dt * rho * EulerForceTotal on every cell of the valid box [vlo, vhi]
// (level indices); cells outside the band get 0.
__global__ void k_ib_export_box(Fab fx, Fab fy, Fab fz, int has_fz, IbArgs a, int vlo0, int vlo1, int vlo2, int nx,
                                int ny, int64_t cells) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cells) return;
  const int i = vlo0 + (int)(t % nx), j = vlo1 + (int)((t / nx) % ny), k = vlo2 + (int)(t / ((int64_t)nx * ny));
  const int loc[3] = {i - a.ioff[0], j - a.ioff[1], k - a.ioff[2]};
  double ox = 0.0, oy = 0.0, oz = 0.0;
  int64_t b;
  if (band_index(a, loc, &b)) {
    const int64_t bc = a.band_cells;
    const double rho = a.Mband[b];
    const double* tot = a.ev_from_macro ? a.EFI : a.EFT;
    ox = a.dt * rho * tot[b];
    oy = a.dt * rho * tot[bc + b];
    oz = a.dt * rho * tot[2 * bc + b];
    a.Fband[b] = ox;
    a.Fband[bc + b] = oy;
    a.Fband[2 * bc + b] = oz;
  }
  fx.at(i, j, k, 0) = ox;
  fy.at(i, j, k, 0) = oy;
  if (has_fz) fz.at(i, j, k, 0) = oz;
}

}  // namespace ilbm

struct ilbm_ib {
  ilbm::Level L;             // the IB half of a level: markers, band, spreading lists
  hipStream_t own = nullptr; // set-up stream (the updates run on the caller's)
  int vlo[3] = {0, 0, 0}, vhi[3] = {0, 0, 0};
  ~ilbm_ib() { L.stream = own; }   // the Level destroys its own stream
};

extern "C" {

int ilbm_ib_create(const ilbm_case_desc* desc, const int valid_lo[3], const int valid_hi[3], ilbm_ib** out) {
  if (!desc || !out) return ILBM_ERR_INVALID;
  *out = nullptr;
  ilbm_ib* h = new (std::nothrow) ilbm_ib();
  if (!h) return ILBM_ERR_ALLOC;
  const int st = guarded([&] {
    using namespace ilbm;
    Level& L = h->L;
    ilbm_case_desc c = *desc;
    if (!c.use_ib) throw StatusError{ILBM_ERR_INVALID, "use_ib is 0"};
    const int dim = c.dim;
    bool sub = false;
    for (int d = 0; d < 3; ++d) {
      const int n = (d < dim) ? c.n[d] : 1;
      h->vlo[d] = (valid_lo && d < dim) ? valid_lo[d] : 0;
      h->vhi[d] = (valid_hi && d < dim) ? valid_hi[d] : n - 1;
      if (h->vlo[d] < 0 || h->vhi[d] > n - 1 || h->vhi[d] < h->vlo[d])
        throw StatusError{ILBM_ERR_INVALID, "valid box outside the level's domain"};
      if (h->vlo[d] != 0 || h->vhi[d] != n - 1) sub = true;
    }
    // a box inside the domain is a refined level's box: its faces are
    // coarse-fine interfaces and the marker stencils must stay inside it
    for (int d = 0; d < dim; ++d) {
      L.ioff[d] = h->vlo[d];
      L.dom_n[d] = c.n[d];
      c.n[d] = h->vhi[d] - h->vlo[d] + 1;
    }
    L.c = c;
    L.cf = sub;
    L.validate();
    HIPCHK(hipStreamCreateWithFlags(&h->own, hipStreamNonBlocking));
    L.stream = h->own;
    L.own_stream = true;
    L.dt = L.c.dx[0];
    L.tau = 3.0 * L.c.nu / L.c.dx[0] + 0.5;
    L.setup_ib();
    HIPCHK(hipStreamSynchronize(L.stream));
  });
  if (st != ILBM_OK) {
    delete h;
    return st;
  }
  *out = h;
  return ILBM_OK;
}

int ilbm_ib_destroy(ilbm_ib* h) {
  if (!h) return ILBM_ERR_INVALID;
  if (h->own) (void)hipStreamSynchronize(h->own);
  delete h;
  return ILBM_OK;
}

int ilbm_ib_update_forcing(ilbm_ib* h, const ilbm_fab* rho, const ilbm_fab* u, const ilbm_fab* v, const ilbm_fab* w,
                           ilbm_fab* fx, ilbm_fab* fy, ilbm_fab* fz, double dt, double time, void* stream) {
  if (!h || !rho || !u || !v || !fx || !fy) return ILBM_ERR_INVALID;
  return guarded([&] {
    using namespace ilbm;
    Level& L = h->L;
    const int dim = L.c.dim;
    if (dim == 3 && (!w || !fz)) throw StatusError{ILBM_ERR_INVALID, "3D update_forcing needs w and Fz"};
    if (!(dt > 0.0)) throw StatusError{ILBM_ERR_INVALID, "dt must be positive"};
    // the band cells are read, the valid box is written
    BoxGeom band{};
    band.dim = dim;
    for (int d = 0; d < 3; ++d) {
      band.vlo[d] = L.band_lo[d] + L.ioff[d];
      band.vhi[d] = L.band_lo[d] + L.band_n[d] - 1 + L.ioff[d];
    }
    BoxGeom valid{};
    valid.dim = dim;
    for (int d = 0; d < 3; ++d) { valid.vlo[d] = h->vlo[d]; valid.vhi[d] = h->vhi[d]; }
    const ilbm_fab* ins[4] = {rho, u, v, w};
    for (int c = 0; c < 4; ++c)
      if (ins[c]) check_fab(ins[c], 1, band, 0);
    const ilbm_fab* outs[3] = {fx, fy, fz};
    for (int c = 0; c < 3; ++c)
      if (outs[c]) check_fab(outs[c], 1, valid, 0);
    const Fab R = to_fab(*rho), U = to_fab(*u), V = to_fab(*v);
    const Fab W = w ? to_fab(*w) : Fab{};
    const Fab FX = to_fab(*fx), FY = to_fab(*fy);
    const Fab FZ = fz ? to_fab(*fz) : Fab{};
    hipStream_t s = (hipStream_t)stream;
    L.stream = s;   // every launch of this call on the caller's stream
    L.dt = dt;
    const IbArgs a0 = L.ib_args();
    hipLaunchKernelGGL(k_ib_gather_box, grid_for(L.band_cells), dim3(256), 0, s, R, U, V, W, w ? 1 : 0, a0, L.Mband);
    L.ib_stages(time);
    const IbArgs a = L.ib_args();
    const int nx = h->vhi[0] - h->vlo[0] + 1, ny = h->vhi[1] - h->vlo[1] + 1;
    const int64_t cells = (int64_t)nx * ny * (h->vhi[2] - h->vlo[2] + 1);
    hipLaunchKernelGGL(k_ib_export_box, grid_for(cells), dim3(256), 0, s, FX, FY, FZ, fz ? 1 : 0, a, h->vlo[0],
                       h->vlo[1], h->vlo[2], nx, ny, cells);
    const hipError_t e = hipGetLastError();
    L.stream = h->own;
    HIPCHK(e);
  });
}

int ilbm_ib_num_markers(ilbm_ib* h, int64_t* out) {
  if (!h || !out) return ILBM_ERR_INVALID;
  *out = h->L.nm;
  return ILBM_OK;
}

static int ib_sums_slice(ilbm_ib* h, int first, double out[3]) {
  if (!h || !out) return ILBM_ERR_INVALID;
  return guarded([&] {
    double s[8];
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(s, h->L.sums_d, sizeof(double) * 8, hipMemcpyDeviceToHost));
    for (int d = 0; d < 3; ++d) out[d] = s[first + d];
  });
}

int ilbm_ib_force_sum(ilbm_ib* h, double out[3]) { return ib_sums_slice(h, 0, out); }

int ilbm_ib_torque(ilbm_ib* h, double out[3]) { return ib_sums_slice(h, 3, out); }

int ilbm_ib_marker_state(ilbm_ib* h, double* xyz, double* uvw, double* fxyz) {
  if (!h) return ILBM_ERR_INVALID;
  return guarded([&] {
    ilbm::Level& L = h->L;
    const int nm = L.nm;
    std::vector<double> t(3 * (size_t)nm);
    HIPCHK(hipDeviceSynchronize());
    auto tr = [&](const double* dev, double* o) {
      if (!o) return;
      HIPCHK(hipMemcpy(t.data(), dev, sizeof(double) * 3 * nm, hipMemcpyDeviceToHost));
      for (int m = 0; m < nm; ++m)
        for (int d = 0; d < 3; ++d) o[3 * m + d] = t[(size_t)d * nm + m];
    };
    tr(L.xm_d, xyz);
    tr(L.U_d, uvw);
    tr(L.F_d, fxyz);
  });
}

}  // extern "C"
