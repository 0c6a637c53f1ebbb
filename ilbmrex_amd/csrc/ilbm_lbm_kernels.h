// Fused lattice kernels of the resident MI355X level (HBM-bound, no MFMA).
//
// Storage.  Populations are kept POST-COLLISION, structure of arrays
// [q][z][y][x] with x fastest.  One step of the reference
// (Source/Core/AdvanceLBMAtLevel.cpp:223-386: collide, FillBoundary,
// bounce-back, pull-stream, forced macros) then becomes one pass over
// memory:
//
//   k_step:  f^{n+1}(x,q) = g^n(x - c_q, q)                 (pull, stream.H:9-22)
//            rho,u = moments(f^{n+1}) + F^{n+1}/2           (calculateMacroForcing.H:9-32)
//            g^{n+1} = f - (f - feq(rho,u))/tau [+ Guo]     (collide.H, collide_forced.H)
//
// reading and writing each population once (2*Q*sizeof(pop) bytes per
// lattice update).  The immersed-boundary force of the next step is only
// known after the IB pass, so on IB cells the Guo term is added afterwards
// by k_guo_fixup; the reference adds it after the BGK update as well
// (collide_forced.H:20-50), so FP64 results are unchanged.
//
// Ghosts.  The arrays carry no ghost cells.  Periodic dimensions wrap
// (FillBoundary of post-collision data, AdvanceLBMAtLevel.cpp:244).  For
// non-periodic faces k_fill_ghosts writes, between steps, dense per-face
// slabs holding what the reference's stream reads in its ghost cells:
// halfway bounce-back for user_1 walls (:265-373) and the pre-collision
// FillPatch values of mesoscopicBcFill / foextrap elsewhere
// (bc_fill.H:303-347).  The pre-collision values of the domain's outer
// layer come from a face buffer that k_step writes while it has f^{n+1} in
// registers.
#pragma once

#include <type_traits>
#include <utility>

#include "ilbm_lattice.h"

namespace ilbm {

constexpr int kBlock = 256;
// ghost-slab fill of a single level: one thread per ghost cell, small blocks
// so that the few thousand ghost cells of a face spread over every CU
constexpr int kGhostBlock = 64;

// Population addressing of the stream-collide kernel:
//   kAddrWall64 a level with non-periodic faces, 64-bit element offsets
//   kAddrWrap   an all-periodic level, 32-bit byte offsets from a plane base
//   kAddrWall32 a level with non-periodic faces whose populations and ghost
//               slabs span under 4 GiB: 32-bit byte offsets from the buffer
//   kAddrWallX32 as kAddrWall32 with the non-periodic faces in x only (the
//               channel layout of the cylinder and sphere cases): a ghost
//               select only for directions with an x component
enum : int { kAddrWall64 = 0, kAddrWrap = 1, kAddrWall32 = 2, kAddrWallX32 = 3 };

// f(std::integral_constant<int, 0>) ... f(std::integral_constant<int, N - 1>):
// a loop over the directions whose index is a constant expression
template <class F, int... I>
__device__ __forceinline__ void static_for_impl(F& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

struct GridGeom {
  int n[3];        // valid extents (n[2] == 1 in 2D)
  int pad;         // ghost layers stored per dimension (0 or 1)
  int dim;
  int64_t js, ks;  // strides of the stored grid
  int64_t qs;      // elements between population q and q+1 of a cell
  int64_t pop_n;   // elements of all populations (the ghost slabs follow)
  int64_t off0;    // offset of valid cell (0,0,0)
  __host__ __device__ int64_t idx(int i, int j, int k) const {
    return off0 + (int64_t)k * ks + (int64_t)j * js + i;
  }
};

// Face buffers of the pre-collision boundary layer: face f = 2*d + side.
struct FaceBuf {
  void* p[6];
  int dims[6][2];  // the two in-plane dims of each face (-1 if unused)
};

// A face cell's Q populations are one record of face_stride elements (the
// record padded to 16 bytes): the boundary cell writes it with 16-byte
// stores instead of Q scattered ones.
__host__ __device__ constexpr int face_stride(int q, int pop_bytes) { return ((q * pop_bytes + 15) / 16 * 16) / pop_bytes; }

// f holds the populations, or for SH_IN the kernel's shifted working
// variable of shifted storage (the record's own format)
template <class LAT, class ST, bool SH_IN = false>
__device__ __forceinline__ void face_store(typename ST::raw* fb, int64_t fi, const typename ST::compute_t* f) {
  using R = typename ST::raw;
  using T = typename ST::compute_t;
  constexpr int QP = face_stride(LAT::Q, (int)sizeof(R));
  alignas(16) R rec[QP];
#pragma unroll
  for (int q = 0; q < QP; ++q) {
    if (q < LAT::Q) {
      if constexpr (SH_IN && ST::shifted) ST::store_s(rec, q, f[q]);
      else ST::store(rec, q, f[q], T(LAT::w[q]));
    } else {
      rec[q] = R(0);
    }
  }
  uint4* dst = reinterpret_cast<uint4*>(fb + fi * QP);
  const uint4* src = reinterpret_cast<const uint4*>(rec);
#pragma unroll
  for (int c = 0; c < QP * (int)sizeof(R) / 16; ++c) dst[c] = src[c];
}

template <class LAT, class ST>
__device__ __forceinline__ void face_load(const typename ST::raw* fb, int64_t fi, typename ST::compute_t* f) {
  using R = typename ST::raw;
  using T = typename ST::compute_t;
  constexpr int QP = face_stride(LAT::Q, (int)sizeof(R));
  alignas(16) R rec[QP];
  const uint4* src = reinterpret_cast<const uint4*>(fb + fi * QP);
  uint4* dst = reinterpret_cast<uint4*>(rec);
#pragma unroll
  for (int c = 0; c < QP * (int)sizeof(R) / 16; ++c) dst[c] = src[c];
#pragma unroll
  for (int q = 0; q < LAT::Q; ++q) f[q] = ST::load(rec, q, T(LAT::w[q]));
}

__host__ __device__ inline int64_t face_index(const GridGeom& g, int d, int i, int j, int k) {
  const int c[3] = {i, j, k};
  if (d == 0) return (int64_t)c[2] * g.n[1] + c[1];
  if (d == 1) return (int64_t)c[2] * g.n[0] + c[0];
  return (int64_t)c[1] * g.n[0] + c[0];
}

// Ghost cells of the non-periodic faces, stored densely per face slab.  A
// cell outside the domain in non-periodic dimensions d1 < d2 belongs to the
// slab of d1; in the other dimensions a slab spans [-1, n] (or [0, n) for a
// lower non-periodic dimension).
struct Slabs {
  int nface;
  int slab_of[3][2];       // face slab index of (dim, side), -1 if periodic
  int face_d[6], face_side[6];
  int face_lo[6][3], face_n[6][3];
  int64_t face_off[7];     // prefix sums of slab sizes; face_off[nface] = total
};

__host__ __device__ inline int64_t slab_index(const Slabs& S, const int c[3], int dim, const int n[3]) {
  int d = -1;
  for (int e = 0; e < dim; ++e)
    if (S.slab_of[e][0] >= 0 && (c[e] < 0 || c[e] > n[e] - 1)) { d = e; break; }
  const int f = S.slab_of[d][c[d] < 0 ? 0 : 1];
  int64_t t = S.face_off[f], mul = 1;
  for (int e = 0; e < 3; ++e) {
    if (e == d) continue;
    t += (int64_t)(c[e] - S.face_lo[f][e]) * mul;
    mul *= S.face_n[f][e];
  }
  return t;
}

// boundary condition types (AMReX BCType and the reference's user kinds)
enum { BC_INT_DIR = 0, BC_FOEXTRAP = 2, BC_EXT_DIR = 3, BC_USER_1 = 1001, BC_USER_2 = 1002 };

struct StepArgs {
  GridGeom g;
  Slabs slabs;
  const void* gin;
  void* gout;
  double tau, inv_tau;
  double Fmac[3];        // uniform F^{n+1} used by the macros outside the IB band
  double Fcol[3];        // uniform F^{n+2} used by the collision (no IB)
  int band_lo[3], band_n[3];
  const double* Fband;   // [3][band cells] exported IB force
  double* Mband;         // [4][band cells] rho,u,v,w written for the IB
  int64_t band_cells;
  double* mout;          // optional [4][cells] macro export (double)
  double* fout;          // optional [Q][cells] f^{n+1} export (double)
  FaceBuf faces;         // pre-collision boundary layer (non-periodic faces)
  int nonper[3];
  double* partial;       // per-block mass partials
  int tile_bx;           // tile width of the batched launch (block = tile_bx x kBlock/tile_bx)
  int tile_nc;           // cells per thread of the batched launch (1 or 2)
  // resident blocks of a persistent launch (k_step_p; 0: one block per tile)
  int persist;
  int band_force;        // the band carries the IB force (else it only exports macros)
  // sub-box split of a level: 0 every cell; 1 the cells outside the box;
  // 2 the cells of the box, the launch grid covering the box from sub_lo;
  // 3 as 2 without the cells of the inner box (in_lo, in_n)
  int sub_mode;
  int sub_lo[3], sub_n[3];
  // x-face ghost slabs written by k_step itself (see ghost_write): the BC
  // kind of each x side and its boundary values (rho, u, v, w)
  int gfuse;
  int gkind[2];
  double gval[2][4];
  int in_lo[3], in_n[3];
  int off32;             // a plane's cells are addressable with 32-bit byte offsets
  // the uniform parameters rounded once to FP32 on the host, for shifted
  // storage (bound by vector ALU issue: no per-wavefront conversion or
  // divide): (float)tau, (float)inv_tau, 1 - 0.5 / (float)tau, (float)Fmac,
  // (float)Fcol
  float ftau, finv_tau, ffac;
  float fFmac[3], fFcol[3];
};

// a uniform parameter in ST's compute type: the host-rounded FP32 copy for
// shifted storage (the same value as converting the FP64 one).  The other
// formats convert on the device: FP32 keeps one load of the band or uniform
// force on every lane, which beat a band branch by 4 % on the 2D cylinder.
template <class ST>
__device__ __forceinline__ typename ST::compute_t arg_t(double d, float f) {
  using T = typename ST::compute_t;
  if constexpr (ST::shifted) return f;
  else return T(d);
}

// Block of a batched launch: case index into the launch's argument table and
// the block's index inside that case's own grid.
struct BlockRef {
  int c;
  int b;
  // stream-collide maps: the first cell of the block's tile (step_ref)
  int i0, j0, k0;
};

// Tile width bx (0: linear tiles) and cells per thread nc of a case's
// batched stream-collide launch: the first cell of block b (its tile's
// corner; for linear tiles cell nc * 256 * b in x-fastest order).  The host
// stores it in the block map, so the kernel needs no division.
__host__ __device__ inline BlockRef step_ref(const int n[3], int bx, int nc, int c, int64_t b) {
  BlockRef r{c, (int)b, 0, 0, 0};
  if (bx == 0) {
    const int64_t lin0 = b * nc * kBlock;
    const int64_t row0 = lin0 / n[0];
    r.i0 = (int)(lin0 - row0 * n[0]);
    r.k0 = (int)(row0 / n[1]);
    r.j0 = (int)(row0 - (int64_t)r.k0 * n[1]);
    return r;
  }
  const int by = kBlock / bx;
  const int gx = (n[0] + bx - 1) / bx, gy = (n[1] + nc * by - 1) / (nc * by);
  const int bxi = (int)(b % gx), rest = (int)(b / gx);
  r.i0 = (bxi) * bx;
  r.j0 = (rest % gy) * (nc * by);
  r.k0 = rest / gy;
  return r;
}

// ---------------------------------------------------------------------------
// Deterministic block reduction (fixed tree for a fixed block shape).
// ---------------------------------------------------------------------------
__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

__device__ inline double block_sum(double v, double* lds) {
  const int t = threadIdx.x + threadIdx.y * blockDim.x;
  const int nt = blockDim.x * blockDim.y;
  v = wave_sum(v);
  if ((t & 63) == 0) lds[t >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (t == 0) {
    for (int wv = 0; wv < (nt + 63) / 64; ++wv) s += lds[wv];
  }
  __syncthreads();
  return s;
}

// Equilibrium as feqFunction (Source/LBM/feq.H:9-29): the 2D form omits w.
template <class LAT, class T>
__device__ inline T feq_q(int q, T rho, T u, T v, T w, T uv) {
  const T cx = T(LAT::cx[q]), cy = T(LAT::cy[q]), cz = T(LAT::cz[q]);
  T eu;
  if constexpr (LAT::D == 3) eu = cx * u + cy * v + cz * w;
  else eu = cx * u + cy * v;
  return T(LAT::w[q]) * rho * (T(1.0) + T(3.0) * eu + T(4.5) * eu * eu - T(1.5) * uv);
}

template <class LAT, class T>
__device__ inline T usq(T u, T v, T w) {
  if constexpr (LAT::D == 3) return u * u + v * v + w * w;
  else return u * u + v * v;
}

// Guo forcing increment of collide_forced.H:27-50.  FP64 keeps the
// reference's divisions by cs2; the FP32 collision multiplies by 1/cs2 = 3.
template <class LAT, class T>
__device__ inline T guo_q(int q, T fac, T ux, T uy, T uz, T fx, T fy, T fz) {
  const T cx = T(LAT::cx[q]), cy = T(LAT::cy[q]), cz = T(LAT::cz[q]);
  const T eiu = cx * ux + cy * uy + cz * uz;
  T termx, termy, termz;
  if constexpr (sizeof(T) == 8) {
    const T cs2 = T(1.0) / T(3.0);
    termx = (cx - ux) / cs2 + (eiu / (cs2 * cs2)) * cx;
    termy = (cy - uy) / cs2 + (eiu / (cs2 * cs2)) * cy;
    if constexpr (LAT::D == 3) termz = (cz - uz) / cs2 + (eiu / (cs2 * cs2)) * cz;
    else termz = T(0.0);
  } else {
    const T e9 = eiu * T(9.0);
    termx = (cx - ux) * T(3.0) + e9 * cx;
    termy = (cy - uy) * T(3.0) + e9 * cy;
    if constexpr (LAT::D == 3) termz = (cz - uz) * T(3.0) + e9 * cz;
    else termz = T(0.0);
  }
  return T(LAT::w[q]) * fac * (termx * fx + termy * fy + termz * fz);
}

// u = (m + F/2) / rho: FP64 divides as the reference; FP32 multiplies by
// the hardware reciprocal.
template <class ST>
__device__ inline void forced_velocity(typename ST::compute_t r, typename ST::compute_t m1,
                                       typename ST::compute_t m2, typename ST::compute_t m3,
                                       typename ST::compute_t fx, typename ST::compute_t fy,
                                       typename ST::compute_t fz, typename ST::compute_t& u,
                                       typename ST::compute_t& v, typename ST::compute_t& w) {
  using T = typename ST::compute_t;
  if constexpr (ST::exact) {
    u = (m1 + T(0.5) * fx) / r;
    v = (m2 + T(0.5) * fy) / r;
    w = (m3 + T(0.5) * fz) / r;
  } else {
    const T ri = __builtin_amdgcn_rcpf(r);
    u = (m1 + T(0.5) * fx) * ri;
    v = (m2 + T(0.5) * fy) * ri;
    w = (m3 + T(0.5) * fz) * ri;
  }
}

// BGK relaxation of collide.H:17-24; FP64 keeps the reference's divide.
template <class ST>
__device__ inline typename ST::compute_t bgk(typename ST::compute_t f, typename ST::compute_t feq,
                                             double tau, double inv_tau) {
  using T = typename ST::compute_t;
  if constexpr (ST::exact) return 1.0 * f - (f - feq) / tau;
  else return f - (f - feq) * T(inv_tau);
}

// Per-block partial of Sum(Rho); k_sum_partials adds them in block order.
// (A last-block-done reduction would need an agent-scope release per block,
// which on a multi-XCD device writes back the XCD's L2 -- far costlier than
// one more small kernel.)
__device__ inline void finish_mass(const StepArgs& a, double s, int64_t blk) {
  if (threadIdx.x == 0 && threadIdx.y == 0) a.partial[blk] = s;
}

// Ghost value of the ext_dir / user_2 mesoscopic functor (bc_fill.H:303-347)
// from the pre-collision populations fnp of its mirror cell: the
// equilibrium of the ghost's macros plus the mirror's non-equilibrium part.
// ext_dir imposes the boundary values bv = (rho, u, v, w); user_2 reflects
// the mirror's macros about them.
template <class LAT, class T>
__device__ __forceinline__ void neep_ghost(int kind, const T* fnp, const double* bv, T* out) {
  T rhonp = T(0), unp = T(0), vnp = T(0), wnp = T(0);
#pragma unroll
  for (int q = 0; q < LAT::Q; ++q) {
    rhonp += fnp[q];
    unp += T(LAT::cx[q]) * fnp[q];
    vnp += T(LAT::cy[q]) * fnp[q];
    wnp += T(LAT::cz[q]) * fnp[q];
  }
  rhonp = rhonp > T(1.0e-20) ? rhonp : T(1.0e-20);
  unp /= rhonp; vnp /= rhonp; wnp /= rhonp;
  const T rho_b = T(bv[0]), u_b = T(bv[1]), v_b = T(bv[2]), w_b = T(bv[3]);
  T rg = rho_b, ug = u_b, vg = v_b, wg = w_b;
  if (kind == BC_USER_2) {
    const T t2 = T(2.0) * rho_b - rhonp;
    rg = T(1.0e-20) > t2 ? T(1.0e-20) : t2;
    ug = T(2.0) * u_b - unp; vg = T(2.0) * v_b - vnp; wg = T(2.0) * w_b - wnp;
  }
#pragma unroll
  for (int q = 0; q < LAT::Q; ++q) {
    const T cx = T(LAT::cx[q]), cy = T(LAT::cy[q]), cz = T(LAT::cz[q]);
    const T eu_np = cx * unp + cy * vnp + cz * wnp;
    const T uv_np = unp * unp + vnp * vnp + wnp * wnp;
    const T feqnp = T(LAT::w[q]) * rhonp * (T(1.0) + T(3.0) * eu_np + T(4.5) * eu_np * eu_np - T(1.5) * uv_np);
    const T eu_g = cx * ug + cy * vg + cz * wg;
    const T uv_g = ug * ug + vg * vg + wg * wg;
    const T feqg = T(LAT::w[q]) * rg * (T(1.0) + T(3.0) * eu_g + T(4.5) * eu_g * eu_g - T(1.5) * uv_g);
    out[q] = feqg + (fnp[q] - feqnp);
  }
}

// The x-face ghost slabs of the next step, written by the boundary cells
// themselves (StepArgs::gfuse: levels walled in x only whose x faces are
// ext_dir, user_2 or foextrap).  Each ghost (-1 or n0, jg, kg) takes its
// FillPatch value from one boundary cell: for ext_dir / user_2, whose
// functor clamps the mirror, (0 or n0 - 1, clamp(jg), clamp(kg)); for
// foextrap (FilccCell, then the periodic copy) the image (wrap(jg),
// wrap(kg)).  fillpatch_value computes the same values from the boundary
// layer.  Only the directions the next pull reads from a slab are written:
// c_x = +1 at the lower face, -1 at the upper.
template <class LAT, class ST, int SIDE>
__device__ __forceinline__ void ghost_write(const StepArgs& a, int j, int k, const typename ST::compute_t* f) {
  using T = typename ST::compute_t;
  using R = typename ST::raw;
  constexpr int CIN = SIDE ? -1 : 1;   // c_x of the directions pulled from this slab
  const GridGeom& g = a.g;
  R* out = static_cast<R*>(a.gout) + g.pop_n;
  const int64_t npad = a.slabs.face_off[a.slabs.nface];
  const int kind = a.gkind[SIDE];
  T fnp[LAT::Q];
#pragma unroll
  for (int q = 0; q < LAT::Q; ++q) fnp[q] = ST::shifted ? f[q] + T(LAT::w[q]) : f[q];
  T val[LAT::Q];
  if (kind == BC_FOEXTRAP) {
#pragma unroll
    for (int q = 0; q < LAT::Q; ++q) val[q] = fnp[q];
  } else {
    neep_ghost<LAT, T>(kind, fnp, a.gval[SIDE], val);
  }
  const bool clamp = kind != BC_FOEXTRAP;
  // the ghost rows of this cell in y (and z): its own, and the rows -1 and
  // n that the rule maps onto it
  const bool jlo = j == (clamp ? 0 : g.n[1] - 1), jhi = j == (clamp ? g.n[1] - 1 : 0);
  const bool klo = k == (clamp ? 0 : g.n[2] - 1), khi = k == (clamp ? g.n[2] - 1 : 0);
  const int fs = a.slabs.slab_of[0][SIDE];
  const int64_t base = a.slabs.face_off[fs];
  const int flo1 = a.slabs.face_lo[fs][1], flo2 = a.slabs.face_lo[fs][2];
  const int64_t fn1 = a.slabs.face_n[fs][1];
#pragma unroll
  for (int bj = 0; bj < 3; ++bj) {
    if ((bj == 1 && !jlo) || (bj == 2 && !jhi)) continue;
    const int jj = (bj == 0) ? j : (bj == 1) ? -1 : g.n[1];
#pragma unroll
    for (int bk = 0; bk < (LAT::D == 3 ? 3 : 1); ++bk) {
      if ((bk == 1 && !klo) || (bk == 2 && !khi)) continue;
      const int kk = (bk == 0) ? k : (bk == 1) ? -1 : g.n[2];
      const int64_t t = base + (jj - flo1) + (int64_t)(kk - flo2) * fn1;
      static_for<LAT::Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (LAT::cx[q] == CIN) ST::store(out, (int64_t)q * npad + t, val[q], T(LAT::w[q]));
      });
    }
  }
}

// ---------------------------------------------------------------------------
// k_step: pull-stream + forced macros + BGK collision, one pass.
// ---------------------------------------------------------------------------
// Index of the first nonzero entry of a lattice component table.
__host__ __device__ constexpr int first_nonzero(const int* c, int n) {
  for (int q = 0; q < n; ++q)
    if (c[q] != 0) return q;
  return -1;
}

// acc (+/-)= x for a term of sign S (0: no term); term K, the first nonzero
// one (FIRST), assigns instead of adding.
template <int FIRST, int K, int S, class T>
__device__ __forceinline__ void accumulate_signed(T& acc, T x) {
  if constexpr (S > 0) {
    if constexpr (K == FIRST) acc = x;
    else acc += x;
  } else if constexpr (S < 0) {
    if constexpr (K == FIRST) acc = -x;
    else acc -= x;
  }
}

// Offsets of the stream-collide kernel: in the 32-bit modes unsigned 32-bit
// arithmetic (every intermediate is exact modulo 2^32 and every final offset
// lies below 2^32), else 64-bit element offsets.
template <int AM>
using StepOff = std::conditional_t<AM != kAddrWall64, uint32_t, int64_t>;

// the cell takes part in this launch (inside the level and its sub-box part)
__device__ __forceinline__ bool step_active(const StepArgs& a, int i, int j, int k) {
  const GridGeom& g = a.g;
  bool active = (i < g.n[0]) && (j < g.n[1]);
  if (a.sub_mode != 0) {
    const bool in_sub = (unsigned)(i - a.sub_lo[0]) < (unsigned)a.sub_n[0] &&
                        (unsigned)(j - a.sub_lo[1]) < (unsigned)a.sub_n[1] &&
                        (unsigned)(k - a.sub_lo[2]) < (unsigned)a.sub_n[2];
    if (a.sub_mode == 3) {
      const bool in_inner = (unsigned)(i - a.in_lo[0]) < (unsigned)a.in_n[0] &&
                            (unsigned)(j - a.in_lo[1]) < (unsigned)a.in_n[1] &&
                            (unsigned)(k - a.in_lo[2]) < (unsigned)a.in_n[2];
      active = active && in_sub && !in_inner;
    } else {
      active = active && (a.sub_mode == 2 ? in_sub : !in_sub);
    }
  }
  return active;
}

// First half of a cell's update: the pull loads of f^{n+1} (into f) and the
// cell's own offset (self).
template <class LAT, class ST, int AM>
__device__ __forceinline__ void step_load(const StepArgs& a, int i, int j, int k, typename ST::compute_t* f,
                                          StepOff<AM>& self) {
  using T = typename ST::compute_t;
  using R = typename ST::raw;
  constexpr bool WRAP = AM == kAddrWrap;
  constexpr bool XO = AM == kAddrWallX32;
  constexpr bool O32 = AM != kAddrWall64;
  using I = StepOff<AM>;
  const GridGeom& g = a.g;
  (void)XO;
  const R* __restrict__ gin = static_cast<const R*>(a.gin);
  // source offsets for c = -1, 0, +1 (source coordinate = x - c); periodic
  // dimensions wrap.  Shifted storage (bound by vector ALU issue) holds
  // byte offsets in the 32-bit modes (SB bytes per element), so that no
  // direction scales its offset again (FP16-shifted 2D cylinder +3 %);
  // the other formats scale each offset by EB where it is used.
  constexpr uint32_t SB = (O32 && ST::shifted) ? (uint32_t)sizeof(R) : 1u;
  constexpr uint32_t EB = O32 ? (uint32_t)sizeof(R) / SB : 1u;
  I ox[3], oy[3], oz[3];
  {
    const int xm = (i == 0) ? g.n[0] - 1 : i - 1, xp = (i == g.n[0] - 1) ? 0 : i + 1;
    const int ym = (j == 0) ? g.n[1] - 1 : j - 1, yp = (j == g.n[1] - 1) ? 0 : j + 1;
    const I jsb = (I)g.js * (I)SB;
    ox[0] = (I)xp * (I)SB; ox[1] = (I)i * (I)SB; ox[2] = (I)xm * (I)SB;
    oy[0] = (I)yp * jsb; oy[1] = (I)j * jsb; oy[2] = (I)ym * jsb;
    if constexpr (LAT::D == 3) {
      const I ksb = (I)g.ks * (I)SB;
      const int zm = (k == 0) ? g.n[2] - 1 : k - 1, zp = (k == g.n[2] - 1) ? 0 : k + 1;
      oz[0] = (I)zp * ksb; oz[1] = (I)k * ksb; oz[2] = (I)zm * ksb;
    } else {
      oz[0] = oz[1] = oz[2] = 0;
    }
  }
  // A source outside a non-periodic face is a ghost cell; its value (the
  // pre-collision FillPatch value, or bounce-back) sits in the slabs that
  // follow the Q population planes of the same buffer.  The BC functor
  // clamps the mirror in every dimension (bc_fill.H:28-40), so a ghost
  // that is also outside a periodic dimension is not wrapped.  All loads
  // are issued unconditionally; only their offsets are selected.
  bool outx[3] = {false, false, false}, outy[3] = {false, false, false}, outz[3] = {false, false, false};
  I tb[3] = {0, 0, 0}, mu[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  if constexpr (XO) {
    // both x faces span the same (y, z) range, so their strides are
    // uniform; only the face's offset differs between the two sides
    outx[0] = (i == g.n[0] - 1);
    outx[2] = (i == 0);
    const int f0 = a.slabs.slab_of[0][0], f1 = a.slabs.slab_of[0][1];
    const I fn1 = (I)a.slabs.face_n[f0][1];
    mu[0][1] = (I)SB;
    mu[0][2] = fn1 * (I)SB;
    const I base = (i == 0) ? (I)a.slabs.face_off[f0] : (I)a.slabs.face_off[f1];
    tb[0] = (base + (I)(j - a.slabs.face_lo[f0][1]) + (I)(k - a.slabs.face_lo[f0][2]) * fn1) * (I)SB;
  } else if constexpr (!WRAP) {
    const int cc[3] = {i, j, k};
    bool* outs[3] = {outx, outy, outz};
#pragma unroll
    for (int d = 0; d < LAT::D; ++d) {
      if (!a.nonper[d]) continue;
      outs[d][0] = (cc[d] == g.n[d] - 1);
      outs[d][2] = (cc[d] == 0);
      const int sd = (cc[d] == 0) ? 0 : 1;
      const int f = a.slabs.slab_of[d][sd];
      I t = (I)a.slabs.face_off[f], mul = 1;
#pragma unroll
      for (int e = 0; e < 3; ++e) {
        if (e == d) continue;
        t += (I)(cc[e] - a.slabs.face_lo[f][e]) * mul;
        mu[d][e] = mul * (I)SB;
        mul *= (I)a.slabs.face_n[f][e];
      }
      tb[d] = t * (I)SB;
    }
  }
  const I gofs = (I)g.pop_n * (I)SB;                  // start of the slabs
  const I npad = (I)a.slabs.face_off[a.slabs.nface] * (I)SB;
#pragma unroll
  for (int q = 0; q < LAT::Q; ++q) {
    if constexpr (WRAP) {
      // all-periodic level with planes under 4 GiB: a 32-bit byte offset
      // from the plane's (uniform) base
      const uint32_t src = (uint32_t)(ox[LAT::cx[q] + 1] + oy[LAT::cy[q] + 1] + oz[LAT::cz[q] + 1]);
      const R* plane = gin + (int64_t)q * g.qs;
      const R* p = reinterpret_cast<const R*>(reinterpret_cast<const char*>(plane) + src * EB);
      // 3D: non-temporal loads (256^3 FP64 5.66 -> 5.94 TB/s, FP32 128^3
      // +1 %); on walled and 2D levels they cost 2-9 %
      if constexpr (LAT::D == 3) f[q] = pop_in_nt<ST>(p, 0, T(LAT::w[q]));
      else f[q] = pop_in<ST>(p, 0, T(LAT::w[q]));
      continue;
    }
    // the source cell inside the level, or for a source beyond a
    // non-periodic face the slab of the lowest such dimension it leaves
    // (a per-lane select: one load per direction on every lane)
    I off = ((I)q * (I)g.qs + (I)g.off0) * (I)SB + ox[LAT::cx[q] + 1] + oy[LAT::cy[q] + 1] + oz[LAT::cz[q] + 1];
    if constexpr (XO) {
      if (LAT::cx[q] != 0) {
        const I t0 = tb[0] - (I)LAT::cy[q] * mu[0][1] - (I)LAT::cz[q] * mu[0][2];
        off = outx[LAT::cx[q] + 1] ? gofs + (I)q * npad + t0 : off;
      }
      const R* p = reinterpret_cast<const R*>(reinterpret_cast<const char*>(gin) + off * EB);
      f[q] = pop_in<ST>(p, 0, T(LAT::w[q]));
      continue;
    }
    const bool o0 = outx[LAT::cx[q] + 1], o1 = outy[LAT::cy[q] + 1], o2 = outz[LAT::cz[q] + 1];
    const I t0 = tb[0] - (I)LAT::cy[q] * mu[0][1] - (I)LAT::cz[q] * mu[0][2];
    const I t1 = tb[1] - (I)LAT::cx[q] * mu[1][0] - (I)LAT::cz[q] * mu[1][2];
    const I t2 = tb[2] - (I)LAT::cx[q] * mu[2][0] - (I)LAT::cy[q] * mu[2][1];
    const I t = o0 ? t0 : (o1 ? t1 : t2);
    off = (o0 | o1 | o2) ? gofs + (I)q * npad + t : off;
    if constexpr (O32) {
      const R* p = reinterpret_cast<const R*>(reinterpret_cast<const char*>(gin) + off * EB);
      f[q] = pop_in<ST>(p, 0, T(LAT::w[q]));
    } else {
      f[q] = pop_in<ST>(gin, off, T(LAT::w[q]));
    }
  }
  self = (I)g.off0 * (I)SB + ox[1] + oy[1] + oz[1];
}

// Second half: macros, exports, boundary layer, collision and stores of the
// loaded f; returns the cell's rho for Sum(Rho).
template <class LAT, class ST, int AM>
__device__ __forceinline__ double step_update(const StepArgs& a, int i, int j, int k,
                                              const typename ST::compute_t* f, StepOff<AM> self) {
  using T = typename ST::compute_t;
  using R = typename ST::raw;
  constexpr bool WRAP = AM == kAddrWrap;
  constexpr bool XO = AM == kAddrWallX32;
  constexpr bool O32 = AM != kAddrWall64;
  using I = StepOff<AM>;
  const GridGeom& g = a.g;
  (void)WRAP;
  R* __restrict__ gout = static_cast<R*>(a.gout);
  constexpr uint32_t SB = (O32 && ST::shifted) ? (uint32_t)sizeof(R) : 1u;
  constexpr uint32_t EB = O32 ? (uint32_t)sizeof(R) / SB : 1u;
  // f holds f_q, or h_q = f_q - w_q for shifted storage (then the first
  // moments are the same sums, since sum_q w_q c_q = 0, and rho = 1 +
  // sum_q h_q)
  constexpr bool SH = ST::shifted;
  // moments in the reference's order (calculateMacroForcing.H:21-27); a
  // term whose direction component is 0 is left out and -1 * f is
  // subtracted, which leaves every finite sum bit-identical
  T r = T(0), m1 = T(0), m2 = T(0), m3 = T(0);
  if constexpr (ST::exact) {
    static_for<LAT::Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      r += f[q];
      if constexpr (LAT::cx[q] > 0) m1 += f[q];
      else if constexpr (LAT::cx[q] < 0) m1 -= f[q];
      if constexpr (LAT::cy[q] > 0) m2 += f[q];
      else if constexpr (LAT::cy[q] < 0) m2 -= f[q];
      if constexpr (LAT::cz[q] > 0) m3 += f[q];
      else if constexpr (LAT::cz[q] < 0) m3 -= f[q];
    });
  } else {
    // reduced precision: each sum starts at its first term (0 + x cannot
    // be folded away in IEEE arithmetic, -0 + 0 = +0)
    static_for<LAT::Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      if constexpr (q == 0) r = f[q];
      else r += f[q];
      accumulate_signed<first_nonzero(LAT::cx, LAT::Q), q, LAT::cx[q]>(m1, f[q]);
      accumulate_signed<first_nonzero(LAT::cy, LAT::Q), q, LAT::cy[q]>(m2, f[q]);
      accumulate_signed<first_nonzero(LAT::cz, LAT::Q), q, LAT::cz[q]>(m3, f[q]);
    });
  }
  if constexpr (SH) r += T(1);
  const int bi = i - a.band_lo[0], bj = j - a.band_lo[1], bk = k - a.band_lo[2];
  const bool in_band = (unsigned)bi < (unsigned)a.band_n[0] && (unsigned)bj < (unsigned)a.band_n[1] &&
                       (unsigned)bk < (unsigned)a.band_n[2];
  const int64_t bidx = ((int64_t)bk * a.band_n[1] + bj) * a.band_n[0] + bi;
  T Fx = arg_t<ST>(a.Fmac[0], a.fFmac[0]), Fy = arg_t<ST>(a.Fmac[1], a.fFmac[1]), Fz = arg_t<ST>(a.Fmac[2], a.fFmac[2]);
  if (in_band && a.band_force) {
    Fx = T(a.Fband[bidx]);
    Fy = T(a.Fband[a.band_cells + bidx]);
    Fz = T(a.Fband[2 * a.band_cells + bidx]);
  }
  T u, v, w;
  forced_velocity<ST>(r, m1, m2, m3, Fx, Fy, Fz, u, v, w);
  const double rho_out = (double)r;
  const int64_t vcell = ((int64_t)k * g.n[1] + j) * g.n[0] + i;
  const int64_t ncell = (int64_t)g.n[0] * g.n[1] * g.n[2];
  if (in_band) {
    a.Mband[bidx] = (double)r;
    a.Mband[a.band_cells + bidx] = (double)u;
    a.Mband[2 * a.band_cells + bidx] = (double)v;
    a.Mband[3 * a.band_cells + bidx] = (double)w;
  }
  if (a.mout) {
    a.mout[vcell] = (double)r;
    a.mout[ncell + vcell] = (double)u;
    a.mout[2 * ncell + vcell] = (double)v;
    a.mout[3 * ncell + vcell] = (double)w;
  }
  if (a.fout) {
#pragma unroll
    for (int q = 0; q < LAT::Q; ++q) a.fout[(int64_t)q * ncell + vcell] = (double)f[q] + (SH ? LAT::w[q] : 0.0);
  }
  // pre-collision boundary layer for the next ghost fill.  A level whose
  // stream-collide kernel writes its x-face slabs itself (gfuse) keeps the
  // layer it had: its slabs are rebuilt from the layer only when they are
  // not current, after the initial state (whose kernel writes the layer)
  // or a restore of a checkpoint taken before the first step.  The stores
  // are lone 2- to 8-byte writes of one lane per wavefront (FP32 sphere
  // 256x128x128: 3 % of the kernel's time, FP16-shifted 8 %).
  const int cc[3] = {i, j, k};
#pragma unroll
  for (int d = 0; d < LAT::D; ++d) {
    if (!a.nonper[d] || (XO && a.gfuse)) continue;
    int side = -1;
    if (cc[d] == 0) side = 0;
    else if (cc[d] == g.n[d] - 1) side = 1;
    if (side < 0) continue;
    R* fb = static_cast<R*>(a.faces.p[2 * d + side]);
    const int64_t fi = face_index(g, d, i, j, k);
    face_store<LAT, ST, true>(fb, fi, f);
    // a cell sitting on the lo and hi face of a 1-cell dimension writes both
    if (g.n[d] == 1) face_store<LAT, ST, true>(static_cast<R*>(a.faces.p[2 * d + 1]), fi, f);
  }
  if constexpr (XO) {
    // (a non-periodic x extent is at least 2: a cell is on one face)
    if (a.gfuse && i == 0) ghost_write<LAT, ST, 0>(a, j, k, f);
    else if (a.gfuse && i == g.n[0] - 1) ghost_write<LAT, ST, 1>(a, j, k, f);
  }
  // BGK collision with the macros just computed (feq uses macro_old of the
  // next step, AdvanceLBMAtLevel.cpp:225-238)
  const T uv = usq<LAT, T>(u, v, w);
  const T cfx = arg_t<ST>(a.Fcol[0], a.fFcol[0]), cfy = arg_t<ST>(a.Fcol[1], a.fFcol[1]), cfz = arg_t<ST>(a.Fcol[2], a.fFcol[2]);
  const bool forced = !(cfx == T(0.0) && cfy == T(0.0) && cfz == T(0.0));
  const T fac = ST::shifted ? T(a.ffac) : T(1.0) - T(0.5) / T(a.tau);
  auto put = [&](int q, T gq) {
    if constexpr (O32) {
      // a 32-bit byte offset from the buffer base, as the loads
      R* p = reinterpret_cast<R*>(reinterpret_cast<char*>(gout) + ((I)q * (I)g.qs * (I)SB + self) * EB);
      pop_out_nt<ST>(p, 0, gq, T(LAT::w[q]));
    } else if (a.off32) {
      R* plane = gout + (int64_t)q * g.qs;
      R* p = reinterpret_cast<R*>(reinterpret_cast<char*>(plane) + (uint32_t)self * (uint32_t)sizeof(R));
      pop_out_nt<ST>(p, 0, gq, T(LAT::w[q]));
    } else {
      pop_out_nt<ST>(gout, (int64_t)q * g.qs + (int64_t)self, gq, T(LAT::w[q]));
    }
  };
  if constexpr (ST::exact) {
#pragma unroll
    for (int q = 0; q < LAT::Q; ++q) {
      const T fe = feq_q<LAT, T>(q, r, u, v, w, uv);
      T gq = bgk<ST>(f[q], fe, a.tau, a.inv_tau);
      if (forced) gq += guo_q<LAT, T>(q, fac, u, v, w, cfx, cfy, cfz);
      put(q, gq);
    }
  } else {
    // reduced-precision collision, a direction and its opposite together:
    // feq = w rho (1 - 1.5 u.u + 4.5 (c.u)^2) +- 3 w rho (c.u), then
    // g = f + (feq - f) / tau with fused multiply-adds.  Shifted storage
    // relaxes h = f - w towards feq - w (the same update, shifted).
    const T base = T(1.0) - T(1.5) * uv;
    const T om = arg_t<ST>(a.inv_tau, a.finv_tau);
    static_for<LAT::Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      constexpr int p = opposite<LAT>(q);
      if constexpr (p == q) {
        const T fe = SH ? __builtin_fmaf(T(LAT::w[q]) * r, base, -T(LAT::w[q])) : T(LAT::w[q]) * r * base;
        T gq = __builtin_fmaf(fe - f[q], om, f[q]);
        if (forced) gq += guo_q<LAT, T>(q, fac, u, v, w, cfx, cfy, cfz);
        put(q, gq);
      } else if constexpr (p > q) {
        // c.u from its nonzero components, starting at the first
        constexpr int c0 = LAT::cx[q], c1 = LAT::cy[q], c2 = (LAT::D == 3) ? LAT::cz[q] : 0;
        constexpr int k0 = c0 ? 0 : (c1 ? 1 : 2);
        T eu = T(0);
        accumulate_signed<k0, 0, c0>(eu, u);
        accumulate_signed<k0, 1, c1>(eu, v);
        accumulate_signed<k0, 2, c2>(eu, w);
        const T wr = T(LAT::w[q]) * r;
        const T A = SH ? __builtin_fmaf(wr, __builtin_fmaf(T(4.5) * eu, eu, base), -T(LAT::w[q]))
                       : wr * __builtin_fmaf(T(4.5) * eu, eu, base);
        const T B = (T(3.0) * wr) * eu;
        T gq = __builtin_fmaf((A + B) - f[q], om, f[q]);
        T gp = __builtin_fmaf((A - B) - f[p], om, f[p]);
        if (forced) {
          gq += guo_q<LAT, T>(q, fac, u, v, w, cfx, cfy, cfz);
          gp += guo_q<LAT, T>(p, fac, u, v, w, cfx, cfy, cfz);
        }
        put(q, gq);
        put(p, gp);
      }
    });
  }
  return rho_out;
}

// NC cells per thread: the loads of every cell are issued before the first
// update, so that a wavefront keeps NC times the bytes in flight.
template <class LAT, class ST, int AM, int NC>
__device__ __forceinline__ void step_body_n(const StepArgs& a, const int* i, const int* j, const int* k, int64_t blk,
                                            double* lds) {
  using T = typename ST::compute_t;
  T f[NC][LAT::Q];
  StepOff<AM> self[NC];
  bool act[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    act[c] = step_active(a, i[c], j[c], k[c]);
    if (act[c]) step_load<LAT, ST, AM>(a, i[c], j[c], k[c], f[c], self[c]);
  }
  double rho_acc = 0.0;
#pragma unroll
  for (int c = 0; c < NC; ++c)
    if (act[c]) rho_acc += step_update<LAT, ST, AM>(a, i[c], j[c], k[c], f[c], self[c]);
  const double s = block_sum(rho_acc, lds);
  finish_mass(a, s, blk);
}

template <class LAT, class ST, int AM>
__device__ __forceinline__ void step_body(const StepArgs& a, int i, int j, int k, int64_t blk, double* lds) {
  using T = typename ST::compute_t;
  double rho_acc = 0.0;
  if (step_active(a, i, j, k)) {
    T f[LAT::Q];
    StepOff<AM> self;
    step_load<LAT, ST, AM>(a, i, j, k, f, self);
    rho_acc = step_update<LAT, ST, AM>(a, i, j, k, f, self);
  }
  const double s = block_sum(rho_acc, lds);
  finish_mass(a, s, blk);
}

// Row of a thread of a level's stream-collide blocks: they are bx x by
// threads with bx a multiple of 64 (the launchers refuse other shapes), so
// every wavefront lies in one row, and the row and the offsets derived from
// it stay in scalar registers (FP16-shifted sphere: fewer vector
// instructions per wavefront).
__device__ __forceinline__ int wave_row(int j) { return __builtin_amdgcn_readfirstlane(j); }

template <class LAT, class ST, int AM>
__global__ void __launch_bounds__(kBlock) k_step(StepArgs a) {
  __shared__ double lds[kBlock / 64];
  const int64_t blk = ((int64_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
  const int o0 = (a.sub_mode >= 2) ? a.sub_lo[0] : 0, o1 = (a.sub_mode >= 2) ? a.sub_lo[1] : 0;
  const int o2 = (a.sub_mode >= 2) ? a.sub_lo[2] : 0;
  step_body<LAT, ST, AM>(a, o0 + blockIdx.x * blockDim.x + threadIdx.x,
                         wave_row(o1 + blockIdx.y * blockDim.y + threadIdx.y), o2 + blockIdx.z, blk, lds);
}

// k_step with a fixed number of resident blocks, each striding over the
// tiles of the (tx, ty, tz) grid k_step would launch; tile t keeps k_step's
// block index as its mass partial.  A launch with fewer blocks than the
// device holds leaves wave slots on every CU to kernels of other streams
// (the IB kernels beside the far cells of an overlapped step), which
// otherwise wait until the whole grid has been dispatched.
template <class LAT, class ST, int AM>
__global__ void __launch_bounds__(kBlock) k_step_p(StepArgs a, uint32_t tx, uint32_t ty, uint32_t ntiles) {
  __shared__ double lds[kBlock / 64];
  const int o0 = (a.sub_mode >= 2) ? a.sub_lo[0] : 0, o1 = (a.sub_mode >= 2) ? a.sub_lo[1] : 0;
  const int o2 = (a.sub_mode >= 2) ? a.sub_lo[2] : 0;
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const uint32_t bz = t / (tx * ty), rem = t - bz * tx * ty, by = rem / tx, bx = rem - by * tx;
    step_body<LAT, ST, AM>(a, o0 + (int)(bx * blockDim.x + threadIdx.x),
                           wave_row(o1 + (int)(by * blockDim.y + threadIdx.y)), o2 + (int)bz, (int64_t)t, lds);
  }
}

// Cell c (of nc per thread) of block b of a case's grid.  bx > 0: the block
// covers a tile of bx x (nc * by) cells, by = kBlock / bx, x fastest as
// k_step's 2D grid, cell c of a thread lying c * by rows below cell 0.
// bx = 0: block b holds cells [nc 256 b, nc 256 (b + 1)) of the case in
// x-fastest order, cell c of thread t at 256 c + t (no row padding; a cell
// past the end gets j = n[1], inactive).
// Cell c (of nc per thread) of the block whose map entry is r: bx > 0, the
// tile of bx x (nc * by) cells at r's corner, cell c of a thread c * by rows
// below cell 0; bx = 0, cells [r, r + nc * 256) in x-fastest order, cell c
// of thread t at 256 c + t (no row padding; a cell past the end gets
// j = n[1], inactive).
// ROW: bx > 0, so every wavefront lies in one row of the tile (wave_row).
template <bool ROW>
__device__ inline void batch_tile(const GridGeom& g, int bx, const BlockRef& r, int c, int* i, int* j, int* k) {
  const int t = threadIdx.x;
  if (!ROW) {
    const int nx = g.n[0], ny = g.n[1];
    // the chunk's first cell (uniform), then the lane's: each steps over at
    // most kBlock / n[0] + 1 row ends
    int i0 = r.i0 + c * kBlock, j0 = r.j0, k0 = r.k0;
    while (i0 >= nx) {
      i0 -= nx;
      if (++j0 == ny) { j0 = 0; ++k0; }
    }
    int ii = i0 + t, jj = j0, kk = k0;
    while (ii >= nx) {
      ii -= nx;
      if (++jj == ny) { jj = 0; ++kk; }
    }
    *i = ii;
    *j = (kk >= g.n[2]) ? g.n[1] : jj;
    *k = kk;
    return;
  }
  const int sh = __builtin_ctz((unsigned)bx);   // bx is 64, 128 or 256
  *i = r.i0 + (t & (bx - 1));
  *j = r.j0 + c * (kBlock >> sh) + wave_row(t >> sh);
  *k = r.k0;
}

// Many cases of one lattice / storage class in one launch.  On 2D lattices
// a case's tile_nc (uniform over its blocks) selects one or two cells per
// thread (nine loads per cell keep too few bytes in flight); 3D lattices
// take one, which keeps their register budget.
template <class LAT, class ST, int AM, bool ROW>
__device__ __forceinline__ void step_b_tile(const StepArgs& a, const BlockRef& r, double* lds) {
  if constexpr (LAT::D == 2) {
    if (a.tile_nc == 2) {
      int i[2], j[2], k[2];
      batch_tile<ROW>(a.g, a.tile_bx, r, 0, &i[0], &j[0], &k[0]);
      batch_tile<ROW>(a.g, a.tile_bx, r, 1, &i[1], &j[1], &k[1]);
      step_body_n<LAT, ST, AM, 2>(a, i, j, k, r.b, lds);
      return;
    }
  }
  int i, j, k;
  batch_tile<ROW>(a.g, a.tile_bx, r, 0, &i, &j, &k);
  step_body<LAT, ST, AM>(a, i, j, k, r.b, lds);
}

template <class LAT, class ST, int AM>
__device__ __forceinline__ void step_b_body(const StepArgs* __restrict__ A, const BlockRef* __restrict__ map,
                                            unsigned blk, double* lds) {
  const BlockRef r = map[blk];
  const StepArgs& a = A[r.c];
  if (a.tile_bx != 0) step_b_tile<LAT, ST, AM, true>(a, r, lds);
  else step_b_tile<LAT, ST, AM, false>(a, r, lds);
}

template <class LAT, class ST, int AM>
__global__ void __launch_bounds__(kBlock) k_step_b(const StepArgs* __restrict__ A, const BlockRef* __restrict__ map) {
  __shared__ double lds[kBlock / 64];
  step_b_body<LAT, ST, AM>(A, map, blockIdx.x, lds);
}

// ---------------------------------------------------------------------------
// k_init: f^0 = feq(M^0) as InitEquilibrium (AmrCoreLBM.cpp:2146-2161), then
// the first BGK collision, band macros, boundary layer and exports.
// ---------------------------------------------------------------------------
struct InitArgs {
  StepArgs s;           // gout = state, Fcol, band, faces, exports, partial
  const double* macro0; // [4][cells] device
};

template <class LAT, class ST>
__global__ void __launch_bounds__(kBlock) k_init(InitArgs ia) {
  using T = typename ST::compute_t;
  using R = typename ST::raw;
  __shared__ double lds[kBlock / 64];
  const StepArgs& a = ia.s;
  const GridGeom& g = a.g;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  double rho_acc = 0.0;
  if (i < g.n[0] && j < g.n[1]) {
    const int64_t ncell = (int64_t)g.n[0] * g.n[1] * g.n[2];
    const int64_t vcell = ((int64_t)k * g.n[1] + j) * g.n[0] + i;
    const double rho0 = ia.macro0[vcell], u0 = ia.macro0[ncell + vcell];
    const double v0 = ia.macro0[2 * ncell + vcell], w0 = ia.macro0[3 * ncell + vcell];
    const double rr = fmax(rho0, 1.0e-12);
    const double u2 = u0 * u0 + v0 * v0 + w0 * w0;
    T f[LAT::Q];
#pragma unroll
    for (int q = 0; q < LAT::Q; ++q) {
      const double cu = double(LAT::cx[q]) * u0 + double(LAT::cy[q]) * v0 + double(LAT::cz[q]) * w0;
      const double fe = LAT::w[q] * rr * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * u2);
      // stored precision, as a state written in that format would hold it
      R tmp;
      ST::store(&tmp, 0, T(fe), T(LAT::w[q]));
      f[q] = ST::load(&tmp, 0, T(LAT::w[q]));
    }
    rho_acc = rho0;
    const int bi = i - a.band_lo[0], bj = j - a.band_lo[1], bk = k - a.band_lo[2];
    const bool in_band = (unsigned)bi < (unsigned)a.band_n[0] && (unsigned)bj < (unsigned)a.band_n[1] &&
                         (unsigned)bk < (unsigned)a.band_n[2];
    const int64_t bidx = ((int64_t)bk * a.band_n[1] + bj) * a.band_n[0] + bi;
    if (in_band) {
      a.Mband[bidx] = rho0;
      a.Mband[a.band_cells + bidx] = u0;
      a.Mband[2 * a.band_cells + bidx] = v0;
      a.Mband[3 * a.band_cells + bidx] = w0;
    }
    if (a.mout) {
      a.mout[vcell] = rho0; a.mout[ncell + vcell] = u0;
      a.mout[2 * ncell + vcell] = v0; a.mout[3 * ncell + vcell] = w0;
    }
    if (a.fout) {
#pragma unroll
      for (int q = 0; q < LAT::Q; ++q) a.fout[(int64_t)q * ncell + vcell] = (double)f[q];
    }
    const int cc[3] = {i, j, k};
#pragma unroll
    for (int d = 0; d < LAT::D; ++d) {
      if (!a.nonper[d]) continue;
      for (int side = 0; side < 2; ++side) {
        if (cc[d] != (side ? g.n[d] - 1 : 0)) continue;
        R* fb = static_cast<R*>(a.faces.p[2 * d + side]);
        face_store<LAT, ST>(fb, face_index(g, d, i, j, k), f);
      }
    }
    const T r = T(rho0), u = T(u0), v = T(v0), w = T(w0);
    const T uv = usq<LAT, T>(u, v, w);
    const T cfx = T(a.Fcol[0]), cfy = T(a.Fcol[1]), cfz = T(a.Fcol[2]);
    const bool forced = !(cfx == T(0.0) && cfy == T(0.0) && cfz == T(0.0));
    const T fac = T(1.0) - T(0.5) / T(a.tau);
    R* gout = static_cast<R*>(a.gout);
    const int64_t self = g.idx(i, j, k);
#pragma unroll
    for (int q = 0; q < LAT::Q; ++q) {
      const T fe = feq_q<LAT, T>(q, r, u, v, w, uv);
      T gq = bgk<ST>(f[q], fe, a.tau, a.inv_tau);
      if (forced) gq += guo_q<LAT, T>(q, fac, u, v, w, cfx, cfy, cfz);
      ST::store(gout, (int64_t)q * g.qs + self, gq, T(LAT::w[q]));
    }
  }
  const double s = block_sum(rho_acc, lds);
  finish_mass(a, s, ((int64_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x);
}

// ---------------------------------------------------------------------------
// k_fill_ghosts: one ghost layer of the padded layout (see header comment).
// ---------------------------------------------------------------------------

struct BcSpec {
  int periodic[3];
  int lo[3], hi[3];
  double lo_val[3][4], hi_val[3][4];
};

struct GhostArgs {
  GridGeom g;
  void* gbuf;            // post-collision populations; slabs follow the Q planes
  FaceBuf faces;
  BcSpec bc;
  Slabs slabs;
};

// the ghost cell of slab entry t (a slab holds under 2^31 cells, checked at
// set-up: 32-bit divisions).  The face is the last one whose offset is at
// most t; its fields are picked with constant indices, so they stay scalar
// kernel-argument loads and selects instead of loads indexed per lane.
__device__ inline void slab_cell_of(const Slabs& S, const int n[3], int64_t t, int c[3]) {
  int fd = S.face_d[0], fs = S.face_side[0];
  int flo[3] = {S.face_lo[0][0], S.face_lo[0][1], S.face_lo[0][2]};
  int fn[3] = {S.face_n[0][0], S.face_n[0][1], S.face_n[0][2]};
  int64_t foff = S.face_off[0];
#pragma unroll
  for (int q = 1; q < 6; ++q) {
    if (q < S.nface && t >= S.face_off[q]) {
      fd = S.face_d[q];
      fs = S.face_side[q];
      foff = S.face_off[q];
#pragma unroll
      for (int e = 0; e < 3; ++e) {
        flo[e] = S.face_lo[q][e];
        fn[e] = S.face_n[q][e];
      }
    }
  }
  uint32_t r = (uint32_t)(t - foff);
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    if (e == fd) { c[e] = fs ? n[e] : -1; continue; }
    const uint32_t nn = (uint32_t)fn[e];
    const uint32_t qn = r / nn;
    c[e] = flo[e] + (int)(r - qn * nn);
    r = qn;
  }
}

__device__ inline void slab_cell(const GhostArgs& a, int64_t t, int c[3]) { slab_cell_of(a.slabs, a.g.n, t, c); }

template <class LAT, class ST>
__device__ inline void load_face(const GhostArgs& a, const int c[3], typename ST::compute_t* out) {
  using T = typename ST::compute_t;
  using R = typename ST::raw;
  const GridGeom& g = a.g;
  for (int d = 0; d < LAT::D; ++d) {
    if (a.bc.periodic[d]) continue;
    int side = -1;
    if (c[d] == 0) side = 0;
    else if (c[d] == g.n[d] - 1) side = 1;
    if (side < 0) continue;
    face_load<LAT, ST>(static_cast<const R*>(a.faces.p[2 * d + side]), face_index(g, d, c[0], c[1], c[2]), out);
    return;
  }
#pragma unroll
  for (int q = 0; q < LAT::Q; ++q) out[q] = T(0);  // unreachable for valid set-ups
}

// FillPatch value of a ghost cell from pre-collision data: FilccCell's
// foextrap chain, then the mesoscopicBcFill override (user_2 > user_1 >
// ext_dir, first matching dimension), bc_fill.H:303-347.
template <class LAT, class ST>
__device__ void fillpatch_value(const GhostArgs& a, const int iv[3], typename ST::compute_t* out) {
  using T = typename ST::compute_t;
  const GridGeom& g = a.g;
  int c[3] = {iv[0], iv[1], iv[2]};
  for (int iter = 0; iter < 8; ++iter) {
    int lo_hit[3] = {0, 0, 0}, hi_hit[3] = {0, 0, 0};
    bool any_out = false, nonper_out = false;
    for (int d = 0; d < LAT::D; ++d) {
      if (c[d] < 0) { lo_hit[d] = 1; any_out = true; if (!a.bc.periodic[d]) nonper_out = true; }
      else if (c[d] > g.n[d] - 1) { hi_hit[d] = 1; any_out = true; if (!a.bc.periodic[d]) nonper_out = true; }
    }
    if (!any_out) { load_face<LAT, ST>(a, c, out); return; }
    if (!nonper_out) {  // ParallelCopy of the periodic image
      for (int d = 0; d < LAT::D; ++d) {
        if (c[d] < 0) c[d] += g.n[d];
        else if (c[d] > g.n[d] - 1) c[d] -= g.n[d];
      }
      continue;
    }
    int kind = 0, bd = -1, lo_side = 1;
    const int types[3] = {BC_USER_2, BC_USER_1, BC_EXT_DIR};
    for (int tt = 0; tt < 3 && kind == 0; ++tt) {
      for (int d = 0; d < LAT::D; ++d) {
        if (lo_hit[d] && a.bc.lo[d] == types[tt]) { kind = types[tt]; bd = d; lo_side = 1; break; }
        if (hi_hit[d] && a.bc.hi[d] == types[tt]) { kind = types[tt]; bd = d; lo_side = 0; break; }
      }
    }
    if (kind != 0) {
      int mir[3] = {c[0], c[1], c[2]};
      for (int d = 0; d < LAT::D; ++d) {
        if (mir[d] < 0) mir[d] = 0;
        else if (mir[d] > g.n[d] - 1) mir[d] = g.n[d] - 1;
      }
      T fnp[LAT::Q];
      load_face<LAT, ST>(a, mir, fnp);
      if (kind == BC_USER_1) {
#pragma unroll
        for (int q = 0; q < LAT::Q; ++q) out[q] = fnp[q];
        return;
      }
      neep_ghost<LAT, T>(kind, fnp, lo_side ? a.bc.lo_val[bd] : a.bc.hi_val[bd], out);
      return;
    }
    // FilccCell: the last foextrap dimension in which c is outside wins
    int dsel = -1;
    for (int d = 0; d < LAT::D; ++d) {
      if ((lo_hit[d] && a.bc.lo[d] == BC_FOEXTRAP) || (hi_hit[d] && a.bc.hi[d] == BC_FOEXTRAP)) dsel = d;
    }
    if (dsel < 0) break;
    c[dsel] = lo_hit[dsel] ? 0 : g.n[dsel] - 1;
  }
#pragma unroll
  for (int q = 0; q < LAT::Q; ++q) out[q] = T(0);
}

template <class LAT, class ST>
__device__ __forceinline__ void ghost_body(const GhostArgs& a, int64_t t) {
  using T = typename ST::compute_t;
  using R = typename ST::raw;
  const int64_t npad = a.slabs.face_off[a.slabs.nface];
  if (t >= npad) return;
  const GridGeom& g = a.g;
  int iv[3];
  slab_cell(a, t, iv);
  const R* gb = static_cast<const R*>(a.gbuf);
  R* out = static_cast<R*>(a.gbuf) + g.pop_n;
  bool all_user1 = true;
  for (int d = 0; d < LAT::D; ++d) {
    const bool lo = iv[d] < 0, hi = iv[d] > g.n[d] - 1;
    if (!lo && !hi) continue;
    const int type = lo ? a.bc.lo[d] : a.bc.hi[d];
    if (type != BC_USER_1) all_user1 = false;
  }
  T val[LAT::Q];
  fillpatch_value<LAT, ST>(a, iv, val);
  if (all_user1) {  // halfway bounce-back (AdvanceLBMAtLevel.cpp:340-372)
#pragma unroll
    for (int q = 0; q < LAT::Q; ++q) {
      const int it = iv[0] + LAT::cx[q], jt = iv[1] + LAT::cy[q], kt = iv[2] + LAT::cz[q];
      if (it < 0 || it >= g.n[0] || jt < 0 || jt >= g.n[1] || kt < 0 || kt >= g.n[2]) continue;
      const int qr = opposite<LAT>(q);
      val[q] = ST::load(gb, (int64_t)qr * g.qs + g.idx(it, jt, kt), T(LAT::w[qr]));
    }
  }
#pragma unroll
  for (int q = 0; q < LAT::Q; ++q) ST::store(out, (int64_t)q * npad + t, val[q], T(LAT::w[q]));
}

template <class LAT, class ST>
__global__ void __launch_bounds__(kBlock) k_fill_ghosts(GhostArgs a) {
  ghost_body<LAT, ST>(a, (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

template <class LAT, class ST>
__global__ void __launch_bounds__(kBlock) k_fill_ghosts_b(const GhostArgs* __restrict__ A,
                                                          const BlockRef* __restrict__ map) {
  const BlockRef r = map[blockIdx.x];
  ghost_body<LAT, ST>(A[r.c], (int64_t)r.b * kBlock + threadIdx.x);
}

// Fixed-order reduction of block partials (Evolve's Sum(Rho)).
__device__ __forceinline__ double sum_partials_body(const double* p, int64_t n) {
  __shared__ double lds[16];
  double s = 0.0;
  for (int64_t t = threadIdx.x; t < n; t += blockDim.x) s += p[t];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = s;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int wv = 0; wv < (int)(blockDim.x + 63) / 64; ++wv) r += lds[wv];
  return r;
}

// Sum(Rho) from two partial arrays (a split step, or a level and its
// average-down corrections), in fixed order
__global__ void __launch_bounds__(1024) k_mass_total(const double* p0, int64_t n0, const double* p1, int64_t n1,
                                                     double* out) {
  const double a = sum_partials_body(p0, n0);
  __syncthreads();
  const double b = (n1 > 0) ? sum_partials_body(p1, n1) : 0.0;
  if (threadIdx.x == 0) *out = a + b;
}

__global__ void __launch_bounds__(1024) k_sum_partials(const double* p, int64_t n, double* out) {
  const double r = sum_partials_body(p, n);
  if (threadIdx.x == 0) *out = r;
}

// Sum(Rho) of consecutive steps whose block partials fill consecutive ring
// slots: block b reduces slot b of each partial array as k_mass_total does
// (the second and third arrays may be empty) and writes out[b]
__global__ void __launch_bounds__(1024) k_mass_ring(const double* p0, int64_t stride0, int64_t n0,
                                                    const double* p1, int64_t stride1, int64_t n1,
                                                    const double* p2, int64_t stride2, int64_t n2, double* out) {
  const int64_t b = blockIdx.x;
  const double a = sum_partials_body(p0 + b * stride0, n0);
  __syncthreads();
  const double c = (n1 > 0) ? sum_partials_body(p1 + b * stride1, n1) : 0.0;
  __syncthreads();
  const double e = (n2 > 0) ? sum_partials_body(p2 + b * stride2, n2) : 0.0;
  if (threadIdx.x == 0) out[b] = (a + c) + e;
}

// End of a batched step, one block per case: Sum(Rho) into the case's mass
// history row, then the case's clock advances (t_new += dt, step += 1) as
// timeStepWithSubcycling does (AmrCoreLBM.cpp:1730-1733).
struct CaseTail {
  const double* partial;
  int64_t nblk;
  double* mass;          // mass history rows
  double* clock;         // [0] t_new, [1] steps taken, [2] history row of step 0
  double dt;
};

__global__ void __launch_bounds__(1024) k_case_tail_b(const CaseTail* __restrict__ T) {
  const CaseTail& c = T[blockIdx.x];
  const double r = sum_partials_body(c.partial, c.nblk);
  if (threadIdx.x == 0) {
    const int row = (int)(c.clock[1] - c.clock[2]);
    c.mass[row] = r;
    c.clock[0] += c.dt;
    c.clock[1] += 1.0;
  }
}

}  // namespace ilbm
