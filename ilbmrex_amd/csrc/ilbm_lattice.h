// Lattice velocity sets and population storage formats for the MI355X path.
//
// Direction order and weights are those of the reference's example inputs,
// with w = wis / wim evaluated in double precision exactly as
// AmrCoreLBM::ReadParameters does (Source/Core/AmrCoreLBM.cpp:1218-1221):
//   D2Q9  Examples/Cylinder_flow/inputs:35-40
//   D3Q19 Examples/SquareDuct_flow/inputs:36-41
//   D3Q27 Examples/Sphere_flow/inputs:34-39
// Keeping the order matters: the macroscopic sums run over q in this order.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ilbm {

struct D2Q9 {
  static constexpr int D = 2, Q = 9, ID = 0;
  static constexpr int cx[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
  static constexpr int cy[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
  static constexpr int cz[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  static constexpr double w[9] = {4.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,
                                  1.0 / 9.0,  1.0 / 9.0,  1.0 / 36.0,
                                  1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
};

struct D3Q19 {
  static constexpr int D = 3, Q = 19, ID = 1;
  static constexpr int cx[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
  static constexpr int cy[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
  static constexpr int cz[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
  static constexpr double w[19] = {
      12.0 / 36.0, 2.0 / 36.0, 2.0 / 36.0, 2.0 / 36.0, 2.0 / 36.0,
      2.0 / 36.0,  2.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
      1.0 / 36.0,  1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
      1.0 / 36.0,  1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
};

struct D3Q27 {
  static constexpr int D = 3, Q = 27, ID = 2;
  static constexpr int cx[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1,
                                 -1, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1};
  static constexpr int cy[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1,
                                 1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1};
  static constexpr int cz[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0,
                                 0, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1};
  static constexpr double w[27] = {
      8.0 / 27.0,  2.0 / 27.0,  2.0 / 27.0,  2.0 / 27.0,  2.0 / 27.0,
      2.0 / 27.0,  2.0 / 27.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,
      1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,
      1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 216.0,
      1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0,
      1.0 / 216.0, 1.0 / 216.0};
};

template <class LAT>
__host__ __device__ constexpr int opposite(int q) {
  for (int p = 0; p < LAT::Q; ++p)
    if (LAT::cx[p] == -LAT::cx[q] && LAT::cy[p] == -LAT::cy[q] && LAT::cz[p] == -LAT::cz[q])
      return p;
  return q;
}

// ---------------------------------------------------------------------------
// Population storage.  compute_t is the collision arithmetic type.
//   StF64  : FP64 storage, FP64 arithmetic (the parity path)
//   StF32  : FP32 storage, FP32 arithmetic
//   StF16S : FP16 storage of the shifted population f_q - w_q, FP32 arithmetic
// load / store take and give f_q.  The stream-collide kernel works in the
// stored variable when `shifted` is set (h_q = f_q - w_q for StF16S):
// load_s / load_s_nt / store_s_nt move it without the +-w_q.
// ---------------------------------------------------------------------------
struct StF64 {
  using raw = double;
  using compute_t = double;
  static constexpr int ID = 0;
  static constexpr bool exact = true;
  static constexpr bool shifted = false;
  __device__ static inline double load(const raw* p, int64_t i, double) { return p[i]; }
  __device__ static inline double load_nt(const raw* p, int64_t i, double) { return __builtin_nontemporal_load(p + i); }
  __device__ static inline void store(raw* p, int64_t i, double v, double) { p[i] = v; }
  __device__ static inline void store_nt(raw* p, int64_t i, double v, double) {
    __builtin_nontemporal_store(v, p + i);
  }
};

struct StF32 {
  using raw = float;
  using compute_t = float;
  static constexpr int ID = 1;
  static constexpr bool exact = false;
  static constexpr bool shifted = false;
  __device__ static inline float load(const raw* p, int64_t i, float) { return p[i]; }
  __device__ static inline float load_nt(const raw* p, int64_t i, float) { return __builtin_nontemporal_load(p + i); }
  __device__ static inline void store(raw* p, int64_t i, float v, float) { p[i] = v; }
  __device__ static inline void store_nt(raw* p, int64_t i, float v, float) {
    __builtin_nontemporal_store(v, p + i);
  }
};

struct StF16S {
  using raw = __half;
  using compute_t = float;
  static constexpr int ID = 2;
  static constexpr bool exact = false;
  static constexpr bool shifted = true;
  __device__ static inline float load_s(const raw* p, int64_t i) { return __half2float(p[i]); }
  __device__ static inline float load_s_nt(const raw* p, int64_t i) {
    return __half2float(__ushort_as_half(__builtin_nontemporal_load(reinterpret_cast<const unsigned short*>(p) + i)));
  }
  __device__ static inline void store_s(raw* p, int64_t i, float h) { p[i] = __float2half(h); }
  __device__ static inline void store_s_nt(raw* p, int64_t i, float h) {
    __builtin_nontemporal_store(__half_as_ushort(__float2half(h)), reinterpret_cast<unsigned short*>(p) + i);
  }
  __device__ static inline float load(const raw* p, int64_t i, float wq) {
    return __half2float(p[i]) + wq;
  }
  __device__ static inline float load_nt(const raw* p, int64_t i, float wq) {
    const unsigned short bits = __builtin_nontemporal_load(reinterpret_cast<const unsigned short*>(p) + i);
    return __half2float(__ushort_as_half(bits)) + wq;
  }
  __device__ static inline void store(raw* p, int64_t i, float v, float wq) {
    p[i] = __float2half(v - wq);
  }
  __device__ static inline void store_nt(raw* p, int64_t i, float v, float wq) {
    __half h = __float2half(v - wq);
    unsigned short bits = __half_as_ushort(h);
    __builtin_nontemporal_store(bits, reinterpret_cast<unsigned short*>(p) + i);
  }
};

template <class ST>
constexpr size_t raw_size() { return sizeof(typename ST::raw); }

// The stream-collide kernel's working variable: f_q, or h_q = f_q - w_q for
// shifted storage.
template <class ST>
__device__ inline typename ST::compute_t pop_in(const typename ST::raw* p, int64_t i, typename ST::compute_t wq) {
  if constexpr (ST::shifted) return ST::load_s(p, i);
  else return ST::load(p, i, wq);
}
template <class ST>
__device__ inline typename ST::compute_t pop_in_nt(const typename ST::raw* p, int64_t i, typename ST::compute_t wq) {
  if constexpr (ST::shifted) return ST::load_s_nt(p, i);
  else return ST::load_nt(p, i, wq);
}
template <class ST>
__device__ inline void pop_out_nt(typename ST::raw* p, int64_t i, typename ST::compute_t v, typename ST::compute_t wq) {
  if constexpr (ST::shifted) ST::store_s_nt(p, i, v);
  else ST::store_nt(p, i, v, wq);
}

}  // namespace ilbm
