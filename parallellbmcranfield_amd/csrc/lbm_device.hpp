// lbm_device.hpp — device-side building blocks of the D2Q9 step.
//
// One step of the reference loop (serial/Iterate.c:263-332) is
//   collide -> UpdateF -> pull-stream -> boundaries -> macroscopic update.
// Stored state here is the post-collision field (the reference's METAF), so
// a fused pass per node does: pull METAF_t from the upstream neighbours
// (stream), apply the node's boundary rules, form rho/u/v (the reference's
// macroscopic update of step t), then collide with them (the collision that
// opens step t+1) and store METAF_{t+1} into the other buffer.  Macros and
// the post-boundary F of step t therefore exist in registers only; "observe"
// passes recompute them from the previous buffer when the host asks.
//
// Storage kinds:
//   F64 : absolute f in FP64; arithmetic in FP64 with the reference's operand
//         order (compiled with -ffp-contract=off) -> bit-identical to the C.
//   F32 : g = f - w_k*rho0 in FP32; FP32 arithmetic on deviations.
//   F16 : g*scale in FP16 (fixed power-of-two scale); FP32 arithmetic.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lbm_kparams.hpp"

namespace lbmdev {

// D2Q9 in the reference's numbering (serial/CellFunctions.c:281-340).
__device__ constexpr int CX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
__device__ constexpr int CY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
__device__ constexpr int OPP[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
__device__ constexpr double W64[9] = {4. / 9., 1. / 9., 1. / 9., 1. / 9., 1. / 9.,
                                      1. / 36., 1. / 36., 1. / 36., 1. / 36.};
__device__ constexpr float W32[9] = {4.f / 9.f, 1.f / 9.f, 1.f / 9.f, 1.f / 9.f, 1.f / 9.f,
                                     1.f / 36.f, 1.f / 36.f, 1.f / 36.f, 1.f / 36.f};

// Non-zero pattern of M^-1 (serial/CellFunctions.c:24-35): the reference's
// dense sums over these rows add exact zeros elsewhere, so skipping them
// leaves every FP64 result unchanged.
__device__ constexpr bool MINV_NZ[9][9] = {
    {1, 1, 1, 0, 0, 0, 0, 0, 0}, {1, 1, 1, 1, 1, 0, 0, 1, 0}, {1, 1, 1, 0, 0, 1, 1, 1, 0},
    {1, 1, 1, 1, 1, 0, 0, 1, 0}, {1, 1, 1, 0, 0, 1, 1, 1, 0}, {1, 1, 1, 1, 1, 1, 1, 0, 1},
    {1, 1, 1, 1, 1, 1, 1, 0, 1}, {1, 1, 1, 1, 1, 1, 1, 0, 1}, {1, 1, 1, 1, 1, 1, 1, 0, 1}};
// Moment matrix M (serial/CellFunctions.c:37-47).
__device__ constexpr int MM[9][9] = {
    {1, 1, 1, 1, 1, 1, 1, 1, 1},     {-4, -1, -1, -1, -1, 2, 2, 2, 2},
    {4, -2, -2, -2, -2, 1, 1, 1, 1}, {0, 1, 0, -1, 0, 1, -1, -1, 1},
    {0, -2, 0, 2, 0, 1, -1, -1, 1},  {0, 0, 1, 0, -1, 1, 1, -1, -1},
    {0, 0, -2, 0, 2, 1, 1, -1, -1},  {0, 1, -1, 1, -1, 0, 0, 0, 0},
    {0, 0, 0, 0, 0, 1, -1, 1, -1}};

// ---------------------------------------------------------------- storage
// Loads of the distributions: every value is read once per step (plus the
// neighbouring tile's edge line), so LBM_NT_LOADS=1 marks them non-temporal.
#ifndef LBM_NT_LOADS
#define LBM_NT_LOADS 0
#endif
template <class T>
__device__ inline T ld(const T* p) {
  if constexpr (LBM_NT_LOADS)
    return __builtin_nontemporal_load(p);
  else
    return *p;
}

// Word i of direction plane k of a state buffer.  The plane's base is the
// same in every lane (scalar registers) and the in-plane byte offset fits in
// 32 bits (a plane holds fewer than 2^31 / 9 words), so the access takes the
// scalar-base plus 32-bit vector-offset form, without 64-bit address
// arithmetic per lane.
template <class T>
__device__ inline T* plane_word(T* base, int plane, int k, int i) {
  T* pk = base + (size_t)k * (size_t)plane;
  return (T*)((char*)pk + (uint32_t)i * (uint32_t)sizeof(T));
}

// store() writes a step's result; store_rn() writes fixed values (the solid
// value, set-up fills) rounded to nearest in every storage kind.
struct StF64 {
  using T = double;
  using R = double;  // arithmetic type
  static constexpr bool DEV = false;
  static constexpr bool SR = false;
  __device__ static R load(const T* p, const KParams&) { return ld(p); }
  __device__ static R round_trip(R v, const KParams&) { return v; }
  __device__ static void store(T* p, R v, const KParams&, uint32_t) { __builtin_nontemporal_store(v, p); }
  __device__ static void store_rn(T* p, R v, const KParams& kp) { store(p, v, kp, 0u); }
  __device__ static void store_cell(T* dst, int idx, const R (&v)[9], const KParams& kp, uint32_t) {
#pragma unroll
    for (int k = 0; k < 9; k++) store(plane_word(dst, kp.plane, k, idx), v[k], kp, 0u);
  }
};
struct StF32 {
  using T = float;
  using R = float;
  static constexpr bool DEV = true;
  static constexpr bool SR = false;
  __device__ static R load(const T* p, const KParams&) { return ld(p); }
  __device__ static R round_trip(R v, const KParams&) { return v; }
  __device__ static void store(T* p, R v, const KParams&, uint32_t) { __builtin_nontemporal_store(v, p); }
  __device__ static void store_rn(T* p, R v, const KParams& kp) { store(p, v, kp, 0u); }
  __device__ static void store_cell(T* dst, int idx, const R (&v)[9], const KParams& kp, uint32_t) {
#pragma unroll
    for (int k = 0; k < 9; k++) store(plane_word(dst, kp.plane, k, idx), v[k], kp, 0u);
  }
};
struct StF16 {
  using T = _Float16;
  using R = float;
  static constexpr bool DEV = true;
  static constexpr bool SR = false;
  __device__ static R load(const T* p, const KParams& kp) { return (float)ld(p) * kp.inv_scale_f; }
  // what load() returns for a value store() wrote
  __device__ static R round_trip(R v, const KParams& kp) {
    return (float)(_Float16)(v * kp.scale_f) * kp.inv_scale_f;
  }
  __device__ static void store(T* p, R v, const KParams& kp, uint32_t) {
    __builtin_nontemporal_store((_Float16)(v * kp.scale_f), p);
  }
  __device__ static void store_rn(T* p, R v, const KParams& kp) { store(p, v, kp, 0u); }
  __device__ static void store_cell(T* dst, int idx, const R (&v)[9], const KParams& kp, uint32_t) {
#pragma unroll
    for (int k = 0; k < 9; k++) store(plane_word(dst, kp.plane, k, idx), v[k], kp, 0u);
  }
};

// Stochastic rounding to FP16.  Round to nearest (StF16) stores every value
// with an error that depends only on the value, so where a population's
// per-step change is below half an FP16 step it is lost the same way step
// after step, and the errors add up coherently; with the neighbour chosen at
// random in proportion to the distance to each (expected value exact) they
// add up as a random walk instead.  The random words hash the value's place
// in the buffer and the step number, so a run repeats bit for bit.
//
// 32-bit integer hash: the "lowbias32" finaliser (xor-shift / multiply
// constants from C. Wellons' hash-prospector search, public domain).
__device__ inline uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}
// The random words of a cell's nine stores in a step: one hash of the cell
// pair (cells 2q and 2q + 1 of a plane share a 4-byte word) and the step
// number, then one xorshift32 step (Marsaglia 2003) per direction, so the
// paired kernel and the one-cell paths draw the same words.  The second
// cell of a pair takes the word rotated by 16 bits, so the two draw on
// different bits whichever of the word's bits the conversion reads.
__device__ inline uint32_t sr_pair_bits(int idx, uint32_t step) {
  return mix32((uint32_t)(idx >> 1) ^ (step * 0x9e3779b9u));
}
__device__ inline uint32_t xorshift32(uint32_t x) {
  x ^= x << 13;
  x ^= x >> 17;
  x ^= x << 5;
  return x;
}
__device__ inline uint32_t rot16(uint32_t x) { return (x >> 16) | (x << 16); }
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
// v rounded to one of its two FP16 neighbours at random, the upper with
// probability (distance to the lower) / (their spacing): gfx950's
// v_cvt_sr_f16_f32 adds random bits below FP16's precision before it
// truncates, with r as the random word; `second` picks the half it writes.
__device__ inline half2_t f16_stochastic_into(half2_t x, float v, uint32_t r, bool second) {
  return second ? __builtin_amdgcn_cvt_sr_f16_f32(x, v, r, true)
                : __builtin_amdgcn_cvt_sr_f16_f32(x, v, r, false);
}
struct StF16S : StF16 {
  static constexpr bool SR = true;
  __device__ static void store_cell(T* dst, int idx, const R (&v)[9], const KParams& kp,
                                    uint32_t step) {
    uint32_t r = sr_pair_bits(idx, step);
    const bool second = (idx & 1) != 0;
#pragma unroll
    for (int k = 0; k < 9; k++) {
      half2_t x = {(_Float16)0, (_Float16)0};
      x = f16_stochastic_into(x, v[k] * kp.scale_f, second ? rot16(r) : r, false);
      __builtin_nontemporal_store(x[0], plane_word(dst, kp.plane, k, idx));
      r = xorshift32(r);
    }
  }
  __device__ static void store_rn(T* p, R v, const KParams& kp) { StF16::store(p, v, kp, 0u); }
};

// The step number a launch stores with (stochastic-rounding storage; 0
// otherwise), read once at the start of the kernel and handed to the stores:
// read inside each store it would be re-read after every store to memory of
// the same type.  A step's launches (MODE 0 / 1) also hand it on to the next
// step from their first workgroup.
template <class S>
__device__ inline uint32_t step_number(const KParams& p) {
  if constexpr (S::SR)
    return p.seed[p.seed_slot];
  else
    return 0u;
}
template <class S>
__device__ inline void advance_seed(const KParams& p, uint32_t step) {
  if constexpr (S::SR) {
    if (blockIdx.x == 0 && threadIdx.x == 0) p.seed[p.seed_slot ^ 1] = step + 1u;
  }
}

// ---------------------------------------------------------------- collision
// Reference-order FP64 kernels (absolute f).
// BGKW: serial/CellFunctions.c:352-365
__device__ inline void collide_bgk_f64(const double F[9], double R, double U, double V,
                                       const KParams& p, double out[9]) {
  const double om = p.omega;
  const double T1 = U * U + V * V;
#pragma unroll
  for (int k = 0; k < 9; k++) {
    const double T2 = U * (double)CX[k] + V * (double)CY[k];
    const double feq = R * W64[k] * (1.0 + 3.0 * T2 + 4.5 * T2 * T2 - 1.5 * T1);
    out[k] = om * feq + (1.0 - om) * F[k];
  }
}

// TRT: serial/CellFunctions.c:372-395 (binary32 parts unless trt_fp64)
__device__ inline void collide_trt_f64(const double F[9], double R, double U, double V,
                                       const KParams& p, double out[9]) {
  const double om = p.omega, oma = p.omega_a;
  const double T1 = U * U + V * V;
  double Feq[9];
#pragma unroll
  for (int k = 0; k < 9; k++) {
    const double T2 = U * (double)CX[k] + V * (double)CY[k];
    Feq[k] = R * W64[k] * (1.0 + 3.0 * T2 + 4.5 * T2 * T2 - 1.5 * T1);
  }
  if (!(p.flags & KF_TRT_FP64)) {
#pragma unroll
    for (int k = 0; k < 9; k++) {
      const int o = OPP[k];
      const float fp = (float)(0.5 * (F[k] + F[o]));
      const float fep = (float)(0.5 * (Feq[k] + Feq[o]));
      const float fm = (float)(0.5 * (F[k] - F[o]));
      const float fem = (float)(0.5 * (Feq[k] - Feq[o]));
      const float dp = fp - fep;
      const float dm = fm - fem;
      out[k] = F[k] - (double)dp * om - (double)dm * oma;
    }
  } else {
#pragma unroll
    for (int k = 0; k < 9; k++) {
      const int o = OPP[k];
      const double fp = 0.5 * (F[k] + F[o]);
      const double fep = 0.5 * (Feq[k] + Feq[o]);
      const double fm = 0.5 * (F[k] - F[o]);
      const double fem = 0.5 * (Feq[k] - Feq[o]);
      out[k] = F[k] - (fp - fep) * om - (fm - fem) * oma;
    }
  }
}

// MRT: serial/CellFunctions.c:401-438, M*F with M's zeros skipped (exact),
// stmiv = M^-1 S prepared on the host exactly as MRTInitializer does.
__device__ inline void collide_mrt_f64(const double F[9], double Rh, double U, double V,
                                       const KParams& p, double out[9]) {
  double meq[9], mom[9];
  meq[0] = Rh;
  if (!(p.flags & KF_MRT_TEXTBOOK)) {
    meq[1] = Rh * (-2.0 + 3.0 * Rh * (U * U + V * V));
    meq[2] = Rh * (1.0 - 3.0 * Rh * (U * U + V * V));
  } else {
    meq[1] = Rh * (-2.0 + 3.0 * (U * U + V * V));
    meq[2] = Rh * (1.0 - 3.0 * (U * U + V * V));
  }
  meq[3] = Rh * U;
  meq[4] = -Rh * U;
  meq[5] = Rh * V;
  meq[6] = -Rh * V;
  meq[7] = Rh * (U * U - V * V);
  meq[8] = Rh * U * V;
#pragma unroll
  for (int k = 0; k < 9; k++) {
    double a = 0.0;
#pragma unroll
    for (int l = 0; l < 9; l++)
      if (MM[k][l] != 0) a = a + (double)MM[k][l] * F[l];
    mom[k] = a;
  }
#pragma unroll
  for (int k = 0; k < 9; k++) {
    double b = 0.0;
#pragma unroll
    for (int l = 0; l < 9; l++)
      if (MINV_NZ[k][l]) b = b + p.stmiv[k][l] * (mom[l] - meq[l]);
    out[k] = F[k] - b;
  }
}

template <int COLL>
__device__ inline void collide_f64(const double F[9], double R, double U, double V,
                                   const KParams& p, double out[9]) {
  if constexpr (COLL == 1) collide_bgk_f64(F, R, U, V, p, out);
  if constexpr (COLL == 2) collide_trt_f64(F, R, U, V, p, out);
  if constexpr (COLL == 3) collide_mrt_f64(F, R, U, V, p, out);
}

// Relaxation parameters of the deviation-form kernels: the solver's global
// values, or one instance's row of the batched parameter table.
struct CollPar {
  float om, oma;  // omega (= 1/tau), TRT omega-minus
  float e[9];     // MRT: s_l / |M_l|^2
  int flags;      // KF_MRT_TEXTBOOK
};

__device__ inline CollPar coll_par(const KParams& p, int inst) {
  CollPar c;
  c.flags = p.flags;
  if (inst < 0) {
    c.om = p.omega_f;
    c.oma = p.omega_a_f;
#pragma unroll
    for (int l = 0; l < 9; l++) c.e[l] = p.mrt_e[l];
  } else {
    const float* r = p.inst_par + (size_t)inst * INST_PAR_W;
    c.om = r[0];
    c.oma = r[1];
#pragma unroll
    for (int l = 0; l < 9; l++) c.e[l] = r[2 + l];
  }
  return c;
}

// Deviation-form FP32 kernels: g_k = f_k - w_k rho0; drho = sum g; rho = rho0 + drho.
// feq_k - w_k rho0 = w_k (drho + rho (3 cu + 4.5 cu^2 - 1.5 u^2)).
__device__ inline void collide_bgk_dev(const float g[9], float drho, float rho, float u, float v,
                                       const CollPar& p, float out[9]) {
  const float om = p.om;
  const float usq15 = 1.5f * (u * u + v * v);
#pragma unroll
  for (int k = 0; k < 9; k++) {
    const float cu = (float)CX[k] * u + (float)CY[k] * v;
    const float geq = W32[k] * (drho + rho * (3.f * cu + 4.5f * cu * cu - usq15));
    out[k] = g[k] + om * (geq - g[k]);
  }
}

__device__ inline void collide_trt_dev(const float g[9], float drho, float rho, float u, float v,
                                       const CollPar& p, float out[9]) {
  const float op = p.om, om = p.oma;
  const float usq15 = 1.5f * (u * u + v * v);
  // rest population: symmetric only
  {
    const float geq0 = W32[0] * (drho - rho * usq15);
    out[0] = g[0] - op * (g[0] - geq0);
  }
  // direction pairs (1,3) (2,4) (5,7) (6,8)
  constexpr int PA[4] = {1, 2, 5, 6};
#pragma unroll
  for (int t = 0; t < 4; t++) {
    const int k = PA[t], o = OPP[k];
    const float cu = (float)CX[k] * u + (float)CY[k] * v;
    const float gsym = 0.5f * (g[k] + g[o]);
    const float gasy = 0.5f * (g[k] - g[o]);
    const float esym = W32[k] * (drho + rho * (4.5f * cu * cu - usq15));
    const float easy = W32[k] * rho * 3.f * cu;
    const float a = op * (gsym - esym);
    const float b = om * (gasy - easy);
    out[k] = g[k] - a - b;
    out[o] = g[o] - a + b;
  }
}

__device__ inline void collide_mrt_dev(const float g[9], float drho, float rho, float u, float v,
                                       const CollPar& p, float out[9]) {
  // No implicit contraction: the fused multiply-adds are written out, so
  // every kernel that inlines this (one cell or two per thread, one step or
  // several per launch) rounds the same way.
#pragma clang fp contract(off)
  // moments of g (rows of M)
  const float s14 = g[1] + g[2] + g[3] + g[4];
  const float s58 = g[5] + g[6] + g[7] + g[8];
  const float jx = (g[1] - g[3]) + (g[5] - g[7]) + (g[8] - g[6]);
  const float jy = (g[2] - g[4]) + (g[5] - g[7]) + (g[6] - g[8]);
  const float m1 = fmaf(2.f, s58, fmaf(-4.f, g[0], -s14));
  const float m2 = fmaf(4.f, g[0], fmaf(-2.f, s14, s58));
  const float m4 = fmaf(2.f, g[3] - g[1], (g[5] - g[7]) + (g[8] - g[6]));
  const float m6 = fmaf(2.f, g[4] - g[2], (g[5] - g[7]) + (g[6] - g[8]));
  const float m7 = (g[1] + g[3]) - (g[2] + g[4]);
  const float m8 = (g[5] + g[7]) - (g[6] + g[8]);
  const float usq = fmaf(u, u, v * v);
  const float r3 = (p.flags & KF_MRT_TEXTBOOK) ? 3.f * rho * usq : 3.f * rho * rho * usq;
  // d_l = m_l(f) - m_l^eq; M w = (1,-2,1,0,...) absorbs rho0
  const float d1 = fmaf(2.f, drho, m1) - r3;
  const float d2 = (m2 - drho) + r3;
  const float d3 = fmaf(-rho, u, jx);
  const float d4 = fmaf(rho, u, m4);
  const float d5 = fmaf(-rho, v, jy);
  const float d6 = fmaf(rho, v, m6);
  const float d7 = fmaf(-rho, fmaf(u, u, -(v * v)), m7);
  const float d8 = fmaf(-rho, u * v, m8);
  // e_l = s_l d_l / |M_l|^2 ; out = g - M^T e, with the terms the columns of
  // M share formed once: columns 1-4 hold -e1 - 2 e2, columns 5-8 2 e1 + e2,
  // and the momentum / energy-flux rows enter as (e3 -+ 2 e4), (e5 -+ 2 e6)
  // or (e3 + e4), (e5 + e6) with the signs of c_x, c_y
  const float e1 = p.e[1] * d1, e2 = p.e[2] * d2, e3 = p.e[3] * d3, e4 = p.e[4] * d4;
  const float e5 = p.e[5] * d5, e6 = p.e[6] * d6, e7 = p.e[7] * d7, e8 = p.e[8] * d8;
  const float a = fmaf(-2.f, e2, -e1), b = fmaf(2.f, e1, e2);
  const float X = fmaf(-2.f, e4, e3), Y = fmaf(-2.f, e6, e5), P = e3 + e4, Q = e5 + e6;
  const float ap = a + e7, am = a - e7, bp = b + e8, bm = b - e8;
  const float PQ = P + Q, PmQ = P - Q;
  out[0] = fmaf(-4.f, e2 - e1, g[0]);
  out[1] = g[1] - (ap + X);
  out[2] = g[2] - (am + Y);
  out[3] = g[3] - (ap - X);
  out[4] = g[4] - (am - Y);
  out[5] = g[5] - (bp + PQ);
  out[6] = g[6] - (bm - PmQ);
  out[7] = g[7] - (bp - PQ);
  out[8] = g[8] - (bm + PmQ);
}

template <int COLL>
__device__ inline void collide_dev(const float g[9], float drho, float rho, float u, float v,
                                   const CollPar& p, float out[9]) {
  if constexpr (COLL == 1) collide_bgk_dev(g, drho, rho, u, v, p, out);
  if constexpr (COLL == 2) collide_trt_dev(g, drho, rho, u, v, p, out);
  if constexpr (COLL == 3) collide_mrt_dev(g, drho, rho, u, v, p, out);
}

// ---------------------------------------------------------------- macros
// serial/CellFunctions.c:462-481, summation order k = 0..8.
__device__ inline void macro_f64(const double F[9], double& R, double& U, double& V) {
  double S = 0.0;
#pragma unroll
  for (int k = 0; k < 9; k++) S = S + F[k];
  R = S;
  double us = 0.0, vs = 0.0;
#pragma unroll
  for (int k = 0; k < 9; k++) {
    us = us + F[k] * (double)CX[k];
    vs = vs + F[k] * (double)CY[k];
  }
  U = us / R;
  V = vs / R;
}

__device__ inline void macro_dev(const float g[9], float rho0, float& drho, float& rho, float& u,
                                 float& v) {
  drho = ((g[0] + g[1]) + (g[2] + g[3])) + ((g[4] + g[5]) + (g[6] + g[7])) + g[8];
  rho = rho0 + drho;
  const float jx = (g[1] - g[3]) + (g[5] - g[7]) + (g[8] - g[6]);
  const float jy = (g[2] - g[4]) + (g[5] - g[7]) + (g[6] - g[8]);
  // the hardware reciprocal (within 1 ulp) and one Newton step instead of the
  // IEEE division's ten-instruction sequence: 343 -> 337 us per three-step
  // launch at 4097^2.  Without the Newton step the reciprocal's error, the
  // same for every node of equal density, drifts the velocity: FP32 BGK at
  // 1025^2 after 2000 steps 1.3e-5 from FP64 against 5.2e-6 with it, as
  // with the division.
  const float ir0 = __builtin_amdgcn_rcpf(rho);
  const float ir = fmaf(fmaf(-rho, ir0, 1.f), ir0, ir0);
  u = jx * ir;
  v = jy * ir;
}

// ---------------------------------------------------------------- XCD-aware block order
// Workgroups are dispatched round-robin over the 8 XCDs (each with its own L2);
// remapping gives each XCD a contiguous run of tiles so that horizontally
// adjacent tiles (which share cache lines through the +-1 pulls) meet in the
// same L2.  Needs gridDim.x % 8 == 0 (the host pads the grid).
__device__ inline unsigned xcd_block(unsigned b, unsigned nb) {
  return (b & 7u) * (nb >> 3) + (b >> 3);
}
// Grouped order: runs of G consecutive tiles go to one XCD, so adjacent tiles
// share their edge lines in one L2 while all XCDs still stream neighbouring
// addresses.  Needs gridDim.x % (8 G) == 0.
__device__ inline unsigned xcd_group_block(unsigned b, unsigned G) {
  const unsigned r = b >> 3, x = b & 7u;
  return (r / G) * (8u * G) + x * G + (r % G);
}

}  // namespace lbmdev
