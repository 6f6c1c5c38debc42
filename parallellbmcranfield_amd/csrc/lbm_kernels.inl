// lbm_kernels.inl — kernel templates, included by one translation unit per
// floating-point contract (kernels_f64.hip: -ffp-contract=off for the
// bit-exact FP64 path; kernels_lp.hip: FP32/FP16 storage).
//
// Passes of one step (all read buffer `src` = METAF_t, write `dst` = METAF_{t+1}):
//   k_bulk      interior fluid nodes (class 1): pull 9, macros, collide, store.
//   k_boundary  boundary-list nodes (class 2): pull with stream mask, the
//               reference's boundary rules, macros, collide, store.
//   k_outlet    open-outlet nodes (class 3), one thread per row, in increasing
//               i, after the boundary pass has left the post-boundary F of
//               their upstream neighbours in the side buffer.
// MODE 0: step.  MODE 1: step + residual partials.  MODE 2: observe (no
// state write; macros / F / METAF of step t to the KOut arrays).
#pragma once
#include <type_traits>

#include "lbm_device.hpp"

namespace lbmdev {

// ---------------------------------------------------------------- helpers
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Deterministic per-workgroup sum of NP values -> part[slot*NP + c].
constexpr int NP = 6;  // 4 residual terms + momentum-exchange force (x, y)
template <int BS>
__device__ inline void block_sum(const double (&a)[NP], double* part, int slot) {
  __shared__ double red[BS / 64][NP];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < NP; c++) {
    const double v = wave_sum(a[c]);
    if (lane == 0) red[wv][c] = v;
  }
  __syncthreads();
  if (threadIdx.x < NP) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < BS / 64; w++) s += red[w][threadIdx.x];
    part[(size_t)slot * NP + threadIdx.x] = s;
  }
}

template <class S>
__device__ inline double absval(typename S::R g, int k, const KParams& p) {
  if constexpr (S::DEV)
    return (double)g + W64[k] * p.rho0;
  else
    return (double)g;
}

// Reference boundary rules on one node's post-streaming F (in place), in the
// order of serial/Iterate.c:499-546: InletBC, then curved or straight walls,
// then (profile 1) the velocity outlet.  Open outlets (profiles 2/3) are
// applied by the caller afterwards.  R is double (absolute values) or float
// (deviations, where every population-weighted density sum gains rho0).
template <class S>
__device__ inline void apply_bc(typename S::R F[9], const BNode& nd, const typename S::T* src,
                                const KParams& p, int idx) {
  using R = typename S::R;
  const R uo = (R)nd.uo;
  const R roff = S::DEV ? (R)p.rho0 : (R)0;
  // InletBC: serial/BoundaryConditions.c:14-58
  if (nd.boundary == 2) {
    if (nd.bc[2] == 2 && nd.corner != 1) {
      R rhoN = F[0] + F[1] + F[3] + 2 * (F[2] + F[6] + F[5]);
      if constexpr (S::DEV) rhoN = rhoN + roff;
      F[4] = F[2];
      F[8] = F[6] + rhoN * uo / (R)6.0;
      F[7] = F[5] - rhoN * uo / (R)6.0;
    }
    if (nd.bc[3] == 2 && nd.corner != 1) {
      R num = F[0] + F[2] + F[4] + (R)2.0 * (F[3] + F[6] + F[7]);
      if constexpr (S::DEV) num = num + roff;
      const R rhoW = num / ((R)1.0 - uo);
      F[1] = F[3] + 2 * rhoW * uo / 3;
      F[5] = F[7] + rhoW * uo / 6;
      F[8] = F[6] + rhoW * uo / 6;
    }
  }
  if (p.curved == 1) {
    // CurvedWallBoundaries: serial/BoundaryConditions.c:127-146 with
    // Fneighbours[k] = METAF[k] of the upstream node (serial/Iterate.c:509-510)
#pragma unroll
    for (int k = 0; k < 9; k++) {
      if (nd.bc[k] == 1) {
        const int o = OPP[k];
        const R q = (R)nd.q[k];
        const R mk = S::load(src + k * p.plane + idx, p);
        if (nd.q[k] < 0.5) {
          const R nk = S::load(src + k * p.plane + idx - CX[k] - CY[k] * p.pitch, p);
          F[o] = 2 * q * mk + (1 - 2 * q) * nk;
        } else {
          const R mo = S::load(src + o * p.plane + idx, p);
          F[o] = mk / 2 / q + (2 * q - 1) / (2 * q) * mo;
        }
      }
    }
  } else if (p.curved == 2) {
    // WallBC: serial/BoundaryConditions.c:104-118 (post-streaming F[k])
#pragma unroll
    for (int k = 0; k < 9; k++)
      if (nd.bc[k] == 1) F[OPP[k]] = F[k];
  }
  if (p.outlet_profile == 4 && nd.bc[1] == 3) {
    // East pressure outlet, Zou & He (1997): rho = rho_out, v = 0; unknown
    // populations 3, 6, 7 from mass and momentum balance.
    const R ro = S::DEV ? (R)p.rho_out_f : (R)p.rho_out;
    R num = F[0] + F[2] + F[4] + (R)2.0 * (F[1] + F[5] + F[8]);
    if constexpr (S::DEV) num = num + roff;
    const R ux = num / ro - (R)1.0;
    const R t = ro * ux / (R)6.0;
    const R h = (R)0.5 * (F[2] - F[4]);
    F[3] = F[1] - (R)4.0 * t;
    F[7] = F[5] + h - t;
    F[6] = F[8] - h - t;
  }
  if (p.outlet_profile == 1 && nd.bc[1] == 3) {
    // OutletBoundaries: serial/BoundaryConditions.c:67-81
    R num = F[0] + F[2] + F[4] + (R)2.0 * (F[1] + F[5] + F[8]);
    if constexpr (S::DEV) num = num + roff;
    const R rhoE = num / (1 - uo);
    F[3] = F[1] - 2 * rhoE * uo / (R)3.0;
    F[7] = F[5] - rhoE * uo / (R)6.0;
    F[6] = F[8] - rhoE * uo / (R)6.0;
  }
}

// Everything after the boundary rules for one node: macros (with the
// outlet's V = 0 and the drag/lift placeholder), residual terms, observe
// outputs, collision and store.  acc[] receives residual terms (MODE 1).
template <class S, int COLL, int MODE, bool BND>
__device__ inline void finish_node(typename S::R F[9], const typename S::T* src,
                                   typename S::T* dst, const KParams& p, const KOut& o, int idx,
                                   const BNode* nd, double (&acc)[NP], int inst, uint32_t step) {
  using R = typename S::R;
  R Rh, U, V, dr = 0;
  if constexpr (S::DEV) {
    macro_dev(F, p.rho0_f, dr, Rh, U, V);
  } else {
    macro_f64(F, Rh, U, V);
  }
  double drag = 0.0, lift = 0.0;
  if constexpr (BND) {
    // serial/CellFunctions.c:483-493
    if (nd->bc[1] == 3) V = (R)0.0;
    if (nd->flags & BN_DRAG) {
      const double rr = (double)Rh;
      drag = rr / 3 * (20 - nd->x) / 5;
      lift = rr / 3 * (20 - nd->y) / 5;
    }
  }
  // momentum exchange on the links of boundary id force_id: the body receives
  // c_k (f*_k(x) + f_opp(k)(x, t+1)) per wall link (Ladd / Mei et al.)
  if constexpr (MODE == 1 && BND) {
    if (p.force_on && nd->boundary_id == p.force_id) {
      double fx = 0.0, fy = 0.0;
#pragma unroll
      for (int k = 1; k < 9; k++) {
        if (nd->bc[k] == 1) {
          const double mk = absval<S>(S::load(src + k * p.plane + idx, p), k, p);
          const double fo = absval<S>(F[OPP[k]], OPP[k], p);
          fx += (double)CX[k] * (mk + fo);
          fy += (double)CY[k] * (mk + fo);
        }
      }
      acc[4] += fx;
      acc[5] += fy;
    }
  }
  // residual terms of step t
  if constexpr (MODE == 1 || BND) {
    const bool need_full = (MODE == 1 && p.res_kind == RK_FULL);
    const bool need_cref = BND && (nd->flags & BN_COMPAT_RES) && o.cres != nullptr;
    if (need_full || need_cref) {
      double d2 = 0.0;
      double* cr = need_cref ? o.cres + (size_t)nd->lin * 12 : nullptr;
#pragma unroll
      for (int k = 0; k < 9; k++) {
        const R mk = S::load(src + k * p.plane + idx, p);
        double d;
        if constexpr (S::DEV)
          d = (double)(F[k] - mk);
        else
          d = F[k] - mk;
        d2 = d2 + d * d;
        if (cr) cr[k] = d;
      }
      if (cr) {
        cr[9] = drag;
        cr[10] = lift;
        cr[11] = 1.0;
      }
      if (need_full) {
        acc[0] += d2;
        acc[2] += drag;
        acc[3] += lift;
      }
    }
    if (MODE == 1 && p.res_kind == RK_VEL) {
      const double rr = (double)Rh, uu = (double)U, vv = (double)V;
      const double r0 = o.snap[idx], u0 = o.snap[p.plane + idx], v0 = o.snap[2 * p.plane + idx];
      acc[0] += (uu - u0) * (uu - u0) + (vv - v0) * (vv - v0);
      acc[1] += uu * uu + vv * vv;
      acc[2] += (rr - r0) * (rr - r0);
      acc[3] += rr * rr;
      o.snap[idx] = rr;
      o.snap[p.plane + idx] = uu;
      o.snap[2 * p.plane + idx] = vv;
      if (p.ninst) {  // per-cell terms for the per-instance sums
        o.snap[3 * (size_t)p.plane + idx] = (uu - u0) * (uu - u0) + (vv - v0) * (vv - v0);
        o.snap[4 * (size_t)p.plane + idx] = uu * uu + vv * vv;
      }
    }
  }
  if constexpr (MODE == 2) {
    o.rho[idx] = (double)Rh;
    o.u[idx] = (double)U;
    o.v[idx] = (double)V;
    if (o.f) {
#pragma unroll
      for (int k = 0; k < 9; k++) o.f[(size_t)k * p.plane + idx] = absval<S>(F[k], k, p);
    }
    if (o.metaf) {
#pragma unroll
      for (int k = 0; k < 9; k++)
        o.metaf[(size_t)k * p.plane + idx] = absval<S>(S::load(src + k * p.plane + idx, p), k, p);
    }
  } else {
    R out[9];
    if constexpr (S::DEV)
      collide_dev<COLL>(F, dr, Rh, U, V, coll_par(p, inst), out);
    else
      collide_f64<COLL>(F, Rh, U, V, p, out);
    S::store_cell(dst, idx, out, p, step);
  }
}

// Pull of a bulk-tile cell.  Class 1: interior, plain pull.  Class 4: a
// boundary node whose only rule is straight-wall bounce-back (curved = 2, no
// inlet/outlet links, no flags, not a force body): StreamingStep with its
// StreamLattice mask (serial/Iterate.c:470-488, bits 0-8 of wmask), then
// WallBC (serial/BoundaryConditions.c:104-118, bits 16-24 = BC_ID[k] == 1) in
// the same order as the boundary-list path, so it runs in its tile with the
// interior cells instead of as a scattered boundary-list thread.
template <class S>
__device__ inline void pull_cell(typename S::R F[9], const typename S::T* src, const KParams& p,
                                 int idx, int c) {
  if (c == 4) {
    const uint32_t w = p.wmask[idx];
#pragma unroll
    for (int k = 0; k < 9; k++) {
      const int off = ((w >> k) & 1) ? -CX[k] - CY[k] * p.pitch : 0;
      F[k] = S::load(src + k * p.plane + idx + off, p);
    }
#pragma unroll
    for (int k = 0; k < 9; k++)
      if ((w >> (16 + k)) & 1) F[OPP[k]] = F[k];
  } else {
#pragma unroll
    for (int k = 0; k < 9; k++) F[k] = S::load(src + k * p.plane + idx - CX[k] - CY[k] * p.pitch, p);
  }
}

// ---------------------------------------------------------------- kernels
// Batched instances: the instance of a bulk node (per tile when the tile is
// all interior, else per cell) and whether it is still being advanced.
__device__ inline int bulk_instance(const KParams& p, unsigned b, int tf, int idx, bool& active) {
  if (!p.ninst) return -1;
  int inst;
  if (tf == 1) {
    inst = p.tile_inst[b];
  } else {
    inst = active ? (int)p.node_inst[idx] : -1;
    if (inst == 0xffff) inst = -1;
  }
  if (inst >= 0 && p.inst_active && !p.inst_active[inst]) active = false;
  return inst;
}

__device__ inline bool bnode_live(const KParams& p, const BNode& nd) {
  return !p.ninst || nd.inst < 0 || !p.inst_active || p.inst_active[nd.inst];
}

// tflag[t] for the tile of BS consecutive cells starting at bulk_begin + t*BS:
// 0 no interior node, 1 all interior, 2 mixed (read the per-cell class).
template <class S, int COLL, int MODE, int BS>
__global__ __launch_bounds__(BS) void k_bulk(const typename S::T* __restrict__ src,
                                             typename S::T* __restrict__ dst,
                                             const uint8_t* __restrict__ cls,
                                             const uint8_t* __restrict__ tflag, KParams p,
                                             KOut o) {
  using R = typename S::R;
  const uint32_t step = step_number<S>(p);
  if constexpr (MODE != 2) advance_seed<S>(p, step);
  unsigned b = blockIdx.x;
  if (p.flags & KF_XCD_REMAP) b = xcd_block(b, gridDim.x);
  else if (p.xcd_group > 0) b = xcd_group_block(b, (unsigned)p.xcd_group);
  if (p.tile_list) b = p.tile_list[b];
  const int tf = b == 0xffffffffu ? 0 : tflag[b];  // uniform per workgroup
  if constexpr (MODE != 1) {
    if (tf == 0) return;
  }
  const int idx = p.bulk_begin + (int)(b * BS + threadIdx.x);
  const int c = tf == 1 ? 1 : (tf == 2 && idx < p.bulk_end ? (int)cls[idx] : 0);
  bool active = c == 1 || c == 4;
  const int inst = bulk_instance(p, b, tf, idx, active);
  double acc[NP] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if constexpr (MODE != 1) {
    if (!active) return;
  }
  if (active) {
    R F[9];
    pull_cell<S>(F, src, p, idx, c);
    finish_node<S, COLL, MODE, false>(F, src, dst, p, o, idx, nullptr, acc, inst, step);
  }
  if constexpr (MODE == 1) block_sum<BS>(acc, o.part, blockIdx.x);
}

template <class S>
__device__ inline void pull_masked(typename S::R F[9], const BNode& nd,
                                   const typename S::T* src, const KParams& p) {
  // StreamingStep: serial/Iterate.c:470-488.  Directions that do not stream
  // keep the node's own METAF (UpdateF, Iterate.c:275-276).
#pragma unroll
  for (int k = 0; k < 9; k++) {
    const int off = ((nd.stream >> k) & 1) ? -CX[k] - CY[k] * p.pitch : 0;
    F[k] = S::load(src + k * p.plane + nd.idx + off, p);
  }
}

template <class S, int COLL, int MODE>
__global__ __launch_bounds__(256) void k_boundary(const typename S::T* __restrict__ src,
                                                  typename S::T* __restrict__ dst,
                                                  const BNode* __restrict__ bn, int nbn, KParams p,
                                                  KOut o) {
  using R = typename S::R;
  const uint32_t step = step_number<S>(p);
  if constexpr (MODE != 2) advance_seed<S>(p, step);
  const int i = blockIdx.x * 256 + threadIdx.x;
  double acc[NP] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (i < nbn && !(bn[i].flags & BN_OUTLET) && bnode_live(p, bn[i])) {
    const BNode& nd = bn[i];  // fields read where used
    R F[9];
    pull_masked<S>(F, nd, src, p);
    apply_bc<S>(F, nd, src, p, nd.idx);
    if (nd.flags & BN_FEEDER) {
      R* side = (R*)o.side + (size_t)nd.side * 9;
#pragma unroll
      for (int k = 0; k < 9; k++) side[k] = F[k];
    }
    finish_node<S, COLL, MODE, true>(F, src, dst, p, o, nd.idx, &nd, acc, nd.inst, step);
  }
  if constexpr (MODE == 1) block_sum<256>(acc, o.part, p.nb_bulk + blockIdx.x);
}

// One launch per step: workgroups [0, nb_bnd) take BS boundary-list nodes
// each (dispatched first, so their longer latency hides under the bulk
// tiles), workgroups [nb_bnd, nb_bnd + nb_bulk) are the bulk tiles.  The
// work of workgroup `bid` of `ngrid`; `tlist` the bulk tiles to run (null:
// all); `step` the stored step number.
template <class S, int COLL, int MODE, int BS>
__device__ inline void step_block(unsigned bid, unsigned ngrid,
                                  const typename S::T* __restrict__ src,
                                  typename S::T* __restrict__ dst,
                                  const uint8_t* __restrict__ cls,
                                  const uint8_t* __restrict__ tflag,
                                  const BNode* __restrict__ bn, int nbn,
                                  const uint32_t* __restrict__ tlist, const KParams& p,
                                  const KOut& o, uint32_t step) {
  using R = typename S::R;
  const int nbb = p.nb_bnd;
  double acc[NP] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if ((int)bid < nbb) {
    const int i = bid * BS + threadIdx.x;
    if (i < nbn && !(bn[i].flags & BN_OUTLET) && bnode_live(p, bn[i])) {
      const BNode& nd = bn[i];  // fields read where used
      R F[9];
      pull_masked<S>(F, nd, src, p);
      apply_bc<S>(F, nd, src, p, nd.idx);
      if (nd.flags & BN_FEEDER) {
        R* side = (R*)o.side + (size_t)nd.side * 9;
#pragma unroll
        for (int k = 0; k < 9; k++) side[k] = F[k];
      }
      finish_node<S, COLL, MODE, true>(F, src, dst, p, o, nd.idx, &nd, acc, nd.inst, step);
    }
  } else {
    unsigned b = bid - nbb;
    if (p.flags & KF_XCD_REMAP) b = xcd_block(b, ngrid - nbb);
    else if (p.xcd_group > 0) b = xcd_group_block(b, (unsigned)p.xcd_group);
    if (tlist) b = tlist[b];
    const int tf = b == 0xffffffffu ? 0 : tflag[b];
    if constexpr (MODE != 1) {
      if (tf == 0) return;
    }
    const int idx = p.bulk_begin + (int)(b * BS + threadIdx.x);
    const int c = tf == 1 ? 1 : (tf == 2 && idx < p.bulk_end ? (int)cls[idx] : 0);
    bool active = c == 1 || c == 4;
    const int inst = bulk_instance(p, b, tf, idx, active);
    if constexpr (MODE != 1) {
      if (!active) return;
    }
    if (active) {
      R F[9];
      pull_cell<S>(F, src, p, idx, c);
      finish_node<S, COLL, MODE, false>(F, src, dst, p, o, idx, nullptr, acc, inst, step);
    }
  }
  if constexpr (MODE == 1) block_sum<BS>(acc, o.part, bid);
}

template <class S, int COLL, int MODE, int BS>
__global__ __launch_bounds__(BS) void k_step(const typename S::T* __restrict__ src,
                                             typename S::T* __restrict__ dst,
                                             const uint8_t* __restrict__ cls,
                                             const uint8_t* __restrict__ tflag,
                                             const BNode* __restrict__ bn, int nbn, KParams p,
                                             KOut o) {
  const uint32_t step = step_number<S>(p);
  if constexpr (MODE != 2) advance_seed<S>(p, step);
  step_block<S, COLL, MODE, BS>(blockIdx.x, gridDim.x, src, dst, cls, tflag, bn, nbn,
                                p.tile_list, p, o, step);
}

// ---------------------------------------------------------------- two steps per launch
// Macros and collision of a bulk node (class 1 or 4) exactly as finish_node
// forms them for MODE 0 outside the boundary list; inst: the node's batched
// instance (-1: the solver's own parameters).
template <class S, int COLL>
__device__ inline void collide_bulk(const typename S::R F[9], const KParams& p,
                                    typename S::R out[9], int inst = -1) {
  using R = typename S::R;
  R Rh, U, V, dr = 0;
  if constexpr (S::DEV) {
    macro_dev(F, p.rho0_f, dr, Rh, U, V);
    collide_dev<COLL>(F, dr, Rh, U, V, coll_par(p, inst), out);
  } else {
    macro_f64(F, Rh, U, V);
    collide_f64<COLL>(F, Rh, U, V, p, out);
  }
}

// WallBC on the links a class-4 node's mask marks (bits 16-24), in the
// order of pull_cell.
template <class R>
__device__ inline void wall_bounce(R F[9], uint32_t w) {
#pragma unroll
  for (int k = 0; k < 9; k++)
    if ((w >> (16 + k)) & 1) F[OPP[k]] = F[k];
}

// Pull loads of one cell for step t+1, issued ahead of their use and
// without branches (so the compiler counts them exactly and the wave waits
// for each row's loads only when it forms that row): class 1 pulls all nine
// from the upstream neighbours, class 4 those its stream mask w marks (its
// own value otherwise), other classes load their own words (valid addresses,
// values unused).
template <class S>
__device__ inline void issue_pulls(typename S::R F[9], const typename S::T* src,
                                   const KParams& p, int idx, int c, uint32_t w) {
  const uint32_t m = c == 1 ? 0x1ffu : (c == 4 ? w : 0u);
#pragma unroll
  for (int k = 0; k < 9; k++) {
    const int off = ((m >> k) & 1) ? -CX[k] - CY[k] * p.pitch : 0;
    F[k] = S::load(plane_word(src, p.plane, k, idx + off), p);
  }
}

// Pulls of one row of a tile without in-tile walls, every lane read as class
// 1: lanes of other classes discard the values (their step is formed as
// zeros), and for rows 1 .. m every address stays inside the allocation (a
// halo lane's pull lands in the next or previous row or plane); rows outside
// load their own words.  The offsets are then the same in every lane and go
// into each plane's scalar base.
template <class S>
__device__ inline void issue_pulls_row(typename S::R F[9], const typename S::T* src,
                                       const KParams& p, int idx, bool inside) {
  using T = typename S::T;
#pragma unroll
  for (int k = 0; k < 9; k++) {
    const int off = inside ? -CX[k] - CY[k] * p.pitch : 0;
    const T* pk = src + (size_t)k * (size_t)p.plane + off;
    F[k] = S::load((const T*)((const char*)pk + (uint32_t)idx * (uint32_t)sizeof(T)), p);
  }
}

// Value of v in lane l-1 (lane_up1) / l+1 (lane_down1) of the wave, by a DPP
// wave shift (a VALU move, no LDS round trip); lane 0 / 63 gets 0.
__device__ inline float lane_up1(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x138, 0xf, 0xf, false));
}
__device__ inline float lane_down1(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x130, 0xf, 0xf, false));
}
__device__ inline double lane_up1(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), 0x138, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x138, 0xf, 0xf, false);
  return __builtin_bit_cast(double, (long long)(((unsigned long long)(unsigned)hi << 32) |
                                                (unsigned)lo));
}
__device__ inline double lane_down1(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), 0x130, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x130, 0xf, 0xf, false);
  return __builtin_bit_cast(double, (long long)(((unsigned long long)(unsigned)hi << 32) |
                                                (unsigned)lo));
}

// Value of v in lane j (j the same in every lane).
__device__ inline float lane_value(float v, int j) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), j));
}
__device__ inline double lane_value(double v, int j) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), j);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), j);
  return __builtin_bit_cast(double, (long long)(((unsigned long long)(unsigned)hi << 32) |
                                                (unsigned)lo));
}

// Step t+1 of one cell from its pulls (bulk classes 1 and 4; zeros for the
// others, which no advanced node reads), rounded as the store would.
template <class S, int COLL, bool WALLS>
__device__ inline void form_t1(typename S::R kn[9], typename S::R F[9], int c, uint32_t w,
                               const KParams& p, int inst = -1) {
  using R = typename S::R;
  if (c == 1 || c == 4) {
    if (WALLS && c == 4) wall_bounce(F, w);
    R o[9];
    collide_bulk<S, COLL>(F, p, o, inst);
#pragma unroll
    for (int k = 0; k < 9; k++) kn[k] = S::round_trip(o[k], p);
  } else {
#pragma unroll
    for (int k = 0; k < 9; k++) kn[k] = 0;
  }
}

// One tile of k_two_steps; WALLS: the tile holds class-4 nodes (their stream
// and wall masks are fetched with the class bytes); BATCH: a packed lattice
// of instances (each node's instance is fetched with its class byte: its
// relaxation rates, and nodes of stopped instances are not stored).
template <class S, int COLL, bool WALLS, bool BATCH>
__device__ inline void two_steps_tile(const typename S::T* __restrict__ src,
                                      typename S::T* __restrict__ dst,
                                      const uint8_t* __restrict__ cls, const int4 tl,
                                      const KParams& p) {
  using R = typename S::R;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int x = tl.x + 64 * wave + lane;
  const bool xin = x < p.pitch;
  const int xc = xin ? x : p.pitch - 1;  // loads of lanes past the row stay in it
  const int nit = tl.z + 2;              // step-t+1 rows y0-1 .. y0+rows
  const int r0 = tl.y - 1, rlast = r0 + nit - 1;
  auto cls_at = [&](int r) -> int { return xin ? (int)cls[(uint32_t)(r * p.pitch + xc)] : 0; };
  auto mask_at = [&](int r) -> uint32_t {
    return WALLS ? p.wmask[(uint32_t)(r * p.pitch + xc)] : 0u;
  };
  auto inst_at = [&](int r) -> int {
    if constexpr (BATCH) {
      const int v = xin ? (int)p.node_inst[(uint32_t)(r * p.pitch + xc)] : 0xffff;
      return v == 0xffff ? -1 : v;
    } else {
      return -1;
    }
  };
  // Edges between the waves: each row, lane 0 of every wave publishes the
  // t+1 values its left neighbour pulls (directions 3, 6, 7) and lane 63 the
  // ones its right neighbour pulls (1, 5, 8); two slots taking turns, one
  // barrier per row.
  __shared__ R edge[2][TWO_WAVES][2][3];
  // The workgroup's outer columns x0-1 (wave 0) and x0+256 (last wave): lane
  // j of a batch forms the cell of row base + j at t+1 (zeros outside the
  // grid or past row rlast); a batch serves the step-t+2 rows base+1 ..
  // base+62.  Its class bytes are fetched two rows before its use and its
  // pulls one row before, ahead of that row's own loads, so forming it waits
  // on nothing else in flight.
  const bool side_owner = wave == 0 || wave == TWO_WAVES - 1;
  const int xs = wave == 0 ? tl.x - 1 : tl.x + TWO_COLS;
  R sd[9];
  R sq[9];
  int sc = 0, si = -1;
  uint32_t sw = 0;
  int sidx = 0;
  auto side_classes = [&](int base) {
    const int row = base + lane;
    const bool ok = xs >= 0 && xs < p.pitch && row <= rlast;
    sidx = (ok ? row : r0) * p.pitch + (ok ? xs : xc);
    sc = ok ? (int)cls[(uint32_t)sidx] & 0x7f : 0;
    sw = (WALLS && ok) ? p.wmask[(uint32_t)sidx] : 0u;
    if constexpr (BATCH) {
      const int v = ok ? (int)p.node_inst[(uint32_t)sidx] : 0xffff;
      si = v == 0xffff ? -1 : v;
    }
  };
  int sbase = r0;
  R kd[3] = {0, 0, 0};                     // row rr-1 at t+1: directions 2, 5, 6
  R kc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};   // row rr at t+1
  int cc = 0;                              // class byte of (x, rr)
  uint32_t wc = 0;                         // masks of (x, rr)
  int ic = -1;                             // instance of (x, rr)
  // neighbour columns' t+1 values: left (1, 5, 8) and right (3, 6, 7) of
  // rows rr (1 / 3), rr-1 (5 / 6) and rr+1 (8 / 7)
  R el1 = 0, el5 = 0, el5m = 0, er3 = 0, er6 = 0, er6m = 0;
  // Software pipeline: while row r is formed, the pulls of rows r+1 and r+2
  // (two buffers taking turns) and the class bytes (and masks) of rows up
  // to r+3 are in flight.
  int c0 = cls_at(r0), c1 = cls_at(min(r0 + 1, rlast)), c2 = cls_at(min(r0 + 2, rlast));
  uint32_t w0 = mask_at(r0), w1 = mask_at(min(r0 + 1, rlast)), w2 = mask_at(min(r0 + 2, rlast));
  int i0 = inst_at(r0), i1 = inst_at(min(r0 + 1, rlast)), i2 = inst_at(min(r0 + 2, rlast));
  auto row_pulls = [&](R (&F)[9], int r, int c, uint32_t w) {
    if constexpr (WALLS || !S::DEV)  // FP64: measured slower (569 -> 582 us per launch)
      issue_pulls<S>(F, src, p, r * p.pitch + xc, c & 0x7f, w);
    else
      issue_pulls_row<S>(F, src, p, r * p.pitch + xc, r >= 1 && r <= p.m);
  };
  R FA[9], FB[9];
  row_pulls(FA, r0, c0, w0);
  row_pulls(FB, min(r0 + 1, rlast), c1, w1);
  // Row r = r0 + it from the buffer Fb that holds its pulls; Fb is refilled
  // with the pulls of row r+2.
  auto row = [&](int it, R (&Fb)[9]) {
    const int r = r0 + it;  // row of step t+1
    const int cb = c0;
    const uint32_t w = w0;
    const int ib = i0;
    if (side_owner) {
      // batch b serves step-t+2 rows from it = 2 + 62 b: classes at
      // it = 62 b, pulls at it = 1 + 62 b
      if (it % 62 == 0) side_classes(r);
      if (it >= 1 && (it - 1) % 62 == 0) issue_pulls<S>(sq, src, p, sidx, sc, sw);
    }
    c0 = c1;
    w0 = w1;
    i0 = i1;
    c1 = c2;
    w1 = w2;
    i1 = i2;
    const int rn2 = min(r + 2, rlast), rn3 = min(r + 3, rlast);
    c2 = cls_at(rn3);
    w2 = mask_at(rn3);
    i2 = inst_at(rn3);
    R kn[9];  // row r at t+1
    form_t1<S, COLL, WALLS>(kn, Fb, cb & 0x7f, w, p, ib);
    row_pulls(Fb, rn2, c1, w1);
    // publish the edge lanes' row-r values, then take the neighbours'
    const int slot = it & 1;
    if (lane == 0) {
      edge[slot][wave][0][0] = kn[3];
      edge[slot][wave][0][1] = kn[6];
      edge[slot][wave][0][2] = kn[7];
    }
    if (lane == 63) {
      edge[slot][wave][1][0] = kn[1];
      edge[slot][wave][1][1] = kn[5];
      edge[slot][wave][1][2] = kn[8];
    }
    __syncthreads();
    const int wl = wave > 0 ? wave - 1 : 0, wr = wave < TWO_WAVES - 1 ? wave + 1 : wave;
    const R nl1 = edge[slot][wl][1][0], nl5 = edge[slot][wl][1][1], nl8 = edge[slot][wl][1][2];
    const R nr3 = edge[slot][wr][0][0], nr6 = edge[slot][wr][0][1], nr7 = edge[slot][wr][0][2];
    if (it >= 2) {
      const int rr = r - 1;  // row of step t+2
      if (side_owner && (it - 2) % 62 == 0) {
        sbase = rr - 1;
        form_t1<S, COLL, WALLS>(sd, sq, sc, sw, p, si);
      }
      // F[k](x) = f_k(x - c_k) at t+1; lanes 0 and 63 take the neighbour
      // columns: the next wave's edge lanes, or the batch at the outer edges
      R nb[9];
      nb[0] = kc[0];
      nb[1] = lane_up1(kc[1]);
      nb[2] = kd[0];
      nb[3] = lane_down1(kc[3]);
      nb[4] = kn[4];
      nb[5] = lane_up1(kd[1]);
      nb[6] = lane_down1(kd[2]);
      nb[7] = lane_down1(kn[7]);
      nb[8] = lane_up1(kn[8]);
      R l1 = el1, l5 = el5m, l8 = nl8, r3 = er3, r6 = er6m, r7 = nr7;
      if (side_owner) {
        const int j = rr - sbase;  // batch lane of row rr
        const R s1 = lane_value(sd[1], j), s5 = lane_value(sd[5], j - 1);
        const R s8 = lane_value(sd[8], j + 1), s3 = lane_value(sd[3], j);
        const R s6 = lane_value(sd[6], j - 1), s7 = lane_value(sd[7], j + 1);
        if (wave == 0) {
          l1 = s1;
          l5 = s5;
          l8 = s8;
        }
        if (wave == TWO_WAVES - 1) {
          r3 = s3;
          r6 = s6;
          r7 = s7;
        }
      }
      if (lane == 0) {
        nb[1] = l1;
        nb[5] = l5;
        nb[8] = l8;
      }
      if (lane == 63) {
        nb[3] = r3;
        nb[6] = r6;
        nb[7] = r7;
      }
      bool store = xin && (cc & 0x80);
      if constexpr (BATCH) {
        if (store && ic >= 0 && p.inst_active && !p.inst_active[ic]) store = false;
      }
      if (store) {
        R G[9];
        if (WALLS && (cc & 0x7f) == 4) {
          // StreamingStep mask, then WallBC, as pull_cell
#pragma unroll
          for (int k = 0; k < 9; k++) G[k] = ((wc >> k) & 1) ? nb[k] : kc[k];
          wall_bounce(G, wc);
        } else {
#pragma unroll
          for (int k = 0; k < 9; k++) G[k] = nb[k];
        }
        R o[9];
        collide_bulk<S, COLL>(G, p, o, ic);
        S::store_cell(dst, rr * p.pitch + x, o, p, 0u);
      }
    }
    el5m = el5;
    er6m = er6;
    el1 = nl1;
    el5 = nl5;
    er3 = nr3;
    er6 = nr6;
    kd[0] = kc[2];
    kd[1] = kc[5];
    kd[2] = kc[6];
#pragma unroll
    for (int k = 0; k < 9; k++) kc[k] = kn[k];
    cc = cb;
    wc = w;
    ic = ib;
  };
  // one copy of the two-row body: unrolled copies only grow the code
#pragma unroll 1
  for (int it = 0; it < nit; it += 2) {
    row(it, FA);
    if (it + 1 < nit) row(it + 1, FB);
  }
}

// Two plain steps of a rectangle of the lattice in one pass over the state:
// METAF_t is read once and METAF_{t+2} written once, where two single steps
// read and write the whole state twice.  One workgroup of four waves per
// tile (x0, y0, rows, walls): lane l of wave w works on padded column
// x0 + 64 w + l (x0 a multiple of 256, so every row of a tile is whole cache
// lines: stores that share a line with another tile's lose half the
// bandwidth); the outputs are rows y0 .. y0 + rows - 1.  The waves march up
// the rows together.  For row r each forms step t+1 of its 64 columns (pull
// from METAF_t, collide, round as the store would) and keeps the values in
// registers; with rows r-2, r-1, r in hand it forms step t+2 of row r-1,
// pulling the vertical directions from its own kept rows, the diagonal and
// horizontal ones from lanes l-1 / l+1 by lane shuffles and, at its edge
// lanes, from the neighbouring waves through LDS, and stores it.  The two
// outer columns x0-1 and x0+256 are formed at t+1 in batches of 64 rows and
// read by readlane; rows y0-1 and y0+rows are formed at t+1 and not stored.
// `walls` is set when any column of the tile or its outer columns holds a
// class-4 node.
// cls carries bit 7 on the nodes the tiles advance: interior (class 1) or
// in-tile wall (class 4) nodes whose fluid neighbours are all of class 1 or
// 4 too, so that every t+1 value they pull is formed here without a
// boundary-list record.  Lanes of other classes form nothing (zeros that no
// advanced node reads).  The remaining nodes take two single steps through a
// third buffer in separate launches.  The arithmetic per cell is k_step's
// (pull_cell, collide_bulk, the store's rounding), so the result equals two
// single steps bit for bit.
// Workgroups [0, nband) run the band's first single step instead (k_step's
// work over the band and its ring, grid bgrid <= nband, into bdst): it reads
// only METAF_t and writes only the third buffer, so it shares the launch;
// nband is a multiple of 8, which keeps the tiles' XCD order.
template <class S, int COLL>
__global__ __launch_bounds__(TWO_COLS) void k_two_steps(
    const typename S::T* __restrict__ src, typename S::T* __restrict__ dst,
    const uint8_t* __restrict__ cls, const int4* __restrict__ tiles, KParams p, int nband,
    int bgrid, typename S::T* __restrict__ bdst, const uint8_t* __restrict__ bcls,
    const uint8_t* __restrict__ btflag, const BNode* __restrict__ bn, int nbn,
    const uint32_t* __restrict__ blist, KOut o) {
  if ((int)blockIdx.x < nband) {
    if ((int)blockIdx.x < bgrid)
      step_block<S, COLL, 0, TWO_COLS>(blockIdx.x, bgrid, src, bdst, bcls, btflag, bn, nbn, blist,
                                       p, o, 0u);
    return;
  }
  const int4 tl = tiles[blockIdx.x - nband];  // x0, y0, rows, walls
  if (tl.z <= 0) return;
  if constexpr (S::DEV) {
    if (p.ninst) {
      if (tl.w)
        two_steps_tile<S, COLL, true, true>(src, dst, cls, tl, p);
      else
        two_steps_tile<S, COLL, false, true>(src, dst, cls, tl, p);
      return;
    }
  }
  if (tl.w)
    two_steps_tile<S, COLL, true, false>(src, dst, cls, tl, p);
  else
    two_steps_tile<S, COLL, false, false>(src, dst, cls, tl, p);
}

// ---------------------------------------------------------------- three steps per launch
// The nine pulls of a cell at row rr from its column's kept rows (prev = row
// rr-1, directions 2, 5, 6 only; cur = row rr; next = row rr+1) and lanes
// l-1 / l+1; lanes 0 and 63 are completed by the caller.  (LDS permutes
// here: DPP wave shifts, which take VALU issue slots, measured 346 -> 352 us
// per launch in this VALU-heavy kernel.)
template <class R>
__device__ inline void gather_rows(R nb[9], const R prev[3], const R cur[9], const R next[9]) {
  nb[0] = cur[0];
  nb[1] = __shfl_up(cur[1], 1, 64);
  nb[2] = prev[0];
  nb[3] = __shfl_down(cur[3], 1, 64);
  nb[4] = next[4];
  nb[5] = __shfl_up(prev[1], 1, 64);
  nb[6] = __shfl_down(prev[2], 1, 64);
  nb[7] = __shfl_down(next[7], 1, 64);
  nb[8] = __shfl_up(next[8], 1, 64);
}

// The next step of a bulk cell (class 1 or 4) from its pulls nb and its own
// values own (what a class-4 cell keeps where its stream mask is clear), as
// pull_cell + finish_node form it, rounded as the store would; zeros for the
// other classes (no advanced node reads them).
template <class S, int COLL, bool WALLS>
__device__ inline void form_next(typename S::R out[9], const typename S::R nb[9],
                                 const typename S::R own[9], int c, uint32_t w,
                                 const KParams& p) {
  using R = typename S::R;
  if (c == 1 || c == 4) {
    R G[9];
    if (WALLS && c == 4) {
#pragma unroll
      for (int k = 0; k < 9; k++) G[k] = ((w >> k) & 1) ? nb[k] : own[k];
      wall_bounce(G, w);
    } else {
#pragma unroll
      for (int k = 0; k < 9; k++) G[k] = nb[k];
    }
    R o[9];
    collide_bulk<S, COLL>(G, p, o);
#pragma unroll
    for (int k = 0; k < 9; k++) out[k] = S::round_trip(o[k], p);
  } else {
#pragma unroll
    for (int k = 0; k < 9; k++) out[k] = 0;
  }
}

// Step-t+3 rows one side batch serves (64 lanes: a batch's first and last two
// rows only feed its own step-t+2 cells).
constexpr int THREE_SIDE_ROWS = 60;

// One tile of k_three_steps: k_two_steps' march one stage deeper.  For row r
// each wave forms step t+1 of row r (pulls of METAF_t), step t+2 of row r-1
// (from its kept t+1 rows r-2 .. r) and step t+3 of row r-2 (from its kept
// t+2 rows r-3 .. r-1), which it stores.  Edge lanes take the neighbouring
// waves' t+1 and t+2 values through LDS (one barrier per stage and row).  The
// outer waves form the tile's outer columns x0-1 and x0+256 at t+1 and t+2 in
// batches of 64 rows (lane j = row sbase + j): t+1 of the columns xs-1, xs,
// xs+1 (classes fetched one row before the pulls, the columns formed a row
// after them), then t+2 of xs from those, all kept in LDS and read there by
// the edge lane.  Rows -1 and m+2, which a tile at the lattice's edge
// reaches, hold nothing (class 0; their loads read rows 0 / m+1).
// LDS of a three-step workgroup: wave edges [slot][wave][0: lane 0 (3, 6, 7),
// 1: lane 63 (1, 5, 8)] at t+1 and t+2; outer columns [side][batch
// parity][slot][batch lane]: slots 0-2 t+1 of xs-1 (directions 1, 5, 8), 3-11
// t+1 of xs (0-8), 12-14 t+1 of xs+1 (3, 6, 7), 15-17 t+2 of xs (1, 5, 8 left;
// 3, 6, 7 right).
template <class R>
struct ThreeLds {
  R edge1[2][TWO_WAVES][2][3];
  R edge2[2][TWO_WAVES][2][3];
  R side[2][2][18][64];
};

template <class S, int COLL, bool WALLS>
__device__ inline void three_steps_tile(const typename S::T* __restrict__ src,
                                        typename S::T* __restrict__ dst,
                                        const uint8_t* __restrict__ cls, const int4 tl,
                                        const KParams& p, ThreeLds<typename S::R>& lds) {
  using R = typename S::R;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int x = tl.x + 64 * wave + lane;
  const bool xin = x < p.pitch;
  const int xc = xin ? x : p.pitch - 1;
  const int nit = tl.z + 4;  // step-t+1 rows y0-2 .. y0+rows+1
  const int r0 = tl.y - 2, rlast = r0 + nit - 1;
  const int rtop = p.m + 1;
  auto rowc = [&](int r) { return min(max(r, 0), rtop); };
  auto cls_at = [&](int r) -> int {
    return (xin && r >= 0 && r <= rtop) ? (int)cls[(uint32_t)(r * p.pitch + xc)] : 0;
  };
  auto mask_at = [&](int r) -> uint32_t {
    return WALLS ? p.wmask[(uint32_t)(rowc(r) * p.pitch + xc)] : 0u;
  };
  auto& edge1 = lds.edge1;
  auto& edge2 = lds.edge2;
  auto& side = lds.side;
  const bool side_owner = wave == 0 || wave == TWO_WAVES - 1;
  const int sn = wave == 0 ? 0 : 1;
  const int xs = wave == 0 ? tl.x - 1 : tl.x + TWO_COLS;
  auto clj = [](int j) { return min(max(j, 0), 63); };
  // side batch state: the three columns' classes, masks and indices, then
  // their pulls, all issued together (the columns share their cache lines,
  // which other tiles' batches would evict between separate passes)
  R sq[3][9];
  int sc[3] = {0, 0, 0}, sidx[3] = {0, 0, 0};
  uint32_t sw[3] = {0, 0, 0};
  auto side_classes = [&](int base) {
    const int row = base + lane;
#pragma unroll
    for (int col = 0; col < 3; col++) {
      const int xx = xs - 1 + col;
      const bool ok = xx >= 0 && xx < p.pitch && row >= 0 && row <= rtop;
      sidx[col] = rowc(row) * p.pitch + min(max(xx, 0), p.pitch - 1);
      sc[col] = ok ? (int)cls[(uint32_t)sidx[col]] & 0x7f : 0;
      sw[col] = (WALLS && ok) ? p.wmask[(uint32_t)sidx[col]] : 0u;
    }
  };
  auto side_pulls = [&]() {
#pragma unroll
    for (int col = 0; col < 3; col++) issue_pulls<S>(sq[col], src, p, sidx[col], sc[col], sw[col]);
  };
  auto side_form = [&](int buf) {
    R* s = &side[sn][buf][0][0];
    R v[9];
    form_t1<S, COLL, WALLS>(v, sq[0], sc[0], sw[0], p);
    s[0 * 64 + lane] = v[1];
    s[1 * 64 + lane] = v[5];
    s[2 * 64 + lane] = v[8];
    form_t1<S, COLL, WALLS>(v, sq[1], sc[1], sw[1], p);
#pragma unroll
    for (int k = 0; k < 9; k++) s[(3 + k) * 64 + lane] = v[k];
    form_t1<S, COLL, WALLS>(v, sq[2], sc[2], sw[2], p);
    s[12 * 64 + lane] = v[3];
    s[13 * 64 + lane] = v[6];
    s[14 * 64 + lane] = v[7];
  };
  auto side_t2 = [&](int buf) {
    R* s = &side[sn][buf][0][0];
    const int jm = clj(lane - 1), jp = clj(lane + 1);
    R own[9], nb[9];
#pragma unroll
    for (int k = 0; k < 9; k++) own[k] = s[(3 + k) * 64 + lane];
    nb[0] = own[0];
    nb[1] = s[0 * 64 + lane];
    nb[5] = s[1 * 64 + jm];
    nb[8] = s[2 * 64 + jp];
    nb[2] = s[(3 + 2) * 64 + jm];
    nb[4] = s[(3 + 4) * 64 + jp];
    nb[3] = s[12 * 64 + lane];
    nb[6] = s[13 * 64 + jm];
    nb[7] = s[14 * 64 + jp];
    R v[9];
    form_next<S, COLL, WALLS>(v, nb, own, sc[1], sw[1], p);
    if (sn == 0) {
      s[15 * 64 + lane] = v[1];
      s[16 * 64 + lane] = v[5];
      s[17 * 64 + lane] = v[8];
    } else {
      s[15 * 64 + lane] = v[3];
      s[16 * 64 + lane] = v[6];
      s[17 * 64 + lane] = v[7];
    }
  };
  // batch 0 (rows r0 .. r0+63), before the march
  if (side_owner) {
    side_classes(r0);
    side_pulls();
    side_form(0);
  }
  __syncthreads();
  if (side_owner) side_t2(0);
  __syncthreads();
  // kept rows in two-slot rings by row parity (the march body is unrolled
  // by two, so every slot index is a constant and no row is copied): K1[P]
  // t+1 of the rows r-2 (then r), K1[1-P] of row r-1; K2[P] t+2 of row r-3
  // (then r-1), K2[1-P] of row r-2; classes and masks of rows r-1 (cc1) and
  // r-2 (cc2)
  R K1[2][9] = {{0, 0, 0, 0, 0, 0, 0, 0, 0}, {0, 0, 0, 0, 0, 0, 0, 0, 0}};
  R K2[2][9] = {{0, 0, 0, 0, 0, 0, 0, 0, 0}, {0, 0, 0, 0, 0, 0, 0, 0, 0}};
  int cc1 = 0, cc2 = 0;
  uint32_t wc1 = 0, wc2 = 0;
  // neighbour waves' edge values (as k_two_steps' el*/er*), t+1 and t+2
  R el1 = 0, el5 = 0, el5m = 0, er3 = 0, er6 = 0, er6m = 0;
  R fl1 = 0, fl5 = 0, fl5m = 0, fr3 = 0, fr6 = 0, fr6m = 0;
  int c0 = cls_at(r0), c1 = cls_at(min(r0 + 1, rlast)), c2 = cls_at(min(r0 + 2, rlast));
  uint32_t w0 = mask_at(r0), w1 = mask_at(min(r0 + 1, rlast)), w2 = mask_at(min(r0 + 2, rlast));
  auto row_pulls = [&](R (&F)[9], int r, int c, uint32_t w) {
    if constexpr (WALLS || !S::DEV)  // FP64: measured slower (569 -> 582 us per two-step launch)
      issue_pulls<S>(F, src, p, rowc(r) * p.pitch + xc, c & 0x7f, w);
    else
      issue_pulls_row<S>(F, src, p, rowc(r) * p.pitch + xc, r >= 1 && r <= p.m);
  };
  R FA[9], FB[9];
  row_pulls(FA, r0, c0, w0);
  row_pulls(FB, min(r0 + 1, rlast), c1, w1);
  auto row = [&](int it, R (&Fb)[9], auto parity) {
    constexpr int P = decltype(parity)::value;  // it & 1
    const int r = r0 + it;  // row of step t+1
    const int cb = c0;
    const uint32_t w = w0;
    // side batch b >= 1 (rows r0 + 60 b ..): classes at iteration 60 b,
    // pulls at 60 b + 1, t+1 formed at 60 b + 2 and t+2 after its first
    // barrier; in use from 60 b + 4
    const int b = it / THREE_SIDE_ROWS, k = it - b * THREE_SIDE_ROWS;
    const bool forming = side_owner && b >= 1 && k <= 2 && b * THREE_SIDE_ROWS + 4 < nit;
    if (forming) {
      if (k == 0) side_classes(r0 + b * THREE_SIDE_ROWS);
      if (k == 1) side_pulls();
      if (k == 2) side_form(b & 1);
    }
    // the batch in use: 0 up to iteration 63, then b from iteration 60 b + 4
    const int bu = it < 64 ? 0 : (it - 4) / THREE_SIDE_ROWS;
    const int ubuf = bu & 1, ubase = r0 + bu * THREE_SIDE_ROWS;
    c0 = c1;
    w0 = w1;
    c1 = c2;
    w1 = w2;
    const int rn2 = min(r + 2, rlast), rn3 = min(r + 3, rlast);
    c2 = cls_at(rn3);
    w2 = mask_at(rn3);
    // row r-2 at t+1 (directions 2, 5, 6) before its slot takes row r
    const R kd1[3] = {K1[P][2], K1[P][5], K1[P][6]};
    R(&kn1)[9] = K1[P];  // row r at t+1
    R(&kc1)[9] = K1[1 - P];  // row r-1
    form_t1<S, COLL, WALLS>(kn1, Fb, cb & 0x7f, w, p);
    row_pulls(Fb, rn2, c1, w1);
    const int slot = it & 1;
    if (lane == 0) {
      edge1[slot][wave][0][0] = kn1[3];
      edge1[slot][wave][0][1] = kn1[6];
      edge1[slot][wave][0][2] = kn1[7];
    }
    if (lane == 63) {
      edge1[slot][wave][1][0] = kn1[1];
      edge1[slot][wave][1][1] = kn1[5];
      edge1[slot][wave][1][2] = kn1[8];
    }
    __syncthreads();
    if (forming && k == 2) side_t2(b & 1);
    const int wl = wave > 0 ? wave - 1 : 0, wr = wave < TWO_WAVES - 1 ? wave + 1 : wave;
    const R nl1 = edge1[slot][wl][1][0], nl5 = edge1[slot][wl][1][1], nl8 = edge1[slot][wl][1][2];
    const R nr3 = edge1[slot][wr][0][0], nr6 = edge1[slot][wr][0][1], nr7 = edge1[slot][wr][0][2];
    if (it >= 2) {
      const int rr = r - 1;  // row of step t+2
      R nb[9];
      gather_rows(nb, kd1, kc1, kn1);
      R l1 = el1, l5 = el5m, l8 = nl8, r3 = er3, r6 = er6m, r7 = nr7;
      if (side_owner) {
        const int j = rr - ubase;
        const R* sv = &side[sn][ubuf][3][0];  // t+1 of xs, direction d at sv[d * 64 + row]
        if (wave == 0) {
          l1 = sv[1 * 64 + clj(j)];
          l5 = sv[5 * 64 + clj(j - 1)];
          l8 = sv[8 * 64 + clj(j + 1)];
        }
        if (wave == TWO_WAVES - 1) {
          r3 = sv[3 * 64 + clj(j)];
          r6 = sv[6 * 64 + clj(j - 1)];
          r7 = sv[7 * 64 + clj(j + 1)];
        }
      }
      if (lane == 0) {
        nb[1] = l1;
        nb[5] = l5;
        nb[8] = l8;
      }
      if (lane == 63) {
        nb[3] = r3;
        nb[6] = r6;
        nb[7] = r7;
      }
      const R kd2[3] = {K2[P][2], K2[P][5], K2[P][6]};  // row r-3 at t+2
      R(&kn2)[9] = K2[P];  // row rr at t+2
      R(&kc2)[9] = K2[1 - P];  // row r-2
      form_next<S, COLL, WALLS>(kn2, nb, kc1, cc1 & 0x7f, wc1, p);
      if (lane == 0) {
        edge2[slot][wave][0][0] = kn2[3];
        edge2[slot][wave][0][1] = kn2[6];
        edge2[slot][wave][0][2] = kn2[7];
      }
      if (lane == 63) {
        edge2[slot][wave][1][0] = kn2[1];
        edge2[slot][wave][1][1] = kn2[5];
        edge2[slot][wave][1][2] = kn2[8];
      }
      __syncthreads();
      const R ml1 = edge2[slot][wl][1][0], ml5 = edge2[slot][wl][1][1], ml8 = edge2[slot][wl][1][2];
      const R mr3 = edge2[slot][wr][0][0], mr6 = edge2[slot][wr][0][1], mr7 = edge2[slot][wr][0][2];
      if (it >= 4) {
        const int rs = r - 2;  // row of step t+3
        R nc[9];
        gather_rows(nc, kd2, kc2, kn2);
        R m1 = fl1, m5 = fl5m, m8 = ml8, m3 = fr3, m6 = fr6m, m7 = mr7;
        if (side_owner) {
          const int j = rs - ubase;
          const R* sv = &side[sn][ubuf][15][0];  // t+2 of xs: three directions
          if (wave == 0) {
            m1 = sv[0 * 64 + clj(j)];
            m5 = sv[1 * 64 + clj(j - 1)];
            m8 = sv[2 * 64 + clj(j + 1)];
          }
          if (wave == TWO_WAVES - 1) {
            m3 = sv[0 * 64 + clj(j)];
            m6 = sv[1 * 64 + clj(j - 1)];
            m7 = sv[2 * 64 + clj(j + 1)];
          }
        }
        if (lane == 0) {
          nc[1] = m1;
          nc[5] = m5;
          nc[8] = m8;
        }
        if (lane == 63) {
          nc[3] = m3;
          nc[6] = m6;
          nc[7] = m7;
        }
        if (xin && (cc2 & 0x80)) {
          R G[9];
          if (WALLS && (cc2 & 0x7f) == 4) {
#pragma unroll
            for (int q = 0; q < 9; q++) G[q] = ((wc2 >> q) & 1) ? nc[q] : kc2[q];
            wall_bounce(G, wc2);
          } else {
#pragma unroll
            for (int q = 0; q < 9; q++) G[q] = nc[q];
          }
          R o[9];
          collide_bulk<S, COLL>(G, p, o);
          S::store_cell(dst, rs * p.pitch + x, o, p, 0u);
        }
      }
      fl5m = fl5;
      fr6m = fr6;
      fl1 = ml1;
      fl5 = ml5;
      fr3 = mr3;
      fr6 = mr6;
      cc2 = cc1;
      wc2 = wc1;
    }
    el5m = el5;
    er6m = er6;
    el1 = nl1;
    el5 = nl5;
    er3 = nr3;
    er6 = nr6;
    cc1 = cb;
    wc1 = w;
  };
#pragma unroll 1
  for (int it = 0; it < nit; it += 2) {
    row(it, FA, std::integral_constant<int, 0>{});
    if (it + 1 < nit) row(it + 1, FB, std::integral_constant<int, 1>{});
  }
}

// Three plain steps of a rectangle of the lattice in one pass over the state
// (METAF_t read once, METAF_{t+3} written once).  Tiles as k_two_steps'; cls
// carries bit 7 on the nodes these tiles advance: class 1 or 4 with every
// fluid node within two nodes (rows and columns) of class 1 or 4, so every
// t+1 and t+2 value they pull is formed here.  The other nodes take three
// single steps through two more buffers; workgroups [0, nband) run the first
// of them (k_step's work over the band and its two rings into bdst), as in
// k_two_steps.  Not for packed lattices of instances.
template <class S, int COLL>
__global__ __launch_bounds__(TWO_COLS) void k_three_steps(
    const typename S::T* __restrict__ src, typename S::T* __restrict__ dst,
    const uint8_t* __restrict__ cls, const int4* __restrict__ tiles, KParams p, int nband,
    int bgrid, typename S::T* __restrict__ bdst, const uint8_t* __restrict__ bcls,
    const uint8_t* __restrict__ btflag, const BNode* __restrict__ bn, int nbn,
    const uint32_t* __restrict__ blist, KOut o) {
  if ((int)blockIdx.x < nband) {
    if ((int)blockIdx.x < bgrid)
      step_block<S, COLL, 0, TWO_COLS>(blockIdx.x, bgrid, src, bdst, bcls, btflag, bn, nbn, blist,
                                       p, o, 0u);
    return;
  }
  __shared__ ThreeLds<typename S::R> lds;
  const int4 tl = tiles[blockIdx.x - nband];  // x0, y0, rows, walls
  if (tl.z <= 0) return;
  if (tl.w)
    three_steps_tile<S, COLL, true>(src, dst, cls, tl, p, lds);
  else
    three_steps_tile<S, COLL, false>(src, dst, cls, tl, p, lds);
}

// Plain steps with FP16 storage: each bulk workgroup of 128 threads takes one
// tile of 256 cells and each thread two neighbouring cells, so the loads and
// stores of all-interior tiles move one 32-bit word per two cells: twice the
// bytes per memory instruction, and as many workgroups as k_step launches
// (MLUPS against k_step: 1025^2 65.6 k / 58.8 k, 2049^2 116.7 k / 111.8 k,
// 4097^2 147.6 k / 144.9 k; pairs of tiles in workgroups of 256 measured
// 65.3 k at 1025^2 and lost at 2049^2 and 4097^2; the same pairing with
// 64-bit words made FP32 7 % slower, so FP32 keeps one cell per thread).
// A pull with cx = +-1 starts on an odd cell; its two halves come from the
// two aligned words around it.  Same arithmetic per cell as k_step
// (macro_dev and collide_dev on the same floats), so results are identical.
// Boundary workgroups (256 list nodes each) and mixed tiles run k_step's
// per-cell code, two cells per thread 128 apart.  Launch: nb_bnd + nb_bulk
// workgroups of 128; XCD grouping by runs of G tiles as k_step.
template <class S>
struct Pair;
template <>
struct Pair<StF16> {
  using W = uint32_t;
  __device__ static float lo(W w, const KParams& p) {
    return (float)__builtin_bit_cast(_Float16, (uint16_t)(w & 0xffffu)) * p.inv_scale_f;
  }
  __device__ static float hi(W w, const KParams& p) {
    return (float)__builtin_bit_cast(_Float16, (uint16_t)(w >> 16)) * p.inv_scale_f;
  }
  __device__ static W pack(float a, float b, uint32_t, const KParams& p) {
    const uint16_t l = __builtin_bit_cast(uint16_t, (_Float16)(a * p.scale_f));
    const uint16_t h = __builtin_bit_cast(uint16_t, (_Float16)(b * p.scale_f));
    return (uint32_t)l | ((uint32_t)h << 16);
  }
};
// Stochastic rounding: the pair's random word r of direction k (the same
// word StF16S::store_cell draws for each of the two cells alone).
template <>
struct Pair<StF16S> : Pair<StF16> {
  __device__ static W pack(float a, float b, uint32_t r, const KParams& p) {
    half2_t x = {(_Float16)0, (_Float16)0};
    x = f16_stochastic_into(x, a * p.scale_f, r, false);
    x = f16_stochastic_into(x, b * p.scale_f, rot16(r), true);
    return __builtin_bit_cast(W, x);
  }
};

template <class S, int COLL>
__global__ __launch_bounds__(128) void k_step_pair(const typename S::T* __restrict__ src,
                                                   typename S::T* __restrict__ dst,
                                                   const uint8_t* __restrict__ cls,
                                                   const uint8_t* __restrict__ tflag,
                                                   const BNode* __restrict__ bn, int nbn,
                                                   KParams p, KOut o) {
  using R = float;
  using PW = Pair<S>;
  using W = typename PW::W;
  const uint32_t step = step_number<S>(p);
  advance_seed<S>(p, step);
  const int nbb = p.nb_bnd;
  double acc[NP] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if ((int)blockIdx.x < nbb) {
    for (int e = 0; e < 256; e += 128) {
      const int i = blockIdx.x * 256 + e + (int)threadIdx.x;
      if (i >= nbn || (bn[i].flags & BN_OUTLET)) continue;
      const BNode& nd = bn[i];
      R F[9];
      pull_masked<S>(F, nd, src, p);
      apply_bc<S>(F, nd, src, p, nd.idx);
      if (nd.flags & BN_FEEDER) {
        R* side = (R*)o.side + (size_t)nd.side * 9;
#pragma unroll
        for (int k = 0; k < 9; k++) side[k] = F[k];
      }
      finish_node<S, COLL, 0, true>(F, src, dst, p, o, nd.idx, &nd, acc, -1, step);
    }
    return;
  }
  unsigned b = blockIdx.x - nbb;
  if (p.xcd_group > 0) b = xcd_group_block(b, (unsigned)p.xcd_group);
  const int tf = tflag[b];
  if (tf == 1) {
    const int c0 = p.bulk_begin + (int)(b * 256 + 2 * threadIdx.x);  // even
    const W* sw = reinterpret_cast<const W*>(src);
    R F0[9], F1[9];
#pragma unroll
    for (int k = 0; k < 9; k++) {
      const int s0 = k * p.plane + c0 - CX[k] - CY[k] * p.pitch;  // source of cell c0
      if (CX[k] == 0) {
        const W w = ld(sw + (s0 >> 1));
        F0[k] = PW::lo(w, p);
        F1[k] = PW::hi(w, p);
      } else {  // s0 odd: high half of the word before, low half of the word after
        F0[k] = PW::hi(ld(sw + ((s0 - 1) >> 1)), p);
        F1[k] = PW::lo(ld(sw + ((s0 + 1) >> 1)), p);
      }
    }
    R out0[9], out1[9];
    {
      R Rh, U, V, dr;
      macro_dev(F0, p.rho0_f, dr, Rh, U, V);
      collide_dev<COLL>(F0, dr, Rh, U, V, coll_par(p, -1), out0);
      macro_dev(F1, p.rho0_f, dr, Rh, U, V);
      collide_dev<COLL>(F1, dr, Rh, U, V, coll_par(p, -1), out1);
    }
    W* dw = reinterpret_cast<W*>(dst);
    uint32_t r = S::SR ? sr_pair_bits(c0, step) : 0u;
#pragma unroll
    for (int k = 0; k < 9; k++) {
      __builtin_nontemporal_store(PW::pack(out0[k], out1[k], r, p), dw + ((k * p.plane + c0) >> 1));
      if constexpr (S::SR) r = xorshift32(r);
    }
    return;
  }
  if (tf == 0) return;
  for (int h = 0; h < 256; h += 128) {  // per-cell path of k_step, two cells 128 apart
    const int idx = p.bulk_begin + (int)(b * 256 + h + threadIdx.x);
    const int c = idx < p.bulk_end ? (int)cls[idx] : 0;
    if (c != 1 && c != 4) continue;
    R F[9];
    pull_cell<S>(F, src, p, idx, c);
    finish_node<S, COLL, 0, false>(F, src, dst, p, o, idx, nullptr, acc, -1, step);
  }
}

// Open outlet (Iterate.c:533-541, 2nd order; cuda_step3 GpuBoundaries.cu:200-206,
// 1st order): one thread walks one row's outlet nodes in increasing i, reading
// the post-boundary F of its (i-1), (i-2) neighbours from the side buffer,
// which holds nodes of earlier columns already finished (as the reference's
// row-major sweep sees them).
template <class S, int COLL, int MODE>
__global__ __launch_bounds__(64) void k_outlet(const typename S::T* __restrict__ src,
                                               typename S::T* __restrict__ dst,
                                               const BNode* __restrict__ on,
                                               const int* __restrict__ row_start, int nrows,
                                               KParams p, KOut o) {
  using R = typename S::R;
  const uint32_t step = step_number<S>(p);
  if constexpr (MODE != 2) advance_seed<S>(p, step);
  const int r = blockIdx.x * 64 + threadIdx.x;
  double acc[NP] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (r < nrows) {
    R* side = (R*)o.side;
    for (int t = row_start[r]; t < row_start[r + 1]; t++) {
      const BNode& nd = on[t];
      R F[9];
      pull_masked<S>(F, nd, src, p);
      apply_bc<S>(F, nd, src, p, nd.idx);
      const R* a = side + (size_t)nd.feed1 * 9;
      const R* b = side + (size_t)nd.feed2 * 9;
      if (p.outlet_profile == 2) {
        F[1] = 2 * a[1] - b[1];
        F[5] = 2 * a[5] - b[5];
        F[8] = 2 * a[8] - b[8];
      } else {
        F[1] = a[1];
        F[5] = a[5];
        F[8] = a[8];
      }
      R* mine = side + (size_t)nd.side * 9;
#pragma unroll
      for (int k = 0; k < 9; k++) mine[k] = F[k];
      finish_node<S, COLL, MODE, true>(F, src, dst, p, o, nd.idx, &nd, acc, -1, step);
    }
  }
  if constexpr (MODE == 1)
    block_sum<64>(acc, o.part, p.nb_bulk + p.nb_bnd + blockIdx.x);
}

// First collision (serial/Iterate.c:268 on the state CellIni leaves): from
// F0 (0 / rest / feq of the given macros) and the macros in o.rho/u/v.
template <class S, int COLL>
__global__ __launch_bounds__(256) void k_init(typename S::T* __restrict__ dst,
                                              const uint8_t* __restrict__ cls, int f0_kind,
                                              KParams p, KOut o) {
  using R = typename S::R;
  const uint32_t step = step_number<S>(p);
  advance_seed<S>(p, step);
  const int idx = p.bulk_begin + (int)(blockIdx.x * 256 + threadIdx.x);
  if (idx >= p.bulk_end) return;
  if (cls[idx] == 0) {
#pragma unroll
    for (int k = 0; k < 9; k++)
      S::store_rn(dst + k * p.plane + idx, S::DEV ? (R)(-W64[k] * p.rho0) : (R)0, p);
    return;
  }
  const double rho = o.rho[idx], u = o.u[idx], v = o.v[idx];
  double F0[9];
#pragma unroll
  for (int k = 0; k < 9; k++) {
    if (f0_kind == 0) {
      F0[k] = 0.0;
    } else if (f0_kind == 1) {
      F0[k] = W64[k] * rho;
    } else {
      const double T1 = u * u + v * v;
      const double T2 = u * (double)CX[k] + v * (double)CY[k];
      F0[k] = rho * W64[k] * (1.0 + 3.0 * T2 + 4.5 * T2 * T2 - 1.5 * T1);
    }
  }
  if (f0_kind == 3) {  // + f^(1) = -3 tau w_k rho Q_k:grad(u), Q = c c - I/3
    const int inst = (S::DEV && p.ninst) ? (int)p.node_inst[idx] : -1;
    const double om = S::DEV ? (double)coll_par(p, inst == 0xffff ? -1 : inst).om : p.omega;
    const double gxx = o.grad[idx], gyy = o.grad[p.plane + idx],
                 gxy = o.grad[2 * (size_t)p.plane + idx];
#pragma unroll
    for (int k = 0; k < 9; k++) {
      const double cx = (double)CX[k], cy = (double)CY[k];
      const double q = (cx * cx - 1.0 / 3.0) * gxx + (cy * cy - 1.0 / 3.0) * gyy + cx * cy * gxy;
      F0[k] -= 3.0 / om * W64[k] * rho * q;
    }
  }
  R out[9];
  if constexpr (S::DEV) {
    float g[9];
#pragma unroll
    for (int k = 0; k < 9; k++) g[k] = (float)(F0[k] - W64[k] * p.rho0);
    const int inst = p.ninst ? (int)p.node_inst[idx] : -1;
    collide_dev<COLL>(g, (float)(rho - p.rho0), (float)rho, (float)u, (float)v,
                      coll_par(p, inst == 0xffff ? -1 : inst), out);
  } else {
    collide_f64<COLL>(F0, rho, u, v, p, out);
  }
  S::store_cell(dst, idx, out, p, step);
}

// Fill a storage buffer with the solid value everywhere (halo included).
template <class S>
__global__ void k_fill_solid(typename S::T* __restrict__ buf, KParams p) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= p.plane) return;
  using R = typename S::R;
#pragma unroll
  for (int k = 0; k < 9; k++)
    S::store_rn(buf + k * p.plane + i, S::DEV ? (R)(-W64[k] * p.rho0) : (R)0, p);
}

}  // namespace lbmdev
