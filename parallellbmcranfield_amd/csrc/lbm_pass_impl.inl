// lbm_pass_impl.inl — grid shapes of the passes (one copy per storage unit).
#pragma once
#include <type_traits>
#include "lbm_kernels.inl"
#include "lbm_launch.hpp"

namespace lbmdev {

template <class S, int COLL, int MODE>
static hipError_t bulk_launch(const Launch& L) {
  using T = typename S::T;
  const T* src = (const T*)L.src;
  T* dst = (T*)L.dst;
  switch (L.block) {
    case 512:
      k_bulk<S, COLL, MODE, 512><<<L.nb_bulk, 512, 0, L.stream>>>(src, dst, L.cls, L.tflag, L.kp, L.out);
      break;
    case 1024:
      k_bulk<S, COLL, MODE, 1024><<<L.nb_bulk, 1024, 0, L.stream>>>(src, dst, L.cls, L.tflag, L.kp, L.out);
      break;
    default:
      k_bulk<S, COLL, MODE, 256><<<L.nb_bulk, 256, 0, L.stream>>>(src, dst, L.cls, L.tflag, L.kp, L.out);
      break;
  }
  return hipGetLastError();
}

template <class S, int COLL, int MODE>
static hipError_t fused_launch(const Launch& L) {
  using T = typename S::T;
  const T* src = (const T*)L.src;
  T* dst = (T*)L.dst;
  if constexpr ((std::is_same_v<S, StF16> || std::is_same_v<S, StF16S>) && MODE == 0) {
    if (L.pairs) {  // paired-cell plain step: one tile per workgroup of 128 threads
      const int grid = L.kp.nb_bnd + L.nb_bulk;
      if (grid <= 0) return hipSuccess;
      k_step_pair<S, COLL><<<grid, 128, 0, L.stream>>>(src, dst, L.cls, L.tflag, L.bn, L.nbn,
                                                       L.kp, L.out);
      return hipGetLastError();
    }
  }
  const int grid = L.kp.nb_bnd + L.nb_bulk;
  if (grid <= 0) return hipSuccess;
  k_step<S, COLL, MODE, 256><<<grid, 256, 0, L.stream>>>(src, dst, L.cls, L.tflag, L.bn, L.nbn,
                                                        L.kp, L.out);
  return hipGetLastError();
}

template <class S, int COLL, int MODE>
static hipError_t pass_mode(int pass, const Launch& L) {
  using T = typename S::T;
  const T* src = (const T*)L.src;
  T* dst = (T*)L.dst;
  switch (pass) {
    case PASS_BULK:
      if (L.nb_bulk <= 0) return hipSuccess;
      return bulk_launch<S, COLL, MODE>(L);
    case PASS_BOUNDARY:
      if (L.nbn <= 0) return hipSuccess;
      k_boundary<S, COLL, MODE><<<(L.nbn + 255) / 256, 256, 0, L.stream>>>(src, dst, L.bn, L.nbn,
                                                                           L.kp, L.out);
      return hipGetLastError();
    case PASS_FUSED:
      return fused_launch<S, COLL, MODE>(L);
    case PASS_TWO:
      if constexpr (MODE == 0 && !S::SR) {
        if (L.ntiles <= 0) return hipSuccess;
        k_two_steps<S, COLL><<<L.nband + L.ntiles, TWO_COLS, 0, L.stream>>>(
            src, dst, L.cls, L.tiles, L.kp, L.nband, L.bgrid, (T*)L.bdst, L.bcls, L.btflag, L.bn,
            L.nbn, L.blist, L.out);
        return hipGetLastError();
      }
      return hipErrorInvalidValue;
    case PASS_THREE:
      if constexpr (MODE == 0 && !S::SR) {
        if (L.ntiles <= 0) return hipSuccess;
        if (L.kp.ninst) return hipErrorInvalidValue;
        k_three_steps<S, COLL><<<L.nband + L.ntiles, TWO_COLS, 0, L.stream>>>(
            src, dst, L.cls, L.tiles, L.kp, L.nband, L.bgrid, (T*)L.bdst, L.bcls, L.btflag, L.bn,
            L.nbn, L.blist, L.out);
        return hipGetLastError();
      }
      return hipErrorInvalidValue;
    case PASS_OUTLET:
      if (L.norows <= 0) return hipSuccess;
      k_outlet<S, COLL, MODE><<<(L.norows + 63) / 64, 64, 0, L.stream>>>(
          src, dst, L.on, L.orow, L.norows, L.kp, L.out);
      return hipGetLastError();
  }
  return hipErrorInvalidValue;
}

template <class S, int COLL>
static hipError_t pass_impl(int mode, int pass, const Launch& L) {
  switch (mode) {
    case MODE_STEP: return pass_mode<S, COLL, 0>(pass, L);
    case MODE_RESID: return pass_mode<S, COLL, 1>(pass, L);
    case MODE_OBSERVE: return pass_mode<S, COLL, 2>(pass, L);
  }
  return hipErrorInvalidValue;
}

// Resident workgroups of k_two_steps per CU; 0 where the kernel does not
// exist.
template <class S, int COLL>
static int two_step_occupancy_impl(int steps) {
  if constexpr (S::SR) {
    return 0;
  } else {
    int nb = 0;
    const hipError_t e =
        steps == 3
            ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_three_steps<S, COLL>, TWO_COLS, 0)
            : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_two_steps<S, COLL>, TWO_COLS, 0);
    return e == hipSuccess ? nb : 0;
  }
}

template <class S>
static int two_step_occupancy_coll(int coll, int steps) {
  switch (coll) {
    case 1: return two_step_occupancy_impl<S, 1>(steps);
    case 2: return two_step_occupancy_impl<S, 2>(steps);
    case 3: return two_step_occupancy_impl<S, 3>(steps);
  }
  return 0;
}

template <class S, int COLL>
static hipError_t init_impl(int f0_kind, const Launch& L) {
  using T = typename S::T;
  const int count = L.kp.bulk_end - L.kp.bulk_begin;
  if (count <= 0) return hipSuccess;
  k_init<S, COLL><<<(count + 255) / 256, 256, 0, L.stream>>>((T*)L.dst, L.cls, f0_kind, L.kp, L.out);
  return hipGetLastError();
}

template <class S>
static hipError_t fill_impl(const Launch& L, void* buf) {
  using T = typename S::T;
  k_fill_solid<S><<<(L.kp.plane + 255) / 256, 256, 0, L.stream>>>((T*)buf, L.kp);
  return hipGetLastError();
}

}  // namespace lbmdev
