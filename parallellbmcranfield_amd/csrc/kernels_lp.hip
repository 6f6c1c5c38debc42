// kernels_lp.hip — FP32 and FP16 storage instantiations (deviation form,
// FP32 arithmetic; the compiler may contract multiply-adds here).
#include "lbm_kernels.inl"
#include "lbm_launch.hpp"
#include "lbm_pass_impl.inl"

namespace lbmdev {

hipError_t launch_pass_lp(int prec, int coll, int mode, int pass, const Launch& L) {
  if (prec == 1) {
    switch (coll) {
      case 1: return pass_impl<StF32, 1>(mode, pass, L);
      case 2: return pass_impl<StF32, 2>(mode, pass, L);
      case 3: return pass_impl<StF32, 3>(mode, pass, L);
    }
  } else if (prec == 2) {
    switch (coll) {
      case 1: return pass_impl<StF16, 1>(mode, pass, L);
      case 2: return pass_impl<StF16, 2>(mode, pass, L);
      case 3: return pass_impl<StF16, 3>(mode, pass, L);
    }
  }
  return hipErrorInvalidValue;
}

hipError_t launch_init_lp(int prec, int coll, int f0_kind, const Launch& L) {
  if (prec == 1) {
    switch (coll) {
      case 1: return init_impl<StF32, 1>(f0_kind, L);
      case 2: return init_impl<StF32, 2>(f0_kind, L);
      case 3: return init_impl<StF32, 3>(f0_kind, L);
    }
  } else if (prec == 2) {
    switch (coll) {
      case 1: return init_impl<StF16, 1>(f0_kind, L);
      case 2: return init_impl<StF16, 2>(f0_kind, L);
      case 3: return init_impl<StF16, 3>(f0_kind, L);
    }
  }
  return hipErrorInvalidValue;
}

hipError_t launch_fill_solid_lp(int prec, const Launch& L, void* buf) {
  if (prec == 1) return fill_impl<StF32>(L, buf);
  if (prec == 2) return fill_impl<StF16>(L, buf);
  return hipErrorInvalidValue;
}

}  // namespace lbmdev
