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

// ---------------------------------------------------------------- copy probe
// A plain streaming copy (16-byte vectors, non-temporal stores): the
// practical HBM ceiling the step kernels are compared with, and the known-size
// kernel that calibrates hardware byte counters.
namespace lbmdev {

typedef float v4f __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_copy(const v4f* __restrict__ src,
                                              v4f* __restrict__ dst, long n4) {
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  const long stride = (long)gridDim.x * 256;
  for (; i < n4; i += stride) __builtin_nontemporal_store(src[i], dst + i);
}

hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t st) {
  const long n4 = (long)(bytes / 16);
  long blocks = (n4 + 255) / 256;
  if (blocks > 256 * 64) blocks = 256 * 64;
  k_copy<<<(unsigned)blocks, 256, 0, st>>>((const v4f*)src, (v4f*)dst, n4);
  return hipGetLastError();
}

}  // namespace lbmdev
