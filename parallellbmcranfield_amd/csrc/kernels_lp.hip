// kernels_lp.hip — FP32 and FP16 storage instantiations (deviation form,
// FP32 arithmetic; the compiler may contract multiply-adds here).  prec 3 is
// FP16 storage with stochastic rounding.
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
  } else if (prec == 3) {
    switch (coll) {
      case 1: return pass_impl<StF16S, 1>(mode, pass, L);
      case 2: return pass_impl<StF16S, 2>(mode, pass, L);
      case 3: return pass_impl<StF16S, 3>(mode, pass, L);
    }
  }
  return hipErrorInvalidValue;
}

int two_step_occupancy_lp(int prec, int coll, int steps) {
  if (prec == 1) return two_step_occupancy_coll<StF32>(coll, steps);
  if (prec == 2) return two_step_occupancy_coll<StF16>(coll, steps);
  return 0;
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
  } else if (prec == 3) {
    switch (coll) {
      case 1: return init_impl<StF16S, 1>(f0_kind, L);
      case 2: return init_impl<StF16S, 2>(f0_kind, L);
      case 3: return init_impl<StF16S, 3>(f0_kind, L);
    }
  }
  return hipErrorInvalidValue;
}

hipError_t launch_fill_solid_lp(int prec, const Launch& L, void* buf) {
  if (prec == 1) return fill_impl<StF32>(L, buf);
  if (prec == 2) return fill_impl<StF16>(L, buf);
  if (prec == 3) return fill_impl<StF16S>(L, buf);
  return hipErrorInvalidValue;
}

}  // namespace lbmdev

// ---------------------------------------------------------------- copy probe
// A plain streaming copy: the practical HBM ceiling the step kernels are
// compared with, and the known-size kernel that calibrates hardware byte
// counters.  Each thread moves U 16-byte vectors per trip, all U loads issued
// before the first store, so U loads are in flight per lane (one per trip
// leaves the copy latency-bound, below the step kernel's own rate).  Vector u
// of a trip sits u * stride elements after the first, so every load and store
// instruction of a wave covers 1 KiB of consecutive bytes.  Stores are
// non-temporal; loads are too when NT_LOAD.
namespace lbmdev {

typedef float v4f __attribute__((ext_vector_type(4)));

template <int U, bool NT_LOAD>
__global__ __launch_bounds__(256) void k_copy(const v4f* __restrict__ src,
                                              v4f* __restrict__ dst, long n4) {
  const long stride = (long)gridDim.x * 256;
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  for (; i + (long)(U - 1) * stride < n4; i += (long)U * stride) {
    v4f r[U];
#pragma unroll
    for (int u = 0; u < U; u++)
      r[u] = NT_LOAD ? __builtin_nontemporal_load(src + i + u * stride) : src[i + u * stride];
#pragma unroll
    for (int u = 0; u < U; u++) __builtin_nontemporal_store(r[u], dst + i + u * stride);
  }
  for (; i < n4; i += stride) __builtin_nontemporal_store(src[i], dst + i);
}

// shape: 0 one vector per trip over a grid of at most 16384 workgroups (the
// original probe); 1..6 U = 2, 4, 8 vectors per thread, normal or
// non-temporal loads, one trip per thread over the whole buffer.
hipError_t launch_copy_shape(const void* src, void* dst, size_t bytes, int shape,
                             hipStream_t st) {
  const long n4 = (long)(bytes / 16);
  auto grid = [&](int U) {
    long b = (n4 + 256L * U - 1) / (256L * U);
    return (unsigned)(b < 1 ? 1 : b);
  };
  const v4f* s = (const v4f*)src;
  v4f* d = (v4f*)dst;
  switch (shape) {
    case 0: {
      long blocks = (n4 + 255) / 256;
      if (blocks > 256 * 64) blocks = 256 * 64;
      if (blocks < 1) blocks = 1;
      k_copy<1, false><<<(unsigned)blocks, 256, 0, st>>>(s, d, n4);
    } break;
    case 1: k_copy<2, false><<<grid(2), 256, 0, st>>>(s, d, n4); break;
    case 2: k_copy<4, false><<<grid(4), 256, 0, st>>>(s, d, n4); break;
    case 3: k_copy<8, false><<<grid(8), 256, 0, st>>>(s, d, n4); break;
    case 4: k_copy<2, true><<<grid(2), 256, 0, st>>>(s, d, n4); break;
    case 5: k_copy<4, true><<<grid(4), 256, 0, st>>>(s, d, n4); break;
    case 6: k_copy<8, true><<<grid(8), 256, 0, st>>>(s, d, n4); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t st) {
  return launch_copy_shape(src, dst, bytes, 2, st);
}

}  // namespace lbmdev

// ---------------------------------------------------------------- pull-stream probe
// The step kernel's memory traffic without its arithmetic: each cell of
// storage rows 1..m loads its nine values from the upstream neighbours (the
// step's offsets, planes, 256-cell tiles and grouped XCD order) and stores
// them with non-temporal stores into the other buffer.  Its bandwidth is the
// practical ceiling of the step's access pattern on the device.  Every pull
// stays inside the nine planes: rows 0 and m+1 are halo, and a plane holds
// at least (m + 2) * pitch elements.
namespace lbmdev {

template <class T>
__global__ __launch_bounds__(256) void k_pull_probe(const T* __restrict__ src,
                                                    T* __restrict__ dst, int plane, int pitch,
                                                    int begin, int end, unsigned G) {
  unsigned b = blockIdx.x;
  if (G > 0) b = xcd_group_block(b, G);
  const int idx = begin + (int)(b * 256 + threadIdx.x);
  if (idx >= end) return;
  T v[9];
#pragma unroll
  for (int k = 0; k < 9; k++) v[k] = ld(src + k * plane + idx - CX[k] - CY[k] * pitch);
#pragma unroll
  for (int k = 0; k < 9; k++) __builtin_nontemporal_store(v[k], dst + k * plane + idx);
}

hipError_t launch_pull_probe(int elem_bytes, const void* src, void* dst, int plane, int pitch,
                             int begin, int end, int G, hipStream_t st) {
  const unsigned g = G > 0 ? (unsigned)G : 0u;
  const long align = 8L * (g > 0 ? g : 1);
  long nb = (long)(end - begin + 255) / 256;
  nb = (nb + align - 1) / align * align;  // the grouped order is a permutation of whole runs
  if (elem_bytes == 4)
    k_pull_probe<float><<<(unsigned)nb, 256, 0, st>>>((const float*)src, (float*)dst, plane,
                                                      pitch, begin, end, g);
  else if (elem_bytes == 8)
    k_pull_probe<double><<<(unsigned)nb, 256, 0, st>>>((const double*)src, (double*)dst, plane,
                                                       pitch, begin, end, g);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace lbmdev

