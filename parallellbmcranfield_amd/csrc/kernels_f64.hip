// kernels_f64.hip — FP64 (reference-order) instantiations and the FP64
// reductions.  Built with -ffp-contract=off: no fused multiply-add, so every
// FP64 result follows the reference's C expressions operation by operation.
#include "lbm_kernels.inl"
#include "lbm_launch.hpp"
#include "lbm_pass_impl.inl"

namespace lbmdev {

hipError_t launch_pass_f64(int coll, int mode, int pass, const Launch& L) {
  switch (coll) {
    case 1: return pass_impl<StF64, 1>(mode, pass, L);
    case 2: return pass_impl<StF64, 2>(mode, pass, L);
    case 3: return pass_impl<StF64, 3>(mode, pass, L);
  }
  return hipErrorInvalidValue;
}

int two_step_occupancy_f64(int coll, int steps) { return two_step_occupancy_coll<StF64>(coll, steps); }

hipError_t launch_init_f64(int coll, int f0_kind, const Launch& L) {
  switch (coll) {
    case 1: return init_impl<StF64, 1>(f0_kind, L);
    case 2: return init_impl<StF64, 2>(f0_kind, L);
    case 3: return init_impl<StF64, 3>(f0_kind, L);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_fill_solid_f64(const Launch& L, void* buf) { return fill_impl<StF64>(L, buf); }

// ---------------------------------------------------------------- reductions
// Sum of nparts x 6 partials in a fixed order: thread t adds parts t, t+256,
// ... in sequence, then a fixed-shape LDS tree.  Same result on every run.
__global__ __launch_bounds__(256) void k_reduce_partials(const double* __restrict__ part,
                                                         int nparts, double* __restrict__ out) {
  constexpr int W = 6;
  __shared__ double red[W][256];
  const int t = threadIdx.x;
  double s[W] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int i = t; i < nparts; i += 256) {
#pragma unroll
    for (int c = 0; c < W; c++) s[c] = s[c] + part[(size_t)i * W + c];
  }
#pragma unroll
  for (int c = 0; c < W; c++) red[c][t] = s[c];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
#pragma unroll
      for (int c = 0; c < W; c++) red[c][t] = red[c][t] + red[c][t + w];
    }
    __syncthreads();
  }
  if (t < W) out[t] = red[t][0];
}

// serial/ComputeResiduals.c:19-30: for i < m, for j < n, for k < 9,
// L2 += (F-METAF)^2 of cell i; drag/lift summed n times per matching cell.
// Cells never written by the boundary pass are solid (all terms zero), and
// adding an exact zero leaves the running sum unchanged, so they are skipped.
__global__ void k_ref_residual(const double* __restrict__ cres, int m, int n,
                               double* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double l2 = 0.0, dr = 0.0, li = 0.0;
  for (int i = 0; i < m; i++) {
    const double* c = cres + (size_t)i * 12;
    if (c[11] == 0.0) continue;
    double d[9];
    bool nz = false;
    for (int k = 0; k < 9; k++) {
      d[k] = c[k];
      nz = nz || (d[k] != 0.0);
    }
    const bool dz = (c[9] != 0.0) || (c[10] != 0.0);
    if (!nz && !dz) continue;
    for (int j = 0; j < n; j++) {
      if (nz)
        for (int k = 0; k < 9; k++) l2 = l2 + d[k] * d[k];
      if (dz) {
        dr += c[9];
        li += c[10];
      }
    }
  }
  out[0] = l2;
  out[1] = dr;
  out[2] = li;
  out[3] = 0.0;
}

// Velocity-change sums of each batched instance over its storage rectangle:
// out[2k] = sum |u(t) - u(t-K)|^2, out[2k+1] = sum |u(t)|^2, from the per-cell
// terms the residual step leaves in planes 3 and 4 of the snapshot.  One
// workgroup per instance, fixed order: the sums are deterministic.
__global__ __launch_bounds__(256) void k_inst_resid(const double* __restrict__ snap, int plane,
                                                    int pitch, const int* __restrict__ rect,
                                                    const uint8_t* __restrict__ active,
                                                    double* __restrict__ out) {
  const int k = blockIdx.x;
  const int x0 = rect[4 * k], y0 = rect[4 * k + 1], w = rect[4 * k + 2], h = rect[4 * k + 3];
  double a = 0.0, b = 0.0;
  if (active[k]) {
    const long cnt = (long)w * h;
    for (long e = threadIdx.x; e < cnt; e += 256) {
      const long r = e / w, c = e - r * w;
      const long idx = (y0 + r) * (long)pitch + x0 + c;
      a += snap[3 * (size_t)plane + idx];
      b += snap[4 * (size_t)plane + idx];
    }
  }
  __shared__ double red[4][2];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  a = wave_sum(a);
  b = wave_sum(b);
  if (lane == 0) {
    red[wv][0] = a;
    red[wv][1] = b;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    double s = 0.0;
    for (int w2 = 0; w2 < 4; w2++) s += red[w2][threadIdx.x];
    out[2 * k + threadIdx.x] = s;
  }
}

hipError_t launch_inst_resid(const double* snap, int plane, int pitch, const int* rect,
                             const uint8_t* active, int ninst, double* out, hipStream_t st) {
  if (ninst <= 0) return hipSuccess;
  k_inst_resid<<<ninst, 256, 0, st>>>(snap, plane, pitch, rect, active, out);
  return hipGetLastError();
}

hipError_t launch_reduce_partials(const double* part, int nparts, double* out, hipStream_t st) {
  k_reduce_partials<<<1, 256, 0, st>>>(part, nparts, out);
  return hipGetLastError();
}

hipError_t launch_ref_residual(const double* cres, int m, int n, double* out, hipStream_t st) {
  k_ref_residual<<<1, 64, 0, st>>>(cres, m, n, out);
  return hipGetLastError();
}

}  // namespace lbmdev
