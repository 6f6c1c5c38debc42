// residuals.hip — ComputeResiduals (serial/ComputeResiduals.c:6-51) over a
// lattice state the caller holds in device memory, in FP64.
//
// Built with -ffp-contract=off: the reference sweep adds d*d terms exactly as
// the C code's pow(d, 2) (correctly rounded, so equal to the product) and in
// its order.
//
// Layout of the inputs (the C ABI's per-cell convention): f and metaf hold
// n*m*9 doubles, cell-major ([idx*9 + k], idx = j*n + i, the reference's
// Cells[idx].F[k] / METAF[k]); boundary_id n*m ints, drag and lift n*m
// doubles (Cells[idx].BoundaryID / DragF / LiftF).
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "lbm_host.hpp"
#include "lbm_launch.hpp"

using lbmhost::check_hip;
using lbmhost::set_error;

namespace {

constexpr int RB = 256;                  // workgroup size of the partial-sum kernel
constexpr int RE = 32;                   // flat elements per thread
constexpr long RCHUNK = (long)RB * RE;   // flat elements per workgroup (one partial)
constexpr int RW = lbmdev::RESID_ROW;    // partial row: sum d^2, drag, lift, unused

__device__ inline double wave_sum64(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Workgroup b: the sum of (F - METAF)^2 over flat elements
// [b*RCHUNK, (b+1)*RCHUNK) and the drag / lift of cells [b*cchunk, (b+1)*cchunk)
// whose boundary id is drag_lift -> part[b*RW ..].  Thread t reads elements
// b*RCHUNK + r*RB + t (r = 0..RE-1): each wave load covers 512 consecutive
// bytes.  Per-thread sums in a fixed order, then a fixed butterfly and LDS
// order: the same partials on every run.
__global__ __launch_bounds__(RB) void k_resid_partials(const double* __restrict__ f,
                                                       const double* __restrict__ metaf,
                                                       const int* __restrict__ bid,
                                                       const double* __restrict__ drag,
                                                       const double* __restrict__ lift,
                                                       long nflat, long ncell, long cchunk,
                                                       int drag_lift, double* __restrict__ part) {
  const long b = blockIdx.x;
  const int t = threadIdx.x;
  double s = 0.0, dr = 0.0, li = 0.0;
  const long e0 = b * RCHUNK + t;
#pragma unroll 8
  for (int r = 0; r < RE; r++) {
    const long e = e0 + (long)r * RB;
    if (e < nflat) {
      const double d = f[e] - metaf[e];
      s += d * d;
    }
  }
  if (bid) {
    const long c1 = (b + 1) * cchunk < ncell ? (b + 1) * cchunk : ncell;
    for (long c = b * cchunk + t; c < c1; c += RB) {
      if (bid[c] != drag_lift) continue;
      if (drag) dr += drag[c];
      if (lift) li += lift[c];
    }
  }
  __shared__ double red[RB / 64][3];
  const int lane = t & 63, wv = t >> 6;
  s = wave_sum64(s);
  dr = wave_sum64(dr);
  li = wave_sum64(li);
  if (lane == 0) {
    red[wv][0] = s;
    red[wv][1] = dr;
    red[wv][2] = li;
  }
  __syncthreads();
  if (t < 3) {
    double a = 0.0;
#pragma unroll
    for (int w = 0; w < RB / 64; w++) a += red[w][t];
    part[b * RW + t] = a;
  }
  if (t == 3) part[b * RW + 3] = 0.0;
}

// Fixed-order sum of nparts partial rows -> out[0..3] as ComputeResiduals
// leaves Residuals[] (serial/ComputeResiduals.c:36-44): sqrt(L2), the
// sqrt(L2 / (m n)) weighted norm, drag, lift.
__global__ __launch_bounds__(256) void k_resid_final(const double* __restrict__ part, long nparts,
                                                     int n, int m, double* __restrict__ out) {
  __shared__ double red[3][256];
  const int t = threadIdx.x;
  double a[3] = {0.0, 0.0, 0.0};
  for (long i = t; i < nparts; i += 256) {
#pragma unroll
    for (int c = 0; c < 3; c++) a[c] = a[c] + part[i * RW + c];
  }
#pragma unroll
  for (int c = 0; c < 3; c++) red[c][t] = a[c];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
#pragma unroll
      for (int c = 0; c < 3; c++) red[c][t] = red[c][t] + red[c][t + w];
    }
    __syncthreads();
  }
  if (t == 0) {
    const double l2 = red[0][0];
    out[1] = sqrt(l2 / (double)(m * n));
    out[0] = sqrt(l2);
    out[2] = red[1][0];
    out[3] = red[2][0];
  }
}

// The reference's sweep as written (serial/ComputeResiduals.c:19-33): for
// i < m, for j < n, for k < 9, L2 += (F - METAF)^2 of cell i — the first m
// cells, n times each — and drag / lift of cell i added n times when its
// BoundaryID equals ComputeDragLift; one thread, the reference's order.  A
// cell whose nine differences are exact zeros (and that adds no drag / lift)
// is skipped: adding +0 leaves each running sum unchanged.
__global__ void k_resid_reference(const double* __restrict__ f, const double* __restrict__ metaf,
                                  const int* __restrict__ bid, const double* __restrict__ drag,
                                  const double* __restrict__ lift, int n, int m, int drag_lift,
                                  double* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double l2 = 0.0, dr = 0.0, li = 0.0;
  for (int i = 0; i < m; i++) {
    double d[9];
    bool nz = false;
    for (int k = 0; k < 9; k++) {
      d[k] = f[(long)i * 9 + k] - metaf[(long)i * 9 + k];
      nz = nz || (d[k] != 0.0);  // NaN compares unequal: kept
    }
    const bool dl = bid && bid[i] == drag_lift;
    const double di = (dl && drag) ? drag[i] : 0.0;
    const double lv = (dl && lift) ? lift[i] : 0.0;
    const bool dz = (di != 0.0) || (lv != 0.0);
    if (!nz && !dz) continue;
    for (int j = 0; j < n; j++) {
      if (nz)
        for (int k = 0; k < 9; k++) l2 = l2 + d[k] * d[k];
      if (dz) {
        dr += di;
        li += lv;
      }
    }
  }
  out[1] = sqrt(l2 / (double)(m * n));
  out[0] = sqrt(l2);
  out[2] = dr;
  out[3] = li;
}

}  // namespace

namespace lbmdev {

long residual_partials(int n, int m) {
  const long nflat = 9L * n * m;
  return (nflat + RCHUNK - 1) / RCHUNK;
}

int check_residual_args(const double* f, const double* metaf, int n, int m, int kind,
                        double* out) {
  if (!f || !metaf || !out || n < 1 || m < 1) {
    set_error("compute_residuals: f, metaf and out must be device arrays, n and m >= 1");
    return LBM_EINVAL;
  }
  if (kind != LBM_RESID_REFERENCE && kind != LBM_RESID_FULL) {
    set_error("compute_residuals: kind must be LBM_RESID_REFERENCE or LBM_RESID_FULL");
    return LBM_EINVAL;
  }
  return LBM_OK;
}

// The launches of lbm_compute_residuals; `part` holds residual_partials(n, m)
// rows of RESID_ROW doubles (LBM_RESID_FULL only).
hipError_t launch_residuals(const double* f, const double* metaf, const int* bid,
                            const double* drag, const double* lift, int n, int m, int drag_lift,
                            int kind, double* out, double* part, hipStream_t st) {
  if (kind == LBM_RESID_REFERENCE) {
    k_resid_reference<<<1, 64, 0, st>>>(f, metaf, bid, drag, lift, n, m, drag_lift, out);
    return hipGetLastError();
  }
  const long nflat = 9L * n * m, ncell = (long)n * m;
  const long P = residual_partials(n, m);
  const long cchunk = (ncell + P - 1) / P;
  k_resid_partials<<<(unsigned)P, RB, 0, st>>>(f, metaf, bid, drag, lift, nflat, ncell, cchunk,
                                                drag_lift, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  k_resid_final<<<1, 256, 0, st>>>(part, P, n, m, out);
  return hipGetLastError();
}

}  // namespace lbmdev

using namespace lbmdev;

extern "C" int lbm_compute_residuals(const double* f, const double* metaf, const int* boundary_id,
                                     const double* drag, const double* lift, int n, int m,
                                     int drag_lift, int kind, double* out, void* stream) {
  int rc = check_residual_args(f, metaf, n, m, kind, out);
  if (rc) return rc;
  const hipStream_t st = (hipStream_t)stream;
  double* part = nullptr;
  if (kind == LBM_RESID_FULL) {
    const size_t bytes = sizeof(double) * RESID_ROW * (size_t)residual_partials(n, m);
    rc = check_hip(hipMallocAsync((void**)&part, bytes, st), "compute_residuals: hipMallocAsync");
    if (rc) return rc;
  }
  rc = check_hip(launch_residuals(f, metaf, boundary_id, drag, lift, n, m, drag_lift, kind, out,
                                  part, st),
                 "compute_residuals: launch");
  if (part) {
    const int rf = check_hip(hipFreeAsync(part, st), "compute_residuals: hipFreeAsync");
    if (!rc) rc = rf;
  }
  return rc;
}
