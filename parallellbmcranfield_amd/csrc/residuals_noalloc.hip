// residuals_noalloc.hip — lbm_compute_residuals without any allocation, for
// callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc).
//
// The partial sums of LBM_RESID_FULL go to a fixed array in the library's
// device image instead of a per-call allocation.  Each GPU the library runs
// on holds its own copy of the array until the process ends.  Every call on
// a GPU shares that GPU's array: two calls that overlap on one GPU (different
// streams or host threads) race on it, so callers serialise their calls per
// GPU.  Each call writes every partial row it reads before reading it (the
// partial kernel fills rows 0..P-1, the final kernel sums exactly those), so
// no clearing is needed and a previous call's contents never leak into a
// result.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "lbm_host.hpp"
#include "lbm_launch.hpp"

using lbmhost::check_hip;
using lbmhost::set_error;
using namespace lbmdev;

// Partial-sum rows for lattices up to LBM_NOALLOC_MAX_CELLS cells: 32768
// rows of RESID_ROW doubles, 1 MiB per GPU.
constexpr long NOALLOC_ROWS = 32768;
static_assert(NOALLOC_ROWS * 8192 / 9 >= LBM_NOALLOC_MAX_CELLS, "scratch below the stated maximum");

static __device__ __align__(256) double scratch[NOALLOC_ROWS * RESID_ROW];

namespace {

// Device address of `scratch` on each GPU, looked up on the first call made
// with that GPU current.  The array's name is a host-side stand-in: only
// these addresses are handed to kernels.
constexpr int MAX_DEVICES = 256;
std::mutex g_mu;
double* g_scratch[MAX_DEVICES] = {};

int scratch_address(double** out) {
  int dev = -1;
  int rc = check_hip(hipGetDevice(&dev), "compute_residuals_noalloc: hipGetDevice");
  if (rc) return rc;
  if (dev < 0 || dev >= MAX_DEVICES) {
    set_error("compute_residuals_noalloc: device ordinal " + std::to_string(dev) +
              " beyond the " + std::to_string(MAX_DEVICES) + " devices it keeps addresses for");
    return LBM_EINVAL;
  }
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    rc = check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)),
                   "compute_residuals_noalloc: hipGetSymbolAddress(scratch)");
    if (rc) return rc;
    if (!p) {
      set_error("compute_residuals_noalloc: no device address for the scratch array");
      return LBM_EHIP;
    }
    g_scratch[dev] = (double*)p;
  }
  *out = g_scratch[dev];
  return LBM_OK;
}

}  // namespace

extern "C" int lbm_compute_residuals_noalloc(const double* f, const double* metaf,
                                             const int* boundary_id, const double* drag,
                                             const double* lift, int n, int m, int drag_lift,
                                             int kind, double* out, void* stream) {
  int rc = check_residual_args(f, metaf, n, m, kind, out);
  if (rc) return rc;
  double* part = nullptr;
  if (kind == LBM_RESID_FULL) {
    if ((long)n * m > LBM_NOALLOC_MAX_CELLS) {
      set_error("compute_residuals_noalloc: " + std::to_string((long)n * m) +
                " cells exceed the fixed scratch's maximum of " +
                std::to_string((long)LBM_NOALLOC_MAX_CELLS) +
                " (use lbm_compute_residuals)");
      return LBM_EINVAL;
    }
    rc = scratch_address(&part);
    if (rc) return rc;
  }
  return check_hip(launch_residuals(f, metaf, boundary_id, drag, lift, n, m, drag_lift, kind, out,
                                    part, (hipStream_t)stream),
                   "compute_residuals_noalloc: launch");
}
