// lbm_launch.hpp — host-side launchers of the kernels (defined in the .hip units).
#pragma once
#include <hip/hip_runtime.h>

#include "lbm_kparams.hpp"

namespace lbmdev {

struct Launch {
  const void* src = nullptr;   // METAF_t
  void* dst = nullptr;         // METAF_{t+1}
  const uint8_t* cls = nullptr;
  const uint8_t* tflag = nullptr;  // per bulk tile: 0 none, 1 all interior, 2 mixed
  const BNode* bn = nullptr;   // boundary-list nodes
  int nbn = 0;
  const BNode* on = nullptr;   // open-outlet nodes, grouped by row
  const int* orow = nullptr;   // row groups: [orow[r], orow[r+1])
  int norows = 0;
  int nb_bulk = 0;             // bulk grid (multiple of 8)
  int block = 256;             // bulk workgroup size
  bool pairs = false;          // FP16 plain steps: paired-cell kernel (k_step_pair)
  const int4* tiles = nullptr; // two-step tiles (x0, y0, rows, walls), one workgroup each
  int ntiles = 0;
  // the band's first single step in the two-step launch: bgrid workgroups of
  // k_step's work (boundary list + bulk tiles `blist` of classes `bcls`) into
  // bdst, padded to nband (a multiple of 8) ahead of the tiles
  int nband = 0, bgrid = 0;
  void* bdst = nullptr;
  const uint8_t* bcls = nullptr;
  const uint8_t* btflag = nullptr;
  const uint32_t* blist = nullptr;
  KParams kp{};
  KOut out{};
  hipStream_t stream = nullptr;
};

// PASS_TWO / PASS_THREE: two / three plain steps of the multi-step tiles
// (MODE_STEP, not with stochastic rounding; three: not for packed lattices
// of instances).
enum Pass {
  PASS_BULK = 0,
  PASS_BOUNDARY = 1,
  PASS_OUTLET = 2,
  PASS_FUSED = 3,
  PASS_TWO = 4,
  PASS_THREE = 5
};
enum Mode { MODE_STEP = 0, MODE_RESID = 1, MODE_OBSERVE = 2 };

// prec: 0 F64, 1 F32, 2 F16.  coll: 1 BGKW, 2 TRT, 3 MRT.
hipError_t launch_pass(int prec, int coll, int mode, int pass, const Launch& L);
hipError_t launch_init(int prec, int coll, int f0_kind, const Launch& L);
hipError_t launch_fill_solid(int prec, const Launch& L, void* buf);

// FP64 launchers live in the -ffp-contract=off unit.
hipError_t launch_pass_f64(int coll, int mode, int pass, const Launch& L);
hipError_t launch_init_f64(int coll, int f0_kind, const Launch& L);
hipError_t launch_fill_solid_f64(const Launch& L, void* buf);
hipError_t launch_pass_lp(int prec, int coll, int mode, int pass, const Launch& L);
hipError_t launch_init_lp(int prec, int coll, int f0_kind, const Launch& L);
hipError_t launch_fill_solid_lp(int prec, const Launch& L, void* buf);
// Resident workgroups of the two-step (steps = 2) or three-step (3) kernel
// per CU (0: no such kernel).
int two_step_occupancy_f64(int coll, int steps = 2);
int two_step_occupancy_lp(int prec, int coll, int steps = 2);

// Deterministic FP64 sum of nparts x 6 partials -> out[6].
hipError_t launch_reduce_partials(const double* part, int nparts, double* out, hipStream_t st);
// The reference residual's sweep (serial/ComputeResiduals.c:19-30) over the
// cres rows of the first m cells -> out[4] = (sum, drag, lift, 0).
hipError_t launch_inst_resid(const double* snap, int plane, int pitch, const int* rect,
                             const uint8_t* active, int ninst, double* out, hipStream_t st);
hipError_t launch_ref_residual(const double* cres, int m, int n, double* out, hipStream_t st);

// ComputeResiduals over device arrays (residuals.hip).  LBM_RESID_FULL sums
// 8192 flat values per partial row of RESID_ROW doubles in `part`
// (residual_partials(n, m) rows); LBM_RESID_REFERENCE needs no `part`.
constexpr int RESID_ROW = 4;
long residual_partials(int n, int m);
int check_residual_args(const double* f, const double* metaf, int n, int m, int kind, double* out);
hipError_t launch_residuals(const double* f, const double* metaf, const int* bid,
                            const double* drag, const double* lift, int n, int m, int drag_lift,
                            int kind, double* out, double* part, hipStream_t st);

// Plain device copy of `bytes` (multiple of 16); launch_copy_shape picks one
// of COPY_SHAPES launch shapes (vectors in flight per thread, load kind).
constexpr int COPY_SHAPES = 7;
hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t st);
hipError_t launch_copy_shape(const void* src, void* dst, size_t bytes, int shape, hipStream_t st);
// The step's loads and stores without arithmetic over an n x m lattice's
// padded layout (plane, pitch; cells [begin, end)), elem_bytes 4 or 8.
hipError_t launch_pull_probe(int elem_bytes, const void* src, void* dst, int plane, int pitch,
                             int begin, int end, int G, hipStream_t st);

}  // namespace lbmdev
