/*
 * lbm.h — C ABI of the MI355X D2Q9 lattice-Boltzmann engine.
 *
 * Drop-in boundary for the serial solver path of mate-szoke/ParallelLbmCranfield.
 * The reference has no library interface: its path sits behind the process
 * contract of serial/main.c (SetUpData.ini, Mesh/D2node.dat,
 * Mesh/BCconnectors.dat in; Results/ out) and the function-level interface
 * of serial/include/Iterate.h:8-27 (Iteration plus CollisionStep,
 * StreamingStep, HandleBoundariesStep, UpdateMacroscopicStep) and
 * serial/include/ComputeResiduals.h:6-7.  Each entry point below names what it
 * replaces.
 *
 * Conventions: plain C types, caller-owned host buffers, library-owned device
 * memory.  Host arrays of per-cell values use the reference's storage order,
 * idx = j*n + i (serial/Iterate.c:499, CellFunctions.c:128), n cells in x and
 * m in y; per-direction values are cell-major, [idx*9 + k], with the
 * reference's direction numbering (serial/CellFunctions.c:281-289).
 * Every call returns an lbm_status; nothing calls exit().  A solver handle is
 * not thread-safe; different handles are independent (each owns a HIP stream).
 */
#ifndef LBM_H
#define LBM_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_ABI_VERSION 1

typedef enum lbm_status {
  LBM_OK = 0,
  LBM_EIO = 1,       /* file cannot be opened / written (FileIO.c:48-52 printed and returned NULL) */
  LBM_EPARSE = 2,    /* malformed ini or mesh */
  LBM_DIVERGED = 3,  /* residual became NaN (ComputeResiduals.c:47-51 printed DIVERGENCE! and exit(1)) */
  LBM_EHIP = 4,      /* HIP runtime error (no device, launch failure, ...) */
  LBM_EINVAL = 5,    /* invalid argument or option combination */
  LBM_ENOMEM = 6
} lbm_status;

/* Collision models, boundary switches: the reference's ini codes
 * (serial/Iterate.c:127-156, cuda/cuda_step3/include/Arguments.h:12-48). */
enum { LBM_BGKW = 1, LBM_TRT = 2, LBM_MRT = 3 };
enum { LBM_INLET_PARABOLIC = 1, LBM_INLET_UNIFORM = 2, LBM_INLET_PULSATILE = 3 };
enum { LBM_CURVED_ON = 1, LBM_CURVED_OFF = 2 };
enum { LBM_OUTLET_VELOCITY = 1, LBM_OUTLET_OPEN2 = 2, LBM_OUTLET_OPEN1 = 3,
       LBM_OUTLET_PRESSURE = 4 /* Zou-He rho = outlet_rho; not in the reference */ };
enum { LBM_FORMAT_CSV = 1, LBM_FORMAT_TECPLOT = 2 };

/* Storage / arithmetic of the distributions. */
enum {
  LBM_PREC_F64 = 0, /* FP64 f, FP64 arithmetic in the reference's evaluation order */
  LBM_PREC_F32 = 1, /* FP32 storage of f - w_k*rho0, FP32 arithmetic */
  LBM_PREC_F16 = 2, /* FP16 storage of (f - w_k*rho0)*scale, FP32 arithmetic */
  LBM_PREC_F16_SR = 3 /* the same with stochastic rounding of the stored values: the FP16
                         neighbour above with probability (distance to the one below) /
                         (their spacing), from a hash of the value's place and the step
                         number (reproducible run to run); where per-step changes fall
                         below half an FP16 spacing (viscous, slowly changing flows) the
                         rounding errors then add up as a random walk instead of
                         coherently */
};
/* Initial state. */
enum {
  LBM_INIT_REFERENCE = 0, /* F = 0, U = Uo, V = 0, Rho = rho_ini (CellFunctions.c:148-150,267) */
  LBM_INIT_REST = 1,      /* f = w*rho_ini, u = 0 */
  LBM_INIT_FIELD = 2,     /* f = feq(rho, u, v) of a caller field (lbm_solver_create_field) */
  LBM_INIT_FIELD_NEQ = 3  /* the same plus the Chapman-Enskog non-equilibrium part
                             -3 tau w_k rho Q_k:grad(u) from the field's velocity gradients
                             (central differences; half-way walls at rest, inlet links at Uo) */
};
/* Residual definitions. */
enum {
  LBM_RESID_REFERENCE = 0, /* ComputeResiduals.c:6-51 as written (first m cells, n times each), every step */
  LBM_RESID_FULL = 1,      /* L2(F - METAF) over every cell (cuda_step3 ComputeResiduals.cu:12-47) */
  LBM_RESID_VELOCITY = 2,  /* ||u^t - u^{t-K}|| / (K ||u^t||) and the same for rho, every K steps */
  LBM_RESID_NONE = 3
};

/* The 16 positional values of SetUpData.ini (serial/FileIO.c:263-289). */
typedef struct lbm_ini {
  float uavg, vavg, rho_ini, viscosity;
  int   inlet_profile, collision_model, curved_boundaries, outlet_profile;
  int   iterations, autosave_every, autosave_after, postproc_prog, calculate_drag_lift;
  float convergence_crit_veloc, convergence_crit_rho;
  int   read_former_data;
} lbm_ini;

/* Solver options.  lbm_options_from_ini() fills the reference behaviour. */
typedef struct lbm_options {
  int    collision;        /* LBM_BGKW / LBM_TRT / LBM_MRT */
  int    curved;           /* LBM_CURVED_ON / LBM_CURVED_OFF */
  int    inlet_profile;    /* LBM_INLET_* */
  int    outlet_profile;   /* LBM_OUTLET_* */
  int    drag_lift;        /* boundary id whose cells carry drag/lift, 0 = none */
  float  uavg, vavg, rho_ini, viscosity;
  int    precision;        /* LBM_PREC_* */
  int    init;             /* LBM_INIT_* */
  int    trt_fp64_parts;   /* 0: TRT parts rounded to binary32 as CellFunctions.c:383-393 (F64 only) */
  double trt_lambda;       /* > 0: omega- from this magic parameter; 0: 8(2-w)/(8-w) (Iterate.c:180) */
  int    mrt_textbook_meq; /* 1: m_eq = rho(-2+3u^2), rho(1-3u^2); 0: the rho^2 forms of CellFunctions.c:412-413 */
  int    mrt_custom_rates; /* 1: mrt_rates[] replace (1,1.4,1.4,1,1.2,1,1.2,w,w) (CellFunctions.c:60-68) */
  double mrt_rates[9];
  int    resid_kind;       /* LBM_RESID_* */
  int    resid_every;      /* K for LBM_RESID_FULL / LBM_RESID_VELOCITY (>= 1) */
  int    use_stop_rule;    /* 1: stop when !(r0 > crit_v && r1 > crit_rho) (Iterate.c:263) */
  float  crit_v, crit_rho;
  int    max_iterations;   /* stop rule's iteration cap (Iterate.c:263 "iter < Iterations") */
  float  f16_scale;        /* LBM_PREC_F16 fixed scale; 0 = 4096 */
  int    block_size;       /* bulk-kernel workgroup size: 0 (tuned default), 256, 512 or 1024 */
  int    tune_flags;       /* LBM_TUNE_* launch-shape switches; 0 = tuned default */
  int    xcd_group;        /* runs of this many adjacent bulk tiles per XCD (1 = hardware order); 0 = tuned default */
  double outlet_rho;       /* LBM_OUTLET_PRESSURE density; 0 = rho_ini */
  int    force_id;         /* momentum-exchange force on the wall links of this boundary id (0 = off) */
  int    force_every;      /* force sampled every K steps (lbm_solver_force_history) */
  int    two_step_rows;    /* output rows per two-step tile (> 0 also turns two-step launches on
                              at any lattice size); 0 = tuned default */
  int    three_step_rows;  /* output rows per three-step tile (> 0 also turns three-step
                              launches on at any lattice size); 0 = tuned default */
} lbm_options;

/* Launch-shape switches (results are identical; only speed changes).
 * LBM_TUNE_XCD_REMAP: give each XCD a contiguous run of bulk tiles (measured
 * slower than the hardware's round-robin order on the 4097^2 lattice).
 * LBM_TUNE_SPLIT_BOUNDARY: boundary-list nodes in their own launch after the
 * bulk pass instead of as the first workgroups of one fused launch.
 * LBM_TUNE_GRAPH: replay captured HIP graphs for runs of plain steps (two
 * single steps, or a two- or three-step launch with its band launches)
 * instead of launching directly.  Off by default: direct launches measured
 * faster at every size (per step, direct against graph: 129^2 4.2 against
 * 7.7 us, 1025^2 FP16 12.8 against 15.3 us, 4097^2 FP32 116.9 against
 * 118.1 us) with identical results.
 * LBM_TUNE_NO_GRAPH: direct launches even where LBM_TUNE_GRAPH is set (the
 * default without it).
 * LBM_TUNE_NO_PAIRS: FP16 storage without the paired-cell plain-step kernel
 * (one cell per thread, 16-bit loads and stores).
 * LBM_TUNE_NO_TILE_WALLS: every boundary node in the boundary list (by
 * default nodes whose only rule is straight-wall bounce-back run in their
 * bulk tile).
 * LBM_TUNE_NO_TWO_STEP: every step its own launch.  By default two plain
 * steps in a row run as one pass over the state where that pays (FP32
 * storage on lattices of at least 2^20 cells, FP64 on lattices of at least
 * 2^22 cells, two thirds of the nodes away from the boundary-list nodes; also
 * on packed lattices of instances): interior and in-tile wall
 * nodes read METAF_t once and write METAF_{t+2} once, the nodes next to
 * boundary-list nodes take two single steps.  two_step_rows > 0 turns them
 * on for any lattice and storage but stochastic rounding.
 * LBM_TUNE_NO_THREE_STEP: no three-step launches.  By default three plain
 * steps in a row run as one pass over the state (METAF_t read once,
 * METAF_{t+3} written once) with FP32 storage on lattices of at least 2^22
 * cells where two-step launches are on (not on packed lattices of
 * instances); the nodes within two nodes of
 * boundary-list nodes take three single steps.  three_step_rows > 0 turns
 * them on for any single lattice and storage but stochastic rounding.  Runs
 * of plain steps take three-step launches first, then a two-step or single
 * one for the rest. */
enum { LBM_TUNE_XCD_REMAP = 1, LBM_TUNE_SPLIT_BOUNDARY = 2, LBM_TUNE_NO_GRAPH = 4,
       LBM_TUNE_NO_PAIRS = 8, LBM_TUNE_NO_TILE_WALLS = 16, LBM_TUNE_NO_TWO_STEP = 32,
       LBM_TUNE_NO_THREE_STEP = 64, LBM_TUNE_GRAPH = 128 };

typedef struct lbm_case lbm_case;
typedef struct lbm_solver lbm_solver;

typedef struct lbm_case_info {
  int    n, m;             /* CompDataNode (FileIO.c:193-194) */
  int    num_nodes, num_conn;
  int    num_fluid, num_boundary;
  int    num_inlet;        /* NumInletNodes (FileIO.c:235-250) */
  double delta;            /* grid spacing (FileIO.c:196-208) */
  double min_inlet_y, max_inlet_y;
} lbm_case_info;

typedef struct lbm_timing {
  double bulk_ms, boundary_ms, outlet_ms, residual_ms, observe_ms, init_ms;
  long   steps;
} lbm_timing;

/* ---------------------------------------------------------------- status */
const char* lbm_last_error(void);
int lbm_abi_version(void);

/* ---------------------------------------------------------------- input  */
/* Replaces ReadIniData (serial/FileIO.c:263-289). */
int lbm_read_ini(const char* path, lbm_ini* out);
/* Reference behaviour for an ini: every switch at its reference value. */
void lbm_options_from_ini(const lbm_ini* ini, lbm_options* out);

/* Replaces ReadNodes + ReadBCconn + CompDataNode + CompDataConn
 * (serial/FileIO.c:14-254, called at serial/Iterate.c:98-110). */
int lbm_case_load(const char* node_file, const char* bc_file, lbm_case** out);
/* Same from arrays already in memory: nodes[num_nodes*5] = (i j x y fluid),
 * conn[num_conn*7] = (i j k type x y boundary_id), as ReadNodes / ReadBCconn
 * store them (all float). */
int lbm_case_from_arrays(const float* nodes, int num_nodes, const float* conn, int num_conn,
                         lbm_case** out);
/* Mesh generator: the shipped cases' lid-driven cavity with N fluid nodes per
 * side (Delta = 1/N, n = N+2, m = N+3; as cases/coarse/Mesh). lid_type 2
 * (inlet, the shipped lid) or 1 (plain wall). */
int lbm_case_cavity(int N, int lid_type, lbm_case** out);
/* Mesh generator: channel NX x NY fluid nodes, west inlet (type 2),
 * east outlet (type 3), walls top/bottom; with an optional circular obstacle
 * (radius r > 0 at (cx, cy) in node units) whose links carry the exact cut
 * distance q and boundary id 5. */
int lbm_case_channel(int NX, int NY, double cx, double cy, double r, lbm_case** out);
/* Mesh generator for batched sweeps: an NX x NY lid-driven cavity (n = NX+2,
 * m = NY+3, Delta = 1/max(NX, NY)) with nrect solid rectangles given in
 * fluid-node units as rects[4r..4r+3] = (i0, j0, i1, j1), inclusive; the
 * rectangles' links are half-way bounce-back walls of boundary id 5.  No
 * counterpart in the reference (generalises cases/coarse/Mesh to any
 * aspect ratio and obstacles). */
int lbm_case_cavity_ex(int NX, int NY, int lid_type, const int* rects, int nrect,
                       lbm_case** out);
/* Packs `count` cases side by side (shelf packing, tallest first) into one
 * lattice for a batched solver.  Instances keep their own coordinates and
 * boundary links; rects (may be NULL) receives (x0, y0, n, m) per instance
 * in the packed lattice.  No counterpart in the reference. */
int lbm_case_pack(const lbm_case* const* cases, int count, lbm_case** out, int* rects);
/* Number of instances of a packed case (0 for an ordinary one). */
int lbm_case_instances(const lbm_case* c);
/* Writes the case as Mesh/D2node.dat + Mesh/BCconnectors.dat text. */
int lbm_case_write(const lbm_case* c, const char* node_file, const char* bc_file);
int lbm_case_info_get(const lbm_case* c, lbm_case_info* out);
/* Per-cell set-up exactly as CellIni (serial/CellFunctions.c:99-270) leaves it.
 * Any pointer may be NULL.  stream/bc: n*m*9 ints; q: n*m*9 doubles. */
int lbm_case_cells(const lbm_case* c, int* fluid, int* corner, int* stream, int* id,
                   int* boundary, int* boundary_id, int* bc, double* q, double* x, double* y);
void lbm_case_free(lbm_case* c);

/* ---------------------------------------------------------------- solver */
/* Replaces Iteration()'s set-up half (serial/Iterate.c:168-228): device state,
 * node classes, boundary-link list, first collision. */
int lbm_solver_create(const lbm_case* c, const lbm_options* o, lbm_solver** out);
/* As lbm_solver_create with LBM_INIT_FIELD: rho/u/v are n*m host arrays. */
int lbm_solver_create_field(const lbm_case* c, const lbm_options* o, const double* rho,
                            const double* u, const double* v, lbm_solver** out);
/* Per-instance parameters of a batched solver. */
typedef struct lbm_instance {
  float viscosity; /* lattice viscosity: omega = 1/(3 nu + 1/2) */
  float uavg;      /* lid speed of the instance (inlet profile 2) */
} lbm_instance;
/* Batched solver over a packed case (lbm_case_pack): every instance advances
 * in the same step launches with its own omega and lid speed, FP32 or FP16
 * storage.  Every resid_every steps each instance's velocity residual
 * ||u(t) - u(t-K)|| / (K ||u(t)||) is computed; an instance below crit_v stops
 * advancing.  With use_stop_rule the batch stops when all have.  No
 * counterpart in the reference (its solver runs one lattice per process). */
int lbm_solver_create_batched(const lbm_case* c, const lbm_options* o, const lbm_instance* inst,
                              int count, lbm_solver** out);
/* Same, every instance starting from f = feq(rho, u, v) of a caller field
 * (n*m host arrays over the packed lattice, as lbm_solver_create_field). */
int lbm_solver_create_batched_field(const lbm_case* c, const lbm_options* o,
                                    const lbm_instance* inst, int count, const double* rho,
                                    const double* u, const double* v, lbm_solver** out);
/* Per-instance state of a batched solver: conv_iter (iteration at which the
 * instance converged, -1 still running, -2 diverged), conv_seconds (wall time
 * from the first step to the detection), resid (last residual).  Any pointer
 * may be NULL; the first min(count, instances) entries are written. */
int lbm_batch_status(lbm_solver* s, int count, int* conv_iter, double* conv_seconds,
                     double* resid);
void lbm_solver_destroy(lbm_solver* s);

/* Replaces one or more passes of the main loop body (serial/Iterate.c:263-332:
 * CollisionStep, UpdateF, StreamingStep, HandleBoundariesStep,
 * UpdateMacroscopicStep, ComputeResiduals).  Runs up to nsteps iterations;
 * with use_stop_rule it stops early when the loop condition fails.
 * *done (may be NULL) receives the number of iterations run. */
int lbm_solver_step(lbm_solver* s, int nsteps, int* done);
/* Same, timing the step kernel with HIP events on the solver's stream around
 * every `sample_every`-th launch (1 = every launch).  total_ms: stream time of
 * the nsteps; kernel_ms: average duration of one sampled step-kernel launch.
 * With multi-step launches on (lbm_solver_two_step_info,
 * lbm_solver_three_step_info) the sampled kernel is the three-step kernel
 * (one launch per three steps) where three-step launches are on, else the
 * two-step kernel (one launch per two steps); launches of fewer steps are
 * not sampled. */
int lbm_solver_step_timed(lbm_solver* s, int nsteps, int sample_every, double* total_ms,
                          double* kernel_ms);
int lbm_solver_sync(lbm_solver* s);
int lbm_solver_iter(const lbm_solver* s);
/* Residuals[0..3] after the last iteration (Iterate.c:300, :259-260 before any). */
int lbm_solver_residuals(lbm_solver* s, double out[4]);
/* Residual rows recorded since the last call (for residuals.dat): iters[k],
 * vals[4k..4k+3].  *count receives how many were written (<= cap). */
int lbm_solver_residual_history(lbm_solver* s, int cap, int* iters, double* vals, int* count);
/* Momentum-exchange force (Fx, Fy) on the wall links of boundary id
 * force_id, sampled every force_every steps, since the last call: iters[k],
 * fxy[2k..2k+1].  Replaces the reference's drag/lift placeholder
 * (serial/CellFunctions.c:489-493) for bodies; no counterpart there. */
int lbm_solver_force_history(lbm_solver* s, int cap, int* iters, double* fxy, int* count);
/* U, V, Rho of every cell after the last iteration (n*m each; may be NULL). */
int lbm_solver_get_macro(lbm_solver* s, double* rho, double* u, double* v);
/* F (post-boundary distributions) and METAF (post-collision) of every cell
 * after the last iteration, absolute values, n*m*9 each (may be NULL). */
int lbm_solver_get_f(lbm_solver* s, double* f, double* metaf);
/* Replaces WriteResults / WriteFinalResults (serial/FileIO.c:298-427).
 * final_layout adds f0..f8; row0_bug reproduces FileIO.c:363-386. */
int lbm_solver_write_results(lbm_solver* s, const char* path, int format, int final_layout,
                             int row0_bug);
int lbm_solver_timing(const lbm_solver* s, lbm_timing* out);
/* Replaces ComputeResiduals (serial/ComputeResiduals.c:6-51, declared at
 * serial/include/ComputeResiduals.h:6-7) for a lattice state held in device
 * memory, reduced in FP64.  f, metaf: DEVICE arrays of n*m*9 doubles,
 * cell-major as above (Cells[idx].F[k], .METAF[k]); boundary_id (n*m ints),
 * drag and lift (n*m doubles): DEVICE arrays, any may be NULL (no drag /
 * lift terms); drag_lift: the ComputeDragLift id whose cells' DragF / LiftF
 * are summed; out: DEVICE array of 4 doubles receiving Residuals[0..3]
 * (sqrt(L2), sqrt(L2 / (m n)), drag, lift).  kind LBM_RESID_REFERENCE sums
 * as the reference's loop is written (the first m cells, n times each, in
 * its order: the reference's values bit for bit); LBM_RESID_FULL sums every
 * cell once (fixed-order partial sums per 8192 values, then a fixed-order
 * total: the same result on every run).  Work is queued on `stream` (a
 * hipStream_t; NULL = the null stream); out is valid once the stream gets
 * there.  A NaN in out[0] is the reference's "DIVERGENCE!" (it printed and
 * called exit(1); ComputeResiduals.c:47-51).  LBM_RESID_FULL holds a
 * stream-ordered allocation of 32 bytes per 8192 values for the call. */
int lbm_compute_residuals(const double* f, const double* metaf, const int* boundary_id,
                          const double* drag, const double* lift, int n, int m, int drag_lift,
                          int kind, double* out, void* stream);
/* lbm_compute_residuals that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The partial sums of LBM_RESID_FULL go to a fixed 1 MiB array
 * in the library's device image (one per GPU, held until the process ends)
 * instead of a per-call allocation, so LBM_RESID_FULL takes at most
 * LBM_NOALLOC_MAX_CELLS cells (n*m) and refuses larger lattices with
 * LBM_EINVAL.  Same arguments, kernels and results as lbm_compute_residuals,
 * bit for bit.  All calls on one GPU share its array: calls that overlap on
 * one GPU (other streams, other host threads) race and must be serialised
 * by the caller. */
#define LBM_NOALLOC_MAX_CELLS 29826161L
int lbm_compute_residuals_noalloc(const double* f, const double* metaf, const int* boundary_id,
                                  const double* drag, const double* lift, int n, int m,
                                  int drag_lift, int kind, double* out, void* stream);
/* Bytes of device memory held by the solver. */
size_t lbm_solver_device_bytes(const lbm_solver* s);
/* Binary checkpoint of the solver state (both distribution buffers, the
 * iteration, residuals, velocity-residual snapshot): the restart the
 * reference declares but never implements (ReadFormerData,
 * serial/FileIO.c:287; ReadResults, serial/include/FilesIO.h:25).  A file
 * loads only into a solver with the same lattice, precision, collision model,
 * rho_ini, FP16 scale and residual kind (and interval, for the velocity
 * residual); any other file, or one of the wrong length, is refused
 * (LBM_EINVAL / LBM_EPARSE) before the solver's state changes.  Should a
 * read fail part-way, the solver refuses to step until a load succeeds. */
int lbm_solver_save(lbm_solver* s, const char* path);
int lbm_solver_load(lbm_solver* s, const char* path);
/* Kernel layout (for profiling): bit 0 set if interior and boundary-list
 * nodes share one launch per step (else separate launches); bit 1 set if
 * plain steps run the paired-cell FP16 kernel; bit 2 set if pairs of plain
 * steps run as two-step launches; bit 3 set if runs of three plain steps run
 * as three-step launches. */
int lbm_solver_fused(const lbm_solver* s);
/* Two-step launches of this solver: *cells = nodes the two-step kernel
 * advances by two steps per launch (0: two-step launches off), *tiles = its
 * tiles (one workgroup each), *rows = output rows per tile, *band = nodes
 * that take two single steps instead.  Any pointer may be NULL.  No
 * counterpart in the reference. */
int lbm_solver_two_step_info(const lbm_solver* s, long long* cells, int* tiles, int* rows,
                             long long* band);
/* Three-step launches, as lbm_solver_two_step_info: nodes the three-step
 * kernel advances by three steps per launch (0: off), its tiles, output rows
 * per tile, nodes that take three single steps instead.  No counterpart in
 * the reference. */
int lbm_solver_three_step_info(const lbm_solver* s, long long* cells, int* tiles, int* rows,
                               long long* band);

/* Bandwidth of a plain device-to-device copy kernel over a buffer of `bytes`
 * (read + write counted), best of `reps` launches of each of several launch
 * shapes (1 to 8 independent 16-byte loads in flight per thread): the
 * practical HBM ceiling the step kernels are compared with.  No counterpart
 * in the reference. */
int lbm_device_copy_gbs(size_t bytes, int reps, double* gbs);
/* Bandwidth of the step kernel's memory traffic without its arithmetic, on
 * the padded layout of an n x m lattice: every cell of the lattice's rows
 * pulls its nine values from its upstream neighbours and stores them (same
 * planes, offsets, tiles and XCD order as the step), elem_bytes 4 (FP32) or
 * 8 (FP64).  *gbs: bytes moved (2 * 9 * elem_bytes per cell) over the average
 * of `reps` launches after a warm-up; *us (may be NULL): that average in
 * microseconds.  The practical ceiling of the step's access pattern.  No
 * counterpart in the reference. */
int lbm_device_stream_probe(int n, int m, int elem_bytes, int reps, double* gbs, double* us);

/* ---------------------------------------------------------------- driver */
/* The whole reference program, serial/main.c:20-93 + Iteration
 * (serial/Iterate.c:17-410): reads ini_file (if NULL, "SetUpData.ini") and the
 * two mesh files, writes Results/ under out_dir with the reference's names
 * and formats.  opt_overrides may be NULL (reference behaviour).  Returns
 * LBM_DIVERGED where the reference printed DIVERGENCE! and exited 1. */
int lbm_run(const char* ini_file, const char* node_file, const char* bc_file,
            const char* out_dir, const lbm_options* opt_overrides, int verbose);
/* lbm_run that first restores restart_file (if not NULL) and finally saves
 * checkpoint_file (if not NULL) with lbm_solver_save. */
int lbm_run_checkpointed(const char* ini_file, const char* node_file, const char* bc_file,
                         const char* out_dir, const lbm_options* opt_overrides,
                         const char* restart_file, const char* checkpoint_file, int verbose);

/* The reference's Iteration() signature (serial/include/Iterate.h:21-27),
 * writing Results/ in the current directory. */
int lbm_iteration(const char* NodeDataFile, const char* BCconnectorDataFile, float Uavg,
                  float Vavg, float rho_ini, float Viscosity, int InletProfile,
                  int CollisionModel, int CurvedBoundaries, int OutletProfile, int Iterations,
                  int AutosaveAfter, int AutosaveEvery, int postproc_prog,
                  int CalculateDragLift, float ConvergenceCritVeloc, float ConvergenceCritRho,
                  int ReadFormerData);

/* ---------------------------------------------------------------- warm start */
/* The warm-start network: a U-Net of 3x3 convolutions (BF16 weights and
 * activations, FP32 accumulation on the matrix cores, hand-written backward
 * and Adam) mapping per-pixel cavity descriptors (fluid mask, normalised
 * log Re, lid speed, wall distances, ...) on an H x W grid to the steady
 * (u, v, rho) there.  `levels` resolutions with base << l channels
 * (l = 0 .. levels-1); base a multiple of 32, base << (levels-1) <= 2048.  Inputs are
 * [sample][H][W][in_channels] FP32 (in_channels <= 32), channel 0 the fluid
 * mask (1/0) that weighs the loss; targets and predictions [sample][H][W][3].
 * No counterpart in the reference. */
typedef struct lbm_net lbm_net;
/* dilation d >= 1: the two bottleneck convolutions use 3x3 stencils dilated
 * by d and 2d (1: plain), widening the receptive field without parameters. */
int lbm_net_create(int H, int W, int batch, int base, int levels, int in_channels,
                   int dilation, unsigned seed, lbm_net** out);
void lbm_net_destroy(lbm_net* n);
long lbm_net_num_params(const lbm_net* n);
/* FP32 master parameters in canonical order: per conv (encoder a/b per
 * level incl. the bottleneck, then decoder a/b from the coarsest level)
 * W[Cout][9][Cin] then bias[Cout] (Cin of the first conv padded to 32,
 * weights of the padding channels stay 0 in training);
 * then the 1x1 head W[3][base], bias[3].  set_params resets Adam. */
int lbm_net_get_params(lbm_net* n, float* params);
int lbm_net_set_params(lbm_net* n, const float* params);
/* Prediction for nb samples (any count; run in batches). */
int lbm_net_predict(lbm_net* n, const float* in, int nb, float* out);
/* Masked mean-square loss and its gradient w.r.t. every parameter
 * (canonical order) for nb <= batch samples; grads may be NULL. */
int lbm_net_gradients(lbm_net* n, const float* in, const float* target, int nb, double* loss,
                      float* grads);
/* Device-resident training set (count samples), then one Adam step
 * (b1 0.9, b2 0.999, eps 1e-8) on the samples idx[0..nb). */
int lbm_net_set_data(lbm_net* n, const float* in, const float* target, int count);
int lbm_net_train(lbm_net* n, const int* idx, int nb, float lr, double* loss);
/* Decoupled weight decay of the Adam step (AdamW): every parameter also
 * moves by -lr * wd * p.  0 (the default) is plain Adam; 0 <= wd <= 1. */
int lbm_net_set_weight_decay(lbm_net* n, float wd);
/* Adds weight * sum d^2 to the loss (same normalisation as the squared
 * error), d the central-difference divergence su dy0/dx + sv dy1/dy of the
 * predicted velocity channels (su, sv: their scales to physical units) at
 * pixels whose four neighbours are fluid.  0 (the default) turns it off. */
int lbm_net_set_div_weight(lbm_net* n, float weight, float su, float sv);
int lbm_net_save(lbm_net* n, const char* path);
int lbm_net_load(lbm_net* n, const char* path);

#ifdef __cplusplus
}
#endif
#endif /* LBM_H */
