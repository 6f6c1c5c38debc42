/*
 * lbm_oracle — CPU restatement of the reference serial D2Q9 solver.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * nothing in the product (parallellbmcranfield_amd/, include/, the CLI) links,
 * loads or calls it.
 *
 * It restates mate-szoke/ParallelLbmCranfield serial/ (C, FP64) operation by
 * operation: the same 528-byte array-of-structs cell, the same six sweeps per
 * step and the same floating-point evaluation order, so that with
 * -ffp-contract=off its results are bit-identical to the reference program.
 * Two deliberate differences from the reference, both of which change no
 * result:
 *   - CellIni's per-cell linear searches (serial/CellFunctions.c:124-131,
 *     162-178) become one dense (i,j) table and one in-order pass over the
 *     connector lines, so set-up is O(N) instead of O(N^2);
 *   - nothing calls exit(): divergence is a return code.
 * Behaviour the reference does not have (rest start, textbook MRT equilibria,
 * a TRT magic parameter, full-field and velocity residuals, the parabolic
 * inlet extent of cuda_step3, first-order outlet) sits behind or_params
 * switches whose zero value is the reference behaviour.
 *
 * Pinning: every function here is checked by tests/test_oracle.py against
 * vectors produced by the reference program itself (oracle/_ref, built from
 * /root/reference/serial by oracle/Makefile; fixtures in tests/golden/) or,
 * for the non-reference switches, against closed forms and invariants.
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Same members, types and order as struct CellProps
 * (serial/include/ShellFunctions.h:14-37); sizeof == 528. */
typedef struct or_cell {
  int    fluid;
  int    corner;
  int    stream[9];
  int    id;
  int    boundary;
  int    boundary_id;
  int    bc[9];
  double q[9];
  double x, y;
  double u, v, rho;
  double uo, vo;
  double drag, lift;
  double f[9];
  double feq[9];
  double metaf[9];
  double fnb[9];
} or_cell;

/* The 16 positional ini scalars (serial/FileIO.c:263-289) plus switches for
 * behaviour outside the reference.  Every switch's zero value reproduces the
 * reference. */
typedef struct or_params {
  float  uavg, vavg, rho_ini, viscosity;
  int    inlet_profile;     /* 1 parabolic, 2 uniform, 3 pulsatile (no-op) */
  int    collision;         /* 1 BGKW, 2 TRT, 3 MRT */
  int    curved;            /* 1 curved (Bouzidi) walls, 2 half-way bounce-back */
  int    outlet_profile;    /* 1 velocity outlet, 2 open 2nd order, 3 open 1st order,
                               4 pressure outlet (Zou-He, rho = outlet_rho) */
  int    iterations;
  int    autosave_every, autosave_after, postproc, drag_lift;
  float  crit_v, crit_rho;
  int    read_former;
  /* ---- switches (0 == reference) ---- */
  int    init_rest;         /* 1: F = w*rho_ini, U = V = 0 at start */
  int    trt_fp64;          /* 1: TRT parts kept in FP64 (no binary32 rounding) */
  double trt_lambda;        /* >0: omega_minus from the magic parameter */
  int    mrt_textbook;      /* 1: m_eq = rho(-2+3|u|^2), rho(1-3|u|^2) */
  int    mrt_custom_s;      /* 1: use mrt_s[] as relaxation rates */
  double mrt_s[9];
  int    resid_kind;        /* 0: reference loop, 1: full field, 2: velocity, 3: none */
  int    resid_every;       /* full-field / velocity residual: K steps between checks */
  int    outlet1_lid;       /* reserved, keep 0 */
  double outlet_rho;        /* outlet_profile 4: outlet density (0 -> rho_ini) */
} or_params;

typedef struct or_solver or_solver;

/* Case construction ------------------------------------------------------ */
/* nodes: num_nodes x 5 floats (i j x y fluid), conn: num_conn x 7 floats
 * (i j k type x y boundary_id), as ReadNodes/ReadBCconn store them. */
or_solver* or_create(const float* nodes, int num_nodes, const float* conn,
                     int num_conn, const or_params* p);
/* Reads the two mesh files with the reference's fscanf scheme. */
or_solver* or_create_from_files(const char* node_file, const char* bc_file,
                                const or_params* p);
int  or_read_ini(const char* path, or_params* p);
void or_destroy(or_solver* s);

/* Runs up to nsteps iterations of the main loop (serial/Iterate.c:263-332,
 * residual, stop rule and iteration counter included; no file output).
 * Returns 0, 1 when the loop's convergence test stopped it, 2 on divergence. */
int  or_step(or_solver* s, int nsteps);

/* Accessors ------------------------------------------------------------- */
void or_dims(const or_solver* s, int* n, int* m, double* delta,
             int* num_inlet, double* min_inlet_y, double* max_inlet_y);
int  or_iter(const or_solver* s);
void or_residuals(const or_solver* s, double out[4]);
void or_omegas(const or_solver* s, double out[2]);
or_cell* or_cells(or_solver* s);          /* n*m cells, idx = j*n + i */
int  or_cell_size(void);                   /* sizeof(or_cell) */
/* Copies macros (n*m each) and distributions (n*m*9, cell-major). */
void or_get_macro(const or_solver* s, double* rho, double* u, double* v);
void or_get_f(const or_solver* s, double* f, double* metaf);

/* Momentum-exchange force on the wall links (BC_ID == 1) of the fluid cells
 * whose BoundaryID is bid, from the current F (post-boundary) and METAF:
 * sum of c_k (METAF_k + F_opp(k)).  Not in the reference (its drag/lift is a
 * placeholder, serial/CellFunctions.c:489-493). */
void or_force(const or_solver* s, int bid, double out[2]);

/* Writers with the reference's formats (serial/FileIO.c:298-427). */
int  or_write_results(const or_solver* s, const char* path, int final_layout,
                      int final_row0_bug);

/* Stand-alone kernels (exported for closed-form tests). */
void or_d2q9(double w[9], int cx[9], int cy[9], int opp[9], int c[9], int n);
void or_mrt_init(double tm[9][9], double stmiv[9][9], const double s[9]);

#ifdef __cplusplus
}
#endif
#endif
