/*
 * lbm_oracle.c — CPU restatement of the reference serial solver (FP64).
 *
 * TEST INFRASTRUCTURE ONLY (see lbm_oracle.h): the checker for the HIP path and
 * the timed CPU baseline; never part of the product.
 *
 * Each function names the reference lines it restates.  Floating-point
 * expressions keep the reference's operand order and association so that,
 * built with -ffp-contract=off, results are bit-identical to the reference
 * program (checked against oracle/_ref goldens in tests/test_oracle.py).
 */
#include "lbm_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

struct or_solver {
  or_params p;
  int       n, m;
  int       num_nodes, num_conn;
  float   (*nodes)[5];
  float   (*conn)[7];
  double    delta;
  int       num_inlet;
  double    min_inlet_y, max_inlet_y;
  double    w[9];
  int       cx[9], cy[9], opp[9], c[9];
  double    omega, omega_a;
  double    tm[9][9], stmiv[9][9];
  or_cell*  cells;
  double    resid[4];
  int       iter;
  /* velocity-residual snapshot (resid_kind 2 only) */
  double   *u_prev, *v_prev, *rho_prev;
};

/* ------------------------------------------------------------------------ */
/* Lattice constants: serial/CellFunctions.c:277-342 (D2Q9Vars).             */
/* Direction order 0 rest, 1 E, 2 N, 3 W, 4 S, 5 NE, 6 NW, 7 SW, 8 SE;        */
/* c[k] is the storage offset of the upstream (pull) neighbour.              */
/* ------------------------------------------------------------------------ */
void or_d2q9(double w[9], int cx[9], int cy[9], int opp[9], int c[9], int n)
{
  static const int ex[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
  static const int ey[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
  static const int op[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
  for (int k = 0; k < 9; k++) {
    w[k]   = (k == 0) ? 4. / 9. : (k < 5 ? 1. / 9. : 1. / 36.);
    cx[k]  = ex[k];
    cy[k]  = ey[k];
    opp[k] = op[k];
    c[k]   = -ex[k] - ey[k] * n;
  }
}

/* ------------------------------------------------------------------------ */
/* MRT matrices: serial/CellFunctions.c:11-92 (MRTInitializer).              */
/* tm is the Lallemand-Luo moment matrix M (rows rho, e, eps, jx, qx, jy,    */
/* qy, pxx, pxy); tminv = M^-1 written with a1 = 1/36; stmiv = M^-1 * S.     */
/* ------------------------------------------------------------------------ */
void or_mrt_init(double tm[9][9], double stmiv[9][9], const double s[9])
{
  static const double M[9][9] = {
      {1, 1, 1, 1, 1, 1, 1, 1, 1},
      {-4, -1, -1, -1, -1, 2, 2, 2, 2},
      {4, -2, -2, -2, -2, 1, 1, 1, 1},
      {0, 1, 0, -1, 0, 1, -1, -1, 1},
      {0, -2, 0, 2, 0, 1, -1, -1, 1},
      {0, 0, 1, 0, -1, 1, 1, -1, -1},
      {0, 0, -2, 0, 2, 1, 1, -1, -1},
      {0, 1, -1, 1, -1, 0, 0, 0, 0},
      {0, 0, 0, 0, 0, 1, -1, 1, -1}};
  /* M^-1 in units of a1 = 1/36; entries are multiplied by a1 in FP64 exactly
   * as the reference's initialiser list does (k*a1). */
  static const int Minv36[9][9] = {
      {4, -4, 4, 0, 0, 0, 0, 0, 0},
      {4, -1, -2, 6, -6, 0, 0, 9, 0},
      {4, -1, -2, 0, 0, 6, -6, -9, 0},
      {4, -1, -2, -6, 6, 0, 0, 9, 0},
      {4, -1, -2, 0, 0, -6, 6, -9, 0},
      {4, 2, 1, 6, 3, 6, 3, 0, 9},
      {4, 2, 1, -6, -3, 6, 3, 0, -9},
      {4, 2, 1, -6, -3, -6, -3, 0, 9},
      {4, 2, 1, 6, 3, -6, -3, 0, -9}};
  const double a1 = 1. / 36.;
  for (int i = 0; i < 9; i++)
    for (int j = 0; j < 9; j++) {
      int k = Minv36[i][j];
      /* The reference writes "-a1" and "a1" without a factor; k*a1 with
       * k = +-1 is the same value. */
      double minv = (k == 0) ? 0.0 : (double)k * a1;
      tm[i][j]    = M[i][j];
      stmiv[i][j] = minv * s[j];
    }
}

/* ------------------------------------------------------------------------ */
/* Mesh-derived constants                                                   */
/* ------------------------------------------------------------------------ */

/* serial/FileIO.c:183-209 (CompDataNode): n, m from the last node line,
 * Delta = |x(1,0) - x(0,0)| (the last matching line wins). */
static void comp_data_node(or_solver* s)
{
  const float* last = s->nodes[s->num_nodes - 1];
  s->n = (int)(last[0] + 1);
  s->m = (int)(last[1] + 1);
  double p1 = 0.0, p2 = 0.0;
  for (int i = 0; i < s->num_nodes; i++) {
    if (s->nodes[i][0] == 0 && s->nodes[i][1] == 0) p1 = s->nodes[i][2];
    if (s->nodes[i][0] == 1 && s->nodes[i][1] == 0) p2 = s->nodes[i][2];
  }
  double hi = p1 > p2 ? p1 : p2;
  double lo = p1 < p2 ? p1 : p2;
  s->delta = hi - lo;
}

/* serial/FileIO.c:218-254 (CompDataConn): the inlet extent is the min/max y
 * of the inlet lines (type 2) whose direction is 1..4, widened by Delta/2,
 * kept in FP64 as cuda/cuda_step3/FilesReading.cu:263-300 does.  The serial
 * build never delivers this value (FilesIO.h:15 declares MyReal pointers,
 * FileIO.c:218-219 defines float pointers, and its running min/max starts from
 * an unset value), so the extent follows the functions' evident intent: the
 * running min/max start at the first qualifying line. */
static void comp_data_conn(or_solver* s)
{
  double mx = 0.0, mn = 0.0;
  int seen = 0;
  s->num_inlet = 0;
  for (int i = 0; i < s->num_conn; i++) {
    const float* c = s->conn[i];
    if (c[3] == 2 && (c[2] == 1 || c[2] == 2 || c[2] == 3 || c[2] == 4)) {
      if (!seen || c[5] > mx) mx = c[5];
      if (!seen || c[5] < mn) mn = c[5];
      seen = 1;
      s->num_inlet++;
    }
  }
  s->max_inlet_y = mx + s->delta / 2;
  s->min_inlet_y = mn - s->delta / 2;
}

/* ------------------------------------------------------------------------ */
/* Cell initialisation: serial/CellFunctions.c:99-270 (CellIni) for all      */
/* cells, with the two linear searches replaced by a dense node table and    */
/* a stable bucket of connector lines per cell (same "last line wins").      */
/* ------------------------------------------------------------------------ */
static int cell_init(or_solver* s)
{
  const int n = s->n, m = s->m;
  const long nc = (long)n * m;
  s->cells = (or_cell*)calloc((size_t)nc, sizeof(or_cell));
  int* node_of = (int*)malloc(sizeof(int) * (size_t)nc);
  int* cstart  = (int*)calloc((size_t)nc + 1, sizeof(int));
  int* clist   = (int*)malloc(sizeof(int) * (size_t)(s->num_conn + 1));
  if (!s->cells || !node_of || !cstart || !clist) return -1;

  /* ID search (:123-131): the last node line with Nod[0]==A, Nod[1]==B. */
  for (long q = 0; q < nc; q++) node_of[q] = -1;
  for (int L = 0; L < s->num_nodes; L++) {
    float fi = s->nodes[L][0], fj = s->nodes[L][1];
    if (fi < 0 || fj < 0 || fi >= (float)n || fj >= (float)m) continue;
    int ii = (int)fi, jj = (int)fj;
    if ((float)ii != fi || (float)jj != fj) continue;
    node_of[(long)jj * n + ii] = L;
  }
  /* Connector lines per cell, in file order (:162-178 visits them so). */
  for (int L = 0; L < s->num_conn; L++) {
    int ci = (int)s->conn[L][0], cj = (int)s->conn[L][1];
    if (ci < 0 || cj < 0 || ci >= n || cj >= m) continue;
    cstart[(long)cj * n + ci + 1]++;
  }
  for (long q = 0; q < nc; q++) cstart[q + 1] += cstart[q];
  {
    int* fill = (int*)malloc(sizeof(int) * (size_t)nc);
    memcpy(fill, cstart, sizeof(int) * (size_t)nc);
    for (int L = 0; L < s->num_conn; L++) {
      int ci = (int)s->conn[L][0], cj = (int)s->conn[L][1];
      if (ci < 0 || cj < 0 || ci >= n || cj >= m) continue;
      clist[fill[(long)cj * n + ci]++] = L;
    }
    free(fill);
  }

  const double qlat[9] = {0., 1., 1., 1., 1., sqrt(2), sqrt(2), sqrt(2), sqrt(2)};
  const or_params* p = &s->p;
  for (int B = 0; B < m; B++)
    for (int A = 0; A < n; A++) {
      long idx = (long)B * n + A;
      or_cell* c = &s->cells[idx];
      if (node_of[idx] >= 0) c->id = node_of[idx];      /* else stays 0 */
      const float* nod = s->nodes[c->id];
      c->x     = nod[2];
      c->y     = nod[3];
      c->fluid = (int)nod[4];
      c->u = 0;
      c->v = 0;
      c->rho = p->rho_ini;
      c->boundary = 0;
      c->boundary_id = 0;
      for (int k = 0; k < 9; k++) {
        c->bc[k] = 0;
        c->q[k]  = 0.5;
      }
      for (int t = cstart[idx]; t < cstart[idx + 1]; t++) {
        const float* con = s->conn[clist[t]];
        for (int j = 1; j < 9; j++)
          if (con[2] == j) {
            c->bc[j]       = (int)con[3];
            c->boundary_id = (int)con[6];
            double dx = con[4] - c->x, dy = con[5] - c->y;
            c->q[j] = sqrt(dx * dx + dy * dy) / (s->delta * qlat[j]);
          }
      }
      /* Boundary / Corner (:180-203) */
      c->corner = 0;
      for (int j = 1; j < 9; j++) {
        if (c->bc[j] == 0) continue;
        if (c->boundary == 0) {
          c->boundary = c->bc[j];
        } else {
          if (c->boundary < c->bc[j]) { c->boundary = 1; c->corner = 1; }
          if (c->boundary > c->bc[j]) { c->boundary = 1; c->corner = 1; }
        }
      }
      /* corner diagonals become wall (:212-231) */
      if (c->corner == 1) {
        if (c->bc[1] != 0 && c->bc[2] != 0) c->bc[5] = 1;
        if (c->bc[1] != 0 && c->bc[4] != 0) c->bc[8] = 1;
        if (c->bc[2] != 0 && c->bc[3] != 0) c->bc[6] = 1;
        if (c->bc[3] != 0 && c->bc[4] != 0) c->bc[7] = 1;
      }
      /* stream everywhere except from outside the domain (:234-249) */
      for (int k = 0; k < 9; k++) c->stream[k] = 1;
      for (int k = 0; k < 9; k++)
        if (c->bc[k] != 0) c->stream[s->opp[k]] = 0;
      /* inlet velocity (:252-267) */
      switch (p->inlet_profile) {
        case 1: {
          double mn = s->min_inlet_y, mx = s->max_inlet_y;
          c->uo = 4 * 1.5 * p->uavg * (c->y - mn) * ((mx - mn) - (c->y - mn)) /
                  ((mx - mn) * (mx - mn));
          c->vo = p->vavg;
        } break;
        case 2:
          c->uo = p->uavg;
          c->vo = p->vavg;
          break;
        default: break; /* pulsatile: not implemented by the reference */
      }
      c->u = c->uo;
      if (p->init_rest && c->fluid == 1) {
        c->u = 0;
        c->v = 0;
        for (int k = 0; k < 9; k++) c->f[k] = s->w[k] * c->rho;
      }
    }
  free(node_of);
  free(cstart);
  free(clist);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Collision: serial/CellFunctions.c:352-438, dispatched per fluid cell as   */
/* serial/Iterate.c:420-467 does.                                            */
/* ------------------------------------------------------------------------ */
static void bgkw(or_solver* s, or_cell* c)
{
  const double om = s->omega;
  double T1 = c->u * c->u + c->v * c->v;
  for (int k = 0; k < 9; k++) {
    double T2   = c->u * s->cx[k] + c->v * s->cy[k];
    c->feq[k]   = c->rho * s->w[k] * (1.0 + 3.0 * T2 + 4.5 * T2 * T2 - 1.5 * T1);
    c->metaf[k] = om * c->feq[k] + (1.0 - om) * c->f[k];
  }
}

static void trt(or_solver* s, or_cell* c)
{
  const double om = s->omega, oma = s->omega_a;
  double T1 = c->u * c->u + c->v * c->v;
  for (int k = 0; k < 9; k++) {
    double T2 = c->u * s->cx[k] + c->v * s->cy[k];
    c->feq[k] = c->rho * s->w[k] * (1.0 + 3.0 * T2 + 4.5 * T2 * T2 - 1.5 * T1);
  }
  if (!s->p.trt_fp64) {
    /* symmetric / antisymmetric parts held in binary32 (:383-393) */
    for (int k = 0; k < 9; k++) {
      int o = s->opp[k];
      float fp  = 0.5 * (c->f[k] + c->f[o]);
      float fep = 0.5 * (c->feq[k] + c->feq[o]);
      float fm  = 0.5 * (c->f[k] - c->f[o]);
      float fem = 0.5 * (c->feq[k] - c->feq[o]);
      float dp = fp - fep, dm = fm - fem;
      c->metaf[k] = c->f[k] - (double)dp * om - (double)dm * oma;
    }
  } else {
    for (int k = 0; k < 9; k++) {
      int o = s->opp[k];
      double fp  = 0.5 * (c->f[k] + c->f[o]);
      double fep = 0.5 * (c->feq[k] + c->feq[o]);
      double fm  = 0.5 * (c->f[k] - c->f[o]);
      double fem = 0.5 * (c->feq[k] - c->feq[o]);
      c->metaf[k] = c->f[k] - (fp - fep) * om - (fm - fem) * oma;
    }
  }
}

static void mrt(or_solver* s, or_cell* c)
{
  double U = c->u, V = c->v, R = c->rho;
  double meq[9], mom[9];
  meq[0] = R;
  if (!s->p.mrt_textbook) { /* rho^2 in e and eps (:412-413) */
    meq[1] = R * (-2.0 + 3.0 * R * (U * U + V * V));
    meq[2] = R * (1.0 - 3.0 * R * (U * U + V * V));
  } else {
    meq[1] = R * (-2.0 + 3.0 * (U * U + V * V));
    meq[2] = R * (1.0 - 3.0 * (U * U + V * V));
  }
  meq[3] = R * U;
  meq[4] = -R * U;
  meq[5] = R * V;
  meq[6] = -R * V;
  meq[7] = R * (U * U - V * V);
  meq[8] = R * U * V;
  for (int k = 0; k < 9; k++) {
    double a = 0.0;
    for (int l = 0; l < 9; l++) a = a + s->tm[k][l] * c->f[l];
    mom[k] = a;
  }
  for (int k = 0; k < 9; k++) {
    double b = 0.0;
    for (int l = 0; l < 9; l++) b = b + s->stmiv[k][l] * (mom[l] - meq[l]);
    c->metaf[k] = c->f[k] - b;
  }
}

static void collision_step(or_solver* s)
{
  const long nc = (long)s->n * s->m;
  for (long i = 0; i < nc; i++) {
    or_cell* c = &s->cells[i];
    if (c->fluid != 1) continue;
    switch (s->p.collision) {
      case 1: bgkw(s, c); break;
      case 2: trt(s, c); break;
      case 3: mrt(s, c); break;
    }
  }
}

/* serial/CellFunctions.c:445-450 over all cells (Iterate.c:275-276). */
static void update_f(or_solver* s)
{
  const long nc = (long)s->n * s->m;
  for (long i = 0; i < nc; i++)
    memcpy(s->cells[i].f, s->cells[i].metaf, sizeof(double) * 9);
}

/* serial/Iterate.c:470-488: pull METAF from the upstream neighbour. */
static void streaming_step(or_solver* s)
{
  const long nc = (long)s->n * s->m;
  for (long i = 0; i < nc; i++) {
    or_cell* c = &s->cells[i];
    if (c->fluid != 1) continue;
    for (int k = 0; k < 9; k++)
      if (c->stream[k] == 1) {
        long src = i + s->c[k];
        c->f[k] = (src >= 0 && src < nc) ? s->cells[src].metaf[k] : 0.0;
      }
  }
}

/* ------------------------------------------------------------------------ */
/* Boundaries: serial/Iterate.c:491-550 calling BoundaryConditions.c.        */
/* ------------------------------------------------------------------------ */
/* serial/BoundaryConditions.c:14-58 (InletBC): top lid (moving-wall form)   */
/* and west Zou-He inlet; corners excluded.                                   */
static void inlet_bc(or_cell* c)
{
  if (c->boundary != 2) return;
  double* F = c->f;
  if (c->bc[2] == 2 && c->corner != 1) {
    double rhoN = F[0] + F[1] + F[3] + 2 * (F[2] + F[6] + F[5]);
    F[4] = F[2];
    F[8] = F[6] + rhoN * c->uo / 6.0;
    F[7] = F[5] - rhoN * c->uo / 6.0;
  }
  if (c->bc[3] == 2 && c->corner != 1) {
    double rhoW = (F[0] + F[2] + F[4] + 2.0 * (F[3] + F[6] + F[7])) / (1.0 - c->uo);
    F[1] = F[3] + 2 * rhoW * c->uo / 3;
    F[5] = F[7] + rhoW * c->uo / 6;
    F[8] = F[6] + rhoW * c->uo / 6;
  }
}

/* serial/BoundaryConditions.c:67-95: east Zou-He velocity outlet. */
static void outlet_bc(or_cell* c)
{
  if (c->bc[1] != 3) return;
  double* F = c->f;
  double rhoE = (F[0] + F[2] + F[4] + 2.0 * (F[1] + F[5] + F[8])) / (1 - c->uo);
  F[3] = F[1] - 2 * rhoE * c->uo / 3.0;
  F[7] = F[5] - rhoE * c->uo / 6.0;
  F[6] = F[8] - rhoE * c->uo / 6.0;
}

/* serial/BoundaryConditions.c:104-118: half-way bounce-back of the
 * post-streaming population, in k order. */
static void wall_bc(or_solver* s, or_cell* c)
{
  for (int k = 0; k < 9; k++)
    if (c->bc[k] == 1) c->f[s->opp[k]] = c->f[k];
}

/* serial/BoundaryConditions.c:127-146: Bouzidi interpolated bounce-back. */
static void curved_bc(or_solver* s, or_cell* c)
{
  for (int k = 0; k < 9; k++) {
    if (c->bc[k] != 1) continue;
    int o = s->opp[k];
    double q = c->q[k];
    if (q < 0.5)
      c->f[o] = 2 * q * c->metaf[k] + (1 - 2 * q) * c->fnb[k];
    else
      c->f[o] = c->metaf[k] / 2 / q + (2 * q - 1) / (2 * q) * c->metaf[o];
  }
}

static void boundaries_step(or_solver* s)
{
  const int n = s->n, m = s->m;
  const long nc = (long)n * m;
  for (int j = 0; j < m; j++)
    for (int i = 0; i < n; i++) {
      long idx = (long)j * n + i;
      or_cell* c = &s->cells[idx];
      if (c->fluid != 1) continue;
      inlet_bc(c);
      switch (s->p.curved) {
        case 1:
          for (int k = 0; k < 9; k++) {
            long src = idx + s->c[k];
            c->fnb[k] = (src >= 0 && src < nc) ? s->cells[src].metaf[k] : 0.0;
          }
          curved_bc(s, c);
          break;
        case 2: wall_bc(s, c); break;
      }
      switch (s->p.outlet_profile) {
        case 1: outlet_bc(c); break;
        case 2: /* open boundary, 2nd order (Iterate.c:534-541) */
          if (c->bc[1] == 3 && idx >= 2) {
            const double* a = s->cells[idx - 1].f;
            const double* b = s->cells[idx - 2].f;
            c->f[1] = 2 * a[1] - b[1];
            c->f[5] = 2 * a[5] - b[5];
            c->f[8] = 2 * a[8] - b[8];
          }
          break;
        case 4: /* pressure outlet, Zou & He (1997), east side; not in the reference */
          if (c->bc[1] == 3) {
            double* F = c->f;
            const double ro = s->p.outlet_rho > 0 ? s->p.outlet_rho : (double)s->p.rho_ini;
            double num = F[0] + F[2] + F[4] + 2.0 * (F[1] + F[5] + F[8]);
            double ux = num / ro - 1.0;
            double t = ro * ux / 6.0;
            double h = 0.5 * (F[2] - F[4]);
            F[3] = F[1] - 4.0 * t;
            F[7] = F[5] + h - t;
            F[6] = F[8] - h - t;
          }
          break;
        case 3: /* 1st order, cuda/cuda_step3/GpuBoundaries.cu:200-206 */
          if (c->bc[1] == 3 && idx >= 1) {
            const double* a = s->cells[idx - 1].f;
            c->f[1] = a[1];
            c->f[5] = a[5];
            c->f[8] = a[8];
          }
          break;
      }
    }
}

/* serial/CellFunctions.c:457-495 via Iterate.c:554-567. */
static void update_macro_step(or_solver* s)
{
  const long nc = (long)s->n * s->m;
  const int cdl = s->p.drag_lift;
  for (long i = 0; i < nc; i++) {
    or_cell* c = &s->cells[i];
    if (c->fluid != 1) continue;
    double S = 0.0;
    for (int k = 0; k < 9; k++) S = S + c->f[k];
    c->rho = S;
    double us = 0.0, vs = 0.0;
    for (int k = 0; k < 9; k++) {
      us = us + c->f[k] * s->cx[k];
      vs = vs + c->f[k] * s->cy[k];
    }
    c->u = us / c->rho;
    c->v = vs / c->rho;
    if (c->bc[1] == 3) c->v = 0.0;
    if (cdl != 0 && c->boundary_id == cdl) {
      c->drag = c->rho / 3 * (20 - c->x) / 5;
      c->lift = c->rho / 3 * (20 - c->y) / 5;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Residuals                                                                */
/* ------------------------------------------------------------------------ */
/* kind 0: serial/ComputeResiduals.c:6-51 exactly, including its index use:
 * the loops run i<m, j<n over cell i, i.e. the first m cells n times.
 * kind 1: the full-field sum cuda/cuda_step3/ComputeResiduals.cu:12-47 does.
 * Returns 1 when the L2 norm is NaN (the reference prints DIVERGENCE and
 * exits there). */
static int residuals_ref(or_solver* s, int full)
{
  const int n = s->n, m = s->m;
  const int cdl = s->p.drag_lift;
  double l2 = 0.0, dr = 0.0, li = 0.0;
  if (!full) {
    for (int i = 0; i < m; i++)
      for (int j = 0; j < n; j++) {
        const or_cell* c = &s->cells[i];
        for (int k = 0; k < 9; k++) {
          double d = c->f[k] - c->metaf[k];
          l2 = l2 + d * d;
        }
        if (c->boundary_id == cdl) {
          dr += c->drag;
          li += c->lift;
        }
      }
  } else {
    const long nc = (long)n * m;
    for (long i = 0; i < nc; i++) {
      const or_cell* c = &s->cells[i];
      for (int k = 0; k < 9; k++) {
        double d = c->f[k] - c->metaf[k];
        l2 = l2 + d * d;
      }
      if (c->boundary_id == cdl) {
        dr += c->drag;
        li += c->lift;
      }
    }
  }
  s->resid[1] = sqrt(l2 / (m * n));
  s->resid[0] = sqrt(l2);
  s->resid[2] = dr;
  s->resid[3] = li;
  return s->resid[0] != s->resid[0];
}

/* kind 2 (not in the reference): relative L2 change of the velocity and
 * density fields over the last K steps, evaluated every K steps:
 *   r_u = ||u^t - u^{t-K}|| / (K ||u^t||),  r_rho likewise.
 * Between evaluations the previous values are kept. */
static int residuals_vel(or_solver* s)
{
  const long nc = (long)s->n * s->m;
  const int K = s->p.resid_every > 0 ? s->p.resid_every : 1;
  if (s->iter % K != 0) return 0;
  double du = 0, uu = 0, dr = 0, rr = 0;
  for (long i = 0; i < nc; i++) {
    const or_cell* c = &s->cells[i];
    if (c->fluid != 1) continue;
    double a = c->u - s->u_prev[i], b = c->v - s->v_prev[i];
    double r = c->rho - s->rho_prev[i];
    du += a * a + b * b;
    uu += c->u * c->u + c->v * c->v;
    dr += r * r;
    rr += c->rho * c->rho;
    s->u_prev[i] = c->u;
    s->v_prev[i] = c->v;
    s->rho_prev[i] = c->rho;
  }
  s->resid[0] = uu > 0 ? sqrt(du) / (K * sqrt(uu)) : sqrt(du);
  s->resid[1] = rr > 0 ? sqrt(dr) / (K * sqrt(rr)) : sqrt(dr);
  s->resid[2] = 0;
  s->resid[3] = 0;
  return s->resid[0] != s->resid[0];
}

/* ------------------------------------------------------------------------ */
/* Public API                                                               */
/* ------------------------------------------------------------------------ */
static or_solver* finish_create(or_solver* s)
{
  if (s->num_nodes <= 0) { or_destroy(s); return NULL; }
  comp_data_node(s);
  if (s->n <= 0 || s->m <= 0) { or_destroy(s); return NULL; }
  comp_data_conn(s);
  or_d2q9(s->w, s->cx, s->cy, s->opp, s->c, s->n);
  /* serial/Iterate.c:179-180 */
  s->omega = 1.0 / (3. * s->p.viscosity + 0.5);
  if (s->p.trt_lambda > 0) {
    double lam = s->p.trt_lambda;
    s->omega_a = 1.0 / (lam / (1.0 / s->omega - 0.5) + 0.5);
  } else {
    s->omega_a = 8 * (2 - s->omega) / (8 - s->omega);
  }
  double sm[9] = {1.0, 1.4, 1.4, 1.0, 1.2, 1.0, 1.2, s->omega, s->omega};
  if (s->p.mrt_custom_s) memcpy(sm, s->p.mrt_s, sizeof sm);
  or_mrt_init(s->tm, s->stmiv, sm);
  if (cell_init(s) != 0) { or_destroy(s); return NULL; }
  s->resid[0] = 100.0;
  s->resid[1] = 100.0;
  s->iter = 0;
  if (s->p.resid_kind == 2) {
    const long nc = (long)s->n * s->m;
    s->u_prev = (double*)calloc((size_t)nc, sizeof(double));
    s->v_prev = (double*)calloc((size_t)nc, sizeof(double));
    s->rho_prev = (double*)calloc((size_t)nc, sizeof(double));
    for (long i = 0; i < nc; i++) {
      s->u_prev[i] = s->cells[i].u;
      s->v_prev[i] = s->cells[i].v;
      s->rho_prev[i] = s->cells[i].rho;
    }
  }
  return s;
}

or_solver* or_create(const float* nodes, int num_nodes, const float* conn,
                     int num_conn, const or_params* p)
{
  or_solver* s = (or_solver*)calloc(1, sizeof(or_solver));
  if (!s) return NULL;
  s->p = *p;
  s->num_nodes = num_nodes;
  s->num_conn = num_conn;
  s->nodes = (float(*)[5])malloc(sizeof(float) * 5 * (size_t)(num_nodes > 0 ? num_nodes : 1));
  s->conn = (float(*)[7])malloc(sizeof(float) * 7 * (size_t)(num_conn > 0 ? num_conn : 1));
  if (num_nodes > 0) memcpy(s->nodes, nodes, sizeof(float) * 5 * (size_t)num_nodes);
  if (num_conn > 0) memcpy(s->conn, conn, sizeof(float) * 7 * (size_t)num_conn);
  return finish_create(s);
}

/* serial/FileIO.c:14-82 and :91-174: count records with the typed pattern,
 * then re-read that many records with all-%f conversions. */
static int read_table(const char* path, int cols, float** out)
{
  FILE* fp = fopen(path, "r");
  if (!fp) return -1;
  int lines = 0, i1, i2, i3, i4, i5;
  float f1, f2;
  if (cols == 5) {
    while (fscanf(fp, "%d %d %f %f %d", &i1, &i2, &f1, &f2, &i3) == 5) lines++;
  } else {
    while (fscanf(fp, "%d %d %d %d %f %f %d", &i1, &i2, &i3, &i4, &f1, &f2, &i5) == 7) lines++;
  }
  fclose(fp);
  float* t = (float*)calloc((size_t)(lines > 0 ? lines : 1) * cols, sizeof(float));
  fp = fopen(path, "r");
  for (int L = 0; L < lines; L++) {
    float* r = t + (size_t)L * cols;
    if (cols == 5)
      (void)!fscanf(fp, "%f %f %f %f %f", r, r + 1, r + 2, r + 3, r + 4);
    else
      (void)!fscanf(fp, "%f %f %f %f %f %f %f", r, r + 1, r + 2, r + 3, r + 4, r + 5, r + 6);
  }
  fclose(fp);
  *out = t;
  return lines;
}

or_solver* or_create_from_files(const char* node_file, const char* bc_file,
                                const or_params* p)
{
  float *nodes = NULL, *conn = NULL;
  int nn = read_table(node_file, 5, &nodes);
  if (nn < 0) return NULL;
  int nc = read_table(bc_file, 7, &conn);
  if (nc < 0) { free(nodes); return NULL; }
  or_solver* s = or_create(nodes, nn, conn, nc, p);
  free(nodes);
  free(conn);
  return s;
}

/* serial/FileIO.c:263-289: 16 positional values. */
int or_read_ini(const char* path, or_params* p)
{
  FILE* f = fopen(path, "r");
  if (!f) return -1;
  memset(p, 0, sizeof *p);
  int ok = 0;
  ok += fscanf(f, "%f", &p->uavg);
  ok += fscanf(f, "%f", &p->vavg);
  ok += fscanf(f, "%f", &p->rho_ini);
  ok += fscanf(f, "%f", &p->viscosity);
  ok += fscanf(f, "%d", &p->inlet_profile);
  ok += fscanf(f, "%d", &p->collision);
  ok += fscanf(f, "%d", &p->curved);
  ok += fscanf(f, "%d", &p->outlet_profile);
  ok += fscanf(f, "%d", &p->iterations);
  ok += fscanf(f, "%d", &p->autosave_every);
  ok += fscanf(f, "%d", &p->autosave_after);
  ok += fscanf(f, "%d", &p->postproc);
  ok += fscanf(f, "%d", &p->drag_lift);
  ok += fscanf(f, "%f", &p->crit_v);
  ok += fscanf(f, "%f", &p->crit_rho);
  ok += fscanf(f, "%d", &p->read_former);
  fclose(f);
  return ok == 16 ? 0 : -2;
}

void or_destroy(or_solver* s)
{
  if (!s) return;
  free(s->nodes);
  free(s->conn);
  free(s->cells);
  free(s->u_prev);
  free(s->v_prev);
  free(s->rho_prev);
  free(s);
}

/* serial/Iterate.c:263-332 without the file writes. */
int or_step(or_solver* s, int nsteps)
{
  for (int t = 0; t < nsteps; t++) {
    if (!(s->resid[0] > s->p.crit_v && s->resid[1] > s->p.crit_rho &&
          s->iter < s->p.iterations))
      return 1;
    collision_step(s);
    update_f(s);
    streaming_step(s);
    boundaries_step(s);
    update_macro_step(s);
    s->iter++;
    int bad = 0;
    if (s->p.resid_kind == 2)
      bad = residuals_vel(s);
    else if (s->p.resid_kind == 0)
      bad = residuals_ref(s, 0);
    else if (s->p.resid_kind == 1 &&
             s->iter % (s->p.resid_every > 0 ? s->p.resid_every : 1) == 0)
      bad = residuals_ref(s, 1);
    if (bad) return 2;
  }
  return 0;
}

void or_dims(const or_solver* s, int* n, int* m, double* delta, int* num_inlet,
             double* min_inlet_y, double* max_inlet_y)
{
  if (n) *n = s->n;
  if (m) *m = s->m;
  if (delta) *delta = s->delta;
  if (num_inlet) *num_inlet = s->num_inlet;
  if (min_inlet_y) *min_inlet_y = s->min_inlet_y;
  if (max_inlet_y) *max_inlet_y = s->max_inlet_y;
}

void or_force(const or_solver* s, int bid, double out[2])
{
  const long nc = (long)s->n * s->m;
  double fx = 0.0, fy = 0.0;
  for (long i = 0; i < nc; i++) {
    const or_cell* c = &s->cells[i];
    if (c->fluid != 1 || c->boundary_id != bid) continue;
    for (int k = 1; k < 9; k++)
      if (c->bc[k] == 1) {
        double g = c->metaf[k] + c->f[s->opp[k]];
        fx += s->cx[k] * g;
        fy += s->cy[k] * g;
      }
  }
  out[0] = fx;
  out[1] = fy;
}

int or_iter(const or_solver* s) { return s->iter; }
void or_residuals(const or_solver* s, double out[4]) { memcpy(out, s->resid, sizeof(double) * 4); }
void or_omegas(const or_solver* s, double out[2]) { out[0] = s->omega; out[1] = s->omega_a; }
or_cell* or_cells(or_solver* s) { return s->cells; }
int or_cell_size(void) { return (int)sizeof(or_cell); }

void or_get_macro(const or_solver* s, double* rho, double* u, double* v)
{
  const long nc = (long)s->n * s->m;
  for (long i = 0; i < nc; i++) {
    if (rho) rho[i] = s->cells[i].rho;
    if (u) u[i] = s->cells[i].u;
    if (v) v[i] = s->cells[i].v;
  }
}

void or_get_f(const or_solver* s, double* f, double* metaf)
{
  const long nc = (long)s->n * s->m;
  for (long i = 0; i < nc; i++) {
    if (f) memcpy(f + 9 * i, s->cells[i].f, sizeof(double) * 9);
    if (metaf) memcpy(metaf + 9 * i, s->cells[i].metaf, sizeof(double) * 9);
  }
}

/* serial/FileIO.c:298-348 (final_layout 0) and :354-427 (final_layout 1,
 * whose loop prints cell i of row 0 for every j unless the bug is off). */
int or_write_results(const or_solver* s, const char* path, int final_layout,
                     int final_row0_bug)
{
  FILE* fp = fopen(path, "w");
  if (!fp) return -1;
  const int n = s->n, m = s->m;
  const int csv = (s->p.postproc != 2);
  if (!final_layout) {
    if (csv)
      fprintf(fp, "x,y,u,v,vel_mag,rho,press,fluid\n");
    else {
      fprintf(fp, "Title = \"LBM results\"\n");
      fprintf(fp, "Variables = \"x\",\"y\",\"u\",\"v\",\"u mag\",\"rho\",\"press\",\"fluid\"\n");
      fprintf(fp, "Zone i=%d, j=%d, f=point\n", n, m);
    }
    for (long i = 0; i < (long)n * m; i++) {
      const or_cell* c = &s->cells[i];
      fprintf(fp, csv ? "%f, %f, %f, %f, %f, %f, %f, %d\n" : "%f %f %f %f %f %f %f %d\n",
              c->x, c->y, c->u, c->v, sqrt(c->u * c->u + c->v * c->v), c->rho,
              c->rho / 3.0, c->fluid);
    }
  } else {
    if (csv)
      fprintf(fp, "x,y,u,v,vel_mag,rho,press,fluid,f0,f1,f2,f3,f4,f5,f6,f7,f8\n");
    else {
      fprintf(fp, "Title = \"LBM results\"\n");
      fprintf(fp, "Variables = \"x\",\"y\",\"u\",\"v\",\"u mag\",\"rho\",\"press\",\"fluid\",\"f0\",\"f1\",\"f2\",\"f3\",\"f4\",\"f5\",\"f6\",\"f7\",\"f8\"\n");
      fprintf(fp, "Zone i=%d, j=%d, f=point\n", n, m);
    }
    for (int j = 0; j < m; j++)
      for (int i = 0; i < n; i++) {
        const or_cell* c = &s->cells[final_row0_bug ? i : (long)j * n + i];
        fprintf(fp,
                csv ? "%f, %f, %f, %f, %f, %f, %f, %d, %f, %f, %f, %f, %f, %f, %f, %f, %f\n"
                    : "%f %f %f %f %f %f %f %d %f %f %f %f %f %f %f %f %f\n",
                c->x, c->y, c->u, c->v, sqrt(c->u * c->u + c->v * c->v), c->rho,
                c->rho / 3, c->fluid, c->f[0], c->f[1], c->f[2], c->f[3], c->f[4],
                c->f[5], c->f[6], c->f[7], c->f[8]);
      }
  }
  fclose(fp);
  return 0;
}
