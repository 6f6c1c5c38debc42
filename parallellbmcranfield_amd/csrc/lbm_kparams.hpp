// lbm_kparams.hpp — plain structs shared by host code and kernels.
#pragma once
#include <stdint.h>

namespace lbmdev {

constexpr int INST_PAR_W = 12;  // batched parameter row: omega, omega-, e[9], unused
// Waves of a two-step workgroup (k_two_steps): 4 waves side by side cover a
// tile of 256 columns.
constexpr int TWO_WAVES = 4;
constexpr int TWO_COLS = 64 * TWO_WAVES;

enum KFlags : int {
  KF_TRT_FP64 = 1,       // TRT parts in FP64 (reference rounds them to binary32)
  KF_MRT_TEXTBOOK = 2,   // textbook MRT e/eps equilibria
  KF_XCD_REMAP = 4,      // grid is a multiple of 8 blocks; use XCD-contiguous order
};

// Residual accumulation requested from a pass.
enum ResKind : int { RK_NONE = 0, RK_FULL = 1, RK_VEL = 2 };

// Kernel parameters (passed by value; lives in the kernarg segment).
struct KParams {
  // layout of the padded storage grid: rows 0 and m+1, columns 0 and >= n+1 are halo
  int n, m;
  int pitch;          // elements per storage row (multiple of 64)
  int plane;          // elements per direction plane
  int bulk_begin;     // first flat element of storage row 1
  int bulk_end;       // one past the last element of storage row m
  int flags;
  int res_kind;       // ResKind for this pass
  int curved;         // 1 interpolated (Bouzidi) walls, else half-way bounce-back
  int outlet_profile; // 1 velocity, 2 open 2nd order, 3 open 1st order, 4 pressure
  int force_id;       // boundary id whose links give the momentum-exchange force
  int force_on;       // this pass accumulates the force (MODE 1 only)
  int drag_lift;      // boundary id for the drag/lift placeholder (0 = off)
  int nb_bulk;        // number of bulk workgroups (partials offset of the boundary pass)
  int nb_bnd;         // number of boundary workgroups
  int xcd_group;      // > 0: grouped XCD tile order with this run length
  // physics
  double omega, omega_a, rho0, rho_out;
  float omega_f, omega_a_f, rho0_f, rho_out_f;
  float scale_f, inv_scale_f;
  double stmiv[9][9];  // FP64 MRT: M^-1 S exactly as MRTInitializer forms it
  float mrt_e[9];      // FP32 MRT: s_l / |M_l|^2
  // batched instances (deviation storage only); ninst == 0 otherwise
  int ninst;
  const float* inst_par;        // [ninst][INST_PAR_W]
  const uint16_t* tile_inst;    // instance of each all-interior bulk tile
  const uint16_t* node_inst;    // instance of each padded cell (0xffff: none)
  const uint8_t* inst_active;   // 0: converged, frozen
  const uint32_t* wmask;        // class-4 cells: StreamLattice bits 0-8, wall links 16-24
  const uint32_t* tile_list;    // bulk tiles to run, in launch order (null: all);
                                // 0xffffffff pads the list to the XCD grouping
  // FP16 storage with stochastic rounding: the step number lives on the
  // device in two slots, so replayed graphs advance it.  A step reads
  // seed[seed_slot]; its first workgroup writes seed[seed_slot ^ 1] = that + 1
  // for the next step (no one reads that slot during the launch).  The random
  // words of a stored value hash its cell and the step number.
  uint32_t* seed;
  int seed_slot;
};

// Boundary-node record (host builds, kernel reads).  One per fluid node that
// is not a plain interior node.
struct BNode {
  int32_t idx;      // padded storage index
  int32_t lin;      // reference cell index j*n + i
  int32_t flags;    // BN_* bits
  int32_t side;     // side-buffer slot (outlet feeders / outlet nodes) or -1
  int8_t bc[9];     // BC_ID[k] (CellIni)
  int8_t corner;    // Corner
  int8_t boundary;  // Boundary
  int8_t pad0;
  uint16_t stream;  // bit k = StreamLattice[k]
  uint16_t pad1;
  int32_t boundary_id;
  int32_t feed1, feed2;  // side slots of the (i-1) and (i-2) feeders (outlet nodes), -1 = solid
  int32_t inst;     // batched instance, -1 when not batched
  double q[9];      // Q[k]
  double uo;        // Uo
  double x, y;      // CoordX, CoordY (drag/lift placeholder)
};

enum BNFlags : int {
  BN_COMPAT_RES = 1,  // lin < m: the reference residual sweep reads this cell
  BN_FEEDER = 2,      // store post-boundary F in side[side]
  BN_OUTLET = 4,      // open-outlet node: handled by the outlet pass
  BN_DRAG = 8,        // boundary_id == drag_lift (placeholder drag/lift)
};

// Outputs of an observe pass / residual pass.
struct KOut {
  double* rho;      // padded [plane] macro outputs (observe)
  double* u;
  double* v;
  double* f;        // padded [9][plane] post-boundary F (observe, optional)
  double* metaf;    // padded [9][plane] own METAF (observe, optional)
  double* part;     // per-workgroup partials [nb][6]: residual terms, force x, y
  double* snap;     // velocity residual snapshot, padded [3][plane] (rho, u, v)
  double* cres;     // reference-residual cells: [m][12] (9 diffs, drag, lift, written flag)
  void* side;       // outlet side buffer [slots][9] (compute type)
  const double* grad;  // field start with non-equilibrium: [3][plane] du/dx, dv/dy, du/dy+dv/dx
};

}  // namespace lbmdev
