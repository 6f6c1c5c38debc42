// lbm_host.hpp — host-side structures behind the C ABI (include/lbm.h).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/lbm.h"
#include "lbm_kparams.hpp"

namespace lbmhost {

void set_error(const std::string& msg);

// Per-cell record for cells that at least one connector line matched
// (the only cells whose CellIni result differs from the defaults
// BC_ID = 0, Q = 0.5, Boundary = BoundaryID = Corner = 0, stream all 1).
struct BCell {
  int bc[9];
  int corner, boundary, boundary_id;
  int stream;  // bit k = StreamLattice[k]
  double q[9];
};

}  // namespace lbmhost

struct lbm_case {
  int num_nodes = 0, num_conn = 0;
  std::vector<float> nodes;  // num_nodes x 5 (i j x y fluid)
  std::vector<float> conn;   // num_conn x 7 (i j k type x y bid)
  int n = 0, m = 0;
  double delta = 0.0;
  int num_inlet = 0;
  double min_inlet_y = 0.0, max_inlet_y = 0.0;
  // CellIni results (options-independent part), reference order j*n+i
  std::vector<int32_t> id;      // node line of each cell (0 if none, as calloc leaves it)
  std::vector<int32_t> fluid;   // (int)Nod[ID][4]
  std::vector<int32_t> bcell;   // index into bcells or -1
  std::vector<lbmhost::BCell> bcells;
  int num_fluid = 0, num_boundary = 0;
  // packed cases (lbm_case_pack): instance of each cell (-1: filler) and each
  // instance's rectangle (x0, y0, n, m) in the packed grid
  std::vector<int16_t> cell_inst;
  std::vector<int> inst_rect;
  int ninst = 0;

  double x(long lin) const { return (double)nodes[(size_t)id[lin] * 5 + 2]; }
  double y(long lin) const { return (double)nodes[(size_t)id[lin] * 5 + 3]; }
};

namespace lbmhost {
// case_io.cpp
int finish_case(lbm_case* c);  // CompDataNode, CompDataConn, CellIni
double inlet_uo(const lbm_case* c, const lbm_options* o, long lin);
}  // namespace lbmhost

struct lbm_solver {
  const lbm_case* cs = nullptr;  // not owned
  lbm_options opt{};
  int prec = 0, coll = 1;
  hipStream_t stream = nullptr;
  int device = 0;
  // layout
  int pitch = 0, rows = 0, plane = 0;
  size_t elem = 8;
  // device buffers
  void* buf[2] = {nullptr, nullptr};
  int cur = 0;  // buf[cur] holds METAF_{t+1}
  uint8_t* d_cls = nullptr;
  uint8_t* d_tflag = nullptr;
  uint32_t* d_wmask = nullptr;  // class-4 (in-tile wall) cells: stream and wall-link bits
  lbmdev::BNode* d_bn = nullptr;
  int nbn = 0;
  lbmdev::BNode* d_on = nullptr;
  int* d_orow = nullptr;
  int norows = 0, non = 0;
  void* d_side = nullptr;
  int nside = 0;
  double *d_rho = nullptr, *d_u = nullptr, *d_v = nullptr;  // padded macro outputs
  double* d_f = nullptr;      // padded [9][plane] (allocated on demand)
  double* d_metaf = nullptr;  // padded [9][plane] (allocated on demand)
  double* d_snap = nullptr;   // velocity residual snapshot [3][plane]
  double* d_cres = nullptr;   // [m][12]
  double* d_part = nullptr;   // [nparts][4]
  double* d_hist = nullptr;   // [HCAP][4]
  uint32_t* d_seed = nullptr; // FP16 stochastic rounding: step number, 2 slots
  int nparts = 0;
  int nb_bulk = 0, nb_bnd = 0, nb_out = 0;
  bool fused = true;          // bulk + boundary-list nodes in one launch
  bool use_graph = true;      // replay captured two-step graphs for plain steps
  hipGraphExec_t gexec[2] = {nullptr, nullptr};  // by starting buffer parity
  // two plain steps per launch: k_two_steps over its tiles, the band (every
  // other node) two single steps through buf3
  bool two = false;
  int4* d_tiles = nullptr;    // [ntiles] (x0, y0, rows, 0), XCD-ordered, rows 0 = padding
  uint8_t* d_cls_two = nullptr;  // classes, bit 7 set on the nodes the tiles advance
  int ntiles = 0, ntiles_real = 0, two_rows = 0;
  std::vector<int4> two_all;           // every two-step tile, row-major
  std::vector<uint32_t> two_first;     // batched: tile t's instances are two_ids[first[t] .. [t+1])
  std::vector<uint16_t> two_ids;
  long long two_cells = 0, band_cells = 0;
  void* buf3 = nullptr;       // METAF_{t+1} of the band and its ring
  uint8_t* d_cls_b[2] = {nullptr, nullptr};    // classes of the band's first / second step
  uint8_t* d_tflag_b[2] = {nullptr, nullptr};
  uint32_t* d_list_b[2] = {nullptr, nullptr};  // bulk tiles of each band step (padded)
  int nlist_b[2] = {0, 0};
  hipGraphExec_t gexec2[2] = {nullptr, nullptr};  // two-step launches by source parity
  // three plain steps per launch: k_three_steps over its tiles, the band
  // three single steps through buf3 (t+1 of the band and two rings) and buf4
  // (t+2 of the band and one ring); single-lattice solvers only
  bool three = false;
  int4* d_tiles3 = nullptr;      // as d_tiles
  uint8_t* d_cls_three = nullptr;  // classes, bit 7 set on the nodes the tiles advance
  int ntiles3 = 0, ntiles3_real = 0, three_rows = 0;
  long long three_cells = 0, band3_cells = 0;
  void* buf4 = nullptr;
  uint8_t* d_cls_b3[3] = {nullptr, nullptr, nullptr};  // classes of the band's three steps
  uint8_t* d_tflag_b3[3] = {nullptr, nullptr, nullptr};
  uint32_t* d_list_b3[3] = {nullptr, nullptr, nullptr};
  int nlist_b3[3] = {0, 0, 0};
  hipGraphExec_t gexec3[2] = {nullptr, nullptr};  // three-step launches by source parity
  // the last launch was a two- or three-step one: buf[cur^1] holds the state
  // prev_lag (2 or 3) steps back, not the one step back that observe passes
  // and checkpoints read
  bool prev_stale = false;
  int prev_lag = 2;
  lbmdev::KParams kp{};
  // host state
  int iter = 0;
  int observed = -1;          // iteration whose macros are in d_rho/u/v
  int observed_f = -1;        // iteration whose F/METAF are in d_f/d_metaf
  double resid[4] = {100.0, 100.0, 0.0, 0.0};
  // residual history not yet handed to the caller
  std::vector<int> hist_iter;
  std::vector<double> hist_val;
  int pending = 0;            // device hist entries not yet copied
  std::vector<int> pending_iter;
  std::vector<int> pending_what;  // bit 0 residual row, bit 1 force row
  std::vector<int> force_iter;
  std::vector<double> force_val;  // fx, fy per row
  bool diverged = false;
  int diverged_iter = -1;
  bool invalid = false;       // a checkpoint load failed after overwriting part of the state
  std::vector<double> rho0, u0, v0;  // initial macros (reference order), for t = 0
  std::vector<double> grad0;         // field start with non-equilibrium: [3][n*m]
  std::vector<float> inst_visc;      // batched: viscosity of each instance
  int f0_kind = 0;
  lbm_timing timing{};
  std::vector<hipEvent_t> ev;  // timing pool
  size_t dev_bytes = 0;
  // batched instances
  int ninst = 0;
  float* d_inst_par = nullptr;
  uint16_t* d_tile_inst = nullptr;
  uint16_t* d_node_inst = nullptr;
  uint8_t* d_inst_active = nullptr;
  int* d_inst_rect = nullptr;        // storage rectangles [ninst][4]
  double* d_inst_res = nullptr;      // [ninst][4]
  uint32_t* d_tile_list = nullptr;   // bulk tiles of running instances (padded)
  int nb_list = 0;
  std::vector<uint32_t> tile_first;  // tile t's instances: tile_ids[tile_first[t] .. [t+1])
  std::vector<uint16_t> tile_ids;
  int list_align = 8;
  std::vector<uint8_t> inst_active;
  std::vector<int> inst_conv_iter;
  std::vector<double> inst_conv_s, inst_res;
  double t_first_step = -1.0;
};

namespace lbmhost {
int check_hip(hipError_t e, const char* what);
}
