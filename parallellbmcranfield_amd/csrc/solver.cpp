// solver.cpp — device state, node classification and the step loop behind
// the lbm_solver_* entry points.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "lbm_host.hpp"
#include "lbm_launch.hpp"

using namespace lbmdev;

namespace lbmhost {

int check_hip(hipError_t e, const char* what) {
  if (e == hipSuccess) return LBM_OK;
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return LBM_EHIP;
}

}  // namespace lbmhost

using namespace lbmhost;

#define HC(x)                                  \
  do {                                         \
    int _rc = check_hip((x), #x);              \
    if (_rc != LBM_OK) return _rc;             \
  } while (0)

static int launch_step_passes(lbm_solver* s, const Launch& L);

namespace {

constexpr int HCAP = 1024;  // device history slots
constexpr int HW = 8;       // slot: raw residual [0..3], force [4..5], unused [6..7]

int dev_alloc(lbm_solver* s, void** p, size_t bytes) {
  if (bytes == 0) bytes = 8;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    set_error(std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? LBM_ENOMEM : LBM_EHIP;
  }
  s->dev_bytes += bytes;
  return LBM_OK;
}

// MRTInitializer (serial/CellFunctions.c:11-92): stmiv = M^-1 S with M^-1's
// entries formed as k * (1/36) in FP64, the values the reference's
// initialiser list produces.
void mrt_matrices(const double sm[9], double stmiv[9][9]) {
  static const int Minv36[9][9] = {
      {4, -4, 4, 0, 0, 0, 0, 0, 0},      {4, -1, -2, 6, -6, 0, 0, 9, 0},
      {4, -1, -2, 0, 0, 6, -6, -9, 0},   {4, -1, -2, -6, 6, 0, 0, 9, 0},
      {4, -1, -2, 0, 0, -6, 6, -9, 0},   {4, 2, 1, 6, 3, 6, 3, 0, 9},
      {4, 2, 1, -6, -3, 6, 3, 0, -9},    {4, 2, 1, -6, -3, -6, -3, 0, 9},
      {4, 2, 1, 6, 3, -6, -3, 0, -9}};
  const double a1 = 1. / 36.;
  for (int i = 0; i < 9; i++)
    for (int j = 0; j < 9; j++) {
      const int k = Minv36[i][j];
      const double minv = (k == 0) ? 0.0 : (double)k * a1;
      stmiv[i][j] = minv * sm[j];
    }
}

Launch make_launch(lbm_solver* s, int src_buf) {
  Launch L;
  L.src = s->buf[src_buf];
  L.dst = s->buf[src_buf ^ 1];
  L.cls = s->d_cls;
  L.tflag = s->d_tflag;
  L.bn = s->d_bn;
  L.nbn = s->nbn;
  L.on = s->d_on;
  L.orow = s->d_orow;
  L.norows = s->norows;
  L.nb_bulk = s->nb_bulk;
  L.block = s->opt.block_size > 0 ? s->opt.block_size : 256;
  L.kp = s->kp;
  if (s->d_tile_list) {  // batched: only the tiles of running instances
    L.nb_bulk = s->nb_list;
    L.kp.tile_list = s->d_tile_list;
  }
  // FP16, fused, single lattice, grouped XCD order (the grouping must stay a
  // permutation of the tiles): paired-cell plain steps
  L.pairs = (s->prec == LBM_PREC_F16 || s->prec == LBM_PREC_F16_SR) && s->fused &&
            !s->d_tile_list && !(s->kp.flags & KF_XCD_REMAP) &&
            s->nb_bulk % (8 * s->kp.xcd_group) == 0 && !(s->opt.tune_flags & LBM_TUNE_NO_PAIRS);
  L.out.rho = s->d_rho;
  L.out.u = s->d_u;
  L.out.v = s->d_v;
  L.out.part = s->d_part;
  L.out.snap = s->d_snap;
  L.out.side = s->d_side;
  L.stream = s->stream;
  // stochastic rounding: the step reads the step number from the slot of its
  // source buffer
  L.kp.seed_slot = src_buf;
  return L;
}

hipError_t pass(lbm_solver* s, int mode, int which, const Launch& L) {
  if (s->prec == 0) return launch_pass_f64(s->coll, mode, which, L);
  return launch_pass_lp(s->prec, s->coll, mode, which, L);
}

// Finalise one raw residual row (device sums) into Residuals[0..3].
bool finalize_resid(lbm_solver* s, const double raw[4], double r[4]) {
  const double nm = (double)(s->cs->m * s->cs->n);
  switch (s->opt.resid_kind) {
    case LBM_RESID_REFERENCE:  // serial/ComputeResiduals.c:33-41
      r[1] = std::sqrt(raw[0] / (s->cs->m * s->cs->n));
      r[0] = std::sqrt(raw[0]);
      r[2] = raw[1];
      r[3] = raw[2];
      break;
    case LBM_RESID_FULL:
      r[1] = std::sqrt(raw[0] / nm);
      r[0] = std::sqrt(raw[0]);
      r[2] = raw[2];
      r[3] = raw[3];
      break;
    case LBM_RESID_VELOCITY: {
      const double K = (double)s->opt.resid_every;
      r[0] = raw[1] > 0 ? std::sqrt(raw[0]) / (K * std::sqrt(raw[1])) : std::sqrt(raw[0]);
      r[1] = raw[3] > 0 ? std::sqrt(raw[2]) / (K * std::sqrt(raw[3])) : std::sqrt(raw[2]);
      r[2] = 0.0;
      r[3] = 0.0;
    } break;
    default: return true;
  }
  return r[0] == r[0];  // NaN -> diverged (ComputeResiduals.c:47-51)
}

int flush_history(lbm_solver* s) {
  if (s->pending == 0) return LBM_OK;
  std::vector<double> raw((size_t)s->pending * HW);
  HC(hipMemcpyAsync(raw.data(), s->d_hist, raw.size() * sizeof(double), hipMemcpyDeviceToHost,
                    s->stream));
  HC(hipStreamSynchronize(s->stream));
  for (int i = 0; i < s->pending; i++) {
    const double* row = &raw[(size_t)i * HW];
    const int what = s->pending_what[i];
    if ((what & 2) && !s->diverged) {
      s->force_iter.push_back(s->pending_iter[i]);
      s->force_val.push_back(row[4]);
      s->force_val.push_back(row[5]);
    }
    if (!(what & 1)) continue;
    double r[4];
    const bool ok = finalize_resid(s, row, r);
    if (s->diverged) continue;  // rows after the divergence are not reported
    std::memcpy(s->resid, r, sizeof r);
    if (!ok) {
      // the reference exits inside ComputeResiduals, before the row is printed
      s->diverged = true;
      s->diverged_iter = s->pending_iter[i];
      continue;
    }
    s->hist_iter.push_back(s->pending_iter[i]);
    s->hist_val.insert(s->hist_val.end(), r, r + 4);
  }
  s->pending = 0;
  s->pending_iter.clear();
  s->pending_what.clear();
  return s->diverged ? LBM_DIVERGED : LBM_OK;
}

// After a two-step launch the state one step back (METAF_t, what observe
// passes pull from) existed only inside the launch: form it again from the
// state two steps back with one plain step into buf3 (the same arithmetic, so
// the same bits) and put it in buf[cur^1]; after a three-step launch, with
// two plain steps from the state three steps back.
int refresh_previous(lbm_solver* s) {
  if (!s->prev_stale) return LBM_OK;
  Launch L = make_launch(s, s->cur ^ 1);
  L.kp.res_kind = RK_NONE;
  L.dst = s->buf3;
  int rc = launch_step_passes(s, L);
  if (rc) return rc;
  if (s->prev_lag == 3) {
    // after a three-step launch: the second step from buf3 straight into
    // buf[cur^1] (the state three steps back is no longer needed)
    Launch M = make_launch(s, s->cur ^ 1);
    M.kp.res_kind = RK_NONE;
    M.src = s->buf3;
    M.dst = s->buf[s->cur ^ 1];
    rc = launch_step_passes(s, M);
    if (rc) return rc;
  } else {
    HC(hipMemcpyAsync(s->buf[s->cur ^ 1], s->buf3, (size_t)9 * s->plane * s->elem,
                      hipMemcpyDeviceToDevice, s->stream));
  }
  s->prev_stale = false;
  return LBM_OK;
}

int observe(lbm_solver* s, bool with_f) {
  if (s->iter == 0) return LBM_OK;
  if (s->observed == s->iter && (!with_f || s->observed_f == s->iter)) return LBM_OK;
  int rc0 = refresh_previous(s);
  if (rc0) return rc0;
  if (with_f && !s->d_f) {
    int rc = dev_alloc(s, (void**)&s->d_f, sizeof(double) * 9 * (size_t)s->plane);
    if (rc) return rc;
    rc = dev_alloc(s, (void**)&s->d_metaf, sizeof(double) * 9 * (size_t)s->plane);
    if (rc) return rc;
  }
  Launch L = make_launch(s, s->cur ^ 1);  // previous buffer holds METAF_t
  L.dst = nullptr;
  L.kp.inst_active = nullptr;  // observe converged instances too
  L.kp.tile_list = nullptr;
  L.nb_bulk = s->nb_bulk;
  if (with_f) {
    L.out.f = s->d_f;
    L.out.metaf = s->d_metaf;
  }
  for (int p = 0; p < 3; p++) HC(pass(s, MODE_OBSERVE, p, L));
  s->observed = s->iter;
  if (with_f) s->observed_f = s->iter;
  return LBM_OK;
}

template <class T>
void depad(const lbm_solver* s, const T* padded, T* out, long plane_off) {
  const int n = s->cs->n, m = s->cs->m;
  for (int j = 0; j < m; j++)
    std::memcpy(out + (size_t)j * n, padded + plane_off + (size_t)(j + 1) * s->pitch + 1,
                sizeof(T) * n);
}

}  // namespace

extern "C" {
static int rebuild_tile_list(lbm_solver* s);

// ---- two plain steps per launch
// Default output rows per two-step tile: as few tiles as fill the GPU once
// (every CU holding as many waves of the kernel as fit), so no tile waits
// for a second round, but at least TWO_STEP_MIN_ROWS (the two ring rows
// cost a step-1 row each).
constexpr int TWO_STEP_MIN_ROWS = 8;
static int two_step_rows(const lbm_solver* s, int nstrip, int steps = 2) {
  const int per_cu = s->prec == 0 ? two_step_occupancy_f64(s->coll, steps)
                                  : two_step_occupancy_lp(s->prec, s->coll, steps);
  hipDeviceProp_t prop;
  int cus = 0;
  if (hipGetDeviceProperties(&prop, s->device) == hipSuccess) cus = prop.multiProcessorCount;
  const long long groups = (long long)std::max(per_cu, 1) * std::max(cus, 1);
  const long long segs = std::max(1LL, groups / std::max(nstrip, 1));
  const int m = s->cs->m;
  return std::max(TWO_STEP_MIN_ROWS, (int)((m + segs - 1) / segs));
}

// Tiles of k_two_steps and the band's launches.  A tile is a strip of 256
// padded columns (256 s .. 256 s + 255, one workgroup of four waves) and a
// run of at most `rows` output rows.
// A node is the tile's to advance (bit 7 of its class byte) when it is an
// interior or in-tile wall node (class 1 or 4) and so are all its fluid
// neighbours (class 0 = no node): the t+1 values it pulls can then be formed
// in the tile without a boundary-list record.  The band (every other node:
// boundary-list and outlet nodes and their neighbours) takes two single
// steps: the first advances the band and its one-cell ring from buf[cur]
// into buf3 (the ring's values at t+1 are what the band's second step
// pulls), the second the band from buf3 into buf[cur^1].  Tiles are sorted
// by row, then column, and dealt to the 8 XCDs in contiguous runs (launch
// slot w goes to XCD w % 8), so the tiles that share edge lines meet in one
// L2.
// The tiles of running instances (all tiles for a single lattice) in XCD
// order into d_tiles; captured two-step graphs are dropped (grid and list
// change).
static int rebuild_two_tiles(lbm_solver* s) {
  std::vector<int4> run;
  const int nt = (int)s->two_all.size();
  for (int t = 0; t < nt; t++) {
    bool live = !s->ninst;
    for (uint32_t e = s->two_first[t]; !live && e < s->two_first[t + 1]; e++)
      live = s->inst_active[s->two_ids[e]] != 0;
    if (live) run.push_back(s->two_all[(size_t)t]);
  }
  const int nr = (int)run.size();
  const int q = (nr + 7) / 8;  // tiles per XCD
  std::vector<int4> order((size_t)8 * std::max(q, 1), make_int4(0, 0, 0, 0));
  for (int x = 0; x < 8; x++)
    for (int k = 0; k < q; k++)
      if (x * q + k < nr) order[(size_t)8 * k + x] = run[(size_t)x * q + k];
  HC(hipMemcpyAsync(s->d_tiles, order.data(), sizeof(int4) * order.size(),
                    hipMemcpyHostToDevice, s->stream));
  HC(hipStreamSynchronize(s->stream));
  s->ntiles = (int)order.size();
  for (hipGraphExec_t& g : s->gexec2)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
  return LBM_OK;
}

static int setup_two_step(lbm_solver* s, const std::vector<uint8_t>& cls, int BS, int align,
                          long long active, const std::vector<uint16_t>& node_inst) {
  const lbm_options* o = &s->opt;
  const int n = s->cs->n, m = s->cs->m, pitch = s->pitch;
  const int R = o->two_step_rows > 0 ? o->two_step_rows : two_step_rows(s, n / TWO_COLS + 1);
  auto ok = [](uint8_t c) { return c == 0 || c == 1 || c == 4; };
  auto bulk = [](uint8_t c) { return c == 1 || c == 4; };
  std::vector<uint8_t> cls2(cls);
  long long cells = 0;
  for (int r = 1; r <= m; r++)
    for (int x = 1; x <= n; x++) {
      const size_t idx = (size_t)r * pitch + x;
      if (!bulk(cls[idx])) continue;
      bool g = true;
      for (int dy = -1; dy <= 1 && g; dy++)
        for (int dx = -1; dx <= 1 && g; dx++) g = ok(cls[idx + (long)dy * pitch + dx]);
      if (g) {
        cls2[idx] |= 0x80;
        cells++;
      }
    }
  std::vector<int4> tiles;
  for (int x0 = 0; x0 <= n; x0 += TWO_COLS) {
    const int o1 = std::min(x0 + TWO_COLS, pitch);  // one past the last output column
    for (int y = 1; y <= m; y += R) {
      const int y1 = std::min(y + R - 1, m);
      bool any = false, walls = false;
      for (int r = y; r <= y1 && !any; r++)
        for (int x = x0; x < o1 && !any; x++) any = cls2[(size_t)r * pitch + x] & 0x80;
      // class-4 nodes anywhere in columns x0-1 .. x0+256 of rows y-1 .. y1+1
      for (int r = y - 1; r <= y1 + 1 && any && !walls; r++)
        for (int x = std::max(x0 - 1, 0); x <= std::min(x0 + TWO_COLS, pitch - 1) && !walls; x++)
          walls = cls[(size_t)r * pitch + x] == 4;
      if (any) tiles.push_back(make_int4(x0, y, y1 - y + 1, walls ? 1 : 0));
    }
  }
  std::vector<uint8_t> covered((size_t)s->plane, 0);
  for (size_t i = 0; i < cls2.size(); i++) covered[i] = (cls2[i] & 0x80) != 0;
  // By default on FP32 lattices of at least 2^20 cells and FP64 ones of at
  // least 2^22 (measured: FP32 at 1025^2 57.6 k against 52.1 k MLUPS for
  // single steps; FP64 at 1025^2, with fewer resident waves, loses);
  // FP16's paired-cell single step moves its bytes faster than the one-cell
  // two-step kernel.
  const long long cells_min = s->prec == LBM_PREC_F32 ? (1LL << 20) : (1LL << 22);
  const bool want = o->two_step_rows > 0 ||
                    ((s->prec == LBM_PREC_F64 || s->prec == LBM_PREC_F32) &&
                     (long long)n * m >= cells_min && 3 * cells >= 2 * active);
  if (tiles.empty() || !want) return LBM_OK;
  std::stable_sort(tiles.begin(), tiles.end(), [](const int4& a, const int4& b) {
    return a.y != b.y ? a.y < b.y : a.x < b.x;
  });
  const int nt = (int)tiles.size();
  // batched: the instances whose nodes each tile advances
  s->two_first.assign(1, 0u);
  s->two_ids.clear();
  for (const int4& t : tiles) {
    if (s->ninst) {
      std::vector<uint16_t> ids;
      for (int r = t.y; r < t.y + t.z; r++)
        for (int x = t.x; x < std::min(t.x + TWO_COLS, pitch); x++) {
          const size_t idx = (size_t)r * pitch + x;
          if ((cls2[idx] & 0x80) && node_inst[idx] != 0xffff &&
              std::find(ids.begin(), ids.end(), node_inst[idx]) == ids.end())
            ids.push_back(node_inst[idx]);
        }
      s->two_ids.insert(s->two_ids.end(), ids.begin(), ids.end());
    }
    s->two_first.push_back((uint32_t)s->two_ids.size());
  }
  s->two_all = tiles;
  // the band, and the band with its ring
  std::vector<uint8_t> band((size_t)s->plane, 0), ring((size_t)s->plane, 0);
  long long nband = 0;
  for (int r = 1; r <= m; r++)
    for (int x = 1; x <= n; x++) {
      const size_t idx = (size_t)r * pitch + x;
      if (cls[idx] == 0 || covered[idx]) continue;
      band[idx] = 1;
      nband++;
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) ring[idx + (long)dy * pitch + dx] = 1;
    }
  std::vector<uint8_t> cb[2];
  std::vector<uint8_t> tf[2];
  std::vector<uint32_t> list[2];
  for (int b = 0; b < 2; b++) {
    const std::vector<uint8_t>& in = b == 0 ? ring : band;
    cb[b].assign((size_t)s->plane, 0);
    for (size_t i = 0; i < cb[b].size(); i++)
      if (in[i] && bulk(cls[i])) cb[b][i] = cls[i];
    tf[b].assign((size_t)s->nb_bulk, 0);
    for (int t = 0; t < s->nb_bulk; t++) {
      for (int e = 0; e < BS; e++) {
        const long idx = (long)s->kp.bulk_begin + (long)t * BS + e;
        if (idx < s->kp.bulk_end && cb[b][idx]) {
          tf[b][t] = 2;
          break;
        }
      }
      if (tf[b][t]) list[b].push_back((uint32_t)t);
    }
    while (list[b].empty() || list[b].size() % align) list[b].push_back(0xffffffffu);
  }
  int rc = dev_alloc(s, (void**)&s->d_tiles, sizeof(int4) * 8 * (size_t)((nt + 7) / 8));
  if (!rc) rc = dev_alloc(s, (void**)&s->d_cls_two, cls2.size());
  if (!rc) rc = check_hip(hipMemcpy(s->d_cls_two, cls2.data(), cls2.size(), hipMemcpyHostToDevice), "upload two-step classes");
  if (!rc) rc = dev_alloc(s, &s->buf3, (size_t)9 * s->plane * s->elem + 256);
  for (int b = 0; b < 2 && !rc; b++) {
    rc = dev_alloc(s, (void**)&s->d_cls_b[b], cb[b].size());
    if (!rc) rc = dev_alloc(s, (void**)&s->d_tflag_b[b], tf[b].size());
    if (!rc) rc = dev_alloc(s, (void**)&s->d_list_b[b], sizeof(uint32_t) * list[b].size());
    if (!rc) rc = check_hip(hipMemcpy(s->d_cls_b[b], cb[b].data(), cb[b].size(), hipMemcpyHostToDevice), "upload band classes");
    if (!rc) rc = check_hip(hipMemcpy(s->d_tflag_b[b], tf[b].data(), tf[b].size(), hipMemcpyHostToDevice), "upload band tile flags");
    if (!rc) rc = check_hip(hipMemcpy(s->d_list_b[b], list[b].data(), sizeof(uint32_t) * list[b].size(), hipMemcpyHostToDevice), "upload band tiles");
    s->nlist_b[b] = (int)list[b].size();
  }
  if (!rc) rc = rebuild_two_tiles(s);
  if (rc) return rc;
  s->two = true;
  s->two_rows = R;
  s->ntiles_real = nt;
  s->two_cells = cells;
  s->band_cells = nband;
  return LBM_OK;
}

// ---- three plain steps per launch (single lattices)
// Tiles of k_three_steps and the band's three single steps, as
// setup_two_step one stage deeper: a node is advanced by the tiles when it is
// of class 1 or 4 and every fluid node within two nodes (rows and columns) is
// of class 1 or 4; the band (every other fluid node) takes three single
// steps: the first over the band and its two-node ring from buf[cur] into
// buf3 (inside the three-step launch), the second over the band and its
// one-node ring from buf3 into buf4, the third over the band from buf4 into
// buf[cur^1].
static int setup_three_step(lbm_solver* s, const std::vector<uint8_t>& cls, int BS, int align,
                            long long active) {
  const lbm_options* o = &s->opt;
  const int n = s->cs->n, m = s->cs->m, pitch = s->pitch;
  auto bulk = [](uint8_t c) { return c == 1 || c == 4; };
  // distance to the nearest fluid node that is not bulk, capped at 3
  std::vector<uint8_t> cls3(cls);
  long long cells = 0;
  for (int r = 1; r <= m; r++)
    for (int x = 1; x <= n; x++) {
      const size_t idx = (size_t)r * pitch + x;
      if (!bulk(cls[idx])) continue;
      bool g = true;
      for (int dy = -2; dy <= 2 && g; dy++) {
        const int rr = r + dy;
        if (rr < 0 || rr > m + 1) continue;
        for (int dx = -2; dx <= 2 && g; dx++) {
          const int xx = x + dx;
          if (xx < 0 || xx >= pitch) continue;
          const uint8_t c = cls[(size_t)rr * pitch + xx];
          g = c == 0 || bulk(c);
        }
      }
      if (g) {
        cls3[idx] |= 0x80;
        cells++;
      }
    }
  // By default with FP32 storage on lattices of at least 2^22 cells beside
  // two-step launches (MRT, MLUPS three-step against two-step: 4097^2 134.4 k
  // / 114.5 k, 3073^2 126.7 k / 112.0 k, 2049^2 111.9 k / 101.6 k; at 1025^2
  // the four ring rows of its 8-row tiles lose, 52.2 k / 56.7 k; FP64 at
  // 4097^2, with two resident workgroups per CU, 55.3 k / 56.9 k)
  const bool want = o->three_step_rows > 0 ||
                    (s->two && s->prec == LBM_PREC_F32 && (long long)n * m >= (1LL << 22) &&
                     3 * cells >= 2 * active);
  if (cells == 0 || !want) return LBM_OK;
  const int R = o->three_step_rows > 0 ? o->three_step_rows : two_step_rows(s, n / TWO_COLS + 1, 3);
  std::vector<int4> tiles;
  for (int x0 = 0; x0 <= n; x0 += TWO_COLS) {
    const int o1 = std::min(x0 + TWO_COLS, pitch);  // one past the last output column
    for (int y = 1; y <= m; y += R) {
      const int y1 = std::min(y + R - 1, m);
      bool any = false, walls = false;
      for (int r = y; r <= y1 && !any; r++)
        for (int x = x0; x < o1 && !any; x++) any = cls3[(size_t)r * pitch + x] & 0x80;
      // class-4 nodes anywhere in columns x0-3 .. x0+258 of rows y-3 .. y1+3
      for (int r = std::max(y - 3, 0); r <= std::min(y1 + 3, m + 1) && any && !walls; r++)
        for (int x = std::max(x0 - 3, 0); x <= std::min(x0 + TWO_COLS + 2, pitch - 1) && !walls; x++)
          walls = cls[(size_t)r * pitch + x] == 4;
      if (any) tiles.push_back(make_int4(x0, y, y1 - y + 1, walls ? 1 : 0));
    }
  }
  std::stable_sort(tiles.begin(), tiles.end(), [](const int4& a, const int4& b) {
    return a.y != b.y ? a.y < b.y : a.x < b.x;
  });
  const int nt = (int)tiles.size();
  // the band, then its rings: reg[0] band + 2 rings, reg[1] band + 1 ring, reg[2] band
  std::vector<uint8_t> reg[3];
  for (auto& v : reg) v.assign((size_t)s->plane, 0);
  long long nband = 0;
  for (int r = 1; r <= m; r++)
    for (int x = 1; x <= n; x++) {
      const size_t idx = (size_t)r * pitch + x;
      if (cls[idx] == 0 || (cls3[idx] & 0x80)) continue;
      nband++;
      reg[2][idx] = 1;
      for (int dy = -2; dy <= 2; dy++)
        for (int dx = -2; dx <= 2; dx++) {
          const int rr = r + dy, xx = x + dx;
          if (rr < 0 || rr > m + 1 || xx < 0 || xx >= pitch) continue;
          const size_t j = (size_t)rr * pitch + xx;
          reg[0][j] = 1;
          if (dy >= -1 && dy <= 1 && dx >= -1 && dx <= 1) reg[1][j] = 1;
        }
    }
  std::vector<uint8_t> cb[3], tf[3];
  std::vector<uint32_t> list[3];
  for (int b = 0; b < 3; b++) {
    cb[b].assign((size_t)s->plane, 0);
    for (size_t i = 0; i < cb[b].size(); i++)
      if (reg[b][i] && bulk(cls[i])) cb[b][i] = cls[i];
    tf[b].assign((size_t)s->nb_bulk, 0);
    for (int t = 0; t < s->nb_bulk; t++) {
      for (int e = 0; e < BS; e++) {
        const long idx = (long)s->kp.bulk_begin + (long)t * BS + e;
        if (idx < s->kp.bulk_end && cb[b][idx]) {
          tf[b][t] = 2;
          break;
        }
      }
      if (tf[b][t]) list[b].push_back((uint32_t)t);
    }
    while (list[b].empty() || list[b].size() % align) list[b].push_back(0xffffffffu);
  }
  // XCD order as rebuild_two_tiles
  const int q = (nt + 7) / 8;
  std::vector<int4> order((size_t)8 * std::max(q, 1), make_int4(0, 0, 0, 0));
  for (int x = 0; x < 8; x++)
    for (int k = 0; k < q; k++)
      if (x * q + k < nt) order[(size_t)8 * k + x] = tiles[(size_t)x * q + k];
  int rc = dev_alloc(s, (void**)&s->d_tiles3, sizeof(int4) * order.size());
  if (!rc) rc = check_hip(hipMemcpy(s->d_tiles3, order.data(), sizeof(int4) * order.size(), hipMemcpyHostToDevice), "upload three-step tiles");
  if (!rc) rc = dev_alloc(s, (void**)&s->d_cls_three, cls3.size());
  if (!rc) rc = check_hip(hipMemcpy(s->d_cls_three, cls3.data(), cls3.size(), hipMemcpyHostToDevice), "upload three-step classes");
  if (!rc && !s->buf3) rc = dev_alloc(s, &s->buf3, (size_t)9 * s->plane * s->elem + 256);
  if (!rc) rc = dev_alloc(s, &s->buf4, (size_t)9 * s->plane * s->elem + 256);
  for (int b = 0; b < 3 && !rc; b++) {
    rc = dev_alloc(s, (void**)&s->d_cls_b3[b], cb[b].size());
    if (!rc) rc = dev_alloc(s, (void**)&s->d_tflag_b3[b], tf[b].size());
    if (!rc) rc = dev_alloc(s, (void**)&s->d_list_b3[b], sizeof(uint32_t) * list[b].size());
    if (!rc) rc = check_hip(hipMemcpy(s->d_cls_b3[b], cb[b].data(), cb[b].size(), hipMemcpyHostToDevice), "upload band classes");
    if (!rc) rc = check_hip(hipMemcpy(s->d_tflag_b3[b], tf[b].data(), tf[b].size(), hipMemcpyHostToDevice), "upload band tile flags");
    if (!rc) rc = check_hip(hipMemcpy(s->d_list_b3[b], list[b].data(), sizeof(uint32_t) * list[b].size(), hipMemcpyHostToDevice), "upload band tiles");
    s->nlist_b3[b] = (int)list[b].size();
  }
  if (rc) return rc;
  s->three = true;
  s->ntiles3 = (int)order.size();
  s->ntiles3_real = nt;
  s->three_rows = R;
  s->three_cells = cells;
  s->band3_cells = nband;
  return LBM_OK;
}
}

static int solver_create(const lbm_case* c, const lbm_options* o, const double* frho,
                         const double* fu, const double* fv, const lbm_instance* inst,
                         int ninst, lbm_solver** out) {
  if (!c || !o || !out) return LBM_EINVAL;
  *out = nullptr;
  if (c->ninst != ninst || (ninst > 0 && !inst)) {
    set_error(c->ninst ? "a packed case needs lbm_solver_create_batched with one parameter "
                         "set per instance"
                       : "instance parameters given for a case that is not packed");
    return LBM_EINVAL;
  }
  if (ninst > 0 && o->precision == LBM_PREC_F64) {
    set_error("batched instances use FP32 or FP16 storage");
    return LBM_EINVAL;
  }
  if (ninst > 0 && o->force_id != 0) {
    set_error("force sampling is not offered on batched solvers");
    return LBM_EINVAL;
  }
  if (o->collision < 1 || o->collision > 3) {
    set_error("collision model must be 1 (BGKW), 2 (TRT) or 3 (MRT)");
    return LBM_EINVAL;
  }
  if (o->precision < 0 || o->precision > 3) {
    set_error("precision must be LBM_PREC_F64/F32/F16/F16_SR");
    return LBM_EINVAL;
  }
  if (o->resid_kind < 0 || o->resid_kind > 3) {
    set_error("unknown residual kind");
    return LBM_EINVAL;
  }
  if ((o->init == LBM_INIT_FIELD || o->init == LBM_INIT_FIELD_NEQ) && (!frho || !fu || !fv)) {
    set_error("LBM_INIT_FIELD needs rho/u/v arrays (lbm_solver_create_field)");
    return LBM_EINVAL;
  }
  if (o->block_size != 0 && o->block_size != 256 && o->block_size != 512 &&
      o->block_size != 1024) {
    set_error("block_size must be 0, 256, 512 or 1024");
    return LBM_EINVAL;
  }
  const int n = c->n, m = c->m;
  lbm_solver* s = new lbm_solver();
  s->cs = c;
  s->opt = *o;
  if (s->opt.resid_every < 1) s->opt.resid_every = 1;
  // batched instances converge one by one on their own velocity residual
  if (ninst > 0) s->opt.resid_kind = LBM_RESID_VELOCITY;
  s->prec = o->precision;
  s->coll = o->collision;
  s->elem = s->prec == 0 ? 8 : (s->prec == 1 ? 4 : 2);  // both FP16 kinds: 2
  s->pitch = ((n + 2) + 63) / 64 * 64;
  s->rows = m + 2;
  const long long plane = ((long long)s->pitch * s->rows + 255) / 256 * 256;
  if (plane * 9 >= (1LL << 31)) {
    set_error("lattice too large for 32-bit plane indexing");
    delete s;
    return LBM_EINVAL;
  }
  s->plane = (int)plane;

  int rc = check_hip(hipGetDevice(&s->device), "hipGetDevice");
  if (rc == LBM_OK) rc = check_hip(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking),
                                   "hipStreamCreate");
  if (rc != LBM_OK) {
    delete s;
    return rc;
  }

  // ---- physics constants (serial/Iterate.c:179-188)
  KParams& kp = s->kp;
  std::memset(&kp, 0, sizeof kp);
  kp.n = n;
  kp.m = m;
  kp.pitch = s->pitch;
  kp.plane = s->plane;
  kp.bulk_begin = s->pitch;
  kp.bulk_end = (m + 1) * s->pitch;
  kp.curved = o->curved;
  kp.outlet_profile = o->outlet_profile;
  kp.drag_lift = o->drag_lift;
  const double omega = 1.0 / (3. * (double)o->viscosity + 0.5);
  double omega_a = 8 * (2 - omega) / (8 - omega);
  if (o->trt_lambda > 0) omega_a = 1.0 / (o->trt_lambda / (1.0 / omega - 0.5) + 0.5);
  kp.omega = omega;
  kp.omega_a = omega_a;
  kp.omega_f = (float)omega;
  kp.omega_a_f = (float)omega_a;
  kp.rho0 = (double)o->rho_ini;
  kp.rho0_f = o->rho_ini;
  kp.rho_out = o->outlet_rho > 0 ? o->outlet_rho : (double)o->rho_ini;
  kp.rho_out_f = (float)kp.rho_out;
  kp.force_id = o->force_id;
  const float scale = o->f16_scale > 0 ? o->f16_scale : 4096.0f;
  kp.scale_f = scale;
  kp.inv_scale_f = 1.0f / scale;
  double sm[9] = {1.0, 1.4, 1.4, 1.0, 1.2, 1.0, 1.2, omega, omega};
  if (o->mrt_custom_rates) std::memcpy(sm, o->mrt_rates, sizeof sm);
  mrt_matrices(sm, kp.stmiv);
  static const double norm[9] = {9, 36, 36, 6, 12, 6, 12, 4, 4};
  for (int l = 0; l < 9; l++) kp.mrt_e[l] = (float)(sm[l] / norm[l]);
  kp.flags = 0;
  if (o->trt_fp64_parts) kp.flags |= KF_TRT_FP64;
  if (o->mrt_textbook_meq) kp.flags |= KF_MRT_TEXTBOOK;

  // Uo of a cell: the inlet profile, or the lid speed of its instance
  auto cell_uo = [&](long lin) -> double {
    if (c->ninst && c->cell_inst[lin] >= 0 && o->inlet_profile == 2)
      return (double)inst[c->cell_inst[lin]].uavg;
    return inlet_uo(c, o, lin);
  };
  // ---- node classes, boundary list, open-outlet rows
  std::vector<uint8_t> cls((size_t)s->plane, 0);
  const long nc = (long)n * m;
  const bool open_outlet = (o->outlet_profile == 2 || o->outlet_profile == 3);
  std::vector<int> side_of((size_t)nc, -1);
  std::vector<uint8_t> is_outlet((size_t)nc, 0);
  int nside = 1;  // slot 0: solid cell (F = 0)
  if (open_outlet) {
    for (long lin = 0; lin < nc; lin++) {
      if (c->fluid[lin] != 1) continue;
      const int b = c->bcell[lin];
      if (b >= 0 && c->bcells[b].bc[1] == 3) is_outlet[lin] = 1;
    }
    for (long lin = 0; lin < nc; lin++) {
      if (!is_outlet[lin]) continue;
      if (side_of[lin] < 0) side_of[lin] = nside++;
      for (int d = 1; d <= (o->outlet_profile == 2 ? 2 : 1); d++) {
        const long f = lin - d;  // Cells + n*j + i - d (serial/Iterate.c:538)
        if (f >= 0 && c->fluid[f] == 1 && side_of[f] < 0) side_of[f] = nside++;
      }
    }
  }
  std::vector<BNode> bn, on;
  std::vector<int> orow;
  std::vector<uint32_t> wmask;  // class-4 cells: StreamLattice and wall-link bits
  int last_row = -1;
  for (long lin = 0; lin < nc; lin++) {
    if (c->fluid[lin] != 1) continue;
    const int i = (int)(lin % n), j = (int)(lin / n);
    const int idx = (j + 1) * s->pitch + (i + 1);
    const int b = c->bcell[lin];
    const BCell* B = b >= 0 ? &c->bcells[b] : nullptr;
    bool special = lin < m || side_of[lin] >= 0;
    if (B) {
      for (int k = 0; k < 9; k++) special = special || B->bc[k] != 0;
      special = special || B->boundary_id != 0;
    }
    if (!special) {
      cls[idx] = 1;
      continue;
    }
    BNode r;
    std::memset(&r, 0, sizeof r);
    r.idx = idx;
    r.lin = (int)lin;
    r.side = side_of[lin];
    r.feed1 = r.feed2 = 0;
    for (int k = 0; k < 9; k++) {
      const int v = B ? B->bc[k] : 0;
      if (v < -128 || v > 127) {
        set_error("boundary type out of range");
        lbm_solver_destroy(s);
        return LBM_EPARSE;
      }
      r.bc[k] = (int8_t)v;
      r.q[k] = B ? B->q[k] : 0.5;
    }
    r.corner = (int8_t)(B ? B->corner : 0);
    r.boundary = (int8_t)(B ? B->boundary : 0);
    r.stream = (uint16_t)(B ? B->stream : 0x1ff);
    r.boundary_id = B ? B->boundary_id : 0;
    r.uo = cell_uo(lin);
    r.inst = c->ninst ? (int32_t)c->cell_inst[lin] : -1;
    r.x = c->x(lin);
    r.y = c->y(lin);
    if (lin < m) r.flags |= BN_COMPAT_RES;
    if (o->drag_lift != 0 && r.boundary_id == o->drag_lift) r.flags |= BN_DRAG;
    if (is_outlet[lin]) {
      r.flags |= BN_OUTLET;
      const long f1 = lin - 1, f2 = lin - 2;
      r.feed1 = (f1 >= 0 && c->fluid[f1] == 1) ? side_of[f1] : 0;
      r.feed2 = (f2 >= 0 && c->fluid[f2] == 1) ? side_of[f2] : 0;
      cls[idx] = 3;
      if (j != last_row) {
        orow.push_back((int)on.size());
        last_row = j;
      }
      on.push_back(r);
    } else {
      if (side_of[lin] >= 0) r.flags |= BN_FEEDER;
      // straight-wall-only node: runs in its bulk tile (class 4, see pull_cell)
      bool wall_only = o->curved == 2 && r.flags == 0 &&
                       !(o->force_id != 0 && r.boundary_id == o->force_id) &&
                       !(o->tune_flags & LBM_TUNE_NO_TILE_WALLS);
      uint32_t wm = r.stream & 0x1ffu;
      for (int k = 0; k < 9; k++) {
        wall_only = wall_only && (r.bc[k] == 0 || r.bc[k] == 1);
        if (r.bc[k] == 1) wm |= 1u << (16 + k);
      }
      if (wall_only) {
        cls[idx] = 4;
        if (wmask.empty()) wmask.assign((size_t)s->plane, 0u);
        wmask[idx] = wm;
        continue;
      }
      cls[idx] = 2;
      bn.push_back(r);
    }
  }
  orow.push_back((int)on.size());
  s->nbn = (int)bn.size();
  s->non = (int)on.size();
  s->norows = (int)orow.size() - 1;
  s->nside = nside;

  // ---- grids
  const int BS = o->block_size > 0 ? o->block_size : 256;
  const int count = kp.bulk_end - kp.bulk_begin;
  // grouped XCD order, runs of 4 tiles (measured best on the 4097^2 lattice)
  const int G = o->xcd_group > 0 ? o->xcd_group : 4;
  const int align = 8 * G;
  s->nb_bulk = ((count + BS - 1) / BS + align - 1) / align * align;
  kp.xcd_group = G;
  if (o->tune_flags & LBM_TUNE_XCD_REMAP) kp.flags |= KF_XCD_REMAP;
  s->fused = BS == 256 && !(o->tune_flags & LBM_TUNE_SPLIT_BOUNDARY);
  s->use_graph = (o->tune_flags & LBM_TUNE_GRAPH) && !(o->tune_flags & LBM_TUNE_NO_GRAPH);
  s->nb_bnd = (s->nbn + 255) / 256;
  s->nb_out = (s->norows + 63) / 64;
  // per-tile summary of the classes (tiles of BS cells in launch order)
  std::vector<uint8_t> tflag((size_t)s->nb_bulk, 0);
  for (int t = 0; t < s->nb_bulk; t++) {
    int nb = 0, nw = 0, other = 0;
    for (int e = 0; e < BS; e++) {
      const long idx = (long)kp.bulk_begin + (long)t * BS + e;
      if (idx < kp.bulk_end && cls[idx] == 1) nb++;
      else if (idx < kp.bulk_end && cls[idx] == 4) nw++;
      else other++;
    }
    tflag[t] = nb + nw == 0 ? 0 : (nw == 0 && other == 0 ? 1 : 2);
  }
  kp.nb_bulk = s->nb_bulk;
  // ---- batched instances: parameter rows, instance of each cell / pure tile
  std::vector<float> inst_par;
  std::vector<uint16_t> node_inst, tile_inst;
  std::vector<int> srect;
  if (ninst > 0) {
    s->ninst = ninst;
    s->inst_visc.resize((size_t)ninst);
    for (int k = 0; k < ninst; k++) s->inst_visc[k] = inst[k].viscosity;
    inst_par.assign((size_t)ninst * INST_PAR_W, 0.0f);
    for (int k = 0; k < ninst; k++) {
      const double om = 1.0 / (3. * (double)inst[k].viscosity + 0.5);
      double oma = 8 * (2 - om) / (8 - om);
      if (o->trt_lambda > 0) oma = 1.0 / (o->trt_lambda / (1.0 / om - 0.5) + 0.5);
      double smk[9];
      std::memcpy(smk, sm, sizeof smk);
      smk[7] = smk[8] = om;  // stress moments relax at 1/tau of the instance
      float* r = &inst_par[(size_t)k * INST_PAR_W];
      r[0] = (float)om;
      r[1] = (float)oma;
      for (int l = 0; l < 9; l++) r[2 + l] = (float)(smk[l] / norm[l]);
    }
    node_inst.assign((size_t)s->plane, 0xffff);
    for (long lin = 0; lin < nc; lin++) {
      const int ii = (int)(lin % n), jj = (int)(lin / n);
      if (c->cell_inst[lin] >= 0)
        node_inst[(size_t)(jj + 1) * s->pitch + (ii + 1)] = (uint16_t)c->cell_inst[lin];
    }
    tile_inst.assign((size_t)s->nb_bulk, 0);
    for (int t = 0; t < s->nb_bulk; t++) {
      const long idx = (long)kp.bulk_begin + (long)t * BS;
      if (tflag[t] != 1) continue;
      tile_inst[t] = node_inst[idx];
      for (int e = 1; e < BS; e++)
        if (node_inst[idx + e] != tile_inst[t]) tflag[t] = 2;  // spans two instances
    }
    srect.resize((size_t)ninst * 4);
    for (int k = 0; k < ninst; k++) {
      srect[4 * k] = c->inst_rect[4 * k] + 1;  // storage column of the rectangle
      srect[4 * k + 1] = c->inst_rect[4 * k + 1] + 1;
      srect[4 * k + 2] = c->inst_rect[4 * k + 2];
      srect[4 * k + 3] = c->inst_rect[4 * k + 3];
    }
    // instances present in each tile (class-1 cells), for the tile list
    s->tile_first.assign((size_t)s->nb_bulk + 1, 0);
    s->tile_ids.clear();
    for (int t = 0; t < s->nb_bulk; t++) {
      s->tile_first[t] = (uint32_t)s->tile_ids.size();
      const size_t e0 = s->tile_ids.size();
      for (int e = 0; e < BS; e++) {
        const long idx = (long)kp.bulk_begin + (long)t * BS + e;
        if (idx >= kp.bulk_end || (cls[idx] != 1 && cls[idx] != 4)) continue;
        const uint16_t k = node_inst[idx];
        bool seen = false;
        for (size_t q = e0; q < s->tile_ids.size(); q++) seen = seen || s->tile_ids[q] == k;
        if (!seen) s->tile_ids.push_back(k);
      }
    }
    s->tile_first[s->nb_bulk] = (uint32_t)s->tile_ids.size();
    s->list_align = align;
    s->inst_active.assign((size_t)ninst, 1);
    s->inst_conv_iter.assign((size_t)ninst, -1);
    s->inst_conv_s.assign((size_t)ninst, -1.0);
    s->inst_res.assign((size_t)ninst, 100.0);
  }
  kp.nb_bnd = s->nb_bnd;
  s->nparts = s->nb_bulk + s->nb_bnd + s->nb_out;

  // ---- device memory
  const size_t sbytes = (size_t)9 * s->plane * s->elem;
  const size_t rsz = s->prec == 0 ? sizeof(double) : sizeof(float);
  rc = LBM_OK;
  // + 256 B: the paired-cell kernel reads whole 2-cell words around a pull,
  // which can reach one word past the last plane's last row
  if (!rc) rc = dev_alloc(s, &s->buf[0], sbytes + 256);
  if (!rc) rc = dev_alloc(s, &s->buf[1], sbytes + 256);
  if (!rc) rc = dev_alloc(s, (void**)&s->d_cls, (size_t)s->plane);
  if (!rc) rc = dev_alloc(s, (void**)&s->d_tflag, (size_t)s->nb_bulk);
  if (!rc) rc = dev_alloc(s, (void**)&s->d_bn, sizeof(BNode) * std::max(1, s->nbn));
  if (!rc) rc = dev_alloc(s, (void**)&s->d_on, sizeof(BNode) * std::max(1, s->non));
  if (!rc) rc = dev_alloc(s, (void**)&s->d_orow, sizeof(int) * orow.size());
  if (!rc) rc = dev_alloc(s, &s->d_side, rsz * 9 * (size_t)nside);
  if (!rc) rc = dev_alloc(s, (void**)&s->d_rho, sizeof(double) * (size_t)s->plane);
  if (!rc) rc = dev_alloc(s, (void**)&s->d_u, sizeof(double) * (size_t)s->plane);
  if (!rc) rc = dev_alloc(s, (void**)&s->d_v, sizeof(double) * (size_t)s->plane);
  if (!rc) rc = dev_alloc(s, (void**)&s->d_part, sizeof(double) * 6 * (size_t)s->nparts);
  if (!rc) rc = dev_alloc(s, (void**)&s->d_hist, sizeof(double) * HW * HCAP);
  if (!rc && s->prec == LBM_PREC_F16_SR) {  // step number of stochastic rounding, 2 slots
    rc = dev_alloc(s, (void**)&s->d_seed, 2 * sizeof(uint32_t));
    if (!rc) rc = check_hip(hipMemset(s->d_seed, 0, 2 * sizeof(uint32_t)), "clear step number");
    kp.seed = s->d_seed;
  }
  if (!rc && s->opt.resid_kind == LBM_RESID_REFERENCE)
    rc = dev_alloc(s, (void**)&s->d_cres, sizeof(double) * 12 * (size_t)m);
  if (!rc && s->opt.resid_kind == LBM_RESID_VELOCITY)
    rc = dev_alloc(s, (void**)&s->d_snap, sizeof(double) * (ninst ? 5 : 3) * (size_t)s->plane);
  if (!rc && ninst > 0)
    rc = check_hip(hipMemset(s->d_snap + 3 * (size_t)s->plane, 0, sizeof(double) * 2 * (size_t)s->plane), "clear snapshot");
  if (!rc && ninst > 0) {
    rc = dev_alloc(s, (void**)&s->d_inst_par, sizeof(float) * inst_par.size());
    if (!rc) rc = dev_alloc(s, (void**)&s->d_node_inst, sizeof(uint16_t) * node_inst.size());
    if (!rc) rc = dev_alloc(s, (void**)&s->d_tile_inst, sizeof(uint16_t) * tile_inst.size());
    if (!rc) rc = dev_alloc(s, (void**)&s->d_inst_active, (size_t)ninst);
    if (!rc) rc = dev_alloc(s, (void**)&s->d_inst_rect, sizeof(int) * srect.size());
    if (!rc) rc = dev_alloc(s, (void**)&s->d_inst_res, sizeof(double) * 4 * (size_t)ninst);
    if (!rc) rc = dev_alloc(s, (void**)&s->d_tile_list, sizeof(uint32_t) * ((size_t)s->nb_bulk + align));
    if (!rc) rc = check_hip(hipMemcpy(s->d_inst_par, inst_par.data(), sizeof(float) * inst_par.size(), hipMemcpyHostToDevice), "upload");
    if (!rc) rc = check_hip(hipMemcpy(s->d_node_inst, node_inst.data(), sizeof(uint16_t) * node_inst.size(), hipMemcpyHostToDevice), "upload");
    if (!rc) rc = check_hip(hipMemcpy(s->d_tile_inst, tile_inst.data(), sizeof(uint16_t) * tile_inst.size(), hipMemcpyHostToDevice), "upload");
    if (!rc) rc = check_hip(hipMemcpy(s->d_inst_active, s->inst_active.data(), (size_t)ninst, hipMemcpyHostToDevice), "upload");
    if (!rc) rc = check_hip(hipMemcpy(s->d_inst_rect, srect.data(), sizeof(int) * srect.size(), hipMemcpyHostToDevice), "upload");
    kp.ninst = ninst;
    kp.inst_par = s->d_inst_par;
    kp.tile_inst = s->d_tile_inst;
    kp.node_inst = s->d_node_inst;
    kp.inst_active = s->d_inst_active;
  }
  if (rc) {
    lbm_solver_destroy(s);
    return rc;
  }

  // ---- initial macros (serial/CellFunctions.c:148-150,267): Rho = rho_ini,
  // U = Uo, V = 0 everywhere; rest / field starts change the fluid cells.
  s->rho0.assign((size_t)nc, (double)o->rho_ini);
  s->u0.assign((size_t)nc, 0.0);
  s->v0.assign((size_t)nc, 0.0);
  for (long lin = 0; lin < nc; lin++) {
    s->u0[lin] = cell_uo(lin);
    if (c->fluid[lin] != 1) continue;
    if (o->init == LBM_INIT_REST) {
      s->u0[lin] = 0.0;
    } else if (o->init == LBM_INIT_FIELD || o->init == LBM_INIT_FIELD_NEQ) {
      s->rho0[lin] = frho[lin];
      s->u0[lin] = fu[lin];
      s->v0[lin] = fv[lin];
    }
  }
  s->f0_kind = o->init == LBM_INIT_REST ? 1
              : o->init == LBM_INIT_FIELD ? 2
              : o->init == LBM_INIT_FIELD_NEQ ? 3 : 0;
  if (s->f0_kind == 3) {
    // velocity gradients of the field by central differences; a neighbour
    // that is not fluid is a half-way wall: ghost = 2 u_wall - u, u_wall = (Uo, 0)
    // on inlet links (type 2), else 0
    s->grad0.assign(3 * (size_t)nc, 0.0);
    for (long lin = 0; lin < nc; lin++) {
      if (c->fluid[lin] != 1) continue;
      const int i = (int)(lin % n), j = (int)(lin / n);
      const int b = c->bcell[lin];
      auto nb = [&](int k, int di, int dj, bool want_u) -> double {
        const int ii = i + di, jj = j + dj;
        const double self = want_u ? s->u0[lin] : s->v0[lin];
        if (ii >= 0 && ii < n && jj >= 0 && jj < m && c->fluid[(long)jj * n + ii] == 1)
          return want_u ? s->u0[(long)jj * n + ii] : s->v0[(long)jj * n + ii];
        const int t = b >= 0 ? c->bcells[b].bc[k] : 0;
        const double uw = (t == 2 && want_u) ? cell_uo(lin) : 0.0;
        return 2.0 * uw - self;
      };
      const double dudx = 0.5 * (nb(1, 1, 0, true) - nb(3, -1, 0, true));
      const double dudy = 0.5 * (nb(2, 0, 1, true) - nb(4, 0, -1, true));
      const double dvdx = 0.5 * (nb(1, 1, 0, false) - nb(3, -1, 0, false));
      const double dvdy = 0.5 * (nb(2, 0, 1, false) - nb(4, 0, -1, false));
      s->grad0[lin] = dudx;
      s->grad0[nc + lin] = dvdy;
      s->grad0[2 * nc + lin] = dudy + dvdx;
    }
  }
  {
    std::vector<double> pr((size_t)s->plane, 0.0), pu((size_t)s->plane, 0.0),
        pv((size_t)s->plane, 0.0);
    for (int j = 0; j < m; j++)
      for (int i = 0; i < n; i++) {
        const size_t idx = (size_t)(j + 1) * s->pitch + (i + 1);
        const size_t lin = (size_t)j * n + i;
        pr[idx] = s->rho0[lin];
        pu[idx] = s->u0[lin];
        pv[idx] = s->v0[lin];
      }
    rc = check_hip(hipMemcpy(s->d_rho, pr.data(), sizeof(double) * s->plane, hipMemcpyHostToDevice), "upload rho");
    if (!rc) rc = check_hip(hipMemcpy(s->d_u, pu.data(), sizeof(double) * s->plane, hipMemcpyHostToDevice), "upload u");
    if (!rc) rc = check_hip(hipMemcpy(s->d_v, pv.data(), sizeof(double) * s->plane, hipMemcpyHostToDevice), "upload v");
    if (!rc && s->d_snap) {
      rc = check_hip(hipMemcpy(s->d_snap, pr.data(), sizeof(double) * s->plane, hipMemcpyHostToDevice), "snap");
      if (!rc) rc = check_hip(hipMemcpy(s->d_snap + s->plane, pu.data(), sizeof(double) * s->plane, hipMemcpyHostToDevice), "snap");
      if (!rc) rc = check_hip(hipMemcpy(s->d_snap + 2 * (size_t)s->plane, pv.data(), sizeof(double) * s->plane, hipMemcpyHostToDevice), "snap");
    }
  }
  if (!rc) rc = check_hip(hipMemcpy(s->d_cls, cls.data(), cls.size(), hipMemcpyHostToDevice), "upload classes");
  if (!rc) rc = check_hip(hipMemcpy(s->d_tflag, tflag.data(), tflag.size(), hipMemcpyHostToDevice), "upload tile flags");
  if (!rc && !wmask.empty()) {
    rc = dev_alloc(s, (void**)&s->d_wmask, sizeof(uint32_t) * wmask.size());
    if (!rc) rc = check_hip(hipMemcpy(s->d_wmask, wmask.data(), sizeof(uint32_t) * wmask.size(), hipMemcpyHostToDevice), "upload wall masks");
    kp.wmask = s->d_wmask;
  }
  if (!rc && s->nbn) rc = check_hip(hipMemcpy(s->d_bn, bn.data(), sizeof(BNode) * bn.size(), hipMemcpyHostToDevice), "upload boundary list");
  if (!rc && s->non) rc = check_hip(hipMemcpy(s->d_on, on.data(), sizeof(BNode) * on.size(), hipMemcpyHostToDevice), "upload outlet list");
  if (!rc) rc = check_hip(hipMemcpy(s->d_orow, orow.data(), sizeof(int) * orow.size(), hipMemcpyHostToDevice), "upload outlet rows");
  if (!rc && s->d_cres) rc = check_hip(hipMemset(s->d_cres, 0, sizeof(double) * 12 * (size_t)m), "clear cres");
  if (!rc) {
    // side slot 0: a solid cell (F = 0)
    std::vector<char> side0(rsz * 9 * (size_t)nside, 0);
    if (s->prec != 0) {
      static const double W[9] = {4. / 9., 1. / 9., 1. / 9., 1. / 9., 1. / 9.,
                                  1. / 36., 1. / 36., 1. / 36., 1. / 36.};
      float* f = (float*)side0.data();
      for (int k = 0; k < 9; k++) f[k] = (float)(-W[k] * (double)o->rho_ini);
    }
    rc = check_hip(hipMemcpy(s->d_side, side0.data(), side0.size(), hipMemcpyHostToDevice), "upload side");
  }
  if (rc) {
    lbm_solver_destroy(s);
    return rc;
  }
  double* d_grad = nullptr;
  if (!rc && s->f0_kind == 3) {
    std::vector<double> pg(3 * (size_t)s->plane, 0.0);
    for (int q = 0; q < 3; q++)
      for (int j = 0; j < m; j++)
        for (int i = 0; i < n; i++)
          pg[(size_t)q * s->plane + (size_t)(j + 1) * s->pitch + (i + 1)] =
              s->grad0[(size_t)q * nc + (size_t)j * n + i];
    rc = check_hip(hipMalloc((void**)&d_grad, pg.size() * sizeof(double)), "hipMalloc gradients");
    if (!rc) rc = check_hip(hipMemcpy(d_grad, pg.data(), pg.size() * sizeof(double), hipMemcpyHostToDevice), "upload gradients");
    if (rc) {
      if (d_grad) (void)hipFree(d_grad);
      lbm_solver_destroy(s);
      return rc;
    }
  }
  s->kp = kp;
  if (ninst > 0) {
    rc = rebuild_tile_list(s);
    if (rc) {
      if (d_grad) (void)hipFree(d_grad);
      lbm_solver_destroy(s);
      return rc;
    }
  }
  if (s->prec != LBM_PREC_F16_SR && !(o->tune_flags & LBM_TUNE_NO_TWO_STEP)) {
    long long active = 0;
    for (uint8_t cv : cls) active += cv != 0;
    rc = setup_two_step(s, cls, BS, align, active, node_inst);
    if (!rc && !s->ninst && !(o->tune_flags & LBM_TUNE_NO_THREE_STEP))
      rc = setup_three_step(s, cls, BS, align, active);
    if (rc) {
      if (d_grad) (void)hipFree(d_grad);
      lbm_solver_destroy(s);
      return rc;
    }
  }
  // ---- fill both buffers with the solid value, then the first collision
  Launch L = make_launch(s, 1);  // dst = buf[0]
  L.out.grad = d_grad;
  hipError_t e = s->prec == 0 ? launch_fill_solid_f64(L, s->buf[0]) : launch_fill_solid_lp(s->prec, L, s->buf[0]);
  if (e == hipSuccess) e = s->prec == 0 ? launch_fill_solid_f64(L, s->buf[1]) : launch_fill_solid_lp(s->prec, L, s->buf[1]);
  if (e == hipSuccess && s->buf3)
    e = s->prec == 0 ? launch_fill_solid_f64(L, s->buf3) : launch_fill_solid_lp(s->prec, L, s->buf3);
  if (e == hipSuccess && s->buf4)
    e = s->prec == 0 ? launch_fill_solid_f64(L, s->buf4) : launch_fill_solid_lp(s->prec, L, s->buf4);
  if (e == hipSuccess) e = s->prec == 0 ? launch_init_f64(s->coll, s->f0_kind, L) : launch_init_lp(s->prec, s->coll, s->f0_kind, L);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  if (d_grad) (void)hipFree(d_grad);
  rc = check_hip(e, "initial collision");
  if (rc) {
    lbm_solver_destroy(s);
    return rc;
  }
  s->cur = 0;
  s->iter = 0;
  s->resid[0] = 100.0;  // serial/Iterate.c:259-260
  s->resid[1] = 100.0;
  *out = s;
  return LBM_OK;
}

// Plain step (no residual work) of one buffer parity: the launches a graph holds.
static int launch_plain_step(lbm_solver* s, int src_buf) {
  Launch L = make_launch(s, src_buf);
  L.kp.res_kind = RK_NONE;
  return launch_step_passes(s, L);
}

// One step of the launches in L (bulk tiles + boundary list, then outlets).
static int launch_step_passes(lbm_solver* s, const Launch& L) {
  if (s->fused) {
    HC(pass(s, MODE_STEP, PASS_FUSED, L));
  } else {
    HC(pass(s, MODE_STEP, PASS_BULK, L));
    HC(pass(s, MODE_STEP, PASS_BOUNDARY, L));
  }
  HC(pass(s, MODE_STEP, PASS_OUTLET, L));
  return LBM_OK;
}

// Two plain steps from buf[src_buf] into buf[src_buf ^ 1]: the two-step
// tiles (between events e0 and e1 when given) and the band's first single
// step into buf3 in one launch (the tiles read only buf[src_buf] and write
// only their own nodes; the band's first step reads buf[src_buf] and writes
// only buf3), then the band's second step from buf3.  (Running the band on
// a second stream beside the tiles measured no faster: its workgroups then
// share the CUs with the tiles.)
static int launch_two_steps(lbm_solver* s, int src_buf, hipEvent_t e0 = nullptr,
                            hipEvent_t e1 = nullptr) {
  Launch L = make_launch(s, src_buf);
  L.kp.res_kind = RK_NONE;
  L.pairs = false;
  L.tiles = s->d_tiles;
  L.ntiles = s->ntiles;
  // with the fused single-step layout (256-node bulk tiles) the band's first
  // step runs as the first workgroups of the two-step launch
  const bool fold = s->fused;
  {
    Launch T = L;
    T.cls = s->d_cls_two;
    if (fold) {
      T.bgrid = s->kp.nb_bnd + s->nlist_b[0];
      T.nband = (T.bgrid + 7) / 8 * 8;
      T.bdst = s->buf3;
      T.bcls = s->d_cls_b[0];
      T.btflag = s->d_tflag_b[0];
      T.blist = s->d_list_b[0];
    }
    if (e0) HC(hipEventRecord(e0, s->stream));
    HC(pass(s, MODE_STEP, PASS_TWO, T));
    if (e1) HC(hipEventRecord(e1, s->stream));
  }
  for (int b = fold ? 1 : 0; b < 2; b++) {
    Launch B = L;
    B.src = b == 0 ? s->buf[src_buf] : s->buf3;
    B.dst = b == 0 ? s->buf3 : s->buf[src_buf ^ 1];
    B.cls = s->d_cls_b[b];
    B.tflag = s->d_tflag_b[b];
    B.kp.tile_list = s->d_list_b[b];
    B.nb_bulk = s->nlist_b[b];
    int rc = LBM_OK;
    if (b == 1 && fold) {
      // the first step's open-outlet rows, after the feeders it folded in
      Launch O = B;
      O.src = s->buf[src_buf];
      O.dst = s->buf3;
      HC(pass(s, MODE_STEP, PASS_OUTLET, O));
    }
    rc = launch_step_passes(s, B);
    if (rc) return rc;
  }
  return LBM_OK;
}

// The launches of launch_two_steps(parity) as a graph: after a replay the
// state is in buf[parity ^ 1] two iterations later.
static int build_graph2(lbm_solver* s, int parity) {
  hipGraph_t g = nullptr;
  HC(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
  int rc = launch_two_steps(s, parity);
  hipError_t e = hipStreamEndCapture(s->stream, &g);
  if (rc != LBM_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  HC(e);
  e = hipGraphInstantiate(&s->gexec2[parity], g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  return check_hip(e, "hipGraphInstantiate");
}

// Three plain steps from buf[src_buf] into buf[src_buf ^ 1]: the three-step
// tiles (between events e0 and e1 when given) and the band's first single
// step (band + two rings, into buf3) in one launch, then the band's second
// (band + one ring, buf3 -> buf4) and third (band, buf4 -> buf[src_buf ^ 1]).
static int launch_three_steps(lbm_solver* s, int src_buf, hipEvent_t e0 = nullptr,
                              hipEvent_t e1 = nullptr) {
  Launch L = make_launch(s, src_buf);
  L.kp.res_kind = RK_NONE;
  L.pairs = false;
  const bool fold = s->fused;
  {
    Launch T = L;
    T.cls = s->d_cls_three;
    T.tiles = s->d_tiles3;
    T.ntiles = s->ntiles3;
    if (fold) {
      T.bgrid = s->kp.nb_bnd + s->nlist_b3[0];
      T.nband = (T.bgrid + 7) / 8 * 8;
      T.bdst = s->buf3;
      T.bcls = s->d_cls_b3[0];
      T.btflag = s->d_tflag_b3[0];
      T.blist = s->d_list_b3[0];
    }
    if (e0) HC(hipEventRecord(e0, s->stream));
    HC(pass(s, MODE_STEP, PASS_THREE, T));
    if (e1) HC(hipEventRecord(e1, s->stream));
  }
  void* bsrc[3] = {s->buf[src_buf], s->buf3, s->buf4};
  void* bdst[3] = {s->buf3, s->buf4, s->buf[src_buf ^ 1]};
  for (int b = fold ? 1 : 0; b < 3; b++) {
    Launch B = L;
    B.src = bsrc[b];
    B.dst = bdst[b];
    B.cls = s->d_cls_b3[b];
    B.tflag = s->d_tflag_b3[b];
    B.kp.tile_list = s->d_list_b3[b];
    B.nb_bulk = s->nlist_b3[b];
    if (b == 1 && fold) {
      // the first step's open-outlet rows, after the feeders it folded in
      Launch O = B;
      O.src = s->buf[src_buf];
      O.dst = s->buf3;
      HC(pass(s, MODE_STEP, PASS_OUTLET, O));
    }
    int rc = launch_step_passes(s, B);
    if (rc) return rc;
  }
  return LBM_OK;
}

static int build_graph3(lbm_solver* s, int parity) {
  hipGraph_t g = nullptr;
  HC(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
  int rc = launch_three_steps(s, parity);
  hipError_t e = hipStreamEndCapture(s->stream, &g);
  if (rc != LBM_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  HC(e);
  e = hipGraphInstantiate(&s->gexec3[parity], g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  return check_hip(e, "hipGraphInstantiate");
}

// Capture steps (parity, parity^1): after replay the state is back in
// buf[parity] two iterations later.  Kernel nodes only (no memset nodes).
static int build_graph(lbm_solver* s, int parity) {
  hipGraph_t g = nullptr;
  HC(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
  int rc = launch_plain_step(s, parity);
  if (rc == LBM_OK) rc = launch_plain_step(s, parity ^ 1);
  hipError_t e = hipStreamEndCapture(s->stream, &g);
  if (rc != LBM_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  HC(e);
  e = hipGraphInstantiate(&s->gexec[parity], g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  return check_hip(e, "hipGraphInstantiate");
}

extern "C" {

int lbm_solver_create(const lbm_case* c, const lbm_options* o, lbm_solver** out) {
  if (o && (o->init == LBM_INIT_FIELD || o->init == LBM_INIT_FIELD_NEQ)) {
    set_error("LBM_INIT_FIELD needs lbm_solver_create_field");
    return LBM_EINVAL;
  }
  return solver_create(c, o, nullptr, nullptr, nullptr, nullptr, c ? c->ninst : 0, out);
}

int lbm_solver_create_field(const lbm_case* c, const lbm_options* o, const double* rho,
                            const double* u, const double* v, lbm_solver** out) {
  if (!o) return LBM_EINVAL;
  lbm_options oo = *o;
  if (oo.init != LBM_INIT_FIELD_NEQ) oo.init = LBM_INIT_FIELD;
  return solver_create(c, &oo, rho, u, v, nullptr, c ? c->ninst : 0, out);
}

int lbm_solver_create_batched(const lbm_case* c, const lbm_options* o, const lbm_instance* inst,
                              int count, lbm_solver** out) {
  if (!c || !o || !inst || count <= 0) return LBM_EINVAL;
  if (o->init == LBM_INIT_FIELD || o->init == LBM_INIT_FIELD_NEQ) {
    set_error("LBM_INIT_FIELD needs rho/u/v arrays (lbm_solver_create_batched_field)");
    return LBM_EINVAL;
  }
  return solver_create(c, o, nullptr, nullptr, nullptr, inst, count, out);
}

int lbm_solver_create_batched_field(const lbm_case* c, const lbm_options* o,
                                    const lbm_instance* inst, int count, const double* rho,
                                    const double* u, const double* v, lbm_solver** out) {
  if (!c || !o || !inst || count <= 0 || !rho || !u || !v) return LBM_EINVAL;
  lbm_options oo = *o;
  if (oo.init != LBM_INIT_FIELD_NEQ) oo.init = LBM_INIT_FIELD;
  return solver_create(c, &oo, rho, u, v, inst, count, out);
}

void lbm_solver_destroy(lbm_solver* s) {
  if (!s) return;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  void* ptrs[] = {s->buf[0], s->buf[1], s->d_cls, s->d_tflag, s->d_bn, s->d_on, s->d_orow, s->d_side,
                  s->d_rho, s->d_u, s->d_v, s->d_f, s->d_metaf, s->d_snap, s->d_cres,
                  s->d_part, s->d_hist, s->d_inst_par, s->d_tile_inst, s->d_node_inst,
                  s->d_inst_active, s->d_inst_rect, s->d_inst_res, s->d_tile_list,
                  s->d_wmask, s->d_seed, s->d_tiles, s->d_cls_two, s->buf3, s->d_cls_b[0], s->d_cls_b[1],
                  s->d_tflag_b[0], s->d_tflag_b[1], s->d_list_b[0], s->d_list_b[1],
                  s->d_tiles3, s->d_cls_three, s->buf4, s->d_cls_b3[0], s->d_cls_b3[1],
                  s->d_cls_b3[2], s->d_tflag_b3[0], s->d_tflag_b3[1], s->d_tflag_b3[2],
                  s->d_list_b3[0], s->d_list_b3[1], s->d_list_b3[2]};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (hipEvent_t e : s->ev) (void)hipEventDestroy(e);
  for (hipGraphExec_t g : s->gexec)
    if (g) (void)hipGraphExecDestroy(g);
  for (hipGraphExec_t g : s->gexec2)
    if (g) (void)hipGraphExecDestroy(g);
  for (hipGraphExec_t g : s->gexec3)
    if (g) (void)hipGraphExecDestroy(g);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

static double wall_seconds() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Tile list of a batched solver: the bulk tiles that hold a node of a running
// instance, padded with 0xffffffff to a multiple of the XCD grouping so the
// grouped launch order stays a permutation.  Captured step graphs are dropped
// (their grid size changes).
static int rebuild_tile_list(lbm_solver* s) {
  std::vector<uint32_t> list;
  list.reserve((size_t)s->nb_bulk);
  for (int t = 0; t < s->nb_bulk; t++) {
    bool run = false;
    for (uint32_t e = s->tile_first[t]; e < s->tile_first[t + 1] && !run; e++)
      run = s->inst_active[s->tile_ids[e]] != 0;
    if (run) list.push_back((uint32_t)t);
  }
  const size_t a = (size_t)s->list_align;
  while (list.size() % a) list.push_back(0xffffffffu);
  s->nb_list = (int)list.size();
  if (!list.empty())
    HC(hipMemcpyAsync(s->d_tile_list, list.data(), list.size() * sizeof(uint32_t),
                      hipMemcpyHostToDevice, s->stream));
  HC(hipStreamSynchronize(s->stream));
  for (hipGraphExec_t& g : s->gexec)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
  if (s->two) return rebuild_two_tiles(s);
  return LBM_OK;
}

// Per-instance velocity residuals of a batched solver after a residual step:
// instances below crit_v stop advancing (their state is copied into the other
// buffer so that later steps and observations see one fixed state).
static int batch_eval(lbm_solver* s) {
  const int ni = s->ninst;
  HC(launch_inst_resid(s->d_snap, s->plane, s->pitch, s->d_inst_rect, s->d_inst_active, ni,
                       s->d_inst_res, s->stream));
  std::vector<double> raw(2 * (size_t)ni);
  HC(hipMemcpyAsync(raw.data(), s->d_inst_res, raw.size() * sizeof(double), hipMemcpyDeviceToHost,
                    s->stream));
  HC(hipStreamSynchronize(s->stream));
  const double now = wall_seconds() - s->t_first_step;
  const double K = (double)s->opt.resid_every;
  const double crit = s->opt.crit_v;
  double worst = 0.0;
  bool changed = false;
  std::vector<int> rect(4 * (size_t)ni);
  for (int k = 0; k < ni; k++) {
    if (!s->inst_active[k]) continue;
    const double a = raw[2 * k], b = raw[2 * k + 1];
    const double r = b > 0 ? std::sqrt(a) / (K * std::sqrt(b)) : std::sqrt(a);
    s->inst_res[k] = r;
    if (!(r == r)) {  // this instance diverged: stop it, the others go on
      s->inst_active[k] = 0;
      s->inst_conv_iter[k] = -2;
      s->inst_conv_s[k] = now;
      changed = true;
      continue;
    }
    if (crit > 0 && r < crit) {
      s->inst_active[k] = 0;
      s->inst_conv_iter[k] = s->iter;
      s->inst_conv_s[k] = now;
      changed = true;
      const int x0 = s->cs->inst_rect[4 * k] + 1, y0 = s->cs->inst_rect[4 * k + 1] + 1;
      const int w = s->cs->inst_rect[4 * k + 2], h = s->cs->inst_rect[4 * k + 3];
      char* from = (char*)s->buf[s->cur];
      char* to = (char*)s->buf[s->cur ^ 1];
      for (int q = 0; q < 9; q++) {
        const size_t off = ((size_t)q * s->plane + (size_t)y0 * s->pitch + x0) * s->elem;
        HC(hipMemcpy2DAsync(to + off, s->pitch * s->elem, from + off, s->pitch * s->elem,
                            (size_t)w * s->elem, (size_t)h, hipMemcpyDeviceToDevice, s->stream));
      }
      continue;
    }
    worst = std::max(worst, r);
  }
  if (changed) {
    HC(hipMemcpyAsync(s->d_inst_active, s->inst_active.data(), (size_t)ni, hipMemcpyHostToDevice,
                      s->stream));
    HC(hipStreamSynchronize(s->stream));
    int rc = rebuild_tile_list(s);
    if (rc) return rc;
  }
  s->resid[0] = s->resid[1] = worst;
  s->resid[2] = s->resid[3] = 0.0;
  s->hist_iter.push_back(s->iter);
  s->hist_val.insert(s->hist_val.end(), s->resid, s->resid + 4);
  return LBM_OK;
}

// evs: timing events, two per sample; rec receives the event-pair index of
// every sample taken.
static int step_impl(lbm_solver* s, int nsteps, int* done, std::vector<hipEvent_t>* evs,
                     int sample_every = 1, std::vector<size_t>* rec = nullptr) {
  if (s->ninst && s->t_first_step < 0) s->t_first_step = wall_seconds();
  if (done) *done = 0;
  if (s->invalid) {
    set_error("solver state invalid: a checkpoint load failed part-way (load one again)");
    return LBM_EINVAL;
  }
  if (s->diverged) {
    set_error("solver diverged");
    return LBM_DIVERGED;
  }
  const lbm_options& o = s->opt;
  const bool stop_can_trigger = o.use_stop_rule && (o.crit_v >= 0 || o.crit_rho >= 0);
  const int K = o.resid_every;
  int ran = 0;
  for (int t = 0; t < nsteps; t++) {
    if (o.use_stop_rule) {
      if (stop_can_trigger) {
        int rc = flush_history(s);
        if (rc) return rc;
      }
      // serial/Iterate.c:263
      if (!(s->resid[0] > o.crit_v && s->resid[1] > o.crit_rho && s->iter < o.max_iterations))
        break;
    }
    const int rk = o.resid_kind;
    auto evals_at = [&](int it) {  // does the step that ends at iteration it+1 evaluate?
      return rk == LBM_RESID_REFERENCE ||
             ((rk == LBM_RESID_FULL || rk == LBM_RESID_VELOCITY) && (it + 1) % K == 0);
    };
    const bool reval = evals_at(s->iter);
    auto forces_at = [&](int it) {
      return o.force_id != 0 && o.force_every > 0 && (it + 1) % o.force_every == 0;
    };
    const bool feval = forces_at(s->iter);
    const bool eval = reval || feval;
    // The stop rule reads only evaluated residuals, which change at
    // evaluation steps alone, and a run of plain steps holds none: the rule
    // cannot end the run inside it, so runs of two- and three-step launches
    // form under a stop rule too (the iteration cap aside), on single
    // lattices that have them.  Otherwise a stop rule keeps single launches:
    // a batched lattice's band lists cover every instance, stopped or not, so
    // in the sweep's tail multi-step launches cost more than they save
    // (256-cavity sweep 25.6 k against 51.4 k MLUPS aggregate), and a small
    // lattice's launches are faster direct than replayed two to a graph (129^2
    // under its stop rule 8.1 against 5.7 us per step).
    const bool runs_ok = !stop_can_trigger || (!s->ninst && (s->two || s->three));
    const bool plain_pair = t + 1 < nsteps && !eval && !evals_at(s->iter + 1) &&
                            !forces_at(s->iter + 1) && runs_ok &&
                            (!o.use_stop_rule || s->iter + 2 <= o.max_iterations);
    const bool sample_pair = evs && (t % sample_every == 0 || (t + 1) % sample_every == 0);
    // Three plain steps in a row with three-step launches: one pass over the state.
    const bool plain_triple = plain_pair && t + 2 < nsteps && !evals_at(s->iter + 2) &&
                              !forces_at(s->iter + 2) &&
                              (!o.use_stop_rule || s->iter + 3 <= o.max_iterations);
    if (s->three && plain_triple) {
      const bool sample = evs && (t % sample_every == 0 || (t + 1) % sample_every == 0 ||
                                  (t + 2) % sample_every == 0);
      if (sample) {
        const int ts = t % sample_every == 0 ? t : ((t + 1) % sample_every == 0 ? t + 1 : t + 2);
        const size_t ei = 2 * (size_t)(ts / sample_every);
        int rc = launch_three_steps(s, s->cur, (*evs)[ei], (*evs)[ei + 1]);
        if (rc) return rc;
        if (rec) rec->push_back(ei);
      } else if (s->use_graph) {
        if (!s->gexec3[s->cur]) {  // both parities at once: the other is next
          int rc = build_graph3(s, s->cur);
          if (!rc && !s->gexec3[s->cur ^ 1]) rc = build_graph3(s, s->cur ^ 1);
          if (rc) return rc;
        }
        HC(hipGraphLaunch(s->gexec3[s->cur], s->stream));
      } else {
        int rc = launch_three_steps(s, s->cur);
        if (rc) return rc;
      }
      s->cur ^= 1;
      s->iter += 3;
      ran += 3;
      t += 2;
      s->prev_stale = true;
      s->prev_lag = 3;
      continue;
    }
    // Two plain steps in a row with two-step launches: one pass over the state.
    if (s->two && plain_pair) {
      if (sample_pair && !s->three) {  // with three-step launches only those are sampled
        const size_t ei = 2 * (size_t)((t % sample_every == 0 ? t : t + 1) / sample_every);
        int rc = launch_two_steps(s, s->cur, (*evs)[ei], (*evs)[ei + 1]);
        if (rc) return rc;
        if (rec) rec->push_back(ei);
      } else if (s->use_graph) {
        if (!s->gexec2[s->cur]) {  // both parities at once
          int rc = build_graph2(s, s->cur);
          if (!rc && !s->gexec2[s->cur ^ 1]) rc = build_graph2(s, s->cur ^ 1);
          if (rc) return rc;
        }
        HC(hipGraphLaunch(s->gexec2[s->cur], s->stream));
      } else {
        int rc = launch_two_steps(s, s->cur);
        if (rc) return rc;
      }
      s->cur ^= 1;
      s->iter += 2;
      ran += 2;
      t += 1;
      s->prev_stale = true;
      s->prev_lag = 2;
      continue;
    }
    s->prev_stale = false;  // the launches below leave the previous state in buf[cur^1]
    // Two plain steps in a row: replay the captured graph.
    if (s->use_graph && plain_pair && !sample_pair) {
      if (!s->gexec[s->cur]) {
        int rc = build_graph(s, s->cur);
        if (rc) return rc;
      }
      HC(hipGraphLaunch(s->gexec[s->cur], s->stream));
      s->iter += 2;
      ran += 2;
      t += 1;
      continue;
    }
    const bool rpart = reval && (rk == LBM_RESID_FULL || rk == LBM_RESID_VELOCITY);
    const int mode = (rpart || feval) ? MODE_RESID : MODE_STEP;
    Launch L = make_launch(s, s->cur);
    L.kp.res_kind = rpart ? (rk == LBM_RESID_FULL ? RK_FULL : RK_VEL) : RK_NONE;
    L.kp.force_on = feval ? 1 : 0;
    if (rk == LBM_RESID_REFERENCE) L.out.cres = s->d_cres;
    // with multi-step launches only those are sampled
    const bool timed = evs && !s->two && !s->three && (t % sample_every == 0);
    const size_t ei = 2 * (size_t)(t / (sample_every > 0 ? sample_every : 1));
    if (timed) HC(hipEventRecord((*evs)[ei], s->stream));
    if (timed && rec) rec->push_back(ei);
    if (s->fused) {
      HC(pass(s, mode, PASS_FUSED, L));  // bulk tiles + boundary-list nodes, one launch
      if (timed) HC(hipEventRecord((*evs)[ei + 1], s->stream));
    } else {
      HC(pass(s, mode, PASS_BULK, L));
      if (timed) HC(hipEventRecord((*evs)[ei + 1], s->stream));
      HC(pass(s, mode, PASS_BOUNDARY, L));
    }
    HC(pass(s, mode, PASS_OUTLET, L));
    const bool row_r = reval && !s->ninst;  // batched residuals are per instance
    if (row_r || feval) {
      double* slot = s->d_hist + (size_t)s->pending * HW;
      if (mode == MODE_RESID) HC(launch_reduce_partials(s->d_part, s->nparts, slot, s->stream));
      if (row_r && rk == LBM_RESID_REFERENCE)
        HC(launch_ref_residual(s->d_cres, s->cs->m, s->cs->n, slot, s->stream));
      s->pending_iter.push_back(s->iter);
      s->pending_what.push_back((row_r ? 1 : 0) | (feval ? 2 : 0));
      s->pending++;
    }
    s->cur ^= 1;
    s->iter++;
    ran++;
    if (s->ninst && reval) {
      int rc = batch_eval(s);
      if (rc) {
        if (done) *done = ran;
        return rc;
      }
    }
    if (s->pending == HCAP || (reval && stop_can_trigger)) {
      int rc = flush_history(s);
      if (rc) {
        if (done) *done = ran;
        return rc;
      }
    }
  }
  int rc = flush_history(s);
  if (done) *done = ran;
  s->timing.steps += ran;
  return rc;
}

int lbm_solver_step(lbm_solver* s, int nsteps, int* done) {
  if (!s || nsteps < 0) return LBM_EINVAL;
  return step_impl(s, nsteps, done, nullptr);
}

int lbm_solver_step_timed(lbm_solver* s, int nsteps, int sample_every, double* total_ms,
                          double* kernel_ms) {
  if (!s || nsteps < 0) return LBM_EINVAL;
  if (sample_every < 1) sample_every = 1;
  const int nsamp = (nsteps + sample_every - 1) / sample_every;
  while ((int)s->ev.size() < 2 * nsamp + 2) {
    hipEvent_t e;
    HC(hipEventCreate(&e));
    s->ev.push_back(e);
  }
  std::vector<hipEvent_t> evs(s->ev.begin(), s->ev.begin() + 2 * nsamp);
  hipEvent_t e0 = s->ev[2 * nsamp], e1 = s->ev[2 * nsamp + 1];
  HC(hipEventRecord(e0, s->stream));
  int done = 0;
  std::vector<size_t> rec;
  int rc = step_impl(s, nsteps, &done, &evs, sample_every, &rec);
  HC(hipEventRecord(e1, s->stream));
  HC(hipEventSynchronize(e1));
  float ms = 0.f;
  HC(hipEventElapsedTime(&ms, e0, e1));
  double bsum = 0.0;
  int nb = 0;
  for (size_t ei : rec) {
    float b = 0.f;
    HC(hipEventElapsedTime(&b, evs[ei], evs[ei + 1]));
    bsum += b;
    nb++;
  }
  if (total_ms) *total_ms = ms;
  if (kernel_ms) *kernel_ms = nb ? bsum / nb : 0.0;
  s->timing.bulk_ms += bsum;
  return rc;
}

int lbm_solver_sync(lbm_solver* s) {
  if (!s) return LBM_EINVAL;
  return check_hip(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
}

int lbm_solver_iter(const lbm_solver* s) { return s ? s->iter : -1; }

int lbm_solver_residuals(lbm_solver* s, double out[4]) {
  if (!s || !out) return LBM_EINVAL;
  int rc = flush_history(s);
  std::memcpy(out, s->resid, sizeof(double) * 4);
  return rc == LBM_DIVERGED ? LBM_OK : rc;
}

int lbm_solver_residual_history(lbm_solver* s, int cap, int* iters, double* vals, int* count) {
  if (!s || cap < 0 || !count) return LBM_EINVAL;
  (void)flush_history(s);
  const int have = (int)s->hist_iter.size();
  const int k = std::min(cap, have);
  for (int i = 0; i < k; i++) {
    if (iters) iters[i] = s->hist_iter[i];
    if (vals) std::memcpy(vals + 4 * i, &s->hist_val[(size_t)4 * i], sizeof(double) * 4);
  }
  s->hist_iter.erase(s->hist_iter.begin(), s->hist_iter.begin() + k);
  s->hist_val.erase(s->hist_val.begin(), s->hist_val.begin() + 4 * (size_t)k);
  *count = k;
  return LBM_OK;
}

int lbm_solver_force_history(lbm_solver* s, int cap, int* iters, double* fxy, int* count) {
  if (!s || cap < 0 || !count) return LBM_EINVAL;
  (void)flush_history(s);
  const int have = (int)s->force_iter.size();
  const int k = std::min(cap, have);
  for (int i = 0; i < k; i++) {
    if (iters) iters[i] = s->force_iter[i];
    if (fxy) {
      fxy[2 * i] = s->force_val[(size_t)2 * i];
      fxy[2 * i + 1] = s->force_val[(size_t)2 * i + 1];
    }
  }
  s->force_iter.erase(s->force_iter.begin(), s->force_iter.begin() + k);
  s->force_val.erase(s->force_val.begin(), s->force_val.begin() + 2 * (size_t)k);
  *count = k;
  return LBM_OK;
}

int lbm_solver_get_macro(lbm_solver* s, double* rho, double* u, double* v) {
  if (!s) return LBM_EINVAL;
  const long nc = (long)s->cs->n * s->cs->m;
  if (s->iter == 0) {
    if (rho) std::memcpy(rho, s->rho0.data(), sizeof(double) * nc);
    if (u) std::memcpy(u, s->u0.data(), sizeof(double) * nc);
    if (v) std::memcpy(v, s->v0.data(), sizeof(double) * nc);
    return LBM_OK;
  }
  int rc = observe(s, false);
  if (rc) return rc;
  std::vector<double> tmp((size_t)s->plane);
  double* outs[3] = {rho, u, v};
  double* dev[3] = {s->d_rho, s->d_u, s->d_v};
  for (int a = 0; a < 3; a++) {
    if (!outs[a]) continue;
    HC(hipMemcpyAsync(tmp.data(), dev[a], sizeof(double) * s->plane, hipMemcpyDeviceToHost, s->stream));
    HC(hipStreamSynchronize(s->stream));
    depad(s, tmp.data(), outs[a], 0);
  }
  return LBM_OK;
}

int lbm_solver_get_f(lbm_solver* s, double* f, double* metaf) {
  if (!s) return LBM_EINVAL;
  const lbm_case* c = s->cs;
  const long nc = (long)c->n * c->m;
  static const double W[9] = {4. / 9., 1. / 9., 1. / 9., 1. / 9., 1. / 9.,
                              1. / 36., 1. / 36., 1. / 36., 1. / 36.};
  if (s->iter == 0) {  // the state CellIni leaves: F0, METAF = 0
    for (long i = 0; i < nc; i++)
      for (int k = 0; k < 9; k++) {
        double v = 0.0;
        if (c->fluid[i] == 1 && s->f0_kind == 1) v = W[k] * s->rho0[i];
        if (c->fluid[i] == 1 && (s->f0_kind == 2 || s->f0_kind == 3)) {
          static const int CX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
          static const int CY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
          const double U = s->u0[i], V = s->v0[i];
          const double T1 = U * U + V * V, T2 = U * CX[k] + V * CY[k];
          v = s->rho0[i] * W[k] * (1.0 + 3.0 * T2 + 4.5 * T2 * T2 - 1.5 * T1);
          if (s->f0_kind == 3) {
            const double cx = CX[k], cy = CY[k];
            const double q = (cx * cx - 1.0 / 3.0) * s->grad0[i] +
                             (cy * cy - 1.0 / 3.0) * s->grad0[nc + i] + cx * cy * s->grad0[2 * nc + i];
            const int in = c->ninst ? c->cell_inst[i] : -1;
            const double om = in >= 0 ? 1.0 / (3. * (double)s->inst_visc[in] + 0.5) : s->kp.omega;
            v -= 3.0 / om * W[k] * s->rho0[i] * q;
          }
        }
        if (f) f[i * 9 + k] = v;
        if (metaf) metaf[i * 9 + k] = 0.0;
      }
    return LBM_OK;
  }
  int rc = observe(s, true);
  if (rc) return rc;
  std::vector<double> tmp((size_t)9 * s->plane), plane_out((size_t)nc);
  double* outs[2] = {f, metaf};
  double* dev[2] = {s->d_f, s->d_metaf};
  for (int a = 0; a < 2; a++) {
    if (!outs[a]) continue;
    HC(hipMemcpyAsync(tmp.data(), dev[a], sizeof(double) * 9 * (size_t)s->plane, hipMemcpyDeviceToHost, s->stream));
    HC(hipStreamSynchronize(s->stream));
    for (int k = 0; k < 9; k++) {
      depad(s, tmp.data(), plane_out.data(), (long)k * s->plane);
      for (long i = 0; i < nc; i++) outs[a][i * 9 + k] = c->fluid[i] == 1 ? plane_out[i] : 0.0;
    }
  }
  return LBM_OK;
}

int lbm_solver_timing(const lbm_solver* s, lbm_timing* out) {
  if (!s || !out) return LBM_EINVAL;
  *out = s->timing;
  return LBM_OK;
}

size_t lbm_solver_device_bytes(const lbm_solver* s) { return s ? s->dev_bytes : 0; }

int lbm_solver_fused(const lbm_solver* s) {
  if (!s) return 0;
  return (s->fused ? 1 : 0) | (make_launch(const_cast<lbm_solver*>(s), 0).pairs ? 2 : 0) |
         (s->two ? 4 : 0) | (s->three ? 8 : 0);
}

int lbm_solver_two_step_info(const lbm_solver* s, long long* cells, int* tiles, int* rows,
                             long long* band) {
  if (!s) return LBM_EINVAL;
  if (cells) *cells = s->two ? s->two_cells : 0;
  if (tiles) *tiles = s->two ? s->ntiles_real : 0;
  if (rows) *rows = s->two ? s->two_rows : 0;
  if (band) *band = s->two ? s->band_cells : 0;
  return LBM_OK;
}

int lbm_solver_three_step_info(const lbm_solver* s, long long* cells, int* tiles, int* rows,
                               long long* band) {
  if (!s) return LBM_EINVAL;
  if (cells) *cells = s->three ? s->three_cells : 0;
  if (tiles) *tiles = s->three ? s->ntiles3_real : 0;
  if (rows) *rows = s->three ? s->three_rows : 0;
  if (band) *band = s->three ? s->band3_cells : 0;
  return LBM_OK;
}

// ---------------------------------------------------------------- checkpoint
namespace {
// Everything that fixes how the stored words map to distributions (lattice,
// layout, storage kind, rho0 of the deviation form, FP16 scale) and what the
// saved counters mean (collision, residual kind and interval).  A file loads
// only into a solver that agrees on all of them.
struct CkptHeader {
  char magic[8];  // "LBMCKPT2"
  int abi, n, m, prec, coll, pitch, plane, iter, cur, resid_kind, has_snap;
  int resid_every;
  double resid[4];
  double rho0;      // rho_ini: FP32/FP16 store f - w_k rho0
  float f16_scale;  // FP16 stores (f - w_k rho0) * scale
  int pad;
};

CkptHeader ckpt_header(const lbm_solver* s) {
  CkptHeader h;
  std::memset(&h, 0, sizeof h);
  std::memcpy(h.magic, "LBMCKPT2", 8);
  h.abi = LBM_ABI_VERSION;
  h.n = s->cs->n;
  h.m = s->cs->m;
  h.prec = s->prec;
  h.coll = s->coll;
  h.pitch = s->pitch;
  h.plane = s->plane;
  h.iter = s->iter;
  h.cur = s->cur;
  h.resid_kind = s->opt.resid_kind;
  h.has_snap = s->d_snap ? 1 : 0;
  h.resid_every = s->opt.resid_every;
  std::memcpy(h.resid, s->resid, sizeof h.resid);
  h.rho0 = s->kp.rho0;
  h.f16_scale = s->kp.scale_f;
  return h;
}

// Bytes that follow the header: both distribution buffers, then the
// velocity-residual snapshot (rho, u, v planes) when the solver keeps one.
size_t ckpt_payload(const lbm_solver* s) {
  const size_t sb = (size_t)9 * s->plane * s->elem;
  return 2 * sb + (s->d_snap ? sizeof(double) * 3 * (size_t)s->plane : 0);
}

int copy_out(lbm_solver* s, FILE* f, const void* dev, size_t bytes) {
  const size_t CH = (size_t)256 << 20;
  std::vector<char> tmp(std::min(bytes, CH));
  for (size_t off = 0; off < bytes; off += CH) {
    const size_t len = std::min(CH, bytes - off);
    HC(hipMemcpy(tmp.data(), (const char*)dev + off, len, hipMemcpyDeviceToHost));
    if (std::fwrite(tmp.data(), 1, len, f) != len) {
      set_error("checkpoint write failed");
      return LBM_EIO;
    }
  }
  return LBM_OK;
}

int copy_in(lbm_solver* s, FILE* f, void* dev, size_t bytes) {
  (void)s;
  const size_t CH = (size_t)256 << 20;
  std::vector<char> tmp(std::min(bytes, CH));
  for (size_t off = 0; off < bytes; off += CH) {
    const size_t len = std::min(CH, bytes - off);
    if (std::fread(tmp.data(), 1, len, f) != len) {
      set_error("checkpoint file truncated");
      return LBM_EPARSE;
    }
    HC(hipMemcpy((char*)dev + off, tmp.data(), len, hipMemcpyHostToDevice));
  }
  return LBM_OK;
}
}  // namespace

int lbm_solver_save(lbm_solver* s, const char* path) {
  if (!s || !path) return LBM_EINVAL;
  if (s->ninst) {
    set_error("checkpoints hold single-lattice solvers only");
    return LBM_EINVAL;
  }
  int rc = flush_history(s);
  if (rc && rc != LBM_DIVERGED) return rc;
  rc = refresh_previous(s);  // both buffers as single steps leave them
  if (rc) return rc;
  HC(hipStreamSynchronize(s->stream));
  FILE* f = std::fopen(path, "wb");
  if (!f) {
    set_error(std::string("cannot write ") + path);
    return LBM_EIO;
  }
  const CkptHeader h = ckpt_header(s);
  rc = std::fwrite(&h, sizeof h, 1, f) == 1 ? LBM_OK : LBM_EIO;
  const size_t sb = (size_t)9 * s->plane * s->elem;
  if (!rc) rc = copy_out(s, f, s->buf[0], sb);
  if (!rc) rc = copy_out(s, f, s->buf[1], sb);
  if (!rc && s->d_snap) rc = copy_out(s, f, s->d_snap, sizeof(double) * 3 * (size_t)s->plane);
  std::fclose(f);
  return rc;
}

int lbm_solver_load(lbm_solver* s, const char* path) {
  if (!s || !path) return LBM_EINVAL;
  if (s->ninst) {
    set_error("checkpoints hold single-lattice solvers only");
    return LBM_EINVAL;
  }
  FILE* f = std::fopen(path, "rb");
  if (!f) {
    set_error(std::string("cannot open ") + path);
    return LBM_EIO;
  }
  CkptHeader h;
  int rc = std::fread(&h, sizeof h, 1, f) == 1 ? LBM_OK : LBM_EPARSE;
  if (rc) set_error(std::string(path) + ": not a checkpoint (short header)");
  const CkptHeader want = ckpt_header(s);
  if (!rc && std::memcmp(h.magic, want.magic, 8) != 0) {
    set_error(std::string(path) + ": not a checkpoint of this format (magic LBMCKPT2)");
    rc = LBM_EPARSE;
  }
  if (!rc) {
    std::string why;
    if (h.abi != want.abi) why += " abi";
    if (h.n != want.n || h.m != want.m || h.pitch != want.pitch || h.plane != want.plane)
      why += " lattice";
    if (h.prec != want.prec) why += " precision";
    if (h.coll != want.coll) why += " collision";
    if (h.resid_kind != want.resid_kind || h.has_snap != want.has_snap) why += " residual-kind";
    if (h.resid_kind == LBM_RESID_VELOCITY && h.resid_every != want.resid_every)
      why += " residual-interval";
    if (std::memcmp(&h.rho0, &want.rho0, sizeof h.rho0) != 0) why += " rho_ini";
    if (s->elem == 2 && std::memcmp(&h.f16_scale, &want.f16_scale, sizeof(float)) != 0)
      why += " f16_scale";
    if (!why.empty()) {
      set_error("checkpoint does not match this solver:" + why);
      rc = LBM_EINVAL;
    }
  }
  // the whole payload must be there before any device memory is touched: a
  // short or over-long file leaves the solver exactly as it was
  if (!rc) {
    const long long need = (long long)sizeof h + (long long)ckpt_payload(s);
    long long have = -1;
    if (fseeko(f, 0, SEEK_END) == 0) have = (long long)ftello(f);
    if (have != need) {
      set_error(std::string(path) + ": checkpoint holds " + std::to_string(have) +
                " bytes, this solver's state needs " + std::to_string(need));
      rc = LBM_EPARSE;
    } else if (fseeko(f, (off_t)sizeof h, SEEK_SET) != 0) {
      set_error(std::string("cannot seek in ") + path);
      rc = LBM_EIO;
    }
  }
  if (rc) {
    std::fclose(f);
    return rc;
  }
  HC(hipStreamSynchronize(s->stream));
  const size_t sb = (size_t)9 * s->plane * s->elem;
  rc = copy_in(s, f, s->buf[0], sb);
  if (!rc) rc = copy_in(s, f, s->buf[1], sb);
  if (!rc && s->d_snap) rc = copy_in(s, f, s->d_snap, sizeof(double) * 3 * (size_t)s->plane);
  std::fclose(f);
  if (rc) {
    // a read failed after part of the state was overwritten: refuse to step
    s->invalid = true;
    set_error(std::string(lbm_last_error()) + " (solver state invalid until a checkpoint loads)");
    return rc;
  }
  if (s->d_seed) {  // the next step (from buffer cur) is step iter + 1
    const uint32_t next = (uint32_t)h.iter + 1u;
    HC(hipMemcpy(s->d_seed + h.cur, &next, sizeof next, hipMemcpyHostToDevice));
  }
  s->invalid = false;
  s->iter = h.iter;
  s->cur = h.cur;
  s->prev_stale = false;
  std::memcpy(s->resid, h.resid, sizeof h.resid);
  s->observed = s->observed_f = -1;
  s->pending = 0;
  s->pending_iter.clear();
  s->pending_what.clear();
  s->hist_iter.clear();
  s->hist_val.clear();
  s->force_iter.clear();
  s->force_val.clear();
  s->diverged = false;
  return LBM_OK;
}

int lbm_device_copy_gbs(size_t bytes, int reps, double* gbs) {
  if (!gbs || bytes < 16 || reps < 1) return LBM_EINVAL;
  bytes = bytes / 16 * 16;
  void *a = nullptr, *b = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = check_hip(hipMalloc(&a, bytes), "hipMalloc");
  if (!rc) rc = check_hip(hipMalloc(&b, bytes), "hipMalloc");
  if (!rc) rc = check_hip(hipMemset(a, 0, bytes), "hipMemset");
  if (!rc) rc = check_hip(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "stream");
  if (!rc) rc = check_hip(hipEventCreate(&e0), "event");
  if (!rc) rc = check_hip(hipEventCreate(&e1), "event");
  // best launch over every copy shape (vectors in flight per thread, load kind)
  double best = 0.0;
  for (int shape = 0; !rc && shape < COPY_SHAPES; shape++)
    for (int r = 0; !rc && r < reps + 1; r++) {
      rc = check_hip(hipEventRecord(e0, st), "record");
      if (!rc) rc = check_hip(launch_copy_shape(a, b, bytes, shape, st), "copy");
      if (!rc) rc = check_hip(hipEventRecord(e1, st), "record");
      if (!rc) rc = check_hip(hipEventSynchronize(e1), "sync");
      float ms = 0.f;
      if (!rc) rc = check_hip(hipEventElapsedTime(&ms, e0, e1), "elapsed");
      const double g = 2.0 * (double)bytes / (ms * 1e-3) / 1e9;
      if (!rc && r > 0 && g > best) best = g;  // first launch of a shape is a warm-up
    }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  *gbs = best;
  return rc;
}

int lbm_device_stream_probe(int n, int m, int elem_bytes, int reps, double* gbs, double* us) {
  if (!gbs || n < 1 || m < 1 || (elem_bytes != 4 && elem_bytes != 8) || reps < 1)
    return LBM_EINVAL;
  // the solver's padded layout (solver_create): pitch, halo rows, planes
  const int pitch = ((n + 2) + 63) / 64 * 64;
  const long long plane = ((long long)pitch * (m + 2) + 255) / 256 * 256;
  if (plane * 9 >= (1LL << 31)) {
    set_error("lattice too large for 32-bit plane indexing");
    return LBM_EINVAL;
  }
  const size_t bytes = (size_t)9 * (size_t)plane * (size_t)elem_bytes + 256;
  void *a = nullptr, *b = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = check_hip(hipMalloc(&a, bytes), "hipMalloc");
  if (!rc) rc = check_hip(hipMalloc(&b, bytes), "hipMalloc");
  if (!rc) rc = check_hip(hipMemset(a, 0, bytes), "hipMemset");
  if (!rc) rc = check_hip(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "stream");
  if (!rc) rc = check_hip(hipEventCreate(&e0), "event");
  if (!rc) rc = check_hip(hipEventCreate(&e1), "event");
  const int begin = pitch, end = (m + 1) * pitch;
  double sum = 0.0;
  for (int r = 0; !rc && r < reps + 1; r++) {
    rc = check_hip(hipEventRecord(e0, st), "record");
    if (!rc) rc = check_hip(launch_pull_probe(elem_bytes, (r & 1) ? b : a, (r & 1) ? a : b,
                                              (int)plane, pitch, begin, end, 4, st), "probe");
    if (!rc) rc = check_hip(hipEventRecord(e1, st), "record");
    if (!rc) rc = check_hip(hipEventSynchronize(e1), "sync");
    float ms = 0.f;
    if (!rc) rc = check_hip(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    if (!rc && r > 0) sum += ms;  // the first launch is a warm-up
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (rc) return rc;
  const double avg_s = sum / reps * 1e-3;
  *gbs = 2.0 * 9.0 * elem_bytes * (double)(end - begin) / avg_s / 1e9;
  if (us) *us = avg_s * 1e6;
  return LBM_OK;
}

}  // extern "C"

int lbm_batch_status(lbm_solver* s, int count, int* conv_iter, double* conv_seconds,
                     double* resid) {
  if (!s || count < 0) return LBM_EINVAL;
  if (!s->ninst) {
    set_error("not a batched solver");
    return LBM_EINVAL;
  }
  const int k = std::min(count, s->ninst);
  for (int i = 0; i < k; i++) {
    if (conv_iter) conv_iter[i] = s->inst_conv_iter[i];
    if (conv_seconds) conv_seconds[i] = s->inst_conv_s[i];
    if (resid) resid[i] = s->inst_res[i];
  }
  return LBM_OK;
}
