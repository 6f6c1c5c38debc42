// case_io.cpp — case input (ini, mesh files), mesh generators and the
// per-cell set-up of the reference's CellIni, restated in O(N).
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "lbm_host.hpp"

namespace lbmhost {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
const char* last_error() { return g_err.c_str(); }

static bool read_file(const char* path, std::string& out) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  long sz = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  out.resize(sz > 0 ? (size_t)sz : 0);
  size_t got = sz > 0 ? std::fread(&out[0], 1, (size_t)sz, f) : 0;
  std::fclose(f);
  out.resize(got);
  return true;
}

// fscanf("%d") / fscanf("%f") equivalents on a NUL-terminated buffer.
static bool scan_int(const char*& p) {
  char* e;
  errno = 0;
  (void)std::strtol(p, &e, 10);
  if (e == p) return false;
  p = e;
  return true;
}
static bool scan_float(const char*& p, float* out) {
  char* e;
  float v = std::strtof(p, &e);
  if (e == p) return false;
  if (out) *out = v;
  p = e;
  return true;
}

// serial/FileIO.c:14-82 (ReadNodes) and :91-174 (ReadBCconn): count the
// records that parse with the typed pattern, then read that many records
// with float conversions for every field.
// The same tables parsed on several threads, for the usual file shape: one
// record per line, integer fields written as integers.  The buffer is cut at
// line ends; each piece parses its lines with the same strtof calls as the
// sequential reader.  Returns the record count, or -2 when a line does not
// hold exactly one clean record (two records on a line, a record across
// lines, "3.5" in an integer field): the sequential reader then decides,
// since fscanf's counting pass and its reading pass can disagree there.
static int read_table_lines(const std::string& buf, int cols, const int* pat,
                            std::vector<float>& t) {
  const long N = (long)buf.size();
  unsigned nt = std::thread::hardware_concurrency();
  nt = nt == 0 ? 1 : (nt > 16 ? 16 : nt);
  std::vector<long> cut(nt + 1, N);
  cut[0] = 0;
  for (unsigned k = 1; k < nt; k++) {  // piece k starts after a newline
    long q = std::max(cut[k - 1], N * (long)k / (long)nt);
    while (q < N && buf[(size_t)q] != '\n') q++;
    cut[k] = q < N ? q + 1 : N;
  }
  struct Piece {
    std::vector<float> v;
    bool stop = false;   // a field failed to parse: the table ends here
    bool clean = true;   // every record was one clean line
  };
  std::vector<Piece> pc(nt);
  auto work = [&](unsigned k) {
    Piece& P = pc[k];
    const char* p = buf.c_str() + cut[k];
    const char* end = buf.c_str() + cut[k + 1];
    P.v.reserve((size_t)((cut[k + 1] - cut[k]) / 6));
    for (;;) {
      while (p < end && std::isspace((unsigned char)*p)) p++;
      if (p >= end) return;
      float rec[8];
      for (int c = 0; c < cols; c++) {
        char* e;
        if (pat[c] == 0) {
          (void)std::strtol(p, &e, 10);
          if (e == p) {
            P.stop = true;
            return;
          }
          if (*e && !std::isspace((unsigned char)*e)) {  // "3.5", "0x10" in an integer field
            P.clean = false;
            return;
          }
        }
        rec[c] = std::strtof(p, &e);
        if (e == p) {
          P.stop = true;
          return;
        }
        p = e;
        // a record must not run past its line, nor stop before it ends
        const bool last = c == cols - 1;
        const char* q = p;
        while (q < end && *q != '\n' && std::isspace((unsigned char)*q)) q++;
        const bool eol = q >= end || *q == '\n';
        if (eol != last || p > end) {
          P.clean = false;
          return;
        }
      }
      P.v.insert(P.v.end(), rec, rec + cols);
    }
  };
  std::vector<std::thread> th;
  for (unsigned k = 1; k < nt; k++) th.emplace_back(work, k);
  work(0);
  for (auto& x : th) x.join();
  t.clear();
  for (unsigned k = 0; k < nt; k++) {
    if (!pc[k].clean) return -2;
    t.insert(t.end(), pc[k].v.begin(), pc[k].v.end());
    if (pc[k].stop) break;  // fscanf's count stops at the first failing field
  }
  return (int)(t.size() / (size_t)cols);
}

static int read_table(const char* path, int cols, std::vector<float>& t) {
  std::string buf;
  if (!read_file(path, buf)) return -1;
  static const int node_pat5[5] = {0, 0, 1, 1, 0};
  static const int conn_pat7[7] = {0, 0, 0, 0, 1, 1, 0};
  if (buf.size() >= (1u << 20)) {
    const int lines = read_table_lines(buf, cols, cols == 5 ? node_pat5 : conn_pat7, t);
    if (lines >= 0) return lines;
  }
  const char* p = buf.c_str();
  // typed pattern: nodes "%d %d %f %f %d", connectors "%d %d %d %d %f %f %d"
  static const int node_pat[5] = {0, 0, 1, 1, 0};
  static const int conn_pat[7] = {0, 0, 0, 0, 1, 1, 0};
  const int* pat = cols == 5 ? node_pat : conn_pat;
  int lines = 0;
  for (;;) {
    bool ok = true;
    for (int c = 0; c < cols && ok; c++) ok = pat[c] ? scan_float(p, nullptr) : scan_int(p);
    if (!ok) break;
    lines++;
  }
  t.assign((size_t)lines * cols, 0.0f);
  p = buf.c_str();
  for (int L = 0; L < lines; L++)
    for (int c = 0; c < cols; c++) scan_float(p, &t[(size_t)L * cols + c]);
  return lines;
}

// serial/FileIO.c:183-209 (CompDataNode) and :218-254 (CompDataConn, with the
// inlet extent as its MyReal declaration intends, see the oracle).
static void comp_data(lbm_case* c) {
  const float* last = &c->nodes[(size_t)(c->num_nodes - 1) * 5];
  c->n = (int)(last[0] + 1);
  c->m = (int)(last[1] + 1);
  double p1 = 0.0, p2 = 0.0;
  for (int i = 0; i < c->num_nodes; i++) {
    const float* r = &c->nodes[(size_t)i * 5];
    if (r[0] == 0 && r[1] == 0) p1 = r[2];
    if (r[0] == 1 && r[1] == 0) p2 = r[2];
  }
  c->delta = (p1 > p2 ? p1 : p2) - (p1 < p2 ? p1 : p2);
  double mx = 0.0, mn = 0.0;
  bool seen = false;
  c->num_inlet = 0;
  for (int i = 0; i < c->num_conn; i++) {
    const float* r = &c->conn[(size_t)i * 7];
    if (r[3] == 2 && (r[2] == 1 || r[2] == 2 || r[2] == 3 || r[2] == 4)) {
      if (!seen || r[5] > mx) mx = r[5];
      if (!seen || r[5] < mn) mn = r[5];
      seen = true;
      c->num_inlet++;
    }
  }
  c->max_inlet_y = mx + c->delta / 2;
  c->min_inlet_y = mn - c->delta / 2;
}

// serial/CellFunctions.c:99-270 (CellIni) for every cell: the node-ID search
// (:123-131) becomes a dense (i,j) table, the connector search (:162-178) a
// stable bucket per cell; "last matching line wins" is kept in both.
static int cell_init(lbm_case* c) {
  const int n = c->n, m = c->m;
  const long nc = (long)n * m;
  if (n <= 0 || m <= 0 || nc > (1L << 30)) {
    set_error("mesh dimensions out of range");
    return LBM_EPARSE;
  }
  c->id.assign((size_t)nc, 0);
  c->fluid.assign((size_t)nc, 0);
  c->bcell.assign((size_t)nc, -1);
  c->bcells.clear();
  for (int L = 0; L < c->num_nodes; L++) {
    const float fi = c->nodes[(size_t)L * 5], fj = c->nodes[(size_t)L * 5 + 1];
    if (!(fi >= 0) || !(fj >= 0) || fi >= (float)n || fj >= (float)m) continue;
    const int ii = (int)fi, jj = (int)fj;
    if ((float)ii != fi || (float)jj != fj) continue;
    c->id[(size_t)jj * n + ii] = L;
  }
  for (long q = 0; q < nc; q++) c->fluid[q] = (int)c->nodes[(size_t)c->id[q] * 5 + 4];

  // connector buckets (counting sort keeps file order)
  std::vector<int> start((size_t)nc + 1, 0), list((size_t)c->num_conn + 1, 0);
  for (int L = 0; L < c->num_conn; L++) {
    const int ci = (int)c->conn[(size_t)L * 7], cj = (int)c->conn[(size_t)L * 7 + 1];
    if (ci < 0 || cj < 0 || ci >= n || cj >= m) continue;
    start[(size_t)cj * n + ci + 1]++;
  }
  for (long q = 0; q < nc; q++) start[q + 1] += start[q];
  {
    std::vector<int> fill(start.begin(), start.end() - 1);
    for (int L = 0; L < c->num_conn; L++) {
      const int ci = (int)c->conn[(size_t)L * 7], cj = (int)c->conn[(size_t)L * 7 + 1];
      if (ci < 0 || cj < 0 || ci >= n || cj >= m) continue;
      list[fill[(size_t)cj * n + ci]++] = L;
    }
  }
  static const int OPP[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
  const double qlat[9] = {0., 1., 1., 1., 1., std::sqrt(2.0), std::sqrt(2.0), std::sqrt(2.0),
                          std::sqrt(2.0)};
  c->num_fluid = 0;
  c->num_boundary = 0;
  for (long q = 0; q < nc; q++) {
    if (c->fluid[q] == 1) c->num_fluid++;
    bool matched = false;
    BCell b;
    for (int k = 0; k < 9; k++) {
      b.bc[k] = 0;
      b.q[k] = 0.5;
    }
    b.boundary_id = 0;
    const double x = c->x(q), y = c->y(q);
    for (int t = start[q]; t < start[q + 1]; t++) {
      const float* con = &c->conn[(size_t)list[t] * 7];
      for (int j = 1; j < 9; j++)
        if (con[2] == j) {
          matched = true;
          b.bc[j] = (int)con[3];
          b.boundary_id = (int)con[6];
          const double dx = con[4] - x, dy = con[5] - y;
          b.q[j] = std::sqrt(dx * dx + dy * dy) / (c->delta * qlat[j]);
        }
    }
    if (!matched) continue;
    // Boundary / Corner (:180-203)
    b.boundary = 0;
    b.corner = 0;
    for (int j = 1; j < 9; j++) {
      if (b.bc[j] == 0) continue;
      if (b.boundary == 0) {
        b.boundary = b.bc[j];
      } else {
        if (b.boundary < b.bc[j]) { b.boundary = 1; b.corner = 1; }
        if (b.boundary > b.bc[j]) { b.boundary = 1; b.corner = 1; }
      }
    }
    if (b.corner == 1) {  // (:212-231)
      if (b.bc[1] != 0 && b.bc[2] != 0) b.bc[5] = 1;
      if (b.bc[1] != 0 && b.bc[4] != 0) b.bc[8] = 1;
      if (b.bc[2] != 0 && b.bc[3] != 0) b.bc[6] = 1;
      if (b.bc[3] != 0 && b.bc[4] != 0) b.bc[7] = 1;
    }
    b.stream = 0x1ff;  // (:234-249)
    for (int k = 0; k < 9; k++)
      if (b.bc[k] != 0) b.stream &= ~(1 << OPP[k]);
    c->bcell[q] = (int32_t)c->bcells.size();
    c->bcells.push_back(b);
    if (c->fluid[q] == 1) c->num_boundary++;
  }
  return LBM_OK;
}

int finish_case(lbm_case* c) {
  if (c->num_nodes <= 0) {
    set_error("no node records");
    return LBM_EPARSE;
  }
  comp_data(c);
  return cell_init(c);
}

// serial/CellFunctions.c:253-267: Uo of a cell for the inlet profile.
double inlet_uo(const lbm_case* c, const lbm_options* o, long lin) {
  switch (o->inlet_profile) {
    case 1: {
      const double y = c->y(lin), mn = c->min_inlet_y, mx = c->max_inlet_y;
      return 4 * 1.5 * (double)o->uavg * (y - mn) * ((mx - mn) - (y - mn)) /
             ((mx - mn) * (mx - mn));
    }
    case 2: return (double)o->uavg;
    default: return 0.0;
  }
}

// ------------------------------------------------------------------ generators
static const int GCX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
static const int GCY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};

struct GenLink {
  int type, bid;
  double q;
};

template <class FluidFn, class ClassFn>
static void gen_mesh(lbm_case* c, int n, int m, double delta, FluidFn fluid, ClassFn classify) {
  c->nodes.resize((size_t)n * m * 5);
  for (int j = 0; j < m; j++)
    for (int i = 0; i < n; i++) {
      float* r = &c->nodes[((size_t)j * n + i) * 5];
      r[0] = (float)i;
      r[1] = (float)j;
      r[2] = (float)((i - 0.5) * delta);
      r[3] = (float)((j - 0.5) * delta);
      r[4] = fluid(i, j) ? 1.0f : 0.0f;
    }
  c->num_nodes = n * m;
  c->conn.clear();
  for (int j = 0; j < m; j++)
    for (int i = 0; i < n; i++) {
      if (!fluid(i, j)) continue;
      for (int k = 1; k < 9; k++) {
        const int ii = i + GCX[k], jj = j + GCY[k];
        if (ii >= 0 && ii < n && jj >= 0 && jj < m && fluid(ii, jj)) continue;
        const GenLink L = classify(i, j, k, ii, jj);
        const double x = (i - 0.5) * delta, y = (j - 0.5) * delta;
        const float rec[7] = {(float)i, (float)j, (float)k, (float)L.type,
                              (float)(x + L.q * GCX[k] * delta), (float)(y + L.q * GCY[k] * delta),
                              (float)L.bid};
        c->conn.insert(c->conn.end(), rec, rec + 7);
      }
    }
  c->num_conn = (int)(c->conn.size() / 7);
}

}  // namespace lbmhost

using namespace lbmhost;

extern "C" {

const char* lbm_last_error(void) { return lbmhost::last_error(); }
int lbm_abi_version(void) { return LBM_ABI_VERSION; }

int lbm_read_ini(const char* path, lbm_ini* o) {
  if (!path || !o) return LBM_EINVAL;
  FILE* f = std::fopen(path, "r");
  if (!f) {
    set_error(std::string("cannot open ") + path);
    return LBM_EIO;
  }
  int ok = 0;
  ok += std::fscanf(f, "%f", &o->uavg);
  ok += std::fscanf(f, "%f", &o->vavg);
  ok += std::fscanf(f, "%f", &o->rho_ini);
  ok += std::fscanf(f, "%f", &o->viscosity);
  ok += std::fscanf(f, "%d", &o->inlet_profile);
  ok += std::fscanf(f, "%d", &o->collision_model);
  ok += std::fscanf(f, "%d", &o->curved_boundaries);
  ok += std::fscanf(f, "%d", &o->outlet_profile);
  ok += std::fscanf(f, "%d", &o->iterations);
  ok += std::fscanf(f, "%d", &o->autosave_every);
  ok += std::fscanf(f, "%d", &o->autosave_after);
  ok += std::fscanf(f, "%d", &o->postproc_prog);
  ok += std::fscanf(f, "%d", &o->calculate_drag_lift);
  ok += std::fscanf(f, "%f", &o->convergence_crit_veloc);
  ok += std::fscanf(f, "%f", &o->convergence_crit_rho);
  ok += std::fscanf(f, "%d", &o->read_former_data);
  std::fclose(f);
  if (ok != 16) {
    set_error(std::string("ini file has fewer than 16 values: ") + path);
    return LBM_EPARSE;
  }
  return LBM_OK;
}

void lbm_options_from_ini(const lbm_ini* ini, lbm_options* o) {
  std::memset(o, 0, sizeof *o);
  o->collision = ini->collision_model;
  o->curved = ini->curved_boundaries;
  o->inlet_profile = ini->inlet_profile;
  o->outlet_profile = ini->outlet_profile;
  o->drag_lift = ini->calculate_drag_lift;
  o->uavg = ini->uavg;
  o->vavg = ini->vavg;
  o->rho_ini = ini->rho_ini;
  o->viscosity = ini->viscosity;
  o->precision = LBM_PREC_F64;
  o->init = LBM_INIT_REFERENCE;
  o->resid_kind = LBM_RESID_REFERENCE;
  o->resid_every = 1;
  o->use_stop_rule = 1;
  o->crit_v = ini->convergence_crit_veloc;
  o->crit_rho = ini->convergence_crit_rho;
  o->max_iterations = ini->iterations;
}

int lbm_case_load(const char* node_file, const char* bc_file, lbm_case** out) {
  if (!node_file || !bc_file || !out) return LBM_EINVAL;
  *out = nullptr;
  lbm_case* c = new lbm_case();
  c->num_nodes = read_table(node_file, 5, c->nodes);
  if (c->num_nodes < 0) {
    set_error(std::string("Can't open the file ") + node_file);
    delete c;
    return LBM_EIO;
  }
  c->num_conn = read_table(bc_file, 7, c->conn);
  if (c->num_conn < 0) {
    set_error(std::string("Can't open the file ") + bc_file);
    delete c;
    return LBM_EIO;
  }
  int rc = finish_case(c);
  if (rc != LBM_OK) {
    delete c;
    return rc;
  }
  *out = c;
  return LBM_OK;
}

int lbm_case_from_arrays(const float* nodes, int num_nodes, const float* conn, int num_conn,
                         lbm_case** out) {
  if (!out || num_nodes <= 0 || !nodes || num_conn < 0 || (num_conn > 0 && !conn))
    return LBM_EINVAL;
  lbm_case* c = new lbm_case();
  c->num_nodes = num_nodes;
  c->num_conn = num_conn;
  c->nodes.assign(nodes, nodes + (size_t)num_nodes * 5);
  if (num_conn) c->conn.assign(conn, conn + (size_t)num_conn * 7);
  int rc = finish_case(c);
  if (rc != LBM_OK) {
    delete c;
    return rc;
  }
  *out = c;
  return LBM_OK;
}

int lbm_case_cavity(int N, int lid_type, lbm_case** out) {
  if (!out || N < 2 || N > 30000) return LBM_EINVAL;
  lbm_case* c = new lbm_case();
  const int n = N + 2, m = N + 3;
  auto fluid = [N](int i, int j) { return i >= 1 && i <= N && j >= 1 && j <= N; };
  auto classify = [N, lid_type](int, int, int, int ii, int jj) -> GenLink {
    if (jj == 0) return {1, 4, 0.5};
    if (ii == N + 1 && jj == N + 1) return {1, 3, 0.5};
    if (jj == N + 1) return {lid_type, 2, 0.5};
    if (ii == 0) return {1, 1, 0.5};
    return {1, 3, 0.5};
  };
  gen_mesh(c, n, m, 1.0 / N, fluid, classify);
  int rc = finish_case(c);
  if (rc != LBM_OK) {
    delete c;
    return rc;
  }
  *out = c;
  return LBM_OK;
}

int lbm_case_cavity_ex(int NX, int NY, int lid_type, const int* rects, int nrect,
                       lbm_case** out) {
  if (!out || NX < 2 || NY < 2 || NX > 30000 || NY > 30000 || nrect < 0 || (nrect && !rects))
    return LBM_EINVAL;
  lbm_case* c = new lbm_case();
  const int n = NX + 2, m = NY + 3;
  auto in_rect = [=](int i, int j) {
    for (int r = 0; r < nrect; r++)
      if (i >= rects[4 * r] && i <= rects[4 * r + 2] && j >= rects[4 * r + 1] && j <= rects[4 * r + 3])
        return true;
    return false;
  };
  auto fluid = [=](int i, int j) {
    return i >= 1 && i <= NX && j >= 1 && j <= NY && !in_rect(i, j);
  };
  auto classify = [=](int, int, int, int ii, int jj) -> GenLink {
    if (ii >= 1 && ii <= NX && jj >= 1 && jj <= NY) return {1, 5, 0.5};  // obstacle
    if (jj == 0) return {1, 4, 0.5};
    if (ii == NX + 1 && jj == NY + 1) return {1, 3, 0.5};
    if (jj == NY + 1) return {lid_type, 2, 0.5};
    if (ii == 0) return {1, 1, 0.5};
    return {1, 3, 0.5};
  };
  gen_mesh(c, n, m, 1.0 / (NX > NY ? NX : NY), fluid, classify);
  int rc = finish_case(c);
  if (rc != LBM_OK) {
    delete c;
    return rc;
  }
  *out = c;
  return LBM_OK;
}

int lbm_case_pack(const lbm_case* const* cases, int count, lbm_case** out, int* rects) {
  if (!cases || count <= 0 || count > 65000 || !out) return LBM_EINVAL;
  // shelf packing, tallest first
  std::vector<int> order(count);
  long area = 0;
  int maxw = 0;
  for (int k = 0; k < count; k++) {
    if (!cases[k]) return LBM_EINVAL;
    order[k] = k;
    area += (long)cases[k]->n * cases[k]->m;
    maxw = std::max(maxw, cases[k]->n);
  }
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return cases[a]->m > cases[b]->m; });
  const int W = std::max(maxw, (int)std::ceil(std::sqrt((double)area * 1.15)));
  std::vector<int> x0(count), y0(count);
  int x = 0, y = 0, shelf = 0;
  for (int k : order) {
    const lbm_case* c = cases[k];
    if (x + c->n > W) {
      y += shelf;
      x = 0;
      shelf = 0;
    }
    x0[k] = x;
    y0[k] = y;
    x += c->n;
    shelf = std::max(shelf, c->m);
  }
  const int n = W, m = y + shelf;
  if ((long)n * m > (1L << 30)) return LBM_EINVAL;
  lbm_case* p = new lbm_case();
  const long nc = (long)n * m;
  p->n = n;
  p->m = m;
  p->num_nodes = (int)nc;
  p->nodes.assign((size_t)nc * 5, 0.0f);
  p->id.resize((size_t)nc);
  p->fluid.assign((size_t)nc, 0);
  p->bcell.assign((size_t)nc, -1);
  p->cell_inst.assign((size_t)nc, -1);
  p->ninst = count;
  p->inst_rect.resize((size_t)count * 4);
  for (long q = 0; q < nc; q++) {
    p->id[q] = (int32_t)q;
    p->nodes[(size_t)q * 5] = (float)(q % n);
    p->nodes[(size_t)q * 5 + 1] = (float)(q / n);
  }
  for (int k = 0; k < count; k++) {
    const lbm_case* c = cases[k];
    p->inst_rect[4 * k] = x0[k];
    p->inst_rect[4 * k + 1] = y0[k];
    p->inst_rect[4 * k + 2] = c->n;
    p->inst_rect[4 * k + 3] = c->m;
    if (rects) {
      rects[4 * k] = x0[k];
      rects[4 * k + 1] = y0[k];
      rects[4 * k + 2] = c->n;
      rects[4 * k + 3] = c->m;
    }
    for (int j = 0; j < c->m; j++)
      for (int i = 0; i < c->n; i++) {
        const long src = (long)j * c->n + i;
        const long dst = (long)(j + y0[k]) * n + (i + x0[k]);
        float* r = &p->nodes[(size_t)dst * 5];
        r[2] = (float)c->x(src);
        r[3] = (float)c->y(src);
        r[4] = (float)c->fluid[src];
        p->fluid[dst] = c->fluid[src];
        p->cell_inst[dst] = (int16_t)k;
        if (c->bcell[src] >= 0) {
          p->bcell[dst] = (int32_t)p->bcells.size();
          p->bcells.push_back(c->bcells[c->bcell[src]]);
        }
      }
    p->num_fluid += c->num_fluid;
    p->num_boundary += c->num_boundary;
    p->num_conn += c->num_conn;
  }
  p->delta = cases[0]->delta;
  *out = p;
  return LBM_OK;
}

int lbm_case_channel(int NX, int NY, double ocx, double ocy, double r, lbm_case** out) {
  if (!out || NX < 3 || NY < 2) return LBM_EINVAL;
  lbm_case* c = new lbm_case();
  const int n = NX + 2, m = NY + 2;
  auto in_obst = [=](int i, int j) {
    return r > 0 && (i - ocx) * (i - ocx) + (j - ocy) * (j - ocy) <= r * r;
  };
  auto fluid = [=](int i, int j) {
    return i >= 1 && i <= NX && j >= 1 && j <= NY && !in_obst(i, j);
  };
  auto classify = [=](int i, int j, int k, int ii, int jj) -> GenLink {
    if (ii >= 1 && ii <= NX && jj >= 1 && jj <= NY) {  // obstacle link: exact cut
      const double dx = i - ocx, dy = j - ocy;
      const double a = GCX[k] * GCX[k] + GCY[k] * GCY[k];
      const double b = 2 * (dx * GCX[k] + dy * GCY[k]);
      const double cc = dx * dx + dy * dy - r * r;
      double disc = b * b - 4 * a * cc;
      if (disc < 0) disc = 0;
      double t = (-b - std::sqrt(disc)) / (2 * a);
      if (t < 1e-3) t = 1e-3;
      if (t > 1.0) t = 1.0;
      return {1, 5, t};
    }
    if (jj == 0) return {1, 4, 0.5};
    if (jj == NY + 1) return {1, 2, 0.5};
    if (ii == 0) return {2, 1, 0.5};
    return {3, 3, 0.5};
  };
  gen_mesh(c, n, m, 1.0 / NY, fluid, classify);
  int rc = finish_case(c);
  if (rc != LBM_OK) {
    delete c;
    return rc;
  }
  *out = c;
  return LBM_OK;
}

int lbm_case_write(const lbm_case* c, const char* node_file, const char* bc_file) {
  if (!c || !node_file || !bc_file) return LBM_EINVAL;
  if (c->ninst) {
    set_error("a packed case has no single reference-format mesh");
    return LBM_EINVAL;
  }
  FILE* f = std::fopen(node_file, "w");
  if (!f) {
    set_error(std::string("cannot write ") + node_file);
    return LBM_EIO;
  }
  for (int L = 0; L < c->num_nodes; L++) {
    const float* r = &c->nodes[(size_t)L * 5];
    std::fprintf(f, "%d    %d    %g    %g    %d\n", (int)r[0], (int)r[1], (double)r[2],
                 (double)r[3], (int)r[4]);
  }
  std::fclose(f);
  f = std::fopen(bc_file, "w");
  if (!f) {
    set_error(std::string("cannot write ") + bc_file);
    return LBM_EIO;
  }
  for (int L = 0; L < c->num_conn; L++) {
    const float* r = &c->conn[(size_t)L * 7];
    std::fprintf(f, "%d    %d    %d  %d   %g   %g   %d\n", (int)r[0], (int)r[1], (int)r[2],
                 (int)r[3], (double)r[4], (double)r[5], (int)r[6]);
  }
  std::fclose(f);
  return LBM_OK;
}

int lbm_case_instances(const lbm_case* c) { return c ? c->ninst : 0; }

int lbm_case_info_get(const lbm_case* c, lbm_case_info* o) {
  if (!c || !o) return LBM_EINVAL;
  o->n = c->n;
  o->m = c->m;
  o->num_nodes = c->num_nodes;
  o->num_conn = c->num_conn;
  o->num_fluid = c->num_fluid;
  o->num_boundary = c->num_boundary;
  o->num_inlet = c->num_inlet;
  o->delta = c->delta;
  o->min_inlet_y = c->min_inlet_y;
  o->max_inlet_y = c->max_inlet_y;
  return LBM_OK;
}

int lbm_case_cells(const lbm_case* c, int* fluid, int* corner, int* stream, int* id,
                   int* boundary, int* boundary_id, int* bc, double* q, double* x, double* y) {
  if (!c) return LBM_EINVAL;
  const long nc = (long)c->n * c->m;
  for (long i = 0; i < nc; i++) {
    const int b = c->bcell[i];
    const BCell* B = b >= 0 ? &c->bcells[b] : nullptr;
    if (fluid) fluid[i] = c->fluid[i];
    if (id) id[i] = c->id[i];
    if (corner) corner[i] = B ? B->corner : 0;
    if (boundary) boundary[i] = B ? B->boundary : 0;
    if (boundary_id) boundary_id[i] = B ? B->boundary_id : 0;
    for (int k = 0; k < 9; k++) {
      if (stream) stream[i * 9 + k] = B ? ((B->stream >> k) & 1) : 1;
      if (bc) bc[i * 9 + k] = B ? B->bc[k] : 0;
      if (q) q[i * 9 + k] = B ? B->q[k] : 0.5;
    }
    if (x) x[i] = c->x(i);
    if (y) y[i] = c->y(i);
  }
  return LBM_OK;
}

void lbm_case_free(lbm_case* c) { delete c; }

}  // extern "C"
