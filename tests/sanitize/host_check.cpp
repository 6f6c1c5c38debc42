// host_check.cpp — drives the host-only code of the library under
// AddressSanitizer + UndefinedBehaviorSanitizer or ThreadSanitizer
// (tests/test_host_sanitizers.py builds it with g++): the mesh readers
// (sequential and line-parallel, clean, joined-line and truncated files),
// CellIni, the mesh generators and packing, the mesh writer, the ini reader,
// the exact "%f" formatter and the threaded ordered result writer.
//
//   host_check SMALL_MESH_DIR LARGE_MESH_DIR OUT_DIR
//
// Prints OK and exits 0 when every check passes.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "fmt_fixed.hpp"
#include "lbm_host.hpp"
#include "result_writer.hpp"

static int failures = 0;
#define CHECK(c)                                                         \
  do {                                                                   \
    if (!(c)) {                                                          \
      std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #c); \
      failures++;                                                        \
    }                                                                    \
  } while (0)

static std::string slurp(const std::string& p) {
  std::ifstream f(p, std::ios::binary);
  std::stringstream ss;
  ss << f.rdbuf();
  return ss.str();
}

struct Cells {
  std::vector<int> fluid, corner, stream, id, boundary, bid, bc;
  std::vector<double> q, x, y;
};

static Cells cells_of(const lbm_case* c) {
  lbm_case_info info;
  CHECK(lbm_case_info_get(c, &info) == LBM_OK);
  const size_t nc = (size_t)info.n * info.m;
  Cells r;
  r.fluid.resize(nc);
  r.corner.resize(nc);
  r.stream.resize(nc * 9);
  r.id.resize(nc);
  r.boundary.resize(nc);
  r.bid.resize(nc);
  r.bc.resize(nc * 9);
  r.q.resize(nc * 9);
  r.x.resize(nc);
  r.y.resize(nc);
  CHECK(lbm_case_cells(c, r.fluid.data(), r.corner.data(), r.stream.data(), r.id.data(),
                       r.boundary.data(), r.bid.data(), r.bc.data(), r.q.data(), r.x.data(),
                       r.y.data()) == LBM_OK);
  return r;
}

static bool same(const Cells& a, const Cells& b) {
  return a.fluid == b.fluid && a.corner == b.corner && a.stream == b.stream && a.id == b.id &&
         a.boundary == b.boundary && a.bid == b.bid && a.bc == b.bc && a.q == b.q;
}

static void mesh_round_trip(const std::string& dir, const std::string& out, const char* tag) {
  lbm_case* c = nullptr;
  CHECK(lbm_case_load((dir + "/D2node.dat").c_str(), (dir + "/BCconnectors.dat").c_str(), &c) ==
        LBM_OK);
  if (!c) return;
  const Cells a = cells_of(c);
  const std::string n2 = out + "/" + tag + "_nodes.dat", b2 = out + "/" + tag + "_bc.dat";
  CHECK(lbm_case_write(c, n2.c_str(), b2.c_str()) == LBM_OK);
  lbm_case* c2 = nullptr;
  CHECK(lbm_case_load(n2.c_str(), b2.c_str(), &c2) == LBM_OK);
  if (c2) {
    CHECK(same(a, cells_of(c2)));
    lbm_case_free(c2);
  }
  // the same file with every other line joined to the next: the sequential
  // fscanf-semantics reader takes over and gives the same case
  const std::string txt = slurp(dir + "/D2node.dat");
  std::string joined;
  int line = 0;
  for (char ch : txt) {
    if (ch == '\n' && (line++ % 2 == 0)) {
      joined += ' ';
      continue;
    }
    joined += ch;
  }
  const std::string nj = out + "/" + tag + "_joined.dat";
  std::ofstream(nj, std::ios::binary) << joined;
  lbm_case* c3 = nullptr;
  CHECK(lbm_case_load(nj.c_str(), (dir + "/BCconnectors.dat").c_str(), &c3) == LBM_OK);
  if (c3) {
    CHECK(same(a, cells_of(c3)));
    lbm_case_free(c3);
  }
  // truncated mid-line: an error or a shorter case, never a crash
  const std::string nt = out + "/" + tag + "_trunc.dat";
  std::ofstream(nt, std::ios::binary) << txt.substr(0, txt.size() / 2 + 3);
  lbm_case* c4 = nullptr;
  const int rc = lbm_case_load(nt.c_str(), (dir + "/BCconnectors.dat").c_str(), &c4);
  CHECK(rc == LBM_OK || rc == LBM_EPARSE || rc == LBM_EIO);
  if (c4) lbm_case_free(c4);
  lbm_case_free(c);
}

static void generators() {
  lbm_case* a = nullptr;
  lbm_case* b = nullptr;
  lbm_case* d = nullptr;
  lbm_case* ch = nullptr;
  CHECK(lbm_case_cavity(60, 2, &a) == LBM_OK);
  const int rects[8] = {5, 5, 12, 9, 20, 14, 26, 20};
  CHECK(lbm_case_cavity_ex(40, 30, 2, rects, 2, &b) == LBM_OK);
  CHECK(lbm_case_channel(120, 40, 30.5, 20.2, 8.3, &ch) == LBM_OK);
  if (a && b) {
    const lbm_case* parts[3] = {a, b, a};
    int r[12];
    CHECK(lbm_case_pack(parts, 3, &d, r) == LBM_OK);
    if (d) {
      CHECK(lbm_case_instances(d) == 3);
      (void)cells_of(d);
    }
  }
  if (ch) (void)cells_of(ch);
  for (lbm_case* c : {a, b, d, ch})
    if (c) lbm_case_free(c);
  lbm_case* bad = nullptr;
  CHECK(lbm_case_load("/nonexistent/D2node.dat", "/nonexistent/BC.dat", &bad) != LBM_OK);
  CHECK(bad == nullptr);
}

static void ini(const std::string& out) {
  const std::string p = out + "/SetUpData.ini";
  std::ofstream(p) << "0.08\n0.0\n5\n0.01\n2\n3\n2\n2\n5000\n100\n10000\n1\n0\n-1.0\n-1.0\n0\n\n:01: note\n";
  lbm_ini v;
  CHECK(lbm_read_ini(p.c_str(), &v) == LBM_OK);
  CHECK(v.collision_model == 3 && v.iterations == 5000);
  lbm_options o;
  lbm_options_from_ini(&v, &o);
  CHECK(o.collision == 3);
  std::ofstream(out + "/short.ini") << "0.08\n0.0\n";
  CHECK(lbm_read_ini((out + "/short.ini").c_str(), &v) != LBM_OK);
}

static void formatter() {
  const double vals[] = {0.0, -0.0, 1.0, -1.0, 0.0000005, 0.0000015, -0.0000025, 1e-300,
                         4.9e-324, 123456.7890125, 1e300, -3.25e-7, 0.1, 2.0 / 3.0};
  for (double v : vals) {
    char a[512], b[512];
    const int la = lbmfmt::fmt_f6(v, a, sizeof a);
    a[la] = 0;
    std::snprintf(b, sizeof b, "%f", v);
    CHECK(std::strcmp(a, b) == 0);
  }
}

static void writer(const std::string& out) {
  // 150 000 rows of eight values and a flag: several threads format, one
  // writes, in order; bytes equal a sequential snprintf loop
  const long rows = 150000;
  std::vector<double> vals((size_t)rows * 8);
  unsigned long long st = 88172645463325252ull;
  for (double& v : vals) {
    st ^= st << 13;
    st ^= st >> 7;
    st ^= st << 17;
    v = ((double)(st >> 11) / 9007199254740992.0 - 0.5) * std::pow(10.0, (double)(st % 9) - 4);
  }
  const std::string p = out + "/rows.csv";
  auto row = [&](long r, char* buf, int) {
    return lbmhost::fmt_row(buf, &vals[(size_t)r * 8], 7, (int)(r % 2), &vals[(size_t)r * 8 + 7],
                            1, ", ");
  };
  CHECK(lbmhost::write_lines(p.c_str(), std::string("head\n"), rows, row) == LBM_OK);
  std::string want = "head\n";
  char b[1024];
  for (long r = 0; r < rows; r++) {
    const double* a = &vals[(size_t)r * 8];
    std::snprintf(b, sizeof b, "%f, %f, %f, %f, %f, %f, %f, %d, %f\n", a[0], a[1], a[2], a[3],
                  a[4], a[5], a[6], (int)(r % 2), a[7]);
    want += b;
  }
  CHECK(slurp(p) == want);
  CHECK(lbmhost::write_lines("/nonexistent/dir/x.csv", std::string(), 10, row) == LBM_EIO);
}

int main(int argc, char** argv) {
  if (argc != 4) {
    std::fprintf(stderr, "usage: host_check SMALL_MESH_DIR LARGE_MESH_DIR OUT_DIR\n");
    return 2;
  }
  mesh_round_trip(argv[1], argv[3], "small");
  mesh_round_trip(argv[2], argv[3], "large");
  generators();
  ini(argv[3]);
  formatter();
  writer(argv[3]);
  if (failures) {
    std::fprintf(stderr, "%d check(s) failed\n", failures);
    return 1;
  }
  std::printf("OK\n");
  return 0;
}
