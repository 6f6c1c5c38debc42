// result_writer.hpp — the ordered, threaded text writer behind the result
// files (WriteResults / WriteFinalResults, serial/FileIO.c:298-427): rows are
// formatted on several threads in 65 536-row chunks and written in order, so
// a file holds the same bytes as the reference's sequential fprintf loop.
// Host code only (no HIP calls); the host sanitizer build compiles it alone.
#pragma once
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/lbm.h"
#include "fmt_fixed.hpp"

namespace lbmhost {
void set_error(const std::string& msg);
}

namespace lbmhost {

// One result-file line: the values a[0..na) as printf "%f", the int flag as
// "%d", then b[0..nb) as "%f", separated by sep, ended by a newline: the
// bytes fprintf gives for the reference's format strings, without the
// printf machinery.  buf holds 8192 characters (a "%f" of a double is at
// most 317).
inline int fmt_row(char* buf, const double* a, int na, int flag, const double* b, int nb,
            const char* sep) {
  const int ls = (int)std::strlen(sep);
  int len = 0;
  for (int k = 0; k < na; k++) {
    len += lbmfmt::fmt_f6(a[k], buf + len, 8192 - len);
    std::memcpy(buf + len, sep, (size_t)ls);
    len += ls;
  }
  len += lbmfmt::fmt_int(flag, buf + len);
  for (int k = 0; k < nb; k++) {
    std::memcpy(buf + len, sep, (size_t)ls);
    len += ls;
    len += lbmfmt::fmt_f6(b[k], buf + len, 8192 - len);
  }
  buf[len++] = '\n';
  return len;
}

// Formats rows [r0, r1) of a result file into `out` (row -> one text line).
template <class RowFn>
void format_rows(long r0, long r1, RowFn row, std::string& out) {
  char line[8192];
  for (long r = r0; r < r1; r++) {
    const int len = row(r, line, sizeof line);
    out.append(line, (size_t)(len < (int)sizeof line ? len : (int)sizeof line - 1));
  }
}

// Writes `nrows` lines produced by `row` after `header`, formatting chunks of
// rows on several threads and writing them in order (same bytes as a
// sequential fprintf loop).  Two sets of chunk buffers alternate: one batch
// is written by a writer thread while the next is formatted.
template <class RowFn>
int write_lines(const char* path, const std::string& header, long nrows, RowFn row) {
  FILE* fp = std::fopen(path, "w");
  if (!fp) {
    set_error(std::string("cannot write ") + path);
    return LBM_EIO;
  }
  std::fwrite(header.data(), 1, header.size(), fp);
  unsigned nt = std::thread::hardware_concurrency();
  if (nt == 0) nt = 1;
  if (nt > 16) nt = 16;
  if (nrows < 20000) nt = 1;
  const long chunk = 1 << 16;  // rows per block
  std::vector<std::string> bufs[2] = {std::vector<std::string>(nt), std::vector<std::string>(nt)};
  std::thread writer;
  int cur = 0;
  for (long base = 0; base < nrows; base += chunk * (long)nt) {
    std::vector<std::string>& bs = bufs[cur];
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++) {
      const long r0 = base + (long)t * chunk;
      const long r1 = std::min(nrows, r0 + chunk);
      bs[t].clear();
      if (r0 >= r1) continue;
      if (nt == 1)
        format_rows(r0, r1, row, bs[t]);
      else
        th.emplace_back([&, t, r0, r1] { format_rows(r0, r1, row, bs[t]); });
    }
    for (auto& x : th) x.join();
    if (writer.joinable()) writer.join();  // the previous batch is on its way out
    writer = std::thread([&bs, fp] {
      for (const std::string& b : bs)
        if (!b.empty()) std::fwrite(b.data(), 1, b.size(), fp);
    });
    cur ^= 1;
  }
  if (writer.joinable()) writer.join();
  const bool bad = std::ferror(fp) != 0;
  std::fclose(fp);
  if (bad) {
    set_error(std::string("write failed: ") + path);
    return LBM_EIO;
  }
  return LBM_OK;
}

}  // namespace lbmhost

