// gary_convert — the reference's graph conversion tool (convert.cpp:13-112) over the C ABI's file
// functions. Same conversions, flags and output files:
//
//   gary_convert -input=<path> -output=<path> -convert=<snap2bin|coo2csr|csr2csc|orient>
//                [-symmetrize=1] [-keepLoops=1]
//
//   snap2bin  SNAP edge list -> binary GraphCOO; self-loops dropped unless -keepLoops=1
//             (convertSNAP2Bin, convert.cpp:13-17; loadFromSNAP, src/Graph.cpp:157-196)
//   coo2csr   binary GraphCOO -> binary GraphCSR, rows sorted, duplicates kept; with
//             -symmetrize=1 every (u, v) also gets (v, u) first (convert.cpp:19-40;
//             GraphCOO::symmetrize and toCSR_, src/Graph.cpp:19-32,63-91)
//   csr2csc   binary GraphCSR -> its transpose (convert.cpp:42-58; GraphCSR::reverse)
//   orient    binary GraphCSR -> the arcs u -> v with u > v (convert.cpp:60-78;
//             GraphCOO::orientation, src/Graph.cpp:34-61)
//
// A flag is `-name=value` or `--name=value`; an integer flag that is absent is 0. Unlike the
// reference, an unknown -convert value, a missing flag or a file of the wrong format ends with a
// message and exit status 1 instead of doing nothing or throwing. Host code only: no GPU needed.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "gary.h"

namespace {

struct Fail {};

void check(int rc) {
  if (rc != GARY_OK) {
    std::fprintf(stderr, "gary_convert: %s\n", gary_last_error());
    throw Fail{};
  }
}

[[noreturn]] void fail(const std::string &msg) {
  std::fprintf(stderr, "gary_convert: %s\n", msg.c_str());
  throw Fail{};
}

// The value of -name=value / --name=value, or nullptr.
const char *flag(int argc, char **argv, const char *name) {
  const size_t len = std::strlen(name);
  for (int i = 1; i < argc; ++i) {
    const char *a = argv[i];
    if (*a != '-') continue;
    a += (a[1] == '-') ? 2 : 1;
    if (std::strncmp(a, name, len) == 0 && a[len] == '=') return a + len + 1;
  }
  return nullptr;
}

int int_flag(int argc, char **argv, const char *name) {
  const char *v = flag(argc, argv, name);
  return v ? std::atoi(v) : 0;
}

struct Csr {
  int32_t m = 0;
  std::vector<int64_t> ptr;
  std::vector<int32_t> idx;
  int64_t nnz() const { return ptr.empty() ? 0 : ptr.back(); }
  int32_t dmax() const {
    int64_t d = 0;
    for (int32_t v = 0; v < m; ++v) d = std::max(d, ptr[v + 1] - ptr[v]);
    return (int32_t)d;
  }
};

struct Coo {
  int32_t m = 0;
  std::vector<int32_t> row, col;
};

void expect_format(const char *path, int want, int32_t *m, int64_t *nnz) {
  int format = -1;
  check(gary_graph_file_info(path, &format, m, nnz));
  if (format != want)
    fail(std::string(path) + ": not a binary " + (want == GARY_GRAPH_FILE_CSR ? "GraphCSR" : "GraphCOO") +
         " file");
}

Coo load_coo(const char *path) {
  Coo g;
  int64_t nnz = 0;
  expect_format(path, GARY_GRAPH_FILE_COO, &g.m, &nnz);
  g.row.resize((size_t)nnz);
  g.col.resize((size_t)nnz);
  check(gary_graph_file_read_coo(path, g.row.data(), g.col.data()));
  return g;
}

Csr load_csr(const char *path) {
  Csr g;
  int64_t nnz = 0;
  expect_format(path, GARY_GRAPH_FILE_CSR, &g.m, &nnz);
  g.ptr.resize((size_t)g.m + 1);
  g.idx.resize((size_t)nnz);
  check(gary_graph_file_read_csr(path, g.ptr.data(), g.idx.data(), nullptr));
  return g;
}

Csr to_csr(const Coo &c) {
  Csr g;
  g.m = c.m;
  g.ptr.resize((size_t)c.m + 1);
  g.idx.resize(c.row.size());
  check(gary_coo_to_csr(c.m, (int64_t)c.row.size(), c.row.data(), c.col.data(), g.ptr.data(),
                        g.idx.data()));
  return g;
}

// GraphCSR::toCOO (src/Graph.cpp:141-155): row ids expanded from ptr, col = idx.
Coo to_coo(const Csr &g) {
  Coo c;
  c.m = g.m;
  c.row.resize(g.idx.size());
  for (int32_t v = 0; v < g.m; ++v)
    std::fill(c.row.begin() + g.ptr[v], c.row.begin() + g.ptr[v + 1], v);
  c.col = g.idx;
  return c;
}

void save(const Csr &g, const char *path) {
  check(gary_graph_file_write_csr(path, g.m, g.nnz(), g.ptr.data(), g.idx.data(), nullptr));
}

void report(const char *path, const Csr &g) {
  std::printf("%s: %d nodes, %lld edges, d_max=%d\n", path, g.m, (long long)g.nnz(), g.dmax());
}

void snap2bin(const char *in, const char *out, bool keep_loops) {
  int64_t nt = 0;
  int32_t n = 0;
  check(gary_snap_read(in, keep_loops ? 0 : 1, nullptr, nullptr, 0, &nt, &n));
  std::vector<int32_t> src((size_t)nt), dst((size_t)nt);
  check(gary_snap_read(in, keep_loops ? 0 : 1, src.data(), dst.data(), nt, &nt, &n));
  std::printf("%s: %d nodes, %lld edges\n", in, n, (long long)nt);
  check(gary_graph_file_write_coo(out, n, nt, src.data(), dst.data()));
}

void coo2csr(const char *in, const char *out, bool symmetrize) {
  Coo c = load_coo(in);
  std::printf("%s: %d nodes, %lld edges\n", in, c.m, (long long)c.row.size());
  if (symmetrize) {  // the original tuples, then each reversed
    const size_t e = c.row.size();
    c.row.resize(2 * e);
    c.col.resize(2 * e);
    std::copy(c.col.begin(), c.col.begin() + e, c.row.begin() + e);
    std::copy(c.row.begin(), c.row.begin() + e, c.col.begin() + e);
  }
  const Csr g = to_csr(c);
  save(g, out);
  report(out, g);
}

void csr2csc(const char *in, const char *out) {
  const Csr g = load_csr(in);
  report(in, g);
  Coo c = to_coo(g);
  std::swap(c.row, c.col);
  const Csr t = to_csr(c);
  save(t, out);
  report(out, t);
}

void orient(const char *in, const char *out) {
  const Csr g = load_csr(in);
  report(in, g);
  const Coo c = to_coo(g);
  Coo o;
  o.m = c.m;
  for (size_t e = 0; e < c.row.size(); ++e)
    if (c.row[e] > c.col[e]) {
      o.row.push_back(c.row[e]);
      o.col.push_back(c.col[e]);
    }
  const Csr t = to_csr(o);
  save(t, out);
  report(out, t);
}

}  // namespace

int main(int argc, char **argv) {
  try {
    const char *in = flag(argc, argv, "input");
    const char *out = flag(argc, argv, "output");
    const char *conv = flag(argc, argv, "convert");
    if (!in) fail("missing -input=<path>");
    if (!out) fail("missing -output=<path>");
    const std::string c = conv ? conv : "";
    if (c == "snap2bin")
      snap2bin(in, out, int_flag(argc, argv, "keepLoops") != 0);
    else if (c == "coo2csr")
      coo2csr(in, out, int_flag(argc, argv, "symmetrize") != 0);
    else if (c == "csr2csc")
      csr2csc(in, out);
    else if (c == "orient")
      orient(in, out);
    else
      fail("-convert must be one of snap2bin, coo2csr, csr2csc, orient");
  } catch (const Fail &) {
    return 1;
  }
  return 0;
}
