// graph_io.cpp — the reference's graph files, read and written byte for byte:
//   * binary graphs: GraphCSR / GraphCOO saved by saveToBinary (include/Graph.h:301-316) through
//     a cereal BinaryOutputArchive, layout V1 (Graph.h:16-23, save/load at :133-148, :270-285);
//   * SNAP edge lists (loadFromSNAP, src/Graph.cpp:157-196).
//
// Layout of a V1 file (cereal's binary archive: native little-endian, no padding, no names):
//   uint32 version (= 1, SerDesVersion::V1)   uint32 format (0 = CSR, 1 = COO)
//   int32 m                                   int64 nnz
//   then each std::shared_ptr<std::vector<T>> in field order (CSR: ptr int64, idx int32,
//   deg int32; COO: row int32, col int32) as
//   uint32 id (0x80000000 | k for the k-th distinct pointer, k from 1)   uint64 size   T[size]
// An id without the top bit refers back to an earlier pointer of the same archive; the
// reference never writes one (its vectors are distinct), and this reader rejects it.
#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

#include "common.h"

namespace {

constexpr uint32_t kVersionV1 = 1;
constexpr uint32_t kFirstPtrId = 0x80000001u;
constexpr uint32_t kPtrNew = 0x80000000u;

struct File {
  FILE *f = nullptr;
  ~File() {
    if (f) std::fclose(f);
  }
};

int read_exact(FILE *f, void *dst, size_t bytes, const char *path, const char *what) {
  if (bytes && std::fread(dst, 1, bytes, f) != bytes)
    return gary::set_error(GARY_ERR_INVALID, "%s: truncated graph file (reading %s)", path, what);
  return GARY_OK;
}

int write_exact(FILE *f, const void *src, size_t bytes, const char *path) {
  if (bytes && std::fwrite(src, 1, bytes, f) != bytes)
    return gary::set_error(GARY_ERR_INVALID, "%s: write failed: %s", path, std::strerror(errno));
  return GARY_OK;
}

struct Header {
  uint32_t version, format;
  int32_t m;
  int64_t nnz;
};

int read_header(FILE *f, const char *path, Header *h) {
  GARY_TRY(read_exact(f, &h->version, 4, path, "version"));
  GARY_TRY(read_exact(f, &h->format, 4, path, "format"));
  GARY_TRY(read_exact(f, &h->m, 4, path, "m"));
  GARY_TRY(read_exact(f, &h->nnz, 8, path, "nnz"));
  if (h->version != kVersionV1)
    return gary::set_error(GARY_ERR_INVALID, "%s: unknown graph file version %u", path,
                           h->version);
  if (h->format > 1)
    return gary::set_error(GARY_ERR_INVALID, "%s: unknown graph format %u", path, h->format);
  if (h->m < 0 || h->nnz < 0)
    return gary::set_error(GARY_ERR_INVALID, "%s: m=%d nnz=%lld", path, h->m,
                           (long long)h->nnz);
  return GARY_OK;
}

// One shared_ptr<vector<T>> of `want` elements into dst (skipped when dst is null).
int read_array(FILE *f, const char *path, const char *what, size_t elem, uint64_t want,
               void *dst) {
  uint32_t id = 0;
  uint64_t size = 0;
  GARY_TRY(read_exact(f, &id, 4, path, what));
  if (id == 0) return gary::set_error(GARY_ERR_INVALID, "%s: %s is a null pointer", path, what);
  if (!(id & kPtrNew))
    return gary::set_error(GARY_ERR_UNSUPPORTED, "%s: %s refers back to pointer %u", path, what,
                           id);
  GARY_TRY(read_exact(f, &size, 8, path, what));
  if (size != want)
    return gary::set_error(GARY_ERR_INVALID, "%s: %s has %llu elements, expected %llu", path,
                           what, (unsigned long long)size, (unsigned long long)want);
  if (dst) return read_exact(f, dst, elem * size, path, what);
  if (std::fseek(f, (long)(elem * size), SEEK_CUR) != 0)
    return gary::set_error(GARY_ERR_INVALID, "%s: truncated graph file (skipping %s)", path, what);
  return GARY_OK;
}

int write_array(FILE *f, const char *path, uint32_t id, size_t elem, uint64_t size,
                const void *src) {
  GARY_TRY(write_exact(f, &id, 4, path));
  GARY_TRY(write_exact(f, &size, 8, path));
  return write_exact(f, src, elem * size, path);
}

int open_file(const char *path, const char *mode, File *file) {
  GARY_CHECK_ARG(path, "null path");
  file->f = std::fopen(path, mode);
  if (!file->f)
    return gary::set_error(GARY_ERR_INVALID, "%s: cannot open: %s", path, std::strerror(errno));
  return GARY_OK;
}

int write_header(FILE *f, const char *path, uint32_t format, int32_t m, int64_t nnz) {
  GARY_TRY(write_exact(f, &kVersionV1, 4, path));
  GARY_TRY(write_exact(f, &format, 4, path));
  GARY_TRY(write_exact(f, &m, 4, path));
  return write_exact(f, &nnz, 8, path);
}

}  // namespace

extern "C" {

int gary_graph_file_info(const char *path, int *format, int32_t *m, int64_t *nnz) {
  GARY_API_BEGIN
  File file;
  GARY_TRY(open_file(path, "rb", &file));
  Header h;
  GARY_TRY(read_header(file.f, path, &h));
  if (format) *format = (int)h.format;
  if (m) *m = h.m;
  if (nnz) *nnz = h.nnz;
  return GARY_OK;
  GARY_API_END
}

int gary_graph_file_read_csr(const char *path, int64_t *indPtr, int32_t *colIdx, int32_t *deg) {
  GARY_API_BEGIN
  File file;
  GARY_TRY(open_file(path, "rb", &file));
  Header h;
  GARY_TRY(read_header(file.f, path, &h));
  if (h.format != GARY_GRAPH_FILE_CSR)
    return gary::set_error(GARY_ERR_INVALID,
                           "%s: not a CSR graph file (format %u); read it as COO", path, h.format);
  GARY_TRY(read_array(file.f, path, "ptr", 8, (uint64_t)h.m + 1, indPtr));
  GARY_TRY(read_array(file.f, path, "idx", 4, (uint64_t)h.nnz, colIdx));
  GARY_TRY(read_array(file.f, path, "deg", 4, (uint64_t)h.m, deg));
  if (indPtr) {
    bool ok = indPtr[0] == 0 && indPtr[h.m] == h.nnz;
    for (int32_t v = 0; ok && v < h.m; ++v) ok = indPtr[v] <= indPtr[v + 1];
    if (!ok) return gary::set_error(GARY_ERR_INVALID, "%s: row pointers are not a CSR", path);
  }
  return GARY_OK;
  GARY_API_END
}

int gary_graph_file_read_coo(const char *path, int32_t *row, int32_t *col) {
  GARY_API_BEGIN
  File file;
  GARY_TRY(open_file(path, "rb", &file));
  Header h;
  GARY_TRY(read_header(file.f, path, &h));
  if (h.format != GARY_GRAPH_FILE_COO)
    return gary::set_error(GARY_ERR_INVALID,
                           "%s: not a COO graph file (format %u); read it as CSR", path, h.format);
  GARY_TRY(read_array(file.f, path, "row", 4, (uint64_t)h.nnz, row));
  GARY_TRY(read_array(file.f, path, "col", 4, (uint64_t)h.nnz, col));
  return GARY_OK;
  GARY_API_END
}

int gary_graph_file_write_csr(const char *path, int32_t m, int64_t nnz, const int64_t *indPtr,
                              const int32_t *colIdx, const int32_t *deg) {
  GARY_API_BEGIN
  GARY_CHECK_ARG(m >= 0 && nnz >= 0 && indPtr && (nnz == 0 || colIdx),
                 "gary_graph_file_write_csr: bad arguments");
  GARY_CHECK_ARG(indPtr[0] == 0 && indPtr[m] == nnz,
                 "gary_graph_file_write_csr: indPtr[0]=%lld indPtr[m]=%lld nnz=%lld",
                 (long long)indPtr[0], (long long)indPtr[m], (long long)nnz);
  std::vector<int32_t> derived;
  if (!deg) {  // GraphCSR's constructor derives deg from ptr (include/Graph.h:242-248)
    derived.resize((size_t)m);
    for (int32_t v = 0; v < m; ++v) derived[(size_t)v] = (int32_t)(indPtr[v + 1] - indPtr[v]);
    deg = derived.data();
  }
  File file;
  GARY_TRY(open_file(path, "wb", &file));
  GARY_TRY(write_header(file.f, path, GARY_GRAPH_FILE_CSR, m, nnz));
  GARY_TRY(write_array(file.f, path, kFirstPtrId, 8, (uint64_t)m + 1, indPtr));
  GARY_TRY(write_array(file.f, path, kFirstPtrId + 1, 4, (uint64_t)nnz, colIdx));
  GARY_TRY(write_array(file.f, path, kFirstPtrId + 2, 4, (uint64_t)m, deg));
  if (std::fflush(file.f) != 0)
    return gary::set_error(GARY_ERR_INVALID, "%s: write failed: %s", path, std::strerror(errno));
  return GARY_OK;
  GARY_API_END
}

int gary_graph_file_write_coo(const char *path, int32_t m, int64_t nnz, const int32_t *row,
                              const int32_t *col) {
  GARY_API_BEGIN
  GARY_CHECK_ARG(m >= 0 && nnz >= 0 && (nnz == 0 || (row && col)),
                 "gary_graph_file_write_coo: bad arguments");
  File file;
  GARY_TRY(open_file(path, "wb", &file));
  GARY_TRY(write_header(file.f, path, GARY_GRAPH_FILE_COO, m, nnz));
  GARY_TRY(write_array(file.f, path, kFirstPtrId, 4, (uint64_t)nnz, row));
  GARY_TRY(write_array(file.f, path, kFirstPtrId + 1, 4, (uint64_t)nnz, col));
  if (std::fflush(file.f) != 0)
    return gary::set_error(GARY_ERR_INVALID, "%s: write failed: %s", path, std::strerror(errno));
  return GARY_OK;
  GARY_API_END
}

int gary_snap_read(const char *path, int remove_loops, int32_t *src, int32_t *dst, int64_t cap,
                   int64_t *ntuples, int32_t *n) {
  GARY_API_BEGIN
  GARY_CHECK_ARG(ntuples && n, "gary_snap_read: null output");
  File file;
  GARY_TRY(open_file(path, "r", &file));
  int64_t count = 0, lineno = 0;
  int64_t max_id = 0;  // numVertex = maxId + 1 with maxId starting at 0 (src/Graph.cpp:162,190)
  std::vector<char> line(1 << 16);
  while (std::fgets(line.data(), (int)line.size(), file.f)) {
    ++lineno;
    const size_t len = std::strlen(line.data());
    if (len + 1 == line.size() && line[len - 1] != '\n')
      return gary::set_error(GARY_ERR_INVALID, "%s:%lld: line too long", path, (long long)lineno);
    const char *p = line.data();
    while (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n' || *p == '\v' || *p == '\f') ++p;
    if (*p == '\0' || *p == '#') continue;  // blank lines and comments
    char *end = nullptr;
    errno = 0;
    const long long u = std::strtoll(p, &end, 10);
    const bool got_u = end != p;
    const char *q = end;
    const long long v = std::strtoll(q, &end, 10);
    if (!got_u || end == q || errno == ERANGE)
      return gary::set_error(GARY_ERR_INVALID, "%s:%lld: expected two vertex ids", path,
                             (long long)lineno);
    if (u < 0 || v < 0 || u > INT32_MAX - 1 || v > INT32_MAX - 1)
      return gary::set_error(GARY_ERR_INVALID, "%s:%lld: vertex id out of the int32 range",
                             path, (long long)lineno);
    if (remove_loops && u == v) continue;
    if (src && dst) {
      if (count >= cap)
        return gary::set_error(GARY_ERR_INVALID, "gary_snap_read: capacity %lld exceeded",
                               (long long)cap);
      src[count] = (int32_t)u;
      dst[count] = (int32_t)v;
    }
    max_id = std::max<long long>(max_id, std::max(u, v));
    ++count;
  }
  if (std::ferror(file.f))
    return gary::set_error(GARY_ERR_INVALID, "%s: read error", path);
  *ntuples = count;
  *n = (int32_t)(max_id + 1);
  return GARY_OK;
  GARY_API_END
}

// GraphCOO::toCSR_ (src/Graph.cpp:63-91): rows by source id (counting sort), each row's columns
// sorted ascending; duplicates and loops stay. Equal (row, col) pairs are indistinguishable, so
// the result is the same whatever the order of the input.
int gary_coo_to_csr(int32_t m, int64_t nnz, const int32_t *row, const int32_t *col,
                    int64_t *indPtr, int32_t *colIdx) {
  GARY_API_BEGIN
  GARY_CHECK_ARG(m >= 0 && nnz >= 0 && indPtr && (nnz == 0 || (row && col && colIdx)),
                 "gary_coo_to_csr: bad arguments");
  std::vector<int64_t> next((size_t)m + 1, 0);
  for (int64_t e = 0; e < nnz; ++e) {
    if (row[e] < 0 || row[e] >= m || col[e] < 0 || col[e] >= m)
      return gary::set_error(GARY_ERR_INVALID, "gary_coo_to_csr: edge %lld (%d, %d) out of [0, %d)",
                             (long long)e, row[e], col[e], m);
    ++next[(size_t)row[e] + 1];
  }
  for (int32_t v = 0; v < m; ++v) next[(size_t)v + 1] += next[v];
  std::copy(next.begin(), next.end(), indPtr);
  for (int64_t e = 0; e < nnz; ++e) colIdx[next[row[e]]++] = col[e];
  for (int32_t v = 0; v < m; ++v) std::sort(colIdx + indPtr[v], colIdx + indPtr[v + 1]);
  return GARY_OK;
  GARY_API_END
}

}  // extern "C"
