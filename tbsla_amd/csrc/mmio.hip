// mmio.hip -- MatrixCOO::readMM (MatrixCOO.cpp:206-278): Matrix Market
// coordinate files into a COO block.
//
// Accepted, as the reference: a first line whose space-separated words are
// "%%MatrixMarket matrix coordinate real general|symmetric" (a trailing
// carriage return is ignored); then lines starting with '%' (and, here,
// blank lines) are skipped; the next line holds "n_row n_col n_values"; then
// n_values entries "row col value", 1-based, separated by any whitespace.
// A symmetric file stores one triangle: every off-diagonal entry (r, c)
// also yields (c, r), right after it.  (The reference writes the mirrored
// entry over the original one, at the same index, and leaves zero entries
// at the end of its arrays; here both are kept.)  Entries keep file order.
//
// Host code only: the file is mapped, the entry region is cut at line ends
// into chunks parsed by threads in parallel (numbers by strtol/strtod, so
// values are the correctly rounded doubles the reference's stream
// extraction gives), and the chunks are joined in order.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"

namespace tbsla_hip {

namespace {

struct Mapped {
  const char *p = nullptr;
  size_t n = 0;
  int fd = -1;
  ~Mapped() {
    if (p && n) munmap(const_cast<char *>(p), n);
    if (fd >= 0) close(fd);
  }
  int open_file(const char *path) {
    fd = ::open(path, O_RDONLY);
    if (fd < 0) return fail(TBSLA_HIP_ERR_INVALID, std::string("cannot open ") + path);
    struct stat st;
    if (fstat(fd, &st) != 0) return fail(TBSLA_HIP_ERR_INVALID, std::string("cannot stat ") + path);
    n = (size_t)st.st_size;
    if (n == 0) return TBSLA_HIP_SUCCESS;
    void *q = mmap(nullptr, n, PROT_READ, MAP_PRIVATE, fd, 0);
    if (q == MAP_FAILED) {
      n = 0;
      return fail(TBSLA_HIP_ERR_INVALID, std::string("cannot map ") + path);
    }
    p = static_cast<const char *>(q);
    return TBSLA_HIP_SUCCESS;
  }
};

struct Header {
  bool symmetric = false;
  long long n_row = 0, n_col = 0, n_values = 0;
  size_t data = 0;  // offset of the first entry
};

// one line [b, e) without its line end
size_t line_end(const char *p, size_t n, size_t b) {
  const void *q = memchr(p + b, '\n', n - b);
  return q ? (size_t)(static_cast<const char *>(q) - p) : n;
}

int parse_header(const char *p, size_t n, Header *h) {
  const char *bad = "This Matrix format import is not implemented!";
  size_t e = line_end(p, n, 0);
  std::string first(p, e);
  if (!first.empty() && first.back() == '\r') first.pop_back();
  std::vector<std::string> w;  // split on ' ', empty words dropped (utils/split.cpp:6-19)
  size_t a = 0;
  while (a <= first.size()) {
    size_t b = first.find(' ', a);
    if (b == std::string::npos) b = first.size();
    if (b > a) w.push_back(first.substr(a, b - a));
    a = b + 1;
  }
  if (w.size() < 5 || w[0] != "%%MatrixMarket" || w[1] != "matrix" || w[2] != "coordinate" ||
      w[3] != "real" || (w[4] != "general" && w[4] != "symmetric"))
    return fail(TBSLA_HIP_ERR_UNSUPPORTED, bad);
  h->symmetric = w[4] == "symmetric";
  // comment lines (and blank lines), then the size line
  size_t b = e < n ? e + 1 : n;
  for (;;) {
    if (b >= n) return fail(TBSLA_HIP_ERR_INVALID, "Matrix Market file without a size line");
    e = line_end(p, n, b);
    size_t k = b;
    while (k < e && (p[k] == ' ' || p[k] == '\t' || p[k] == '\r')) k++;
    if (p[b] != '%' && k < e) break;
    b = e + 1;
  }
  std::string size_line(p + b, e - b);
  char *end = nullptr;
  const char *s = size_line.c_str();
  errno = 0;
  h->n_row = strtoll(s, &end, 10);
  const char *s2 = end;
  h->n_col = strtoll(s2, &end, 10);
  const char *s3 = end;
  h->n_values = strtoll(s3, &end, 10);
  if (errno || end == s3 || s2 == s || s3 == s2 || h->n_row < 0 || h->n_col < 0 ||
      h->n_values < 0 || h->n_row > INT32_MAX || h->n_col > INT32_MAX)
    return fail(TBSLA_HIP_ERR_INVALID, "bad Matrix Market size line: " + size_line);
  h->data = e < n ? e + 1 : n;
  return TBSLA_HIP_SUCCESS;
}

struct Chunk {
  std::vector<long long> r, c;
  std::vector<double> v;
  bool ok = true;
};

inline bool space(char ch) { return ch == ' ' || ch == '\t' || ch == '\n' || ch == '\r' || ch == '\f' || ch == '\v'; }

// whitespace-separated (row, col, value) triples of [b, e); false on a token
// that is not a number or on a trailing incomplete triple
void parse_chunk(const char *p, size_t b, size_t e, Chunk *out) {
  std::string buf;  // strtod needs a terminated token
  size_t i = b;
  int field = 0;
  long long r = 0, c = 0;
  while (i < e) {
    while (i < e && space(p[i])) i++;
    if (i >= e) break;
    size_t j = i;
    while (j < e && !space(p[j])) j++;
    buf.assign(p + i, j - i);
    char *end = nullptr;
    if (field < 2) {
      errno = 0;
      const long long x = strtoll(buf.c_str(), &end, 10);
      if (errno || *end) {
        out->ok = false;
        return;
      }
      (field == 0 ? r : c) = x;
    } else {
      const double x = strtod(buf.c_str(), &end);
      if (*end) {
        out->ok = false;
        return;
      }
      out->r.push_back(r);
      out->c.push_back(c);
      out->v.push_back(x);
    }
    field = (field + 1) % 3;
    i = j;
  }
  if (field != 0) out->ok = false;
}

}  // namespace

// Parses a file into host triplets (0-based, reference order).
int read_mm_host(const char *path, long long *n_row, long long *n_col, std::vector<int> *rows,
                 std::vector<int> *cols, std::vector<double> *vals) {
  Mapped f;
  TBSLA_TRY(f.open_file(path));
  if (f.n == 0) return fail(TBSLA_HIP_ERR_UNSUPPORTED, "This Matrix format import is not implemented!");
  Header h;
  TBSLA_TRY(parse_header(f.p, f.n, &h));
  // chunks cut at line ends
  const size_t bytes = f.n - h.data;
  unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  int T = (int)std::min<size_t>(std::min(16u, hw), bytes / (4u << 20) + 1);
  std::vector<size_t> cut(T + 1);
  cut[0] = h.data;
  cut[T] = f.n;
  for (int t = 1; t < T; t++) {
    size_t x = std::max(cut[t - 1], h.data + bytes * t / T);
    x = line_end(f.p, f.n, x);
    cut[t] = x < f.n ? x + 1 : f.n;
  }
  std::vector<Chunk> ch(T);
  {
    std::vector<std::thread> th;
    for (int t = 1; t < T; t++) th.emplace_back(parse_chunk, f.p, cut[t], cut[t + 1], &ch[t]);
    parse_chunk(f.p, cut[0], cut[1], &ch[0]);
    for (auto &x : th) x.join();
  }
  // a chunk that ends inside a triple means entries span lines: parse the
  // whole region in one piece
  auto all_ok = [&ch]() {
    return std::all_of(ch.begin(), ch.end(), [](const Chunk &x) { return x.ok; });
  };
  if (!all_ok() && T > 1) {
    ch.assign(1, Chunk());
    parse_chunk(f.p, h.data, f.n, &ch[0]);
  }
  if (!all_ok()) return fail(TBSLA_HIP_ERR_INVALID, "malformed Matrix Market entries");
  long long total = 0;
  for (auto &x : ch) total += (long long)x.v.size();
  if (total < h.n_values)
    return fail(TBSLA_HIP_ERR_INVALID, "Matrix Market file holds " + std::to_string(total) +
                                           " entries, the size line declares " +
                                           std::to_string(h.n_values));
  // the first n_values entries (MatrixCOO.cpp:233,258), mirrored for symmetric
  rows->clear();
  cols->clear();
  vals->clear();
  const size_t cap = (size_t)h.n_values * (h.symmetric ? 2 : 1);
  if (cap > (size_t)INT32_MAX) return fail(TBSLA_HIP_ERR_RANGE, "too many entries for int32 indexing");
  rows->reserve(cap);
  cols->reserve(cap);
  vals->reserve(cap);
  long long seen = 0;
  for (auto &x : ch) {
    for (size_t k = 0; k < x.v.size() && seen < h.n_values; k++, seen++) {
      const long long r = x.r[k], c = x.c[k];
      if (r < 1 || r > h.n_row || c < 1 || c > h.n_col)
        return fail(TBSLA_HIP_ERR_INVALID, "Matrix Market entry " + std::to_string(seen + 1) +
                                               " outside the matrix");
      rows->push_back((int)(r - 1));
      cols->push_back((int)(c - 1));
      vals->push_back(x.v[k]);
      if (h.symmetric && r != c) {
        rows->push_back((int)(c - 1));
        cols->push_back((int)(r - 1));
        vals->push_back(x.v[k]);
      }
    }
  }
  *n_row = h.n_row;
  *n_col = h.n_col;
  return TBSLA_HIP_SUCCESS;
}

}  // namespace tbsla_hip

using namespace tbsla_hip;

extern "C" int tbsla_hip_mm_parse(const char *path, int *n_row, int *n_col, long long *nnz,
                                  int *h_rows, int *h_cols, double *h_values,
                                  long long capacity) {
  if (!path || !n_row || !n_col || !nnz) return fail(TBSLA_HIP_ERR_INVALID, "null argument");
  long long nr = 0, nc = 0;
  std::vector<int> r, c;
  std::vector<double> v;
  TBSLA_TRY(read_mm_host(path, &nr, &nc, &r, &c, &v));
  *n_row = (int)nr;
  *n_col = (int)nc;
  *nnz = (long long)v.size();
  if (h_rows || h_cols || h_values) {
    if (capacity < (long long)v.size())
      return fail(TBSLA_HIP_ERR_INVALID, "capacity below the entry count");
    if (h_rows) std::copy(r.begin(), r.end(), h_rows);
    if (h_cols) std::copy(c.begin(), c.end(), h_cols);
    if (h_values) std::copy(v.begin(), v.end(), h_values);
  }
  return TBSLA_HIP_SUCCESS;
}

extern "C" int tbsla_hip_read_mm(tbsla_hip_matrix_t m, const char *path, void *stream) {
  if (!m || !path) return fail(TBSLA_HIP_ERR_INVALID, "null argument");
  if (m->format != TBSLA_HIP_COO)
    return fail(TBSLA_HIP_ERR_UNSUPPORTED, "readMM fills a COO block (convert afterwards)");
  long long nr = 0, nc = 0;
  std::vector<int> r, c;
  std::vector<double> v;
  TBSLA_TRY(read_mm_host(path, &nr, &nc, &r, &c, &v));
  // MatrixCOO::init (MatrixCOO.cpp:14-34): one block covering the matrix
  tbsla_hip_geometry g{};
  g.n_row = g.ln_row = (int)nr;
  g.n_col = g.ln_col = (int)nc;
  g.f_row = g.f_col = g.pr = g.pc = 0;
  g.NR = g.NC = 1;
  g.nnz = g.stored = (long long)v.size();
  g.format = TBSLA_HIP_COO;
  return tbsla_hip_import(m, TBSLA_HIP_COO, &g, r.empty() ? nullptr : r.data(),
                          c.empty() ? nullptr : c.data(), v.empty() ? nullptr : v.data(), stream);
}
