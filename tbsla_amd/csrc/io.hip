// io.hip -- the binary image of a matrix block as tbsla::cpp::Matrix::write
// writes it and Matrix::read reads it (MatrixCSR.cpp:236-295,
// MatrixELL.cpp:208-259, MatrixCOO.cpp:144-204, MatrixSCOO.cpp:166-226,
// MatrixDENSE.cpp:88-127), so files pass between the reference and this
// library in both directions.
//
// Image: n_row n_col ln_row ln_col f_row f_col (int32), nnz (int64), pr pc
// NR NC (int32), ELL only: max_col (int32); then each array as a size_t
// element count followed by its elements, in the reference's array layouts:
//   CSR    values[nnz] rowptr[ln_row+1] colidx[nnz]
//   COO    values[nnz] row[nnz] col[nnz]          (SCOO the same)
//   ELL    values[ln_row*max_col] columns[ln_row*max_col]   (row-major)
//   DENSE  values[ln_row*ln_col]                   (row-major)
// CSR writes nnz entries of values and colidx as the reference does; where
// the nnz field counts more entries than the block holds (fill_cqmat's count
// pass, MatrixCSR.cpp:423-434) the reference writes whatever its arrays hold
// past rowptr[ln_row] (never initialised); this library writes zeros there.
// Reading keeps the CSR arrays as long as the image says and holds
// rowptr[ln_row] entries.
#include <cstdio>
#include <cstring>
#include <vector>

#include "common.hpp"

namespace tbsla_hip {
namespace {

struct Writer {
  std::vector<char> &out;
  template <typename T>
  void put(const T &v) {
    const char *p = reinterpret_cast<const char *>(&v);
    out.insert(out.end(), p, p + sizeof(T));
  }
  // an array of `count` elements, of which the first `have` come from p and
  // the rest are zero
  template <typename T>
  void array(const T *p, size_t have, size_t count) {
    put<size_t>(count);
    const size_t n = std::min(have, count);
    const char *b = reinterpret_cast<const char *>(p);
    if (n) out.insert(out.end(), b, b + n * sizeof(T));
    out.insert(out.end(), (count - n) * sizeof(T), 0);
  }
};

struct Reader {
  const char *p;
  size_t left;
  bool ok = true;
  template <typename T>
  T get() {
    T v{};
    if (left < sizeof(T)) {
      ok = false;
      return v;
    }
    std::memcpy(&v, p, sizeof(T));
    p += sizeof(T);
    left -= sizeof(T);
    return v;
  }
  template <typename T>
  void array(std::vector<T> &v) {
    const size_t n = get<size_t>();
    if (!ok || n > left / sizeof(T)) {
      ok = false;
      return;
    }
    v.resize(n);
    if (n) std::memcpy(v.data(), p, n * sizeof(T));
    p += n * sizeof(T);
    left -= n * sizeof(T);
  }
};

int image_of(tbsla_hip_matrix_s *m, std::vector<char> &out, void *stream) {
  tbsla_hip_geometry g{};
  TBSLA_TRY(tbsla_hip_get_geometry(m, &g));
  std::vector<int> a, b;
  std::vector<double> v;
  switch (g.format) {
    case TBSLA_HIP_CSR:
      a.resize((size_t)g.ln_row + 1);
      b.resize((size_t)g.stored);
      v.resize((size_t)g.stored);
      break;
    case TBSLA_HIP_COO:
    case TBSLA_HIP_SCOO:
      a.resize((size_t)g.stored);
      b.resize((size_t)g.stored);
      v.resize((size_t)g.stored);
      break;
    case TBSLA_HIP_ELL:
      a.resize((size_t)g.stored);
      v.resize((size_t)g.stored);
      break;
    default:
      v.resize((size_t)g.stored);
      break;
  }
  if (m->has_arrays)
    TBSLA_TRY(tbsla_hip_export(m, a.empty() ? nullptr : a.data(), b.empty() ? nullptr : b.data(),
                               v.empty() ? nullptr : v.data(), stream));
  else
    std::fill(a.begin(), a.end(), 0);
  out.clear();
  Writer w{out};
  w.put<int>(g.n_row);
  w.put<int>(g.n_col);
  w.put<int>(g.ln_row);
  w.put<int>(g.ln_col);
  w.put<int>(g.f_row);
  w.put<int>(g.f_col);
  w.put<long>((long)g.nnz);
  w.put<int>(g.pr);
  w.put<int>(g.pc);
  w.put<int>(g.NR);
  w.put<int>(g.NC);
  const size_t nnz = (size_t)std::max(g.nnz, 0LL);
  switch (g.format) {
    case TBSLA_HIP_CSR:
      w.array(v.data(), v.size(), nnz);
      w.array(a.data(), a.size(), (size_t)g.ln_row + 1);
      w.array(b.data(), b.size(), nnz);
      break;
    case TBSLA_HIP_COO:
    case TBSLA_HIP_SCOO:
      w.array(v.data(), v.size(), nnz);
      w.array(a.data(), a.size(), nnz);
      w.array(b.data(), b.size(), nnz);
      break;
    case TBSLA_HIP_ELL: {
      w.put<int>(g.max_col);
      const size_t slots = (size_t)g.ln_row * (size_t)std::max(g.max_col, 0);
      w.array(v.data(), v.size(), slots);
      w.array(a.data(), a.size(), slots);
      break;
    }
    default: {
      const size_t cells = (size_t)g.ln_row * (size_t)g.ln_col;
      w.array(v.data(), v.size(), cells);
      break;
    }
  }
  return TBSLA_HIP_SUCCESS;
}

int from_image(tbsla_hip_matrix_s *m, const char *p, size_t len, void *stream) {
  Reader r{p, len};
  tbsla_hip_geometry g{};
  g.format = m->format;
  g.n_row = r.get<int>();
  g.n_col = r.get<int>();
  g.ln_row = r.get<int>();
  g.ln_col = r.get<int>();
  g.f_row = r.get<int>();
  g.f_col = r.get<int>();
  g.nnz = r.get<long>();
  g.pr = r.get<int>();
  g.pc = r.get<int>();
  g.NR = r.get<int>();
  g.NC = r.get<int>();
  if (g.format == TBSLA_HIP_ELL) g.max_col = r.get<int>();
  std::vector<int> a, b;
  std::vector<double> v;
  switch (g.format) {
    case TBSLA_HIP_CSR:
      r.array(v);
      r.array(a);
      r.array(b);
      break;
    case TBSLA_HIP_COO:
    case TBSLA_HIP_SCOO:
      r.array(v);
      r.array(a);
      r.array(b);
      g.stored = (long long)v.size();
      if (a.size() != v.size() || b.size() != v.size())
        return fail(TBSLA_HIP_ERR_INVALID, "image: triplet arrays of different lengths");
      break;
    case TBSLA_HIP_ELL:
      r.array(v);
      r.array(a);
      if (v.size() != (size_t)g.ln_row * (size_t)std::max(g.max_col, 0) || a.size() != v.size())
        return fail(TBSLA_HIP_ERR_INVALID, "image: ELL arrays are not ln_row x max_col");
      break;
    default:
      r.array(v);
      if (v.size() != (size_t)g.ln_row * (size_t)g.ln_col)
        return fail(TBSLA_HIP_ERR_INVALID, "image: DENSE array is not ln_row x ln_col");
      break;
  }
  if (!r.ok) return fail(TBSLA_HIP_ERR_INVALID, "image truncated");
  if (g.format == TBSLA_HIP_CSR) {
    if (a.size() != (size_t)g.ln_row + 1)
      return fail(TBSLA_HIP_ERR_INVALID, "image: CSR rowptr is not ln_row + 1 long");
    if (a.back() < 0 || (size_t)a.back() > v.size() || v.size() != b.size())
      return fail(TBSLA_HIP_ERR_INVALID, "image: CSR arrays shorter than rowptr[ln_row]");
  }
  return tbsla_hip_import(m, g.format, &g, a.empty() ? nullptr : a.data(),
                          b.empty() ? nullptr : b.data(), v.empty() ? nullptr : v.data(), stream);
}

}  // namespace
}  // namespace tbsla_hip

using namespace tbsla_hip;

extern "C" int tbsla_hip_write(tbsla_hip_matrix_t m, void *h_buf, size_t capacity,
                               size_t *written, void *stream) {
  if (!m || !written) return fail(TBSLA_HIP_ERR_INVALID, "null argument");
  std::vector<char> img;
  TBSLA_TRY(image_of(m, img, stream));
  *written = img.size();
  if (!h_buf) return TBSLA_HIP_SUCCESS;
  if (capacity < img.size()) return fail(TBSLA_HIP_ERR_INVALID, "buffer smaller than the image");
  std::memcpy(h_buf, img.data(), img.size());
  return TBSLA_HIP_SUCCESS;
}

extern "C" int tbsla_hip_read(tbsla_hip_matrix_t m, const void *h_buf, size_t bytes,
                              void *stream) {
  if (!m || (!h_buf && bytes)) return fail(TBSLA_HIP_ERR_INVALID, "null argument");
  return from_image(m, static_cast<const char *>(h_buf), bytes, stream);
}

extern "C" int tbsla_hip_write_file(tbsla_hip_matrix_t m, const char *path, void *stream) {
  if (!m || !path) return fail(TBSLA_HIP_ERR_INVALID, "null argument");
  std::vector<char> img;
  TBSLA_TRY(image_of(m, img, stream));
  FILE *f = std::fopen(path, "wb");
  if (!f) return fail(TBSLA_HIP_ERR_INVALID, std::string("cannot open ") + path);
  const size_t n = std::fwrite(img.data(), 1, img.size(), f);
  const int closed = std::fclose(f);
  if (n != img.size() || closed != 0) return fail(TBSLA_HIP_ERR_INVALID, std::string("cannot write ") + path);
  return TBSLA_HIP_SUCCESS;
}

extern "C" int tbsla_hip_read_file(tbsla_hip_matrix_t m, const char *path, long long seek,
                                   void *stream) {
  if (!m || !path || seek < 0) return fail(TBSLA_HIP_ERR_INVALID, "null argument or negative seek");
  FILE *f = std::fopen(path, "rb");
  if (!f) return fail(TBSLA_HIP_ERR_INVALID, std::string("cannot open ") + path);
  std::vector<char> img;
  int st = TBSLA_HIP_SUCCESS;
  if (std::fseek(f, 0, SEEK_END) != 0) st = fail(TBSLA_HIP_ERR_INVALID, "cannot seek");
  long long size = st ? 0 : (long long)std::ftell(f);
  if (!st && (size < seek || std::fseek(f, (long)seek, SEEK_SET) != 0))
    st = fail(TBSLA_HIP_ERR_INVALID, "seek beyond the end of the file");
  if (!st) {
    img.resize((size_t)(size - seek));
    if (!img.empty() && std::fread(img.data(), 1, img.size(), f) != img.size())
      st = fail(TBSLA_HIP_ERR_INVALID, std::string("cannot read ") + path);
  }
  std::fclose(f);
  if (st) return st;
  return from_image(m, img.data(), img.size(), stream);
}
