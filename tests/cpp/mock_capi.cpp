// mock_capi.cpp -- the handful of C-ABI calls that the host-side code under
// test makes (handles, geometry, import/export, error strings), implemented
// in host memory so that the binary-image code (io.hip), the Matrix Market
// parser (mmio.hip) and the C++ class (include/tbsla/hip/Matrix.hpp) can be
// built with AddressSanitizer and UBSan and run without a GPU.  Arrays are
// kept in the reference layouts (ELL row-major), so export is a copy.
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"

namespace tbsla_hip {
thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }
int fail(int status, const std::string &msg) {
  g_err = msg;
  return status;
}
}  // namespace tbsla_hip

using namespace tbsla_hip;

namespace {
template <typename T>
T *host_copy(const T *src, size_t n) {
  T *p = static_cast<T *>(std::malloc(std::max<size_t>(n, 1) * sizeof(T)));
  if (src && n) std::memcpy(p, src, n * sizeof(T));
  else std::memset(p, 0, std::max<size_t>(n, 1) * sizeof(T));
  return p;
}
void drop(tbsla_hip_matrix_s *m) {
  std::free(m->rowptr);
  std::free(m->colidx);
  std::free(m->rowidx);
  std::free(m->values);
  m->rowptr = m->colidx = m->rowidx = nullptr;
  m->values = nullptr;
  m->has_arrays = 0;
}
size_t slots_of(const tbsla_hip_matrix_s *m) {
  switch (m->format) {
    case TBSLA_HIP_ELL: return (size_t)m->ln_row * (size_t)std::max(m->max_col, 0);
    case TBSLA_HIP_DENSE: return (size_t)m->ln_row * (size_t)m->ln_col;
    default: return (size_t)m->stored;
  }
}
}  // namespace

extern "C" {
const char *tbsla_hip_last_error(void) { return g_err.c_str(); }
const char *tbsla_hip_status_string(int status) { return status ? "error" : "success"; }

int tbsla_hip_create(tbsla_hip_matrix_t *out, int format) {
  if (!out) return fail(TBSLA_HIP_ERR_INVALID, "null argument");
  *out = new tbsla_hip_matrix_s();
  (*out)->format = format;
  return TBSLA_HIP_SUCCESS;
}

int tbsla_hip_destroy(tbsla_hip_matrix_t m) {
  if (m) {
    drop(m);
    delete m;
  }
  return TBSLA_HIP_SUCCESS;
}

int tbsla_hip_get_geometry(tbsla_hip_matrix_t m, tbsla_hip_geometry *g) {
  if (!m || !g) return fail(TBSLA_HIP_ERR_INVALID, "null argument");
  *g = tbsla_hip_geometry{};
  g->n_row = m->n_row;
  g->n_col = m->n_col;
  g->f_row = m->f_row;
  g->f_col = m->f_col;
  g->ln_row = m->ln_row;
  g->ln_col = m->ln_col;
  g->pr = m->pr;
  g->pc = m->pc;
  g->NR = m->NR;
  g->NC = m->NC;
  g->nnz = m->nnz;
  g->stored = (long long)slots_of(m);
  g->max_col = m->max_col;
  g->format = m->format;
  return TBSLA_HIP_SUCCESS;
}

int tbsla_hip_import(tbsla_hip_matrix_t m, int format, const tbsla_hip_geometry *g,
                     const int *h_a, const int *h_b, const double *h_v, void *) {
  if (!m || !g) return fail(TBSLA_HIP_ERR_INVALID, "null argument");
  drop(m);
  m->format = format;
  m->n_row = g->n_row;
  m->n_col = g->n_col;
  m->f_row = g->f_row;
  m->f_col = g->f_col;
  m->ln_row = g->ln_row;
  m->ln_col = g->ln_col;
  m->pr = g->pr;
  m->pc = g->pc;
  m->NR = g->NR;
  m->NC = g->NC;
  m->nnz = g->nnz;
  m->max_col = format == TBSLA_HIP_ELL ? g->max_col : 0;
  switch (format) {
    case TBSLA_HIP_CSR:
      if (!h_a) return fail(TBSLA_HIP_ERR_INVALID, "CSR import needs rowptr");
      m->stored = h_a[g->ln_row];
      m->rowptr = host_copy(h_a, (size_t)g->ln_row + 1);
      m->colidx = host_copy(h_b, (size_t)m->stored);
      m->values = host_copy(h_v, (size_t)m->stored);
      break;
    case TBSLA_HIP_COO:
    case TBSLA_HIP_SCOO:
      m->stored = g->stored;
      m->rowidx = host_copy(h_a, (size_t)m->stored);
      m->colidx = host_copy(h_b, (size_t)m->stored);
      m->values = host_copy(h_v, (size_t)m->stored);
      break;
    case TBSLA_HIP_ELL:
      m->stored = (long long)slots_of(m);
      m->colidx = host_copy(h_a, (size_t)m->stored);
      m->values = host_copy(h_v, (size_t)m->stored);
      break;
    default:
      m->stored = (long long)slots_of(m);
      m->values = host_copy(h_v, (size_t)m->stored);
      break;
  }
  m->has_arrays = 1;
  return TBSLA_HIP_SUCCESS;
}

int tbsla_hip_export(tbsla_hip_matrix_t m, int *h_a, int *h_b, double *h_v, void *) {
  if (!m) return fail(TBSLA_HIP_ERR_INVALID, "null argument");
  const size_t n = slots_of(m);
  switch (m->format) {
    case TBSLA_HIP_CSR:
      if (h_a) std::memcpy(h_a, m->rowptr, ((size_t)m->ln_row + 1) * sizeof(int));
      if (h_b && n) std::memcpy(h_b, m->colidx, n * sizeof(int));
      break;
    case TBSLA_HIP_COO:
    case TBSLA_HIP_SCOO:
      if (h_a && n) std::memcpy(h_a, m->rowidx, n * sizeof(int));
      if (h_b && n) std::memcpy(h_b, m->colidx, n * sizeof(int));
      break;
    case TBSLA_HIP_ELL:
      if (h_a && n) std::memcpy(h_a, m->colidx, n * sizeof(int));
      break;
    default:
      break;
  }
  if (h_v && n) std::memcpy(h_v, m->values, n * sizeof(double));
  return TBSLA_HIP_SUCCESS;
}
}

// Entry points the class's virtual methods name but the host checks never
// call: refused, so a call by mistake is reported instead of running.
extern "C" {
int tbsla_hip_malloc(void **p, size_t bytes) {
  *p = std::malloc(std::max<size_t>(bytes, 1));
  return TBSLA_HIP_SUCCESS;
}
int tbsla_hip_free(void *p) {
  std::free(p);
  return TBSLA_HIP_SUCCESS;
}
int tbsla_hip_memcpy_dtoh(void *dst, const void *src, size_t bytes, void *) {
  if (bytes) std::memcpy(dst, src, bytes);
  return TBSLA_HIP_SUCCESS;
}
int tbsla_hip_page_rank(tbsla_hip_matrix_t, double, double, int, double *, int *, double *, void *) {
  return fail(TBSLA_HIP_ERR_UNSUPPORTED, "not provided by the host mock");
}
}
