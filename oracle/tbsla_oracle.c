/*
 * tbsla_oracle.c -- CPU restatement of TBSLA's sparse hot path.
 *
 * TEST INFRASTRUCTURE.  This is the oracle of the HIP library: tests/, the
 * smoke check and the CPU-baseline leg of bench.py load it (liboracle.so);
 * nothing in the product links, loads or calls it.
 *
 * Each routine restates the reference routine cited beside it, in plain C,
 * serially (the reference's OpenMP loops are run single-threaded so the
 * floating-point order is fixed), with the reference's quirks kept where they
 * change results:
 *   - the tail-row integer division in cqmat_value (every tail entry is
 *     rewired when q > 0, values_generation.cpp:45);
 *   - CSR fill_cqmat's count pass over n*c generator indices, which makes the
 *     nnz field larger than rowptr[ln_row] (MatrixCSR.cpp:423-434);
 *   - ELL fill_cqmat's max_col tail-row quirk (MatrixELL.cpp:408-412);
 *   - seed_mult == 0 behaving like seed_mult == 1 (values_generation.cpp:29-32).
 * Documented deviations (the reference is undefined or racy there):
 *   - COO -> CSR/ELL sorts are stable (the reference uses std::sort, whose
 *     tie order between equal (row, col) keys is unspecified);
 *   - get_col_sums runs serially (MatrixCSR.cpp:1200-1206 races under OpenMP);
 *   - arrays the reference leaves unallocated or uninitialised are zeroed;
 *   - CG uses the corrected argument order of Ax (CG.cpp:61 swaps them).
 *
 * Compiled with -ffp-contract=off so that every a*b+c is two roundings, as
 * in an x86-64 build of the reference without FMA.
 */
#include "tbsla_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_RAND_MAX 2147483647 /* glibc RAND_MAX, <stdlib.h> */

static long lmin(long a, long b) { return a < b ? a : b; }
static long lmax(long a, long b) { return a > b ? a : b; }
static int imin(int a, int b) { return a < b ? a : b; }
static int imin3(int a, int b, int c) { return imin(imin(a, b), c); }

static void *zalloc(size_t count, size_t size) {
  if (count == 0) count = 1;
  return calloc(count, size);
}

/* ------------------------------------------------------------------------ */
/* Generators and partition helpers                                          */
/* ------------------------------------------------------------------------ */

/* glibc rand_r (stdlib/rand_r.c, glibc 2.35): three steps of the LCG
 * next = next * 1103515245 + 12345 (mod 2^32); bits 16..26 of the first
 * state, then bits 16..25 of the next two, concatenated. */
int orc_rand_r(unsigned *seed) {
  unsigned next = *seed;
  int result;
  next = next * 1103515245u + 12345u;
  result = (int)((next >> 16) & 2047u);
  next = next * 1103515245u + 12345u;
  result = (result << 10) ^ (int)((next >> 16) & 1023u);
  next = next * 1103515245u + 12345u;
  result = (result << 10) ^ (int)((next >> 16) & 1023u);
  *seed = next;
  return result;
}

/* values_generation::cqmat_value (src/cpp/utils/values_generation.cpp:28-47).
 * Generator index k -> (row, col).  Full rows hold c consecutive columns
 * starting at the row index; each entry is rewired to rand_r % nc with
 * probability q.  Tail rows are decoded by the while loop at :40-43 and
 * use an integer division (:45) that makes the rewire test `0 < q`. */
void orc_cqmat_value(uint64_t k, uint64_t nr, uint64_t nc, uint64_t c_in, double q,
                     unsigned seed_mult, int64_t *row, int64_t *col) {
  unsigned seed = (unsigned)k;
  if (seed_mult > 0) seed = (unsigned)((uint64_t)seed_mult * k);
  uint64_t c = nc < c_in ? nc : c_in;
  uint64_t nfull = (nc - c + 1) < nr ? (nc - c + 1) : nr;
  if (k < c * nfull) {
    uint64_t r = k / c;
    int draw = orc_rand_r(&seed);
    *row = (int64_t)r;
    if ((double)draw / ORC_RAND_MAX < q)
      *col = (int64_t)((uint64_t)orc_rand_r(&seed) % nc);
    else
      *col = (int64_t)(r + k % c);
  } else {
    uint64_t it = k - c * nfull;
    uint64_t cur = 1;
    while (cur < nr - nfull && it >= c - cur) {
      it -= c - cur;
      cur++;
    }
    cur += nfull - 1;
    int draw = orc_rand_r(&seed);
    *row = (int64_t)cur;
    if ((double)(draw / ORC_RAND_MAX) < q)
      *col = (int64_t)((uint64_t)orc_rand_r(&seed) % nc);
    else
      *col = (int64_t)(cur + it);
  }
}

/* values_generation::cdiag_value (values_generation.cpp:6-26): the i-th
 * entry of the two diagonals at distance cdiag, interleaved by row. */
void orc_cdiag_value(uint64_t i, uint64_t nv, uint64_t nr, uint64_t nc, uint64_t cdiag,
                     int64_t *row, int64_t *col) {
  (void)nv;
  (void)nr;
  int64_t ii = (int64_t)i, cd = (int64_t)cdiag, n = (int64_t)nc;
  if (cdiag == 0) {
    *row = ii;
    *col = ii;
    return;
  }
  int64_t lim1 = lmax(lmin(n - cd, cd), 0);
  if (ii < lim1) {
    *row = ii;
    *col = ii + cd;
    return;
  }
  int64_t lim2 = lmax(cd + 2 * (n - 2 * cd), 0);
  if (ii < lim2) {
    int64_t it = (int64_t)((i + cdiag) / 2);
    *row = it;
    *col = (i % 2 == 0) ? it - cd : it + cd;
    return;
  }
  int64_t it = ii - (n - 2 * cd);
  if (cdiag > nc) it -= cd - n;
  *row = it;
  *col = it - cd;
}

/* range::lnv / range::pflv (src/cpp/utils/range.cpp:7-26): block length and
 * first index when `size` is split into `npos` blocks, the first size%npos
 * blocks one longer. */
long orc_lnv(long size, int pos, int npos) {
  long n = size / npos;
  long mod = size % npos;
  if (pos < mod) n++;
  return n;
}

long orc_pflv(long size, int pos, int npos) {
  long mod = size % npos;
  long n = orc_lnv(size, pos, npos) * pos;
  if (pos >= mod) n += mod;
  return n;
}

/* Project value generator (no reference counterpart: the reference's values
 * are all 1.0).  V(seed, k) = top 53 bits of splitmix64(splitmix64(seed) ^ k)
 * times 2^-53, uniform in [0, 1).  Integer-only, so host and device agree
 * bit for bit.  splitmix64 is Steele, Lea & Flood, OOPSLA 2014. */
static uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

double orc_unit_value(uint64_t seed, uint64_t k) {
  uint64_t h = mix64(mix64(seed) ^ k);
  return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

static double entry_value(int64_t value_seed, uint64_t k) {
  return value_seed < 0 ? 1.0 : orc_unit_value((uint64_t)value_seed, k);
}

/* Global cqmat entry count (MatrixCOO.cpp:321-327). */
long orc_cqmat_gnv(int n_row, int n_col, int c) {
  int cc = imin(c, n_col);
  long gnv = 0;
  int full = imin(n_col - cc + 1, n_row);
  for (long i = 0; i < full; i++) gnv += cc;
  for (long i = 0; i < imin(n_row, n_col) - full; i++) gnv += cc - i - 1;
  return gnv;
}

/* Generator offset of row f_row and entry count of the block's rows
 * (MatrixCSR.cpp:393-417; the same loops appear in every format). */
void orc_cqmat_prescan(int n_row, int n_col, int c, int f_row, int ln_row, long *incr_out,
                       long *nv_out) {
  int cc = imin(c, n_col);
  int min_ = imin(n_col - cc + 1, n_row);
  long incr = 0, nv = 0;
  for (long i = 0; i < min_; i++) {
    if (i < f_row) incr += cc;
    if (i >= f_row && i < f_row + ln_row) nv += cc;
    if (i >= f_row + ln_row) break;
  }
  for (long i = 0; i < imin(n_row, n_col) - min_; i++) {
    if (i + min_ < f_row) incr += cc - i - 1;
    if (i + min_ >= f_row && i + min_ < f_row + ln_row) nv += cc - i - 1;
    if (i + min_ >= f_row + ln_row) break;
  }
  *incr_out = incr;
  *nv_out = nv;
}

static void set_block(orc_geom *g, int n_row, int n_col, int pr, int pc, int NR, int NC) {
  g->n_row = n_row;
  g->n_col = n_col;
  g->pr = pr;
  g->pc = pc;
  g->NR = NR;
  g->NC = NC;
  g->ln_row = (int)orc_lnv(n_row, pr, NR);
  g->f_row = (int)orc_pflv(n_row, pr, NR);
  g->ln_col = (int)orc_lnv(n_col, pc, NC);
  g->f_col = (int)orc_pflv(n_col, pc, NC);
}

static int in_block(const orc_geom *g, int64_t ii, int64_t jj) {
  return ii >= g->f_row && ii < (int64_t)g->f_row + g->ln_row && jj >= g->f_col &&
         jj < (int64_t)g->f_col + g->ln_col;
}

#define DECODE(k)                                                                            \
  orc_cqmat_value((uint64_t)(k), (uint64_t)n_row, (uint64_t)n_col, (uint64_t)c, q, seed_mult, \
                  &ii, &jj)

/* ------------------------------------------------------------------------ */
/* fill_cqmat                                                                */
/* ------------------------------------------------------------------------ */

/* MatrixCSR::fill_cqmat (MatrixCSR.cpp:373-486). */
int orc_csr_fill_cqmat(orc_csr *m, int n_row, int n_col, int c, double q, unsigned seed_mult,
                       int pr, int pc, int NR, int NC, int64_t value_seed) {
  memset(m, 0, sizeof *m);
  orc_geom *g = &m->g;
  set_block(g, n_row, n_col, pr, pc, NR, NC);
  int cc = imin(c, n_col);
  int min_ = imin(n_col - cc + 1, n_row);
  long incr, nv;
  orc_cqmat_prescan(n_row, n_col, c, g->f_row, g->ln_row, &incr, &nv);
  int64_t ii, jj;
  /* count pass: c generator indices for every row of the block (:423-434) */
  long nnz = 0, k = incr;
  int row_end = imin3(n_row, n_col, g->f_row + g->ln_row);
  for (long i = g->f_row; i < row_end; i++)
    for (long j = 0; j < cc; j++, k++) {
      DECODE(k);
      if (in_block(g, ii, jj)) nnz++;
    }
  g->nnz = nnz;
  m->rowptr = (int *)zalloc((size_t)g->ln_row + 1, sizeof(int));
  if (nv == 0) return 0;
  m->colidx = (int *)zalloc((size_t)nnz, sizeof(int));
  m->values = (double *)zalloc((size_t)nnz, sizeof(double));
  /* fill pass: the block's own generator entries, in generation order */
  long lincr = 0, i;
  k = incr;
  for (i = g->f_row; i < imin(min_, g->f_row + g->ln_row); i++) {
    for (long j = 0; j < cc; j++, k++) {
      DECODE(k);
      if (in_block(g, ii, jj)) {
        m->colidx[lincr] = (int)jj;
        m->values[lincr] = entry_value(value_seed, (uint64_t)k);
        lincr++;
      }
    }
    m->rowptr[i - g->f_row + 1] = (int)lincr;
  }
  for (; i < row_end; i++) {
    for (long j = 0; j < cc - i + min_ - 1; j++, k++) {
      DECODE(k);
      if (in_block(g, ii, jj)) {
        m->colidx[lincr] = (int)jj;
        m->values[lincr] = entry_value(value_seed, (uint64_t)k);
        lincr++;
      }
    }
    m->rowptr[i - g->f_row + 1] = (int)lincr;
  }
  for (; i < g->f_row + g->ln_row; i++) m->rowptr[i - g->f_row + 1] = (int)lincr;
  m->stored = lincr;
  return 0;
}

/* MatrixELL::fill_cqmat (MatrixELL.cpp:335-466), row-major storage. */
int orc_ell_fill_cqmat(orc_ell *m, int n_row, int n_col, int c, double q, unsigned seed_mult,
                       int pr, int pc, int NR, int NC, int64_t value_seed) {
  memset(m, 0, sizeof *m);
  orc_geom *g = &m->g;
  set_block(g, n_row, n_col, pr, pc, NR, NC);
  int cc = imin(c, n_col);
  int min_ = imin(n_col - cc + 1, n_row);
  long incr, nv;
  orc_cqmat_prescan(n_row, n_col, c, g->f_row, g->ln_row, &incr, &nv);
  int64_t ii, jj;
  long nnz = 0, k = incr, i;
  int max_col = 0;
  for (i = g->f_row; i < imin(min_, g->f_row + g->ln_row); i++) {
    int nbc = 0;
    for (long j = 0; j < cc; j++, k++) {
      DECODE(k);
      if (in_block(g, ii, jj)) {
        nnz++;
        nbc++;
      }
    }
    if (nbc > max_col) max_col = nbc;
  }
  for (; i < imin3(n_row, n_col, g->f_row + g->ln_row); i++) {
    int nbc = 0;
    long j = 0;
    for (; j < cc - i + min_ - 1; j++, k++) {
      DECODE(k);
      if (in_block(g, ii, jj)) {
        nbc++;
        nnz++;
      }
    }
    /* quirk (:408-412): the remaining slot indices j are tested against the
     * block's column range and counted */
    for (; j < cc; j++)
      if (j >= g->f_col && j < (long)g->f_col + g->ln_col) nbc++;
    if (nbc > max_col) max_col = nbc;
  }
  g->nnz = nnz;
  m->max_col = max_col;
  if (nv == 0) return 0;
  size_t slots = (size_t)g->ln_row * (size_t)max_col;
  m->values = (double *)zalloc(slots, sizeof(double));
  m->columns = (int *)zalloc(slots, sizeof(int));
  k = incr;
  for (i = g->f_row; i < imin(min_, g->f_row + g->ln_row); i++) {
    long lincr = 0;
    for (long j = 0; j < cc; j++, k++) {
      DECODE(k);
      if (in_block(g, ii, jj)) {
        size_t s = (size_t)(i - g->f_row) * max_col + lincr;
        m->columns[s] = (int)jj;
        m->values[s] = entry_value(value_seed, (uint64_t)k);
        lincr++;
      }
    }
  }
  for (; i < imin(n_row, g->f_row + g->ln_row); i++) {
    long lincr = 0;
    for (long j = 0; j < cc - i + min_ - 1; j++, k++) {
      DECODE(k);
      if (in_block(g, ii, jj)) {
        size_t s = (size_t)(i - g->f_row) * max_col + lincr;
        m->columns[s] = (int)jj;
        m->values[s] = entry_value(value_seed, (uint64_t)k);
        lincr++;
      }
    }
  }
  return 0;
}

/* MatrixCOO::fill_cqmat (MatrixCOO.cpp:320-354): the (pr*NC+pc)-th slice of
 * the global generation order; geometry stays global. */
int orc_coo_fill_cqmat(orc_coo *m, int n_row, int n_col, int c, double q, unsigned seed_mult,
                       int pr, int pc, int NR, int NC, int64_t value_seed) {
  memset(m, 0, sizeof *m);
  orc_geom *g = &m->g;
  long gnv = orc_cqmat_gnv(n_row, n_col, c);
  g->pr = pr;
  g->pc = pc;
  g->NR = NR;
  g->NC = NC;
  g->ln_row = g->n_row = n_row;
  g->ln_col = g->n_col = n_col;
  if (gnv == 0) return 0;
  long s = orc_pflv(gnv, pr * NC + pc, NR * NC);
  long n = orc_lnv(gnv, pr * NC + pc, NR * NC);
  g->nnz = n;
  /* MatrixCOO::init (MatrixCOO.cpp:14-34), called here without restoring the
   * grid fields afterwards (fill_cdiag does restore them, :303-306): the
   * block records pr = pc = 0, NR = NC = 1. */
  g->pr = g->pc = 0;
  g->NR = g->NC = 1;
  m->len = n;
  m->row = (int *)zalloc((size_t)n, sizeof(int));
  m->col = (int *)zalloc((size_t)n, sizeof(int));
  m->values = (double *)zalloc((size_t)n, sizeof(double));
  int64_t ii, jj;
  for (long i = 0; i < n; i++) {
    DECODE(i + s);
    m->row[i] = (int)ii;
    m->col[i] = (int)jj;
    m->values[i] = entry_value(value_seed, (uint64_t)(i + s));
  }
  return 0;
}

/* MatrixSCOO::fill_cqmat (MatrixSCOO.cpp:295-403): block-filtered triplets
 * with global indices, in generation order; nnz ends as the stored count. */
int orc_scoo_fill_cqmat(orc_coo *m, int n_row, int n_col, int c, double q, unsigned seed_mult,
                        int pr, int pc, int NR, int NC, int64_t value_seed) {
  memset(m, 0, sizeof *m);
  orc_geom *g = &m->g;
  set_block(g, n_row, n_col, pr, pc, NR, NC);
  int cc = imin(c, n_col);
  int min_ = imin(n_col - cc + 1, n_row);
  long incr, nv;
  orc_cqmat_prescan(n_row, n_col, c, g->f_row, g->ln_row, &incr, &nv);
  int64_t ii, jj;
  long nnz = 0, k = incr, i;
  int row_end = imin3(n_row, n_col, g->f_row + g->ln_row);
  for (i = g->f_row; i < row_end; i++)
    for (long j = 0; j < cc; j++, k++) {
      DECODE(k);
      if (in_block(g, ii, jj)) nnz++;
    }
  g->nnz = nnz;
  if (nv == 0) return 0;
  m->row = (int *)zalloc((size_t)nv, sizeof(int));
  m->col = (int *)zalloc((size_t)nv, sizeof(int));
  m->values = (double *)zalloc((size_t)nv, sizeof(double));
  long idx = 0;
  k = incr;
  for (i = g->f_row; i < imin(min_, g->f_row + g->ln_row); i++)
    for (long j = 0; j < cc; j++, k++) {
      DECODE(k);
      if (in_block(g, ii, jj)) {
        m->row[idx] = (int)ii;
        m->col[idx] = (int)jj;
        m->values[idx] = entry_value(value_seed, (uint64_t)k);
        idx++;
      }
    }
  for (; i < row_end; i++)
    for (long j = 0; j < cc - i + min_ - 1; j++, k++) {
      DECODE(k);
      if (in_block(g, ii, jj)) {
        m->row[idx] = (int)ii;
        m->col[idx] = (int)jj;
        m->values[idx] = entry_value(value_seed, (uint64_t)k);
        idx++;
      }
    }
  g->nnz = idx;
  m->len = idx;
  return 0;
}

/* MatrixDENSE::fill_cqmat (MatrixDENSE.cpp:187-270): duplicates accumulate
 * in generation order and are each counted in nnz. */
int orc_dense_fill_cqmat(orc_dense *m, int n_row, int n_col, int c, double q, unsigned seed_mult,
                         int pr, int pc, int NR, int NC, int64_t value_seed) {
  memset(m, 0, sizeof *m);
  orc_geom *g = &m->g;
  set_block(g, n_row, n_col, pr, pc, NR, NC);
  int cc = imin(c, n_col);
  int min_ = imin(n_col - cc + 1, n_row);
  long incr, nv;
  orc_cqmat_prescan(n_row, n_col, c, g->f_row, g->ln_row, &incr, &nv);
  if (nv == 0) return 0;
  m->values = (double *)zalloc((size_t)g->ln_row * (size_t)g->ln_col, sizeof(double));
  int64_t ii, jj;
  long nnz = 0, k = incr, i;
  for (i = g->f_row; i < imin(min_, g->f_row + g->ln_row); i++)
    for (long j = 0; j < cc; j++, k++) {
      DECODE(k);
      if (in_block(g, ii, jj)) {
        m->values[(size_t)(ii - g->f_row) * g->ln_col + (jj - g->f_col)] +=
            entry_value(value_seed, (uint64_t)k);
        nnz++;
      }
    }
  for (; i < imin(n_row, g->f_row + g->ln_row); i++)
    for (long j = 0; j < cc - i + min_ - 1; j++, k++) {
      DECODE(k);
      if (in_block(g, ii, jj)) {
        m->values[(size_t)(ii - g->f_row) * g->ln_col + (jj - g->f_col)] +=
            entry_value(value_seed, (uint64_t)k);
        nnz++;
      }
    }
  g->nnz = nnz;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* fill_cdiag                                                                */
/* ------------------------------------------------------------------------ */

/* Number of entries of the block on the diagonals at distance cdiag, as the
 * count loops of every fill_cdiag (e.g. MatrixCSR.cpp:323-338). */
static long cdiag_count(const orc_geom *g, int cdiag) {
  long nnz = 0;
  for (long i = g->f_row; i < (long)g->f_row + g->ln_row; i++) {
    if (in_block(g, i, i - cdiag)) nnz++;
    if (cdiag != 0 && in_block(g, i, i + cdiag)) nnz++;
  }
  return nnz;
}

/* MatrixCSR::fill_cdiag (MatrixCSR.cpp:297-371). */
int orc_csr_fill_cdiag(orc_csr *m, int n_row, int n_col, int cdiag, int pr, int pc, int NR,
                       int NC) {
  memset(m, 0, sizeof *m);
  orc_geom *g = &m->g;
  set_block(g, n_row, n_col, pr, pc, NR, NC);
  g->nnz = cdiag_count(g, cdiag);
  m->rowptr = (int *)zalloc((size_t)g->ln_row + 1, sizeof(int));
  if (g->nnz == 0) return 0;
  m->colidx = (int *)zalloc((size_t)g->nnz, sizeof(int));
  m->values = (double *)zalloc((size_t)g->nnz, sizeof(double));
  long incr = 0;
  for (long i = g->f_row; i < (long)g->f_row + g->ln_row; i++) {
    if (in_block(g, i, i - cdiag)) {
      m->colidx[incr] = (int)(i - cdiag);
      m->values[incr++] = 1;
    }
    if (cdiag != 0 && in_block(g, i, i + cdiag)) {
      m->colidx[incr] = (int)(i + cdiag);
      m->values[incr++] = 1;
    }
    m->rowptr[i - g->f_row + 1] = (int)incr;
  }
  m->stored = incr;
  return 0;
}

/* MatrixELL::fill_cdiag (MatrixELL.cpp:261-333): slot 0 holds the i-cdiag
 * entry, slot 1 the i+cdiag entry; a missing entry leaves (col 0, val 0). */
int orc_ell_fill_cdiag(orc_ell *m, int n_row, int n_col, int cdiag, int pr, int pc, int NR,
                       int NC) {
  memset(m, 0, sizeof *m);
  orc_geom *g = &m->g;
  set_block(g, n_row, n_col, pr, pc, NR, NC);
  m->max_col = cdiag == 0 ? 1 : 2;
  g->nnz = cdiag_count(g, cdiag);
  if (g->nnz == 0) return 0;
  size_t slots = (size_t)g->ln_row * m->max_col;
  m->values = (double *)zalloc(slots, sizeof(double));
  m->columns = (int *)zalloc(slots, sizeof(int));
  size_t idx = 0;
  for (long i = g->f_row; i < (long)g->f_row + g->ln_row; i++) {
    if (in_block(g, i, i - cdiag)) {
      m->columns[idx] = (int)(i - cdiag);
      m->values[idx] = 1;
    }
    idx++;
    if (cdiag != 0) {
      if (in_block(g, i, i + cdiag)) {
        m->columns[idx] = (int)(i + cdiag);
        m->values[idx] = 1;
      }
      idx++;
    }
  }
  return 0;
}

/* MatrixCOO::fill_cdiag (MatrixCOO.cpp:280-318). */
int orc_coo_fill_cdiag(orc_coo *m, int n_row, int n_col, int cdiag, int pr, int pc, int NR,
                       int NC) {
  memset(m, 0, sizeof *m);
  orc_geom *g = &m->g;
  long gnv = lmax(lmin(n_row, (long)n_col - cdiag), 0) + lmax(lmin((long)n_row - cdiag, n_col), 0);
  if (cdiag == 0) gnv /= 2;
  g->pr = pr;
  g->pc = pc;
  g->NR = NR;
  g->NC = NC;
  g->ln_row = g->n_row = n_row;
  g->ln_col = g->n_col = n_col;
  if (gnv == 0) return 0;
  long s = orc_pflv(gnv, pr * NC + pc, NR * NC);
  long n = orc_lnv(gnv, pr * NC + pc, NR * NC);
  g->nnz = n;
  m->len = n;
  m->row = (int *)zalloc((size_t)n, sizeof(int));
  m->col = (int *)zalloc((size_t)n, sizeof(int));
  m->values = (double *)zalloc((size_t)n, sizeof(double));
  for (long i = 0; i < n; i++) {
    int64_t ii, jj;
    orc_cdiag_value((uint64_t)(i + s), (uint64_t)gnv, (uint64_t)n_row, (uint64_t)n_col,
                    (uint64_t)cdiag, &ii, &jj);
    m->row[i] = (int)ii;
    m->col[i] = (int)jj;
    m->values[i] = 1;
  }
  return 0;
}

/* MatrixSCOO::fill_cdiag (MatrixSCOO.cpp:228-293). */
int orc_scoo_fill_cdiag(orc_coo *m, int n_row, int n_col, int cdiag, int pr, int pc, int NR,
                        int NC) {
  memset(m, 0, sizeof *m);
  orc_geom *g = &m->g;
  set_block(g, n_row, n_col, pr, pc, NR, NC);
  g->nnz = cdiag_count(g, cdiag);
  if (g->nnz == 0) return 0;
  m->len = g->nnz;
  m->row = (int *)zalloc((size_t)g->nnz, sizeof(int));
  m->col = (int *)zalloc((size_t)g->nnz, sizeof(int));
  m->values = (double *)zalloc((size_t)g->nnz, sizeof(double));
  long idx = 0;
  for (long i = g->f_row; i < (long)g->f_row + g->ln_row; i++) {
    if (in_block(g, i, i - cdiag)) {
      m->row[idx] = (int)i;
      m->col[idx] = (int)(i - cdiag);
      m->values[idx++] = 1;
    }
    if (cdiag != 0 && in_block(g, i, i + cdiag)) {
      m->row[idx] = (int)i;
      m->col[idx] = (int)(i + cdiag);
      m->values[idx++] = 1;
    }
  }
  return 0;
}

/* MatrixDENSE::fill_cdiag (MatrixDENSE.cpp:129-185). */
int orc_dense_fill_cdiag(orc_dense *m, int n_row, int n_col, int cdiag, int pr, int pc, int NR,
                         int NC) {
  memset(m, 0, sizeof *m);
  orc_geom *g = &m->g;
  set_block(g, n_row, n_col, pr, pc, NR, NC);
  g->nnz = cdiag_count(g, cdiag);
  if (g->nnz == 0) return 0;
  m->values = (double *)zalloc((size_t)g->ln_row * (size_t)g->ln_col, sizeof(double));
  for (long i = g->f_row; i < (long)g->f_row + g->ln_row; i++) {
    if (in_block(g, i, i - cdiag))
      m->values[(size_t)(i - g->f_row) * g->ln_col + (i - cdiag - g->f_col)] = 1;
    if (cdiag != 0 && in_block(g, i, i + cdiag))
      m->values[(size_t)(i - g->f_row) * g->ln_col + (i + cdiag - g->f_col)] = 1;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Conversions from COO                                                      */
/* ------------------------------------------------------------------------ */

typedef struct {
  int64_t key;
  long pos;
} keyed;

static int keyed_cmp(const void *a, const void *b) {
  const keyed *x = (const keyed *)a, *y = (const keyed *)b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  return x->pos < y->pos ? -1 : (x->pos > y->pos);
}

/* Permutation sorting the triplets by (row, col) (MatrixCSR.cpp:62-64 with
 * csr::compare_row, src/cpp/utils/csr.cpp:3-8); ties keep input order. */
static long *row_col_order(const orc_coo *src) {
  long n = src->g.nnz;
  keyed *t = (keyed *)zalloc((size_t)n, sizeof(keyed));
  for (long i = 0; i < n; i++) {
    t[i].key = ((int64_t)src->row[i] << 32) | (uint32_t)src->col[i];
    t[i].pos = i;
  }
  qsort(t, (size_t)n, sizeof(keyed), keyed_cmp);
  long *p = (long *)zalloc((size_t)n, sizeof(long));
  for (long i = 0; i < n; i++) p[i] = t[i].pos;
  free(t);
  return p;
}

static void geom_from_coo(orc_geom *g, const orc_geom *s) {
  *g = *s;
  g->ln_row = s->n_row;
  g->ln_col = s->n_col;
}

/* MatrixCSR(const MatrixCOO&) (MatrixCSR.cpp:49-96). */
int orc_csr_from_coo(orc_csr *m, const orc_coo *src) {
  memset(m, 0, sizeof *m);
  geom_from_coo(&m->g, &src->g);
  long n = src->g.nnz;
  m->rowptr = (int *)zalloc((size_t)m->g.n_row + 1, sizeof(int));
  m->colidx = (int *)zalloc((size_t)n, sizeof(int));
  m->values = (double *)zalloc((size_t)n, sizeof(double));
  if (n == 0) return 0;
  long *p = row_col_order(src);
  for (long i = 0; i < n; i++) {
    m->colidx[i] = src->col[p[i]];
    m->values[i] = src->values[p[i]];
    m->rowptr[src->row[p[i]] + 1]++;
  }
  for (int i = 0; i < m->g.n_row; i++) m->rowptr[i + 1] += m->rowptr[i];
  m->stored = n;
  free(p);
  return 0;
}

/* MatrixELL(const MatrixCOO&) (MatrixELL.cpp:25-79). */
int orc_ell_from_coo(orc_ell *m, const orc_coo *src) {
  memset(m, 0, sizeof *m);
  geom_from_coo(&m->g, &src->g);
  long n = src->g.nnz;
  if (n == 0) {
    m->g.nnz = 0;
    m->max_col = 0;
    return 0;
  }
  long *p = row_col_order(src);
  int *nvrow = (int *)zalloc((size_t)m->g.n_row, sizeof(int));
  for (long i = 0; i < n; i++) nvrow[src->row[i]]++;
  int max_col = 0;
  for (int i = 0; i < m->g.n_row; i++)
    if (nvrow[i] > max_col) max_col = nvrow[i];
  m->max_col = max_col;
  size_t slots = (size_t)m->g.n_row * max_col;
  m->values = (double *)zalloc(slots, sizeof(double));
  m->columns = (int *)zalloc(slots, sizeof(int));
  long incr = 0;
  for (int i = 0; i < m->g.n_row; i++)
    for (int j = 0; j < nvrow[i]; j++, incr++) {
      m->values[(size_t)i * max_col + j] = src->values[p[incr]];
      m->columns[(size_t)i * max_col + j] = src->col[p[incr]];
    }
  free(nvrow);
  free(p);
  return 0;
}

/* MatrixSCOO(const MatrixCOO&) (MatrixSCOO.cpp:75-101): a copy. */
int orc_scoo_from_coo(orc_coo *m, const orc_coo *src) {
  memset(m, 0, sizeof *m);
  geom_from_coo(&m->g, &src->g);
  long n = src->g.nnz;
  m->len = n;
  m->row = (int *)zalloc((size_t)n, sizeof(int));
  m->col = (int *)zalloc((size_t)n, sizeof(int));
  m->values = (double *)zalloc((size_t)n, sizeof(double));
  for (long i = 0; i < n; i++) {
    m->row[i] = src->row[i];
    m->col[i] = src->col[i];
    m->values[i] = src->values[i];
  }
  return 0;
}

/* MatrixDENSE(const MatrixCOO&) (MatrixDENSE.cpp:15-37): scatter-add in
 * triplet order. */
int orc_dense_from_coo(orc_dense *m, const orc_coo *src) {
  memset(m, 0, sizeof *m);
  geom_from_coo(&m->g, &src->g);
  m->values = (double *)zalloc((size_t)m->g.ln_row * (size_t)m->g.ln_col, sizeof(double));
  for (long i = 0; i < src->g.nnz; i++)
    m->values[(size_t)src->row[i] * m->g.ln_col + src->col[i]] += src->values[i];
  return 0;
}

/* Single-block containers for matrices built outside the generators (tests
 * build Laplacians and stochastic matrices this way); arrays are zeroed. */
int orc_csr_alloc(orc_csr *m, int n_row, int n_col, long nnz) {
  memset(m, 0, sizeof *m);
  set_block(&m->g, n_row, n_col, 0, 0, 1, 1);
  m->g.nnz = nnz;
  m->stored = nnz;
  m->rowptr = (int *)zalloc((size_t)n_row + 1, sizeof(int));
  m->colidx = (int *)zalloc((size_t)nnz, sizeof(int));
  m->values = (double *)zalloc((size_t)nnz, sizeof(double));
  return 0;
}

int orc_coo_alloc(orc_coo *m, int n_row, int n_col, long nnz) {
  memset(m, 0, sizeof *m);
  set_block(&m->g, n_row, n_col, 0, 0, 1, 1);
  m->g.nnz = nnz;
  m->len = nnz;
  m->row = (int *)zalloc((size_t)nnz, sizeof(int));
  m->col = (int *)zalloc((size_t)nnz, sizeof(int));
  m->values = (double *)zalloc((size_t)nnz, sizeof(double));
  return 0;
}

/* MatrixCOO::readMM (MatrixCOO.cpp:206-278), the symmetric branch's
 * overwrite (:256-269) fixed: the header's space-separated words
 * (utils/split.cpp:6-19) must be "%%MatrixMarket matrix coordinate real"
 * and "general" or "symmetric"; '%' lines are skipped (blank ones too);
 * "n_row n_col n_values"; then n_values whitespace-separated triples
 * "row col value", 1-based, read in order, each off-diagonal entry of a
 * symmetric file followed by its mirror.  Returns 0; -1 when the file cannot
 * be opened; -2 for a header the reference rejects
 * (MatrixFormatReadException); -3 for fewer or unparsable entries or an
 * index outside the matrix. */
int orc_coo_read_mm(orc_coo *m, const char *path) {
  FILE *f = fopen(path, "rb");
  if (!f) return -1;
  fseek(f, 0, SEEK_END);
  long size = ftell(f);
  fseek(f, 0, SEEK_SET);
  char *buf = (char *)malloc((size_t)size + 1);
  size_t got = fread(buf, 1, (size_t)size, f);
  fclose(f);
  buf[got] = 0;
  int rc = 0;
  /* header words */
  char *eol = strchr(buf, '\n');
  size_t hl = eol ? (size_t)(eol - buf) : got;
  if (hl > 0 && buf[hl - 1] == '\r') hl--;
  char *words[5];
  int nw = 0;
  size_t a = 0;
  char *first = (char *)malloc(hl + 1);
  memcpy(first, buf, hl);
  first[hl] = 0;
  while (a < hl && nw < 5) {
    size_t b = a;
    while (b < hl && first[b] != ' ') b++;
    if (b > a) {
      first[b] = 0;
      words[nw++] = first + a;
    }
    a = b + 1;
  }
  if (nw < 5 || strcmp(words[0], "%%MatrixMarket") || strcmp(words[1], "matrix") ||
      strcmp(words[2], "coordinate") || strcmp(words[3], "real") ||
      (strcmp(words[4], "general") && strcmp(words[4], "symmetric"))) {
    free(first);
    free(buf);
    return -2;
  }
  const int sym = strcmp(words[4], "symmetric") == 0;
  free(first);
  /* comments and blank lines, then the size line */
  char *p = eol ? eol + 1 : buf + got;
  for (;;) {
    if (!*p) {
      free(buf);
      return -3;
    }
    char *e = strchr(p, '\n');
    char *k = p;
    while (*k == ' ' || *k == '\t' || *k == '\r') k++;
    if (*p != '%' && *k != '\n' && *k != 0) break;
    if (!e) {
      free(buf);
      return -3;
    }
    p = e + 1;
  }
  char *q;
  long nr = strtol(p, &q, 10);
  long nc = strtol(q, &q, 10);
  long nv = strtol(q, &q, 10);
  if (nr < 0 || nc < 0 || nv < 0) {
    free(buf);
    return -3;
  }
  q = strchr(q, '\n');
  p = q ? q + 1 : buf + got;
  orc_coo_alloc(m, (int)nr, (int)nc, sym ? 2 * nv : nv);
  long n = 0;
  for (long i = 0; i < nv; i++) {
    char *e1, *e2, *e3;
    long r = strtol(p, &e1, 10);
    long c = strtol(e1, &e2, 10);
    double v = strtod(e2, &e3);
    if (e1 == p || e2 == e1 || e3 == e2 || r < 1 || r > nr || c < 1 || c > nc) {
      rc = -3;
      break;
    }
    p = e3;
    m->row[n] = (int)(r - 1);
    m->col[n] = (int)(c - 1);
    m->values[n] = v;
    n++;
    if (sym && r != c) {
      m->row[n] = (int)(c - 1);
      m->col[n] = (int)(r - 1);
      m->values[n] = v;
      n++;
    }
  }
  free(buf);
  m->g.nnz = n;
  m->len = n;
  return rc;
}

void orc_csr_free(orc_csr *m) {
  free(m->rowptr);
  free(m->colidx);
  free(m->values);
  memset(m, 0, sizeof *m);
}
void orc_ell_free(orc_ell *m) {
  free(m->columns);
  free(m->values);
  memset(m, 0, sizeof *m);
}
void orc_coo_free(orc_coo *m) {
  free(m->row);
  free(m->col);
  free(m->values);
  memset(m, 0, sizeof *m);
}
void orc_dense_free(orc_dense *m) {
  free(m->values);
  memset(m, 0, sizeof *m);
}

/* ------------------------------------------------------------------------ */
/* SpMV                                                                      */
/* ------------------------------------------------------------------------ */

static const orc_geom *geom_of(int fmt, const void *m) {
  switch (fmt) {
    case ORC_CSR: return &((const orc_csr *)m)->g;
    case ORC_ELL: return &((const orc_ell *)m)->g;
    case ORC_DENSE: return &((const orc_dense *)m)->g;
    default: return &((const orc_coo *)m)->g;
  }
}

/* MatrixCSR::Ax (MatrixCSR.cpp:209-233): overwrites r. */
static void csr_Ax(const orc_csr *m, double *r, const double *v) {
  const orc_geom *g = &m->g;
  if (g->nnz == 0) return;
  for (int i = 0; i < g->ln_row; i++) {
    double tmp = 0;
    for (int j = m->rowptr[i]; j < m->rowptr[i + 1]; j++)
      tmp += m->values[j] * v[m->colidx[j] - g->f_col];
    r[i] = tmp;
  }
}

/* Threads of the OpenMP loops below (the CPU baseline's thread count). */
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

/* The same loop under `omp parallel for schedule(static)`, the CPU baseline. */
void orc_csr_Ax_omp(const orc_csr *m, double *r, const double *v) {
  const orc_geom *g = &m->g;
  if (g->nnz == 0) return;
  const int *rp = m->rowptr, *ci = m->colidx;
  const double *va = m->values;
  const double *vv = v - g->f_col;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < g->ln_row; i++) {
    double tmp = 0;
    for (int j = rp[i]; j < rp[i + 1]; j++) tmp += va[j] * vv[ci[j]];
    r[i] = tmp;
  }
}

/* MatrixELL::Ax (MatrixELL.cpp:181-205): overwrites r; with f_col != 0 a
 * negative local column index is clamped to 0. */
static void ell_Ax(const orc_ell *m, double *r, const double *v) {
  const orc_geom *g = &m->g;
  if (g->nnz == 0 || m->max_col == 0 || m->values == NULL || m->columns == NULL) return;
  for (int i = 0; i < g->ln_row; i++) {
    double tmp = 0;
    for (int j = 0; j < m->max_col; j++) {
      size_t s = (size_t)i * m->max_col + j;
      int idx = m->columns[s] - g->f_col;
      if (g->f_col != 0 && idx < 0) idx = 0;
      tmp += m->values[s] * v[idx];
    }
    r[i] = tmp;
  }
}

/* MatrixCOO::Ax (MatrixCOO.cpp:135-142): accumulates into r, global
 * column index, row offset by vect_incr. */
static void coo_Ax(const orc_coo *m, double *r, const double *v, int vect_incr) {
  for (long i = 0; i < m->g.nnz; i++) r[m->row[i] + vect_incr] += m->values[i] * v[m->col[i]];
}

/* MatrixSCOO::Ax (MatrixSCOO.cpp:154-164): accumulates, local indices. */
static void scoo_Ax(const orc_coo *m, double *r, const double *v, int vect_incr) {
  const orc_geom *g = &m->g;
  if (g->nnz == 0) return;
  for (long i = 0; i < g->nnz; i++)
    r[m->row[i] + vect_incr - g->f_row] += m->values[i] * v[m->col[i] - g->f_col];
}

/* MatrixDENSE::Ax (MatrixDENSE.cpp:77-86): r[i] += A[i,:] . v. */
static void dense_Ax(const orc_dense *m, double *r, const double *v) {
  const orc_geom *g = &m->g;
  if (g->nnz == 0 || m->values == NULL) return;
  for (int i = 0; i < g->ln_row; i++)
    for (int j = 0; j < g->ln_col; j++) r[i] += m->values[(size_t)i * g->ln_col + j] * v[j];
}

void orc_Ax(int fmt, const void *m, double *r, const double *v, int vect_incr) {
  switch (fmt) {
    case ORC_CSR: csr_Ax((const orc_csr *)m, r, v); break;
    case ORC_ELL: ell_Ax((const orc_ell *)m, r, v); break;
    case ORC_COO: coo_Ax((const orc_coo *)m, r, v, vect_incr); break;
    case ORC_SCOO: scoo_Ax((const orc_coo *)m, r, v, vect_incr); break;
    case ORC_DENSE: dense_Ax((const orc_dense *)m, r, v); break;
  }
}

/* Length of the vector spmv returns: n_row for COO (MatrixCOO.cpp:126),
 * ln_row otherwise. */
int orc_result_len(int fmt, const void *m) {
  const orc_geom *g = geom_of(fmt, m);
  return fmt == ORC_COO ? g->n_row : g->ln_row;
}

/* <Format>::spmv, e.g. MatrixCSR.cpp:144-152: zeroed result, then Ax. */
void orc_spmv(int fmt, const void *m, double *r, const double *v, int vect_incr) {
  memset(r, 0, sizeof(double) * (size_t)orc_result_len(fmt, m));
  orc_Ax(fmt, m, r, v, vect_incr);
}

/* Matrix::a_axpx_ (src/cpp/matrix/Matrix.cpp:5-14): A(Av + v). */
void orc_a_axpx_(int fmt, const void *m, double *r, const double *v, int vect_incr) {
  const orc_geom *g = geom_of(fmt, m);
  int len = orc_result_len(fmt, m);
  double *t = (double *)zalloc((size_t)len, sizeof(double));
  orc_spmv(fmt, m, t, v, vect_incr);
  for (int i = 0; i < g->ln_col; i++) t[i] += v[i];
  orc_spmv(fmt, m, r, t, vect_incr);
  free(t);
}

/* Matrix::AAxpAx (Matrix.cpp:16-23): Ax into the caller's buffer (no
 * zeroing), buffer += v, then Ax into r. */
void orc_AAxpAx(int fmt, const void *m, double *r, const double *v, double *buf, int vect_incr) {
  const orc_geom *g = geom_of(fmt, m);
  orc_Ax(fmt, m, buf, v, vect_incr);
  for (int i = 0; i < g->ln_col; i++) buf[i] += v[i];
  orc_Ax(fmt, m, r, buf, vect_incr);
}

/* Matrix::AAxpAxpx (Matrix.cpp:25-31). */
void orc_AAxpAxpx(int fmt, const void *m, double *r, const double *v, double *buf, int vect_incr) {
  const orc_geom *g = geom_of(fmt, m);
  orc_AAxpAx(fmt, m, r, v, buf, vect_incr);
  for (int i = 0; i < g->ln_col; i++) r[i] += v[i];
}

/* ------------------------------------------------------------------------ */
/* Column normalisation                                                      */
/* ------------------------------------------------------------------------ */

/* get_col_sums: CSR (MatrixCSR.cpp:1192-1213, run serially), ELL
 * (MatrixELL.cpp:803-816, positive values only), SCOO
 * (MatrixSCOO.cpp:713-724); COO and DENSE are no-ops (MatrixCOO.cpp:390,
 * MatrixDENSE.cpp:312).  s is accumulated into, not zeroed. */
void orc_get_col_sums(int fmt, const void *m, double *s) {
  if (fmt == ORC_CSR) {
    const orc_csr *a = (const orc_csr *)m;
    if (a->g.nnz == 0) return;
    for (int i = 0; i < a->g.ln_row; i++)
      for (int j = a->rowptr[i]; j < a->rowptr[i + 1]; j++)
        s[a->colidx[j] - a->g.f_col] += a->values[j];
  } else if (fmt == ORC_ELL) {
    const orc_ell *a = (const orc_ell *)m;
    if (a->g.nnz == 0) return;
    for (int i = 0; i < a->g.ln_row; i++)
      for (int j = 0; j < a->max_col; j++) {
        size_t k = (size_t)i * a->max_col + j;
        if (a->values[k] > 0) s[a->columns[k] - a->g.f_col] += a->values[k];
      }
  } else if (fmt == ORC_SCOO) {
    const orc_coo *a = (const orc_coo *)m;
    if (a->g.nnz == 0) return;
    for (long k = 0; k < a->g.nnz; k++) s[a->col[k] - a->g.f_col] += a->values[k];
  }
}

/* normalize_cols: CSR (MatrixCSR.cpp:1215-1232), ELL (MatrixELL.cpp:818-833),
 * SCOO (MatrixSCOO.cpp:726-738); division only where the sum is > 0 (and,
 * for ELL, the value is > 0).  ELL tests the value before it reads s: the
 * reference reads s[columns - f_col] first (:828), which for a padding slot
 * (column 0) of a block with f_col > 0 lies before s; that value is never
 * used, so the results are the same without the out-of-bounds read. */
void orc_normalize_cols(int fmt, void *m, const double *s) {
  if (fmt == ORC_CSR) {
    orc_csr *a = (orc_csr *)m;
    if (a->g.nnz == 0) return;
    for (int i = 0; i < a->g.ln_row; i++)
      for (int j = a->rowptr[i]; j < a->rowptr[i + 1]; j++) {
        double sv = s[a->colidx[j] - a->g.f_col];
        if (sv > 0) a->values[j] /= sv;
      }
  } else if (fmt == ORC_ELL) {
    orc_ell *a = (orc_ell *)m;
    if (a->g.nnz == 0) return;
    for (int i = 0; i < a->g.ln_row; i++)
      for (int j = 0; j < a->max_col; j++) {
        size_t k = (size_t)i * a->max_col + j;
        if (!(a->values[k] > 0)) continue;
        double sv = s[a->columns[k] - a->g.f_col];
        if (sv > 0) a->values[k] /= sv;
      }
  } else if (fmt == ORC_SCOO) {
    orc_coo *a = (orc_coo *)m;
    if (a->g.nnz == 0) return;
    for (long k = 0; k < a->g.nnz; k++) {
      double sv = s[a->col[k] - a->g.f_col];
      if (sv > 0) a->values[k] /= sv;
    }
  }
}

/* Matrix::make_stochastic (src/cpp/matrix/Matrix.cpp:33-38); s must hold
 * ln_col zeros on entry and returns the column sums. */
void orc_make_stochastic(int fmt, void *m, double *s) {
  orc_get_col_sums(fmt, m, s);
  orc_normalize_cols(fmt, m, s);
}

/* ------------------------------------------------------------------------ */
/* Row operations                                                            */
/* ------------------------------------------------------------------------ */

/* get_row_sums: CSR (MatrixCSR.cpp:1147-1163: s[i] for local rows, zeroed
 * first), ELL (MatrixELL.cpp:783-792: s[i + f_row], every slot), SCOO
 * (MatrixSCOO.cpp:701-705: s[row[k]] += values[k]), DENSE
 * (MatrixDENSE.cpp:292-301: s[i]); COO a no-op (MatrixCOO.cpp:382-384).
 * A block without arrays reads as all zeros. */
void orc_get_row_sums(int fmt, const void *m, double *s) {
  if (fmt == ORC_CSR) {
    const orc_csr *a = (const orc_csr *)m;
    for (int i = 0; i < a->g.ln_row; i++) s[i] = 0.0;
    if (!a->rowptr) return;
    for (int i = 0; i < a->g.ln_row; i++) {
      double sum = 0;
      for (int j = a->rowptr[i]; j < a->rowptr[i + 1]; j++) sum += a->values[j];
      s[i] = sum;
    }
  } else if (fmt == ORC_ELL) {
    const orc_ell *a = (const orc_ell *)m;
    for (int i = 0; i < a->g.ln_row; i++) {
      double sum = 0;
      if (a->values)
        for (int j = 0; j < a->max_col; j++) sum += a->values[(size_t)i * a->max_col + j];
      s[i + a->g.f_row] = sum;
    }
  } else if (fmt == ORC_SCOO) {
    const orc_coo *a = (const orc_coo *)m;
    for (long k = 0; k < a->g.nnz; k++) s[a->row[k]] += a->values[k];
  } else if (fmt == ORC_DENSE) {
    const orc_dense *a = (const orc_dense *)m;
    for (int i = 0; i < a->g.ln_row; i++) {
      double sum = 0;
      if (a->values)
        for (int j = 0; j < a->g.ln_col; j++) sum += a->values[(size_t)i * a->g.ln_col + j];
      s[i] = sum;
    }
  }
}

/* normalize_rows: CSR (MatrixCSR.cpp:1177-1190: divide, or 0 where the sum
 * is 0), ELL (MatrixELL.cpp:794-801: every slot by s[i + f_row]), SCOO
 * (MatrixSCOO.cpp:707-711), DENSE (MatrixDENSE.cpp:303-310); COO a no-op. */
void orc_normalize_rows(int fmt, void *m, const double *s) {
  if (fmt == ORC_CSR) {
    orc_csr *a = (orc_csr *)m;
    if (!a->rowptr) return;
    for (int i = 0; i < a->g.ln_row; i++)
      for (int j = a->rowptr[i]; j < a->rowptr[i + 1]; j++) {
        if (s[i] != 0) a->values[j] /= s[i];
        else a->values[j] = 0;
      }
  } else if (fmt == ORC_ELL) {
    orc_ell *a = (orc_ell *)m;
    if (!a->values) return;
    for (int i = 0; i < a->g.ln_row; i++)
      for (int j = 0; j < a->max_col; j++) a->values[(size_t)i * a->max_col + j] /= s[i + a->g.f_row];
  } else if (fmt == ORC_SCOO) {
    orc_coo *a = (orc_coo *)m;
    for (long k = 0; k < a->g.nnz; k++) a->values[k] /= s[a->row[k]];
  } else if (fmt == ORC_DENSE) {
    orc_dense *a = (orc_dense *)m;
    if (!a->values) return;
    for (int i = 0; i < a->g.ln_row; i++)
      for (int j = 0; j < a->g.ln_col; j++) a->values[(size_t)i * a->g.ln_col + j] /= s[i];
  }
}

/* get_row_max_abs: CSR only (MatrixCSR.cpp:1285-1315): the largest |z| over
 * the nonzero values of each row, 0 for an empty row; std::max keeps its
 * first argument when the second is NaN, as the comparison below.  Every
 * other format is a no-op (MatrixELL.cpp:835, MatrixSCOO.cpp:740,
 * MatrixCOO.cpp:398, MatrixDENSE.cpp:320). */
void orc_get_row_max_abs(int fmt, const void *m, double *max_abs) {
  if (fmt != ORC_CSR) return;
  const orc_csr *a = (const orc_csr *)m;
  for (int i = 0; i < a->g.ln_row; i++) max_abs[i] = 0.0;
  if (!a->rowptr) return;
  for (int i = 0; i < a->g.ln_row; i++) {
    double mx = 0.0;
    for (int j = a->rowptr[i]; j < a->rowptr[i + 1]; j++) {
      double z = a->values[j];
      if (z != 0) mx = (mx < fabs(z)) ? fabs(z) : mx;
    }
    max_abs[i] = mx;
  }
}

/* apply_exponential: CSR only (MatrixCSR.cpp:1318-1334): each nonzero z
 * becomes exp(z - max_abs[i]) for base <= 0, pow(base, z - max_abs[i])
 * otherwise; max_abs NULL means no shift.  Other formats: no-ops. */
void orc_apply_exponential(int fmt, void *m, const double *max_abs, int base) {
  if (fmt != ORC_CSR) return;
  orc_csr *a = (orc_csr *)m;
  if (!a->rowptr) return;
  for (int i = 0; i < a->g.ln_row; i++)
    for (int j = a->rowptr[i]; j < a->rowptr[i + 1]; j++) {
      double z = a->values[j];
      if (z == 0) continue;
      double e = max_abs ? z - max_abs[i] : z;
      a->values[j] = base <= 0 ? exp(e) : pow((double)base, e);
    }
}

/* ------------------------------------------------------------------------ */
/* PageRank                                                                  */
/* ------------------------------------------------------------------------ */

/* tbsla::mpi::Matrix::page_rank (src/mpi/matrix/Matrix.cpp:540-602) on a
 * single block (NR = NC = 1, every collective a copy): b starts at 1; each
 * iteration y = A b_t (Ax_local :90-103), b = y copied over the overlap of
 * row and column ranges (redistribute_vector :266-292), then
 * pagerank_normalization (:106-146): S = sum(b_t), b = beta*b + S(1-beta)/n,
 * b /= S, err = sum|b - b_t| / NC; stop when err < epsilon.  Finally
 * b /= sum(b) (pagerank_norma_end :149-171).  Returns the iteration count. */
int orc_page_rank(int fmt, const void *m, double beta, double epsilon, int max_iterations,
                  double *b, double *err_out) {
  return orc_page_rank_source(fmt, m, beta, epsilon, max_iterations, -1, b, err_out);
}

/* The same iteration with the teleport mass S(1 - beta) put on one vertex
 * (`source` >= 0) instead of spread uniformly: personalised PageRank, one
 * column of the multi-vector form.  No reference counterpart. */
int orc_page_rank_source(int fmt, const void *m, double beta, double epsilon, int max_iterations,
                         int source, double *b, double *err_out) {
  const orc_geom *g = geom_of(fmt, m);
  int ln_col = g->ln_col, ln_row = g->ln_row;
  double *b_t = (double *)zalloc((size_t)ln_col, sizeof(double));
  double *buf = (double *)zalloc((size_t)ln_row, sizeof(double));
  for (int i = 0; i < ln_col; i++) b[i] = 1;
  int recv_start = g->f_col, recv_end = g->f_col + ln_col;
  int blk_start = g->f_row, blk_end = g->f_row + ln_row;
  int s_start = 0, r_start = 0, length = 0;
  if (recv_start <= blk_end && recv_end >= blk_start) {
    if (blk_start <= recv_start) {
      r_start = 0;
      s_start = recv_start - blk_start;
    } else {
      r_start = blk_start - recv_start;
      s_start = 0;
    }
    length = imin(blk_end, recv_end) - (recv_start + r_start);
  }
  int it = 0, converged = 0;
  double error = 0;
  while (!converged && it < max_iterations) {
    for (int i = 0; i < ln_col; i++) {
      b_t[i] = b[i];
      b[i] = 0;
    }
    memset(buf, 0, sizeof(double) * (size_t)ln_row);
    orc_Ax(fmt, m, buf, b_t, 0);
    for (int z = 0; z < length; z++) b[r_start + z] = buf[s_start + z];
    double sum = 0;
    for (int k = 0; k < ln_col; k++) sum += b_t[k];
    if (source < 0) {
      double t_sum = sum * ((1 - beta) / g->n_col); /* t_sum *= (1-beta)/n_col, :123 */
      for (int k = 0; k < ln_col; k++) b[k] = beta * b[k] + t_sum;
    } else {
      double tel = sum * (1 - beta);
      for (int k = 0; k < ln_col; k++) b[k] = beta * b[k] + (k + g->f_col == source ? tel : 0.0);
    }
    error = 0;
    for (int k = 0; k < ln_col; k++) {
      b[k] /= sum;
      error += fabs(b[k] - b_t[k]);
    }
    error /= g->NC;
    if (error < epsilon) converged = 1;
    it++;
  }
  double tot = 0;
  for (int k = 0; k < ln_col; k++) tot += b[k];
  for (int k = 0; k < ln_col; k++) b[k] /= tot;
  if (err_out) *err_out = error;
  free(b_t);
  free(buf);
  return it;
}

/* tbsla::mpi::MatrixCOO::page_rank (src/mpi/matrix/MatrixCOO.cpp:108-175)
 * at one rank, on a block covering a square matrix: b = 1 (:115-118); per
 * iteration b_t = b, b = A b_t (:121-129), T = (T + sum(b_t))(1-beta)/n --
 * the OpenMP reduction adds into T's previous value (:130-134), which starts
 * at 0 here (uninitialised in the reference) -- then b = beta b + T, max over
 * b (:136-142), b /= max, error = sum |b - b_t| (:144-149); stop when
 * error < epsilon (:150-152).  Finally b /= sum(b) (:158-168). */
int orc_page_rank_max_norm(int fmt, const void *m, double beta, double epsilon,
                           int max_iterations, double *b, double *err_out) {
  const orc_geom *g = geom_of(fmt, m);
  int n = g->n_col;
  double *b_t = (double *)zalloc((size_t)n, sizeof(double));
  for (int i = 0; i < n; i++) b[i] = 1;
  int it = 0, converged = 0;
  double error = 0, T = 0;
  while (!converged && it < max_iterations) {
    for (int i = 0; i < n; i++) {
      b_t[i] = b[i];
      b[i] = 0;
    }
    orc_Ax(fmt, m, b, b_t, 0);
    double sum = 0;
    for (int i = 0; i < n; i++) sum += b_t[i];
    T = (T + sum) * ((1 - beta) / n);
    double mx = 0;
    for (int i = 0; i < n; i++) {
      b[i] = beta * b[i] + T;
      if (i == 0 || mx < b[i]) mx = b[i];
    }
    error = 0;
    for (int i = 0; i < n; i++) {
      b[i] = b[i] / mx;
      error += fabs(b[i] - b_t[i]);
    }
    if (error < epsilon) converged = 1;
    it++;
  }
  double tot = 0;
  for (int i = 0; i < n; i++) tot += b[i];
  for (int i = 0; i < n; i++) b[i] /= tot;
  if (err_out) *err_out = error;
  free(b_t);
  return it;
}

/* ------------------------------------------------------------------------ */
/* CG                                                                        */
/* ------------------------------------------------------------------------ */

static double dot(const double *a, const double *b, int n) {
  double s = 0;
  for (int k = 0; k < n; k++) s += a[k] * b[k];
  return s;
}

/* tbsla::mpi::Matrix::CG (src/mpi/matrix/CG.cpp:32-73) with the Ax call
 * corrected to AP = A * P (:61 passes them swapped).  R starts at the
 * right-hand side (the caller's x is assumed zero, as the app sets it,
 * tbsla_perf_cg_mpi.cpp:124-127); alpha = (R.R)/max(P.AP, 1e-10);
 * x += alpha P; R -= alpha AP; stop when |R| < tol (absolute, :34, :67) or
 * |R| < tol*|rhs| when `relative`; beta = (R.R)/max(Rold.Rold, 1e-10);
 * P = R + beta P.  Returns the reference's nb_iterations_done (iterations
 * that did not converge). */
int orc_cg(int fmt, const void *m, const double *rhs, double *x, int max_iterations,
           double tolerance, int relative, double *res_norm_out) {
  return orc_cg_clamp(fmt, m, rhs, x, max_iterations, tolerance, relative,
                      relative ? 1e-300 : 1.0e-10, res_norm_out);
}

/* The same iteration with the divisor clamp (NEARZERO, CG.cpp:35) given:
 * the reference's absolute 1e-10 stalls the iteration once r.r drops below
 * it, so the relative mode uses a guard against exact zero only. */
int orc_cg_clamp(int fmt, const void *m, const double *rhs, double *x, int max_iterations,
                 double tolerance, int relative, double near_zero, double *res_norm_out) {
  const double NEARZERO = near_zero;
  const orc_geom *g = geom_of(fmt, m);
  int n = g->n_col;
  double *R = (double *)zalloc((size_t)n, sizeof(double));
  double *P = (double *)zalloc((size_t)n, sizeof(double));
  double *AP = (double *)zalloc((size_t)n, sizeof(double));
  memcpy(R, rhs, sizeof(double) * (size_t)n);
  memcpy(P, R, sizeof(double) * (size_t)n);
  double stop = tolerance;
  if (relative) stop = tolerance * sqrt(dot(rhs, rhs, n));
  int it = 0;
  double rr = dot(R, R, n), rnorm = sqrt(rr);
  while (it < max_iterations) {
    double rr_old = rr;
    memset(AP, 0, sizeof(double) * (size_t)n);
    orc_Ax(fmt, m, AP, P, 0);
    double alpha = rr / fmax(dot(P, AP, n), NEARZERO);
    for (int k = 0; k < n; k++) x[k] = 1.0 * x[k] + alpha * P[k];
    for (int k = 0; k < n; k++) R[k] = 1.0 * R[k] + (-alpha) * AP[k];
    rr = dot(R, R, n);
    rnorm = sqrt(rr);
    if (rnorm < stop) break;
    double beta = rr / fmax(rr_old, NEARZERO);
    for (int k = 0; k < n; k++) P[k] = 1.0 * R[k] + beta * P[k];
    it++;
  }
  if (res_norm_out) *res_norm_out = rnorm;
  free(R);
  free(P);
  free(AP);
  return it;
}

/* ------------------------------------------------------------------------ */
/* SpMM                                                                      */
/* ------------------------------------------------------------------------ */

/* MatrixCSR::dense_multiply (MatrixCSR.cpp:1270-1283) in its block form
 * (src/mpi/matrix/MatrixCSR.cpp:11-30): C = A B with B (ln_col x k) and C
 * (ln_row x k) row-major, columns outside the block skipped; C is zeroed. */
void orc_csr_spmm(const orc_csr *m, const double *B, double *C, int k) {
  const orc_geom *g = &m->g;
  memset(C, 0, sizeof(double) * (size_t)g->ln_row * (size_t)k);
  for (int i = 0; i < g->ln_row; i++)
    for (int j = m->rowptr[i]; j < m->rowptr[i + 1]; j++) {
      int col = m->colidx[j] - g->f_col;
      if (col < 0 || col >= g->ln_col) continue; /* src/mpi/matrix/MatrixCSR.cpp:19 */
      double v = m->values[j];
      for (int kk = 0; kk < k; kk++) C[(size_t)i * k + kk] += v * B[(size_t)col * k + kk];
    }
}
