/*
 * tbsla_hip.h -- C ABI of libtbsla_hip, the MI355X (gfx950) implementation of
 * TBSLA's sparse hot path.
 *
 * One opaque handle holds one matrix block on the GPU in one of TBSLA's five
 * formats.  It replaces a tbsla::cpp::Matrix object
 * (src/cpp/include/tbsla/cpp/Matrix.hpp:11-57) and the stale opaque-handle C
 * interface C_CPP_Matrix<FMT>_* (src/cpp/interface/c/include/tbsla/cpp/
 * MatrixCSR.h:10-21), which does not compile.
 *
 * Conventions
 *   - Every function returns a status (TBSLA_HIP_SUCCESS = 0); on failure
 *     tbsla_hip_last_error() describes the cause (per thread).
 *   - Vector arguments named d_* are device pointers (hipMalloc memory, or
 *     tbsla_hip_malloc); h_* are host pointers.
 *   - `stream` is a hipStream_t passed as void* (NULL: the null stream).
 *     Work is enqueued on it; functions that return scalars to the host
 *     (iteration counts, norms) synchronise that stream.  The only streams
 *     the library creates are the one a handle's CG uses and the one a
 *     batched PageRank call uses when given the null stream (which cannot
 *     be captured into a hipGraph).
 *   - Thread-compatible: one handle per thread, or external synchronisation.
 *   - Geometry follows the reference exactly (n_row .. NC, nnz), including
 *     the count-pass nnz of CSR fill_cqmat (MatrixCSR.cpp:423-434) that can
 *     exceed the stored entries; `stored` reports what is held.
 *   - Device layouts: CSR rowptr/colidx/values; ELL column-major
 *     (columns[j * ln_row + i], the transpose of the reference's row-major
 *     layout, so that a wavefront's loads coalesce); COO/SCOO triplets;
 *     DENSE row-major.  tbsla_hip_export returns the reference's layouts.
 */
#ifndef TBSLA_HIP_H
#define TBSLA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tbsla_hip_matrix_s *tbsla_hip_matrix_t;

/* Formats, in the order of the reference perf app's --format switch
 * (src/cpp/apps/tbsla_perf.cpp:217-230). */
enum {
  TBSLA_HIP_COO = 0,
  TBSLA_HIP_SCOO = 1,
  TBSLA_HIP_CSR = 2,
  TBSLA_HIP_ELL = 3,
  TBSLA_HIP_DENSE = 4
};

enum {
  TBSLA_HIP_SUCCESS = 0,
  TBSLA_HIP_ERR_INVALID = 1,     /* bad argument or handle */
  TBSLA_HIP_ERR_HIP = 2,         /* a HIP runtime call failed */
  TBSLA_HIP_ERR_ALLOC = 3,       /* device allocation failed */
  TBSLA_HIP_ERR_UNSUPPORTED = 4, /* operation not defined for this format */
  TBSLA_HIP_ERR_RANGE = 5,       /* size beyond int32 indexing */
  TBSLA_HIP_ERR_BREAKDOWN = 6    /* CG: r.z <= 0 or not finite, or no progress (see stall_window) */
};

/* Geometry of a block: the protected fields of tbsla::cpp::Matrix
 * (Matrix.hpp:54-55) and their getters (:21-31), plus what the GPU holds. */
typedef struct {
  int n_row, n_col, f_row, f_col, ln_row, ln_col, pr, pc, NR, NC;
  long long nnz;    /* the reference's nnz field (get_nnz) */
  long long stored; /* entries held (CSR rowptr[ln_row]; COO/SCOO triplets) */
  int max_col;      /* ELL row width */
  int format;
} tbsla_hip_geometry;

const char *tbsla_hip_status_string(int status);
const char *tbsla_hip_last_error(void);
int tbsla_hip_version(void);

/* ---- memory helpers (for callers without their own allocator) ----------- */
int tbsla_hip_malloc(void **d_ptr, size_t bytes);
int tbsla_hip_free(void *d_ptr);
int tbsla_hip_memcpy_htod(void *d_dst, const void *h_src, size_t bytes, void *stream);
int tbsla_hip_memcpy_dtoh(void *h_dst, const void *d_src, size_t bytes, void *stream);
int tbsla_hip_memcpy_dtod(void *d_dst, const void *d_src, size_t bytes, void *stream);
int tbsla_hip_stream_synchronize(void *stream);
/* Device-to-device copy by a plain streaming kernel (16-byte loads and
 * stores), e.g. as a calibration dispatch for hardware counters. */
int tbsla_hip_copy_kernel(void *d_dst, const void *d_src, size_t bytes, void *stream);
/* One streaming pass over bytes (a multiple of 16; d_src 16-byte aligned)
 * with 16-byte loads, as a read-bandwidth ceiling; 8192 x 8 bytes are
 * written to d_sink. */
int tbsla_hip_read_kernel(const void *d_src, size_t bytes, void *d_sink, void *stream);
int tbsla_hip_device_count(int *count);

/* ---- handles ------------------------------------------------------------ */
/* C_CPP_Matrix<FMT>_create / _destroy (cpp/interface/c/include/tbsla/cpp/
 * MatrixCSR.h:10-11). */
int tbsla_hip_create(tbsla_hip_matrix_t *out, int format);
int tbsla_hip_destroy(tbsla_hip_matrix_t m);
int tbsla_hip_get_geometry(tbsla_hip_matrix_t m, tbsla_hip_geometry *out);

/* ---- generators --------------------------------------------------------- */
/* Matrix::fill_cdiag (Matrix.hpp:41); per format MatrixCSR.cpp:297-371,
 * MatrixELL.cpp:261-333, MatrixCOO.cpp:280-318, MatrixSCOO.cpp:228-293,
 * MatrixDENSE.cpp:129-185. */
int tbsla_hip_fill_cdiag(tbsla_hip_matrix_t m, int n_row, int n_col, int cdiag, int pr, int pc,
                         int NR, int NC, void *stream);
/* Matrix::fill_cqmat (Matrix.hpp:42); per format MatrixCSR.cpp:373-486,
 * MatrixELL.cpp:335-466, MatrixCOO.cpp:320-354, MatrixSCOO.cpp:295-403,
 * MatrixDENSE.cpp:187-270.  value_seed < 0 stores 1.0 for every entry, as
 * the reference; value_seed >= 0 stores V(value_seed, k), the project's
 * uniform [0,1) value of generator index k (no reference counterpart). */
int tbsla_hip_fill_cqmat(tbsla_hip_matrix_t m, int n_row, int n_col, int c, double q,
                         unsigned seed_mult, int pr, int pc, int NR, int NC,
                         long long value_seed, void *stream);
/* x_i = V(seed, i) for i < n into a device vector (the synthetic inputs). */
int tbsla_hip_fill_unit_vector(double *d_x, long long n, long long seed, void *stream);
/* d_v[i] = lo + (hi - lo) V(seed, i) (lo = 0, hi = 1 equals fill_unit_vector). */
int tbsla_hip_fill_uniform_vector(double *d_v, long long n, unsigned long long seed, double lo,
                                  double hi, void *stream);

/* ---- SPD problems (benchmark inputs; no counterpart in the reference) ----
 * The reference's CG driver (src/mpi/matrix/CG.cpp:32-73) is fed cdiag /
 * cqmat blocks; these generators give it SPD systems.  All fill a square
 * single-block CSR handle with columns sorted within each row. */
/* -div(k grad u) on an nx*ny*nz cell grid, 7-point, Dirichlet boundaries;
 * cell (x,y,z) is row x + nx (y + ny z).  Face coefficient 2 k_a k_b /
 * (k_a + k_b); a boundary face adds k_cell to the diagonal.  d_kappa (device,
 * nx*ny*nz, nullable): NULL means k = 1 (diagonal 6, off-diagonal -1). */
int tbsla_hip_fill_poisson3d(tbsla_hip_matrix_t m, int nx, int ny, int nz, const double *d_kappa,
                             void *stream);
/* k = exp(mu + sigma z) on the grid: z is standard normal noise (Box-Muller
 * on V(seed, .)) under a separable Gaussian filter of standard deviation
 * corr_len cells (radius ceil(3 corr_len), reflecting boundaries), scaled to
 * unit variance away from the boundary. */
int tbsla_hip_fill_lognormal_field(double *d_kappa, int nx, int ny, int nz, double mu,
                                   double sigma, double corr_len, unsigned long long seed,
                                   void *stream);
/* A = (B + B^T)/2 with diag(A) = sum_{j != i} |A_ij| + 1, where B is
 * fill_cqmat(n, n, c, q, seed_mult, 0, 0, 1, 1, value_seed) as CSR.  Each
 * A_ij = (sum of B_ij + sum of B_ji) * 0.5, each sum in generation order. */
int tbsla_hip_fill_sym_cqmat(tbsla_hip_matrix_t m, int n, int c, double q, unsigned seed_mult,
                             long long value_seed, void *stream);
/* dst = src^T for a CSR block (geometry mirrored; rows keep source-row
 * order).  MatrixCSR has no transpose in the reference. */
int tbsla_hip_transpose(tbsla_hip_matrix_t dst, tbsla_hip_matrix_t src, void *stream);
/* dst = diag(parts[0], ..., parts[k-1]) as one CSR block (local columns of
 * each part shifted by the columns before it): the layout of the batched
 * solvers. */
int tbsla_hip_block_diag(tbsla_hip_matrix_t dst, const tbsla_hip_matrix_t *parts, int k,
                         void *stream);

/* ---- import / export ---------------------------------------------------- */
/* The arrays of the reference layout, copied to the host (the fields a
 * Matrix::write would emit, MatrixCSR.cpp:236-261 etc.):
 *   CSR:   h_a = rowptr (ln_row+1 int), h_b = colidx (stored int), h_v = values
 *   ELL:   h_a = columns (ln_row*max_col int, row-major), h_v = values (row-major)
 *   COO/SCOO: h_a = row, h_b = col (stored int), h_v = values
 *   DENSE: h_v = values (ln_row*ln_col, row-major)
 * Unused pointers may be NULL. */
int tbsla_hip_export(tbsla_hip_matrix_t m, int *h_a, int *h_b, double *h_v, void *stream);
/* Builds a block from host arrays in the reference layout (the state a
 * Matrix::read restores, MatrixCSR.cpp:263-295; `geometry` supplies
 * n_row .. NC, nnz, stored and, for ELL, max_col).  Array roles as export. */
int tbsla_hip_import(tbsla_hip_matrix_t m, int format, const tbsla_hip_geometry *geometry,
                     const int *h_a, const int *h_b, const double *h_v, void *stream);
/* ---- binary images ------------------------------------------------------ */
/* Matrix::write (Matrix.hpp:17; MatrixCSR.cpp:236-261, MatrixELL.cpp:208-228,
 * MatrixCOO.cpp:144-169, MatrixSCOO.cpp:166-191, MatrixDENSE.cpp:88-105):
 * the block's binary image, byte for byte the reference's: header
 * n_row n_col ln_row ln_col f_row f_col (int32), nnz (int64), pr pc NR NC
 * (int32), ELL max_col (int32), then each array as a size_t count and its
 * elements in the reference layout (CSR values[nnz] rowptr[ln_row+1]
 * colidx[nnz]; COO/SCOO values row col, nnz each; ELL values columns,
 * ln_row*max_col each, row-major; DENSE values ln_row*ln_col).  CSR entries
 * past rowptr[ln_row] that the nnz field still counts (fill_cqmat's count
 * pass) are written as zeros (the reference leaves them uninitialised).
 * h_buf NULL: *written receives the image's size only. */
int tbsla_hip_write(tbsla_hip_matrix_t m, void *h_buf, size_t capacity, size_t *written,
                    void *stream);
/* Matrix::read (Matrix.hpp:18; MatrixCSR.cpp:263-295 and its twins): the
 * block from an image in the handle's format.  A truncated image or arrays
 * of inconsistent lengths give TBSLA_HIP_ERR_INVALID. */
int tbsla_hip_read(tbsla_hip_matrix_t m, const void *h_buf, size_t bytes, void *stream);
/* The file forms of the reference C interface's C_CPP_MatrixCSR_write(m,
 * file) and C_CPP_MatrixCSR_read(m, file, seek)
 * (cpp/interface/c/include/tbsla/cpp/MatrixCSR.h:18-19): the image written
 * to `path`, or read from `path` starting `seek` bytes in. */
int tbsla_hip_write_file(tbsla_hip_matrix_t m, const char *path, void *stream);
int tbsla_hip_read_file(tbsla_hip_matrix_t m, const char *path, long long seek, void *stream);

/* Conversion constructors <FMT>(const MatrixCOO&): MatrixCSR.cpp:49-96,
 * MatrixELL.cpp:25-79, MatrixSCOO.cpp:75-101, MatrixDENSE.cpp:15-37.  The
 * (row, col) sort is stable (the reference's std::sort leaves the order of
 * equal keys unspecified). */
int tbsla_hip_convert(tbsla_hip_matrix_t dst, int format, tbsla_hip_matrix_t coo, void *stream);

/* Device pointers of the block's arrays (owned by the handle; valid until
 * the next fill/import/convert or destroy), in the device layout: CSR
 * a = rowptr, b = colidx; ELL b = columns (column-major); COO/SCOO a = row,
 * b = col; v = values.  For zero-copy interoperation: read them freely;
 * writes through them bypass the copies the library derives from the
 * arrays (the split and diagonal CSR layouts, the mixed CG's FP32 values,
 * the COO row pointer): call tbsla_hip_values_changed after them. */
int tbsla_hip_device_arrays(tbsla_hip_matrix_t m, void **d_a, void **d_b, void **d_v);
/* After writes through tbsla_hip_device_arrays (values or indices): rebuilds
 * the split/diagonal execution layout from the arrays, drops the COO row
 * pointer (remade, and the row order checked, at the next SpMV) and marks
 * every FP32 copy of the old values stale.  Ordered after the earlier work
 * on `stream`; synchronises it.  No reference counterpart: the reference's
 * arrays are host members with no derived copies. */
int tbsla_hip_values_changed(tbsla_hip_matrix_t m, void *stream);

/* ---- SpMV and composites ------------------------------------------------ */
/* Matrix::Ax (Matrix.hpp:15): CSR and ELL overwrite r (MatrixCSR.cpp:209-233,
 * MatrixELL.cpp:181-205); COO, SCOO and DENSE add into r
 * (MatrixCOO.cpp:135-142, MatrixSCOO.cpp:154-164, MatrixDENSE.cpp:77-86).
 * d_v has ln_col entries (n_col for COO); d_r has ln_row (n_row for COO). */
int tbsla_hip_Ax(tbsla_hip_matrix_t m, double *d_r, const double *d_v, int vect_incr,
                 void *stream);
/* Matrix::Ax without device allocation: no counterpart in the reference;
 * for callers that cannot allocate (code under hipGraph capture, serving
 * loops, allocation-free harnesses).  Same result as tbsla_hip_Ax bit for
 * bit (unsorted COO triplets, summed with atomics in both, agree within
 * rounding).  The one path of tbsla_hip_Ax that allocates -- the segmented
 * accumulation of row-sorted COO/SCOO triplets, which keeps per-chunk
 * partials -- uses a fixed 16 MiB array per GPU instead: at most about
 * 358 M triplets (more gives TBSLA_HIP_ERR_RANGE).  Every call on one GPU
 * shares that array: calls that overlap on one GPU race. */
int tbsla_hip_Ax_noalloc(tbsla_hip_matrix_t m, double *d_r, const double *d_v, int vect_incr,
                         void *stream);
/* Matrix::spmv (Matrix.hpp:14): r = 0, then Ax. */
int tbsla_hip_spmv(tbsla_hip_matrix_t m, double *d_r, const double *d_v, int vect_incr,
                   void *stream);
/* Matrix::a_axpx_ (src/cpp/matrix/Matrix.cpp:5-14): r = A(Av + v); d_tmp
 * holds ln_row doubles of scratch. */
int tbsla_hip_a_axpx(tbsla_hip_matrix_t m, double *d_r, const double *d_v, double *d_tmp,
                     int vect_incr, void *stream);
/* Matrix::AAxpAx / AAxpAxpx (Matrix.cpp:16-31); d_buf is the caller's
 * buffer, used without zeroing as the reference does. */
int tbsla_hip_AAxpAx(tbsla_hip_matrix_t m, double *d_r, const double *d_v, double *d_buf,
                     int vect_incr, void *stream);
int tbsla_hip_AAxpAxpx(tbsla_hip_matrix_t m, double *d_r, const double *d_v, double *d_buf,
                       int vect_incr, void *stream);
/* Times `iters` back-to-back tbsla_hip_spmv calls between two events on
 * `stream`; *ms = total milliseconds.  For measurement only. */
int tbsla_hip_time_spmv(tbsla_hip_matrix_t m, double *d_r, const double *d_v, int iters,
                        float *ms, void *stream);
/* ---- column normalisation, PageRank, CG --------------------------------- */
/* Matrix::get_col_sums / normalize_cols / make_stochastic (Matrix.hpp:46-49,
 * src/cpp/matrix/Matrix.cpp:33-38): CSR MatrixCSR.cpp:1192-1232, ELL
 * MatrixELL.cpp:803-833 (positive values only), SCOO MatrixSCOO.cpp:713-738;
 * COO and DENSE are no-ops as in the reference.  d_s (ln_col doubles) is
 * accumulated into, as the reference does; the caller zeroes it.  Column
 * sums are formed in the reference's serial loop order (stable sort by
 * column), so they are bit-identical to a serial run. */
int tbsla_hip_get_col_sums(tbsla_hip_matrix_t m, double *d_s, void *stream);
int tbsla_hip_normalize_cols(tbsla_hip_matrix_t m, const double *d_s, void *stream);
int tbsla_hip_make_stochastic(tbsla_hip_matrix_t m, double *d_s, void *stream);

/* ---- Matrix Market import ----------------------------------------------- */
/* MatrixCOO::readMM (MatrixCOO.cpp:206-278; the converter app
 * tbsla_mm_to_binCOO.cpp:10): a "%%MatrixMarket matrix coordinate real
 * general|symmetric" file into the COO block m (one block covering the
 * matrix, as MatrixCOO::init).  Entries keep file order, 0-based; in a
 * symmetric file each off-diagonal entry (r, c) is followed by (c, r) (the
 * reference overwrites the first with the second, MatrixCOO.cpp:256-269;
 * here both are kept).  Other headers: TBSLA_HIP_ERR_UNSUPPORTED (the
 * reference throws MatrixFormatReadException); a missing file, fewer
 * entries than declared or an index outside the matrix:
 * TBSLA_HIP_ERR_INVALID.  Convert with tbsla_hip_convert for other formats. */
int tbsla_hip_read_mm(tbsla_hip_matrix_t m, const char *path, void *stream);
/* The same parse into host arrays, with no device work: *n_row, *n_col and
 * *nnz are set; h_rows/h_cols/h_values (nullable, `capacity` entries each)
 * receive the triplets when given. */
int tbsla_hip_mm_parse(const char *path, int *n_row, int *n_col, long long *nnz, int *h_rows,
                       int *h_cols, double *h_values, long long capacity);

/* ---- row operations ------------------------------------------------------ */
/* Matrix::get_row_sums / normalize_rows (Matrix.hpp:46-47).  Each format
 * keeps the reference's indexing, which differs between formats:
 *   CSR   (MatrixCSR.cpp:1147-1190): d_s[i], i < ln_row, is set to row i's
 *         sum; normalize divides by d_s[i], or sets 0 where d_s[i] == 0;
 *   ELL   (MatrixELL.cpp:783-801): d_s[f_row + i] is set to the sum over
 *         all max_col slots (padding included); normalize divides every
 *         slot by it;
 *   SCOO  (MatrixSCOO.cpp:701-711): d_s[row[k]] += values[k] (global row
 *         index, not zeroed first); normalize divides values[k] by
 *         d_s[row[k]];
 *   DENSE (MatrixDENSE.cpp:292-310): d_s[i] set; normalize divides by it;
 *   COO   (MatrixCOO.cpp:382-388): nothing.
 * So d_s holds ln_row doubles for CSR and DENSE and n_row (at least
 * f_row + ln_row) for ELL and SCOO.  Sums follow the reference's serial
 * loop order and are bit-identical to it. */
int tbsla_hip_get_row_sums(tbsla_hip_matrix_t m, double *d_s, void *stream);
int tbsla_hip_normalize_rows(tbsla_hip_matrix_t m, const double *d_s, void *stream);
/* Matrix::get_row_max_abs / apply_exponential (Matrix.hpp:50-51); only CSR
 * acts (MatrixCSR.cpp:1285-1334), the other formats are no-ops as in the
 * reference.  get_row_max_abs: d_max_abs[i] (ln_row doubles) = the largest
 * |z| over row i's nonzero values, 0 for an empty row.  apply_exponential:
 * every nonzero z of row i becomes exp(z - d_max_abs[i]) when base <= 0,
 * pow(base, z - d_max_abs[i]) otherwise; with d_max_abs null, exp(z) or
 * pow(base, z).  Device exp/pow are within a few ulp of the C library's. */
int tbsla_hip_get_row_max_abs(tbsla_hip_matrix_t m, double *d_max_abs, void *stream);
int tbsla_hip_apply_exponential(tbsla_hip_matrix_t m, const double *d_max_abs, int base,
                                void *stream);
/* The attention step's row softmax (tbsla_attention_mpi.cpp:183-191) on one
 * CSR block, fused: the same result, bit for bit, as get_row_max_abs,
 * apply_exponential(max_abs, base), get_row_sums and normalize_rows called
 * in turn (the row reductions over the process row are identities on one
 * block), in one pass over the values.  d_max_abs and d_s (ln_row doubles
 * each, nullable) receive the row maxima and the row sums.  CSR only. */
int tbsla_hip_softmax_rows(tbsla_hip_matrix_t m, int base, double *d_max_abs, double *d_s,
                           void *stream);

/* tbsla::mpi::Matrix::page_rank (src/mpi/matrix/Matrix.cpp:540-602) on this
 * block: b starts at 1; per iteration y = A b_t, b = (beta y + S(1-beta)/n)/S
 * with S = sum(b_t), error = sum|b - b_t| / NC; stops when error < epsilon
 * or after max_iterations; finally b /= sum(b).  d_b receives ln_col
 * doubles; *nb_iterations and *final_error as the reference reports them. */
int tbsla_hip_page_rank(tbsla_hip_matrix_t m, double beta, double epsilon, int max_iterations,
                        double *d_b, int *nb_iterations, double *final_error, void *stream);
/* tbsla::mpi::MatrixCOO::page_rank (src/mpi/matrix/MatrixCOO.cpp:108-175),
 * the COO override of the above, on a block covering a square matrix (any
 * COO block; other formats with NR = NC = 1): b starts at 1; per iteration
 * y = A b_t, T = (T + sum(b_t))(1-beta)/n (T carries over between
 * iterations as in the reference's reduction, starting at 0), b = beta y + T,
 * b /= max(b), error = sum|b - b_t|; stops when error < epsilon or after
 * max_iterations; finally b /= sum(b).  d_b receives n doubles. */
int tbsla_hip_page_rank_max_norm(tbsla_hip_matrix_t m, double beta, double epsilon,
                                 int max_iterations, double *d_b, int *nb_iterations,
                                 double *final_error, void *stream);
/* Many PageRank instances at once (no counterpart in the reference): the
 * instances are the square diagonal blocks of the CSR block m at (0, 0),
 * instance k owning rows [h_row_offsets[k], h_row_offsets[k+1])
 * (tbsla_hip_block_diag builds such an m from column-stochastic parts).
 * Each runs tbsla_hip_page_rank's iteration on its own block, with its own
 * sum, error and stop (error < epsilon, or max_iterations), and its own
 * final normalisation; one SpMV per iteration serves every instance.  The
 * iterations go in batches of 16 captured once into a hipGraph (on a
 * stream of the library's own when `stream` is NULL).  Instance k follows a
 * tbsla_hip_page_rank of its block alone bit for bit when both SpMVs use
 * the same row kernel (tbsla_hip_set_csr_kernel).  d_b receives every
 * instance's vector (ln_row doubles); h_iterations / h_errors / h_time_us
 * (nullable, ninst entries) the per-instance counts, final errors and solve
 * times (device wall clock from the first iteration to the test that ended
 * the instance, microseconds; -1 if it did not end). */
int tbsla_hip_page_rank_batched(tbsla_hip_matrix_t m, int ninst, const int *h_row_offsets,
                                double beta, double epsilon, int max_iterations, double *d_b,
                                int *h_iterations, double *h_errors, double *h_time_us,
                                void *stream);

enum { TBSLA_HIP_CG_FP64 = 0, TBSLA_HIP_CG_MIXED = 1 };
enum { TBSLA_HIP_PRECOND_NONE = 0, TBSLA_HIP_PRECOND_JACOBI = 1, TBSLA_HIP_PRECOND_FACTORED = 2 };

typedef struct {
  int max_iterations;    /* reference default 10000 */
  double tolerance;      /* stop when |r| < tolerance (x |rhs| if relative); CG.cpp:34 */
  int relative;          /* 0: absolute test as the reference; 1: relative */
  int precision;         /* TBSLA_HIP_CG_FP64, or _MIXED: FP32 matrix and preconditioner
                            values with FP32 accumulation; FP64 vectors and dot products */
  int preconditioner;    /* TBSLA_HIP_PRECOND_NONE, _JACOBI (D^-1) or _FACTORED (M = G G^T) */
  int replace_every;     /* > 0: replace r by the true FP64 residual every k iterations */
  int use_initial_guess; /* 0: x starts at 0 (as the reference app); 1: from d_x */
  double near_zero;      /* clamp of alpha's and beta's divisors (CG.cpp:35): 0 selects the
                            reference's 1e-10 in its own setting (absolute test, FP64, no
                            preconditioner) and none otherwise; > 0 that value; < 0 none */
  tbsla_hip_matrix_t factor;   /* _FACTORED: G, CSR, same shape as the matrix */
  tbsla_hip_matrix_t factor_t; /* _FACTORED: G^T as CSR, or NULL (built per call) */
  int graph_batch;       /* iterations per captured hipGraph: 0 picks (16, or a multiple
                            of replace_every); < 0 launches every kernel directly */
  int lanes_per_row;     /* lanes per matrix row in the iteration's SpMV (power of two
                            <= 64); 0 picks from the mean row length */
  int stall_window;      /* > 0: breakdown when |r| >= |r0| after that many iterations */
  int factor_form;       /* _FACTORED: how M = G G^T is applied.  TBSLA_HIP_FACTOR_AUTO: from G's
                            diagonals when its entries lie on at most 8 diagonals of offset
                            <= 0 (a stencil's lower factor: no column indices read), else
                            from the CSR rows of G and G^T; _ROWS: always the CSR rows;
                            _DIAGONALS: the diagonals or TBSLA_HIP_ERR_UNSUPPORTED.  Both
                            forms sum G^T's and G's rows in ascending column, so they give
                            bitwise the same iterates */
  int max_blocks_per_system; /* workgroups of the iteration's kernels per system, at most
                            (1..1024); 0 picks one per 512 rows, 256 to 1024, for one
                            system and 256 when tbsla_hip_cg_batched solves more
                            than one.  Every workgroup reads all of its system's
                            dot-product partials, so a batch, whose systems fill the GPU
                            together, takes fewer; a system solved alone and in a batch
                            gives bitwise the same iterates when both use the same value
                            (and the same lanes_per_row) */
  int spmv_passes;       /* K1 (p = z + beta p, then A p) on A's CSR rows: 1 forms p at every
                            gathered entry from z and p (one kernel); 2 writes p first and
                            then gathers it alone (two kernels, 24 more bytes per row, half
                            the gathers: for matrices whose gathers miss the caches); 0 picks
                            2 from 2^24 stored entries.  Both give bitwise the same iterates;
                            A in the diagonal layout always takes its own one-kernel form */
} tbsla_hip_cg_options;
enum {
  TBSLA_HIP_FACTOR_AUTO = 0,
  TBSLA_HIP_FACTOR_ROWS = 1,
  TBSLA_HIP_FACTOR_DIAGONALS = 2
};

/* tbsla::mpi::Matrix::CG (src/mpi/matrix/CG.cpp:32-73) with its Ax call
 * corrected (:61 passes output and input swapped, so the reference never
 * moves): alpha = r.z / max(p.Ap, nz), x += alpha p, r -= alpha Ap, z = M r,
 * stop on |r|, beta = r.z / max(r_old.z_old, nz), p = z + beta p (z = r
 * without a preconditioner, as the reference).  opt NULL: the reference's
 * settings.  *nb_iterations counts the iterations that did not converge
 * (:71); *res_norm is |r| at exit (the true residual's in the mixed form).
 * Returns TBSLA_HIP_ERR_BREAKDOWN when the iteration broke down. */
int tbsla_hip_cg(tbsla_hip_matrix_t m, const double *d_rhs, double *d_x,
                 const tbsla_hip_cg_options *opt, int *nb_iterations, double *res_norm,
                 void *stream);

/* Outcome of one system of tbsla_hip_cg_batched. */
enum {
  TBSLA_HIP_CG_CONVERGED = 0,
  TBSLA_HIP_CG_MAX_ITERATIONS = 1,
  TBSLA_HIP_CG_BREAKDOWN = 2,
  TBSLA_HIP_CG_INACTIVE = 3
};
typedef struct {
  int iterations;  /* as tbsla_hip_cg's *nb_iterations */
  int status;      /* TBSLA_HIP_CG_* */
  double res_norm; /* |r| at exit */
  double time_us;  /* device wall clock from the first iteration to the test that ended
                      this system's iteration (the iteration it converged at; a true-residual
                      check after it, which may resume the system, is not counted); -1 if
                      it did not finish */
} tbsla_hip_cg_result;

/* Many independent CG solves at once (no counterpart in the reference): the
 * systems are the diagonal blocks of m, system k owning rows [h_row_offsets[k],
 * h_row_offsets[k+1]) (tbsla_hip_block_diag builds such an m).  One kernel
 * sequence per iteration serves all systems; each has its own scalars,
 * threshold and convergence, and stops on its own.  h_active (nullable, nsys
 * ints): 0 skips a system.  d_rhs / d_x cover all rows. */
int tbsla_hip_cg_batched(tbsla_hip_matrix_t m, int nsys, const int *h_row_offsets,
                         const int *h_active, const double *d_rhs, double *d_x,
                         const tbsla_hip_cg_options *opt, tbsla_hip_cg_result *h_results,
                         void *stream);

/* ---- learned factored preconditioner (no counterpart in the reference) ---
 * A 3-layer message-passing network (width 64; BF16 MFMA products with FP32
 * accumulation) maps an SPD CSR block A (at (0,0), symmetric pattern) to a
 * lower-triangular G on tril(A)'s pattern with a positive diagonal, for CG
 * with M = G G^T (TBSLA_HIP_PRECOND_FACTORED).  With s_i = A_ii^-1/2 and
 * A^ = S A S the network predicts G^ and returns G = S G^; its output heads
 * start at zero, which is G^ = I (Jacobi).  Training minimises
 * |Z - G^^T A^ G^ Z|_F^2 / (n k) over k Rademacher probes. */
typedef struct tbsla_hip_precond_net_s *tbsla_hip_precond_net_t;
/* number of FP32 parameters */
int tbsla_hip_precond_net_param_count(void);
/* parameters drawn from seed; Adam state zero */
int tbsla_hip_precond_net_create(tbsla_hip_precond_net_t *out, unsigned long long seed);
int tbsla_hip_precond_net_destroy(tbsla_hip_precond_net_t net);
int tbsla_hip_precond_net_get_params(tbsla_hip_precond_net_t net, float *h_theta);
/* also resets the Adam state */
int tbsla_hip_precond_net_set_params(tbsla_hip_precond_net_t net, const float *h_theta);
/* G (and G^T if GT is not NULL) for A: CSR, FP64 values, columns sorted */
int tbsla_hip_precond_net_apply(tbsla_hip_precond_net_t net, tbsla_hip_matrix_t A,
                                tbsla_hip_matrix_t G, tbsla_hip_matrix_t GT, void *stream);
/* loss and its gradient (h_grad: param_count floats, nullable) for probes
 * 1, 2, 4, 8 or 16 drawn from probe_seed */
int tbsla_hip_precond_net_loss_grad(tbsla_hip_precond_net_t net, tbsla_hip_matrix_t A,
                                    int probes, unsigned long long probe_seed, double *loss,
                                    float *h_grad, void *stream);
/* one Adam step (beta 0.9 / 0.999, eps 1e-8) on the gradient for A, after
 * clipping its global norm to clip (0: none); *loss before the step */
int tbsla_hip_precond_net_train_step(tbsla_hip_precond_net_t net, tbsla_hip_matrix_t A,
                                     int probes, unsigned long long probe_seed, float lr,
                                     float clip, double *loss, void *stream);

/* ---- SpMM and personalised PageRank ------------------------------------- */
/* MatrixCSR::dense_multiply (MatrixCSR.cpp:1270-1283; block form
 * src/mpi/matrix/MatrixCSR.cpp:11-30): C = A B, B ln_col x k and C
 * ln_row x k row-major device arrays, any k >= 1.  CSR only.  Products are
 * accumulated in the reference's order without fused multiply-add, so C is
 * bit-identical to a serial run of the reference loop. */
int tbsla_hip_spmm(tbsla_hip_matrix_t m, const double *d_B, double *d_C, int k, void *stream);
/* Personalised PageRank for 64 one-hot teleport vectors solved together as
 * one SpMM per iteration (no reference counterpart): column j follows the
 * page_rank iteration of tbsla_hip_page_rank with the teleport mass
 * S_j(1-beta) placed on vertex h_sources[j] (global index).  d_X receives
 * ln_col x 64 doubles, row-major, each column normalised to sum 1.  Stops
 * when every column's L1 change is below epsilon. */
int tbsla_hip_personalized_page_rank(tbsla_hip_matrix_t m, double beta, double epsilon,
                                     int max_iterations, const int *h_sources, double *d_X,
                                     int *nb_iterations, double *max_error, void *stream);

/* CSR SpMV schedule (no reference counterpart): `lanes` cooperating on one
 * row (1..64, power of two) and `rows` each lane group keeps in flight
 * (1, 2 or 4); 0 restores the choice from the mean row length. */
int tbsla_hip_set_csr_schedule(tbsla_hip_matrix_t m, int lanes, int rows);
int tbsla_hip_get_csr_schedule(tbsla_hip_matrix_t m, int *lanes, int *rows);
/* The full CSR row-kernel choice (no reference counterpart): the schedule
 * above plus `vec` entries per lane load (1, 2 or 4; loads of vec column
 * indices and vec values, aligned) and `nontemporal` (1: entry loads do not
 * keep their lines in L2).  Compiled kernels: vec 1 with lanes 1..64 and
 * rows 1, 2, 4; vec 2 with lanes 4..32 and vec 4 with lanes 4..16, both with
 * rows 1, 2.  tbsla_hip_set_csr_schedule resets vec to 1 and nontemporal to
 * 1. */
int tbsla_hip_set_csr_kernel(tbsla_hip_matrix_t m, int lanes, int rows, int vec,
                             int nontemporal);
int tbsla_hip_get_csr_kernel(tbsla_hip_matrix_t m, int *lanes, int *rows, int *vec,
                             int *nontemporal);
/* CSR / ELL / COO / SCOO execution layout (no counterpart in the
 * reference; results within rounding of the row-major sum; ELL padding
 * slots count as entries).  COO and SCOO take it over the rows from their
 * first to their last triplet's row through a row pointer derived from the
 * row array, and only when the triplets are sorted by row (else
 * TBSLA_HIP_ERR_UNSUPPORTED); their diagonals are measured from the global
 * row (COO) or the row less f_row (SCOO), as CSR's, and the layout serves
 * their SpMV (Ax accumulates through its own path).  DENSE has none.
 * TBSLA_HIP_CSR_ROWS: the arrays as they are.
 * TBSLA_HIP_CSR_SPLIT: entries with |col - row| <= window stay in row order;
 * the others ("remote") are grouped by column slice (slices a multiple of 8,
 * at most 64) and multiplied by one XCD per slice, whose L2 then holds that
 * slice of x; their products are summed per row in slice order and added to
 * the row kernel's sum.  TBSLA_HIP_CSR_DIAG: the diagonals c - r within
 * +-window (at most 1048576) holding nonzero entries in at least a quarter
 * of the rows (up to 64, the fullest) are stored densely, tiled by 64 rows
 * (8 bytes a slot, no column indices); every other entry -- off those
 * diagonals, explicit zeros, a second entry on a row's slot -- is remote.
 * The diagonal layout multiplies its remote entries through LDS: column
 * slices of 16384 (x's piece of 128 KB held by one workgroup), products
 * summed per row in CSR order inside the diagonal kernel, whenever every
 * block of 64..4096 rows holds at most 14336 remote entries (slices is then
 * unused); otherwise through L2 slices as the split layout.
 * Inf and NaN in x reach the same rows as in the row-major sum.  Both copy
 * the entries once, are rebuilt by normalize_cols and values_changed and
 * dropped when the arrays change; a block without columns cannot take the
 * diagonal layout. */
enum { TBSLA_HIP_CSR_ROWS = 0, TBSLA_HIP_CSR_SPLIT = 1, TBSLA_HIP_CSR_DIAG = 2 };
int tbsla_hip_set_csr_layout(tbsla_hip_matrix_t m, int layout, int window, int slices,
                             void *stream);
int tbsla_hip_get_csr_layout(tbsla_hip_matrix_t m, int *layout, int *window, int *slices,
                             long long *remote_entries);
/* The remote form of the current layout (no reference counterpart):
 * *block_rows = the rows per block of the diagonal layout's LDS-slice form
 * (kernels k_slice_products + k_dia_rows_remote), or 0 when the remote
 * entries go through L2 slices (k_remote_products + k_remote_combine), the
 * layout has no remote entries, or there is no derived layout. */
int tbsla_hip_get_csr_remote_form(tbsla_hip_matrix_t m, int *block_rows);
/* Autotuner (no reference counterpart): times every compiled CSR row
 * kernel (tbsla_hip_set_csr_kernel: lanes x rows x vec x nontemporal) on the
 * row layout and on the split layout (window 64, 8 and 16 slices), and the
 * diagonal layout (window 65536; 8 and 16 slices unless its LDS form is
 * taken), on this block with d_v /
 * d_r (`iters` launches each; the three fastest row kernels are timed again
 * with 3 x `iters` launches before the choice) and keeps the fastest; a
 * split or diagonal layout is kept only when it is at least 3 % faster than
 * the row layout.
 * ELL compares its column-major kernel, and row-sorted COO / SCOO their row
 * kernel, with the split and diagonal layouts by the same rule; DENSE has a
 * single schedule.  *best_us
 * receives its time per SpMV. */
int tbsla_hip_autotune(tbsla_hip_matrix_t m, double *d_r, const double *d_v, int iters,
                       float *best_us, void *stream);

#ifdef __cplusplus
}
#endif
#endif
