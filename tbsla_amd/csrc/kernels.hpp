// kernels.hpp -- internal interfaces between the .hip translation units.
#pragma once

#include "common.hpp"

namespace tbsla_hip {

// scan.hip: prefix sum of n values (in may equal out); *d_total (device,
// nullable) receives the sum.  workspace: scan_workspace_bytes(n) bytes.
template <typename T>
int scan(const T *in, T *out, long long n, bool inclusive, long long *d_total, void *workspace,
         hipStream_t s);
size_t scan_workspace_bytes(long long n);

// spmv.hip
// d_done (device, nullable): CSR and ELL launches return at once when it
// holds a nonzero value (converged solver iterations queued in a batch).
int spmv_launch(tbsla_hip_matrix_s *m, double *d_r, const double *d_v, int vect_incr,
                bool zero_first, hipStream_t s, const int *d_done = nullptr,
                const int *seg_offs = nullptr, int nseg = 0, const int *seg_done = nullptr);
int pick_csr_lanes(const tbsla_hip_matrix_s *m);
int pick_csr_rows(const tbsla_hip_matrix_s *m, int lanes);
int pick_csr_vec(const tbsla_hip_matrix_s *m);
int pick_csr_nt(const tbsla_hip_matrix_s *m);

// vec.hip
int launch_fill(double *d, long long n, double v, hipStream_t s);
int launch_axpy_inplace(double *y, const double *x, long long n, hipStream_t s);  // y += x

// sort.hip: stable LSD radix sort of 64-bit keys carrying a 32-bit payload
// (the original position); key_bits significant bits.  Buffers of n
// elements; the result lands in keys/payload.
int radix_sort_pairs(unsigned long long *keys, int *payload, long long n, int key_bits,
                     hipStream_t s);

// problems.hip: row pointers from per-row counts (exclusive scan; *total,
// host, receives the sum; synchronises the stream)
int counts_to_rowptr(const int *cnt, int *rowptr, long long n, long long *total, hipStream_t s);

// problems.hip: CSR transpose into new arrays (rows = src's local columns)
int csr_transpose_arrays(const tbsla_hip_matrix_s *src, int **rowptr_t, int **colidx_t,
                         double **values_t, hipStream_t s);

// split.hip: the split CSR layout (near-diagonal + column-sliced remote)
int split_launch(tbsla_hip_matrix_s *m, double *y, const double *x, const int *done,
                 hipStream_t s);
int split_build(tbsla_hip_matrix_s *m, int window, int slices, hipStream_t s,
                int layout = TBSLA_HIP_CSR_SPLIT);
int split_params(const tbsla_hip_matrix_s *m, int *window, int *slices, int *layout);
// rows per block of the diagonal layout's LDS-slice remote form, or 0 (no
// layout, no remote entries, or the L2-slice form, whose slice count matters)
int split_lds_block_rows(const tbsla_hip_matrix_s *m);
// PageRank's SpMV and update in one pass of the diagonal layout's LDS-slice
// form (CSR / ELL blocks with that form only; split_pr_blocks gives its
// workgroup count, the number of partials, or 0): b = (beta A b_t + S (1 -
// beta) / n_col) / S over the block's rows with b_t = x, as k_pr_update.
struct PrFuse {
  double *b;
  const double *S;
  double beta, n_col;
  double *p_err, *p_sum;
};
int split_pr_blocks(const tbsla_hip_matrix_s *m);
int split_launch_pr(tbsla_hip_matrix_s *m, const double *bt, const PrFuse &pf, const int *done,
                    hipStream_t s);
// The row pointer of a COO / SCOO block's n row-sorted triplets over their
// row range (built once and kept with the block); *ok false when the
// triplets are not sorted by row
int coo_rowptr(tbsla_hip_matrix_s *m, long long n, bool *ok, hipStream_t s);
// The diagonal layout's arrays when it holds every entry (no remote part):
// offsets, tiled values (dia_at), count, values' length, and the build's
// id (unique in the process: a copy made from one build is stale for any
// other, even one whose arrays reuse the same addresses); else 0.
int split_diag_only(const tbsla_hip_matrix_s *m, const int **doff, const double **dv, int *ndiag,
                    long long *dv_len, unsigned long long *build_id);

// solvers.hip: after the values of m changed in place: rebuild the split
// layout (it holds a copy of the values), mark copies of the old values
// stale (values_version), synchronise
int values_changed(tbsla_hip_matrix_s *m, hipStream_t s);

}  // namespace tbsla_hip
