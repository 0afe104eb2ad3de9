/* qk_launch.h -- the thin internal C ABI between the C++ host layer and the
 * HIP kernels (qk_kernels.hip). Plain pointers, sizes and streams only. */
#ifndef QK_LAUNCH_H
#define QK_LAUNCH_H

#include <stddef.h>
#include <stdint.h>

#include "qk_device.hpp"

#ifdef __cplusplus
extern "C" {
#endif

typedef void* qkl_stream;

/* Number of partial sums (per reduced quantity) the reducing kernels write:
 * qkl_apply_partials for qkl_lanczos_apply on these tables (wide: 64-bit
 * states), qkl_blas_partials for the BLAS-1 and update kernels. */
int qkl_apply_partials(const qk::QkTables* T, int wide);
int qkl_blas_partials(void);

/* LDS plan of the single-vector apply (host only, no device call): the
 * widest workgroup of at most pref_waves waves (16, 8, 4, 2, 1) whose dynamic
 * LDS (tables bytes of lo-rank table, branch tables, per-wave tables) plus
 * the kernel's static LDS fits the 160 KB of a gfx950 workgroup; -1 when
 * not even one wave fits. */
int qkl_apply_lds_plan(size_t tables, int nB, int nA, int cplx, int pref_waves, int* waves, size_t* dyn_lds);

/* y = H x (pull form). cplx: interleaved complex x/y. wide: 64-bit states.
 * The Lanczos form is qkl_lanczos_apply below. */
int qkl_apply(const qk::QkTables* T, int cplx, int wide, const double* x, double* y, qkl_stream s);

/* Block apply over nc <= 64 real columns: X[row * ld + c0 + c]. Lanczos
 * epilogue per column when lz_k >= 0 (scalars per column, stride ld_sc):
 *   w = HX / nrm[k][c] - nrm[k][c] * P / nrm[k-1][c];
 *   partials[c * npart + b] = sum (X / nrm) * w                             */
int qkl_apply_block(const qk::QkTables* T, int wide, const double* X, double* Y, int ld, int c0, int nc,
                    int lz_k, const double* P, const double* nrm, int ld_sc, double* partials, int npart,
                    qkl_stream s);
int qkl_block_partials(const qk::QkTables* T);

/* The real single-vector Lanczos step as two launches, each totalling the
 * partial sums the other wrote instead of a separate reduction launch:
 *   qkl_lanczos_apply:   nrm[k] = sqrt(total of nrm_part) (the previous
 *                        update's, or the start vector's, sums of squares;
 *                        workgroup 0 stores it), then y = H x with the
 *                        Lanczos epilogue
 *                          w = (H x) / nrm[k] - (k > 0 ? nrm[k] prev / nrm[k-1] : 0)
 *                          partials[b] = sum_rows (x / nrm[k]) w  (of alpha);
 *   qkl_lanczos_update1: alpha[k] = total of a_part (workgroup 0 stores it),
 *                        then qkl_lanczos_update for one vector; with knext
 *                        set, workgroup 0 also stores k - koff + 3 there.
 * Totals are fixed-order (qkl_total computes the same bits). k is taken
 * from *kdev + koff when kdev is set (launches captured once in a graph and
 * replayed), else from the argument. n_part is the room in partials; it
 * must equal qkl_apply_partials(T, wide) or the launch is refused. When
 * nrm_host / alpha_host is set (coherent pinned host memory), workgroup 0
 * also stores nrm[k] / alpha[k] there, so the host reads them after an
 * event without a copy. */
int qkl_lanczos_apply(const qk::QkTables* T, int wide, const double* x, double* y, int k, const int* kdev, int koff,
                      const double* prev, double* nrm, const double* nrm_part, int n_nrm_part, double* partials,
                      int n_part, double* nrm_host, qkl_stream s);
int qkl_lanczos_update1(size_t nrows, double* w, const double* q, double* alpha, const double* nrm, int k,
                        const int* kdev, int koff, const double* a_part, int n_a_part, double* partials, int* knext,
                        double* alpha_host, qkl_stream s);
/* out[idx] = total of partials[0..npart) (sqrt when root), in the order of
 * the two launches above */
int qkl_total(const double* partials, int npart, int root, double* out, int idx, qkl_stream s);

/* q_{k+1} = w - alpha[k] * q_k / nrm[k]; partials of |q_{k+1}|^2 (ncol columns
 * interleaved with stride ld; ncol = 1, ld = 1 for a single vector; cplx for
 * interleaved complex single vector). */
int qkl_lanczos_update(int cplx, size_t nrows, int ncol, int ld, double* w_inout, const double* q,
                       const double* alpha, const double* nrm, int k, int ld_sc, double* partials, qkl_stream s);

/* Fixed-order sum of partials. mode 0: out[idx] = sum (real part);
 * mode 1: out[idx] = sqrt(sum); mode 2: out[idx] = sum, out[idx+1] = sum of
 * imaginary partials (cplx). For ncol columns (grouped by `group` columns
 * which share one result): partials[c * npart + b]; results to
 * out[idx * ld_sc + c]. */
int qkl_reduce(const double* partials, int npart, int ncol, int group, int mode, int cplx,
               double* out, int idx, int ld_sc, qkl_stream s);

/* Multi-column apply: nv in {1, 2, 4, 8} real columns stored row-major
 * (X[row * nv + c]); column c uses H with its diagonal scaled by dscale[c]
 * (device array of nv). Real spin tables without parity masks, 32-bit
 * states (wide = 0), else -1. Plain form (Y = H_c X) for k < 0. Lanczos
 * form for k >= 0, per column: nrm[k * nv + c] = sqrt(total of
 * nrm_part[c * n_nrm_part + i]) (the previous update's partial sums, fixed
 * order; workgroup 0 stores it), then W = H_c X / nrm[k] - nrm[k] P /
 * nrm[k - 1] and partials[c * npart + b] of <X / nrm[k], W> with npart =
 * qkl_cols_partials. A column's bits do not depend on nv or on the other
 * columns.
 * qkl_update_cols: the Lanczos update of nv row-major columns,
 * w -= (alpha[k][c] / nrm[k][c]) q, with alpha[k][c] first totalled from
 * a_part (the apply's partials, same order as above; workgroup 0 stores
 * it) when a_part is given, else read; partial sums of |w_c|^2 at
 * partials[c * npart + b], npart = qkl_cols_update_partials(nrows).
 * qkl_total_cols: out[idx * nv + c] = sqrt(total of partials[c * npart +
 * i]) in the same order (the value the next apply would form).
 * qkl_axpy_cols: Y[r][c] += coef[c] X[r][c] for every column whose
 * coefficient is not zero. qkl_scale_rows_cols: Y[r][c] = d[r] X[r][c]. */
int qkl_cols_partials(const qk::QkTables* T);
int qkl_cols_update_partials(size_t nrows);
int qkl_apply_cols(const qk::QkTables* T, int wide, int nv, const double* X, double* Y, const double* dscale, int k,
                   const double* P, double* nrm, double* partials, int npart, const double* nrm_part, int n_nrm_part,
                   qkl_stream s);
int qkl_update_cols(int nv, size_t nrows, double* w, const double* q, double* alpha, const double* nrm, int k,
                    const double* a_part, int n_a_part, double* partials, qkl_stream s);
int qkl_total_cols(int nv, const double* partials, int npart, double* out, int idx, qkl_stream s);
int qkl_axpy_cols(int nv, size_t nrows, const double* coef, const double* X, double* Y, qkl_stream s);
int qkl_scale_rows_cols(int nv, size_t nrows, const double* d, const double* X, double* Y, qkl_stream s);

/* BLAS-1 with deterministic partials (npart = qkl_blas_partials()) */
int qkl_dot_partials(int cplx, size_t n, const double* x, const double* y, double* partials, qkl_stream s);
int qkl_axpy(int cplx, size_t n, double are, double aim, const double* x, double* y, qkl_stream s);
int qkl_scal(int cplx, size_t n, double are, double aim, double* x, qkl_stream s);
int qkl_copy_scaled(size_t n_doubles, const double* src, double* dst, const double* scale_inv, int idx, qkl_stream s);
int qkl_fill_random(int kind, size_t n_doubles, uint64_t seed, uint64_t counter0, int ncol, int ld, int c0,
                    double* x, qkl_stream s);
int qkl_fill_zero(size_t n_doubles, double* x, qkl_stream s);
/* *p = v on the stream (a one-thread kernel: no host staging, graph-safe) */
int qkl_set_int(int* p, int v, qkl_stream s);
/* dst[i] = vals[i], i < n <= 8, on the stream; the values travel as kernel
 * arguments, so the call never waits for a host-to-device copy */
int qkl_set_vals(double* dst, int n, const double* vals, qkl_stream s);
int qkl_copy(size_t n_doubles, const double* src, double* dst, qkl_stream s);
/* gather-rate probe: blocks x 8 waves, each `tiles` rounds of 16 coalesced
 * gathers of random 64-lane segments (8 or 16 bytes per lane) of x[n];
 * y holds blocks * 512 doubles */
int qkl_gather_probe(const double* x, size_t n, int bytes_per_lane, double* y, int blocks, int tiles, qkl_stream s);
/* dst[r * dst_ld + j] = src[r * src_ld + j], r < nrows, j < width <= both ld */
int qkl_copy_strided(size_t nrows, int width, const double* src, size_t src_ld, double* dst, size_t dst_ld,
                     qkl_stream s);

/* h[j] partials: partials[j * npart + b] = sum_i V[j][i] u[i], j < m (real) */
int qkl_multidot(size_t n, int m, const double* V, size_t ldv, const double* u, double* partials, qkl_stream s);
/* u[i] += sum_j coef[j] * V[j][i]; coef on device (coef_scale: multiply by -1 when negate) */
int qkl_multiaxpy(size_t n, int m, const double* V, size_t ldv, const double* coef, double coef_sign, double* u,
                  qkl_stream s);

/* diagonal of H into d (complex interleaved when cplx) */
int qkl_diagonal(const qk::QkTables* T, int cplx, int wide, double* d, qkl_stream s);

/* entries of rows [r0, r0 + nrows): cols[r * cap + j], vals (complex
 * interleaved) and counts[r]; diagonal first (col = row), then branches */
int qkl_row_entries(const qk::QkTables* T, int wide, uint64_t r0, uint64_t nrows, int cap, int64_t* cols,
                    double* vals, int* counts, qkl_stream s);

/* states of rows [r0, r0 + n) */
int qkl_states(const qk::QkTables* T, int wide, uint64_t r0, uint64_t n, uint64_t* out, qkl_stream s);

/* diagonal operator: y[row] = f(state) * x[row], f = sum of diag terms
 * (used to build S^z_q seeds): cplx x/y interleaved. */
int qkl_apply_diag_terms(const qk::QkTables* T, int wide, int xcplx, const double* x, double* y, qkl_stream s);

/* Second pass of the split apply (qk_split.hpp): y = acc + H x over these
 * tables, plain or in the Lanczos form of qkl_lanczos_apply (w formed from
 * acc + H x). Real vectors, 32-bit states, no parity masks, one-level lo
 * rank; otherwise -1. qkl_apply_partials_acc: partial sums of the Lanczos
 * form. */
int qkl_apply_acc(const qk::QkTables* T, const double* x, double* y, const double* acc, qkl_stream s);
int qkl_lanczos_apply_acc(const qk::QkTables* T, const double* x, double* y, int k, const int* kdev, int koff,
                          const double* prev, double* nrm, const double* nrm_part, int n_nrm_part, double* partials,
                          int n_part, const double* acc, double* nrm_host, qkl_stream s);
int qkl_apply_partials_acc(const qk::QkTables* T);
/* rows of the relabelled basis of Tint <-> rows of the public basis of Tpub
 * (32-bit states; bytemap: 4 x 256 public-bit tables of the internal bytes) */
int qkl_permute(const qk::QkTables* Tint, const uint32_t* bytemap, const qk::QkTables* Tpub, int to_pub,
                const double* src, double* dst, qkl_stream s);

/* Static per-device workspace of the no-allocation entry points: its
 * device address on the current device (looked up once per device) and its
 * size in bytes. */
int qkl_scratch(void** ptr, size_t* bytes);

/* Davidson preconditioner t = r / (diag - lambda) with the reference clamp */
int qkl_davidson_precond(int cplx, size_t n, const double* r, const double* diag, double lambda, double* t,
                         qkl_stream s);

#ifdef __cplusplus
}
#endif
#endif
