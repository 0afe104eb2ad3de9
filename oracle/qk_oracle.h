/*
 * qk_oracle.h -- CPU restatement of the sjp95/qkrylov algorithm (TEST INFRASTRUCTURE).
 *
 * This library is the checker for the HIP engine in qkrylov_amd/. It is never
 * linked into, loaded by, or called from the product path: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * It restates, in plain C99 + OpenMP, the reference's
 *   - bases        (src/basis/{spinhalf,fermion,hubbard,tj}_basis.cpp): ascending
 *                  enumeration of valid bitstrings, index lookup;
 *   - site actions (src/sites/{spinhalf,fermion,hubbard,tj}_site.cpp);
 *   - CSR build    (src/hamiltonian/matrix_free_hamiltonian.cpp:45-92, Phase 1);
 *   - gather SpMV  (matrix_free_hamiltonian.cpp:156-167), complex FP64;
 *   - solvers      (src/solvers/lanczos.cpp:25-202, ftlm.cpp:24-170,
 *                   dynamics.cpp:12-89, davidson.cpp:22-227), complex FP64.
 * Random start vectors are inputs here (the reference draws them from
 * std::mt19937 inside the solver); everything downstream of the start vector
 * follows the reference step by step.
 *
 * Pinning (tests/test_oracle.py): basis lists, site actions and CSR triplets
 * against fixtures emitted by the reference's own compiled basis/site/opsum
 * sources (oracle/_ref); closed-form dimensions; the reference tests' known
 * answers (-0.75, -1.6160254038, -2.0, 0.3125, JW signs); dense numpy ED built
 * from independent Kronecker/Jordan-Wigner matrices; Bethe-ansatz energies.
 */
#ifndef QK_ORACLE_H
#define QK_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { QO_SPIN = 0, QO_FERMION = 1, QO_HUBBARD = 2, QO_TJ = 3 };

typedef struct {
    int use_sz, sz2;
    int use_nup, nup, use_ndn, ndn;
    int use_n, n;
    int use_nb, nb;
} qo_sector;

typedef struct qo_model qo_model;

/* ---- site actions (reference src/sites/*.cpp) ---- */
/* returns 1 valid / 0 invalid / -1 unknown operator (reference throws) */
int qo_site_apply(int site_kind, const char* op, int site, uint64_t state,
                  uint64_t* new_state, double* amp_re, double* amp_im);

/* ---- basis (reference src/basis/*.cpp) ---- */
/* enumerates; returns dim, or -1 when the reference constructor throws */
int64_t qo_basis_enumerate(int basis_kind, int nsites, const qo_sector* sec,
                           uint64_t* out_states, int64_t cap);
/* closed-form counting rank (the oracle's own; pinned against the enumeration) */
int64_t qo_count_dimension(int basis_kind, int nsites, const qo_sector* sec);
int     qo_count_contains(int basis_kind, int nsites, const qo_sector* sec, uint64_t s);
int64_t qo_count_rank(int basis_kind, int nsites, const qo_sector* sec, uint64_t s);
uint64_t qo_count_unrank(int basis_kind, int nsites, const qo_sector* sec, int64_t r);

/* ---- model = basis + site + opsum ---- */
/* enumerate=1: build the state list like the reference; 0: counting rank only */
qo_model* qo_model_create(int basis_kind, int nsites, const qo_sector* sec,
                          int site_kind, int enumerate);
void      qo_model_destroy(qo_model* m);
int       qo_model_add_term(qo_model* m, double re, double im, int nf,
                            const char** ops, const int* sites);
int64_t   qo_model_dim(const qo_model* m);
uint64_t  qo_model_state(const qo_model* m, int64_t i);
int64_t   qo_model_index(const qo_model* m, uint64_t s);     /* -1 if absent */

/* Phase-1 CSR build; returns nnz or -1 on error (unknown operator) */
int64_t   qo_build_csr(qo_model* m);
int64_t   qo_csr_nnz(const qo_model* m);
void      qo_csr_export(const qo_model* m, int64_t* row_offsets, int64_t* cols,
                        double* vals_interleaved, double* diag_interleaved);
/* entries of one row, computed directly (no CSR): returns count or -1 */
int       qo_row_entries(const qo_model* m, int64_t row, int64_t* cols,
                         double* vals_interleaved, int cap);

/* complex gather SpMV over the CSR; x, y interleaved re/im; OpenMP */
void      qo_apply(const qo_model* m, const double* x, double* y);
/* same over a contiguous row range [r0, r0+nrows) */
void      qo_apply_rows(const qo_model* m, const double* x, double* y,
                        int64_t r0, int64_t nrows);
/* time-only helper: CSR of a row range (for a row sample of a huge basis);
   returns nnz. Uses counting rank, so needs no enumeration. */
int64_t   qo_build_csr_rows(qo_model* m, int64_t r0, int64_t nrows);

/* ---- solvers (complex FP64, reference algorithm) ---- */
typedef struct {
    double energy;
    int iterations;
    int converged;
    int n_alpha;
    int n_beta;
} qo_lanczos_result;

/* start: interleaved complex, dim entries (normalised inside as reference).
   reorth=1 is the reference behaviour (lanczos.cpp:149-151).
   alphas/betas: capacity maxiter. eigvec: interleaved, may be NULL. */
int qo_lanczos(const qo_model* m, int maxiter, double tol, const double* start,
               int reorth, qo_lanczos_result* res, double* alphas, double* betas,
               double* eigvec);

/* reference tridiag_ground_state_full (lanczos.cpp:25-98) */
double qo_tridiag_ground(const double* alpha, const double* beta, int n, double* evec);
/* reference diagonalize_tridiag_components (ftlm.cpp:61-119) */
void   qo_tridiag_components(const double* alpha, const double* beta, int n,
                             double* evals, double* first_comp);

/* compute_tridiag (ftlm.cpp:24-53): returns number of alphas; nbeta_out set */
int qo_compute_tridiag(const qo_model* m, const double* v_start, int n_steps,
                       double* alphas, double* betas, int* nbeta_out);

/* ftlm (ftlm.cpp:122-170) with caller-provided random vectors
   (rvecs: n_random x dim interleaved complex) */
int qo_ftlm(const qo_model* m, double beta, int n_random, int n_steps,
            const double* rvecs, double* Z, double* E, double* Cv);

/* continued_fraction_coeffs (dynamics.cpp:12-59) */
int qo_cf_coeffs(const qo_model* m, const double* phi0, int n_iter,
                 double* alphas, double* betas, int* n_alpha, int* n_beta,
                 double* norm_phi0);
/* evaluate_spectral_function (dynamics.cpp:61-89) */
double qo_spectral(const double* alphas, const double* betas, size_t n,
                   double norm_phi0, double omega, double E0, double eta);

/* davidson_lowest (davidson.cpp:85-227); eigvecs interleaved n_eig x dim */
int qo_davidson(const qo_model* m, int n_eig, int max_subspace, double tol,
                double* evals, double* eigvecs, int* iterations, int* converged);

/* counter-based random numbers shared with the engine's start-vector
   generator (splitmix64 of seed ^ golden*counter) */
uint64_t qo_rng_u64(uint64_t seed, uint64_t counter);
double   qo_rng_uniform(uint64_t seed, uint64_t counter);   /* [-1, 1) */

#ifdef __cplusplus
}
#endif
#endif
