/* qkrylov_amd.h -- extensions of the drop-in C ABI (qkrylov/c_api.h) for
 * the MI355X engine: device memory helpers, device-resident and batched
 * applies, the full Lanczos driver, batched FTLM / continued fraction,
 * operator application, diagnostics and the serving loop. None of these has
 * a counterpart in the reference ABI; they extend it for callers that keep
 * vectors in HBM (see the per-function notes). All return 0 on success,
 * -1 for an invalid argument, -2 for a runtime failure (message in
 * qkrylov_amd_last_error()).
 */
#ifndef QKRYLOV_AMD_H
#define QKRYLOV_AMD_H

#include "qkrylov/c_api.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- runtime ---- */
const char* qkrylov_amd_last_error(void);
int qkrylov_amd_device_count(void);
int qkrylov_amd_set_device(int device_id);
void* qkrylov_amd_malloc(size_t bytes);                 /* device memory */
int qkrylov_amd_free(void* ptr);
int qkrylov_amd_memcpy_h2d(void* dst_dev, const void* src_host, size_t bytes);
int qkrylov_amd_memcpy_d2h(void* dst_host, const void* src_dev, size_t bytes);
int qkrylov_amd_synchronize(void);
/* x[i] = uniform[-1,1) (kind 0) or +-1 (kind 1) from the counter RNG
 * (splitmix64 of seed and index i); on the library stream, synchronous */
int qkrylov_amd_fill_random(double* x_dev, size_t n, uint64_t seed, int kind);
/* dst[i] = src[i] for n doubles with a plain streaming kernel (synchronous);
 * also the reference kernel for calibrating HBM byte counters */
int qkrylov_amd_copy(double* dst_dev, const double* src_dev, size_t n);
/* Measurement of the vector memory path, not part of any solver (no
 * counterpart in the reference): wave-wide coalesced gathers of random
 * 64-lane segments, 8 or 16 bytes per lane, from a device array of
 * array_bytes; reports the mean time of one launch over `reps` launches and
 * the wave gathers one launch issues. */
int qkrylov_amd_gather_probe(size_t array_bytes, int bytes_per_lane, int reps, double* ms_per_launch,
                             double* gathers_per_launch);

/* ---- sector ---- */
/* Fixes one species of a Hubbard / t-J sector alone (species 0: up
 * electrons, 1: down electrons), leaving the other free; the reference's
 * Python Sector sets use_nup / use_ndn independently
 * (bindings/python/qkrylov/basis.py:136-149), its C ABI only both at once
 * (qkrylov_sector_set_hubbard_particles). */
int qkrylov_amd_sector_set_species(qkrylov_sector_h sector, int species, int count);

/* ---- Hamiltonian construction options and info ---- */
/* NL: number of low state bits ranked per lane (-1: automatic) */
qkrylov_hamiltonian_h qkrylov_hamiltonian_create_ex(qkrylov_basis_h basis, qkrylov_site_h site,
                                                    qkrylov_opsum_h opsum, int NL);
/* Operator O whose apply gives y = O x (the Hamiltonian apply of the
 * reference computes y = O^dagger x, matrix_free_hamiltonian.cpp:75-84):
 * built from the adjoint OpSum. Used to form seeds such as S^z_q |psi0>. */
qkrylov_hamiltonian_h qkrylov_operator_create(qkrylov_basis_h basis, qkrylov_site_h site, qkrylov_opsum_h opsum);

typedef struct {
    uint64_t dim;
    int nbits, NL, is_real, n_offdiag_branches, n_diag_terms;
    uint32_t ntiles;
    uint64_t table_bytes;
} qkrylov_hamiltonian_info_t;
int qkrylov_hamiltonian_info(qkrylov_hamiltonian_h h, qkrylov_hamiltonian_info_t* info);
/* Launch tuning of the apply kernels. The apply's output is bitwise the
 * same under every setting; only speed differs (the Lanczos drivers' fused
 * per-workgroup partial sums follow the tile-to-workgroup assignment, so
 * alpha can change in the last bits). Not to be changed while applies or a
 * solver on this handle are running.
 *   QKRYLOV_TUNE_SCHEDULE        tile schedule of the single-vector apply:
 *                                0 one contiguous tile segment per XCD
 *                                (default), 1 one front swept by all XCDs,
 *                                2 one front of XCD-contiguous chunks
 *   QKRYLOV_TUNE_BLOCK_SCHEDULE  the same for the multi-vector apply
 *                                (default 1)
 *   QKRYLOV_TUNE_APPLY_WAVES     waves per workgroup of the single-vector
 *                                apply, sharing one LDS copy of the tables:
 *                                4, 8 or 16 (default)
 *   QKRYLOV_TUNE_BLOCK_WAVES     waves per workgroup of the multi-vector
 *                                apply: 16, 8 (default) or 4, with 8, 16 or
 *                                32 gathers in flight per wave
 *   QKRYLOV_TUNE_BLOCK_ROWS      rows per work item of the multi-vector
 *                                apply: 64 (default), 32 or 16; fewer rows
 *                                in flight share more of their gathers in
 *                                the caches */
enum {
    QKRYLOV_TUNE_SCHEDULE = 0,
    QKRYLOV_TUNE_BLOCK_SCHEDULE = 1,
    QKRYLOV_TUNE_APPLY_WAVES = 2,
    QKRYLOV_TUNE_BLOCK_WAVES = 3,
    QKRYLOV_TUNE_BLOCK_ROWS = 4,
    QKRYLOV_TUNE_SPLIT_SCHEDULE = 5,   /* tile schedule of both passes of the split apply (default 0) */
    QKRYLOV_TUNE_SPLIT_WAVES = 6       /* waves per workgroup of both passes (default 16) */
};
int qkrylov_hamiltonian_set_tuning(qkrylov_hamiltonian_h h, int knob, int value);

/* Autotuner: times y = H x (nvec columns, random vectors) under the launch
 * tunings (for nvec > 1 also the work-item widths) and, when allow_split,
 * the lo-bit splits around the current one, and keeps the fastest for this
 * handle. Launch tunings never change the apply's output; another split
 * changes the order in which a row's entries are added, i.e. the last bits.
 * Allocates two vectors of nvec columns. */
typedef struct {
    int split_bits, apply_waves, schedule, block_waves, block_schedule, candidates;
    double ms_before, ms_after;
    int block_rows;
} qkrylov_tune_report_t;
int qkrylov_hamiltonian_autotune(qkrylov_hamiltonian_h h, int nvec, int allow_split, qkrylov_tune_report_t* report);
int qkrylov_hamiltonian_get_tuning(qkrylov_hamiltonian_h h, int knob, int* value);

/* ---- device-resident applies (pointers from qkrylov_amd_malloc or any
 * device allocation; stream may be NULL for the library stream) ---- */
/* real vectors; nvec > 1: row-major block x[row * nvec + c] */
int qkrylov_hamiltonian_apply_device_f64(qkrylov_hamiltonian_h h, const double* x_dev, double* y_dev, int nvec,
                                         void* stream);
/* nv in {1, 2, 4, 8} real row-major columns x[row * nv + c]; column c is
 * multiplied by H with its diagonal scaled by diag_scale[c] (host array of
 * nv), every decoded matrix element serving all columns: the batched apply
 * of requests that share the off-diagonal part of H and differ in the weight
 * of its diagonal (the serving loop's XXZ anisotropies). Column c's result
 * does not depend on nv or on the other columns. Real Hamiltonians without
 * fermion signs on at most 32 bits; others are refused (-1). No counterpart
 * in the reference. */
int qkrylov_hamiltonian_apply_columns_device_f64(qkrylov_hamiltonian_h h, const double* x_dev, double* y_dev, int nv,
                                                 const double* diag_scale, void* stream);
/* interleaved complex vectors */
int qkrylov_hamiltonian_apply_device_complex_f64(qkrylov_hamiltonian_h h, const double* x_dev, double* y_dev,
                                                 void* stream);
/* diagonal into a device array (complex interleaved when cplx) */
int qkrylov_hamiltonian_diagonal_device_f64(qkrylov_hamiltonian_h h, double* d_dev, int cplx, void* stream);

/* entries of rows [row0, row0 + nrows) as produced by the device kernels:
 * diagonal first (col = row), then off-diagonal entries; host buffers
 * cols[nrows * cap], vals[2 * nrows * cap] (complex), counts[nrows] */
int qkrylov_hamiltonian_row_entries(qkrylov_hamiltonian_h h, uint64_t row0, uint64_t nrows, int cap, int64_t* cols,
                                    double* vals, int* counts);
/* basis states of rows [row0, row0 + n) computed on the device (host out) */
int qkrylov_hamiltonian_states(qkrylov_hamiltonian_h h, uint64_t row0, uint64_t n, uint64_t* out);

/* timing helper: `reps` applies of y = H x on random device vectors
 * (nvec columns), timed with HIP events on the library stream; reports the
 * mean per-apply time of the events bracket (ms) */
int qkrylov_hamiltonian_time_apply(qkrylov_hamiltonian_h h, int nvec, int warmup, int reps, double* mean_ms);

/* ---- two-pass apply (real spin-1/2 models, 32 sites at most): the sites
 * are relabelled into groups [B | C | A] (A the low bits) such that every
 * off-diagonal term acts inside A u B or inside A u C; pass 1 applies the
 * A u B terms and the diagonal with the C bits fixed, pass 2 the rest with
 * the B bits fixed, so each pass walks units of at most 21 sites that stay
 * in one XCD's L2. Vectors of the relabelled basis are a permutation of the
 * public ones; the Lanczos driver keeps its vectors in that order. ---- */
/* NL, nB, nC of the split (0, 0, 0 when the model admits none; below 2^24
 * rows only when any_size is set) */
int qkrylov_hamiltonian_split_info(qkrylov_hamiltonian_h h, int any_size, int* NL, int* nB, int* nC);
/* y = H x through the two passes, public row order in and out (permutes
 * both; allocates three vectors) */
int qkrylov_hamiltonian_apply_split_device_f64(qkrylov_hamiltonian_h h, const double* x_dev, double* y_dev,
                                               void* stream);
/* mean time of the two passes on vectors in the internal order (HIP events
 * on the library stream, as qkrylov_hamiltonian_time_apply) */
int qkrylov_hamiltonian_time_split_apply(qkrylov_hamiltonian_h h, int warmup, int reps, double* mean_ms);

/* ---- Lanczos driver ---- */
typedef struct {
    int maxiter;          /* reference default 200 */
    double tol;           /* |E_k - E_{k-1}| < tol, reference default 1e-12 */
    int reorthogonalize;  /* 1 = reference behaviour */
    int fixed_steps;      /* 1: run min(maxiter, dim) steps, no test */
    int want_vector;
    uint64_t seed;        /* start vector uniform[-1,1) from (seed, index) */
    int check_every;      /* 0: automatic */
    int graph_mode;       /* 0: automatic (replay a captured six-step graph for
                           * fixed-step runs below 2^22 rows when no basis is
                           * stored), 1: always, 2: never */
    int split_mode;       /* two-pass apply over a relabelled basis (real spin-1/2
                           * models whose terms admit it): 1 use it when admitted;
                           * 0 (automatic) and 2 the single-pass apply, which
                           * measured faster on the J1-J2 4x8 and 32-site ring */
    int start_mode;       /* 0: `start`, else uniform[-1,1) from (seed, index),
                           * real for real H. 1 / 2: the reference's start
                           * (src/solvers/lanczos.cpp:115-119): complex entries
                           * of two std::mt19937(seed) uniform[-1,1) draws, the
                           * imaginary part drawn first (1: GCC's evaluation of
                           * its KComplex(dist(rng), dist(rng))) or the real
                           * part first (2: Clang's); the recurrence runs on
                           * complex vectors for real H too, a given `start` is
                           * interleaved complex and `vector` receives 2*dim
                           * doubles. -1: the process default
                           * (qkrylov_amd_set_start_mode) */
} qkrylov_lanczos_options_t;

typedef struct {
    double energy;
    int iterations;
    int converged;
    int steps_run;
    int n_alpha, n_beta;
    double loop_seconds;
    double ritz_seconds;
    int split;            /* 1: the two-pass apply ran */
    int vector_complex;   /* 1: `vector` holds 2*dim doubles (interleaved complex) */
} qkrylov_lanczos_report_t;

/* Start vector of the solvers that take no options: the reference-named
 * Lanczos calls (qkrylov_lanczos_ground_state*, C++ lanczos_ground_state,
 * Python lanczos_ground_state) and start_mode -1 above. mode 0 (default):
 * counter RNG; 1 / 2: the reference's std::mt19937(1234) complex draws as
 * described for start_mode, so that a drop-in run follows the reference's
 * Krylov trajectory (iteration count, and the Ritz vector up to a sign).
 * qkrylov_ftlm always draws the reference's complex normal vectors
 * (src/solvers/ftlm.cpp:133-143, std::mt19937(42)); mode 1 makes the first
 * draw of an entry its imaginary part (GCC's argument order), any other
 * mode its real part (Clang's). Process-wide; returns -1 for other values. */
int qkrylov_amd_set_start_mode(int mode);
int qkrylov_amd_get_start_mode(void);
/* the reference's start vector itself: n complex entries interleaved into
 * out[2n] (host memory; no device needed), mode 1 or 2 as above */
int qkrylov_amd_reference_start_vector(size_t n, uint64_t seed, int mode, double* out);
/* the reference's FTLM draws (std::mt19937(seed), std::normal_distribution
 * (0, 1), one stream over all random vectors): n complex entries
 * interleaved into out[2n], in the order of mode 1 or 2 (host memory; no
 * device needed). Replaces the draws of src/solvers/ftlm.cpp:133-143. */
int qkrylov_amd_reference_normal_vector(size_t n, uint64_t seed, int mode, double* out);

/* start: optional host start vector (real for real H, interleaved complex
 * otherwise); alphas/betas: host arrays of maxiter entries (optional);
 * vector: host output (dim doubles for real H with start_mode 0, 2*dim
 * otherwise; optional) */
int qkrylov_lanczos_run_f64(qkrylov_hamiltonian_h h, const qkrylov_lanczos_options_t* opt, const double* start,
                            double* alphas, double* betas, double* vector, qkrylov_lanczos_report_t* report);

/* The driver of qkrylov_lanczos_run_f64 without any device allocation, for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). No counterpart in the reference.
 * Every device buffer (three Krylov vectors, alpha/norm arrays, partial
 * sums, the stored basis when reorthogonalize is set, the Ritz pass) comes
 * from one static 64 MiB workspace per GPU that the library holds until the
 * process ends; sizes that do not fit are refused with -2 and a message
 * (roughly dim <= 2.0e6 with want_vector, 2.7e6 without, far less with
 * reorthogonalize). Real Hamiltonians only; no graph replay (graph_mode is
 * ignored). Results equal qkrylov_lanczos_run_f64's. Calls on one GPU share
 * its workspace: this library serialises them within a process, so calls
 * that overlap on one GPU wait for each other; the workspace keeps the last
 * call's contents and every buffer is written before it is read. */
int qkrylov_lanczos_run_noalloc_f64(qkrylov_hamiltonian_h h, const qkrylov_lanczos_options_t* opt, const double* start,
                                    double* alphas, double* betas, double* vector, qkrylov_lanczos_report_t* report);

/* ---- batched FTLM: R Rademacher vectors (counter RNG: column r, row i)
 * run as one multi-vector recurrence of M steps (sequential = 1 runs the
 * same vectors one at a time); observables for nT inverse temperatures.
 * alphas/betas_l: R*M host arrays (row r at r*M), nsteps[R]. ---- */
int qkrylov_ftlm_batched_f64(qkrylov_hamiltonian_h h, int R, int M, uint64_t seed, const double* betas_T, int nT,
                             int sequential, double* Z, double* E, double* Cv, double* alphas, double* betas_l,
                             int* nsteps, double* seconds);

/* ---- batched continued fraction: nseeds complex host seeds (each 2*dim
 * doubles, seed s at s*2*dim); alphas/betas: nseeds*n_iter; norms[nseeds];
 * n_alpha[nseeds] ---- */
int qkrylov_cf_batched_f64(qkrylov_hamiltonian_h h, int nseeds, const double* seeds_complex, int n_iter,
                           double* alphas, double* betas, double* norms, int* n_alpha, double* seconds);

/* ---- dynamical structure factor S_q(omega) for nops operators O_q (handles
 * from qkrylov_operator_create, e.g. S^z_q = N^-1/2 sum_r e^{iq.r} S^z_r):
 * seeds phi_q = O_q psi0 formed on the device, one block continued fraction
 * of n_iter steps over all seeds, then
 * S_q(omega_w) = -1/pi Im <phi_q|(omega_w + E0 + i eta - H)^-1|phi_q>.
 * psi0: host, dim doubles (psi0_complex = 0) or 2*dim interleaved.
 * spectrum: nops*nomega (row q at q*nomega); alphas/betas (optional):
 * nops*n_iter; norms, n_alpha (optional): nops. The reference has no such
 * driver: it builds phi0 by hand (tests/test_dynamics.cpp:41-46) and calls
 * continued_fraction_coeffs / evaluate_spectral_function
 * (include/qkrylov/solvers/dynamics.hpp:20-38) per seed. ---- */
int qkrylov_structure_factor_f64(qkrylov_hamiltonian_h h, const qkrylov_hamiltonian_h* operators, int nops,
                                 const double* psi0, int psi0_complex, double E0, int n_iter, const double* omega,
                                 int nomega, double eta, double* spectrum, double* alphas, double* betas,
                                 double* norms, int* n_alpha, double* seconds);

/* ---- serving loop (continuous batching of ground-state and S^zz(omega)
 * solves of XXZ rings): see qkrylov_amd/csrc/qk_serving.cpp ---- */
typedef struct {
    int L;            /* sites */
    int sz2;          /* 2 Sz */
    double delta;     /* anisotropy */
    int kind;         /* 0 ground state, 1 ground state + S^zz continued fraction */
    double arrival;   /* seconds from trace start */
    double tol;       /* ground-state tolerance */
    int cf_steps;     /* continued-fraction steps */
    int maxiter;      /* Lanczos cap */
} qkrylov_serve_request_t;

typedef struct {
    double energy;
    int iterations;
    int converged;
    int n_cf;
    double cf_norm;
    double start, finish;   /* seconds from trace start */
    double latency;         /* finish - arrival */
    int batch;              /* requests in the batch this one ran in */
    int ritz_stored;        /* kind 1: the Ritz vector was summed from stored
                             * Krylov vectors (1) or from the recurrence run
                             * again (0, when the store could not hold them);
                             * the same numbers either way */
} qkrylov_serve_result_t;

typedef struct {
    size_t hbm_budget_bytes;   /* admission budget for the batches' device
                                * footprints (0: 80% of free memory). Stored
                                * Krylov vectors of kind-1 requests use a
                                * separate buffer: up to a quarter of the
                                * budget, within the free memory the budget
                                * leaves less 1 GiB and at most 64 GiB,
                                * allocated by the first run on the device that
                                * stores vectors and kept by the process */
    int max_active;            /* concurrent batches on the device */
    int realtime;              /* 1: honour arrival times; 0: all arrive at t=0 */
    double* const* gs_vectors; /* optional, nreq host pointers (each NULL or dim
                                * doubles): the normalised ground state a kind-1
                                * request's continued fraction started from */
    int max_batch;             /* requests of one (L, Sz) sector run as the columns
                                * of one batch of 1, 2, 4 or 8, up to this many
                                * (0: 4; at most 8) */
} qkrylov_serve_options_t;

/* Runs the trace. cf_alphas/cf_betas: nreq * cf_cap host arrays (optional).
 * totals: [0] wall seconds, [1] solves/s, [2] p50 latency, [3] p99 latency.
 * Runs on one device are serialised (they share its vector store). */
int qkrylov_serve_run(const qkrylov_serve_request_t* reqs, int nreq, const qkrylov_serve_options_t* opt,
                      qkrylov_serve_result_t* results, double* cf_alphas, double* cf_betas, int cf_cap,
                      double* totals);

/* ---- host eigensolvers used by the Krylov drivers (no device work; usable
 * without a GPU) ---- */
/* Symmetric tridiagonal (alpha[0..n), beta[0..n-1)): eigenvalues ascending
 * in evals; want 0: none, 1: evecs[i] = first component of eigenvector i,
 * 2: evecs[r * n + i] = component r of eigenvector i (implicit QL). The
 * quantities of the reference's tridiag_ground_state_full and
 * diagonalize_tridiag_components (src/solvers/lanczos.cpp:25-98,
 * src/solvers/ftlm.cpp:61-119). */
int qkrylov_amd_tridiag_eigen(int n, const double* alpha, const double* beta, double* evals, double* evecs, int want);
/* lowest eigenvalue only (Sturm bisection) */
int qkrylov_amd_tridiag_lowest(int n, const double* alpha, const double* beta, double* lowest);
/* Hermitian n x n (row-major, interleaved complex): ascending eigenvalues,
 * evecs[2 * (r * n + i)] = component r of eigenvector i (Householder to
 * tridiagonal + QL); the small eigenproblem of the reference's
 * solve_small_hermitian (src/solvers/davidson.cpp:22-82). */
int qkrylov_amd_hermitian_eigen(int n, const double* a, double* evals, double* evecs);

#ifdef __cplusplus
}
#endif
#endif /* QKRYLOV_AMD_H */
