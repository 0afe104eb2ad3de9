// qk_engine.hpp -- device-resident Hamiltonian engine and solver kernels
// orchestration (host side). Calls the HIP kernels only through qk_launch.h.
#pragma once
#include "qk_launch.h"
#include "qk_model.hpp"

#include <hip/hip_runtime_api.h>

#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <vector>

namespace qk {

class SplitEngines;

void hip_check(hipError_t e, const char* what);

// RAII device allocation (hipMalloc / hipFree)
class DevBuf {
public:
    DevBuf() = default;
    explicit DevBuf(size_t bytes);
    // stream-ordered allocation: hipMallocAsync now, hipFreeAsync on release,
    // so freeing never waits for work queued on other streams
    DevBuf(size_t bytes, hipStream_t async_stream);
    ~DevBuf();
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept
        : p_(o.p_), bytes_(o.bytes_), stream_(o.stream_), async_(o.async_), owned_(o.owned_)
    {
        o.p_ = nullptr;
        o.bytes_ = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept;
    // non-owning view of memory that someone else keeps (never freed here)
    static DevBuf view(void* p, size_t bytes)
    {
        DevBuf b;
        b.p_ = p;
        b.bytes_ = bytes;
        b.owned_ = false;
        return b;
    }
    void* get() const { return p_; }
    double* d() const { return static_cast<double*>(p_); }
    size_t bytes() const { return bytes_; }
    void reset();

private:
    void* p_ = nullptr;
    size_t bytes_ = 0;
    hipStream_t stream_ = nullptr;
    bool async_ = false;
    bool owned_ = true;
};

// Bump allocator over device memory the caller keeps (a fixed workspace):
// take() hands out 256 B aligned non-owning views and throws when the
// workspace is exhausted; mark()/rewind() reuse a region after its users are
// gone. Used by the entry points that must not allocate.
struct Arena {
    char* base = nullptr;
    size_t cap = 0, used = 0;
    DevBuf take(size_t bytes);
    size_t mark() const { return used; }
    void rewind(size_t m) { used = m; }
};

// one library stream per host thread (default priority)
hipStream_t default_stream();
void ensure_device(int id);

// Device copies of the rank / unrank tables of a basis (tiles, hoff, lrank,
// lolist). They depend only on the basis and the bit split, so every
// Hamiltonian over the same basis shares one copy.
struct BasisTables {
    HostTables host;   // NL, NH, lrank_u16 (the large host arrays are released)
    DevBuf tiles, hoff, lrank, lolist, tl, th;
    size_t ntiles = 0;
};
std::shared_ptr<const BasisTables> basis_tables(const BasisSpec& b, int NL);

struct EngineOptions {
    int NL = -1;                   // lo-bit split (-1: automatic)
    hipStream_t stream = nullptr;  // non-null: stream-ordered term-table allocation
};

class Engine {
public:
    Engine(std::shared_ptr<const BasisSpec> basis, Kind site_kind, const OpSum& ops, EngineOptions opt = {});

    int64_t dim() const { return basis_->dim(); }
    bool is_real() const { return comp_.is_real; }
    bool wide() const { return basis_->nbits() > 32; }
    const QkTables* tables() const { return &T_; }
    const BasisSpec& basis() const { return *basis_; }
    const Compiled& compiled() const { return comp_; }
    int NL() const { return tables_->host.NL; }
    size_t table_bytes() const { return table_bytes_; }
    // launch tuning of the apply kernels (knobs in qkrylov_amd.h); results
    // do not depend on it. Not to be changed while applies are in flight or
    // a solver on this engine is running.
    int tuning(int knob) const;
    void set_tuning(int knob, int value) const;

    // y = H^dagger-convention apply of the reference (y[a] = sum conj(amp) x[b])
    void apply(const double* x, double* y, bool cplx, hipStream_t s) const;
    void apply_block(const double* X, double* Y, int ld, int ncols, hipStream_t s) const;
    // Y[r * nv + c] = (H_c X_c)[r] for nv in {1, 2, 4, 8} row-major real
    // columns, H_c = H with its diagonal scaled by dscale[c] (device array)
    void apply_cols(const double* X, double* Y, int nv, const double* dscale, hipStream_t s) const;
    void diagonal(double* d, bool cplx, hipStream_t s) const;
    // y = O x for a diagonal operator compiled into this engine
    void apply_diag_operator(const double* x, bool xcplx, double* y, hipStream_t s) const;

    // The two-pass apply over a relabelled basis (qk_split.hpp) when this
    // model admits one and has at least min_rows rows, else null; built on
    // the first call and kept with the engine.
    static constexpr int64_t kSplitMinRows = int64_t(1) << 24;
    const SplitEngines* split_engines(int64_t min_rows = kSplitMinRows) const;

private:
    std::shared_ptr<const BasisSpec> basis_;
    Kind kind_;
    OpSum ops_;
    mutable std::mutex split_mu_;
    mutable std::shared_ptr<const SplitEngines> split_;
    mutable bool split_tried_ = false;
    Compiled comp_;
    std::shared_ptr<const BasisTables> tables_;
    DevBuf br_, brk_, dg_;
    mutable QkTables T_{};
    size_t table_bytes_ = 0;
};

double now_s();

// message returned by qkrylov_amd_last_error() on this thread
void set_last_error(const char* msg);

// Autotuner: times y = H x (nvec columns) on random vectors under candidate
// launch tunings and, when allow_split, lo-bit splits around the current
// one; returns the fastest engine (the current one when nothing beats it).
// Launch tunings never change results; a different split changes the order
// in which a row's entries are added, i.e. the last bits.
struct TuneResult {
    int NL = -1, apply_waves = 0, sched = 0, blk_waves = 0, blk_sched = 0, blk_rows = 64;
    double ms = 0.0, ms_before = 0.0;
    int tried = 0;
};
std::shared_ptr<const Engine> autotune(const std::shared_ptr<const BasisSpec>& basis, Kind kind, const OpSum& ops,
                                       const std::shared_ptr<const Engine>& current, int nvec, bool allow_split,
                                       hipStream_t s, TuneResult& out);

// Replays the reference's convergence logic of lanczos_ground_state
// (src/solvers/lanczos.cpp:135-183) on alpha_k and beta_k = |q_{k+1}|.
struct LanczosReplay {
    int kmax;
    double tol;
    bool fixed;
    std::vector<double> alphas, betas;
    double e_old = 1e100;
    int next = 0;
    bool done = false, converged = false;
    void feed(double a, double b);
};

// Device state of the fused real three-term recurrence: three rotating
// vectors q (stored unnormalised with their norms nrm[k] on the device),
// alpha[k], and the partial-sum buffers. One step = fused apply + alpha
// reduction + update + norm reduction, all on stream s, no host sync.
struct RealCore {
    const Engine& E;
    hipStream_t s;
    size_t D;
    int kcap;
    DevBuf Q[3];
    DevBuf alpha, nrm, hcoef, partA, partB, partM;
    int npA, NB;
    int ip = 0, ic = 1, iw = 2;
    // graph of three plain steps (one full turn of the vector rotation) whose
    // kernels read the step from the device counter kdev
    DevBuf kdev;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int cap_ic = -1;   // rotation state the graph was captured in
    int pending = -1;  // nrm index whose partial sums are not yet totalled
    // set: coherent pinned host memory laid out as alpha and nrm on the
    // device (alpha at 0, nrm at kcap + 2), into which the kernels that
    // total them store them too
    double* host_mirror = nullptr;

    Arena* arena = nullptr;   // set: every device buffer comes from it
    bool stream_ordered = false;   // buffers from hipMallocAsync on s
    // set: the apply runs as the two passes of the split engines, in their
    // internal row order (tmp holds the first pass)
    const SplitEngines* split = nullptr;
    DevBuf tmp;

    RealCore(const Engine& e, int kcap, hipStream_t st, bool reorth, bool stream_ordered = false,
             Arena* arena = nullptr, const SplitEngines* split = nullptr);
    ~RealCore();
    RealCore(const RealCore&) = delete;
    RealCore& operator=(const RealCore&) = delete;
    double* cur() { return Q[ic].d(); }
    double* prv() { return Q[ip].d(); }
    double* wrk() { return Q[iw].d(); }
    void init_norm();                        // nrm[0] = |cur|, prv = 0
    // step k leaves alpha[k] in memory and the partial sums of nrm[k+1] in
    // partB: the next step's apply totals them; flush() does it now (before
    // nrm[k+1] is read on the host or by another kernel)
    void step(int k, const double* basis);   // basis: normalised v_0..v_k for reorthogonalisation
    // the Lanczos-form apply of step k (k from *kdev + koff when kdev is set)
    int lanczos_apply(int k, const int* kdev, int koff);
    void flush();
    // steps k .. k + n - 1 (n a multiple of 3, k + n <= kcap) as replays of
    // the captured six-step graph (a remaining three steps eagerly); the
    // same kernels in the same order as step(), so the same numbers
    void steps_graph(int k, int n);
};

// Device state of nv independent real three-term recurrences run as the
// columns of row-major blocks (nv in 1, 2, 4, 8): column c uses H with its
// diagonal scaled by dscale[c]. One step = the multi-column Lanczos-form
// apply, the per-column alpha totals, the update and the per-column norm
// totals, on stream s without host synchronisation. Column c computes the
// same bits whatever nv is and whatever the other columns hold; a column
// whose recurrence broke down (beta = 0) holds non-finite numbers from then
// on without touching the others. Vectors stored unnormalised with their
// norms nrm[k * nv + c], as in RealCore.
struct ColsCore {
    const Engine& E;
    hipStream_t s;
    size_t D;
    int nv, kcap;
    DevBuf Q[3], alpha, nrm, partA, partB, dsc, zero, one;
    int npA, NB;
    int ip = 0, ic = 1, iw = 2;
    // dscale: nv host values; stream_ordered: buffers from hipMallocAsync on st
    ColsCore(const Engine& e, int nv, int kcap, const double* dscale, hipStream_t st, bool stream_ordered);
    double* cur() { return Q[ic].d(); }
    double* prv() { return Q[ip].d(); }
    // a new recurrence from cur (cur unchanged; step 0 reads no previous
    // vector): the partial sums of |cur_c|^2, which step 0 totals into nrm[0]
    void init_norm();
    // step k: nrm[k][c] (totalled by its apply), alpha[k][c], and the partial
    // sums of nrm[k + 1][c]; cur becomes q_{k+1}
    void step(int k);
    // nrm[k][c] now (after step k - 1 or init_norm at k = 0), before the host
    // or another kernel reads it; the same bits step k would form
    void flush(int k);
};
// widths ColsCore and qkl_apply_cols take: the smallest of 1, 2, 4, 8 holding n columns (0 if n > 8)
int cols_width(int n);

// ---------------------------------------------------------------- tridiagonal
// Lowest eigenvalue of the Lanczos tridiagonal (alpha, beta) and, when evec
// is given, its unit eigenvector: the quantities of the reference's
// tridiag_ground_state_full (src/solvers/lanczos.cpp:25-98), computed by
// Sturm bisection (value only) or implicit QL (qk_eigen.hpp).
double tridiag_min(const double* alpha, const double* beta, int n, std::vector<double>* evec);

// ---------------------------------------------------------------- Lanczos
struct LanczosOptions {
    int maxiter = 200;
    double tol = 1e-12;
    bool reorth = true;          // reference behaviour: full reorthogonalisation
    bool want_vector = false;
    bool fixed_steps = false;    // run min(maxiter, dim) steps, no convergence test
    uint64_t seed = 1234;
    int check_every = 0;         // 0: automatic
    int graph = -1;              // replay captured step graphs: -1 below 2^22 rows, 0 never, 1 always
    int split = -1;              // two-pass apply (real spin models that admit it, qk_split.hpp): 1 use it,
                                 // -1 (automatic) and 0 the single pass
    const double* start = nullptr;   // host start vector (real, or interleaved complex for complex H)
    const double* start_dev = nullptr;   // device start vector (same layout)
    // 0: the start above, else uniform [-1, 1) from the counter RNG (real for
    // real H). 1, 2: the reference's start (lanczos.cpp:115-119): complex
    // entries of two std::mt19937(seed) uniform [-1, 1) draws each, the
    // imaginary part drawn first (1, as GCC evaluates the reference's
    // KComplex(dist(rng), dist(rng))) or the real part first (2, as Clang
    // does); the recurrence then runs on complex vectors whatever H is, a
    // given start is read as interleaved complex, and the Ritz vector is
    // complex
    int start_mode = 0;
};

// the reference's start vector of lanczos_ground_state (start_mode 1 or 2
// above), D complex entries interleaved into out[2 D]
void reference_start_vector(size_t D, uint64_t seed, int start_mode, double* out);

// The reference's FTLM draws (src/solvers/ftlm.cpp:133-143): one
// std::mt19937(seed) and one std::normal_distribution(0, 1) for all random
// vectors, entry i of a vector = KComplex(dist(rng), dist(rng)). The
// distribution keeps state between calls (the polar method's second
// deviate), so the draws are one stream: fill() continues it. start_mode 1
// stores the first draw of an entry as its imaginary part (GCC's order of
// the two arguments), anything else as its real part (Clang's).
class ReferenceNormal {
public:
    ReferenceNormal(uint64_t seed, int start_mode);
    void fill(size_t n, double* out);   // n complex entries, interleaved
private:
    std::mt19937 rng_;
    std::normal_distribution<double> dist_;
    size_t first_;
};

struct LanczosOutput {
    double energy = 0.0;
    int iterations = 0;
    bool converged = false;
    std::vector<double> alphas, betas;
    std::vector<double> vector;      // real, or interleaved complex for complex H
    double loop_seconds = 0.0;       // wall time of the iteration loop
    double ritz_seconds = 0.0;
    int steps_run = 0;               // device steps computed (>= iterations)
    bool split = false;              // the two-pass apply ran (qk_split.hpp)
};

// arena: when given, every device buffer (Krylov vectors, scalars, partial
// sums, the stored basis) is carved from it and nothing is allocated; real
// Hamiltonians only, no graph capture.
LanczosOutput lanczos_run(const Engine& E, const LanczosOptions& o, hipStream_t s, Arena* arena = nullptr);
// device bytes lanczos_run needs with an arena for these options
size_t lanczos_arena_bytes(const Engine& E, const LanczosOptions& o);

// three-term recurrence from a start vector without reorthogonalisation;
// alphas/betas truncated at the first beta < 1e-15 as the reference's
// compute_tridiag / continued_fraction_coeffs (ftlm.cpp:24-53, dynamics.cpp:12-59)
struct TridiagOutput {
    std::vector<double> alphas, betas;
    double norm0 = 0.0;
};
TridiagOutput tridiag_run(const Engine& E, const double* start_host, bool start_cplx, int n_steps, hipStream_t s);
// real H, real device start vector (not modified); seconds: recurrence wall time
TridiagOutput tridiag_run_dev(const Engine& E, const double* start_dev, int n_steps, hipStream_t s,
                              double* seconds = nullptr);

// R independent recurrences on a row-major block (columns grouped `group`
// at a time into one complex column pair when group == 2)
struct BlockTridiag {
    int ncol = 0, group = 1;
    std::vector<std::vector<double>> alphas, betas;   // per group
    std::vector<double> norm0;                         // per group
};
// X0: device block start (nrows x ncol, row-major, ld = ncol); consumed.
BlockTridiag block_tridiag_run(const Engine& E, double* X0_dev, int ncol, int group, int n_steps, hipStream_t s,
                               double* seconds = nullptr);

// FTLM observables from tridiagonals (reference ftlm.cpp:122-170, for a grid
// of inverse temperatures); weights use the squared start-vector norms.
void ftlm_observables(const std::vector<std::vector<double>>& alphas, const std::vector<std::vector<double>>& betas,
                      const std::vector<double>& norm2, const std::vector<double>& betas_T, std::vector<double>& Z,
                      std::vector<double>& E, std::vector<double>& Cv);

// ---------------------------------------------------------------- Davidson
struct DavidsonOutput {
    std::vector<double> eigenvalues;
    std::vector<std::vector<double>> eigenvectors;   // interleaved complex
    int iterations = 0;
    bool converged = false;
};
DavidsonOutput davidson_run(const Engine& E, int n_eig, int max_subspace, double tol, bool want_vectors,
                            hipStream_t s);

double spectral_value(const double* alphas, const double* betas, size_t n, double norm_phi0, double omega, double E0,
                      double eta);

} // namespace qk
