// qkrylov.hpp -- C++ drop-in surface of the MI355X engine, mirroring the
// reference's public C++ names (namespace qkrylov::fp64):
//   Sector                       include/qkrylov/symmetry/sector.hpp:10-41
//   Basis + 4 kinds              include/qkrylov/basis/*.hpp
//   Site + 4 kinds, LocalAction  include/qkrylov/sites/*.hpp, operators/local_action.hpp
//   OperatorTerm/LocalOp/OpSum   include/qkrylov/operators/{operator_term,opsum,local_op}.hpp
//   Device                       include/qkrylov/core/device.hpp:27-84
//   MatrixFreeHamiltonian        include/qkrylov/hamiltonian/matrix_free_hamiltonian.hpp:24-77
//   lanczos_ground_state, davidson_lowest, continued_fraction_coeffs,
//   evaluate_spectral_function, ftlm   include/qkrylov/solvers/*.hpp
// The headers at the reference's paths (qkrylov/basis/spinhalf_basis.hpp,
// qkrylov/solvers/lanczos.hpp, ...) include this file, so code written
// against the reference compiles unchanged. Differences: the execution-space
// template parameter of MatrixFreeHamiltonian and the solvers is kept for
// that spelling but every space runs the one HIP engine; the Kokkos View
// overload of apply() becomes device-pointer overloads; qkrylov::fp32
// (qkrylov_fp32.hpp) has float scalars and vectors at its boundary over the
// same FP64 engine; batched entry points (ftlm_batched,
// continued_fraction_batched) are additions. QKRYLOV_PRECISION_NAMESPACE is
// fp64 with QKRYLOV_DOUBLE_PRECISION defined and fp32 otherwise, as in the
// reference (core/types.hpp:15-19).
#pragma once

#include <complex>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#ifndef QKRYLOV_PRECISION_NAMESPACE
#ifdef QKRYLOV_DOUBLE_PRECISION
#define QKRYLOV_PRECISION_NAMESPACE fp64
#else
#define QKRYLOV_PRECISION_NAMESPACE fp32
#endif
#endif

namespace qk {
class BasisSpec;
class Engine;
}

namespace qkrylov {

// Execution-space tag of the engine: HIP on the current device.
struct HipSpace {};

// Start vector of lanczos_ground_state when its options leave it to the
// process (start_mode -1, and the (H, maxiter, tol) form): 0 (default) the
// engine's counter RNG; 1 / 2 the reference's std::mt19937(seed) complex
// uniform [-1, 1) draws (src/solvers/lanczos.cpp:115-119), imaginary part
// drawn first (1, GCC's order) or real part first (2, Clang's), run on
// complex vectors. Throws std::invalid_argument for other values.
void set_default_start_mode(int mode);
int default_start_mode();

}  // namespace qkrylov

// Code written against the reference names its execution space with Kokkos
// types (MatrixFreeHamiltonian<Kokkos::DefaultExecutionSpace>,
// tests/test_workflow.cpp:36). Without Kokkos in the translation unit those
// names denote the engine's tag; with Kokkos included first, its own types
// are accepted as the (unused) template argument.
#ifndef KOKKOS_VERSION
namespace Kokkos {
using DefaultExecutionSpace = qkrylov::HipSpace;
using DefaultHostExecutionSpace = qkrylov::HipSpace;
}  // namespace Kokkos
#endif

namespace qkrylov {
namespace fp64 {

using Real = double;
using Complex = std::complex<double>;
using StateID = uint64_t;
using Index = std::size_t;
using HostVector = std::vector<Complex>;

inline int popcount(uint64_t x) noexcept { return __builtin_popcountll(x); }

struct Sector {
    bool use_sz = false;
    int sz2 = 0;
    bool use_nup = false;
    bool use_ndn = false;
    int nup = 0;
    int ndn = 0;
    bool use_n = false;
    int n = 0;
    bool use_nb = false;
    int nb = 0;
};

struct Device {
    int id = 0;
    Device() = default;
    explicit Device(int device_id) : id(device_id) {}
    explicit Device(const std::string& s);
    static bool is_gpu_build() { return true; }
    static std::string backend_name() { return "hip"; }
    static int gpu_count();
};

class Basis {
public:
    virtual ~Basis();
    virtual Index size() const;
    virtual StateID state(Index i) const;     // throws std::out_of_range
    virtual Index index(StateID s) const;     // throws std::runtime_error if absent
    virtual bool contains(StateID s) const;
    int nsites() const noexcept;
    const Sector& sector() const noexcept { return sector_; }
    std::shared_ptr<const qk::BasisSpec> spec() const { return spec_; }

protected:
    Basis(int kind, int N, const Sector& sector);
    std::shared_ptr<const qk::BasisSpec> spec_;
    Sector sector_;
};

class SpinHalfBasis : public Basis {
public:
    SpinHalfBasis(int N, const Sector& sector = Sector{});
};
class FermionBasis : public Basis {
public:
    FermionBasis(int N, const Sector& sector = Sector{});
};
class HubbardBasis : public Basis {
public:
    HubbardBasis(int N, const Sector& sector = Sector{});
};
class TJBasis : public Basis {
public:
    TJBasis(int N, const Sector& sector = Sector{});
};

struct LocalAction {
    bool valid = false;
    StateID new_state = 0;
    Complex matrix_element = 0.0;
};

class Site {
public:
    virtual ~Site() = default;
    virtual LocalAction apply(const std::string& op, int site, StateID state) const = 0;
    virtual int kind() const = 0;   // 0 spin-1/2, 1 fermion, 2 Hubbard, 3 t-J
};
class SpinHalfSite : public Site {
public:
    LocalAction apply(const std::string& op, int site, StateID state) const override;
    int kind() const override { return 0; }
};
class FermionSite : public Site {
public:
    LocalAction apply(const std::string& op, int site, StateID state) const override;
    int kind() const override { return 1; }
};
class HubbardSite : public Site {
public:
    LocalAction apply(const std::string& op, int site, StateID state) const override;
    int kind() const override { return 2; }
};
class TJSite : public Site {
public:
    LocalAction apply(const std::string& op, int site, StateID state) const override;
    int kind() const override { return 3; }
};

struct OperatorFactor {
    std::string op;
    int site = 0;
    OperatorFactor() = default;
    OperatorFactor(std::string op_, int site_) : op(std::move(op_)), site(site_) {}
};

struct OperatorTerm {
    Complex coeff = 1.0;
    std::vector<OperatorFactor> factors;
    OperatorTerm() = default;
    OperatorTerm(Complex c, std::initializer_list<OperatorFactor> f) : coeff(c), factors(f) {}
};

struct LocalOp {
    std::string name;
    int site = 0;
    LocalOp() = default;
    LocalOp(std::string n, int s) : name(std::move(n)), site(s) {}
};

inline OperatorTerm operator*(Complex c, const LocalOp& op)
{
    OperatorTerm t;
    t.coeff = c;
    t.factors.push_back({op.name, op.site});
    return t;
}
inline OperatorTerm operator*(Real c, const LocalOp& op) { return Complex(c, 0.0) * op; }
inline OperatorTerm operator*(const LocalOp& a, const LocalOp& b)
{
    OperatorTerm t;
    t.coeff = 1.0;
    t.factors.push_back({a.name, a.site});
    t.factors.push_back({b.name, b.site});
    return t;
}
inline OperatorTerm operator*(OperatorTerm t, const LocalOp& op)
{
    t.factors.push_back({op.name, op.site});
    return t;
}
inline OperatorTerm operator*(OperatorTerm a, const OperatorTerm& b)
{
    a.coeff *= b.coeff;
    a.factors.insert(a.factors.end(), b.factors.begin(), b.factors.end());
    return a;
}

struct OpSumExpr {
    std::vector<OperatorTerm> terms;
    OpSumExpr(const OperatorTerm& t) : terms({t}) {}
    OpSumExpr(const std::vector<OperatorTerm>& ts) : terms(ts) {}
};
inline OpSumExpr operator+(const OperatorTerm& a, const OperatorTerm& b) { return OpSumExpr({a, b}); }
inline OpSumExpr operator+(OpSumExpr e, const OperatorTerm& t)
{
    e.terms.push_back(t);
    return e;
}
inline OpSumExpr operator+(const OperatorTerm& t, OpSumExpr e)
{
    e.terms.insert(e.terms.begin(), t);
    return e;
}
inline OpSumExpr operator+(OpSumExpr a, const OpSumExpr& b)
{
    a.terms.insert(a.terms.end(), b.terms.begin(), b.terms.end());
    return a;
}
inline OpSumExpr operator-(const OperatorTerm& a, OperatorTerm b)
{
    b.coeff = -b.coeff;
    return OpSumExpr({a, b});
}
inline OpSumExpr operator-(OpSumExpr e, OperatorTerm t)
{
    t.coeff = -t.coeff;
    e.terms.push_back(t);
    return e;
}

class OpSum {
public:
    void add_term(const OperatorTerm& term) { terms_.push_back(term); }
    OpSum& operator+=(const OperatorTerm& t)
    {
        add_term(t);
        return *this;
    }
    OpSum& operator+=(const OpSumExpr& e)
    {
        for (const auto& t : e.terms) add_term(t);
        return *this;
    }
    void clear() { terms_.clear(); }
    std::size_t size() const { return terms_.size(); }
    const OperatorTerm& operator[](std::size_t i) const { return terms_[i]; }
    const std::vector<OperatorTerm>& terms() const { return terms_; }

private:
    std::vector<OperatorTerm> terms_;
};

inline LocalOp Sz(int i) { return {"Sz", i}; }
inline LocalOp Sp(int i) { return {"Sp", i}; }
inline LocalOp Sm(int i) { return {"Sm", i}; }
inline LocalOp Sx(int i) { return {"Sx", i}; }
inline LocalOp Sy(int i) { return {"Sy", i}; }
inline LocalOp CdagUp(int i) { return {"CdagUp", i}; }
inline LocalOp CUp(int i) { return {"CUp", i}; }
inline LocalOp CdagDn(int i) { return {"CdagDn", i}; }
inline LocalOp CDn(int i) { return {"CDn", i}; }
inline LocalOp Nup(int i) { return {"Nup", i}; }
inline LocalOp Ndn(int i) { return {"Ndn", i}; }
inline LocalOp Nupdn(int i) { return {"Nupdn", i}; }
inline LocalOp Bdag(int i) { return {"Bdag", i}; }
inline LocalOp B(int i) { return {"B", i}; }
inline LocalOp N(int i) { return {"N", i}; }

// The engine-backed operator; MatrixFreeHamiltonian<ExecSpace> below is the
// reference's spelling of it.
class MatrixFreeHamiltonianBase {
public:
    // split_bits: number of low state bits ranked per lane (-1: automatic)
    MatrixFreeHamiltonianBase(std::shared_ptr<Basis> basis, std::shared_ptr<Site> site, const OpSum& ops,
                              Device device = Device(), int split_bits = -1);
    virtual ~MatrixFreeHamiltonianBase();

    // host pointers: y = H x as the reference (y[a] = sum conj(amp) x[b])
    void apply(const Complex* x, Complex* y) const;
    // device pointers, real vectors; nvec > 1: row-major block x[row * nvec + c]
    void apply_device(const double* x, double* y, int nvec = 1, void* stream = nullptr) const;
    // device pointers, interleaved complex vectors
    void apply_device_complex(const double* x, double* y, void* stream = nullptr) const;

    HostVector diagonal_host() const;
    Index dimension() const;
    const Device& device() const { return device_; }
    bool is_real() const;
    std::shared_ptr<const qk::Engine> engine() const { return engine_; }
    std::shared_ptr<Basis> basis() const { return basis_; }

    // Times the apply (nvec columns) under candidate launch tunings and, when
    // allow_split, lo-bit splits around the current one, and keeps the
    // fastest. Launch tunings never change results; another split changes
    // the order in which a row's entries are added (the last bits).
    struct TuneReport {
        int split_bits, apply_waves, schedule, block_waves, block_schedule, candidates;
        double ms_before, ms_after;
        int block_rows;
    };
    TuneReport autotune(int nvec = 1, bool allow_split = true);

private:
    Device device_;
    std::shared_ptr<Basis> basis_;
    std::shared_ptr<Site> site_;
    OpSum ops_;
    std::shared_ptr<const qk::Engine> engine_;
};

// matrix_free_hamiltonian.hpp:24-25 of the reference: a class template over
// the execution space. Every space is served by the same engine.
template <class ExecSpace = HipSpace>
class MatrixFreeHamiltonian : public MatrixFreeHamiltonianBase {
public:
    MatrixFreeHamiltonian(std::shared_ptr<Basis> basis, std::shared_ptr<Site> site, const OpSum& ops,
                          Device device = Device(), int split_bits = -1)
        : MatrixFreeHamiltonianBase(std::move(basis), std::move(site), ops, device, split_bits)
    {
    }
};

struct LanczosResult {
    Real energy = 0.0;
    int iterations = 0;
    bool converged = false;
    HostVector eigenvector;
};

struct LanczosOptions {
    int maxiter = 200;
    Real tol = 1e-12;
    bool reorthogonalize = true;    // reference behaviour (lanczos.cpp:149-151)
    bool want_eigenvector = true;   // the reference always builds the Ritz vector
    uint64_t seed = 1234;
    int start_mode = -1;            // 0 counter RNG, 1 / 2 the reference's draws, -1 default_start_mode()
};

LanczosResult lanczos_ground_state(const MatrixFreeHamiltonianBase& H, int maxiter = 200, Real tol = 1.0e-12);
LanczosResult lanczos_ground_state(const MatrixFreeHamiltonianBase& H, const LanczosOptions& opt);

struct DavidsonResult {
    std::vector<Real> eigenvalues;
    std::vector<HostVector> eigenvectors;
    int iterations = 0;
    bool converged = false;
};
DavidsonResult davidson_lowest(const MatrixFreeHamiltonianBase& H, int n_eig = 1, int max_subspace = 20,
                               Real tol = 1.0e-8);

struct DynamicsResult {
    std::vector<Real> alphas;
    std::vector<Real> betas;
    Real norm_phi0 = 0.0;
};
DynamicsResult continued_fraction_coeffs(const MatrixFreeHamiltonianBase& H, const HostVector& phi0, int n_iter = 100);
Real evaluate_spectral_function(const Real* alphas, const Real* betas, size_t n, Real norm_phi0, Real omega, Real E0,
                                Real eta = 0.1);

struct FTLMResult {
    Real beta;
    Real partition_function = 0.0;
    Real internal_energy = 0.0;
    Real specific_heat = 0.0;
};
FTLMResult ftlm(const MatrixFreeHamiltonianBase& H, Real beta, int n_random = 50, int n_steps = 100);

// ---- additions: batched solves (one multi-vector recurrence for R vectors)
struct FTLMGridResult {
    std::vector<Real> betas, partition_function, internal_energy, specific_heat;
    std::vector<std::vector<Real>> alphas, betas_lanczos;   // per random vector
    double seconds = 0.0;                                   // recurrence wall time
};
// R real Rademacher (+-1) vectors from the counter RNG (seed, column, row);
// sequential = true runs them one at a time through the plain path.
FTLMGridResult ftlm_batched(const MatrixFreeHamiltonianBase& H, int R, int M, uint64_t seed,
                            const std::vector<Real>& betas, bool sequential = false);

// continued fractions for several seeds at once; seeds are device arrays,
// complex seeds occupy two real columns (group = 2)
std::vector<DynamicsResult> continued_fraction_batched(const MatrixFreeHamiltonianBase& H, const std::vector<HostVector>& seeds,
                                                       int n_iter, double* seconds = nullptr);

// Dynamical structure factor driver: for each operator O_q (built with
// operator semantics, y = O_q x, e.g. S^z_q = N^-1/2 sum_r e^{iq.r} S^z_r),
// the seed phi_q = O_q psi0 is formed on the device, all seeds run as one
// block continued fraction of n_iter steps (a complex seed is two real
// columns), and S_q(omega) = -1/pi Im <phi_q| (omega + E0 + i eta - H)^-1
// |phi_q> is evaluated on the grid. The reference leaves seed building to
// the caller (tests/test_dynamics.cpp:41-46) and provides the two halves as
// continued_fraction_coeffs / evaluate_spectral_function
// (solvers/dynamics.hpp:20-38).
struct StructureFactorResult {
    std::vector<DynamicsResult> cf;               // per operator
    std::vector<std::vector<Real>> spectrum;      // spectrum[q][w]
    double seconds = 0.0;                          // block recurrence wall time
};
StructureFactorResult structure_factor(const MatrixFreeHamiltonianBase& H,
                                       const std::vector<const MatrixFreeHamiltonianBase*>& operators,
                                       const HostVector& psi0, Real E0, int n_iter, const std::vector<Real>& omega,
                                       Real eta);

// The reference's solver templates (solvers/lanczos.hpp:20-24,
// davidson.hpp:21-27, dynamics.hpp:21-26, ftlm.hpp:21-27) with their
// defaults, forwarding to the engine-backed functions above.
template <class ExecSpace>
LanczosResult lanczos_ground_state(const MatrixFreeHamiltonian<ExecSpace>& H, int maxiter = 200, Real tol = 1.0e-12)
{
    return lanczos_ground_state(static_cast<const MatrixFreeHamiltonianBase&>(H), maxiter, tol);
}
template <class ExecSpace>
DavidsonResult davidson_lowest(const MatrixFreeHamiltonian<ExecSpace>& H, int n_eig = 1, int max_subspace = 20,
                               Real tol = 1.0e-8)
{
    return davidson_lowest(static_cast<const MatrixFreeHamiltonianBase&>(H), n_eig, max_subspace, tol);
}
template <class ExecSpace>
DynamicsResult continued_fraction_coeffs(const MatrixFreeHamiltonian<ExecSpace>& H, const HostVector& phi0,
                                         int n_iter = 100)
{
    return continued_fraction_coeffs(static_cast<const MatrixFreeHamiltonianBase&>(H), phi0, n_iter);
}
template <class ExecSpace>
FTLMResult ftlm(const MatrixFreeHamiltonian<ExecSpace>& H, Real beta, int n_random = 50, int n_steps = 100)
{
    return ftlm(static_cast<const MatrixFreeHamiltonianBase&>(H), beta, n_random, n_steps);
}

} // namespace fp64
} // namespace qkrylov

#include "qkrylov/qkrylov_fp32.hpp"
