// qkrylov_fp32.hpp -- the reference's single-precision C++ names
// (namespace qkrylov::fp32, include/qkrylov/core/types.hpp:15-44 of the
// reference: Real = float, Complex = std::complex<float>) over the FP64
// engine. Scalars, coefficients, vectors and results are float at this
// boundary; everything behind it (the device apply and the solvers) runs in
// FP64, and values are rounded to float on the way out. Included by
// qkrylov/qkrylov.hpp.
#pragma once

#include <algorithm>
#include <initializer_list>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace qkrylov {
namespace fp32 {

using Real = float;
using Complex = std::complex<float>;
using StateID = uint64_t;
using Index = std::size_t;
using HostVector = std::vector<Complex>;

using fp64::popcount;
using Sector = fp64::Sector;
using Device = fp64::Device;
using Basis = fp64::Basis;
using SpinHalfBasis = fp64::SpinHalfBasis;
using FermionBasis = fp64::FermionBasis;
using HubbardBasis = fp64::HubbardBasis;
using TJBasis = fp64::TJBasis;

namespace detail {
inline Complex narrow(const std::complex<double>& z) { return Complex((float)z.real(), (float)z.imag()); }
inline std::complex<double> widen(const Complex& z) { return std::complex<double>(z.real(), z.imag()); }
inline HostVector narrow(const fp64::HostVector& v)
{
    HostVector o(v.size());
    std::transform(v.begin(), v.end(), o.begin(), [](const std::complex<double>& z) { return narrow(z); });
    return o;
}
inline fp64::HostVector widen(const HostVector& v)
{
    fp64::HostVector o(v.size());
    std::transform(v.begin(), v.end(), o.begin(), [](const Complex& z) { return widen(z); });
    return o;
}
inline std::vector<Real> narrow(const std::vector<double>& v) { return std::vector<Real>(v.begin(), v.end()); }
} // namespace detail

// ---------------------------------------------------------------- sites
struct LocalAction {
    bool valid = false;
    StateID new_state = 0;
    Complex matrix_element = 0.0f;
};

class Site {
public:
    virtual ~Site() = default;
    virtual LocalAction apply(const std::string& op, int site, StateID state) const = 0;
    virtual int kind() const = 0;
    // the FP64 site the engine is built from
    virtual std::shared_ptr<fp64::Site> engine_site() const = 0;
};

template <class S64>
class SiteOf : public Site {
public:
    LocalAction apply(const std::string& op, int site, StateID state) const override
    {
        const fp64::LocalAction a = impl_.apply(op, site, state);
        return LocalAction{a.valid, a.new_state, detail::narrow(a.matrix_element)};
    }
    int kind() const override { return impl_.kind(); }
    std::shared_ptr<fp64::Site> engine_site() const override { return std::make_shared<S64>(); }

private:
    S64 impl_;
};
class SpinHalfSite : public SiteOf<fp64::SpinHalfSite> {};
class FermionSite : public SiteOf<fp64::FermionSite> {};
class HubbardSite : public SiteOf<fp64::HubbardSite> {};
class TJSite : public SiteOf<fp64::TJSite> {};

// ---------------------------------------------------------------- operators
using OperatorFactor = fp64::OperatorFactor;

struct OperatorTerm {
    Complex coeff = 1.0f;
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
inline OperatorTerm operator*(double c, const LocalOp& op) { return Complex((float)c, 0.0f) * op; }
inline OperatorTerm operator*(const LocalOp& a, const LocalOp& b)
{
    OperatorTerm t;
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
    // the same terms with FP64 coefficients (the float values, exactly)
    fp64::OpSum widened() const
    {
        fp64::OpSum o;
        for (const auto& t : terms_) {
            fp64::OperatorTerm w;
            w.coeff = detail::widen(t.coeff);
            w.factors = t.factors;
            o.add_term(w);
        }
        return o;
    }

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

// ---------------------------------------------------------------- Hamiltonian
class MatrixFreeHamiltonianBase {
public:
    MatrixFreeHamiltonianBase(std::shared_ptr<Basis> basis, std::shared_ptr<Site> site, const OpSum& ops,
                              Device device = Device(), int split_bits = -1)
        : impl_(std::move(basis), site ? site->engine_site() : nullptr, ops.widened(), device, split_bits)
    {
    }
    virtual ~MatrixFreeHamiltonianBase() = default;

    // host pointers, `dimension()` entries each
    void apply(const Complex* x, Complex* y) const
    {
        const Index n = dimension();
        fp64::HostVector xd(x, x + n), yd(n);
        impl_.apply(xd.data(), yd.data());
        std::transform(yd.begin(), yd.end(), y, [](const std::complex<double>& z) { return detail::narrow(z); });
    }
    HostVector diagonal_host() const { return detail::narrow(impl_.diagonal_host()); }
    Index dimension() const { return impl_.dimension(); }
    const Device& device() const { return impl_.device(); }
    bool is_real() const { return impl_.is_real(); }
    std::shared_ptr<Basis> basis() const { return impl_.basis(); }
    // the FP64 operator behind this one (device-pointer applies, tuning)
    const fp64::MatrixFreeHamiltonianBase& engine() const { return impl_; }
    fp64::MatrixFreeHamiltonianBase& engine() { return impl_; }

private:
    fp64::MatrixFreeHamiltonianBase impl_;
};

template <class ExecSpace = HipSpace>
class MatrixFreeHamiltonian : public MatrixFreeHamiltonianBase {
public:
    MatrixFreeHamiltonian(std::shared_ptr<Basis> basis, std::shared_ptr<Site> site, const OpSum& ops,
                          Device device = Device(), int split_bits = -1)
        : MatrixFreeHamiltonianBase(std::move(basis), std::move(site), ops, device, split_bits)
    {
    }
};

// ---------------------------------------------------------------- solvers
struct LanczosResult {
    Real energy = 0.0f;
    int iterations = 0;
    bool converged = false;
    HostVector eigenvector;
};
struct DavidsonResult {
    std::vector<Real> eigenvalues;
    std::vector<HostVector> eigenvectors;
    int iterations = 0;
    bool converged = false;
};
struct DynamicsResult {
    std::vector<Real> alphas;
    std::vector<Real> betas;
    Real norm_phi0 = 0.0f;
};
struct FTLMResult {
    Real beta;
    Real partition_function = 0.0f;
    Real internal_energy = 0.0f;
    Real specific_heat = 0.0f;
};

inline LanczosResult lanczos_ground_state(const MatrixFreeHamiltonianBase& H, int maxiter = 200, Real tol = 1.0e-12f)
{
    const fp64::LanczosResult r = fp64::lanczos_ground_state(H.engine(), maxiter, tol);
    return LanczosResult{(Real)r.energy, r.iterations, r.converged, detail::narrow(r.eigenvector)};
}
inline DavidsonResult davidson_lowest(const MatrixFreeHamiltonianBase& H, int n_eig = 1, int max_subspace = 20,
                                      Real tol = 1.0e-8f)
{
    const fp64::DavidsonResult r = fp64::davidson_lowest(H.engine(), n_eig, max_subspace, tol);
    DavidsonResult o;
    o.eigenvalues = detail::narrow(r.eigenvalues);
    for (const auto& v : r.eigenvectors) o.eigenvectors.push_back(detail::narrow(v));
    o.iterations = r.iterations;
    o.converged = r.converged;
    return o;
}
inline DynamicsResult continued_fraction_coeffs(const MatrixFreeHamiltonianBase& H, const HostVector& phi0,
                                                int n_iter = 100)
{
    const fp64::DynamicsResult r = fp64::continued_fraction_coeffs(H.engine(), detail::widen(phi0), n_iter);
    return DynamicsResult{detail::narrow(r.alphas), detail::narrow(r.betas), (Real)r.norm_phi0};
}
inline Real evaluate_spectral_function(const Real* alphas, const Real* betas, size_t n, Real norm_phi0, Real omega,
                                       Real E0, Real eta = 0.1f)
{
    const std::vector<double> a(alphas, alphas + n), b(betas, betas + (n ? n - 1 : 0));
    return (Real)fp64::evaluate_spectral_function(a.data(), b.empty() ? nullptr : b.data(), n, norm_phi0, omega, E0,
                                                  eta);
}
inline FTLMResult ftlm(const MatrixFreeHamiltonianBase& H, Real beta, int n_random = 50, int n_steps = 100)
{
    const fp64::FTLMResult r = fp64::ftlm(H.engine(), beta, n_random, n_steps);
    return FTLMResult{(Real)r.beta, (Real)r.partition_function, (Real)r.internal_energy, (Real)r.specific_heat};
}

// the solver templates of the reference, as in namespace fp64
template <class ExecSpace>
LanczosResult lanczos_ground_state(const MatrixFreeHamiltonian<ExecSpace>& H, int maxiter = 200, Real tol = 1.0e-12f)
{
    return lanczos_ground_state(static_cast<const MatrixFreeHamiltonianBase&>(H), maxiter, tol);
}
template <class ExecSpace>
DavidsonResult davidson_lowest(const MatrixFreeHamiltonian<ExecSpace>& H, int n_eig = 1, int max_subspace = 20,
                               Real tol = 1.0e-8f)
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

} // namespace fp32
} // namespace qkrylov
