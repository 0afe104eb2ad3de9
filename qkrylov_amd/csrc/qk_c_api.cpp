// qk_c_api.cpp -- the extern "C" boundary. Reference ABI semantics
// (src/c_api.cpp:56-624): NULL handle -> -1, exceptions -> -2 (creators
// return NULL), *_destroy(NULL) is a no-op, basis and site handles are shared
// by the Hamiltonian, the OpSum is copied into it.
#include "qkrylov/c_api.h"
#include "qkrylov/qkrylov.hpp"
#include "qkrylov_amd.h"

#include "qk_eigen.hpp"
#include "qk_engine.hpp"
#include "qk_split.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

using namespace qkrylov::fp64;

struct qkrylov_sector_t {
    Sector sector;
};
struct qkrylov_basis_t {
    std::shared_ptr<Basis> ptr;
};
struct qkrylov_site_t {
    std::shared_ptr<Site> ptr;
};
struct qkrylov_opsum_t {
    OpSum opsum;
};
struct qkrylov_hamiltonian_t {
    std::unique_ptr<MatrixFreeHamiltonianBase> ptr;
};

namespace {

thread_local std::string g_last_error;

int fail(const std::exception& e)
{
    g_last_error = e.what();
    return QKRYLOV_ERROR_EXCEPTION;
}
int fail_unknown()
{
    g_last_error = "unknown exception";
    return QKRYLOV_ERROR_EXCEPTION;
}
int invalid(const char* what)
{
    g_last_error = what;
    return QKRYLOV_ERROR_INVALID_ARG;
}

#define QK_TRY try {
#define QK_CATCH                                     \
    }                                                \
    catch (const std::exception& e) { return fail(e); } \
    catch (...) { return fail_unknown(); }

template <class T, class F>
T* make_handle(F&& f)
{
    try {
        return f();
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return nullptr;
    } catch (...) {
        g_last_error = "unknown exception";
        return nullptr;
    }
}

std::string adjoint_name(const std::string& op)
{
    if (op == "Sp") return "Sm";
    if (op == "Sm") return "Sp";
    if (op == "C") return "Cdag";
    if (op == "Cdag") return "C";
    if (op == "CUp") return "CdagUp";
    if (op == "CdagUp") return "CUp";
    if (op == "CDn") return "CdagDn";
    if (op == "CdagDn") return "CDn";
    return op;   // Sz, Sx, Sy, N, Id, Nup, Ndn, Nupdn are Hermitian
}

OpSum adjoint(const OpSum& o)
{
    OpSum a;
    for (const auto& t : o.terms()) {
        OperatorTerm at;
        at.coeff = std::conj(t.coeff);
        for (auto it = t.factors.rbegin(); it != t.factors.rend(); ++it) at.factors.push_back({adjoint_name(it->op), it->site});
        a.add_term(at);
    }
    return a;
}

const MatrixFreeHamiltonianBase* H_of(qkrylov_hamiltonian_h h) { return (h && h->ptr) ? h->ptr.get() : nullptr; }

} // namespace

extern "C" {

// ---------------------------------------------------------------- sector
qkrylov_sector_h qkrylov_sector_create(void)
{
    return make_handle<qkrylov_sector_t>([] { return new qkrylov_sector_t(); });
}
void qkrylov_sector_destroy(qkrylov_sector_h s) { delete s; }
int qkrylov_sector_set_sz(qkrylov_sector_h s, int sz2)
{
    if (!s) return QKRYLOV_ERROR_INVALID_ARG;
    s->sector.use_sz = true;
    s->sector.sz2 = sz2;
    return QKRYLOV_SUCCESS;
}
int qkrylov_sector_set_hubbard_particles(qkrylov_sector_h s, int nup, int ndn)
{
    if (!s) return QKRYLOV_ERROR_INVALID_ARG;
    s->sector.use_nup = s->sector.use_ndn = true;
    s->sector.nup = nup;
    s->sector.ndn = ndn;
    return QKRYLOV_SUCCESS;
}
int qkrylov_amd_sector_set_species(qkrylov_sector_h s, int species, int count)
{
    if (!s || species < 0 || species > 1) return QKRYLOV_ERROR_INVALID_ARG;
    if (species == 0) {
        s->sector.use_nup = true;
        s->sector.nup = count;
    } else {
        s->sector.use_ndn = true;
        s->sector.ndn = count;
    }
    return QKRYLOV_SUCCESS;
}
int qkrylov_sector_set_n(qkrylov_sector_h s, int n)
{
    if (!s) return QKRYLOV_ERROR_INVALID_ARG;
    s->sector.use_n = true;
    s->sector.n = n;
    return QKRYLOV_SUCCESS;
}
int qkrylov_sector_set_nb(qkrylov_sector_h s, int nb)
{
    if (!s) return QKRYLOV_ERROR_INVALID_ARG;
    s->sector.use_nb = true;
    s->sector.nb = nb;
    return QKRYLOV_SUCCESS;
}
int qkrylov_sector_get_sz(qkrylov_sector_h s, int* v, int* a)
{
    if (!s) return QKRYLOV_ERROR_INVALID_ARG;
    if (a) *a = s->sector.use_sz ? 1 : 0;
    if (v) *v = s->sector.sz2;
    return QKRYLOV_SUCCESS;
}
int qkrylov_sector_get_hubbard_particles(qkrylov_sector_h s, int* nu, int* nd, int* a)
{
    if (!s) return QKRYLOV_ERROR_INVALID_ARG;
    if (a) *a = (s->sector.use_nup && s->sector.use_ndn) ? 1 : 0;
    if (nu) *nu = s->sector.nup;
    if (nd) *nd = s->sector.ndn;
    return QKRYLOV_SUCCESS;
}
int qkrylov_sector_get_n(qkrylov_sector_h s, int* v, int* a)
{
    if (!s) return QKRYLOV_ERROR_INVALID_ARG;
    if (a) *a = s->sector.use_n ? 1 : 0;
    if (v) *v = s->sector.n;
    return QKRYLOV_SUCCESS;
}
int qkrylov_sector_get_nb(qkrylov_sector_h s, int* v, int* a)
{
    if (!s) return QKRYLOV_ERROR_INVALID_ARG;
    if (a) *a = s->sector.use_nb ? 1 : 0;
    if (v) *v = s->sector.nb;
    return QKRYLOV_SUCCESS;
}

// ---------------------------------------------------------------- basis
#define QK_BASIS_CREATE(fname, Cls)                                              \
    qkrylov_basis_h fname(int n, qkrylov_sector_h sec)                           \
    {                                                                            \
        return make_handle<qkrylov_basis_t>([&] {                                \
            auto h = std::make_unique<qkrylov_basis_t>();                        \
            h->ptr = std::make_shared<Cls>(n, sec ? sec->sector : Sector());     \
            return h.release();                                                  \
        });                                                                      \
    }
QK_BASIS_CREATE(qkrylov_spinhalf_basis_create, SpinHalfBasis)
QK_BASIS_CREATE(qkrylov_fermion_basis_create, FermionBasis)
QK_BASIS_CREATE(qkrylov_hubbard_basis_create, HubbardBasis)
QK_BASIS_CREATE(qkrylov_tj_basis_create, TJBasis)
#undef QK_BASIS_CREATE

void qkrylov_basis_destroy(qkrylov_basis_h b) { delete b; }
uint64_t qkrylov_basis_dimension(qkrylov_basis_h b) { return (b && b->ptr) ? (uint64_t)b->ptr->size() : 0; }
int qkrylov_basis_nsites(qkrylov_basis_h b) { return (b && b->ptr) ? b->ptr->nsites() : 0; }
uint64_t qkrylov_basis_state(qkrylov_basis_h b, uint64_t i)
{
    if (!b || !b->ptr) return 0;
    try {
        return b->ptr->state((Index)i);
    } catch (...) {
        return 0;
    }
}
int64_t qkrylov_basis_index(qkrylov_basis_h b, uint64_t s)
{
    if (!b || !b->ptr) return -1;
    if (!b->ptr->contains(s)) return -1;
    return (int64_t)b->ptr->index(s);
}
int qkrylov_basis_contains(qkrylov_basis_h b, uint64_t s) { return (b && b->ptr && b->ptr->contains(s)) ? 1 : 0; }

// ---------------------------------------------------------------- site
#define QK_SITE_CREATE(fname, Cls)                             \
    qkrylov_site_h fname(void)                                 \
    {                                                          \
        return make_handle<qkrylov_site_t>([] {                \
            auto h = std::make_unique<qkrylov_site_t>();       \
            h->ptr = std::make_shared<Cls>();                  \
            return h.release();                                \
        });                                                    \
    }
QK_SITE_CREATE(qkrylov_spinhalf_site_create, SpinHalfSite)
QK_SITE_CREATE(qkrylov_fermion_site_create, FermionSite)
QK_SITE_CREATE(qkrylov_hubbard_site_create, HubbardSite)
QK_SITE_CREATE(qkrylov_tj_site_create, TJSite)
#undef QK_SITE_CREATE
void qkrylov_site_destroy(qkrylov_site_h s) { delete s; }

// ---------------------------------------------------------------- opsum
qkrylov_opsum_h qkrylov_opsum_create(void)
{
    return make_handle<qkrylov_opsum_t>([] { return new qkrylov_opsum_t(); });
}
void qkrylov_opsum_destroy(qkrylov_opsum_h o) { delete o; }
int qkrylov_opsum_clear(qkrylov_opsum_h o)
{
    if (!o) return QKRYLOV_ERROR_INVALID_ARG;
    o->opsum.clear();
    return QKRYLOV_SUCCESS;
}
int qkrylov_opsum_add_term_nbody_f64(qkrylov_opsum_h o, double re, double im, int nf, const char** ops,
                                     const int* sites)
{
    if (!o || !ops || !sites || nf <= 0) return QKRYLOV_ERROR_INVALID_ARG;
    QK_TRY
    OperatorTerm t;
    t.coeff = Complex(re, im);
    for (int i = 0; i < nf; ++i) {
        if (!ops[i]) return QKRYLOV_ERROR_INVALID_ARG;
        t.factors.push_back({std::string(ops[i]), sites[i]});
    }
    o->opsum.add_term(t);
    return QKRYLOV_SUCCESS;
    QK_CATCH
}
int qkrylov_opsum_add_term_1body(qkrylov_opsum_h o, float re, float im, const char* op1, int s1)
{
    if (!o || !op1) return QKRYLOV_ERROR_INVALID_ARG;
    const char* ops[1] = {op1};
    const int sites[1] = {s1};
    return qkrylov_opsum_add_term_nbody_f64(o, (double)re, (double)im, 1, ops, sites);
}
int qkrylov_opsum_add_term_2body(qkrylov_opsum_h o, float re, float im, const char* op1, int s1, const char* op2,
                                 int s2)
{
    if (!o || !op1 || !op2) return QKRYLOV_ERROR_INVALID_ARG;
    const char* ops[2] = {op1, op2};
    const int sites[2] = {s1, s2};
    return qkrylov_opsum_add_term_nbody_f64(o, (double)re, (double)im, 2, ops, sites);
}
int qkrylov_opsum_add_term_nbody(qkrylov_opsum_h o, float re, float im, int nf, const char** ops, const int* sites)
{
    return qkrylov_opsum_add_term_nbody_f64(o, (double)re, (double)im, nf, ops, sites);
}

// ---------------------------------------------------------------- hamiltonian
qkrylov_hamiltonian_h qkrylov_hamiltonian_create(qkrylov_basis_h b, qkrylov_site_h s, qkrylov_opsum_h o)
{
    if (!b || !b->ptr || !s || !s->ptr || !o) return nullptr;
    return make_handle<qkrylov_hamiltonian_t>([&] {
        auto h = std::make_unique<qkrylov_hamiltonian_t>();
        h->ptr = std::make_unique<MatrixFreeHamiltonianBase>(b->ptr, s->ptr, o->opsum, Device());
        return h.release();
    });
}
void qkrylov_hamiltonian_destroy(qkrylov_hamiltonian_h h) { delete h; }
uint64_t qkrylov_hamiltonian_dimension(qkrylov_hamiltonian_h h)
{
    const auto* H = H_of(h);
    return H ? (uint64_t)H->dimension() : 0;
}

int qkrylov_hamiltonian_apply_f64(qkrylov_hamiltonian_h h, const double* xr, const double* xi, double* yr, double* yi)
{
    const auto* H = H_of(h);
    if (!H || !xr || !yr) return QKRYLOV_ERROR_INVALID_ARG;
    QK_TRY
    const size_t D = H->dimension();
    std::vector<Complex> x(D), y(D);
    for (size_t i = 0; i < D; ++i) x[i] = Complex(xr[i], xi ? xi[i] : 0.0);
    H->apply(x.data(), y.data());
    for (size_t i = 0; i < D; ++i) {
        yr[i] = y[i].real();
        if (yi) yi[i] = y[i].imag();
    }
    return QKRYLOV_SUCCESS;
    QK_CATCH
}
int qkrylov_hamiltonian_apply(qkrylov_hamiltonian_h h, const float* xr, const float* xi, float* yr, float* yi)
{
    const auto* H = H_of(h);
    if (!H || !xr || !yr) return QKRYLOV_ERROR_INVALID_ARG;
    QK_TRY
    const size_t D = H->dimension();
    std::vector<Complex> x(D), y(D);
    for (size_t i = 0; i < D; ++i) x[i] = Complex((double)xr[i], xi ? (double)xi[i] : 0.0);
    H->apply(x.data(), y.data());
    for (size_t i = 0; i < D; ++i) {
        yr[i] = (float)y[i].real();
        if (yi) yi[i] = (float)y[i].imag();
    }
    return QKRYLOV_SUCCESS;
    QK_CATCH
}
int qkrylov_hamiltonian_apply_complex_f64(qkrylov_hamiltonian_h h, const double* x, double* y)
{
    const auto* H = H_of(h);
    if (!H || !x || !y) return QKRYLOV_ERROR_INVALID_ARG;
    QK_TRY
    H->apply(reinterpret_cast<const Complex*>(x), reinterpret_cast<Complex*>(y));
    return QKRYLOV_SUCCESS;
    QK_CATCH
}
int qkrylov_hamiltonian_apply_complex(qkrylov_hamiltonian_h h, const float* x, float* y)
{
    const auto* H = H_of(h);
    if (!H || !x || !y) return QKRYLOV_ERROR_INVALID_ARG;
    QK_TRY
    const size_t D = H->dimension();
    std::vector<Complex> xc(D), yc(D);
    for (size_t i = 0; i < D; ++i) xc[i] = Complex((double)x[2 * i], (double)x[2 * i + 1]);
    H->apply(xc.data(), yc.data());
    for (size_t i = 0; i < D; ++i) {
        y[2 * i] = (float)yc[i].real();
        y[2 * i + 1] = (float)yc[i].imag();
    }
    return QKRYLOV_SUCCESS;
    QK_CATCH
}
int qkrylov_hamiltonian_diagonal_f64(qkrylov_hamiltonian_h h, double* d)
{
    const auto* H = H_of(h);
    if (!H || !d) return QKRYLOV_ERROR_INVALID_ARG;
    QK_TRY
    const auto v = H->diagonal_host();
    for (size_t i = 0; i < v.size(); ++i) d[i] = v[i].real();
    return QKRYLOV_SUCCESS;
    QK_CATCH
}
int qkrylov_hamiltonian_diagonal(qkrylov_hamiltonian_h h, float* d)
{
    const auto* H = H_of(h);
    if (!H || !d) return QKRYLOV_ERROR_INVALID_ARG;
    QK_TRY
    const auto v = H->diagonal_host();
    for (size_t i = 0; i < v.size(); ++i) d[i] = (float)v[i].real();
    return QKRYLOV_SUCCESS;
    QK_CATCH
}

// ---------------------------------------------------------------- solvers (FP64)
int qkrylov_lanczos_ground_state_f64(qkrylov_hamiltonian_h h, int maxiter, double tol,
                                     qkrylov_lanczos_result_f64_t* res, double* evec)
{
    const auto* H = H_of(h);
    if (!H || !res) return QKRYLOV_ERROR_INVALID_ARG;
    QK_TRY
    LanczosOptions o;
    o.maxiter = maxiter;
    o.tol = tol;
    o.want_eigenvector = evec != nullptr;
    auto r = lanczos_ground_state(*H, o);
    res->energy = r.energy;
    res->iterations = r.iterations;
    res->converged = r.converged ? 1 : 0;
    if (evec)
        for (size_t i = 0; i < r.eigenvector.size(); ++i) {
            evec[2 * i] = r.eigenvector[i].real();
            evec[2 * i + 1] = r.eigenvector[i].imag();
        }
    return QKRYLOV_SUCCESS;
    QK_CATCH
}
int qkrylov_lanczos_ground_state_complex(qkrylov_hamiltonian_h h, int maxiter, float tol,
                                         qkrylov_lanczos_result_c_t* res, float* evec)
{
    const auto* H = H_of(h);
    if (!H || !res) return QKRYLOV_ERROR_INVALID_ARG;
    QK_TRY
    LanczosOptions o;
    o.maxiter = maxiter;
    o.tol = (double)tol;
    o.want_eigenvector = evec != nullptr;
    auto r = lanczos_ground_state(*H, o);
    res->energy = (float)r.energy;
    res->iterations = r.iterations;
    res->converged = r.converged ? 1 : 0;
    if (evec)
        for (size_t i = 0; i < r.eigenvector.size(); ++i) {
            evec[2 * i] = (float)r.eigenvector[i].real();
            evec[2 * i + 1] = (float)r.eigenvector[i].imag();
        }
    return QKRYLOV_SUCCESS;
    QK_CATCH
}
int qkrylov_lanczos_ground_state(qkrylov_hamiltonian_h h, int maxiter, float tol, qkrylov_lanczos_result_c_t* res)
{
    return qkrylov_lanczos_ground_state_complex(h, maxiter, tol, res, nullptr);
}

int qkrylov_davidson_lowest_complex_f64(qkrylov_hamiltonian_h h, int n_eig, int max_subspace, double tol,
                                        double* evals, double* evecs, qkrylov_davidson_result_f64_t* info)
{
    const auto* H = H_of(h);
    if (!H || !evals || n_eig <= 0) return QKRYLOV_ERROR_INVALID_ARG;
    QK_TRY
    auto r = davidson_lowest(*H, n_eig, max_subspace, tol);
    const size_t k = std::min((size_t)n_eig, r.eigenvalues.size());
    for (size_t i = 0; i < k; ++i) evals[i] = r.eigenvalues[i];
    if (info) {
        info->iterations = r.iterations;
        info->converged = r.converged ? 1 : 0;
    }
    if (evecs) {
        const size_t D = H->dimension();
        for (size_t v = 0; v < k && v < r.eigenvectors.size(); ++v)
            for (size_t i = 0; i < D; ++i) {
                evecs[v * 2 * D + 2 * i] = r.eigenvectors[v][i].real();
                evecs[v * 2 * D + 2 * i + 1] = r.eigenvectors[v][i].imag();
            }
    }
    return QKRYLOV_SUCCESS;
    QK_CATCH
}
int qkrylov_davidson_lowest_complex(qkrylov_hamiltonian_h h, int n_eig, int max_subspace, float tol, float* evals,
                                    float* evecs, qkrylov_davidson_result_c_t* info)
{
    const auto* H = H_of(h);
    if (!H || !evals || n_eig <= 0) return QKRYLOV_ERROR_INVALID_ARG;
    QK_TRY
    auto r = davidson_lowest(*H, n_eig, max_subspace, (double)tol);
    const size_t k = std::min((size_t)n_eig, r.eigenvalues.size());
    for (size_t i = 0; i < k; ++i) evals[i] = (float)r.eigenvalues[i];
    if (info) {
        info->iterations = r.iterations;
        info->converged = r.converged ? 1 : 0;
    }
    if (evecs) {
        const size_t D = H->dimension();
        for (size_t v = 0; v < k && v < r.eigenvectors.size(); ++v)
            for (size_t i = 0; i < D; ++i) {
                evecs[v * 2 * D + 2 * i] = (float)r.eigenvectors[v][i].real();
                evecs[v * 2 * D + 2 * i + 1] = (float)r.eigenvectors[v][i].imag();
            }
    }
    return QKRYLOV_SUCCESS;
    QK_CATCH
}

int qkrylov_continued_fraction_coeffs_complex_f64(qkrylov_hamiltonian_h h, const double* phi0, int n_iter,
                                                  double* al, double* be, double* nrm, int* ncoef)
{
    const auto* H = H_of(h);
    if (!H || !phi0 || !al || !be || !nrm || n_iter <= 0) return QKRYLOV_ERROR_INVALID_ARG;
    QK_TRY
    const size_t D = H->dimension();
    HostVector p(D);
    for (size_t i = 0; i < D; ++i) p[i] = Complex(phi0[2 * i], phi0[2 * i + 1]);
    auto r = continued_fraction_coeffs(*H, p, n_iter);
    *nrm = r.norm_phi0;
    for (size_t i = 0; i < r.alphas.size() && i < (size_t)n_iter; ++i) al[i] = r.alphas[i];
    for (size_t i = 0; i < r.betas.size() && i < (size_t)n_iter; ++i) be[i] = r.betas[i];
    if (ncoef) *ncoef = (int)r.alphas.size();
    return QKRYLOV_SUCCESS;
    QK_CATCH
}
int qkrylov_continued_fraction_coeffs_complex(qkrylov_hamiltonian_h h, const float* phi0, int n_iter, float* al,
                                              float* be, float* nrm, int* ncoef)
{
    const auto* H = H_of(h);
    if (!H || !phi0 || !al || !be || !nrm || n_iter <= 0) return QKRYLOV_ERROR_INVALID_ARG;
    QK_TRY
    const size_t D = H->dimension();
    HostVector p(D);
    for (size_t i = 0; i < D; ++i) p[i] = Complex((double)phi0[2 * i], (double)phi0[2 * i + 1]);
    auto r = continued_fraction_coeffs(*H, p, n_iter);
    *nrm = (float)r.norm_phi0;
    for (size_t i = 0; i < r.alphas.size() && i < (size_t)n_iter; ++i) al[i] = (float)r.alphas[i];
    for (size_t i = 0; i < r.betas.size() && i < (size_t)n_iter; ++i) be[i] = (float)r.betas[i];
    if (ncoef) *ncoef = (int)r.alphas.size();
    return QKRYLOV_SUCCESS;
    QK_CATCH
}

double qkrylov_evaluate_spectral_function_f64(const double* al, const double* be, size_t n, double nrm, double omega,
                                              double E0, double eta)
{
    if (!al || !be || n == 0) return 0.0;
    return evaluate_spectral_function(al, be, n, nrm, omega, E0, eta);
}
float qkrylov_evaluate_spectral_function(const float* al, const float* be, size_t n, float nrm, float omega, float E0,
                                         float eta)
{
    if (!al || !be || n == 0) return 0.0f;
    std::vector<double> a(n), b(n > 1 ? n - 1 : 1, 0.0);
    for (size_t i = 0; i < n; ++i) a[i] = al[i];
    for (size_t i = 0; i + 1 < n; ++i) b[i] = be[i];
    return (float)evaluate_spectral_function(a.data(), b.data(), n, nrm, omega, E0, eta);
}

int qkrylov_ftlm_f64(qkrylov_hamiltonian_h h, double beta, int n_random, int n_steps, qkrylov_ftlm_result_f64_t* res)
{
    const auto* H = H_of(h);
    if (!H || !res) return QKRYLOV_ERROR_INVALID_ARG;
    QK_TRY
    auto r = ftlm(*H, beta, n_random, n_steps);
    res->beta = r.beta;
    res->partition_function = r.partition_function;
    res->internal_energy = r.internal_energy;
    res->specific_heat = r.specific_heat;
    return QKRYLOV_SUCCESS;
    QK_CATCH
}
int qkrylov_ftlm(qkrylov_hamiltonian_h h, float beta, int n_random, int n_steps, qkrylov_ftlm_result_c_t* res)
{
    if (!res) return QKRYLOV_ERROR_INVALID_ARG;
    qkrylov_ftlm_result_f64_t r{};
    const int rc = qkrylov_ftlm_f64(h, (double)beta, n_random, n_steps, &r);
    if (rc != QKRYLOV_SUCCESS) return rc;
    res->beta = (float)r.beta;
    res->partition_function = (float)r.partition_function;
    res->internal_energy = (float)r.internal_energy;
    res->specific_heat = (float)r.specific_heat;
    return QKRYLOV_SUCCESS;
}

// ---------------------------------------------------------------- extensions
const char* qkrylov_amd_last_error(void) { return g_last_error.c_str(); }

} // extern "C"

namespace qk {
void set_last_error(const char* msg) { g_last_error = msg ? msg : ""; }
} // namespace qk

extern "C" {


int qkrylov_amd_device_count(void) { return Device::gpu_count(); }
int qkrylov_amd_set_start_mode(int mode)
{
    if (mode < 0 || mode > 2) {
        g_last_error = "start mode must be 0, 1 or 2";
        return QKRYLOV_ERROR_INVALID_ARG;
    }
    qkrylov::set_default_start_mode(mode);
    return QKRYLOV_SUCCESS;
}
int qkrylov_amd_get_start_mode(void) { return qkrylov::default_start_mode(); }
int qkrylov_amd_reference_start_vector(size_t n, uint64_t seed, int mode, double* out)
{
    if ((n > 0 && !out) || (mode != 1 && mode != 2)) {
        g_last_error = "reference start vector: mode must be 1 or 2 and out non-null";
        return QKRYLOV_ERROR_INVALID_ARG;
    }
    QK_TRY
    qk::reference_start_vector(n, seed, mode, out);
    return QKRYLOV_SUCCESS;
    QK_CATCH
}

int qkrylov_amd_reference_normal_vector(size_t n, uint64_t seed, int mode, double* out)
{
    if ((n > 0 && !out) || (mode != 1 && mode != 2)) {
        g_last_error = "reference normal vector: mode must be 1 or 2 and out non-null";
        return QKRYLOV_ERROR_INVALID_ARG;
    }
    QK_TRY
    qk::ReferenceNormal(seed, mode).fill(n, out);
    return QKRYLOV_SUCCESS;
    QK_CATCH
}

int qkrylov_amd_set_device(int id)
{
    QK_TRY
    qk::ensure_device(id);
    return QKRYLOV_SUCCESS;
    QK_CATCH
}
void* qkrylov_amd_malloc(size_t bytes)
{
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        g_last_error = "hipMalloc failed";
        return nullptr;
    }
    return p;
}
int qkrylov_amd_free(void* p)
{
    if (!p) return QKRYLOV_SUCCESS;
    return hipFree(p) == hipSuccess ? QKRYLOV_SUCCESS : QKRYLOV_ERROR_EXCEPTION;
}
int qkrylov_amd_memcpy_h2d(void* d, const void* s, size_t n)
{
    if ((!d || !s) && n) return QKRYLOV_ERROR_INVALID_ARG;
    return hipMemcpy(d, s, n, hipMemcpyHostToDevice) == hipSuccess ? QKRYLOV_SUCCESS : QKRYLOV_ERROR_EXCEPTION;
}
int qkrylov_amd_memcpy_d2h(void* d, const void* s, size_t n)
{
    if ((!d || !s) && n) return QKRYLOV_ERROR_INVALID_ARG;
    return hipMemcpy(d, s, n, hipMemcpyDeviceToHost) == hipSuccess ? QKRYLOV_SUCCESS : QKRYLOV_ERROR_EXCEPTION;
}
int qkrylov_amd_synchronize(void)
{
    return hipDeviceSynchronize() == hipSuccess ? QKRYLOV_SUCCESS : QKRYLOV_ERROR_EXCEPTION;
}
int qkrylov_amd_copy(double* dst, const double* src, size_t n)
{
    if ((!dst || !src) && n) return QKRYLOV_ERROR_INVALID_ARG;
    QK_TRY
    hipStream_t s = qk::default_stream();
    if (qkl_copy(n, src, dst, s) != 0) throw std::runtime_error("copy launch failed");
    qk::hip_check(hipStreamSynchronize(s), "sync");
    return QKRYLOV_SUCCESS;
    QK_CATCH
}
int qkrylov_amd_gather_probe(size_t array_bytes, int bytes_per_lane, int reps, double* ms_per_launch,
                             double* gathers_per_launch)
{
    if (array_bytes < 2048 || (bytes_per_lane != 8 && bytes_per_lane != 16) || reps <= 0 || !ms_per_launch ||
        !gathers_per_launch)
        return QKRYLOV_ERROR_INVALID_ARG;
    QK_TRY
    hipStream_t s = qk::default_stream();
    const size_t n = array_bytes / sizeof(double);
    const int blocks = 2048, tiles = 64;
    qk::DevBuf x(n * sizeof(double)), y((size_t)blocks * 512 * sizeof(double));
    if (qkl_fill_zero(n, x.d(), s) != 0) throw std::runtime_error("fill launch failed");
    hipEvent_t a = nullptr, b = nullptr;
    qk::hip_check(hipEventCreate(&a), "event");
    qk::hip_check(hipEventCreate(&b), "event");
    struct Ev {
        hipEvent_t a, b;
        ~Ev()
        {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } ev{a, b};
    if (qkl_gather_probe(x.d(), n, bytes_per_lane, y.d(), blocks, tiles, s) != 0)
        throw std::runtime_error("probe launch failed");
    qk::hip_check(hipEventRecord(a, s), "event");
    for (int i = 0; i < reps; ++i)
        if (qkl_gather_probe(x.d(), n, bytes_per_lane, y.d(), blocks, tiles, s) != 0)
            throw std::runtime_error("probe launch failed");
    qk::hip_check(hipEventRecord(b, s), "event");
    qk::hip_check(hipEventSynchronize(b), "event sync");
    float ms = 0.0f;
    qk::hip_check(hipEventElapsedTime(&ms, a, b), "elapsed");
    *ms_per_launch = (double)ms / reps;
    *gathers_per_launch = (double)blocks * 8 * tiles * 16;
    return QKRYLOV_SUCCESS;
    QK_CATCH
}

int qkrylov_amd_fill_random(double* x, size_t n, uint64_t seed, int kind)
{
    if (!x || (kind != 0 && kind != 1)) return QKRYLOV_ERROR_INVALID_ARG;
    QK_TRY
    hipStream_t s = qk::default_stream();
    if (qkl_fill_random(kind, n, seed, 0, 1, 1, 0, x, s) != 0) throw std::runtime_error("fill launch failed");
    qk::hip_check(hipStreamSynchronize(s), "sync");
    return QKRYLOV_SUCCESS;
    QK_CATCH
}

qkrylov_hamiltonian_h qkrylov_hamiltonian_create_ex(qkrylov_basis_h b, qkrylov_site_h s, qkrylov_opsum_h o, int NL)
{
    if (!b || !b->ptr || !s || !s->ptr || !o) return nullptr;
    if (NL < 0) return qkrylov_hamiltonian_create(b, s, o);
    // an explicit split goes through the engine directly
    return make_handle<qkrylov_hamiltonian_t>([&] {
        auto h = std::make_unique<qkrylov_hamiltonian_t>();
        h->ptr = std::make_unique<MatrixFreeHamiltonianBase>(b->ptr, s->ptr, o->opsum, Device(), NL);
        return h.release();
    });
}

qkrylov_hamiltonian_h qkrylov_operator_create(qkrylov_basis_h b, qkrylov_site_h s, qkrylov_opsum_h o)
{
    if (!b || !b->ptr || !s || !s->ptr || !o) return nullptr;
    return make_handle<qkrylov_hamiltonian_t>([&] {
        auto h = std::make_unique<qkrylov_hamiltonian_t>();
        h->ptr = std::make_unique<MatrixFreeHamiltonianBase>(b->ptr, s->ptr, adjoint(o->opsum), Device());
        return h.release();
    });
}

int qkrylov_hamiltonian_info(qkrylov_hamiltonian_h h, qkrylov_hamiltonian_info_t* info)
{
    const auto* H = H_of(h);
    if (!H || !info) return QKRYLOV_ERROR_INVALID_ARG;
    const auto& E = *H->engine();
    info->dim = (uint64_t)E.dim();
    info->nbits = E.basis().nbits();
    info->NL = E.NL();
    info->is_real = E.is_real() ? 1 : 0;
    info->n_offdiag_branches = (int)E.compiled().offdiag.size();
    info->n_diag_terms = (int)E.compiled().diag.size();
    info->ntiles = E.tables()->ntiles;
    info->table_bytes = E.table_bytes();
    return QKRYLOV_SUCCESS;
}

int qkrylov_hamiltonian_set_tuning(qkrylov_hamiltonian_h h, int knob, int value)
{
    const auto* H = H_of(h);
    if (!H) return QKRYLOV_ERROR_INVALID_ARG;
    try {
        H->engine()->set_tuning(knob, value);
    } catch (const std::invalid_argument& e) {
        g_last_error = e.what();
        return QKRYLOV_ERROR_INVALID_ARG;
    }
    return QKRYLOV_SUCCESS;
}

int qkrylov_hamiltonian_autotune(qkrylov_hamiltonian_h h, int nvec, int allow_split, qkrylov_tune_report_t* report)
{
    if (!h || !h->ptr || nvec <= 0 || nvec > 4096) return QKRYLOV_ERROR_INVALID_ARG;
    QK_TRY
    const auto r = h->ptr->autotune(nvec, allow_split != 0);
    if (report) {
        report->split_bits = r.split_bits;
        report->apply_waves = r.apply_waves;
        report->schedule = r.schedule;
        report->block_waves = r.block_waves;
        report->block_schedule = r.block_schedule;
        report->candidates = r.candidates;
        report->ms_before = r.ms_before;
        report->ms_after = r.ms_after;
        report->block_rows = r.block_rows;
    }
    return QKRYLOV_SUCCESS;
    QK_CATCH
}

int qkrylov_hamiltonian_get_tuning(qkrylov_hamiltonian_h h, int knob, int* value)
{
    const auto* H = H_of(h);
    if (!H || !value) return QKRYLOV_ERROR_INVALID_ARG;
    try {
        *value = H->engine()->tuning(knob);
    } catch (const std::invalid_argument& e) {
        g_last_error = e.what();
        return QKRYLOV_ERROR_INVALID_ARG;
    }
    return QKRYLOV_SUCCESS;
}

int qkrylov_hamiltonian_apply_device_f64(qkrylov_hamiltonian_h h, const double* x, double* y, int nvec, void* stream)
{
    const auto* H = H_of(h);
    if (!H || !x || !y || nvec <= 0) return QKRYLOV_ERROR_INVALID_ARG;
    QK_TRY
    H->apply_device(x, y, nvec, stream);
    return QKRYLOV_SUCCESS;
    QK_CATCH
}
int qkrylov_hamiltonian_apply_columns_device_f64(qkrylov_hamiltonian_h h, const double* x, double* y, int nv,
                                                 const double* diag_scale, void* stream)
{
    const auto* H = H_of(h);
    if (!H || !x || !y || !diag_scale || qk::cols_width(nv) != nv) return QKRYLOV_ERROR_INVALID_ARG;
    const auto& E = *H->engine();
    if (!E.is_real() || E.wide() || E.tables()->signs != 0) {
        g_last_error = "apply_columns: needs a real Hamiltonian without fermion signs on at most 32 bits";
        return QKRYLOV_ERROR_INVALID_ARG;
    }
    QK_TRY
    hipStream_t s = stream ? (hipStream_t)stream : qk::default_stream();
    qk::DevBuf sc(nv * sizeof(double));
    qk::hip_check(hipMemcpyAsync(sc.get(), diag_scale, nv * sizeof(double), hipMemcpyHostToDevice, s), "h2d");
    E.apply_cols(x, y, nv, sc.d(), s);
    qk::hip_check(hipStreamSynchronize(s), "sync");   // before the scale buffer goes
    return QKRYLOV_SUCCESS;
    QK_CATCH
}
int qkrylov_hamiltonian_apply_device_complex_f64(qkrylov_hamiltonian_h h, const double* x, double* y, void* stream)
{
    const auto* H = H_of(h);
    if (!H || !x || !y) return QKRYLOV_ERROR_INVALID_ARG;
    QK_TRY
    H->apply_device_complex(x, y, stream);
    return QKRYLOV_SUCCESS;
    QK_CATCH
}
int qkrylov_hamiltonian_diagonal_device_f64(qkrylov_hamiltonian_h h, double* d, int cplx, void* stream)
{
    const auto* H = H_of(h);
    if (!H || !d) return QKRYLOV_ERROR_INVALID_ARG;
    QK_TRY
    H->engine()->diagonal(d, cplx != 0, stream ? (hipStream_t)stream : qk::default_stream());
    return QKRYLOV_SUCCESS;
    QK_CATCH
}

int qkrylov_hamiltonian_row_entries(qkrylov_hamiltonian_h h, uint64_t r0, uint64_t n, int cap, int64_t* cols,
                                    double* vals, int* counts)
{
    const auto* H = H_of(h);
    if (!H || !cols || !vals || !counts || cap <= 0) return QKRYLOV_ERROR_INVALID_ARG;
    if (r0 + n > (uint64_t)H->dimension()) return QKRYLOV_ERROR_INVALID_ARG;
    QK_TRY
    if (n == 0) return QKRYLOV_SUCCESS;
    hipStream_t s = qk::default_stream();
    qk::DevBuf dc(n * cap * sizeof(int64_t)), dv(2 * n * cap * sizeof(double)), dn(n * sizeof(int));
    const auto& E = *H->engine();
    if (qkl_row_entries(E.tables(), E.wide() ? 1 : 0, r0, n, cap, static_cast<int64_t*>(dc.get()), dv.d(),
                        static_cast<int*>(dn.get()), s) != 0)
        throw std::runtime_error("row_entries launch failed");
    qk::hip_check(hipMemcpyAsync(cols, dc.get(), n * cap * sizeof(int64_t), hipMemcpyDeviceToHost, s), "d2h");
    qk::hip_check(hipMemcpyAsync(vals, dv.get(), 2 * n * cap * sizeof(double), hipMemcpyDeviceToHost, s), "d2h");
    qk::hip_check(hipMemcpyAsync(counts, dn.get(), n * sizeof(int), hipMemcpyDeviceToHost, s), "d2h");
    qk::hip_check(hipStreamSynchronize(s), "sync");
    return QKRYLOV_SUCCESS;
    QK_CATCH
}

int qkrylov_hamiltonian_states(qkrylov_hamiltonian_h h, uint64_t r0, uint64_t n, uint64_t* out)
{
    const auto* H = H_of(h);
    if (!H || !out) return QKRYLOV_ERROR_INVALID_ARG;
    if (r0 + n > (uint64_t)H->dimension()) return QKRYLOV_ERROR_INVALID_ARG;
    QK_TRY
    if (n == 0) return QKRYLOV_SUCCESS;
    hipStream_t s = qk::default_stream();
    qk::DevBuf d(n * sizeof(uint64_t));
    const auto& E = *H->engine();
    if (qkl_states(E.tables(), E.wide() ? 1 : 0, r0, n, static_cast<uint64_t*>(d.get()), s) != 0)
        throw std::runtime_error("states launch failed");
    qk::hip_check(hipMemcpyAsync(out, d.get(), n * sizeof(uint64_t), hipMemcpyDeviceToHost, s), "d2h");
    qk::hip_check(hipStreamSynchronize(s), "sync");
    return QKRYLOV_SUCCESS;
    QK_CATCH
}

int qkrylov_hamiltonian_time_apply(qkrylov_hamiltonian_h h, int nvec, int warmup, int reps, double* mean_ms)
{
    const auto* H = H_of(h);
    if (!H || nvec <= 0 || reps <= 0 || !mean_ms) return QKRYLOV_ERROR_INVALID_ARG;
    QK_TRY
    const auto& E = *H->engine();
    const size_t D = (size_t)E.dim();
    hipStream_t s = qk::default_stream();
    qk::DevBuf x(D * nvec * sizeof(double)), y(D * nvec * sizeof(double));
    if (qkl_fill_random(0, D, 7, 0, nvec, nvec, 0, x.d(), s) != 0) throw std::runtime_error("fill failed");
    for (int i = 0; i < warmup; ++i) H->apply_device(x.d(), y.d(), nvec, s);
    hipEvent_t e0, e1;
    qk::hip_check(hipEventCreate(&e0), "event");
    qk::hip_check(hipEventCreate(&e1), "event");
    qk::hip_check(hipEventRecord(e0, s), "record");
    for (int i = 0; i < reps; ++i) H->apply_device(x.d(), y.d(), nvec, s);
    qk::hip_check(hipEventRecord(e1, s), "record");
    qk::hip_check(hipEventSynchronize(e1), "sync");
    float ms = 0.0f;
    qk::hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *mean_ms = (double)ms / reps;
    return QKRYLOV_SUCCESS;
    QK_CATCH
}

int qkrylov_hamiltonian_split_info(qkrylov_hamiltonian_h h, int any_size, int* NL, int* nB, int* nC)
{
    const auto* H = H_of(h);
    if (!H || !NL || !nB || !nC) return QKRYLOV_ERROR_INVALID_ARG;
    QK_TRY
    const auto* sp = H->engine()->split_engines(any_size ? 0 : qk::Engine::kSplitMinRows);
    *NL = sp ? sp->plan().NL : 0;
    *nB = sp ? sp->plan().nB : 0;
    *nC = sp ? sp->plan().nC : 0;
    return QKRYLOV_SUCCESS;
    QK_CATCH
}

int qkrylov_hamiltonian_apply_split_device_f64(qkrylov_hamiltonian_h h, const double* x, double* y, void* stream)
{
    const auto* H = H_of(h);
    if (!H || !x || !y) return QKRYLOV_ERROR_INVALID_ARG;
    QK_TRY
    const auto* sp = H->engine()->split_engines(0);
    if (!sp) throw std::invalid_argument("this model admits no split apply");
    const size_t D = H->dimension();
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : qk::default_stream();
    qk::DevBuf xi(D * sizeof(double)), yi(D * sizeof(double)), t(D * sizeof(double));
    sp->to_internal(x, xi.d(), s);
    sp->apply(xi.d(), yi.d(), t.d(), s);
    sp->to_public(yi.d(), y, s);
    qk::hip_check(hipStreamSynchronize(s), "sync");
    return QKRYLOV_SUCCESS;
    QK_CATCH
}

int qkrylov_hamiltonian_time_split_apply(qkrylov_hamiltonian_h h, int warmup, int reps, double* mean_ms)
{
    const auto* H = H_of(h);
    if (!H || reps <= 0 || !mean_ms) return QKRYLOV_ERROR_INVALID_ARG;
    QK_TRY
    const auto* sp = H->engine()->split_engines(0);
    if (!sp) throw std::invalid_argument("this model admits no split apply");
    const size_t D = H->dimension();
    hipStream_t s = qk::default_stream();
    qk::DevBuf x(D * sizeof(double)), y(D * sizeof(double)), t(D * sizeof(double));
    if (qkl_fill_random(0, D, 7, 0, 1, 1, 0, x.d(), s) != 0) throw std::runtime_error("fill failed");
    for (int i = 0; i < warmup; ++i) sp->apply(x.d(), y.d(), t.d(), s);
    hipEvent_t e0, e1;
    qk::hip_check(hipEventCreate(&e0), "event");
    qk::hip_check(hipEventCreate(&e1), "event");
    qk::hip_check(hipEventRecord(e0, s), "record");
    for (int i = 0; i < reps; ++i) sp->apply(x.d(), y.d(), t.d(), s);
    qk::hip_check(hipEventRecord(e1, s), "record");
    qk::hip_check(hipEventSynchronize(e1), "sync");
    float ms = 0.0f;
    qk::hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *mean_ms = (double)ms / reps;
    return QKRYLOV_SUCCESS;
    QK_CATCH
}

int qkrylov_lanczos_run_f64(qkrylov_hamiltonian_h h, const qkrylov_lanczos_options_t* opt, const double* start,
                            double* alphas, double* betas, double* vector, qkrylov_lanczos_report_t* rep)
{
    const auto* H = H_of(h);
    if (!H || !opt || !rep) return QKRYLOV_ERROR_INVALID_ARG;
    QK_TRY
    qk::LanczosOptions o;
    o.maxiter = opt->maxiter;
    o.tol = opt->tol;
    o.reorth = opt->reorthogonalize != 0;
    o.fixed_steps = opt->fixed_steps != 0;
    o.want_vector = (opt->want_vector != 0) && vector != nullptr;
    o.seed = opt->seed;
    o.check_every = opt->check_every;
    if (opt->graph_mode < 0 || opt->graph_mode > 2) return invalid("lanczos: graph_mode must be 0, 1 or 2");
    o.graph = opt->graph_mode == 0 ? -1 : (opt->graph_mode == 1 ? 1 : 0);
    if (opt->split_mode < 0 || opt->split_mode > 2) return invalid("lanczos: split_mode must be 0, 1 or 2");
    o.split = opt->split_mode == 0 ? -1 : (opt->split_mode == 1 ? 1 : 0);
    if (opt->start_mode < -1 || opt->start_mode > 2) return invalid("lanczos: start_mode must be -1, 0, 1 or 2");
    o.start_mode = opt->start_mode < 0 ? qkrylov::default_start_mode() : opt->start_mode;
    o.start = start;
    auto r = qk::lanczos_run(*H->engine(), o, qk::default_stream());
    rep->split = r.split ? 1 : 0;
    rep->vector_complex = (!H->is_real() || o.start_mode != 0) ? 1 : 0;
    rep->energy = r.energy;
    rep->iterations = r.iterations;
    rep->converged = r.converged ? 1 : 0;
    rep->steps_run = r.steps_run;
    rep->n_alpha = (int)r.alphas.size();
    rep->n_beta = (int)r.betas.size();
    rep->loop_seconds = r.loop_seconds;
    rep->ritz_seconds = r.ritz_seconds;
    if (alphas) std::memcpy(alphas, r.alphas.data(), r.alphas.size() * sizeof(double));
    if (betas) std::memcpy(betas, r.betas.data(), r.betas.size() * sizeof(double));
    if (o.want_vector) std::memcpy(vector, r.vector.data(), r.vector.size() * sizeof(double));
    return QKRYLOV_SUCCESS;
    QK_CATCH
}

int qkrylov_lanczos_run_noalloc_f64(qkrylov_hamiltonian_h h, const qkrylov_lanczos_options_t* opt, const double* start,
                                    double* alphas, double* betas, double* vector, qkrylov_lanczos_report_t* rep)
{
    const auto* H = H_of(h);
    if (!H || !opt || !rep) return QKRYLOV_ERROR_INVALID_ARG;
    if (opt->graph_mode < 0 || opt->graph_mode > 2) return invalid("lanczos: graph_mode must be 0, 1 or 2");
    QK_TRY
    if (!H->is_real()) throw std::invalid_argument("the no-allocation Lanczos driver needs a real Hamiltonian");
    qk::LanczosOptions o;
    o.maxiter = opt->maxiter;
    o.tol = opt->tol;
    o.reorth = opt->reorthogonalize != 0;
    o.fixed_steps = opt->fixed_steps != 0;
    o.want_vector = (opt->want_vector != 0) && vector != nullptr;
    o.seed = opt->seed;
    o.check_every = opt->check_every;
    o.graph = 0;   // graph instantiation may allocate
    if (opt->start_mode < -1 || opt->start_mode > 2) return invalid("lanczos: start_mode must be -1, 0, 1 or 2");
    o.start_mode = opt->start_mode < 0 ? qkrylov::default_start_mode() : opt->start_mode;
    if (o.start_mode != 0)
        throw std::invalid_argument("the no-allocation Lanczos driver runs real vectors only (start_mode 0)");
    o.start = start;
    int dev = 0;
    qk::hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= 64) throw std::runtime_error("device id out of range");
    // one caller at a time per device: every call shares the device's array
    static std::mutex mu[64];
    std::lock_guard<std::mutex> lk(mu[dev]);
    void* base = nullptr;
    size_t cap = 0;
    if (qkl_scratch(&base, &cap) != 0) throw std::runtime_error("static workspace lookup failed");
    const size_t need = qk::lanczos_arena_bytes(*H->engine(), o);
    if (need > cap)
        throw std::length_error("no-allocation Lanczos: needs " + std::to_string(need) + " bytes of workspace, the static "
                                "workspace holds " + std::to_string(cap) + " (dimension " +
                                std::to_string(H->dimension()) + ")");
    qk::Arena arena;
    arena.base = static_cast<char*>(base);
    arena.cap = cap;
    auto r = qk::lanczos_run(*H->engine(), o, qk::default_stream(), &arena);
    rep->split = 0;
    rep->vector_complex = 0;
    rep->energy = r.energy;
    rep->iterations = r.iterations;
    rep->converged = r.converged ? 1 : 0;
    rep->steps_run = r.steps_run;
    rep->n_alpha = (int)r.alphas.size();
    rep->n_beta = (int)r.betas.size();
    rep->loop_seconds = r.loop_seconds;
    rep->ritz_seconds = r.ritz_seconds;
    if (alphas) std::memcpy(alphas, r.alphas.data(), r.alphas.size() * sizeof(double));
    if (betas) std::memcpy(betas, r.betas.data(), r.betas.size() * sizeof(double));
    if (o.want_vector) std::memcpy(vector, r.vector.data(), r.vector.size() * sizeof(double));
    return QKRYLOV_SUCCESS;
    QK_CATCH
}

int qkrylov_ftlm_batched_f64(qkrylov_hamiltonian_h h, int R, int M, uint64_t seed, const double* bT, int nT,
                             int sequential, double* Z, double* E, double* Cv, double* alphas, double* betas_l,
                             int* nsteps, double* seconds)
{
    const auto* H = H_of(h);
    if (!H || R <= 0 || M <= 0 || nT < 0 || (nT > 0 && (!bT || !Z || !E || !Cv))) return QKRYLOV_ERROR_INVALID_ARG;
    QK_TRY
    std::vector<double> bv(bT, bT + nT);
    auto r = ftlm_batched(*H, R, M, seed, bv, sequential != 0);
    for (int t = 0; t < nT; ++t) {
        Z[t] = r.partition_function[t];
        E[t] = r.internal_energy[t];
        Cv[t] = r.specific_heat[t];
    }
    for (int v = 0; v < R && v < (int)r.alphas.size(); ++v) {
        if (alphas) for (size_t k = 0; k < r.alphas[v].size() && k < (size_t)M; ++k) alphas[(size_t)v * M + k] = r.alphas[v][k];
        if (betas_l) for (size_t k = 0; k < r.betas_lanczos[v].size() && k < (size_t)M; ++k) betas_l[(size_t)v * M + k] = r.betas_lanczos[v][k];
        if (nsteps) nsteps[v] = (int)r.alphas[v].size();
    }
    if (seconds) *seconds = r.seconds;
    return QKRYLOV_SUCCESS;
    QK_CATCH
}

int qkrylov_cf_batched_f64(qkrylov_hamiltonian_h h, int nseeds, const double* seeds, int n_iter, double* alphas,
                           double* betas, double* norms, int* n_alpha, double* seconds)
{
    const auto* H = H_of(h);
    if (!H || nseeds <= 0 || !seeds || n_iter <= 0 || !alphas || !betas || !norms) return QKRYLOV_ERROR_INVALID_ARG;
    QK_TRY
    const size_t D = H->dimension();
    std::vector<HostVector> sv(nseeds, HostVector(D));
    for (int s = 0; s < nseeds; ++s)
        for (size_t i = 0; i < D; ++i) sv[s][i] = Complex(seeds[(size_t)s * 2 * D + 2 * i], seeds[(size_t)s * 2 * D + 2 * i + 1]);
    auto r = continued_fraction_batched(*H, sv, n_iter, seconds);
    for (int s = 0; s < nseeds; ++s) {
        for (size_t k = 0; k < r[s].alphas.size() && k < (size_t)n_iter; ++k) alphas[(size_t)s * n_iter + k] = r[s].alphas[k];
        for (size_t k = 0; k < r[s].betas.size() && k < (size_t)n_iter; ++k) betas[(size_t)s * n_iter + k] = r[s].betas[k];
        norms[s] = r[s].norm_phi0;
        if (n_alpha) n_alpha[s] = (int)r[s].alphas.size();
    }
    return QKRYLOV_SUCCESS;
    QK_CATCH
}

int qkrylov_structure_factor_f64(qkrylov_hamiltonian_h h, const qkrylov_hamiltonian_h* operators, int nops,
                                 const double* psi0, int psi0_complex, double E0, int n_iter, const double* omega,
                                 int nomega, double eta, double* spectrum, double* alphas, double* betas,
                                 double* norms, int* n_alpha, double* seconds)
{
    const auto* H = H_of(h);
    if (!H || !operators || nops <= 0 || !psi0 || n_iter <= 0 || nomega < 0 || (nomega > 0 && (!omega || !spectrum)))
        return QKRYLOV_ERROR_INVALID_ARG;
    std::vector<const MatrixFreeHamiltonianBase*> ops(nops);
    for (int q = 0; q < nops; ++q) {
        ops[q] = H_of(operators[q]);
        if (!ops[q]) return QKRYLOV_ERROR_INVALID_ARG;
    }
    QK_TRY
    const size_t D = H->dimension();
    HostVector p(D);
    for (size_t i = 0; i < D; ++i) p[i] = psi0_complex ? Complex(psi0[2 * i], psi0[2 * i + 1]) : Complex(psi0[i], 0.0);
    std::vector<Real> om(omega, omega + nomega);
    auto r = structure_factor(*H, ops, p, E0, n_iter, om, eta);
    for (int q = 0; q < nops; ++q) {
        const auto& d = r.cf[q];
        for (int w = 0; w < nomega; ++w) spectrum[(size_t)q * nomega + w] = r.spectrum[q][w];
        if (alphas) for (size_t k = 0; k < d.alphas.size() && k < (size_t)n_iter; ++k) alphas[(size_t)q * n_iter + k] = d.alphas[k];
        if (betas) for (size_t k = 0; k < d.betas.size() && k < (size_t)n_iter; ++k) betas[(size_t)q * n_iter + k] = d.betas[k];
        if (norms) norms[q] = d.norm_phi0;
        if (n_alpha) n_alpha[q] = (int)d.alphas.size();
    }
    if (seconds) *seconds = r.seconds;
    return QKRYLOV_SUCCESS;
    QK_CATCH
}

/* ---- host eigensolvers of the Krylov drivers (no device work) ---- */
int qkrylov_amd_tridiag_eigen(int n, const double* alpha, const double* beta, double* evals, double* evecs, int want)
{
    if (n < 0 || (n > 0 && (!alpha || !evals)) || (n > 1 && !beta) || want < 0 || want > 2 ||
        (want > 0 && n > 0 && !evecs))
        return QKRYLOV_ERROR_INVALID_ARG;
    QK_TRY
    qk::symtri_eig(alpha, beta, n, evals, evecs, want);
    return QKRYLOV_SUCCESS;
    QK_CATCH
}

int qkrylov_amd_tridiag_lowest(int n, const double* alpha, const double* beta, double* lowest)
{
    if (n <= 0 || !alpha || (n > 1 && !beta) || !lowest) return QKRYLOV_ERROR_INVALID_ARG;
    QK_TRY
    *lowest = qk::symtri_lowest(alpha, beta, n);
    return QKRYLOV_SUCCESS;
    QK_CATCH
}

int qkrylov_amd_hermitian_eigen(int n, const double* a, double* evals, double* evecs)
{
    if (n < 0 || (n > 0 && (!a || !evals || !evecs))) return QKRYLOV_ERROR_INVALID_ARG;
    QK_TRY
    std::vector<std::complex<double>> A((size_t)n * n), V;
    for (size_t i = 0; i < (size_t)n * n; ++i) A[i] = {a[2 * i], a[2 * i + 1]};
    std::vector<double> w;
    qk::herm_eig(n, A, w, V);
    for (int i = 0; i < n; ++i) evals[i] = w[i];
    for (size_t i = 0; i < (size_t)n * n; ++i) {
        evecs[2 * i] = V[i].real();
        evecs[2 * i + 1] = V[i].imag();
    }
    return QKRYLOV_SUCCESS;
    QK_CATCH
}

} // extern "C"
