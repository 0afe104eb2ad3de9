// qk_api.cpp -- implementation of the C++ drop-in surface (qkrylov::fp64)
// over the engine. Site::apply keeps the reference's per-call semantics
// (src/sites/*_site.cpp) for callers that use sites directly.
#include "qkrylov/qkrylov.hpp"

#include "qk_engine.hpp"

#include <hip/hip_runtime.h>

#include <atomic>
#include <random>
#include <stdexcept>

namespace qkrylov {

namespace {
std::atomic<int> g_start_mode{0};
}

void set_default_start_mode(int mode)
{
    if (mode < 0 || mode > 2) throw std::invalid_argument("start mode must be 0, 1 or 2");
    g_start_mode.store(mode);
}

int default_start_mode() { return g_start_mode.load(); }

namespace fp64 {

using qk::DevBuf;
using qk::hip_check;

// ---------------------------------------------------------------- Device
Device::Device(const std::string& s)
{
    if (s == "cpu") {
        id = 0;
        return;
    }
    const auto colon = s.find(':');
    id = (colon == std::string::npos) ? 0 : std::stoi(s.substr(colon + 1));
}

int Device::gpu_count()
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

// ---------------------------------------------------------------- Basis
static qk::Sector to_qk(const Sector& s)
{
    qk::Sector q;
    q.use_sz = s.use_sz;
    q.sz2 = s.sz2;
    q.use_nup = s.use_nup;
    q.use_ndn = s.use_ndn;
    q.nup = s.nup;
    q.ndn = s.ndn;
    q.use_n = s.use_n;
    q.n = s.n;
    q.use_nb = s.use_nb;
    q.nb = s.nb;
    return q;
}

Basis::Basis(int kind, int N, const Sector& sector)
    : spec_(std::make_shared<qk::BasisSpec>(static_cast<qk::Kind>(kind), N, to_qk(sector))), sector_(sector)
{
}
Basis::~Basis() = default;
Index Basis::size() const { return static_cast<Index>(spec_->dim()); }
StateID Basis::state(Index i) const
{
    if (i >= static_cast<Index>(spec_->dim())) throw std::out_of_range("basis index out of range");
    return spec_->unrank(static_cast<int64_t>(i));
}
Index Basis::index(StateID s) const
{
    const int64_t r = spec_->rank(s);
    if (r < 0) throw std::runtime_error("State not present in basis");
    return static_cast<Index>(r);
}
bool Basis::contains(StateID s) const { return spec_->contains(s); }
int Basis::nsites() const noexcept { return spec_->nsites(); }

SpinHalfBasis::SpinHalfBasis(int N, const Sector& s) : Basis(0, N, s) {}
FermionBasis::FermionBasis(int N, const Sector& s) : Basis(1, N, s) {}
HubbardBasis::HubbardBasis(int N, const Sector& s) : Basis(2, N, s) {}
TJBasis::TJBasis(int N, const Sector& s) : Basis(3, N, s) {}

// ---------------------------------------------------------------- Sites
static Real jw(StateID st, int bit) { return (popcount(st & ((1ULL << bit) - 1ULL)) % 2 == 0) ? 1.0 : -1.0; }

LocalAction SpinHalfSite::apply(const std::string& op, int site, StateID st) const
{
    LocalAction a;
    const bool up = (st >> site) & 1ULL;
    if (op == "Sz") { a.valid = true; a.new_state = st; a.matrix_element = up ? 0.5 : -0.5; return a; }
    if (op == "Sp") { if (up) return a; a.valid = true; a.new_state = st | (1ULL << site); a.matrix_element = 1.0; return a; }
    if (op == "Sm") { if (!up) return a; a.valid = true; a.new_state = st & ~(1ULL << site); a.matrix_element = 1.0; return a; }
    if (op == "Sx") { a.valid = true; a.new_state = st ^ (1ULL << site); a.matrix_element = 0.5; return a; }
    if (op == "Sy") {
        a.valid = true;
        a.new_state = st ^ (1ULL << site);
        a.matrix_element = up ? Complex(0.0, -0.5) : Complex(0.0, 0.5);
        return a;
    }
    throw std::runtime_error("Unknown spin operator: " + op);
}

LocalAction FermionSite::apply(const std::string& op, int site, StateID st) const
{
    LocalAction a;
    const bool occ = (st >> site) & 1ULL;
    if (op == "N") { a.valid = true; a.new_state = st; a.matrix_element = occ ? 1.0 : 0.0; return a; }
    if (op == "Id") { a.valid = true; a.new_state = st; a.matrix_element = 1.0; return a; }
    if (op == "C") { if (!occ) return a; a.valid = true; a.new_state = st & ~(1ULL << site); a.matrix_element = jw(st, site); return a; }
    if (op == "Cdag") { if (occ) return a; a.valid = true; a.new_state = st | (1ULL << site); a.matrix_element = jw(st, site); return a; }
    throw std::runtime_error("Unknown fermion operator: " + op);
}

static LocalAction two_species(bool tj, const std::string& op, int site, StateID st)
{
    LocalAction a;
    const int bu = 2 * site, bd = 2 * site + 1;
    const bool up = (st >> bu) & 1ULL, dn = (st >> bd) & 1ULL;
    if (op == "Nup") { a.valid = true; a.new_state = st; a.matrix_element = up ? 1.0 : 0.0; return a; }
    if (op == "Ndn") { a.valid = true; a.new_state = st; a.matrix_element = dn ? 1.0 : 0.0; return a; }
    if (!tj && op == "Nupdn") { a.valid = true; a.new_state = st; a.matrix_element = (up && dn) ? 1.0 : 0.0; return a; }
    if (op == "CUp") { if (!up) return a; a.valid = true; a.new_state = st & ~(1ULL << bu); a.matrix_element = jw(st, bu); return a; }
    if (op == "CdagUp") { if (up || (tj && dn)) return a; a.valid = true; a.new_state = st | (1ULL << bu); a.matrix_element = jw(st, bu); return a; }
    if (op == "CDn") { if (!dn) return a; a.valid = true; a.new_state = st & ~(1ULL << bd); a.matrix_element = jw(st, bd); return a; }
    if (op == "CdagDn") { if (dn || (tj && up)) return a; a.valid = true; a.new_state = st | (1ULL << bd); a.matrix_element = jw(st, bd); return a; }
    throw std::runtime_error(std::string(tj ? "Unknown TJ operator: " : "Unknown Hubbard operator: ") + op);
}
LocalAction HubbardSite::apply(const std::string& op, int site, StateID st) const { return two_species(false, op, site, st); }
LocalAction TJSite::apply(const std::string& op, int site, StateID st) const { return two_species(true, op, site, st); }

// ---------------------------------------------------------------- Hamiltonian
static qk::OpSum to_qk(const OpSum& ops)
{
    qk::OpSum q;
    for (const auto& t : ops.terms()) {
        qk::Term qt;
        qt.coeff = t.coeff;
        for (const auto& f : t.factors) qt.factors.push_back({f.op, f.site});
        q.terms.push_back(std::move(qt));
    }
    return q;
}

MatrixFreeHamiltonianBase::MatrixFreeHamiltonianBase(std::shared_ptr<Basis> basis, std::shared_ptr<Site> site,
                                             const OpSum& ops, Device device, int split_bits)
    : device_(device), basis_(std::move(basis)), site_(std::move(site)), ops_(ops)
{
    if (!basis_ || !site_) throw std::invalid_argument("null basis or site");
    qk::ensure_device(device_.id);
    qk::EngineOptions eo;
    eo.NL = split_bits;
    engine_ = std::make_shared<qk::Engine>(basis_->spec(), static_cast<qk::Kind>(site_->kind()), to_qk(ops_), eo);
}

MatrixFreeHamiltonianBase::~MatrixFreeHamiltonianBase() = default;

MatrixFreeHamiltonianBase::TuneReport MatrixFreeHamiltonianBase::autotune(int nvec, bool allow_split)
{
    qk::TuneResult r;
    engine_ = qk::autotune(basis_->spec(), static_cast<qk::Kind>(site_->kind()), to_qk(ops_), engine_, nvec,
                           allow_split, qk::default_stream(), r);
    return TuneReport{r.NL, r.apply_waves, r.sched, r.blk_waves, r.blk_sched, r.tried, r.ms_before, r.ms, r.blk_rows};
}

Index MatrixFreeHamiltonianBase::dimension() const { return static_cast<Index>(engine_->dim()); }
bool MatrixFreeHamiltonianBase::is_real() const { return engine_->is_real(); }

void MatrixFreeHamiltonianBase::apply(const Complex* x, Complex* y) const
{
    const size_t D = dimension();
    if (D == 0) return;
    hipStream_t s = qk::default_stream();
    DevBuf dx(2 * D * sizeof(double)), dy(2 * D * sizeof(double));
    hip_check(hipMemcpyAsync(dx.get(), x, 2 * D * sizeof(double), hipMemcpyHostToDevice, s), "h2d");
    engine_->apply(dx.d(), dy.d(), true, s);
    hip_check(hipMemcpyAsync(y, dy.get(), 2 * D * sizeof(double), hipMemcpyDeviceToHost, s), "d2h");
    hip_check(hipStreamSynchronize(s), "sync");
}

void MatrixFreeHamiltonianBase::apply_device(const double* x, double* y, int nvec, void* stream) const
{
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : qk::default_stream();
    if (nvec == 1) engine_->apply(x, y, false, s);
    else engine_->apply_block(x, y, nvec, nvec, s);
}

void MatrixFreeHamiltonianBase::apply_device_complex(const double* x, double* y, void* stream) const
{
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : qk::default_stream();
    engine_->apply(x, y, true, s);
}

HostVector MatrixFreeHamiltonianBase::diagonal_host() const
{
    const size_t D = dimension();
    HostVector out(D);
    if (D == 0) return out;
    hipStream_t s = qk::default_stream();
    DevBuf d(2 * D * sizeof(double));
    engine_->diagonal(d.d(), true, s);
    hip_check(hipMemcpyAsync(out.data(), d.get(), 2 * D * sizeof(double), hipMemcpyDeviceToHost, s), "d2h");
    hip_check(hipStreamSynchronize(s), "sync");
    return out;
}

// ---------------------------------------------------------------- solvers
LanczosResult lanczos_ground_state(const MatrixFreeHamiltonianBase& H, const LanczosOptions& opt)
{
    qk::LanczosOptions o;
    o.maxiter = opt.maxiter;
    o.tol = opt.tol;
    o.reorth = opt.reorthogonalize;
    o.want_vector = opt.want_eigenvector;
    o.seed = opt.seed;
    o.start_mode = opt.start_mode < 0 ? default_start_mode() : opt.start_mode;
    auto r = qk::lanczos_run(*H.engine(), o, qk::default_stream());
    LanczosResult res;
    res.energy = r.energy;
    res.iterations = r.iterations;
    res.converged = r.converged;
    if (opt.want_eigenvector) {
        const size_t D = H.dimension();
        res.eigenvector.resize(D);
        if (H.is_real() && r.vector.size() == D)
            for (size_t i = 0; i < D && i < r.vector.size(); ++i) res.eigenvector[i] = r.vector[i];
        else
            for (size_t i = 0; i < D && 2 * i + 1 < r.vector.size(); ++i)
                res.eigenvector[i] = Complex(r.vector[2 * i], r.vector[2 * i + 1]);
    }
    return res;
}

LanczosResult lanczos_ground_state(const MatrixFreeHamiltonianBase& H, int maxiter, Real tol)
{
    LanczosOptions o;
    o.maxiter = maxiter;
    o.tol = tol;
    return lanczos_ground_state(H, o);
}

DavidsonResult davidson_lowest(const MatrixFreeHamiltonianBase& H, int n_eig, int max_subspace, Real tol)
{
    auto r = qk::davidson_run(*H.engine(), n_eig, max_subspace, tol, true, qk::default_stream());
    DavidsonResult out;
    out.eigenvalues = r.eigenvalues;
    out.iterations = r.iterations;
    out.converged = r.converged;
    for (const auto& v : r.eigenvectors) {
        HostVector h(v.size() / 2);
        for (size_t i = 0; i < h.size(); ++i) h[i] = Complex(v[2 * i], v[2 * i + 1]);
        out.eigenvectors.push_back(std::move(h));
    }
    return out;
}

DynamicsResult continued_fraction_coeffs(const MatrixFreeHamiltonianBase& H, const HostVector& phi0, int n_iter)
{
    DynamicsResult out;
    if (H.dimension() == 0) return out;
    if (phi0.size() != H.dimension()) throw std::invalid_argument("phi0 has the wrong length");
    auto r = qk::tridiag_run(*H.engine(), reinterpret_cast<const double*>(phi0.data()), true, n_iter,
                             qk::default_stream());
    out.alphas = r.alphas;
    out.betas = r.betas;
    out.norm_phi0 = (r.norm0 < 1e-15) ? 0.0 : r.norm0;
    return out;
}

Real evaluate_spectral_function(const Real* alphas, const Real* betas, size_t n, Real norm_phi0, Real omega, Real E0,
                                Real eta)
{
    return qk::spectral_value(alphas, betas, n, norm_phi0, omega, E0, eta);
}

FTLMResult ftlm(const MatrixFreeHamiltonianBase& H, Real beta, int n_random, int n_steps)
{
    FTLMResult res;
    res.beta = beta;
    const size_t D = H.dimension();
    if (D == 0 || n_random <= 0) return res;
    // complex normal start vectors as in the reference (ftlm.cpp:129-144):
    // std::mt19937(42), std::normal_distribution(0, 1), one stream for all
    // vectors; the first draw of an entry is its imaginary part under start
    // mode 1 (GCC's argument order), its real part otherwise (Clang's)
    qk::ReferenceNormal draws(42, default_start_mode());
    std::vector<std::vector<double>> al, be;
    std::vector<double> n2;
    hipStream_t s = qk::default_stream();
    const auto& E = *H.engine();
    if (E.is_real()) {
        // complex vectors as column pairs of one real block
        for (int r0 = 0; r0 < n_random; r0 += 32) {
            const int nv = std::min(32, n_random - r0), nc = 2 * nv;
            std::vector<double> hb(D * (size_t)nc), hv(2 * D);
            for (int v = 0; v < nv; ++v) {
                draws.fill(D, hv.data());
                for (size_t i = 0; i < D; ++i) {
                    hb[i * nc + 2 * v] = hv[2 * i];
                    hb[i * nc + 2 * v + 1] = hv[2 * i + 1];
                }
            }
            DevBuf X(hb.size() * sizeof(double));
            hip_check(hipMemcpyAsync(X.get(), hb.data(), hb.size() * sizeof(double), hipMemcpyHostToDevice, s), "h2d");
            auto bt = qk::block_tridiag_run(E, X.d(), nc, 2, n_steps, s);
            for (int v = 0; v < nv; ++v) {
                al.push_back(bt.alphas[v]);
                be.push_back(bt.betas[v]);
                n2.push_back(bt.norm0[v] * bt.norm0[v]);
            }
        }
    } else {
        std::vector<double> hv(2 * D);
        for (int r = 0; r < n_random; ++r) {
            draws.fill(D, hv.data());
            auto t = qk::tridiag_run(E, hv.data(), true, n_steps, s);
            al.push_back(t.alphas);
            be.push_back(t.betas);
            n2.push_back(t.norm0 * t.norm0);
        }
    }
    std::vector<double> Z, En, Cv;
    qk::ftlm_observables(al, be, n2, {beta}, Z, En, Cv);
    res.partition_function = Z[0];
    res.internal_energy = En[0];
    res.specific_heat = Cv[0];
    return res;
}

FTLMGridResult ftlm_batched(const MatrixFreeHamiltonianBase& H, int R, int M, uint64_t seed, const std::vector<Real>& bT,
                            bool sequential)
{
    FTLMGridResult out;
    out.betas = bT;
    const size_t D = H.dimension();
    if (D == 0 || R <= 0) return out;
    const auto& E = *H.engine();
    if (!E.is_real()) throw std::invalid_argument("batched FTLM needs a real Hamiltonian");
    hipStream_t s = qk::default_stream();
    std::vector<double> n2;
    if (!sequential) {
        for (int r0 = 0; r0 < R; r0 += 64) {
            const int nc = std::min(64, R - r0);
            DevBuf X(D * (size_t)nc * sizeof(double));
            if (qkl_fill_random(1, D, seed, (uint64_t)r0 * D, nc, nc, 0, X.d(), s) != 0)
                throw std::runtime_error("random fill failed");
            double sec = 0.0;
            auto bt = qk::block_tridiag_run(E, X.d(), nc, 1, M, s, &sec);
            out.seconds += sec;
            for (int c = 0; c < nc; ++c) {
                out.alphas.push_back(bt.alphas[c]);
                out.betas_lanczos.push_back(bt.betas[c]);
                n2.push_back(bt.norm0[c] * bt.norm0[c]);
            }
        }
    } else {
        // one plain recurrence per vector (the reference's compute_tridiag,
        // ftlm.cpp:24-53), same vectors as the block columns
        DevBuf X(D * sizeof(double));
        for (int r = 0; r < R; ++r) {
            if (qkl_fill_random(1, D, seed, (uint64_t)r * D, 1, 1, 0, X.d(), s) != 0)
                throw std::runtime_error("random fill failed");
            double sec = 0.0;
            auto t = qk::tridiag_run_dev(E, X.d(), M, s, &sec);
            out.seconds += sec;
            out.alphas.push_back(t.alphas);
            out.betas_lanczos.push_back(t.betas);
            n2.push_back(t.norm0 * t.norm0);
        }
    }
    qk::ftlm_observables(out.alphas, out.betas_lanczos, n2, bT, out.partition_function, out.internal_energy,
                         out.specific_heat);
    return out;
}

std::vector<DynamicsResult> continued_fraction_batched(const MatrixFreeHamiltonianBase& H,
                                                       const std::vector<HostVector>& seeds, int n_iter,
                                                       double* seconds)
{
    std::vector<DynamicsResult> out;
    const size_t D = H.dimension();
    const auto& E = *H.engine();
    hipStream_t s = qk::default_stream();
    if (seconds) *seconds = 0.0;
    if (!E.is_real()) {
        for (const auto& phi : seeds) out.push_back(continued_fraction_coeffs(H, phi, n_iter));
        return out;
    }
    for (size_t g0 = 0; g0 < seeds.size(); g0 += 32) {
        const int ng = (int)std::min<size_t>(32, seeds.size() - g0), nc = 2 * ng;
        std::vector<double> hb(D * (size_t)nc);
        for (int g = 0; g < ng; ++g) {
            if (seeds[g0 + g].size() != D) throw std::invalid_argument("seed has the wrong length");
            for (size_t i = 0; i < D; ++i) {
                hb[i * nc + 2 * g] = seeds[g0 + g][i].real();
                hb[i * nc + 2 * g + 1] = seeds[g0 + g][i].imag();
            }
        }
        DevBuf X(hb.size() * sizeof(double));
        hip_check(hipMemcpyAsync(X.get(), hb.data(), hb.size() * sizeof(double), hipMemcpyHostToDevice, s), "h2d");
        double sec = 0.0;
        auto bt = qk::block_tridiag_run(E, X.d(), nc, 2, n_iter, s, &sec);
        if (seconds) *seconds += sec;
        for (int g = 0; g < ng; ++g) {
            DynamicsResult d;
            if (bt.norm0[g] >= 1e-15) {
                d.alphas = bt.alphas[g];
                d.betas = bt.betas[g];
                d.norm_phi0 = bt.norm0[g];
            }
            out.push_back(std::move(d));
        }
    }
    return out;
}

StructureFactorResult structure_factor(const MatrixFreeHamiltonianBase& H,
                                       const std::vector<const MatrixFreeHamiltonianBase*>& ops,
                                       const HostVector& psi0, Real E0, int n_iter, const std::vector<Real>& omega,
                                       Real eta)
{
    StructureFactorResult out;
    const size_t D = H.dimension();
    if (psi0.size() != D) throw std::invalid_argument("psi0 has the wrong length");
    for (const auto* op : ops) {
        if (!op) throw std::invalid_argument("null operator");
        if (op->dimension() != H.dimension()) throw std::invalid_argument("operator acts on another basis");
    }
    if (n_iter <= 0) throw std::invalid_argument("n_iter must be positive");
    if (D == 0 || ops.empty()) return out;
    const auto& E = *H.engine();
    hipStream_t s = qk::default_stream();
    DevBuf psi(2 * D * sizeof(double)), tmp(2 * D * sizeof(double));
    hip_check(hipMemcpyAsync(psi.get(), psi0.data(), 2 * D * sizeof(double), hipMemcpyHostToDevice, s), "h2d");
    auto lc = [](int rc, const char* what) {
        if (rc != 0) throw std::runtime_error(std::string("kernel launch failed: ") + what);
    };
    if (E.is_real()) {
        // up to 32 complex seeds per block (64 real columns)
        for (size_t g0 = 0; g0 < ops.size(); g0 += 32) {
            const int ng = (int)std::min<size_t>(32, ops.size() - g0), nc = 2 * ng;
            DevBuf X(D * (size_t)nc * sizeof(double));
            for (int g = 0; g < ng; ++g) {
                ops[g0 + g]->engine()->apply(psi.d(), tmp.d(), true, s);
                lc(qkl_copy_strided(D, 2, tmp.d(), 2, X.d() + 2 * g, (size_t)nc, s), "pack seed");
            }
            double sec = 0.0;
            auto bt = qk::block_tridiag_run(E, X.d(), nc, 2, n_iter, s, &sec);
            out.seconds += sec;
            for (int g = 0; g < ng; ++g) {
                DynamicsResult d;
                if (bt.norm0[g] >= 1e-15) {
                    d.alphas = bt.alphas[g];
                    d.betas = bt.betas[g];
                    d.norm_phi0 = bt.norm0[g];
                }
                out.cf.push_back(std::move(d));
            }
        }
    } else {
        HostVector seed(D);
        for (const auto* op : ops) {
            op->engine()->apply(psi.d(), tmp.d(), true, s);
            hip_check(hipMemcpyAsync(seed.data(), tmp.get(), 2 * D * sizeof(double), hipMemcpyDeviceToHost, s), "d2h");
            hip_check(hipStreamSynchronize(s), "sync");
            const double t0 = qk::now_s();
            out.cf.push_back(continued_fraction_coeffs(H, seed, n_iter));
            out.seconds += qk::now_s() - t0;
        }
    }
    for (const auto& d : out.cf) {
        std::vector<Real> sp(omega.size());
        for (size_t w = 0; w < omega.size(); ++w)
            sp[w] = qk::spectral_value(d.alphas.data(), d.betas.data(), d.alphas.size(), d.norm_phi0, omega[w], E0, eta);
        out.spectrum.push_back(std::move(sp));
    }
    return out;
}

} // namespace fp64
} // namespace qkrylov
