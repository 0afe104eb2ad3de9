// qk_engine.cpp -- Engine (device tables), the tridiagonal helpers, and the Krylov
// drivers: Lanczos ground state, plain three-term recurrences (FTLM / CF),
// the block recurrence for batched FTLM / CF, and Davidson.
#include "qk_engine.hpp"
#include "qk_eigen.hpp"
#include "qk_split.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <random>
#include <stdexcept>

namespace qk {

namespace {
inline int popc(uint64_t x) { return __builtin_popcountll(x); }
} // namespace


void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

static void lcheck(int rc, const char* what)
{
    if (rc != 0) throw std::runtime_error(std::string("kernel launch failed: ") + what + " (" + std::to_string(rc) + ")");
}

DevBuf::DevBuf(size_t bytes) : bytes_(bytes)
{
    if (bytes == 0) return;
    hipError_t e = hipMalloc(&p_, bytes);
    if (e != hipSuccess) {
        p_ = nullptr;
        throw std::runtime_error("device allocation of " + std::to_string(bytes) + " bytes failed: " +
                                 hipGetErrorString(e));
    }
}

DevBuf::~DevBuf() { reset(); }

void DevBuf::reset()
{
    if (p_ && owned_) {
        if (async_) (void)hipFreeAsync(p_, stream_);
        else (void)hipFree(p_);
    }
    p_ = nullptr;
    bytes_ = 0;
    owned_ = true;
}

DevBuf Arena::take(size_t bytes)
{
    const size_t start = (used + 255) & ~(size_t)255;
    const size_t len = std::max<size_t>(bytes, 8);
    if (start > cap || len > cap - start)
        throw std::length_error("workspace of " + std::to_string(cap) + " bytes exhausted (needs " +
                                std::to_string(start + len) + ")");
    used = start + len;
    return DevBuf::view(base + start, bytes);
}

DevBuf& DevBuf::operator=(DevBuf&& o) noexcept
{
    if (this != &o) {
        reset();
        p_ = o.p_;
        bytes_ = o.bytes_;
        stream_ = o.stream_;
        async_ = o.async_;
        owned_ = o.owned_;
        o.p_ = nullptr;
        o.bytes_ = 0;
    }
    return *this;
}

hipStream_t default_stream()
{
    struct Holder {
        hipStream_t s = nullptr;
        ~Holder() { if (s) (void)hipStreamDestroy(s); }
    };
    thread_local Holder h;
    if (!h.s) hip_check(hipStreamCreateWithFlags(&h.s, hipStreamNonBlocking), "hipStreamCreate");
    return h.s;
}

void ensure_device(int id)
{
    int n = 0;
    hip_check(hipGetDeviceCount(&n), "hipGetDeviceCount");
    if (n <= 0) throw std::runtime_error("no HIP device available");
    if (id < 0 || id >= n) throw std::invalid_argument("device id out of range");
    hip_check(hipSetDevice(id), "hipSetDevice");
}

DevBuf::DevBuf(size_t bytes, hipStream_t async_stream) : bytes_(bytes), stream_(async_stream), async_(true)
{
    if (bytes == 0) return;
    hipError_t e = hipMallocAsync(&p_, bytes, async_stream);
    if (e != hipSuccess) {
        p_ = nullptr;
        throw std::runtime_error("stream-ordered allocation of " + std::to_string(bytes) + " bytes failed: " +
                                 hipGetErrorString(e));
    }
}

template <class T>
static DevBuf upload(const std::vector<T>& v, hipStream_t s)
{
    const size_t bytes = std::max<size_t>(v.size() * sizeof(T), 8);
    DevBuf b = s ? DevBuf(bytes, s) : DevBuf(bytes);
    if (!v.empty()) {
        if (s) {
            hip_check(hipMemcpyAsync(b.get(), v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s), "upload");
            hip_check(hipStreamSynchronize(s), "upload sync");
        } else {
            hip_check(hipMemcpy(b.get(), v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "upload");
        }
    }
    return b;
}

// ---------------------------------------------------------------- shared basis tables

std::shared_ptr<const BasisTables> basis_tables(const BasisSpec& b, int NL)
{
    static std::mutex mu;
    static std::map<std::string, std::weak_ptr<const BasisTables>> cache;
    const Sector& s = b.sector();
    const std::string key = std::to_string((int)b.kind()) + "/" + std::to_string(b.nsites()) + "/" +
                            std::to_string(b.target_a()) + "/" + std::to_string(b.target_b()) + "/" +
                            std::to_string(s.use_nup || s.use_ndn || s.use_sz || s.use_n) + "/" + std::to_string(NL);
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end())
        if (auto sp = it->second.lock()) return sp;
    auto t = std::make_shared<BasisTables>();
    t->host = build_tables(b, NL);
    t->tiles = upload(t->host.tiles, nullptr);
    t->hoff = upload(t->host.hoff, nullptr);
    t->lrank = upload(t->host.lrank, nullptr);
    t->lolist = upload(t->host.lolist, nullptr);
    if (t->host.split) {
        t->tl = upload(t->host.tl, nullptr);
        t->th = upload(t->host.th, nullptr);
    }
    // the host copies of the large arrays are not needed after upload
    t->ntiles = t->host.tiles.size();
    std::vector<QkTile>().swap(t->host.tiles);
    std::vector<uint32_t>().swap(t->host.hoff);
    std::vector<uint32_t>().swap(t->host.lolist);
    std::vector<uint32_t>().swap(t->host.lrank);
    std::vector<uint16_t>().swap(t->host.tl);
    std::vector<uint32_t>().swap(t->host.th);
    cache[key] = t;
    return t;
}

// ---------------------------------------------------------------- Engine

Engine::Engine(std::shared_ptr<const BasisSpec> basis, Kind site_kind, const OpSum& ops, EngineOptions opt)
    : basis_(std::move(basis)), kind_(site_kind), ops_(ops)
{
    if (basis_->dim() > 0) comp_ = compile_opsum(*basis_, site_kind, ops);
    const int NL = opt.NL >= 0 ? opt.NL : choose_NL_model(*basis_, comp_);
    tables_ = basis_tables(*basis_, NL);
    br_ = upload(comp_.offdiag, opt.stream);
    dg_ = upload(comp_.diag, opt.stream);
    // branches split at NL: hi-only first (padded to 4), then the rest
    // (padded to 8); padding branches can never be taken (hcm = 0, hcv = 1)
    {
        const int nl = tables_->host.NL;
        const uint64_t lm = (nl >= 64) ? ~0ULL : ((1ULL << nl) - 1ULL);
        std::vector<QkBranchK> A, B;
        T_.signs = 0;
        for (const auto& b : comp_.offdiag) {
            if (b.M != 0) T_.signs = 1;
            QkBranchK k;
            k.hcm = (uint32_t)(b.cm >> nl);
            k.hcv = (uint32_t)(b.cv >> nl);
            k.halt = (uint32_t)(b.alt >> nl);
            k.hF = (uint32_t)(b.F >> nl);
            k.hM = (uint32_t)(b.M >> nl);
            k.lcm = (uint32_t)(b.cm & lm);
            k.lcv = (uint32_t)(b.cv & lm);
            k.lalt = (uint32_t)(b.alt & lm);
            k.lF = (uint32_t)(b.F & lm);
            k.lM = (uint32_t)(b.M & lm);
            k.re = b.re;
            k.im = b.im;
            const bool hi_only = (k.lcm | k.lcv | k.lalt | k.lF | k.lM) == 0;
            (hi_only ? A : B).push_back(k);
        }
        const QkBranchK pad{0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0.0, 0.0};
        while (A.size() % 4) A.push_back(pad);
        // with the two-level lo rank the per-lane branches are grouped by how
        // they move lo: flips inside the low half that keep its class (the
        // high-half term of the rank is the row's own), flips inside the high
        // half (the low-half term is the row's own), the rest; each group is
        // padded to 8
        std::vector<QkBranchK> B1, B2, B0;
        const HostTables& ht = tables_->host;
        for (const auto& k : B) {
            int kind = 0;
            if (ht.split) {
                const uint32_t m2 = (1u << ht.NL2) - 1u;
                if ((k.lF & ~m2) == 0 && (k.lcm & k.lF) == k.lF) {
                    auto keeps = [&](uint32_t pre) {
                        const uint32_t post = pre ^ k.lF;
                        return popc((uint64_t)(pre & ht.upm2)) == popc((uint64_t)(post & ht.upm2)) &&
                               popc((uint64_t)(pre & ht.dnm2)) == popc((uint64_t)(post & ht.dnm2));
                    };
                    if (keeps(k.lcv & k.lF) && keeps((k.lcv ^ k.lalt) & k.lF)) kind = 1;
                } else if ((k.lF & m2) == 0) {
                    kind = 2;
                }
            }
            (kind == 1 ? B1 : kind == 2 ? B2 : B0).push_back(k);
        }
        for (auto* v : {&B1, &B2, &B0})
            while (v->size() % 8) v->push_back(pad);
        T_.nA = (int32_t)A.size();
        T_.nB1 = (int32_t)B1.size();
        T_.nB2 = (int32_t)B2.size();
        T_.nB = (int32_t)(B1.size() + B2.size() + B0.size());
        for (auto* v : {&B1, &B2, &B0}) A.insert(A.end(), v->begin(), v->end());
        brk_ = upload(A, opt.stream);
    }
    T_.tiles = static_cast<const QkTile*>(tables_->tiles.get());
    T_.hoff = static_cast<const uint32_t*>(tables_->hoff.get());
    T_.lrank = static_cast<const uint32_t*>(tables_->lrank.get());
    T_.lolist = static_cast<const uint32_t*>(tables_->lolist.get());
    T_.br = static_cast<const QkBranch*>(br_.get());
    T_.brk = static_cast<const QkBranchK*>(brk_.get());
    T_.dg = static_cast<const QkDiagTerm*>(dg_.get());
    T_.lomask = (tables_->host.NL >= 64) ? ~0ULL : ((1ULL << tables_->host.NL) - 1ULL);
    T_.dim = static_cast<uint64_t>(basis_->dim());
    T_.dconst_re = comp_.dconst.real();
    T_.dconst_im = comp_.dconst.imag();
    T_.ntiles = static_cast<uint32_t>(tables_->ntiles);
    T_.nbr = static_cast<int32_t>(comp_.offdiag.size());
    T_.ndg = static_cast<int32_t>(comp_.diag.size());
    T_.NL = tables_->host.NL;
    T_.lrank_u16 = tables_->host.lrank_u16 ? 1 : 0;
    T_.nlo = 1 << tables_->host.NL;
    {
        const HostTables& h = tables_->host;
        T_.rank_mode = h.split ? 2 : (h.lrank_u16 ? 1 : 0);
        T_.tl = h.split ? static_cast<const uint16_t*>(tables_->tl.get()) : nullptr;
        T_.th = h.split ? static_cast<const uint32_t*>(tables_->th.get()) : nullptr;
        T_.NL2 = h.NL2;
        T_.nc = h.nc;
        T_.KB = h.KB;
        T_.ntl = h.split ? (1 << h.NL2) : 0;
        T_.nth = h.split ? (int32_t)((1 << (h.NL - h.NL2)) * h.nc) : 0;
        T_.upm2 = h.upm2;
        T_.dnm2 = h.dnm2;
    }
    T_.blk_waves = 8;
    T_.blk_rows = 64;
    T_.apply_waves = 16;
    // one XCD segment each for the single-vector apply (its x mostly fits the
    // caches); one front for the block apply, whose 512 B rows do not
    T_.sched = 0;
    T_.sched_block = 1;
    table_bytes_ = tables_->tiles.bytes() + tables_->hoff.bytes() + tables_->lrank.bytes() + tables_->lolist.bytes() +
                   br_.bytes() + brk_.bytes() + dg_.bytes();
}

namespace {

double time_engine(const Engine& E, const double* x, double* y, int nvec, hipStream_t s, int reps)
{
    auto run = [&]() {
        if (nvec == 1) E.apply(x, y, false, s);
        else E.apply_block(x, y, nvec, nvec, s);
    };
    run();   // warm-up
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "event");
    hip_check(hipEventCreate(&e1), "event");
    hip_check(hipEventRecord(e0, s), "record");
    for (int i = 0; i < reps; ++i) run();
    hip_check(hipEventRecord(e1, s), "record");
    hip_check(hipEventSynchronize(e1), "sync");
    float ms = 0.0f;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return (double)ms / reps;
}

} // namespace

std::shared_ptr<const Engine> autotune(const std::shared_ptr<const BasisSpec>& basis, Kind kind, const OpSum& ops,
                                       const std::shared_ptr<const Engine>& current, int nvec, bool allow_split,
                                       hipStream_t s, TuneResult& out)
{
    constexpr int kKnobs = 5;
    if (nvec < 1) throw std::invalid_argument("autotune: nvec must be positive");
    if (nvec > 1 && !current->is_real()) throw std::invalid_argument("autotune: block apply needs a real Hamiltonian");
    const size_t D = (size_t)current->dim();
    const bool cplx1 = !current->is_real();
    out = TuneResult{};
    if (D == 0) return current;
    const size_t n = D * (size_t)nvec * (cplx1 ? 2 : 1);
    DevBuf x(n * sizeof(double)), y(n * sizeof(double));
    lcheck(qkl_fill_random(0, n, 11, 0, 1, 1, 0, x.d(), s), "fill");
    const int reps = D > ((size_t)1 << 24) ? 2 : 5;
    auto timed = [&](const Engine& E) {
        if (cplx1) {
            E.apply(x.d(), y.d(), true, s);
            hipEvent_t e0, e1;
            hip_check(hipEventCreate(&e0), "event");
            hip_check(hipEventCreate(&e1), "event");
            hip_check(hipEventRecord(e0, s), "record");
            for (int i = 0; i < reps; ++i) E.apply(x.d(), y.d(), true, s);
            hip_check(hipEventRecord(e1, s), "record");
            hip_check(hipEventSynchronize(e1), "sync");
            float ms = 0.0f;
            hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            return (double)ms / reps;
        }
        return time_engine(E, x.d(), y.d(), nvec, s, reps);
    };
    out.ms_before = timed(*current);
    double best_ms = out.ms_before;
    int best_NL = current->NL(), best_tune[kKnobs];
    for (int k = 0; k < kKnobs; ++k) best_tune[k] = current->tuning(k);
    bool improved = false;
    std::vector<int> splits{current->NL()};
    if (allow_split) {
        // around the current split, and the wide splits ranked by two-level
        // tables (more branches inside the low bits, fewer far gathers)
        const int nb = basis->nbits(), m = current->NL();
        const int stepNL = basis->spinless() ? 1 : 2;
        std::vector<int> cand;
        for (int d = -2; d <= 2; ++d) cand.push_back(m + d * stepNL);
        for (int w : {16, 18, 20}) cand.push_back(w);
        for (int NL : cand) {
            if (NL < 4 || NL > std::min(nb, 20) || nb - NL > 28) continue;
            if (!basis->spinless() && (NL & 1)) continue;
            if (std::find(splits.begin(), splits.end(), NL) == splits.end()) splits.push_back(NL);
        }
    }
    const int knob_w = nvec == 1 ? 2 : 3, knob_s = nvec == 1 ? 0 : 1;
    auto try_shape = [&](const std::shared_ptr<Engine>& E, int NL) {
        double ms;
        try {
            ms = timed(*E);
        } catch (const std::exception&) {
            return;   // a launch shape that does not fit this split
        }
        ++out.tried;
        if (ms < best_ms) {
            best_ms = ms;
            best_NL = NL;
            for (int k = 0; k < kKnobs; ++k) best_tune[k] = E->tuning(k);
            improved = true;
        }
    };
    auto make = [&](int NL) -> std::shared_ptr<Engine> {
        EngineOptions eo;
        eo.NL = NL;
        try {
            auto E = std::make_shared<Engine>(basis, kind, ops, eo);
            for (int k = 0; k < kKnobs; ++k) E->set_tuning(k, current->tuning(k));
            return E;
        } catch (const std::exception&) {
            return nullptr;   // a split whose tables cannot be built is not a candidate
        }
    };
    // stage 1: split x waves x schedule
    for (int NL : splits) {
        auto E = make(NL);
        if (!E) continue;
        for (int w : {8, 16}) {
            for (int sc : {0, 1, 2}) {
                E->set_tuning(knob_w, w);
                E->set_tuning(knob_s, sc);
                try_shape(E, NL);
            }
        }
    }
    // stage 2 (block apply): narrower work items, fewer waves, at the best split
    if (nvec > 1) {
        auto E = make(best_NL);
        if (E) {
            for (int k = 0; k < kKnobs; ++k) E->set_tuning(k, best_tune[k]);
            for (int rows : {32, 16}) {
                for (int w : {4, 8}) {
                    for (int sc : {1, 2}) {
                        E->set_tuning(4, rows);
                        E->set_tuning(knob_w, w);
                        E->set_tuning(knob_s, sc);
                        try_shape(E, best_NL);
                    }
                }
            }
        }
    }
    std::shared_ptr<const Engine> best = current;
    if (improved) {
        EngineOptions eo;
        eo.NL = best_NL;
        auto E = std::make_shared<Engine>(basis, kind, ops, eo);
        for (int k = 0; k < kKnobs; ++k) E->set_tuning(k, best_tune[k]);
        best = E;
    }
    out.ms = best_ms;
    out.NL = best->NL();
    out.sched = best->tuning(0);
    out.blk_sched = best->tuning(1);
    out.apply_waves = best->tuning(2);
    out.blk_waves = best->tuning(3);
    out.blk_rows = best->tuning(4);
    return best;
}

int Engine::tuning(int knob) const
{
    if (knob == 5 || knob == 6) {
        const SplitEngines* sp = split_engines(0);
        if (!sp) throw std::invalid_argument("this model admits no split apply");
        return sp->tuning(knob == 5 ? 0 : 2);
    }
    switch (knob) {
    case 0: return T_.sched;
    case 1: return T_.sched_block;
    case 2: return T_.apply_waves;
    case 3: return T_.blk_waves;
    case 4: return T_.blk_rows;
    default: throw std::invalid_argument("unknown tuning knob");
    }
}

void Engine::set_tuning(int knob, int value) const
{
    if (knob == 5 || knob == 6) {
        const SplitEngines* sp = split_engines(0);
        if (!sp) throw std::invalid_argument("this model admits no split apply");
        sp->set_tuning(knob == 5 ? 0 : 2, value);
        return;
    }
    const bool sched_ok = value >= 0 && value <= 2;
    const bool waves_ok = value == 4 || value == 8 || value == 16;
    const bool rows_ok = value == 16 || value == 32 || value == 64;
    switch (knob) {
    case 0: if (!sched_ok) break; T_.sched = value; return;
    case 1: if (!sched_ok) break; T_.sched_block = value; return;
    case 2: if (!waves_ok) break; T_.apply_waves = value; return;
    case 3: if (!waves_ok) break; T_.blk_waves = value; return;
    case 4: if (!rows_ok) break; T_.blk_rows = value; return;
    default: throw std::invalid_argument("unknown tuning knob");
    }
    throw std::invalid_argument("tuning value out of range");
}

void Engine::apply(const double* x, double* y, bool cplx, hipStream_t s) const
{
    if (!cplx && !is_real()) throw std::invalid_argument("complex Hamiltonian needs complex vectors");
    lcheck(qkl_apply(&T_, cplx ? 1 : 0, wide() ? 1 : 0, x, y, s), "apply");
}

void Engine::apply_cols(const double* X, double* Y, int nv, const double* dscale, hipStream_t s) const
{
    if (cols_width(nv) != nv) throw std::invalid_argument("column block width must be 1, 2, 4 or 8");
    if (!is_real() || wide()) throw std::invalid_argument("column blocks need a real Hamiltonian on <= 32 bits");
    if (T_.signs != 0) throw std::invalid_argument("column blocks: operators with parity masks are not supported");
    lcheck(qkl_apply_cols(&T_, 0, nv, X, Y, dscale, -1, nullptr, nullptr, nullptr, 0, nullptr, 0, s), "apply_cols");
}

void Engine::apply_block(const double* X, double* Y, int ld, int ncols, hipStream_t s) const
{
    if (!is_real()) throw std::invalid_argument("block apply needs a real Hamiltonian");
    for (int c0 = 0; c0 < ncols; c0 += 64) {
        const int nc = std::min(64, ncols - c0);
        lcheck(qkl_apply_block(&T_, wide() ? 1 : 0, X, Y, ld, c0, nc, -1, nullptr, nullptr, 0, nullptr, 0, s),
               "apply_block");
    }
}

void Engine::diagonal(double* d, bool cplx, hipStream_t s) const
{
    lcheck(qkl_diagonal(&T_, cplx ? 1 : 0, wide() ? 1 : 0, d, s), "diagonal");
}

const SplitEngines* Engine::split_engines(int64_t min_rows) const
{
    if (dim() < min_rows || !is_real() || kind_ != Kind::SpinHalf || basis_->kind() != Kind::SpinHalf) return nullptr;
    std::lock_guard<std::mutex> lk(split_mu_);
    if (!split_tried_) {
        split_tried_ = true;
        SplitPlan plan;
        // units of at most 21 sites: C(21, 10) rows = 2.8 MB of a vector
        if (find_split(*basis_, ops_, 21, plan)) split_ = std::make_shared<SplitEngines>(*this, std::move(plan));
    }
    return split_.get();
}

void Engine::apply_diag_operator(const double* x, bool xcplx, double* y, hipStream_t s) const
{
    if (!comp_.diag_only) throw std::invalid_argument("operator is not diagonal");
    lcheck(qkl_apply_diag_terms(&T_, wide() ? 1 : 0, xcplx ? 1 : 0, x, y, s), "apply_diag_operator");
}

// ---------------------------------------------------------------- tridiagonal

double tridiag_min(const double* alpha, const double* beta, int n, std::vector<double>* evec)
{
    if (n == 0) {
        if (evec) evec->clear();
        return 0.0;
    }
    if (!evec) return symtri_lowest(alpha, beta, n);
    std::vector<double> w(n), z((size_t)n * n);
    symtri_eig(alpha, beta, n, w.data(), z.data(), kAllVectors);
    evec->resize(n);
    for (int i = 0; i < n; ++i) (*evec)[i] = z[(size_t)i * n];
    return w[0];
}

// ---------------------------------------------------------------- helpers

namespace {

// <x, y> (complex conj-linear when cplx) -> host
std::complex<double> dot_host(bool cplx, size_t n, const double* x, const double* y, double* part, double* out,
                              hipStream_t s)
{
    lcheck(qkl_dot_partials(cplx ? 1 : 0, n, x, y, part, s), "dot");
    lcheck(qkl_reduce(part, qkl_blas_partials(), 1, 1, cplx ? 2 : 0, cplx ? 1 : 0, out, 0, 2, s), "reduce");
    double v[2] = {0, 0};
    hip_check(hipMemcpyAsync(v, out, (cplx ? 2 : 1) * sizeof(double), hipMemcpyDeviceToHost, s), "d2h");
    hip_check(hipStreamSynchronize(s), "sync");
    return {v[0], cplx ? v[1] : 0.0};
}

} // namespace

void LanczosReplay::feed(double a, double b)
{
    const int it = next++;
    alphas.push_back(a);
    if (b < 1e-15) {
        converged = true;
        done = true;
        return;
    }
    if (it + 1 == kmax) {
        done = true;
        return;
    }
    betas.push_back(b);
    if (!fixed && it > 0) {
        const double e = tridiag_min(alphas.data(), betas.data(), (int)alphas.size(), nullptr);
        if (std::abs(e - e_old) < tol) {
            e_old = e;
            converged = true;
            done = true;
            return;
        }
        e_old = e;
    }
}

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

RealCore::RealCore(const Engine& e, int kcap_, hipStream_t st, bool reorth, bool so, Arena* ar,
                   const SplitEngines* sp)
    : E(e), s(st), D((size_t)e.dim()), kcap(kcap_), arena(ar), stream_ordered(so), split(sp)
{
    auto mk = [&](size_t bytes) { return ar ? ar->take(bytes) : (so ? DevBuf(bytes, st) : DevBuf(bytes)); };
    for (auto& q : Q) q = mk(D * sizeof(double));
    if (split) tmp = mk(D * sizeof(double));
    // alpha[kcap + 2] and nrm[kcap + 3] in one allocation: one copy reads both back
    alpha = mk((size_t)(2 * kcap + 5) * sizeof(double));
    nrm = DevBuf::view(alpha.d() + kcap + 2, (size_t)(kcap + 3) * sizeof(double));
    npA = split ? qkl_apply_partials_acc(split->tables2()) : qkl_apply_partials(E.tables(), E.wide() ? 1 : 0);
    if (npA <= 0) throw std::runtime_error("apply: the operator tables do not fit in LDS");
    NB = qkl_blas_partials();
    partA = mk(npA * sizeof(double));
    partB = mk(NB * sizeof(double));
    if (reorth) {
        hcoef = mk((kcap + 2) * sizeof(double));
        partM = mk((size_t)NB * (kcap + 2) * sizeof(double));
    }
}

int cols_width(int n)
{
    return n <= 1 ? 1 : n <= 2 ? 2 : n <= 4 ? 4 : n <= 8 ? 8 : 0;
}

ColsCore::ColsCore(const Engine& e, int nv_, int kcap_, const double* dscale, hipStream_t st, bool so)
    : E(e), s(st), D((size_t)e.dim()), nv(nv_), kcap(kcap_)
{
    if (cols_width(nv) != nv) throw std::invalid_argument("column block width must be 1, 2, 4 or 8");
    if (!E.is_real() || E.wide()) throw std::invalid_argument("column blocks need a real Hamiltonian on <= 32 bits");
    if (E.tables()->signs != 0)
        throw std::invalid_argument("column blocks: operators with parity masks are not supported");
    auto mk = [&](size_t bytes) { return so ? DevBuf(bytes, st) : DevBuf(bytes); };
    for (auto& q : Q) q = mk(D * nv * sizeof(double));
    alpha = mk((size_t)(kcap + 2) * nv * sizeof(double));
    nrm = mk((size_t)(kcap + 3) * nv * sizeof(double));
    npA = qkl_cols_partials(E.tables());
    if (npA <= 0) throw std::runtime_error("column apply: the operator tables do not fit in LDS");
    NB = qkl_cols_update_partials(D);
    partA = mk((size_t)npA * nv * sizeof(double));
    partB = mk((size_t)NB * nv * sizeof(double));
    dsc = mk(nv * sizeof(double));
    zero = mk(nv * sizeof(double));
    one = mk(nv * sizeof(double));
    // the per-column scalars as kernel arguments: a copy from pageable host
    // memory would make the host wait for the stream's hardware queue
    std::vector<double> z(nv, 0.0), o(nv, 1.0);
    lcheck(qkl_set_vals(dsc.d(), nv, dscale, s), "set diagonal scales");
    lcheck(qkl_set_vals(zero.d(), nv, z.data(), s), "set zeros");
    lcheck(qkl_set_vals(one.d(), nv, o.data(), s), "set ones");
}

void ColsCore::init_norm()
{
    // an update with coefficient 0 / 1 leaves cur as it is and yields the
    // partial sums of its squared column norms, which step 0 totals
    lcheck(qkl_update_cols(nv, D, cur(), cur(), zero.d(), one.d(), 0, nullptr, 0, partB.d(), s), "norm0");
}

void ColsCore::flush(int k)
{
    lcheck(qkl_total_cols(nv, partB.d(), NB, nrm.d(), k, s), "total norms");
}

void ColsCore::step(int k)
{
    if (k < 0 || k > kcap) throw std::logic_error("column recurrence: step beyond the capacity");
    double* w = Q[iw].d();
    // the apply totals nrm[k] from partB, the update alpha[k] from partA
    lcheck(qkl_apply_cols(E.tables(), 0, nv, cur(), w, dsc.d(), k, prv(), nrm.d(), partA.d(), npA, partB.d(), NB, s),
           "apply_cols");
    lcheck(qkl_update_cols(nv, D, w, cur(), alpha.d(), nrm.d(), k, partA.d(), npA, partB.d(), s), "update_cols");
    const int t = ip;
    ip = ic;
    ic = iw;
    iw = t;
}

RealCore::~RealCore()
{
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
}

void RealCore::steps_graph(int k, int n)
{
    if (n % 3 != 0 || k + n > kcap) throw std::logic_error("graph steps: n must be a multiple of 3 within the cap");
    if (!split && qkl_apply_partials(E.tables(), E.wide() ? 1 : 0) != npA)
        throw std::logic_error("apply launch shape changed during a run");
    if (gexec && ic != cap_ic) {
        // eager steps moved the rotation: capture again for the current roles
        (void)hipGraphExecDestroy(gexec);
        (void)hipGraphDestroy(graph);
        gexec = nullptr;
        graph = nullptr;
    }
    if (!gexec) {
        if (kdev.bytes() == 0)
            kdev = arena ? arena->take(2 * sizeof(int))
                         : (stream_ordered ? DevBuf(2 * sizeof(int), s) : DevBuf(2 * sizeof(int)));
        cap_ic = ic;
        hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
        int rc = 0;
        // six steps, two turns of the vector rotation: steps 0-2 read
        // counter 0 and the third update stores counter 1 = counter 0 + 3;
        // steps 3-5 read counter 1 and the sixth update stores counter 0 for
        // the next replay. No launch writes a counter it reads.
        int* kd = static_cast<int*>(kdev.get());
        for (int i = 0; i < 6 && rc == 0; ++i) {
            int* kc = kd + i / 3;
            int* kn = (i % 3 == 2) ? kd + (1 - i / 3) : nullptr;
            rc |= lanczos_apply(0, kc, i % 3);
            rc |= qkl_lanczos_update1(D, wrk(), cur(), alpha.d(), nrm.d(), 0, kc, i % 3, partA.d(), npA, partB.d(), kn,
                                      host_mirror, s);
            const int t = ip;
            ip = ic;
            ic = iw;
            iw = t;
        }
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &g);
        if (rc != 0 || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            throw std::runtime_error("capturing the Lanczos step graph failed");
        }
        graph = g;
        hip_check(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0), "graph instantiate");
    }
    // the counter is set by a one-thread kernel outside the graph (no host
    // staging, so the caller never waits); six rotations leave the vector
    // roles as they were; a remaining turn of three steps runs eagerly
    lcheck(qkl_set_int(static_cast<int*>(kdev.get()), k, s), "set step");
    for (int r = 0; r < n / 6; ++r) hip_check(hipGraphLaunch(gexec, s), "graph launch");
    pending = k + n / 6 * 6;
    for (int i = k + n / 6 * 6; i < k + n; ++i) step(i, nullptr);
}

void RealCore::init_norm()
{
    lcheck(qkl_fill_zero(D, prv(), s), "zero");
    lcheck(qkl_dot_partials(0, D, cur(), cur(), partB.d(), s), "dot");
    lcheck(qkl_total(partB.d(), NB, 1, nrm.d(), 0, s), "total");
    if (host_mirror) lcheck(qkl_total(partB.d(), NB, 1, host_mirror + kcap + 2, 0, s), "total");
    pending = -1;
}

void RealCore::flush()
{
    if (pending >= 0) {
        lcheck(qkl_total(partB.d(), NB, 1, nrm.d(), pending, s), "total");
        if (host_mirror) lcheck(qkl_total(partB.d(), NB, 1, host_mirror + kcap + 2, pending, s), "total");
    }
    pending = -1;
}

// one step k: nrm[k] from the previous partials; w = H v_k - beta v_{k-1};
// alpha_k; q_{k+1} and the partial sums of |q_{k+1}|^2 (nrm[k+1] is totalled
// by the next step's apply, or by flush())
int RealCore::lanczos_apply(int k, const int* kdev, int koff)
{
    if (split) {
        const int rc = qkl_apply(split->tables1(), 0, 0, cur(), tmp.d(), s);
        if (rc != 0) return rc;
        return qkl_lanczos_apply_acc(split->tables2(), cur(), wrk(), k, kdev, koff, prv(), nrm.d(), partB.d(), NB,
                                     partA.d(), npA, tmp.d(), host_mirror ? host_mirror + kcap + 2 : nullptr, s);
    }
    return qkl_lanczos_apply(E.tables(), E.wide() ? 1 : 0, cur(), wrk(), k, kdev, koff, prv(), nrm.d(), partB.d(), NB,
                             partA.d(), npA, host_mirror ? host_mirror + kcap + 2 : nullptr, s);
}

void RealCore::step(int k, const double* basis)
{
    // the partial buffer was sized for the launch shape of the tuning at
    // construction; a retuned engine would write past it
    if (!split && qkl_apply_partials(E.tables(), E.wide() ? 1 : 0) != npA)
        throw std::logic_error("apply launch shape changed during a run");
    lcheck(lanczos_apply(k, nullptr, 0), "apply");
    lcheck(qkl_lanczos_update1(D, wrk(), cur(), alpha.d(), nrm.d(), k, nullptr, 0, partA.d(), npA, partB.d(), nullptr,
                               host_mirror, s),
           "update");
    if (basis) {
        // classical Gram-Schmidt against v_0..v_k (stored normalised)
        lcheck(qkl_multidot(D, k + 1, basis, D, wrk(), partM.d(), s), "multidot");
        lcheck(qkl_reduce(partM.d(), NB, k + 1, 1, 0, 0, hcoef.d(), 0, 0, s), "reduce h");
        lcheck(qkl_multiaxpy(D, k + 1, basis, D, hcoef.d(), -1.0, wrk(), s), "multiaxpy");
        lcheck(qkl_dot_partials(0, D, wrk(), wrk(), partB.d(), s), "dot");
    }
    pending = k + 1;
    const int t = ip;
    ip = ic;
    ic = iw;
    iw = t;
}

// ---------------------------------------------------------------- Lanczos (real, fused)

static LanczosOutput lanczos_complex(const Engine& E, const LanczosOptions& o, hipStream_t s);

void reference_start_vector(size_t D, uint64_t seed, int start_mode, double* out)
{
    if (start_mode != 1 && start_mode != 2) throw std::invalid_argument("reference start: start_mode must be 1 or 2");
    // the library's own engine and distribution, as the reference declares
    // them; the two draws of an entry go to (imag, real) or (real, imag)
    std::mt19937 rng(static_cast<std::mt19937::result_type>(seed));
    std::uniform_real_distribution<double> dist(-1.0, 1.0);
    const size_t first = start_mode == 1 ? 1 : 0;
    for (size_t i = 0; i < D; ++i) {
        const double a = dist(rng);
        const double b = dist(rng);
        out[2 * i + first] = a;
        out[2 * i + (1 - first)] = b;
    }
}

ReferenceNormal::ReferenceNormal(uint64_t seed, int start_mode)
    : rng_(static_cast<std::mt19937::result_type>(seed)), dist_(0.0, 1.0), first_(start_mode == 1 ? 1 : 0)
{
}

void ReferenceNormal::fill(size_t n, double* out)
{
    for (size_t i = 0; i < n; ++i) {
        const double a = dist_(rng_);
        const double b = dist_(rng_);
        out[2 * i + first_] = a;
        out[2 * i + (1 - first_)] = b;
    }
}

size_t lanczos_arena_bytes(const Engine& E, const LanczosOptions& o)
{
    const size_t D = (size_t)E.dim();
    const int kmax = (int)std::min<int64_t>(o.maxiter, (int64_t)D);
    if (D == 0 || kmax <= 0) return 0;
    auto al = [](size_t b) { return (std::max<size_t>(b, 8) + 255) & ~(size_t)255; };
    const size_t npA = (size_t)std::max(qkl_apply_partials(E.tables(), E.wide() ? 1 : 0), 1);
    const size_t NB = (size_t)qkl_blas_partials();
    size_t core = 3 * al(D * sizeof(double)) + al((kmax + 2) * sizeof(double)) + al((kmax + 3) * sizeof(double)) +
                  al(npA * sizeof(double)) + al(NB * sizeof(double));
    size_t basis = 0;
    if (o.reorth) {
        basis = al((size_t)(kmax + 1) * D * sizeof(double));
        core += al((kmax + 2) * sizeof(double)) + al(NB * (kmax + 2) * sizeof(double));
    }
    size_t ritz = 0;
    if (o.want_vector) {
        ritz = al(D * sizeof(double)) + al(2 * sizeof(double)) + al(NB * 2 * sizeof(double));
        if (o.reorth) ritz += al((size_t)(kmax + 1) * sizeof(double));
        else ritz += 3 * al(D * sizeof(double)) + al((kmax + 2) * sizeof(double)) + al((kmax + 3) * sizeof(double)) +
                     al(npA * sizeof(double)) + al(NB * sizeof(double));
    }
    return basis + std::max(core, ritz);
}

// Pinned host memory for the Lanczos driver's read-backs, kept per host
// thread and only grown, at the start of a run (freeing pinned memory
// synchronises the whole device, so it is never freed while a run is on it)
static double* pinned_readback(size_t n)
{
    thread_local double* p = nullptr;
    thread_local size_t cap = 0;
    if (n > cap) {
        double* q = nullptr;
        // coherent: kernels store into it directly
        hip_check(hipHostMalloc((void**)&q, n * sizeof(double), hipHostMallocCoherent), "hipHostMalloc");
        if (p) (void)hipHostFree(p);
        p = q;
        cap = n;
    }
    return p;
}

LanczosOutput lanczos_run(const Engine& E, const LanczosOptions& o, hipStream_t s, Arena* arena)
{
    LanczosOutput out;
    const int64_t D = E.dim();
    if (D == 0) return out;
    if (o.start_mode < 0 || o.start_mode > 2) throw std::invalid_argument("start_mode must be 0, 1 or 2");
    if (o.start_mode != 0) {
        // the reference's complex start: complex vectors for real H as well
        if (arena) throw std::invalid_argument("the workspace Lanczos driver runs real vectors only (start_mode 0)");
        if (o.maxiter <= 0) return out;
        std::vector<double> st;
        LanczosOptions oc = o;
        if (!o.start && !o.start_dev) {
            st.resize(2 * (size_t)D);
            reference_start_vector((size_t)D, o.seed, o.start_mode, st.data());
            oc.start = st.data();
        }
        return lanczos_complex(E, oc, s);
    }
    if (!E.is_real()) {
        if (arena) throw std::invalid_argument("the workspace Lanczos driver needs a real Hamiltonian");
        return lanczos_complex(E, o, s);
    }
    const int kmax = (int)std::min<int64_t>(o.maxiter, D);
    if (kmax <= 0) return out;
    const bool reorth = o.reorth;
    const size_t basis_bytes = (size_t)(kmax + 1) * (size_t)D * sizeof(double);
    bool store = reorth;
    if (arena) {
        if (lanczos_arena_bytes(E, o) > arena->cap - std::min(arena->cap, arena->used))
            throw std::length_error("Lanczos needs " + std::to_string(lanczos_arena_bytes(E, o)) +
                                    " bytes of workspace, " + std::to_string(arena->cap - arena->used) + " available");
    } else {
        size_t free_b = 0, total_b = 0;
        hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
        store = reorth || (o.want_vector && basis_bytes < free_b / 3);
        if (reorth && basis_bytes > (free_b / 10) * 9)
            throw std::runtime_error("full reorthogonalisation needs more device memory than is free");
    }
    auto mk = [&](size_t bytes) { return arena ? arena->take(bytes) : DevBuf(bytes); };
    // the two-pass apply keeps every Krylov vector in its internal row
    // order: a given start vector is permuted in, the Ritz vector out; a
    // random start is drawn directly in that order
    // (automatic = the single pass: on the J1-J2 4x8 and 32-site-ring
    // workloads the two passes measured slower, 62 vs 53 ms and 25 vs 16 ms)
    const SplitEngines* sp = nullptr;
    if (!arena && o.split > 0) sp = E.split_engines(0);
    DevBuf basis;
    if (store) basis = mk(basis_bytes);
    const size_t core_mark = arena ? arena->mark() : 0;
    // launch-bound sizes running a fixed number of steps replay a captured
    // six-step graph; runs to a tolerance launch eagerly, ahead of the
    // host's checks (below), which keeps the GPU as busy without the
    // capture's fixed cost
    const bool graph = !arena && !store && !reorth && kmax >= 3 &&
                       (o.graph > 0 || (o.graph < 0 && o.fixed_steps && D < (1 << 22)));
    // host convergence checks: every 4 steps up to 2^25 rows, every step
    // beyond; the replay fixes the iteration count whatever the cadence.
    // A check costs an event (about 3 us, a fifth of a step of the 20-site
    // ring) and a sparse one runs more steps past convergence: that ring to
    // 1e-12 took 1.12, 1.04, 1.04, 1.02 and 1.10 ms checking every 1, 2, 3,
    // 4 and 6 steps
    int check = o.check_every > 0 ? o.check_every : (D < (1 << 25) ? (graph ? 12 : 4) : 1);
    if (o.fixed_steps) check = kmax;
    // runs to a tolerance check while the device runs ahead; the kernels
    // that total alpha and the norms store them into pinned host memory as
    // well, so a check reads them after an event, with no copy
    const bool ahead = !(arena || o.fixed_steps || kmax <= check);
    double* hb = ahead ? pinned_readback(2 * (size_t)kmax + 6) : nullptr;
    auto Cp = std::make_unique<RealCore>(E, kmax, s, reorth, false, arena, sp);
    RealCore& C = *Cp;
    C.host_mirror = hb;
    out.split = sp != nullptr;

    // dst: the recurrence's start; stage: a scratch vector for the permutation
    auto load_start = [&](double* dst, double* stage) {
        double* tgt = (sp && (o.start_dev || o.start)) ? stage : dst;
        if (o.start_dev) hip_check(hipMemcpyAsync(tgt, o.start_dev, D * sizeof(double), hipMemcpyDeviceToDevice, s), "start");
        else if (o.start) hip_check(hipMemcpyAsync(tgt, o.start, D * sizeof(double), hipMemcpyHostToDevice, s), "start");
        else lcheck(qkl_fill_random(0, (size_t)D, o.seed, 0, 1, 1, 0, dst, s), "random start");
        if (tgt != dst) sp->to_internal(tgt, dst, s);
    };
    load_start(C.cur(), C.wrk());
    C.init_norm();
    if (store) lcheck(qkl_copy_scaled((size_t)D, C.cur(), basis.d(), C.nrm.d(), 0, s), "store v0");

    LanczosReplay rp{kmax, o.tol, o.fixed_steps, {}, {}};
    const double t0 = now_s();
    int k = 0;
    std::vector<double> hnorm(kmax + 3);   // the norms read back so far (the Ritz pass divides by them)
    // enqueue the steps up to the next check and one read-back of alpha and
    // the norms so far into hbuf (alpha at 0, the norms at kmax + 2, as on
    // the device); returns the steps whose alpha and following norm it holds
    auto enqueue_chunk = [&](double* hbuf) {
        const int kend = std::min(kmax, k + check);
        if (graph && kend - k >= 3) {
            const int n3 = (kend - k) / 3 * 3;
            C.steps_graph(k, n3);
            k += n3;
        }
        for (; k < kend; ++k) {
            C.step(k, reorth ? basis.d() : nullptr);
            if (store) {
                C.flush();
                lcheck(qkl_copy_scaled((size_t)D, C.cur(), basis.d() + (size_t)(k + 1) * D, C.nrm.d(), k + 1, s),
                       "store v");
            }
        }
        // the norm of the newest vector is totalled by the next chunk's
        // first apply, which the replay waits for; only the last chunk
        // totals it here (a total and a second copy per check cost as much
        // as a step of the 20-site ring)
        if (k >= kmax) C.flush();
        if (!C.host_mirror)
            hip_check(hipMemcpyAsync(hbuf, C.alpha.d(), (size_t)(kmax + 2 + k + 1) * sizeof(double),
                                     hipMemcpyDeviceToHost, s),
                      "d2h");
        return C.pending >= 0 ? C.pending - 1 : k;
    };
    if (!ahead) {
        // one check at a time (no pinned read-back buffers: the workspace
        // entry point allocates nothing)
        std::vector<double> hbuf(2 * (size_t)kmax + 6);
        while (!rp.done) {
            const int avail = enqueue_chunk(hbuf.data());
            hip_check(hipStreamSynchronize(s), "sync");
            const double* ha = hbuf.data();
            const double* hn = ha + kmax + 2;
            std::copy(hn, hn + avail + 1, hnorm.begin());
            while (!rp.done && rp.next < avail) rp.feed(ha[rp.next], hn[rp.next + 1]);
            if (k >= kmax) rp.done = true;
        }
    } else {
        // the device runs one check ahead: while the host replays the
        // convergence test on chunk i, chunk i + 1 is already queued, so the
        // GPU does not idle through the replay; steps past convergence are
        // run and not used (steps_run counts them). The kernels store alpha
        // and the norms into hb as they total them; the host reads only
        // entries that its event orders before it.
        const double* ha = hb;
        const double* hn = ha + kmax + 2;
        hipEvent_t ev[2] = {nullptr, nullptr};
        struct Events {
            hipEvent_t* e;
            ~Events()
            {
                for (int i = 0; i < 2; ++i)
                    if (e[i]) (void)hipEventDestroy(e[i]);
            }
        } ev_guard{ev};
        for (auto& e : ev) hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
        int pending[2] = {-1, -1}, avail[2] = {0, 0};
        auto enqueue_into = [&](int b) {
            avail[b] = enqueue_chunk(hb);
            hip_check(hipEventRecord(ev[b], s), "event");
            pending[b] = k;
        };
        enqueue_into(0);
        if (k < kmax) enqueue_into(1);
        int b = 0;
        while (!rp.done) {
            if (pending[b] < 0) throw std::logic_error("lanczos: no read-back in flight");
            hip_check(hipEventSynchronize(ev[b]), "event sync");
            const int kb = pending[b];
            pending[b] = -1;
            std::copy(hn, hn + avail[b] + 1, hnorm.begin());
            while (!rp.done && rp.next < avail[b]) rp.feed(ha[rp.next], hn[rp.next + 1]);
            if (kb >= kmax) rp.done = true;
            if (rp.done) break;
            if (k < kmax) enqueue_into(b);
            b ^= 1;
        }
        hip_check(hipStreamSynchronize(s), "sync");   // the chunk run ahead
    }
    out.loop_seconds = now_s() - t0;
    out.steps_run = k;
    const int n = (int)rp.alphas.size();
    out.iterations = rp.next;
    out.converged = rp.converged;
    out.alphas = rp.alphas;
    out.betas = rp.betas;
    std::vector<double> z;
    out.energy = tridiag_min(rp.alphas.data(), rp.betas.data(), n, o.want_vector ? &z : nullptr);
    // the recurrence's buffers are done with: the Ritz pass reuses their room
    hip_check(hipStreamSynchronize(s), "sync");
    Cp.reset();
    if (arena) arena->rewind(core_mark);

    if (o.want_vector) {
        const double t1 = now_s();
        DevBuf ritz = mk(D * sizeof(double));
        lcheck(qkl_fill_zero((size_t)D, ritz.d(), s), "zero");
        if (store) {
            DevBuf zc = mk(n * sizeof(double));
            hip_check(hipMemcpyAsync(zc.get(), z.data(), n * sizeof(double), hipMemcpyHostToDevice, s), "h2d");
            lcheck(qkl_multiaxpy((size_t)D, n, basis.d(), (size_t)D, zc.d(), 1.0, ritz.d(), s), "ritz");
            hip_check(hipStreamSynchronize(s), "sync");
        } else {
            // second pass: rerun the same deterministic recurrence and
            // accumulate sum_k z_k v_k (v_k = q_k / nrm_k)
            RealCore R(E, n, s, false, false, arena, sp);
            load_start(R.cur(), R.wrk());
            R.init_norm();
            for (int kk = 0; kk < n; ++kk) {
                lcheck(qkl_axpy(0, (size_t)D, z[kk] / hnorm[kk], 0.0, R.cur(), ritz.d(), s), "axpy");
                if (kk == n - 1) break;
                R.step(kk, nullptr);
            }
            hip_check(hipStreamSynchronize(s), "sync");
        }
        DevBuf tmp = mk(2 * sizeof(double));
        DevBuf part = mk(qkl_blas_partials() * 2 * sizeof(double));
        const double nr = std::sqrt(dot_host(false, (size_t)D, ritz.d(), ritz.d(), part.d(), tmp.d(), s).real());
        if (nr != 0.0) lcheck(qkl_scal(0, (size_t)D, 1.0 / nr, 0.0, ritz.d(), s), "scal");
        const double* vout = ritz.d();
        DevBuf pubv;
        if (sp) {
            pubv = mk(D * sizeof(double));
            sp->to_public(ritz.d(), pubv.d(), s);
            vout = pubv.d();
        }
        out.vector.resize(D);
        hip_check(hipMemcpyAsync(out.vector.data(), vout, D * sizeof(double), hipMemcpyDeviceToHost, s), "d2h");
        hip_check(hipStreamSynchronize(s), "sync");
        out.ritz_seconds = now_s() - t1;
    }
    return out;
}

// complex Hamiltonian: reference-shaped loop with host scalars
static LanczosOutput lanczos_complex(const Engine& E, const LanczosOptions& o, hipStream_t s)
{
    LanczosOutput out;
    const size_t D = (size_t)E.dim();
    const int kmax = (int)std::min<int64_t>(o.maxiter, (int64_t)D);
    DevBuf vp(2 * D * sizeof(double)), vc(2 * D * sizeof(double)), w(2 * D * sizeof(double));
    DevBuf part(qkl_blas_partials() * 2 * sizeof(double)), tmp(4 * sizeof(double));
    std::vector<DevBuf> basis;
    if (o.start_dev) hip_check(hipMemcpyAsync(vc.get(), o.start_dev, 2 * D * sizeof(double), hipMemcpyDeviceToDevice, s), "start");
    else if (o.start) hip_check(hipMemcpyAsync(vc.get(), o.start, 2 * D * sizeof(double), hipMemcpyHostToDevice, s), "start");
    else lcheck(qkl_fill_random(0, 2 * D, o.seed, 0, 1, 1, 0, vc.d(), s), "random start");
    lcheck(qkl_fill_zero(2 * D, vp.d(), s), "zero");
    auto dotc = [&](const double* a, const double* b) { return dot_host(true, D, a, b, part.d(), tmp.d(), s); };
    double n0 = std::sqrt(dotc(vc.d(), vc.d()).real());
    if (n0 != 0.0) lcheck(qkl_scal(1, D, 1.0 / n0, 0.0, vc.d(), s), "scal");
    const bool store = o.reorth || o.want_vector;
    auto push_basis = [&](const double* v) {
        basis.emplace_back(2 * D * sizeof(double));
        hip_check(hipMemcpyAsync(basis.back().get(), v, 2 * D * sizeof(double), hipMemcpyDeviceToDevice, s), "copy");
    };
    if (store) push_basis(vc.d());
    LanczosReplay rp{kmax, o.tol, o.fixed_steps, {}, {}};
    const double t0 = now_s();
    for (int it = 0; it < kmax && !rp.done; ++it) {
        E.apply(vc.d(), w.d(), true, s);
        const double a = dotc(vc.d(), w.d()).real();
        lcheck(qkl_axpy(1, D, -a, 0.0, vc.d(), w.d(), s), "axpy");
        if (it > 0) lcheck(qkl_axpy(1, D, -rp.betas.back(), 0.0, vp.d(), w.d(), s), "axpy");
        if (o.reorth)
            for (auto& b : basis) {
                const auto h = dotc(b.d(), w.d());
                lcheck(qkl_axpy(1, D, -h.real(), -h.imag(), b.d(), w.d(), s), "axpy");
            }
        const double b = std::sqrt(dotc(w.d(), w.d()).real());
        rp.feed(a, b);
        if (rp.done) break;
        std::swap(vp, vc);
        std::swap(vc, w);
        lcheck(qkl_scal(1, D, 1.0 / b, 0.0, vc.d(), s), "scal");
        if (store) push_basis(vc.d());
    }
    out.loop_seconds = now_s() - t0;
    const int n = (int)rp.alphas.size();
    out.iterations = rp.next;
    out.steps_run = rp.next;
    out.converged = rp.converged;
    out.alphas = rp.alphas;
    out.betas = rp.betas;
    std::vector<double> z;
    out.energy = tridiag_min(rp.alphas.data(), rp.betas.data(), n, o.want_vector ? &z : nullptr);
    if (o.want_vector) {
        DevBuf ritz(2 * D * sizeof(double));
        lcheck(qkl_fill_zero(2 * D, ritz.d(), s), "zero");
        for (int i = 0; i < n; ++i) lcheck(qkl_axpy(1, D, z[i], 0.0, basis[i].d(), ritz.d(), s), "axpy");
        const double nr = std::sqrt(dotc(ritz.d(), ritz.d()).real());
        if (nr != 0.0) lcheck(qkl_scal(1, D, 1.0 / nr, 0.0, ritz.d(), s), "scal");
        out.vector.resize(2 * D);
        hip_check(hipMemcpyAsync(out.vector.data(), ritz.get(), 2 * D * sizeof(double), hipMemcpyDeviceToHost, s), "d2h");
        hip_check(hipStreamSynchronize(s), "sync");
    }
    return out;
}

// ---------------------------------------------------------------- plain recurrences

TridiagOutput tridiag_run_dev(const Engine& E, const double* start_dev, int n_steps, hipStream_t s, double* seconds)
{
    TridiagOutput out;
    const size_t D = (size_t)E.dim();
    if (D == 0 || n_steps <= 0) return out;
    if (!E.is_real()) throw std::invalid_argument("real recurrence needs a real Hamiltonian");
    RealCore C(E, n_steps, s, false);
    hip_check(hipMemcpyAsync(C.cur(), start_dev, D * sizeof(double), hipMemcpyDeviceToDevice, s), "start");
    hip_check(hipStreamSynchronize(s), "sync");
    const double t0 = now_s();
    C.init_norm();
    for (int k = 0; k < n_steps; ++k) C.step(k, nullptr);
    C.flush();
    std::vector<double> ha(n_steps), hn(n_steps + 1);
    hip_check(hipMemcpyAsync(ha.data(), C.alpha.d(), n_steps * sizeof(double), hipMemcpyDeviceToHost, s), "d2h");
    hip_check(hipMemcpyAsync(hn.data(), C.nrm.d(), (n_steps + 1) * sizeof(double), hipMemcpyDeviceToHost, s), "d2h");
    hip_check(hipStreamSynchronize(s), "sync");
    if (seconds) *seconds = now_s() - t0;
    out.norm0 = hn[0];
    if (hn[0] < 1e-15) return out;
    for (int k = 0; k < n_steps; ++k) {
        out.alphas.push_back(ha[k]);
        if (hn[k + 1] < 1e-15) break;
        out.betas.push_back(hn[k + 1]);
    }
    return out;
}

TridiagOutput tridiag_run(const Engine& E, const double* start, bool start_cplx, int n_steps, hipStream_t s)
{
    TridiagOutput out;
    const size_t D = (size_t)E.dim();
    if (D == 0 || n_steps <= 0) return out;
    if (E.is_real() && !start_cplx) {
        DevBuf st(D * sizeof(double));
        hip_check(hipMemcpyAsync(st.get(), start, D * sizeof(double), hipMemcpyHostToDevice, s), "start");
        return tridiag_run_dev(E, st.d(), n_steps, s, nullptr);
    }
    // complex start (or complex H): host-scalar loop
    DevBuf vp(2 * D * sizeof(double)), vc(2 * D * sizeof(double)), w(2 * D * sizeof(double));
    DevBuf part(qkl_blas_partials() * 2 * sizeof(double)), tmp(4 * sizeof(double));
    if (start_cplx) hip_check(hipMemcpyAsync(vc.get(), start, 2 * D * sizeof(double), hipMemcpyHostToDevice, s), "start");
    else {
        std::vector<double> c(2 * D, 0.0);
        for (size_t i = 0; i < D; ++i) c[2 * i] = start[i];
        hip_check(hipMemcpyAsync(vc.get(), c.data(), 2 * D * sizeof(double), hipMemcpyHostToDevice, s), "start");
        hip_check(hipStreamSynchronize(s), "sync");
    }
    lcheck(qkl_fill_zero(2 * D, vp.d(), s), "zero");
    auto dotc = [&](const double* a, const double* b) { return dot_host(true, D, a, b, part.d(), tmp.d(), s); };
    const double n0 = std::sqrt(dotc(vc.d(), vc.d()).real());
    out.norm0 = n0;
    if (n0 < 1e-15) return out;
    lcheck(qkl_scal(1, D, 1.0 / n0, 0.0, vc.d(), s), "scal");
    for (int it = 0; it < n_steps; ++it) {
        E.apply(vc.d(), w.d(), true, s);
        const double a = dotc(vc.d(), w.d()).real();
        out.alphas.push_back(a);
        lcheck(qkl_axpy(1, D, -a, 0.0, vc.d(), w.d(), s), "axpy");
        if (it > 0) lcheck(qkl_axpy(1, D, -out.betas.back(), 0.0, vp.d(), w.d(), s), "axpy");
        const double b = std::sqrt(dotc(w.d(), w.d()).real());
        if (b < 1e-15) break;
        out.betas.push_back(b);
        std::swap(vp, vc);
        std::swap(vc, w);
        lcheck(qkl_scal(1, D, 1.0 / b, 0.0, vc.d(), s), "scal");
    }
    return out;
}

BlockTridiag block_tridiag_run(const Engine& E, double* X0, int ncol, int group, int n_steps, hipStream_t s,
                               double* seconds)
{
    BlockTridiag out;
    const size_t D = (size_t)E.dim();
    if (ncol <= 0 || ncol > 64) throw std::invalid_argument("block width must be 1..64");
    if (group <= 0 || ncol % group) throw std::invalid_argument("columns must split into whole groups");
    if (!E.is_real()) throw std::invalid_argument("block recurrences need a real Hamiltonian");
    out.ncol = ncol;
    out.group = group;
    const int ng = ncol / group;
    out.alphas.assign(ng, {});
    out.betas.assign(ng, {});
    out.norm0.assign(ng, 0.0);
    if (D == 0 || n_steps <= 0) return out;
    const size_t nel = D * (size_t)ncol;
    DevBuf Pb(nel * sizeof(double)), Wb(nel * sizeof(double));
    double* P = Pb.d();
    double* Cc = X0;
    double* W = Wb.d();
    const int npA = qkl_block_partials(E.tables()), NB = qkl_blas_partials();
    if (npA <= 0) throw std::runtime_error("block recurrence: the operator tables do not fit in LDS");
    DevBuf partA((size_t)npA * ncol * sizeof(double)), partB((size_t)NB * ncol * sizeof(double));
    DevBuf alpha((size_t)(n_steps + 1) * ncol * sizeof(double)), nrm((size_t)(n_steps + 2) * ncol * sizeof(double));
    // initial norms: an update with zero coefficient leaves X unchanged and
    // yields the column sums of squares
    {
        std::vector<double> z0(ncol, 0.0), one(ncol, 1.0);
        DevBuf za(ncol * sizeof(double)), on(ncol * sizeof(double));
        hip_check(hipMemcpyAsync(za.get(), z0.data(), ncol * sizeof(double), hipMemcpyHostToDevice, s), "h2d");
        hip_check(hipMemcpyAsync(on.get(), one.data(), ncol * sizeof(double), hipMemcpyHostToDevice, s), "h2d");
        lcheck(qkl_lanczos_update(0, D, ncol, ncol, Cc, Cc, za.d(), on.d(), 0, ncol, partB.d(), s), "norm0");
        lcheck(qkl_reduce(partB.d(), NB, ncol, group, 1, 0, nrm.d(), 0, ncol, s), "reduce");
        hip_check(hipStreamSynchronize(s), "sync");
    }
    lcheck(qkl_fill_zero(nel, P, s), "zero");
    const double t0 = now_s();
    for (int k = 0; k < n_steps; ++k) {
        lcheck(qkl_apply_block(E.tables(), E.wide() ? 1 : 0, Cc, W, ncol, 0, ncol, k, P, nrm.d(), ncol, partA.d(), npA, s),
               "apply_block");
        lcheck(qkl_reduce(partA.d(), npA, ncol, group, 0, 0, alpha.d(), k, ncol, s), "reduce alpha");
        lcheck(qkl_lanczos_update(0, D, ncol, ncol, W, Cc, alpha.d(), nrm.d(), k, ncol, partB.d(), s), "update");
        lcheck(qkl_reduce(partB.d(), NB, ncol, group, 1, 0, nrm.d(), k + 1, ncol, s), "reduce beta");
        double* t = P;
        P = Cc;
        Cc = W;
        W = t;
    }
    hip_check(hipStreamSynchronize(s), "sync");
    if (seconds) *seconds = now_s() - t0;
    std::vector<double> ha((size_t)n_steps * ncol), hn((size_t)(n_steps + 1) * ncol);
    hip_check(hipMemcpy(ha.data(), alpha.get(), ha.size() * sizeof(double), hipMemcpyDeviceToHost), "d2h");
    hip_check(hipMemcpy(hn.data(), nrm.get(), hn.size() * sizeof(double), hipMemcpyDeviceToHost), "d2h");
    for (int g = 0; g < ng; ++g) {
        const int c = g * group;
        out.norm0[g] = hn[c];
        if (hn[c] < 1e-15) continue;
        for (int k = 0; k < n_steps; ++k) {
            out.alphas[g].push_back(ha[(size_t)k * ncol + c]);
            const double b = hn[(size_t)(k + 1) * ncol + c];
            if (b < 1e-15) break;
            out.betas[g].push_back(b);
        }
    }
    return out;
}

void ftlm_observables(const std::vector<std::vector<double>>& alphas, const std::vector<std::vector<double>>& betas,
                      const std::vector<double>& norm2, const std::vector<double>& bT, std::vector<double>& Zo,
                      std::vector<double>& Eo, std::vector<double>& Cvo)
{
    const size_t R = alphas.size(), nT = bT.size();
    std::vector<double> Z(nT, 0.0), E(nT, 0.0), E2(nT, 0.0);
    for (size_t r = 0; r < R; ++r) {
        const int n = (int)alphas[r].size();
        if (n == 0) continue;
        // eigenvalues and first eigenvector components of the r-th tridiagonal
        std::vector<double> ev(n), fc(n);
        symtri_eig(alphas[r].data(), betas[r].data(), n, ev.data(), fc.data(), kFirstRow);
        for (size_t t = 0; t < nT; ++t)
            for (int i = 0; i < n; ++i) {
                const double w = norm2[r] * fc[i] * fc[i] * std::exp(-bT[t] * ev[i]);
                Z[t] += w;
                E[t] += ev[i] * w;
                E2[t] += ev[i] * ev[i] * w;
            }
    }
    Zo.assign(nT, 0.0);
    Eo.assign(nT, 0.0);
    Cvo.assign(nT, 0.0);
    for (size_t t = 0; t < nT; ++t) {
        const double z = Z[t] / R, e = E[t] / R, e2 = E2[t] / R;
        Zo[t] = z;
        Eo[t] = e / z;
        Cvo[t] = bT[t] * bT[t] * (e2 / z - (e / z) * (e / z));
    }
}

double spectral_value(const double* alphas, const double* betas, size_t n, double norm_phi0, double omega, double E0,
                      double eta)
{
    if (n == 0) return 0.0;
    const std::complex<double> z(omega + E0, eta);
    std::complex<double> f = 0.0;
    for (int i = (int)n - 1; i >= 0; --i) {
        if (i == (int)n - 1) f = 1.0 / (z - alphas[i]);
        else f = 1.0 / (z - alphas[i] - betas[i] * betas[i] * f);
    }
    return -1.0 / M_PI * std::imag(norm_phi0 * norm_phi0 * f);
}

// ---------------------------------------------------------------- Davidson

namespace {

using cd = std::complex<double>;

} // namespace

DavidsonOutput davidson_run(const Engine& E, int n_eig, int max_subspace, double tol, bool want_vectors, hipStream_t s)
{
    DavidsonOutput out;
    const size_t D = (size_t)E.dim();
    if (D == 0) return out;
    n_eig = (int)std::min<size_t>((size_t)std::max(n_eig, 1), D);
    max_subspace = std::max(max_subspace, 2 * n_eig);
    const size_t vb = 2 * D * sizeof(double);
    DevBuf diag(vb), part(qkl_blas_partials() * 2 * sizeof(double)), tmp(4 * sizeof(double));
    E.diagonal(diag.d(), true, s);
    std::vector<double> dh(2 * D);
    hip_check(hipMemcpyAsync(dh.data(), diag.get(), vb, hipMemcpyDeviceToHost, s), "d2h");
    hip_check(hipStreamSynchronize(s), "sync");
    auto dotc = [&](const double* a, const double* b) { return dot_host(true, D, a, b, part.d(), tmp.d(), s); };
    auto axpy = [&](cd a, const double* x, double* y) { lcheck(qkl_axpy(1, D, a.real(), a.imag(), x, y, s), "axpy"); };
    auto nrm = [&](const double* x) { return std::sqrt(dotc(x, x).real()); };
    // n_eig smallest diagonal entries (ties by index)
    std::vector<size_t> di(D);
    std::iota(di.begin(), di.end(), 0);
    std::partial_sort(di.begin(), di.begin() + n_eig, di.end(), [&](size_t i, size_t j) {
        return dh[2 * i] < dh[2 * j] || (dh[2 * i] == dh[2 * j] && i < j);
    });
    std::vector<DevBuf> V, HV;
    for (int i = 0; i < n_eig; ++i) {
        DevBuf v(vb);
        lcheck(qkl_fill_zero(2 * D, v.d(), s), "zero");
        const double one = 1.0;
        hip_check(hipMemcpyAsync(v.d() + 2 * di[i], &one, sizeof(double), hipMemcpyHostToDevice, s), "h2d");
        for (auto& vp : V) axpy(-dotc(vp.d(), v.d()), vp.d(), v.d());
        double n = nrm(v.d());
        if (n < 1e-10) {
            lcheck(qkl_fill_random(0, 2 * D, 1234, 0, 1, 1, 0, v.d(), s), "random");
            for (auto& vp : V) axpy(-dotc(vp.d(), v.d()), vp.d(), v.d());
            n = nrm(v.d());
        }
        lcheck(qkl_scal(1, D, 1.0 / n, 0.0, v.d(), s), "scal");
        DevBuf hv(vb);
        E.apply(v.d(), hv.d(), true, s);
        V.push_back(std::move(v));
        HV.push_back(std::move(hv));
    }
    for (int iter = 0; iter < 100; ++iter) {
        const int m = (int)V.size();
        std::vector<cd> Hs((size_t)m * m);
        for (int i = 0; i < m; ++i)
            for (int j = 0; j <= i; ++j) {
                Hs[(size_t)i * m + j] = dotc(V[i].d(), HV[j].d());
                if (i != j) Hs[(size_t)j * m + i] = std::conj(Hs[(size_t)i * m + j]);
            }
        // Ritz pairs of the projection: es[k][i] = component i of the k-th
        std::vector<double> ev;
        std::vector<cd> Vs;
        herm_eig(m, Hs, ev, Vs);
        std::vector<std::vector<cd>> es(m, std::vector<cd>(m));
        for (int k = 0; k < m; ++k)
            for (int i = 0; i < m; ++i) es[k][i] = Vs[(size_t)i * m + k];
        bool all_conv = true;
        std::vector<DevBuf> corr;
        for (int k = 0; k < n_eig; ++k) {
            const double lam = ev[k];
            DevBuf r(vb);
            lcheck(qkl_fill_zero(2 * D, r.d(), s), "zero");
            for (int i = 0; i < m; ++i) {
                axpy(es[k][i], HV[i].d(), r.d());
                axpy(-lam * es[k][i], V[i].d(), r.d());
            }
            if (nrm(r.d()) > tol) {
                all_conv = false;
                DevBuf t(vb);
                lcheck(qkl_davidson_precond(1, D, r.d(), diag.d(), lam, t.d(), s), "precond");
                corr.push_back(std::move(t));
            }
        }
        if (all_conv || iter == 99) {
            out.iterations = iter + 1;
            out.converged = all_conv;
            for (int k = 0; k < n_eig; ++k) {
                out.eigenvalues.push_back(ev[k]);
                if (want_vectors) {
                    DevBuf e(vb);
                    lcheck(qkl_fill_zero(2 * D, e.d(), s), "zero");
                    for (int i = 0; i < m; ++i) axpy(es[k][i], V[i].d(), e.d());
                    std::vector<double> h(2 * D);
                    hip_check(hipMemcpyAsync(h.data(), e.get(), vb, hipMemcpyDeviceToHost, s), "d2h");
                    hip_check(hipStreamSynchronize(s), "sync");
                    out.eigenvectors.push_back(std::move(h));
                }
            }
            return out;
        }
        if (V.size() + corr.size() > (size_t)max_subspace) {
            std::vector<DevBuf> nV, nHV;
            for (int k = 0; k < n_eig; ++k) {
                DevBuf rv(vb), rh(vb);
                lcheck(qkl_fill_zero(2 * D, rv.d(), s), "zero");
                lcheck(qkl_fill_zero(2 * D, rh.d(), s), "zero");
                for (int i = 0; i < m; ++i) {
                    axpy(es[k][i], V[i].d(), rv.d());
                    axpy(es[k][i], HV[i].d(), rh.d());
                }
                nV.push_back(std::move(rv));
                nHV.push_back(std::move(rh));
            }
            V = std::move(nV);
            HV = std::move(nHV);
        }
        for (auto& t : corr) {
            for (auto& v : V) axpy(-dotc(v.d(), t.d()), v.d(), t.d());
            const double n = nrm(t.d());
            if (n > 1e-10) {
                lcheck(qkl_scal(1, D, 1.0 / n, 0.0, t.d(), s), "scal");
                DevBuf ht(vb);
                E.apply(t.d(), ht.d(), true, s);
                V.push_back(std::move(t));
                HV.push_back(std::move(ht));
            }
        }
    }
    return out;
}

} // namespace qk
