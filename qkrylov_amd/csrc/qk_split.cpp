// qk_split.cpp -- partition search, the two pass engines and the row
// permutation of the split apply (see qk_split.hpp).
#include "qk_split.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <string>

namespace qk {

namespace {

int popc64(uint64_t x) { return __builtin_popcountll(x); }

bool diagonal_op(const std::string& op) { return op == "Sz" || op == "Id"; }

void lc(int rc, const char* what)
{
    if (rc != 0) throw std::runtime_error(std::string("kernel launch failed: ") + what + " (" + std::to_string(rc) + ")");
}

} // namespace

bool find_split(const BasisSpec& basis, const OpSum& ops, int max_unit, SplitPlan& out)
{
    if (basis.kind() != Kind::SpinHalf) return false;
    const int N = basis.nsites();
    if (N < 16 || N > 32) return false;
    // supports of the off-diagonal terms
    std::vector<uint64_t> supp;
    for (const auto& t : ops.terms) {
        uint64_t m = 0;
        bool diag = true;
        for (const auto& f : t.factors) {
            if (f.site < 0 || f.site >= N) return false;
            m |= 1ULL << f.site;
            if (!diagonal_op(f.op)) diag = false;
        }
        if (!diag && t.coeff != cplx(0.0)) supp.push_back(m);
    }
    auto window = [&](int start, int len) {
        uint64_t m = 0;
        for (int k = 0; k < len; ++k) m |= 1ULL << ((start + k) % N);
        return m;
    };
    for (int w : {5, 4}) {
        int best = 1 << 30;
        uint64_t bA = 0, bB = 0, bC = 0;
        for (int i = 0; i < N; ++i) {
            for (int g = 1; g <= N - 2 * w - 1; ++g) {
                const uint64_t mA = window(i, w) | window(i + w + g, w);
                const uint64_t mB = window(i + w, g);
                const uint64_t mC = window(i + 2 * w + g, N - 2 * w - g);
                const int uB = popc64(mA | mB), uC = popc64(mA | mC);
                if (uB > max_unit || uC > max_unit) continue;
                bool ok = true;
                for (uint64_t m : supp)
                    if ((m & mB) && (m & mC)) {
                        ok = false;
                        break;
                    }
                if (!ok) continue;
                const int score = std::max(uB, uC) * 64 + std::abs(uB - uC);
                if (score < best) {
                    best = score;
                    bA = mA;
                    bB = mB;
                    bC = mC;
                }
            }
        }
        if (best == (1 << 30)) continue;
        // internal bits: A low (site order), then C, then B on top
        out = SplitPlan{};
        out.perm.assign(N, -1);
        out.NL = popc64(bA);
        out.nC = popc64(bC);
        out.nB = popc64(bB);
        int bit = 0;
        for (uint64_t m : {bA, bC, bB})
            for (int s = 0; s < N; ++s)
                if ((m >> s) & 1ULL) out.perm[s] = bit++;
        for (const auto& t : ops.terms) {
            uint64_t m = 0;
            bool diag = true;
            Term r;
            r.coeff = t.coeff;
            for (const auto& f : t.factors) {
                m |= 1ULL << f.site;
                if (!diagonal_op(f.op)) diag = false;
                r.factors.push_back({f.op, out.perm[f.site]});
            }
            if (diag) out.ops1.terms.push_back(r);
            else if (t.coeff == cplx(0.0)) continue;
            else if ((m & bC) == 0) out.ops1.terms.push_back(r);
            else out.ops2.terms.push_back(r);
        }
        return true;
    }
    return false;
}

SplitEngines::SplitEngines(const Engine& pub, SplitPlan plan) : pub_(pub), plan_(std::move(plan))
{
    const BasisSpec& pb = pub.basis();
    if (pb.nbits() > 32 || !pub.is_real()) throw std::invalid_argument("split apply: real 32-bit spin models only");
    basis_ = std::make_shared<BasisSpec>(pb.kind(), pb.nsites(), pb.sector());
    EngineOptions eo;
    eo.NL = plan_.NL;
    e1_ = std::make_unique<Engine>(basis_, Kind::SpinHalf, plan_.ops1, eo);
    e2_ = std::make_unique<Engine>(basis_, Kind::SpinHalf, plan_.ops2, eo);
    if (e1_->NL() != plan_.NL || e2_->NL() != plan_.NL) throw std::logic_error("split apply: split mismatch");
    if (!e1_->is_real() || !e2_->is_real() || e1_->tables()->signs || e2_->tables()->signs)
        throw std::invalid_argument("split apply: real terms without parity masks only");
    if (e1_->tables()->rank_mode == 2) throw std::invalid_argument("split apply: lo field too wide");
    // both passes: one contiguous tile segment per XCD, 16 waves per workgroup
    for (Engine* e : {e1_.get(), e2_.get()}) {
        e->set_tuning(0, 0);
        e->set_tuning(2, 16);
    }
    // pass-1 tiles: the same tiles ordered by (C bits, B bits) of hi
    const QkTables* T2 = e2_->tables();
    const size_t nt = T2->ntiles;
    std::vector<QkTile> t(nt);
    hip_check(hipMemcpy(t.data(), T2->tiles, nt * sizeof(QkTile), hipMemcpyDeviceToHost), "d2h tiles");
    const uint32_t cmask = (plan_.nC >= 32) ? ~0u : ((1u << plan_.nC) - 1u);
    std::stable_sort(t.begin(), t.end(), [&](const QkTile& a, const QkTile& b) {
        const uint64_t ka = ((uint64_t)(a.hi & cmask) << 32) | (a.hi >> plan_.nC);
        const uint64_t kb = ((uint64_t)(b.hi & cmask) << 32) | (b.hi >> plan_.nC);
        return ka < kb;
    });
    tiles1_ = DevBuf(std::max<size_t>(nt, 1) * sizeof(QkTile));
    hip_check(hipMemcpy(tiles1_.get(), t.data(), nt * sizeof(QkTile), hipMemcpyHostToDevice), "h2d tiles");
    T1_ = *e1_->tables();
    T1_.tiles = static_cast<const QkTile*>(tiles1_.get());
    // internal byte -> public bits
    std::vector<uint32_t> bm(4 * 256, 0);
    std::vector<int> inv(32, -1);
    for (int s = 0; s < (int)plan_.perm.size(); ++s) inv[plan_.perm[s]] = s;
    for (int k = 0; k < 4; ++k)
        for (int v = 0; v < 256; ++v) {
            uint32_t m = 0;
            for (int j = 0; j < 8; ++j)
                if ((v >> j) & 1) {
                    const int ib = 8 * k + j;
                    if (inv[ib] >= 0) m |= 1u << inv[ib];
                }
            bm[256 * k + v] = m;
        }
    bitmap_ = DevBuf(bm.size() * sizeof(uint32_t));
    hip_check(hipMemcpy(bitmap_.get(), bm.data(), bm.size() * sizeof(uint32_t), hipMemcpyHostToDevice), "h2d map");
}

void SplitEngines::set_tuning(int knob, int value) const
{
    if (knob != 0 && knob != 2) throw std::invalid_argument("split apply: only the schedule and the waves are tunable");
    e1_->set_tuning(knob, value);
    e2_->set_tuning(knob, value);
    if (knob == 0) T1_.sched = value;
    else T1_.apply_waves = value;
}

void SplitEngines::apply(const double* x, double* y, double* tmp, hipStream_t s) const
{
    lc(qkl_apply(&T1_, 0, 0, x, tmp, s), "split pass 1");
    lc(qkl_apply_acc(tables2(), x, y, tmp, s), "split pass 2");
}

void SplitEngines::to_internal(const double* x_pub, double* x_int, hipStream_t s) const
{
    lc(qkl_permute(tables2(), static_cast<const uint32_t*>(bitmap_.get()), pub_.tables(), 0, x_pub, x_int, s),
       "permute to internal");
}

void SplitEngines::to_public(const double* x_int, double* x_pub, hipStream_t s) const
{
    lc(qkl_permute(tables2(), static_cast<const uint32_t*>(bitmap_.get()), pub_.tables(), 1, x_int, x_pub, s),
       "permute to public");
}

} // namespace qk
