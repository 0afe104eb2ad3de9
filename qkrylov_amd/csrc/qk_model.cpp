// qk_model.cpp -- closed-form bases, OpSum compiler, device-table builder.
#include "qk_model.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <tuple>
#include <unordered_map>

namespace qk {

namespace {

struct Binom {
    uint64_t c[65][65];
    Binom() {
        for (int n = 0; n <= 64; ++n) {
            c[n][0] = 1;
            for (int k = 1; k <= 64; ++k) c[n][k] = (k > n) ? 0 : (n == 0 ? 0 : c[n - 1][k - 1] + (k <= n - 1 ? c[n - 1][k] : 0));
        }
    }
};
const Binom& binom_table()
{
    static Binom b;
    return b;
}
inline uint64_t C(int n, int k)
{
    if (n < 0 || k < 0 || k > n || n > 64) return 0;
    return binom_table().c[n][k];
}
inline uint64_t pow2(int e) { return e >= 64 ? 0 : (1ULL << e); }
inline uint64_t pow3(int e)
{
    uint64_t r = 1;
    while (e-- > 0) r *= 3;
    return r;
}
inline int popc(uint64_t x) { return __builtin_popcountll(x); }

// t-J: fillings of i complete sites with a ups and b dns (-1 = any)
uint64_t tj_sites(int i, int a, int b)
{
    if (a == -1 && b == -1) return pow3(i);
    if (a == -1) return C(i, b) * pow2(i - b);
    if (b == -1) return C(i, a) * pow2(i - a);
    if (a < 0 || b < 0) return 0;
    return C(i, a) * C(i - a, b);
}

} // namespace

// ---------------------------------------------------------------- basis

BasisSpec::BasisSpec(Kind kind, int nsites, const Sector& sec) : kind_(kind), nsites_(nsites), sec_(sec)
{
    // limits and messages of src/basis/*_basis.cpp constructors
    switch (kind) {
    case Kind::SpinHalf:
        if (nsites <= 0 || nsites > 63) throw std::runtime_error("SpinHalfBasis: N must satisfy 1 <= N <= 63");
        break;
    case Kind::Fermion:
        if (nsites <= 0 || nsites > 63) throw std::runtime_error("FermionBasis: N must satisfy 1 <= N <= 63");
        break;
    case Kind::Hubbard:
        if (nsites <= 0 || nsites > 31) throw std::runtime_error("HubbardBasis: N must satisfy 1 <= N <= 31");
        break;
    case Kind::TJ:
        if (nsites <= 0 || nsites > 31) throw std::runtime_error("TJBasis: N must satisfy 1 <= N <= 31");
        break;
    }
    nbits_ = spinless() ? nsites : 2 * nsites;
    for (int p = 0; p < nbits_; ++p) {
        if (is_up_bit(p)) up_mask_ |= 1ULL << p;
        else dn_mask_ |= 1ULL << p;
    }
    switch (kind) {
    case Kind::SpinHalf:
        if (sec.use_sz) {
            if (((sec.sz2 + nsites) & 1) != 0) empty_ = true;
            else {
                A_ = (sec.sz2 + nsites) / 2;
                if (A_ < 0 || A_ > nsites) empty_ = true;
            }
        }
        break;
    case Kind::Fermion:
        if (sec.use_n) {
            A_ = sec.n;
            if (A_ < 0 || A_ > nsites) empty_ = true;
        }
        break;
    case Kind::Hubbard:
    case Kind::TJ:
        if (sec.use_nup) {
            A_ = sec.nup;
            if (A_ < 0 || A_ > nsites) empty_ = true;
        }
        if (sec.use_ndn) {
            B_ = sec.ndn;
            if (B_ < 0 || B_ > nsites) empty_ = true;
        }
        break;
    }
    dim_ = empty_ ? 0 : static_cast<int64_t>(completions(nbits_, A_, B_));
}

uint64_t BasisSpec::completions(int p, int a, int b) const
{
    if (spinless()) return (a == -1) ? pow2(p) : C(p, a);
    if (kind_ == Kind::Hubbard) {
        const int nu = (p + 1) / 2, nd = p / 2;
        const uint64_t cu = (a == -1) ? pow2(nu) : C(nu, a);
        const uint64_t cd = (b == -1) ? pow2(nd) : C(nd, b);
        return cu * cd;
    }
    const int i = p / 2;
    if ((p & 1) == 0) return tj_sites(i, a, b);
    // bits [0, 2i+1): sites 0..i-1 plus the lone up bit of site i
    if (a == -1) return 2 * tj_sites(i, -1, b);
    return tj_sites(i, a, b) + (a >= 1 ? tj_sites(i, a - 1, b) : 0);
}

bool BasisSpec::contains(uint64_t s) const
{
    if (empty_) return false;
    if (nbits_ < 64 && (s >> nbits_) != 0) return false;
    if (kind_ == Kind::TJ && (s & (s >> 1) & up_mask_) != 0) return false;
    if (A_ != -1 && popc(s & up_mask_) != A_) return false;
    if (B_ != -1 && popc(s & dn_mask_) != B_) return false;
    return true;
}

int64_t BasisSpec::rank(uint64_t s) const
{
    if (!contains(s)) return -1;
    int a = A_, b = B_;
    uint64_t r = 0;
    for (int p = nbits_ - 1; p >= 0; --p) {
        if (!((s >> p) & 1ULL)) continue;
        r += completions(p, a, b);
        if (is_up_bit(p)) { if (a != -1) --a; }
        else { if (b != -1) --b; }
    }
    return static_cast<int64_t>(r);
}

uint64_t BasisSpec::unrank(int64_t rr) const
{
    uint64_t r = static_cast<uint64_t>(rr);
    int a = A_, b = B_;
    uint64_t s = 0;
    for (int p = nbits_ - 1; p >= 0; --p) {
        const uint64_t c = completions(p, a, b);
        if (r >= c) {
            r -= c;
            s |= 1ULL << p;
            if (is_up_bit(p)) { if (a != -1) --a; }
            else { if (b != -1) --b; }
        }
    }
    return s;
}

// ---------------------------------------------------------------- compiler

namespace {

struct Partial {
    uint64_t cm, cv, F, M;
    int csign;
    cplx amp;
};

inline uint64_t bitv(int b) { return 1ULL << b; }
inline uint64_t below(int b) { return b >= 64 ? ~0ULL : (bitv(b) - 1ULL); }
inline bool known(const Partial& p, int b) { return (p.cm >> b) & 1ULL; }
inline int cur(const Partial& p, int b) { return (int)(((p.cv ^ p.F) >> b) & 1ULL); }

// require the current value of bit b to be v; false if impossible
bool require(Partial& p, int b, int v)
{
    if (known(p, b)) return cur(p, b) == v;
    const int orig = v ^ (int)((p.F >> b) & 1ULL);
    p.cm |= bitv(b);
    p.cv = (p.cv & ~bitv(b)) | ((uint64_t)orig << b);
    return true;
}

// make bit b's original value known, splitting if needed
void split_on(std::vector<Partial>& v, int b)
{
    std::vector<Partial> out;
    out.reserve(v.size() * 2);
    for (auto& p : v) {
        if (known(p, b)) { out.push_back(p); continue; }
        Partial p0 = p, p1 = p;
        p0.cm |= bitv(b);
        p0.cv &= ~bitv(b);
        p1.cm |= bitv(b);
        p1.cv |= bitv(b);
        out.push_back(p0);
        out.push_back(p1);
    }
    v.swap(out);
}

void jw(Partial& p, int b)
{
    // phase on the pre-application state: (-1)^popcount((s ^ F) & below(b))
    p.M ^= below(b);
    p.csign ^= popc(p.F & below(b)) & 1;
}

void check_bit(int b)
{
    if (b < 0 || b >= 64) throw std::invalid_argument("operator site index out of the 64-bit state range");
}

// apply one factor to every live partial branch
void apply_factor(Kind sk, const Factor& f, std::vector<Partial>& v)
{
    const std::string& op = f.op;
    std::vector<Partial> out;
    auto diag_split = [&](int b, cplx a0, cplx a1) {
        check_bit(b);
        split_on(v, b);
        for (auto& p : v) {
            p.amp *= cur(p, b) ? a1 : a0;
            if (p.amp != cplx(0.0)) out.push_back(p);
        }
    };
    auto flip_req = [&](int b, int need, bool with_jw, int also_zero) {
        check_bit(b);
        if (also_zero >= 0) check_bit(also_zero);
        for (auto p : v) {
            if (!require(p, b, need)) continue;
            if (also_zero >= 0 && !require(p, also_zero, 0)) continue;
            if (with_jw) jw(p, b);
            p.F ^= bitv(b);
            out.push_back(p);
        }
    };
    switch (sk) {
    case Kind::SpinHalf: {
        const int b = f.site;
        if (op == "Sz") diag_split(b, -0.5, 0.5);
        else if (op == "Sp") flip_req(b, 0, false, -1);
        else if (op == "Sm") flip_req(b, 1, false, -1);
        else if (op == "Sx") {
            check_bit(b);
            for (auto p : v) { p.F ^= bitv(b); p.amp *= 0.5; out.push_back(p); }
        } else if (op == "Sy") {
            // spinhalf_site.cpp:104-113: up -> -i/2, down -> +i/2
            check_bit(b);
            split_on(v, b);
            for (auto p : v) {
                p.amp *= cur(p, b) ? cplx(0.0, -0.5) : cplx(0.0, 0.5);
                p.F ^= bitv(b);
                out.push_back(p);
            }
        } else throw std::runtime_error("Unknown spin operator: " + op);
        break;
    }
    case Kind::Fermion: {
        const int b = f.site;
        if (op == "N") diag_split(b, 0.0, 1.0);
        else if (op == "Id") { check_bit(b); out = v; }
        else if (op == "C") flip_req(b, 1, true, -1);
        else if (op == "Cdag") flip_req(b, 0, true, -1);
        else throw std::runtime_error("Unknown fermion operator: " + op);
        break;
    }
    case Kind::Hubbard:
    case Kind::TJ: {
        const bool tj = (sk == Kind::TJ);
        const int bu = 2 * f.site, bd = 2 * f.site + 1;
        if (op == "Nup") diag_split(bu, 0.0, 1.0);
        else if (op == "Ndn") diag_split(bd, 0.0, 1.0);
        else if (!tj && op == "Nupdn") {
            check_bit(bu);
            check_bit(bd);
            split_on(v, bu);
            split_on(v, bd);
            for (auto& p : v)
                if (cur(p, bu) && cur(p, bd)) out.push_back(p);
        } else if (op == "CUp") flip_req(bu, 1, true, -1);
        else if (op == "CdagUp") flip_req(bu, 0, true, tj ? bd : -1);
        else if (op == "CDn") flip_req(bd, 1, true, -1);
        else if (op == "CdagDn") flip_req(bd, 0, true, tj ? bu : -1);
        else throw std::runtime_error(std::string(tj ? "Unknown TJ operator: " : "Unknown Hubbard operator: ") + op);
        break;
    }
    }
    v.swap(out);
}

} // namespace

Compiled compile_opsum(const BasisSpec& basis, Kind sk, const OpSum& ops)
{
    Compiled out;
    const uint64_t valid = basis.valid_mask();
    const int nb = basis.nbits();
    std::map<uint64_t, cplx> monos;
    std::vector<QkDiagTerm> cond_terms;
    // merge key for paired branches
    std::map<std::tuple<uint64_t, uint64_t, uint64_t, uint64_t, double, double>, size_t> pair_idx;

    for (const auto& t : ops.terms) {
        std::vector<Partial> live;
        live.push_back(Partial{~valid, 0, 0, 0, 0, t.coeff});
        for (const auto& f : t.factors) {
            if (live.empty()) break;   // no state reaches later factors
            apply_factor(sk, f, live);
        }
        for (auto p : live) {
            if (p.F & ~valid) continue;   // target leaves the bit range
            // make every flipped bit's original value known when the sector
            // constrains its species, then drop non-conserving branches
            std::vector<Partial> vv{p};
            for (int b = 0; b < nb; ++b) {
                if (!((p.F >> b) & 1ULL)) continue;
                const bool up = basis.is_up_bit(b);
                if ((up && basis.target_a() != -1) || (!up && basis.target_b() != -1)) split_on(vv, b);
                if (basis.no_double()) {
                    const int site = b / 2;
                    split_on(vv, 2 * site);
                    split_on(vv, 2 * site + 1);
                }
            }
            for (auto& q : vv) {
                int da = 0, db = 0;
                bool ok = true;
                for (int b = 0; b < nb; ++b) {
                    if (!((q.F >> b) & 1ULL) || !known(q, b)) continue;
                    const int d = ((q.cv >> b) & 1ULL) ? -1 : 1;
                    if (basis.is_up_bit(b)) da += d; else db += d;
                }
                if (basis.target_a() != -1 && da != 0) ok = false;
                if (basis.target_b() != -1 && db != 0) ok = false;
                if (ok && basis.no_double()) {
                    for (int site = 0; site < basis.nsites() && ok; ++site) {
                        const uint64_t sm = bitv(2 * site) | bitv(2 * site + 1);
                        // source must itself be allowed
                        if ((q.cm & sm) == sm && (q.cv & sm) == sm) ok = false;
                        if ((q.F & sm) && (q.cm & sm) == sm && (((q.cv ^ q.F) & sm) == sm)) ok = false;
                    }
                }
                if (!ok) continue;
                if (q.amp == cplx(0.0)) continue;
                cplx val = std::conj(q.amp) * (q.csign ? -1.0 : 1.0);
                const uint64_t cm = q.cm & valid, cv = q.cv & valid, M = q.M & valid;
                if (q.F == 0 && M == 0) {
                    const uint64_t ones = cm & cv, zeros = cm & ~cv;
                    const int nz = popc(zeros);
                    if (nz <= 8) {
                        std::vector<int> zb;
                        for (int b = 0; b < 64; ++b) if ((zeros >> b) & 1ULL) zb.push_back(b);
                        for (uint32_t sub = 0; sub < (1u << nz); ++sub) {
                            uint64_t mono = ones;
                            for (int k = 0; k < nz; ++k) if ((sub >> k) & 1u) mono |= bitv(zb[k]);
                            monos[mono] += (popc(sub) & 1) ? -val : val;
                        }
                    } else {
                        cond_terms.push_back(QkDiagTerm{cm, cv, 0, 2, val.real(), val.imag()});
                    }
                    continue;
                }
                QkBranch br{cm, cv, 0, q.F, M, val.real(), val.imag()};
                // pair merge: same structure and value, condition flipped on F
                if ((q.F & ~cm) == 0) {
                    const uint64_t lowcv = std::min(cv, cv ^ q.F);
                    auto key = std::make_tuple(cm, q.F, M, lowcv, val.real(), val.imag());
                    auto it = pair_idx.find(key);
                    if (it != pair_idx.end()) {
                        QkBranch& e = out.offdiag[it->second];
                        if (e.alt == 0 && e.cv == (cv ^ q.F)) {
                            e.alt = q.F;
                            pair_idx.erase(it);
                            continue;
                        }
                    }
                    pair_idx[key] = out.offdiag.size();
                }
                out.offdiag.push_back(br);
            }
        }
    }

    // diagonal polynomial -> grouped terms
    std::map<std::tuple<double, double, int>, uint64_t> pair_groups;   // (re, im, d) -> mask
    std::vector<std::pair<uint64_t, cplx>> higher;
    for (const auto& kv : monos) {
        const uint64_t mono = kv.first;
        const cplx c = kv.second;
        if (c == cplx(0.0)) continue;
        const int deg = popc(mono);
        if (deg == 0) { out.dconst += c; continue; }
        if (deg == 1) {
            pair_groups[std::make_tuple(c.real(), c.imag(), 0)] |= mono;
        } else if (deg == 2) {
            const int i = __builtin_ctzll(mono);
            const int j = 63 - __builtin_clzll(mono);
            auto& m = pair_groups[std::make_tuple(c.real(), c.imag(), j - i)];
            if (m & bitv(i)) {
                higher.push_back({mono, c});  // cannot happen (unique mono), kept for safety
            } else {
                m |= bitv(i);
            }
        } else {
            higher.push_back({mono, c});
        }
    }
    for (const auto& kv : pair_groups) {
        out.diag.push_back(QkDiagTerm{kv.second, 0, std::get<2>(kv.first), 0, std::get<0>(kv.first), std::get<1>(kv.first)});
    }
    for (const auto& h : higher) out.diag.push_back(QkDiagTerm{h.first, 0, 0, 1, h.second.real(), h.second.imag()});
    for (const auto& c : cond_terms) out.diag.push_back(c);

    out.is_real = (out.dconst.imag() == 0.0);
    for (const auto& b : out.offdiag) if (b.im != 0.0) out.is_real = false;
    for (const auto& d : out.diag) if (d.im != 0.0) out.is_real = false;
    out.diag_only = out.offdiag.empty();
    out.max_hops = static_cast<int>(out.offdiag.size());
    return out;
}

// ---------------------------------------------------------------- tables

double tile_utilisation(const BasisSpec& b, int NL)
{
    const int nb = b.nbits(), NH = nb - NL;
    const bool two = !b.spinless(), tj = b.no_double();
    auto cnt = [&](int bits, int a, int c) -> double {
        if (a < 0 || c < 0) return 0.0;
        if (!two) return c == 0 ? (double)C(bits, a) : 0.0;
        const int n = bits / 2;
        if (tj) return (double)C(n, a) * (double)C(n - a, c);
        return (double)C(n, a) * (double)C(n, c);
    };
    const int amax = two ? NL / 2 : NL, hmax = two ? NH / 2 : NH;
    auto size_lo = [&](int ar, int br) {
        double s = 0.0;
        for (int a = 0; a <= amax; ++a)
            for (int c = 0; c <= (two ? amax : 0); ++c)
                if ((ar == -1 || a == ar) && (br == -1 || c == br)) s += cnt(NL, a, c);
        return s;
    };
    const int A = b.target_a(), B = b.target_b();
    double rows = 0.0, tiles = 0.0;
    for (int a = 0; a <= hmax; ++a)
        for (int c = 0; c <= (two ? hmax : 0); ++c) {
            const double nh = cnt(NH, a, c);
            if (nh == 0.0) continue;
            const int ar = A == -1 ? -1 : A - a, br = B == -1 ? -1 : B - c;
            if ((A != -1 && (ar < 0 || ar > amax)) || (B != -1 && (br < 0 || br > amax))) continue;
            const double s = size_lo(ar, br);
            rows += nh * s;
            tiles += nh * std::ceil(s / 64.0);
        }
    return tiles > 0 ? rows / (64.0 * tiles) : 1.0;
}

int choose_NL_model(const BasisSpec& basis, const Compiled& comp)
{
    const int nb = basis.nbits();
    if (nb <= 8) return nb - ((!basis.spinless() && (nb & 1)) ? 1 : 0);
    int best = choose_NL(basis);
    double best_cost = 1e300;
    for (int NL = std::min(nb, 8); NL <= std::min(nb, 15); ++NL) {
        if (!basis.spinless() && (NL & 1)) continue;
        if (nb - NL > 28) continue;
        const uint64_t lomask = (1ULL << NL) - 1ULL;
        double w = 0.5 * comp.diag.size();
        for (const auto& br : comp.offdiag) {
            const bool hi_only = ((br.cm | br.cv | br.alt | br.F | br.M) & lomask) == 0;
            w += hi_only ? 1.0 : 5.0;
        }
        const double cost = w / tile_utilisation(basis, NL);
        if (cost < best_cost * 0.999) {
            best_cost = cost;
            best = NL;
        }
    }
    return best;
}

int choose_NL(const BasisSpec& basis)
{
    int nl = std::min(basis.nbits(), 14);
    if (!basis.spinless()) nl &= ~1;
    if (basis.nbits() - nl > 28) nl = basis.nbits() - 28;
    if (!basis.spinless() && (nl & 1)) ++nl;
    return nl;
}

// Two-level rank of the lo field (see HostTables): the low half of NL2 bits
// (even for two-species bases, so no site straddles the halves) is ranked
// inside its own class, the high half contributes the count of valid lo of
// the same total class with a smaller high half. Checked against the
// one-level table for every valid lo.
static void build_split_rank(const BasisSpec& basis, HostTables& T)
{
    const int NL = T.NL;
    const bool two = !basis.spinless();
    const bool tj = basis.no_double();
    const bool ca = basis.target_a() != -1, cb = basis.target_b() != -1;
    // the low-half width with the smallest pair of tables
    auto n_classes = [&](int n2) {
        return (ca ? (two ? n2 / 2 : n2) + 1 : 1) * (cb ? n2 / 2 + 1 : 1);
    };
    int NL2 = 0;
    size_t best = ~(size_t)0;
    for (int n2 = (NL + 1) / 2; n2 < NL && n2 <= 16; ++n2) {
        if (two && (n2 & 1)) continue;
        const size_t bytes = ((size_t)1 << n2) * 2 + ((size_t)1 << (NL - n2)) * (size_t)n_classes(n2) * 4;
        if (bytes < best) {
            best = bytes;
            NL2 = n2;
        }
    }
    const int NLH = NL - NL2;
    const uint32_t m2 = (1u << NL2) - 1u;
    uint32_t up = 0, dn = 0;
    for (int p = 0; p < NL2; ++p) (basis.is_up_bit(p) ? up : dn) |= 1u << p;
    const int maxa = two ? NL2 / 2 : NL2, maxb = two ? NL2 / 2 : 0;
    const int KA = ca ? maxa + 1 : 1, KB = cb ? maxb + 1 : 1;
    const uint32_t upm2 = ca ? up : 0u, dnm2 = cb ? dn : 0u;
    auto valid = [&](uint32_t v) { return !tj || (v & (v >> 1) & 0x55555555u) == 0; };
    auto cls = [&](uint32_t l) { return popc((uint64_t)(l & upm2)) * KB + popc((uint64_t)(l & dnm2)); };
    const int nc = KA * KB;
    // low halves: rank inside their class, and class sizes
    std::vector<uint32_t> nlow(nc, 0);
    T.tl.assign(1u << NL2, 0);
    for (uint32_t l = 0; l <= m2; ++l) {
        if (!valid(l)) continue;
        T.tl[l] = static_cast<uint16_t>(nlow[cls(l)]++);
    }
    // high halves, ascending: for every total class (a, b) the number of
    // valid lo seen so far (with a smaller high half)
    const int TA = ca ? 2 * maxa + 2 + (two ? NLH / 2 : NLH) : 1;
    const int TB = cb ? 2 * maxb + 2 + NLH / 2 : 1;
    std::vector<uint64_t> run((size_t)TA * TB, 0);
    const uint32_t nh = 1u << NLH;
    T.th.assign((size_t)nh * nc, 0);
    for (uint32_t h = 0; h < nh; ++h) {
        const uint32_t hs = h << NL2;
        if (!valid(hs)) continue;
        const int ah = ca ? popc((uint64_t)(hs & (basis.up_mask() & ((1ULL << NL) - 1)))) : 0;
        const int bh = cb ? popc((uint64_t)(hs & (basis.dn_mask() & ((1ULL << NL) - 1)))) : 0;
        for (int c = 0; c < nc; ++c) {
            const int a = ah + c / KB, b = bh + c % KB;
            const uint64_t v = run[(size_t)a * TB + b];
            if (v > 0xFFFFFFFFull) throw std::logic_error("split rank overflow");
            T.th[(size_t)h * nc + c] = static_cast<uint32_t>(v);
        }
        // lo with this high half add to every total class they reach
        for (int c = 0; c < nc; ++c) {
            const int a = ah + c / KB, b = bh + c % KB;
            run[(size_t)a * TB + b] += nlow[c];
        }
    }
    for (uint32_t lo = 0; lo < (1u << NL); ++lo) {
        if (T.lrank[lo] == 0xFFFFFFFFu) continue;
        const uint32_t l = lo & m2, h = lo >> NL2;
        if (T.th[(size_t)h * nc + cls(l)] + T.tl[l] != T.lrank[lo]) throw std::logic_error("split rank mismatch");
    }
    T.split = true;
    T.NL2 = NL2;
    T.nc = nc;
    T.KB = KB;
    T.upm2 = upm2;
    T.dnm2 = dnm2;
}

HostTables build_tables(const BasisSpec& basis, int NL)
{
    HostTables T;
    const int nbits = basis.nbits();
    if (NL > nbits) NL = nbits;
    if (!basis.spinless() && (NL & 1)) throw std::invalid_argument("NL must be even for two-species bases");
    if (NL > 20) throw std::invalid_argument("NL too large");
    const int NH = nbits - NL;
    if (NH > 28)
        throw std::invalid_argument("split NL = " + std::to_string(NL) + " leaves " + std::to_string(NH) +
                                    " high bits; at most 28 are supported (the block-offset table has "
                                    "2^(bits - NL) entries): use a larger split");
    if (basis.dim() >= (int64_t(1) << 32)) throw std::invalid_argument("dimension exceeds 2^32 rows");
    T.NL = NL;
    T.NH = NH;
    const int A = basis.target_a(), B = basis.target_b();
    const bool tj = basis.no_double();
    // classes of lo values: (a_lo, b_lo), collapsed where unconstrained
    const int K = NL + 1;
    auto up_cnt = [&](uint64_t v, int off) {
        int c = 0;
        for (uint64_t w = v; w; w &= w - 1) if (basis.is_up_bit(__builtin_ctzll(w) + off)) ++c;
        return c;
    };
    auto lo_key = [&](uint64_t lo) {
        const int a = up_cnt(lo, 0), b = popc(lo) - a;
        return (A == -1 ? 0 : a) * K + (B == -1 ? 0 : b);
    };
    const uint64_t nlo = 1ULL << NL;
    std::vector<uint32_t> class_count(K * K, 0);
    T.lrank.assign(nlo, 0xFFFFFFFFu);
    for (uint64_t lo = 0; lo < nlo; ++lo) {
        if (tj && (lo & (lo >> 1) & basis.up_mask()) != 0) continue;
        const int k = lo_key(lo);
        T.lrank[lo] = class_count[k]++;
    }
    std::vector<uint32_t> class_start(K * K + 1, 0);
    for (int k = 0; k < K * K; ++k) class_start[k + 1] = class_start[k] + class_count[k];
    T.lolist.assign(class_start[K * K], 0);
    {
        std::vector<uint32_t> fill(class_start.begin(), class_start.end() - 1);
        for (uint64_t lo = 0; lo < nlo; ++lo) {
            if (T.lrank[lo] == 0xFFFFFFFFu) continue;
            T.lolist[fill[lo_key(lo)]++] = static_cast<uint32_t>(lo);
        }
    }
    uint32_t maxc = 0;
    for (auto c : class_count) maxc = std::max(maxc, c);
    T.lrank_u16 = (maxc <= 65536);
    if (NL > 14) build_split_rank(basis, T);
    const uint64_t nhi = 1ULL << NH;
    T.hoff.assign(nhi, 0);
    uint64_t rows = 0;
    T.tiles.reserve(static_cast<size_t>(basis.dim() / 64 + 1024));
    for (uint64_t hi = 0; hi < nhi; ++hi) {
        T.hoff[hi] = static_cast<uint32_t>(rows);
        const uint64_t hs = hi << NL;
        if (tj && (hs & (hs >> 1) & basis.up_mask()) != 0) continue;
        const int a = up_cnt(hs, 0), b = popc(hs) - a;
        int ka = 0, kb = 0;
        if (A != -1) { ka = A - a; if (ka < 0 || ka > NL) continue; }
        if (B != -1) { kb = B - b; if (kb < 0 || kb > NL) continue; }
        const int k = ka * K + kb;
        const uint32_t cnt = class_count[k];
        for (uint32_t t = 0; t < cnt; t += 64) {
            T.tiles.push_back(QkTile{static_cast<uint32_t>(hi), class_start[k] + t, static_cast<uint32_t>(rows + t),
                                     std::min<uint32_t>(64, cnt - t)});
        }
        rows += cnt;
    }
    if (static_cast<int64_t>(rows) != basis.dim()) throw std::logic_error("table build: row count mismatch");
    return T;
}

} // namespace qk
