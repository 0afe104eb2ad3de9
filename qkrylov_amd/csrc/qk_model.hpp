// qk_model.hpp -- host-side model layer: sectors, closed-form bases, site
// kinds, OpSum, and the OpSum -> device-table compiler.
//
// Reference behaviour mirrored here:
//   Sector                 include/qkrylov/symmetry/sector.hpp:10-41
//   SpinHalf/Fermion/Hubbard/TJ bases   src/basis/*_basis.cpp (ascending order
//                          of valid bitstrings; up at bit 2i, dn at 2i+1)
//   site actions           src/sites/*_site.cpp
//   OpSum / OperatorTerm   include/qkrylov/operators/{opsum,operator_term}.hpp
//   Phase-1 semantics      src/hamiltonian/matrix_free_hamiltonian.cpp:45-92
//                          (factor[0] acts first, conj storage, intermediate
//                          validity, amplitude product in term order)
#pragma once
#include "qk_device.hpp"

#include <complex>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace qk {

using cplx = std::complex<double>;

enum class Kind : int { SpinHalf = 0, Fermion = 1, Hubbard = 2, TJ = 3 };

struct Sector {
    bool use_sz = false;
    int sz2 = 0;
    bool use_nup = false, use_ndn = false;
    int nup = 0, ndn = 0;
    bool use_n = false;
    int n = 0;
    bool use_nb = false;
    int nb = 0;
};

// Closed-form basis. The reference enumerates s = 0 .. 2^nbits-1 and keeps the
// valid ones; index(s) is the position in that ascending list. Here the same
// index is computed by counting: rank(s) = sum over set bits p of the number
// of valid states that agree with s above p and have 0 at p.
class BasisSpec {
public:
    BasisSpec(Kind kind, int nsites, const Sector& sec);

    Kind kind() const { return kind_; }
    int nsites() const { return nsites_; }
    int nbits() const { return nbits_; }
    const Sector& sector() const { return sec_; }
    int64_t dim() const { return dim_; }
    bool no_double() const { return kind_ == Kind::TJ; }
    int target_a() const { return A_; }   // -1: unconstrained
    int target_b() const { return B_; }
    bool is_up_bit(int p) const { return spinless() || (p & 1) == 0; }
    bool spinless() const { return kind_ == Kind::SpinHalf || kind_ == Kind::Fermion; }
    uint64_t valid_mask() const { return nbits_ >= 64 ? ~0ULL : ((1ULL << nbits_) - 1ULL); }
    uint64_t up_mask() const { return up_mask_; }
    uint64_t dn_mask() const { return dn_mask_; }

    bool contains(uint64_t s) const;
    int64_t rank(uint64_t s) const;        // -1 if not contained
    uint64_t unrank(int64_t r) const;      // r must be in [0, dim)
    // number of valid fillings of bits [0, p) with a ups / b dns left
    // (-1 = unconstrained)
    uint64_t completions(int p, int a, int b) const;

private:
    Kind kind_;
    int nsites_, nbits_;
    Sector sec_;
    int A_ = -1, B_ = -1;
    bool empty_ = false;
    int64_t dim_ = 0;
    uint64_t up_mask_ = 0, dn_mask_ = 0;
};

struct Factor {
    std::string op;
    int site;
};

struct Term {
    cplx coeff;
    std::vector<Factor> factors;
};

struct OpSum {
    std::vector<Term> terms;
};

// Output of the compiler for one (basis, site kind, OpSum).
struct Compiled {
    std::vector<QkBranch> offdiag;
    std::vector<QkDiagTerm> diag;
    cplx dconst = 0.0;
    bool is_real = true;        // every value has zero imaginary part
    bool diag_only = false;
    int max_hops = 0;           // upper bound of off-diagonal entries per row
};

// Compile the OpSum against a basis and site kind. Throws std::runtime_error
// for an unknown operator name reached by a live branch (the reference's
// Site::apply throws) and std::invalid_argument for a site index that does
// not fit a 64-bit state.
Compiled compile_opsum(const BasisSpec& basis, Kind site_kind, const OpSum& ops);

// Split of the state bits into hi (uniform per tile) and lo (per lane).
struct HostTables {
    int NL = 0, NH = 0;
    std::vector<uint32_t> hoff, lrank, lolist;
    std::vector<QkTile> tiles;
    bool lrank_u16 = false;
    // Two-level lo rank for wide lo fields (NL > 14), small enough for LDS:
    // lo = (h << NL2) | l, rank(lo) = th[h * nc + c(l)] + tl[l], where
    // c(l) = popcount(l & upm2) * KB + popcount(l & dnm2) is the class of
    // the low half (a zero mask for a count the sector leaves free), tl[l]
    // the rank of l among the valid low halves of its class, and th[h][c]
    // the number of valid lo of the same total class whose high half is
    // below h.
    bool split = false;
    int NL2 = 0, nc = 0, KB = 1;
    uint32_t upm2 = 0, dnm2 = 0;
    std::vector<uint16_t> tl;
    std::vector<uint32_t> th;
};

HostTables build_tables(const BasisSpec& basis, int NL);
int choose_NL(const BasisSpec& basis);
// fraction of lanes holding a row when rows are tiled per hi block
double tile_utilisation(const BasisSpec& basis, int NL);
// split minimising (uniform + 5 x per-lane branch work) / utilisation
int choose_NL_model(const BasisSpec& basis, const Compiled& comp);

} // namespace qk
