// ref_harness.cpp -- C entry points over the reference's OWN compiled basis,
// site and opsum sources (TEST INFRASTRUCTURE; built by oracle/Makefile into
// oracle/_ref/libqkref.so from the files where they lie under the reference
// checkout; nothing of the reference is copied into this repository).
//
// The reference's hamiltonian/solver/c_api sources need Kokkos (fetched from
// the network at configure time, CMakeLists.txt:57-82) and are unbuildable
// here. This harness therefore restates only the Phase-1 loop of
// src/hamiltonian/matrix_free_hamiltonian.cpp:45-92, driving the reference's
// compiled Basis::state/contains/index and Site::apply, so the emitted CSR
// triplets carry the reference's own state order, site actions and signs.
#include "qkrylov/basis/fermion_basis.hpp"
#include "qkrylov/basis/hubbard_basis.hpp"
#include "qkrylov/basis/spinhalf_basis.hpp"
#include "qkrylov/basis/tj_basis.hpp"
#include "qkrylov/operators/opsum.hpp"
#include "qkrylov/sites/fermion_site.hpp"
#include "qkrylov/sites/hubbard_site.hpp"
#include "qkrylov/sites/spinhalf_site.hpp"
#include "qkrylov/sites/tj_site.hpp"
#include "qkrylov/symmetry/sector.hpp"

#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

using namespace qkrylov;
using namespace qkrylov::fp64;

namespace {

std::shared_ptr<Basis> make_basis(int kind, int n, const int* sec)
{
    Sector s;
    s.use_sz = sec[0]; s.sz2 = sec[1];
    s.use_nup = sec[2]; s.nup = sec[3];
    s.use_ndn = sec[4]; s.ndn = sec[5];
    s.use_n = sec[6]; s.n = sec[7];
    switch (kind) {
    case 0: return std::make_shared<SpinHalfBasis>(n, s);
    case 1: return std::make_shared<FermionBasis>(n, s);
    case 2: return std::make_shared<HubbardBasis>(n, s);
    case 3: return std::make_shared<TJBasis>(n, s);
    }
    return nullptr;
}

std::shared_ptr<Site> make_site(int kind)
{
    switch (kind) {
    case 0: return std::make_shared<SpinHalfSite>();
    case 1: return std::make_shared<FermionSite>();
    case 2: return std::make_shared<HubbardSite>();
    case 3: return std::make_shared<TJSite>();
    }
    return nullptr;
}

} // namespace

extern "C" {

// sec = {use_sz, sz2, use_nup, nup, use_ndn, ndn, use_n, n}
// returns dim (all states written when cap >= dim), -1 if the ctor throws
int64_t qkref_basis_states(int kind, int n, const int* sec, uint64_t* out, int64_t cap)
{
    try {
        auto b = make_basis(kind, n, sec);
        const int64_t d = static_cast<int64_t>(b->size());
        for (int64_t i = 0; i < d && i < cap; ++i) out[i] = b->state(static_cast<Index>(i));
        return d;
    } catch (...) {
        return -1;
    }
}

// 1 valid / 0 invalid / -1 exception (unknown operator)
int qkref_site_apply(int kind, const char* op, int site, uint64_t state,
                     uint64_t* new_state, double* re, double* im)
{
    try {
        auto s = make_site(kind);
        LocalAction a = s->apply(op, site, state);
        if (!a.valid) return 0;
        *new_state = a.new_state;
        *re = a.matrix_element.real();
        *im = a.matrix_element.imag();
        return 1;
    } catch (...) {
        return -1;
    }
}

// Phase-1 CSR (matrix_free_hamiltonian.cpp:45-92) over reference objects.
// Terms: nterms, coeff re/im, nf per term, flattened ops/sites.
// Two-call protocol: first with cols == NULL returns nnz; second fills.
int64_t qkref_csr(int bkind, int n, const int* sec, int skind,
                  int nterms, const double* cre, const double* cim, const int* nf,
                  const char** ops, const int* sites,
                  int64_t* row_offsets, int64_t* cols, double* vals, double* diag)
{
    try {
        auto basis = make_basis(bkind, n, sec);
        auto site = make_site(skind);
        OpSum os;
        int off = 0;
        for (int t = 0; t < nterms; ++t) {
            OperatorTerm term;
            term.coeff = Complex(cre[t], cim[t]);
            for (int k = 0; k < nf[t]; ++k) term.factors.push_back({std::string(ops[off + k]), sites[off + k]});
            off += nf[t];
            os.add_term(term);
        }
        const Index dim = basis->size();
        int64_t nnz = 0;
        for (Index alpha = 0; alpha < dim; ++alpha) {
            if (row_offsets) row_offsets[alpha] = nnz;
            const StateID s0 = basis->state(alpha);
            for (const auto& term : os.terms()) {
                StateID st = s0;
                Complex amp = term.coeff;
                bool valid = true;
                for (const auto& f : term.factors) {
                    auto a = site->apply(f.op, f.site, st);
                    if (!a.valid) { valid = false; break; }
                    st = a.new_state;
                    amp *= a.matrix_element;
                }
                if (!valid) continue;
                if (!basis->contains(st)) continue;
                const Index beta = basis->index(st);
                if (cols) {
                    cols[nnz] = static_cast<int64_t>(beta);
                    vals[2 * nnz] = amp.real();
                    vals[2 * nnz + 1] = -amp.imag();
                    if (beta == alpha && diag) {
                        diag[2 * alpha] += amp.real();
                        diag[2 * alpha + 1] += -amp.imag();
                    }
                }
                ++nnz;
            }
        }
        if (row_offsets) row_offsets[dim] = nnz;
        return nnz;
    } catch (...) {
        return -1;
    }
}

} // extern "C"
