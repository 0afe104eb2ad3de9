// qk_device.hpp -- plain-old-data tables shared by the host compiler and the
// HIP kernels. No HIP or STL types: this header is included on both sides.
#pragma once
#include <stdint.h>

namespace qk {

// One compiled off-diagonal branch of an OpSum term, acting on source state s:
//   taken when (s & cm) == cv, or (s & cm) == (cv ^ alt) when alt != 0
//   target t = s ^ F,   value = (re + i im) * (-1)^popcount(s & M)
// The value already holds conj(coeff * product of factor amplitudes) and the
// constant Jordan-Wigner sign, so the pull-form row sum is
//   y[row(s)] += value * x[row(t)]
// which is exactly the reference's stored CSR entry for that (row, col)
// (src/hamiltonian/matrix_free_hamiltonian.cpp:75-84).
struct QkBranch {
    uint64_t cm, cv, alt, F, M;
    double re, im;
};

// Diagonal contribution, evaluated per state:
//   kind 0: c * popcount(s & (s >> d) & mask)   (d = 0: linear term)
//   kind 1: c * [(s & mask) == mask]             (higher-order monomial)
//   kind 2: c * [(s & mask) == val]              (unexpanded condition)
struct QkDiagTerm {
    uint64_t mask, val;
    int32_t d, kind;
    double re, im;
};

// A run of <= 64 consecutive basis rows that share the same high part `hi`.
// Row `row_start + j` has state (hi << NL) | lolist[lo_base + j].
struct QkTile {
    uint32_t hi, lo_base, row_start, count;
};

struct QkTables {
    const QkTile* tiles;
    const uint32_t* hoff;     // [2^NH]  rows before the block of each hi
    const uint32_t* lrank;    // [2^NL]  rank of lo inside its class
    const uint32_t* lolist;   // ascending valid lo values, grouped by class
    const QkBranch* br;       // off-diagonal branches
    const QkDiagTerm* dg;     // diagonal terms
    uint64_t lomask;
    uint64_t dim;             // number of basis rows
    double dconst_re, dconst_im;
    uint32_t ntiles;
    int32_t nbr, ndg, NL;
    int32_t lrank_u16;        // 1: lrank values fit 16 bits (LDS copy is u16)
    int32_t nlo;              // 2^NL
};

} // namespace qk
