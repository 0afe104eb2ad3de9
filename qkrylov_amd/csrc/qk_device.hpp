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

// A branch split at the lo/hi boundary NL, as the apply kernels use it:
// the h* fields act on hi = s >> NL (wave-uniform inside a tile), the l*
// fields on lo = s & ((1 << NL) - 1) (per lane). Branches whose lo fields
// are all zero ("hi-only") come first: their condition, target block and
// sign are uniform, and their gather is one coalesced row segment.
struct QkBranchK {
    uint32_t hcm, hcv, halt, hF, hM;
    uint32_t lcm, lcv, lalt, lF, lM;
    double re, im;
};

struct QkTables {
    const QkTile* tiles;
    const uint32_t* hoff;     // [2^NH]  rows before the block of each hi
    const uint32_t* lrank;    // [2^NL]  rank of lo inside its class
    const uint32_t* lolist;   // ascending valid lo values, grouped by class
    const QkBranch* br;       // off-diagonal branches (full-width form)
    const QkBranchK* brk;     // the same branches split at NL: nA hi-only, then nB others (padded)
    const QkDiagTerm* dg;     // diagonal terms
    int32_t nA, nB;           // padded counts (multiples of the kernel's unroll)
    int32_t nB1, nB2;         // the first nB1 per-lane branches flip only low-half bits of lo and keep its
                              // class, the next nB2 flip only high-half bits (two-level rank only)
    uint64_t lomask;
    uint64_t dim;             // number of basis rows
    double dconst_re, dconst_im;
    uint32_t ntiles;
    int32_t nbr, ndg, NL;
    int32_t lrank_u16;        // 1: lrank values fit 16 bits (LDS copy is u16)
    int32_t nlo;              // 2^NL
    int32_t sched;            // tile schedule of the single-vector apply: 0 one contiguous segment per XCD,
                              // 1 one front over all XCDs, 2 one front of XCD-contiguous chunks
    int32_t sched_block;      // tile schedule of the multi-vector apply
    int32_t blk_rows;         // rows per work item of the multi-vector apply (64, 32 or 16; 0: 64)
    int32_t blk_waves;        // waves per workgroup of the multi-vector apply (16, 8 or 4; 0: 16)
    int32_t signs;            // 1: some branch carries a parity mask (fermionic signs)
    int32_t apply_waves;      // waves per workgroup of the single-vector apply (4, 8 or 16; 0: 4)
    // lo rank in LDS: 0 one-level uint32 table, 1 one-level uint16 table,
    // 2 two-level (tl[2^NL2] uint16, th[2^(NL-NL2) * nc] uint32):
    // rank(lo) = th[(lo >> NL2) * nc + c] + tl[lo & (2^NL2 - 1)],
    // c = popcount(l & upm2) * KB + popcount(l & dnm2)
    int32_t rank_mode;
    const uint16_t* tl;
    const uint32_t* th;
    int32_t NL2, nc, KB, ntl, nth;
    uint32_t upm2, dnm2;
};

} // namespace qk
