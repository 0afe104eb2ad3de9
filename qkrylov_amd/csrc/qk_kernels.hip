// qk_kernels.hip -- CDNA4 (gfx950) kernels of the matrix-free engine.
//
// Pull form: every output row is produced by exactly one lane, so there are
// no atomics and results are deterministic. A wavefront owns a tile: up to 64
// consecutive basis rows that share the high bits `hi` of their state, so
// everything that depends on hi alone (term conditions on high bits, the
// target block offset hoff[hi ^ F_hi], the high part of the Jordan-Wigner
// parity) is wave-uniform and runs on the scalar unit. Each lane decodes its
// low bits from a coalesced lolist load, applies the compiled branches with
// bit operations, ranks the target through the LDS-resident lo-rank table,
// and gathers x. Reductions (Lanczos dot products and norms) are two-pass and
// fixed-order: per-workgroup partials, then one workgroup sums them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>

#include <math.h>

#include "qk_launch.h"

using namespace qk;

#define QK_TPB 256
#define QK_WPB 4
#define QK_BLAS_BLOCKS 1024

namespace {

struct QkScal {
    uint32_t lomask;
    uint32_t ntiles;
    int32_t nbr, ndg, NL, nlo, nA, nB, sched;
    int32_t tsh;              // block apply: each tile is 2^tsh work items of 64 >> tsh rows
    int32_t NL2, nc, KB, ntl, nB1, nB2;
    uint32_t upm2, dnm2;
    double dcr, dci;
};

QkScal make_scal(const QkTables* T)
{
    QkScal s;
    s.lomask = static_cast<uint32_t>(T->lomask);
    s.ntiles = T->ntiles;
    s.nA = T->nA;
    s.nB = T->nB;
    s.nbr = T->nbr;
    s.ndg = T->ndg;
    s.NL = T->NL;
    s.nlo = T->nlo;
    s.sched = T->sched;
    s.tsh = 0;
    s.NL2 = T->NL2;
    s.nc = T->nc;
    s.KB = T->KB;
    s.ntl = T->ntl;
    s.upm2 = T->upm2;
    s.dnm2 = T->dnm2;
    s.nB1 = T->nB1;
    s.nB2 = T->nB2;
    s.dcr = T->dconst_re;
    s.dci = T->dconst_im;
    return s;
}

__device__ __forceinline__ uint32_t rfl(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// two doubles as one 16 B vector access
typedef double qk_double2 __attribute__((ext_vector_type(2)));

template <typename ST>
__device__ __forceinline__ int popc(ST x);
template <>
__device__ __forceinline__ int popc<uint32_t>(uint32_t x) { return __builtin_popcount(x); }
template <>
__device__ __forceinline__ int popc<uint64_t>(uint64_t x) { return __builtin_popcountll(x); }

// Tile schedules. Workgroups are dispatched round-robin over the 8 XCDs, so
// workgroup b runs on XCD b % 8.
//  sched 0: each XCD gets one contiguous segment of tiles and its waves
//           sweep that segment together, keeping the rows (and the in-block
//           gather targets) in flight on one XCD inside one window of its L2;
//  sched 1: all waves of the grid sweep the tiles together in one front
//           (wave g takes tiles g, g + G, ...), so the rows in flight on all
//           XCDs share one window of the memory-side cache, which then also
//           holds the targets of branches that leave the block;
//  sched 2: one front as in 1, cut into XCD-contiguous chunks of as many
//           tiles as an XCD has waves: chunk j runs on XCD j % 8, so a block
//           of consecutive tiles stays on one XCD (its L2 serves the in-block
//           and repeated targets) while the eight XCDs still sweep one narrow
//           window of the memory-side cache.
__device__ __forceinline__ void tile_schedule(uint32_t ntiles, int sched, uint32_t& t0, uint32_t& tend,
                                              uint32_t& step)
{
    const uint32_t nb = gridDim.x, b = blockIdx.x;
    const uint32_t w = rfl(threadIdx.x >> 6), wpb = blockDim.x >> 6;
    if (sched != 1 && nb >= 8 && (nb & 7u) == 0) {
        const uint32_t xcd = b & 7u, loc = b >> 3, nloc = nb >> 3;
        if (sched == 0) {
            const uint32_t seg = (ntiles + 7u) / 8u;
            const uint32_t s0 = min(xcd * seg, ntiles), s1 = min(s0 + seg, ntiles);
            t0 = s0 + loc * wpb + w;
            tend = s1;
            step = nloc * wpb;
        } else {
            const uint32_t chunk = nloc * wpb;
            t0 = xcd * chunk + loc * wpb + w;
            tend = ntiles;
            step = 8u * chunk;
        }
    } else {
        t0 = b * wpb + w;
        tend = ntiles;
        step = nb * wpb;
    }
}

template <typename ST, bool CPLX>
__device__ __forceinline__ void eval_diag(const QkDiagTerm* __restrict__ dg, int ndg, double dcr, double dci, ST s,
                                          double& dre, double& dim)
{
    dre = dcr;
    dim = dci;
    for (int i = 0; i < ndg; ++i) {
        const QkDiagTerm g = dg[i];
        const ST m = static_cast<ST>(g.mask);
        int v;
        if (g.kind == 0) v = popc<ST>(s & (s >> g.d) & m);
        else if (g.kind == 1) v = ((s & m) == m) ? 1 : 0;
        else v = ((s & m) == static_cast<ST>(g.val)) ? 1 : 0;
        dre = fma(g.re, static_cast<double>(v), dre);
        if (CPLX) dim = fma(g.im, static_cast<double>(v), dim);
    }
}

// Lo rank from the LDS tables. RM 0 / 1: one-level uint32 / uint16 table of
// 2^NL entries. RM 2: two-level, lo = (h << NL2) | l,
// rank = th[h * nc + c(l)] + tl[l] with the class c(l) of the low half.
__host__ __device__ __forceinline__ size_t tl_bytes(int ntl) { return ((size_t)ntl * 2 + 15) & ~(size_t)15; }

template <int RM>
__device__ __forceinline__ uint32_t lds_rank(const unsigned char* smem, uint32_t lo, const QkScal& sc)
{
    if (RM == 1) return reinterpret_cast<const uint16_t*>(smem)[lo];
    if (RM == 0) return reinterpret_cast<const uint32_t*>(smem)[lo];
    const uint32_t l = lo & ((1u << sc.NL2) - 1u), h = lo >> sc.NL2;
    const uint32_t c = (uint32_t)__builtin_popcount(l & sc.upm2) * (uint32_t)sc.KB + (uint32_t)__builtin_popcount(l & sc.dnm2);
    const uint16_t* tl = reinterpret_cast<const uint16_t*>(smem);
    const uint32_t* th = reinterpret_cast<const uint32_t*>(smem + tl_bytes(sc.ntl));
    return (uint32_t)tl[l] + th[h * (uint32_t)sc.nc + c];
}

template <int RM>
__device__ __forceinline__ void load_lrank(unsigned char* smem, const uint32_t* __restrict__ lrank,
                                           const uint16_t* __restrict__ tl, const uint32_t* __restrict__ th,
                                           int nlo, int ntl, int nth)
{
    if (RM == 2) {
        for (int i = threadIdx.x; i < ntl; i += blockDim.x) reinterpret_cast<uint16_t*>(smem)[i] = tl[i];
        uint32_t* d = reinterpret_cast<uint32_t*>(smem + tl_bytes(ntl));
        for (int i = threadIdx.x; i < nth; i += blockDim.x) d[i] = th[i];
    } else {
        for (int i = threadIdx.x; i < nlo; i += blockDim.x) {
            if (RM == 1) reinterpret_cast<uint16_t*>(smem)[i] = static_cast<uint16_t>(lrank[i]);
            else reinterpret_cast<uint32_t*>(smem)[i] = lrank[i];
        }
    }
    __syncthreads();
}

// Rank of a target lo = own lo ^ lF. KIND 1: lF inside the low half and the
// low half keeps its class, so the high-half term is the row's own; KIND 2:
// lF inside the high half, so the low-half term is the row's own; KIND 0:
// the full lookup. With one-level tables every kind is one lookup.
struct OwnRank {
    uint32_t h, c, th_own, tl_own;
};

template <int RM>
__device__ __forceinline__ OwnRank own_rank(const unsigned char* smem, uint32_t lo, const QkScal& sc)
{
    OwnRank o{0, 0, 0, 0};
    if (RM == 2) {
        const uint32_t l = lo & ((1u << sc.NL2) - 1u);
        o.h = lo >> sc.NL2;
        o.c = (uint32_t)__builtin_popcount(l & sc.upm2) * (uint32_t)sc.KB + (uint32_t)__builtin_popcount(l & sc.dnm2);
        const uint32_t* th = reinterpret_cast<const uint32_t*>(smem + tl_bytes(sc.ntl));
        o.th_own = th[o.h * (uint32_t)sc.nc + o.c];
        o.tl_own = reinterpret_cast<const uint16_t*>(smem)[l];
    }
    return o;
}

template <int RM, int KIND>
__device__ __forceinline__ uint32_t target_rank(const unsigned char* smem, uint32_t lo, uint32_t lF, const OwnRank& o,
                                                const QkScal& sc)
{
    if (RM != 2 || KIND == 0) return lds_rank<RM>(smem, lo ^ lF, sc);
    if (KIND == 1) return o.th_own + reinterpret_cast<const uint16_t*>(smem)[(lo ^ lF) & ((1u << sc.NL2) - 1u)];
    const uint32_t* th = reinterpret_cast<const uint32_t*>(smem + tl_bytes(sc.ntl));
    return th[(o.h ^ (lF >> sc.NL2)) * (uint32_t)sc.nc + o.c] + o.tl_own;
}

// acc += sum_u c_u * x[rt_u], all U loads issued before the first FMA;
// the FMAs run in u order, so the sum order is the branch order
template <bool CPLX, int U>
__device__ __forceinline__ void accumulate(const double* __restrict__ x, const uint32_t* rt, const double* cr,
                                           const double* ci, double& ar, double& ai)
{
    if (CPLX) {
        double tr[U], ti[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            tr[u] = x[2 * (size_t)rt[u]];
            ti[u] = x[2 * (size_t)rt[u] + 1];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            ar = fma(cr[u], tr[u], ar);
            ar = fma(-ci[u], ti[u], ar);
            ai = fma(cr[u], ti[u], ai);
            ai = fma(ci[u], tr[u], ai);
        }
    } else {
        double t[U];
#pragma unroll
        for (int u = 0; u < U; ++u) t[u] = x[rt[u]];
#pragma unroll
        for (int u = 0; u < U; ++u) ar = fma(cr[u], t[u], ar);
    }
}

// as accumulate, but only the lanes whose ok[u] is set load x[rt[u]]: the
// others take 0.0 (their coefficient is 0 as well) and issue no memory access
// for that entry, which spares the vector memory unit the cache lines of
// the lanes a branch does not reach
template <bool CPLX, int U>
__device__ __forceinline__ void accumulate_masked(const double* __restrict__ x, const uint32_t* rt, const bool* ok,
                                                  const double* cr, const double* ci, double& ar, double& ai)
{
    if (CPLX) {
        double tr[U], ti[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            tr[u] = 0.0;
            ti[u] = 0.0;
            if (ok[u]) {
                tr[u] = x[2 * (size_t)rt[u]];
                ti[u] = x[2 * (size_t)rt[u] + 1];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            ar = fma(cr[u], tr[u], ar);
            ar = fma(-ci[u], ti[u], ar);
            ai = fma(cr[u], ti[u], ai);
            ai = fma(ci[u], tr[u], ai);
        }
    } else {
        double t[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            t[u] = 0.0;
            if (ok[u]) t[u] = x[rt[u]];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) ar = fma(cr[u], t[u], ar);
    }
}

__device__ __forceinline__ double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Fixed-order total of n partial sums, by the first wave of the workgroup:
// lane l adds p[l], p[l + 64], ... in index order, then the butterfly of
// wave_sum, and lane 0's value is broadcast through *sh. Every workgroup
// (and k_total) that totals the same partials gets the same bits. All
// threads of the workgroup must call it.
__device__ __forceinline__ double block_total(const double* __restrict__ p, int n, double* sh)
{
    if (threadIdx.x < 64) {
        const int l = (int)threadIdx.x;
        double s = 0.0;
        if (n <= 64 * 32) {
            // batches of 16 loads in flight, then their adds in index order
            // (absent terms add +0.0); a batch holds 1024 partials, which
            // covers every launch shape but the smallest workgroups. Sixteen
            // doubles per lane keep this prologue's register peak below the
            // apply loop's, so it does not lower the apply's occupancy.
            for (int j0 = 0; j0 < 32 && 64 * j0 < n; j0 += 16) {
                double v[16];
#pragma unroll
                for (int j = 0; j < 16; ++j) {
                    const int i = l + 64 * (j0 + j);
                    v[j] = p[i < n ? i : 0];
                    v[j] = i < n ? v[j] : 0.0;
                }
#pragma unroll
                for (int j = 0; j < 16; ++j) s += v[j];
            }
        } else {
            for (int i = l; i < n; i += 64) s += p[i];
        }
        s = wave_sum(s);
        if (l == 0) *sh = s;
    }
    __syncthreads();
    return *sh;
}

// ------------------------------------------------------------------ apply
// Per-wave LDS scratch of the apply kernel: for every per-lane branch the
// tile's uniform data (target block offset; the two values lo & lcm may take
// for the branch to apply, given what the tile's hi allows, or an impossible
// value; hi parity), and the compacted list of the hi-only branches the
// tile's hi takes (value, imaginary part for complex vectors, block offset).
__host__ __device__ __forceinline__ size_t apply_wave_bytes(int nB, int nA, bool cplx)
{
    return (((size_t)nB * 16 + (size_t)nA * (cplx ? 20 : 12)) + 15) & ~(size_t)15;
}

// Per-workgroup LDS copy of the lo side of the per-lane branches: 16 B of
// {lcm, lcv, lalt, lF}, then lM, re, im; read as broadcasts in the hot loop.
__host__ __device__ __forceinline__ size_t apply_branch_bytes(int nB)
{
    return (((size_t)nB * (16 + 4 + 8 + 8)) + 15) & ~(size_t)15;
}

// Per-lane branches [gbeg, gend) of one tile, eight gathers in flight. A
// branch that does not apply to a lane (or a lane past the tile's rows)
// contributes with coefficient 0. Spin models (no parity masks) then issue
// no load for that lane: the vector memory unit is their bound, and the
// lines of lanes a branch does not reach cost it time (the 32-site ring's
// apply 16 % faster, J1-J2 4x8 4 %). Fermion models, bound by their VALU
// work, gather the row's own x instead, which needs no exec-mask changes.
// The rank lookups and value loads of the eight branches overlap either way.
template <bool CPLX, int RM, bool SIGNS, int KIND>
__device__ __forceinline__ void lane_branches(int gbeg, int gend, const unsigned char* smem, const uint4* lb4,
                                              const uint32_t* lbm, const double* lbre, const double* lbim,
                                              const uint4* tb, uint32_t lo, uint32_t row, bool active,
                                              const OwnRank& orank, const QkScal& sc, const double* __restrict__ x,
                                              double& ar, double& ai)
{
    for (int g = gbeg; g < gend; g += 8) {
        uint32_t rt[8];
        bool ok[8];
        double cr[8], ci[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const uint4 L = lb4[g + u];            // lcm, lcv, lalt, lF
            const uint4 tv = tb[g + u];            // block offset, va, vb, hpar
            const uint32_t lf = lo & L.x;
            const uint32_t m = 0u - (uint32_t)((lf == tv.y) | (lf == tv.z));
            const uint32_t lr = target_rank<RM, KIND>(smem, lo, L.w, orank, sc);
            uint64_t br = __builtin_bit_cast(uint64_t, lbre[g + u]);
            uint64_t bi = CPLX ? __builtin_bit_cast(uint64_t, lbim[g + u]) : 0ull;
            if (SIGNS) {
                const uint64_t neg = (uint64_t)(((uint32_t)__builtin_popcount(lo & lbm[g + u]) ^ tv.w) & 1u) << 63;
                br ^= neg;
                bi ^= neg;
            }
            const uint64_t m64 = ((uint64_t)m << 32) | m;
            rt[u] = SIGNS ? row ^ ((row ^ (tv.x + lr)) & m) : tv.x + lr;
            ok[u] = active && m != 0u;
            cr[u] = __builtin_bit_cast(double, br & m64);
            ci[u] = __builtin_bit_cast(double, bi & m64);
        }
        if (SIGNS) accumulate<CPLX, 8>(x, rt, cr, ci, ar, ai);
        else accumulate_masked<CPLX, 8>(x, rt, ok, cr, ci, ar, ai);
    }
}

// The real Lanczos form is held to 80 VGPRs (six waves per SIMD): left
// alone, its epilogue's partial dot raises the register count past the
// plain apply's and costs waves per CU in the 8- and 4-wave shapes. Other
// instances keep the compiler's choice.
// ACC: y = acc + (this pass's sum), the second pass of the split apply
// (qk_split.hpp); real 32-bit spin tables only.
template <typename ST, bool CPLX, int RM, bool LZ, bool SIGNS, bool ACC>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu((LZ && !CPLX) ? 6 : 1))) void k_apply(const QkTile* __restrict__ tiles, const uint32_t* __restrict__ hoff,
                                                   const uint32_t* __restrict__ lrank, const uint16_t* __restrict__ tl,
                                                   const uint32_t* __restrict__ th, int nth,
                                                   const uint32_t* __restrict__ lolist,
                                                   const QkBranchK* __restrict__ brk,
                                                   const QkDiagTerm* __restrict__ dg, QkScal sc,
                                                   const double* __restrict__ x, double* __restrict__ y, int lz_k,
                                                   const double* __restrict__ prev, double* __restrict__ nrm,
                                                   double* __restrict__ partials, int table_bytes,
                                                   const int* __restrict__ kdev, int koff,
                                                   const double* __restrict__ nrm_part, int n_nrm_part,
                                                   const double* __restrict__ acc, double* __restrict__ nrm_host)
{
    extern __shared__ __align__(16) unsigned char smem[];
    // static LDS: QK_APPLY_STATIC_LDS bytes, counted by apply_lds_plan
    __shared__ double wsum[16];
    __shared__ double tot_sh;
    const int nA = sc.nA, nB = sc.nB, nbk = sc.nA + sc.nB;
    uint4* lb4 = reinterpret_cast<uint4*>(smem + table_bytes);
    uint32_t* lbm = reinterpret_cast<uint32_t*>(lb4 + nB);
    double* lbre = reinterpret_cast<double*>(smem + table_bytes + ((size_t)nB * 20 + 7) / 8 * 8);
    double* lbim = lbre + nB;
    for (int i = threadIdx.x; i < nB; i += blockDim.x) {
        const QkBranchK B = brk[nA + i];
        lb4[i] = make_uint4(B.lcm, B.lcv, B.lalt, B.lF);
        lbm[i] = B.lM;
        lbre[i] = B.re;
        lbim[i] = B.im;
    }
    load_lrank<RM>(smem, lrank, tl, th, sc.nlo, sc.ntl, nth);

    const uint32_t lane = threadIdx.x & 63u;
    const int NL = sc.NL;
    unsigned char* wb = smem + table_bytes + apply_branch_bytes(nB) + (threadIdx.x >> 6) * apply_wave_bytes(nB, nA, CPLX);
    uint4* tb = reinterpret_cast<uint4*>(wb);                          // [nB] {block offset, va, vb, hpar}
    double* a_cr = reinterpret_cast<double*>(wb + (size_t)nB * 16);    // [nA] taken hi-only values
    double* a_ci = a_cr + nA;                                          // [nA] (complex only)
    uint32_t* a_base = reinterpret_cast<uint32_t*>(a_cr + (CPLX ? 2 : 1) * nA);   // [nA] their block offsets
    const uint64_t lt_mask = (lane == 0) ? 0ULL : (~0ULL >> (64 - lane));

    // Lanczos step: from the argument, or from the device step counter
    // (plus koff) when the launch is replayed from a captured graph;
    // nrm[k] = |q_k| is totalled here from the previous update's partial
    // sums (and stored by workgroup 0) instead of by a separate reduction
    // launch.
    const int lzk = (LZ && kdev) ? *kdev + koff : lz_k;
    double inv_n = 1.0, cprev = 0.0;
    if (LZ) {
        const double nk = sqrt(block_total(nrm_part, n_nrm_part, &tot_sh));
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            nrm[lzk] = nk;
            if (nrm_host) nrm_host[lzk] = nk;   // the host's copy (coherent pinned memory)
        }
        inv_n = 1.0 / nk;
        if (lzk > 0) cprev = nk / nrm[lzk - 1];
    }
    double part = 0.0;

    uint32_t t0, tend, step;
    tile_schedule(sc.ntiles, sc.sched, t0, tend, step);
    for (uint32_t t = t0; t < tend; t += step) {
        const QkTile tl = tiles[t];
        const uint32_t hi = tl.hi;
        const bool active = lane < tl.count;
        const uint32_t base_self = hoff[hi];
        const uint32_t lidx = active ? lane : 0u;
        const uint32_t my_lr = tl.row_start - base_self + lidx;
        const uint32_t row = tl.row_start + lidx;
        const uint32_t lo = lolist[tl.lo_base + lidx];
        // the Lanczos epilogue's q_{k-1} entry, loaded now so that its
        // latency hides behind the tile's gathers
        const double pv = (LZ && lzk > 0) ? __builtin_nontemporal_load(prev + row) : 0.0;
        const double av = ACC ? __builtin_nontemporal_load(acc + row) : 0.0;
        const ST s = (static_cast<ST>(hi) << NL) | static_cast<ST>(lo);

        // (1) the uniform part of every branch, one branch per lane
        uint32_t ntaken = 0;
        for (int b0 = 0; b0 < nbk; b0 += 64) {
            const int b = b0 + (int)lane;
            bool takeA = false;
            uint32_t base = 0;
            double vr = 0.0, vi = 0.0;
            if (b < nbk) {
                const QkBranchK B = brk[b];
                const uint32_t hm = hi & B.hcm;
                const uint32_t u0 = (hm == B.hcv) ? 1u : 0u;
                const uint32_t u1 = (((B.halt | B.lalt) != 0) && (hm == (B.hcv ^ B.halt))) ? 2u : 0u;
                const uint32_t hpar = SIGNS ? (uint32_t)(__builtin_popcount(hi & B.hM) & 1) : 0u;
                // lo & lcm < 2^NL <= 2^20, so ~lcm is a value it never takes
                const uint32_t never = ~B.lcm;
                base = hoff[hi ^ B.hF];
                if (b >= nA) tb[b - nA] = make_uint4(base, u0 ? B.lcv : never, u1 ? (B.lcv ^ B.lalt) : never, hpar);
                takeA = (b < nA) && (u0 | u1);
                vr = hpar ? -B.re : B.re;
                vi = hpar ? -B.im : B.im;
            }
            // (2) compact the taken hi-only branches
            const uint64_t m = __ballot(takeA);
            if (takeA) {
                const uint32_t pos = ntaken + (uint32_t)__popcll(m & lt_mask);
                a_base[pos] = base;
                a_cr[pos] = vr;
                if (CPLX) a_ci[pos] = vi;
            }
            ntaken += (uint32_t)__popcll(m);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");

        double dre, dim;
        eval_diag<ST, CPLX>(dg, sc.ndg, sc.dcr, sc.dci, s, dre, dim);
        double ar, ai = 0.0, xr, xi = 0.0;
        if (CPLX) {
            xr = x[2 * (size_t)row];
            xi = x[2 * (size_t)row + 1];
            ar = dre * xr - dim * xi;
            ai = dre * xi + dim * xr;
        } else {
            xr = x[row];
            ar = dre * xr;
        }

        // (3) taken hi-only branches: coalesced gathers of the tile's own
        // lo-rank range in each target block, four in flight
        for (uint32_t i = 0; i < ntaken; i += 4) {
            uint32_t rt[4];
            bool ok[4];
            double cr[4], ci[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const bool in = (i + u) < ntaken;
                const uint32_t k = in ? i + u : 0u;
                rt[u] = in ? a_base[k] + my_lr : row;
                ok[u] = in && active;
                cr[u] = in ? a_cr[k] : 0.0;
                ci[u] = (CPLX && in) ? a_ci[k] : 0.0;
            }
            if (SIGNS) accumulate<CPLX, 4>(x, rt, cr, ci, ar, ai);
            else accumulate_masked<CPLX, 4>(x, rt, ok, cr, ci, ar, ai);
        }

        // (4) remaining branches: per-lane test of lo & lcm against the two
        // tile-uniform values, lo rank from LDS, eight in flight; with the
        // two-level rank, three runs by rank kind
        const OwnRank orank = own_rank<RM>(smem, lo, sc);
        if (RM == 2) {
            lane_branches<CPLX, RM, SIGNS, 1>(0, sc.nB1, smem, lb4, lbm, lbre, lbim, tb, lo, row, active, orank, sc,
                                              x, ar, ai);
            lane_branches<CPLX, RM, SIGNS, 2>(sc.nB1, sc.nB1 + sc.nB2, smem, lb4, lbm, lbre, lbim, tb, lo, row,
                                              active, orank, sc, x, ar, ai);
            lane_branches<CPLX, RM, SIGNS, 0>(sc.nB1 + sc.nB2, nB, smem, lb4, lbm, lbre, lbim, tb, lo, row, active,
                                              orank, sc, x, ar, ai);
        } else {
            lane_branches<CPLX, RM, SIGNS, 0>(0, nB, smem, lb4, lbm, lbre, lbim, tb, lo, row, active, orank, sc, x, ar,
                                              ai);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");

        // y is written once and not read again by this kernel: stream it
        // past the caches so L2 keeps x
        if (ACC) ar += av;
        if (active) {
            if (LZ) {
                double w = ar * inv_n;
                if (lzk > 0) w = fma(-cprev, pv, w);
                __builtin_nontemporal_store(w, y + row);
                part = fma(xr * inv_n, w, part);
            } else if (CPLX) {
                __builtin_nontemporal_store(ar, y + 2 * (size_t)row);
                __builtin_nontemporal_store(ai, y + 2 * (size_t)row + 1);
            } else {
                __builtin_nontemporal_store(ar, y + row);
            }
        }
    }
    if (LZ) {
        part = wave_sum(part);
        if (lane == 0) wsum[threadIdx.x >> 6] = part;
        __syncthreads();
        if (threadIdx.x == 0) {
            double s = 0.0;
            for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += wsum[i];
            partials[blockIdx.x] = s;
        }
    }
}

// ------------------------------------------------------------------ gather-rate probe
// The rate at which waves issue coalesced gathers of random 64-lane
// segments (W bytes per lane) from an array of nseg segments: every wave
// runs `tiles` rounds of 16 gathers, four in flight, as the apply's hi-only
// gathers do. A measurement of the vector memory path, not part of any
// solver.
__device__ __forceinline__ uint64_t probe_mix(uint64_t v)
{
    v ^= v >> 33;
    v *= 0xff51afd7ed558ccdULL;
    v ^= v >> 33;
    v *= 0xc4ceb9fe1a85ec53ULL;
    v ^= v >> 33;
    return v;
}

template <int W>
__global__ __launch_bounds__(512) void k_gather_probe(const double* __restrict__ x, size_t nseg,
                                                      double* __restrict__ y, uint32_t tiles)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t wave = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    double acc = 0.0;
    for (uint32_t t = 0; t < tiles; ++t) {
        const uint64_t key = wave * tiles + t;
#pragma unroll
        for (int g0 = 0; g0 < 16; g0 += 4) {
            double v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const uint64_t seg = probe_mix(key * 131 + (uint64_t)(g0 + u)) % nseg;
                if (W == 8) {
                    v[u] = x[seg * 64 + lane];
                } else {
                    const qk_double2 p2 = reinterpret_cast<const qk_double2*>(x)[seg * 64 + lane];
                    v[u] = p2.x + p2.y;
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) acc += v[u];
        }
    }
    y[wave * 64 + lane] = acc;
}

// ------------------------------------------------------------------ block apply
// Y = H X for up to 64 real columns of a row-major block (row stride ld).
// One wave per tile, up to 16 waves per workgroup sharing the lo-rank table.
// Phase A (lane = row) decodes the rows of the tile into one flat LDS list
// of (target row, value id) entries in row order: the row's diagonal first
// (target = the row itself, id = kDiag + j), then, in the Lanczos form, the
// row of the previous vector (id = kPrev), then the off-diagonal entries in
// branch order. Phase B (lane = column) walks the list eight entries at a
// time with the next eight gathers already in flight (two register sets
// used in turn), so row boundaries do not drain the memory pipeline. Every
// decoded entry is reused by all columns, and every gather is one
// contiguous segment of a row of X.

// per-workgroup LDS after the lo-rank table: lo fields of the per-lane
// branches ({lcm, lcv, lalt, lF} as 16 B, then lM), then +-value of every
// branch (value id 2b + sign)
__host__ __device__ __forceinline__ size_t block_branch_bytes(int nB) { return ((size_t)nB * 20 + 15) & ~(size_t)15; }

// per-wave LDS (cap a multiple of 8): entry targets [cap] uint32, entry
// value ids [cap] uint16, the rows' diagonals [64] double, the tile data of
// every branch [nbk] uint2, the taken hi-only list [nA] x (block offset,
// value id)
__host__ __device__ __forceinline__ size_t block_wave_bytes(int nbk, int nA, int cap)
{
    return ((size_t)cap * 6 + 64 * 8 + (size_t)nbk * 8 + (size_t)nA * 8 + 15) & ~(size_t)15;
}

// issue the gathers of entries [e0, e0 + G) of the list (entries at or past
// `total` gather row `tfill` of X and carry id kNop, so the walk can always
// keep the next group in flight); e0 is a multiple of G and the list is
// 16-byte aligned, so targets and ids are read as 16-byte LDS vectors
template <int G>
__device__ __forceinline__ void block_gather(const uint32_t* l_tgt, const uint16_t* l_vid, uint32_t e0,
                                             uint32_t total, uint32_t tfill, const double* __restrict__ Xc,
                                             const double* __restrict__ Pc, size_t ld, uint32_t kPrev,
                                             uint32_t kNop, double* xg, uint32_t* vg)
{
    const uint32_t eb = e0 < total ? e0 : 0u;
    uint32_t tg[G], vd[G];
#pragma unroll
    for (int q = 0; q < G / 4; ++q) {
        const uint4 t4 = *reinterpret_cast<const uint4*>(l_tgt + eb + 4 * q);
        tg[4 * q] = t4.x;
        tg[4 * q + 1] = t4.y;
        tg[4 * q + 2] = t4.z;
        tg[4 * q + 3] = t4.w;
    }
#pragma unroll
    for (int q = 0; q < G / 8; ++q) {
        const uint4 v4 = *reinterpret_cast<const uint4*>(l_vid + eb + 8 * q);
        const uint32_t w4[4] = {v4.x, v4.y, v4.z, v4.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            vd[8 * q + 2 * k] = w4[k] & 0xFFFFu;
            vd[8 * q + 2 * k + 1] = w4[k] >> 16;
        }
    }
#pragma unroll
    for (int u = 0; u < G; ++u) {
        const bool in = e0 + (uint32_t)u < total;
        uint32_t t = __builtin_amdgcn_readfirstlane(tg[u]);
        uint32_t v = __builtin_amdgcn_readfirstlane(vd[u]);
        t = in ? t : tfill;
        v = in ? v : kNop;
        const double* src = (v == kPrev) ? Pc : Xc;
        xg[u] = src[(size_t)t * ld];
        vg[u] = v;
    }
}

template <typename ST, int RM, bool LZ, int NW>
__global__ __launch_bounds__(64 * NW) void k_apply_block(
    const QkTile* __restrict__ tiles, const uint32_t* __restrict__ hoff, const uint32_t* __restrict__ lrank,
    const uint16_t* __restrict__ tl, const uint32_t* __restrict__ th, int nth,
    const uint32_t* __restrict__ lolist, const QkBranchK* __restrict__ brk, const QkDiagTerm* __restrict__ dg,
    QkScal sc, const double* __restrict__ X, double* __restrict__ Y, int ld, int c0, int nc, int lz_k,
    const double* __restrict__ P, const double* __restrict__ nrm, int ld_sc, double* __restrict__ partials,
    int npart, int cap, int table_bytes)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int nA = sc.nA, nB = sc.nB, nbk = sc.nA + sc.nB;
    uint4* lb4 = reinterpret_cast<uint4*>(smem + table_bytes);
    uint32_t* lbm = reinterpret_cast<uint32_t*>(lb4 + nB);
    double* l_vals = reinterpret_cast<double*>(smem + table_bytes + block_branch_bytes(nB));
    for (int i = threadIdx.x; i < nB; i += blockDim.x) {
        const QkBranchK B = brk[nA + i];
        lb4[i] = make_uint4(B.lcm, B.lcv, B.lalt, B.lF);
        lbm[i] = B.lM;
    }
    for (int i = threadIdx.x; i < nbk; i += blockDim.x) {
        const double v = brk[i].re;
        l_vals[2 * i] = v;
        l_vals[2 * i + 1] = -v;
    }
    load_lrank<RM>(smem, lrank, tl, th, sc.nlo, sc.ntl, nth);

    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wv = threadIdx.x >> 6;
    const size_t wave0 = (size_t)table_bytes + block_branch_bytes(nB) + (size_t)nbk * 16;
    unsigned char* wb = smem + wave0 + (size_t)wv * block_wave_bytes(nbk, nA, cap);
    uint32_t* l_tgt = reinterpret_cast<uint32_t*>(wb);                                   // [cap]
    uint16_t* l_vid = reinterpret_cast<uint16_t*>(wb + (size_t)cap * 4);                 // [cap]
    double* l_diag = reinterpret_cast<double*>(wb + (size_t)cap * 6);                    // [64]
    uint2* tb = reinterpret_cast<uint2*>(l_diag + 64);                                   // [nbk]
    uint32_t* a_base = reinterpret_cast<uint32_t*>(tb + nbk);                            // [nA]
    uint32_t* a_vid = a_base + nA;                                                       // [nA]
    const uint64_t lt_mask = (lane == 0) ? 0ULL : (~0ULL >> (64 - lane));
    const uint32_t kDiag = 2u * (uint32_t)nbk, kPrev = kDiag + 64u, kNop = 0xFFFFu, kNone = 0xFFFFFFFFu;

    const int NL = sc.NL;
    const bool colok = lane < (uint32_t)nc;
    const int col = c0 + (int)(colok ? lane : 0u);
    const bool use_prev = LZ && lz_k > 0;
    double inv_n = 1.0, cprev = 0.0;
    if (LZ) {
        inv_n = 1.0 / nrm[(size_t)lz_k * ld_sc + col];
        if (lz_k > 0) cprev = nrm[(size_t)lz_k * ld_sc + col] / nrm[(size_t)(lz_k - 1) * ld_sc + col];
    }
    const double* Xc = X + col;
    const double* Pc = use_prev ? P + col : Xc;
    double* Yc = Y + col;
    const size_t lds = (size_t)ld;
    double part = 0.0;

    // work item it = rows [wbeg, wbeg + (64 >> tsh)) of tile it >> tsh: smaller
    // items narrow the rows in flight (and the window of targets they share)
    // at the cost of repeating the tile's uniform branch preparation
    uint32_t t0, tend, step;
    tile_schedule(sc.ntiles << sc.tsh, sc.sched, t0, tend, step);
    const uint32_t rsub = 64u >> sc.tsh;
    for (uint32_t it = t0; it < tend; it += step) {
        const uint32_t t = it >> sc.tsh;
        const uint32_t wbeg = (it & ((1u << sc.tsh) - 1u)) * rsub;
        const QkTile tl = tiles[t];
        const uint32_t hi = tl.hi;
        const uint32_t count = tl.count;
        if (wbeg >= count) continue;
        const uint32_t wend = min(count, wbeg + rsub);
        const bool active = lane >= wbeg && lane < wend;
        const uint32_t base_self = hoff[hi];
        const uint32_t lidx = lane < count ? lane : 0u;
        const uint32_t my_lr = tl.row_start - base_self + lidx;
        const uint32_t lo = lolist[tl.lo_base + lidx];
        const ST s = (static_cast<ST>(hi) << NL) | static_cast<ST>(lo);

        // ---- phase A.1: the uniform half of every branch, one per lane,
        // and the compacted list of the hi-only branches the tile takes
        uint32_t ntaken = 0;
        for (int b0 = 0; b0 < nbk; b0 += 64) {
            const int b = b0 + (int)lane;
            bool takeA = false;
            uint32_t base = 0, vid = 0;
            if (b < nbk) {
                const QkBranchK B = brk[b];
                const uint32_t hm = hi & B.hcm;
                const uint32_t u0 = (hm == B.hcv) ? 1u : 0u;
                const uint32_t u1 = (((B.halt | B.lalt) != 0) && (hm == (B.hcv ^ B.halt))) ? 2u : 0u;
                const uint32_t hpar = (uint32_t)(__builtin_popcount(hi & B.hM) & 1);
                base = hoff[hi ^ B.hF];
                tb[b] = make_uint2(base, u0 | u1 | (hpar << 2));
                takeA = (b < nA) && (u0 | u1);
                vid = 2u * (uint32_t)b + hpar;
            }
            const uint64_t m = __ballot(takeA);
            if (takeA) {
                const uint32_t pos = ntaken + (uint32_t)__popcll(m & lt_mask);
                a_base[pos] = base;
                a_vid[pos] = vid;
            }
            ntaken += (uint32_t)__popcll(m);
        }
        double dre, dim;
        eval_diag<ST, false>(dg, sc.ndg, sc.dcr, sc.dci, s, dre, dim);
        l_diag[lane] = dre;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");

        // ---- phase A.2: entries per row
        uint32_t n_j = 1u + (use_prev ? 1u : 0u) + ntaken;
        for (int g = 0; g < nB; g += 8) {
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const uint4 L = lb4[g + u];
                const uint32_t fl = tb[nA + g + u].y;
                const uint32_t lf = lo & L.x;
                n_j += ((fl & 1u) & (uint32_t)(lf == L.y)) | ((fl >> 1) & 1u & (uint32_t)(lf == (L.y ^ L.z)));
            }
        }
        if (!active) n_j = 0;
        uint32_t incl = n_j;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t o = __shfl_up(incl, off, 64);
            if ((int)lane >= off) incl += o;
        }
        const uint32_t excl = incl - n_j;

        uint32_t j0 = wbeg;
        while (j0 < wend) {
            const uint32_t ebase = __builtin_amdgcn_readfirstlane(__shfl(excl, (int)j0, 64));
            const bool inr = (lane >= j0) && active && (incl - ebase <= (uint32_t)cap);
            const uint64_t bal = __ballot(inr);
            const uint32_t j1 = j0 + (uint32_t)__popcll(bal);
            const uint32_t total = __builtin_amdgcn_readfirstlane(__shfl(incl, (int)j1 - 1, 64)) - ebase;
            // ---- phase A.3: write the entries of rows [j0, j1)
            if (inr) {
                uint32_t k = excl - ebase;
                const uint32_t row = tl.row_start + lane;
                l_tgt[k] = row;
                l_vid[k] = (uint16_t)(kDiag + lane);
                ++k;
                if (use_prev) {
                    l_tgt[k] = row;
                    l_vid[k] = (uint16_t)kPrev;
                    ++k;
                }
                for (uint32_t i = 0; i < ntaken; ++i) {
                    l_tgt[k] = a_base[i] + my_lr;
                    l_vid[k] = (uint16_t)a_vid[i];
                    ++k;
                }
                const OwnRank orank = own_rank<RM>(smem, lo, sc);
                auto write_run = [&](auto kind_tag, int gbeg, int gend) {
                    constexpr int KIND = decltype(kind_tag)::value;
                    for (int g = gbeg; g < gend; g += 8) {
#pragma unroll
                        for (int u = 0; u < 8; ++u) {
                            const uint4 L = lb4[g + u];
                            const uint2 tv = tb[nA + g + u];
                            const uint32_t lf = lo & L.x;
                            const bool ok = ((tv.y & 1u) && lf == L.y) || (((tv.y >> 1) & 1u) && lf == (L.y ^ L.z));
                            if (ok) {
                                const uint32_t lr = target_rank<RM, KIND>(smem, lo, L.w, orank, sc);
                                const uint32_t neg =
                                    ((uint32_t)__builtin_popcount(lo & lbm[g + u]) ^ (tv.y >> 2)) & 1u;
                                l_tgt[k] = tv.x + lr;
                                l_vid[k] = (uint16_t)(2u * (uint32_t)(nA + g + u) + neg);
                                ++k;
                            }
                        }
                    }
                };
                if (RM == 2) {
                    write_run(std::integral_constant<int, 1>{}, 0, sc.nB1);
                    write_run(std::integral_constant<int, 2>{}, sc.nB1, sc.nB1 + sc.nB2);
                    write_run(std::integral_constant<int, 0>{}, sc.nB1 + sc.nB2, nB);
                } else {
                    write_run(std::integral_constant<int, 0>{}, 0, nB);
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");

            // ---- phase B: lane = column, flat walk over the list
            uint32_t cur = kNone;
            double acc = 0.0, xs = 0.0, pv = 0.0;
            auto finish = [&]() {
                const size_t r = (size_t)(tl.row_start + cur) * lds;
                if (LZ) {
                    double w = acc * inv_n;
                    if (use_prev) w = fma(-cprev, pv, w);
                    if (colok) __builtin_nontemporal_store(w, Yc + r);
                    part = fma(xs * inv_n, w, part);
                } else {
                    if (colok) __builtin_nontemporal_store(acc, Yc + r);
                }
            };
            constexpr int G = 128 / NW;   // gathers per group: 8, 16 or 32 with 16, 8 or 4 waves
            auto consume = [&](const double* xg, const uint32_t* vg) {
#pragma unroll
                for (int u = 0; u < G; ++u) {
                    const uint32_t v = vg[u];
                    if (v < kDiag) {
                        acc = fma(l_vals[v], xg[u], acc);
                    } else if (v < kPrev) {
                        if (cur != kNone) finish();
                        cur = v - kDiag;
                        xs = xg[u];
                        acc = l_diag[cur] * xs;
                    } else if (v == kPrev) {
                        pv = xg[u];
                    }
                }
            };
            const uint32_t tfill = __builtin_amdgcn_readfirstlane(l_tgt[0]);
            double xa[G], xb[G];
            uint32_t va[G], vb[G];
            block_gather<G>(l_tgt, l_vid, 0, total, tfill, Xc, Pc, lds, kPrev, kNop, xa, va);
            for (uint32_t e0 = 0; e0 < total; e0 += 2 * G) {
                block_gather<G>(l_tgt, l_vid, e0 + G, total, tfill, Xc, Pc, lds, kPrev, kNop, xb, vb);
                consume(xa, va);
                block_gather<G>(l_tgt, l_vid, e0 + 2 * G, total, tfill, Xc, Pc, lds, kPrev, kNop, xa, va);
                consume(xb, vb);
            }
            if (cur != kNone) finish();
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            j0 = j1;
        }
    }
    if (LZ) {
        // per-column partial of <q_k, w>, waves added in index order
        __syncthreads();
        l_diag[lane] = part;
        __syncthreads();
        if (wv == 0 && colok) {
            double s = 0.0;
            const int nw = (int)(blockDim.x >> 6);
            for (int i = 0; i < nw; ++i)
                s += reinterpret_cast<const double*>(smem + wave0 + (size_t)i * block_wave_bytes(nbk, nA, cap) +
                                                     (size_t)cap * 6)[lane];
            partials[(size_t)lane * npart + blockIdx.x] = s;
        }
    }
}

// ------------------------------------------------------------------ Lanczos update
// q_{k+1} = w - alpha_k q_k / nrm_k (in place on w); partial sums of |q_{k+1}|^2.
// Single vector: thread-strided. Block: lane = column within a 64-wide row.
template <bool CPLX>
__global__ __launch_bounds__(QK_TPB) void k_update1(size_t n, double* __restrict__ w, const double* __restrict__ q,
                                                     double* __restrict__ alpha, const double* __restrict__ nrm,
                                                     int k, double* __restrict__ partials,
                                                     const int* __restrict__ kdev, int koff,
                                                     const double* __restrict__ a_part, int n_a_part,
                                                     int* __restrict__ knext, double* __restrict__ alpha_host)
{
    __shared__ double wsum[QK_WPB];
    __shared__ double tot_sh;
    const int kk = kdev ? *kdev + koff : k;
    // with a_part, alpha[k] is totalled here from the apply's partial sums
    // (workgroup 0 stores it) instead of by a separate reduction launch
    double a;
    if (a_part) {
        a = block_total(a_part, n_a_part, &tot_sh);
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            alpha[kk] = a;
            if (alpha_host) alpha_host[kk] = a;   // the host's copy (coherent pinned memory)
            // a graph's last step of a turn advances the other step counter
            // (one this launch does not read) to the next turn
            if (knext) *knext = kk - koff + 3;
        }
    } else {
        a = alpha[kk];
    }
    const double c = a / nrm[kk];
    double part = 0.0;
    const size_t nel = CPLX ? 2 * n : n;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (((reinterpret_cast<uintptr_t>(w) | reinterpret_cast<uintptr_t>(q)) & 15u) == 0) {
        // 16 B per lane, then an odd last element
        const size_t npair = nel / 2;
        qk_double2* w2 = reinterpret_cast<qk_double2*>(w);
        const qk_double2* q2 = reinterpret_cast<const qk_double2*>(q);
        for (size_t i = tid; i < npair; i += stride) {
            const qk_double2 wv = w2[i], qv = q2[i];
            qk_double2 v;
            v.x = fma(-c, qv.x, wv.x);
            v.y = fma(-c, qv.y, wv.y);
            w2[i] = v;
            part = fma(v.x, v.x, part);
            part = fma(v.y, v.y, part);
        }
        if ((nel & 1) && tid == 0) {
            const double v = fma(-c, q[nel - 1], w[nel - 1]);
            w[nel - 1] = v;
            part = fma(v, v, part);
        }
    } else {
        for (size_t i = tid; i < nel; i += stride) {
            const double v = fma(-c, q[i], w[i]);
            w[i] = v;
            part = fma(v, v, part);
        }
    }
    part = wave_sum(part);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = part;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += wsum[i];
        partials[blockIdx.x] = s;
    }
}

__global__ __launch_bounds__(QK_TPB) void k_update_block(size_t nrows, int ncol, int ld, double* __restrict__ w,
                                                          const double* __restrict__ q,
                                                          const double* __restrict__ alpha,
                                                          const double* __restrict__ nrm, int k, int ld_sc,
                                                          double* __restrict__ partials)
{
    __shared__ double wsum[QK_WPB][64];
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    const bool colok = lane < (uint32_t)ncol;
    const int c = colok ? (int)lane : 0;
    const double coef = alpha[(size_t)k * ld_sc + c] / nrm[(size_t)k * ld_sc + c];
    double part = 0.0;
    if (colok) {
        for (size_t r = (size_t)blockIdx.x * (blockDim.x >> 6) + wv; r < nrows; r += (size_t)gridDim.x * (blockDim.x >> 6)) {
            const size_t i = r * ld + c;
            const double v = fma(-coef, q[i], w[i]);
            w[i] = v;
            part = fma(v, v, part);
        }
    }
    wsum[wv][lane] = part;
    __syncthreads();
    if (wv == 0 && colok) {
        double s = 0.0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += wsum[i][lane];
        partials[(size_t)lane * gridDim.x + blockIdx.x] = s;
    }
}

// ------------------------------------------------------------------ multi-column apply
// Y = H_c X for NV real columns stored row-major (X[row * NV + c]), where H_c
// is H with its diagonal scaled by dscale[c] (for instance the anisotropy of
// an XXZ chain whose exchange part the columns share). One lane per row as
// in k_apply: every decoded entry (target, value) serves all NV columns, and
// every gather reads the NV consecutive doubles of one target row. In the
// Lanczos form (LZ) each column c runs its own recurrence: scalars
// nrm[k * NV + c], previous block P, partial <q_k, w> of column c at
// partials[c * npart + workgroup]. A column's arithmetic (entries in branch
// order, the FMA order, the reduction tree) depends neither on NV nor on the
// other columns, and the launch grid depends only on the tables, so a column
// gets the same bits in a block of any width. Real spin tables (no parity
// masks), 32-bit states. In the Lanczos form the kernel totals nrm[k][c]
// itself from the previous update's partial sums (nrm_part[c * n_nrm_part
// + i], fixed order, workgroup 0 stores it), so a step needs no separate
// reduction launch.
template <int NV>
__device__ __forceinline__ void ld_cols(const double* __restrict__ X, size_t r, double* v)
{
    if (NV == 1) {
        v[0] = X[r];
    } else {
        const qk_double2* p = reinterpret_cast<const qk_double2*>(X + r * NV);
#pragma unroll
        for (int q = 0; q < NV / 2; ++q) {
            const qk_double2 t = p[q];
            v[2 * q] = t.x;
            v[2 * q + 1] = t.y;
        }
    }
}

template <int NV>
__device__ __forceinline__ void st_cols(double* __restrict__ Y, size_t r, const double* v, bool nt)
{
    if (NV == 1) {
        if (nt) __builtin_nontemporal_store(v[0], Y + r);
        else Y[r] = v[0];
    } else {
        qk_double2* p = reinterpret_cast<qk_double2*>(Y + r * NV);
#pragma unroll
        for (int q = 0; q < NV / 2; ++q) {
            qk_double2 t;
            t.x = v[2 * q];
            t.y = v[2 * q + 1];
            if (nt) __builtin_nontemporal_store(t, p + q);
            else p[q] = t;
        }
    }
}

// acc[c] += sum_u cr[u] * X[rt[u]][c]: all U row loads first (only for the
// lanes whose ok[u] is set; the others take 0.0 under a zero coefficient and
// issue no load), then the FMAs in u order for every column
template <int NV, int U>
__device__ __forceinline__ void accumulate_cols(const double* __restrict__ X, const uint32_t* rt, const bool* ok,
                                                const double* cr, double* acc)
{
    double t[U][NV];
#pragma unroll
    for (int u = 0; u < U; ++u) {
#pragma unroll
        for (int c = 0; c < NV; ++c) t[u][c] = 0.0;
        if (ok[u]) ld_cols<NV>(X, rt[u], t[u]);
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
        for (int c = 0; c < NV; ++c) acc[c] = fma(cr[u], t[u][c], acc[c]);
}

template <int RM, int KIND, int NV>
__device__ __forceinline__ void lane_branches_cols(int gbeg, int gend, const unsigned char* smem, const uint4* lb4,
                                                   const double* lbre, const uint4* tb, uint32_t lo, bool active,
                                                   const OwnRank& orank, const QkScal& sc,
                                                   const double* __restrict__ X, double* acc)
{
    constexpr int U = NV >= 4 ? 4 : 8;
    for (int g = gbeg; g < gend; g += U) {
        uint32_t rt[U];
        bool ok[U];
        double cr[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint4 L = lb4[g + u];   // lcm, lcv, lalt, lF
            const uint4 tv = tb[g + u];   // block offset, va, vb, hpar
            const uint32_t lf = lo & L.x;
            const uint32_t m = 0u - (uint32_t)((lf == tv.y) | (lf == tv.z));
            const uint32_t lr = target_rank<RM, KIND>(smem, lo, L.w, orank, sc);
            const uint64_t m64 = ((uint64_t)m << 32) | m;
            rt[u] = tv.x + lr;
            ok[u] = active && m != 0u;
            cr[u] = __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, lbre[g + u]) & m64);
        }
        accumulate_cols<NV, U>(X, rt, ok, cr, acc);
    }
}

template <int RM, bool LZ, int NV>
__global__ __launch_bounds__(512) void k_apply_cols(const QkTile* __restrict__ tiles, const uint32_t* __restrict__ hoff,
                                                     const uint32_t* __restrict__ lrank,
                                                     const uint16_t* __restrict__ tl,
                                                     const uint32_t* __restrict__ th, int nth,
                                                     const uint32_t* __restrict__ lolist,
                                                     const QkBranchK* __restrict__ brk,
                                                     const QkDiagTerm* __restrict__ dg, QkScal sc,
                                                     const double* __restrict__ X, double* __restrict__ Y,
                                                     const double* __restrict__ dscale, int lz_k,
                                                     const double* __restrict__ P, double* __restrict__ nrm,
                                                     double* __restrict__ partials, int npart, int table_bytes,
                                                     const double* __restrict__ nrm_part, int n_nrm_part)
{
    extern __shared__ __align__(16) unsigned char smem[];
    // static LDS: QK_COLS_STATIC_LDS bytes
    __shared__ double wsum[8][NV];
    __shared__ double tot_sh[NV];
    const int nA = sc.nA, nB = sc.nB, nbk = sc.nA + sc.nB;
    uint4* lb4 = reinterpret_cast<uint4*>(smem + table_bytes);
    double* lbre = reinterpret_cast<double*>(smem + table_bytes + ((size_t)nB * 20 + 7) / 8 * 8);
    for (int i = threadIdx.x; i < nB; i += blockDim.x) {
        const QkBranchK B = brk[nA + i];
        lb4[i] = make_uint4(B.lcm, B.lcv, B.lalt, B.lF);
        lbre[i] = B.re;
    }
    load_lrank<RM>(smem, lrank, tl, th, sc.nlo, sc.ntl, nth);

    const uint32_t lane = threadIdx.x & 63u;
    const int NL = sc.NL;
    unsigned char* wb = smem + table_bytes + apply_branch_bytes(nB) + (threadIdx.x >> 6) * apply_wave_bytes(nB, nA, false);
    uint4* tb = reinterpret_cast<uint4*>(wb);                          // [nB] {block offset, va, vb, 0}
    double* a_cr = reinterpret_cast<double*>(wb + (size_t)nB * 16);    // [nA] taken hi-only values
    uint32_t* a_base = reinterpret_cast<uint32_t*>(a_cr + nA);         // [nA] their block offsets
    const uint64_t lt_mask = (lane == 0) ? 0ULL : (~0ULL >> (64 - lane));

    double dsc[NV], inv_n[NV], cprev[NV], part[NV];
#pragma unroll
    for (int c = 0; c < NV; ++c) {
        dsc[c] = dscale[c];
        inv_n[c] = 1.0;
        cprev[c] = 0.0;
        part[c] = 0.0;
        if (LZ) {
            const double nk = sqrt(block_total(nrm_part + (size_t)c * n_nrm_part, n_nrm_part, &tot_sh[c]));
            if (blockIdx.x == 0 && threadIdx.x == 0) nrm[(size_t)lz_k * NV + c] = nk;
            inv_n[c] = 1.0 / nk;
            if (lz_k > 0) cprev[c] = nk / nrm[(size_t)(lz_k - 1) * NV + c];
        }
    }

    uint32_t t0, tend, step;
    tile_schedule(sc.ntiles, sc.sched, t0, tend, step);
    for (uint32_t t = t0; t < tend; t += step) {
        const QkTile tl = tiles[t];
        const uint32_t hi = tl.hi;
        const bool active = lane < tl.count;
        const uint32_t base_self = hoff[hi];
        const uint32_t lidx = active ? lane : 0u;
        const uint32_t my_lr = tl.row_start - base_self + lidx;
        const uint32_t row = tl.row_start + lidx;
        const uint32_t lo = lolist[tl.lo_base + lidx];
        double pv[NV];
        if (LZ && lz_k > 0) ld_cols<NV>(P, row, pv);
        const uint32_t s = (hi << NL) | lo;

        // (1) the uniform part of every branch, one branch per lane
        uint32_t ntaken = 0;
        for (int b0 = 0; b0 < nbk; b0 += 64) {
            const int b = b0 + (int)lane;
            bool takeA = false;
            uint32_t base = 0;
            double vr = 0.0;
            if (b < nbk) {
                const QkBranchK B = brk[b];
                const uint32_t hm = hi & B.hcm;
                const uint32_t u0 = (hm == B.hcv) ? 1u : 0u;
                const uint32_t u1 = (((B.halt | B.lalt) != 0) && (hm == (B.hcv ^ B.halt))) ? 2u : 0u;
                const uint32_t never = ~B.lcm;
                base = hoff[hi ^ B.hF];
                if (b >= nA) tb[b - nA] = make_uint4(base, u0 ? B.lcv : never, u1 ? (B.lcv ^ B.lalt) : never, 0u);
                takeA = (b < nA) && (u0 | u1);
                vr = B.re;
            }
            // (2) compact the taken hi-only branches
            const uint64_t m = __ballot(takeA);
            if (takeA) {
                const uint32_t pos = ntaken + (uint32_t)__popcll(m & lt_mask);
                a_base[pos] = base;
                a_cr[pos] = vr;
            }
            ntaken += (uint32_t)__popcll(m);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");

        double dre, dim;
        eval_diag<uint32_t, false>(dg, sc.ndg, sc.dcr, sc.dci, s, dre, dim);
        double xv[NV], acc[NV];
        ld_cols<NV>(X, row, xv);
#pragma unroll
        for (int c = 0; c < NV; ++c) acc[c] = (dre * dsc[c]) * xv[c];

        // (3) taken hi-only branches, four in flight
        for (uint32_t i = 0; i < ntaken; i += 4) {
            uint32_t rt[4];
            bool ok[4];
            double cr[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const bool in = (i + u) < ntaken;
                const uint32_t k = in ? i + u : 0u;
                rt[u] = in ? a_base[k] + my_lr : row;
                ok[u] = in && active;
                cr[u] = in ? a_cr[k] : 0.0;
            }
            accumulate_cols<NV, 4>(X, rt, ok, cr, acc);
        }

        // (4) per-lane branches
        const OwnRank orank = own_rank<RM>(smem, lo, sc);
        if (RM == 2) {
            lane_branches_cols<RM, 1, NV>(0, sc.nB1, smem, lb4, lbre, tb, lo, active, orank, sc, X, acc);
            lane_branches_cols<RM, 2, NV>(sc.nB1, sc.nB1 + sc.nB2, smem, lb4, lbre, tb, lo, active, orank, sc, X,
                                          acc);
            lane_branches_cols<RM, 0, NV>(sc.nB1 + sc.nB2, nB, smem, lb4, lbre, tb, lo, active, orank, sc, X, acc);
        } else {
            lane_branches_cols<RM, 0, NV>(0, nB, smem, lb4, lbre, tb, lo, active, orank, sc, X, acc);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");

        if (active) {
            if (LZ) {
#pragma unroll
                for (int c = 0; c < NV; ++c) {
                    double w = acc[c] * inv_n[c];
                    if (lz_k > 0) w = fma(-cprev[c], pv[c], w);
                    acc[c] = w;
                    part[c] = fma(xv[c] * inv_n[c], w, part[c]);
                }
            }
            st_cols<NV>(Y, row, acc, true);
        }
    }
    if (LZ) {
        const uint32_t wv = threadIdx.x >> 6;
#pragma unroll
        for (int c = 0; c < NV; ++c) {
            const double v = wave_sum(part[c]);
            if (lane == 0) wsum[wv][c] = v;
        }
        __syncthreads();
        if (threadIdx.x < (unsigned)NV) {
            double s = 0.0;
            for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += wsum[i][threadIdx.x];
            partials[(size_t)threadIdx.x * npart + blockIdx.x] = s;
        }
    }
}

// Lanczos update of NV row-major columns, each with its own coefficient:
// w[r][c] -= (alpha[k][c] / nrm[k][c]) q[r][c]; partials[c * grid + b] of
// |w_c|^2. With a_part, alpha[k][c] is first totalled from the apply's
// partial sums (a_part[c * n_a_part + i], fixed order; workgroup 0 stores
// it); without, it is read. Rows are thread-strided over a grid that
// depends only on nrows, so a column's partial sums do not depend on NV.
template <int NV>
__global__ __launch_bounds__(QK_TPB) void k_update_cols(size_t nrows, double* __restrict__ w,
                                                         const double* __restrict__ q, double* __restrict__ alpha,
                                                         const double* __restrict__ nrm, int k,
                                                         const double* __restrict__ a_part, int n_a_part,
                                                         double* __restrict__ partials)
{
    __shared__ double wsum[QK_WPB][NV];
    __shared__ double tot_sh[NV];
    double coef[NV], part[NV];
#pragma unroll
    for (int c = 0; c < NV; ++c) {
        double a;
        if (a_part) {
            a = block_total(a_part + (size_t)c * n_a_part, n_a_part, &tot_sh[c]);
            if (blockIdx.x == 0 && threadIdx.x == 0) alpha[(size_t)k * NV + c] = a;
        } else {
            a = alpha[(size_t)k * NV + c];
        }
        coef[c] = a / nrm[(size_t)k * NV + c];
        part[c] = 0.0;
    }
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrows; r += stride) {
        double wv[NV], qv[NV];
        ld_cols<NV>(w, r, wv);
        ld_cols<NV>(q, r, qv);
#pragma unroll
        for (int c = 0; c < NV; ++c) {
            wv[c] = fma(-coef[c], qv[c], wv[c]);
            part[c] = fma(wv[c], wv[c], part[c]);
        }
        st_cols<NV>(w, r, wv, false);
    }
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < NV; ++c) {
        const double v = wave_sum(part[c]);
        if (lane == 0) wsum[wv][c] = v;
    }
    __syncthreads();
    if (threadIdx.x < (unsigned)NV) {
        double s = 0.0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += wsum[i][threadIdx.x];
        partials[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = s;
    }
}

// Y[r][c] += coef[c] X[r][c] for the columns whose coefficient is not zero
// (a zero coefficient leaves the column untouched, whatever X holds)
template <int NV>
__global__ __launch_bounds__(QK_TPB) void k_axpy_cols(size_t nrows, const double* __restrict__ coef,
                                                       const double* __restrict__ X, double* __restrict__ Y)
{
    double a[NV];
#pragma unroll
    for (int c = 0; c < NV; ++c) a[c] = coef[c];
    for (size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrows; r += (size_t)gridDim.x * blockDim.x) {
        double xv[NV], yv[NV];
        ld_cols<NV>(X, r, xv);
        ld_cols<NV>(Y, r, yv);
#pragma unroll
        for (int c = 0; c < NV; ++c) yv[c] = (a[c] != 0.0) ? fma(a[c], xv[c], yv[c]) : yv[c];
        st_cols<NV>(Y, r, yv, false);
    }
}

// Y[r][c] = d[r] X[r][c] (a diagonal operator applied to every column)
template <int NV>
__global__ __launch_bounds__(QK_TPB) void k_scale_rows_cols(size_t nrows, const double* __restrict__ d,
                                                             const double* __restrict__ X, double* __restrict__ Y)
{
    for (size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrows; r += (size_t)gridDim.x * blockDim.x) {
        double xv[NV];
        ld_cols<NV>(X, r, xv);
        const double dr = d[r];
#pragma unroll
        for (int c = 0; c < NV; ++c) xv[c] = dr * xv[c];
        st_cols<NV>(Y, r, xv, false);
    }
}

// ------------------------------------------------------------------ reductions
// One workgroup per result group. Fixed order: thread-strided sums, then a
// fixed tree, then columns of the group added in index order.
__global__ __launch_bounds__(QK_TPB) void k_reduce(const double* __restrict__ partials, int npart, int group,
                                                    int mode, int cplx, double* __restrict__ out, int idx, int ld_sc)
{
    __shared__ double red[QK_TPB];
    __shared__ double red_i[QK_TPB];
    const int g = blockIdx.x;
    double tot = 0.0, tot_i = 0.0;
    for (int cc = 0; cc < group; ++cc) {
        const int c = g * group + cc;
        const double* p = partials + (size_t)c * npart * (cplx ? 2 : 1);
        double s = 0.0, si = 0.0;
        for (int i = threadIdx.x; i < npart; i += blockDim.x) {
            s += p[i];
            if (cplx) si += p[npart + i];
        }
        red[threadIdx.x] = s;
        red_i[threadIdx.x] = si;
        __syncthreads();
        for (int off = blockDim.x / 2; off > 0; off >>= 1) {
            if ((int)threadIdx.x < off) {
                red[threadIdx.x] += red[threadIdx.x + off];
                red_i[threadIdx.x] += red_i[threadIdx.x + off];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            tot += red[0];
            tot_i += red_i[0];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        for (int cc = 0; cc < group; ++cc) {
            const int c = g * group + cc;
            double* o = out + (size_t)idx * ld_sc + c * (mode == 2 ? 2 : 1);
            if (mode == 0) o[0] = tot;
            else if (mode == 1) o[0] = sqrt(tot);
            else {
                o[0] = tot;
                o[1] = tot_i;
            }
        }
    }
}

// ------------------------------------------------------------------ BLAS-1
template <bool CPLX>
__global__ __launch_bounds__(QK_TPB) void k_dot(size_t n, const double* __restrict__ x, const double* __restrict__ y,
                                                 double* __restrict__ partials)
{
    __shared__ double wr[QK_WPB], wi[QK_WPB];
    double sr = 0.0, si = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        if (CPLX) {
            const double xr = x[2 * i], xi = x[2 * i + 1], yr = y[2 * i], yi = y[2 * i + 1];
            sr = fma(xr, yr, sr);
            sr = fma(xi, yi, sr);
            si = fma(xr, yi, si);
            si = fma(-xi, yr, si);
        } else {
            sr = fma(x[i], y[i], sr);
        }
    }
    sr = wave_sum(sr);
    if (CPLX) si = wave_sum(si);
    if ((threadIdx.x & 63) == 0) {
        wr[threadIdx.x >> 6] = sr;
        wi[threadIdx.x >> 6] = si;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) {
            a += wr[i];
            b += wi[i];
        }
        partials[blockIdx.x] = a;
        if (CPLX) partials[gridDim.x + blockIdx.x] = b;
    }
}

template <bool CPLX>
__global__ __launch_bounds__(QK_TPB) void k_axpy(size_t n, double ar, double ai, const double* __restrict__ x,
                                                  double* __restrict__ y)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        if (CPLX) {
            const double xr = x[2 * i], xi = x[2 * i + 1];
            y[2 * i] = fma(ar, xr, fma(-ai, xi, y[2 * i]));
            y[2 * i + 1] = fma(ar, xi, fma(ai, xr, y[2 * i + 1]));
        } else {
            y[i] = fma(ar, x[i], y[i]);
        }
    }
}

template <bool CPLX>
__global__ __launch_bounds__(QK_TPB) void k_scal(size_t n, double ar, double ai, double* __restrict__ x)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        if (CPLX) {
            const double xr = x[2 * i], xi = x[2 * i + 1];
            x[2 * i] = ar * xr - ai * xi;
            x[2 * i + 1] = ar * xi + ai * xr;
        } else {
            x[i] *= ar;
        }
    }
}

__global__ __launch_bounds__(QK_TPB) void k_copy_scaled(size_t n, const double* __restrict__ src,
                                                         double* __restrict__ dst, const double* __restrict__ sc,
                                                         int idx)
{
    const double inv = 1.0 / sc[idx];
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i] * inv;
}

// 16 B per lane with streaming (nontemporal) loads and stores: the fastest
// of the plain copy shapes measured on MI355X (about 5.7 TB/s for 2 GiB)
__global__ __launch_bounds__(QK_TPB) void k_copy(size_t n, const double* __restrict__ src, double* __restrict__ dst)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    if (((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) & 15u) != 0) {
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
        return;
    }
    const size_t n2 = n / 2;
    const qk_double2* s2 = reinterpret_cast<const qk_double2*>(src);
    qk_double2* d2 = reinterpret_cast<qk_double2*>(dst);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += stride)
        __builtin_nontemporal_store(__builtin_nontemporal_load(s2 + i), d2 + i);
    if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) dst[n - 1] = src[n - 1];
}

// dst[r * dst_ld + j] = src[r * src_ld + j] for r < nrows, j < width (packs
// vectors into columns of a row-major block and back)
__global__ __launch_bounds__(QK_TPB) void k_copy_strided(size_t nrows, int width, const double* __restrict__ src,
                                                          size_t src_ld, double* __restrict__ dst, size_t dst_ld)
{
    const size_t n = nrows * (size_t)width;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const size_t r = i / (size_t)width, j = i - r * (size_t)width;
        dst[r * dst_ld + j] = src[r * src_ld + j];
    }
}

// out[idx] = total of n partial sums (sqrt of it when root), in the order
// of block_total
__global__ __launch_bounds__(64) void k_total(const double* __restrict__ p, int n, int root, double* __restrict__ out,
                                              int idx)
{
    __shared__ double tot_sh;
    const double t = block_total(p, n, &tot_sh);
    if (threadIdx.x == 0) out[idx] = root ? sqrt(t) : t;
}

// out[idx * ncol + c] = sqrt(total of p[c * n + i]) for column c = blockIdx.x,
// in the order of block_total (what the column apply's prologue computes)
__global__ __launch_bounds__(64) void k_total_cols(const double* __restrict__ p, int n, double* __restrict__ out,
                                                   int idx, int ncol)
{
    __shared__ double tot_sh;
    const int c = blockIdx.x;
    const double t = block_total(p + (size_t)c * n, n, &tot_sh);
    if (threadIdx.x == 0) out[(size_t)idx * ncol + c] = sqrt(t);
}

// up to eight doubles handed over as kernel arguments
struct QkVals8 {
    double v[8];
};

__global__ void k_set_vals(double* __restrict__ dst, int n, QkVals8 vals)
{
    const int i = (int)threadIdx.x;
    if (i < n) dst[i] = vals.v[i];
}

__global__ void k_set_int(int* __restrict__ p, int v)
{
    if (threadIdx.x == 0) *p = v;
}

__global__ __launch_bounds__(QK_TPB) void k_fill_zero(size_t n, double* __restrict__ x)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        x[i] = 0.0;
}

// splitmix64 finaliser (Steele, Lea, Flood 2014) over a (seed, counter) mix
__device__ __forceinline__ uint64_t rng_u64(uint64_t seed, uint64_t counter)
{
    uint64_t z = seed * 0xD1B54A32D192ED03ULL + (counter + 1ULL) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(QK_TPB) void k_fill_random(int kind, size_t nrows, int ncol, int ld, int c0,
                                                         uint64_t seed, uint64_t counter0, double* __restrict__ x)
{
    const size_t nel = nrows * (size_t)ncol;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nel; i += (size_t)gridDim.x * blockDim.x) {
        const size_t row = i / (size_t)ncol;
        const int c = (int)(i - row * (size_t)ncol);
        const uint64_t z = rng_u64(seed, counter0 + (uint64_t)(c0 + c) * nrows + row);
        double v;
        if (kind == 0) v = (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
        else v = (z >> 63) ? -1.0 : 1.0;
        x[row * (size_t)ld + c0 + c] = v;
    }
}

// h partials for classical Gram-Schmidt: partials[j * nb + b] = sum_i V_j[i] u[i]
__global__ __launch_bounds__(QK_TPB) void k_multidot(size_t n, int m, const double* __restrict__ V, size_t ldv,
                                                      const double* __restrict__ u, double* __restrict__ partials)
{
    __shared__ double wsum[QK_WPB];
    const size_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const size_t i0 = (size_t)blockIdx.x * chunk, i1 = min(n, i0 + chunk);
    for (int j = 0; j < m; ++j) {
        const double* v = V + (size_t)j * ldv;
        double s = 0.0;
        for (size_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) s = fma(v[i], u[i], s);
        s = wave_sum(s);
        if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            double a = 0.0;
            for (int w = 0; w < (int)(blockDim.x >> 6); ++w) a += wsum[w];
            partials[(size_t)j * gridDim.x + blockIdx.x] = a;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(QK_TPB) void k_multiaxpy(size_t n, int m, const double* __restrict__ V, size_t ldv,
                                                       const double* __restrict__ coef, double sgn,
                                                       double* __restrict__ u)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        double a = u[i];
        for (int j = 0; j < m; ++j) a = fma(sgn * coef[j], V[(size_t)j * ldv + i], a);
        u[i] = a;
    }
}

// ------------------------------------------------------------------ row helpers
__device__ __forceinline__ uint32_t find_tile(const QkTile* __restrict__ tiles, uint32_t ntiles, uint32_t row)
{
    uint32_t lo = 0, hi = ntiles;   // last tile with row_start <= row
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (tiles[mid].row_start <= row) lo = mid; else hi = mid;
    }
    return lo;
}

template <typename ST>
__global__ __launch_bounds__(QK_TPB) void k_diagonal(const QkTile* __restrict__ tiles,
                                                      const uint32_t* __restrict__ lolist,
                                                      const QkDiagTerm* __restrict__ dg, QkScal sc, int cplx,
                                                      size_t dim, double* __restrict__ d)
{
    for (size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x; r < dim; r += (size_t)gridDim.x * blockDim.x) {
        const QkTile tl = tiles[find_tile(tiles, sc.ntiles, (uint32_t)r)];
        const uint32_t lo = lolist[tl.lo_base + ((uint32_t)r - tl.row_start)];
        const ST s = (static_cast<ST>(tl.hi) << sc.NL) | static_cast<ST>(lo);
        double dre, dim_;
        eval_diag<ST, true>(dg, sc.ndg, sc.dcr, sc.dci, s, dre, dim_);
        if (cplx) {
            d[2 * r] = dre;
            d[2 * r + 1] = dim_;
        } else {
            d[r] = dre;
        }
    }
}

template <typename ST>
__global__ __launch_bounds__(QK_TPB) void k_states(const QkTile* __restrict__ tiles,
                                                    const uint32_t* __restrict__ lolist, QkScal sc, size_t r0,
                                                    size_t n, uint64_t* __restrict__ out)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const uint32_t r = (uint32_t)(r0 + i);
        const QkTile tl = tiles[find_tile(tiles, sc.ntiles, r)];
        const uint32_t lo = lolist[tl.lo_base + (r - tl.row_start)];
        out[i] = (static_cast<uint64_t>(tl.hi) << sc.NL) | lo;
    }
}

template <typename ST>
__global__ __launch_bounds__(QK_TPB) void k_row_entries(const QkTile* __restrict__ tiles,
                                                         const uint32_t* __restrict__ hoff,
                                                         const uint32_t* __restrict__ lrank,
                                                         const uint32_t* __restrict__ lolist,
                                                         const QkBranch* __restrict__ br,
                                                         const QkDiagTerm* __restrict__ dg, QkScal sc, size_t r0,
                                                         size_t n, int cap, int64_t* __restrict__ cols,
                                                         double* __restrict__ vals, int* __restrict__ counts)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const uint32_t r = (uint32_t)(r0 + i);
        const QkTile tl = tiles[find_tile(tiles, sc.ntiles, r)];
        const uint32_t hi = tl.hi;
        const uint32_t lo = lolist[tl.lo_base + (r - tl.row_start)];
        const ST s = (static_cast<ST>(hi) << sc.NL) | static_cast<ST>(lo);
        const uint32_t my_lr = r - hoff[hi];
        int k = 0;
        double dre, dim_;
        eval_diag<ST, true>(dg, sc.ndg, sc.dcr, sc.dci, s, dre, dim_);
        if (dre != 0.0 || dim_ != 0.0) {
            if (k < cap) {
                cols[i * cap + k] = r;
                vals[2 * (i * cap + k)] = dre;
                vals[2 * (i * cap + k) + 1] = dim_;
            }
            ++k;
        }
        for (int b = 0; b < sc.nbr; ++b) {
            const QkBranch B = br[b];
            const uint64_t sm = (uint64_t)s & B.cm;
            const bool okc = (sm == B.cv) || (B.alt != 0 && sm == (B.cv ^ B.alt));
            if (!okc) continue;
            const uint64_t t = (uint64_t)s ^ B.F;
            const uint32_t thi = (uint32_t)(t >> sc.NL), tlo = (uint32_t)t & sc.lomask;
            const uint32_t lr = (B.F & sc.lomask) ? lrank[tlo] : my_lr;
            const bool neg = (__builtin_popcountll((uint64_t)s & B.M) & 1) != 0;
            if (k < cap) {
                cols[i * cap + k] = (int64_t)(hoff[thi] + lr);
                vals[2 * (i * cap + k)] = neg ? -B.re : B.re;
                vals[2 * (i * cap + k) + 1] = neg ? -B.im : B.im;
            }
            ++k;
        }
        counts[i] = k;
    }
}

template <typename ST>
__global__ __launch_bounds__(QK_TPB) void k_apply_diag_terms(const QkTile* __restrict__ tiles,
                                                              const uint32_t* __restrict__ lolist,
                                                              const QkDiagTerm* __restrict__ dg, QkScal sc,
                                                              int xcplx, size_t dim, const double* __restrict__ x,
                                                              double* __restrict__ y)
{
    for (size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x; r < dim; r += (size_t)gridDim.x * blockDim.x) {
        const QkTile tl = tiles[find_tile(tiles, sc.ntiles, (uint32_t)r)];
        const uint32_t lo = lolist[tl.lo_base + ((uint32_t)r - tl.row_start)];
        const ST s = (static_cast<ST>(tl.hi) << sc.NL) | static_cast<ST>(lo);
        double fr, fi;
        eval_diag<ST, true>(dg, sc.ndg, sc.dcr, sc.dci, s, fr, fi);
        // stored diagonal values are conj(amp); O x needs amp = conj(stored)
        fi = -fi;
        const double xr = xcplx ? x[2 * r] : x[r];
        const double xi = xcplx ? x[2 * r + 1] : 0.0;
        y[2 * r] = fr * xr - fi * xi;
        y[2 * r + 1] = fr * xi + fi * xr;
    }
}

template <bool CPLX>
__global__ __launch_bounds__(QK_TPB) void k_precond(size_t n, const double* __restrict__ r,
                                                     const double* __restrict__ dg, double lambda,
                                                     double* __restrict__ t)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        if (CPLX) {
            double dr = dg[2 * i] - lambda, di = dg[2 * i + 1];
            if (sqrt(dr * dr + di * di) < 1e-10) {
                dr = (dr >= 0) ? 1e-10 : -1e-10;
                di = 0.0;
            }
            const double den = dr * dr + di * di;
            const double rr = r[2 * i], ri = r[2 * i + 1];
            t[2 * i] = (rr * dr + ri * di) / den;
            t[2 * i + 1] = (ri * dr - rr * di) / den;
        } else {
            double d = dg[i] - lambda;
            if (fabs(d) < 1e-10) d = (d >= 0) ? 1e-10 : -1e-10;
            t[i] = r[i] / d;
        }
    }
}

// ------------------------------------------------------------------ basis permutation
// Rows of a relabelled spin basis <-> rows of the public basis. The internal
// state (hi << NL) | lo of row `row` maps to the public state bit by bit
// through four byte tables (bytemap[256 k + v]: the public bits of the
// internal bits 8k..8k+7 holding v); its public row is
// hoff_p[s >> NLp] + lrank_p[s & (2^NLp - 1)]. TO_PUB: dst[public row] =
// src[row], else dst[row] = src[public row]. One wave per tile.
template <bool TO_PUB>
__global__ __launch_bounds__(QK_TPB) void k_permute(const QkTile* __restrict__ tiles, uint32_t ntiles,
                                                     const uint32_t* __restrict__ lolist, int NL,
                                                     const uint32_t* __restrict__ bytemap,
                                                     const uint32_t* __restrict__ hoff_p,
                                                     const uint32_t* __restrict__ lrank_p, int NLp,
                                                     const double* __restrict__ src, double* __restrict__ dst)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nw = (gridDim.x * blockDim.x) >> 6;
    const uint32_t lmask = (1u << NLp) - 1u;
    for (uint32_t t = wave; t < ntiles; t += nw) {
        const QkTile tl = tiles[t];
        if (lane >= tl.count) continue;
        const uint32_t row = tl.row_start + lane;
        const uint32_t si = (tl.hi << NL) | lolist[tl.lo_base + lane];
        const uint32_t sp = bytemap[si & 255u] | bytemap[256u + ((si >> 8) & 255u)] |
                            bytemap[512u + ((si >> 16) & 255u)] | bytemap[768u + (si >> 24)];
        const uint32_t r = hoff_p[sp >> NLp] + lrank_p[sp & lmask];
        if (TO_PUB) dst[r] = src[row];
        else dst[row] = src[r];
    }
}

// ------------------------------------------------------------------ static workspace
// The workspace of the entry points that must not allocate (for callers
// under graph capture or in loops that forbid hipMalloc): one array per GPU,
// alive until the process ends. Host code never uses the symbol's name as
// an address; qkl_scratch looks up the device address once per device.
#define QK_SCRATCH_BYTES ((size_t)64 << 20)
static __device__ __align__(256) double scratch[QK_SCRATCH_BYTES / sizeof(double)];

// ------------------------------------------------------------------ launch helpers
// LDS bytes of the lo-rank table, rounded to 16 so what follows it is aligned
inline size_t table_bytes(const QkTables* T)
{
    if (T->rank_mode == 2) return tl_bytes(T->ntl) + (((size_t)T->nth * 4 + 15) & ~(size_t)15);
    return ((size_t)T->nlo * (T->rank_mode == 1 ? 2 : 4) + 15) & ~(size_t)15;
}

// Workgroups of kernel f (64 * wpb threads, lds dynamic LDS bytes) that one
// CU holds at once, by the runtime's occupancy calculator (registers, LDS,
// wave slots), and the number of CUs; cached per kernel, shape and device.
inline void residency(const void* f, int wpb, size_t lds, int& per_cu, int& ncu)
{
    int dev = 0;
    (void)hipGetDevice(&dev);
    static std::mutex mu;
    static std::map<std::tuple<const void*, int, size_t, int>, std::pair<int, int>> cache;
    const auto key = std::make_tuple(f, wpb, lds, dev);
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find(key);
    if (it == cache.end()) {
        // on a failed query: one workgroup per CU and the MI355X's 256 CUs
        // (the grid size changes only the speed); the query's own error is
        // cleared so that the launch after it does not report it
        int n = 0, cus = 0;
        bool failed = false;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, f, 64 * wpb, lds) != hipSuccess) failed = true;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) failed = true;
        if (failed) (void)hipGetLastError();
        if (n < 1) n = 1;
        if (cus < 1) cus = 256;
        it = cache.emplace(key, std::make_pair(n, cus)).first;
    }
    per_cu = it->second.first;
    ncu = it->second.second;
}

// Single-vector apply launch shape: the preferred number of waves per
// workgroup (4, 8 or 16; they share one LDS copy of the tables), or fewer
// when the LDS of that many waves exceeds 160 KB. The grid is a whole
// number of full residencies of the chip (so no partial last round of
// workgroups leaves a tail) of at most 8192 waves, a multiple of 8
// workgroups (XCD segments). f is the kernel instance to be launched.
struct ApplyCfg {
    int wpb, nb;
    size_t lds;
};

// LDS of a gfx950 workgroup, and the static LDS k_apply declares beside its
// dynamic tables (wsum[16] and tot_sh)
#define QK_LDS_LIMIT ((size_t)160 * 1024)
#define QK_APPLY_STATIC_LDS ((size_t)17 * sizeof(double))

// static LDS of k_apply_cols (wsum[8][NV] and tot_sh[NV] at the widest NV = 8)
#define QK_COLS_STATIC_LDS ((size_t)9 * 8 * sizeof(double))

inline int apply_lds_plan(size_t tables, int nB, int nA, bool cplx, int pref_waves, size_t static_lds, int* waves,
                          size_t* dyn_lds)
{
    for (int wpb = pref_waves; wpb >= 1; wpb /= 2) {
        const size_t lds = tables + apply_branch_bytes(nB) + (size_t)wpb * apply_wave_bytes(nB, nA, cplx);
        if (lds + static_lds > QK_LDS_LIMIT) continue;
        *waves = wpb;
        *dyn_lds = lds;
        return 0;
    }
    return -1;
}

extern "C" int qkl_apply_lds_plan(size_t tables, int nB, int nA, int cplx, int pref_waves, int* waves,
                                  size_t* dyn_lds)
{
    return apply_lds_plan(tables, nB, nA, cplx != 0, pref_waves, QK_APPLY_STATIC_LDS, waves, dyn_lds);
}

// Multi-column apply launch shape: the single-vector apply's preferred
// waves per workgroup up to 8 (the kernel's bound, which leaves the widest
// instance room for its registers), within LDS, and a grid of at most 4096
// waves; both depend only on the tables (never on the number of columns).
// The grid is a multiple of 8 workgroups.
inline bool cols_config(const QkTables* T, ApplyCfg& c)
{
    int wpb = (T->apply_waves == 8 || T->apply_waves == 16) ? 8 : 4;
    size_t lds = 0;
    if (apply_lds_plan(table_bytes(T), T->nB, T->nA, false, wpb, QK_COLS_STATIC_LDS, &wpb, &lds) != 0) return false;
    uint32_t nb = (T->ntiles + wpb - 1) / wpb;
    const uint32_t maxb = 4096u / (uint32_t)wpb;
    if (nb > maxb) nb = maxb;
    if (nb >= 8) nb = (nb + 7u) & ~7u;
    if (nb == 0) nb = 1;
    c.wpb = wpb;
    c.nb = (int)nb;
    c.lds = lds;
    return true;
}

inline bool apply_config(const QkTables* T, bool cplx, const void* f, ApplyCfg& c)
{
    int wpb = (T->apply_waves == 8 || T->apply_waves == 16) ? T->apply_waves : 4;
    size_t lds = 0;
    if (qkl_apply_lds_plan(table_bytes(T), T->nB, T->nA, cplx ? 1 : 0, wpb, &wpb, &lds) != 0) return false;
    int per_cu = 1, ncu = 256;
    residency(f, wpb, lds, per_cu, ncu);
    const uint32_t resident = (uint32_t)per_cu * (uint32_t)ncu;
    const uint32_t rounds = std::max(1u, (8192u / (uint32_t)wpb) / resident);
    uint32_t nb = (T->ntiles + wpb - 1) / wpb;
    const uint32_t maxb = resident * rounds;
    if (nb > maxb) nb = maxb;
    if (nb >= 8) nb = (nb + 7u) & ~7u;
    if (nb == 0) nb = 1;
    c.wpb = wpb;
    c.nb = (int)nb;
    c.lds = lds;
    return true;
}

// Block-apply launch shape: one workgroup per CU (up to 150 KB of LDS) of
// 16, 8 or 4 waves, walking 8, 16 or 32 gathers per group, so every shape
// keeps about 128 gathers per CU in flight while fewer waves keep fewer
// tiles (and their rows) live in each XCD's L2. The preferred width is
// tried first, then the others, until the entry list of a wave holds at
// least one row's entries.
struct BlockCfg {
    int nw, cap, nb;
    size_t lds;
};

// work items per tile of the block apply, as a shift: 64 rows 0, 32 rows 1, 16 rows 2
inline int block_tsh(const QkTables* T) { return T->blk_rows == 16 ? 2 : (T->blk_rows == 32 ? 1 : 0); }

inline bool block_config(const QkTables* T, BlockCfg& c)
{
    const int nbk = T->nA + T->nB;
    const size_t shared = table_bytes(T) + block_branch_bytes(T->nB) + (size_t)nbk * 16;
    const size_t budget = (size_t)150 * 1024;
    const int pref = (T->blk_waves == 4 || T->blk_waves == 8) ? T->blk_waves : 16;
    const int widths[3] = {pref, pref == 16 ? 8 : 16, pref == 4 ? 8 : 4};
    for (int w = 0; w < 3; ++w) {
        const int nw = widths[w];
        if (shared >= budget) break;
        const size_t per = (budget - shared) / nw;
        long cap = ((long)per - (long)block_wave_bytes(nbk, T->nA, 0) - 32) / 6;
        if (cap > 8192) cap = 8192;
        cap &= ~31L;
        if (cap < nbk + 2) continue;
        c.nw = nw;
        c.cap = (int)cap;
        c.lds = shared + (size_t)nw * block_wave_bytes(nbk, T->nA, (int)cap);
        uint32_t nb = ((T->ntiles << block_tsh(T)) + nw - 1) / nw;
        if (nb > 256u) nb = 256u;
        if (nb >= 8) nb = (nb + 7u) & ~7u;
        if (nb == 0) nb = 1;
        c.nb = (int)nb;
        return true;
    }
    return false;
}

inline int grid_for(size_t n)
{
    size_t nb = (n + QK_TPB - 1) / QK_TPB;
    if (nb > QK_BLAS_BLOCKS) nb = QK_BLAS_BLOCKS;
    return nb == 0 ? 1 : (int)nb;
}

inline int rc(hipError_t e) { return e == hipSuccess ? 0 : -(int)e; }

// Lanczos-form extras of the single-vector apply: the device step counter
// (plus offset) of graph replays, and the partial sums that total to nrm[k]
struct LzArgs {
    const int* kdev = nullptr;
    int koff = 0;
    const double* nrm_part = nullptr;
    int n_nrm_part = 0;
    int n_part = 0;           // partial sums the caller holds room for
    double* nrm_host = nullptr;   // also store nrm[k] here (host memory), when set
};

template <typename ST, bool CPLX, int RM, bool LZ, bool SIGNS, bool ACC = false>
const void* apply_fn()
{
    return reinterpret_cast<const void*>(&k_apply<ST, CPLX, RM, LZ, SIGNS, ACC>);
}

template <typename ST, bool CPLX, int RM, bool LZ, bool SIGNS, bool ACC = false>
int launch_apply(const QkTables* T, const double* x, double* y, int k, const double* prev, double* nrm,
                 double* partials, hipStream_t s, const LzArgs& lz, const double* acc = nullptr)
{
    ApplyCfg c;
    if (!apply_config(T, CPLX, apply_fn<ST, CPLX, RM, LZ, SIGNS, ACC>(), c)) return -2;
    if (LZ && lz.n_part != c.nb) return -1;   // one partial sum per workgroup
    hipLaunchKernelGGL((k_apply<ST, CPLX, RM, LZ, SIGNS, ACC>), dim3(c.nb), dim3(64 * c.wpb), c.lds, s, T->tiles,
                       T->hoff, T->lrank, T->tl, T->th, T->nth,
                       T->lolist, T->brk, T->dg, make_scal(T), x, y, k, prev, nrm, partials, (int)table_bytes(T),
                       lz.kdev, lz.koff, lz.nrm_part, lz.n_nrm_part, acc, lz.nrm_host);
    return rc(hipGetLastError());
}

// the accumulating (second-pass) instances: real vectors, 32-bit states,
// no parity masks, one-level lo rank
template <bool LZ>
int launch_apply_acc(const QkTables* T, const double* x, double* y, int k, const double* prev, double* nrm,
                     double* partials, hipStream_t s, const LzArgs& lz, const double* acc)
{
    if (T->signs != 0) return -1;
    if (T->rank_mode == 1)
        return launch_apply<uint32_t, false, 1, LZ, false, true>(T, x, y, k, prev, nrm, partials, s, lz, acc);
    if (T->rank_mode == 0)
        return launch_apply<uint32_t, false, 0, LZ, false, true>(T, x, y, k, prev, nrm, partials, s, lz, acc);
    return -1;
}

template <typename ST, int RM, bool LZ, int NW>
int launch_block_nw(const QkTables* T, const BlockCfg& c, const double* X, double* Y, int ld, int c0, int nc, int k,
                 const double* P, const double* nrm, int ld_sc, double* partials, int npart, hipStream_t s)
{
    if (LZ && npart != c.nb) return -1;
    QkScal sc = make_scal(T);
    sc.sched = T->sched_block;
    sc.tsh = block_tsh(T);
    hipLaunchKernelGGL((k_apply_block<ST, RM, LZ, NW>), dim3(c.nb), dim3(64 * NW), c.lds, s, T->tiles, T->hoff,
                       T->lrank, T->tl, T->th, T->nth, T->lolist, T->brk, T->dg, sc, X, Y, ld, c0, nc, k, P, nrm, ld_sc, partials,
                       npart, c.cap, (int)table_bytes(T));
    return rc(hipGetLastError());
}

template <typename ST, int RM, bool LZ>
int launch_block(const QkTables* T, const BlockCfg& c, const double* X, double* Y, int ld, int c0, int nc, int k,
                 const double* P, const double* nrm, int ld_sc, double* partials, int npart, hipStream_t s)
{
    if (c.nw == 4) return launch_block_nw<ST, RM, LZ, 4>(T, c, X, Y, ld, c0, nc, k, P, nrm, ld_sc, partials, npart, s);
    if (c.nw == 8) return launch_block_nw<ST, RM, LZ, 8>(T, c, X, Y, ld, c0, nc, k, P, nrm, ld_sc, partials, npart, s);
    return launch_block_nw<ST, RM, LZ, 16>(T, c, X, Y, ld, c0, nc, k, P, nrm, ld_sc, partials, npart, s);
}

template <typename ST, bool C, bool LZ, bool SG>
int apply_rm(const QkTables* T, const double* x, double* y, int k, const double* prev, double* nrm,
             double* partials, hipStream_t s, const LzArgs& lz)
{
    switch (T->rank_mode) {
    case 2: return launch_apply<ST, C, 2, LZ, SG>(T, x, y, k, prev, nrm, partials, s, lz);
    case 1: return launch_apply<ST, C, 1, LZ, SG>(T, x, y, k, prev, nrm, partials, s, lz);
    default: return launch_apply<ST, C, 0, LZ, SG>(T, x, y, k, prev, nrm, partials, s, lz);
    }
}

template <typename ST, bool C>
int apply_any(const QkTables* T, bool lzf, const double* x, double* y, int k, const double* prev, double* nrm,
              double* partials, hipStream_t s, const LzArgs& lz)
{
    const bool sg = T->signs != 0;
    if (lzf) return sg ? apply_rm<ST, C, true, true>(T, x, y, k, prev, nrm, partials, s, lz)
                       : apply_rm<ST, C, true, false>(T, x, y, k, prev, nrm, partials, s, lz);
    return sg ? apply_rm<ST, C, false, true>(T, x, y, k, prev, nrm, partials, s, lz)
              : apply_rm<ST, C, false, false>(T, x, y, k, prev, nrm, partials, s, lz);
}

template <typename ST, bool LZ>
int block_rm(const QkTables* T, const BlockCfg& c, const double* X, double* Y, int ld, int c0, int nc, int k,
             const double* P, const double* nrm, int ld_sc, double* partials, int npart, hipStream_t s)
{
    switch (T->rank_mode) {
    case 2: return launch_block<ST, 2, LZ>(T, c, X, Y, ld, c0, nc, k, P, nrm, ld_sc, partials, npart, s);
    case 1: return launch_block<ST, 1, LZ>(T, c, X, Y, ld, c0, nc, k, P, nrm, ld_sc, partials, npart, s);
    default: return launch_block<ST, 0, LZ>(T, c, X, Y, ld, c0, nc, k, P, nrm, ld_sc, partials, npart, s);
    }
}

// the Lanczos-form instance apply_rm launches for these tables
template <typename ST>
const void* lz_apply_fn(const QkTables* T)
{
    const bool sg = T->signs != 0;
    switch (T->rank_mode) {
    case 2: return sg ? apply_fn<ST, false, 2, true, true>() : apply_fn<ST, false, 2, true, false>();
    case 1: return sg ? apply_fn<ST, false, 1, true, true>() : apply_fn<ST, false, 1, true, false>();
    default: return sg ? apply_fn<ST, false, 0, true, true>() : apply_fn<ST, false, 0, true, false>();
    }
}

} // namespace

struct ColsLz {
    const double* P = nullptr;        // previous block
    double* nrm = nullptr;            // nrm[k * nv + c]
    double* partials = nullptr;       // this apply's partial dots
    int npart = 0;
    const double* nrm_part = nullptr; // the previous update's partial squared norms
    int n_nrm_part = 0;
};

template <int RM, bool LZ, int NV>
int launch_cols(const QkTables* T, const ApplyCfg& c, const double* X, double* Y, const double* dscale, int k,
                const ColsLz& lz, hipStream_t s)
{
    hipLaunchKernelGGL((k_apply_cols<RM, LZ, NV>), dim3(c.nb), dim3(64 * c.wpb), c.lds, s, T->tiles, T->hoff, T->lrank,
                       T->tl, T->th, T->nth, T->lolist, T->brk, T->dg, make_scal(T), X, Y, dscale, k, lz.P, lz.nrm,
                       lz.partials, lz.npart, (int)table_bytes(T), lz.nrm_part, lz.n_nrm_part);
    return rc(hipGetLastError());
}

template <int RM, bool LZ>
int cols_nv(int nv, const QkTables* T, const ApplyCfg& c, const double* X, double* Y, const double* dscale, int k,
            const ColsLz& lz, hipStream_t s)
{
    switch (nv) {
    case 1: return launch_cols<RM, LZ, 1>(T, c, X, Y, dscale, k, lz, s);
    case 2: return launch_cols<RM, LZ, 2>(T, c, X, Y, dscale, k, lz, s);
    case 4: return launch_cols<RM, LZ, 4>(T, c, X, Y, dscale, k, lz, s);
    case 8: return launch_cols<RM, LZ, 8>(T, c, X, Y, dscale, k, lz, s);
    default: return -1;
    }
}

template <int NV>
int launch_cols_blas(int op, size_t nrows, double* w, const double* q, double* a, const double* nrm, int k,
                     const double* a_part, int n_a_part, double* partials, hipStream_t s)
{
    if (op == 0)
        hipLaunchKernelGGL(k_update_cols<NV>, dim3(grid_for(nrows)), dim3(QK_TPB), 0, s, nrows, w, q, a, nrm, k, a_part,
                           n_a_part, partials);
    else if (op == 1)
        hipLaunchKernelGGL(k_axpy_cols<NV>, dim3(grid_for(nrows)), dim3(QK_TPB), 0, s, nrows, a, q, w);
    else
        hipLaunchKernelGGL(k_scale_rows_cols<NV>, dim3(grid_for(nrows)), dim3(QK_TPB), 0, s, nrows, a, q, w);
    return rc(hipGetLastError());
}

int cols_blas(int nv, int op, size_t nrows, double* w, const double* q, double* a, const double* nrm, int k,
              const double* a_part, int n_a_part, double* partials, hipStream_t s)
{
    switch (nv) {
    case 1: return launch_cols_blas<1>(op, nrows, w, q, a, nrm, k, a_part, n_a_part, partials, s);
    case 2: return launch_cols_blas<2>(op, nrows, w, q, a, nrm, k, a_part, n_a_part, partials, s);
    case 4: return launch_cols_blas<4>(op, nrows, w, q, a, nrm, k, a_part, n_a_part, partials, s);
    case 8: return launch_cols_blas<8>(op, nrows, w, q, a, nrm, k, a_part, n_a_part, partials, s);
    default: return -1;
    }
}

extern "C" {

int qkl_apply_partials(const QkTables* T, int wide)
{
    ApplyCfg c;
    const void* f = wide ? lz_apply_fn<uint64_t>(T) : lz_apply_fn<uint32_t>(T);
    return apply_config(T, false, f, c) ? c.nb : -2;
}
int qkl_apply_partials_acc(const QkTables* T)
{
    ApplyCfg c;
    const void* f = T->rank_mode == 1 ? apply_fn<uint32_t, false, 1, true, false, true>()
                                      : apply_fn<uint32_t, false, 0, true, false, true>();
    return apply_config(T, false, f, c) ? c.nb : -2;
}
int qkl_block_partials(const QkTables* T)
{
    BlockCfg c;
    return block_config(T, c) ? c.nb : -2;
}
int qkl_blas_partials(void) { return QK_BLAS_BLOCKS; }

int qkl_cols_partials(const QkTables* T)
{
    ApplyCfg c;
    return cols_config(T, c) ? c.nb : -2;
}

int qkl_apply_cols(const QkTables* T, int wide, int nv, const double* X, double* Y, const double* dscale, int k,
                   const double* P, double* nrm, double* partials, int npart, const double* nrm_part,
                   int n_nrm_part, qkl_stream sv)
{
    hipStream_t s = (hipStream_t)sv;
    if (wide || T->signs != 0 || !dscale) return -1;
    if (T->ntiles == 0) return 0;
    ApplyCfg c;
    if (!cols_config(T, c)) return -2;
    if (k >= 0 && (npart != c.nb || !nrm || !partials || !nrm_part || n_nrm_part <= 0 || (k > 0 && !P))) return -1;
    const bool lz = k >= 0;
    ColsLz a;
    a.P = P;
    a.nrm = nrm;
    a.partials = partials;
    a.npart = npart;
    a.nrm_part = nrm_part;
    a.n_nrm_part = n_nrm_part;
    switch (T->rank_mode) {
    case 2: return lz ? cols_nv<2, true>(nv, T, c, X, Y, dscale, k, a, s) : cols_nv<2, false>(nv, T, c, X, Y, dscale, k, a, s);
    case 1: return lz ? cols_nv<1, true>(nv, T, c, X, Y, dscale, k, a, s) : cols_nv<1, false>(nv, T, c, X, Y, dscale, k, a, s);
    default: return lz ? cols_nv<0, true>(nv, T, c, X, Y, dscale, k, a, s) : cols_nv<0, false>(nv, T, c, X, Y, dscale, k, a, s);
    }
}

int qkl_cols_update_partials(size_t nrows) { return grid_for(nrows); }

int qkl_update_cols(int nv, size_t nrows, double* w, const double* q, double* alpha, const double* nrm, int k,
                    const double* a_part, int n_a_part, double* partials, qkl_stream sv)
{
    if (a_part && n_a_part <= 0) return -1;
    return cols_blas(nv, 0, nrows, w, q, alpha, nrm, k, a_part, n_a_part, partials, (hipStream_t)sv);
}

int qkl_total_cols(int nv, const double* partials, int npart, double* out, int idx, qkl_stream sv)
{
    if (npart <= 0 || nv <= 0) return -1;
    hipLaunchKernelGGL(k_total_cols, dim3(nv), dim3(64), 0, (hipStream_t)sv, partials, npart, out, idx, nv);
    return rc(hipGetLastError());
}

int qkl_axpy_cols(int nv, size_t nrows, const double* coef, const double* X, double* Y, qkl_stream sv)
{
    return cols_blas(nv, 1, nrows, Y, X, const_cast<double*>(coef), nullptr, 0, nullptr, 0, nullptr, (hipStream_t)sv);
}

int qkl_scale_rows_cols(int nv, size_t nrows, const double* d, const double* X, double* Y, qkl_stream sv)
{
    return cols_blas(nv, 2, nrows, Y, X, const_cast<double*>(d), nullptr, 0, nullptr, 0, nullptr, (hipStream_t)sv);
}

int qkl_apply(const QkTables* T, int cplx, int wide, const double* x, double* y, qkl_stream sv)
{
    hipStream_t s = (hipStream_t)sv;
    if (T->ntiles == 0) return 0;
    const LzArgs none;
    if (wide) {
        return cplx ? apply_any<uint64_t, true>(T, false, x, y, -1, nullptr, nullptr, nullptr, s, none)
                    : apply_any<uint64_t, false>(T, false, x, y, -1, nullptr, nullptr, nullptr, s, none);
    }
    return cplx ? apply_any<uint32_t, true>(T, false, x, y, -1, nullptr, nullptr, nullptr, s, none)
                : apply_any<uint32_t, false>(T, false, x, y, -1, nullptr, nullptr, nullptr, s, none);
}

// The two kernels of a real single-vector Lanczos step with the
// reductions folded in (see qk_launch.h).
int qkl_lanczos_apply(const QkTables* T, int wide, const double* x, double* y, int k, const int* kdev, int koff,
                      const double* prev, double* nrm, const double* nrm_part, int n_nrm_part, double* partials,
                      int n_part, double* nrm_host, qkl_stream sv)
{
    hipStream_t s = (hipStream_t)sv;
    if (T->ntiles == 0 || !nrm_part || n_nrm_part <= 0) return -1;
    LzArgs lz;
    lz.nrm_host = nrm_host;
    lz.n_part = n_part;
    lz.kdev = kdev;
    lz.koff = koff;
    lz.nrm_part = nrm_part;
    lz.n_nrm_part = n_nrm_part;
    return wide ? apply_any<uint64_t, false>(T, true, x, y, k, prev, nrm, partials, s, lz)
                : apply_any<uint32_t, false>(T, true, x, y, k, prev, nrm, partials, s, lz);
}

int qkl_apply_acc(const QkTables* T, const double* x, double* y, const double* acc, qkl_stream sv)
{
    if (T->ntiles == 0) return 0;
    if (!acc) return -1;
    const LzArgs none;
    return launch_apply_acc<false>(T, x, y, -1, nullptr, nullptr, nullptr, (hipStream_t)sv, none, acc);
}

int qkl_lanczos_apply_acc(const QkTables* T, const double* x, double* y, int k, const int* kdev, int koff,
                          const double* prev, double* nrm, const double* nrm_part, int n_nrm_part, double* partials,
                          int n_part, const double* acc, double* nrm_host, qkl_stream sv)
{
    if (T->ntiles == 0 || !nrm_part || n_nrm_part <= 0 || !acc) return -1;
    LzArgs lz;
    lz.nrm_host = nrm_host;
    lz.n_part = n_part;
    lz.kdev = kdev;
    lz.koff = koff;
    lz.nrm_part = nrm_part;
    lz.n_nrm_part = n_nrm_part;
    return launch_apply_acc<true>(T, x, y, k, prev, nrm, partials, (hipStream_t)sv, lz, acc);
}

int qkl_lanczos_update1(size_t nrows, double* w, const double* q, double* alpha, const double* nrm, int k,
                        const int* kdev, int koff, const double* a_part, int n_a_part, double* partials, int* knext,
                        double* alpha_host, qkl_stream sv)
{
    if (!a_part || n_a_part <= 0) return -1;
    hipLaunchKernelGGL(k_update1<false>, dim3(QK_BLAS_BLOCKS), dim3(QK_TPB), 0, (hipStream_t)sv, nrows, w, q, alpha,
                       nrm, k, partials, kdev, koff, a_part, n_a_part, knext, alpha_host);
    return rc(hipGetLastError());
}

int qkl_total(const double* partials, int npart, int root, double* out, int idx, qkl_stream sv)
{
    hipLaunchKernelGGL(k_total, dim3(1), dim3(64), 0, (hipStream_t)sv, partials, npart, root, out, idx);
    return rc(hipGetLastError());
}

int qkl_apply_block(const QkTables* T, int wide, const double* X, double* Y, int ld, int c0, int nc, int lz_k,
                    const double* P, const double* nrm, int ld_sc, double* partials, int npart, qkl_stream sv)
{
    hipStream_t s = (hipStream_t)sv;
    if (nc <= 0 || nc > 64 || ld < c0 + nc) return -1;
    if (T->ntiles == 0) return 0;
    if (2 * (T->nA + T->nB) + 64 >= 0xFFFF) return -3;   // value ids are 16-bit
    BlockCfg c;
    if (!block_config(T, c)) return -2;
    const bool lz = lz_k >= 0;
    if (wide) return lz ? block_rm<uint64_t, true>(T, c, X, Y, ld, c0, nc, lz_k, P, nrm, ld_sc, partials, npart, s)
                        : block_rm<uint64_t, false>(T, c, X, Y, ld, c0, nc, lz_k, P, nrm, ld_sc, partials, npart, s);
    return lz ? block_rm<uint32_t, true>(T, c, X, Y, ld, c0, nc, lz_k, P, nrm, ld_sc, partials, npart, s)
              : block_rm<uint32_t, false>(T, c, X, Y, ld, c0, nc, lz_k, P, nrm, ld_sc, partials, npart, s);
}

int qkl_lanczos_update(int cplx, size_t nrows, int ncol, int ld, double* w, const double* q, const double* alpha,
                       const double* nrm, int k, int ld_sc, double* partials, qkl_stream sv)
{
    hipStream_t s = (hipStream_t)sv;
    if (ncol == 1 && ld == 1) {
        // without a_part the kernel only reads alpha
        double* aw = const_cast<double*>(alpha);
        if (cplx)
            hipLaunchKernelGGL(k_update1<true>, dim3(QK_BLAS_BLOCKS), dim3(QK_TPB), 0, s, nrows, w, q, aw, nrm, k,
                               partials, (const int*)nullptr, 0, (const double*)nullptr, 0, (int*)nullptr,
                               (double*)nullptr);
        else
            hipLaunchKernelGGL(k_update1<false>, dim3(QK_BLAS_BLOCKS), dim3(QK_TPB), 0, s, nrows, w, q, aw, nrm, k,
                               partials, (const int*)nullptr, 0, (const double*)nullptr, 0, (int*)nullptr,
                               (double*)nullptr);
    } else {
        if (cplx || ncol > 64) return -1;
        hipLaunchKernelGGL(k_update_block, dim3(QK_BLAS_BLOCKS), dim3(QK_TPB), 0, s, nrows, ncol, ld, w, q, alpha, nrm,
                           k, ld_sc, partials);
    }
    return rc(hipGetLastError());
}

int qkl_reduce(const double* partials, int npart, int ncol, int group, int mode, int cplx, double* out, int idx,
               int ld_sc, qkl_stream sv)
{
    hipStream_t s = (hipStream_t)sv;
    if (group <= 0 || ncol % group) return -1;
    hipLaunchKernelGGL(k_reduce, dim3(ncol / group), dim3(QK_TPB), 0, s, partials, npart, group, mode, cplx, out, idx,
                       ld_sc);
    return rc(hipGetLastError());
}

int qkl_dot_partials(int cplx, size_t n, const double* x, const double* y, double* partials, qkl_stream sv)
{
    hipStream_t s = (hipStream_t)sv;
    if (cplx) hipLaunchKernelGGL(k_dot<true>, dim3(QK_BLAS_BLOCKS), dim3(QK_TPB), 0, s, n, x, y, partials);
    else hipLaunchKernelGGL(k_dot<false>, dim3(QK_BLAS_BLOCKS), dim3(QK_TPB), 0, s, n, x, y, partials);
    return rc(hipGetLastError());
}

int qkl_axpy(int cplx, size_t n, double ar, double ai, const double* x, double* y, qkl_stream sv)
{
    hipStream_t s = (hipStream_t)sv;
    if (cplx) hipLaunchKernelGGL(k_axpy<true>, dim3(grid_for(n)), dim3(QK_TPB), 0, s, n, ar, ai, x, y);
    else hipLaunchKernelGGL(k_axpy<false>, dim3(grid_for(n)), dim3(QK_TPB), 0, s, n, ar, ai, x, y);
    return rc(hipGetLastError());
}

int qkl_scal(int cplx, size_t n, double ar, double ai, double* x, qkl_stream sv)
{
    hipStream_t s = (hipStream_t)sv;
    if (cplx) hipLaunchKernelGGL(k_scal<true>, dim3(grid_for(n)), dim3(QK_TPB), 0, s, n, ar, ai, x);
    else hipLaunchKernelGGL(k_scal<false>, dim3(grid_for(n)), dim3(QK_TPB), 0, s, n, ar, ai, x);
    return rc(hipGetLastError());
}

int qkl_copy_scaled(size_t n, const double* src, double* dst, const double* sc, int idx, qkl_stream sv)
{
    hipLaunchKernelGGL(k_copy_scaled, dim3(grid_for(n)), dim3(QK_TPB), 0, (hipStream_t)sv, n, src, dst, sc, idx);
    return rc(hipGetLastError());
}

int qkl_fill_random(int kind, size_t nrows, uint64_t seed, uint64_t counter0, int ncol, int ld, int c0, double* x,
                    qkl_stream sv)
{
    hipLaunchKernelGGL(k_fill_random, dim3(grid_for(nrows * (size_t)ncol)), dim3(QK_TPB), 0, (hipStream_t)sv, kind,
                       nrows, ncol, ld, c0, seed, counter0, x);
    return rc(hipGetLastError());
}

int qkl_gather_probe(const double* x, size_t n, int bytes_per_lane, double* y, int blocks, int tiles, qkl_stream sv)
{
    // 8 waves per workgroup; y holds blocks * 512 doubles
    const size_t seg_doubles = bytes_per_lane == 16 ? 128 : 64;
    if ((bytes_per_lane != 8 && bytes_per_lane != 16) || n < seg_doubles || blocks <= 0 || tiles <= 0 || !x || !y)
        return -1;
    const size_t nseg = n / seg_doubles;
    if (bytes_per_lane == 8)
        hipLaunchKernelGGL(k_gather_probe<8>, dim3(blocks), dim3(512), 0, (hipStream_t)sv, x, nseg, y, (uint32_t)tiles);
    else
        hipLaunchKernelGGL(k_gather_probe<16>, dim3(blocks), dim3(512), 0, (hipStream_t)sv, x, nseg, y, (uint32_t)tiles);
    return rc(hipGetLastError());
}

int qkl_copy(size_t n, const double* src, double* dst, qkl_stream sv)
{
    hipLaunchKernelGGL(k_copy, dim3(grid_for(n)), dim3(QK_TPB), 0, (hipStream_t)sv, n, src, dst);
    return rc(hipGetLastError());
}

int qkl_copy_strided(size_t nrows, int width, const double* src, size_t src_ld, double* dst, size_t dst_ld,
                     qkl_stream sv)
{
    if (width <= 0 || (size_t)width > src_ld || (size_t)width > dst_ld) return -1;
    hipLaunchKernelGGL(k_copy_strided, dim3(grid_for(nrows * (size_t)width)), dim3(QK_TPB), 0, (hipStream_t)sv, nrows,
                       width, src, src_ld, dst, dst_ld);
    return rc(hipGetLastError());
}

int qkl_set_vals(double* dst, int n, const double* vals, qkl_stream sv)
{
    if (n < 0 || n > 8 || (n > 0 && (!dst || !vals))) return -1;
    if (n == 0) return 0;
    QkVals8 v{};
    for (int i = 0; i < n; ++i) v.v[i] = vals[i];
    hipLaunchKernelGGL(k_set_vals, dim3(1), dim3(64), 0, (hipStream_t)sv, dst, n, v);
    return rc(hipGetLastError());
}

int qkl_set_int(int* p, int v, qkl_stream sv)
{
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, (hipStream_t)sv, p, v);
    return rc(hipGetLastError());
}

int qkl_fill_zero(size_t n, double* x, qkl_stream sv)
{
    hipLaunchKernelGGL(k_fill_zero, dim3(grid_for(n)), dim3(QK_TPB), 0, (hipStream_t)sv, n, x);
    return rc(hipGetLastError());
}

int qkl_multidot(size_t n, int m, const double* V, size_t ldv, const double* u, double* partials, qkl_stream sv)
{
    if (m <= 0) return 0;
    hipLaunchKernelGGL(k_multidot, dim3(QK_BLAS_BLOCKS), dim3(QK_TPB), 0, (hipStream_t)sv, n, m, V, ldv, u, partials);
    return rc(hipGetLastError());
}

int qkl_multiaxpy(size_t n, int m, const double* V, size_t ldv, const double* coef, double sgn, double* u,
                  qkl_stream sv)
{
    if (m <= 0) return 0;
    hipLaunchKernelGGL(k_multiaxpy, dim3(grid_for(n)), dim3(QK_TPB), 0, (hipStream_t)sv, n, m, V, ldv, coef, sgn, u);
    return rc(hipGetLastError());
}

int qkl_diagonal(const QkTables* T, int cplx, int wide, double* d, qkl_stream sv)
{
    const size_t dim = (size_t)T->dim;
    if (dim == 0) return 0;
    if (wide) hipLaunchKernelGGL(k_diagonal<uint64_t>, dim3(grid_for(dim)), dim3(QK_TPB), 0, (hipStream_t)sv, T->tiles, T->lolist, T->dg, make_scal(T), cplx, dim, d);
    else hipLaunchKernelGGL(k_diagonal<uint32_t>, dim3(grid_for(dim)), dim3(QK_TPB), 0, (hipStream_t)sv, T->tiles, T->lolist, T->dg, make_scal(T), cplx, dim, d);
    return rc(hipGetLastError());
}

int qkl_row_entries(const QkTables* T, int wide, uint64_t r0, uint64_t nrows, int cap, int64_t* cols, double* vals,
                    int* counts, qkl_stream sv)
{
    if (nrows == 0) return 0;
    if (wide) hipLaunchKernelGGL(k_row_entries<uint64_t>, dim3(grid_for(nrows)), dim3(QK_TPB), 0, (hipStream_t)sv, T->tiles, T->hoff, T->lrank, T->lolist, T->br, T->dg, make_scal(T), (size_t)r0, (size_t)nrows, cap, cols, vals, counts);
    else hipLaunchKernelGGL(k_row_entries<uint32_t>, dim3(grid_for(nrows)), dim3(QK_TPB), 0, (hipStream_t)sv, T->tiles, T->hoff, T->lrank, T->lolist, T->br, T->dg, make_scal(T), (size_t)r0, (size_t)nrows, cap, cols, vals, counts);
    return rc(hipGetLastError());
}

int qkl_states(const QkTables* T, int wide, uint64_t r0, uint64_t n, uint64_t* out, qkl_stream sv)
{
    if (n == 0) return 0;
    if (wide) hipLaunchKernelGGL(k_states<uint64_t>, dim3(grid_for(n)), dim3(QK_TPB), 0, (hipStream_t)sv, T->tiles, T->lolist, make_scal(T), (size_t)r0, (size_t)n, out);
    else hipLaunchKernelGGL(k_states<uint32_t>, dim3(grid_for(n)), dim3(QK_TPB), 0, (hipStream_t)sv, T->tiles, T->lolist, make_scal(T), (size_t)r0, (size_t)n, out);
    return rc(hipGetLastError());
}

int qkl_apply_diag_terms(const QkTables* T, int wide, int xcplx, const double* x, double* y, qkl_stream sv)
{
    const size_t dim = (size_t)T->dim;
    if (dim == 0) return 0;
    if (wide) hipLaunchKernelGGL(k_apply_diag_terms<uint64_t>, dim3(grid_for(dim)), dim3(QK_TPB), 0, (hipStream_t)sv, T->tiles, T->lolist, T->dg, make_scal(T), xcplx, dim, x, y);
    else hipLaunchKernelGGL(k_apply_diag_terms<uint32_t>, dim3(grid_for(dim)), dim3(QK_TPB), 0, (hipStream_t)sv, T->tiles, T->lolist, T->dg, make_scal(T), xcplx, dim, x, y);
    return rc(hipGetLastError());
}

int qkl_davidson_precond(int cplx, size_t n, const double* r, const double* diag, double lambda, double* t,
                         qkl_stream sv)
{
    if (cplx) hipLaunchKernelGGL(k_precond<true>, dim3(grid_for(n)), dim3(QK_TPB), 0, (hipStream_t)sv, n, r, diag, lambda, t);
    else hipLaunchKernelGGL(k_precond<false>, dim3(grid_for(n)), dim3(QK_TPB), 0, (hipStream_t)sv, n, r, diag, lambda, t);
    return rc(hipGetLastError());
}

// device address and size of the static workspace on the current device:
// hipGetSymbolAddress once per device (with that device current), kept
int qkl_scratch(void** ptr, size_t* bytes)
{
    static std::mutex mu;
    static void* addr[64] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return -(int)e;
    if (dev < 0 || dev >= 64) return -1;
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return e != hipSuccess ? -(int)e : -1;
        addr[dev] = p;
    }
    *ptr = addr[dev];
    *bytes = QK_SCRATCH_BYTES;
    return 0;
}

int qkl_permute(const QkTables* Tint, const uint32_t* bytemap, const QkTables* Tpub, int to_pub, const double* src,
                double* dst, qkl_stream sv)
{
    if (Tint->dim != Tpub->dim || Tint->NL <= 0 || Tint->NL >= 32 || Tpub->NL <= 0 || Tpub->NL >= 32) return -1;
    if (Tint->ntiles == 0) return 0;
    uint32_t nb = (Tint->ntiles + 3u) / 4u;
    if (nb > 16384u) nb = 16384u;
    if (to_pub)
        hipLaunchKernelGGL(k_permute<true>, dim3(nb), dim3(QK_TPB), 0, (hipStream_t)sv, Tint->tiles, Tint->ntiles,
                           Tint->lolist, Tint->NL, bytemap, Tpub->hoff, Tpub->lrank, Tpub->NL, src, dst);
    else
        hipLaunchKernelGGL(k_permute<false>, dim3(nb), dim3(QK_TPB), 0, (hipStream_t)sv, Tint->tiles, Tint->ntiles,
                           Tint->lolist, Tint->NL, bytemap, Tpub->hoff, Tpub->lrank, Tpub->NL, src, dst);
    return rc(hipGetLastError());
}

} // extern "C"
