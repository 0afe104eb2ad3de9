// qk_serving.cpp -- continuous-batching request loop for ground-state and
// S^zz(pi, omega) continued-fraction solves of XXZ rings.
//
// Requests are admitted in arrival order while the HBM footprint of what is
// on the device stays under the budget and a stream slot is free. Waiting
// requests of the same (L, Sz) sector are admitted together as one batch:
// they share the exchange part of H and differ only in the anisotropy,
// which scales the Ising diagonal, so they run as the columns of one block
// recurrence (ColsCore) in which every decoded matrix element serves all
// columns. Each batch owns a stream, its Krylov blocks (stream-ordered
// allocations) and its alpha/norm arrays; all batches in flight enqueue
// their steps without host synchronisation, so their kernels overlap on the
// device, and the host only polls events. A batch runs three phases:
//   GS    the Lanczos recurrence of every column until each column's replay
//         of the reference's convergence test is done (columns that finish
//         early keep being stepped, their numbers unused);
//   Ritz  for the columns that also want S^zz(pi, omega): the recurrence
//         again from the same start vectors (a narrower block of just those
//         columns), accumulating each column's Ritz vector;
//   CF    S^z_pi applied to the Ritz vectors, then the continued-fraction
//         recurrence of those columns.
// While the GS phase runs, the Krylov vectors of the S^zz columns are kept
// (in chunks of steps, taken as the run goes on from a store reserved beside
// the budget), and the Ritz vectors are then summed from them; when the
// store cannot hold them, the Ritz phase recomputes them instead. Both give
// the same bits. Each phase enqueues a few milliseconds of steps per poll,
// a small request's phase often in one poll; the polls of the largest
// requests go onto the device at most two at a time, oldest first.
// Nothing in the loop waits for the device: engines are built per sector
// once per run with plain uploads, the store is carved on the host, the
// per-column scalars are set by kernels, and every free is stream-ordered or
// deferred to the end of the run. A ground-state request is finished when
// the GS phase ends. Every kernel and reduction of a column has a fixed grid
// and order that depend neither on the block width nor on the other columns,
// so a request's result is the same bits alone (a batch of one) and in any
// batch.
#include "qkrylov_amd.h"

#include "qk_engine.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace qk {
namespace {

// Widest batch by default, and the widest column block (ColsCore). Batches
// take power-of-two sizes, so no column of a block is padding. On an MI355X
// the XXZ ring's column apply is bound by the bytes its gathers move, which
// grow with the width: per column it costs 0.63x the single-vector apply at
// two and four columns and 0.83x at eight (L = 28), so four is the default.
constexpr int kDefaultBatch = 4, kMaxBatch = 8;

// bonds (i, i + 1 mod L); a two-site ring has its one bond once
OpSum xxz_ring(int L, double delta)
{
    OpSum o;
    for (int i = 0; i < (L == 2 ? 1 : L); ++i) {
        const int j = (i + 1) % L;
        o.terms.push_back({delta, {{"Sz", i}, {"Sz", j}}});
        o.terms.push_back({0.5, {{"Sp", i}, {"Sm", j}}});
        o.terms.push_back({0.5, {{"Sm", i}, {"Sp", j}}});
    }
    return o;
}

// S^z at q = pi: L^-1/2 sum_j (-1)^j Sz_j (real, diagonal)
OpSum szpi(int L)
{
    OpSum o;
    const double c = 1.0 / std::sqrt((double)L);
    for (int j = 0; j < L; ++j) o.terms.push_back({(j & 1) ? -c : c, {{"Sz", j}}});
    return o;
}

// Device memory for the stored Krylov vectors, reserved once per run. The
// chunks are requested as a run goes on, in many sizes and up to gigabytes;
// growing the stream-ordered pool by such blocks while batches run held the
// host for seconds at a time (one measured call took 2.2 s), stalling every
// batch it had to feed. Here a chunk is a first-fit range of one reserved
// buffer, handed out and returned on the host without any HIP call; a
// returned range is reused only after the batch that held it has passed an
// event recorded after its last use.
// The buffer is allocated by the first run on a device that stores vectors
// and kept by the process (allocating and freeing tens of gigabytes per run
// made later runs slower and slower); each run uses at most its own limit
// of it. Runs on one device are serialised, since they share it.
struct StoreArena {
    char* base = nullptr;
    size_t cap = 0;
    std::map<size_t, size_t> free_;   // offset -> bytes, disjoint and merged
    // start of a run: the first `limit` bytes (at most the buffer) are free
    void reset(size_t limit)
    {
        free_.clear();
        limit = std::min(limit, cap) & ~(size_t)255;
        if (limit) free_[0] = limit;
    }
    bool take(size_t bytes, size_t& off)
    {
        bytes = (bytes + 255) & ~(size_t)255;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            off = it->first;
            const size_t rest = it->second - bytes;
            free_.erase(it);
            if (rest) free_[off + bytes] = rest;
            return true;
        }
        return false;
    }
    void give(size_t off, size_t bytes)
    {
        bytes = (bytes + 255) & ~(size_t)255;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

constexpr size_t kStoreArenaMax = (size_t)64 << 30;

struct DeviceArena {
    std::mutex run_mu;   // one serving run at a time on the device
    StoreArena arena;
};

DeviceArena& device_arena(int dev)
{
    static std::mutex mu;
    static std::map<int, std::unique_ptr<DeviceArena>> all;
    std::lock_guard<std::mutex> lk(mu);
    auto& p = all[dev];
    if (!p) p = std::make_unique<DeviceArena>();
    return *p;
}

// the device's arena, allocated by the first call that can give it room
// (at most kStoreArenaMax, within `room`); call with run_mu held
StoreArena& store_arena(DeviceArena& da, size_t want)
{
    StoreArena& a = da.arena;
    want = std::min(want, kStoreArenaMax) & ~(size_t)255;
    if (!a.base && want) {
        if (hipMalloc(reinterpret_cast<void**>(&a.base), want) == hipSuccess) {
            a.cap = want;
        } else {
            (void)hipGetLastError();
            a.base = nullptr;
        }
    }
    return a;
}

// pinned host buffers for the alpha / norm read-backs, one pair per stream
// slot, allocated once per run (freeing pinned memory synchronises the
// device, so nothing is freed while requests are in flight)
struct Pinned {
    double* p = nullptr;
    explicit Pinned(size_t n)
    {
        if (hipHostMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(double), hipHostMallocDefault) != hipSuccess)
            throw std::runtime_error("hipHostMalloc failed");
    }
    ~Pinned()
    {
        if (p) (void)hipHostFree(p);
    }
    Pinned(const Pinned&) = delete;
    Pinned& operator=(const Pinned&) = delete;
};

size_t dim_of(int L, int sz2)
{
    Sector s;
    s.use_sz = true;
    s.sz2 = sz2;
    return (size_t)BasisSpec(Kind::SpinHalf, L, s).dim();
}

int kcap_of(const qkrylov_serve_request_t& r)
{
    return (int)std::min<size_t>((size_t)r.maxiter, dim_of(r.L, r.sz2));
}

// device bytes of a batch at its widest phase: the GS block recurrence
// (three blocks of nv columns), then, for the S^zz columns (nvr of them,
// padded), their recurrence plus the Ritz block; the GS blocks are released
// before the second phase allocates
size_t footprint_of(const qkrylov_serve_request_t* const* rq, int n)
{
    const size_t D = dim_of(rq[0]->L, rq[0]->sz2);
    int ncf = 0, kmax = 0;
    for (int i = 0; i < n; ++i) {
        ncf += rq[i]->kind == 1 ? 1 : 0;
        kmax = std::max(kmax, std::max(rq[i]->maxiter, rq[i]->cf_steps));
    }
    const size_t nv = (size_t)cols_width(n), nvr = ncf ? (size_t)cols_width(ncf) : 0;
    const size_t vecs = std::max(3 * nv, 4 * nvr);
    return vecs * D * sizeof(double) + (size_t)(kmax + 8) * 16 * nv + (2 << 20);
}

constexpr int kChunk = 12;   // recurrence steps per stored-vector chunk

enum Phase { kGS, kRitz, kCF, kDone };
// the batch's next enqueue, held back while the window of large polls is full
enum Next { kNone, kGsChunk, kStartRitz, kRitzChunk, kStartCF, kCFChunk };

struct Column {
    int id = 0;
    qkrylov_serve_request_t rq{};
    LanczosReplay rp{0, 0.0, false, {}, {}};
    int kcap = 0, n_gs = 0;
    std::vector<double> z, nrm_gs;   // Ritz coefficients and the GS norms they divide by
    double* gs_out = nullptr;        // caller's host buffer for the ground state (or null)
    bool ritz_stored = false;        // the Ritz vector came from stored Krylov vectors
};

struct Batch {
    std::vector<Column> cols;        // the S^zz columns first
    std::vector<int> cf;             // indices into cols of the S^zz columns (0 .. ncf - 1)
    int L = 0, sz2 = 0;
    hipStream_t s = nullptr;
    hipEvent_t ev = nullptr;
    Pinned *ha = nullptr, *hn = nullptr;   // the slot's read-back buffers
    std::shared_ptr<BasisSpec> basis;
    std::shared_ptr<const Engine> H, Op;   // the sector's, shared (SectorEngines)
    std::unique_ptr<ColsCore> C;
    DevBuf ritz, coef, psi2;
    // stored Krylov vectors of the S^zz columns: step k in chunks[k / kChunk]
    // (offset, bytes in the run's StoreArena) at row-major width
    // cols_width(ncf); `store` is dropped when the arena cannot take the
    // next chunk. Dropped chunks wait in `retired` for the batch's next
    // event.
    bool store = false;
    std::vector<std::pair<size_t, size_t>> chunks, retired;
    StoreArena* arena = nullptr;
    Phase phase = kGS;
    Next next = kNone;
    bool large = false;              // the poll in flight counts against the window of large polls
    int k = 0, kend = 0;             // steps enqueued; steps the phase runs at most
    int rn = 0;                      // Ritz phase without stored vectors: terms of the sum
    bool inflight = false;
    size_t footprint = 0;
    uint64_t seed = 1234;
};

// column c of a width-nv block gets the start vector a single request uses
// (uniform [-1, 1) from the counter RNG, counter = row)
void fill_start(Batch& b, int nv)
{
    const size_t D = (size_t)b.basis->dim();
    for (int c = 0; c < nv; ++c)
        if (qkl_fill_random(0, D, b.seed, (uint64_t)0 - (uint64_t)c * D, 1, nv, c, b.C->cur(), b.s) != 0)
            throw std::runtime_error("start vector");
}

void copy_scalars(Batch& b, int nsteps)
{
    const size_t nv = (size_t)b.C->nv;
    b.C->flush(nsteps);
    hip_check(hipMemcpyAsync(b.ha->p, b.C->alpha.d(), nsteps * nv * sizeof(double), hipMemcpyDeviceToHost, b.s), "d2h");
    hip_check(hipMemcpyAsync(b.hn->p, b.C->nrm.d(), (nsteps + 1) * nv * sizeof(double), hipMemcpyDeviceToHost, b.s),
              "d2h");
}

void drop_store(Batch& b)
{
    b.store = false;
    b.retired.insert(b.retired.end(), b.chunks.begin(), b.chunks.end());
    b.chunks.clear();
}

// the batch's retired chunks go back to the arena (call after an event the
// batch recorded behind their last use has completed)
void give_back(Batch& b)
{
    for (const auto& c : b.retired) b.arena->give(c.first, c.second);
    b.retired.clear();
}

// Recurrence steps one poll enqueues: about 4 ms of device work. The stream
// slots share a few hardware queues that run their packets in order, so a
// long run of one batch's steps holds back every batch queued behind it;
// polls of similar length keep the slots sharing the device evenly. Each
// poll of a batch in turn waits behind what the others have queued, so a
// small request is served in as few polls as its steps fit: alone an L = 22
// request takes 7 ms, and at 12 steps per poll it took 0.39 s in the
// Poisson stream. The step cost is modelled from the ring's column apply
// (about 1.6e-8 ms per row and column on an MI355X, and at least 15 us for
// the step's two launches).
double step_ms(const Batch& b, int nv)
{
    return 1.6e-8 * (double)b.basis->dim() * (double)nv;
}

int steps_per_poll(const Batch& b, int nv)
{
    return std::max(1, (int)(4.0 / std::max(0.015, step_ms(b, nv))));
}

double* stored_step(Batch& b, int k)
{
    const size_t D = (size_t)b.basis->dim(), w = (size_t)cols_width((int)b.cf.size());
    return reinterpret_cast<double*>(b.arena->base + b.chunks[(size_t)k / kChunk].first) + (size_t)(k % kChunk) * D * w;
}

void enqueue_gs_chunk(Batch& b)
{
    // the first poll runs up to 64 steps (requests of the stream converge
    // after 50-70), later ones up to 16; steps past a column's convergence
    // are run and not used
    const int want = std::max(16, 64 - b.k);
    const int kend = std::min(b.kend, b.k + std::min(want, steps_per_poll(b, b.C->nv)));
    const size_t D = (size_t)b.basis->dim();
    const int ncf = (int)b.cf.size(), w = ncf ? cols_width(ncf) : 0;
    for (; b.k < kend; ++b.k) {
        if (b.store && (size_t)b.k / kChunk >= b.chunks.size()) {
            const size_t bytes = (size_t)kChunk * D * w * sizeof(double);
            size_t off = 0;
            if (b.arena && b.arena->take(bytes, off)) b.chunks.emplace_back(off, bytes);
            else drop_store(b);
        }
        // q_k of the S^zz columns (the block's first ncf columns) before step k
        if (b.store && qkl_copy_strided(D, ncf, b.C->cur(), b.C->nv, stored_step(b, b.k), w, b.s) != 0)
            throw std::runtime_error("store Krylov vector");
        b.C->step(b.k);
    }
    copy_scalars(b, b.k);
    hip_check(hipEventRecord(b.ev, b.s), "event");
    b.inflight = true;
}

// The engines of one (L, Sz) sector, shared by every batch of the sector
// for the whole run: H (the exchange part and the Ising diagonal at
// Delta = 1; column c scales the diagonal by its own Delta) and the S^z_pi
// operator of the continued fraction. They are built with plain (not
// stream-ordered) uploads: a copy ordered on a batch's stream would make the
// host wait for that stream's hardware queue, which the other batches' work
// shares, and the run keeps them to its end, so that no hipFree (which
// waits for the whole device) happens while requests are in flight.
struct SectorEngines {
    std::shared_ptr<BasisSpec> basis;
    std::shared_ptr<const Engine> H, Op;
};
using EngineCache = std::map<std::pair<int, int>, SectorEngines>;

const SectorEngines& sector_engines(EngineCache& cache, int L, int sz2, bool want_op)
{
    SectorEngines& e = cache[{L, sz2}];
    if (!e.H) {
        Sector sec;
        sec.use_sz = true;
        sec.sz2 = sz2;
        e.basis = std::make_shared<BasisSpec>(Kind::SpinHalf, L, sec);
        EngineOptions eo;
        eo.NL = choose_NL(*e.basis);   // the split whose tables the run keeps resident
        auto H = std::make_shared<Engine>(e.basis, Kind::SpinHalf, xxz_ring(L, 1.0), eo);
        // one front of XCD-contiguous tile chunks: 2-6 % faster column
        // applies than one segment per XCD for rings of L = 24-28 (launch
        // shapes never change results)
        H->set_tuning(0, 2);
        e.H = H;
    }
    if (want_op && !e.Op) {
        EngineOptions eo;
        eo.NL = e.H->NL();
        e.Op = std::make_shared<Engine>(e.basis, Kind::SpinHalf, szpi(L), eo);
    }
    return e;
}

void start_batch(Batch& b, EngineCache& engines, StoreArena& arena)
{
    bool any_cf = false;
    for (const auto& col : b.cols) any_cf = any_cf || col.rq.kind == 1;
    const SectorEngines& se = sector_engines(engines, b.L, b.sz2, any_cf);
    b.basis = se.basis;
    b.H = se.H;
    b.Op = se.Op;
    b.arena = &arena;
    const int n = (int)b.cols.size(), nv = cols_width(n);
    std::vector<double> dsc(nv);
    b.kend = 0;
    for (int c = 0; c < nv; ++c) dsc[c] = b.cols[c < n ? c : 0].rq.delta;   // padding repeats column 0
    for (auto& col : b.cols) {
        col.kcap = kcap_of(col.rq);
        col.rp = LanczosReplay{col.kcap, col.rq.tol, false, {}, {}};
        b.kend = std::max(b.kend, col.kcap);
        if (col.rq.kind == 1) b.cf.push_back((int)(&col - b.cols.data()));
    }
    b.C = std::make_unique<ColsCore>(*b.H, nv, b.kend + 2, dsc.data(), b.s, true);
    fill_start(b, nv);
    b.C->init_norm();
    b.phase = kGS;
    b.k = 0;
    b.store = !b.cf.empty();
    b.next = kGsChunk;
}

void start_cf(Batch& b);
void enqueue_ritz_chunk(Batch& b);
void enqueue_cf_chunk(Batch& b);

// the CF columns' Ritz vectors: each column adds z_k q_k / |q_k| for its own
// k < n_gs, from the stored Krylov vectors or, without them, from the
// recurrence run again from the same start vectors
void start_ritz(Batch& b)
{
    const size_t D = (size_t)b.basis->dim();
    const int ncf = (int)b.cf.size(), nvr = cols_width(ncf);
    int nmax = 0;
    for (int j : b.cf) nmax = std::max(nmax, b.cols[j].n_gs);
    // the coefficients go through the slot's pinned buffer, whose GS
    // read-back has been consumed
    double* cf = b.hn->p;
    std::fill(cf, cf + (size_t)nmax * nvr, 0.0);
    std::vector<double> dsc(nvr);
    int cfmax = 0;
    for (int j = 0; j < nvr; ++j) {
        const Column& col = b.cols[b.cf[j < ncf ? j : 0]];
        dsc[j] = col.rq.delta;
        if (j >= ncf) continue;   // padding: coefficient 0 keeps its Ritz column at zero
        cfmax = std::max(cfmax, col.rq.cf_steps);
        for (int kk = 0; kk < col.n_gs; ++kk) cf[(size_t)kk * nvr + j] = col.z[kk] / col.nrm_gs[kk];
    }
    b.C.reset();   // stream-ordered release of the GS blocks
    b.C = std::make_unique<ColsCore>(*b.H, nvr, std::max(nmax, cfmax) + 2, dsc.data(), b.s, true);
    b.ritz = DevBuf(D * nvr * sizeof(double), b.s);
    b.coef = DevBuf(std::max<size_t>((size_t)nmax * nvr, 1) * sizeof(double), b.s);
    b.psi2 = DevBuf(nvr * sizeof(double), b.s);
    hip_check(hipMemcpyAsync(b.coef.get(), cf, (size_t)nmax * nvr * sizeof(double), hipMemcpyHostToDevice, b.s),
              "h2d");
    if (qkl_fill_zero(D * nvr, b.ritz.d(), b.s) != 0) throw std::runtime_error("zero");
    b.kend = cfmax;
    if (b.store) {
        for (int kk = 0; kk < nmax; ++kk)
            if (qkl_axpy_cols(nvr, D, b.coef.d() + (size_t)kk * nvr, stored_step(b, kk), b.ritz.d(), b.s) != 0)
                throw std::runtime_error("axpy");
        drop_store(b);   // back to the arena after the next event, behind the sums that read them
        for (int j : b.cf) b.cols[j].ritz_stored = true;
        start_cf(b);
        return;
    }
    fill_start(b, nvr);
    b.C->init_norm();
    b.rn = nmax;
    b.k = 0;
    b.phase = kRitz;
    enqueue_ritz_chunk(b);
}

// terms [k, ...) of the recomputed Ritz sum: add term k, then step k
void enqueue_ritz_chunk(Batch& b)
{
    const size_t D = (size_t)b.basis->dim();
    const int nvr = b.C->nv;
    const int kend = std::min(b.rn, b.k + steps_per_poll(b, nvr));
    for (; b.k < kend; ++b.k) {
        if (qkl_axpy_cols(nvr, D, b.coef.d() + (size_t)b.k * nvr, b.C->cur(), b.ritz.d(), b.s) != 0)
            throw std::runtime_error("axpy");
        if (b.k + 1 < b.rn) b.C->step(b.k);
    }
    hip_check(hipEventRecord(b.ev, b.s), "event");
    b.inflight = true;
}

// phi_c = S^z_pi psi_c, |psi_c|^2, then cf_steps steps of the plain recurrence from phi
void start_cf(Batch& b)
{
    const size_t D = (size_t)b.basis->dim();
    const int nvr = b.C->nv;
    b.Op->apply_cols(b.ritz.d(), b.C->cur(), nvr, b.C->one.d(), b.s);
    // an update with coefficient 0 / 1 leaves the Ritz block unchanged and
    // yields its column sums of squares
    if (qkl_update_cols(nvr, D, b.ritz.d(), b.ritz.d(), b.C->zero.d(), b.C->one.d(), 0, nullptr, 0, b.C->partB.d(),
                        b.s) != 0 ||
        qkl_reduce(b.C->partB.d(), b.C->NB, nvr, 1, 0, 0, b.psi2.d(), 0, nvr, b.s) != 0)
        throw std::runtime_error("psi norms");
    b.C->init_norm();
    b.k = 0;
    b.phase = kCF;
    enqueue_cf_chunk(b);
}

// CF steps [k, ...); after the last one the read-back of the scalars and of
// |psi|^2
void enqueue_cf_chunk(Batch& b)
{
    const int nvr = b.C->nv;
    const int kend = std::min(b.kend, b.k + steps_per_poll(b, nvr));
    for (; b.k < kend; ++b.k) b.C->step(b.k);
    if (b.k == b.kend) {
        copy_scalars(b, b.kend);
        hip_check(hipMemcpyAsync(b.ha->p + (size_t)(b.kend + 1) * nvr, b.psi2.d(), nvr * sizeof(double),
                                 hipMemcpyDeviceToHost, b.s),
                  "d2h");
    }
    hip_check(hipEventRecord(b.ev, b.s), "event");
    b.inflight = true;
}

// called when the batch's last event completed: takes the results it
// brings and enqueues the next phase; returns the requests it finished
std::vector<int> advance(Batch& b, qkrylov_serve_result_t* results, double* cf_a, double* cf_b, int cf_cap)
{
    std::vector<int> finished;
    const size_t D = (size_t)b.basis->dim();
    if (b.phase == kGS) {
        const int nv = b.C->nv;
        bool all = true;
        for (int c = 0; c < (int)b.cols.size(); ++c) {
            Column& col = b.cols[c];
            if (col.rp.done) continue;
            while (!col.rp.done && col.rp.next < b.k)
                col.rp.feed(b.ha->p[(size_t)col.rp.next * nv + c], b.hn->p[(size_t)(col.rp.next + 1) * nv + c]);
            if (!col.rp.done && b.k >= col.kcap) col.rp.done = true;
            if (!col.rp.done) {
                all = false;
                continue;
            }
            col.n_gs = (int)col.rp.alphas.size();
            qkrylov_serve_result_t& r = results[col.id];
            r.iterations = col.rp.next;
            r.converged = col.rp.converged ? 1 : 0;
            r.energy = tridiag_min(col.rp.alphas.data(), col.rp.betas.data(), col.n_gs,
                                   col.rq.kind == 1 ? &col.z : nullptr);
            if (col.rq.kind == 1) {
                col.nrm_gs.resize(col.n_gs);
                for (int kk = 0; kk < col.n_gs; ++kk) col.nrm_gs[kk] = b.hn->p[(size_t)kk * nv + c];
            } else {
                finished.push_back(col.id);
            }
        }
        if (!all) {
            b.next = kGsChunk;
            return finished;
        }
        if (b.cf.empty()) {
            b.phase = kDone;
            return finished;
        }
        b.next = kStartRitz;
        return finished;
    }
    if (b.phase == kRitz) {
        b.next = b.k < b.rn ? kRitzChunk : kStartCF;
        return finished;
    }
    if (b.phase == kCF && b.k < b.kend) {
        b.next = kCFChunk;
        return finished;
    }
    if (b.phase == kCF) {
        const int nvr = b.C->nv;
        for (int j = 0; j < (int)b.cf.size(); ++j) {
            Column& col = b.cols[b.cf[j]];
            qkrylov_serve_result_t& res = results[col.id];
            const double n0 = b.hn->p[j];
            const double psi2 = b.ha->p[(size_t)(b.kend + 1) * nvr + j];
            res.cf_norm = (n0 < 1e-15 || psi2 <= 0) ? 0.0 : n0 / std::sqrt(psi2);
            int na = 0;
            if (n0 >= 1e-15) {
                for (int kk = 0; kk < col.rq.cf_steps; ++kk) {
                    if (cf_a && na < cf_cap) cf_a[(size_t)col.id * cf_cap + na] = b.ha->p[(size_t)kk * nvr + j];
                    ++na;
                    const double be = b.hn->p[(size_t)(kk + 1) * nvr + j];
                    if (be < 1e-15) break;
                    if (cf_b && kk < cf_cap) cf_b[(size_t)col.id * cf_cap + kk] = be;
                }
            }
            res.n_cf = na;
            res.ritz_stored = col.ritz_stored ? 1 : 0;
            if (col.gs_out) {
                // the normalised Ritz vector the continued fraction started from
                hip_check(hipMemcpy2DAsync(col.gs_out, sizeof(double), b.ritz.d() + j, nvr * sizeof(double),
                                           sizeof(double), D, hipMemcpyDeviceToHost, b.s),
                          "d2h");
                hip_check(hipStreamSynchronize(b.s), "sync");
                const double inv = (psi2 > 0) ? 1.0 / std::sqrt(psi2) : 0.0;
                for (size_t i = 0; i < D; ++i) col.gs_out[i] *= inv;
            }
            finished.push_back(col.id);
        }
        b.phase = kDone;
    }
    return finished;
}

// enqueues the batch's next poll (each ends with the batch's event)
void launch(Batch& b)
{
    const Next n = b.next;
    b.next = kNone;
    switch (n) {
    case kGsChunk: enqueue_gs_chunk(b); break;
    case kStartRitz: start_ritz(b); break;
    case kRitzChunk: enqueue_ritz_chunk(b); break;
    case kStartCF: start_cf(b); break;
    case kCFChunk: enqueue_cf_chunk(b); break;
    case kNone: break;
    }
}

// Large polls, those of batches whose step costs at least 0.1 ms (rings of
// L = 26-28), are let onto the device at most two at a time, oldest batch
// first; smaller polls always go. Left to share the device evenly, a large
// request is slowed by everything that overlaps it (L = 28 requests ran
// 2.4-3.5x their time alone in the Poisson stream at about two-thirds
// load); two large polls keep the device busy, so the oldest large
// requests run close to their time alone and the small ones still pass.
constexpr int kLargeWindow = 2;

bool large_poll(const Batch& b)
{
    return step_ms(b, b.C->nv) >= 0.1;
}

// pinned read-back room of one request's column (alpha and norm arrays, the psi norms)
size_t readback_doubles(const qkrylov_serve_request_t& r)
{
    return (size_t)std::max(kcap_of(r), r.cf_steps) + 2 + 3;
}

} // namespace
} // namespace qk

extern "C" int qkrylov_serve_run(const qkrylov_serve_request_t* reqs, int nreq, const qkrylov_serve_options_t* opt,
                                 qkrylov_serve_result_t* results, double* cf_a, double* cf_b, int cf_cap,
                                 double* totals)
{
    using namespace qk;
    if (!reqs || nreq <= 0 || !results) return QKRYLOV_ERROR_INVALID_ARG;
    try {
        int dev = 0;
        hip_check(hipGetDevice(&dev), "hipGetDevice");
        size_t free_b = 0, total_b = 0;
        hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
        // the budget admits batches by their footprint; beside it, up to a
        // quarter of it (within the free memory the budget leaves, less a
        // margin) is reserved for the S^zz requests' stored Krylov vectors
        const size_t budget = (opt && opt->hbm_budget_bytes) ? opt->hbm_budget_bytes : (size_t)(0.8 * free_b);
        bool any_cf = false;
        for (int i = 0; i < nreq; ++i) any_cf = any_cf || reqs[i].kind == 1;
        const size_t margin = (size_t)1 << 30;
        const size_t room = free_b > budget + margin ? free_b - budget - margin : 0;
        const size_t store_bytes = any_cf ? std::min(budget / 4, room) : 0;
        const int max_active = (opt && opt->max_active > 0) ? opt->max_active : 32;
        const bool realtime = opt ? opt->realtime != 0 : true;
        const int max_batch = (opt && opt->max_batch > 0) ? std::min(opt->max_batch, kMaxBatch) : kDefaultBatch;
        for (int i = 0; i < nreq; ++i) {
            if (reqs[i].L < 2 || reqs[i].L > 32 || reqs[i].maxiter <= 0 || reqs[i].cf_steps < 0) {
                set_last_error("serve: request with L outside 2..32, maxiter <= 0 or cf_steps < 0");
                return QKRYLOV_ERROR_INVALID_ARG;
            }
            if (dim_of(reqs[i].L, reqs[i].sz2) == 0) {
                set_last_error("serve: request whose Sz sector is empty (sz2 out of range or of the wrong parity "
                               "for L)");
                return QKRYLOV_ERROR_INVALID_ARG;
            }
            const qkrylov_serve_request_t* one = &reqs[i];
            if (footprint_of(&one, 1) > budget) {   // could never be admitted
                set_last_error("serve: a request's device footprint exceeds the memory budget");
                return QKRYLOV_ERROR_INVALID_ARG;
            }
            std::memset(&results[i], 0, sizeof(results[i]));
        }
        // keep every basis table alive for the whole run: no hipFree while
        // requests are in flight
        std::map<std::pair<int, int>, std::shared_ptr<const BasisTables>> keep;
        for (int i = 0; i < nreq; ++i) {
            const auto key = std::make_pair(reqs[i].L, reqs[i].sz2);
            if (keep.count(key)) continue;
            Sector s;
            s.use_sz = true;
            s.sz2 = reqs[i].sz2;
            BasisSpec b(Kind::SpinHalf, reqs[i].L, s);
            keep[key] = basis_tables(b, choose_NL(b));
        }
        // streams, events and batches are released on every exit; on an
        // error the streams are drained first, so no batch's memory is freed
        // while its kernels may still run; the read-back buffers outlive
        // `release` below
        std::vector<std::unique_ptr<Pinned>> pin_a, pin_n;
        // every sector's engines (and through them its basis tables), kept
        // to the end of the run; declared before `release`, so they go after
        // the streams are drained
        EngineCache engines;
        DeviceArena& darena = device_arena(dev);
        std::lock_guard<std::mutex> run_lock(darena.run_mu);
        StoreArena& arena = store_arena(darena, any_cf ? room : 0);
        arena.reset(store_bytes);
        std::vector<hipStream_t> streams;
        std::vector<hipEvent_t> events;
        std::vector<std::unique_ptr<Batch>> active;
        struct Release {
            std::vector<hipStream_t>& st;
            std::vector<hipEvent_t>& ev;
            std::vector<std::unique_ptr<Batch>>& jobs;
            ~Release()
            {
                for (auto s : st) (void)hipStreamSynchronize(s);
                jobs.clear();
                for (auto s : st) (void)hipStreamDestroy(s);
                for (auto e : ev) (void)hipEventDestroy(e);
            }
        } release{streams, events, active};
        size_t rb = 1;
        for (int i = 0; i < nreq; ++i) rb = std::max(rb, readback_doubles(reqs[i]));
        rb *= (size_t)cols_width(max_batch);
        for (int i = 0; i < max_active; ++i) {
            hipStream_t st = nullptr;
            hip_check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "stream");
            streams.push_back(st);
            hipEvent_t ev = nullptr;
            hip_check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "event");
            events.push_back(ev);
            pin_a.push_back(std::make_unique<Pinned>(rb));
            pin_n.push_back(std::make_unique<Pinned>(rb));
        }
        // the device's stream-ordered pool keeps what the run frees (Krylov
        // blocks, stored vectors) for the next batch instead of returning it
        // at every synchronisation, and may not hand a block freed on one
        // stream to another stream by making the second wait for the first:
        // such inserted waits chain unrelated batches together (stalls of
        // seconds under load were measured). Restored when the run ends.
        struct PoolHold {
            hipMemPool_t pool = nullptr;
            uint64_t old = 0;
            int old_dep = 1;
            bool dep_set = false;
            explicit PoolHold(int dev)
            {
                if (hipDeviceGetDefaultMemPool(&pool, dev) != hipSuccess ||
                    hipMemPoolGetAttribute(pool, hipMemPoolAttrReleaseThreshold, &old) != hipSuccess) {
                    pool = nullptr;
                    (void)hipGetLastError();
                    return;
                }
                uint64_t keep = UINT64_MAX;
                if (hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep) != hipSuccess) {
                    pool = nullptr;
                    (void)hipGetLastError();
                    return;
                }
                int no = 0;
                if (hipMemPoolGetAttribute(pool, hipMemPoolReuseAllowInternalDependencies, &old_dep) == hipSuccess &&
                    hipMemPoolSetAttribute(pool, hipMemPoolReuseAllowInternalDependencies, &no) == hipSuccess)
                    dep_set = true;
                else
                    (void)hipGetLastError();
            }
            ~PoolHold()
            {
                if (pool) {
                    if (dep_set) (void)hipMemPoolSetAttribute(pool, hipMemPoolReuseAllowInternalDependencies, &old_dep);
                    (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &old);
                    (void)hipMemPoolTrimTo(pool, 0);
                }
            }
        } pool_hold(dev);
        std::vector<int> order(nreq);
        for (int i = 0; i < nreq; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return reqs[a].arrival < reqs[b].arrival; });
        std::list<int> pending(order.begin(), order.end());
        std::vector<int> free_slots;
        for (int i = max_active - 1; i >= 0; --i) free_slots.push_back(i);
        std::vector<int> slot_of;
        size_t used = 0;
        int done = 0, large_inflight = 0;
        const double t0 = now_s();
        auto finish = [&](int id) {
            auto& r = results[id];
            r.finish = now_s() - t0;
            r.latency = r.finish - (realtime ? reqs[id].arrival : 0.0);
            ++done;
        };
        while (done < nreq) {
            bool progress = false;
            const double now = now_s() - t0;
            // admission: first fit in arrival order; a request takes along
            // the waiting requests of its sector that fit
            for (auto it = pending.begin(); it != pending.end() && !free_slots.empty();) {
                const int id = *it;
                if (realtime && reqs[id].arrival > now) break;
                std::vector<const qkrylov_serve_request_t*> mem{&reqs[id]};
                if (used + footprint_of(mem.data(), 1) > budget) {
                    ++it;
                    continue;
                }
                std::vector<std::list<int>::iterator> take{it};
                for (auto jt = std::next(it); jt != pending.end() && (int)mem.size() < max_batch; ++jt) {
                    const int o = *jt;
                    if (realtime && reqs[o].arrival > now) break;
                    if (reqs[o].L != reqs[id].L || reqs[o].sz2 != reqs[id].sz2) continue;
                    mem.push_back(&reqs[o]);
                    if (used + footprint_of(mem.data(), (int)mem.size()) > budget) {
                        mem.pop_back();
                        continue;
                    }
                    take.push_back(jt);
                }
                // a power-of-two batch (the rest stays waiting for the next one)
                while (cols_width((int)take.size()) != (int)take.size()) {
                    take.pop_back();
                    mem.pop_back();
                }
                auto b = std::make_unique<Batch>();
                b->L = reqs[id].L;
                b->sz2 = reqs[id].sz2;
                for (int pass = 0; pass < 2; ++pass)   // the S^zz columns first
                    for (auto t : take) {
                        if ((reqs[*t].kind == 1) != (pass == 0)) continue;
                        Column col;
                        col.id = *t;
                        col.rq = reqs[*t];
                        col.gs_out = (opt && opt->gs_vectors && reqs[*t].kind == 1) ? opt->gs_vectors[*t] : nullptr;
                        b->cols.push_back(std::move(col));
                    }
                const int slot = free_slots.back();
                free_slots.pop_back();
                b->s = streams[slot];
                b->ev = events[slot];
                b->ha = pin_a[slot].get();
                b->hn = pin_n[slot].get();
                b->footprint = footprint_of(mem.data(), (int)mem.size());
                const double st = now_s() - t0;
                for (const auto& col : b->cols) {
                    results[col.id].start = st;
                    results[col.id].batch = (int)b->cols.size();
                }
                // the iterator `it` is among those erased: continue after the last one
                auto next = std::next(take.back());
                for (auto t : take) pending.erase(t);
                it = (take.size() == 1) ? next : pending.begin();
                used += b->footprint;
                start_batch(*b, engines, arena);
                slot_of.push_back(slot);
                active.push_back(std::move(b));
                progress = true;
            }
            // batches in admission order, so the oldest take the large-poll window first
            for (size_t a = 0; a < active.size();) {
                Batch& b = *active[a];
                if (b.inflight) {
                    const hipError_t q = hipEventQuery(b.ev);
                    if (q == hipErrorNotReady) {
                        ++a;
                        continue;
                    }
                    hip_check(q, "event query");
                    b.inflight = false;
                    if (b.large) {
                        b.large = false;
                        --large_inflight;
                    }
                    progress = true;
                    give_back(b);
                    for (int id : advance(b, results, cf_a, cf_b, cf_cap)) finish(id);
                }
                if (b.next != kNone) {
                    const bool large = large_poll(b);
                    if (!large || large_inflight < kLargeWindow) {
                        launch(b);
                        if (large) {
                            b.large = true;
                            ++large_inflight;
                        }
                        progress = true;
                    }
                }
                if (b.phase == kDone && !b.inflight && b.next == kNone) {
                    used -= b.footprint;
                    drop_store(b);
                    give_back(b);
                    free_slots.push_back(slot_of[a]);
                    active.erase(active.begin() + (long)a);
                    slot_of.erase(slot_of.begin() + (long)a);
                    continue;
                }
                ++a;
            }
            if (!progress) std::this_thread::yield();
        }
        const double wall = now_s() - t0;
        if (totals) {
            std::vector<double> lat(nreq);
            for (int i = 0; i < nreq; ++i) lat[i] = results[i].latency;
            std::sort(lat.begin(), lat.end());
            auto pct = [&](double p) {
                const double pos = p * (nreq - 1);
                const int lo = (int)std::floor(pos);
                const int hi = std::min(nreq - 1, lo + 1);
                return lat[lo] + (lat[hi] - lat[lo]) * (pos - lo);
            };
            totals[0] = wall;
            totals[1] = nreq / wall;
            totals[2] = pct(0.50);
            totals[3] = pct(0.99);
        }
        return QKRYLOV_SUCCESS;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return QKRYLOV_ERROR_EXCEPTION;
    }
}
