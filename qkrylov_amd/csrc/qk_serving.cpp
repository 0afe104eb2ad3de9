// qk_serving.cpp -- continuous-batching request loop for ground-state and
// S^zz(pi, omega) continued-fraction solves of XXZ rings.
//
// Each admitted request owns a stream, its Krylov vectors (stream-ordered
// allocations) and its alpha/beta arrays; all requests in flight enqueue
// their steps without host synchronisation, so their kernels overlap on the
// device, and the host only polls events. Requests are admitted first-fit
// in arrival order while the HBM footprint of the active set stays under
// the budget and a stream slot is free. Every kernel and reduction has a
// fixed grid and order, so a request's result does not depend on what else
// is running.
#include "qkrylov_amd.h"

#include "qk_engine.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <vector>

namespace qk {
namespace {

OpSum xxz_ring(int L, double delta)
{
    OpSum o;
    for (int i = 0; i < L; ++i) {
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

enum Phase { kGS, kRitz, kCF, kDone };

struct Job {
    int id = 0;
    qkrylov_serve_request_t rq{};
    hipStream_t s = nullptr;
    hipEvent_t ev = nullptr;
    std::shared_ptr<BasisSpec> basis;
    std::unique_ptr<Engine> H, Op;
    std::unique_ptr<RealCore> C;
    std::unique_ptr<LanczosReplay> rp;
    Pinned *ha = nullptr, *hn = nullptr;   // the slot's read-back buffers
    double* gs_out = nullptr;              // caller's host buffer for the ground state (or null)
    DevBuf ritz, psi2;
    Phase phase = kGS;
    int kcap = 0, k = 0, n_gs = 0;
    bool inflight = false;
    std::vector<double> z, nrm_gs;
    size_t footprint = 0;
    double start = 0.0;
    uint64_t seed = 0;
};

size_t dim_of(const qkrylov_serve_request_t& r)
{
    Sector s;
    s.use_sz = true;
    s.sz2 = r.sz2;
    return (size_t)BasisSpec(Kind::SpinHalf, r.L, s).dim();
}

size_t footprint_of(const qkrylov_serve_request_t& r)
{
    const size_t D = dim_of(r);
    const size_t vecs = 3 + (r.kind == 1 ? 1 : 0);
    return vecs * D * sizeof(double) + (size_t)(r.maxiter + r.cf_steps + 8) * 16 + (1 << 20);
}

constexpr int kChunk = 12;   // Lanczos steps enqueued per poll

void copy_scalars(Job& j, int nsteps)
{
    hip_check(hipMemcpyAsync(j.ha->p, j.C->alpha.d(), nsteps * sizeof(double), hipMemcpyDeviceToHost, j.s), "d2h");
    hip_check(hipMemcpyAsync(j.hn->p, j.C->nrm.d(), (nsteps + 1) * sizeof(double), hipMemcpyDeviceToHost, j.s), "d2h");
}

void enqueue_gs_chunk(Job& j)
{
    const int kend = std::min(j.kcap, j.k + kChunk);
    // whole turns of the vector rotation as graph replays, the rest eager
    const int n3 = (kend - j.k) / 3 * 3;
    if (n3 > 0) {
        j.C->steps_graph(j.k, n3);
        j.k += n3;
    }
    for (; j.k < kend; ++j.k) j.C->step(j.k, nullptr);
    j.C->flush();
    copy_scalars(j, j.k);
    hip_check(hipEventRecord(j.ev, j.s), "event");
    j.inflight = true;
}

void start_job(Job& j)
{
    Sector sec;
    sec.use_sz = true;
    sec.sz2 = j.rq.sz2;
    j.basis = std::make_shared<BasisSpec>(Kind::SpinHalf, j.rq.L, sec);
    EngineOptions eo;
    eo.stream = j.s;
    eo.NL = choose_NL(*j.basis);   // the split whose tables the run keeps resident
    j.H = std::make_unique<Engine>(j.basis, Kind::SpinHalf, xxz_ring(j.rq.L, j.rq.delta), eo);
    j.kcap = (int)std::min<int64_t>(j.rq.maxiter, j.basis->dim());
    const int cap = std::max(j.kcap, j.rq.cf_steps) + 2;
    j.C = std::make_unique<RealCore>(*j.H, cap, j.s, false, true);
    j.rp = std::make_unique<LanczosReplay>(LanczosReplay{j.kcap, j.rq.tol, false, {}, {}});
    j.seed = 1234;
    if (qkl_fill_random(0, (size_t)j.basis->dim(), j.seed, 0, 1, 1, 0, j.C->cur(), j.s) != 0)
        throw std::runtime_error("start vector");
    j.C->init_norm();
    j.phase = kGS;
    j.k = 0;
    enqueue_gs_chunk(j);
}

// called when the job's last event completed; enqueues the next phase
void advance(Job& j, qkrylov_serve_result_t& res, double* cf_a, double* cf_b, int cf_cap)
{
    const size_t D = (size_t)j.basis->dim();
    if (j.phase == kGS) {
        while (!j.rp->done && j.rp->next < j.k) j.rp->feed(j.ha->p[j.rp->next], j.hn->p[j.rp->next + 1]);
        if (!j.rp->done && j.k >= j.kcap) j.rp->done = true;
        if (!j.rp->done) {
            enqueue_gs_chunk(j);
            return;
        }
        j.n_gs = (int)j.rp->alphas.size();
        res.iterations = j.rp->next;
        res.converged = j.rp->converged ? 1 : 0;
        res.energy = tridiag_min(j.rp->alphas.data(), j.rp->betas.data(), j.n_gs, j.rq.kind == 1 ? &j.z : nullptr);
        if (j.rq.kind != 1) {
            j.phase = kDone;
            return;
        }
        // Ritz vector by a second pass over the same deterministic recurrence
        j.nrm_gs.assign(j.hn->p, j.hn->p + j.n_gs);
        j.ritz = DevBuf(D * sizeof(double), j.s);
        j.psi2 = DevBuf(2 * sizeof(double), j.s);
        if (qkl_fill_zero(D, j.ritz.d(), j.s) != 0) throw std::runtime_error("zero");
        if (qkl_fill_random(0, D, j.seed, 0, 1, 1, 0, j.C->cur(), j.s) != 0) throw std::runtime_error("start");
        j.C->init_norm();
        for (int kk = 0; kk < j.n_gs; ++kk) {
            if (qkl_axpy(0, D, j.z[kk] / j.nrm_gs[kk], 0.0, j.C->cur(), j.ritz.d(), j.s) != 0)
                throw std::runtime_error("axpy");
            if (kk + 1 < j.n_gs) j.C->step(kk, nullptr);
        }
        hip_check(hipEventRecord(j.ev, j.s), "event");
        j.inflight = true;
        j.phase = kRitz;
        return;
    }
    if (j.phase == kRitz) {
        // phi = S^z_pi psi, then the plain recurrence from phi
        Sector sec;
        sec.use_sz = true;
        sec.sz2 = j.rq.sz2;
        EngineOptions eo;
        eo.stream = j.s;
        eo.NL = choose_NL(*j.basis);
        j.Op = std::make_unique<Engine>(j.basis, Kind::SpinHalf, szpi(j.rq.L), eo);
        j.Op->apply(j.ritz.d(), j.C->cur(), false, j.s);
        if (qkl_dot_partials(0, D, j.ritz.d(), j.ritz.d(), j.C->partB.d(), j.s) != 0) throw std::runtime_error("dot");
        if (qkl_reduce(j.C->partB.d(), j.C->NB, 1, 1, 0, 0, j.psi2.d(), 0, 1, j.s) != 0)
            throw std::runtime_error("reduce");
        j.C->init_norm();
        {
            const int n3 = j.rq.cf_steps / 3 * 3;
            if (n3 > 0) j.C->steps_graph(0, n3);
            for (int kk = n3; kk < j.rq.cf_steps; ++kk) j.C->step(kk, nullptr);
        }
        j.C->flush();
        copy_scalars(j, j.rq.cf_steps);
        hip_check(hipMemcpyAsync(j.ha->p + j.rq.cf_steps + 1, j.psi2.d(), sizeof(double), hipMemcpyDeviceToHost, j.s),
                  "d2h");
        hip_check(hipEventRecord(j.ev, j.s), "event");
        j.inflight = true;
        j.phase = kCF;
        return;
    }
    if (j.phase == kCF) {
        const double n0 = j.hn->p[0];
        const double psi2 = j.ha->p[j.rq.cf_steps + 1];
        res.cf_norm = (n0 < 1e-15 || psi2 <= 0) ? 0.0 : n0 / std::sqrt(psi2);
        int na = 0;
        if (n0 >= 1e-15) {
            for (int kk = 0; kk < j.rq.cf_steps; ++kk) {
                if (cf_a && na < cf_cap) cf_a[(size_t)j.id * cf_cap + na] = j.ha->p[kk];
                ++na;
                const double b = j.hn->p[kk + 1];
                if (b < 1e-15) break;
                if (cf_b && kk < cf_cap) cf_b[(size_t)j.id * cf_cap + kk] = b;
            }
        }
        res.n_cf = na;
        if (j.gs_out) {
            // the normalised Ritz vector the continued fraction started from
            hip_check(hipMemcpyAsync(j.gs_out, j.ritz.get(), D * sizeof(double), hipMemcpyDeviceToHost, j.s), "d2h");
            hip_check(hipStreamSynchronize(j.s), "sync");
            const double inv = (psi2 > 0) ? 1.0 / std::sqrt(psi2) : 0.0;
            for (size_t i = 0; i < D; ++i) j.gs_out[i] *= inv;
        }
        j.phase = kDone;
    }
}

// pinned read-back room a request needs (alpha and norm arrays)
size_t readback_doubles(const qkrylov_serve_request_t& r)
{
    const size_t D = dim_of(r);
    const int kcap = (int)std::min<size_t>((size_t)r.maxiter, D);
    return (size_t)std::max(kcap, r.cf_steps) + 2 + 3;
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
        const size_t budget = (opt && opt->hbm_budget_bytes) ? opt->hbm_budget_bytes : (size_t)(0.8 * free_b);
        const int max_active = (opt && opt->max_active > 0) ? opt->max_active : 32;
        const bool realtime = opt ? opt->realtime != 0 : true;
        for (int i = 0; i < nreq; ++i) {
            if (reqs[i].L < 2 || reqs[i].L > 32 || reqs[i].maxiter <= 0 || reqs[i].cf_steps < 0) {
                set_last_error("serve: request with L outside 2..32, maxiter <= 0 or cf_steps < 0");
                return QKRYLOV_ERROR_INVALID_ARG;
            }
            if (dim_of(reqs[i]) == 0) {
                set_last_error("serve: request whose Sz sector is empty (sz2 out of range or of the wrong parity "
                               "for L)");
                return QKRYLOV_ERROR_INVALID_ARG;
            }
            if (footprint_of(reqs[i]) > budget) {   // could never be admitted
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
        // streams, events and jobs are released on every exit; on an error
        // the streams are drained first, so no job's memory is freed while
        // its kernels may still run
        // the read-back buffers outlive `release` below, which drains the
        // streams before anything else is freed
        std::vector<std::unique_ptr<Pinned>> pin_a, pin_n;
        std::vector<hipStream_t> streams;
        std::vector<hipEvent_t> events;
        std::vector<std::unique_ptr<Job>> active;
        struct Release {
            std::vector<hipStream_t>& st;
            std::vector<hipEvent_t>& ev;
            std::vector<std::unique_ptr<Job>>& jobs;
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
        std::vector<int> order(nreq);
        for (int i = 0; i < nreq; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return reqs[a].arrival < reqs[b].arrival; });
        std::deque<int> pending(order.begin(), order.end());
        std::vector<int> free_slots;
        for (int i = max_active - 1; i >= 0; --i) free_slots.push_back(i);
        std::vector<int> slot_of;
        size_t used = 0;
        int done = 0;
        const double t0 = now_s();
        while (done < nreq) {
            bool progress = false;
            const double now = now_s() - t0;
            // admission: first fit in arrival order
            for (auto it = pending.begin(); it != pending.end() && !free_slots.empty();) {
                const int id = *it;
                if (realtime && reqs[id].arrival > now) break;
                const size_t fp = footprint_of(reqs[id]);
                if (used + fp > budget) {
                    ++it;
                    continue;
                }
                auto j = std::make_unique<Job>();
                j->id = id;
                j->rq = reqs[id];
                const int slot = free_slots.back();
                free_slots.pop_back();
                j->s = streams[slot];
                j->ev = events[slot];
                j->ha = pin_a[slot].get();
                j->hn = pin_n[slot].get();
                j->gs_out = (opt && opt->gs_vectors && reqs[id].kind == 1) ? opt->gs_vectors[id] : nullptr;
                j->footprint = fp;
                j->start = now_s() - t0;
                results[id].start = j->start;
                start_job(*j);
                used += fp;
                slot_of.push_back(slot);
                active.push_back(std::move(j));
                it = pending.erase(it);
                progress = true;
            }
            for (size_t a = 0; a < active.size();) {
                Job& j = *active[a];
                if (j.inflight) {
                    const hipError_t q = hipEventQuery(j.ev);
                    if (q == hipErrorNotReady) {
                        ++a;
                        continue;
                    }
                    hip_check(q, "event query");
                    j.inflight = false;
                    progress = true;
                    advance(j, results[j.id], cf_a, cf_b, cf_cap);
                }
                if (j.phase == kDone && !j.inflight) {
                    auto& r = results[j.id];
                    r.finish = now_s() - t0;
                    r.latency = r.finish - (realtime ? reqs[j.id].arrival : 0.0);
                    used -= j.footprint;
                    free_slots.push_back(slot_of[a]);
                    active.erase(active.begin() + (long)a);
                    slot_of.erase(slot_of.begin() + (long)a);
                    ++done;
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
