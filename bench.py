"""Benchmark of the matrix-free H.x apply and the Lanczos loop on one MI355X.

Default workload: J1-J2 Heisenberg model on the 4x8 periodic square lattice
(Lx = 4, Ly = 8, site index x + 4y; J1 = 1, J2 = 0.5), Sz = 0 sector,
D = 601,080,390, real FP64. A "step" is
one full apply y = H x with x and y resident in HBM. The JSON line reports
the apply throughput in algorithmic bytes (16 B per row: x read once, y
written once) over the kernel's mean time from HIP events on the engine's
stream, Lanczos iterations per second (200 plain three-term steps, no
reorthogonalisation), and the oracle (a C/OpenMP restatement of the
reference's CSR build + complex gather SpMV) timed on a row sample of the
same workload on this host.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "j1j2_4x8": dict(desc="J1-J2 Heisenberg 4x8 periodic square lattice (Lx=4, Ly=8, site x + 4y), J1=1 J2=0.5, Sz=0",
                     kind="spin", n=32, sz2=0, lanczos=200),
    "j1j2_8x4": dict(desc="J1-J2 Heisenberg 4x8 periodic square lattice labelled along the long side (Lx=8, Ly=4, "
                          "site x + 8y), J1=1 J2=0.5, Sz=0", kind="spin", n=32, sz2=0, lanczos=200),
    "heis_ring32": dict(desc="Heisenberg ring L=32 periodic, J=1, Sz=0", kind="spin", n=32, sz2=0, lanczos=200),
    "heis_ring20": dict(desc="Heisenberg ring L=20 periodic, J=1, Sz=0", kind="spin", n=20, sz2=0, lanczos=300),
    "hubbard16": dict(desc="Fermi-Hubbard 16-site periodic chain, t=1 U=4, Nup=Ndn=8", kind="hubbard", n=16,
                      lanczos=250),
    "tj4x5": dict(desc="t-J 4x5 open lattice, t=1 J=0.4, 9 up 9 dn (2 holes)", kind="tj", n=20, lanczos=100),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_terms(name):
    from qkrylov_amd import models
    if name == "j1j2_4x8":
        return models.j1j2_square(8, 4, 1.0, 0.5)     # 8 rows of 4 sites
    if name == "j1j2_8x4":
        return models.j1j2_square(4, 8, 1.0, 0.5)     # 4 rows of 8 sites
    if name == "heis_ring32":
        return models.heisenberg_chain(32)
    if name == "heis_ring20":
        return models.heisenberg_chain(20)
    if name == "hubbard16":
        return models.hubbard_chain(16, 1.0, 4.0)
    if name == "tj4x5":
        return models.tj_model(4, 5, 1.0, 0.4, periodic=False)
    raise ValueError(name)


def make_hamiltonian(name, split_bits=-1):
    import qkrylov_amd as qa
    w = WORKLOADS[name]
    terms = build_terms(name)
    if w["kind"] == "spin":
        basis = qa.SpinHalfBasis(w["n"], sz2=w["sz2"])
    elif w["kind"] == "hubbard":
        basis = qa.HubbardBasis(w["n"], nup=8, ndn=8)
    else:
        basis = qa.TJBasis(w["n"], nup=9, ndn=9)
    return qa.MatrixFreeHamiltonian(basis, None, qa.OpSum(terms), split_bits=split_bits), terms


def oracle_sector(name):
    import oracle_lib as O
    w = WORKLOADS[name]
    if w["kind"] == "spin":
        return O.SPIN, O.sector(use_sz=1, sz2=w["sz2"])
    if w["kind"] == "hubbard":
        return O.HUBBARD, O.sector(use_nup=1, nup=8, use_ndn=1, ndn=8)
    return O.TJ, O.sector(use_nup=1, nup=9, use_ndn=1, ndn=9)


def cpu_baseline(name, terms, D, budget_s):
    """Reference algorithm (oracle): Phase-1 CSR build of a contiguous row
    sample from the middle of the basis, then the complex gather SpMV over
    those rows, x spanning the full basis. Returns rows/s as 16 B/row GB/s."""
    import ctypes

    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    kind, sec = oracle_sector(name)
    w = WORKLOADS[name]
    m = O.Model(kind, w["n"], sec, kind, terms, enumerate=False)
    # half a million rows: the sample's CSR (about 2 GB at C2) is far larger
    # than the host caches, as the full matrix would be
    nrows = min(D, 1 << 19)
    r0 = (D - nrows) // 2
    t0 = time.perf_counter()
    nnz = m.L.qo_build_csr_rows(m.h, r0, nrows)
    t_build = time.perf_counter() - t0
    cols = np.zeros(nnz, dtype=np.int64)
    m.L.qo_csr_export(m.h, None, cols.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), None, None)
    x = np.zeros(2 * D)          # complex, lazily committed
    rng = np.random.default_rng(1)
    x[2 * cols] = rng.uniform(-1, 1, nnz)
    x[2 * cols + 1] = rng.uniform(-1, 1, nnz)
    y = np.zeros(2 * D)
    px = x.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    py = y.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    m.L.qo_apply_rows(m.h, px, py, r0, nrows)   # warm
    reps = 0
    t0 = time.perf_counter()
    while True:
        m.L.qo_apply_rows(m.h, px, py, r0, nrows)
        reps += 1
        el = time.perf_counter() - t0
        if el > budget_s:
            break
    rate_rows = reps * nrows / el
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(value=round(rate_rows * 16 / 1e9, 4), unit="GB/s (16 B per row equivalent)", cores=threads,
                kind="port",
                sample=(f"oracle CSR gather SpMV (complex FP64, reference algorithm) over rows [{r0}, {r0 + nrows}) "
                        f"of D={D}, nnz={nnz}, {reps} passes in {el:.1f} s; CSR build of the sample {t_build:.1f} s"),
                rows_per_s=round(rate_rows, 1))


def ftlm_leg(args, with_cf=True):
    """Batched FTLM + S^zz(q, omega) on the t-J 4x5 open lattice (2 holes):
    R = 64 Rademacher vectors, M = 100 steps, 64 temperatures T = 0.05..5,
    one multi-vector recurrence vs the same 64 vectors run one at a time;
    then the ground state and the 20 plane-wave S^z_q seeds through the
    structure-factor driver (one 40-column block continued fraction, 100
    steps, eta = 0.05)."""
    import numpy as np

    import qkrylov_amd as qa
    from qkrylov_amd import models
    R, M = args.ftlm_R, 100
    T = np.linspace(0.05, 5.0, 64)
    betas = 1.0 / T
    H, terms = make_hamiltonian("tj4x5")
    D = H.dimension
    tune = None
    Hs = H
    if args.autotune:
        # each mode runs its own best configuration: one handle tuned for
        # the R-column block, one for the single-vector runs
        Hs, _ = make_hamiltonian("tj4x5")
        tune = {"block": H.autotune(nvec=R, allow_split=True), "single": Hs.autotune(nvec=1, allow_split=True)}
        log(f"[bench-ftlm] autotune: {tune}")
    log(f"[bench-ftlm] D={D} R={R} M={M}")
    b = qa.ftlm_batched(H, R, M, betas, seed=42)
    log(f"[bench-ftlm] batched: {b['seconds']:.3f}s")
    seq = None
    if not args.no_sequential:
        seq = qa.ftlm_batched(Hs, R, M, betas, seed=42, sequential=True)
        log(f"[bench-ftlm] sequential: {seq['seconds']:.3f}s")
    rel = None
    if seq is not None:
        rel = max(float(np.max(np.abs(b[k] - seq[k]) / np.maximum(np.abs(seq[k]), 1e-300))) for k in ("Z", "E", "Cv"))
    out = {
        "workload": WORKLOADS["tj4x5"]["desc"] + ", FTLM R=64 Rademacher vectors, M=100, 64 T in 0.05..5",
        "D": D, "R": R, "M": M, "temperatures": 64, "NL": H.info["NL"], "tuning": H.tuning(), "autotune": tune,
        "solves_per_s": round(R / b["seconds"], 3),
        "batched_seconds": round(b["seconds"], 4),
        "sequential_seconds": round(seq["seconds"], 4) if seq else None,
        "speedup_vs_sequential": round(seq["seconds"] / b["seconds"], 3) if seq else None,
        "max_rel_diff_batched_vs_sequential": rel,
        "Z_E_Cv_at_T=0.05,1,5": [[float(b[k][i]) for k in ("Z", "E", "Cv")] for i in (0, 12, 63)],
        "roofline": {"bound": "hbm", "achieved": round(16.0 * D * R * M / b["seconds"] / 1e9, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(16.0 * D * R * M / b["seconds"] / 1e9 / HBM_PEAK_GBS, 5), "traffic": None,
                     "note": "block recurrence wall time incl. reductions; 16 B x R per row per step"},
    }
    if with_cf:
        g = qa.lanczos_ground_state(Hs, maxiter=300, tol=1e-10, reorthogonalize=False, return_state=True)
        psi = g.eigenvector
        ops = [qa.OpSum(models.szq_terms(4, 5, q)) for q in models.lattice_momenta(4, 5)]
        omega = np.linspace(0.0, 4.0, 201)
        t0 = time.perf_counter()
        sf = qa.structure_factor(H, psi, ops, omega, g.energy, eta=0.05, n_iter=100)
        t_sf = time.perf_counter() - t0
        log(f"[bench-ftlm] E0={g.energy:.10f} ({g.iterations} it), S(q, omega) 20 q in {sf.seconds:.3f}s "
            f"({t_sf:.3f}s with seeds and spectra)")
        out["ground_state"] = {"energy": g.energy, "iterations": g.iterations,
                               "seconds": round(g.report["loop_seconds"], 4)}
        out["szz_structure_factor"] = {"q_points": 20, "steps": 100, "eta": 0.05, "omega_points": 201,
                                       "block_cf_seconds": round(sf.seconds, 4),
                                       "driver_seconds": round(t_sf, 4),
                                       "S_peak_q_pi": float(np.max(sf.spectrum[len(ops) // 2]))}
    return out


def bench_ftlm(args):
    o = ftlm_leg(args)
    res = {
        "metric": "batched FTLM solves/s (R=64 vectors x M=100 steps, FP64)",
        "value": o["solves_per_s"], "unit": "random-vector solves/s", "n_gpus": 1, "steps": o["M"],
        "warmup": 0, "ms_per_step": round(1e3 * o["batched_seconds"] / o["M"], 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (Rademacher vectors from the counter RNG)",
        "config": {"workload": o["workload"], "D": o["D"], "R": o["R"], "M": o["M"], "temperatures": 64,
                   "NL": o["NL"], "tuning": o["tuning"], "autotune": o["autotune"]},
    }
    res.update({k: v for k, v in o.items() if k not in ("workload", "D", "R", "M", "NL", "tuning", "autotune")})
    print(json.dumps(res), flush=True)


def serve_leg(args):
    """Serving stream of XXZ-ring solves: Poisson arrivals at 50 req/s,
    L in 12..28, Delta in [0.5, 1.5], |Sz| <= 2, 70 % ground states
    (tol 1e-10), 30 % ground state + S^zz(pi, omega) (100 steps); then the
    same requests all present at t = 0 (capacity), batched by sector and,
    for comparison, one request per batch."""
    from qkrylov_amd import serving
    reqs = serving.make_trace(args.serve_n, rate=50.0, seed=2024)
    out, tot = serving.run(reqs, realtime=True, max_active=args.serve_active)
    log(f"[bench-serve] {len(reqs)} requests: {tot}")
    sat_out, sat = serving.run(reqs, realtime=False, max_active=args.serve_active)
    log(f"[bench-serve] saturated: {sat}")
    _, sat1 = serving.run(reqs, realtime=False, max_active=args.serve_active, max_batch=1)
    log(f"[bench-serve] saturated, unbatched: {sat1}")

    def mean_batch(res):
        return round(sum(r["batch"] for r in res) / max(len(res), 1), 3)
    return {
        "workload": "XXZ ring serving stream, 50 req/s Poisson, L 12..28, Delta 0.5..1.5, |Sz| <= 2, "
                    "70% ground state / 30% + S^zz CF (100 steps)",
        "requests": len(reqs), "max_active": args.serve_active,
        "solves_per_s": round(tot["solves_per_s"], 3),
        "p50_latency_s": round(tot["p50_s"], 5), "p99_latency_s": round(tot["p99_s"], 5),
        "wall_s": round(tot["wall_s"], 3),
        "converged": int(sum(o["converged"] for o in out)),
        "mean_batch": mean_batch(out),
        "saturated": {"solves_per_s": round(float(sat["solves_per_s"]), 3), "p50_latency_s": round(float(sat["p50_s"]), 5),
                      "p99_latency_s": round(float(sat["p99_s"]), 5), "wall_s": round(float(sat["wall_s"]), 3),
                      "mean_batch": mean_batch(sat_out)},
        "saturated_unbatched": {"solves_per_s": round(float(sat1["solves_per_s"]), 3),
                                "p50_latency_s": round(float(sat1["p50_s"]), 5),
                                "p99_latency_s": round(float(sat1["p99_s"]), 5), "wall_s": round(float(sat1["wall_s"]), 3)},
    }


def bench_gather_probe(args):
    """Diagnostic, not a reference metric: the rate at which waves issue
    coalesced gathers of random 64-lane segments (8 or 16 B per lane) from
    device arrays of 2 MB to 4.6 GB -- the instruction rate that bounds the
    apply's gathers whatever the array size."""
    import ctypes
    from qkrylov_amd import _lib
    L = _lib.load()
    rows = []
    for nbytes in (2 << 20, 16 << 20, 256 << 20, 4600 << 20):
        for w in (8, 16):
            ms, g = ctypes.c_double(), ctypes.c_double()
            _lib.check(L.qkrylov_amd_gather_probe(nbytes, w, max(args.steps, 1), ctypes.byref(ms), ctypes.byref(g)),
                       "gather probe")
            rate = g.value / ms.value
            rows.append({"array_bytes": nbytes, "bytes_per_lane": w, "ms_per_launch": round(ms.value, 4),
                         "wave_gathers_per_launch": int(g.value), "wave_gathers_per_ms": round(rate, 0),
                         "GB_per_s": round(rate * 64 * w / 1e6, 1)})
            log(f"[bench] gather probe {nbytes >> 20} MiB, {w} B/lane: {ms.value:.4f} ms, "
                f"{rate / 1e6:.2f} M wave gathers/ms, {rate * 64 * w / 1e6:.0f} GB/s")
    best = max(r["wave_gathers_per_ms"] for r in rows)
    print(json.dumps({"metric": "wave gathers per ms (diagnostic)", "value": best, "unit": "wave gathers/ms",
                      "n_gpus": 1, "steps": args.steps, "warmup": 1, "higher_is_better": True, "scaling": "weak",
                      "vs_baseline": None, "dtype": "f64", "data": "synthetic random segments",
                      "config": {"workload": "gather_probe", "grid": "2048 x 512 threads, 64 rounds of 16 gathers"},
                      "rows": rows}), flush=True)


def bench_serve(args):
    o = serve_leg(args)
    res = {
        "metric": "serving solves/s with p50/p99 latency (FP64)",
        "value": o["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": o["requests"], "warmup": 0,
        "ms_per_step": round(1e3 * o["wall_s"] / o["requests"], 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic seeded Poisson trace",
        "config": {"workload": o["workload"], "requests": o["requests"], "max_active": o["max_active"]},
    }
    res.update({k: v for k, v in o.items() if k not in ("workload", "requests", "max_active")})
    print(json.dumps(res), flush=True)


def lanczos_c1_leg():
    """Plain Lanczos ground state of the Heisenberg ring L = 20 (D = 184,756):
    start uniform[-1,1) from seed 1234, tol 1e-12 on |E_k - E_{k-1}|, at most
    300 iterations; without reorthogonalisation (graph replay) and with the
    reference's full reorthogonalisation. Energies against the Bethe ansatz."""
    import qkrylov_amd as qa
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import textbook as TB
    H, _ = make_hamiltonian("heis_ring20")
    e_bethe = float(TB.bethe_xxx_ring(20))
    res = {"workload": WORKLOADS["heis_ring20"]["desc"] + ", tol 1e-12, <= 300 iterations", "D": H.dimension,
           "bethe_energy": e_bethe}
    for name, reorth in (("plain", False), ("reorthogonalized", True)):
        qa.lanczos_ground_state(H, maxiter=300, tol=1e-12, reorthogonalize=reorth, return_state=False)   # warm
        r = qa.lanczos_ground_state(H, maxiter=300, tol=1e-12, reorthogonalize=reorth, return_state=False)
        rep = r.report
        res[name] = {"iterations": r.iterations, "converged": bool(r.converged), "energy": r.energy,
                     "rel_err_vs_bethe": abs(r.energy - e_bethe) / abs(e_bethe),
                     "loop_seconds": round(rep["loop_seconds"], 5), "steps_run": rep["steps_run"],
                     "iters_per_s": round(r.iterations / rep["loop_seconds"], 1)}
        log(f"[bench-c1] {name}: {r.iterations} it, E={r.energy:.12f} (Bethe {e_bethe:.12f}), "
            f"{res[name]['iters_per_s']} it/s")
    return res


def cpu_lanczos_c1(budget_s=20.0):
    """The oracle's restatement of the reference's lanczos_ground_state
    (complex CSR gather SpMV + classical Gram-Schmidt against every stored
    vector + tridiagonal QR each step) on the full Heisenberg ring L = 20,
    tol 1e-12, at most 300 iterations, on this host's cores."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    terms = build_terms("heis_ring20")
    kind, sec = oracle_sector("heis_ring20")
    m = O.Model(kind, 20, sec, kind, terms, enumerate=False)
    t0 = time.perf_counter()
    m.build_csr()
    t_build = time.perf_counter() - t0
    start = O.start_vector(m.dim, 1234) + 0j
    t0 = time.perf_counter()
    r = m.lanczos(300, 1e-12, start=start, reorth=True)
    el = time.perf_counter() - t0
    return {"iterations": r["iterations"], "energy": r["energy"], "seconds": round(el, 3),
            "iters_per_s": round(r["iterations"] / el, 2), "csr_build_seconds": round(t_build, 3),
            "algorithm": "oracle restatement of src/solvers/lanczos.cpp:101-202 (CGS reorthogonalisation)"}


def cpu_ftlm_cf_tj4x4(n_vec=4, n_steps=100):
    """The oracle's restatement of the reference's FTLM (ftlm.cpp:122-170:
    one three-term recurrence per random vector, no reorthogonalisation,
    one inverse temperature per call) and continued fraction
    (dynamics.cpp:12-59) on the scaled C4 instance the reference can build,
    t-J 4x4 open, 2 holes (D = 411,840), on this host's cores: n_vec
    Rademacher vectors x n_steps steps, and one 100-step fraction."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import oracle_lib as O
    from qkrylov_amd import models
    m = O.Model(O.TJ, 16, O.sector(use_nup=1, nup=7, use_ndn=1, ndn=7), O.TJ, models.tj_model(4, 4, 1.0, 0.4),
                enumerate=False)
    t0 = time.perf_counter()
    m.build_csr()
    t_build = time.perf_counter() - t0
    rng = np.random.default_rng(42)
    rv = np.where(rng.integers(0, 2, (n_vec, m.dim)) == 1, 1.0, -1.0) + 0j
    t0 = time.perf_counter()
    Z, E, Cv = m.ftlm(1.0, rv, n_steps)
    t_ftlm = time.perf_counter() - t0
    phi = rng.uniform(-1, 1, m.dim) + 0j
    t0 = time.perf_counter()
    al, be, _ = m.cf(phi, 100)
    t_cf = time.perf_counter() - t0
    return {"workload": "t-J 4x4 open, t=1 J=0.4, 7 up 7 dn (2 holes), the scaled C4 instance", "D": m.dim,
            "csr_build_seconds": round(t_build, 3),
            "ftlm": {"vectors": n_vec, "steps": n_steps, "seconds": round(t_ftlm, 3),
                     "seconds_per_vector": round(t_ftlm / n_vec, 4), "E_at_beta_1": E},
            "cf_100_steps_seconds": round(t_cf, 3),
            "algorithm": "oracle restatements of src/solvers/ftlm.cpp:24-53,122-170 and dynamics.cpp:12-59"}


def host_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(),
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS"), "OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND"),
            "OMP_PLACES": os.environ.get("OMP_PLACES")}


def stream_copy_gbs(qa, L, gib=2.0, reps=5):
    """Measured STREAM-copy rate of this GPU: k_copy of `gib` GiB (read +
    write counted), synchronous calls timed on the host, best of `reps`."""
    from qkrylov_amd import _lib
    n = int(gib * (1 << 30) / 8)
    a, b = qa.DeviceArray(n), qa.DeviceArray(n)
    _lib.check(L.qkrylov_amd_fill_random(a.ptr, n, 5, 0), "fill")
    _lib.check(L.qkrylov_amd_copy(b.ptr, a.ptr, n), "copy")
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        _lib.check(L.qkrylov_amd_copy(b.ptr, a.ptr, n), "copy")
        best = min(best, time.perf_counter() - t0)
    a.free()
    b.free()
    return 2 * n * 8 / best / 1e9


def config_flags(args):
    """Command-line flags that reproduce this run's split and launch tuning."""
    f = ["--split-bits", str(args.split_bits), "--nvec", str(args.nvec), "--apply", args.apply]
    if args.sched is not None:
        f += ["--sched", str(args.sched)]
    if args.block_waves is not None:
        f += ["--block-waves", str(args.block_waves)]
    if args.apply_waves is not None:
        f += ["--apply-waves", str(args.apply_waves)]
    if args.block_rows is not None:
        f += ["--block-rows", str(args.block_rows)]
    return f


def measure_traffic(args, flags):
    """HBM-side bytes per apply launch from two child runs of this script
    under `rocprofv3 --pmc` (FETCH_SIZE, then WRITE_SIZE) with the given
    configuration flags, each also running one plain copy kernel of known
    size whose counter reading gives the scale factor.  Returns (bytes per
    launch, detail dict) or (None, reason)."""
    import shutil
    import subprocess
    import tempfile
    from qkrylov_amd import profiling
    prof = shutil.which("rocprofv3")
    if prof is None:
        return None, "rocprofv3 not on PATH"
    calib_gib = 1.0
    tmp = tempfile.mkdtemp(prefix="qk_pmc_", dir="/tmp")
    env = dict(os.environ, TMPDIR="/tmp")
    dbs = {}
    try:
        for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
            out = os.path.join(tmp, ctr)
            cmd = ["timeout", "-s", "KILL", "300", prof, "--pmc", ctr, "-d", out, "-o", "run", "--",
                   sys.executable, os.path.abspath(__file__), "--workload", args.workload] + flags + [
                   "--steps", "2", "--warmup", "0", "--no-lanczos",
                   "--no-cpu", "--no-traffic", "--no-stream", "--no-extras", "--no-autotune",
                   "--copy-calib-gib", str(calib_gib)]
            log(f"[bench] traffic pass {ctr}: {' '.join(cmd[3:])}")
            r = subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE,
                               text=True)
            db = os.path.join(out, "run_results.db")
            if r.returncode != 0 or not os.path.exists(db):
                return None, f"rocprofv3 --pmc {ctr} exited {r.returncode}: {r.stderr.strip()[-300:]}"
            dbs[ctr] = db
        mode = flags[flags.index("--apply") + 1] if "--apply" in flags else "auto"
        res = None
        if mode != "public" and args.nvec == 1:
            try:
                res = profiling.calibrated_traffic_split(dbs["FETCH_SIZE"], dbs["WRITE_SIZE"], calib_gib * (1 << 30))
            except RuntimeError:
                if mode == "split":
                    raise
        if res is None:
            res = profiling.calibrated_traffic(dbs["FETCH_SIZE"], dbs["WRITE_SIZE"], calib_gib * (1 << 30),
                                                 "k_apply_block<" if args.nvec > 1 else "k_apply<")
        return res["bytes_per_dispatch"], res
    except Exception as e:          # a counter run that cannot be read leaves traffic unmeasured
        return None, f"{type(e).__name__}: {e}"
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="j1j2_4x8",
                    choices=sorted(WORKLOADS) + ["ftlm_tj4x5", "serve_xxz", "gather_probe", "lanczos_c1"])
    ap.add_argument("--split-bits", type=int, default=-1,
                    help="low state bits ranked per lane (-1: the engine's cost model)")
    ap.add_argument("--ftlm-R", type=int, default=64)
    ap.add_argument("--no-sequential", action="store_true")
    ap.add_argument("--serve-n", type=int, default=300)
    ap.add_argument("--serve-active", type=int, default=32)
    ap.add_argument("--lanczos-iters", type=int, default=None)
    ap.add_argument("--no-lanczos", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--apply", choices=["auto", "split", "public"], default="auto",
                    help="timed apply: the single-pass apply in the public row order (what the Lanczos loop runs; "
                         "'auto') or the two-pass split apply in its relabelled row order ('split', real spin "
                         "models that admit it)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the C1 Lanczos, C4 batched FTLM / S(q, omega) and C5 serving legs of the default run")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--nvec", type=int, default=1,
                    help="columns of the applied block (1: single-vector kernel; >1: row-major block kernel)")
    ap.add_argument("--sched", type=int, default=None,
                    help="tile schedule (0: one segment per XCD, 1: one front, 2: one front of XCD-contiguous "
                         "chunks; default: the engine's choice)")
    ap.add_argument("--block-rows", type=int, default=None,
                    help="rows per work item of the block apply (64, 32, 16)")
    ap.add_argument("--block-waves", type=int, default=None, help="waves per workgroup of the block apply (16, 8, 4)")
    ap.add_argument("--apply-waves", type=int, default=None, help="waves per workgroup of the apply (4, 8, 16)")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="skip the autotuner (which times the candidate launch tunings and lo-bit splits before the "
                         "timed region and keeps the fastest)")
    ap.add_argument("--no-stream", action="store_true", help="skip the STREAM-copy measurement")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 --pmc child runs that measure HBM bytes per apply")
    ap.add_argument("--copy-calib-gib", type=float, default=0.0,
                    help="also run one plain copy kernel of this many GiB (counter calibration)")
    args = ap.parse_args()
    # the reference's recommended OpenMP placement for its CPU path
    # (bindings/julia/README.md:343-347), used by the oracle's timed runs
    os.environ.setdefault("OMP_PROC_BIND", "spread")
    os.environ.setdefault("OMP_PLACES", "threads")
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU per process; multi-GPU runs are not part of this "
                                   "build (the apply would shard by ranges of output rows cut at block boundaries, "
                                   "with a halo of x exchanged directly between peers)"}))
        sys.exit(2)

    traffic, traffic_info = None, "not measured (--no-traffic)"
    want_traffic = not args.no_traffic and args.workload in WORKLOADS
    if want_traffic and not args.autotune:
        # the counter runs go first, before this process touches the GPU
        traffic, traffic_info = measure_traffic(args, config_flags(args))
        log(f"[bench] traffic: {traffic_info if traffic is None else f'{traffic / 1e9:.3f} GB per apply'}")

    import qkrylov_amd as qa
    from qkrylov_amd import _lib
    L = _lib.load()
    if args.workload == "ftlm_tj4x5":
        return bench_ftlm(args)
    if args.workload == "serve_xxz":
        return bench_serve(args)
    if args.workload == "gather_probe":
        return bench_gather_probe(args)
    if args.workload == "lanczos_c1":
        o = lanczos_c1_leg()
        p = o["plain"]
        print(json.dumps({"metric": "Lanczos iters/s (FP64), run to tol 1e-12", "value": p["iters_per_s"],
                          "unit": "it/s", "n_gpus": 1, "steps": p["steps_run"], "warmup": 1,
                          "ms_per_step": round(1e3 * p["loop_seconds"] / p["steps_run"], 5), "higher_is_better": True,
                          "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded start)",
                          "config": {"workload": o["workload"], "D": o["D"]}, "lanczos_c1": o}), flush=True)
        return
    w = WORKLOADS[args.workload]
    if args.copy_calib_gib > 0:
        n = int(args.copy_calib_gib * (1 << 30) / 8)
        a, b = qa.DeviceArray(n), qa.DeviceArray(n)
        _lib.check(L.qkrylov_amd_fill_random(a.ptr, n, 3, 0), "fill")
        _lib.check(L.qkrylov_amd_copy(b.ptr, a.ptr, n), "copy")
        log(f"[bench] copy calibration: {n * 8} bytes read + {n * 8} bytes written by k_copy")
        a.free()
        b.free()
    t0 = time.perf_counter()
    H, terms = make_hamiltonian(args.workload, args.split_bits)
    t_setup = time.perf_counter() - t0
    D = H.dimension
    log(f"[bench] {args.workload}: D={D} branches={H.info['n_offdiag_branches']} "
        f"diag_terms={H.info['n_diag_terms']} NL={H.info['NL']} setup={t_setup:.2f}s")

    # ---- timed applies (x, y resident in HBM; HIP events on the engine stream)
    L.qkrylov_amd_synchronize()
    t0 = time.perf_counter()
    tune = {}
    if args.sched is not None:
        tune.update(schedule=args.sched, block_schedule=args.sched)
    if args.block_waves is not None:
        tune["block_waves"] = args.block_waves
    if args.apply_waves is not None:
        tune["apply_waves"] = args.apply_waves
    if args.block_rows is not None:
        tune["block_rows"] = args.block_rows
    H.set_tuning(**tune)
    tune_report = None
    want_split = args.nvec == 1 and (args.apply == "split" or (args.apply == "auto" and not args.no_extras))
    split = H.split_info(any_size=True) if want_split else None
    if args.apply == "split" and split is None:
        raise SystemExit(f"--apply split: {args.workload} admits no split")
    if args.apply == "auto":
        # the single pass: on this image the two passes measured slower on
        # the J1-J2 4x8 (62 vs 53 ms) and the 32-site ring (25 vs 16 ms)
        args.apply = "public"
    if args.autotune:
        t0 = time.perf_counter()
        tune_report = H.autotune(nvec=args.nvec, allow_split=args.split_bits < 0)
        tune_report["seconds"] = round(time.perf_counter() - t0, 2)
        log(f"[bench] autotune: {tune_report}")
        if want_traffic:
            # counter runs as child processes with the configuration chosen here
            t = H.tuning()
            flags = ["--split-bits", str(H.info["NL"]), "--nvec", str(args.nvec), "--apply", args.apply,
                     "--apply-waves", str(t["apply_waves"]), "--block-waves", str(t["block_waves"]),
                     "--block-rows", str(t["block_rows"])]
            sched = t["block_schedule"] if args.nvec > 1 else t["schedule"]
            flags += ["--sched", str(sched)]
            traffic, traffic_info = measure_traffic(args, flags)
            log(f"[bench] traffic: {traffic_info if traffic is None else f'{traffic / 1e9:.3f} GB per apply'}")
    ms_public = ms_split = None
    if args.apply == "public" and split is not None:
        ms_split = H.time_split_apply(args.warmup, args.steps)
        log(f"[bench] two-pass split apply (relabelled row order): {ms_split:.3f} ms/apply")
    if args.apply == "split":
        if not args.no_extras:
            ms_public = H.time_apply(1, args.warmup, args.steps)
            log(f"[bench] single-pass apply (public row order): {ms_public:.3f} ms/apply")
        t0 = time.perf_counter()
        ms = H.time_split_apply(args.warmup, args.steps)
    else:
        ms = H.time_apply(args.nvec, args.warmup, args.steps)
    L.qkrylov_amd_synchronize()
    wall = time.perf_counter() - t0
    alg_bytes = 16.0 * D * args.nvec
    gbs = alg_bytes / (ms * 1e-3) / 1e9
    log(f"[bench] apply ({args.apply}): {ms:.3f} ms/apply -> {gbs:.1f} GB/s algorithmic "
        f"({wall:.2f}s wall incl. warmup)")

    out = {
        "metric": "H.x apply HBM throughput (algorithmic bytes, FP64)",
        "value": round(gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (x uniform[-1,1) from the counter RNG)",
        "config": {"workload": w["desc"], "D": D, "terms": len(terms), "columns": args.nvec, "NL": H.info["NL"],
                   "tuning": H.tuning(), "autotune": tune_report,
                   "offdiag_branches": H.info["n_offdiag_branches"], "setup_s": round(t_setup, 3),
                   "apply": ("two-pass split apply in the relabelled basis' row order (the Lanczos loop's apply), "
                             f"site groups NL/nB/nC = {split}") if args.apply == "split"
                            else "single-pass apply in the public row order"},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 5),
                     "traffic": None if traffic is None else int(traffic),
                     "algorithmic_bytes_per_launch": int(alg_bytes)},
    }

    if args.nvec == 1:
        # diagnostics (not the bar): off-diagonal entries per row from 4 x 8192
        # sampled rows (the device's own row listing), the gather-inclusive
        # bytes 8 B x (2 + entries) per row, and branch evaluations per second
        # (every row tests every off-diagonal branch)
        import numpy as np
        n_off, n_rows = 0, 0
        for f in (0.11, 0.37, 0.61, 0.89):
            r0 = min(int(f * D), max(D - 8192, 0))
            for _, v in H.row_entries(r0, min(8192, D - r0)):
                n_off += int(np.count_nonzero(v[1:]))   # the row's diagonal entry comes first
                n_rows += 1
        nbar = n_off / max(n_rows, 1)
        out["diagnostics"] = {
            "offdiag_entries_per_row": round(nbar, 3),
            "gather_inclusive_bytes_per_launch": int(8 * D * (2 + nbar)),
            "gather_inclusive_GBs": round(8 * D * (2 + nbar) / (ms * 1e-3) / 1e9, 1),
            "branch_evaluations_per_s": round(D * H.info["n_offdiag_branches"] / (ms * 1e-3), 1),
        }
        log(f"[bench] diagnostics: {out['diagnostics']}")
    if ms_split is not None:
        out["apply_split_two_pass"] = {"ms": round(ms_split, 4),
                                       "GBs_alg": round(alg_bytes / (ms_split * 1e-3) / 1e9, 2),
                                       "site_groups_NL_nB_nC": list(split),
                                       "note": "the same apply as two passes over the relabelled basis "
                                               "(qkrylov_hamiltonian_time_split_apply); not the default, slower here"}
    if ms_public is not None:
        out["apply_public_order"] = {"ms": round(ms_public, 4),
                                     "GBs_alg": round(alg_bytes / (ms_public * 1e-3) / 1e9, 2),
                                     "note": "single-pass apply on vectors in the public (reference) row order, "
                                             "the kernel of qkrylov_hamiltonian_apply_device_f64"}
    if not args.no_stream:
        copy_gbs = stream_copy_gbs(qa, L)
        out["roofline"]["stream_copy_GBs"] = round(copy_gbs, 1)
        out["roofline"]["frac_of_stream_copy"] = round(gbs / copy_gbs, 5)
        log(f"[bench] stream copy: {copy_gbs:.1f} GB/s (read + write of 2 GiB, same box)")

    if traffic is not None:
        out["roofline"]["traffic_detail"] = {
            "source": "rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE (KiB per dispatch, mean over the apply launches)",
            "fetch_factor": round(traffic_info["FETCH_SIZE"]["factor"], 4),
            "write_factor": round(traffic_info["WRITE_SIZE"]["factor"], 4),
            "fetch_bytes": int(traffic_info["FETCH_SIZE"]["bytes_per_dispatch"]),
            "write_bytes": int(traffic_info["WRITE_SIZE"]["bytes_per_dispatch"]),
            "traffic_over_algorithmic": round(traffic / alg_bytes, 2)}
        if "pass_bytes" in traffic_info["FETCH_SIZE"]:
            out["roofline"]["traffic_detail"]["per_pass_bytes"] = [
                int(traffic_info["FETCH_SIZE"]["pass_bytes"][i] + traffic_info["WRITE_SIZE"]["pass_bytes"][i])
                for i in range(2)]
    else:
        out["roofline"]["traffic_detail"] = {"source": traffic_info}

    if not args.no_lanczos and args.nvec == 1:
        iters = args.lanczos_iters or w["lanczos"]
        r = qa.lanczos_ground_state(H, maxiter=iters, reorthogonalize=False, fixed_steps=True, seed=1234,
                                   return_state=False)
        ips = r.report["steps_run"] / r.report["loop_seconds"]
        log(f"[bench] lanczos: {r.report['steps_run']} steps in {r.report['loop_seconds']:.3f}s "
            f"-> {ips:.2f} it/s, E({r.iterations})={r.energy:.12f}")
        out["lanczos"] = {"iterations": r.report["steps_run"], "iters_per_s": round(ips, 3),
                          "seconds": round(r.report["loop_seconds"], 4), "energy": r.energy,
                          "reorthogonalize": False,
                          "bytes_per_iter_alg": int(48 * D),
                          "achieved_GBs_alg": round(48 * D * ips / 1e9, 2)}

    # the rest of the metric on its own workloads: C1 Lanczos it/s, C4
    # batched FTLM + S(q, omega), C5 serving (each with its own setup,
    # outside the apply's timed region)
    if not args.no_extras and args.nvec == 1:
        del H
        out["lanczos_c1"] = lanczos_c1_leg()
        out["ftlm_c4"] = ftlm_leg(args)
        out["serving_c5"] = serve_leg(args)

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.workload, terms, D, args.cpu_seconds)
        out["cpu_baseline"]["host"] = host_info()
        log(f"[bench] cpu baseline: {out['cpu_baseline']['value']} GB/s-eq on {out['cpu_baseline']['cores']} threads")
        if not args.no_extras:
            out["cpu_baseline"]["lanczos_c1_reference"] = cpu_lanczos_c1()
            log(f"[bench] cpu reference Lanczos (C1): {out['cpu_baseline']['lanczos_c1_reference']}")
            out["cpu_baseline"]["ftlm_cf_tj4x4_reference"] = cpu_ftlm_cf_tj4x4()
            log(f"[bench] cpu reference FTLM / CF (t-J 4x4): {out['cpu_baseline']['ftlm_cf_tj4x4_reference']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
