"""Serving: a seeded Poisson trace of XXZ-ring requests (ground state, or
ground state + S^zz(pi, omega) continued fraction) run through the engine's
continuous-batching loop (qkrylov_serve_run)."""
import ctypes
from math import comb

import numpy as np

from . import _lib


def xxz_dim(L, sz2):
    if (L + sz2) % 2:
        return 0
    return comb(L, (L + sz2) // 2)


def make_trace(n, rate=50.0, L_range=(12, 28), delta_range=(0.5, 1.5), max_abs_sz=2, cf_fraction=0.3, tol=1e-10,
               cf_steps=100, maxiter=300, seed=2024):
    """Poisson arrivals (mean `rate` per second); L uniform; Delta uniform;
    a random allowed sector with |Sz| <= max_abs_sz; 30 % of requests also
    ask for the S^zz continued fraction."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    reqs = []
    for i in range(n):
        L = int(rng.integers(L_range[0], L_range[1] + 1))
        allowed = [s for s in range(-2 * max_abs_sz, 2 * max_abs_sz + 1) if (s + L) % 2 == 0]
        sz2 = int(rng.choice(allowed))
        reqs.append(dict(L=L, sz2=sz2, delta=float(rng.uniform(*delta_range)),
                         kind=1 if rng.uniform() < cf_fraction else 0, arrival=float(t[i]), tol=tol,
                         cf_steps=cf_steps, maxiter=maxiter))
    return reqs


def run(reqs, realtime=True, max_active=32, hbm_budget_bytes=0, cf_cap=None, want_ground_states=False,
        max_batch=0):
    """Runs the trace; with want_ground_states every kind-1 result also
    carries "ground_state", the normalised vector its continued fraction
    started from. Requests of one (L, Sz) sector that wait together run as
    the columns of one batch of 1, 2, 4 or 8, at most max_batch (0: the
    library's default of 4; 1: every request alone); each result's "batch"
    says how many shared its batch."""
    n = len(reqs)
    cf_cap = cf_cap or max([r["cf_steps"] for r in reqs] + [1])
    arr = (_lib.ServeRequest * n)()
    for i, r in enumerate(reqs):
        arr[i] = _lib.ServeRequest(r["L"], r["sz2"], r["delta"], r["kind"], r["arrival"], r["tol"], r["cf_steps"],
                                   r["maxiter"])
    res = (_lib.ServeResult * n)()
    cfa = np.zeros(n * cf_cap)
    cfb = np.zeros(n * cf_cap)
    tot = np.zeros(4)
    gs = [np.zeros(xxz_dim(r["L"], r["sz2"])) if (want_ground_states and r["kind"] == 1) else None for r in reqs]
    ptrs = (ctypes.c_void_p * n)(*[g.ctypes.data if g is not None else None for g in gs])
    opt = _lib.ServeOptions(int(hbm_budget_bytes), int(max_active), 1 if realtime else 0,
                            ctypes.cast(ptrs, ctypes.POINTER(ctypes.c_void_p)) if want_ground_states else None,
                            int(max_batch))
    p = _lib.p_dbl
    _lib.check(_lib.load().qkrylov_serve_run(arr, n, ctypes.byref(opt), res, cfa.ctypes.data_as(p),
                                             cfb.ctypes.data_as(p), cf_cap, tot.ctypes.data_as(p)), "serve_run")
    out = []
    for i in range(n):
        r = res[i]
        d = {f: getattr(r, f) for f, _ in r._fields_}
        d["cf_alphas"] = cfa[i * cf_cap:i * cf_cap + r.n_cf].copy()
        d["cf_betas"] = cfb[i * cf_cap:i * cf_cap + max(r.n_cf - 1, 0)].copy()
        if gs[i] is not None:
            d["ground_state"] = gs[i]
        out.append(d)
    return out, dict(wall_s=tot[0], solves_per_s=tot[1], p50_s=tot[2], p99_s=tot[3])
