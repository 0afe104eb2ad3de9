"""Serving loop: every request gives the same result alone and in the mix
(bitwise: fixed kernels and reduction order per request), energies agree
with the oracle, admission respects the HBM budget."""
import numpy as np
import pytest

import oracle_lib as O
from qkrylov_amd import models, serving

pytestmark = pytest.mark.gpu


def small_trace():
    return serving.make_trace(14, rate=400.0, L_range=(8, 16), cf_steps=40, maxiter=200, seed=5)


def test_alone_equals_mixed_bitwise():
    reqs = small_trace()
    mixed, tot = serving.run(reqs, realtime=True, max_active=8)
    assert tot["solves_per_s"] > 0 and tot["p99_s"] >= tot["p50_s"] > 0
    for i, r in enumerate(reqs):
        alone, _ = serving.run([dict(r, arrival=0.0)], realtime=False, max_active=1)
        a, m = alone[0], mixed[i]
        assert a["energy"] == m["energy"] and a["iterations"] == m["iterations"]
        assert a["n_cf"] == m["n_cf"]
        assert np.array_equal(a["cf_alphas"], m["cf_alphas"]) and np.array_equal(a["cf_betas"], m["cf_betas"])
        assert a["cf_norm"] == m["cf_norm"]


def test_energies_and_cf_vs_oracle():
    reqs = small_trace()
    out, _ = serving.run(reqs, realtime=False, max_active=6)
    for r, o in zip(reqs, out):
        if r["L"] > 14:
            continue
        om = O.Model(O.SPIN, r["L"], O.sector(use_sz=1, sz2=r["sz2"]), O.SPIN, models.xxz_chain(r["L"], r["delta"]))
        w, V = np.linalg.eigh(om.dense())
        assert abs(o["energy"] - w[0]) < 1e-10 * abs(w[0])
        assert o["converged"] == 1
        if r["kind"] == 1 and abs(w[1] - w[0]) > 1e-6:
            psi = V[:, 0]
            st = O.enumerate_basis(O.SPIN, r["L"], O.sector(use_sz=1, sz2=r["sz2"]))
            sign = np.array([sum(((-1) ** j) * (0.5 if (int(s) >> j) & 1 else -0.5) for j in range(r["L"]))
                             for s in st]) / np.sqrt(r["L"])
            phi = sign * psi
            # the Ritz vector of a run stopped at |dE| < 1e-10 carries an
            # O(sqrt(1e-10)) vector error, so the seed agrees to ~1e-5
            assert abs(o["cf_norm"] - np.linalg.norm(phi)) < 1e-4 * np.linalg.norm(phi)
            # later coefficients amplify the seed error: compare the first three
            al, be, nrm = om.cf(phi + 0j, 3)
            assert np.allclose(o["cf_alphas"][:3], al[:3], rtol=1e-5, atol=1e-5)


def test_admission_respects_budget():
    reqs = serving.make_trace(6, rate=1000.0, L_range=(16, 16), cf_fraction=0.0, maxiter=100, seed=9)
    d = serving.xxz_dim(16, 0)
    budget = int(3.5 * d * 8 + (1 << 21) + 8 * 200)      # room for exactly one request at a time
    out, tot = serving.run(reqs, realtime=False, max_active=4, hbm_budget_bytes=budget)
    spans = sorted((o["start"], o["finish"]) for o in out)
    for (s0, f0), (s1, f1) in zip(spans, spans[1:]):
        assert s1 >= f0 - 1e-9            # never two at once


def test_invalid_requests_are_refused_with_a_message():
    reqs = serving.make_trace(2, rate=1000.0, L_range=(12, 12), cf_fraction=0.0, maxiter=50, seed=3)
    bad = [dict(r) for r in reqs]
    bad[0]["L"] = 40
    with pytest.raises(RuntimeError, match="L outside"):
        serving.run(bad, realtime=False)
    with pytest.raises(RuntimeError, match="footprint exceeds"):
        serving.run(reqs, realtime=False, hbm_budget_bytes=1024)
    for sz2 in (1, 14):   # wrong parity for 12 sites; beyond the all-up state
        bad = [dict(r) for r in reqs]
        bad[1]["sz2"] = sz2
        with pytest.raises(RuntimeError, match="sector is empty"):
            serving.run(bad, realtime=False)
    # the loop still serves after refused runs
    out, tot = serving.run(reqs, realtime=False)
    assert len(out) == 2 and all(r["iterations"] > 0 for r in out) and tot["solves_per_s"] > 0
