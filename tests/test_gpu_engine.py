"""Parity of the HIP engine with the oracle, through the C ABI.

Bit-exact: basis states of every row, and every row's column set and
signed values against the oracle's CSR (which itself equals the reference's
Phase-1 output on the fixtures). Floating point: applies within 1e-13
relative (only the summation order differs), eigenvalues within 1e-10
relative, FTLM / continued-fraction observables within 1e-8."""
import json
import math
import os

import numpy as np
import pytest

import oracle_lib as O
import qkrylov_amd as qa
from qkrylov_amd import models

pytestmark = pytest.mark.gpu

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ref_fixtures.json")))

BASIS_CLS = {O.SPIN: qa.SpinHalfBasis, O.FERMION: qa.FermionBasis, O.HUBBARD: qa.HubbardBasis, O.TJ: qa.TJBasis}
SITE_CLS = {O.SPIN: qa.SpinHalfSite, O.FERMION: qa.FermionSite, O.HUBBARD: qa.HubbardSite, O.TJ: qa.TJSite}


def engine_basis(kind, n, sec):
    if kind == O.SPIN:
        return qa.SpinHalfBasis(n, sz2=sec["sz2"]) if sec.get("use_sz") else qa.SpinHalfBasis(n)
    if kind == O.FERMION:
        return qa.FermionBasis(n, conserve_n=True, n=sec["n"]) if sec.get("use_n") else qa.FermionBasis(n)
    return BASIS_CLS[kind](n, conserve_nup=bool(sec.get("use_nup")), nup=sec.get("nup", 0),
                           conserve_ndn=bool(sec.get("use_ndn")), ndn=sec.get("ndn", 0))


def pair(name, split_bits=-1):
    m = FIX["models"][name]
    terms = [(complex(*c), [tuple(f) for f in fs]) for c, fs in m["terms"]]
    om = O.Model(m["bkind"], m["n"], O.sector(**m["sector"]), m["skind"], terms)
    b = engine_basis(m["bkind"], m["n"], m["sector"])
    H = qa.MatrixFreeHamiltonian(b, SITE_CLS[m["skind"]](), qa.OpSum(terms), split_bits=split_bits)
    return H, om


CASES = [n for n in sorted(FIX["models"]) if not FIX["models"][n]["sector"].get("use_ndn", 1) == 0 or True]


def merged(cols, vals):
    d = {}
    for c, v in zip(cols.tolist(), vals.tolist()):
        d[c] = d.get(c, 0.0) + v
    return {c: v for c, v in d.items() if v != 0}


@pytest.mark.parametrize("name", sorted(FIX["models"]))
@pytest.mark.parametrize("split", [-1, 2])
def test_states_and_row_entries_bit_exact(name, split):
    m = FIX["models"][name]
    if m["bkind"] in (O.HUBBARD, O.TJ) and split % 2:
        split += 1
    H, om = pair(name, split)
    D = H.dimension
    assert D == om.dim
    st = H.states(0, D)
    ref_states = O.enumerate_basis(m["bkind"], m["n"], O.sector(**m["sector"]))
    assert np.array_equal(st, ref_states)
    ro, cols, vals, _ = om.csr()
    rows = H.row_entries(0, D)
    for r in range(D):
        a = merged(*rows[r])
        b = merged(cols[ro[r]:ro[r + 1]], vals[ro[r]:ro[r + 1]])
        assert set(a) == set(b), f"row {r}: columns differ"
        for c in a:
            assert abs(a[c] - b[c]) <= 1e-15 * max(1.0, abs(b[c])), f"row {r} col {c}: {a[c]} vs {b[c]}"


@pytest.mark.parametrize("name", sorted(FIX["models"]))
def test_apply_matches_oracle(name):
    H, om = pair(name)
    rng = np.random.default_rng(5)
    x = rng.standard_normal(H.dimension) + 1j * rng.standard_normal(H.dimension)
    y = H.apply(x)
    y_ref = om.apply(x)
    assert np.linalg.norm(y - y_ref) <= 1e-13 * max(np.linalg.norm(y_ref), 1e-300)


def test_apply_is_deterministic_and_split_invariant():
    terms = models.j1j2_square(4, 4, 1.0, 0.5)
    ys = []
    for split in (-1, 6, 10, 16):
        H = qa.MatrixFreeHamiltonian(qa.SpinHalfBasis(16, sz2=0), qa.SpinHalfSite(), qa.OpSum(terms), split_bits=split)
        x = np.cos(np.arange(H.dimension) * 0.37) + 0j
        y1, y2 = H.apply(x), H.apply(x)
        # same configuration: bitwise identical
        assert np.array_equal(y1.view(np.float64), y2.view(np.float64))
        ys.append(y1)
    # different splits order the branch sums differently: equal to rounding
    for y in ys[1:]:
        assert np.linalg.norm(y - ys[0]) <= 1e-14 * np.linalg.norm(ys[0])


@pytest.mark.parametrize("kind,n,sec,terms", [
    (O.SPIN, 16, dict(use_sz=1, sz2=0), models.j1j2_square(4, 4, 1.0, 0.5)),
    (O.SPIN, 18, dict(use_sz=1, sz2=2), models.xxz_chain(18, 0.7)),
    (O.HUBBARD, 8, dict(use_nup=1, nup=4, use_ndn=1, ndn=4), models.hubbard_chain(8, 1.0, 4.0)),
    (O.HUBBARD, 8, dict(use_nup=1, nup=3, use_ndn=1, ndn=4), models.hubbard_chain(8, 1.0, 2.0, periodic=False)),
    (O.TJ, 9, dict(use_nup=1, nup=4, use_ndn=1, ndn=3), models.tj_model(3, 3, 1.0, 0.4)),
    (O.TJ, 8, dict(use_nup=1, nup=3, use_ndn=1, ndn=3), models.tj_model(2, 4, 1.0, 0.4, periodic=True)),
])
def test_larger_models_rows_and_apply(kind, n, sec, terms):
    om = O.Model(kind, n, O.sector(**sec), kind, terms)
    for split in (-1, 4):
        H = qa.MatrixFreeHamiltonian(engine_basis(kind, n, sec), SITE_CLS[kind](), qa.OpSum(terms), split_bits=split)
        assert H.dimension == om.dim
        sample = np.unique(np.linspace(0, om.dim - 1, 300).astype(np.int64))
        rows = H.row_entries(0, om.dim)
        for r in sample:
            c, v = om.row_entries(int(r))
            assert merged(*rows[r]) == pytest.approx(merged(c, v), rel=1e-15, abs=1e-15)
            assert set(merged(*rows[r])) == set(merged(c, v))
        rng = np.random.default_rng(1)
        x = rng.standard_normal(om.dim) + 0j
        assert np.linalg.norm(H.apply(x) - om.apply(x)) <= 1e-13 * np.linalg.norm(om.apply(x))


def test_lanczos_matches_oracle_trajectory_and_energy():
    terms = models.heisenberg_chain(14)
    om = O.Model(O.SPIN, 14, O.sector(use_sz=1, sz2=0), O.SPIN, terms)
    H = qa.MatrixFreeHamiltonian(qa.SpinHalfBasis(14, sz2=0), qa.SpinHalfSite(), qa.OpSum(terms))
    start = O.start_vector(om.dim, 99)
    for reorth in (True, False):
        r = qa.lanczos_ground_state(H, 200, 1e-12, reorthogonalize=reorth, start=start, return_state=True)
        rr = om.lanczos(200, 1e-12, start=start + 0j, reorth=reorth, want_vec=True)
        assert abs(r.energy - rr["energy"]) < 1e-10 * abs(rr["energy"])
        n = min(20, len(rr["alphas"]))
        assert np.allclose(r.alphas[:n], rr["alphas"][:n], rtol=1e-11, atol=1e-11)
        assert np.allclose(r.betas[:n], rr["betas"][:n], rtol=1e-10, atol=1e-11)
        if reorth:
            assert r.iterations == rr["iterations"] and r.converged == rr["converged"]
        v = r.eigenvector
        hv = H.apply(v + 0j).real
        assert np.linalg.norm(hv - r.energy * v) < 1e-6
        assert abs(np.linalg.norm(v) - 1) < 1e-12


def test_lanczos_known_answers_and_maxiter():
    H = qa.MatrixFreeHamiltonian(qa.SpinHalfBasis(4), qa.SpinHalfSite(), qa.OpSum(models.heisenberg_chain(4)))
    r = qa.lanczos_ground_state(H, maxiter=50, tol=1e-12)
    assert abs(r.energy + 2.0) < 1e-12 and r.converged
    r = qa.lanczos_ground_state(H, maxiter=2, tol=1e-15)
    assert r.iterations == 2 and not r.converged
    H4 = qa.MatrixFreeHamiltonian(qa.SpinHalfBasis(4, sz2=0), qa.SpinHalfSite(),
                                  qa.OpSum(models.heisenberg_chain(4, periodic=False)))
    assert abs(qa.lanczos_ground_state(H4).energy + 1.6160254037844386) < 1e-12


def test_lanczos_complex_hamiltonian():
    terms = models.heisenberg_chain(10) + [(0.3, [("Sy", i), ("Sy", (i + 2) % 10)]) for i in range(10)] + \
        [(0.2j, [("Sp", i), ("Sm", (i + 3) % 10)]) for i in range(10)] + \
        [(-0.2j, [("Sm", i), ("Sp", (i + 3) % 10)]) for i in range(10)]
    om = O.Model(O.SPIN, 10, O.sector(use_sz=1, sz2=0), O.SPIN, terms)
    H = qa.MatrixFreeHamiltonian(qa.SpinHalfBasis(10, sz2=0), qa.SpinHalfSite(), qa.OpSum(terms))
    assert not H.is_real
    ex = np.linalg.eigvalsh(om.dense())[0]
    r = qa.lanczos_ground_state(H, 300, 1e-13)
    assert abs(r.energy - ex) < 1e-10 * abs(ex)


def _rademacher(seed, r, D):
    L = O.lib()
    return np.array([-1.0 if (L.qo_rng_u64(seed, r * D + i) >> 63) else 1.0 for i in range(D)])


def test_ftlm_batched_vs_sequential_vs_oracle():
    terms = models.tj_model(2, 4, 1.0, 0.4, periodic=False)
    sec = dict(use_nup=1, nup=3, use_ndn=1, ndn=3)
    H = qa.MatrixFreeHamiltonian(qa.TJBasis(8, nup=3, ndn=3), qa.TJSite(), qa.OpSum(terms))
    om = O.Model(O.TJ, 8, O.sector(**sec), O.TJ, terms)
    betas = np.linspace(0.2, 20, 16)
    R, M = 12, 40
    b = qa.ftlm_batched(H, R, M, betas, seed=42)
    s = qa.ftlm_batched(H, R, M, betas, seed=42, sequential=True)
    for key in ("Z", "E", "Cv"):
        assert np.allclose(b[key], s[key], rtol=1e-8, atol=1e-10)
    rv = np.array([_rademacher(42, r, om.dim) for r in range(R)]) + 0j
    for i, beta in enumerate(betas[::5]):
        Z, E, Cv = om.ftlm(beta, rv, M)
        j = i * 5
        assert abs(b["Z"][j] - Z) < 1e-8 * Z and abs(b["E"][j] - E) < 1e-8 * max(1, abs(E))
        print(f"beta={beta:.3f}: Z {abs(b['Z'][j] - Z) / Z:.2e} E {abs(b['E'][j] - E):.2e} "
              f"Cv {abs(b['Cv'][j] - Cv):.2e} (Cv={Cv:.3e})", flush=True)
        assert abs(b["Cv"][j] - Cv) < 1e-8 * max(1, abs(Cv))


def test_ftlm_reference_entry_point():
    """ftlm() keeps the reference's complex normal vectors; with H = Sz0 Sz1
    on two sites the exact Z is 4 cosh(beta/4) times E|r|^2 / D = 2."""
    H = qa.MatrixFreeHamiltonian(qa.SpinHalfBasis(2), qa.SpinHalfSite(), qa.OpSum([(1.0, [("Sz", 0), ("Sz", 1)])]))
    r = qa.ftlm(H, 1.0, 200, 10)
    assert r.partition_function > 0 and abs(r.internal_energy) < 1.0
    assert abs(r.partition_function / (2 * 4 * math.cosh(0.25)) - 1) < 0.25


def test_cf_single_and_batched_vs_oracle():
    terms = models.heisenberg_chain(12)
    om = O.Model(O.SPIN, 12, O.sector(use_sz=1, sz2=0), O.SPIN, terms)
    H = qa.MatrixFreeHamiltonian(qa.SpinHalfBasis(12, sz2=0), qa.SpinHalfSite(), qa.OpSum(terms))
    rng = np.random.default_rng(7)
    seeds = rng.standard_normal((5, om.dim)) + 1j * rng.standard_normal((5, om.dim))
    # 25 steps: before the plain recurrence loses orthogonality on D = 924,
    # where two correct implementations still agree to rounding
    nst = 25
    batched, _ = qa.continued_fraction_batched(H, seeds, nst)
    for k in range(5):
        al, be, nrm = om.cf(seeds[k], nst)
        single = qa.continued_fraction_coeffs(H, seeds[k], nst)
        for res in (single, batched[k]):
            assert abs(res.norm_phi0 - nrm) < 1e-12 * nrm
            n = min(len(al), len(res.alphas), 30)
            assert np.allclose(res.alphas[:n], al[:n], rtol=1e-9, atol=1e-9)
            for om_ in (0.5, 1.5):
                E0 = -5.0
                s1 = qa.evaluate_spectral_function(res, om_, E0, 0.2)
                s0 = O.spectral(al, be, nrm, om_, E0, 0.2)
                assert abs(s1 - s0) < 1e-8 * max(1.0, abs(s0))


def test_cf_long_run_against_dense_resolvent():
    """Past the loss of orthogonality the coefficients are no longer
    comparable between implementations, but S(omega) still converges to the
    exact resolvent (ghost copies of converged Ritz values carry no weight)."""
    terms = models.heisenberg_chain(12)
    om = O.Model(O.SPIN, 12, O.sector(use_sz=1, sz2=0), O.SPIN, terms)
    H = qa.MatrixFreeHamiltonian(qa.SpinHalfBasis(12, sz2=0), qa.SpinHalfSite(), qa.OpSum(terms))
    Hd = om.dense()
    E0 = np.linalg.eigvalsh(Hd)[0]
    rng = np.random.default_rng(8)
    phi = rng.standard_normal(om.dim) + 1j * rng.standard_normal(om.dim)
    (res,), _ = qa.continued_fraction_batched(H, phi[None, :], 150)
    import textbook as TB
    for w in (0.2, 1.0, 2.0, 3.5):
        ex = TB.resolvent_spectrum(Hd, phi, w, E0, 0.3)
        assert abs(qa.evaluate_spectral_function(res, w, E0, 0.3) - ex) < 1e-6 * max(1.0, abs(ex))


def test_davidson_vs_dense():
    terms = models.heisenberg_chain(10)
    om = O.Model(O.SPIN, 10, O.sector(use_sz=1, sz2=0), O.SPIN, terms)
    w = np.linalg.eigvalsh(om.dense())
    H = qa.MatrixFreeHamiltonian(qa.SpinHalfBasis(10, sz2=0), qa.SpinHalfSite(), qa.OpSum(terms))
    r = qa.davidson_lowest(H, n_eig=2, max_subspace=20, tol=1e-9)
    assert r.converged
    assert abs(r.eigenvalues[0] - w[0]) < 1e-10 * abs(w[0])
    v = r.eigenvectors[0]
    assert np.linalg.norm(H.apply(v) - r.eigenvalues[0] * v) < 1e-6


def test_operator_apply_is_O_not_Odagger():
    terms = [(0.7 + 0.2j, [("Sp", 0), ("Sm", 2)]), (0.3, [("Sz", 1)])]
    b = qa.SpinHalfBasis(4, sz2=0)
    op = qa.MatrixFreeHamiltonian(b, qa.SpinHalfSite(), qa.OpSum(terms), operator=True)
    hm = qa.MatrixFreeHamiltonian(b, qa.SpinHalfSite(), qa.OpSum(terms))
    rng = np.random.default_rng(2)
    x = rng.standard_normal(6) + 1j * rng.standard_normal(6)
    y = rng.standard_normal(6) + 1j * rng.standard_normal(6)
    # <y, O x> = <O^dagger y, x>
    assert abs(np.vdot(y, op.apply(x)) - np.vdot(hm.apply(y), x)) < 1e-14


def _long_range_xxz(L):
    """All-to-all exchange with distinct couplings: every row has ~L^2/4
    off-diagonal entries, so one tile's entry list spans several LDS chunks."""
    terms = []
    for i in range(L):
        for j in range(i + 1, L):
            J = 1.0 / (j - i)
            terms += [(J, [("Sz", i), ("Sz", j)]), (0.5 * J, [("Sp", i), ("Sm", j)]),
                      (0.5 * J, [("Sm", i), ("Sp", j)])]
    return terms


BLOCK_MODELS = {
    "heis12": lambda: (qa.SpinHalfBasis(12, sz2=0), qa.SpinHalfSite(), models.heisenberg_chain(12)),
    "hubbard8": lambda: (qa.HubbardBasis(8, nup=4, ndn=3), qa.HubbardSite(), models.hubbard_chain(8, 1.0, 4.0)),
    "tj2x4": lambda: (qa.TJBasis(8, nup=3, ndn=3), qa.TJSite(), models.tj_model(2, 4, 1.0, 0.4, periodic=False)),
    "longrange16": lambda: (qa.SpinHalfBasis(16, sz2=0), qa.SpinHalfSite(), _long_range_xxz(16)),
}


@pytest.mark.parametrize("name", sorted(BLOCK_MODELS))
@pytest.mark.parametrize("split", [-1, 4])
def test_block_apply_equals_single_columns(name, split):
    """Multi-vector apply (row-major block, lane = column) against the
    single-vector kernel column by column: same entries in the same order,
    so the results are equal bit for bit; ragged column counts and a block
    wider than 64 columns (two launches) included."""
    basis, site, terms = BLOCK_MODELS[name]()
    H = qa.MatrixFreeHamiltonian(basis, site, qa.OpSum(terms), split_bits=split)
    D = H.dimension
    rng = np.random.default_rng(11)
    for waves, nvec in ((16, 2), (16, 5), (8, 40), (4, 40), (16, 64), (8, 64), (4, 64), (4, 70)):
        H.set_tuning(block_waves=waves)
        Xh = rng.uniform(-1, 1, (D, nvec))
        X, Y = qa.DeviceArray.from_host(Xh), qa.DeviceArray(D * nvec)
        print(f"block apply {name} split={split} D={D} nvec={nvec} waves={waves}", flush=True)
        H.apply_device(X.ptr, Y.ptr, nvec)
        Yh = Y.to_host().reshape(D, nvec)
        for c in (0, nvec // 2, nvec - 1):
            x1, y1 = qa.DeviceArray.from_host(Xh[:, c]), qa.DeviceArray(D)
            H.apply_device(x1.ptr, y1.ptr, 1)
            assert np.array_equal(Yh[:, c], y1.to_host()), (name, split, nvec, waves, c)
        X.free()
        Y.free()


def test_block_apply_long_range_vs_oracle():
    basis, site, terms = BLOCK_MODELS["longrange16"]()
    H = qa.MatrixFreeHamiltonian(basis, site, qa.OpSum(terms))
    om = O.Model(O.SPIN, 16, O.sector(use_sz=1, sz2=0), O.SPIN, terms)
    D = H.dimension
    Xh = np.random.default_rng(12).uniform(-1, 1, (D, 3))
    X, Y = qa.DeviceArray.from_host(Xh), qa.DeviceArray(D * 3)
    H.apply_device(X.ptr, Y.ptr, 3)
    Yh = Y.to_host().reshape(D, 3)
    for c in range(3):
        ref = om.apply(Xh[:, c] + 0j).real
        assert np.max(np.abs(Yh[:, c] - ref)) < 1e-13 * np.max(np.abs(ref))


@pytest.mark.parametrize("name", ["heis12", "tj2x4", "longrange16"])
def test_schedules_give_identical_results(name):
    """The launch tuning only changes which wave handles which rows: single
    and block applies are bitwise the same under every schedule, workgroup
    width and block work-item width."""
    basis, site, terms = BLOCK_MODELS[name]()
    H = qa.MatrixFreeHamiltonian(basis, site, qa.OpSum(terms))
    D = H.dimension
    Xh = np.random.default_rng(13).uniform(-1, 1, (D, 7))
    X = qa.DeviceArray.from_host(Xh)
    out = []
    for sched, waves, rows in ((0, 4, 64), (1, 4, 32), (0, 8, 16), (1, 16, 64), (2, 8, 16), (2, 16, 32), (2, 4, 64)):
        H.set_tuning(schedule=sched, block_schedule=sched, apply_waves=waves, block_waves=waves, block_rows=rows)
        assert H.tuning() == dict(schedule=sched, block_schedule=sched, apply_waves=waves, block_waves=waves,
                                  block_rows=rows)
        Y, y = qa.DeviceArray(D * 7), qa.DeviceArray(D)
        x1 = qa.DeviceArray.from_host(Xh[:, 0])
        H.apply_device(X.ptr, Y.ptr, 7)
        H.apply_device(x1.ptr, y.ptr, 1)
        out.append((Y.to_host(), y.to_host()))
    for o in out[1:]:
        assert np.array_equal(out[0][0], o[0]) and np.array_equal(out[0][1], o[1])
    with pytest.raises(RuntimeError):
        H.set_tuning(schedule=3)
    with pytest.raises(RuntimeError):
        H.set_tuning(apply_waves=5)
    with pytest.raises(RuntimeError):
        H.set_tuning(block_rows=8)


@pytest.mark.parametrize("name", ["heis22", "hubbard12"])
def test_lanczos_under_every_launch_shape(name):
    """The Lanczos-form apply writes one partial sum per workgroup, and its
    grid follows the occupancy of the launched instance (more tiles than one
    full residency of the chip here, for a spin and a fermion model): under
    every workgroup width and schedule the run completes and reaches the same
    energy; only the order in which the alpha and beta totals add differs."""
    basis, site, terms = {
        "heis22": lambda: (qa.SpinHalfBasis(22, sz2=0), qa.SpinHalfSite(), models.heisenberg_chain(22)),
        "hubbard12": lambda: (qa.HubbardBasis(12, nup=6, ndn=6), qa.HubbardSite(), models.hubbard_chain(12, 1.0, 4.0)),
    }[name]()
    H = qa.MatrixFreeHamiltonian(basis, site, qa.OpSum(terms))
    assert H.dimension > 600_000
    es = []
    for sched, waves in ((0, 4), (1, 8), (2, 16), (2, 8), (1, 4), (0, 16)):
        H.set_tuning(schedule=sched, apply_waves=waves)
        r = qa.lanczos_ground_state(H, 300, 1e-12)
        assert r.converged
        es.append(r.energy)
    assert max(es) - min(es) < 1e-11 * abs(es[0])


def test_autotune_keeps_results():
    """Autotuning launch shapes leaves the apply bitwise unchanged; letting it
    move the lo-bit split changes only the summation order of a row."""
    basis, site, terms = BLOCK_MODELS["longrange16"]()
    H = qa.MatrixFreeHamiltonian(basis, site, qa.OpSum(terms))
    D = H.dimension
    xh = np.random.default_rng(14).uniform(-1, 1, D)
    x, y = qa.DeviceArray.from_host(xh), qa.DeviceArray(D)
    H.apply_device(x.ptr, y.ptr, 1)
    y0 = y.to_host()
    r = H.autotune(nvec=1, allow_split=False)
    assert r["split_bits"] == H.info["NL"] and r["candidates"] >= 4 and r["ms_after"] <= r["ms_before"]
    H.apply_device(x.ptr, y.ptr, 1)
    assert np.array_equal(y.to_host(), y0)
    r = H.autotune(nvec=1, allow_split=True)
    assert r["split_bits"] == H.info["NL"]
    H.apply_device(x.ptr, y.ptr, 1)
    assert np.max(np.abs(y.to_host() - y0)) <= 1e-14 * np.max(np.abs(y0))
    Xh = np.random.default_rng(15).uniform(-1, 1, (D, 16))
    X, Y = qa.DeviceArray.from_host(Xh), qa.DeviceArray(D * 16)
    H.apply_device(X.ptr, Y.ptr, 16)
    Y0 = Y.to_host()
    H.autotune(nvec=16, allow_split=False)
    H.apply_device(X.ptr, Y.ptr, 16)
    assert np.array_equal(Y.to_host(), Y0)


@pytest.mark.parametrize("check_every", [0, 8])
def test_lanczos_graph_replay_equals_eager(check_every):
    """The captured six-step graph replays the eager launches exactly: the
    same alpha/beta bit for bit, with host reads and eager steps between the
    replays (check_every = 8 leaves the vector rotation off the graph's
    capture state, which forces a second capture)."""
    terms = models.heisenberg_chain(16)
    H = qa.MatrixFreeHamiltonian(qa.SpinHalfBasis(16, sz2=0), qa.SpinHalfSite(), qa.OpSum(terms))
    for fixed, tol in ((True, 1e-12), (False, 1e-13)):
        kw = dict(maxiter=100, tol=tol, reorthogonalize=False, fixed_steps=fixed, seed=5, check_every=check_every,
                  return_state=False)
        a = qa.lanczos_ground_state(H, graph=False, **kw)
        b = qa.lanczos_ground_state(H, graph=True, **kw)
        c = qa.lanczos_ground_state(H, graph=True, **kw)
        for r in (b, c):
            assert np.array_equal(a.alphas, r.alphas) and np.array_equal(a.betas, r.betas)
            assert a.energy == r.energy and a.iterations == r.iterations


def test_lanczos_run_ahead_changes_no_number():
    """Runs to a tolerance keep the device one convergence check ahead of
    the host: whatever the check interval, the same alpha/beta, energy and
    iteration count (the reference's convergence rule replayed on the host),
    and the oracle's energy and first coefficients on the same start vector;
    the device runs at most two check intervals past the last iteration."""
    L = 14
    terms = models.heisenberg_chain(L)
    H = qa.MatrixFreeHamiltonian(qa.SpinHalfBasis(L, sz2=0), qa.SpinHalfSite(), qa.OpSum(terms))
    om = O.Model(O.SPIN, L, O.sector(use_sz=1, sz2=0), O.SPIN, terms)
    ref = om.lanczos(200, 1e-12, start=O.start_vector(om.dim, 1234) + 0j, reorth=False)
    runs = {}
    for ce in (1, 3, 7, 12):
        r = qa.lanczos_ground_state(H, maxiter=200, tol=1e-12, reorthogonalize=False, check_every=ce,
                                    return_state=False)
        assert r.iterations <= r.report["steps_run"] <= min(200, r.iterations + 2 * ce)
        runs[ce] = r
    a = runs[1]
    for r in runs.values():
        assert np.array_equal(a.alphas, r.alphas) and np.array_equal(a.betas, r.betas)
        assert a.energy == r.energy and a.iterations == r.iterations
    assert abs(a.energy - ref["energy"]) < 1e-10 * abs(ref["energy"])
    n = min(20, len(ref["alphas"]))
    assert np.allclose(a.alphas[:n], ref["alphas"][:n], rtol=1e-11, atol=1e-11)


@pytest.mark.parametrize("kind,n,sec,terms,splits", [
    (O.SPIN, 20, dict(use_sz=1, sz2=0), models.j1j2_square(5, 4, 1.0, 0.5), (15, 16, 17, 18, 20)),
    (O.SPIN, 18, dict(use_sz=1, sz2=4), models.xxz_chain(18, 0.7), (16, 18)),
    (O.HUBBARD, 10, dict(use_nup=1, nup=5, use_ndn=1, ndn=4), models.hubbard_chain(10, 1.0, 4.0), (16, 18, 20)),
    (O.TJ, 10, dict(use_nup=1, nup=4, use_ndn=1, ndn=4), models.tj_model(2, 5, 1.0, 0.4), (16, 18, 20)),
])
def test_two_level_lo_rank(kind, n, sec, terms, splits):
    """Splits above 14 bits rank lo with two half-width LDS tables: rows
    bit-exact against the oracle, applies within rounding, block applies
    equal to the single-vector kernel."""
    om = O.Model(kind, n, O.sector(**sec), kind, terms)
    sample = np.unique(np.linspace(0, om.dim - 1, 200).astype(np.int64))
    rng = np.random.default_rng(3)
    x = rng.standard_normal(om.dim)
    y_ref = om.apply(x + 0j).real
    Xh = rng.uniform(-1, 1, (om.dim, 6))
    for split in splits:
        H = qa.MatrixFreeHamiltonian(engine_basis(kind, n, sec), SITE_CLS[kind](), qa.OpSum(terms), split_bits=split)
        assert H.info["NL"] == split and H.dimension == om.dim
        rows = H.row_entries(0, om.dim)
        for r in sample:
            c, v = om.row_entries(int(r))
            assert set(merged(*rows[r])) == set(merged(c, v)), (split, r)
            # off-diagonal values are single products (exact); the diagonal
            # sums its terms in another order than the oracle
            assert merged(*rows[r]) == pytest.approx(merged(c, v), rel=4e-15, abs=1e-15)
        xd, yd = qa.DeviceArray.from_host(x), qa.DeviceArray(om.dim)
        H.apply_device(xd.ptr, yd.ptr, 1)
        y = yd.to_host()
        assert np.linalg.norm(y - y_ref) <= 1e-13 * np.linalg.norm(y_ref), split
        X, Y = qa.DeviceArray.from_host(Xh), qa.DeviceArray(om.dim * 6)
        H.apply_device(X.ptr, Y.ptr, 6)
        Yh = Y.to_host().reshape(om.dim, 6)
        for cidx in (0, 5):
            x1, y1 = qa.DeviceArray.from_host(Xh[:, cidx]), qa.DeviceArray(om.dim)
            H.apply_device(x1.ptr, y1.ptr, 1)
            assert np.array_equal(Yh[:, cidx], y1.to_host()), (split, cidx)


def test_degenerate_sizes_follow_the_reference():
    """lanczos.cpp:109, davidson.cpp:94, dynamics.cpp:20, ftlm.cpp:131: an
    empty basis returns the default result. With one state Lanczos stops
    after one step (beta = 0 < 1e-15, converged) at E = H_00; with
    maxiter = 1 it stops after alpha_0, unconverged (lanczos.cpp:155-161)."""
    site, terms = qa.SpinHalfSite(), models.heisenberg_chain(4)
    H0 = qa.MatrixFreeHamiltonian(qa.SpinHalfBasis(4, sz2=10), site, qa.OpSum(terms))
    assert H0.dimension == 0
    r = qa.lanczos_ground_state(H0)
    assert (r.energy, r.iterations, r.converged, len(r.eigenvector)) == (0.0, 0, False, 0)
    d = qa.davidson_lowest(H0)
    assert d.iterations == 0 and not d.converged
    cf = qa.continued_fraction_coeffs(H0, np.zeros(0, dtype=np.complex128), 20)
    assert len(cf.alphas) == 0 and len(cf.betas) == 0 and cf.norm_phi0 == 0.0
    f = qa.ftlm(H0, 0.5, 4, 10)
    assert (f.beta, f.partition_function, f.internal_energy, f.specific_heat) == (0.5, 0.0, 0.0, 0.0)
    assert len(H0.apply(np.zeros(0))) == 0

    H1 = qa.MatrixFreeHamiltonian(qa.SpinHalfBasis(4, sz2=4), site, qa.OpSum(terms))
    assert H1.dimension == 1
    h00 = H1.apply(np.ones(1))[0].real
    r = qa.lanczos_ground_state(H1)
    assert r.iterations == 1 and r.converged and r.energy == h00
    assert abs(abs(r.eigenvector[0]) - 1.0) < 1e-15

    H = qa.MatrixFreeHamiltonian(qa.SpinHalfBasis(8, sz2=0), site, qa.OpSum(models.heisenberg_chain(8)))
    r = qa.lanczos_ground_state(H, maxiter=1)
    assert r.iterations == 1 and not r.converged and len(r.alphas) == 1 and r.energy == r.alphas[0]


def test_wide_states_and_size_limit():
    """64-bit states: a 40-site chain with three down spins (D = C(40, 3))
    and an 18-site Hubbard ring with 2 + 2 electrons (36 bits, signed hops),
    states bit-exact against the oracle's counting unrank and the apply
    against its CSR, for two splits; a basis of 2^32 rows or more, and a
    split leaving more than 28 high bits, are refused with a message."""
    terms = models.heisenberg_chain(40)
    om = O.Model(O.SPIN, 40, O.sector(use_sz=1, sz2=34), O.SPIN, terms, enumerate=False)
    ref = np.array([om.state(i) for i in range(om.dim)], dtype=np.uint64)
    xh = np.random.default_rng(17).uniform(-1, 1, om.dim)
    yref = om.apply(xh + 0j).real
    for split in (-1, 14):
        H = qa.MatrixFreeHamiltonian(qa.SpinHalfBasis(40, sz2=34), qa.SpinHalfSite(), qa.OpSum(terms),
                                     split_bits=split)
        assert H.dimension == om.dim == 9880
        assert np.array_equal(H.states(0, H.dimension).astype(np.uint64), ref)
        assert np.max(np.abs(H.apply(xh).real - yref)) < 1e-13 * np.max(np.abs(yref))
    # fermions on 36 bits: the periodic hops cross both 32-bit halves with
    # Jordan-Wigner signs
    terms = models.hubbard_chain(18, 1.0, 4.0)
    om = O.Model(O.HUBBARD, 18, O.sector(use_nup=1, nup=2, use_ndn=1, ndn=2), O.HUBBARD, terms, enumerate=False)
    ref = np.array([om.state(i) for i in range(om.dim)], dtype=np.uint64)
    xh = np.random.default_rng(18).uniform(-1, 1, om.dim)
    yref = om.apply(xh + 0j).real
    for split in (-1, 14):
        H = qa.MatrixFreeHamiltonian(qa.HubbardBasis(18, nup=2, ndn=2), qa.HubbardSite(), qa.OpSum(terms),
                                     split_bits=split)
        assert H.dimension == om.dim == 23409
        assert np.array_equal(H.states(0, H.dimension).astype(np.uint64), ref)
        assert np.max(np.abs(H.apply(xh).real - yref)) < 1e-13 * np.max(np.abs(yref))
    with pytest.raises(ValueError, match="2\\^32"):
        qa.MatrixFreeHamiltonian(qa.SpinHalfBasis(36, sz2=0), qa.SpinHalfSite(), qa.OpSum(models.heisenberg_chain(36)))
    with pytest.raises(ValueError, match="at most 28"):
        qa.MatrixFreeHamiltonian(qa.SpinHalfBasis(40, sz2=34), qa.SpinHalfSite(), qa.OpSum(models.heisenberg_chain(40)),
                                 split_bits=8)
