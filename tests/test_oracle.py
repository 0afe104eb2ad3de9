"""Pins the oracle (oracle/qk_oracle.c) before it is trusted as the checker:
  - against fixtures emitted by the reference's own compiled basis/site
    sources and its Phase-1 loop (tests/golden/ref_fixtures.json), and
    live against oracle/_ref when it was built here;
  - against the reference tests' known answers;
  - against independent textbook constructions (tests/textbook.py):
    dense Kronecker / Jordan-Wigner operators, Bethe ansatz, exact thermal
    sums and resolvents.
Runs on the CPU."""
import json
import math
import os

import numpy as np
import pytest

import oracle_lib as O
import textbook as TB
from qkrylov_amd import models

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ref_fixtures.json")))


def sec_of(d):
    return O.sector(**d)


# ------------------------------------------------------------------ bases
@pytest.mark.parametrize("case", FIX["bases"], ids=lambda c: f"k{c['kind']}n{c['n']}{c['sector']}")
def test_basis_matches_reference_fixture(case):
    st = O.enumerate_basis(case["kind"], case["n"], sec_of(case["sector"]))
    assert [int(s) for s in st] == case["states"]


@pytest.mark.parametrize("case", FIX["bases"], ids=lambda c: f"k{c['kind']}n{c['n']}{c['sector']}")
def test_counting_rank_equals_enumeration(case):
    L = O.lib()
    sec = sec_of(case["sector"])
    st = case["states"]
    assert L.qo_count_dimension(case["kind"], case["n"], sec) == len(st)
    for i, s in enumerate(st):
        assert L.qo_count_rank(case["kind"], case["n"], sec, s) == i
        assert L.qo_count_unrank(case["kind"], case["n"], sec, i) == s
    nb = case["n"] if case["kind"] in (O.SPIN, O.FERMION) else 2 * case["n"]
    ss = set(st)
    for s in range(1 << nb):
        assert bool(L.qo_count_contains(case["kind"], case["n"], sec, s)) == (s in ss)


@pytest.mark.parametrize("kind,n,sec,dim", [
    (O.SPIN, 20, dict(use_sz=1, sz2=0), math.comb(20, 10)),
    (O.SPIN, 32, dict(use_sz=1, sz2=0), 601080390),
    (O.HUBBARD, 16, dict(use_nup=1, nup=8, use_ndn=1, ndn=8), math.comb(16, 8) ** 2),
    (O.TJ, 20, dict(use_nup=1, nup=9, use_ndn=1, ndn=9), 9237800),
    (O.FERMION, 4, dict(), 16), (O.FERMION, 4, dict(use_n=1, n=2), 6), (O.TJ, 2, dict(), 9),
])
def test_closed_form_dimensions(kind, n, sec, dim):
    # reference dims 16, 6 (test_fermion_basis.cpp:12,20), 9 (test_tj_basis.cpp:14)
    assert O.lib().qo_count_dimension(kind, n, sec_of(sec)) == dim


def test_counting_rank_large_brute_windows():
    """Counting rank against brute-force ascending enumeration over windows of
    the 32-bit Sz=0 sector and the 16-site Hubbard sector."""
    L = O.lib()
    for kind, n, sec, start in [(O.SPIN, 32, dict(use_sz=1, sz2=0), (1 << 31) + 12345),
                                (O.HUBBARD, 16, dict(use_nup=1, nup=8, use_ndn=1, ndn=8), 0xAAAA5555)]:
        s_ = sec_of(sec)
        s = start
        while not L.qo_count_contains(kind, n, s_, s):
            s += 1
        r = L.qo_count_rank(kind, n, s_, s)
        k = 0
        t = s + 1
        while k < 2000:
            if L.qo_count_contains(kind, n, s_, t):
                k += 1
                assert L.qo_count_rank(kind, n, s_, t) == r + k
            t += 1


# ------------------------------------------------------------------ sites
@pytest.mark.parametrize("table", FIX["sites"], ids=lambda t: f"kind{t['kind']}")
def test_site_actions_match_reference_fixture(table):
    kind = table["kind"]
    seen = {}
    for op, site, s, v, ns, re, im in table["rows"]:
        seen[(op, site, s)] = (v, ns, complex(re, im))
    for (op, site, s), (v, ns, a) in seen.items():
        v2, ns2, a2 = O.site_apply(kind, op, site, s)
        assert v2 == v
        if v == 1:
            assert ns2 == ns and a2 == a
    # every state not listed as valid must be invalid (unless the op threw)
    ops = {r[0] for r in table["rows"]}
    for op in ops:
        for site in range(3 if kind in (O.HUBBARD, O.TJ) else 4):
            if seen.get((op, site, 0), (0,))[0] == -1:
                continue
            for s in range(1 << table["nbits"]):
                if (op, site, s) not in seen:
                    assert O.site_apply(kind, op, site, s)[0] == 0


def test_reference_site_known_answers():
    # tests/test_fermion_site.cpp:14-26
    assert O.site_apply(O.FERMION, "N", 1, 0b1010) == (1, 0b1010, 1.0)
    assert O.site_apply(O.FERMION, "N", 0, 0b1010)[2] == 0.0
    assert O.site_apply(O.FERMION, "C", 1, 0b1010) == (1, 0b1000, 1.0)
    assert O.site_apply(O.FERMION, "C", 3, 0b1010) == (1, 0b0010, -1.0)
    # tests/test_tj_site.cpp:14,18
    assert O.site_apply(O.TJ, "CUp", 0, 0b0001) == (1, 0, 1.0)
    assert O.site_apply(O.TJ, "CdagDn", 0, 0b0001)[0] == 0


# ------------------------------------------------------------------ CSR
@pytest.mark.parametrize("name", sorted(FIX["models"]))
def test_csr_matches_reference_fixture(name):
    m = FIX["models"][name]
    terms = [(complex(*c), [tuple(f) for f in fs]) for c, fs in m["terms"]]
    mod = O.Model(m["bkind"], m["n"], sec_of(m["sector"]), m["skind"], terms)
    ro, cols, vals, diag = mod.csr()
    assert ro.tolist() == m["row_offsets"]
    assert cols.tolist() == m["cols"]
    ref_vals = np.array([complex(*v) for v in m["vals"]])
    assert np.array_equal(vals, ref_vals)
    assert np.allclose(diag, np.array([complex(*v) for v in m["diag"]]), rtol=0, atol=1e-15)


@pytest.mark.parametrize("name", sorted(FIX["models"]))
def test_csr_matches_textbook_dense(name):
    m = FIX["models"][name]
    terms = [(complex(*c), [tuple(f) for f in fs]) for c, fs in m["terms"]]
    if name == "spin_out_of_range_site":
        pytest.skip("site beyond the basis width: no finite dense embedding")
    nbits = m["n"] if m["bkind"] in (O.SPIN, O.FERMION) else 2 * m["n"]
    if nbits > 10:
        # dense 2^12 products are too slow for the CPU suite: check the same
        # model family on a 2x2 plaquette instead
        terms = models.tj_model(2, 2, 1.0, 0.4)
        m = dict(m, n=4, sector=dict(use_nup=1, nup=1, use_ndn=1, ndn=2))
        nbits = 8
    mod = O.Model(m["bkind"], m["n"], sec_of(m["sector"]), m["skind"], terms)
    st = np.array(O.enumerate_basis(m["bkind"], m["n"], sec_of(m["sector"])), dtype=np.int64)
    # the site kind decides the operator semantics; the basis decides the rows
    ref = TB.stored_matrix(m["skind"], terms, nbits, st)
    # basis projection drops transitions that leave the sector; compare there
    assert np.allclose(mod.dense(), ref, atol=1e-14)


def test_known_answer_diagonal_0p3125():
    """bindings/julia/test/runtests.jl:102-106: <111|H|111> = 0.3125 for
    Sz0Sz1 + 0.5 Sz0Sz1Sz2 on a 3-site spin basis."""
    terms = [(1.0, [("Sz", 0), ("Sz", 1)]), (0.5, [("Sz", 0), ("Sz", 1), ("Sz", 2)])]
    m = O.Model(O.SPIN, 3, O.sector(), O.SPIN, terms)
    x = np.zeros(8, complex)
    x[7] = 1.0
    assert abs(m.apply(x)[7] - 0.3125) < 1e-15


# ------------------------------------------------------------------ solvers
def test_lanczos_known_answers():
    # -0.75 (test_lanczos_ritz.cpp:46), -1.6160254038 (test_workflow.cpp:46), -2.0 (runtests.jl:122)
    m2 = O.Model(O.SPIN, 2, O.sector(), O.SPIN, models.heisenberg_chain(2, periodic=False))
    r = m2.lanczos(200, 1e-12, start=O.start_vector(4, 3) + 0j, want_vec=True)
    assert abs(r["energy"] + 0.75) < 1e-12
    Hv = m2.apply(r["eigenvector"])
    assert np.linalg.norm(Hv - r["energy"] * r["eigenvector"]) < 1e-10
    m4 = O.Model(O.SPIN, 4, O.sector(use_sz=1, sz2=0), O.SPIN, models.heisenberg_chain(4, periodic=False))
    assert abs(m4.lanczos(200, 1e-12, start=O.start_vector(6, 1) + 0j)["energy"] + (0.75 + math.sqrt(3) / 2)) < 1e-12
    m4p = O.Model(O.SPIN, 4, O.sector(), O.SPIN, models.heisenberg_chain(4))
    r = m4p.lanczos(50, 1e-12, start=O.start_vector(16, 5) + 0j)
    assert abs(r["energy"] + 2.0) < 1e-12 and r["converged"]
    r = m4p.lanczos(2, 1e-15, start=O.start_vector(16, 5) + 0j)
    assert r["iterations"] == 2 and not r["converged"]    # runtests.jl:172-175


@pytest.mark.parametrize("L", [8, 10, 12])
def test_lanczos_vs_bethe_ansatz(L):
    m = O.Model(O.SPIN, L, O.sector(use_sz=1, sz2=0), O.SPIN, models.heisenberg_chain(L))
    r = m.lanczos(300, 1e-13, start=O.start_vector(m.dim, 11) + 0j)
    e_b = TB.bethe_xxx_ring(L)
    assert abs(r["energy"] - e_b) < 1e-10 * abs(e_b)


def test_bethe_helper_small_rings():
    for L in (4, 6, 8):
        st = TB.sector_states(TB.SPIN, L, use_sz=1, sz2=0)
        H = TB.stored_matrix(TB.SPIN, models.heisenberg_chain(L), L, st)
        assert abs(np.linalg.eigvalsh(H)[0] - TB.bethe_xxx_ring(L)) < 1e-12


@pytest.mark.parametrize("L,sz2", [(6, 0), (8, 0), (8, 2), (10, 0), (10, 4)])
def test_xxz_bethe_helper_vs_textbook_dense(L, sz2):
    """The XXZ Bethe-ansatz solver the serving tests use, against dense
    matrices built from the definitions (no oracle involved)."""
    st = TB.sector_states(TB.SPIN, L, use_sz=1, sz2=sz2)
    for delta in (0.5, 0.73, 0.95):
        H = TB.stored_matrix(TB.SPIN, models.xxz_chain(L, delta), L, st)
        assert abs(np.linalg.eigvalsh(H)[0] - TB.bethe_xxz_ring(L, delta, sz2)) < 1e-12 * L
    # sector symmetry Sz <-> -Sz
    assert TB.bethe_xxz_ring(L, 0.6, sz2) == TB.bethe_xxz_ring(L, 0.6, -sz2)


@pytest.mark.parametrize("L", [4])
def test_lieb_wu_helper_vs_textbook_dense(L):
    """The Lieb-Wu solver the full-size Hubbard test uses (half filling),
    against dense Jordan-Wigner matrices built from the definitions, for
    two values of U."""
    h = L // 2
    st = TB.sector_states(TB.HUBBARD, L, use_nup=1, nup=h, use_ndn=1, ndn=h)
    for U in (4.0, 2.5):
        H = TB.stored_matrix(TB.HUBBARD, models.hubbard_chain(L, 1.0, U), 2 * L, st)
        assert abs(np.linalg.eigvalsh(H)[0] - TB.lieb_wu_ring(L, h, h, U)) < 1e-12 * L


def test_lieb_wu_helper_vs_oracle_L6_to_L10():
    """Beyond the textbook matrices' reach: the oracle's sector matrices at
    L = 6 and 8 (D = 400, 4,900, dense) and its reorthogonalised Lanczos at L = 10
    (D = 63,504) against Lieb-Wu."""
    for L in (6, 8):
        m = O.Model(O.HUBBARD, L, O.sector(use_nup=1, nup=L // 2, use_ndn=1, ndn=L // 2), O.HUBBARD,
                    models.hubbard_chain(L, 1.0, 4.0), enumerate=False)
        assert abs(np.linalg.eigvalsh(m.dense())[0] - TB.lieb_wu_ring(L, L // 2, L // 2, 4.0)) < 1e-12 * L
    m = O.Model(O.HUBBARD, 10, O.sector(use_nup=1, nup=5, use_ndn=1, ndn=5), O.HUBBARD,
                models.hubbard_chain(10, 1.0, 4.0), enumerate=False)
    r = m.lanczos(300, 1e-13)
    e = TB.lieb_wu_ring(10, 5, 5, 4.0)
    assert abs(r["energy"] - e) < 1e-11 * abs(e)


def test_lanczos_hubbard_tj_vs_dense():
    for bk, n, sec, terms, st_kw in [
        (O.HUBBARD, 4, dict(use_nup=1, nup=2, use_ndn=1, ndn=2), models.hubbard_chain(4, 1.0, 4.0),
         dict(use_nup=1, nup=2, use_ndn=1, ndn=2)),
        (O.TJ, 6, dict(use_nup=1, nup=2, use_ndn=1, ndn=2), models.tj_model(2, 3, 1.0, 0.4),
         dict(use_nup=1, nup=2, use_ndn=1, ndn=2)),
    ]:
        m = O.Model(bk, n, O.sector(**sec), bk, terms)
        ex = np.linalg.eigvalsh(m.dense())[0]
        r = m.lanczos(300, 1e-13, start=O.start_vector(m.dim, 2) + 0j)
        assert abs(r["energy"] - ex) < 1e-10 * abs(ex)


def test_ftlm_closed_form_two_site():
    """tests/test_ftlm.cpp: H = Sz0 Sz1 on 2 sites, Z = 4 cosh(beta/4). With the
    full orthonormal basis as the random vectors, FTLM is exact:
    Z_ftlm * D = Tr exp(-beta H)."""
    m = O.Model(O.SPIN, 2, O.sector(), O.SPIN, [(1.0, [("Sz", 0), ("Sz", 1)])])
    rv = np.eye(4, dtype=complex)
    for beta in (0.5, 1.0, 3.0):
        Z, E, Cv = m.ftlm(beta, rv, 10)
        assert abs(Z * 4 - 4 * math.cosh(beta / 4)) < 1e-12
        ex = TB.thermal([0.25, -0.25, -0.25, 0.25], [beta])[0]
        assert abs(E - ex[1]) < 1e-12 and abs(Cv - ex[2]) < 1e-12


def test_ftlm_exact_with_full_basis_heisenberg():
    m = O.Model(O.SPIN, 6, O.sector(use_sz=1, sz2=0), O.SPIN, models.heisenberg_chain(6))
    w = np.linalg.eigvalsh(m.dense())
    rv = np.eye(m.dim, dtype=complex)
    for beta in (0.2, 1.0, 4.0):
        Z, E, Cv = m.ftlm(beta, rv, m.dim + 5)
        ex = TB.thermal(w, [beta])[0]
        assert abs(Z * m.dim - ex[0]) < 1e-10 * ex[0]
        assert abs(E - ex[1]) < 1e-10 and abs(Cv - ex[2]) < 1e-9


def test_cf_against_dense_resolvent_and_moments():
    m = O.Model(O.SPIN, 8, O.sector(use_sz=1, sz2=0), O.SPIN, models.heisenberg_chain(8))
    H = m.dense()
    rng = np.random.default_rng(3)
    phi = rng.standard_normal(m.dim) + 1j * rng.standard_normal(m.dim)
    al, be, nrm = m.cf(phi, 200)
    assert abs(nrm - np.linalg.norm(phi)) < 1e-12
    u = phi / nrm
    a0 = np.vdot(u, H @ u).real
    assert abs(al[0] - a0) < 1e-12
    assert abs(be[0] - np.linalg.norm(H @ u - a0 * u)) < 1e-12
    E0 = np.linalg.eigvalsh(H)[0]
    for omega in (0.3, 1.0, 2.5):
        s = O.spectral(al, be, nrm, omega, E0, 0.2)
        ex = TB.resolvent_spectrum(H, phi, omega, E0, 0.2)
        assert abs(s - ex) < 1e-8 * max(1.0, abs(ex))
    # sum rule: integral of S over omega = |phi|^2
    om = np.linspace(-30, 30, 60001)
    S = np.array([O.spectral(al, be, nrm, w, 0.0, 0.05) for w in om[::100]])
    assert abs(np.trapezoid(S, om[::100]) - nrm ** 2) < 0.02 * nrm ** 2


def test_davidson_against_dense():
    m = O.Model(O.SPIN, 8, O.sector(use_sz=1, sz2=0), O.SPIN, models.heisenberg_chain(8))
    w = np.linalg.eigvalsh(m.dense())
    r = m.davidson(2, 20, 1e-9, want_vecs=True)
    assert r["converged"]
    assert abs(r["eigenvalues"][0] - w[0]) < 1e-9
    # test_c_api.cpp:110-127 semantics: E0 <= E1, residuals small
    assert r["eigenvalues"][0] <= r["eigenvalues"][1] + 1e-12
    v = r["eigenvectors"][0]
    assert np.linalg.norm(m.apply(v) - r["eigenvalues"][0] * v) < 1e-6


def test_planted_sign_error_is_caught():
    """A plausible mistake planted in a copy of the oracle's data (a dropped
    Jordan-Wigner sign) must fail the textbook comparison."""
    m = FIX["models"]["hubbard4_22"]
    terms = [(complex(*c), [tuple(f) for f in fs]) for c, fs in m["terms"]]
    mod = O.Model(m["bkind"], m["n"], sec_of(m["sector"]), m["skind"], terms)
    A = mod.dense()
    st = np.array(O.enumerate_basis(m["bkind"], m["n"], sec_of(m["sector"])), dtype=np.int64)
    ref = TB.stored_matrix(m["skind"], terms, 8, st)
    assert np.allclose(A, ref, atol=1e-14)
    planted = np.abs(A)          # every Jordan-Wigner sign dropped
    assert not np.allclose(planted, ref, atol=1e-14)
    flipped = A.copy()           # one transposed operand: H instead of H^dagger storage
    flipped[np.tril_indices(A.shape[0], -1)] *= -1
    assert not np.allclose(flipped, ref, atol=1e-14)
