"""Lanczos from the reference's own start vector (start_mode "reference":
std::mt19937(1234) complex uniform draws, src/solvers/lanczos.cpp:115-119)
against the oracle's Lanczos (full CGS reorthogonalisation, as
lanczos.cpp:149-151) started from the same vector: the whole alpha/beta
trajectory, the iteration count and convergence flag, the energy, and the
Ritz vector (equal up to the sign the tridiagonal eigensolver gives its
eigenvector). The start vector itself is pinned on the CPU
(test_start_vector_cpu.py). Likewise ftlm() from the reference's own complex
normal draws against the oracle's FTLM from the same vectors."""
import ctypes

import numpy as np
import pytest

import oracle_lib as O
import qkrylov_amd as qa
from qkrylov_amd import _lib, models

pytestmark = pytest.mark.gpu

def heis(n):
    terms = models.heisenberg_chain(n)
    return (qa.MatrixFreeHamiltonian(qa.SpinHalfBasis(n, sz2=0), qa.SpinHalfSite(), qa.OpSum(terms)),
            O.Model(O.SPIN, n, O.sector(use_sz=1, sz2=0), O.SPIN, terms))


def complex_spin(n):
    terms = models.heisenberg_chain(n) + [(0.3, [("Sy", i), ("Sy", (i + 2) % n)]) for i in range(n)] + \
        [(0.2j, [("Sp", i), ("Sm", (i + 3) % n)]) for i in range(n)] + \
        [(-0.2j, [("Sm", i), ("Sp", (i + 3) % n)]) for i in range(n)]
    return (qa.MatrixFreeHamiltonian(qa.SpinHalfBasis(n, sz2=0), qa.SpinHalfSite(), qa.OpSum(terms)),
            O.Model(O.SPIN, n, O.sector(use_sz=1, sz2=0), O.SPIN, terms))


def hubbard(n, nup, ndn):
    terms = models.hubbard_chain(n, 1.0, 4.0)
    return (qa.MatrixFreeHamiltonian(qa.HubbardBasis(n, conserve_nup=True, nup=nup, conserve_ndn=True, ndn=ndn),
                                     qa.HubbardSite(), qa.OpSum(terms)),
            O.Model(O.HUBBARD, n, O.sector(use_nup=1, nup=nup, use_ndn=1, ndn=ndn), O.HUBBARD, terms))


CASES = {"heis12": lambda: heis(12), "complex10": lambda: complex_spin(10), "hubbard6": lambda: hubbard(6, 3, 3)}


def check_same_run(r_energy, r_iters, r_conv, r_alphas, r_betas, vec, rr):
    assert r_iters == rr["iterations"] and r_conv == rr["converged"]
    assert abs(r_energy - rr["energy"]) < 1e-10 * abs(rr["energy"])
    if r_alphas is not None:
        assert len(r_alphas) == len(rr["alphas"]) and len(r_betas) == len(rr["betas"])
        assert np.allclose(r_alphas, rr["alphas"], rtol=1e-11, atol=1e-11)
        assert np.allclose(r_betas, rr["betas"], rtol=1e-10, atol=1e-11)
    if vec is not None:
        ref = rr["eigenvector"]
        s = np.vdot(ref, vec)
        assert abs(abs(s) - 1.0) < 1e-8
        assert abs(s.imag) < 1e-8 and abs(abs(s.real) - 1.0) < 1e-8   # the same phase, up to a sign
        assert np.linalg.norm(vec - np.sign(s.real) * ref) < 1e-8


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("mode", ["reference", "reference-clang"])
def test_reference_start_follows_the_oracle_trajectory(name, mode):
    H, om = CASES[name]()
    start = qa.reference_start_vector(om.dim, 1234, mode)
    rr = om.lanczos(200, 1e-12, start=start, reorth=True, want_vec=True)
    r = qa.lanczos_ground_state(H, 200, 1e-12, start_mode=mode)
    print(f"{name} {mode}: D={om.dim} iterations={r.iterations} (oracle {rr['iterations']})", flush=True)
    assert np.iscomplexobj(r.eigenvector) and np.linalg.norm(r.eigenvector.imag) > 1e-3
    check_same_run(r.energy, r.iterations, r.converged, r.alphas, r.betas, r.eigenvector, rr)


def test_process_default_drives_the_reference_named_c_call():
    H, om = heis(12)
    start = qa.reference_start_vector(om.dim, 1234, "reference")
    rr = om.lanczos(200, 1e-12, start=start, reorth=True, want_vec=True)
    L = _lib.load()
    old = qa.get_start_mode()
    try:
        qa.set_start_mode("reference")
        res = _lib.LanczosResultF64()
        vec = np.zeros(2 * om.dim)
        vp = vec.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        assert L.qkrylov_lanczos_ground_state_f64(H._h, 200, 1e-12, ctypes.byref(res), vp) == 0
        check_same_run(res.energy, res.iterations, bool(res.converged), None, None, vec.view(np.complex128), rr)
        # the Python name with no start_mode takes the process default too
        r = qa.lanczos_ground_state(H, 200, 1e-12)
        check_same_run(r.energy, r.iterations, r.converged, r.alphas, r.betas, r.eigenvector, rr)
    finally:
        qa.set_start_mode(old)
    # back to the counter RNG: a real run, which differs from the reference's
    r = qa.lanczos_ground_state(H, 200, 1e-12)
    assert abs(r.energy - rr["energy"]) < 1e-10 * abs(rr["energy"])
    assert np.linalg.norm(r.eigenvector.imag) == 0.0


def test_reference_start_edge_cases():
    # one-state basis: the reference returns the diagonal element after one step
    H1 = qa.MatrixFreeHamiltonian(qa.SpinHalfBasis(4, sz2=4), qa.SpinHalfSite(), qa.OpSum(models.heisenberg_chain(4)))
    assert H1.dimension == 1
    r = qa.lanczos_ground_state(H1, 200, 1e-12, start_mode="reference")
    assert abs(r.energy - 1.0) < 1e-14 and r.iterations == 1
    assert abs(abs(r.eigenvector[0]) - 1.0) < 1e-14
    # a given start vector takes precedence over the draws and is read as complex
    H, om = heis(10)
    start = qa.reference_start_vector(om.dim, 7, "reference-clang")
    rr = om.lanczos(200, 1e-12, start=start, reorth=True, want_vec=True)
    r = qa.lanczos_ground_state(H, 200, 1e-12, start=start, start_mode="reference")
    check_same_run(r.energy, r.iterations, r.converged, r.alphas, r.betas, r.eigenvector, rr)
    # the no-allocation driver runs real vectors only
    o = _lib.LanczosOptions(maxiter=50, tol=1e-12, reorthogonalize=0, fixed_steps=0, want_vector=0, seed=1234,
                            check_every=0, graph_mode=0, split_mode=0, start_mode=1)
    rep = _lib.LanczosReport()
    assert _lib.load().qkrylov_lanczos_run_noalloc_f64(H._h, ctypes.byref(o), None, None, None, None,
                                                       ctypes.byref(rep)) != 0
    # out-of-range modes are refused with a message naming the field
    for field, bad in (("start_mode", 3), ("graph_mode", -1), ("split_mode", 5)):
        ob = _lib.LanczosOptions(maxiter=50, tol=1e-12, reorthogonalize=0, fixed_steps=0, want_vector=0, seed=1234,
                                 check_every=0, graph_mode=0, split_mode=0, start_mode=0)
        setattr(ob, field, bad)
        assert _lib.load().qkrylov_lanczos_run_f64(H._h, ctypes.byref(ob), None, None, None, None,
                                                   ctypes.byref(rep)) == -1
        assert field in _lib.last_error()


@pytest.mark.parametrize("mode", ["reference", "reference-clang", "counter"])
@pytest.mark.parametrize("name", ["heis10", "complex10"])
def test_ftlm_from_the_reference_draws_matches_the_oracle(name, mode):
    """ftlm() draws the reference's complex normal vectors (std::mt19937(42),
    one normal_distribution stream, src/solvers/ftlm.cpp:133-143) in the
    order the start mode names (GCC's for "reference", Clang's otherwise);
    the oracle's FTLM (ftlm.cpp:122-170) from those same vectors gives the
    same Z, E and C_v."""
    H, om = heis(10) if name == "heis10" else complex_spin(10)
    n_random, M = 4, 15
    order = "reference" if mode == "reference" else "reference-clang"
    rv = qa.reference_normal_vector(n_random * om.dim, 42, order).reshape(n_random, om.dim)
    old = qa.get_start_mode()
    try:
        qa.set_start_mode(mode)
        got = {beta: qa.ftlm(H, beta, n_random, M) for beta in (0.3, 1.0)}
    finally:
        qa.set_start_mode(old)
    for beta, r in got.items():
        Z, E, Cv = om.ftlm(beta, rv, M)
        print(f"{name} {mode} beta={beta}: Z {r.partition_function:.12g} (oracle {Z:.12g})", flush=True)
        assert abs(r.partition_function - Z) < 1e-8 * abs(Z)
        assert abs(r.internal_energy - E) < 1e-8 * max(1.0, abs(E))
        assert abs(r.specific_heat - Cv) < 1e-8 * max(1.0, abs(Cv))
