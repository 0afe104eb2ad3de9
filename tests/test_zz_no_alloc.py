"""The no-allocation Lanczos entry point (qkrylov_lanczos_run_noalloc_f64)
against the allocating one (qkrylov_lanczos_run_f64), through the C ABI:
two calls in a row on one GPU, the second on a smaller Hamiltonian (the
static workspace keeps the first call's contents), plus the refusal of a
size beyond the workspace."""
import ctypes

import numpy as np
import pytest

import qkrylov_amd as qa
from qkrylov_amd import _lib, models

pytestmark = pytest.mark.gpu


def ring(L):
    return qa.MatrixFreeHamiltonian(qa.SpinHalfBasis(L, sz2=0), qa.SpinHalfSite(),
                                    qa.OpSum(models.heisenberg_chain(L)))


def run(fn, H, maxiter=300, tol=1e-12, reorth=0, want_vector=1, seed=1234):
    L = _lib.load()
    D = H.dimension
    opt = _lib.LanczosOptions(maxiter=maxiter, tol=tol, reorthogonalize=reorth, fixed_steps=0,
                              want_vector=want_vector, seed=seed, check_every=0, graph_mode=0)
    al, be = np.zeros(maxiter), np.zeros(maxiter)
    vec = np.zeros(D)
    rep = _lib.LanczosReport()
    p = lambda a: a.ctypes.data_as(_lib.p_dbl)  # noqa: E731
    rc = getattr(L, fn)(H._h, ctypes.byref(opt), None, p(al), p(be), p(vec) if want_vector else None,
                        ctypes.byref(rep))
    return rc, rep, al[:rep.n_alpha].copy(), be[:rep.n_beta].copy(), vec


def compare(H, **kw):
    rc0, r0, a0, b0, v0 = run("qkrylov_lanczos_run_f64", H, **kw)
    rc1, r1, a1, b1, v1 = run("qkrylov_lanczos_run_noalloc_f64", H, **kw)
    assert rc0 == 0 and rc1 == 0, _lib.last_error()
    assert r1.iterations == r0.iterations and r1.converged == r0.converged
    assert abs(r1.energy - r0.energy) <= 1e-10 * abs(r0.energy)
    # same kernels in the same order: the same recurrence
    assert np.array_equal(a1, a0) and np.array_equal(b1, b0)
    if kw.get("want_vector", 1):
        assert abs(np.linalg.norm(v1) - 1.0) < 1e-12
        assert abs(abs(v1 @ v0) - 1.0) < 1e-10
        assert np.linalg.norm(H.apply(v1).real - r1.energy * v1) < 1e-6
    return r1


def test_noalloc_lanczos_equals_allocating_twice_then_smaller():
    H20 = ring(20)                       # D = 184,756
    r = compare(H20)
    assert r.converged and abs(r.energy - (-8.904386529876442)) < 1e-9   # Bethe ansatz
    H14 = ring(14)                       # D = 3,432: second call, smaller input
    compare(H14)
    compare(H14, reorth=1)               # stored basis inside the workspace
    compare(H14, want_vector=0, maxiter=40, tol=0.0)


def test_noalloc_lanczos_refuses_sizes_beyond_the_workspace():
    H = ring(26)                         # D = 10,400,600: 3 vectors = 250 MB
    rc, *_ = run("qkrylov_lanczos_run_noalloc_f64", H, maxiter=5)
    assert rc == -2
    assert "workspace" in _lib.last_error()
    # the refused call leaves the entry point usable
    compare(ring(12))
