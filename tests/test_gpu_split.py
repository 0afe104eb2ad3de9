"""The two-pass apply over a relabelled basis (qk_split.hpp): the sites are
split into groups [B | C | A] such that every off-diagonal term acts inside
A u B or inside A u C, and the apply runs as two passes in the relabelled
basis' row order. Checked against the single-pass apply in the public row
order (the permutation in and out included): within 1e-13 relative (the
two paths add a row's entries in different orders) and bitwise
reproducible; the Lanczos energies of both paths agree within 1e-10, and
the split path's Ritz vector, returned in the public order, is an
eigenvector of H."""
import ctypes

import numpy as np
import pytest

import qkrylov_amd as qa
import textbook as TB
from qkrylov_amd import _lib, models

pytestmark = pytest.mark.gpu


def spin_h(terms, n, sz2=0):
    return qa.MatrixFreeHamiltonian(qa.SpinHalfBasis(n, sz2=sz2), qa.SpinHalfSite(), qa.OpSum(terms))


CASES = {
    "j1j2_6x4": (lambda: models.j1j2_square(6, 4, 1.0, 0.5), 24, 0, (10, 7, 7)),
    "j1j2_6x4_sz1": (lambda: models.j1j2_square(6, 4, 1.0, 0.5), 24, 2, (10, 7, 7)),
    "xxz_ring24": (lambda: models.xxz_chain(24, 0.7), 24, 0, (10, 7, 7)),
    "ring26_bonds_weighted": (lambda: models.heisenberg_bonds([(i, (i + 1) % 26, 1.0 + 0.01 * i) for i in range(26)]),
                              26, 0, (10, 8, 8)),
}


def apply_both(H, x):
    D = H.dimension
    xd, y1, y2 = qa.DeviceArray.from_host(x), qa.DeviceArray(D), qa.DeviceArray(D)
    H.apply_device(xd.ptr, y1.ptr)
    H.apply_split_device(xd.ptr, y2.ptr)
    _lib.check(_lib.load().qkrylov_amd_synchronize(), "sync")
    return y1.to_host(), y2.to_host()


@pytest.mark.parametrize("name", sorted(CASES))
def test_split_apply_equals_single_pass(name):
    terms_fn, n, sz2, groups = CASES[name]
    H = spin_h(terms_fn(), n, sz2)
    assert H.split_info(any_size=True) == groups
    assert H.split_info(any_size=False) is None          # below the automatic size
    rng = np.random.default_rng(3)
    x = rng.uniform(-1, 1, H.dimension)
    y1, y2 = apply_both(H, x)
    scale = np.max(np.abs(y1))
    assert np.max(np.abs(y2 - y1)) <= 1e-13 * scale
    # the same launches give the same bits
    _, y3 = apply_both(H, x)
    assert np.array_equal(y2, y3)
    # a basis vector probes the permutation alone: H e_j is column j
    j = H.dimension // 3
    e = np.zeros(H.dimension)
    e[j] = 1.0
    c1, c2 = apply_both(H, e)
    assert np.array_equal(np.nonzero(c1)[0], np.nonzero(c2)[0])
    assert np.max(np.abs(c1 - c2)) <= 1e-15 * np.max(np.abs(c1))


def test_models_without_a_split_are_refused():
    H = spin_h(models.j1j2_square(4, 8, 1.0, 0.5), 32)          # 8-site rows: no 10-site separator
    assert H.split_info(any_size=True) is None
    Hh = qa.MatrixFreeHamiltonian(qa.HubbardBasis(8, nup=4, ndn=4), qa.HubbardSite(),
                                  qa.OpSum(models.hubbard_chain(8)))
    assert Hh.split_info(any_size=True) is None
    x = qa.DeviceArray(H.dimension)
    with pytest.raises(RuntimeError, match="admits no split"):
        H.apply_split_device(x.ptr, x.ptr)


@pytest.mark.parametrize("name", ["j1j2_6x4", "xxz_ring24"])
def test_split_lanczos_equals_single_pass(name):
    terms_fn, n, sz2, _ = CASES[name]
    H = spin_h(terms_fn(), n, sz2)
    a = qa.lanczos_ground_state(H, maxiter=300, tol=1e-12, reorthogonalize=False, split=False)
    b = qa.lanczos_ground_state(H, maxiter=300, tol=1e-12, reorthogonalize=False, split=True)
    assert a.report["split"] == 0 and b.report["split"] == 1
    assert a.converged and b.converged
    assert abs(a.energy - b.energy) < 1e-10 * abs(a.energy)
    v = np.asarray(b.eigenvector).real
    assert abs(np.linalg.norm(v) - 1.0) < 1e-12
    # |dE| < 1e-12 leaves a Ritz residual of order 1e-6
    assert np.linalg.norm(H.apply(v + 0j).real - b.energy * v) < 1e-5
    assert abs(abs(v @ np.asarray(a.eigenvector).real) - 1.0) < 1e-8
    # a given start vector (public order) is permuted in: same trajectory
    # as the single pass up to rounding
    st = np.random.default_rng(5).uniform(-1, 1, H.dimension)
    c = qa.lanczos_ground_state(H, maxiter=40, tol=0.0, reorthogonalize=False, split=False, start=st,
                                return_state=False)
    d = qa.lanczos_ground_state(H, maxiter=40, tol=0.0, reorthogonalize=False, split=True, start=st,
                                return_state=False)
    assert np.allclose(c.alphas[:20], d.alphas[:20], rtol=1e-11, atol=1e-11)
    # the reorthogonalised (reference) mode through the split path
    r = qa.lanczos_ground_state(H, maxiter=300, tol=1e-12, reorthogonalize=True, split=True, return_state=True)
    assert r.report["split"] == 1 and abs(r.energy - a.energy) < 1e-10 * abs(a.energy)
    if name == "xxz_ring24":
        e = TB.bethe_xxz_ring(24, 0.7, 0)
        assert abs(b.energy - e) < 1e-10 * abs(e)


def test_full_size_ring32_split_lanczos_vs_bethe():
    """The 32-site Heisenberg ring (D = 601,080,390) through the split path
    against the Bethe ansatz; the automatic choice is the single pass."""
    H = spin_h(models.heisenberg_chain(32), 32)
    assert H.split_info(any_size=False) == (10, 11, 11)
    r = qa.lanczos_ground_state(H, maxiter=300, tol=1e-12, reorthogonalize=False, return_state=False, split=True)
    assert r.report["split"] == 1 and r.converged
    e = TB.bethe_xxx_ring(32)
    print(f"ring32 split: {r.energy!r} vs Bethe {e!r}, {r.iterations} it", flush=True)
    assert abs(r.energy - e) < 1e-10 * abs(e)
    a = qa.lanczos_ground_state(H, maxiter=5, tol=0.0, reorthogonalize=False, return_state=False)
    assert a.report["split"] == 0
