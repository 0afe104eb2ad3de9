"""The reference's own Python tests and examples, run against this package
on the GPU (bindings/python/tests/test_basic.py, test_workflow.py,
tests/python/test_scipy.py, examples/heisenberg_python.py,
dynamics_example.py, hdf5_storage.py),
plus to_sparse against the textbook Kronecker matrix, dtype handling and
storage of solver results."""
import math

import numpy as np
import pytest

import qkrylov_amd as qk
import textbook as TB

pytestmark = pytest.mark.gpu


def heis_obc(n, use_enum=False):
    os_ = qk.OpSum()
    for i in range(n - 1):
        if use_enum:
            os_ += 1.0, qk.Op.Sz, i, qk.Op.Sz, i + 1
            os_ += 0.5, qk.Op.Sp, i, qk.Op.Sm, i + 1
            os_ += 0.5, qk.Op.Sm, i, qk.Op.Sp, i + 1
        else:
            os_ += 1.0 * qk.Sz(i) * qk.Sz(i + 1) + 0.5 * (qk.Sp(i) * qk.Sm(i + 1) + qk.Sm(i) * qk.Sp(i + 1))
    return os_


def test_reference_basic():
    """test_basic.py:14-48: dimension, diagonal length, wrong-length apply."""
    os_ = qk.OpSum()
    os_.add_term(1.0, "Sz", 0, "Sz", 1)
    H = qk.MatrixFreeHamiltonian(qk.SpinHalfBasis(2), qk.SpinHalfSite(), os_)
    assert H.dimension == 4 and len(H.diagonal()) == 4
    assert np.array_equal(H.diagonal(), [0.25, -0.25, -0.25, 0.25])
    with pytest.raises(ValueError):
        H.apply(np.zeros(10, dtype=np.complex128))


def test_reference_workflow():
    """test_workflow.py:5-35: 4-site OBC chain, Sz = 0, E0 = -(3/4 + sqrt(3)/2),
    positional (maxiter, tol), eigenvector of the sector's length."""
    basis = qk.SpinHalfBasis(N=4, sz=0)
    os_ = heis_obc(4)
    assert os_.size == 9
    H = qk.MatrixFreeHamiltonian(basis, qk.SpinHalfSite(), os_)
    assert H.dimension == 6
    res = qk.lanczos_ground_state(H, 200, 1e-12)
    exact = -(0.75 + math.sqrt(3) / 2)
    assert abs(res.energy - exact) < 1e-12
    v = res.eigenvector
    assert len(v) == 6 and v.dtype == np.complex128
    assert np.linalg.norm(H @ v - res.energy * v) < 1e-8
    assert repr(res).startswith("LanczosResult(energy=")


def test_dynamics_example_flow():
    """examples/dynamics_example.py: Op-enum OpSum, GS, continued fraction
    started from the eigenvector, S(omega) on a grid; compared with the
    dense resolvent of the textbook matrix."""
    n = 4
    os_ = heis_obc(n, use_enum=True)
    H = qk.MatrixFreeHamiltonian(qk.SpinHalfBasis(N=n), qk.SpinHalfSite(), os_)
    gs = qk.lanczos_ground_state(H)
    res = qk.continued_fraction_coeffs(H, gs.eigenvector, n_iter=20)
    dense = TB.opsum_matrix(0, os_.terms, n)
    E0 = np.linalg.eigvalsh(dense)[0]
    assert abs(gs.energy - E0) < 1e-12
    for w in np.linspace(0, 5, 11):
        s = qk.evaluate_spectral_function(res, w, gs.energy, eta=0.1)
        want = TB.resolvent_spectrum(dense, gs.eigenvector, w, gs.energy, 0.1)
        assert abs(s - want) < 1e-8 * max(1.0, abs(want))
    with pytest.raises(ValueError):
        qk.continued_fraction_coeffs(H, np.zeros(3), n_iter=5)


def test_davidson_reference_defaults():
    """examples/heisenberg_python.py: davidson_lowest(H, n_eig=1)."""
    H = qk.MatrixFreeHamiltonian(qk.SpinHalfBasis(N=8), qk.SpinHalfSite(), heis_obc(8))
    d = qk.davidson_lowest(H, n_eig=2)
    ev = np.linalg.eigvalsh(TB.opsum_matrix(0, heis_obc(8).terms, 8))
    assert np.allclose(d.eigenvalues, ev[:2], atol=1e-7)
    assert repr(d).startswith("DavidsonResult(")


@pytest.mark.parametrize("kind,n,sec,site", [
    (0, 6, dict(use_sz=1, sz2=0), qk.SpinHalfSite),
    (2, 3, dict(use_nup=1, nup=2), qk.HubbardSite),
    (3, 4, dict(use_ndn=1, ndn=1), qk.TJSite),
])
def test_to_sparse_equals_textbook(kind, n, sec, site):
    """to_sparse (hamiltonian.py:147-180) is the stored matrix conj(<b|O|a>)
    on the sector, one-species sectors included."""
    from test_python_api import python_basis
    if kind == 0:
        terms = [(1.0, [("Sz", i), ("Sz", (i + 1) % n)]) for i in range(n)] + \
                [(0.5 + 0.25j, [("Sp", i), ("Sm", (i + 2) % n)]) for i in range(n)]
    else:
        terms = [(-1.0, [("CdagUp", i), ("CUp", (i + 1) % n)]) for i in range(n)] + \
                [(-0.7, [("CdagDn", (i + 1) % n), ("CDn", i)]) for i in range(n)] + \
                [(0.3j, [("CdagUp", i), ("CDn", (i + 2) % n)]) for i in range(n)]
        if kind == 2:
            terms += [(4.0, [("Nupdn", i)]) for i in range(n)]
    b = python_basis(kind, n, sec)
    H = qk.MatrixFreeHamiltonian(b, site(), qk.OpSum(terms))
    nbits = n if kind in (0, 1) else 2 * n
    st = np.array([b.state(i) for i in range(b.size)], dtype=np.int64)
    want = TB.stored_matrix(kind, terms, nbits, st)
    got = H.to_sparse().toarray()
    assert np.allclose(got, want, atol=1e-14)
    x = np.random.default_rng(1).standard_normal(b.size) + 0j
    assert np.allclose(H.apply(x), want @ x, atol=1e-13)


def test_dtype_float32_and_to():
    """dtype=float32 returns complex64 vectors as the reference's FP32 build;
    to() makes a new handle with the same terms."""
    b = qk.SpinHalfBasis(10, sz=0, dtype=np.float32)
    os_ = qk.OpSum(dtype=np.float32)
    for i in range(10):
        os_ += 1.0, "Sz", i, "Sz", (i + 1) % 10
        os_ += 0.5, "Sp", i, "Sm", (i + 1) % 10
        os_ += 0.5, "Sm", i, "Sp", (i + 1) % 10
    H = qk.MatrixFreeHamiltonian(b, qk.SpinHalfSite(dtype=np.float32), os_, device="cpu", dtype=np.float32)
    assert H.dtype == np.float32 and H.device == "cpu"
    x = np.random.default_rng(2).standard_normal(H.dimension).astype(np.complex64)
    y = H.apply(x)
    assert y.dtype == np.complex64 and H.diagonal().dtype == np.float32
    r = qk.lanczos_ground_state(H)
    assert r.eigenvector.dtype == np.complex64
    H64 = H.to(dtype=np.float64)
    assert H64.dtype == np.float64 and H64.dimension == H.dimension
    y64 = H64.apply(x)
    assert y64.dtype == np.complex128 and np.array_equal(y64.astype(np.complex64), y)
    assert abs(qk.lanczos_ground_state(H64).energy - TB.bethe_xxx_ring(10)) < 1e-11
    Hg = H.to(device="hip:0")
    assert Hg.device == "hip:0" and np.array_equal(Hg.apply(x), y)


def test_storage_example_flow(tmp_path):
    """examples/hdf5_storage.py: tuple OpSum, Lanczos, save; then load back."""
    n = 4
    os_ = qk.OpSum()
    for i in range(n - 1):
        os_ += 1.0, "Sz", i, "Sz", i + 1
    H = qk.MatrixFreeHamiltonian(qk.SpinHalfBasis(n), qk.SpinHalfSite(), os_)
    res = qk.lanczos_ground_state(H)
    assert abs(res.energy + 0.75) < 1e-12
    qk.save_result(tmp_path / "result.h5", res)
    back = qk.load_result(tmp_path / "result.h5")
    assert back.energy == res.energy and np.array_equal(back.eigenvector, res.eigenvector)
    d = qk.davidson_lowest(H, n_eig=2)
    qk.save_result(tmp_path / "dav.npz", d)
    back = qk.load_result(tmp_path / "dav.npz")
    assert np.array_equal(back.eigenvalues, d.eigenvalues)
    assert all(np.array_equal(a, b) for a, b in zip(back.eigenvectors, d.eigenvectors))


def test_reference_scipy_flow():
    """tests/python/test_scipy.py: the two-site Sz = 0 Hamiltonian from
    operator algebra, as a scipy LinearOperator and a sparse matrix, a host
    copy by to(device="cpu") and H @ x equal to H.apply(x); plus the values
    those objects must hold (singlet -3/4, triplet 1/4)."""
    basis = qk.SpinHalfBasis(2, sz=0)
    site = qk.SpinHalfSite()
    os_ = qk.OpSum()
    os_ += 1.0 * qk.Sz(0) * qk.Sz(1) + 0.5 * (qk.Sp(0) * qk.Sm(1) + qk.Sm(0) * qk.Sp(1))
    H = qk.MatrixFreeHamiltonian(basis, site, os_)
    assert H.dimension > 0
    H_op = H.aslinearoperator()
    assert H_op.shape == (H.dimension, H.dimension)
    H_mat = H.to_sparse()
    assert H_mat.shape == (H.dimension, H.dimension)
    H2 = H.to(device="cpu")
    assert H2.dimension == H.dimension
    x = np.random.rand(H.dimension)
    y1 = H.apply(x)
    y2 = H2 @ x
    np.testing.assert_allclose(y1, y2)
    np.testing.assert_allclose(H_op @ x, y1)
    np.testing.assert_allclose(H_mat @ x, y1)
    np.testing.assert_allclose(np.linalg.eigvalsh(H_mat.toarray()), [-0.75, 0.25], atol=1e-15)
