"""The Python package's reference-facing surface on the CPU: operator
expressions and OpSum construction in every form the reference's Python
package and examples use (bindings/python/qkrylov/operators.py,
bindings/python/tests/*.py, examples/*.py), basis construction through the
Python classes against the reference-compiled basis fixtures (including
sectors that fix one species only), and result storage round trips. No
Hamiltonian is built here (that needs a GPU)."""
import json
import os

import numpy as np
import pytest

import qkrylov_amd as qk
from qkrylov_amd import storage

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ref_fixtures.json")))


# ------------------------------------------------------------------ operators
def test_op_enum_is_the_reference_name_set():
    names = ["Sz", "Sp", "Sm", "Sx", "Sy", "CdagUp", "CUp", "CdagDn", "CDn", "Nup", "Ndn", "Nupdn", "Bdag", "B", "N"]
    assert [o.value for o in qk.Op] == names
    assert qk.Op.Sz == "Sz" and isinstance(qk.Op.CdagUp, str)


def heis_terms(n):
    t = []
    for i in range(n - 1):
        t += [(1.0, [("Sz", i), ("Sz", i + 1)]), (0.5, [("Sp", i), ("Sm", i + 1)]), (0.5, [("Sm", i), ("Sp", i + 1)])]
    return [(complex(c), f) for c, f in t]


def test_opsum_forms_agree():
    """Tuple +=, Op-enum tuples, add_term and algebraic expressions build the
    same ordered term list (bindings/python/tests/test_workflow.py:17-20,
    examples/dynamics_example.py:11-14)."""
    n = 4
    a = qk.OpSum()
    for i in range(n - 1):
        a += 1.0 * qk.Sz(i) * qk.Sz(i + 1) + 0.5 * (qk.Sp(i) * qk.Sm(i + 1) + qk.Sm(i) * qk.Sp(i + 1))
    b = qk.OpSum()
    for i in range(n - 1):
        b += 1.0, qk.Op.Sz, i, qk.Op.Sz, i + 1
        b += 0.5, qk.Op.Sp, i, qk.Op.Sm, i + 1
        b += 0.5, qk.Op.Sm, i, qk.Op.Sp, i + 1
    c = qk.OpSum(dtype=np.float64)
    for i in range(n - 1):
        c.add_term(1.0, "Sz", i, "Sz", i + 1)
        c.add_term(0.5, "Sp", i, "Sm", i + 1)
        c.add_term(0.5, qk.Op.Sm, i, qk.Op.Sp, i + 1)
    want = heis_terms(n)
    for o in (a, b, c):
        assert o.size == 9 and len(o) == 9 and repr(o) == "OpSum(terms=9)"
        assert o.terms == want


def test_opsum_expression_algebra():
    """Term * Term concatenates factors and multiplies coefficients; a - b
    negates b (operator_term.hpp:73-77,105-114)."""
    o = qk.OpSum()
    o += (2.0 * qk.Sz(0)) * (3.0 * qk.Sz(1)) - 0.5 * qk.Sx(2)
    o += 1j * qk.CdagUp(0) * qk.CUp(1) - 1.0 * qk.N(0)
    o += -1.0 * (qk.B(1) * qk.Bdag(2))
    assert o.terms == [(6.0, [("Sz", 0), ("Sz", 1)]), (-0.5, [("Sx", 2)]), (1j, [("CdagUp", 0), ("CUp", 1)]),
                       (-1.0, [("N", 0)]), (-1.0, [("B", 1), ("Bdag", 2)])]


def test_opsum_invalid_and_clear():
    """bindings/python/tests/test_basic.py:35-40 and operators.py:151-182."""
    o = qk.OpSum()
    o += (1.0, "Sz", 0, "Sz", 1)
    with pytest.raises(ValueError):
        o += (1.0, "Sz")
    with pytest.raises(ValueError):
        o += (1.0, "Sz", 0, "Sz")
    with pytest.raises(ValueError):
        o.add_term(1.0, "Sz", 0, "Sz")
    with pytest.raises(ValueError):
        o += 3.0
    assert o.size == 1
    o.clear()
    assert o.size == 0 and o.terms == []


def test_opsum_handle_roundtrip_through_c_abi():
    """The term list reaches the library's OpSum unchanged in count."""
    from qkrylov_amd import _lib
    o = qk.OpSum()
    o += 1.0 * qk.Sz(0) * qk.Sz(1) + 0.25 * qk.Sp(0) * qk.Sm(1) * qk.Sz(2)
    h = o._handle()
    try:
        assert h
    finally:
        _lib.load().qkrylov_opsum_destroy(h)


# ------------------------------------------------------------------ bases
def python_basis(kind, n, sec):
    if kind == 0:
        return qk.SpinHalfBasis(n, sz=sec["sz2"] / 2) if sec.get("use_sz") else qk.SpinHalfBasis(n)
    if kind == 1:
        return qk.FermionBasis(n, conserve_n=True, n=sec["n"]) if sec.get("use_n") else qk.FermionBasis(n)
    cls = qk.HubbardBasis if kind == 2 else qk.TJBasis
    return cls(n, conserve_nup=bool(sec.get("use_nup")), nup=sec.get("nup", 0),
               conserve_ndn=bool(sec.get("use_ndn")), ndn=sec.get("ndn", 0), dtype=np.float32)


@pytest.mark.parametrize("case", FIX["bases"], ids=lambda c: f"k{c['kind']}n{c['n']}{c['sector']}")
def test_python_bases_match_reference_fixtures(case):
    """Every fixture basis, one-species Hubbard / t-J sectors included, built
    through the Python classes: same states in the same order, bit exact."""
    b = python_basis(case["kind"], case["n"], case["sector"])
    st = case["states"]
    assert b.size == len(st) and b.nsites == case["n"]
    assert [b.state(i) for i in range(len(st))] == st
    assert all(b.index(s) == i for i, s in enumerate(st))


def test_reference_python_basis_sizes():
    """bindings/python/tests/test_basic.py:4-30, test_workflow.py:7-11."""
    b = qk.SpinHalfBasis(4)
    assert b.size == 16 and b.nsites == 4
    assert qk.FermionBasis(4, conserve_n=True, n=2).size == 6
    assert qk.TJBasis(2).size == 9
    assert qk.SpinHalfBasis(N=4, sz=0).size == 6
    assert qk.SpinHalfBasis(4, conserve_sz=True).size == 6
    assert qk.SpinHalfBasis(4, sz=1).size == 4 and qk.SpinHalfBasis(4, sz=2).size == 1
    assert qk.SpinHalfBasis(5, sz=0.5).size == 10
    assert qk.HubbardBasis(4, nup=2).size == 6 * 16           # down electrons free
    assert qk.HubbardBasis(4, conserve_ndn=True, ndn=0).size == 16
    assert repr(qk.SpinHalfBasis(4, sz=0)) == "SpinHalfBasis(N=4, dim=6, sz=0)"
    assert repr(qk.HubbardBasis(3, nup=1, ndn=2)) == "HubbardBasis(N=3, dim=9, nup=1, ndn=2)"
    assert repr(qk.SpinHalfSite(dtype=np.float32)) == "SpinHalfSite()"


def test_gpu_helpers_consistent():
    n = qk.gpu_count()
    assert n >= 0
    assert qk.find_gpu() == ("hip" if n > 0 else None)
    assert isinstance(qk.__version__, str)


# ------------------------------------------------------------------ storage
def _results(rng):
    v = rng.standard_normal(6) + 1j * rng.standard_normal(6)
    lz = qk.LanczosResult(-1.6160254037844386, 5, True, v, np.array([0.1, -0.2]), np.array([0.9]))
    dv = qk.DavidsonResult(np.array([-2.0, -1.0]), [v, v.conj()], 7, True)
    dy = qk.DynamicsResult(np.array([0.5, 1.5, 2.5]), np.array([0.25, 0.125]), 0.75)
    ft = qk.FTLMResult(np.array([0.1, 1.0]), np.array([4.0, 3.0]), np.array([-0.1, -0.5]), np.array([0.01, 0.2]))
    return lz, dv, dy, ft


@pytest.mark.parametrize("suffix", [".npz", ".h5", ""])
def test_storage_roundtrip(tmp_path, suffix):
    lz, dv, dy, ft = _results(np.random.default_rng(3))
    p = storage.save_result(tmp_path / f"lz{suffix}", lz)
    r = storage.load_result(tmp_path / f"lz{suffix}")
    assert os.path.exists(p)
    assert r.energy == lz.energy and r.iterations == 5 and r.converged is True
    assert np.array_equal(r.eigenvector, lz.eigenvector) and r.eigenvector.dtype == np.complex128
    assert np.array_equal(r.alphas, lz.alphas) and np.array_equal(r.betas, lz.betas)
    storage.save_result(tmp_path / f"dv{suffix}", dv)
    r = storage.load_result(tmp_path / f"dv{suffix}")
    assert np.array_equal(r.eigenvalues, dv.eigenvalues) and len(r.eigenvectors) == 2
    assert all(np.array_equal(a, b) for a, b in zip(r.eigenvectors, dv.eigenvectors))
    storage.save_result(tmp_path / f"dy{suffix}", dy)
    r = storage.load_result(tmp_path / f"dy{suffix}")
    assert np.array_equal(r.alphas, dy.alphas) and np.array_equal(r.betas, dy.betas) and r.norm_phi0 == 0.75
    storage.save_result(tmp_path / f"ft{suffix}", ft)
    r = storage.load_result(tmp_path / f"ft{suffix}")
    for k in ("beta", "partition_function", "internal_energy", "specific_heat"):
        assert np.array_equal(getattr(r, k), getattr(ft, k))


def test_storage_dataset_names(tmp_path):
    """The archive holds the reference example's dataset names
    (examples/hdf5_storage.py:8-13)."""
    lz, dv, _, _ = _results(np.random.default_rng(4))
    with np.load(storage.save_result(tmp_path / "a.npz", lz)) as z:
        assert {"energy", "eigenvector"} <= set(z.files)
    with np.load(storage.save_result(tmp_path / "b.npz", dv)) as z:
        assert {"eigenvalues", "eigenvector_0", "eigenvector_1"} <= set(z.files)
    with pytest.raises(TypeError):
        storage.save_result(tmp_path / "c.npz", {"energy": 1.0})
