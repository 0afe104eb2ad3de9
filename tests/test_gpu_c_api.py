"""The reference ABI in FP32, as its callers use it.

1. The reference's own C-ABI test program (tests/test_c_api.cpp of
   sjp95/qkrylov) compiled unchanged against include/qkrylov/c_api.h and
   linked to libqkrylov_amd.so (oracle/Makefile builds it into oracle/_ref
   when the reference checkout is present at build time), run as is.
2. The same flow through ctypes with the FP32 prototypes, with its
   assertions (4-site open Heisenberg chain: E0 = -(3/4 + sqrt(3)/2))."""
import ctypes
import math
import os
import subprocess

import numpy as np
import pytest

from qkrylov_amd import _lib

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "test_c_api")


def test_reference_c_api_program_runs_unchanged():
    if not os.path.exists(BIN):
        pytest.skip("the reference's C-ABI test was not built (needs the reference checkout at build time)")
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "C API tests passed successfully!" in r.stdout
    line = [x for x in r.stdout.splitlines() if "Lanczos Ground State Energy" in x][0]
    assert abs(float(line.split(":")[1]) - (-(0.75 + math.sqrt(3) / 2))) < 1e-5


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def test_fp32_abi_flow():
    L = _lib.load()
    sec = L.qkrylov_sector_create()
    assert L.qkrylov_sector_set_sz(sec, 0) == 0
    basis = L.qkrylov_spinhalf_basis_create(4, sec)
    dim = L.qkrylov_basis_dimension(basis)
    assert dim == 6 and L.qkrylov_basis_nsites(basis) == 4
    states = [L.qkrylov_basis_state(basis, i) for i in range(dim)]
    assert states == sorted(states) and all(bin(st).count("1") == 2 for st in states)
    s0 = L.qkrylov_basis_state(basis, 0)
    assert L.qkrylov_basis_contains(basis, s0) == 1 and L.qkrylov_basis_index(basis, s0) == 0
    site = L.qkrylov_spinhalf_site_create()
    ops = L.qkrylov_opsum_create()
    for i in range(3):
        assert L.qkrylov_opsum_add_term_2body(ops, 1.0, 0.0, b"Sz", i, b"Sz", i + 1) == 0
        assert L.qkrylov_opsum_add_term_2body(ops, 0.5, 0.0, b"Sp", i, b"Sm", i + 1) == 0
        assert L.qkrylov_opsum_add_term_2body(ops, 0.5, 0.0, b"Sm", i, b"Sp", i + 1) == 0
    H = L.qkrylov_hamiltonian_create(basis, site, ops)
    assert H and L.qkrylov_hamiltonian_dimension(H) == dim
    # the handles the Hamiltonian shares may go first
    L.qkrylov_opsum_destroy(ops)
    L.qkrylov_site_destroy(site)
    L.qkrylov_basis_destroy(basis)
    L.qkrylov_sector_destroy(sec)

    xr, xi = np.ones(dim, np.float32), np.zeros(dim, np.float32)
    yr, yi = np.zeros(dim, np.float32), np.zeros(dim, np.float32)
    assert L.qkrylov_hamiltonian_apply(H, _fp(xr), _fp(xi), _fp(yr), _fp(yi)) == 0
    xc = np.ones(dim, np.complex64)
    yc = np.zeros(dim, np.complex64)
    assert L.qkrylov_hamiltonian_apply_complex(H, _fp(xc.view(np.float32)), _fp(yc.view(np.float32))) == 0
    assert np.array_equal(yc.real, yr) and np.array_equal(yc.imag, yi)
    # x_imag / y_imag may be NULL
    y2 = np.zeros(dim, np.float32)
    assert L.qkrylov_hamiltonian_apply(H, _fp(xr), None, _fp(y2), None) == 0
    assert np.array_equal(y2, yr)
    diag = np.zeros(dim, np.float32)
    assert L.qkrylov_hamiltonian_diagonal(H, _fp(diag)) == 0
    sz = [[0.5 if (st >> i) & 1 else -0.5 for i in range(4)] for st in states]
    assert np.allclose(diag, [sum(z[i] * z[i + 1] for i in range(3)) for z in sz], atol=1e-7)

    res = _lib.LanczosResultC()
    psi = np.zeros(dim, np.complex64)
    assert L.qkrylov_lanczos_ground_state_complex(H, 200, 1e-12, ctypes.byref(res), _fp(psi.view(np.float32))) == 0
    e0 = -(0.75 + math.sqrt(3) / 2)
    assert abs(res.energy - e0) < 1e-5
    assert abs(np.vdot(psi, psi).real - 1.0) < 1e-4

    n_eig = 2
    ev = np.zeros(n_eig, np.float32)
    evecs = np.zeros(n_eig * dim, np.complex64)
    dinfo = _lib.DavidsonResultC()
    assert L.qkrylov_davidson_lowest_complex(H, n_eig, 20, 1e-6, _fp(ev), _fp(evecs.view(np.float32)),
                                             ctypes.byref(dinfo)) == 0
    assert dinfo.converged == 1 and abs(ev[0] - res.energy) < 1e-4 and ev[0] <= ev[1]
    for k in range(n_eig):
        vk = evecs[k * dim:(k + 1) * dim].copy()
        hv = np.zeros(dim, np.complex64)
        assert L.qkrylov_hamiltonian_apply_complex(H, _fp(vk.view(np.float32)), _fp(hv.view(np.float32))) == 0
        assert np.max(np.abs(hv - ev[k] * vk)) < 1e-3

    n_iter = 20
    al, be = np.zeros(n_iter, np.float32), np.zeros(n_iter, np.float32)
    nrm = ctypes.c_float()
    nc = ctypes.c_int()
    assert L.qkrylov_continued_fraction_coeffs_complex(H, _fp(psi.view(np.float32)), n_iter, _fp(al), _fp(be),
                                                       ctypes.byref(nrm), ctypes.byref(nc)) == 0
    assert nc.value > 0 and abs(nrm.value - 1.0) < 1e-3
    sval = L.qkrylov_evaluate_spectral_function(_fp(al), _fp(be), nc.value, nrm.value, 0.5, res.energy, 0.1)
    assert sval > 0.0

    fr = _lib.FTLMResultC()
    assert L.qkrylov_ftlm(H, 1.0, 10, 20, ctypes.byref(fr)) == 0
    assert fr.partition_function > 0.0
    L.qkrylov_hamiltonian_destroy(H)
    L.qkrylov_hamiltonian_destroy(None)


def test_fp32_abi_error_codes():
    L = _lib.load()
    assert L.qkrylov_hamiltonian_apply(None, None, None, None, None) == -1
    res = _lib.LanczosResultC()
    assert L.qkrylov_lanczos_ground_state(None, 10, 1e-6, ctypes.byref(res)) == -1
    ops = L.qkrylov_opsum_create()
    assert L.qkrylov_opsum_add_term_1body(ops, 1.0, 0.0, b"Sz", 0) == 0
    sec = L.qkrylov_sector_create()
    basis = L.qkrylov_spinhalf_basis_create(2, sec)
    site = L.qkrylov_spinhalf_site_create()
    bad = L.qkrylov_opsum_create()
    assert L.qkrylov_opsum_add_term_1body(bad, 1.0, 0.0, b"Bogus", 0) in (0, -1, -2)
    assert not L.qkrylov_hamiltonian_create(basis, site, bad)   # unknown operator: NULL, like the reference
    for h in (bad, ops):
        L.qkrylov_opsum_destroy(h)
    L.qkrylov_site_destroy(site)
    L.qkrylov_basis_destroy(basis)
    L.qkrylov_sector_destroy(sec)


def test_amd_copy_exact_for_any_length_and_alignment():
    """qkrylov_amd_copy (the STREAM-copy and counter-calibration kernel)
    copies bit for bit: odd lengths, and pointers off 16-byte alignment."""
    L = _lib.load()
    n = 100003
    src = np.random.default_rng(4).standard_normal(n + 2)
    a = L.qkrylov_amd_malloc((n + 2) * 8)
    b = L.qkrylov_amd_malloc((n + 2) * 8)
    assert a and b
    try:
        assert L.qkrylov_amd_memcpy_h2d(a, src.ctypes.data, (n + 2) * 8) == 0
        for off_s, off_d, m in ((0, 0, n), (8, 0, n), (0, 8, n - 1), (8, 8, 1), (0, 0, 0)):
            out = np.full(n + 2, np.nan)
            assert L.qkrylov_amd_memcpy_h2d(b, out.ctypes.data, (n + 2) * 8) == 0
            assert L.qkrylov_amd_copy(ctypes.c_void_p(b + off_d), ctypes.c_void_p(a + off_s), m) == 0
            assert L.qkrylov_amd_memcpy_d2h(out.ctypes.data, b, (n + 2) * 8) == 0
            got = out[off_d // 8: off_d // 8 + m]
            assert np.array_equal(got, src[off_s // 8: off_s // 8 + m])
            assert np.isnan(out[off_d // 8 + m:]).all()
    finally:
        L.qkrylov_amd_free(a)
        L.qkrylov_amd_free(b)
