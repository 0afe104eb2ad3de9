"""CPU checks of the C-ABI library: it loads, exports every symbol the public
headers declare, and its host-side basis functions (closed-form rank /
unrank, no device involved) agree with the oracle. No kernel is launched."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", "qkrylov", "c_api.h"), os.path.join(ROOT, "include", "qkrylov_amd.h")]


def declared_symbols():
    names = set()
    for h in HEADERS:
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b(qkrylov_[a-z0-9_]+)\s*\(", txt):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    from qkrylov_amd import _lib
    L = _lib.load()
    names = declared_symbols()
    assert len(names) > 60
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    # the ctypes layer declares exactly the header's functions
    assert names == set(_lib.SIGNATURES), names ^ set(_lib.SIGNATURES)


def test_reference_abi_symbols_unchanged():
    """Every symbol of the reference ABI is present with the same spelling
    (include/qkrylov/c_api.h:30-168 of the reference)."""
    ref = {"qkrylov_sector_create", "qkrylov_sector_destroy", "qkrylov_sector_set_sz",
           "qkrylov_sector_set_hubbard_particles", "qkrylov_sector_set_n", "qkrylov_sector_set_nb",
           "qkrylov_sector_get_sz", "qkrylov_sector_get_hubbard_particles", "qkrylov_sector_get_n",
           "qkrylov_sector_get_nb", "qkrylov_spinhalf_basis_create", "qkrylov_fermion_basis_create",
           "qkrylov_hubbard_basis_create", "qkrylov_tj_basis_create", "qkrylov_basis_destroy",
           "qkrylov_basis_dimension", "qkrylov_basis_nsites", "qkrylov_basis_state", "qkrylov_basis_index",
           "qkrylov_basis_contains", "qkrylov_spinhalf_site_create", "qkrylov_fermion_site_create",
           "qkrylov_hubbard_site_create", "qkrylov_tj_site_create", "qkrylov_site_destroy", "qkrylov_opsum_create",
           "qkrylov_opsum_destroy", "qkrylov_opsum_clear", "qkrylov_opsum_add_term_1body",
           "qkrylov_opsum_add_term_2body", "qkrylov_opsum_add_term_nbody", "qkrylov_hamiltonian_create",
           "qkrylov_hamiltonian_destroy", "qkrylov_hamiltonian_dimension", "qkrylov_hamiltonian_apply",
           "qkrylov_hamiltonian_apply_complex", "qkrylov_hamiltonian_diagonal", "qkrylov_lanczos_ground_state",
           "qkrylov_lanczos_ground_state_complex", "qkrylov_davidson_lowest_complex",
           "qkrylov_continued_fraction_coeffs_complex", "qkrylov_evaluate_spectral_function", "qkrylov_ftlm"}
    assert ref <= declared_symbols()


def _engine_basis(L, kind, n, sec):
    s = L.qkrylov_sector_create()
    if sec.get("use_sz"):
        L.qkrylov_sector_set_sz(s, sec["sz2"])
    if sec.get("use_n"):
        L.qkrylov_sector_set_n(s, sec["n"])
    if sec.get("use_nup") and sec.get("use_ndn"):
        L.qkrylov_sector_set_hubbard_particles(s, sec["nup"], sec["ndn"])
    f = {O.SPIN: L.qkrylov_spinhalf_basis_create, O.FERMION: L.qkrylov_fermion_basis_create,
         O.HUBBARD: L.qkrylov_hubbard_basis_create, O.TJ: L.qkrylov_tj_basis_create}[kind]
    b = f(n, s)
    L.qkrylov_sector_destroy(s)
    return b


@pytest.mark.parametrize("kind,n,sec", [
    (O.SPIN, 10, dict(use_sz=1, sz2=0)), (O.SPIN, 9, dict(use_sz=1, sz2=3)), (O.SPIN, 7, dict()),
    (O.FERMION, 9, dict(use_n=1, n=4)), (O.HUBBARD, 5, dict(use_nup=1, nup=2, use_ndn=1, ndn=3)),
    (O.HUBBARD, 3, dict()), (O.TJ, 6, dict(use_nup=1, nup=2, use_ndn=1, ndn=2)), (O.TJ, 4, dict()),
])
def test_basis_abi_matches_oracle(kind, n, sec):
    from qkrylov_amd import _lib
    L = _lib.load()
    b = _engine_basis(L, kind, n, sec)
    assert b
    ref = O.enumerate_basis(kind, n, O.sector(**sec))
    assert L.qkrylov_basis_dimension(b) == len(ref)
    assert L.qkrylov_basis_nsites(b) == n
    for i, s in enumerate(ref):
        assert L.qkrylov_basis_state(b, i) == s
        assert L.qkrylov_basis_index(b, int(s)) == i
        assert L.qkrylov_basis_contains(b, int(s)) == 1
    refset = set(int(s) for s in ref)
    nb = n if kind in (O.SPIN, O.FERMION) else 2 * n
    for s in range(min(1 << nb, 4096)):
        if s not in refset:
            assert L.qkrylov_basis_index(b, s) == -1 and L.qkrylov_basis_contains(b, s) == 0
    assert L.qkrylov_basis_state(b, len(ref)) == 0          # out of range -> 0 (reference catches)
    L.qkrylov_basis_destroy(b)


def test_abi_error_conventions_on_cpu():
    from qkrylov_amd import _lib
    L = _lib.load()
    assert L.qkrylov_spinhalf_basis_create(0, None) is None          # N <= 0 throws -> NULL
    assert L.qkrylov_hubbard_basis_create(32, None) is None          # N > 31
    assert L.qkrylov_sector_set_sz(None, 0) == -1
    assert L.qkrylov_opsum_clear(None) == -1
    assert L.qkrylov_opsum_add_term_1body(None, 1.0, 0.0, b"Sz", 0) == -1
    o = L.qkrylov_opsum_create()
    assert L.qkrylov_opsum_add_term_nbody(o, 1.0, 0.0, 0, None, None) == -1
    assert L.qkrylov_hamiltonian_create(None, None, o) is None
    assert L.qkrylov_hamiltonian_dimension(None) == 0
    assert L.qkrylov_basis_dimension(None) == 0 and L.qkrylov_basis_index(None, 1) == -1
    L.qkrylov_opsum_destroy(o)
    L.qkrylov_opsum_destroy(None)
    L.qkrylov_basis_destroy(None)
    L.qkrylov_sector_destroy(None)
    sec = L.qkrylov_sector_create()
    L.qkrylov_sector_set_hubbard_particles(sec, 3, 2)
    nu, nd, act = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    assert L.qkrylov_sector_get_hubbard_particles(sec, ctypes.byref(nu), ctypes.byref(nd), ctypes.byref(act)) == 0
    assert (nu.value, nd.value, act.value) == (3, 2, 1)
    L.qkrylov_sector_destroy(sec)


def test_large_basis_rank_without_enumeration():
    """601,080,390-state sector: ABI rank/unrank against the oracle's
    counting rank on a sample (the reference would enumerate 2^32 states)."""
    from qkrylov_amd import _lib
    L = _lib.load()
    b = _engine_basis(L, O.SPIN, 32, dict(use_sz=1, sz2=0))
    assert L.qkrylov_basis_dimension(b) == 601080390
    OL = O.lib()
    sec = O.sector(use_sz=1, sz2=0)
    for r in np.linspace(0, 601080389, 257).astype(np.int64):
        s = L.qkrylov_basis_state(b, int(r))
        assert s == OL.qo_count_unrank(O.SPIN, 32, sec, int(r))
        assert L.qkrylov_basis_index(b, s) == r
    L.qkrylov_basis_destroy(b)


def test_apply_lds_plan_counts_the_static_lds():
    """The single-vector apply's launch plan (host arithmetic only) keeps the
    dynamic tables plus the kernel's 136 B of static LDS within the 160 KB of
    a workgroup: a table set that fits 16 waves only without the static part
    falls back to 8 waves instead of failing at launch."""
    from qkrylov_amd import _lib
    L = _lib.load()
    f = L.qkl_apply_lds_plan
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_size_t, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                  ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_size_t)]
    nB, nA = 40, 30
    branch = (nB * 36 + 15) & ~15              # lo fields, lM, re, im of the per-lane branches
    wave = (nB * 16 + nA * 12 + 15) & ~15      # real vectors
    limit, static = 160 * 1024, 17 * 8
    w, lds = ctypes.c_int(), ctypes.c_size_t()

    def plan(tables, pref=16):
        rc = f(tables, nB, nA, 0, pref, ctypes.byref(w), ctypes.byref(lds))
        return rc, w.value, lds.value

    exact = limit - static - branch - 16 * wave            # 16 waves fit with nothing to spare
    assert plan(exact) == (0, 16, exact + branch + 16 * wave)
    for over in (8, 64, static):                           # within the static bytes of the limit
        rc, waves, dyn = plan(exact + over)
        assert (rc, waves) == (0, 8) and dyn + static <= limit
    assert plan(1024, pref=4)[:2] == (0, 4)
    assert plan(limit)[0] == -1                            # nothing fits
