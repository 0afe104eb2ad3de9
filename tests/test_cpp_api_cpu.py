"""The reference's C++ API tests that need no GPU (bases, sites, OpSum:
tests/test_{spinhalf,fermion,tj}_basis.cpp, test_{spinhalf,fermion,tj}_site.cpp,
test_opsum.cpp of sjp95/qkrylov), compiled unchanged against this engine's
headers at the reference's include paths and linked to libqkrylov_amd.so
(oracle/Makefile target `ref`), run on the CPU. Their own assertions run;
the printed states and matrix elements are compared with the oracle's
enumeration and with the reference's own compiled site sources
(oracle/_ref/libqkref.so) where that was built. Each program runs built
with QKRYLOV_DOUBLE_PRECISION (fp64) and without it (fp32); the site
amplitudes printed are exact in both."""
import os
import re
import subprocess

import pytest

import oracle_lib as O

REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref")


PRECISIONS = pytest.mark.parametrize("prefix", ["cpp_", "cpp32_"], ids=["fp64", "fp32"])


def run(name, prefix="cpp_"):
    exe = os.path.join(REF, prefix + name)
    if not os.path.exists(exe):
        pytest.skip("the reference's C++ tests were not built (needs the reference checkout at build time)")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


@PRECISIONS
def test_reference_basis_programs(prefix):
    out = run("test_spinhalf_basis", prefix)
    assert "Basis dimension = 252" in out
    printed = [tuple(map(int, ln.split())) for ln in out.splitlines() if re.fullmatch(r"\d+\s+\d+", ln.strip())]
    want = [int(s) for s in O.enumerate_basis(O.SPIN, 10, O.sector(use_sz=1, sz2=0))[:10]]
    assert printed == list(enumerate(want))
    out = run("test_fermion_basis", prefix)
    states = [int(m) for m in re.findall(r"State \d+: (\d+)", out)]
    assert states == [int(s) for s in O.enumerate_basis(O.FERMION, 4, O.sector(use_n=1, n=2))]
    run("test_tj_basis", prefix)


@PRECISIONS
def test_reference_site_programs(prefix):
    out = run("test_spinhalf_site", prefix)
    blocks = re.findall(r"(\w+):\s*valid = (\d)\s*new_state = (\d+)\s*matrix_element = \(([-+0-9.eE]+),([-+0-9.eE]+)\)",
                        out)
    got = {b[0]: (int(b[1]), int(b[2]), complex(float(b[3]), float(b[4]))) for b in blocks}
    assert set(got) == {"Sx", "Sy", "Sz", "Sp", "Sm"}
    for op, (valid, new, amp) in got.items():
        state = 1 if op == "Sm" else 0
        ok, ns, a = O.site_apply(O.SPIN, op, 0, state)
        assert (valid, new, amp) == (int(ok), ns, a)
        if O.ref_available():
            rok, rns, ra = O.ref_site_apply(O.SPIN, op, 0, state)
            assert (valid, new, amp) == (int(rok), rns, ra)
    assert "FermionSite tests passed!" in run("test_fermion_site", prefix)
    assert "TJSite tests passed!" in run("test_tj_site", prefix)


@PRECISIONS
def test_reference_opsum_program(prefix):
    assert run("test_opsum", prefix).strip().splitlines()[-1].strip().endswith("1")
