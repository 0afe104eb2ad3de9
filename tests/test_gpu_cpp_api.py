"""The reference's C++ API tests (tests/*.cpp of sjp95/qkrylov that build a
MatrixFreeHamiltonian or call a solver), compiled unchanged against this
engine's headers at the reference's include paths (include/qkrylov/...)
and linked to libqkrylov_amd.so by oracle/Makefile (target `ref`, when the
reference checkout is present at build time), run as they are. Besides each
program's own assertions, the printed numbers are checked against closed
forms: 2-site singlet -3/4, 4-site open chain -(3/4 + sqrt(3)/2), the 2-site
Heisenberg matrix, Sz on one site, sum of Sz_i on four sites -2 (to the
six significant digits the programs print). Each program runs twice: built
with QKRYLOV_DOUBLE_PRECISION (namespace fp64) and without it (fp32)."""
import math
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref")
E4 = -(0.75 + math.sqrt(3) / 2)


PRECISIONS = pytest.mark.parametrize("prefix", ["cpp_", "cpp32_"], ids=["fp64", "fp32"])


def run(name, prefix="cpp_"):
    exe = os.path.join(REF, prefix + name)
    if not os.path.exists(exe):
        pytest.skip("the reference's C++ tests were not built (needs the reference checkout at build time)")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def number_after(text, label):
    m = re.search(re.escape(label) + r"\s*([-+0-9.eE]+)", text)
    assert m, text
    return float(m.group(1))


@PRECISIONS
def test_workflow(prefix):
    out = run("test_workflow", prefix)
    assert "test_heisenberg_workflow PASSED!" in out
    assert abs(number_after(out, "Computed Energy:") - E4) < 1e-5      # printed to 6 digits


@PRECISIONS
def test_lanczos_ritz(prefix):
    out = run("test_lanczos_ritz", prefix)
    assert "Ritz vector verification passed!" in out
    assert abs(number_after(out, "Lanczos Energy:") + 0.75) < 1e-5


@PRECISIONS
def test_lanczos_2site(prefix):
    assert abs(number_after(run("test_lanczos_2site", prefix), "Energy =") + 0.75) < 1e-5


@PRECISIONS
def test_user_friendly(prefix):
    assert abs(number_after(run("user_friendly_test", prefix), "Ground state energy:") - E4) < 1e-5   # 6 digits


@PRECISIONS
def test_heisenberg_2site_matrix(prefix):
    out = run("test_heisenberg_2site", prefix)
    rows = [ln for ln in out.splitlines() if ln.strip().startswith("(")]
    M = np.array([[complex(float(a), float(b)) for a, b in re.findall(r"\(([-+0-9.eE]+),([-+0-9.eE]+)\)", ln)]
                  for ln in rows])
    want = np.array([[0.25, 0, 0, 0], [0, -0.25, 0.5, 0], [0, 0.5, -0.25, 0], [0, 0, 0, 0.25]])
    assert M.shape == (4, 4) and np.allclose(M, want, atol=1e-7)


@PRECISIONS
def test_hamiltonian_sz(prefix):
    out = run("test_hamiltonian_sz", prefix)
    ys = [complex(*map(float, m)) for m in re.findall(r"y\[\d\] = \(([-+0-9.eE]+),([-+0-9.eE]+)\)", out)]
    assert ys == [-0.5, 0.5]


@PRECISIONS
def test_davidson(prefix):
    assert abs(number_after(run("test_davidson", prefix), "Davidson Energy:") + 2.0) < 1e-5


@PRECISIONS
def test_dynamics(prefix):
    out = run("test_dynamics", prefix)
    assert "Dynamics test passed!" in out and number_after(out, "S(0.5) =") >= 0.0


@PRECISIONS
def test_ftlm(prefix):
    out = run("test_ftlm", prefix)
    assert "FTLM test passed!" in out
    assert number_after(out, "FTLM Partition Function Z:") > 0.0
