"""The reference's Lanczos start vector (src/solvers/lanczos.cpp:115-119:
std::mt19937 rng(1234), std::uniform_real_distribution<Real> dist(-1, 1),
v(i) = KComplex(dist(rng), dist(rng))), as the library draws it on the host;
and, at the end, the reference's FTLM normal draws.

Pinned three ways, none of which reuses the library's code:
* numpy's legacy RandomState(seed) is MT19937 seeded by init_genrand, the
  seeding std::mt19937(seed) uses; its raw 32-bit outputs, combined the way
  libstdc++'s generate_canonical<double, 53> does (two draws,
  (g1 + g2 * 2^32) / 2^64), then mapped to [-1, 1), must give every entry
  bit for bit;
* the order of the two draws in an entry is unspecified in C++: a small
  program with the reference's two-argument constructor call, compiled here
  with g++ and with clang++, fixes which compiler's order each mode follows;
* a known answer: the first entry for seed 1234.
No GPU is used."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import qkrylov_amd as qa


def mt_uniform(seed, n):
    """2n uniform [-1, 1) draws of std::mt19937(seed) + libstdc++'s
    uniform_real_distribution<double>(-1, 1), from numpy's MT19937."""
    rs = np.random.RandomState(seed)
    raw = rs._bit_generator.random_raw(4 * n) & 0xFFFFFFFF
    g1 = raw[0::2].astype(np.float64)
    g2 = raw[1::2].astype(np.float64)
    u = (g1 + g2 * 4294967296.0) / 18446744073709551616.0
    u = np.where(u >= 1.0, np.nextafter(1.0, 0.0), u)
    return (1.0 - -1.0) * u + -1.0


@pytest.mark.parametrize("seed", [0, 1, 1234, 42, 2**32 - 1])
def test_draws_equal_mt19937_and_generate_canonical(seed):
    n = 1000
    d = mt_uniform(seed, n)
    gcc = qa.reference_start_vector(n, seed, "reference")
    clang = qa.reference_start_vector(n, seed, "reference-clang")
    # mode 1: the first draw of an entry is its imaginary part
    assert np.array_equal(gcc.imag, d[0::2]) and np.array_equal(gcc.real, d[1::2])
    # mode 2: the first draw is the real part
    assert np.array_equal(clang.real, d[0::2]) and np.array_equal(clang.imag, d[1::2])
    assert np.all(np.abs(d) <= 1.0) and np.any(d < 0) and np.any(d > 0)


def test_known_first_entry_for_seed_1234():
    v = qa.reference_start_vector(3, 1234, "reference")
    assert v[0] == complex(0.63567688577906645, -0.0046726673880966851)


PROBE = r"""
#include <cstdio>
#include <random>
struct Cplx {
    double re, im;
    Cplx(double a, double b) : re(a), im(b) {}
};
int main() {
    std::mt19937 rng(1234);
    std::uniform_real_distribution<double> dist(-1.0, 1.0);
    for (int i = 0; i < 4; ++i) {
        Cplx z(dist(rng), dist(rng));
        std::printf("%.17g %.17g\n", z.re, z.im);
    }
    return 0;
}
"""


def _compiler_order(cxx, tmp_path):
    src = tmp_path / "probe.cpp"
    src.write_text(PROBE)
    exe = tmp_path / ("probe_" + os.path.basename(cxx))
    subprocess.run([cxx, "-O2", str(src), "-o", str(exe)], check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    return np.array([complex(float(out[2 * i]), float(out[2 * i + 1])) for i in range(4)])


@pytest.mark.parametrize("cxx,mode", [("g++", "reference"), ("/opt/rocm/llvm/bin/clang++", "reference-clang")])
def test_modes_follow_each_compilers_argument_order(cxx, mode, tmp_path):
    path = shutil.which(cxx) or (cxx if os.path.exists(cxx) else None)
    if path is None:
        pytest.skip(f"{cxx} not available")
    z = _compiler_order(path, tmp_path)
    assert np.array_equal(z, qa.reference_start_vector(4, 1234, mode))


def test_start_mode_setting_round_trip_and_errors():
    old = qa.get_start_mode()
    try:
        for m in ("reference-clang", "reference", "counter"):
            qa.set_start_mode(m)
            assert qa.get_start_mode() == m
        with pytest.raises(ValueError):
            qa.set_start_mode("mt19937")
        with pytest.raises(ValueError):
            qa.reference_start_vector(4, 1234, 0)
    finally:
        qa.set_start_mode(old)


# The reference's FTLM draws (src/solvers/ftlm.cpp:133-143): std::mt19937(42),
# std::normal_distribution(0, 1), one stream over all random vectors, entry
# i = KComplex(dist(rng), dist(rng)). The polar method behind libstdc++'s
# normal_distribution keeps its second deviate for the next call, so the
# stream is checked whole, across many entries, against the same declarations
# compiled by g++ and by clang++.
NORMAL_PROBE = r"""
#include <cstdio>
#include <cstdlib>
#include <random>
struct Cplx {
    double re, im;
    Cplx(double a, double b) : re(a), im(b) {}
};
int main(int argc, char** argv) {
    const unsigned seed = (unsigned)std::strtoul(argv[1], nullptr, 10);
    const int n = std::atoi(argv[2]);
    std::mt19937 rng(seed);
    std::normal_distribution<double> dist(0.0, 1.0);
    for (int i = 0; i < n; ++i) {
        Cplx z(dist(rng), dist(rng));
        std::printf("%.17g %.17g\n", z.re, z.im);
    }
    return 0;
}
"""


def _normal_stream(cxx, tmp_path, seed, n):
    src = tmp_path / "normal_probe.cpp"
    src.write_text(NORMAL_PROBE)
    exe = tmp_path / ("normal_probe_" + os.path.basename(cxx))
    subprocess.run([cxx, "-O2", str(src), "-o", str(exe)], check=True, capture_output=True)
    out = subprocess.run([str(exe), str(seed), str(n)], check=True, capture_output=True, text=True).stdout.split()
    v = np.array([float(t) for t in out])
    return v[0::2] + 1j * v[1::2]


@pytest.mark.parametrize("cxx,mode", [("g++", "reference"), ("/opt/rocm/llvm/bin/clang++", "reference-clang")])
@pytest.mark.parametrize("seed", [42, 7])
def test_ftlm_normal_draws_follow_each_compilers_stream(cxx, mode, seed, tmp_path):
    path = shutil.which(cxx) or (cxx if os.path.exists(cxx) else None)
    if path is None:
        pytest.skip(f"{cxx} not available")
    n = 3001    # odd: the stream ends with a deviate held back by the polar method
    z = _normal_stream(path, tmp_path, seed, n)
    assert np.array_equal(z, qa.reference_normal_vector(n, seed, mode))


def test_ftlm_normal_draws_known_answer_and_order():
    g = qa.reference_normal_vector(2, 42, "reference")
    c = qa.reference_normal_vector(2, 42, "reference-clang")
    # the first two deviates of the stream, and the two orders swap them
    assert g[0] == complex(0.51543306969120128, -0.55023449442049355)
    assert c[0] == complex(-0.55023449442049355, 0.51543306969120128)
    assert g[1] == complex(c[1].imag, c[1].real)
    # a shorter request is a prefix of a longer one (one stream)
    assert np.array_equal(qa.reference_normal_vector(1000, 42)[:999], qa.reference_normal_vector(999, 42))
    with pytest.raises(ValueError):
        qa.reference_normal_vector(4, 42, "counter")
