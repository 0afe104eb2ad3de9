"""The gather-rate probe (qkrylov_amd_gather_probe): a measurement of the
vector memory path, not part of any solver. Checks that it runs within its
bounds on small arrays and refuses invalid arguments."""
import ctypes

import pytest

from qkrylov_amd import _lib

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nbytes,width", [(2 << 20, 8), (2 << 20, 16), (4096, 16), (2048, 8)])
def test_gather_probe_runs(nbytes, width):
    L = _lib.load()
    ms, g = ctypes.c_double(), ctypes.c_double()
    assert L.qkrylov_amd_gather_probe(nbytes, width, 2, ctypes.byref(ms), ctypes.byref(g)) == 0
    # 2048 workgroups x 8 waves x 64 rounds x 16 gathers
    assert g.value == 2048 * 8 * 64 * 16
    assert 0.0 < ms.value < 1000.0
    print(f"{nbytes} B, {width} B/lane: {ms.value:.4f} ms per launch", flush=True)


def test_gather_probe_refuses_invalid_arguments():
    L = _lib.load()
    ms, g = ctypes.c_double(), ctypes.c_double()
    for args in ((1024, 8, 2), (2 << 20, 4, 2), (2 << 20, 8, 0)):
        assert L.qkrylov_amd_gather_probe(*args, ctypes.byref(ms), ctypes.byref(g)) == -1
    assert L.qkrylov_amd_gather_probe(2 << 20, 8, 1, None, ctypes.byref(g)) == -1
