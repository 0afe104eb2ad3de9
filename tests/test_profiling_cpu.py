"""The vector-memory access model of qkrylov_amd.profiling (CPU only): the
per-four-lane sector count it uses, checked on hand-made address patterns,
and its invariants on a small J1-J2 lattice."""
import numpy as np

from qkrylov_amd import profiling


def test_sector_counts_of_known_patterns():
    """Known answers of the per-four-lane sector count."""
    row = np.arange(64) * 8                                   # 64 consecutive doubles from an aligned base
    assert profiling.quad_sectors(row) == 16                  # each quad reads 32 B of one sector
    assert profiling.quad_sectors(row + 40) == 16 + 8         # a base 40 B past a boundary splits every other quad
    assert profiling.quad_sectors(np.arange(64) * 64) == 64   # one sector per lane
    act = np.zeros(64, bool)
    act[::4] = True
    assert profiling.quad_sectors(np.arange(64) * 64, act) == 16
    assert profiling.quad_sectors(np.arange(64) * 64, np.zeros(64, bool)) == 0
    assert profiling.quad_sectors(np.full(64, 4096)) == 16    # one address for all: one access per quad


def test_quad_model_invariants_small_lattice():
    r = profiling.quad_model(rows=4, cols=4, NL=8, tiles=30, seed=3)
    for kind in ("hionly", "inblock", "mixed"):
        assert 0 < r[(kind, "masked")] <= r[(kind, "own")]
    # a hi-only branch reaches every lane of a tile it applies to
    assert r[("hionly", "masked")] == r[("hionly", "own")]
    # every instruction costs at least one access per quad of lanes that load
    assert r[("hionly", "own")] >= 16
