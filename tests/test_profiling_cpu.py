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


def test_tj_ranks_against_enumeration():
    """The counting rank equals the position in the ascending enumeration of
    the sector, and the hole-major rank is a bijection onto 0..D-1 (3x2
    lattice, 2 up, 2 down: D = 6!/(2! 2! 2!) = 90)."""
    n, nup, ndn = 6, 2, 2
    states = []
    for s in range(1 << (2 * n)):
        vals = [(s >> (2 * i)) & 3 for i in range(n)]
        if 3 in vals or vals.count(1) != nup or vals.count(2) != ndn:
            continue
        states.append(s)
    assert len(states) == 90
    cnt = profiling._tj_rank_tables(n, nup, ndn)
    assert cnt[n][nup][ndn] == 90
    assert [profiling.tj_public_rank(s, n, nup, ndn, cnt) for s in states] == list(range(90))
    hm = sorted(profiling.tj_hole_major_rank(s, n, n - nup - ndn, nup) for s in states)
    assert hm == list(range(90))


def test_tj_locality_counts_entries():
    r = profiling.tj_locality(rows=2, cols=3, nup=2, ndn=2, samples=200, windows=(1, 100))
    assert r["D"] == 90
    assert 0 < r["entries_per_row"] <= 7            # at most one entry per bond
    for pub, hm in r["windows"].values():
        assert 0.0 <= pub <= 1.0 and 0.0 <= hm <= 1.0
    assert r["windows"][100] == (1.0, 1.0)          # every target lies within D rows
