"""Writes tests/golden/ref_fixtures.json from the reference's OWN compiled
basis / site / opsum sources (oracle/_ref/libqkref.so, built by
oracle/Makefile from the reference checkout) and the Phase-1 loop of
oracle/ref_harness.cpp that drives them.

Run in the build container (where the reference checkout exists):
    make -C oracle && python tests/golden/make_golden.py
The fixtures are data only: inputs and the outputs the reference produced.
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle_lib as O  # noqa: E402
from qkrylov_amd import models  # noqa: E402

BASES = [
    # (kind, N, sector kwargs)
    (O.SPIN, 4, dict(use_sz=1, sz2=0)),
    (O.SPIN, 6, dict(use_sz=1, sz2=2)),
    (O.SPIN, 7, dict(use_sz=1, sz2=-1)),
    (O.SPIN, 5, dict(use_sz=1, sz2=0)),        # empty: parity mismatch
    (O.SPIN, 5, dict()),
    (O.FERMION, 4, dict()),
    (O.FERMION, 4, dict(use_n=1, n=2)),
    (O.FERMION, 7, dict(use_n=1, n=3)),
    (O.HUBBARD, 3, dict(use_nup=1, nup=2, use_ndn=1, ndn=1)),
    (O.HUBBARD, 4, dict(use_nup=1, nup=2, use_ndn=1, ndn=2)),
    (O.HUBBARD, 4, dict(use_nup=1, nup=1)),
    (O.HUBBARD, 2, dict()),
    (O.TJ, 2, dict()),
    (O.TJ, 4, dict(use_nup=1, nup=1, use_ndn=1, ndn=2)),
    (O.TJ, 5, dict(use_nup=1, nup=2, use_ndn=1, ndn=2)),
    (O.TJ, 3, dict(use_ndn=1, ndn=1)),
]

SITE_OPS = {
    O.SPIN: ["Sz", "Sp", "Sm", "Sx", "Sy", "Foo"],
    O.FERMION: ["N", "Id", "C", "Cdag", "Sz"],
    O.HUBBARD: ["Nup", "Ndn", "Nupdn", "CUp", "CdagUp", "CDn", "CdagDn", "Sz"],
    O.TJ: ["Nup", "Ndn", "Nupdn", "CUp", "CdagUp", "CDn", "CdagDn"],
}

MODELS = {
    "heis4_obc": (O.SPIN, 4, dict(use_sz=1, sz2=0), O.SPIN, models.heisenberg_chain(4, periodic=False)),
    "heis6_pbc_sz2": (O.SPIN, 6, dict(use_sz=1, sz2=2), O.SPIN, models.heisenberg_chain(6)),
    "spin5_full_xy": (O.SPIN, 5, dict(), O.SPIN,
                      [(0.3, [("Sx", 0)]), (0.7 + 0.2j, [("Sy", 2), ("Sz", 3)]), (1.0, [("Sy", 1), ("Sy", 4)]),
                       (-0.4j, [("Sp", 4), ("Sm", 1)]), (0.25, [("Sz", 0), ("Sz", 1), ("Sz", 2)])]),
    "fermion6_n3": (O.FERMION, 6, dict(use_n=1, n=3), O.FERMION,
                    [(-1.0, [("C", (i + 1) % 6), ("Cdag", i)]) for i in range(6)] +
                    [(-1.0, [("C", i), ("Cdag", (i + 1) % 6)]) for i in range(6)] +
                    [(0.5, [("N", i), ("N", (i + 1) % 6)]) for i in range(6)] +
                    [(0.1, [("Cdag", 0), ("C", 0)])]),
    "hubbard4_22": (O.HUBBARD, 4, dict(use_nup=1, nup=2, use_ndn=1, ndn=2), O.HUBBARD,
                    models.hubbard_chain(4, 1.0, 4.0)),
    "hubbard3_full": (O.HUBBARD, 3, dict(), O.HUBBARD,
                      models.hubbard_chain(3, 1.0, 2.0) + [(0.3, [("CdagUp", 0)]), (0.2, [("CDn", 2), ("CUp", 1)])]),
    "tj_2x3": (O.TJ, 6, dict(use_nup=1, nup=2, use_ndn=1, ndn=2), O.TJ, models.tj_model(2, 3, 1.0, 0.4)),
    "tj3_free": (O.TJ, 3, dict(), O.TJ, models.tj_model(1, 3, 1.0, 0.4, periodic=False)),
    "tj_with_hubbard_site": (O.TJ, 4, dict(use_nup=1, nup=1, use_ndn=1, ndn=2), O.HUBBARD,
                             models.hubbard_chain(4, 1.0, 3.0)),
    "spin_out_of_range_site": (O.SPIN, 4, dict(use_sz=1, sz2=0), O.SPIN,
                               [(1.0, [("Sz", 0), ("Sz", 9)]), (0.5, [("Sp", 9), ("Sm", 1)]),
                                (0.5, [("Sp", 0), ("Sm", 1)])]),
}


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref/libqkref.so not built (needs the reference checkout)")
    out = {"bases": [], "sites": [], "models": {}}
    for kind, n, sec in BASES:
        st = O.ref_basis_states(kind, n, O.sector(**sec))
        out["bases"].append(dict(kind=kind, n=n, sector=sec, states=[int(s) for s in st]))
    for kind, ops in SITE_OPS.items():
        nb = 8
        rows = []
        for op in ops:
            for site in range(3 if kind in (O.HUBBARD, O.TJ) else 4):
                for s in range(1 << nb):
                    v, ns, a = O.ref_site_apply(kind, op, site, s)
                    if v == 1:
                        rows.append([op, site, s, 1, int(ns), a.real, a.imag])
                    elif s == 0:
                        rows.append([op, site, s, v, 0, 0.0, 0.0])
                    if v == -1:
                        break
        out["sites"].append(dict(kind=kind, nbits=nb, rows=rows))
    for name, (bk, n, sec, sk, terms) in MODELS.items():
        r = O.ref_csr(bk, n, O.sector(**sec), sk, terms)
        ro, cols, vals, diag = r
        out["models"][name] = dict(
            bkind=bk, n=n, sector=sec, skind=sk,
            terms=[[[c.real, c.imag], [[f[0], f[1]] for f in fs]] for c, fs in [(complex(c), fs) for c, fs in terms]],
            row_offsets=[int(x) for x in ro], cols=[int(x) for x in cols],
            vals=[[v.real, v.imag] for v in vals], diag=[[d.real, d.imag] for d in diag])
    path = os.path.join(HERE, "ref_fixtures.json")
    with open(path, "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print(f"wrote {path} ({os.path.getsize(path) // 1024} KiB)")


if __name__ == "__main__":
    main()
