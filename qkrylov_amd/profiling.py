"""Readers for rocprofv3 result databases (the SQLite `run_results.db` that
`rocprofv3 -d DIR -o run` writes) and the HBM-traffic calibration used by
`bench.py`.

    python -m qkrylov_amd.profiling kernels DIR/run_results.db
    python -m qkrylov_amd.profiling counters FETCH.db WRITE.db --copy-bytes N [--match k_apply<]
    python -m qkrylov_amd.profiling pmc DB [DB ...] [--match k_apply<]
    python -m qkrylov_amd.profiling quad-model [--rows 8 --cols 4 --nl 14 --tiles 200]
    python -m qkrylov_amd.profiling tj-locality [--samples 1500]

`kernels` prints a per-kernel table (calls, mean/min/max duration) from a
`--kernel-trace` run; `counters` prints FETCH_SIZE / WRITE_SIZE per dispatch
(KiB) for the kernels whose name contains `--match`, scaled by what the same
counter reads for the plain copy kernel `k_copy` of known size; `pmc`
prints the mean per dispatch of every counter collected in the databases
for the matching kernels; `quad-model` counts, for sampled 64-row tiles of
the J1-J2 apply, the vector-memory accesses of its gathers as the L1 counts
them (per group of four lanes, one access per distinct 64 B sector), with
and without loads from the lanes a branch does not reach; `tj-locality`
compares how close (in rows) the targets of the t-J 4x5 entries lie to their
rows in the public basis order and in a hole-major order.
"""
import argparse
import sqlite3
import statistics
import sys


def kernel_table(db_path):
    """{kernel name: dict(calls, total_ns, mean_ns, min_ns, max_ns)} from the
    `kernels` view of a --kernel-trace database."""
    con = sqlite3.connect(db_path)
    rows = con.execute("select name, duration from kernels").fetchall()
    con.close()
    out = {}
    for name, dur in rows:
        out.setdefault(name, []).append(int(dur))
    return {n: dict(calls=len(d), total_ns=sum(d), mean_ns=statistics.fmean(d), min_ns=min(d), max_ns=max(d))
            for n, d in out.items()}


def counter_values(db_path, counter):
    """[(dispatch id, kernel name, value)] of one counter, one entry per
    dispatch (values of several rows of a dispatch are summed)."""
    con = sqlite3.connect(db_path)
    rows = con.execute("select dispatch_id, kernel_name, value from counters_collection where counter_name = ?",
                       (counter,)).fetchall()
    con.close()
    acc = {}
    for did, name, val in rows:
        k = (int(did), name)
        acc[k] = acc.get(k, 0.0) + float(val)
    return [(d, n, v) for (d, n), v in sorted(acc.items())]


def counter_means(db_path, match):
    """{counter: (mean per dispatch, dispatches)} over the kernels whose name
    contains `match`."""
    con = sqlite3.connect(db_path)
    rows = con.execute("select dispatch_id, kernel_name, counter_name, value from counters_collection").fetchall()
    con.close()
    acc = {}
    for did, name, ctr, val in rows:
        if match in name:
            acc.setdefault(ctr, {}).setdefault(int(did), 0.0)
            acc[ctr][int(did)] += float(val)
    return {c: (statistics.fmean(d.values()), len(d)) for c, d in acc.items()}


def short_name(name):
    """Kernel name without the return type and the argument list."""
    if name.startswith("void "):
        name = name[5:]
    depth = 0
    for i, ch in enumerate(name):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0 and i > 0:
            return name[:i]
    return name


def calibrated_traffic(fetch_db, write_db, copy_bytes, match):
    """Mean HBM-side bytes per dispatch of the kernels whose name contains
    `match`.  FETCH_SIZE and WRITE_SIZE (KiB per dispatch) are each scaled by
    the factor that maps the counter's reading of `k_copy` (which reads and
    writes `copy_bytes` bytes) onto those byte counts."""
    res = {}
    for db, counter in ((fetch_db, "FETCH_SIZE"), (write_db, "WRITE_SIZE")):
        vals = counter_values(db, counter)
        copy = [v for _, n, v in vals if "k_copy(" in n]
        hits = [v for _, n, v in vals if match in n]
        if not copy or not hits:
            raise RuntimeError(f"{counter}: no k_copy or no '{match}' dispatch in {db}")
        factor = (copy_bytes / 1024.0) / copy[-1]
        res[counter] = dict(kib_per_dispatch=statistics.fmean(hits), dispatches=len(hits),
                            copy_kib=copy[-1], factor=factor,
                            bytes_per_dispatch=statistics.fmean(hits) * factor * 1024.0)
    res["bytes_per_dispatch"] = res["FETCH_SIZE"]["bytes_per_dispatch"] + res["WRITE_SIZE"]["bytes_per_dispatch"]
    return res


def apply_acc_flag(name):
    """For a k_apply kernel name, its last template argument (the
    accumulating second pass of the split apply): 'true' / 'false', else
    None."""
    i = name.find("k_apply<")
    if i < 0:
        return None
    depth, j0 = 0, i + len("k_apply<")
    for j in range(j0, len(name)):
        if name[j] == "<":
            depth += 1
        elif name[j] == ">":
            if depth == 0:
                return name[j0:j].split(",")[-1].strip()
            depth -= 1
    return None


def calibrated_traffic_split(fetch_db, write_db, copy_bytes):
    """HBM-side bytes of one split apply: the mean per dispatch of its first
    pass (k_apply<..., false>) plus that of its second (k_apply<..., true>),
    each counter scaled by its k_copy factor as in calibrated_traffic."""
    res = {}
    for db, counter in ((fetch_db, "FETCH_SIZE"), (write_db, "WRITE_SIZE")):
        vals = counter_values(db, counter)
        copy = [v for _, n, v in vals if "k_copy(" in n]
        p1 = [v for _, n, v in vals if apply_acc_flag(n) == "false"]
        p2 = [v for _, n, v in vals if apply_acc_flag(n) == "true"]
        if not copy or not p1 or not p2:
            raise RuntimeError(f"{counter}: no k_copy or no split-apply dispatches in {db}")
        factor = (copy_bytes / 1024.0) / copy[-1]
        kib = statistics.fmean(p1) + statistics.fmean(p2)
        res[counter] = dict(kib_per_dispatch=kib, dispatches=len(p1) + len(p2), copy_kib=copy[-1], factor=factor,
                            bytes_per_dispatch=kib * factor * 1024.0,
                            pass_bytes=[statistics.fmean(p1) * factor * 1024.0, statistics.fmean(p2) * factor * 1024.0])
    res["bytes_per_dispatch"] = res["FETCH_SIZE"]["bytes_per_dispatch"] + res["WRITE_SIZE"]["bytes_per_dispatch"]
    return res


def quad_sectors(addr, active=None):
    """Accesses of one 64-lane gather as the L1 counts them: per group of
    four lanes, the number of distinct 64 B sectors its lanes (those with
    active set, when given) load from byte addresses addr."""
    total = 0
    for q in range(len(addr) // 4):
        a = addr[4 * q:4 * q + 4]
        if active is not None:
            a = a[active[4 * q:4 * q + 4]]
        total += len({int(v) >> 6 for v in a})
    return total


def quad_model(rows=8, cols=4, NL=14, tiles=200, seed=0):
    """Vector-memory accesses per 64-row tile of the pull-form apply of the
    J1-J2 Heisenberg model on a rows x cols torus (sites r * cols + c), Sz = 0,
    split at NL low bits: one gather instruction per branch a tile issues
    (hi-only branches only where the tile's hi takes them), costed per group
    of four lanes as the number of distinct 64 B sectors its lanes load.
    Returns {(kind, variant): accesses per tile} with kind in hionly /
    inblock / mixed and variant 'own' (lanes a branch does not reach load
    their own row) or 'masked' (they load nothing)."""
    import numpy as np
    from math import comb
    from . import models
    N = rows * cols
    NH, nup = N - NL, N // 2
    bonds = sorted({(min(f[0][1], f[1][1]), max(f[0][1], f[1][1]))
                    for c, f in models.j1j2_square(rows, cols) if f[0][0] == "Sp"})
    pc_lo = np.array([bin(v).count("1") for v in range(1 << NL)])
    lrank = np.zeros(1 << NL, dtype=np.int64)
    for k in range(NL + 1):
        idx = np.where(pc_lo == k)[0]
        lrank[idx] = np.arange(len(idx))
    pc_hi = np.array([bin(v).count("1") for v in range(1 << NH)])
    size = np.array([comb(NL, nup - p) if 0 <= nup - p <= NL else 0 for p in pc_hi])
    hoff = np.concatenate([[0], np.cumsum(size)[:-1]])
    LM = (1 << NL) - 1
    rng = np.random.default_rng(seed)
    acc, n = {}, 0
    while n < tiles:
        hi = int(rng.integers(0, 1 << NH))
        c = nup - pc_hi[hi]
        if not 0 <= c <= NL:
            continue
        los = np.where(pc_lo == c)[0]
        if len(los) < 64:
            continue
        t0 = int(rng.integers(0, len(los) - 63))
        s = (hi << NL) | los[t0:t0 + 64]
        own = hoff[hi] + lrank[s & LM]
        n += 1
        for a, b in bonds:
            act = ((s >> a) & 1) != ((s >> b) & 1)
            hi_only = a >= NL and b >= NL
            if hi_only and not act.any():
                continue
            kind = "hionly" if hi_only else ("inblock" if a < NL and b < NL else "mixed")
            t = s ^ ((1 << a) | (1 << b))
            tgt = np.where(act, hoff[t >> NL] + lrank[t & LM], own) * 8
            acc[(kind, "own")] = acc.get((kind, "own"), 0) + quad_sectors(tgt)
            acc[(kind, "masked")] = acc.get((kind, "masked"), 0) + quad_sectors(tgt, act)
    return {k: v / n for k, v in acc.items()}


def _tj_rank_tables(n, nup, ndn):
    """cnt[i][u][d]: number of t-J configurations of sites 0..i-1 (each site
    empty, up or down) with u up and d down electrons."""
    cnt = [[[0] * (ndn + 1) for _ in range(nup + 1)] for _ in range(n + 1)]
    cnt[0][0][0] = 1
    for i in range(n):
        for u in range(nup + 1):
            for d in range(ndn + 1):
                c = cnt[i][u][d]
                if not c:
                    continue
                cnt[i + 1][u][d] += c
                if u < nup:
                    cnt[i + 1][u + 1][d] += c
                if d < ndn:
                    cnt[i + 1][u][d + 1] += c
    return cnt


def tj_public_rank(state, n, nup, ndn, cnt):
    """Rank of a t-J state (up at bit 2i, down at bit 2i + 1 of site i) among
    all states of the sector in ascending bit-string order: the order of the
    engine's (and the reference's) t-J basis. Site values ascend as empty
    (0) < up (1) < down (2); walking from the top site, every smaller value at
    a site adds the completions of the sites below it."""
    r, u, d = 0, nup, ndn
    for i in range(n - 1, -1, -1):
        v = (state >> (2 * i)) & 3
        if v >= 1:
            r += cnt[i][u][d]              # site i empty instead
        if v == 2 and u >= 1:
            r += cnt[i][u - 1][d]          # site i up instead
        if v == 1:
            u -= 1
        elif v == 2:
            d -= 1
    return r


def tj_hole_major_rank(state, n, nholes, nup):
    """Rank with the hole positions slowest (combinatorial rank of the hole
    set) and the spin string of the occupied sites, in site order, fastest."""
    from math import comb
    vals = [(state >> (2 * i)) & 3 for i in range(n)]
    holes = [i for i, v in enumerate(vals) if v == 0]
    spins = [v for v in vals if v != 0]
    ups = [k for k, v in enumerate(spins) if v == 1]
    rh = sum(comb(p, j + 1) for j, p in enumerate(holes))
    rs = sum(comb(p, j + 1) for j, p in enumerate(ups))
    return rh * comb(n - nholes, nup) + rs


def tj_locality(rows=4, cols=5, nup=9, ndn=9, samples=1500, windows=(64, 8000, 64000, 500000), seed=0):
    """Fraction of the off-diagonal entries of sampled rows of the t-J
    Hamiltonian (open rows x cols lattice, nearest-neighbour hops and
    exchange) whose target row lies within w rows of the source, for the
    public (ascending bit-string) order and for a hole-major order. With R =
    64 columns a row of the block is 512 B, so w = 8000 rows is one XCD's 4 MB
    L2 and 500000 the 256 MB memory-side cache."""
    import numpy as np
    from . import models
    n = rows * cols
    cnt = _tj_rank_tables(n, nup, ndn)
    D = cnt[n][nup][ndn]
    bonds = models.square_bonds(rows, cols, periodic=False)
    rng = np.random.default_rng(seed)
    # random states of the sector: a random arrangement of holes, ups and downs
    base = [0] * (n - nup - ndn) + [1] * nup + [2] * ndn
    dp, dh = [], []
    for _ in range(samples):
        sites = rng.permutation(base)
        s = sum(int(v) << (2 * i) for i, v in enumerate(sites))
        rp, rh = tj_public_rank(s, n, nup, ndn, cnt), tj_hole_major_rank(s, n, n - nup - ndn, nup)
        for i, j in bonds:
            a, b = (s >> (2 * i)) & 3, (s >> (2 * j)) & 3
            if a == b:
                continue
            # a hop (one site empty) or an exchange (opposite spins) swaps the two sites
            t = s & ~((3 << (2 * i)) | (3 << (2 * j))) | (b << (2 * i)) | (a << (2 * j))
            dp.append(abs(tj_public_rank(t, n, nup, ndn, cnt) - rp))
            dh.append(abs(tj_hole_major_rank(t, n, n - nup - ndn, nup) - rh))
    dp, dh = np.array(dp), np.array(dh)
    return dict(D=D, entries_per_row=len(dp) / samples,
                windows={w: (float(np.mean(dp < w)), float(np.mean(dh < w))) for w in windows})


def _main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m qkrylov_amd.profiling")
    sub = ap.add_subparsers(dest="cmd", required=True)
    k = sub.add_parser("kernels")
    k.add_argument("db")
    c = sub.add_parser("counters")
    c.add_argument("fetch_db")
    c.add_argument("write_db")
    c.add_argument("--copy-bytes", type=float, required=True)
    c.add_argument("--match", default="k_apply<")
    m = sub.add_parser("pmc")
    m.add_argument("dbs", nargs="+")
    m.add_argument("--match", default="k_apply<")
    q = sub.add_parser("quad-model")
    q.add_argument("--rows", type=int, default=8)
    q.add_argument("--cols", type=int, default=4)
    q.add_argument("--nl", type=int, default=14)
    q.add_argument("--tiles", type=int, default=200)
    tj = sub.add_parser("tj-locality")
    tj.add_argument("--samples", type=int, default=1500)
    a = ap.parse_args(argv)
    if a.cmd == "tj-locality":
        r = tj_locality(samples=a.samples)
        print(f"D = {r['D']}, off-diagonal entries per row {r['entries_per_row']:.2f}")
        for w, (pub, hm) in r["windows"].items():
            print(f"within {w:7d} rows: public order {pub:.3f}, hole-major order {hm:.3f}")
        return 0
    if a.cmd == "quad-model":
        r = quad_model(a.rows, a.cols, a.nl, a.tiles)
        for variant in ("own", "masked"):
            parts = {k: round(r[(k, variant)], 1) for k in ("hionly", "inblock", "mixed")}
            print(f"{variant:7s} {parts} total {sum(parts.values()):.1f} accesses per 64-row tile")
        return 0
    if a.cmd == "pmc":
        for db in a.dbs:
            for ctr, (v, n) in sorted(counter_means(db, a.match).items()):
                print(f"{ctr:40s} {v:20.1f}   ({n} dispatches)")
        return 0
    if a.cmd == "kernels":
        tab = kernel_table(a.db)
        tot = sum(v["total_ns"] for v in tab.values())
        print("| kernel | calls | mean us | min us | max us | % time |")
        print("|---|---:|---:|---:|---:|---:|")
        for name, v in sorted(tab.items(), key=lambda kv: -kv[1]["total_ns"]):
            print(f"| `{short_name(name)}` | {v['calls']} | {v['mean_ns'] / 1e3:.1f} | {v['min_ns'] / 1e3:.1f} | "
                  f"{v['max_ns'] / 1e3:.1f} | {100.0 * v['total_ns'] / tot:.2f} |")
    else:
        r = calibrated_traffic(a.fetch_db, a.write_db, a.copy_bytes, a.match)
        for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
            v = r[ctr]
            print(f"{ctr}: {v['kib_per_dispatch']:.1f} KiB/dispatch over {v['dispatches']} dispatches; "
                  f"k_copy reads {v['copy_kib']:.1f} KiB for {a.copy_bytes:.0f} B -> factor {v['factor']:.4f}; "
                  f"{v['bytes_per_dispatch'] / 1e9:.3f} GB/dispatch")
        print(f"traffic: {r['bytes_per_dispatch'] / 1e9:.3f} GB per dispatch")
    return 0


if __name__ == "__main__":
    sys.exit(_main())
