"""Readers for rocprofv3 result databases (the SQLite `run_results.db` that
`rocprofv3 -d DIR -o run` writes) and the HBM-traffic calibration used by
`bench.py`.

    python -m qkrylov_amd.profiling kernels DIR/run_results.db
    python -m qkrylov_amd.profiling counters FETCH.db WRITE.db --copy-bytes N [--match k_apply<]
    python -m qkrylov_amd.profiling pmc DB [DB ...] [--match k_apply<]

`kernels` prints a per-kernel table (calls, mean/min/max duration) from a
`--kernel-trace` run; `counters` prints FETCH_SIZE / WRITE_SIZE per dispatch
(KiB) for the kernels whose name contains `--match`, scaled by what the same
counter reads for the plain copy kernel `k_copy` of known size; `pmc`
prints the mean per dispatch of every counter collected in the databases
for the matching kernels.
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
    a = ap.parse_args(argv)
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
