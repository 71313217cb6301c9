"""Summarises a rocprofv3 kernel-trace database (``*_results.db``) as text.

    python -m pycpet_amd.profsummary <results.db> [--pmc]
    python -m pycpet_amd.profsummary --copy-mib 256

Prints one row per kernel: calls, mean / min / max duration and total time,
and with --pmc each collected counter per dispatch (summed over the
counter's instances).  The profile
summaries kept with the repository are produced with this.

--copy-mib N runs three plain device copies of N MiB (torch copy_) and
nothing else: run under `rocprofv3 --pmc FETCH_SIZE` (then WRITE_SIZE), it
gives the counter's reading for a known number of bytes moved, the scale
that turns a kernel's counter reading into HBM bytes.
"""
from __future__ import annotations

import argparse
import sqlite3
import sys


def kernel_rows(db: str):
    con = sqlite3.connect(db)
    cur = con.cursor()
    rows = cur.execute(
        "select name, count(*), avg(end-start), min(end-start), max(end-start), "
        "sum(end-start) from kernels group by name order by sum(end-start) desc").fetchall()
    con.close()
    return rows


def pmc_rows(db: str):
    """(kernel, counter, dispatches, counter instances per dispatch, value
    per dispatch) rows; the value per dispatch is the sum over the counter's
    instances (XCDs, shader engines), averaged over the dispatches."""
    con = sqlite3.connect(db)
    cur = con.cursor()
    cols = [r[1] for r in cur.execute("pragma table_info(pmc_events)")]
    if {"name", "counter_name", "counter_value", "dispatch_id"} <= set(cols):
        q = ("select name, counter_name, count(distinct dispatch_id), count(*), "
             "sum(counter_value) from pmc_events group by name, counter_name order by name")
    elif {"kernel_name", "counter_name", "value", "dispatch_id"} <= set(cols):
        q = ("select kernel_name, counter_name, count(distinct dispatch_id), count(*), "
             "sum(value) from pmc_events group by kernel_name, counter_name order by kernel_name")
    else:
        con.close()
        raise SystemExit(f"unexpected pmc_events layout: {cols}")
    rows = [(k, c, nd, n / nd, tot / nd) for k, c, nd, n, tot in cur.execute(q).fetchall()]
    con.close()
    return rows


def short_name(name: str, width: int = 72) -> str:
    """A kernel's name without its namespace, return type and argument list,
    with HIP's vector types spelled float4 / double4."""
    s = name.replace("void ", "", 1) if name.startswith("void ") else name
    s = s.replace("cpet::(anonymous namespace)::", "")
    s = s.replace("HIP_vector_type<float, 4u>", "float4").replace("HIP_vector_type<double, 4u>",
                                                                    "double4")
    depth = 0
    for i, ch in enumerate(s):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            s = s[:i]
            break
    return s[:width]


def copy_calibration(mib: int) -> None:
    import torch

    n = (mib << 20) // 4
    a = torch.empty(n, dtype=torch.float32, device="cuda")
    a.fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print("copied", n * 4, "bytes x3", flush=True)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("db", nargs="?")
    ap.add_argument("--pmc", action="store_true")
    ap.add_argument("--copy-mib", type=int, default=0)
    a = ap.parse_args(argv)
    if a.copy_mib:
        copy_calibration(a.copy_mib)
        return 0
    if not a.db:
        ap.error("a results database is required")
    print(f"{'kernel':60s} {'calls':>5s} {'mean_ms':>10s} {'min_ms':>10s} {'max_ms':>10s} {'total_ms':>10s}")
    for name, n, avg, mn, mx, tot in kernel_rows(a.db):
        print(f"{short_name(name, 60):60s} {n:5d} {avg / 1e6:10.3f} {mn / 1e6:10.3f} "
              f"{mx / 1e6:10.3f} {tot / 1e6:10.2f}")
    if a.pmc:
        print()
        print(f"{'kernel':60s} {'counter':24s} {'dispatches':>10s} {'instances':>9s} "
              f"{'per_dispatch':>16s}")
        for kname, cname, nd, inst, per in pmc_rows(a.db):
            print(f"{short_name(kname, 60):60s} {cname:24s} {nd:10d} {inst:9.0f} {per:16.1f}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
