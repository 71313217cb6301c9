"""Summarises a rocprofv3 kernel-trace database (``*_results.db``) as text.

    python -m pycpet_amd.profsummary <results.db> [--pmc]
    python -m pycpet_amd.profsummary --copy-mib 256

Prints one row per kernel: calls, mean / min / max duration and total time,
and with --pmc the mean of each collected counter per dispatch.  The profile
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
    """(kernel, counter, dispatches, mean value per dispatch) rows."""
    con = sqlite3.connect(db)
    cur = con.cursor()
    cols = [r[1] for r in cur.execute("pragma table_info(pmc_events)")]
    if {"name", "counter_name", "counter_value"} <= set(cols):
        q = ("select name, counter_name, count(*), avg(counter_value) from pmc_events "
             "group by name, counter_name order by name")
    elif {"kernel_name", "counter_name", "value"} <= set(cols):
        q = ("select kernel_name, counter_name, count(*), avg(value) from pmc_events "
             "group by kernel_name, counter_name order by kernel_name")
    else:
        con.close()
        raise SystemExit(f"unexpected pmc_events layout: {cols}")
    rows = cur.execute(q).fetchall()
    con.close()
    return rows


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
    print(f"{'kernel':72s} {'calls':>5s} {'mean_ms':>10s} {'min_ms':>10s} {'max_ms':>10s} {'total_ms':>10s}")
    for name, n, avg, mn, mx, tot in kernel_rows(a.db):
        print(f"{name[:72]:72s} {n:5d} {avg / 1e6:10.3f} {mn / 1e6:10.3f} {mx / 1e6:10.3f} {tot / 1e6:10.2f}")
    if a.pmc:
        print()
        print(f"{'kernel':72s} {'counter':24s} {'dispatches':>10s} {'mean':>16s}")
        for kname, cname, n, avg in pmc_rows(a.db):
            print(f"{kname[:72]:72s} {cname:24s} {n:10d} {avg:16.1f}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
