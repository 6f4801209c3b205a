"""Summarise rocprofv3 SQLite output (rocpd databases) as text.

    python -m deorbit_amd.rocpd_summary kernels <results.db> [--match NAME]
    python -m deorbit_amd.rocpd_summary counters <results.db> [--match NAME]

``kernels`` prints per-kernel dispatch count, total and mean duration (the
``--stats`` table); ``counters`` prints, per kernel, the mean per-dispatch value of
each collected counter (``--pmc`` runs).
"""
from __future__ import annotations

import argparse
import re
import sqlite3
from collections import defaultdict


def short(name: str, width: int = 70) -> str:
    name = re.sub(r"\s+", " ", name)
    return name if len(name) <= width else name[: width - 3] + "..."


def kernels(db: str, match: str | None = None):
    con = sqlite3.connect(db)
    rows = con.execute("select name, count(*), sum(duration), avg(duration) from kernels group by name "
                       "order by sum(duration) desc").fetchall()
    total = sum(r[2] for r in rows) or 1
    out = []
    for name, cnt, tot, avg in rows:
        if match and match not in name:
            continue
        out.append((short(name), cnt, tot / 1e6, avg / 1e3, 100.0 * tot / total))
    return out


def counters(db: str, match: str | None = None):
    con = sqlite3.connect(db)
    rows = con.execute("select kernel_name, counter_name, dispatch_id, sum(value) from counters_collection "
                       "group by kernel_name, counter_name, dispatch_id").fetchall()
    acc = defaultdict(list)
    for kname, cname, _, v in rows:
        if match and match not in kname:
            continue
        acc[(short(kname), cname)].append(v)
    return [(k, c, len(v), sum(v) / len(v)) for (k, c), v in sorted(acc.items())]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["kernels", "counters"])
    ap.add_argument("db")
    ap.add_argument("--match")
    a = ap.parse_args()
    if a.what == "kernels":
        print(f"{'kernel':72s} {'calls':>6s} {'total ms':>10s} {'mean us':>10s} {'%':>6s}")
        for name, cnt, tot, avg, pct in kernels(a.db, a.match):
            print(f"{name:72s} {cnt:6d} {tot:10.3f} {avg:10.2f} {pct:6.2f}")
    else:
        print(f"{'kernel':72s} {'counter':14s} {'dispatches':>10s} {'mean value':>16s}")
        for k, c, n, v in counters(a.db, a.match):
            print(f"{k:72s} {c:14s} {n:10d} {v:16.1f}")


if __name__ == "__main__":
    main()
