"""Summaries of rocprofv3 result databases (SQLite, `<name>_results.db`).

    python -m paracfd_amd.profiling trace DB            per-kernel calls, mean and total duration
    python -m paracfd_amd.profiling last DB N TEXT      mean duration of the last N dispatches of each
                                                        kernel whose name contains TEXT
    python -m paracfd_amd.profiling counters DB [TEXT]  per-kernel mean of each collected counter
                                                        (kernels whose name contains TEXT)
    python -m paracfd_amd.profiling range DB NAME       per-kernel calls and mean duration of the
                                                        dispatches inside the roctx range(s) NAME
                                                        (rocprofv3 --marker-trace)
    python -m paracfd_amd.profiling tables DB           tables and views of the database
"""
from __future__ import annotations

import collections
import sqlite3
import sys


def _short(name: str, width: int = 72) -> str:
    return name if len(name) <= width else name[:width - 3] + "..."


def kernel_table(db: str) -> list[tuple[str, int, float, float]]:
    """(kernel, calls, mean us, total ms) sorted by total time."""
    con = sqlite3.connect(db)
    cols = [r[1] for r in con.execute("pragma table_info(kernels)")]
    name_col = "kernel_name" if "kernel_name" in cols else "name"
    agg = collections.defaultdict(list)
    for name, start, end in con.execute(f"select {name_col}, start, end from kernels"):
        agg[name].append((end - start) / 1000.0)
    rows = [(k, len(v), sum(v) / len(v), sum(v) / 1000.0) for k, v in agg.items()]
    return sorted(rows, key=lambda r: -r[3])


def last_dispatches(db: str, n: int, text: str) -> dict[str, tuple[float, int]]:
    """kernel -> (mean us of its last n dispatches in start order, dispatches used)."""
    con = sqlite3.connect(db)
    cols = [r[1] for r in con.execute("pragma table_info(kernels)")]
    name_col = "kernel_name" if "kernel_name" in cols else "name"
    agg = collections.defaultdict(list)
    for name, start, end in con.execute(f"select {name_col}, start, end from kernels order by start"):
        if text in name:
            agg[name].append((end - start) / 1000.0)
    return {k: (sum(v[-n:]) / len(v[-n:]), len(v[-n:])) for k, v in agg.items()}


def _columns(con, table):
    return [r[1] for r in con.execute(f"pragma table_info('{table}')")]


def marker_ranges(db: str, name: str) -> list[tuple[int, int]]:
    """(start, end) of every roctx range called `name`: rows of the `regions` view whose JSON
    extdata carries {"message": name} (rocprofv3 --marker-trace, rocpd database)."""
    import json
    con = sqlite3.connect(db)
    out = []
    for start, end, ext in con.execute("select start, end, extdata from regions"):
        try:
            msg = json.loads(ext or "{}").get("message")
        except ValueError:
            msg = None
        if msg == name:
            out.append((start, end))
    return out


def range_table(db: str, name: str) -> list[tuple[str, int, float, float]]:
    """(kernel, calls, mean us, total ms) of the dispatches inside the ranges called `name`."""
    ranges = marker_ranges(db, name)
    if not ranges:
        raise SystemExit(f"no roctx range {name!r} in {db}")
    con = sqlite3.connect(db)
    cols = _columns(con, "kernels")
    name_col = "kernel_name" if "kernel_name" in cols else "name"
    agg = collections.defaultdict(list)
    for kname, start, end in con.execute(f"select {name_col}, start, end from kernels"):
        if any(a <= start and end <= b for a, b in ranges):
            agg[kname].append((end - start) / 1000.0)
    rows = [(k, len(v), sum(v) / len(v), sum(v) / 1000.0) for k, v in agg.items()]
    return sorted(rows, key=lambda r: -r[3])


def counter_table(db: str, text: str = "") -> dict[str, dict[str, tuple[float, int]]]:
    """kernel -> counter -> (mean per dispatch, dispatches)."""
    con = sqlite3.connect(db)
    agg = collections.defaultdict(lambda: collections.defaultdict(list))
    for kernel, counter, value in con.execute("select kernel_name, counter_name, value from counters_collection"):
        if text in kernel:
            agg[kernel][counter].append(value)
    return {k: {c: (sum(v) / len(v), len(v)) for c, v in cs.items()} for k, cs in agg.items()}


def main(argv: list[str]) -> int:
    if (len(argv) < 2 or argv[0] not in ("trace", "counters", "last", "range", "tables")
            or (argv[0] == "last" and len(argv) < 4) or (argv[0] == "range" and len(argv) < 3)):
        print(__doc__, file=sys.stderr)
        return 2
    if argv[0] == "tables":
        con = sqlite3.connect(argv[1])
        for t, kind in con.execute("select name, type from sqlite_master where type in ('table', 'view')"):
            print(f"{kind:5s} {t}: {', '.join(_columns(con, t))}")
        return 0
    if argv[0] == "range":
        rows = range_table(argv[1], argv[2])
        total = sum(r[3] for r in rows) or 1.0
        print(f"dispatches inside roctx range {argv[2]!r}")
        print(f"{'kernel':72s} {'calls':>6s} {'avg_us':>9s} {'total_ms':>9s} {'pct':>6s}")
        for name, calls, avg, tot in rows:
            print(f"{_short(name):72s} {calls:6d} {avg:9.2f} {tot:9.2f} {100 * tot / total:6.1f}")
        return 0
    if argv[0] == "last":
        for name, (mean, n) in last_dispatches(argv[1], int(argv[2]), argv[3]).items():
            print(f"{_short(name):72s} last {n:4d} dispatches: {mean:9.2f} us")
        return 0
    if argv[0] == "trace":
        rows = kernel_table(argv[1])
        total = sum(r[3] for r in rows) or 1.0
        print(f"{'kernel':72s} {'calls':>6s} {'avg_us':>9s} {'total_ms':>9s} {'pct':>6s}")
        for name, calls, avg, tot in rows:
            print(f"{_short(name):72s} {calls:6d} {avg:9.1f} {tot:9.2f} {100 * tot / total:6.1f}")
    else:
        for kernel, cs in counter_table(argv[1], argv[2] if len(argv) > 2 else "").items():
            print(_short(kernel, 110))
            for counter, (mean, n) in sorted(cs.items()):
                print(f"    {counter:32s} mean {mean:.6g} (dispatches {n})")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
