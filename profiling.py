"""Summaries of rocprofv3 result databases for the pixel pass.

The pixel-pass kernel is VALU-bound, and its HBM traffic is read from two
``rocprofv3 --pmc`` runs, FETCH_SIZE then WRITE_SIZE.  Each byte counter is
scaled by what it reads for a plain copy of known size: on MI355X a 1 GiB
device copy reads FETCH_SIZE = 0.5 GiB and WRITE_SIZE = 1 GiB.

Usage (each ``<dir>`` is a rocprofv3 ``-d`` output directory):

    python profiling.py stats <dir> [--limit 12]
    python profiling.py api <dir> [--match Graph]
    python profiling.py counters <dir> [--min-ms 10]
    python profiling.py traffic --fetch <dir> --write <dir> \\
        --cal-fetch <dir> --cal-write <dir> --launch K,W,P,G --command "..."

``traffic`` prints the summary that ``bench.py`` reads for ``roofline.traffic``.
It carries a digest of the kernel sources' code (``forcepho_amd/csrc/kernels.hip`` and
the two headers it compiles with), so a summary taken on other kernel code
is recognised as stale.
"""

import argparse
import collections
import glob
import hashlib
import os
import re
import sqlite3
import statistics

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "forcepho_amd", "csrc")
KERNEL_SOURCE = os.path.join(CSRC, "kernels.hip")
# the sources whose code fixes the pixel-pass machine code
KERNEL_SOURCES = [KERNEL_SOURCE, os.path.join(CSRC, "kernels.h"),
                  os.path.join(CSRC, "fpho_internal.h")]
COPY_BYTES = 1 << 30   # bench.py --copy-calibration copies 1 GiB per dispatch

__all__ = ["kernel_source_digest", "kernel_stats", "api_stats", "counter_medians",
           "read_summary"]


def _code_only(text):
    """C/C++ source without comments, whitespace runs collapsed."""
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    return " ".join(text.split())


def kernel_source_digest(paths=None):
    """First 16 hex digits of the SHA-256 of the kernel sources' code
    (``kernels.hip``, ``kernels.h``, ``fpho_internal.h``; comments and
    whitespace do not count, so a profile stays valid across comment edits)."""
    h = hashlib.sha256()
    for p in paths or KERNEL_SOURCES:
        with open(p) as f:
            h.update(_code_only(f.read()).encode())
        h.update(b"\0")
    return h.hexdigest()[:16]


def _dbs(path):
    if os.path.isfile(path):
        return [path]
    return sorted(glob.glob(os.path.join(path, "**", "*.db"), recursive=True))


def kernel_stats(path, limit=12):
    """Rows (name, calls, total_us, avg_us, percent) of the top_kernels view."""
    rows = []
    for db in _dbs(path):
        con = sqlite3.connect(db)
        rows += list(con.execute("select name, total_calls, total_duration, average, "
                                 "percentage from top_kernels limit ?", (limit,)))
        con.close()
    return rows


def api_stats(path, match=None):
    """Rows (name, calls, total_us, avg_us) of the traced host API calls (the
    ``regions`` view of a ``--hip-runtime-trace`` run), largest total first;
    with ``match``, only the names containing it."""
    acc = collections.defaultdict(lambda: [0, 0.0])
    for db in _dbs(path):
        con = sqlite3.connect(db)
        for name, dur in con.execute("select name, duration from regions"):
            if match is None or match in name:
                acc[name][0] += 1
                acc[name][1] += (dur or 0) / 1e3
        con.close()
    rows = [(n, c, t, t / c) for n, (c, t) in acc.items()]
    return sorted(rows, key=lambda r: -r[2])


def counter_medians(path, min_ms=0.0):
    """{(kernel, counter): (dispatches, median value, median duration ns)} over
    dispatches that ran at least ``min_ms`` milliseconds."""
    vals = collections.defaultdict(list)
    durs = collections.defaultdict(list)
    for db in _dbs(path):
        con = sqlite3.connect(db)
        for k, d, c, v, du in con.execute("select kernel_name, dispatch_id, counter_name, "
                                          "value, duration from counters_collection"):
            if du is not None and du < min_ms * 1e6:
                continue
            name = k.split("(")[0]
            vals[(name, c)].append(v)
            durs[(name, c)].append(du)
        con.close()
    return {key: (len(v), statistics.median(v), statistics.median(durs[key]))
            for key, v in vals.items()}


def _one(path, counter, pattern, min_ms):
    hits = [(k, v) for k, v in counter_medians(path, min_ms).items()
            if k[1] == counter and pattern in k[0]]
    if len(hits) != 1:
        raise SystemExit(f"{path}: expected one {pattern!r} kernel with {counter}, found "
                         f"{[h[0] for h in hits]}")
    return hits[0]


def read_summary(path):
    """The key: value lines of a traffic summary (comments skipped)."""
    kv = {}
    with open(path) as f:
        for line in f:
            if line.startswith("#") or ":" not in line:
                continue
            k, v = line.split(":", 1)
            kv[k.strip()] = v.strip()
    return kv


def _traffic(a):
    (kf, (nf, fetch_kib, dur_f)) = _one(a.fetch, "FETCH_SIZE", a.kernel, a.min_ms)
    (kw, (nw, write_kib, dur_w)) = _one(a.write, "WRITE_SIZE", a.kernel, a.min_ms)
    (_, (_, cal_f, _)) = _one(a.cal_fetch, "FETCH_SIZE", "copy", 0.0)
    (_, (_, cal_w, _)) = _one(a.cal_write, "WRITE_SIZE", "copy", 0.0)
    f_scale = COPY_BYTES / (cal_f * 1024.0)
    w_scale = COPY_BYTES / (cal_w * 1024.0)
    traffic = fetch_kib * 1024.0 * f_scale + write_kib * 1024.0 * w_scale
    print("# HBM traffic of one pixel-pass launch, from rocprofv3 byte counters")
    print(f"# command: {a.command}")
    print(f"kernel: {kf}")
    print(f"launch: {a.launch}")
    print(f"kernel_source_sha256_16: {kernel_source_digest()}")
    print(f"dispatches_fetch_run: {nf} (median duration {dur_f / 1e6:.3f} ms)")
    print(f"dispatches_write_run: {nw} (median duration {dur_w / 1e6:.3f} ms)")
    print(f"fetch_size_kib_per_dispatch: {fetch_kib:.1f}")
    print(f"write_size_kib_per_dispatch: {write_kib:.1f}")
    print(f"copy_calibration_fetch_kib: {cal_f:.1f} (1 GiB copy; scale {f_scale:.4f})")
    print(f"copy_calibration_write_kib: {cal_w:.1f} (1 GiB copy; scale {w_scale:.4f})")
    print(f"traffic_bytes_per_launch: {traffic:.6e}")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("path")
    s.add_argument("--limit", type=int, default=12)
    ap_ = sub.add_parser("api")
    ap_.add_argument("path")
    ap_.add_argument("--match", default=None)
    c = sub.add_parser("counters")
    c.add_argument("path")
    c.add_argument("--min-ms", type=float, default=0.0)
    t = sub.add_parser("traffic")
    for opt in ("--fetch", "--write", "--cal-fetch", "--cal-write", "--launch", "--command"):
        t.add_argument(opt, required=True)
    t.add_argument("--kernel", default="eval_")
    t.add_argument("--min-ms", type=float, default=10.0)
    a = ap.parse_args(argv)
    if a.cmd == "stats":
        print(f"{'avg_us':>12s} {'calls':>6s} {'total_us':>14s} {'pct':>6s}  kernel")
        for name, calls, tot, avg, pct in kernel_stats(a.path, a.limit):
            print(f"{avg:12.3f} {calls:6d} {tot:14.1f} {pct:6.2f}  {name[:110]}")
    elif a.cmd == "api":
        print(f"{'avg_us':>12s} {'calls':>6s} {'total_us':>14s}  call")
        for name, calls, tot, avg in api_stats(a.path, a.match):
            print(f"{avg:12.3f} {calls:6d} {tot:14.1f}  {name}")
    elif a.cmd == "counters":
        for (k, cn), (n, med, dur) in sorted(counter_medians(a.path, a.min_ms).items()):
            print(f"{k[:70]:70s} {cn:28s} n={n:5d} median={med:.6e} dur_ms={dur / 1e6:.3f}")
    else:
        _traffic(a)


if __name__ == "__main__":
    main()
