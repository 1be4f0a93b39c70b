"""HBM traffic per kernel launch from two rocprofv3 counter runs of bench.py.

The L2 (TCC) counts its requests to memory by size, so the bytes follow
without a calibration factor:

  read  = 32 * TCC_EA0_RDREQ_32B + 64 * TCC_EA0_RDREQ_64B + 128 * TCC_EA0_RDREQ_128B
  write = 64 * TCC_EA0_WRREQ_64B + 32 * (TCC_EA0_WRREQ - TCC_EA0_WRREQ_64B)

(rocprofv3's FETCH_SIZE counts a 128-byte read as 64 bytes on this chip: a
copy of N bytes reads FETCH_SIZE = N/2.)  The library's copy kernel, run
first by `bench.py --copy-calibration MIB`, checks the formula: it must read
and write exactly MIB MiB per launch.

Usage (the two counter runs, then the summary):

  rocprofv3 --pmc TCC_EA0_RDREQ_32B_sum TCC_EA0_RDREQ_64B_sum TCC_EA0_RDREQ_128B_sum \\
      --output-format csv -d OUT/rd -o run -- python3 bench.py <args> --copy-calibration 256 > OUT/rd.json
  rocprofv3 --pmc TCC_EA0_WRREQ_sum TCC_EA0_WRREQ_64B_sum \\
      --output-format csv -d OUT/wr -o run -- python3 bench.py <args> --copy-calibration 256 > OUT/wr.json
  python -m gsbn_amd.traffic --read OUT/rd --write OUT/wr --bench OUT/rd.json \\
      --workload 100x100 --out profiles/traffic_100x100.json

Each kernel of the bench line's per-kernel table is averaged over its last
`launches` dispatches (the replayed window whose algorithmic bytes the line
reports).  bench.py reads the summary back as `roofline.traffic` when the
workload, the kernel and its plan match.
"""

from __future__ import annotations

import argparse
import collections
import csv
import glob
import json
import os
import re

READ = {"TCC_EA0_RDREQ_32B_sum": 32, "TCC_EA0_RDREQ_64B_sum": 64, "TCC_EA0_RDREQ_128B_sum": 128}
WRITES = "TCC_EA0_WRREQ_sum"
WRITES_64B = "TCC_EA0_WRREQ_64B_sum"
COPY_KERNEL = "copy_kernel"


def kernel_short_name(name: str) -> str:
    m = re.search(r"::(\w+)(<[^>]*>)?\(", name)
    return m.group(1) if m else name.split("(")[0].strip()


def per_dispatch(run_dir: str) -> dict:
    """{kernel: {counter: [value per dispatch, in dispatch order]}}"""
    files = glob.glob(os.path.join(run_dir, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit("no counter_collection.csv under " + run_dir)
    rows = collections.defaultdict(lambda: collections.defaultdict(dict))
    for f in files:
        for r in csv.DictReader(open(f)):
            k = kernel_short_name(r["Kernel_Name"])
            rows[k][int(r["Dispatch_Id"])][r["Counter_Name"]] = float(r["Counter_Value"])
    out = {}
    for k, by_id in rows.items():
        ids = sorted(by_id)
        names = set().union(*(by_id[i].keys() for i in ids))
        out[k] = {c: [by_id[i].get(c, 0.0) for i in ids] for c in names}
    return out


def read_bytes(c: dict, i: int) -> float:
    return sum(w * c[n][i] for n, w in READ.items() if n in c)


def write_bytes(c: dict, i: int) -> float:
    n64 = c[WRITES_64B][i]
    return 64.0 * n64 + 32.0 * (c[WRITES][i] - n64)


def last_line(path: str) -> dict:
    lines = [l for l in open(path).read().splitlines() if l.strip().startswith("{")]
    if not lines:
        raise SystemExit("no JSON line in " + path)
    return json.loads(lines[-1])


def summarize(read_dir: str, write_dir: str, bench: dict, workload: str, copy_mib: int) -> dict:
    rd = per_dispatch(read_dir)
    wr = per_dispatch(write_dir)
    roof = bench["roofline"]
    out = {"workload": workload, "window": roof.get("window"), "plan": roof.get("step_plan"),
           "deferred_columns": roof.get("deferred_columns"),
           "method": "TCC_EA0 read requests by size (32/64/128 B) and write requests (32/64 B)"}
    if COPY_KERNEL in rd and COPY_KERNEL in wr:
        n = len(rd[COPY_KERNEL][next(iter(READ))])
        rb = sum(read_bytes(rd[COPY_KERNEL], i) for i in range(n)) / n
        m = len(wr[COPY_KERNEL][WRITES])
        wb = sum(write_bytes(wr[COPY_KERNEL], i) for i in range(m)) / m
        want = float(copy_mib) * (1 << 20)
        out["copy_check"] = {"bytes_per_copy": want, "read_bytes": rb, "write_bytes": wb,
                             "read_ratio": rb / want, "write_ratio": wb / want}
    kernels = {}
    for name, info in roof["kernels"].items():
        if name not in rd or name not in wr:
            continue
        n = int(info["launches"])
        r = rd[name]
        w = wr[name]
        nr = len(r[next(iter(READ))])
        nw = len(w[WRITES])
        if nr < n or nw < n:
            raise SystemExit("%s: %d/%d dispatches with counters, %d in the window" % (name, nr, nw, n))
        rb = sum(read_bytes(r, i) for i in range(nr - n, nr)) / n
        wb = sum(write_bytes(w, i) for i in range(nw - n, nw)) / n
        alg = float(info["bytes_per_launch"])
        rowcol = not name.startswith("pop_kernel")
        kernels[name] = {"plan": roof.get("step_plan") if rowcol else "any",
                         "deferred_columns": roof.get("deferred_columns"),
                         "read_bytes_per_launch": int(rb), "write_bytes_per_launch": int(wb),
                         "hbm_bytes_per_launch": int(rb + wb),
                         "algorithmic_bytes_per_launch": alg,
                         "ratio": (rb + wb) / alg if alg > 0 else None,
                         "launches_averaged": n}
    out["kernels"] = kernels
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--read", required=True, help="rocprofv3 output directory of the read-request run")
    ap.add_argument("--write", required=True, help="rocprofv3 output directory of the write-request run")
    ap.add_argument("--bench", required=True, help="bench.py output of the read-request run")
    ap.add_argument("--workload", required=True)
    ap.add_argument("--copy-mib", type=int, default=256)
    ap.add_argument("--out", required=True)
    a = ap.parse_args()
    s = summarize(a.read, a.write, last_line(a.bench), a.workload, a.copy_mib)
    with open(a.out, "w") as f:
        json.dump(s, f, indent=1)
    print(json.dumps(s, indent=1))


if __name__ == "__main__":
    main()
