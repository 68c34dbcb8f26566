"""Counter bookkeeping for the RHS roofline: HBM traffic per launch from rocprofv3 counter runs, and the lookup the
benchmark uses to report it.

Traffic of a kernel = FETCH_SIZE x a + WRITE_SIZE x b per dispatch (KiB), where a and b are what the same counters
read for the library's streaming copy (``k_probe_copy``, known bytes read and written).  A traffic record holds the
hash of the HIP sources it was measured on (``source_hash``); the benchmark reports a record only for the tree
that built the library, so an edited kernel never inherits an old measurement.

Command line (run on the counter outputs of one workload):

    python -m ablate_amd.rocprof --fetch DIR --write DIR --workload NAME --n 256 --strategy fused \
        --cells 16777216 --kernels k_fused [--copy-bytes 4294967296]

prints one JSON record (append it to ``profiles/traffic.json``) and a markdown table of the per-kernel figures.
"""
from __future__ import annotations

import argparse
import csv
import glob
import hashlib
import json
import os
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
TRAFFIC_FILE = ROOT / "profiles" / "traffic.json"
CALIBRATION_KERNEL = "k_probe_copy"


def source_hash() -> str:
    """SHA-256 (first 16 hex digits) of the device sources: every .hip / .hpp file of the native library."""
    h = hashlib.sha256()
    for p in sorted((HERE / "csrc").glob("*")):
        if p.suffix in (".hip", ".hpp"):
            h.update(p.name.encode())
            h.update(p.read_bytes())
    return h.hexdigest()[:16]


def _counter_csvs(d: str) -> list[str]:
    return sorted(glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True))


def counter_per_dispatch(d: str, counter: str) -> dict[str, list[float]]:
    """Values of `counter` per dispatch, grouped by kernel name, from a rocprofv3 --pmc output directory."""
    out: dict[str, list[float]] = {}
    for f in _counter_csvs(d):
        with open(f, newline="") as fh:
            for row in csv.DictReader(fh):
                if row.get("Counter_Name") != counter:
                    continue
                out.setdefault(row["Kernel_Name"], []).append(float(row["Counter_Value"]))
    if not out:
        raise FileNotFoundError(f"no {counter} rows under {d}")
    return out


def _match(values: dict[str, list[float]], key: str) -> list[float]:
    hits = [v for name, v in values.items() if key in name]
    if not hits:
        raise KeyError(f"kernel {key!r} not in the counter output ({sorted(values)})")
    return [x for v in hits for x in v]


def _steady(vals: list[float]) -> float:
    """Median over the dispatches (the first launches of a kernel can include cold-cache effects)."""
    s = sorted(vals)
    return s[len(s) // 2]


def traffic_record(fetch_dir: str, write_dir: str, workload: str, n: int, strategy: str, cells: int,
                   kernels: list[str], copy_bytes: int = 4 << 30) -> dict:
    fetch = counter_per_dispatch(fetch_dir, "FETCH_SIZE")
    write = counter_per_dispatch(write_dir, "WRITE_SIZE")
    kib = 1024.0
    a = copy_bytes / (_steady(_match(fetch, CALIBRATION_KERNEL)) * kib)
    b = copy_bytes / (_steady(_match(write, CALIBRATION_KERNEL)) * kib)
    per_kernel = {}
    total = 0.0
    for k in kernels:
        fr = _steady(_match(fetch, k)) * kib * a
        wr = _steady(_match(write, k)) * kib * b
        per_kernel[k] = {"read_bytes": round(fr), "write_bytes": round(wr),
                         "read_B_per_cell": round(fr / cells, 1), "write_B_per_cell": round(wr / cells, 1)}
        total += fr + wr
    return {"workload": workload, "n": n, "strategy": strategy, "source_hash": source_hash(),
            "fetch_scale": round(a, 4), "write_scale": round(b, 4), "kernels": per_kernel,
            "traffic_bytes": round(total), "B_per_cell": round(total / cells, 1)}


def lookup(workload: str, n: int, strategy: str) -> dict | None:
    """The committed traffic record for this workload measured on the current device sources, or None."""
    if not TRAFFIC_FILE.exists():
        return None
    h = source_hash()
    for rec in json.loads(TRAFFIC_FILE.read_text()):
        if rec["workload"] == workload and rec["n"] == n and rec["strategy"] == strategy and rec["source_hash"] == h:
            return rec
    return None


def main() -> None:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--fetch", required=True)
    ap.add_argument("--write", required=True)
    ap.add_argument("--workload", required=True)
    ap.add_argument("--n", type=int, required=True)
    ap.add_argument("--strategy", required=True)
    ap.add_argument("--cells", type=int, required=True)
    ap.add_argument("--kernels", nargs="+", required=True)
    ap.add_argument("--copy-bytes", type=int, default=4 << 30)
    args = ap.parse_args()
    rec = traffic_record(args.fetch, args.write, args.workload, args.n, args.strategy, args.cells, args.kernels,
                         args.copy_bytes)
    print(json.dumps(rec))
    print()
    print("| kernel | read B/cell | write B/cell | GB |")
    print("|---|---|---|---|")
    for k, v in rec["kernels"].items():
        print(f"| `{k}` | {v['read_B_per_cell']} | {v['write_B_per_cell']} | "
              f"{(v['read_bytes'] + v['write_bytes']) / 1e9:.3f} |")
    print(f"| **total** | | | **{rec['traffic_bytes'] / 1e9:.3f}** ({rec['B_per_cell']} B/cell) |")


if __name__ == "__main__":
    main()
