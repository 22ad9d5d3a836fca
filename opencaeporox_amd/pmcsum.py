"""HBM traffic per dispatch from rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE).

    python -m opencaeporox_amd.pmcsum --fetch F.csv --write W.csv \
        --copy-fetch CF.csv --copy-write CW.csv --copy-mib 1024 [--last N] [--json out.json]
        [--workload KEY]

Each CSV is the counter_collection.csv of one rocprofv3 --pmc pass.  The
copy passes profile a device-to-device copy of --copy-mib MiB (bench.py
--copy-mib); the ratio of the bytes it moves to what each counter reports
for it is the calibration factor applied to every other kernel:

    fetch_factor = copy bytes read    / FETCH_SIZE(copy)
    write_factor = copy bytes written / WRITE_SIZE(copy)

Both counters are in KiB per dispatch.  --last N averages the last N
dispatches of each kernel (the timed steps of a bench run).  Prints a table
and, with --json, writes {kernel: {dispatches, fetch_kib, write_kib,
hbm_bytes}} plus the factors.
"""
import argparse
import csv
import json
import sys


def per_dispatch(path, counter):
    """{kernel name: [value per dispatch in dispatch order]} for one counter."""
    rows = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] != counter:
                continue
            key = int(r.get("Dispatch_Id") or r.get("Correlation_Id") or 0)
            rows.setdefault(r["Kernel_Name"], {}).setdefault(key, 0.0)
            rows[r["Kernel_Name"]][key] += float(r["Counter_Value"])
    return {k: [v[d] for d in sorted(v)] for k, v in rows.items()}


def copy_factor(path, counter, mib):
    """Expected KiB / reported KiB for the largest copy dispatch in a pass."""
    best = 0.0
    for name, vals in per_dispatch(path, counter).items():
        if "copy" in name.lower():
            best = max(best, max(vals))
    if best <= 0:
        raise ValueError(f"no copy kernel with {counter} in {path}")
    return mib * 1024.0 / best


def summarize(fetch, write, ff, wf, last=0):
    F, W = per_dispatch(fetch, "FETCH_SIZE"), per_dispatch(write, "WRITE_SIZE")
    out = {}
    for name in sorted(set(F) & set(W)):
        f, w = F[name], W[name]
        if last > 0:
            f, w = f[-last:], w[-last:]
        fk, wk = sum(f) / len(f), sum(w) / len(w)
        out[name] = {"dispatches": len(f), "fetch_kib": fk, "write_kib": wk,
                     "hbm_bytes": (fk * ff + wk * wf) * 1024.0}
    return out


def main(argv=None):
    p = argparse.ArgumentParser()
    p.add_argument("--fetch", required=True)
    p.add_argument("--write", required=True)
    p.add_argument("--copy-fetch", required=True)
    p.add_argument("--copy-write", required=True)
    p.add_argument("--copy-mib", type=int, default=1024)
    p.add_argument("--last", type=int, default=0)
    p.add_argument("--json", default="")
    p.add_argument("--workload", default="", help="configuration the runs measured (stored in the JSON)")
    a = p.parse_args(argv)
    ff = copy_factor(a.copy_fetch, "FETCH_SIZE", a.copy_mib)
    wf = copy_factor(a.copy_write, "WRITE_SIZE", a.copy_mib)
    rows = summarize(a.fetch, a.write, ff, wf, a.last)
    print(f"calibration: fetch x{ff:.3f}, write x{wf:.3f} (copy of {a.copy_mib} MiB)")
    print(f"{'kernel':<72} {'n':>4} {'FETCH KiB':>12} {'WRITE KiB':>12} {'HBM GB':>9}")
    for name, r in sorted(rows.items(), key=lambda kv: -kv[1]["hbm_bytes"]):
        print(f"{name[:72]:<72} {r['dispatches']:>4} {r['fetch_kib']:>12.1f} {r['write_kib']:>12.1f} "
              f"{r['hbm_bytes'] / 1e9:>9.4f}")
    if a.json:
        with open(a.json, "w") as f:
            json.dump({"workload": a.workload, "fetch_factor": ff, "write_factor": wf, "copy_mib": a.copy_mib,
                       "kernels": rows}, f, indent=1)


if __name__ == "__main__":
    sys.exit(main())
