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

    python -m opencaeporox_amd.pmcsum --counters A.csv [B.csv ...] [--last N]
        [--match SUBSTR,SUBSTR] [--trace kernel_stats.csv] [--json out.json]

prints every counter of the given passes per kernel (the mean of the last N
dispatches) with the derived figures that need them: FP64 VALU operations
(64 lanes x (2 FMA + ADD + MUL + TRANS) wave instructions, an upper bound:
inactive lanes count too), VALU lane utilisation (SQ_THREAD_CYCLES_VALU /
(64 SQ_ACTIVE_INST_VALU)), the L2 hit rate (TCC_HIT / (TCC_HIT + TCC_MISS))
and, with a --kernel-trace --stats table of the same command, FP64 GFLOP/s
over the kernel's mean duration.
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


def counter_means(paths, last=0):
    """{kernel: {counter: mean of the last `last` dispatches}} over several passes."""
    out = {}
    for path in paths:
        names = set()
        with open(path) as f:
            for r in csv.DictReader(f):
                names.add(r["Counter_Name"])
        for c in sorted(names):
            for k, vals in per_dispatch(path, c).items():
                v = vals[-last:] if last > 0 else vals
                out.setdefault(k, {})[c] = sum(v) / len(v)
                out[k]["dispatches"] = len(vals)
    return out


def trace_means(path):
    """{kernel: mean duration in ns} from a rocprofv3 kernel_stats.csv."""
    out = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            out[r["Name"]] = float(r["AverageNs"])
    return out


def derived(c, ns=None):
    d = {}
    if all(k in c for k in ("SQ_INSTS_VALU_FMA_F64", "SQ_INSTS_VALU_ADD_F64", "SQ_INSTS_VALU_MUL_F64")):
        ops = 64 * (2 * c["SQ_INSTS_VALU_FMA_F64"] + c["SQ_INSTS_VALU_ADD_F64"] + c["SQ_INSTS_VALU_MUL_F64"]
                    + c.get("SQ_INSTS_VALU_TRANS_F64", 0))
        d["fp64_ops_upper"] = ops
        if ns:
            d["fp64_gflops_upper"] = ops / ns
    if c.get("SQ_ACTIVE_INST_VALU"):
        d["valu_lane_util"] = c.get("SQ_THREAD_CYCLES_VALU", 0) / (64 * c["SQ_ACTIVE_INST_VALU"])
    if "TCC_HIT_sum" in c and "TCC_MISS_sum" in c and c["TCC_HIT_sum"] + c["TCC_MISS_sum"] > 0:
        d["l2_hit_rate"] = c["TCC_HIT_sum"] / (c["TCC_HIT_sum"] + c["TCC_MISS_sum"])
    if ns:
        d["mean_ms"] = ns / 1e6
    return d


def counters_main(a):
    rows = counter_means(a.counters, a.last)
    ns = trace_means(a.trace) if a.trace else {}
    pats = [m for m in a.match.split(",") if m]
    res = {}
    for k in sorted(rows):
        if pats and not any(m in k for m in pats):
            continue
        res[k] = dict(rows[k])
        res[k].update(derived(rows[k], ns.get(k)))
    for k, r in res.items():
        print(k[:110])
        for c in sorted(r):
            v = r[c]
            print(f"    {c:<34} {v:.6g}" if isinstance(v, float) else f"    {c:<34} {v}")
    if a.json:
        with open(a.json, "w") as f:
            json.dump(res, f, indent=1)
    return 0


def main(argv=None):
    p = argparse.ArgumentParser()
    p.add_argument("--counters", nargs="+", default=None)
    p.add_argument("--match", default="")
    p.add_argument("--trace", default="")
    p.add_argument("--fetch")
    p.add_argument("--write")
    p.add_argument("--copy-fetch")
    p.add_argument("--copy-write")
    p.add_argument("--copy-mib", type=int, default=1024)
    p.add_argument("--last", type=int, default=0)
    p.add_argument("--json", default="")
    p.add_argument("--workload", default="", help="configuration the runs measured (stored in the JSON)")
    a = p.parse_args(argv)
    if a.counters:
        return counters_main(a)
    if not (a.fetch and a.write and a.copy_fetch and a.copy_write):
        p.error("--fetch, --write, --copy-fetch and --copy-write are needed without --counters")
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
