"""Per-kernel summary of a rocprofv3 kernel trace: the rocpd SQLite
database (RUN_results.db) or the CSV trace (RUN_kernel_trace.csv, from
--output-format csv).

    python -m opencaeporox_amd.profsum RUN_results.db|RUN_kernel_trace.csv [--last N] [--kernel SUBSTR]

Prints, per kernel name: calls, total / average / min / max duration (ns)
and the register and scratch footprint of the code object.  --last N keeps
only the last N dispatches of each kernel (e.g. the timed steps of a bench
run that also launched the kernel while preparing its input).
"""
import argparse
import csv
import sqlite3
import sys


def kernel_rows(db):
    if db.endswith(".csv"):
        with open(db) as f:
            rows = [(r["Kernel_Name"], int(r["Start_Timestamp"]), int(r["End_Timestamp"]), int(r["VGPR_Count"]),
                     int(r["Accum_VGPR_Count"]), int(r["SGPR_Count"]), int(r["Scratch_Size"]),
                     int(r["LDS_Block_Size"]), int(r["Grid_Size_X"]), int(r["Workgroup_Size_X"]))
                    for r in csv.DictReader(f)]
        rows.sort(key=lambda r: r[1])
        return rows
    con = sqlite3.connect(db)
    q = ("select name, start, end, vgpr_count, accum_vgpr_count, sgpr_count, scratch_size, "
         "lds_size, grid_x, workgroup_x from kernels order by start")
    rows = con.execute(q).fetchall()
    con.close()
    return rows


def summarize(db, last=0, substr=""):
    per = {}
    for name, s, e, v, a, sg, scr, lds, gx, wx in kernel_rows(db):
        if substr and substr not in name:
            continue
        per.setdefault(name, []).append((e - s, v, a, sg, scr, lds, gx, wx))
    out = []
    for name, lst in per.items():
        if last > 0:
            lst = lst[-last:]
        d = [x[0] for x in lst]
        v, a, sg, scr, lds, gx, wx = lst[-1][1:]
        out.append({"name": name, "calls": len(d), "total_ns": sum(d), "avg_ns": sum(d) / len(d),
                    "min_ns": min(d), "max_ns": max(d), "vgpr": v, "agpr": a, "sgpr": sg,
                    "scratch": scr, "lds": lds, "grid": gx, "block": wx})
    out.sort(key=lambda r: -r["total_ns"])
    return out


def main(argv=None):
    p = argparse.ArgumentParser()
    p.add_argument("db")
    p.add_argument("--last", type=int, default=0)
    p.add_argument("--kernel", default="")
    a = p.parse_args(argv)
    rows = summarize(a.db, a.last, a.kernel)
    tot = sum(r["total_ns"] for r in rows) or 1
    print(f"{'kernel':<72} {'calls':>5} {'avg ms':>10} {'min ms':>10} {'max ms':>10} {'%':>6} "
          f"{'vgpr':>4} {'agpr':>4} {'scratch':>7} {'grid':>9} {'block':>5}")
    for r in rows:
        print(f"{r['name'][:72]:<72} {r['calls']:>5} {r['avg_ns'] / 1e6:>10.4f} {r['min_ns'] / 1e6:>10.4f} "
              f"{r['max_ns'] / 1e6:>10.4f} {100 * r['total_ns'] / tot:>6.2f} {r['vgpr']:>4} {r['agpr']:>4} "
              f"{r['scratch']:>7} {r['grid']:>9} {r['block']:>5}")


if __name__ == "__main__":
    sys.exit(main())
