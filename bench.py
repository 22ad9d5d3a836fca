#!/usr/bin/env python3
"""Flash + FIM-assembly throughput on one MI355X.

Workload (per step, FP64): 1,048,576 independent cells of a 9-component
Peng-Robinson mixture plus water at 344 K, compositions Dirichlet(0.5), 30%
of the pressures within +-5% of the cell's saturation pressure (found by
bisection on the device's own phase count), the rest U(10, 28) MPa.  One
step = ResetToLastTimeStep + flash with dXsdXp + kr/Pc + rock + residual +
Jacobian assembly over every cell, i.e. one IsoT_FIM Newton pass
(src/IsoThermalMethod.cpp:811-836, 1016-1269) restarted from the same state
so every step does identical work.  Inputs are resident in HBM before the
timed region.  value = cells/s of the persistent work-queue ("batched")
flash path; the one-cell-per-lane ("plain") path is timed beside it.

Also reported: the 100x100x50 (500,000 cells, Nc = 6) assembly pass with its
HBM roofline fraction, and the C oracle timed on the host on a sample of the
same cells (cpu_baseline).

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "flash+FIM-assembly cells/s (FP64, 1 GPU) batched vs plain; % FP64/HBM roofline"


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--cells", type=int, default=1 << 20)
    p.add_argument("--Nc", type=int, default=9)
    p.add_argument("--seed", type=int, default=2)
    p.add_argument("--cpu-cells", type=int, default=32768,
                   help="oracle sample size for cpu_baseline (0 = skip)")
    p.add_argument("--assembly", type=str, default="100x100x50",
                   help="grid of the assembly measurement, '' to skip")
    p.add_argument("--modes", type=str, default="plain,batched")
    p.add_argument("--multimodel", type=int, default=256,
                   help="models of the uneven multi-model batch (256 = the full batch; 0 = skip)")
    p.add_argument("--adjoint", type=str, default="60x220x85",
                   help="grid of the adjoint (J^T lambda) measurement, '' to skip")
    p.add_argument("--cpu-procs", type=int, default=0,
                   help="host threads for the parallel CPU baseline, one contiguous cell block "
                        "each (0 = min(16, cores))")
    p.add_argument("--assembly-b7", type=int, default=1,
                   help="also measure the assembly at b = 7 (Nc = 5) on the same grid")
    p.add_argument("--peaks", type=int, default=1,
                   help="measure achievable copy bandwidth and FP64 DGEMM rate on this GPU")
    p.add_argument("--queue-order", choices=("auto", "longest-first", "natural"), default="auto",
                   help="cell order of the flash work queue: the autotuner's pick, or pinned (profiling "
                        "passes that must see the same dispatches pin it)")
    p.add_argument("--copy-mib", type=int, default=0,
                   help="only copy a buffer of this many MiB on the GPU (calibrates the "
                        "rocprofv3 FETCH_SIZE / WRITE_SIZE counters) and exit")
    p.add_argument("--wells", type=str, default="",
                   help="only run the well workload on this grid (e.g. 100x100x50: five vertical wells "
                        "through every layer, four Newton passes with the well terms, one solve) and exit")
    return p.parse_args()


def wells_run(dims):
    """Five vertical wells (BHP producers at the corners, a rate injector at
    the centre) through every layer of the heterogeneous grid: four Newton
    passes with the well terms and one device solve with the well rows.
    The kernel times come from a kernel trace of this run."""
    from opencaeporox_amd import wells as WL
    from opencaeporox_amd import workload as W
    from opencaeporox_amd.fim import BATCHED, IsoTFIM
    nx, ny, nz = (int(v) for v in dims.split("x"))
    wl = W.heterogeneous_grid(nx, ny, nz, Nc=6)
    wells = WL.corner_and_centre_wells(wl)
    dev = IsoTFIM.from_workload(wl)
    dev.SetWells(wells)
    for k in range(4):
        dev.newton_pass(1.0, BATCHED)
        dev.sync()
        t = time.perf_counter()
        dev.CalResWells(1.0, jacobian=True)
        log(f"pass {k}: well terms {1e3 * (time.perf_counter() - t):.3f} ms (host clock, with the norm copies)")
    u, it, rr = dev.LinearSolve(rtol=1e-3, maxit=50)
    log(f"solve with {len(wells)} well rows: {it} iterations, ||r||/||b|| {rr:.3e}")
    dev.close()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def flash_workload(n, Nc, seed):
    """The 1M-cell flash workload; saturation pressures from the device."""
    from opencaeporox_amd import workload as W
    from opencaeporox_amd.fim import BATCHED, IsoTFIM
    rng = np.random.Generator(np.random.PCG64(seed))
    wl, z = W.flash_cells(n, Nc, seed=seed + 1000, alpha=0.5)
    near = rng.uniform(0, 1, n) < 0.3
    idx = np.where(near)[0]
    # bisection on the phase count of the near cells (device phase equilibrium)
    sub, _ = W.flash_cells(idx.size, Nc, seed=seed + 1000, alpha=0.5)
    sub.state["NI"] = wl.state["NI"].reshape(n, -1)[idx].ravel()
    sub.state["P"] = wl.state["P"][idx]
    dev = IsoTFIM.from_workload(sub)

    def nphase(P, zz, T):
        dev.set("P", P)
        dev.PhaseEquilibrium(BATCHED)
        return dev.pe_result()[0]

    t = time.perf_counter()
    psat = W.saturation_pressure(nphase, z[idx], np.full(idx.size, 344.0))
    dev.close()
    log(f"saturation pressures of {idx.size} cells in {time.perf_counter() - t:.1f} s "
        f"({np.isnan(psat).sum()} without a two-phase region)")
    P = rng.uniform(100.0, 280.0, n)
    ok = np.isfinite(psat)
    P[idx[ok]] = psat[ok] * (1 + rng.uniform(-0.05, 0.05, ok.sum()))
    st = W.initial_state(wl.fluid, wl.rock, wl.grid, P, z, rng, perturb=False)
    wl.state = st
    wl.name = (f"{n} independent cells, Nc={Nc} PR + H2O, T=344 K, z~Dirichlet(0.5), "
               f"{ok.sum()} cells within 5% of P_sat")
    return wl, int(ok.sum())


def time_steps(dev, mode, dt, steps, warmup, sync):
    for _ in range(warmup):
        dev.ResetToLastTimeStep()
        dev.newton_pass(dt, mode)
    sync()
    dev.timing(True)
    sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        dev.ResetToLastTimeStep()
        dev.newton_pass(dt, mode)
    sync()
    t1 = time.perf_counter()
    k = {name: dev.timing_get(i)[0] for i, name in enumerate(("pe", "tail", "krpc_rock", "res", "asm"))}
    dev.timing(False)
    return (t1 - t0) * 1e3 / steps, k


def _oracle_block(wl, lo, hi):
    """Oracle state and a connection-free grid of cells lo..hi-1."""
    from opencaeporox_amd import workload as W
    from oracle.oracle import OracleState
    nb, g, n = wl.grid.nb, wl.grid, hi - lo
    sub = W.Grid(n, 1, 1, g.v[lo:hi], g.depth[lo:hi], g.poro[lo:hi],
                 np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0))
    init = {k: np.asarray(v).reshape(nb, -1)[lo:hi].ravel() for k, v in wl.state.items()}
    return sub, OracleState(wl.fluid.nc, n, sub.v, sub.poro, sub.depth, init=init)


def _oracle_pass(o, wl, sub, st, dt):
    o.cal_flash(wl.fluid, st)
    o.cal_krpc(wl.sat, st)
    o.cal_rock(wl.rock, st)
    o.cal_res(wl.fluid, st, sub, dt)
    o.assemble(wl.fluid, st, sub, dt)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(wl, ncell, dt, procs):
    """The oracle (the C restatement of the reference loops) on the same
    workload: flash + kr/Pc + rock + residual + assembly.  MPI-rank style:
    `procs` host threads, each on its own contiguous block of cells (the C
    calls release the GIL and share no state), timed together; and one core
    on the first `ncell` cells."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle.oracle import Oracle
    o = Oracle()
    sub, st = _oracle_block(wl, 0, ncell)
    t = time.perf_counter()
    _oracle_pass(o, wl, sub, st, dt)
    sec1 = time.perf_counter() - t
    one = {"value": ncell / sec1, "cores": 1, "sample": f"first {ncell} cells, one full pass, {sec1:.1f} s"}
    per = max(1, ncell // 2)
    n = min(wl.grid.nb, per * procs)
    edges = np.linspace(0, n, procs + 1).astype(int)
    blocks = [_oracle_block(wl, edges[k], edges[k + 1]) for k in range(procs)]
    with ThreadPoolExecutor(procs) as ex:
        t = time.perf_counter()
        list(ex.map(lambda b: _oracle_pass(o, wl, b[0], b[1], dt), blocks))
        sec = time.perf_counter() - t
    return {"value": n / sec, "unit": "cells/s", "cores": procs, "kind": "port",
            "sample": f"first {n} cells of the same workload in {procs} contiguous blocks, one thread "
                      f"each (MPI-rank style), one full pass, {sec:.1f} s",
            "cpu": cpu_model(), "single_core": one}


def measured_peaks(sync):
    """Achievable rates on this GPU, beside the published peaks: a 4 GiB
    device copy (read + write bytes) and an FP64 8192^3 GEMM (2 n^3 flops)."""
    import torch
    x = torch.ones(1 << 29, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    y.copy_(x)
    sync()
    e0.record()
    for _ in range(5):
        y.copy_(x)
    e1.record()
    sync()
    gbs = 5 * 2 * x.numel() * 8 / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del x, y
    n = 8192
    a = torch.randn(n, n, dtype=torch.float64, device="cuda")
    b = torch.randn(n, n, dtype=torch.float64, device="cuda")
    c = a @ b
    sync()
    e0.record()
    for _ in range(3):
        c = a @ b
    e1.record()
    sync()
    tf = 3 * 2 * n ** 3 / (e0.elapsed_time(e1) * 1e-3) / 1e12
    del a, b, c
    torch.cuda.empty_cache()
    return {"copy_GBs": gbs, "dgemm_fp64_TFLOPs": tf,
            "method": "torch copy_ of 4 GiB (2 x bytes), torch matmul FP64 8192^3 (2 n^3 flops)"}


TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "hbm_traffic.json")


def traffic_key(a):
    """The bench configuration a committed traffic measurement must match."""
    return f"cells={a.cells} Nc={a.Nc} seed={a.seed} assembly={a.assembly}"


def measured_traffic(kernel, key):
    """HBM bytes per launch of `kernel` (a substring of its name) from the
    committed rocprofv3 --pmc summary (FETCH_SIZE and WRITE_SIZE per
    dispatch, scaled by the factors a plain copy gives), or None when that
    summary was not taken on this configuration or lacks the kernel."""
    try:
        with open(TRAFFIC_FILE) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    if d.get("workload") != key:
        return None
    for name, r in d.get("kernels", {}).items():
        if kernel in name:
            return r["hbm_bytes"]
    return None


def assembly_measure(spec, steps, warmup, sync, dt, key=None, Nc=6, solve=True):
    from opencaeporox_amd import workload as W
    from opencaeporox_amd.fim import BATCHED, IsoTFIM
    from opencaeporox_amd.roofline import PEAK_HBM_GBS, assembly_bytes
    nx, ny, nz = (int(v) for v in spec.split("x"))
    wl = W.heterogeneous_grid(nx, ny, nz, Nc=Nc)
    dev = IsoTFIM.from_workload(wl)
    dev.UpdateLastTimeStep()
    dev.autotune_flash()
    dev.ResetToLastTimeStep()
    ms, k = time_steps(dev, BATCHED, dt, steps, warmup, sync)
    nb, nconn, nc = dev.nb, dev.nconn, dev.nc
    byts = assembly_bytes(nb, nconn, nc)
    gbs = byts / (k["asm"] * 1e-3) / 1e9
    out = {"workload": wl.name, "cells": nb, "conns": nconn, "block": nc + 1,
           "ms_per_pass": ms, "cells_per_s": nb / (ms * 1e-3), "kernel_ms": k,
           "assembly_conns_per_s": nconn / (k["asm"] * 1e-3),
           "roofline": {"bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": gbs / PEAK_HBM_GBS, "traffic": measured_traffic(f"res_asm<{nc}>", key),
                        "kernel": f"res_asm<{nc}>", "algorithmic_bytes": byts}}
    dev.close()
    if solve:
        # the linear solve on the same grid with one composition everywhere
        # (per-cell random compositions give a Newton system no preconditioner
        # tried converges on)
        wl = W.heterogeneous_grid(nx, ny, nz, Nc=Nc, composition="uniform")
        dev = IsoTFIM.from_workload(wl)
        dev.newton_pass(dt, BATCHED)
        out["linear_solve"] = solve_measure(dev, sync)
        out["linear_solve"]["workload"] = wl.name
        dev.close()
    return out


def solve_measure(dev, sync, rtol=1e-3, maxit=200):
    """Device solve of the assembled system with the reference's solver
    settings (1e-3 on ||r|| / ||b||, 200 iterations): ABF decoupling +
    FGMRES(30) with the CPR preconditioner; the first call includes the
    pressure AMG's one-time set-up.  And the HBM roofline of the BSR
    matrix-vector product."""
    from opencaeporox_amd.roofline import PEAK_HBM_GBS
    sync()
    t = time.perf_counter()
    dev.LinearSolve(rtol, maxit, fetch=False)        # set-up: Krylov vectors, AMG hierarchy, colouring
    sync()
    first_ms = (time.perf_counter() - t) * 1e3
    dev.timing(True)
    _, it, rr = dev.LinearSolve(rtol, maxit, fetch=False)
    sync()
    total = dev.timing_get(7)[0]
    spmv, calls = dev.timing_get(8)
    dev.timing(False)
    nb, B = dev.nb, dev.bs
    nnzb = nb + 2 * dev.nconn
    byts = nnzb * (8 * B * B + 4) + 8 * (nb + 1) + 16 * nb * B     # blocks, columns, rowptr, x and y once
    gbs = byts / (spmv * 1e-3) / 1e9
    return {"method": "ABF decoupling + FGMRES(30), CPR: aggregation-AMG pressure V-cycle, multicolour block "
                      "Gauss-Seidel", "rtol": rtol, "maxit": maxit, "iterations": it,
            "relres": rr, "relres_of": "the decoupled system, as FASP's stop_type 1",
            "ms": total, "first_call_ms": first_ms, "spmv_calls": calls, "spmv_ms": spmv,
            "spmv_roofline": {"bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                              "frac": gbs / PEAK_HBM_GBS, "kernel": f"bsr_spmv<{B}>",
                              "algorithmic_bytes": byts}}


def adjoint_measure(spec, steps, sync, dt):
    """Jᵀλ and ∂(λ·R)/∂(trans, Vp) after a Newton pass on the channelized grid."""
    import torch
    from opencaeporox_amd import workload as W
    from opencaeporox_amd.fim import BATCHED, IsoTFIM
    from opencaeporox_amd.roofline import PEAK_HBM_GBS
    nx, ny, nz = (int(v) for v in spec.split("x"))
    wl = W.channelized_grid(nx, ny, nz, Nc=4)
    dev = IsoTFIM.from_workload(wl)
    dev.UpdateLastTimeStep()
    dev.autotune_flash()
    dev.ResetToLastTimeStep()
    nb, B, nconn = dev.nb, dev.bs, dev.nconn
    lam = torch.randn(nb * B, dtype=torch.float64, device="cuda")
    out = torch.empty_like(lam)
    dT = torch.empty(nconn, dtype=torch.float64, device="cuda")
    dV = torch.empty(nb, dtype=torch.float64, device="cuda")
    sync()
    for it in range(steps + 1):
        if it == 1:
            dev.timing(True)
            sync()
            t0 = time.perf_counter()
        dev.ResetToLastTimeStep()
        dev.newton_pass(dt, BATCHED)
        dev.jt_lambda_dev(lam.data_ptr(), out.data_ptr())
        dev.dR_dparams_dev(dt, lam.data_ptr(), dT.data_ptr(), dV.data_ptr())
    sync()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    k = {name: dev.timing_get(i)[0] for i, name in enumerate(("pe", "tail", "krpc_rock", "res", "asm", "jt", "dR"))}
    nnzb = nb + 2 * nconn
    jt_bytes = nnzb * (B * B * 8 + 4 + 4 + 4) + nb * 8 + 2 * nb * B * 8   # blocks, colind/tslot, rowptr, lambda, out
    gbs = jt_bytes / (k["jt"] * 1e-3) / 1e9
    dev.close()
    return {"workload": wl.name, "cells": nb, "conns": nconn, "block": B, "ms_per_pass": ms,
            "cells_per_s": nb / (ms * 1e-3), "kernel_ms": k,
            "jt_roofline": {"bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                            "frac": gbs / PEAK_HBM_GBS, "algorithmic_bytes": jt_bytes}}


def multimodel_measure(nmodels, steps, sync, dt):
    """Uneven batch: cell counts log-uniform 5k-200k, Nc uniform in 3..12,
    one context per Nc, each Newton step replayed from captured graphs."""
    from opencaeporox_amd import workload as W
    from opencaeporox_amd.fim import MultiModelBatch
    t = time.perf_counter()
    wls = W.uneven_batch(nmodels)
    batch = MultiModelBatch(wls)
    log(f"multi-model batch: {len(wls)} models, {batch.nb} cells, {len(batch.groups)} Nc groups, "
        f"set up in {time.perf_counter() - t:.1f} s")
    tune = batch.autotune()
    batch.UpdateLastTimeStep()
    for it in range(steps + 1):
        if it == 1:
            sync()
            t0 = time.perf_counter()
        batch.ResetToLastTimeStep()
        batch.newton_pass(dt, graph=True)
    sync()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    out = {"workload": f"{len(wls)} models, cells log-uniform 5k-200k, Nc uniform 3..12",
           "models": len(wls), "cells": batch.nb, "conns": batch.nconn, "ms_per_step": ms,
           "cells_per_s": batch.nb / (ms * 1e-3),
           "groups": {int(Nc): {"models": len(idx), "cells": int(d.nb), "lanes_per_cell": d.flash_mapping(),
                                "autotune_ms": tune[Nc]} for Nc, (d, idx) in batch.groups.items()}}
    batch.close()
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported "
              "(cells are independent: a multi-GPU run would split them into contiguous blocks)",
              file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    from opencaeporox_amd.fim import BATCHED, PLAIN, IsoTFIM
    from opencaeporox_amd.roofline import PEAK_FP64_TFLOPS, flash_flops

    def sync():
        torch.cuda.synchronize()

    if a.wells:
        wells_run(a.wells)
        return

    if a.copy_mib:
        x = torch.ones(a.copy_mib << 17, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        sync()
        log(f"copied {a.copy_mib} MiB three times")
        return

    dt = 1.0
    t = time.perf_counter()
    wl, n_near = flash_workload(a.cells, a.Nc, a.seed)
    log(f"workload ready in {time.perf_counter() - t:.1f} s: {wl.name}")
    dev = IsoTFIM.from_workload(wl)
    dev.UpdateLastTimeStep()
    out = {}
    modes = [m for m in a.modes.split(",") if m]
    tune = None
    for name in modes:
        mode = PLAIN if name == "plain" else BATCHED
        if mode == BATCHED:
            dev.ResetToLastTimeStep()
            tune = dev.autotune_flash()
            if a.queue_order != "auto":
                dev.set_flash_queue_order(a.queue_order == "longest-first")
            log(f"autotuner (ms per mapping, lanes per cell): {tune} -> {dev.flash_mapping()}")
        ms, k = time_steps(dev, mode, dt, a.steps, a.warmup, sync)
        it = dev.get("ITERS")
        fe = it.reshape(-1, 8)[:, 5]
        out[name] = {"ms_per_step": ms, "cells_per_s": dev.nb / (ms * 1e-3), "kernel_ms": k,
                     "flops": flash_flops(it, a.Nc),
                     "fugacity_evals_per_cell": {q: float(np.percentile(fe, p)) for q, p in
                                                 (("p50", 50), ("p99", 99), ("p99.9", 99.9))}
                                                | {"mean": float(fe.mean()), "max": int(fe.max())}}
        log(f"{name}: {ms:.2f} ms/step, {dev.nb / (ms * 1e-3):.3e} cells/s, kernels {k}")
    head = out.get("batched", out[modes[-1]])
    lanes = dev.flash_mapping() if "batched" in out else 1
    stages = dev.flash_stages() if "batched" in out else 0
    if "batched" not in out:
        pe_kernel = f"pe_plain<{a.Nc}>"
    elif lanes > 1:
        pe_kernel = f"pe_team_batched<{a.Nc}, {lanes}, false>"
    else:
        pe_kernel = f"pe_batched<{a.Nc}>"
    # the batched flash is one launch group: the stages, then the mapped kernel
    # the stage kernels of a context of one model (template flag false)
    pe_group = ([f"pe_sta_ssm<{a.Nc}, false>", f"pe_sta_nr<{a.Nc}, false>", f"pe_sta_join<{a.Nc}, false>"]
                + ([f"pe_sta_split<{a.Nc}, false>"] if stages >= 2 else [])
                + ([f"pe_sta_snr<{a.Nc}, false>"] if stages >= 3 else []) if stages else []) + [pe_kernel]
    group_traffic = [measured_traffic(k, traffic_key(a)) for k in pe_group]
    pe_traffic = sum(group_traffic) if all(v is not None for v in group_traffic) else None
    tflops = head["flops"] / (head["kernel_ms"]["pe"] * 1e-3) / 1e12
    line = {
        "metric": METRIC, "value": head["cells_per_s"], "unit": "cells/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": head["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": wl.name, "cells": dev.nb, "Nc": a.Nc, "near_saturation_cells": n_near,
                   "fugacity_evals_per_cell": head["fugacity_evals_per_cell"],
                   "pass": "reset + flash_der + kr/Pc + rock + residual + assembly",
                   "flash": "batched" if "batched" in out else modes[-1],
                   "lanes_per_cell": dev.flash_mapping() if "batched" in out else 1,
                   "flash_stages": stages,
                   "flash_long_evals": dev.flash_long() if "batched" in out else 0,
                   "queue_longest_first": dev.flash_queue_order() if "batched" in out else False,
                   "autotune_ms": tune},
        "roofline": {"bound": "mfma", "achieved": tflops, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": tflops / PEAK_FP64_TFLOPS, "traffic": pe_traffic,
                     "kernel": " + ".join(pe_group), "kernels": pe_group,
                     "flops_per_launch": head["flops"], "kernel_ms": head["kernel_ms"]["pe"]},
    }
    if "plain" in out and "batched" in out:
        line["plain"] = {"value": out["plain"]["cells_per_s"], "ms_per_step": out["plain"]["ms_per_step"],
                         "kernel_ms": out["plain"]["kernel_ms"]}
        line["batched_over_plain"] = out["batched"]["cells_per_s"] / out["plain"]["cells_per_s"]
        line["kernel_ms"] = out["batched"]["kernel_ms"]
    dev.close()
    if a.assembly:
        line["assembly"] = assembly_measure(a.assembly, max(3, a.steps // 2), 1, sync, dt, traffic_key(a))
        log(f"assembly: {line['assembly']['kernel_ms']}")
        if a.assembly_b7:
            line["assembly_b7"] = assembly_measure(a.assembly, max(3, a.steps // 2), 1, sync, dt, traffic_key(a),
                                                   Nc=5)
            log(f"assembly b=7: {line['assembly_b7']['kernel_ms']}")
    if a.multimodel:
        line["multimodel"] = multimodel_measure(a.multimodel, 3, sync, dt)
        log(f"multi-model: {line['multimodel']['cells_per_s']:.3e} cells/s")
    if a.adjoint:
        line["adjoint"] = adjoint_measure(a.adjoint, max(3, a.steps // 2), sync, dt)
        log(f"adjoint: {line['adjoint']['kernel_ms']}")
    if a.peaks:
        pk = measured_peaks(sync)
        line["measured_peaks"] = pk
        line["roofline"]["frac_of_measured"] = line["roofline"]["achieved"] / pk["dgemm_fp64_TFLOPs"]
        if "assembly" in line:
            r = line["assembly"]["roofline"]
            r["frac_of_measured"] = r["achieved"] / pk["copy_GBs"]
        log(f"measured peaks: {pk}")
    if a.cpu_cells > 0:
        procs = a.cpu_procs or min(16, os.cpu_count() or 1)
        line["cpu_baseline"] = cpu_baseline(wl, min(a.cpu_cells, wl.grid.nb), dt, procs)
        log(f"cpu baseline: {line['cpu_baseline']}")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
