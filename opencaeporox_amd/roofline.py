"""Algorithmic work of the hot path, for roofline fractions.

Flash (phase stability + split, FP64 vector arithmetic): the kernel counts,
per cell, how many times it evaluates each building block (the ITERS field:
SSM/NR iterations of stability and split, Rachford-Rice iterations,
fugacity evaluations, d ln f / dn evaluations and dense solves).  One
evaluation of a block costs, counting +, -, * and / as 1 flop, a fused
multiply-add as 2 and each exp/log/sqrt as 1:

    fugacity (PR ln phi of one phase, a_ik = sqrt(a_i a_k)(1-k_ik) cached
    per cell)                         2 Nc^2 + 16 Nc + 60
    d ln f_i / d n_k (one phase)      6 Nc^2 + 30 Nc
    dense Nc x Nc solve               Nc^3 / 3 + 2 Nc^2
    Rachford-Rice iteration           6 Nc

Assembly (HBM-bound): every byte a pass must move once, with b = nc + 1
primary unknowns per cell and np = 3 phases:

    per cell   R_c = 8 * (np (9 + 2 np + 4 nc) + 3 + nc + np (nc + 1)^2)   read
               8 (b^2 + b)                                                  written
    per conn   24 (two cell indices, transmissibility and upwind flags)     read
               16 b^2 (the two off-diagonal blocks)                         written
"""
import numpy as np

NPH = 3
# MI355X published peaks
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
SOLVE_FLAGS = 0x0FFFFFFF   # the top bits of the solve counter carry cap flags


def flash_flops_per_block(Nc):
    return {
        "fug": 2 * Nc * Nc + 16 * Nc + 60,
        "lnfn": 6 * Nc * Nc + 30 * Nc,
        "solve": Nc ** 3 / 3.0 + 2 * Nc * Nc,
        "rr": 6 * Nc,
    }


def flash_flops(iters, Nc):
    """Total algorithmic flops of a phase-equilibrium launch from the ITERS
    field (cell-major [nb, 8]: SSMsta, NRsta, SSMsp, NRsp, RR iterations,
    fugacity evaluations, dlnf/dn evaluations, dense solves)."""
    it = np.asarray(iters).reshape(-1, 8).astype(np.int64)
    w = flash_flops_per_block(Nc)
    n_rr = it[:, 4].sum()
    n_fug = it[:, 5].sum()
    n_lnfn = it[:, 6].sum()
    n_solve = (it[:, 7] & SOLVE_FLAGS).sum()
    return float(n_fug * w["fug"] + n_lnfn * w["lnfn"] + n_solve * w["solve"] + n_rr * w["rr"])


def cell_record_doubles(nc):
    return NPH * (9 + 2 * NPH + 4 * nc) + 3 + nc + NPH * (nc + 1) ** 2


def assembly_bytes(nb, nconn, nc):
    b = nc + 1
    return float(nb * (8 * cell_record_doubles(nc) + 8 * (b * b + b)) + nconn * (24 + 16 * b * b))
