"""Peaceman wells (host logic): well descriptions packed for
ocpx_set_wells, the Peaceman well index, and the reference's OCPMatrix with
the well rows assembled from the bulk BSR and the well blocks.

PeacemanWell keeps, per well, the fields PeacemanWell::InputPerfo / Setup
(src/WellPeaceman.cpp:15-105) fill and the per-step inputs the reference
computes with extra flashes (perforation heads dG from CaldG, the injected
fluid's molar density from XiPhase, rate weights from CalFactor).
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp

from . import abi

INJ, PROD = 0, 1
BHP, RATE = 0, 1


@dataclass
class PeacemanWell:
    name: str
    type: int                    # INJ / PROD
    mode: int                    # BHP / RATE
    bhp: float                   # bar
    target: float                # tarBHP (bar) or tarRate
    cells: list                  # perforated cells, in perforation order
    WI: list                     # WI * multiplier per perforation (CONV_DARCY included)
    dG: list = None              # perforation heads (bar), zeros by default
    xi: list = None              # injectors: injected fluid's molar density per perforation
    zi: list = None              # injectors: composition over the nc components
    factor: list = field(default=None)   # rate weights per component


def pack(wells, nc):
    """(ocpx_well array, perf_cell, perf_WI, perf_dG, perf_xi) for ocpx_set_wells."""
    arr = (abi.Well * max(1, len(wells)))()
    cells, WI, dG, xi = [], [], [], []
    for k, w in enumerate(wells):
        n = len(w.cells)
        if len(w.WI) != n:
            raise ValueError(f"well {w.name}: {n} cells but {len(w.WI)} WI values")
        s = arr[k]
        s.type, s.mode, s.perf0, s.nperf = w.type, w.mode, len(cells), n
        s.bhp, s.target = w.bhp, w.target
        for i, v in enumerate(w.zi if w.zi is not None else [0.0] * nc):
            s.zi[i] = v
        for i, v in enumerate(w.factor if w.factor is not None else [0.0] * nc):
            s.factor[i] = v
        cells += list(w.cells)
        WI += list(w.WI)
        dG += list(w.dG) if w.dG is not None else [0.0] * n
        xi += list(w.xi) if w.xi is not None else [0.0] * n
    return (arr, np.array(cells, np.int64), np.array(WI, float), np.array(dG, float), np.array(xi, float))


def peaceman_wi(direction, dx, dy, dz, kx, ky, kz, radius, skin=0.0, kh=-1.0):
    """PeacemanWell::CalWI for one perforation (METRIC, CONV_DARCY included);
    direction 'x', 'y' or 'z'; dz already times NTG."""
    d = {"x": 0, "y": 1, "z": 2}[direction]
    return abi.load().ocpx_peaceman_wi(d, dx, dy, dz, kx, ky, kz, radius, skin, kh)


def xi_phase(fluid, P, T, z=None, water=False):
    """Molar density of the injected fluid at a perforation (ocpx_xi_phase):
    P bar, T degC, z over the Nc hydrocarbon components."""
    zz = None if z is None else np.ascontiguousarray(z, float)
    v = abi.load().ocpx_xi_phase(C.byref(fluid), P, T, None if zz is None else zz.ctypes.data, 1 if water else 0)
    if v < 0:
        raise ValueError("ocpx_xi_phase: bad argument")
    return v


def full_system(rowptr, col, val, rhs, wells, Bnw, Cwn, Dw, wres, b):
    """The reference's OCPMatrix and right-hand side with the well rows
    (scalar CSR of size (nb + nwell) * b).  Well w is block row nb + w; a
    perforated cell's row gets column 0 of block (n, w), the well row gets
    row 0 of block (w, n) and its diagonal block Dw; the well rhs is -wres
    in entry 0."""
    nb = len(rowptr) - 1
    nw = len(wells)
    N = (nb + nw) * b
    V = val.reshape(-1, b, b)
    r_idx, c_idx, vals = [], [], []
    for n in range(nb):
        for s in range(rowptr[n], rowptr[n + 1]):
            rr, cc = np.meshgrid(np.arange(b), np.arange(b), indexing="ij")
            r_idx.append((n * b + rr).ravel()); c_idx.append((col[s] * b + cc).ravel()); vals.append(V[s].ravel())
    p = 0
    for w, well in enumerate(wells):
        wr = (nb + w) * b
        rr, cc = np.meshgrid(np.arange(b), np.arange(b), indexing="ij")
        r_idx.append((wr + rr).ravel()); c_idx.append((wr + cc).ravel()); vals.append(Dw[w].ravel())
        for n in well.cells:
            r_idx.append(n * b + np.arange(b)); c_idx.append(np.full(b, wr)); vals.append(Bnw[p])
            r_idx.append(np.full(b, wr)); c_idx.append(n * b + np.arange(b)); vals.append(Cwn[p])
            p += 1
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(r_idx), np.concatenate(c_idx))), shape=(N, N))
    f = np.zeros(N)
    f[:nb * b] = rhs
    for w in range(nw):
        f[(nb + w) * b] = -wres[w]
    return A, f
