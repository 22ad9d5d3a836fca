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


PSURF, TSURF = 1.01325, 20.0          # METRIC standard conditions (include/OCPUnits.hpp:97-99)


def rate_factors(fluid, qi, weights, Psurf=PSURF, Tsurf=TSURF):
    """Rate-equation weights of producers, PeacemanWell::CalFactor
    (src/WellPeaceman.cpp:654-718) for compositional fluids: the produced
    moles qi [nwell, nc] are flashed at (Psurf, Tsurf) on the device
    (CalVStd), the surface volumes of oil, gas and water are weighted by
    weights [nwell, 3] (prodPhaseWeight: ORAT (1, 0, 0), GRAT (0, 1, 0),
    WRAT (0, 0, 1), LRAT (1, 0, 1)), and every component gets
    factor = qv / qt.  METRIC: the surface-unit conversions are 1 and water's
    surface volume is its moles (stdVw = 1).  Requires qi >= 0 with
    sum > 1e-8 (the reference then falls back to perforation mobilities)."""
    from .fim import PLAIN, IsoTFIM
    qi = np.atleast_2d(np.asarray(qi, float))
    weights = np.atleast_2d(np.asarray(weights, float))
    nw, nc = qi.shape
    if nc != fluid.nc + 1 or weights.shape != (nw, 3):
        raise ValueError("qi must be [nwell, Nc + 1] and weights [nwell, 3]")
    qt = qi.sum(axis=1)
    if np.any(qi < 0) or np.any(qt <= 1e-8):
        raise ValueError("rate_factors needs non-negative produced moles")
    wl = surface_workload(fluid, qi, Psurf, Tsurf)
    dev = IsoTFIM.from_workload(wl)
    try:
        dev.CalFlash(PLAIN)
        S = dev.get("S")
        vf = dev.get("VF")
    finally:
        dev.close()
    return factors_from_flash(S, vf, qi, weights)


def factor_stream(qw, cell, well):
    """The moles CalFactor flashes for a producer (src/WellPeaceman.cpp:
    665-688): the well's component rates qw [nc] when they are all
    non-negative and sum above 1e-8, otherwise sum_p sum_j transj * xi_j *
    x_ij over the open phases of the perforated cells (cell_state)."""
    qw = np.asarray(qw, float)
    if qw.sum() > 1e-8 and np.all(qw >= 0):
        return qw.copy()
    ex = cell["PHASEEXIST"]
    tj = np.where(ex, np.asarray(well.WI)[:, None] * cell["KR"] / np.where(ex, cell["MU"], 1.0), 0.0)
    out = np.zeros_like(qw)
    for p in range(len(well.cells)):
        for j in range(3):
            if ex[p, j]:
                out += tj[p, j] * cell["XI"][p, j] * cell["XIJ"][p, j]
    return out


def surface_workload(fluid, qi, Psurf=PSURF, Tsurf=TSURF):
    """One cell per well holding the produced moles at (Psurf, Tsurf)."""
    from . import workload as W
    nw, nc = qi.shape
    fl = abi.Fluid.from_buffer_copy(fluid)
    fl.skipPSA = 0
    grid = W.Grid(nw, 1, 1, np.ones(nw), np.zeros(nw), np.full(nw, 0.2), np.zeros(0, np.int32),
                  np.zeros(0, np.int32), np.zeros(0))
    st = {"P": np.full(nw, float(Psurf)), "T": np.full(nw, float(Tsurf)), "NI": qi.ravel().copy(),
          "LNI": qi.ravel().copy(), "PHASEEXIST": np.tile(np.array([0, 0, 1], np.uint32), nw),
          "S": np.tile(np.array([0.0, 0.0, 1.0]), nw), "XIJ": np.zeros(nw * 3 * nc)}
    st["XIJ"].reshape(nw, 3, nc)[:, 2, nc - 1] = 1.0
    return W.Workload("surface", fl, W.make_sat(), W.make_rock(), grid, st)


def factors_from_flash(S, vf, qi, weights):
    """qv / qt from a surface flash: oil and gas volumes S_j * Vf, water's
    surface volume its moles."""
    nw, nc = qi.shape
    S = np.asarray(S).reshape(nw, 3)
    vj = S[:, :2] * np.asarray(vf)[:, None]
    qv = (vj * weights[:, :2]).sum(axis=1) + qi[:, nc - 1] * weights[:, 2]
    return qv / qi.sum(axis=1)


class SegmentFlash:
    """Flash of one mixture at (P, T) on the device (a one-cell context,
    reused): returns (phase exists [3], phase volumes [3], densities [3]),
    as PVT->Flash + GetVj / GetRho in the perforation-head loops."""

    def __init__(self, fluid):
        from .fim import IsoTFIM
        nc = fluid.nc + 1
        self.wl = surface_workload(fluid, np.ones((1, nc)))
        self.dev = IsoTFIM.from_workload(self.wl)
        self.base = {k: np.array(v) for k, v in self.wl.state.items()}

    def __call__(self, P, T, Ni):
        from .fim import PLAIN
        d = self.dev
        for k in ("PHASEEXIST", "S", "XIJ"):
            d.set(k, self.base[k])
        d.set("P", np.array([float(P)]))
        d.set("T", np.array([float(T)]))
        d.set("NI", np.asarray(Ni, float))
        d.CalFlash(PLAIN)
        S, vf = d.get("S"), d.get("VF")
        return d.get("PHASEEXIST").astype(bool), S * vf[0], d.get("RHO")

    def close(self):
        self.dev.close()


def _segments(d0, d1, maxlen):
    n = int(np.ceil(abs((d1 - d0) / maxlen)))
    return n, ((d1 - d0) / n if n else 0.0)


def producer_heads(cell, well, well_depth, perf_depth, flash, maxlen=10.0):
    """Perforation heads of a producer, PeacemanWell::CalProddG01
    (src/WellPeaceman.cpp:802-923).  cell: dict of the perforated cells'
    state (P, T [degC], PHASEEXIST [n, 3], KR, MU, XI [n, 3], XIJ [n, 3, nc],
    NI [n, nc]) in perforation order; perforation pressures are bhp + the
    current dG; the produced stream accumulates from the far end of the well
    and is flashed once per segment of at most maxlen (deck length units,
    10) with flash(P, T, Ni).  Returns the new dG; a well head between the
    perforations leaves dG as it was (the reference does the same)."""
    npf = len(well.cells)
    dG_old = np.asarray(well.dG if well.dG is not None else np.zeros(npf), float)
    perfP = well.bhp + dG_old
    nc = cell["NI"].shape[1]
    ex = cell["PHASEEXIST"]
    tj = np.where(ex, np.asarray(well.WI)[:, None] * cell["KR"] / np.where(ex, cell["MU"], 1.0), 0.0)
    g = 0.0000980665
    dGp = np.zeros(npf)
    tmpNi = np.zeros(nc)

    def stream(p):
        for j in range(3):
            if cell["PHASEEXIST"][p, j]:
                for k in range(nc):
                    tmpNi[k] += (cell["P"][p] - perfP[p]) * tj[p, j] * cell["XI"][p, j] * cell["XIJ"][p, j, k]
        if np.abs(tmpNi).sum() < 1e-8:                # no inflow yet: the cell's own moles
            tmpNi[:] = cell["NI"][p]
        return tmpNi

    def run(p, n, h, sign):
        Ni = stream(p)
        Pt = perfP[p]
        for _ in range(n):
            ex, vj, rho = flash(Pt, cell["T"][p], Ni)
            qt = vj[ex].sum()
            Pt += sign * (vj[ex] * rho[ex]).sum() / qt * h * g
        return Pt

    if well_depth <= perf_depth[0]:                  # head above the perforations
        for p in range(npf - 1, -1, -1):
            n, h = _segments(well_depth if p == 0 else perf_depth[p - 1], perf_depth[p], maxlen)
            if n == 0:                                 # the reference skips the perforation's inflow too
                continue
            dGp[p] = perfP[p] - run(p, n, h, -1.0)
        return np.cumsum(dGp)
    if well_depth >= perf_depth[-1]:                 # head below the perforations
        for p in range(npf):
            n, h = _segments(perf_depth[p], well_depth if p == npf - 1 else perf_depth[p + 1], maxlen)
            if n == 0:
                continue
            dGp[p] = run(p, n, h, +1.0) - perfP[p]
        return np.cumsum(dGp[::-1])[::-1]
    return dG_old


def injector_heads(fluid, well, well_depth, perf_depth, T, water=False, maxlen=10.0):
    """Perforation heads of an injector, PeacemanWell::CalInjdG
    (src/WellPeaceman.cpp:733-799): the injected fluid's density
    (RhoPhase: MW * xi from ocpx_xi_phase, water rhoW_std / Bw) integrated
    over segments of at most maxlen from each perforation pressure."""
    npf = len(well.cells)
    dG_old = np.asarray(well.dG if well.dG is not None else np.zeros(npf), float)
    perfP = well.bhp + dG_old
    Nc = fluid.nc
    z = np.asarray(well.zi, float)[:Nc]
    mw = sum(z[i] * fluid.MW[i] for i in range(Nc))
    g = 0.0000980665

    def rho(P):
        return fluid.rhoW_std * xi_phase(fluid, P, T, water=True) if water else mw * xi_phase(fluid, P, T, z)

    dGp = np.zeros(npf)
    if well_depth <= perf_depth[0]:
        for p in range(npf - 1, -1, -1):
            n, h = _segments(well_depth if p == 0 else perf_depth[p - 1], perf_depth[p], maxlen)
            Pt = perfP[p]
            for _ in range(n):
                Pt -= rho(Pt) * g * h
            dGp[p] = perfP[p] - Pt
        return np.cumsum(dGp)
    if well_depth >= perf_depth[-1]:
        for p in range(npf):
            n, h = _segments(perf_depth[p], well_depth if p == npf - 1 else perf_depth[p + 1], maxlen)
            Pt = perfP[p]
            for _ in range(n):
                Pt += rho(Pt) * g * h
            dGp[p] = Pt - perfP[p]
        return np.cumsum(dGp[::-1])[::-1]
    return dG_old


def cell_state(dev, cells):
    """The perforated cells' fields producer_heads and check_opt_mode read,
    from the device."""
    nb, nc = dev.nb, dev.nc
    c = np.asarray(cells)
    return {"P": dev.get("P")[c], "T": dev.get("T")[c], "PJ": dev.get("PJ").reshape(nb, 3)[c],
            "PHASEEXIST": dev.get("PHASEEXIST").reshape(nb, 3)[c].astype(bool),
            "KR": dev.get("KR").reshape(nb, 3)[c], "MU": dev.get("MU").reshape(nb, 3)[c],
            "XI": dev.get("XI").reshape(nb, 3)[c], "XIJ": dev.get("XIJ").reshape(nb, 3, nc)[c],
            "NI": dev.get("NI").reshape(nb, nc)[c]}


@dataclass
class WellLimits:
    """The control data CheckOptMode reads (WellOpt): the mode the schedule
    asks for, the rate limit, the limiting BHP (maxBHP for injectors, minBHP
    for producers), the BHP target, the producer's phase weights and the
    injected fluid's temperature (degC) and phase."""
    init_mode: int
    max_rate: float
    limit_bhp: float
    tar_bhp: float
    weights: tuple = (1.0, 0.0, 0.0)
    inj_T: float = 20.0
    water: bool = False


def check_opt_mode(fluid, cell, well, lim, flash, Psurf=PSURF, Tsurf=TSURF):
    """PeacemanWell::CheckOptMode (src/WellPeaceman.cpp:117-159) with
    CalInjRateMaxBHP / CalProdRateMinBHP (:468-530): the well's rate at its
    limiting BHP decides between its scheduled rate control and BHP control.
    cell is the perforated cells' state (cell_state); the producer's stream
    is flashed at (Psurf, Tsurf) with flash(P, T, Ni) (SegmentFlash).
    Returns (mode, bhp, q): bhp becomes tarBHP under BHP control, else
    stays; q is the rate at the limiting BHP (surface units)."""
    npf = len(well.cells)
    dG = np.asarray(well.dG if well.dG is not None else np.zeros(npf), float)
    ex = cell["PHASEEXIST"]
    tj = np.where(ex, np.asarray(well.WI)[:, None] * cell["KR"] / np.where(ex, cell["MU"], 1.0), 0.0)
    Nc = fluid.nc
    if well.type == INJ:
        zi = np.asarray(well.zi, float)[:Nc]
        q = 0.0
        for p in range(npf):
            Pperf = lim.limit_bhp + dG[p]
            tinj = 0.0
            for j in range(3):
                if ex[p, j]:
                    tinj += tj[p, j]
            xi = xi_phase(fluid, Pperf, lim.inj_T, zi, water=lim.water)
            q += tinj * xi * (Pperf - cell["P"][p])
        vstd = 1.0 if lim.water else 1.0 / xi_phase(fluid, Psurf, Tsurf, zi)      # CalVmStd
        q *= vstd
    else:
        nc = Nc + 1
        qi = np.zeros(nc)
        for p in range(npf):
            Pperf = lim.limit_bhp + dG[p]
            for j in range(3):
                if ex[p, j]:
                    temp = tj[p, j] * cell["XI"][p, j] * (cell["PJ"][p, j] - Pperf)
                    qi += cell["XIJ"][p, j] * temp
        if qi.sum() > 0:                                                         # CalVStd
            exf, vj, _ = flash(Psurf, Tsurf, qi)
            vj = np.where(exf, vj, 0.0)
            w = np.asarray(lim.weights, float)
            q = vj[0] * w[0] + vj[1] * w[1] + qi[nc - 1] * w[2]
        else:
            q = 0.0                                                              # no inflow at minBHP
    if lim.init_mode == BHP:
        if well.type == INJ and q > lim.max_rate:
            return RATE, well.bhp, q
        return BHP, lim.tar_bhp, q
    if q > lim.max_rate:
        return lim.init_mode, well.bhp, q
    return BHP, lim.tar_bhp, q


def check_cross_flow(cell_P, well, perf_open):
    """PeacemanWell::CheckCrossFlow (src/WellPeaceman.cpp:573-652): the first
    open perforation whose pressure difference points the wrong way (bulk
    below the perforation pressure for a producer, above it for an
    injector) is closed, closed perforations that point the right way again
    reopen, and the deepest perforation opens if none is left.  Returns
    (open flags, crossflow); a closed perforation has multiplier 0, so the
    WI passed to ocpx_set_wells is WI * open."""
    npf = len(well.cells)
    dG = np.asarray(well.dG if well.dG is not None else np.zeros(npf), float)
    perfP = well.bhp + dG
    op = list(bool(v) for v in perf_open)
    cross = False
    for p in range(npf):
        wrong = cell_P[p] < perfP[p] if well.type == PROD else cell_P[p] > perfP[p]
        right = cell_P[p] > perfP[p] if well.type == PROD else cell_P[p] < perfP[p]
        if op[p] and wrong:
            op[p] = False
            cross = True
            break
        if not op[p] and right:
            op[p] = True
    if not any(op):
        op[-1] = True
    return op, cross


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


def corner_and_centre_wells(wl, WI=5.0):
    """The five vertical wells of the heterogeneous-grid workload, through
    every layer: BHP producers at the four corners 10 bar below the lowest
    pressure they perforate, and a rate injector of the first component at
    the centre (molar density 8, 100 units per day), with hydrostatic-like
    perforation heads."""
    g, nc = wl.grid, wl.Nc + 1
    nx, ny, nz = g.nx, g.ny, g.nz
    col = lambda i, j: [i + j * nx + k * nx * ny for k in range(nz)]
    P = wl.state["P"]
    out = []
    for name, (i, j) in (("P1", (0, 0)), ("P2", (nx - 1, 0)), ("P3", (0, ny - 1)), ("P4", (nx - 1, ny - 1))):
        c = col(i, j)
        out.append(PeacemanWell(name, PROD, BHP, bhp=float(P[c].min()) - 10, target=float(P[c].min()) - 10,
                                cells=c, WI=[WI] * nz, dG=list(0.07 * np.arange(nz))))
    c = col(nx // 2, ny // 2)
    zi = np.zeros(nc)
    zi[0] = 1.0
    out.append(PeacemanWell("I", INJ, RATE, bhp=float(P[c].max()) + 10, target=100.0, cells=c, WI=[WI] * nz,
                            dG=list(0.05 * np.arange(nz)), xi=[8.0] * nz, zi=list(zi), factor=[1.0] * nc))
    return out
