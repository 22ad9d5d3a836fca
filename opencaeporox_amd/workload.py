"""Synthetic inputs for the flash + FIM-assembly workloads (host logic).

The reference ships no benchmark inputs for this path, so every workload is
generated here from a fixed numpy PCG64 seed, in the reference's METRIC
units (bar, K / degC, m, mD, day; include/OCPUnits.hpp:78-100):

* fluids: Nc Peng-Robinson components with Tc ~ U(190, 800) K,
  Pc ~ U(1.0, 4.6) MPa, acentric factor ~ U(0.01, 0.9), binary interaction
  ~ U(0, 0.15) (symmetric, zero diagonal); MW grows monotonically with Tc;
  Zc = 0.2905 - 0.085*omega gives Vc = R*Zc*Tc/Pc; LBC coefficients are the
  reference defaults (src/ParamReservoir.cpp:710-718);
* water (PVTW), SWOF/SGOF and ROCK from data/spe5/spe5.data:87-127 converted
  from FIELD to METRIC, with a small seeded capillary-pressure curve added
  (the SPE5 tables have Pc = 0, which would leave the dPc/dS terms of the
  Jacobian unexercised);
* flash controls from data/spe5/EGOIL.in:70-97;
* Cartesian TPFA grids in the reference's connection order
  (src/PreParamGridWell.cpp:1140-1214) with BulkConnTransMethod01
  transmissibilities (src/BulkConnFunc.cpp:15-67).
"""
from dataclasses import dataclass, field

import numpy as np

from . import abi

R_GAS = 0.083143
DARCY = 0.008527
CONV4 = 273.15
PSI2BAR = 0.0689475729
D1, D2 = 2.41421356237, -0.41421356237
LBC_DEFAULT = (0.1023, 0.023364, 0.058533, -0.040758, 0.0093324)

# data/spe5/spe5.data:87-113 (SWOF Sw krw krow, SGOF Sg krg krog)
SPE5_SWOF = [(0.2, 0, 1), (0.2899, 0.0022, 0.6769), (0.3778, 0.018, 0.4153),
             (0.4667, 0.0607, 0.2178), (0.5556, 0.1438, 0.0835), (0.6444, 0.2809, 0.0123),
             (0.7, 0.4089, 0), (0.7333, 0.4855, 0), (0.8222, 0.7709, 0), (0.9111, 1, 0),
             (1, 1, 0)]
SPE5_SGOF = [(0.0, 0.0, 1.0), (0.05, 0.0, 0.88), (0.0889, 0.001, 0.7023),
             (0.1778, 0.01, 0.4705), (0.2667, 0.03, 0.2963), (0.3556, 0.05, 0.1715),
             (0.4444, 0.1, 0.0878), (0.5333, 0.2, 0.037), (0.6222, 0.35, 0.011),
             (0.65, 0.39, 0.0), (0.7111, 0.56, 0.0), (0.8, 1.0, 0.0)]


def make_fluid(Nc, seed, skip_psa=True):
    """Seeded Nc-component PR fluid with SPE5 water and flash controls."""
    rng = np.random.Generator(np.random.PCG64(seed))
    # heavier components: higher Tc, lower Pc, larger acentric factor
    Tc = np.sort(rng.uniform(190.0, 800.0, Nc))
    Pc = np.sort(rng.uniform(1.0, 4.6, Nc))[::-1] * 10.0            # MPa -> bar
    acf = np.sort(rng.uniform(0.01, 0.9, Nc))
    MW = 16.04 * (Tc / 190.6) ** 1.8
    Zc = 0.2905 - 0.085 * acf
    Vc = R_GAS * Zc * Tc / Pc
    kij = np.zeros((Nc, Nc))
    iu = np.triu_indices(Nc, 1)
    kij[iu] = rng.uniform(0.0, 0.15, len(iu[0]))
    kij = kij + kij.T
    return fluid_from_arrays(Tc, Pc, acf, MW, Vc, kij, skip_psa=skip_psa)


def fluid_from_arrays(Tc, Pc, acf, MW, Vc, kij, skip_psa=True, Vshift=None):
    Nc = len(Tc)
    f = abi.Fluid()
    f.nc = Nc
    for i in range(Nc):
        f.Tc[i], f.Pc[i], f.acf[i], f.MW[i] = Tc[i], Pc[i], acf[i], MW[i]
        f.Vc[i] = f.Vcvis[i] = Vc[i]
        f.OmegaA[i], f.OmegaB[i] = 0.457235529, 0.077796074
        f.Vshift[i] = 0.0 if Vshift is None else Vshift[i]
        for k in range(Nc):
            f.BIC[i * abi.MAX_NC + k] = kij[i, k]
    for k in range(5):
        f.LBCcoef[k] = LBC_DEFAULT[k]
    # PVTW 14.7 psi, 1.0, 3.3e-6 /psi, 0.7 cP, 0  (data/spe5/spe5.data:117-119)
    f.PVTW[0], f.PVTW[1], f.PVTW[2], f.PVTW[3], f.PVTW[4] = (
        14.7 * PSI2BAR, 1.0, 3.3e-6 / PSI2BAR, 0.7, 0.0)
    f.rhoW_std = 999.014
    # data/spe5/EGOIL.in:70-97
    f.RR_maxIt, f.RR_tol = 30, 1e-12
    f.SSMsta_maxIt, f.SSMsta_tol, f.SSMsta_eYt = 100, 1e-12, 1e-8
    f.NRsta_maxIt, f.NRsta_tol = 55, 1e-12
    f.SSMsp_maxIt, f.SSMsp_tol = 100, 1e-6
    f.NRsp_maxIt, f.NRsp_tol = 55, 1e-12
    f.skipPSA = 1 if skip_psa else 0
    return f


def fluid_arrays(f):
    n = f.nc
    g = lambda a: np.array(a[:n])
    kij = np.array(f.BIC[:]).reshape(abi.MAX_NC, abi.MAX_NC)[:n, :n].copy()
    return dict(Tc=g(f.Tc), Pc=g(f.Pc), acf=g(f.acf), MW=g(f.MW), Vc=g(f.Vc),
                OmegaA=g(f.OmegaA), OmegaB=g(f.OmegaB), kij=kij)


def make_sat(pc_scale=1.0):
    """SPE5 SWOF/SGOF with a seeded-shape capillary pressure (bar)."""
    s = abi.Sat()
    sw = np.array([r[0] for r in SPE5_SWOF])
    pcow = pc_scale * 0.6 * ((1.0 - sw) / 0.8) ** 2            # decreasing, 0 at Sw=1
    for r, row in enumerate(SPE5_SWOF):
        s.swof[4 * r:4 * r + 4] = [row[0], row[1], row[2], pcow[r]]
    s.nswof = len(SPE5_SWOF)
    sg = np.array([r[0] for r in SPE5_SGOF])
    pcgo = pc_scale * 0.3 * (sg / 0.8) ** 1.5                    # increasing, 0 at Sg=0
    for r, row in enumerate(SPE5_SGOF):
        s.sgof[4 * r:4 * r + 4] = [row[0], row[1], row[2], pcgo[r]]
    s.nsgof = len(SPE5_SGOF)
    return s


def make_rock():
    # ROCK LINEAR01 3990.3 psi 5e-6 /psi  (data/spe5/spe5.data:125-127)
    r = abi.Rock()
    r.Pref, r.cp1, r.cp2 = 3990.3 * PSI2BAR, 5e-6 / PSI2BAR, 0.0
    return r


# ---------------------------------------------------------------- PR helpers
def pr_aibi(fa, P, T):
    """Ai, Bi per cell and component (src/OCPEoS.cpp:68-84), P,T shape (n,)."""
    w = fa["acf"]
    m = np.where(w <= 0.49, 0.37464 + 1.54226 * w - 0.26992 * w ** 2,
                 0.379642 + 1.48503 * w - 0.164423 * w ** 2 + 0.016667 * w ** 3)
    Pr = P[:, None] / fa["Pc"][None, :]
    Tr = T[:, None] / fa["Tc"][None, :]
    Ai = fa["OmegaA"] * Pr / Tr ** 2 * (1 + m * (1 - np.sqrt(Tr))) ** 2
    Bi = fa["OmegaB"] * Pr / Tr
    return Ai, Bi


def pr_single_phase_vm(fa, P, T, z):
    """Molar volume of composition z with the reference's root choice."""
    Ai, Bi = pr_aibi(fa, P, T)
    sq = np.sqrt(Ai)
    A = np.einsum("ni,nk,ik,ni,nk->n", z, z, 1 - fa["kij"], sq, sq)
    B = np.sum(z * Bi, axis=1)
    a = (D1 + D2 - 1) * B - 1
    b = A + D1 * D2 * B * B - (D1 + D2) * B * (B + 1)
    c = -(A * B + D1 * D2 * B * B * (B + 1))
    comp = np.zeros((len(P), 3, 3))
    comp[:, 0, :] = np.stack([-a, -b, -c], axis=1)
    comp[:, 1, 0] = 1
    comp[:, 2, 1] = 1
    r = np.linalg.eigvals(comp)
    real = np.abs(r.imag) < 1e-9 * (1 + np.abs(r.real))
    rr = np.where(real & (r.real > B[:, None]), r.real, np.nan)
    z1 = np.nanmin(rr, axis=1)
    z2 = np.nanmax(rr, axis=1)
    with np.errstate(invalid="ignore", divide="ignore"):
        dG = (z2 - z1) + np.log((z1 - B) / (z2 - B)) - A / (B * (D2 - D1)) * np.log(
            (z1 + D1 * B) * (z2 + D2 * B) / ((z1 + D2 * B) * (z2 + D1 * B)))
    Z = np.where(z1 == z2, z1, np.where(dG > 0, z1, z2))
    return Z * R_GAS * T / P


# --------------------------------------------------------------------- grids
@dataclass
class Grid:
    nx: int
    ny: int
    nz: int
    v: np.ndarray
    depth: np.ndarray
    poro: np.ndarray
    conn_b: np.ndarray
    conn_e: np.ndarray
    trans: np.ndarray

    @property
    def nb(self):
        return self.v.size

    @property
    def nconn(self):
        return self.conn_b.size

    def to_abi(self):
        self._keep = [np.ascontiguousarray(a) for a in
                      (self.v, self.depth, self.poro, self.conn_b.astype(np.int32),
                       self.conn_e.astype(np.int32), self.trans)]
        g = abi.Grid()
        g.nb, g.nconn = self.nb, self.nconn
        g.v, g.depth, g.poroInit, g.conn_b, g.conn_e, g.trans = [a.ctypes.data for a in self._keep]
        return g


def cartesian_grid(nx, ny, nz, d, kx, ky, kz, poro, ntg=None, top=2000.0):
    """Uniform cubes of edge d (m); permeabilities per cell (mD)."""
    nb = nx * ny * nz
    ntg = np.ones(nb) if ntg is None else ntg
    v = np.full(nb, d * d * d)
    k = np.arange(nb) // (nx * ny)
    depth = top + (k + 0.5) * d
    area = 2 * d * d / d
    idx = np.arange(nb)
    i, j = idx % nx, (idx // nx) % ny
    cbs, ces, trs, dirs = [], [], [], []
    for dcode, mask, off, kk, use_ntg in ((0, i < nx - 1, 1, kx, True),
                                          (1, j < ny - 1, nx, ky, True),
                                          (2, k < nz - 1, nx * ny, kz, False)):
        b = idx[mask]
        e = b + off
        t1 = (ntg[b] if use_ntg else 1.0) * kk[b] * area
        t2 = (ntg[e] if use_ntg else 1.0) * kk[e] * area
        cbs.append(b); ces.append(e); trs.append(1.0 / (1 / t1 + 1 / t2) * DARCY)
        dirs.append(np.full(b.size, dcode))
    cb, ce, tr, dr = (np.concatenate(a) for a in (cbs, ces, trs, dirs))
    order = np.lexsort((dr, cb))
    return Grid(nx, ny, nz, v, depth, np.asarray(poro, float), cb[order].astype(np.int32),
                ce[order].astype(np.int32), tr[order])


# --------------------------------------------------------------------- state
@dataclass
class Workload:
    name: str
    fluid: object
    sat: object
    rock: object
    grid: Grid
    state: dict = field(default_factory=dict)
    dt: float = 1.0

    @property
    def Nc(self):
        return self.fluid.nc


def initial_state(fluid, rock, grid, P, z, rng, Sw=0.2, perturb=True, T=344.0):
    """BulkVarSet-layout arrays for the first Newton flash of a time step:
    water-only previous phase state, Ni filling the pore volume at P
    (InitFlashDer semantics with the single-phase molar volume,
    src/OCPMixtureMethodComp.cpp:732-743, 821-837), lNi = Ni and then a
    seeded +-1% Ni and +-0.1 MPa P perturbation so residuals are non-zero.
    T is the temperature in K (scalar or per cell)."""
    fa = fluid_arrays(fluid)
    Nc = fluid.nc
    nc, np_ = Nc + 1, 3
    nb = grid.nb
    T = np.broadcast_to(np.asarray(T, float), (nb,)).copy()
    dP = P - rock.Pref
    Vp = grid.v * grid.poro * (1 + (rock.cp1 + rock.cp2 / 2 * dP) * dP)
    vm = pr_single_phase_vm(fa, P, T, z)
    Nhc = Vp * (1 - Sw) / vm
    Bw = fluid.PVTW[1] * (1 - fluid.PVTW[2] * (P - fluid.PVTW[0]))
    Nw = Vp * Sw / Bw
    Ni = np.concatenate([z * Nhc[:, None], Nw[:, None]], axis=1)
    lNi = Ni.copy()
    if perturb:
        Ni = Ni * (1 + rng.uniform(-0.01, 0.01, Ni.shape))
        P = P + rng.uniform(-1.0, 1.0, nb)
    st = {
        "P": P.astype(float), "T": T - CONV4, "NI": Ni.ravel(), "LNI": lNi.ravel(),
        "PHASEEXIST": np.tile(np.array([0, 0, 1], np.uint32), nb),
        "S": np.tile(np.array([0.0, 0.0, 1.0]), nb),
        "XIJ": np.zeros(nb * np_ * nc),
    }
    x = st["XIJ"].reshape(nb, np_, nc)
    x[:, 2, Nc] = 1.0
    return st


def cube_grid(seed=1):
    """10x10x10 cubes of 20 ft, ln k ~ N(ln 100, 1), phi ~ U(0.1, 0.3), Nc = 6."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nx = ny = nz = 10
    nb = nx * ny * nz
    k = np.exp(rng.normal(np.log(100.0), 1.0, nb))
    poro = rng.uniform(0.1, 0.3, nb)
    grid = cartesian_grid(nx, ny, nz, 6.096, k, k, k, poro)
    fl = make_fluid(6, seed + 100)
    rock = make_rock()
    P = rng.uniform(100.0, 280.0, nb)
    z = rng.dirichlet(np.ones(6), nb)
    st = initial_state(fl, rock, grid, P, z, rng)
    return Workload("10x10x10 cubes, Nc=6", fl, make_sat(), rock, grid, st)


def heterogeneous_grid(nx=100, ny=100, nz=50, Nc=6, seed=3, lbc_exact=0, composition="per-cell"):
    """Hydrostatic P + N(0, 1 MPa), ln kx ~ N(ln 100, 2), kz = 0.1 kx.
    composition "per-cell": an independent Dirichlet(1) draw per cell (every
    phase state side by side, for parity); "uniform": one Dirichlet(1) draw
    for the whole grid (a reservoir's fluid; the linear-solve workload, whose
    Newton system is far better posed: with per-cell draws neighbouring cells
    hold unrelated fluids and no preconditioner tried converges)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nb = nx * ny * nz
    kx = np.exp(rng.normal(np.log(100.0), 2.0, nb))
    poro = rng.uniform(0.1, 0.3, nb)
    grid = cartesian_grid(nx, ny, nz, 6.096, kx, kx, 0.1 * kx, poro)
    fl = make_fluid(Nc, seed + 100)
    fl.lbc_exact = lbc_exact
    rock = make_rock()
    P = 200.0 + 0.0980665 * (grid.depth - grid.depth.min()) + rng.normal(0, 10.0, nb)
    if composition == "per-cell":
        z = rng.dirichlet(np.ones(Nc), nb)
    elif composition == "uniform":
        z = np.tile(rng.dirichlet(np.ones(Nc)), (nb, 1))
    else:
        raise ValueError("composition is 'per-cell' or 'uniform'")
    st = initial_state(fl, rock, grid, P, z, rng)
    tag = "" if composition == "per-cell" else ", one composition"
    return Workload(f"{nx}x{ny}x{nz} heterogeneous grid, Nc={Nc}{tag}", fl, make_sat(), rock, grid, st)


def channel_facies(nx, ny, nz, rng, nchan=6, width=0.035):
    """Seeded binary facies: sinuous channels along x, one set per layer
    drifting slowly with depth.  1 = channel."""
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    fac = np.zeros((nz, ny, nx), np.int8)
    c0 = rng.uniform(0, 1, nchan)
    amp = rng.uniform(0.03, 0.12, nchan)
    wl = rng.uniform(0.3, 1.0, nchan)
    ph = rng.uniform(0, 2 * np.pi, nchan)
    for k in range(nz):
        drift = 0.15 * np.sin(2 * np.pi * k / max(nz, 1) + ph)
        for c in range(nchan):
            yc = (c0[c] + drift[c] + amp[c] * np.sin(2 * np.pi * x / wl[c] + ph[c])) % 1.0
            d = np.abs(y[:, None] - yc[None, :])
            fac[k] |= (np.minimum(d, 1 - d) < width).astype(np.int8)
    return fac.ravel()


def channelized_grid(nx=60, ny=220, nz=85, Nc=4, seed=4, lbc_exact=0):
    """Channelized facies (2000 / 1 mD means, log-normal sigma = 1 inside
    each facies), Nc = 4, hydrostatic P + N(0, 1 MPa)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nb = nx * ny * nz
    fac = channel_facies(nx, ny, nz, rng)
    k = np.exp(np.where(fac == 1, np.log(2000.0), np.log(1.0)) + rng.normal(0, 1.0, nb))
    poro = np.where(fac == 1, rng.uniform(0.2, 0.3, nb), rng.uniform(0.05, 0.15, nb))
    grid = cartesian_grid(nx, ny, nz, 6.096, k, k, 0.1 * k, poro)
    fl = make_fluid(Nc, seed + 100)
    fl.lbc_exact = lbc_exact
    rock = make_rock()
    P = 200.0 + 0.0980665 * (grid.depth - grid.depth.min()) + rng.normal(0, 10.0, nb)
    z = rng.dirichlet(np.ones(Nc), nb)
    st = initial_state(fl, rock, grid, P, z, rng)
    return Workload(f"{nx}x{ny}x{nz} channelized grid, Nc={Nc}", fl, make_sat(), rock, grid, st)


def model_box(ncell, nz=10):
    """Cartesian box of about ncell cells: nz layers of a square plane."""
    side = max(2, int(round(np.sqrt(ncell / nz))))
    return side, side, nz


def uneven_batch(nmodels=256, seed=5, cmin=5000, cmax=200000):
    """Uneven multi-model batch: cell counts log-uniform in [cmin, cmax],
    Nc uniform in {3..12}, each model with its own seeded PR fluid,
    log-normal permeability, porosity and pressures."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for m in range(nmodels):
        ncell = int(np.exp(rng.uniform(np.log(cmin), np.log(cmax))))
        Nc = int(rng.integers(3, 13))
        nx, ny, nz = model_box(ncell)
        nb = nx * ny * nz
        r = np.random.Generator(np.random.PCG64(seed * 100003 + m))
        k = np.exp(r.normal(np.log(100.0), 1.0, nb))
        poro = r.uniform(0.1, 0.3, nb)
        grid = cartesian_grid(nx, ny, nz, 6.096, k, k, 0.1 * k, poro)
        fl = make_fluid(Nc, seed * 1000 + m)
        rock = make_rock()
        P = r.uniform(100.0, 280.0, nb)
        z = r.dirichlet(np.ones(Nc), nb)
        st = initial_state(fl, rock, grid, P, z, r)
        out.append(Workload(f"model {m} {nx}x{ny}x{nz} Nc={Nc}", fl, make_sat(), rock, grid, st))
    return out


def flash_cells(n, Nc, seed, P=None, alpha=0.5):
    """Independent cells (no connections) for the flash-throughput workload."""
    rng = np.random.Generator(np.random.PCG64(seed))
    fl = make_fluid(Nc, seed + 100)
    rock = make_rock()
    grid = Grid(n, 1, 1, np.full(n, 226.5), np.full(n, 2000.0), np.full(n, 0.2),
                np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0))
    if P is None:
        P = rng.uniform(100.0, 280.0, n)
    z = rng.dirichlet(np.full(Nc, alpha), n)
    st = initial_state(fl, rock, grid, P, z, rng, perturb=False)
    return Workload(f"flash {n} cells Nc={Nc}", fl, make_sat(), rock, grid, st), z


def saturation_pressure(nphase, z, T, lo=5.0, hi=700.0, coarse=24, iters=30):
    """Upper phase-boundary pressure per composition by bisection on the
    phase count returned by nphase(P, z, T) (vectorised).  Cells that are
    single-phase on the whole coarse grid get NaN."""
    n = z.shape[0]
    grid = np.linspace(lo, hi, coarse)
    top2 = np.full(n, np.nan)
    for p in grid:
        two = nphase(np.full(n, p), z, T) >= 2
        top2 = np.where(two, p, top2)
    ok = ~np.isnan(top2)
    a = np.where(ok, top2, lo)
    b = np.minimum(a + (grid[1] - grid[0]), hi)
    for _ in range(iters):
        m = 0.5 * (a + b)
        two = nphase(m, z, T) >= 2
        a = np.where(two, m, a)
        b = np.where(two, b, m)
    return np.where(ok, 0.5 * (a + b), np.nan)


def near_saturation_pressures(n, rng, psat, frac_near=0.3, band=0.05):
    """30% of the cells within +-5% of their saturation pressure, the rest
    (and cells with no saturation pressure) U(10, 28) MPa."""
    P = rng.uniform(100.0, 280.0, n)
    near = (rng.uniform(0, 1, n) < frac_near) & np.isfinite(psat)
    P[near] = psat[near] * (1 + rng.uniform(-band, band, near.sum()))
    return P, near
