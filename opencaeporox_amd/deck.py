"""Reader for the reference's input decks (host logic).

Reads the keywords the flash + assembly path consumes and turns them into
the structures the C ABI takes (ocpx_fluid, ocpx_sat, ocpx_rock and a
Cartesian grid), in the METRIC units the kernels compute in:

* fluid: COMPS, CNAMES, TCRIT, PCRIT, VCRIT, ZCRIT, MW, ACF, OMEGAA, OMEGAB,
  SSHIFT, PARACHOR, VCRITVIS, ZCRITVIS, BIC, LBCCOEF and the flash controls
  RR, SSMSTA, NRSTA, SSMSP, NRSP (src/ParamRead.cpp:73-365,
  src/ParamReservoir.cpp:907-1119);
* PVT / SAT / rock: RTEMP, PVTW, SWOF, SGOF, ZMFVD, PBVD (and the other
  tables of InputTABLE, kept as parsed), ROCK, GRAVITY, DENSITY, EQUIL,
  TABDIMS (src/ParamReservoir.cpp:178-595);
* grid: DIMENS, EQUALS, COPY, MULTIPLY and the per-cell arrays DX, DY, DZ,
  TOPS, NTG, PORO, PERMX, PERMY, PERMZ (src/PreParamGridWell.cpp:240-500),
  with depth and volume as CalDepthVOrthogonalGrid (:1088-1120) and
  transmissibilities as BulkConnTransMethod01 (src/BulkConnFunc.cpp:15-67);
* INCLUDE, relative to the deck's folder.

Lines are tokenised as the reference's ReadLine / DealDefault
(src/UtilInput.cpp:14-76).  Keywords this reader does not know are skipped
one line at a time, as the reference's keyword switch skips them.  WELSPECS
and COMPDAT records are kept as token lists.

Units.  The reference computes in the deck's own unit system
(include/OCPUnits.hpp); the kernels here compute in METRIC, so a FIELD deck
is converted on the way in: degR -> K, degF -> degC, psia -> bar, ft -> m,
ft3/lb-mol -> m3/kmol, 1/psi -> 1/bar.  The Peng-Robinson quantities
(A, B, Z, ln phi) depend on reduced P and T only and come out the same in
either system; molar volumes differ by the ratio of the two gas constants
(10.73159 FIELD vs 0.083143 METRIC, 2.1e-5 apart after conversion).  The
reference evaluates the LBC dilute-gas term 5.4402 Tc^(1/6) / Pc^(2/3) in
the deck's units (src/OCPFuncPVT.cpp:457); after conversion it is evaluated
in K and bar, as a METRIC deck would give it.
"""
import os
import re
from dataclasses import dataclass, field

import numpy as np

from . import abi
from . import workload as W

# include/OCPUnits.hpp:45-100
FIELD_GAS_CONSTANT = 10.73159
METRIC_GAS_CONSTANT = 0.083143
METRIC_RHOW_STD = 999.014
METRIC_GRAVITY_FACTOR = 0.0000980665
PSI = 0.0689475729            # bar per psi
FT = 0.3048                   # m per ft
FT3_PER_LBMOL = 0.028316846592 / 0.45359237   # m3/kmol per ft3/lb-mol

# src/ParamReservoir.cpp:710-718 and src/OCPEoS.cpp:26-29
LBC_DEFAULT = (0.1023, 0.023364, 0.058533, -0.040758, 0.0093324)
OMEGA_A, OMEGA_B = 0.457235529, 0.077796074

# component keywords held per PVT region (src/ParamReservoir.cpp:721-856)
COMPONENT_KEYWORDS = ("TCRIT", "PCRIT", "VCRIT", "ZCRIT", "MW", "ACF", "OMEGAA", "OMEGAB", "SSHIFT",
                      "PARACHOR", "VCRITVIS", "ZCRITVIS")
# tables read by InputTABLE, with their column counts (src/ParamReservoir.cpp:142-176)
TABLE_COLUMNS = {"SWFN": 3, "SWOF": 4, "SWGF": 4, "SGFN": 3, "SGOF": 4, "SOF3": 3, "PBVD": 2,
                 "PVCO": 6, "PVDO": 3, "PVCDO": 5, "PVDG": 3, "PVTW": 5, "TEMPVD": 2}
GRID_ARRAYS = ("DX", "DY", "DZ", "TOPS", "NTG", "PORO", "PERMX", "PERMY", "PERMZ")
FLASH_KEYWORDS = {"RR": 2, "SSMSTA": 3, "NRSTA": 2, "SSMSP": 2, "NRSP": 2}

_BLANKS = re.compile(r"[ \t\n\v\f\r]+")


class DeckError(ValueError):
    pass


def tokens(line):
    """One line as ReadLine gives it, or None for a line it skips: leading
    blanks dropped; blank, '#' and '--' lines skipped; everything after the
    first '/' dropped and '/' kept as the last token; quotes and commas are
    blanks."""
    s = line.lstrip(" \t\r").rstrip("\n")
    if not s or s[0] == "#" or s.startswith("--"):
        return None
    p = s.find("/")
    if p >= 0:
        s = s[:p] + "/"
    s = s.replace("'", " ").replace(",", " ")
    t = [x for x in _BLANKS.split(s) if x]
    return t or None


def deal_default(vbuf):
    """n* -> n x 'DEFAULT', n*v -> n x v; a token whose count is not a
    positive integer is kept as it is (DealDefault)."""
    out = []
    for s in vbuf:
        p = s.find("*")
        if p < 0:
            out.append(s)
            continue
        m = re.match(r"\s*[+-]?\d+", s[:p])
        num = int(m.group(0)) if m else 0
        if num <= 0:
            out.append(s)
        else:
            out.extend([s[p + 1:] if p != len(s) - 1 else "DEFAULT"] * num)
    return out


class _Lines:
    """Token lines of a deck file, INCLUDE files read in place."""

    def __init__(self, path):
        self.stack = [self._open(path)]
        self.workdir = os.path.dirname(os.path.abspath(path))

    @staticmethod
    def _open(path):
        with open(path, encoding="latin-1") as f:
            text = f.read()
        return iter(text.splitlines())

    def include(self, name):
        self.stack.append(self._open(os.path.join(self.workdir, name)))

    def next(self):
        """The next token line, or None at the end of the deck."""
        while self.stack:
            for line in self.stack[-1]:
                t = tokens(line)
                if t is not None:
                    return t
            self.stack.pop()
        return None

    def need(self, keyword):
        t = self.next()
        if t is None:
            raise DeckError(f"{keyword}: the deck ends inside the keyword")
        return t


def _float(s, keyword):
    try:
        return float(s.replace("D", "E").replace("d", "e"))
    except ValueError:
        raise DeckError(f"{keyword}: '{s}' is not a number") from None


@dataclass
class Deck:
    """The parsed deck, in the deck's own units (see to_* for METRIC)."""
    path: str = ""
    units: str = "FIELD"
    num_com: int = 0
    ntsfun: int = 1
    ntpvt: int = 1
    ntrooc: int = 1
    cnames: list = field(default_factory=list)
    comp: dict = field(default_factory=dict)           # keyword -> [per region [numCom]]
    bic: list = field(default_factory=list)            # per region, flat
    lbccoef: list = field(default_factory=lambda: list(LBC_DEFAULT))
    flash: dict = field(default_factory=dict)          # keyword -> [str]
    rtemp: float = 60.0
    tables: dict = field(default_factory=dict)         # name -> [table: [columns]]
    rock: list = field(default_factory=list)           # [{type, Pref, cp1, cp2}]
    gravity: list = field(default_factory=lambda: [45.5, 1.0, 0.7773])
    gravity_active: bool = False
    density: list = field(default_factory=lambda: [37.457, 62.366416, 0.062428])
    density_active: bool = False
    equil: list = field(default_factory=list)
    dimens: tuple = None
    arrays: dict = field(default_factory=dict)         # grid keyword -> np.ndarray
    welspecs: list = field(default_factory=list)
    compdat: list = field(default_factory=list)
    schedule: list = field(default_factory=list)       # (step, keyword, record) of WCONPROD / WCONINJE
    wellstre: dict = field(default_factory=dict)       # injected stream name -> composition
    step: int = 0                                      # TSTEP keywords read so far
    skipped: list = field(default_factory=list)        # keywords not read

    # ------------------------------------------------------------ parsing
    @classmethod
    def read(cls, path):
        d = cls(path=os.path.abspath(path))
        L = _Lines(path)
        while True:
            t = L.next()
            if t is None:
                break
            kw = t[0]
            h = _HANDLERS.get(kw)
            if h is not None:
                h(d, L, t)
            elif kw in COMPONENT_KEYWORDS:
                d._components(L, kw)
            elif kw in TABLE_COLUMNS or kw == "ZMFVD":
                d._table(L, kw)
            elif kw in GRID_ARRAYS:
                d._array(L, kw)
            elif kw != "/":
                d.skipped.append(kw)
        return d

    def _need_comps(self, kw):
        if self.num_com <= 0:
            raise DeckError(f"{kw}: the number of components (COMPS or NCOMPS) has not been read")

    def _components(self, L, kw):
        # InputCOMPONENTS: NTPVT records of numCom values, each ended by '/'
        self._need_comps(kw)
        data, tmp = [], []
        while True:
            v = L.need(kw)
            if v[0] == "/":
                data.append((tmp + [0.0] * self.num_com)[:self.num_com])
                tmp = []
                if len(data) >= self.ntpvt:
                    break
                continue
            tmp += [_float(x, kw) for x in v if x != "/"]
            if v[-1] == "/":
                data.append((tmp + [0.0] * self.num_com)[:self.num_com])
                tmp = []
                if len(data) >= self.ntpvt:
                    break
        self.comp[kw] = data

    def _table(self, L, kw):
        # InputTABLE: rows of colNum values; '/' at a row's end closes a
        # table, a '/' line ends the keyword
        if kw == "ZMFVD":
            self._need_comps(kw)
        col = TABLE_COLUMNS.get(kw, self.num_com + 1)
        tabs, cur = [], [[] for _ in range(col)]
        while True:
            v = L.next()
            if v is None or v[0] == "/":
                break
            if len(v) - (v[-1] == "/") < col:
                raise DeckError(f"{kw}: row {' '.join(v)} has fewer than {col} values")
            for i in range(col):
                cur[i].append(_float(v[i], kw))
            if v[-1] == "/":
                tabs.append(cur)
                cur = [[] for _ in range(col)]
        if cur[0]:
            tabs.append(cur)
        self.tables.setdefault(kw, []).extend(tabs)

    def _array(self, L, kw):
        # InputGridParam: values with m*v repetition until a '/' line
        vals = []
        while True:
            v = L.next()
            if v is None or v[0] == "/":
                break
            for s in v:
                if s == "/":
                    break
                p = s.find("*")
                if p < 0:
                    vals.append(_float(s, kw))
                else:
                    vals.extend([_float(s[p + 1:], kw)] * int(s[:p]))
            if v[-1] == "/":
                break
        a = np.array(vals)
        if kw == "TOPS" and self.dimens and a.size == self.nxny:
            full = np.zeros(self.nb)
            full[:a.size] = a
            a = full
        self.arrays[kw] = a

    # ----------------------------------------------------------- grid
    @property
    def nxny(self):
        return self.dimens[0] * self.dimens[1]

    @property
    def nb(self):
        return self.dimens[0] * self.dimens[1] * self.dimens[2]

    def _grid_array(self, kw):
        if self.dimens is None:
            raise DeckError(f"{kw}: DIMENS has not been read")
        a = self.arrays.get(kw)
        if a is None or a.size != self.nb:
            a = np.zeros(self.nb)
            self.arrays[kw] = a
        return a

    def _box(self, vbuf, kw):
        nx, ny, nz = self.dimens
        index = [0, nx - 1, 0, ny - 1, 0, nz - 1]
        vb = deal_default(vbuf)
        for n in range(2, 8):
            if n < len(vb) and vb[n] not in ("DEFAULT", "/"):
                index[n - 2] = int(vb[n]) - 1
        i0, i1, j0, j1, k0, k1 = index
        if i0 < 0 or j0 < 0 or k0 < 0 or i1 > nx - 1 or j1 > ny - 1 or k1 > nz - 1:
            raise DeckError(f"{kw}: box {[x + 1 for x in index]} outside the grid {self.dimens}")
        return index

    def _cells(self, index):
        i0, i1, j0, j1, k0, k1 = index
        nx, ny, _ = self.dimens
        k, j, i = np.meshgrid(np.arange(k0, k1 + 1), np.arange(j0, j1 + 1), np.arange(i0, i1 + 1), indexing="ij")
        return (k * ny * nx + j * nx + i).ravel()

    def _equals(self, L, kw, op):
        while True:
            v = L.next()
            if v is None or v[0] == "/":
                break
            name, val = v[0], _float(v[1], kw)
            if name not in GRID_ARRAYS:
                raise DeckError(f"{kw}: unknown array {name}")
            index = self._box(v, kw)
            if name == "TOPS":
                index[4] = index[5] = 0
            a = self._grid_array(name)
            c = self._cells(index)
            if op == "set":
                a[c] = val
            else:
                a[c] *= val

    def _copy(self, L, kw):
        while True:
            v = L.next()
            if v is None or v[0] == "/":
                break
            src, dst = v[0], v[1]
            if src not in self.arrays or dst not in GRID_ARRAYS:
                raise DeckError(f"COPY: {src} -> {dst}: unknown or unset array")
            c = self._cells(self._box(v, kw))
            a = self._grid_array(dst)
            a[c] = self.arrays[src][c]

    # -------------------------------------------------------- conversion
    @property
    def field_units(self):
        return self.units == "FIELD"

    def _region(self, kw, region):
        d = self.comp.get(kw)
        return None if d is None else np.array(d[region], float)

    def to_fluid(self, region=0, skip_psa=True, lbc_exact=0):
        """ocpx_fluid in METRIC, with the reference's defaults and derived
        values: Omega_A/B (src/OCPEoS.cpp:26-29); Vc from ZCRIT when VCRIT
        is absent (src/OCPMixtureMethodComp.cpp:34-43); Vcvis from VCRITVIS,
        ZCRITVIS, VCRIT or ZCRIT in that order (src/OCPFuncPVT.cpp:428-448);
        BIC as a lower triangle or a full matrix (src/OCPEoS.cpp:33-45);
        water from PVTW and GRAVITY / DENSITY
        (src/OCPMixtureMethodComp.cpp:655-664)."""
        n = self.num_com
        if n <= 0:
            raise DeckError("no COMPS / NCOMPS in the deck")
        if n > abi.MAX_NC:
            raise DeckError(f"{n} components; the kernels are built for at most {abi.MAX_NC}")
        for kw in ("TCRIT", "PCRIT", "MW", "ACF"):
            if kw not in self.comp:
                raise DeckError(f"{kw} hasn't been input")
        if "VCRIT" not in self.comp and "ZCRIT" not in self.comp:
            raise DeckError("VCRIT or ZCRIT hasn't been input")
        for kw, m in FLASH_KEYWORDS.items():
            if len(self.flash.get(kw, [])) < m:
                raise DeckError(f"{kw} hasn't been input (the flash controls have no defaults)")
        fi = self.field_units
        R = FIELD_GAS_CONSTANT if fi else METRIC_GAS_CONSTANT
        Tc, Pc = self._region("TCRIT", region), self._region("PCRIT", region)
        vol = FT3_PER_LBMOL if fi else 1.0
        if "VCRIT" in self.comp:
            Vc = self._region("VCRIT", region)
        else:
            Vc = R * self._region("ZCRIT", region) * Tc / Pc
        if "VCRITVIS" in self.comp:
            Vcvis = self._region("VCRITVIS", region)
        elif "ZCRITVIS" in self.comp:
            Vcvis = R * self._region("ZCRITVIS", region) * Tc / Pc
        elif "VCRIT" in self.comp:
            Vcvis = Vc
        else:
            # the reference uses the FIELD gas constant here in every unit system
            Vcvis = FIELD_GAS_CONSTANT * self._region("ZCRIT", region) * Tc / Pc
        OA = self._region("OMEGAA", region)
        OB = self._region("OMEGAB", region)
        OA = np.full(n, OMEGA_A) if OA is None else OA
        OB = np.full(n, OMEGA_B) if OB is None else OB
        bic = list(self.bic[region]) if region < len(self.bic) else []
        kij = np.zeros((n, n))
        if len(bic) == n * n:
            kij = np.array(bic).reshape(n, n)
        else:
            it = iter(bic + [0.0] * (n * (n - 1) // 2))
            for i in range(1, n):
                for j in range(i):
                    kij[i, j] = kij[j, i] = next(it)
        f = abi.Fluid()
        f.nc = n
        f.lbc_exact = lbc_exact
        TcK = Tc * (5.0 / 9.0) if fi else Tc
        Pcb = Pc * PSI if fi else Pc
        sh = self._region("SSHIFT", region)
        for i in range(n):
            f.Tc[i], f.Pc[i] = TcK[i], Pcb[i]
            f.Vc[i], f.Vcvis[i] = Vc[i] * vol, Vcvis[i] * vol
            f.MW[i] = self.comp["MW"][region][i]
            f.acf[i] = self.comp["ACF"][region][i]
            f.OmegaA[i], f.OmegaB[i] = OA[i], OB[i]
            f.Vshift[i] = 0.0 if sh is None else sh[i]
            for k in range(n):
                f.BIC[i * abi.MAX_NC + k] = kij[i, k]
        for k in range(5):
            f.LBCcoef[k] = self.lbccoef[k]
        pvtw = self.tables.get("PVTW")
        if not pvtw:
            raise DeckError("PVTW is missing")
        row = [c[0] for c in pvtw[min(region, len(pvtw) - 1)]]
        if fi:
            row = [row[0] * PSI, row[1], row[2] / PSI, row[3], row[4] / PSI]
        for k in range(5):
            f.PVTW[k] = row[k]
        f.rhoW_std = METRIC_RHOW_STD * (1.0 if self.density_active else self.gravity[1])
        g = lambda kw, i: self.flash[kw][i]
        f.RR_maxIt, f.RR_tol = int(g("RR", 0)), float(g("RR", 1))
        f.SSMsta_maxIt, f.SSMsta_tol, f.SSMsta_eYt = int(g("SSMSTA", 0)), float(g("SSMSTA", 1)), float(g("SSMSTA", 2))
        f.NRsta_maxIt, f.NRsta_tol = int(g("NRSTA", 0)), float(g("NRSTA", 1))
        f.SSMsp_maxIt, f.SSMsp_tol = int(g("SSMSP", 0)), float(g("SSMSP", 1))
        f.NRsp_maxIt, f.NRsp_tol = int(g("NRSP", 0)), float(g("NRSP", 1))
        f.skipPSA = 1 if skip_psa else 0
        return f

    def to_sat(self, region=0):
        """ocpx_sat from SWOF and SGOF (capillary pressure in bar)."""
        s = abi.Sat()
        for kw, dst, cnt in (("SWOF", s.swof, "nswof"), ("SGOF", s.sgof, "nsgof")):
            tabs = self.tables.get(kw)
            if not tabs:
                raise DeckError(f"{kw} is missing")
            cols = tabs[min(region, len(tabs) - 1)]
            nr = len(cols[0])
            if nr > abi.MAX_TABROWS:
                raise DeckError(f"{kw}: {nr} rows, at most {abi.MAX_TABROWS}")
            pc = PSI if self.field_units else 1.0
            for r in range(nr):
                dst[4 * r:4 * r + 4] = [cols[0][r], cols[1][r], cols[2][r], cols[3][r] * pc]
            setattr(s, cnt, nr)
        return s

    def to_rock(self, region=0):
        """ocpx_rock (reference pressure in bar, compressibilities per bar)."""
        rk = self.rock[min(region, len(self.rock) - 1)] if self.rock else \
            {"type": "LINEAR", "Pref": 14.7, "cp1": 3.406e-6, "cp2": 0.0}
        r = abi.Rock()
        if self.field_units:
            r.Pref, r.cp1, r.cp2 = rk["Pref"] * PSI, rk["cp1"] / PSI, rk["cp2"] / PSI ** 2
        else:
            r.Pref, r.cp1, r.cp2 = rk["Pref"], rk["cp1"], rk["cp2"]
        return r

    def temperature_c(self):
        """Reservoir temperature (RTEMP) in degC."""
        return (self.rtemp - 32.0) * 5.0 / 9.0 if self.field_units else self.rtemp

    def to_grid(self):
        """Cartesian grid in METRIC: depth and volume as CalDepthVOrthogonalGrid,
        connections x+, y+, z+ per cell, BulkConnTransMethod01 transmissibilities."""
        if self.dimens is None:
            raise DeckError("DIMENS is missing")
        nx, ny, nz = self.dimens
        L = FT if self.field_units else 1.0
        get = lambda kw: self.arrays.get(kw)
        for kw in ("DX", "DY", "DZ", "TOPS", "PORO", "PERMX"):
            a = get(kw)
            if a is None or a.size < (self.nxny if kw == "TOPS" else self.nb):
                raise DeckError(f"{kw} does not cover the grid")
        dx, dy, dz = (get(k)[:self.nb] * L for k in ("DX", "DY", "DZ"))
        tops = get("TOPS")[:self.nxny] * L
        kx = get("PERMX")[:self.nb]
        ky = get("PERMY")[:self.nb] if get("PERMY") is not None else kx
        kz = get("PERMZ")[:self.nb] if get("PERMZ") is not None else kx
        ntg = get("NTG")[:self.nb] if get("NTG") is not None else np.ones(self.nb)
        depth = np.zeros(self.nb)
        depth[:self.nxny] = tops + dz[:self.nxny] / 2
        for k in range(1, nz):
            a, b = k * self.nxny, (k + 1) * self.nxny
            depth[a:b] = depth[a - self.nxny:b - self.nxny] + dz[a - self.nxny:b - self.nxny] / 2 + dz[a:b] / 2
        v = dx * dy * dz
        idx = np.arange(self.nb)
        i, j, k = idx % nx, (idx // nx) % ny, idx // self.nxny
        cbs, ces, trs, dirs = [], [], [], []
        for dcode, mask, off, kk, area, use_ntg in (
                (0, i < nx - 1, 1, kx, lambda c: 2 * dy[c] * dz[c] / dx[c], True),
                (1, j < ny - 1, nx, ky, lambda c: 2 * dz[c] * dx[c] / dy[c], True),
                (2, k < nz - 1, self.nxny, kz, lambda c: 2 * dx[c] * dy[c] / dz[c], False)):
            b = idx[mask]
            e = b + off
            t1 = (ntg[b] if use_ntg else 1.0) * kk[b] * area(b)
            t2 = (ntg[e] if use_ntg else 1.0) * kk[e] * area(e)
            cbs.append(b); ces.append(e); trs.append(1.0 / (1 / t1 + 1 / t2) * W.DARCY)
            dirs.append(np.full(b.size, dcode))
        cb, ce, tr, dr = (np.concatenate(a) for a in (cbs, ces, trs, dirs))
        order = np.lexsort((dr, cb))
        return W.Grid(nx, ny, nz, v, depth, get("PORO")[:self.nb].astype(float).copy(),
                      cb[order].astype(np.int32), ce[order].astype(np.int32), tr[order])

    def well_perforations(self, name):
        """Perforations of a well from WELSPECS and COMPDAT on the Cartesian
        grid (WellParam, src/ParamWell.cpp:50-127; PeacemanWell::InputPerfo,
        src/WellPeaceman.cpp:15-53): [(cell, WI)] with WI in METRIC, CONV_DARCY
        included, by PeacemanWell::CalWI unless COMPDAT gives it (a given WI
        is taken times CONV1 as the reference does, then converted)."""
        from . import wells as WL
        spec = [w for w in self.welspecs if w[0] == name]
        if not spec:
            raise DeckError(f"well {name} is not in WELSPECS")
        I, J = int(spec[-1][2]), int(spec[-1][3])
        nx, ny, nz = self.dimens
        fi = self.field_units
        L = FT if fi else 1.0
        out = []
        for rec in self.compdat:
            pat = rec[0]
            star = pat.find("*")
            if not (name == pat or (star >= 0 and name[:star] == pat[:star])):
                continue
            v = rec + ["DEFAULT"] * (10 - len(rec))
            v = ["DEFAULT" if x == "/" else x for x in v]
            i, j = (I, J) if v[1] == "DEFAULT" or v[2] == "DEFAULT" else (int(v[1]), int(v[2]))
            for k in range(int(v[3]), int(v[4]) + 1):
                n = (k - 1) * nx * ny + (j - 1) * nx + (i - 1)
                if v[5] != "DEFAULT":
                    WI = _float(v[5], "COMPDAT") * (5.61458 * FT ** 3 / PSI if fi else 1.0)
                else:
                    diam = _float(v[6], "COMPDAT") if v[6] != "DEFAULT" else 1.0
                    kh = _float(v[7], "COMPDAT") * L if v[7] != "DEFAULT" else -1.0
                    skin = _float(v[8], "COMPDAT") if v[8] != "DEFAULT" else 0.0
                    d = v[9].lower() if v[9] != "DEFAULT" else "z"
                    a = lambda kw: self.arrays[kw][n]
                    ntg = self.arrays["NTG"][n] if "NTG" in self.arrays else 1.0
                    ky = a("PERMY") if "PERMY" in self.arrays else a("PERMX")
                    kz = a("PERMZ") if "PERMZ" in self.arrays else a("PERMX")
                    WI = WL.peaceman_wi(d, a("DX") * L, a("DY") * L, a("DZ") * L * ntg, a("PERMX"), ky, kz,
                                        diam / 2 * L, skin, kh)
                out.append((n, WI))
        if not out:
            raise DeckError(f"well {name} has no COMPDAT perforation")
        return out

    def well_controls(self, name, step=0):
        """The WCONPROD / WCONINJE record of a well in force at a schedule step,
        the step counting TSTEP values as the reference's criticalTime does.  A
        name with '*' matches every well whose name contains the text before the
        '*' (ParamWell::InputWCONPROD / InputWCONINJE, src/ParamWell.cpp:222-286:
        name.find(src))."""
        rec = None
        for st, kw, r in self.schedule:
            if st > step:
                break
            pat = r[0]
            star = pat.find("*")
            if name == pat or (star >= 0 and pat[:star] in name):
                rec = (kw, r)
        return rec

    def to_wells(self, step=0):
        """(PeacemanWell, WellLimits) for every well with a control at the
        step, in METRIC: WellOpt (src/WellOpt.cpp:14-70) -- producers ORAT /
        GRAT / WRAT / LRAT (rate control, phase weights) or BHP with the
        minimum BHP as target; injectors of water or a WELLSTRE stream, RATE
        or BHP with the maximum BHP as target.  FIELD rates are stb/day
        (oil, water) or Mscf/day (gas), converted to m3/day.  bhp starts at
        the target BHP; dG is zero and the rate weights are left to
        wells.rate_factors."""
        from . import wells as WL
        fi = self.field_units
        stb, mscf = (0.158987294928, 28.316846592) if fi else (1.0, 1.0)
        pconv = PSI if fi else 1.0
        weights = {"ORAT": (1.0, 0.0, 0.0), "GRAT": (0.0, 1.0, 0.0), "WRAT": (0.0, 0.0, 1.0),
                   "LRAT": (1.0, 0.0, 1.0)}
        nc = self.num_com + 1
        out = []
        names = []
        for w in self.welspecs:
            if w[0] not in names:
                names.append(w[0])
        for name in names:
            ctl = self.well_controls(name, step)
            if ctl is None:
                continue
            kw, r = ctl
            cells, wi = zip(*self.well_perforations(name))
            if kw == "WCONPROD":
                state, mode = r[1], r[2]
                if state != "OPEN":
                    continue
                rate = 1e10 if r[3] == "DEFAULT" else _float(r[3], kw)
                minbhp = _float(r[4], kw) * pconv
                if mode == "BHP":
                    lim = WL.WellLimits(WL.BHP, rate, minbhp, minbhp)
                    target = minbhp
                else:
                    wgt = weights[mode]
                    rate *= mscf if mode == "GRAT" else stb
                    lim = WL.WellLimits(WL.RATE, rate, minbhp, minbhp, weights=wgt)
                    target = rate
                well = WL.PeacemanWell(name, WL.PROD, lim.init_mode, bhp=minbhp, target=target, cells=list(cells),
                                       WI=list(wi), dG=[0.0] * len(cells))
            else:
                fluid, state, mode = r[1].upper(), r[2], r[3]
                if state != "OPEN":
                    continue
                water = fluid in ("WAT", "WATER")
                zi = [0.0] * nc
                if water:
                    zi[nc - 1] = 1.0
                else:
                    comp = {k.upper(): v for k, v in self.wellstre.items()}.get(fluid)
                    if comp is None:
                        raise DeckError(f"{name}: injected fluid {r[1]} is neither water nor a WELLSTRE stream")
                    for i, v in enumerate(comp[:nc - 1]):
                        zi[i] = v
                rate = 1e10 if r[4] == "DEFAULT" else _float(r[4], kw) * (stb if water else mscf)
                maxbhp = _float(r[5], kw) * pconv
                init = WL.BHP if mode == "BHP" else WL.RATE
                lim = WL.WellLimits(init, rate, maxbhp, maxbhp, inj_T=self.temperature_c(), water=water)
                well = WL.PeacemanWell(name, WL.INJ, init, bhp=maxbhp, target=maxbhp if init == WL.BHP else rate,
                                       cells=list(cells), WI=list(wi), dG=[0.0] * len(cells), zi=zi,
                                       factor=[1.0] * nc)
            out.append((well, lim))
        return out

    def composition(self, depth_m, region=0):
        """Hydrocarbon mole fractions per cell from ZMFVD (depth-interpolated,
        constant outside the table, normalised)."""
        tabs = self.tables.get("ZMFVD")
        if not tabs:
            raise DeckError("ZMFVD is missing")
        cols = tabs[min(region, len(tabs) - 1)]
        d = np.array(cols[0]) * (FT if self.field_units else 1.0)
        z = np.stack([np.interp(depth_m, d, np.array(c)) for c in cols[1:]], axis=1)
        return z / z.sum(axis=1, keepdims=True)

    def datum(self):
        """(datum depth m, pressure bar) of the first EQUIL record."""
        if not self.equil:
            raise DeckError("EQUIL is missing")
        dd, pd = self.equil[0][0], self.equil[0][1]
        return (dd * FT, pd * PSI) if self.field_units else (dd, pd)

    def to_workload(self, seed=0, perturb=True, skip_psa=True, P=None):
        """The deck's grid, fluid, tables and rock as a workload.  Pressure
        (bar) is P if given, else the EQUIL datum pressure plus a
        water-gradient hydrostatic term; the composition is ZMFVD at each
        cell's depth (the reference's equilibration is not reproduced)."""
        rng = np.random.Generator(np.random.PCG64(seed))
        fl = self.to_fluid(skip_psa=skip_psa)
        rock = self.to_rock()
        grid = self.to_grid()
        if P is None:
            dd, pd = self.datum()
            P = pd + fl.rhoW_std * METRIC_GRAVITY_FACTOR * (grid.depth - dd)
        P = np.broadcast_to(np.asarray(P, float), (grid.nb,)).copy()
        z = self.composition(grid.depth)
        st = W.initial_state(fl, rock, grid, P, z, rng, perturb=perturb,
                             T=self.temperature_c() + W.CONV4)
        name = f"{os.path.basename(self.path)} {grid.nx}x{grid.ny}x{grid.nz}, Nc={self.num_com}"
        return W.Workload(name, fl, self.to_sat(), rock, grid, st)


# ---------------------------------------------------------------- handlers
def _units(name):
    def h(d, L, t):
        d.units = name
    return h


def _comps(d, L, t):
    v = L.need(t[0])
    d.num_com = int(v[0])


def _cnames(d, L, t):
    """Names on the keyword line are taken too (the reference starts
    reading on the next line, src/ParamReservoir.cpp:958-973, so a one-line
    CNAMES there swallows the keyword that follows it)."""
    d._need_comps("CNAMES")
    names = [x for x in t[1:] if x != "/"]
    done = bool(t[1:]) and t[-1] == "/"
    while not done:
        v = L.need("CNAMES")
        if v[0] == "/":
            break
        names += [x for x in v if x != "/"]
        done = v[-1] == "/"
    d.cnames = names


def _bic(d, L, t):
    d._need_comps("BIC")
    regs = [[] for _ in range(d.ntpvt)]
    r = 0
    while r < d.ntpvt:
        v = L.need("BIC")
        if v[0] == "/":
            r += 1
            continue
        regs[r] += [_float(x, "BIC") for x in v if x != "/"]
        if v[-1] == "/":
            r += 1
    d.bic = regs


def _lbccoef(d, L, t):
    v = deal_default(L.need("LBCCOEF"))
    for i in range(min(5, len(v))):
        if v[i] not in ("DEFAULT", "/"):
            d.lbccoef[i] = _float(v[i], "LBCCOEF")


def _flash_ctrl(d, L, t):
    v = L.need(t[0])
    d.flash[t[0]] = [x for x in v if x != "/"]


def _rtemp(d, L, t):
    v = L.need("RTEMP")
    if v[0] != "/":
        d.rtemp = _float(v[0], "RTEMP")


def _rock(d, L, t):
    while True:
        v = L.need("ROCK")
        if v[0] == "/":
            break
        rk = {"type": v[0], "Pref": _float(v[1], "ROCK"), "cp1": _float(v[2], "ROCK"), "cp2": 0.0}
        if v[0] == "LINEAR02":
            rk["cp2"] = _float(v[3], "ROCK") if len(v) > 3 and v[3] != "/" else rk["cp1"]
        d.rock.append(rk)


def _three(attr):
    def h(d, L, t):
        if attr == "gravity":                    # active even without values
            d.gravity_active = True
        v = L.need(t[0])
        if v[0] == "/":
            return
        v = deal_default(v)
        if v[-1] == "/":
            v = v[:-1]
        if len(v) != 3:
            raise DeckError(f"Wrong Keyword {t[0]}!")
        vals = getattr(d, attr)
        for i in range(3):
            if v[i] != "DEFAULT":
                vals[i] = _float(v[i], t[0])
                if attr == "density":
                    d.density_active = True
    return h


def _equil(d, L, t):
    v = L.need("EQUIL")
    if v[0] == "/":
        return
    v = deal_default(v)
    d.equil.append([_float(v[i], "EQUIL") if i < len(v) and v[i] not in ("DEFAULT", "/") else 0.0
                    for i in range(6)])


def _tabdims(d, L, t):
    v = L.need("TABDIMS")
    if len(v) < 3:
        raise DeckError("TABDIMS: input the number of saturation, PVT and rock tables in turn")
    d.ntsfun, d.ntpvt, d.ntrooc = int(v[0]), int(v[1]), int(v[2])


def _dimens(d, L, t):
    v = L.need("DIMENS")
    d.dimens = (int(v[0]), int(v[1]), int(v[2]))


def _include(d, L, t):
    v = deal_default(L.need("INCLUDE"))
    L.include(v[0])


def _records(attr):
    def h(d, L, t):
        while True:
            v = L.need(t[0])
            if v[0] == "/":
                break
            getattr(d, attr).append(deal_default(v))
    return h


def _controls(d, L, t):
    # WellOptParam (src/ParamWell.cpp:14-47): the records apply from the current schedule step
    while True:
        v = L.need(t[0])
        if v[0] == "/":
            break
        d.schedule.append((d.step, t[0], deal_default(v)))


def _wellstre(d, L, t):
    while True:
        v = L.need("WELLSTRE")
        if v[0] == "/":
            break
        d.wellstre[v[0]] = [_float(x, "WELLSTRE") for x in v[1:] if x != "/"]


def _tstep(d, L, t):
    # ParamWell::InputTSTEP (src/ParamWell.cpp:288-307): every time value adds one
    # critical time, and the controls read next apply from index criticalTime.size() - 1
    while True:
        v = L.need("TSTEP")
        if v[0] == "/":
            break
        v = deal_default(v)
        d.step += sum(1 for x in v if x != "/")
        if v[-1] == "/":
            break


_HANDLERS = {
    "FIELD": _units("FIELD"), "METRIC": _units("METRIC"),
    "COMPS": _comps, "NCOMPS": _comps, "CNAMES": _cnames, "BIC": _bic, "LBCCOEF": _lbccoef,
    "RR": _flash_ctrl, "SSMSTA": _flash_ctrl, "NRSTA": _flash_ctrl, "SSMSP": _flash_ctrl, "NRSP": _flash_ctrl,
    "RTEMP": _rtemp, "ROCK": _rock, "GRAVITY": _three("gravity"), "DENSITY": _three("density"),
    "EQUIL": _equil, "TABDIMS": _tabdims, "DIMENS": _dimens, "INCLUDE": _include,
    "EQUALS": lambda d, L, t: d._equals(L, "EQUALS", "set"),
    "MULTIPLY": lambda d, L, t: d._equals(L, "MULTIPLY", "mul"),
    "COPY": lambda d, L, t: d._copy(L, "COPY"),
    "WELSPECS": _records("welspecs"), "COMPDAT": _records("compdat"),
    "WCONPROD": _controls, "WCONINJE": _controls, "WELLSTRE": _wellstre, "TSTEP": _tstep,
}


def read_deck(path):
    """Parse a deck file (and its INCLUDE files)."""
    return Deck.read(path)
