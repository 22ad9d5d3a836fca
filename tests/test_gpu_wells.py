"""Peaceman wells on the device (ocpx_set_wells / ocpx_well_terms) against
the oracle (orc_wells) from identical flash states: the residual and
right-hand side with the perforation rates, the norms, the bulk diagonal
blocks with the perforation terms, the well blocks, the well rows and the
component rates.  Wells: a BHP producer and a rate injector through every
layer, a one-perforation rate producer, and a BHP producer sharing a cell
with the first one; then the SPE5 deck's own two wells (WELSPECS /
COMPDAT, Peaceman index by ocpx_peaceman_wi)."""
import os

import numpy as np
import pytest

from opencaeporox_amd import deck
from opencaeporox_amd import wells as WL
from opencaeporox_amd.fim import IsoTFIM
from oracle.oracle import state_from_workload
from tests.conftest import close
from tests.test_gpu_parity import _clean_grid, _upload_flash
from tests.test_oracle_wells import _wells

pytestmark = pytest.mark.gpu

DT = 1.0
DECKS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "decks")


def _check(dev, oracle, wl, wells, tag):
    b = wl.Nc + 2
    st = state_from_workload(wl)
    oracle.cal_flash(wl.fluid, st)
    _upload_flash(dev, st)
    dev.CalKrPc(); dev.CalRock()
    oracle.cal_krpc(wl.sat, st); oracle.cal_rock(wl.rock, st)
    dev.SetWells(wells)
    # residual only
    dev.CalRes(DT, fetch=False)
    nd, wd = dev.CalResWells(DT, jacobian=False)
    res_o, _ = oracle.cal_res(wl.fluid, st, wl.grid, DT)
    no, wo, qo, *_ = oracle.wells(wl.fluid, st, DT, wells, res_o)
    rhs = dev.bsr()[3]
    for name, a, e in (("rhs", rhs, res_o), ("well rows", wd, wo), ("rates", dev.wells_state()[1], qo)):
        ok, worst = close(a, e)
        assert ok, f"{tag} residual pass: {name} worst {worst}"
    assert nd[1] == no[1] and nd[3] == no[3], f"{tag}: norm cells {nd} vs {no}"
    assert np.allclose(nd[[0, 2]], no[[0, 2]], rtol=1e-10)
    # residual + Jacobian
    dev.AssembleMatBulks(DT)
    nd, wd = dev.CalResWells(DT, jacobian=True)
    res_o, _ = oracle.cal_res(wl.fluid, st, wl.grid, DT)
    A = oracle.assemble(wl.fluid, st, wl.grid, DT)
    no, wo, qo, Bo, Co, Do = oracle.wells(wl.fluid, st, DT, wells, res_o, A)
    rp, col, val, rhs = dev.bsr()
    Bd, Cd, Dd = dev.well_blocks()
    assert np.array_equal(rp, A[0]) and np.array_equal(col, A[1])
    for name, a, e in (("rhs", rhs, res_o), ("well rows", wd, wo), ("Bnw", Bd, Bo), ("Cwn", Cd, Co),
                       ("Dw", Dd, Do)):
        ok, worst = close(a, e)
        assert ok, f"{tag} Jacobian pass: {name} worst {worst}"
    V, Vo = val.reshape(-1, b, b), A[2]
    for r in range(len(rp) - 1):
        ok, worst = close(V[rp[r]:rp[r + 1]], Vo[rp[r]:rp[r + 1]])
        assert ok, f"{tag}: block row {r} worst {worst}"
    # the assembled system with its well rows is the oracle's
    Jd, fd = WL.full_system(rp, col, val, rhs, wells, Bd, Cd, Dd, wd, b)
    Jo, fo = WL.full_system(A[0], A[1], A[2], res_o, wells, Bo, Co, Do, wo, b)
    assert close(Jd.toarray(), Jo.toarray())[0] and close(fd, fo)[0]


@pytest.mark.parametrize("mode_prod", [WL.BHP, WL.RATE])
def test_wells_match_oracle(oracle, mode_prod):
    wl = _clean_grid(seed=31, Nc=3, nx=4, ny=3, nz=3)
    wells = _wells(wl, wl.Nc + 1, mode_prod)
    first = wells[0].cells[0]
    wells.append(WL.PeacemanWell("P3", WL.PROD, WL.BHP, bhp=float(wl.state["P"][first]) - 3.0, target=100.0,
                                 cells=[first], WI=[7.5], dG=[0.0]))
    dev = IsoTFIM.from_workload(wl)
    _check(dev, oracle, wl, wells, f"grid mode={mode_prod}")
    dev.close()


def test_spe5_deck_wells(oracle):
    """The SPE5 deck's own wells at schedule step 3 (criticalTime index) (WELSPECS /
    COMPDAT / WCONPROD / WCONINJE / WELLSTRE): the ORAT producer with rate
    weights from the device surface flash of its mobility stream, the
    solvent injector with its molar density from ocpx_xi_phase."""
    d = deck.read_deck(os.path.join(DECKS, "spe5", "spe5.data"))
    _, pd = d.datum()
    rng = np.random.Generator(np.random.PCG64(3))
    wl = d.to_workload(seed=3, P=pd * rng.uniform(0.3, 1.1, d.nb))
    nc = wl.Nc + 1
    wells = [w for w, _ in d.to_wells(step=3)]
    assert [w.name for w in wells] == ["PROD1", "I"]
    dev = IsoTFIM.from_workload(wl)
    st = state_from_workload(wl)
    oracle.cal_flash(wl.fluid, st)
    _upload_flash(dev, st)
    dev.CalKrPc()
    prod, inj = wells
    stream = WL.factor_stream(np.zeros(nc), WL.cell_state(dev, prod.cells), prod)
    prod.factor = list(np.full(nc, WL.rate_factors(wl.fluid, stream[None, :], np.array([[1.0, 0.0, 0.0]]))[0]))
    P = wl.state["P"]
    prod.bhp = float(P[prod.cells[0]]) - 20.0
    inj.bhp = float(P[inj.cells[0]]) + 30.0
    inj.xi = [WL.xi_phase(wl.fluid, inj.bhp, d.temperature_c(), inj.zi[:nc - 1])]
    _check(dev, oracle, wl, wells, "spe5")
    dev.close()


def test_set_wells_rejects_bad_input():
    from opencaeporox_amd.fim import OcpxError
    wl = _clean_grid(seed=31, Nc=3, nx=4, ny=3, nz=2)
    dev = IsoTFIM.from_workload(wl)
    bad = WL.PeacemanWell("X", WL.PROD, WL.BHP, bhp=100.0, target=90.0, cells=[0, 0], WI=[1.0, 1.0])
    with pytest.raises(OcpxError, match="twice"):
        dev.SetWells([bad])
    bad = WL.PeacemanWell("X", WL.PROD, WL.BHP, bhp=100.0, target=90.0, cells=[wl.grid.nb], WI=[1.0])
    with pytest.raises(OcpxError, match="outside the grid"):
        dev.SetWells([bad])
    with pytest.raises(OcpxError, match="no wells"):
        dev.CalResWells(DT)
    dev.close()


def test_well_terms_once_per_residual():
    """ocpx_well_terms adds the wells to the residual and Jacobian in place: a
    second call on the same residual, or a call after the state moved, is an
    error instead of a double count."""
    from opencaeporox_amd.fim import BATCHED, OcpxError
    wl, wells = _newton_case()
    dev = IsoTFIM.from_workload(wl)
    dev.SetWells(wells)
    dev.newton_pass(DT, BATCHED)
    _, w1 = dev.CalResWells(DT, jacobian=False)
    with pytest.raises(OcpxError, match="already"):
        dev.CalResWells(DT, jacobian=True)
    dev.AssembleMatBulks(DT)                      # a new residual + Jacobian takes them again
    _, w2 = dev.CalResWells(DT, jacobian=True)
    assert np.array_equal(w1, w2)
    with pytest.raises(OcpxError, match="already"):
        dev.CalResWells(DT, jacobian=False)
    u, it, _ = dev.LinearSolve(rtol=1e-6, maxit=500)
    assert it > 0
    dev.GetSolution(None)
    with pytest.raises(OcpxError, match="state changed"):
        dev.CalResWells(DT, jacobian=False)
    dev.close()


def _newton_case():
    """The hydrostatic grid of the device Newton test with a rate producer
    (90% of its initial rate) and a BHP injector of the reservoir's own
    composition 0.5 bar above the highest pressure it perforates.  With a
    direct solve the oracle's Newton sequence on this step is max |R| 4.5e3,
    28, 8.7e-2, 1.2e-4, 1.7e-8, the well rows falling to 2.5e-10."""
    from tests.test_gpu_solver import _smooth_grid
    from oracle.oracle import Oracle
    wl = _smooth_grid()
    nc = wl.Nc + 1
    g = wl.grid
    col = lambda i, j: [i + j * g.nx + k * g.nx * g.ny for k in range(g.nz)]
    P = wl.state["P"]
    Ni = wl.state["NI"].reshape(-1, nc)[0]
    zi = np.zeros(nc)
    zi[:nc - 1] = Ni[:nc - 1] / Ni[:nc - 1].sum()
    top = float(P[col(7, 7)].max()) + 0.5
    wells = [WL.PeacemanWell("P", WL.PROD, WL.RATE, bhp=float(P[col(0, 0)].min()) - 2, target=0.0, cells=col(0, 0),
                             WI=[2.0] * g.nz, dG=list(0.07 * np.arange(g.nz)), factor=[0.05] * nc),
             WL.PeacemanWell("I", WL.INJ, WL.BHP, bhp=top, target=top, cells=col(7, 7), WI=[2.0] * g.nz,
                             dG=list(0.05 * np.arange(g.nz)), xi=[8.0] * g.nz, zi=list(zi))]
    o = Oracle()
    st = state_from_workload(wl)
    o.cal_flash(wl.fluid, st); o.cal_krpc(wl.sat, st); o.cal_rock(wl.rock, st)
    res, _ = o.cal_res(wl.fluid, st, wl.grid, DT)
    _, _, qw, *_ = o.wells(wl.fluid, st, DT, wells, res)
    wells[0].target = 0.9 * float(qw[0] @ np.array(wells[0].factor))
    return wl, wells


def test_linear_solve_with_well_rows():
    import scipy.sparse.linalg as spla
    from opencaeporox_amd.fim import BATCHED
    wl, wells = _newton_case()
    dev = IsoTFIM.from_workload(wl)
    dev.SetWells(wells)
    dev.newton_pass(DT, BATCHED)
    _, wres = dev.CalResWells(DT, jacobian=True)
    rp, col, val, rhs = dev.bsr()
    Bd, Cd, Dd = dev.well_blocks()
    J, f = WL.full_system(rp, col, val, rhs, wells, Bd, Cd, Dd, wres, dev.bs)
    u, it, rr = dev.LinearSolve(rtol=1e-12, maxit=2000)
    assert u.size == (dev.nb + len(wells)) * dev.bs
    true = np.linalg.norm(f - J @ u) / np.linalg.norm(f)
    ref = spla.spsolve(J.tocsc(), f)
    err = np.linalg.norm(u - ref) / np.linalg.norm(ref)
    print(f"with wells: {it} iterations, true residual {true:.2e}, |u - u_direct| / |u_direct| {err:.2e}")
    assert it > 0 and true < 1e-11 and err < 1e-8
    dev.close()


def test_newton_iterations_with_wells_on_device():
    """assemble -> well terms -> device solve -> device GetSolution (bulk
    update and bhp += u of the well rows), repeated; the same iterations with
    u through the host and ocpx_set_well_bhp give the same bits."""
    from opencaeporox_amd.fim import BATCHED
    wl, wells = _newton_case()
    a, b = IsoTFIM.from_workload(wl), IsoTFIM.from_workload(wl)
    a.SetWells(wells)
    b.SetWells(wells)
    nb, bs = a.nb, a.bs
    bulk, wrow = [], []
    for k in range(5):
        a.newton_pass(DT, BATCHED)
        b.newton_pass(DT, BATCHED)
        _, wa = a.CalResWells(DT, jacobian=True)
        _, wb = b.CalResWells(DT, jacobian=True)
        bulk.append(np.abs(a.bsr()[3]).max())
        wrow.append(np.abs(wa).max())
        u, it, rr = a.LinearSolve(rtol=1e-10, maxit=2000)
        assert it > 0
        a.GetSolution(None, 0.2)
        b.GetSolution(u[:nb * bs], 0.2)
        b.set_well_bhp(b.wells_state()[0] + u[nb * bs::bs])
        for f in ("P", "NI", "S"):
            assert np.array_equal(a.get(f), b.get(f)), f"iteration {k} field {f}"
        assert np.array_equal(a.wells_state()[0], b.wells_state()[0]), f"iteration {k} bhp"
    print("max |R| bulk:", " ".join(f"{v:.2e}" for v in bulk), "| well rows:", " ".join(f"{v:.2e}" for v in wrow))
    assert bulk[-1] < 1e-6 * bulk[0] and wrow[-1] < 1e-6 * max(wrow[0], 1.0)
    a.close()
    b.close()


@pytest.mark.parametrize("name", ["spe5", "spe3"])
def test_rate_factors_match_oracle(oracle, name):
    """CalFactor: the produced stream flashed at surface conditions on the
    device against the oracle's flash of the same moles."""
    d = deck.read_deck(os.path.join(DECKS, name, f"{name}.data"))
    fl = d.to_fluid()
    z = d.composition(np.array([2300.0]))[0]
    qi = np.array([list(z * 100.0) + [5.0], list(z * 40.0) + [0.0], list(z * 7.0) + [20.0]])
    w = np.array([[1, 0, 0], [0, 1, 0], [1, 0, 1]], float)       # ORAT, GRAT, LRAT
    fd = WL.rate_factors(fl, qi, w)
    wl = WL.surface_workload(fl, qi)
    st = state_from_workload(wl)
    oracle.cal_flash(wl.fluid, st)
    fo = WL.factors_from_flash(st.a["S"], st.a["VF"], qi, w)
    print(name, "factors", fd)
    assert np.all(fd > 0)
    ok, worst = close(fd, fo)
    assert ok, f"{name}: worst {worst}"


def test_producer_heads_match_oracle(oracle):
    """CalProddG01: the produced stream flashed segment by segment up the
    well (device flash) against the same loop over the oracle's states and
    flash; the heads grow with depth."""
    from opencaeporox_amd.fim import BATCHED
    wl, wells = _newton_case()
    prod = wells[0]
    nb, nc = wl.grid.nb, wl.Nc + 1
    dev = IsoTFIM.from_workload(wl)
    dev.CalFlash(BATCHED)
    dev.CalKrPc()
    st = state_from_workload(wl)
    oracle.cal_flash(wl.fluid, st)
    oracle.cal_krpc(wl.sat, st)
    depth = wl.grid.depth[prod.cells]
    head = float(depth[0]) - 25.0
    seg = WL.SegmentFlash(wl.fluid)
    dG_d = WL.producer_heads(WL.cell_state(dev, prod.cells), prod, head, depth, seg)
    seg.close()
    c = np.asarray(prod.cells)
    cs = {"P": st.a["P"][c], "T": st.a["T"][c], "PHASEEXIST": st.a["PHASEEXIST"].reshape(nb, 3)[c].astype(bool),
          "KR": st.a["KR"].reshape(nb, 3)[c], "MU": st.a["MU"].reshape(nb, 3)[c], "XI": st.a["XI"].reshape(nb, 3)[c],
          "XIJ": st.a["XIJ"].reshape(nb, 3, nc)[c], "NI": st.a["NI"].reshape(nb, nc)[c]}

    def oflash(P, T, Ni):
        w2 = WL.surface_workload(wl.fluid, np.asarray(Ni, float)[None, :], P, T)
        s2 = state_from_workload(w2)
        oracle.cal_flash(w2.fluid, s2)
        return s2.a["PHASEEXIST"].astype(bool), s2.a["S"] * s2.a["VF"][0], s2.a["RHO"]

    dG_o = WL.producer_heads(cs, prod, head, depth, oflash)
    print("producer heads (bar):", dG_d)
    assert np.all(np.diff(dG_d) > 0) and dG_d[0] > 0
    ok, worst = close(dG_d, dG_o)
    assert ok, f"worst {worst}"
    dev.close()


def test_producer_mode_check_matches_oracle(oracle):
    """CheckOptMode for a producer: the surface oil rate at minBHP from the
    device flash against the same check over the oracle's states and flash;
    the limit decides the control."""
    from opencaeporox_amd.fim import BATCHED
    wl, wells = _newton_case()
    prod = wells[0]
    nb, nc = wl.grid.nb, wl.Nc + 1
    dev = IsoTFIM.from_workload(wl)
    dev.CalFlash(BATCHED)
    dev.CalKrPc()
    st = state_from_workload(wl)
    oracle.cal_flash(wl.fluid, st)
    oracle.cal_krpc(wl.sat, st)
    c = np.asarray(prod.cells)
    cs = {"P": st.a["P"][c], "PJ": st.a["PJ"].reshape(nb, 3)[c],
          "PHASEEXIST": st.a["PHASEEXIST"].reshape(nb, 3)[c].astype(bool), "KR": st.a["KR"].reshape(nb, 3)[c],
          "MU": st.a["MU"].reshape(nb, 3)[c], "XI": st.a["XI"].reshape(nb, 3)[c],
          "XIJ": st.a["XIJ"].reshape(nb, 3, nc)[c]}

    def oflash(P, T, Ni):
        w2 = WL.surface_workload(wl.fluid, np.asarray(Ni, float)[None, :], P, T)
        s2 = state_from_workload(w2)
        oracle.cal_flash(w2.fluid, s2)
        return s2.a["PHASEEXIST"].astype(bool), s2.a["S"] * s2.a["VF"][0], s2.a["RHO"]

    lim = WL.WellLimits(init_mode=WL.RATE, max_rate=0.0, limit_bhp=float(cs["P"].min()) - 20.0,
                        tar_bhp=float(cs["P"].min()) - 20.0, weights=(1.0, 0.0, 0.0))
    seg = WL.SegmentFlash(wl.fluid)
    md, bd, qd = WL.check_opt_mode(wl.fluid, WL.cell_state(dev, prod.cells), prod, lim, seg)
    mo, bo, qo = WL.check_opt_mode(wl.fluid, cs, prod, lim, oflash)
    print("producer rate at minBHP:", qd, qo)
    assert qd > 0 and (md, bd) == (mo, bo) == (WL.RATE, prod.bhp)
    assert close(np.array([qd]), np.array([qo]))[0]
    lim.max_rate = 2 * qd
    assert WL.check_opt_mode(wl.fluid, WL.cell_state(dev, prod.cells), prod, lim, seg)[:2] == (WL.BHP, lim.tar_bhp)
    seg.close()
    dev.close()


def test_spe5_time_step_with_its_producer(oracle):
    """One SPE5 time step driven by the deck: the ORAT producer's schedule
    control, CheckOptMode at minBHP (device surface flash), the rate weights
    from the mobility stream, then Newton iterations with the well row on
    the device (assemble -> well terms -> solve -> update).  The same step
    on the oracle with a direct solve converges to bhp 268.6245 bar with the
    rate on its 12000 stb/day target."""
    import scipy.sparse.linalg as spla
    from opencaeporox_amd.fim import BATCHED
    d = deck.read_deck(os.path.join(DECKS, "spe5", "spe5.data"))
    wl = d.to_workload(seed=3)
    (prod, lim), = d.to_wells(step=0)
    nb, b, nc = wl.grid.nb, wl.Nc + 2, wl.Nc + 1
    dev = IsoTFIM.from_workload(wl)
    dev.newton_pass(DT, BATCHED)
    cs = WL.cell_state(dev, prod.cells)
    seg = WL.SegmentFlash(wl.fluid)
    mode, _, q = WL.check_opt_mode(wl.fluid, cs, prod, lim, seg)
    seg.close()
    assert mode == WL.RATE and q > prod.target
    prod.mode = mode
    prod.bhp = float(cs["P"][0]) - 5.0
    stream = WL.factor_stream(np.zeros(nc), cs, prod)
    prod.factor = list(np.full(nc, WL.rate_factors(wl.fluid, stream[None, :], np.array([lim.weights]))[0]))
    # the oracle's step with a direct solve
    ow = WL.PeacemanWell(**{**prod.__dict__})
    st = state_from_workload(wl)
    for _ in range(6):
        oracle.cal_flash(wl.fluid, st); oracle.cal_krpc(wl.sat, st); oracle.cal_rock(wl.rock, st)
        res, _ = oracle.cal_res(wl.fluid, st, wl.grid, DT)
        A = oracle.assemble(wl.fluid, st, wl.grid, DT)
        _, wres, _, Bo, Co, Do = oracle.wells(wl.fluid, st, DT, [ow], res, A)
        J, f = WL.full_system(A[0], A[1], A[2], res, [ow], Bo, Co, Do, wres, b)
        u = spla.spsolve(J.tocsc(), f)
        oracle.get_solution(wl.fluid, st, u[:nb * b], 0.2)
        ow.bhp += u[nb * b]
    # the device step
    dev.SetWells([prod])
    bulk, wrow = [], []
    for k in range(6):
        dev.newton_pass(DT, BATCHED)
        _, w = dev.CalResWells(DT, jacobian=True)
        bulk.append(np.abs(dev.bsr()[3]).max())
        wrow.append(abs(w[0]))
        u, it, rr = dev.LinearSolve(rtol=1e-12, maxit=2000)
        assert it > 0
        dev.GetSolution(None, 0.2)
    bhp, qw = dev.wells_state()
    print("max |R|:", " ".join(f"{v:.2e}" for v in bulk), "| well row:", " ".join(f"{v:.2e}" for v in wrow),
          f"| bhp {bhp[0]:.7f} (oracle {ow.bhp:.7f})")
    assert bulk[-1] < 1e-9 * bulk[0] and wrow[-1] < 1e-8 * wrow[0]
    assert bhp[0] == pytest.approx(ow.bhp, rel=1e-9)
    assert qw[0] @ np.array(prod.factor) == pytest.approx(prod.target, rel=1e-8)
    dev.close()


def test_spe3_time_step_with_its_wells(oracle):
    """One SPE3 time step driven by the deck (gas condensate near its dew
    point, 272 of 324 cells two-phase): the solvent injector and the GRAT
    producer under rate control after CheckOptMode, the producer's weights
    from its mobility stream, the injector's from the stream's surface molar
    volume; device Newton iterations with both well rows against the
    oracle's step with a direct solve.  The intermediate iterates pass
    through rounding-sensitive flash cells, so the converged state is
    compared."""
    import scipy.sparse.linalg as spla
    from opencaeporox_amd.fim import BATCHED
    d = deck.read_deck(os.path.join(DECKS, "spe3", "spe3.data"))
    wl = d.to_workload(seed=3)
    nb, b, nc = wl.grid.nb, wl.Nc + 2, wl.Nc + 1
    dev = IsoTFIM.from_workload(wl)
    dev.newton_pass(DT, BATCHED)
    seg = WL.SegmentFlash(wl.fluid)
    wells = []
    for w, lim in d.to_wells(step=0):
        cs = WL.cell_state(dev, w.cells)
        w.mode, _, q = WL.check_opt_mode(wl.fluid, cs, w, lim, seg)
        assert w.mode == WL.RATE and q > w.target
        if w.type == WL.PROD:
            w.bhp = float(cs["P"].min()) - 5.0
            stream = WL.factor_stream(np.zeros(nc), cs, w)
            w.factor = list(np.full(nc, WL.rate_factors(wl.fluid, stream[None, :], np.array([lim.weights]))[0]))
        else:
            zhc = np.array(w.zi[:nc - 1])
            w.bhp = float(cs["P"].max()) + 5.0
            w.xi = [WL.xi_phase(wl.fluid, w.bhp, d.temperature_c(), zhc)] * len(w.cells)
            w.factor = [1.0 / WL.xi_phase(wl.fluid, WL.PSURF, WL.TSURF, zhc)] * nc
        wells.append(w)
    seg.close()
    ows = [WL.PeacemanWell(**{**w.__dict__}) for w in wells]
    st = state_from_workload(wl)
    for _ in range(8):
        oracle.cal_flash(wl.fluid, st); oracle.cal_krpc(wl.sat, st); oracle.cal_rock(wl.rock, st)
        res, _ = oracle.cal_res(wl.fluid, st, wl.grid, DT)
        A = oracle.assemble(wl.fluid, st, wl.grid, DT)
        _, wres, _, Bo, Co, Do = oracle.wells(wl.fluid, st, DT, ows, res, A)
        J, f = WL.full_system(A[0], A[1], A[2], res, ows, Bo, Co, Do, wres, b)
        u = spla.spsolve(J.tocsc(), f)
        oracle.get_solution(wl.fluid, st, u[:nb * b], 0.2)
        for k, w in enumerate(ows):
            w.bhp += u[(nb + k) * b]
    dev.SetWells(wells)
    bulk = []
    for k in range(8):
        dev.newton_pass(DT, BATCHED)
        _, wres = dev.CalResWells(DT, jacobian=True)
        bulk.append(np.abs(dev.bsr()[3]).max())
        u, it, rr = dev.LinearSolve(rtol=1e-12, maxit=3000)
        assert it > 0
        dev.GetSolution(None, 0.2)
    bhp, qw = dev.wells_state()
    print("max |R|:", " ".join(f"{v:.2e}" for v in bulk), "| bhp", bhp, "oracle", [w.bhp for w in ows])
    assert bulk[-1] < 1e-9 * bulk[0] and np.abs(wres).max() < 1e-6
    assert np.allclose(bhp, [w.bhp for w in ows], rtol=1e-8)
    for k, w in enumerate(wells):
        rate = qw[k] @ np.array(w.factor)
        assert abs(rate) == pytest.approx(w.target, rel=1e-8)
    assert np.allclose(dev.get("P"), st.a["P"], rtol=1e-9)
    dev.close()
