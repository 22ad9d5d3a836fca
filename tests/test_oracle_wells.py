"""Peaceman wells in the oracle (orc_wells) on the CPU: the assembled
system with its well rows against central finite differences of the
residual, mass conservation between the wells and the bulk rows, the
reference's well-row conventions, and the Peaceman index against its
closed form.  The well terms follow src/WellPeaceman.cpp:365-462,
1226-1612; the reference's WellPeaceman sources need the whole Bulk /
LinearSystem stack and are not built here, so these checks pin them."""
import numpy as np
import pytest

from opencaeporox_amd import abi
from opencaeporox_amd import wells as WL
from opencaeporox_amd import workload as W
from tests.test_oracle_math import _dense, _pass, _small_grid_workload

DT = 1.0


def _wells(wl, nc, mode_prod=WL.BHP, mode_inj=WL.RATE):
    g = wl.grid
    nxny = g.nx * g.ny
    col = lambda i, j: [i + j * g.nx + k * nxny for k in range(g.nz)]
    rng = np.random.Generator(np.random.PCG64(9))
    P = wl.state["P"]
    zi = np.zeros(nc)
    zi[0] = 0.9
    zi[1] = 0.1
    prod = WL.PeacemanWell("P1", WL.PROD, mode_prod, bhp=float(P[col(0, 0)].min()) - 8.0, target=0.0,
                           cells=col(0, 0), WI=list(rng.uniform(5, 20, g.nz)),
                           dG=list(0.07 * np.arange(g.nz)), factor=list(rng.uniform(0.5, 1.5, nc)))
    if mode_prod == WL.BHP:
        prod.target = prod.bhp - 0.5
    else:
        prod.target = 50.0
    inj = WL.PeacemanWell("I1", WL.INJ, mode_inj, bhp=float(P[col(g.nx - 1, g.ny - 1)].max()) + 10.0,
                          target=40.0, cells=col(g.nx - 1, g.ny - 1), WI=list(rng.uniform(5, 20, g.nz)),
                          dG=list(0.05 * np.arange(g.nz)), xi=list(rng.uniform(4, 8, g.nz)), zi=list(zi),
                          factor=list(np.full(nc, 0.2)))
    prod2 = WL.PeacemanWell("P2", WL.PROD, WL.RATE, bhp=float(P[col(g.nx - 1, 0)].min()) - 5.0, target=30.0,
                            cells=col(g.nx - 1, 0)[:1], WI=[12.0], dG=[0.0], factor=list(rng.uniform(0.5, 1.5, nc)))
    return [prod, inj, prod2]


def _system(oracle, wl, wells, dt, P=None, Ni=None, bhp=None):
    """F (the residual R with the well rows) at a state, from the oracle."""
    st, res, _ = _pass(oracle, wl, dt, P, Ni)
    ws = [WL.PeacemanWell(**{**w.__dict__}) for w in wells]
    if bhp is not None:
        for w, v in zip(ws, bhp):
            w.bhp = v
    norms, wres, qw, *_ = oracle.wells(wl.fluid, st, dt, ws, res)
    return st, np.concatenate([-res, np.concatenate([[r] + [0.0] * (wl.Nc + 1) for r in wres])]), qw


@pytest.mark.parametrize("mode_prod", [WL.BHP, WL.RATE])
def test_well_system_matches_finite_differences(oracle, mode_prod):
    Nc, seed = 3, 30
    wl = _small_grid_workload(Nc, seed, nx=3, ny=3, nz=2, Plo=5.0, Phi=120.0)
    nb, nc, b = wl.grid.nb, Nc + 1, Nc + 2
    wells = _wells(wl, nc, mode_prod)
    nw = len(wells)
    st, res, _ = _pass(oracle, wl, DT)
    lab = st.a["PHASEEXIST"].copy()
    assert st.ill.max() == 0
    A = oracle.assemble(wl.fluid, st, wl.grid, DT)
    norms, wres, qw, Bnw, Cwn, Dw = oracle.wells(wl.fluid, st, DT, wells, res, A)
    rp, col, val = A
    J, f = WL.full_system(rp, col, val, res, wells, Bnw, Cwn, Dw, wres, b)
    J = J.toarray()
    assert np.allclose(f[:nb * b], res)
    # the reference adds one identity per perforation to a BHP well's diagonal
    for w, well in enumerate(wells):
        if well.mode == WL.BHP:
            assert np.array_equal(Dw[w], len(well.cells) * np.eye(b))
        else:
            assert np.array_equal(np.diag(Dw[w])[1:], np.full(b - 1, len(well.cells)))
    P0, N0 = st.a["P"].copy(), st.a["NI"].reshape(nb, nc).copy()
    bhp0 = np.array([w.bhp for w in wells])
    N = (nb + nw) * b
    Jfd = np.zeros((N, N))
    checked = np.zeros(N, bool)
    for c in range(nb):
        for k in range(b):
            h = 1e-3 if k == 0 else 1e-5 * N0[c, k - 1]
            out = []
            for s in (1, -1):
                P, Nn = P0.copy(), N0.copy()
                if k == 0:
                    P[c] += s * h
                else:
                    Nn[c, k - 1] += s * h
                sp_, F, _ = _system(oracle, wl, wells, DT, P=P, Ni=Nn.ravel())
                out.append((np.array_equal(sp_.a["PHASEEXIST"], lab) and sp_.ill.max() == 0, F))
            if out[0][0] and out[1][0]:
                Jfd[:, c * b + k] = (out[0][1] - out[1][1]) / (2 * h)
                checked[c * b + k] = True
    for w in range(nw):
        h = 1e-3
        Fp = _system(oracle, wl, wells, DT, bhp=bhp0 + h * (np.arange(nw) == w))[1]
        Fm = _system(oracle, wl, wells, DT, bhp=bhp0 - h * (np.arange(nw) == w))[1]
        Jfd[:, (nb + w) * b] = (Fp - Fm) / (2 * h)
        checked[(nb + w) * b] = True
    assert checked[:nb * b].mean() > 0.9
    # rows: every bulk row and each well's equation row; the other well rows
    # are the reference's placeholder equations (identity per perforation)
    rows = np.r_[np.arange(nb * b), (nb + np.arange(nw)) * b]
    skip = [(nb + w) * b for w, well in enumerate(wells) if well.mode == WL.BHP]
    Jc, Fc = J[np.ix_(rows, checked)], Jfd[np.ix_(rows, checked)]
    for r in skip:                       # d(bhp - target)/d bhp = 1, the reference assembles nperf
        i = np.where(rows == r)[0][0]
        j = np.where(np.where(checked)[0] == r)[0][0]
        assert Fc[i, j] == pytest.approx(1.0) and Jc[i, j] == len(wells[(r // b) - nb].cells)
        Jc[i, j] = Fc[i, j]
    scale = np.max(np.abs(Fc), axis=1, keepdims=True) + 1e-300
    err = np.abs(Jc - Fc) / scale
    assert np.max(err) < 1e-6, f"worst {np.max(err):.2e} at {np.unravel_index(np.argmax(err), err.shape)}"
    # the wells couple every perforated cell to their bhp column
    for w, well in enumerate(wells):
        for n in well.cells:
            assert np.abs(J[n * b + 1:(n + 1) * b, (nb + w) * b]).max() > 0


def test_wells_conserve_moles(oracle):
    Nc, seed = 3, 30
    wl = _small_grid_workload(Nc, seed, nx=3, ny=3, nz=2, Plo=5.0, Phi=120.0)
    nb, nc, b = wl.grid.nb, Nc + 1, Nc + 2
    wells = _wells(wl, nc)
    st, res0, _ = _pass(oracle, wl, 2.0)
    res = res0.copy()
    _, wres, qw, *_ = oracle.wells(wl.fluid, st, 2.0, wells, res)
    d = (-res + res0).reshape(nb, b)                 # what the wells added to R
    assert np.all(d[:, 0] == 0)
    assert np.allclose(d[:, 1:].sum(axis=0), 2.0 * qw.sum(axis=0), rtol=1e-12, atol=1e-14)
    # producers take moles out (q > 0 adds to R), the injector puts its composition in
    assert np.all(qw[0] > 0) and np.all(qw[1] <= 0)
    zi = np.array(wells[1].zi)
    assert np.allclose(qw[1] / qw[1].sum(), zi)
    # the BHP well row is bhp - target; the rate rows +-target + sum q_i factor_i
    assert wres[0] == wells[0].bhp - wells[0].target
    assert wres[1] == pytest.approx(wells[1].target + qw[1] @ np.array(wells[1].factor))
    assert wres[2] == pytest.approx(-wells[2].target + qw[2] @ np.array(wells[2].factor))


def test_peaceman_index_closed_form(oracle):
    """Isotropic square cell: r_o = 0.14 sqrt(dx^2 + dy^2) (0.198 dx, Peaceman
    1978), WI = 2 pi k dz / (ln(r_o / r_w) + s), times CONV_DARCY."""
    dx, dz, k, rw, s = 100.0, 5.0, 250.0, 0.1, 1.5
    ro = 0.14 * np.sqrt(2 * dx * dx)
    expect = 2 * np.pi * k * dz / (np.log(ro / rw) + s) * W.DARCY
    got = oracle.peaceman_wi(2, dx, dx, dz, k, k, 0.1 * k, rw, s, -1.0)
    assert got == pytest.approx(expect, rel=1e-9)       # the reference's pi has 10 digits
    # x-direction perforation, anisotropic: swapping (dy, ky) with (dz, kz) changes nothing
    a = oracle.peaceman_wi(0, 50.0, 80.0, 10.0, 100.0, 300.0, 30.0, 0.1, 0.0, -1.0)
    b2 = oracle.peaceman_wi(0, 50.0, 10.0, 80.0, 100.0, 30.0, 300.0, 0.1, 0.0, -1.0)
    assert a == pytest.approx(b2, rel=1e-14)
    # the C ABI's host helper is the same function
    lib = abi.load()
    for d, args in ((2, (dx, dx, dz, k, k, 0.1 * k, rw, s, -1.0)), (0, (50.0, 80.0, 10.0, 100.0, 300.0, 30.0, 0.1, 0.0, -1.0)),
                    (1, (30.0, 60.0, 4.0, 10.0, 20.0, 5.0, 0.08, -0.5, 7.0))):
        assert lib.ocpx_peaceman_wi(d, *args) == oracle.peaceman_wi(d, *args)
    assert lib.ocpx_peaceman_wi(5, *args) == -1.0


def test_xi_phase_is_the_single_phase_molar_density(oracle):
    """ocpx_xi_phase (the injected fluid's perf_xi) against the oracle's molar
    volume (with a volume shift) and its water PVTW."""
    fl = W.make_fluid(6, 101, skip_psa=True)
    for i in range(6):
        fl.Vshift[i] = 0.02 * (i - 2)
    rng = np.random.Generator(np.random.PCG64(12))
    for P, T in ((50.0, 60.0), (150.0, 71.0), (400.0, 120.0)):
        z = rng.dirichlet(np.ones(6))
        vm = oracle.vmder(fl, P, T + W.CONV4, z)[0]
        assert WL.xi_phase(fl, P, T, z) == pytest.approx(1 / vm, rel=1e-13)
        assert WL.xi_phase(fl, P, T, water=True) == pytest.approx(oracle.pvtw(fl, P)[1], rel=1e-14)
    with pytest.raises(ValueError):
        WL.xi_phase(fl, -1.0, 60.0, z)


def test_rate_factors_reject_bad_input():
    fl = W.make_fluid(3, 101)
    with pytest.raises(ValueError, match="non-negative"):
        WL.rate_factors(fl, np.array([[1.0, -1.0, 0.5, 0.0]]), np.array([[1.0, 0.0, 0.0]]))
    with pytest.raises(ValueError, match="nwell"):
        WL.rate_factors(fl, np.array([[1.0, 1.0, 0.5]]), np.array([[1.0, 0.0, 0.0]]))


def test_injector_heads_integrate_the_fluid_column():
    """CalInjdG: incompressible water gives dG = rho g (depth - head depth)
    exactly; a gas injector's heads approach the solution of
    dP/dz = rho(P) g as the segments shrink."""
    from scipy.integrate import solve_ivp
    fl = W.make_fluid(4, 101)
    fl.PVTW[2] = 0.0                                  # incompressible water
    depths = np.array([2010.0, 2030.0, 2055.0])
    w = WL.PeacemanWell("I", WL.INJ, WL.BHP, bhp=200.0, target=200.0, cells=[0, 1, 2], WI=[1.0] * 3,
                        dG=[0.0] * 3, zi=[0.0] * 5)
    dG = WL.injector_heads(fl, w, 2000.0, depths, T=70.0, water=True)
    rho = fl.rhoW_std / fl.PVTW[1]
    assert np.allclose(dG, rho * 0.0000980665 * (depths - 2000.0), rtol=1e-12)
    zi = [0.9, 0.06, 0.03, 0.01, 0.0]
    w = WL.PeacemanWell("I", WL.INJ, WL.BHP, bhp=200.0, target=200.0, cells=[0, 1, 2], WI=[1.0] * 3,
                        dG=[0.0] * 3, zi=zi)
    mw = sum(zi[i] * fl.MW[i] for i in range(4))
    ode = lambda z, P: [mw * WL.xi_phase(fl, P[0], 70.0, zi[:4]) * 0.0000980665]     # dP/dz = rho g
    fine = WL.injector_heads(fl, w, 2000.0, depths, T=70.0, maxlen=0.05)
    for p, d in enumerate(depths):
        # each perforation integrates upwards from its own pressure (bhp + old dG = 200 bar)
        top = solve_ivp(ode, (d, 2000.0 if p == 0 else depths[p - 1]), [200.0], rtol=1e-11, atol=1e-9).y[0, -1]
        assert 200.0 - top == pytest.approx(np.diff(np.r_[0.0, fine])[p], rel=2e-5)


def test_injector_mode_check(oracle):
    """CheckOptMode for an injector: the rate at maxBHP is sum_p transINJ *
    xi(Pperf) * (Pperf - P) times the injected fluid's surface molar volume
    (here from the oracle's EoS), and it decides between rate and BHP control."""
    fl = W.make_fluid(4, 101)
    zi = [0.9, 0.06, 0.03, 0.01, 0.0]
    cell = {"P": np.array([150.0, 152.0]), "PHASEEXIST": np.array([[1, 0, 1], [1, 1, 1]], bool),
            "KR": np.array([[0.6, 0.0, 0.1], [0.3, 0.2, 0.05]]), "MU": np.array([[0.8, 0.0, 0.5], [0.9, 0.02, 0.5]])}
    w = WL.PeacemanWell("I", WL.INJ, WL.RATE, bhp=170.0, target=50.0, cells=[0, 1], WI=[3.0, 4.0], dG=[0.0, 0.7],
                        zi=zi)
    lim = WL.WellLimits(init_mode=WL.RATE, max_rate=0.0, limit_bhp=180.0, tar_bhp=180.0, inj_T=60.0)
    mode, bhp, q = WL.check_opt_mode(fl, cell, w, lim, flash=None)
    expect = 0.0
    for p in range(2):
        Pp = 180.0 + w.dG[p]
        tinj = sum(w.WI[p] * cell["KR"][p, j] / cell["MU"][p, j] for j in range(3) if cell["PHASEEXIST"][p, j])
        expect += tinj / oracle.vmder(fl, Pp, 60.0 + W.CONV4, zi[:4])[0] * (Pp - cell["P"][p])
    expect *= oracle.vmder(fl, WL.PSURF, WL.TSURF + W.CONV4, zi[:4])[0]
    assert q == pytest.approx(expect, rel=1e-12)
    assert (mode, bhp) == (WL.RATE, 170.0)                     # the potential exceeds the limit
    lim.max_rate = 1.01 * q
    assert WL.check_opt_mode(fl, cell, w, lim, flash=None)[:2] == (WL.BHP, 180.0)
    lim.init_mode = WL.BHP
    lim.max_rate = 0.99 * q
    assert WL.check_opt_mode(fl, cell, w, lim, flash=None)[:2] == (WL.RATE, 170.0)


def test_cross_flow_check():
    """CheckCrossFlow: the first reversed open perforation closes (and stops
    the scan), closed ones pointing the right way reopen, and the deepest
    opens when none is left."""
    prod = WL.PeacemanWell("P", WL.PROD, WL.BHP, bhp=100.0, target=100.0, cells=[0, 1, 2, 3], WI=[1.0] * 4,
                           dG=[0.0, 1.0, 2.0, 3.0])
    # perforation pressures 100, 101, 102, 103
    op, cross = WL.check_cross_flow(np.array([105.0, 100.5, 101.0, 110.0]), prod, [True, True, True, True])
    assert cross and op == [True, False, True, True]        # perforation 1 reversed; the scan stops there
    op, cross = WL.check_cross_flow(np.array([105.0, 102.0, 101.0, 110.0]), prod, [True, False, True, True])
    assert cross and op == [True, True, False, True]        # 1 reopens, then 2 (reversed) closes
    op, cross = WL.check_cross_flow(np.array([105.0, 102.0, 103.0, 110.0]), prod, [True, True, False, True])
    assert not cross and op == [True, True, True, True]
    inj = WL.PeacemanWell("I", WL.INJ, WL.BHP, bhp=100.0, target=100.0, cells=[0, 1], WI=[1.0] * 2, dG=[0.0, 1.0])
    op, cross = WL.check_cross_flow(np.array([120.0, 130.0]), inj, [False, False])
    assert not cross and op == [False, True]                # none open: the deepest opens
    op, cross = WL.check_cross_flow(np.array([99.0, 130.0]), inj, [True, True])
    assert cross and op == [True, False]


def test_factor_stream_fallback():
    """CalFactor flashes the well's rates, or the perforation mobility stream
    when a rate is negative or the total is not positive."""
    nc = 3
    cell = {"PHASEEXIST": np.array([[1, 1, 0]], bool), "KR": np.array([[0.5, 0.2, 0.0]]),
            "MU": np.array([[1.0, 0.05, 0.0]]), "XI": np.array([[8.0, 2.0, 55.0]]),
            "XIJ": np.array([[[0.2, 0.8, 0.0], [0.9, 0.1, 0.0], [0.0, 0.0, 1.0]]])}
    w = WL.PeacemanWell("P", WL.PROD, WL.RATE, bhp=100.0, target=10.0, cells=[0], WI=[2.0])
    q = np.array([1.0, 2.0, 0.5])
    assert np.array_equal(WL.factor_stream(q, cell, w), q)
    fb = 2.0 * 0.5 / 1.0 * 8.0 * np.array([0.2, 0.8, 0.0]) + 2.0 * 0.2 / 0.05 * 2.0 * np.array([0.9, 0.1, 0.0])
    assert np.allclose(WL.factor_stream(np.array([1.0, -0.1, 0.5]), cell, w), fb, rtol=1e-15)
    assert np.allclose(WL.factor_stream(np.zeros(nc), cell, w), fb, rtol=1e-15)
