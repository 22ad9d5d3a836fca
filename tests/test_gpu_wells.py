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
    d = deck.read_deck(os.path.join(DECKS, "spe5", "spe5.data"))
    _, pd = d.datum()
    rng = np.random.Generator(np.random.PCG64(3))
    wl = d.to_workload(seed=3, P=pd * rng.uniform(0.3, 1.1, d.nb))
    nc = wl.Nc + 1
    (pc, pwi), = d.well_perforations("PROD1")
    (ic, iwi), = d.well_perforations("I")
    P = wl.state["P"]
    zi = np.zeros(nc)
    zi[:3] = [0.77, 0.20, 0.03]                       # the deck's WELLSTRE solvent
    wells = [WL.PeacemanWell("PROD1", WL.PROD, WL.RATE, bhp=float(P[pc]) - 20.0, target=12000 * 0.159,
                             cells=[pc], WI=[pwi], factor=list(np.full(nc, 0.12))),
             WL.PeacemanWell("I", WL.INJ, WL.BHP, bhp=float(P[ic]) + 30.0, target=float(P[ic]) + 30.0,
                             cells=[ic], WI=[iwi], xi=[6.5], zi=list(zi))]
    dev = IsoTFIM.from_workload(wl)
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
