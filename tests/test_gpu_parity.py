"""HIP path vs the C oracle on the same inputs (FP64, 1e-10 relative,
identical phase labels).  All compute goes through libocpx.so."""
import numpy as np
import pytest

from opencaeporox_amd import workload as W
from opencaeporox_amd.fim import BATCHED, PLAIN, IsoTFIM
from oracle.oracle import state_from_workload
from tests.conftest import (PARITY_RTOL, SPREAD_MULT, close, perturbed_oracle_states, rel_err,
                            rounding_sensitive, rounding_spread)

pytestmark = pytest.mark.gpu

FLASH_FIELDS = ["S", "XIJ", "RHO", "XI", "MU", "RHOP", "XIP", "MUP", "RHOX", "XIX", "MUX",
                "NT", "VF", "VFP", "VFI"]
FLASH_OUT = FLASH_FIELDS + ["PHASEEXIST", "DSEC_DPRI"]


def _bits(v):
    return "|".join(str(1 << b) for b in range(10) if (v >> b) & 1) or "0"


def _compare_flash(dev, st, Nc, tag, others, max_mismatch=0.02):
    """Labels identical on every cell whose reference trajectory is not
    rounding-sensitive (no split / Rachford-Rice ill bit in the oracle); all
    fields within PARITY_RTOL on those cells, or within SPREAD_MULT times the
    oracle's own spread under rounding-level perturbations (others) where
    that is larger.  Elsewhere the reference's own outcome depends on
    rounding: label mismatches there are counted, must stay rare, and are
    listed with their oracle bits."""
    n = dev.nb
    pe_d = dev.get("PHASEEXIST").reshape(n, 3)
    pe_o = st.a["PHASEEXIST"].reshape(n, 3)
    mism = np.any(pe_d != pe_o, axis=1)
    sens = rounding_sensitive(st.ill)
    bad = np.where(mism & ~sens)[0]
    print(f"{tag}: {mism.sum()} label mismatches of {n} ({sens.sum()} rounding-sensitive cells); "
          f"oracle bits of mismatches: {[_bits(int(st.ill[c])) for c in np.where(mism)[0][:20]]}")
    assert bad.size == 0, (f"{tag}: labels differ in {bad.size} cells that are not rounding-sensitive, "
                           f"e.g. {bad[:10]} with bits {[_bits(int(st.ill[c])) for c in bad[:10]]}")
    assert mism.mean() <= max_mismatch, f"{tag}: {mism.sum()} label mismatches"
    keep = ~mism & ~sens
    fo_d, fo_o = dev.get("FTYPE_OUT"), st.a["FTYPE_OUT"]
    assert np.array_equal(fo_d[keep], fo_o[keep]), f"{tag}: ftype differs"
    loose = np.zeros(n, bool)
    for f in FLASH_FIELDS + ["DSEC_DPRI"]:
        rowwise = f == "DSEC_DPRI"
        b = st.a[f].reshape(n, -1)
        a = dev.get(f).reshape(n, -1)
        err = np.zeros_like(b)
        err[keep] = rel_err(a[keep], b[keep], rowwise)
        spread = rounding_spread(st, others, f, n, rowwise)
        tol = np.maximum(PARITY_RTOL, SPREAD_MULT * spread)
        over = np.any(err > PARITY_RTOL, axis=1)
        loose |= over
        fail = np.any(err > tol, axis=1)
        if over.any():
            c = np.where(over)[0]
            print(f"{tag}: {f} above {PARITY_RTOL:g} in {c.size} cells (worst {err.max():.2e}; oracle's own "
                  f"rounding spread there up to {spread[c].max():.2e}), e.g. {c[:6]}")
        assert not fail.any(), (f"{tag}: {f} differs beyond tolerance in cells {np.where(fail)[0][:8]}: "
                                f"err {err[fail].max():.2e}, spread {spread[fail].max():.2e}")
    print(f"{tag}: {keep.sum() - loose[keep].sum()} of {keep.sum()} compared cells within {PARITY_RTOL:g}, "
          f"{loose[keep].sum()} within {SPREAD_MULT:g}x the oracle's rounding spread")
    return ~mism


@pytest.mark.parametrize("mode,lanes", [(PLAIN, 1), (BATCHED, 1), (BATCHED, 4), (BATCHED, 8)])
def test_flash_config1(oracle, oracle_fma, mode, lanes):
    wl = W.cube_grid()
    dev = IsoTFIM.from_workload(wl)
    dev.set_flash_mapping(lanes)
    st = state_from_workload(wl)
    others = perturbed_oracle_states(oracle, oracle_fma, wl.fluid, st)
    dev.CalFlash(mode)
    oracle.cal_flash(wl.fluid, st)
    same = _compare_flash(dev, st, wl.Nc, f"10x10x10 mode {mode} lanes {lanes}", others)
    # SkipPSA next-step data
    keep = same & ~rounding_sensitive(st.ill)
    assert np.array_equal(dev.get("SKIP_FLAG")[keep], st.a["SKIP_FLAG"][keep])
    ok, worst = close(dev.get("SKIP_MINEIG")[keep], st.a["SKIP_MINEIG"][keep], rtol=1e-6)
    assert ok, f"minEig worst {worst}"


# team sizes built per Nc: 4 always, 8 from Nc = 5, 16 from Nc = 9
PE_CASES = [(Nc, m, lanes) for Nc in (3, 4, 5, 9, 12)
            for m, lanes in ((PLAIN, 1), (BATCHED, 1), (BATCHED, 4), (BATCHED, 8), (BATCHED, 16))
            if lanes < 8 or (lanes == 8 and Nc >= 5) or (lanes == 16 and Nc >= 9)]


@pytest.mark.parametrize("Nc,mode,lanes", PE_CASES)
def test_phase_equilibrium_cells(oracle, oracle_fma, Nc, mode, lanes):
    """Phase count, split and flash type per cell against the oracle, for
    the one-cell-per-lane kernels and the team mapping (one component per
    lane)."""
    # pressures across both phase regions; seed 44 gives the 3-component
    # fluid a two-phase region there
    wl, z = W.flash_cells(3000, Nc, seed=44 if Nc == 3 else 10 + Nc, P=np.linspace(5.0, 250.0, 3000))
    dev = IsoTFIM.from_workload(wl)
    dev.set_flash_mapping(lanes)
    dev.PhaseEquilibrium(mode)
    NP, nu, x = dev.pe_result()
    P = wl.state["P"]
    T = wl.state["T"] + W.CONV4
    NPo, nuo, xo, fo, ito = oracle.phase_equilibrium(wl.fluid, P, T, z)
    ill = oracle.last_ill
    sens = rounding_sensitive(ill)
    mism = NP != NPo
    print(f"Nc={Nc}: {mism.sum()} NP mismatches of {len(P)}, rounding-sensitive cells {sens.sum()}")
    bad = np.where(mism & ~sens)[0]
    assert bad.size == 0, f"Nc={Nc}: NP differs in {bad.size} cells {bad[:10]}"
    assert mism.mean() < 0.02
    two = (NP == 2) & (NPo == 2) & ~sens
    assert two.sum() > 100, "workload must contain well-conditioned two-phase cells"
    runs = [oracle_fma.phase_equilibrium(wl.fluid, P, T, z),
            oracle.phase_equilibrium(wl.fluid, P * (1 + 2.0 ** -50), T, z)]
    for name, dv, ov, k in (("nu", nu, nuo, 1), ("x", x, xo, 2)):
        dv, ov = dv[two].reshape(two.sum(), -1), ov[two].reshape(two.sum(), -1)
        err = rel_err(dv, ov)
        spread = np.zeros_like(err)
        for r in runs:
            same = (r[0] == NPo)[two]
            e = rel_err(r[k][two].reshape(two.sum(), -1), ov)
            e[~same] = 0
            spread = np.maximum(spread, e)
        tol = np.maximum(PARITY_RTOL, SPREAD_MULT * spread)
        over = np.any(err > PARITY_RTOL, axis=1)
        print(f"Nc={Nc}: {name} above {PARITY_RTOL:g} in {over.sum()} of {two.sum()} two-phase cells "
              f"(worst {err.max():.2e}, oracle spread there {spread[over].max() if over.any() else 0:.2e})")
        assert np.all(err <= tol), f"{name}: err {err.max():.2e}"
    assert np.array_equal(dev.get("FTYPE_OUT")[~sens], fo[~sens])
    # iteration counts agree on clean cells (same algorithm; a converged
    # fixed-point iteration may take one more or one fewer step under
    # different rounding)
    clean = ill == 0
    it = dev.get("ITERS").reshape(-1, 8)
    same = np.mean(np.all(it[clean, :5] == ito[clean, :5], axis=1))
    print(f"iteration counts identical in {100 * same:.1f}% of clean cells")
    assert same > 0.9


def test_batched_equals_plain():
    """The work queue changes the order cells are processed in, nothing else."""
    wl, z = W.flash_cells(5000, 9, seed=77)
    a = IsoTFIM.from_workload(wl)
    b = IsoTFIM.from_workload(wl)
    b.set_flash_mapping(1)
    a.CalFlash(PLAIN)
    b.CalFlash(BATCHED)
    for f in FLASH_OUT:
        assert np.array_equal(a.get(f), b.get(f), equal_nan=True), f
    ia, ib = a.get("ITERS").reshape(-1, 8), b.get("ITERS").reshape(-1, 8)
    diff = np.any(ia != ib, axis=1)
    if diff.any():
        slots = np.where(np.any(ia != ib, axis=0))[0]
        print(f"ITERS differ in {diff.sum()} cells, slots {slots}, e.g.")
        for c in np.where(diff)[0][:5]:
            print("  ", c, ia[c], ib[c])
    assert diff.mean() < 0.01


def _upload_flash(dev, st):
    for f in FLASH_OUT:
        dev.set(f, st.a[f])


def _compare_assembly(dev, oracle, wl, st, dt, tag):
    dev.CalKrPc()
    dev.CalRock()
    res_d, norms_d = dev.CalRes(dt)
    dev.AssembleMatBulks(dt)
    oracle.cal_krpc(wl.sat, st)
    oracle.cal_rock(wl.rock, st)
    res_o, norms_o = oracle.cal_res(wl.fluid, st, wl.grid, dt)
    rp, col, val = oracle.assemble(wl.fluid, st, wl.grid, dt)
    for f in ["KR", "PC", "PJ", "DKRDS", "DPCDS", "PORO", "POROP", "ROCKVP"]:
        ok, worst = close(dev.get(f), st.a[f])
        assert ok, f"{tag}: {f} worst {worst}"
    ok, worst = close(res_d, res_o)
    assert ok, f"{tag}: resAbs worst {worst}"
    assert norms_d[1] == norms_o[1] and norms_d[3] == norms_o[3]
    ok, worst = close(norms_d[[0, 2]], norms_o[[0, 2]])
    assert ok
    rp_d, col_d, val_d, rhs_d = dev.bsr()
    assert np.array_equal(rp_d, rp) and np.array_equal(col_d, col)
    scale = 1e-2 * np.max(np.abs(val), axis=(1, 2), keepdims=True)
    err = np.abs(val_d - val) / (np.abs(val) + scale + 1e-300)
    assert np.max(err) <= 1e-10, (f"{tag}: BSR worst {np.max(err):.3e} at block "
                                  f"{np.unravel_index(np.argmax(err), err.shape)}")
    ok, worst = close(rhs_d, res_o)
    assert ok


def test_assembly_config1_same_flash(oracle):
    """Residual and Jacobian from identical flash outputs: the oracle's flash
    state is uploaded, then kr/Pc, rock, residual and assembly run on the
    device."""
    wl = W.cube_grid()
    st = state_from_workload(wl)
    oracle.cal_flash(wl.fluid, st)
    dev = IsoTFIM.from_workload(wl)
    _upload_flash(dev, st)
    _compare_assembly(dev, oracle, wl, st, 1.0, "10x10x10")


def _clean_grid(seed=31, Nc=3, nx=4, ny=3, nz=2):
    rng = np.random.Generator(np.random.PCG64(seed))
    nb = nx * ny * nz
    k = np.exp(rng.normal(np.log(100.0), 1.0, nb))
    grid = W.cartesian_grid(nx, ny, nz, 6.096, k, k, 0.1 * k, rng.uniform(0.1, 0.3, nb))
    fl = W.make_fluid(Nc, 130)
    rock = W.make_rock()
    stt = W.initial_state(fl, rock, grid, rng.uniform(5.0, 120.0, nb),
                          rng.dirichlet(np.ones(Nc), nb), rng)
    return W.Workload("clean grid", fl, W.make_sat(), rock, grid, stt)


@pytest.mark.parametrize("mode,lanes", [(PLAIN, 1), (BATCHED, 1), (BATCHED, 4)])
def test_newton_passes_end_to_end(oracle, oracle_fma, mode, lanes):
    """Device flash + assembly against the oracle's on a grid whose cells are
    all well-conditioned, then GetSolution and a second pass from the
    two-phase history."""
    wl = _clean_grid()
    dt = 1.0
    dev = IsoTFIM.from_workload(wl)
    dev.set_flash_mapping(lanes)
    st = state_from_workload(wl)
    others = perturbed_oracle_states(oracle, oracle_fma, wl.fluid, st)
    dev.CalFlash(mode)
    oracle.cal_flash(wl.fluid, st)
    assert st.ill.max() == 0
    _compare_flash(dev, st, wl.Nc, "pass1", others, max_mismatch=0)
    _compare_assembly(dev, oracle, wl, st, dt, "pass1")
    rng = np.random.Generator(np.random.PCG64(5))
    u = rng.normal(0, 1, dev.nb * dev.bs)
    u.reshape(-1, dev.bs)[:, 0] *= 2.0                        # bar
    u.reshape(-1, dev.bs)[:, 1:] *= 1e-3 * np.abs(st.a["NI"].reshape(-1, dev.nc))
    dev.GetSolution(u, 0.1)
    oracle.get_solution(wl.fluid, st, u, 0.1)
    for f in ["P", "NI", "S", "XIJ"]:
        ok, worst = close(dev.get(f), st.a[f])
        assert ok, f"GetSolution {f} worst {worst}"
    others = perturbed_oracle_states(oracle, oracle_fma, wl.fluid, st)
    dev.CalFlash(mode)
    oracle.cal_flash(wl.fluid, st)
    _compare_flash(dev, st, wl.Nc, "pass2", others, max_mismatch=0)
    _compare_assembly(dev, oracle, wl, st, dt, "pass2")


def test_reset_to_last_step_repeats_the_pass():
    wl = W.cube_grid()
    dev = IsoTFIM.from_workload(wl)
    dev.UpdateLastTimeStep()
    dev.newton_pass(1.0, BATCHED)
    ref = [dev.get(f) for f in ("PHASEEXIST", "S", "DSEC_DPRI", "ITERS")] + [dev.bsr()[2]]
    for _ in range(2):
        dev.ResetToLastTimeStep()
        assert np.array_equal(dev.get("NI"), dev.get("LNI"))
        dev.newton_pass(1.0, BATCHED)
        got = [dev.get(f) for f in ("PHASEEXIST", "S", "DSEC_DPRI", "ITERS")] + [dev.bsr()[2]]
        for x, y in zip(got, ref):
            assert np.array_equal(x, y, equal_nan=True)


def test_jt_lambda_dot_product():
    wl = W.cube_grid()
    dev = IsoTFIM.from_workload(wl)
    dev.newton_pass(1.0, BATCHED)
    rp, col, val, _ = dev.bsr()
    rng = np.random.Generator(np.random.PCG64(9))
    n, b = dev.nb, dev.bs
    v = rng.normal(size=n * b)
    lam = rng.normal(size=n * b)
    Jv = np.zeros(n * b)
    for r in range(n):
        for s in range(rp[r], rp[r + 1]):
            Jv[r * b:(r + 1) * b] += val[s] @ v[col[s] * b:(col[s] + 1) * b]
    JTl = dev.jt_lambda(lam)
    assert abs(Jv @ lam - v @ JTl) <= 1e-10 * (abs(Jv @ lam) + np.abs(Jv).dot(np.abs(lam)))


def test_autotuner_picks_a_valid_mapping():
    wl, _ = W.flash_cells(20000, 6, seed=3)
    dev = IsoTFIM.from_workload(wl)
    ms = dev.autotune_flash()
    nat = ms.pop("natural_order")
    assert set(ms) == {1, 4, 8} and all(v > 0 for v in ms.values()) and nat > 0
    assert dev.flash_mapping() == min(ms, key=ms.get)
    assert dev.flash_queue_order() == (ms[dev.flash_mapping()] <= nat)
    dev.set_flash_mapping(0)
    assert dev.flash_mapping() == min(ms, key=ms.get)


@pytest.mark.parametrize("lanes", [1, 8])
def test_queue_order_does_not_change_results(lanes):
    """The longest-first queue only reorders work: a second batched flash
    (ordered by the first one's counts) gives the same bits as a flash in
    natural order."""
    wl, _ = W.flash_cells(20000, 6, seed=11, P=np.linspace(5.0, 250.0, 20000))
    a = IsoTFIM.from_workload(wl)
    b = IsoTFIM.from_workload(wl)
    for d in (a, b):
        d.set_flash_mapping(lanes)
    a.set_flash_queue_order(True)
    b.set_flash_queue_order(False)
    for _ in range(2):
        a.set_state(wl.state)
        b.set_state(wl.state)
        a.CalFlash(BATCHED)
        b.CalFlash(BATCHED)
    for f in ("PHASEEXIST", "S", "XIJ", "DSEC_DPRI", "ITERS"):
        assert np.array_equal(a.get(f), b.get(f)), f
    a.close()
    b.close()


def test_graph_replay_matches_eager():
    wl = W.cube_grid()
    a = IsoTFIM.from_workload(wl)
    b = IsoTFIM.from_workload(wl)
    a.newton_pass(1.0, BATCHED, graph=False)
    ref = [a.get(f) for f in ("S", "DSEC_DPRI")] + [a.bsr()[2]]
    for rep in range(3):
        b.set_state(wl.state)
        b.set("SKIP_FLAG", np.zeros(b.nb, np.uint32))
        b.newton_pass(1.0, BATCHED, graph=True)
        got = [b.get(f) for f in ("S", "DSEC_DPRI")] + [b.bsr()[2]]
        for x, y in zip(got, ref):
            assert np.array_equal(x, y), f"replay {rep} differs"
