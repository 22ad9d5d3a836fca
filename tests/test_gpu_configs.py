"""The heterogeneous-grid, channelized-grid and uneven multi-model workloads on
the device against the oracle.

* Heterogeneous grid (ln k ~ N(ln 100, 2), kz = 0.1 kx, hydrostatic P + noise,
  five vertical Peaceman wells): at 20x20x10, flash, residual, BSR and well
  blocks against the oracle at b = 8 (Nc = 6) and b = 7 (Nc = 5); at the full
  100x100x50, sampled dR/dx blocks against central differences of the device
  residual.
* Channelized grid (Nc = 4): J^T lambda against the oracle's BSR transposed,
  d(lambda.R)/d(trans) against the oracle's per-connection fluxes, d/dVp
  against the oracle's accumulation; at the full 60x220x85, 200 sampled
  adjoint gradients against central differences of the device residual.
* Uneven multi-model batch (Nc from 3 to 12): each model's flash, residual and
  BSR against the oracle's run of that model alone.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from opencaeporox_amd import wells as WL
from opencaeporox_amd import workload as W
from opencaeporox_amd.fim import BATCHED, IsoTFIM, MultiModelBatch
from oracle.oracle import OracleState, state_from_workload
from tests.conftest import close, perturbed_oracle_states
from tests.test_gpu_parity import FLASH_OUT, _compare_assembly, _compare_flash, _upload_flash
from tests.test_gpu_wells import _check

pytestmark = pytest.mark.gpu

DT = 1.0


# ------------------------------------------------------------- heterogeneous grid
@pytest.mark.parametrize("Nc", [6, 5])
def test_heterogeneous_grid_with_five_wells(oracle, oracle_fma, Nc):
    wl = W.heterogeneous_grid(20, 20, 10, Nc=Nc)
    tag = f"heterogeneous 20x20x10 Nc={Nc} (b={Nc + 2})"
    print(tag, flush=True)
    dev = IsoTFIM.from_workload(wl)
    st = state_from_workload(wl)
    others = perturbed_oracle_states(oracle, oracle_fma, wl.fluid, st)
    dev.CalFlash(BATCHED)
    oracle.cal_flash(wl.fluid, st)
    _compare_flash(dev, st, Nc, tag, others, min_cells=int(0.98 * dev.nb), min_two_phase=20)
    # residual, BSR and the five wells' terms from identical flash outputs
    _check(dev, oracle, wl, WL.corner_and_centre_wells(wl), tag)
    dev.close()


def _lattice(nx, ny, nz, step=3):
    """Cells on a lattice of spacing `step`: any two differ by >= step in some
    coordinate, so no cell's 7-point stencil holds two of them."""
    i, j, k = np.meshgrid(np.arange(0, nx, step), np.arange(0, ny, step), np.arange(0, nz, step), indexing="ij")
    return np.sort((i + j * nx + k * nx * ny).ravel())


def _stencil(c, nx, ny, nz):
    """Rows whose blocks hold column c: c and its in-grid face neighbours."""
    i, j, k = c % nx, (c // nx) % ny, c // (nx * ny)
    out = [(c, np.ones_like(c, bool))]
    for off, ok in ((-1, i > 0), (1, i < nx - 1), (-nx, j > 0), (nx, j < ny - 1),
                    (-nx * ny, k > 0), (nx * ny, k < nz - 1)):
        out.append((c + off, ok))
    return out


def _slot(rp, col, r, c):
    """BSR slot of block (r, c) (rows hold at most 7 blocks), -1 if absent."""
    s = np.full(r.shape, -1, np.int64)
    for t in range(7):
        idx = rp[r] + t
        inrow = idx < rp[r + 1]
        hit = inrow & (col[np.where(inrow, idx, 0)] == c)
        s[hit & (s < 0)] = idx[hit & (s < 0)]
    return s


def test_heterogeneous_full_grid_blocks_match_central_differences(oracle):
    """100x100x50, Nc = 6 (b = 8), with the exact LBC derivative: the device
    Jacobian's blocks in the columns of 19,652 lattice cells (every column of
    each, all rows of its stencil) against central differences of the device
    residual, the flash recomputed; cells whose flash trajectory the oracle
    marks capped, or whose labels move under the step, are not checked."""
    nx, ny, nz, Nc = 100, 100, 50, 6
    wl = W.heterogeneous_grid(nx, ny, nz, Nc=Nc, lbc_exact=1)
    dev = IsoTFIM.from_workload(wl)
    dev.UpdateLastTimeStep()                 # the previous step every evaluation starts from
    dev.set("LNI", wl.state["LNI"])
    nb, nc, b = dev.nb, dev.nc, dev.bs
    P0 = wl.state["P"].copy()
    N0 = wl.state["NI"].reshape(nb, nc).copy()

    def run(P, N, assemble=False):
        dev.ResetToLastTimeStep()
        dev.set("P", P)
        dev.set("NI", N.ravel())
        dev.CalFlash(BATCHED)
        dev.CalKrPc()
        dev.CalRock()
        res, _ = dev.CalRes(DT)
        if assemble:
            dev.AssembleMatBulks(DT)
        return -res.reshape(nb, b), dev.get("PHASEEXIST").reshape(nb, 3).copy()

    _, lab0 = run(P0, N0, assemble=True)
    rp, col, val, _ = dev.bsr()
    c = _lattice(nx, ny, nz)
    # the oracle's conditioning marks for the lattice cells (the flash is per cell)
    sub = OracleState(Nc, c.size, wl.grid.v[c], wl.grid.poro[c], wl.grid.depth[c],
                      init={k: np.asarray(v).reshape(nb, -1)[c].ravel() for k, v in wl.state.items()})
    oracle.cal_flash(wl.fluid, sub)
    clean = sub.ill == 0
    rows = _stencil(c, nx, ny, nz)
    slots = [np.where(ok, _slot(rp, col, np.where(ok, r, 0), c), -1) for r, ok in rows]
    scale = np.abs(val).max(axis=2)                                    # [slot][equation]
    # largest entry of each equation row over the row's blocks
    rowscale = np.zeros((nb, b))
    blk_row = np.repeat(np.arange(nb), np.diff(rp))
    np.maximum.at(rowscale, blk_row, scale)
    worst, checked, total, kinked = 0.0, 0, 0, 0
    for k in range(b):
        h = np.full(c.size, 1e-3) if k == 0 else 1e-5 * N0[c, k - 1]
        fds = []
        for mult in (1.0, 4.0):
            out = []
            for s in (1, -1):
                P, N = P0.copy(), N0.copy()
                if k == 0:
                    P[c] += s * mult * h
                else:
                    N[c, k - 1] += s * mult * h
                out.append(run(P, N))
            (Fp, lp), (Fm, lm) = out
            ok = np.all(lp[c] == lab0[c], axis=1) & np.all(lm[c] == lab0[c], axis=1)
            fds.append(([(Fp[r] - Fm[r]) / (2 * mult * h[:, None]) for r, _ in rows], ok))
        ok = clean & fds[0][1] & fds[1][1]
        # a column is differentiable at this state when both step sizes agree
        # (upwind switches and table kinks of kr / Pc break that)
        smooth = ok.copy()
        for (r, inside), f1, f4 in zip(rows, fds[0][0], fds[1][0]):
            rs = rowscale[np.where(inside, r, 0)] + 1e-300
            smooth &= ~inside | np.all(np.abs(f1 - f4) / rs <= 1e-6, axis=1)
        kinked += int((ok & ~smooth).sum())
        total += c.size
        checked += int(smooth.sum())
        for (r, inside), sl, f1 in zip(rows, slots, fds[0][0]):
            m = smooth & inside
            J = val[sl[m], :, k]
            err = np.abs(J - f1[m]) / (rowscale[r[m]] + 1e-300)
            worst = max(worst, float(err.max()) if err.size else 0.0)
    print(f"100x100x50: {checked} of {total} lattice columns checked, worst |J - FD| / row scale {worst:.2e}; "
          f"{kinked} columns not differentiable at this state (step sizes h and 4h disagree)")
    assert checked >= 0.5 * total
    assert worst < 1e-6
    dev.close()


# ------------------------------------------------------------- channelized grid
def test_channelized_adjoint_matches_oracle(oracle):
    """J^T lambda, d(lambda.R)/d(trans) and d(lambda.R)/dVp on the device
    against the oracle's BSR, fluxes and accumulation, from identical flash
    outputs (12x44x17, Nc = 4)."""
    wl = W.channelized_grid(12, 44, 17, Nc=4)
    nb, b = wl.grid.nb, wl.Nc + 2
    st = state_from_workload(wl)
    oracle.cal_flash(wl.fluid, st)
    dev = IsoTFIM.from_workload(wl)
    _upload_flash(dev, st)
    dev.CalKrPc()
    dev.CalRock()
    dev.CalRes(DT, fetch=False)
    dev.AssembleMatBulks(DT)
    oracle.cal_krpc(wl.sat, st)
    oracle.cal_rock(wl.rock, st)
    rp, col, val = oracle.assemble(wl.fluid, st, wl.grid, DT)
    J = sp.bsr_matrix((val, col, rp), shape=(nb * b, nb * b)).tocsr()
    rng = np.random.Generator(np.random.PCG64(21))
    lam = rng.normal(size=nb * b)
    ok, worst = close(dev.jt_lambda(lam), J.T @ lam)
    print(f"channelized 12x44x17: J^T lambda worst {worst:.2e}")
    assert ok
    dT, dV = dev.dR_dparams(DT, lam)
    # the residual is linear in each transmissibility and in each pore volume:
    # the oracle's residual with one connection alone, minus the accumulation
    # alone, is that connection's dt * flux
    g = wl.grid
    empty = W.Grid(nb, 1, 1, g.v, g.depth, g.poro, np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0))
    res0, _ = oracle.cal_res(wl.fluid, st, empty, DT)
    pick = rng.choice(g.nconn, 1500, replace=False)
    gT = np.zeros(pick.size)
    for n, cidx in enumerate(pick):
        one = W.Grid(nb, 1, 1, g.v, g.depth, g.poro, g.conn_b[cidx:cidx + 1], g.conn_e[cidx:cidx + 1],
                     g.trans[cidx:cidx + 1])
        r1, _ = oracle.cal_res(wl.fluid, st, one, DT)
        gT[n] = lam @ (-(r1 - res0)) / g.trans[cidx]
    ok, worst = close(dT[pick], gT)
    print(f"channelized 12x44x17: d(lambda.R)/dtrans worst {worst:.2e} on {pick.size} connections")
    assert ok
    st2 = st.copy()
    dVp = 1e-3 * st.a["ROCKVP"]
    st2.a["ROCKVP"] += dVp
    r2, _ = oracle.cal_res(wl.fluid, st2, empty, DT)
    gV = (lam.reshape(nb, b) * (-(r2 - res0)).reshape(nb, b)).sum(axis=1) / dVp
    ok, worst = close(dV, gV, rtol=1e-9)
    print(f"channelized 12x44x17: d(lambda.R)/dVp worst {worst:.2e}")
    assert ok
    dev.close()


def test_channelized_full_grid_adjoint_matches_central_differences(oracle):
    """60x220x85 (1.12M cells, Nc = 4), exact LBC derivative: 140 primary
    variables, 60 transmissibilities and 60 bulk volumes, each gradient of
    lambda.F from the adjoint (recomputed flash, J^T lambda, dR/dparams)
    against central differences of the device residual.  Perturbations that
    touch disjoint rows are taken in one residual evaluation: primary
    variables of cells 3 or more apart, connections and volumes without a
    shared cell."""
    import torch
    from opencaeporox_amd.autodiff import FIMProblem, fim_residual
    nx, ny, nz = 60, 220, 85
    wl = W.channelized_grid(nx, ny, nz, Nc=4, seed=4, lbc_exact=1)
    dev = IsoTFIM.from_workload(wl)
    prob = FIMProblem(dev, dt=DT, lNi=wl.state["LNI"])
    nb, b = dev.nb, dev.bs
    X0 = prob.primary()
    T0 = wl.grid.trans.copy()
    v0 = wl.grid.v.copy()
    rng = np.random.Generator(np.random.PCG64(11))
    lam = rng.normal(size=nb * b)
    X = torch.tensor(X0, requires_grad=True)
    Tt = torch.tensor(T0, requires_grad=True)
    vt = torch.tensor(v0, requires_grad=True)
    F = fim_residual(X, Tt, vt, prob)
    (F * torch.tensor(lam)).sum().backward()
    gX, gT, gv = X.grad.numpy(), Tt.grad.numpy(), vt.grad.numpy()
    lab0 = dev.get("PHASEEXIST").reshape(nb, 3).copy()
    lamr = lam.reshape(nb, b)
    # picks: 140 lattice cells (one primary variable each), 60 connections and
    # 60 cells with disjoint row sets
    lat = _lattice(nx, ny, nz)
    sub = OracleState(4, lat.size, wl.grid.v[lat], wl.grid.poro[lat], wl.grid.depth[lat],
                      init={k: np.asarray(v).reshape(nb, -1)[lat].ravel() for k, v in wl.state.items()})
    oracle.cal_flash(wl.fluid, sub)
    cells = rng.choice(lat[sub.ill == 0], 140, replace=False)
    var = rng.integers(0, b, cells.size)
    used = set()
    conns = []
    for cidx in rng.permutation(wl.grid.nconn):
        e = (int(wl.grid.conn_b[cidx]), int(wl.grid.conn_e[cidx]))
        if e[0] in used or e[1] in used:
            continue
        used.update(e)
        conns.append(cidx)
        if len(conns) == 60:
            break
    conns = np.array(conns)
    vcells = rng.choice(np.setdiff1d(np.arange(nb), np.array(sorted(used))), 60, replace=False)

    def fd_pass(Xd=None, Td=None, vd=None):
        out = []
        for s in (1, -1):
            X1 = X0 if Xd is None else X0 + s * Xd
            T1 = T0 if Td is None else T0 + s * Td
            v1 = v0 if vd is None else v0 + s * vd
            Fs = prob.run(X1, T1, v1).reshape(nb, b)
            out.append((Fs, dev.get("PHASEEXIST").reshape(nb, 3).copy()))
        return out

    worst, checked, noisy = 0.0, 0, 0
    for k in range(b):
        sel = cells[var == k]
        if sel.size == 0:
            continue
        idx = sel * b + k
        h = 1e-4 * np.maximum(np.abs(X0[idx]), 1e-8)
        res = {}
        for mult in (1.0, 4.0):
            Xd = np.zeros_like(X0)
            Xd[idx] = mult * h
            (Fp, lp), (Fm, lm) = fd_pass(Xd=Xd)
            ok = np.all(lp[sel] == lab0[sel], axis=1) & np.all(lm[sel] == lab0[sel], axis=1)
            g = np.zeros(sel.size)
            for r, inside in _stencil(sel, nx, ny, nz):
                rr = np.where(inside, r, 0)
                g += np.where(inside, np.sum(lamr[rr] * (Fp[rr] - Fm[rr]), axis=1), 0.0)
            res[mult] = (g / (2 * mult * h), ok)
        (f1, ok1), (f4, ok4) = res[1.0], res[4.0]
        a = gX[idx]
        conv = ok1 & ok4 & (np.abs(f4 - f1) <= 1e-6 * np.maximum(np.maximum(np.abs(f1), np.abs(a)), 1e-300))
        noisy += int((ok1 & ~conv).sum())
        err = np.abs(f1 - a) / np.maximum(np.maximum(np.abs(a), np.abs(f1)), 1e-300)
        assert np.all(err[conv] <= 1e-6), f"primary variable {k}: worst {err[conv].max():.2e}"
        worst = max(worst, float(err[conv].max()) if conv.any() else 0.0)
        checked += int(conv.sum())
    # transmissibilities and volumes: F is linear in them, a large step keeps
    # rounding out of the difference
    hT = np.zeros_like(T0)
    hT[conns] = 1e-2 * T0[conns]
    (Fp, _), (Fm, _) = fd_pass(Td=hT)
    for cidx in conns:
        rows = [int(wl.grid.conn_b[cidx]), int(wl.grid.conn_e[cidx])]
        f = np.sum(lamr[rows] * (Fp[rows] - Fm[rows])) / (2 * hT[cidx])
        err = abs(f - gT[cidx]) / max(abs(f), abs(gT[cidx]), 1e-300)
        assert err <= 1e-6, f"trans[{cidx}]: adjoint {gT[cidx]:.12e} vs FD {f:.12e}"
        worst = max(worst, err)
        checked += 1
    hv = np.zeros_like(v0)
    hv[vcells] = 1e-2 * v0[vcells]
    (Fp, _), (Fm, _) = fd_pass(vd=hv)
    f = np.sum(lamr[vcells] * (Fp[vcells] - Fm[vcells]), axis=1) / (2 * hv[vcells])
    err = np.abs(f - gv[vcells]) / np.maximum(np.maximum(np.abs(f), np.abs(gv[vcells])), 1e-300)
    assert np.all(err <= 1e-6), f"v: worst {err.max():.2e}"
    worst = max(worst, float(err.max()))
    checked += vcells.size
    print(f"60x220x85: {checked} gradients checked, worst relative difference {worst:.2e}; "
          f"{noisy} primary variables skipped (finite differences not converged)")
    assert checked >= 200
    dev.close()


# ------------------------------------------------------------- uneven multi-model batch
class _ModelView:
    """One model's fields of a multi-model batch, as _compare_flash reads them."""

    def __init__(self, batch, m):
        self.batch, self.m = batch, m
        self.nb = batch.models[m].grid.nb

    def get(self, f):
        return self.batch.get(self.m, f)


def test_uneven_batch_models_match_oracle(oracle, oracle_fma):
    """A reduced uneven batch (14 models of 150 to 1,500 cells, Nc from 3 to
    12): every model's flash against the oracle's run of that model, then, from
    the oracle's flash outputs uploaded into the merged contexts, every model's
    residual and BSR rows against the oracle's."""
    wls = W.uneven_batch(nmodels=14, seed=6, cmin=150, cmax=1500)
    assert len({w.fluid.nc for w in wls}) >= 6
    batch = MultiModelBatch(wls)
    batch.autotune()
    batch.newton_pass(DT, graph=True)
    batch.sync()
    states = []
    for m, wl in enumerate(wls):
        st = state_from_workload(wl)
        others = perturbed_oracle_states(oracle, oracle_fma, wl.fluid, st)
        oracle.cal_flash(wl.fluid, st)
        _compare_flash(_ModelView(batch, m), st, wl.Nc, f"model {m} Nc={wl.Nc} ({wl.grid.nb} cells)", others,
                       min_cells=int(0.95 * wl.grid.nb))
        states.append(st)
    # assembly from identical flash outputs, per merged context
    for Nc, (d, idx) in batch.groups.items():
        for f in FLASH_OUT:
            d.set(f, np.concatenate([states[m].a[f] for m in idx]))
        d.CalKrPc()
        d.CalRock()
        res_d, _ = d.CalRes(DT)
        d.AssembleMatBulks(DT)
        rp, col, val, _ = d.bsr()
        for k, m in enumerate(idx):
            wl, st = wls[m], states[m]
            a, e = d.offsets[k], d.offsets[k + 1]
            oracle.cal_krpc(wl.sat, st)
            oracle.cal_rock(wl.rock, st)
            res_o, _ = oracle.cal_res(wl.fluid, st, wl.grid, DT)
            rpo, colo, valo = oracle.assemble(wl.fluid, st, wl.grid, DT)
            bs = Nc + 2
            ok, worst = close(res_d[a * bs:e * bs], res_o)
            assert ok, f"model {m}: residual worst {worst}"
            s0, s1 = rp[a], rp[e]
            assert np.array_equal(rp[a:e + 1] - s0, rpo) and np.array_equal(col[s0:s1] - a, colo)
            for r in range(e - a):
                ok, worst = close(val[s0 + rpo[r]:s0 + rpo[r + 1]], valo[rpo[r]:rpo[r + 1]])
                assert ok, f"model {m}: block row {r} worst {worst}"
        print(f"Nc={Nc}: models {idx} residual and BSR match the oracle")
    batch.close()
