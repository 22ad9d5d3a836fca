"""Adjoint gradients (Jᵀλ, ∂(λ·F)/∂trans, ∂(λ·F)/∂v) through the
torch.autograd.Function over the C ABI, against central finite differences
of the device residual on sampled parameters, and the device dXs/dXp
against finite differences of the device flash."""
import numpy as np
import pytest
import torch

from opencaeporox_amd import workload as W
from opencaeporox_amd.autodiff import FIMProblem, fim_residual
from opencaeporox_amd.fim import BATCHED, IsoTFIM
from oracle.parity import rounding_sensitive

pytestmark = pytest.mark.gpu

GRAD_RTOL = 1e-6


def _problem(nx=12, ny=44, nz=17):
    # exact LBC derivative: the reference's d(mu)/dx drops a term (integer
    # division, src/OCPFuncPVT.cpp:537), so its Jacobian is not the exact
    # tangent; the gradient check needs the exact one
    wl = W.channelized_grid(nx, ny, nz, Nc=4, seed=4, lbc_exact=1)
    dev = IsoTFIM.from_workload(wl)
    prob = FIMProblem(dev, dt=1.0, lNi=wl.state["LNI"])
    return wl, dev, prob


def _clean_cells(oracle, wl):
    """Cells whose flash trajectory the oracle marks well-conditioned."""
    from oracle.oracle import state_from_workload
    st = state_from_workload(wl)
    oracle.cal_flash(wl.fluid, st)
    return st.ill == 0


def test_adjoint_gradients_match_finite_differences(oracle):
    wl, dev, prob = _problem()
    clean = _clean_cells(oracle, wl)
    nb, b, nconn = dev.nb, dev.bs, dev.nconn
    X0 = prob.primary()
    T0 = wl.grid.trans.copy()
    v0 = wl.grid.v.copy()
    X = torch.tensor(X0, requires_grad=True)
    Tt = torch.tensor(T0, requires_grad=True)
    vt = torch.tensor(v0, requires_grad=True)
    rng = np.random.Generator(np.random.PCG64(11))
    lam = torch.tensor(rng.normal(size=nb * b))
    F = fim_residual(X, Tt, vt, prob)
    (F * lam).sum().backward()
    gX, gT, gv = X.grad.numpy(), Tt.grad.numpy(), vt.grad.numpy()
    F0 = F.detach().numpy()
    labels0 = dev.get("PHASEEXIST").copy()
    lam_np = lam.numpy()

    def fd(kind, idx, h):
        out = []
        for s in (1, -1):
            X1, T1, v1 = X0.copy(), T0.copy(), v0.copy()
            if kind == "X":
                X1[idx] += s * h
            elif kind == "T":
                T1[idx] += s * h
            else:
                v1[idx] += s * h
            Fs = prob.run(X1, T1, v1)
            same = np.array_equal(dev.get("PHASEEXIST"), labels0)
            out.append((Fs, same))
        if not (out[0][1] and out[1][1]):
            return None
        return lam_np @ (out[0][0] - out[1][0]) / (2 * h)

    checked, noisy, worst = 0, 0, 0.0
    picks = [("X", i) for i in rng.choice(np.where(np.repeat(clean, b))[0], 140, replace=False)]
    picks += [("T", i) for i in rng.choice(nconn, 60, replace=False)]
    picks += [("v", i) for i in rng.choice(nb, 60, replace=False)]
    for kind, i in picks:
        if kind == "X":
            h = 1e-4 * max(abs(X0[i]), 1e-8)
            g = gX[i]
        elif kind == "T":
            # F is linear in trans and v: no truncation error, so a large step
            # keeps rounding out of the difference (low-permeability fluxes are
            # ~1e-4 of the cell residuals they are added to)
            h = 1e-2 * T0[i]
            g = gT[i]
        else:
            h = 1e-2 * v0[i]
            g = gv[i]
        f = fd(kind, i, h)
        if f is None:
            continue
        if kind == "X":
            # a converged flash is smooth only to its stopping tolerance: where
            # two step sizes disagree, finite differences cannot check anything
            f4 = fd(kind, i, 4 * h)
            if f4 is None or abs(f4 - f) > GRAD_RTOL * max(abs(f), abs(g), 1e-300):
                noisy += 1
                continue
        err = abs(f - g) / max(abs(g), abs(f), 1e-300)
        worst = max(worst, err)
        assert err <= GRAD_RTOL, f"{kind}[{i}]: adjoint {g:.12e} vs FD {f:.12e} (rel {err:.2e})"
        checked += 1
    print(f"{checked} parameters checked, worst relative difference {worst:.2e}; "
          f"{noisy} state entries skipped (finite differences not converged)")
    assert checked >= 200


def test_device_dXsdXp_matches_finite_differences_of_device_flash(oracle):
    """Same workload and steps as the oracle's own check
    (tests/test_oracle_math.py::test_dXsdXp_matches_finite_differences)."""
    from tests.test_oracle_math import _small_grid_workload
    wl = _small_grid_workload(5, 25, nx=8, ny=8, nz=2, Plo=5.0, Phi=120.0)
    clean = _clean_cells(oracle, wl)
    dev = IsoTFIM.from_workload(wl)
    dev.UpdateLastTimeStep()
    nb, nc = dev.nb, dev.nc

    def flash(P=None, Ni=None):
        dev.ResetToLastTimeStep()
        dev.set("P", wl.state["P"] if P is None else P)
        dev.set("NI", wl.state["NI"] if Ni is None else Ni)
        dev.CalFlash(BATCHED)
        S = dev.get("S").reshape(nb, 3)
        x = dev.get("XIJ").reshape(nb, -1)
        return np.concatenate([S, x], axis=1), dev.get("PHASEEXIST").reshape(nb, 3).copy()

    base, lab = flash()
    D = dev.get("DSEC_DPRI").reshape(nb, 3 * (nc + 1), nc + 1)
    P0 = wl.state["P"].copy()
    N0 = wl.state["NI"].reshape(nb, nc).copy()
    assert ((lab[:, 0] & lab[:, 1]) == 1)[clean].sum() >= 5
    worst = 0.0
    for k in range(nc + 1):
        if k == 0:
            h = 1e-3 * np.ones(nb)
            (op, lp), (om, lm) = flash(P=P0 + h), flash(P=P0 - h)
        else:
            h = 1e-5 * N0[:, k - 1]
            Np, Nm = N0.copy(), N0.copy()
            Np[:, k - 1] += h
            Nm[:, k - 1] -= h
            (op, lp), (om, lm) = flash(Ni=Np.ravel()), flash(Ni=Nm.ravel())
        same = np.all(lp == lab, axis=1) & np.all(lm == lab, axis=1) & clean
        fdv = (op - om) / (2 * h[:, None])
        scale = np.max(np.abs(fdv), axis=1, keepdims=True) + 1e-12
        err = np.abs(D[:, :, k] - fdv) / scale
        worst = max(worst, float(np.max(err[same])))
    print(f"device dXsdXp vs FD of the device flash: worst {worst:.2e}")
    assert worst < 2e-6
