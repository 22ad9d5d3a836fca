"""Differentiable FIM residual for gradient-based history matching.

    F(X, trans, v) = accumulation + dt * flux

is the un-negated residual of IsoT_FIM::CalRes (src/IsoThermalMethod.cpp:
1090-1185; the reference stores -F).  X = (P, N_0 .. N_{nc-1}) per cell, trans
the connection transmissibilities, v the bulk volumes (pore volume = v * phi(P)).

FIMResidual is a torch.autograd.Function over the C ABI:
  forward   restores the previous time step (ocpx_reset_to_last_step), sets X,
            runs flash + kr/Pc + rock + residual on the device;
  backward  recomputes the same pass (no flash iterate is stored), assembles
            J = dF/dX and returns
              dL/dX     = J^T lambda                    (ocpx_jt_lambda)
              dL/dtrans = lambda . dF/dtrans            (ocpx_dR_dparams)
              dL/dv     = lambda_n[0] * phi_n           (dF_0/dv = phi)
            for lambda = dL/dF.
J is the reference's FIM Jacobian, with the flash's implicit-function
derivatives dXs/dXp (tangent) inside.  Everything runs through libocpx.so.
"""
import numpy as np
import torch

from .fim import BATCHED


class FIMProblem:
    """One device context and the fixed data of a residual evaluation: the
    previous time step (phase state the flash starts from, SkipPSA data and
    lNi) and dt."""

    def __init__(self, dev, dt, lNi=None, mode=BATCHED):
        self.dev, self.dt, self.mode = dev, float(dt), mode
        dev.UpdateLastTimeStep()
        if lNi is not None:
            dev.set("LNI", lNi)

    def run(self, X, trans, v, assemble=False):
        """F(X, trans, v) as a numpy vector [nb * (nc + 1)]."""
        d = self.dev
        d.ResetToLastTimeStep()
        d.set_geometry(v, trans)
        Xn = np.ascontiguousarray(X, np.float64).reshape(d.nb, d.bs)
        d.set("P", Xn[:, 0])
        d.set("NI", Xn[:, 1:])
        d.CalFlash(self.mode)
        d.CalKrPc()
        d.CalRock()
        res, _ = d.CalRes(self.dt)
        if assemble:
            d.AssembleMatBulks(self.dt)
        return -res

    def primary(self):
        d = self.dev
        return np.concatenate([d.get("P")[:, None], d.get("NI").reshape(d.nb, d.nc)], axis=1).ravel()


def _np(t):
    return t.detach().to("cpu", torch.float64).numpy()


class FIMResidual(torch.autograd.Function):
    @staticmethod
    def forward(ctx, X, trans, v, prob):
        F = prob.run(_np(X), _np(trans), _np(v))
        ctx.prob = prob
        ctx.save_for_backward(X, trans, v)
        return torch.from_numpy(F).to(X.device)

    @staticmethod
    def backward(ctx, gF):
        X, trans, v = ctx.saved_tensors
        prob = ctx.prob
        d = prob.dev
        prob.run(_np(X), _np(trans), _np(v), assemble=True)     # recompute, then J
        lam = np.ascontiguousarray(_np(gF))
        gX = d.jt_lambda(lam) if ctx.needs_input_grad[0] else None
        gT = gv = None
        if ctx.needs_input_grad[1] or ctx.needs_input_grad[2]:
            dT, dVp = d.dR_dparams(prob.dt, lam)
            gT = dT
            gv = dVp * d.get("PORO")
        out = []
        for g, ref, need in ((gX, X, 0), (gT, trans, 1), (gv, v, 2)):
            out.append(torch.from_numpy(np.ascontiguousarray(g)).to(ref.device)
                       if (g is not None and ctx.needs_input_grad[need]) else None)
        return out[0], out[1], out[2], None


def fim_residual(X, trans, v, prob):
    return FIMResidual.apply(X, trans, v, prob)
