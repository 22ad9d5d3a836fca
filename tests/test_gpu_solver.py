"""Device linear solve (ocpx_linear_solve: ABF decoupling + FGMRES(30) with the
CPR-type preconditioner by default, BiCGStab + block Jacobi as the other
method) against a direct sparse solve of the same system, the reference's
solver settings on the heterogeneous grid with and without its five wells,
and a Newton iteration that stays on the device (assemble -> solve ->
GetSolution)."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from opencaeporox_amd import workload as W
from opencaeporox_amd.fim import BATCHED, IsoTFIM

pytestmark = pytest.mark.gpu


def _system(dev):
    rowptr, col, val, rhs = dev.bsr()
    B = dev.bs
    J = sp.bsr_matrix((val.reshape(-1, B, B), col, rowptr), shape=(dev.nb * B, dev.nb * B)).tocsr()
    return J, rhs


@pytest.mark.parametrize("method", [1, 0])
def test_linear_solve_matches_direct_solve(method):
    wl = W.cube_grid()
    dev = IsoTFIM.from_workload(wl)
    dev.set_linear_solver(method)
    dev.newton_pass(1.0, BATCHED)
    J, rhs = _system(dev)
    u, it, rr = dev.LinearSolve(rtol=1e-12, maxit=2000)
    print(f"10x10x10, method {method}: {it} iterations, ||r||/||b|| {rr:.2e}")
    assert it > 0 and rr < 1e-12
    true = np.linalg.norm(rhs - J @ u) / np.linalg.norm(rhs)
    ref = spla.spsolve(J.tocsc(), rhs)
    err = np.linalg.norm(u - ref) / np.linalg.norm(ref)
    print(f"10x10x10: true residual {true:.2e}, |u - u_direct| / |u_direct| {err:.2e}")
    assert true < 1e-11
    # forward error within cond(J) times the residual: cond(J) is about 2e4 here
    assert err < 1e-7
    dev.close()


def test_linear_solve_reference_tolerance_on_heterogeneous_grid():
    # the reference's solver settings (1e-3 on ||r|| / ||b||, 200 iterations)
    wl = W.heterogeneous_grid(20, 20, 10, Nc=6)
    dev = IsoTFIM.from_workload(wl)
    dev.newton_pass(1.0, BATCHED)
    J, rhs = _system(dev)
    u, it, rr = dev.LinearSolve(rtol=1e-3, maxit=200)
    true = np.linalg.norm(rhs - J @ u) / np.linalg.norm(rhs)
    print(f"20x20x10: {it} iterations, recursive {rr:.2e}, true {true:.2e}")
    assert it > 0 and rr < 1e-3
    assert true < 2e-3
    dev.close()


def test_cpr_on_the_heterogeneous_grid_with_wells():
    """The heterogeneous grid (lognormal permeability, sigma = 2, kv/kh =
    0.1) with one composition everywhere: at 20x20x10 with its five vertical
    wells in the system the reference's settings converge, and so they do on
    the full 100x100x50 grid.  (Per-cell random compositions put unrelated
    fluids side by side; their Newton system defeats every preconditioner
    tried, an exact pressure solve included, from 50x50x25 on.)"""
    from opencaeporox_amd import wells as WL
    wl = W.heterogeneous_grid(20, 20, 10, Nc=6, composition="uniform")
    dev = IsoTFIM.from_workload(wl)
    dev.SetWells(WL.corner_and_centre_wells(wl))
    dev.newton_pass(1.0, BATCHED)
    dev.CalResWells(1.0, jacobian=True)
    u, it, rr = dev.LinearSolve(rtol=1e-3, maxit=200)
    print(f"20x20x10 with wells: {it} iterations, decoupled ||r||/||b|| {rr:.2e}")
    assert it > 0 and rr < 1e-3
    assert u.size == (dev.nb + 5) * dev.bs and np.all(np.isfinite(u))
    dev.close()
    wl = W.heterogeneous_grid(100, 100, 50, Nc=6, composition="uniform")
    dev = IsoTFIM.from_workload(wl)
    dev.newton_pass(1.0, BATCHED)
    u, it, rr = dev.LinearSolve(rtol=1e-3, maxit=200, fetch=False)
    print(f"100x100x50: {it} iterations, decoupled ||r||/||b|| {rr:.2e}")
    assert it > 0 and rr < 1e-3
    dev.close()


def test_nonconvergence_is_reported():
    wl = W.heterogeneous_grid(20, 20, 10, Nc=6)
    dev = IsoTFIM.from_workload(wl)
    dev.newton_pass(1.0, BATCHED)
    u, it, rr = dev.LinearSolve(rtol=1e-30, maxit=3)
    assert it == -3 and rr > 1e-30
    # a solve that did not converge leaves nothing to apply: the state is unchanged
    P0 = dev.get("P").copy()
    from opencaeporox_amd.fim import OcpxError
    with pytest.raises(OcpxError, match="needs ocpx_linear_solve"):
        dev.GetSolution(None)
    assert np.array_equal(dev.get("P"), P0)
    # after a converged solve it applies once
    u, it, rr = dev.LinearSolve(rtol=1e-1, maxit=200)
    assert it > 0
    dev.GetSolution(None)
    assert not np.array_equal(dev.get("P"), P0)
    with pytest.raises(OcpxError, match="needs ocpx_linear_solve"):
        dev.GetSolution(None)
    dev.close()


def _smooth_grid(nx=8, ny=8, nz=4, Nc=4, seed=7):
    """A time step a Newton iteration can converge: hydrostatic pressure, one
    composition everywhere, the residual coming from the +-1% Ni and
    +-0.1 MPa perturbation of the previous step.  (The parity configurations'
    random pressure fields give Newton steps of thousands of bar.)"""
    rng = np.random.Generator(np.random.PCG64(seed))
    nb = nx * ny * nz
    kx = np.exp(rng.normal(np.log(100.0), 1.0, nb))
    grid = W.cartesian_grid(nx, ny, nz, 6.096, kx, kx, 0.1 * kx, rng.uniform(0.1, 0.3, nb))
    fl = W.make_fluid(Nc, 107)
    P = 150.0 + 0.0980665 * (grid.depth - grid.depth.min())
    z = np.tile(rng.dirichlet(np.ones(Nc)), (nb, 1))
    st = W.initial_state(fl, W.make_rock(), grid, P, z, rng)
    return W.Workload("smooth", fl, W.make_sat(), W.make_rock(), grid, st)


def test_newton_iterations_on_device():
    """Assemble -> device solve -> device GetSolution, repeated: the residual
    falls quadratically (the oracle with a direct solve goes 4.5e3, 18,
    3e-2, 2.5e-6 on this step), and the device-solution path matches
    GetSolution with the same u handed over from the host."""
    wl = _smooth_grid()
    a = IsoTFIM.from_workload(wl)
    b = IsoTFIM.from_workload(wl)
    norms = []
    for k in range(4):
        a.newton_pass(1.0, BATCHED)
        b.newton_pass(1.0, BATCHED)
        res, nrm = a.CalRes(1.0)
        norms.append(np.abs(res).max())
        u, it, rr = a.LinearSolve(rtol=1e-10, maxit=2000)
        assert it > 0
        a.GetSolution(None, 0.2)
        b.GetSolution(u, 0.2)
        for f in ("P", "NI", "S"):
            assert np.array_equal(a.get(f), b.get(f)), f"iteration {k} field {f}"
    print("max |R| per Newton iteration:", " ".join(f"{v:.3e}" for v in norms))
    assert norms[-1] < 1e-6 * norms[0]
    a.close()
    b.close()
