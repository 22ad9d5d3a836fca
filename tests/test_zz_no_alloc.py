"""ocpx_phase_equilibrium_noalloc (scratch in a file-scope device array, no
allocation) against ocpx_phase_equilibrium on the same cells: identical bits,
on two calls in a row, the second on fewer cells."""
import ctypes as C

import numpy as np
import pytest

from opencaeporox_amd import abi
from opencaeporox_amd import workload as W
from opencaeporox_amd.fim import BATCHED, IsoTFIM

pytestmark = pytest.mark.gpu


def _noalloc(lib, wl, Nc):
    import torch
    n = wl.grid.nb
    nc = Nc + 1
    dev = torch.device("cuda", 0)
    P = torch.tensor(wl.state["P"], dtype=torch.float64, device=dev)
    T = torch.tensor(wl.state["T"], dtype=torch.float64, device=dev)
    z = torch.tensor(np.ascontiguousarray(wl.state["NI"].reshape(n, nc)[:, :Nc].T), dtype=torch.float64,
                     device=dev)
    NP = torch.full((n,), -1, dtype=torch.int32, device=dev)
    nu = torch.full((n,), -1.0, dtype=torch.float64, device=dev)
    x = torch.full((2 * Nc, n), -1.0, dtype=torch.float64, device=dev)
    ft = torch.full((n,), -1, dtype=torch.int32, device=dev)
    st = torch.cuda.current_stream().cuda_stream
    rc = lib.ocpx_phase_equilibrium_noalloc(C.byref(wl.fluid), n, C.c_void_p(P.data_ptr()), C.c_void_p(T.data_ptr()),
                                            C.c_void_p(z.data_ptr()), C.c_void_p(NP.data_ptr()),
                                            C.c_void_p(nu.data_ptr()), C.c_void_p(x.data_ptr()),
                                            C.c_void_p(ft.data_ptr()), C.c_void_p(st))
    assert rc == 0, lib.ocpx_last_error(None).decode()
    torch.cuda.synchronize()
    return NP.cpu().numpy(), nu.cpu().numpy(), x.cpu().numpy().T.reshape(n, 2, Nc), ft.cpu().numpy()


def test_noalloc_matches_phase_equilibrium():
    import torch
    assert torch.cuda.is_available()
    lib = abi.load()
    Nc = 9
    for n, seed in ((6000, 41), (1500, 42)):
        wl, _ = W.flash_cells(n, Nc, seed=seed, P=np.linspace(5.0, 250.0, n))
        ctx = IsoTFIM.from_workload(wl)
        ctx.set_flash_mapping(1)
        ctx.set_flash_stages(False)
        ctx.PhaseEquilibrium(BATCHED)
        NP, nu, x = ctx.pe_result()
        ft = ctx.get("FTYPE_OUT")
        ctx.close()
        NPa, nua, xa, fta = _noalloc(lib, wl, Nc)
        print(f"{n} cells: {np.sum(NP == 2)} two-phase")
        assert np.sum(NP == 2) > 100
        assert np.array_equal(NPa, NP)
        assert np.array_equal(nua, nu[:, 0])
        assert np.array_equal(xa, x)
        assert np.array_equal(fta, ft)


def test_noalloc_rejects_bad_sizes():
    lib = abi.load()
    wl, _ = W.flash_cells(8, 3, seed=1)
    fl = wl.fluid
    keep = fl.nc
    fl.nc = 13
    rc = lib.ocpx_phase_equilibrium_noalloc(C.byref(fl), 8, *([None] * 7), None)
    fl.nc = keep
    assert rc == -1 and b"Nc must be" in lib.ocpx_last_error(None)
    rc = lib.ocpx_phase_equilibrium_noalloc(C.byref(fl), 1 << 31, *([None] * 7), None)
    assert rc == -1 and b"cells" in lib.ocpx_last_error(None)
