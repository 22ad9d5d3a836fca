"""Python face of the C ABI, shaped like the reference's IsoT_FIM loops.

`IsoTFIM` owns one ocpx context (device-resident BulkVarSet mirror + BSR)
and exposes the per-Newton-step entry points under the reference's names:

    CalFlash        IsoT_FIM::CalFlash + PassFlashValue  (src/IsoThermalMethod.cpp:1016-1071)
    CalKrPc         IsoT_FIM::CalKrPc                     (:1073-1088)
    CalRock         IsothermalMethod::CalRock             (:19-30)
    CalRes          IsoT_FIM::CalRes                      (:1090-1185, bulk part)
    CalResWells     its well part (src/WellPeaceman.cpp:1226-1612, after CalRes / AssembleMatBulks)
    AssembleMatBulks IsoT_FIM::AssembleMatBulks + CopyRhs (:1187-1269, :742-752)
    GetSolution     IsoT_FIM::GetSolution                 (:1282-1418, bulk part)
    UpdateLastTimeStep / ResetToLastTimeStep              (:1420-1515)

Everything runs through libocpx.so on the GPU; there is no CPU path here.
"""
import ctypes as C

import numpy as np

from . import abi

PLAIN, BATCHED = 0, 1


class OcpxError(RuntimeError):
    pass


class IsoTFIM:
    def __init__(self, fluid, sat, rock, grid, stream=None):
        self.lib = abi.load()
        self.fluid, self.sat, self.rock, self.grid = fluid, sat, rock, grid
        self.Nc = fluid.nc
        self.nc = self.Nc + 1
        self.bs = self.nc + 1
        self.nb = grid.nb
        self.nconn = grid.nconn
        self._g = grid.to_abi()
        h = C.c_void_p()
        rc = self.lib.ocpx_ctx_create(C.byref(fluid), C.byref(sat), C.byref(rock), C.byref(self._g),
                                      stream, C.byref(h))
        self.h = h
        if rc != 0:
            msg = self.lib.ocpx_last_error(h if h.value else None).decode()
            if h.value:
                self.lib.ocpx_ctx_destroy(h)
                self.h = C.c_void_p()
            raise OcpxError(f"ocpx_ctx_create failed ({rc}): {msg}")

    @classmethod
    def from_workload(cls, wl, stream=None):
        obj = cls(wl.fluid, wl.sat, wl.rock, wl.grid, stream)
        obj.set_state(wl.state)
        return obj

    @classmethod
    def from_workloads(cls, wls, stream=None):
        """Several models of one Nc in one context (ocpx_ctx_create_multi):
        cells and connections concatenated in model order, each cell with its
        own model's fluid, tables and rock."""
        obj = cls.__new__(cls)
        obj.lib = abi.load()
        n = len(wls)
        if n == 0 or any(w.fluid.nc != wls[0].fluid.nc for w in wls):
            raise ValueError("models of one context must share Nc")
        obj.fluid, obj.sat, obj.rock = wls[0].fluid, wls[0].sat, wls[0].rock
        obj.Nc = wls[0].fluid.nc
        obj.nc = obj.Nc + 1
        obj.bs = obj.nc + 1
        obj.nb = sum(w.grid.nb for w in wls)
        obj.nconn = sum(w.grid.nconn for w in wls)
        obj.offsets = np.cumsum([0] + [w.grid.nb for w in wls])
        fl = (abi.Fluid * n)(*[w.fluid for w in wls])
        sa = (abi.Sat * n)(*[w.sat for w in wls])
        ro = (abi.Rock * n)(*[w.rock for w in wls])
        gr = (abi.Grid * n)(*[w.grid.to_abi() for w in wls])
        obj._keep = (fl, sa, ro, gr, [w.grid for w in wls])
        h = C.c_void_p()
        rc = obj.lib.ocpx_ctx_create_multi(n, fl, sa, ro, gr, stream, C.byref(h))
        obj.h = h
        if rc != 0:
            msg = obj.lib.ocpx_last_error(h if h.value else None).decode()
            raise OcpxError(f"ocpx_ctx_create_multi failed ({rc}): {msg}")
        for k in wls[0].state:
            obj.set(k, np.concatenate([np.asarray(w.state[k]).ravel() for w in wls]))
        return obj

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.ocpx_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != 0:
            raise OcpxError(f"{what} failed ({rc}): {self.lib.ocpx_last_error(self.h).decode()}")

    # ---- state exchange (BulkVarSet layout)
    def set(self, name, arr):
        fid = abi.FIELD_ID[name]
        dt = abi.INT_FIELDS.get(name, np.float64)
        a = np.ascontiguousarray(np.asarray(arr, dtype=dt).ravel())
        n = self.nb * abi.field_len(name, self.Nc)
        if a.size != n:
            raise ValueError(f"{name}: expected {n} entries, got {a.size}")
        self._chk(self.lib.ocpx_set_field(self.h, fid, a.ctypes.data), f"set {name}")

    def get(self, name):
        fid = abi.FIELD_ID[name]
        dt = abi.INT_FIELDS.get(name, np.float64)
        a = np.zeros(self.nb * abi.field_len(name, self.Nc), dt)
        self._chk(self.lib.ocpx_get_field(self.h, fid, a.ctypes.data), f"get {name}")
        return a

    def set_state(self, st):
        for k, v in st.items():
            self.set(k, v)

    def sync(self):
        self._chk(self.lib.ocpx_sync(self.h), "sync")

    # ---- hot path
    def PhaseEquilibrium(self, mode=PLAIN):
        self._chk(self.lib.ocpx_phase_equilibrium(self.h, mode), "phase equilibrium")

    def pe_result(self):
        NP = np.zeros(self.nb, np.int32)
        nu = np.zeros(2 * self.nb)
        x = np.zeros(2 * self.nb * self.Nc)
        self._chk(self.lib.ocpx_get_pe_result(self.h, NP.ctypes.data, nu.ctypes.data, x.ctypes.data),
                  "pe result")
        return NP, nu.reshape(-1, 2), x.reshape(-1, 2, self.Nc)

    def set_flash_mapping(self, lanes_per_cell):
        self._chk(self.lib.ocpx_set_flash_mapping(self.h, lanes_per_cell), "flash mapping")

    def flash_mapping(self):
        v = C.c_int32()
        self._chk(self.lib.ocpx_get_flash_mapping(self.h, C.byref(v)), "flash mapping")
        return v.value

    def autotune_flash(self):
        """Times the batched phase-equilibrium kernel per mapping on the
        current state; keeps the fastest.  Returns {lanes_per_cell: ms}."""
        ms = np.zeros(5)
        self._chk(self.lib.ocpx_autotune_flash(self.h, ms.ctypes.data), "autotune")
        out = {k: v for k, v in zip((1, 4, 8, 16), ms[:4]) if v > 0}
        if ms[4] > 0:
            out["natural_order"] = ms[4]
        return out

    def set_flash_queue_order(self, longest_first):
        self._chk(self.lib.ocpx_set_flash_queue_order(self.h, 1 if longest_first else 0), "queue order")

    def flash_queue_order(self):
        v = C.c_int32()
        self._chk(self.lib.ocpx_get_flash_queue_order(self.h, C.byref(v)), "queue order")
        return bool(v.value)

    def set_flash_stages(self, on):
        """Stages of the batched flash: 0 (or False) none, 1 the strict stability
        test in its own kernels, 2 also the split's first SSM pass, 3 (default, or
        True) also its Newton steps."""
        v = 3 if on is True else (0 if on is False else int(on))
        self._chk(self.lib.ocpx_set_flash_stages(self.h, v), "flash stages")

    def set_flash_long(self, evals):
        """Cells over `evals` fugacity evaluations in the previous batched flash
        run beside the stages from the start (0 = off)."""
        self._chk(self.lib.ocpx_set_flash_long(self.h, int(evals)), "flash long cells")

    def flash_long(self):
        v = C.c_int32()
        self._chk(self.lib.ocpx_get_flash_long(self.h, C.byref(v)), "flash long cells")
        return v.value

    def flash_stages(self):
        v = C.c_int32()
        self._chk(self.lib.ocpx_get_flash_stages(self.h, C.byref(v)), "flash stages")
        return v.value

    def CalFlash(self, mode=PLAIN):
        self._chk(self.lib.ocpx_flash_der(self.h, mode), "flash")

    def CalKrPc(self):
        self._chk(self.lib.ocpx_krpc_der(self.h), "krpc")

    def CalRock(self):
        self._chk(self.lib.ocpx_cal_rock(self.h), "rock")

    def CalRes(self, dt, fetch=True):
        res = np.zeros(self.nb * self.bs) if fetch else None
        norms = abi.ResNorms()
        self._chk(self.lib.ocpx_cal_res(self.h, dt, res.ctypes.data if fetch else None,
                                        C.byref(norms) if fetch else None), "residual")
        self._wells_in_system = 0
        if not fetch:
            return None
        return res, np.array([norms.maxRelRes_V, norms.maxId_V, norms.maxRelRes_N, norms.maxId_N])

    def AssembleMatBulks(self, dt):
        self._chk(self.lib.ocpx_assemble(self.h, dt), "assemble")
        self._wells_in_system = 0

    def bsr(self):
        nnz = C.c_int64()
        self._chk(self.lib.ocpx_get_bsr(self.h, C.byref(nnz), None, None, None, None), "bsr size")
        n = nnz.value
        rowptr = np.zeros(self.nb + 1, np.int64)
        col = np.zeros(n, np.int32)
        val = np.zeros(n * self.bs * self.bs)
        rhs = np.zeros(self.nb * self.bs)
        self._chk(self.lib.ocpx_get_bsr(self.h, C.byref(nnz), rowptr.ctypes.data, col.ctypes.data,
                                        val.ctypes.data, rhs.ctypes.data), "bsr")
        return rowptr, col, val.reshape(n, self.bs, self.bs), rhs

    def GetSolution(self, u=None, dSmax=0.1):
        """Newton update from a host u, or (u None) from the device solution
        of the last LinearSolve."""
        if u is None:
            self._chk(self.lib.ocpx_get_solution(self.h, None, dSmax), "get solution")
            return
        u = np.ascontiguousarray(u, np.float64)
        self._chk(self.lib.ocpx_get_solution(self.h, u.ctypes.data, dSmax), "get solution")

    def set_linear_solver(self, method):
        """1 (default): ABF decoupling + FGMRES(30) with the CPR-type preconditioner;
        0: BiCGStab + block Jacobi."""
        self._chk(self.lib.ocpx_set_linear_solver(self.h, int(method)), "linear solver")

    def LinearSolve(self, rtol=1e-3, maxit=200, fetch=True):
        """Device solve of the assembled system (ocpx_linear_solve, the method of
        set_linear_solver).  Returns (u or None, iterations (negative if not
        converged), the final relative residual)."""
        rows = self.nb + getattr(self, "_wells_in_system", 0)     # well rows after ocpx_well_terms
        u = np.zeros(rows * self.bs) if fetch else None
        it, rr = C.c_int32(), C.c_double()
        self._chk(self.lib.ocpx_linear_solve(self.h, rtol, maxit, None if u is None else u.ctypes.data,
                                             C.byref(it), C.byref(rr)), "linear solve")
        return u, it.value, rr.value

    def set_geometry(self, v=None, trans=None):
        va = None if v is None else np.ascontiguousarray(v, np.float64)
        ta = None if trans is None else np.ascontiguousarray(trans, np.float64)
        if va is not None and va.size != self.nb:
            raise ValueError("v must have one entry per cell")
        if ta is not None and ta.size != self.nconn:
            raise ValueError("trans must have one entry per connection")
        self._chk(self.lib.ocpx_set_geometry(self.h, None if va is None else va.ctypes.data,
                                             None if ta is None else ta.ctypes.data), "set geometry")

    def UpdateLastTimeStep(self):
        self._chk(self.lib.ocpx_update_last_step(self.h), "update last step")

    def ResetToLastTimeStep(self):
        self._chk(self.lib.ocpx_reset_to_last_step(self.h), "reset to last step")

    # ---- Peaceman wells
    def SetWells(self, wells):
        """ocpx_set_wells from a list of wells.PeacemanWell."""
        from . import wells as WL
        arr, cell, WI, dG, xi = WL.pack(wells, self.nc)
        self._wells = list(wells)
        self._chk(self.lib.ocpx_set_wells(self.h, len(wells), C.cast(arr, C.c_void_p), cell.ctypes.data,
                                          WI.ctypes.data, dG.ctypes.data, xi.ctypes.data), "set wells")

    def CalResWells(self, dt, jacobian=False):
        """Well terms on top of CalRes (jacobian=False) or AssembleMatBulks
        (jacobian=True): returns (norms, well-row residuals)."""
        nrm = abi.ResNorms()
        wres = np.zeros(len(self._wells))
        self._chk(self.lib.ocpx_well_terms(self.h, dt, 1 if jacobian else 0, C.byref(nrm), wres.ctypes.data),
                  "well terms")
        if jacobian:
            self._wells_in_system = len(self._wells)
        return np.array([nrm.maxRelRes_V, nrm.maxId_V, nrm.maxRelRes_N, nrm.maxId_N]), wres

    def well_blocks(self):
        """(Bnw [nperf, b], Cwn [nperf, b], Dw [nwell, b, b])."""
        npf = sum(len(w.cells) for w in self._wells)
        Bnw, Cwn = np.zeros((npf, self.bs)), np.zeros((npf, self.bs))
        Dw = np.zeros((len(self._wells), self.bs, self.bs))
        self._chk(self.lib.ocpx_get_well_blocks(self.h, Bnw.ctypes.data, Cwn.ctypes.data, Dw.ctypes.data),
                  "well blocks")
        return Bnw, Cwn, Dw

    def wells_state(self):
        """(bhp [nwell], component rates [nwell, nc]) of the last well terms."""
        bhp, qw = np.zeros(len(self._wells)), np.zeros((len(self._wells), self.nc))
        self._chk(self.lib.ocpx_get_wells(self.h, bhp.ctypes.data, qw.ctypes.data), "wells")
        return bhp, qw

    def set_well_bhp(self, bhp):
        b = np.ascontiguousarray(bhp, float)
        self._chk(self.lib.ocpx_set_well_bhp(self.h, b.ctypes.data), "well bhp")

    def newton_pass(self, dt, mode=BATCHED, graph=False):
        self._chk(self.lib.ocpx_newton_pass(self.h, dt, mode, 1 if graph else 0), "newton pass")
        self._wells_in_system = 0

    # ---- differentiation
    def jt_lambda(self, lam):
        lam = np.ascontiguousarray(lam, np.float64)
        out = np.zeros_like(lam)
        self._chk(self.lib.ocpx_jt_lambda(self.h, lam.ctypes.data, out.ctypes.data), "J^T lambda")
        return out

    def dR_dparams(self, dt, lam):
        lam = np.ascontiguousarray(lam, np.float64)
        dT = np.zeros(max(self.nconn, 1))
        dV = np.zeros(self.nb)
        self._chk(self.lib.ocpx_dR_dparams(self.h, dt, lam.ctypes.data, dT.ctypes.data, dV.ctypes.data),
                  "dR/dparams")
        return dT[: self.nconn], dV

    def jt_lambda_dev(self, lam_ptr, out_ptr):
        """J^T lambda on device arrays (raw device pointers, [nb*(nc+1)])."""
        self._chk(self.lib.ocpx_jt_lambda_dev(self.h, lam_ptr, out_ptr), "J^T lambda (device)")

    def dR_dparams_dev(self, dt, lam_ptr, dT_ptr, dV_ptr):
        self._chk(self.lib.ocpx_dR_dparams_dev(self.h, dt, lam_ptr, dT_ptr, dV_ptr), "dR/dparams (device)")

    # ---- timing
    def timing(self, on=True):
        self._chk(self.lib.ocpx_timing_enable(self.h, 1 if on else 0), "timing")

    def timing_get(self, which):
        ms = C.c_double()
        calls = C.c_int64()
        self._chk(self.lib.ocpx_timing_get(self.h, which, C.byref(ms), C.byref(calls)), "timing get")
        return ms.value, calls.value


class MultiModelBatch:
    """Independent models of uneven size and component count: one context per
    Nc holding all models of that Nc (ocpx_ctx_create_multi), each context
    with its own autotuned flash mapping; a Newton step of every model is one
    graph launch (ocpx_batch_newton_pass)."""

    def __init__(self, wls, stream=None):
        groups = {}
        for m, w in enumerate(wls):
            groups.setdefault(w.fluid.nc, []).append(m)
        self.models = wls
        self.groups = {}          # Nc -> (context, model indices)
        self._batch = None
        self.lib = abi.load()
        for Nc in sorted(groups):
            idx = groups[Nc]
            self.groups[Nc] = (IsoTFIM.from_workloads([wls[m] for m in idx], stream), idx)

    @property
    def nb(self):
        return sum(d.nb for d, _ in self.groups.values())

    @property
    def nconn(self):
        return sum(d.nconn for d, _ in self.groups.values())

    def contexts(self):
        return [d for d, _ in self.groups.values()]

    def autotune(self):
        return {Nc: d.autotune_flash() for Nc, (d, _) in self.groups.items()}

    def UpdateLastTimeStep(self):
        for d in self.contexts():
            d.UpdateLastTimeStep()

    def ResetToLastTimeStep(self):
        for d in self.contexts():
            d.ResetToLastTimeStep()

    def newton_pass(self, dt, graph=False):
        """graph=True: every context's pass in one captured graph, one launch
        (ocpx_batch_newton_pass); False: the contexts' passes one after another."""
        if not graph:
            for d in self.contexts():
                d.newton_pass(dt, BATCHED, graph=False)
            return
        if self._batch is None:
            ctxs = self.contexts()
            arr = (C.c_void_p * len(ctxs))(*[d.h.value for d in ctxs])
            out = C.c_void_p()
            rc = self.lib.ocpx_batch_create(arr, len(ctxs), None, C.byref(out))
            if rc != 0:
                raise OcpxError(f"ocpx_batch_create failed ({rc}): {self.lib.ocpx_last_error(None).decode()}")
            self._batch = out
        rc = self.lib.ocpx_batch_newton_pass(self._batch, dt)
        if rc != 0:
            raise OcpxError(f"ocpx_batch_newton_pass failed ({rc}): {self.lib.ocpx_last_error(None).decode()}")

    def sync(self):
        for d in self.contexts():
            d.sync()

    def get(self, m, name):
        """Field of model m (BulkVarSet layout)."""
        Nc = self.models[m].fluid.nc
        d, idx = self.groups[Nc]
        k = idx.index(m)
        L = abi.field_len(name, Nc)
        a = d.get(name)
        return a[d.offsets[k] * L:d.offsets[k + 1] * L]

    def close(self):
        if self._batch is not None:
            self.lib.ocpx_batch_destroy(self._batch)
            self._batch = None
        for d in self.contexts():
            d.close()
