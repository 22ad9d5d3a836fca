"""The C-ABI library loads and exports every entry point include/ocpx.h
declares (no GPU needed: nothing is called)."""
import ctypes as C

from opencaeporox_amd import abi


def test_library_exports_declared_symbols():
    lib = abi.load()
    names = abi.declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layout_matches_header():
    # sizes of the ABI structures, computed from the header's member list
    assert C.sizeof(abi.Fluid) == 8 + 9 * 16 * 8 + 256 * 8 + 11 * 8 + 6 * 4 + 6 * 8
    assert C.sizeof(abi.Sat) == 8 + 2 * 4 * 64 * 8
    assert C.sizeof(abi.Grid) == 8 * 8


def test_field_len_matches_python_table():
    lib = abi.load()
    for Nc in (3, 6, 9):
        for name, fid in abi.FIELD_ID.items():
            assert lib.ocpx_field_len(fid, Nc) == abi.field_len(name, Nc), (name, Nc)


def test_creation_rejects_bad_arguments_before_touching_the_gpu():
    """Argument checks run before any HIP call, so they can be exercised here;
    the reason is available from ocpx_last_error(NULL)."""
    import numpy as np
    from opencaeporox_amd import workload as W
    lib = abi.load()
    fl = W.make_fluid(6, 101)
    sat, rock = W.make_sat(), W.make_rock()
    empty = W.Grid(0, 1, 1, np.zeros(0), np.zeros(0), np.zeros(0), np.zeros(0, np.int32),
                   np.zeros(0, np.int32), np.zeros(0)).to_abi()
    h = C.c_void_p()
    rc = lib.ocpx_ctx_create(C.byref(fl), C.byref(sat), C.byref(rock), C.byref(empty), None, C.byref(h))
    assert rc == -1 and not h.value
    assert lib.ocpx_last_error(None).decode() == "grid with no cells"
    bad = W.make_fluid(6, 101)
    bad.nc = 13
    one = W.Grid(1, 1, 1, np.ones(1), np.ones(1), np.full(1, 0.2), np.zeros(0, np.int32),
                 np.zeros(0, np.int32), np.zeros(0)).to_abi()
    rc = lib.ocpx_ctx_create(C.byref(bad), C.byref(sat), C.byref(rock), C.byref(one), None, C.byref(h))
    assert rc == -5 and "2 to 12" in lib.ocpx_last_error(None).decode()
