"""Every index the assembly, kr/Pc and well kernels form, checked on the CPU
against the buffers they address, on the full 100x100x50 and 60x220x85
layouts.  The block-CSR pattern, the assembly row order and the perforation
grouping come from the library's own set-up code (csrc/layout.hpp, compiled
here for the host); the index formulas are those of csrc/fim.hip (krpc_aux,
res_only, res_asm, xcd_chunk, block_norms) and csrc/wells.hip, evaluated for
every block and lane of their launch grids."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from opencaeporox_amd import wells as WL
from opencaeporox_amd import workload as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "opencaeporox_amd", "csrc")

HARNESS = r'''
#include "layout.hpp"
extern "C" {
void layout(int64_t nb, int64_t nconn, const int32_t* cb, const int32_t* ce, int64_t* rp, int32_t* col,
            int32_t* ts, int32_t* nbc) { ocpx::bsr_layout(nb, nconn, cb, ce, rp, col, ts, nbc); }
int64_t order(int64_t nb, int64_t nconn, const int32_t* cb, const int32_t* ce, int32_t* out)
{
    std::vector<int32_t> o = ocpx::pencil_order(nb, nconn, cb, ce);
    for (size_t k = 0; k < o.size(); k++) out[k] = o[k];
    return (int64_t)o.size();
}
int64_t wells(int32_t nwell, const int32_t* p0, const int32_t* np, int64_t nperf, const int64_t* cell,
              int32_t* pwell, int64_t* ucell, int32_t* uptr, int32_t* uperf)
{
    std::vector<int32_t> pw, up, upf;
    std::vector<int64_t> uc;
    ocpx::well_layout(nwell, p0, np, nperf, cell, pw, uc, up, upf);
    for (size_t k = 0; k < pw.size(); k++) { pwell[k] = pw[k]; uperf[k] = upf[k]; }
    for (size_t k = 0; k < uc.size(); k++) ucell[k] = uc[k];
    for (size_t k = 0; k < up.size(); k++) uptr[k] = up[k];
    return (int64_t)uc.size();
}
}
'''


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    d = tmp_path_factory.mktemp("layout")
    src = d / "h.cpp"
    src.write_text(HARNESS)
    so = d / "libh.so"
    subprocess.run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", f"-I{CSRC}", str(src), "-o", str(so)],
                   check=True)
    L = C.CDLL(str(so))
    L.order.restype = C.c_int64
    L.wells.restype = C.c_int64
    return L


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def aux_len(b, nc):                    # csrc/ocpx_kernels.hpp
    return (24 + 9 * b + 3 * nc + 1) & ~1


def xcd_chunk(b, nblk):                # csrc/fim.hip
    X = 8
    x, k, q, r = b % X, b // X, nblk // X, nblk % X
    return x * q + np.minimum(x, r) + k


def _grid(nx, ny, nz):
    nb = nx * ny * nz
    one = np.ones(nb)
    return W.cartesian_grid(nx, ny, nz, 6.096, one, one, one, one * 0.2)


@pytest.mark.parametrize("dims", [(100, 100, 50), (60, 220, 85)])
def test_assembly_indices_in_bounds(lib, dims):
    g = _grid(*dims)
    nb, nconn = g.nb, g.nconn
    cb, ce = np.ascontiguousarray(g.conn_b, np.int32), np.ascontiguousarray(g.conn_e, np.int32)
    nnzb = nb + 2 * nconn
    rp = np.zeros(nb + 1, np.int64)
    col, ts, nbc = (np.zeros(nnzb, np.int32) for _ in range(3))
    lib.layout(C.c_int64(nb), C.c_int64(nconn), _p(cb), _p(ce), _p(rp), _p(col), _p(ts), _p(nbc))
    order = np.zeros(nb, np.int32)
    no = lib.order(C.c_int64(nb), C.c_int64(nconn), _p(cb), _p(ce), _p(order))
    assert no == nb, "a natural Cartesian numbering gets the pencil order"
    # the pattern: rowptr, diagonal first, connection of every off-diagonal slot
    assert rp[0] == 0 and rp[-1] == nnzb and np.all(np.diff(rp) >= 1)
    row = np.repeat(np.arange(nb), np.diff(rp))
    diag = rp[:-1]
    assert np.array_equal(col[diag], np.arange(nb)) and np.all(nbc[diag] == -1)
    off = np.ones(nnzb, bool)
    off[diag] = False
    k = nbc[off]
    assert k.min() >= 0 and k.max() < nconn
    r = row[off]
    assert np.all((cb[k] == r) | (ce[k] == r)), "a slot's connection touches its row"
    assert np.array_equal(np.where(cb[k] == r, ce[k], cb[k]), col[off])
    assert ts.min() >= 0 and ts.max() < nnzb and np.array_equal(ts[ts], np.arange(nnzb))
    assert np.array_equal(row[ts], col), "the transposed slot sits in the block column's row"
    # the row order is a permutation
    assert np.array_equal(np.sort(order), np.arange(nb))
    for B in range(3, 15):                              # nc = 2 .. 13 (Nc = 1 .. 12)
        nc, nph = B - 1, 3
        G = 64 // B
        nblk = (nb + G - 1) // G
        # xcd_chunk renumbers the blocks one to one
        blk = xcd_chunk(np.arange(nblk), nblk)
        assert np.array_equal(np.sort(blk), np.arange(nblk)), f"B={B}: xcd_chunk is not a permutation"
        # rows of valid lanes: p = blk * G + g < nb, g < G (lanes g >= G and p >= nb only
        # take part in block_norms' barriers and write nothing)
        p = (blk[:, None] * G + np.arange(G)[None, :]).ravel()
        p = p[p < nb]
        assert p.size == nb and np.array_equal(np.sort(order[p]), np.arange(nb))
        # records and dXs/dXp rows: the largest offsets the kernels form in a cell
        RL, REC = aux_len(B, nc), nph * B * B
        SV, SX = 24, 24 + 9 * B
        rec = [8 * j + c for j in range(nph) for c in range(8)]                      # scalars, 16-byte pairs
        rec += [SV + 3 * B * j + q * B + lane for j in range(nph) for q in range(3) for lane in range(B)]
        rec += [SX + j * nc + lane for j in range(nph) for lane in range(nc)]
        assert RL % 2 == 0 and max(rec) < RL and min(rec) >= 0, f"B={B}: record offset {max(rec)} of {RL}"
        # res_asm's LDS copy of a record: 16-byte chunk ch of the copy comes from record
        # offset 2ch (the 24 scalars) or XA + 2(ch - 12) (x_ij to the record's end); lane r
        # of a row group takes chunks r, r + B, ...; the scalars and x_ij are read at
        # [0, 24) and OX + j nc + i; two records per row group, G groups per block
        XA = SX & ~1
        SR, OX = 24 + (RL - XA), 24 + (SX - XA)
        NCH = SR // 2
        assert SR % 2 == 0 and XA % 2 == 0
        src = np.array([2 * ch if ch < 12 else XA + 2 * (ch - 12) for ch in range(NCH)])
        assert src.min() >= 0 and src.max() + 1 < RL, f"B={B}: staged chunk past the record"
        dst = np.arange(NCH) * 2
        assert dst.max() + 1 < SR
        lanes = sorted(ch for rr in range(B) for ch in range(rr, NCH, B))
        assert lanes == list(range(NCH)), f"B={B}: chunks not covered once each"
        staged = np.zeros(RL, bool)                      # record offsets the copy holds, at copy index
        where = np.full(RL, -1)
        for ch in range(NCH):
            for h in range(2):
                staged[src[ch] + h] = True
                where[src[ch] + h] = dst[ch] + h
        read = [8 * j + q for j in range(nph) for q in range(8)] + \
               [SX + j * nc + i for j in range(nph) for i in range(nc)]
        assert staged[read].all(), f"B={B}: a read record entry is not staged"
        assert np.array_equal(where[[SX + j * nc + i for j in range(nph) for i in range(nc)]],
                              [OX + j * nc + i for j in range(nph) for i in range(nc)])
        assert np.array_equal(where[:24], np.arange(24))
        assert G * 2 * SR * 8 + 2048 <= 65536, f"B={B}: LDS per block"
        x_rows = [k * B + lane for k in range(nph) for lane in range(B)]            # dS columns
        x_rows += [(nph + j * nc + i) * B + lane for j in range(nph) for i in range(nc) for lane in range(B)]
        assert max(x_rows) < REC, f"B={B}: dXs/dXp offset {max(x_rows)} of {REC}"
        # blocks: column lane of every slot of the row, the diagonal's B rows
        assert (nnzb - 1) * B * B + (B - 1) * B + (B - 1) < nnzb * B * B
        # block_norms writes bmax / barg [2 * blockIdx + 1] < 2 * nblk, the size allocated


def test_well_indices_in_bounds(lib):
    """The five vertical wells of the 100x100x50 workload: every perforation,
    perforated cell and well index the well kernels form."""
    wl = W.heterogeneous_grid(10, 10, 5, Nc=3)
    nb = 100 * 100 * 50
    g = _grid(100, 100, 50)
    wl.grid = g
    P = np.full(nb, 200.0)
    wl.state = {"P": P}
    wells = WL.corner_and_centre_wells(wl)
    arr, cell, WI, dG, xi = WL.pack(wells, 4)
    nw, npf = len(wells), len(cell)
    p0 = np.array([arr[k].perf0 for k in range(nw)], np.int32)
    npw = np.array([arr[k].nperf for k in range(nw)], np.int32)
    pwell, uperf = np.zeros(npf, np.int32), np.zeros(npf, np.int32)
    ucell, uptr = np.zeros(npf, np.int64), np.zeros(npf + 1, np.int32)
    nu = lib.wells(nw, _p(p0), _p(npw), C.c_int64(npf), _p(np.ascontiguousarray(cell)), _p(pwell), _p(ucell),
                   _p(uptr), _p(uperf))
    ucell, uptr = ucell[:nu], uptr[:nu + 1]
    assert cell.min() >= 0 and cell.max() < nb
    assert np.array_equal(np.sort(uperf), np.arange(npf)) and uptr[0] == 0 and uptr[-1] == npf
    assert np.all(np.diff(uptr) >= 1) and np.all(np.diff(ucell) > 0)
    assert np.array_equal(ucell[np.repeat(np.arange(nu), np.diff(uptr))], cell[uperf])
    assert pwell.min() >= 0 and pwell.max() < nw
    for w in range(nw):
        assert np.all(pwell[p0[w]:p0[w] + npw[w]] == w)
    for B in range(3, 15):
        nc = B - 1
        # well_perf / well_apply / well_rows / well_spmv: qperf [nperf][nc], bdiag
        # [nperf][B][B], Bnw / Cwn [nperf][B], Dw [nwell][B][B], x / y [(nb + nwell) B]
        assert (npf - 1) * nc + nc - 1 < npf * nc
        assert (nb + nw - 1) * B < (nb + nw) * B
        assert ucell.max() * B + B - 1 < nb * B
