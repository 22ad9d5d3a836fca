"""The pressure AMG's hierarchy and the Gauss-Seidel colouring, checked on
the CPU: csrc/amg_layout.hpp (the set-up code the library runs) is compiled
here for the host, the hierarchy of the 100x100x50 grid's connection graph
is built level by level, and every index array cpr.hip's kernels read is
checked against the graphs it addresses: aggregates partition the rows, the
members lists are those partitions, every fine slot is gathered into the
one coarse slot of (agg[row], agg[col]), the coarse weights are the
Galerkin sums, the diagonal comes first in each row, and no two rows of a
colour are neighbours."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp

from opencaeporox_amd import workload as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "opencaeporox_amd", "csrc")

HARNESS = r'''
#include "amg_layout.hpp"
#include "layout.hpp"
using namespace ocpx;
static amg::Graph G, GC;
static amg::LevelMaps M;
extern "C" {
void layout(int64_t nb, int64_t nconn, const int32_t* cb, const int32_t* ce, int64_t* rp, int32_t* col,
            int32_t* ts, int32_t* nbc) { bsr_layout(nb, nconn, cb, ce, rp, col, ts, nbc); }
int coarsen(int64_t n, const int64_t* rp, const int32_t* ci, const double* w)
{
    G.n = n;
    G.rp.assign(rp, rp + n + 1);
    G.ci.assign(ci, ci + rp[n]);
    G.w.assign(w, w + rp[n]);
    return amg::coarsen(G, GC, M) ? 1 : 0;
}
void sizes(int64_t* out) { out[0] = GC.n; out[1] = (int64_t)GC.ci.size(); }
void fetch(int64_t* crp, int32_t* cci, double* cw, int32_t* agg, int64_t* gp, int64_t* gs, int64_t* mp, int32_t* mi)
{
    for (size_t k = 0; k < GC.rp.size(); k++) crp[k] = GC.rp[k];
    for (size_t k = 0; k < GC.ci.size(); k++) { cci[k] = GC.ci[k]; cw[k] = GC.w[k]; }
    for (size_t k = 0; k < M.agg.size(); k++) agg[k] = M.agg[k];
    for (size_t k = 0; k < M.gp.size(); k++) gp[k] = M.gp[k];
    for (size_t k = 0; k < M.gs.size(); k++) gs[k] = M.gs[k];
    for (size_t k = 0; k < M.mp.size(); k++) mp[k] = M.mp[k];
    for (size_t k = 0; k < M.mi.size(); k++) mi[k] = M.mi[k];
}
int64_t colors(int64_t n, const int64_t* rp, const int32_t* ci, int32_t* cptr, int32_t* rows)
{
    std::vector<int64_t> r(rp, rp + n + 1);
    std::vector<int32_t> c(ci, ci + rp[n]), cp, rw;
    amg::color_rows(n, r, c, cp, rw);
    for (size_t k = 0; k < cp.size(); k++) cptr[k] = cp[k];
    for (size_t k = 0; k < rw.size(); k++) rows[k] = rw[k];
    return (int64_t)cp.size() - 1;
}
}
'''

DENSE_MAX = 64                      # PressureAmg::DENSE_MAX


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    d = tmp_path_factory.mktemp("amg")
    src = d / "h.cpp"
    src.write_text(HARNESS)
    so = d / "libh.so"
    subprocess.run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", f"-I{CSRC}", str(src), "-o", str(so)],
                   check=True)
    L = C.CDLL(str(so))
    L.coarsen.restype = C.c_int
    L.colors.restype = C.c_int64
    return L


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _level0(lib, dims):
    one = np.ones(int(np.prod(dims)))
    rng = np.random.Generator(np.random.PCG64(5))
    k = np.exp(rng.normal(np.log(100.0), 2.0, one.size))
    g = W.cartesian_grid(*dims, 6.096, k, k, 0.1 * k, one * 0.2)
    nb, nconn = g.nb, g.nconn
    cb, ce = np.ascontiguousarray(g.conn_b, np.int32), np.ascontiguousarray(g.conn_e, np.int32)
    nnz = nb + 2 * nconn
    rp = np.zeros(nb + 1, np.int64)
    col, ts, nbc = (np.zeros(nnz, np.int32) for _ in range(3))
    lib.layout(C.c_int64(nb), C.c_int64(nconn), _p(cb), _p(ce), _p(rp), _p(col), _p(ts), _p(nbc))
    w = np.where(nbc >= 0, np.abs(g.trans[np.maximum(nbc, 0)]), 0.0)     # as cpr_setup forms it
    return nb, rp, col, w


def test_hierarchy_indices(lib):
    n, rp, ci, w = _level0(lib, (100, 100, 50))
    sizes = [n]
    while n > DENSE_MAX:
        if not lib.coarsen(C.c_int64(n), _p(rp), _p(ci), _p(w)):
            break
        sz = np.zeros(2, np.int64)
        lib.sizes(_p(sz))
        nc, nnzc = int(sz[0]), int(sz[1])
        crp, cci, cw = np.zeros(nc + 1, np.int64), np.zeros(nnzc, np.int32), np.zeros(nnzc)
        agg, mi = np.zeros(n, np.int32), np.zeros(n, np.int32)
        gp, gs, mp = np.zeros(nnzc + 1, np.int64), np.zeros(rp[-1], np.int64), np.zeros(nc + 1, np.int64)
        lib.fetch(_p(crp), _p(cci), _p(cw), _p(agg), _p(gp), _p(gs), _p(mp), _p(mi))
        # aggregates: of up to four rows, a partition, members in ascending order
        assert agg.min() == 0 and agg.max() == nc - 1
        cnt = np.bincount(agg, minlength=nc)
        assert cnt.min() >= 1 and cnt.max() <= 4
        assert mp[0] == 0 and mp[-1] == n and np.array_equal(np.diff(mp), cnt)
        assert np.array_equal(np.sort(mi), np.arange(n))
        assert np.array_equal(agg[mi], np.repeat(np.arange(nc), cnt))
        for I in range(0, nc, max(1, nc // 997)):
            assert np.all(np.diff(mi[mp[I]:mp[I + 1]]) > 0)
        # coarse pattern: diagonal first, unique columns, in range
        assert crp[0] == 0 and crp[-1] == nnzc and np.all(np.diff(crp) >= 1)
        crow = np.repeat(np.arange(nc), np.diff(crp))
        assert np.array_equal(cci[crp[:-1]], np.arange(nc))
        assert cci.min() >= 0 and cci.max() < nc
        assert np.unique(crow.astype(np.int64) * nc + cci).size == nnzc
        # gather list: a permutation of the fine slots grouped by coarse slot
        assert gp[0] == 0 and gp[-1] == rp[-1] and np.all(np.diff(gp) >= 1)
        assert np.array_equal(np.sort(gs), np.arange(rp[-1]))
        frow = np.repeat(np.arange(n), np.diff(rp))
        q = np.repeat(np.arange(nnzc), np.diff(gp))       # coarse slot of fine slot gs[k]
        assert np.array_equal(crow[q], agg[frow[gs]]) and np.array_equal(cci[q], agg[ci[gs]])
        # coarse weights: the Galerkin sums of the off-diagonal fine weights
        P = sp.csr_matrix((np.ones(n), (np.arange(n), agg)), shape=(n, nc))
        Wf = sp.csr_matrix((w, (frow, ci)), shape=(n, n))
        Wc = (P.T @ Wf @ P).tocsr()
        Wc.setdiag(0)
        ref = np.asarray(Wc[crow, cci]).ravel()
        assert np.allclose(cw, ref, rtol=1e-12, atol=0) and np.all(cw[crp[:-1]] == 0)
        n, rp, ci, w = nc, crp, cci, cw
        sizes.append(n)
    print("levels:", sizes)
    assert sizes[-1] <= DENSE_MAX and len(sizes) < 25


def test_colouring(lib):
    n, rp, ci, _ = _level0(lib, (60, 22, 9))
    cptr, rows = np.zeros(n + 1, np.int32), np.zeros(n, np.int32)
    ncol = lib.colors(C.c_int64(n), _p(rp), _p(ci), _p(cptr), _p(rows))
    cptr = cptr[:ncol + 1]
    assert ncol == 2, "a 7-point Cartesian graph in natural order takes two colours"
    assert cptr[0] == 0 and cptr[-1] == n and np.array_equal(np.sort(rows), np.arange(n))
    color = np.zeros(n, np.int64)
    for c in range(ncol):
        seg = rows[cptr[c]:cptr[c + 1]]
        assert np.all(np.diff(seg) > 0)
        color[seg] = c
    frow = np.repeat(np.arange(n), np.diff(rp))
    off = frow != ci
    assert np.all(color[frow[off]] != color[ci[off]])
