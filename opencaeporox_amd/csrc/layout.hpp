// layout.hpp — the block-CSR pattern and the assembly row order, plain C++
// (no HIP), shared by the context set-up and the CPU checks of the kernels'
// indexing.
#pragma once
#include <algorithm>
#include <cstdint>
#include <vector>

namespace ocpx {

// OCPMatrix push order (src/OCPMatrix.cpp:92-137): per row the diagonal
// first, then, connection by connection in index order, row e gets (e, b)
// and row b gets (b, e).  rp [nb + 1], col / ts / nbc [nb + 2 nconn]: block
// column, slot of the transposed block, connection of the slot (-1 on the
// diagonal).  Connections satisfy 0 <= b < e < nb.
inline void bsr_layout(int64_t nb, int64_t nconn, const int32_t* cb, const int32_t* ce, int64_t* rp, int32_t* col,
                       int32_t* ts, int32_t* nbc)
{
    std::vector<int64_t> cnt(nb, 1);
    for (int64_t k = 0; k < nconn; k++) { cnt[cb[k]]++; cnt[ce[k]]++; }
    rp[0] = 0;
    for (int64_t n = 0; n < nb; n++) rp[n + 1] = rp[n] + cnt[n];
    std::vector<int64_t> fill(nb);
    for (int64_t n = 0; n < nb; n++) {
        col[rp[n]] = (int32_t)n; ts[rp[n]] = (int32_t)rp[n]; nbc[rp[n]] = -1; fill[n] = rp[n] + 1;
    }
    for (int64_t k = 0; k < nconn; k++) {
        const int32_t b = cb[k], e = ce[k];
        const int64_t se = fill[e]++, sb = fill[b]++;
        col[se] = b; nbc[se] = (int32_t)k;
        col[sb] = e; nbc[sb] = (int32_t)k;
        ts[se] = (int32_t)sb; ts[sb] = (int32_t)se;
    }
}

// Row order for the assembly.  When the connections are those of a natural
// Cartesian numbering (every e - b is 1, nx or nx*ny, x and y connections do
// not wrap), rows go in pencils: 16 x 16 tiles of the xy plane, each swept
// through z, so the z-neighbours a row reads were touched a plane earlier
// and are still in the XCD's L2.  Otherwise the natural order is kept.
inline std::vector<int32_t> pencil_order(int64_t nb, int64_t nconn, const int32_t* cb, const int32_t* ce)
{
    int64_t d[3] = {0, 0, 0};
    int nd = 0;
    for (int64_t k = 0; k < nconn; k++) {
        const int64_t v = (int64_t)ce[k] - cb[k];
        bool seen = false;
        for (int q = 0; q < nd; q++) seen = seen || d[q] == v;
        if (seen) continue;
        if (nd == 3 || v <= 0) return {};
        d[nd++] = v;
    }
    if (nd != 3) return {};
    std::sort(d, d + 3);
    const int64_t nx = d[1], nxy = d[2];
    if (d[0] != 1 || nxy % nx != 0 || nb % nxy != 0) return {};
    const int64_t ny = nxy / nx, nz = nb / nxy;
    if (nz < 2 || ny < 2 || nx < 2) return {};
    for (int64_t k = 0; k < nconn; k++) {
        const int64_t b = cb[k], e = ce[k];
        if (e - b == 1 && b % nx == nx - 1) return {};
        if (e - b == nx && (b / nx) % ny == ny - 1) return {};
    }
    const int64_t TX = 16, TY = 16;
    std::vector<int32_t> ord;
    ord.reserve(nb);
    for (int64_t y0 = 0; y0 < ny; y0 += TY)
        for (int64_t x0 = 0; x0 < nx; x0 += TX)
            for (int64_t z = 0; z < nz; z++)
                for (int64_t y = y0; y < std::min(ny, y0 + TY); y++)
                    for (int64_t x = x0; x < std::min(nx, x0 + TX); x++) ord.push_back((int32_t)(z * nxy + y * nx + x));
    return ord;
}


// Perforations of the wells (well w owns [perf0[w], perf0[w] + nperf[w])):
// pwell [nperf] the well of each perforation; the perforated cells ucell in
// ascending order, their perforations uperf[uptr[u] .. uptr[u + 1]) in
// perforation order (the order the reference adds them to a shared cell).
inline void well_layout(int32_t nwell, const int32_t* perf0, const int32_t* nperf, int64_t np_,
                        const int64_t* perf_cell, std::vector<int32_t>& pwell, std::vector<int64_t>& ucell,
                        std::vector<int32_t>& uptr, std::vector<int32_t>& uperf)
{
    pwell.assign(np_, 0);
    for (int32_t w = 0; w < nwell; w++)
        for (int32_t p = perf0[w]; p < perf0[w] + nperf[w]; p++) pwell[p] = w;
    std::vector<std::pair<int64_t, int32_t>> byc(np_);
    for (int64_t p = 0; p < np_; p++) byc[p] = {perf_cell[p], (int32_t)p};
    std::sort(byc.begin(), byc.end());
    ucell.clear();
    uptr.clear();
    uperf.assign(np_, 0);
    for (int64_t k = 0; k < np_; k++) {
        if (k == 0 || byc[k].first != byc[k - 1].first) { ucell.push_back(byc[k].first); uptr.push_back((int32_t)k); }
        uperf[k] = byc[k].second;
    }
    uptr.push_back((int32_t)np_);
}

}  // namespace ocpx
