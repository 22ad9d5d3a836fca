// amg_layout.hpp — host-side set-up of the pressure AMG's hierarchy and of
// the Gauss-Seidel colouring (plain C++, no device code): the index arrays
// cpr.hip's kernels read.  Used by cpr.hip and ocpx_host.cpp, and compiled
// for the host by the tests that check every index on the CPU.
//
// Graphs are CSR with the diagonal first in each row and a weight per slot
// (0 on the diagonal).  Level 0 is the bulk block pattern weighted by the
// connections' transmissibilities; a coarse level is the Galerkin graph of
// piecewise-constant aggregation (summed off-diagonal weights).
#pragma once
#include <algorithm>
#include <cstdint>
#include <vector>

namespace ocpx {
namespace amg {

// pairing candidates: w_ij >= STRONG max_k w_ik (0: the strongest unmatched
// neighbour, whatever its weight)
constexpr double STRONG = 0.0;

struct Graph {
    int64_t n = 0;
    std::vector<int64_t> rp;
    std::vector<int32_t> ci;
    std::vector<double> w;
};

// Pairwise matching in row order: an unmatched row pairs with its unmatched
// neighbour of largest (positive) weight among those within STRONG of the
// row's largest; rows without one stay alone.  Returns the number of groups.
inline int64_t pair_rows(const Graph& g, std::vector<int32_t>& map)
{
    map.assign(g.n, -1);
    int64_t m = 0;
    for (int64_t i = 0; i < g.n; i++) {
        if (map[i] >= 0) continue;
        double wmax = 0;
        for (int64_t s = g.rp[i]; s < g.rp[i + 1]; s++)
            if (g.ci[s] != i) wmax = std::max(wmax, g.w[s]);
        int64_t best = -1;
        double bw = 0;
        for (int64_t s = g.rp[i]; s < g.rp[i + 1]; s++) {
            const int64_t j = g.ci[s];
            if (j == i || map[j] >= 0) continue;
            if (g.w[s] >= STRONG * wmax && g.w[s] > bw) { best = j; bw = g.w[s]; }
        }
        map[i] = (int32_t)m;
        if (best >= 0) map[best] = (int32_t)m;
        m++;
    }
    return m;
}

// Galerkin graph of the aggregation map (m groups): the coarse pattern
// (diagonal first), summed off-diagonal weights, the coarse slot of every
// fine slot, and the members of every group (mi[mp[I] .. mp[I+1]) in
// ascending order).
inline void coarse_graph(const Graph& f, const std::vector<int32_t>& map, int64_t m, Graph& c,
                         std::vector<int64_t>& slot_of, std::vector<int64_t>& mp, std::vector<int32_t>& mi)
{
    mp.assign(m + 1, 0);
    for (int64_t i = 0; i < f.n; i++) mp[map[i] + 1]++;
    for (int64_t I = 0; I < m; I++) mp[I + 1] += mp[I];
    mi.assign(f.n, 0);
    {
        std::vector<int64_t> pos(mp.begin(), mp.end() - 1);
        for (int64_t i = 0; i < f.n; i++) mi[pos[map[i]]++] = (int32_t)i;
    }
    c.n = m;
    c.rp.assign(1, 0);
    c.ci.clear();
    c.w.clear();
    slot_of.assign(f.ci.size(), -1);
    std::vector<int64_t> marker(m, -1);
    for (int64_t I = 0; I < m; I++) {
        const int64_t start = (int64_t)c.ci.size();
        marker[I] = start;
        c.ci.push_back((int32_t)I);
        c.w.push_back(0);
        for (int64_t k = mp[I]; k < mp[I + 1]; k++) {
            const int64_t i = mi[k];
            for (int64_t s = f.rp[i]; s < f.rp[i + 1]; s++) {
                const int64_t J = map[f.ci[s]];
                if (marker[J] < start) {
                    marker[J] = (int64_t)c.ci.size();
                    c.ci.push_back((int32_t)J);
                    c.w.push_back(0);
                }
                const int64_t q = marker[J];
                if (J != I) c.w[q] += f.w[s];
                slot_of[s] = q;
            }
        }
        c.rp.push_back((int64_t)c.ci.size());
    }
}

// What one level hands to the next: agg [n] (coarse row of each row), the
// gather list (coarse slot s sums the fine slots gs[gp[s] .. gp[s+1])) and
// the members of each coarse row (mi[mp[I] .. mp[I+1])).
struct LevelMaps {
    std::vector<int32_t> agg, mi;
    std::vector<int64_t> gp, gs, mp;
};

// Two rounds of pairing (aggregates of up to four rows) and the coarse
// graph.  False, with nothing built, when that would keep more than 9/10 of
// the rows (no real coarsening left).
inline bool coarsen(const Graph& g, Graph& gc, LevelMaps& L)
{
    std::vector<int32_t> m1, m2;
    const int64_t n1 = pair_rows(g, m1);
    Graph g1;
    std::vector<int64_t> so;
    coarse_graph(g, m1, n1, g1, so, L.mp, L.mi);
    const int64_t n2 = pair_rows(g1, m2);
    if (n2 > (g.n * 9) / 10) return false;
    L.agg.resize(g.n);
    for (int64_t i = 0; i < g.n; i++) L.agg[i] = m2[m1[i]];
    coarse_graph(g, L.agg, n2, gc, so, L.mp, L.mi);
    L.gp.assign(gc.ci.size() + 1, 0);
    L.gs.assign(g.ci.size(), 0);
    for (size_t s = 0; s < so.size(); s++) L.gp[so[s] + 1]++;
    for (size_t q = 0; q + 1 < L.gp.size(); q++) L.gp[q + 1] += L.gp[q];
    std::vector<int64_t> pos(L.gp.begin(), L.gp.end() - 1);
    for (size_t s = 0; s < so.size(); s++) L.gs[pos[so[s]]++] = (int64_t)s;
    return true;
}

// Greedy colouring of a block graph (rowptr, col) in row order: colour
// lists rows[cptr[c] .. cptr[c+1]) in ascending row order; no two rows of one
// colour are neighbours.
inline void color_rows(int64_t n, const std::vector<int64_t>& rp, const std::vector<int32_t>& col,
                       std::vector<int32_t>& cptr, std::vector<int32_t>& rows)
{
    std::vector<int32_t> color(n, -1);
    std::vector<int64_t> seen;               // colour -> last row that found it taken
    int ncol = 0;
    for (int64_t i = 0; i < n; i++) {
        for (int64_t s = rp[i]; s < rp[i + 1]; s++) {
            const int32_t c = color[col[s]];
            if (c >= 0) seen[c] = i;
        }
        int c = 0;
        while (c < ncol && seen[c] == i) c++;
        if (c == ncol) { seen.push_back(-1); ncol++; }
        color[i] = c;
    }
    cptr.assign(ncol + 1, 0);
    for (int64_t i = 0; i < n; i++) cptr[color[i] + 1]++;
    for (int c = 0; c < ncol; c++) cptr[c + 1] += cptr[c];
    rows.assign(n, 0);
    std::vector<int32_t> pos(cptr.begin(), cptr.end() - 1);
    for (int64_t i = 0; i < n; i++) rows[pos[color[i]]++] = (int32_t)i;
}

}  // namespace amg
}  // namespace ocpx
