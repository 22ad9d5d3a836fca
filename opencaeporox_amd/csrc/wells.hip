// wells.hip — Peaceman wells in the FIM residual and Jacobian, for gfx950.
//
// The per-iteration well terms of PeacemanWellIsoT (src/WellPeaceman.cpp):
// CalTrans (:365-406), CalFlux (:409-462), CalResFIM (:1226-1265) and
// AssembleMatInjFIM / AssembleMatProdFIM (:1268-1360, :1488-1612), applied
// after the bulk residual / assembly kernels of fim.hip, in three passes:
//
//  1. well_perf: one workgroup per perforation computes its component rates
//     and, for the Jacobian, its bulk-diagonal block dt*(dQ/dXp + dQ/dXs *
//     dXs/dXp) (one entry per lane), column 0 of block (n, w), row 0 of
//     block (w, n) and its share of the well's diagonal block, into
//     per-perforation buffers.
//  2. well_apply: one workgroup per perforated cell adds its perforations'
//     rates and blocks to the bulk rows in perforation order, so a cell two
//     wells share is updated in the reference's sequence, without atomics.
//  3. well_rows: one workgroup per well sums its perforations in order into
//     the well row (residual, rates, diagonal block).
#include <algorithm>

#include "ocpx_dev.hpp"
#include "ocpx_kernels.hpp"

namespace ocpx {

template <int NCB>
__global__ void __launch_bounds__(64) well_perf(CellIO c, WellIO w, double dt, int do_jac)
{
    constexpr int nc = NCB, B = nc + 1, np = 3, C2 = np * B, BB = B * B;
    __shared__ double sQ[B * C2];        // dQ/dXs of the perforation, row-major b x np*b
    __shared__ double sPB[B], sPW[B];    // column 0 of dQ/dXp (bulk) and dQ/dXp (well)
    __shared__ double sM[BB];            // dQ/dXp + dQ/dXs * dXs/dXp (no dt), rate wells
    const int lane = threadIdx.x;
    const size_t nb = c.nb;
    const int p = blockIdx.x;
    const ocpx_well& W = w.wells[w.pwell[p]];
    const bool inj = W.type == 0, rate = W.mode == 1;
    const int64_t n = w.cell[p];
    const double WI = w.WI[p], dG = w.dG[p], xp = inj ? w.xi[p] : 0.0;
    const double perfP = W.bhp + dG;                                 // CalPerfP
    double tj[np];
    bool ex[np];
    double tINJ = 0;
    UNROLL for (int j = 0; j < np; j++) {                            // CalTrans
        ex[j] = c.phaseExist[j * nb + n] != 0;
        tj[j] = 0;
        if (ex[j]) { tj[j] = WI * c.kr[j * nb + n] / c.mu[j * nb + n]; tINJ += tj[j]; }
    }
    if (lane < nc) {                                                 // CalFlux
        const int i = lane;
        double q = 0;
        if (inj) {
            const double qt = tINJ * (c.P[n] - perfP);
            q = qt * xp * W.zi[i];
        } else {
            UNROLL for (int j = 0; j < np; j++) {
                if (!ex[j]) continue;
                const double qj = tj[j] * (c.Pj[j * nb + n] - perfP);
                q += qj * c.xi[j * nb + n] * c.xij[(size_t)(j * nc + i) * nb + n];
            }
        }
        w.qperf[(size_t)p * nc + i] = q;
    }
    if (!do_jac) return;
    // dQ/dXs, one entry (row r = 1 + i, column q) per lane
    for (int e = lane; e < B * C2; e += 64) {
        const int r = e / C2, q = e % C2;
        double acc = 0;
        if (r > 0) {
            const int i = r - 1;
            if (q < np) {
                UNROLL for (int j = 0; j < np; j++) {
                    if (!ex[j]) continue;
                    const double mu = c.mu[j * nb + n];
                    const double dkr = c.dKrdS[(j * np + q) * nb + n];
                    if (inj) {
                        const double dP = c.P[n] - W.bhp - dG;
                        acc += WI * xp * W.zi[i] * dkr * dP / mu;
                    } else {
                        const double dP = c.Pj[j * nb + n] - W.bhp - dG;
                        const double xi = c.xi[j * nb + n];
                        const double xij = c.xij[(size_t)(j * nc + i) * nb + n];
                        double t = WI * dP / mu * xi * xij * dkr;
                        t += tj[j] * xi * xij * c.dPcdS[(j * np + q) * nb + n];
                        acc += t;
                    }
                }
            } else {
                const int j = (q - np) / nc, k = (q - np) % nc;
                if (ex[j]) {
                    const double mu = c.mu[j * nb + n];
                    const double mux = c.mux[(size_t)(j * nc + k) * nb + n];
                    if (inj) {
                        const double dP = c.P[n] - W.bhp - dG;
                        const double tIJ = tj[j] * xp * W.zi[i];
                        acc += -tIJ * dP / mu * mux;
                    } else {
                        const double dP = c.Pj[j * nb + n] - W.bhp - dG;
                        const double xi = c.xi[j * nb + n];
                        const double xij = c.xij[(size_t)(j * nc + i) * nb + n];
                        acc += dP * tj[j] * xij * (c.xix[(size_t)(j * nc + k) * nb + n] - xi / mu * mux);
                        if (k == i) acc += tj[j] * xi * dP;
                    }
                }
            }
        }
        sQ[e] = acc;
    }
    if (lane < B) {                                                  // column 0 of dQ/dXp
        double pb = 0, pw = 0;
        if (lane > 0) {
            const int i = lane - 1;
            UNROLL for (int j = 0; j < np; j++) {
                if (!ex[j]) continue;
                const double mu = c.mu[j * nb + n], muP = c.muP[j * nb + n];
                if (inj) {
                    const double dP = c.P[n] - W.bhp - dG;
                    const double tIJ = tj[j] * xp * W.zi[i];
                    pb += tIJ * (1 - dP * muP / mu);
                    pw += -tIJ;
                } else {
                    const double dP = c.Pj[j * nb + n] - W.bhp - dG;
                    const double xi = c.xi[j * nb + n];
                    const double xij = c.xij[(size_t)(j * nc + i) * nb + n];
                    const double tIJ = tj[j] * xi * xij;
                    pb += tIJ * (1 - dP * muP / mu) + dP * tj[j] * xij * c.xiP[j * nb + n];
                    pw += -tIJ;
                }
            }
        }
        sPB[lane] = pb;
        sPW[lane] = pw;
        w.Bnw[(size_t)p * B + lane] = pw * dt;                       // bulk - well, column 0
    }
    __syncthreads();
    const double* X = c.dXsdXp + n * C2 * B;                        // row q, column r: X[q * B + r]
    for (int e = lane; e < BB; e += 64) {
        const int r = e / B, col = e % B;
        double s = 0;
        for (int q = 0; q < C2; q++) s += sQ[r * C2 + q] * X[q * B + col];
        const double pb = col == 0 ? sPB[r] : 0.0;
        w.bdiag[(size_t)p * BB + e] = dt * s + dt * pb;              // AddDiag(n) block
        if (rate) sM[e] = 1.0 * s + 1.0 * pb;
    }
    __syncthreads();
    if (rate) {
        if (lane < B) {                                              // well - bulk, row 0
            double a = 0;
            for (int i = 0; i < nc; i++) a += W.factor[i] * sM[(i + 1) * B + lane];
            w.Cwn[(size_t)p * B + lane] = a;
        }
        if (lane == 0) {
            double b00 = 0;
            for (int i = 0; i < nc; i++) b00 += sPW[i + 1] * W.factor[i];
            w.dD[p] = b00;
        }
    } else if (lane < B) {
        w.Cwn[(size_t)p * B + lane] = 0;
        if (lane == 0) w.dD[p] = 1.0;
    }
}

// bulk rows of the perforated cells, perforations of a cell in order
template <int NCB>
__global__ void __launch_bounds__(64) well_apply(BsrIO bsr, WellIO w, double dt, int do_jac,
                                                 double* __restrict__ resAbs)
{
    constexpr int nc = NCB, B = nc + 1, BB = B * B;
    const int lane = threadIdx.x;
    const int u = blockIdx.x;
    const int64_t n = w.ucell[u];
    double* diag = bsr.val + (size_t)bsr.rowptr[n] * BB;
    for (int k = w.uptr[u]; k < w.uptr[u + 1]; k++) {
        const int p = w.uperf[k];
        if (lane < nc) {                                             // CalResFIM: resAbs holds -R
            const size_t e = (size_t)n * B + 1 + lane;
            const double out = -(-resAbs[e] + w.qperf[(size_t)p * nc + lane] * dt);
            resAbs[e] = out;
            if (bsr.rhs) bsr.rhs[e] = out;
        }
        if (do_jac)
            for (int e = lane; e < BB; e += 64) diag[e] += w.bdiag[(size_t)p * BB + e];
    }
}

// the well rows: residual, rates, diagonal block (one workgroup per well)
template <int NCB>
__global__ void __launch_bounds__(64) well_rows(WellIO w, int do_jac)
{
    constexpr int nc = NCB, B = nc + 1, BB = B * B;
    const int lane = threadIdx.x;
    const int iw = blockIdx.x;
    const ocpx_well& W = w.wells[iw];
    const bool inj = W.type == 0, rate = W.mode == 1;
    __shared__ double sqw[B];
    if (lane < nc) {
        double q = 0;
        for (int p = W.perf0; p < W.perf0 + W.nperf; p++) q += w.qperf[(size_t)p * nc + lane];
        sqw[lane] = q;
        w.qw[(size_t)iw * nc + lane] = q;
    }
    if (do_jac) {
        for (int e = lane; e < BB; e += 64) {
            const int r = e / B, col = e % B;
            double v = 0;
            if (r == col) {
                if (!rate || r > 0) v = (double)W.nperf;             // one identity per perforation
                else for (int p = W.perf0; p < W.perf0 + W.nperf; p++) v += w.dD[p];
            }
            w.Dw[(size_t)iw * BB + e] = v;
        }
    }
    __syncthreads();
    if (lane == 0) {
        double r0;
        if (rate) {
            r0 = inj ? W.target : -W.target;
            for (int i = 0; i < nc; i++) r0 += sqw[i] * W.factor[i];
        } else {
            r0 = W.bhp - W.target;
        }
        w.wres[iw] = r0;
        w.wrhs[(size_t)iw * B] = -r0;
    } else if (lane < B) {
        w.wrhs[(size_t)iw * B + lane] = 0.0;
    }
}

// y[bulk] += (n, w) blocks x[well] (workgroups 0 .. nucell-1, one per
// perforated cell, perforations in order); y[well] = Dw x[well] + (w, n)
// blocks x[bulk] (workgroups nucell .., one per well).  The bulk part of y
// comes from bsr_spmv before this kernel runs.
__global__ void __launch_bounds__(64) well_spmv(WellIO w, int B, int64_t nb, const double* __restrict__ x,
                                                double* __restrict__ y)
{
    __shared__ double part[64];
    const int lane = threadIdx.x;
    const int g = blockIdx.x;
    if (g < w.nucell) {                                              // uniform per workgroup
        const int64_t n = w.ucell[g];
        if (lane < B) {
            double acc = y[(size_t)n * B + lane];
            for (int k = w.uptr[g]; k < w.uptr[g + 1]; k++) {
                const int p = w.uperf[k];
                acc += w.Bnw[(size_t)p * B + lane] * x[(size_t)(nb + w.pwell[p]) * B];
            }
            y[(size_t)n * B + lane] = acc;
        }
        return;
    }
    const int iw = g - w.nucell;
    const ocpx_well& W = w.wells[iw];
    const double* xw = x + (size_t)(nb + iw) * B;
    // row 0's coupling to the perforated cells, perforations split over lanes
    double c0 = 0;
    for (int p = W.perf0 + lane; p < W.perf0 + W.nperf; p += 64) {
        const int64_t n = w.cell[p];
        for (int col = 0; col < B; col++) c0 += w.Cwn[(size_t)p * B + col] * x[(size_t)n * B + col];
    }
    part[lane] = c0;
    __syncthreads();
    for (int s = 32; s > 0; s >>= 1) {
        if (lane < s) part[lane] += part[lane + s];
        __syncthreads();
    }
    if (lane < B) {
        double acc = 0;
        for (int col = 0; col < B; col++) acc += w.Dw[((size_t)iw * B + lane) * B + col] * xw[col];
        if (lane == 0) acc += part[0];
        y[(size_t)(nb + iw) * B + lane] = acc;
    }
}

// PeacemanWellIsoT::GetSolutionFIM (src/WellPeaceman.cpp:1198-1204): bhp += u of the well row
__global__ void __launch_bounds__(64) well_bhp_update(ocpx_well* wells, int32_t nwell, int B, int64_t nb,
                                                      const double* __restrict__ u)
{
    const int iw = threadIdx.x;
    if (iw < nwell) wells[iw].bhp += u[(size_t)(nb + iw) * B];
}

hipError_t run_well_spmv(const WellIO& w, int B, int64_t nb, const double* x, double* y, hipStream_t st)
{
    if (w.nwell == 0) return hipSuccess;
    hipLaunchKernelGGL(well_spmv, dim3(w.nucell + w.nwell), dim3(64), 0, st, w, B, nb, x, y);
    return hipGetLastError();
}

hipError_t run_well_bhp_update(ocpx_well* wells, int32_t nwell, int B, int64_t nb, const double* u, hipStream_t st)
{
    for (int32_t w0 = 0; w0 < nwell; w0 += 64) {
        hipLaunchKernelGGL(well_bhp_update, dim3(1), dim3(64), 0, st, wells + w0, std::min<int32_t>(64, nwell - w0),
                           B, nb + w0, u);
    }
    return hipGetLastError();
}

template <int NCB>
static hipError_t launch_wells(const CellIO& c, const BsrIO& bsr, const WellIO& w, double dt, int dj, double* resAbs,
                               hipStream_t st)
{
    hipLaunchKernelGGL(well_perf<NCB>, dim3(w.nperf), dim3(64), 0, st, c, w, dt, dj);
    hipLaunchKernelGGL(well_apply<NCB>, dim3(w.nucell), dim3(64), 0, st, bsr, w, dt, dj, resAbs);
    hipLaunchKernelGGL(well_rows<NCB>, dim3(w.nwell), dim3(64), 0, st, w, dj);
    return hipGetLastError();
}

#define OCPX_NC_SWITCH(X)                                                                           \
    switch (nc) {                                                                                   \
    X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13)                                      \
    default: return hipErrorInvalidValue;                                                           \
    }

hipError_t run_well_terms(int nc, const CellIO& c, const BsrIO& bsr, const WellIO& w, double dt, bool do_jac,
                          double* resAbs, hipStream_t st)
{
    if (w.nwell == 0) return hipSuccess;
#define X(N) case N: return launch_wells<N>(c, bsr, w, dt, do_jac ? 1 : 0, resAbs, st);
    OCPX_NC_SWITCH(X)
#undef X
}
#undef OCPX_NC_SWITCH

}  // namespace ocpx
