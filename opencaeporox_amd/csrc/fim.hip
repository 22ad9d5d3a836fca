// fim.hip — kr/Pc + rock, residual + Jacobian assembly, Newton update and
// the adjoint products of the FIM system, for gfx950.
//
// Assembly (IsoT_FIM::CalRes + AssembleMatBulks, src/IsoThermalMethod.cpp:
// 1090-1269, with OCPConvection01::CalFlux/AssembleMatFIM,
// src/OCPConvection.cpp:22-187, and BulkAccumuTerm01, src/BulkAccumuModule.cpp:
// 24-70) is row-owner and atomic-free: each block row of the BSR is produced
// by b = nc+1 lanes, lane r owning column r of every block of the row.  The
// row's slots come in OCPMatrix push order (ascending connection index), so
// the diagonal accumulates its terms in the reference's order.  A flux block
// dt*(dF/dXp + dF/dXs * dXs/dXp) is never formed as a GEMM: per phase, dF/dXs
// is an outer product x_up (x) q plus a scaled identity, so a column costs
// O(np*b) instead of O(np*b^2).  Each connection's two blocks are computed by
// both rows that need them; the doubled arithmetic is cheaper than the HBM
// round trip of a connection-ordered scatter.
#include "ocpx_dev.hpp"
#include "ocpx_kernels.hpp"

namespace ocpx {

constexpr double GRAVF = 0.0000980665;   // METRIC gravity factor

// ------------------------------------------------------------------ kr / Pc
// OCPTable::Eval_All with slopes (src/OCPTable.cpp:26-63) on a 4-column table
OCPX_HD void tab_eval(const double* tab, int nrow, double v, double* out, double* sl)
{
    if (v >= tab[0]) {
        for (int i = 1; i < nrow; i++) {
            if (v < tab[i * 4]) {
                const int k = i - 1;
                const double dx = tab[(k + 1) * 4] - tab[k * 4];
                UNROLL for (int c = 0; c < 4; c++) {
                    sl[c] = (tab[(k + 1) * 4 + c] - tab[k * 4 + c]) / dx;
                    out[c] = tab[k * 4 + c] + sl[c] * (v - tab[k * 4]);
                }
                return;
            }
        }
        UNROLL for (int c = 0; c < 4; c++) { sl[c] = 0; out[c] = tab[(nrow - 1) * 4 + c]; }
    } else {
        UNROLL for (int c = 0; c < 4; c++) { sl[c] = 0; out[c] = tab[c]; }
    }
}

// IsoT_FIM::CalKrPc (:1073-1088) via OCPFlowMethod_OGW01::CalKrPcDer
// (src/OCPFlowMethod.cpp:63-73) and Stone II (src/OCPFuncSAT.cpp:162-186);
// IsothermalMethod::CalRock (:19-30) via OCPRockIsoT_Linear (src/OCPRock.cpp:19-24)
__global__ void __launch_bounds__(256) krpc_rock(const DevSat* __restrict__ sp, double Pref, double cp1,
                                                 double cp2, CellIO c)
{
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= c.nb) return;
    const DevSat& s = *sp;
    const size_t nb = c.nb;
    const double Sg = c.S[nb + n], Sw = c.S[2 * nb + n];
    double d[4], sl[4];
    double kr[3], pc[3], dk[9], dp[9];
    UNROLL for (int k = 0; k < 9; k++) { dk[k] = 0; dp[k] = 0; }
    tab_eval(s.swof, s.nswof, Sw, d, sl);
    kr[2] = d[1];
    const double krow = d[2];
    pc[2] = -d[3];
    dk[8] = sl[1];
    const double dkrowdSw = sl[2];
    dp[8] = -sl[3];
    tab_eval(s.sgof, s.nsgof, Sg, d, sl);
    kr[1] = d[1];
    const double krog = d[2];
    pc[1] = d[3];
    dk[4] = sl[1];
    const double dkrogdSg = sl[2];
    dp[4] = sl[3];
    pc[0] = 0;
    const double krocw = s.swof[2];
    kr[0] = krocw * ((krow / krocw + kr[2]) * (krog / krocw + kr[1]) - (kr[2] + kr[1]));
    if (kr[0] < 0) {
        kr[0] = 0;
    } else {
        dk[0] = 0;   // dKrowdSo = dKrogdSo = 0 in OGW01
        dk[2] = krocw * ((dkrowdSw / krocw + dk[8]) * (krog / krocw + kr[1]) - dk[8]);
        dk[1] = krocw * ((krow / krocw + kr[2]) * (dkrogdSg / krocw + dk[4]) - dk[4]);
    }
    const double P = c.P[n];
    UNROLL for (int j = 0; j < 3; j++) {
        c.kr[j * nb + n] = kr[j];
        c.Pc[j * nb + n] = pc[j];
        c.Pj[j * nb + n] = P + pc[j];
    }
    UNROLL for (int k = 0; k < 9; k++) { c.dKrdS[k * nb + n] = dk[k]; c.dPcdS[k * nb + n] = dp[k]; }
    const double dP = P - Pref;
    const double poro = c.poroInit[n] * (1 + (cp1 + cp2 / 2 * dP) * dP);
    c.poro[n] = poro;
    c.poroP[n] = c.poroInit[n] * (cp1 + cp2 * dP);
    c.rockVp[n] = c.v[n] * poro;
}

// --------------------------------------------------------- residual + Jacobian
// Upwinding of one phase (OCPConvection01::CalFlux :37-78).  Returns false
// when the phase carries no flux.
struct Upw {
    double dP;
    bool upIsB, downExists;
};

OCPX_HD bool upwind(const CellIO& c, size_t nb, int64_t b, int64_t e, int j, Upw& u)
{
    const bool exb = c.phaseExist[j * nb + b] != 0, exe = c.phaseExist[j * nb + e] != 0;
    if (!exb && !exe) return false;
    double rho;
    if (exb && exe) rho = (c.rho[j * nb + b] + c.rho[j * nb + e]) / 2;
    else if (exb) rho = c.rho[j * nb + b];
    else rho = c.rho[j * nb + e];
    u.dP = (c.Pj[j * nb + b] - GRAVF * rho * c.depth[b]) - (c.Pj[j * nb + e] - GRAVF * rho * c.depth[e]);
    u.upIsB = !(u.dP < 0);
    const bool exu = u.upIsB ? exb : exe;
    u.downExists = u.upIsB ? exe : exb;
    return exu;
}

template <int NCB>
__global__ void __launch_bounds__(64) res_asm(CellIO c, ConnIO cn, BsrIO bsr, double dt, int do_res,
                                              int do_jac, double* __restrict__ resAbs,
                                              double* __restrict__ bmax, int64_t* __restrict__ barg)
{
    constexpr int nc = NCB, B = nc + 1, np = 3, G = 64 / B, XR = np * B;   // XR: dXsdXp rows
    const int lane = threadIdx.x;
    const int r = lane / G, g = lane % G;
    const size_t nb = c.nb;
    const int64_t n = (int64_t)blockIdx.x * G + g;
    const bool valid = (r < B) && (n < c.nb);
    __shared__ double sV[64], sN[64];
    __shared__ int64_t iV[64], iN[64];

    double res = 0, absV = -1, absN = -1;
    if (valid) {
        double Xo[XR];                     // column r of own dXsdXp
        if (do_jac) UNROLL for (int k = 0; k < XR; k++) Xo[k] = c.dXsdXp[(size_t)(k * B + r) * nb + n];
        // accumulation (BulkAccumuTerm01::CalResFIM / CaldFdXpFIM)
        double diag[B];
        if (do_jac) {
            UNROLL for (int i = 0; i < B; i++) diag[i] = (i == r && i > 0) ? 1.0 : 0.0;
            diag[0] = r == 0 ? c.v[n] * c.poroP[n] - c.vfP[n] : -c.vfi[(size_t)(r - 1) * nb + n];
        }
        if (do_res) res = r == 0 ? c.rockVp[n] - c.vf[n] : c.Ni[(size_t)(r - 1) * nb + n] - c.lNi[(size_t)(r - 1) * nb + n];

        const int64_t s0 = bsr.rowptr[n], s1 = bsr.rowptr[n + 1];
        for (int64_t s = s0 + 1; s < s1; s++) {
            const int cc = bsr.nbr_conn[s];
            const int64_t bI = cn.cb[cc], eI = cn.ce[cc];
            const bool ownB = (bI == n);
            const int64_t m = ownB ? eI : bI;
            const double T = cn.trans[cc];
            double Xm[XR];
            if (do_jac) UNROLL for (int k = 0; k < XR; k++) Xm[k] = c.dXsdXp[(size_t)(k * B + r) * nb + m];
            double colB[B], colE[B];
            UNROLL for (int i = 0; i < B; i++) { colB[i] = 0; colE[i] = 0; }
            double flux = 0;
            const double dGamma = GRAVF * (c.depth[bI] - c.depth[eI]);
            UNROLL for (int j = 0; j < np; j++) {
                Upw u;
                if (!upwind(c, nb, bI, eI, j, u)) continue;
                const int64_t U = u.upIsB ? bI : eI, D = u.upIsB ? eI : bI;
                const double xi = c.xi[j * nb + U], kr = c.kr[j * nb + U], mu = c.mu[j * nb + U];
                const double tJ = T * kr / mu;
                double xu[nc];
                UNROLL for (int i = 0; i < nc; i++) xu[i] = c.xij[(size_t)(j * nc + i) * nb + U];
                if (do_res && r > 0) flux += tJ * u.dP * xi * c.xij[(size_t)(j * nc + r - 1) * nb + U];
                if (!do_jac) continue;
                const double muP = c.muP[j * nb + U], xiP = c.xiP[j * nb + U];
                const double wU = u.downExists ? 0.5 : 1.0, wD = u.downExists ? 0.5 : 0.0;
                const double xt = xi * tJ, dP = u.dP;
                // pressure-column coefficients (dFdXp col 0)
                const double pU = tJ * xiP * dP - xt * muP / mu * dP - xt * wU * c.rhoP[j * nb + U] * dGamma;
                const double pD = -xt * wD * c.rhoP[j * nb + D] * dGamma;
                // q = coefficient vector dotted with the X column, per side
                double qB = 0, qE = 0;
                if (r == 0) {
                    qB = xt + (u.upIsB ? pU : pD);
                    qE = -xt + (u.upIsB ? pD : pU);
                }
                // the two register arrays are chosen element-wise with
                // selects, never through a pointer, so both stay in VGPRs
                const bool upB = u.upIsB;
                const double kcoef = T / mu * xi * dP;
                UNROLL for (int k = 0; k < np; k++) {
                    const double xb = ownB ? Xo[k] : Xm[k], xe = ownB ? Xm[k] : Xo[k];
                    const double cb = xt * c.dPcdS[(j * np + k) * nb + bI] + (upB ? kcoef * c.dKrdS[(j * np + k) * nb + U] : 0.0);
                    const double ce = -xt * c.dPcdS[(j * np + k) * nb + eI] + (upB ? 0.0 : kcoef * c.dKrdS[(j * np + k) * nb + U]);
                    qB += cb * xb;
                    qE += ce * xe;
                }
                // upstream is the own cell when (own is b) == (up is b)
                const bool upOwn = (ownB == upB);
                double qU = 0, qD = 0;
                UNROLL for (int mm = 0; mm < nc; mm++) {
                    const int k = np + j * nc + mm;
                    const double xU = upOwn ? Xo[k] : Xm[k], xD = upOwn ? Xm[k] : Xo[k];
                    const double eu = -xt * wU * c.rhox[(size_t)(j * nc + mm) * nb + U] * dGamma +
                                      tJ * c.xix[(size_t)(j * nc + mm) * nb + U] * dP -
                                      xt * c.mux[(size_t)(j * nc + mm) * nb + U] / mu * dP;
                    const double ed = -xt * wD * c.rhox[(size_t)(j * nc + mm) * nb + D] * dGamma;
                    qU += eu * xU;
                    qD += ed * xD;
                }
                if (upB) { qB += qU; qE += qD; } else { qE += qU; qB += qD; }
                const double sI = tJ * xi * dP;   // k == i term, upstream side
                UNROLL for (int i = 0; i < nc; i++) {
                    const int k = np + j * nc + i;
                    const double xU = upOwn ? Xo[k] : Xm[k];
                    colB[i + 1] += xu[i] * qB + (upB ? sI * xU : 0.0);
                    colE[i + 1] += xu[i] * qE + (upB ? 0.0 : sI * xU);
                }
            }
            if (do_res && r > 0) res += ownB ? dt * flux : -dt * flux;
            if (do_jac) {
                double* off = bsr.val + (size_t)s * B * B + r;
                if (ownB) {                       // AddDiag(b, B); NewOffDiag(b, e, E)
                    UNROLL for (int i = 0; i < B; i++) { diag[i] += dt * colB[i]; off[i * B] = dt * colE[i]; }
                } else {                          // NewOffDiag(e, b, -B); AddDiag(e, -E)
                    UNROLL for (int i = 0; i < B; i++) { off[i * B] = -(dt * colB[i]); diag[i] += -(dt * colE[i]); }
                }
            }
        }
        if (do_jac) {
            double* dg = bsr.val + (size_t)s0 * B * B + r;
            UNROLL for (int i = 0; i < B; i++) dg[i * B] = diag[i];
        }
        if (do_res) {
            const double out = -res;             // Dscalar(-1) :1170
            resAbs[(size_t)n * B + r] = out;
            if (bsr.rhs) bsr.rhs[(size_t)n * B + r] = out;   // CopyRhs
            absV = fabs(res / c.rockVp[n]);
            if (r > 0) absN = fabs(res / c.Nt[n]);
        }
    }
    if (!do_res) return;
    // block reduction of max |R/Vp| and max |R_i/Nt| (first cell on ties)
    sV[lane] = absV; iV[lane] = valid ? n : INT64_MAX;
    sN[lane] = absN; iN[lane] = valid ? n : INT64_MAX;
    __syncthreads();
    for (int w = 32; w > 0; w >>= 1) {
        if (lane < w) {
            const int o = lane + w;
            if (sV[o] > sV[lane] || (sV[o] == sV[lane] && iV[o] < iV[lane])) { sV[lane] = sV[o]; iV[lane] = iV[o]; }
            if (sN[o] > sN[lane] || (sN[o] == sN[lane] && iN[o] < iN[lane])) { sN[lane] = sN[o]; iN[lane] = iN[o]; }
        }
        __syncthreads();
    }
    if (lane == 0) {
        bmax[2 * blockIdx.x] = sV[0]; barg[2 * blockIdx.x] = iV[0];
        bmax[2 * blockIdx.x + 1] = sN[0]; barg[2 * blockIdx.x + 1] = iN[0];
    }
}

__global__ void __launch_bounds__(256) norms_final(int64_t nblk, const double* __restrict__ bmax,
                                                   const int64_t* __restrict__ barg, double* out4)
{
    __shared__ double sv[256], sn[256];
    __shared__ int64_t iv[256], in_[256];
    const int t = threadIdx.x;
    double v = -1, w = -1;
    int64_t a = INT64_MAX, b = INT64_MAX;
    for (int64_t k = t; k < nblk; k += 256) {
        const double v2 = bmax[2 * k], w2 = bmax[2 * k + 1];
        const int64_t a2 = barg[2 * k], b2 = barg[2 * k + 1];
        if (v2 > v || (v2 == v && a2 < a)) { v = v2; a = a2; }
        if (w2 > w || (w2 == w && b2 < b)) { w = w2; b = b2; }
    }
    sv[t] = v; iv[t] = a; sn[t] = w; in_[t] = b;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) {
            if (sv[t + s] > sv[t] || (sv[t + s] == sv[t] && iv[t + s] < iv[t])) { sv[t] = sv[t + s]; iv[t] = iv[t + s]; }
            if (sn[t + s] > sn[t] || (sn[t + s] == sn[t] && in_[t + s] < in_[t])) { sn[t] = sn[t + s]; in_[t] = in_[t + s]; }
        }
        __syncthreads();
    }
    if (t == 0) {
        // the reference starts from max = 0, id = 0 and replaces on strict >
        out4[0] = sv[0] > 0 ? sv[0] : 0.0;
        out4[1] = sv[0] > 0 ? (double)iv[0] : 0.0;
        out4[2] = sn[0] > 0 ? sn[0] : 0.0;
        out4[3] = sn[0] > 0 ? (double)in_[0] : 0.0;
    }
}

// ------------------------------------------------- Newton update (GetSolution)
// src/IsoThermalMethod.cpp:1346-1400: dXs = dXsdXp * u, saturation chop,
// then S, xij, P, Ni updates.
template <int NCB>
__global__ void __launch_bounds__(256) get_solution(double* P, double* Ni, double* S, double* xij,
                                                    const double* __restrict__ X, const double* __restrict__ u,
                                                    double dSmax, int64_t nb)
{
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nb) return;
    constexpr int nc = NCB, col = nc + 1, np = 3, row = np * col;
    double uu[col];
    UNROLL for (int j = 0; j < col; j++) uu[j] = u[(size_t)n * col + j];
    double d[row];
    UNROLL for (int i = 0; i < row; i++) {
        double s = 0;
        UNROLL for (int j = 0; j < col; j++) s += 1.0 * X[(size_t)(i * col + j) * nb + n] * uu[j];
        d[i] = s;
    }
    double chop = 1;
    UNROLL for (int j = 0; j < np; j++) {
        double ct = 1;
        const double Sj = S[(size_t)j * nb + n];
        if (fabs(d[j]) > dSmax) ct = dSmax / fabs(d[j]);
        else if (Sj + d[j] < 0.0) ct = 0.9 * Sj / fabs(d[j]);
        chop = fmin(chop, ct);
    }
    UNROLL for (int j = 0; j < np; j++) S[(size_t)j * nb + n] += chop * d[j];
    UNROLL for (int j = 0; j < np; j++)
        UNROLL for (int i = 0; i < nc; i++) xij[(size_t)(j * nc + i) * nb + n] += chop * d[np + j * nc + i];
    P[n] += uu[0];
    UNROLL for (int i = 0; i < nc; i++) Ni[(size_t)i * nb + n] += chop * uu[1 + i];
}

// ------------------------------------------------------------ adjoint pieces
// out = J^T lambda: output block n gathers A(m, n)^T lambda_m over the row's
// slots through the transposed-slot map; lane r owns output entry r.
template <int NCB>
__global__ void __launch_bounds__(64) jt_lambda(BsrIO bsr, int64_t nb, const double* __restrict__ lam,
                                                double* __restrict__ out)
{
    constexpr int B = NCB + 1, G = 64 / B;
    const int lane = threadIdx.x, r = lane / G, g = lane % G;
    const int64_t n = (int64_t)blockIdx.x * G + g;
    if (r >= B || n >= nb) return;
    double acc = 0;
    for (int64_t s = bsr.rowptr[n]; s < bsr.rowptr[n + 1]; s++) {
        const int64_t m = bsr.colind[s];
        const double* blk = bsr.val + (size_t)bsr.tslot[s] * B * B;
        UNROLL for (int i = 0; i < B; i++) acc += blk[i * B + r] * lam[(size_t)m * B + i];
    }
    out[(size_t)n * B + r] = acc;
}

// d(lambda . R)/d trans_c = dt * sum_i (lambda_b - lambda_e)[1+i] * F_i / T_c
// and d(lambda . R)/d Vp_n = lambda_n[0] (R_0 = Vp - Vf)
template <int NCB>
__global__ void __launch_bounds__(256) dR_dtrans(CellIO c, ConnIO cn, double dt, const double* __restrict__ lam,
                                                 double* __restrict__ dT, double* __restrict__ dVp)
{
    constexpr int nc = NCB, B = nc + 1;
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nb = c.nb;
    if (k < c.nb && dVp) dVp[k] = lam[(size_t)k * B];
    if (k >= cn.nconn) return;
    const int64_t b = cn.cb[k], e = cn.ce[k];
    double g = 0;
    UNROLL for (int j = 0; j < 3; j++) {
        Upw u;
        if (!upwind(c, nb, b, e, j, u)) continue;
        const int64_t U = u.upIsB ? b : e;
        const double v = c.kr[j * nb + U] / c.mu[j * nb + U] * u.dP * c.xi[j * nb + U];
        UNROLL for (int i = 0; i < nc; i++)
            g += v * c.xij[(size_t)(j * nc + i) * nb + U] * (lam[(size_t)b * B + 1 + i] - lam[(size_t)e * B + 1 + i]);
    }
    dT[k] = dt * g;
}

__global__ void fill_u32(unsigned* p, unsigned v, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// --------------------------------------------------------------- launchers
hipError_t run_fill_u32(unsigned* p, unsigned v, int64_t n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

hipError_t run_krpc_rock(const DevSat* s, double Pref, double cp1, double cp2, const CellIO& c, hipStream_t st)
{
    if (c.nb == 0) return hipSuccess;
    hipLaunchKernelGGL(krpc_rock, dim3((unsigned)((c.nb + 255) / 256)), dim3(256), 0, st, s, Pref, cp1, cp2, c);
    return hipGetLastError();
}

template <int NCB>
static hipError_t launch_res_asm(const CellIO& c, const ConnIO& cn, const BsrIO& bsr, double dt, bool dr,
                                 bool dj, double* res, double* bmax, int64_t* barg, hipStream_t st)
{
    constexpr int G = 64 / (NCB + 1);
    const unsigned nblk = (unsigned)((c.nb + G - 1) / G);
    hipLaunchKernelGGL(res_asm<NCB>, dim3(nblk), dim3(64), 0, st, c, cn, bsr, dt, dr ? 1 : 0, dj ? 1 : 0, res,
                       bmax, barg);
    return hipGetLastError();
}

int64_t res_asm_blocks(int nc, int64_t nb) { const int G = 64 / (nc + 1); return (nb + G - 1) / G; }

#define OCPX_NCB_SWITCH(X)                                                                          \
    switch (nc) {                                                                                   \
    X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13)                                      \
    default: return hipErrorInvalidValue;                                                           \
    }

hipError_t run_res_asm(int nc, const CellIO& c, const ConnIO& cn, const BsrIO& bsr, double dt, bool dr,
                       bool dj, double* res, double* bmax, int64_t* barg, hipStream_t st)
{
    if (c.nb == 0) return hipSuccess;
#define X(N) case N: return launch_res_asm<N>(c, cn, bsr, dt, dr, dj, res, bmax, barg, st);
    OCPX_NCB_SWITCH(X)
#undef X
}

hipError_t run_norms_final(int64_t nblocks, const double* bmax, const int64_t* barg, double* out4, hipStream_t st)
{
    hipLaunchKernelGGL(norms_final, dim3(1), dim3(256), 0, st, nblocks, bmax, barg, out4);
    return hipGetLastError();
}

hipError_t run_get_solution(int nc, double* P, double* Ni, double* S, double* xij, const double* X,
                            const double* u, double dSmax, int64_t nb, hipStream_t st)
{
    if (nb == 0) return hipSuccess;
    const unsigned nblk = (unsigned)((nb + 255) / 256);
#define X_(N) case N: hipLaunchKernelGGL(get_solution<N>, dim3(nblk), dim3(256), 0, st, P, Ni, S, xij, X, u, dSmax, nb); return hipGetLastError();
    switch (nc) {
    X_(3) X_(4) X_(5) X_(6) X_(7) X_(8) X_(9) X_(10) X_(11) X_(12) X_(13)
    default: return hipErrorInvalidValue;
    }
#undef X_
}

hipError_t run_jt_lambda(int nc, int64_t nb, const BsrIO& bsr, const double* lam, double* out, hipStream_t st)
{
    if (nb == 0) return hipSuccess;
    const int G = 64 / (nc + 1);
    const unsigned nblk = (unsigned)((nb + G - 1) / G);
#define X_(N) case N: hipLaunchKernelGGL(jt_lambda<N>, dim3(nblk), dim3(64), 0, st, bsr, nb, lam, out); return hipGetLastError();
    switch (nc) {
    X_(3) X_(4) X_(5) X_(6) X_(7) X_(8) X_(9) X_(10) X_(11) X_(12) X_(13)
    default: return hipErrorInvalidValue;
    }
#undef X_
}

hipError_t run_dR_dtrans(int nc, const CellIO& c, const ConnIO& cn, double dt, const double* lam, double* dT,
                         double* dVp, hipStream_t st)
{
    const int64_t n = c.nb > cn.nconn ? c.nb : cn.nconn;
    if (n == 0) return hipSuccess;
    const unsigned nblk = (unsigned)((n + 255) / 256);
#define X_(N) case N: hipLaunchKernelGGL(dR_dtrans<N>, dim3(nblk), dim3(256), 0, st, c, cn, dt, lam, dT, dVp); return hipGetLastError();
    switch (nc) {
    X_(3) X_(4) X_(5) X_(6) X_(7) X_(8) X_(9) X_(10) X_(11) X_(12) X_(13)
    default: return hipErrorInvalidValue;
    }
#undef X_
}

}  // namespace ocpx
