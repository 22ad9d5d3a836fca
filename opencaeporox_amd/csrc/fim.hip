// fim.hip — kr/Pc + rock, residual + Jacobian assembly, Newton update and
// the adjoint products of the FIM system, for gfx950.
//
// Assembly (IsoT_FIM::CalRes + AssembleMatBulks, src/IsoThermalMethod.cpp:
// 1090-1269, with OCPConvection01::CalFlux/AssembleMatFIM,
// src/OCPConvection.cpp:22-187, and BulkAccumuTerm01, src/BulkAccumuModule.cpp:
// 24-70) runs in two passes, both atomic-free:
//
//  1. krpc_aux (one cell per b = nc+1 lanes, lane r = primary variable r):
//     kr/Pc and rock as in the reference, then the cell's connection record:
//     per phase j the scalars a connection reads (rho, Pj, xi, kr, mu,
//     existence, depth), x_ij, and the total derivatives with respect to
//     the primary variables of the mobility lambda_j = xi_j kr_j / mu_j, the
//     phase pressure P_j = P + Pc_j and the mass density rho_j, chained
//     through dXs/dXp once per cell.
//  2. res_asm (one block row per b lanes, lane r = block column r): for each
//     connection of the row, in OCPMatrix push order, the flux
//     F_i = T lambda_U x_ijU dPhi_j and column r of dF/dX on either side,
//         dF_i/dX_s = T dPhi_j [s == U] (x_ij dlambda_U + lambda_U dx_ij,U)
//                   + T lambda_U x_ij dPhi_j/dX_s,
//     with dPhi_j/dX_b = dP_jb - g dz w_b drho_jb, dPhi_j/dX_e = -dP_je -
//     g dz w_e drho_je.  That is the reference's dt (dF/dXp + dF/dXs dXs/dXp)
//     with the product taken once per cell instead of once per connection.
//     A lane reads one element of each derivative vector (the B lanes of a
//     row read B contiguous doubles) and writes column r of each block.
//     Both rows of a connection compute its flux; that doubled arithmetic is
//     cheaper than a connection-ordered scatter.
//  res_only is the residual alone (CalRes without a Jacobian).
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
// (src/OCPFlowMethod.cpp:63-73) and Stone II (src/OCPFuncSAT.cpp:162-186)
OCPX_HD void krpc_cell(const DevSat& s, double Sg, double Sw, double* kr, double* pc, double* dk, double* dp)
{
    double d[4], sl[4];
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
}

// Also writes the cell's connection record (aux_len layout) for the assembly.
template <int NCB>
__global__ void __launch_bounds__(64) krpc_aux(const DevSat* __restrict__ sp, const DevRock* __restrict__ rp,
                                               CellIO c)
{
    constexpr int nc = NCB, B = nc + 1, np = 3, G = 64 / B, REC = np * B * B, RL = aux_len(B, nc);
    constexpr int SV = 24, SX = 24 + 9 * B;
    const int lane = threadIdx.x, g = lane / B, r = lane % B;
    const int64_t n = (int64_t)blockIdx.x * G + g;
    if (g >= G || n >= c.nb) return;
    const size_t nb = c.nb;
    const int model = c.mid ? c.mid[n] : 0;
    const double Pref = rp[model].Pref, cp1 = rp[model].cp1, cp2 = rp[model].cp2;
    double kr[3], pc[3], dk[9], dp[9];
    krpc_cell(sp[model], c.S[nb + n], c.S[2 * nb + n], kr, pc, dk, dp);
    const double P = c.P[n];
    if (r == 0) {
        UNROLL for (int j = 0; j < 3; j++) {
            c.kr[j * nb + n] = kr[j];
            c.Pc[j * nb + n] = pc[j];
            c.Pj[j * nb + n] = P + pc[j];
        }
        UNROLL for (int k = 0; k < 9; k++) { c.dKrdS[k * nb + n] = dk[k]; c.dPcdS[k * nb + n] = dp[k]; }
        // IsothermalMethod::CalRock (:19-30) via OCPRockIsoT_Linear (src/OCPRock.cpp:19-24)
        const double dP = P - Pref;
        const double poro = c.poroInit[n] * (1 + (cp1 + cp2 / 2 * dP) * dP);
        c.poro[n] = poro;
        c.poroP[n] = c.poroInit[n] * (cp1 + cp2 * dP);
        c.rockVp[n] = c.v[n] * poro;
    }
    if (!c.aux) return;
    const double* X = c.dXsdXp + (size_t)n * REC;     // row k, column r: X[k * B + r]
    double dS[np];
    UNROLL for (int k = 0; k < np; k++) dS[k] = X[k * B + r];
    double* A = c.aux + (size_t)n * RL;
    const double one = r == 0 ? 1.0 : 0.0;
    const double depth = c.depth[n];
    UNROLL for (int j = 0; j < np; j++) {
        const bool ex = c.phaseExist[j * nb + n] != 0;
        double drho = 0, dxi = 0, dmu = 0;
        if (ex) {
            drho = one * c.rhoP[j * nb + n];
            dxi = one * c.xiP[j * nb + n];
            dmu = one * c.muP[j * nb + n];
            UNROLL for (int m = 0; m < nc; m++) {
                const double xm = X[(np + j * nc + m) * B + r];
                drho += c.rhox[(size_t)(j * nc + m) * nb + n] * xm;
                dxi += c.xix[(size_t)(j * nc + m) * nb + n] * xm;
                dmu += c.mux[(size_t)(j * nc + m) * nb + n] * xm;
            }
        }
        double dkr = 0, dPj = one;
        UNROLL for (int k = 0; k < np; k++) { dkr += dk[j * 3 + k] * dS[k]; dPj += dp[j * 3 + k] * dS[k]; }
        const double xi = c.xi[j * nb + n], mu = c.mu[j * nb + n];
        double lam = 0, dlam = 0;
        if (ex) {
            lam = xi * kr[j] / mu;
            dlam = kr[j] / mu * dxi + xi / mu * dkr - xi * kr[j] / (mu * mu) * dmu;
        }
        A[SV + 3 * B * j + r] = dlam;
        A[SV + 3 * B * j + B + r] = dPj;
        A[SV + 3 * B * j + 2 * B + r] = drho;
        if (r < nc) A[SX + j * nc + r] = c.xij[(size_t)(j * nc + r) * nb + n];
        if (r < 4) {                          // the phase's 8 scalars, one 16-byte pair per lane
            const double v0 = r == 0 ? c.rho[j * nb + n] : (r == 1 ? xi : (r == 2 ? mu : lam));
            const double v1 = r == 0 ? P + pc[j] : (r == 1 ? kr[j] : (r == 2 ? (ex ? 1.0 : 0.0) : depth));
            reinterpret_cast<double2*>(A + 8 * j)[r] = make_double2(v0, v1);
        }
    }
}

// Blocks are dealt to the 8 XCDs round-robin (block b runs on XCD b % 8).
// Renumber them so each XCD works through one contiguous chunk of rows and
// the neighbour records a row reads stay in that XCD's L2.
__device__ __forceinline__ int64_t xcd_chunk(int64_t b, int64_t nblk)
{
    constexpr int64_t X = 8;
    const int64_t x = b % X, k = b / X, q = nblk / X, r = nblk % X;
    return x * q + (x < r ? x : r) + k;
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

// Block reduction of max |R/Vp| and max |R_i/Nt| (first cell on ties) into
// bmax/barg[2 * block + {0, 1}].
__device__ __forceinline__ void block_norms(double absV, double absN, int64_t idx, double* bmax, int64_t* barg)
{
    __shared__ double sV[64], sN[64];
    __shared__ int64_t iV[64], iN[64];
    const int lane = threadIdx.x;
    sV[lane] = absV; iV[lane] = idx;
    sN[lane] = absN; iN[lane] = idx;
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

// Residual only (IsoT_FIM::CalRes), one block row per b lanes, lane i =
// equation i; reads the cell state arrays directly.
template <int NCB>
__global__ void __launch_bounds__(64) res_only(CellIO c, ConnIO cn, BsrIO bsr, double dt,
                                               double* __restrict__ resAbs, double* __restrict__ bmax,
                                               int64_t* __restrict__ barg)
{
    constexpr int nc = NCB, B = nc + 1, np = 3, G = 64 / B;
    const int lane = threadIdx.x;
    const int g = lane / B, i = lane % B;
    const size_t nb = c.nb;
    const int64_t blk = xcd_chunk(blockIdx.x, gridDim.x);
    const int64_t p = blk * G + g;
    const bool valid = (g < G) && (p < c.nb);
    const int64_t n = valid ? (bsr.order ? (int64_t)bsr.order[p] : p) : 0;
    double absV = -1, absN = -1;
    if (valid) {
        double res = i == 0 ? c.rockVp[n] - c.vf[n] : c.Ni[(size_t)(i - 1) * nb + n] - c.lNi[(size_t)(i - 1) * nb + n];
        const int64_t s0 = bsr.rowptr[n], s1 = bsr.rowptr[n + 1];
        for (int64_t s = s0 + 1; s < s1 && i > 0; s++) {
            const int cc = bsr.nbr_conn[s];
            const int64_t bI = cn.cb[cc], eI = cn.ce[cc];
            const double T = cn.trans[cc];
            double flux = 0;
            UNROLL for (int j = 0; j < np; j++) {
                Upw u;
                if (!upwind(c, nb, bI, eI, j, u)) continue;
                const int64_t U = u.upIsB ? bI : eI;
                const double tJ = T * c.kr[j * nb + U] / c.mu[j * nb + U];
                flux += tJ * u.dP * c.xi[j * nb + U] * c.xij[(size_t)(j * nc + i - 1) * nb + U];
            }
            res += bI == n ? dt * flux : -dt * flux;
        }
        const double out = -res;             // Dscalar(-1) :1170
        resAbs[(size_t)n * B + i] = out;
        if (bsr.rhs) bsr.rhs[(size_t)n * B + i] = out;   // CopyRhs
        absV = fabs(res / c.rockVp[n]);
        if (i > 0) absN = fabs(res / c.Nt[n]);
    }
    block_norms(absV, absN, valid ? n : INT64_MAX, bmax, barg);
}

// A Jacobian block is written once and not read again by the assembly.  When
// a block is whole 128-byte lines (b = 4, 8) its stores bypass the caches'
// reuse (nontemporal); otherwise (b = 7: 392-byte blocks straddle lines)
// plain stores let L2 merge the partial lines of neighbouring blocks.  On
// 100x100x50: b = 8 1.28 ms nontemporal vs 1.31 plain, b = 7 1.37 vs 1.33.
template <int B>
__device__ __forceinline__ void put_block(double v, double* p)
{
    if constexpr ((B * B * sizeof(double)) % 128 == 0) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// Residual + Jacobian (CalRes + AssembleMatBulks), one block row per b
// lanes, lane r = block column r (and equation r of the residual).  Every
// load of a connection's vectors is one element per lane, 64 contiguous
// bytes per row; the cell scalars come from the connection record.
template <int NCB>
__global__ void __launch_bounds__(64) res_asm(CellIO c, ConnIO cn, BsrIO bsr, double dt,
                                              double* __restrict__ resAbs, double* __restrict__ bmax,
                                              int64_t* __restrict__ barg)
{
    constexpr int nc = NCB, B = nc + 1, np = 3, G = 64 / B, REC = np * B * B, RL = aux_len(B, nc);
    constexpr int SV = 24, SX = 24 + 9 * B;
    // The part of a record every lane of a row reads alike (the phase scalars
    // and x_ij) is staged in LDS by the row's b lanes with 16-byte loads:
    // slot 0 the row's own cell (once), slot 1 the other cell of the current
    // connection.  Staged copy: [0, 24) the scalars, then the record's
    // [XA, RL) (x_ij from its 16-byte aligned start to the record's end).
    constexpr int XA = SX & ~1, SR = 24 + (RL - XA), OX = 24 + (SX - XA), NCH = SR / 2;
    static_assert(RL % 2 == 0 && SR % 2 == 0, "records and staged copies are whole 16-byte chunks");
    __shared__ __align__(16) double stg[G][2][SR];
    const int lane = threadIdx.x;
    const int g = lane / B, r = lane % B;
    const int gs = g < G ? g : 0;                      // staging group (the lanes past G stage nothing)
    const size_t nb = c.nb;
    const int64_t blk = xcd_chunk(blockIdx.x, gridDim.x);
    const int64_t p = blk * G + g;                     // position in the row order
    const bool valid = (g < G) && (p < c.nb);
    const int64_t n = valid ? (bsr.order ? (int64_t)bsr.order[p] : p) : 0;
    auto stage = [&](int slot, int64_t cell) {
        const double* src = c.aux + (size_t)cell * RL;
        for (int ch = r; ch < NCH; ch += B) {
            const int so = ch < 12 ? 2 * ch : XA + 2 * (ch - 12);
            *reinterpret_cast<double2*>(&stg[gs][slot][2 * ch]) = *reinterpret_cast<const double2*>(src + so);
        }
    };
    double absV = -1, absN = -1;
    double diag[B];
    double res = 0;
    int64_t s0 = 0, s1 = 0;
    if (valid) {
        // accumulation (BulkAccumuTerm01::CalResFIM / CaldFdXpFIM), column r
        diag[0] = r == 0 ? c.v[n] * c.poroP[n] - c.vfP[n] : -c.vfi[(size_t)(r - 1) * nb + n];
        UNROLL for (int i = 1; i < B; i++) diag[i] = (r == i) ? 1.0 : 0.0;
        res = r == 0 ? c.rockVp[n] - c.vf[n] : c.Ni[(size_t)(r - 1) * nb + n] - c.lNi[(size_t)(r - 1) * nb + n];
        s0 = bsr.rowptr[n];
        s1 = bsr.rowptr[n + 1];
        if (s1 - s0 > 1) stage(0, n);
    }
    // every lane runs the wavefront's largest connection count (the block is
    // one wavefront), so the barriers below are reached by all lanes
    const int cnt = valid ? (int)(s1 - s0 - 1) : 0;
    int mc = cnt;
    for (int o = 32; o > 0; o >>= 1) mc = max(mc, __shfl_xor(mc, o));
    __syncthreads();
    for (int k = 0; k < mc; k++) {
        const bool act = k < cnt;
        const int64_t s = s0 + 1 + k;
        int cc = 0;
        int64_t bI = 0, eI = 0;
        double T = 0;
        if (act) {
            cc = bsr.nbr_conn[s];
            bI = cn.cb[cc];
            eI = cn.ce[cc];
            T = cn.trans[cc];
        }
        const bool ownB = (bI == n);
        if (act) stage(1, ownB ? eI : bI);
        __syncthreads();
        if (act) {
            const double* Lb = stg[gs][ownB ? 0 : 1];
            const double* Le = stg[gs][ownB ? 1 : 0];
            const double* Ab = c.aux + (size_t)bI * RL;
            const double* Ae = c.aux + (size_t)eI * RL;
            const double zb = Lb[7], ze = Le[7];
            const double dGamma = GRAVF * (zb - ze);
            double rb[nc], re[nc];                     // column r of dF/dX_b and dF/dX_e, rows 1..nc
            UNROLL for (int i = 0; i < nc; i++) { rb[i] = 0; re[i] = 0; }
            double flux = 0;
            UNROLL for (int j = 0; j < np; j++) {
                // upwinding (OCPConvection01::CalFlux :37-78)
                const double* Sb = Lb + 8 * j;
                const double* Se = Le + 8 * j;
                const bool exb = Sb[5] != 0, exe = Se[5] != 0;
                if (!exb && !exe) continue;
                double rho;
                if (exb && exe) rho = (Sb[0] + Se[0]) / 2;
                else if (exb) rho = Sb[0];
                else rho = Se[0];
                const double dP = (Sb[1] - GRAVF * rho * zb) - (Se[1] - GRAVF * rho * ze);
                const bool upIsB = !(dP < 0);
                if (!(upIsB ? exb : exe)) continue;
                const bool downExists = upIsB ? exe : exb;
                const double* LU = upIsB ? Lb : Le;
                const int64_t U = upIsB ? bI : eI;
                const double xi = LU[8 * j + 2], kr = LU[8 * j + 3], mu = LU[8 * j + 4], lam = LU[8 * j + 6];
                const double tJ = T * kr / mu;
                double x[nc];
                UNROLL for (int i = 0; i < nc; i++) x[i] = LU[OX + j * nc + i];
                double xr = 0;                          // component of equation r
                UNROLL for (int i = 0; i < nc; i++) xr = (r == i + 1) ? x[i] : xr;
                flux += tJ * dP * xi * xr;
                const double wU = downExists ? 0.5 : 1.0, wD = downExists ? 0.5 : 0.0;
                const double wb = upIsB ? wU : wD, we = upIsB ? wD : wU;
                const double* Vb = Ab + SV + 3 * B * j;
                const double* Ve = Ae + SV + 3 * B * j;
                const double vb = Vb[B + r] - dGamma * (wb * Vb[2 * B + r]);
                const double ve = -Ve[B + r] - dGamma * (we * Ve[2 * B + r]);
                const double dl = (upIsB ? Vb : Ve)[r];
                const double Tdp = T * dP;
                const double* XU = c.dXsdXp + (size_t)U * REC + (size_t)(np + j * nc) * B + r;
                UNROLL for (int i = 0; i < nc; i++) {
                    const double tIJ = x[i] * xi * tJ;
                    const double up = Tdp * (x[i] * dl + lam * XU[i * B]);
                    rb[i] += tIJ * vb + (upIsB ? up : 0.0);
                    re[i] += tIJ * ve + (upIsB ? 0.0 : up);
                }
            }
            if (r > 0) res += ownB ? dt * flux : -dt * flux;
            double* off = bsr.val + (size_t)s * B * B + r;        // column r of the slot's block
            // the volume row carries no flux
            put_block<B>(ownB ? dt * 0.0 : -(dt * 0.0), &off[0]);
            UNROLL for (int i = 0; i < nc; i++) {
                if (ownB) {                       // AddDiag(b, B); NewOffDiag(b, e, E)
                    diag[i + 1] += dt * rb[i];
                    put_block<B>(dt * re[i], &off[(i + 1) * B]);
                } else {                          // NewOffDiag(e, b, -B); AddDiag(e, -E)
                    put_block<B>(-(dt * rb[i]), &off[(i + 1) * B]);
                    diag[i + 1] += -(dt * re[i]);
                }
            }
        }
        __syncthreads();                          // slot 1 is staged again next connection
    }
    if (valid) {
        double* dg = bsr.val + (size_t)s0 * B * B + r;
        UNROLL for (int i = 0; i < B; i++) put_block<B>(diag[i], &dg[i * B]);
        const double out = -res;             // Dscalar(-1) :1170
        resAbs[(size_t)n * B + r] = out;
        if (bsr.rhs) bsr.rhs[(size_t)n * B + r] = out;   // CopyRhs
        absV = fabs(res / c.rockVp[n]);
        if (r > 0) absN = fabs(res / c.Nt[n]);
    }
    block_norms(absV, absN, valid ? n : INT64_MAX, bmax, barg);
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
        UNROLL for (int j = 0; j < col; j++) s += 1.0 * X[((size_t)n * row + i) * col + j] * uu[j];
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

hipError_t run_krpc_rock(const DevSat* s, const DevRock* r, const CellIO& c, hipStream_t st)
{
    if (c.nb == 0) return hipSuccess;
    const int nc = c.nc;
    const int G = 64 / (nc + 1);
    const unsigned nblk = (unsigned)((c.nb + G - 1) / G);
#define X_(N) case N: hipLaunchKernelGGL(krpc_aux<N>, dim3(nblk), dim3(64), 0, st, s, r, c); return hipGetLastError();
    switch (nc) {
    X_(3) X_(4) X_(5) X_(6) X_(7) X_(8) X_(9) X_(10) X_(11) X_(12) X_(13)
    default: return hipErrorInvalidValue;
    }
#undef X_
}

// Norms of a residual already in resAbs (-R), as the norm loop of
// IsoT_FIM::CalRes (src/IsoThermalMethod.cpp:1147-1168): used after terms
// are added to the bulk rows (wells).
template <int NCB>
__global__ void __launch_bounds__(64) res_norms(CellIO c, const double* __restrict__ resAbs,
                                                double* __restrict__ bmax, int64_t* __restrict__ barg)
{
    constexpr int B = NCB + 1, G = 64 / B;
    const int lane = threadIdx.x;
    const int g = lane / B, i = lane % B;
    const int64_t n = (int64_t)blockIdx.x * G + g;
    const bool valid = (g < G) && (n < c.nb);
    double absV = -1, absN = -1;
    if (valid) {
        const double r = resAbs[(size_t)n * B + i];
        absV = fabs(r / c.rockVp[n]);
        if (i > 0) absN = fabs(r / c.Nt[n]);
    }
    block_norms(absV, absN, valid ? n : INT64_MAX, bmax, barg);
}

template <int NCB>
static hipError_t launch_res_asm(const CellIO& c, const ConnIO& cn, const BsrIO& bsr, double dt, bool dj,
                                 double* res, double* bmax, int64_t* barg, hipStream_t st)
{
    constexpr int G = 64 / (NCB + 1);
    const unsigned nblk = (unsigned)((c.nb + G - 1) / G);
    if (dj) hipLaunchKernelGGL(res_asm<NCB>, dim3(nblk), dim3(64), 0, st, c, cn, bsr, dt, res, bmax, barg);
    else hipLaunchKernelGGL(res_only<NCB>, dim3(nblk), dim3(64), 0, st, c, cn, bsr, dt, res, bmax, barg);
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
    if (!dr) return hipErrorInvalidValue;      // the Jacobian pass always forms the residual too
    if (dj && !c.aux && cn.nconn > 0) return hipErrorInvalidValue;   // records are read only through connections
#define X(N) case N: return launch_res_asm<N>(c, cn, bsr, dt, dj, res, bmax, barg, st);
    OCPX_NCB_SWITCH(X)
#undef X
}

hipError_t run_res_norms(int nc, const CellIO& c, const double* res, double* bmax, int64_t* barg, hipStream_t st)
{
    if (c.nb == 0) return hipSuccess;
    const unsigned nblk = (unsigned)res_asm_blocks(nc, c.nb);
#define X(N) case N: hipLaunchKernelGGL(res_norms<N>, dim3(nblk), dim3(64), 0, st, c, res, bmax, barg); return hipGetLastError();
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
