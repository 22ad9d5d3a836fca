// pe.hip — per-cell Peng-Robinson phase equilibrium on gfx950.
//
// Algorithm: OCPPhaseEquilibrium::PhaseEquilibrium
// (src/OCPPhaseEquilibrium.cpp:71-112): Wilson K-values, strict stability
// (StableSSM + StableNR, 4 trial sets, :397-541) or the relaxed test
// (StableSSM01, :299-395) when the previous step had phases, then two-phase
// split (SplitSSM2 + RachfordRice2 + SplitNR + CheckSplit, :549-1010), with
// SkipPSA's flash-type choice (src/AcceleratePEC.cpp:64-124) at the start.
//
// Each cell is a resumable state machine whose every step evaluates ln(phi)
// of one composition y at the same program point (the "common evaluator").
// That lets one kernel body serve two mappings:
//   * pe_plain   — one cell per lane, stepping until the cell is done;
//   * pe_batched — persistent lanes that take a new cell from a global work
//                  queue as soon as their cell is done, so lanes stuck on
//                  near-saturation cells (hundreds of SSM steps) do not idle
//                  the rest of the wavefront.
// The residuals and updates are the reference's literal expressions
// (fugacities phi*x*P, their ratios and logs), so degenerate trajectories
// (K far from 1, Rachford-Rice fractions outside [0, 1]) overflow and
// stop the way the reference's do.  In pe_batched a lane that reaches a
// Newton step waits for a batch of them (NR_BATCH) before the wavefront
// runs the Newton branch.
#include "ocpx_dev.hpp"
#include "ocpx_kernels.hpp"

namespace ocpx {

enum : int {
    ST_DONE = 0,
    ST_FEED,      // evaluate ln phi of the feed
    ST_T_START,   // strict stability: trial k start   (:418-433)
    ST_T_SSM,     //   SSM iteration                    (:437-458)
    ST_T_NR,      //   NR iteration                     (:488-516)
    ST_A_START,   // relaxed stability: trial k start   (:322-335)
    ST_A_ITER,    //   iteration                        (:328-375)
    ST_S_SSM_A,   // split SSM: RR + evaluate phase 0   (:624-629)
    ST_S_SSM_B,   //   evaluate phase 1 and update K    (:629-643)
    ST_S_NR,      // split NR iteration                 (:823-891)
};

constexpr int MAX_SPLIT_RESTARTS = 64;   // the reference's restart loop has no cap
constexpr int MAX_STEPS = 200000;        // hard bound on a cell's steps (flagged if hit)
constexpr int NR_BATCH = 16;             // lanes a Newton step waits for (pe_batched)
constexpr double KTOL = 1E-4, DYTOL = 1E-6;   // include/OCPPhaseEquilibrium.hpp:43-44

template <int NC>
struct PeLane {
    int cell, st, k, iter, maxIt, NP, ftype, flag, stage, restarts, ssm_conv, nr_conv, steps;
    uint32_t lNP;
    double P, T, logP, Yt, tol2, Sk, nu0, eNR0;
    int cnt[8];
    double z[NC];
    EosCell<NC> ec;
    double y[NC];
    double a1[NC], a2[NC], a3[NC], a4[NC], a5[NC];
};

// log of the Wilson K of trial set k: Kw[k] = Kw0^c_k, c = 1, -1, 1/3, -1/3
// (src/OCPPhaseEquilibrium.cpp:244-252)
OCPX_HD double trial_coef(int k)
{
    return k == 0 ? 1.0 : (k == 1 ? -1.0 : (k == 2 ? (1.0 / 3) : -(1.0 / 3)));
}
template <int NC>
OCPX_HD double log_kw0(const DevFluid& f, const PeLane<NC>& L, int i)
{
    return f.logPc[i] - L.logP + f.kwA[i] * (1 - f.Tc[i] / L.T);
}

template <int NC>
OCPX_HD void normalise_y(PeLane<NC>& L)
{
    double yt = 0;
    UNROLL for (int i = 0; i < NC; i++) yt += L.y[i];
    L.Yt = yt;
    const double r = 1 / yt;
    UNROLL for (int i = 0; i < NC; i++) L.y[i] *= r;
}

// ---------------------------------------------------------------- helpers
template <int NC>
OCPX_HD void rr2(const DevFluid& f, PeLane<NC>& L)
{   // RachfordRice2 :651-705 on K = a4
    const double* K = L.a4;
    double Kmin = K[0], Kmax = K[0];
    UNROLL for (int i = 1; i < NC; i++) { Kmin = fmin(Kmin, K[i]); Kmax = fmax(Kmax, K[i]); }
    const double numin = 1 / (1 - Kmax), numax = 1 / (1 - Kmin);
    double nu = 0.5 * (numin + numax);
    int it = 0;
    while (true) {
        double rj = 0, J = 0;
        UNROLL for (int i = 0; i < NC; i++) {
            const double km = K[i] - 1;
            const double t = 1 / (1 + nu * km);
            rj += L.z[i] * km * t;
            J -= L.z[i] * km * km * t * t;
        }
        if (fabs(rj) < f.RR_tol || it > f.RR_maxIt) break;
        const double d = -rj / J, tn = nu + d;
        if (tn < numax && tn > numin) nu = tn;
        else nu = d > 0 ? (nu + numax) / 2 : (nu + numin) / 2;
        it++;
    }
    L.cnt[4] += it;
    L.nu0 = nu;
}

template <int NC>
OCPX_HD void update_x_rr(PeLane<NC>& L)
{   // UpdateXRR :778-791
    UNROLL for (int i = 0; i < NC; i++) {
        L.a3[i] = L.z[i] / (1 + L.nu0 * (L.a4[i] - 1));
        L.a2[i] = L.a4[i] * L.a3[i];
    }
}

template <int NC>
OCPX_HD void split_start(const DevFluid& f, PeLane<NC>& L, const PeIO& io, bool restart)
{   // SplitSSM2 :586-622
    L.ssm_conv = 1;
    L.tol2 = f.SSMsp_tol2;
    L.maxIt = f.SSMsp_maxIt;
    if (!restart) {
        if (L.lNP == 2) {
            const size_t n = io.nb;
            const int nc = NC + 1;
            UNROLL for (int i = 0; i < NC; i++)
                L.a4[i] = io.xij[(size_t)i * n + L.cell] / io.xij[(size_t)(nc + i) * n + L.cell];
        } else {
            UNROLL for (int i = 0; i < NC; i++) L.a4[i] = exp(log_kw0(f, L, i));
        }
    } else {
        L.tol2 *= 1E-1;
        L.maxIt *= 2;
    }
    if (L.Yt < 1.1) {
        L.tol2 *= 1E-1;
        L.maxIt *= 2;
    }
    L.iter = 0;
    L.st = ST_S_SSM_A;
}

template <int NC>
OCPX_HD bool check_split(PeLane<NC>& L)
{   // CheckSplit :960-1010
    if (L.NP != 2) return true;
    const double nu1 = 1 - L.nu0;
    const double nuMax = fmax(L.nu0, nu1);
    double eX = 0;
    UNROLL for (int i = 0; i < NC; i++) eX += (L.a2[i] - L.a3[i]) * (L.a2[i] - L.a3[i]);
    if (nuMax < 1 && L.nr_conv && finite_d(eX)) return true;
    if (!finite_d(eX) || 1 - nuMax < 1E-3) {
        L.NP = 1;
        L.nu0 = 1;
        UNROLL for (int i = 0; i < NC; i++) L.a2[i] = L.z[i];
        return false;
    }
    return true;
}

template <int NC>
OCPX_HD void finish_split(PeLane<NC>& L)
{
    check_split(L);
    L.st = ST_DONE;
}

// after SplitNR ends: the PhaseSplit restart loop (:558-566)
template <int NC>
OCPX_HD void after_split_nr(const DevFluid& f, PeLane<NC>& L, const PeIO& io)
{
    if (L.stage == 0) {
        if (L.nr_conv) { finish_split(L); return; }
        L.stage = 1;
    } else {
        if (!check_split(L)) { finish_split(L); return; }
        if (L.ssm_conv || L.nr_conv) { finish_split(L); return; }
    }
    if (++L.restarts > MAX_SPLIT_RESTARTS) { L.cnt[7] |= 0x40000000; finish_split(L); return; }
    split_start(f, L, io, true);
}

template <int NC>
OCPX_HD void split_nr_begin(const DevFluid& f, PeLane<NC>& L, const PeIO& io,
                                               const double* lnphi1)
{   // SplitNR :807-823; a5 (= ln phi0 of x0) and lnphi1 at the current x0, x1
    double e2 = 0;
    UNROLL for (int i = 0; i < NC; i++) {
        L.a1[i] = L.nu0 * L.a2[i];                                  // x2n
        L.a5[i] = log((exp(lnphi1[i]) * L.a3[i] * L.P) / (exp(L.a5[i]) * L.a2[i] * L.P));   // ln(f1/f0)
        e2 += L.a5[i] * L.a5[i];
    }
    const double eNR = sqrt(e2);
    L.eNR0 = eNR;
    L.iter = 0;
    L.nr_conv = 0;
    if (eNR > f.NRsp_tol) { L.st = ST_S_NR; return; }
    L.nr_conv = eNR < f.NRsp_tol;
    after_split_nr(f, L, io);
}

// decision after StableNR in the strict test (:460-470)
template <int NC>
OCPX_HD void strict_decide(const DevFluid& f, PeLane<NC>& L, const PeIO& io)
{
    if (L.flag && L.Yt > 1 + f.eYt) {           // unstable -> split
        L.NP = 2;
        L.stage = 0;
        L.restarts = 0;
        split_start(f, L, io, false);
        return;
    }
    if (++L.k < 4) { L.st = ST_T_START; return; }
    L.st = ST_DONE;                               // stable, NP = 1
}

template <int NC>
OCPX_HD void nr_sta_begin(const DevFluid& f, PeLane<NC>& L, const PeIO& io)
{   // StableNR :478-488, resSTA in a3
    double e2 = 0;
    UNROLL for (int i = 0; i < NC; i++) e2 += L.a3[i] * L.a3[i];
    L.iter = 0;
    if (sqrt(e2) > f.NRsta_tol) { L.st = ST_T_NR; return; }
    L.flag = 1;
    strict_decide(f, L, io);
}

// end of one relaxed trial (:377-392) and of the relaxed test (PhaseStable :283-291)
template <int NC>
OCPX_HD void relaxed_trial_end(const DevFluid& f, PeLane<NC>& L, const PeIO& io)
{
    if (L.flag && L.Yt > 1 - 0.1 && L.Sk > 1) L.ftype = 2;
    bool stable;
    if (L.flag && L.Yt > 1 + f.eYt) stable = false;
    else if (++L.k < 2) { L.st = ST_A_START; return; }
    else stable = true;
    const uint32_t tmpNP = 1u + (stable ? 0u : 1u);
    if (tmpNP != L.lNP) {                         // rerun the strict test
        L.ftype = 2;
        L.k = 0;
        L.st = ST_T_START;
        return;
    }
    if (stable) { L.st = ST_DONE; return; }
    L.NP = 2;
    L.stage = 0;
    L.restarts = 0;
    split_start(f, L, io, false);
}

// ------------------------------------------------------------ cell begin/end
template <int NC>
OCPX_HD void pe_begin(const DevFluid& f, const PeIO& io, int cell, PeLane<NC>& L)
{
    const size_t n = io.nb;
    const int nc = NC + 1;
    L.cell = cell;
    L.P = io.P[cell];
    L.T = io.T[cell] + CONV4;                     // SetVarSet :674
    L.logP = log(L.P);
    double Ni[NC];
    double Nt = 0;
    UNROLL for (int i = 0; i < NC; i++) { Ni[i] = io.Ni[(size_t)i * n + cell]; Nt += Ni[i]; }
    UNROLL for (int i = 0; i < NC; i++) L.z[i] = Ni[i] / Nt;         // InitNtZ :812-818
    uint32_t phaseNum = 1;                         // no previous step: one phase
    if (io.phaseExist) {
        phaseNum = 0;
        UNROLL for (int j = 0; j < 3; j++) phaseNum += io.phaseExist[(size_t)j * n + cell] != 0;
    }
    L.lNP = phaseNum - 1u;
    // flash type: explicit override or SkipPSA::CalFtype02 (src/AcceleratePEC.cpp:103-124)
    int ft = io.ftype_in ? io.ftype_in[cell] : -1;
    if (ft < 0) {
        ft = 0;
        if (f.skip) {
            bool skip = false;
            if (io.skip_flag[cell]) {
                const double me = io.skip_minEig[cell];
                skip = fabs(1 - io.skip_P[cell] / L.P) < me / 10 && fabs(io.skip_T[cell] - L.T) < me * 10;
                UNROLL for (int i = 0; i < NC; i++)
                    skip = skip && fabs(Ni[i] / Nt - io.skip_zi[(size_t)i * n + cell]) < me / 10;
            }
            const uint32_t npPE = phaseNum - 1u;
            if (skip) ft = 1;
            else if (npPE >= 2) {
                uint32_t t = (io.S[cell] >= 1E-4) + (io.S[n + cell] >= 1E-4);
                ft = t == npPE ? (int)npPE : 0;
            }
        }
    }
    L.ftype = ft;
    UNROLL for (int c = 0; c < 8; c++) L.cnt[c] = 0;
    eos_cell<NC>(f, L.P, L.T, L.ec);
    L.NP = 1;
    L.nu0 = 1;
    L.k = 0;
    L.flag = 0;
    L.stage = 0;
    L.restarts = 0;
    L.Yt = 0;
    L.Sk = 0;
    L.steps = 0;
    if (ft == 1) { L.st = ST_DONE; return; }
    if (ft == 2) {                                 // :98-104
        L.NP = 2;
        L.Yt = 1.01;
        split_start(f, L, io, false);
        return;
    }
    if (io.sta_route && io.sta_route[cell] == 2) {   // unstable by the stability stages
        L.cnt[0] = io.rs_cnt[cell];
        L.cnt[1] = io.rs_cnt[n + cell];
        L.cnt[5] = io.rs_cnt[2 * n + cell];
        L.cnt[6] = L.cnt[7] = io.rs_cnt[3 * n + cell];
        L.Yt = io.rs_Yt[cell];
        L.NP = 2;                                  // strict_decide's unstable branch
        L.stage = 0;
        L.restarts = 0;
        split_start(f, L, io, false);
        return;
    }
    if (io.sta_route && io.sta_route[cell] == 3) {   // SplitSSM2's first pass done by the split stage
        L.cnt[0] = io.rs_cnt[cell];
        L.cnt[1] = io.rs_cnt[n + cell];
        L.cnt[5] = io.rs_cnt[2 * n + cell];
        L.cnt[6] = L.cnt[7] = io.rs_cnt[3 * n + cell];
        L.cnt[2] = io.rs_cnt[4 * n + cell];
        L.cnt[4] = io.rs_cnt[5 * n + cell];
        L.ssm_conv = io.rs_cnt[6 * n + cell];
        L.Yt = io.rs_Yt[cell];
        L.NP = 2;
        L.stage = 0;
        L.restarts = 0;
        L.nu0 = io.rs_nu[cell];
        double l1[NC];
        UNROLL for (int i = 0; i < NC; i++) {
            L.a2[i] = io.rs_v[(size_t)i * n + cell];
            L.a3[i] = io.rs_v[(size_t)(NC + i) * n + cell];
            L.a4[i] = io.rs_v[(size_t)(2 * NC + i) * n + cell];
            L.a5[i] = io.rs_v[(size_t)(3 * NC + i) * n + cell];
            l1[i] = io.rs_v[(size_t)(4 * NC + i) * n + cell];
        }
        split_nr_begin(f, L, io, l1);
        return;
    }
    if (io.sta_route && io.sta_route[cell] == 4) {   // SplitNR did not converge in the split-Newton stage
        L.cnt[0] = io.rs_cnt[cell];
        L.cnt[1] = io.rs_cnt[n + cell];
        L.cnt[5] = io.rs_cnt[2 * n + cell];
        L.cnt[6] = io.rs_cnt[3 * n + cell];
        L.cnt[2] = io.rs_cnt[4 * n + cell];
        L.cnt[4] = io.rs_cnt[5 * n + cell];
        L.cnt[3] = io.rs_cnt[7 * n + cell];
        L.cnt[7] = L.cnt[6] - L.cnt[3];
        L.ssm_conv = io.rs_cnt[6 * n + cell];
        L.Yt = io.rs_Yt[cell];
        L.NP = 2;
        L.stage = 0;
        L.restarts = 0;
        L.nr_conv = 0;
        UNROLL for (int i = 0; i < NC; i++) L.a4[i] = io.rs_v[(size_t)(2 * NC + i) * n + cell];
        after_split_nr(f, L, io);                  // the restart loop (:558-566)
        return;
    }
    UNROLL for (int i = 0; i < NC; i++) L.y[i] = L.z[i];
    L.st = ST_FEED;
}

template <int NC>
OCPX_HD void pe_end(const PeIO& io, PeLane<NC>& L)
{
    const size_t n = io.nb;
    const int c = L.cell;
    if (L.NP > 1) L.ftype = 2;                     // :111
    io.pe_NP[c] = L.NP;
    io.pe_nu[c] = L.NP > 1 ? L.nu0 : 1.0;
    UNROLL for (int i = 0; i < NC; i++) {
        io.pe_x[(size_t)i * n + c] = L.NP > 1 ? L.a2[i] : L.z[i];
        io.pe_x[(size_t)(NC + i) * n + c] = L.NP > 1 ? L.a3[i] : 0.0;
    }
    io.ftype_out[c] = L.ftype;
    if (io.iters)
        UNROLL for (int q = 0; q < 8; q++) io.iters[(size_t)q * n + c] = L.cnt[q];
}

// --------------------------------------------------------------- one step
template <int NC>
OCPX_HD void pe_step(const DevFluid& f, const PeIO& io, PeLane<NC>& L)
{
    // every cell ends: a legitimate flash needs far fewer steps than this
    if (++L.steps > MAX_STEPS) { L.cnt[7] |= 0x20000000; L.st = ST_DONE; return; }
    // mixture terms and ln phi of the composition evaluated in this step
    // (step-local, so they are not carried in registers across steps)
    EosMix<NC> mx;
    double ev[NC];
    // PRE: state-specific preparation of y (and the NR solves)
    bool eval = true;
    switch (L.st) {
    case ST_T_START: {
        const double ck = trial_coef(L.k);
        UNROLL for (int i = 0; i < NC; i++) L.y[i] = L.z[i] / exp(ck * log_kw0(f, L, i));
        normalise_y(L);
        break;
    }
    case ST_T_SSM:   // Y = di / phiSta (:439-444), a2 = ln phiSta
        UNROLL for (int i = 0; i < NC; i++) L.y[i] = L.a1[i] * exp(-L.a2[i]);
        normalise_y(L);
        break;
    case ST_A_ITER: {   // Y = x ks, dY against the previous Y (:330-346)
        UNROLL for (int i = 0; i < NC; i++) L.y[i] = L.z[i] * L.a4[i];
        normalise_y(L);
        double dY = 0;
        UNROLL for (int i = 0; i < NC; i++) dY += (L.y[i] - L.a3[i]) * (L.y[i] - L.a3[i]);
        if (dY < DYTOL) { L.flag = 1; eval = false; }
        break;
    }
    case ST_A_START: {  // ks = Kw[k] (:324)
        const double ck = trial_coef(L.k);
        UNROLL for (int i = 0; i < NC; i++) {
            L.a4[i] = exp(ck * log_kw0(f, L, i));
            L.y[i] = L.z[i] * L.a4[i];
        }
        normalise_y(L);
        L.iter = 0;
        L.flag = 0;
        break;
    }
    case ST_T_NR: {   // Jacobian (:523-541) solved through its lower triangle (dsysv_ 'U', :493)
        eos_mix<NC>(f, L.ec, L.y, mx);              // mixture at the current Y (as evaluated last step)
        double Zx[NC];
        const double den = z_den(mx), bc = z_bcoef(mx);
        UNROLL for (int k = 0; k < NC; k++) Zx[k] = ((mx.B - mx.Z) * (2 * mx.s[k]) + bc * L.ec.Bi[k]) / den;
        double S[NC][NC], rhs[1][NC];
        UNROLL for (int i = 0; i < NC; i++) {
            double row[NC];
            lnfugx_row<NC>(f, L.ec, mx, Zx, L.y, L.P, i, row);
            double t = 0;
            UNROLL for (int k = 0; k < NC; k++) t += L.y[k] * row[k];
            UNROLL for (int j = 0; j <= i; j++) {
                const double v = (row[j] - t + 1) / L.Yt;
                S[i][j] = v;
                S[j][i] = v;
            }
            rhs[0][i] = L.a3[i];
        }
        gauss_solve<NC, 1>(S, rhs);
        L.cnt[6]++;
        L.cnt[7]++;
        const double lYt = L.Yt;
        UNROLL for (int i = 0; i < NC; i++) L.y[i] = L.y[i] * lYt + rhs[0][i];
        normalise_y(L);
        break;
    }
    case ST_S_SSM_A:
        rr2(f, L);
        update_x_rr(L);
        UNROLL for (int i = 0; i < NC; i++) L.y[i] = L.a2[i];
        break;
    case ST_S_SSM_B:
        UNROLL for (int i = 0; i < NC; i++) L.y[i] = L.a3[i];
        break;
    case ST_S_NR: {   // :826-868
        const double nu1 = 1 - L.nu0;
        double J[NC][NC], rhs[1][NC];
        {
            EosMix<NC> m0, m1;
            LnfN<NC> d0, d1;
            eos_mix<NC>(f, L.ec, L.a2, m0);
            eos_mix<NC>(f, L.ec, L.a3, m1);
            lnfn_prep<NC>(L.ec, m0, L.nu0, d0);
            lnfn_prep<NC>(L.ec, m1, nu1, d1);
            UNROLL for (int i = 0; i < NC; i++) {
                UNROLL for (int k = 0; k <= i; k++) {
                    const double v = lnfn_entry<NC, false>(f, L.ec, m0, d0, L.a2, L.P, L.nu0, i, k) +
                                     lnfn_entry<NC, false>(f, L.ec, m1, d1, L.a3, L.P, nu1, i, k);
                    J[i][k] = v;
                    J[k][i] = v;
                }
                rhs[0][i] = L.a5[i];
            }
        }
        L.cnt[6] += 2;
        L.cnt[7]++;
        gauss_solve<NC, 1>(J, rhs);
        double alpha = 1;                          // CalStepNRsp :939-957
        UNROLL for (int i = 0; i < NC; i++)
            if (L.a1[i] + alpha * rhs[0][i] < 0) alpha *= 0.9 * fabs(L.a1[i] / rhs[0][i]);
        double nu0 = 0, en = 0;
        UNROLL for (int i = 0; i < NC; i++) {
            const double d = alpha * rhs[0][i];
            L.a1[i] += d;
            nu0 += L.a1[i];
            en += 2 * d * d;
        }
        L.nu0 = nu0;
        L.Sk = en;                                 // carry sum (ln - n)^2 to POST
        const double r0 = 1 / nu0, r1 = 1 / (1 - nu0);
        UNROLL for (int i = 0; i < NC; i++) {
            L.a2[i] = fabs(L.a1[i] * r0);
            L.a3[i] = fabs((L.z[i] - L.a1[i]) * r1);
            L.y[i] = L.a2[i];
        }
        break;
    }
    default:   // ST_FEED: y = z already
        break;
    }

    // COMMON EVALUATOR: ln phi at y
    if (eval) {
        eos_mix<NC>(f, L.ec, L.y, mx);
        eos_lnphi<NC>(L.ec, mx, ev);
        L.cnt[5]++;
    }

    // POST
    switch (L.st) {
    case ST_FEED:   // di = phi[Id] * x (:403-408), fug[Id] = di * P
        UNROLL for (int i = 0; i < NC; i++) L.a1[i] = exp(ev[i]) * L.z[i];
        L.k = 0;
        L.st = L.lNP == 0 ? ST_T_START : ST_A_START;   // PhaseStable :279-291
        break;
    case ST_T_START:
    case ST_T_SSM: {   // Se = sum ln(fugSta / fugId * Yt)^2 (:429-432, :448-451)
        // After an SSM update Y Yt = di / phiSta of the previous iterate, so
        // ln(fugSta / fugId * Yt) is the change of ln phiSta.
        double Se = 0;
        const bool first = L.st == ST_T_START;
        UNROLL for (int i = 0; i < NC; i++) {
            const double r = first ? log((exp(ev[i]) * L.y[i] * L.P) / (L.a1[i] * L.P) * L.Yt) : ev[i] - L.a2[i];
            Se += r * r;
            L.a3[i] = -r;                          // resSTA for StableNR
            L.a2[i] = ev[i];
        }
        if (L.st == ST_T_START) {
            L.iter = 0;
            if (Se > f.SSMsta_tol2) { L.st = ST_T_SSM; break; }
            nr_sta_begin(f, L, io);
            break;
        }
        L.cnt[0]++;
        if (Se > f.SSMsta_tol2 && ++L.iter <= f.SSMsta_maxIt) break;
        nr_sta_begin(f, L, io);
        break;
    }
    case ST_T_NR: {
        double e2 = 0;
        UNROLL for (int i = 0; i < NC; i++) {
            L.a3[i] = log((L.a1[i] * L.P) / ((exp(ev[i]) * L.y[i] * L.P) * L.Yt));
            e2 += L.a3[i] * L.a3[i];
        }
        L.cnt[1]++;
        if (++L.iter > f.NRsta_maxIt) { L.flag = 0; strict_decide(f, L, io); break; }
        if (sqrt(e2) > f.NRsta_tol) break;
        L.flag = 1;
        strict_decide(f, L, io);
        break;
    }
    case ST_A_START:
    case ST_A_ITER: {
        if (!eval) { relaxed_trial_end(f, L, io); break; }
        // di = fugId / (fugSta Yt), ks *= di, Se, Sk (:347-356)
        double Se = 0, Sk = 0;
        UNROLL for (int i = 0; i < NC; i++) {
            const double di = (L.a1[i] * L.P) / ((exp(ev[i]) * L.y[i] * L.P) * L.Yt);
            L.a4[i] *= di;
            Se += (di - 1) * (di - 1);
            const double lk = log(L.a4[i]);
            Sk += lk * lk;
            L.a3[i] = L.y[i];
            L.a2[i] = ev[i];
        }
        L.Sk = Sk;
        L.cnt[0]++;
        L.iter++;
        if (Se < f.SSMsta_tol2 || Sk < KTOL) { L.flag = 1; relaxed_trial_end(f, L, io); break; }
        if (L.iter > f.SSMsta_maxIt) { L.flag = 0; relaxed_trial_end(f, L, io); break; }
        L.st = ST_A_ITER;
        break;
    }
    case ST_S_SSM_A:
        UNROLL for (int i = 0; i < NC; i++) L.a5[i] = ev[i];
        L.st = ST_S_SSM_B;
        break;
    case ST_S_SSM_B: {
        double Se = 0;
        bool last;
        L.cnt[2]++;
        L.iter++;
        // literal forms of :629-635: fugacities phi*x*P of both phases, their
        // ratio, K <- phi1/phi0 (identical overflow behaviour included)
        double Knew[NC];
        UNROLL for (int i = 0; i < NC; i++) {
            const double phi0 = exp(L.a5[i]), phi1 = exp(ev[i]);
            const double q = (phi1 * L.a3[i] * L.P) / (phi0 * L.a2[i] * L.P) - 1;
            Se += q * q;
            Knew[i] = phi1 / phi0;
        }
        if (L.iter > L.maxIt) { L.ssm_conv = 0; last = true; }
        else last = !(Se > L.tol2);
        UNROLL for (int i = 0; i < NC; i++) L.a4[i] = Knew[i];
        // SplitNR starts from the x0, x1 just evaluated: ln(f1/f0) from
        // ln phi1 = ev and ln phi0 = a5
        if (last) split_nr_begin(f, L, io, ev);
        else L.st = ST_S_SSM_A;
        break;
    }
    case ST_S_NR: {
        double lnphi0[NC];
        UNROLL for (int i = 0; i < NC; i++) { lnphi0[i] = ev[i]; L.y[i] = L.a3[i]; }
        eos_mix<NC>(f, L.ec, L.y, mx);
        eos_lnphi<NC>(L.ec, mx, ev);
        L.cnt[5]++;
        double e2 = 0;
        UNROLL for (int i = 0; i < NC; i++) {
            L.a5[i] = log((exp(ev[i]) * L.a3[i] * L.P) / (exp(lnphi0[i]) * L.a2[i] * L.P));
            e2 += L.a5[i] * L.a5[i];
        }
        const double eNR = sqrt(e2);
        L.cnt[3]++;
        L.iter++;
        bool end = false;
        if (eNR > L.eNR0 || L.iter > f.NRsp_maxIt) end = true;
        else if (L.Sk < 1E-16 * (2.0 * NC) * (2.0 * NC)) { L.nr_conv = 1; end = true; }
        else if (!(eNR > f.NRsp_tol)) end = true;
        if (end) {
            if (eNR < f.NRsp_tol) L.nr_conv = 1;
            after_split_nr(f, L, io);
        }
        break;
    }
    default:
        break;
    }
}

// --------------------------------------------------------------- kernels
// The fluid constants are loop-invariant; left alone, the compiler hoists
// every one of them (Nc^2 interaction terms included) out of the step loop
// into registers.  Passing the pointer through an empty asm each step keeps
// them as uniform scalar loads at their point of use.
OCPX_HD const DevFluid* opaque(const DevFluid* p)
{
    asm volatile("" : "+s"(p));
    return p;
}

template <int NC>
__global__ void __launch_bounds__(256) pe_plain(const DevFluid* __restrict__ fp, PeIO io)
{
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= io.nb) return;
    const DevFluid& f = *fp;
    PeLane<NC> L;
    pe_begin<NC>(f, io, cell, L);
    while (L.st != ST_DONE) pe_step<NC>(*opaque(fp), io, L);
    pe_end<NC>(io, L);
}

// Persistent lanes + global work queue (counter at io.queue, zeroed before
// the launch).  A lane whose cell is done writes it back and takes the next
// cell index; the wavefront exits when the queue is empty and every lane is
// done.  Every cell's state machine is bounded, so every lane reaches the
// exit.
template <int NC>
__global__ void __launch_bounds__(256) pe_batched(const DevFluid* __restrict__ fp, PeIO io)
{
    const DevFluid& f = *fp;
    PeLane<NC> L;
    L.st = ST_DONE;
    bool have = false, drained = false;
    const int lane = __lane_id();
    const unsigned nq = io.nq ? *io.nq : (unsigned)io.nb;
    while (true) {
        const bool need = (L.st == ST_DONE) && !drained;
        if (L.st == ST_DONE && have) { pe_end<NC>(io, L); have = false; }
        const unsigned long long m = __ballot(need);
        if (m) {
            const int leader = __ffsll((long long)m) - 1;
            unsigned base = 0;
            if (lane == leader) base = atomicAdd(io.queue, (unsigned)__popcll(m));
            base = __shfl(base, leader);
            if (need) {
                const unsigned idx = base + (unsigned)__popcll(m & ((1ull << lane) - 1ull));
                if (idx < nq) {
                    pe_begin<NC>(f, io, io.qorder ? io.qorder[idx] : (int)idx, L);
                    have = true;
                }
                else drained = true;
            }
        }
        if (!__any(have)) break;
        // Newton steps (matrix build + solve) cost several SSM steps; a lane
        // that wants one waits until NR_BATCH lanes of the wavefront do (or
        // nothing else is left to run), so the wavefront pays for the Newton
        // branch once per batch instead of on nearly every step.
        const bool live = have && L.st != ST_DONE;
        const bool nr = live && (L.st == ST_T_NR || L.st == ST_S_NR);
        const unsigned long long mnr = __ballot(nr), mlive = __ballot(live);
        const bool run_nr = __popcll(mnr) >= NR_BATCH || mnr == mlive;
        if (live && (!nr || run_nr)) pe_step<NC>(*opaque(fp), io, L);
    }
}

template <int NC>
static hipError_t launch_pe(const DevFluid* f, const PeIO& io, int mode, int grid_waves,
                            hipStream_t st)
{
    if (io.nb == 0) return hipSuccess;
    if (mode == 0) {
        const int bs = 64;
        const int nblk = (int)((io.nb + bs - 1) / bs);
        hipLaunchKernelGGL(pe_plain<NC>, dim3(nblk), dim3(bs), 0, st, f, io);
    } else {
        const int bs = 64;
        int nblk = grid_waves;
        const long need = (io.nb + bs - 1) / bs;
        if (nblk > need) nblk = (int)need;
        hipLaunchKernelGGL(pe_batched<NC>, dim3(nblk), dim3(bs), 0, st, f, io);
    }
    return hipGetLastError();
}

// ------------------------------------------------------------ queue order
// Longest-first order for the next batched flash: a counting sort of the
// cells by their fugacity-evaluation count in the last flash (counts at or
// above QB - 1 share the first bucket).  Cells with equal counts sit next to
// each other, so a wavefront takes cells that tend to follow the same path,
// and the long cells start first instead of extending the tail.  The order
// inside a bucket is the atomics' order; it only schedules work.
constexpr int QB = 4096;

__device__ __forceinline__ int qbucket(int32_t e) { return QB - 1 - (e < 0 ? 0 : (e > QB - 1 ? QB - 1 : e)); }

// Counts go through a block histogram in LDS first, so the global atomics
// are one per block and bucket instead of one per cell (most cells share a
// handful of counts).
constexpr int QTILE = 2048;                      // cells per block

__global__ void __launch_bounds__(256) qorder_hist(const int32_t* __restrict__ evals, int64_t nb, unsigned* hist)
{
    __shared__ unsigned sh[QB];
    const int t = threadIdx.x;
    for (int b = t; b < QB; b += 256) sh[b] = 0;
    __syncthreads();
    const int64_t c0 = (int64_t)blockIdx.x * QTILE;
    for (int k = t; k < QTILE; k += 256) {
        const int64_t i = c0 + k;
        if (i < nb) atomicAdd(&sh[qbucket(evals[i])], 1u);
    }
    __syncthreads();
    for (int b = t; b < QB; b += 256)
        if (sh[b]) atomicAdd(&hist[b], sh[b]);
}

// exclusive prefix sum of hist[QB] in place, one block of 256 threads
__global__ void __launch_bounds__(256) qorder_scan(unsigned* hist)
{
    constexpr int PER = QB / 256;
    __shared__ unsigned part[256];
    const int t = threadIdx.x;
    unsigned loc[PER], s = 0;
    UNROLL for (int k = 0; k < PER; k++) { loc[k] = hist[t * PER + k]; s += loc[k]; }
    part[t] = s;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {          // inclusive scan of the partial sums
        const unsigned v = t >= o ? part[t - o] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    unsigned run = part[t] - s;
    UNROLL for (int k = 0; k < PER; k++) { hist[t * PER + k] = run; run += loc[k]; }
}

// each block reserves one range per bucket it holds, then ranks its cells
// inside the range with LDS atomics
__global__ void __launch_bounds__(256) qorder_scatter(const int32_t* __restrict__ evals, int64_t nb, unsigned* offs,
                                                      int32_t* __restrict__ order)
{
    __shared__ unsigned cnt[QB], base[QB];
    const int t = threadIdx.x;
    for (int b = t; b < QB; b += 256) cnt[b] = 0;
    __syncthreads();
    const int64_t c0 = (int64_t)blockIdx.x * QTILE;
    for (int k = t; k < QTILE; k += 256) {
        const int64_t i = c0 + k;
        if (i < nb) atomicAdd(&cnt[qbucket(evals[i])], 1u);
    }
    __syncthreads();
    for (int b = t; b < QB; b += 256) {
        base[b] = cnt[b] ? atomicAdd(&offs[b], cnt[b]) : 0u;
        cnt[b] = 0;
    }
    __syncthreads();
    for (int k = t; k < QTILE; k += 256) {
        const int64_t i = c0 + k;
        if (i < nb) {
            const int b = qbucket(evals[i]);
            order[base[b] + atomicAdd(&cnt[b], 1u)] = (int32_t)i;
        }
    }
}

int queue_order_buckets() { return QB; }

hipError_t run_queue_order(const int32_t* evals, int64_t nb, unsigned* hist, int32_t* order, hipStream_t st)
{
    if (nb == 0) return hipSuccess;
    const unsigned g = (unsigned)((nb + QTILE - 1) / QTILE);
    hipLaunchKernelGGL(qorder_hist, dim3(g), dim3(256), 0, st, evals, nb, hist);
    hipLaunchKernelGGL(qorder_scan, dim3(1), dim3(256), 0, st, hist);
    hipLaunchKernelGGL(qorder_scatter, dim3(g), dim3(256), 0, st, evals, nb, hist, order);
    return hipGetLastError();
}

hipError_t run_pe(int NC, const DevFluid* f, const PeIO& io, int mode, int grid_waves, hipStream_t st)
{
    switch (NC) {
#define OCPX_CASE(N) case N: return launch_pe<N>(f, io, mode, grid_waves, st);
        OCPX_CASE(2) OCPX_CASE(3) OCPX_CASE(4) OCPX_CASE(5) OCPX_CASE(6) OCPX_CASE(7)
        OCPX_CASE(8) OCPX_CASE(9) OCPX_CASE(10) OCPX_CASE(11) OCPX_CASE(12)
#undef OCPX_CASE
    default: return hipErrorInvalidValue;
    }
}

}  // namespace ocpx

namespace ocpx {
// resident 64-lane blocks of pe_batched per CU, for sizing its persistent grid
int pe_batched_blocks_per_cu(int NC)
{
    int nblk = 0;
    hipError_t e = hipErrorInvalidValue;
    switch (NC) {
#define OCPX_CASE(N) case N: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nblk, pe_batched<N>, 64, 0); break;
        OCPX_CASE(2) OCPX_CASE(3) OCPX_CASE(4) OCPX_CASE(5) OCPX_CASE(6) OCPX_CASE(7)
        OCPX_CASE(8) OCPX_CASE(9) OCPX_CASE(10) OCPX_CASE(11) OCPX_CASE(12)
#undef OCPX_CASE
    default: break;
    }
    return e == hipSuccess ? nblk : 0;
}
}  // namespace ocpx
