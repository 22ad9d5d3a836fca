// pe_team.hip — phase equilibrium with a team of T lanes per cell.
//
// Same algorithm and state machine as pe.hip (OCPPhaseEquilibrium::
// PhaseEquilibrium, src/OCPPhaseEquilibrium.cpp:71-112: stability by
// StableSSM/StableNR or StableSSM01, split by SplitSSM2 + RachfordRice2 +
// SplitNR + CheckSplit), mapped differently: lane t of a team owns component
// t, so a cell's state is a handful of registers per lane instead of
// ~7 Nc doubles in one lane.  That keeps Nc = 5..16 free of private-memory
// spills and leaves room for several waves per SIMD.
//
//   * sums over components are butterfly reductions inside the team
//     (every lane ends with the same bits, so control flow stays uniform
//     across the team);
//   * the mixture term s_i = sqrt(Ai) sum_k (1-k_ik) sqrt(Ak) x_k takes the
//     other lanes' sqrt(Ak) x_k by shuffle;
//   * Newton systems are held one row per lane and solved by Gaussian
//     elimination with partial pivoting across the team (pivot search by
//     butterfly argmax, pivot row by shuffle, column-oriented back
//     substitution).
//
// A wavefront holds 64 / T teams.  pe_team_plain gives each team one cell;
// pe_team_batched keeps teams resident and refills them from the global
// work queue as their cells finish.
#include "ocpx_dev.hpp"
#include "ocpx_kernels.hpp"

namespace ocpx {
namespace team {

constexpr int MAX_SPLIT_RESTARTS = 64;
constexpr int MAX_STEPS = 200000;
constexpr double KTOL = 1E-4, DYTOL = 1E-6;   // include/OCPPhaseEquilibrium.hpp:43-44

enum : int {
    ST_DONE = 0, ST_FEED, ST_T_START, ST_T_SSM, ST_T_NR, ST_A_START, ST_A_ITER,
    ST_S_SSM_A, ST_S_SSM_B, ST_S_NR,
};

#define DEV __device__ __forceinline__

// ----------------------------------------------------------- team helpers
template <int T>
DEV double tsum(double v)
{
    UNROLL for (int o = T / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, T);
    return v;
}
template <int T>
DEV double tmin(double v)
{
    UNROLL for (int o = T / 2; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, T));
    return v;
}
template <int T>
DEV double tmax(double v)
{
    UNROLL for (int o = T / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, T));
    return v;
}
template <int T>
DEV double bc(double v, int k) { return __shfl(v, k, T); }

// Fluid constants of component t, loaded when the team takes a cell.
template <int NC>
struct LaneFluid {
    double Tc, Pc, logPc, mw, OA, OB, kwA;
    double omk[NC];          // 1 - k_tk
};

template <int NC>
DEV void lane_fluid(const DevFluid& f, int t, LaneFluid<NC>& lf)
{
    const int i = t < NC ? t : 0;
    lf.Tc = f.Tc[i]; lf.Pc = f.Pc[i]; lf.logPc = f.logPc[i]; lf.mw = f.mw[i];
    lf.OA = f.OA[i]; lf.OB = f.OB[i]; lf.kwA = f.kwA[i];
    UNROLL for (int k = 0; k < NC; k++) lf.omk[k] = f.omk[i * OCPX_NCMAX + k];
}

template <int NC, int T>
struct Cell {
    int cell, model, st, k, iter, maxIt, NP, ftype, flag, stage, restarts, ssm_conv, nr_conv, steps;
    uint32_t lNP;
    double P, Tk, logP, Yt, tol2, Sk, nu0, eNR0;
    int cnt[8];
    // this lane's component
    double z, y, a1, a2, a3, a4, a5;
    double sA, Bi;              // sqrt(Ai), Bi at (P, T)
    double s, A, B, Z;          // mixture at the last evaluated y (s of this component)
};

// ------------------------------------------------------------- EoS pieces
template <int NC, int T>
DEV void mix(const LaneFluid<NC>& lf, const Cell<NC, T>& C, bool act, double x, double& s,
             double& A, double& B, double& Z)
{
    const double w = act ? C.sA * x : 0.0;
    double t = 0;
    UNROLL for (int k = 0; k < NC; k++) t += lf.omk[k] * bc<T>(w, k);
    s = C.sA * t;
    A = tsum<T>(act ? x * s : 0.0);
    B = tsum<T>(act ? C.Bi * x : 0.0);
    Z = pr_z(A, B);
}

DEV double lnphi_of(double s, double Bi, double A, double B, double Z)
{
    const double lz = log(Z - B);
    const double lg = log((Z + D1 * B) / (Z + D2 * B));
    const double cA = A / (D1 - D2) / B * lg;
    const double iB = 1 / B;
    return Bi * iB * (Z - 1) - lz - cA * (2 * s / A - Bi * iB);
}

DEV double zden(double A, double B, double Z)
{
    return 3 * Z * Z + 2 * ((D1 + D2 - 1) * B - 1) * Z + (A + D1 * D2 * B * B - (D1 + D2) * B * (B + 1));
}
DEV double zbcoef(double A, double B, double Z)
{
    return (A + D1 * D2 * (3 * B * B + 2 * B)) + ((D1 + D2) * (2 * B + 1) - 2 * D1 * D2 * B) * Z -
           (D1 + D2 - 1) * Z * Z;
}

// Entry (i,k) of d ln f_i / d n_k for one phase (src/OCPEoS.cpp:254-300);
// i-quantities of this lane, k-quantities broadcast.
DEV double lnfn_ik(double A, double B, double Z, double lG, double G, double P, double nt,
                   double si, double Bii, double xi, double aik, bool diag, double Bik, double Ank,
                   double Bnk, double Znk)
{
    const double ZB = Z - B;
    const double E = -A / ((D1 - D2) * B) * (2 * si / A - Bii / B);
    const double C = xi * P / ZB;
    const double Cn = P / ZB / ZB * (ZB / nt * ((diag ? 1.0 : 0.0) - xi) - xi * (Znk - Bnk));
    const double Dn = Bii / B * (Znk - (Bik - B) * (Z - 1) / (nt * B));
    const double Gn = (D1 - D2) / ((Z + D2 * B) * (Z + D2 * B)) * (Bnk * Z - Znk * B);
    double En = -1 / (D1 - D2) / (B * B) * (2 * (B * aik / nt + Bnk * (Bii * A / B - si)) - Ank * Bii - A * Bii / nt);
    En -= E / nt;
    return 1 / C * Cn + Dn + En * lG + E / G * Gn;
}

// Row t of a symmetric matrix whose lower triangle (entries k <= t) this lane
// computed: entries k > t are taken from lane k's row, element t.
template <int N, int T>
DEV void mirror_lower(double (&M)[N], int t)
{
    double out[N];
    UNROLL for (int k = 0; k < N; k++) out[k] = M[k];
    UNROLL for (int d = 1; d < N; d++) {
        double u = 0;
        UNROLL for (int k = 0; k < N; k++) u = (k == t - d) ? M[k] : u;
        const double v = bc<T>(u, t + d);
        UNROLL for (int k = 0; k < N; k++) out[k] = (k == t + d) ? v : out[k];
    }
    UNROLL for (int k = 0; k < N; k++) M[k] = out[k];
}

// Solve M x = b, row r of M and b[r] on lane r < N (partial pivoting).
// Returns x_t on lane t < N.
template <int N, int T>
DEV double team_solve(double (&M)[N], double b, int t)
{
    bool used = t >= N;
    int ord = N;
    int piv[N];
    UNROLL for (int c = 0; c < N; c++) {
        double v = used ? -1.0 : fabs(M[c]);
        v = (v == v) ? v : -0.5;                         // NaN rows pivot last
        int p = t;
        UNROLL for (int o = T / 2; o > 0; o >>= 1) {
            const double v2 = __shfl_xor(v, o, T);
            const int p2 = __shfl_xor(p, o, T);
            if (v2 > v || (v2 == v && p2 < p)) { v = v2; p = p2; }
        }
        piv[c] = p;
        const double id = 1 / bc<T>(M[c], p);
        const double bp = bc<T>(b, p);
        double prow[N];
        UNROLL for (int k = c + 1; k < N; k++) prow[k] = bc<T>(M[k], p);
        if (t == p) { used = true; ord = c; }
        else if (!used) {
            const double l = M[c] * id;
            UNROLL for (int k = c + 1; k < N; k++) M[k] -= l * prow[k];
            b -= l * bp;
        }
    }
    double sol = 0;
    UNROLL for (int c = N - 1; c >= 0; c--) {
        const int p = piv[c];
        const double xc = bc<T>(b, p) / bc<T>(M[c], p);
        if (t == c) sol = xc;
        if (ord < c) b -= M[c] * xc;
    }
    return sol;
}

// ------------------------------------------------------------ algorithm
DEV double trial_coef(int k)
{
    return k == 0 ? 1.0 : (k == 1 ? -1.0 : (k == 2 ? (1.0 / 3) : -(1.0 / 3)));
}

template <int NC>
DEV double log_kw0(const LaneFluid<NC>& lf, double logP, double Tk)
{
    return lf.logPc - logP + lf.kwA * (1 - lf.Tc / Tk);
}

template <int NC, int T>
DEV void normalise_y(Cell<NC, T>& C, bool act)
{
    C.Yt = tsum<T>(act ? C.y : 0.0);
    C.y *= 1 / C.Yt;
}

template <int NC, int T>
DEV void rr2(const DevFluid& f, Cell<NC, T>& C, bool act)
{   // RachfordRice2 :651-705 on K = a4
    const double Kmin = tmin<T>(act ? C.a4 : INFINITY);
    const double Kmax = tmax<T>(act ? C.a4 : -INFINITY);
    const double numin = 1 / (1 - Kmax), numax = 1 / (1 - Kmin);
    double nu = 0.5 * (numin + numax);
    int it = 0;
    const double km = C.a4 - 1;
    while (true) {
        const double t = 1 / (1 + nu * km);
        const double rj = tsum<T>(act ? C.z * km * t : 0.0);
        const double J = -tsum<T>(act ? C.z * km * km * t * t : 0.0);
        if (fabs(rj) < f.RR_tol || it > f.RR_maxIt) break;
        const double d = -rj / J, tn = nu + d;
        if (tn < numax && tn > numin) nu = tn;
        else nu = d > 0 ? (nu + numax) / 2 : (nu + numin) / 2;
        it++;
    }
    C.cnt[4] += it;
    C.nu0 = nu;
}

template <int NC, int T>
DEV void update_x_rr(Cell<NC, T>& C)
{   // UpdateXRR :778-791
    C.a3 = C.z / (1 + C.nu0 * (C.a4 - 1));
    C.a2 = C.a4 * C.a3;
}

template <int NC, int T>
DEV void split_start(const DevFluid& f, const LaneFluid<NC>& lf, Cell<NC, T>& C, const PeIO& io,
                     bool restart, int t)
{   // SplitSSM2 :586-622
    C.ssm_conv = 1;
    C.tol2 = f.SSMsp_tol2;
    C.maxIt = f.SSMsp_maxIt;
    if (!restart) {
        if (C.lNP == 2) {
            const size_t n = io.nb;
            const int nc = NC + 1;
            const int i = t < NC ? t : 0;
            C.a4 = io.xij[(size_t)i * n + C.cell] / io.xij[(size_t)(nc + i) * n + C.cell];
        } else {
            C.a4 = exp(log_kw0(lf, C.logP, C.Tk));
        }
    } else {
        C.tol2 *= 1E-1;
        C.maxIt *= 2;
    }
    if (C.Yt < 1.1) {
        C.tol2 *= 1E-1;
        C.maxIt *= 2;
    }
    C.iter = 0;
    C.st = ST_S_SSM_A;
}

template <int NC, int T>
DEV bool check_split(Cell<NC, T>& C, bool act)
{   // CheckSplit :960-1010
    if (C.NP != 2) return true;
    const double nu1 = 1 - C.nu0;
    const double nuMax = fmax(C.nu0, nu1);
    const double eX = tsum<T>(act ? (C.a2 - C.a3) * (C.a2 - C.a3) : 0.0);
    if (nuMax < 1 && C.nr_conv && finite_d(eX)) return true;
    if (!finite_d(eX) || 1 - nuMax < 1E-3) {
        C.NP = 1;
        C.nu0 = 1;
        C.a2 = C.z;
        return false;
    }
    return true;
}

template <int NC, int T>
DEV void finish_split(Cell<NC, T>& C, bool act)
{
    check_split(C, act);
    C.st = ST_DONE;
}

template <int NC, int T>
DEV void after_split_nr(const DevFluid& f, const LaneFluid<NC>& lf, Cell<NC, T>& C, const PeIO& io,
                        bool act, int t)
{   // PhaseSplit restart loop :558-566
    if (C.stage == 0) {
        if (C.nr_conv) { finish_split(C, act); return; }
        C.stage = 1;
    } else {
        if (!check_split(C, act)) { finish_split(C, act); return; }
        if (C.ssm_conv || C.nr_conv) { finish_split(C, act); return; }
    }
    if (++C.restarts > MAX_SPLIT_RESTARTS) { C.cnt[7] |= 0x40000000; finish_split(C, act); return; }
    split_start(f, lf, C, io, true, t);
}

template <int NC, int T>
DEV void split_nr_begin(const DevFluid& f, const LaneFluid<NC>& lf, Cell<NC, T>& C, const PeIO& io,
                        double lnphi1, bool act, int t)
{   // SplitNR :807-823; a5 = ln phi0 of x0
    C.a1 = C.nu0 * C.a2;
    C.a5 = log((exp(lnphi1) * C.a3 * C.P) / (exp(C.a5) * C.a2 * C.P));
    const double eNR = sqrt(tsum<T>(act ? C.a5 * C.a5 : 0.0));
    C.eNR0 = eNR;
    C.iter = 0;
    C.nr_conv = 0;
    if (eNR > f.NRsp_tol) { C.st = ST_S_NR; return; }
    C.nr_conv = eNR < f.NRsp_tol;
    after_split_nr(f, lf, C, io, act, t);
}

template <int NC, int T>
DEV void strict_decide(const DevFluid& f, const LaneFluid<NC>& lf, Cell<NC, T>& C, const PeIO& io, int t)
{   // :460-470
    if (C.flag && C.Yt > 1 + f.eYt) {
        C.NP = 2;
        C.stage = 0;
        C.restarts = 0;
        split_start(f, lf, C, io, false, t);
        return;
    }
    if (++C.k < 4) { C.st = ST_T_START; return; }
    C.st = ST_DONE;
}

template <int NC, int T>
DEV void nr_sta_begin(const DevFluid& f, const LaneFluid<NC>& lf, Cell<NC, T>& C, const PeIO& io,
                      bool act, int t)
{   // StableNR :478-488, resSTA in a3
    const double e2 = tsum<T>(act ? C.a3 * C.a3 : 0.0);
    C.iter = 0;
    if (sqrt(e2) > f.NRsta_tol) { C.st = ST_T_NR; return; }
    C.flag = 1;
    strict_decide(f, lf, C, io, t);
}

template <int NC, int T>
DEV void relaxed_trial_end(const DevFluid& f, const LaneFluid<NC>& lf, Cell<NC, T>& C, const PeIO& io,
                           int t)
{   // :377-392 and PhaseStable :283-291
    if (C.flag && C.Yt > 1 - 0.1 && C.Sk > 1) C.ftype = 2;
    bool stable;
    if (C.flag && C.Yt > 1 + f.eYt) stable = false;
    else if (++C.k < 2) { C.st = ST_A_START; return; }
    else stable = true;
    const uint32_t tmpNP = 1u + (stable ? 0u : 1u);
    if (tmpNP != C.lNP) {
        C.ftype = 2;
        C.k = 0;
        C.st = ST_T_START;
        return;
    }
    if (stable) { C.st = ST_DONE; return; }
    C.NP = 2;
    C.stage = 0;
    C.restarts = 0;
    split_start(f, lf, C, io, false, t);
}

// ------------------------------------------------------------ begin / end
template <int NC, int T>
DEV void begin(const DevFluid& f, LaneFluid<NC>& lf, const PeIO& io, int cell, Cell<NC, T>& C,
               bool act, int t)
{
    const size_t n = io.nb;
    const int i = act ? t : 0;
    C.cell = cell;
    lane_fluid<NC>(f, t, lf);
    C.P = io.P[cell];
    C.Tk = io.T[cell] + CONV4;
    C.logP = log(C.P);
    const double Ni = act ? io.Ni[(size_t)i * n + cell] : 0.0;
    const double Nt = tsum<T>(Ni);
    C.z = act ? Ni / Nt : 0.0;
    uint32_t phaseNum = 0;
    UNROLL for (int j = 0; j < 3; j++) phaseNum += io.phaseExist[(size_t)j * n + cell] != 0;
    C.lNP = phaseNum - 1u;
    int ft = io.ftype_in ? io.ftype_in[cell] : -1;
    if (ft < 0) {
        ft = 0;
        if (f.skip) {
            bool skip = false;
            if (io.skip_flag[cell]) {
                const double me = io.skip_minEig[cell];
                skip = fabs(1 - io.skip_P[cell] / C.P) < me / 10 && fabs(io.skip_T[cell] - C.Tk) < me * 10;
                const bool mine = !act || fabs(Ni / Nt - io.skip_zi[(size_t)i * n + cell]) < me / 10;
                skip = skip && tmin<T>(mine ? 1.0 : 0.0) > 0.5;
            }
            const uint32_t npPE = phaseNum - 1u;
            if (skip) ft = 1;
            else if (npPE >= 2) {
                uint32_t s = (io.S[cell] >= 1E-4) + (io.S[n + cell] >= 1E-4);
                ft = s == npPE ? (int)npPE : 0;
            }
        }
    }
    C.ftype = ft;
    UNROLL for (int c = 0; c < 8; c++) C.cnt[c] = 0;
    // EoS constants of this component (src/OCPEoS.cpp:68-84)
    {
        const double Pr = C.P / lf.Pc, Tr = C.Tk / lf.Tc;
        const double q = 1 + lf.mw * (1 - sqrt(Tr));
        const double Ai = lf.OA * Pr / (Tr * Tr) * (q * q);
        C.sA = act ? sqrt(Ai) : 0.0;
        C.Bi = act ? lf.OB * Pr / Tr : 0.0;
    }
    C.NP = 1;
    C.nu0 = 1;
    C.k = 0;
    C.flag = 0;
    C.stage = 0;
    C.restarts = 0;
    C.Yt = 0;
    C.Sk = 0;
    C.steps = 0;
    C.a1 = C.a2 = C.a3 = C.a4 = C.a5 = 0;
    if (ft == 1) { C.st = ST_DONE; return; }
    if (ft == 2) {
        C.NP = 2;
        C.Yt = 1.01;
        split_start(f, lf, C, io, false, t);
        return;
    }
    C.y = C.z;
    C.st = ST_FEED;
}

template <int NC, int T>
DEV void end(const PeIO& io, const Cell<NC, T>& C, bool act, int t)
{
    const size_t n = io.nb;
    const int c = C.cell;
    const int NP = C.NP;
    if (act) {
        io.pe_x[(size_t)t * n + c] = NP > 1 ? C.a2 : C.z;
        io.pe_x[(size_t)(NC + t) * n + c] = NP > 1 ? C.a3 : 0.0;
    }
    if (t == 0) {
        io.pe_NP[c] = NP;
        io.pe_nu[c] = NP > 1 ? C.nu0 : 1.0;
        io.ftype_out[c] = NP > 1 ? 2 : C.ftype;            // :111
        if (io.iters)
            UNROLL for (int q = 0; q < 8; q++) io.iters[(size_t)q * n + c] = C.cnt[q];
    }
}

// ------------------------------------------------------------ one step
template <int NC, int T>
DEV void step(const DevFluid& f, const LaneFluid<NC>& lf, const PeIO& io, Cell<NC, T>& C, bool act, int t)
{
    if (++C.steps > MAX_STEPS) { C.cnt[7] |= 0x20000000; C.st = ST_DONE; return; }
    bool eval = true;
    switch (C.st) {
    case ST_T_START: {
        const double ck = trial_coef(C.k);
        C.y = act ? C.z / exp(ck * log_kw0(lf, C.logP, C.Tk)) : 0.0;
        normalise_y(C, act);
        break;
    }
    case ST_T_SSM:
        C.y = act ? C.a1 / C.a2 : 0.0;
        normalise_y(C, act);
        break;
    case ST_A_ITER: {
        C.y = C.z * C.a4;
        normalise_y(C, act);
        const double dY = tsum<T>(act ? (C.y - C.a3) * (C.y - C.a3) : 0.0);
        if (dY < DYTOL) { C.flag = 1; eval = false; }
        break;
    }
    case ST_A_START: {
        const double ck = trial_coef(C.k);
        C.a4 = act ? exp(ck * log_kw0(lf, C.logP, C.Tk)) : 0.0;
        C.y = C.z * C.a4;
        normalise_y(C, act);
        C.iter = 0;
        C.flag = 0;
        break;
    }
    case ST_T_NR: {   // Jacobian :523-541, symmetric from its lower triangle (dsysv_ 'U', :493)
        const double A = C.A, B = C.B, Z = C.Z, P = C.P;
        const double den = zden(A, B, Z), bcf = zbcoef(A, B, Z);
        const double Zx = ((B - Z) * (2 * C.s) + bcf * C.Bi) / den;
        const double G = (Z + D1 * B) / (Z + D2 * B), lG = log(G);
        const double ZB = Z - B;
        const double Axi = 2 * C.s, Bxi = C.Bi;
        const double Cc = C.y * P / ZB;
        const double E = -A / ((D1 - D2) * B) * (Axi / A - Bxi / B);
        const double g2 = (D1 - D2) / (Z + D2 * B) / (Z + D2 * B);
        double row[NC];
        double tr = 0;
        UNROLL for (int k = 0; k < NC; k++) {
            const double sAk = bc<T>(C.sA, k), Bxk = bc<T>(C.Bi, k), Axk = 2 * bc<T>(C.s, k);
            const double Zxk = bc<T>(Zx, k), yk = bc<T>(C.y, k);
            const double aik = lf.omk[k] * C.sA * sAk;
            const double Cx = (ZB * (t == k ? 1.0 : 0.0) - C.y * (Zxk - Bxk)) * P / (ZB * ZB);
            const double Dx = Bxi / B * (Zxk - Bxk * (Z - 1) / B);
            double Ex = (2 * (A / B * Bxk * Bxi + B * aik) - Axi * Bxk - Axk * C.Bi) / (B * B);
            Ex /= -(D1 - D2);
            const double Gx = g2 * (Z * Bxk - Zxk * B);
            row[k] = 1 / Cc * Cx + Dx + Ex * lG + E / G * Gx;
            tr += yk * row[k];
        }
        UNROLL for (int k = 0; k < NC; k++) row[k] = (row[k] - tr + 1) / C.Yt;
        mirror_lower<NC, T>(row, t);
        const double sol = team_solve<NC, T>(row, C.a3, t);
        C.cnt[6]++;
        C.cnt[7]++;
        C.y = act ? C.y * C.Yt + sol : 0.0;
        normalise_y(C, act);
        break;
    }
    case ST_S_SSM_A:
        rr2(f, C, act);
        update_x_rr(C);
        C.y = C.a2;
        break;
    case ST_S_SSM_B:
        C.y = C.a3;
        break;
    case ST_S_NR: {   // :826-868
        const double nu0 = C.nu0, nu1 = 1 - C.nu0, P = C.P;
        double s0, A0, B0, Z0, s1, A1, B1, Z1;
        mix<NC, T>(lf, C, act, C.a2, s0, A0, B0, Z0);
        mix<NC, T>(lf, C, act, C.a3, s1, A1, B1, Z1);
        // d ln f / d n precomputations of both phases (src/OCPEoS.cpp:153-168)
        const double den0 = zden(A0, B0, Z0), bc0 = zbcoef(A0, B0, Z0);
        const double den1 = zden(A1, B1, Z1), bc1 = zbcoef(A1, B1, Z1);
        const double An0 = 2 / nu0 * (s0 - A0), Bn0 = 1 / nu0 * (C.Bi - B0);
        const double Zn0 = ((B0 - Z0) * An0 + bc0 * Bn0) / den0;
        const double An1 = 2 / nu1 * (s1 - A1), Bn1 = 1 / nu1 * (C.Bi - B1);
        const double Zn1 = ((B1 - Z1) * An1 + bc1 * Bn1) / den1;
        const double G0 = (Z0 + D1 * B0) / (Z0 + D2 * B0), lG0 = log(G0);
        const double G1 = (Z1 + D1 * B1) / (Z1 + D2 * B1), lG1 = log(G1);
        double J[NC];
        UNROLL for (int k = 0; k < NC; k++) {
            const double sAk = bc<T>(C.sA, k), Bik = bc<T>(C.Bi, k);
            const double aik = lf.omk[k] * C.sA * sAk;
            const double v0 = lnfn_ik(A0, B0, Z0, lG0, G0, P, nu0, s0, C.Bi, C.a2, aik, t == k, Bik,
                                      bc<T>(An0, k), bc<T>(Bn0, k), bc<T>(Zn0, k));
            const double v1 = lnfn_ik(A1, B1, Z1, lG1, G1, P, nu1, s1, C.Bi, C.a3, aik, t == k, Bik,
                                      bc<T>(An1, k), bc<T>(Bn1, k), bc<T>(Zn1, k));
            J[k] = v0 + v1;
        }
        mirror_lower<NC, T>(J, t);
        C.cnt[6] += 2;
        C.cnt[7]++;
        const double dn = team_solve<NC, T>(J, C.a5, t);
        // CalStepNRsp :939-957: alpha shrinks over the components in order
        double alpha = 1;
        UNROLL for (int i = 0; i < NC; i++) {
            const double a1i = bc<T>(C.a1, i), dni = bc<T>(dn, i);
            if (a1i + alpha * dni < 0) alpha *= 0.9 * fabs(a1i / dni);
        }
        const double d = alpha * dn;
        C.a1 = act ? C.a1 + d : 0.0;
        C.nu0 = tsum<T>(C.a1);
        C.Sk = tsum<T>(act ? 2 * d * d : 0.0);
        const double r0 = 1 / C.nu0, r1 = 1 / (1 - C.nu0);
        C.a2 = act ? fabs(C.a1 * r0) : 0.0;
        C.a3 = act ? fabs((C.z - C.a1) * r1) : 0.0;
        C.y = C.a2;
        break;
    }
    default:   // ST_FEED: y = z
        break;
    }

    // common evaluator: ln phi at y
    double ev = 0;
    if (eval) {
        mix<NC, T>(lf, C, act, C.y, C.s, C.A, C.B, C.Z);
        ev = act ? lnphi_of(C.s, C.Bi, C.A, C.B, C.Z) : 0.0;
        C.cnt[5]++;
    }

    switch (C.st) {
    case ST_FEED:
        C.a1 = act ? exp(ev) * C.z : 0.0;
        C.k = 0;
        C.st = C.lNP == 0 ? ST_T_START : ST_A_START;
        break;
    case ST_T_START:
    case ST_T_SSM: {
        const double phS = exp(ev);
        const double r = act ? log((phS * C.y * C.P) / (C.a1 * C.P) * C.Yt) : 0.0;
        const double Se = tsum<T>(r * r);
        C.a3 = -r;
        C.a2 = act ? phS : 0.0;
        if (C.st == ST_T_START) {
            C.iter = 0;
            if (Se > f.SSMsta_tol2) { C.st = ST_T_SSM; break; }
            nr_sta_begin(f, lf, C, io, act, t);
            break;
        }
        C.cnt[0]++;
        if (Se > f.SSMsta_tol2 && ++C.iter <= f.SSMsta_maxIt) break;
        nr_sta_begin(f, lf, C, io, act, t);
        break;
    }
    case ST_T_NR: {
        C.a3 = act ? log((C.a1 * C.P) / ((exp(ev) * C.y * C.P) * C.Yt)) : 0.0;
        const double e2 = tsum<T>(C.a3 * C.a3);
        C.cnt[1]++;
        if (++C.iter > f.NRsta_maxIt) { C.flag = 0; strict_decide(f, lf, C, io, t); break; }
        if (sqrt(e2) > f.NRsta_tol) break;
        C.flag = 1;
        strict_decide(f, lf, C, io, t);
        break;
    }
    case ST_A_START:
    case ST_A_ITER: {
        if (!eval) { relaxed_trial_end(f, lf, C, io, t); break; }
        const double di = (C.a1 * C.P) / ((exp(ev) * C.y * C.P) * C.Yt);
        C.a4 = act ? C.a4 * di : 0.0;
        const double lk = act ? log(C.a4) : 0.0;
        const double Se = tsum<T>(act ? (di - 1) * (di - 1) : 0.0);
        const double Sk = tsum<T>(lk * lk);
        C.a3 = C.y;
        C.a2 = ev;
        C.Sk = Sk;
        C.cnt[0]++;
        C.iter++;
        if (Se < f.SSMsta_tol2 || Sk < KTOL) { C.flag = 1; relaxed_trial_end(f, lf, C, io, t); break; }
        if (C.iter > f.SSMsta_maxIt) { C.flag = 0; relaxed_trial_end(f, lf, C, io, t); break; }
        C.st = ST_A_ITER;
        break;
    }
    case ST_S_SSM_A:
        C.a5 = ev;
        C.st = ST_S_SSM_B;
        break;
    case ST_S_SSM_B: {
        C.cnt[2]++;
        C.iter++;
        const double phi0 = exp(C.a5), phi1 = exp(ev);
        const double q = act ? (phi1 * C.a3 * C.P) / (phi0 * C.a2 * C.P) - 1 : 0.0;
        const double Se = tsum<T>(q * q);
        bool last;
        if (C.iter > C.maxIt) { C.ssm_conv = 0; last = true; }
        else last = !(Se > C.tol2);
        C.a4 = act ? phi1 / phi0 : 0.0;
        if (last) split_nr_begin(f, lf, C, io, ev, act, t);
        else C.st = ST_S_SSM_A;
        break;
    }
    case ST_S_NR: {
        const double lnphi0 = ev;
        C.y = C.a3;
        mix<NC, T>(lf, C, act, C.y, C.s, C.A, C.B, C.Z);
        const double lnphi1 = act ? lnphi_of(C.s, C.Bi, C.A, C.B, C.Z) : 0.0;
        C.cnt[5]++;
        C.a5 = act ? log((exp(lnphi1) * C.a3 * C.P) / (exp(lnphi0) * C.a2 * C.P)) : 0.0;
        const double eNR = sqrt(tsum<T>(C.a5 * C.a5));
        C.cnt[3]++;
        C.iter++;
        bool fin = false;
        if (eNR > C.eNR0 || C.iter > f.NRsp_maxIt) fin = true;
        else if (C.Sk < 1E-16 * (2.0 * NC) * (2.0 * NC)) { C.nr_conv = 1; fin = true; }
        else if (!(eNR > f.NRsp_tol)) fin = true;
        if (fin) {
            if (eNR < f.NRsp_tol) C.nr_conv = 1;
            after_split_nr(f, lf, C, io, act, t);
        }
        break;
    }
    default:
        break;
    }
}

// ------------------------------------------------------------ kernels
DEV const DevFluid* opaque(const DevFluid* p)
{
    asm volatile("" : "+s"(p));
    return p;
}

// MULTI: cells of several models (io.mid), each with its own fluid; the
// fluid pointer is then per team instead of uniform across the wavefront.
template <int NC, int T, bool MULTI>
DEV const DevFluid& fluid_of(const DevFluid* fp, const Cell<NC, T>& C)
{
    if constexpr (MULTI) return fp[C.model];
    else return *opaque(fp);
}

template <int NC, int T, bool MULTI>
__global__ void __launch_bounds__(64) pe_team_plain(const DevFluid* __restrict__ fp, PeIO io)
{
    const int lane = __lane_id();
    const int t = lane & (T - 1);
    const int cell = (blockIdx.x * 64 + threadIdx.x) / T;
    if (cell >= io.nb) return;                   // whole teams leave together
    const bool act = t < NC;
    LaneFluid<NC> lf;
    Cell<NC, T> C;
    C.model = MULTI ? io.mid[cell] : 0;
    begin<NC, T>(fp[C.model], lf, io, cell, C, act, t);
    while (C.st != ST_DONE) step<NC, T>(fluid_of<NC, T, MULTI>(fp, C), lf, io, C, act, t);
    end<NC, T>(io, C, act, t);
}

// Persistent teams over the global work queue (counter io.queue, zeroed
// before the launch).  Only team leaders vote; a team whose cell is done
// writes it back and takes the next index.  The wavefront leaves when the
// queue is drained and no team has a cell.
template <int NC, int T, bool MULTI>
__global__ void __launch_bounds__(64) pe_team_batched(const DevFluid* __restrict__ fp, PeIO io)
{
    const int lane = __lane_id();
    const int t = lane & (T - 1);
    const int lead = lane & ~(T - 1);
    const bool act = t < NC;
    LaneFluid<NC> lf;
    Cell<NC, T> C;
    C.st = ST_DONE;
    C.model = 0;
    bool have = false, drained = false;
    while (true) {
        if (C.st == ST_DONE && have) { end<NC, T>(io, C, act, t); have = false; }
        const bool need = !have && !drained;
        const unsigned long long m = __ballot(need && t == 0);
        if (m) {
            const int first = __ffsll((long long)m) - 1;
            unsigned base = 0;
            if (lane == first) base = atomicAdd(io.queue, (unsigned)__popcll(m));
            base = __shfl(base, first);
            if (need) {
                const unsigned idx = base + (unsigned)__popcll(m & ((1ull << lead) - 1ull));
                if (idx < (unsigned)io.nb) {
                    C.model = MULTI ? io.mid[idx] : 0;
                    begin<NC, T>(fp[C.model], lf, io, (int)idx, C, act, t);
                    have = true;
                }
                else drained = true;
            }
        }
        if (!__any(have)) break;
        if (have && C.st != ST_DONE) step<NC, T>(fluid_of<NC, T, MULTI>(fp, C), lf, io, C, act, t);
    }
}

template <int NC, int T>
hipError_t launch(const DevFluid* f, const PeIO& io, bool batched, int grid_blocks, hipStream_t st)
{
    if (io.nb == 0) return hipSuccess;
    const long lanes = io.nb * (long)T;
    long need = (lanes + 63) / 64;
    const bool multi = io.mid != nullptr;
    if (!batched) {
        if (multi) hipLaunchKernelGGL((pe_team_plain<NC, T, true>), dim3((unsigned)need), dim3(64), 0, st, f, io);
        else hipLaunchKernelGGL((pe_team_plain<NC, T, false>), dim3((unsigned)need), dim3(64), 0, st, f, io);
    } else {
        const long nblk = grid_blocks < need ? grid_blocks : need;
        if (multi) hipLaunchKernelGGL((pe_team_batched<NC, T, true>), dim3((unsigned)nblk), dim3(64), 0, st, f, io);
        else hipLaunchKernelGGL((pe_team_batched<NC, T, false>), dim3((unsigned)nblk), dim3(64), 0, st, f, io);
    }
    return hipGetLastError();
}

template <int NC, int T>
int blocks_per_cu()
{
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, pe_team_batched<NC, T, true>, 64, 0) != hipSuccess) return 0;
    return n;
}

}  // namespace team

// T = 8 for Nc <= 8, 16 above (teams tile a 64-lane wavefront exactly)
hipError_t run_pe_team(int NC, int T, const DevFluid* f, const PeIO& io, bool batched, int grid_blocks,
                       hipStream_t st)
{
#define OCPX_T8(N) case N: return T == 8 ? team::launch<N, 8>(f, io, batched, grid_blocks, st) \
                                         : team::launch<N, 16>(f, io, batched, grid_blocks, st);
#define OCPX_T16(N) case N: return team::launch<N, 16>(f, io, batched, grid_blocks, st);
    if (T != 8 && T != 16) return hipErrorInvalidValue;
    if (T == 8 && NC > 8) return hipErrorInvalidValue;
    switch (NC) {
        OCPX_T8(2) OCPX_T8(3) OCPX_T8(4) OCPX_T8(5) OCPX_T8(6) OCPX_T8(7) OCPX_T8(8)
        OCPX_T16(9) OCPX_T16(10) OCPX_T16(11) OCPX_T16(12)
    default: return hipErrorInvalidValue;
    }
#undef OCPX_T8
#undef OCPX_T16
}

int pe_team_blocks_per_cu(int NC, int T)
{
    switch (NC) {
#define OCPX_B8(N) case N: return T == 8 ? team::blocks_per_cu<N, 8>() : team::blocks_per_cu<N, 16>();
#define OCPX_B16(N) case N: return T == 16 ? team::blocks_per_cu<N, 16>() : 0;
        OCPX_B8(2) OCPX_B8(3) OCPX_B8(4) OCPX_B8(5) OCPX_B8(6) OCPX_B8(7) OCPX_B8(8)
        OCPX_B16(9) OCPX_B16(10) OCPX_B16(11) OCPX_B16(12)
#undef OCPX_B8
#undef OCPX_B16
    default: return 0;
    }
}

}  // namespace ocpx
