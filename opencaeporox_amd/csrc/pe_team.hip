// pe_team.hip — phase equilibrium with a team of T lanes per cell.
//
// Same algorithm and state machine as pe.hip (OCPPhaseEquilibrium::
// PhaseEquilibrium, src/OCPPhaseEquilibrium.cpp:71-112: stability by
// StableSSM/StableNR or StableSSM01, split by SplitSSM2 + RachfordRice2 +
// SplitNR + CheckSplit), mapped differently: a team of T lanes holds one
// cell and lane t owns components t, t+T, t+2T, ... (L = ceil(Nc/T) slots),
// so a cell's state is a few registers per lane instead of ~7 Nc doubles in
// one lane, and no Nc fills private memory.
//
//   * sums over components are butterfly reductions inside the team (DPP
//     lane permutations) after a local sum over the lane's slots; every lane
//     ends with the same bits, so control flow stays uniform across the team;
//   * the mixture term s_i = sqrt(Ai) sum_k (1-k_ik) sqrt(Ak) x_k takes the
//     other components' sqrt(Ak) x_k by broadcast (DPP quad_perm for teams
//     of 4, ds_swizzle for 8 and 16), summed in k order;
//   * Newton matrices are built row by row on the lanes owning the rows; the
//     entries above the diagonal, which the reference takes from the lower
//     triangle (dsysv_ 'U' on row-major storage), are evaluated on the row's
//     own lane from the other component's broadcast quantities;
//   * they are solved by Gaussian elimination with partial pivoting across
//     the team (pivot search by butterfly argmax, pivot row by shuffle,
//     column-oriented back substitution).
//
// Scalar work per cell (cubic root, Rachford-Rice scalars) is shared by the
// T lanes of a team, so a small T serves more cells per instruction; a large
// T spreads the component work thinner.  The autotuner picks T per Nc.
// pe_team_plain gives each team one cell; pe_team_batched keeps teams
// resident and refills them from the global work queue as cells finish.
#include "ocpx_dev.hpp"
#include "ocpx_kernels.hpp"

#include <type_traits>
#include <utility>

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
// Reductions are butterflies of DPP lane permutations (quad_perm [1,0,3,2],
// quad_perm [2,3,0,1], row_half_mirror, row_mirror): each stage pairs every
// lane with one partner in the other half of its group of 2, 4, 8 or 16
// lanes, and the combining operation is commutative, so every lane of a
// team ends with the same bits.
template <int CTRL>
DEV int dpp_i(int v) { return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false); }
template <int CTRL>
DEV double dpp_d(double v)
{
    const long long u = __double_as_longlong(v);
    const int lo = dpp_i<CTRL>((int)(u & 0xffffffffll));
    const int hi = dpp_i<CTRL>((int)(u >> 32));
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
template <int T, typename Op>
DEV double tred(double v, Op op)
{
    static_assert(T == 4 || T == 8 || T == 16, "team size");
    v = op(v, dpp_d<0xB1>(v));
    v = op(v, dpp_d<0x4E>(v));
    if constexpr (T >= 8) v = op(v, dpp_d<0x141>(v));
    if constexpr (T >= 16) v = op(v, dpp_d<0x140>(v));
    return v;
}
template <int T>
DEV double tsum(double v) { return tred<T>(v, [](double a, double b) { return a + b; }); }
template <int T>
DEV double tmin(double v) { return tred<T>(v, [](double a, double b) { return fmin(a, b); }); }
template <int T>
DEV double tmax(double v) { return tred<T>(v, [](double a, double b) { return fmax(a, b); }); }

// lane k of the team (k uniform across the team, not compile-time)
template <int T>
DEV double bc(double v, int k) { return __shfl(v, k, T); }

// lane K of the team, K compile-time: a quad_perm broadcast for teams of 4,
// else ds_swizzle in bit mode (source lane = (lane & ~(T-1)) | K within each
// 32-lane half).
template <int T, int K>
DEV double bck(double v)
{
    if constexpr (T == 4) {
        return dpp_d<K | (K << 2) | (K << 4) | (K << 6)>(v);
    } else {
        constexpr int pat = ((32 - T) & 0x1F) | (K << 5);
        const long long u = __double_as_longlong(v);
        const int lo = __builtin_amdgcn_ds_swizzle((int)(u & 0xffffffffll), pat);
        const int hi = __builtin_amdgcn_ds_swizzle((int)(u >> 32), pat);
        return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
    }
}

// f(std::integral_constant<int, k>) for k = 0 .. N-1, in order
template <typename F, int... K>
DEV void static_for_impl(F&& f, std::integer_sequence<int, K...>) { (f(std::integral_constant<int, K>{}), ...); }
template <int N, typename F>
DEV void static_for(F&& f) { static_for_impl(f, std::make_integer_sequence<int, N>{}); }

// component k of a slot array: lane k % T, slot k / T (k a constant
// expression: the loop variable of a static_for)
#define BCK(arr, k) bck<T, (k) % T>((arr)[(k) / T])
#define SLOTS UNROLL for (int c = 0; c < L; c++)

// slot s of an array, s uniform across the team but not compile-time
template <int L>
DEV double pick(const double (&v)[L], int s)
{
    double u = v[0];
    UNROLL for (int c = 1; c < L; c++) u = (s == c) ? v[c] : u;
    return u;
}

template <int NC, int T>
struct Lane {
    static constexpr int L = (NC + T - 1) / T;
    int t;                 // lane within the team
    bool act[L];           // slot c holds component c*T + t < NC
};

// Fluid constants of the lane's components, loaded when the team takes a cell.
template <int NC, int T>
struct LaneFluid {
    static constexpr int L = (NC + T - 1) / T;
    double Tc[L], Pc[L], logPc[L], mw[L], OA[L], OB[L], kwA[L];
    double omk[L][NC];     // 1 - k_ik
};

template <int NC, int T>
DEV void lane_fluid(const DevFluid& f, const Lane<NC, T>& ln, LaneFluid<NC, T>& lf)
{
    constexpr int L = LaneFluid<NC, T>::L;
    SLOTS {
        const int i = ln.act[c] ? c * T + ln.t : 0;
        lf.Tc[c] = f.Tc[i]; lf.Pc[c] = f.Pc[i]; lf.logPc[c] = f.logPc[i]; lf.mw[c] = f.mw[i];
        lf.OA[c] = f.OA[i]; lf.OB[c] = f.OB[i]; lf.kwA[c] = f.kwA[i];
        UNROLL for (int k = 0; k < NC; k++) lf.omk[c][k] = f.omk[i * OCPX_NCMAX + k];
    }
}

template <int NC, int T>
struct Cell {
    static constexpr int L = (NC + T - 1) / T;
    int cell, model, st, k, iter, maxIt, NP, ftype, flag, stage, restarts, ssm_conv, nr_conv, steps;
    uint32_t lNP;
    double P, Tk, logP, Yt, tol2, Sk, nu0, eNR0;
    int cnt[8];
    double z[L], y[L], a1[L], a2[L], a3[L], a4[L], a5[L];
    double sA[L], Bi[L];        // sqrt(Ai), Bi at (P, T)
    double s[L], A, B, Z;       // mixture at the last evaluated y
};

// masked team sum of a slot expression
#define TSUM(expr) ([&]() { double acc_ = 0; SLOTS acc_ += ln.act[c] ? (expr) : 0.0; return tsum<T>(acc_); }())

// ------------------------------------------------------------- EoS pieces
template <int NC, int T>
DEV void mix(const LaneFluid<NC, T>& lf, const Lane<NC, T>& ln, const Cell<NC, T>& C,
             const double (&x)[Cell<NC, T>::L], double (&s)[Cell<NC, T>::L], double& A, double& B,
             double& Z)
{
    constexpr int L = Cell<NC, T>::L;
    double w[L], acc[L];
    SLOTS { w[c] = ln.act[c] ? C.sA[c] * x[c] : 0.0; acc[c] = 0; }
    static_for<NC>([&](auto kc_) {
        constexpr int k = decltype(kc_)::value;
        const double wk = BCK(w, k);
        SLOTS acc[c] += lf.omk[c][k] * wk;
    });
    SLOTS s[c] = C.sA[c] * acc[c];
    A = TSUM(x[c] * s[c]);
    B = TSUM(C.Bi[c] * x[c]);
    Z = pr_z(A, B);
}

DEV double lnphi_of(double s, double Bi, double A, double B, double Z)
{
    const double lz = log(Z - B);
    const double lg = log((Z + D1 * B) / (Z + D2 * B));
    const double cA = A / (D1 - D2) / B * lg;
    const double iB = 1 / B, iA = 1 / A;
    return Bi * iB * (Z - 1) - lz - cA * (2 * s * iA - Bi * iB);
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

// Entry (i,k) of d ln f_i / d x_k (src/OCPEoS.cpp:217-251): row i quantities
// (Axi = 2 s_i, Bxi = B_i, x_i), column k quantities (Axk, Bxk, Zxk).
DEV double lnfx_ik(double A, double B, double Z, double lG, double G, double ZB, double g2, double P,
                   double Axi, double Bxi, double xi, double aik, bool diag, double Axk, double Bxk, double Zxk)
{
    const double C = xi * P / ZB;
    const double E = -A / ((D1 - D2) * B) * (Axi / A - Bxi / B);
    const double Cx = (ZB * (diag ? 1.0 : 0.0) - xi * (Zxk - Bxk)) * P / (ZB * ZB);
    const double Dx = Bxi / B * (Zxk - Bxk * (Z - 1) / B);
    double Ex = (2 * (A / B * Bxk * Bxi + B * aik) - Axi * Bxk - Axk * Bxi) / (B * B);
    Ex /= -(D1 - D2);
    const double Gx = g2 * (Z * Bxk - Zxk * B);
    return 1 / C * Cx + Dx + Ex * lG + E / G * Gx;
}

// Entry (i,k) of d ln f_i / d n_k for one phase (src/OCPEoS.cpp:254-300):
// row i quantities (s_i, B_i, x_i), column k quantities (B_k, An_k, Bn_k, Zn_k).
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

// One butterfly stage of the pivot search: larger |value|, ties to the lower
// row; a commutative choice, so the team agrees on the pivot.
template <int CTRL>
DEV void argmax_step(double& v, int& p)
{
    const double v2 = dpp_d<CTRL>(v);
    const int p2 = dpp_i<CTRL>(p);
    if (v2 > v || (v2 == v && p2 < p)) { v = v2; p = p2; }
}

// Solve M x = b with rows i = c*T + t on the team (partial pivoting, ties to
// the lower row).  Returns x of the lane's components.
template <int N, int T, int L>
DEV void team_solve(double (&M)[L][N], double (&b)[L], const bool (&act)[L], int t, double (&sol)[L])
{
    bool used[L];
    int ord[L];
    SLOTS { used[c] = !act[c]; ord[c] = N; sol[c] = 0; }
    int piv[N];
    UNROLL for (int col = 0; col < N; col++) {
        double v = -1.0;
        int p = 1 << 30;
        SLOTS {
            double vc = used[c] ? -1.0 : fabs(M[c][col]);
            vc = (vc == vc) ? vc : -0.5;                 // NaN rows pivot last
            const int rc = c * T + t;
            if (vc > v || (vc == v && rc < p)) { v = vc; p = rc; }
        }
        argmax_step<0xB1>(v, p);
        argmax_step<0x4E>(v, p);
        if constexpr (T >= 8) argmax_step<0x141>(v, p);
        if constexpr (T >= 16) argmax_step<0x140>(v, p);
        piv[col] = p;
        const int pl = p % T, ps = p / T;
        double colv[L];
        SLOTS colv[c] = M[c][col];
        const double id = 1 / bc<T>(pick<L>(colv, ps), pl);
        const double bp = bc<T>(pick<L>(b, ps), pl);
        double prow[N];
        UNROLL for (int k = col + 1; k < N; k++) {
            double mk[L];
            SLOTS mk[c] = M[c][k];
            prow[k] = bc<T>(pick<L>(mk, ps), pl);
        }
        SLOTS {
            const bool mine = (t == pl) && (c == ps);
            if (mine) { used[c] = true; ord[c] = col; }
            else if (!used[c]) {
                const double l = M[c][col] * id;
                UNROLL for (int k = col + 1; k < N; k++) M[c][k] -= l * prow[k];
                b[c] -= l * bp;
            }
        }
    }
    UNROLL for (int col = N - 1; col >= 0; col--) {
        const int p = piv[col];
        const int pl = p % T, ps = p / T;
        double colv[L];
        SLOTS colv[c] = M[c][col];
        const double xc = bc<T>(pick<L>(b, ps), pl) / bc<T>(pick<L>(colv, ps), pl);
        if (t == col % T) sol[col / T] = xc;
        SLOTS if (ord[c] < col) b[c] -= M[c][col] * xc;
    }
}

// ------------------------------------------------------------ algorithm
DEV double trial_coef(int k)
{
    return k == 0 ? 1.0 : (k == 1 ? -1.0 : (k == 2 ? (1.0 / 3) : -(1.0 / 3)));
}

template <int NC, int T>
DEV double log_kw0(const LaneFluid<NC, T>& lf, int c, double logP, double Tk)
{
    return lf.logPc[c] - logP + lf.kwA[c] * (1 - lf.Tc[c] / Tk);
}

template <int NC, int T>
DEV void normalise_y(Cell<NC, T>& C, const Lane<NC, T>& ln)
{
    constexpr int L = Cell<NC, T>::L;
    C.Yt = TSUM(C.y[c]);
    const double r = 1 / C.Yt;
    SLOTS C.y[c] *= r;
}

template <int NC, int T>
DEV void rr2(const DevFluid& f, Cell<NC, T>& C, const Lane<NC, T>& ln)
{   // RachfordRice2 :651-705 on K = a4
    constexpr int L = Cell<NC, T>::L;
    double kmn = INFINITY, kmx = -INFINITY;
    SLOTS if (ln.act[c]) { kmn = fmin(kmn, C.a4[c]); kmx = fmax(kmx, C.a4[c]); }
    const double Kmin = tmin<T>(kmn), Kmax = tmax<T>(kmx);
    const double numin = 1 / (1 - Kmax), numax = 1 / (1 - Kmin);
    double nu = 0.5 * (numin + numax);
    int it = 0;
    double km[L];
    SLOTS km[c] = C.a4[c] - 1;
    while (true) {
        double tt[L];
        SLOTS tt[c] = 1 / (1 + nu * km[c]);
        const double rj = TSUM(C.z[c] * km[c] * tt[c]);
        const double J = -TSUM(C.z[c] * km[c] * km[c] * tt[c] * tt[c]);
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
DEV void update_x_rr(Cell<NC, T>& C, const Lane<NC, T>& ln)
{   // UpdateXRR :778-791
    constexpr int L = Cell<NC, T>::L;
    SLOTS {
        C.a3[c] = ln.act[c] ? C.z[c] / (1 + C.nu0 * (C.a4[c] - 1)) : 0.0;
        C.a2[c] = C.a4[c] * C.a3[c];
    }
}

template <int NC, int T>
DEV void split_start(const DevFluid& f, const LaneFluid<NC, T>& lf, Cell<NC, T>& C, const PeIO& io,
                     bool restart, const Lane<NC, T>& ln)
{   // SplitSSM2 :586-622
    constexpr int L = Cell<NC, T>::L;
    C.ssm_conv = 1;
    C.tol2 = f.SSMsp_tol2;
    C.maxIt = f.SSMsp_maxIt;
    if (!restart) {
        if (C.lNP == 2) {
            const size_t n = io.nb;
            const int nc = NC + 1;
            SLOTS {
                const int i = ln.act[c] ? c * T + ln.t : 0;
                C.a4[c] = io.xij[(size_t)i * n + C.cell] / io.xij[(size_t)(nc + i) * n + C.cell];
            }
        } else {
            SLOTS C.a4[c] = exp(log_kw0(lf, c, C.logP, C.Tk));
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
DEV bool check_split(Cell<NC, T>& C, const Lane<NC, T>& ln)
{   // CheckSplit :960-1010
    constexpr int L = Cell<NC, T>::L;
    if (C.NP != 2) return true;
    const double nu1 = 1 - C.nu0;
    const double nuMax = fmax(C.nu0, nu1);
    const double eX = TSUM((C.a2[c] - C.a3[c]) * (C.a2[c] - C.a3[c]));
    if (nuMax < 1 && C.nr_conv && finite_d(eX)) return true;
    if (!finite_d(eX) || 1 - nuMax < 1E-3) {
        C.NP = 1;
        C.nu0 = 1;
        SLOTS C.a2[c] = C.z[c];
        return false;
    }
    return true;
}

template <int NC, int T>
DEV void finish_split(Cell<NC, T>& C, const Lane<NC, T>& ln)
{
    check_split(C, ln);
    C.st = ST_DONE;
}

template <int NC, int T>
DEV void after_split_nr(const DevFluid& f, const LaneFluid<NC, T>& lf, Cell<NC, T>& C, const PeIO& io,
                        const Lane<NC, T>& ln)
{   // PhaseSplit restart loop :558-566
    if (C.stage == 0) {
        if (C.nr_conv) { finish_split(C, ln); return; }
        C.stage = 1;
    } else {
        if (!check_split(C, ln)) { finish_split(C, ln); return; }
        if (C.ssm_conv || C.nr_conv) { finish_split(C, ln); return; }
    }
    if (++C.restarts > MAX_SPLIT_RESTARTS) { C.cnt[7] |= 0x40000000; finish_split(C, ln); return; }
    split_start(f, lf, C, io, true, ln);
}

template <int NC, int T>
DEV void split_nr_begin(const DevFluid& f, const LaneFluid<NC, T>& lf, Cell<NC, T>& C, const PeIO& io,
                        const double (&lnphi1)[Cell<NC, T>::L], const Lane<NC, T>& ln)
{   // SplitNR :807-823; a5 = ln phi0 of x0
    constexpr int L = Cell<NC, T>::L;
    SLOTS {
        C.a1[c] = C.nu0 * C.a2[c];
        C.a5[c] = ln.act[c] ? log((exp(lnphi1[c]) * C.a3[c] * C.P) / (exp(C.a5[c]) * C.a2[c] * C.P)) : 0.0;
    }
    const double eNR = sqrt(TSUM(C.a5[c] * C.a5[c]));
    C.eNR0 = eNR;
    C.iter = 0;
    C.nr_conv = 0;
    if (eNR > f.NRsp_tol) { C.st = ST_S_NR; return; }
    C.nr_conv = eNR < f.NRsp_tol;
    after_split_nr(f, lf, C, io, ln);
}

template <int NC, int T>
DEV void strict_decide(const DevFluid& f, const LaneFluid<NC, T>& lf, Cell<NC, T>& C, const PeIO& io,
                       const Lane<NC, T>& ln)
{   // :460-470
    if (C.flag && C.Yt > 1 + f.eYt) {
        C.NP = 2;
        C.stage = 0;
        C.restarts = 0;
        split_start(f, lf, C, io, false, ln);
        return;
    }
    if (++C.k < 4) { C.st = ST_T_START; return; }
    C.st = ST_DONE;
}

template <int NC, int T>
DEV void nr_sta_begin(const DevFluid& f, const LaneFluid<NC, T>& lf, Cell<NC, T>& C, const PeIO& io,
                      const Lane<NC, T>& ln)
{   // StableNR :478-488, resSTA in a3
    constexpr int L = Cell<NC, T>::L;
    const double e2 = TSUM(C.a3[c] * C.a3[c]);
    C.iter = 0;
    if (sqrt(e2) > f.NRsta_tol) { C.st = ST_T_NR; return; }
    C.flag = 1;
    strict_decide(f, lf, C, io, ln);
}

template <int NC, int T>
DEV void relaxed_trial_end(const DevFluid& f, const LaneFluid<NC, T>& lf, Cell<NC, T>& C, const PeIO& io,
                           const Lane<NC, T>& ln)
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
    split_start(f, lf, C, io, false, ln);
}

// ------------------------------------------------------------ begin / end
template <int NC, int T>
DEV void begin(const DevFluid& f, LaneFluid<NC, T>& lf, const PeIO& io, int cell, Cell<NC, T>& C,
               const Lane<NC, T>& ln)
{
    constexpr int L = Cell<NC, T>::L;
    const size_t n = io.nb;
    C.cell = cell;
    lane_fluid<NC, T>(f, ln, lf);
    C.P = io.P[cell];
    C.Tk = io.T[cell] + CONV4;
    C.logP = log(C.P);
    double Ni[L];
    SLOTS Ni[c] = ln.act[c] ? io.Ni[(size_t)(c * T + ln.t) * n + cell] : 0.0;
    const double Nt = TSUM(Ni[c]);
    SLOTS C.z[c] = ln.act[c] ? Ni[c] / Nt : 0.0;
    uint32_t phaseNum = 1;                         // no previous step: one phase
    if (io.phaseExist) {
        phaseNum = 0;
        UNROLL for (int j = 0; j < 3; j++) phaseNum += io.phaseExist[(size_t)j * n + cell] != 0;
    }
    C.lNP = phaseNum - 1u;
    int ft = io.ftype_in ? io.ftype_in[cell] : -1;
    if (ft < 0) {
        ft = 0;
        if (f.skip) {
            bool skip = false;
            if (io.skip_flag[cell]) {
                const double me = io.skip_minEig[cell];
                skip = fabs(1 - io.skip_P[cell] / C.P) < me / 10 && fabs(io.skip_T[cell] - C.Tk) < me * 10;
                bool mine = true;
                SLOTS if (ln.act[c])
                    mine = mine && fabs(Ni[c] / Nt - io.skip_zi[(size_t)(c * T + ln.t) * n + cell]) < me / 10;
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
    UNROLL for (int q = 0; q < 8; q++) C.cnt[q] = 0;
    // EoS constants of the lane's components (src/OCPEoS.cpp:68-84)
    SLOTS {
        const double Pr = C.P / lf.Pc[c], Tr = C.Tk / lf.Tc[c];
        const double q = 1 + lf.mw[c] * (1 - sqrt(Tr));
        const double Ai = lf.OA[c] * Pr / (Tr * Tr) * (q * q);
        C.sA[c] = ln.act[c] ? sqrt(Ai) : 0.0;
        C.Bi[c] = ln.act[c] ? lf.OB[c] * Pr / Tr : 0.0;
        C.a1[c] = C.a2[c] = C.a3[c] = C.a4[c] = C.a5[c] = 0;
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
    if (ft == 1) { C.st = ST_DONE; return; }
    if (ft == 2) {
        C.NP = 2;
        C.Yt = 1.01;
        split_start(f, lf, C, io, false, ln);
        return;
    }
    if (io.sta_route && io.sta_route[cell] == 2) {   // unstable by the stability stages
        C.cnt[0] = io.rs_cnt[cell];
        C.cnt[1] = io.rs_cnt[n + cell];
        C.cnt[5] = io.rs_cnt[2 * n + cell];
        C.cnt[6] = C.cnt[7] = io.rs_cnt[3 * n + cell];
        C.Yt = io.rs_Yt[cell];
        C.NP = 2;                                  // strict_decide's unstable branch
        C.stage = 0;
        C.restarts = 0;
        split_start(f, lf, C, io, false, ln);
        return;
    }
    if (io.sta_route && io.sta_route[cell] == 3) {   // SplitSSM2's first pass done by the split stage
        C.cnt[0] = io.rs_cnt[cell];
        C.cnt[1] = io.rs_cnt[n + cell];
        C.cnt[5] = io.rs_cnt[2 * n + cell];
        C.cnt[6] = C.cnt[7] = io.rs_cnt[3 * n + cell];
        C.cnt[2] = io.rs_cnt[4 * n + cell];
        C.cnt[4] = io.rs_cnt[5 * n + cell];
        C.ssm_conv = io.rs_cnt[6 * n + cell];
        C.Yt = io.rs_Yt[cell];
        C.NP = 2;
        C.stage = 0;
        C.restarts = 0;
        C.nu0 = io.rs_nu[cell];
        double l1[L];
        SLOTS {
            const size_t i = ln.act[c] ? (size_t)(c * T + ln.t) : 0;
            C.a2[c] = ln.act[c] ? io.rs_v[i * n + cell] : 0.0;
            C.a3[c] = ln.act[c] ? io.rs_v[(NC + i) * n + cell] : 0.0;
            C.a4[c] = ln.act[c] ? io.rs_v[(2 * NC + i) * n + cell] : 0.0;
            C.a5[c] = ln.act[c] ? io.rs_v[(3 * NC + i) * n + cell] : 0.0;
            l1[c] = ln.act[c] ? io.rs_v[(4 * NC + i) * n + cell] : 0.0;
        }
        split_nr_begin(f, lf, C, io, l1, ln);
        return;
    }
    if (io.sta_route && io.sta_route[cell] == 4) {   // SplitNR did not converge in the split-Newton stage
        C.cnt[0] = io.rs_cnt[cell];
        C.cnt[1] = io.rs_cnt[n + cell];
        C.cnt[5] = io.rs_cnt[2 * n + cell];
        C.cnt[6] = io.rs_cnt[3 * n + cell];
        C.cnt[2] = io.rs_cnt[4 * n + cell];
        C.cnt[4] = io.rs_cnt[5 * n + cell];
        C.cnt[3] = io.rs_cnt[7 * n + cell];
        C.cnt[7] = C.cnt[6] - C.cnt[3];
        C.ssm_conv = io.rs_cnt[6 * n + cell];
        C.Yt = io.rs_Yt[cell];
        C.NP = 2;
        C.stage = 0;
        C.restarts = 0;
        C.nr_conv = 0;
        SLOTS C.a4[c] = ln.act[c] ? io.rs_v[(2 * NC + (size_t)(c * T + ln.t)) * n + cell] : 0.0;
        after_split_nr(f, lf, C, io, ln);           // the restart loop (:558-566)
        return;
    }
    SLOTS C.y[c] = C.z[c];
    C.st = ST_FEED;
}

template <int NC, int T>
DEV void end(const PeIO& io, const Cell<NC, T>& C, const Lane<NC, T>& ln)
{
    constexpr int L = Cell<NC, T>::L;
    const size_t n = io.nb;
    const int cl = C.cell;
    const int NP = C.NP;
    SLOTS if (ln.act[c]) {
        const int i = c * T + ln.t;
        io.pe_x[(size_t)i * n + cl] = NP > 1 ? C.a2[c] : C.z[c];
        io.pe_x[(size_t)(NC + i) * n + cl] = NP > 1 ? C.a3[c] : 0.0;
    }
    if (ln.t == 0) {
        io.pe_NP[cl] = NP;
        io.pe_nu[cl] = NP > 1 ? C.nu0 : 1.0;
        io.ftype_out[cl] = NP > 1 ? 2 : C.ftype;            // :111
        if (io.iters)
            UNROLL for (int q = 0; q < 8; q++) io.iters[(size_t)q * n + cl] = C.cnt[q];
    }
}

// ------------------------------------------------------------ Newton steps
// StableNR (:488-516): Jacobian JmatSTA (:523-541), solved as the symmetric
// matrix of its lower triangle (dsysv_ 'U' on row-major storage, :493).
template <int NC, int T>
DEV void stable_nr_pre(const LaneFluid<NC, T>& lf, Cell<NC, T>& C, const Lane<NC, T>& ln)
{
    constexpr int L = Cell<NC, T>::L;
    const double A = C.A, B = C.B, Z = C.Z, P = C.P;
    const double den = zden(A, B, Z), bcf = zbcoef(A, B, Z);
    const double G = (Z + D1 * B) / (Z + D2 * B), lG = log(G);
    const double ZB = Z - B;
    const double g2 = (D1 - D2) / (Z + D2 * B) / (Z + D2 * B);
    double Zx[L], Ax[L];
    SLOTS { Zx[c] = ((B - Z) * (2 * C.s[c]) + bcf * C.Bi[c]) / den; Ax[c] = 2 * C.s[c]; }
    // own rows: entries (i, k) for every k and the row sums t_i = sum_k y_k S_ik
    double row[L][NC], tr[L];
    SLOTS tr[c] = 0;
    static_for<NC>([&](auto kc_) {
        constexpr int k = decltype(kc_)::value;
        const double sAk = BCK(C.sA, k), Bxk = BCK(C.Bi, k), Axk = BCK(Ax, k);
        const double Zxk = BCK(Zx, k), yk = BCK(C.y, k);
        SLOTS {
            const double aik = lf.omk[c][k] * C.sA[c] * sAk;
            row[c][k] = lnfx_ik(A, B, Z, lG, G, ZB, g2, P, Ax[c], C.Bi[c], C.y[c], aik, c * T + ln.t == k,
                                Axk, Bxk, Zxk);
            tr[c] += yk * row[c][k];
        }
    });
    // entries above the diagonal come from row k's lower triangle:
    // S_ki = (lnfx(k, i) - t_k + 1) / Yt, evaluated here from k's quantities
    double M[L][NC];
    static_for<NC>([&](auto kc_) {
        constexpr int k = decltype(kc_)::value;
        const double sAk = BCK(C.sA, k), Bxk = BCK(C.Bi, k), Axk = BCK(Ax, k), yk = BCK(C.y, k);
        const double tk = BCK(tr, k);
        SLOTS {
            const int i = c * T + ln.t;
            const double own = (row[c][k] - tr[c] + 1) / C.Yt;
            const double aki = lf.omk[c][k] * sAk * C.sA[c];
            const double other = (lnfx_ik(A, B, Z, lG, G, ZB, g2, P, Axk, Bxk, yk, aki, false, Ax[c], C.Bi[c],
                                          Zx[c]) - tk + 1) / C.Yt;
            M[c][k] = (k <= i) ? own : other;
        }
    });
    double rhs[L], sol[L];
    SLOTS rhs[c] = C.a3[c];
    team_solve<NC, T, L>(M, rhs, ln.act, ln.t, sol);
    C.cnt[6]++;
    C.cnt[7]++;
    SLOTS C.y[c] = ln.act[c] ? C.y[c] * C.Yt + sol[c] : 0.0;
    normalise_y(C, ln);
}

// SplitNR (:826-868): Jacobian of both phases' d ln f / d n (:839-854),
// symmetric from its lower triangle (dsysv_ 'U', :835), CalStepNRsp (:939-957).
template <int NC, int T>
DEV void split_nr_pre(const LaneFluid<NC, T>& lf, Cell<NC, T>& C, const Lane<NC, T>& ln)
{
    constexpr int L = Cell<NC, T>::L;
    const double nu0 = C.nu0, nu1 = 1 - C.nu0, P = C.P;
    double s0[L], s1[L], A0, B0, Z0, A1, B1, Z1;
    mix<NC, T>(lf, ln, C, C.a2, s0, A0, B0, Z0);
    mix<NC, T>(lf, ln, C, C.a3, s1, A1, B1, Z1);
    const double den0 = zden(A0, B0, Z0), bc0 = zbcoef(A0, B0, Z0);
    const double den1 = zden(A1, B1, Z1), bc1 = zbcoef(A1, B1, Z1);
    const double G0 = (Z0 + D1 * B0) / (Z0 + D2 * B0), lG0 = log(G0);
    const double G1 = (Z1 + D1 * B1) / (Z1 + D2 * B1), lG1 = log(G1);
    double An0[L], Bn0[L], Zn0[L], An1[L], Bn1[L], Zn1[L];
    SLOTS {   // src/OCPEoS.cpp:153-168
        An0[c] = 2 / nu0 * (s0[c] - A0); Bn0[c] = 1 / nu0 * (C.Bi[c] - B0);
        Zn0[c] = ((B0 - Z0) * An0[c] + bc0 * Bn0[c]) / den0;
        An1[c] = 2 / nu1 * (s1[c] - A1); Bn1[c] = 1 / nu1 * (C.Bi[c] - B1);
        Zn1[c] = ((B1 - Z1) * An1[c] + bc1 * Bn1[c]) / den1;
    }
    double J[L][NC];
    static_for<NC>([&](auto kc_) {
        constexpr int k = decltype(kc_)::value;
        const double sAk = BCK(C.sA, k), Bik = BCK(C.Bi, k);
        const double An0k = BCK(An0, k), Bn0k = BCK(Bn0, k), Zn0k = BCK(Zn0, k);
        const double An1k = BCK(An1, k), Bn1k = BCK(Bn1, k), Zn1k = BCK(Zn1, k);
        const double s0k = BCK(s0, k), s1k = BCK(s1, k), x0k = BCK(C.a2, k), x1k = BCK(C.a3, k);
        SLOTS {
            const int i = c * T + ln.t;
            const double aik = lf.omk[c][k] * C.sA[c] * sAk;
            const double aki = lf.omk[c][k] * sAk * C.sA[c];
            // own row (i, k) for k <= i; row k's entry (k, i) above the diagonal
            const double own = lnfn_ik(A0, B0, Z0, lG0, G0, P, nu0, s0[c], C.Bi[c], C.a2[c], aik, i == k, Bik,
                                       An0k, Bn0k, Zn0k) +
                               lnfn_ik(A1, B1, Z1, lG1, G1, P, nu1, s1[c], C.Bi[c], C.a3[c], aik, i == k, Bik,
                                       An1k, Bn1k, Zn1k);
            const double other = lnfn_ik(A0, B0, Z0, lG0, G0, P, nu0, s0k, Bik, x0k, aki, false, C.Bi[c],
                                         An0[c], Bn0[c], Zn0[c]) +
                                 lnfn_ik(A1, B1, Z1, lG1, G1, P, nu1, s1k, Bik, x1k, aki, false, C.Bi[c],
                                         An1[c], Bn1[c], Zn1[c]);
            J[c][k] = (k <= i) ? own : other;
        }
    });
    C.cnt[6] += 2;
    C.cnt[7]++;
    double rhs[L], dn[L];
    SLOTS rhs[c] = C.a5[c];
    team_solve<NC, T, L>(J, rhs, ln.act, ln.t, dn);
    double alpha = 1;
    static_for<NC>([&](auto kc_) {
        constexpr int i = decltype(kc_)::value;
        const double a1i = BCK(C.a1, i), dni = BCK(dn, i);
        if (a1i + alpha * dni < 0) alpha *= 0.9 * fabs(a1i / dni);
    });
    double d[L];
    SLOTS {
        d[c] = alpha * dn[c];
        C.a1[c] = ln.act[c] ? C.a1[c] + d[c] : 0.0;
    }
    C.nu0 = TSUM(C.a1[c]);
    C.Sk = TSUM(2 * d[c] * d[c]);
    const double r0 = 1 / C.nu0, r1 = 1 / (1 - C.nu0);
    SLOTS {
        C.a2[c] = ln.act[c] ? fabs(C.a1[c] * r0) : 0.0;
        C.a3[c] = ln.act[c] ? fabs((C.z[c] - C.a1[c]) * r1) : 0.0;
        C.y[c] = C.a2[c];
    }
}

// ------------------------------------------------------------ one step
template <int NC, int T>
DEV void step(const DevFluid& f, const LaneFluid<NC, T>& lf, const PeIO& io, Cell<NC, T>& C,
              const Lane<NC, T>& ln)
{
    constexpr int L = Cell<NC, T>::L;
    if (++C.steps > MAX_STEPS) { C.cnt[7] |= 0x20000000; C.st = ST_DONE; return; }
    bool eval = true;
    double ev[L];                      // ln phi of this step's evaluation
    SLOTS ev[c] = 0;
    switch (C.st) {
    case ST_T_START: {
        const double ck = trial_coef(C.k);
        SLOTS C.y[c] = ln.act[c] ? C.z[c] / exp(ck * log_kw0(lf, c, C.logP, C.Tk)) : 0.0;
        normalise_y(C, ln);
        break;
    }
    case ST_T_SSM:   // Y = di / phiSta with a2 = ln phiSta
        SLOTS C.y[c] = ln.act[c] ? C.a1[c] * exp(-C.a2[c]) : 0.0;
        normalise_y(C, ln);
        break;
    case ST_A_ITER: {
        SLOTS C.y[c] = C.z[c] * C.a4[c];
        normalise_y(C, ln);
        const double dY = TSUM((C.y[c] - C.a3[c]) * (C.y[c] - C.a3[c]));
        if (dY < DYTOL) { C.flag = 1; eval = false; }
        break;
    }
    case ST_A_START: {
        const double ck = trial_coef(C.k);
        SLOTS {
            C.a4[c] = ln.act[c] ? exp(ck * log_kw0(lf, c, C.logP, C.Tk)) : 0.0;
            C.y[c] = C.z[c] * C.a4[c];
        }
        normalise_y(C, ln);
        C.iter = 0;
        C.flag = 0;
        break;
    }
    case ST_T_NR:
        stable_nr_pre(lf, C, ln);
        break;
    case ST_S_SSM_A: {
        // one SplitSSM2 iteration per step: the Rachford-Rice split, then
        // ln phi of both phases side by side (two independent evaluations
        // the compiler can overlap, which shortens a lone cell's iteration),
        // phase 0 at x = a2 into a5, phase 1 at a3 into ev and C's mixture
        rr2(f, C, ln);
        update_x_rr(C, ln);
        double s0[L], A0, B0, Z0;
        SLOTS C.y[c] = C.a3[c];
        mix<NC, T>(lf, ln, C, C.a2, s0, A0, B0, Z0);
        mix<NC, T>(lf, ln, C, C.y, C.s, C.A, C.B, C.Z);
        SLOTS {
            C.a5[c] = ln.act[c] ? lnphi_of(s0[c], C.Bi[c], A0, B0, Z0) : 0.0;
            ev[c] = ln.act[c] ? lnphi_of(C.s[c], C.Bi[c], C.A, C.B, C.Z) : 0.0;
        }
        C.cnt[5] += 2;
        C.st = ST_S_SSM_B;
        eval = false;
        break;
    }
    case ST_S_SSM_B:   // entered only from ST_S_SSM_A in the same step
        break;
    case ST_S_NR:
        split_nr_pre(lf, C, ln);
        break;
    default:   // ST_FEED: y = z
        break;
    }

    // common evaluator: ln phi at y
    if (eval) {
        mix<NC, T>(lf, ln, C, C.y, C.s, C.A, C.B, C.Z);
        SLOTS ev[c] = ln.act[c] ? lnphi_of(C.s[c], C.Bi[c], C.A, C.B, C.Z) : 0.0;
        C.cnt[5]++;
    }

    switch (C.st) {
    case ST_FEED:
        SLOTS C.a1[c] = ln.act[c] ? exp(ev[c]) * C.z[c] : 0.0;
        C.k = 0;
        C.st = C.lNP == 0 ? ST_T_START : ST_A_START;
        break;
    case ST_T_START:
    case ST_T_SSM: {
        // r = ln(fugSta / fugId * Yt).  After an SSM update Y Yt = di / phiSta
        // of the previous iterate, so r is the change of ln phiSta.
        double r[L];
        if (C.st == ST_T_START) {
            SLOTS r[c] = ln.act[c] ? log((exp(ev[c]) * C.y[c] * C.P) / (C.a1[c] * C.P) * C.Yt) : 0.0;
        } else {
            SLOTS r[c] = ln.act[c] ? ev[c] - C.a2[c] : 0.0;
        }
        SLOTS {
            C.a3[c] = -r[c];
            C.a2[c] = ln.act[c] ? ev[c] : 0.0;
        }
        const double Se = TSUM(r[c] * r[c]);
        if (C.st == ST_T_START) {
            C.iter = 0;
            if (Se > f.SSMsta_tol2) { C.st = ST_T_SSM; break; }
            nr_sta_begin(f, lf, C, io, ln);
            break;
        }
        C.cnt[0]++;
        if (Se > f.SSMsta_tol2 && ++C.iter <= f.SSMsta_maxIt) break;
        nr_sta_begin(f, lf, C, io, ln);
        break;
    }
    case ST_T_NR: {
        SLOTS C.a3[c] = ln.act[c] ? log((C.a1[c] * C.P) / ((exp(ev[c]) * C.y[c] * C.P) * C.Yt)) : 0.0;
        const double e2 = TSUM(C.a3[c] * C.a3[c]);
        C.cnt[1]++;
        if (++C.iter > f.NRsta_maxIt) { C.flag = 0; strict_decide(f, lf, C, io, ln); break; }
        if (sqrt(e2) > f.NRsta_tol) break;
        C.flag = 1;
        strict_decide(f, lf, C, io, ln);
        break;
    }
    case ST_A_START:
    case ST_A_ITER: {
        if (!eval) { relaxed_trial_end(f, lf, C, io, ln); break; }
        double di[L], lk[L];
        SLOTS {
            di[c] = (C.a1[c] * C.P) / ((exp(ev[c]) * C.y[c] * C.P) * C.Yt);
            C.a4[c] = ln.act[c] ? C.a4[c] * di[c] : 0.0;
            lk[c] = ln.act[c] ? log(C.a4[c]) : 0.0;
        }
        const double Se = TSUM((di[c] - 1) * (di[c] - 1));
        const double Sk = TSUM(lk[c] * lk[c]);
        SLOTS { C.a3[c] = C.y[c]; C.a2[c] = ev[c]; }
        C.Sk = Sk;
        C.cnt[0]++;
        C.iter++;
        if (Se < f.SSMsta_tol2 || Sk < KTOL) { C.flag = 1; relaxed_trial_end(f, lf, C, io, ln); break; }
        if (C.iter > f.SSMsta_maxIt) { C.flag = 0; relaxed_trial_end(f, lf, C, io, ln); break; }
        C.st = ST_A_ITER;
        break;
    }
    case ST_S_SSM_B: {
        C.cnt[2]++;
        C.iter++;
        double q[L], Kn[L];
        SLOTS {   // literal forms of :629-635 (overflow behaviour included)
            const double phi0 = exp(C.a5[c]), phi1 = exp(ev[c]);
            q[c] = ln.act[c] ? (phi1 * C.a3[c] * C.P) / (phi0 * C.a2[c] * C.P) - 1 : 0.0;
            Kn[c] = ln.act[c] ? phi1 / phi0 : 0.0;
        }
        const double Se = TSUM(q[c] * q[c]);
        bool last;
        if (C.iter > C.maxIt) { C.ssm_conv = 0; last = true; }
        else last = !(Se > C.tol2);
        SLOTS C.a4[c] = Kn[c];
        if (last) split_nr_begin(f, lf, C, io, ev, ln);
        else C.st = ST_S_SSM_A;
        break;
    }
    case ST_S_NR: {
        double lnphi1[L];
        SLOTS C.y[c] = C.a3[c];
        mix<NC, T>(lf, ln, C, C.y, C.s, C.A, C.B, C.Z);
        SLOTS lnphi1[c] = ln.act[c] ? lnphi_of(C.s[c], C.Bi[c], C.A, C.B, C.Z) : 0.0;
        C.cnt[5]++;
        SLOTS C.a5[c] = ln.act[c] ? log((exp(lnphi1[c]) * C.a3[c] * C.P) / (exp(ev[c]) * C.a2[c] * C.P)) : 0.0;
        const double eNR = sqrt(TSUM(C.a5[c] * C.a5[c]));
        C.cnt[3]++;
        C.iter++;
        bool fin = false;
        if (eNR > C.eNR0 || C.iter > f.NRsp_maxIt) fin = true;
        else if (C.Sk < 1E-16 * (2.0 * NC) * (2.0 * NC)) { C.nr_conv = 1; fin = true; }
        else if (!(eNR > f.NRsp_tol)) fin = true;
        if (fin) {
            if (eNR < f.NRsp_tol) C.nr_conv = 1;
            after_split_nr(f, lf, C, io, ln);
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

template <int NC, int T>
DEV Lane<NC, T> make_lane()
{
    Lane<NC, T> ln;
    ln.t = __lane_id() & (T - 1);
    UNROLL for (int c = 0; c < Lane<NC, T>::L; c++) ln.act[c] = c * T + ln.t < NC;
    return ln;
}

template <int NC, int T, bool MULTI>
__global__ void __launch_bounds__(64) pe_team_plain(const DevFluid* __restrict__ fp, PeIO io)
{
    const Lane<NC, T> ln = make_lane<NC, T>();
    const int cell = (blockIdx.x * 64 + threadIdx.x) / T;
    if (cell >= io.nb) return;                   // whole teams leave together
    LaneFluid<NC, T> lf;
    Cell<NC, T> C;
    C.model = MULTI ? io.mid[cell] : 0;
    begin<NC, T>(fp[C.model], lf, io, cell, C, ln);
    while (C.st != ST_DONE) step<NC, T>(fluid_of<NC, T, MULTI>(fp, C), lf, io, C, ln);
    end<NC, T>(io, C, ln);
}

// Persistent teams over the global work queue (counter io.queue, zeroed
// before the launch; io.qorder maps queue positions to cells).  Only team
// leaders vote; a team whose cell is done writes it back and takes the next
// index.  The wavefront leaves when the queue is drained and no team has a
// cell; a team waiting for a Newton batch steps as soon as only Newton
// teams are left, so every team reaches the exit.
template <int NC, int T, bool MULTI>
__global__ void __launch_bounds__(64) pe_team_batched(const DevFluid* __restrict__ fp, PeIO io)
{
    const Lane<NC, T> ln = make_lane<NC, T>();
    const int lane = __lane_id();
    const int lead = lane & ~(T - 1);
    constexpr int NR_BATCH = (64 / T) / 4;
    LaneFluid<NC, T> lf;
    Cell<NC, T> C;
    C.st = ST_DONE;
    C.model = 0;
    bool have = false, drained = false;
    const unsigned nq = io.nq ? *io.nq : (unsigned)io.nb;
    while (true) {
        if (C.st == ST_DONE && have) { end<NC, T>(io, C, ln); have = false; }
        const bool need = !have && !drained;
        const unsigned long long m = __ballot(need && ln.t == 0);
        if (m) {
            const int first = __ffsll((long long)m) - 1;
            unsigned base = 0;
            if (lane == first) base = atomicAdd(io.queue, (unsigned)__popcll(m));
            base = __shfl(base, first);
            if (need) {
                const unsigned idx = base + (unsigned)__popcll(m & ((1ull << lead) - 1ull));
                if (idx < nq) {
                    const int cell = io.qorder ? io.qorder[idx] : (int)idx;
                    C.model = MULTI ? io.mid[cell] : 0;
                    begin<NC, T>(fp[C.model], lf, io, cell, C, ln);
                    have = true;
                }
                else drained = true;
            }
        }
        if (!__any(have)) break;
        // Newton steps (matrix build + solve) cost several SSM steps: with 8
        // or more teams in a wavefront, a team that wants one waits until a
        // quarter of the teams do (or nothing else is left), so the Newton
        // branch runs for several teams at once.
        const bool live = have && C.st != ST_DONE;
        if constexpr (NR_BATCH > 1) {
            const bool nr = live && (C.st == ST_T_NR || C.st == ST_S_NR);
            const unsigned long long mnr = __ballot(nr && ln.t == 0), mlive = __ballot(live && ln.t == 0);
            const bool run_nr = __popcll(mnr) >= NR_BATCH || mnr == mlive;
            if (live && (!nr || run_nr)) step<NC, T>(fluid_of<NC, T, MULTI>(fp, C), lf, io, C, ln);
        } else if (live) {
            step<NC, T>(fluid_of<NC, T, MULTI>(fp, C), lf, io, C, ln);
        }
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
        if (multi) return hipErrorInvalidValue;
        hipLaunchKernelGGL((pe_team_plain<NC, T, false>), dim3((unsigned)need), dim3(64), 0, st, f, io);
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

// Team sizes built per Nc: 4 for every Nc (1-3 components per lane), 8 from
// Nc = 5 (1-2 per lane), 16 from Nc = 9 (one per lane).
bool pe_team_size_ok(int NC, int T)
{
    if (NC < 2 || NC > 12) return false;
    return T == 4 || (T == 8 && NC >= 5) || (T == 16 && NC >= 9);
}

hipError_t run_pe_team(int NC, int T, const DevFluid* f, const PeIO& io, bool batched, int grid_blocks,
                       hipStream_t st)
{
    if (!pe_team_size_ok(NC, T)) return hipErrorInvalidValue;
#define OCPX_T4(N) case N: return team::launch<N, 4>(f, io, batched, grid_blocks, st);
#define OCPX_T48(N) case N: return T == 4 ? team::launch<N, 4>(f, io, batched, grid_blocks, st) \
                                          : team::launch<N, 8>(f, io, batched, grid_blocks, st);
#define OCPX_T4816(N) case N: return T == 4 ? team::launch<N, 4>(f, io, batched, grid_blocks, st) \
                                 : (T == 8 ? team::launch<N, 8>(f, io, batched, grid_blocks, st) \
                                           : team::launch<N, 16>(f, io, batched, grid_blocks, st));
    switch (NC) {
        OCPX_T4(2) OCPX_T4(3) OCPX_T4(4) OCPX_T48(5) OCPX_T48(6) OCPX_T48(7) OCPX_T48(8)
        OCPX_T4816(9) OCPX_T4816(10) OCPX_T4816(11) OCPX_T4816(12)
    default: return hipErrorInvalidValue;
    }
#undef OCPX_T4
#undef OCPX_T48
#undef OCPX_T4816
}

int pe_team_blocks_per_cu(int NC, int T)
{
    if (!pe_team_size_ok(NC, T)) return 0;
#define OCPX_B4(N) case N: return team::blocks_per_cu<N, 4>();
#define OCPX_B48(N) case N: return T == 4 ? team::blocks_per_cu<N, 4>() : team::blocks_per_cu<N, 8>();
#define OCPX_B4816(N) case N: return T == 4 ? team::blocks_per_cu<N, 4>() \
                                 : (T == 8 ? team::blocks_per_cu<N, 8>() : team::blocks_per_cu<N, 16>());
    switch (NC) {
        OCPX_B4(2) OCPX_B4(3) OCPX_B4(4) OCPX_B48(5) OCPX_B48(6) OCPX_B48(7) OCPX_B48(8)
        OCPX_B4816(9) OCPX_B4816(10) OCPX_B4816(11) OCPX_B4816(12)
    default: return 0;
    }
#undef OCPX_B4
#undef OCPX_B48
#undef OCPX_B4816
}

}  // namespace ocpx
