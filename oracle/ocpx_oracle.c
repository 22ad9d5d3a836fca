/*
 * ocpx_oracle.c — plain-C restatement of the reference's flash + FIM-assembly
 * hot path (OpenCAEPoroX).
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker the HIP implementation is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product path never does.
 *
 * Each function cites the reference file:line it restates.  It keeps the
 * reference's control flow, iteration caps, tolerances and constant values
 * (PI = 3.141592653, delta1/2 = +/-(sqrt2 +/- 1) truncated to 11 digits,
 * Wilson 5.373, ...) so that discrete decisions (phase counts, labels,
 * SkipPSA) follow the reference.  The reference's LAPACK calls are replaced
 * by Gaussian elimination with partial pivoting on the same (mirrored)
 * matrices, and its float ssyevd_ by a cyclic Jacobi iteration on the same
 * float-rounded matrix.
 *
 * Parity pins (see tests/): EoS, cubic, PVTW, LBC, kr/Pc and rock are
 * checked against the reference's own sources built in oracle/_ref; the
 * phase equilibrium, FlashDer tail and FIM loops against closed forms,
 * invariants, brute-force Gibbs minimisation and central finite
 * differences.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include "oracle.h"

#define MC OCPX_MAX_NC
#define MB (OCPX_MAX_NC + 1)          /* bulk components incl. water */
#define NPH 3

/* METRIC constants, include/OCPUnits.hpp:78-100, include/OCPConst.hpp:30-32 */
static const double R_GAS = 0.083143;
static const double GRAV = 0.0000980665;
static const double DARCY = 0.008527;
static const double CONV4 = 273.15;
static const double PI_REF = 3.141592653;
static const double D1 = 2.41421356237;     /* include/OCPEoS.hpp:111 */
static const double D2 = -0.41421356237;    /* include/OCPEoS.hpp:112 */

/* ------------------------------------------------------------------------ */
/* Fluid set-up: EoS_PR ctor src/OCPEoS.cpp:19-65, ViscosityMethod03 ctor     */
/* src/OCPFuncPVT.cpp:415-459                                                 */
/* ------------------------------------------------------------------------ */
typedef struct {
    const ocpx_fluid* in;
    int NC;
    double vsh[MC];        /* Vshift * R*OmegaB*Tc/Pc                    */
    double lbc[5];         /* LBC coefficients x10                        */
    double auxB[MC], sqrtMW[MC];
} fl_t;

static void fl_init(fl_t* f, const ocpx_fluid* in)
{
    f->in = in;
    f->NC = in->nc;
    for (int i = 0; i < in->nc; i++) {
        f->vsh[i] = in->Vshift[i] * (R_GAS * in->OmegaB[i] * in->Tc[i] / in->Pc[i]);
        f->auxB[i] = 5.4402 * pow(in->Tc[i], 1.0 / 6) / pow(in->Pc[i], 2.0 / 3);
        f->sqrtMW[i] = sqrt(in->MW[i]);
    }
    for (int k = 0; k < 5; k++) f->lbc[k] = in->LBCcoef[k] * 10;
}

static double kij(const fl_t* f, int i, int k) { return f->in->BIC[i * MC + k]; }

/* ------------------------------------------------------------------------ */
/* Cubic roots: src/UtilMath.cpp:15-80                                       */
/* ------------------------------------------------------------------------ */
static void newton_polish(double* root, double a, double b, double c)
{
    double r = *root;
    double e = r * (r * (r + a) + b) + c;
    double best = r, beste = fabs(e);
    int it = 0;
    while (fabs(e) > 1E-8) {
        double df = r * (3 * r + 2 * a) + b;
        r = r - e / df;
        if (++it > 10) break;
        e = r * (r * (r + a) + b) + c;
        if (fabs(e) <= beste) { beste = fabs(e); best = r; }
    }
    *root = best;
}

static double sgn(double d) { return d > 0 ? 1.0 : (d < 0 ? -1.0 : 0.0); }

int orc_cubic_root(double a, double b, double c, double* z)
{
    double Q = (a * a - 3 * b) / 9;
    double R = (2 * a * a * a - 9 * a * b + 27 * c) / 54;
    double Q3 = Q * Q * Q;
    double M = R * R - Q3;
    if (M <= 0) {
        double th = acos(R / sqrt(Q3));
        z[0] = -2 * sqrt(Q) * cos(th / 3) - a / 3;
        z[1] = -2 * sqrt(Q) * cos((th + 2 * PI_REF) / 3) - a / 3;
        z[2] = -2 * sqrt(Q) * cos((th - 2 * PI_REF) / 3) - a / 3;
        for (int k = 0; k < 3; k++) newton_polish(&z[k], a, b, c);
        /* ascending sort of 3 values */
        for (int p = 0; p < 2; p++)
            for (int q = 0; q < 2 - p; q++)
                if (z[q] > z[q + 1]) { double t = z[q]; z[q] = z[q + 1]; z[q + 1] = t; }
        return 3;
    }
    double t1 = -R + sqrt(M), t2 = R + sqrt(M);
    double S = sgn(t1) * pow(fabs(t1), 1.0 / 3);
    double T = -sgn(t2) * pow(fabs(t2), 1.0 / 3);
    z[0] = S + T - a / 3;
    newton_polish(&z[0], a, b, c);
    return 1;
}

/* ------------------------------------------------------------------------ */
/* Peng-Robinson EoS: src/OCPEoS.cpp:68-414                                   */
/* ------------------------------------------------------------------------ */
typedef struct {
    const fl_t* f;
    double Ai[MC], Bi[MC], Aj, Bj, Zj;
    double Ax[MC], Bx[MC], Zx[MC], An[MC], Bn[MC], Zn[MC], Ap, Bp, Zp;
    long nfug, nder;
} eos_t;

static void eos_aibi(eos_t* e, double P, double T)          /* :68-84 */
{
    const ocpx_fluid* in = e->f->in;
    for (int i = 0; i < e->f->NC; i++) {
        double w = in->acf[i], m;
        if (w <= 0.49) m = 0.37464 + 1.54226 * w - 0.26992 * pow(w, 2);
        else m = 0.379642 + 1.48503 * w - 0.164423 * pow(w, 2) + 0.016667 * pow(w, 3);
        double Pr = P / in->Pc[i], Tr = T / in->Tc[i];
        e->Ai[i] = in->OmegaA[i] * Pr / pow(Tr, 2) * pow((1 + m * (1 - sqrt(Tr))), 2);
        e->Bi[i] = in->OmegaB[i] * Pr / Tr;
    }
}

static void eos_ajbj(eos_t* e, const double* x)            /* :87-99 */
{
    int n = e->f->NC;
    double A = 0, B = 0;
    for (int p = 0; p < n; p++) {
        B += e->Bi[p] * x[p];
        A += x[p] * x[p] * e->Ai[p] * (1 - kij(e->f, p, p));
        for (int q = 0; q < p; q++)
            A += 2 * x[p] * x[q] * sqrt(e->Ai[p] * e->Ai[q]) * (1 - kij(e->f, p, q));
    }
    e->Aj = A; e->Bj = B;
}

static void eos_zj(eos_t* e)                                /* :102-124 */
{
    double A = e->Aj, B = e->Bj, z[3];
    double a = (D1 + D2 - 1) * B - 1;
    double b = (A + D1 * D2 * B * B - (D1 + D2) * B * (B + 1));
    double c = -(A * B + D1 * D2 * B * B * (B + 1));
    if (orc_cubic_root(a, b, c, z) == 1) { e->Zj = z[0]; return; }
    double z1 = z[0], z2 = z[2];
    double dG = (z2 - z1) + log((z1 - B) / (z2 - B)) -
                A / (B * (D2 - D1)) *
                log((z1 + D1 * B) * (z2 + D2 * B) / ((z1 + D2 * B) * (z2 + D1 * B)));
    e->Zj = dG > 0 ? z1 : z2;
}

static void eos_abz(eos_t* e, double P, double T, const double* x)
{
    eos_aibi(e, P, T);
    eos_ajbj(e, x);
    eos_zj(e);
}

/* denominator shared by Zx, Zn, Zp (:144-179) */
static double eos_zden(const eos_t* e)
{
    double A = e->Aj, B = e->Bj, Z = e->Zj;
    return 3 * Z * Z + 2 * ((D1 + D2 - 1) * B - 1) * Z + (A + D1 * D2 * B * B - (D1 + D2) * B * (B + 1));
}
static double eos_znum(const eos_t* e, double dA, double dB)
{
    double A = e->Aj, B = e->Bj, Z = e->Zj;
    return (B - Z) * dA + ((A + D1 * D2 * (3 * B * B + 2 * B)) +
                           ((D1 + D2) * (2 * B + 1) - 2 * D1 * D2 * B) * Z -
                           (D1 + D2 - 1) * Z * Z) * dB;
}

static void eos_axbxzx(eos_t* e, const double* x)          /* :135-150 */
{
    int n = e->f->NC;
    double den = eos_zden(e);
    for (int i = 0; i < n; i++) {
        double s = 0;
        for (int k = 0; k < n; k++) s += x[k] * (1 - kij(e->f, i, k)) * sqrt(e->Ai[i] * e->Ai[k]);
        e->Ax[i] = 2 * s;
        e->Bx[i] = e->Bi[i];
        e->Zx[i] = eos_znum(e, e->Ax[i], e->Bx[i]) / den;
    }
}

static void eos_anbnzn(eos_t* e, const double* x, double nt)   /* :153-168 */
{
    int n = e->f->NC;
    double den = eos_zden(e);
    for (int i = 0; i < n; i++) {
        double s = 0;
        for (int m = 0; m < n; m++) s += (1 - kij(e->f, i, m)) * sqrt(e->Ai[i] * e->Ai[m]) * x[m];
        e->An[i] = 2 / nt * (s - e->Aj);
        e->Bn[i] = 1 / nt * (e->Bi[i] - e->Bj);
        e->Zn[i] = eos_znum(e, e->An[i], e->Bn[i]) / den;
    }
}

static void eos_apbpzp(eos_t* e, double P)                 /* :171-180 */
{
    e->Ap = e->Aj / P;
    e->Bp = e->Bj / P;
    e->Zp = eos_znum(e, e->Ap, e->Bp) / eos_zden(e);
}

/* sum_k (1-k_ik) sqrt(Ai Ak) x_k */
static double eos_asum(const eos_t* e, int i, const double* x)
{
    double s = 0;
    for (int k = 0; k < e->f->NC; k++) s += (1 - kij(e->f, i, k)) * sqrt(e->Ai[i] * e->Ai[k]) * x[k];
    return s;
}

/* CalFug / CalFugPhi :183-214 */
static void eos_fugphi(eos_t* e, double P, double T, const double* x, double* fug, double* phi)
{
    eos_abz(e, P, T, x);
    e->nfug++;
    double A = e->Aj, B = e->Bj, Z = e->Zj;
    double lz = log(Z - B), lg = log((Z + D1 * B) / (Z + D2 * B));
    for (int i = 0; i < e->f->NC; i++) {
        double s = 2 * eos_asum(e, i, x);
        double ph = exp(e->Bi[i] / B * (Z - 1) - lz - A / (D1 - D2) / B * (s / A - e->Bi[i] / B) * lg);
        if (phi) phi[i] = ph;
        fug[i] = ph * x[i] * P;
    }
}

/* CalLnFugX :217-251, row-major [i*NC + k] = d ln f_i / d x_k */
static void eos_lnfugx(eos_t* e, double P, double T, const double* x, double* out)
{
    int n = e->f->NC;
    eos_abz(e, P, T, x);
    eos_axbxzx(e, x);
    e->nder++;
    double A = e->Aj, B = e->Bj, Z = e->Zj;
    double G = (Z + D1 * B) / (Z + D2 * B);
    for (int i = 0; i < n; i++) {
        double C = x[i] * P / (Z - B);
        double E = -A / ((D1 - D2) * B) * (e->Ax[i] / A - e->Bx[i] / B);
        for (int k = 0; k < n; k++) {
            double aik = (1 - kij(e->f, i, k)) * sqrt(e->Ai[i] * e->Ai[k]);
            double Cx = ((Z - B) * (i == k ? 1.0 : 0.0) - x[i] * (e->Zx[k] - e->Bx[k])) * P / ((Z - B) * (Z - B));
            double Dx = e->Bx[i] / B * (e->Zx[k] - e->Bx[k] * (Z - 1) / B);
            double Ex = (2 * (A / B * e->Bx[k] * e->Bx[i] + B * aik) - e->Ax[i] * e->Bx[k] - e->Ax[k] * e->Bi[i]) / (B * B);
            Ex /= -(D1 - D2);
            double Gx = (D1 - D2) / (Z + D2 * B) / (Z + D2 * B) * (Z * e->Bx[k] - e->Zx[k] * B);
            out[i * n + k] = 1 / C * Cx + Dx + Ex * log(G) + E / G * Gx;
        }
    }
}

/* CalLnFugN :254-300 (flag 0) and CalLnPhiN :343-382 (flag 1) */
static void eos_lnfugn_impl(eos_t* e, double P, double T, const double* x, double nt, double* out, int phi)
{
    int n = e->f->NC;
    eos_abz(e, P, T, x);
    eos_anbnzn(e, x, nt);
    e->nder++;
    double A = e->Aj, B = e->Bj, Z = e->Zj;
    double G = (Z + D1 * B) / (Z + D2 * B);
    for (int i = 0; i < n; i++) {
        double C = phi ? 1 / (Z - B) : x[i] * P / (Z - B);
        double s = eos_asum(e, i, x);
        double E = -A / ((D1 - D2) * B) * (2 * s / A - e->Bi[i] / B);
        for (int k = 0; k <= i; k++) {
            double aik = (1 - kij(e->f, i, k)) * sqrt(e->Ai[i] * e->Ai[k]);
            double Cn;
            if (phi) Cn = (e->Bn[k] - e->Zn[k]) / ((Z - B) * (Z - B));
            else Cn = P / (Z - B) / (Z - B) * ((Z - B) / nt * ((i == k ? 1.0 : 0.0) - x[i]) - x[i] * (e->Zn[k] - e->Bn[k]));
            double Dn = e->Bi[i] / B * (e->Zn[k] - (e->Bi[k] - B) * (Z - 1) / (nt * B));
            double Gn = (D1 - D2) / ((Z + D2 * B) * (Z + D2 * B)) * (e->Bn[k] * Z - e->Zn[k] * B);
            double En = -1 / (D1 - D2) / (B * B) *
                        (2 * (B * aik / nt + e->Bn[k] * (e->Bi[i] * A / B - s)) - e->An[k] * e->Bi[i] - A * e->Bi[i] / nt);
            En -= E / nt;
            out[i * n + k] = 1 / C * Cn + Dn + En * log(G) + E / G * Gn;
            out[k * n + i] = out[i * n + k];
        }
    }
}

/* CalLnFugP :303-340 */
static void eos_lnfugp(eos_t* e, double P, double T, const double* x, double* out)
{
    int n = e->f->NC;
    eos_abz(e, P, T, x);
    eos_apbpzp(e, P);
    e->nder++;
    double A = e->Aj, B = e->Bj, Z = e->Zj;
    double G = (Z + D1 * B) / (Z + D2 * B);
    double Gp = (D1 - D2) / ((Z + D2 * B) * (Z + D2 * B)) * (e->Bp * Z - e->Zp * B);
    for (int i = 0; i < n; i++) {
        double C = P / (Z - B);
        double s = eos_asum(e, i, x);
        double E = -A / ((D1 - D2) * B) * (2 * s / A - e->Bi[i] / B);
        double Cp = ((Z - B) - P * (e->Zp - e->Bp)) / ((Z - B) * (Z - B));
        double Dp = e->Bi[i] / B * e->Zp;
        out[i] = Cp / C + Dp + E / G * Gp;
    }
}

/* CalVmDer :396-414 */
static double eos_vmder(eos_t* e, double P, double T, const double* x, double* vmP, double* vmx)
{
    int n = e->f->NC;
    eos_abz(e, P, T, x);
    eos_axbxzx(e, x);
    eos_apbpzp(e, P);
    double c = R_GAS * T / P;
    double v = e->Zj * c;
    for (int i = 0; i < n; i++) v -= x[i] * e->f->vsh[i];
    for (int i = 0; i < n; i++) vmx[i] = c * e->Zx[i] - e->f->vsh[i];
    *vmP = c * (e->Zp - e->Zj / P);
    return v;
}

/* ------------------------------------------------------------------------ */
/* Dense helpers replacing src/DenseMat.cpp LAPACK/BLAS calls                */
/* ------------------------------------------------------------------------ */
static double dnorm2(int n, const double* x)
{
    double s = 0;
    for (int i = 0; i < n; i++) s += x[i] * x[i];
    return sqrt(s);
}

/* Solve A X = B, A n x n row-major (destroyed), B n x nrhs stored column by
 * column (B[c*n + i]), Gaussian elimination with partial pivoting.  Returns
 * 0, or 1 when a zero pivot appears (LAPACK info > 0). */
static int gauss_solve(int n, double* A, int nrhs, double* B)
{
    int info = 0;
    for (int c = 0; c < n; c++) {
        int p = c;
        double best = fabs(A[c * n + c]);
        for (int r = c + 1; r < n; r++)
            if (fabs(A[r * n + c]) > best) { best = fabs(A[r * n + c]); p = r; }
        if (p != c) {
            for (int k = 0; k < n; k++) { double t = A[c * n + k]; A[c * n + k] = A[p * n + k]; A[p * n + k] = t; }
            for (int q = 0; q < nrhs; q++) { double t = B[q * n + c]; B[q * n + c] = B[q * n + p]; B[q * n + p] = t; }
        }
        double d = A[c * n + c];
        if (d == 0) { info = 1; continue; }
        for (int r = c + 1; r < n; r++) {
            double l = A[r * n + c] / d;
            if (l == 0) continue;
            for (int k = c + 1; k < n; k++) A[r * n + k] -= l * A[c * n + k];
            for (int q = 0; q < nrhs; q++) B[q * n + r] -= l * B[q * n + c];
        }
    }
    for (int q = 0; q < nrhs; q++) {
        double* b = B + q * n;
        for (int r = n - 1; r >= 0; r--) {
            double s = b[r];
            for (int k = r + 1; k < n; k++) s -= A[r * n + k] * b[k];
            b[r] = s / A[r * n + r];
        }
    }
    return info;
}

/* dsysv_ with uplo 'U' on a row-major array reads the row-major lower
 * triangle (column-major upper == row-major lower): mirror it, then solve. */
static int sys_solve_lower(int n, double* A, double* b)
{
    double S[MC * MC];
    for (int r = 0; r < n; r++)
        for (int c = 0; c < n; c++)
            S[r * n + c] = r >= c ? A[r * n + c] : A[c * n + r];
    return gauss_solve(n, S, 1, b);
}

/* ------------------------------------------------------------------------ */
/* Phase equilibrium: src/OCPPhaseEquilibrium.cpp                             */
/* ------------------------------------------------------------------------ */
typedef struct {
    const fl_t* f;
    eos_t e;
    int NC;
    double P, T, zi[MC];
    uint32_t lNP;
    int NP, ftype, testPId;
    double nu[2], x[2][MC], n[2][MC], ln[2][MC], MW[2], phi[2][MC], fug[2][MC];
    double Kw[4][MC], Ks[MC], lKs[MC], phiSta[MC], fugSta[MC], Y[MC], Yt, di[MC];
    double resSTA[MC], JmatSTA[MC * MC], lnfugX[MC * MC];
    double resSP[MC], JmatSP[MC * MC], lnfugN[2][MC * MC];
    int ssmsta_conflag, nrsta_conflag, ssmsp_conflag, nrsp_conflag;
    long it_ssmsta, it_nrsta, it_ssmsp, it_nrsp, it_rr, nsolve;
    /* ill-conditioned path markers (not part of the reference's state):
     * 1 stability SSM hit maxIt, 2 stability NR hit maxIt, 4 split SSM hit
     * maxIt, 8 split NR ended unconverged, 16 non-finite residual,
     * 32 RR phase fraction outside [-10, 11], 64 CheckSplit fallback,
     * 128 RR hit maxIt, 256 split restart cap (the reference would loop) */
    int ill;
} pe_t;

static void pe_kwilson(pe_t* s)                             /* :244-252 */
{
    const ocpx_fluid* in = s->f->in;
    for (int i = 0; i < s->NC; i++) {
        s->Kw[0][i] = (in->Pc[i] / s->P) * exp(5.373 * (1 + in->acf[i]) * (1 - in->Tc[i] / s->T));
        s->Kw[1][i] = 1 / s->Kw[0][i];
        s->Kw[2][i] = pow(s->Kw[0][i], 1.0 / 3);
        s->Kw[3][i] = pow(s->Kw[1][i], 1.0 / 3);
    }
}

static void pe_calmw(pe_t* s)                               /* :176-185 */
{
    for (int j = 0; j < s->NP; j++) {
        s->MW[j] = 0;
        for (int i = 0; i < s->NC; i++) s->MW[j] += s->x[j][i] * s->f->in->MW[i];
    }
}

static void pe_normalise_Y(pe_t* s)
{
    double yt = 0;
    for (int i = 0; i < s->NC; i++) yt += s->Y[i];
    s->Yt = yt;
    for (int i = 0; i < s->NC; i++) s->Y[i] *= 1 / yt;
}

/* StableNR :475-521; JmatSTA :523-541 */
static int pe_stable_nr(pe_t* s, int id)
{
    int n = s->NC;
    for (int i = 0; i < n; i++) s->resSTA[i] = log(s->fug[id][i] / (s->fugSta[i] * s->Yt));
    int maxIt = s->f->in->NRsta_maxIt;
    double tol = s->f->in->NRsta_tol;
    double Se = dnorm2(n, s->resSTA);
    int iter = 0;
    while (Se > tol) {
        eos_lnfugx(&s->e, s->P, s->T, s->Y, s->lnfugX);
        for (int i = 0; i < n; i++) {
            double t = 0;
            for (int k = 0; k < n; k++) t += s->Y[k] * s->lnfugX[i * n + k];
            for (int j = 0; j < n; j++) s->JmatSTA[i * n + j] = (s->lnfugX[i * n + j] - t + 1) / s->Yt;
        }
        sys_solve_lower(n, s->JmatSTA, s->resSTA);
        s->nsolve++;
        double lYt = s->Yt;
        for (int i = 0; i < n; i++) s->Y[i] = s->Y[i] * lYt + 1.0 * s->resSTA[i];
        pe_normalise_Y(s);
        eos_fugphi(&s->e, s->P, s->T, s->Y, s->fugSta, NULL);
        for (int i = 0; i < n; i++) s->resSTA[i] = log(s->fug[id][i] / (s->fugSta[i] * s->Yt));
        Se = dnorm2(n, s->resSTA);
        iter++;
        if (iter > maxIt) {
            s->it_nrsta += iter;
            s->nrsta_conflag = 0;
            s->ill |= 2;
            return 0;
        }
    }
    s->it_nrsta += iter;
    s->nrsta_conflag = 1;
    return 1;
}

/* StableSSM (strict, 4 Wilson trial sets) :397-473 */
static int pe_stable_ssm(pe_t* s, int id)
{
    int n = s->NC;
    const double* xj = s->x[id];
    eos_fugphi(&s->e, s->P, s->T, xj, s->fug[id], s->phi[id]);
    for (int i = 0; i < n; i++) s->di[i] = s->phi[id][i] * xj[i];
    double tol2 = s->f->in->SSMsta_tol * s->f->in->SSMsta_tol;
    int maxIt = s->f->in->SSMsta_maxIt;
    double eYt = s->f->in->SSMsta_eYt;
    for (int k = 0; k < 4; k++) {
        for (int i = 0; i < n; i++) s->Y[i] = xj[i] / s->Kw[k][i];
        pe_normalise_Y(s);
        eos_fugphi(&s->e, s->P, s->T, s->Y, s->fugSta, s->phiSta);
        double Se = 0;
        for (int i = 0; i < n; i++) Se += pow(log(s->fugSta[i] / s->fug[id][i] * s->Yt), 2);
        int iter = 0;
        while (Se > tol2) {
            for (int i = 0; i < n; i++) s->Y[i] = s->di[i] / s->phiSta[i];
            pe_normalise_Y(s);
            eos_fugphi(&s->e, s->P, s->T, s->Y, s->fugSta, s->phiSta);
            Se = 0;
            for (int i = 0; i < n; i++) Se += pow(log(s->fugSta[i] / s->fug[id][i] * s->Yt), 2);
            if (++iter > maxIt) { s->ill |= 1; break; }
        }
        s->it_ssmsta += iter;
        int flag = pe_stable_nr(s, id);
        if (flag && s->Yt > 1 + eYt) {
            s->ssmsta_conflag = 1;
            return 0;
        }
    }
    return 1;
}

/* StableSSM01 (relaxed, 2 trial sets) :299-395 */
static int pe_stable_ssm01(pe_t* s, int id)
{
    int n = s->NC;
    const double* xj = s->x[id];
    double tol2 = s->f->in->SSMsta_tol * s->f->in->SSMsta_tol;
    int maxIt = s->f->in->SSMsta_maxIt;
    double eYt = s->f->in->SSMsta_eYt;
    const double Ktol = 1E-4, dYtol = 1E-6;   /* include/OCPPhaseEquilibrium.hpp:43-44 */
    double Se = 0, Sk = 0;
    eos_fugphi(&s->e, s->P, s->T, xj, s->fug[id], NULL);
    for (int k = 0; k < 2; k++) {
        double* ks = s->Ks;
        memcpy(ks, s->Kw[k], sizeof(double) * n);
        int iter = 0, flag = 0;
        while (1) {
            for (int i = 0; i < n; i++) s->Y[i] = xj[i] * ks[i];
            pe_normalise_Y(s);
            if (iter > 0) {
                double dY = 0;
                for (int i = 0; i < n; i++) dY += pow((s->Y[i] - s->di[i]), 2);
                if (dY < dYtol) { flag = 1; break; }
            }
            eos_fugphi(&s->e, s->P, s->T, s->Y, s->fugSta, NULL);
            Se = 0; Sk = 0;
            for (int i = 0; i < n; i++) {
                s->di[i] = s->fug[id][i] / (s->fugSta[i] * s->Yt);
                ks[i] *= s->di[i];
                Se += pow(s->di[i] - 1, 2);
                Sk += pow(log(ks[i]), 2);
            }
            iter++;
            if (Se < tol2) { flag = 1; break; }
            if (Sk < Ktol) { flag = 1; break; }
            if (iter > maxIt) { s->ill |= 1; break; }
            memcpy(s->di, s->Y, sizeof(double) * n);
        }
        s->it_ssmsta += iter;
        if (flag && s->Yt > 1 - 0.1 && Sk > 1) s->ftype = 2;
        if (flag && s->Yt > 1 + eYt) {
            s->ssmsta_conflag = 1;
            return 0;
        }
    }
    return 1;
}

/* PhaseStable :260-297 */
static int pe_phase_stable(pe_t* s)
{
    if (s->NP == 1) s->testPId = 0;
    else {
        pe_calmw(s);
        int id = 0;
        double mw = s->MW[0];
        for (int j = 1; j < s->NP; j++) if (mw < s->MW[j]) { mw = s->MW[j]; id = j; }
        s->testPId = id;
    }
    s->ssmsta_conflag = 0;
    s->nrsta_conflag = 0;
    int flag;
    uint32_t tmpNP = (uint32_t)s->NP;
    if (s->lNP == 0) {
        flag = pe_stable_ssm(s, s->testPId);
    } else {
        flag = pe_stable_ssm01(s, s->testPId);
        if (!flag) tmpNP++;
        if (tmpNP != s->lNP) {
            flag = pe_stable_ssm(s, s->testPId);
            s->ftype = 2;
        }
    }
    return flag;
}

/* RachfordRice2 :651-705 */
static void pe_rr2(pe_t* s)
{
    int n = s->NC;
    const double* K = s->Ks;
    double Kmin = K[0], Kmax = K[0];
    for (int i = 1; i < n; i++) {
        if (K[i] < Kmin) Kmin = K[i];
        if (K[i] > Kmax) Kmax = K[i];
    }
    const double numin = 1 / (1 - Kmax), numax = 1 / (1 - Kmin);
    double nu = 0.5 * (numin + numax);
    const double tol = s->f->in->RR_tol;
    const double maxIt = s->f->in->RR_maxIt;
    int iter = 0;
    while (1) {
        double rj = 0, J = 0;
        for (int i = 0; i < n; i++) {
            double t = 1 + nu * (K[i] - 1);
            rj += s->zi[i] * (K[i] - 1) / t;
            J -= s->zi[i] * (K[i] - 1) * (K[i] - 1) / (t * t);
        }
        if (fabs(rj) < tol) break;
        if (iter > maxIt) { s->ill |= 128; break; }
        double d = -rj / J, tn = nu + d;
        if (tn < numax && tn > numin) nu = tn;
        else nu = d > 0 ? (nu + numax) / 2 : (nu + numin) / 2;
        iter++;
    }
    s->it_rr += iter;
    if (!(nu > -10 && nu < 11)) s->ill |= 32;
    s->nu[0] = nu;
    s->nu[1] = 1 - nu;
}

/* UpdateXRR :778-791 (NP = 2) */
static void pe_update_x_rr(pe_t* s)
{
    for (int i = 0; i < s->NC; i++) {
        double t = 1 + s->nu[0] * (s->Ks[i] - 1);
        s->x[1][i] = s->zi[i] / t;
        s->x[0][i] = s->Ks[i] * s->x[1][i];
    }
}

/* SplitSSM2 :586-647 */
static void pe_split_ssm2(pe_t* s, int restart)
{
    int n = s->NC;
    s->ssmsp_conflag = 1;
    double Se = 1;
    double tol2 = s->f->in->SSMsp_tol * s->f->in->SSMsp_tol;
    int maxIt = s->f->in->SSMsp_maxIt;
    if (!restart) {
        if (s->lNP == 2) memcpy(s->Ks, s->lKs, sizeof(double) * n);
        else memcpy(s->Ks, s->Kw[0], sizeof(double) * n);
    } else {
        tol2 *= 1E-1;
        maxIt *= 2;
    }
    if (s->Yt < 1.1) {
        tol2 *= 1E-1;
        maxIt *= 2;
    }
    int iter = 0;
    while (Se > tol2) {
        pe_rr2(s);
        pe_update_x_rr(s);
        for (int j = 0; j < 2; j++) eos_fugphi(&s->e, s->P, s->T, s->x[j], s->fug[j], s->phi[j]);
        Se = 0;
        for (int i = 0; i < n; i++) {
            Se += pow(s->fug[1][i] / s->fug[0][i] - 1, 2);
            s->Ks[i] = s->phi[1][i] / s->phi[0][i];
        }
        if (!isfinite(Se)) s->ill |= 16;
        if (++iter > maxIt) { s->ssmsp_conflag = 0; s->ill |= 4; break; }
    }
    s->it_ssmsp += iter;
}

static void pe_res_sp(pe_t* s)                              /* :899-907 */
{
    for (int i = 0; i < s->NC; i++) s->resSP[i] = log(s->fug[1][i] / s->fug[0][i]);
}

/* SplitNR :805-897 with AssembleJmatSP :911-937, CalStepNRsp :939-957 */
static void pe_split_nr(pe_t* s)
{
    int n = s->NC;
    s->nrsp_conflag = 0;
    for (int j = 0; j < 2; j++)
        for (int i = 0; i < n; i++) s->n[j][i] = s->nu[j] * s->x[j][i];
    pe_res_sp(s);
    double eNR = dnorm2(n, s->resSP), eNR0 = eNR;
    double tol = s->f->in->NRsp_tol;
    int iter = 0;
    while (eNR > tol) {
        memcpy(s->ln, s->n, sizeof(s->n));
        for (int j = 0; j < 2; j++) eos_lnfugn_impl(&s->e, s->P, s->T, s->x[j], s->nu[j], s->lnfugN[j], 0);
        for (int i = 0; i < n; i++)
            for (int k = 0; k < n; k++) s->JmatSP[i * n + k] = s->lnfugN[0][i * n + k] + s->lnfugN[1][i * n + k];
        sys_solve_lower(n, s->JmatSP, s->resSP);
        s->nsolve++;
        double alpha = 1;
        for (int i = 0; i < n; i++) {
            double t = s->n[0][i] + alpha * s->resSP[i];
            if (t < 0) alpha *= 0.9 * fabs(s->n[0][i] / s->resSP[i]);
        }
        s->nu[1] = 1;
        s->nu[0] = 0;
        for (int i = 0; i < n; i++) {
            s->n[1][i] = s->zi[i];
        }
        for (int i = 0; i < n; i++) {
            s->n[0][i] += alpha * s->resSP[i];
            s->n[1][i] -= s->n[0][i];
            s->nu[0] += s->n[0][i];
        }
        s->nu[1] -= s->nu[0];
        for (int i = 0; i < n; i++) s->x[0][i] = fabs(s->n[0][i] / s->nu[0]);
        for (int i = 0; i < n; i++) s->x[1][i] = fabs(s->n[1][i] / s->nu[1]);
        for (int j = 0; j < 2; j++) eos_fugphi(&s->e, s->P, s->T, s->x[j], s->fug[j], NULL);
        pe_res_sp(s);
        eNR = dnorm2(n, s->resSP);
        iter++;
        if (eNR > eNR0 || iter > s->f->in->NRsp_maxIt) break;
        double en = 0;
        for (int j = 0; j < 2; j++)
            for (int i = 0; i < n; i++) en += (s->ln[j][i] - s->n[j][i]) * (s->ln[j][i] - s->n[j][i]);
        if (en < 1E-16 * (2.0 * n) * (2.0 * n)) { s->nrsp_conflag = 1; break; }
    }
    if (eNR < tol) s->nrsp_conflag = 1;
    if (!s->nrsp_conflag) s->ill |= 8;
    if (!isfinite(eNR)) s->ill |= 16;
    s->it_nrsp += iter;
}

/* CheckSplit :960-1010 */
static int pe_check_split(pe_t* s)
{
    if (s->NP == 2) {
        double eX = 0, nuMax = s->nu[0] > s->nu[1] ? s->nu[0] : s->nu[1];
        for (int i = 0; i < s->NC; i++) eX += (s->x[0][i] - s->x[1][i]) * (s->x[0][i] - s->x[1][i]);
        if (nuMax < 1 && s->nrsp_conflag && isfinite(eX)) {
            /* accept */
        } else if (!isfinite(eX) || 1 - nuMax < 1E-3) {
            s->ill |= 64;
            s->NP = 1;
            memcpy(s->x[0], s->zi, sizeof(double) * s->NC);
            s->nu[0] = 1;
            s->ssmsta_conflag = 0;
            s->nrsta_conflag = 0;
            return 0;
        }
    }
    return 1;
}

/* PhaseSplit :549-573 */
static void pe_phase_split(pe_t* s)
{
    s->ssmsp_conflag = 0;
    s->nrsp_conflag = 0;
    pe_split_ssm2(s, 0);
    pe_split_nr(s);
    int restarts = 0;
    while (!s->nrsp_conflag) {
        /* the reference's loop has no cap and can cycle forever; stop after
         * 64 restarts like the HIP kernel and mark the cell */
        if (++restarts > 64) { s->ill |= 256; break; }
        pe_split_ssm2(s, 1);
        pe_split_nr(s);
        if (!pe_check_split(s)) break;
        if (s->ssmsp_conflag) break;
    }
    pe_check_split(s);
}

/* PhaseEquilibrium :71-112 with SetInitalValue :121-138.  lx uses the bulk
 * stride nb (OCPMixtureVarSet::x, nc = Nc + 1). */
static void pe_run(pe_t* s, double P, double T, const double* z, int ftype, uint32_t lNP,
                   const double* lx, int nb)
{
    int n = s->NC;
    s->P = P; s->T = T;
    memcpy(s->zi, z, sizeof(double) * n);
    s->ftype = ftype;
    s->lNP = lNP;
    if (lNP == 2)
        for (int i = 0; i < n; i++) s->lKs[i] = lx[i] / lx[i + nb];
    switch (ftype) {
    case 0:
        s->NP = 1;
        s->nu[0] = 1;
        memcpy(s->x[0], s->zi, sizeof(double) * n);
        pe_kwilson(s);
        while (!pe_phase_stable(s)) {
            s->NP++;
            pe_phase_split(s);
            if (s->NP == 2 || s->NP == 1) break;
        }
        break;
    case 1:
        s->NP = 1;
        s->nu[0] = 1;
        memcpy(s->x[0], s->zi, sizeof(double) * n);
        break;
    default:
        s->NP = 2;
        s->Yt = 1.01;
        pe_kwilson(s);
        pe_phase_split(s);
        break;
    }
    if (s->NP > 1) s->ftype = 2;
}

static void pe_init(pe_t* s, const fl_t* f)
{
    memset(s, 0, sizeof(*s));
    s->f = f;
    s->e.f = f;
    s->NC = f->NC;
}

static void pe_counters(const pe_t* s, int32_t* it)
{
    it[0] = (int32_t)s->it_ssmsta; it[1] = (int32_t)s->it_nrsta;
    it[2] = (int32_t)s->it_ssmsp;  it[3] = (int32_t)s->it_nrsp;
    it[4] = (int32_t)s->it_rr;     it[5] = (int32_t)s->e.nfug;
    it[6] = (int32_t)s->e.nder;    it[7] = (int32_t)s->nsolve;
}

void orc_phase_equilibrium(const ocpx_fluid* fin, int64_t ncell, const double* P,
                           const double* T, const double* z, const int32_t* ftype,
                           const int32_t* lNP, const double* lx, int32_t* NP,
                           double* nu, double* x, int32_t* ftype_out, int32_t* iters,
                           int32_t* ill)
{
    fl_t f;
    fl_init(&f, fin);
    int n = f.NC, nb = n + 1;
    pe_t* s = (pe_t*)malloc(sizeof(pe_t));
    for (int64_t c = 0; c < ncell; c++) {
        pe_init(s, &f);
        pe_run(s, P[c], T[c], z + c * n, ftype ? ftype[c] : 0,
               lNP ? (uint32_t)lNP[c] : 0u, lx ? lx + c * 2 * nb : NULL, nb);
        NP[c] = s->NP;
        nu[2 * c] = s->nu[0];
        nu[2 * c + 1] = s->NP > 1 ? s->nu[1] : 0;
        for (int j = 0; j < 2; j++)
            for (int i = 0; i < n; i++) x[(c * 2 + j) * n + i] = (j < s->NP) ? s->x[j][i] : 0;
        if (ftype_out) ftype_out[c] = s->ftype;
        if (iters) pe_counters(s, iters + 8 * c);
        if (ill) ill[c] = s->ill;
    }
    free(s);
}

/* ------------------------------------------------------------------------ */
/* Water and viscosity: src/OCPFuncPVT.cpp:32-57, 462-575                     */
/* ------------------------------------------------------------------------ */
void orc_pvtw(const ocpx_fluid* f, double P, double* o)
{
    /* single PVTW row: GetCloseRow returns row 0 (src/OCPTable.cpp:312-345) */
    const double Pref = f->PVTW[0], bref = f->PVTW[1], Cb = f->PVTW[2];
    const double muref = f->PVTW[3], Cmu = f->PVTW[4];
    double b = bref * (1 - Cb * (P - Pref));
    double bp = -Cb * bref;
    double mu = muref * (1 + Cmu * (P - Pref));
    double muP = Cmu * muref;
    const double stdVw = 1.0, CONV1 = 1.0;
    o[1] = 1 / CONV1 / (b * stdVw);                       /* xi   */
    o[4] = -1 / CONV1 * bp / ((b * b) * stdVw);           /* xiP  */
    o[0] = f->rhoW_std / b;                               /* rho  */
    o[3] = -f->rhoW_std * bp / (b * b);                   /* rhoP */
    o[2] = mu;
    o[5] = muP;
}

typedef struct { double MW, xPc, xTc, xVc, a0, a1, a2, a3, a4; } lbc_aux;

static double lbc_mu(const fl_t* f, double T, const double* x, double xi, lbc_aux* A)  /* :462-498 */
{
    const ocpx_fluid* in = f->in;
    int n = f->NC;
    double MW = 0;
    for (int i = 0; i < n; i++) MW += x[i] * in->MW[i];
    double a0 = 0, a1 = 0, xPc = 0, xTc = 0, xVc = 0;
    for (int i = 0; i < n; i++) {
        double t = f->auxB[i] / sqrt(MW);
        double Tr = T / in->Tc[i];
        if (Tr <= 1.5) t = 34 * 1E-5 * pow(Tr, 0.94) / t;
        else t = 17.78 * 1E-5 * pow((4.58 * Tr - 1.67), 0.625) / t;
        a0 += x[i] * f->sqrtMW[i] * t;
        a1 += x[i] * f->sqrtMW[i];
        xPc += x[i] * in->Pc[i];
        xTc += x[i] * in->Tc[i];
        xVc += x[i] * in->Vcvis[i];
    }
    double a2 = 5.4402 * pow(xTc, 1.0 / 6) / sqrt(MW) / pow(xPc, 2.0 / 3);
    double a3 = xi * xVc;
    const double* c = f->lbc;
    double a4 = a3 * (a3 * (a3 * (c[4] * a3 + c[3]) + c[2]) + c[1]) + c[0];
    A->MW = MW; A->xPc = xPc; A->xTc = xTc; A->xVc = xVc;
    A->a0 = a0; A->a1 = a1; A->a2 = a2; A->a3 = a3; A->a4 = a4;
    return a0 / a1 + 1E-4 * (pow(a4, 4) - 1) / a2;
}

/* :501-575; the d(sqrt MW) term is zero because -(1/2) is integer division
 * (:537), and the auxA[3] <= 0.18 branch is dead (:491, 515, 556). */
static double lbc_der(const fl_t* f, double T, const double* x, double xi, double xiP,
                      const double* xix, double* muP, double* mux)
{
    const ocpx_fluid* in = f->in;
    int n = f->NC;
    lbc_aux A;
    double mu = lbc_mu(f, T, x, xi, &A);
    const double* c = f->lbc;
    double d8 = 0;
    {
        double d7 = A.xVc * xiP;
        d8 = d7 * (c[1] + A.a3 * (2 * c[2] + A.a3 * (3 * c[3] + A.a3 * 4 * c[4])));
        *muP = (4 * 1E-4) * pow(A.a4, 3) * d8 / A.a2;
    }
    for (int k = 0; k < n; k++) {
        double dTc = in->Tc[k], dPc = in->Pc[k], dMW = in->MW[k];
        double d3 = 0;
        for (int i = 0; i < n; i++) {
            double v1 = f->auxB[i] / sqrt(A.MW);
            /* the reference's -(1/2)*... is 0 (integer division); lbc_exact
             * restores the true derivative of 1/sqrt(MW) */
            double d1 = (in->lbc_exact ? -0.5 : 0.0) * f->auxB[i] * pow(A.MW, -1.5) * dMW;
            double Tr = T / in->Tc[i], t;
            if (Tr <= 1.5) t = 34 * 1E-5 * pow(Tr, 0.94);
            else t = 17.78 * 1E-5 * pow(4.58 * Tr - 1.67, 0.625);
            double v2 = t / v1;
            double d2 = -t * d1 / (v1 * v1);
            d3 += x[i] * f->sqrtMW[i] * d2 + (i == k ? 1.0 : 0.0) * f->sqrtMW[k] * v2;
        }
        double d4 = f->sqrtMW[k];
        double d6 = 5.4402 *
                    (1.0 / 6 * pow(A.xTc, -5.0 / 6) * dTc -
                     pow(A.xTc, 1.0 / 6) * (0.5 / A.MW * dMW + 2.0 / 3 / A.xPc * dPc)) /
                    (sqrt(A.MW) * pow(A.xPc, 2.0 / 3));
        double d7 = xix[k] * A.xVc + xi * in->Vcvis[k];
        d8 = d7 * (c[1] + A.a3 * (2 * c[2] + A.a3 * (3 * c[3] + A.a3 * 4 * c[4])));
        mux[k] = (d3 * A.a1 - A.a0 * d4) / (A.a1 * A.a1) +
                 1E-4 * (4 * pow(A.a4, 3) * d8 * A.a2 - (pow(A.a4, 4) - 1) * d6) / (A.a2 * A.a2);
    }
    return mu;
}

double orc_lbc(const ocpx_fluid* fin, double T, const double* x, double xi, double xiP,
               const double* xix, double* muP, double* mux)
{
    fl_t f;
    fl_init(&f, fin);
    return lbc_der(&f, T, x, xi, xiP, xix, muP, mux);
}

/* ------------------------------------------------------------------------ */
/* OCPMixtureMethodComp01::FlashDer  src/OCPMixtureMethodComp.cpp:746-755     */
/* ------------------------------------------------------------------------ */
typedef struct {
    int np, nc;
    double P, T, Ni[MB], S[NPH];
    uint32_t phaseExist[NPH];
    uint32_t phaseNum;
    double x[NPH * MB], Nt, Vf, nj[NPH], vj[NPH], rho[NPH], xi[NPH], mu[NPH];
    double vfP, vfi[MB], vjP[NPH], vji[NPH][MB];
    double rhoP[NPH], xiP[NPH], muP[NPH], rhox[NPH * MB], xix[NPH * MB], mux[NPH * MB];
    double dXsdXp[NPH * (MB + 1) * (MB + 1)];
} mvs_t;

typedef struct {
    pe_t pe;
    int NC, NP;
    double Nt, zi[MC], MW[2], vm[2], vmP[2], vmx[2][MC];
    int label0_is_oil, nep, ep[2];
    double lnfugP[2][MC], lnfugN[2][MC * MC], Jmat[MC * MC], rhs[MC * (MC + 1)];
} mix_t;

static void swapd(double* a, double* b, int n)
{
    for (int i = 0; i < n; i++) { double t = a[i]; a[i] = b[i]; b[i] = t; }
}

static void mix_flash_der(mix_t* m, mvs_t* vs, int ftype)
{
    const fl_t* f = m->pe.f;
    const ocpx_fluid* in = f->in;
    const int NC = m->NC, nc = vs->nc, np = vs->np;
    const int wP = np - 1, wC = nc - 1;
    eos_t* e = &m->pe.e;

    /* InitNtZ :812-818 */
    m->Nt = 0;
    for (int i = 0; i < NC; i++) m->Nt += vs->Ni[i];
    for (int i = 0; i < NC; i++) m->zi[i] = vs->Ni[i] / m->Nt;
    vs->Nt = m->Nt + vs->Ni[wC];

    /* PE.PhaseEquilibrium(P, T, zi, ftype, phaseNum - 1, x, nc) :749 */
    pe_run(&m->pe, vs->P, vs->T, m->zi, ftype, vs->phaseNum - 1u, vs->x, nc);

    /* CalPropertyDer :134-142 -> CopyPhaseFromPE :82-89 */
    m->NP = m->pe.NP;
    for (int j = 0; j < m->NP; j++) {
        vs->nj[j] = m->Nt * m->pe.nu[j];
        memcpy(&vs->x[j * nc], m->pe.x[j], sizeof(double) * NC);
    }
    /* CalMW :97-105 */
    for (int j = 0; j < m->NP; j++) {
        m->MW[j] = 0;
        for (int i = 0; i < NC; i++) m->MW[j] += vs->x[j * nc + i] * in->MW[i];
    }
    /* CalVmVj :107-113 */
    for (int j = 0; j < m->NP; j++) {
        m->vm[j] = eos_vmder(e, vs->P, vs->T, &vs->x[j * nc], &m->vmP[j], m->vmx[j]);
        vs->vj[j] = vs->nj[j] * m->vm[j];
    }
    /* CalXiRhoMuDer :144-166 */
    for (int j = 0; j < m->NP; j++) {
        vs->xi[j] = 1 / m->vm[j];
        vs->xiP[j] = -vs->xi[j] * vs->xi[j] * m->vmP[j];
        for (int i = 0; i < NC; i++) vs->xix[j * nc + i] = -vs->xi[j] * vs->xi[j] * m->vmx[j][i];
        vs->rho[j] = m->MW[j] * vs->xi[j];
        vs->rhoP[j] = m->MW[j] * vs->xiP[j];
        for (int i = 0; i < NC; i++) vs->rhox[j * nc + i] = vs->xix[j * nc + i] * m->MW[j] + vs->xi[j] * in->MW[i];
        vs->mu[j] = lbc_der(f, vs->T, &vs->x[j * nc], vs->xi[j], vs->xiP[j], &vs->xix[j * nc],
                            &vs->muP[j], &vs->mux[j * nc]);
    }
    /* IdentifyPhase :174-214 */
    vs->phaseExist[0] = 0;
    vs->phaseExist[1] = 0;
    if (m->NP == 1) {
        double A = 0, B = 0;
        for (int i = 0; i < NC; i++) {
            A += m->zi[i] * in->Vc[i] * in->Tc[i];
            B += m->zi[i] * in->Vc[i];
        }
        double Tcm = A / B;
        if (vs->T > Tcm) { m->label0_is_oil = 0; vs->phaseExist[1] = 1; }
        else { m->label0_is_oil = 1; vs->phaseExist[0] = 1; }
    } else {
        m->label0_is_oil = m->MW[0] > m->MW[1];
        vs->phaseExist[0] = 1;
        vs->phaseExist[1] = 1;
    }
    m->nep = 0;
    for (int j = 0; j < 2; j++) if (vs->phaseExist[j]) m->ep[m->nep++] = j;   /* NPmax = 2 */
    /* ReOrderPhaseDer :235-256 */
    if (!m->label0_is_oil) {
        swapd(&vs->nj[0], &vs->nj[1], 1);
        swapd(&vs->x[0], &vs->x[nc], NC);
        swapd(&m->MW[0], &m->MW[1], 1);
        swapd(&vs->vj[0], &vs->vj[1], 1);
        swapd(&m->vm[0], &m->vm[1], 1);
        swapd(&m->vmP[0], &m->vmP[1], 1);
        swapd(m->vmx[0], m->vmx[1], NC);
        swapd(&vs->xi[0], &vs->xi[1], 1);
        swapd(&vs->xiP[0], &vs->xiP[1], 1);
        swapd(&vs->xix[0], &vs->xix[nc], NC);
        swapd(&vs->rho[0], &vs->rho[1], 1);
        swapd(&vs->rhoP[0], &vs->rhoP[1], 1);
        swapd(&vs->rhox[0], &vs->rhox[nc], NC);
        swapd(&vs->mu[0], &vs->mu[1], 1);
        swapd(&vs->muP[0], &vs->muP[1], 1);
        swapd(&vs->mux[0], &vs->mux[nc], NC);
    }

    /* CalPropertyW :846-860 with vw = -1 */
    {
        double o6[6];
        vs->phaseExist[wP] = 1;
        vs->x[wP * nc + wC] = 1.0;
        orc_pvtw(in, vs->P, o6);
        vs->rho[wP] = o6[0]; vs->xi[wP] = o6[1]; vs->mu[wP] = o6[2];
        vs->rhoP[wP] = o6[3]; vs->xiP[wP] = o6[4]; vs->muP[wP] = o6[5];
        vs->nj[wP] = vs->Ni[wC];
        vs->vj[wP] = vs->nj[wP] / vs->xi[wP];
        vs->vji[wP][wC] = 1 / vs->xi[wP];
        vs->vjP[wP] = -vs->nj[wP] * vs->xiP[wP] / (vs->xi[wP] * vs->xi[wP]);
    }
    /* CalVfS include/OCPMixtureVarSet.hpp:143-156 */
    vs->Vf = 0;
    for (int j = 0; j < np; j++) if (vs->phaseExist[j]) vs->Vf += vs->vj[j];
    for (int j = 0; j < np; j++) vs->S[j] = vs->phaseExist[j] ? vs->vj[j] / vs->Vf : 0;

    /* CalVfiVfp_full02 :462-524 + AddVfpVfiW :896-901 */
    if (m->NP == 1) {
        int j = m->ep[0];
        vs->vjP[j] = vs->nj[j] * m->vmP[j];
        vs->vfP = vs->vjP[j];
        for (int i = 0; i < NC; i++) {
            vs->vji[j][i] = m->vm[j] + m->vmx[j][i];
            for (int k = 0; k < NC; k++) vs->vji[j][i] -= m->vmx[j][k] * vs->x[j * nc + k];
            vs->vfi[i] = vs->vji[j][i];
        }
    } else {
        for (int q = 0; q < 2; q++) {
            int j = m->ep[q];
            eos_lnfugn_impl(e, vs->P, vs->T, &vs->x[j * nc], vs->nj[j], m->lnfugN[j], 0);
            eos_lnfugp(e, vs->P, vs->T, &vs->x[j * nc], m->lnfugP[j]);
        }
        for (int i = 0; i < NC; i++)
            for (int k = 0; k < NC; k++) m->Jmat[i * NC + k] = m->lnfugN[0][i * NC + k] + m->lnfugN[1][i * NC + k];
        for (int i = 0; i < NC; i++) m->rhs[i] = m->lnfugP[1][i] - m->lnfugP[0][i];
        for (int k = 0; k < NC; k++)
            for (int i = 0; i < NC; i++) m->rhs[(k + 1) * NC + i] = m->lnfugN[1][k * NC + i];
        gauss_solve(NC, m->Jmat, NC + 1, m->rhs);
        m->pe.nsolve++;
        double t0 = 0, t1 = 0;
        for (int i = 0; i < NC; i++) {
            t0 -= m->vmx[0][i] * vs->x[i];
            t1 -= m->vmx[1][i] * vs->x[nc + i];
        }
        const double* dP = m->rhs;
        vs->vjP[0] = vs->nj[0] * m->vmP[0];
        vs->vjP[1] = vs->nj[1] * m->vmP[1];
        for (int k = 0; k < NC; k++) {
            vs->vjP[0] += (m->vm[0] + t0 + m->vmx[0][k]) * dP[k];
            vs->vjP[1] -= (m->vm[1] + t1 + m->vmx[1][k]) * dP[k];
        }
        vs->vfP = vs->vjP[0] + vs->vjP[1];
        const double* dN = dP + NC;
        for (int i = 0; i < NC; i++) {
            vs->vji[0][i] = 0;
            vs->vji[1][i] = 0;
            for (int k = 0; k < NC; k++) {
                vs->vji[0][i] += (m->vm[0] + t0 + m->vmx[0][k]) * dN[k];
                vs->vji[1][i] += (m->vm[1] + t1 + m->vmx[1][k]) * ((i == k ? 1.0 : 0.0) - dN[k]);
            }
            vs->vfi[i] = vs->vji[0][i] + vs->vji[1][i];
            dN += NC;
        }
    }
    vs->vfP += vs->vjP[wP];
    vs->vfi[wC] = vs->vji[wP][wC];

    /* CaldXsdXp02 :565-642 + CaldXsdXpW :904-920 */
    const int ncol = nc + 1;
    memset(vs->dXsdXp, 0, sizeof(double) * np * ncol * ncol);
    for (int q = 0; q < m->nep; q++) {
        int j = m->ep[q];
        double* r = &vs->dXsdXp[j * ncol];
        r[0] = (vs->vjP[j] - vs->vfP * vs->S[j]) / vs->Vf;
        for (int mm = 0; mm < NC; mm++) r[1 + mm] = (vs->vji[j][mm] - vs->vfi[mm] * vs->S[j]) / vs->Vf;
    }
    if (m->NP == 1) {
        double* r = &vs->dXsdXp[(np + m->ep[0] * nc) * ncol + 1];
        for (int i = 0; i < NC; i++) {
            for (int mm = 0; mm < NC; mm++)
                r[mm] = ((i == mm ? 1.0 : 0.0) * m->Nt - vs->Ni[i]) / (m->Nt * m->Nt);
            r += ncol;
        }
    } else {
        double* b0 = &vs->dXsdXp[np * ncol];
        double* b1 = b0 + nc * ncol;
        const double* d = m->rhs;
        const double* x0 = &vs->x[0];
        const double* x1 = &vs->x[nc];
        double sum = 0;
        for (int i = 0; i < NC; i++) sum += d[i];
        for (int i = 0; i < NC; i++) {
            b0[i * ncol] = (d[i] - sum * x0[i]) / vs->nj[0];
            b1[i * ncol] = (-d[i] + sum * x1[i]) / vs->nj[1];
        }
        d += NC;
        for (int mm = 0; mm < NC; mm++) {
            sum = 0;
            for (int i = 0; i < NC; i++) sum += d[i];
            for (int i = 0; i < NC; i++) {
                b0[i * ncol + mm + 1] = (d[i] - sum * x0[i]) / vs->nj[0];
                b1[i * ncol + mm + 1] = (((i == mm) ? 1.0 : 0.0) - d[i] - (1 - sum) * x1[i]) / vs->nj[1];
            }
            d += NC;
        }
    }
    for (int q = 0; q < m->nep; q++) {
        int j = m->ep[q];
        vs->dXsdXp[j * ncol + wC + 1] = (vs->vji[j][wC] - vs->vfi[wC] * vs->S[j]) / vs->Vf;
    }
    {
        double* r = &vs->dXsdXp[wP * ncol];
        r[0] = (vs->vjP[wP] - vs->vfP * vs->S[wP]) / vs->Vf;
        for (int i = 0; i < nc; i++) r[1 + i] = (vs->vji[wP][i] - vs->vfi[i] * vs->S[wP]) / vs->Vf;
    }
}

/* ------------------------------------------------------------------------ */
/* SkipPSA: src/AcceleratePEC.cpp:64-172                                      */
/* ------------------------------------------------------------------------ */
/* Smallest eigenvalue of a symmetric float matrix (stands for ssyevd_ 'N',
 * src/DenseMat.cpp:126-145): cyclic Jacobi in double on the float entries,
 * result rounded to float. */
static float min_eig_sym(int n, const float* Af)
{
    double A[MC * MC];
    for (int i = 0; i < n * n; i++) A[i] = Af[i];
    for (int sweep = 0; sweep < 60; sweep++) {
        double off = 0;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) off += A[p * n + q] * A[p * n + q];
        if (off < 1e-30) break;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) {
                double apq = A[p * n + q];
                if (fabs(apq) < 1e-300) continue;
                double th = (A[q * n + q] - A[p * n + p]) / (2 * apq);
                double t = sgn(th) / (fabs(th) + sqrt(th * th + 1));
                if (th == 0) t = 1;
                double c = 1 / sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < n; k++) {
                    double akp = A[k * n + p], akq = A[k * n + q];
                    A[k * n + p] = c * akp - s * akq;
                    A[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {
                    double apk = A[p * n + k], aqk = A[q * n + k];
                    A[p * n + k] = c * apk - s * aqk;
                    A[q * n + k] = s * apk + c * aqk;
                }
            }
    }
    double mn = A[0];
    for (int i = 1; i < n; i++) if (A[i * n + i] < mn) mn = A[i * n + i];
    return (float)mn;
}

static int skip_if_skip(const orc_bulk* b, int64_t c, int NC, const mvs_t* vs)   /* :64-89 */
{
    if (!b->skip_flag[c]) return 0;
    double Nt = 0;
    for (int i = 0; i < NC; i++) Nt += vs->Ni[i];
    double me = b->skip_minEig[c];
    if (fabs(1 - b->skip_P[c] / vs->P) >= me / 10) return 0;
    if (fabs(b->skip_T[c] - vs->T) >= me * 10) return 0;
    for (int i = 0; i < NC; i++)
        if (fabs(vs->Ni[i] / Nt - b->skip_zi[c * NC + i]) >= me / 10) return 0;
    return 1;
}

static int skip_ftype02(const orc_bulk* b, int64_t c, int NC, const mvs_t* vs)  /* :103-124 */
{
    uint32_t npPE = vs->phaseNum - 1u;
    if (skip_if_skip(b, c, NC, vs)) return 1;
    if (npPE >= 2) {   /* unsigned: phaseNum = 0 wraps and falls through to 0 */
        uint32_t t = 0;
        for (int j = 0; j < 2; j++) if (vs->S[j] >= 1E-4) t++;
        return t == npPE ? (int)npPE : 0;
    }
    return 0;
}

static void skip_next(orc_bulk* b, int64_t c, mix_t* m, const mvs_t* vs)       /* :127-172 */
{
    int NC = m->NC, ft = m->pe.ftype;
    if (ft == 0) {
        double lnphiN[MC * MC];
        float M[MC * MC];
        eos_lnfugn_impl(&m->pe.e, vs->P, vs->T, m->zi, m->Nt, lnphiN, 1);
        for (int i = 0; i < NC; i++)
            for (int j = 0; j <= i; j++) {
                M[i * NC + j] = (float)((i == j ? 1.0 : 0.0) + m->Nt * sqrt(m->zi[i] * m->zi[j]) * lnphiN[i * NC + j]);
                M[j * NC + i] = M[i * NC + j];
            }
        b->skip_flag[c] = 1;
        b->skip_minEig[c] = (double)min_eig_sym(NC, M);
        b->skip_P[c] = vs->P;
        b->skip_T[c] = vs->T;
        memcpy(&b->skip_zi[c * NC], m->zi, sizeof(double) * NC);
    } else if (ft == 1) {
        b->skip_flag[c] = 1;
    } else {
        b->skip_flag[c] = 0;
    }
}

/* ------------------------------------------------------------------------ */
/* IsoT_FIM::CalFlash + PassFlashValue  src/IsoThermalMethod.cpp:1016-1071   */
/* ------------------------------------------------------------------------ */
void orc_cal_flash(const ocpx_fluid* fin, orc_bulk* b)
{
    fl_t f;
    fl_init(&f, fin);
    const int NC = f.NC, nc = NC + 1, np = NPH;
    const int lend = np * (nc + 1) * (nc + 1);
    mix_t* m = (mix_t*)malloc(sizeof(mix_t));
    mvs_t* vs = (mvs_t*)malloc(sizeof(mvs_t));
    for (int64_t c = 0; c < b->nb; c++) {
        memset(m, 0, sizeof(*m));
        pe_init(&m->pe, &f);
        m->NC = NC;
        memset(vs, 0, sizeof(*vs));
        vs->np = np; vs->nc = nc;
        /* SetVarSet :671-687 */
        vs->P = b->P[c];
        vs->T = b->T[c] + CONV4;
        memcpy(vs->Ni, &b->Ni[c * nc], sizeof(double) * nc);
        memcpy(vs->S, &b->S[c * np], sizeof(double) * np);
        memcpy(vs->phaseExist, &b->phaseExist[c * np], sizeof(uint32_t) * np);
        memcpy(vs->x, &b->xij[c * np * nc], sizeof(double) * np * nc);
        vs->phaseNum = 0;
        for (int j = 0; j < np; j++) if (vs->phaseExist[j]) vs->phaseNum++;
        int ftype;
        if (b->ftype_in && b->ftype_in[c] >= 0) ftype = b->ftype_in[c];
        else ftype = fin->skipPSA ? skip_ftype02(b, c, NC, vs) : 0;
        mix_flash_der(m, vs, ftype);
        if (fin->skipPSA) skip_next(b, c, m, vs);
        if (b->ftype_out) b->ftype_out[c] = m->pe.ftype;
        if (b->iters) pe_counters(&m->pe, b->iters + 8 * c);
        if (b->pe_ill) b->pe_ill[c] = m->pe.ill;
        /* PassFlashValue :1027-1071 */
        b->Nt[c] = vs->Nt;
        b->vf[c] = vs->Vf;
        for (int j = 0; j < np; j++) {
            b->S[c * np + j] = vs->S[j];
            b->phaseExist[c * np + j] = vs->phaseExist[j];
            if (vs->phaseExist[j]) {
                b->rho[c * np + j] = vs->rho[j];
                b->xi[c * np + j] = vs->xi[j];
                b->mu[c * np + j] = vs->mu[j];
                b->rhoP[c * np + j] = vs->rhoP[j];
                b->xiP[c * np + j] = vs->xiP[j];
                b->muP[c * np + j] = vs->muP[j];
                for (int i = 0; i < nc; i++) {
                    b->xij[(c * np + j) * nc + i] = vs->x[j * nc + i];
                    b->rhox[(c * np + j) * nc + i] = vs->rhox[j * nc + i];
                    b->xix[(c * np + j) * nc + i] = vs->xix[j * nc + i];
                    b->mux[(c * np + j) * nc + i] = vs->mux[j * nc + i];
                }
            }
        }
        b->vfP[c] = vs->vfP;
        for (int i = 0; i < nc; i++) b->vfi[c * nc + i] = vs->vfi[i];
        memcpy(&b->dSec_dPri[c * lend], vs->dXsdXp, sizeof(double) * lend);
    }
    free(vs);
    free(m);
}

/* ------------------------------------------------------------------------ */
/* kr / Pc: OCPFlowMethod_OGW01::CalKrPcDer src/OCPFlowMethod.cpp:63-73,       */
/* OCPTable::Eval_All src/OCPTable.cpp:26-63, Stone II src/OCPFuncSAT.cpp:162  */
/* ------------------------------------------------------------------------ */
static void table_eval(const double* tab, int nrow, double v, double* out, double* slope)
{
    /* tab row-major [nrow][4]; bracket data[k] <= v < data[k+1]; clamp with
     * zero slope outside (the reference's row hint does not change the
     * bracket of a strictly increasing column). */
    if (v >= tab[0]) {
        for (int i = 1; i < nrow; i++) {
            if (v < tab[i * 4]) {
                int k = i - 1;
                for (int c = 0; c < 4; c++) {
                    slope[c] = (tab[(k + 1) * 4 + c] - tab[k * 4 + c]) / (tab[(k + 1) * 4] - tab[k * 4]);
                    out[c] = tab[k * 4 + c] + slope[c] * (v - tab[k * 4]);
                }
                return;
            }
        }
        for (int c = 0; c < 4; c++) { slope[c] = 0; out[c] = tab[(nrow - 1) * 4 + c]; }
    } else {
        for (int c = 0; c < 4; c++) { slope[c] = 0; out[c] = tab[c]; }
    }
}

void orc_kr_pc(const ocpx_sat* s, const double* S, double* kr, double* pc, double* dkrds, double* dpcds)
{
    const int o = 0, g = 1, w = 2, np = 3;
    double d[4], sl[4];
    for (int k = 0; k < np * np; k++) { dkrds[k] = 0; dpcds[k] = 0; }
    pc[o] = 0;
    table_eval(s->swof, s->nswof, S[w], d, sl);
    kr[w] = d[1];
    double krow = d[2];
    pc[w] = -d[3];
    dkrds[w * np + w] = sl[1];
    double dkrowdSw = sl[2];
    dpcds[w * np + w] = -sl[3];
    table_eval(s->sgof, s->nsgof, S[g], d, sl);
    kr[g] = d[1];
    double krog = d[2];
    pc[g] = d[3];
    dkrds[g * np + g] = sl[1];
    double dkrogdSg = sl[2];
    dpcds[g * np + g] = sl[3];
    double krocw = s->swof[2];
    const double dKrowdSo = 0, dKrogdSo = 0;     /* stay 0 in OGW01 */
    kr[o] = krocw * ((krow / krocw + kr[w]) * (krog / krocw + kr[g]) - (kr[w] + kr[g]));
    if (kr[o] < 0) {
        kr[o] = 0;
        dkrds[o * np + o] = 0; dkrds[o * np + g] = 0; dkrds[o * np + w] = 0;
    } else {
        dkrds[o * np + o] = dKrowdSo * (krog / krocw + kr[g]) + dKrogdSo * (krow / krocw + kr[w]);
        dkrds[o * np + w] = krocw * ((dkrowdSw / krocw + dkrds[w * np + w]) * (krog / krocw + kr[g]) - dkrds[w * np + w]);
        dkrds[o * np + g] = krocw * ((krow / krocw + kr[w]) * (dkrogdSg / krocw + dkrds[g * np + g]) - dkrds[g * np + g]);
    }
}

void orc_cal_krpc(const ocpx_sat* s, orc_bulk* b)            /* IsoThermalMethod.cpp:1073-1088 */
{
    const int np = 3;
    for (int64_t n = 0; n < b->nb; n++) {
        orc_kr_pc(s, &b->S[n * np], &b->kr[n * np], &b->Pc[n * np], &b->dKrdS[n * np * np], &b->dPcdS[n * np * np]);
        for (int j = 0; j < np; j++) b->Pj[n * np + j] = b->P[n] + b->Pc[n * np + j];
    }
}

void orc_cal_rock(const ocpx_rock* r, orc_bulk* b)            /* :19-30, OCPRock.cpp:19-24 */
{
    for (int64_t n = 0; n < b->nb; n++) {
        double dP = b->P[n] - r->Pref;
        b->poro[n] = b->poroInit[n] * (1 + (r->cp1 + r->cp2 / 2 * dP) * dP);
        b->poroP[n] = b->poroInit[n] * (r->cp1 + r->cp2 * dP);
        b->rockVp[n] = b->v[n] * b->poro[n];
    }
}

/* ------------------------------------------------------------------------ */
/* Flux: OCPConvection01::CalFlux src/OCPConvection.cpp:22-79                 */
/* ------------------------------------------------------------------------ */
static void conv_flux(const orc_bulk* b, int nc, int64_t bId, int64_t eId, double Akd,
                      int64_t* up, double* dPj, double* flux)
{
    const int np = 3;
    for (int i = 0; i < nc; i++) flux[i] = 0;
    for (int j = 0; j < np; j++) {
        int64_t bj = bId * np + j, ej = eId * np + j;
        int exb = b->phaseExist[bj] != 0, exe = b->phaseExist[ej] != 0;
        double rho;
        if (exb && exe) rho = (b->rho[bj] + b->rho[ej]) / 2;
        else if (exb) rho = b->rho[bj];
        else if (exe) rho = b->rho[ej];
        else { up[j] = bId; dPj[j] = 0; continue; }
        up[j] = bId;
        dPj[j] = (b->Pj[bj] - GRAV * rho * b->depth[bId]) - (b->Pj[ej] - GRAV * rho * b->depth[eId]);
        if (dPj[j] < 0) up[j] = eId;
        int64_t uj = up[j] * np + j;
        if (!b->phaseExist[uj]) continue;
        double vj = Akd * b->kr[uj] / b->mu[uj] * dPj[j];
        for (int i = 0; i < nc; i++) flux[i] += vj * b->xi[uj] * b->xij[uj * nc + i];
    }
}

/* IsoT_FIM::CalRes src/IsoThermalMethod.cpp:1090-1185 (no wells) */
void orc_cal_res(const ocpx_fluid* fin, const orc_bulk* b, int64_t nconn, const int32_t* cb,
                 const int32_t* ce, const double* trans, double dt, double* res, double* norms)
{
    const int nc = fin->nc + 1, len = nc + 1;
    const int64_t nb = b->nb;
    for (int64_t n = 0; n < nb; n++) {
        res[n * len] = b->rockVp[n] - b->vf[n];
        for (int i = 0; i < nc; i++) res[n * len + 1 + i] = b->Ni[n * nc + i] - b->lNi[n * nc + i];
    }
    double flux[MB], dPj[3];
    int64_t up[3];
    for (int64_t c = 0; c < nconn; c++) {
        int64_t bId = cb[c], eId = ce[c];
        conv_flux(b, nc, bId, eId, trans[c], up, dPj, flux);
        for (int i = 0; i < nc; i++) {
            res[bId * len + 1 + i] += dt * flux[i];
            res[eId * len + 1 + i] -= dt * flux[i];
        }
    }
    double mV = 0, mN = 0;
    int64_t iV = 0, iN = 0;
    for (int64_t n = 0; n < nb; n++) {
        for (int i = 0; i < len; i++) {
            double t = fabs(res[n * len + i] / b->rockVp[n]);
            if (mV < t) { mV = t; iV = n; }
        }
        for (int i = 1; i < len; i++) {
            double t = fabs(res[n * len + i] / b->Nt[n]);
            if (mN < t) { mN = t; iN = n; }
        }
    }
    for (int64_t k = 0; k < nb * len; k++) res[k] = -res[k];
    if (norms) { norms[0] = mV; norms[1] = (double)iV; norms[2] = mN; norms[3] = (double)iN; }
}

/* OCPConvection01::AssembleMatFIM src/OCPConvection.cpp:82-187 */
static void conv_jac(const orc_bulk* b, int nc, int64_t bId, int64_t eId, double Akd,
                     const int64_t* upb, const double* dPc, double* dFdXpB, double* dFdXpE,
                     double* dFdXsB, double* dFdXsE)
{
    const int np = 3, n1 = nc + 1, n2 = (nc + 1) * np;
    memset(dFdXpB, 0, sizeof(double) * n1 * n1);
    memset(dFdXpE, 0, sizeof(double) * n1 * n1);
    memset(dFdXsB, 0, sizeof(double) * n1 * n2);
    memset(dFdXsE, 0, sizeof(double) * n1 * n2);
    const double dGamma = GRAV * (b->depth[bId] - b->depth[eId]);
    for (int j = 0; j < np; j++) {
        int64_t uj = upb[j] * np + j;
        if (!b->phaseExist[uj]) continue;
        int64_t bj = bId * np + j, ej = eId * np + j, dj;
        double *pU, *pD, *sU, *sD;
        if (bj == uj) { dj = ej; pU = dFdXpB; pD = dFdXpE; sU = dFdXsB; sD = dFdXsE; }
        else { dj = bj; pU = dFdXpE; pD = dFdXpB; sU = dFdXsE; sD = dFdXsB; }
        double wU, wD;
        if (b->phaseExist[dj]) { wU = 0.5; wD = 0.5; } else { wU = 1; wD = 0; }
        double dP = dPc[j], xi = b->xi[uj], kr = b->kr[uj], mu = b->mu[uj];
        double muP = b->muP[uj], xiP = b->xiP[uj];
        double tJ = Akd * kr / mu;
        for (int i = 0; i < nc; i++) {
            double xij = b->xij[uj * nc + i];
            double tIJ = xij * xi * tJ;
            dFdXpB[(i + 1) * n1] += tIJ;
            dFdXpE[(i + 1) * n1] -= tIJ;
            double t = tJ * xiP * xij * dP;
            t += -tIJ * muP / mu * dP;
            pU[(i + 1) * n1] += (t - tIJ * wU * b->rhoP[uj] * dGamma);
            pD[(i + 1) * n1] += -tIJ * wD * b->rhoP[dj] * dGamma;
            for (int k = 0; k < np; k++) {
                dFdXsB[(i + 1) * n2 + k] += tIJ * b->dPcdS[bj * np + k];
                dFdXsE[(i + 1) * n2 + k] -= tIJ * b->dPcdS[ej * np + k];
                sU[(i + 1) * n2 + k] += Akd * b->dKrdS[uj * np + k] / mu * xi * xij * dP;
            }
            for (int k = 0; k < nc; k++) {
                double rhox = b->rhox[uj * nc + k], xix = b->xix[uj * nc + k], mux = b->mux[uj * nc + k];
                t = -tIJ * wU * rhox * dGamma;
                t += tJ * xix * xij * dP;
                t += -tIJ * mux / mu * dP;
                sU[(i + 1) * n2 + np + j * nc + k] += t;
                sD[(i + 1) * n2 + np + j * nc + k] += -tIJ * wD * b->rhox[dj * nc + k] * dGamma;
            }
            sU[(i + 1) * n2 + np + j * nc + i] += tJ * xi * dP;
        }
    }
}

/* C = dt * A B + dt * C (DaABpbC, src/DenseMat.cpp:61-85) */
static void gemm_dt(int m, int n, int k, double dt, const double* A, const double* B, double* C)
{
    for (int i = 0; i < m; i++)
        for (int j = 0; j < n; j++) {
            double s = 0;
            for (int q = 0; q < k; q++) s += A[i * k + q] * B[q * n + j];
            C[i * n + j] = dt * s + dt * C[i * n + j];
        }
}

/* IsoT_FIM::AssembleMatBulks src/IsoThermalMethod.cpp:1187-1269 + the
 * OCPMatrix push semantics src/OCPMatrix.cpp:92-137 (rows: diagonal block
 * first, off-diagonal blocks in push order). */
void orc_assemble(const ocpx_fluid* fin, const orc_bulk* b, int64_t nconn, const int32_t* cb,
                  const int32_t* ce, const double* trans, double dt, int64_t* rowptr,
                  int32_t* colind, double* val)
{
    const int nc = fin->nc + 1, np = 3, n1 = nc + 1, n2 = np * nc + np, bs = n1 * n1;
    const int64_t nb = b->nb;
    const int lend = n1 * n2;
    /* count blocks per row: 1 + number of connections touching the row */
    int64_t* cnt = (int64_t*)calloc((size_t)nb, sizeof(int64_t));
    for (int64_t c = 0; c < nconn; c++) { cnt[cb[c]]++; cnt[ce[c]]++; }
    rowptr[0] = 0;
    for (int64_t n = 0; n < nb; n++) rowptr[n + 1] = rowptr[n] + 1 + cnt[n];
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)nb);
    /* NewDiag: CaldFdXpFIM src/BulkAccumuModule.cpp:47-70 */
    for (int64_t n = 0; n < nb; n++) {
        double* d = &val[rowptr[n] * bs];
        memset(d, 0, sizeof(double) * bs);
        d[0] = b->v[n] * b->poroP[n] - b->vfP[n];
        for (int i = 0; i < nc; i++) d[i + 1] = -b->vfi[n * nc + i];
        for (int i = 1; i < n1; i++) d[i * n1 + i] = 1;
        colind[rowptr[n]] = (int32_t)n;
        fill[n] = rowptr[n] + 1;
    }
    double *pB = malloc(sizeof(double) * bs), *pE = malloc(sizeof(double) * bs);
    double *sB = malloc(sizeof(double) * n1 * n2), *sE = malloc(sizeof(double) * n1 * n2);
    double flux[MB], dPj[3];
    int64_t up[3];
    for (int64_t c = 0; c < nconn; c++) {
        int64_t bId = cb[c], eId = ce[c];
        conv_flux(b, nc, bId, eId, trans[c], up, dPj, flux);      /* upblock/dP from CalRes */
        conv_jac(b, nc, bId, eId, trans[c], up, dPj, pB, pE, sB, sE);
        gemm_dt(n1, n1, n2, dt, sB, &b->dSec_dPri[bId * lend], pB);
        double* db = &val[rowptr[bId] * bs];
        for (int k = 0; k < bs; k++) db[k] += pB[k];              /* AddDiag(bId)     */
        {
            int64_t s = fill[eId]++;                               /* NewOffDiag(e, b) */
            colind[s] = (int32_t)bId;
            for (int k = 0; k < bs; k++) val[s * bs + k] = -pB[k];
        }
        gemm_dt(n1, n1, n2, dt, sE, &b->dSec_dPri[eId * lend], pE);
        {
            int64_t s = fill[bId]++;                               /* NewOffDiag(b, e) */
            colind[s] = (int32_t)eId;
            for (int k = 0; k < bs; k++) val[s * bs + k] = pE[k];
        }
        double* de = &val[rowptr[eId] * bs];
        for (int k = 0; k < bs; k++) de[k] += -pE[k];             /* AddDiag(eId)     */
    }
    free(pB); free(pE); free(sB); free(sE);
    free(fill); free(cnt);
}

/* ---- Peaceman wells ---------------------------------------------------
 * Per-iteration well terms of the FIM step: CalTrans (src/WellPeaceman.cpp:
 * 365-406), CalFlux (:409-462), CalResFIM (:1226-1265) and
 * AssembleMatInjFIM / AssembleMatProdFIM (:1268-1360, :1488-1612), on top of
 * the bulk residual of orc_cal_res (res: -R, updated to -(R + well terms),
 * norms recomputed over the bulk rows as :1147-1168) and, when rowptr is
 * given, the bulk BSR of orc_assemble (diagonal blocks updated in place).
 * Outputs: wres [nw] well-row residual (before the sign flip), qw [nw*nc]
 * the well's component rates; Bnw [nperf*len] column 0 of block (n, w),
 * Cwn [nperf*len] row 0 of block (w, n), Dw [nw*len*len] the well's
 * diagonal block (every other entry of those blocks is zero). */
void orc_wells(const ocpx_fluid* fin, const orc_bulk* b, double dt, int32_t nw, const ocpx_well* w,
               const int64_t* pcell, const double* pWI, const double* pdG, const double* pxi,
               double* res, double* norms, double* wres, double* qw,
               const int64_t* rowptr, double* val, double* Bnw, double* Cwn, double* Dw)
{
    const int nc = fin->nc + 1, np = 3, n1 = nc + 1, n2 = np * nc + np, bs = n1 * n1;
    const int64_t nb = b->nb;
    double *pB = malloc(sizeof(double) * bs), *pW = malloc(sizeof(double) * bs);
    double *sB = malloc(sizeof(double) * n1 * n2), *bm = malloc(sizeof(double) * bs);
    for (int32_t iw = 0; iw < nw; iw++) {
        const ocpx_well* W = &w[iw];
        const int inj = W->type == OCPX_WELL_INJ;
        double qwell[MB];
        for (int i = 0; i < nc; i++) qwell[i] = 0;
        double* D = Dw ? &Dw[(size_t)iw * bs] : NULL;
        if (D) memset(D, 0, sizeof(double) * bs);
        for (int32_t p = W->perf0; p < W->perf0 + W->nperf; p++) {
            const int64_t n = pcell[p];
            const double WI = pWI[p], perfP = W->bhp + pdG[p];          /* CalPerfP */
            /* CalTrans */
            double tj[3], tINJ = 0;
            for (int j = 0; j < np; j++) {
                const int64_t id = n * np + j;
                tj[j] = 0;
                if (b->phaseExist[id]) { tj[j] = WI * b->kr[id] / b->mu[id]; tINJ += tj[j]; }
            }
            /* CalFlux */
            double qi[MB];
            for (int i = 0; i < nc; i++) qi[i] = 0;
            if (inj) {
                const double qt = tINJ * (b->P[n] - perfP);
                for (int i = 0; i < nc; i++) qi[i] = qt * pxi[p] * W->zi[i];
            } else {
                for (int j = 0; j < np; j++) {
                    const int64_t id = n * np + j;
                    if (!b->phaseExist[id]) continue;
                    const double qj = tj[j] * (b->Pj[id] - perfP);
                    for (int i = 0; i < nc; i++) qi[i] += qj * b->xi[id] * b->xij[id * nc + i];
                }
            }
            /* CalResFIM: well to bulk (res holds -R) */
            for (int i = 0; i < nc; i++) {
                qwell[i] += qi[i];
                res[n * n1 + 1 + i] = -(-res[n * n1 + 1 + i] + qi[i] * dt);
            }
            if (!rowptr) continue;
            /* AssembleMat{Inj,Prod}FIM */
            memset(pB, 0, sizeof(double) * bs);
            memset(pW, 0, sizeof(double) * bs);
            memset(sB, 0, sizeof(double) * n1 * n2);
            for (int j = 0; j < np; j++) {
                const int64_t id = n * np + j;
                if (!b->phaseExist[id]) continue;
                const double mu = b->mu[id], muP = b->muP[id];
                if (inj) {
                    const double dP = b->P[n] - W->bhp - pdG[p];
                    for (int i = 0; i < nc; i++) {
                        const double tIJ = tj[j] * pxi[p] * W->zi[i];
                        pB[(i + 1) * n1] += tIJ * (1 - dP * muP / mu);
                        pW[(i + 1) * n1] += -tIJ;
                        for (int k = 0; k < np; k++)
                            sB[(i + 1) * n2 + k] += WI * pxi[p] * W->zi[i] * b->dKrdS[id * np + k] * dP / mu;
                        for (int k = 0; k < nc; k++)
                            sB[(i + 1) * n2 + np + j * nc + k] += -tIJ * dP / mu * b->mux[id * nc + k];
                    }
                } else {
                    const double dP = b->Pj[id] - W->bhp - pdG[p];
                    const double xi = b->xi[id], xiP = b->xiP[id];
                    for (int i = 0; i < nc; i++) {
                        const double xij = b->xij[id * nc + i];
                        const double tIJ = tj[j] * xi * xij;
                        pB[(i + 1) * n1] += tIJ * (1 - dP * muP / mu) + dP * tj[j] * xij * xiP;
                        pW[(i + 1) * n1] += -tIJ;
                        for (int k = 0; k < np; k++) {
                            double t = WI * dP / mu * xi * xij * b->dKrdS[id * np + k];
                            t += tIJ * b->dPcdS[id * np + k];
                            sB[(i + 1) * n2 + k] += t;
                        }
                        for (int k = 0; k < nc; k++)
                            sB[(i + 1) * n2 + np + j * nc + k] +=
                                dP * tj[j] * xij * (b->xix[id * nc + k] - xi / mu * b->mux[id * nc + k]);
                        sB[(i + 1) * n2 + np + j * nc + i] += tj[j] * xi * dP;
                    }
                }
            }
            /* bulk - bulk: AddDiag(n, dt*(dQdXpB + dQdXsB*dXsdXp)) */
            memcpy(bm, pB, sizeof(double) * bs);
            gemm_dt(n1, n1, n2, dt, sB, &b->dSec_dPri[n * n1 * n2], bm);
            double* d = &val[rowptr[n] * bs];
            for (int k = 0; k < bs; k++) d[k] += bm[k];
            /* bulk - well: dt * dQdXpW (column 0) */
            for (int r = 0; r < n1; r++) Bnw[(size_t)p * n1 + r] = pW[r * n1] * dt;
            if (W->mode == OCPX_WELL_RATE) {
                double b00 = 0;
                for (int i = 0; i < nc; i++) b00 += pW[(i + 1) * n1] * W->factor[i];
                D[0] += b00;
                for (int i = 0; i < nc; i++) D[(i + 1) * n1 + i + 1] += 1;
                memcpy(bm, pB, sizeof(double) * bs);
                gemm_dt(n1, n1, n2, 1.0, sB, &b->dSec_dPri[n * n1 * n2], bm);
                double* c = &Cwn[(size_t)p * n1];
                for (int k = 0; k < n1; k++) c[k] = 0;
                for (int i = 0; i < nc; i++)
                    for (int k = 0; k < n1; k++) c[k] += W->factor[i] * bm[(i + 1) * n1 + k];
            } else {
                for (int r = 0; r < n1; r++) D[r * n1 + r] += 1;
                for (int k = 0; k < n1; k++) Cwn[(size_t)p * n1 + k] = 0;
            }
        }
        double r0;
        if (W->mode == OCPX_WELL_RATE) {
            r0 = inj ? W->target : -W->target;
            for (int i = 0; i < nc; i++) r0 += qwell[i] * W->factor[i];
        } else {
            r0 = W->bhp - W->target;
        }
        if (wres) wres[iw] = r0;
        if (qw) for (int i = 0; i < nc; i++) qw[(size_t)iw * nc + i] = qwell[i];
    }
    free(pB); free(pW); free(sB); free(bm);
    if (norms) {
        double mV = 0, mN = 0;
        int64_t iV = 0, iN = 0;
        for (int64_t n = 0; n < nb; n++) {
            for (int i = 0; i < n1; i++) {
                double t = fabs(res[n * n1 + i] / b->rockVp[n]);
                if (mV < t) { mV = t; iV = n; }
            }
            for (int i = 1; i < n1; i++) {
                double t = fabs(res[n * n1 + i] / b->Nt[n]);
                if (mN < t) { mN = t; iN = n; }
            }
        }
        norms[0] = mV; norms[1] = (double)iV; norms[2] = mN; norms[3] = (double)iN;
    }
}

/* Peaceman well index of a perforation, CalWI (src/WellPeaceman.cpp:277-362)
 * for the x, y, z directions (0, 1, 2): dz is already multiplied by NTG;
 * kh < 0 selects the geometric kh; the result includes CONV_DARCY. */
double orc_peaceman_wi(int dir, double dx, double dy, double dz, double kx, double ky, double kz,
                       double radius, double skin, double kh)
{
    double ro = 0;
    if (dir == 0) {
        const double kykz = ky * kz, ky_kz = ky / kz;
        ro = 0.28 * pow((dy * dy * pow(1 / ky_kz, 0.5) + dz * dz * pow(ky_kz, 0.5)), 0.5);
        ro /= (pow(ky_kz, 0.25) + pow(1 / ky_kz, 0.25));
        if (kh < 0) kh = (dx * pow(kykz, 0.5));
    } else if (dir == 1) {
        const double kzkx = kz * kx, kz_kx = kz / kx;
        ro = 0.28 * pow((dz * dz * pow(1 / kz_kx, 0.5) + dx * dx * pow(kz_kx, 0.5)), 0.5);
        ro /= (pow(kz_kx, 0.25) + pow(1 / kz_kx, 0.25));
        if (kh < 0) kh = (dy * pow(kzkx, 0.5));
    } else {
        const double kxky = kx * ky, kx_ky = kx / ky;
        ro = 0.28 * pow((dx * dx * pow(1 / kx_ky, 0.5) + dy * dy * pow(kx_ky, 0.5)), 0.5);
        ro /= (pow(kx_ky, 0.25) + pow(1 / kx_ky, 0.25));
        if (kh < 0) kh = (dz * pow(kxky, 0.5));
    }
    double WI = (2 * PI_REF) * kh / (log(ro / radius) + skin);
    return WI * DARCY;
}

/* IsoT_FIM::GetSolution src/IsoThermalMethod.cpp:1330-1400 (bulk part) */
void orc_get_solution(const ocpx_fluid* fin, orc_bulk* b, const double* u, double dSmax)
{
    const int nc = fin->nc + 1, np = 3, col = nc + 1, row = np * (nc + 1);
    double dt[NPH * (MB + 1)];
    for (int64_t n = 0; n < b->nb; n++) {
        const double* A = &b->dSec_dPri[n * row * col];
        const double* x = &u[n * col];
        for (int i = 0; i < row; i++) {
            dt[i] = 0;
            for (int j = 0; j < col; j++) dt[i] += 1.0 * A[i * col + j] * x[j];
        }
        double chop = 1;
        for (int j = 0; j < np; j++) {
            double ct = 1;
            if (fabs(dt[j]) > dSmax) ct = dSmax / fabs(dt[j]);
            else if (b->S[n * np + j] + dt[j] < 0.0) ct = 0.9 * b->S[n * np + j] / fabs(dt[j]);
            if (ct < chop) chop = ct;
        }
        for (int j = 0; j < np; j++) b->S[n * np + j] += chop * dt[j];
        int js = np;
        for (int j = 0; j < np; j++)
            for (int i = 0; i < nc; i++) b->xij[(n * np + j) * nc + i] += chop * dt[js++];
        b->P[n] += x[0];
        for (int i = 0; i < nc; i++) b->Ni[n * nc + i] += chop * x[1 + i];
    }
}

/* ------------------------------------------------------------------------ */
/* Cartesian connections: src/PreParamGridWell.cpp:1140-1214 (order, area)   */
/* and src/BulkConnFunc.cpp:15-67 (transmissibility)                          */
/* ------------------------------------------------------------------------ */
int64_t orc_cartesian_conns(int nx, int ny, int nz, const double* dx, const double* dy,
                            const double* dz, const double* kx, const double* ky,
                            const double* kz, const double* ntg, int32_t* cb, int32_t* ce,
                            double* trans)
{
    int64_t nxny = (int64_t)nx * ny, c = 0;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                int64_t b = k * nxny + (int64_t)j * nx + i;
                if (i < nx - 1) {
                    int64_t e = b + 1;
                    double aB = 2 * dy[b] * dz[b] / dx[b], aE = 2 * dy[e] * dz[e] / dx[e];
                    double T1 = ntg[b] * kx[b] * aB, T2 = ntg[e] * kx[e] * aE;
                    cb[c] = (int32_t)b; ce[c] = (int32_t)e;
                    trans[c++] = 1.0 / (1 / T1 + 1 / T2) * DARCY;
                }
                if (j < ny - 1) {
                    int64_t e = b + nx;
                    double aB = 2 * dz[b] * dx[b] / dy[b], aE = 2 * dz[e] * dx[e] / dy[e];
                    double T1 = ntg[b] * ky[b] * aB, T2 = ntg[e] * ky[e] * aE;
                    cb[c] = (int32_t)b; ce[c] = (int32_t)e;
                    trans[c++] = 1.0 / (1 / T1 + 1 / T2) * DARCY;
                }
                if (k < nz - 1) {
                    int64_t e = b + nxny;
                    double aB = 2 * dx[b] * dy[b] / dz[b], aE = 2 * dx[e] * dy[e] / dz[e];
                    double T1 = kz[b] * aB, T2 = kz[e] * aE;
                    cb[c] = (int32_t)b; ce[c] = (int32_t)e;
                    trans[c++] = 1.0 / (1 / T1 + 1 / T2) * DARCY;
                }
            }
    return c;
}

/* ------------------------------------------------------------------------ */
/* Single-composition EoS probes for tests                                    */
/* ------------------------------------------------------------------------ */
void orc_eos_fug(const ocpx_fluid* fin, double P, double T, const double* x, double* fug,
                 double* phi, double* ABZ)
{
    fl_t f; eos_t e;
    fl_init(&f, fin);
    memset(&e, 0, sizeof(e)); e.f = &f;
    eos_fugphi(&e, P, T, x, fug, phi);
    if (ABZ) { ABZ[0] = e.Aj; ABZ[1] = e.Bj; ABZ[2] = e.Zj; }
}
void orc_eos_lnfugx(const ocpx_fluid* fin, double P, double T, const double* x, double* out)
{
    fl_t f; eos_t e;
    fl_init(&f, fin);
    memset(&e, 0, sizeof(e)); e.f = &f;
    eos_lnfugx(&e, P, T, x, out);
}
void orc_eos_lnfugn(const ocpx_fluid* fin, double P, double T, const double* x, double nt, double* out)
{
    fl_t f; eos_t e;
    fl_init(&f, fin);
    memset(&e, 0, sizeof(e)); e.f = &f;
    eos_lnfugn_impl(&e, P, T, x, nt, out, 0);
}
void orc_eos_lnphin(const ocpx_fluid* fin, double P, double T, const double* x, double nt, double* out)
{
    fl_t f; eos_t e;
    fl_init(&f, fin);
    memset(&e, 0, sizeof(e)); e.f = &f;
    eos_lnfugn_impl(&e, P, T, x, nt, out, 1);
}
void orc_eos_lnfugp(const ocpx_fluid* fin, double P, double T, const double* x, double* out)
{
    fl_t f; eos_t e;
    fl_init(&f, fin);
    memset(&e, 0, sizeof(e)); e.f = &f;
    eos_lnfugp(&e, P, T, x, out);
}
double orc_eos_vmder(const ocpx_fluid* fin, double P, double T, const double* x, double* vmP, double* vmx)
{
    fl_t f; eos_t e;
    fl_init(&f, fin);
    memset(&e, 0, sizeof(e)); e.f = &f;
    return eos_vmder(&e, P, T, x, vmP, vmx);
}
