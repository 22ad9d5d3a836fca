// ocpx_dev.hpp — device-side data and the Peng-Robinson evaluator for gfx950.
//
// All per-cell working arrays are NC-sized register arrays indexed only
// with compile-time indices (every loop over components is unrolled), so
// nothing is indexed dynamically and nothing needs private memory for
// addressing.  Fluid constants live in one DevFluid block in global memory;
// every lane reads the same address, so the loads are uniform.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/ocpx.h"

#define OCPX_NCMAX OCPX_MAX_NC
#define UNROLL _Pragma("unroll")
// per-cell device functions are also host-callable, so the exact kernel
// logic can be stepped on the CPU
#define OCPX_HD __host__ __device__ __forceinline__

namespace ocpx {

// METRIC constants, include/OCPUnits.hpp:78-100 and include/OCPConst.hpp:30-32
constexpr double R_GAS = 0.083143;
constexpr double GRAV = 0.0000980665;
constexpr double CONV4 = 273.15;
constexpr double PI_REF = 3.141592653;
constexpr double D1 = 2.41421356237;   // include/OCPEoS.hpp:111
constexpr double D2 = -0.41421356237;  // include/OCPEoS.hpp:112

struct DevFluid {
    int32_t NC, lbc_exact, skip;
    int32_t RR_maxIt, SSMsta_maxIt, NRsta_maxIt, SSMsp_maxIt, NRsp_maxIt;
    double RR_tol, SSMsta_tol2, eYt, NRsta_tol, SSMsp_tol2, NRsp_tol;
    double Tc[OCPX_NCMAX], Pc[OCPX_NCMAX], logPc[OCPX_NCMAX], MW[OCPX_NCMAX];
    double Vc[OCPX_NCMAX], Vcvis[OCPX_NCMAX], OA[OCPX_NCMAX], OB[OCPX_NCMAX];
    double vsh[OCPX_NCMAX];      // Vshift * R*OmegaB*Tc/Pc
    double mw[OCPX_NCMAX];       // m(omega), src/OCPEoS.cpp:72-77
    double kwA[OCPX_NCMAX];      // 5.373*(1+omega), src/OCPPhaseEquilibrium.cpp:247
    double omk[OCPX_NCMAX * OCPX_NCMAX];   // 1 - k_ij, stride OCPX_NCMAX
    double lbc[5];               // LBC coefficients x10
    double auxB[OCPX_NCMAX], sqrtMW[OCPX_NCMAX];
    double pvtw[5], rhoW;
    double bT[OCPX_NCMAX];       // OmegaB Tc / Pc: Bi = bT * P / T
};

struct DevSat {
    int32_t nswof, nsgof;
    double swof[4 * OCPX_MAX_TABROWS];
    double sgof[4 * OCPX_MAX_TABROWS];
};

// finite test usable on host and device (v - v is NaN for inf and NaN)
OCPX_HD bool finite_d(double v) { return (v - v) == 0.0; }

// ---------------------------------------------------------------------------
// Cubic Z^3 + a Z^2 + b Z + c: trigonometric (3 roots) or Cardano (1 root)
// followed by Newton polishing to |e| <= 1e-8 with at most 10 steps,
// keeping the best iterate (src/UtilMath.cpp:15-80).
// ---------------------------------------------------------------------------
OCPX_HD double polish(double r, double a, double b, double c)
{
    double e = r * (r * (r + a) + b) + c;
    double best = r, beste = fabs(e);
    for (int it = 1; fabs(e) > 1E-8; ++it) {
        double df = r * (3 * r + 2 * a) + b;
        r = r - e / df;
        if (it > 10) break;
        e = r * (r * (r + a) + b) + c;
        if (fabs(e) <= beste) { beste = fabs(e); best = r; }
    }
    return best;
}

// Root selected by the reference (src/OCPEoS.cpp:102-124): the single real
// root, or between the smallest and largest of three the one with lower
// Gibbs energy.
OCPX_HD double pr_z(double A, double B)
{
    const double a = (D1 + D2 - 1) * B - 1;
    const double b = A + D1 * D2 * B * B - (D1 + D2) * B * (B + 1);
    const double c = -(A * B + D1 * D2 * B * B * (B + 1));
    const double Q = (a * a - 3 * b) / 9;
    const double R = (2 * a * a * a - 9 * a * b + 27 * c) / 54;
    const double Q3 = Q * Q * Q;
    const double M = R * R - Q3;
    if (M <= 0) {
        const double th = acos(R / sqrt(Q3));
        const double sq = -2 * sqrt(Q);
        double z0 = polish(sq * cos(th / 3) - a / 3, a, b, c);
        double z1 = polish(sq * cos((th + 2 * PI_REF) / 3) - a / 3, a, b, c);
        double z2 = polish(sq * cos((th - 2 * PI_REF) / 3) - a / 3, a, b, c);
        const double lo = fmin(z0, fmin(z1, z2));
        const double hi = fmax(z0, fmax(z1, z2));
        const double dG = (hi - lo) + log((lo - B) / (hi - B)) -
                          A / (B * (D2 - D1)) *
                          log((lo + D1 * B) * (hi + D2 * B) / ((lo + D2 * B) * (hi + D1 * B)));
        return dG > 0 ? lo : hi;
    }
    const double sM = sqrt(M);
    const double t1 = -R + sM, t2 = R + sM;
    const double S = copysign(cbrt(fabs(t1)), t1) * (t1 != 0);
    const double T = -copysign(cbrt(fabs(t2)), t2) * (t2 != 0);
    return polish(S + T - a / 3, a, b, c);
}

// ---------------------------------------------------------------------------
// Per-cell EoS constants at fixed (P, T): sqrt(Ai), Bi (src/OCPEoS.cpp:68-84).
// The reference recomputes them in every fugacity call; they depend only on
// (P, T), so they are computed once per cell here.
// ---------------------------------------------------------------------------
template <int NC>
struct EosCell {
    double sA[NC], Bi[NC];
};

template <int NC>
OCPX_HD void eos_cell(const DevFluid& f, double P, double T, EosCell<NC>& e)
{
    UNROLL for (int i = 0; i < NC; i++) {
        const double Pr = P / f.Pc[i], Tr = T / f.Tc[i];
        const double q = 1 + f.mw[i] * (1 - sqrt(Tr));
        const double Ai = f.OA[i] * Pr / (Tr * Tr) * (q * q);
        e.sA[i] = sqrt(Ai);
        e.Bi[i] = f.OB[i] * Pr / Tr;
    }
}

// Mixture evaluation at composition x: s_i = sum_k (1-k_ik) sqrt(Ai Ak) x_k,
// A = sum_i x_i s_i, B = sum_i x_i Bi, Z.
template <int NC>
struct EosMix {
    double s[NC];
    double A, B, Z;
};

template <int NC>
OCPX_HD void eos_mix(const DevFluid& f, const EosCell<NC>& e, const double* x,
                                        EosMix<NC>& m)
{
    double w[NC];
    UNROLL for (int k = 0; k < NC; k++) w[k] = e.sA[k] * x[k];
    double A = 0, B = 0;
    UNROLL for (int i = 0; i < NC; i++) {
        double t = 0;
        UNROLL for (int k = 0; k < NC; k++) t += f.omk[i * OCPX_NCMAX + k] * w[k];
        m.s[i] = e.sA[i] * t;
        A += x[i] * m.s[i];
        B += e.Bi[i] * x[i];
    }
    m.A = A;
    m.B = B;
    m.Z = pr_z(A, B);
}

// ln(phi_i) (src/OCPEoS.cpp:199-214 in log form)
template <int NC>
OCPX_HD void eos_lnphi(const EosCell<NC>& e, const EosMix<NC>& m, double* lnphi)
{
    const double A = m.A, B = m.B, Z = m.Z;
    const double lz = log(Z - B);
    const double lg = log((Z + D1 * B) / (Z + D2 * B));
    const double cA = A / (D1 - D2) / B * lg;
    const double iB = 1 / B, iA = 1 / A;
    UNROLL for (int i = 0; i < NC; i++)
        lnphi[i] = e.Bi[i] * iB * (Z - 1) - lz - cA * (2 * m.s[i] * iA - e.Bi[i] * iB);
}

// Shared pieces of the Z derivatives (src/OCPEoS.cpp:135-180)
template <int NC>
OCPX_HD double z_den(const EosMix<NC>& m)
{
    const double A = m.A, B = m.B, Z = m.Z;
    return 3 * Z * Z + 2 * ((D1 + D2 - 1) * B - 1) * Z + (A + D1 * D2 * B * B - (D1 + D2) * B * (B + 1));
}
template <int NC>
OCPX_HD double z_bcoef(const EosMix<NC>& m)
{
    const double A = m.A, B = m.B, Z = m.Z;
    return (A + D1 * D2 * (3 * B * B + 2 * B)) + ((D1 + D2) * (2 * B + 1) - 2 * D1 * D2 * B) * Z -
           (D1 + D2 - 1) * Z * Z;
}

// d ln f_i / d x_k, row i (src/OCPEoS.cpp:217-251).  Ax = 2 s, Bx = Bi.
template <int NC>
OCPX_HD void lnfugx_row(const DevFluid& f, const EosCell<NC>& e,
                                           const EosMix<NC>& m, const double* Zx, const double* x,
                                           double P, int i, double* row)
{
    const double A = m.A, B = m.B, Z = m.Z;
    const double G = (Z + D1 * B) / (Z + D2 * B), lG = log(G);
    const double ZB = Z - B;
    const double Axi = 2 * m.s[i], Bxi = e.Bi[i];
    const double C = x[i] * P / ZB;
    const double E = -A / ((D1 - D2) * B) * (Axi / A - Bxi / B);
    const double g2 = (D1 - D2) / (Z + D2 * B) / (Z + D2 * B);
    UNROLL for (int k = 0; k < NC; k++) {
        const double aik = f.omk[i * OCPX_NCMAX + k] * e.sA[i] * e.sA[k];
        const double Bxk = e.Bi[k], Axk = 2 * m.s[k];
        const double Cx = (ZB * (i == k ? 1.0 : 0.0) - x[i] * (Zx[k] - Bxk)) * P / (ZB * ZB);
        const double Dx = Bxi / B * (Zx[k] - Bxk * (Z - 1) / B);
        double Ex = (2 * (A / B * Bxk * Bxi + B * aik) - Axi * Bxk - Axk * e.Bi[i]) / (B * B);
        Ex /= -(D1 - D2);
        const double Gx = g2 * (Z * Bxk - Zx[k] * B);
        row[k] = 1 / C * Cx + Dx + Ex * lG + E / G * Gx;
    }
}

// Precomputations of d ln f / d n at (x, nt) (src/OCPEoS.cpp:153-168)
template <int NC>
struct LnfN {
    double An[NC], Bn[NC], Zn[NC];
};
template <int NC>
OCPX_HD void lnfn_prep(const EosCell<NC>& e, const EosMix<NC>& m, double nt,
                                          LnfN<NC>& d)
{
    const double den = z_den(m), bc = z_bcoef(m);
    UNROLL for (int i = 0; i < NC; i++) {
        d.An[i] = 2 / nt * (m.s[i] - m.A);
        d.Bn[i] = 1 / nt * (e.Bi[i] - m.B);
        d.Zn[i] = ((m.B - m.Z) * d.An[i] + bc * d.Bn[i]) / den;
    }
}

// Entry (i,k) of d ln f_i / d n_k (phi=false, src/OCPEoS.cpp:254-300) or of
// d ln phi_i / d n_k (phi=true, :343-382).  Symmetric in (i,k).
template <int NC, bool PHI>
OCPX_HD double lnfn_entry(const DevFluid& f, const EosCell<NC>& e,
                                             const EosMix<NC>& m, const LnfN<NC>& d,
                                             const double* x, double P, double nt, int i, int k)
{
    const double A = m.A, B = m.B, Z = m.Z;
    const double ZB = Z - B;
    const double G = (Z + D1 * B) / (Z + D2 * B);
    const double si = m.s[i];
    const double E = -A / ((D1 - D2) * B) * (2 * si / A - e.Bi[i] / B);
    const double aik = f.omk[i * OCPX_NCMAX + k] * e.sA[i] * e.sA[k];
    double C, Cn;
    if (PHI) {
        C = 1 / ZB;
        Cn = (d.Bn[k] - d.Zn[k]) / (ZB * ZB);
    } else {
        C = x[i] * P / ZB;
        Cn = P / ZB / ZB * (ZB / nt * ((i == k ? 1.0 : 0.0) - x[i]) - x[i] * (d.Zn[k] - d.Bn[k]));
    }
    const double Dn = e.Bi[i] / B * (d.Zn[k] - (e.Bi[k] - B) * (Z - 1) / (nt * B));
    const double Gn = (D1 - D2) / ((Z + D2 * B) * (Z + D2 * B)) * (d.Bn[k] * Z - d.Zn[k] * B);
    double En = -1 / (D1 - D2) / (B * B) *
                (2 * (B * aik / nt + d.Bn[k] * (e.Bi[i] * A / B - si)) - d.An[k] * e.Bi[i] - A * e.Bi[i] / nt);
    En -= E / nt;
    return 1 / C * Cn + Dn + En * log(G) + E / G * Gn;
}

// d ln f_i / d P (src/OCPEoS.cpp:303-340)
template <int NC>
OCPX_HD void lnfugp(const EosCell<NC>& e, const EosMix<NC>& m, double P,
                                       double* out)
{
    const double A = m.A, B = m.B, Z = m.Z;
    const double Ap = A / P, Bp = B / P;
    const double Zp = ((B - Z) * Ap + z_bcoef(m) * Bp) / z_den(m);
    const double G = (Z + D1 * B) / (Z + D2 * B);
    const double Gp = (D1 - D2) / ((Z + D2 * B) * (Z + D2 * B)) * (Bp * Z - Zp * B);
    const double C = P / (Z - B);
    const double Cp = ((Z - B) - P * (Zp - Bp)) / ((Z - B) * (Z - B));
    UNROLL for (int i = 0; i < NC; i++) {
        const double E = -A / ((D1 - D2) * B) * (2 * m.s[i] / A - e.Bi[i] / B);
        out[i] = Cp / C + e.Bi[i] / B * Zp + E / G * Gp;
    }
}

// Molar volume and derivatives (src/OCPEoS.cpp:396-414)
template <int NC>
OCPX_HD double vm_der(const DevFluid& f, const EosCell<NC>& e,
                                         const EosMix<NC>& m, const double* x, double P, double T,
                                         double& vmP, double* vmx)
{
    const double den = z_den(m), bc = z_bcoef(m);
    const double cg = R_GAS * T / P;
    double v = m.Z * cg;
    UNROLL for (int i = 0; i < NC; i++) v -= x[i] * f.vsh[i];
    UNROLL for (int i = 0; i < NC; i++) {
        const double Zx = ((m.B - m.Z) * (2 * m.s[i]) + bc * e.Bi[i]) / den;
        vmx[i] = cg * Zx - f.vsh[i];
    }
    const double Ap = m.A / P, Bp = m.B / P;
    const double Zp = ((m.B - m.Z) * Ap + bc * Bp) / den;
    vmP = cg * (Zp - m.Z / P);
    return v;
}

// ---------------------------------------------------------------------------
// Dense solve with partial pivoting on an N x N register matrix, K RHS
// columns.  Row swaps are done with predicated moves so every index stays a
// compile-time constant.  Returns false on a zero pivot.
// ---------------------------------------------------------------------------
template <int N, int K>
OCPX_HD bool gauss_solve(double (&A)[N][N], double (&b)[K][N])
{
    bool ok = true;
    UNROLL for (int c = 0; c < N; c++) {
        double best = fabs(A[c][c]);
        int p = c;
        UNROLL for (int r = c + 1; r < N; r++) {
            const double v = fabs(A[r][c]);
            if (v > best) { best = v; p = r; }
        }
        UNROLL for (int r = c + 1; r < N; r++) {
            if (p == r) {
                UNROLL for (int k = c; k < N; k++) { double t = A[c][k]; A[c][k] = A[r][k]; A[r][k] = t; }
                UNROLL for (int q = 0; q < K; q++) { double t = b[q][c]; b[q][c] = b[q][r]; b[q][r] = t; }
            }
        }
        const double d = A[c][c];
        ok = ok && (d != 0);
        const double id = 1 / d;
        UNROLL for (int r = c + 1; r < N; r++) {
            const double l = A[r][c] * id;
            UNROLL for (int k = c + 1; k < N; k++) A[r][k] -= l * A[c][k];
            UNROLL for (int q = 0; q < K; q++) b[q][r] -= l * b[q][c];
        }
    }
    UNROLL for (int q = 0; q < K; q++) {
        UNROLL for (int r = N - 1; r >= 0; r--) {
            double s = b[q][r];
            UNROLL for (int k = r + 1; k < N; k++) s -= A[r][k] * b[q][k];
            b[q][r] = s / A[r][r];
        }
    }
    return ok;
}

}  // namespace ocpx
