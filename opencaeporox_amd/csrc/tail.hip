// tail.hip — flash derivative tail, one lane per cell.
//
// Restates OCPMixtureMethodComp01::FlashDer after the phase equilibrium
// (src/OCPMixtureMethodComp.cpp:746-755): CalPropertyDer (:134-166),
// IdentifyPhase / ReOrderPhaseDer (:174-256), CalPropertyW (:846-860),
// CalVfS (include/OCPMixtureVarSet.hpp:143-156), CalVfiVfp_full02 +
// AddVfpVfiW (:462-562, :896-901), CaldXsdXp02 + CaldXsdXpW (:565-642,
// :904-920), SkipPSA::CalSkipForNextStep (src/AcceleratePEC.cpp:127-172) and
// IsoT_FIM::PassFlashValue (src/IsoThermalMethod.cpp:1027-1071).
//
// The phases are labelled before their properties are computed, so each
// phase's derivative vectors are written straight to their final slot.
// The implicit-function system for dn/d(P, N) is factorised once and its
// nc+1 right-hand sides are generated, solved and consumed one column at a
// time, which keeps one Nc x Nc matrix live instead of three.
#include "ocpx_dev.hpp"
#include "ocpx_kernels.hpp"

namespace ocpx {

template <int N>
OCPX_HD void lu_factor(double (&A)[N][N], int (&piv)[N])
{
    UNROLL for (int c = 0; c < N; c++) {
        double best = fabs(A[c][c]);
        int p = c;
        UNROLL for (int r = c + 1; r < N; r++) {
            const double v = fabs(A[r][c]);
            if (v > best) { best = v; p = r; }
        }
        piv[c] = p;
        UNROLL for (int r = c + 1; r < N; r++)
            if (p == r) UNROLL for (int k = c; k < N; k++) { double t = A[c][k]; A[c][k] = A[r][k]; A[r][k] = t; }
        const double id = 1 / A[c][c];
        UNROLL for (int r = c + 1; r < N; r++) {
            const double l = A[r][c] * id;
            A[r][c] = l;
            UNROLL for (int k = c + 1; k < N; k++) A[r][k] -= l * A[c][k];
        }
    }
}

template <int N>
OCPX_HD void lu_solve(const double (&A)[N][N], const int (&piv)[N], double (&b)[N])
{
    UNROLL for (int c = 0; c < N; c++) {
        UNROLL for (int r = c + 1; r < N; r++)
            if (piv[c] == r) { double t = b[c]; b[c] = b[r]; b[r] = t; }
        UNROLL for (int r = c + 1; r < N; r++) b[r] -= A[r][c] * b[c];
    }
    UNROLL for (int r = N - 1; r >= 0; r--) {
        double s = b[r];
        UNROLL for (int k = r + 1; k < N; k++) s -= A[r][k] * b[k];
        b[r] = s / A[r][r];
    }
}

// LBC viscosity and derivatives (ViscosityMethod03, src/OCPFuncPVT.cpp:462-575).
// tt[i] is the per-component reduced-temperature term (depends on T only).
template <int NC>
OCPX_HD double lbc(const DevFluid& f, const double* tt, const double* x, double xi,
                                      double xiP, const double* xix, double& muP, double* mux)
{
    double MW = 0, a0 = 0, a1 = 0, xPc = 0, xTc = 0, xVc = 0;
    UNROLL for (int i = 0; i < NC; i++) MW += x[i] * f.MW[i];
    const double sMW = sqrt(MW);
    UNROLL for (int i = 0; i < NC; i++) {
        const double t = f.auxB[i] / sMW;
        a0 += x[i] * f.sqrtMW[i] * (tt[i] / t);
        a1 += x[i] * f.sqrtMW[i];
        xPc += x[i] * f.Pc[i];
        xTc += x[i] * f.Tc[i];
        xVc += x[i] * f.Vcvis[i];
    }
    const double xTc16 = pow(xTc, 1.0 / 6), xPc23 = pow(xPc, 2.0 / 3);
    const double a2 = 5.4402 * xTc16 / sMW / xPc23;
    const double a3 = xi * xVc;
    const double* c = f.lbc;
    const double a4 = a3 * (a3 * (a3 * (c[4] * a3 + c[3]) + c[2]) + c[1]) + c[0];
    const double a4_3 = a4 * a4 * a4, a4_4 = a4_3 * a4;
    const double mu = a0 / a1 + 1E-4 * (a4_4 - 1) / a2;
    const double dpoly = c[1] + a3 * (2 * c[2] + a3 * (3 * c[3] + a3 * 4 * c[4]));
    muP = (4 * 1E-4) * a4_3 * (xVc * xiP * dpoly) / a2;
    // d(sqrt MW) part of d(a0)/dx_k: 0 in the reference (-(1/2) integer
    // division, :537), 0.5 * MW_k * a0 / MW when lbc_exact is set
    const double ex = f.lbc_exact ? 0.5 * a0 / MW : 0.0;
    const double xTcm56 = pow(xTc, -5.0 / 6);
    UNROLL for (int k = 0; k < NC; k++) {
        const double v2 = tt[k] / (f.auxB[k] / sMW);
        const double d3 = ex * f.MW[k] + f.sqrtMW[k] * v2;
        const double d4 = f.sqrtMW[k];
        const double d6 = 5.4402 * (1.0 / 6 * xTcm56 * f.Tc[k] - xTc16 * (0.5 / MW * f.MW[k] + 2.0 / 3 / xPc * f.Pc[k])) /
                          (sMW * xPc23);
        const double d8 = (xix[k] * xVc + xi * f.Vcvis[k]) * dpoly;
        mux[k] = (d3 * a1 - a0 * d4) / (a1 * a1) + 1E-4 * (4 * a4_3 * d8 * a2 - (a4_4 - 1) * d6) / (a2 * a2);
    }
    return mu;
}

// Smallest eigenvalue of the float-rounded symmetric matrix (stands for
// ssyevd_, src/DenseMat.cpp:126-145): cyclic Jacobi in double, rounded to
// float.
template <int NC>
OCPX_HD float min_eig(double (&A)[NC][NC])
{
    for (int sweep = 0; sweep < 60; sweep++) {
        double off = 0;
        UNROLL for (int p = 0; p < NC; p++) UNROLL for (int q = p + 1; q < NC; q++) off += A[p][q] * A[p][q];
        if (off < 1e-30) break;
        UNROLL for (int p = 0; p < NC; p++) {
            UNROLL for (int q = p + 1; q < NC; q++) {
                const double apq = A[p][q];
                if (fabs(apq) < 1e-300) continue;
                const double th = (A[q][q] - A[p][p]) / (2 * apq);
                double t = (th > 0 ? 1.0 : (th < 0 ? -1.0 : 0.0)) / (fabs(th) + sqrt(th * th + 1));
                if (th == 0) t = 1;
                const double cs = 1 / sqrt(t * t + 1), sn = t * cs;
                UNROLL for (int k = 0; k < NC; k++) {
                    const double akp = A[k][p], akq = A[k][q];
                    A[k][p] = cs * akp - sn * akq;
                    A[k][q] = sn * akp + cs * akq;
                }
                UNROLL for (int k = 0; k < NC; k++) {
                    const double apk = A[p][k], aqk = A[q][k];
                    A[p][k] = cs * apk - sn * aqk;
                    A[q][k] = sn * apk + cs * aqk;
                }
            }
        }
    }
    double mn = A[0][0];
    UNROLL for (int i = 1; i < NC; i++) mn = fmin(mn, A[i][i]);
    return (float)mn;
}

template <int NC>
__global__ void __launch_bounds__(64) flash_tail(const DevFluid* __restrict__ fp, TailIO io)
{
    const int64_t pos = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (io.list ? pos >= (int64_t)*io.nlist : pos >= io.nb) return;
    const int64_t cell = io.list ? (int64_t)io.list[pos] : pos;
    if (!io.list && io.route) {
        const int r = io.route[cell];
        if ((r == 1 || r == 5) != (io.route_pass == 1)) return;
    }
    // one fluid: a uniform pointer; several models: the cell's own fluid
    const DevFluid& f = io.mid ? fp[io.mid[cell]] : *fp;
    const size_t n = io.nb;
    constexpr int nc = NC + 1, np = 3, ncol = nc + 1, wC = NC;
    auto at = [&](int k) { return (size_t)k * n + cell; };

    const double P = io.P[cell];
    const double T = io.T[cell] + CONV4;
    double Ni[nc];
    UNROLL for (int i = 0; i < nc; i++) Ni[i] = io.Ni[at(i)];
    double Nt = 0;
    UNROLL for (int i = 0; i < NC; i++) Nt += Ni[i];
    double zi[NC];
    UNROLL for (int i = 0; i < NC; i++) zi[i] = Ni[i] / Nt;             // InitNtZ
    const int NP = io.pe_NP[cell];
    const double nuPE0 = io.pe_nu[cell];

    EosCell<NC> ec;
    eos_cell<NC>(f, P, T, ec);
    double tt[NC];                                   // LBC reduced-temperature terms
    UNROLL for (int i = 0; i < NC; i++) {
        const double Tr = T / f.Tc[i];
        tt[i] = Tr <= 1.5 ? 34 * 1E-5 * pow(Tr, 0.94) : 17.78 * 1E-5 * pow((4.58 * Tr - 1.67), 0.625);
    }

    // ---- phase identification (IdentifyPhase / ReOrderPhaseDer)
    double x[2][NC];           // compositions in final slots: 0 oil, 1 gas
    double nuF[2];
    bool ex[2];
    {
        double xa[NC], xb[NC];
        UNROLL for (int i = 0; i < NC; i++) { xa[i] = io.pe_x[at(i)]; xb[i] = io.pe_x[at(NC + i)]; }
        if (NP == 1) {
            double A = 0, B = 0;
            UNROLL for (int i = 0; i < NC; i++) { A += zi[i] * f.Vc[i] * f.Tc[i]; B += zi[i] * f.Vc[i]; }
            const bool gas = T > A / B;
            ex[0] = !gas;
            ex[1] = gas;
            UNROLL for (int i = 0; i < NC; i++) { x[0][i] = xa[i]; x[1][i] = xa[i]; }
            nuF[0] = nuF[1] = 1;
        } else {
            double MWa = 0, MWb = 0;
            UNROLL for (int i = 0; i < NC; i++) { MWa += xa[i] * f.MW[i]; MWb += xb[i] * f.MW[i]; }
            const bool aOil = MWa > MWb;
            ex[0] = ex[1] = true;
            UNROLL for (int i = 0; i < NC; i++) { x[0][i] = aOil ? xa[i] : xb[i]; x[1][i] = aOil ? xb[i] : xa[i]; }
            nuF[0] = aOil ? nuPE0 : 1 - nuPE0;
            nuF[1] = aOil ? 1 - nuPE0 : nuPE0;
        }
    }

    // ---- hydrocarbon phase properties (CalVmVj, CalXiRhoMuDer)
    double nj[2], vj[2], vm[2], vmP[2], vmx[2][NC];
    UNROLL for (int j = 0; j < 2; j++) {
        nj[j] = Nt * nuF[j];
        vj[j] = 0;
        if (!ex[j]) continue;
        EosMix<NC> m;
        eos_mix<NC>(f, ec, x[j], m);
        vm[j] = vm_der<NC>(f, ec, m, x[j], P, T, vmP[j], vmx[j]);
        vj[j] = nj[j] * vm[j];
        const double xi = 1 / vm[j];
        const double xiP = -xi * xi * vmP[j];
        double xix[NC], rhox[NC], mux[NC];
        double MW = 0;
        UNROLL for (int i = 0; i < NC; i++) MW += x[j][i] * f.MW[i];
        UNROLL for (int i = 0; i < NC; i++) {
            xix[i] = -xi * xi * vmx[j][i];
            rhox[i] = xix[i] * MW + xi * f.MW[i];
        }
        double muP;
        const double mu = lbc<NC>(f, tt, x[j], xi, xiP, xix, muP, mux);
        io.rho[at(j)] = MW * xi;
        io.xi[at(j)] = xi;
        io.mu[at(j)] = mu;
        io.rhoP[at(j)] = MW * xiP;
        io.xiP[at(j)] = xiP;
        io.muP[at(j)] = muP;
        UNROLL for (int i = 0; i < NC; i++) {
            io.xij[at(j * nc + i)] = x[j][i];
            io.rhox[at(j * nc + i)] = rhox[i];
            io.xix[at(j * nc + i)] = xix[i];
            io.mux[at(j * nc + i)] = mux[i];
        }
        io.rhox[at(j * nc + wC)] = 0;
        io.xix[at(j * nc + wC)] = 0;
        io.mux[at(j * nc + wC)] = 0;
    }

    // ---- water (CalPropertyW with vw = -1, OCP_PVTW::CalRhoXiMuDer)
    double xiw, xiPw;
    {
        const double Pref = f.pvtw[0], bref = f.pvtw[1], Cb = f.pvtw[2], muref = f.pvtw[3], Cmu = f.pvtw[4];
        const double b = bref * (1 - Cb * (P - Pref));
        const double bp = -Cb * bref;
        xiw = 1 / b;
        xiPw = -bp / (b * b);
        io.rho[at(2)] = f.rhoW / b;
        io.xi[at(2)] = xiw;
        io.mu[at(2)] = muref * (1 + Cmu * (P - Pref));
        io.rhoP[at(2)] = -f.rhoW * bp / (b * b);
        io.xiP[at(2)] = xiPw;
        io.muP[at(2)] = Cmu * muref;
        io.xij[at(2 * nc + wC)] = 1.0;
        UNROLL for (int i = 0; i < nc; i++) {
            io.rhox[at(2 * nc + i)] = 0;
            io.xix[at(2 * nc + i)] = 0;
            io.mux[at(2 * nc + i)] = 0;
        }
    }
    const double njw = Ni[wC];
    const double vjw = njw / xiw;
    const double vjiw = 1 / xiw;
    const double vjPw = -njw * xiPw / (xiw * xiw);

    // ---- CalVfS
    const double Vf = (ex[0] ? vj[0] : 0.0) + (ex[1] ? vj[1] : 0.0) + vjw;
    double S[3];
    S[0] = ex[0] ? vj[0] / Vf : 0;
    S[1] = ex[1] ? vj[1] / Vf : 0;
    S[2] = vjw / Vf;
    UNROLL for (int j = 0; j < 3; j++) {
        io.S[at(j)] = S[j];
        io.phaseExist[at(j)] = j < 2 ? (ex[j] ? 1u : 0u) : 1u;
    }
    io.Nt[cell] = Nt + Ni[wC];
    io.vf[cell] = Vf;
    const double iVf = 1 / Vf;

    // dXsdXp rows: S_j (j < 3), then x_ij at row np + j*nc + i; all written
    auto W = [&](int row, int col, double v) { io.dXsdXp[((size_t)cell * 3 * ncol + row) * ncol + col] = v; };
    double vfP;
    if (NP == 1) {
        // the existing phase j is a runtime value: pick its data with selects
        const bool j0 = ex[0];
        const int j = j0 ? 0 : 1;
        double xj[NC], vmxj[NC];
        UNROLL for (int k = 0; k < NC; k++) { xj[k] = j0 ? x[0][k] : x[1][k]; vmxj[k] = j0 ? vmx[0][k] : vmx[1][k]; }
        const double njj = j0 ? nj[0] : nj[1], vmj = j0 ? vm[0] : vm[1], vmPj = j0 ? vmP[0] : vmP[1];
        const double Sj = j0 ? S[0] : S[1];
        double t = 0;
        UNROLL for (int k = 0; k < NC; k++) t += vmxj[k] * xj[k];
        const double vjP = njj * vmPj;
        vfP = vjP + vjPw;                                      // + AddVfpVfiW
        io.vfP[cell] = vfP;
        io.vfi[at(wC)] = vjiw;
        W(j, 0, (vjP - vfP * Sj) * iVf);
        W(1 - j, 0, 0);
        W(2, 0, (vjPw - vfP * S[2]) * iVf);
        UNROLL for (int m = 0; m < NC; m++) {
            const double vji = vmj + vmxj[m] - t;
            io.vfi[at(m)] = vji;
            W(j, 1 + m, (vji - vji * Sj) * iVf);
            W(1 - j, 1 + m, 0);
            W(2, 1 + m, (0 - vji * S[2]) * iVf);
        }
        W(j, 1 + wC, (0 - vjiw * Sj) * iVf);
        W(1 - j, 1 + wC, 0);
        W(2, 1 + wC, (vjiw - vjiw * S[2]) * iVf);
        // x rows: dx_ij/dN_m of the existing phase, zero elsewhere
        UNROLL for (int jj = 0; jj < 3; jj++)
            UNROLL for (int i = 0; i < nc; i++)
                UNROLL for (int cc = 0; cc < ncol; cc++) {
                    double v = 0;
                    if (i < NC && cc >= 1 && cc <= NC)
                        v = (jj == j) ? ((i == cc - 1 ? 1.0 : 0.0) * Nt - Ni[i]) / (Nt * Nt) : 0.0;
                    W(np + jj * nc + i, cc, v);
                }
    } else {
        // implicit-function system: (dlnf0/dn + dlnf1/dn) dn0 = rhs
        double J[NC][NC];
        int piv[NC];
        EosMix<NC> mg;
        LnfN<NC> dg;
        double lfp[NC];
        {
            EosMix<NC> mo;
            LnfN<NC> dO;
            eos_mix<NC>(f, ec, x[0], mo);
            eos_mix<NC>(f, ec, x[1], mg);
            lnfn_prep<NC>(ec, mo, nj[0], dO);
            lnfn_prep<NC>(ec, mg, nj[1], dg);
            UNROLL for (int i = 0; i < NC; i++)
                UNROLL for (int k = 0; k <= i; k++) {
                    const double v = lnfn_entry<NC, false>(f, ec, mo, dO, x[0], P, nj[0], i, k) +
                                     lnfn_entry<NC, false>(f, ec, mg, dg, x[1], P, nj[1], i, k);
                    J[i][k] = v;
                    J[k][i] = v;
                }
            double lpo[NC];
            lnfugp<NC>(ec, mo, P, lpo);
            lnfugp<NC>(ec, mg, P, lfp);
            UNROLL for (int i = 0; i < NC; i++) lfp[i] -= lpo[i];
        }
        lu_factor<NC>(J, piv);
        double t0 = 0, t1 = 0;
        UNROLL for (int i = 0; i < NC; i++) { t0 -= vmx[0][i] * x[0][i]; t1 -= vmx[1][i] * x[1][i]; }
        double vfi[NC];
        // column 0: d/dP
        {
            double d[NC];
            UNROLL for (int i = 0; i < NC; i++) d[i] = lfp[i];
            lu_solve<NC>(J, piv, d);
            double vjP0 = nj[0] * vmP[0], vjP1 = nj[1] * vmP[1], sum = 0;
            UNROLL for (int k = 0; k < NC; k++) {
                vjP0 += (vm[0] + t0 + vmx[0][k]) * d[k];
                vjP1 -= (vm[1] + t1 + vmx[1][k]) * d[k];
                sum += d[k];
            }
            vfP = vjP0 + vjP1 + vjPw;
            io.vfP[cell] = vfP;
            W(0, 0, (vjP0 - vfP * S[0]) * iVf);
            W(1, 0, (vjP1 - vfP * S[1]) * iVf);
            W(2, 0, (vjPw - vfP * S[2]) * iVf);
            UNROLL for (int i = 0; i < NC; i++) {
                W(np + i, 0, (d[i] - sum * x[0][i]) / nj[0]);
                W(np + nc + i, 0, (-d[i] + sum * x[1][i]) / nj[1]);
            }
        }
        // columns 1..NC: d/dN_m, right-hand side = row m of dlnf1/dn
        UNROLL for (int m = 0; m < NC; m++) {
            double d[NC];
            // entry (m, i) as the reference stores it: computed at (max, min), mirrored
            UNROLL for (int i = 0; i < NC; i++)
                d[i] = lnfn_entry<NC, false>(f, ec, mg, dg, x[1], P, nj[1], m > i ? m : i, m > i ? i : m);
            lu_solve<NC>(J, piv, d);
            double v0 = 0, v1 = 0, sum = 0;
            UNROLL for (int k = 0; k < NC; k++) {
                v0 += (vm[0] + t0 + vmx[0][k]) * d[k];
                v1 += (vm[1] + t1 + vmx[1][k]) * ((m == k ? 1.0 : 0.0) - d[k]);
                sum += d[k];
            }
            vfi[m] = v0 + v1;
            io.vfi[at(m)] = vfi[m];
            W(0, 1 + m, (v0 - vfi[m] * S[0]) * iVf);
            W(1, 1 + m, (v1 - vfi[m] * S[1]) * iVf);
            W(2, 1 + m, (0 - vfi[m] * S[2]) * iVf);
            UNROLL for (int i = 0; i < NC; i++) {
                W(np + i, 1 + m, (d[i] - sum * x[0][i]) / nj[0]);
                W(np + nc + i, 1 + m, ((m == i ? 1.0 : 0.0) - d[i] - (1 - sum) * x[1][i]) / nj[1]);
            }
        }
        io.vfi[at(wC)] = vjiw;
        W(0, 1 + wC, (0 - vjiw * S[0]) * iVf);
        W(1, 1 + wC, (0 - vjiw * S[1]) * iVf);
        W(2, 1 + wC, (vjiw - vjiw * S[2]) * iVf);
        UNROLL for (int i = 0; i < NC; i++) {
            W(np + i, 1 + wC, 0);
            W(np + nc + i, 1 + wC, 0);
        }
        UNROLL for (int cc = 0; cc < ncol; cc++) {
            W(np + wC, cc, 0);
            W(np + nc + wC, cc, 0);
            UNROLL for (int i = 0; i < nc; i++) W(np + 2 * nc + i, cc, 0);
        }
    }

    // ---- SkipPSA::CalSkipForNextStep
    if (f.skip) {
        const int ft = io.ftype_out[cell];
        if (ft == 0) {
            EosMix<NC> m;
            LnfN<NC> d;
            eos_mix<NC>(f, ec, zi, m);
            lnfn_prep<NC>(ec, m, Nt, d);
            double M[NC][NC];
            UNROLL for (int i = 0; i < NC; i++)
                UNROLL for (int j = 0; j <= i; j++) {
                    const double v = (i == j ? 1.0 : 0.0) +
                                     Nt * sqrt(zi[i] * zi[j]) * lnfn_entry<NC, true>(f, ec, m, d, zi, P, Nt, i, j);
                    const double vf = (double)(float)v;
                    M[i][j] = vf;
                    M[j][i] = vf;
                }
            io.skip_flag[cell] = 1;
            io.skip_minEig[cell] = (double)min_eig<NC>(M);
            io.skip_P[cell] = P;
            io.skip_T[cell] = T;
            UNROLL for (int i = 0; i < NC; i++) io.skip_zi[at(i)] = zi[i];
        } else {
            io.skip_flag[cell] = ft == 1 ? 1u : 0u;
        }
    }
}

template <int NC>
static hipError_t launch_tail(const DevFluid* f, const TailIO& io, hipStream_t st)
{
    if (io.nb == 0) return hipSuccess;
    const int bs = 64;
    hipLaunchKernelGGL(flash_tail<NC>, dim3((unsigned)((io.nb + bs - 1) / bs)), dim3(bs), 0, st, f, io);
    return hipGetLastError();
}

hipError_t run_tail(int NC, const DevFluid* f, const TailIO& io, hipStream_t st)
{
    switch (NC) {
#define OCPX_CASE(N) case N: return launch_tail<N>(f, io, st);
        OCPX_CASE(2) OCPX_CASE(3) OCPX_CASE(4) OCPX_CASE(5) OCPX_CASE(6) OCPX_CASE(7)
        OCPX_CASE(8) OCPX_CASE(9) OCPX_CASE(10) OCPX_CASE(11) OCPX_CASE(12)
#undef OCPX_CASE
    default: return hipErrorInvalidValue;
    }
}

}  // namespace ocpx
