// ref_flash.cpp — C entry points over the reference's own flash classes,
// compiled together with the reference sources of the compositional flash:
//   src/OCPPhaseEquilibrium.cpp, src/OCPMixtureMethodComp.cpp,
//   src/AcceleratePEC.cpp, src/DenseMat.cpp and what they call (EoS, cubic,
//   PVT, tables, units, parameter records)
// and the LAPACK/BLAS routines DenseMat calls (dgesv_, dsysv_, ssyevd_,
// dnrm2_, dscal_, dgemm_ ...) from the MKL runtime this image ships
// (recipe: oracle/Makefile target `ref`, output oracle/_ref/libocpref_flash.so).
//
// TEST INFRASTRUCTURE ONLY.  It pins the C oracle's restatement of the flash
// (PhaseEquilibrium, FlashDer with dXsdXp, SkipPSA, IdentifyPhase) against
// the reference itself, on the same inputs; nothing in the product path
// loads it.
#include <cstdio>
#include <string>
#include <vector>

#include "AcceleratePEC.hpp"
#include "BulkVarSet.hpp"
#include "OCPMixtureMethodComp.hpp"
#include "OCPPhaseEquilibrium.hpp"
#include "ParamReservoir.hpp"
#include "../include/ocpx.h"
#include "oracle.h"

using std::string;
using std::vector;

namespace {

bool units_set = false;
void metric()
{
    if (!units_set) { SetUnit("METRIC"); units_set = true; }
}

string num(double v)
{
    char b[64];
    snprintf(b, sizeof(b), "%.17g", v);
    return b;
}

// ParamReservoir of one compositional PVT region as the deck keywords give
// it (COMPS, TCRIT, ..., PVTW, SSMSTA, NRSTA, SSMSP, NRSP, RR)
void make_params(const ocpx_fluid* f, ParamReservoir& rs)
{
    metric();
    ComponentParam& p = rs.comsParam;
    const int n = f->nc;
    p.numCom = n;
    p.numPhase = 2;
    p.NTPVT = 1;
    auto set = [&](Type_A_r<vector<double>>& t, const double* a) {
        t.activity = 1;
        t.data.assign(1, vector<double>(a, a + n));
    };
    set(p.Tc, f->Tc);
    set(p.Pc, f->Pc);
    set(p.Vc, f->Vc);
    set(p.Vcvis, f->Vcvis);
    set(p.MW, f->MW);
    set(p.Acf, f->acf);
    set(p.OmegaA, f->OmegaA);
    set(p.OmegaB, f->OmegaB);
    set(p.Vshift, f->Vshift);
    vector<double> bic(n * n);
    for (int i = 0; i < n; i++)
        for (int k = 0; k < n; k++) bic[i * n + k] = f->BIC[i * OCPX_MAX_NC + k];
    p.BIC.assign(1, bic);
    p.LBCcoef.assign(f->LBCcoef, f->LBCcoef + 5);
    p.SSMparamSTA = {std::to_string(f->SSMsta_maxIt), num(f->SSMsta_tol), num(f->SSMsta_eYt)};
    p.NRparamSTA = {std::to_string(f->NRsta_maxIt), num(f->NRsta_tol)};
    p.SSMparamSP = {std::to_string(f->SSMsp_maxIt), num(f->SSMsp_tol)};
    p.NRparamSP = {std::to_string(f->NRsp_maxIt), num(f->NRsp_tol)};
    p.RRparam = {std::to_string(f->RR_maxIt), num(f->RR_tol)};
    rs.numCom = n;
    rs.numPhase = 3;
    rs.comps = OCP_TRUE;
    vector<vector<double>> pvtw(5);
    for (int k = 0; k < 5; k++) pvtw[k].assign(1, f->PVTW[k]);
    rs.PVTW_T.data.assign(1, pvtw);
    // standard water density: RHOW_STD (DENSITY given) or RHOW_STD * gravity
    // (GRAVITY given), OCPMixtureMethodComp01's constructor
    if (f->rhoW_std == RHOW_STD) {
        rs.density.activity = 1;
        rs.density.data.assign(3, 0.0);
        rs.gravity.activity = 0;
    } else {
        rs.gravity.activity = 1;
        rs.gravity.data.assign(3, 1.0);
        rs.gravity.data[1] = f->rhoW_std / RHOW_STD;
        rs.density.activity = 0;
    }
}

// access to the flash's protected parts: the phase-equilibrium object and
// its iteration totals, and the SkipPSA variable set
struct MixProbe : public OCPMixtureMethodComp01 {
    MixProbe(const ParamReservoir& rs, OCPMixtureVarSet& vs) : OCPMixtureMethodComp01(rs, 0, vs) {}
    OCPPhaseEquilibrium& pe() { return PE; }
};

struct PEProbe : public OCPPhaseEquilibrium {
    static void totals(const OCPPhaseEquilibrium& p, unsigned long long* t)
    {
        const PEProbe& q = static_cast<const PEProbe&>(p);
        t[0] = q.itersSSMSTA; t[1] = q.itersNRSTA; t[2] = q.itersSSMSP; t[3] = q.itersNRSP; t[4] = q.itersRR;
    }
};

struct SkipProbe : public SkipPSA {
    SkipPSAVarset& set() { return vs; }
};

}  // namespace

extern "C" {

// OCPPhaseEquilibrium::PhaseEquilibrium (src/OCPPhaseEquilibrium.cpp:71-112)
// for n compositions, through the reference's own Comp01 set-up: T in K, z
// normalised over the Nc hydrocarbon components; ftype / lNP / lx as the
// C oracle's orc_phase_equilibrium (lx with the bulk nc = Nc + 1 stride).
// Outputs NP [n], nu [n*2], x [n*2*Nc] (PE order), ftype_out [n], and the
// per-cell increments of the reference's SSMSTA, NRSTA, SSMSP, NRSP and RR
// iteration totals in iters5 [n*5] (any output may be NULL).
void ref_phase_equilibrium(const ocpx_fluid* f, int64_t n, const double* P, const double* T, const double* z,
                           const int32_t* ftype, const int32_t* lNP, const double* lx, int32_t* NP, double* nu,
                           double* x, int32_t* ftype_out, int64_t* iters5)
{
    ParamReservoir rs;
    make_params(f, rs);
    OCPMixtureVarSet vs;
    MixProbe m(rs, vs);
    OCPPhaseEquilibrium& pe = m.pe();
    const int NC = f->nc, nc = NC + 1;
    unsigned long long t0[5], t1[5];
    for (int64_t c = 0; c < n; c++) {
        PEProbe::totals(pe, t0);
        const USI ft = ftype ? (USI)ftype[c] : 0;
        const double lnp = lNP ? (double)lNP[c] : 0.0;
        pe.PhaseEquilibrium(P[c], T[c], z + c * NC, ft, lnp, lx ? lx + c * 2 * nc : nullptr, nc);
        PEProbe::totals(pe, t1);
        if (NP) NP[c] = (int32_t)pe.GetNP();
        for (int j = 0; j < 2; j++) {
            if (nu) nu[c * 2 + j] = j < (int)pe.GetNP() ? pe.GetNu(j) : 0.0;
            if (x)
                for (int i = 0; i < NC; i++) x[(c * 2 + j) * NC + i] = j < (int)pe.GetNP() ? pe.GetX(j)[i] : 0.0;
        }
        if (ftype_out) ftype_out[c] = (int32_t)pe.GetFtype();
        if (iters5)
            for (int k = 0; k < 5; k++) iters5[c * 5 + k] = (int64_t)(t1[k] - t0[k]);
    }
}

// IsoT_FIM::CalFlash + PassFlashValue (src/IsoThermalMethod.cpp:1016-1071)
// with OCPMixtureComp::FlashDer (include/OCPMixture.hpp:85-90): SetVarSet,
// SkipPSA::CalFtype02 (entries of ftype_in >= 0 override it, as the C
// oracle's orc_cal_flash), OCPMixtureMethodComp01::FlashDer,
// SkipPSA::CalSkipForNextStep, on the cells of b in order.  The SkipPSA
// state is read from and written back to b's skip_* arrays, ftype_out
// receives the phase-equilibrium ftype.  iters and pe_ill are not written.
void ref_cal_flash(const ocpx_fluid* f, orc_bulk* b)
{
    ParamReservoir rs;
    make_params(f, rs);
    OCPMixtureVarSet vs;
    MixProbe m(rs, vs);
    const int NC = f->nc, nc = NC + 1, np = 3;
    const int64_t nb = b->nb;
    const int lend = np * (nc + 1) * (nc + 1);
    SkipProbe skip;
    skip.SetUseSkip(f->skipPSA ? OCP_TRUE : OCP_FALSE);
    const USI sidx = skip.Setup(nb, &m);
    if (f->skipPSA) {
        SkipPSAVarset& s = skip.set();
        for (int64_t c = 0; c < nb; c++) {
            s.flag[c] = b->skip_flag[c];
            s.minEigen[c] = b->skip_minEig[c];
            s.P[c] = b->skip_P[c];
            s.T[c] = b->skip_T[c];
            for (int i = 0; i < NC; i++) s.zi[c * NC + i] = b->skip_zi[c * NC + i];
        }
    }
    // one-cell BulkVarSet for SetVarSet (bId 0)
    BulkVarSet bv;
    bv.np = np;
    bv.nc = nc;
    bv.P.assign(1, 0.0);
    bv.T.assign(1, 0.0);
    bv.Pj.assign(np, 0.0);
    bv.Ni.assign(nc, 0.0);
    bv.S.assign(np, 0.0);
    bv.phaseExist.assign(np, 0);
    bv.xij.assign(np * nc, 0.0);
    for (int64_t c = 0; c < nb; c++) {
        bv.P[0] = b->P[c];
        bv.T[0] = b->T[c];
        for (int j = 0; j < np; j++) {
            bv.Pj[j] = b->Pj[c * np + j];
            bv.S[j] = b->S[c * np + j];
            bv.phaseExist[j] = b->phaseExist[c * np + j];
        }
        for (int i = 0; i < nc; i++) bv.Ni[i] = b->Ni[c * nc + i];
        for (int k = 0; k < np * nc; k++) bv.xij[k] = b->xij[c * np * nc + k];
        m.SetVarSet(0, bv, vs);
        USI ftype;
        if (b->ftype_in && b->ftype_in[c] >= 0) ftype = (USI)b->ftype_in[c];
        else ftype = skip.CalFtype02(c, sidx, vs);
        m.FlashDer(vs, ftype);
        skip.CalSkipForNextStep(c, sidx, vs);
        if (b->ftype_out) b->ftype_out[c] = (int32_t)m.GetFtype();
        // PassFlashValue
        b->Nt[c] = vs.Nt;
        b->vf[c] = vs.Vf;
        for (int j = 0; j < np; j++) {
            b->S[c * np + j] = vs.S[j];
            b->phaseExist[c * np + j] = vs.phaseExist[j];
            if (vs.phaseExist[j]) {
                b->rho[c * np + j] = vs.rho[j];
                b->xi[c * np + j] = vs.xi[j];
                b->mu[c * np + j] = vs.mu[j];
                b->rhoP[c * np + j] = vs.rhoP[j];
                b->xiP[c * np + j] = vs.xiP[j];
                b->muP[c * np + j] = vs.muP[j];
                for (int i = 0; i < nc; i++) {
                    b->xij[(c * np + j) * nc + i] = vs.x[j * nc + i];
                    b->rhox[(c * np + j) * nc + i] = vs.rhox[j * nc + i];
                    b->xix[(c * np + j) * nc + i] = vs.xix[j * nc + i];
                    b->mux[(c * np + j) * nc + i] = vs.mux[j * nc + i];
                }
            }
        }
        b->vfP[c] = vs.vfP;
        for (int i = 0; i < nc; i++) b->vfi[c * nc + i] = vs.vfi[i];
        for (int k = 0; k < lend; k++) b->dSec_dPri[c * lend + k] = vs.dXsdXp[k];
    }
    if (f->skipPSA) {
        SkipPSAVarset& s = skip.set();
        for (int64_t c = 0; c < nb; c++) {
            b->skip_flag[c] = s.flag[c];
            b->skip_minEig[c] = s.minEigen[c];
            b->skip_P[c] = s.P[c];
            b->skip_T[c] = s.T[c];
            for (int i = 0; i < NC; i++) b->skip_zi[c * NC + i] = s.zi[c * NC + i];
        }
    }
}

}  // extern "C"
