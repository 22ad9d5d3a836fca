// ref_harness.cpp — C entry points over the reference's own classes, compiled
// together with the reference sources that build without external libraries:
//   src/OCPEoS.cpp, src/UtilMath.cpp, src/OCPUnits.cpp, src/OCPFuncPVT.cpp,
//   src/OCPTable.cpp, src/OCPFuncSAT.cpp, src/OCPFlowMethod.cpp, src/OCPRock.cpp
// (recipe: oracle/Makefile target `ref`, output oracle/_ref/libocpref.so).
//
// TEST INFRASTRUCTURE ONLY.  It is used to pin the C oracle's restatement of
// these functions against the reference itself; nothing in the product path
// loads it.  The phase-equilibrium and mixture-method sources need LAPACK /
// BLAS (src/DenseMat.cpp) and are not built here.
#include <vector>
#include <string>
#include "OCPEoS.hpp"
#include "UtilMath.hpp"
#include "OCPFuncPVT.hpp"
#include "OCPFlowMethod.hpp"
#include "OCPRock.hpp"
#include "../include/ocpx.h"

using std::vector;

namespace {

// SetUnit writes globals defined in src/OCPUnits.cpp (std::string among
// them), so it must run after static initialisation: call it lazily.
bool units_set = false;
void metric()
{
    if (!units_set) { SetUnit("METRIC"); units_set = true; }
}

vector<double> take(const double* a, int n) { return vector<double>(a, a + n); }

ComponentParam make_param(const ocpx_fluid* f)
{
    metric();
    ComponentParam p;
    const int n = f->nc;
    p.numCom = n;
    p.numPhase = 2;
    auto set = [&](Type_A_r<vector<double>>& t, const double* a) {
        t.activity = 1;
        t.data.assign(1, take(a, n));
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
    p.LBCcoef = take(f->LBCcoef, 5);
    return p;
}

// expose the protected EoS workspace
struct EoSProbe : public EoS_PR {
    EoSProbe(const ComponentParam& p) : EoS_PR(p, 0) {}
    double A() const { return Aj; }
    double B() const { return Bj; }
    double Z() const { return Zj; }
};

vector<vector<double>> columns(const double* rows, int nrow)
{
    vector<vector<double>> c(4, vector<double>(nrow));
    for (int r = 0; r < nrow; r++)
        for (int k = 0; k < 4; k++) c[k][r] = rows[r * 4 + k];
    return c;
}

}  // namespace

extern "C" {

int ref_cubic(double a, double b, double c, double* z)
{
    metric();
    vector<double> zz(3, 0.0);
    unsigned k = CubicRoot(a, b, c, OCP_TRUE, zz);
    for (int i = 0; i < 3; i++) z[i] = zz[i];
    return (int)k;
}

void ref_eos_fug(const ocpx_fluid* f, double P, double T, const double* x, double* fug,
                 double* phi, double* ABZ)
{
    EoSProbe e(make_param(f));
    e.CalFugPhi(P, T, x, fug, phi);
    if (ABZ) { ABZ[0] = e.A(); ABZ[1] = e.B(); ABZ[2] = e.Z(); }
}

void ref_eos_lnfugx(const ocpx_fluid* f, double P, double T, const double* x, double* out)
{
    EoSProbe e(make_param(f));
    e.CalLnFugX(P, T, x, out);
}

void ref_eos_lnfugn(const ocpx_fluid* f, double P, double T, const double* x, double nt, double* out)
{
    EoSProbe e(make_param(f));
    e.CalLnFugN(P, T, x, nt, out);
}

void ref_eos_lnphin(const ocpx_fluid* f, double P, double T, const double* x, double nt, double* out)
{
    EoSProbe e(make_param(f));
    e.CalLnPhiN(P, T, x, nt, out);
}

void ref_eos_lnfugp(const ocpx_fluid* f, double P, double T, const double* x, double* out)
{
    EoSProbe e(make_param(f));
    e.CalLnFugP(P, T, x, out);
}

double ref_eos_vmder(const ocpx_fluid* f, double P, double T, const double* x, double* vmP, double* vmx)
{
    EoSProbe e(make_param(f));
    double vp = 0;
    double v = e.CalVmDer(P, T, x, vp, vmx);
    *vmP = vp;
    return v;
}

// A fluid as a deck gives it, in the deck's own unit system ("FIELD" or
// "METRIC"), through the reference's EoS set-up (src/OCPEoS.cpp:20-59: the
// Omega defaults, the flat lower-triangle BIC expanded to the symmetric
// matrix).  OmegaA / OmegaB may be null (defaults).  Returns phi, (A, B, Z)
// and the molar volume Z*R*T/P in that unit system; the units are METRIC
// again afterwards.
void ref_eos_deck(const char* units, int nc, const double* Tc, const double* Pc, const double* acf,
                  const double* OmegaA, const double* OmegaB, const double* bic, int nbic,
                  double P, double T, const double* x, double* phi, double* ABZ, double* vm)
{
    SetUnit(units);
    units_set = true;
    ComponentParam p;
    p.numCom = nc;
    p.numPhase = 2;
    auto set = [&](Type_A_r<vector<double>>& t, const double* a) {
        if (!a) return;
        t.activity = 1;
        t.data.assign(1, take(a, nc));
    };
    set(p.Tc, Tc);
    set(p.Pc, Pc);
    set(p.Acf, acf);
    set(p.OmegaA, OmegaA);
    set(p.OmegaB, OmegaB);
    p.BIC.assign(1, take(bic, nbic));
    {
        EoSProbe e(p);
        vector<double> fug(nc);
        e.CalFugPhi(P, T, x, fug.data(), phi);
        ABZ[0] = e.A(); ABZ[1] = e.B(); ABZ[2] = e.Z();
        *vm = e.Z() * GAS_CONSTANT * T / P;
    }
    SetUnit("METRIC");
}

void ref_pvtw(const ocpx_fluid* f, double P, double* o)
{
    metric();
    OCP_PVTW w;
    vector<vector<double>> cols(5);
    for (int k = 0; k < 5; k++) cols[k].assign(1, f->PVTW[k]);
    w.Setup(cols, f->rhoW_std, 1.0);
    w.CalRhoXiMuDer(P, o[0], o[1], o[2], o[3], o[4], o[5]);
}

double ref_lbc(const ocpx_fluid* f, double T, const double* x, double xi, double xiP,
               const double* xix, double* muP, double* mux)
{
    ComponentParam p = make_param(f);
    ViscosityMethod03 v(p, 0);
    double P = 0, muT = 0;
    ViscosityParams vp(&P, &T, x, &xi, &xiP, nullptr, xix);
    return v.CalViscosity(vp, *muP, muT, mux);
}

void ref_kr_pc(const ocpx_sat* s, const double* S, double* kr, double* pc, double* dkrds, double* dpcds)
{
    metric();
    OCPFlowVarSet vs;
    OCPFlowMethod_OGW01 m(columns(s->sgof, s->nsgof), columns(s->swof, s->nswof), 0, vs);
    for (int j = 0; j < 3; j++) vs.S[j] = S[j];
    m.CalKrPcDer(vs);
    for (int j = 0; j < 3; j++) { kr[j] = vs.kr[j]; pc[j] = vs.Pc[j]; }
    for (int k = 0; k < 9; k++) { dkrds[k] = vs.dKrdS[k]; dpcds[k] = vs.dPcdS[k]; }
}

void ref_rock(const ocpx_rock* r, double P, double poroInit, double* out2)
{
    metric();
    RockParam rp;
    rp.Pref = r->Pref;
    rp.cp1 = r->cp1;
    rp.cp2 = r->cp2;
    OCPRockIsoT_Linear rock(rp);
    rock.CalPoro(P, 0.0, poroInit, BulkContent::rf);
    out2[0] = rock.GetPoro();
    out2[1] = rock.GetdPorodP();
}

}  // extern "C"
