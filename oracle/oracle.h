/*
 * oracle.h — CPU restatement of the reference's flash + FIM-assembly path.
 *
 * TEST INFRASTRUCTURE ONLY.  The functions declared here are the checker for
 * the HIP implementation: only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load this library.  The product path
 * (opencaeporox_amd/, libocpx.so) never links or calls it.
 *
 * Layouts are the reference's BulkVarSet layouts (cell-major), see
 * include/ocpx.h for the field list.
 */
#ifndef OCPX_ORACLE_H
#define OCPX_ORACLE_H

#include <stdint.h>
#include "../include/ocpx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Host arrays in BulkVarSet layout.  nb cells, np = 3, nc = Nc + 1. */
typedef struct orc_bulk {
    int64_t nb;
    double *P, *T, *Ni, *lNi, *S, *xij, *rho, *xi, *mu, *rhoP, *xiP, *muP;
    double *rhox, *xix, *mux, *Nt, *vf, *vfP, *vfi, *dSec_dPri;
    double *kr, *Pc, *Pj, *dKrdS, *dPcdS, *poro, *poroP, *rockVp;
    double *v, *poroInit, *depth;
    uint32_t *phaseExist;
    int32_t *ftype_in;      /* may be NULL; entries < 0 use SkipPSA   */
    int32_t *ftype_out;     /* may be NULL                            */
    uint32_t *skip_flag;
    double *skip_minEig, *skip_P, *skip_T, *skip_zi;
    int32_t *iters;         /* [nb*8] may be NULL                     */
    int32_t *pe_ill;        /* [nb] ill-conditioned path markers, may be NULL */
} orc_bulk;

/* --- EoS_PR (src/OCPEoS.cpp) single-composition probes ----------------- */
int orc_cubic_root(double a, double b, double c, double* z);  /* returns count */
void orc_eos_fug(const ocpx_fluid* f, double P, double T, const double* x,
                 double* fug, double* phi, double* ABZ /* Aj,Bj,Zj */);
void orc_eos_lnfugx(const ocpx_fluid* f, double P, double T, const double* x, double* out);
void orc_eos_lnfugn(const ocpx_fluid* f, double P, double T, const double* x, double nt, double* out);
void orc_eos_lnfugp(const ocpx_fluid* f, double P, double T, const double* x, double* out);
void orc_eos_lnphin(const ocpx_fluid* f, double P, double T, const double* x, double nt, double* out);
double orc_eos_vmder(const ocpx_fluid* f, double P, double T, const double* x, double* vmP, double* vmx);

/* --- Phase equilibrium (src/OCPPhaseEquilibrium.cpp:71-112) -------------
 * For n cells with T in K and z normalised: NP [n], nu [n*2], x [n*2*Nc],
 * ftype_out [n], iters [n*8], ill [n] (ill-conditioned path markers, see
 * pe_t in ocpx_oracle.c).  ftype [n], lNP [n], lx [n*2*nc] (bulk nc stride
 * as in OCPMixtureVarSet::x) may be NULL -> 0 / 0 / unused. */
void orc_phase_equilibrium(const ocpx_fluid* f, int64_t n, const double* P,
                           const double* T, const double* z, const int32_t* ftype,
                           const int32_t* lNP, const double* lx, int32_t* NP,
                           double* nu, double* x, int32_t* ftype_out, int32_t* iters,
                           int32_t* ill);

/* --- Viscosity LBC (src/OCPFuncPVT.cpp:462-575) and PVTW (:32-57) ------ */
double orc_lbc(const ocpx_fluid* f, double T, const double* x, double xi, double xiP,
               const double* xix, double* muP, double* mux);
void orc_pvtw(const ocpx_fluid* f, double P, double* out6 /* rho xi mu rhoP xiP muP */);

/* --- FIM loops (src/IsoThermalMethod.cpp:1016-1269) -------------------- */
void orc_cal_flash(const ocpx_fluid* f, orc_bulk* b);
void orc_cal_krpc(const ocpx_sat* s, orc_bulk* b);
void orc_kr_pc(const ocpx_sat* s, const double* S, double* kr, double* pc,
               double* dkrds, double* dpcds);
void orc_cal_rock(const ocpx_rock* r, orc_bulk* b);
/* trans, conn arrays as in ocpx_grid; resAbs [nb*(nc+1)] negated;
 * norms [4] = maxRelRes_V, maxId_V, maxRelRes_N, maxId_N. */
void orc_cal_res(const ocpx_fluid* f, const orc_bulk* b, int64_t nconn,
                 const int32_t* cb, const int32_t* ce, const double* trans,
                 double dt, double* resAbs, double* norms);
/* Row-segmented BSR in OCPMatrix push order, flattened:
 * rowptr [nb+1], colind [nb+2*nconn], val [(nb+2*nconn)*bs*bs]. */
void orc_assemble(const ocpx_fluid* f, const orc_bulk* b, int64_t nconn,
                  const int32_t* cb, const int32_t* ce, const double* trans,
                  double dt, int64_t* rowptr, int32_t* colind, double* val);
/* GetSolution (src/IsoThermalMethod.cpp:1330-1400), no wells, no halos */
void orc_get_solution(const ocpx_fluid* f, orc_bulk* b, const double* u, double dSmax);

/* Cartesian TPFA connections (src/PreParamGridWell.cpp:1140-1214 order,
 * BulkConnTransMethod01::CalTrans src/BulkConnFunc.cpp:15-67).
 * Returns the connection count; arrays sized 3*nx*ny*nz. */
int64_t orc_cartesian_conns(int nx, int ny, int nz, const double* dx, const double* dy,
                            const double* dz, const double* kx, const double* ky,
                            const double* kz, const double* ntg, int32_t* cb,
                            int32_t* ce, double* trans);

/* --- Peaceman wells (src/WellPeaceman.cpp), see ocpx_oracle.c ---------- */
void orc_wells(const ocpx_fluid* fin, const orc_bulk* b, double dt, int32_t nw, const ocpx_well* w,
               const int64_t* pcell, const double* pWI, const double* pdG, const double* pxi,
               double* res, double* norms, double* wres, double* qw,
               const int64_t* rowptr, double* val, double* Bnw, double* Cwn, double* Dw);
double orc_peaceman_wi(int dir, double dx, double dy, double dz, double kx, double ky, double kz,
                       double radius, double skin, double kh);

#ifdef __cplusplus
}
#endif
#endif
