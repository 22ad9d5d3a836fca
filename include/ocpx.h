/*
 * ocpx.h — C ABI of the MI355X flash + FIM-assembly hot path.
 *
 * One context = one reservoir region on one GPU, bound to one HIP stream.
 * The context owns a device-resident structure-of-arrays mirror of the
 * reference's BulkVarSet / BulkConnVarSet (include/BulkVarSet.hpp:34-294,
 * include/BulkConnVarSet.hpp:27-172) and a device block-CSR matrix that
 * replaces OCPMatrix (include/OCPMatrix.hpp:26-101).
 *
 * Units follow the reference's METRIC system (include/OCPUnits.hpp:78-100):
 * pressure in bar, temperature in degC in the state arrays (the flash adds
 * 273.15, src/OCPMixtureMethodComp.cpp:674), length in m, permeability in
 * mD, time in days.  Arithmetic is FP64 throughout.
 *
 * Naming: Nc = number of hydrocarbon (Peng-Robinson) components, nc = Nc+1
 * bulk components (water is always the last component), np = 3 phases in
 * the order oil=0, gas=1, water=2, block size b = nc+1.
 *
 * Every entry point returns 0 on success and a negative OCPX_E_* code on
 * failure; ocpx_last_error() then describes it.  Nothing aborts.
 * Calls are asynchronous on the context's stream unless stated otherwise;
 * ocpx_sync() waits.  A context is thread-compatible, not thread-safe.
 */
#ifndef OCPX_H
#define OCPX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OCPX_MAX_NC      16   /* max hydrocarbon components per fluid   */
#define OCPX_MAX_TABROWS 64   /* max rows of a SWOF / SGOF table        */

/* status codes */
#define OCPX_OK            0
#define OCPX_E_ARG        -1   /* bad argument / size                    */
#define OCPX_E_HIP        -2   /* HIP runtime error                      */
#define OCPX_E_STATE      -3   /* call order (e.g. assemble before flash) */
#define OCPX_E_NOMEM      -4
#define OCPX_E_UNSUPPORTED -5

/*
 * Fluid description of one PVT region.  Replaces the ComponentParam +
 * PVTW + flash-control keywords read by ParamReservoir
 * (src/ParamReservoir.cpp; consumed by EoS_PR::EoS_PR src/OCPEoS.cpp:19-65,
 * OCPPhaseEquilibrium::Setup src/OCPPhaseEquilibrium.cpp:15-68,
 * OCPMixtureMethodComp::Setup src/OCPMixtureMethodComp.cpp:25-74,
 * ViscosityMethod03 src/OCPFuncPVT.cpp:415-459,
 * OCPMixtureMethodComp01 ctor src/OCPMixtureMethodComp.cpp:650-668).
 * Arrays are indexed by hydrocarbon component; BIC is full, symmetric,
 * row-major with a fixed row stride of OCPX_MAX_NC.
 */
typedef struct ocpx_fluid {
    int32_t nc;                      /* Nc, 1..OCPX_MAX_NC               */
    int32_t lbc_exact;               /* 0: reference LBC d(mu)/dx, whose
                                        d(sqrt MW) term is 0 through integer
                                        division (src/OCPFuncPVT.cpp:537);
                                        1: include that term (exact tangent) */
    double Tc[OCPX_MAX_NC];          /* K                                */
    double Pc[OCPX_MAX_NC];          /* bar                              */
    double Vc[OCPX_MAX_NC];          /* m3/kmol (phase identification)   */
    double Vcvis[OCPX_MAX_NC];       /* m3/kmol (LBC viscosity)          */
    double MW[OCPX_MAX_NC];
    double acf[OCPX_MAX_NC];         /* acentric factor                  */
    double OmegaA[OCPX_MAX_NC];      /* default 0.457235529              */
    double OmegaB[OCPX_MAX_NC];      /* default 0.077796074              */
    double Vshift[OCPX_MAX_NC];      /* dimensionless deck value; scaled
                                        by R*OmegaB*Tc/Pc internally     */
    double BIC[OCPX_MAX_NC * OCPX_MAX_NC];
    double LBCcoef[5];               /* deck values; x10 internally      */
    double PVTW[5];                  /* Pref, Bw(Pref), Cw, muw(Pref), Cvw */
    double rhoW_std;                 /* water density at std conditions  */
    /* flash control (deck keywords RR, SSMSTA, NRSTA, SSMSP, NRSP) */
    int32_t RR_maxIt, SSMsta_maxIt, NRsta_maxIt, SSMsp_maxIt, NRsp_maxIt;
    int32_t skipPSA;                 /* 1 = SkipPSA on (reference default) */
    double RR_tol, SSMsta_tol, SSMsta_eYt, NRsta_tol, SSMsp_tol, NRsp_tol;
} ocpx_fluid;

/*
 * Saturation tables for OCPFlowMethod_OGW01 (src/OCPFlowMethod.cpp:19-73):
 * SWOF rows (Sw, krw, krow, Pcow) and SGOF rows (Sg, krg, krog, Pcgo),
 * row-major, 4 doubles per row.  Oil kr by Stone II
 * (OCP3POilPerMethod01, src/OCPFuncSAT.cpp:149-186).
 */
typedef struct ocpx_sat {
    int32_t nswof, nsgof;
    double swof[4 * OCPX_MAX_TABROWS];
    double sgof[4 * OCPX_MAX_TABROWS];
} ocpx_sat;

/* OCPRockIsoT_Linear (src/OCPRock.cpp:19-24) */
typedef struct ocpx_rock {
    double Pref, cp1, cp2;
} ocpx_rock;

/*
 * Cells and connections.  Connections follow BulkConn::iteratorConn
 * (src/Reservoir.cpp:456-531): bId < eId, sorted by bId, and per cell in
 * x+, y+, z+ order, so that each block row's columns come out ascending
 * with the diagonal block first (OCPMatrix push order,
 * src/OCPMatrix.cpp:92-137).  trans already includes CONV_DARCY
 * (BulkConnTransMethod01::CalTrans, src/BulkConnFunc.cpp:15-67).
 * The arrays are copied at ocpx_ctx_create.
 */
typedef struct ocpx_grid {
    int64_t nb;                      /* cells                            */
    int64_t nconn;                   /* connections (may be 0)           */
    const double*  v;                /* bulk volume [nb]                 */
    const double*  depth;            /* cell-centre depth [nb]           */
    const double*  poroInit;         /* [nb]                             */
    const int32_t* conn_b;           /* [nconn]                          */
    const int32_t* conn_e;           /* [nconn]                          */
    const double*  trans;            /* [nconn]                          */
} ocpx_grid;

/*
 * Peaceman wells (PeacemanWell / PeacemanWellIsoT, src/WellPeaceman.cpp).
 * One well unknown (the bottom-hole pressure) per open well, as the
 * reference's FIM well rows.  Perforations of well w are entries
 * [perf0, perf0 + nperf) of the perforation arrays of ocpx_set_wells.
 * The per-time-step quantities the reference computes with extra flashes
 * are inputs here: the perforation heads dG (CaldG), the injected fluid's
 * molar density at each perforation (XiPhase in CalFlux) and, in rate
 * mode, the component weights of the rate equation (CalFactor).
 */
enum ocpx_well_type { OCPX_WELL_INJ = 0, OCPX_WELL_PROD = 1 };
enum ocpx_well_mode { OCPX_WELL_BHP = 0, OCPX_WELL_RATE = 1 };

typedef struct ocpx_well {
    int32_t type;                    /* ocpx_well_type                   */
    int32_t mode;                    /* ocpx_well_mode                   */
    int32_t perf0, nperf;            /* perforation range                */
    double  bhp;                     /* bottom-hole pressure, bar        */
    double  target;                  /* tarBHP (bar) or tarRate          */
    double  zi[OCPX_MAX_NC + 1];     /* injected composition (nc = Nc+1) */
    double  factor[OCPX_MAX_NC + 1]; /* rate weight per component        */
} ocpx_well;

/* Residual norms of IsoT_FIM::CalRes (src/IsoThermalMethod.cpp:1145-1168) */
typedef struct ocpx_resnorms {
    double maxRelRes_V;   int64_t maxId_V;
    double maxRelRes_N;   int64_t maxId_N;
} ocpx_resnorms;

/* Per-cell state fields, exchanged in the reference's BulkVarSet layout
 * (cell-major: field[n*len + k]).  len is given in the comment. */
enum ocpx_field {
    OCPX_F_P = 0,         /* [1]   bvs.P                                 */
    OCPX_F_T,             /* [1]   bvs.T (degC)                          */
    OCPX_F_NI,            /* [nc]  bvs.Ni                                */
    OCPX_F_LNI,           /* [nc]  bvs.lNi (Ni at the last time step)    */
    OCPX_F_PHASEEXIST,    /* [np]  uint32 (OCP_BOOL)                     */
    OCPX_F_S,             /* [np]                                        */
    OCPX_F_XIJ,           /* [np*nc]                                     */
    OCPX_F_RHO,           /* [np]                                        */
    OCPX_F_XI,            /* [np]                                        */
    OCPX_F_MU,            /* [np]                                        */
    OCPX_F_RHOP,          /* [np]                                        */
    OCPX_F_XIP,           /* [np]                                        */
    OCPX_F_MUP,           /* [np]                                        */
    OCPX_F_RHOX,          /* [np*nc]                                     */
    OCPX_F_XIX,           /* [np*nc]                                     */
    OCPX_F_MUX,           /* [np*nc]                                     */
    OCPX_F_NT,            /* [1]                                         */
    OCPX_F_VF,            /* [1]                                         */
    OCPX_F_VFP,           /* [1]                                         */
    OCPX_F_VFI,           /* [nc]                                        */
    OCPX_F_DSEC_DPRI,     /* [np*(nc+1)*(nc+1)]  dXsdXp, row-major       */
    OCPX_F_KR,            /* [np]                                        */
    OCPX_F_PC,            /* [np]                                        */
    OCPX_F_PJ,            /* [np]                                        */
    OCPX_F_DKRDS,         /* [np*np]                                     */
    OCPX_F_DPCDS,         /* [np*np]                                     */
    OCPX_F_PORO,          /* [1]                                         */
    OCPX_F_POROP,         /* [1]                                         */
    OCPX_F_ROCKVP,        /* [1]                                         */
    OCPX_F_FTYPE_IN,      /* [1]   int32; -1 = let SkipPSA decide        */
    OCPX_F_FTYPE_OUT,     /* [1]   int32; PE ftype after the flash       */
    OCPX_F_SKIP_FLAG,     /* [1]   uint32 SkipPSAVarset::flag            */
    OCPX_F_SKIP_MINEIG,   /* [1]   SkipPSAVarset::minEigen               */
    OCPX_F_SKIP_P,        /* [1]                                         */
    OCPX_F_SKIP_T,        /* [1]                                         */
    OCPX_F_SKIP_ZI,       /* [Nc]                                        */
    OCPX_F_ITERS,         /* [8]   int32 per-cell counters, see below    */
    OCPX_F_NFIELDS
};

/* OCPX_F_ITERS layout: SSMsta iters, NRsta iters, SSMsp iters, NRsp iters,
 * RR iters, fugacity evaluations, ln-fugacity-derivative evaluations,
 * dense solves. */

typedef struct ocpx_ctx ocpx_ctx;

/* Per-cell entry size (doubles or 4-byte ints) of a field for a fluid with
 * Nc hydrocarbon components. */
int64_t ocpx_field_len(int field, int32_t Nc);

/* Create a context (≙ IsoT_FIM::Setup/AllocateReservoir,
 * src/IsoThermalMethod.cpp:711-715, 905-998).  stream may be NULL (default
 * stream).  Allocates all device memory up front. */
int  ocpx_ctx_create(const ocpx_fluid* fluid, const ocpx_sat* sat,
                     const ocpx_rock* rock, const ocpx_grid* grid,
                     void* hip_stream, ocpx_ctx** out);
/* Several independent models of one component count Nc in one context (no
 * counterpart in the reference, which runs one reservoir per process):
 * fluid, sat, rock and grids are arrays of nmodel; the models' cells and
 * connections are concatenated in model order (cell indices of model m are
 * offset by the cells of models 0..m-1), each cell keeps its own model's
 * fluid, tables and rock.  Fields, BSR and residuals are exchanged for the
 * concatenated cells.  The flash runs in OCPX_FLASH_BATCHED mode with a team
 * mapping. */
int  ocpx_ctx_create_multi(int32_t nmodel, const ocpx_fluid* fluid, const ocpx_sat* sat,
                           const ocpx_rock* rock, const ocpx_grid* grids, void* stream,
                           ocpx_ctx** out);
void ocpx_ctx_destroy(ocpx_ctx* ctx);
/* The context's last error; with ctx == NULL, why the last ocpx_ctx_create or
 * ocpx_ctx_create_multi on this thread failed (a grid needs at least one
 * cell). */
const char* ocpx_last_error(const ocpx_ctx* ctx);
int  ocpx_sync(ocpx_ctx* ctx);

/* Host <-> device state exchange in BulkVarSet layout (synchronous). */
int  ocpx_set_field(ocpx_ctx* ctx, int field, const void* host);
int  ocpx_get_field(ocpx_ctx* ctx, int field, void* host);
/* Device pointer of a field (component-major SoA: f[k*nb + n]). */
int  ocpx_field_device_ptr(ocpx_ctx* ctx, int field, void** dptr);

/* Flash modes */
#define OCPX_FLASH_PLAIN    0   /* one cell per lane, one launch          */
#define OCPX_FLASH_BATCHED  1   /* persistent kernel with work queue      */

/* ≙ IsoT_FIM::CalFlash + PassFlashValue (src/IsoThermalMethod.cpp:1016-1071):
 * SkipPSA::CalFtype02 (src/AcceleratePEC.cpp:103-124) unless FTYPE_IN >= 0,
 * OCPMixtureMethodComp01::FlashDer (src/OCPMixtureMethodComp.cpp:746-755),
 * SkipPSA::CalSkipForNextStep (src/AcceleratePEC.cpp:127-172). */
int  ocpx_flash_der(ocpx_ctx* ctx, int mode);
/* Phase equilibrium only (OCPPhaseEquilibrium::PhaseEquilibrium,
 * src/OCPPhaseEquilibrium.cpp:71-112), for the batched-flash throughput
 * workload.  Results: OCPX_F_FTYPE_OUT, ITERS and the split below. */
int  ocpx_phase_equilibrium(ocpx_ctx* ctx, int mode);
/* NP [nb] int32, nu [nb*2], x [nb*2*Nc] (PE order, before relabelling). */
int  ocpx_get_pe_result(ocpx_ctx* ctx, int32_t* NP, double* nu, double* x);
/* Phase equilibrium of n cells that allocates nothing.  No counterpart in the
 * reference; it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It runs the kernel of ocpx_phase_equilibrium's batched mode
 * with one cell per lane (no stability stages) on every cell from one phase
 * (OCPPhaseEquilibrium::PhaseEquilibrium(P, T, z) with ftype 0,
 * src/OCPPhaseEquilibrium.cpp:71-112), on device arrays the caller owns:
 * P [n] bar, T [n] degC, z [Nc][n] (component-major, normalised by the
 * kernel); outputs NP [n], nu [n] (nu of PE phase 0; phase 1 has 1 - nu),
 * x [2*Nc][n] (phase 0's Nc fractions, then phase 1's) and ftype [n],
 * device-side and component-major.  Its scratch (the work-queue counter and the fluid
 * constants) is a fixed 8 KiB device array per GPU, held until the process
 * ends; calls that overlap on one GPU race.  Limits: Nc 2..12, n up to
 * 2^31 - 1 cells; anything else returns OCPX_E_ARG, and ocpx_last_error(NULL)
 * says why.  Asynchronous on hip_stream (NULL: the default stream). */
int  ocpx_phase_equilibrium_noalloc(const ocpx_fluid* fluid, int64_t n, const double* P, const double* T,
                                    const double* z, int32_t* NP, double* nu, double* x, int32_t* ftype,
                                    void* hip_stream);

/* Cell-to-lane mapping of the batched flash's phase-equilibrium kernel.  No
 * counterpart in the reference (its flash is a serial loop over cells,
 * src/IsoThermalMethod.cpp:1016-1071); same results either way.
 * lanes_per_cell: 1 = one cell per lane; 4, 8 or 16 = a team of that many
 * lanes per cell, lane t holding components t, t+T, ... (4 for any Nc, 8 for
 * Nc >= 5, 16 for Nc >= 9); 0 = default (the autotuner's choice if it ran,
 * else 1 for Nc <= 4 and a team of 8 or 16 above; contexts of several models
 * take 4 for Nc <= 4).  OCPX_FLASH_PLAIN is always one cell per lane. */
int  ocpx_set_flash_mapping(ocpx_ctx* ctx, int lanes_per_cell);
int  ocpx_get_flash_mapping(ocpx_ctx* ctx, int32_t* lanes_per_cell);
/* Autotuner: runs the batched phase-equilibrium kernel on the current state
 * once per mapping built for Nc (1, 4, 8, 16 lanes per cell) with the
 * longest-first queue, keeps the fastest as the context's mapping, then runs
 * that mapping with the cells in natural order and keeps the faster queue
 * order.  ms[5] (may be NULL) receives the kernel times: the four mappings
 * (0 where one does not apply), then the chosen one in natural order.
 * Synchronous. */
int  ocpx_autotune_flash(ocpx_ctx* ctx, double* ms);
/* Stages of the batched flash.  No counterpart in the reference; the same
 * algorithm either way (sums may round differently).  In a context of one
 * model, 1: the strict stability test (StableSSM + StableNR,
 * src/OCPPhaseEquilibrium.cpp:397-521) runs in its own stages -- the SSM
 * loops of every cell's trial compositions in a lean kernel, the Newton steps
 * of the trials that need one in a second, then per cell the first unstable
 * trial in the reference's order -- and the mapped kernel of
 * ocpx_set_flash_mapping takes only the cells left (splits, flash types 1
 * and 2, the relaxed test); 2: as 1, and the first SplitSSM2 pass of the
 * unstable cells (:586-649) runs in a lean kernel of its own, the mapped
 * kernel resuming them at SplitNR (:807-823); 3 (default): as 2, and SplitNR
 * with CheckSplit (:805-1010) in a lean kernel too, the mapped kernel taking
 * only the cells whose Newton steps do not converge, at the restart loop
 * (:558-566); 0: the mapped kernel takes every cell.  In a context of several
 * models every stage reads each cell's own model's fluid. */
int  ocpx_set_flash_stages(ocpx_ctx* ctx, int on);
int  ocpx_get_flash_stages(ocpx_ctx* ctx, int32_t* on);
/* Long cells of the staged batched flash.  No counterpart in the reference;
 * the same algorithm either way.  Cells whose previous batched flash took more
 * than `evals` fugacity evaluations (ITERS row 5; default 1000, 0 = off) are
 * flashed from the start by the mapped kernel on a branch of their own beside
 * the stages (a second stream forked from and joined to the context's,
 * inside the context's own captured graph too; off inside a batch graph,
 * ocpx_batch_newton_pass): their serial Newton and restart work, not the
 * bulk, is the flash's critical path.  Such a cell runs the mapped kernel's
 * arithmetic instead of the stages', so its last bits follow the counts in
 * OCPX_F_ITERS when the flash starts; those are saved and restored with the
 * last time step, so a pass repeated after ocpx_reset_to_last_step repeats
 * its bits. */
int  ocpx_set_flash_long(ocpx_ctx* ctx, int32_t evals);
int  ocpx_get_flash_long(ocpx_ctx* ctx, int32_t* evals);
/* Queue order of the batched flash.  No counterpart in the reference (its
 * flash loops over cells in order); same results either way.  1 (default):
 * each batched flash hands out the cells longest first, by their fugacity-
 * evaluation counts in the previous batched flash (ITERS), so cells with the
 * same count share wavefronts and the longest start first; 0: natural order. */
int  ocpx_set_flash_queue_order(ocpx_ctx* ctx, int longest_first);
int  ocpx_get_flash_queue_order(ocpx_ctx* ctx, int32_t* longest_first);

/* ≙ IsoT_FIM::CalKrPc (src/IsoThermalMethod.cpp:1073-1088) */
int  ocpx_krpc_der(ocpx_ctx* ctx);
/* ≙ IsothermalMethod::CalRock (src/IsoThermalMethod.cpp:19-30) */
int  ocpx_cal_rock(ocpx_ctx* ctx);
/* ≙ IsoT_FIM::CalRes (src/IsoThermalMethod.cpp:1090-1185), bulk part (the
 * wells' part is ocpx_well_terms).  Keeps resAbs (negated, [(nb)*(nc+1)])
 * on the device; norms and resAbs are copied to the host when the pointers
 * are non-NULL (synchronous then). */
int  ocpx_cal_res(ocpx_ctx* ctx, double dt, double* resAbs, ocpx_resnorms* norms);
/* ≙ IsoT_FIM::AssembleMatBulks + LinearSystem::CopyRhs
 * (src/IsoThermalMethod.cpp:1187-1269, 742-752) into the device BSR. */
int  ocpx_assemble(ocpx_ctx* ctx, double dt);
/* BSR pattern and values (host copies).  rowptr [nb+1], colind [nnzb],
 * val [nnzb*b*b] (each block row-major), rhs [nb*b].  Any pointer may be
 * NULL.  nnzb = nb + 2*nconn. */
int  ocpx_get_bsr(ocpx_ctx* ctx, int64_t* nnzb, int64_t* rowptr,
                  int32_t* colind, double* val, double* rhs);
/* Device pointers of the BSR arrays (valid for the context lifetime). */
int  ocpx_bsr_device_ptrs(ocpx_ctx* ctx, void** rowptr, void** colind,
                          void** val, void** rhs);

/* ≙ IsoT_FIM::GetSolution (src/IsoThermalMethod.cpp:1282-1418) without
 * wells or halos: dSmax chop, S/xij/P/Ni update.  u is a host array
 * [nb*(nc+1)], or NULL for the device solution of the last
 * ocpx_linear_solve (no copy). */
int  ocpx_get_solution(ocpx_ctx* ctx, const double* u, double dSmax);

/* ---- Peaceman wells ------------------------------------------------------
 * ocpx_set_wells: the wells and their perforations (perf_cell: cell index,
 * perf_WI: WI * multiplier as PeacemanWell::CalWI leaves it, CONV_DARCY
 * included; perf_dG: perforation head of CaldG, bar; perf_xi: molar density
 * of the injected fluid at the perforation, injectors only, may be NULL
 * without injectors).  A well perforates a cell at most once.  Replaces any
 * earlier set; nwell = 0 removes the wells.  Copies the arrays.
 * (PeacemanWell::InputPerfo / Setup, src/WellPeaceman.cpp:15-105.) */
int  ocpx_set_wells(ocpx_ctx* ctx, int32_t nwell, const ocpx_well* wells, const int64_t* perf_cell,
                    const double* perf_WI, const double* perf_dG, const double* perf_xi);
/* ocpx_well_terms: the wells' part of IsoT_FIM::CalRes and AssembleMat
 * (PeacemanWell::CalTrans + CalFlux, PeacemanWellIsoT::CalResFIM and
 * AssembleMatInjFIM / AssembleMatProdFIM, src/WellPeaceman.cpp:365-462,
 * 1226-1612) on the current state.  Call it once after ocpx_cal_res
 * (jacobian = 0) or ocpx_assemble (jacobian = 1) with the same dt; a second
 * call on the same residual, or a call after the state changed (flash,
 * ocpx_get_solution, ocpx_set_field, reset, geometry), returns OCPX_E_STATE
 * instead of adding the wells twice or to a stale system.  It adds
 * the perforation rates to the bulk residual and right-hand side and
 * recomputes the norms (norms may be NULL); with jacobian = 1 it adds the
 * perforation terms to the bulk diagonal blocks and forms the well blocks
 * (ocpx_get_well_blocks).  well_res [nwell] (may be NULL): each well row's
 * residual, bhp - tarBHP or +-tarRate + sum_i q_i factor_i; the
 * right-hand side of a well row is its negative. */
int  ocpx_well_terms(ocpx_ctx* ctx, double dt, int32_t jacobian, ocpx_resnorms* norms, double* well_res);
/* The blocks the reference's OCPMatrix gets from the wells beyond the bulk
 * diagonal: Bnw [nperf][b] column 0 of block (perforated cell, well),
 * Cwn [nperf][b] row 0 of block (well, perforated cell), Dw [nwell][b][b]
 * the well's diagonal block.  The other entries of those blocks are zero.
 * In OCPMatrix push order a perforated row's well columns follow its bulk
 * columns, and a well row holds its diagonal and then its perforated cells
 * in perforation order (well rows after the nb bulk rows). */
int  ocpx_get_well_blocks(ocpx_ctx* ctx, double* Bnw, double* Cwn, double* Dw);
/* bhp [nwell] and the component rates qw [nwell][nc] of the last
 * ocpx_well_terms (either may be NULL); ocpx_set_well_bhp sets the
 * bottom-hole pressures (PeacemanWellIsoT::GetSolutionFIM adds u of the
 * well row to bhp, src/WellPeaceman.cpp:1198-1204). */
int  ocpx_get_wells(ocpx_ctx* ctx, double* bhp, double* qw);
int  ocpx_set_well_bhp(ocpx_ctx* ctx, const double* bhp);
/* Peaceman well index of one perforation, PeacemanWell::CalWI
 * (src/WellPeaceman.cpp:277-362): dir 0/1/2 = x/y/z, dz already times NTG,
 * kh < 0 for the geometric kh; METRIC units, CONV_DARCY included.  Host
 * function, no context.  Returns -1 for an unknown direction. */
double ocpx_peaceman_wi(int32_t dir, double dx, double dy, double dz, double kx, double ky, double kz,
                        double radius, double skin, double kh);
/* Molar density of a single-phase fluid, the injected fluid's perf_xi
 * (MixtureUnit::XiPhase -> OCPMixtureMethodComp01::CalXi,
 * src/OCPMixtureMethodComp.cpp:775-786; evaluated per time step, as
 * PeacemanWell::CalFluxInit does): water = 1 / Bw(P) from PVTW, otherwise
 * 1 / Vm(P, T + 273.15, z) of the Peng-Robinson fluid (z over the Nc
 * hydrocarbon components).  P in bar, T in degC.  Host function, no
 * context.  Returns -1 on a bad argument. */
double ocpx_xi_phase(const ocpx_fluid* fluid, double P, double T, const double* z, int32_t water);

/* ≙ LinearSystem::AssembleMatLinearSolver + Solve (src/LinearSystem.cpp:96-119)
 * as IsoT_FIM::SolveLinearSystem calls them (src/IsoThermalMethod.cpp:754-
 * 768), on the device BSR of ocpx_assemble: J u = rhs (rhs = -R, CopyRhs).
 * The method is ocpx_set_linear_solver's (default: ABF decoupling + FGMRES(30)
 * with a CPR-type preconditioner, as the reference's FASP set-up) in place
 * of the reference's external FASP/PETSc solvers; stops at ||r|| / ||b|| <
 * rtol (the reference's stop_type 1, on the decoupled system as FASP does;
 * its decks use 1e-3 and 200 iterations, data/spe5/bsr.fasp).  *iters = iterations, negative when
 * not converged (the reference's convention); *relres = final ||r||/||b||.
 * u (host [nb*(nc+1)], may be NULL) receives the last iterate; only a
 * converged solution stays on the device for ocpx_get_solution(ctx, NULL,
 * dSmax), which returns OCPX_E_STATE after a solve that did not converge
 * (the reference never applies such a solution: it cuts dt and resets,
 * src/IsoThermalMethod.cpp:776-781).  When
 * ocpx_well_terms(jacobian = 1) ran after the last residual / assembly pass
 * the system has the well rows too (AddDim, src/WellPeaceman.cpp:1285): u is
 * then [(nb + nwell)*(nc+1)], and ocpx_get_solution(ctx, NULL, dSmax) also
 * adds each well row's entry 0 to its bhp (GetSolutionFIM).  Synchronous. */
int  ocpx_linear_solve(ocpx_ctx* ctx, double rtol, int32_t maxit, double* u, int32_t* iters,
                       double* relres);
/* Method of ocpx_linear_solve.  1 (default), as the reference's FASP set-up
 * (data/spe5/bsr.fasp: decoup_type 1, solver_type 6, restart 30): the system
 * is decoupled by ABF (each block row times the inverse of its diagonal
 * block, src/Decoupling.cpp:23-57) and solved by FGMRES(30), right-
 * preconditioned by a CPR-type two-stage preconditioner -- one V-cycle of
 * an aggregation AMG on the decoupled pressure block, then one multicolour
 * block Gauss-Seidel sweep -- stopping on the decoupled ||r|| / ||b||.
 * 0: BiCGStab with block Jacobi on the system as assembled. */
int  ocpx_set_linear_solver(ocpx_ctx* ctx, int method);
int  ocpx_get_linear_solver(ocpx_ctx* ctx, int32_t* method);

/* ≙ IsoT_FIM::UpdateLastTimeStep (src/IsoThermalMethod.cpp:1473-1515) with
 * SkipPSAVarset::UpdateLastTimeStep (src/AcceleratePEC.cpp:24-31): keeps a
 * device copy of the bulk state (rock, Nt, vf, P, Pj, Pc, phaseExist, S,
 * xij, rho, xi, mu, kr, their derivatives, dSec_dPri, the SkipPSA data, and
 * the flash counters OCPX_F_ITERS, which pick the long cells of the next
 * batched flash) and sets lNi (OCPX_F_LNI) = Ni.  Device-to-device, on the
 * context stream. */
int  ocpx_update_last_step(ocpx_ctx* ctx);
/* ≙ IsoT_FIM::ResetToLastTimeStep (:1420-1464) with
 * SkipPSAVarset::ResetToLastTimeStep (src/AcceleratePEC.cpp:15-22):
 * restores that copy and Ni = lNi (no wells, no residual recomputation).
 * OCPX_E_STATE before the first ocpx_update_last_step. */
int  ocpx_reset_to_last_step(ocpx_ctx* ctx);

/* Replace the bulk volumes [nb] and/or the connection transmissibilities
 * [nconn] (host arrays; NULL keeps the current values).  The reference sets
 * them once from the grid (Bulk::v, BulkConnTransMethod01::CalTrans,
 * src/BulkConnFunc.cpp:15-67); here they are also the parameters of the
 * adjoint gradients (ocpx_dR_dparams). */
int  ocpx_set_geometry(ocpx_ctx* ctx, const double* v, const double* trans);

/* One full Newton-step pass: flash_der + krpc + rock + res + assemble.
 * With use_graph != 0 the pass is captured once into a hipGraph and
 * replayed afterwards (dt is baked in at capture; re-captured if dt or
 * mode changes). */
int  ocpx_newton_pass(ocpx_ctx* ctx, double dt, int mode, int use_graph);

/* A batch of contexts -- for example one ocpx_ctx_create_multi context per
 * component count -- whose Newton passes run as one graph launch.  No
 * counterpart in the reference (one simulator instance per model).
 * ocpx_batch_newton_pass(batch, dt) is ocpx_newton_pass(ctx, dt,
 * OCPX_FLASH_BATCHED, 1) for every context at once: each context's pass is
 * a branch of one captured hipGraph (forked from and joined to the batch's
 * stream, so the contexts' kernels may overlap), launched once; it is
 * ordered after the work already queued on each context's stream, and each
 * context's later work after it.  Captured at the first call, again when dt
 * or a context's flash settings change.  Contexts stay owned by the caller
 * and must outlive the batch; hip_stream NULL: the batch makes its own.
 * Errors: ocpx_last_error(NULL). */
typedef struct ocpx_batch ocpx_batch;
int  ocpx_batch_create(ocpx_ctx* const* ctxs, int32_t n, void* hip_stream, ocpx_batch** out);
int  ocpx_batch_newton_pass(ocpx_batch* batch, double dt);
void ocpx_batch_destroy(ocpx_batch* batch);

/* Differentiation of the assembled system.
 * ocpx_jt_lambda: out = J^T lambda, J the BSR assembled last
 * (lambda, out: host [nb*(nc+1)]).
 * ocpx_dR_dparams: gradients of lambda . R with respect to the connection
 * transmissibilities [nconn] and the cell rock pore volumes [nb]
 * (∂R/∂T from OCPConvection01::CalFlux, src/OCPConvection.cpp:22-79;
 * ∂R/∂Vp from BulkAccumuTerm01::CalResFIM, src/BulkAccumuModule.cpp:24-44),
 * sign convention of the un-negated residual R = accumulation + dt*flux. */
int  ocpx_jt_lambda(ocpx_ctx* ctx, const double* lambda, double* out);
int  ocpx_dR_dparams(ocpx_ctx* ctx, double dt, const double* lambda,
                     double* dTrans, double* dVp);
/* Device-pointer variants of the two adjoint entry points (arrays of the
 * same sizes in device memory; asynchronous on the context stream). */
int  ocpx_jt_lambda_dev(ocpx_ctx* ctx, const double* lambda_d, double* out_d);
int  ocpx_dR_dparams_dev(ocpx_ctx* ctx, double dt, const double* lambda_d,
                         double* dTrans_d, double* dVp_d);

/* Kernel timing: average duration (ms) of the kernels of one entry point
 * over the calls since the last reset, from HIP events recorded on the
 * context stream.  which: 0 PE, 1 flash tail, 2 krpc+rock, 3 residual,
 * 4 assembly, 5 J^T lambda, 6 dR/dparams, 7 linear solve (whole, host
 * round trips included), 8 BSR matrix-vector product inside the solve. */
int  ocpx_timing_enable(ocpx_ctx* ctx, int on);
int  ocpx_timing_get(ocpx_ctx* ctx, int which, double* avg_ms, int64_t* calls);

#ifdef __cplusplus
}
#endif
#endif /* OCPX_H */
