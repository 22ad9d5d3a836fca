// ocpx_kernels.hpp — kernel argument blocks and launchers shared by the
// HIP translation units and the host code (ocpx_host.cpp).
//
// Every per-cell array is component-major structure-of-arrays:
// field[k * nb + cell].
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/ocpx.h"

namespace ocpx {

struct DevFluid;
struct DevSat;

// phase-equilibrium kernel (pe.hip)
struct PeIO {
    int64_t nb;
    const double* P;            // [nb]
    const double* T;            // [nb] degC
    const double* Ni;           // [nc][nb]
    const uint32_t* phaseExist; // [3][nb]  previous step, or null (one phase: the strict test)
    const double* S;            // [3][nb]  previous step
    const double* xij;          // [3*nc][nb] previous step
    const int32_t* ftype_in;    // [nb] or null; < 0 -> SkipPSA
    const uint32_t* skip_flag;  // SkipPSAVarset
    const double* skip_minEig;
    const double* skip_P;
    const double* skip_T;
    const double* skip_zi;      // [Nc][nb]
    int32_t* pe_NP;             // [nb]
    double* pe_nu;              // [nb] nu of PE phase 0
    double* pe_x;               // [2*Nc][nb]
    int32_t* ftype_out;         // [nb]
    int32_t* iters;             // [8][nb] or null
    unsigned* queue;            // work-queue counter (batched mode)
    const int32_t* qorder;      // [nb] cell handed out at each queue position (batched mode), or null
    const int32_t* mid;         // [nb] model of each cell (multi-model contexts), or null
    const unsigned* nq;         // queue length on the device (batched mode), or null (nb)
    // cells the stability stages found unstable start at the split (route 2)
    // with the deciding trial's Yt and counters, or at SplitNR (route 3) from
    // what the split stage left (pe_sta.hip); null: no stages
    const int32_t* sta_route;   // [nb]
    const double* rs_Yt;        // [nb]
    const int32_t* rs_cnt;      // [8][nb]: ITERS rows 0, 1, 5, 6, 2, 4; SSM converged; row 3
                                // (row 7 = row 6 - row 3)
    const double* rs_v;         // [5][Nc][nb]: x0, x1, K, ln phi(x0), ln phi(x1)
    const double* rs_nu;        // [nb]
};

// device fluid constants of one PVT region (ocpx_host.cpp)
void build_dev_fluid(const ocpx_fluid* fl, DevFluid& h);
// resident 64-lane blocks of pe_batched per compute unit (pe.hip)
int pe_batched_blocks_per_cu(int NC);

// stability stages of the batched flash (pe_sta.hip)
struct StaIO {
    unsigned* qhead;            // counters, consecutive words: stage-1 queue head,
    unsigned* gcount;           //   cells for the general kernel,
    unsigned* ncount;           //   queued Newton trials,
    unsigned* scount;           //   cells for the split stage,
    unsigned* shead;            //   split-stage queue head,
    unsigned* tcount;           //   cells for the split-Newton stage,
    unsigned* thead;            //   split-Newton queue head
    int32_t* glist;             // [nb] cells for the general kernel
    int32_t* slist;             // [nb] cells for the split stage
    int32_t* tlist;             // [nb] cells for the split-Newton stage
    int split;                  // 1: unstable cells take the split stage before the general kernel
    int snr;                    // 1: and then the split-Newton stage
    // cells whose previous flash took more than long_evals fugacity
    // evaluations (route 6): the general kernel takes them at once, beside
    // the stages, over llist with its own queue counter
    int long_evals;             // 0: none
    const int32_t* prev_evals;  // [nb] ITERS row 5 of the previous flash
    unsigned* lcount;           // counters, continued: long cells,
    unsigned* lqueue;           //   their work-queue counter,
    unsigned* nhead;            //   the Newton-trial queue head
    int32_t* llist;             // [nb]
    int32_t* route;             // [nb] 0 general kernel, 1 stable/undecided by stage 1, 2 unstable,
                                //      3 resumed at SplitNR after the split stage, 4 resumed at
                                //      the restart loop after the split-Newton stage, 5 written out,
                                //      6 long (taken by the general kernel beside the stages)
    int32_t* decided;           // [nb] first trial stage 1 found unstable, 4 none
    int32_t* defer;             // [nb] bit k: trial k queued for its Newton step
    double* Ytd;                // [nb] Yt of the decided trial
    int32_t* evals;             // [nb] stage-1 fugacity evaluations
    int32_t* its;               // [4][nb] SSM iterations per trial
    int32_t* item;              // [4][nb] Newton item of a queued trial
    double* di;                 // [Nc][nb] phi * z of the feed
    int64_t nr_cap;             // Newton items (4 nb)
    int32_t *nr_cell, *nr_k, *nr_flag, *nr_its, *nr_evals;   // [nr_cap]
    double *nr_Yt, *nr_y;       // [nr_cap], [Nc][nr_cap]
    double* rs_Yt;              // [nb]
    int32_t* rs_cnt;            // [8][nb] counters of the stages: SSMsta, NRsta, evaluations,
                                //   d ln f evaluations (ITERS row 6), SSMsp, Rachford-Rice, SSM
                                //   converged, split Newton steps
    double* rs_v;               // [5][Nc][nb] after the split stage: x0, x1, K, ln phi(x0), ln phi(x1)
    double* rs_nu;              // [nb] nu of the split stage's last Rachford-Rice solve
};
// stages of the batched flash for one model on persistent grids of grid[0]
// (stability SSM loops), grid[1] (stability Newton steps), grid[2] (split
// SSM loops) and grid[3] (split Newton steps) 64-lane blocks
hipError_t run_pe_sta(int NC, const DevFluid* f, const PeIO& io, const StaIO& s, const int* grid, hipStream_t st);
// before run_pe_sta: zero the stage counters and, with s.long_evals, mark the
// long cells and list them
hipError_t run_pe_sta_prep(const StaIO& s, int64_t nb, hipStream_t st);
// resident blocks per compute unit of the stability SSM (which = 0), Newton
// (1), split (2) and split-Newton (3) stages
int pe_sta_blocks_per_cu(int NC, int which, bool multi);

// queue order for the next batched flash (pe.hip): cells by descending count
// of fugacity evaluations in the last one; hist has queue_order_buckets()
// entries, zeroed by the caller
int queue_order_buckets();
hipError_t run_queue_order(const int32_t* evals, int64_t nb, unsigned* hist, int32_t* order, hipStream_t st);

// flash-derivative tail (tail.hip): properties, labels, dVf, dXsdXp,
// SkipPSA next-step data and the PassFlashValue writes
struct TailIO {
    int64_t nb;
    const double* P;
    const double* T;
    const double* Ni;
    const int32_t* pe_NP;
    const double* pe_nu;
    const double* pe_x;
    const int32_t* ftype_out;
    // outputs (BulkVarSet mirror)
    uint32_t* phaseExist;
    double *S, *xij, *rho, *xi, *mu, *rhoP, *xiP, *muP, *rhox, *xix, *mux;
    double *Nt, *vf, *vfP, *vfi, *dXsdXp;   // dXsdXp cell-major [nb][np*b][b]
    uint32_t* skip_flag;
    double *skip_minEig, *skip_P, *skip_T, *skip_zi;
    const int32_t* mid;         // [nb] model of each cell, or null (one model)
    // tail in two passes around the general flash kernel: with route set, pass 1
    // takes the cells the flash stages finished (route 1 or 5), pass 2 the others
    const int32_t* route;
    int route_pass;
    // with list set, the cells are list[0 .. *nlist) (route is not read): pass 2
    // runs over the general kernel's and the long cells' lists, packed in
    // wavefronts instead of scattered over all cells
    const int32_t* list;
    const unsigned* nlist;
};

// team mapping (pe_team.hip): T lanes per cell, T = 4, 8 or 16
bool pe_team_size_ok(int NC, int T);
hipError_t run_pe_team(int NC, int T, const DevFluid* f, const PeIO& io, bool batched, int grid_blocks,
                       hipStream_t st);
int pe_team_blocks_per_cu(int NC, int T);
hipError_t run_pe(int NC, const DevFluid* f, const PeIO& io, int mode, int grid_waves,
                  hipStream_t st);
hipError_t run_tail(int NC, const DevFluid* f, const TailIO& io, hipStream_t st);

// FIM kernels (fim.hip)

// Connection record of a cell, aux_len(b, nc) doubles at aux + n * aux_len
// (written by krpc_aux, read by the Jacobian assembly):
//   [8j, 8j + 8)            phase j: rho, Pj, xi, kr, mu, exists (1/0), lambda, depth
//   [24 + 3bj, 24 + 3b(j+1)) phase j: dlambda[b], dPj[b], drho[b] over the primary variables
//   [24 + 9b + nc j, ...)    phase j: x_ij, i < nc
// Even length keeps every record 16-byte aligned.
constexpr int aux_len(int b, int nc) { return (24 + 9 * b + 3 * nc + 1) & ~1; }

struct CellIO {
    int64_t nb;
    int32_t nc;                 // bulk components (Nc + 1)
    const double *P, *Ni, *lNi, *S, *xij, *rho, *xi, *mu, *rhoP, *xiP, *muP;
    const double *rhox, *xix, *mux, *Nt, *vf, *vfP, *vfi, *dXsdXp;
    const uint32_t* phaseExist;
    double *kr, *Pc, *Pj, *dKrdS, *dPcdS, *poro, *poroP, *rockVp;
    const double *v, *depth, *poroInit;
    double* aux;                // [nb][aux_len] connection record (krpc_aux)
    const int32_t* mid;         // [nb] model of each cell, or null (one model)
};

// rock compressibility of one model (OCPRockIsoT_Linear, src/OCPRock.cpp:19-24)
struct DevRock {
    double Pref, cp1, cp2;
};

// block-CSR and cell adjacency
struct BsrIO {
    const int64_t* rowptr;      // [nb+1]
    const int32_t* colind;      // [nnzb]
    const int32_t* tslot;       // [nnzb] slot of the transposed block
    const int32_t* nbr_conn;    // [nnzb] connection of the slot (-1 = diagonal)
    double* val;                // [nnzb][b][b]
    double* rhs;                // [nb][b]
    const int32_t* order;       // [nb] row processing order, or null (natural)
};

struct ConnIO {
    int64_t nconn;
    const int32_t *cb, *ce;
    const double* trans;
};

hipError_t run_krpc_rock(const DevSat* s, const DevRock* r, const CellIO& c, hipStream_t st);
hipError_t run_res_asm(int nc, const CellIO& c, const ConnIO& cn, const BsrIO& bsr, double dt,
                       bool do_res, bool do_jac, double* res, double* blockmax, int64_t* blockarg,
                       hipStream_t st);
hipError_t run_norms_final(int64_t nblocks, const double* blockmax, const int64_t* blockarg,
                           double* out4, hipStream_t st);
// norms of the residual in res (-R) into blockmax/blockarg (res_asm_blocks(nc, nb) blocks)
hipError_t run_res_norms(int nc, const CellIO& c, const double* res, double* blockmax, int64_t* blockarg,
                         hipStream_t st);

// Peaceman wells (wells.hip): per perforation in parallel, then the bulk
// rows per perforated cell and the well rows per well, each in perforation
// order (the reference's update sequence, no atomics)
struct WellIO {
    int32_t nwell, nperf, nucell;
    const ocpx_well* wells;     // [nwell] device copy; bhp is the current iterate
    const int64_t* cell;        // [nperf] perforated cell
    const double* WI;           // [nperf] WI * multiplier (CONV_DARCY included)
    const double* dG;           // [nperf] perforation head
    const double* xi;           // [nperf] injected fluid's molar density (injectors)
    const int32_t* pwell;       // [nperf] well of each perforation
    const int64_t* ucell;       // [nucell] perforated cells, ascending
    const int32_t* uptr;        // [nucell + 1] their perforations in uperf
    const int32_t* uperf;       // [nperf] perforations grouped by cell, in perforation order
    double* qperf;              // [nperf][nc] perforation component rates
    double* bdiag;              // [nperf][b][b] perforation block added to the bulk diagonal
    double* dD;                 // [nperf] the perforation's share of the well diagonal (0, 0)
    double* Bnw;                // [nperf][b] column 0 of block (n, w)
    double* Cwn;                // [nperf][b] row 0 of block (w, n)
    double* Dw;                 // [nwell][b][b]
    double* wres;               // [nwell] well-row residual (before the sign flip)
    double* wrhs;               // [nwell][b] right-hand side of the well rows (-wres in entry 0)
    double* qw;                 // [nwell][nc] well component rates
};
hipError_t run_well_terms(int nc, const CellIO& c, const BsrIO& bsr, const WellIO& w, double dt, bool do_jac,
                          double* resAbs, hipStream_t st);
// the well rows and columns of y = A x after the bulk bsr_spmv (x, y of length (nb + nwell) * B)
hipError_t run_well_spmv(const WellIO& w, int B, int64_t nb, const double* x, double* y, hipStream_t st);
// bhp += u[(nb + w) * B] for every well (device copy of the wells)
hipError_t run_well_bhp_update(ocpx_well* wells, int32_t nwell, int B, int64_t nb, const double* u, hipStream_t st);
hipError_t run_get_solution(int nc, double* P, double* Ni, double* S, double* xij,
                            const double* dXsdXp, const double* u, double dSmax, int64_t nb,
                            hipStream_t st);
// linear solve (solver.hip)
hipError_t run_bsr_diag_inv(int B, const int64_t* rowptr, const double* val, int64_t nb, double* dinv,
                            hipStream_t st);
hipError_t run_bsr_spmv(int B, const int64_t* rowptr, const int32_t* col, const double* val, const double* x,
                        double* y, int64_t nb, hipStream_t st);
hipError_t run_dot(const double* x, const double* y, int64_t len, double* part, double* out, hipStream_t st);
int dot_scratch_len();
hipError_t run_axpbypcz(double* z, const double* x, double a, const double* y, double b, const double* w,
                        int64_t len, hipStream_t st);

hipError_t run_jt_lambda(int nc, int64_t nb, const BsrIO& bsr, const double* lambda, double* out,
                         hipStream_t st);
hipError_t run_dR_dtrans(int nc, const CellIO& c, const ConnIO& cn, double dt,
                         const double* lambda, double* dTrans, double* dVp, hipStream_t st);
hipError_t run_fill_u32(unsigned* p, unsigned v, int64_t n, hipStream_t st);

}  // namespace ocpx
