// ocpx_host.cpp — C ABI (include/ocpx.h) over the gfx950 kernels.
//
// The context mirrors what IsoT_FIM owns in the reference
// (IsoT_FIM::AllocateReservoir, src/IsoThermalMethod.cpp:905-998): the
// BulkVarSet arrays, the connection list and the linear system, here all
// device-resident and component-major.  Host arrays cross the boundary in
// the reference's BulkVarSet layout (cell-major) and are transposed here.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/ocpx.h"
#include "ocpx_dev.hpp"
#include "ocpx_kernels.hpp"
#include "layout.hpp"
#include "amg_layout.hpp"
#include "cpr.hpp"

namespace ocpx {
int64_t res_asm_blocks(int nc, int64_t nb);
int pe_batched_blocks_per_cu(int NC);
}

using namespace ocpx;

namespace {

struct Field {
    void* d = nullptr;   // device [len][nb], or [nb][len] when aos
    int len = 0;         // entries per cell
    int esz = 8;         // element size
    bool aos = false;    // cell-major on the device too (dXsdXp)
};

struct TimingAcc {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    double total_ms = 0;
    int64_t calls = 0;
};

}  // namespace

struct ocpx_ctx {
    int NC = 0, nc = 0;
    int64_t nb = 0, nconn = 0, nnzb = 0;
    hipStream_t st = nullptr;
    bool own_stream = false;
    std::string err;
    DevFluid hf{};               // fluid of model 0 (host copy)
    int nmodel = 1;
    DevFluid* df = nullptr;      // [nmodel]
    DevSat* ds = nullptr;        // [nmodel]
    DevRock* drock = nullptr;    // [nmodel]
    int32_t* mid = nullptr;      // [nb] model of each cell (nmodel > 1), else null
    Field f[OCPX_F_NFIELDS];
    Field lf[OCPX_F_NFIELDS];   // last-time-step copies (UpdateLastTimeStep)
    bool have_last = false;
    bool ftype_override = false;
    // geometry
    double *v = nullptr, *depth = nullptr, *poroInit = nullptr, *trans = nullptr;
    double* aux = nullptr;       // connection record [nb][aux_len(b, nc)]
    int32_t *cb = nullptr, *ce = nullptr;
    // PE results
    int32_t* pe_NP = nullptr;
    double *pe_nu = nullptr, *pe_x = nullptr;
    unsigned* queue = nullptr;
    int qmode = 1;               // batched flash queue: 1 longest first, 0 natural order
    int32_t* qorder = nullptr;   // [nb] batched flash: cell at each queue position
    unsigned* qhist = nullptr;   // counting-sort buckets for the next order
    // linear system
    int64_t* rowptr = nullptr;
    int32_t *colind = nullptr, *tslot = nullptr, *nbr_conn = nullptr;
    int32_t* order = nullptr;    // assembly row order (null = natural)
    double *val = nullptr, *rhs = nullptr, *resAbs = nullptr;
    double* bmax = nullptr;
    int64_t* barg = nullptr;
    double* norms4 = nullptr;
    double *lam = nullptr, *out = nullptr, *gT = nullptr, *gV = nullptr;
    bool flashed = false, assembled = false;
    // linear solve (allocated on the first ocpx_linear_solve)
    double *sol = nullptr, *dinv = nullptr, *kv[8] = {}, *dpart = nullptr, *dout = nullptr;
    bool solved = false;
    bool aux_valid = false;      // krpc_aux ran on the current flash state
    int grid_waves = 0;
    int pe_lanes = 1;            // batched-flash mapping: lanes per cell (1, 4, 8, 16)
    int pe_lanes_auto = 0;       // the autotuner's choice, 0 if it has not run
    int team_blocks[3] = {0, 0, 0}; // persistent grid of the team kernels, T = 4 / 8 / 16
    int ncu = 256;               // compute units of the device
    int sta_grid[4] = {0, 0, 0, 0}; // persistent grids of the flash stages (stability SSM, Newton, split SSM, split Newton)
    int stages = 3;              // batched flash: 1 strict stability test in its own stages, 2 and
                                 // the split's first SSM pass, 3 and its Newton steps (pe_sta.hip), 0 none
    bool sta_ready = false;      // the stage buffers are allocated
    StaIO sta{};
    int long_evals = 1000;       // staged flash: cells over this many evaluations last time run beside the stages
    hipStream_t st2 = nullptr;   // the branch that runs them, forked from and joined to st by events
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    hipStream_t st3 = nullptr;   // the tail of the cells the stages finished, beside the general kernel
    hipEvent_t ev_fork3 = nullptr, ev_join3 = nullptr;
    bool tail_rest = false;      // the next tail takes only the cells that pass left
    // Peaceman wells (ocpx_set_wells)
    int32_t nwell = 0, nperf = 0;
    ocpx_well* wells = nullptr;  // [nwell] device copy
    int64_t* pcell = nullptr;
    double *pWI = nullptr, *pdG = nullptr, *pxi = nullptr;
    double *Bnw = nullptr, *Cwn = nullptr, *Dw = nullptr, *wres = nullptr, *qw = nullptr;
    double *wrhs = nullptr, *dinv_w = nullptr, *rhsf = nullptr;
    double *qperf = nullptr, *bdiag = nullptr, *dD = nullptr;
    int32_t *pwell = nullptr, *uptr = nullptr, *uperf = nullptr;
    int64_t* ucell = nullptr;
    int32_t nucell = 0;
    int64_t* wrow = nullptr;     // [nwell + 1] 0, 1, ..: block-row pointer of the well diagonal blocks
    bool wells_jac = false;      // the assembled system holds the well rows (ocpx_well_terms, jacobian = 1)
    bool wells_applied = false;  // the current residual already holds the well terms
    bool res_current = false;    // residual / Jacobian were formed from the current state
    int32_t solved_wells = 0;    // well rows in the last ocpx_linear_solve
    int64_t sol_cap = 0;         // length of the solver vectors
    // FGMRES(30) with the CPR-type preconditioner (cpr.hpp)
    int lsolver = 1;             // 1 FGMRES + CPR, 0 BiCGStab + block Jacobi
    std::unique_ptr<PressureAmg> amg;   // built at the first solve from the connection graph
    std::vector<int32_t> gs_cptr;        // colour lists of the block Gauss-Seidel stage
    int32_t* gs_rows = nullptr;
    double *fg_V = nullptr, *fg_Z = nullptr, *fg_w = nullptr, *fg_t = nullptr, *fg_x1 = nullptr;
    double *fg_r2 = nullptr, *fg_pb = nullptr, *fg_px = nullptr, *fg_h = nullptr, *fg_part = nullptr;
    double* fg_out = nullptr;
    int64_t fg_cap = 0;
    // timing
    bool timing = false;
    TimingAcc tacc[9];
    // graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    double gdt = 0;
    int gmode = -1;
    uint64_t gen = 0;            // settings generation: changes whenever a captured pass goes stale
    std::vector<void*> allocs;
};

namespace {

int fail(ocpx_ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(c, call)                                                                            \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail(c, OCPX_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));        \
    } while (0)

template <class T>
int dalloc(ocpx_ctx* c, T** p, size_t count)
{
    if (count == 0) count = 1;
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, count * sizeof(T));
    if (e != hipSuccess) return fail(c, OCPX_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    e = hipMemset(q, 0, count * sizeof(T));
    if (e != hipSuccess) return fail(c, OCPX_E_HIP, std::string("hipMemset: ") + hipGetErrorString(e));
    c->allocs.push_back(q);
    *p = (T*)q;
    return OCPX_OK;
}

int64_t flen(int field, int Nc)
{
    const int nc = Nc + 1, np = 3;
    switch (field) {
    case OCPX_F_NI: case OCPX_F_LNI: case OCPX_F_VFI: return nc;
    case OCPX_F_PHASEEXIST: case OCPX_F_S: case OCPX_F_RHO: case OCPX_F_XI: case OCPX_F_MU:
    case OCPX_F_RHOP: case OCPX_F_XIP: case OCPX_F_MUP: case OCPX_F_KR: case OCPX_F_PC:
    case OCPX_F_PJ: return np;
    case OCPX_F_XIJ: case OCPX_F_RHOX: case OCPX_F_XIX: case OCPX_F_MUX: return np * nc;
    case OCPX_F_DSEC_DPRI: return np * (nc + 1) * (nc + 1);
    case OCPX_F_DKRDS: case OCPX_F_DPCDS: return np * np;
    case OCPX_F_SKIP_ZI: return Nc;
    case OCPX_F_ITERS: return 8;
    default: return (field >= 0 && field < OCPX_F_NFIELDS) ? 1 : -1;
    }
}

int fesz(int field)
{
    switch (field) {
    case OCPX_F_PHASEEXIST: case OCPX_F_FTYPE_IN: case OCPX_F_FTYPE_OUT: case OCPX_F_SKIP_FLAG:
    case OCPX_F_ITERS: return 4;
    default: return 8;
    }
}

double* D(ocpx_ctx* c, int field) { return (double*)c->f[field].d; }

// BulkVarSet layout (cell-major, [n][L]) <-> device layout (component-major,
// [L][n]), in tiles of 64 cells so both sides stream.
template <class T>
void to_soa(const T* aos, T* soa, size_t n, size_t L)
{
    for (size_t c0 = 0; c0 < n; c0 += 64) {
        const size_t c1 = c0 + 64 < n ? c0 + 64 : n;
        for (size_t k = 0; k < L; k++)
            for (size_t c = c0; c < c1; c++) soa[k * n + c] = aos[c * L + k];
    }
}

template <class T>
void to_aos(const T* soa, T* aos, size_t n, size_t L)
{
    for (size_t c0 = 0; c0 < n; c0 += 64) {
        const size_t c1 = c0 + 64 < n ? c0 + 64 : n;
        for (size_t c = c0; c < c1; c++)
            for (size_t k = 0; k < L; k++) aos[c * L + k] = soa[k * n + c];
    }
}

CellIO cell_io(ocpx_ctx* c)
{
    CellIO io{};
    io.nb = c->nb;
    io.nc = c->nc;
    io.P = D(c, OCPX_F_P); io.Ni = D(c, OCPX_F_NI); io.lNi = D(c, OCPX_F_LNI); io.S = D(c, OCPX_F_S);
    io.xij = D(c, OCPX_F_XIJ); io.rho = D(c, OCPX_F_RHO); io.xi = D(c, OCPX_F_XI); io.mu = D(c, OCPX_F_MU);
    io.rhoP = D(c, OCPX_F_RHOP); io.xiP = D(c, OCPX_F_XIP); io.muP = D(c, OCPX_F_MUP);
    io.rhox = D(c, OCPX_F_RHOX); io.xix = D(c, OCPX_F_XIX); io.mux = D(c, OCPX_F_MUX);
    io.Nt = D(c, OCPX_F_NT); io.vf = D(c, OCPX_F_VF); io.vfP = D(c, OCPX_F_VFP); io.vfi = D(c, OCPX_F_VFI);
    io.dXsdXp = D(c, OCPX_F_DSEC_DPRI);
    io.phaseExist = (const uint32_t*)c->f[OCPX_F_PHASEEXIST].d;
    io.kr = D(c, OCPX_F_KR); io.Pc = D(c, OCPX_F_PC); io.Pj = D(c, OCPX_F_PJ);
    io.dKrdS = D(c, OCPX_F_DKRDS); io.dPcdS = D(c, OCPX_F_DPCDS);
    io.poro = D(c, OCPX_F_PORO); io.poroP = D(c, OCPX_F_POROP); io.rockVp = D(c, OCPX_F_ROCKVP);
    io.v = c->v; io.depth = c->depth; io.poroInit = c->poroInit;
    io.aux = c->aux;               // null without connections: no record is ever read
    io.mid = c->mid;
    return io;
}

BsrIO bsr_io(ocpx_ctx* c)
{
    BsrIO b{};
    b.rowptr = c->rowptr; b.colind = c->colind; b.tslot = c->tslot; b.nbr_conn = c->nbr_conn;
    b.val = c->val; b.rhs = c->rhs;
    b.order = c->order;
    return b;
}

ConnIO conn_io(ocpx_ctx* c)
{
    ConnIO n{};
    n.nconn = c->nconn; n.cb = c->cb; n.ce = c->ce; n.trans = c->trans;
    return n;
}

struct TimerScope {
    ocpx_ctx* c;
    int which;
    hipEvent_t a = nullptr, b = nullptr;
    TimerScope(ocpx_ctx* cc, int w) : c(cc), which(w)
    {
        if (!c->timing) return;
        hipEventCreate(&a);
        hipEventCreate(&b);
        hipEventRecord(a, c->st);
    }
    ~TimerScope()
    {
        if (!a) return;
        hipEventRecord(b, c->st);
        c->tacc[which].pending.push_back({a, b});
    }
};

inline int team_slot(int T) { return T == 4 ? 0 : (T == 8 ? 1 : 2); }

// buffers of the stability stages, allocated at the first staged flash
int sta_alloc(ocpx_ctx* c)
{
    if (c->sta_ready) return OCPX_OK;
    const int64_t n = c->nb, cap = 4 * c->nb;
    StaIO& s = c->sta;
    unsigned* cnt = nullptr;
    int rc;
    if ((rc = dalloc(c, &cnt, 16)) || (rc = dalloc(c, &s.glist, n)) || (rc = dalloc(c, &s.route, n)) ||
        (rc = dalloc(c, &s.decided, n)) || (rc = dalloc(c, &s.defer, n)) || (rc = dalloc(c, &s.Ytd, n)) ||
        (rc = dalloc(c, &s.evals, n)) || (rc = dalloc(c, &s.its, 4 * n)) || (rc = dalloc(c, &s.item, 4 * n)) ||
        (rc = dalloc(c, &s.di, (size_t)c->NC * n)) || (rc = dalloc(c, &s.nr_cell, cap)) ||
        (rc = dalloc(c, &s.nr_k, cap)) || (rc = dalloc(c, &s.nr_flag, cap)) || (rc = dalloc(c, &s.nr_its, cap)) ||
        (rc = dalloc(c, &s.nr_evals, cap)) || (rc = dalloc(c, &s.nr_Yt, cap)) ||
        (rc = dalloc(c, &s.nr_y, (size_t)c->NC * cap)) || (rc = dalloc(c, &s.rs_Yt, n)) ||
        (rc = dalloc(c, &s.rs_cnt, 8 * n)) || (rc = dalloc(c, &s.slist, n)) || (rc = dalloc(c, &s.tlist, n)) ||
        (rc = dalloc(c, &s.llist, n)) ||
        (rc = dalloc(c, &s.rs_v, (size_t)5 * c->NC * n)) || (rc = dalloc(c, &s.rs_nu, n)))
        return rc;
    s.qhead = cnt;
    s.gcount = cnt + 1;
    s.ncount = cnt + 2;
    s.scount = cnt + 3;
    s.shead = cnt + 4;
    s.tcount = cnt + 5;
    s.thead = cnt + 6;
    s.lcount = cnt + 7;
    s.lqueue = cnt + 8;
    s.nhead = cnt + 9;
    if (!c->st2) {
        if (hipStreamCreate(&c->st2) != hipSuccess || hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess)
            return fail(c, OCPX_E_HIP, "flash branch stream or events");
    }
    if (!c->st3) {
        if (hipStreamCreate(&c->st3) != hipSuccess || hipEventCreateWithFlags(&c->ev_fork3, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&c->ev_join3, hipEventDisableTiming) != hipSuccess)
            return fail(c, OCPX_E_HIP, "tail branch stream or events");
    }
    s.nr_cap = cap;
    c->sta_ready = true;
    return OCPX_OK;
}

TailIO tail_io(ocpx_ctx* c);

// branch_ok: the long cells may run on the context's second stream and the
// finished cells' tail on its third (off when the context's own stream is
// itself a branch of a batch capture); tail_follows: launch_tail comes next
int launch_pe(ocpx_ctx* c, int mode, bool branch_ok = true, bool tail_follows = false)
{
    PeIO io{};
    io.nb = c->nb;
    io.P = D(c, OCPX_F_P); io.T = D(c, OCPX_F_T); io.Ni = D(c, OCPX_F_NI);
    io.phaseExist = (const uint32_t*)c->f[OCPX_F_PHASEEXIST].d;
    io.S = D(c, OCPX_F_S); io.xij = D(c, OCPX_F_XIJ);
    io.ftype_in = c->ftype_override ? (const int32_t*)c->f[OCPX_F_FTYPE_IN].d : nullptr;
    io.skip_flag = (const uint32_t*)c->f[OCPX_F_SKIP_FLAG].d;
    io.skip_minEig = D(c, OCPX_F_SKIP_MINEIG); io.skip_P = D(c, OCPX_F_SKIP_P);
    io.skip_T = D(c, OCPX_F_SKIP_T); io.skip_zi = D(c, OCPX_F_SKIP_ZI);
    io.pe_NP = c->pe_NP; io.pe_nu = c->pe_nu; io.pe_x = c->pe_x;
    io.ftype_out = (int32_t*)c->f[OCPX_F_FTYPE_OUT].d;
    io.iters = (int32_t*)c->f[OCPX_F_ITERS].d;
    io.queue = c->queue;
    io.qorder = (mode == OCPX_FLASH_BATCHED && c->qmode) ? c->qorder : nullptr;
    if (mode == OCPX_FLASH_BATCHED) HIPCHK(c, run_fill_u32(c->queue, 0u, 1, c->st));
    io.mid = c->mid;
    if (c->mid && (mode != OCPX_FLASH_BATCHED || c->pe_lanes == 1))
        return fail(c, OCPX_E_UNSUPPORTED, "contexts of several models flash in batched mode with a team mapping");
    // single-model batched flash: the strict stability test in stages, then the
    // general kernel on the cells that remain (the split, other flash types)
    const bool staged = mode == OCPX_FLASH_BATCHED && c->stages && c->nb > 0;
    if (staged) {
        int rc = sta_alloc(c);
        if (rc) return rc;
    }
    bool branched = false, early_tail = false;
    {
        TimerScope t(c, 0);
        if (staged) {
            c->sta.split = c->stages >= 2;
            c->sta.snr = c->stages >= 3;
            // the cells that took longest last time run beside the stages from the
            // start (their serial Newton and restart work is the flash's critical path)
            const bool branch = branch_ok && c->long_evals > 0 && mode == OCPX_FLASH_BATCHED;
            c->sta.long_evals = branch ? c->long_evals : 0;
            c->sta.prev_evals = io.iters + (size_t)5 * c->nb;
            HIPCHK(c, run_pe_sta_prep(c->sta, c->nb, c->st));
            if (branch) {
                PeIO lio = io;
                lio.qorder = c->sta.llist;
                lio.nq = c->sta.lcount;
                lio.queue = c->sta.lqueue;
                HIPCHK(c, hipEventRecord(c->ev_fork, c->st));
                HIPCHK(c, hipStreamWaitEvent(c->st2, c->ev_fork, 0));
                const int T = c->pe_lanes;
                if (T > 1) HIPCHK(c, run_pe_team(c->NC, T, c->df, lio, true, std::min(c->team_blocks[team_slot(T)], 512), c->st2));
                else HIPCHK(c, run_pe(c->NC, c->df, lio, OCPX_FLASH_BATCHED, std::min(c->grid_waves, 512), c->st2));
                HIPCHK(c, hipEventRecord(c->ev_join, c->st2));
                branched = true;
            }
            HIPCHK(c, run_pe_sta(c->NC, c->df, io, c->sta, c->sta_grid, c->st));
            if (tail_follows && branch_ok) {
                // the cells the stages finished take their FlashDer tail now, beside
                // the general kernel, whose few long cells leave the device idle
                TailIO tio = tail_io(c);
                tio.route = c->sta.route;
                tio.route_pass = 1;
                HIPCHK(c, hipEventRecord(c->ev_fork3, c->st));
                HIPCHK(c, hipStreamWaitEvent(c->st3, c->ev_fork3, 0));
                HIPCHK(c, run_tail(c->NC, c->df, tio, c->st3));
                HIPCHK(c, hipEventRecord(c->ev_join3, c->st3));
                early_tail = true;
                c->tail_rest = true;
            }
            io.qorder = c->sta.glist;
            io.nq = c->sta.gcount;
            io.sta_route = c->sta.route;
            io.rs_Yt = c->sta.rs_Yt;
            io.rs_cnt = c->sta.rs_cnt;
            io.rs_v = c->sta.rs_v;
            io.rs_nu = c->sta.rs_nu;
        }
        if (mode == OCPX_FLASH_BATCHED && c->pe_lanes > 1) {
            const int T = c->pe_lanes;
            HIPCHK(c, run_pe_team(c->NC, T, c->df, io, true, c->team_blocks[team_slot(T)], c->st));
        } else {
            HIPCHK(c, run_pe(c->NC, c->df, io, mode, c->grid_waves, c->st));
        }
        if (branched) HIPCHK(c, hipStreamWaitEvent(c->st, c->ev_join, 0));
        if (early_tail) HIPCHK(c, hipStreamWaitEvent(c->st, c->ev_join3, 0));
    }
    if (mode == OCPX_FLASH_BATCHED) {
        // the next batched flash takes the cells longest first, by this one's
        // fugacity-evaluation counts (ITERS row 5)
        HIPCHK(c, run_fill_u32(c->qhist, 0u, queue_order_buckets(), c->st));
        HIPCHK(c, run_queue_order(io.iters + (size_t)5 * c->nb, c->nb, c->qhist, c->qorder, c->st));
    }
    return OCPX_OK;
}

// Lanes per cell before the autotuner has run: one lane for a few components,
// a team of 8 or 16 (one or two components per lane) above.  Contexts of
// several models need a team (4 lanes below Nc = 5).
int default_lanes(int NC, bool multi)
{
    if (NC <= 4) return multi ? 4 : 1;
    return NC <= 8 ? 8 : 16;
}

TailIO tail_io(ocpx_ctx* c)
{
    TailIO io{};
    io.nb = c->nb;
    io.P = D(c, OCPX_F_P); io.T = D(c, OCPX_F_T); io.Ni = D(c, OCPX_F_NI);
    io.pe_NP = c->pe_NP; io.pe_nu = c->pe_nu; io.pe_x = c->pe_x;
    io.ftype_out = (const int32_t*)c->f[OCPX_F_FTYPE_OUT].d;
    io.phaseExist = (uint32_t*)c->f[OCPX_F_PHASEEXIST].d;
    io.S = D(c, OCPX_F_S); io.xij = D(c, OCPX_F_XIJ); io.rho = D(c, OCPX_F_RHO); io.xi = D(c, OCPX_F_XI);
    io.mu = D(c, OCPX_F_MU); io.rhoP = D(c, OCPX_F_RHOP); io.xiP = D(c, OCPX_F_XIP); io.muP = D(c, OCPX_F_MUP);
    io.rhox = D(c, OCPX_F_RHOX); io.xix = D(c, OCPX_F_XIX); io.mux = D(c, OCPX_F_MUX);
    io.Nt = D(c, OCPX_F_NT); io.vf = D(c, OCPX_F_VF); io.vfP = D(c, OCPX_F_VFP); io.vfi = D(c, OCPX_F_VFI);
    io.dXsdXp = D(c, OCPX_F_DSEC_DPRI);
    io.skip_flag = (uint32_t*)c->f[OCPX_F_SKIP_FLAG].d;
    io.skip_minEig = D(c, OCPX_F_SKIP_MINEIG); io.skip_P = D(c, OCPX_F_SKIP_P);
    io.skip_T = D(c, OCPX_F_SKIP_T); io.skip_zi = D(c, OCPX_F_SKIP_ZI);
    io.mid = c->mid;
    return io;
}

int launch_tail(ocpx_ctx* c)
{
    TailIO io = tail_io(c);
    TimerScope t(c, 1);
    if (c->tail_rest) {
        // the flash ran the tail of the cells its stages finished (route 1 or 5)
        // already; every other cell is on the general kernel's list or on the
        // long cells' list, so the rest of the tail runs over those two lists
        c->tail_rest = false;
        for (int k = 0; k < 2; k++) {
            io.list = k == 0 ? c->sta.glist : c->sta.llist;
            io.nlist = k == 0 ? c->sta.gcount : c->sta.lcount;
            HIPCHK(c, run_tail(c->NC, c->df, io, c->st));
        }
        return OCPX_OK;
    }
    HIPCHK(c, run_tail(c->NC, c->df, io, c->st));
    return OCPX_OK;
}

int launch_krpc_rock(ocpx_ctx* c)
{
    TimerScope t(c, 2);
    HIPCHK(c, run_krpc_rock(c->ds, c->drock, cell_io(c), c->st));
    c->aux_valid = c->flashed;
    return OCPX_OK;
}

int launch_res_asm(ocpx_ctx* c, double dt, bool dr, bool dj)
{
    c->wells_jac = false;              // a new residual / Jacobian holds no well terms yet
    c->wells_applied = false;
    c->res_current = true;
    {
        TimerScope t(c, dj ? 4 : 3);       // the residual / assembly kernel alone
        HIPCHK(c, run_res_asm(c->nc, cell_io(c), conn_io(c), bsr_io(c), dt, dr, dj, c->resAbs, c->bmax, c->barg,
                              c->st));
    }
    if (dr) HIPCHK(c, run_norms_final(res_asm_blocks(c->nc, c->nb), c->bmax, c->barg, c->norms4, c->st));
    return OCPX_OK;
}

int destroy_graph(ocpx_ctx* c)
{
    c->gen++;                    // graphs of batches holding this context are stale too
    if (c->gexec) hipGraphExecDestroy(c->gexec);
    if (c->graph) hipGraphDestroy(c->graph);
    c->gexec = nullptr;
    c->graph = nullptr;
    c->gmode = -1;
    return OCPX_OK;
}

}  // namespace

namespace ocpx {
// Device fluid constants from the ABI description (EoS_PR ctor
// src/OCPEoS.cpp:19-65, ViscosityMethod03 ctor src/OCPFuncPVT.cpp:415-459,
// OCPPhaseEquilibrium::Setup src/OCPPhaseEquilibrium.cpp:45-64)
void build_dev_fluid(const ocpx_fluid* fl, DevFluid& h)
{
    h = DevFluid{};
    const double R = R_GAS;
    
    h.NC = fl->nc;
    h.lbc_exact = fl->lbc_exact;
    h.skip = fl->skipPSA;
    h.RR_maxIt = fl->RR_maxIt; h.SSMsta_maxIt = fl->SSMsta_maxIt; h.NRsta_maxIt = fl->NRsta_maxIt;
    h.SSMsp_maxIt = fl->SSMsp_maxIt; h.NRsp_maxIt = fl->NRsp_maxIt;
    h.RR_tol = fl->RR_tol; h.SSMsta_tol2 = fl->SSMsta_tol * fl->SSMsta_tol; h.eYt = fl->SSMsta_eYt;
    h.NRsta_tol = fl->NRsta_tol; h.SSMsp_tol2 = fl->SSMsp_tol * fl->SSMsp_tol; h.NRsp_tol = fl->NRsp_tol;
    for (int i = 0; i < fl->nc; i++) {
        const double w = fl->acf[i];
        h.Tc[i] = fl->Tc[i]; h.Pc[i] = fl->Pc[i]; h.logPc[i] = std::log(fl->Pc[i]); h.MW[i] = fl->MW[i];
        h.Vc[i] = fl->Vc[i]; h.Vcvis[i] = fl->Vcvis[i]; h.OA[i] = fl->OmegaA[i]; h.OB[i] = fl->OmegaB[i];
        h.vsh[i] = fl->Vshift[i] * (R * fl->OmegaB[i] * fl->Tc[i] / fl->Pc[i]);
        h.mw[i] = w <= 0.49 ? 0.37464 + 1.54226 * w - 0.26992 * w * w
                            : 0.379642 + 1.48503 * w - 0.164423 * w * w + 0.016667 * w * w * w;
        h.kwA[i] = 5.373 * (1 + w);
        h.auxB[i] = 5.4402 * std::pow(fl->Tc[i], 1.0 / 6) / std::pow(fl->Pc[i], 2.0 / 3);
        h.sqrtMW[i] = std::sqrt(fl->MW[i]);
        h.bT[i] = fl->OmegaB[i] * fl->Tc[i] / fl->Pc[i];
        for (int k = 0; k < fl->nc; k++) h.omk[i * OCPX_NCMAX + k] = 1 - fl->BIC[i * OCPX_MAX_NC + k];
    }
    for (int k = 0; k < 5; k++) { h.lbc[k] = fl->LBCcoef[k] * 10; h.pvtw[k] = fl->PVTW[k]; }
    h.rhoW = fl->rhoW_std;
}
}  // namespace ocpx

namespace {

}  // namespace

extern "C" {

int64_t ocpx_field_len(int field, int32_t Nc) { return flen(field, Nc); }

}  // extern "C"

namespace {
// nmodel models share one context: fl, sat, rock are arrays of nmodel, g is
// the concatenated grid and mid (nmodel > 1) the model of each cell.
// Why the last context creation on this thread failed (no context exists to
// hold the message); ocpx_last_error(NULL) returns it.
thread_local std::string g_create_err;

int create_fail(int code, const char* msg)
{
    g_create_err = msg;
    return code;
}
}  // namespace

namespace ocpx {
// the message ocpx_last_error(NULL) returns, for calls without a context
void set_call_error(const std::string& msg) { g_create_err = msg; }
void clear_call_error() { g_create_err.clear(); }
}  // namespace ocpx

namespace {

int ctx_create_impl(int nmodel, const ocpx_fluid* fl, const ocpx_sat* sat, const ocpx_rock* rock,
                    const ocpx_grid* g, const int32_t* mid, void* stream, ocpx_ctx** out)
{
    g_create_err.clear();
    if (!out) return create_fail(OCPX_E_ARG, "null output pointer");
    *out = nullptr;
    if (!fl || !sat || !rock || !g || nmodel < 1) return create_fail(OCPX_E_ARG, "null fluid, tables, rock or grid");
    for (int m = 0; m < nmodel; m++) {
        if (fl[m].nc < 2 || fl[m].nc > 12)
            return create_fail(OCPX_E_UNSUPPORTED, "hydrocarbon components must be 2 to 12");
        if (fl[m].nc != fl[0].nc) return create_fail(OCPX_E_ARG, "models of one context must share Nc");
        if (sat[m].nswof < 1 || sat[m].nswof > OCPX_MAX_TABROWS || sat[m].nsgof < 1 ||
            sat[m].nsgof > OCPX_MAX_TABROWS)
            return create_fail(OCPX_E_ARG, "SWOF/SGOF row count outside 1..OCPX_MAX_TABROWS");
    }
    if (g->nb <= 0) return create_fail(OCPX_E_ARG, "grid with no cells");
    if (g->nb > INT32_MAX || g->nconn < 0 || g->nconn > INT32_MAX)
        return create_fail(OCPX_E_ARG, "cell or connection count outside int32");
    ocpx_ctx* c = new ocpx_ctx();
    c->NC = fl->nc;
    c->nc = fl->nc + 1;
    c->nb = g->nb;
    c->nconn = g->nconn;
    c->nmodel = nmodel;
    if (stream) c->st = (hipStream_t)stream;
    else {
        if (hipStreamCreate(&c->st) != hipSuccess) { delete c; return OCPX_E_HIP; }
        c->own_stream = true;
    }
    // connections must be valid and ordered by bId (iteratorConn order)
    for (int64_t k = 0; k < g->nconn; k++) {
        const int32_t b = g->conn_b[k], e = g->conn_e[k];
        if (b < 0 || e < 0 || b >= g->nb || e >= g->nb || b >= e || (k > 0 && b < g->conn_b[k - 1])) {
            c->err = "connections must satisfy 0 <= bId < eId < nb and be sorted by bId";
            *out = c;
            return OCPX_E_ARG;
        }
    }
    // --- fluid constants (EoS_PR ctor, ViscosityMethod03 ctor, flash controls),
    // saturation tables and rock, one set per model
    std::vector<DevFluid> hfs(nmodel);
    std::vector<DevSat> hss(nmodel);
    std::vector<DevRock> hrs(nmodel);
    for (int m = 0; m < nmodel; m++) {
        build_dev_fluid(&fl[m], hfs[m]);
        hss[m] = DevSat{};
        hss[m].nswof = sat[m].nswof; hss[m].nsgof = sat[m].nsgof;
        std::memcpy(hss[m].swof, sat[m].swof, sizeof(hss[m].swof));
        std::memcpy(hss[m].sgof, sat[m].sgof, sizeof(hss[m].sgof));
        hrs[m] = DevRock{rock[m].Pref, rock[m].cp1, rock[m].cp2};
    }
    c->hf = hfs[0];

    int rc;
#define ALLOC(p, n) if ((rc = dalloc(c, &(p), (n))) != OCPX_OK) { *out = c; return rc; }
    ALLOC(c->df, nmodel);
    ALLOC(c->ds, nmodel);
    ALLOC(c->drock, nmodel);
    if (hipMemcpy(c->df, hfs.data(), sizeof(DevFluid) * nmodel, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->ds, hss.data(), sizeof(DevSat) * nmodel, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->drock, hrs.data(), sizeof(DevRock) * nmodel, hipMemcpyHostToDevice) != hipSuccess) {
        *out = c;
        return fail(c, OCPX_E_HIP, "copy of fluid constants failed");
    }
    if (nmodel > 1) {
        ALLOC(c->mid, c->nb);
        if (hipMemcpy(c->mid, mid, (size_t)c->nb * 4, hipMemcpyHostToDevice) != hipSuccess) {
            *out = c;
            return fail(c, OCPX_E_HIP, "copy of the cell models failed");
        }
    }
    for (int fi = 0; fi < OCPX_F_NFIELDS; fi++) {
        c->f[fi].len = (int)flen(fi, c->NC);
        c->f[fi].esz = fesz(fi);
        c->f[fi].aos = fi == OCPX_F_DSEC_DPRI;
        if (c->f[fi].esz == 8) { double* p; ALLOC(p, (size_t)c->f[fi].len * c->nb); c->f[fi].d = p; }
        else { int32_t* p; ALLOC(p, (size_t)c->f[fi].len * c->nb); c->f[fi].d = p; }
    }
    ALLOC(c->v, c->nb); ALLOC(c->depth, c->nb); ALLOC(c->poroInit, c->nb);
    ALLOC(c->cb, c->nconn); ALLOC(c->ce, c->nconn); ALLOC(c->trans, c->nconn);
    ALLOC(c->pe_NP, c->nb); ALLOC(c->pe_nu, c->nb); ALLOC(c->pe_x, (size_t)2 * c->NC * c->nb);
    ALLOC(c->queue, 4);
    ALLOC(c->qorder, (size_t)c->nb);
    ALLOC(c->qhist, (size_t)queue_order_buckets());
    {
        std::vector<int32_t> iota((size_t)c->nb);
        for (int64_t k = 0; k < c->nb; k++) iota[(size_t)k] = (int32_t)k;
        if (c->nb && hipMemcpy(c->qorder, iota.data(), iota.size() * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess) {
            *out = c;
            return fail(c, OCPX_E_HIP, "queue order upload");
        }
    }
    if (c->nconn > 0) ALLOC(c->aux, (size_t)c->nb * aux_len(c->nc + 1, c->nc));
    // BSR in OCPMatrix push order: diagonal first, then per connection in
    // index order, row e gets (e, b) and row b gets (b, e)
    c->nnzb = c->nb + 2 * c->nconn;
    std::vector<int64_t> rp(c->nb + 1, 0);
    std::vector<int32_t> col(c->nnzb), ts(c->nnzb), nbc(c->nnzb);
    bsr_layout(c->nb, c->nconn, g->conn_b, g->conn_e, rp.data(), col.data(), ts.data(), nbc.data());
    const int B = c->nc + 1;
    ALLOC(c->rowptr, c->nb + 1); ALLOC(c->colind, c->nnzb); ALLOC(c->tslot, c->nnzb); ALLOC(c->nbr_conn, c->nnzb);
    ALLOC(c->val, (size_t)c->nnzb * B * B); ALLOC(c->rhs, (size_t)c->nb * B); ALLOC(c->resAbs, (size_t)c->nb * B);
    const int64_t nblk = res_asm_blocks(c->nc, c->nb);
    ALLOC(c->bmax, 2 * nblk); ALLOC(c->barg, 2 * nblk); ALLOC(c->norms4, 4);
    ALLOC(c->lam, (size_t)c->nb * B); ALLOC(c->out, (size_t)c->nb * B);
    ALLOC(c->gT, c->nconn); ALLOC(c->gV, c->nb);
    const std::vector<int32_t> ord =
        nmodel == 1 ? pencil_order(c->nb, c->nconn, g->conn_b, g->conn_e) : std::vector<int32_t>();
    if (!ord.empty()) ALLOC(c->order, c->nb);
#undef ALLOC
    bool okc = hipMemcpy(c->rowptr, rp.data(), rp.size() * 8, hipMemcpyHostToDevice) == hipSuccess &&
               hipMemcpy(c->colind, col.data(), col.size() * 4, hipMemcpyHostToDevice) == hipSuccess &&
               hipMemcpy(c->tslot, ts.data(), ts.size() * 4, hipMemcpyHostToDevice) == hipSuccess &&
               hipMemcpy(c->nbr_conn, nbc.data(), nbc.size() * 4, hipMemcpyHostToDevice) == hipSuccess &&
               hipMemcpy(c->v, g->v, c->nb * 8, hipMemcpyHostToDevice) == hipSuccess &&
               hipMemcpy(c->depth, g->depth, c->nb * 8, hipMemcpyHostToDevice) == hipSuccess &&
               hipMemcpy(c->poroInit, g->poroInit, c->nb * 8, hipMemcpyHostToDevice) == hipSuccess;
    if (okc && c->order)
        okc = hipMemcpy(c->order, ord.data(), ord.size() * 4, hipMemcpyHostToDevice) == hipSuccess;
    if (okc && c->nconn > 0)
        okc = hipMemcpy(c->cb, g->conn_b, c->nconn * 4, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(c->ce, g->conn_e, c->nconn * 4, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(c->trans, g->trans, c->nconn * 8, hipMemcpyHostToDevice) == hipSuccess;
    if (!okc) { *out = c; return fail(c, OCPX_E_HIP, "upload of grid data failed"); }
    // FTYPE_IN defaults to -1 (SkipPSA decides)
    {
        std::vector<int32_t> m1(c->nb, -1);
        if (hipMemcpy(c->f[OCPX_F_FTYPE_IN].d, m1.data(), c->nb * 4, hipMemcpyHostToDevice) != hipSuccess) {
            *out = c;
            return fail(c, OCPX_E_HIP, "init of FTYPE_IN failed");
        }
    }
    int dev = 0, ncu = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    c->ncu = ncu > 0 ? ncu : 256;
    for (int q = 0; q < 4; q++) c->sta_grid[q] = c->ncu * pe_sta_blocks_per_cu(c->NC, q, c->mid != nullptr);
    const int per_cu = pe_batched_blocks_per_cu(c->NC);
    c->grid_waves = (ncu > 0 ? ncu : 256) * (per_cu > 0 ? per_cu : 4);
    for (int q = 0; q < 3; q++) {
        const int T = 4 << q;
        if (!pe_team_size_ok(c->NC, T)) continue;
        const int b = pe_team_blocks_per_cu(c->NC, T);
        c->team_blocks[q] = (ncu > 0 ? ncu : 256) * (b > 0 ? b : 4);
    }
    c->pe_lanes = default_lanes(c->NC, nmodel > 1);
    *out = c;
    return OCPX_OK;
}
}  // namespace

extern "C" {

int ocpx_ctx_create(const ocpx_fluid* fl, const ocpx_sat* sat, const ocpx_rock* rock, const ocpx_grid* g,
                    void* stream, ocpx_ctx** out)
{
    return ctx_create_impl(1, fl, sat, rock, g, nullptr, stream, out);
}

int ocpx_ctx_create_multi(int32_t nmodel, const ocpx_fluid* fl, const ocpx_sat* sat, const ocpx_rock* rock,
                          const ocpx_grid* grids, void* stream, ocpx_ctx** out)
{
    if (!out) return OCPX_E_ARG;
    *out = nullptr;
    if (nmodel < 1 || !fl || !sat || !rock || !grids) return OCPX_E_ARG;
    if (nmodel == 1) return ctx_create_impl(1, fl, sat, rock, grids, nullptr, stream, out);
    int64_t nb = 0, nconn = 0;
    for (int m = 0; m < nmodel; m++) {
        if (grids[m].nb <= 0 || grids[m].nconn < 0) return OCPX_E_ARG;
        nb += grids[m].nb;
        nconn += grids[m].nconn;
    }
    if (nb > INT32_MAX || nconn > INT32_MAX) return OCPX_E_ARG;
    std::vector<double> v(nb), depth(nb), poro(nb), trans(nconn);
    std::vector<int32_t> cb(nconn), ce(nconn), mid(nb);
    int64_t o = 0, k = 0;
    for (int m = 0; m < nmodel; m++) {
        const ocpx_grid& gm = grids[m];
        std::memcpy(&v[o], gm.v, gm.nb * 8);
        std::memcpy(&depth[o], gm.depth, gm.nb * 8);
        std::memcpy(&poro[o], gm.poroInit, gm.nb * 8);
        for (int64_t q = 0; q < gm.nb; q++) mid[o + q] = m;
        for (int64_t q = 0; q < gm.nconn; q++) {
            cb[k + q] = (int32_t)(gm.conn_b[q] + o);
            ce[k + q] = (int32_t)(gm.conn_e[q] + o);
            trans[k + q] = gm.trans[q];
        }
        o += gm.nb;
        k += gm.nconn;
    }
    ocpx_grid g{nb, nconn, v.data(), depth.data(), poro.data(), cb.data(), ce.data(), trans.data()};
    return ctx_create_impl(nmodel, fl, sat, rock, &g, mid.data(), stream, out);
}

void ocpx_ctx_destroy(ocpx_ctx* c)
{
    if (!c) return;
    hipStreamSynchronize(c->st);
    destroy_graph(c);
    for (auto& t : c->tacc)
        for (auto& p : t.pending) { hipEventDestroy(p.first); hipEventDestroy(p.second); }
    for (void* p : c->allocs) hipFree(p);
    if (c->own_stream) hipStreamDestroy(c->st);
    if (c->st2) hipStreamDestroy(c->st2);
    if (c->ev_fork) hipEventDestroy(c->ev_fork);
    if (c->ev_join) hipEventDestroy(c->ev_join);
    if (c->st3) hipStreamDestroy(c->st3);
    if (c->ev_fork3) hipEventDestroy(c->ev_fork3);
    if (c->ev_join3) hipEventDestroy(c->ev_join3);
    delete c;
}

const char* ocpx_last_error(const ocpx_ctx* c)
{
    if (c) return c->err.c_str();
    return g_create_err.empty() ? "null context" : g_create_err.c_str();
}

int ocpx_sync(ocpx_ctx* c)
{
    if (!c) return OCPX_E_ARG;
    HIPCHK(c, hipStreamSynchronize(c->st));
    return OCPX_OK;
}

int ocpx_set_field(ocpx_ctx* c, int field, const void* host)
{
    if (!c || !host || field < 0 || field >= OCPX_F_NFIELDS) return fail(c, OCPX_E_ARG, "bad field");
    const Field& F = c->f[field];
    const size_t n = c->nb, L = F.len;
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (F.aos) {
        HIPCHK(c, hipMemcpy(F.d, host, n * L * F.esz, hipMemcpyHostToDevice));
    } else {
        std::vector<char> tmp(n * L * F.esz);
        if (F.esz == 8) to_soa((const uint64_t*)host, (uint64_t*)tmp.data(), n, L);
        else to_soa((const uint32_t*)host, (uint32_t*)tmp.data(), n, L);
        HIPCHK(c, hipMemcpy(F.d, tmp.data(), tmp.size(), hipMemcpyHostToDevice));
    }
    if (field == OCPX_F_DSEC_DPRI) c->flashed = true;   // dXsdXp supplied by the caller
    c->aux_valid = false;
    c->res_current = false;
    if (field == OCPX_F_FTYPE_IN) {
        const int32_t* fi = (const int32_t*)host;
        bool any = false;
        for (size_t i = 0; i < n; i++) any = any || fi[i] >= 0;
        c->ftype_override = any;
    }
    return OCPX_OK;
}

int ocpx_get_field(ocpx_ctx* c, int field, void* host)
{
    if (!c || !host || field < 0 || field >= OCPX_F_NFIELDS) return fail(c, OCPX_E_ARG, "bad field");
    const Field& F = c->f[field];
    const size_t n = c->nb, L = F.len;
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (F.aos) {
        HIPCHK(c, hipMemcpy(host, F.d, n * L * F.esz, hipMemcpyDeviceToHost));
        return OCPX_OK;
    }
    std::vector<char> tmp(n * L * F.esz);
    HIPCHK(c, hipMemcpy(tmp.data(), F.d, tmp.size(), hipMemcpyDeviceToHost));
    if (F.esz == 8) to_aos((const uint64_t*)tmp.data(), (uint64_t*)host, n, L);
    else to_aos((const uint32_t*)tmp.data(), (uint32_t*)host, n, L);
    return OCPX_OK;
}

int ocpx_field_device_ptr(ocpx_ctx* c, int field, void** dptr)
{
    if (!c || !dptr || field < 0 || field >= OCPX_F_NFIELDS) return fail(c, OCPX_E_ARG, "bad field");
    *dptr = c->f[field].d;
    return OCPX_OK;
}

int ocpx_phase_equilibrium(ocpx_ctx* c, int mode)
{
    if (!c || (mode != OCPX_FLASH_PLAIN && mode != OCPX_FLASH_BATCHED)) return fail(c, OCPX_E_ARG, "bad mode");
    return launch_pe(c, mode);
}

int ocpx_set_flash_mapping(ocpx_ctx* c, int lanes)
{
    if (!c) return OCPX_E_ARG;
    if (lanes == 0) lanes = c->pe_lanes_auto ? c->pe_lanes_auto : default_lanes(c->NC, c->mid != nullptr);
    if (lanes != 1 && lanes != 4 && lanes != 8 && lanes != 16)
        return fail(c, OCPX_E_ARG, "lanes per cell must be 1, 4, 8 or 16");
    if (lanes == 1 && c->mid) return fail(c, OCPX_E_UNSUPPORTED, "contexts of several models need a team mapping");
    if (lanes > 1 && !pe_team_size_ok(c->NC, lanes))
        return fail(c, OCPX_E_ARG, "team size not built for this Nc (4: all; 8: Nc >= 5; 16: Nc >= 9)");
    if (lanes != c->pe_lanes) destroy_graph(c);
    c->pe_lanes = lanes;
    return OCPX_OK;
}

int ocpx_get_flash_mapping(ocpx_ctx* c, int32_t* lanes)
{
    if (!c || !lanes) return OCPX_E_ARG;
    *lanes = c->pe_lanes;
    return OCPX_OK;
}

int ocpx_autotune_flash(ocpx_ctx* c, double* ms)
{
    if (!c) return OCPX_E_ARG;
    static const int cand[4] = {1, 4, 8, 16};
    double t[4] = {0, 0, 0, 0};
    const int keep = c->pe_lanes;
    const bool tm = c->timing;
    c->timing = false;
    int best = keep;
    double best_ms = 0;
    hipEvent_t a = nullptr, b = nullptr;
    HIPCHK(c, hipEventCreate(&a));
    HIPCHK(c, hipEventCreate(&b));
    int rc = OCPX_OK;
    for (int q = 0; q < 4 && !rc; q++) {
        if (cand[q] > 1 && !pe_team_size_ok(c->NC, cand[q])) continue;
        if (cand[q] == 1 && c->mid) continue;
        c->pe_lanes = cand[q];
        for (int rep = 0; rep < 2 && !rc; rep++) {      // first run warms caches and code
            hipEventRecord(a, c->st);
            rc = launch_pe(c, OCPX_FLASH_BATCHED);
            hipEventRecord(b, c->st);
        }
        if (rc) break;
        if (hipEventSynchronize(b) != hipSuccess) { rc = fail(c, OCPX_E_HIP, "autotune sync"); break; }
        float e = 0;
        hipEventElapsedTime(&e, a, b);
        t[q] = e;
        if (best_ms == 0 || e < best_ms) { best_ms = e; best = cand[q]; }
    }
    // the best mapping once more with the cells in natural order: the
    // longest-first queue stays only if it was faster
    double nat = 0;
    const int keepq = c->qmode;
    if (!rc) {
        c->pe_lanes = best;
        c->qmode = 0;
        for (int rep = 0; rep < 2 && !rc; rep++) {
            hipEventRecord(a, c->st);
            rc = launch_pe(c, OCPX_FLASH_BATCHED);
            hipEventRecord(b, c->st);
        }
        if (!rc && hipEventSynchronize(b) != hipSuccess) rc = fail(c, OCPX_E_HIP, "autotune sync");
        if (!rc) {
            float e = 0;
            hipEventElapsedTime(&e, a, b);
            nat = e;
        }
    }
    hipEventDestroy(a);
    hipEventDestroy(b);
    c->timing = tm;
    c->pe_lanes = keep;
    c->qmode = keepq;
    if (rc) return rc;
    if (ms) {
        for (int q = 0; q < 4; q++) ms[q] = t[q];
        ms[4] = nat;
    }
    c->pe_lanes_auto = best;
    c->qmode = nat > 0 && nat < best_ms ? 0 : 1;
    return ocpx_set_flash_mapping(c, best);
}

int ocpx_set_flash_long(ocpx_ctx* c, int32_t evals)
{
    if (!c || evals < 0) return OCPX_E_ARG;
    if (evals != c->long_evals) destroy_graph(c);   // (also marks batch graphs stale)
    c->long_evals = evals;
    return OCPX_OK;
}

int ocpx_get_flash_long(ocpx_ctx* c, int32_t* evals)
{
    if (!c || !evals) return OCPX_E_ARG;
    *evals = c->long_evals;
    return OCPX_OK;
}

int ocpx_set_linear_solver(ocpx_ctx* c, int method)
{
    if (!c || (method != 0 && method != 1)) return OCPX_E_ARG;
    c->lsolver = method;
    return OCPX_OK;
}

int ocpx_get_linear_solver(ocpx_ctx* c, int32_t* method)
{
    if (!c || !method) return OCPX_E_ARG;
    *method = c->lsolver;
    return OCPX_OK;
}

int ocpx_set_flash_stages(ocpx_ctx* c, int on)
{
    if (!c || on < 0 || on > 3) return OCPX_E_ARG;
    if (on != c->stages) destroy_graph(c);
    c->stages = on;
    return OCPX_OK;
}

int ocpx_get_flash_stages(ocpx_ctx* c, int32_t* on)
{
    if (!c || !on) return OCPX_E_ARG;
    *on = c->stages;
    return OCPX_OK;
}

int ocpx_set_flash_queue_order(ocpx_ctx* c, int longest_first)
{
    if (!c || (longest_first != 0 && longest_first != 1)) return OCPX_E_ARG;
    if (longest_first != c->qmode) destroy_graph(c);
    c->qmode = longest_first;
    return OCPX_OK;
}

int ocpx_get_flash_queue_order(ocpx_ctx* c, int32_t* longest_first)
{
    if (!c || !longest_first) return OCPX_E_ARG;
    *longest_first = c->qmode;
    return OCPX_OK;
}

int ocpx_get_pe_result(ocpx_ctx* c, int32_t* NP, double* nu, double* x)
{
    if (!c) return OCPX_E_ARG;
    HIPCHK(c, hipStreamSynchronize(c->st));
    const size_t n = c->nb;
    if (NP) HIPCHK(c, hipMemcpy(NP, c->pe_NP, n * 4, hipMemcpyDeviceToHost));
    std::vector<double> a(n), b(2 * c->NC * n);
    HIPCHK(c, hipMemcpy(a.data(), c->pe_nu, n * 8, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(b.data(), c->pe_x, b.size() * 8, hipMemcpyDeviceToHost));
    std::vector<int32_t> np_(n);
    HIPCHK(c, hipMemcpy(np_.data(), c->pe_NP, n * 4, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; i++) {
        if (nu) { nu[2 * i] = a[i]; nu[2 * i + 1] = np_[i] > 1 ? 1 - a[i] : 0.0; }
        if (x)
            for (int j = 0; j < 2; j++)
                for (int k = 0; k < c->NC; k++) x[(i * 2 + j) * c->NC + k] = b[(size_t)(j * c->NC + k) * n + i];
    }
    return OCPX_OK;
}

int ocpx_flash_der(ocpx_ctx* c, int mode)
{
    if (!c || (mode != OCPX_FLASH_PLAIN && mode != OCPX_FLASH_BATCHED)) return fail(c, OCPX_E_ARG, "bad mode");
    int rc = launch_pe(c, mode, true, true);
    if (rc) return rc;
    rc = launch_tail(c);
    if (rc) return rc;
    c->flashed = true;
    c->aux_valid = false;
    c->res_current = false;
    return OCPX_OK;
}

int ocpx_krpc_der(ocpx_ctx* c)
{
    if (!c) return OCPX_E_ARG;
    return launch_krpc_rock(c);   // rock is cheap and fused; CalRock is idempotent
}

int ocpx_cal_rock(ocpx_ctx* c)
{
    if (!c) return OCPX_E_ARG;
    return launch_krpc_rock(c);
}

int ocpx_cal_res(ocpx_ctx* c, double dt, double* resAbs, ocpx_resnorms* norms)
{
    if (!c) return OCPX_E_ARG;
    int rc = launch_res_asm(c, dt, true, false);
    if (rc) return rc;
    if (resAbs || norms) HIPCHK(c, hipStreamSynchronize(c->st));
    if (resAbs) HIPCHK(c, hipMemcpy(resAbs, c->resAbs, (size_t)c->nb * (c->nc + 1) * 8, hipMemcpyDeviceToHost));
    if (norms) {
        double h[4];
        HIPCHK(c, hipMemcpy(h, c->norms4, sizeof(h), hipMemcpyDeviceToHost));
        norms->maxRelRes_V = h[0]; norms->maxId_V = (int64_t)h[1];
        norms->maxRelRes_N = h[2]; norms->maxId_N = (int64_t)h[3];
    }
    return OCPX_OK;
}

namespace {

void dfree(ocpx_ctx* c, void* p)
{
    if (!p) return;
    for (auto it = c->allocs.begin(); it != c->allocs.end(); ++it)
        if (*it == p) { c->allocs.erase(it); break; }
    hipFree(p);
}

WellIO well_io(ocpx_ctx* c)
{
    WellIO w{};
    w.nwell = c->nwell; w.wells = c->wells; w.cell = c->pcell; w.WI = c->pWI; w.dG = c->pdG; w.xi = c->pxi;
    w.Bnw = c->Bnw; w.Cwn = c->Cwn; w.Dw = c->Dw; w.wres = c->wres; w.qw = c->qw; w.wrhs = c->wrhs;
    w.nperf = c->nperf; w.nucell = c->nucell; w.pwell = c->pwell; w.ucell = c->ucell; w.uptr = c->uptr;
    w.uperf = c->uperf; w.qperf = c->qperf; w.bdiag = c->bdiag; w.dD = c->dD;
    return w;
}

}  // namespace

int ocpx_set_wells(ocpx_ctx* c, int32_t nwell, const ocpx_well* wells, const int64_t* perf_cell,
                   const double* perf_WI, const double* perf_dG, const double* perf_xi)
{
    if (!c) return OCPX_E_ARG;
    if (nwell < 0 || (nwell > 0 && (!wells || !perf_cell || !perf_WI || !perf_dG)))
        return fail(c, OCPX_E_ARG, "ocpx_set_wells: null well or perforation arrays");
    int64_t np_ = 0;
    std::vector<std::pair<int64_t, int32_t>> seen;
    for (int32_t w = 0; w < nwell; w++) {
        const ocpx_well& W = wells[w];
        if ((W.type != OCPX_WELL_INJ && W.type != OCPX_WELL_PROD) || (W.mode != OCPX_WELL_BHP && W.mode != OCPX_WELL_RATE))
            return fail(c, OCPX_E_ARG, "ocpx_set_wells: well " + std::to_string(w) + " has an unknown type or mode");
        if (W.perf0 < 0 || W.nperf < 1)
            return fail(c, OCPX_E_ARG, "ocpx_set_wells: well " + std::to_string(w) + " has no perforation");
        if (W.type == OCPX_WELL_INJ && !perf_xi)
            return fail(c, OCPX_E_ARG, "ocpx_set_wells: injector " + std::to_string(w) + " needs perf_xi");
        np_ = std::max<int64_t>(np_, (int64_t)W.perf0 + W.nperf);
        for (int32_t p = W.perf0; p < W.perf0 + W.nperf; p++) {
            if (perf_cell[p] < 0 || perf_cell[p] >= c->nb)
                return fail(c, OCPX_E_ARG, "ocpx_set_wells: perforation " + std::to_string(p) + " outside the grid");
            seen.push_back({perf_cell[p], w});
        }
    }
    {
        std::vector<int32_t> owner(np_, -1);
        for (int32_t w = 0; w < nwell; w++)
            for (int32_t p = wells[w].perf0; p < wells[w].perf0 + wells[w].nperf; p++) {
                if (owner[p] >= 0)
                    return fail(c, OCPX_E_ARG, "ocpx_set_wells: perforation " + std::to_string(p) + " in two wells");
                owner[p] = w;
            }
        for (int64_t p = 0; p < np_; p++)
            if (owner[p] < 0)
                return fail(c, OCPX_E_ARG, "ocpx_set_wells: perforation " + std::to_string(p) + " in no well");
    }
    std::sort(seen.begin(), seen.end());
    for (size_t k = 1; k < seen.size(); k++)
        if (seen[k] == seen[k - 1])
            return fail(c, OCPX_E_ARG, "ocpx_set_wells: well " + std::to_string(seen[k].second) +
                                           " perforates cell " + std::to_string(seen[k].first) + " twice");
    void* old[20] = {c->wells, c->pcell, c->pWI, c->pdG, c->pxi, c->Bnw, c->Cwn, c->Dw, c->wres, c->qw,
                     c->wrhs, c->dinv_w, c->wrow, c->qperf, c->bdiag, c->dD, c->pwell, c->uptr, c->uperf, c->ucell};
    for (void* p : old) dfree(c, p);
    c->wells = nullptr; c->pcell = nullptr; c->pWI = c->pdG = c->pxi = nullptr;
    c->Bnw = c->Cwn = c->Dw = c->wres = c->qw = c->wrhs = c->dinv_w = nullptr;
    c->qperf = c->bdiag = c->dD = nullptr;
    c->pwell = c->uptr = c->uperf = nullptr;
    c->ucell = nullptr;
    c->nucell = 0;
    c->wrow = nullptr;
    c->nwell = 0; c->nperf = 0;
    c->wells_jac = false;
    c->solved_wells = 0;
    if (nwell == 0) return OCPX_OK;
    const int B = c->nc + 1;
    int rc;
    if ((rc = dalloc(c, &c->wells, nwell)) || (rc = dalloc(c, &c->pcell, np_)) || (rc = dalloc(c, &c->pWI, np_)) ||
        (rc = dalloc(c, &c->pdG, np_)) || (rc = dalloc(c, &c->pxi, np_)) || (rc = dalloc(c, &c->Bnw, np_ * B)) ||
        (rc = dalloc(c, &c->Cwn, np_ * B)) || (rc = dalloc(c, &c->Dw, (size_t)nwell * B * B)) ||
        (rc = dalloc(c, &c->wres, nwell)) || (rc = dalloc(c, &c->qw, (size_t)nwell * c->nc)) ||
        (rc = dalloc(c, &c->wrhs, (size_t)nwell * B)) || (rc = dalloc(c, &c->dinv_w, (size_t)nwell * B * B)) ||
        (rc = dalloc(c, &c->wrow, (size_t)nwell + 1)) || (rc = dalloc(c, &c->qperf, (size_t)np_ * c->nc)) ||
        (rc = dalloc(c, &c->bdiag, (size_t)np_ * B * B)) || (rc = dalloc(c, &c->dD, (size_t)np_)) ||
        (rc = dalloc(c, &c->pwell, (size_t)np_)) || (rc = dalloc(c, &c->uperf, (size_t)np_)))
        return rc;
    {
        // perforation -> well, and the perforated cells with their perforations in order
        std::vector<int32_t> pw, up, upf;
        std::vector<int64_t> uc;
        std::vector<int32_t> w0(nwell), wn(nwell);
        for (int32_t w = 0; w < nwell; w++) { w0[w] = wells[w].perf0; wn[w] = wells[w].nperf; }
        well_layout(nwell, w0.data(), wn.data(), np_, perf_cell, pw, uc, up, upf);
        if ((rc = dalloc(c, &c->ucell, uc.size())) || (rc = dalloc(c, &c->uptr, up.size()))) return rc;
        HIPCHK(c, hipMemcpy(c->pwell, pw.data(), 4 * pw.size(), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->uperf, upf.data(), 4 * upf.size(), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->ucell, uc.data(), 8 * uc.size(), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->uptr, up.data(), 4 * up.size(), hipMemcpyHostToDevice));
        c->nucell = (int32_t)uc.size();
    }
    {
        std::vector<int64_t> wr(nwell + 1);
        for (int32_t w = 0; w <= nwell; w++) wr[w] = w;
        HIPCHK(c, hipMemcpy(c->wrow, wr.data(), 8 * wr.size(), hipMemcpyHostToDevice));
    }
    HIPCHK(c, hipMemcpy(c->wells, wells, sizeof(ocpx_well) * nwell, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->pcell, perf_cell, 8 * np_, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->pWI, perf_WI, 8 * np_, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->pdG, perf_dG, 8 * np_, hipMemcpyHostToDevice));
    if (perf_xi) HIPCHK(c, hipMemcpy(c->pxi, perf_xi, 8 * np_, hipMemcpyHostToDevice));
    c->nwell = nwell;
    c->nperf = (int32_t)np_;
    return OCPX_OK;
}

int ocpx_well_terms(ocpx_ctx* c, double dt, int32_t jacobian, ocpx_resnorms* norms, double* well_res)
{
    if (!c) return OCPX_E_ARG;
    if (c->nwell == 0) return fail(c, OCPX_E_STATE, "ocpx_well_terms: no wells (ocpx_set_wells)");
    if (jacobian && !c->assembled) return fail(c, OCPX_E_STATE, "ocpx_well_terms: the Jacobian needs ocpx_assemble first");
    if (!c->res_current)
        return fail(c, OCPX_E_STATE, "ocpx_well_terms: the state changed since the last ocpx_cal_res / ocpx_assemble");
    if (c->wells_applied)
        return fail(c, OCPX_E_STATE, "ocpx_well_terms: the wells are already in this residual / Jacobian");
    HIPCHK(c, run_well_terms(c->nc, cell_io(c), bsr_io(c), well_io(c), dt, jacobian != 0, c->resAbs, c->st));
    c->wells_applied = true;
    if (jacobian) c->wells_jac = true;
    HIPCHK(c, run_res_norms(c->nc, cell_io(c), c->resAbs, c->bmax, c->barg, c->st));
    HIPCHK(c, run_norms_final(res_asm_blocks(c->nc, c->nb), c->bmax, c->barg, c->norms4, c->st));
    if (norms || well_res) HIPCHK(c, hipStreamSynchronize(c->st));
    if (norms) {
        double h[4];
        HIPCHK(c, hipMemcpy(h, c->norms4, sizeof(h), hipMemcpyDeviceToHost));
        norms->maxRelRes_V = h[0]; norms->maxId_V = (int64_t)h[1];
        norms->maxRelRes_N = h[2]; norms->maxId_N = (int64_t)h[3];
    }
    if (well_res) HIPCHK(c, hipMemcpy(well_res, c->wres, 8 * (size_t)c->nwell, hipMemcpyDeviceToHost));
    return OCPX_OK;
}

int ocpx_get_well_blocks(ocpx_ctx* c, double* Bnw, double* Cwn, double* Dw)
{
    if (!c) return OCPX_E_ARG;
    const size_t B = c->nc + 1;
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (Bnw && c->nperf) HIPCHK(c, hipMemcpy(Bnw, c->Bnw, 8 * B * c->nperf, hipMemcpyDeviceToHost));
    if (Cwn && c->nperf) HIPCHK(c, hipMemcpy(Cwn, c->Cwn, 8 * B * c->nperf, hipMemcpyDeviceToHost));
    if (Dw && c->nwell) HIPCHK(c, hipMemcpy(Dw, c->Dw, 8 * B * B * c->nwell, hipMemcpyDeviceToHost));
    return OCPX_OK;
}

int ocpx_get_wells(ocpx_ctx* c, double* bhp, double* qw)
{
    if (!c) return OCPX_E_ARG;
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (bhp && c->nwell) {
        std::vector<ocpx_well> h(c->nwell);
        HIPCHK(c, hipMemcpy(h.data(), c->wells, sizeof(ocpx_well) * c->nwell, hipMemcpyDeviceToHost));
        for (int32_t w = 0; w < c->nwell; w++) bhp[w] = h[w].bhp;
    }
    if (qw && c->nwell) HIPCHK(c, hipMemcpy(qw, c->qw, 8 * (size_t)c->nwell * c->nc, hipMemcpyDeviceToHost));
    return OCPX_OK;
}

int ocpx_set_well_bhp(ocpx_ctx* c, const double* bhp)
{
    if (!c || !bhp) return OCPX_E_ARG;
    if (c->nwell == 0) return OCPX_OK;
    HIPCHK(c, hipStreamSynchronize(c->st));
    std::vector<ocpx_well> h(c->nwell);
    HIPCHK(c, hipMemcpy(h.data(), c->wells, sizeof(ocpx_well) * c->nwell, hipMemcpyDeviceToHost));
    for (int32_t w = 0; w < c->nwell; w++) h[w].bhp = bhp[w];
    HIPCHK(c, hipMemcpy(c->wells, h.data(), sizeof(ocpx_well) * c->nwell, hipMemcpyHostToDevice));
    return OCPX_OK;
}

double ocpx_peaceman_wi(int32_t dir, double dx, double dy, double dz, double kx, double ky, double kz,
                        double radius, double skin, double kh)
{
    // PeacemanWell::CalWI, src/WellPeaceman.cpp:277-362 (PI and CONV_DARCY as the reference's METRIC constants)
    const double PI_ = 3.141592653, DARCY_ = 0.008527;
    double ro;
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
    } else if (dir == 2) {
        const double kxky = kx * ky, kx_ky = kx / ky;
        ro = 0.28 * pow((dx * dx * pow(1 / kx_ky, 0.5) + dy * dy * pow(kx_ky, 0.5)), 0.5);
        ro /= (pow(kx_ky, 0.25) + pow(1 / kx_ky, 0.25));
        if (kh < 0) kh = (dz * pow(kxky, 0.5));
    } else {
        return -1.0;
    }
    return (2 * PI_) * kh / (log(ro / radius) + skin) * DARCY_;
}

}  // extern "C"

namespace {
template <int NC>
double hc_vm(const DevFluid& f, double P, double T, const double* z)
{
    EosCell<NC> e;
    EosMix<NC> m;
    eos_cell<NC>(f, P, T, e);
    eos_mix<NC>(f, e, z, m);
    double v = m.Z * R_GAS * T / P;
    for (int i = 0; i < NC; i++) v -= z[i] * f.vsh[i];
    return v;
}
}  // namespace

extern "C" {

double ocpx_xi_phase(const ocpx_fluid* f, double P, double T, const double* z, int32_t water)
{
    // OCPMixtureMethodComp01::CalXi (src/OCPMixtureMethodComp.cpp:775-786)
    if (!f || f->nc < 2 || f->nc > 12 || !(P > 0)) return -1.0;
    if (water) {                                   // OCP_PVTW::CalXiW
        const double b = f->PVTW[1] * (1 - f->PVTW[2] * (P - f->PVTW[0]));
        return 1 / b;
    }
    if (!z) return -1.0;
    DevFluid h;
    build_dev_fluid(f, h);
    const double Tk = T + CONV4;
    double v;
    switch (f->nc) {
#define X(N) case N: v = hc_vm<N>(h, P, Tk, z); break;
    X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12)
#undef X
    default: return -1.0;
    }
    return 1 / v;
}

int ocpx_assemble(ocpx_ctx* c, double dt)
{
    if (!c) return OCPX_E_ARG;
    if (!c->flashed) return fail(c, OCPX_E_STATE, "ocpx_assemble before ocpx_flash_der");
    if (!c->aux_valid) return fail(c, OCPX_E_STATE, "ocpx_assemble needs ocpx_krpc_der after the flash");
    int rc = launch_res_asm(c, dt, true, true);
    if (rc) return rc;
    c->assembled = true;
    return OCPX_OK;
}

int ocpx_get_bsr(ocpx_ctx* c, int64_t* nnzb, int64_t* rowptr, int32_t* colind, double* val, double* rhs)
{
    if (!c) return OCPX_E_ARG;
    HIPCHK(c, hipStreamSynchronize(c->st));
    const int B = c->nc + 1;
    if (nnzb) *nnzb = c->nnzb;
    if (rowptr) HIPCHK(c, hipMemcpy(rowptr, c->rowptr, (c->nb + 1) * 8, hipMemcpyDeviceToHost));
    if (colind) HIPCHK(c, hipMemcpy(colind, c->colind, c->nnzb * 4, hipMemcpyDeviceToHost));
    if (val) HIPCHK(c, hipMemcpy(val, c->val, (size_t)c->nnzb * B * B * 8, hipMemcpyDeviceToHost));
    if (rhs) HIPCHK(c, hipMemcpy(rhs, c->rhs, (size_t)c->nb * B * 8, hipMemcpyDeviceToHost));
    return OCPX_OK;
}

int ocpx_bsr_device_ptrs(ocpx_ctx* c, void** rowptr, void** colind, void** val, void** rhs)
{
    if (!c) return OCPX_E_ARG;
    if (rowptr) *rowptr = c->rowptr;
    if (colind) *colind = c->colind;
    if (val) *val = c->val;
    if (rhs) *rhs = c->rhs;
    return OCPX_OK;
}

int ocpx_get_solution(ocpx_ctx* c, const double* u, double dSmax)
{
    if (!c) return OCPX_E_ARG;
    if (!c->flashed) return fail(c, OCPX_E_STATE, "ocpx_get_solution needs dXsdXp from ocpx_flash_der");
    const size_t B = c->nc + 1;
    const double* ud = c->sol;
    if (u) {
        HIPCHK(c, hipStreamSynchronize(c->st));
        HIPCHK(c, hipMemcpy(c->lam, u, c->nb * B * 8, hipMemcpyHostToDevice));
        ud = c->lam;
    } else if (!c->solved) {
        return fail(c, OCPX_E_STATE, "ocpx_get_solution(u = NULL) needs ocpx_linear_solve");
    }
    HIPCHK(c, run_get_solution(c->nc, D(c, OCPX_F_P), D(c, OCPX_F_NI), D(c, OCPX_F_S), D(c, OCPX_F_XIJ),
                               D(c, OCPX_F_DSEC_DPRI), ud, dSmax, c->nb, c->st));
    if (!u && c->solved_wells)
        HIPCHK(c, run_well_bhp_update(c->wells, c->solved_wells, (int)B, c->nb, c->sol, c->st));
    c->aux_valid = false;
    c->solved = false;
    c->res_current = false;
    return OCPX_OK;
}

}  // extern "C"

namespace {

// The pressure AMG and the colouring of the block graph, from the bulk
// pattern and the connections' transmissibilities (once per context and
// transmissibility set).
int cpr_setup(ocpx_ctx* c)
{
    if (c->amg && c->amg->ready()) return OCPX_OK;
    const int64_t nb = c->nb;
    std::vector<int64_t> rp(nb + 1);
    std::vector<int32_t> col(c->nnzb), nbc(c->nnzb);
    std::vector<double> trans(c->nconn > 0 ? c->nconn : 1), w(c->nnzb, 0.0);
    HIPCHK(c, hipStreamSynchronize(c->st));
    HIPCHK(c, hipMemcpy(rp.data(), c->rowptr, (nb + 1) * 8, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(col.data(), c->colind, c->nnzb * 4, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(nbc.data(), c->nbr_conn, c->nnzb * 4, hipMemcpyDeviceToHost));
    if (c->nconn) HIPCHK(c, hipMemcpy(trans.data(), c->trans, c->nconn * 8, hipMemcpyDeviceToHost));
    for (int64_t k = 0; k < c->nnzb; k++)
        if (nbc[k] >= 0) w[k] = std::fabs(trans[nbc[k]]);
    auto amg = std::make_unique<PressureAmg>();
    std::string err;
    if (amg->setup(nb, rp, col, w, err)) return fail(c, OCPX_E_NOMEM, "pressure AMG set-up: " + err);
    c->amg = std::move(amg);
    if (!c->gs_rows) {
        std::vector<int32_t> rows;
        amg::color_rows(nb, rp, col, c->gs_cptr, rows);
        int rc = dalloc(c, &c->gs_rows, rows.size());
        if (rc) return rc;
        HIPCHK(c, hipMemcpy(c->gs_rows, rows.data(), rows.size() * 4, hipMemcpyHostToDevice));
    }
    return OCPX_OK;
}

// FGMRES(30) (Saad, SIAM J. Sci. Comput. 14 (1993) 461-469) on the
// ABF-decoupled system D^{-1} J x = D^{-1} rhs (D the diagonal blocks, the
// well rows' too; src/Decoupling.cpp:23-57), right-preconditioned by two
// stages: the pressure AMG V-cycle on the decoupled pressure block, then one
// multicolour block Gauss-Seidel sweep on the remaining residual.  Classical
// Gram-Schmidt with one reorthogonalisation; stops on the decoupled
// ||r|| / ||b||, as FASP's stop_type 1 on the decoupled system.  x0 = 0,
// the solution in c->sol.
int fgmres_cpr(ocpx_ctx* c, const double* rhs, int64_t len, int32_t nwj, double rtol, int32_t maxit, int& it,
               double& res, bool& conv)
{
    constexpr int M = 30;
    const int B = c->nc + 1;
    const int64_t nb = c->nb;
    hipStream_t st = c->st;
    int rc = cpr_setup(c);
    if (rc) return rc;
    if (c->fg_cap < len) {
        for (double** q : {&c->fg_V, &c->fg_Z, &c->fg_w, &c->fg_t, &c->fg_x1, &c->fg_r2}) { dfree(c, *q); *q = nullptr; }
        if ((rc = dalloc(c, &c->fg_V, (size_t)(M + 1) * len)) || (rc = dalloc(c, &c->fg_Z, (size_t)M * len)) ||
            (rc = dalloc(c, &c->fg_w, (size_t)len)) || (rc = dalloc(c, &c->fg_t, (size_t)len)) ||
            (rc = dalloc(c, &c->fg_x1, (size_t)len)) || (rc = dalloc(c, &c->fg_r2, (size_t)len)))
            return rc;
        c->fg_cap = len;
    }
    if (!c->fg_pb) {
        if ((rc = dalloc(c, &c->fg_pb, (size_t)nb)) || (rc = dalloc(c, &c->fg_px, (size_t)nb)) ||
            (rc = dalloc(c, &c->fg_h, 64)) || (rc = dalloc(c, &c->fg_part, (size_t)mdot_scratch_len())) ||
            (rc = dalloc(c, &c->fg_out, 64)))
            return rc;
    }
    const BsrIO io = bsr_io(c);
    const WellIO wio = well_io(c);
    double *V = c->fg_V, *Z = c->fg_Z, *w = c->fg_w, *t = c->fg_t, *x = c->sol;
    // y = D^{-1} J x
    auto aop = [&](const double* in, double* out) -> int {
        {
            TimerScope ts(c, 8);
            HIPCHK(c, run_bsr_spmv(B, io.rowptr, io.colind, io.val, in, t, nb, st));
        }
        if (nwj) HIPCHK(c, run_well_spmv(wio, B, nb, in, t, st));
        HIPCHK(c, run_bsr_spmv(B, nullptr, nullptr, c->dinv, t, out, nb, st));
        if (nwj) HIPCHK(c, run_bsr_spmv(B, nullptr, nullptr, c->dinv_w, t + nb * B, out + nb * B, nwj, st));
        return OCPX_OK;
    };
    // z = M^{-1} v
    auto precond = [&](const double* v, double* z) -> int {
        HIPCHK(c, run_gather_p(B, v, c->fg_pb, nb, st));
        HIPCHK(c, c->amg->vcycle(c->fg_pb, c->fg_px, st));
        HIPCHK(c, run_scatter_p(B, c->fg_px, c->fg_x1, nb, len, st));
        int rr = aop(c->fg_x1, c->fg_r2);
        if (rr) return rr;
        HIPCHK(c, run_axpbypcz(c->fg_r2, v, -1.0, c->fg_r2, 0, nullptr, len, st));    // r2 = v - A x1
        // z = r2 on every row, then the colours of the bulk rows in turn
        // (rows of one colour read only rows of other colours)
        HIPCHK(c, run_axpbypcz(z, nullptr, 0, nullptr, 0, nullptr, nb * B, st));
        if (nwj) HIPCHK(c, run_axpbypcz(z + nb * B, c->fg_r2 + nb * B, 0, nullptr, 0, nullptr, (int64_t)nwj * B, st));
        for (size_t k = 0; k + 1 < c->gs_cptr.size(); k++)
            HIPCHK(c, run_gs_color(B, c->gs_rows + c->gs_cptr[k], c->gs_cptr[k + 1] - c->gs_cptr[k], io.rowptr,
                                   io.colind, io.val, c->dinv, c->fg_r2, z, st));
        HIPCHK(c, run_axpbypcz(z, z, 1.0, c->fg_x1, 0, nullptr, len, st));
        return OCPX_OK;
    };
    auto dots = [&](const double* Vb, int m, const double* vec, double* out) -> int {
        HIPCHK(c, run_mdot(Vb, len, m, vec, len, c->fg_part, c->fg_out, st));
        HIPCHK(c, hipMemcpyAsync(out, c->fg_out, (size_t)m * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(c, hipStreamSynchronize(st));
        return OCPX_OK;
    };
    HIPCHK(c, c->amg->values(B, io.val, c->dinv, st));
    // b = D^{-1} rhs into V_0's slot for now; x = 0
    double* bh = c->fg_r2;
    HIPCHK(c, run_bsr_spmv(B, nullptr, nullptr, c->dinv, rhs, V, nb, st));
    if (nwj) HIPCHK(c, run_bsr_spmv(B, nullptr, nullptr, c->dinv_w, rhs + nb * B, V + nb * B, nwj, st));
    HIPCHK(c, run_axpbypcz(x, nullptr, 0, nullptr, 0, nullptr, len, st));
    double bb = 0;
    if ((rc = dots(V, 1, V, &bb))) return rc;
    const double bnorm = std::sqrt(bb);
    double beta = bnorm;
    it = 0;
    res = 0;
    conv = bnorm == 0;
    std::vector<double> H((M + 1) * M), cs(M), sn(M), g(M + 1), y(M), h(M + 1), h2(M + 1);
    while (!conv && it < maxit && std::isfinite(beta) && beta > 0) {
        HIPCHK(c, run_axpbypcz(V, nullptr, 1.0 / beta, V, 0, nullptr, len, st));        // V_0 = r / beta
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int k = 0;
        bool stop = false;
        for (int j = 0; j < M && it < maxit; j++) {
            double* Vj = V + (size_t)j * len;
            double* Zj = Z + (size_t)j * len;
            if ((rc = precond(Vj, Zj)) || (rc = aop(Zj, w))) return rc;
            if ((rc = dots(V, j + 1, w, h.data()))) return rc;
            for (int i = 0; i <= j; i++) h2[i] = -h[i];
            HIPCHK(c, hipMemcpyAsync(c->fg_h, h2.data(), (size_t)(j + 1) * 8, hipMemcpyHostToDevice, st));
            HIPCHK(c, run_maxpy(w, V, len, j + 1, c->fg_h, len, st));
            if ((rc = dots(V, j + 1, w, h2.data()))) return rc;
            for (int i = 0; i <= j; i++) { h[i] += h2[i]; h2[i] = -h2[i]; }
            HIPCHK(c, hipMemcpyAsync(c->fg_h, h2.data(), (size_t)(j + 1) * 8, hipMemcpyHostToDevice, st));
            HIPCHK(c, run_maxpy(w, V, len, j + 1, c->fg_h, len, st));
            double ww = 0;
            HIPCHK(c, run_dot(w, w, len, c->dpart, c->dout, st));
            HIPCHK(c, hipMemcpyAsync(&ww, c->dout, sizeof(double), hipMemcpyDeviceToHost, st));
            HIPCHK(c, hipStreamSynchronize(st));
            const double hn = std::sqrt(ww);
            for (int i = 0; i <= j; i++) H[i * M + j] = h[i];
            H[(j + 1) * M + j] = hn;
            for (int i = 0; i < j; i++) {             // earlier rotations
                const double a = H[i * M + j], b2 = H[(i + 1) * M + j];
                H[i * M + j] = cs[i] * a + sn[i] * b2;
                H[(i + 1) * M + j] = -sn[i] * a + cs[i] * b2;
            }
            const double a = H[j * M + j], b2 = H[(j + 1) * M + j];
            const double den = std::hypot(a, b2);
            cs[j] = den > 0 ? a / den : 1.0;
            sn[j] = den > 0 ? b2 / den : 0.0;
            H[j * M + j] = den;
            H[(j + 1) * M + j] = 0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            it++;
            k = j + 1;
            res = std::fabs(g[j + 1]) / bnorm;
            if (!std::isfinite(res)) return fail(c, OCPX_E_HIP, "FGMRES: non-finite residual");
            if (res < rtol || hn == 0) { stop = true; break; }
            HIPCHK(c, run_axpbypcz(V + (size_t)(j + 1) * len, nullptr, 1.0 / hn, w, 0, nullptr, len, st));
        }
        for (int i = k - 1; i >= 0; i--) {           // H y = g, upper triangular
            double v = g[i];
            for (int q = i + 1; q < k; q++) v -= H[i * M + q] * y[q];
            y[i] = H[i * M + i] != 0 ? v / H[i * M + i] : 0.0;
        }
        HIPCHK(c, hipMemcpyAsync(c->fg_h, y.data(), (size_t)k * 8, hipMemcpyHostToDevice, st));
        HIPCHK(c, run_maxpy(x, Z, len, k, c->fg_h, len, st));
        // the true decoupled residual: V_0 = D^{-1} rhs - D^{-1} J x
        if ((rc = aop(x, w))) return rc;
        HIPCHK(c, run_bsr_spmv(B, nullptr, nullptr, c->dinv, rhs, bh, nb, st));
        if (nwj) HIPCHK(c, run_bsr_spmv(B, nullptr, nullptr, c->dinv_w, rhs + nb * B, bh + nb * B, nwj, st));
        HIPCHK(c, run_axpbypcz(V, bh, -1.0, w, 0, nullptr, len, st));
        double rr2 = 0;
        if ((rc = dots(V, 1, V, &rr2))) return rc;
        beta = std::sqrt(rr2);
        res = beta / bnorm;
        if (res < rtol) conv = true;
        if (stop && !conv && k == 0) break;
    }
    return OCPX_OK;
}

}  // namespace

extern "C" {

// BiCGStab, right-preconditioned with block Jacobi, x0 = 0, stopping on
// ||r|| / ||b|| (the reference's stop_type 1); scalars come back to the host
// once per dot product.  Or (the default) FGMRES with the CPR-type
// preconditioner, fgmres_cpr.
int ocpx_linear_solve(ocpx_ctx* c, double rtol, int32_t maxit, double* u, int32_t* iters, double* relres)
{
    if (!c || !(rtol > 0) || maxit <= 0) return OCPX_E_ARG;
    if (!c->assembled) return fail(c, OCPX_E_STATE, "ocpx_linear_solve needs ocpx_assemble");
    const int B = c->nc + 1;
    const int32_t nwj = c->wells_jac ? c->nwell : 0;        // well rows after the nb bulk rows
    const int64_t nb = c->nb, len = (nb + nwj) * B;
    int rc = OCPX_OK;
    if (!c->dinv) {
        if ((rc = dalloc(c, &c->dinv, (size_t)nb * B * B)) || (rc = dalloc(c, &c->dpart, (size_t)dot_scratch_len())) ||
            (rc = dalloc(c, &c->dout, 1)))
            return rc;
    }
    if (c->sol_cap < len) {
        dfree(c, c->sol); dfree(c, c->rhsf);
        for (auto& v : c->kv) { dfree(c, v); v = nullptr; }
        c->sol = c->rhsf = nullptr;
        if ((rc = dalloc(c, &c->sol, (size_t)len)) || (rc = dalloc(c, &c->rhsf, (size_t)len))) return rc;
        for (auto& v : c->kv)
            if ((rc = dalloc(c, &v, (size_t)len))) return rc;
        c->sol_cap = len;
    }
    double *r = c->kv[0], *rh = c->kv[1], *p = c->kv[2], *ph = c->kv[3], *v = c->kv[4], *s = c->kv[5],
           *sh = c->kv[6], *t = c->kv[7], *x = c->sol;
    const BsrIO io = bsr_io(c);
    const WellIO wio = well_io(c);
    hipStream_t st = c->st;
    TimerScope whole(c, 7);
    auto dot = [&](const double* a, const double* b, double& out) -> int {
        HIPCHK(c, run_dot(a, b, len, c->dpart, c->dout, st));
        HIPCHK(c, hipMemcpyAsync(&out, c->dout, sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(c, hipStreamSynchronize(st));
        return OCPX_OK;
    };
    auto spmv = [&](const double* in, double* outv) -> int {
        TimerScope ts(c, 8);
        HIPCHK(c, run_bsr_spmv(B, io.rowptr, io.colind, io.val, in, outv, nb, st));
        if (nwj) HIPCHK(c, run_well_spmv(wio, B, nb, in, outv, st));
        return OCPX_OK;
    };
    auto precond = [&](const double* in, double* outv) -> int {
        HIPCHK(c, run_bsr_spmv(B, nullptr, nullptr, c->dinv, in, outv, nb, st));
        if (nwj) HIPCHK(c, run_bsr_spmv(B, nullptr, nullptr, c->dinv_w, in + nb * B, outv + nb * B, nwj, st));
        return OCPX_OK;
    };
    const double* rhs = io.rhs;
    if (nwj) {                                               // [bulk rhs; well rhs]
        HIPCHK(c, hipMemcpyAsync(c->rhsf, io.rhs, (size_t)nb * B * 8, hipMemcpyDeviceToDevice, st));
        HIPCHK(c, hipMemcpyAsync(c->rhsf + nb * B, c->wrhs, (size_t)nwj * B * 8, hipMemcpyDeviceToDevice, st));
        HIPCHK(c, run_bsr_diag_inv(B, c->wrow, c->Dw, nwj, c->dinv_w, st));
        rhs = c->rhsf;
    }
    HIPCHK(c, run_bsr_diag_inv(B, io.rowptr, io.val, nb, c->dinv, st));
    if (c->lsolver == 1) {
        int it = 0;
        double res = 0;
        bool conv = false;
        if ((rc = fgmres_cpr(c, rhs, len, nwj, rtol, maxit, it, res, conv))) return rc;
        if (iters) *iters = conv ? it : -std::max(it, 1);
        if (relres) *relres = res;
        if (u) {
            HIPCHK(c, hipMemcpyAsync(u, x, (size_t)len * sizeof(double), hipMemcpyDeviceToHost, st));
            HIPCHK(c, hipStreamSynchronize(st));
        }
        c->solved = conv;
        c->solved_wells = conv ? nwj : 0;
        return OCPX_OK;
    }
    HIPCHK(c, run_axpbypcz(x, nullptr, 0, nullptr, 0, nullptr, len, st));
    HIPCHK(c, run_axpbypcz(r, rhs, 0, nullptr, 0, nullptr, len, st));
    HIPCHK(c, run_axpbypcz(rh, rhs, 0, nullptr, 0, nullptr, len, st));
    HIPCHK(c, run_axpbypcz(p, nullptr, 0, nullptr, 0, nullptr, len, st));
    HIPCHK(c, run_axpbypcz(v, nullptr, 0, nullptr, 0, nullptr, len, st));
    double bb = 0;
    if ((rc = dot(rhs, rhs, bb))) return rc;
    const double bnorm = std::sqrt(bb);
    int it = 0;
    bool conv = bnorm == 0;
    double res = 0, rho = 1, alpha = 1, omega = 1;
    while (!conv && it < maxit && std::isfinite(bnorm)) {
        it++;
        double rho1 = 0, rv = 0, ss = 0, tt = 0, ts = 0, rr = 0;
        if ((rc = dot(rh, r, rho1))) return rc;
        if (rho1 == 0 || !std::isfinite(rho1)) break;                     // breakdown
        const double beta = (rho1 / rho) * (alpha / omega);
        HIPCHK(c, run_axpbypcz(p, r, beta, p, -beta * omega, v, len, st));  // p = r + beta (p - omega v)
        if ((rc = precond(p, ph)) || (rc = spmv(ph, v)) || (rc = dot(rh, v, rv))) return rc;
        if (rv == 0 || !std::isfinite(rv)) break;
        alpha = rho1 / rv;
        HIPCHK(c, run_axpbypcz(s, r, -alpha, v, 0, nullptr, len, st));    // s = r - alpha v
        if ((rc = dot(s, s, ss))) return rc;
        if (std::sqrt(ss) / bnorm < rtol) {
            HIPCHK(c, run_axpbypcz(x, x, alpha, ph, 0, nullptr, len, st));
            res = std::sqrt(ss) / bnorm;
            conv = true;
            break;
        }
        if ((rc = precond(s, sh)) || (rc = spmv(sh, t)) || (rc = dot(t, t, tt)) || (rc = dot(t, s, ts))) return rc;
        if (tt == 0 || !std::isfinite(tt)) break;
        omega = ts / tt;
        HIPCHK(c, run_axpbypcz(x, x, alpha, ph, omega, sh, len, st));    // x += alpha ph + omega sh
        HIPCHK(c, run_axpbypcz(r, s, -omega, t, 0, nullptr, len, st));   // r = s - omega t
        if ((rc = dot(r, r, rr))) return rc;
        res = std::sqrt(rr) / bnorm;
        if (res < rtol) { conv = true; break; }
        if (omega == 0 || !std::isfinite(res)) break;
        rho = rho1;
    }
    if (iters) *iters = conv ? it : -std::max(it, 1);        // negative on failure, as LinearSystem::Solve
    if (relres) *relres = res;
    if (u) {
        HIPCHK(c, hipMemcpyAsync(u, x, (size_t)len * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(c, hipStreamSynchronize(st));
    }
    // a solve that did not converge leaves nothing to apply (the reference cuts dt and
    // resets instead, src/IsoThermalMethod.cpp:776-781)
    c->solved = conv;
    c->solved_wells = conv ? nwj : 0;
    return OCPX_OK;
}

}  // extern "C"

namespace {
// The bulk fields IsoT_FIM::UpdateLastTimeStep / ResetToLastTimeStep carry
// (src/IsoThermalMethod.cpp:1420-1515) plus the SkipPSA variable set
// (src/AcceleratePEC.cpp:15-31).  Ni's last-step copy is the lNi field itself.
const int kLastStepFields[] = {
    OCPX_F_PORO, OCPX_F_POROP, OCPX_F_ROCKVP, OCPX_F_NT, OCPX_F_VF, OCPX_F_P, OCPX_F_PJ,
    OCPX_F_PC, OCPX_F_PHASEEXIST, OCPX_F_S, OCPX_F_XIJ, OCPX_F_RHO, OCPX_F_XI, OCPX_F_MU,
    OCPX_F_KR, OCPX_F_VFP, OCPX_F_VFI, OCPX_F_RHOP, OCPX_F_RHOX, OCPX_F_XIP, OCPX_F_XIX,
    OCPX_F_MUP, OCPX_F_MUX, OCPX_F_DPCDS, OCPX_F_DKRDS, OCPX_F_DSEC_DPRI, OCPX_F_SKIP_FLAG,
    OCPX_F_SKIP_MINEIG, OCPX_F_SKIP_P, OCPX_F_SKIP_T, OCPX_F_SKIP_ZI,
    OCPX_F_ITERS};   // the counters route the next staged flash's long cells

size_t field_bytes(const ocpx_ctx* c, int fid) { return (size_t)c->nb * c->f[fid].len * c->f[fid].esz; }
}  // namespace

extern "C" {

int ocpx_update_last_step(ocpx_ctx* c)
{
    if (!c) return OCPX_E_ARG;
    for (int fid : kLastStepFields) {
        if (!c->lf[fid].d) {
            char* p = nullptr;
            int rc = dalloc(c, &p, field_bytes(c, fid));
            if (rc) return rc;
            c->lf[fid] = c->f[fid];
            c->lf[fid].d = p;
        }
        HIPCHK(c, hipMemcpyAsync(c->lf[fid].d, c->f[fid].d, field_bytes(c, fid), hipMemcpyDeviceToDevice, c->st));
    }
    HIPCHK(c, hipMemcpyAsync(c->f[OCPX_F_LNI].d, c->f[OCPX_F_NI].d, field_bytes(c, OCPX_F_NI),
                             hipMemcpyDeviceToDevice, c->st));
    c->have_last = true;
    return OCPX_OK;
}

int ocpx_reset_to_last_step(ocpx_ctx* c)
{
    if (!c) return OCPX_E_ARG;
    if (!c->have_last) return fail(c, OCPX_E_STATE, "ocpx_reset_to_last_step before ocpx_update_last_step");
    c->aux_valid = false;
    c->res_current = false;
    for (int fid : kLastStepFields)
        HIPCHK(c, hipMemcpyAsync(c->f[fid].d, c->lf[fid].d, field_bytes(c, fid), hipMemcpyDeviceToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->f[OCPX_F_NI].d, c->f[OCPX_F_LNI].d, field_bytes(c, OCPX_F_NI),
                             hipMemcpyDeviceToDevice, c->st));
    return OCPX_OK;
}

int ocpx_set_geometry(ocpx_ctx* c, const double* v, const double* trans)
{
    if (!c) return OCPX_E_ARG;
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (v) HIPCHK(c, hipMemcpy(c->v, v, (size_t)c->nb * 8, hipMemcpyHostToDevice));
    if (trans && c->nconn) HIPCHK(c, hipMemcpy(c->trans, trans, (size_t)c->nconn * 8, hipMemcpyHostToDevice));
    if (trans) c->amg.reset();          // the pressure AMG's aggregates follow the transmissibilities
    c->aux_valid = false;
    c->res_current = false;
    return OCPX_OK;
}

int ocpx_newton_pass(ocpx_ctx* c, double dt, int mode, int use_graph)
{
    if (!c || (mode != OCPX_FLASH_PLAIN && mode != OCPX_FLASH_BATCHED)) return fail(c, OCPX_E_ARG, "bad mode");
    if (!use_graph) {
        int rc = ocpx_flash_der(c, mode);
        if (rc) return rc;
        rc = launch_krpc_rock(c);
        if (rc) return rc;
        rc = launch_res_asm(c, dt, true, true);
        if (rc) return rc;
        c->assembled = true;
        return OCPX_OK;
    }
    if (c->own_stream == false && c->st == nullptr) return fail(c, OCPX_E_STATE, "graph capture needs a stream");
    if (!c->gexec || c->gdt != dt || c->gmode != mode) {
        destroy_graph(c);
        if (mode == OCPX_FLASH_BATCHED && c->stages) {   // no allocation inside the capture
            int rc = sta_alloc(c);
            if (rc) return rc;
        }
        const bool tm = c->timing;
        c->timing = false;
        HIPCHK(c, hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
        int rc = launch_pe(c, mode, true, true);
        if (!rc) rc = launch_tail(c);
        if (!rc) rc = launch_krpc_rock(c);
        if (!rc) rc = launch_res_asm(c, dt, true, true);
        hipGraph_t gph = nullptr;
        hipError_t e = hipStreamEndCapture(c->st, &gph);
        c->timing = tm;
        if (rc) { if (gph) hipGraphDestroy(gph); return rc; }
        if (e != hipSuccess) return fail(c, OCPX_E_HIP, std::string("capture: ") + hipGetErrorString(e));
        c->graph = gph;
        HIPCHK(c, hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
        c->gdt = dt;
        c->gmode = mode;
    }
    HIPCHK(c, hipGraphLaunch(c->gexec, c->st));
    c->flashed = c->assembled = true;
    c->wells_jac = false;
    c->wells_applied = false;
    c->res_current = true;
    return OCPX_OK;
}

}  // extern "C"

// Several contexts whose Newton passes are captured into one graph: each
// context's flash + kr/Pc + rock + residual + assembly is a branch forked
// from and joined to the batch stream, so one graph launch runs every
// context's pass, the branches free to overlap on the device.
struct ocpx_batch {
    std::vector<ocpx_ctx*> ctx;
    std::vector<uint64_t> gen;   // the contexts' settings generations at capture
    hipStream_t st = nullptr;
    bool own_stream = false;
    hipEvent_t fork = nullptr, done = nullptr;
    std::vector<hipEvent_t> ev;  // per context: its join into the capture, its work before a launch
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    double gdt = 0;
};

namespace {

void batch_drop_graph(ocpx_batch* b)
{
    if (b->gexec) hipGraphExecDestroy(b->gexec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
}

int batch_fail(int code, const std::string& msg)
{
    set_call_error("ocpx_batch: " + msg);
    return code;
}

#define BCHK(call)                                                                                 \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) return batch_fail(OCPX_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

int batch_capture(ocpx_batch* b, double dt)
{
    batch_drop_graph(b);
    for (ocpx_ctx* c : b->ctx)
        if (c->stages) {                  // no allocation inside the capture
            int rc = sta_alloc(c);
            if (rc) return batch_fail(rc, c->err);
        }
    std::vector<bool> tm(b->ctx.size());
    for (size_t k = 0; k < b->ctx.size(); k++) { tm[k] = b->ctx[k]->timing; b->ctx[k]->timing = false; }
    BCHK(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipEventRecord(b->fork, b->st);
    int rc = e == hipSuccess ? OCPX_OK : OCPX_E_HIP;
    std::string why = e == hipSuccess ? "" : std::string("fork: ") + hipGetErrorString(e);
    for (size_t k = 0; k < b->ctx.size() && !rc; k++) {
        ocpx_ctx* c = b->ctx[k];
        if ((e = hipStreamWaitEvent(c->st, b->fork, 0)) != hipSuccess) {
            rc = OCPX_E_HIP;
            why = std::string("branch: ") + hipGetErrorString(e);
            break;
        }
        // the long-cell branch stays off here: its fork from a context stream
        // that is itself a branch of this capture is not supported (the
        // contexts' branches of the batch overlap on the device instead)
        rc = launch_pe(c, OCPX_FLASH_BATCHED, false);
        if (!rc) rc = launch_tail(c);
        if (!rc) rc = launch_krpc_rock(c);
        if (!rc) rc = launch_res_asm(c, dt, true, true);
        if (rc) { why = c->err; break; }
        if ((e = hipEventRecord(b->ev[k], c->st)) != hipSuccess || (e = hipStreamWaitEvent(b->st, b->ev[k], 0)) != hipSuccess) {
            rc = OCPX_E_HIP;
            why = std::string("join: ") + hipGetErrorString(e);
        }
    }
    hipGraph_t gph = nullptr;
    e = hipStreamEndCapture(b->st, &gph);
    for (size_t k = 0; k < b->ctx.size(); k++) b->ctx[k]->timing = tm[k];
    if (rc) { if (gph) hipGraphDestroy(gph); return batch_fail(rc, why); }
    if (e != hipSuccess) return batch_fail(OCPX_E_HIP, std::string("capture: ") + hipGetErrorString(e));
    b->graph = gph;
    BCHK(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
    b->gdt = dt;
    for (size_t k = 0; k < b->ctx.size(); k++) b->gen[k] = b->ctx[k]->gen;
    return OCPX_OK;
}

}  // namespace

extern "C" {

int ocpx_batch_create(ocpx_ctx* const* ctxs, int32_t n, void* hip_stream, ocpx_batch** out)
{
    clear_call_error();
    if (!ctxs || n <= 0 || !out) return batch_fail(OCPX_E_ARG, "needs at least one context and an output");
    for (int32_t k = 0; k < n; k++) {
        if (!ctxs[k]) return batch_fail(OCPX_E_ARG, "null context");
        for (int32_t q = 0; q < k; q++)
            if (ctxs[q] == ctxs[k]) return batch_fail(OCPX_E_ARG, "a context appears twice");
        if (ctxs[k]->mid && ctxs[k]->pe_lanes == 1)
            return batch_fail(OCPX_E_UNSUPPORTED, "contexts of several models flash with a team mapping");
    }
    auto* b = new ocpx_batch();
    b->ctx.assign(ctxs, ctxs + n);
    b->gen.assign(n, ~0ull);
    b->ev.assign(n, nullptr);
    b->st = (hipStream_t)hip_stream;
    bool ok = true;
    if (!b->st) {
        ok = hipStreamCreate(&b->st) == hipSuccess;
        b->own_stream = ok;
    }
    ok = ok && hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) == hipSuccess &&
         hipEventCreateWithFlags(&b->done, hipEventDisableTiming) == hipSuccess;
    for (int32_t k = 0; k < n && ok; k++) ok = hipEventCreateWithFlags(&b->ev[k], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        ocpx_batch_destroy(b);
        return batch_fail(OCPX_E_HIP, "stream or event creation failed");
    }
    *out = b;
    return OCPX_OK;
}

int ocpx_batch_newton_pass(ocpx_batch* b, double dt)
{
    clear_call_error();
    if (!b) return batch_fail(OCPX_E_ARG, "null batch");
    bool stale = !b->gexec || b->gdt != dt;
    for (size_t k = 0; k < b->ctx.size(); k++) stale = stale || b->gen[k] != b->ctx[k]->gen;
    if (stale) {
        int rc = batch_capture(b, dt);
        if (rc) return rc;
    }
    // the launch follows what was queued on each context's stream, and each
    // context's later work follows the launch
    for (size_t k = 0; k < b->ctx.size(); k++) {
        BCHK(hipEventRecord(b->ev[k], b->ctx[k]->st));
        BCHK(hipStreamWaitEvent(b->st, b->ev[k], 0));
    }
    BCHK(hipGraphLaunch(b->gexec, b->st));
    BCHK(hipEventRecord(b->done, b->st));
    for (ocpx_ctx* c : b->ctx) {
        BCHK(hipStreamWaitEvent(c->st, b->done, 0));
        c->flashed = c->assembled = true;
        c->wells_jac = false;
        c->wells_applied = false;
        c->res_current = true;
    }
    return OCPX_OK;
}

void ocpx_batch_destroy(ocpx_batch* b)
{
    if (!b) return;
    if (b->st) hipStreamSynchronize(b->st);
    batch_drop_graph(b);
    if (b->fork) hipEventDestroy(b->fork);
    if (b->done) hipEventDestroy(b->done);
    for (hipEvent_t e : b->ev)
        if (e) hipEventDestroy(e);
    if (b->own_stream) hipStreamDestroy(b->st);
    delete b;
}

int ocpx_jt_lambda_dev(ocpx_ctx* c, const double* lam, double* out)
{
    if (!c || !lam || !out) return OCPX_E_ARG;
    if (!c->assembled) return fail(c, OCPX_E_STATE, "ocpx_jt_lambda before ocpx_assemble");
    TimerScope t(c, 5);
    HIPCHK(c, run_jt_lambda(c->nc, c->nb, bsr_io(c), lam, out, c->st));
    return OCPX_OK;
}

int ocpx_jt_lambda(ocpx_ctx* c, const double* lam, double* out)
{
    if (!c || !lam || !out) return OCPX_E_ARG;
    const size_t B = c->nc + 1;
    HIPCHK(c, hipStreamSynchronize(c->st));
    HIPCHK(c, hipMemcpy(c->lam, lam, c->nb * B * 8, hipMemcpyHostToDevice));
    int rc = ocpx_jt_lambda_dev(c, c->lam, c->out);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->st));
    HIPCHK(c, hipMemcpy(out, c->out, c->nb * B * 8, hipMemcpyDeviceToHost));
    return OCPX_OK;
}

int ocpx_dR_dparams_dev(ocpx_ctx* c, double dt, const double* lam, double* dT, double* dVp)
{
    if (!c || !lam) return OCPX_E_ARG;
    if (!c->flashed) return fail(c, OCPX_E_STATE, "ocpx_dR_dparams before ocpx_flash_der");
    TimerScope t(c, 6);
    HIPCHK(c, run_dR_dtrans(c->nc, cell_io(c), conn_io(c), dt, lam, dT ? dT : c->gT, dVp ? dVp : c->gV, c->st));
    return OCPX_OK;
}

int ocpx_dR_dparams(ocpx_ctx* c, double dt, const double* lam, double* dT, double* dVp)
{
    if (!c || !lam) return OCPX_E_ARG;
    const size_t B = c->nc + 1;
    HIPCHK(c, hipStreamSynchronize(c->st));
    HIPCHK(c, hipMemcpy(c->lam, lam, c->nb * B * 8, hipMemcpyHostToDevice));
    int rc = ocpx_dR_dparams_dev(c, dt, c->lam, c->gT, c->gV);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (dT && c->nconn) HIPCHK(c, hipMemcpy(dT, c->gT, c->nconn * 8, hipMemcpyDeviceToHost));
    if (dVp) HIPCHK(c, hipMemcpy(dVp, c->gV, c->nb * 8, hipMemcpyDeviceToHost));
    return OCPX_OK;
}

int ocpx_timing_enable(ocpx_ctx* c, int on)
{
    if (!c) return OCPX_E_ARG;
    HIPCHK(c, hipStreamSynchronize(c->st));
    for (auto& t : c->tacc) {
        for (auto& p : t.pending) { hipEventDestroy(p.first); hipEventDestroy(p.second); }
        t.pending.clear();
        t.total_ms = 0;
        t.calls = 0;
    }
    c->timing = on != 0;
    return OCPX_OK;
}

int ocpx_timing_get(ocpx_ctx* c, int which, double* avg_ms, int64_t* calls)
{
    if (!c || which < 0 || which > 8) return OCPX_E_ARG;
    TimingAcc& t = c->tacc[which];
    for (auto& p : t.pending) {
        HIPCHK(c, hipEventSynchronize(p.second));
        float ms = 0;
        HIPCHK(c, hipEventElapsedTime(&ms, p.first, p.second));
        t.total_ms += ms;
        t.calls++;
        hipEventDestroy(p.first);
        hipEventDestroy(p.second);
    }
    t.pending.clear();
    if (avg_ms) *avg_ms = t.calls ? t.total_ms / t.calls : 0.0;
    if (calls) *calls = t.calls;
    return OCPX_OK;
}

}  // extern "C"
