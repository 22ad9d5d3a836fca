// pe_sta.hip — the strict stability test of the batched flash as its own
// stages, cells grouped by what they are doing.
//
// OCPPhaseEquilibrium::StableSSM (src/OCPPhaseEquilibrium.cpp:397-473) runs
// four trial compositions in turn, each an SSM loop to SSMSTA's tolerance or
// cap followed by StableNR (:475-521); the first trial whose NR converges
// with Yt > 1 + eYt declares the feed unstable.  The SSM loops are most of a
// flash's work, and they share one short body.  So:
//
//   stage 1  pe_sta_ssm   persistent lanes, one cell each: SkipPSA's flash
//                         type, the feed's ln phi, then the trials' SSM loops
//                         only.  A trial whose residual is already below
//                         NRSTA's tolerance needs no Newton step and is
//                         decided on the spot; any other is queued for stage 2
//                         and the lane carries on with the next trial.  A lane
//                         stops at the first trial it decides unstable.  Cells
//                         that take another path (flash type 1 or 2, or the
//                         relaxed test after a step with phases) are queued
//                         for the general kernel at once.
//   stage 2  pe_sta_nr    StableNR for the queued trials, one lane each.
//   stage 3  pe_sta_join  per cell, the first trial in order that is unstable
//                         (the reference's early exit), the counters of the
//                         trials up to it; stable cells are written out,
//                         unstable ones queued for the split with that
//                         trial's Yt.
//   stage 4  pe_sta_split persistent lanes, one unstable cell each:
//                         SplitSSM2's first pass (Wilson K, RachfordRice2,
//                         UpdateXRR, ln phi of both phases, K update) to
//                         SSMSP's tolerance or cap.  The general kernel
//                         resumes every such cell at SplitNR and carries on
//                         (Newton steps, CheckSplit, the restart loop).
//   stage 5  pe_sta_snr   (with it) persistent lanes over the cells stage 4
//                         leaves: SplitNR and CheckSplit, results written
//                         out; a cell whose Newton steps do not converge
//                         goes to the general kernel, which resumes it at
//                         the restart loop.
//
// Trials run past an undecided one are speculative: their results are not
// used when the earlier trial turns out unstable, exactly as the reference
// would not have run them.  The plain mode skips these stages; in a context of
// several models each cell reads its own model's fluid (MULTI).
#include "ocpx_dev.hpp"
#include "ocpx_kernels.hpp"

namespace ocpx {
namespace sta {

constexpr int NONE = 4;          // decided: no trial found unstable
#ifndef STA_WAVES
#define STA_WAVES 2               // stage-1 wavefronts per SIMD the register budget targets
#endif
#ifndef SPLIT_WAVES
#define SPLIT_WAVES 2             // the same for stage 4
#endif
#ifndef SNR_WAVES
#define SNR_WAVES 1               // and for stage 5
#endif
#define DEV __device__ __forceinline__

// SkipPSA::CalFtype02 (src/AcceleratePEC.cpp:103-124) and the previous phase
// count, as the general kernels' begin step computes them
template <int NC>
OCPX_HD int flash_type(const DevFluid& f, const PeIO& io, int cell, const double* Ni, double Nt, double P,
                       double Tk, uint32_t& lNP)
{
    const size_t n = io.nb;
    uint32_t phaseNum = 1;                         // no previous step: one phase
    if (io.phaseExist) {
        phaseNum = 0;
        UNROLL for (int j = 0; j < 3; j++) phaseNum += io.phaseExist[(size_t)j * n + cell] != 0;
    }
    lNP = phaseNum - 1u;
    int ft = io.ftype_in ? io.ftype_in[cell] : -1;
    if (ft >= 0) return ft;
    ft = 0;
    if (f.skip) {
        bool skip = false;
        if (io.skip_flag[cell]) {
            const double me = io.skip_minEig[cell];
            skip = fabs(1 - io.skip_P[cell] / P) < me / 10 && fabs(io.skip_T[cell] - Tk) < me * 10;
            UNROLL for (int i = 0; i < NC; i++)
                skip = skip && fabs(Ni[i] / Nt - io.skip_zi[(size_t)i * n + cell]) < me / 10;
        }
        const uint32_t npPE = phaseNum - 1u;
        if (skip) ft = 1;
        else if (npPE >= 2) {
            const uint32_t t = (io.S[cell] >= 1E-4) + (io.S[n + cell] >= 1E-4);
            ft = t == npPE ? (int)npPE : 0;
        }
    }
    return ft;
}

// Wilson K of trial k, Kw[k] = Kw0^c_k with c = 1, -1, 1/3, -1/3
// (CalKwilson, src/OCPPhaseEquilibrium.cpp:244-252)
OCPX_HD double trial_coef(int k)
{
    return k == 0 ? 1.0 : (k == 1 ? -1.0 : (k == 2 ? (1.0 / 3) : -(1.0 / 3)));
}

OCPX_HD const DevFluid* opaque(const DevFluid* p)
{
    asm volatile("" : "+s"(p));
    return p;
}

// The fluid constants seen through the constant address space: the lean
// kernels' uniform reads of them become scalar loads into SGPRs instead of
// vector loads into VGPRs.  The asm makes the pointer opaque where it is
// taken, so each use reloads what it needs rather than holding the
// binary-interaction matrix in registers across the loop.
using CFluid = __attribute__((address_space(4))) const DevFluid;
__device__ __forceinline__ CFluid* cfl(const DevFluid* p)
{
    CFluid* q = (CFluid*)p;
    asm volatile("" : "+s"(q));
    return q;
}

// Append to a device list with one atomic per wavefront.
__device__ __forceinline__ void append(bool want, unsigned* count, int32_t* list, int32_t value)
{
    const unsigned long long m = __ballot(want);
    if (!m) return;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)m) - 1;
    unsigned base = 0;
    if (lane == leader) base = atomicAdd(count, (unsigned)__popcll(m));
    base = __shfl(base, leader);
    if (want) list[base + __popcll(m & ((1ull << lane) - 1ull))] = value;
}

// The fluid of a lane's cell: the one fluid through the opaque pointer, or,
// in a context of several models (MULTI), the cell's model's entry of the
// per-model array (a per-lane address, so vector loads).
template <bool MULTI>
DEV const DevFluid& fl_of(const DevFluid* fp, const PeIO& io, int cell)
{
    if constexpr (MULTI) return fp[io.mid[cell]];
    else return *opaque(fp);
}

enum : int { S_IDLE = 0, S_FEED, S_START, S_SSM };

// ln phi at y (src/OCPEoS.cpp:68-124, 199-214 in log form) with the cell's
// sqrt(Ai) in registers and Bi = bT P / T formed where used, so a lane holds
// only what the SSM loop carries
template <int NC, class F>
DEV void lean_lnphi(const F& f, const double* sA, double PT, const double* y, double* ev)
{
    double w[NC];
    UNROLL for (int k = 0; k < NC; k++) w[k] = sA[k] * y[k];
    double A = 0, B = 0;
    UNROLL for (int i = 0; i < NC; i++) {
        double t = 0;
        UNROLL for (int k = 0; k < NC; k++) t += f.omk[i * OCPX_NCMAX + k] * w[k];
        ev[i] = sA[i] * t;                                   // s_i for now
        A += y[i] * ev[i];
        B += (f.bT[i] * PT) * y[i];
    }
    const double Z = pr_z(A, B);
    const double lz = log(Z - B);
    const double lg = log((Z + D1 * B) / (Z + D2 * B));
    const double cA = A / (D1 - D2) / B * lg;
    const double iB = 1 / B, iA = 1 / A;
    UNROLL for (int i = 0; i < NC; i++) {
        const double Bi = f.bT[i] * PT;
        ev[i] = Bi * iB * (Z - 1) - lz - cA * (2 * ev[i] * iA - Bi * iB);
    }
}

// stage 1: one lane per cell, persistent over the work queue
template <int NC, bool MULTI>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(STA_WAVES)))
pe_sta_ssm(const DevFluid* __restrict__ fp, PeIO io, StaIO s)
{
    const size_t n = io.nb;
    const int lane = __lane_id();
    int cell = -1, st = S_IDLE, k = 0, iter = 0, decided = NONE, defer = 0;
    int its[4] = {0, 0, 0, 0};
    double P = 0, PT = 0, logP = 0, Tk = 0, Yt = 0, Ytd = 0;
    double y[NC], di[NC], lp[NC], sA[NC];
    bool drained = false;
    while (true) {
        // cells for idle lanes; cells off the strict path go to the general list
        while (true) {
            const bool need = st == S_IDLE && !drained;
            const unsigned long long m = __ballot(need);
            if (!m) break;
            const int leader = __ffsll((long long)m) - 1;
            unsigned base = 0;
            if (lane == leader) base = atomicAdd(s.qhead, (unsigned)__popcll(m));
            base = __shfl(base, leader);
            bool general = false;
            if (need) {
                const unsigned idx = base + (unsigned)__popcll(m & ((1ull << lane) - 1ull));
                if (idx < (unsigned)n && s.long_evals > 0 && s.route[io.qorder ? io.qorder[idx] : (int)idx] == 6) {
                    // a long cell: the general kernel beside the stages has it
                } else if (idx < (unsigned)n) {
                    cell = io.qorder ? io.qorder[idx] : (int)idx;
                    const DevFluid& f0 = fl_of<MULTI>(fp, io, cell);
                    P = io.P[cell];
                    Tk = io.T[cell] + CONV4;                          // SetVarSet :674
                    double Ni[NC], Nt = 0;
                    UNROLL for (int i = 0; i < NC; i++) { Ni[i] = io.Ni[(size_t)i * n + cell]; Nt += Ni[i]; }
                    uint32_t lNP;
                    const int ft = flash_type<NC>(f0, io, cell, Ni, Nt, P, Tk, lNP);
                    if (ft == 0 && lNP == 0) {
                        st = S_FEED;
                        logP = log(P);
                        PT = P / Tk;
                        UNROLL for (int i = 0; i < NC; i++) {   // sqrt(Ai), src/OCPEoS.cpp:68-84
                            const double Pr = P / f0.Pc[i], Tr = Tk / f0.Tc[i];
                            const double q = 1 + f0.mw[i] * (1 - sqrt(Tr));
                            sA[i] = sqrt(f0.OA[i] * Pr / (Tr * Tr) * (q * q));
                            y[i] = Ni[i] / Nt;                             // InitNtZ :812-818
                        }
                        k = 0; iter = 0; decided = NONE; defer = 0;
                        UNROLL for (int t = 0; t < 4; t++) its[t] = 0;
                    } else {
                        general = true;
                        s.route[cell] = 0;
                    }
                } else {
                    drained = true;
                }
            }
            append(general, s.gcount, s.glist, cell);
        }
        if (!__any(st != S_IDLE)) break;
        if (st == S_IDLE) continue;

        auto step = [&](const auto& f) {
        // the composition evaluated in this step (the feed's is in y already)
        if (st == S_START) {       // Y = x / Kw[k] (:420-425), x re-read from Ni
            const double ck = trial_coef(k);
            double Nt = 0;
            UNROLL for (int i = 0; i < NC; i++) { y[i] = io.Ni[(size_t)i * n + cell]; Nt += y[i]; }
            UNROLL for (int i = 0; i < NC; i++)
                y[i] = (y[i] / Nt) / exp(ck * (f.logPc[i] - logP + f.kwA[i] * (1 - f.Tc[i] / Tk)));
        } else if (st == S_SSM) {  // Y = di / phiSta (:439-444), lp = ln phiSta
            UNROLL for (int i = 0; i < NC; i++) y[i] = di[i] * exp(-lp[i]);
        }
        if (st != S_FEED) {
            double yt = 0;
            UNROLL for (int i = 0; i < NC; i++) yt += y[i];
            Yt = yt;
            const double r = 1 / yt;
            UNROLL for (int i = 0; i < NC; i++) y[i] *= r;
        }
        double ev[NC];
        lean_lnphi<NC>(f, sA, PT, y, ev);

        bool trial_end = false;
        double Se = 0;
        if (st == S_FEED) {        // di = phi[Id] x (:403-408), fug[Id] = di P
            UNROLL for (int i = 0; i < NC; i++) {
                di[i] = exp(ev[i]) * y[i];
                s.di[(size_t)i * n + cell] = di[i];
            }
            st = S_START;
        } else {
            // Se = sum ln(fugSta / fugId * Yt)^2 (:429-432, :448-451); after an SSM
            // update Y Yt = di / phiSta of the previous iterate, so the log is the
            // change of ln phiSta
            const bool first = st == S_START;
            if (first) {
                UNROLL for (int i = 0; i < NC; i++) {
                    const double r = log((exp(ev[i]) * y[i] * P) / (di[i] * P) * Yt);
                    Se += r * r;
                }
            } else {
                UNROLL for (int i = 0; i < NC; i++) {
                    const double r = ev[i] - lp[i];
                    Se += r * r;
                }
            }
            UNROLL for (int i = 0; i < NC; i++) lp[i] = ev[i];
            if (first) {
                iter = 0;
                if (Se > f.SSMsta_tol2) st = S_SSM;
                else trial_end = true;
            } else {
                its[k]++;
                if (!(Se > f.SSMsta_tol2 && ++iter <= f.SSMsta_maxIt)) trial_end = true;
            }
        }
        if (trial_end) {
            // StableNR (:478-488): its residual is the last SSM residual
            bool done = false;
            if (sqrt(Se) > f.NRsta_tol) {
                defer |= 1 << k;
                const unsigned q = atomicAdd(s.ncount, 1u);
                s.nr_cell[q] = cell;
                s.nr_k[q] = k;
                s.nr_Yt[q] = Yt;
                UNROLL for (int i = 0; i < NC; i++) s.nr_y[(size_t)i * s.nr_cap + q] = y[i];
                s.item[(size_t)k * n + cell] = (int32_t)q;
            } else if (Yt > 1 + f.eYt) {                  // flag && Yt > 1 + eYt (:463)
                decided = k;
                Ytd = Yt;
                done = true;
            }
            if (!done && ++k == 4) done = true;
            if (done) {
                s.route[cell] = 1;
                s.decided[cell] = decided;
                s.defer[cell] = defer;
                s.Ytd[cell] = Ytd;
                UNROLL for (int t = 0; t < 4; t++) s.its[(size_t)t * n + cell] = its[t];
                st = S_IDLE;
            } else {
                st = S_START;
            }
        }
        };
        if constexpr (MULTI) step(fp[io.mid[cell]]);
        else step(*cfl(fp));
    }
}

// stage 2: StableNR (:475-521) of the queued trials, persistent lanes: a lane
// takes a trial (one wave-aggregated atomic for all idle lanes), evaluates it
// at its SSM's last Y and Yt, and then does one convergence test and at most
// one Newton step per trip, so lanes whose trials converge early take new ones
// while the others iterate.  Jacobian JmatSTA (:523-541) solved as the
// symmetric matrix of its lower triangle (dsysv_ 'U' on row-major storage,
// :493).  The queue is complete before the launch (stage 1 filled it), and a
// lane leaves only when it is drained and the lane holds no trial.
template <int NC, bool MULTI>
__global__ void __launch_bounds__(64) pe_sta_nr(const DevFluid* __restrict__ fp, PeIO io, StaIO s)
{
    const unsigned total = *s.ncount;
    const size_t n = io.nb;
    const int lane = __lane_id();
    bool busy = false, drained = false;
    unsigned q = 0;
    int cell = 0, iter = 0, evals = 0, flag = 1;
    double P = 0, Yt = 0, e2 = 0;
    EosCell<NC> ec;
    EosMix<NC> mx;
    double y[NC], di[NC], res[NC];
    while (true) {
        const bool need = !busy && !drained;
        const unsigned long long m = __ballot(need);
        if (m) {
            const int leader = __ffsll((long long)m) - 1;
            unsigned base = 0;
            if (lane == leader) base = atomicAdd(s.nhead, (unsigned)__popcll(m));
            base = __shfl(base, leader);
            if (need) {
                const unsigned idx = base + (unsigned)__popcll(m & ((1ull << lane) - 1ull));
                if (idx < total) {
                    q = idx;
                    cell = s.nr_cell[q];
                    const DevFluid& f = fl_of<MULTI>(fp, io, cell);
                    P = io.P[cell];
                    const double Tk = io.T[cell] + CONV4;
                    eos_cell<NC>(f, P, Tk, ec);
                    Yt = s.nr_Yt[q];
                    UNROLL for (int i = 0; i < NC; i++) {
                        y[i] = s.nr_y[(size_t)i * s.nr_cap + q];
                        di[i] = s.di[(size_t)i * n + cell];
                    }
                    double ev[NC];
                    eos_mix<NC>(f, ec, y, mx);
                    eos_lnphi<NC>(ec, mx, ev);
                    evals = 1;
                    e2 = 0;
                    UNROLL for (int i = 0; i < NC; i++) {       // resSTA = ln(fug[Id] / (fugSta Yt))
                        res[i] = log((di[i] * P) / ((exp(ev[i]) * y[i] * P) * Yt));
                        e2 += res[i] * res[i];
                    }
                    iter = 0;
                    flag = 1;
                    busy = true;
                } else {
                    drained = true;
                }
            }
        }
        if (!__any(busy)) break;
        if (busy) {
            const DevFluid& f = fl_of<MULTI>(fp, io, cell);
            bool fin = false;
            if (sqrt(e2) > f.NRsta_tol) {
                double Zx[NC];
                const double den = z_den(mx), bc = z_bcoef(mx);
                UNROLL for (int j = 0; j < NC; j++) Zx[j] = ((mx.B - mx.Z) * (2 * mx.s[j]) + bc * ec.Bi[j]) / den;
                double S[NC][NC], rhs[1][NC];
                UNROLL for (int i = 0; i < NC; i++) {
                    double row[NC];
                    lnfugx_row<NC>(f, ec, mx, Zx, y, P, i, row);
                    double t = 0;
                    UNROLL for (int j = 0; j < NC; j++) t += y[j] * row[j];
                    UNROLL for (int j = 0; j <= i; j++) {
                        const double v = (row[j] - t + 1) / Yt;
                        S[i][j] = v;
                        S[j][i] = v;
                    }
                    rhs[0][i] = res[i];
                }
                gauss_solve<NC, 1>(S, rhs);
                const double lYt = Yt;
                double yt = 0;
                UNROLL for (int i = 0; i < NC; i++) { y[i] = y[i] * lYt + rhs[0][i]; yt += y[i]; }
                Yt = yt;
                const double r = 1 / yt;
                UNROLL for (int i = 0; i < NC; i++) y[i] *= r;
                double ev[NC];
                eos_mix<NC>(f, ec, y, mx);
                eos_lnphi<NC>(ec, mx, ev);
                evals++;
                e2 = 0;
                UNROLL for (int i = 0; i < NC; i++) {
                    res[i] = log((di[i] * P) / ((exp(ev[i]) * y[i] * P) * Yt));
                    e2 += res[i] * res[i];
                }
                if (++iter > f.NRsta_maxIt) { flag = 0; fin = true; }
            } else {
                fin = true;
            }
            if (fin) {
                s.nr_flag[q] = flag;
                s.nr_Yt[q] = Yt;
                s.nr_its[q] = iter;
                s.nr_evals[q] = evals - 1;          // the re-evaluation at the SSM's Y is not the reference's
                busy = false;
            }
        }
    }
}

// stage 3: the reference's order of trials, the counters up to the deciding
// trial, and the split queue
template <int NC, bool MULTI>
__global__ void __launch_bounds__(256) pe_sta_join(const DevFluid* __restrict__ fp, PeIO io, StaIO s)
{
    const size_t n = io.nb;
    const int64_t cell = (int64_t)blockIdx.x * 256 + threadIdx.x;
    bool split = false;
    if (cell < (int64_t)n && s.route[cell] == 1) {
        const DevFluid& f = MULTI ? fp[io.mid[cell]] : *fp;
        const int dec = s.decided[cell], def = s.defer[cell];
        int kstar = NONE;
        double Yt = 0;
        int c0 = 0, c1 = 0, c5 = 1, c6 = 0;      // c5: the feed's evaluation
        for (int k = 0; k < 4; k++) {
            if (k > dec) break;
            const int it = s.its[(size_t)k * n + cell];
            c0 += it;
            c5 += 1 + it;
            if (k == dec) { kstar = k; Yt = s.Ytd[cell]; break; }
            if (def & (1 << k)) {
                const unsigned q = (unsigned)s.item[(size_t)k * n + cell];
                const int nit = s.nr_its[q];
                c1 += nit;
                c5 += s.nr_evals[q];
                c6 += nit;
                if (s.nr_flag[q] && s.nr_Yt[q] > 1 + f.eYt) { kstar = k; Yt = s.nr_Yt[q]; break; }
            }
        }
        if (kstar == NONE) {                       // stable: NP = 1, x = z
            double Ni[NC], Nt = 0;
            UNROLL for (int i = 0; i < NC; i++) { Ni[i] = io.Ni[(size_t)i * n + cell]; Nt += Ni[i]; }
            io.pe_NP[cell] = 1;
            io.pe_nu[cell] = 1.0;
            UNROLL for (int i = 0; i < NC; i++) {
                io.pe_x[(size_t)i * n + cell] = Ni[i] / Nt;
                io.pe_x[(size_t)(NC + i) * n + cell] = 0.0;
            }
            io.ftype_out[cell] = 0;
            if (io.iters) {
                const int cnt[8] = {c0, c1, 0, 0, 0, c5, c6, c6};
                UNROLL for (int q = 0; q < 8; q++) io.iters[(size_t)q * n + cell] = cnt[q];
            }
        } else {                                   // unstable at trial kstar: split
            s.rs_Yt[cell] = Yt;
            s.rs_cnt[(size_t)0 * n + cell] = c0;
            s.rs_cnt[(size_t)1 * n + cell] = c1;
            s.rs_cnt[(size_t)2 * n + cell] = c5;
            s.rs_cnt[(size_t)3 * n + cell] = c6;
            s.route[cell] = 2;
            split = true;
        }
    }
    if (s.split) append(split, s.scount, s.slist, (int32_t)cell);
    else append(split, s.gcount, s.glist, (int32_t)cell);
}

// stage 4: the first pass of SplitSSM2 (:586-649) for a cell stage 3 found
// unstable.  Such a cell had one phase before (strict test, flash type 0), so
// K starts from Wilson's (:598-603) and the pass takes SSMSP's tolerance and
// cap, both tightened when Yt < 1.1 (:611-614).  Each iteration:
// RachfordRice2 (:651-705) for nu, UpdateXRR (:778-791) for x1 = z / (1 +
// nu (K - 1)) and x0 = K x1, ln phi of x0 and of x1, the residual
// sum (f1 / f0 - 1)^2 in the literal form of :629-635, and K = phi1 / phi0.
// x0, x1 are formed again from z, K and nu wherever they are needed, so a
// lane carries z, K, sqrt(Ai) and ln phi(x0) across the iteration.
template <int NC, bool MULTI>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(SPLIT_WAVES)))
pe_sta_split(const DevFluid* __restrict__ fp, PeIO io, StaIO s)
{
    const size_t n = io.nb;
    const int lane = __lane_id();
    const unsigned total = *s.scount;
    int cell = -1, iter = 0, maxIt = 0, rr = 0;
    bool busy = false, drained = false;
    double P = 0, PT = 0, tol2 = 0, nu = 0;
    double z[NC], K[NC], sA[NC];
    while (true) {
        while (true) {
            const bool need = !busy && !drained;
            const unsigned long long m = __ballot(need);
            if (!m) break;
            const int leader = __ffsll((long long)m) - 1;
            unsigned base = 0;
            if (lane == leader) base = atomicAdd(s.shead, (unsigned)__popcll(m));
            base = __shfl(base, leader);
            if (need) {
                const unsigned idx = base + (unsigned)__popcll(m & ((1ull << lane) - 1ull));
                if (idx < total) {
                    cell = s.slist[idx];
                    const DevFluid& f0 = fl_of<MULTI>(fp, io, cell);
                    P = io.P[cell];
                    const double Tk = io.T[cell] + CONV4, logP = log(P);
                    PT = P / Tk;
                    double Nt = 0;
                    UNROLL for (int i = 0; i < NC; i++) { z[i] = io.Ni[(size_t)i * n + cell]; Nt += z[i]; }
                    UNROLL for (int i = 0; i < NC; i++) {
                        z[i] = z[i] / Nt;
                        const double Pr = P / f0.Pc[i], Tr = Tk / f0.Tc[i];
                        const double q = 1 + f0.mw[i] * (1 - sqrt(Tr));
                        sA[i] = sqrt(f0.OA[i] * Pr / (Tr * Tr) * (q * q));
                        K[i] = exp(f0.logPc[i] - logP + f0.kwA[i] * (1 - f0.Tc[i] / Tk));
                    }
                    tol2 = f0.SSMsp_tol2;
                    maxIt = f0.SSMsp_maxIt;
                    if (s.rs_Yt[cell] < 1.1) {
                        tol2 *= 1E-1;
                        maxIt *= 2;
                    }
                    iter = 0;
                    rr = 0;
                    busy = true;
                } else {
                    drained = true;
                }
            }
        }
        if (!__any(busy)) break;
        if (!busy) continue;

        const DevFluid& f = fl_of<MULTI>(fp, io, cell);
        {   // RachfordRice2 on K
            double Kmin = K[0], Kmax = K[0];
            UNROLL for (int i = 1; i < NC; i++) { Kmin = fmin(Kmin, K[i]); Kmax = fmax(Kmax, K[i]); }
            const double numin = 1 / (1 - Kmax), numax = 1 / (1 - Kmin);
            nu = 0.5 * (numin + numax);
            int it = 0;
            while (true) {
                double rj = 0, J = 0;
                UNROLL for (int i = 0; i < NC; i++) {
                    const double km = K[i] - 1;
                    const double t = 1 / (1 + nu * km);
                    rj += z[i] * km * t;
                    J -= z[i] * km * km * t * t;
                }
                if (fabs(rj) < f.RR_tol || it > f.RR_maxIt) break;
                const double d = -rj / J, tn = nu + d;
                if (tn < numax && tn > numin) nu = tn;
                else nu = d > 0 ? (nu + numax) / 2 : (nu + numin) / 2;
                it++;
            }
            rr += it;
        }
        // each evaluation reads the fluid constants through its own opaque
        // pointer, so the binary-interaction matrix is not held in registers
        // from one evaluation to the next
        double x[NC], l0[NC], l1[NC];
        UNROLL for (int i = 0; i < NC; i++) x[i] = K[i] * (z[i] / (1 + nu * (K[i] - 1)));     // x0
        if constexpr (MULTI) lean_lnphi<NC>(fp[io.mid[cell]], sA, PT, x, l0);
        else lean_lnphi<NC>(*cfl(fp), sA, PT, x, l0);
        UNROLL for (int i = 0; i < NC; i++) x[i] = z[i] / (1 + nu * (K[i] - 1));              // x1
        if constexpr (MULTI) lean_lnphi<NC>(fp[io.mid[cell]], sA, PT, x, l1);
        else lean_lnphi<NC>(*cfl(fp), sA, PT, x, l1);
        iter++;
        double Se = 0;
        UNROLL for (int i = 0; i < NC; i++) {
            const double x1 = z[i] / (1 + nu * (K[i] - 1)), x0 = K[i] * x1;
            const double phi0 = exp(l0[i]), phi1 = exp(l1[i]);
            const double q = (phi1 * x1 * P) / (phi0 * x0 * P) - 1;
            Se += q * q;
        }
        bool last, conv = true;
        if (iter > maxIt) { conv = false; last = true; }
        else last = !(Se > tol2);
        if (last) {     // x0, x1 of this iteration's K, then the updated K
            UNROLL for (int i = 0; i < NC; i++) {
                const double x1 = z[i] / (1 + nu * (K[i] - 1)), x0 = K[i] * x1;
                s.rs_v[(size_t)(0 * NC + i) * n + cell] = x0;
                s.rs_v[(size_t)(1 * NC + i) * n + cell] = x1;
                s.rs_v[(size_t)(3 * NC + i) * n + cell] = l0[i];
                s.rs_v[(size_t)(4 * NC + i) * n + cell] = l1[i];
            }
        }
        UNROLL for (int i = 0; i < NC; i++) K[i] = exp(l1[i]) / exp(l0[i]);
        if (last) {
            UNROLL for (int i = 0; i < NC; i++) s.rs_v[(size_t)(2 * NC + i) * n + cell] = K[i];
            s.rs_nu[cell] = nu;
            s.rs_cnt[(size_t)2 * n + cell] += 2 * iter;
            s.rs_cnt[(size_t)4 * n + cell] = iter;
            s.rs_cnt[(size_t)5 * n + cell] = rr;
            s.rs_cnt[(size_t)6 * n + cell] = conv ? 1 : 0;
            s.route[cell] = 3;
            busy = false;
        }
        if (s.snr) append(last, s.tcount, s.tlist, cell);
        else append(last, s.gcount, s.glist, cell);
    }
}

// stage 5: SplitNR (:805-897, AssembleJmatSP :911-937, CalStepNRsp
// :939-957) from what stage 4 left, then, when it converges, CheckSplit
// (:960-1010) and the cell's results; otherwise PhaseSplit's restart loop
// (:558-566) follows in the general kernel (route 4).  One lane per cell,
// persistent over the list; the Newton system is solved as the symmetric
// matrix of its lower triangle (dsysv_ 'U' on row-major storage, :835).
// The fluid constants come through one opaque pointer per step: with one per
// Jacobian entry (about 80 per step) this kernel made wrong convergence
// decisions at Nc = 12 under its register spills, where its steps should
// match the general kernels'.
template <int NC, bool MULTI>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(SNR_WAVES)))
pe_sta_snr(const DevFluid* __restrict__ fp, PeIO io, StaIO s)
{
    const size_t n = io.nb;
    const int lane = __lane_id();
    const unsigned total = *s.tcount;
    int cell = -1, iter = 0;
    bool busy = false, drained = false, newton = false, conv = false;
    double P = 0, nu0 = 0, eNR0 = 0;
    double z[NC], a1[NC], a2[NC], a3[NC], a5[NC];
    EosCell<NC> ec;
    while (true) {
        while (true) {
            const bool need = !busy && !drained;
            const unsigned long long m = __ballot(need);
            if (!m) break;
            const int leader = __ffsll((long long)m) - 1;
            unsigned base = 0;
            if (lane == leader) base = atomicAdd(s.thead, (unsigned)__popcll(m));
            base = __shfl(base, leader);
            if (need) {
                const unsigned idx = base + (unsigned)__popcll(m & ((1ull << lane) - 1ull));
                if (idx < total) {
                    cell = s.tlist[idx];
                    const DevFluid& f0 = fl_of<MULTI>(fp, io, cell);
                    P = io.P[cell];
                    eos_cell<NC>(f0, P, io.T[cell] + CONV4, ec);
                    double Nt = 0;
                    UNROLL for (int i = 0; i < NC; i++) { z[i] = io.Ni[(size_t)i * n + cell]; Nt += z[i]; }
                    UNROLL for (int i = 0; i < NC; i++) z[i] = z[i] / Nt;
                    nu0 = s.rs_nu[cell];
                    // SplitNR :807-823: n0 = nu0 x0, resSP = ln(f1 / f0)
                    double e2 = 0;
                    UNROLL for (int i = 0; i < NC; i++) {
                        a2[i] = s.rs_v[(size_t)i * n + cell];
                        a3[i] = s.rs_v[(size_t)(NC + i) * n + cell];
                        const double l0 = s.rs_v[(size_t)(3 * NC + i) * n + cell];
                        const double l1 = s.rs_v[(size_t)(4 * NC + i) * n + cell];
                        a1[i] = nu0 * a2[i];
                        a5[i] = log((exp(l1) * a3[i] * P) / (exp(l0) * a2[i] * P));
                        e2 += a5[i] * a5[i];
                    }
                    const double eNR = sqrt(e2);
                    eNR0 = eNR;
                    iter = 0;
                    newton = eNR > f0.NRsp_tol;
                    conv = !newton && eNR < f0.NRsp_tol;
                    busy = true;
                } else {
                    drained = true;
                }
            }
        }
        if (!__any(busy)) break;
        if (!busy) continue;

        bool end = !newton;
        if (newton) {
            const DevFluid& f = fl_of<MULTI>(fp, io, cell);
            const double nu1 = 1 - nu0;
            double J[NC][NC], rhs[1][NC];
            {
                EosMix<NC> m0, m1;
                LnfN<NC> d0, d1;
                eos_mix<NC>(f, ec, a2, m0);
                eos_mix<NC>(f, ec, a3, m1);
                lnfn_prep<NC>(ec, m0, nu0, d0);
                lnfn_prep<NC>(ec, m1, nu1, d1);
                UNROLL for (int i = 0; i < NC; i++) {
                    UNROLL for (int k = 0; k <= i; k++) {
                        const double v = lnfn_entry<NC, false>(f, ec, m0, d0, a2, P, nu0, i, k) +
                                         lnfn_entry<NC, false>(f, ec, m1, d1, a3, P, nu1, i, k);
                        J[i][k] = v;
                        J[k][i] = v;
                    }
                    rhs[0][i] = a5[i];
                }
            }
            gauss_solve<NC, 1>(J, rhs);
            double alpha = 1;                          // CalStepNRsp
            UNROLL for (int i = 0; i < NC; i++)
                if (a1[i] + alpha * rhs[0][i] < 0) alpha *= 0.9 * fabs(a1[i] / rhs[0][i]);
            double nn = 0, en = 0;
            UNROLL for (int i = 0; i < NC; i++) {
                const double d = alpha * rhs[0][i];
                a1[i] += d;
                nn += a1[i];
                en += 2 * d * d;
            }
            nu0 = nn;
            const double r0 = 1 / nu0, r1 = 1 / (1 - nu0);
            UNROLL for (int i = 0; i < NC; i++) {
                a2[i] = fabs(a1[i] * r0);
                a3[i] = fabs((z[i] - a1[i]) * r1);
            }
            double l0[NC], l1[NC];
            {
                EosMix<NC> m;
                eos_mix<NC>(f, ec, a2, m);
                eos_lnphi<NC>(ec, m, l0);
                eos_mix<NC>(f, ec, a3, m);
                eos_lnphi<NC>(ec, m, l1);
            }
            double e2 = 0;
            UNROLL for (int i = 0; i < NC; i++) {
                a5[i] = log((exp(l1[i]) * a3[i] * P) / (exp(l0[i]) * a2[i] * P));
                e2 += a5[i] * a5[i];
            }
            const double eNR = sqrt(e2);
            iter++;
            if (eNR > eNR0 || iter > f.NRsp_maxIt) end = true;
            else if (en < 1E-16 * (2.0 * NC) * (2.0 * NC)) { conv = true; end = true; }
            else if (!(eNR > f.NRsp_tol)) end = true;
            if (end && eNR < f.NRsp_tol) conv = true;
        }
        bool restart = false;
        if (end) {
            const int c0 = s.rs_cnt[cell], c1 = s.rs_cnt[n + cell], c5 = s.rs_cnt[2 * n + cell];
            const int c6 = s.rs_cnt[3 * n + cell], c2 = s.rs_cnt[4 * n + cell], c4 = s.rs_cnt[5 * n + cell];
            if (conv) {                                // CheckSplit, then the results (PhaseEquilibrium :106-111)
                int NP = 2;
                const double nuMax = fmax(nu0, 1 - nu0);
                double eX = 0;
                UNROLL for (int i = 0; i < NC; i++) eX += (a2[i] - a3[i]) * (a2[i] - a3[i]);
                if (!(nuMax < 1 && finite_d(eX)) && (!finite_d(eX) || 1 - nuMax < 1E-3)) NP = 1;
                io.pe_NP[cell] = NP;
                io.pe_nu[cell] = NP > 1 ? nu0 : 1.0;
                UNROLL for (int i = 0; i < NC; i++) {
                    io.pe_x[(size_t)i * n + cell] = NP > 1 ? a2[i] : z[i];
                    io.pe_x[(size_t)(NC + i) * n + cell] = NP > 1 ? a3[i] : 0.0;
                }
                io.ftype_out[cell] = NP > 1 ? 2 : 0;
                if (io.iters) {
                    const int cnt[8] = {c0, c1, c2, iter, c4, c5 + 2 * iter, c6 + 2 * iter, c6 + iter};
                    UNROLL for (int q = 0; q < 8; q++) io.iters[(size_t)q * n + cell] = cnt[q];
                }
                s.route[cell] = 5;
            } else {                                   // the restart loop, in the general kernel
                s.rs_cnt[2 * n + cell] = c5 + 2 * iter;
                s.rs_cnt[3 * n + cell] = c6 + 2 * iter;
                s.rs_cnt[7 * n + cell] = iter;
                s.route[cell] = 4;
                restart = true;
            }
            busy = false;
        }
        append(restart, s.gcount, s.glist, cell);
    }
}

__global__ void __launch_bounds__(64) sta_reset(unsigned* c)
{
    if (threadIdx.x < 10) c[threadIdx.x] = 0;
}

// route 6 and the list for the cells whose previous flash took more than
// long_evals fugacity evaluations, route 0 for the others (stage 1 sets theirs)
__global__ void __launch_bounds__(256) sta_pick_long(int64_t nb, StaIO s)
{
    const int64_t cell = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool isl = cell < nb && s.prev_evals[cell] > s.long_evals;
    if (cell < nb) s.route[cell] = isl ? 6 : 0;
    append(isl, s.lcount, s.llist, (int32_t)cell);
}

template <int NC, bool MULTI>
int blocks_per_cu(int which)
{
    int b = 0;
    const hipError_t e = which == 0   ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, pe_sta_ssm<NC, MULTI>, 64, 0)
                         : which == 1 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, pe_sta_nr<NC, MULTI>, 64, 0)
                         : which == 2 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, pe_sta_split<NC, MULTI>, 64, 0)
                                      : hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, pe_sta_snr<NC, MULTI>, 64, 0);
    return e == hipSuccess && b > 0 ? b : 2;
}

template <int NC, bool MULTI>
hipError_t launch(const DevFluid* f, const PeIO& io, const StaIO& s, const int* grid, hipStream_t st)
{
    if (io.nb == 0) return hipSuccess;
    const long need = (io.nb + 63) / 64;
    long g[4];
    for (int q = 0; q < 4; q++) g[q] = grid[q] < need ? grid[q] : need;
    hipLaunchKernelGGL((pe_sta_ssm<NC, MULTI>), dim3((unsigned)g[0]), dim3(64), 0, st, f, io, s);
    hipLaunchKernelGGL((pe_sta_nr<NC, MULTI>), dim3((unsigned)g[1]), dim3(64), 0, st, f, io, s);
    hipLaunchKernelGGL((pe_sta_join<NC, MULTI>), dim3((unsigned)((io.nb + 255) / 256)), dim3(256), 0, st, f, io, s);
    if (s.split) hipLaunchKernelGGL((pe_sta_split<NC, MULTI>), dim3((unsigned)g[2]), dim3(64), 0, st, f, io, s);
    if (s.split && s.snr) hipLaunchKernelGGL((pe_sta_snr<NC, MULTI>), dim3((unsigned)g[3]), dim3(64), 0, st, f, io, s);
    return hipGetLastError();
}

}  // namespace sta

hipError_t run_pe_sta_prep(const StaIO& s, int64_t nb, hipStream_t st)
{
    hipLaunchKernelGGL(sta::sta_reset, dim3(1), dim3(64), 0, st, s.qhead);
    if (s.long_evals > 0 && nb > 0)
        hipLaunchKernelGGL(sta::sta_pick_long, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st, nb, s);
    return hipGetLastError();
}

hipError_t run_pe_sta(int NC, const DevFluid* f, const PeIO& io, const StaIO& s, const int* grid, hipStream_t st)
{
    switch (NC) {
#define OCPX_CASE(N) case N: return io.mid ? sta::launch<N, true>(f, io, s, grid, st) : sta::launch<N, false>(f, io, s, grid, st);
        OCPX_CASE(2) OCPX_CASE(3) OCPX_CASE(4) OCPX_CASE(5) OCPX_CASE(6) OCPX_CASE(7)
        OCPX_CASE(8) OCPX_CASE(9) OCPX_CASE(10) OCPX_CASE(11) OCPX_CASE(12)
#undef OCPX_CASE
    default: return hipErrorInvalidValue;
    }
}

int pe_sta_blocks_per_cu(int NC, int which, bool multi)
{
    switch (NC) {
#define OCPX_CASE(N) case N: return multi ? sta::blocks_per_cu<N, true>(which) : sta::blocks_per_cu<N, false>(which);
        OCPX_CASE(2) OCPX_CASE(3) OCPX_CASE(4) OCPX_CASE(5) OCPX_CASE(6) OCPX_CASE(7)
        OCPX_CASE(8) OCPX_CASE(9) OCPX_CASE(10) OCPX_CASE(11) OCPX_CASE(12)
#undef OCPX_CASE
    default: return 2;
    }
}

}  // namespace ocpx
