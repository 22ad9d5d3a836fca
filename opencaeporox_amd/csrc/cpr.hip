// cpr.hip — pressure AMG, multicolour block Gauss-Seidel and FGMRES pieces
// of the device linear solve, for gfx950 (declarations and the method in
// cpr.hpp).
//
// Every kernel here is a plain gather: a thread (or a group of B lanes for a
// block row) owns its output entries, so no atomics and the same bits on
// every run.  The AMG level matrices are scalar CSR with the diagonal first
// in each row.
#include <algorithm>
#include <cstring>

#include "amg_layout.hpp"
#include "cpr.hpp"
#include "ocpx_dev.hpp"

namespace ocpx {

namespace {

constexpr double OMEGA = 2.0 / 3.0;          // weighted-Jacobi smoother
constexpr int MAX_LEVELS = 25;

// ------------------------------------------------------------------ kernels
// level 0: a[s] = sum_m Dinv_i(0, m) A_s(m, 0), one thread per block row i,
// made an M-matrix for the hierarchy: positive off-diagonal entries go to
// the diagonal, and the diagonal is at least the sum of the remaining
// off-diagonal magnitudes (the decoupled pressure block of a state far from
// equilibrium is indefinite, and a Jacobi-smoothed aggregation cycle on it
// diverges)
__global__ void __launch_bounds__(256) amg_l0(int B, int64_t n, const int64_t* __restrict__ rp,
                                              const double* __restrict__ val, const double* __restrict__ dinv,
                                              double* __restrict__ a)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double* d = dinv + (size_t)i * B * B;     // row 0 of Dinv_i
    const size_t BB = (size_t)B * B;
    const int64_t s0 = rp[i], s1 = rp[i + 1];       // the diagonal slot first
    double diag = 0, pos = 0, neg = 0;
    for (int64_t s = s0; s < s1; s++) {
        const double* blk = val + (size_t)s * BB;
        double acc = 0;
        for (int m = 0; m < B; m++) acc += d[m] * blk[(size_t)m * B];
        if (s == s0) {
            diag = acc;
        } else if (acc > 0) {
            pos += acc;
            acc = 0;
        } else {
            neg -= acc;
        }
        a[s] = acc;
    }
    a[s0] = fmax(diag + pos, neg);
}

__global__ void __launch_bounds__(256) amg_gather(int64_t nnz, const int64_t* __restrict__ gp,
                                                  const int64_t* __restrict__ gs, const double* __restrict__ af,
                                                  double* __restrict__ ac)
{
    const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= nnz) return;
    double acc = 0;
    for (int64_t t = gp[s]; t < gp[s + 1]; t++) acc += af[gs[t]];
    ac[s] = acc;
}

__global__ void __launch_bounds__(256) amg_dinv(int64_t n, const int64_t* __restrict__ rp,
                                                const double* __restrict__ a, double* __restrict__ dinv)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double d = a[rp[i]];
    dinv[i] = d != 0 ? 1 / d : 0.0;
}

// x = OMEGA Dinv b (a sweep from x = 0)
__global__ void __launch_bounds__(256) amg_smooth0(int64_t n, const double* __restrict__ dinv,
                                                   const double* __restrict__ b, double* __restrict__ x)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    x[i] = OMEGA * dinv[i] * b[i];
}

// xo = x + OMEGA Dinv (b - A x)
__global__ void __launch_bounds__(256) amg_smooth(int64_t n, const int64_t* __restrict__ rp,
                                                  const int32_t* __restrict__ ci, const double* __restrict__ a,
                                                  const double* __restrict__ dinv, const double* __restrict__ b,
                                                  const double* __restrict__ x, double* __restrict__ xo)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double r = b[i];
    for (int64_t s = rp[i]; s < rp[i + 1]; s++) r -= a[s] * x[ci[s]];
    xo[i] = x[i] + OMEGA * dinv[i] * r;
}

// bc[I] = sum over the rows i of aggregate I of (b - A x)[i]
__global__ void __launch_bounds__(256) amg_resid_restrict(int64_t nc, const int64_t* __restrict__ mp,
                                                          const int32_t* __restrict__ mi,
                                                          const int64_t* __restrict__ rp,
                                                          const int32_t* __restrict__ ci,
                                                          const double* __restrict__ a,
                                                          const double* __restrict__ b,
                                                          const double* __restrict__ x, double* __restrict__ bc)
{
    const int64_t I = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (I >= nc) return;
    double acc = 0;
    for (int64_t k = mp[I]; k < mp[I + 1]; k++) {
        const int64_t i = mi[k];
        double r = b[i];
        for (int64_t s = rp[i]; s < rp[i + 1]; s++) r -= a[s] * x[ci[s]];
        acc += r;
    }
    bc[I] = acc;
}

__global__ void __launch_bounds__(256) amg_prolong(int64_t n, const int32_t* __restrict__ agg,
                                                   const double* __restrict__ xc, double* __restrict__ x)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    x[i] += xc[agg[i]];
}

// Dense LU with partial pivoting of the coarsest level (n <= 64), one
// 64-lane block; lane i holds row i.
__global__ void __launch_bounds__(64) amg_dense_lu(int n, const int64_t* __restrict__ rp,
                                                   const int32_t* __restrict__ ci, const double* __restrict__ a,
                                                   double* __restrict__ lu, int32_t* __restrict__ piv)
{
    constexpr int N = PressureAmg::DENSE_MAX;
    __shared__ double M[N][N + 1];
    __shared__ int prow;
    const int t = threadIdx.x;
    for (int j = 0; j < N; j++) M[t][j] = (t == j && t >= n) ? 1.0 : 0.0;
    if (t < n)
        for (int64_t s = rp[t]; s < rp[t + 1]; s++) M[t][ci[s]] += a[s];
    __syncthreads();
    for (int k = 0; k < n; k++) {
        if (t == 0) {
            int p = k;
            double best = fabs(M[k][k]);
            for (int i = k + 1; i < n; i++)
                if (fabs(M[i][k]) > best) { best = fabs(M[i][k]); p = i; }
            prow = p;
            piv[k] = p;
        }
        __syncthreads();
        const int p = prow;
        if (p != k && t < N) {           // lane t swaps column t of rows k and p
            const double v = M[k][t];
            M[k][t] = M[p][t];
            M[p][t] = v;
        }
        __syncthreads();
        const double d = M[k][k];
        if (t > k && t < n) {
            const double f = d != 0 ? M[t][k] / d : 0.0;
            M[t][k] = f;
            for (int j = k + 1; j < n; j++) M[t][j] -= f * M[k][j];
        }
        __syncthreads();
    }
    for (int j = 0; j < N; j++) lu[t * N + j] = M[t][j];
}

__global__ void __launch_bounds__(64) amg_dense_solve(int n, const double* __restrict__ lu,
                                                      const int32_t* __restrict__ piv, const double* __restrict__ b,
                                                      double* __restrict__ x)
{
    constexpr int N = PressureAmg::DENSE_MAX;
    __shared__ double y[N];
    const int t = threadIdx.x;
    y[t] = t < n ? b[t] : 0.0;
    __syncthreads();
    if (t == 0)
        for (int k = 0; k < n; k++) {
            const int p = piv[k];
            if (p != k) { const double v = y[k]; y[k] = y[p]; y[p] = v; }
        }
    __syncthreads();
    for (int k = 0; k < n; k++) {        // L (unit diagonal)
        if (t > k && t < n) y[t] -= lu[t * N + k] * y[k];
        __syncthreads();
    }
    for (int k = n - 1; k >= 0; k--) {   // U
        if (t == k) y[k] = lu[k * N + k] != 0 ? y[k] / lu[k * N + k] : 0.0;
        __syncthreads();
        if (t < k) y[t] -= lu[t * N + k] * y[k];
        __syncthreads();
    }
    if (t < n) x[t] = y[t];
}

// One colour of the block Gauss-Seidel sweep; a group of B lanes per row,
// lane r = row r of the block.  The diagonal slot is first in each row.
template <int B>
__global__ void __launch_bounds__(64) gs_color(const int32_t* __restrict__ rows, int64_t nrows,
                                               const int64_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                               const double* __restrict__ val, const double* __restrict__ dinv,
                                               const double* __restrict__ rhat, double* __restrict__ x)
{
    constexpr int G = 64 / B;
    const int lane = threadIdx.x, g = lane / B, r = lane % B;
    const int64_t k = (int64_t)blockIdx.x * G + g;
    const bool valid = g < G && k < nrows;
    const int64_t i = valid ? rows[k] : 0;
    double t = 0;
    if (valid)
        for (int64_t s = rowptr[i] + 1; s < rowptr[i + 1]; s++) {
            const double* blk = val + ((size_t)s * B + r) * B;
            const double* xm = x + (size_t)col[s] * B;
            UNROLL for (int c = 0; c < B; c++) t += blk[c] * xm[c];
        }
    const int base = (g < G ? g : 0) * B;
    double y = valid ? rhat[(size_t)i * B + r] : 0.0;
    const double* d = dinv + ((size_t)i * B + r) * B;
    UNROLL for (int c = 0; c < B; c++) {
        const double tc = __shfl(t, base + c);
        if (valid) y -= d[c] * tc;
    }
    if (valid) x[(size_t)i * B + r] = y;
}

constexpr int MD_BLOCKS = 512, MD_THREADS = 256, MD_MAX = 32;

// part[k][blockIdx.x] = partial sums of V_k . w over a fixed slice
__global__ void __launch_bounds__(MD_THREADS) mdot_part(const double* __restrict__ V, int64_t ldv, int m,
                                                        const double* __restrict__ w, int64_t len,
                                                        double* __restrict__ part)
{
    __shared__ double sh[MD_THREADS];
    const int t = threadIdx.x;
    for (int k = 0; k < m; k++) {
        const double* v = V + (size_t)k * ldv;
        double acc = 0;
        for (int64_t i = (int64_t)blockIdx.x * MD_THREADS + t; i < len; i += (int64_t)MD_BLOCKS * MD_THREADS)
            acc += v[i] * w[i];
        sh[t] = acc;
        __syncthreads();
        for (int s = MD_THREADS / 2; s > 0; s >>= 1) {
            if (t < s) sh[t] += sh[t + s];
            __syncthreads();
        }
        if (t == 0) part[(size_t)k * MD_BLOCKS + blockIdx.x] = sh[0];
        __syncthreads();
    }
}

__global__ void __launch_bounds__(MD_THREADS) mdot_final(int m, const double* __restrict__ part,
                                                         double* __restrict__ out)
{
    __shared__ double sh[MD_THREADS];
    const int t = threadIdx.x;
    for (int k = 0; k < m; k++) {
        double acc = 0;
        for (int j = t; j < MD_BLOCKS; j += MD_THREADS) acc += part[(size_t)k * MD_BLOCKS + j];
        sh[t] = acc;
        __syncthreads();
        for (int s = MD_THREADS / 2; s > 0; s >>= 1) {
            if (t < s) sh[t] += sh[t + s];
            __syncthreads();
        }
        if (t == 0) out[k] = sh[0];
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) maxpy(double* __restrict__ w, const double* __restrict__ V, int64_t ldv,
                                             int m, const double* __restrict__ h, int64_t len)
{
    __shared__ double hs[MD_MAX];
    if (threadIdx.x < m) hs[threadIdx.x] = h[threadIdx.x];
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= len) return;
    double acc = w[i];
    for (int k = 0; k < m; k++) acc += hs[k] * V[(size_t)k * ldv + i];
    w[i] = acc;
}

__global__ void __launch_bounds__(256) gather_p(int B, const double* __restrict__ v, double* __restrict__ p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v[(size_t)i * B];
}

__global__ void __launch_bounds__(256) scatter_p(int B, const double* __restrict__ p, double* __restrict__ v,
                                                 int64_t n, int64_t len)
{
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= len) return;
    const int64_t i = k / B;
    v[k] = (i < n && k == i * B) ? p[i] : 0.0;
}

unsigned grid256(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

// ------------------------------------------------------------- PressureAmg
template <class T>
int PressureAmg::dev(T** p, size_t count, const T* host, std::string& err)
{
    void* q = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(&q, count * sizeof(T));
    if (e != hipSuccess) { err = std::string("hipMalloc: ") + hipGetErrorString(e); return -1; }
    mem_.push_back(q);
    if (host) e = hipMemcpy(q, host, count * sizeof(T), hipMemcpyHostToDevice);
    else e = hipMemset(q, 0, count * sizeof(T));
    if (e != hipSuccess) { err = std::string("hipMemcpy: ") + hipGetErrorString(e); return -1; }
    *p = (T*)q;
    return 0;
}

PressureAmg::~PressureAmg()
{
    for (void* p : mem_) (void)hipFree(p);
}

int PressureAmg::setup(int64_t n, const std::vector<int64_t>& rp, const std::vector<int32_t>& col,
                       const std::vector<double>& w, std::string& err)
{
    amg::Graph g;
    g.n = n;
    g.rp = rp;
    g.ci = col;
    g.w = w;
    while (true) {
        Level L;
        L.n = g.n;
        L.nnz = (int64_t)g.ci.size();
        if (dev(&L.rp, g.rp.size(), g.rp.data(), err) || dev(&L.ci, g.ci.size(), g.ci.data(), err) ||
            dev(&L.a, (size_t)L.nnz, (const double*)nullptr, err) ||
            dev(&L.dinv, (size_t)L.n, (const double*)nullptr, err) ||
            dev(&L.b, (size_t)L.n, (const double*)nullptr, err) ||
            dev(&L.x, (size_t)L.n, (const double*)nullptr, err) ||
            dev(&L.r, (size_t)L.n, (const double*)nullptr, err))
            return -1;
        amg::Graph gc;
        amg::LevelMaps M;
        if (g.n <= DENSE_MAX || (int)lv_.size() + 1 >= MAX_LEVELS || !amg::coarsen(g, gc, M)) {
            lv_.push_back(L);
            break;
        }
        if (dev(&L.agg, M.agg.size(), M.agg.data(), err) || dev(&L.gp, M.gp.size(), M.gp.data(), err) ||
            dev(&L.gs, M.gs.size(), M.gs.data(), err) || dev(&L.mp, M.mp.size(), M.mp.data(), err) ||
            dev(&L.mi, M.mi.size(), M.mi.data(), err))
            return -1;
        lv_.push_back(L);
        g = std::move(gc);
    }
    if (lv_.back().n <= DENSE_MAX &&
        (dev(&lu_, (size_t)DENSE_MAX * DENSE_MAX, (const double*)nullptr, err) ||
         dev(&piv_, (size_t)DENSE_MAX, (const int32_t*)nullptr, err)))
        return -1;
    rp0_ = lv_[0].rp;
    return 0;
}

hipError_t PressureAmg::values(int B, const double* val, const double* dinv, hipStream_t st)
{
    Level& L0 = lv_[0];
    hipLaunchKernelGGL(amg_l0, dim3(grid256(L0.n)), dim3(256), 0, st, B, L0.n, L0.rp, val, dinv, L0.a);
    for (size_t l = 0; l < lv_.size(); l++) {
        Level& L = lv_[l];
        hipLaunchKernelGGL(amg_dinv, dim3(grid256(L.n)), dim3(256), 0, st, L.n, L.rp, L.a, L.dinv);
        if (l + 1 < lv_.size()) {
            Level& C = lv_[l + 1];
            hipLaunchKernelGGL(amg_gather, dim3(grid256(C.nnz)), dim3(256), 0, st, C.nnz, L.gp, L.gs, L.a, C.a);
        }
    }
    const Level& Lc = lv_.back();
    if (Lc.n <= DENSE_MAX)
        hipLaunchKernelGGL(amg_dense_lu, dim3(1), dim3(64), 0, st, (int)Lc.n, Lc.rp, Lc.ci, Lc.a, lu_, piv_);
    return hipGetLastError();
}

hipError_t PressureAmg::cycle(int l, const double* b, double* x, hipStream_t st)
{
    Level& L = lv_[l];
    if (l + 1 == (int)lv_.size()) {
        if (L.n <= DENSE_MAX) {
            hipLaunchKernelGGL(amg_dense_solve, dim3(1), dim3(64), 0, st, (int)L.n, lu_, piv_, b, x);
        } else {                              // no dense level: a few sweeps
            hipLaunchKernelGGL(amg_smooth0, dim3(grid256(L.n)), dim3(256), 0, st, L.n, L.dinv, b, L.r);
            double *in = L.r, *out = x;
            for (int k = 0; k < 9; k++) {
                hipLaunchKernelGGL(amg_smooth, dim3(grid256(L.n)), dim3(256), 0, st, L.n, L.rp, L.ci, L.a, L.dinv,
                                   b, in, out);
                std::swap(in, out);
            }
            if (in != x) hipMemcpyAsync(x, in, (size_t)L.n * sizeof(double), hipMemcpyDeviceToDevice, st);
        }
        return hipGetLastError();
    }
    Level& C = lv_[l + 1];
    hipLaunchKernelGGL(amg_smooth0, dim3(grid256(L.n)), dim3(256), 0, st, L.n, L.dinv, b, L.x);
    hipLaunchKernelGGL(amg_resid_restrict, dim3(grid256(C.n)), dim3(256), 0, st, C.n, L.mp, L.mi, L.rp, L.ci, L.a,
                       b, L.x, C.b);
    hipError_t e = cycle(l + 1, C.b, C.x, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(amg_prolong, dim3(grid256(L.n)), dim3(256), 0, st, L.n, L.agg, C.x, L.x);
    hipLaunchKernelGGL(amg_smooth, dim3(grid256(L.n)), dim3(256), 0, st, L.n, L.rp, L.ci, L.a, L.dinv, b, L.x, x);
    return hipGetLastError();
}

hipError_t PressureAmg::vcycle(const double* b, double* x, hipStream_t st)
{
    return cycle(0, b, x, st);
}

// --------------------------------------------------------------- launchers
#define OCPX_B_SWITCH(X)                                                                            \
    switch (B) {                                                                                    \
    X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14)                                     \
    default: return hipErrorInvalidValue;                                                           \
    }

hipError_t run_gs_color(int B, const int32_t* rows, int64_t nrows, const int64_t* rowptr, const int32_t* col,
                        const double* val, const double* dinv, const double* rhat, double* x, hipStream_t st)
{
    if (nrows == 0) return hipSuccess;
    const unsigned nblk = (unsigned)((nrows + 64 / B - 1) / (64 / B));
#define X(N) case N: hipLaunchKernelGGL(gs_color<N>, dim3(nblk), dim3(64), 0, st, rows, nrows, rowptr, col, val, \
                                        dinv, rhat, x); break;
    OCPX_B_SWITCH(X)
#undef X
    return hipGetLastError();
}
#undef OCPX_B_SWITCH

hipError_t run_mdot(const double* V, int64_t ldv, int m, const double* w, int64_t len, double* part, double* out,
                    hipStream_t st)
{
    if (m <= 0 || m > MD_MAX) return hipErrorInvalidValue;
    hipLaunchKernelGGL(mdot_part, dim3(MD_BLOCKS), dim3(MD_THREADS), 0, st, V, ldv, m, w, len, part);
    hipLaunchKernelGGL(mdot_final, dim3(1), dim3(MD_THREADS), 0, st, m, part, out);
    return hipGetLastError();
}

int mdot_scratch_len() { return MD_BLOCKS * MD_MAX; }

hipError_t run_maxpy(double* w, const double* V, int64_t ldv, int m, const double* h, int64_t len, hipStream_t st)
{
    if (m <= 0 || m > MD_MAX) return hipErrorInvalidValue;
    hipLaunchKernelGGL(maxpy, dim3(grid256(len)), dim3(256), 0, st, w, V, ldv, m, h, len);
    return hipGetLastError();
}

hipError_t run_gather_p(int B, const double* v, double* p, int64_t n, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(gather_p, dim3(grid256(n)), dim3(256), 0, st, B, v, p, n);
    return hipGetLastError();
}

hipError_t run_scatter_p(int B, const double* p, double* v, int64_t n, int64_t len, hipStream_t st)
{
    if (len == 0) return hipSuccess;
    hipLaunchKernelGGL(scatter_p, dim3(grid256(len)), dim3(256), 0, st, B, p, v, n, len);
    return hipGetLastError();
}

}  // namespace ocpx
