// solver.hip — device solve of the assembled FIM system J u = rhs, for gfx950.
//
// The reference hands its OCPMatrix to an external Krylov solver
// (LinearSystem::Solve, src/LinearSystem.cpp:103-119; FASP FGMRES(30) with a
// decoupling preconditioner, tolerance 1e-3 on ||r|| / ||b||, at most 200
// iterations, data/spe5/bsr.fasp).  Here the solve stays on the device next
// to the BSR it reads: BiCGStab (van der Vorst, SIAM J. Sci. Stat. Comput.
// 13 (1992) 631-644), right-preconditioned with block Jacobi, i.e. the
// inverse of every b x b diagonal block.
//
//  - bsr_diag_inv: one block per b lanes, lane r holding column r of the
//    block and of the identity; Gauss-Jordan with partial pivoting, the
//    pivot column broadcast from its lane.
//  - bsr_spmv (also the block-diagonal apply): one block row per b lanes,
//    lane r = row r; a lane reads row r of each block (b contiguous
//    doubles, the b lanes of a row read b*b contiguous doubles) and the b
//    entries of x it multiplies.
//  - dot: fixed-shape block partial sums, then one block sums them in
//    order, so a dot product has the same bits on every run.
#include "ocpx_dev.hpp"
#include "ocpx_kernels.hpp"

namespace ocpx {

template <int B>
__global__ void __launch_bounds__(64) bsr_diag_inv(const int64_t* __restrict__ rowptr, const double* __restrict__ val,
                                                   int64_t nb, double* __restrict__ dinv)
{
    constexpr int G = 64 / B;
    const int lane = threadIdx.x, g = lane / B, r = lane % B;
    const int64_t n = (int64_t)blockIdx.x * G + g;
    const bool valid = (g < G) && (n < nb);
    const int base = g * B;                      // first lane of this block's group
    double a[B], e[B];
    const double* blk = val + (size_t)(valid ? rowptr[n] : 0) * B * B;
    UNROLL for (int i = 0; i < B; i++) {
        a[i] = valid ? blk[i * B + r] : (i == r ? 1.0 : 0.0);
        e[i] = (i == r) ? 1.0 : 0.0;
    }
    UNROLL for (int k = 0; k < B; k++) {
        double ck[B];                            // column k, held by lane base + k
        UNROLL for (int i = 0; i < B; i++) ck[i] = __shfl(a[i], (g < G ? base : 0) + k);
        int p = k;
        double best = fabs(ck[k]);
        UNROLL for (int i = k + 1; i < B; i++)
            if (fabs(ck[i]) > best) { best = fabs(ck[i]); p = i; }
        UNROLL for (int i = k + 1; i < B; i++) {  // swap rows k and p
            if (i == p) {
                const double ta = a[k], te = e[k], tc = ck[k];
                a[k] = a[i]; e[k] = e[i]; ck[k] = ck[i];
                a[i] = ta; e[i] = te; ck[i] = tc;
            }
        }
        const double ip = 1 / ck[k];
        a[k] *= ip;
        e[k] *= ip;
        UNROLL for (int i = 0; i < B; i++) {
            if (i == k) continue;
            const double f = ck[i];
            a[i] -= f * a[k];
            e[i] -= f * e[k];
        }
    }
    if (!valid) return;
    double* out = dinv + (size_t)n * B * B;
    UNROLL for (int i = 0; i < B; i++) out[i * B + r] = e[i];
}

// y = A x over the block rows (A the BSR: rowptr, col, val), or, with
// rowptr == nullptr, y = blockdiag(val) x (val holding one block per row).
template <int B>
__global__ void __launch_bounds__(64) bsr_spmv(const int64_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                               const double* __restrict__ val, const double* __restrict__ x,
                                               double* __restrict__ y, int64_t nb)
{
    constexpr int G = 64 / B;
    const int lane = threadIdx.x, g = lane / B, r = lane % B;
    const int64_t n = (int64_t)blockIdx.x * G + g;
    if (g >= G || n >= nb) return;
    const int64_t s0 = rowptr ? rowptr[n] : n, s1 = rowptr ? rowptr[n + 1] : n + 1;
    double acc = 0;
    for (int64_t s = s0; s < s1; s++) {
        const int64_t m = rowptr ? (int64_t)col[s] : n;
        const double* blk = val + ((size_t)s * B + r) * B;
        const double* xm = x + (size_t)m * B;
        UNROLL for (int c = 0; c < B; c++) acc += blk[c] * xm[c];
    }
    y[(size_t)n * B + r] = acc;
}

constexpr int DOT_BLOCKS = 1024, DOT_THREADS = 256;

// part[blockIdx.x] = partial sum of x.y over a fixed grid-stride slice
__global__ void __launch_bounds__(DOT_THREADS) dot_part(const double* __restrict__ x, const double* __restrict__ y,
                                                        int64_t len, double* __restrict__ part)
{
    __shared__ double sh[DOT_THREADS];
    const int t = threadIdx.x;
    double acc = 0;
    for (int64_t i = (int64_t)blockIdx.x * DOT_THREADS + t; i < len; i += (int64_t)DOT_BLOCKS * DOT_THREADS)
        acc += x[i] * y[i];
    sh[t] = acc;
    __syncthreads();
    for (int w = DOT_THREADS / 2; w > 0; w >>= 1) {
        if (t < w) sh[t] += sh[t + w];
        __syncthreads();
    }
    if (t == 0) part[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(DOT_THREADS) dot_final(const double* __restrict__ part, double* __restrict__ out)
{
    __shared__ double sh[DOT_THREADS];
    const int t = threadIdx.x;
    double acc = 0;
    for (int k = t; k < DOT_BLOCKS; k += DOT_THREADS) acc += part[k];
    sh[t] = acc;
    __syncthreads();
    for (int w = DOT_THREADS / 2; w > 0; w >>= 1) {
        if (t < w) sh[t] += sh[t + w];
        __syncthreads();
    }
    if (t == 0) out[0] = sh[0];
}

// z = x + a y + b w   (any of y, w may alias z's inputs; a pointer may be null when its scalar is 0)
__global__ void __launch_bounds__(256) axpbypcz(double* z, const double* x, double a, const double* y,
                                                double b, const double* w, int64_t len)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    double v = x ? x[i] : 0.0;
    if (y) v += a * y[i];
    if (w) v += b * w[i];
    z[i] = v;
}

// ------------------------------------------------------------- launchers
#define OCPX_B_SWITCH(X)                                                                            \
    switch (B) {                                                                                    \
    X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14)                                     \
    default: return hipErrorInvalidValue;                                                           \
    }

hipError_t run_bsr_diag_inv(int B, const int64_t* rowptr, const double* val, int64_t nb, double* dinv,
                            hipStream_t st)
{
    if (nb == 0) return hipSuccess;
    const unsigned nblk = (unsigned)((nb + 64 / B - 1) / (64 / B));
#define X(N) case N: hipLaunchKernelGGL(bsr_diag_inv<N>, dim3(nblk), dim3(64), 0, st, rowptr, val, nb, dinv); break;
    OCPX_B_SWITCH(X)
#undef X
    return hipGetLastError();
}

hipError_t run_bsr_spmv(int B, const int64_t* rowptr, const int32_t* col, const double* val, const double* x,
                        double* y, int64_t nb, hipStream_t st)
{
    if (nb == 0) return hipSuccess;
    const unsigned nblk = (unsigned)((nb + 64 / B - 1) / (64 / B));
#define X(N) case N: hipLaunchKernelGGL(bsr_spmv<N>, dim3(nblk), dim3(64), 0, st, rowptr, col, val, x, y, nb); break;
    OCPX_B_SWITCH(X)
#undef X
    return hipGetLastError();
}
#undef OCPX_B_SWITCH

hipError_t run_dot(const double* x, const double* y, int64_t len, double* part, double* out, hipStream_t st)
{
    hipLaunchKernelGGL(dot_part, dim3(DOT_BLOCKS), dim3(DOT_THREADS), 0, st, x, y, len, part);
    hipLaunchKernelGGL(dot_final, dim3(1), dim3(DOT_THREADS), 0, st, part, out);
    return hipGetLastError();
}

int dot_scratch_len() { return DOT_BLOCKS; }

hipError_t run_axpbypcz(double* z, const double* x, double a, const double* y, double b, const double* w,
                        int64_t len, hipStream_t st)
{
    if (len == 0) return hipSuccess;
    hipLaunchKernelGGL(axpbypcz, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, st, z, x, a, y, b, w, len);
    return hipGetLastError();
}

}  // namespace ocpx
