// cpr.hpp — the CPR-type preconditioner of the device linear solve and the
// FGMRES pieces around it (cpr.hip).
//
// The reference decouples the FIM system with ABF (each block row times the
// inverse of its diagonal block, src/Decoupling.cpp:23-57, 426-499) and
// solves it with FGMRES(30) preconditioned by FASP's two-stage pressure /
// full-system preconditioner (data/spe5/bsr.fasp: unsmoothed-aggregation AMG
// with pairwise matching on the pressure block, one V-cycle).  Here:
//
//  * PressureAmg: the pressure block of the decoupled system, A_p(i, j) =
//    row 0 of D_i^{-1} times column 0 of A_ij, made an M-matrix (positive
//    off-diagonal entries moved to the diagonal, the diagonal at least the
//    off-diagonal magnitudes' sum).  Its aggregation hierarchy is
//    built once per context on the host from the transmissibility graph
//    (pairwise matching on the strongest connection, twice per level, so
//    aggregates of up to four cells); every solve forms the level matrices
//    on the device (level 0 from the BSR, coarser ones by Galerkin sums over
//    a precomputed gather list) and applies one V-cycle: one weighted-Jacobi
//    sweep before and after the coarse correction, a dense LU on the
//    coarsest level.
//  * the second stage: one multicolour block Gauss-Seidel sweep on the
//    decoupled system (whose diagonal blocks are the identity), colours
//    from a greedy colouring of the block graph.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace ocpx {

class PressureAmg {
public:
    PressureAmg() = default;
    PressureAmg(const PressureAmg&) = delete;
    PressureAmg& operator=(const PressureAmg&) = delete;
    ~PressureAmg();

    // Hierarchy from the level-0 pattern (block rows of the bulk system:
    // rowptr, col with the diagonal first in each row) and the weight of
    // every off-diagonal slot (its connection's transmissibility).
    int setup(int64_t n, const std::vector<int64_t>& rp, const std::vector<int32_t>& col,
              const std::vector<double>& w, std::string& err);
    bool ready() const { return !lv_.empty(); }
    int levels() const { return (int)lv_.size(); }
    int64_t rows(int l) const { return lv_[l].n; }

    // Level matrices of the current system: val the bulk BSR (B x B blocks,
    // row-major), dinv the inverses of its diagonal blocks.
    hipError_t values(int B, const double* val, const double* dinv, hipStream_t st);
    // x = one V-cycle applied to b (level-0 vectors of n doubles)
    hipError_t vcycle(const double* b, double* x, hipStream_t st);

    static constexpr int DENSE_MAX = 64;     // the coarsest level is solved densely up to this size

private:
    struct Level {
        int64_t n = 0, nnz = 0;
        int64_t* rp = nullptr;
        int32_t* ci = nullptr;
        double *a = nullptr, *dinv = nullptr, *b = nullptr, *x = nullptr, *r = nullptr;
        // to the next level (all but the last)
        int32_t* agg = nullptr;              // [n] coarse row of each row
        int64_t *gp = nullptr, *gs = nullptr;   // coarse slot s = sum of a[gs[gp[s] .. gp[s+1])]
        int64_t* mp = nullptr;               // coarse row I: fine rows mi[mp[I] .. mp[I+1])
        int32_t* mi = nullptr;
    };
    std::vector<Level> lv_;
    double* lu_ = nullptr;                   // dense LU of the coarsest level (DENSE_MAX^2)
    int32_t* piv_ = nullptr;
    const int64_t* rp0_ = nullptr;           // level-0 pattern (device, the bulk BSR's)
    std::vector<void*> mem_;
    template <class T>
    int dev(T** p, size_t count, const T* host, std::string& err);
    hipError_t cycle(int l, const double* b, double* x, hipStream_t st);
};

// x_i = rhat_i - Dinv_i sum_{j != i} A_ij x_j for the rows of one colour
hipError_t run_gs_color(int B, const int32_t* rows, int64_t nrows, const int64_t* rowptr, const int32_t* col,
                        const double* val, const double* dinv, const double* rhat, double* x, hipStream_t st);
// out[k] = V_k . w for k < m (V: m vectors of len doubles, stride ldv)
hipError_t run_mdot(const double* V, int64_t ldv, int m, const double* w, int64_t len, double* part, double* out,
                    hipStream_t st);
int mdot_scratch_len();
// w += sum_{k < m} h[k] V_k (h on the device, m <= 32)
hipError_t run_maxpy(double* w, const double* V, int64_t ldv, int m, const double* h, int64_t len, hipStream_t st);
// the pressure entries of a block vector and back: p[i] = v[i B], v = 0 except v[i B] = p[i]
hipError_t run_gather_p(int B, const double* v, double* p, int64_t n, hipStream_t st);
hipError_t run_scatter_p(int B, const double* p, double* v, int64_t n, int64_t len, hipStream_t st);

}  // namespace ocpx
