// ocpx_isot_fim.cpp — the reference-side binding of the C ABI (include/ocpx.h):
// what IsoT_FIM (src/IsoThermalMethod.cpp:711-1517) calls in place of its
// per-cell and per-connection loops.  It is compiled into the reference, not
// into libocpx.so, and uses only the reference's own types:
//   BulkVarSet   (include/BulkVarSet.hpp)     the bulk state, cell-major
//   BulkConnPair (include/BulkConnVarSet.hpp) BulkConn::iteratorConn's entries
//   OCPNRresidual (include/OCPNRresidual.hpp) NR.res: resAbs and the norms
// and, for the block-CSR hand-over, any linear system with the
// LinearSystem interface NewDiag / NewOffDiag / CopyRhs
// (include/LinearSystem.hpp:59-78).
//
// The call sites in IsoT_FIM (member names as the reference has them):
//   Setup:            ocpx_shim::Setup(dev, fl, sat, rock, rs.bulk.vs, rs.conn.iteratorConn)
//   UpdateProperty:   ocpx_shim::PushState(dev, rs.bulk.vs); ocpx_shim::UpdateProperty(dev);
//                     ocpx_shim::PullState(dev, rs.bulk.vs)
//   CalRes:           ocpx_shim::CalRes(dev, dt, NR.res)
//   AssembleMatBulks: ocpx_shim::AssembleMat(dev, dt, ls)
//   GetSolution:      ocpx_shim::GetSolution(dev, u, ctrlNR.DSmax())
//   Update/ResetToLastTimeStep: ocpx_update_last_step / ocpx_reset_to_last_step
//
// Every function returns the library's status (0, or a negative OCPX_E_*);
// the caller turns a failure into OCP_ABORT(ocpx_last_error(dev.ctx)) or a
// time-step cut, as it does for its own failures.
#include <cstdint>
#include <vector>

#include "BulkConnVarSet.hpp"
#include "BulkVarSet.hpp"
#include "OCPNRresidual.hpp"
#include "ocpx.h"

namespace ocpx_shim {

struct Binding {
    ocpx_ctx* ctx = nullptr;
    int64_t nb = 0, nconn = 0;
    int bsize = 0;               // nc + 1
};

// IsoT_FIM::Setup / AllocateReservoir (src/IsoThermalMethod.cpp:711-715,
// 905-998): the rank's cells and its connections in iteratorConn order
// (bId < eId, per cell x+, y+, z+), transmissibilities as
// BulkConnTransMethod01::CalTrans left them (src/BulkConnFunc.cpp:15-67).
inline int Setup(Binding& dev, const ocpx_fluid& fl, const ocpx_sat& sat, const ocpx_rock& rock,
                 const BulkVarSet& bvs, const std::vector<BulkConnPair>& iteratorConn, void* hip_stream = nullptr)
{
    std::vector<int32_t> cb, ce;
    std::vector<double> trans;
    cb.reserve(iteratorConn.size());
    ce.reserve(iteratorConn.size());
    trans.reserve(iteratorConn.size());
    for (const BulkConnPair& c : iteratorConn) {
        cb.push_back((int32_t)c.BId());
        ce.push_back((int32_t)c.EId());
        trans.push_back(c.Trans());
    }
    ocpx_grid g{};
    g.nb = (int64_t)bvs.nb;
    g.nconn = (int64_t)cb.size();
    g.v = bvs.v.data();
    g.depth = bvs.depth.data();
    g.poroInit = bvs.poroInit.data();
    g.conn_b = cb.data();
    g.conn_e = ce.data();
    g.trans = trans.data();
    const int rc = ocpx_ctx_create(&fl, &sat, &rock, &g, hip_stream, &dev.ctx);
    if (rc != OCPX_OK) return rc;
    dev.nb = g.nb;
    dev.nconn = g.nconn;
    dev.bsize = (int)bvs.nc + 1;
    return ocpx_autotune_flash(dev.ctx, nullptr);
}

// the BulkVarSet arrays the flash reads (SetVarSet, src/OCPMixtureMethodComp.cpp:671-687)
inline int PushState(Binding& dev, const BulkVarSet& bvs)
{
    int rc;
    if ((rc = ocpx_set_field(dev.ctx, OCPX_F_P, bvs.P.data())) ||
        (rc = ocpx_set_field(dev.ctx, OCPX_F_T, bvs.T.data())) ||
        (rc = ocpx_set_field(dev.ctx, OCPX_F_NI, bvs.Ni.data())) ||
        (rc = ocpx_set_field(dev.ctx, OCPX_F_LNI, bvs.lNi.data())) ||
        (rc = ocpx_set_field(dev.ctx, OCPX_F_S, bvs.S.data())) ||
        (rc = ocpx_set_field(dev.ctx, OCPX_F_XIJ, bvs.xij.data())))
        return rc;
    const std::vector<uint32_t> pe(bvs.phaseExist.begin(), bvs.phaseExist.end());   // OCP_BOOL
    return ocpx_set_field(dev.ctx, OCPX_F_PHASEEXIST, pe.data());
}

// IsoT_FIM::UpdateProperty's CalFlash + CalKrPc + CalRock (src/IsoThermalMethod.cpp:811-836)
inline int UpdateProperty(Binding& dev)
{
    int rc;
    if ((rc = ocpx_flash_der(dev.ctx, OCPX_FLASH_BATCHED))) return rc;
    return ocpx_krpc_der(dev.ctx);          // kr / Pc and the rock (CalRock is fused)
}

// the flash outputs the rest of the simulator reads back (PassFlashValue,
// src/IsoThermalMethod.cpp:1027-1071)
inline int PullState(Binding& dev, BulkVarSet& bvs)
{
    int rc;
    std::vector<uint32_t> pe(bvs.phaseExist.size());
    if ((rc = ocpx_get_field(dev.ctx, OCPX_F_PHASEEXIST, pe.data()))) return rc;
    for (size_t k = 0; k < pe.size(); k++) bvs.phaseExist[k] = (OCP_BOOL)pe[k];
    struct { int f; std::vector<OCP_DBL>* v; } out[] = {
        {OCPX_F_S, &bvs.S}, {OCPX_F_XIJ, &bvs.xij}, {OCPX_F_NT, &bvs.Nt}, {OCPX_F_VF, &bvs.vf},
        {OCPX_F_RHO, &bvs.rho}, {OCPX_F_XI, &bvs.xi}, {OCPX_F_MU, &bvs.mu}, {OCPX_F_KR, &bvs.kr},
        {OCPX_F_PC, &bvs.Pc}, {OCPX_F_PJ, &bvs.Pj}, {OCPX_F_DSEC_DPRI, &bvs.dSec_dPri},
        {OCPX_F_PORO, &bvs.poro}, {OCPX_F_ROCKVP, &bvs.rockVp}};
    for (const auto& o : out)
        if ((rc = ocpx_get_field(dev.ctx, o.f, o.v->data()))) return rc;
    return OCPX_OK;
}

// IsoT_FIM::CalRes (src/IsoThermalMethod.cpp:1090-1185), bulk part: the
// negated resAbs and the relative norms into NR.res
inline int CalRes(Binding& dev, double dt, OCPNRresidual& res)
{
    ocpx_resnorms n{};
    const int rc = ocpx_cal_res(dev.ctx, dt, res.resAbs.data(), &n);
    if (rc) return rc;
    res.maxRelRes_V = n.maxRelRes_V;
    res.maxId_V = (OCP_INT)n.maxId_V;
    res.maxRelRes_N = n.maxRelRes_N;
    res.maxId_N = (OCP_INT)n.maxId_N;
    return OCPX_OK;
}

// IsoT_FIM::AssembleMatBulks + CopyRhs (src/IsoThermalMethod.cpp:1187-1269,
// 742-752): the device BSR in OCPMatrix push order (diagonal first, then the
// row's connections) handed to the linear system block by block
template <class LinearSystemT>
int AssembleMat(Binding& dev, double dt, LinearSystemT& ls)
{
    int rc;
    if ((rc = ocpx_assemble(dev.ctx, dt))) return rc;
    int64_t nnzb = 0;
    if ((rc = ocpx_get_bsr(dev.ctx, &nnzb, nullptr, nullptr, nullptr, nullptr))) return rc;
    const int64_t b2 = (int64_t)dev.bsize * dev.bsize;
    std::vector<int64_t> rowptr(dev.nb + 1);
    std::vector<int32_t> col(nnzb);
    std::vector<OCP_DBL> val(nnzb * b2), rhs(dev.nb * dev.bsize);
    if ((rc = ocpx_get_bsr(dev.ctx, &nnzb, rowptr.data(), col.data(), val.data(), rhs.data()))) return rc;
    for (int64_t n = 0; n < dev.nb; n++)
        for (int64_t s = rowptr[n]; s < rowptr[n + 1]; s++) {
            const std::vector<OCP_DBL> blk(val.begin() + s * b2, val.begin() + (s + 1) * b2);
            if (s == rowptr[n]) ls.NewDiag((OCP_USI)n, blk);
            else ls.NewOffDiag((OCP_USI)n, (OCP_USI)col[s], blk);
        }
    ls.CopyRhs(rhs);
    return OCPX_OK;
}

// IsoT_FIM::GetSolution (src/IsoThermalMethod.cpp:1282-1418), bulk part:
// the dSmax chop and the S / xij / P / Ni update, dSmax = ctrlNR.DSmax()
inline int GetSolution(Binding& dev, const std::vector<OCP_DBL>& u, OCP_DBL dSmax)
{
    return ocpx_get_solution(dev.ctx, u.data(), dSmax);
}

// SolveLinearSystem with the system left on the device: 1e-3 and 200
// iterations as the reference's decks (data/spe5/bsr.fasp).  *iters is the
// iteration count, negative when the solve did not converge; nothing is then
// applied and the caller cuts dt (src/IsoThermalMethod.cpp:776-781).
inline int SolveAndUpdate(Binding& dev, OCP_DBL dSmax, int32_t* iters)
{
    double relres = 0;
    const int rc = ocpx_linear_solve(dev.ctx, 1e-3, 200, nullptr, iters, &relres);
    if (rc || *iters < 0) return rc;
    return ocpx_get_solution(dev.ctx, nullptr, dSmax);
}

inline void Destroy(Binding& dev)
{
    ocpx_ctx_destroy(dev.ctx);
    dev.ctx = nullptr;
}

}  // namespace ocpx_shim
