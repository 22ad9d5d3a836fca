"""The reference-side binding (opencaeporox_amd/shim/ocpx_isot_fim.cpp)
compiles against the reference's own headers, its AssembleMat hand-over
instantiates with a sink of LinearSystem's signatures, and the member names
its call sites use exist in the reference.  Runs where the reference is
present; nothing is built or linked from the reference here."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
SHIM = os.path.join(ROOT, "opencaeporox_amd", "shim", "ocpx_isot_fim.cpp")

pytestmark = pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "include")),
                                reason="the reference headers are present only in the build container")

# a linear system with the NewDiag / NewOffDiag / CopyRhs overloads that
# include/LinearSystem.hpp:59-78 declares
SINK = r'''
#include "%s"
struct Sink {
    void NewDiag(const OCP_USI& n, const OCP_DBL& v) {}
    void NewDiag(const OCP_USI& n, const std::vector<OCP_DBL>& v) {}
    void NewOffDiag(const OCP_USI& b, const OCP_USI& e, const OCP_DBL& v) {}
    void NewOffDiag(const OCP_USI& b, const OCP_USI& e, const std::vector<OCP_DBL>& v) {}
    void CopyRhs(const std::vector<OCP_DBL>& rhs) {}
};
int use(ocpx_shim::Binding& d, Sink& s) { return ocpx_shim::AssembleMat(d, 1.0, s); }
'''


def _syntax(src, tmp_path):
    f = tmp_path / "t.cpp"
    f.write_text(src)
    r = subprocess.run(["g++", "-std=c++14", "-fsyntax-only", "-Wall", "-Werror", f"-I{REF}/include",
                        f"-I{ROOT}/include", str(f)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]


def test_shim_compiles_against_the_reference_headers(tmp_path):
    _syntax(f'#include "{SHIM}"\n', tmp_path)


def test_assemble_hand_over_instantiates_with_linear_system_signatures(tmp_path):
    _syntax(SINK % SHIM, tmp_path)


def test_call_site_member_names_exist():
    """rs.bulk.vs, rs.conn.iteratorConn, NR.res, ctrlNR.DSmax() and the
    BulkVarSet fields the shim moves are the reference's names."""
    def has(header, pattern):
        return re.search(pattern, open(os.path.join(REF, "include", header)).read()) is not None
    assert has("Bulk.hpp", r"\bBulkVarSet\s+vs;")
    assert has("BulkConn.hpp", r"\bvector<BulkConnPair>\s+iteratorConn;")
    assert has("OCPNRsuite.hpp", r"\bOCPNRresidual\s+res;")
    assert has("OCPControlNR.hpp", r"\bDSmax\(\)")
    assert has("IsoThermalMethod.hpp", r"GetSolution\(Reservoir& rs, vector<OCP_DBL>& u, const ControlNR& ctrlNR\)")
    shim = open(SHIM).read()
    for m in re.findall(r"bvs\.([A-Za-z_]+)", shim):
        assert has("BulkVarSet.hpp", rf"\b{m}\b"), m
