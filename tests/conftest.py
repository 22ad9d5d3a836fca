import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def oracle():
    from oracle.oracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def oracle_fma():
    """The oracle built with fused multiply-add contraction."""
    import oracle.oracle as OO
    path = os.path.join(OO.HERE, "build", "libocpx_oracle_fma.so")
    if not os.path.exists(path):
        pytest.skip("oracle FMA build missing")
    return OO.Oracle(path)


@pytest.fixture(scope="session")
def ref():
    from oracle.oracle import Ref
    if not Ref.available():
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    return Ref()


def close(a, b, rtol=1e-10, scale=None):
    """|a-b| <= rtol*(|b| + scale); scale defaults to FLOOR*max|b| (entries
    more than 1/FLOOR below the field's largest are held to rtol*FLOOR of it,
    see oracle/parity.py)."""
    from oracle.parity import FLOOR
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    if scale is None:
        scale = FLOOR * (np.max(np.abs(b)) if b.size else 0.0)
    err = np.abs(a - b)
    bound = rtol * (np.abs(b) + scale)
    ok = (err <= bound) | (a == b)
    return bool(np.all(ok)), float(np.max(err / np.maximum(np.abs(b) + scale, 1e-300))) if b.size else 0.0


from oracle.parity import (FLOOR, PARITY_RTOL, SPLIT_ILL, SPREAD_MULT, cell_spread, degenerate,  # noqa: E402,F401
                           label_stable, perturbed_oracle_states, rel_err, rounding_sensitive, rounding_spread)
