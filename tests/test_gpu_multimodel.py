"""Several models in one context: each model's flash, residual and
Jacobian equal those of the model in a context of its own (same kernels,
same mapping), and the merged BSR is block-diagonal by model."""
import numpy as np
import pytest

from opencaeporox_amd import workload as W
from opencaeporox_amd.fim import BATCHED, IsoTFIM, MultiModelBatch

pytestmark = pytest.mark.gpu


def _models():
    out = []
    for m, (dims, seed, pcs) in enumerate((((5, 4, 3), 1, 1.0), ((7, 3, 2), 2, 0.5), ((4, 4, 4), 3, 2.0))):
        rng = np.random.Generator(np.random.PCG64(seed))
        nb = int(np.prod(dims))
        k = np.exp(rng.normal(np.log(100.0), 1.0, nb))
        grid = W.cartesian_grid(*dims, 6.096, k, k, 0.1 * k, rng.uniform(0.1, 0.3, nb))
        fl = W.make_fluid(5, 200 + seed)
        rock = W.make_rock()
        rock.cp1 *= 1 + 0.2 * m
        st = W.initial_state(fl, rock, grid, rng.uniform(20.0, 200.0, nb), rng.dirichlet(np.ones(5), nb), rng)
        out.append(W.Workload(f"m{m}", fl, W.make_sat(pcs), rock, grid, st))
    return out


@pytest.mark.parametrize("lanes", [4, 8])
def test_merged_context_equals_separate_contexts(lanes):
    wls = _models()
    merged = IsoTFIM.from_workloads(wls)
    merged.set_flash_mapping(lanes)
    merged.newton_pass(1.0, BATCHED)
    rp, col, val, rhs = merged.bsr()
    off = merged.offsets
    for m, wl in enumerate(wls):
        one = IsoTFIM.from_workload(wl)
        one.set_flash_mapping(lanes)
        one.set_flash_stages(False)                 # the merged context's kernel alone
        one.newton_pass(1.0, BATCHED)
        a, b = off[m], off[m + 1]
        for f in ("PHASEEXIST", "S", "XIJ", "DSEC_DPRI", "KR", "PORO"):
            L = len(one.get(f)) // wl.grid.nb
            x = merged.get(f)[a * L:b * L]
            y = one.get(f)
            assert np.allclose(x, y, rtol=1e-12, atol=0), f"model {m} field {f}"
        rp1, col1, val1, rhs1 = one.bsr()
        # model m's rows of the merged BSR = its own BSR with columns offset
        s0, s1 = rp[a], rp[b]
        assert np.array_equal(rp[a:b + 1] - s0, rp1)
        assert np.array_equal(col[s0:s1] - a, col1)
        assert np.allclose(val[s0:s1], val1, rtol=1e-12, atol=1e-300)
        assert np.allclose(rhs[a * one.bs:b * one.bs], rhs1, rtol=1e-12, atol=1e-300)
        one.close()


def test_multimodel_batch_groups_by_component_count():
    wls = W.uneven_batch(nmodels=6, seed=3, cmin=200, cmax=2000)
    batch = MultiModelBatch(wls)
    assert sorted(batch.groups) == sorted({w.fluid.nc for w in wls})
    batch.autotune()
    batch.UpdateLastTimeStep()
    batch.newton_pass(1.0, graph=True)
    batch.sync()
    for m, wl in enumerate(wls):
        one = IsoTFIM.from_workload(wl)
        d, _ = batch.groups[wl.fluid.nc]
        one.set_flash_mapping(d.flash_mapping())
        one.set_flash_stages(False)
        one.newton_pass(1.0, BATCHED)
        for f in ("PHASEEXIST", "S", "DSEC_DPRI"):
            assert np.allclose(batch.get(m, f), one.get(f), rtol=1e-12, atol=0), f"model {m} {f}"
        one.close()
    batch.close()


def test_batch_graph_replays_equal_the_contexts_passes():
    """One graph launch per Newton step for every Nc group (ocpx_batch_newton_pass),
    replayed three times from the last time step with eager work in between:
    each replay equals the groups' own passes run one after another."""
    wls = W.uneven_batch(nmodels=8, seed=5, cmin=300, cmax=3000)
    a, b = MultiModelBatch(wls), MultiModelBatch(wls)
    for d, _ in a.groups.values():
        d.set_flash_mapping(4)
    for d, _ in b.groups.values():
        d.set_flash_mapping(4)
    a.UpdateLastTimeStep()
    b.UpdateLastTimeStep()
    b.newton_pass(1.0, graph=False)
    b.sync()
    ref = {Nc: [d.get(f) for f in ("S", "DSEC_DPRI")] + list(d.bsr()[2:]) for Nc, (d, _) in b.groups.items()}
    for rep in range(3):
        a.ResetToLastTimeStep()
        a.newton_pass(1.0, graph=True)
        for Nc, (d, _) in a.groups.items():
            got = [d.get(f) for f in ("S", "DSEC_DPRI")] + list(d.bsr()[2:])
            for x, y in zip(got, ref[Nc]):
                assert np.array_equal(x, y), f"replay {rep}, Nc = {Nc}"
    a.close()
    b.close()
