"""The SPE3 and SPE5 decks (tests/golden/decks, the reference's data/spe3 and
data/spe5) read by opencaeporox_amd.deck, then the device flash and
assembly against the oracle on the same converted inputs.  Pressures are
spread over 30-110% of the EQUIL datum pressure so both decks have single-
and two-phase cells (SPE5: 47 of 147 two-phase, SPE3: 282 of 324)."""
import os

import numpy as np
import pytest

from opencaeporox_amd import deck
from opencaeporox_amd.fim import BATCHED, PLAIN, IsoTFIM
from oracle.oracle import state_from_workload
from tests.conftest import perturbed_oracle_states
from tests.test_gpu_parity import _compare_assembly, _compare_flash, _upload_flash

pytestmark = pytest.mark.gpu

DECKS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "decks")


def _deck_workload(name, seed=3):
    d = deck.read_deck(os.path.join(DECKS, name, f"{name}.data"))
    _, pd = d.datum()
    rng = np.random.Generator(np.random.PCG64(seed))
    P = pd * rng.uniform(0.3, 1.1, d.nb)
    return d.to_workload(seed=seed, P=P)


@pytest.mark.parametrize("name,mode,lanes", [("spe5", PLAIN, 1), ("spe5", BATCHED, 4),
                                             ("spe3", BATCHED, 1), ("spe3", BATCHED, 8)])
def test_deck_flash_and_assembly(oracle, oracle_fma, name, mode, lanes):
    wl = _deck_workload(name)
    tag = f"{name} mode={mode} lanes={lanes}"
    print(tag, flush=True)
    dev = IsoTFIM.from_workload(wl)
    dev.set_flash_mapping(lanes)
    st = state_from_workload(wl)
    others = perturbed_oracle_states(oracle, oracle_fma, wl.fluid, st)
    dev.CalFlash(mode)
    oracle.cal_flash(wl.fluid, st)
    two = (st.a["PHASEEXIST"].reshape(-1, 3)[:, :2].sum(axis=1) == 2).sum()
    assert two > 0, "no two-phase cell: the case would not exercise the split"
    _compare_flash(dev, st, wl.Nc, tag, others, min_cells=int(0.98 * wl.grid.nb), min_two_phase=two - 2)
    _upload_flash(dev, st)
    _compare_assembly(dev, oracle, wl, st, 1.0, tag)
    dev.close()
