"""ph_brush_amd.titration runs end to end (surrogate training, equilibration,
plain and delayed-acceptance sampling from the shared state) on a small brush."""
import math

import pytest

from ph_brush_amd import system as S
from ph_brush_amd import titration as T

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("local_pairs", [True, False])
def test_titration_tool_runs(gpu, local_pairs):
    """Both forms of the surrogate; alpha of the two chains agrees within five
    standard errors of the difference of the block means (12 blocks of 20,000
    trials), or within 0.005 (the accuracy bar of alpha(pH)) when that is wider."""
    res = T.titration(cfg=S.config_small_fp64, ph_values=(7.5,), burn=20000, blocks=12, block_trials=20000,
                      samples_per_ph=4096, epochs=4, log=lambda m: print(m, flush=True), local_pairs=local_pairs)
    (row,) = res["rows"]
    assert row["pH"] == 7.5
    for k in ("alpha_plain", "alpha_da"):
        assert 0.0 < row[k] < 1.0
    sem = math.hypot(row["sem_plain"], row["sem_da"])
    assert math.isfinite(row["diff"]) and abs(row["diff"]) <= max(5.0 * sem, 0.005), row
    assert row["acceptance_da"] > 0.0 and row["speedup_accepted"] > 0.0
