"""ph_brush_amd.titration runs end to end (surrogate training, equilibration,
plain and delayed-acceptance sampling from the shared state) on a small brush."""
import math

import pytest

from ph_brush_amd import system as S
from ph_brush_amd import titration as T

pytestmark = pytest.mark.gpu


def test_titration_tool_runs(gpu):
    res = T.titration(cfg=S.config_small_fp64, ph_values=(7.5,), burn=20000, blocks=4, block_trials=5000,
                      samples_per_ph=4096, epochs=2, log=lambda m: print(m, flush=True))
    (row,) = res["rows"]
    assert row["pH"] == 7.5
    for k in ("alpha_plain", "alpha_da"):
        assert 0.0 < row[k] < 1.0
    assert math.isfinite(row["diff"]) and abs(row["diff"]) < 0.1   # a loose bound: a few blocks only
