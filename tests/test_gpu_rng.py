"""The chains' device engine (MT19937 with libstdc++'s uniform_int_distribution
and generate_canonical<double,53>, MC/src/mc.cu:62-65, :368-369) against
std::mt19937 draw by draw, through phb_rng_draws: 1.2 M draws mixing canonical
doubles, small ranges and ranges near 2^32/3 where a third of the Lemire draws
fall in the rejection zone, so both the fixed-position rounds and the serial
redraw of a round with a rejection run many times.  Bar: every draw and the
engine state afterwards bit-identical."""
import numpy as np
import pytest

from oracle import pyoracle as O
from ph_brush_amd import system as S

pytestmark = pytest.mark.gpu


def _his(n, seed, p_hard):
    rng = np.random.default_rng(seed)
    his = rng.integers(0, 40000, size=n).astype(np.int32)
    his[rng.random(n) < 0.35] = -1
    hard = rng.random(n) < p_hard
    his[hard] = rng.choice(np.array([1431655765, 1431655766, 2 ** 31 - 1, 1500000000], np.int32), hard.sum())
    return his


@pytest.mark.parametrize("seed,n,p_hard", [(1, 1_200_000, 0.002), (2025, 200_000, 0.05)])
def test_device_engine_stream_equals_std_mt19937(gpu, seed, n, p_hard):
    his = _his(n, seed, p_hard)
    st0 = S.rng_seed_state(seed)
    got, after, nser = S.rng_draws(st0, his)
    exp, exp_after = O.mixed_draws(st0, his)
    bad = np.flatnonzero(got != exp)
    assert bad.size == 0, (bad[:5], got[bad[:5]], exp[bad[:5]])
    assert np.array_equal(after, exp_after)
    assert nser > 100, nser            # rounds that took the serial path
    assert nser < (n + 63) // 64       # ... and rounds that did not


def test_device_engine_resumes_mid_stream(gpu):
    """A state taken after an odd number of outputs (index not a multiple of the
    block) resumes exactly: two calls equal one."""
    his = _his(50_001, 7, 0.01)
    st0 = S.rng_seed_state(7)
    a, mid, _ = S.rng_draws(st0, his[:12_345])
    b, end, _ = S.rng_draws(mid, his[12_345:])
    c, end1, _ = S.rng_draws(st0, his)
    assert np.array_equal(np.concatenate([a, b]), c) and np.array_equal(end, end1)
    assert mid[624] not in (0, 624)
