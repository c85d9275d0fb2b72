"""bench.py's command line: more than one GPU is refused with a message (the
build runs one replica chain on one GPU; independent replicas would shard
across GPUs without communication), and the traffic record it reports is read
from the committed counter-pass summary."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, cwd=ROOT, timeout=120)
    assert r.returncode == 2 and r.stdout == ""
    assert "--gpus 2 is not supported" in r.stderr and "one GPU" in r.stderr


def test_traffic_record_is_read():
    sys.path.insert(0, ROOT)
    import bench
    with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
        rec = json.load(f)["k_da_chain"]
    assert bench._traffic("k_da_chain") == rec["bytes_per_launch"] > 0
    assert bench._traffic("no_such_kernel") is None


def test_integrated_autocorrelation_of_ar1():
    """tau = 1/2 + sum rho(t) of an AR(1) series x_t = phi x_{t-1} + e_t is
    (1 + phi) / (2 (1 - phi)) in closed form."""
    import numpy as np
    from scipy.signal import lfilter
    import bench
    rng = np.random.default_rng(0)
    for phi in (0.0, 0.5, 0.9, 0.98):
        x = lfilter([1.0], [1.0, -phi], rng.standard_normal(400000))
        tau, m = bench.integrated_autocorr(x)
        exact = (1 + phi) / (2 * (1 - phi))
        assert abs(tau - exact) <= 0.06 * exact, (phi, tau, exact)
        assert m >= 5 * tau - 1


def test_ess_of_chain_reconstructs_the_series():
    """N_prot and E rebuilt from a record: +-1 per accepted forward/backward move,
    E summed over the accepted trials' Delta-E; rejected trials change nothing."""
    import numpy as np
    import bench
    rng = np.random.default_rng(1)
    n = 50000
    fwd = rng.integers(0, 2, n).astype(np.int8)
    acc = (rng.random(n) < 0.3).astype(np.int8)
    dE = np.where(acc != 0, rng.standard_normal(n), np.nan)
    out = bench.ess_of_chain(100, 5.0, fwd, acc, dE, 1e5, "synthetic")
    assert out["trials"] == n
    for k in ("nprot", "E"):
        assert out[k]["ess"] > 0 and out[k]["ess_per_s"] == 1e5 / (2 * out[k]["tau_int_trials"])
    # a random walk: tau grows with the series length (N_prot is not stationary here), ESS << n
    assert out["nprot"]["ess"] < n / 10
