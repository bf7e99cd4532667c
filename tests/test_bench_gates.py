"""The FP32-vs-FP64 equal-accuracy gate bench.py reports: two runs drawn
from one distribution pass, a 1% dose bias or a 0.6% integral energy offset
fail; voxels are chosen on the mean of both runs, so the selection does not
bias z; the summed-variance form (beamlet and multi-beam sums) gives the
same answer as the batch-sum form."""
import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
import bench  # noqa: E402


def _run(rng, mean, rel_sd, nb, bias=1.0):
    """Batch sums of nb batches with per-voxel means `mean` (region 0 first)."""
    x = rng.normal(mean * bias, rel_sd * mean, size=(nb, mean.size))
    s = np.concatenate([[0.0], x.sum(0)])
    s2 = np.concatenate([[0.0], (x * x).sum(0)])
    return s, s2, nb


def test_gates_pass_for_one_distribution_and_fail_for_a_bias():
    rng = np.random.default_rng(5)
    nvox = 40000
    rho = np.where(np.arange(nvox) % 7 == 0, 0.001, 1.0)       # some air voxels
    mean = np.exp(-np.linspace(0, 4, nvox)) + 0.05
    ok = bench.equal_accuracy(_run(rng, mean, 0.05, 150), _run(rng, mean, 0.05, 150), rho=rho)
    assert ok["pass"], ok
    assert abs(ok["expected_if_identical"] - 0.9537) < 1e-3
    assert abs(ok["fraction_within_2sigma"] - ok["expected_if_identical"]) < 0.005
    assert abs(ok["z_mean"]) < 0.02
    biased = bench.equal_accuracy(_run(rng, mean, 0.05, 150, bias=1.01), _run(rng, mean, 0.05, 150), rho=rho)
    assert not biased["pass"] and biased["fraction_within_2sigma"] < 0.94
    # a uniform 0.6% energy offset with tiny noise: voxels disagree and the integral is off
    shifted = bench.equal_accuracy(_run(rng, mean, 1e-5, 150, bias=1.006), _run(rng, mean, 1e-5, 150), rho=rho)
    assert not shifted["pass"]
    assert abs(shifted["integral_energy_fp32_over_fp64"] - 1.006) < 1e-3


def test_summed_variance_form_matches_batch_form():
    rng = np.random.default_rng(9)
    mean = np.exp(-np.linspace(0, 3, 30000)) + 0.1
    a, b = _run(rng, mean, 0.05, 120), _run(rng, mean, 0.05, 120)

    def mv(r):
        s, s2, nb = r
        m = s[1:] / nb
        return ("mv", m, np.maximum(s2[1:] / nb - m * m, 0.0) / (nb - 1), nb - 1)
    x = bench.equal_accuracy(a, b)
    y = bench.equal_accuracy(mv(a), mv(b))
    assert x["fraction_within_2sigma"] == y["fraction_within_2sigma"] and x["pass"] == y["pass"]
