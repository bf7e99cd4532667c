"""The FP32-vs-FP64 accuracy gates bench.py reports: two runs drawn from one
distribution pass, a 1% dose bias or a 0.6% integral energy offset fail."""
import ctypes as C
import sys
from pathlib import Path
from types import SimpleNamespace

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
import bench  # noqa: E402


def _prob(rho):
    buf = (C.c_double * rho.size)(*rho)
    ph = SimpleNamespace(isize=rho.size, jsize=1, ksize=1, rho=C.cast(buf, C.POINTER(C.c_double)))
    return SimpleNamespace(ph=ph, _keep=buf)


def _run(rng, mean, rel_sd, nb, bias=1.0):
    """Batch sums of nb batches with per-voxel means `mean` (region 0 first)."""
    x = rng.normal(mean * bias, rel_sd * mean, size=(nb, mean.size))
    s = np.concatenate([[0.0], x.sum(0)])
    s2 = np.concatenate([[0.0], (x * x).sum(0)])
    return s, s2, nb


def test_gates_pass_for_one_distribution_and_fail_for_a_bias():
    rng = np.random.default_rng(5)
    nvox = 20000
    rho = np.where(np.arange(nvox) % 7 == 0, 0.001, 1.0)       # some air voxels
    mean = np.exp(-np.linspace(0, 4, nvox)) + 0.05
    prob = _prob(rho)
    ok = bench.accuracy_gates(prob, _run(rng, mean, 0.05, 30), _run(rng, mean, 0.05, 30))
    assert ok["pass"], ok
    assert abs(ok["expected_if_identical"] - 0.9498) < 1e-3
    assert abs(ok["fraction_within_2sigma"] - ok["expected_if_identical"]) < 0.01
    biased = bench.accuracy_gates(prob, _run(rng, mean, 0.05, 30, bias=1.01), _run(rng, mean, 0.05, 30))
    assert not biased["pass"] and biased["fraction_within_2sigma"] < 0.94
    # a uniform 0.6% energy offset with tiny noise: voxels disagree and the integral is off
    shifted = bench.accuracy_gates(prob, _run(rng, mean, 1e-5, 30, bias=1.006), _run(rng, mean, 1e-5, 30))
    assert not shifted["pass"]
    assert abs(shifted["integral_energy_fp32_over_fp64"] - 1.006) < 1e-3
