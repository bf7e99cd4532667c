"""The .3ddose comparator (omc_compare_3ddose / bin/omc_compare3ddose):
z = (d_a - d_b) / sqrt((u_a d_a)^2 + (u_b d_b)^2) over the voxels above 10%
of the maximum of the two doses' mean, the 95% / 2 sigma and 0.5% integral
gates, on files written in the reference's .3ddose format."""
import subprocess
from pathlib import Path

import numpy as np

from cases import write_case
from ompmc_amd import _native as N

ROOT = Path(__file__).resolve().parents[1]
CLI = ROOT / "ompmc_amd" / "bin" / "omc_compare3ddose"


def _write(oracle, prob, path, dose, unc):
    oracle.write_3ddose(prob, path, np.ascontiguousarray(dose), np.ascontiguousarray(unc))


def test_known_answers(oracle, tmp_path):
    """Identical files: every voxel at z = 0.  A copy whose doses are shifted
    by exactly 3 combined sigma in a known set of voxels: those voxels fall
    outside 2 sigma and the fraction drops by their share; a 1% scale fails
    the integral gate."""
    prob = oracle.load_problem(write_case("water2MeV", tmp_path / "c", ncase=100))
    try:
        n = prob.ph.isize * prob.ph.jsize * prob.ph.ksize
        rng = np.random.default_rng(1)
        dose = 1.0 + rng.random(n)
        unc = np.full(n, 0.01)
        _write(oracle, prob, tmp_path / "a.3ddose", dose, unc)
        r = N.compare_3ddose(tmp_path / "a.3ddose", tmp_path / "a.3ddose")
        assert r["voxels"] == n and r["frac_within_2sigma"] == 1.0 and r["z_max_abs"] == 0.0 and r["pass"]
        # shift 3% of the voxels by +3 sigma of the pair: d_b = d_a (1 + 3*0.01*sqrt(2)) / ...
        k = np.arange(n) % 33 == 0
        d2 = dose.copy()
        f = 3.0 * 0.01 * np.sqrt(2.0)
        d2[k] = dose[k] * (1.0 + f) / 1.0
        _write(oracle, prob, tmp_path / "b.3ddose", d2, unc)
        r = N.compare_3ddose(tmp_path / "a.3ddose", tmp_path / "b.3ddose")
        assert abs(r["frac_within_2sigma"] - (1.0 - k.mean())) < 1e-12
        assert r["pass"] == (1.0 - k.mean() >= 0.95) and 2.0 < r["z_max_abs"] < 3.1
        _write(oracle, prob, tmp_path / "c.3ddose", dose * 1.01, unc)
        r = N.compare_3ddose(tmp_path / "a.3ddose", tmp_path / "c.3ddose")
        assert abs(r["integral_ratio"] - 1 / 1.01) < 1e-5 and not r["pass"]
        out = subprocess.run([str(CLI), str(tmp_path / "a.3ddose"), str(tmp_path / "c.3ddose")], capture_output=True,
                             text=True)
        assert out.returncode == 1 and "pass = 0" in out.stdout
        out = subprocess.run([str(CLI), str(tmp_path / "a.3ddose"), str(tmp_path / "a.3ddose")], capture_output=True,
                             text=True)
        assert out.returncode == 0 and "fraction_within_2sigma = 1.000000" in out.stdout
    finally:
        oracle.free_problem(prob)


def test_two_reference_runs_agree(oracle, tmp_path):
    """Two RANMAR runs of the oracle (the reference's .3ddose byte for byte)
    with different seeds on 6 MeV electrons (whose deposited energy varies
    little from history to history), weighted by the phantom's masses, and
    the same run against itself scaled by 2%."""
    stem = write_case("electron6MeV", tmp_path / "c", ncase=20000)
    inp = Path(str(stem) + ".inp")
    inp.write_text(inp.read_text().replace("nbatch = 10", "nbatch = 40"))
    outs = []
    for seeds in ((97, 33), (11, 5)):
        inp.write_text(inp.read_text().replace("rng seeds = 97 33", f"rng seeds = {seeds[0]} {seeds[1]}"))
        prob = oracle.load_problem(stem)
        try:
            s, s2, _ = oracle.run_dosxyz_ranmar(prob)
            d, u = oracle.dose_from_sums(prob, s, s2)
            _write(oracle, prob, tmp_path / f"r{seeds[0]}.3ddose", d, u)
            outs.append((d, u))
            ph = prob.ph
            r = None
            if len(outs) == 2:
                r = N.compare_3ddose(tmp_path / "r97.3ddose", tmp_path / "r11.3ddose", phantom=ph)
                _write(oracle, prob, tmp_path / "scaled.3ddose", d * 1.02, u)
                bad = N.compare_3ddose(tmp_path / "r11.3ddose", tmp_path / "scaled.3ddose", phantom=ph)
        finally:
            oracle.free_problem(prob)
    assert r["voxels"] > 100
    assert r["frac_within_2sigma"] > 0.9 and abs(r["z_mean"]) < 0.2 and abs(r["integral_ratio"] - 1) < 0.005
    assert not bad["pass"] and abs(bad["integral_ratio"] - 1 / 1.02) < 1e-4
