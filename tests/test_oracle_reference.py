"""The CPU oracle against the reference.

The oracle in RANMAR mode, driven by the reference's serial history loop,
must reproduce the reference omc_dosxyz .3ddose file byte for byte.  The
expected hashes were produced by the reference build itself
(tests/golden/make_golden.py); where the reference build is present
(oracle/_ref) it is also run live."""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from cases import CASES, write_case

GOLDEN = Path(__file__).resolve().parent / "golden"


def oracle_3ddose(oracle, stem, out):
    prob = oracle.load_problem(stem)
    try:
        s, s2, cnt = oracle.run_dosxyz_ranmar(prob)
        d, u = oracle.dose_from_sums(prob, s, s2)
        oracle.write_3ddose(prob, out, d, u)
    finally:
        oracle.free_problem(prob)
    return Path(out).read_bytes(), cnt


@pytest.mark.parametrize("case", sorted(CASES))
def test_oracle_ranmar_reproduces_reference_3ddose(oracle, tmp_path, case):
    gold = json.loads((GOLDEN / f"ref_dose_{case}.json").read_text())
    stem = write_case(case, tmp_path)
    text, cnt = oracle_3ddose(oracle, stem, tmp_path / "oracle.3ddose")
    assert hashlib.sha256(text).hexdigest() == gold["sha256"], f"{case}: oracle output differs from the reference"
    if oracle.ref_available():
        oracle.run_reference(stem, "live")
        assert (tmp_path / "live.3ddose").read_bytes() == text


def test_oracle_per_history_counts_match_survey_c1(oracle, tmp_path):
    # C1 per-history means measured on the reference by instrumentation:
    # ~35.4 ausgab calls, ~54.0 photon and ~0.95 charged region changes,
    # ~516 random numbers per history
    stem = write_case("C1", tmp_path, ncase=20000)
    prob = oracle.load_problem(stem)
    try:
        _, _, cnt = oracle.run_dosxyz_ranmar(prob)
    finally:
        oracle.free_problem(prob)
    n = cnt.n_hist
    assert abs(cnt.n_dep / n - 35.4) < 1.0
    assert abs(cnt.n_cross_photon / n - 54.0) < 1.5
    assert abs(cnt.n_cross_charged / n - 0.95) < 0.1
    assert abs(cnt.n_draws / n - 516) < 15


def test_oracle_philox_energy_conservation(oracle, tmp_path):
    """Per history: energy deposited (phantom + region 0) + photon energy
    leaving = incident kinetic energy (+2 m c^2 for positrons)."""
    import ctypes as C
    for case, e_in in (("water2MeV", 2.0), ("electron6MeV", 6.0), ("positron6MeV", 6.0 + 2 * 0.5109989461)):
        stem = write_case(case, tmp_path / case, ncase=300)
        prob = oracle.load_problem(stem)
        L = oracle.lib()
        ctx = L.ora_create(C.byref(prob.tab), C.byref(prob.reg), C.byref(prob.ph), 1, oracle.PHILOX)
        try:
            L.ora_seed_philox(ctx, 1, 2)
            led = np.zeros(4 * prob.nhist)
            L.ora_set_ledger(ctx, led.ctypes.data_as(C.POINTER(C.c_double)), prob.nhist)
            s = np.zeros(prob.reg.nreg)
            s2 = np.zeros(prob.reg.nreg)
            P = C.POINTER(C.c_double)
            assert L.ora_run_dosxyz(ctx, C.byref(prob.src), prob.nperbatch, prob.nbatch,
                                    s.ctypes.data_as(P), s2.ctypes.data_as(P)) == 0
        finally:
            L.ora_destroy(ctx)
            oracle.free_problem(prob)
        tot = led.reshape(-1, 4)[:, :3].sum(1)
        assert np.allclose(tot, e_in, rtol=1e-12, atol=1e-12), case


def test_oracle_philox_is_order_independent(oracle, tmp_path):
    """Per-particle streams: the same histories on 1 or 4 threads give the
    same per-voxel energy up to FP64 summation order."""
    stem = write_case("slab6MV", tmp_path, ncase=2000)
    prob = oracle.load_problem(stem)
    try:
        a, _, _ = oracle.run_dosxyz_philox(prob, seed=(7, 9), nthreads=1)
        b, _, _ = oracle.run_dosxyz_philox(prob, seed=(7, 9), nthreads=4)
        c, _, _ = oracle.run_dosxyz_philox(prob, seed=(7, 10), nthreads=1)
    finally:
        oracle.free_problem(prob)
    assert np.allclose(a, b, rtol=1e-12, atol=1e-12 * a.max())
    assert not np.allclose(a, c)


def _batches(tmp_path, case, ncase, nbatch):
    stem = write_case(case, tmp_path, ncase=ncase)
    inp = Path(str(stem) + ".inp")
    inp.write_text(inp.read_text().replace("nbatch = 10", f"nbatch = {nbatch}"))
    return stem


def test_oracle_philox_statistically_equals_ranmar(oracle, tmp_path):
    """Changing only the random-number source keeps the dose distribution:
    voxels above 10% Dmax agree within 2 combined batch sigma at the rate
    expected for identical distributions (50 batches, so the sigma estimate
    is close to normal), and the integral energy of an electron beam
    (small per-history variance) agrees to 0.5%."""
    stem = _batches(tmp_path / "w", "water2MeV", 200000, 50)
    prob = oracle.load_problem(stem)
    try:
        r, r2, _ = oracle.run_dosxyz_ranmar(prob)
        p, p2, _ = oracle.run_dosxyz_philox(prob, seed=(3, 3), nthreads=0)
        nb = prob.nbatch
    finally:
        oracle.free_problem(prob)

    def mv(s, s2):
        m = s / nb
        return m, np.sqrt(np.maximum(s2 / nb - m * m, 0) / (nb - 1))
    mr, er = mv(r, r2)
    mp, ep = mv(p, p2)
    sel = np.where(mr[1:] > 0.1 * mr[1:].max())[0] + 1
    z = (mr[sel] - mp[sel]) / np.sqrt(er[sel] ** 2 + ep[sel] ** 2)
    assert (np.abs(z) < 2).mean() >= 0.90
    assert abs(np.mean(z)) < 1.0

    stem = _batches(tmp_path / "e", "electron6MeV", 20000, 50)
    prob = oracle.load_problem(stem)
    try:
        r, _, _ = oracle.run_dosxyz_ranmar(prob)
        p, _, _ = oracle.run_dosxyz_philox(prob, seed=(3, 3), nthreads=0)
    finally:
        oracle.free_problem(prob)
    assert abs(r[1:].sum() / p[1:].sum() - 1) < 0.005
