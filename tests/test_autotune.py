"""omc_dev_tune through the C ABI.  The autotuner times the dosxyz loop under
candidate engine settings and leaves the fastest in the context.  Histories
do not depend on the schedule, so a run with the same seed after tuning must
give the sums it gave before (to FP64 summation order)."""
import numpy as np
import pytest

from cases import write_case

pytestmark = pytest.mark.gpu

CANDIDATES = {
    "pipelines": {2, 4, 6, 8, 10, 12},
    "yield_trips": {1024, 4096, 16384},
    "drain_lanes": {16, 32, 48},
    "drain_trips": {8, 16, 32},
    "event_div": {1, 2, 4},
    "fetch_min": {1, 4, 8, 16},
}


@pytest.mark.parametrize("precision", [0, 1])
def test_tune_picks_a_candidate_and_keeps_histories(oracle, tmp_path, precision):
    from ompmc_amd import device as D
    prob = oracle.load_problem(write_case("slab6MV", tmp_path, ncase=20000))
    try:
        with D.Engine(precision, hbm_budget=4 << 30, pipelines=2, chunk=1 << 20) as eng:
            eng.upload_media(prob.tab)
            pid = eng.upload_phantom(prob.ph, prob.reg)
            nreg = prob.reg.nreg
            s0, q0, _, _ = eng.run_dosxyz(pid, prob.src, prob.nperbatch, prob.nbatch, seed=(3, 4), nreg=nreg)
            best, hps = eng.tune(pid, prob.src, 100000, nbatch=10, seed=(5, 6))
            s1, q1, st, _ = eng.run_dosxyz(pid, prob.src, prob.nperbatch, prob.nbatch, seed=(3, 4), nreg=nreg)
    finally:
        oracle.free_problem(prob)
    assert hps > 0
    for key, values in CANDIDATES.items():
        assert best[key] in values, (key, best[key])
    # the chunk is a candidate unless the budget lowered it
    assert 0 < best["chunk"] <= 1 << 23
    assert best["precision"] == precision
    scale = np.abs(s0).max()
    assert np.allclose(s1, s0, rtol=0, atol=1e-9 * scale)
    assert np.allclose(q1, q0, rtol=0, atol=1e-9 * np.abs(q0).max())
    assert st.killed == 0


def test_tune_rejects_bad_arguments(oracle, tmp_path):
    from ompmc_amd import device as D
    prob = oracle.load_problem(write_case("water2MeV", tmp_path, ncase=1000))
    try:
        with D.Engine(1, hbm_budget=2 << 30, pipelines=2, chunk=1 << 18) as eng:
            eng.upload_media(prob.tab)
            pid = eng.upload_phantom(prob.ph, prob.reg)
            with pytest.raises(RuntimeError):
                eng.tune(pid, prob.src, 5, nbatch=10)          # fewer histories than batches
            with pytest.raises(RuntimeError):
                eng.tune(pid + 7, prob.src, 10000, nbatch=10)  # no such phantom
    finally:
        oracle.free_problem(prob)
