"""Jobs whose batch tally ([nbatch][nreg] FP64) exceeds the context's
tally_bytes are admitted batch by batch: each batch is a job with one grid,
folded into the job's sum and sum-of-squares grids when its last particle is
gone.  Each batch's histories keep the streams they have in the whole job
(the stream index starts at the batch's first history), so every history is
the same: per-history ledger rows agree to FP64 summation order, and so do
the sums (batches are summed in the order they end).  A deposit-order or
batch-offset mistake moves whole histories and fails the ledger check."""
import numpy as np
import pytest

from ompmc_amd import _native as N, configs as cfg

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def D():
    from ompmc_amd import device
    return device


@pytest.fixture
def problem(oracle, tmp_path):
    from cases import write_case
    pp = oracle.load_problem(write_case("patient6MV", tmp_path / "p"))
    yield pp
    oracle.free_problem(pp)


def small_cap(nreg):
    """A tally limit of four grids: ten batches go two at a time, or as many
    as keep half the refill room busy (two pipelines of 4096 here: 4096
    histories, two batches of 1500 or more)."""
    return 4 * 8 * nreg


ENGINE = dict(pipelines=2, chunk=4096)


def close(a, b, rel=1e-12):
    return np.abs(a - b).max() <= rel * max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("precision", [0, 1])
def test_dosxyz_batch_by_batch_equals_whole_tally(D, problem, precision):
    pp = problem
    nreg = pp.reg.nreg
    out = {}
    for name, cap in (("whole", -1), ("split", small_cap(nreg))):
        with D.Engine(precision, tally_bytes=cap, **ENGINE) as eng:
            eng.upload_media(pp.tab)
            pid = eng.upload_phantom(pp.ph, pp.reg)
            out[name] = eng.run_dosxyz(pid, pp.src, 1500, 10, seed=(11, 4), nreg=nreg, ledger=True)
    (s0, q0, st0, l0), (s1, q1, st1, l1) = out["whole"], out["split"]
    assert st1.histories == st0.histories == 15000
    assert (st1.deposits, st1.electron_steps, st1.photon_steps) == (st0.deposits, st0.electron_steps,
                                                                    st0.photon_steps)
    l0, l1 = l0.reshape(-1, 4), l1.reshape(-1, 4)
    assert np.all(np.abs(l1 - l0) <= 1e-9 * (1 + np.abs(l0)))
    assert close(s1, s0) and close(q1, q0)
    assert s0[1:].sum() > 0


def test_beamlets_batch_by_batch_equals_whole_tally(D, problem):
    pp = problem
    nreg = pp.reg.nreg
    bs = cfg.beamlets(gantry_deg=(0.0,), field=2.0, bixel=1.0)
    src = N.beamlet_source(pp.src.spec, bs.ibeam, bs.src, bs.corner, bs.side1, bs.side2)
    nh = np.array([40000, 16000, 24000, 32000], np.int64)
    out = {}
    for name, cap in (("whole", -1), ("split", small_cap(nreg))):
        with D.Engine(1, tally_bytes=cap, **ENGINE) as eng:
            eng.upload_media(pp.tab)
            pid = eng.upload_phantom(pp.ph, pp.reg)
            out[name] = eng.run_beamlets(pid, src, nh, 8, seed=(5, 9))
    (s0, q0, st0), (s1, q1, st1) = out["whole"], out["split"]
    assert st1.histories == st0.histories == int(nh.sum())
    assert st1.deposits == st0.deposits
    for b in range(len(nh)):
        assert close(s1[b], s0[b]) and close(q1[b], q0[b]), b


def test_request_batch_by_batch_equals_whole_tally(D, problem):
    pp = problem
    nreg = pp.reg.nreg
    res = {}
    for name, cap in (("whole", -1), ("split", small_cap(nreg))):
        with D.Engine(0, tally_bytes=cap, **ENGINE) as eng:
            eng.upload_media(pp.tab)
            pid = eng.upload_phantom(pp.ph, pp.reg)
            t = eng.submit(pid, pp.src, 12000, nbatch=10, seed=(7, 1))
            for _ in range(200000):
                r, s, s2 = eng.poll(t, nreg)
                if r.done:
                    break
                eng.step()
            assert r.done and r.status == 0 and r.nhist == 12000
            res[name] = (s, s2)
    assert close(res["split"][0], res["whole"][0]) and close(res["split"][1], res["whole"][1])
