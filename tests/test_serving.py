"""Serving: continuous batching of dose requests on resident phantoms.

Requests share the device's transport generations; every history keeps its
own Philox stream, so a request's result must not depend on what else is in
flight: alone and in a mix it gives the same batch sums up to FP64
summation order.  Admission follows the HBM budget."""
import ctypes as C

import numpy as np
import pytest

from ompmc_amd import _native as N, configs as cfg

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def D():
    from ompmc_amd import device
    return device


@pytest.fixture
def problems(oracle, tmp_path):
    """Patient phantom with a 6 MV field and four beamlets."""
    from cases import write_case
    pp = oracle.load_problem(write_case("patient6MV", tmp_path / "p"))
    bs = cfg.beamlets(gantry_deg=(0.0,), field=2.0, bixel=1.0)
    src = N.beamlet_source(pp.src.spec, bs.ibeam, bs.src, bs.corner, bs.side1, bs.side2)
    yield pp, src
    oracle.free_problem(pp)


def drain(eng, tickets, nreg, max_steps=200000):
    out = {}
    for _ in range(max_steps):
        for t in list(tickets):
            res, s, s2 = eng.poll(t, nreg[t])
            if res.done:
                out[t] = (res, s, s2)
                tickets.remove(t)
        if not tickets:
            return out
        eng.step()
    raise AssertionError("requests did not finish")


@pytest.mark.parametrize("precision", [0, 1])
def test_request_result_same_alone_and_in_mix(D, problems, precision):
    pp, src = problems
    with D.Engine(precision) as eng:
        eng.upload_media(pp.tab)
        pid = eng.upload_phantom(pp.ph, pp.reg)
        nreg = pp.reg.nreg
        # alone
        ta = eng.submit(pid, pp.src, 20000, nbatch=10, seed=(1, 2))
        alone = drain(eng, [ta], {ta: nreg})[ta]
        assert alone[0].status == 0 and alone[0].nhist == 20000
        tc = eng.submit(pid, src, 8000, nbatch=4, seed=(3, 3), kind=1, beamlet=2)
        alone_c = drain(eng, [tc], {tc: nreg})[tc]
        # in a mix: B first (large: still being injected when the others
        # arrive), A and C arrive while B is on the device
        tb = eng.submit(pid, pp.src, 12000000, nbatch=10, seed=(7, 7))
        for _ in range(3):
            eng.step()
        ta2 = eng.submit(pid, pp.src, 20000, nbatch=10, seed=(1, 2))
        tc2 = eng.submit(pid, src, 8000, nbatch=4, seed=(3, 3), kind=1, beamlet=2)
        mix = drain(eng, [tb, ta2, tc2], {tb: nreg, ta2: nreg, tc2: nreg})
    for (ref, got) in ((alone, mix[ta2]), (alone_c, mix[tc2])):
        assert got[0].status == 0
        scale = np.abs(ref[1]).max()
        assert scale > 0
        assert np.allclose(got[1], ref[1], rtol=1e-12, atol=1e-12 * scale)
        assert np.allclose(got[2], ref[2], rtol=1e-12, atol=1e-12 * np.abs(ref[2]).max())
    rb, ra = mix[tb][0], mix[ta2][0]
    assert rb.submit_ms <= rb.admit_ms <= rb.finish_ms
    assert ra.admit_ms < rb.finish_ms, "a request arriving mid-run must be admitted while others run"
    assert ra.finish_ms < rb.finish_ms, "the small request must not wait for the large one"


def test_admission_follows_hbm_budget(D, problems):
    pp, _ = problems
    nreg = pp.reg.nreg
    tally = 8 * nreg                            # bytes per batch
    # ~80 MB of batch tallies per request: one fits beside the queues (half
    # of the 256 MiB budget), two do not
    nb_big = int(80e6 // tally)
    with D.Engine(0, hbm_budget=256 << 20) as eng:
        eng.upload_media(pp.tab)
        pid = eng.upload_phantom(pp.ph, pp.reg)
        t1 = eng.submit(pid, pp.src, 4 * nb_big, nbatch=nb_big, seed=(1, 1))
        t2 = eng.submit(pid, pp.src, 4 * nb_big, nbatch=nb_big, seed=(1, 2))
        t3 = eng.submit(pid, pp.src, 1000, nbatch=4 * nb_big, seed=(1, 3))   # can never fit
        out = drain(eng, [t1, t2, t3], {t1: nreg, t2: nreg, t3: nreg})
    r1, r2, r3 = out[t1][0], out[t2][0], out[t3][0]
    assert r1.status == 0 and r2.status == 0
    assert r2.admit_ms >= r1.finish_ms, "two requests must not be resident beyond the budget"
    assert r3.status == -9 and out[t3][1] is None


def test_media_uploads_do_not_leak_budget(D, problems):
    """Each media upload replaces the table buffers and returns their bytes
    to the budget: many uploads under a tight budget still leave room for
    a request (without that, the table bytes of every upload would add up
    and admission would refuse it)."""
    pp, _ = problems
    nreg = pp.reg.nreg
    with D.Engine(0, hbm_budget=256 << 20) as eng:
        for _ in range(80):             # about 3 MB of tables each
            eng.upload_media(pp.tab)
        pid = eng.upload_phantom(pp.ph, pp.reg)
        t = eng.submit(pid, pp.src, 4000, nbatch=10, seed=(9, 9))
        r = drain(eng, [t], {t: nreg})[t][0]
    assert r.status == 0 and r.nhist == 4000


def test_request_errors(D, problems):
    pp, src = problems
    with D.Engine(0) as eng:
        eng.upload_media(pp.tab)
        pid = eng.upload_phantom(pp.ph, pp.reg)
        with pytest.raises(RuntimeError):
            eng.submit(pid, pp.src, 0)
        with pytest.raises(RuntimeError):
            eng.submit(pid + 3, pp.src, 100)
        with pytest.raises(RuntimeError, match="beamlet"):
            eng.submit(pid, src, 100, kind=1, beamlet=99)
        with pytest.raises(RuntimeError, match="ticket"):
            eng.poll(12345)
        assert eng.step() == 0                  # idle engine


def _first_slices(ph, reg, k):
    """The phantom's first k slices along z, sharing its arrays (voxel index
    ix + iy*isize + iz*isize*jsize, so they are a prefix)."""
    sub = N.Phantom()
    C.pointer(sub)[0] = ph
    sub.ksize = k
    r = N.Regions()
    C.pointer(r)[0] = reg
    r.nreg = ph.isize * ph.jsize * k + 1
    return sub, r


def test_larger_phantom_uploaded_while_requests_finish(D, problems):
    """Uploading a phantom larger than any resident one regrows the scratch
    that finished requests are collected through; requests already finishing
    must be delivered intact first, and the new phantom must serve at once."""
    pp, _ = problems
    small, sreg = _first_slices(pp.ph, pp.reg, 3)
    nreg_s = sreg.nreg
    seeds = [(11, k) for k in range(6)]
    with D.Engine(0) as eng:                        # each request alone
        eng.upload_media(pp.tab)
        sid = eng.upload_phantom(small, sreg)
        ref = {}
        for sd in seeds:
            t = eng.submit(sid, pp.src, 4000, nbatch=4, seed=sd)
            ref[sd] = drain(eng, [t], {t: nreg_s})[t]
    with D.Engine(0) as eng:
        eng.upload_media(pp.tab)
        sid = eng.upload_phantom(small, sreg)
        tick = {eng.submit(sid, pp.src, 4000, nbatch=4, seed=sd): sd for sd in seeds}
        got = {}
        for _ in range(200000):                     # until the first request is delivered
            eng.step()
            for t in list(tick):
                res, s, s2 = eng.poll(t, nreg_s)
                if res.done:
                    got[tick.pop(t)] = (res, s, s2)
            if got:
                break
        assert got and tick, "the upload must happen with requests still on the device"
        pid = eng.upload_phantom(pp.ph, pp.reg)
        tp = eng.submit(pid, pp.src, 6000, nbatch=4, seed=(5, 5))
        rest = drain(eng, list(tick) + [tp], {**{t: nreg_s for t in tick}, tp: pp.reg.nreg})
        got.update({tick[t]: r for t, r in rest.items() if t != tp})
    assert sorted(got) == sorted(seeds)
    for sd in seeds:
        r, s, s2 = got[sd]
        assert r.status == 0
        scale = np.abs(ref[sd][1]).max()
        assert scale > 0
        assert np.allclose(s, ref[sd][1], rtol=1e-12, atol=1e-12 * scale)
        assert np.allclose(s2, ref[sd][2], rtol=1e-12, atol=1e-12 * np.abs(ref[sd][2]).max())
    rp, sp_, _ = rest[tp]
    assert rp.status == 0 and sp_[1:].sum() > 0
