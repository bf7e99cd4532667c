"""The large configurations, run whole on the GPU.

C5, the multi-beam plan: the 256x256x160 phantom at 0.2 cm (the patient
geometry scaled, air / lung / water / bone) with five coplanar 8x8 cm^2 6 MV
beams at gantry 0/72/144/216/288 through the matRad beam geometry
(omc_matrad.c:1084-1254).  FP32 transport against the FP64 kernel under the
equal-accuracy gate, energy conservation history by history, and on the
same geometry at a coarser grid the FP64 engine replays the oracle's
beamlet histories.

C4, the serving stream: three phantoms resident at once (water 40^3, the
slab, the patient) with one table set; requests on all three arrive while
others run, and each gives the same result alone and in the mix."""
import numpy as np
import pytest

import dosestats as DS
from ompmc_amd import _native as N, configs as cfg

pytestmark = pytest.mark.gpu

RM = 0.5109989461


@pytest.fixture(scope="module")
def D():
    from ompmc_amd import device
    return device


def beam_problem(oracle, d, phantom, mono=None):
    d.mkdir(parents=True, exist_ok=True)
    pf = phantom.write_egsphant(d / "phantom.egsphant")
    kw = dict(mono=mono) if mono else dict(spectrum_file=cfg.write_spectrum(d / "6mv.spectrum"))
    cfg.write_inp(d / "c5", cfg.dosxyz_inp_items(d, pf, ncase=1000, **kw))
    prob = oracle.load_problem(d / "c5")
    bs = cfg.c5_beams()
    src = N.beamlet_source(prob.src.spec, bs.ibeam, bs.src, bs.corner, bs.side1, bs.side2)
    return prob, src


def summed(s, s2, nb):
    """Dose of all beams together: mean and variance of the mean per voxel
    (beams are independent, so their variances add)."""
    m = s[:, 1:] / nb
    v = np.maximum(s2[:, 1:] / nb - m * m, 0.0) / (nb - 1)
    return m.sum(0), v.sum(0)


def test_c5_fp32_matches_fp64(oracle, D, tmp_path):
    """1e7 histories per beam and 100 batches in each precision,
    independent streams: >= 95% of the voxels above 10% of the summed Dmax
    within 2 sigma, integral within 0.5%."""
    prob, src = beam_problem(oracle, tmp_path, cfg.config_phantom("C5"))
    try:
        assert (prob.ph.isize, prob.ph.jsize, prob.ph.ksize) == (256, 256, 160)
        nb = 100
        nh = np.full(src.nbeamlets, 10_000_000, np.int64)
        res = {}
        for prec, seed in ((1, (71, 1)), (0, (72, 1))):
            with D.Engine(prec) as eng:
                eng.upload_media(prob.tab)
                pid = eng.upload_phantom(prob.ph, prob.reg)
                s, s2, st = eng.run_beamlets(pid, src, nh, nb, seed=seed)
            assert st.killed == 0 and st.woodcock_over == 0
            assert st.histories == int(nh.sum())
            res[prec] = summed(s, s2, nb)
            del s, s2
    finally:
        oracle.free_problem(prob)
    (m32, v32), (m64, v64) = res[1], res[0]
    # the helper takes (sum, sum2, nbatch); the summed-beam form has mean and
    # variance of the mean, with at least nb - 1 degrees of freedom
    avg = 0.5 * (m32 + m64)
    sel = avg > 0.1 * avg.max()
    z = (m32 - m64)[sel] / np.sqrt(v32 + v64)[sel]
    r = {"frac": float((np.abs(z) < 2).mean()), "n": int(sel.sum()), "dof": 2 * (nb - 1),
         "ratio": float(m32.sum() / m64.sum()), "zmean": float(z.mean())}
    DS.assert_equal_accuracy(r, "C5 FP32 vs FP64")


@pytest.mark.parametrize("precision,tol", [(0, 1e-12), (1, 1e-5)])
def test_c5_energy_conserved_per_history(oracle, D, tmp_path, precision, tol):
    """Mono 6 MeV photons in the five C5 beams: every history's energy is
    deposited in the phantom, in region 0 or leaves it (the ledger), and
    the deposits in the phantom are the tallies.  FP32 carries a particle's
    energy in FP32 and rounds it at every energy-loss step (half an ulp of
    4-8 MeV is 2.4e-7 MeV); the electrons of a 6 MeV history take a few
    hundred steps through the 2 mm voxels, so its total may drift by some
    1e-6 and at most ~1e-4: the FP32 bound is 1e-5 relative."""
    prob, src = beam_problem(oracle, tmp_path, cfg.config_phantom("C5"), mono=6.0)
    try:
        nh = np.full(src.nbeamlets, 40000, np.int64)
        with D.Engine(precision) as eng:
            eng.upload_media(prob.tab)
            pid = eng.upload_phantom(prob.ph, prob.reg)
            s, _, st, led = eng.run_beamlets(pid, src, nh, 10, seed=(5, 5), ledger=True)
    finally:
        oracle.free_problem(prob)
    assert st.killed == 0 and st.woodcock_over == 0
    led = led.reshape(-1, 4)
    assert led.shape[0] == int(nh.sum())
    dev = np.abs(led[:, :3].sum(1) - 6.0)
    assert dev.max() <= tol * 6.0, f"max deviation {dev.max():.3e} MeV, {np.count_nonzero(dev > tol * 6)} histories"
    assert abs(led[:, :2].sum() - s.sum()) < 1e-9 * s.sum()


def test_c5_geometry_replays_oracle_beams(oracle, D, tmp_path):
    """The C5 geometry on a 64x64x40 grid (0.8 cm, same extent): the FP64
    engine's dose per beam equals the oracle's beamlet loop on the same
    histories, to FP64 summation order."""
    ph = cfg.patient_phantom(64, 64, 40, 0.8, scale=(1.6, 1.6, 32.0 / 24.0))
    prob, src = beam_problem(oracle, tmp_path, ph)
    try:
        nb, nper = 10, 150
        nh = np.full(src.nbeamlets, nb * nper, np.int64)
        with D.Engine(0) as eng:
            eng.upload_media(prob.tab)
            pid = eng.upload_phantom(prob.ph, prob.reg)
            sd, sd2, st = eng.run_beamlets(pid, src, nh, nb, seed=(8, 8), hist0=0)
        assert st.killed == 0
        h0 = 0
        for b in range(src.nbeamlets):
            so, so2 = oracle.run_beamlet_philox(prob, src, b, nper, nb, seed=(8, 8), hist0=h0)
            h0 += nb * nper
            assert so[1:].sum() > 0
            assert np.abs(sd[b] - so).max() <= 1e-9 * np.abs(so).max(), f"beam {b}"
            assert np.abs(sd2[b] - so2).max() <= 1e-9 * np.abs(so2).max(), f"beam {b}"
    finally:
        oracle.free_problem(prob)


def _c4_problems(oracle, d):
    d.mkdir(parents=True, exist_ok=True)
    spec = cfg.write_spectrum(d / "6mv.spectrum")
    probs = []
    for k, ph in enumerate(cfg.c4_phantoms()):
        pf = ph.write_egsphant(d / f"p{k}.egsphant")
        cfg.write_inp(d / f"p{k}", cfg.dosxyz_inp_items(d, pf, ncase=1000, spectrum_file=spec))
        probs.append(oracle.load_problem(d / f"p{k}"))
    return probs


def _field(prob, side, energy):
    import ctypes as C
    src = N.SourceDosxyz()
    h = side / 2
    N.check(N.host().omc_source_dosxyz_setup(C.byref(src), C.byref(prob.ph), -h, h, -h, h, 100.0, 0))
    if energy:
        N.host().omc_spectrum_mono(C.byref(src.spec), energy)
    else:
        src.spec = prob.src.spec
    return src


def _drain(eng, tickets, nreg):
    out = {}
    for _ in range(10_000_000):
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
def test_c4_three_resident_phantoms_alone_and_in_mix(oracle, D, tmp_path, precision):
    """Requests on each of the three resident phantoms (fields of 2 to 18 cm,
    mono and 6 MV) give the same batch sums alone and when they arrive while
    a large request on another phantom and each other are on the device."""
    probs = _c4_problems(oracle, tmp_path)
    try:
        reqs = [(0, 18.0, 0.0, 30000), (1, 6.0, 3.5, 20000), (2, 12.0, 0.0, 25000), (2, 2.0, 1.25, 15000)]
        with D.Engine(precision) as eng:
            eng.upload_media(probs[2].tab)
            pids = [eng.upload_phantom(p.ph, p.reg) for p in probs]
            nreg = [p.reg.nreg for p in probs]
            srcs = [_field(probs[ph], side, e) for ph, side, e, _ in reqs]
            alone = []
            for k, (ph, _, _, n) in enumerate(reqs):
                t = eng.submit(pids[ph], srcs[k], n, nbatch=10, seed=(40 + k, 1))
                alone.append(_drain(eng, [t], {t: nreg[ph]})[t])
            big_src = _field(probs[1], 10.0, 0.0)
            tb = eng.submit(pids[1], big_src, 8_000_000, nbatch=10, seed=(9, 9))
            for _ in range(3):
                eng.step()
            ts = [eng.submit(pids[ph], srcs[k], n, nbatch=10, seed=(40 + k, 1)) for k, (ph, _, _, n) in
                  enumerate(reqs)]
            mix = _drain(eng, [tb] + ts, {tb: nreg[1], **{t: nreg[reqs[k][0]] for k, t in enumerate(ts)}})
    finally:
        for p in probs:
            oracle.free_problem(p)
    for k, t in enumerate(ts):
        ref, got = alone[k], mix[t]
        assert got[0].status == 0 and ref[0].status == 0
        scale = np.abs(ref[1]).max()
        assert scale > 0
        assert np.allclose(got[1], ref[1], rtol=1e-12, atol=1e-12 * scale), f"request {k}"
        assert np.allclose(got[2], ref[2], rtol=1e-12, atol=1e-12 * np.abs(ref[2]).max()), f"request {k}"
        assert got[0].admit_ms < mix[tb][0].finish_ms, f"request {k} waited for the large one"
