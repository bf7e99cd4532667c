"""Device engine (libomc_hip.so, through its C ABI) against the oracle.

Replay parity: the device and the oracle in PHILOX mode draw every particle's
random numbers from the same per-particle Philox stream, so with the same
seed they follow the same histories; FP64 results must agree to rounding.
The oracle's RANMAR mode is bit-identical to the reference, and the Philox
mode is statistically identical to it (test_oracle_reference.py), which
closes the chain device -> oracle -> reference.
"""
import ctypes as C

import numpy as np
import pytest

from cases import write_case

pytestmark = pytest.mark.gpu

RM = 0.5109989461


@pytest.fixture(scope="module")
def D():
    from ompmc_amd import device
    return device


def run_both(oracle, D, stem, seed=(21, 5), precision=0, chunk=0):
    prob = oracle.load_problem(stem)
    try:
        so, so2, cnt = oracle.run_dosxyz_philox(prob, seed=seed, nthreads=1)
        L = oracle.lib()
        ctx = L.ora_create(C.byref(prob.tab), C.byref(prob.reg), C.byref(prob.ph), 1, oracle.PHILOX)
        L.ora_seed_philox(ctx, *seed)
        led_o = np.zeros(3 * prob.nhist)
        L.ora_set_ledger(ctx, led_o.ctypes.data_as(C.POINTER(C.c_double)), prob.nhist)
        t1, t2 = np.zeros(prob.reg.nreg), np.zeros(prob.reg.nreg)
        P = C.POINTER(C.c_double)
        assert L.ora_run_dosxyz(ctx, C.byref(prob.src), prob.nperbatch, prob.nbatch, t1.ctypes.data_as(P),
                                t2.ctypes.data_as(P)) == 0
        L.ora_destroy(ctx)
        with D.Engine(precision, chunk=chunk) as eng:
            eng.upload_media(prob.tab)
            pid = eng.upload_phantom(prob.ph, prob.reg)
            sd, sd2, st, led_d = eng.run_dosxyz(pid, prob.src, prob.nperbatch, prob.nbatch, seed=seed,
                                                nreg=prob.reg.nreg, ledger=True)
        return dict(so=so, so2=so2, sd=sd, sd2=sd2, st=st, led_o=led_o.reshape(-1, 3),
                    led_d=led_d.reshape(-1, 3), nb=prob.nbatch, nhist=prob.nhist)
    finally:
        oracle.free_problem(prob)


REPLAY = ["water2MeV", "slab6MV", "patient6MV", "electron6MeV", "positron6MeV", "pencil", "oversize"]


@pytest.mark.parametrize("case", REPLAY)
def test_fp64_replays_oracle_histories(oracle, D, tmp_path, case):
    r = run_both(oracle, D, write_case(case, tmp_path, ncase=1000))
    so, sd = r["so"], r["sd"]
    scale = np.abs(so).max()
    close = np.abs(sd - so) <= 1e-9 * scale
    assert close.mean() >= 0.999, f"{case}: {np.count_nonzero(~close)} voxels differ"
    same = np.all(np.abs(r["led_d"] - r["led_o"]) <= 1e-9 * (1 + np.abs(r["led_o"])), axis=1)
    assert same.mean() >= 0.995, f"{case}: {np.count_nonzero(~same)} histories differ"
    assert abs(sd.sum() / so.sum() - 1) < 1e-6
    assert r["st"].killed == 0


@pytest.mark.parametrize("case,e_in", [("water2MeV", 2.0), ("electron6MeV", 6.0), ("positron6MeV", 6.0 + 2 * RM)])
@pytest.mark.parametrize("precision,tol", [(0, 1e-12), (1, 2e-6)])
def test_energy_conservation_per_history(oracle, D, tmp_path, case, e_in, precision, tol):
    stem = write_case(case, tmp_path, ncase=2000)
    prob = oracle.load_problem(stem)
    try:
        with D.Engine(precision) as eng:
            eng.upload_media(prob.tab)
            pid = eng.upload_phantom(prob.ph, prob.reg)
            s, _, st, led = eng.run_dosxyz(pid, prob.src, prob.nperbatch, prob.nbatch, nreg=prob.reg.nreg,
                                           ledger=True)
    finally:
        oracle.free_problem(prob)
    tot = led.reshape(-1, 3).sum(1)
    assert np.allclose(tot, e_in, rtol=tol, atol=tol)
    assert abs(led.reshape(-1, 3)[:, :2].sum() - s.sum()) < 1e-9 * s.sum()


def _stat(m1, e1, m2, e2):
    sel = np.where(m1[1:] > 0.1 * m1[1:].max())[0] + 1
    z = (m1[sel] - m2[sel]) / np.sqrt(e1[sel] ** 2 + e2[sel] ** 2 + 1e-300)
    return (np.abs(z) < 2).mean(), len(sel)


def _mv(s, s2, nb):
    m = s / nb
    return m, np.sqrt(np.maximum(s2 / nb - m * m, 0) / (nb - 1))


@pytest.mark.parametrize("case", ["slab6MV", "patient6MV", "electron6MeV"])
def test_fp32_agrees_with_fp64_within_2_sigma(oracle, D, tmp_path, case):
    """FP32 transport with FP64 tallies against the FP64 kernel: >= 95% of
    voxels above 10% Dmax within 2 combined sigma (50 batches), integral
    energy within 0.5%.  Both precisions draw from the same per-particle
    streams (same seed) and step photons voxel by voxel (the FP32 engine's
    delta tracking is tested on its own); a history diverges only once
    rounding changes a branch, so the integral comparison measures
    precision, not the 0.5% relative standard deviation of one
    200k-history photon run."""
    stem = write_case(case, tmp_path, ncase=200000)
    inp = tmp_path / f"{case}.inp"
    inp.write_text(inp.read_text().replace("nbatch = 10", "nbatch = 50"))
    prob = oracle.load_problem(stem)
    try:
        res = {}
        for prec in (0, 1):
            with D.Engine(prec, woodcock=-1) as eng:
                eng.upload_media(prob.tab)
                pid = eng.upload_phantom(prob.ph, prob.reg)
                res[prec] = eng.run_dosxyz(pid, prob.src, prob.nperbatch, prob.nbatch, seed=(8, 1),
                                           nreg=prob.reg.nreg)
        nb = prob.nbatch
    finally:
        oracle.free_problem(prob)
    m64, e64 = _mv(res[0][0], res[0][1], nb)
    m32, e32 = _mv(res[1][0], res[1][1], nb)
    frac, n = _stat(m64, e64, m32, e32)
    assert n > 50 and frac >= 0.95, f"{case}: {frac:.3f} of {n} voxels within 2 sigma"
    assert abs(res[1][0][1:].sum() / res[0][0][1:].sum() - 1) < 0.005


def test_device_matches_reference_rng_statistically(oracle, D, tmp_path):
    """Device FP64 (Philox streams) against the oracle in RANMAR mode, which
    is bit-identical to the reference: same dose distribution."""
    stem = write_case("slab6MV", tmp_path, ncase=200000)
    inp = tmp_path / "slab6MV.inp"
    inp.write_text(inp.read_text().replace("nbatch = 10", "nbatch = 50"))
    prob = oracle.load_problem(stem)
    try:
        r, r2, _ = oracle.run_dosxyz_ranmar(prob)
        with D.Engine(0) as eng:
            eng.upload_media(prob.tab)
            pid = eng.upload_phantom(prob.ph, prob.reg)
            d, d2, _, _ = eng.run_dosxyz(pid, prob.src, prob.nperbatch, prob.nbatch, nreg=prob.reg.nreg)
        nb = prob.nbatch
    finally:
        oracle.free_problem(prob)
    mr, er = _mv(r, r2, nb)
    md, ed = _mv(d, d2, nb)
    frac, n = _stat(mr, er, md, ed)
    assert frac >= 0.93, f"{frac:.3f} of {n}"
    assert abs(d[1:].sum() / r[1:].sum() - 1) < 0.01


@pytest.mark.parametrize("precision", [0, 1])
def test_photon_splitting_matches_reference_statistically(oracle, D, tmp_path, precision):
    """nsplit = 5 (6 MeV photons, water): device against the oracle in
    RANMAR mode, which gives the reference's .3ddose byte for byte for this
    case (ref_dose_split5.json).  Splitting has no Philox-mode counterpart in
    the oracle, so the check is statistical: the dose per history in >= 93%
    of the voxels above 10% of Dmax within 2 combined sigma, and the
    deposited energy within 1%."""
    stem = write_case("split5", tmp_path, ncase=60000)
    inp = tmp_path / "split5.inp"
    inp.write_text(inp.read_text().replace("nbatch = 10", "nbatch = 50"))
    prob = oracle.load_problem(stem)
    try:
        assert prob.nsplit == 5
        r, r2, _ = oracle.run_dosxyz_ranmar(prob)
        with D.Engine(precision) as eng:
            eng.upload_media(prob.tab)
            pid = eng.upload_phantom(prob.ph, prob.reg)
            L = D.lib()
            st = D.DevStats()
            d = np.zeros(prob.reg.nreg)
            d2 = np.zeros(prob.reg.nreg)
            P = C.POINTER(C.c_double)
            rp = D.RunParams(pid, prob.nsplit, (C.c_uint32 * 2)(11, 12), 0, prob.nperbatch, prob.nbatch, None,
                             C.pointer(st))
            assert L.omc_dev_run_dosxyz(eng.ctx, C.byref(prob.src), C.byref(rp), d.ctypes.data_as(P),
                                        d2.ctypes.data_as(P)) == 0
        nb = prob.nbatch
    finally:
        oracle.free_problem(prob)
    assert st.killed == 0
    mr, er = _mv(r, r2, nb)
    md, ed = _mv(d, d2, nb)
    frac, n = _stat(mr, er, md, ed)
    assert n > 50
    assert frac >= 0.93, f"{frac:.3f} of {n}"
    assert abs(d[1:].sum() / r[1:].sum() - 1) < 0.01


@pytest.mark.parametrize("case,precision", [("patient6MV", 0), ("electron6MeV", 0), ("patient6MV", 1)])
def test_schedule_and_reruns_do_not_change_histories(oracle, D, tmp_path, case, precision):
    """Histories are identified by (seed, index) only: a different number of
    pipelines, a small refill chunk (many injections, primaries mixed with
    older secondaries), particles handed to the next launch after every
    one or three loop trips (their lane state travels with them), handed
    on after every trip once a launch's input is used up, not drained at
    all, or a repeat reproduce the same energies up to FP64 summation
    order."""
    stem = write_case(case, tmp_path, ncase=3000)
    prob = oracle.load_problem(stem)
    try:
        out = []
        for chunk, pipes, ytrips, drain in ((0, 0, 0, -1), (0, 1, 0, 0), (777, 3, 0, 0), (0, 2, 1, 0),
                                            (5000, 0, 3, 0), (2000, 2, 0, 64), (0, 0, 0, None)):
            # drain_min -1: these small launches drain too (by default, the
            # last configuration, a launch smaller than the device's
            # resident lanes does not)
            dm = 0 if drain is None or drain < 0 else -1
            with D.Engine(precision, chunk=chunk, pipelines=pipes, yield_trips=ytrips, drain_lanes=drain or 0,
                          drain_trips=1 if drain == 64 else 0, drain_min=dm) as eng:
                eng.upload_media(prob.tab)
                pid = eng.upload_phantom(prob.ph, prob.reg)
                s, _, st, led = eng.run_dosxyz(pid, prob.src, prob.nperbatch, prob.nbatch, seed=(4, 4),
                                               nreg=prob.reg.nreg, ledger=True)
                out.append((s, led, st.injections, st.yields))
    finally:
        oracle.free_problem(prob)
    assert out[2][2] >= 4                       # 3000 histories, chunk 777
    assert out[3][3] > 1000 and out[4][3] > 100 and out[5][3] > 100
    assert out[6][3] <= 5, "a job this small must not drain by default"
    for s, led, _, _ in out[1:]:
        assert np.allclose(s, out[0][0], rtol=1e-12, atol=1e-12 * out[0][0].max())
        assert np.allclose(led, out[0][1], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("energy,charge", [(3.0, 0), (6.0, -1), (4.0, 1)])
def test_vacuum_voxels_inside_phantom(oracle, D, tmp_path, energy, charge):
    """Medium 0 voxels (vacuum) inside the grid: straight-line transport,
    same histories as the oracle, including charged particles that cross
    the gap and re-enter the medium they left (where the reference goes on
    to the fictitious-sigma test without restarting the step)."""
    from ompmc_amd import configs as cfg
    ph = cfg.water_cube(10, 5.0)
    ph.med[4:6, 3:7, 3:7] = 0
    ph.rho[4:6, 3:7, 3:7] = 0.0
    pf = ph.write_egsphant(tmp_path / "vac.egsphant")
    items = cfg.dosxyz_inp_items(tmp_path, pf, ncase=1000, mono=energy, charge=charge, field=(-1, 1, -1, 1))
    cfg.write_inp(tmp_path / "vac", items)
    r = run_both(oracle, D, tmp_path / "vac")
    scale = np.abs(r["so"]).max()
    assert (np.abs(r["sd"] - r["so"]) <= 1e-9 * scale).mean() >= 0.999
    same = np.all(np.abs(r["led_d"] - r["led_o"]) <= 1e-9 * (1 + np.abs(r["led_o"])), axis=1)
    assert same.mean() >= 0.995, f"{np.count_nonzero(~same)} histories differ"


@pytest.mark.parametrize("charge,edge", [(0, 0.01), (-1, 0.7 - RM)])
@pytest.mark.parametrize("ulps", [1, 2, 3, 5, 8])
def test_fp32_energies_on_table_lower_edge(oracle, D, tmp_path, charge, edge, ulps):
    """Particles a few FP32 ulps above a table's lower edge: a photon just
    above pcut = ap = 10 keV, an electron just above ecut - m.  There the
    table index l*c1 + c0 nearly cancels (c0 ~ 1070 for the photon tables),
    and FP32 rounding put it one bin below the table, reading another
    medium's coefficients; at 10 keV that made a photon Rayleigh-scatter
    until the trip watchdog killed it.  Every history must now end by itself
    and conserve its energy."""
    from ompmc_amd import configs as cfg
    e = np.float32(edge)
    for _ in range(ulps):
        e = np.nextafter(e, np.float32(1))
    ph = cfg.water_cube(10, 5.0)
    pf = ph.write_egsphant(tmp_path / "edge.egsphant")
    items = cfg.dosxyz_inp_items(tmp_path, pf, ncase=2000, mono=f"{float(e):.9g}", charge=charge,
                                 field=(-1, 1, -1, 1))
    cfg.write_inp(tmp_path / "edge", items)
    prob = oracle.load_problem(tmp_path / "edge")
    try:
        with D.Engine(1) as eng:
            eng.upload_media(prob.tab)
            pid = eng.upload_phantom(prob.ph, prob.reg)
            s, _, st, led = eng.run_dosxyz(pid, prob.src, prob.nperbatch, prob.nbatch, nreg=prob.reg.nreg,
                                           ledger=True)
    finally:
        oracle.free_problem(prob)
    assert st.killed == 0
    tot = led.reshape(-1, 3).sum(1)
    # FP32 carries an electron's total energy T + m (m = 0.511 MeV), so each
    # electron set in motion rounds by up to half an ulp of m (3e-8 MeV):
    # an absolute bound, 3e-6 of a 10 keV history.  A history cut short
    # loses most of its energy.
    assert np.allclose(tot, float(e), rtol=0, atol=8 * 2.0**-24)


def test_error_paths_return_codes(oracle, D, tmp_path):
    stem = write_case("water2MeV", tmp_path, ncase=100)
    prob = oracle.load_problem(stem)
    try:
        with D.Engine(0) as eng:
            with pytest.raises(RuntimeError, match="media"):
                eng.upload_phantom(prob.ph, prob.reg)
            eng.upload_media(prob.tab)
            pid = eng.upload_phantom(prob.ph, prob.reg)
            L = D.lib()
            st = D.DevStats()
            s = np.zeros(prob.reg.nreg)
            P = C.POINTER(C.c_double)
            rp = D.RunParams(pid, 0, (C.c_uint32 * 2)(1, 2), 0, 10, 2, None, C.pointer(st))
            assert L.omc_dev_run_dosxyz(eng.ctx, C.byref(prob.src), C.byref(rp), s.ctypes.data_as(P),
                                        s.ctypes.data_as(P)) == -4           # nsplit < 1
            rp = D.RunParams(pid + 7, 1, (C.c_uint32 * 2)(1, 2), 0, 10, 2, None, C.pointer(st))
            assert L.omc_dev_run_dosxyz(eng.ctx, C.byref(prob.src), C.byref(rp), s.ctypes.data_as(P),
                                        s.ctypes.data_as(P)) == -4           # unknown phantom
            rp = D.RunParams(pid, 1, (C.c_uint32 * 2)(1, 2), 0, 0, 2, None, C.pointer(st))
            assert L.omc_dev_run_dosxyz(eng.ctx, C.byref(prob.src), C.byref(rp), s.ctypes.data_as(P),
                                        s.ctypes.data_as(P)) == -4           # empty batch
        with pytest.raises(RuntimeError, match="budget"):
            D.Engine(0, hbm_budget=1 << 20)
    finally:
        oracle.free_problem(prob)


def test_single_history_and_single_batch(oracle, D, tmp_path):
    stem = write_case("electron6MeV", tmp_path, ncase=1)
    prob = oracle.load_problem(stem)
    try:
        so, _, _ = oracle.run_dosxyz_philox(prob, seed=(1, 1), nperbatch=1, nbatch=1, nthreads=1)
        with D.Engine(0) as eng:
            eng.upload_media(prob.tab)
            pid = eng.upload_phantom(prob.ph, prob.reg)
            sd, sd2, st, _ = eng.run_dosxyz(pid, prob.src, 1, 1, seed=(1, 1), nreg=prob.reg.nreg)
    finally:
        oracle.free_problem(prob)
    assert st.histories == 1
    assert np.allclose(sd, so, rtol=1e-12, atol=1e-12) and np.allclose(sd2, sd * sd, rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("case", ["slab6MV", "electron6MeV"])
def test_device_counters_equal_oracle_counts(oracle, D, tmp_path, case):
    """FP64 replay: the device's per-kernel counters of ausgab calls and voxel
    crossings (the units of the algorithmic-bytes roofline) equal the
    oracle's counts for the same histories."""
    stem = write_case(case, tmp_path, ncase=2000)
    prob = oracle.load_problem(stem)
    try:
        _, _, cnt = oracle.run_dosxyz_philox(prob, seed=(3, 4), nthreads=1)
        with D.Engine(0) as eng:
            eng.upload_media(prob.tab)
            pid = eng.upload_phantom(prob.ph, prob.reg)
            _, _, st, _ = eng.run_dosxyz(pid, prob.src, prob.nperbatch, prob.nbatch, seed=(3, 4),
                                         nreg=prob.reg.nreg)
    finally:
        oracle.free_problem(prob)
    assert st.deposits == st.photon_deposits + st.electron_deposits
    for dev, ora in ((st.deposits, cnt.n_dep), (st.photon_crossings, cnt.n_cross_photon),
                     (st.electron_crossings, cnt.n_cross_charged)):
        assert abs(dev - ora) <= 1e-3 * ora + 2, (dev, ora)


def test_media_upload_must_cover_resident_phantoms(oracle, D, tmp_path):
    """New tables with fewer media than a resident phantom uses would let
    the kernels index past them: the upload is refused until the phantom
    is released."""
    pp = oracle.load_problem(write_case("patient6MV", tmp_path / "p", ncase=100))
    wp = oracle.load_problem(write_case("water2MeV", tmp_path / "w", ncase=100))
    try:
        assert wp.tab.nmed < pp.tab.nmed
        with D.Engine(1) as eng:
            eng.upload_media(pp.tab)
            pid = eng.upload_phantom(pp.ph, pp.reg)
            with pytest.raises(RuntimeError, match="resident phantom"):
                eng.upload_media(wp.tab)
            s, _, st, _ = eng.run_dosxyz(pid, pp.src, pp.nperbatch, pp.nbatch, nreg=pp.reg.nreg)
            assert st.killed == 0 and s[1:].sum() > 0
            eng.release_phantom(pid)
            eng.upload_media(wp.tab)
            wid = eng.upload_phantom(wp.ph, wp.reg)
            s, _, st, _ = eng.run_dosxyz(wid, wp.src, wp.nperbatch, wp.nbatch, nreg=wp.reg.nreg)
            assert st.killed == 0 and s[1:].sum() > 0
    finally:
        oracle.free_problem(pp)
        oracle.free_problem(wp)
