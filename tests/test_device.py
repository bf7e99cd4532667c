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

import dosestats as DS
from cases import write_case

pytestmark = pytest.mark.gpu

RM = 0.5109989461


@pytest.fixture(scope="module")
def D():
    from ompmc_amd import device
    return device


def run_both(oracle, D, stem, seed=(21, 5), precision=0, chunk=0, **engine):
    prob = oracle.load_problem(stem)
    try:
        so, so2, cnt = oracle.run_dosxyz_philox(prob, seed=seed, nthreads=1)
        L = oracle.lib()
        ctx = L.ora_create(C.byref(prob.tab), C.byref(prob.reg), C.byref(prob.ph), 1, oracle.PHILOX)
        L.ora_seed_philox(ctx, *seed)
        led_o = np.zeros(4 * prob.nhist)
        L.ora_set_ledger(ctx, led_o.ctypes.data_as(C.POINTER(C.c_double)), prob.nhist)
        t1, t2 = np.zeros(prob.reg.nreg), np.zeros(prob.reg.nreg)
        P = C.POINTER(C.c_double)
        assert L.ora_run_dosxyz(ctx, C.byref(prob.src), prob.nperbatch, prob.nbatch, t1.ctypes.data_as(P),
                                t2.ctypes.data_as(P)) == 0
        L.ora_destroy(ctx)
        with D.Engine(precision, chunk=chunk, **engine) as eng:
            eng.upload_media(prob.tab)
            pid = eng.upload_phantom(prob.ph, prob.reg)
            sd, sd2, st, led_d = eng.run_dosxyz(pid, prob.src, prob.nperbatch, prob.nbatch, seed=seed,
                                                nreg=prob.reg.nreg, ledger=True)
        return dict(so=so, so2=so2, sd=sd, sd2=sd2, st=st, cnt=cnt, led_o=led_o.reshape(-1, 4),
                    led_d=led_d.reshape(-1, 4), nb=prob.nbatch, nhist=prob.nhist)
    finally:
        oracle.free_problem(prob)


# device statistic <- oracle counter, for the same histories
COUNTERS = [("histories", "n_hist"), ("deposits", "n_dep"), ("photon_crossings", "n_cross_photon"),
            ("electron_crossings", "n_cross_charged"), ("photon_steps", "n_photon_steps"),
            ("electron_steps", "n_electron_steps"), ("compton", "n_compton"), ("photo", "n_photo"),
            ("pair", "n_pair"), ("rayleigh", "n_rayleigh"), ("brems", "n_brems"), ("moller", "n_moller"),
            ("bhabha", "n_bhabha"), ("annih_flight", "n_annih"), ("annih_rest", "n_rannih"),
            ("fict_rejections", "n_fict"), ("msdist", "n_msdist"), ("sscat", "n_sscat"),
            ("howfar_charged", "n_howfar_charged")]


def replay_report(r):
    """Differences between a device run and the oracle run of the same
    histories: per-history ledger rows that differ (energy in the phantom,
    in region 0, carried out, and the path checksum sum(energy * region),
    which moves as soon as any deposit lands in another voxel), event
    counters that differ, and voxels beyond FP64 summation order (deposits
    reach a voxel in a different order on the device, so its sum may differ
    in the last bits: 1e-9 of Dmax)."""
    led_o, led_d = r["led_o"], r["led_d"]
    same = np.all(np.abs(led_d - led_o) <= 1e-9 * (1 + np.abs(led_o)), axis=1)
    bad = np.flatnonzero(~same)
    st, cnt = r["st"], r["cnt"]
    diff = {d: (getattr(st, d), getattr(cnt, o)) for d, o in COUNTERS if getattr(st, d) != getattr(cnt, o)}
    so, sd = r["so"], r["sd"]
    far = np.abs(sd - so) > 1e-9 * np.abs(so).max()
    return {"histories": int(same.size), "histories_differ": int(bad.size), "first": bad[:5].tolist(),
            "first_rows": [(led_d[b].tolist(), led_o[b].tolist()) for b in bad[:2]],
            "counters_differ": diff, "voxels_differ": int(np.count_nonzero(far)),
            "max_voxel_diff_of_dmax": float(np.abs(sd - so).max() / np.abs(so).max()), "killed": int(st.killed)}


def howfar_tie_allowance(n):
    """Charged-particle howfar calls the device and the oracle may count
    differently without any other difference.  The device's FP64 log, exp,
    pow, sin and cos differ from glibc's by one ulp on 0.3-24% of arguments
    (test_math.py), so a step length or lateral distance can differ in its
    last bit; where that flips the reference's `ustep < tperp` test
    (ompmc.c:5057) on a step that howfar then finds inside the voxel (the
    distance along the direction is at least the perpendicular tperp), one
    side calls howfar and the other does not, and the history is the same.
    Measured: 1-2 calls in 1e4-5e4 per case; allowed: 1e-4 of the calls + 2."""
    return int(1e-4 * n) + 2


def assert_same_histories(r, what):
    """Exact replay: every history's ledger row (energies and the path
    checksum), every physics event count, and every voxel (to summation
    order) equal the oracle's; the count of charged-particle howfar calls
    within the allowance howfar_tie_allowance names the cause of."""
    rep = replay_report(r)
    diff = dict(rep["counters_differ"])
    hf = diff.pop("howfar_charged", None)
    assert (rep["histories_differ"], diff, rep["voxels_differ"], rep["killed"]) == (0, {}, 0, 0), f"{what}: {rep}"
    if hf is not None:
        assert abs(hf[0] - hf[1]) <= howfar_tie_allowance(hf[1]), f"{what}: {rep}"


REPLAY = ["water2MeV", "slab6MV", "patient6MV", "electron6MeV", "positron6MeV", "pencil", "oversize", "C1",
          "pairAg8MeV", "pairAg52MeV", "C2", "C3"]
REPLAY_NCASE = {"pairAg8MeV": 3000, "pairAg52MeV": 300, "C2": 4000, "C3": 2000}


@pytest.mark.parametrize("case", REPLAY)
def test_fp64_replays_oracle_histories(oracle, D, tmp_path, case):
    """The FP64 engine follows the oracle's histories exactly: same energy
    per history, same number of each interaction, multiple-scattering step,
    single scattering, charged-particle howfar call, photon and electron
    loop iteration, voxel crossing and deposit.  A wrong pair() or
    annihilation, an off-by-one table bin or voxel index changes a count or
    a history.  C1, C2 and C3 are those configurations at their full grids
    (40^3 water / 2 MeV, the 64^3 slab and the 128x128x96 patient under
    6 MV); pairAg* put photons of 8 and 52 MeV into silver, where pair
    production, Bhabha scattering and annihilation run hundreds of times."""
    r = run_both(oracle, D, write_case(case, tmp_path, ncase=REPLAY_NCASE.get(case, 1000)))
    assert_same_histories(r, case)
    assert abs(r["sd"].sum() / r["so"].sum() - 1) < 1e-12


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
    tot = led.reshape(-1, 4)[:, :3].sum(1)
    assert np.allclose(tot, e_in, rtol=tol, atol=tol)
    assert abs(led.reshape(-1, 4)[:, :2].sum() - s.sum()) < 1e-9 * s.sum()


def _device_run(D, prob, precision, seed, nperbatch=None, nbatch=None, woodcock=0, nsplit=None):
    with D.Engine(precision, woodcock=woodcock) as eng:
        eng.upload_media(prob.tab)
        pid = eng.upload_phantom(prob.ph, prob.reg)
        s, s2, st, _ = eng.run_dosxyz(pid, prob.src, nperbatch or prob.nperbatch, nbatch or prob.nbatch,
                                      seed=seed, nreg=prob.reg.nreg, nsplit=nsplit or prob.nsplit)
    assert st.killed == 0
    return s, s2, st


@pytest.mark.parametrize("case", ["statSlab", "statPatient", "statElectron"])
def test_fp32_agrees_with_fp64_within_2_sigma(oracle, D, tmp_path, case):
    """FP32 transport with FP64 tallies against the FP64 kernel, independent
    streams: >= 95% of the voxels above 10% of Dmax within 2 combined sigma
    (1e7 histories and 100 batches each), integral energy within 0.5%.
    Photons are stepped voxel by voxel in both (the FP32 engine's delta
    tracking is tested on its own)."""
    prob = oracle.load_problem(write_case(case, tmp_path, ncase=10_000_000, nbatch=100))
    try:
        a = _device_run(D, prob, 1, (8, 1), woodcock=-1)
        b = _device_run(D, prob, 0, (9, 1), woodcock=-1)
        nb = prob.nbatch
    finally:
        oracle.free_problem(prob)
    DS.assert_equal_accuracy(DS.compare((a[0], a[1], nb), (b[0], b[1], nb)), f"{case} FP32 vs FP64")


def test_device_matches_reference_rng_statistically(oracle, D, tmp_path):
    """Device FP64 (Philox streams) against the oracle in RANMAR mode, which
    is bit-identical to the reference: 20 independent reference-generator
    runs (seeds jxx + k) pooled to 200 batches of 2e4 histories, against
    one device run of 200 batches: the equal-accuracy gate."""
    prob = oracle.load_problem(write_case("statSlab", tmp_path, nbatch=200))
    try:
        ref = oracle.run_dosxyz_ranmar_runs(prob, 20, prob.nperbatch, 10)
        s, s2, _ = _device_run(D, prob, 0, (7, 7))
        nb = prob.nbatch
    finally:
        oracle.free_problem(prob)
    DS.assert_equal_accuracy(DS.compare(ref, (s, s2, nb)), "device FP64 vs reference RANMAR")


@pytest.mark.parametrize("precision", [0, 1])
def test_photon_splitting_matches_reference_statistically(oracle, D, tmp_path, precision):
    """nsplit = 5 (6 MeV photons, water): device against the oracle in
    RANMAR mode, which gives the reference's .3ddose byte for byte for
    splitting (ref_dose_split5.json).  Splitting has no Philox-mode
    counterpart in the oracle, so the check is the equal-accuracy gate on
    2e6 histories (20 pooled reference-generator runs, 200 batches)."""
    prob = oracle.load_problem(write_case("statSplit", tmp_path, nbatch=200))
    try:
        assert prob.nsplit == 5
        ref = oracle.run_dosxyz_ranmar_runs(prob, 20, prob.nperbatch, 10)
        s, s2, _ = _device_run(D, prob, precision, (11, 12))
        nb = prob.nbatch
    finally:
        oracle.free_problem(prob)
    DS.assert_equal_accuracy(DS.compare(ref, (s, s2, nb)), f"nsplit 5, precision {precision}")


@pytest.mark.parametrize("precision", [0, 1])
def test_splitting_with_small_queues_keeps_histories(oracle, D, tmp_path, precision):
    """nsplit = 20 (the reference example's value) with queues of 4096 slots
    per pipeline: a splitting primary counts as nsplit particles against the
    refill room, so its sub-photons' secondaries fit; the run completes and
    gives the same histories as with the default queues."""
    stem = write_case("split5", tmp_path, ncase=100000)
    inp = tmp_path / "split5.inp"
    inp.write_text(inp.read_text().replace("nsplit = 5", "nsplit = 20"))
    prob = oracle.load_problem(stem)
    try:
        assert prob.nsplit == 20
        out = []
        for chunk in (1024, 0):
            with D.Engine(precision, chunk=chunk) as eng:
                eng.upload_media(prob.tab)
                pid = eng.upload_phantom(prob.ph, prob.reg)
                s, _, st, led = eng.run_dosxyz(pid, prob.src, prob.nperbatch, prob.nbatch, seed=(3, 9),
                                               nreg=prob.reg.nreg, ledger=True, nsplit=20)
            assert st.killed == 0
            out.append((s, led, st))
    finally:
        oracle.free_problem(prob)
    assert out[0][2].injections > 50 * out[1][2].injections
    assert np.allclose(out[0][0], out[1][0], rtol=1e-12, atol=1e-12 * out[1][0].max())
    assert np.allclose(out[0][1], out[1][1], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("case,precision", [("patient6MV", 0), ("electron6MeV", 0), ("patient6MV", 1)])
def test_schedule_and_reruns_do_not_change_histories(oracle, D, tmp_path, case, precision):
    """Histories are identified by (seed, index) only: a different number of
    pipelines, a small refill chunk (many injections, primaries mixed with
    older secondaries), particles handed to the next launch after every
    one or three loop trips (their lane state travels with them), handed
    on after every trip once a launch's input is used up, not drained at
    all, free lanes waiting until 16 of a wavefront are free before they
    take new particles, or a repeat reproduce the same energies up to FP64
    summation order."""
    stem = write_case(case, tmp_path, ncase=3000)
    prob = oracle.load_problem(stem)
    try:
        out = []
        for chunk, pipes, ytrips, drain, fmin in ((0, 0, 0, -1, 0), (0, 1, 0, 0, 0), (777, 3, 0, 0, 0),
                                                  (0, 2, 1, 0, 0), (5000, 0, 3, 0, 0), (2000, 2, 0, 64, 0),
                                                  (0, 0, 0, None, 0), (777, 2, 3, 0, 16)):
            # drain_min -1: these small launches drain too (by default, the
            # seventh configuration, a launch smaller than the device's
            # resident lanes does not)
            dm = 0 if drain is None or drain < 0 else -1
            with D.Engine(precision, chunk=chunk, pipelines=pipes, yield_trips=ytrips, drain_lanes=drain or 0,
                          drain_trips=1 if drain == 64 else 0, drain_min=dm, fetch_min=fmin) as eng:
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


@pytest.mark.parametrize("case,ytrips", [("slab6MV", 0), ("positron6MeV", 0), ("slab6MV", 6)])
def test_energy_ordered_fetch_keeps_histories(oracle, D, tmp_path, case, ytrips):
    """Charged-particle launches taking their input in order of falling
    energy (sort_min=1: every launch) run the same histories: each one's
    ledger row, every event count and every voxel equal the oracle's.  With
    a yield budget of 6 trips most charged particles yield and resume, so
    the sort moves their resume state as well."""
    r = run_both(oracle, D, write_case(case, tmp_path, ncase=1000), sort_min=1, yield_trips=ytrips)
    assert_same_histories(r, case + " sorted")
    if ytrips:
        assert r["st"].yields > 1000


@pytest.mark.parametrize("case", ["electron6MeV", "slab6MV"])
def test_longest_track_statistic(oracle, D, tmp_path, case):
    """longest_electron_trips: the loop trips of the run's longest charged
    track lie between the mean per track and the run's total.  A trip takes
    one or two condensed-history steps (a zero-length step runs fused with
    the next), and a track's last trip may take none (cut off on arrival),
    so steps / (2 tracks) <= longest <= steps + tracks.  A charged track
    starts with each charged primary, Moller or Bhabha secondary, Compton
    or photo electron and pair member; the counts come from the replayed
    run, whose events equal the oracle's."""
    r = run_both(oracle, D, write_case(case, tmp_path, ncase=500))
    assert_same_histories(r, case)
    st = r["st"]
    primaries_charged = st.histories if case.startswith("electron") else 0
    tracks = primaries_charged + st.moller + st.bhabha + st.compton + st.photo + 2 * st.pair
    assert 2 * st.longest_electron_trips * tracks >= st.electron_steps > 0
    assert st.longest_electron_trips <= st.electron_steps + tracks
    assert (st.longest_photon_trips > 0) == (st.photon_steps > 0)


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
    assert_same_histories(r, f"vacuum {energy} MeV charge {charge}")


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
    tot = led.reshape(-1, 4)[:, :3].sum(1)
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
