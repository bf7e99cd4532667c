"""Photon flights by delta (Woodcock) tracking.

The reference steps a photon voxel by voxel, decrementing its mean free
paths (ompmc.c:1951-2019).  Photons deposit nothing in flight, so sampling
the interaction sites against a majorant attenuation coefficient and
accepting each with probability mu/mu_max gives the same dose
distribution with a handful of sites per flight.  The FP32 engine does
this by default where a phantom allows it; the FP64 engine keeps voxel
stepping (history-by-history parity with the oracle) unless asked.

Checks: the dose agrees with voxel stepping within statistics (voxels
within 2 sigma, integral energy within 4 standard errors of the per-history
ledger), no site ever exceeds the majorant, energy is conserved history by
history, and a phantom the tracking cannot serve (non-uniform grid) falls
back to voxel stepping exactly."""
import numpy as np
import pytest

import dosestats as DS
from cases import write_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def D():
    from ompmc_amd import device
    return device


def _run(D, prob, precision, woodcock, seed, ledger=False):
    with D.Engine(precision, woodcock=woodcock) as eng:
        eng.upload_media(prob.tab)
        pid = eng.upload_phantom(prob.ph, prob.reg)
        return eng.run_dosxyz(pid, prob.src, prob.nperbatch, prob.nbatch, seed=seed, nreg=prob.reg.nreg,
                              ledger=ledger)


def _mv(s, s2, nb):
    m = s / nb
    return m, np.sqrt(np.maximum(s2 / nb - m * m, 0) / (nb - 1))


@pytest.mark.parametrize("case,precision", [("statSlab", 0), ("statPatient", 0), ("statPatient", 1)])
def test_delta_tracking_matches_voxel_stepping(oracle, D, tmp_path, case, precision):
    """The equal-accuracy gate (>= 95% of the voxels above 10% of Dmax within
    2 sigma, integral within 0.5%) between delta tracking and FP64 voxel
    stepping, 1e7 histories and 100 batches each, independent streams."""
    prob = oracle.load_problem(write_case(case, tmp_path, ncase=10_000_000, nbatch=100))
    try:
        nb = prob.nbatch
        sv, sv2, stv, _ = _run(D, prob, 0, -1, (31, 7))               # voxel stepping, FP64
        sw, sw2, stw, _ = _run(D, prob, precision, 1, (32, 7))       # delta tracking
    finally:
        oracle.free_problem(prob)
    assert stw.woodcock_over == 0 and stw.killed == 0
    # a few sites per flight instead of a stop at every voxel boundary
    assert stw.photon_steps < 0.25 * stv.photon_steps, (stw.photon_steps, stv.photon_steps)
    DS.assert_equal_accuracy(DS.compare((sw, sw2, nb), (sv, sv2, nb)), f"{case} delta tracking, precision {precision}")


@pytest.mark.parametrize("case,precision", [("slab6MV", 0), ("patient6MV", 1)])
def test_delta_tracking_integral_energy_per_history(oracle, D, tmp_path, case, precision):
    """Mean energy deposited per history (from the per-history ledger) of
    delta tracking against voxel stepping: within 4 standard errors of the
    independent per-history estimates."""
    prob = oracle.load_problem(write_case(case, tmp_path, ncase=200000))
    try:
        nh = prob.nhist
        _, _, _, lv = _run(D, prob, 0, -1, (31, 7), ledger=True)
        _, _, stw, lw = _run(D, prob, precision, 1, (32, 7), ledger=True)
    finally:
        oracle.free_problem(prob)
    assert stw.woodcock_over == 0 and stw.killed == 0
    dv, dw = lv.reshape(-1, 4)[:, 0], lw.reshape(-1, 4)[:, 0]
    se = np.sqrt(dv.var() / nh + dw.var() / nh)
    assert abs(dw.mean() - dv.mean()) < 4 * se, (dw.mean(), dv.mean(), se)


@pytest.mark.parametrize("precision,tol", [(0, 1e-12), (1, 2e-6)])
def test_delta_tracking_conserves_energy_per_history(oracle, D, tmp_path, precision, tol):
    """Every 2 MeV history's energy is deposited in the phantom, in region 0
    or leaves it (ledger columns), and the ledger's deposits are the
    tallies."""
    stem = write_case("water2MeV", tmp_path, ncase=20000)
    prob = oracle.load_problem(stem)
    try:
        s, _, st, led = _run(D, prob, precision, 1, (5, 9), ledger=True)
    finally:
        oracle.free_problem(prob)
    led = led.reshape(-1, 4)
    assert st.woodcock_over == 0 and st.killed == 0
    assert st.photon_steps < 10 * prob.nhist
    assert np.allclose(led[:, :3].sum(1), 2.0, rtol=tol, atol=tol)
    assert abs(led[:, :2].sum() - s.sum()) < 1e-9 * s.sum()


def test_non_uniform_grid_falls_back_to_voxel_stepping(D, tmp_path):
    """A phantom whose voxels differ in size cannot be tracked by position:
    asking for delta tracking there must give exactly the voxel-stepping
    histories (FP64, same streams)."""
    from ompmc_amd import configs as cfg
    from ompmc_amd import problem as P
    ph = cfg.water_cube(10, 5.0)
    ph.xb = np.concatenate([ph.xb[:5], ph.xb[5:] + np.linspace(0.0, 0.3, len(ph.xb) - 5)])
    ph.med[:, :, 6:] = 1
    ph.rho[:, :, 6:] = 1.4
    pf = ph.write_egsphant(tmp_path / "nonuni.egsphant")
    items = cfg.dosxyz_inp_items(tmp_path, pf, ncase=4000, mono=2.0, field=(-1, 1, -1, 1))
    cfg.write_inp(tmp_path / "nonuni", items)
    prob = P.load_problem(tmp_path / "nonuni")
    try:
        so, so2, sto, _ = _run(D, prob, 0, -1, (3, 3))
        sw, sw2, stw, _ = _run(D, prob, 0, 1, (3, 3))
    finally:
        P.free_problem(prob)
    assert stw.photon_steps == sto.photon_steps
    scale = np.abs(so).max()
    assert scale > 0
    assert np.allclose(sw, so, rtol=0, atol=1e-12 * scale)


def test_delta_tracking_beamlets_match_voxel_stepping(oracle, D, tmp_path):
    """matRad beamlets (primaries start on the phantom's surface, coming in
    from two gantry angles): per-beamlet dose of the FP32 engine with delta
    tracking against FP64 voxel stepping, pooled over the beamlets: the
    equal-accuracy gate over the voxels above 10% of each beamlet's
    maximum."""
    from ompmc_amd import _native as N, configs as cfg
    ph = cfg.patient_phantom(64, 64, 48, 0.5)
    pf = ph.write_egsphant(tmp_path / "phantom.egsphant")
    spec = cfg.write_spectrum(tmp_path / "6mv.spectrum")
    cfg.write_inp(tmp_path / "bl", cfg.dosxyz_inp_items(tmp_path, pf, ncase=100, spectrum_file=spec))
    prob = oracle.load_problem(tmp_path / "bl")
    try:
        bs = cfg.beamlets(gantry_deg=(0.0, 90.0), field=8.0, bixel=4.0, sad=100.0)
        src = N.beamlet_source(prob.src.spec, bs.ibeam, bs.src, bs.corner, bs.side1, bs.side2)
        nh, nb = [2_000_000] * src.nbeamlets, 100
        out = {}
        for prec, wc, seed in ((0, -1, (41, 1)), (1, 0, (42, 1))):
            with D.Engine(prec, woodcock=wc) as eng:
                eng.upload_media(prob.tab)
                pid = eng.upload_phantom(prob.ph, prob.reg)
                s, s2, st = eng.run_beamlets(pid, src, nh, nb, seed=seed)
                out[prec] = (s, s2, st)
    finally:
        oracle.free_problem(prob)
    assert out[1][2].woodcock_over == 0 and out[1][2].killed == 0
    pooled = {"frac": 0.0, "n": 0, "dof": 2 * nb - 2, "ratio": 1.0}
    inside = 0
    for b in range(src.nbeamlets):
        r = DS.compare((out[1][0][b], out[1][1][b], nb), (out[0][0][b], out[0][1][b], nb))
        inside += r["frac"] * r["n"]
        pooled["n"] += r["n"]
        assert abs(r["ratio"] - 1) <= 0.005, (b, r)
    pooled["frac"] = inside / pooled["n"]
    DS.assert_equal_accuracy(pooled, "beamlets, delta tracking FP32 vs voxel stepping FP64")


def test_delta_tracking_through_vacuum_voxels(D, tmp_path):
    """A vacuum slab inside the grid (no attenuation, every site there is
    fictitious) and a denser region behind it: the equal-accuracy gate
    against voxel stepping, energy conserved per history."""
    from ompmc_amd import configs as cfg
    from ompmc_amd import problem as P
    ph = cfg.water_cube(48, 12.0)
    ph.med[16:24, :, :] = 0
    ph.rho[16:24, :, :] = 0.0
    ph.rho[32:, :, :] = 1.8
    pf = ph.write_egsphant(tmp_path / "vac.egsphant")
    items = cfg.dosxyz_inp_items(tmp_path, pf, ncase=10_000_000, nbatch=100, mono=3.0, field=(-5, 5, -5, 5))
    cfg.write_inp(tmp_path / "vac", items)
    prob = P.load_problem(tmp_path / "vac")
    try:
        nb = prob.nbatch
        sv, sv2, stv, _ = _run(D, prob, 0, -1, (51, 2))
        sw, sw2, stw, _ = _run(D, prob, 1, 1, (52, 2))
    finally:
        P.free_problem(prob)
    assert stw.woodcock_over == 0 and stw.killed == 0
    DS.assert_equal_accuracy(DS.compare((sw, sw2, nb), (sv, sv2, nb)), "vacuum slab, delta tracking")
    items = cfg.dosxyz_inp_items(tmp_path, pf, ncase=100000, nbatch=10, mono=3.0, field=(-5, 5, -5, 5))
    cfg.write_inp(tmp_path / "vac_small", items)
    prob = P.load_problem(tmp_path / "vac_small")
    try:
        _, _, stw, lw = _run(D, prob, 1, 1, (52, 2), ledger=True)
    finally:
        P.free_problem(prob)
    assert stw.woodcock_over == 0 and stw.killed == 0
    assert np.allclose(lw.reshape(-1, 4)[:, :3].sum(1), 3.0, rtol=2e-6, atol=2e-6)
