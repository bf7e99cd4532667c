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


@pytest.mark.parametrize("case,precision", [("slab6MV", 0), ("patient6MV", 0), ("patient6MV", 1)])
def test_delta_tracking_matches_voxel_stepping(oracle, D, tmp_path, case, precision):
    stem = write_case(case, tmp_path, ncase=200000)
    inp = tmp_path / f"{case}.inp"
    inp.write_text(inp.read_text().replace("nbatch = 10", "nbatch = 50"))
    prob = oracle.load_problem(stem)
    try:
        nb, nh = prob.nbatch, prob.nhist
        sv, sv2, stv, lv = _run(D, prob, 0, -1, (31, 7), ledger=True)               # voxel stepping, FP64
        sw, sw2, stw, lw = _run(D, prob, precision, 1, (32, 7), ledger=True)       # delta tracking
    finally:
        oracle.free_problem(prob)
    assert stw.woodcock_over == 0 and stw.killed == 0
    # a few sites per flight instead of a stop at every voxel boundary
    assert stw.photon_steps < 0.25 * stv.photon_steps, (stw.photon_steps, stv.photon_steps)
    mv, ev = _mv(sv, sv2, nb)
    mw, ew = _mv(sw, sw2, nb)
    sel = np.where(mv[1:] > 0.1 * mv[1:].max())[0] + 1
    z = (mw[sel] - mv[sel]) / np.sqrt(ev[sel] ** 2 + ew[sel] ** 2 + 1e-300)
    frac = (np.abs(z) < 2).mean()
    assert len(sel) > 50 and frac >= 0.93, f"{case}: {frac:.3f} of {len(sel)} voxels within 2 sigma"
    # integral: mean energy deposited per history, independent histories
    dv, dw = lv.reshape(-1, 3)[:, 0], lw.reshape(-1, 3)[:, 0]
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
    led = led.reshape(-1, 3)
    assert st.woodcock_over == 0 and st.killed == 0
    assert st.photon_steps < 10 * prob.nhist
    assert np.allclose(led.sum(1), 2.0, rtol=tol, atol=tol)
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
    tracking against FP64 voxel stepping, voxels above 10% of each
    beamlet's maximum within 2 sigma."""
    from ompmc_amd import _native as N, configs as cfg
    ph = cfg.patient_phantom(24, 20, 16, 1.0)
    pf = ph.write_egsphant(tmp_path / "phantom.egsphant")
    spec = cfg.write_spectrum(tmp_path / "6mv.spectrum")
    cfg.write_inp(tmp_path / "bl", cfg.dosxyz_inp_items(tmp_path, pf, ncase=100, spectrum_file=spec))
    prob = oracle.load_problem(tmp_path / "bl")
    try:
        bs = cfg.beamlets(gantry_deg=(0.0, 90.0), field=2.0, bixel=1.0, sad=100.0)
        src = N.beamlet_source(prob.src.spec, bs.ibeam, bs.src, bs.corner, bs.side1, bs.side2)
        nh, nb = [40000] * src.nbeamlets, 20
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
    inside, total = 0, 0
    for b in range(src.nbeamlets):
        mv, ev = _mv(out[0][0][b], out[0][1][b], nb)
        mw, ew = _mv(out[1][0][b], out[1][1][b], nb)
        sel = np.where(mv[1:] > 0.1 * mv[1:].max())[0] + 1
        z = (mw[sel] - mv[sel]) / np.sqrt(ev[sel] ** 2 + ew[sel] ** 2 + 1e-300)
        inside += int((np.abs(z) < 2).sum())
        total += len(sel)
    assert total > 100 and inside / total >= 0.93, f"{inside / total:.3f} of {total} voxels within 2 sigma"


def test_delta_tracking_through_vacuum_voxels(D, tmp_path):
    """A vacuum slab inside the grid (no attenuation, every site there is
    fictitious) and a denser region behind it: the same dose as voxel
    stepping within statistics, energy conserved per history."""
    from ompmc_amd import configs as cfg
    from ompmc_amd import problem as P
    ph = cfg.water_cube(12, 6.0)
    ph.med[:, :, 4:6] = 0
    ph.rho[:, :, 4:6] = 0.0
    ph.rho[:, :, 8:] = 1.8
    pf = ph.write_egsphant(tmp_path / "vac.egsphant")
    items = cfg.dosxyz_inp_items(tmp_path, pf, ncase=200000, nbatch=20, mono=3.0, field=(-1, 1, -1, 1))
    cfg.write_inp(tmp_path / "vac", items)
    prob = P.load_problem(tmp_path / "vac")
    try:
        nb = prob.nbatch
        sv, sv2, stv, lv = _run(D, prob, 0, -1, (51, 2), ledger=True)
        sw, sw2, stw, lw = _run(D, prob, 1, 1, (52, 2), ledger=True)
        nh = prob.nperbatch * prob.nbatch
    finally:
        P.free_problem(prob)
    assert stw.woodcock_over == 0 and stw.killed == 0
    assert np.allclose(lw.reshape(-1, 3).sum(1), 3.0, rtol=2e-6, atol=2e-6)
    mv, ev = _mv(sv, sv2, nb)
    mw, ew = _mv(sw, sw2, nb)
    sel = np.where(mv[1:] > 0.1 * mv[1:].max())[0] + 1
    z = (mw[sel] - mv[sel]) / np.sqrt(ev[sel] ** 2 + ew[sel] ** 2 + 1e-300)
    frac = (np.abs(z) < 2).mean()
    assert len(sel) > 50 and frac >= 0.93, f"{frac:.3f} of {len(sel)} voxels within 2 sigma"
    dv, dw = lv.reshape(-1, 3)[:, 0], lw.reshape(-1, 3)[:, 0]
    se = np.sqrt(dv.var() / nh + dw.var() / nh)
    assert abs(dw.mean() - dv.mean()) < 4 * se, (dw.mean(), dv.mean(), se)
