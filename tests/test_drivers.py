"""The omc_dosxyz and omc_matrad executables: command line and input errors
(CPU, before any device call) and end-to-end runs on the GPU checked
against the oracle (dosxyz) and the C ABI (matrad)."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from cases import write_case
from ompmc_amd import configs as cfg, formats as F

BIN = Path(__file__).resolve().parents[1] / "ompmc_amd" / "bin"


def run(exe, *args, timeout=300):
    return subprocess.run([str(BIN / exe), *map(str, args)], capture_output=True, text=True, timeout=timeout)


def matrad_inp(d, ncase=200, nbatch=5, thr=0.01, beamlet_rows=None):
    d.mkdir(parents=True, exist_ok=True)
    ph = cfg.patient_phantom(24, 20, 16, 1.0)
    pf = ph.write_egsphant(d / "phantom.egsphant")
    spec = cfg.write_spectrum(d / "6mv.spectrum")
    items = cfg.dosxyz_inp_items(d, pf, ncase=ncase, nbatch=nbatch, spectrum_file=spec, seeds=(11, 12))
    for k in ("collimator bounds", "ssd", "charge"):
        items.pop(k, None)
    bs = cfg.beamlets(gantry_deg=(0.0, 90.0), field=2.0, bixel=1.0)
    nh = np.array([150, 40, 300, 90, 120, 60, 10, 200])
    items["beamlet file"] = str(F.write_beamlet_file(d / "beamlets.txt", bs.ibeam, bs.src, bs.corner, bs.side1,
                                                     bs.side2, nh))
    items["relative dose threshold"] = thr
    cfg.write_inp(d / "plan", items)
    return d / "plan", bs, nh


def test_usage_error_without_arguments():
    for exe in ("omc_dosxyz", "omc_matrad"):
        r = run(exe)
        assert r.returncode != 0 and "usage" in r.stderr


@pytest.mark.gpu
def test_dosxyz_runs_photon_splitting(tmp_path):
    stem = write_case("split5", tmp_path)
    r = run("omc_dosxyz", "-i", stem, "-o", "out")
    assert r.returncode == 0, r.stderr
    assert "Photon splitting enabled, nsplit = 5" in r.stdout
    assert (tmp_path / "out.3ddose").stat().st_size > 0


def test_dosxyz_reports_missing_keys(tmp_path):
    stem = write_case("water2MeV", tmp_path)
    inp = Path(str(stem) + ".inp")
    inp.write_text(inp.read_text().replace("phantom file", "phantom fil"))
    r = run("omc_dosxyz", "-i", stem, "-o", "out")
    assert r.returncode != 0 and "phantom file" in (r.stdout + r.stderr)


def test_matrad_rejects_malformed_beamlet_file(tmp_path):
    stem, _, _ = matrad_inp(tmp_path)
    (tmp_path / "beamlets.txt").write_text("2 1\n0 -100 0\n1 0 0 0 1 0 0 0 0 1\n3 0 0 0 1 0 0 0 0 1\n")
    r = run("omc_matrad", "-i", stem, "-o", "dij")
    assert r.returncode != 0 and "malformed" in r.stdout


@pytest.mark.gpu
def test_dosxyz_end_to_end_matches_oracle(oracle, tmp_path):
    stem = write_case("water2MeV", tmp_path)
    r = run("omc_dosxyz", "-i", stem, "-o", "out", "--verbose")
    assert r.returncode == 0, r.stdout + r.stderr
    ni, nj, nk, xb, yb, zb, dose, unc = F.read_3ddose(tmp_path / "out.3ddose")
    prob = oracle.load_problem(stem)
    try:
        s, s2, _ = oracle.run_dosxyz_philox(prob, seed=tuple(prob.seeds), nthreads=0)
        d_o, u_o = oracle.dose_from_sums(prob, s, s2)
    finally:
        oracle.free_problem(prob)
    assert (ni, nj, nk) == (12, 12, 12)
    # the file prints %e (7 significant digits) and %f
    assert np.allclose(dose, d_o, rtol=2e-6, atol=1e-30)
    assert np.allclose(unc, u_o, atol=2e-6)


@pytest.mark.gpu
def test_matrad_end_to_end_matches_device_api(oracle, tmp_path):
    from ompmc_amd import _native as N, device as D
    stem, bs, nh = matrad_inp(tmp_path)
    r = run("omc_matrad", "-i", stem, "-o", "dij")
    assert r.returncode == 0, r.stdout + r.stderr
    nrows, ncols, jcs, irs, vals = F.read_dij(tmp_path / "dij.dij")
    assert (nrows, ncols) == (24 * 20 * 16, len(nh))
    # same problem through the C ABI
    items = (Path(str(stem) + ".inp")).read_text()
    dd = tmp_path / "api"
    dd.mkdir()
    ph = cfg.patient_phantom(24, 20, 16, 1.0)
    pf = ph.write_egsphant(dd / "phantom.egsphant")
    spec = cfg.write_spectrum(dd / "6mv.spectrum")
    cfg.write_inp(dd / "p", cfg.dosxyz_inp_items(dd, pf, ncase=100, spectrum_file=spec))
    prob = oracle.load_problem(dd / "p")
    try:
        src = N.beamlet_source(prob.src.spec, bs.ibeam, bs.src, bs.corner, bs.side1, bs.side2)
        with D.Engine(0) as eng:
            eng.upload_media(prob.tab)
            pid = eng.upload_phantom(prob.ph, prob.reg)
            j2, i2, v2, _ = eng.beamlets_sparse(pid, src, nh, 5, 0.01, seed=(11, 12))
    finally:
        oracle.free_problem(prob)
    assert "relative dose threshold" in items
    assert np.array_equal(jcs, j2) and np.array_equal(irs, i2)
    assert np.allclose(vals, v2, rtol=1e-12, atol=0)


@pytest.mark.gpu
def test_dosxyz_photon_flight_mode_follows_environment(tmp_path):
    """OMC_PRECISION=fp32 gives delta-tracked photon flights unless
    OMC_WOODCOCK=0 asks for the reference's voxel stepping; FP64 steps
    voxel by voxel unless OMC_WOODCOCK=1."""
    import os
    stem = write_case("water2MeV", tmp_path, ncase=2000)
    for env, expect in (({"OMC_PRECISION": "fp32"}, "delta tracking"),
                        ({"OMC_PRECISION": "fp32", "OMC_WOODCOCK": "0"}, "voxel stepping"),
                        ({}, "voxel stepping"),
                        ({"OMC_WOODCOCK": "on"}, "delta tracking")):
        e = {k: v for k, v in os.environ.items() if k not in ("OMC_PRECISION", "OMC_WOODCOCK")}
        e.update(env)
        r = subprocess.run([str(BIN / "omc_dosxyz"), "-i", str(stem), "-o", "out"], capture_output=True, text=True,
                           timeout=300, env=e, cwd=tmp_path)
        assert r.returncode == 0, r.stderr
        assert f"photon flights by {expect}" in r.stdout, (env, r.stdout[-500:])
