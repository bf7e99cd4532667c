"""matRad beamlets: the device's batch sums and its device-assembled sparse
Dij against the oracle's beamlet loop (omc_matrad.c:1389-1414 with
per-particle Philox streams) and the oracle's restatement of the per-beamlet
dose and threshold (omc_matrad.c:839-918, 1416-1479).

The restatement itself is pinned on the CPU against omc_batch_results, whose
output is byte-identical to the reference's .3ddose (test_oracle_reference):
the matRad conversion is the same expression with the incident fluence set
to the beamlet's total history count."""
import ctypes as C

import numpy as np
import pytest

from ompmc_amd import _native as N, configs as cfg

RM = 0.5109989461


def beamlet_problem(oracle, d, gantry=(0.0, 90.0), field=2.0, bixel=1.0):
    d.mkdir(parents=True, exist_ok=True)
    ph = cfg.patient_phantom(24, 20, 16, 1.0)
    pf = ph.write_egsphant(d / "phantom.egsphant")
    spec = cfg.write_spectrum(d / "6mv.spectrum")
    items = cfg.dosxyz_inp_items(d, pf, ncase=100, spectrum_file=spec)
    cfg.write_inp(d / "bl", items)
    prob = oracle.load_problem(d / "bl")
    bs = cfg.beamlets(gantry_deg=gantry, field=field, bixel=bixel, sad=100.0)
    src = N.beamlet_source(prob.src.spec, bs.ibeam, bs.src, bs.corner, bs.side1, bs.side2)
    return prob, src, bs


def phantom_arrays(prob):
    ph = prob.ph
    nvox = ph.isize * ph.jsize * ph.ksize
    xb = np.ctypeslib.as_array(ph.xbounds, shape=(ph.isize + 1,)).copy()
    yb = np.ctypeslib.as_array(ph.ybounds, shape=(ph.jsize + 1,)).copy()
    zb = np.ctypeslib.as_array(ph.zbounds, shape=(ph.ksize + 1,)).copy()
    rho = np.ctypeslib.as_array(ph.rho, shape=(nvox,)).copy()
    return xb, yb, zb, rho


def nper(nh, nbatch):
    return max(1, int(nh) // nbatch)


# ------------------------------------------------------------------ CPU ----
def test_matrad_column_restatement_equals_reference_dose(oracle, tmp_path):
    prob, _, _ = beamlet_problem(oracle, tmp_path)
    try:
        xb, yb, zb, rho = phantom_arrays(prob)
        nreg = prob.reg.nreg
        rng = np.random.default_rng(3)
        s = rng.gamma(0.3, 2.0, size=nreg) * (rng.random(nreg) < 0.4)
        s2 = s * s * 0.3
        nbatch, nhist = 7, 700
        # reference conversion (accumulateResults, iout = 1, fluence = nhist)
        nvox = nreg - 1
        dose = np.zeros(nvox)
        unc = np.zeros(nvox)
        P = C.POINTER(C.c_double)
        N.check(N.host().omc_batch_results(C.byref(prob.ph), s.ctypes.data_as(P), s2.ctypes.data_as(P), nbatch,
                                           float(nhist), 1, dose.ctypes.data_as(P), unc.ctypes.data_as(P)))
        for thr in (0.0, 0.01, 0.5):
            rows, vals = oracle.matrad_column(xb, yb, zb, rho, s, nbatch, nhist, thr)
            expect = np.nonzero(dose > max(dose.max(), 0.0) * thr)[0]
            assert np.array_equal(rows, expect)
            assert np.array_equal(vals, dose[expect])          # bit-identical
        # closed form for one water voxel: E/nbatch * 1.602e-10 / (rho V nhist)
        v = int(np.nonzero((rho == 1.0))[0][0])
        s1 = np.zeros(nreg)
        s1[v + 1] = 3.5
        rows, vals = oracle.matrad_column(xb, yb, zb, rho, s1, nbatch, nhist, 0.01)
        assert rows.tolist() == [v]
        assert vals[0] == pytest.approx(3.5 / nbatch * 1.602e-10 / (1.0 * 1.0 * nhist), rel=1e-14)
        # dose in air (density < 0.044) is never reported
        va = int(np.nonzero(rho < 0.044)[0][0])
        s1[va + 1] = 1e6
        rows, _ = oracle.matrad_column(xb, yb, zb, rho, s1, nbatch, nhist, 0.0)
        assert va not in rows
    finally:
        oracle.free_problem(prob)


def test_beamlet_primaries_geometry(oracle, tmp_path):
    """The oracle's beamlet initHistory (omc_matrad.c:1084-1254; the MEX
    cannot be built here) against geometry computed independently: the
    primary starts on the phantom's bounding surface, moves along the unit
    vector from the beam's source, on a line that crosses the bixel
    parallelogram, in the voxel that contains its position."""
    prob, src, bs = beamlet_problem(oracle, tmp_path, gantry=(0.0, 90.0, 135.0), field=4.0, bixel=2.0)
    try:
        xb, yb, zb, _ = phantom_arrays(prob)
        lo = np.array([xb[0], yb[0], zb[0]])
        hi = np.array([xb[-1], yb[-1], zb[-1]])
        emax = 6.0
        for b in range(src.nbeamlets):
            s0 = bs.src[bs.ibeam[b]]
            for p in oracle.beamlet_primaries(prob, src, b, range(40), seed=(3, b + 1)):
                pos = np.array([p.x, p.y, p.z])
                d = np.array([p.u, p.v, p.w])
                assert p.iq == 0 and 0.0 < p.e <= emax and p.wt == 1.0
                assert abs(np.linalg.norm(d) - 1.0) < 1e-12
                # on the bounding box surface, inside its extent
                assert np.all(pos >= lo - 1e-9) and np.all(pos <= hi + 1e-9)
                assert np.min(np.minimum(np.abs(pos - lo), np.abs(pos - hi))) < 1e-9
                # along the ray from the source
                ray = pos - s0
                assert np.linalg.norm(np.cross(ray / np.linalg.norm(ray), d)) < 1e-12 and ray @ d > 0
                # the ray crosses the bixel: solve s0 + t d = corner + a side1 + c side2
                M = np.column_stack([d, -bs.side1[b], -bs.side2[b]])
                t, a, c = np.linalg.solve(M, bs.corner[b] - s0)
                assert -1e-9 <= a <= 1 + 1e-9 and -1e-9 <= c <= 1 + 1e-9
                # region index of the containing voxel (interior side of the face)
                q = pos + 1e-9 * d
                ix = int(np.clip(np.searchsorted(xb, q[0]) - 1, 0, len(xb) - 2))
                iy = int(np.clip(np.searchsorted(yb, q[1]) - 1, 0, len(yb) - 2))
                iz = int(np.clip(np.searchsorted(zb, q[2]) - 1, 0, len(zb) - 2))
                ni, nj = len(xb) - 1, len(yb) - 1
                assert p.ir == 1 + ix + iy * ni + iz * ni * nj
    finally:
        oracle.free_problem(prob)


def test_beamlet_file_roundtrip_and_dij_reader(tmp_path):
    from ompmc_amd import formats as F
    bs = cfg.beamlets(gantry_deg=(0.0, 72.0), field=1.0, bixel=0.5)
    nh = np.arange(len(bs.ibeam)) + 10
    path = F.write_beamlet_file(tmp_path / "b.txt", bs.ibeam, bs.src, bs.corner, bs.side1, bs.side2, nh)
    lines = [ln for ln in path.read_text().splitlines() if ln.strip()]
    assert lines[0].split()[:2] == [str(len(bs.ibeam)), "2"]
    assert int(lines[3].split()[0]) == bs.ibeam[0] + 1 and int(lines[3].split()[-1]) == 10
    body = np.array([[float(t) for t in ln.split()[1:10]] for ln in lines[3:]])
    assert np.array_equal(body[:, 0:3], bs.corner)
    # .dij reader against a hand-made file
    jcs = np.array([0, 2, 2, 3], np.int64)
    irs = np.array([4, 9, 1], np.int64)
    vals = np.array([1.5, 2.5, 3.5])
    raw = F.DIJ_MAGIC + np.array([10, 3, 3], "<i8").tobytes() + jcs.tobytes() + irs.tobytes() + vals.tobytes()
    (tmp_path / "x.dij").write_bytes(raw)
    nr, nc, j2, i2, v2 = F.read_dij(tmp_path / "x.dij")
    assert (nr, nc) == (10, 3) and np.array_equal(j2, jcs) and np.array_equal(i2, irs) and np.array_equal(v2, vals)
    (tmp_path / "y.dij").write_bytes(raw[:-8])
    with pytest.raises(ValueError):
        F.read_dij(tmp_path / "y.dij")


# ------------------------------------------------------------------ GPU ----
NH = [40, 95, 3, 60, 25, 70, 10, 50]
NBATCH = 5


@pytest.fixture(scope="module")
def D():
    from ompmc_amd import device
    return device


@pytest.mark.gpu
def test_beamlet_sums_replay_oracle(oracle, D, tmp_path):
    prob, src, bs = beamlet_problem(oracle, tmp_path)
    try:
        assert src.nbeamlets == len(NH)
        with D.Engine(0) as eng:
            eng.upload_media(prob.tab)
            pid = eng.upload_phantom(prob.ph, prob.reg)
            sd, sd2, st = eng.run_beamlets(pid, src, NH, NBATCH, seed=(5, 6), hist0=100)
        assert st.histories == sum(nper(n, NBATCH) * NBATCH for n in NH)
        h0 = 100
        for b, nh in enumerate(NH):
            npb = nper(nh, NBATCH)
            so, so2 = oracle.run_beamlet_philox(prob, src, b, npb, NBATCH, seed=(5, 6), hist0=h0)
            h0 += npb * NBATCH
            scale = max(np.abs(so).max(), 1e-300)
            assert np.abs(sd[b] - so).max() <= 1e-9 * scale, f"beamlet {b}"
            assert np.abs(sd2[b] - so2).max() <= 1e-9 * max(np.abs(so2).max(), 1e-300), f"beamlet {b}"
            assert so[1:].sum() > 0
    finally:
        oracle.free_problem(prob)


def _compare_columns(jcs, irs, vals, cols, tol=1e-9):
    assert len(jcs) == len(cols) + 1 and jcs[0] == 0
    for b, (rows, v) in enumerate(cols):
        r_d = irs[jcs[b]:jcs[b + 1]]
        v_d = vals[jcs[b]:jcs[b + 1]]
        assert np.all(np.diff(r_d) > 0), "rows must be in voxel order"
        common, i_d, i_o = np.intersect1d(r_d, rows, return_indices=True)
        assert np.allclose(v_d[i_d], v[i_o], rtol=tol, atol=0), f"column {b}"
        # rows present on one side only must sit at the threshold
        only = len(r_d) + len(rows) - 2 * len(common)
        assert only <= max(1, len(rows) // 1000), f"column {b}: {only} rows differ"


@pytest.mark.gpu
def test_sparse_dij_matches_oracle_columns(oracle, D, tmp_path):
    prob, src, bs = beamlet_problem(oracle, tmp_path)
    try:
        xb, yb, zb, rho = phantom_arrays(prob)
        with D.Engine(0) as eng:
            eng.upload_media(prob.tab)
            pid = eng.upload_phantom(prob.ph, prob.reg)
            jcs, irs, vals, st = eng.beamlets_sparse(pid, src, NH, NBATCH, 0.01, seed=(9, 1))
        cols = []
        h0 = 0
        for b, nh in enumerate(NH):
            npb = nper(nh, NBATCH)
            so, _ = oracle.run_beamlet_philox(prob, src, b, npb, NBATCH, seed=(9, 1), hist0=h0)
            h0 += npb * NBATCH
            cols.append(oracle.matrad_column(xb, yb, zb, rho, so, NBATCH, npb * NBATCH, 0.01))
        _compare_columns(jcs, irs, vals, cols)
        assert len(irs) == jcs[-1] > 0
    finally:
        oracle.free_problem(prob)


@pytest.mark.gpu
def test_sparse_dij_fp32_consistent_with_its_sums_and_capacity_retry(oracle, D, tmp_path):
    """FP32 transport: the device-assembled Dij equals the restatement
    applied to the device's own batch sums of the same histories, and a Dij
    larger than the caller's arrays is collected without rerunning."""
    prob, src, bs = beamlet_problem(oracle, tmp_path)
    try:
        xb, yb, zb, rho = phantom_arrays(prob)
        nh = [400] * src.nbeamlets
        with D.Engine(1) as eng:
            eng.upload_media(prob.tab)
            pid = eng.upload_phantom(prob.ph, prob.reg)
            sd, _, _ = eng.run_beamlets(pid, src, nh, NBATCH, seed=(2, 2))
            jcs, irs, vals, _ = eng.beamlets_sparse(pid, src, nh, NBATCH, 0.05, seed=(2, 2), cap=1)
        cols = [oracle.matrad_column(xb, yb, zb, rho, sd[b], NBATCH, 400, 0.05) for b in range(src.nbeamlets)]
        _compare_columns(jcs, irs, vals, cols, tol=1e-9)
    finally:
        oracle.free_problem(prob)


@pytest.mark.gpu
def test_beamlet_argument_errors(oracle, D, tmp_path):
    prob, src, bs = beamlet_problem(oracle, tmp_path)
    try:
        with D.Engine(0) as eng:
            eng.upload_media(prob.tab)
            pid = eng.upload_phantom(prob.ph, prob.reg)
            with pytest.raises(RuntimeError, match="histories"):
                eng.run_beamlets(pid, src, [10, 0] + [10] * (src.nbeamlets - 2), 2)
            bad = N.beamlet_source(prob.src.spec, np.full(src.nbeamlets, 5), bs.src, bs.corner, bs.side1,
                                   bs.side2)
            with pytest.raises(RuntimeError, match="beam"):
                eng.run_beamlets(pid, bad, [10] * src.nbeamlets, 2)
    finally:
        oracle.free_problem(prob)
