"""howfar / hownear of the egsphant voxel geometry (omc_dosxyz.c:187-334)
against brute force: region decoding, distance to the next plane with the
reference's z, x, y order and strict '<', and the perpendicular distance."""
import ctypes as C

import numpy as np

from ompmc_amd import _native as N, configs as cfg


def _problem(oracle, tmp_path):
    from cases import write_case
    stem = write_case("patient6MV", tmp_path, ncase=10)
    return oracle.load_problem(stem)


def brute_howfar(xb, yb, zb, ir, pos, d, ustep):
    nx, ny = len(xb) - 1, len(yb) - 1
    nz = len(zb) - 1
    v = ir - 1
    ix, iy, iz = v % nx, (v // nx) % ny, v // (nx * ny)
    irnew = ir
    best = ustep
    for axis, (b, i, n, stride) in ((2, (zb, iz, nz, nx * ny)), (0, (xb, ix, nx, 1)), (1, (yb, iy, ny, nx))):
        if d[axis] > 0:
            dist = (b[i + 1] - pos[axis]) / d[axis]
            if dist < best:
                best = dist
                irnew = ir + stride if i != n - 1 else 0
        elif d[axis] < 0:
            dist = -(pos[axis] - b[i]) / d[axis]
            if dist < best:
                best = dist
                irnew = ir - stride if i != 0 else 0
    return irnew, best


def test_howfar_hownear_against_brute_force(oracle, tmp_path):
    prob = _problem(oracle, tmp_path)
    L = oracle.lib()
    ctx = L.ora_create(C.byref(prob.tab), C.byref(prob.reg), C.byref(prob.ph), 1, oracle.PHILOX)
    try:
        nx, ny, nz = prob.ph.isize, prob.ph.jsize, prob.ph.ksize
        xb = np.ctypeslib.as_array(prob.ph.xbounds, shape=(nx + 1,))
        yb = np.ctypeslib.as_array(prob.ph.ybounds, shape=(ny + 1,))
        zb = np.ctypeslib.as_array(prob.ph.zbounds, shape=(nz + 1,))
        rng = np.random.default_rng(1)
        for _ in range(3000):
            ix, iy, iz = rng.integers(nx), rng.integers(ny), rng.integers(nz)
            pos = np.array([rng.uniform(xb[ix], xb[ix + 1]), rng.uniform(yb[iy], yb[iy + 1]),
                            rng.uniform(zb[iz], zb[iz + 1])])
            d = rng.normal(size=3)
            if rng.random() < 0.2:
                d[rng.integers(3)] = 0.0
            d /= np.linalg.norm(d)
            ir = 1 + ix + iy * nx + iz * nx * ny
            p = oracle.Particle(iq=0, ir=ir, e=1.0, x=pos[0], y=pos[1], z=pos[2], u=d[0], v=d[1], w=d[2])
            ustep0 = rng.choice([1e8, rng.uniform(0, 3.0)])
            idisc, irnew, us = C.c_int(0), C.c_int(ir), C.c_double(ustep0)
            L.ora_howfar(ctx, C.byref(p), C.byref(idisc), C.byref(irnew), C.byref(us))
            want_ir, want_us = brute_howfar(xb, yb, zb, ir, pos, d, ustep0)
            assert idisc.value == 0
            assert irnew.value == want_ir and us.value == want_us
            tperp = L.ora_hownear(ctx, C.byref(p))
            want = min(1e10, xb[ix + 1] - pos[0], pos[0] - xb[ix], yb[iy + 1] - pos[1], pos[1] - yb[iy],
                       zb[iz + 1] - pos[2], pos[2] - zb[iz])
            assert tperp == want
        # region 0 (outside) is discarded and has zero distance to the boundary
        p = oracle.Particle(iq=0, ir=0, e=1.0, x=0, y=0, z=0, u=0, v=0, w=1)
        idisc, irnew, us = C.c_int(0), C.c_int(0), C.c_double(5.0)
        L.ora_howfar(ctx, C.byref(p), C.byref(idisc), C.byref(irnew), C.byref(us))
        assert idisc.value == 1 and L.ora_hownear(ctx, C.byref(p)) == 0.0
    finally:
        L.ora_destroy(ctx)
        oracle.free_problem(prob)


def test_compute_drange_matches_numerical_integral(oracle, tmp_path):
    """computeDrange (path length between two energies in one table bin) is
    the integral of 1/(dE/dx) with log-linear dE/dx; check against a
    fine numerical quadrature of the same stopping-power model."""
    prob = _problem(oracle, tmp_path)
    L = oracle.lib()
    ctx = L.ora_create(C.byref(prob.tab), C.byref(prob.reg), C.byref(prob.ph), 1, oracle.PHILOX)
    try:
        t = prob.tab
        m = 2  # water in the patient media set
        d1 = np.ctypeslib.as_array(t.ededx1, shape=(t.nmed * 500,))
        d0 = np.ctypeslib.as_array(t.ededx0, shape=(t.nmed * 500,))
        for k in (10, 60, 120):
            e_hi = np.exp((k + 1 + 1 - t.eke0[m]) / t.eke1[m]) * 0.999
            e_lo = np.exp((k + 1 - t.eke0[m]) / t.eke1[m]) * 1.001
            lelke = int(np.log(e_hi) * t.eke1[m] + t.eke0[m]) - 1
            assert lelke == k
            got = L.ora_compute_drange(ctx, m, -1, lelke, e_hi, e_lo, np.log(e_hi), np.log(e_lo))
            es = np.linspace(e_lo, e_hi, 200001)
            dedx = d1[m * 500 + k] * np.log(es) + d0[m * 500 + k]
            want = np.trapezoid(1.0 / dedx, es)
            assert abs(got / want - 1) < 1e-5
    finally:
        L.ora_destroy(ctx)
        oracle.free_problem(prob)
