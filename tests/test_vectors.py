"""Per-function parity vectors of the reference (tests/golden/
vectors_slab6MV.npz, written by `ref_dump vectors` from the reference build
on the slab case's geometry and media; make_golden.py).

The oracle's howfar / hownear (omc_dosxyz.c:187-334), computeDrange /
computeEloss (ompmc.c:3979-4108) and the photon table lookups pwlfInterval /
pwlfEval (ompmc.c:203-211 at 1952-1969) must reproduce every recorded output
bit for bit: positions on and between voxel boundaries, axis-parallel and
oblique directions, the outside region, step limits below and above the
next boundary, both charges, energies across each medium's table range."""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

from cases import write_case

VEC = Path(__file__).resolve().parent / "golden" / "vectors_slab6MV.npz"


@pytest.fixture(scope="module")
def slab(oracle, tmp_path_factory):
    prob = oracle.load_problem(write_case("slab6MV", tmp_path_factory.mktemp("slab")))
    L = oracle.lib()
    ctx = L.ora_create(C.byref(prob.tab), C.byref(prob.reg), C.byref(prob.ph), 1, oracle.PHILOX)
    yield prob, ctx, L
    L.ora_destroy(ctx)
    oracle.free_problem(prob)


def test_geometry_vectors(oracle, slab):
    prob, ctx, L = slab
    v = np.load(VEC)["geometry"]
    assert len(v) >= 500
    for r in v:
        p = oracle.Particle()
        p.ir = int(r[0])
        p.x, p.y, p.z, p.u, p.v, p.w = r[1:7]
        idisc, irnew, ustep = C.c_int(0), C.c_int(int(r[0])), C.c_double(r[7])
        L.ora_howfar(ctx, C.byref(p), C.byref(idisc), C.byref(irnew), C.byref(ustep))
        assert (idisc.value, irnew.value) == (int(r[8]), int(r[9])), r
        assert ustep.value == r[10], r                       # bit for bit
        assert L.ora_hownear(ctx, C.byref(p)) == r[11], r
    # the vectors cover what they claim to
    assert (v[:, 0] == 0).any() and (v[:, 8] == 1).any()
    assert (v[:, 10] < v[:, 7]).any() and (v[:, 10] == v[:, 7]).any()
    assert ((v[:, 4] == 0) & (v[:, 5] == 0)).any()


def test_csda_vectors(oracle, slab):
    prob, ctx, L = slab
    v = np.load(VEC)["csda"]
    for r in v:
        imed, iq, lelke = int(r[0]), int(r[1]), int(r[2])
        eke, elke, ekei, elkei, rhof, tustep, rng = r[3:10]
        dr = L.ora_compute_drange(ctx, imed, iq, lelke, eke, ekei, elke, elkei)
        assert dr == r[10], r
        de = L.ora_compute_eloss(ctx, imed, iq, rhof, tustep, rng, eke, elke, lelke)
        assert de == r[11], r
    assert set(np.unique(v[:, 1])) == {-1.0, 1.0} and len(np.unique(v[:, 0])) == prob.tab.nmed


def test_photon_table_lookup_vectors(slab):
    prob, _, _ = slab
    tab = prob.tab
    n = tab.nmed * 2000
    arr = {k: np.ctypeslib.as_array(getattr(tab, k), shape=(n,)) for k in
           ("gmfp0", "gmfp1", "cohe0", "cohe1", "gbr10", "gbr11", "gbr20", "gbr21")}
    ge0 = np.ctypeslib.as_array(tab.ge0)
    ge1 = np.ctypeslib.as_array(tab.ge1)
    for r in np.load(VEC)["pwlf"]:
        m = int(r[0])
        gle = np.log(r[1])
        lgle = int(gle * ge1[m] + ge0[m]) - 1
        assert lgle == int(r[2])
        ix = m * 2000 + lgle
        for k, (a1, a0) in enumerate((("gmfp1", "gmfp0"), ("cohe1", "cohe0"), ("gbr11", "gbr10"),
                                      ("gbr21", "gbr20"))):
            assert gle * arr[a1][ix] + arr[a0][ix] == r[3 + k], (r, a1)
