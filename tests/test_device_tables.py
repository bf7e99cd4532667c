"""initPhotonData on the GPU (omc_dev_build_photon) against the host builder
omc_build_photon, whose arrays are bit-identical to the reference's own
tables (test_tables.py).  The device evaluates the same arithmetic in the
same order; an entry may differ only where the device's exp or log rounds
differently from glibc's.  Checked per table as the transport uses it: the
fitted value intercept + slope * log(E) at both edges of every bin agrees to
1e-12 of the table's scale, and the log-energy grid (ge0, ge1) exactly.  A wrong branch, element order or threshold factor
moves values by far more."""
import ctypes as C

import numpy as np
import pytest

from ompmc_amd import _native as N

pytestmark = pytest.mark.gpu

MXGE = 2000
PAIRS = [("gmfp1", "gmfp0"), ("gbr11", "gbr10"), ("gbr21", "gbr20"), ("cohe1", "cohe0")]


@pytest.fixture(scope="module")
def D():
    from ompmc_amd import device
    return device


def host_tables(host, media):
    t = N.Tables()
    names = (C.c_char * 64 * 9)()
    for i, n in enumerate(media):
        names[i].value = n.encode()
    p = N.default_paths()
    N.check(host.omc_tables_build(C.byref(t), len(media), names, C.byref(p)))
    return t


@pytest.mark.parametrize("media", [["H2O700ICRU"], ["H2O700ICRU", "LUNG700ICRU", "ICRPBONE700ICRU"],
                                   ["AIR700ICRU", "H2O700ICRU", "AG700ICRU"]])
def test_device_photon_tables_match_host_builder(host, D, media):
    t = host_tables(host, media)
    nmed = len(media)
    G = nmed * MXGE
    want = {f: np.ctypeslib.as_array(getattr(t, f), shape=(G,)).copy() for pair in PAIRS for f in pair}
    ge0, ge1 = np.array(t.ge0[:nmed]), np.array(t.ge1[:nmed])
    # the same media with fresh photon arrays and the grid cleared
    d = N.Tables()
    C.memmove(C.byref(d), C.byref(t), C.sizeof(N.Tables))
    bufs = {f: np.full(G, np.nan) for f in want}
    for f, b in bufs.items():
        setattr(d, f, b.ctypes.data_as(C.POINTER(C.c_double)))
    for i in range(nmed):
        d.ge0[i] = d.ge1[i] = 0.0
    D.build_photon(d, N.default_paths().data_folder.decode())
    assert np.array_equal(np.array(d.ge0[:nmed]), ge0) and np.array_equal(np.array(d.ge1[:nmed]), ge1)
    same = 0
    for slope, icpt in PAIRS:
        for b in (bufs[slope], bufs[icpt]):
            assert np.isfinite(b).all()
        same += int(np.count_nonzero(bufs[slope] == want[slope]) + np.count_nonzero(bufs[icpt] == want[icpt]))
        for m in range(nmed):
            j = np.arange(MXGE)
            lo = (j + 1 - ge0[m]) / ge1[m]
            hi = (j + 2 - ge0[m]) / ge1[m]
            s = slice(m * MXGE, (m + 1) * MXGE)
            for x in (lo, hi):
                got = bufs[icpt][s] + bufs[slope][s] * x
                ref = want[icpt][s] + want[slope][s] * x
                scale = np.abs(ref).max()
                assert np.abs(got - ref).max() <= 1e-12 * scale, (media[m], slope)
    print(f"{media}: {same} of {8 * G} entries bit-identical to the host builder's")
    host.omc_tables_free(C.byref(t))
