"""omc_dev_run_dosxyz_noalloc, the allocation-free form of the dosxyz batch
loop, against omc_dev_run_dosxyz through the C ABI: two calls in a row on
one context, the second smaller (fewer batches and histories, so stale
tallies from the first would show), each equal to the allocating entry
point's run of the same histories to FP64 summation order; a run whose
tallies exceed the scratch array is refused."""
import ctypes as C

import numpy as np
import pytest

from cases import write_case

pytestmark = pytest.mark.gpu


def _run(D, eng, fn, pid, src, nper, nb, seed, nreg):
    L = D.lib()
    st = D.DevStats()
    s = np.zeros(nreg)
    s2 = np.zeros(nreg)
    rp = D.RunParams(pid, 1, (C.c_uint32 * 2)(*seed), 0, nper, nb, None, C.pointer(st))
    P = C.POINTER(C.c_double)
    rc = getattr(L, fn)(eng.ctx, C.byref(src), C.byref(rp), s.ctypes.data_as(P), s2.ctypes.data_as(P))
    return rc, s, s2, st


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_matches_allocating_entry_point(oracle, tmp_path, precision):
    from ompmc_amd import device as D
    L = D.lib()
    L.omc_dev_run_dosxyz_noalloc.argtypes = L.omc_dev_run_dosxyz.argtypes
    prob = oracle.load_problem(write_case("slab6MV", tmp_path, ncase=20000))
    try:
        nreg = prob.reg.nreg
        with D.Engine(precision) as eng:
            eng.upload_media(prob.tab)
            pid = eng.upload_phantom(prob.ph, prob.reg)
            for nper, nb, seed in ((3000, 10, (5, 1)), (700, 3, (6, 2))):
                rc, a, a2, sa = _run(D, eng, "omc_dev_run_dosxyz_noalloc", pid, prob.src, nper, nb, seed, nreg)
                assert rc == 0, D.lib().omc_dev_last_error()
                rc, b, b2, sb = _run(D, eng, "omc_dev_run_dosxyz", pid, prob.src, nper, nb, seed, nreg)
                assert rc == 0
                assert sa.histories == sb.histories == nper * nb and sa.killed == 0
                assert sa.deposits == sb.deposits and sa.electron_steps == sb.electron_steps
                scale = np.abs(b).max()
                assert scale > 0
                assert np.allclose(a, b, rtol=1e-12, atol=1e-12 * scale)
                assert np.allclose(a2, b2, rtol=1e-12, atol=1e-12 * np.abs(b2).max())
            # tallies beyond the scratch array: refused with OMC_ERANGE
            big = (64 << 20) // (8 * nreg) + 1
            rc, _, _, _ = _run(D, eng, "omc_dev_run_dosxyz_noalloc", pid, prob.src, 10, big, (1, 1), nreg)
            assert rc == -6 and b"scratch" in D.lib().omc_dev_last_error()
    finally:
        oracle.free_problem(prob)
