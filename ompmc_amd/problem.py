"""omc_dosxyz problems from .inp files through the host library
(omc_dosxyz_problem_load: initPhantom, initMediaData, initSource,
initRegions, initVrt -- omc_dosxyz.c:1155-1231) and the dose and
uncertainty of batch sums (accumulateResults, omc_dosxyz.c:719-799)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N

_dp = C.POINTER(C.c_double)


class DosxyzProblem(C.Structure):
    _fields_ = [("inp_n", C.c_int), ("inp_key", (C.c_char * 256) * 80), ("inp_value", (C.c_char * 256) * 80),
                ("ph", N.Phantom), ("tab", N.Tables), ("reg", N.Regions), ("src", N.SourceDosxyz),
                ("nsplit", C.c_int), ("seeds", C.c_int * 2), ("nhist", C.c_int), ("nbatch", C.c_int),
                ("nperbatch", C.c_int), ("output_folder", C.c_char * 256)]


def load_problem(stem) -> DosxyzProblem:
    """Reads <stem>.inp and everything it names; raises RuntimeError with the
    library's message on failure."""
    h = N.host()
    h.omc_dosxyz_problem_load.argtypes = [C.POINTER(DosxyzProblem), C.c_char_p]
    p = DosxyzProblem()
    N.check(h.omc_dosxyz_problem_load(C.byref(p), str(stem).encode()))
    return p


def free_problem(p: DosxyzProblem):
    h = N.host()
    h.omc_dosxyz_problem_free.argtypes = [C.POINTER(DosxyzProblem)]
    h.omc_dosxyz_problem_free(C.byref(p))


def dose_from_sums(prob: DosxyzProblem, s, s2, nbatch=None, fluence=None):
    """(dose, relative uncertainty) per voxel as omc_dosxyz writes them
    (iout = 1; the reference divides by the histories per batch)."""
    h = N.host()
    nvox = prob.ph.isize * prob.ph.jsize * prob.ph.ksize
    dose = np.zeros(nvox)
    unc = np.zeros(nvox)
    s = np.ascontiguousarray(s, dtype=np.float64)
    s2 = np.ascontiguousarray(s2, dtype=np.float64)
    N.check(h.omc_batch_results(C.byref(prob.ph), s.ctypes.data_as(_dp), s2.ctypes.data_as(_dp),
                                nbatch or prob.nbatch, float(fluence or prob.nperbatch), 1,
                                dose.ctypes.data_as(_dp), unc.ctypes.data_as(_dp)))
    return dose, unc
