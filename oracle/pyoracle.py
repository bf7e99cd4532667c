"""ctypes binding of the CPU oracle (oracle/build/liboracle.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

from ompmc_amd import _native as N

HERE = Path(__file__).resolve().parent
LIB = HERE / "build" / "liboracle.so"
REF_DIR = HERE / "_ref"

RANMAR, PHILOX, TAPE = 0, 1, 2
_dp = C.POINTER(C.c_double)


class Stream(C.Structure):
    _fields_ = [("k0", C.c_uint32), ("k1", C.c_uint32), ("ctr", C.c_uint32)]


class Particle(C.Structure):
    _fields_ = [("iq", C.c_int), ("ir", C.c_int), ("e", C.c_double), ("x", C.c_double),
                ("y", C.c_double), ("z", C.c_double), ("u", C.c_double), ("v", C.c_double),
                ("w", C.c_double), ("dnear", C.c_double), ("wt", C.c_double), ("rs", Stream)]


class Counters(C.Structure):
    _fields_ = [("n_hist", C.c_int64), ("n_dep", C.c_int64), ("n_cross_photon", C.c_int64),
                ("n_cross_charged", C.c_int64), ("n_draws", C.c_int64)]


class DosxyzProblem(C.Structure):
    _fields_ = [("inp_n", C.c_int), ("inp_key", (C.c_char * 256) * 80), ("inp_value", (C.c_char * 256) * 80),
                ("ph", N.Phantom), ("tab", N.Tables), ("reg", N.Regions), ("src", N.SourceDosxyz),
                ("nsplit", C.c_int), ("seeds", C.c_int * 2), ("nhist", C.c_int), ("nbatch", C.c_int),
                ("nperbatch", C.c_int), ("output_folder", C.c_char * 256)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} not built (make -C oracle)")
        N.host()
        L = C.CDLL(str(LIB))
        L.ora_create.restype = C.c_void_p
        L.ora_create.argtypes = [C.POINTER(N.Tables), C.POINTER(N.Regions), C.POINTER(N.Phantom), C.c_int, C.c_int]
        L.ora_destroy.argtypes = [C.c_void_p]
        L.ora_seed_ranmar.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.ora_seed_philox.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32]
        L.ora_set_tape.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.c_int64]
        L.ora_tape_pos.argtypes = [C.c_void_p]
        L.ora_tape_pos.restype = C.c_int64
        L.ora_random.argtypes = [C.c_void_p]
        L.ora_random.restype = C.c_double
        L.ora_get_counters.argtypes = [C.c_void_p, C.POINTER(Counters)]
        L.ora_set_ledger.argtypes = [C.c_void_p, _dp, C.c_int64]
        L.ora_run_dosxyz.argtypes = [C.c_void_p, C.POINTER(N.SourceDosxyz), C.c_int, C.c_int, _dp, _dp]
        L.ora_run_dosxyz_parallel.argtypes = [C.POINTER(N.Tables), C.POINTER(N.Regions), C.POINTER(N.Phantom),
                                              C.POINTER(N.SourceDosxyz), C.c_uint32, C.c_uint32, C.c_int64,
                                              C.c_int, C.c_int, C.c_int, _dp, _dp, C.POINTER(Counters)]
        L.ora_run_beamlet.argtypes = [C.c_void_p, C.POINTER(N.SourceBeamlets), C.c_int, C.c_int64,
                                      C.c_int, C.c_int, _dp, _dp]
        L.ora_howfar.argtypes = [C.c_void_p, C.POINTER(Particle), C.POINTER(C.c_int), C.POINTER(C.c_int), _dp]
        L.ora_hownear.argtypes = [C.c_void_p, C.POINTER(Particle)]
        L.ora_hownear.restype = C.c_double
        L.ora_compute_drange.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                         C.c_double, C.c_double]
        L.ora_compute_drange.restype = C.c_double
        L.ora_compute_eloss.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                        C.c_double, C.c_double, C.c_int]
        L.ora_compute_eloss.restype = C.c_double
        L.ora_msdist.argtypes = [C.c_void_p, C.POINTER(Particle), C.c_int, C.c_double, C.c_double, C.c_double,
                                 C.c_double, C.POINTER(Particle)]
        L.ora_msdist.restype = C.c_double
        L.ora_interaction.argtypes = [C.c_void_p, C.c_int, C.POINTER(Particle), C.POINTER(Particle)]
        L.ora_shower_one.argtypes = [C.c_void_p, C.POINTER(Particle), _dp]
        L.ora_philox4x32.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        _lib = L
    return _lib


def load_problem(stem) -> DosxyzProblem:
    h = N.host()
    h.omc_dosxyz_problem_load.argtypes = [C.POINTER(DosxyzProblem), C.c_char_p]
    p = DosxyzProblem()
    N.check(h.omc_dosxyz_problem_load(C.byref(p), str(stem).encode()))
    return p


def free_problem(p: DosxyzProblem):
    h = N.host()
    h.omc_dosxyz_problem_free.argtypes = [C.POINTER(DosxyzProblem)]
    h.omc_dosxyz_problem_free(C.byref(p))


def run_dosxyz_ranmar(prob: DosxyzProblem, nperbatch=None, nbatch=None):
    """Serial RANMAR run with the problem's seeds: returns (sum, sum2, counters)."""
    L = lib()
    ctx = L.ora_create(C.byref(prob.tab), C.byref(prob.reg), C.byref(prob.ph), prob.nsplit, RANMAR)
    try:
        L.ora_seed_ranmar(ctx, prob.seeds[0], prob.seeds[1])
        nreg = prob.reg.nreg
        s = np.zeros(nreg)
        s2 = np.zeros(nreg)
        rc = L.ora_run_dosxyz(ctx, C.byref(prob.src), nperbatch or prob.nperbatch, nbatch or prob.nbatch,
                              s.ctypes.data_as(_dp), s2.ctypes.data_as(_dp))
        if rc:
            raise RuntimeError(f"oracle run failed {rc}")
        cnt = Counters()
        L.ora_get_counters(ctx, C.byref(cnt))
        return s, s2, cnt
    finally:
        L.ora_destroy(ctx)


def run_dosxyz_philox(prob: DosxyzProblem, seed=(97, 33), nperbatch=None, nbatch=None, nthreads=0,
                      hist0=0):
    L = lib()
    nreg = prob.reg.nreg
    s = np.zeros(nreg)
    s2 = np.zeros(nreg)
    cnt = Counters()
    rc = L.ora_run_dosxyz_parallel(C.byref(prob.tab), C.byref(prob.reg), C.byref(prob.ph), C.byref(prob.src),
                                   seed[0], seed[1], hist0, nperbatch or prob.nperbatch, nbatch or prob.nbatch,
                                   nthreads, s.ctypes.data_as(_dp), s2.ctypes.data_as(_dp), C.byref(cnt))
    if rc:
        raise RuntimeError(f"oracle run failed {rc}")
    return s, s2, cnt


def dose_from_sums(prob: DosxyzProblem, s, s2, nbatch=None, nperbatch=None):
    h = N.host()
    nvox = prob.ph.isize * prob.ph.jsize * prob.ph.ksize
    dose = np.zeros(nvox)
    unc = np.zeros(nvox)
    N.check(h.omc_batch_results(C.byref(prob.ph), s.ctypes.data_as(_dp), s2.ctypes.data_as(_dp),
                                nbatch or prob.nbatch, float(nperbatch or prob.nperbatch), 1,
                                dose.ctypes.data_as(_dp), unc.ctypes.data_as(_dp)))
    return dose, unc


def write_3ddose(prob: DosxyzProblem, path, dose, unc):
    N.check(N.host().omc_write_3ddose(str(path).encode(), C.byref(prob.ph), dose.ctypes.data_as(_dp),
                                      unc.ctypes.data_as(_dp)))


def ref_available() -> bool:
    return (REF_DIR / "omc_dosxyz_ref").exists()


def run_reference(stem, out_name, omp=False, threads=None, timeout=600):
    """Runs the reference omc_dosxyz built in oracle/_ref; returns stdout."""
    exe = REF_DIR / ("omc_dosxyz_ref_omp" if omp else "omc_dosxyz_ref")
    cmd = [str(exe), "-i", str(stem), "-o", out_name]
    if omp and threads:
        cmd = ["taskset", "-c", f"0-{threads - 1}"] + cmd
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"reference failed: {r.stdout[-2000:]} {r.stderr[-2000:]}")
    return r.stdout + r.stderr
