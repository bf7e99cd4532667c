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
from ompmc_amd.problem import DosxyzProblem, free_problem, load_problem  # noqa: F401  (host loader)

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
    _fields_ = [(n, C.c_int64) for n in (
        "n_hist", "n_dep", "n_cross_photon", "n_cross_charged", "n_draws",
        "n_compton", "n_photo", "n_pair", "n_rayleigh", "n_brems", "n_moller", "n_bhabha", "n_annih",
        "n_rannih", "n_fict", "n_msdist", "n_sscat", "n_howfar_charged", "n_photon_steps",
        "n_electron_steps")]


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
        L.ora_beamlet_primary.argtypes = [C.c_void_p, C.POINTER(N.SourceBeamlets), C.c_int, C.c_int64,
                                          C.POINTER(Particle)]
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


def run_dosxyz_ranmar_runs(prob: DosxyzProblem, nruns, nperbatch, nbatch, threads=None):
    """`nruns` independent serial RANMAR runs of the problem (the reference's
    own generator and history loop), each with seeds (ixx, jxx + k), run on
    up to `threads` host threads at once; their batches are pooled.  Returns
    (sum, sum2, total batches) over all runs."""
    import concurrent.futures as cf
    import os
    L = lib()
    nreg = prob.reg.nreg
    threads = threads or min(nruns, len(os.sched_getaffinity(0)), 16)

    def one(k):
        ctx = L.ora_create(C.byref(prob.tab), C.byref(prob.reg), C.byref(prob.ph), prob.nsplit, RANMAR)
        try:
            L.ora_seed_ranmar(ctx, prob.seeds[0], (prob.seeds[1] + k) % 30081)
            s = np.zeros(nreg)
            s2 = np.zeros(nreg)
            rc = L.ora_run_dosxyz(ctx, C.byref(prob.src), nperbatch, nbatch, s.ctypes.data_as(_dp),
                                  s2.ctypes.data_as(_dp))
            if rc:
                raise RuntimeError(f"oracle run failed {rc}")
            return s, s2
        finally:
            L.ora_destroy(ctx)

    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        res = list(ex.map(one, range(nruns)))
    return sum(r[0] for r in res), sum(r[1] for r in res), nruns * nbatch


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


def run_beamlet_philox(prob: DosxyzProblem, src, beamlet, nperbatch, nbatch, seed=(97, 33), hist0=0):
    """matRad beamlet loop for one beamlet (omc_matrad.c:1389-1414) in
    PHILOX mode: histories hist0 .. hist0 + nperbatch*nbatch - 1; returns the
    batch sums (sum, sum2) over all regions."""
    L = lib()
    ctx = L.ora_create(C.byref(prob.tab), C.byref(prob.reg), C.byref(prob.ph), 1, PHILOX)
    try:
        L.ora_seed_philox(ctx, seed[0], seed[1])
        nreg = prob.reg.nreg
        s = np.zeros(nreg)
        s2 = np.zeros(nreg)
        rc = L.ora_run_beamlet(ctx, C.byref(src), beamlet, hist0, nperbatch, nbatch, s.ctypes.data_as(_dp),
                               s2.ctypes.data_as(_dp))
        if rc:
            raise RuntimeError(f"oracle beamlet run failed {rc}")
        return s, s2
    finally:
        L.ora_destroy(ctx)


def matrad_column(xb, yb, zb, rho, s, nbatch, nhist, rel_thr):
    """One column of the matRad dose influence matrix from a beamlet's batch
    sums `s` ([nvox+1], region 0 first): accumulateResults with iout = 1
    (omc_matrad.c:839-918: mean over batches, Gy per incident history,
    zero below density 0.044), then the threshold at rel_thr * max and the
    rows in voxel order (omc_matrad.c:1420-1479).  Same operation order as
    the reference, element by element.  Returns (rows, values)."""
    xb, yb, zb = (np.asarray(a, dtype=np.float64) for a in (xb, yb, zb))
    ni, nj, nk = len(xb) - 1, len(yb) - 1, len(zb) - 1
    rho = np.asarray(rho, dtype=np.float64).reshape(nk, nj, ni)
    e = np.asarray(s, dtype=np.float64)[1:].reshape(nk, nj, ni) / float(nbatch)
    dx, dy, dz = np.diff(xb), np.diff(yb), np.diff(zb)
    mass = (dx[None, None, :] * dy[None, :, None]) * dz[:, None, None]
    mass = mass * rho
    e = e * (1.602E-10 / (mass * float(nhist)))
    e = np.where(rho < 0.044, 0.0, e).reshape(-1)
    dmax = 0.0
    if e.size and e.max() > dmax:
        dmax = float(e.max())
    thresh = dmax * rel_thr
    rows = np.nonzero(e > thresh)[0].astype(np.int64)
    return rows, e[rows]


def beamlet_primaries(prob: DosxyzProblem, src, beamlet, hids, seed=(97, 33)):
    """Primary particles of the given histories of one beamlet (PHILOX mode)."""
    L = lib()
    ctx = L.ora_create(C.byref(prob.tab), C.byref(prob.reg), C.byref(prob.ph), 1, PHILOX)
    try:
        L.ora_seed_philox(ctx, seed[0], seed[1])
        out = []
        for h in hids:
            p = Particle()
            if L.ora_beamlet_primary(ctx, C.byref(src), beamlet, int(h), C.byref(p)):
                raise RuntimeError("bad beamlet")
            out.append(p)
        return out
    finally:
        L.ora_destroy(ctx)
