"""ctypes binding of libomc_hip.so (include/omc_dev.h) for tests and the
benchmark.  Loading fails loudly when the HIP library has not been built;
there is no CPU fallback."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N

FP64, FP32 = 0, 1
_dp = C.POINTER(C.c_double)

# entries per history of the ledger (OMC_LEDGER): deposited in the phantom,
# deposited in region 0, carried out by photons, sum(energy * region)
LEDGER = 4


class DevOpts(C.Structure):
    _fields_ = [("device", C.c_int), ("precision", C.c_int), ("hbm_budget", C.c_int64), ("chunk", C.c_int64),
                ("pipelines", C.c_int), ("yield_trips", C.c_int), ("drain_lanes", C.c_int),
                ("drain_trips", C.c_int), ("event_div", C.c_int), ("fetch_min", C.c_int),
                ("drain_min", C.c_int64), ("woodcock", C.c_int), ("tally_bytes", C.c_int64),
                ("sort_min", C.c_int64)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class DevStats(C.Structure):
    _fields_ = [("histories", C.c_int64), ("photon_steps", C.c_int64), ("electron_steps", C.c_int64),
                ("deposits", C.c_int64), ("photons_pushed", C.c_int64), ("electrons_pushed", C.c_int64),
                ("generations", C.c_int64), ("injections", C.c_int64), ("transport_ms", C.c_double),
                ("photon_kernel_ms", C.c_double), ("electron_kernel_ms", C.c_double),
                ("source_kernel_ms", C.c_double), ("photon_launches", C.c_int64),
                ("electron_launches", C.c_int64), ("killed", C.c_int64), ("photon_crossings", C.c_int64),
                ("electron_crossings", C.c_int64), ("photon_deposits", C.c_int64),
                ("electron_deposits", C.c_int64), ("yields", C.c_int64), ("host_ms", C.c_double),
                ("pipe_idle_ms", C.c_double), ("woodcock_over", C.c_int64)] + [
                (n, C.c_int64) for n in ("compton", "photo", "pair", "rayleigh", "brems", "moller", "bhabha",
                                         "annih_flight", "annih_rest", "fict_rejections", "msdist", "sscat",
                                         "howfar_charged", "longest_photon_trips", "longest_electron_trips")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class RunParams(C.Structure):
    _fields_ = [("phantom_id", C.c_int), ("nsplit", C.c_int), ("seed", C.c_uint32 * 2), ("hist0", C.c_int64),
                ("nperbatch", C.c_int64), ("nbatch", C.c_int), ("ledger", _dp),
                ("stats", C.POINTER(DevStats))]


class Request(C.Structure):
    _fields_ = [("phantom_id", C.c_int), ("kind", C.c_int), ("src_dosxyz", C.POINTER(N.SourceDosxyz)),
                ("src_beamlets", C.POINTER(N.SourceBeamlets)), ("beamlet", C.c_int), ("nhist", C.c_int64),
                ("nbatch", C.c_int), ("seed", C.c_uint32 * 2)]


class Result(C.Structure):
    _fields_ = [("ticket", C.c_int), ("done", C.c_int), ("status", C.c_int), ("submit_ms", C.c_double),
                ("admit_ms", C.c_double), ("finish_ms", C.c_double), ("nhist", C.c_int64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = N.LIBDIR / "libomc_hip.so"
        if not path.exists():
            raise RuntimeError(f"{path} is not built; the device path has no fallback")
        N.host()
        L = C.CDLL(str(path))
        L.omc_dev_last_error.restype = C.c_char_p
        L.omc_dev_create.argtypes = [C.POINTER(C.c_void_p), C.POINTER(DevOpts)]
        L.omc_dev_destroy.argtypes = [C.c_void_p]
        L.omc_dev_upload_media.argtypes = [C.c_void_p, C.POINTER(N.Tables)]
        L.omc_dev_upload_phantom.argtypes = [C.c_void_p, C.POINTER(N.Phantom), C.POINTER(N.Regions),
                                             C.POINTER(C.c_int)]
        L.omc_dev_release_phantom.argtypes = [C.c_void_p, C.c_int]
        L.omc_dev_run_dosxyz.argtypes = [C.c_void_p, C.POINTER(N.SourceDosxyz), C.POINTER(RunParams), _dp, _dp]
        L.omc_dev_tune.argtypes = [C.c_void_p, C.POINTER(N.SourceDosxyz), C.POINTER(RunParams), C.c_int64,
                                   C.POINTER(DevOpts), C.POINTER(C.c_double)]
        L.omc_dev_math_eval.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int64, _dp, _dp, _dp]
        L.omc_dev_build_photon.argtypes = [C.c_int, C.POINTER(N.Tables), C.c_char_p]
        L.omc_dev_tune_beamlets.argtypes = [C.c_void_p, C.POINTER(N.SourceBeamlets), C.POINTER(C.c_int64),
                                            C.POINTER(RunParams), C.POINTER(DevOpts), C.POINTER(C.c_double)]
        _i64p = C.POINTER(C.c_int64)
        L.omc_dev_run_beamlets.argtypes = [C.c_void_p, C.POINTER(N.SourceBeamlets), _i64p, C.POINTER(RunParams),
                                           _dp, _dp]
        L.omc_dev_beamlets_sparse.argtypes = [C.c_void_p, C.POINTER(N.SourceBeamlets), _i64p,
                                              C.POINTER(RunParams), C.c_double, _i64p, _i64p, _dp, C.c_int64,
                                              _i64p]
        L.omc_dev_submit.argtypes = [C.c_void_p, C.POINTER(Request)]
        L.omc_dev_step.argtypes = [C.c_void_p]
        L.omc_dev_poll.argtypes = [C.c_void_p, C.c_int, C.POINTER(Result), _dp, _dp]
        _lib = L
    return _lib


def check(rc):
    if rc < 0:
        raise RuntimeError(f"omc_dev error {rc}: {lib().omc_dev_last_error().decode()}")
    return rc


def build_photon(tables, data_folder, device=0):
    """omc_dev_build_photon: the photon tables of `tables` (an N.Tables with
    its pegs4 fields filled and its photon arrays allocated), computed on
    the GPU from the XCOM data in data_folder."""
    folder = str(data_folder)
    if not folder.endswith("/"):
        folder += "/"
    check(lib().omc_dev_build_photon(device, C.byref(tables), folder.encode()))


class Engine:
    """One device context: tables + resident phantoms."""

    def __init__(self, precision=FP64, device=0, hbm_budget=0, chunk=0, pipelines=0, yield_trips=0,
                 drain_lanes=0, drain_trips=0, event_div=0, fetch_min=0, drain_min=0, woodcock=0,
                 tally_bytes=0, sort_min=0):
        L = lib()
        self.ctx = C.c_void_p()
        self.nreg = {}
        opts = DevOpts(device, precision, hbm_budget, chunk, pipelines, yield_trips, drain_lanes, drain_trips,
                       event_div, fetch_min, drain_min, woodcock, tally_bytes, sort_min)
        check(L.omc_dev_create(C.byref(self.ctx), C.byref(opts)))
        self.precision = precision

    def close(self):
        if self.ctx:
            lib().omc_dev_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload_media(self, tables):
        check(lib().omc_dev_upload_media(self.ctx, C.byref(tables)))

    def upload_phantom(self, phantom, regions) -> int:
        pid = C.c_int(-1)
        check(lib().omc_dev_upload_phantom(self.ctx, C.byref(phantom), C.byref(regions), C.byref(pid)))
        self.nreg[pid.value] = regions.nreg
        return pid.value

    def release_phantom(self, pid):
        """omc_dev_release_phantom: frees a resident phantom no work uses."""
        check(lib().omc_dev_release_phantom(self.ctx, int(pid)))
        self.nreg.pop(pid, None)

    def run_dosxyz(self, pid, src, nperbatch, nbatch, seed=(97, 33), hist0=0, nreg=None, ledger=False, nsplit=1):
        stats = DevStats()
        led = np.zeros(LEDGER * nperbatch * nbatch) if ledger else None
        rp = RunParams(pid, nsplit, (C.c_uint32 * 2)(*seed), hist0, nperbatch, nbatch,
                       led.ctypes.data_as(_dp) if ledger else None, C.pointer(stats))
        s = np.zeros(nreg)
        s2 = np.zeros(nreg)
        check(lib().omc_dev_run_dosxyz(self.ctx, C.byref(src), C.byref(rp), s.ctypes.data_as(_dp),
                                       s2.ctypes.data_as(_dp)))
        return s, s2, stats, led

    def tune(self, pid, src, probe, nbatch=10, seed=(97, 33)):
        """omc_dev_tune: time the dosxyz loop of `src` on `probe` histories
        under each candidate setting; the engine keeps the fastest.  Returns
        (settings dict, histories/s)."""
        rp = RunParams(pid, 1, (C.c_uint32 * 2)(*seed), 0, 0, nbatch, None, None)
        best = DevOpts()
        hps = C.c_double(0)
        check(lib().omc_dev_tune(self.ctx, C.byref(src), C.byref(rp), int(probe), C.byref(best), C.byref(hps)))
        return best.as_dict(), hps.value

    def tune_beamlets(self, pid, src, nhist, nbatch=10, seed=(97, 33)):
        """omc_dev_tune_beamlets: the autotuner on a beamlet batch (nhist[b]
        histories for beamlet b).  Returns (settings dict, histories/s)."""
        nh = np.ascontiguousarray(nhist, dtype=np.int64)
        rp = RunParams(pid, 1, (C.c_uint32 * 2)(*seed), 0, 0, nbatch, None, None)
        best = DevOpts()
        hps = C.c_double(0)
        check(lib().omc_dev_tune_beamlets(self.ctx, C.byref(src), nh.ctypes.data_as(C.POINTER(C.c_int64)),
                                          C.byref(rp), C.byref(best), C.byref(hps)))
        return best.as_dict(), hps.value

    def run_beamlets(self, pid, src, nhist, nbatch, seed=(97, 33), hist0=0, nsplit=1, ledger=False):
        """Batch sums of every beamlet: arrays (nbeamlets, nreg); with
        ledger=True also the per-history ledger (LEDGER values per history,
        beamlet after beamlet) as a fourth value."""
        nreg = self.nreg[pid]
        nh = np.ascontiguousarray(nhist, dtype=np.int64)
        stats = DevStats()
        nrun = int(sum(max(1, int(n) // nbatch) * nbatch for n in nh))
        led = np.zeros(LEDGER * nrun) if ledger else None
        rp = RunParams(pid, nsplit, (C.c_uint32 * 2)(*seed), hist0, 0, nbatch,
                       led.ctypes.data_as(_dp) if ledger else None, C.pointer(stats))
        s = np.zeros((src.nbeamlets, nreg))
        s2 = np.zeros((src.nbeamlets, nreg))
        check(lib().omc_dev_run_beamlets(self.ctx, C.byref(src), nh.ctypes.data_as(C.POINTER(C.c_int64)),
                                         C.byref(rp), s.ctypes.data_as(_dp), s2.ctypes.data_as(_dp)))
        return (s, s2, stats, led) if ledger else (s, s2, stats)

    def beamlets_sparse(self, pid, src, nhist, nbatch, rel_thr, seed=(97, 33), hist0=0, cap=None, nsplit=1):
        """matRad Dij as CSC arrays (jcs, irs, vals) plus the run statistics."""
        nh = np.ascontiguousarray(nhist, dtype=np.int64)
        stats = DevStats()
        rp = RunParams(pid, nsplit, (C.c_uint32 * 2)(*seed), hist0, 0, nbatch, None, C.pointer(stats))
        P64 = C.POINTER(C.c_int64)
        nb = src.nbeamlets
        cap = int(cap) if cap is not None else max(1, (self.nreg[pid] - 1) * nb // 50)
        first = True
        while True:
            jcs = np.zeros(nb + 1, np.int64)
            irs = np.zeros(cap, np.int64)
            vals = np.zeros(cap)
            nnz = C.c_int64(0)
            rc = lib().omc_dev_beamlets_sparse(self.ctx, C.byref(src) if first else None,
                                               nh.ctypes.data_as(P64), C.byref(rp), rel_thr,
                                               jcs.ctypes.data_as(P64), irs.ctypes.data_as(P64),
                                               vals.ctypes.data_as(_dp), cap, C.byref(nnz))
            if rc == -6 and nnz.value > cap:          # OMC_ERANGE: collect the kept matrix
                cap = nnz.value
                first = False
                continue
            check(rc)
            return jcs, irs[:nnz.value], vals[:nnz.value], stats

    # ---- serving -------------------------------------------------------
    def submit(self, pid, src, nhist, nbatch=10, seed=(97, 33), kind=0, beamlet=0):
        req = Request(pid, kind, C.pointer(src) if kind == 0 else None, C.pointer(src) if kind == 1 else None,
                      beamlet, int(nhist), nbatch, (C.c_uint32 * 2)(*seed))
        return check(lib().omc_dev_submit(self.ctx, C.byref(req)))

    def step(self):
        return check(lib().omc_dev_step(self.ctx))

    def poll(self, ticket, nreg=None):
        """(Result, sum, sum2) -- sums are None until the ticket is done."""
        res = Result()
        if nreg is None:
            check(lib().omc_dev_poll(self.ctx, ticket, C.byref(res), None, None))
            return res, None, None
        s = np.empty(nreg)        # written in full once the ticket is done
        s2 = np.empty(nreg)
        check(lib().omc_dev_poll(self.ctx, ticket, C.byref(res), s.ctypes.data_as(_dp), s2.ctypes.data_as(_dp)))
        return (res, s, s2) if res.done and res.status == 0 else (res, None, None)


MATH = {"log": 0, "exp": 1, "sqrt": 2, "sin": 3, "cos": 4, "pow": 5}


def math_eval(fn, x, y=None, precision=FP64, device=0):
    """omc_dev_math_eval: the transport's elementary function `fn` on the
    device, for the arguments x (and y for pow)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    yy = np.ascontiguousarray(y if y is not None else x, dtype=np.float64)
    out = np.empty_like(x)
    check(lib().omc_dev_math_eval(device, precision, MATH[fn], x.size, x.ctypes.data_as(_dp),
                                  yy.ctypes.data_as(_dp), out.ctypes.data_as(_dp)))
    return out

