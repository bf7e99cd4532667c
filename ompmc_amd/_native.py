"""ctypes mirror of include/omc_host.h and include/omc_dev.h.

The C libraries are the product; this module only loads them for the Python
test and benchmark harness.  Layouts here must match the headers exactly.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIBDIR = PKG / "lib"
DATA = PKG / "data"

MXGE, MXEKE, MXMED, MXELEMENT = 2000, 500, 9, 50
MXRAYFF, RAYCDFSIZE = 100, 100
SPIN_PER_MED = 2 * 32 * 16 * 32
MS_SIZE = 64 * 8 * 32
INVDIM = 1000

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class Element(C.Structure):
    _fields_ = [("symbol", C.c_char * 3), ("z", C.c_double), ("wa", C.c_double),
                ("pz", C.c_double), ("rhoz", C.c_double)]


def _arr(t, n):
    return t * n


class Tables(C.Structure):
    _fields_ = [
        ("nmed", C.c_int), ("names", (C.c_char * 64) * MXMED),
        ("ne", _arr(C.c_int, MXMED)), ("iunrst", _arr(C.c_int, MXMED)),
        ("epstfl", _arr(C.c_int, MXMED)), ("iaprim", _arr(C.c_int, MXMED)),
        ("msge", _arr(C.c_int, MXMED)), ("mge", _arr(C.c_int, MXMED)),
        ("mseke", _arr(C.c_int, MXMED)), ("meke", _arr(C.c_int, MXMED)),
        ("mleke", _arr(C.c_int, MXMED)), ("mcmfp", _arr(C.c_int, MXMED)),
        ("mrange", _arr(C.c_int, MXMED)),
        ("rho", _arr(C.c_double, MXMED)), ("rlc", _arr(C.c_double, MXMED)),
        ("ae", _arr(C.c_double, MXMED)), ("ap", _arr(C.c_double, MXMED)),
        ("ue", _arr(C.c_double, MXMED)), ("up", _arr(C.c_double, MXMED)),
        ("te", _arr(C.c_double, MXMED)), ("thmoll", _arr(C.c_double, MXMED)),
        ("pegs_delcm", _arr(C.c_double, MXMED)),
        ("elements", (Element * MXELEMENT) * MXMED),
        ("ge0", _arr(C.c_double, MXMED)), ("ge1", _arr(C.c_double, MXMED)),
        ("gmfp0", _dp), ("gmfp1", _dp), ("gbr10", _dp), ("gbr11", _dp),
        ("gbr20", _dp), ("gbr21", _dp), ("cohe0", _dp), ("cohe1", _dp),
        ("xgrid", _dp), ("fcum", _dp), ("b_array", _dp), ("c_array", _dp),
        ("i_array", _ip), ("pmax0", _dp), ("pmax1", _dp),
        ("dl1", _arr(C.c_double, MXMED * 8)), ("dl2", _arr(C.c_double, MXMED * 8)),
        ("dl3", _arr(C.c_double, MXMED * 8)), ("dl4", _arr(C.c_double, MXMED * 8)),
        ("dl5", _arr(C.c_double, MXMED * 8)), ("dl6", _arr(C.c_double, MXMED * 8)),
        ("bpar0", _arr(C.c_double, MXMED)), ("bpar1", _arr(C.c_double, MXMED)),
        ("delcm", _arr(C.c_double, MXMED)), ("zbrang", _arr(C.c_double, MXMED)),
        ("eke0", _arr(C.c_double, MXMED)), ("eke1", _arr(C.c_double, MXMED)),
        ("blcc", _arr(C.c_double, MXMED)), ("xcc", _arr(C.c_double, MXMED)),
        ("esig_e", _arr(C.c_double, MXMED)), ("psig_e", _arr(C.c_double, MXMED)),
        ("expeke1", _arr(C.c_double, MXMED)),
        ("sig_ismonotone", _arr(C.c_int, MXMED) * 2),
        ("esig0", _dp), ("esig1", _dp), ("psig0", _dp), ("psig1", _dp),
        ("ededx0", _dp), ("ededx1", _dp), ("pdedx0", _dp), ("pdedx1", _dp),
        ("ebr10", _dp), ("ebr11", _dp), ("pbr10", _dp), ("pbr11", _dp),
        ("pbr20", _dp), ("pbr21", _dp), ("tmxs0", _dp), ("tmxs1", _dp),
        ("etae_ms0", _dp), ("etae_ms1", _dp), ("etap_ms0", _dp), ("etap_ms1", _dp),
        ("q1ce_ms0", _dp), ("q1ce_ms1", _dp), ("q1cp_ms0", _dp), ("q1cp_ms1", _dp),
        ("q2ce_ms0", _dp), ("q2ce_ms1", _dp), ("q2cp_ms0", _dp), ("q2cp_ms1", _dp),
        ("blcce0", _dp), ("blcce1", _dp), ("range_ep", _dp), ("e_array", _dp),
        ("b2spin_min", C.c_double), ("dbeta2i", C.c_double), ("espml", C.c_double),
        ("dleneri", C.c_double), ("dqq1i", C.c_double), ("spin_rej", _dp),
        ("ums", _dp), ("fms", _dp), ("wms", _dp), ("ims", _ip),
        ("dllambi", C.c_double), ("dqmsi", C.c_double),
    ]


class DataPaths(C.Structure):
    _fields_ = [("pegs_file", C.c_char_p), ("pgs4form_file", C.c_char_p),
                ("data_folder", C.c_char_p)]


class Phantom(C.Structure):
    _fields_ = [("nmed", C.c_int), ("names", (C.c_char * 64) * MXMED),
                ("isize", C.c_int), ("jsize", C.c_int), ("ksize", C.c_int),
                ("xbounds", _dp), ("ybounds", _dp), ("zbounds", _dp),
                ("med", _ip), ("rho", _dp)]


class Regions(C.Structure):
    _fields_ = [("nreg", C.c_int), ("med", _ip), ("rhof", _dp), ("pcut", _dp), ("ecut", _dp)]


class Spectrum(C.Structure):
    _fields_ = [("spectrum", C.c_int), ("energy", C.c_double), ("deltak", C.c_double),
                ("cdfinv1", C.c_double * INVDIM), ("cdfinv2", C.c_double * INVDIM)]


class SourceDosxyz(C.Structure):
    _fields_ = [("spec", Spectrum), ("charge", C.c_int), ("ssd", C.c_double),
                ("xinl", C.c_double), ("xinu", C.c_double), ("yinl", C.c_double),
                ("yinu", C.c_double), ("xsize", C.c_double), ("ysize", C.c_double),
                ("ixinl", C.c_int), ("ixinu", C.c_int), ("iyinl", C.c_int), ("iyinu", C.c_int)]


class SourceBeamlets(C.Structure):
    _fields_ = [("spec", Spectrum), ("charge", C.c_int), ("nbeamlets", C.c_int),
                ("nbeams", C.c_int), ("ibeam", _ip), ("src", _dp), ("corner", _dp),
                ("side1", _dp), ("side2", _dp)]


class DoseComparison(C.Structure):
    _fields_ = [("voxels", C.c_int64), ("frac_within_2sigma", C.c_double), ("z_mean", C.c_double),
                ("z_rms", C.c_double), ("z_max_abs", C.c_double), ("integral_ratio", C.c_double),
                ("pass_", C.c_int)]

    def as_dict(self):
        return {("pass" if k == "pass_" else k): getattr(self, k) for k, _ in self._fields_}


def compare_3ddose(path_a, path_b, phantom=None, rel_thr=0.1):
    """omc_compare_3ddose: equal-accuracy statistics of two .3ddose files."""
    r = DoseComparison()
    check(host().omc_compare_3ddose(str(path_a).encode(), str(path_b).encode(),
                                    C.byref(phantom) if phantom is not None else None, rel_thr, C.byref(r)))
    return r.as_dict()


def beamlet_source(spec, ibeam, src, corner, side1, side2, charge=0):
    """omc_source_beamlets over numpy arrays; the arrays are kept alive on
    the returned structure."""
    arrs = dict(ibeam=np.ascontiguousarray(ibeam, dtype=np.int32),
                src=np.ascontiguousarray(src, dtype=np.float64).reshape(-1),
                corner=np.ascontiguousarray(corner, dtype=np.float64).reshape(-1),
                side1=np.ascontiguousarray(side1, dtype=np.float64).reshape(-1),
                side2=np.ascontiguousarray(side2, dtype=np.float64).reshape(-1))
    s = SourceBeamlets()
    s.spec = spec
    s.charge = charge
    s.nbeamlets = len(arrs["ibeam"])
    s.nbeams = len(arrs["src"]) // 3
    s.ibeam = arrs["ibeam"].ctypes.data_as(_ip)
    for k in ("src", "corner", "side1", "side2"):
        setattr(s, k, arrs[k].ctypes.data_as(_dp))
    s._keep = arrs
    return s


_host = None


def host():
    """Loads lib/libomc_host.so (raises if it has not been built)."""
    global _host
    if _host is None:
        path = LIBDIR / "libomc_host.so"
        if not path.exists():
            raise RuntimeError(f"{path} not built: run __graft_entry__.build()")
        lib = C.CDLL(str(path))
        lib.omc_last_error.restype = C.c_char_p
        lib.omc_tables_build.argtypes = [C.POINTER(Tables), C.c_int, C.c_void_p, C.POINTER(DataPaths)]
        lib.omc_tables_free.argtypes = [C.POINTER(Tables)]
        lib.omc_write_synthetic_spinms.argtypes = [C.c_char_p]
        lib.omc_kn_sigma0.argtypes = [C.c_double]
        lib.omc_kn_sigma0.restype = C.c_double
        lib.omc_spline_setup.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, C.c_int]
        lib.omc_spline_eval.argtypes = [C.c_double, _dp, _dp, _dp, _dp, _dp, C.c_int]
        lib.omc_spline_eval.restype = C.c_double
        lib.omc_phantom_read.argtypes = [C.POINTER(Phantom), C.c_char_p]
        lib.omc_phantom_write.argtypes = [C.POINTER(Phantom), C.c_char_p]
        lib.omc_phantom_alloc.argtypes = [C.POINTER(Phantom), C.c_int, C.c_int, C.c_int, C.c_int]
        lib.omc_phantom_free.argtypes = [C.POINTER(Phantom)]
        lib.omc_regions_build.argtypes = [C.POINTER(Regions), C.POINTER(Phantom), C.POINTER(Tables),
                                          C.c_double, C.c_double]
        lib.omc_regions_free.argtypes = [C.POINTER(Regions)]
        lib.omc_spectrum_read.argtypes = [C.POINTER(Spectrum), C.c_char_p]
        lib.omc_spectrum_from_bins.argtypes = [C.POINTER(Spectrum), C.c_int, C.c_double, _dp, _dp, C.c_int]
        lib.omc_spectrum_mono.argtypes = [C.POINTER(Spectrum), C.c_double]
        lib.omc_source_dosxyz_setup.argtypes = [C.POINTER(SourceDosxyz), C.POINTER(Phantom), C.c_double,
                                                C.c_double, C.c_double, C.c_double, C.c_double, C.c_int]
        lib.omc_batch_results.argtypes = [C.POINTER(Phantom), _dp, _dp, C.c_int, C.c_double, C.c_int, _dp, _dp]
        lib.omc_write_3ddose.argtypes = [C.c_char_p, C.POINTER(Phantom), _dp, _dp]
        lib.omc_compare_doses.argtypes = [C.c_int64, _dp, _dp, _dp, _dp, _dp, C.c_double,
                                          C.POINTER(DoseComparison)]
        lib.omc_compare_3ddose.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(Phantom), C.c_double,
                                           C.POINTER(DoseComparison)]
        lib.omc_sparse_column.argtypes = [_dp, C.c_int64, C.c_double, C.POINTER(C.c_int64), _dp, C.c_int64]
        lib.omc_sparse_column.restype = C.c_int64
        lib.omc_inp_parse_text.argtypes = [C.c_void_p, C.c_char_p]
        lib.omc_inp_get.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p, C.c_size_t]
        lib.omc_inp_get_path.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p, C.c_size_t]
        _host = lib
    return _host


def check(rc: int, lib=None) -> int:
    if rc < 0:
        lib = lib or host()
        raise RuntimeError(f"omc error {rc}: {lib.omc_last_error().decode()}")
    return rc


def default_paths(data_root: Path | None = None) -> DataPaths:
    d = Path(data_root) if data_root else DATA
    return DataPaths(str(d / "pegs4" / "700icru.pegs4dat").encode(),
                     str(d / "pegs4" / "pgs4form.dat").encode(),
                     (str(d / "xs") + "/").encode())


def ensure_spinms(data_root: Path | None = None) -> Path:
    d = Path(data_root) if data_root else DATA
    p = d / "xs" / "spinms.data"
    if not p.exists() or p.stat().st_size != 7066704:
        check(host().omc_write_synthetic_spinms(str(p).encode()))
    return p
