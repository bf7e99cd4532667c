"""Small, deterministic omc_dosxyz cases shared by the fixture generator
(make_golden.py) and the tests.  Each case writes an egsphant phantom, an
optional spectrum and an .inp file into a directory and returns the .inp
stem.  The generated text is byte-for-byte reproducible."""
from __future__ import annotations

from pathlib import Path

from ompmc_amd import configs as cfg

# Silver (Z = 47) is the highest-Z medium the reference can build tables
# for: it reads the XCOM data of elements 1-50 only (MXELEMENT, ompmc.h:359)
# and indexes them by Z - 1 (ompmc.c:365), so gold or lead read outside them.
SILVER = "AG700ICRU"


def silver_slab(n=12, size=6.0):
    """Water with a 2 cm silver slab at depth 1-3 cm: photons of several MeV
    interact there mostly by pair production."""
    ph = cfg.water_cube(n, size)
    zc = 0.5 * (ph.zb[1:] + ph.zb[:-1])
    layer = (zc >= 1.0) & (zc < 3.0)
    ph.names = [cfg.WATER, SILVER]
    ph.med[layer] = 2
    ph.rho[layer] = 10.5
    return ph


CASES = {
    # name: (phantom builder, source kwargs, ncase)
    "water2MeV": (lambda: cfg.water_cube(12, 6.0), dict(mono=2.0, field=(-1, 1, -1, 1)), 2000),
    "slab6MV": (lambda: cfg.slab_phantom(16, 16.0), dict(spectrum=True, field=(-5, 5, -5, 5)), 3000),
    "patient6MV": (lambda: cfg.patient_phantom(32, 32, 24, 1.0), dict(spectrum=True, field=(-5, 5, -5, 5)), 3000),
    "electron6MeV": (lambda: cfg.water_cube(12, 6.0), dict(mono=6.0, charge=-1, field=(-1, 1, -1, 1)), 1500),
    "positron6MeV": (lambda: cfg.water_cube(12, 6.0), dict(mono=6.0, charge=1, field=(-1, 1, -1, 1)), 1500),
    "split5": (lambda: cfg.water_cube(12, 6.0), dict(mono=6.0, nsplit=5, field=(-1, 1, -1, 1)), 1500),
    "pencil": (lambda: cfg.water_cube(12, 6.0), dict(mono=2.0, field=(0, 0, 0, 0)), 2000),
    "oversize": (lambda: cfg.water_cube(8, 4.0), dict(mono=1.0, field=(-20, 20, -20, 20)), 2000),
    "C1": (lambda: cfg.water_cube(), dict(mono=2.0), 5000),
    # photons of 8 MeV in the silver slab: about one pair production per
    # three histories, with Bhabha scattering and annihilation of the
    # positrons it makes
    "pairAg8MeV": (silver_slab, dict(mono=8.0, field=(-1, 1, -1, 1)), 1500),
}

# Cases without a reference fixture: 52 MeV photons take pair()'s
# Coulomb-corrected branch (E >= 50 MeV) and brems()'s high-energy one, but
# positrons above ~40 MeV make the reference read its positron eta spline
# past the end of the table (ompmc.c:2697, 2725), so its output there is not
# defined.  The device is compared with the oracle on them.
DEVICE_CASES = {
    "pairAg52MeV": (silver_slab, dict(mono=52.0, field=(-1, 1, -1, 1)), 300),
    # the C2 and C3 configurations at their full grids: the 64^3 slab at
    # 0.25 cm and the 128x128x96 patient at 0.25 cm, 6 MV, 10x10 cm^2
    "C2": (lambda: cfg.config_phantom("C2"), dict(spectrum=True, field=(-5, 5, -5, 5)), 2000),
    "C3": (lambda: cfg.config_phantom("C3"), dict(spectrum=True, field=(-5, 5, -5, 5)), 1000),
}


# Problems for the statistical gates (>= 95% of the voxels above 10% of the
# maximum within 2 sigma): tens of thousands of voxels in the field, so that
# the fraction measures agreement rather than chance.
STAT_CASES = {
    "statSlab": (lambda: cfg.slab_phantom(48, 16.0), dict(spectrum=True, field=(-7, 7, -7, 7)), 4000000),
    "statPatient": (lambda: cfg.patient_phantom(64, 64, 48, 0.5), dict(spectrum=True, field=(-8, 8, -8, 8)),
                    4000000),
    "statElectron": (lambda: cfg.water_cube(48, 12.0), dict(mono=6.0, charge=-1, field=(-5, 5, -5, 5)), 4000000),
    "statSplit": (lambda: cfg.water_cube(48, 12.0), dict(mono=6.0, nsplit=5, field=(-5, 5, -5, 5)), 2000000),
}


def write_case(name: str, d: Path, ncase: int | None = None, seeds=(97, 33), nbatch: int = 10) -> Path:
    d = Path(d)
    d.mkdir(parents=True, exist_ok=True)
    build, kw, n = CASES.get(name) or DEVICE_CASES.get(name) or STAT_CASES[name]
    kw = dict(kw)
    ph = build()
    pf = ph.write_egsphant(d / "phantom.egsphant")
    if kw.pop("spectrum", False):
        kw["spectrum_file"] = cfg.write_spectrum(d / "6mv.spectrum")
    items = cfg.dosxyz_inp_items(d, pf, ncase=ncase or n, seeds=seeds, nbatch=nbatch, **kw)
    stem = d / name
    cfg.write_inp(stem, items)
    return stem
