"""Small, deterministic omc_dosxyz cases shared by the fixture generator
(make_golden.py) and the tests.  Each case writes an egsphant phantom, an
optional spectrum and an .inp file into a directory and returns the .inp
stem.  The generated text is byte-for-byte reproducible."""
from __future__ import annotations

from pathlib import Path

from ompmc_amd import configs as cfg

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
}


def write_case(name: str, d: Path, ncase: int | None = None, seeds=(97, 33)) -> Path:
    d = Path(d)
    d.mkdir(parents=True, exist_ok=True)
    build, kw, n = CASES[name]
    kw = dict(kw)
    ph = build()
    pf = ph.write_egsphant(d / "phantom.egsphant")
    if kw.pop("spectrum", False):
        kw["spectrum_file"] = cfg.write_spectrum(d / "6mv.spectrum")
    items = cfg.dosxyz_inp_items(d, pf, ncase=ncase or n, seeds=seeds, **kw)
    stem = d / name
    cfg.write_inp(stem, items)
    return stem
