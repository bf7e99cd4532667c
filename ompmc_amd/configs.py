"""Synthetic workloads: phantoms, spectra, beamlet sets and .inp files.

Config names follow the benchmark definitions:
  C1  water cube 40^3 at 0.5 cm, 2 MeV photons, 5x5 cm^2 at SSD 100
  C2  64^3 slab phantom at 0.25 cm (water / lung 0.26 / bone 1.85 / water),
      synthetic 6 MV spectrum, 10x10 cm^2 at SSD 100
  C3  128x128x96 patient-like phantom at 0.25 cm, 400 beamlets 0.5x0.5 cm^2
      tiling 10x10 cm^2 at gantry 0, SAD 100
  C5  256x256x160 at 0.2 cm, the C3 geometry scaled, 5 beams 8x8 cm^2

Geometry choices the config text leaves open are fixed here:
  * the isocentre is the phantom centre; gantry rotates about the z axis,
    gantry 0 places the source at (0, -SAD, 0) so beams travel along +y;
  * lung ellipsoid centres at x = +-6 cm (y = z = 0);
  * bone cylinder (spine) along z at x = 0, y = +7 cm (posterior, far side);
  * outside the body: AIR700ICRU at 0.0012 g/cm^3.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import _native as N

WATER, LUNG, BONE, AIR = "H2O700ICRU", "LUNG700ICRU", "ICRPBONE700ICRU", "AIR700ICRU"


@dataclass
class PhantomSpec:
    names: list
    xb: np.ndarray
    yb: np.ndarray
    zb: np.ndarray
    med: np.ndarray      # (nz, ny, nx) int, 1-based medium numbers
    rho: np.ndarray      # (nz, ny, nx) float64

    @property
    def shape(self):
        return len(self.xb) - 1, len(self.yb) - 1, len(self.zb) - 1

    @property
    def nvox(self):
        i, j, k = self.shape
        return i * j * k

    def write_egsphant(self, path):
        """egsphant text in the layout read by omc_dosxyz (omc_dosxyz.c:85-152)."""
        path = Path(path)
        nx, ny, nz = self.shape
        with open(path, "w") as f:
            f.write(f"{len(self.names)}\n")
            for n in self.names:
                f.write(n + "\n")
            f.write("  ".join(["0.25"] * len(self.names)) + "\n")
            f.write(f"{nx:5d}{ny:5d}{nz:5d}\n")
            for b in (self.xb, self.yb, self.zb):
                f.write(" ".join(repr(float(v)) for v in b) + "\n")
            digits = (self.med + ord("0")).astype(np.uint8)
            for k in range(nz):
                rows = [bytes(digits[k, j]).decode() for j in range(ny)]
                f.write("\n".join(rows) + "\n\n")
            for k in range(nz):
                sl = self.rho[k].reshape(-1)
                f.write("\n".join(" ".join(repr(float(v)) for v in sl[i:i + 16])
                                  for i in range(0, sl.size, 16)))
                f.write("\n\n")
        return path

    def to_native(self):
        """omc_phantom owned by the host library (free with omc_phantom_free)."""
        lib = N.host()
        p = N.Phantom()
        nx, ny, nz = self.shape
        N.check(lib.omc_phantom_alloc(C_ref(p), len(self.names), nx, ny, nz))
        for i, n in enumerate(self.names):
            p.names[i].value = n.encode()
        for arr, ptr in ((self.xb, p.xbounds), (self.yb, p.ybounds), (self.zb, p.zbounds)):
            np.ctypeslib.as_array(ptr, shape=(arr.size,))[:] = arr
        np.ctypeslib.as_array(p.med, shape=(self.nvox,))[:] = self.med.reshape(-1)
        np.ctypeslib.as_array(p.rho, shape=(self.nvox,))[:] = self.rho.reshape(-1)
        return p


def C_ref(obj):
    import ctypes
    return ctypes.byref(obj)


def _grid(n, lo, hi):
    return np.linspace(lo, hi, n + 1, dtype=np.float64)


def water_cube(n=40, size=20.0) -> PhantomSpec:
    h = size / 2
    xb, yb, zb = _grid(n, -h, h), _grid(n, -h, h), _grid(n, 0.0, size)
    med = np.ones((n, n, n), dtype=np.int32)
    rho = np.ones((n, n, n), dtype=np.float64)
    return PhantomSpec([WATER], xb, yb, zb, med, rho)


def slab_phantom(n=64, size=16.0) -> PhantomSpec:
    h = size / 2
    xb, yb, zb = _grid(n, -h, h), _grid(n, -h, h), _grid(n, 0.0, size)
    zc = 0.5 * (zb[1:] + zb[:-1])
    med = np.empty((n, n, n), dtype=np.int32)
    rho = np.empty((n, n, n), dtype=np.float64)
    for k, z in enumerate(zc):
        if z < 3.0:
            m, r = 1, 1.0
        elif z < 9.0:
            m, r = 2, 0.26
        elif z < 11.0:
            m, r = 3, 1.85
        else:
            m, r = 1, 1.0
        med[k] = m
        rho[k] = r
    return PhantomSpec([WATER, LUNG, BONE], xb, yb, zb, med, rho)


def patient_phantom(nx=128, ny=128, nz=96, dx=0.25, scale=(1.0, 1.0, 1.0)) -> PhantomSpec:
    """Water ellipsoid body with two lungs and a spine, air outside."""
    sx, sy, sz = scale
    X, Y, Z = nx * dx / 2, ny * dx / 2, nz * dx / 2
    xb, yb, zb = _grid(nx, -X, X), _grid(ny, -Y, Y), _grid(nz, -Z, Z)
    xc = 0.5 * (xb[1:] + xb[:-1])
    yc = 0.5 * (yb[1:] + yb[:-1])
    zc = 0.5 * (zb[1:] + zb[:-1])
    zz, yy, xx = np.meshgrid(zc, yc, xc, indexing="ij")
    body = (xx / (16 * sx)) ** 2 + (yy / (11 * sy)) ** 2 + (zz / (20 * sz)) ** 2 <= 1.0
    lung = np.zeros_like(body)
    for cx in (-6.0 * sx, 6.0 * sx):
        lung |= ((xx - cx) / (5 * sx)) ** 2 + (yy / (4 * sy)) ** 2 + (zz / (10 * sz)) ** 2 <= 1.0
    spine = (xx / sx) ** 2 + ((yy - 7.0 * sy) / sy) ** 2 <= 1.2 ** 2
    med = np.full(body.shape, 1, dtype=np.int32)            # air
    rho = np.full(body.shape, 0.0012, dtype=np.float64)
    med[body] = 3
    rho[body] = 1.0
    m = body & lung
    med[m] = 2
    rho[m] = 0.3
    m = body & spine
    med[m] = 4
    rho[m] = 1.6
    return PhantomSpec([AIR, LUNG, WATER, BONE], xb, yb, zb, med, rho)


def config_phantom(name: str) -> PhantomSpec:
    if name == "C1":
        return water_cube()
    if name == "C2":
        return slab_phantom()
    if name == "C3":
        return patient_phantom()
    if name == "C5":
        return patient_phantom(256, 256, 160, 0.2, scale=(1.6, 1.6, 32.0 / 24.0))
    raise KeyError(name)


def spectrum_6mv():
    """Synthetic 6 MV: 24 bins of 0.25 MeV, weight ~ Ec^-0.5 (1 - Ec/6)."""
    edges = np.arange(1, 25) * 0.25
    ec = edges - 0.125
    w = ec ** -0.5 * (1.0 - ec / 6.0)
    w = w / w.sum()
    return edges, w


def write_spectrum(path, edges=None, weights=None, title="synthetic 6 MV"):
    if edges is None:
        edges, weights = spectrum_6mv()
    with open(path, "w") as f:
        f.write(title + "\n")
        f.write(f"   {len(edges)}  0.0  0\n")
        for e, w in zip(edges, weights):
            f.write(f"  {float(e)!r}  {float(w)!r}\n")
    return Path(path)


def write_inp(path_stem, items: dict):
    """`key = value` lines as read by parseInputFile (omc_utilities.c:63-105)."""
    path = Path(str(path_stem) + ".inp")
    with open(path, "w") as f:
        for k, v in items.items():
            f.write(f"{k} = {v}\n")
    return path


def dosxyz_inp_items(workdir: Path, phantom_file, *, ncase, nbatch=10, seeds=(97, 33),
                     mono=None, spectrum_file=None, field=(-2.5, 2.5, -2.5, 2.5), ssd=100.0,
                     charge=0, ecut=0.7, pcut=0.01, nsplit=1, data_root=None):
    d = Path(data_root) if data_root else N.DATA
    items = {}
    if spectrum_file is not None:
        items["spectrum file"] = str(spectrum_file)
    items["mono energy"] = mono if mono is not None else 2.0
    items["charge"] = charge
    items["collimator bounds"] = " ".join(str(v) for v in field)
    items["ssd"] = ssd
    items["ncase"] = int(ncase)
    items["nbatch"] = int(nbatch)
    items["rng seeds"] = f"{seeds[0]} {seeds[1]}"
    items["phantom file"] = str(phantom_file)
    items["global ecut"] = ecut
    items["global pcut"] = pcut
    items["pegs file"] = str(d / "pegs4" / "700icru.pegs4dat")
    items["pgs4form file"] = str(d / "pegs4" / "pgs4form.dat")
    items["nsplit"] = nsplit
    items["data folder"] = str(d / "xs") + "/"
    items["output folder"] = str(workdir) + "/"
    return items


@dataclass
class BeamletSet:
    ibeam: np.ndarray        # (nb,) int32, 0-based
    src: np.ndarray          # (nbeams, 3)
    corner: np.ndarray       # (nb, 3)
    side1: np.ndarray        # (nb, 3)
    side2: np.ndarray        # (nb, 3)


def beamlets(gantry_deg=(0.0,), field=10.0, bixel=0.5, sad=100.0, iso=(0.0, 0.0, 0.0)) -> BeamletSet:
    """Bixels tiling a field x field square at the isocentre for each gantry
    angle (rotation about z; gantry 0 = source at -y)."""
    iso = np.asarray(iso, dtype=np.float64)
    n = int(round(field / bixel))
    ib, corners, s1, s2, srcs = [], [], [], [], []
    for b, g in enumerate(gantry_deg):
        th = math.radians(g)
        d = np.array([math.sin(th), math.cos(th), 0.0])
        e1 = np.array([math.cos(th), -math.sin(th), 0.0])
        e2 = np.array([0.0, 0.0, 1.0])
        srcs.append(iso - sad * d)
        for j in range(n):
            for i in range(n):
                c = iso + (-field / 2 + i * bixel) * e1 + (-field / 2 + j * bixel) * e2
                ib.append(b)
                corners.append(c)
                s1.append(bixel * e1)
                s2.append(bixel * e2)
    return BeamletSet(np.asarray(ib, np.int32), np.asarray(srcs), np.asarray(corners),
                      np.asarray(s1), np.asarray(s2))


def loguniform_histories(n, lo, hi, seed=12345):
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(math.log(lo), math.log(hi), size=n)).astype(np.int64)


def c5_beams(sad=100.0, field=8.0, gantry=(0.0, 72.0, 144.0, 216.0, 288.0)) -> BeamletSet:
    """The large multi-beam plan: five coplanar beams, each one field x field
    square at the isocentre (one beamlet per beam, matRad beam geometry)."""
    return beamlets(gantry_deg=gantry, field=field, bixel=field, sad=sad)


def c4_phantoms():
    """The three phantoms the serving stream keeps resident -- water 40^3 at
    0.5 cm, the 64^3 slab, the 128x128x96 patient -- on one media list
    (AIR, LUNG, WATER, BONE), so one table upload serves all three."""
    names = [AIR, LUNG, WATER, BONE]
    out = []
    for ph in (water_cube(40, 20.0), slab_phantom(64, 16.0), patient_phantom()):
        lut = np.array([0] + [names.index(n) + 1 for n in ph.names], dtype=np.int32)
        ph.med = lut[ph.med]
        ph.names = list(names)
        out.append(ph)
    return out


@dataclass
class ServingRequest:
    arrival_s: float         # time after the stream starts
    phantom: int             # 0 water 40^3, 1 slab 64^3, 2 patient 128x128x96
    side_cm: float           # square field at the phantom surface (SSD 100)
    energy: float            # mono energy in MeV, or 0 for the 6 MV spectrum
    nhist: int


def serving_stream(n, rate, seed=4242, hist_lo=1e5, hist_hi=1e8):
    """The serving stream: Poisson arrivals at `rate` per second, phantom
    uniform over the three resident ones, square field side U(1, 20) cm,
    energy mono U(1, 6) MeV or the 6 MV spectrum with probability 1/2 each,
    histories log-uniform in [hist_lo, hist_hi]."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n))
    out = []
    for i in range(n):
        ph = int(rng.integers(0, 3))
        side = float(rng.uniform(1.0, 20.0))
        energy = float(rng.uniform(1.0, 6.0)) if rng.random() < 0.5 else 0.0
        nh = int(math.exp(rng.uniform(math.log(hist_lo), math.log(hist_hi))))
        out.append(ServingRequest(float(t[i]), ph, side, energy, nh))
    return out

