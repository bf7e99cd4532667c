"""Readers and writers for the files of the omc_matrad driver: the beamlet
file (matRad's mcSrc fields as text) and the .dij sparse dose-influence
matrix.  The layouts are described in csrc/host/matrad_main.c."""
from __future__ import annotations

from pathlib import Path

import numpy as np

DIJ_MAGIC = b"OMCDIJ1\0"


def write_beamlet_file(path, ibeam, src, corner, side1, side2, nhist=None) -> Path:
    """ibeam 0-based (written 1-based, as matRad numbers beams)."""
    ibeam = np.asarray(ibeam)
    src = np.asarray(src, dtype=np.float64).reshape(-1, 3)
    corner, side1, side2 = (np.asarray(a, dtype=np.float64).reshape(-1, 3) for a in (corner, side1, side2))
    lines = [f"{len(ibeam)} {len(src)}  # nBixels nBeams"]
    lines += [" ".join(repr(float(v)) for v in s) for s in src]
    for i in range(len(ibeam)):
        vals = [str(int(ibeam[i]) + 1)] + [repr(float(v)) for v in (*corner[i], *side1[i], *side2[i])]
        if nhist is not None:
            vals.append(str(int(nhist[i])))
        lines.append(" ".join(vals))
    path = Path(path)
    path.write_text("\n".join(lines) + "\n")
    return path


def read_dij(path):
    """(nrows, ncols, jcs, irs, vals) of a .dij file."""
    raw = Path(path).read_bytes()
    if raw[:8] != DIJ_MAGIC:
        raise ValueError(f"{path}: not a .dij file")
    nrows, ncols, nnz = np.frombuffer(raw, dtype="<i8", count=3, offset=8)
    off = 32
    jcs = np.frombuffer(raw, dtype="<i8", count=ncols + 1, offset=off)
    off += 8 * (ncols + 1)
    irs = np.frombuffer(raw, dtype="<i8", count=nnz, offset=off)
    off += 8 * nnz
    vals = np.frombuffer(raw, dtype="<f8", count=nnz, offset=off)
    if off + 8 * nnz != len(raw):
        raise ValueError(f"{path}: size does not match its header")
    return int(nrows), int(ncols), jcs.copy(), irs.copy(), vals.copy()


def read_3ddose(path):
    """(isize, jsize, ksize, xb, yb, zb, dose, unc) of a .3ddose file, dose and
    unc as flat arrays in file order (x fastest)."""
    tok = Path(path).read_text().split()
    ni, nj, nk = (int(t) for t in tok[:3])
    at = 3
    xb = np.array(tok[at:at + ni + 1], dtype=np.float64); at += ni + 1
    yb = np.array(tok[at:at + nj + 1], dtype=np.float64); at += nj + 1
    zb = np.array(tok[at:at + nk + 1], dtype=np.float64); at += nk + 1
    n = ni * nj * nk
    dose = np.array(tok[at:at + n], dtype=np.float64); at += n
    unc = np.array(tok[at:at + n], dtype=np.float64)
    return ni, nj, nk, xb, yb, zb, dose, unc
