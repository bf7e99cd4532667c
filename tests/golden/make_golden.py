#!/usr/bin/env python3
"""Regenerates the golden fixtures from the reference build in oracle/_ref.

Run in a container where the ompMC sources are present (make -C oracle ref).
Writes, per case in cases.py:
  ref_dose_<case>.json   SHA-256 of the reference omc_dosxyz (serial build)
                         .3ddose output plus summary numbers
per-function vectors of the reference on the slab case:
  vectors_slab6MV.npz    howfar / hownear, computeDrange / computeEloss and
                         the photon table lookups, inputs and outputs
                         (ref_dump vectors)
and per media set:
  tables_<set>.json      SHA-256 of every table the transport reads, as
                         dumped from the reference's globals by ref_dump,
                         with the entries the reference leaves undefined
                         (see UNDEFINED below) zeroed before hashing
"""
from __future__ import annotations

import hashlib
import json
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(HERE))

from ompmc_amd import configs as cfg  # noqa: E402
from cases import CASES, write_case  # noqa: E402

REF = ROOT / "oracle" / "_ref"

# arrays as (name, dtype, per-medium length or None for fixed size)
TABLES = (
    [(n, "f8", 2000) for n in ("gmfp0", "gmfp1", "gbr10", "gbr11", "gbr20", "gbr21", "cohe0", "cohe1",
                               "pmax0", "pmax1")]
    + [(n, "f8", 100) for n in ("xgrid", "fcum", "b_array", "c_array")] + [("i_array", "i4", 100)]
    + [(n, "f8", 1) for n in ("ge0", "ge1", "bpar0", "bpar1", "delcm", "zbrang", "eke0", "eke1", "blcc",
                              "xcc", "esig_e", "psig_e", "expeke1")]
    + [(n, "f8", 8) for n in ("dl1", "dl2", "dl3", "dl4", "dl5", "dl6")]
    + [("sig_ismonotone", "i4", 2)]
    + [(n, "f8", 500) for n in ("esig0", "esig1", "psig0", "psig1", "ededx0", "ededx1", "pdedx0", "pdedx1",
                                "ebr10", "ebr11", "pbr10", "pbr11", "pbr20", "pbr21", "tmxs0", "tmxs1",
                                "etae_ms0", "etae_ms1", "etap_ms0", "etap_ms1", "q1ce_ms0", "q1ce_ms1",
                                "q1cp_ms0", "q1cp_ms1", "q2ce_ms0", "q2ce_ms1", "q2cp_ms0", "q2cp_ms1",
                                "blcce0", "blcce1", "e_array")]
    + [("range_ep", "f8", 1000), ("spin_rej", "f8", 32768)]
    + [(n, "f8", None) for n in ("ums", "fms", "wms")] + [("ims", "i4", None)]
    + [("spin_scalars", "f8", None), ("ms_scalars", "f8", None)]
)

MEDIA_SETS = {
    "water": [cfg.WATER],
    "slab": [cfg.WATER, cfg.LUNG, cfg.BONE],
    "patient": [cfg.AIR, cfg.LUNG, cfg.WATER, cfg.BONE],
}


def mask_undefined(name, a, nmed, eke0, eke1):
    """Entries the reference never writes or computes from out-of-bounds
    reads (src/ompmc.c:2697 and 2725 read past eta_array for positrons above
    ~48 MeV; b_array/c_array[99], wms/ims[k=31] are never written)."""
    a = a.copy()
    if name in ("b_array", "c_array"):
        a[99::100] = 0
    if name in ("wms", "ims"):
        a[31::32] = 0
    if name in ("etap_ms0", "etap_ms1"):
        for m in range(nmed):
            k = np.arange(500)
            node = np.exp((k + 2 - eke0[m]) / eke1[m])
            a[m * 500 + k[node > 40.0]] = 0
    return a


def table_hashes(arrays, nmed):
    eke0, eke1 = arrays["eke0"], arrays["eke1"]
    out = {}
    for name, dt, per in TABLES:
        a = np.asarray(arrays[name]).astype(dt)
        a = mask_undefined(name, a, nmed, eke0, eke1)
        out[name] = {"sha256": hashlib.sha256(a.astype("<" + dt).tobytes()).hexdigest(), "n": int(a.size),
                     "sample": [float(x) for x in a[:: max(1, a.size // 4)][:4]]}
    return out


def reference_tables(media, work: Path):
    ph = cfg.water_cube(4, 4.0)
    ph.names = list(media)
    ph.med[:] = 1
    pf = ph.write_egsphant(work / "t.egsphant")
    cfg.write_inp(work / "t", cfg.dosxyz_inp_items(work, pf, ncase=10, mono=1.0))
    dump = work / "dump"
    dump.mkdir(exist_ok=True)
    subprocess.run([str(REF / "ref_dump"), "tables", str(work / "t"), str(dump)], check=True,
                   capture_output=True)
    arrays = {}
    for name, dt, _ in TABLES:
        arrays[name] = np.fromfile(dump / f"{name}.bin", dtype="<" + dt)
    return arrays


def main():
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--cases", nargs="*", help="regenerate only these dose cases (no table sets)")
    args = ap.parse_args()
    if not (REF / "ref_dump").exists():
        sys.exit("oracle/_ref not built (make -C oracle ref)")
    for set_name, media in (MEDIA_SETS.items() if args.cases is None else ()):
        with tempfile.TemporaryDirectory() as td:
            arrays = reference_tables(media, Path(td))
            js = {"media": media, "source": "reference ref_dump (oracle/_ref)",
                  "tables": table_hashes(arrays, len(media))}
            (HERE / f"tables_{set_name}.json").write_text(json.dumps(js, indent=1))
            print("tables", set_name)
    if args.cases is None:
        # per-function vectors of the reference (ref_dump vectors) on the
        # slab problem's geometry and media
        with tempfile.TemporaryDirectory() as td:
            stem = write_case("slab6MV", Path(td))
            out = Path(td) / "vec"
            out.mkdir()
            subprocess.run([str(REF / "ref_dump"), "vectors", str(stem), str(out), "600"], check=True,
                           capture_output=True)
            arrs = {n: np.fromfile(out / f"vec_{n}.bin", dtype="<f8").reshape(-1, w)
                    for n, w in (("geometry", 12), ("csda", 12), ("pwlf", 7))}
            np.savez_compressed(HERE / "vectors_slab6MV.npz", **arrs)
            print("vectors", {k: v.shape for k, v in arrs.items()})
    for case in (args.cases if args.cases else CASES):
        with tempfile.TemporaryDirectory() as td:
            stem = write_case(case, Path(td))
            r = subprocess.run([str(REF / "omc_dosxyz_ref"), "-i", str(stem), "-o", "ref"],
                               capture_output=True, text=True, check=True)
            text = (Path(td) / "ref.3ddose").read_bytes()
            vals = text.split(b"\n")
            n = int(vals[0][:5]) * int(vals[0][5:10]) * int(vals[0][10:15])
            dose = np.array(vals[4].split(), dtype=float)
            js = {"case": case, "sha256": hashlib.sha256(text).hexdigest(), "bytes": len(text), "nvox": n,
                  "dose_sum": float(dose.sum()), "dose_max": float(dose.max()),
                  "source": "reference omc_dosxyz, serial build (gcc -O2 -fcommon)"}
            (HERE / f"ref_dose_{case}.json").write_text(json.dumps(js, indent=1))
            print("dose", case, js["sha256"][:12])


if __name__ == "__main__":
    main()
