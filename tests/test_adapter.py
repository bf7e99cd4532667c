"""The ompMC-globals adapter (include/omc_ompmc_adapter.h).

An ompMC program that already builds its tables can hand them to the device
engine through omc_tables_from_ompmc().  The checker oracle/_ref/ref_adapter
is the reference's own initialisation (initPhantom, initMediaData, built
from its sources) with the adapter on top; every array it fills must equal
what omc_tables_build makes for the same media, and hash to the committed
fixtures of the reference's globals.  Entries the reference leaves undefined
(never written, or read past its arrays) are masked on both sides, as the
fixtures do."""
import hashlib
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

from cases import write_case
from make_golden import TABLES, mask_undefined

ROOT = Path(__file__).resolve().parents[1]
EXE = ROOT / "oracle" / "_ref" / "ref_adapter"
GOLDEN = Path(__file__).resolve().parent / "golden"

pytestmark = pytest.mark.skipif(not EXE.exists(), reason="reference build (oracle/_ref) not present")


def _field_len(name, dt, per, nmed):
    if per is None:
        return {"ums": 16384, "fms": 16384, "wms": 16384, "ims": 16384, "spin_scalars": 5,
                "ms_scalars": 2}[name]
    return per * nmed


def _built(tab, name, n):
    import ctypes as C
    if name == "sig_ismonotone":
        return np.array([tab.sig_ismonotone[q][m] for q in (0, 1) for m in range(tab.nmed)], dtype=np.int32)
    if name == "spin_scalars":
        return np.array([tab.b2spin_min, tab.dbeta2i, tab.espml, tab.dleneri, tab.dqq1i])
    if name == "ms_scalars":
        return np.array([tab.dllambi, tab.dqmsi])
    v = getattr(tab, name)
    if isinstance(v, C.Array):
        return np.ctypeslib.as_array(v)[:n].copy()
    return np.ctypeslib.as_array(v, shape=(n,)).copy()


@pytest.mark.parametrize("case,media_set", [("water2MeV", "water"), ("slab6MV", "slab"),
                                            ("patient6MV", "patient"), ("pairAg8MeV", None)])
def test_adapter_fills_tables_equal_to_omc_tables_build(oracle, tmp_path, case, media_set):
    stem = write_case(case, tmp_path / "case")
    out = tmp_path / "dump"
    out.mkdir()
    subprocess.run([str(EXE), str(stem), str(out)], check=True, capture_output=True, timeout=300)
    prob = oracle.load_problem(stem)
    try:
        tab = prob.tab
        nmed = tab.nmed
        assert np.fromfile(out / "nmed.bin", dtype="<i4")[0] == nmed
        eke0 = np.ctypeslib.as_array(tab.eke0)[:nmed].copy()
        eke1 = np.ctypeslib.as_array(tab.eke1)[:nmed].copy()
        gold = json.loads((GOLDEN / f"tables_{media_set}.json").read_text())["tables"] if media_set else {}
        for name, dt, per in TABLES:
            n = _field_len(name, dt, per, nmed)
            got = np.fromfile(out / f"{name}.bin", dtype="<" + dt)
            assert got.size == n, name
            a = mask_undefined(name, got, nmed, eke0, eke1)
            b = mask_undefined(name, _built(tab, name, n).astype(dt), nmed, eke0, eke1)
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), f"{case}: {name} differs"
            if name in gold:
                assert hashlib.sha256(a.astype("<" + dt).tobytes()).hexdigest() == gold[name]["sha256"], name
        for name in ("rho", "ae", "ap", "ue", "up", "te", "thmoll"):
            got = np.fromfile(out / f"{name}.bin", dtype="<f8")
            assert np.array_equal(got, np.ctypeslib.as_array(getattr(tab, name))[:nmed]), name
    finally:
        oracle.free_problem(prob)
