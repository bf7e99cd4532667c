"""The C-ABI libraries load and export every function the public headers
declare (no device calls: this runs without a GPU)."""
import ctypes as C
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared(header):
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_ ]*\*?\s*\**(omc_[a-z0-9_]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


@pytest.mark.parametrize("header,lib", [("omc_host.h", "libomc_host.so"), ("omc_dev.h", "libomc_hip.so")])
def test_library_exports_every_declared_symbol(header, lib):
    names = declared(header)
    assert len(names) > 5
    L = C.CDLL(str(ROOT / "ompmc_amd" / "lib" / lib))
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, f"{lib} lacks {missing}"


def test_device_library_reports_missing_gpu_cleanly():
    """Without a visible device the engine returns OMC_EDEVICE (never exits)."""
    import os
    from ompmc_amd import device as D
    L = D.lib()
    ctx = C.c_void_p()
    opts = D.DevOpts(999, 0, 0, 0)
    rc = L.omc_dev_create(C.byref(ctx), C.byref(opts))
    assert rc == -7 and not ctx.value
    assert L.omc_dev_last_error()
