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


PAIRS = [("omc_dev.h", "omc_dev_opts", "device", "DevOpts"), ("omc_dev.h", "omc_dev_stats", "device", "DevStats"),
         ("omc_dev.h", "omc_dev_run_params", "device", "RunParams"),
         ("omc_dev.h", "omc_dev_request", "device", "Request"), ("omc_dev.h", "omc_dev_result", "device", "Result"),
         ("omc_host.h", "omc_element", "_native", "Element"), ("omc_host.h", "omc_tables", "_native", "Tables"),
         ("omc_host.h", "omc_data_paths", "_native", "DataPaths"), ("omc_host.h", "omc_phantom", "_native", "Phantom"),
         ("omc_host.h", "omc_regions", "_native", "Regions"), ("omc_host.h", "omc_spectrum", "_native", "Spectrum"),
         ("omc_host.h", "omc_source_dosxyz", "_native", "SourceDosxyz"),
         ("omc_host.h", "omc_source_beamlets", "_native", "SourceBeamlets"),
         ("omc_host.h", "omc_dose_comparison", "_native", "DoseComparison")]


def test_ctypes_layouts_match_headers(tmp_path):
    """The Python bindings' structures have the headers' size and, for every
    field, the header's offset (a C program built with gcc prints them)."""
    import importlib
    import shutil
    import subprocess
    if not shutil.which("gcc"):
        pytest.skip("no gcc")
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "omc_dev.h"', "int main(void) {"]
    expect = []
    for header, cname, mod, pyname in PAIRS:
        text = (ROOT / "include" / header).read_text()
        body = text[:text.index(f"}} {cname};")]
        body = body[body.rindex("typedef struct"):]
        body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
        cls = getattr(importlib.import_module(f"ompmc_amd.{mod}"), pyname)
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        expect.append(f"{cname} size {C.sizeof(cls)}")
        for f in cls._fields_:
            pyfield = f[0]
            name = pyfield.rstrip("_")          # `pass_`: a Python keyword in C
            assert re.search(rf"\b{name}\b", body), f"{pyname}.{pyfield} is not a member of {cname}"
            lines.append(f'printf("{cname}.{name} %zu\\n", offsetof({cname}, {name}));')
            expect.append(f"{cname}.{name} {getattr(cls, pyfield).offset}")
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), "-o", str(exe), str(src)], check=True)
    got = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    bad = [(g, e) for g, e in zip(got, expect) if g != e]
    assert not bad, bad[:10]
