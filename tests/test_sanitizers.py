"""The host library and the CPU oracle under AddressSanitizer,
UndefinedBehaviorSanitizer and LeakSanitizer (CPU only): a build of both
from their sources (make -C oracle sanitize) loads each problem, runs the
transport in both random-number modes, writes and reads the .3ddose and
frees everything; any out-of-bounds access, undefined operation or leak
ends it with an error."""
import os
import subprocess
from pathlib import Path

import pytest

from cases import write_case

ROOT = Path(__file__).resolve().parents[1]
EXE = ROOT / "oracle" / "build" / "omc_sanitize"


@pytest.fixture(scope="module")
def sanitized():
    r = subprocess.run(["make", "-C", str(ROOT / "oracle"), "sanitize"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return EXE


@pytest.mark.parametrize("case", ["water2MeV", "slab6MV", "patient6MV", "positron6MeV", "split5", "pairAg8MeV",
                                  "oversize"])
def test_host_and_oracle_clean_under_sanitizers(sanitized, tmp_path, case):
    stem = write_case(case, tmp_path, ncase=200)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(sanitized), str(stem), str(tmp_path / "o.3ddose")], capture_output=True, text=True,
                       timeout=300, env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0 and "ok" in r.stdout, out[-4000:]
    assert "runtime error" not in out and "AddressSanitizer" not in out and "LeakSanitizer" not in out, out[-4000:]
