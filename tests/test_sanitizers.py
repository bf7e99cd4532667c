"""The host library and the CPU oracle under AddressSanitizer,
UndefinedBehaviorSanitizer and LeakSanitizer (CPU only): a build of both
from their sources (make -C oracle sanitize) loads each problem, runs the
transport in both random-number modes, writes and reads the .3ddose and
frees everything; any out-of-bounds access, undefined operation or leak
ends it with an error."""
import fcntl
import os
import subprocess
from contextlib import contextmanager
from pathlib import Path

import pytest

from cases import write_case

ROOT = Path(__file__).resolve().parents[1]
EXE = ROOT / "oracle" / "build" / "omc_sanitize"
LOCK = ROOT / "oracle" / "build" / ".sanitize.lock"


@contextmanager
def locked(exclusive):
    """Test processes running in parallel share the build: it is (re)made
    under an exclusive lock and run under a shared one, so no process
    relinks it while another executes it."""
    LOCK.parent.mkdir(parents=True, exist_ok=True)
    with open(LOCK, "a") as f:
        fcntl.flock(f, fcntl.LOCK_EX if exclusive else fcntl.LOCK_SH)
        try:
            yield
        finally:
            fcntl.flock(f, fcntl.LOCK_UN)


@pytest.fixture(scope="module")
def sanitized():
    with locked(True):
        r = subprocess.run(["make", "-C", str(ROOT / "oracle"), "sanitize"], capture_output=True, text=True,
                           timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return EXE


@pytest.mark.parametrize("case", ["water2MeV", "slab6MV", "patient6MV", "positron6MeV", "split5", "pairAg8MeV",
                                  "oversize"])
def test_host_and_oracle_clean_under_sanitizers(sanitized, tmp_path, case):
    stem = write_case(case, tmp_path, ncase=200)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    with locked(False):
        r = subprocess.run([str(sanitized), str(stem), str(tmp_path / "o.3ddose")], capture_output=True,
                           text=True, timeout=300, env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0 and "ok" in r.stdout, out[-4000:]
    assert "runtime error" not in out and "AddressSanitizer" not in out and "LeakSanitizer" not in out, out[-4000:]
