"""The elementary functions of the FP64 transport on the device against the
host C library the oracle and the reference use (glibc libm), through
omc_dev_math_eval on arguments of the ranges the transport evaluates them
on.  sqrt is IEEE and must match exactly; log, exp, sin, cos and pow are
faithfully rounded on both sides (within one ulp of each other) but need not
round the same way.  Where they differ, a history that evaluates one of
them at such an argument can leave the oracle's path: this test measures
how often, per function, which is the cause the exact-replay tests name."""
import ctypes as C
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LIBM = C.CDLL("libm.so.6")
for _f in ("log", "exp", "sqrt", "sin", "cos"):
    getattr(LIBM, _f).restype = C.c_double
    getattr(LIBM, _f).argtypes = [C.c_double]
LIBM.pow.restype = C.c_double
LIBM.pow.argtypes = [C.c_double, C.c_double]


def ulps(a, b):
    """Distance in units in the last place between doubles of one sign."""
    ia = a.view(np.int64)
    ib = b.view(np.int64)
    return np.abs(ia - ib)


RANGES = {
    "log": lambda r, n: (np.exp(r.uniform(np.log(1e-7), np.log(60.0), n)), None),
    "exp": lambda r, n: (r.uniform(-30.0, 5.0, n), None),
    "sqrt": lambda r, n: (np.exp(r.uniform(np.log(1e-12), np.log(1e4), n)), None),
    "sin": lambda r, n: (r.uniform(0.0, 2 * math.pi, n), None),
    "cos": lambda r, n: (r.uniform(0.0, 2 * math.pi, n), None),
    "pow": lambda r, n: (r.uniform(0.01, 50.0, n), r.uniform(-3.0, 3.0, n)),
}


@pytest.mark.parametrize("fn", sorted(RANGES))
def test_fp64_functions_against_host_libm(fn, capsys):
    from ompmc_amd import device as D
    rng = np.random.default_rng(7)
    n = 200000
    x, y = RANGES[fn](rng, n)
    dev = D.math_eval(fn, x, y, precision=D.FP64)
    f = getattr(LIBM, fn)
    host = np.array([f(a) for a in x] if y is None else [f(a, b) for a, b in zip(x, y)])
    d = ulps(dev, host)
    frac = float(np.mean(d > 0))
    with capsys.disabled():
        print(f"\n[device libm] {fn}: {np.count_nonzero(d)} of {n} arguments differ from glibc "
              f"({100 * frac:.4f}%), max {int(d.max())} ulp")
    if fn == "sqrt":
        assert d.max() == 0
    else:
        assert d.max() <= 1, f"{fn}: {int(d.max())} ulp"
