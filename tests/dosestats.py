"""Statistical comparison of two dose estimates given as batch sums.

Two runs of the same problem with independent random numbers are compared
voxel by voxel: z = (mean_a - mean_b) / sqrt(se_a^2 + se_b^2), with each
standard error estimated from the run's batches (omc_dosxyz.c:741-751).
The acceptance gate is the one the equal-accuracy target states: at least
95% of the voxels above 10% of the maximum within |z| < 2, and the integral
energies within 0.5%.

For identical distributions the expected fraction within 2 is P(|t| < 2)
at the pooled degrees of freedom (95.2% at 98, 95.4% in the normal limit),
so a 95% gate only separates equal engines from biased ones when many
voxels take part and the batches are many: below GATE_MIN_VOXELS voxels
or GATE_MIN_DOF degrees of freedom a test asserts less than it claims, and
the helpers refuse to evaluate the gate."""
from __future__ import annotations

import numpy as np

GATE_MIN_VOXELS = 20000
GATE_MIN_DOF = 150


def mean_se(s, s2, nb):
    m = np.asarray(s, dtype=np.float64) / nb
    v = np.maximum(np.asarray(s2, dtype=np.float64) / nb - m * m, 0.0) / (nb - 1)
    return m, np.sqrt(v)


def compare(a, b, rel=0.1, select=None):
    """a, b = (sum, sum2, nbatch) over regions (region 0 first, ignored).
    Returns dict with the fraction of the voxels above rel * max within 2
    combined sigma, their number, the degrees of freedom, the mean and rms
    of z, and the integral ratio a/b.  Voxels are selected on the mean of
    both estimates: selecting on one of them favours voxels where that one
    came out high, which biases z upwards near the threshold."""
    ma, ea = mean_se(*a)
    mb, eb = mean_se(*b)
    ma, ea, mb, eb = ma[1:], ea[1:], mb[1:], eb[1:]
    avg = 0.5 * (ma + mb)
    sel = avg > rel * avg.max()
    if select is not None:
        sel &= select
    z = (ma[sel] - mb[sel]) / np.sqrt(ea[sel] ** 2 + eb[sel] ** 2 + 1e-300)
    return {"frac": float((np.abs(z) < 2).mean()), "n": int(sel.sum()), "dof": int(a[2] + b[2] - 2),
            "zmean": float(z.mean()), "zrms": float(np.sqrt((z * z).mean())),
            "ratio": float(np.asarray(a[0])[1:].sum() / np.asarray(b[0])[1:].sum())}


def assert_equal_accuracy(r, what, integral_tol=0.005):
    """The 95% / 2 sigma / 0.5% gate on a compare() result."""
    assert r["n"] >= GATE_MIN_VOXELS, f"{what}: only {r['n']} voxels above 10% of the maximum"
    assert r["dof"] >= GATE_MIN_DOF, f"{what}: only {r['dof']} degrees of freedom"
    assert r["frac"] >= 0.95, f"{what}: {r['frac']:.4f} of {r['n']} voxels within 2 sigma ({r})"
    assert abs(r["ratio"] - 1.0) <= integral_tol, f"{what}: integral ratio {r['ratio']:.5f} ({r})"
