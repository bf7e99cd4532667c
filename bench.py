#!/usr/bin/env python3
"""Benchmark of the shower() path on one MI355X.

value: histories/s of the FP32-transport engine (FP64 tallies) on the
heterogeneous slab phantom -- 64^3 voxels at 0.25 cm, water 0-3 cm / lung
(0.26) 3-9 cm / bone (1.85) 9-11 cm / water to 16 cm -- irradiated by the
synthetic 6 MV spectrum, 10x10 cm^2 field at SSD 100, ECUT 0.7 MeV, PCUT
0.01 MeV, 10 batches.  One step = one device run of --hist histories (source,
transport, batch fold); tables and phantom are resident before the timed
region.  The plain FP64 kernel runs the same steps beside it.

Also reported:
  roofline     the dominant transport kernel's algorithmic HBM bytes (16 B per
               FP64 tally update, 1 B medium + 4/8 B density per voxel
               entered, counted on the device) over its launch durations
               (HIP events on the launching stream)
  beamlets     matRad-style dose influence on the 128x128x96 patient phantom:
               400 beamlets of 0.5x0.5 cm^2 at gantry 0, histories
               log-uniform in [1e4, 1e6], all in one device pass with the
               sparse Dij (1% threshold) assembled on the device; FP32 and FP64,
               and the plain FP64 method (one beamlet per job, in turn)
  latency      Poisson stream of single-beamlet requests, each sized for 2%
               relative sigma at its Dmax voxel, served by continuous batching;
               p50/p99 of submit -> result, FP32 and FP64 at the same load
  autotune     what omc_dev_tune picks for the FP32 engine on the slab
               workload (probe = one step's histories) and the rate it measured
  cpu_baseline the reference's OpenMP omc_dosxyz on a bounded sample of the
               slab workload, on this host's cores

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--hist H]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import re
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

# The engine drives 6 pipeline streams plus a helper stream; with HIP's
# default of 4 hardware queues per process two of them can share a queue and
# serialise.  Give the process 8 before the HIP runtime starts.
if int(os.environ.get("GPU_MAX_HW_QUEUES", "0") or 0) < 8:
    os.environ["GPU_MAX_HW_QUEUES"] = "8"

HBM_PEAK_GBS = 8000.0
METRIC = "histories/s and p50/p99 beamlet-dose latency at 2% σ, FP32-batched vs FP64 GPU"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# ------------------------------------------------------------- workloads --
def slab_problem(workdir: Path, ncase: int):
    from ompmc_amd import configs as cfg, problem as P
    ph = cfg.slab_phantom()
    pf = ph.write_egsphant(workdir / "slab.egsphant")
    spec = cfg.write_spectrum(workdir / "6mv.spectrum")
    items = cfg.dosxyz_inp_items(workdir, pf, ncase=ncase, nbatch=10, spectrum_file=spec,
                                 field=(-5.0, 5.0, -5.0, 5.0), ssd=100.0)
    cfg.write_inp(workdir / "c2", items)
    return P.load_problem(workdir / "c2"), workdir / "c2"


def patient_problem(workdir: Path):
    """128x128x96 patient phantom at 0.25 cm with the 6 MV spectrum and its
    400 beamlets at gantry 0 (isocentre = phantom centre, SAD 100)."""
    from ompmc_amd import _native as N, configs as cfg, problem as P
    ph = cfg.patient_phantom()
    pf = ph.write_egsphant(workdir / "patient.egsphant")
    spec = cfg.write_spectrum(workdir / "6mv.spectrum")
    items = cfg.dosxyz_inp_items(workdir, pf, ncase=1000, nbatch=10, spectrum_file=spec)
    cfg.write_inp(workdir / "c3", items)
    prob = P.load_problem(workdir / "c3")
    bs = cfg.beamlets(gantry_deg=(0.0,), field=10.0, bixel=0.5, sad=100.0)
    src = N.beamlet_source(prob.src.spec, bs.ibeam, bs.src, bs.corner, bs.side1, bs.side2)
    nhist = cfg.loguniform_histories(len(bs.ibeam), 1e4, 1e6, seed=2024)
    return prob, src, nhist


# -------------------------------------------------------------- slab run --
def time_slab(prob, precision, hist, steps, warmup, pipelines=0, chunk=0, seed0=2000):
    from ompmc_amd import device as D
    nreg = prob.reg.nreg
    nbatch = 10
    nper = hist // nbatch
    with D.Engine(precision, pipelines=pipelines, chunk=chunk) as eng:
        eng.upload_media(prob.tab)
        pid = eng.upload_phantom(prob.ph, prob.reg)
        # untimed warm-up: the requested steps, and at least 3 s of transport
        # so that clocks and caches have settled before the timed region
        t_w, w = time.perf_counter(), 0
        while w < warmup or time.perf_counter() - t_w < 3.0:
            eng.run_dosxyz(pid, prob.src, nper, nbatch, seed=(1000 + w, 7), nreg=nreg)
            w += 1
        stats, etot, sums = [], [], []
        t0 = time.perf_counter()
        for k in range(steps):
            s, s2, st, _ = eng.run_dosxyz(pid, prob.src, nper, nbatch, seed=(seed0 + k, 7), nreg=nreg)
            stats.append(st.as_dict())
            sums.append((s, s2))
        el = time.perf_counter() - t0
    etot = [float(s[1:].sum()) for s, _ in sums]
    # the steps' batches together: steps x nbatch batches of nper histories
    batches = (sum(s for s, _ in sums), sum(s2 for _, s2 in sums), steps * nbatch)
    return el, stats, etot, nper * nbatch, batches


def accuracy_gates(prob, a, b):
    """FP32 run `a` against FP64 run `b` (independent seeds), each as
    (sum, sum of squares, batches) per region: the fraction of voxels above
    10% of the FP64 Dmax whose batch means agree within two combined
    standard errors, and the ratio of the integral energies."""
    rho = density(prob)

    def mean_var(s, s2, nb):
        m = s[1:] / nb
        return m, np.maximum(s2[1:] / nb - m * m, 0.0) / (nb - 1)

    from scipy import stats
    ma, va = mean_var(*a)
    mb, vb = mean_var(*b)
    dose_b = np.where(rho < 0.044, 0.0, mb / np.maximum(rho, 1e-30))
    sel = dose_b > 0.1 * dose_b.max()
    n = int(sel.sum())
    frac = float(np.mean(np.abs(ma - mb)[sel] <= 2.0 * np.sqrt(va + vb)[sel]))
    ratio = float(a[0][1:].sum() / b[0][1:].sum())
    # sigma is estimated from the batches, so even two identical engines put
    # only P(|t| <= 2) of the voxels within 2 sigma (94.98% at 58 degrees of
    # freedom); below 95% the gate allows three binomial standard deviations
    # under that expectation
    expect = float(2.0 * stats.t.cdf(2.0, a[2] + b[2] - 2) - 1.0)
    floor = min(0.95, expect - 3.0 * math.sqrt(expect * (1.0 - expect) / max(n, 1)))
    return {"voxels_above_10pct_dmax": n, "fraction_within_2sigma": frac,
            "expected_if_identical": expect, "integral_energy_fp32_over_fp64": ratio,
            "batches": [a[2], b[2]],
            "gates": "at least 95% of those voxels within 2 sigma (or within 3 binomial sd of the coverage "
                     "expected with batch-estimated sigma, if that is below 95%), integral energy within 0.5%",
            "pass": bool(frac >= floor and abs(ratio - 1.0) <= 0.005)}


def autotune_section(prob, precision, probe=4_000_000):
    """omc_dev_tune on this workload: the settings it picks (from the engine
    defaults) and the histories/s it measured for them on the probe."""
    from ompmc_amd import device as D
    with D.Engine(precision) as eng:
        eng.upload_media(prob.tab)
        pid = eng.upload_phantom(prob.ph, prob.reg)
        eng.run_dosxyz(pid, prob.src, 100000, 10, seed=(3, 3), nreg=prob.reg.nreg)      # warm-up
        t0 = time.perf_counter()
        best, hps = eng.tune(pid, prob.src, probe, nbatch=10, seed=(4, 4))
        el = time.perf_counter() - t0
    keys = ("pipelines", "chunk", "yield_trips", "drain_lanes", "drain_trips", "drain_min", "event_div", "fetch_min")
    return {"probe_histories": probe, "seconds": el, "histories_per_s": hps,
            "chosen": {k: best[k] for k in keys}}


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (copy-calibrated), with the algorithmic bytes per
    launch of that same profiled run."""
    f = ROOT / "profiles" / f"{kernel}_slab6mv_fp32_traffic.md"
    if not f.exists():
        return None, None
    text = f.read_text()
    t = re.search(r"traffic_bytes_per_launch: ([0-9.e+]+)", text)
    a = re.search(r"algorithmic_bytes_per_launch: ([0-9.e+]+)", text)
    return (float(t.group(1)), float(a.group(1))) if t and a else (None, None)


def roofline(stats, precision):
    """Dominant kernel of the transport pass and its achieved algorithmic
    bandwidth: bytes = 16 N_dep + (9 | 5) N_cross of the particles it ran."""
    per_cross = 9 if precision == 0 else 5
    tot = {k: sum(s[k] for s in stats) for k in stats[0]}
    kern = {}
    for name, pre in (("k_photon", "photon"), ("k_electron", "electron")):
        ms = tot[f"{pre}_kernel_ms"]
        launches = tot[f"{pre}_launches"]
        byts = 16 * tot[f"{pre}_deposits"] + per_cross * tot[f"{pre}_crossings"]
        kern[name] = dict(ms=ms, launches=launches, bytes=byts)
    name = max(kern, key=lambda k: kern[k]["ms"])
    k = kern[name]
    achieved = k["bytes"] / (k["ms"] / 1e3) / 1e9 if k["ms"] > 0 else 0.0
    hist = tot["histories"]
    measured, alg_profiled = profiled_traffic(name) if precision == 1 else (None, None)
    alg_launch = k["bytes"] / max(k["launches"], 1)
    # the profiled run's measured/algorithmic ratio applied to this run's launches
    ratio = measured / alg_profiled if measured else None
    return {
        "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS, "traffic": ratio * alg_launch if ratio else None,
        "traffic_over_algorithmic": ratio,
        "traffic_source": ({"file": f"profiles/{name}_slab6mv_fp32_traffic.md",
                            "profiled_run": "this slab workload at the same size (1e8 histories per step)",
                            "measured_bytes_per_launch": measured,
                            "algorithmic_bytes_per_launch": alg_profiled,
                            "here": "measured/algorithmic of the profiled run x this run's algorithmic bytes per launch"}
                           if ratio else None),
        "kernel": name, "launches": k["launches"],
        "algorithmic_bytes_per_launch": alg_launch,
        "avg_launch_ms": k["ms"] / max(k["launches"], 1),
        "per_history": {"deposits": tot["deposits"] / hist,
                        "photon_crossings": tot["photon_crossings"] / hist,
                        "electron_crossings": tot["electron_crossings"] / hist,
                        "bytes": (16 * tot["deposits"] + per_cross *
                                  (tot["photon_crossings"] + tot["electron_crossings"])) / hist},
        "all_kernels_GBs": {n: (v["bytes"] / (v["ms"] / 1e3) / 1e9 if v["ms"] else 0.0) for n, v in kern.items()},
    }


# --------------------------------------------------------- beamlet batch --
def beamlet_batch(prob, src, nhist, precision, nbatch=10, engine=None):
    from ompmc_amd import device as D
    with D.Engine(precision, **(engine or {})) as eng:
        eng.upload_media(prob.tab)
        pid = eng.upload_phantom(prob.ph, prob.reg)
        small = np.minimum(nhist, 2000)
        eng.beamlets_sparse(pid, src, small, nbatch, 0.01, seed=(5, 5))          # warm-up
        t0 = time.perf_counter()
        jcs, irs, vals, st = eng.beamlets_sparse(pid, src, nhist, nbatch, 0.01, seed=(6, 6))
        el = time.perf_counter() - t0
    h = int(st.histories)
    return {"seconds": el, "histories": h, "histories_per_s": h / el, "nnz": int(len(irs)),
            "device_transport_ms": st.transport_ms}, (jcs, irs, vals)


def beamlet_plain(prob, src, nhist, precision, nbatch=10):
    """The plain method: one beamlet per device job, each run to completion
    before the next is submitted (no batching across beamlets).  Only the
    job status is polled (no dose copied back), which favours this path."""
    from ompmc_amd import device as D
    with D.Engine(precision) as eng:
        eng.upload_media(prob.tab)
        pid = eng.upload_phantom(prob.ph, prob.reg)
        wait_done(eng, eng.submit(pid, src, 20000, nbatch=nbatch, seed=(5, 5), kind=1, beamlet=0))   # warm-up
        t0 = time.perf_counter()
        total = 0
        for b in range(src.nbeamlets):
            res = wait_done(eng, eng.submit(pid, src, int(nhist[b]), nbatch=nbatch, seed=(6, 6 + b), kind=1,
                                            beamlet=b))
            if res.status != 0:
                raise RuntimeError(f"beamlet {b} failed: {res.status}")
            total += res.nhist
        el = time.perf_counter() - t0
    return {"seconds": el, "histories": total, "histories_per_s": total / el,
            "method": "one beamlet per job, run to completion in turn"}


# --------------------------------------------------------------- latency --
def wait_done(eng, ticket):
    """Steps the engine until the ticket is done; returns its result record."""
    while True:
        res, _, _ = eng.poll(ticket)
        if res.done:
            return res
        eng.step()


def wait_ticket(eng, ticket, nreg):
    """Steps the engine until the ticket is done; returns (result, sum, sum2)."""
    while True:
        res, s, s2 = eng.poll(ticket, nreg)
        if res.done:
            return res, s, s2
        eng.step()


def sigma_at_dmax(s, s2, nbatch, rho):
    """Relative batch sigma (omc_dosxyz.c:741-751 estimator) at the voxel of
    highest dose (energy / density; dose below density 0.044 not reported)."""
    m = s[1:] / nbatch
    v = np.maximum(s2[1:] / nbatch - m * m, 0.0) / (nbatch - 1)
    dose = np.where(rho < 0.044, 0.0, m / np.maximum(rho, 1e-30))
    i = int(np.argmax(dose))
    return math.sqrt(v[i]) / m[i] if m[i] > 0 else float("inf")


def density(prob):
    n = prob.ph.isize * prob.ph.jsize * prob.ph.ksize
    return np.ctypeslib.as_array(prob.ph.rho, shape=(n,)).copy()


def serve_stream(prob, src, beamlets, nh, precision, rate, arrivals, nbatch=10, engine=None):
    from ompmc_amd import device as D
    nreg = prob.reg.nreg
    rho = density(prob)
    lat, done = [], []
    with D.Engine(precision, **(engine or {})) as eng:
        eng.upload_media(prob.tab)
        pid = eng.upload_phantom(prob.ph, prob.reg)
        # warm-up request
        t = eng.submit(pid, src, 20000, nbatch=nbatch, seed=(1, 1), kind=1, beamlet=int(beamlets[0]))
        wait_ticket(eng, t, nreg)
        open_t = {}
        i = 0
        t0 = time.perf_counter()
        total = 0
        while i < len(arrivals) or open_t:
            now = time.perf_counter() - t0
            while i < len(arrivals) and arrivals[i] <= now:
                tk = eng.submit(pid, src, nh, nbatch=nbatch, seed=(100 + i, 3), kind=1, beamlet=int(beamlets[i]))
                open_t[tk] = i
                i += 1
            left = eng.step() if open_t else 0
            for tk in list(open_t):
                res, s, s2 = eng.poll(tk, nreg)
                if res.done:
                    if res.status != 0:
                        raise RuntimeError(f"request {tk} failed: {res.status}")
                    lat.append(res.finish_ms - res.submit_ms)
                    done.append((s, s2))      # analysed after the stream: the loop only drives the engine
                    total += res.nhist
                    del open_t[tk]
            if not left and not open_t and i < len(arrivals):
                time.sleep(max(0.0, min(0.002, arrivals[i] - (time.perf_counter() - t0))))
        el = time.perf_counter() - t0
    sig = [sigma_at_dmax(s, s2, nbatch, rho) for s, s2 in done]
    lat = np.array(lat)
    if os.environ.get("OMC_BENCH_LATENCIES"):
        log(f"latencies ms ({'fp32' if precision else 'fp64'}): " + " ".join(f"{v:.0f}" for v in lat))
    return {"requests": len(lat), "histories_per_request": nh, "offered_rate_per_s": rate,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "histories_per_s": total / el,
            "sigma_at_dmax_median": float(np.median(sig)), "sigma_at_dmax_max": float(np.max(sig))}


def latency_section(prob, src, nreq=48, load=0.7, seed=77, engine=None):
    """Requests sized for 2% sigma at Dmax; Poisson arrivals at `load` of the
    FP64 engine's closed-loop capacity, the same schedule for both precisions."""
    from ompmc_amd import device as D
    nreg = prob.reg.nreg
    rng = np.random.default_rng(seed)
    centre = src.nbeamlets // 2 + int(math.sqrt(src.nbeamlets)) // 2
    # pilot: sigma at Dmax of the central beamlet
    with D.Engine(D.FP32) as eng:
        eng.upload_media(prob.tab)
        pid = eng.upload_phantom(prob.ph, prob.reg)
        npil = 50000
        t = eng.submit(pid, src, npil, nbatch=10, seed=(9, 9), kind=1, beamlet=centre)
        _, s, s2 = wait_ticket(eng, t, nreg)
    sig0 = sigma_at_dmax(s, s2, 10, density(prob))
    nh = int(min(max(npil * (sig0 / 0.02) ** 2, 1e4), 5e6))
    nh -= nh % 10
    beamlets = rng.integers(0, src.nbeamlets, size=nreq)
    # capacity of the FP64 engine on these requests (closed loop)
    with D.Engine(D.FP64, **(engine or {})) as eng:
        eng.upload_media(prob.tab)
        pid = eng.upload_phantom(prob.ph, prob.reg)
        ts = [eng.submit(pid, src, nh, nbatch=10, seed=(50 + k, 1), kind=1, beamlet=int(beamlets[k]))
              for k in range(8)]
        t0 = time.perf_counter()
        while eng.step() > 0:
            pass
        cap64 = 8 / (time.perf_counter() - t0)
        for t in ts:
            eng.poll(t)
    rate = load * cap64
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=nreq))
    out = {"pilot_sigma_at_dmax": sig0, "pilot_histories": npil, "fp64_capacity_requests_per_s": cap64,
           "load_of_fp64_capacity": load}
    for name, prec in (("fp32", D.FP32), ("fp64", D.FP64)):
        log(f"latency stream {name}: {nreq} requests of {nh} histories at {rate:.2f}/s")
        out[name] = serve_stream(prob, src, beamlets, nh, prec, rate, arrivals, engine=engine)
    return out


# ------------------------------------------------------------ CPU baseline --
def cpu_baseline(stem: Path, workdir: Path, ncase: int):
    """The reference's own OpenMP build (oracle/_ref) on a bounded sample of
    the same workload; the oracle's CPU restatement when that is absent."""
    cores = min(len(os.sched_getaffinity(0)), 16)
    exe = ROOT / "oracle" / "_ref" / "omc_dosxyz_ref_omp"
    text = Path(str(stem) + ".inp").read_text()
    text = re.sub(r"ncase = \d+", f"ncase = {ncase}", text)
    sample = workdir / "cpu_sample"
    Path(str(sample) + ".inp").write_text(text)
    if exe.exists():
        cmd = ["taskset", "-c", f"0-{cores - 1}", str(exe), "-i", str(sample), "-o", "cpu_sample"]
        t0 = time.perf_counter()
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
        wall = time.perf_counter() - t0
        out = r.stdout + r.stderr
        ts = [float(x) for x in re.findall(r"Execution time up to this point :\s+([0-9.]+)", out)]
        transport = ts[1] - ts[0] if len(ts) >= 2 else wall
        return {"value": ncase / transport, "unit": "histories/s", "cores": cores, "kind": "reference",
                "sample": f"{ncase} histories of the same slab/6 MV workload; reference omc_dosxyz "
                          f"(gcc -O2 -mtune=native -fopenmp -fcommon), {cores} OpenMP threads, transport "
                          f"{transport:.2f} s of {wall:.2f} s total"}
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as O
    prob = O.load_problem(sample)
    t0 = time.perf_counter()
    O.run_dosxyz_philox(prob, seed=(97, 33), nthreads=cores)
    el = time.perf_counter() - t0
    O.free_problem(prob)
    return {"value": ncase / el, "unit": "histories/s", "cores": cores, "kind": "port",
            "sample": f"{ncase} histories, CPU restatement with {cores} OpenMP threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--hist", type=int, default=100_000_000)
    ap.add_argument("--pipelines", type=int, default=0)
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--cpu-sample", type=int, default=10_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-beamlets", action="store_true")
    ap.add_argument("--no-latency", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU per process; --gpus must be 1 "
                                   "(sharding over several GPUs is designed, not launched)"}))
        sys.exit(2)

    from ompmc_amd import device as D
    work = Path(tempfile.mkdtemp(prefix="omc_bench_"))
    prob, stem = slab_problem(work, args.hist)
    log(f"slab FP32: {args.steps} steps of {args.hist} histories")
    el32, st32, e32, nh, b32 = time_slab(prob, D.FP32, args.hist, args.steps, args.warmup, args.pipelines,
                                         args.chunk)
    hps32 = args.steps * nh / el32
    out = {
        "metric": METRIC,
        "value": hps32,
        "unit": "histories/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * el32 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (phantoms, spectrum and beamlets generated in-process)",
        "config": {"workload": "slab phantom 64^3 @ 0.25 cm (water/lung 0.26/bone 1.85/water), synthetic 6 MV, "
                               "10x10 cm2 at SSD 100, ECUT 0.7 MeV, PCUT 0.01 MeV; FP32 transport, FP64 tallies",
                   "histories_per_step": nh, "nbatch": 10,
                   "photon_flights": {"fp32": "delta tracking (engine default)",
                                      "fp64": "voxel stepping, as the reference (engine default)"}},
        "roofline": roofline(st32, 1),
    }
    if not args.no_fp64:
        log("slab FP64")
        el64, st64, e64, _, b64 = time_slab(prob, D.FP64, args.hist, args.steps, args.warmup, args.pipelines,
                                            args.chunk, seed0=3000)
        hps64 = args.steps * nh / el64
        out["fp64"] = {"value": hps64, "ms_per_step": 1000.0 * el64 / args.steps, "roofline": roofline(st64, 0)}
        out["fp32_over_fp64"] = hps32 / hps64
        out["energy_per_history_MeV"] = {"fp32": float(np.mean(e32)) / nh, "fp64": float(np.mean(e64)) / nh}
        out["accuracy_fp32_vs_fp64"] = accuracy_gates(prob, b32, b64)
    out["device_stats_fp32_last_step"] = st32[-1]
    if not args.no_tune:
        log("autotune FP32 on the slab workload")
        out["autotune"] = autotune_section(prob, D.FP32, probe=min(args.hist, 20_000_000))
    if not args.no_beamlets or not args.no_latency:
        pprob, src, bl_nh = patient_problem(work)
        if not args.no_beamlets:
            res = {}
            for name, prec in (("fp32", D.FP32), ("fp64", D.FP64)):
                log(f"beamlets {name}: {src.nbeamlets} beamlets, {int(bl_nh.sum())} histories")
                res[name], _ = beamlet_batch(pprob, src, bl_nh, prec)
            log("beamlets fp32, photon flights stepped voxel by voxel as in the reference")
            res["fp32_voxel_stepping"], _ = beamlet_batch(pprob, src, bl_nh, D.FP32, engine={"woodcock": -1})
            log(f"beamlets plain fp64: {src.nbeamlets} jobs in turn")
            res["plain_fp64"] = beamlet_plain(pprob, src, bl_nh, D.FP64)
            res["fp32_over_fp64"] = res["fp32"]["histories_per_s"] / res["fp64"]["histories_per_s"]
            res["fp32_batched_over_plain_fp64"] = res["fp32"]["histories_per_s"] / res["plain_fp64"]["histories_per_s"]
            res["fp32_voxel_stepping_batched_over_plain_fp64"] = (res["fp32_voxel_stepping"]["histories_per_s"] /
                                                                  res["plain_fp64"]["histories_per_s"])
            res["workload"] = ("patient 128x128x96 @ 0.25 cm, 400 beamlets 0.5x0.5 cm2 at gantry 0, histories "
                               "log-uniform [1e4, 1e6] per beamlet, nbatch 10, sparse Dij at 1% of beamlet max; "
                               "fp32 = delta-tracked photon flights (engine default), fp32_voxel_stepping and the "
                               "fp64 rows step photons voxel by voxel as the reference does")
            out["beamlets"] = res
        if not args.no_latency:
            out["latency"] = latency_section(pprob, src)
    if not args.no_cpu:
        log(f"CPU baseline: {args.cpu_sample} histories")
        out["cpu_baseline"] = cpu_baseline(stem, work, args.cpu_sample)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
