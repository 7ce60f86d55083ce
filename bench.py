#!/usr/bin/env python3
"""Benchmark of the RK3 time advance on one MI355X.

Workload: the single production run 512x513x256 (x, y, z), density ratio 4,
Re = 800, seeded perturbation of amplitude 1e-3 with a Gaussian envelope in y
(synthetic initial condition, loma_phd_amd.ic).  A "step" is one full RK3
step (3 substeps) over the whole grid, state resident in HBM.  Both forms
are timed: plain FP64 and mixed (FP32 transforms, FP64 y-solves and RK
accumulation).  `value` is the mixed form's grid-point*RK3-steps/s; the FP64
form, their ratio, p50/p99 step latencies, the whole-step fraction of the HBM
roofline and the dominant kernel's roofline are reported beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config B|A]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E

# HBM bytes per launch measured with rocprofv3 on the 512x513x256 step
# without Z: read = 32/64/128 x the TCC read requests of each size, write =
# WRITE_SIZE x 1024; a plain 1 GiB device copy in the same runs reads and
# writes 1 GiB by these sums (python3 -m loma_phd_amd.profile_step; summary
# in profiles/rk3_step_512x513x256_hbm_traffic.md).
MEASURED_TRAFFIC = {
    "fp64": {"phaseA": 4.262e9, "zfft_inv": 4.775e9, "xline": 8.034e9,
             "zfft_fwd": 6.582e9, "phaseC": 4.328e9, "phaseC2": 0.978e9},
    "mixed": {"phaseA": 3.309e9, "zfft_inv": 2.408e9, "xline": 4.553e9,
              "zfft_fwd": 3.327e9, "phaseC": 3.112e9, "phaseC2": 0.853e9},
}


_TRAFFIC_KEYS = {"k_phaseA": "phaseA", "k_zfft_inv": "zfft_inv", "k_xline": "xline",
                 "k_zfft_fwd": "zfft_fwd", "k_phaseC": "phaseC", "k_phaseC32": "phaseC",
                 "k_phaseC2": "phaseC2"}


def _kernel_key(name):
    """Kernel class of a rocprofv3 kernel name (None for other kernels)."""
    base = name.replace("void ", "").replace("loma::", "").split("(")[0].split("<")[0]
    return _TRAFFIC_KEYS.get(base.strip())


def _pmc_bytes(path, counters):
    """Per-dispatch bytes of one counter pass (csv of rocprofv3 --pmc)."""
    import collections
    import csv
    per, names = collections.defaultdict(float), {}
    with open(path) as f:
        for r in csv.DictReader(f):
            k = int(r["Dispatch_Id"])
            names[k] = r["Kernel_Name"]
            per[k] += counters.get(r["Counter_Name"], 0.0) * float(r["Counter_Value"])
    return per, names


def measure_traffic(precision_name, timeout=240):
    """HBM bytes per launch of each kernel class, measured in this run: two
    rocprofv3 counter passes over one RK3 step of the 512x513x256 case
    (python3 -m loma_phd_amd.profile_step, child processes).  Read bytes =
    32/64/128 x the TCC read requests of each size, write bytes = WRITE_SIZE
    x 1024, each divided by what the same sum reads for the step's trailing
    1 GiB device copy (the calibration factor, 1.0 on MI355X).  Returns
    (bytes per launch by kernel class, calibration) or None when rocprofv3 is
    missing or a pass gives no counters."""
    import shutil
    import signal
    import subprocess
    import tempfile
    exe = shutil.which("rocprofv3")
    if exe is None:
        return None
    passes = (("rd", {"TCC_EA0_RDREQ_32B_sum": 32.0, "TCC_EA0_RDREQ_64B_sum": 64.0,
                      "TCC_EA0_RDREQ_128B_sum": 128.0, "TCC_EA0_RDREQ_sum": 0.0}),
              ("wr", {"WRITE_SIZE": 1024.0}))
    env = dict(os.environ, TMPDIR="/tmp",
               PYTHONPATH=HERE + os.pathsep + os.environ.get("PYTHONPATH", ""))
    got = {}
    with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
        for name, counters in passes:
            d = os.path.join(tmp, name)
            cmd = [exe, "--pmc", *counters, "--output-format", "csv", "-d", d, "-o", "p", "--",
                   sys.executable, "-m", "loma_phd_amd.profile_step", precision_name]
            proc = subprocess.Popen(cmd, cwd="/tmp", env=env, stdout=subprocess.DEVNULL,
                                    stderr=subprocess.DEVNULL, start_new_session=True)
            try:
                proc.wait(timeout=timeout)
            except subprocess.TimeoutExpired:
                os.killpg(proc.pid, signal.SIGKILL)  # the group this call started
                proc.wait()
                return None
            # rocprofv3 may end non-zero after the program and its files are
            # complete; the counters are what count
            path = next((os.path.join(r, f) for r, _, fs in os.walk(d) for f in fs
                         if f.endswith("counter_collection.csv")), None)
            if path is None:
                return None
            got[name] = _pmc_bytes(path, counters)
    calib, out = {}, {}
    for name, (per, names) in got.items():
        copies = [v for k, v in per.items() if "copyBuffer" in names[k]]
        if not copies or max(copies) <= 0:
            return None
        calib[name] = float(1 << 30) / max(copies)
        by = {}
        for k, v in per.items():
            key = _kernel_key(names[k])
            if key is not None:
                by.setdefault(key, []).append(v)
        for key, vals in by.items():
            out[key] = out.get(key, 0.0) + calib[name] * sum(vals) / len(vals)
    return out, {k: round(v, 4) for k, v in calib.items()}


def algorithmic_bytes(run, with_z: bool):
    """Compulsory HBM bytes per launch of each kernel class (one substep)
    and per RK3 step: one read of every field a kernel consumes and one write
    of every field it produces (S, W, P sized; es bytes per complex transform
    entry), nothing else."""
    S, W, P = run.S, run.W, run.P
    es = 8 if run.precision == 1 else 16
    z = 1 if with_z else 0
    nin, nout = (12, 12) if with_z else (8, 11)
    k = {
        # reads phi, ten13, vor, ten12, psi, T (+Z); writes phi, vor; writes
        # the transform inputs
        "phaseA": (6 + z) * 16 * S + 2 * 16 * S + nin * es * S,
        "zfft_inv": nin * es * S + nin * es * W,
        "xline": nin * es * W + nout * es * W + (1 + z) * 2 * 8 * P,
        "zfft_fwd": nout * es * W + nout * es * S,
        # reads the transform outputs but drho; reads T (+Z), vor, phi;
        # writes T (+Z), vor, phi, RHS(vor), RHS(phi)
        "phaseC": (nout - 1) * es * S + (3 + z) * 16 * S + (5 + z) * 16 * S,
        # reads drho, writes psi
        "phaseC2": es * S + 16 * S,
    }
    step = 3 * sum(k.values())
    return k, step


def model_bytes_per_step(S, P, mixed: bool):
    """Three-phase compulsory-traffic model of the reference dataflow:
    55 S-sized complex passes + 2 P-sized real passes per substep."""
    return (1728 * S + 48 * P) if mixed else (2640 * S + 48 * P)


def bench_precision(cfg, st, precision, steps, warmup, prof_steps):
    import torch
    from loma_phd_amd import capi
    flags = capi.LOMA_FLAG_SCALAR_Z if cfg.scalar_z else 0
    run = capi.Run(cfg.mgalx, cfg.my, cfg.mgalz, st["y"], cfg.Re, cfg.alp, cfg.bet, cfg.pr,
                   cfg.cfl, cfg.ncfl, cfg.dt0, precision, flags)
    run.upload(st["vor"], st["phi"], st["psi"], st["scal"], st["u00"], st["v00"], st["w00"],
               z=st["z"] if cfg.scalar_z else None)
    run.advance(warmup)
    stream = torch.cuda.ExternalStream(run.stream()) if run.stream() else \
        torch.cuda.default_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev[0].record(stream)
    for q in range(steps):
        run.advance_async(1)
        ev[q + 1].record(stream)
    run.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per = np.array([ev[q].elapsed_time(ev[q + 1]) for q in range(steps)])
    total_ms = ev[0].elapsed_time(ev[steps])
    # per-kernel durations, HIP events on the run's stream
    run.profile(True)
    run.advance(prof_steps)
    prof = run.profile_read()
    run.profile(False)
    kb, step_bytes = algorithmic_bytes(run, cfg.scalar_z)
    out = dict(
        ms_per_step=total_ms / steps,
        wall_ms_per_step=wall * 1e3 / steps,
        p50_ms=float(np.percentile(per, 50)),
        p99_ms=float(np.percentile(per, 99)),
        gpts_per_s=cfg.points * steps / (total_ms * 1e-3),
        kernels_ms={k: (v[0] / v[1] if v[1] else 0.0) for k, v in prof.items()},
        kernel_launches={k: v[1] for k, v in prof.items()},
        kernel_bytes=kb,
        step_bytes=step_bytes,
        S=run.S, W=run.W, P=run.P, device_bytes=run.bytes,
        nsteps_total=warmup + steps + prof_steps,
    )
    out["final_state"] = run.download()
    run.close()
    return out


def bench_band(nsys=16384, sizes=(257, 385, 513), reps=20, seed=3):
    """Batched banded solves in isolation: nsys = 128 x 128 (kx, kz) modes of
    n-row pentadiagonal systems with seeded diagonally dominant bands
    (|diag| >= 2.5 x the off-diagonal sum) plus a log-uniform shift in
    [1, 1e4], complex N(0,1) right-hand sides in FP32 and FP64 (FP64 factor
    and solve), through loma_band5_solve_dev.  Algorithmic bytes per system:
    n x (40 band + 16 or 8 rhs + 16 solution); the factor scratch the kernel
    writes and reads back (2 x 40 bytes per row) is not counted."""
    import torch
    from loma_phd_amd import capi
    dev = torch.device("cuda:0")
    g = torch.Generator(device="cpu").manual_seed(seed)
    out = {}
    for n in sizes:
        band = torch.rand((nsys, n, 5), dtype=torch.float64, generator=g) * 2.0 - 1.0
        shift = 10.0 ** (4.0 * torch.rand((nsys, 1), dtype=torch.float64, generator=g))
        off = band[:, :, [0, 1, 3, 4]].abs().sum(dim=2)
        band[:, :, 2] = 2.5 * off + 0.1 + shift
        for j, d in ((0, 0), (0, 1), (1, 0), (n - 1, 4), (n - 1, 3), (n - 2, 4)):
            band[:, j, d] = 0.0
        bd = band.to(dev)
        for rhs_fp32 in (False, True):
            dt = torch.complex64 if rhs_fp32 else torch.complex128
            rhs = torch.complex(torch.randn((nsys, n), dtype=torch.float64, generator=g),
                                torch.randn((nsys, n), dtype=torch.float64, generator=g))
            r = rhs.to(dt).to(dev)
            x = torch.empty((nsys, n), dtype=torch.complex128, device=dev)
            stream = torch.cuda.current_stream()
            capi.band5_solve_dev(bd.data_ptr(), r.data_ptr(), x.data_ptr(), n, nsys,
                                 rhs_fp32, stream.cuda_stream)
            ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            ev0.record(stream)
            for _ in range(reps):
                capi.band5_solve_dev(bd.data_ptr(), r.data_ptr(), x.data_ptr(), n, nsys,
                                     rhs_fp32, stream.cuda_stream)
            ev1.record(stream)
            torch.cuda.synchronize()
            ms = ev0.elapsed_time(ev1) / reps
            alg = nsys * n * (40 + (8 if rhs_fp32 else 16) + 16)
            out[f"n{n}_{'fp32' if rhs_fp32 else 'fp64'}_rhs"] = dict(
                ms=round(ms, 4), algorithmic_bytes=alg,
                achieved_GBps=round(alg / (ms * 1e-3) / 1e9, 1),
                frac=round(alg / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
                scratch_bytes=nsys * n * 80)
    return dict(systems=nsys, form="per-system bands (bandec/banbks), two wavefronts per 64 systems (real and imaginary right-hand side)",
                results=out)


def roofline_of(res, traffic=None):
    # dominant kernel = largest total time per step
    per_step = {k: res["kernels_ms"][k] * (3 if k != "finish" else 1)
                for k in res["kernels_ms"]}
    dom = max((k for k in per_step if k in res["kernel_bytes"]), key=lambda k: per_step[k])
    t = res["kernels_ms"][dom] * 1e-3
    achieved = res["kernel_bytes"][dom] / t / 1e9
    tr = traffic.get(dom) if traffic else None
    return dom, dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS,
                     unit="GB/s", frac=round(achieved / HBM_PEAK_GBS, 4),
                     traffic=int(tr) if tr else None,
                     kernel=dom, bytes_per_launch=int(res["kernel_bytes"][dom]),
                     avg_launch_ms=round(res["kernels_ms"][dom], 4))


def reference_timing_here():
    """The reference solver and the oracle timed on one machine (the build
    container, where the reference exists): per-step seconds of each at
    192x453x96 (the reference's compile-time grid), stored beside the
    reference run's outputs by tests/golden/make_ref_cross1.py."""
    path = os.path.join(HERE, "tests", "golden", "ref_cross1_192x453x96.npz")
    if not os.path.exists(path):
        return None
    with np.load(path, allow_pickle=False) as z:
        if "oracle_step_s" not in z.files:
            return None
        ref_s = float(np.median(z["ref_step_wall_s"]))
        orc_s = float(np.median(z["oracle_step_s"]))
        ranks = int(z["ranks"])
        host = str(z["host"])
    return dict(grid="192x453x96", reference_step_s=round(ref_s, 3),
                reference_ranks=ranks, oracle_step_s=round(orc_s, 3),
                restatement_over_reference_time=round(orc_s / ref_s, 3), host=host)


def cpu_baseline(cfg, st, max_threads=16):
    """FP64 oracle (C restatement) on the host cores: one full RK3 step of
    the same workload, as grid-pt*RK3-steps/s, with the reference/oracle
    timing of the build machine beside it."""
    from oracle import oracle as O
    O.build()
    threads = max(1, min(max_threads, os.cpu_count() or 1))
    O.set_threads(threads)
    g = O.Grid(cfg.mgalx, cfg.my, cfg.mgalz, st["y"], cfg.Re, cfg.alp, cfg.bet, cfg.pr)
    S = O.State(g, dt0=cfg.dt0, cfl=cfg.cfl, ncfl=cfg.ncfl, flags=2 if cfg.scalar_z else 0)
    S.set(st["vor"], st["phi"], st["psi"], st["scal"], st["u00"], st["v00"], st["w00"],
          z=st["z"] if cfg.scalar_z else None)
    t0 = time.perf_counter()
    S.step(1)
    dt = time.perf_counter() - t0
    out = dict(value=cfg.points / dt, unit="grid-pt*RK3-steps/s", cores=threads,
               kind="port",
               sample=f"1 RK3 step of {cfg.mgalx}x{cfg.my}x{cfg.mgalz}, "
                      f"FP64 C oracle, {threads} OpenMP threads, {dt:.1f} s")
    here = reference_timing_here()
    if here:
        out["reference_vs_restatement_on_build_host"] = here
    return out


def bench_ensemble(nruns, steps, warmup, seed=7, workers=8):
    """Ensemble batch: nruns independent mixed-precision runs (grids drawn
    from 128x129x64, 192x193x96, 256x257x128; density ratio U[1,8]; Re
    U[160,800]) advanced by one graph launch per RK3 step (loma_advance_many),
    tile shapes from the autotuner."""
    import concurrent.futures as cf
    from loma_phd_amd import capi, grid, ic
    cfgs = grid.ensemble_configs(nruns, seed)
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        sts = list(ex.map(ic.initial_condition, cfgs))
    flags = capi.LOMA_FLAG_SCALAR_Z | capi.LOMA_FLAG_AUTOTUNE
    runs = []
    for c, st in zip(cfgs, sts):
        r = capi.Run(c.mgalx, c.my, c.mgalz, st["y"], c.Re, c.alp, c.bet, c.pr, c.cfl,
                     c.ncfl, c.dt0, capi.LOMA_MIXED, flags)
        r.upload(st["vor"], st["phi"], st["psi"], st["scal"], st["u00"], st["v00"],
                 st["w00"], z=st["z"])
        r.use_graph(True)
        runs.append(r)
    del sts
    capi.advance_many(runs, warmup)
    t0 = time.perf_counter()
    capi.advance_many(runs, steps)
    wall = time.perf_counter() - t0
    lat = []
    for _ in range(10):
        t1 = time.perf_counter()
        capi.advance_many(runs, 1)
        lat.append((time.perf_counter() - t1) * 1e3)
    pts = sum(c.points for c in cfgs)
    # each grid of the ensemble alone (one run, same precision, flags and
    # tiles): the ensemble step against its runs stepped one after another
    single_ms = {}
    for c, r in zip(cfgs, runs):
        key = (c.mgalx, c.my, c.mgalz)
        if key in single_ms:
            continue
        r.advance(1)
        t1 = time.perf_counter()
        r.advance(3)
        single_ms[key] = (time.perf_counter() - t1) * 1e3 / 3
    seq_ms = sum(single_ms[(c.mgalx, c.my, c.mgalz)] for c in cfgs)
    out = dict(runs=nruns, steps=steps, grid_points=pts,
               gpts_per_s=round(pts * steps / wall, 1),
               ms_per_ensemble_step=round(wall * 1e3 / steps, 3),
               p50_ms=round(float(np.percentile(lat, 50)), 3),
               p99_ms=round(float(np.percentile(lat, 99)), 3),
               device_bytes=int(sum(r.bytes for r in runs)),
               grids={f"{g[0]}x{g[1]}x{g[2]}": sum(1 for c in cfgs
                                                  if (c.mgalx, c.my, c.mgalz) == g)
                      for g in grid.ENSEMBLE_GRIDS},
               tiles=sorted({(c.mgalx, c.my, c.mgalz, *r.tiles()) for c, r in zip(cfgs, runs)}),
               precision="mixed",
               launch="one hipGraph launch per RK3 step; one launch per kernel, substage and "
                      "group of runs with the same instantiation, one capture lane per group",
               single_ms_per_step={f"{k[0]}x{k[1]}x{k[2]}": round(v, 3)
                                   for k, v in single_ms.items()},
               runs_in_sequence_ms=round(seq_ms, 3),
               vs_runs_in_sequence=round(seq_ms / (wall * 1e3 / steps), 3))
    for r in runs:
        r.close()
    return out


def bench_stats(cfg, st, steps=3):
    """Cost of the statistics window (addstats1/2, the phys2 sums) at the
    bench workload, mixed: ms per step over `steps` steps with the window
    closed (captured step graph) and open (direct launches plus the
    statistics kernels at rkstep 1)."""
    import torch
    from loma_phd_amd import capi
    run = capi.Run(cfg.mgalx, cfg.my, cfg.mgalz, st["y"], cfg.Re, cfg.alp, cfg.bet, cfg.pr,
                   cfg.cfl, cfg.ncfl, cfg.dt0, capi.LOMA_MIXED, 0)
    run.upload(st["vor"], st["phi"], st["psi"], st["scal"], st["u00"], st["v00"], st["w00"])
    run.use_graph(True)
    rows = np.unique(np.linspace(cfg.my // 8, cfg.my - cfg.my // 8, 30).astype(int))
    run.stats_config(rows)
    run.advance(2)

    def timed(n):
        torch.cuda.synchronize()
        t = time.perf_counter()
        run.advance(n)
        torch.cuda.synchronize()
        return (time.perf_counter() - t) * 1e3 / n

    closed = timed(steps)
    run.stats_accumulate(True)
    opened = timed(steps)
    nacum = run.stats_read()["nacum"]
    run.close()
    return dict(steps=steps, spectra_rows=int(rows.size) + 1,
                ms_per_step_window_closed=round(closed, 3),
                ms_per_step_window_open=round(opened, 3),
                extra_ms_per_step=round(opened - closed, 3), nacum=nacum)


def bench_served(window, rate=2.0, seed=11, max_slots=32, workers=8):
    """Served ensemble: Poisson arrivals (rate per s) over `window` seconds,
    grids from 64x129x64 to 384x385x192, density ratios U[1,8], step budgets
    U[200,2000], admission by device footprint; runs still resident at the
    end of the window are cut after their step in flight."""
    import torch
    from loma_phd_amd.serve import HipBackend, Server, poisson_requests
    free, _ = torch.cuda.mem_get_info()
    budget = int(0.8 * free)
    reqs = poisson_requests(seed, rate, window)
    be = HipBackend()
    try:
        rep = Server(be, mem_budget=budget, max_slots=max_slots,
                     workers=workers).run(reqs, window=window)
    finally:
        be.close()
    prep = sorted(b - a for n, a, b in be.spans if n in ("prepare", "seed"))
    return dict(window_s=window, arrival_rate_per_s=rate, requests=len(reqs),
                prep_s=dict(n=len(prep), p50=round(prep[len(prep) // 2], 2) if prep else None,
                            max=round(prep[-1], 2) if prep else None),
                gpts_per_s=round(rep["gp_steps_per_s"], 1), steps=rep["steps"],
                p50_ms=round(rep["p50_ms"], 3), p99_ms=round(rep["p99_ms"], 3),
                device_p50_ms=round(rep["device_p50_ms"], 3),
                by_grid={g: {k: (round(v, 3) if isinstance(v, float) else v)
                             for k, v in d.items()} for g, d in rep["by_grid"].items()},
                peak_active=rep["peak_active"], peak_bytes=rep["peak_bytes"],
                slow_loop_passes=rep["slow_loop_passes"],
                slowest_loop_pass_ms=round(rep["slowest_loop_pass_ms"], 1),
                slowest_call_ms={k: round(v, 1) for k, v in rep["slowest_call_ms"].items()},
                slow_calls=rep["slow_calls"], slow_call_overlaps=rep["slow_call_overlaps"],
                slow_pass_overlaps=rep["slow_pass_overlaps"],
                reuse="finished runs kept per grid and re-targeted",
                preparation="seeded draws on a host thread, state built on the device (loma_seed)",
                budget_bytes=budget, outcomes=rep["counts"], precision="mixed",
                latency="host wall time from step enqueue to observed completion")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="B", choices=["A", "B"])
    ap.add_argument("--profile-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--band", type=int, default=1,
                    help="1: time the batched banded solves in isolation (16,384 systems)")
    ap.add_argument("--precision", default="both", choices=["both", "fp64", "mixed"])
    ap.add_argument("--ensemble-runs", type=int, default=32,
                    help="runs in the ensemble-batch leg (0 skips it)")
    ap.add_argument("--ensemble-steps", type=int, default=30)
    ap.add_argument("--stats-steps", type=int, default=3,
                    help="steps timed with the statistics window closed and open (0: skip)")
    ap.add_argument("--traffic", default="live", choices=["live", "const"],
                    help="roofline.traffic from rocprofv3 counter passes of this run (live) "
                         "or the published per-kernel constants")
    ap.add_argument("--served-window", type=float, default=30.0,
                    help="seconds of Poisson arrivals in the served leg (0 skips it)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: this build has no multi-GPU path "
              f"(asked for --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    from loma_phd_amd import grid, ic
    import dataclasses
    cfg = dataclasses.replace(grid.CONFIG_B if args.config == "B" else grid.CONFIG_A)
    if args.config == "B":
        cfg.scalar_z = False
    print(f"[bench] building initial condition {cfg.mgalx}x{cfg.my}x{cfg.mgalz}",
          file=sys.stderr, flush=True)
    st = ic.initial_condition(cfg)
    res = {}
    for prec in (["fp64", "mixed"] if args.precision == "both" else [args.precision]):
        print(f"[bench] timing {prec}", file=sys.stderr, flush=True)
        res[prec] = bench_precision(cfg, st, 0 if prec == "fp64" else 1, args.steps,
                                    args.warmup, args.profile_steps)
        print(f"[bench] {prec}: {res[prec]['ms_per_step']:.3f} ms/step", file=sys.stderr,
              flush=True)
    head = res.get("mixed") or res["fp64"]
    mixed = "mixed" in res
    measured = (args.config == "B" and not cfg.scalar_z)
    tr = {p: (MEASURED_TRAFFIC[p] if measured else None) for p in ("fp64", "mixed")}
    tr_source, tr_calib = "profiles/rk3_step_512x513x256_hbm_traffic.md (earlier run)", None
    if measured and args.traffic == "live":
        live = {}
        for prec in res:
            print(f"[bench] HBM traffic counters ({prec})", file=sys.stderr, flush=True)
            m = measure_traffic(prec)
            if m is None:
                live = None
                break
            live[prec], tr_calib = m[0], m[1]
        if live:
            tr.update(live)
            tr_source = "rocprofv3 counters of this run (loma_phd_amd.profile_step)"
    dom, roof = roofline_of(head, tr["mixed" if mixed else "fp64"])
    if measured:
        roof["traffic_source"] = tr_source
        if tr_calib:
            roof["traffic_calibration"] = tr_calib
    S, P = head["S"], head["P"]
    line = {
        "metric": "grid-pt*RK3-steps/s (mixed vs FP64, 1 GPU)",
        "value": round(head["gpts_per_s"], 1),
        "unit": "grid-pt*RK3-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if mixed else "f64",
        "data": "synthetic (seeded mixing-layer initial condition)",
        "config": {"workload": f"temporal mixing layer {cfg.mgalx}x{cfg.my}x{cfg.mgalz}, "
                               f"s={cfg.s:g}, Re={cfg.Re:g}, one RK3 step",
                   "grid": [cfg.mgalx, cfg.my, cfg.mgalz], "precision": "mixed" if mixed
                   else "fp64"},
        "p50_ms": round(head["p50_ms"], 4),
        "p99_ms": round(head["p99_ms"], 4),
        "hbm_roof_frac_step": round(model_bytes_per_step(S, P, mixed) /
                                    (head["ms_per_step"] * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
        "roofline": roof,
    }
    for prec, r in res.items():
        line[prec] = {"gpts_per_s": round(r["gpts_per_s"], 1),
                      "ms_per_step": round(r["ms_per_step"], 4),
                      "p50_ms": round(r["p50_ms"], 4), "p99_ms": round(r["p99_ms"], 4),
                      "hbm_roof_frac_step": round(
                          model_bytes_per_step(S, P, prec == "mixed") /
                          (r["ms_per_step"] * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
                      "kernel_ms": {k: round(v, 4) for k, v in r["kernels_ms"].items()},
                      "dominant": roofline_of(r, tr[prec])[1]}
    if "fp64" in res and "mixed" in res:
        line["speedup_mixed_vs_fp64"] = round(res["mixed"]["gpts_per_s"] /
                                              res["fp64"]["gpts_per_s"], 3)
        # per-field error of the mixed state against the FP64 state after the
        # same number of steps from the same initial condition
        from loma_phd_amd import diagnostics
        errs = diagnostics.field_errors(res["mixed"]["final_state"], res["fp64"]["final_state"],
                                        ic.HostOps(st["y"]), cfg.alp, cfg.bet)
        line["mixed_vs_fp64_field_errors"] = dict(
            steps=res["mixed"]["nsteps_total"], **{k: float(f"{v:.3e}") for k, v in errs.items()})
    if args.stats_steps > 0:
        print("[bench] statistics window", file=sys.stderr, flush=True)
        line["stats_window"] = bench_stats(cfg, st, args.stats_steps)
    if args.band:
        print("[bench] batched banded solves", file=sys.stderr, flush=True)
        line["band"] = bench_band()
    if args.ensemble_runs > 0:
        print(f"[bench] ensemble batch of {args.ensemble_runs} runs", file=sys.stderr,
              flush=True)
        e = bench_ensemble(args.ensemble_runs, args.ensemble_steps, args.warmup)
        if "fp64" in res:
            e["vs_fp64_single"] = round(e["gpts_per_s"] / res["fp64"]["gpts_per_s"], 3)
        line["ensemble"] = e
    if args.served_window > 0:
        print(f"[bench] served ensemble, {args.served_window:g} s window", file=sys.stderr,
              flush=True)
        sv = bench_served(args.served_window)
        if "fp64" in res:
            sv["vs_fp64_single"] = round(sv["gpts_per_s"] / res["fp64"]["gpts_per_s"], 3)
        line["served"] = sv
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline (oracle, bounded sample)", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(cfg, st)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
