#!/usr/bin/env python3
"""Benchmark of the MI355X RK3 (magneto)hydrodynamics step.

Headline workload (one GPU): plain FP64 MHD with rotation on a 512x512x256
grid (synthetic stratified convection: polytrope m=1, density contrast 10,
Ra 1e6, Pr 1, Ta 1e4, Pm 1, tanh z grid), CFL time step every step -- the
reference's ``step()`` (3dmhdsub.h:31-153) as a drop-in on the GPU, at the
largest single-GPU configuration.  State and metrics are resident in HBM
before the timed region.

Prints one JSON line: Mcell-RK3-steps/s, the HBM roofline of the stage
kernel (algorithmic bytes / event-timed kernel duration, and the HBM traffic
per launch measured by rocprofv3), the drop-in rate (one step() call per
step, as the reference's main loop calls it), the streaming ceiling of the
memory system, the CPU baseline (the C restatement in oracle/, timed on this
host on a bounded sample) and the mixed / ensemble / served workloads.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "rast-3dmhd-cpp_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E


def alg_bytes_per_cell_step(nf, precision=0):
    """Algorithmic HBM bytes per cell per RK3 step with the 2-register Wray
    scheme.  FP64: stage 1 reads U, writes U and Z; stage 2 reads U, Z, writes
    U, Z; stage 3 reads U, Z, writes U -> 10 field sweeps of 8 bytes.  Mixed:
    stage 1 reads U (8) and writes the FP32 copy (4) and Z (8); stage 2 reads
    the copy and Z, writes a copy and Z (24); stage 3 reads the copy and Z and
    writes U (20) -> 64 bytes per field."""
    return (64 if precision else 80) * nf


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def host_cores():
    """Cores this process may run on (its CPU affinity), and the CPU quota of
    its cgroup when one is set (cgroup v2 cpu.max)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()[:2]
            if q != "max":
                quota = max(1, int(int(q) / int(period)))
    except (OSError, ValueError):
        pass
    return n, quota


def run_cpu_baseline(cfg, state, budget_s=15.0):
    """Time the oracle (C restatement of the reference path) on all the host
    cores this process may use: steady state, differencing two step counts."""
    from oracle import oracle as om
    from mhd3d.desc import Descriptor

    n_aff, quota = host_cores()
    threads = min(n_aff, quota) if quota else n_aff
    desc = Descriptor(cfg)
    o = om.Oracle(desc, threads=threads)
    o.set_state(state)
    o.communicate()
    o.step()                       # warm-up (page-in)
    t0 = time.perf_counter()
    o.step()
    t1 = time.perf_counter()
    n = 1
    while time.perf_counter() - t0 < budget_s and n < 50:
        o.step()
        n += 1
    el = time.perf_counter() - t1  # steps 2..n: the difference of two step counts
    o.close()
    cells = cfg.nPx * cfg.nPy * cfg.nPz
    return {
        "value": cells * (n - 1) / el / 1e6 if n > 1 else cells / (t1 - t0) / 1e6,
        "unit": "Mcell-RK3-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{n - 1} CFL steps timed (after 1 warm-up step and 1 step subtracted) of the same "
                  f"{cfg.nPx}x{cfg.nPy}x{cfg.nPz} workload, {threads} OpenMP threads = every core "
                  f"this process may use (affinity {n_aff}, cgroup quota {quota or 'none'}) on {cpu_model()}",
    }


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_stage_kernel.json")


def measured_traffic(mag, grid, precision=0):
    """HBM bytes per launch of the interior stage kernel (mean of the three
    stages) from the rocprofv3 FETCH_SIZE / WRITE_SIZE passes recorded in
    profiles/traffic_stage_kernel.json (each counter scaled by its reading
    for a copy kernel of known size), or None for an unprofiled workload."""
    try:
        with open(TRAFFIC_FILE) as f:
            table = json.load(f)
    except (OSError, ValueError):
        return None
    key = f"{'mhd' if mag else 'hydro'} {'x'.join(map(str, grid))} {'mixed' if precision else 'fp64'}"
    rec = table.get(key)
    return None if rec is None else rec.get("bytes_per_launch")


def measure_ceiling(bytes_per_array=1 << 30, reps=10):
    """Streaming ceiling of the memory system for the read/write mixes of the
    RK stages (mhd3d_stream_ceiling): GB/s per mix."""
    import ctypes as C
    from mhd3d import _lib as L
    out = {}
    for nr, nw in ((1, 0), (1, 1), (1, 2), (2, 2), (2, 1)):
        ms = C.c_double()
        rc = L.lib().mhd3d_stream_ceiling(int(bytes_per_array), nr, nw, int(reps), C.byref(ms))
        if rc != L.OK:
            raise L.MhdError(rc, "stream ceiling failed")
        out[f"read{nr}_write{nw}"] = round((nr + nw) * bytes_per_array / (ms.value * 1e-3) / 1e9, 1)
    return {"gbs": out, "note": f"{bytes_per_array >> 20} MiB per array, 16-byte vector loads, "
                                "4 in flight per thread and array, 2048 x 256 threads"}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _relerr(g, c):
    g = g[1:-1, 1:-1, 1:-1]
    c = c[1:-1, 1:-1, 1:-1]
    den = float(abs(c).max())
    return float(abs(g - c).max()) / den if den > 0 else float(abs(g - c).max())


def measure_mixed(size=(512, 512, 256), steps=20, err_steps=100):
    """Mixed-precision MHD workload: FP64 and mixed forms from the same device
    initial state; throughput over `steps` timed steps and the per-field /
    integral error of the mixed form against the FP64 form after err_steps."""
    from mhd3d.desc import Descriptor
    from mhd3d.domain import Domain
    from mhd3d.initial import stratified_config

    out = {"workload": f"MHD+rotation {size[0]}x{size[1]}x{size[2]}, Ra 1e6, CFL dt",
           "steps_timed": steps, "error_after_steps": err_steps}
    cells = size[0] * size[1] * size[2]
    doms = []
    for prec in (0, 1):
        cfg = stratified_config(*size, Ra=1e6, lMag=True, lRot=True, precision=prec)
        dom = Domain(Descriptor(cfg))
        dom.init_stratified(seed=4242)
        dom.step(3)
        dom.synchronize()
        dom.timing(True)
        t0 = time.perf_counter()
        dom.step(steps)
        dom.synchronize()
        el = time.perf_counter() - t0
        tm = dom.timing()
        dom.timing(False)
        nf = dom.nf
        interior = size[0] * size[1] * (size[2] - 2)
        avg = tm["stage_ms_total"] / max(1, tm["stage_launches"])
        achieved = alg_bytes_per_cell_step(nf, prec) / 3.0 * interior / (avg * 1e-3) / 1e9
        out["fp64" if prec == 0 else "mixed"] = {
            "mcell_steps_per_s": round(cells * steps / el / 1e6, 1), "ms_per_step": round(el / steps * 1e3, 3),
            "stage_kernel_ms": round(avg, 4), "stage_kernel_gbs": round(achieved, 1)}
        doms.append(dom)
        log(f"mixed workload: form {prec} {out['fp64' if prec == 0 else 'mixed']}")
    out["speedup"] = round(out["mixed"]["mcell_steps_per_s"] / out["fp64"]["mcell_steps_per_s"], 3)
    # error of the mixed form after err_steps from the same initial state
    for dom in doms:
        dom.init_stratified(seed=4242)
        dom.step(err_steps)
    s64 = doms[0].state()
    g64 = doms[0].diagnostics()
    smx = doms[1].state()
    gmx = doms[1].diagnostics()
    out["field_error"] = {k: float("%.3e" % _relerr(smx[k], s64[k])) for k in s64}
    out["mass_rel_error"] = float("%.3e" % (abs(gmx["mass"] - g64["mass"]) / abs(g64["mass"])))
    out["energy_rel_error"] = float("%.3e" % (abs(gmx["energy"] - g64["energy"]) / abs(g64["energy"])))
    out["max_divB"] = {"fp64": float("%.6e" % g64["max_divB"]), "mixed": float("%.6e" % gmx["max_divB"])}
    for dom in doms:
        dom.close()
    return out


def _tolerant_step(obj, n):
    """step(n) under the reference driver's rules: a non-finite member or a
    DT-guard halt ends nothing but that run."""
    from mhd3d import _lib as L
    try:
        obj.step(n)
    except L.MhdError as err:
        if err.code not in (L.ENONFINITE, L.EHALTED):
            raise


def measure_ensemble(n=64, steps=200):
    """Ensemble workload: n runs of uneven grids / flags / parameters advanced
    together (graph replay, autotuned variant), FP64 and mixed, against the
    same runs stepped one after another in plain FP64.  Every run has the
    reference driver's DT guard; throughput counts the steps each run took
    (a run the guard halts stops early)."""
    from mhd3d.desc import Descriptor
    from mhd3d.domain import Domain
    from mhd3d.ensemble import Ensemble
    from mhd3d.workloads import ensemble_mix

    specs = ensemble_mix(n)
    cells = sum(s.cells for s in specs)
    out = {"workload": f"{n} runs, nx=ny in {{32..128}}, nz in {{32..96}}, Ra 1e4-1e6, half lMag",
           "cells": cells, "steps_timed": steps}
    descs0 = []
    for prec in (0, 1):
        descs = [Descriptor(s.config(prec)) for s in specs]
        if prec == 0:
            descs0 = descs
        ens = Ensemble(descs)
        _init_ensemble(ens, descs, specs)
        ens.set_dt_guard()
        _tolerant_step(ens, 2)
        best, trials = ens.autotune(steps=3)
        ens.graph(True)
        _tolerant_step(ens, 2)
        ens.synchronize()
        n0 = [ens.clock(i)[0] for i in range(n)]
        t0 = time.perf_counter()
        _tolerant_step(ens, steps)
        ens.synchronize()
        el = time.perf_counter() - t0
        done = [ens.clock(i)[0] - n0[i] for i in range(n)]
        cs = sum(s.cells * d for s, d in zip(specs, done))
        out["fp64" if prec == 0 else "mixed"] = {
            "mcell_steps_per_s": round(cs / el / 1e6, 1), "ms_per_step": round(el / steps * 1e3, 3),
            "runs_stopped_early": sum(1 for d in done if d < steps),
            "variant": {"waves": best[0], "sched": best[1], "chunk": best[2]}}
        log(f"ensemble: form {prec} {out['fp64' if prec == 0 else 'mixed']}")
        ens.close()
    tot = 0.0
    cs = 0
    for d, s in zip(descs0, specs):
        dom = Domain(d)
        dom.init_stratified(seed=s.seed)
        dom.set_dt_guard()
        _tolerant_step(dom, 4)
        dom.synchronize()
        n0 = dom.clock()[0]
        t0 = time.perf_counter()
        _tolerant_step(dom, steps)
        dom.synchronize()
        tot += time.perf_counter() - t0
        cs += s.cells * (dom.clock()[0] - n0)
        dom.close()
    out["sequential_fp64"] = {"mcell_steps_per_s": round(cs / tot / 1e6, 1),
                              "ms_per_step": round(tot / steps * 1e3, 3)}
    for k in ("fp64", "mixed"):
        out[k]["vs_sequential_fp64"] = round(out[k]["mcell_steps_per_s"] / out["sequential_fp64"]["mcell_steps_per_s"], 3)
    log(f"ensemble: sequential {out['sequential_fp64']}")
    return out


def _init_ensemble(ens, descs, specs):
    """Initial state of every member: the synthetic stratified state of the
    request server's initializer (its host copy)."""
    from mhd3d.initial import stratified_state_hashed
    for i, (d, s) in enumerate(zip(descs, specs)):
        ens.set_state(i, stratified_state_hashed(d, seed=s.seed))


def measure_served(n=40, rate=50.0, quantum=10):
    """Served workload: a Poisson trace replayed against the request server
    (FP64 and mixed) and the same runs one at a time in plain FP64."""
    from mhd3d.server import sequential_trace, serve_trace
    from mhd3d.workloads import served_trace

    specs = served_trace(n, rate=rate)
    out = {"workload": f"{n} runs, Poisson arrivals {rate:g}/s, grids 32^3..256x256x128, "
                       f"steps 50-2000, random lMag/lRot", "quantum_steps": quantum,
           "hbm_budget_gib": 64}
    for prec in (0, 1):
        r = serve_trace(specs, precision=prec, quantum_steps=quantum)
        out["fp64" if prec == 0 else "mixed"] = {
            "p50_s": round(r["p50_s"], 3), "p99_s": round(r["p99_s"], 3),
            "mcell_steps_per_s": round(r["mcell_steps_per_s"], 1), "makespan_s": round(r["makespan_s"], 3),
            "halted_by_dt_guard": r["failed"], "evicted_nonfinite": r["evicted_nonfinite"]}
        log(f"served: form {prec} {out['fp64' if prec == 0 else 'mixed']}")
    q = sequential_trace(specs, 0)
    out["sequential_fp64"] = {"p50_s": round(q["p50_s"], 3), "p99_s": round(q["p99_s"], 3),
                              "mcell_steps_per_s": round(q["mcell_steps_per_s"], 1),
                              "halted_by_dt_guard": q["failed"]}
    for k in ("fp64", "mixed"):
        out[k]["vs_sequential_fp64"] = round(out[k]["mcell_steps_per_s"] / out["sequential_fp64"]["mcell_steps_per_s"], 3)
    log(f"served: sequential {out['sequential_fp64']}")
    return out


def branch_config(name, size, precision=0):
    """The branch workloads of the plane-sum kernels (stratified polytrope,
    Ra 3e4): "iD1" hydro with density diffusion iD 1 (RO means); "lRem1"
    hydro with the Rempel conduction, iTCon 1, izCon 1 (TT means); "rlax"
    MHD+rotation with the RLAX relaxation, iTCon 0, izCon 1, two z ranks
    (TT means, the convective-flux means, the relaxation pass)."""
    import dataclasses
    import numpy as np
    from mhd3d.initial import stratified_config

    mag = name == "rlax"
    cfg = stratified_config(*size, lMag=mag, lRot=mag, Ra=3e4, precision=precision)
    par = cfg.par.copy()
    zs = np.linspace(0.0, 1.0, size[2] + 2)
    if name == "iD1":
        par[55], par[56] = 0.5, 0.4
        return dataclasses.replace(cfg, par=par, iD=1)
    if name == "lRem1":
        return dataclasses.replace(cfg, par=par, lRem=True, iTCon=1, izCon=1, TU=1.0, DZTU=-0.02,
                                   DZTB=0.03, kappa=(1.0 + 0.5 * zs ** 2, 1.0 * zs))
    if name == "rlax":
        par[57], par[58], par[59] = 1.0, 0.5, 0.2
        return dataclasses.replace(cfg, par=par, iTCon=0, izCon=1, TIMI=3.7, relax_nPEz=2)
    raise ValueError(f"unknown branch workload {name!r}")


def profile_run(steps, size=(512, 512, 256), mixed=False, branch=None):
    """For rocprofv3: `steps` steps of the MHD+rotation workload at `size`
    (the headline workload at 512x512x256, Ra 1e6; Ra 1e5 at other sizes;
    FP64 or the mixed form) or of a branch workload (branch_config), then a
    plain device copy of 1 GiB whose counters calibrate FETCH/WRITE_SIZE."""
    from mhd3d.desc import Descriptor
    from mhd3d.domain import Domain
    from mhd3d.initial import stratified_config

    ra = 1e6 if tuple(size) == (512, 512, 256) else 1e5
    if branch:
        cfg = branch_config(branch, tuple(size), precision=1 if mixed else 0)
    else:
        cfg = stratified_config(*size, Ra=ra, lMag=True, lRot=True, precision=1 if mixed else 0)
    dom = Domain(Descriptor(cfg))
    dom.init_stratified(seed=12345)
    dom.step(steps)
    dom.synchronize()
    dom.close()
    import ctypes as C
    from mhd3d import _lib as L
    ms = C.c_double()
    rc = L.lib().mhd3d_calibration_copy(1 << 30, 3, C.byref(ms))
    if rc != L.OK:
        raise L.MhdError(rc, "calibration copy failed")
    log(f"profile run: {steps} steps of {'x'.join(map(str, size))} ({'mixed' if mixed else 'FP64'}"
        f"{', ' + branch if branch else ''}), "
        f"1 + 3 copies of {1 << 30} bytes, "
        f"{ms.value:.4f} ms per copy ({2 * (1 << 30) / ms.value / 1e6:.0f} GB/s)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--size", default="512x512x256")
    ap.add_argument("--hydro", action="store_true", help="lMag off, lRot off")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workloads", default="mixed,ensemble,served",
                    help="extra workloads measured beside the headline ('' for none)")
    ap.add_argument("--profile", type=int, default=0, metavar="STEPS",
                    help="run STEPS steps and a calibration copy for rocprofv3, print nothing "
                         "(MHD+rotation at --size; --mixed for the mixed form)")
    ap.add_argument("--mixed", action="store_true", help="with --profile: the mixed form")
    ap.add_argument("--branch", default=None, choices=["iD1", "lRem1", "rlax"],
                    help="with --profile: a branch workload (plane means, the relaxation)")
    ap.add_argument("--ceiling", action="store_true", help="print the streaming ceiling only")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU z-slab halo design is documented but not built)", file=sys.stderr)
        sys.exit(2)

    if args.profile:
        profile_run(args.profile, tuple(int(v) for v in args.size.split("x")), args.mixed, args.branch)
        return
    if args.ceiling:
        print(json.dumps(measure_ceiling()), flush=True)
        return

    from mhd3d.desc import Descriptor
    from mhd3d.domain import Domain
    from mhd3d.initial import stratified_config

    nx, ny, nz = (int(v) for v in args.size.split("x"))
    mag = not args.hydro
    ra = 1e6 if (nx, ny, nz) == (512, 512, 256) else 1e5
    cfg = stratified_config(nx, ny, nz, Ra=ra, Pr=1.0, lMag=mag, lRot=mag, Ta=1e4, Pm=1.0)
    desc = Descriptor(cfg)
    dom = Domain(desc)
    dom.init_stratified(seed=12345)
    state = None if args.no_cpu_baseline else dom.state()   # host copy for the CPU baseline
    nf = dom.nf

    if args.warmup:
        dom.step(args.warmup)
    dom.synchronize()
    dom.timing(True)
    t0 = time.perf_counter()
    dt = dom.step(args.steps)
    dom.synchronize()
    el = time.perf_counter() - t0
    tm = dom.timing()
    dom.timing(False)

    cells = nx * ny * nz
    value = cells * args.steps / el / 1e6
    # dominant kernel: the interior stage kernel (planes between the walls)
    interior_cells = nx * ny * (nz - 2)
    bytes_per_launch = alg_bytes_per_cell_step(nf) / 3.0 * interior_cells
    avg_ms = tm["stage_ms_total"] / max(1, tm["stage_launches"])
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    traffic = measured_traffic(mag, (nx, ny, nz))
    out = {
        "metric": "Mcell-RK3-steps/s",
        "value": round(value, 3),
        "unit": "Mcell-RK3-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"plain FP64 {'MHD+rotation' if mag else 'hydro'} RK3 step, "
                        f"{nx}x{ny}x{nz}, CFL dt, stratified polytrope Ra {ra:g}",
            "grid": [nx, ny, nz],
            "lMag": mag, "lRot": mag,
            "fields": nf,
            "parallelism": "single GPU",
            "last_dt": dt,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": "interior stage kernel",
            "kernel_avg_ms": round(avg_ms, 5),
            "alg_bytes_per_launch": int(bytes_per_launch),
        },
    }
    log(f"headline: {value:.1f} Mcell-steps/s, stage kernel {avg_ms:.4f} ms, {achieved:.0f} GB/s")
    # the drop-in call pattern: one step() per call, DT back to the host after
    # each (the reference's main loop prints it, 3dmhd.cpp:278-290)
    dom.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dom.step(1)
    el1 = time.perf_counter() - t0
    out["dropin"] = {"mcell_steps_per_s": round(cells * args.steps / el1 / 1e6, 1),
                     "ms_per_step": round(el1 / args.steps * 1e3, 4),
                     "vs_batched": round(el / el1, 4),
                     "note": f"{args.steps} calls of mhd3d_step(ctx, 1, 0, &DT)"}
    log(f"drop-in: {out['dropin']}")
    # the single-domain autotuner over the streaming variants (register budget
    # x schedule x launch size); the headline above runs the default variant,
    # which the traffic profile describes
    w0, s0, _ = dom.variant()
    trials = dom.autotune(steps=2)
    ok = [t for t in trials if t[3] > 0]
    best = min(ok, key=lambda t: t[3]) if ok else None
    dflt = [t for t in ok if (t[0], t[1], t[2]) == (w0, s0, 64.0)]
    out["autotune"] = {
        "candidates": len(trials),
        "best": None if best is None else {"waves": int(best[0]), "sched": int(best[1]),
                                           "planes_per_chunk": int(best[2]), "ms_per_step": round(best[3], 4)},
        "default_ms_per_step": round(dflt[0][3], 4) if dflt else None,
        "note": "mhd3d_autotune, 2 timed steps per candidate on the headline state; planes_per_chunk 0 = "
                "one round of the resident blocks"}
    log(f"autotune: {out['autotune']}")
    # the drop-in boundary hands over host arrays: time one upload and one
    # download of the whole state (host layout, ghosts included)
    if state is not None:
        t0 = time.perf_counter()
        dom.set_state(state)
        t_set = time.perf_counter() - t0
        t0 = time.perf_counter()
        dom.state()
        t_get = time.perf_counter() - t0
        out["boundary"] = {
            "set_state_ms": round(t_set * 1e3, 2), "get_state_ms": round(t_get * 1e3, 2),
            "pcie_inclusive_mcell_steps_per_s": round(cells * args.steps / (el + t_set + t_get) / 1e6, 1),
            "note": f"{args.steps} steps between one set_state and one get_state"}
    dom.close()
    out["ceiling"] = measure_ceiling()
    log(f"ceiling: {out['ceiling']}")
    if state is not None:
        out["cpu_baseline"] = run_cpu_baseline(cfg, state)
        state = None
        log(f"cpu baseline: {out['cpu_baseline']['value']:.2f} Mcell-steps/s")
    extra = {}
    for w in [w for w in args.workloads.split(",") if w]:
        fn = {"mixed": measure_mixed, "ensemble": measure_ensemble, "served": measure_served}.get(w)
        if fn is None:
            extra[w] = {"error": "unknown workload"}
            continue
        try:
            extra[w] = fn()
        except Exception as err:   # a failed side measurement must not lose the headline
            extra[w] = {"error": f"{type(err).__name__}: {err}"}
            log(f"{w}: {extra[w]['error']}")
    if extra:
        out["workloads"] = extra
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
