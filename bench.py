#!/usr/bin/env python3
"""Benchmark of the MI355X RK3 (magneto)hydrodynamics step.

Workload (one GPU): plain FP64 MHD with rotation on a 256x256x128 grid
(synthetic stratified convection: polytrope m=1, density contrast 10, Ra 1e5,
Pr 1, Ta 1e4, Pm 1, tanh z grid), CFL time step every step -- the reference's
``step()`` (3dmhdsub.h:31-153) as a drop-in on the GPU.  State and metrics are
resident in HBM before the timed region.

Prints one JSON line: Mcell-RK3-steps/s, the HBM roofline of the stage
kernel (algorithmic bytes / event-timed kernel duration) and the CPU baseline
(the C restatement in oracle/, timed on this host on a bounded sample).
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
    scheme: stage 1 reads U, writes U and Z; stage 2 reads U, Z, writes U, Z;
    stage 3 reads U, Z, writes U -> 10 field sweeps of 8 bytes."""
    return 10 * nf * 8


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def run_cpu_baseline(cfg, state, budget_s=12.0):
    """Time the oracle (C restatement of the reference path) on host cores."""
    from oracle import oracle as om
    from mhd3d.desc import Descriptor

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    desc = Descriptor(cfg)
    o = om.Oracle(desc, threads=threads)
    o.set_state(state)
    o.communicate()
    o.step()                       # warm-up (page-in)
    n = 0
    t0 = time.perf_counter()
    while True:
        o.step()
        n += 1
        el = time.perf_counter() - t0
        if el >= budget_s or n >= 50:
            break
    o.close()
    cells = cfg.nPx * cfg.nPy * cfg.nPz
    return {
        "value": cells * n / el / 1e6,
        "unit": "Mcell-RK3-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{n} CFL steps of the same {cfg.nPx}x{cfg.nPy}x{cfg.nPz} workload after 1 warm-up "
                  f"step, {threads} OpenMP threads on {cpu_model()}",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--size", default="256x256x128")
    ap.add_argument("--hydro", action="store_true", help="lMag off, lRot off")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU z-slab halo design is documented but not built)", file=sys.stderr)
        sys.exit(2)

    from mhd3d.desc import Descriptor
    from mhd3d.domain import Domain
    from mhd3d.initial import stratified_config, stratified_state

    nx, ny, nz = (int(v) for v in args.size.split("x"))
    mag = not args.hydro
    cfg = stratified_config(nx, ny, nz, Ra=1e5, Pr=1.0, lMag=mag, lRot=mag, Ta=1e4, Pm=1.0)
    desc = Descriptor(cfg)
    state = stratified_state(desc, seed=12345)
    dom = Domain(desc)
    dom.set_state(state)
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
                        f"{nx}x{ny}x{nz}, CFL dt, stratified polytrope Ra 1e5",
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
            "traffic": None,
            "kernel": "interior stage kernel",
            "kernel_avg_ms": round(avg_ms, 5),
            "alg_bytes_per_launch": int(bytes_per_launch),
        },
    }
    dom.close()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = run_cpu_baseline(cfg, state)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
