#!/usr/bin/env python3
"""Throughput of the per-step pseudo-particle loop on one MI355X.

Workload: the Na196 + Xe25+ (524.8 keV, b = 45 a0) crossing at production
shape: 8e6 Float32 pseudo-particles, nfine = 64 (130^3 fine and coarse
grids), dt = 0.5 a.u.; a "step" is one full step of the resident path (sort,
smoothed and CIC deposits, nested Poisson, mean field, field gather with the
projectile, Verlet).  By default the 700 steps of the crossing are run: 40
untimed warm-up steps, then 660 timed.  Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--particles N] [--nfine n]
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FIELD_FLOP_PER_PARTICLE = 4900.0      # 2300 FMA + 300 MUL of contract_spline_10
METRIC = "ms/step & particle-steps/s @8M,130^3; capture-q err vs 80M plain run"


def load_package():
    pkg = ROOT / "vlasov.jl_amd"
    spec = importlib.util.spec_from_file_location("vlasov_amd", pkg / "__init__.py",
                                                  submodule_search_locations=[str(pkg)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["vlasov_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(va, nfine: int, npart: int, steps: int):
    """The oracle (FP64 restatement, C + numpy) on a bounded sample of the
    same workload: same geometry, `npart` particles, `steps` steps."""
    sys.path.insert(0, str(ROOT))
    from oracle import oracle as O
    O.build()
    p = va.xenon_parameters(npart, nfine)
    fine, coarse = va.simulation.build_meshes(p)
    sm = va.grids.GaussianSmoothing.of(fine.axis)
    w = p.nelectrons / npart
    pos, mom = va.sample_thomas_fermi(va.read_radial_profile(), npart, w)
    x0 = float(fine.axis.knots[0])
    h = float((fine.axis.knots[-1] - fine.axis.knots[0]) / (fine.axis.nknots - 1))
    rec = va.physics.packed_records(pos, pos - (p.dt / w) * mom, x0, h)
    st = O.OracleStep(fine, coarse, sm, va.Jellium.of(p.nions), w, w, p.dt, p.rcmax)
    proj = va.xenon_projectile()
    t0 = time.perf_counter()
    for _ in range(steps):
        r = st.update_forces(rec, proj)
        rec, _ = st.verlet(r["srec"], r["force"])
    dt = time.perf_counter() - t0
    return {"value": npart * steps / dt, "unit": "particle-steps/s",
            "cores": int(O.lib().or_num_threads()), "kind": "port",
            "sample": f"oracle (FP64 C/numpy restatement), {npart} particles, {steps} steps, "
                      f"nfine={nfine}, same Xe25+ crossing; {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=660)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--particles", type=int, default=8_000_000)
    ap.add_argument("--nfine", type=int, default=64)
    ap.add_argument("--mode", default="device", choices=["device", "ball"])
    ap.add_argument("--cpu-particles", type=int, default=400_000)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the path has no "
              "multi-GPU decomposition in this build)", file=sys.stderr)
        sys.exit(2)

    va = load_package()
    t0 = time.perf_counter()
    p = va.xenon_parameters(args.particles, args.nfine)
    proj = va.xenon_projectile()
    sim = va.Simulation(p, va.read_radial_profile(), projectile=proj, softening_mode=args.mode)
    log(f"built in {time.perf_counter() - t0:.1f} s: {args.particles} particles, "
        f"{sim.fine.n}^3 grids")
    for i in range(args.warmup):
        sim.step()
    sim.ctx.sync()
    log(f"warm-up {args.warmup} steps done")
    va._lib.call("vh_profile_enable", sim.ctx.handle, 1)
    ms = np.zeros(len(va._lib.STAGES))
    cnt = np.zeros(len(va._lib.STAGES), dtype=np.int64)
    sim.ctx.sync()
    t1 = time.perf_counter()
    for i in range(args.steps):
        sim.step()
        if (i + 1) % 100 == 0:
            log(f"step {args.warmup + i + 1}: x_ion = {sim.projectile.position[0]:.1f}")
    sim.ctx.sync()
    elapsed = time.perf_counter() - t1
    va._lib.call("vh_profile_read", sim.ctx.handle, ms, cnt)
    va._lib.call("vh_profile_enable", sim.ctx.handle, 0)
    ms_step = 1000.0 * elapsed / args.steps
    stages = {s: (ms[i] / cnt[i] if cnt[i] else 0.0) for i, s in enumerate(va._lib.STAGES)}
    q = sim.enclosed_charge(8.0)
    nocc, outcount = sim.acc.sort()
    # roofline of the dominant kernel: the smoothed-field contraction (FP32 VALU)
    field_ms = stages["field_kernel"]
    inside = args.particles - outcount
    achieved = inside * FIELD_FLOP_PER_PARTICLE / (field_ms * 1e-3) / 1e12 if field_ms else 0.0
    line = {
        "metric": METRIC,
        "value": args.particles * args.steps / elapsed,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: Thomas-Fermi Na196 cloud sampled from the reference's hm1.dat/rhoinit.dat "
                "with its ran2 seed -1",
        "config": {"workload": f"Na196 + Xe25+ 524.8 keV b=45 a0 crossing, {args.particles} "
                               f"pseudo-particles, nfine={args.nfine} ({sim.fine.n}^3 fine + coarse), "
                               f"dt=0.5, steps {args.warmup + 1}-{args.warmup + args.steps} of 700",
                   "particles": args.particles, "grid": sim.fine.n, "softening": args.mode},
        "roofline": {"bound": "mfma", "engine": "VALU FP32 fma (packed FP32 disabled)",
                     "kernel": "k_field",
                     "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None,
                     "per_launch_flop": inside * FIELD_FLOP_PER_PARTICLE,
                     "avg_launch_ms": field_ms},
        "stages_ms": stages,
        "capture_q8": q,
        "step": sim.step_count,
        "outcount": outcount,
        "occupied_cells": nocc,
    }
    sim.close()
    if not args.no_cpu_baseline:
        log("cpu baseline...")
        line["cpu_baseline"] = cpu_baseline(va, args.nfine, args.cpu_particles, args.cpu_steps)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
