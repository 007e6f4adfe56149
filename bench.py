#!/usr/bin/env python3
"""Throughput of the per-step pseudo-particle loop on one MI355X.

Workload: the Na196 + Xe25+ (524.8 keV, b = 45 a0) crossing at production
shape: 8e6 Float32 pseudo-particles, nfine = 64 (130^3 fine and coarse
grids), dt = 0.5 a.u.; a "step" is one full step of the resident path (sort,
smoothed and CIC deposits, nested Poisson, mean field, field gather with the
projectile, Verlet).  By default the 700 steps of the crossing are run: 40
untimed warm-up steps, then 660 timed.  The crossing is then run again, untimed
in the metric, at 8e7 particles, the converged run the capture charge
q(< 8 a0) is compared with.  Every crossing runs to step 700 (the steps
after the timed ones untimed), so the capture charges are those of step 700
whatever --steps is.  Then the crossing from three more Ran2 seeds (the
spread of q), the isolated cluster's energy conservation, and the CPU
baseline.  Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--particles N] [--nfine n]
                    [--reference-particles M] [--no-cpu-baseline]
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
# HBM bytes per k_field_mfma launch of the step (forces and push) at 8e6 particles
# on 130^3, from two rocprofv3 counter passes over `vlasov.jl_amd/profile_step.py
# --steps 20` (the average of the 19 launches that push: FETCH_SIZE 331 153 KiB x 2.000
# + WRITE_SIZE 534 149 KiB x 1.000, the factors that make a 1 GiB device copy read and
# write 1 GiB); recorded with the commands in profiles/hbm_traffic_xenon_8M_130.md
FIELD_TRAFFIC_BYTES_8M = (331153 * 2.0 + 534149 * 1.0) * 1024
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
    CPU-runnable workload: the same Xe25+ crossing with `npart` particles on
    the `nfine` grid, its first `steps` steps (projectile advanced on the host
    as the reference's host path does)."""
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
    jel = va.Jellium.of(p.nions)
    t0 = time.perf_counter()
    for _ in range(steps):
        r = st.update_forces(rec, proj)
        fjel, _ = va.physics.jellium_force(proj, jel)
        proj.step(tuple(a + b for a, b in zip(fjel, r["red"][:3])), p.dt)
        rec, _ = st.verlet(r["srec"], r["force"])
    dt = time.perf_counter() - t0
    return {"value": npart * steps / dt, "unit": "particle-steps/s",
            "cores": int(O.lib().or_num_threads()), "kind": "port",
            "sample": f"oracle (FP64 C/numpy restatement; the reference's Julia CPU path cannot "
                      f"run here) on the CPU-runnable workload: {npart} particles, nfine={nfine} "
                      f"({2 * nfine + 2}^3), the same Xe25+ crossing, steps 1-{steps} of 700; "
                      f"{dt:.1f} s"}


def copy_calibration(mib: int):
    """A plain copy of `mib` MiB between two device buffers, five times: its
    byte counters per dispatch calibrate those of the step's kernels."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    log(f"copied {mib} MiB five times")


CROSSING_STEPS = 700     # ceil(140 / (0.4 * 0.5)), scripts/xenon.jl:306-309


def run_crossing(va, npart: int, nfine: int, mode: str, warmup: int, steps: int, profile: bool,
                 seed: int = -1, precision: str = "fp32", denoise=None):
    """The crossing at `npart` particles: `warmup` untimed steps, then `steps`
    timed ones, then, untimed, the rest of the 700 steps, after which the
    capture charge q(< 8 a0) is read (step 700 whatever `steps` is).  The
    energy budget is taken on the first warm-up step and on the last timed
    step (two interaction-energy sweeps, inside the timed region for the
    second).  With `profile`, the field kernel is timed with HIP events on the
    library's stream over the timed steps (two events per step), and every
    stage over the last 20 warm-up steps (outside the timed region: a region's
    two events cost about 4 us of device time)."""
    t0 = time.perf_counter()
    p = va.xenon_parameters(npart, nfine)
    proj = va.xenon_projectile()
    sim = va.Simulation(p, va.read_radial_profile(), projectile=proj, softening_mode=mode, seed=seed,
                        poisson_precision=precision)
    den = None
    if denoise is not None:
        # the learned denoiser between the deposit and the Poisson solve
        weights, tol = denoise
        den = va.denoiser.Denoiser(sim.ctx, weights, rho0=va.denoiser.jellium_density(sim.jellium))
        sim.attach_denoiser(den, tol)
    log(f"built in {time.perf_counter() - t0:.1f} s: {npart} particles, {sim.fine.n}^3 grids")
    nst = len(va._lib.STAGES)
    field_bit = 1 << va._lib.STAGES.index("field_kernel")

    def read_profile(nsteps):
        """ms per step of every stage timed over the last `nsteps` steps"""
        ms = np.zeros(nst)
        cnt = np.zeros(nst, dtype=np.int64)
        va._lib.call("vh_profile_read", sim.ctx.handle, ms, cnt)
        va._lib.call("vh_profile_enable", sim.ctx.handle, 0)
        return {s: ms[i] / nsteps for i, s in enumerate(va._lib.STAGES)}

    e_first = None
    stage_from = max(0, warmup - 20)
    for i in range(warmup):
        if profile and i == stage_from:
            sim.ctx.sync()
            va._lib.call("vh_profile_enable", sim.ctx.handle, va._lib.VH_PROF_ALL)
        b = sim.step(energy=(i == 0))
        if b is not None:
            e_first = b
    sim.ctx.sync()
    stages = read_profile(warmup - stage_from) if profile else {}
    log(f"warm-up {warmup} steps done")
    if profile:
        va._lib.call("vh_profile_enable", sim.ctx.handle, field_bit)
    sim.ctx.sync()
    t1 = time.perf_counter()
    e_last = None
    for i in range(steps):
        b = sim.step(energy=(i == steps - 1))
        if b is not None:
            e_last = b
        if (i + 1) % 100 == 0:
            log(f"step {warmup + i + 1}: x_ion = {sim.projectile.position[0]:.1f}")
    sim.ctx.sync()
    elapsed = time.perf_counter() - t1
    field_ms = read_profile(steps)["field_kernel"] if profile else 0.0
    timed_end = sim.step_count
    nocc, outcount = sim.acc.sort()
    # the rest of the crossing, untimed
    t2 = time.perf_counter()
    while sim.step_count < CROSSING_STEPS:
        sim.step()
    sim.ctx.sync()
    rest = time.perf_counter() - t2
    q = sim.enclosed_charge(8.0)
    nocc_end, outcount_end = sim.acc.sort()
    out = dict(elapsed=elapsed, stages=stages, field_ms=field_ms, q=q, nocc=nocc,
               outcount=outcount, timed_end=timed_end, step=sim.step_count,
               nocc_end=nocc_end, outcount_end=outcount_end, rest_s=rest,
               rest_steps=sim.step_count - timed_end, e_first=e_first, e_last=e_last,
               eloss_ev=sim.projectile.energy_loss() * va.physics.HARTREE_TO_EV,
               fallbacks=den.fallbacks if den is not None else None)
    sim.close()
    if den is not None:
        den.close()
    return out


def capture_spread(va, npart: int, nfine: int, mode: str, seeds, q_main: float):
    """q(< 8 a0) at step 700 of the same crossing sampled from other Ran2
    seeds: the run-to-run spread of the capture charge.  The step is
    deterministic (every reduction has a fixed order), so the spread the
    reference attributes to FP32 atomic ordering (README.md:149-152) is
    measured here over the sampling of the initial cloud instead."""
    qs = [q_main]
    for sd in seeds:
        log(f"capture spread: seed {sd}...")
        qs.append(run_crossing(va, npart, nfine, mode, 0, 0, False, seed=sd)["q"])
    a = np.array(qs)
    return {"particles": npart, "seeds": [-1] + list(seeds), "q8": qs, "mean": float(a.mean()),
            "std": float(a.std(ddof=1)), "range": float(a.max() - a.min())}


def precision_extension(va, npart: int, nfine: int, mode: str, warmup: int, steps: int):
    """Config 3: the crossing at `npart` particles on the nfine grid, once
    with FP32 rotating products and once with the Poisson solve's
    contractions in split-BF16 x3 on the matrix cores: ms/step over the timed
    steps and q(< 8 a0) at step 700 of each (the errors of the split solve
    against the FP64 oracle are the GPU tests' at this size)."""
    out = {"particles": npart, "grid": 2 * nfine + 2}
    for prec in ("fp32", "bf16x3"):
        log(f"precision extension: {prec}, {npart} particles, {2 * nfine + 2}^3...")
        r = run_crossing(va, npart, nfine, mode, warmup, steps, False, precision=prec)
        out[prec] = {"ms_per_step": 1000.0 * r["elapsed"] / steps, "capture_q8": r["q"],
                     "capture_step": r["step"], "projectile_energy_loss_eV": r["eloss_ev"]}
    out["capture_q_diff"] = abs(out["bf16x3"]["capture_q8"] - out["fp32"]["capture_q8"])
    out["speedup"] = out["fp32"]["ms_per_step"] / out["bf16x3"]["ms_per_step"]
    out["error_vs_fp64_at_258"] = ("tests/test_gpu_config3.py: fine potential 2.1e-5 (FP32 1.0e-6), "
                                   "field 2.5e-3 (FP32 9.8e-5), max |diff| / max |ref|")
    return out


def batching_extension(va, count: int, steps: int, nfine: int = 64):
    """Config 4: `count` independent collisions of the family (three
    projectiles, impact parameters uniform in 0-60 a0, particle counts
    log-uniform in 1e6-8e6) on the nfine grid.  Aggregate particle-steps/s
    of `steps` steps: every collision run alone one after the other (each
    step synchronised, the plain runs), then all of them as one batch whose
    step is one replayed HIP graph, after the autotuner has picked each
    member's work-unit size (and, between the two, the batch without the
    graph: every member's launches on its own streams, interleaved)."""
    fam = va.batch.collision_family(count=count, nfine=nfine)
    t0 = time.perf_counter()
    sims = [va.Simulation(c.params, va.read_radial_profile(), projectile=c.projectile) for c in fam]
    log(f"batch of {count} collisions built in {time.perf_counter() - t0:.1f} s")
    total = sum(c.params.nparticles for c in fam)
    out = {"collisions": count, "particles": total,
           "members": [{"projectile": c.label, "particles": c.params.nparticles,
                        "impact": round(c.impact, 2)} for c in fam]}
    try:
        for s in sims:      # one warm step each
            s.step()
            s.ctx.sync()
        t_seq = 0.0
        for s in sims:
            s.ctx.sync()
            t1 = time.perf_counter()
            for _ in range(steps):
                s.step()
            s.ctx.sync()
            t_seq += time.perf_counter() - t1
        out["sequential"] = {"particle_steps_per_s": total * steps / t_seq,
                             "ms_per_batch_step": 1000.0 * t_seq / steps}
        eager = va.batch.Batch(sims, graph=False)
        try:
            eager.step()
            eager.sync()
            t1 = time.perf_counter()
            for _ in range(steps):
                eager.step()
            eager.sync()
            t_e = time.perf_counter() - t1
        finally:
            eager.close()
        out["batched_streams"] = {"particle_steps_per_s": total * steps / t_e,
                                  "ms_per_batch_step": 1000.0 * t_e / steps}
        batch = va.batch.Batch(sims, graph=True)
        try:
            t1 = time.perf_counter()
            units, ms = batch.autotune(steps=4)
            out["autotune"] = {"candidates": [128, 256, 512, 1024], "units": units,
                               "s": time.perf_counter() - t1}
            batch.step()          # recorded
            batch.step()          # the other parity
            batch.sync()
            t1 = time.perf_counter()
            for _ in range(steps):
                batch.step()
            batch.sync()
            t_b = time.perf_counter() - t1
        finally:
            batch.close()
        out["batched_graph"] = {"particle_steps_per_s": total * steps / t_b,
                                "ms_per_batch_step": 1000.0 * t_b / steps}
        out["speedup"] = t_seq / t_b
        out["steps"] = steps
    finally:
        for s in sims:
            s.close()
    return out


def isolated_drift(va, npart: int, nfine: int, steps: int, every: int, denoise=None):
    """Energy conservation of the isolated Na196 cluster, the reference's
    validation observable (src/energy.jl:1-7): the budget every `every` steps
    over `steps` steps, no projectile (with `denoise`, the learned denoiser
    on every step's fine density)."""
    p = va.xenon_parameters(npart, nfine)
    sim = va.Simulation(p, va.read_radial_profile())
    den = None
    if denoise is not None:
        den = va.denoiser.Denoiser(sim.ctx, denoise[0], rho0=va.denoiser.jellium_density(sim.jellium))
        sim.attach_denoiser(den, denoise[1])
    hist = sim.run(steps, energy_every=every)
    sim.close()
    if den is not None:
        den.close()
    tot = np.array([b.total for b in hist])
    b1 = hist[0]
    # the reference's measure (test/runtests.jl:1138-1148): the range of the
    # total over the scale of the terms that make it up, which must stay
    # below 1e-3, and a total that neither only rises nor only falls
    scale = abs(b1.hartree) + abs(b1.meanfield) + b1.ions
    dif = np.diff(tot)
    return {"particles": npart, "steps": steps, "every": every, "first": float(tot[0]),
            "last": float(tot[-1]), "max_abs_drift": float(np.max(np.abs(tot - tot[0]))),
            "drift_rel": float((tot[-1] - tot[0]) / abs(tot[0])), "scale": float(scale),
            "range_over_scale": float((tot.max() - tot.min()) / scale),
            "oscillates": bool(not np.all(dif < 0) and not np.all(dif > 0)),
            "reference_bar": "range/scale < 1e-3 and not monotonic (test/runtests.jl:1138-1148)"}


def denoiser_extension(va, weights_path, npart: int, nfine: int, mode: str, warmup: int,
                       steps: int, seeds, tol: float, plain: dict, reference: dict | None,
                       drift_steps: int, reference_particles: int):
    """Config 5: the crossing at `npart` particles with the trained denoiser
    run on every step's fine density (matrix-core inference, residual charge
    check at `tol`, the plain density as fallback): ms/step over the timed
    steps, q(< 8 a0) at step 700 over the seeds of the plain spread, against
    the plain runs and the converged plain run; the isolated cluster's energy
    drift with the denoiser, beside the plain one at `npart` and at the
    converged run's count; the wall-clock fraction of the denoised run
    against the converged plain run."""
    w = np.load(weights_path, allow_pickle=False)
    report = Path(weights_path).with_suffix(".json")
    out = {"weights": Path(weights_path).name, "particles": npart, "grid": 2 * nfine + 2,
           "tol": tol}
    if report.exists():
        t = json.loads(report.read_text())
        out["training"] = {k: t[k] for k in ("pairs", "training_pairs", "collisions", "iters",
                                             "holdout_mse_plain", "holdout_mse_net", "holdout_gain")
                           if k in t}
    log("denoiser extension: timed crossing...")
    r = run_crossing(va, npart, nfine, mode, warmup, steps, False, denoise=(w, tol))
    qs = [r["q"]]
    fb = [r["fallbacks"]]
    for sd in seeds:
        log(f"denoiser extension: seed {sd}...")
        rs = run_crossing(va, npart, nfine, mode, 0, 0, False, seed=sd, denoise=(w, tol))
        qs.append(rs["q"])
        fb.append(rs["fallbacks"])
    a = np.array(qs)
    out["ms_per_step"] = 1000.0 * r["elapsed"] / steps
    out["plain_ms_per_step"] = plain["ms_per_step"]
    out["capture_q8"] = {"seeds": [-1] + list(seeds), "denoised": qs, "mean": float(a.mean()),
                         "std": float(a.std(ddof=1)) if len(a) > 1 else None,
                         "plain": plain["q8"], "plain_mean": float(np.mean(plain["q8"]))}
    out["fallbacks"] = fb
    out["projectile_energy_loss_eV"] = r["eloss_ev"]
    if reference is not None:
        q_ref = reference["capture_q8"]
        out["capture_q_err"] = {"denoised_seed_-1": abs(qs[0] - q_ref),
                                "denoised_mean": abs(float(a.mean()) - q_ref),
                                "plain_seed_-1": abs(plain["q8"][0] - q_ref),
                                "plain_mean": abs(float(np.mean(plain["q8"])) - q_ref),
                                "reference_q8": q_ref, "reference_particles": reference["particles"]}
        out["wall_clock_fraction"] = out["ms_per_step"] / reference["ms_per_step"]
    if drift_steps > 0:
        log("denoiser extension: isolated cluster energy drift...")
        out["isolated_energy_drift"] = isolated_drift(va, npart, nfine, drift_steps, 10,
                                                      denoise=(w, tol))
        if reference_particles > 0:
            log(f"isolated cluster energy drift at {reference_particles} particles (plain)...")
            out["isolated_energy_drift_reference"] = isolated_drift(va, reference_particles, nfine,
                                                                    drift_steps, 10)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=660)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--particles", type=int, default=8_000_000)
    ap.add_argument("--nfine", type=int, default=64)
    ap.add_argument("--mode", default="device", choices=["device", "ball"])
    ap.add_argument("--reference-particles", type=int, default=80_000_000,
                    help="particle count of the converged run the capture charge is compared "
                         "with (0: skip)")
    ap.add_argument("--drift-steps", type=int, default=700,
                    help="steps of the isolated-cluster energy-conservation run (0: skip)")
    ap.add_argument("--spread-seeds", type=int, default=3,
                    help="extra Ran2 seeds (-2, -3, ...) of the crossing whose step-700 capture "
                         "charges give the run-to-run spread (0: skip)")
    ap.add_argument("--precision-particles", type=int, default=80_000_000,
                    help="particles of the precision extension's crossings on nfine=128 "
                         "(FP32 and split-BF16 x3 Poisson; 0: skip)")
    ap.add_argument("--batch-collisions", type=int, default=16,
                    help="collisions of the batching extension's run (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=30)
    ap.add_argument("--denoiser-weights", type=Path, default=ROOT / "vlasov.jl_amd" / "data" /
                    "denoiser_na196_130.npy",
                    help="trained denoiser of the learned extension (missing file: skip)")
    ap.add_argument("--denoiser-tol", type=float, default=0.05,
                    help="residual charge check of the denoiser (relative)")
    ap.add_argument("--cpu-particles", type=int, default=100_000)
    ap.add_argument("--cpu-nfine", type=int, default=28)
    ap.add_argument("--cpu-steps", type=int, default=150)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run a plain device-to-device copy of this many MiB, five times "
                         "(the reference kernel for scaling HBM byte counters), and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the path has no "
              "multi-GPU decomposition in this build)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration_mib > 0:
        copy_calibration(args.copy_calibration_mib)
        return

    va = load_package()
    r = run_crossing(va, args.particles, args.nfine, args.mode, args.warmup, args.steps, True)
    elapsed, stages, outcount = r["elapsed"], r["stages"], r["outcount"]
    ms_step = 1000.0 * elapsed / args.steps
    # roofline of the dominant kernel: the smoothed-field contraction (FP32),
    # timed with HIP events over the timed steps
    field_ms = r["field_ms"]
    inside = args.particles - outcount
    achieved = inside * FIELD_FLOP_PER_PARTICLE / (field_ms * 1e-3) / 1e12 if field_ms else 0.0
    e0, e1 = r["e_first"], r["e_last"]
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
                               f"pseudo-particles, nfine={args.nfine} (130^3 fine + coarse), "
                               f"dt=0.5, steps {args.warmup + 1}-{args.warmup + args.steps} of 700",
                   "particles": args.particles, "grid": 2 * args.nfine + 2, "softening": args.mode},
        "roofline": {"bound": "mfma",
                     "engine": "x contraction on v_mfma_f32_16x16x32_f16 over fp16 splits (hi*hi + hi*lo + lo*hi, FP32 accumulation), y and z contractions on the FP32 VALU; achieved counts the FP32 algorithmic work of contract_spline_10 against the FP32 matrix-core peak. The launch also pushes its particles (Verlet, keys of the next sort), reading the records through the sort's permutation and writing them in slot order",
                     "kernel": "k_field_mfma",
                     "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS,
                     "traffic": FIELD_TRAFFIC_BYTES_8M if (args.particles, args.nfine) == (8_000_000, 64) else None,
                     "per_launch_flop": inside * FIELD_FLOP_PER_PARTICLE,
                     "avg_launch_ms": field_ms},
        "stages_ms": stages,
        "stages_ms_steps": f"{max(0, args.warmup - 20) + 1}-{args.warmup} (warm-up, every stage timed)",
        "capture_q8": r["q"],
        "capture_step": r["step"],
        "untimed_rest": {"steps": r["rest_steps"], "s": r["rest_s"]},
        "outcount": {f"step {r['timed_end']}": outcount, f"step {r['step']}": r["outcount_end"]},
        "occupied_cells": {f"step {r['timed_end']}": r["nocc"], f"step {r['step']}": r["nocc_end"]},
        "projectile_energy_loss_eV": r["eloss_ev"],
        # electrons only: the projectile's interaction is not in the budget
        "electron_energy": {"step1": e0.total if e0 else None, "last": e1.total if e1 else None},
    }
    seeds = [-(2 + i) for i in range(args.spread_seeds)]
    if args.spread_seeds > 0:
        line["capture_spread"] = capture_spread(va, args.particles, args.nfine, args.mode, seeds,
                                                r["q"])
    if args.drift_steps > 0:
        log("isolated cluster energy drift...")
        line["isolated_energy_drift"] = isolated_drift(va, args.particles, args.nfine,
                                                       args.drift_steps, 10)
    if args.reference_particles > 0:
        log(f"converged run: {args.reference_particles} particles...")
        rr = run_crossing(va, args.reference_particles, args.nfine, args.mode, args.warmup,
                          args.steps, False)
        if rr["step"] != r["step"]:
            raise RuntimeError("the two crossings must be compared at the same step")
        line["capture_reference"] = {
            "particles": args.reference_particles, "capture_q8": rr["q"], "capture_step": rr["step"],
            "capture_q_err": abs(r["q"] - rr["q"]),
            "ms_per_step": 1000.0 * rr["elapsed"] / args.steps,
            "projectile_energy_loss_eV": rr["eloss_ev"],
            "published_q8": {"8e6 (RTX 4070 / M1 Max)": [0.842, 0.840], "8e7 (M1 Max)": 0.81}}
    if args.denoiser_weights is not None and args.denoiser_weights.exists():
        plain = {"ms_per_step": ms_step,
                 "q8": line["capture_spread"]["q8"] if "capture_spread" in line else [r["q"]]}
        line["denoiser_extension"] = denoiser_extension(
            va, args.denoiser_weights, args.particles, args.nfine, args.mode, args.warmup,
            args.steps, seeds, args.denoiser_tol, plain, line.get("capture_reference"),
            args.drift_steps, args.reference_particles)
    else:
        log(f"denoiser extension skipped: no weights at {args.denoiser_weights}")
    if args.precision_particles > 0:
        line["precision_extension"] = precision_extension(va, args.precision_particles, 128,
                                                          args.mode, args.warmup, args.steps)
    if args.batch_collisions > 0:
        log(f"batching extension: {args.batch_collisions} collisions...")
        line["batching_extension"] = batching_extension(va, args.batch_collisions, args.batch_steps)
    if not args.no_cpu_baseline:
        log("cpu baseline...")
        line["cpu_baseline"] = cpu_baseline(va, args.cpu_nfine, args.cpu_particles, args.cpu_steps)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
