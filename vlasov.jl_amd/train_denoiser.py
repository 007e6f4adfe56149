"""Trains the learned denoiser on pairs the method generates, on the GPU.

    python vlasov.jl_amd/train_denoiser.py --collisions 24 --out weights.npy

For each collision of the family (Na196 at 8e6 pseudo-particles on the
130^3 grids; protons at 100 keV, Ar8+ at 200 keV and Xe25+ at 500 keV in
turn, impact parameters uniform in 0-60 a0, from x = -70 a0 at dt = 0.5),
an 8e6 run and an 8e7 run with their own seeds are stepped side by side
through the 700-step crossing and their fine densities, after every tenth
step's deposit, are kept in device memory as a training pair.  The last
`--holdout` collisions are not trained on: they measure what the network
gains (mean squared error against the 8e7 density, plain and denoised).
Training is Adam on random crops of the training pairs, with the kernels of
the library.  Writes the FP32 master weights (.npy) and a JSON report beside
them.

The recipe (measured on 4 collisions: held-out error halved after 600 Adam
steps at learning rate 3e-4, where 1e-3 collapsed the net to the identity):
crops of 48^3, four a step; two chunks in three centred within 12 points of
the grid's centre (the cluster), one in three within 40 (the projectile's
path); learning rate 3e-4, then 1e-4 from 60 % and 3e-5 from 85 % of the
steps.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import math
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def load_package():
    if "vlasov_amd" in sys.modules:
        return sys.modules["vlasov_amd"]
    spec = importlib.util.spec_from_file_location("vlasov_amd", HERE / "__init__.py",
                                                  submodule_search_locations=[str(HERE)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["vlasov_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(msg, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--collisions", type=int, default=24, help="collisions of the family")
    ap.add_argument("--holdout", type=int, default=2, help="last collisions kept for validation")
    ap.add_argument("--particles", type=float, default=8e6, help="N (the pairs are N and 10 N)")
    ap.add_argument("--nfine", type=int, default=64)
    ap.add_argument("--steps", type=int, default=700, help="steps of each crossing")
    ap.add_argument("--every", type=int, default=10, help="a pair every this many steps")
    ap.add_argument("--iters", type=int, default=3000, help="Adam steps")
    ap.add_argument("--crop", type=int, default=48)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=3e-4)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--out", type=Path, required=True)
    args = ap.parse_args(argv)

    va = load_package()
    D_ = va.denoiser
    t0 = time.perf_counter()
    family = D_.training_family(args.collisions, int(args.particles), args.nfine, seed=args.seed)
    ctx = va.Context(0)
    jel = va.Jellium.of(196.0)
    den = D_.Denoiser(ctx, seed=args.seed, rho0=D_.jellium_density(jel))
    n = 2 * args.nfine + 2
    per = args.steps // args.every
    den.dataset_reserve(n, per * len(family))
    spans = D_.generate_pairs(den, family, steps=args.steps, every=args.every, log=log)
    ntrain = spans[len(family) - args.holdout - 1][1] if args.holdout < len(family) else spans[-1][1]
    log(f"{den.dataset_size} pairs ({ntrain} for training) in {time.perf_counter() - t0:.0f} s")

    held = list(range(ntrain, den.dataset_size, max(1, (den.dataset_size - ntrain) // 16)))

    def validate():
        e = [den.dataset_error(k, 96) for k in held]
        return float(np.mean([a for a, _ in e])), float(np.mean([b for _, b in e]))

    before = validate() if held else (math.nan, math.nan)
    log(f"held-out error before training: plain {before[0]:.4e}, net {before[1]:.4e}")
    losses, curve = [], []
    chunk = 150
    t1 = time.perf_counter()
    done = 0
    k = 0
    while done < args.iters:
        it = min(chunk, args.iters - done)
        radius = 40 if k % 3 == 2 else 12
        frac = done / args.iters
        lr = args.lr * (1.0 if frac < 0.6 else 1 / 3 if frac < 0.85 else 0.1)
        losses += list(den.train(it, crop=args.crop, batch=args.batch, first=0, last=ntrain,
                                 radius=radius, lr=lr, seed=args.seed + k))
        done += it
        k += 1
        if k % 4 == 0 or done == args.iters:
            v = validate() if held else (math.nan, math.nan)
            curve.append({"iter": done, "holdout_mse_net": v[1]})
            log(f"iter {done}: loss {np.mean(losses[-it:]):.4e}, held-out gain {v[0] / v[1]:.3f} "
                f"({time.perf_counter() - t1:.0f} s)")
    after = validate() if held else (math.nan, math.nan)
    log(f"held-out error after training: plain {after[0]:.4e}, net {after[1]:.4e}")
    den.save(args.out)
    report = {
        "pairs": den.dataset_size, "training_pairs": ntrain, "collisions": len(family),
        "holdout_collisions": args.holdout, "particles": int(args.particles), "ratio": 10,
        "grid": n, "every": args.every, "steps": args.steps,
        "members": [{"projectile": c.label, "impact": round(c.impact, 2)} for c in family],
        "iters": args.iters, "crop": args.crop, "batch": args.batch, "lr": args.lr,
        "loss_first_250": float(np.mean(losses[:250])), "loss_last_250": float(np.mean(losses[-250:])),
        "holdout_mse_plain": after[0], "holdout_mse_net": after[1],
        "holdout_gain": after[0] / after[1] if after[1] > 0 else None,
        "holdout_crop": 96, "curve": curve,
        "train_s": time.perf_counter() - t1, "total_s": time.perf_counter() - t0,
    }
    args.out.with_suffix(".json").write_text(json.dumps(report, indent=1) + "\n")
    log(json.dumps(report))
    den.close()
    ctx.close()


if __name__ == "__main__":
    main()
