"""Steps the Na196 + Xe25+ crossing with every stage timed, for profilers.

    python vlasov.jl_amd/profile_step.py --particles 8000000 --nfine 64 --steps 25

With stage timing on, a step runs its stages one after the other on the
context's main stream (the form hardware-counter passes of rocprofv3 need);
the kernels are the step's own.  Prints the mean ms of every stage.
"""
from __future__ import annotations

import argparse
import importlib.util
import sys
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


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--particles", type=int, default=8_000_000)
    ap.add_argument("--nfine", type=int, default=64)
    ap.add_argument("--steps", type=int, default=25)
    ap.add_argument("--seed", type=int, default=-1)
    ap.add_argument("--field-fp32", action="store_true",
                    help="the field contraction's first stage in FP32 products (VH_OPT_FIELD_FP32)")
    args = ap.parse_args(argv)
    va = load_package()
    L = va._lib
    p = va.xenon_parameters(args.particles, args.nfine)
    sim = va.Simulation(p, va.read_radial_profile(), projectile=va.xenon_projectile(), seed=args.seed)
    try:
        sim.acc.set_field_fp32(args.field_fp32)
        L.call("vh_profile_enable", sim.ctx.handle, L.VH_PROF_ALL)
        for _ in range(args.steps):
            sim.step()
        ms = np.zeros(len(L.STAGES))
        cnt = np.zeros(len(L.STAGES), dtype=np.int64)
        L.call("vh_profile_read", sim.ctx.handle, ms, cnt)
        for i, s in enumerate(L.STAGES):
            print(f"{s:15s} {ms[i] / args.steps:8.4f} ms/step", flush=True)
    finally:
        sim.close()


if __name__ == "__main__":
    main()
