"""The batching extension: many independent collisions on one GPU, stepped
together as one HIP graph, each with its work-unit size picked by an
autotuner.

No counterpart in the reference, which runs one collision per process
(scripts/xenon.jl, scripts/film_*.jl).  The family of collisions follows
its scripts' projectiles: protons with a Gaussian softening of width 1
(scripts/film_proton.jl:77-79), Ar8+ with a ball of radius 2
(scripts/film_argon.jl:87-89) and Xe25+ with a ball of radius 5
(scripts/xenon.jl:337-340), all from x0 = -70 a0 at dt = 0.5 on the Na196
cluster of scripts/xenon.jl:330-333.

    sims = [Simulation(p, profile, projectile=proj) for p, proj in family]
    batch = Batch(sims)
    batch.autotune()
    for _ in range(steps):
        batch.step()
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _lib
from .physics import HARTREE_TO_EV, BallSoftening, GaussianSoftening, Projectile
from .simulation import Simulation, SimulationParameters, xenon_parameters

KEV = 1000.0 / HARTREE_TO_EV
AMU_ME = 1836.154

# (name, mass in electron masses, charge, energy in keV, softening)
PROJECTILES = (
    ("proton 100 keV", AMU_ME, 1.0, 100.0, lambda: GaussianSoftening(1.0)),
    ("Ar8+ 200 keV", 40.0 * AMU_ME, 8.0, 200.0, lambda: BallSoftening(2.0)),
    ("Xe25+ 500 keV", 131.3 * AMU_ME, 25.0, 500.0, lambda: BallSoftening(5.0)),
)


@dataclass
class Collision:
    """One member of the family: its parameters and projectile."""
    params: SimulationParameters
    projectile: Projectile
    label: str
    impact: float


def collision_family(count: int = 16, nmin: float = 1e6, nmax: float = 8e6, nfine: int = 64,
                     bmax: float = 60.0, seed: int = 2024, dt: float = 0.5, x0: float = -70.0):
    """`count` collisions: impact parameters uniform in [0, bmax] a0, the
    three projectiles in turn, particle counts log-uniform in [nmin, nmax]
    (a seeded numpy generator), on the nfine grid."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        npart = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        b = float(rng.uniform(0.0, bmax))
        name, mass, charge, kev, soft = PROJECTILES[i % len(PROJECTILES)]
        proj = Projectile.of(mass=mass, charge=charge, energy=kev * KEV, impact=b, x0=x0, dt=dt,
                             softening=soft())
        out.append(Collision(xenon_parameters(npart, nfine, dt), proj, name, b))
    return out


class Batch:
    """Independent simulations on one device stepped together (vh_batch_*):
    with `graph`, one HIP graph per step that forks into every member's
    streams and joins back, recorded once and replayed."""

    def __init__(self, sims: list[Simulation], graph: bool = True):
        if not sims:
            raise ValueError("an empty batch")
        self.sims = list(sims)
        n = len(sims)
        accs = (C.c_void_p * n)(*[s.acc.handle.value for s in sims])
        fines = (C.c_void_p * n)(*[s.dmf.handle.value for s in sims])
        coarses = (C.c_void_p * n)(*[s.dmc.handle.value for s in sims])
        params = np.array([[s.weight, s.params.dt, s.mass, s.params.rcmax, s.jellium.nions,
                            s.jellium.radius] for s in sims], dtype=np.float64).ravel()
        h = C.c_void_p()
        _lib.call("vh_batch_create", n, accs, fines, coarses, params, 1 if graph else 0, C.byref(h))
        self.handle = h
        self.units = [256] * n

    @property
    def particles(self) -> int:
        return sum(s.params.nparticles for s in self.sims)

    def step(self):
        """One step of every member; nothing waits for the device."""
        _lib.call("vh_batch_step", self.handle)
        for s in self.sims:
            s.step_count += 1
            s._proj_stale = s._proj_on_device

    def sync(self):
        _lib.call("vh_batch_sync", self.handle)

    def autotune(self, candidates=(128, 256, 512, 1024), steps: int = 3):
        """Each member's work-unit size, picked by timing `steps` real steps
        with each candidate (vh_autotune_unit); returns (chosen sizes, ms per
        step of every candidate per member).  The tuning steps advance the
        simulations like any other step."""
        cand = np.ascontiguousarray(candidates, dtype=np.int32)
        timings = []
        for i, s in enumerate(self.sims):
            ms = np.zeros(len(cand))
            chosen = C.c_int()
            _lib.call("vh_autotune_unit", s.acc.handle, s.dmf.handle, s.dmc.handle,
                      float(s.weight), float(s.params.dt), float(s.mass), float(s.params.rcmax),
                      float(s.jellium.nions), float(s.jellium.radius), cand, len(cand),
                      int(steps), C.byref(chosen), ms)
            s.step_count += len(cand) * (steps + 1)
            s._proj_stale = s._proj_on_device
            self.units[i] = int(chosen.value)
            timings.append(ms)
        return list(self.units), timings

    def close(self):
        if self.handle:
            _lib.call("vh_batch_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
