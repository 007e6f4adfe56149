"""The learned denoiser: a small 3-D U-Net that maps the fine density
deposited by N pseudo-particles to the density 10 N deposit, trained on the
device on pairs the method itself generates, and run by `vh_step` between the
deposit and the Poisson solve with the plain density as its fallback.

No counterpart in the reference.  The network, its kernels (forward,
backward, Adam) and the residual charge check live in libvlasov_hip.so
(vh_denoiser_*); this module holds the handle, the weight files and the
generation of training pairs from seeded collisions.

    den = Denoiser.load(sim.ctx, DEFAULT_WEIGHTS, rho0=jellium_density(sim.jellium))
    sim.attach_denoiser(den, tol=0.02)
    for _ in range(700):
        sim.step()
"""
from __future__ import annotations

import ctypes as C
import math
from pathlib import Path

import numpy as np

from . import _lib

DEFAULT_WEIGHTS = Path(__file__).resolve().parent / "data" / "denoiser_na196_130.npy"


def param_count() -> int:
    n = C.c_int64()
    _lib.call("vh_denoiser_param_count", C.byref(n))
    return int(n.value)


def jellium_density(jellium) -> float:
    """rho0, the input scale: the jellium's mean density nions / (4/3 pi R^3)."""
    return float(jellium.nions / (4.0 / 3.0 * math.pi * jellium.radius ** 3))


class Denoiser:
    """One network on one context (vh_denoiser_*)."""

    def __init__(self, ctx, weights=None, *, seed: int = 0, rho0: float = 1.0):
        self.ctx = ctx
        self.nparams = param_count()
        self.rho0 = float(rho0)
        h = C.c_void_p()
        if weights is None:
            _lib.call("vh_denoiser_create", ctx.handle, None, 0, int(seed), self.rho0, C.byref(h))
        else:
            w = np.ascontiguousarray(weights, dtype=np.float32)
            if w.size != self.nparams:
                raise ValueError(f"{w.size} weights for a network of {self.nparams}")
            _lib.call("vh_denoiser_create", ctx.handle, w.ctypes.data, w.size, int(seed), self.rho0,
                      C.byref(h))
        self.handle = h

    @classmethod
    def load(cls, ctx, path=DEFAULT_WEIGHTS, rho0: float = 1.0):
        """Weights saved by `save` (a float32 .npy of the master weights)."""
        return cls(ctx, np.load(path, allow_pickle=False), rho0=rho0)

    def save(self, path):
        Path(path).parent.mkdir(parents=True, exist_ok=True)
        np.save(path, self.weights())

    def weights(self, effective: bool = False) -> np.ndarray:
        """The FP32 master weights, or (effective) the bf16 roundings the
        products use."""
        w = np.empty(self.nparams, dtype=np.float32)
        if effective:
            _lib.call("vh_denoiser_get_weights", self.handle, None, w.ctypes.data)
        else:
            _lib.call("vh_denoiser_get_weights", self.handle, w.ctypes.data, None)
        return w

    def set_weights(self, weights):
        """New master weights; the optimiser restarts."""
        w = np.ascontiguousarray(weights, dtype=np.float32).ravel()
        _lib.call("vh_denoiser_set_weights", self.handle, w, w.size)

    def eval(self, x, t, grad: bool = True):
        """One cube through the net: (o, loss, gradient or None)."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        D = x.shape[0]
        if x.shape != (D, D, D) or t.shape != x.shape:
            raise ValueError(f"cubes expected, got {x.shape} and {t.shape}")
        o = np.empty_like(x)
        g = np.empty(self.nparams, dtype=np.float32) if grad else None
        loss = C.c_double()
        _lib.call("vh_denoiser_eval", self.handle, D, x.ravel(), t.ravel(), o.ctypes.data,
                  C.byref(loss), g.ctypes.data if grad else None)
        return o, float(loss.value), g

    # --- the training set ---------------------------------------------------------
    def dataset_reserve(self, n: int, capacity: int):
        _lib.call("vh_denoiser_dataset_reserve", self.handle, int(n), int(capacity))

    def dataset_add(self, low_mesh, high_mesh):
        """Appends the pair (low_mesh.rho, high_mesh.rho) (DeviceMesh handles)."""
        _lib.call("vh_denoiser_dataset_add", self.handle, low_mesh.handle, high_mesh.handle)

    @property
    def dataset_size(self) -> int:
        n = C.c_int64()
        _lib.call("vh_denoiser_dataset_size", self.handle, C.byref(n))
        return int(n.value)

    def train(self, iters: int, crop: int = 32, batch: int = 4, first: int = 0, last=None,
              radius: int = 1 << 20, lr: float = 1e-3, seed: int = 0) -> np.ndarray:
        """`iters` Adam steps on random crops of pairs [first, last); the
        losses of every step."""
        last = self.dataset_size if last is None else last
        losses = np.zeros(iters)
        _lib.call("vh_denoiser_train", self.handle, int(iters), int(crop), int(batch), int(first),
                  int(last), int(radius), float(lr), int(seed), losses.ctypes.data)
        return losses

    def dataset_error(self, k: int, crop: int):
        """(plain, denoised) mean squared error against the 10 N density over
        the central cube of side `crop` of pair k."""
        a, b = C.c_double(), C.c_double()
        _lib.call("vh_denoiser_dataset_error", self.handle, int(k), int(crop), C.byref(a), C.byref(b))
        return float(a.value), float(b.value)

    # --- inference ----------------------------------------------------------------
    def apply(self, mesh, tol: float):
        """The net on mesh.rho in place, with the residual charge check."""
        _lib.call("vh_denoiser_apply", self.handle, mesh.handle, float(tol))

    def set_inference(self, path: str):
        """"mfma" (default): bf16 activations on the matrix cores; "fp32":
        the training kernels."""
        if path not in ("mfma", "fp32"):
            raise ValueError(f"unknown inference path {path!r}")
        _lib.call("vh_denoiser_set_inference", self.handle, 1 if path == "fp32" else 0)

    def set_support(self, xmin: float):
        """No correction where |rho / rho0| <= xmin (default 1e-4)."""
        _lib.call("vh_denoiser_set_support", self.handle, float(xmin))

    @property
    def fallbacks(self) -> int:
        n = C.c_int32()
        _lib.call("vh_denoiser_fallbacks", self.handle, C.byref(n))
        return int(n.value)

    def close(self):
        if self.handle:
            _lib.call("vh_denoiser_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def training_family(count: int, npart: int = 8_000_000, nfine: int = 64, seed: int = 2024,
                    bmax: float = 60.0, dt: float = 0.5, x0: float = -70.0):
    """The collisions the pairs come from: Na196 at `npart` pseudo-particles,
    the batching family's three projectiles in turn (protons at 100 keV,
    Ar8+ at 200 keV, Xe25+ at 500 keV), impact parameters uniform in
    [0, bmax] a0 from a seeded numpy generator."""
    from .batch import KEV, PROJECTILES, Collision
    from .physics import Projectile
    from .simulation import xenon_parameters

    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        b = float(rng.uniform(0.0, bmax))
        name, mass, charge, kev, soft = PROJECTILES[i % len(PROJECTILES)]
        proj = Projectile.of(mass=mass, charge=charge, energy=kev * KEV, impact=b, x0=x0, dt=dt,
                             softening=soft())
        out.append(Collision(xenon_parameters(npart, nfine, dt), proj, name, b))
    return out


def generate_pairs(den: Denoiser, collisions, steps: int = 700, every: int = 10, ratio: int = 10,
                   profile=None, seed0: int = 100, log=None):
    """Training pairs from the method itself: for each collision (a
    batch.Collision), an N-particle run and a ratio N-particle run of the same
    seeded collision, stepped side by side; every `every` steps the pair of
    their fine densities after the step's deposit goes into the dataset.
    Returns the pair indices of each collision."""
    import dataclasses

    from .physics import read_radial_profile
    from .simulation import Simulation, SimulationParameters

    profile = read_radial_profile() if profile is None else profile
    spans = []
    for ci, col in enumerate(collisions):
        p = col.params
        hp = SimulationParameters(**{**p.__dict__, "nparticles": p.nparticles * ratio})
        first = den.dataset_size
        lo = Simulation(p, profile, projectile=dataclasses.replace(col.projectile),
                        seed=-(seed0 + ci))
        hi = Simulation(hp, profile, projectile=dataclasses.replace(col.projectile),
                        seed=-(seed0 + 1000 + ci))
        try:
            for s in range(steps):
                lo.step()
                hi.step()
                if (s + 1) % every == 0:
                    den.dataset_add(lo.dmf, hi.dmf)
        finally:
            lo.close()
            hi.close()
        spans.append((first, den.dataset_size))
        if log is not None:
            log(f"collision {ci} ({col.label}, b={col.impact:.1f}, N={p.nparticles}): "
                f"pairs {first}..{den.dataset_size}")
    return spans
