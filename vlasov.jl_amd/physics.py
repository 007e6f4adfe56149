"""Host-side physics of the step: the jellium, the projectile and its two
softening laws, the packed (cell, offset) form of a position, and the initial
Thomas-Fermi sampling.  Everything here is FP64 scalar or small-array work,
as it is on the reference's host (src/meanfield.jl, src/projectile.jl,
src/particles.jl:135-138, src/initial.jl).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import _lib

WIGNER_SEITZ_NA = 4.0                  # src/meanfield.jl:195
ELECTRON_MASS = 1.0                    # src/particles.jl:13
ELECTRON_CHARGE = -1.0                 # src/particles.jl:16
HARTREE_TO_EV = 27.2116                # src/projectile.jl:188
def _cbrt(x: float) -> float:
    return float(np.cbrt(np.float64(x)))


FERMI_COEFFICIENT = _cbrt(3.0 * (math.pi * math.pi))   # src/initial.jl:13
XC_EXCHANGE = -_cbrt(3.0 / math.pi)                    # src/meanfield.jl:250
XC_CORRELATION = -0.0333
XC_CORRELATION_SCALE = 11.4 * _cbrt(4.0 * math.pi / 3.0)

DATA_DIR = Path(__file__).resolve().parent / "data"

PACKED_DTYPE = np.dtype([("knode", "<i4", (3,)), ("delta", "<f4", (3,)),
                         ("pknode", "<i4", (3,)), ("pdelta", "<f4", (3,))])
assert PACKED_DTYPE.itemsize == 48


# ---------------------------------------------------------------------------
# jellium (src/meanfield.jl:211-244)
# ---------------------------------------------------------------------------
@dataclass
class Jellium:
    nions: float
    radius: float

    @classmethod
    def of(cls, nions: float, rs: float = WIGNER_SEITZ_NA):
        return cls(float(nions), float(rs) * _cbrt(float(nions)))

    def potential(self, r: float) -> float:
        return uniform_sphere_potential(self.nions, self.radius, r)


def ion_self_energy(jel: Jellium) -> float:
    """3 N^2 / 5 r0, the uniformly charged ball (src/energy.jl:51)."""
    return 3 * jel.nions ** 2 / (5 * jel.radius)


@dataclass
class EnergyBudget:
    """src/energy.jl:69-80: total = ions + kinetic + hartree + meanfield;
    `escaped` is the kinetic share beyond rcmax and enters no sum."""
    total: float
    kinetic: float
    hartree: float
    meanfield: float
    ions: float
    escaped: float = 0.0


def energy_budget(jel: Jellium, kinetic: float, hartree: float, total_interaction: float,
                  escaped: float = 0.0) -> EnergyBudget:
    """energy_budget (src/energy.jl:114-120): the mean-field term by difference,
    total interaction - 2 * hartree."""
    meanfield = total_interaction - 2 * hartree
    ions = ion_self_energy(jel)
    return EnergyBudget(ions + kinetic + hartree + meanfield, kinetic, hartree, meanfield, ions,
                        escaped)


def uniform_sphere_potential(Q, R, r):
    """-Q (3 - (r/R)^2) / 2R inside, -Q/r outside (src/meanfield.jl:235-237)."""
    return -Q * (3 - (r / R) ** 2) / (2 * R) if r < R else -Q / r


def xc_potential(rho):
    """LDA Dirac exchange + Gunnarsson-Lundqvist correlation (src/meanfield.jl:262-265)."""
    c = np.cbrt(rho)
    return XC_EXCHANGE * c + XC_CORRELATION * np.log1p(XC_CORRELATION_SCALE * c)


# ---------------------------------------------------------------------------
# softenings and projectile (src/projectile.jl)
# ---------------------------------------------------------------------------
@dataclass
class BallSoftening:
    radius: float

    @property
    def scale(self):
        return self.radius


@dataclass
class GaussianSoftening:
    sigma: float

    @property
    def scale(self):
        return self.sigma


_GAUSS_SERIES = (1 / 3, -1 / 10, 1 / 56, -1 / 432, 1 / 4224, -1 / 49920, 1 / 685440)
_SQRT_2_OVER_PI = math.sqrt(2 / math.pi)


def gaussian_force_kernel(r2: float, s: float) -> float:
    """src/projectile.jl:95-108."""
    u2 = r2 / (s * s)
    if u2 <= 0.25:
        p = 0.0
        for c in reversed(_GAUSS_SERIES):
            p = c + u2 * p
        return _SQRT_2_OVER_PI * p / (s * s * s)
    r = math.sqrt(r2)
    x = r / (math.sqrt(2.0) * s)
    return (math.erf(x) - _SQRT_2_OVER_PI * (r / s) * math.exp(-u2 / 2)) / (r2 * r)


def force_kernel(soft, r2: float) -> float:
    """Factor m with force = Q Q' m r (src/projectile.jl:79-80, 110-111)."""
    if isinstance(soft, BallSoftening):
        R = soft.radius
        return 1.0 / (r2 * math.sqrt(r2)) if r2 > R ** 2 else 1.0 / R ** 3
    return gaussian_force_kernel(r2, soft.sigma)


@dataclass
class Projectile:
    """Incident ion, Verlet-integrated (src/projectile.jl:141-168, 298-304)."""
    mass: float
    charge: float
    softening: object
    initial_energy: float
    position: tuple
    previous: tuple
    velocity: tuple

    @classmethod
    def of(cls, *, mass, charge, energy, impact=0.0, x0, dt, softening):
        v = (math.sqrt(2 * float(energy) / float(mass)), 0.0, 0.0)
        pos = (float(x0), float(impact), 0.0)
        prev = tuple(p - float(dt) * vv for p, vv in zip(pos, v))
        return cls(float(mass), float(charge), softening, float(energy), pos, prev, v)

    @property
    def scale(self) -> float:
        return float(self.softening.scale)

    def kinetic_energy(self) -> float:
        return self.mass * sum(v * v for v in self.velocity) / 2

    def energy_loss(self) -> float:
        return self.initial_energy - self.kinetic_energy()

    def step(self, force, dt: float):
        """new = 2q - q_prev + (dt^2/M) F; v = (new - q_prev)/2dt (src/projectile.jl:298-304)."""
        c = dt ** 2 / self.mass
        new = tuple(2 * p - q + c * f for p, q, f in zip(self.position, self.previous, force))
        self.velocity = tuple((a - b) / (2 * dt) for a, b in zip(new, self.previous))
        self.previous = self.position
        self.position = new
        return self


def jellium_force(proj: Projectile, jel: Jellium):
    """Force of the ionic background on the projectile, and its energy
    (src/projectile.jl:210-216; src/accelerator.jl:746-756)."""
    p = proj.position
    q = proj.charge
    r2 = p[0] * p[0] + p[1] * p[1] + p[2] * p[2]
    modf = jel.nions * q / r2 ** 1.5 if r2 > jel.radius ** 2 else q / WIGNER_SEITZ_NA ** 3
    return tuple(modf * x for x in p), q * jel.potential(math.sqrt(r2))


# ---------------------------------------------------------------------------
# packed positions (src/particles.jl:115-158)
# ---------------------------------------------------------------------------
def split(x: np.ndarray, x0: float, h: float):
    """Cell index (nearest knot, ties to even) and FP32 offset of each
    coordinate: k = round((x - x0)/h) + 1, delta = E(x - (x0 + (k-1) h))
    (src/particles.jl:135-138)."""
    x = np.asarray(x, dtype=np.float64)
    k = np.rint((x - x0) / h).astype(np.int64) + 1
    d = (x - (x0 + (k - 1) * h)).astype(np.float32)
    return k.astype(np.int32), d


def unsplit(k: np.ndarray, d: np.ndarray, x0: float, h: float) -> np.ndarray:
    """x = x0 + (k - 1) h + T(delta) (src/particles.jl:128-132)."""
    return x0 + (k.astype(np.float64) - 1.0) * h + d.astype(np.float64)


def packed_records(positions, previous, x0: float, h: float) -> np.ndarray:
    """PackedParticle{Float32} records of (current, previous) absolute positions."""
    n = len(positions)
    rec = np.zeros(n, dtype=PACKED_DTYPE)
    rec["knode"], rec["delta"] = split(positions, x0, h)
    rec["pknode"], rec["pdelta"] = split(previous, x0, h)
    return rec


def record_positions(rec: np.ndarray, x0: float, h: float, previous: bool = False) -> np.ndarray:
    if previous:
        return unsplit(rec["pknode"], rec["pdelta"], x0, h)
    return unsplit(rec["knode"], rec["delta"], x0, h)


# ---------------------------------------------------------------------------
# initial state (src/initial.jl)
# ---------------------------------------------------------------------------
@dataclass
class RadialProfile:
    quantiles: np.ndarray
    density: np.ndarray
    rmax: float


def read_radial_profile(directory: Path | str = DATA_DIR) -> RadialProfile:
    """hm1.dat (count, then radius quantiles) and rhoinit.dat (count, rmax,
    then the density) (src/initial.jl:53-68)."""
    d = Path(directory)
    with open(d / "hm1.dat") as f:
        n = int(f.readline())
        q = np.array([float(f.readline()) for _ in range(n)])
    with open(d / "rhoinit.dat") as f:
        n = int(f.readline())
        rmax = float(f.readline())
        rho = np.array([float(f.readline()) for _ in range(n)])
    return RadialProfile(q, rho, rmax)


@dataclass
class PotentialProfile:
    grid: np.ndarray
    potential: np.ndarray
    rmax: float
    pmax: float
    fermi: float


def read_potential_profile(path: Path | str) -> PotentialProfile:
    """pot.dat: intervals, `rmax pmax E_F`, then n+1 rows of which the first
    and third columns are kept (src/initial.jl:144-151)."""
    with open(path) as f:
        n = int(f.readline())
        rmax, pmax, fermi = (float(v) for v in f.readline().split())
        cols = [[float(v) for v in f.readline().split()] for _ in range(n + 1)]
    return PotentialProfile(np.array([c[0] for c in cols]), np.array([c[-1] for c in cols]),
                            rmax, pmax, fermi)


def sample_thomas_fermi(profile, npart: int, weight: float, seed: int = -1):
    """(positions, momenta), each npart x 3, from Ran2(seed), six uniforms per
    particle in the Fortran's order (src/initial.jl:94-113, 204-229)."""
    pos = np.zeros((npart, 3))
    mom = np.zeros((npart, 3))
    if isinstance(profile, RadialProfile):
        q = np.ascontiguousarray(profile.quantiles, dtype=np.float64)
        r = np.ascontiguousarray(profile.density, dtype=np.float64)
        _lib.call("vh_sample_radial", q, len(q), r, len(r), float(profile.rmax), int(npart),
                  float(weight), int(seed), pos, mom)
    else:
        g = np.ascontiguousarray(profile.grid, dtype=np.float64)
        v = np.ascontiguousarray(profile.potential, dtype=np.float64)
        _lib.call("vh_sample_rejection", g, v, len(g), float(profile.rmax), float(profile.pmax),
                  float(profile.fermi), int(npart), float(weight), int(seed), pos, mom)
    return pos, mom


class RadialSampler:
    """sample_thomas_fermi(::RadialProfile, npart, weight; rng = Ran2(seed))
    (src/initial.jl:94-113) drawn in slices: `next(count)` returns the next
    `count` (positions, momenta) of the sequence one call would draw, so that
    a cloud is built without holding all of its FP64 positions and momenta on
    the host at once (docs/src/device.md:196-200)."""

    def __init__(self, profile: "RadialProfile", weight: float, seed: int = -1):
        q = np.ascontiguousarray(profile.quantiles, dtype=np.float64)
        r = np.ascontiguousarray(profile.density, dtype=np.float64)
        h = C.c_void_p()
        _lib.call("vh_sampler_create", q, len(q), r, len(r), float(profile.rmax), float(weight),
                  int(seed), C.byref(h))
        self.handle = h

    def next(self, count: int):
        pos = np.zeros((count, 3))
        mom = np.zeros((count, 3))
        _lib.call("vh_sampler_next", self.handle, int(count), pos, mom)
        return pos, mom

    def close(self):
        if self.handle:
            _lib.call("vh_sampler_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ran2_sequence(count: int, seed: int = -1, consistent: bool = False) -> np.ndarray:
    out = np.zeros(count, dtype=np.float32)
    _lib.call("vh_ran2_sequence", int(seed), 1 if consistent else 0, int(count), out)
    return out
