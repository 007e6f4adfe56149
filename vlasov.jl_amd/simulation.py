"""The simulation driver: the reference's `Simulation` on its resident device
path (src/simulation.jl:171-277, 382-416, 472-531), with every stage of the
step running in libvlasov_hip.so.

    sim = Simulation(params, read_radial_profile(), projectile=proj)
    for s in range(nsteps):
        sim.step()
    q = sim.enclosed_charge(8.0)

The constructor does what the reference's does once on the host in FP64:
axes, meshes, eigen-factorisations, smoothing tables, the Thomas-Fermi
sample, and the leapfrog priming (which needs one force evaluation).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import grids
from .device import Context, DeviceAccelerator, DeviceMesh, poisson_nested
from . import _lib
from .physics import (ELECTRON_MASS, Jellium, Projectile, RadialProfile, RadialSampler,
                      energy_budget, jellium_force, packed_records, record_positions,
                      sample_thomas_fermi)


@dataclass
class SimulationParameters:
    """src/simulation.jl:26-38."""
    nfine: int = 28
    ninner: int = 14
    nouter: int = 14
    rcluster: float = 50.0
    rbox: float = 150.0
    nions: float = 196.0
    nelectrons: float = 196.0
    nparticles: int = 20_000
    nsteps: int = 10
    dt: float = 1.0
    rcmax: float = 100.0


def xenon_parameters(nparticles: int, nfine: int, dt: float = 0.5) -> SimulationParameters:
    """The Na196 + Xe25+ crossing of scripts/xenon.jl:303-333."""
    ninner, nouter = grids.coarse_grid_sizes(nfine)
    return SimulationParameters(nfine=nfine, ninner=ninner, nouter=nouter, rcluster=78.0,
                                rbox=235.0, nions=196.0, nelectrons=196.0,
                                nparticles=nparticles, nsteps=0, dt=dt)


def xenon_projectile(dt: float = 0.5, v: float = 0.40, x0: float = -70.0, impact: float = 45.0,
                     softening=None) -> Projectile:
    """Xe25+ at v = 0.40 a.u., b = 45 a0, from x = -70 a0, ball softening of
    radius 5 (scripts/xenon.jl:306-340)."""
    from .physics import BallSoftening
    mass = 131.3 * 1836.154
    return Projectile.of(mass=mass, charge=25.0, energy=0.5 * mass * v ** 2, impact=impact, x0=x0,
                         dt=dt, softening=softening or BallSoftening(5.0))


def build_meshes(p: SimulationParameters):
    fine_ax = grids.uniform_axis(-p.rcluster, p.rcluster, p.nfine)
    coarse_ax = grids.stretched_axis(p.rcluster, p.rbox, p.ninner // 2, (p.nouter + 2) // 2)
    fine, coarse = grids.SplineMesh.of(fine_ax), grids.SplineMesh.of(coarse_ax)
    grids.nested_meshes(fine, coarse)
    if fine.n != coarse.n:
        raise ValueError("rho must cover the whole collocation grid: the two levels carry "
                         f"{fine.n} and {coarse.n} basis functions (nfine must be even)")
    return fine, coarse


class Simulation:
    """State of one run on one device."""

    def __init__(self, params: SimulationParameters, profile, *, projectile: Projectile | None = None,
                 seed: int = -1, softening_mode: str = "device", consistent_startup: bool = False,
                 device: int = 0, tables=None, positions=None, momenta=None,
                 resident_projectile: bool = True, slice_size: int = 1 << 20,
                 poisson_precision: str = "fp32"):
        self.params = params
        self._projectile = projectile
        self._proj_on_device = False
        self._proj_stale = False
        self.softening_mode = softening_mode
        if tables is None:
            fine, coarse = build_meshes(params)
            sm = grids.GaussianSmoothing.of(fine.axis)
        else:
            fine, coarse, sm = tables
        self.fine, self.coarse, self.smoothing = fine, coarse, sm
        self.jellium = Jellium.of(params.nions)
        self.weight = params.nelectrons / params.nparticles
        self.mass = ELECTRON_MASS * self.weight
        if positions is None and not isinstance(profile, RadialProfile):
            positions, momenta = sample_thomas_fermi(profile, params.nparticles, self.weight, seed)
        self.ctx = Context(device)
        self.dmf = DeviceMesh(self.ctx, fine)
        self.dmc = DeviceMesh(self.ctx, coarse)
        # the precision extension: the Poisson solve's contractions in
        # split-BF16 x3 on the matrix cores (the S^-1 transforms stay FP32)
        if poisson_precision not in ("fp32", "bf16x3"):
            raise ValueError(f"unknown Poisson precision {poisson_precision!r}")
        if poisson_precision == "bf16x3":
            for d in (self.dmf, self.dmc):
                d.set_bf16x3(solve=True, transforms=False)
        self.acc = DeviceAccelerator(self.ctx, fine.axis, sm, params.nparticles)
        self.x0, self.h = self.acc.x0, self.acc.h
        self.nout = 0
        self.outcount = 0
        self.red = np.zeros(4)
        self.step_count = 0
        self.diagnostics = None
        self._denoiser, self._denoise_tol = None, 0.0
        if positions is None:
            # the Thomas-Fermi sample drawn and uploaded slice by slice
            sampler = RadialSampler(profile, self.weight, seed)
            try:
                slices = (sampler.next(min(slice_size, params.nparticles - i0))
                          for i0 in range(0, params.nparticles, slice_size))
                self._prime_leapfrog(slices, consistent_startup, slice_size)
            finally:
                sampler.close()
        else:
            n = params.nparticles
            slices = ((positions[i0:i0 + slice_size], momenta[i0:i0 + slice_size])
                      for i0 in range(0, n, slice_size))
            self._prime_leapfrog(slices, consistent_startup, slice_size)
        if projectile is not None and resident_projectile:
            # from here the projectile is integrated on the device, and a step
            # waits for nothing on the host
            self.acc.set_projectile(projectile, softening_mode, self.jellium)
            self._proj_on_device = True

    @property
    def projectile(self) -> Projectile | None:
        """The projectile, brought up to date from the device when it lives
        there (one synchronisation when it has moved since the last read)."""
        p = self._projectile
        if p is not None and self._proj_on_device and self._proj_stale:
            st = self.acc.projectile_state()
            p.position, p.previous, p.velocity = tuple(st[0:3]), tuple(st[3:6]), tuple(st[6:9])
            self._proj_stale = False
        return p

    # --- construction ------------------------------------------------------------
    def _prime_leapfrog(self, slices, consistent, slice_size):
        """prime_leapfrog! (src/simulation.jl:233-277): q(-dt/2) from the
        momenta (half_step_back, src/particles.jl:441-442), with q(0) carried
        in the previous half through the sort; the forces there; then q(-dt)
        by the thesis's (dimensionally dubious) coefficient dt/M, or dt^2/4M
        with `consistent` (full_step_back, src/particles.jl:466-469).  The
        cloud goes up slice by slice (each slice's half step on the host in
        FP64, as the reference's map does), and the second stage runs on the
        device (vh_prime), so the whole cloud never crosses back."""
        M, dt = self.mass, self.params.dt
        i0 = 0
        for pos, mom in slices:
            half = pos - (dt / (2 * M)) * mom
            rec = packed_records(half, pos, self.x0, self.h)
            _lib.call("vh_particles_upload_range", self.acc.handle, i0, len(rec), rec.ctypes.data)
            i0 += len(rec)
        if i0 != self.params.nparticles:
            raise ValueError(f"{i0} particles for an accelerator of {self.params.nparticles}")
        self.update_forces(advance=False)
        coef = dt ** 2 / (4 * M) if consistent else dt / M
        _lib.call("vh_prime", self.acc.handle, float(coef))

    # --- the step ------------------------------------------------------------------
    def update_forces(self, advance: bool = True, energy: bool = False, counts: bool = False):
        """_update_forces_resident! (src/simulation.jl:472-531).  With `energy`,
        returns the Hartree energy (half the interaction energy on the
        coefficients of the Hartree potential alone, before the mean field is
        added to them); otherwise None.  With `counts`, records nout, outcount
        and the projectile reduction (which waits for the device); otherwise,
        with the projectile on the device, nothing waits."""
        w = self.weight
        nout = self.acc.deposit_smoothed(self.dmf, w, count=counts)
        if counts:
            self.nout = nout
        if self._denoiser is not None:
            self._denoiser.apply(self.dmf, self._denoise_tol)
        self.acc.deposit_cic(self.dmc, w)
        poisson_nested(self.dmf, self.dmc)        # leaves csol_c = S^-1 phi_c
        hartree = None
        if energy:
            # the Hartree energy needs the coefficients of phi alone
            self.dmf.spline_coefficients()
            hartree = self.acc.interaction_energy(self.dmf, self.dmc, w) / 2
            self.dmf.effective_potential(self.jellium)
        else:
            self.dmf.potential_coefficients(self.jellium)
        self.dmc.effective_potential(self.jellium)
        if self._proj_on_device:
            self.acc.forces_resident(self.dmf, self.dmc, w)
            if advance:
                self.acc.projectile_step(self.params.dt)
                self._proj_stale = True
        else:
            host_outputs = counts or self._projectile is not None
            outcount, red = self.acc.compute_forces(self.dmf, self.dmc, w, self._projectile,
                                                    self.softening_mode, outputs=host_outputs)
            if host_outputs:
                self.outcount, self.red = outcount, red
            if advance and self._projectile is not None:
                self.advance_projectile()
        return hartree

    def attach_denoiser(self, den, tol: float = 0.02):
        """Runs the learned denoiser `den` (a denoiser.Denoiser on this
        simulation's context) on the fine density of every later step,
        between the deposit and the Poisson solve, with the residual charge
        check at `tol` (None detaches)."""
        _lib.call("vh_accel_set_denoiser", self.acc.handle, den.handle if den is not None else None,
                  float(tol))
        self._denoiser, self._denoise_tol = den, float(tol)

    def projectile_forces(self):
        """projectile_forces!(cloud, acc, proj, jel) (src/accelerator.jl:746-756)."""
        p = self.projectile
        fjel, e_jel = jellium_force(p, self.jellium)
        force = tuple(a + b for a, b in zip(fjel, self.red[:3]))
        return force, self.weight * p.charge * self.red[3], e_jel

    def advance_projectile(self):
        force, _, _ = self.projectile_forces()
        self.projectile.step(force, self.params.dt)

    def step(self, energy: bool = False):
        """step!(sim; energy) (src/simulation.jl:382-396).  With `energy`,
        returns the EnergyBudget: the Hartree energy measured before the
        push, the kinetic energy of the push, and the total interaction on the
        step's total potential at the pushed positions (the host path's
        order, src/simulation.jl:309-351).  Without, returns None."""
        if not energy and (self._proj_on_device or self._projectile is None):
            # nothing comes back to the host: the whole step in one native call
            self.acc.step(self.dmf, self.dmc, self.weight, self.params.dt, self.mass,
                          self.params.rcmax, self.jellium)
            self._proj_stale = self._proj_on_device
            self.diagnostics = None
            self.step_count += 1
            return None
        hartree = self.update_forces(advance=True, energy=energy)
        self.diagnostics = self.acc.verlet(self.params.dt, self.mass, self.params.rcmax,
                                           diagnostics=energy)
        self.step_count += 1
        if not energy:
            return None
        total = self.acc.interaction_energy(self.dmf, self.dmc, self.weight)
        return energy_budget(self.jellium, self.diagnostics["kinetic"], hartree, total,
                             self.diagnostics["escaped"])

    def run(self, nsteps: int, energy_every: int = 0, callback=None):
        """run! (src/simulation.jl:429-441): the budget history of the steps
        where it is taken (every `energy_every` steps, none when 0)."""
        if energy_every < 0:
            raise ValueError("`energy_every` must be at least 0")
        history = []
        for i in range(nsteps):
            b = self.step(energy=energy_every > 0 and i % energy_every == 0)
            if b is not None:
                history.append(b)
            if callback is not None:
                callback(i + 1, b)
        return history

    # --- observables -------------------------------------------------------------
    def enclosed_charge(self, radius: float = 8.0) -> float:
        """q(< radius) around the projectile (src/projectile.jl:243-247)."""
        if self.projectile is None:
            raise ValueError("no projectile")
        return self.weight * self.acc.enclosed_count(self.projectile.position, radius)

    def net_charge(self, radius: float = 35.0) -> float:
        """The film's q_net: electrons minus those within `radius` of the
        cluster's centre (scripts/xenon.jl:405-406), counted on the device."""
        return self.params.nelectrons - self.weight * self.acc.enclosed_count(np.zeros(3), radius)

    def frame(self) -> np.ndarray:
        """The film frame: the fine density's mean over the planes nfine and
        nfine + 1, the reaction plane z = 0 (scripts/xenon.jl:370, 403),
        reduced on the device (n^2 floats read back instead of n^3)."""
        return self.dmf.density_slice(self.params.nfine, self.params.nfine + 1)

    def positions(self) -> np.ndarray:
        return record_positions(self.acc.download(), self.x0, self.h)

    def sync_host(self):
        """sync_host!: the grids of both levels (src/simulation.jl:544-552)."""
        return {"rho": (self.dmf.rho, self.dmc.rho), "phi": (self.dmf.phi, self.dmc.phi),
                "csol": (self.dmf.csol, self.dmc.csol)}

    def close(self):
        if getattr(self, "_denoiser", None) is not None:
            self.attach_denoiser(None)
        for o in (getattr(self, "acc", None), getattr(self, "dmf", None), getattr(self, "dmc", None)):
            if o is not None:
                o.close()
        if getattr(self, "ctx", None) is not None:
            self.ctx.close()
