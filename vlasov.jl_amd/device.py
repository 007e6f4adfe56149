"""The reference's device layer, backed by libvlasov_hip.so.

`DeviceMesh` and `DeviceAccelerator` mirror src/devicemesh.jl and
src/accelerator.jl: same constructors' inputs, same method names (with `!`
dropped), same argument meaning and the same refusals (ValueError where the
reference throws ArgumentError / DimensionMismatch, before any launch).
Device memory is owned by the library behind opaque handles; numpy arrays
cross the boundary only through explicit upload/download calls, as
`DualBuffer` does (src/accelerator.jl:28-68).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import VH_GRID_CSOL, VH_GRID_PHI, VH_GRID_RHO, VH_SOFT_BALL, VH_SOFT_DEVICE
from .grids import GaussianSmoothing, SplineAxis, SplineMesh
from .physics import PACKED_DTYPE, BallSoftening, Jellium, Projectile

SOFTENING_MODES = {"device": VH_SOFT_DEVICE, "ball": VH_SOFT_BALL}


class Context:
    """One device and one stream."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _lib.call("vh_context_create", int(device), C.byref(h))
        self.handle = h
        self.device = device

    def sync(self):
        _lib.call("vh_sync", self.handle)

    def close(self):
        if self.handle:
            _lib.call("vh_context_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _colmajor(m: np.ndarray) -> np.ndarray:
    """Julia column-major flattening of a math-indexed matrix."""
    return np.ascontiguousarray(np.asarray(m, dtype=np.float64).T).ravel()


class DeviceMesh:
    """Device mirror of a SplineMesh (src/devicemesh.jl:21-81)."""

    def __init__(self, ctx: Context, mesh: SplineMesh):
        ax = mesh.axis
        n, nk = ax.nbasis, ax.nknots
        op = mesh.op
        dual = np.ascontiguousarray(np.concatenate(mesh.dual_moments))
        h = C.c_void_p()
        _lib.call("vh_mesh_create", ctx.handle, n, nk,
                  np.ascontiguousarray(ax.colloc), np.ascontiguousarray(ax.knots),
                  _colmajor(mesh.laplacian), _colmajor(op.M), _colmajor(op.Minv),
                  np.ascontiguousarray(op.lam), _colmajor(mesh.coll.Sinv), dual,
                  np.ascontiguousarray(mesh.dual_len), float(mesh.locator.x0),
                  float(mesh.locator.invwidth), np.ascontiguousarray(mesh.locator.cell, dtype=np.int32),
                  len(mesh.locator.cell), C.byref(h))
        self.handle = h
        self.ctx = ctx
        self.mesh = mesh
        self.n = n

    # --- grids ---------------------------------------------------------------
    def download(self, which: int) -> np.ndarray:
        out = np.empty(self.n ** 3, dtype=np.float32)
        _lib.call("vh_mesh_grid_download", self.handle, int(which), out)
        return out

    def upload(self, which: int, data: np.ndarray):
        a = np.ascontiguousarray(np.asarray(data, dtype=np.float32).ravel())
        if a.size != self.n ** 3:
            raise ValueError(f"grid must hold {self.n}^3 values, got {a.size}")
        _lib.call("vh_mesh_grid_upload", self.handle, int(which), a)

    def set_bf16x3(self, solve: bool = True, transforms: bool | None = None):
        """Rotating products in split-BF16 x3 on the matrix cores (FP32
        accumulation) instead of FP32 (VH_MESH_OPT_BF16X3): the Poisson
        solve's M^-1 and M contractions with `solve`, the S^-1 transforms
        (spline coefficients, mean field) with `transforms` (default: as
        `solve`)."""
        if transforms is None:
            transforms = solve
        _lib.call("vh_mesh_set_option", self.handle, _lib.VH_MESH_OPT_BF16X3,
                  (1 if solve else 0) | (2 if transforms else 0))

    def copy_grid(self, dst: int, src: int):
        """copyto!(dst grid, src grid) on the device (src/simulation.jl:527-530)."""
        _lib.call("vh_mesh_grid_copy", self.handle, int(dst), int(src))

    def density_slice(self, iz1: int, iz2: int) -> np.ndarray:
        """A film frame, Float32.((rho[:, :, iz1] .+ rho[:, :, iz2]) ./ 2) with
        1-based planes (scripts/xenon.jl:396-410), reduced on the device:
        shape (n, n), indexed [j, i] (i fastest in memory)."""
        out = np.empty(self.n * self.n, dtype=np.float32)
        _lib.call("vh_density_slice", self.handle, int(iz1), int(iz2), out)
        return out.reshape(self.n, self.n)

    @property
    def rho(self):
        return self.download(VH_GRID_RHO)

    @property
    def phi(self):
        return self.download(VH_GRID_PHI)

    @property
    def csol(self):
        return self.download(VH_GRID_CSOL)

    # --- the chain -----------------------------------------------------------
    def poisson(self):
        """poisson!(phi, rho, dm), multipole faces (src/devicemesh.jl:148-151)."""
        _lib.call("vh_poisson", self.handle)

    def spline_coefficients(self):
        """csol = S^-1 phi along every axis (src/devicemesh.jl:126-127)."""
        _lib.call("vh_spline_coefficients", self.handle)

    def effective_potential(self, jel: Jellium):
        """csol += S^-1 (V_xc(rho) + V_jel) (src/devicemesh.jl:171-183)."""
        _lib.call("vh_effective_potential", self.handle, float(jel.nions), float(jel.radius))

    def potential_coefficients(self, jel: Jellium):
        """spline_coefficients then effective_potential as one transform,
        csol = S^-1 (phi + V_xc(rho) + V_jel): three products instead of six."""
        _lib.call("vh_potential_coefficients", self.handle, float(jel.nions), float(jel.radius))

    def total_charge(self) -> float:
        q = C.c_double()
        _lib.call("vh_total_charge", self.handle, C.byref(q))
        return float(q.value)

    def multipole(self) -> np.ndarray:
        out = np.zeros(10)
        _lib.call("vh_multipole", self.handle, out)
        return out

    def close(self):
        if self.handle:
            _lib.call("vh_mesh_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def poisson_nested(fine: DeviceMesh, coarse: DeviceMesh):
    """poisson!(phis, rhos, dms) (src/devicemesh.jl:160-169)."""
    _lib.call("vh_poisson_nested", fine.handle, coarse.handle)


class DeviceAccelerator:
    """src/accelerator.jl:126-331, for a cloud resident on the device
    (`packed = true`)."""

    def __init__(self, ctx: Context, fine: SplineAxis, sm: GaussianSmoothing, npart: int):
        knots = np.ascontiguousarray(fine.knots, dtype=np.float64)
        h = (knots[-1] - knots[0]) / (len(knots) - 1)
        if np.max(np.abs(np.diff(knots) - h)) > 1e-9 * abs(h):
            raise ValueError("the accelerated path assumes a uniform fine grid")
        hdl = C.c_void_p()
        _lib.call("vh_accel_create", ctx.handle, int(npart), len(knots), knots, float(sm.spacing),
                  int(sm.nbdt), np.ascontiguousarray(sm.overlap), np.ascontiguousarray(sm.gradient),
                  np.ascontiguousarray(sm.nodes), C.byref(hdl))
        self.handle = hdl
        self.ctx = ctx
        self.npart = int(npart)
        self.x0 = float(knots[0])
        self.h = float(h)
        self.nk = len(knots)
        self.n = 2 * len(knots)

    def set_exact_field(self, exact: bool):
        """True: the field contraction in the reference's exact FMA order
        (bit-identical to contract_spline_10 in FP32); False (default): its
        first stage on the matrix cores."""
        _lib.call("vh_accel_set_option", self.handle, _lib.VH_OPT_EXACT_FIELD, 1 if exact else 0)

    def set_field_fp32(self, fp32: bool):
        """True: the matrix-core contraction's first stage in FP32 products;
        False (default): in fp16 splits (hi*hi + hi*lo + lo*hi, FP32
        accumulation, about 2^-21 of the largest term)."""
        _lib.call("vh_accel_set_option", self.handle, _lib.VH_OPT_FIELD_FP32, 1 if fp32 else 0)

    def set_poisson_unfused(self, unfused: bool):
        """True: vh_step solves and transforms as the entry points do
        (bit-identical to calling them); False (default): each level's
        coefficients straight from the density and the boundary."""
        _lib.call("vh_accel_set_option", self.handle, _lib.VH_OPT_POISSON_UNFUSED, 1 if unfused else 0)

    def set_step_graph(self, on: bool = True):
        """vh_step recorded once into a HIP graph per particle buffer and
        replayed (VH_OPT_STEP_GRAPH)."""
        _lib.call("vh_accel_set_option", self.handle, _lib.VH_OPT_STEP_GRAPH, 1 if on else 0)

    # --- buffers -------------------------------------------------------------
    def upload(self, records: np.ndarray):
        r = np.ascontiguousarray(records)
        if r.dtype != PACKED_DTYPE or r.shape != (self.npart,):
            raise ValueError(f"expected {self.npart} PackedParticle{{Float32}} records")
        _lib.call("vh_particles_upload", self.handle, r.ctypes.data)

    def download(self) -> np.ndarray:
        r = np.zeros(self.npart, dtype=PACKED_DTYPE)
        _lib.call("vh_particles_download", self.handle, r.ctypes.data)
        return r

    def forces(self) -> np.ndarray:
        """(N, 3) forces in slot order."""
        f = np.zeros(3 * self.npart, dtype=np.float32)
        _lib.call("vh_forces_download", self.handle, f)
        return f.reshape(self.npart, 3)

    def upload_forces(self, f: np.ndarray):
        a = np.ascontiguousarray(np.asarray(f, dtype=np.float32).reshape(-1))
        if a.size != 3 * self.npart:
            raise ValueError("force buffer must be (3, npart)")
        _lib.call("vh_forces_upload", self.handle, a)

    # --- methods -------------------------------------------------------------
    def sort(self):
        """_sort_and_list!: returns (n_occupied, outcount)."""
        nocc, nout = C.c_int32(), C.c_int32()
        _lib.call("vh_sort", self.handle, C.byref(nocc), C.byref(nout))
        return int(nocc.value), int(nout.value)

    def sort_lists(self, nocc: int, outcount: int):
        """(occupied keys 1-based, bounds, outlist 1-based slots) of the last sort."""
        occ = np.zeros(max(nocc, 1), dtype=np.int32)
        bnd = np.zeros(nocc + 1, dtype=np.int32)
        out = np.zeros(max(outcount, 1), dtype=np.int32)
        _lib.call("vh_sort_lists", self.handle, occ, bnd, out)
        return occ[:nocc], bnd, out[:outcount]

    def columns(self):
        """_fill_columns!: ((N, 3) 1-based columns, nout)."""
        cols = np.zeros(3 * self.npart, dtype=np.int32)
        nout = C.c_int32()
        _lib.call("vh_columns", self.handle, cols, C.byref(nout))
        return cols.reshape(self.npart, 3), int(nout.value)

    def deposit_smoothed(self, mesh: DeviceMesh, charge: float, count: bool = True):
        """deposit_smoothed!(rho, acc, mesh, sm, positions; charge) -> nout
        (None without `count`: then nothing waits for the device)."""
        if mesh.n != self.n:
            raise ValueError("rho must cover the whole collocation grid")
        if not count:
            _lib.call("vh_deposit_smoothed", self.handle, mesh.handle, float(charge), None)
            return None
        nout = C.c_int32()
        _lib.call("vh_deposit_smoothed", self.handle, mesh.handle, float(charge), C.byref(nout))
        return int(nout.value)

    def deposit_cic(self, mesh: DeviceMesh, charge: float):
        """deposit_cic!(rho, dm, acc, npart, charge, x0f, hf)."""
        _lib.call("vh_deposit_cic", self.handle, mesh.handle, float(charge))

    def compute_forces(self, fine: DeviceMesh, coarse: DeviceMesh, weight: float,
                       projectile: Projectile | None = None, mode: str = "device",
                       outputs: bool = True):
        """forces!(cloud, acc, ...; projectile, packed=true): returns
        (outcount, red[4]) with red the projectile reduction; (None, None)
        without `outputs` (then nothing waits for the device)."""
        if mode not in SOFTENING_MODES:
            raise ValueError(f"unknown softening mode {mode!r}")
        if fine.n != self.n:
            raise ValueError(f"accelerator sized for n={self.n}")
        red = np.zeros(4)
        nout = C.c_int32()
        if projectile is None:
            pp, radius = None, 0.0
        else:
            pp = np.array([*projectile.position, projectile.charge, projectile.scale], dtype=np.float64)
            radius = projectile.softening.radius if isinstance(projectile.softening, BallSoftening) else 0.0
        _lib.call("vh_forces", self.handle, fine.handle, coarse.handle, float(weight),
                  None if pp is None else pp.ctypes.data, SOFTENING_MODES[mode], float(radius),
                  red.ctypes.data if outputs else None, C.byref(nout) if outputs else None)
        return (int(nout.value), red) if outputs else (None, None)

    def verlet(self, dt: float, mass: float, rcmax: float = float("inf"), diagnostics: bool = True):
        """step!(cloud, dt, acc; rcmax) -> (kinetic, escaped, angular); None
        without `diagnostics` (then nothing waits for the device)."""
        if not diagnostics:
            _lib.call("vh_verlet", self.handle, float(dt), float(mass), float(rcmax), None)
            return None
        d = np.zeros(5)
        _lib.call("vh_verlet", self.handle, float(dt), float(mass), float(rcmax), d.ctypes.data)
        return {"kinetic": d[0], "escaped": d[1], "angular": (d[2], d[3], d[4])}

    def step(self, fine: DeviceMesh, coarse: DeviceMesh, weight: float, dt: float, mass: float,
             rcmax: float, jel: Jellium):
        """One resident step in one native call (vh_step): deposits, nested
        Poisson, coefficients, mean field, forces with the device projectile
        (or none), its step, push.  Nothing waits for the device."""
        _lib.call("vh_step", self.handle, fine.handle, coarse.handle, float(weight), float(dt),
                  float(mass), float(rcmax), float(jel.nions), float(jel.radius))

    # --- the projectile kept on the device -----------------------------------
    def set_projectile(self, proj: Projectile, mode: str, jel: Jellium):
        """Places the projectile's state on the device; from then on
        forces_resident and projectile_step integrate it there."""
        if mode not in SOFTENING_MODES:
            raise ValueError(f"unknown softening mode {mode!r}")
        st = np.array([*proj.position, *proj.previous, *proj.velocity], dtype=np.float64)
        radius = proj.softening.radius if isinstance(proj.softening, BallSoftening) else 0.0
        _lib.call("vh_projectile_set", self.handle, st, float(proj.mass), float(proj.charge),
                  float(proj.scale), SOFTENING_MODES[mode], float(radius), float(jel.nions),
                  float(jel.radius))

    def projectile_state(self) -> np.ndarray:
        """(position, previous, velocity) of the device projectile, 9 values."""
        st = np.zeros(9)
        _lib.call("vh_projectile_get", self.handle, st)
        return st

    def forces_resident(self, fine: DeviceMesh, coarse: DeviceMesh, weight: float):
        """forces! with the device projectile; nothing waits for the device."""
        if fine.n != self.n:
            raise ValueError(f"accelerator sized for n={self.n}")
        _lib.call("vh_forces_resident", self.handle, fine.handle, coarse.handle, float(weight))

    def projectile_step(self, dt: float):
        """projectile_forces! and step!(proj) on the device."""
        _lib.call("vh_projectile_step", self.handle, float(dt))

    def interaction_energy(self, fine: DeviceMesh, coarse: DeviceMesh, weight: float,
                           enclosed: float = 0.0) -> float:
        """interaction_energy(cloud, fine, csol_f, coarse, csol_c, sm; enclosed)
        (src/energy.jl:19-43) on the csol grids the two meshes hold now."""
        if fine.n != self.n:
            raise ValueError(f"accelerator sized for n={self.n}")
        e = C.c_double()
        _lib.call("vh_interaction_energy", self.handle, fine.handle, coarse.handle, float(weight),
                  float(enclosed), C.byref(e))
        return float(e.value)

    def enclosed_count(self, center, radius: float) -> int:
        c = np.ascontiguousarray(np.asarray(center, dtype=np.float64))
        cnt = C.c_int64()
        _lib.call("vh_enclosed_count", self.handle, c, float(radius), C.byref(cnt))
        return int(cnt.value)

    def close(self):
        if self.handle:
            _lib.call("vh_accel_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ForceAccelerator:
    """The legacy accelerator route, `ForceAccelerator(ArrayT, fine, sm, npart,
    n)` (ext/VlasovMetalExt.jl:51-54) as `update_forces!`'s accelerator branch
    uses it (src/simulation.jl:318-345; scripts/film_xenon_80M.jl:83-120):
    the cloud stays on the host as FP64 positions, and so do the CIC deposit,
    the FP64 Poisson chain and the Verlet, which are the caller's; the
    smoothed deposit and the field run on the device.  Methods mirror
    `deposit_smoothed!(ρ::Array, acc, mesh, sm, positions; charge)`,
    `forces!(cloud, acc, fine, csol_f, coarse, csol_c, sm; escaped,
    projectile, packed)` and `projectile_forces!(cloud, acc, proj, jel)`
    (src/accelerator.jl:531, 607, 746)."""

    def __init__(self, ctx: Context, fine: SplineMesh, sm: GaussianSmoothing, npart: int):
        self.mesh = DeviceMesh(ctx, fine)
        self.acc = DeviceAccelerator(ctx, fine.axis, sm, npart)
        self.npart = int(npart)
        self.n = self.acc.n
        self.reduction = np.zeros(4)

    def _positions(self, positions) -> np.ndarray:
        pos = np.ascontiguousarray(np.asarray(positions, dtype=np.float64))
        if pos.shape != (self.npart, 3):
            raise ValueError(f"accelerator sized for {self.npart} particles")
        return pos

    def deposit_smoothed(self, rho: np.ndarray, positions, charge: float) -> int:
        """deposit_smoothed!(ρ, acc, mesh, sm, positions; charge) -> nout:
        packs and sorts the host cloud on the device, deposits, renormalises,
        and writes the density into the host array `rho` (n^3, FP64)."""
        if rho.size != self.n ** 3:
            raise ValueError("ρ must cover the whole collocation grid")
        _lib.call("vh_positions_upload", self.acc.handle, self._positions(positions))
        nout = self.acc.deposit_smoothed(self.mesh, charge)
        rho.reshape(-1)[:] = self.mesh.rho
        return nout

    def forces(self, forces: np.ndarray, positions, csol_fine, coarse_knots, csol_coarse,
               weight: float, projectile: Projectile | None = None, escaped: float = 0.0,
               packed: bool = True) -> int:
        """forces!(cloud, acc, fine, csol_f, coarse, csol_c, sm; escaped,
        projectile, packed) -> outcount: the forces of the cloud, in its own
        order and FP64, written into `forces` (N x 3).  `packed` as the
        driver passes it: the deposit of the same step has already packed the
        cloud (otherwise `positions` are packed again first)."""
        if not packed:
            _lib.call("vh_positions_upload", self.acc.handle, self._positions(positions))
        f = np.zeros((self.npart, 3))
        kc = np.ascontiguousarray(coarse_knots, dtype=np.float64)
        cc = np.ascontiguousarray(np.asarray(csol_coarse, dtype=np.float64).ravel())
        cf = np.ascontiguousarray(np.asarray(csol_fine, dtype=np.float64).ravel())
        if cf.size != self.n ** 3:
            raise ValueError(f"fine coefficients must hold {self.n}^3 values")
        nc = 2 * len(kc)
        if cc.size != nc ** 3:
            raise ValueError(f"coarse coefficients must hold {nc}^3 values")
        if projectile is None:
            pp, radius = None, 0.0
        else:
            pp = np.array([*projectile.position, projectile.charge, projectile.scale])
            radius = projectile.softening.radius if isinstance(projectile.softening, BallSoftening) else 0.0
        out = C.c_int32()
        _lib.call("vh_forces_host", self.acc.handle, self.mesh.handle, cf, kc, len(kc), cc, nc,
                  float(weight), None if pp is None else pp.ctypes.data, float(radius),
                  float(escaped), f, self.reduction, C.byref(out))
        forces[...] = f.reshape(forces.shape)
        return int(out.value)

    def projectile_forces(self, weight: float, proj: Projectile, jel: Jellium):
        """projectile_forces!(cloud, acc, proj, jel) on the reduction of the
        last forces (src/accelerator.jl:746-756)."""
        from .physics import jellium_force
        fjel, e_jel = jellium_force(proj, jel)
        force = tuple(a + b for a, b in zip(fjel, self.reduction[:3]))
        return force, weight * proj.charge * self.reduction[3], e_jel

    def close(self):
        self.acc.close()
        self.mesh.close()
