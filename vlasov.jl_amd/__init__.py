"""MI355X-native implementation of Vlasov.jl's per-step pseudo-particle loop.

The hot path (cell sort, smoothed and cloud-in-cell deposits, two-level
spline Poisson solve with the mean field, smoothed field gather with the
fused projectile, packed Verlet) runs in libvlasov_hip.so, hand-written HIP
for gfx950 behind the C ABI of include/vlasov_hip.h.  This package is the
host side: FP64 construction of the tables, the device-layer objects of the
reference (DeviceMesh, DeviceAccelerator) and the simulation driver.
"""
from . import _lib, batch, denoiser, grids, physics
from ._lib import VlasovHIPError, device_count
from .device import Context, DeviceAccelerator, DeviceMesh, ForceAccelerator, poisson_nested
from .grids import GaussianSmoothing, SplineAxis, SplineMesh, stretched_axis, uniform_axis
from .physics import (BallSoftening, GaussianSoftening, Jellium, PACKED_DTYPE, Projectile,
                      RadialProfile, read_radial_profile, sample_thomas_fermi)
from .simulation import Simulation, SimulationParameters, xenon_parameters, xenon_projectile

__all__ = [
    "VlasovHIPError", "device_count", "Context", "DeviceAccelerator", "DeviceMesh", "ForceAccelerator",
    "poisson_nested", "GaussianSmoothing", "SplineAxis", "SplineMesh", "stretched_axis",
    "uniform_axis", "BallSoftening", "GaussianSoftening", "Jellium", "PACKED_DTYPE",
    "Projectile", "RadialProfile", "read_radial_profile", "sample_thomas_fermi", "Simulation",
    "SimulationParameters", "xenon_parameters", "xenon_projectile", "grids", "physics", "batch",
    "denoiser",
]
