"""ctypes binding of libvlasov_hip.so (the C ABI declared in include/vlasov_hip.h).

There is no fallback: if the library is missing or a call fails, an exception
says so.  Build the library with `python vlasov.jl_amd/build.py`.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libvlasov_hip.so"

VH_GRID_RHO, VH_GRID_PHI, VH_GRID_CSOL = 0, 1, 2
VH_PROF_ALL = -1
VH_SOFT_DEVICE, VH_SOFT_BALL = 0, 1
VH_OPT_EXACT_FIELD = 0
VH_OPT_STEP_GRAPH = 1
VH_OPT_UNIT_SLOTS = 2
VH_OPT_FIELD_FP32 = 3
VH_OPT_POISSON_UNFUSED = 4
VH_MESH_OPT_BF16X3 = 0
STAGES = ("sort", "deposit_fine", "deposit_cic", "poisson", "spline", "meanfield", "field",
          "verlet", "field_kernel", "deposit_kernel")

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_vp = C.c_void_p


class VlasovHIPError(RuntimeError):
    """A call into libvlasov_hip.so returned an error."""


# name -> argtypes; every function returns int status
_SIGNATURES = {
    "vh_device_count": [C.POINTER(C.c_int)],
    "vh_context_create": [C.c_int, C.POINTER(_vp)],
    "vh_context_destroy": [_vp],
    "vh_sync": [_vp],
    "vh_mesh_create": [_vp, C.c_int, C.c_int, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p,
                       _f64p, _f64p, C.c_double, C.c_double, _i32p, C.c_int, C.POINTER(_vp)],
    "vh_mesh_destroy": [_vp],
    "vh_mesh_set_option": [_vp, C.c_int, C.c_int],
    "vh_mesh_grid_download": [_vp, C.c_int, _f32p],
    "vh_mesh_grid_upload": [_vp, C.c_int, _f32p],
    "vh_mesh_grid_copy": [_vp, C.c_int, C.c_int],
    "vh_density_slice": [_vp, C.c_int, C.c_int, _f32p],
    "vh_poisson": [_vp],
    "vh_poisson_nested": [_vp, _vp],
    "vh_spline_coefficients": [_vp],
    "vh_effective_potential": [_vp, C.c_double, C.c_double],
    "vh_total_charge": [_vp, C.POINTER(C.c_double)],
    "vh_multipole": [_vp, _f64p],
    "vh_multipole_noalloc": [C.c_int, _vp, _vp, _vp, C.c_int, _vp],
    "vh_accel_create": [_vp, C.c_int64, C.c_int, _f64p, C.c_double, C.c_int, _f64p, _f64p, _f64p,
                        C.POINTER(_vp)],
    "vh_accel_destroy": [_vp],
    "vh_accel_set_option": [_vp, C.c_int, C.c_int],
    "vh_particles_upload": [_vp, _vp],
    "vh_particles_download": [_vp, _vp],
    "vh_forces_download": [_vp, _f32p],
    "vh_particles_upload_range": [_vp, C.c_int64, C.c_int64, _vp],
    "vh_prime": [_vp, C.c_double],
    "vh_sampler_create": [_f64p, C.c_int, _f64p, C.c_int, C.c_double, C.c_double, C.c_int,
                          C.POINTER(_vp)],
    "vh_sampler_next": [_vp, C.c_int64, _f64p, _f64p],
    "vh_sampler_destroy": [_vp],
    "vh_forces_upload": [_vp, _f32p],
    "vh_sort": [_vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32)],
    "vh_sort_lists": [_vp, _i32p, _i32p, _i32p],
    "vh_columns": [_vp, _i32p, C.POINTER(C.c_int32)],
    "vh_deposit_smoothed": [_vp, _vp, C.c_double, C.POINTER(C.c_int32)],
    "vh_deposit_cic": [_vp, _vp, C.c_double],
    "vh_forces": [_vp, _vp, _vp, C.c_double, _vp, C.c_int, C.c_double, _vp, C.POINTER(C.c_int32)],
    "vh_verlet": [_vp, C.c_double, C.c_double, C.c_double, _vp],
    "vh_positions_upload": [_vp, _f64p],
    "vh_forces_host": [_vp, _vp, _f64p, _f64p, C.c_int, _f64p, C.c_int, C.c_double, _vp, C.c_double,
                       C.c_double, _f64p, _f64p, C.POINTER(C.c_int32)],
    "vh_enclosed_count": [_vp, _f64p, C.c_double, C.POINTER(C.c_int64)],
    "vh_projectile_set": [_vp, _f64p, C.c_double, C.c_double, C.c_double, C.c_int, C.c_double,
                          C.c_double, C.c_double],
    "vh_projectile_get": [_vp, _f64p],
    "vh_forces_resident": [_vp, _vp, _vp, C.c_double],
    "vh_projectile_step": [_vp, C.c_double],
    "vh_potential_coefficients": [_vp, C.c_double, C.c_double],
    "vh_step": [_vp, _vp, _vp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                C.c_double],
    "vh_autotune_unit": [_vp, _vp, _vp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                         C.c_double, _i32p, C.c_int, C.c_int, C.POINTER(C.c_int), _f64p],
    "vh_batch_create": [C.c_int, _vp, _vp, _vp, _f64p, C.c_int, C.POINTER(_vp)],
    "vh_batch_step": [_vp],
    "vh_batch_sync": [_vp],
    "vh_batch_destroy": [_vp],
    "vh_denoiser_param_count": [C.POINTER(C.c_int64)],
    "vh_denoiser_create": [_vp, _vp, C.c_int64, C.c_uint64, C.c_double, C.POINTER(_vp)],
    "vh_denoiser_destroy": [_vp],
    "vh_denoiser_get_weights": [_vp, _vp, _vp],
    "vh_denoiser_set_weights": [_vp, _f32p, C.c_int64],
    "vh_denoiser_eval": [_vp, C.c_int, _f32p, _f32p, _vp, C.POINTER(C.c_double), _vp],
    "vh_denoiser_dataset_reserve": [_vp, C.c_int, C.c_int64],
    "vh_denoiser_dataset_add": [_vp, _vp, _vp],
    "vh_denoiser_dataset_size": [_vp, C.POINTER(C.c_int64)],
    "vh_denoiser_train": [_vp, C.c_int, C.c_int, C.c_int, C.c_int64, C.c_int64, C.c_int, C.c_double,
                          C.c_uint64, _vp],
    "vh_denoiser_dataset_error": [_vp, C.c_int64, C.c_int, C.POINTER(C.c_double),
                                  C.POINTER(C.c_double)],
    "vh_denoiser_apply": [_vp, _vp, C.c_double],
    "vh_denoiser_fallbacks": [_vp, C.POINTER(C.c_int32)],
    "vh_denoiser_set_inference": [_vp, C.c_int],
    "vh_denoiser_set_support": [_vp, C.c_double],
    "vh_accel_set_denoiser": [_vp, _vp, C.c_double],
    "vh_interaction_energy": [_vp, _vp, _vp, C.c_double, C.c_double, C.POINTER(C.c_double)],
    "vh_profile_enable": [_vp, C.c_int],
    "vh_profile_read": [_vp, _f64p, np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")],
    "vh_ran2_sequence": [C.c_int, C.c_int, C.c_int64, _f32p],
    "vh_sample_radial": [_f64p, C.c_int, _f64p, C.c_int, C.c_double, C.c_int64, C.c_double, C.c_int,
                         _f64p, _f64p],
    "vh_sample_rejection": [_f64p, _f64p, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int64,
                            C.c_double, C.c_int, _f64p, _f64p],
}

EXPORTED = sorted(list(_SIGNATURES) + ["vh_last_error", "vh_version"])

_lib = None


def load(path: Path | str | None = None):
    """Loads the library once; raises if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise VlasovHIPError(f"{p} is missing: build it with `python vlasov.jl_amd/build.py` "
                             "(there is no CPU fallback)")
    lib = C.CDLL(str(p))
    for name, args in _SIGNATURES.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = C.c_int
    lib.vh_last_error.argtypes = []
    lib.vh_last_error.restype = C.c_char_p
    lib.vh_version.argtypes = []
    lib.vh_version.restype = C.c_char_p
    _lib = lib
    return lib


def call(name: str, *args):
    """Calls `name`, raising VlasovHIPError with the library's message on failure."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.vh_last_error().decode(errors="replace")
        raise VlasovHIPError(f"{name} failed ({rc}): {msg}")
    return rc


def last_error() -> str:
    return load().vh_last_error().decode(errors="replace")


def device_count() -> int:
    n = C.c_int(0)
    call("vh_device_count", C.byref(n))
    return int(n.value)
