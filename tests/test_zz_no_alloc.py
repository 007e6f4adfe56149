"""vh_multipole_noalloc, the multipole without allocation, against
vh_multipole through the C ABI: the same kernels in the same order, so the
same bits, on two calls in a row, the second on a smaller grid.

Device memory for the caller's side comes from the HIP runtime the library
itself links (the same runtime instance in this process), through ctypes."""
import ctypes as C

import numpy as np
import pytest

from conftest import load_package

pytestmark = pytest.mark.gpu
va = load_package()


class HipMemory:
    """hipMalloc / hipMemcpy / hipFree of the runtime libvlasov_hip.so uses."""

    def __init__(self):
        va._lib.load()
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.hip.hipDeviceSynchronize.argtypes = []
        self.hip.hipSetDevice.argtypes = [C.c_int]
        self.ptrs = []

    def check(self, rc, what):
        assert rc == 0, f"{what} failed ({rc})"

    def upload(self, a: np.ndarray) -> int:
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        self.check(self.hip.hipMalloc(C.byref(p), a.nbytes), "hipMalloc")
        self.ptrs.append(p.value)
        self.check(self.hip.hipMemcpy(p.value, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")
        return p.value

    def download(self, ptr: int, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        self.check(self.hip.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2), "hipMemcpy D2H")
        return out

    def close(self):
        self.check(self.hip.hipDeviceSynchronize(), "hipDeviceSynchronize")
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def cluster_density(mesh, seed):
    g = mesh.axis.colloc
    kk, jj, ii = np.meshgrid(g, g, g, indexing="ij")
    rng = np.random.default_rng(seed)
    r2 = (ii - 3.0) ** 2 + (jj + 1.5) ** 2 + (kk - 0.5) ** 2
    rho = np.exp(-r2 / (2 * 15.0 ** 2)) * (1 + 0.1 * rng.random(r2.shape))
    return rho.astype(np.float32).ravel()


def test_multipole_noalloc_matches_vh_multipole():
    ctx = va.Context(0)
    mem = HipMemory()
    try:
        for nfine, seed in ((64, 1), (28, 2)):      # the second call on a smaller grid
            mesh = va.grids.SplineMesh.of(va.grids.uniform_axis(-78.0, 78.0, nfine))
            n = mesh.n
            rho = cluster_density(mesh, seed)
            dual = np.ascontiguousarray(np.concatenate(mesh.dual_moments))
            dm = va.DeviceMesh(ctx, mesh)
            try:
                dm.upload(va._lib.VH_GRID_RHO, rho)
                want = dm.multipole()
            finally:
                dm.close()
            d_rho, d_dual = mem.upload(rho), mem.upload(dual)
            d_out = mem.upload(np.full(10, np.nan))
            va._lib.call("vh_multipole_noalloc", 0, None, d_rho, d_dual, n, d_out)
            got = mem.download(d_out, np.zeros(10))
            assert got.tobytes() == want.tobytes(), (nfine, got, want)
            assert abs(got[0] - want[0]) == 0 and np.isfinite(got).all()
        # sizes beyond the scratch are refused before any launch
        with pytest.raises(va.VlasovHIPError, match="outside"):
            va._lib.call("vh_multipole_noalloc", 0, None, d_rho, d_dual, 300, d_out)
    finally:
        mem.close()
        ctx.close()
