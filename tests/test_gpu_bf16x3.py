"""The precision extension's rotating products: split-BF16 x3 on the matrix
cores (hi = bf16(x), lo = bf16(x - hi); A S ~ A_lo S_hi + A_hi S_lo +
A_hi S_hi, FP32 accumulation), against the FP64 oracle, at n = 58, 130, 258.

Bars, stated as the error each path carries (one split-BF16 x3 product
keeps ~17 bits: |A - (hi + lo)| <= 2^-17 |A| and the dropped lo lo term is
~2^-18), relative (max |diff| / max |ref|):

| n | path | S^-1 (random grid) | Poisson solve | mean field S^-1 V_eff added to csol |
|---|---|---|---|---|
| 58 | FP32 | 2.3e-7 | 5.6e-7 | 4.3e-7 |
| 58 | split-BF16 x3 | 1.1e-5 | 1.7e-5 | 3.3e-5 |
| 130 | FP32 | 2.3e-7 | 8.3e-7 | 2.1e-6 |
| 130 | split-BF16 x3 | 5.9e-6 | 2.6e-5 | 1.1e-4 |
| 258 | FP32 | 2.7e-7 | 7.2e-7 | 2.3e-5 |
| 258 | split-BF16 x3 | 7.6e-6 | 2.7e-5 | 8.9e-4 |

(the measured values; the mean field is the S^-1 transform of the large,
smooth jellium and exchange terms, whose rounding S^-1 amplifies more as the
grid refines, in FP32 as well).  Bars: FP32 S^-1 and solve 1e-5, mean field
5e-5; split-BF16 x3 S^-1 3e-5, solve 1e-4, mean field 2e-3; and the split
path is measurably another one (errors above 1e-6): a plain BF16 product
would be ~4e-3."""
import numpy as np
import pytest

from conftest import load_package

pytestmark = pytest.mark.gpu
va = load_package()
L = va._lib


def relerr(got, ref):
    return float(np.max(np.abs(np.asarray(got, np.float64) - ref)) / np.max(np.abs(ref)))


@pytest.mark.parametrize("nfine", [28, 64, 128])
def test_bf16x3_products_against_fp64(orc, nfine):
    mesh = va.grids.SplineMesh.of(va.grids.uniform_axis(-78.0, 78.0, nfine))
    n = mesh.n
    rng = np.random.default_rng(100 + nfine)
    ctx = va.Context(0)
    dm = va.DeviceMesh(ctx, mesh)
    try:
        g = rng.standard_normal(n ** 3).astype(np.float32)
        want_s = orc.spline_coefficients(g.astype(np.float64), mesh)
        gv = mesh.axis.colloc
        kk, jj, ii = np.meshgrid(gv, gv, gv, indexing="ij")
        rho = (np.exp(-(ii ** 2 + jj ** 2 + kk ** 2) / (2 * 15.0 ** 2))
               * (1 + 0.05 * rng.random(ii.shape))).astype(np.float32).ravel()
        want_phi, _ = orc.poisson_single(rho.astype(np.float64), mesh)
        jel = va.Jellium.of(196.0)
        base = rng.standard_normal(n ** 3).astype(np.float32)
        extra = orc.spline_coefficients(orc.effective_extra(rho.astype(np.float64), mesh, jel), mesh)
        errs = {}
        for b3 in (False, True):
            dm.set_bf16x3(b3)
            dm.upload(L.VH_GRID_PHI, g)
            dm.spline_coefficients()
            e_s = relerr(dm.csol, want_s)
            dm.upload(L.VH_GRID_RHO, rho)
            dm.poisson()
            e_p = relerr(dm.phi, want_phi)
            dm.upload(L.VH_GRID_CSOL, base)
            dm.effective_potential(jel)
            scale = max(np.max(np.abs(base)), np.max(np.abs(extra)))
            e_m = float(np.max(np.abs(dm.csol - (base + extra))) / scale)
            errs[b3] = (e_s, e_p, e_m)
        print(f"n={n}: FP32 {errs[False]}, split-BF16x3 {errs[True]}")
        e_s, e_p, e_m = errs[False]
        assert e_s < 1e-5 and e_p < 1e-5 and e_m < 5e-5
        e_s, e_p, e_m = errs[True]
        assert e_s < 3e-5 and e_p < 1e-4 and e_m < 2e-3
        assert e_s > 1e-6 and e_p > 1e-6        # the split path did run
    finally:
        dm.close()
        ctx.close()


def test_bf16x3_nested_poisson_config2(orc):
    """The nested solve at config 2's shape (130^3, both levels in split-BF16
    x3) on a cluster-like density pair, against the FP64 chain: phi of both
    levels and the coarse coefficients within 1e-4."""
    p = va.xenon_parameters(8_000_000, 64)
    fine, coarse = va.simulation.build_meshes(p)
    rng = np.random.default_rng(7)

    def cluster(mesh):
        gv = mesh.axis.colloc
        kk, jj, ii = np.meshgrid(gv, gv, gv, indexing="ij")
        r = np.exp(-((ii - 3) ** 2 + jj ** 2 + kk ** 2) / (2 * 14.0 ** 2))
        r = r * (1 + 0.02 * rng.random(ii.shape))
        return (r * (196.0 / orc.total_charge(r.ravel(), mesh.dual_moments[0]))).astype(np.float32).ravel()

    rf, rc = cluster(fine), cluster(coarse)
    wf, wc, _ = orc.poisson_nested(rf.astype(np.float64), rc.astype(np.float64), fine, coarse)
    ctx = va.Context(0)
    dmf, dmc = va.DeviceMesh(ctx, fine), va.DeviceMesh(ctx, coarse)
    try:
        for d in (dmf, dmc):
            d.set_bf16x3(True)
        dmf.upload(L.VH_GRID_RHO, rf)
        dmc.upload(L.VH_GRID_RHO, rc)
        va.poisson_nested(dmf, dmc)
        errs = [relerr(dmf.phi, wf), relerr(dmc.phi, wc),
                relerr(dmc.csol, orc.spline_coefficients(wc, coarse))]
        print(f"nested split-BF16x3 at 130^3: phi_f {errs[0]:.2e}, phi_c {errs[1]:.2e}, "
              f"csol_c {errs[2]:.2e}")
        assert max(errs) < 1e-4
    finally:
        dmf.close(); dmc.close(); ctx.close()
