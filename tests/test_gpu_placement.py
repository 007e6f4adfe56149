"""The placement known answers of tests/test_placement.py, on the HIP path
(config-1 grids: nfine = 28, 58^3).

The clouds put every particle but the probes beyond both grids (x = 400 a0):
those deposit nothing, so the densities hold the probes' contributions alone,
at the accelerator's ordinary sizes and launch shapes."""
import numpy as np
import pytest

from conftest import load_package
from test_placement import (QUAD_KNOTS, cic_cases, mirror_map, mirror_pair, quadratic_cloud,
                            quadratic_potentials, record_at, tabulated_x, field_ratio_error,
                            check_mirror)

pytestmark = pytest.mark.gpu
va = load_package()
NPART = 1024


def far_cloud(x0, h, n=NPART):
    pos = np.zeros((n, 3))
    pos[:, 0] = 400.0
    pos[:, 1] = np.linspace(-100.0, 100.0, n)
    return va.physics.packed_records(pos, pos, x0, h)


@pytest.fixture(scope="module")
def dev(config1):
    p, fine, coarse, sm = config1
    ctx = va.Context(0)
    dmf = va.DeviceMesh(ctx, fine)
    dmc = va.DeviceMesh(ctx, coarse)
    acc = va.DeviceAccelerator(ctx, fine.axis, sm, NPART)
    kn = fine.axis.knots
    yield dict(fine=fine, coarse=coarse, sm=sm, ctx=ctx, dmf=dmf, dmc=dmc, acc=acc,
               x0=float(kn[0]), h=float(kn[1] - kn[0]))
    acc.close(); dmf.close(); dmc.close(); ctx.close()


def test_gpu_cic_single_particle_placement(dev):
    """A particle on coarse collocation point 5 (each axis) puts its whole
    charge there; one half-way between points 7 and 8 splits it 0.5/0.5
    (locate's conventions, test/runtests.jl:299-306)."""
    coarse, acc, dmc = dev["coarse"], dev["acc"], dev["dmc"]
    rec = far_cloud(dev["x0"], dev["h"])
    want = {}
    for i, (target, w) in enumerate(cic_cases(coarse)):
        rec[i] = record_at(target, dev["x0"], dev["h"])[0]
        want.update(w)
    acc.upload(rec)
    acc.sort()
    acc.deposit_cic(dmc, 1.0)
    n = coarse.n
    wl = coarse.dual_len
    vol = (wl[None, None, :] * wl[None, :, None]) * wl[:, None, None]
    raw = (dmc.rho.astype(np.float64).reshape(n, n, n)) * vol
    nz = {(int(i) + 1, int(j) + 1, int(k) + 1): raw[k, j, i] for k, j, i in zip(*np.nonzero(raw))}
    assert set(nz) == set(want), nz
    for key, wv in want.items():
        assert abs(nz[key] - wv) < 2e-5, (key, nz[key])


def test_gpu_smoothed_deposit_mirror_stencils(dev):
    """One particle at knot + d, then one at knot - d: mirror-image 8^3
    stencils about the knot (src/fields.jl:278-287)."""
    fine, sm, acc, dmf = dev["fine"], dev["sm"], dev["acc"], dev["dmf"]
    ci, (ra, rb) = mirror_pair(fine, sm)
    rhos = []
    for r in (ra, rb):
        rec = far_cloud(dev["x0"], dev["h"])
        rec[0] = r[0]
        acc.upload(rec)
        nout = acc.deposit_smoothed(dmf, 1.0)
        assert nout == NPART - 1
        rhos.append(dmf.rho.astype(np.float64))
    mirror = mirror_map(fine.axis.colloc, fine.axis.knots[ci - 1])
    # FP32 stencil values: mirror images to Float32 rounding
    a, b = rhos
    n = fine.n
    A, B = a.reshape(n, n, n), b.reshape(n, n, n)
    sup = np.nonzero(np.any(A != 0, axis=(0, 1)))[0]
    assert len(sup) == 8 and np.all(mirror[sup] >= 0)
    err = np.max(np.abs(A[:, :, sup] - B[:, :, mirror[sup]])) / np.max(np.abs(A))
    assert err < 1e-6, err
    assert np.count_nonzero(A) == np.count_nonzero(B) == 8 ** 3


@pytest.mark.parametrize("exact", [False, True])
def test_gpu_field_of_a_quadratic_potential(dev, exact):
    """Phi = (x - X)^2/2 about knot X, divided by the field of Phi = x - X:
    the tabulated offset x_tab - X to 0.1 spacing/nbdt (a one-column slip is
    a whole spacing/nbdt), with either contraction."""
    fine, coarse, sm, acc, dmf, dmc = (dev[k] for k in ("fine", "coarse", "sm", "acc", "dmf", "dmc"))
    acc.set_exact_field(exact)
    try:
        dmc.upload(va._lib.VH_GRID_CSOL, np.zeros(coarse.n ** 3, np.float32))
        for ci in QUAD_KNOTS:
            rec, x0, h = quadratic_cloud(fine, ci, n=NPART)
            acc.upload(rec)
            acc.sort()
            X, quad, lin = quadratic_potentials(fine, ci)
            f = []
            for csol in (quad, lin):
                dmf.upload(va._lib.VH_GRID_CSOL, csol.astype(np.float32))
                outcount, _ = acc.compute_forces(dmf, dmc, 1.0)
                assert outcount == 0
                f.append(acc.forces().astype(np.float64))
            srec = acc.download()
            err = field_ratio_error(f[0][:, 0], f[1][:, 0], tabulated_x(srec, x0, h, sm) - X, sm)
            assert err < 0.1, (ci, err)
    finally:
        acc.set_exact_field(False)
