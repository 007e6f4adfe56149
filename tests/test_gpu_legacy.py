"""The legacy accelerator route (ForceAccelerator(ArrayT, fine, sm, npart, n),
update_forces!'s accelerator branch, src/simulation.jl:318-345) on the HIP
path, against the oracle at config-1 shape (1e5 particles, nfine = 28).

The cloud is held on the host as FP64 positions, some of them outside the
fine grid and some beyond both grids.  The oracle plays the caller's FP64
host half (CIC deposit, nested Poisson, coefficients, mean field) on the
density the device deposited, as the reference's host code would.

Bars: the packed and sorted records bit-exact; the density within 1e-5 of
the oracle's FP64 deposit of the same records; the forces, in the cloud's
own order, within 1e-5 (in-stencil FP32 field) and 1e-12 (the FP64 fallback
at the particles' FP64 positions); the projectile reduction within 1e-5."""
import numpy as np
import pytest

from conftest import load_package

pytestmark = pytest.mark.gpu
va = load_package()
NPART = 100_000


def relerr(got, ref):
    return float(np.max(np.abs(np.asarray(got, np.float64) - ref)) / np.max(np.abs(ref)))


@pytest.fixture(scope="module")
def legacy(config1, orc):
    p, fine, coarse, sm = config1
    w = p.nelectrons / NPART
    pos, _ = va.sample_thomas_fermi(va.read_radial_profile(), NPART, w, seed=-13)
    pos[:40, 0] += 70.0          # outside the fine grid's stencils
    pos[40:45, 1] = 400.0        # beyond both grids
    ctx = va.Context(0)
    fa = va.ForceAccelerator(ctx, fine, sm, NPART)
    kn = fine.axis.knots
    x0, h = float(kn[0]), float((kn[-1] - kn[0]) / (len(kn) - 1))
    yield dict(p=p, fine=fine, coarse=coarse, sm=sm, w=w, pos=pos, ctx=ctx, fa=fa, x0=x0, h=h,
               st=orc.OracleStep(fine, coarse, sm, va.Jellium.of(p.nions), w, w, p.dt, p.rcmax))
    fa.close()
    ctx.close()


@pytest.mark.parametrize("mode", ["ball", "gaussian"])
def test_legacy_route_against_oracle(legacy, orc, mode):
    s = legacy
    fine, coarse, sm, w, pos, fa, st = (s[k] for k in ("fine", "coarse", "sm", "w", "pos", "fa", "st"))
    n = fine.n
    # deposit_smoothed!(ρf, accelerator, fine, sm, cloud.positions; charge = w)
    rho = np.zeros(n ** 3)
    nout = fa.deposit_smoothed(rho, pos, w)
    rec = orc.pack_kd(pos, s["x0"], s["h"], fine.axis.nknots)
    perm, occ, bnd = orc.cellsort(rec, fine.axis.nknots)
    srec = rec[perm].copy()
    assert fa.acc.download().tobytes() == srec.tobytes()
    rho_ref, nout_ref = st.deposit_fine(srec, occ, bnd)
    assert nout == nout_ref
    assert relerr(rho, rho_ref) < 1e-5
    # the caller's FP64 host half, on the device's density
    rho_c = st.deposit_coarse(srec)
    phi_f, phi_c, _ = orc.poisson_nested(rho, rho_c, fine, coarse)
    jel = va.Jellium.of(s["p"].nions)
    csol_f = orc.effective_potential(orc.spline_coefficients(phi_f, fine), rho, fine, jel)
    csol_c = orc.effective_potential(orc.spline_coefficients(phi_c, coarse), rho_c, coarse, jel)
    soft = va.BallSoftening(5.0) if mode == "ball" else va.GaussianSoftening(5.0)
    proj = va.xenon_projectile(x0=-20.0, impact=10.0, softening=soft)
    # forces!(cloud, accelerator, fine.axes, csolf, coarse.axes, csolc, sm;
    #         projectile, packed = true)
    f = np.zeros((NPART, 3))
    outcount = fa.forces(f, pos, csol_f, coarse.axis.knots, csol_c, w, proj)
    # the oracle: the field of the sorted records (in-stencil, device law of
    # the projectile), mapped back to the cloud's order; the fallback at the
    # cloud's own FP64 positions with the projectile's law
    pj = np.array([*proj.position, -w * proj.charge, proj.scale])
    fs, red = orc.field(srec, occ, bnd, fine.axis.nknots, n, csol_f, sm.overlap, sm.gradient,
                        s["x0"], s["h"], st.spacing, sm.nbdt, sm.nbdt + 1, w, pj, 0, 0.0)
    outl = orc.outside(srec, n)
    assert outcount == len(outl)
    want = np.zeros((NPART, 3))
    want[perm] = fs
    out_idx = perm[outl]
    # the five beyond both grids are among them, and some of the shifted ones
    assert set(range(40, 45)) <= set(out_idx.tolist()) and np.sum(out_idx < 40) >= 10
    ball = mode == "ball"
    radius = 5.0 if ball else proj.scale
    want[out_idx] = orc.fallback_points(pos[out_idx], coarse.axis.knots, csol_c, w,
                                        proj=np.array([*proj.position, proj.charge]), ball=ball,
                                        radius=radius, escaped=0.0)
    ins = np.ones(NPART, bool)
    ins[out_idx] = False
    assert relerr(f[ins], want[ins]) < 1e-5
    assert np.max(np.abs(f[out_idx] - want[out_idx])) <= 1e-12 * np.max(np.abs(want[out_idx]))
    assert relerr(fa.reduction, red) < 1e-5
    # projectile_forces!(cloud, accelerator, proj, jel)
    F, e_el, _ = fa.projectile_forces(w, proj, jel)
    assert np.all(np.isfinite(F)) and np.isfinite(e_el)


def test_legacy_forces_need_a_deposit_first(config1):
    """forces! of a host-held cloud reads the sort its deposit made: without
    one the call is refused before any launch."""
    p, fine, coarse, sm = config1
    ctx = va.Context(0)
    fa = va.ForceAccelerator(ctx, fine, sm, 1000)
    try:
        f = np.zeros((1000, 3))
        with pytest.raises(va.VlasovHIPError, match="vh_positions_upload"):
            fa.forces(f, np.zeros((1000, 3)), np.zeros(fine.n ** 3), coarse.axis.knots,
                      np.zeros(coarse.n ** 3), 0.1)
    finally:
        fa.close()
        ctx.close()
