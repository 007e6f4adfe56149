"""The oracle pinned by invariants and closed forms, before it is trusted as
the checker (test/runtests.jl:283-398, 400-466, 637-695, 745-847, 887-931).
CPU only."""
import math

import numpy as np
import pytest

from conftest import load_package

va = load_package()
G = va.grids


def cloud(n, x0, h, seed=1, spread=8.0):
    rng = np.random.default_rng(seed)
    pos = rng.standard_normal((n, 3)) * spread
    prev = pos - rng.standard_normal((n, 3)) * 0.01
    return va.physics.packed_records(pos, prev, x0, h), pos


def test_sort_invariants_and_stability(orc):
    ax = G.uniform_axis(-10.0, 10.0, 8)
    x0, h, nk = ax.knots[0], ax.knots[1] - ax.knots[0], ax.nknots
    rec, _ = cloud(500, x0, h, spread=8.0)
    perm, occ, bnd = orc.cellsort(rec, nk)
    assert np.array_equal(np.sort(perm), np.arange(500))
    keys = np.array([orc_key(r, nk) for r in rec])
    assert np.all(np.diff(keys[perm]) >= 0)
    assert np.array_equal(perm, np.argsort(keys, kind="stable"))         # stable
    assert bnd[0] == 0 and bnd[-1] == 500 and len(bnd) == len(occ) + 1
    for g in range(len(occ)):
        assert np.all(keys[perm[bnd[g]:bnd[g + 1]]] == occ[g]) and bnd[g + 1] > bnd[g]
    assert np.all(np.diff(occ) > 0)
    # idempotent
    p2, o2, b2 = orc.cellsort(rec[perm], nk)
    assert np.array_equal(p2, np.arange(500)) and np.array_equal(o2, occ)


def test_oracle_refuses_strided_views(orc):
    # the C loops read records densely: a strided view must be refused, not
    # read from its base pointer (which would give the wrong records); its
    # contiguous copy sorts like the records it holds
    ax = G.uniform_axis(-10.0, 10.0, 8)
    rec, _ = cloud(400, ax.knots[0], ax.knots[1] - ax.knots[0])
    with pytest.raises(ValueError, match="contiguous"):
        orc.cellsort(rec[::3], ax.nknots)
    sub = np.ascontiguousarray(rec[::3])
    perm, _, _ = orc.cellsort(sub, ax.nknots)
    keys = np.array([orc_key(r, ax.nknots) for r in sub])
    assert np.array_equal(perm, np.argsort(keys, kind="stable"))


def orc_key(r, nk):
    k = np.clip(r["knode"], 1, nk) - 1
    return k[0] + nk * (k[1] + nk * k[2])


def test_outside_list_brute_force(orc):
    ax = G.uniform_axis(-10.0, 10.0, 8)
    x0, h, nk = ax.knots[0], ax.knots[1] - ax.knots[0], ax.nknots
    rec, _ = cloud(2000, x0, h, spread=9.0)
    n = 2 * nk
    got = orc.outside(rec, n)
    k = rec["knode"]
    b = 2 * k - 5
    fits = np.all((b >= 1) & (b + 9 <= n), axis=1)
    assert np.array_equal(got, np.nonzero(~fits)[0])


def test_smoothed_deposit_conservation_and_support(orc, config1):
    p, fine, coarse, sm = config1
    k = fine.axis.knots
    x0, h, nk, n = k[0], k[1] - k[0], fine.axis.nknots, fine.n
    st = orc.OracleStep(fine, coarse, sm, va.Jellium.of(196.0), 196.0 / 5000, 196.0 / 5000, 0.5)
    rec, _ = cloud(5000, x0, h, seed=99)
    srec, occ, bnd = st.sort(rec)
    rho, nout = st.deposit_fine(srec, occ, bnd)
    assert nout == 0
    assert abs(orc.total_charge(rho, fine.dual_moments[0]) - 196.0) < 1e-12 * 196.0
    one, _ = cloud(1, x0, h, seed=3, spread=1.0)
    srec, occ, bnd = st.sort(one)
    raw = orc.deposit_smoothed_raw(srec, occ, bnd, nk, n, sm.nodes, st.spacing, sm.nbdt, sm.nbdt + 1)
    assert np.count_nonzero(raw) == 8 ** 3
    raw = orc.deposit_cic_raw(srec, np.float32(x0), np.float32(h), coarse.axis.colloc.astype(np.float32),
                              np.float32(coarse.locator.x0), np.float32(coarse.locator.invwidth),
                              coarse.locator.cell)
    assert np.count_nonzero(raw) == 2 ** 3


def test_cic_conservation(orc, config1):
    p, fine, coarse, sm = config1
    k = fine.axis.knots
    st = orc.OracleStep(fine, coarse, sm, va.Jellium.of(196.0), 196.0 / 50000, 196.0 / 50000, 0.5)
    rec, _ = cloud(50000, k[0], k[1] - k[0], seed=1234)
    rho = st.deposit_coarse(rec)
    assert abs(orc.total_charge(rho, coarse.dual_moments[0]) - 196.0) < 1e-12 * 196.0


def test_multipole_of_a_gaussian(orc):
    """A centred Gaussian charge: q, no dipole, no quadrupole; shifted: the
    barycentre (test/runtests.jl:339-366)."""
    ax = G.uniform_axis(-20.0, 20.0, 16)
    m = G.SplineMesh.of(ax)
    g = ax.colloc
    kk, jj, ii = np.meshgrid(g, g, g, indexing="ij")
    sig = 3.0
    rho = np.exp(-(ii ** 2 + jj ** 2 + kk ** 2) / (2 * sig ** 2)).ravel()
    rho *= 10.0 / orc.total_charge(rho, m.dual_moments[0])
    mp = orc.multipole(orc.moments(rho, m.dual_moments))
    assert abs(mp[0] - 10.0) < 1e-10 * 10.0
    assert np.max(np.abs(mp[1:4])) < 1e-8
    assert np.max(np.abs(mp[4:])) < 1e-6
    for r in (1e3, 1e4):
        v = orc.multipole_potential(mp, np.array([[r, 0.0, 0.0]]))[0]
        assert abs(v - 10.0 / r) < 1e-6 * 10.0 / r
    shifted = np.exp(-((ii - 4) ** 2 + jj ** 2 + kk ** 2) / (2 * sig ** 2)).ravel()
    mps = orc.multipole(orc.moments(shifted, m.dual_moments))
    assert abs(mps[1] - 4.0) < 1e-6 * 4.0 and abs(mps[2]) < 1e-8


def test_poisson_far_field_and_nested_faces(orc, config1):
    """A Gaussian charge: the solution approaches q/r away from it; the fine
    faces are the coarse spline there (test/runtests.jl:368-398, 887-931)."""
    p, fine, coarse, sm = config1
    gf, gc = fine.axis.colloc, coarse.axis.colloc

    def gauss(mesh, sig=12.0):
        g = mesh.axis.colloc
        kk, jj, ii = np.meshgrid(g, g, g, indexing="ij")
        r = np.exp(-(ii ** 2 + jj ** 2 + kk ** 2) / (2 * sig * sig)).ravel()
        return r * (10.0 / orc.total_charge(r, mesh.dual_moments[0]))

    rf, rc = gauss(fine), gauss(coarse)
    phi_f, phi_c, mp = orc.poisson_nested(rf, rc, fine, coarse)
    assert abs(mp[0] - 10.0) < 1e-10 * 10.0
    # faces of the coarse level: the multipole potential itself
    idx, pts = orc.face_points(gc)
    assert np.allclose(phi_c[idx], orc.multipole_potential(mp, pts), rtol=1e-14, atol=0)
    # interior: the operator applied to the solution gives back the rhs
    for mesh, phi, rho in ((coarse, phi_c, rc), (fine, phi_f, rf)):
        n, ni = mesh.n, mesh.ni
        rhs = orc.poisson_rhs(rho, phi, mesh.laplacian, n).reshape(ni, ni, ni)
        inner = phi.reshape(n, n, n)[1:-1, 1:-1, 1:-1]
        Dint = mesh.laplacian[1:-1, 1:-1]
        lap = inner @ Dint.T + np.einsum("ab,kbi->kai", Dint, inner) + np.einsum("ab,bji->aji", Dint, inner)
        assert np.linalg.norm(lap - rhs) / np.linalg.norm(rhs) < 1e-10
    # far field along +x: q/r to 1 %
    n = fine.n
    r = gf[n - 3]
    got = phi_f.reshape(n, n, n)[n // 2, n // 2, n - 3]
    rr = math.sqrt(r * r + 2 * gf[n // 2] ** 2)
    assert abs(got - 10.0 / rr) < 1e-2 * 10.0 / rr
    # the fine faces are the coarse spline there
    csol_c = orc.spline_coefficients(phi_c, coarse)
    idx, pts = orc.face_points(gf)
    assert np.allclose(phi_f[idx], orc.spline_potential(coarse.axis.knots, csol_c, pts), rtol=0, atol=1e-12)


def test_field_of_a_linear_potential(orc, config1):
    """Phi = x: the unsmoothed spline field is exactly (-1, 0, 0); the smoothed
    one to 1e-4 (test/runtests.jl:637-695)."""
    p, fine, coarse, sm = config1
    k = fine.axis.knots
    x0, h, nk, n = k[0], k[1] - k[0], fine.axis.nknots, fine.n
    # coefficients of phi = x: value x at each knot, slope 1
    def lin(ax):
        a = np.zeros(ax.nbasis)          # x: value x_k, slope 1
        a[0::2] = ax.knots
        a[1::2] = 1.0
        one = np.zeros(ax.nbasis)        # 1: value 1, slope 0
        one[0::2] = 1.0
        return (one[:, None, None] * one[None, :, None] * a[None, None, :]).ravel()
    csol_f, csol_c = lin(fine.axis), lin(coarse.axis)
    rec, pos = cloud(3000, x0, h, seed=4, spread=15.0)
    st = orc.OracleStep(fine, coarse, sm, va.Jellium.of(196.0), 1.0, 1.0, 0.5)
    srec, occ, bnd = st.sort(rec)
    f, red = orc.field(srec, occ, bnd, nk, n, csol_f, sm.overlap, sm.gradient, x0, h, st.spacing,
                       sm.nbdt, sm.nbdt + 1, 1.0)
    ins = ~np.isnan(f[:, 0])
    assert ins.sum() > 2500
    # away from the origin the tabulated kernel's normalisation defect (1.3e-5
    # on a constant's derivative) grows with the potential: a percent bound
    assert np.max(np.abs(f[ins] - np.array([-1.0, 0.0, 0.0]))) < 1e-2
    # the reference's own case: 50 a0 box, 28 intervals, 200-point tables,
    # three points, 1e-4 (test/runtests.jl:637-686)
    ax = G.uniform_axis(-50.0, 50.0, 28)
    sm2 = G.GaussianSmoothing.of(ax, nbdt=200, quadrature=200)
    pts = np.array([(2.3, -1.7, 4.1), (-8.0, 0.5, -3.3), (11.2, 6.6, -9.9)])
    x0b, hb = ax.knots[0], ax.knots[1] - ax.knots[0]
    r3 = va.physics.packed_records(pts, pts, x0b, hb)
    perm, occ, bnd = orc.cellsort(r3, ax.nknots)
    f3, _ = orc.field(r3[perm], occ, bnd, ax.nknots, ax.nbasis, lin(ax), sm2.overlap, sm2.gradient,
                      x0b, hb, np.float32(sm2.spacing), 200, 201, 1.0)
    assert np.max(np.abs(f3 - np.array([-1.0, 0.0, 0.0]))) < 1e-4
    outl = np.arange(len(srec), dtype=np.int32)
    fb = orc.fallback(srec, outl, x0, h, coarse.axis.knots, csol_c, 1.0, None)
    assert np.max(np.abs(fb - np.array([-1.0, 0.0, 0.0]))) < 1e-12


def test_projectile_laws_closed_forms(orc):
    """Ball: m = 1/r^3 outside, 1/R^3 inside; Gaussian: series and direct
    forms agree at the switch and tend to 1/r^3 (test/runtests.jl:1343-1401)."""
    P = va.physics
    b = P.BallSoftening(5.0)
    assert P.force_kernel(b, 36.0) == 1 / (36.0 * 6.0)
    assert P.force_kernel(b, 4.0) == 1 / 125.0
    s = 5.0
    r2 = 0.25 * s * s
    lo = P.gaussian_force_kernel(r2 * (1 - 1e-12), s)
    hi = P.gaussian_force_kernel(r2 * (1 + 1e-12), s)
    assert abs(lo - hi) < 1e-9 * lo
    assert abs(P.gaussian_force_kernel(1e4, s) - 1e-6) < 1e-12
    # series vs mpmath-free direct form in long double style: at u = 0.4
    r = 0.4 * s
    direct = (math.erf(r / (math.sqrt(2) * s)) - math.sqrt(2 / math.pi) * (r / s) * math.exp(-0.08)) / r ** 3
    assert abs(P.gaussian_force_kernel(r * r, s) - direct) < 1e-9 * direct


def test_verlet_free_flight_and_packed_equivalence(orc):
    """No force: the packed integrator moves every particle by exactly its
    previous displacement, and the FP32 packed step agrees with a FP64
    absolute-coordinate step to Float32 resolution (test/runtests.jl:400-596)."""
    x0, h = -78.0, 2.4375
    rng = np.random.default_rng(2)
    pos = rng.uniform(-60, 60, (1000, 3))
    vel = rng.standard_normal((1000, 3)) * 0.3
    rec = va.physics.packed_records(pos, pos - vel, x0, h)
    f = np.zeros((1000, 3), np.float32)
    r1 = rec.copy()
    orc.verlet_f32(r1, f, 0.5, 1.0, h, x0, 100.0)
    q1 = va.physics.record_positions(r1, x0, h)
    q0 = va.physics.record_positions(rec, x0, h)
    qm = va.physics.record_positions(rec, x0, h, previous=True)
    assert np.max(np.abs(q1 - (2 * q0 - qm))) < 4e-6
    assert np.array_equal(r1["pknode"], rec["knode"]) and np.array_equal(r1["pdelta"], rec["delta"])
    # with a force: q1 = 2 q0 - qm + dt^2/M F
    F = rng.standard_normal((1000, 3)).astype(np.float32)
    r2 = rec.copy()
    d = orc.verlet_f32(r2, F, 0.5, 2.0, h, x0, 100.0)
    q2 = va.physics.record_positions(r2, x0, h)
    assert np.max(np.abs(q2 - (2 * q0 - qm + 0.125 * F))) < 4e-6
    pgap = (q2 - qm) * (2.0 / (2 * 0.5))
    ek = np.sum(pgap ** 2) / (2 * 2.0)
    assert abs(d[0] - ek) < 1e-4 * ek
    L = np.sum(np.cross(q0, pgap), axis=0)
    assert np.allclose(d[2:], L, rtol=1e-4, atol=1e-3)


def test_oracle_step_runs_and_conserves(orc, config1):
    p, fine, coarse, sm = config1
    n = 20_000
    w = 196.0 / n
    pos, mom = va.sample_thomas_fermi(va.read_radial_profile(), n, w)
    k = fine.axis.knots
    rec = va.physics.packed_records(pos, pos - (0.5 / w) * mom, k[0], k[1] - k[0])
    st = orc.OracleStep(fine, coarse, sm, va.Jellium.of(196.0), w, w, 0.5)
    r = st.update_forces(rec, va.xenon_projectile())
    assert abs(orc.total_charge(r["rho_f"], fine.dual_moments[0]) - (n - r["nout"]) * w) < 1e-9
    assert abs(orc.total_charge(r["rho_c"], coarse.dual_moments[0]) - 196.0) < 1e-9
    assert np.all(np.isfinite(r["force"]))
    assert abs(r["mp"][0] - 196.0) < 1e-9


def _hermite_dofs(ax, f, df):
    """Coefficients of the cubic Hermite interpolant of f: value, slope per knot."""
    a = np.zeros(ax.nbasis)
    a[0::2] = f(ax.knots)
    a[1::2] = df(ax.knots)
    return a


def _tensor(fx, fy, fz):
    """Julia-order flat tensor product (x fastest)."""
    return (fz[:, None, None] * fy[None, :, None] * fx[None, None, :]).ravel()


def _separable(ax, terms):
    """sum over terms of fx(x) fy(y) fz(z), each factor given as (f, df)."""
    return sum(_tensor(*(_hermite_dofs(ax, *t) for t in term)) for term in terms)


ONE = (lambda x: np.ones_like(x), lambda x: np.zeros_like(x))
LIN = (lambda x: x, lambda x: np.ones_like(x))
CUB = (lambda x: x ** 3 / 1e4, lambda x: 3 * x ** 2 / 1e4)
SQR = (lambda x: -x ** 2 / 50, lambda x: -x / 25)


def test_interaction_energy_closed_forms(orc, config1):
    """interaction_energy (src/energy.jl:19-43), branch by branch:
    * inside (knots[3], knots[end-2])^3 the smoothed potential of a linear
      Phi is Phi itself (a symmetric kernel), up to the tables' normalisation
      defect -- an error of one table column (h/1000) or one knot fails it;
    * elsewhere the coarse spline reproduces a cubic exactly;
    * beyond both grids, enclosed / |r|;
    * the region test is the strict one on FP64 positions."""
    p, fine, coarse, sm = config1
    kf, kc = fine.axis.knots, coarse.axis.knots
    x0, h = float(kf[0]), float((kf[-1] - kf[0]) / (len(kf) - 1))
    lin3 = [(LIN, ONE, ONE), (ONE, LIN, ONE), (ONE, ONE, LIN)]
    csol_f = _separable(fine.axis, [(LIN, ONE, ONE)]) + 2 * _separable(fine.axis, [(ONE, LIN, ONE)]) \
        + 3 * _separable(fine.axis, [(ONE, ONE, LIN)])
    cubic = [(CUB, ONE, ONE), (ONE, LIN, ONE), (ONE, ONE, SQR), (ONE, ONE, ONE)]
    csol_c = _separable(coarse.axis, cubic)

    def cubic_at(q):
        return q[:, 0] ** 3 / 1e4 + q[:, 1] - q[:, 2] ** 2 / 50 + 1

    rng = np.random.default_rng(17)
    pos = np.concatenate([rng.uniform(-70, 70, (3000, 3)),          # mostly inside
                          rng.uniform(-230, 230, (3000, 3)),        # mostly coarse
                          rng.uniform(-500, 500, (500, 3)) + 600])  # beyond both
    rec = va.physics.packed_records(pos, pos, x0, h)
    q = va.physics.unsplit(rec["knode"], rec["delta"], x0, h)
    w, encl = 0.37, 7.5
    tot, phi = orc.interaction_energy(rec, x0, h, kf, csol_f, sm.overlap, float(sm.spacing), sm.nbdt,
                                      kc, csol_c, w, encl, per_particle=True)
    inside = np.all((q > kf[2]) & (q < kf[-3]), axis=1)
    beyond = np.any((q < kc[0]) | (q > kc[-1]), axis=1)
    coarse_pts = ~inside & ~beyond
    assert inside.sum() > 1000 and coarse_pts.sum() > 1000 and beyond.sum() > 400
    assert np.max(np.abs(phi[coarse_pts] - cubic_at(q[coarse_pts]))) < 1e-9
    assert np.allclose(phi[beyond], encl / np.linalg.norm(q[beyond], axis=1), rtol=1e-14)
    assert math.isclose(tot, w * phi.sum(), rel_tol=1e-12)
    # the smoothed potential of a constant is the kernel's normalisation (~1)
    one_f = _separable(fine.axis, [(ONE, ONE, ONE)])
    _, phi1 = orc.interaction_energy(rec, x0, h, kf, one_f, sm.overlap, float(sm.spacing), sm.nbdt,
                                     kc, np.zeros_like(csol_c), 1.0, 0.0, per_particle=True)
    assert np.max(np.abs(phi1[inside] - 1.0)) < 5e-5
    assert np.all(phi1[~inside] == 0.0)
    # normalised by it, the linear potential comes back to Phi at the
    # tabulated offset (the table samples the offset from the nearest knot
    # every spacing/nbdt), up to the kernel's first-moment defect: far below
    # the spacing/nbdt x |grad Phi| of a one-column slip or a knot slip
    qi = q[inside]
    ci = np.argmin(np.abs(qi[:, :, None] - kf[None, None, :]), axis=2)   # ties: first (left)
    sp = float(sm.spacing)
    col = np.floor((qi - kf[ci] + sp / 2) / sp * sm.nbdt + 0.5)
    xq = kf[ci] + col * sp / sm.nbdt - sp / 2
    lin_q = xq[:, 0] + 2 * xq[:, 1] + 3 * xq[:, 2]
    err = np.max(np.abs(phi[inside] / phi1[inside] - lin_q))
    assert err < 0.05 * sp / sm.nbdt
