"""Parity of every stage of the HIP path with the oracle, at config-1 shape
(nfine = 28: 58^3 fine and coarse grids) on a seeded Thomas-Fermi cloud.

Bars: bit-exact for the sort (records, occupied cells, bounds, out-of-stencil
list), the table columns, the Float32 Verlet and the exact-order Float32
contraction of the field; 1e-5 relative (max |diff| / max |ref|) for FP32
densities, potentials and forces (the default matrix-core contraction
included) against the FP64 oracle; 1e-12 for the FP64 interaction energy."""
import numpy as np
import pytest

from conftest import load_package

pytestmark = pytest.mark.gpu
va = load_package()

NPART = 100_000
REL = 1e-5


def relerr(got, ref):
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    return float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))


@pytest.fixture(scope="module")
def setup(config1, orc):
    p, fine, coarse, sm = config1
    w = p.nelectrons / NPART
    pos, mom = va.sample_thomas_fermi(va.read_radial_profile(), NPART, w, seed=-1)
    # a cloud in motion: previous positions half a step back along the momenta
    prev = pos - (p.dt / w) * mom
    x0 = float(fine.axis.knots[0])
    h = float((fine.axis.knots[-1] - fine.axis.knots[0]) / (fine.axis.nknots - 1))
    rec = va.physics.packed_records(pos, prev, x0, h)
    # add a few particles outside the fine grid and outside both grids
    rec[:50]["knode"][:, 0] = 1
    rec[50:60]["knode"][:, 1] = fine.axis.nknots + 3
    rec[60:62]["knode"][:, 2] = -200
    ctx = va.Context(0)
    dmf = va.DeviceMesh(ctx, fine)
    dmc = va.DeviceMesh(ctx, coarse)
    acc = va.DeviceAccelerator(ctx, fine.axis, sm, NPART)
    jel = va.Jellium.of(p.nions)
    ost = orc.OracleStep(fine, coarse, sm, jel, w, w, p.dt, p.rcmax)
    yield dict(p=p, fine=fine, coarse=coarse, sm=sm, w=w, rec=rec, ctx=ctx, dmf=dmf, dmc=dmc,
               acc=acc, jel=jel, ost=ost, x0=x0, h=h)
    acc.close(); dmf.close(); dmc.close(); ctx.close()


def test_sort_bit_exact_and_idempotent(setup, orc):
    s = setup
    acc, rec = s["acc"], s["rec"]
    acc.upload(rec)
    nocc, nout = acc.sort()
    got = acc.download()
    perm, occ, bnd = orc.cellsort(rec, s["fine"].axis.nknots)
    want = rec[perm]
    assert got.tobytes() == want.tobytes()
    assert nocc == len(occ)
    occ_d, bnd_d, out_d = acc.sort_lists(nocc, nout)
    assert np.array_equal(occ_d, occ + 1)
    assert np.array_equal(bnd_d, bnd)
    outl = orc.outside(want, s["fine"].n)
    assert nout == len(outl) and np.array_equal(out_d, outl + 1)
    # sorting a sorted cloud moves nothing
    acc.sort()
    assert acc.download().tobytes() == want.tobytes()


def test_columns_bit_exact(setup, orc):
    s = setup
    acc = s["acc"]
    acc.upload(s["rec"])
    acc.sort()
    srec = acc.download()
    cols, nout = acc.columns()
    ost = s["ost"]
    wc, wn = orc.columns(srec, np.float32(ost.x0), np.float32(ost.h), ost.lo, ost.hi, ost.spacing,
                         s["sm"].nbdt, s["sm"].nbdt + 1)
    assert nout == wn and np.array_equal(cols, wc)


def test_fine_deposit(setup, orc):
    s = setup
    acc, dmf = s["acc"], s["dmf"]
    acc.upload(s["rec"])
    nout = acc.deposit_smoothed(dmf, s["w"])
    rho = dmf.rho
    srec = acc.download()
    perm, occ, bnd = orc.cellsort(srec, s["fine"].axis.nknots)
    assert np.array_equal(perm, np.arange(NPART))
    want, wn = s["ost"].deposit_fine(srec, occ, bnd)
    assert nout == wn
    assert relerr(rho, want) < REL
    # renormalised charge, and deterministic to the bit
    q = dmf.total_charge()
    assert abs(q - (NPART - nout) * s["w"]) < 1e-5 * (NPART - nout) * s["w"]
    acc.deposit_smoothed(dmf, s["w"])
    assert dmf.rho.tobytes() == rho.tobytes()


def test_coarse_deposit(setup, orc):
    s = setup
    acc, dmc = s["acc"], s["dmc"]
    acc.upload(s["rec"])
    acc.sort()
    acc.deposit_cic(dmc, s["w"])
    rho = dmc.rho
    want = s["ost"].deposit_coarse(acc.download())
    assert relerr(rho, want) < REL
    acc.deposit_cic(dmc, s["w"])
    assert dmc.rho.tobytes() == rho.tobytes()


def test_fresh_objects_start_clean(setup, orc):
    """Objects created after others were used and released start from zero:
    the allocator hands back memory holding the earlier owner's grids, and
    the first coarse deposit of a new mesh sums into its accumulator.  Three
    generations of context, meshes and accelerator, each run through the
    chain, each first deposit checked against the oracle."""
    s = setup
    rec = s["rec"]
    want = None
    for _ in range(3):
        ctx = va.Context(0)
        dmf, dmc = va.DeviceMesh(ctx, s["fine"]), va.DeviceMesh(ctx, s["coarse"])
        acc = va.DeviceAccelerator(ctx, s["fine"].axis, s["sm"], NPART)
        try:
            acc.upload(rec)
            acc.deposit_smoothed(dmf, s["w"], count=False)
            acc.deposit_cic(dmc, s["w"])
            rho = dmc.rho
            if want is None:
                want = s["ost"].deposit_coarse(acc.download())
            assert relerr(rho, want) < REL
            va.poisson_nested(dmf, dmc)
            dmf.spline_coefficients()
            acc.compute_forces(dmf, dmc, s["w"])
        finally:
            acc.close(); dmf.close(); dmc.close(); ctx.close()


def test_poisson_chain(setup, orc):
    """nested poisson!, spline_coefficients! and effective_potential! on the
    oracle's own densities (narrowed to Float32)."""
    s = setup
    fine, coarse, dmf, dmc = s["fine"], s["coarse"], s["dmf"], s["dmc"]
    acc = s["acc"]
    acc.upload(s["rec"])
    acc.sort()
    srec = acc.download()
    perm, occ, bnd = orc.cellsort(srec, fine.axis.nknots)
    rho_f, _ = s["ost"].deposit_fine(srec, occ, bnd)
    rho_c = s["ost"].deposit_coarse(srec)
    rho_f32, rho_c32 = rho_f.astype(np.float32), rho_c.astype(np.float32)
    dmf.upload(va._lib.VH_GRID_RHO, rho_f32)
    dmc.upload(va._lib.VH_GRID_RHO, rho_c32)
    # multipole moments in FP64
    mp = dmc.multipole()
    want_mp = orc.multipole(orc.moments(rho_c32.astype(np.float64), coarse.dual_moments))
    assert np.allclose(mp, want_mp, rtol=1e-9, atol=1e-9 * abs(want_mp[0]))
    va.poisson_nested(dmf, dmc)
    phi_f, phi_c, csol_c = dmf.phi, dmc.phi, dmc.csol
    wf, wc, _ = orc.poisson_nested(rho_f32.astype(np.float64), rho_c32.astype(np.float64), fine, coarse)
    assert relerr(phi_c, wc) < REL
    assert relerr(phi_f, wf) < REL
    assert relerr(csol_c, orc.spline_coefficients(wc, coarse)) < REL
    dmf.spline_coefficients()
    csol_f = dmf.csol
    want_csol_f = orc.spline_coefficients(wf, fine)
    assert relerr(csol_f, want_csol_f) < REL
    dmf.effective_potential(s["jel"])
    extra = orc.spline_coefficients(orc.effective_extra(rho_f32.astype(np.float64), fine, s["jel"]), fine)
    want_eff = want_csol_f + extra
    # the Hartree and jellium terms cancel to a few percent of either: the
    # error is held relative to the magnitude of the terms being summed
    scale = max(np.max(np.abs(want_csol_f)), np.max(np.abs(extra)))
    assert np.max(np.abs(dmf.csol - want_eff)) / scale < REL


def test_potential_coefficients_fused(setup, orc):
    """vh_potential_coefficients (csol = S^-1 V, then S^-1 phi added) against
    spline_coefficients followed by effective_potential (S^-1 phi, then
    S^-1 V added: the same two FP32 terms, added once), and against the FP64
    oracle."""
    s = setup
    acc, dmf, fine = s["acc"], s["dmf"], s["fine"]
    acc.upload(s["rec"])
    acc.deposit_smoothed(dmf, s["w"])
    acc.deposit_cic(s["dmc"], s["w"])
    va.poisson_nested(dmf, s["dmc"])
    dmf.spline_coefficients()
    hartree = dmf.csol
    dmf.effective_potential(s["jel"])
    separate = dmf.csol
    dmf.potential_coefficients(s["jel"])
    fused = dmf.csol
    rho = dmf.rho.astype(np.float64)
    extra = orc.spline_coefficients(orc.effective_extra(rho, fine, s["jel"]), fine)
    want = orc.spline_coefficients(dmf.phi.astype(np.float64), fine) + extra
    scale = max(np.max(np.abs(hartree)), np.max(np.abs(extra)))
    assert np.max(np.abs(fused - want)) / scale < REL
    assert fused.tobytes() == separate.tobytes()


def test_spline_coefficients_mfma_layout(setup, orc):
    """The rotating product on a random grid, FP32 MFMA vs FP64."""
    s = setup
    rng = np.random.default_rng(11)
    g = rng.standard_normal(s["fine"].n ** 3).astype(np.float32)
    s["dmf"].upload(va._lib.VH_GRID_PHI, g)
    s["dmf"].spline_coefficients()
    want = orc.spline_coefficients(g.astype(np.float64), s["fine"])
    assert relerr(s["dmf"].csol, want) < REL


@pytest.mark.parametrize("nfine", [16, 17, 33, 64, 80, 99, 128])
def test_rotating_product_shapes(setup, orc, nfine):
    """Every epilogue of the rotating product at the launch shapes it takes:
    one panel per block (n = 34, 36), several (n = 68, 130), a last row tile
    with two rows (n = 34, 66, 130, 162), two block rows with four source
    pairs per thread (n = 162, 200, 256, 258; at 256 and 258 the 65-step K
    slices and three panels in flight), against the FP64 oracle.  S^-1 on a random grid (plain stores), the single-level
    Poisson solve (Hadamard and interior-fill epilogues, at n - 2) and the
    mean field added to a random csol (accumulating epilogue).  The mean
    field is not added to S^-1 of the solved potential here: that potential
    (|phi| ~ 4e3 for a random density) loses 2.5e-5 of its scale at n = 130
    to FP32 rounding in S^-1 even with numpy's own FP32 products, so a bar at
    FP32 rounding of the product would measure the operator, not the kernel."""
    s = setup
    mesh = va.grids.SplineMesh.of(va.grids.uniform_axis(-30.0, 30.0, nfine))
    n = mesh.n
    rng = np.random.default_rng(nfine)
    dm = va.DeviceMesh(s["ctx"], mesh)
    try:
        g = rng.standard_normal(n ** 3).astype(np.float32)
        dm.upload(va._lib.VH_GRID_PHI, g)
        dm.spline_coefficients()
        assert relerr(dm.csol, orc.spline_coefficients(g.astype(np.float64), mesh)) < REL
        rho = rng.random(n ** 3).astype(np.float32)
        dm.upload(va._lib.VH_GRID_RHO, rho)
        dm.poisson()
        want_phi, _ = orc.poisson_single(rho.astype(np.float64), mesh)
        assert relerr(dm.phi, want_phi) < REL
        base = rng.standard_normal(n ** 3).astype(np.float32)
        dm.upload(va._lib.VH_GRID_CSOL, base)
        dm.effective_potential(s["jel"])
        extra = orc.spline_coefficients(orc.effective_extra(rho.astype(np.float64), mesh, s["jel"]),
                                         mesh)
        scale = max(np.max(np.abs(base)), np.max(np.abs(extra)))
        assert np.max(np.abs(dm.csol - (base + extra))) / scale < REL
    finally:
        dm.close()


def _forces_setup(s, orc, seed=5):
    rng = np.random.default_rng(seed)
    fine, coarse = s["fine"], s["coarse"]
    acc = s["acc"]
    acc.upload(s["rec"])
    acc.sort()
    srec = acc.download()
    perm, occ, bnd = orc.cellsort(srec, fine.axis.nknots)
    # smooth potentials: a cluster-like well plus noise, as coefficients
    csol_f = (rng.standard_normal(fine.n ** 3) * 0.01 + 1.0).astype(np.float32)
    csol_c = (rng.standard_normal(coarse.n ** 3) * 0.01 + 1.0).astype(np.float32)
    s["dmf"].upload(va._lib.VH_GRID_CSOL, csol_f)
    s["dmc"].upload(va._lib.VH_GRID_CSOL, csol_c)
    return srec, occ, bnd, csol_f, csol_c


def test_field_contraction_bit_exact(setup, orc):
    """The exact-order variant: bit-identical to contract_spline_10 in FP32."""
    s = setup
    srec, occ, bnd, csol_f, csol_c = _forces_setup(s, orc)
    sm, ost = s["sm"], s["ost"]
    s["acc"].set_exact_field(True)
    try:
        nout, red = s["acc"].compute_forces(s["dmf"], s["dmc"], s["w"], None)
    finally:
        s["acc"].set_exact_field(False)
    f = s["acc"].forces()
    want, _ = orc.field(srec, occ, bnd, ost.nk, s["fine"].n, csol_f.astype(np.float64),
                        sm.overlap, sm.gradient, ost.x0, ost.h, ost.spacing, sm.nbdt, sm.nbdt + 1,
                        s["w"], None, 0, 0.0, f32=True)
    inside = ~np.isnan(want[:, 0])
    assert inside.sum() > 0.9 * NPART
    assert np.array_equal(f[inside].view(np.uint32), want[inside].astype(np.float32).view(np.uint32))
    # the others took the coarse fallback
    outl = orc.outside(srec, s["fine"].n)
    assert nout == len(outl) and np.array_equal(np.nonzero(~inside)[0], outl)
    wfb = orc.fallback(srec, outl, ost.x0, ost.h, s["coarse"].axis.knots, csol_c.astype(np.float64),
                       s["w"], None)
    assert relerr(f[outl], wfb[outl]) < REL


@pytest.mark.parametrize("products", ["split16", "fp32"])
def test_field_matrix_cores(setup, orc, products):
    """The matrix-core variant (first stage on MFMA; fp16 split products by
    default, FP32 products as an option): within FP32 rounding of the FP64
    contraction, identical to the bit from one call to the next, and within
    1e-6 of the exact-order variant (a few FP32 ulps; the split's dropped
    terms are about 2^-21 of the largest product)."""
    s = setup
    srec, occ, bnd, _, csol_c = _forces_setup(s, orc, seed=8)
    sm, ost = s["sm"], s["ost"]
    acc = s["acc"]
    acc.set_field_fp32(products == "fp32")
    try:
        _field_matrix_cores(s, orc, srec, occ, bnd, sm, ost, acc)
    finally:
        acc.set_field_fp32(False)


def _field_matrix_cores(s, orc, srec, occ, bnd, sm, ost, acc):
    # a zero-mean rough potential: no constant part whose cancelling
    # contributions would make the rounding of any summation order look large
    rng = np.random.default_rng(18)
    csol_f = (rng.standard_normal(s["fine"].n ** 3) * 0.01).astype(np.float32)
    s["dmf"].upload(va._lib.VH_GRID_CSOL, csol_f)
    acc.compute_forces(s["dmf"], s["dmc"], s["w"], None)
    f = acc.forces()
    want, _ = orc.field(srec, occ, bnd, ost.nk, s["fine"].n, csol_f.astype(np.float64),
                        sm.overlap, sm.gradient, ost.x0, ost.h, ost.spacing, sm.nbdt, sm.nbdt + 1,
                        s["w"], None, 0, 0.0, f32=False)
    inside = ~np.isnan(want[:, 0])
    assert np.all(np.isfinite(f))
    assert relerr(f[inside], want[inside]) < REL
    acc.compute_forces(s["dmf"], s["dmc"], s["w"], None)
    assert acc.forces().tobytes() == f.tobytes()
    acc.set_exact_field(True)
    try:
        acc.compute_forces(s["dmf"], s["dmc"], s["w"], None)
        fe = acc.forces()
    finally:
        acc.set_exact_field(False)
    assert relerr(f[inside], fe[inside]) < 1e-6


@pytest.mark.parametrize("mode", ["device", "ball"])
def test_field_with_projectile(setup, orc, mode):
    s = setup
    srec, occ, bnd, csol_f, csol_c = _forces_setup(s, orc, seed=6)
    proj = va.xenon_projectile(dt=0.5, x0=-20.0, impact=10.0)
    nout, red = s["acc"].compute_forces(s["dmf"], s["dmc"], s["w"], proj, mode)
    f = s["acc"].forces()
    want, wred, _ = s["ost"].forces(srec, occ, bnd, csol_f.astype(np.float64),
                                    csol_c.astype(np.float64), proj, mode)
    assert relerr(f, want) < REL
    assert np.allclose(red, wred, rtol=1e-5, atol=1e-6 * np.max(np.abs(wred)))


def test_verlet_bit_exact(setup, orc):
    s = setup
    acc = s["acc"]
    acc.upload(s["rec"])
    rng = np.random.default_rng(9)
    f = (rng.standard_normal((NPART, 3)) * 1e-3).astype(np.float32)
    acc.upload_forces(f)
    d = acc.verlet(0.5, s["w"], 100.0)
    got = acc.download()
    want = s["rec"].copy()
    wd = orc.verlet_f32(want, f, 0.5, s["w"], s["h"], s["x0"], 100.0)
    assert got.tobytes() == want.tobytes()
    assert np.allclose([d["kinetic"], d["escaped"], *d["angular"]], wd, rtol=1e-10, atol=1e-12)
    # the push leaves the cell keys and first-digit histograms of the moved
    # records; the sort that starts from them is the stable sort to the bit
    nocc, nout = acc.sort()
    perm, occ, bnd = orc.cellsort(want, s["fine"].axis.nknots)
    assert acc.download().tobytes() == want[perm].tobytes()
    occ_d, bnd_d, out_d = acc.sort_lists(nocc, nout)
    assert np.array_equal(occ_d, occ + 1) and np.array_equal(bnd_d, bnd)


def test_records_left_in_place_by_a_sort(config1, orc):
    # After the field kernel's push the sort moves no record: slot s's record
    # is read through the sort's permutation until the next push moves it or
    # another reader (download, a second sort, vh_verlet) gathers it first.
    p, fine, coarse, sm = config1
    n = p.nparticles
    w = p.nelectrons / n
    nk = fine.axis.nknots
    x0 = float(fine.axis.knots[0])
    h = float((fine.axis.knots[-1] - fine.axis.knots[0]) / (nk - 1))
    pos, mom = va.sample_thomas_fermi(va.read_radial_profile(), n, w, seed=-5)
    sim = va.Simulation(p, None, tables=(fine, coarse, sm), positions=pos, momenta=mom)
    try:
        acc = sim.acc
        rng = np.random.default_rng(11)
        f = (rng.standard_normal((n, 3)) * 1e-3).astype(np.float32)
        for case in ("download", "second sort", "verlet"):
            sim.step()
            rec = acc.download()
            want = rec[orc.cellsort(rec, nk)[0]].copy()
            acc.sort()
            if case == "second sort":
                acc.sort()
            if case == "verlet":
                acc.upload_forces(f)
                acc.verlet(0.5, w, 100.0)
                orc.verlet_f32(want, f, 0.5, w, h, x0, 100.0)
            assert acc.download().tobytes() == want.tobytes(), case
    finally:
        sim.close()


def test_enclosed_count(setup, orc):
    s = setup
    acc = s["acc"]
    acc.upload(s["rec"])
    c = np.array([3.0, -2.0, 1.0])
    assert acc.enclosed_count(c, 8.0) == orc.enclosed(s["rec"], s["x0"], s["h"], c, 8.0)


def test_density_slice_frame(setup):
    # the film frame of scripts/xenon.jl:396-410: Float32 of the FP64 mean of
    # two z planes of the fine density, to the bit
    s = setup
    acc, dmf = s["acc"], s["dmf"]
    acc.upload(s["rec"])
    acc.deposit_smoothed(dmf, s["w"])
    n = dmf.n
    rho = dmf.rho.reshape(n, n, n).astype(np.float64)   # [k, j, i]
    for iz1, iz2 in ((n // 2, n // 2 + 1), (1, n), (7, 7)):
        want = ((rho[iz1 - 1] + rho[iz2 - 1]) / 2).astype(np.float32)
        got = dmf.density_slice(iz1, iz2)
        assert got.tobytes() == want.tobytes(), (iz1, iz2)
    for bad in ((0, 1), (1, n + 1)):
        with pytest.raises(va.VlasovHIPError, match="outside"):
            dmf.density_slice(*bad)


def test_film_observables(config1):
    # frame and q_net of a film step read from the device against the same
    # quantities computed from the downloaded density and cloud
    p, fine, coarse, sm = config1
    n = p.nparticles
    w = p.nelectrons / n
    pos, mom = va.sample_thomas_fermi(va.read_radial_profile(), n, w, seed=-3)
    sim = va.Simulation(p, None, tables=(fine, coarse, sm), positions=pos, momenta=mom)
    try:
        sim.step()
        nf, ng = p.nfine, sim.dmf.n
        rho = sim.dmf.rho.reshape(ng, ng, ng).astype(np.float64)
        want = ((rho[nf - 1] + rho[nf]) / 2).astype(np.float32)
        assert sim.frame().tobytes() == want.tobytes()
        r = sim.positions()
        inside = int(np.count_nonzero(np.sum(r.astype(np.float64) ** 2, axis=1) < 35.0 ** 2))
        assert sim.net_charge(35.0) == p.nelectrons - sim.weight * inside
    finally:
        sim.close()


def test_bad_arguments_refused_before_launch(setup):
    s = setup
    with pytest.raises(va.VlasovHIPError):
        va._lib.call("vh_forces", s["acc"].handle, s["dmf"].handle, s["dmc"].handle, 1.0, None, 7, 0.0,
                     None, None)
    with pytest.raises(ValueError):
        s["acc"].upload(np.zeros(3, dtype=va.PACKED_DTYPE))
    with pytest.raises(va.VlasovHIPError):
        va._lib.call("vh_accel_set_option", s["acc"].handle, 99, 1)
    with pytest.raises(va.VlasovHIPError):
        va._lib.call("vh_accel_set_option", s["acc"].handle, va._lib.VH_OPT_EXACT_FIELD, 2)


def test_interaction_energy(setup, orc):
    """vh_interaction_energy against the FP64 oracle on the same coefficients:
    the same per-particle arithmetic, summed in another order.  The cloud has
    particles in all three regions (smoothed fine, coarse spline, beyond both
    grids)."""
    s = setup
    srec, occ, bnd, csol_f, csol_c = _forces_setup(s, orc, seed=7)
    fine, coarse, sm = s["fine"], s["coarse"], s["sm"]
    for enclosed in (0.0, 3.0):
        got = s["acc"].interaction_energy(s["dmf"], s["dmc"], s["w"], enclosed)
        want = orc.interaction_energy(srec, s["x0"], s["h"], fine.axis.knots, csol_f.astype(np.float64),
                                      sm.overlap, float(sm.spacing), sm.nbdt, coarse.axis.knots,
                                      csol_c.astype(np.float64), s["w"], enclosed)
        assert abs(got - want) <= 1e-12 * abs(want)


def test_step_with_energy_budget(config1, orc):
    """One resident step with the budget: the Hartree energy on the step's
    own Hartree potential, the kinetic energy of the push and the total
    interaction at the pushed positions, against the FP64 oracle."""
    p, fine, coarse, sm = config1
    n = NPART
    w = p.nelectrons / n
    pos, mom = va.sample_thomas_fermi(va.read_radial_profile(), n, w, seed=-3)
    proj = va.xenon_projectile(x0=-20.0, impact=10.0)
    sim = va.Simulation(p, None, projectile=proj, tables=(fine, coarse, sm), positions=pos,
                        momenta=mom)
    try:
        rec0 = sim.acc.download()
        st = orc.OracleStep(fine, coarse, sm, va.Jellium.of(p.nions), w, sim.mass, p.dt, p.rcmax)
        proj_before = va.physics.Projectile(**vars(sim.projectile))
        ref = st.update_forces(rec0, proj_before, energy=True)
        b = sim.step(energy=True)
        f_gpu = sim.acc.forces()
        assert relerr(f_gpu, ref["force"]) < 1e-4
        # the push on the device's own forces is the oracle's to the bit
        rec1, d = st.verlet(ref["srec"], f_gpu)
        assert sim.acc.download().tobytes() == rec1.tobytes()
        assert abs(b.kinetic - d[0]) <= 1e-10 * abs(d[0])
        assert abs(b.hartree - ref["hartree"]) <= REL * abs(ref["hartree"])
        total = st.energy(rec1, ref["csol_f"], ref["csol_c"])
        scale = max(abs(total), abs(b.hartree))
        # the mean field by difference, as the budget takes it; the Hartree and
        # jellium parts of the total potential cancel to a fraction of either
        assert abs(b.meanfield - (total - 2 * ref["hartree"])) <= 2 * REL * scale
        assert b.total == b.ions + b.kinetic + b.hartree + b.meanfield
        assert b.ions == 3 * p.nions ** 2 / (5 * va.Jellium.of(p.nions).radius)
    finally:
        sim.close()


def test_resident_projectile_matches_host(config1):
    """The projectile integrated on the device (so that a step waits for
    nothing on the host) follows the host integration of the same step:
    the cloud identical to the bit after four steps, the projectile's state to
    FP64 rounding (the jellium term's r2^1.5 is the device's pow)."""
    p, fine, coarse, sm = config1
    n = p.nparticles
    w = p.nelectrons / n
    pos, mom = va.sample_thomas_fermi(va.read_radial_profile(), n, w, seed=-5)
    sims = []
    try:
        for resident in (True, False):
            proj = va.xenon_projectile(x0=-20.0, impact=10.0)
            sims.append(va.Simulation(p, None, projectile=proj, tables=(fine, coarse, sm),
                                      positions=pos, momenta=mom, resident_projectile=resident))
        # the resident one steps in one native call, solving in the entry
        # points' order (the host one calls them)
        sims[0].acc.set_poisson_unfused(True)
        for _ in range(4):
            for s in sims:
                s.step()
        a, b = sims[0].projectile, sims[1].projectile
        assert np.allclose(a.position, b.position, rtol=1e-13, atol=1e-12)
        assert np.allclose(a.previous, b.previous, rtol=1e-13, atol=1e-12)
        assert np.allclose(a.velocity, b.velocity, rtol=1e-10, atol=1e-14)
        assert sims[0].acc.download().tobytes() == sims[1].acc.download().tobytes()
    finally:
        for s in sims:
            s.close()


@pytest.mark.parametrize("with_projectile", [False, True])
def test_native_step_matches_separate_calls(config1, with_projectile):
    """vh_step (one native call per step, the field and fallback kernels
    pushing the particles they computed the force of) against the same step
    made of the separate entry points (forces, projectile step, vh_verlet),
    without and with the projectile on the device: identical clouds, forces
    and projectile states after three steps, each step's sort starting from
    the keys and histograms the push left."""
    p, fine, coarse, sm = config1
    n = p.nparticles
    w = p.nelectrons / n
    pos, mom = va.sample_thomas_fermi(va.read_radial_profile(), n, w, seed=-7)
    sims = []
    try:
        for _ in range(2):
            proj = va.xenon_projectile(x0=-20.0, impact=10.0) if with_projectile else None
            sims.append(va.Simulation(p, None, projectile=proj, tables=(fine, coarse, sm),
                                      positions=pos, momenta=mom))
        a, b = sims
        a.acc.set_poisson_unfused(True)  # vh_step solving in the entry points' order
        for _ in range(3):
            a.step()
            b.update_forces(advance=True)
            b.acc.verlet(p.dt, b.mass, p.rcmax, diagnostics=False)
        assert a.acc.download().tobytes() == b.acc.download().tobytes()
        assert a.acc.forces().tobytes() == b.acc.forces().tobytes()
        if with_projectile:
            pa, pb = a.projectile, b.projectile
            assert np.array_equal(pa.position, pb.position)
            assert np.array_equal(pa.velocity, pb.velocity)
    finally:
        for s in sims:
            s.close()


def test_step_graph_replays_match_eager(config1):
    """vh_step recorded into HIP graphs (one per particle buffer) and
    replayed: every replay gives the cloud of the eager step to the bit, with
    host reads, an energy step (the separate entry points) and a projectile
    read between the replays."""
    p, fine, coarse, sm = config1
    n = p.nparticles
    w = p.nelectrons / n
    pos, mom = va.sample_thomas_fermi(va.read_radial_profile(), n, w, seed=-9)
    sims = []
    try:
        for graph in (True, False):
            proj = va.xenon_projectile(x0=-20.0, impact=10.0)
            s = va.Simulation(p, None, projectile=proj, tables=(fine, coarse, sm),
                              positions=pos, momenta=mom)
            s.acc.set_step_graph(graph)
            sims.append(s)
        for i in range(7):
            energy = i == 3
            for s in sims:
                s.step(energy=energy)
            a, b = (s.acc.download() for s in sims)
            assert a.tobytes() == b.tobytes(), f"step {i}"
            if i % 2:
                assert sims[0].projectile.position == sims[1].projectile.position
    finally:
        for s in sims:
            s.close()
