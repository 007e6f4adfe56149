"""Properties that do not depend on the size, checked at the benchmark's
full size: 8e6 pseudo-particles of the Na196 cloud on the 130^3 grids of
nfine = 64, with the Xe25+ projectile on the device.

Bars: the sort bit-exact against the oracle's stable counting sort (records,
occupied cells, bounds); the charge of the fine deposit equal to
(N - nout) w to FP32 summation accuracy, and every deposit reproducible to
the bit; one vh_step (forces and push fused) bit-identical to the separate
entry points (forces, projectile step, vh_verlet), records, forces and
projectile alike, over two steps; the matrix-core field within 1e-6 of the
exact-order kernel on every particle, and the exact-order kernel equal to the
oracle's FP32 emulation of the reference's FMA chain, bit for bit, and within
1e-5 of its FP64 contraction, on every 97th particle; both densities within
1e-5 of the oracle's FP64 deposits; the nested Poisson solve and the coarse
coefficients within 1e-5 of the oracle's FP64 chain; vh_verlet bit-exact
against the oracle's FP32 restatement; the interaction energy within 1e-11
and the enclosed counts exact against the oracle."""
import numpy as np
import pytest

from conftest import load_package

pytestmark = pytest.mark.gpu
va = load_package()

NPART = 8_000_000
NFINE = 64


@pytest.fixture(scope="module")
def full():
    p = va.xenon_parameters(NPART, NFINE)
    sim = va.Simulation(p, va.read_radial_profile(), projectile=va.xenon_projectile())
    yield p, sim
    sim.close()


def test_full_size_sort_bit_exact(full, orc):
    p, sim = full
    acc = sim.acc
    rec = acc.download()
    # an unsorted cloud: the sorted one shuffled
    shuffled = rec[np.random.default_rng(5).permutation(NPART)]
    acc.upload(shuffled)
    nocc, nout = acc.sort()
    perm, occ, bnd = orc.cellsort(shuffled, sim.fine.axis.nknots)
    assert acc.download().tobytes() == shuffled[perm].tobytes()
    assert nocc == len(occ)
    occ_d, bnd_d, _ = acc.sort_lists(nocc, nout)
    assert np.array_equal(occ_d, occ + 1) and np.array_equal(bnd_d, bnd)
    acc.upload(rec)


def test_full_size_deposits_conserve_and_repeat(full):
    p, sim = full
    acc, dmf, dmc = sim.acc, sim.dmf, sim.dmc
    nout = acc.deposit_smoothed(dmf, sim.weight)
    rho = dmf.rho
    q = dmf.total_charge()
    want = (NPART - nout) * sim.weight
    assert abs(q - want) < 1e-5 * want
    acc.deposit_smoothed(dmf, sim.weight)
    assert dmf.rho.tobytes() == rho.tobytes()
    acc.deposit_cic(dmc, sim.weight)
    rc = dmc.rho
    acc.deposit_cic(dmc, sim.weight)
    assert dmc.rho.tobytes() == rc.tobytes()


def test_full_size_step_matches_separate_calls(full):
    p, sim = full
    pos = sim.positions()
    # the momenta that give the device's previous positions back are not
    # needed: both runs start from the same records
    rec = sim.acc.download()
    runs = []
    try:
        for _ in range(2):
            s = va.Simulation(p, None, projectile=va.xenon_projectile(),
                              tables=(sim.fine, sim.coarse, sim.smoothing),
                              positions=pos, momenta=np.zeros_like(pos))
            s.acc.upload(rec)
            runs.append(s)
        a, b = runs
        a.acc.set_poisson_unfused(True)  # vh_step solving in the entry points' order
        for _ in range(2):
            a.step()
            b.update_forces(advance=True)
            b.acc.verlet(p.dt, b.mass, p.rcmax, diagnostics=False)
        assert a.acc.download().tobytes() == b.acc.download().tobytes()
        assert a.acc.forces().tobytes() == b.acc.forces().tobytes()
        assert np.array_equal(a.projectile.position, b.projectile.position)
        assert np.array_equal(a.projectile.velocity, b.projectile.velocity)
    finally:
        for s in runs:
            s.close()


def test_full_size_field_matrix_cores_against_exact_order(full):
    # the dominant kernel on the full cloud and grid, against the kernel
    # that runs the reference's FMA order (pinned to it bit for bit at the
    # parity size): a rough zero-mean fine potential, every particle whose
    # stencil fits, 1e-6 relative (the two orders' FP32 rounding)
    p, sim = full
    acc, dmf, dmc = sim.acc, sim.dmf, sim.dmc
    acc.sort()
    rng = np.random.default_rng(19)
    csol = (rng.standard_normal(sim.fine.n ** 3) * 0.01).astype(np.float32)
    dmf.upload(va._lib.VH_GRID_CSOL, csol)
    nout, _ = acc.compute_forces(dmf, dmc, sim.weight, None)
    f = acc.forces().astype(np.float64)
    acc.set_exact_field(True)
    try:
        acc.compute_forces(dmf, dmc, sim.weight, None)
        fe = acc.forces().astype(np.float64)
    finally:
        acc.set_exact_field(False)
    assert np.all(np.isfinite(f)) and nout < NPART // 10
    assert np.max(np.abs(f - fe)) / np.max(np.abs(fe)) < 1e-6


def test_full_size_field_against_oracle(full, orc):
    # the field on the full cloud and grid against the oracle on every 97th
    # particle (the oracle's field of a particle depends only on its own
    # record): a rough zero-mean fine potential, the exact-order kernel bit
    # for bit against the FP32 emulation, the matrix-core kernel at 1e-5
    # relative against the FP64 contraction
    p, sim = full
    acc, dmf, dmc = sim.acc, sim.dmf, sim.dmc
    fine, sm = sim.fine, sim.smoothing
    acc.sort()
    srec = acc.download()
    rng = np.random.default_rng(19)
    csol = (rng.standard_normal(fine.n ** 3) * 0.01).astype(np.float32)
    dmf.upload(va._lib.VH_GRID_CSOL, csol)
    acc.compute_forces(dmf, dmc, sim.weight, None)
    f = acc.forces()[::97]
    acc.set_exact_field(True)
    try:
        acc.compute_forces(dmf, dmc, sim.weight, None)
        fe = acc.forces()[::97]
    finally:
        acc.set_exact_field(False)
    sub = np.ascontiguousarray(srec[::97])
    _, occ, bnd = orc.cellsort(sub, fine.axis.nknots)
    ost = orc.OracleStep(fine, sim.coarse, sm, va.Jellium.of(p.nions), sim.weight, sim.weight,
                         p.dt, p.rcmax)
    args = (sub, occ, bnd, ost.nk, fine.n, csol.astype(np.float64), sm.overlap, sm.gradient,
            ost.x0, ost.h, ost.spacing, sm.nbdt, sm.nbdt + 1, sim.weight, None, 0, 0.0)
    w32, _ = orc.field(*args, f32=True)
    w64, _ = orc.field(*args, f32=False)
    inside = ~np.isnan(w64[:, 0])
    assert inside.sum() > 0.9 * len(sub)
    assert fe[inside].tobytes() == w32[inside].astype(np.float32).tobytes()
    ref = w64[inside]
    assert np.max(np.abs(f[inside] - ref)) / np.max(np.abs(ref)) < 1e-5


def test_full_size_deposits_against_oracle(full, orc):
    # both densities of the full cloud against the oracle's FP64 deposits,
    # 1e-5 relative (max |diff| / max |ref|)
    p, sim = full
    acc, dmf, dmc = sim.acc, sim.dmf, sim.dmc
    nout = acc.deposit_smoothed(dmf, sim.weight)
    acc.deposit_cic(dmc, sim.weight)
    srec = acc.download()
    _, occ, bnd = orc.cellsort(srec, sim.fine.axis.nknots)
    ost = orc.OracleStep(sim.fine, sim.coarse, sim.smoothing, va.Jellium.of(p.nions), sim.weight,
                         sim.weight, p.dt, p.rcmax)
    want_f, wn = ost.deposit_fine(srec, occ, bnd)
    assert nout == wn
    got_f = dmf.rho.astype(np.float64)
    assert np.max(np.abs(got_f - want_f)) / np.max(np.abs(want_f)) < 1e-5
    want_c = ost.deposit_coarse(srec)
    got_c = dmc.rho.astype(np.float64)
    assert np.max(np.abs(got_c - want_c)) / np.max(np.abs(want_c)) < 1e-5


def test_full_size_nested_poisson_against_oracle(full, orc):
    # the two-level solve on the full grids from the device's own densities
    # (narrowed to FP32 on both sides), against the oracle's FP64 chain at
    # 1e-5 relative; the coarse coefficients likewise
    p, sim = full
    acc, dmf, dmc = sim.acc, sim.dmf, sim.dmc
    acc.deposit_smoothed(dmf, sim.weight)
    acc.deposit_cic(dmc, sim.weight)
    rho_f, rho_c = dmf.rho.astype(np.float64), dmc.rho.astype(np.float64)
    va.poisson_nested(dmf, dmc)
    wf, wc, _ = orc.poisson_nested(rho_f, rho_c, sim.fine, sim.coarse)
    for got, want in ((dmc.phi, wc), (dmf.phi, wf),
                      (dmc.csol, orc.spline_coefficients(wc, sim.coarse))):
        got = got.astype(np.float64)
        assert np.max(np.abs(got - want)) / np.max(np.abs(want)) < 1e-5


def test_full_size_verlet_bit_exact(full, orc):
    # vh_verlet on the full cloud with random forces against the oracle's
    # FP32 restatement of _verlet_packed, to the bit; its FP64 diagnostics
    p, sim = full
    acc = sim.acc
    acc.sort()
    rec = acc.download()
    knots = sim.fine.axis.knots
    x0 = float(knots[0])
    h = float((knots[-1] - knots[0]) / (sim.fine.axis.nknots - 1))
    rng = np.random.default_rng(23)
    f = (rng.standard_normal((NPART, 3)) * 1e-3).astype(np.float32)
    acc.upload_forces(f)
    d = acc.verlet(p.dt, sim.mass, p.rcmax)
    want = rec.copy()
    wd = orc.verlet_f32(want, f, p.dt, sim.mass, h, x0, p.rcmax)
    assert acc.download().tobytes() == want.tobytes()
    assert np.allclose([d["kinetic"], d["escaped"], *d["angular"]], wd, rtol=1e-9, atol=1e-12)


def test_full_size_energy_and_enclosed(full, orc):
    # the interaction energy of the full cloud on the step's own coefficients
    # against the oracle's FP64 sum (same per-particle arithmetic, another
    # summation order: 1e-11 relative), and the enclosed count, exact
    p, sim = full
    acc, dmf, dmc = sim.acc, sim.dmf, sim.dmc
    sim.step()
    rec = acc.download()
    fine, coarse, sm = sim.fine, sim.coarse, sim.smoothing
    knots = fine.axis.knots
    x0 = float(knots[0])
    h = float((knots[-1] - knots[0]) / (fine.axis.nknots - 1))
    csol_f, csol_c = dmf.csol.astype(np.float64), dmc.csol.astype(np.float64)
    got = acc.interaction_energy(dmf, dmc, sim.weight, 3.0)
    want = orc.interaction_energy(rec, x0, h, knots, csol_f, sm.overlap, float(sm.spacing),
                                  sm.nbdt, coarse.axis.knots, csol_c, sim.weight, 3.0)
    assert abs(got - want) <= 1e-11 * abs(want)
    c = np.array(sim.projectile.position, dtype=np.float64)
    for r in (8.0, 35.0):
        assert acc.enclosed_count(c, r) == orc.enclosed(rec, x0, h, c, r)


def test_full_size_fused_step_coefficients_against_oracle(full, orc):
    # vh_step's default solve (each level's coefficients by linearity, phi's
    # interior never formed) on the full grids: the fine and coarse
    # coefficients (Hartree plus mean field) against the oracle's FP64 chain
    # on the step's own densities, at 1e-5 of the magnitude of the terms
    p, sim = full
    sim.step()
    rho_f, rho_c = sim.dmf.rho.astype(np.float64), sim.dmc.rho.astype(np.float64)
    wf, wc, _ = orc.poisson_nested(rho_f, rho_c, sim.fine, sim.coarse)
    for mesh, dm, phi, rho in ((sim.fine, sim.dmf, wf, rho_f), (sim.coarse, sim.dmc, wc, rho_c)):
        hartree = orc.spline_coefficients(phi, mesh)
        extra = orc.spline_coefficients(orc.effective_extra(rho, mesh, sim.jellium), mesh)
        scale = max(np.max(np.abs(hartree)), np.max(np.abs(extra)))
        assert np.max(np.abs(dm.csol - (hartree + extra))) / scale < 1e-5
