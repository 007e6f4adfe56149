"""The batching extension: independent collisions stepped as one HIP graph
(vh_batch_*), and the work-unit autotuner (vh_autotune_unit).

* A batch of three collisions of the config-4 family (three projectiles,
  three particle counts) stepped as one recorded graph, with an energy step of
  one member between the batch steps, gives every member the cloud and the
  projectile of the same member stepped alone, to the bit; so does the batch
  without the graph.
* The work-unit size is a tile size of the per-cell kernels only: the sort
  and the field do not depend on it (bit-identical), the deposits within FP32
  summation; the autotuner keeps one of its candidates and times every one."""
import numpy as np
import pytest

from conftest import load_package

pytestmark = pytest.mark.gpu
va = load_package()


def family_small(config1):
    p, fine, coarse, sm = config1
    fam = va.batch.collision_family(count=3, nmin=2e4, nmax=6e4, nfine=28, seed=3)
    return fam, (fine, coarse, sm)


def make(fam, tables):
    return [va.Simulation(c.params, va.read_radial_profile(),
                          projectile=va.physics.Projectile(**vars(c.projectile)), tables=tables)
            for c in fam]


@pytest.mark.parametrize("graph", [True, False])
def test_batch_matches_members_alone(config1, graph):
    fam, tables = family_small(config1)
    alone, batched = make(fam, tables), make(fam, tables)
    batch = va.batch.Batch(batched, graph=graph)
    try:
        for i in range(7):
            if i == 3:
                # one member's energy step between batch steps (separate
                # entry points: another host state, recorded anew)
                alone[0].step(energy=True)
                batched[0].step(energy=True)
            batch.step()
            for s in alone:
                s.step()
        batch.sync()
        for a, b in zip(alone, batched):
            assert a.acc.download().tobytes() == b.acc.download().tobytes()
            assert np.array_equal(a.projectile.position, b.projectile.position)
            assert np.array_equal(a.projectile.velocity, b.projectile.velocity)
    finally:
        batch.close()
        for s in alone + batched:
            s.close()


def test_unit_size_and_autotuner(config1, orc):
    p, fine, coarse, sm = config1
    n = 100_000
    w = p.nelectrons / n
    pos, mom = va.sample_thomas_fermi(va.read_radial_profile(), n, w, seed=-17)
    sims = [va.Simulation(p, None, tables=(fine, coarse, sm), positions=pos, momenta=mom)
            for _ in range(2)]
    try:
        a, b = sims
        va._lib.call("vh_accel_set_option", b.acc.handle, va._lib.VH_OPT_UNIT_SLOTS, 128)
        with pytest.raises(va.VlasovHIPError):
            va._lib.call("vh_accel_set_option", b.acc.handle, va._lib.VH_OPT_UNIT_SLOTS, 200)
        for s in sims:
            s.acc.sort()
        assert a.acc.download().tobytes() == b.acc.download().tobytes()
        for s in sims:
            s.acc.deposit_smoothed(s.dmf, w)
            s.acc.deposit_cic(s.dmc, w)
        for g in ("rho",):
            ra, rb = getattr(a.dmf, g).astype(np.float64), getattr(b.dmf, g).astype(np.float64)
            assert np.max(np.abs(ra - rb)) <= 1e-6 * np.max(np.abs(ra))
            ca, cb = a.dmc.rho.astype(np.float64), b.dmc.rho.astype(np.float64)
            assert np.max(np.abs(ca - cb)) <= 1e-6 * np.max(np.abs(ca))
        csol = (np.random.default_rng(3).standard_normal(fine.n ** 3) * 0.01).astype(np.float32)
        for s in sims:
            s.dmf.upload(va._lib.VH_GRID_CSOL, csol)
            s.dmc.upload(va._lib.VH_GRID_CSOL, np.zeros(coarse.n ** 3, np.float32))
            s.acc.compute_forces(s.dmf, s.dmc, w)
        assert a.acc.forces().tobytes() == b.acc.forces().tobytes()
        batch = va.batch.Batch([a])
        try:
            units, ms = batch.autotune(candidates=(128, 256, 512), steps=2)
        finally:
            batch.close()
        assert units[0] in (128, 256, 512)
        assert np.all(ms[0] > 0) and ms[0][(128, 256, 512).index(units[0])] == ms[0].min()
        a.step()
        assert np.all(np.isfinite(a.acc.forces()))
    finally:
        for s in sims:
            s.close()
