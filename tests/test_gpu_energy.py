"""Energy conservation of the isolated Na196 cluster on the HIP path, by the
reference's own measure (test/runtests.jl:1101-1148): the range of the total
energy over the scale of the terms that make it up, |hartree| + |meanfield| +
ions of the first budget, stays below 1e-3, and the total oscillates (it
neither only rises nor only falls: a symplectic scheme)."""
import numpy as np
import pytest

from conftest import load_package

pytestmark = pytest.mark.gpu
va = load_package()


def conservation(hist):
    tot = np.array([b.total for b in hist])
    b1 = hist[0]
    scale = abs(b1.hartree) + abs(b1.meanfield) + b1.ions
    d = np.diff(tot)
    return (tot.max() - tot.min()) / scale, bool(not np.all(d < 0) and not np.all(d > 0))


def test_reference_isolated_cluster_case():
    """The reference's case: SimulationParameters(nparticles = 2000, nsteps = 6,
    dt = 1.0) (nfine 28, rcluster 50, rbox 150), budget every step; and the
    budget observes without feeding back: 12 steps with it every step or
    every 4 steps leave the same cloud (test/runtests.jl:1109-1133)."""
    p = va.SimulationParameters(nparticles=2000, nsteps=6, dt=1.0)
    prof = va.read_radial_profile()
    sim = va.Simulation(p, prof)
    try:
        hist = sim.run(6, energy_every=1)
        assert len(hist) == 6 and all(np.isfinite(b.total) for b in hist)
        rel, osc = conservation(hist)
        assert rel < 1e-3, rel
        assert osc
        # the deposited charge stays that of the electrons (FP32 density
        # renormalised in FP64, summed back in FP64)
        assert abs(sim.dmf.total_charge() - 196.0) < 1e-5 * 196.0
    finally:
        sim.close()
    a = va.Simulation(p, prof)
    b = va.Simulation(p, prof)
    # b's steps without the budget are native calls: solving in the entry
    # points' order, as a's budget steps do
    b.acc.set_poisson_unfused(True)
    try:
        a.run(12, energy_every=1)
        h = b.run(12, energy_every=4)
        assert len(h) == 3
        ra, rb = a.acc.download(), b.acc.download()
        assert ra.tobytes() == rb.tobytes()
        assert b.step(energy=False) is None
        assert isinstance(b.step(energy=True), va.physics.EnergyBudget)
        with pytest.raises(ValueError):
            b.run(1, energy_every=-1)
    finally:
        a.close()
        b.close()


def test_isolated_cluster_config1_700_steps():
    """Config 1's cluster (1e5 particles, nfine 28, rcluster 78, rbox 235,
    dt 0.5) without projectile over the 700 steps of a crossing, the budget
    every 10 steps as the Fortran took it."""
    p = va.xenon_parameters(100_000, 28)
    sim = va.Simulation(p, va.read_radial_profile())
    try:
        hist = sim.run(700, energy_every=10)
    finally:
        sim.close()
    assert len(hist) == 70
    rel, osc = conservation(hist)
    print(f"isolated cluster, 1e5 particles, 700 steps: range/scale = {rel:.3e}, oscillates = {osc}")
    assert rel < 1e-3, rel
    assert osc
