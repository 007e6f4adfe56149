"""The reference's resident driver, replayed call for call through the C ABI,
against vh_step.

`replay_step` makes, in order, the calls `step!(sim; energy)` makes on the
resident path (src/simulation.jl:382-396): `_update_forces_resident!`
(src/simulation.jl:472-531) -- deposit_smoothed!, deposit_cic!, poisson!,
spline_coefficients! on both levels, the Hartree energy on energy steps,
effective_potential! on both levels, the coarse coefficients brought back,
forces! with the projectile, advance_projectile! on the host, the copies of
the total potential into phi -- then `_advance_cloud!`, the device step!
(src/accelerator.jl:580-605), and on energy steps the total interaction.
Each call is the ABI entry point the Julia extension binds to that method.

Against it, the same cloud stepped by vh_step (one call per step, the
projectile integrated on the device, the fine coefficients as one transform,
the solve in the driver's order: VH_OPT_POISSON_UNFUSED):
the clouds and forces identical to the bit after six steps with the budget
taken on two of them, the projectile's state to FP64 rounding (the device's
pow in the jellium term)."""
import numpy as np
import pytest

from conftest import load_package

pytestmark = pytest.mark.gpu
va = load_package()
L = va._lib


def replay_step(sim, energy):
    acc, dmf, dmc, w, jel = sim.acc, sim.dmf, sim.dmc, sim.weight, sim.jellium
    p, proj = sim.params, sim._projectile
    # deposit_smoothed!(dev.rho[1], accelerator, fine, sm, positions; charge = w)
    acc.deposit_smoothed(dmf, w)
    # deposit_cic!(dev.rho[2], dmc, accelerator, N, w, knots[1], hf)
    acc.deposit_cic(dmc, w)
    # poisson!(dev.phi, dev.rho, dev.meshes)
    va.poisson_nested(dmf, dmc)
    # csolf = spline_coefficients!(dev.csol[1], dev.phi[1], dmf), then coarse
    dmf.spline_coefficients()
    dmc.spline_coefficients()
    hartree = acc.interaction_energy(dmf, dmc, w) / 2 if energy else None
    # effective_potential!(csolf, dev.rho[1], dmf, jel), then coarse
    dmf.effective_potential(jel)
    dmc.effective_potential(jel)
    # copyto!(dev.csolc_host, csolc)
    csolc_host = dmc.csol
    # forces!(cloud, accelerator, fine.axes, csolf, coarse.axes, csolc_host, sm;
    #         projectile, packed = true)
    outcount, red = acc.compute_forces(dmf, dmc, w, proj, "device")
    # advance_projectile!(sim, accelerator): projectile_forces!(cloud, acc, proj, jel)
    # then step!(proj, force, dt), on the host
    fjel, _ = va.physics.jellium_force(proj, jel)
    proj.step(tuple(a + b for a, b in zip(fjel, red[:3])), p.dt)
    # copyto!(dev.phi[1], csolf); copyto!(dev.phi[2], csolc)
    dmf.copy_grid(L.VH_GRID_PHI, L.VH_GRID_CSOL)
    dmc.copy_grid(L.VH_GRID_PHI, L.VH_GRID_CSOL)
    # _advance_cloud!(sim) -> step!(cloud, dt, acc; rcmax)
    diag = acc.verlet(p.dt, sim.mass, p.rcmax, diagnostics=energy)
    total = acc.interaction_energy(dmf, dmc, w) if energy else None
    return dict(outcount=outcount, csolc_host=csolc_host, hartree=hartree, diag=diag, total=total)


def test_driver_call_order_matches_vh_step(config1):
    p, fine, coarse, sm = config1
    n = p.nparticles
    w = p.nelectrons / n
    pos, mom = va.sample_thomas_fermi(va.read_radial_profile(), n, w, seed=-11)
    sims = []
    try:
        for resident in (True, False):
            proj = va.xenon_projectile(x0=-20.0, impact=10.0)
            sims.append(va.Simulation(p, None, projectile=proj, tables=(fine, coarse, sm),
                                      positions=pos, momenta=mom, resident_projectile=resident))
        native, replay = sims
        native.acc.set_poisson_unfused(True)  # vh_step solving in the driver's order
        for i in range(6):
            energy = i in (0, 3)
            native.step()
            r = replay_step(replay, energy)
            if energy:
                assert np.isfinite(r["hartree"]) and np.isfinite(r["total"])
                assert r["diag"]["kinetic"] > 0
            # the copies of the total potential into phi
            assert replay.dmf.phi.tobytes() == replay.dmf.csol.tobytes()
            assert r["csolc_host"].tobytes() == replay.dmc.csol.tobytes()
        assert native.acc.download().tobytes() == replay.acc.download().tobytes()
        assert native.acc.forces().tobytes() == replay.acc.forces().tobytes()
        a, b = native.projectile, replay._projectile
        assert np.allclose(a.position, b.position, rtol=1e-13, atol=1e-12)
        assert np.allclose(a.velocity, b.velocity, rtol=1e-10, atol=1e-14)
    finally:
        for s in sims:
            s.close()
