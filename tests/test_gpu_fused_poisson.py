"""vh_step's Poisson solve by linearity (the default form of the step): each
level's spline coefficients computed straight from its density and its
boundary as P~ Z~ (P~ = S^-1 M~ formed in FP64), with M^-1 of the density
part before the boundary is known and three products after it, phi's
interior never formed.

Bars: the step's fine and coarse coefficients (Hartree plus mean field)
against the oracle's FP64 chain on the step's own densities (poisson_nested,
spline_coefficients, effective_potential; src/devicemesh.jl:126-183,
src/poisson.jl:275-326) at 1e-5 of the magnitude of the terms summed, at
config 1's and at the benchmark's size; the forces of a step within 1e-5 of
the same step solving in the entry points' order (VH_OPT_POISSON_UNFUSED),
which is pinned to the oracle by tests/test_gpu_parity.py."""
import numpy as np
import pytest

from conftest import load_package

pytestmark = pytest.mark.gpu
va = load_package()
REL = 1e-5


def relerr(got, ref):
    return float(np.max(np.abs(np.asarray(got, np.float64) - ref)) / np.max(np.abs(ref)))


def check_coefficients(sim, orc, fine, coarse):
    """The coefficients vh_step left on both meshes against the oracle's FP64
    chain on the densities it deposited: csol = S^-1 phi + S^-1 V_eff."""
    rho_f = sim.dmf.rho.astype(np.float64)
    rho_c = sim.dmc.rho.astype(np.float64)
    wf, wc, _ = orc.poisson_nested(rho_f, rho_c, fine, coarse)
    errs = []
    for mesh, dm, phi, rho in ((fine, sim.dmf, wf, rho_f), (coarse, sim.dmc, wc, rho_c)):
        hartree = orc.spline_coefficients(phi, mesh)
        extra = orc.spline_coefficients(orc.effective_extra(rho, mesh, sim.jellium), mesh)
        # the Hartree and jellium terms cancel to a few percent of either:
        # the error is held relative to the magnitude of the terms summed
        scale = max(np.max(np.abs(hartree)), np.max(np.abs(extra)))
        errs.append(float(np.max(np.abs(dm.csol - (hartree + extra))) / scale))
    return errs


def test_fused_step_coefficients_against_oracle(config1, orc):
    p, fine, coarse, sm = config1
    n = p.nparticles
    w = p.nelectrons / n
    pos, mom = va.sample_thomas_fermi(va.read_radial_profile(), n, w, seed=-13)
    sim = va.Simulation(p, None, projectile=va.xenon_projectile(x0=-20.0, impact=10.0),
                        tables=(fine, coarse, sm), positions=pos, momenta=mom)
    try:
        for _ in range(2):
            sim.step()
            ef, ec = check_coefficients(sim, orc, fine, coarse)
            print(f"fused solve: fine coefficients {ef:.2e}, coarse {ec:.2e}")
            assert ef < REL and ec < REL
    finally:
        sim.close()


def test_fused_step_forces_match_unfused(config1):
    p, fine, coarse, sm = config1
    n = p.nparticles
    w = p.nelectrons / n
    pos, mom = va.sample_thomas_fermi(va.read_radial_profile(), n, w, seed=-17)
    sims = []
    try:
        for unfused in (False, True):
            s = va.Simulation(p, None, projectile=va.xenon_projectile(x0=-20.0, impact=10.0),
                              tables=(fine, coarse, sm), positions=pos, momenta=mom)
            s.acc.set_poisson_unfused(unfused)
            sims.append(s)
        for s in sims:
            s.step()
        a, b = (s.acc.forces().astype(np.float64) for s in sims)
        err = relerr(a, b)
        print(f"forces, fused against unfused solve: {err:.2e}")
        assert err < REL
    finally:
        for s in sims:
            s.close()
