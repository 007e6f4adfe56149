"""prime_leapfrog! (src/simulation.jl:233-277) on the HIP path, against the
oracle's restatement of it.

* The device's second stage (vh_prime) on given records and forces is the
  restatement's loop to the bit, for both coefficients.
* A Simulation built in slices (sampler and upload slice by slice, the
  second stage on the device) against the restatement run on the oracle's
  FP64 forces: q(0) (the current halves, in the sort's order) bit-exact, q(-dt)
  within the forces' FP32-against-FP64 difference times the coefficient,
  far below the change the other coefficient makes."""
import numpy as np
import pytest

from conftest import load_package

pytestmark = pytest.mark.gpu
va = load_package()


def test_device_second_stage_bit_exact(config1, orc):
    p, fine, coarse, sm = config1
    n = 50_000
    kn = fine.axis.knots
    x0, h = float(kn[0]), float((kn[-1] - kn[0]) / (len(kn) - 1))
    rng = np.random.default_rng(21)
    q0 = rng.standard_normal((n, 3)) * 12.0
    hp = q0 + rng.standard_normal((n, 3)) * 0.2
    rec = np.zeros(n, dtype=va.PACKED_DTYPE)
    rec["knode"], rec["delta"] = orc.split_positions(hp, x0, h)
    rec["pknode"], rec["pdelta"] = orc.split_positions(q0, x0, h)
    f = (rng.standard_normal((n, 3)) * 1e-3).astype(np.float32)
    ctx = va.Context(0)
    acc = va.DeviceAccelerator(ctx, fine.axis, sm, n)
    try:
        for consistent in (False, True):
            acc.upload(rec)
            acc.sort()
            srec = acc.download()
            acc.upload_forces(f)
            M, dt = 0.002, 0.5
            coef = dt ** 2 / (4 * M) if consistent else dt / M
            va._lib.call("vh_prime", acc.handle, coef)
            want = orc.prime_records(srec, f, x0, h, M, dt, consistent)
            assert acc.download().tobytes() == want.tobytes()
    finally:
        acc.close()
        ctx.close()


def test_sliced_simulation_priming_against_oracle(config1, orc):
    p0, fine, coarse, sm = config1
    n = 20_000
    p = va.xenon_parameters(n, 28)
    prof = va.read_radial_profile()
    proj = va.xenon_projectile(x0=-20.0, impact=10.0)
    proj_start = va.physics.Projectile(**vars(proj))
    sim = va.Simulation(p, prof, projectile=proj, tables=(fine, coarse, sm), slice_size=4096)
    try:
        got = sim.acc.download()
        M, dt, w = sim.mass, p.dt, sim.weight
    finally:
        sim.close()
    x0, h = float(fine.axis.knots[0]), float(fine.axis.knots[1] - fine.axis.knots[0])
    pos, mom = va.sample_thomas_fermi(prof, n, w, seed=-1)
    half = orc.half_step_back(pos, mom, M, dt)
    rec = np.zeros(n, dtype=orc.REC)
    rec["knode"], rec["delta"] = orc.split_positions(half, x0, h)
    rec["pknode"], rec["pdelta"] = orc.split_positions(pos, x0, h)
    st = orc.OracleStep(fine, coarse, sm, va.Jellium.of(p.nions), w, M, dt, p.rcmax)
    r = st.update_forces(rec, proj_start)
    want = orc.prime_records(r["srec"], r["force"], x0, h, M, dt)
    assert np.array_equal(got["knode"], want["knode"]) and np.array_equal(got["delta"], want["delta"])
    qg = orc.unsplit_positions(got["pknode"], got["pdelta"], x0, h)
    qw = orc.unsplit_positions(want["pknode"], want["pdelta"], x0, h)
    coef = dt / M
    bound = coef * 3e-5 * np.max(np.abs(r["force"])) + 1e-5
    err = np.max(np.abs(qg - qw))
    assert err < bound, (err, bound)
    other = orc.prime_records(r["srec"], r["force"], x0, h, M, dt, consistent=True)
    qo = orc.unsplit_positions(other["pknode"], other["pdelta"], x0, h)
    assert np.max(np.abs(qo - qw)) > 100 * bound
