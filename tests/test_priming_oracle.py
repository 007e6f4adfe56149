"""The oracle's restatement of prime_leapfrog! (src/simulation.jl:233-277)
pinned by hand-computed values and a closed form (CPU only):
half_step_back q - (dt/2M) p and full_step_back -q0 + 2 hp + coef f with the
thesis's coef = dt/M or the consistent dt^2/4M (src/particles.jl:441-469);
without forces the primed previous position is free flight's q0 - dt p/M;
the split is the nearest knot, ties to even."""
import numpy as np

from conftest import load_package

va = load_package()


def test_priming_known_values(orc):
    q0 = np.array([[1.0, 2.0, 3.0]])
    p = np.array([[0.5, -1.0, 2.0]])
    f = np.array([[0.1, 0.2, -0.3]])
    M, dt = 2.0, 0.5
    hp = orc.half_step_back(q0, p, M, dt)
    assert np.array_equal(hp, [[0.9375, 2.125, 2.75]])
    assert np.allclose(orc.full_step_back(q0, hp, f, M, dt), [[0.9, 2.3, 2.425]], rtol=0, atol=1e-15)
    assert np.allclose(orc.full_step_back(q0, hp, f, M, dt, consistent=True),
                       [[0.878125, 2.25625, 2.490625]], rtol=0, atol=1e-15)
    # no force: q(-dt) of free flight
    rng = np.random.default_rng(4)
    q = rng.uniform(-50, 50, (100, 3))
    mom = rng.standard_normal((100, 3))
    back = orc.full_step_back(q, orc.half_step_back(q, mom, M, dt), np.zeros_like(q), M, dt)
    assert np.allclose(back, q - dt * mom / M, rtol=0, atol=1e-12)


def test_split_ties_to_even_and_round_trip(orc):
    # exact ties (x0, h and the points are binary fractions): 2.5 -> 2, 3.5 -> 4
    x0, h = -80.0, 4.0
    k, d = orc.split_positions(np.array([x0 + 2.5 * h, x0 + 3.5 * h, -500.0]), x0, h)
    assert list(k) == [3, 5, -104]
    assert d[0] == 0.5 * h and d[1] == -0.5 * h and d[2] == 0.0
    x0, h = -78.0, 5.571428571428571
    rng = np.random.default_rng(6)
    x = rng.uniform(-300, 300, 10000)
    k, d = orc.split_positions(x, x0, h)
    assert np.all(np.abs(d) <= h / 2 * (1 + 1e-6))
    back = orc.unsplit_positions(k, d, x0, h)
    assert np.all(np.abs(back - x) <= np.spacing(np.abs(d)) + 1e-12)


def test_prime_records_order_of_halves(orc):
    """The loop writes q(0) to the current half and q(-dt) to the previous
    one, reading q(0) from the previous half and q(-dt/2) from the current:
    swapping them (a plausible slip) gives another cloud."""
    x0, h = -78.0, 5.571428571428571
    rng = np.random.default_rng(9)
    q0 = rng.uniform(-20, 20, (50, 3))
    hp = q0 + rng.standard_normal((50, 3)) * 0.1
    rec = np.zeros(50, dtype=orc.REC)
    rec["knode"], rec["delta"] = orc.split_positions(hp, x0, h)
    rec["pknode"], rec["pdelta"] = orc.split_positions(q0, x0, h)
    f = rng.standard_normal((50, 3)).astype(np.float32)
    out = orc.prime_records(rec, f, x0, h, 2.0, 0.5)
    assert np.array_equal(out["knode"], rec["pknode"]) and np.array_equal(out["delta"], rec["pdelta"])
    want = orc.full_step_back(orc.unsplit_positions(rec["pknode"], rec["pdelta"], x0, h),
                              orc.unsplit_positions(rec["knode"], rec["delta"], x0, h),
                              f.astype(np.float64), 2.0, 0.5)
    got = orc.unsplit_positions(out["pknode"], out["pdelta"], x0, h)
    assert np.max(np.abs(got - want)) < 1e-5
