"""How a step orders its two streams.

A step forks its coarse chain onto the context's side stream and joins it
back.  Alone in the process, a context orders the two streams by device flags
(a spinning wait kernel); with other contexts alive, or when its two streams
share a hardware queue, by events.

* Twenty contexts alive at once (forty streams over the runtime's few
  hardware queues) each step and synchronise in turn without a wait timing
  out (vh_sync reports timed-out waits as an error).
* A collision stepped by a context alone in the process (flags) and the same
  collision stepped while another context is alive (events) end with the same
  cloud and projectile, to the bit."""
import numpy as np
import pytest

from conftest import load_package

pytestmark = pytest.mark.gpu
va = load_package()


def _sims(config1, count, seed=5):
    p, fine, coarse, sm = config1
    fam = va.batch.collision_family(count=count, nmin=2e4, nmax=4e4, nfine=28, seed=seed)
    return [va.Simulation(c.params, va.read_radial_profile(),
                          projectile=va.physics.Projectile(**vars(c.projectile)),
                          tables=(fine, coarse, sm))
            for c in fam]


def test_many_contexts_step_without_timeouts(config1):
    sims = _sims(config1, 20)
    try:
        for _ in range(2):
            for s in sims:
                s.step()
                s.ctx.sync()
    finally:
        for s in sims:
            s.close()


def _run(config1, other_alive):
    (s,) = _sims(config1, 1, seed=9)
    other = va.Context(0) if other_alive else None
    try:
        for i in range(6):
            s.step(energy=(i == 2))
        s.ctx.sync()
        return (s.acc.download().tobytes(), np.array(s.projectile.position, dtype=np.float64),
                np.array(s.projectile.velocity, dtype=np.float64))
    finally:
        s.close()
        if other is not None:
            other.close()


def test_flags_and_events_agree(config1):
    alone = _run(config1, False)
    shared = _run(config1, True)
    assert alone[0] == shared[0]
    assert np.array_equal(alone[1], shared[1])
    assert np.array_equal(alone[2], shared[2])
