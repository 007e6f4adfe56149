"""What the trained denoiser gains over the method without it.

The network shipped in vlasov.jl_amd/data/denoiser_na196_130.npy (trained by
vlasov.jl_amd/train_denoiser.py on 1540 pairs of 24 collisions; held-out
mean squared error against the 8e7 density halved) is applied, through the
matrix-core inference that vh_step runs, to the fine density of an 8e6 run
of a collision and seed none of its pairs came from (Xe25+ at v = 0.40, b =
45 a0, seed -7), after 200 steps; the reference is the fine density of an
8e7 run of the same collision (seed -1007) at the same step.  Over the
central 96^3 points the denoised density is closer to the 8e7 density than
the deposited one, by at least 1.5x in mean squared error (the training's
held-out gain was 2.0x), and it carries the deposited total charge."""
import numpy as np
import pytest

from conftest import load_package

pytestmark = pytest.mark.gpu
va = load_package()
L = va._lib
D_ = va.denoiser


def central(a, n, c):
    o = (n - c) // 2
    return a.reshape(n, n, n)[o:o + c, o:o + c, o:o + c]


def test_trained_net_gains_on_an_unseen_collision(orc):
    assert D_.DEFAULT_WEIGHTS.exists()
    steps, crop = 200, 96
    lo = va.Simulation(va.xenon_parameters(8_000_000, 64), va.read_radial_profile(),
                       projectile=va.xenon_projectile(), seed=-7)
    hi = va.Simulation(va.xenon_parameters(80_000_000, 64), va.read_radial_profile(),
                       projectile=va.xenon_projectile(), seed=-1007)
    den = None
    scratch = None
    try:
        for _ in range(steps):
            lo.step()
            hi.step()
        n = lo.fine.n
        x, t = lo.dmf.rho.astype(np.float64), hi.dmf.rho.astype(np.float64)
        rho0 = D_.jellium_density(lo.jellium)
        den = D_.Denoiser.load(lo.ctx, D_.DEFAULT_WEIGHTS, rho0=rho0)
        scratch = va.DeviceMesh(lo.ctx, lo.fine)
        scratch.upload(L.VH_GRID_RHO, x.astype(np.float32))
        den.apply(scratch, 1.0)
        y = scratch.rho.astype(np.float64)
        assert den.fallbacks == 0
        mse_plain = np.mean((central(x, n, crop) - central(t, n, crop)) ** 2) / rho0 ** 2
        mse_net = np.mean((central(y, n, crop) - central(t, n, crop)) ** 2) / rho0 ** 2
        m0 = lo.fine.dual_moments[0]
        dq = abs(orc.total_charge(y, m0) / orc.total_charge(x, m0) - 1)
        print(f"step {steps}, central {crop}^3, in rho0^2: 8e6 deposited {mse_plain:.3e}, "
              f"denoised {mse_net:.3e} (gain {mse_plain / mse_net:.2f}); charge kept to {dq:.1e}")
        assert mse_net < mse_plain / 1.5
        assert dq < 1e-6
    finally:
        if den is not None:
            den.close()
        if scratch is not None:
            scratch.close()
        lo.close()
        hi.close()
