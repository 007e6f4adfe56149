"""The learned denoiser's kernels against the FP64 oracle (oracle/unet.py):
forward, loss and the gradient of every parameter on small cubes, the
gradient again against FP64 central differences, the Adam training loop on
a synthetic denoising task, the residual charge check of the inference (both
outcomes) and the step with the denoiser attached (recorded graph against
eager launches).

Bars, relative (max |diff| / max |ref|), FP32 accumulation over at most
27 x 64 products a voxel and D^3 voxels a weight: correction o and loss 1e-5,
each layer's gradient 1e-4 against the oracle's reverse mode and against
central differences (h = 1e-6 on the FP64 loss of the bf16 weights)."""
import numpy as np
import pytest

from conftest import load_package
from oracle import unet

pytestmark = pytest.mark.gpu
va = load_package()
L = va._lib
D_ = va.denoiser


def relerr(got, ref):
    return float(np.max(np.abs(np.asarray(got, np.float64) - ref)) / np.max(np.abs(ref)))


def small_correction(seed, last=0.1):
    """He-scale weights with the last layer scaled by `last`: a correction of
    a tenth of the input's scale or so."""
    p = unet.random_params(seed)
    p[-unet.layer_sizes()[-1]:] *= last
    return p.astype(np.float32)


@pytest.fixture(scope="module")
def ctx():
    c = va.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("D", [8, 12])
def test_forward_and_gradient_against_oracle(ctx, D):
    p = unet.random_params(40 + D).astype(np.float32)
    den = D_.Denoiser(ctx, p)
    try:
        weff = den.weights(effective=True).astype(np.float64)
        assert np.array_equal(weff, unet.bf16(p))
        assert np.array_equal(den.weights(), p)
        rng = np.random.default_rng(50 + D)
        x = rng.standard_normal((D, D, D)).astype(np.float32)
        t = (x + 0.3 * rng.standard_normal((D, D, D))).astype(np.float32)
        o, loss, g = den.eval(x, t)
        o64, _ = unet.forward(weff, x.astype(np.float64))
        l64, g64 = unet.gradient(weff, x.astype(np.float64), t.astype(np.float64))
        e_o, e_l = relerr(o, o64), abs(loss - l64) / l64
        errs, off = [], 0
        for size in unet.layer_sizes():
            errs.append(relerr(g[off:off + size], g64[off:off + size]))
            off += size
        print(f"D={D}: o {e_o:.1e}, loss {e_l:.1e}, gradient per layer "
              + " ".join(f"{e:.1e}" for e in errs))
        assert e_o < 1e-5 and e_l < 1e-5
        assert max(errs) < 1e-4
        # central differences of the FP64 loss on three weights a layer
        idx, off = [], 0
        for size in unet.layer_sizes():
            idx += list(off + rng.choice(size, size=3, replace=False))
            off += size
        fd = unet.fd_gradient(weff, x.astype(np.float64), t.astype(np.float64), idx, h=1e-6)
        off, k = 0, 0
        for size in unet.layer_sizes():
            scale = np.max(np.abs(g64[off:off + size]))
            assert np.max(np.abs(g[idx[k:k + 3]] - fd[k:k + 3])) < 1e-4 * scale
            off += size
            k += 3
    finally:
        den.close()


def test_untrained_net_is_the_identity(ctx):
    den = D_.Denoiser(ctx, seed=3)
    try:
        x = np.random.default_rng(1).standard_normal((8, 8, 8)).astype(np.float32)
        o, loss, _ = den.eval(x, x, grad=False)
        assert np.array_equal(o, np.zeros_like(o)) and loss == 0.0
    finally:
        den.close()


def _blobs(mesh, rng, nblob=4):
    gv = mesh.axis.colloc
    kk, jj, ii = np.meshgrid(gv, gv, gv, indexing="ij")
    r = np.zeros(ii.shape)
    for _ in range(nblob):
        c = rng.uniform(-0.3, 0.3, 3) * (gv[-1] - gv[0])
        r += np.exp(-((ii - c[0]) ** 2 + (jj - c[1]) ** 2 + (kk - c[2]) ** 2) / (2 * 12.0 ** 2))
    return r


def test_training_learns_a_denoiser(ctx):
    """Adam on pairs (smooth blobs + white noise, smooth blobs) made on a
    30^3 grid: the loss falls and the trained net halves the held-out pair's
    error against the clean density."""
    mesh = va.grids.SplineMesh.of(va.grids.uniform_axis(-40.0, 40.0, 14))
    n = mesh.n
    rng = np.random.default_rng(7)
    lo, hi = va.DeviceMesh(ctx, mesh), va.DeviceMesh(ctx, mesh)
    den = D_.Denoiser(ctx, seed=11, rho0=1.0)
    try:
        den.dataset_reserve(n, 9)
        for _ in range(9):
            clean = _blobs(mesh, rng)
            noisy = clean + 0.15 * rng.standard_normal(clean.shape)
            lo.upload(L.VH_GRID_RHO, noisy.astype(np.float32).ravel())
            hi.upload(L.VH_GRID_RHO, clean.astype(np.float32).ravel())
            den.dataset_add(lo, hi)
        assert den.dataset_size == 9
        before = den.dataset_error(8, 24)
        losses = den.train(300, crop=16, batch=2, first=0, last=8, lr=2e-3, seed=5)
        after = den.dataset_error(8, 24)
        print(f"loss {losses[:10].mean():.4f} -> {losses[-10:].mean():.4f}; held-out pair "
              f"plain {after[0]:.4f}, net before {before[1]:.4f}, after {after[1]:.4f}")
        assert before[0] == after[0] and abs(before[1] - before[0]) < 1e-9
        assert losses[-10:].mean() < 0.6 * losses[:10].mean()
        assert after[1] < 0.5 * after[0]
    finally:
        den.close()
        lo.close()
        hi.close()


def test_apply_and_residual_charge_check(ctx, orc):
    """On a 30^3 density (padded to 32^3), the FP32-activation inference:
    within the tolerance the density becomes rho0 (x + o) scaled to the
    deposited charge (against the oracle's forward, 1e-5); at tolerance 0 the
    density stays as deposited, bit for bit, and the fallback is counted."""
    mesh = va.grids.SplineMesh.of(va.grids.uniform_axis(-40.0, 40.0, 14))
    n = mesh.n
    P = (n + 3) // 4 * 4
    rng = np.random.default_rng(9)
    rho = (_blobs(mesh, rng) * 0.004 + 1e-4 * rng.standard_normal((n, n, n))).astype(np.float32)
    p = small_correction(60)
    rho0 = 0.004
    dm = va.DeviceMesh(ctx, mesh)
    den = D_.Denoiser(ctx, p, rho0=rho0)
    den.set_inference("fp32")
    try:
        weff = den.weights(effective=True).astype(np.float64)
        dm.upload(L.VH_GRID_RHO, rho.ravel())
        den.apply(dm, 10.0)
        got = dm.rho.astype(np.float64)
        x = np.zeros((P, P, P))
        x[:n, :n, :n] = rho.astype(np.float64) / rho0
        o, _ = unet.forward(weff, x)
        m = unet.support_mask(rho * np.float32(1.0 / rho0))
        net = (rho0 * (x[:n, :n, :n] + m * o[:n, :n, :n])).ravel()
        q = orc.total_charge(rho.astype(np.float64).ravel(), mesh.dual_moments[0])
        qn = orc.total_charge(net, mesh.dual_moments[0])
        want = net * (q / qn)
        print(f"net charge ratio {qn / q:.4f}, error {relerr(got, want):.1e}")
        assert relerr(got, want) < 1e-5
        assert abs(orc.total_charge(got, mesh.dual_moments[0]) / q - 1) < 1e-5
        assert den.fallbacks == 0
        dm.upload(L.VH_GRID_RHO, rho.ravel())
        den.apply(dm, 0.0)
        assert dm.rho.tobytes() == rho.ravel().tobytes()
        assert den.fallbacks == 1
    finally:
        den.close()
        dm.close()


def _path_params(entries):
    """A network that routes single channels: w[layer][tap][ci][co] = value
    for each entry, every other weight and bias zero."""
    p = np.zeros(unet.param_count())
    offs = np.concatenate([[0], np.cumsum(unet.layer_sizes())])
    for layer, tap, ci, co, v in entries:
        cin, cout, _ = unet.LAYERS[layer]
        p[offs[layer] + (tap * cin + ci) * cout + co] = v
    return p.astype(np.float32)


C_ = 13   # the centre tap
PATHS = {
    # x -> every layer's channel 0 through the centre taps, the up branches
    "up branches": [(l, C_, 0, 0, 1.0) for l in range(11)],
    # the level-1 skip: o = a1[0] = relu(x)
    "level-1 skip": [(0, C_, 0, 0, 1.0), (1, C_, 0, 0, 1.0), (9, C_, 16, 0, 1.0), (10, C_, 0, 0, 1.0)],
    # the level-2 skip through L7's second half, then up
    "level-2 skip": [(0, C_, 0, 0, 1.0), (1, C_, 0, 0, 1.0), (2, C_, 0, 0, 1.0), (3, C_, 0, 0, 1.0),
                     (7, C_, 32, 0, 1.0), (8, C_, 0, 0, 1.0), (9, C_, 0, 0, 1.0), (10, C_, 0, 0, 1.0)],
    # channels in every k group and both halves of 64, off-centre taps
    "channels and taps": [(0, 5, 0, 7, 1.0), (1, 21, 7, 9, 1.0), (2, C_, 9, 20, 1.0),
                          (3, 1, 20, 31, 1.0), (4, C_, 31, 40, 1.0), (5, 25, 40, 63, 1.0),
                          (6, C_, 63, 5, 1.0), (7, 3, 5, 9, 1.0), (8, 17, 9, 15, 1.0),
                          (9, 9, 15, 3, 1.0), (10, 26, 3, 0, 0.5)],
}


@pytest.mark.parametrize("path", sorted(PATHS))
def test_matrix_core_plumbing_exact(ctx, orc, path):
    """Single-channel routes through the network on a 0/1 input, where every
    value is exact in bf16 (pooled means of 0/1 values), the correction kept
    where the input is 1 (the support): the matrix-core inference equals the
    oracle's to the charge renormalisation's FP32
    rounding (1e-6), so a wrong tap, channel group, concatenation half,
    pooling or upsampling index shows as an O(1) difference."""
    mesh = va.grids.SplineMesh.of(va.grids.uniform_axis(-40.0, 40.0, 14))
    n = mesh.n
    P = (n + 3) // 4 * 4
    rng = np.random.default_rng(123)
    rho0 = 0.25
    rho = (rho0 * (rng.random((n, n, n)) < 0.5)).astype(np.float32)
    p = _path_params(PATHS[path])
    dm = va.DeviceMesh(ctx, mesh)
    den = D_.Denoiser(ctx, p, rho0=rho0)
    try:
        x = np.zeros((P, P, P))
        x[:n, :n, :n] = rho / rho0
        o, _ = unet.forward(p.astype(np.float64), x, act_bf16=True)
        assert np.array_equal(unet.bf16(o), o), "the route leaves bf16"
        m = unet.support_mask(rho * np.float32(1.0 / rho0))
        net = (rho.astype(np.float64) + rho0 * m * o[:n, :n, :n]).ravel()
        q = orc.total_charge(rho.astype(np.float64).ravel(), mesh.dual_moments[0])
        qn = orc.total_charge(net, mesh.dual_moments[0])
        want = net * (q / qn)
        dm.upload(L.VH_GRID_RHO, rho.ravel())
        den.apply(dm, 100.0)
        got = dm.rho.astype(np.float64)
        err = np.abs(got - want) / np.max(np.abs(want))
        bad = np.argwhere(err.reshape(n, n, n) > 1e-6)
        print(f"{path}: max {err.max():.1e}, {len(bad)} voxels off, first {bad[:4].tolist()}")
        assert err.max() < 1e-6
    finally:
        den.close()
        dm.close()


@pytest.mark.parametrize("nfine", [14, 28])
def test_matrix_core_inference_against_oracle(ctx, orc, nfine):
    """The matrix-core inference (bf16 activations, FP32 accumulation) of a
    random network on a 30^3 and a 58^3 density, against the oracle's forward
    with the same roundings (input, ReLU outputs and pooled tensors to bf16)
    and against the FP32-activation path, both renormalised: within 5e-2 of
    the largest correction max |rho0 o|.  The index mapping is pinned exactly
    by the routing tests above; what remains is rounding: an activation that
    lies within the FP32 accumulation's error of a bf16 rounding boundary
    rounds the other way, and the oracle's own output moves by 1.7e-2 of max
    |o| (9e-3 at the 99.9th percentile) when every value it rounds is
    perturbed by 1e-7 relative (measured on the CPU for the n = 30 case)."""
    mesh = va.grids.SplineMesh.of(va.grids.uniform_axis(-40.0, 40.0, nfine))
    n = mesh.n
    P = (n + 3) // 4 * 4
    rng = np.random.default_rng(90 + nfine)
    rho = (_blobs(mesh, rng) * 0.004 + 1e-4 * rng.standard_normal((n, n, n))).astype(np.float32)
    p = small_correction(61)
    rho0 = 0.004
    dm = va.DeviceMesh(ctx, mesh)
    den = D_.Denoiser(ctx, p, rho0=rho0)
    try:
        weff = den.weights(effective=True).astype(np.float64)
        x32 = np.zeros((P, P, P), np.float32)
        x32[:n, :n, :n] = rho * np.float32(1.0 / rho0)
        o, _ = unet.forward(weff, x32.astype(np.float64), act_bf16=True)
        corr = rho0 * np.abs(o).max()
        m = unet.support_mask(x32[:n, :n, :n])
        net = (rho.astype(np.float64) + rho0 * m * o[:n, :n, :n]).ravel()
        q = orc.total_charge(rho.astype(np.float64).ravel(), mesh.dual_moments[0])
        qn = orc.total_charge(net, mesh.dual_moments[0])
        want = net * (q / qn)
        dm.upload(L.VH_GRID_RHO, rho.ravel())
        den.apply(dm, 10.0)
        got = dm.rho.astype(np.float64)
        e = np.abs(got - want) / corr
        den.set_inference("fp32")
        dm.upload(L.VH_GRID_RHO, rho.ravel())
        den.apply(dm, 10.0)
        f32 = dm.rho.astype(np.float64)
        e32 = np.max(np.abs(got - f32)) / corr
        print(f"n={n}: against the bf16 oracle {e.max():.1e} of max |rho0 o| "
              f"(99.9% {np.quantile(e, 0.999):.1e}); against the FP32-activation path {e32:.1e}")
        assert e.max() < 5e-2 and e32 < 5e-2
        assert den.fallbacks == 0
    finally:
        den.close()
        dm.close()


def test_step_with_denoiser_graph_matches_eager(config1):
    """vh_step with the denoiser between the deposit and the solve, recorded
    into graphs and replayed, against eager launches: the cloud to the bit
    over 5 steps (an energy step among them, where the Python driver applies
    the net itself); the net's correction is nonzero and its fallbacks zero."""
    p, fine, coarse, sm = config1
    n = p.nparticles
    w = p.nelectrons / n
    pos, mom = va.sample_thomas_fermi(va.read_radial_profile(), n, w, seed=-9)
    params = small_correction(70, last=0.02)
    sims, dens = [], []
    try:
        for graph in (True, False):
            s = va.Simulation(p, None, projectile=va.xenon_projectile(x0=-20.0, impact=10.0),
                              tables=(fine, coarse, sm), positions=pos, momenta=mom)
            s.acc.set_step_graph(graph)
            d = D_.Denoiser(s.ctx, params, rho0=D_.jellium_density(s.jellium))
            s.attach_denoiser(d, tol=0.5)
            sims.append(s)
            dens.append(d)
        plain = va.Simulation(p, None, projectile=va.xenon_projectile(x0=-20.0, impact=10.0),
                              tables=(fine, coarse, sm), positions=pos, momenta=mom)
        sims.append(plain)
        for i in range(5):
            for s in sims:
                s.step(energy=(i == 2))
            a, b = (s.acc.download() for s in sims[:2])
            assert a.tobytes() == b.tobytes(), f"step {i}"
        assert sims[0].dmf.rho.tobytes() == sims[1].dmf.rho.tobytes()
        assert not np.array_equal(sims[0].dmf.rho, plain.dmf.rho)
        assert dens[0].fallbacks == 0 and dens[1].fallbacks == 0
    finally:
        for s in sims:
            s.close()       # detaches the denoiser, then frees the context
        for d in dens:
            d.close()       # needs no context
