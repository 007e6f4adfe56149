"""The denoiser's FP64 oracle pinned on the CPU: each layer against an
independent computation, the network's reverse-mode gradient against central
finite differences, and a planted mistake that the convolution's check
catches."""
import numpy as np
import pytest
from scipy import ndimage

from oracle import unet


def test_param_count_matches_the_layer_table():
    # 27 cin cout + cout per layer of 1-16-16 | 16-32-32 | 32-64-64-32 | 64-32-16 | 32-16-1
    assert unet.param_count() == 353_697
    assert len(unet.unpack(np.zeros(unet.param_count()))) == 11


def _conv_by_correlate(x, w, b):
    """Channel by channel with scipy: out[.., co] = b + sum_ci correlate(x[.., ci], w[:, ci, co])
    with the taps laid out as a 3x3x3 [dz][dy][dx] kernel."""
    D, cin = x.shape[0], x.shape[-1]
    out = np.zeros((D, D, D, w.shape[2]))
    for co in range(w.shape[2]):
        acc = np.full((D, D, D), b[co])
        for ci in range(cin):
            acc += ndimage.correlate(x[..., ci], w[:, ci, co].reshape(3, 3, 3), mode="constant",
                                     cval=0.0)
        out[..., co] = acc
    return out


def test_conv_against_scipy_correlate():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((6, 6, 6, 3))
    w = rng.standard_normal((27, 3, 4))
    b = rng.standard_normal(4)
    np.testing.assert_allclose(unet.conv3(x, w, b), _conv_by_correlate(x, w, b), rtol=0, atol=1e-12)


def test_conv_tap_order_single_offsets():
    """One nonzero tap moves the input by its offset: tap 9 dz + 3 dy + dx
    reads x[z + dz - 1, y + dy - 1, x + dx - 1]."""
    x = np.zeros((5, 5, 5, 1))
    x[2, 2, 2, 0] = 1.0
    for tap, (dz, dy, dx) in ((1, (0, 0, 1)), (5, (0, 1, 2)), (21, (2, 1, 0))):
        w = np.zeros((27, 1, 1))
        w[tap] = 1.0
        out = unet.conv3(x, w)[..., 0]
        want = np.zeros((5, 5, 5))
        want[2 - (dz - 1), 2 - (dy - 1), 2 - (dx - 1)] = 1.0
        assert np.array_equal(out, want)


def test_planted_tap_reversal_fails_the_correlate_check():
    """A convolution that reads its taps reversed (a true convolution
    instead of a correlation) fails the scipy check."""
    rng = np.random.default_rng(2)
    x = rng.standard_normal((5, 5, 5, 2))
    w = rng.standard_normal((27, 2, 3))
    b = np.zeros(3)
    planted = unet.conv3(x, w[::-1].copy(), b)
    assert np.max(np.abs(planted - _conv_by_correlate(x, w, b))) > 1e-2


def test_pool_and_up_closed_forms_and_adjoint():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((4, 4, 4, 2))
    p = unet.pool(x)
    assert p.shape == (2, 2, 2, 2)
    assert np.isclose(p[1, 0, 1, 1], x[2:4, 0:2, 2:4, 1].mean())
    u = unet.up(p)
    assert u.shape == x.shape and np.all(u[2:4, 0:2, 2:4, 1] == p[1, 0, 1, 1])
    # <up(a), b> = <a, up_back(b)> and up_back = 8 pool
    a = rng.standard_normal((2, 2, 2, 2))
    assert np.isclose(np.sum(unet.up(a) * x), np.sum(a * unet.up_back(x)))
    np.testing.assert_allclose(unet.up_back(x), 8 * unet.pool(x), atol=1e-14)
    # pool_back is pool's adjoint
    assert np.isclose(np.sum(unet.pool(x) * a), np.sum(x * unet.pool_back(a)))


def test_untrained_identity_and_shapes():
    """A zero last layer gives o = 0 whatever the rest."""
    p = unet.random_params(4)
    sizes = unet.layer_sizes()
    p[-sizes[-1]:] = 0.0
    x = np.random.default_rng(5).standard_normal((8, 8, 8))
    o, c = unet.forward(p, x)
    assert np.array_equal(o, np.zeros_like(x))
    assert c["a4"].shape == (2, 2, 2, 64) and c["cat1"].shape == (8, 8, 8, 32)


@pytest.mark.parametrize("D", [4, 8])
def test_gradient_against_central_differences(D):
    """Every layer's reverse-mode gradient against FP64 central differences
    on a handful of its weights and its biases (h = 1e-6, relative 1e-5 of
    the layer's largest gradient)."""
    p = unet.random_params(10 + D)
    rng = np.random.default_rng(20 + D)
    x = rng.standard_normal((D, D, D))
    t = x + 0.3 * rng.standard_normal((D, D, D))
    l, g = unet.gradient(p, x, t)
    assert np.isclose(l, unet.loss(p, x, t), rtol=1e-14)
    off = 0
    for li, size in enumerate(unet.layer_sizes()):
        cout = unet.LAYERS[li][1]
        idx = list(off + rng.choice(size - cout, size=6, replace=False)) + \
            list(off + size - cout + np.arange(min(cout, 3)))
        fd = unet.fd_gradient(p, x, t, idx)
        scale = np.max(np.abs(g[off:off + size]))
        assert scale > 0, f"layer {li} has no gradient"
        err = np.max(np.abs(fd - g[idx])) / scale
        assert err < 1e-5, f"layer {li}: {err:.2e}"
        off += size


def test_planted_skip_swap_fails_the_difference_check():
    """A backward that takes the level-1 skip's gradient from the wrong half
    of the concatenation fails the finite differences of layer 1."""
    p = unet.random_params(30)
    rng = np.random.default_rng(31)
    x = rng.standard_normal((4, 4, 4))
    t = rng.standard_normal((4, 4, 4))
    off = unet.layer_sizes()[0]
    idx = list(off + rng.choice(unet.layer_sizes()[1] - 16, size=8, replace=False))
    fd = unet.fd_gradient(p, x, t, idx)
    errs = {}
    for plant in (None, "SKIP_SWAP"):
        _, g = unet.gradient(p, x, t, plant=plant)
        errs[plant] = np.max(np.abs(fd - g[idx])) / np.max(np.abs(g[off:off + unet.layer_sizes()[1]]))
    assert errs[None] < 1e-5 and errs["SKIP_SWAP"] > 1e-2


def test_support_mask_removes_the_correction_outside():
    """Where |x| <= support the loss does not see the net: its gradient is
    the finite differences' with half the input zeroed, and a network with
    only a last-layer bias changes the loss only inside the support."""
    p = unet.random_params(40)
    rng = np.random.default_rng(41)
    x = rng.standard_normal((4, 4, 4))
    x[:, :2] = 0.0
    t = rng.standard_normal((4, 4, 4))
    l, g = unet.gradient(p, x, t)
    idx = list(rng.choice(unet.param_count(), size=12, replace=False))
    fd = unet.fd_gradient(p, x, t, idx)
    assert np.max(np.abs(fd - g[idx])) < 1e-5 * np.max(np.abs(g))
    q = np.zeros(unet.param_count())
    q[-1] = 0.5                           # o = 0.5 everywhere
    inside = unet.support_mask(x)
    want = np.mean((x + 0.5 * inside - t) ** 2)
    assert np.isclose(unet.loss(q, x, t), want, rtol=1e-14)
    assert inside.sum() == 32
