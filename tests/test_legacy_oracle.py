"""The oracle pieces of the legacy route (a cloud held on the host as FP64
positions, src/simulation.jl:318-345), pinned by closed forms (CPU only):

* `pack_kd` (src/kernels.jl:134-145): on a knot the offset is 0 (to the
  FP64 rounding of x0 + (k - 1) h against the knot); half-way
  between knots k and k+1 the knot is the even-rounded one (Julia's `round`);
  beyond the grid the knot is clamped to 1 or nk and the offset carries the
  rest; everywhere x0 + (k - 1) h + delta gives back p to the offset's
  Float32 rounding.
* `fallback_points` (src/accelerator.jl:705-730): the coarse spline field of
  Phi = x at the particles' own FP64 positions is exactly (-w, 0, 0); beyond
  both grids it is the monopole -w^2 escaped p / |p|^3; the projectile's ball
  law adds w q (p_proj - p) / max(r, R)^3."""
import numpy as np

from conftest import load_package

va = load_package()
G = va.grids


def test_pack_kd_known_answers(orc):
    ax = G.uniform_axis(-78.0, 78.0, 28)
    kn = ax.knots
    x0, h, nk = float(kn[0]), float(kn[1] - kn[0]), len(kn)
    pts = np.array([[kn[4], kn[10], kn[-1]],                      # on knots
                    [x0 + 2.5 * h, x0 + 3.5 * h, x0 + 10.5 * h],   # half-way: even k - 1
                    [-200.0, 500.0, kn[0]]])                       # beyond the grid
    rec = orc.pack_kd(pts, x0, h, nk)
    assert np.array_equal(rec["knode"][0], [5, 11, nk])
    # the knots are correctly rounded values, x0 + (k - 1) h is not: the
    # offset on a knot is that FP64 rounding difference
    assert np.all(np.abs(rec["delta"][0]) < 1e-12)
    assert np.array_equal(rec["knode"][1], [3, 5, 11])             # rint(2.5)=2, 3.5->4, 10.5->10
    assert np.allclose(rec["delta"][1], [0.5 * h, -0.5 * h, 0.5 * h], rtol=1e-6)
    assert np.array_equal(rec["knode"][2], [1, nk, 1])
    rng = np.random.default_rng(3)
    p = rng.uniform(-300, 300, (5000, 3))
    r = orc.pack_kd(p, x0, h, nk)
    back = x0 + (r["knode"].astype(np.float64) - 1) * h + r["delta"].astype(np.float64)
    inside = (p > x0 - h / 2) & (p < kn[-1] + h / 2)
    assert np.all(np.abs(r["delta"][inside]) <= h / 2 * (1 + 1e-6))
    assert np.all(np.abs(back - p) <= np.spacing(np.abs(r["delta"]).astype(np.float32)) + 1e-12)
    assert np.array_equal(r["pknode"], r["knode"]) and np.array_equal(r["pdelta"], r["delta"])


def test_fallback_points_closed_forms(orc, config1):
    p, fine, coarse, sm = config1
    ax = coarse.axis
    one = np.zeros(ax.nbasis)
    one[0::2] = 1.0
    lin = np.zeros(ax.nbasis)
    lin[0::2] = ax.knots
    lin[1::2] = 1.0
    csol = (one[:, None, None] * one[None, :, None] * lin[None, None, :]).ravel()
    rng = np.random.default_rng(8)
    inside = rng.uniform(-220, 220, (500, 3))
    beyond = rng.uniform(300, 500, (50, 3))
    w, esc = 0.3, 17.0
    f = orc.fallback_points(inside, ax.knots, csol, w, escaped=esc)
    assert np.max(np.abs(f - np.array([-w, 0.0, 0.0]))) < 1e-12
    fb = orc.fallback_points(beyond, ax.knots, csol, w, escaped=esc)
    r3 = np.linalg.norm(beyond, axis=1) ** 3
    assert np.allclose(fb, -(w * w) * esc * beyond / r3[:, None], rtol=1e-14, atol=0)
    proj = np.array([1.0, 2.0, 3.0, 25.0])
    fp = orc.fallback_points(inside, ax.knots, csol, w, proj=proj, ball=True, radius=5.0,
                             escaped=esc)
    d = proj[:3] - inside
    r = np.linalg.norm(d, axis=1)
    k = np.where(r > 5.0, 1 / r ** 3, 1 / 125.0)
    assert np.allclose(fp - f, (w * 25.0 * k)[:, None] * d, rtol=1e-12, atol=1e-15)
