"""FP64 host construction, pinned by closed forms and invariants the
reference's own suite checks (test/runtests.jl:53-281, 468-501, 722-743,
1480-1512)."""
import math

import numpy as np
import pytest

from conftest import load_package

va = load_package()
G = va.grids


def make_testaxis(n=8, L=1.0):
    """The reference's non-uniform test grid (test/runtests.jl:9-21)."""
    t = np.linspace(0, 1, n + 1)
    knots = L * (t + 0.15 * np.sin(2 * np.pi * t))
    colloc = np.empty(2 * len(knots))
    colloc[0], colloc[-1] = knots[0], knots[-1]
    for j in range(n):
        a, b = knots[j], knots[j + 1]
        colloc[2 * j + 1] = a + (b - a) / 3
        colloc[2 * j + 2] = a + 2 * (b - a) / 3
    return G.SplineAxis(knots, colloc)


def test_hermite_basis_at_knots():
    ax = make_testaxis()
    g = ax.knots
    for k in range(1, ax.nknots - 1):
        v, dv = G.evaluate(g, 2 * k, g[k], 1)
        s, ds = G.evaluate(g, 2 * k + 1, g[k], 1)
        assert abs(v - 1) < 1e-14 and abs(s) < 1e-14
        assert abs(dv) < 1e-12 and abs(ds - 1) < 1e-12
        assert abs(G.value(g, 2 * k, g[k - 1])) < 1e-14 and abs(G.value(g, 2 * k, g[k + 1])) < 1e-14
    assert G.value(g, 6, g[1]) == 0 and G.value(g, 6, g[5]) == 0


def test_cubic_reproduction():
    ax = make_testaxis(10)
    p = lambda x: 2.0 - 0.7 * x + 1.3 * x ** 2 - 0.4 * x ** 3
    dp = lambda x: -0.7 + 2.6 * x - 1.2 * x ** 2
    for x in np.linspace(ax.knots[0], ax.knots[-1], 37):
        s = sum(p(ax.knots[k]) * G.value(ax.knots, 2 * k, x) + dp(ax.knots[k]) * G.value(ax.knots, 2 * k + 1, x)
                for k in range(ax.nknots))
        assert abs(s - p(x)) < 1e-12


def test_axes():
    ax = G.uniform_axis(-50.0, 50.0, 28)
    assert ax.nknots == 29 and ax.nbasis == 58
    assert ax.knots[0] == -50.0 and abs(ax.knots[-1] - 50.0) < 1e-13
    h = ax.knots[1] - ax.knots[0]
    assert abs(ax.colloc[1] - (ax.knots[0] + h * (1 - 1 / math.sqrt(3)) / 2)) < 1e-13
    a = G.stretch_ratio(50.0 / 6, 100.0, 8)
    assert a > 1 and abs(100.0 * (1 - a) / (1 - a ** 8) - 50.0 / 6) < 1e-13
    axs = G.stretched_axis(50.0, 150.0, 7, 8)
    assert axs.nknots == 29
    assert abs(axs.knots[0] + 150) < 1e-12 and abs(axs.knots[-1] - 150) < 1e-12
    assert abs(axs.knots[14]) < 1e-12
    assert np.allclose(axs.knots, -axs.knots[::-1])
    assert abs(np.diff(axs.knots)[7] - 50.0 / 6) < 1e-9 * 50 / 6
    # the exactly representable production step
    f = G.uniform_axis(-78.0, 78.0, 64)
    assert np.all(np.diff(f.knots) == 2.4375)


def test_coarse_sizes_match_fine():
    for nfine in (28, 64, 128):
        ni, no = G.coarse_grid_sizes(nfine)
        c = G.stretched_axis(78.0, 235.0, ni // 2, (no + 2) // 2)
        assert c.nbasis == 2 * nfine + 2
    with pytest.raises(ValueError):
        G.coarse_grid_sizes(27)


def test_moments_exact_identity():
    coeffs = (2.0, -0.7, 1.3, -0.4)
    p = lambda x: sum(c * x ** j for j, c in enumerate(coeffs))
    dp = lambda x: sum(j * c * x ** (j - 1) for j, c in enumerate(coeffs) if j)
    exact = lambda k, a, b: sum(c * (b ** (j + k + 1) - a ** (j + k + 1)) / (j + k + 1) for j, c in enumerate(coeffs))
    for ax in (G.uniform_axis(-2.0, 3.0, 9), G.stretched_axis(1.0, 4.0, 5, 6), make_testaxis(11)):
        for k in range(3):
            m = G.moments(ax, k)
            got = sum(p(ax.knots[kn]) * m[2 * kn] + dp(ax.knots[kn]) * m[2 * kn + 1] for kn in range(ax.nknots))
            assert abs(got - exact(k, ax.knots[0], ax.knots[-1])) <= 1e-11 * abs(exact(k, ax.knots[0], ax.knots[-1]))


def test_collocation_and_second_derivative():
    ax = make_testaxis(12)
    cm = G.collocation_matrices(ax)
    m = ax.nbasis
    i, j = np.indices((m, m))
    assert np.all(cm.S[np.abs(i - j) > 2] == 0)
    assert not np.allclose(cm.S, cm.S.T)
    D = G.laplacian1d_full(cm)
    assert np.linalg.norm(D @ cm.S - cm.S2) / np.linalg.norm(cm.S2) < 1e-12
    lam = np.linalg.eigvals(D[1:-1, 1:-1])
    assert np.max(np.abs(lam.imag)) < 1e-10 * np.max(np.abs(lam.real))
    assert np.all(lam.real < 0)
    # D is the exact second derivative of cubics at the interior points
    ax = G.uniform_axis(-2.0, 3.0, 10)
    Df = G.laplacian1d_full(G.collocation_matrices(ax))
    f = lambda x: 2.0 - 0.7 * x + 1.3 * x ** 2 - 0.4 * x ** 3
    f2 = lambda x: 2.6 - 2.4 * x
    got = Df @ f(ax.colloc)
    assert np.max(np.abs(got[1:-1] - f2(ax.colloc)[1:-1])) < 1e-11


def _solve(mesh, B):
    from oracle import oracle as O
    ni = mesh.ni
    w = O.apply_all(mesh.op.Minv, B, ni) * mesh.invlsum
    return O.apply_all(mesh.op.M, w, ni)


def test_tensor_solver_residual_and_convergence():
    L = 2.0

    def err(n):
        a = G.uniform_axis(0.0, L, n)
        m = G.SplineMesh.of(a)
        c = a.colloc[1:-1]
        k, j, i = np.meshgrid(c, c, c, indexing="ij")
        phi = (np.sin(np.pi * i / L) * np.sin(np.pi * j / L) * np.sin(np.pi * k / L)).ravel()
        x = _solve(m, (-3 * (np.pi / L) ** 2) * phi)
        return np.linalg.norm(x - phi) / np.linalg.norm(phi)

    a = G.uniform_axis(0.0, L, 8)
    m = G.SplineMesh.of(a)
    rng = np.random.default_rng(3)
    rho = rng.standard_normal(m.ni ** 3)
    phi = _solve(m, rho)
    # forward operator: sum_d D_d along each axis
    from oracle import oracle as O
    Dint = m.laplacian[1:-1, 1:-1]
    n = m.ni
    P = phi.reshape(n, n, n)
    lap = (P @ Dint.T) + np.einsum("ab,kbi->kai", Dint, P) + np.einsum("ab,bji->aji", Dint, P)
    assert np.linalg.norm(lap.ravel() - rho) / np.linalg.norm(rho) < 1e-10
    errs = [err(n) for n in (4, 8, 16)]
    assert errs[0] > errs[1] > errs[2]
    for e0, e1 in zip(errs, errs[1:]):
        assert math.log2(e0 / e1) > 3.8
    assert errs[-1] < 1e-6


def test_singular_and_nonreal_refused():
    with pytest.raises(ValueError):
        G.DiagonalizedOperator.of(np.array([[0.0, -1.0], [1.0, 0.0]]))
    with pytest.raises(ValueError):
        G.inv_lambda_sum(np.array([1.0]), np.array([-1.0]), np.array([0.0]))


@pytest.mark.parametrize("nfine", [28, 64])
def test_smoothing_tables_against_closed_form(nfine):
    """Tables by 1000-point trapezoid agree with the closed form to 1e-9 /
    1e-8 (test/runtests.jl:468-501); the value functions sum to 1 to 1e-5."""
    ax = G.uniform_axis(-78.0, 78.0, nfine)
    sm = G.GaussianSmoothing.of(ax)
    h = ax.knots[1] - ax.knots[0]
    for i in (0, 137, 500, 863, 1000):
        d = -h / 2 + i * h / 1000
        ov, gr = G.smoothing_columns(d, h)
        assert np.max(np.abs(ov - sm.overlap[i])) < 1e-9
        assert np.max(np.abs(gr - sm.gradient[i])) < 1e-8 / h * 2.4375
    assert np.max(np.abs(sm.overlap[:, 0::2].sum(axis=1) - 1)) < 1e-5
    # the node table is the 1-D Gaussian at the 8 collocation points of the window
    sig = h / 3
    r = (ax.knots[2] + ax.knots[1]) / 2 + 250 * h / 1000
    want = np.exp(-(r - ax.colloc[1:9]) ** 2 / (2 * sig * sig)) / (math.sqrt(2 * math.pi) * sig)
    assert np.allclose(sm.nodes[250], want, rtol=1e-14, atol=0)


def test_locate_table_is_exact():
    for ax in (G.stretched_axis(78.0, 235.0, 7, 8), G.stretched_axis(78.0, 235.0, 16, 17), make_testaxis(12, 5.0)):
        tbl = G.LocateTable.of(ax)
        rng = np.random.default_rng(7)
        xs = np.concatenate([rng.uniform(ax.colloc[0] - 1, ax.colloc[-1] + 1, 4000), ax.colloc])
        for x in xs:
            assert G.locate_table(tbl, ax, float(x)) == G.locate(ax, float(x))


def test_dual_lengths_cover_domain():
    for ax in (G.uniform_axis(-50.0, 50.0, 28), G.stretched_axis(78.0, 235.0, 16, 17)):
        l = G.dual_lengths(ax)
        assert np.all(l > 0)
        assert abs(l.sum() - (ax.knots[-1] - ax.knots[0])) < 1e-10
