"""The oracle's placement conventions pinned by known answers, each shown to
fail when a plausible mistake is planted in the oracle (CPU only).

* Coarse cell location and CIC weight orientation: the reference's own known
  answers for `locate` (test/runtests.jl:299-306), then a single particle on a
  collocation point puts its whole charge there and one at the midpoint of two
  points splits it 0.5/0.5.  Swapping the left and right weights (a mistake
  that charge conservation cannot see) fails.
* Smoothed-deposit stencil base: one particle at +d and one at -d about a knot
  deposit mirror-image 8^3 stencils (the window 2k-4 ... 2k+3 is symmetric
  about knot k, src/fields.jl:278-287).  A stencil shifted by one collocation
  point, or a table column off by one, fails.
* Table column of the field: for Phi = x^2/2 the smoothed field divided by
  that of Phi = x is the particle's offset rounded to the table, x_tab =
  knot + ((col - 1)/nbdt - 1/2) spacing (src/fields.jl:241-245;
  src/kernels.jl:414-419), to a small fraction of spacing/nbdt.  A column off
  by one fails.
"""
import numpy as np
import pytest

from conftest import load_package

va = load_package()
G = va.grids
f32 = np.float32


def record_at(target, x0, h):
    """A PackedParticle{Float32} record whose position, rebuilt as the device
    rebuilds it (x0 + Float32(k - 1) h + delta, src/kernels.jl:1217-1219),
    is exactly the Float32 values `target` (3,)."""
    x0f, hf = f32(x0), f32(h)
    rec = np.zeros(1, dtype=va.PACKED_DTYPE)
    for d in range(3):
        t = f32(target[d])
        k = int(np.rint((float(t) - float(x0f)) / float(hf))) + 1
        base = f32(x0f + f32(f32(k - 1) * hf))
        delta = f32(t - base)
        for _ in range(8):
            got = f32(base + delta)
            if got == t:
                break
            delta = np.nextafter(delta, f32(np.inf) if got < t else f32(-np.inf))
        assert f32(base + delta) == t
        rec["knode"][0, d] = rec["pknode"][0, d] = k
        rec["delta"][0, d] = rec["pdelta"][0, d] = delta
    return rec


def rebuilt(rec, x0, h):
    x0f, hf = f32(x0), f32(h)
    return np.array([f32(f32(x0f + f32(f32(rec["knode"][0, d] - 1) * hf)) + rec["delta"][0, d])
                     for d in range(3)])


def cic_cases(coarse):
    """(record target, expected {(i, j, k) 1-based: weight}) on the coarse axis."""
    g = coarse.axis.colloc.astype(np.float32)
    mid = f32((g[6] + g[7]) / f32(2))
    return [((g[4], g[4], g[4]), {(5, 5, 5): 1.0}),
            ((mid, g[4], g[4]), {(7, 5, 5): 0.5, (8, 5, 5): 0.5}),
            ((g[0], mid, g[-1]), {(1, 7, len(g)): 0.5, (1, 8, len(g)): 0.5})]


def cic_one(orc, rec, fine, coarse):
    k = fine.axis.knots
    c = coarse
    return orc.deposit_cic_raw(rec, f32(k[0]), f32(k[1] - k[0]), c.axis.colloc.astype(np.float32),
                               f32(c.locator.x0), f32(c.locator.invwidth), c.locator.cell)


def check_cic_cases(orc, fine, coarse):
    n = coarse.n
    k = fine.axis.knots
    for target, want in cic_cases(coarse):
        rho = cic_one(orc, record_at(target, k[0], k[1] - k[0]), fine, coarse).reshape(n, n, n)
        nz = {(int(i) + 1, int(j) + 1, int(kk) + 1): rho[kk, j, i] for kk, j, i in zip(*np.nonzero(rho))}
        assert set(nz) == set(want), (target, nz)
        for key, wv in want.items():
            assert abs(nz[key] - wv) < 2e-5, (target, key, nz[key])


def test_locate_known_answers(orc):
    """locate's conventions (test/runtests.jl:299-306) on the oracle's Float32
    table walk: on point 5 the cell is 4 with weight 0 (all of it to point 5),
    on point 1 the cell is 1 with weight 1, outside the domain 0, half-way
    between points 7 and 8 the cell is 7 with weight 1/2."""
    for ax in (G.uniform_axis(-50.0, 50.0, 28), G.stretched_axis(78.0, 235.0, 14, 15)):
        tbl = G.LocateTable.of(ax)
        g = ax.colloc.astype(np.float32)
        args = (g, tbl.x0, tbl.invwidth, tbl.cell)
        assert orc.locate(g[4], *args) == (4, f32(0.0))
        assert orc.locate(g[0], *args) == (1, f32(1.0))
        assert orc.locate(f32(g[0] - f32(10.0)), *args)[0] == 0
        assert orc.locate(f32(g[-1] + f32(10.0)), *args)[0] == 0
        c, w = orc.locate(f32((g[6] + g[7]) / f32(2)), *args)
        # the Float32 midpoint is within an ulp of the middle of the cell
        assert c == 7 and abs(w - 0.5) < 2e-5
        # the last point: cell n-1, weight 0 (all of it to point n)
        assert orc.locate(g[-1], *args) == (len(g) - 1, f32(0.0))
        # the bisection of locate (src/deposition.jl:42-47),
        # max(1, searchsortedfirst(gt, x) - 1), on the same Float32 points
        # gives the same cells as the table walk
        for x in np.concatenate([g, (g[1:] + g[:-1]) / f32(2)]):
            c = max(1, int(np.searchsorted(g, x, side="left")))
            assert orc.locate(x, *args)[0] == c


def test_cic_single_particle_placement(orc, config1):
    _, fine, coarse, _ = config1
    check_cic_cases(orc, fine, coarse)


def test_cic_placement_fails_with_swapped_weights(orc, config1):
    _, fine, coarse, _ = config1
    with orc.planted("CIC_SWAP"):
        with pytest.raises(AssertionError):
            check_cic_cases(orc, fine, coarse)
    check_cic_cases(orc, fine, coarse)


def mirror_pair(fine, sm, j=137):
    """Records at knot ci + d and ci - d (x only), d = spacing (j + 0.3)/nbdt,
    so that their table columns are c and nbdt + 2 - c."""
    kn = fine.axis.knots
    h = float(kn[1] - kn[0])
    ci = len(kn) // 2 + 1                    # 1-based knot near the centre
    d = f32(sm.spacing * (j + 0.3) / sm.nbdt)
    dy, dz = f32(0.21 * h), f32(-0.37 * h)
    recs = []
    for s in (1, -1):
        r = np.zeros(1, dtype=va.PACKED_DTYPE)
        r["knode"][0] = r["pknode"][0] = (ci, ci, ci)
        r["delta"][0] = r["pdelta"][0] = (f32(s) * d, dy, dz)
        recs.append(r)
    return ci, recs


def mirror_map(colloc, centre):
    """For every collocation point i, the point i' with colloc[i'] = 2 centre - colloc[i]
    (-1 when there is none)."""
    m = np.full(len(colloc), -1)
    for i, x in enumerate(colloc):
        j = int(np.argmin(np.abs(colloc - (2 * centre - x))))
        if abs(colloc[j] - (2 * centre - x)) < 1e-9 * (1 + abs(x)):
            m[i] = j
    return m


def check_mirror(rho_a, rho_b, n, mirror):
    a = rho_a.reshape(n, n, n)
    b = rho_b.reshape(n, n, n)
    sup = np.nonzero(np.any(a != 0, axis=(0, 1)))[0]
    assert len(sup) == 8
    assert np.all(mirror[sup] >= 0)
    b_mirrored = b[:, :, mirror[sup]]
    err = np.max(np.abs(a[:, :, sup] - b_mirrored)) / np.max(np.abs(a))
    assert err < 1e-9, err
    assert np.count_nonzero(b) == np.count_nonzero(a) == 8 ** 3


def smoothed_one(orc, rec, fine, sm):
    kn = fine.axis.knots
    nk = len(kn)
    perm, occ, bnd = orc.cellsort(rec, nk)
    return orc.deposit_smoothed_raw(rec, occ, bnd, nk, fine.n, sm.nodes, f32(sm.spacing), sm.nbdt,
                                    sm.nbdt + 1)


def check_smoothed_mirror(orc, fine, sm):
    ci, (ra, rb) = mirror_pair(fine, sm)
    mirror = mirror_map(fine.axis.colloc, fine.axis.knots[ci - 1])
    check_mirror(smoothed_one(orc, ra, fine, sm), smoothed_one(orc, rb, fine, sm), fine.n, mirror)


def test_smoothed_deposit_mirror_stencils(orc, config1):
    _, fine, _, sm = config1
    check_smoothed_mirror(orc, fine, sm)


@pytest.mark.parametrize("mistake", ["STENCIL_SHIFT", "COLUMN_SLIP"])
def test_smoothed_mirror_fails_with_planted_mistake(orc, config1, mistake):
    _, fine, _, sm = config1
    with orc.planted(mistake):
        with pytest.raises(AssertionError):
            check_smoothed_mirror(orc, fine, sm)


def hermite_dofs(ax, f, df):
    a = np.zeros(ax.nbasis)
    a[0::2] = f(ax.knots)
    a[1::2] = df(ax.knots)
    return a


def x_potential(ax, f, df):
    """csol of Phi(x, y, z) = f(x), Julia order (x fastest)."""
    one = np.zeros(ax.nbasis)
    one[0::2] = 1.0
    fx = hermite_dofs(ax, f, df)
    return (one[:, None, None] * one[None, :, None] * fx[None, None, :]).ravel()


def tabulated_x(rec, x0, h, sm):
    """knot + ((col - 1)/nbdt - 1/2) spacing, with col the Float32 column of
    the record's offset (src/kernels.jl:414-419)."""
    d = rec["delta"][:, 0].astype(np.float32)
    sp = f32(sm.spacing)
    t = (d + sp * f32(0.5)) / sp * f32(sm.nbdt) + f32(0.5)
    col = np.clip(np.floor(t).astype(np.int64) + 1, 1, sm.nbdt + 1)
    knot = x0 + (rec["knode"][:, 0].astype(np.float64) - 1) * h
    return knot + ((col - 1) / sm.nbdt - 0.5) * float(sm.spacing)


QUAD_KNOTS = (8, 15, 22)      # 1-based knots the quadratic potentials are centred on


def quadratic_cloud(fine, ci, n=2000, seed=5):
    """Particles whose nearest knot along x is knot ci (offsets over the whole
    table), y and z anywhere inside the fine grid's stencil-fitting region."""
    kn = fine.axis.knots
    x0, h = float(kn[0]), float(kn[1] - kn[0])
    rng = np.random.default_rng(seed + ci)
    pos = rng.uniform(-30.0, 30.0, (n, 3))
    pos[:, 0] = kn[ci - 1] + rng.uniform(-0.4999, 0.4999, n) * h
    return va.physics.packed_records(pos, pos, x0, h), x0, h


def field_ratio_error(fx_quad, fx_lin, xtab, sm):
    ratio = fx_quad / fx_lin
    return float(np.max(np.abs(ratio - xtab)) / (float(sm.spacing) / sm.nbdt))


def quadratic_potentials(fine, ci):
    X = float(fine.axis.knots[ci - 1])
    quad = x_potential(fine.axis, lambda x: (x - X) ** 2 / 2, lambda x: x - X)
    lin = x_potential(fine.axis, lambda x: x - X, np.ones_like)
    return X, quad, lin


def check_quadratic_field(orc, fine, sm):
    """Phi = (x - X)^2/2 about a knot X: the field divided by that of
    Phi = x - X is x_tab - X.  Centring on the particles' knot keeps the tabulated
    kernel's truncation (its derivative integrates to ~1e-4 at the table's
    end columns, times (x - X)^2/2) far below the bound; a one-column slip
    moves the ratio by a whole spacing/nbdt."""
    nk = fine.axis.nknots
    for ci in QUAD_KNOTS:
        rec, x0, h = quadratic_cloud(fine, ci)
        perm, occ, bnd = orc.cellsort(rec, nk)
        srec = rec[perm].copy()
        X, quad, lin = quadratic_potentials(fine, ci)
        args = (x0, h, f32(sm.spacing), sm.nbdt, sm.nbdt + 1, 1.0)
        fq, _ = orc.field(srec, occ, bnd, nk, fine.n, quad, sm.overlap, sm.gradient, *args)
        fl, _ = orc.field(srec, occ, bnd, nk, fine.n, lin, sm.overlap, sm.gradient, *args)
        assert np.all(np.isfinite(fq)) and np.all(np.isfinite(fl))
        # F = -w grad: Phi = x - X gives -N (N ~ 1 the kernel's normalisation)
        assert np.max(np.abs(fl[:, 0] + 1.0)) < 1e-3
        err = field_ratio_error(fq[:, 0], fl[:, 0], tabulated_x(srec, x0, h, sm) - X, sm)
        assert err < 0.1, (ci, err)
        # nothing along y and z beyond the kernel's truncation
        assert np.max(np.abs(fq[:, 1:])) < 1e-2 * np.max(np.abs(fq[:, 0]))


def test_field_of_a_quadratic_potential(orc, config1):
    _, fine, _, sm = config1
    check_quadratic_field(orc, fine, sm)


def test_quadratic_field_fails_with_a_column_slip(orc, config1):
    _, fine, _, sm = config1
    with orc.planted("COLUMN_SLIP"):
        with pytest.raises(AssertionError):
            check_quadratic_field(orc, fine, sm)
