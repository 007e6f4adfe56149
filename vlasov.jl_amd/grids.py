"""FP64 host-side construction of everything the per-step kernels read.

This is the analogue of the reference's Julia constructors, which run once on
the host in Float64 before the time loop (src/simulation.jl:175-183):

* spline axes and collocation points       (src/splines.jl:159-287)
* Hermite basis evaluation                  (src/splines.jl:74-139)
* exact Gauss-3 moments                     (src/splines.jl:302-340)
* locate table of the stretched axis        (src/splines.jl:362-396)
* collocation matrices S, S', S'' and S^-1  (src/collocation.jl:34-54)
* D = S'' S^-1, eigen-factorisation of its interior block, 1/(l_i+l_j+l_k)
                                            (src/collocation.jl:94, src/tensorsolver.jl:128-168)
* dual moments S^-T m_k and dual lengths    (src/mesh.jl:61-62, src/deposition.jl:185-195)
* Gaussian smoothing tables (trapezoid)     (src/fields.jl:162-217)

Indices are 0-based here; the docstrings say where the reference's 1-based
formula was shifted.  Everything is returned as numpy float64 arrays; the
device library narrows the tables to float32 when it uploads them, exactly as
the reference's `dev(x) = copyto!(zeros(E, ...), E.(x))` does.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from fractions import Fraction

import numpy as np

GAUSS2_NODES = ((1.0 - 1.0 / math.sqrt(3.0)) / 2.0, (1.0 + 1.0 / math.sqrt(3.0)) / 2.0)
GAUSS3_NODES = (-math.sqrt(3.0 / 5.0), 0.0, math.sqrt(3.0 / 5.0))
GAUSS3_WEIGHTS = (5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0)

VALUE, SLOPE = 0, 1


# ---------------------------------------------------------------------------
# axes
# ---------------------------------------------------------------------------

def julia_range(x0: float, xn: float, length: int) -> np.ndarray:
    """`range(x0, xn; length)` for Float64 endpoints: each element is the
    correctly rounded value of x0 + i (xn - x0)/(length-1), which is what the
    reference's TwicePrecision ranges deliver (src/splines.jl:220, 281)."""
    a, b = Fraction(x0), Fraction(xn)
    d = length - 1
    return np.array([float(a + (b - a) * i / d) for i in range(length)], dtype=np.float64)


def collocation_points(knots: np.ndarray) -> np.ndarray:
    """Ends of the domain plus the two Gauss points of every interval
    (src/splines.jl:170-174)."""
    out = [knots[0]]
    for j in range(len(knots) - 1):
        h = knots[j + 1] - knots[j]
        for u in GAUSS2_NODES:
            out.append(knots[j] + h * u)
    out.append(knots[-1])
    return np.array(out, dtype=np.float64)


@dataclass
class SplineAxis:
    """Knots and collocation points of one axis (src/splines.jl:46-56)."""
    knots: np.ndarray
    colloc: np.ndarray

    def __post_init__(self):
        if np.any(np.diff(self.knots) < 0):
            raise ValueError("knots must be increasing")
        if len(self.colloc) != 2 * len(self.knots):
            raise ValueError("2 collocation points per knot are required")

    @property
    def nknots(self) -> int:
        return len(self.knots)

    @property
    def nbasis(self) -> int:
        return 2 * len(self.knots)


def uniform_axis(x0: float, xn: float, nintervals: int) -> SplineAxis:
    """Constant-step axis (src/splines.jl:219-222)."""
    k = julia_range(float(x0), float(xn), nintervals + 1)
    return SplineAxis(k, collocation_points(k))


def stretch_ratio(h1: float, L: float, n: int) -> float:
    """Bisection to the last float for L(1-a)/(1-a^n) = h1 (src/splines.jl:239-253)."""
    def f(a):
        return L * (1.0 - a) / (1.0 - a ** n) - h1
    lo, hi = math.nextafter(1.0, 2.0), 10.0
    flo = f(lo)
    if math.copysign(1.0, flo) == math.copysign(1.0, f(hi)):
        raise ValueError(f"no geometric ratio in ]1, 10] for h1={h1}, L={L}, n={n}")
    while math.nextafter(lo, math.inf) < hi:
        mid = (lo + hi) / 2.0
        if mid == lo or mid == hi:
            break
        if math.copysign(1.0, f(mid)) == math.copysign(1.0, flo):
            lo = mid
        else:
            hi = mid
    return (lo + hi) / 2.0


def stretched_axis(xinner: float, xouter: float, n_inner: int, n_outer: int) -> SplineAxis:
    """Constant step on [0, xinner], geometric steps up to xouter, mirrored
    about zero (src/splines.jl:272-287)."""
    h1 = xinner / (n_inner - 1)
    L = xouter - xinner
    a = stretch_ratio(h1, L, n_outer)
    step = L * (1.0 - a) / (1.0 - a ** n_outer)
    inner = julia_range(0.0, float(xinner), n_inner)
    acc, outer = 0.0, []
    for k in range(n_outer):
        acc += step * a ** k
        outer.append(xinner + acc)
    half = np.concatenate([inner, np.array(outer)])
    knots = np.concatenate([-half[1:][::-1], half])
    return SplineAxis(knots, collocation_points(knots))


def coarse_grid_sizes(nfine: int) -> tuple[int, int]:
    """`ninner, nouter` derived from an even `nfine`, so that both levels
    carry 2 nfine + 2 basis functions (scripts/xenon.jl:162-167)."""
    if nfine % 2:
        raise ValueError(f"nfine must be even (got {nfine})")
    m = nfine // 2 + 1
    n1 = m // 2
    return 2 * n1, 2 * (m - n1) - 2


# ---------------------------------------------------------------------------
# Hermite basis (0-based: basis b <-> knot b // 2, kind b % 2)
# ---------------------------------------------------------------------------

def _hermite(kind, a, h, s, d):
    """Value (d=0), first (d=1) and second (d=2) derivatives on the reduced
    variable a in [0, 1], orientation s = +-1 (src/splines.jl:74-84)."""
    if d == 0:
        return 3.0 * (a * a) - 2.0 * (a * a * a) if kind == VALUE else h * ((a * a) - (a * a * a)) * s
    if d == 1:
        return -6.0 * (a - a * a) * s / h if kind == VALUE else -2.0 * a + 3.0 * (a * a)
    return 6.0 * (1.0 - 2.0 * a) / (h * h) if kind == VALUE else (2.0 - 6.0 * a) * s / h


def _localcoords(g, knot, x):
    """Reduced coordinate of x in the support of the basis of `knot`
    (0-based), or None (src/splines.jl:110-122)."""
    n = len(g)
    if knot > 0 and x <= g[knot]:
        if x < g[knot - 1]:
            return None
        h = g[knot] - g[knot - 1]
        return (x - g[knot - 1]) / h, h, -1.0
    if knot < n - 1 and x >= g[knot]:
        if x > g[knot + 1]:
            return None
        h = g[knot + 1] - g[knot]
        return (g[knot + 1] - x) / h, h, 1.0
    return None


def evaluate(g, basis: int, x: float, nder: int = 0):
    """Basis `basis` (0-based linear index) and its first `nder` derivatives
    at x; zeros outside the support (src/splines.jl:130-139)."""
    loc = _localcoords(g, basis // 2, x)
    if loc is None:
        return tuple(0.0 for _ in range(nder + 1))
    a, h, s = loc
    return tuple(_hermite(basis % 2, a, h, s, d) for d in range(nder + 1))


def value(g, basis: int, x: float) -> float:
    return evaluate(g, basis, x, 0)[0]


def _gauss3(g, basis, a, c, k):
    mid, half = (a + c) / 2.0, (c - a) / 2.0
    tot = 0.0
    for xi, w in zip(GAUSS3_NODES, GAUSS3_WEIGHTS):
        x = mid + half * xi
        tot += w * x ** k * value(g, basis, x)
    return half * tot


def moments(ax: SplineAxis, k: int) -> np.ndarray:
    """Exact integral of x^k phi_b over each basis support, per half-support
    (src/splines.jl:325-340)."""
    g, n = ax.knots, ax.nknots
    out = np.zeros(ax.nbasis)
    for b in range(ax.nbasis):
        kn = b // 2
        tot = 0.0
        if kn > 0:
            tot += _gauss3(g, b, g[kn - 1], g[kn], k)
        if kn < n - 1:
            tot += _gauss3(g, b, g[kn], g[kn + 1], k)
        out[b] = tot
    return out


# ---------------------------------------------------------------------------
# collocation matrices and the 1-D operator
# ---------------------------------------------------------------------------

@dataclass
class Collocation:
    axis: SplineAxis
    S: np.ndarray
    S1: np.ndarray
    S2: np.ndarray
    Sinv: np.ndarray


def collocation_matrices(ax: SplineAxis) -> Collocation:
    """Pentadiagonal S, S', S'' (rows: points, columns: basis functions) and
    the dense inverse of S (src/collocation.jl:34-54)."""
    m = ax.nbasis
    S, S1, S2 = np.zeros((m, m)), np.zeros((m, m)), np.zeros((m, m))
    g, tau = ax.knots, ax.colloc
    for j in range(ax.nknots - 1):          # interval j joins knots j and j+1
        for lin in range(2 * j, 2 * j + 4):
            for k in (2 * j + 1, 2 * j + 2):
                S[k, lin], S1[k, lin], S2[k, lin] = evaluate(g, lin, tau[k], 2)
    S[0, 0], S1[0, 0], S2[0, 0] = evaluate(g, 0, tau[0], 2)
    S[m - 1, m - 2], S1[m - 1, m - 2], S2[m - 1, m - 2] = evaluate(g, m - 2, tau[m - 1], 2)
    return Collocation(ax, S, S1, S2, np.linalg.inv(S))


def laplacian1d_full(cm: Collocation) -> np.ndarray:
    """D = S'' S^-1 by an LU solve, boundary rows and columns kept
    (src/collocation.jl:94)."""
    return np.linalg.solve(cm.S.T, cm.S2.T).T


@dataclass
class DiagonalizedOperator:
    """D = M diag(l) M^-1 with a real spectrum (src/tensorsolver.jl:123-138).
    Eigenpairs are ordered by increasing eigenvalue, as LinearAlgebra.eigen
    sorts them."""
    M: np.ndarray
    Minv: np.ndarray
    lam: np.ndarray

    @classmethod
    def of(cls, D: np.ndarray, atol: float = 1e-10):
        w, V = np.linalg.eig(D)
        imagmax = float(np.max(np.abs(w.imag)))
        scale = float(np.max(np.abs(w.real)))
        if imagmax > atol * max(scale, 1.0):
            raise ValueError(f"non-real spectrum (max|Im l| = {imagmax}): "
                             "the real tensor solver does not apply")
        order = np.lexsort((w.imag, w.real))
        M = np.ascontiguousarray(V[:, order].real)
        return cls(M, np.linalg.inv(M), np.ascontiguousarray(w.real[order]))


def inv_lambda_sum(lx, ly, lz) -> np.ndarray:
    """1/(lx_i + ly_j + lz_k), stored [i + n j + n^2 k] (Julia column-major);
    refuses a vanishing sum (src/tensorsolver.jl:157-168)."""
    s = (lx[:, None, None] + ly[None, :, None]) + lz[None, None, :]
    if np.any(s == 0):
        raise ValueError("eigenvalue sum vanishes: singular operator")
    return np.ascontiguousarray((1.0 / s).transpose(2, 1, 0)).ravel()   # [k][j][i]: i fastest


# ---------------------------------------------------------------------------
# deposition helpers
# ---------------------------------------------------------------------------

def dual_lengths(ax: SplineAxis) -> np.ndarray:
    """Share of the domain owned by each collocation point
    (src/deposition.jl:185-195)."""
    gt, g = ax.colloc, ax.knots
    m = len(gt)
    out = np.empty(m)
    for j in range(m):        # 0-based; the reference's j == 1 is j == 0 here
        if j == 0:
            out[j] = (gt[1] - g[0]) / 2
        elif j == 1:
            out[j] = (gt[2] - g[0]) / 2
        elif j == m - 1:
            out[j] = (g[-1] - gt[m - 2]) / 2
        elif j == m - 2:
            out[j] = (g[-1] - gt[m - 3]) / 2
        else:
            out[j] = (gt[j + 1] - gt[j - 1]) / 2
    return out


@dataclass
class LocateTable:
    """Uniform bins over the collocation points giving a candidate cell
    (src/splines.jl:362-376).  `cell` holds the reference's 1-based cell
    indices (the left collocation point of the interval)."""
    x0: float
    invwidth: float
    cell: np.ndarray

    @classmethod
    def of(cls, ax: SplineAxis):
        gt = ax.colloc
        n = len(gt)
        width = float(np.min(np.diff(gt))) / 2.0
        nbin = math.ceil((gt[-1] - gt[0]) / width) + 2
        cells = np.empty(nbin, dtype=np.int32)
        for b in range(nbin):
            x = gt[0] + b * width
            c = int(np.searchsorted(gt, x, side="right"))   # searchsortedlast, 1-based
            cells[b] = min(max(c, 1), n - 1)
        return cls(float(gt[0]), 1.0 / width, cells)


def locate(ax: SplineAxis, x: float):
    """(cell, weight) of x between collocation points, 1-based cell, or None
    (src/deposition.jl:204-209)."""
    gt = ax.colloc
    if x < gt[0] or x > gt[-1]:
        return None
    c = max(1, int(np.searchsorted(gt, x, side="left")) + 1 - 1)
    return c, (gt[c] - x) / (gt[c] - gt[c - 1])


def locate_table(tbl: LocateTable, ax: SplineAxis, x: float):
    """Same contract as `locate`, through the table (src/splines.jl:384-396)."""
    gt = ax.colloc
    if x < gt[0] or x > gt[-1]:
        return None
    b = min(math.floor((x - tbl.x0) * tbl.invwidth) + 1, len(tbl.cell))
    c = int(tbl.cell[b - 1])
    last = len(gt) - 1
    while c < last and gt[c] < x:        # gt[c+1] in 1-based indexing
        c += 1
    return c, (gt[c] - x) / (gt[c] - gt[c - 1])


# ---------------------------------------------------------------------------
# meshes
# ---------------------------------------------------------------------------

@dataclass
class SplineMesh:
    """Everything a grid level contributes to a step: the three (identical)
    axes and the tables of src/mesh.jl:21-76, src/devicemesh.jl:21-81."""
    axis: SplineAxis
    coll: Collocation
    laplacian: np.ndarray          # full D, n x n (row = point, column = function)
    op: DiagonalizedOperator       # interior block
    invlsum: np.ndarray            # (n-2)^3, i fastest
    dual_moments: tuple            # (m0, m1, m2), S^-T m_k
    dual_len: np.ndarray
    locator: LocateTable

    @property
    def n(self) -> int:
        return self.axis.nbasis

    @property
    def ni(self) -> int:
        return self.axis.nbasis - 2

    @classmethod
    def of(cls, ax: SplineAxis):
        cm = collocation_matrices(ax)
        D = laplacian1d_full(cm)
        op = DiagonalizedOperator.of(D[1:-1, 1:-1])
        duals = tuple(cm.Sinv.T @ moments(ax, k) for k in range(3))
        return cls(ax, cm, D, op, inv_lambda_sum(op.lam, op.lam, op.lam),
                   duals, dual_lengths(ax), LocateTable.of(ax))


def nested_meshes(fine: "SplineMesh", coarse: "SplineMesh"):
    """Refuses a fine level that sticks out of the coarse one
    (src/mesh.jl:94-104)."""
    if not (coarse.axis.knots[0] <= fine.axis.knots[0] and fine.axis.knots[-1] <= coarse.axis.knots[-1]):
        raise ValueError("level 1 sticks out of level 2")
    return fine, coarse


# ---------------------------------------------------------------------------
# Gaussian smoothing tables
# ---------------------------------------------------------------------------

SMOOTHING_SUPPORTS = ((1, 2), (1, 2), (1, 3), (1, 3), (2, 4),
                      (2, 4), (3, 5), (3, 5), (4, 5), (4, 5))


def _basis_values(g, b, xs):
    return np.array([value(g, b, float(x)) for x in xs])


@dataclass
class GaussianSmoothing:
    """Overlap and gradient (10 x nbdt+1) and node (8 x nbdt+1) tables of a
    constant-step axis, sigma = h/3, trapezoid quadrature with accumulated
    abscissae (src/fields.jl:162-217).  Stored as [column][a] (a fastest),
    i.e. the reference's column-major 10 x 1001 layout."""
    sigma: float
    spacing: float
    nbdt: int
    overlap: np.ndarray
    gradient: np.ndarray
    nodes: np.ndarray

    @classmethod
    def of(cls, ax: SplineAxis, nbdt: int = 1000, quadrature: int = 1000):
        g = ax.knots
        h = g[1] - g[0]
        sigma = h / 3.0
        norm1d = 1.0 / (math.sqrt(2.0 * math.pi) * sigma)
        rmin, rmax = (g[2] + g[1]) / 2.0, (g[3] + g[2]) / 2.0
        r = np.array([rmin + i * (rmax - rmin) / nbdt for i in range(nbdt + 1)])
        s2 = 2.0 * (sigma * sigma)
        overlap = np.zeros((nbdt + 1, 10))
        gradient = np.zeros((nbdt + 1, 10))
        for a in range(10):
            lo, hi = SMOOTHING_SUPPORTS[a]
            lo_x, hi_x = g[lo - 1], g[hi - 1]
            step = (hi_x - lo_x) / quadrature
            xs = [lo_x]
            for _ in range(quadrature):          # accumulated abscissae
                xs.append(xs[-1] + step)
            xs = np.array(xs)
            phi = _basis_values(g, a, xs)         # basis a (0-based) of the window
            # integrands for every r at once; the sum runs in the reference's order
            d = r[:, None] - xs[None, :]
            gauss = norm1d * np.exp(-(d * d) / s2)
            fo = phi[None, :] * gauss
            fg = (-d / (sigma * sigma)) * phi[None, :] * gauss
            to = np.zeros(nbdt + 1)
            tg = np.zeros(nbdt + 1)
            for q in range(quadrature):
                to += (fo[:, q] + fo[:, q + 1]) * step / 2.0
                tg += (fg[:, q] + fg[:, q + 1]) * step / 2.0
            overlap[:, a] = to
            gradient[:, a] = tg
        nodes = np.zeros((nbdt + 1, 8))
        for a in range(8):
            d = r - ax.colloc[a + 1]
            nodes[:, a] = norm1d * np.exp(-(d * d) / s2)
        return cls(sigma, h, nbdt, overlap, gradient, nodes)


def smoothing_moments(u: float):
    """Primitives of u^n exp(-u^2/2), n = 0..4 (src/fields.jl:79-83)."""
    e = math.exp(-u * u / 2.0)
    s = math.sqrt(math.pi / 2.0) * math.erf(u / math.sqrt(2.0))
    return (s, -e, s - u * e, -(u * u + 2.0) * e, 3.0 * s - (u * u * u + 3.0 * u) * e)


def smoothing_columns(delta: float, h: float):
    """Closed-form overlap and gradient columns at offset delta
    (src/fields.jl:124-160); the exact reference the tables are held to."""
    t = 3.0 * delta / h
    M = [smoothing_moments(3.0 * (j - 3) - t) for j in range(1, 6)]
    ov, gr = np.zeros(10), np.zeros(10)
    third = 1.0 / 3.0
    for a in range(10):
        kind, k = a % 2, a // 2 + 1
        lo, hi = SMOOTHING_SUPPORTS[a]
        o = dg = 0.0
        for j, direc in ((k - 1, 1), (k, -1)):
            if not (1 <= j <= 4 and j >= lo and j + 1 <= hi):
                continue
            d = [M[j][n] - M[j - 1][n] for n in range(5)]
            al = -(3.0 * (j - 3) - t) * third if direc > 0 else (3.0 * (j - 2) - t) * third
            be = third if direc > 0 else -third
            q2 = al * al * d[0] + 2 * al * be * d[1] + be * be * d[2]
            q3 = al ** 3 * d[0] + 3 * al ** 2 * be * d[1] + 3 * al * be ** 2 * d[2] + be ** 3 * d[3]
            r2 = al * al * d[1] + 2 * al * be * d[2] + be * be * d[3]
            r3 = al ** 3 * d[1] + 3 * al ** 2 * be * d[2] + 3 * al * be ** 2 * d[3] + be ** 3 * d[4]
            sgn = -1.0 if direc > 0 else 1.0
            cv, cs = (3.0, -2.0) if kind == VALUE else (sgn, -sgn)
            sc = 1.0 if kind == VALUE else h
            o += sc * (cv * q2 + cs * q3)
            dg += sc * (cv * r2 + cs * r3)
        ov[a], gr[a] = o, dg
    sigma = h / 3.0
    norm = 1.0 / (math.sqrt(2.0 * math.pi) * sigma)
    return norm * sigma * ov, norm * gr
