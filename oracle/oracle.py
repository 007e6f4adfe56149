"""CPU oracle of the per-step loop -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline
leg, as the checker and the timed CPU baseline; the product never imports it.

It restates the reference's algorithm for the path in FP64:
  * particle loops in plain C (oracle/vlasov_oracle.c, built by `build()`),
  * the grid chain in numpy, following src/poisson.jl, src/tensorsolver.jl,
    src/deposition.jl, src/meanfield.jl and src/devicemesh.jl,
  * the stage order of one resident step, _update_forces_resident!
    (src/simulation.jl:472-531) followed by the device Verlet
    (src/accelerator.jl:580-605).
Grids are flat float64 arrays in Julia column-major order ([i + n j + n^2 k]).
"""
from __future__ import annotations

import contextlib
import ctypes as C
import math
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
SRC = HERE / "vlasov_oracle.c"
LIB = HERE / "_build" / "liboracle.so"

REC = np.dtype([("knode", "<i4", (3,)), ("delta", "<f4", (3,)),
                ("pknode", "<i4", (3,)), ("pdelta", "<f4", (3,))])

_lib = None


PLANTS = ("COLUMN_SLIP", "STENCIL_SHIFT", "CIC_SWAP")


def build(force: bool = False, plant: str | None = None) -> Path:
    """Compiles the C restatement with gcc (OpenMP, no FP contraction).
    `plant` names one ORACLE_PLANT_* mistake: that copy goes to its own file
    and is only ever loaded by `planted()`."""
    LIB.parent.mkdir(exist_ok=True)
    if plant is not None and plant not in PLANTS:
        raise ValueError(f"unknown planted mistake {plant!r}")
    out = LIB if plant is None else LIB.with_name(f"liboracle_plant_{plant.lower()}.so")
    if force or not out.exists() or out.stat().st_mtime < SRC.stat().st_mtime:
        tmp = out.with_suffix(".so.tmp")
        cmd = ["gcc", "-O2", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-fno-fast-math",
               "-std=c11", str(SRC), "-o", str(tmp), "-lm"]
        if plant is not None:
            cmd.insert(1, f"-DORACLE_PLANT_{plant}")
        subprocess.run(cmd, check=True)
        os.replace(tmp, out)
    return out


@contextlib.contextmanager
def planted(mistake: str):
    """Within the block, every oracle function runs a copy of the C
    restatement built with the planted mistake `mistake` (one of PLANTS)."""
    global _lib
    saved = lib()
    _lib = _bind(C.CDLL(str(build(plant=mistake))))
    try:
        yield
    finally:
        _lib = saved


def lib():
    global _lib
    if _lib is None:
        build()
        _lib = _bind(C.CDLL(str(LIB)))
    return _lib


def _bind(L):
    """Declares the argument and result types of the C restatement's functions."""
    i64, i32, f32, f64 = C.c_int64, C.c_int, C.c_float, C.c_double
    vp = C.c_void_p
    L.or_cellsort.argtypes = [vp, i64, i32, vp, vp, vp]
    L.or_cellsort.restype = i64
    L.or_outside.argtypes = [vp, i64, i32, vp]
    L.or_outside.restype = i64
    L.or_columns.argtypes = [vp, i64, f32, f32, f32, f32, f32, f32, i32, vp]
    L.or_columns.restype = i64
    L.or_deposit_smoothed.argtypes = [vp, vp, vp, i64, i32, i32, vp, f32, f32, i32, vp]
    L.or_deposit_cic.argtypes = [vp, i64, f32, f32, vp, i32, f32, f32, vp, i32, vp]
    L.or_field.argtypes = [vp, vp, vp, i64, i32, i32, vp, vp, vp, f64, f64, f32, f32, i32, f64,
                           vp, i32, f64, vp, vp, i32]
    L.or_spline_potential.argtypes = [vp, i32, vp, i32, vp, i64, vp]
    L.or_fallback.argtypes = [vp, vp, i64, f64, f64, vp, i32, vp, i32, f64, vp, i32, f64, f64, vp]
    L.or_verlet_f32.argtypes = [vp, vp, i64, f32, f32, f32, f32, f32, f32, vp]
    L.or_interaction_energy.argtypes = [vp, i64, f64, f64, vp, i32, vp, i32, vp, f64, i32, i32,
                                        vp, i32, vp, i32, f64, f64, vp]
    L.or_interaction_energy.restype = f64
    L.or_enclosed.argtypes = [vp, i64, f64, f64, vp, f64]
    L.or_enclosed.restype = i64
    L.or_num_threads.restype = i32
    L.or_fallback_points.argtypes = [vp, i64, vp, i32, vp, i32, f64, vp, i32, f64, f64, vp]
    L.or_locate.argtypes = [vp, i32, f32, f32, vp, i32, f32, C.POINTER(f32)]
    L.or_locate.restype = i32
    return L


def _p(a: np.ndarray):
    # the C loops read the buffer densely: a strided view (rec[::k]) would
    # hand them its base pointer and they would read the wrong records
    if not a.flags.c_contiguous:
        raise ValueError("oracle arrays must be C-contiguous (use np.ascontiguousarray)")
    return a.ctypes.data


# ---------------------------------------------------------------------------
# particle loops
# ---------------------------------------------------------------------------
def cellsort(rec: np.ndarray, nk: int):
    """Stable counting sort (src/sorting.jl:163-204): perm, occupied (0-based
    keys), bounds."""
    n = len(rec)
    perm = np.zeros(n, np.int32)
    occ = np.zeros(min(n, nk ** 3) + 1, np.int32)
    bnd = np.zeros(min(n, nk ** 3) + 2, np.int32)
    nocc = lib().or_cellsort(_p(rec), n, nk, _p(perm), _p(occ), _p(bnd))
    return perm, occ[:nocc].copy(), bnd[:nocc + 1].copy()


def outside(rec: np.ndarray, ngrid: int) -> np.ndarray:
    out = np.zeros(len(rec), np.int32)
    m = lib().or_outside(_p(rec), len(rec), ngrid, _p(out))
    return out[:m].copy()


def columns(rec, x0, h, lo, hi, spacing, nbdt, ncol):
    cols = np.zeros(3 * len(rec), np.int32)
    nout = lib().or_columns(_p(rec), len(rec), x0, h, lo, hi, spacing, float(nbdt), ncol, _p(cols))
    return cols.reshape(-1, 3), int(nout)


def deposit_smoothed_raw(srec, occ, bnd, nk, ngrid, nodes, spacing, nbdt, ncol):
    rho = np.zeros(ngrid ** 3)
    nodes = np.ascontiguousarray(nodes, dtype=np.float64)
    lib().or_deposit_smoothed(_p(srec), _p(occ), _p(bnd), len(occ), nk, ngrid, _p(nodes),
                              spacing, float(nbdt), ncol, _p(rho))
    return rho


def locate(x, colloc32, x0t, iwt, cells):
    """_locate_cell in Float32 (src/kernels.jl:1069-1081): (1-based left
    collocation point, weight of that point), (0, 0) outside."""
    g = np.ascontiguousarray(colloc32, dtype=np.float32)
    c = np.ascontiguousarray(cells, dtype=np.int32)
    w = C.c_float()
    cell = lib().or_locate(_p(g), len(g), np.float32(x0t), np.float32(iwt), _p(c), len(c),
                           np.float32(x), C.byref(w))
    return int(cell), np.float32(w.value)


def deposit_cic_raw(rec, x0f, hf, colloc32, x0t, iwt, cells):
    ng = len(colloc32)
    rho = np.zeros(ng ** 3)
    g = np.ascontiguousarray(colloc32, dtype=np.float32)
    c = np.ascontiguousarray(cells, dtype=np.int32)
    lib().or_deposit_cic(_p(rec), len(rec), x0f, hf, _p(g), ng, x0t, iwt, _p(c), len(c), _p(rho))
    return rho


def field(srec, occ, bnd, nk, ngrid, csol, ovl, grad, x0, h, spacing, nbdt, ncol, w,
          proj=None, law=0, radius=0.0, f32=False):
    n = len(srec)
    force = np.zeros(3 * n)
    red = np.zeros(4)
    cs = np.ascontiguousarray(csol, dtype=np.float64)
    pj = None if proj is None else np.ascontiguousarray(proj, dtype=np.float64)
    lib().or_field(_p(srec), _p(occ), _p(bnd), len(occ), nk, ngrid, _p(cs),
                   _p(np.ascontiguousarray(ovl)), _p(np.ascontiguousarray(grad)), x0, h,
                   spacing, float(nbdt), ncol, w, None if pj is None else _p(pj), law, radius,
                   _p(force), _p(red), 1 if f32 else 0)
    return force.reshape(-1, 3), red


def fallback(srec, outl, x0, h, knots_c, csolc, w, proj=None, ball=True, radius=5.0, escaped=0.0,
             force=None):
    n = len(srec)
    f = np.zeros(3 * n) if force is None else force.reshape(-1).copy()
    kc = np.ascontiguousarray(knots_c, dtype=np.float64)
    cs = np.ascontiguousarray(csolc, dtype=np.float64)
    nc = round(len(cs) ** (1 / 3))
    pj = None if proj is None else np.ascontiguousarray(proj, dtype=np.float64)
    ol = np.ascontiguousarray(outl, dtype=np.int32)
    lib().or_fallback(_p(srec), _p(ol), len(ol), x0, h, _p(kc), len(kc), _p(cs), nc, w,
                      None if pj is None else _p(pj), 1 if ball else 0, radius, escaped, _p(f))
    return f.reshape(-1, 3)


def fallback_points(pts, knots_c, csolc, w, proj=None, ball=True, radius=5.0, escaped=0.0):
    """The out-of-stencil force at FP64 points (a host-held cloud's own
    positions), src/accelerator.jl:705-730.  proj = (x, y, z, charge)."""
    p = np.ascontiguousarray(pts, dtype=np.float64)
    kc = np.ascontiguousarray(knots_c, dtype=np.float64)
    cs = np.ascontiguousarray(csolc, dtype=np.float64)
    nc = round(len(cs) ** (1 / 3))
    pj = None if proj is None else np.ascontiguousarray(proj, dtype=np.float64)
    f = np.zeros(3 * len(p))
    lib().or_fallback_points(_p(p), len(p), _p(kc), len(kc), _p(cs), nc, w,
                             None if pj is None else _p(pj), 1 if ball else 0, radius, escaped, _p(f))
    return f.reshape(-1, 3)


def pack_kd(pos, x0, h, nk):
    """_pack_kd_kernel! (src/kernels.jl:134-145): FP64 absolute positions ->
    PackedParticle{Float32} records, knode = round((p - x0)/h) + 1 (half to
    even) clamped to [1, nk], delta = Float32(p - (x0 + (k - 1) h)) with the
    subtraction in FP64; the previous half set equal to the current one."""
    p = np.asarray(pos, dtype=np.float64)
    k = np.clip(np.rint((p - x0) / h) + 1, 1, nk).astype(np.int32)
    d = (p - (x0 + (k.astype(np.float64) - 1) * h)).astype(np.float32)
    rec = np.zeros(len(p), dtype=REC)
    rec["knode"] = rec["pknode"] = k
    rec["delta"] = rec["pdelta"] = d
    return rec


def split_positions(x, x0, h):
    """_split (src/particles.jl:135-138): k = round((x - x0)/h) + 1 (half to
    even, unclamped), delta = Float32(x - (x0 + (k - 1) h))."""
    x = np.asarray(x, dtype=np.float64)
    k = np.rint((x - x0) / h) + 1
    return k.astype(np.int32), (x - (x0 + (k - 1) * h)).astype(np.float32)


def unsplit_positions(k, d, x0, h):
    """getindex of PackedPositions (src/particles.jl:128-132): x0 + (k - 1) h + T(delta)."""
    return (x0 + (k.astype(np.float64) - 1) * h) + d.astype(np.float64)


def half_step_back(positions, momenta, M, dt):
    """half_step_back (src/particles.jl:441-442): q .- (dt / 2M) .* p."""
    return np.asarray(positions, np.float64) - (dt / (2 * M)) * np.asarray(momenta, np.float64)


def full_step_back(q0, hp, f, M, dt, consistent=False):
    """full_step_back (src/particles.jl:466-469): .-q .+ 2 .* h .+ coef .* f,
    coef = dt/M (the thesis's) or dt^2/4M."""
    coef = dt ** 2 / (4 * M) if consistent else dt / M
    return (-q0 + 2 * hp) + coef * f


def prime_records(srec, force, x0, h, M, dt, consistent=False):
    """The loop of prime_leapfrog! (src/simulation.jl:266-273) over records
    in the sort's order holding q(-dt/2) (current) and q(0) (previous), with
    the forces at q(-dt/2) in the same order (widened to FP64)."""
    q0 = unsplit_positions(srec["pknode"], srec["pdelta"], x0, h)
    hp = unsplit_positions(srec["knode"], srec["delta"], x0, h)
    back = full_step_back(q0, hp, np.asarray(force, np.float32).astype(np.float64), M, dt, consistent)
    out = srec.copy()
    out["knode"], out["delta"] = split_positions(q0, x0, h)
    out["pknode"], out["pdelta"] = split_positions(back, x0, h)
    return out


def spline_potential(knots, csol, pts):
    kc = np.ascontiguousarray(knots, dtype=np.float64)
    cs = np.ascontiguousarray(csol, dtype=np.float64)
    nc = round(len(cs) ** (1 / 3))
    p = np.ascontiguousarray(pts, dtype=np.float64)
    out = np.zeros(len(p))
    lib().or_spline_potential(_p(kc), len(kc), _p(cs), nc, _p(p), len(p), _p(out))
    return out


def interaction_energy(rec, x0, h, knots_f, csolf, ovl64, spacing, nbdt, knots_c, csolc, w,
                       enclosed=0.0, per_particle=False):
    """interaction_energy (src/energy.jl:19-43) in FP64: the total, and each
    particle's phi if `per_particle`."""
    kf = np.ascontiguousarray(knots_f, dtype=np.float64)
    kc = np.ascontiguousarray(knots_c, dtype=np.float64)
    cf = np.ascontiguousarray(csolf, dtype=np.float64)
    cc = np.ascontiguousarray(csolc, dtype=np.float64)
    ov = np.ascontiguousarray(ovl64, dtype=np.float64)
    nf, nc = round(len(cf) ** (1 / 3)), round(len(cc) ** (1 / 3))
    r = np.ascontiguousarray(rec)
    phi = np.zeros(len(r)) if per_particle else None
    tot = lib().or_interaction_energy(_p(r), len(r), x0, h, _p(kf), len(kf), _p(cf), nf, _p(ov),
                                      spacing, int(nbdt), int(nbdt) + 1, _p(kc), len(kc), _p(cc), nc,
                                      w, enclosed, None if phi is None else _p(phi))
    return (tot, phi) if per_particle else tot


def verlet_f32(rec, force, dt, mass, h, x0, rcmax):
    """In place on `rec`; returns (kinetic, escaped, angular)."""
    f = np.ascontiguousarray(force, dtype=np.float32).reshape(-1)
    d = np.zeros(5)
    lib().or_verlet_f32(_p(rec), _p(f), len(rec), np.float32(dt * dt / mass), np.float32(h),
                        np.float32(x0), np.float32(mass / (2 * dt)), np.float32(rcmax * rcmax),
                        np.float32(1 / (2 * mass)), _p(d))
    return d


def enclosed(rec, x0, h, center, r):
    c = np.ascontiguousarray(center, dtype=np.float64)
    return int(lib().or_enclosed(_p(rec), len(rec), x0, h, _p(c), r))


# ---------------------------------------------------------------------------
# grid chain, FP64 numpy
# ---------------------------------------------------------------------------
def rotate(A: np.ndarray, src: np.ndarray, n: int) -> np.ndarray:
    """mul!(reshape(dest, m, n), transpose(reshape(src, n, m)), transpose(A))
    (src/tensorsolver.jl:217-223)."""
    m = src.size // n
    return (A @ src.reshape(m, n).T).ravel()


def apply_all(A: np.ndarray, src: np.ndarray, n: int) -> np.ndarray:
    out = src
    for _ in range(3):
        out = rotate(A, out, n)
    return out


def total_charge(rho, m0):
    n = len(m0)
    r = rho.reshape(n, n, n)          # [k][j][i]
    s = r @ m0                          # sum over i
    return float(np.einsum("kj,j,k->", s, m0, m0))


def moments(rho, dual):
    n = len(dual[0])
    p0, p1, p2 = dual
    r = rho.reshape(n, n, n)
    s0, s1, s2 = r @ p0, r @ p1, r @ p2          # [k][j]
    def c(s, a, b):
        return float(np.einsum("kj,j,k->", s, a, b))
    return (c(s0, p0, p0), c(s1, p0, p0), c(s0, p1, p0), c(s0, p0, p1), c(s2, p0, p0),
            c(s0, p2, p0), c(s0, p0, p2), c(s1, p1, p0), c(s1, p0, p1), c(s0, p1, p1))


def multipole(mom):
    """_multipole (src/poisson.jl:168-189) -> [q, c(3), quad(6)]."""
    q, d100, d010, d001, m200, m020, m002, mxy, mxz, myz = mom
    bx, by, bz = (0.0, 0.0, 0.0) if q == 0 else (d100 / q, d010 / q, d001 / q)
    quad = (2 * m200 - m020 - m002, 2 * m020 - m200 - m002, 2 * m002 - m200 - m020,
            3 * mxy, 3 * mxz, 3 * myz)
    b2 = bx * bx + by * by + bz * bz
    shift = (q * (3 * bx * bx - b2), q * (3 * by * by - b2), q * (3 * bz * bz - b2),
             3 * q * bx * by, 3 * q * bx * bz, 3 * q * by * bz)
    return np.array([q, bx, by, bz] + [a - b for a, b in zip(quad, shift)])


def face_mask(n: int) -> np.ndarray:
    m = np.zeros((n, n, n), bool)
    m[0, :, :] = m[-1, :, :] = m[:, 0, :] = m[:, -1, :] = m[:, :, 0] = m[:, :, -1] = True
    return m.ravel()                   # symmetric, so the flat order does not matter


def face_points(colloc):
    n = len(colloc)
    mask = face_mask(n)
    idx = np.nonzero(mask)[0]
    i, j, k = idx % n, (idx // n) % n, idx // (n * n)
    return idx, np.stack([colloc[i], colloc[j], colloc[k]], axis=1)


def multipole_potential(mp, pts):
    """potential(mp, x, y, z) (src/poisson.jl:198-206)."""
    px, py, pz = (pts[:, d] - mp[1 + d] for d in range(3))
    r2 = px * px + py * py + pz * pz
    invr = 1 / np.sqrt(r2)
    qxx, qyy, qzz, qxy, qxz, qyz = mp[4:]
    quad = qxx * px ** 2 + qyy * py ** 2 + qzz * pz ** 2 + 2 * (qxy * px * py + qxz * px * pz + qyz * py * pz)
    return mp[0] * invr + quad * invr ** 5 / 2


def poisson_rhs(rho, phi, D, n):
    """_poisson_rhs_kernel! (src/poisson.jl:275-294): interior rhs, (n-2)^3."""
    r = rho.reshape(n, n, n)[1:-1, 1:-1, 1:-1]        # [k][j][i]
    p = phi.reshape(n, n, n)
    c = -4 * math.pi
    dl, dr = D[1:-1, 0], D[1:-1, n - 1]
    out = c * r
    out = out - dl[None, None, :] * p[1:-1, 1:-1, 0][:, :, None] - dr[None, None, :] * p[1:-1, 1:-1, n - 1][:, :, None]
    out = out - dl[None, :, None] * p[1:-1, 0, 1:-1][:, None, :] - dr[None, :, None] * p[1:-1, n - 1, 1:-1][:, None, :]
    out = out - dl[:, None, None] * p[0, 1:-1, 1:-1][None, :, :] - dr[:, None, None] * p[n - 1, 1:-1, 1:-1][None, :, :]
    return out.ravel()


def solve_interior(phi, rho, mesh):
    """solve_interior! (src/poisson.jl:366-372) with _solve!
    (src/tensorsolver.jl:267-296)."""
    n, ni = mesh.n, mesh.ni
    rhs = poisson_rhs(rho, phi, mesh.laplacian, n)
    w1 = apply_all(mesh.op.Minv, rhs, ni) * mesh.invlsum
    x = apply_all(mesh.op.M, w1, ni)
    out = phi.reshape(n, n, n).copy()
    out[1:-1, 1:-1, 1:-1] = x.reshape(ni, ni, ni)
    return out.ravel()


def poisson_single(rho, mesh):
    """poisson!(phi, rho, mesh) with multipole faces (src/poisson.jl:351-354)."""
    n = mesh.n
    mp = multipole(moments(rho, mesh.dual_moments))
    phi = np.zeros(n ** 3)
    idx, pts = face_points(mesh.axis.colloc)
    phi[idx] = multipole_potential(mp, pts)
    return solve_interior(phi, rho, mesh), mp


def spline_coefficients(phi, mesh):
    return apply_all(mesh.coll.Sinv, phi, mesh.n)


def poisson_nested(rho_f, rho_c, fine, coarse):
    """poisson!(phis, rhos, nested) (src/poisson.jl:427-436)."""
    phi_c, mp = poisson_single(rho_c, coarse)
    csol_c = spline_coefficients(phi_c, coarse)
    phi_f = np.zeros(fine.n ** 3)
    idx, pts = face_points(fine.axis.colloc)
    phi_f[idx] = spline_potential(coarse.axis.knots, csol_c, pts)
    phi_f = solve_interior(phi_f, rho_f, fine)
    return phi_f, phi_c, mp


def effective_extra(rho, mesh, jel):
    """V_xc(rho) + V_jel(r) at the collocation points (src/meanfield.jl:116-123)."""
    from importlib import import_module
    n = mesh.n
    g = mesh.axis.colloc
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    r = np.sqrt(g[i] ** 2 + g[j] ** 2 + g[k] ** 2).ravel()
    c = np.cbrt(rho)
    vxc = -np.cbrt(3 / math.pi) * c + (-0.0333) * np.log1p(11.4 * np.cbrt(4 * math.pi / 3) * c)
    R, N = jel.radius, jel.nions
    vj = np.where(r < R, -N * (3 - (r / R) ** 2) / (2 * R), -N / np.where(r == 0, 1.0, r))
    return vxc + vj


def effective_potential(csol, rho, mesh, jel):
    return csol + spline_coefficients(effective_extra(rho, mesh, jel), mesh)


# ---------------------------------------------------------------------------
# one resident step (src/simulation.jl:472-531 + src/accelerator.jl:580-605)
# ---------------------------------------------------------------------------
class OracleStep:
    """FP64 restatement of one step on the device path, for a cloud held as
    PackedParticle{Float32} records (the device's own representation)."""

    def __init__(self, fine, coarse, sm, jel, weight, mass, dt, rcmax=100.0):
        self.fine, self.coarse, self.sm, self.jel = fine, coarse, sm, jel
        self.w, self.mass, self.dt, self.rcmax = weight, mass, dt, rcmax
        k = fine.axis.knots
        self.nk = len(k)
        self.x0 = float(k[0])
        self.h = float((k[-1] - k[0]) / (len(k) - 1))
        self.spacing = np.float32(sm.spacing)
        self.lo = np.float32(k[1] + sm.spacing / 2)
        self.hi = np.float32(k[-2] - sm.spacing / 2)

    def sort(self, rec):
        perm, occ, bnd = cellsort(rec, self.nk)
        return rec[perm].copy(), occ, bnd

    def deposit_fine(self, srec, occ, bnd):
        sm = self.sm
        raw = deposit_smoothed_raw(srec, occ, bnd, self.nk, self.fine.n, sm.nodes, self.spacing,
                                   sm.nbdt, sm.nbdt + 1)
        _, nout = columns(srec, np.float32(self.x0), np.float32(self.h), self.lo, self.hi,
                          self.spacing, sm.nbdt, sm.nbdt + 1)
        q = total_charge(raw, self.fine.dual_moments[0])
        return raw * ((len(srec) - nout) * self.w / q), nout

    def deposit_coarse(self, srec):
        c = self.coarse
        raw = deposit_cic_raw(srec, np.float32(self.x0), np.float32(self.h),
                              c.axis.colloc.astype(np.float32), np.float32(c.locator.x0),
                              np.float32(c.locator.invwidth), c.locator.cell)
        wl = c.dual_len
        vol = (wl[None, None, :] * wl[None, :, None]) * wl[:, None, None]
        return raw * (self.w / vol.ravel())

    def forces(self, srec, occ, bnd, csol_f, csol_c, projectile=None, mode="device"):
        sm = self.sm
        law = 1 if mode == "ball" else 0
        proj = None
        radius = 0.0
        if projectile is not None:
            from_ball = hasattr(projectile.softening, "radius")
            radius = projectile.softening.radius if from_ball else projectile.softening.sigma
            proj = np.array([*projectile.position, -self.w * projectile.charge, projectile.scale])
        f, red = field(srec, occ, bnd, self.nk, self.fine.n, csol_f, sm.overlap, sm.gradient,
                       self.x0, self.h, self.spacing, sm.nbdt, sm.nbdt + 1, self.w, proj, law, radius)
        outl = outside(srec, self.fine.n)
        pj = None if projectile is None else np.array([*projectile.position, projectile.charge])
        ball = projectile is not None and hasattr(projectile.softening, "radius")
        f = fallback(srec, outl, self.x0, self.h, self.coarse.axis.knots, csol_c, self.w, pj,
                     ball, radius, 0.0, force=f)
        return f, red, len(outl)

    def energy(self, rec, csol_f, csol_c, enclosed=0.0):
        """interaction_energy of the cloud on the given coefficients."""
        sm = self.sm
        return interaction_energy(rec, self.x0, self.h, self.fine.axis.knots, csol_f, sm.overlap,
                                  float(sm.spacing), sm.nbdt, self.coarse.axis.knots, csol_c, self.w,
                                  enclosed)

    def update_forces(self, rec, projectile=None, mode="device", energy=False):
        srec, occ, bnd = self.sort(rec)
        rho_f, nout = self.deposit_fine(srec, occ, bnd)
        rho_c = self.deposit_coarse(srec)
        phi_f, phi_c, mp = poisson_nested(rho_f, rho_c, self.fine, self.coarse)
        csol_f = spline_coefficients(phi_f, self.fine)
        csol_c = spline_coefficients(phi_c, self.coarse)
        hartree = self.energy(srec, csol_f, csol_c) / 2 if energy else None
        csol_f = effective_potential(csol_f, rho_f, self.fine, self.jel)
        csol_c = effective_potential(csol_c, rho_c, self.coarse, self.jel)
        f, red, outcount = self.forces(srec, occ, bnd, csol_f, csol_c, projectile, mode)
        return dict(srec=srec, occ=occ, bnd=bnd, rho_f=rho_f, rho_c=rho_c, phi_f=phi_f,
                    phi_c=phi_c, csol_f=csol_f, csol_c=csol_c, force=f, red=red, nout=nout,
                    outcount=outcount, mp=mp, hartree=hartree)

    def verlet(self, srec, force):
        rec = srec.copy()
        d = verlet_f32(rec, force.astype(np.float32), self.dt, self.mass, self.h, self.x0, self.rcmax)
        return rec, d
