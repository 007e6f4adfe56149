/*
 * vlasov_oracle.c -- CPU restatement of the reference's per-step particle
 * loops, used ONLY as the checker of the HIP path (tests, smoke test) and as
 * the timed CPU baseline of bench.py.  Nothing in the product links it.
 *
 * Every function cites the reference code it restates.  Two kinds of
 * arithmetic appear:
 *   - the discrete decisions of the reference's device path (cell keys,
 *     stable order, table columns, stencil-fit tests, coarse cell location)
 *     are evaluated exactly as the Float32 device kernels evaluate them, so
 *     that they can be compared bit for bit;
 *   - the accumulations (deposits, contractions, reductions) run in FP64,
 *     the reference's CPU precision, and are compared within tolerance.
 * `or_field(..., f32=1)` and `or_verlet_f32` are additionally exact Float32
 * emulations (explicit fmaf, no contraction: build with -ffp-contract=off).
 *
 * The ORACLE_PLANT_* macros are never defined in the oracle the tests check
 * against.  A test builds a separate copy with one of them, a plausible
 * mistake in a placement convention, to show that its known-answer check
 * fails on that mistake.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
  int32_t k[3];
  float d[3];
  int32_t pk[3];
  float pd[3];
} rec_t;

int or_num_threads(void);

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* clamped 0-based cell key: src/kernels.jl:902-908, src/sorting.jl:134-151 */
static inline int64_t key_of(const rec_t *r, int nk) {
  int kx = clampi(r->k[0], 1, nk), ky = clampi(r->k[1], 1, nk), kz = clampi(r->k[2], 1, nk);
  return (kx - 1) + (int64_t)nk * ((ky - 1) + (int64_t)nk * (kz - 1));
}

/* Stable counting sort, cellsort! with one chunk (src/sorting.jl:163-204):
 * perm[s] = index of the s-th sorted particle, occupied = increasing
 * 0-based keys, bounds = exclusive prefix (nocc + 1 entries). */
int64_t or_cellsort(const rec_t *parts, int64_t n, int nk, int32_t *perm, int32_t *occupied,
                    int32_t *bounds) {
  int64_t ncell = (int64_t)nk * nk * nk;
  int32_t *cnt = (int32_t *)calloc((size_t)ncell, sizeof(int32_t));
  int32_t *cur = (int32_t *)calloc((size_t)ncell, sizeof(int32_t));
  for (int64_t i = 0; i < n; ++i) cnt[key_of(parts + i, nk)]++;
  int64_t nocc = 0;
  int32_t acc = 0;
  bounds[0] = 0;
  for (int64_t c = 0; c < ncell; ++c) {
    if (cnt[c] == 0) continue;
    occupied[nocc] = (int32_t)c;
    cur[c] = acc;
    acc += cnt[c];
    bounds[++nocc] = acc;
  }
  for (int64_t i = 0; i < n; ++i) {
    int64_t c = key_of(parts + i, nk);
    perm[cur[c]++] = (int32_t)i;
  }
  free(cnt);
  free(cur);
  return nocc;
}

/* particles whose 10^3 stencil leaves the fine grid, from the unclamped
 * knode (src/kernels.jl:1041-1058); increasing slot order */
int64_t or_outside(const rec_t *parts, int64_t n, int ngrid, int32_t *list) {
  int64_t m = 0;
  for (int64_t i = 0; i < n; ++i) {
    int bx = 2 * parts[i].k[0] - 5, by = 2 * parts[i].k[1] - 5, bz = 2 * parts[i].k[2] - 5;
    int ok = bx >= 1 && by >= 1 && bz >= 1 && bx + 9 <= ngrid && by + 9 <= ngrid && bz + 9 <= ngrid;
    if (!ok) list[m++] = (int32_t)i;
  }
  return m;
}

/* table column, Float32 (src/kernels.jl:770-772, src/fields.jl:241-245) */
static inline int col_f32(float d, float spacing, float nbdt, int ncol) {
  float half = spacing * 0.5f;
  float t = (d + half) / spacing * nbdt + 0.5f;
  int c = (int)floorf(t) + 1;
#ifdef ORACLE_PLANT_COLUMN_SLIP
  c += 1; /* planted mistake: one table column off */
#endif
  return clampi(c, 1, ncol);
}

/* _columns_kernel! (src/kernels.jl:754-779), Float32 */
int64_t or_columns(const rec_t *parts, int64_t n, float x0, float h, float lo, float hi,
                   float spacing, float nbdt, int ncol, int32_t *cols) {
  int64_t nout = 0;
  for (int64_t i = 0; i < n; ++i) {
    int inside = 1;
    for (int d = 0; d < 3; ++d) {
      float p = x0 + (float)(parts[i].k[d] - 1) * h + parts[i].d[d];
      inside = inside && (lo <= p) && (p <= hi);
    }
    if (inside) {
      for (int d = 0; d < 3; ++d) cols[3 * i + d] = col_f32(parts[i].d[d], spacing, nbdt, ncol);
    } else {
      cols[3 * i] = cols[3 * i + 1] = cols[3 * i + 2] = 1;
      ++nout;
    }
  }
  return nout;
}

/* stencil corner (1-based) and whether the 8^3 deposit stencil fits
 * (_cell_base, src/kernels.jl:106-117) */
static inline int cell_fits8(int64_t key, int nk, int ngrid, int *b) {
  int kx = (int)(key % nk) + 1, ky = (int)((key / nk) % nk) + 1, kz = (int)(key / ((int64_t)nk * nk)) + 1;
  b[0] = 2 * kx - 5;
  b[1] = 2 * ky - 5;
  b[2] = 2 * kz - 5;
#ifdef ORACLE_PLANT_STENCIL_SHIFT
  b[0] += 1; /* planted mistake: the x stencil one collocation point right */
#endif
  return b[0] >= 0 && b[1] >= 0 && b[2] >= 0 && b[0] + 8 <= ngrid && b[1] + 8 <= ngrid &&
         b[2] + 8 <= ngrid;
}

/* sorted smoothed deposit, device semantics (src/kernels.jl:601-726), sums in
 * FP64; rho is the raw deposit (n^3 doubles, i fastest), not renormalised */
/* Threads take contiguous ranges of occupied cells, each into its own FP64
 * grid; the grids are summed in thread order (so the sum depends on the
 * thread count only through FP64 rounding). */
static double **thread_grids(int nt, int64_t npts) {
  double **g = (double **)calloc((size_t)nt, sizeof(double *));
  for (int t = 0; t < nt; ++t) g[t] = (double *)calloc((size_t)npts, sizeof(double));
  return g;
}

static void sum_thread_grids(double **g, int nt, int64_t npts, double *out) {
#pragma omp parallel for
  for (int64_t i = 0; i < npts; ++i) {
    double s = 0.0;
    for (int t = 0; t < nt; ++t) s += g[t][i];
    out[i] = s;
  }
  for (int t = 0; t < nt; ++t) free(g[t]);
  free(g);
}

void or_deposit_smoothed(const rec_t *parts, const int32_t *occupied, const int32_t *bounds,
                         int64_t nocc, int nk, int ngrid, const double *nodes, float spacing,
                         float nbdt, int ncol, double *rho) {
  int64_t nn = ngrid, npts = nn * nn * nn;
  int nt = or_num_threads();
  double **tg = thread_grids(nt, npts);
#pragma omp parallel num_threads(nt)
  {
#ifdef _OPENMP
    const int t = omp_get_thread_num();
#else
    const int t = 0;
#endif
    double *dst = tg[t];
    const int64_t lo = nocc * t / nt, hi = nocc * (t + 1) / nt;
    for (int64_t g = lo; g < hi; ++g) {
      int b[3];
      if (!cell_fits8(occupied[g], nk, ngrid, b)) continue;
      for (int32_t s = bounds[g]; s < bounds[g + 1]; ++s) {
        const double *gx = nodes + 8 * (col_f32(parts[s].d[0], spacing, nbdt, ncol) - 1);
        const double *gy = nodes + 8 * (col_f32(parts[s].d[1], spacing, nbdt, ncol) - 1);
        const double *gz = nodes + 8 * (col_f32(parts[s].d[2], spacing, nbdt, ncol) - 1);
        for (int kk = 0; kk < 8; ++kk)
          for (int jj = 0; jj < 8; ++jj) {
            double w = gy[jj] * gz[kk];
            double *row = dst + b[0] + nn * ((b[1] + jj) + nn * (b[2] + kk));
            for (int ii = 0; ii < 8; ++ii) row[ii] += gx[ii] * w;
          }
      }
    }
  }
  sum_thread_grids(tg, nt, npts, rho);
}

/* _locate_cell in Float32 (src/kernels.jl:1069-1081): 1-based cell, 0 outside */
static inline int locate_f32(const float *gt, int ng, float x0t, float iwt, const int32_t *cells,
                             int nbin, float x, float *w) {
  if (x < gt[0] || x > gt[ng - 1]) {
    *w = 0.f;
    return 0;
  }
  int b = (int)floorf((x - x0t) * iwt) + 1;
  if (b > nbin) b = nbin;
  int c = cells[b - 1];
  while (c < ng - 1 && gt[c] < x) ++c;
  *w = (gt[c] - x) / (gt[c] - gt[c - 1]);
  return c;
}

/* locate through the table (src/kernels.jl:1069-1081) for one position:
 * 1-based left collocation point (0 outside) and the weight of that point */
int or_locate(const float *gt, int ng, float x0t, float iwt, const int32_t *cells, int nbin,
              float x, float *w) {
  return locate_f32(gt, ng, x0t, iwt, cells, nbin, x, w);
}

/* coarse cloud-in-cell of every particle (src/kernels.jl:1207-1273), cell and
 * weights in Float32 as the device computes them, products and sums in FP64.
 * rho raw (not divided by the dual volumes). */
void or_deposit_cic(const rec_t *parts, int64_t n, float x0f, float hf, const float *gt, int ng,
                    float x0t, float iwt, const int32_t *cells, int nbin, double *out) {
  int64_t nn = ng, npts = nn * nn * nn;
  int nt = or_num_threads();
  double **tg = thread_grids(nt, npts);
#pragma omp parallel num_threads(nt)
  {
#ifdef _OPENMP
  const int t = omp_get_thread_num();
#else
  const int t = 0;
#endif
  double *rho = tg[t];
  for (int64_t s = n * t / nt; s < n * (t + 1) / nt; ++s) {
    int ix[3];
    float a[3];
    int in = 1;
    for (int d = 0; d < 3; ++d) {
      float p = x0f + (float)(parts[s].k[d] - 1) * hf + parts[s].d[d];
      ix[d] = locate_f32(gt, ng, x0t, iwt, cells, nbin, p, &a[d]);
      in = in && ix[d] > 0;
    }
    if (!in) continue;
    double ax = a[0], ay = a[1], az = a[2];
    double bx = 1.0 - ax, by = 1.0 - ay, bz = 1.0 - az;
#ifdef ORACLE_PLANT_CIC_SWAP
    { double t = ax; ax = bx; bx = t; } /* planted mistake: x weights swapped */
#endif
    int64_t i = ix[0] - 1, j = ix[1] - 1, k = ix[2] - 1;
    rho[i + nn * (j + nn * k)] += ax * ay * az;
    rho[i + nn * (j + nn * (k + 1))] += ax * ay * bz;
    rho[i + nn * ((j + 1) + nn * k)] += ax * by * az;
    rho[i + nn * ((j + 1) + nn * (k + 1))] += ax * by * bz;
    rho[(i + 1) + nn * (j + nn * k)] += bx * ay * az;
    rho[(i + 1) + nn * (j + nn * (k + 1))] += bx * ay * bz;
    rho[(i + 1) + nn * ((j + 1) + nn * k)] += bx * by * az;
    rho[(i + 1) + nn * ((j + 1) + nn * (k + 1))] += bx * by * bz;
  }
  }
  sum_thread_grids(tg, nt, npts, out);
}

/* --- projectile laws (src/projectile.jl:79-108) ----------------------------- */
static const double GS[7] = {1.0 / 3.0, -1.0 / 10.0, 1.0 / 56.0, -1.0 / 432.0,
                             1.0 / 4224.0, -1.0 / 49920.0, 1.0 / 685440.0};
static const double SQ2PI = 0.79788456080286535588;

static double gaussian_fk(double r2, double s) {
  double u2 = r2 / (s * s);
  if (u2 <= 0.25) {
    double p = 0.0;
    for (int k = 6; k >= 0; --k) p = GS[k] + u2 * p;
    return SQ2PI * p / (s * s * s);
  }
  double r = sqrt(r2), x = r / (sqrt(2.0) * s);
  return (erf(x) - SQ2PI * (r / s) * exp(-u2 / 2.0)) / (r2 * r);
}

static double ball_fk(double r2, double R) { return r2 > R * R ? 1.0 / (r2 * sqrt(r2)) : 1.0 / (R * R * R); }

/* --- smoothed field (src/kernels.jl:339-470; src/fields.jl:305-326) -------- */
/* force: N x 3 (slot order), NaN for particles of cells whose stencil does
 * not fit.  proj = {x, y, z, coef = -w q, sigma} or NULL.  law 0: the device
 * kernel's Gaussian at the cell-rebuilt position; law 1: ball of `radius` at
 * the particle's own position.  f32 != 0 evaluates the contraction and the
 * force exactly as the Float32 kernel does (projectile excluded). */
void or_field(const rec_t *parts, const int32_t *occupied, const int32_t *bounds, int64_t nocc,
              int nk, int ngrid, const double *csol, const double *ovl, const double *grad,
              double x0, double h, float spacing, float nbdt, int ncol, double w,
              const double *proj, int law, double radius, double *force, double *red, int f32) {
  int64_t nn = ngrid;
  double sx = 0, sy = 0, sz = 0, se = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : sx, sy, sz, se)
  for (int64_t g = 0; g < nocc; ++g) {
    int64_t key = occupied[g];
    int kx = (int)(key % nk) + 1, ky = (int)((key / nk) % nk) + 1, kz = (int)(key / ((int64_t)nk * nk)) + 1;
    int bx = 2 * kx - 5, by = 2 * ky - 5, bz = 2 * kz - 5;
    int ok = bx >= 1 && by >= 1 && bz >= 1 && bx + 9 <= ngrid && by + 9 <= ngrid && bz + 9 <= ngrid;
    for (int32_t s = bounds[g]; s < bounds[g + 1]; ++s) {
      double f[3] = {NAN, NAN, NAN};
      if (ok) {
        int cx = col_f32(parts[s].d[0], spacing, nbdt, ncol) - 1;
        int cy = col_f32(parts[s].d[1], spacing, nbdt, ncol) - 1;
        int cz = col_f32(parts[s].d[2], spacing, nbdt, ncol) - 1;
        if (f32) {
          float fx = 0.f, fy = 0.f, fz = 0.f;
          for (int kk = 0; kk < 10; ++kk) {
            float oz = (float)ovl[cz * 10 + kk], gz = (float)grad[cz * 10 + kk];
            for (int jj = 0; jj < 10; ++jj) {
              float oy = (float)ovl[cy * 10 + jj], gy = (float)grad[cy * 10 + jj];
              float dxp = 0.f, val = 0.f;
              const double *row = csol + (bx - 1) + nn * ((by - 1 + jj) + nn * (bz - 1 + kk));
              for (int ii = 0; ii < 10; ++ii) {
                float c = (float)row[ii];
                dxp = fmaf(c, (float)grad[cx * 10 + ii], dxp);
                val = fmaf(c, (float)ovl[cx * 10 + ii], val);
              }
              fx = fmaf(oy * oz, dxp, fx);
              fy = fmaf(gy * oz, val, fy);
              fz = fmaf(oy * gz, val, fz);
            }
          }
          float wf = (float)w;
          f[0] = -wf * fx;
          f[1] = -wf * fy;
          f[2] = -wf * fz;
        } else {
          double fx = 0, fy = 0, fz = 0;
          for (int kk = 0; kk < 10; ++kk) {
            double oz = ovl[cz * 10 + kk], gz = grad[cz * 10 + kk];
            for (int jj = 0; jj < 10; ++jj) {
              double oy = ovl[cy * 10 + jj], gy = grad[cy * 10 + jj];
              double dxp = 0, val = 0;
              const double *row = csol + (bx - 1) + nn * ((by - 1 + jj) + nn * (bz - 1 + kk));
              for (int ii = 0; ii < 10; ++ii) {
                dxp = fma(row[ii], grad[cx * 10 + ii], dxp);
                val = fma(row[ii], ovl[cx * 10 + ii], val);
              }
              fx = fma(oy * oz, dxp, fx);
              fy = fma(gy * oz, val, fy);
              fz = fma(oy * gz, val, fz);
            }
          }
          f[0] = -w * fx;
          f[1] = -w * fy;
          f[2] = -w * fz;
        }
      }
      if (proj && proj[3] != 0.0) {
        double p[3];
        for (int d = 0; d < 3; ++d) {
          int kd = law ? parts[s].k[d] : (d == 0 ? kx : (d == 1 ? ky : kz));
          p[d] = x0 + (double)(kd - 1) * h + (double)parts[s].d[d];
        }
        double dx = proj[0] - p[0], dy = proj[1] - p[1], dz = proj[2] - p[2];
        double d2 = dx * dx + dy * dy + dz * dz;
        double mm = proj[3] * (law ? ball_fk(d2, radius) : gaussian_fk(d2, proj[4]));
        double f2[3] = {mm * dx, mm * dy, mm * dz};
        if (ok)
          for (int d = 0; d < 3; ++d) f[d] -= f2[d];
        double r = sqrt(d2);
        sx += f2[0];
        sy += f2[1];
        sz += f2[2];
        if (law) {
          double R = radius;
          se += r < R ? -(3.0 - (r / R) * (r / R)) / (2.0 * R) : -1.0 / r;
        } else {
          double sg = proj[4];
          se += r < 1.0e-4 * sg ? SQ2PI / sg : erf(r / sqrt(2.0) / sg) / r;
        }
      }
      force[3 * s] = f[0];
      force[3 * s + 1] = f[1];
      force[3 * s + 2] = f[2];
    }
  }
  if (red) {
    red[0] = sx;
    red[1] = sy;
    red[2] = sz;
    red[3] = se;
  }
}

/* --- Hermite basis, FP64 (src/splines.jl:74-139) ---------------------------- */
static void hermite_vd(const double *g, int nk, int b, double x, double *v, double *dv) {
  int kn = b >> 1, kind = b & 1;
  double a, h, s;
  *v = 0.0;
  *dv = 0.0;
  if (kn > 0 && x <= g[kn]) {
    if (x < g[kn - 1]) return;
    h = g[kn] - g[kn - 1];
    a = (x - g[kn - 1]) / h;
    s = -1.0;
  } else if (kn < nk - 1 && x >= g[kn]) {
    if (x > g[kn + 1]) return;
    h = g[kn + 1] - g[kn];
    a = (g[kn + 1] - x) / h;
    s = 1.0;
  } else {
    return;
  }
  if (kind == 0) {
    *v = 3.0 * (a * a) - 2.0 * (a * a * a);
    *dv = -6.0 * (a - a * a) * s / h;
  } else {
    *v = h * ((a * a) - (a * a * a)) * s;
    *dv = -2.0 * a + 3.0 * (a * a);
  }
}

/* cell_index (src/fields.jl:235-238): 0-based interval or -1 */
static int cell_index(const double *g, int nk, double x) {
  if (x < g[0] || x > g[nk - 1]) return -1;
  int lo = 0, hi = nk;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (g[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  int c = lo - 1;
  return c < 0 ? 0 : c;
}

/* spline_potential at npts points (src/fields.jl:440-469); NaN outside */
void or_spline_potential(const double *knots, int nk, const double *csol, int nc, const double *pts,
                         int64_t npts, double *out) {
  int64_t n2 = nc;
#pragma omp parallel for
  for (int64_t t = 0; t < npts; ++t) {
    const double *p = pts + 3 * t;
    int cell[3], bad = 0;
    for (int d = 0; d < 3; ++d) {
      cell[d] = cell_index(knots, nk, p[d]);
      bad |= cell[d] < 0;
    }
    if (bad) {
      out[t] = NAN;
      continue;
    }
    double v[3][4], dv;
    for (int d = 0; d < 3; ++d)
      for (int a = 0; a < 4; ++a) hermite_vd(knots, nk, 2 * cell[d] + a, p[d], &v[d][a], &dv);
    double phi = 0.0;
    for (int ak = 0; ak < 4; ++ak) {
      int k = 2 * cell[2] + ak;
      for (int aj = 0; aj < 4; ++aj) {
        int j = 2 * cell[1] + aj;
        double w = v[1][aj] * v[2][ak];
        for (int ai = 0; ai < 4; ++ai) phi += csol[2 * cell[0] + ai + n2 * (j + n2 * k)] * v[0][ai] * w;
      }
    }
    out[t] = phi;
  }
}

/* nearest_knot (src/fields.jl:224-227), 0-based.  searchsortedlast by a
 * linear scan (the number of knots <= x), clamped to [1, nk-1], then the
 * closer end of that interval, the left one on a tie. */
static int nearest_knot(const double *g, int nk, double x) {
  int i = 0;
  while (i < nk && g[i] <= x) ++i;
  i = clampi(i, 1, nk - 1);
  return (g[i] - x) < (x - g[i - 1]) ? i : i - 1;
}

/* smoothed_potential (src/fields.jl:405-428): the potential convolved with
 * the particle's Gaussian, FP64, the reference's loop order and roundings
 * (s += c * ox, then phi += s * (oy * oz)). */
static double smoothed_potential(const double *kf, int nkf, const double *csol, int nf,
                                 const double *ovl, double spacing, int nbdt, int ncol,
                                 const double *p) {
  int64_t b[3];
  const double *o[3];
  for (int d = 0; d < 3; ++d) {
    int ci = nearest_knot(kf, nkf, p[d]);
    b[d] = 2 * (int64_t)(ci + 1) - 5 - 1;
    int col = (int)floor((p[d] - kf[ci] + spacing / 2) / spacing * (double)nbdt + 0.5) + 1;
    o[d] = ovl + 10 * (clampi(col, 1, ncol) - 1);
  }
  int64_t n = nf;
  double phi = 0.0;
  for (int kk = 0; kk < 10; ++kk) {
    for (int jj = 0; jj < 10; ++jj) {
      double w = o[1][jj] * o[2][kk];
      double s = 0.0;
      for (int ii = 0; ii < 10; ++ii) s += csol[b[0] + ii + n * ((b[1] + jj) + n * (b[2] + kk))] * o[0][ii];
      phi += s * w;
    }
  }
  return phi;
}

/* interaction_energy (src/energy.jl:19-43): sum_i w phi(r_i) with r from the
 * current (k, delta) as T = Float64 (src/particles.jl:128-132).  Inside
 * (knots[3], knots[end-2])^3 of the fine axis: smoothed_potential of the fine
 * csol; otherwise spline_potential of the coarse csol, or enclosed / |r|
 * beyond both grids.  phi_out (may be NULL) receives each particle's phi. */
double or_interaction_energy(const rec_t *parts, int64_t n, double x0, double h, const double *kf,
                             int nkf, const double *csolf, int nf, const double *ovl,
                             double spacing, int nbdt, int ncol, const double *kc, int nkc,
                             const double *csolc, int nc, double w, double enclosed,
                             double *phi_out) {
  const double lo = kf[2], hi = kf[nkf - 3];
  double total = 0.0;
#pragma omp parallel for reduction(+ : total)
  for (int64_t i = 0; i < n; ++i) {
    double p[3];
    int inside = 1;
    for (int d = 0; d < 3; ++d) {
      p[d] = x0 + (double)(parts[i].k[d] - 1) * h + (double)parts[i].d[d];
      inside = inside && lo < p[d] && p[d] < hi;
    }
    double phi;
    if (inside) {
      phi = smoothed_potential(kf, nkf, csolf, nf, ovl, spacing, nbdt, ncol, p);
    } else {
      or_spline_potential(kc, nkc, csolc, nc, p, 1, &phi);
      if (isnan(phi)) phi = enclosed / sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    }
    if (phi_out) phi_out[i] = phi;
    total += w * phi;
  }
  return total;
}

/* spline_field (src/fields.jl:358-393); returns 0 outside (with out = 0) */
static int spline_field(const double *knots, int nk, const double *csol, int nc, const double *p,
                        double *e) {
  int cell[3];
  for (int d = 0; d < 3; ++d) {
    cell[d] = cell_index(knots, nk, p[d]);
    if (cell[d] < 0) return 0;
  }
  double v[3][4], dv[3][4];
  for (int d = 0; d < 3; ++d)
    for (int a = 0; a < 4; ++a) hermite_vd(knots, nk, 2 * cell[d] + a, p[d], &v[d][a], &dv[d][a]);
  int64_t n2 = nc;
  double ex = 0, ey = 0, ez = 0;
  for (int ak = 0; ak < 4; ++ak) {
    int k = 2 * cell[2] + ak;
    for (int aj = 0; aj < 4; ++aj) {
      int j = 2 * cell[1] + aj;
      for (int ai = 0; ai < 4; ++ai) {
        int i = 2 * cell[0] + ai;
        double c = csol[i + n2 * (j + n2 * k)];
        ex -= c * dv[0][ai] * v[1][aj] * v[2][ak];
        ey -= c * v[0][ai] * dv[1][aj] * v[2][ak];
        ez -= c * v[0][ai] * v[1][aj] * dv[2][ak];
      }
    }
  }
  e[0] = ex;
  e[1] = ey;
  e[2] = ez;
  return 1;
}

/* the host fallback at given FP64 points (the positions of a host-held
 * cloud, src/accelerator.jl:705-730 with p = cloud.positions[i]): coarse
 * spline field, or the monopole of `escaped` beyond both grids, plus the
 * projectile's own law.  proj = {x, y, z, charge} or NULL; force n x 3. */
void or_fallback_points(const double *pts, int64_t n, const double *knots, int nkc,
                        const double *csolc, int nc, double w, const double *proj, int ball,
                        double radius, double escaped, double *force) {
#pragma omp parallel for
  for (int64_t u = 0; u < n; ++u) {
    const double *p = pts + 3 * u;
    double f[3], e[3];
    if (spline_field(knots, nkc, csolc, nc, p, e)) {
      for (int d = 0; d < 3; ++d) f[d] = w * e[d];
    } else {
      double r3 = pow(p[0] * p[0] + p[1] * p[1] + p[2] * p[2], 1.5);
      double cf = -(w * w) * escaped / r3;
      for (int d = 0; d < 3; ++d) f[d] = cf * p[d];
    }
    if (proj) {
      double dd[3] = {proj[0] - p[0], proj[1] - p[1], proj[2] - p[2]};
      double r2 = dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2];
      double m = -w * proj[3] * (ball ? ball_fk(r2, radius) : gaussian_fk(r2, radius));
      for (int d = 0; d < 3; ++d) f[d] = f[d] - m * dd[d];
    }
    for (int d = 0; d < 3; ++d) force[3 * u + d] = f[d];
  }
}

/* the host fallback of the device forces! (src/accelerator.jl:705-730):
 * coarse spline field, or the enclosed-charge monopole beyond both grids, plus
 * the projectile's own law.  proj = {x, y, z, charge}. */
void or_fallback(const rec_t *parts, const int32_t *list, int64_t nlist, double x0, double h,
                 const double *knots, int nkc, const double *csolc, int nc, double w,
                 const double *proj, int ball, double radius, double escaped, double *force) {
#pragma omp parallel for
  for (int64_t u = 0; u < nlist; ++u) {
    int32_t s = list[u];
    double p[3], f[3], e[3];
    for (int d = 0; d < 3; ++d) p[d] = x0 + (double)(parts[s].k[d] - 1) * h + (double)parts[s].d[d];
    if (spline_field(knots, nkc, csolc, nc, p, e)) {
      for (int d = 0; d < 3; ++d) f[d] = w * e[d];
    } else {
      double r3 = pow(p[0] * p[0] + p[1] * p[1] + p[2] * p[2], 1.5);
      double cf = -(w * w) * escaped / r3;
      for (int d = 0; d < 3; ++d) f[d] = cf * p[d];
    }
    if (proj) {
      double dd[3] = {proj[0] - p[0], proj[1] - p[1], proj[2] - p[2]};
      double r2 = dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2];
      double m = -w * proj[3] * (ball ? ball_fk(r2, radius) : gaussian_fk(r2, radius));
      for (int d = 0; d < 3; ++d) f[d] = f[d] - m * dd[d];
    }
    for (int d = 0; d < 3; ++d) force[3 * s + d] = f[d];
  }
}

/* --- packed Verlet, exact Float32 (src/kernels.jl:985-1025,
 * src/particles.jl:337-349); diagnostics summed in FP64 ---------------------- */
void or_verlet_f32(rec_t *parts, const float *force, int64_t n, float a, float h, float x0,
                   float pfac, float r2max, float inv2M, double *diag) {
  double ek = 0, eo = 0, lx = 0, ly = 0, lz = 0;
  for (int64_t i = 0; i < n; ++i) {
    rec_t q = parts[i];
    int kn[3];
    float dn[3], gp[3];
    for (int d = 0; d < 3; ++d) {
      int k = 2 * q.k[d] - q.pk[d];
      float dd = 2.0f * q.d[d] - q.pd[d] + a * force[3 * i + d];
      int m = (int)rintf(dd / h);
      k += m;
      dd -= (float)m * h;
      kn[d] = k;
      dn[d] = dd;
      gp[d] = (float)(k - q.pk[d]) * h + (dd - q.pd[d]);
    }
    float px = gp[0] * pfac, py = gp[1] * pfac, pz = gp[2] * pfac;
    float e = (px * px + py * py + pz * pz) * inv2M;
    ek += e;
    float qx = x0 + (float)(q.k[0] - 1) * h + q.d[0];
    float qy = x0 + (float)(q.k[1] - 1) * h + q.d[1];
    float qz = x0 + (float)(q.k[2] - 1) * h + q.d[2];
    if (qx * qx + qy * qy + qz * qz > r2max) eo += e;
    lx += qy * pz - qz * py;
    ly += qz * px - qx * pz;
    lz += qx * py - qy * px;
    for (int d = 0; d < 3; ++d) {
      parts[i].pk[d] = q.k[d];
      parts[i].pd[d] = q.d[d];
      parts[i].k[d] = kn[d];
      parts[i].d[d] = dn[d];
    }
  }
  diag[0] = ek;
  diag[1] = eo;
  diag[2] = lx;
  diag[3] = ly;
  diag[4] = lz;
}

/* enclosed count around a point, FP64 positions (src/projectile.jl:243-247) */
int64_t or_enclosed(const rec_t *parts, int64_t n, double x0, double h, const double *c, double r) {
  int64_t m = 0;
  for (int64_t i = 0; i < n; ++i) {
    double d2 = 0;
    double dd[3];
    for (int d = 0; d < 3; ++d) dd[d] = (x0 + (double)(parts[i].k[d] - 1) * h + (double)parts[i].d[d]) - c[d];
    d2 = dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2];
    if (d2 < r * r) ++m;
  }
  return m;
}

int or_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
