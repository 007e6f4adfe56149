// Charge deposition: the smoothed 8^3 deposit on the fine grid and the
// cloud-in-cell deposit on the stretched coarse grid.
//
// Fine (src/kernels.jl:601-726, src/accelerator.jl:531-565).  One wave per
// work unit (at most VH_UNIT particles of one occupied cell whose stencil
// fits): the unit's 8^3 sums are rank-4 updates on the matrix cores, written
// to a tile of its own.  The reference commits each cell with float atomics;
// here a second kernel builds every grid point from the tiles that cover it,
// units of a cell in slot order and cells in a fixed order: no float
// atomics, bit-reproducible, and the total charge (FP64) is reduced in the
// same pass.
//
// Coarse (src/kernels.jl:1207-1273, src/devicemesh.jl:93-106).  One wave per
// work unit of an occupied fine cell: the unit's particles are summed into a
// 4^3 window of coarse collocation points on the matrix cores; windows are
// committed, and particles that leave their window are deposited, through
// 64-bit fixed-point atomics, which are exact and therefore
// order-independent.
#include "vh_internal.h"

#define DEP_STAGE 64   // particles staged at a time (one per lane)
#define DEP_ROW 25     // LDS row of a particle: gx, gy, gz (8 each), odd stride

// table column of src/kernels.jl:770-772, FP32, 1-based, clamped
__device__ __forceinline__ int table_col(float d, float half, float spacing, float nbdt, int ncol) {
  float t = (d + half) / spacing * nbdt + 0.5f;
  int c = (int)floorf(t) + 1;
  return c < 1 ? 1 : (c > ncol ? ncol : c);
}

typedef float vh_f4d __attribute__((ext_vector_type(4)));

// The unit's 8^3 sums as a rank-P update on the matrix cores:
//   C[kk][ii + 8 jj] = sum_p gz[p][kk] * (gx[p][ii] gy[p][jj])
// four particles per v_mfma_f32_16x16x4_f32 (rows kk, 8 of 16 used), four
// 16-column tiles for the 64 (ii, jj) pairs.  The node values of 64
// particles at a time (one per lane) are staged in LDS, rows of missing
// particles zero, so that every operand read is unconditional; the next
// chunk's offsets are loaded while the current chunk is multiplied.
__global__ void __launch_bounds__(64)
    k_dep_tiles(const int32_t *__restrict__ P, const uint32_t *__restrict__ perm,
                const float *__restrict__ nodes,
                const vh_unit *__restrict__ units, const int32_t *__restrict__ cnt, int nk, int n,
                float spacing, float nbdt, int ncol, float *__restrict__ tiles, vh_sig sig) {
  __shared__ float vals[DEP_ROW * DEP_STAGE];
  vh_sig_store(sig);
  const int l = threadIdx.x;
  const int row = l & 15, kq = l >> 4;       // A: row kk, particle kq of the group
  const int ii = l & 7, jh = (l >> 3) & 1;   // B: column l & 15 -> (ii, jj = 2 t + jh)
  const float amask = row < 8 ? 1.f : 0.f;   // rows 8..15 of A are padding
  const float half = spacing * 0.5f;
  const int nunits = cnt[CNT_NUNITS];
  for (int u = blockIdx.x; u < nunits; u += gridDim.x) {
    const vh_unit un = units[u];
    // units of cells whose 10^3 stencil leaves the grid deposit nothing
    const int kx = un.key % nk + 1, ky = (un.key / nk) % nk + 1, kz = un.key / (nk * nk) + 1;
    const int bx = 2 * kx - 5, by = 2 * ky - 5, bz = 2 * kz - 5;
    if (!(bx >= 1 && by >= 1 && bz >= 1 && bx + 9 <= n && by + 9 <= n && bz + 9 <= n)) continue;
    const int lo = un.lo, hi = un.hi;
    vh_f4d acc[4];
#pragma unroll
    for (int tt = 0; tt < 4; ++tt) acc[tt] = (vh_f4d){0.f, 0.f, 0.f, 0.f};
    // slot s's record is record perm[s]; the index is loaded two chunks
    // ahead of its products, the offsets and the table rows one chunk ahead
    // (in flight during the previous chunk's products)
    auto rows = [&](float x, float y, float z, float4 (&g)[6]) {
      const int c1 = table_col(x, half, spacing, nbdt, ncol) - 1;
      const int c2 = table_col(y, half, spacing, nbdt, ncol) - 1;
      const int c3 = table_col(z, half, spacing, nbdt, ncol) - 1;
      const float4 *n1 = (const float4 *)(nodes + c1 * 8), *n2 = (const float4 *)(nodes + c2 * 8),
                   *n3 = (const float4 *)(nodes + c3 * 8);
      g[0] = n1[0];
      g[1] = n1[1];
      g[2] = n2[0];
      g[3] = n2[1];
      g[4] = n3[0];
      g[5] = n3[1];
    };
    float dx, dy, dz;
    vh_load_delta(P, perm[lo + l < hi ? lo + l : lo], dx, dy, dz);
    float4 g[6];
    rows(dx, dy, dz, g);
    vh_load_delta(P, perm[lo + DEP_STAGE + l < hi ? lo + DEP_STAGE + l : lo], dx, dy, dz);
    uint32_t inext = perm[lo + 2 * DEP_STAGE + l < hi ? lo + 2 * DEP_STAGE + l : lo];
    for (int cur = lo; cur < hi; cur += DEP_STAGE) {
      const int chunk = min(DEP_STAGE, hi - cur);
      const float live = l < chunk ? 1.f : 0.f;
      __syncthreads();   // the previous chunk's products have read vals
      float *v = vals + DEP_ROW * l;
#pragma unroll
      for (int a = 0; a < 6; ++a) {
        v[4 * a] = g[a].x * live;
        v[4 * a + 1] = g[a].y * live;
        v[4 * a + 2] = g[a].z * live;
        v[4 * a + 3] = g[a].w * live;
      }
      __syncthreads();
      if (cur + DEP_STAGE < hi) {   // the next chunk's rows and the offsets after it
        rows(dx, dy, dz, g);
        const int sn = cur + 3 * DEP_STAGE + l;
        vh_load_delta(P, inext, dx, dy, dz);
        inext = perm[sn < hi ? sn : lo];
      }
      const int ngroups = (chunk + 3) >> 2;
#pragma unroll 2
      for (int q = 0; q < ngroups; ++q) {
        const float *vp = vals + DEP_ROW * (4 * q + kq);
        const float a = vp[16 + (row & 7)] * amask;
        const float gxv = vp[ii];
#pragma unroll
        for (int tt = 0; tt < 4; ++tt) {
          const float b = gxv * vp[8 + 2 * tt + jh];
          acc[tt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[tt], 0, 0, 0);
        }
      }
    }
    // lane l holds rows 4 (l >> 4) + r, column l & 15 of each tile; rows >= 8 are padding
    if ((l >> 4) < 2) {
      float *out = tiles + (int64_t)u * 512;
#pragma unroll
      for (int tt = 0; tt < 4; ++tt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int kk = 4 * (l >> 4) + r;
          const int jj = 2 * tt + jh;
          out[ii + 8 * (jj + 8 * kk)] = acc[tt][r];
        }
    }
  }
}

// A cell cut into several units: its tiles summed in slot order into the
// tile of its first unit.  One block per fitting cell, a float2 per thread.
__global__ void __launch_bounds__(256)
    k_dep_merge(float *__restrict__ tiles, const int32_t *__restrict__ fitlist,
                const int32_t *__restrict__ unitstart, const int32_t *__restrict__ cnt) {
  const int nfit = cnt[CNT_NFIT];
  for (int f = blockIdx.x; f < nfit; f += gridDim.x) {
    const int g = fitlist[f];
    const int u0 = unitstart[g], u1 = unitstart[g + 1];
    if (u1 - u0 < 2) continue;
    float2 *t0 = (float2 *)(tiles + (int64_t)u0 * 512) + threadIdx.x;
    float2 v = *t0;
    for (int u = u0 + 1; u < u1; ++u) {
      const float2 w = ((const float2 *)(tiles + (int64_t)u * 512))[threadIdx.x];
      v.x += w.x;
      v.y += w.y;
    }
    *t0 = v;
  }
}

// every grid point from the tiles that cover it; thread = 2x2x2 block of
// points (2m-1, 2m) per axis, whose covering cells are m-2 .. m+1 (0-based),
// 64 cells.  A block takes 8 x 8 x 4 such blocks; the tile offsets of the
// 11 x 11 x 7 cells around them are staged in LDS (-1: no particle), each
// thread keeps a 64-bit mask of its occupied cells and reads the tiles of
// four cells at a time (all loads issued before the sums).  Each point adds
// its cells in (z, y, x) order: a fixed order, so the density is
// reproducible.  Also the FP64 charge sum(rho m0x m0y m0z) per block.
#define CMB_TX 8
#define CMB_TY 8
#define CMB_TZ 4
__global__ void __launch_bounds__(256)
    k_dep_combine(const float *__restrict__ tiles, const int32_t *__restrict__ cellmap,
                  const int32_t *__restrict__ unitstart, int nk, int n,
                  const double *__restrict__ m0, float *__restrict__ rho,
                  double *__restrict__ part) {
  constexpr int EX = CMB_TX + 3, EY = CMB_TY + 3, EZ = CMB_TZ + 3;
  __shared__ int32_t offs[EX * EY * EZ];
  __shared__ double lds[8];
  const int nb = nk + 1;
  const int tbx = (nb + CMB_TX - 1) / CMB_TX, tby = (nb + CMB_TY - 1) / CMB_TY;
  const int bt = blockIdx.x;
  const int bx0 = (bt % tbx) * CMB_TX, by0 = ((bt / tbx) % tby) * CMB_TY;
  const int bz0 = (bt / (tbx * tby)) * CMB_TZ;
  for (int e = threadIdx.x; e < EX * EY * EZ; e += blockDim.x) {
    const int ex = e % EX, ey = (e / EX) % EY, ez = e / (EX * EY);
    const int cx = bx0 - 2 + ex, cy = by0 - 2 + ey, cz = bz0 - 2 + ez;
    int off = -1;
    if (cx >= 0 && cy >= 0 && cz >= 0 && cx < nk && cy < nk && cz < nk) {
      const int g = cellmap[cx + nk * (cy + nk * cz)];
      // the cell's sums, merged into the tile of its first unit
      if (g >= 0) off = unitstart[g];
    }
    offs[e] = off;
  }
  __syncthreads();
  const int tx = threadIdx.x % CMB_TX, ty = (threadIdx.x / CMB_TX) % CMB_TY;
  const int tz = threadIdx.x / (CMB_TX * CMB_TY);
  const int mx = bx0 + tx, my = by0 + ty, mz = bz0 + tz;
  double q = 0.0;
  if (mx < nb && my < nb && mz < nb) {
    // bit (z 4 + y) 4 + x: cell (mx - 2 + x, my - 2 + y, mz - 2 + z), LDS
    // entry (tx + x, ty + y, tz + z); ascending bits = (z, y, x) order
    uint64_t mask = 0;
#pragma unroll
    for (int z = 0; z < 4; ++z)
#pragma unroll
      for (int y = 0; y < 4; ++y)
#pragma unroll
        for (int x = 0; x < 4; ++x)
          if (offs[(tx + x) + EX * ((ty + y) + EY * (tz + z))] >= 0)
            mask |= 1ull << ((z * 4 + y) * 4 + x);
    // the dual moments of the thread's points, loaded ahead (clamped indices:
    // points outside the grid are not stored)
    double wx[2], wy[2], wz[2];
#pragma unroll
    for (int d = 0; d < 2; ++d) {
      const int px = min(max(2 * mx - 1 + d, 0), n - 1), py = min(max(2 * my - 1 + d, 0), n - 1);
      const int pz = min(max(2 * mz - 1 + d, 0), n - 1);
      wx[d] = m0[px];
      wy[d] = m0[py];
      wz[d] = m0[pz];
    }
    float acc[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) acc[i] = 0.f;
    while (mask) {
      // the next four occupied cells (slots past the last repeat the first
      // and add nothing): tile of cell x, y, z at local (6-2x, 6-2y, 6-2z)
      const float *tp[4];
      bool ok[4];
      int b0 = 0;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        ok[u] = mask != 0;
        const int b = ok[u] ? __builtin_ctzll(mask) : b0;
        if (u == 0) b0 = b;
        mask &= mask - 1;
        const int x = b & 3, y = (b >> 2) & 3, z = b >> 4;
        const int off = offs[(tx + x) + EX * ((ty + y) + EY * (tz + z))];
        tp[u] = tiles + (int64_t)off * 512 + (6 - 2 * x) + 8 * ((6 - 2 * y) + 8 * (6 - 2 * z));
      }
      float2 v[4][2][2];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int dz = 0; dz < 2; ++dz)
#pragma unroll
          for (int dy = 0; dy < 2; ++dy) v[u][dz][dy] = *(const float2 *)(tp[u] + 8 * dy + 64 * dz);
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int dz = 0; dz < 2; ++dz)
#pragma unroll
          for (int dy = 0; dy < 2; ++dy) {
            if (ok[u]) {
              acc[(dz * 2 + dy) * 2 + 0] += v[u][dz][dy].x;
              acc[(dz * 2 + dy) * 2 + 1] += v[u][dz][dy].y;
            }
          }
    }
#pragma unroll
    for (int dz = 0; dz < 2; ++dz)
#pragma unroll
      for (int dy = 0; dy < 2; ++dy)
#pragma unroll
        for (int dx = 0; dx < 2; ++dx) {
          const int px = 2 * mx - 1 + dx, py = 2 * my - 1 + dy, pz = 2 * mz - 1 + dz;
          if (px < 0 || py < 0 || pz < 0 || px >= n || py >= n || pz >= n) continue;
          const float val = acc[(dz * 2 + dy) * 2 + dx];
          rho[px + (int64_t)n * (py + (int64_t)n * pz)] = val;
          q += (double)val * wx[dx] * wy[dy] * wz[dz];
        }
  }
  double sq = vh_block_sum_d(q, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = sq;
}

// fixed-order sum of nparts partials (one block)
__global__ void k_reduce_partials(const double *__restrict__ part, int nparts, int stride, int nvals,
                                  double *__restrict__ out) {
  __shared__ double lds[16];
  for (int v = 0; v < nvals; ++v) {
    double r = vh_block_sum_d(vh_thread_sum(part + (int64_t)v * stride, nparts), lds);
    if (threadIdx.x == 0) out[v] = r;
  }
}

// rho *= E((N - nout) * charge / q)  (src/accelerator.jl:557)
// nout: the rejection count; clear: the count the next push makes, zeroed
__global__ void k_dep_scale(float *__restrict__ rho, int64_t len, const double *__restrict__ q,
                            const int32_t *__restrict__ nout, int32_t *__restrict__ clear,
                            int64_t npart, double charge) {
  const float sc = (float)((double)(npart - *nout) * charge / q[0]);
  if (blockIdx.x == 0 && threadIdx.x == 0) *clear = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < len;
       i += (int64_t)gridDim.x * blockDim.x)
    rho[i] *= sc;
}


int vh_deposit_fine_impl(vh_accel *a, vh_mesh *fine, double charge, vh_sig sig) {
  vh_context *c = a->ctx;
  hipStream_t st = c->stream;
  // one wave per work unit up to 65536 units, then a stride over units
  const int grid = (int)(a->maxunits < 65536 ? a->maxunits : 65536);
  int tok = vh_prof_begin(c, VH_ST_DEPOSIT_KERNEL);
  hipLaunchKernelGGL(k_dep_tiles, dim3(grid), dim3(64), 0, st, a->parts[a->cur], vh_rec_perm(a), a->nodes,
                     a->units, a->counters, a->nk, a->n, a->spacing, (float)a->nbdt, a->ncol,
                     a->tiles, sig);
  vh_prof_end(c, tok);
  VH_LAUNCH_CHECK();
  const int mgrid = (int)(a->maxocc < 16384 ? a->maxocc : 16384);
  hipLaunchKernelGGL(k_dep_merge, dim3(mgrid), dim3(256), 0, st, a->tiles, a->fitlist, a->unitstart,
                     a->counters);
  VH_LAUNCH_CHECK();
  const int nb = a->nk + 1;
  const int cblocks = ((nb + CMB_TX - 1) / CMB_TX) * ((nb + CMB_TY - 1) / CMB_TY) *
                      ((nb + CMB_TZ - 1) / CMB_TZ);
  VH_REQUIRE(cblocks <= a->dpart_len, "deposit: %d combine blocks beyond %lld", cblocks,
             (long long)a->dpart_len);
  hipLaunchKernelGGL(k_dep_combine, dim3(cblocks), dim3(CMB_TX * CMB_TY * CMB_TZ), 0, st, a->tiles,
                     a->cellmap, a->unitstart, a->nk, a->n, fine->dual, fine->rho, a->dpart);
  VH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(1024), 0, st, a->dpart, cblocks, 0, 1,
                     a->dscal + 16);
  VH_LAUNCH_CHECK();
  int64_t len = (int64_t)a->n * a->n * a->n;
  hipLaunchKernelGGL(k_dep_scale, dim3(2048), dim3(256), 0, st, fine->rho, len, a->dscal + 16,
                     a->counters + a->nout_idx, a->counters + CNT_PUSH_NOUT + 1 - a->cur,
                     a->npart, charge);
  VH_LAUNCH_CHECK();
  return VH_OK;
}

// ---------------------------------------------------------------------------
// coarse cloud-in-cell
// ---------------------------------------------------------------------------
// _locate_cell of src/kernels.jl:1069-1081 in FP32: 1-based left point, 0 outside
__device__ __forceinline__ int locate_cell(const float *__restrict__ gt, int n, float x0t, float iwt,
                                           const int32_t *__restrict__ cells, int nbin, float x,
                                           float &w) {
  if (x < gt[0] || x > gt[n - 1]) {
    w = 0.f;
    return 0;
  }
  int b = (int)floorf((x - x0t) * iwt) + 1;
  if (b > nbin) b = nbin;
  int cc = cells[b - 1];
  const int lastc = n - 1;
  while (cc < lastc && gt[cc] < x) ++cc;  // gt[c+1] of the 1-based form
  w = (gt[cc] - x) / (gt[cc] - gt[cc - 1]);
  return cc;
}

__device__ __forceinline__ unsigned long long fx_of(float v, double scale) {
  return (unsigned long long)(long long)rint((double)v * scale);
}

__device__ __forceinline__ void fx_add(unsigned long long *acc, int64_t idx, float v, double scale) {
  if (v != 0.f) atomicAdd(acc + idx, fx_of(v, scale));
}

// The window of every fine cell index of an axis (vh_cicwin): the coarse
// interval of the cell's lower corner by locate_cell, as the windows of the
// deposit are anchored.  The three axes share the fine and coarse axes.
__global__ void k_cic_window_table(int nkf, float x0f, float hf, const float *__restrict__ gt, int n,
                                   float x0t, float iwt, const int32_t *__restrict__ cells,
                                   int nbin, vh_cicwin *__restrict__ tab) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nkf) return;
  float w;
  const float xl = x0f + ((float)t - 0.5f) * hf;   // fine cell index kc = t + 1
  vh_cicwin e;
  e.a = locate_cell(gt, n, x0t, iwt, cells, nbin, xl, w);
  const int a = e.a > 0 ? e.a : 1;
#pragma unroll
  for (int j = 0; j < 4; ++j) e.g[j] = gt[min(a - 1 + j, n - 1)];
#pragma unroll
  for (int j = 0; j < 3; ++j) e.r[j] = e.g[j + 1] > e.g[j] ? 1.f / (e.g[j + 1] - e.g[j]) : 0.f;
  tab[t] = e;
}

// Position of a particle in its cell's window along one axis: the linear
// weights X[0..3] on the window's four points (left point weight
// (g[q + 1] - p) / (g[q + 1] - g[q]) at q, the rest at q + 1, q the interval
// of p); in: g[0] < p <= g[3], i.e. the interval locate_cell finds for p is
// one of the window's three.  Outside, X holds finite values to be ignored.
// g = the window's points, r = (r[0], r[1], r[2], -): the entry of
// vh_cicwin after its anchor, in registers.
__device__ __forceinline__ void cic_axis(float p, float4 g, float4 r, float (&X)[4], bool &in) {
  in = p > g.x && p <= g.w;
  const bool s1 = p > g.y, s2 = p > g.z;
  const float hi = s2 ? g.w : (s1 ? g.z : g.y);
  const float rr = s2 ? r.z : (s1 ? r.y : r.x);
  const float a = in ? (hi - p) * rr : 0.f;
  const float b = 1.f - a;
  X[0] = s1 ? 0.f : a;
  X[1] = s2 ? 0.f : (s1 ? a : b);
  X[2] = s2 ? a : (s1 ? b : 0.f);
  X[3] = s2 ? b : 0.f;
}

// a window entry (vh_cicwin, 32 bytes) as anchor, points, reciprocals
struct cic_win {
  int a;
  float4 g, r;
};
__device__ __forceinline__ cic_win cic_load(const vh_cicwin *__restrict__ tab, int i) {
  const float4 *q = reinterpret_cast<const float4 *>(tab + i);
  const float4 lo = q[0], hi = q[1];
  cic_win w;
  w.a = __float_as_int(lo.x);
  w.g = make_float4(lo.y, lo.z, lo.w, hi.x);
  w.r = make_float4(hi.y, hi.z, hi.w, 0.f);
  return w;
}

// One wave per work unit (at most VH_UNIT particles of one occupied fine
// cell), one wave per block: beside the fine deposit's one-wave blocks,
// larger blocks of this kernel waited for four free wave slots on one CU
// and started late.  The unit's particles fall into a 4^3
// window of coarse points anchored at the coarse interval of the cell's
// lower corner (vh_cicwin, one table entry per axis); the window's sums are
// a rank-P update
//   W[x + 4 y][z] = sum_p (X_p[x] Y_p[y]) Z_p[z],
// X_p = the particle's two linear weights at its window offsets along x (zero
// elsewhere), run on v_mfma_f32_16x16x4_f32: A = the 16 products X Y of four
// particles, B = their z weights, each particle in its own four-column group
// (B[k][4 g + z] = Z_k[z] when g = k), so that a product takes four particles
// and the four groups are four partial sums.  The unit's FP32 sums are added
// across the groups and converted to 64-bit fixed point once.  The window
// of a cell's only unit is committed with global 64-bit atomics; a cell cut
// into several units writes one window per unit, and k_cic_flush adds them
// and commits once per cell, which keeps the atomics on the hot points of
// the cluster's core few.  Particles outside their window go straight to
// the global accumulator.  Integer sums do not depend on order, and the
// FP32 sums of a unit are formed in a fixed order: the result is
// reproducible.
#define CIC_STRIDE 21   // 16 products + 4 z weights, odd row stride
#define CIC_WAVES 1     // waves per block, each on its own units

__global__ void __launch_bounds__(64 * CIC_WAVES)
    k_cic_tiles(const int32_t *__restrict__ P, const uint32_t *__restrict__ perm,
                const vh_unit *__restrict__ units,
                const int32_t *__restrict__ unitstart, const int32_t *__restrict__ cnt, int nkf,
                float x0f, float hf, const vh_cicwin *__restrict__ wtab,
                const float *__restrict__ gt, int n, float x0t, float iwt,
                const int32_t *__restrict__ cells, int nbin, double scale,
                unsigned long long *__restrict__ acc, long long *__restrict__ wins) {
  __shared__ float lds_cic[CIC_WAVES * 64 * CIC_STRIDE];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  float *st = lds_cic + wv * 64 * CIC_STRIDE;
  const int l = threadIdx.x & 63;
  const int kq = l >> 4, xy = l & 15, grp = (l & 15) >> 2, zc = l & 3;
  const int nunits = cnt[CNT_NUNITS];
  for (int u = blockIdx.x * CIC_WAVES + wv; u < nunits; u += gridDim.x * CIC_WAVES) {
    const vh_unit un = units[u];
    const int key = un.key;
    const cic_win wx = cic_load(wtab, key % nkf), wy = cic_load(wtab, (key / nkf) % nkf),
                  wz = cic_load(wtab, key / (nkf * nkf));
    const bool anchored = wx.a > 0 && wy.a > 0 && wz.a > 0;
    vh_f4d d0 = {0.f, 0.f, 0.f, 0.f}, d1 = {0.f, 0.f, 0.f, 0.f};
    // slot s's record is record perm[s]; the index is loaded a chunk ahead
    // of the record
    vh_pos rn = vh_load_pos(P, perm[un.lo + l < un.hi ? un.lo + l : un.lo]);
    uint32_t inext = perm[un.lo + 64 + l < un.hi ? un.lo + 64 + l : un.lo];
    for (int c0 = un.lo; c0 < un.hi; c0 += 64) {
      const bool live = c0 + l < un.hi;
      const vh_pos r = rn;
      {  // the next chunk's records, in flight during this chunk's products
        const int sn = c0 + 128 + l;
        rn = vh_load_pos(P, inext);
        inext = perm[sn < un.hi ? sn : un.lo];
      }
      float p[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) p[d] = x0f + (float)(r.k[d] - 1) * hf + r.d[d];
      float X[4], Y[4], Z[4];
      bool inx, iny, inz;
      cic_axis(p[0], wx.g, wx.r, X, inx);
      cic_axis(p[1], wy.g, wy.r, Y, iny);
      cic_axis(p[2], wz.g, wz.r, Z, inz);
      const bool fit = live && anchored && inx && iny && inz;
      if (live && !fit) {
        // outside the window: the particle's eight points by the full
        // locate (src/kernels.jl:1207-1273), straight to the accumulator
        int ix[3];
        float ax[3];
        bool inside = true;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          ix[d] = locate_cell(gt, n, x0t, iwt, cells, nbin, p[d], ax[d]);
          inside = inside && ix[d] > 0;
        }
        if (inside) {
          const float bx = 1.f - ax[0], by = 1.f - ax[1], bz = 1.f - ax[2];
          const int64_t i0 = ix[0] - 1, j0 = ix[1] - 1, k0 = ix[2] - 1;
          const int64_t nn = n;
          fx_add(acc, i0 + nn * (j0 + nn * k0), ax[0] * ax[1] * ax[2], scale);
          fx_add(acc, i0 + nn * (j0 + nn * (k0 + 1)), ax[0] * ax[1] * bz, scale);
          fx_add(acc, i0 + nn * ((j0 + 1) + nn * k0), ax[0] * by * ax[2], scale);
          fx_add(acc, i0 + nn * ((j0 + 1) + nn * (k0 + 1)), ax[0] * by * bz, scale);
          fx_add(acc, (i0 + 1) + nn * (j0 + nn * k0), bx * ax[1] * ax[2], scale);
          fx_add(acc, (i0 + 1) + nn * (j0 + nn * (k0 + 1)), bx * ax[1] * bz, scale);
          fx_add(acc, (i0 + 1) + nn * ((j0 + 1) + nn * k0), bx * by * ax[2], scale);
          fx_add(acc, (i0 + 1) + nn * ((j0 + 1) + nn * (k0 + 1)), bx * by * bz, scale);
        }
      }
      // the particle's row: X[x] Y[y] at x + 4 y, then Z (zero unless the
      // particle is in its window); the wave's own rows only, so a
      // wave-level order suffices between the writes and the reads
      float *row = st + CIC_STRIDE * l;
#pragma unroll
      for (int y = 0; y < 4; ++y)
#pragma unroll
        for (int x = 0; x < 4; ++x) row[x + 4 * y] = X[x] * Y[y];
#pragma unroll
      for (int z = 0; z < 4; ++z) row[16 + z] = fit ? Z[z] : 0.f;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      // groups of four particles, alternately into two accumulators (two
      // independent product chains); rows past the chunk have zero z weights
      const int ng = (min(64, un.hi - c0) + 3) >> 2;
      for (int m = 0; m < ng; m += 2) {
        const float *pa = st + CIC_STRIDE * (4 * m + kq), *pb = pa + 4 * CIC_STRIDE;
        const float a0 = pa[xy], z0 = pa[16 + zc];
        const float a1 = pb[xy], z1 = pb[16 + zc];   // group m + 1 (zero weights past ng)
        d0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, grp == kq ? z0 : 0.f, d0, 0, 0, 0);
        d1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, grp == kq ? z1 : 0.f, d1, 0, 0, 0);
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
    // lane l holds W[4 (l >> 4) + v][4 grp + zc]: rows x = v, y = l >> 4; the
    // four groups (lanes l ^ 4, l ^ 8, l ^ 12 share y and z) added in FP32,
    // then the window point in fixed point
    long long sum[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      float s = d0[v] + d1[v];
      s += __shfl_xor(s, 4, 64);
      s += __shfl_xor(s, 8, 64);
      sum[v] = (long long)rint((double)s * scale);
    }
    if (unitstart[un.g + 1] - unitstart[un.g] > 1) {
      // window point x + 4 y + 16 z of this unit, for k_cic_flush
      if (grp == 0) {
#pragma unroll
        for (int v = 0; v < 4; ++v) wins[(int64_t)u * 64 + v + 4 * kq + 16 * zc] = sum[v];
      }
    } else if (grp == 0 && anchored) {
      const int64_t j = wy.a - 1 + kq, k = wz.a - 1 + zc;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t i = wx.a - 1 + v;
        if (sum[v] != 0 && i < n && j < n && k < n)
          atomicAdd(acc + i + (int64_t)n * (j + (int64_t)n * k), (unsigned long long)sum[v]);
      }
    }
  }
}

// The windows of a cell cut into several units: added (exact integers) and
// committed once.  One wave per occupied cell, a window point per lane.
__global__ void __launch_bounds__(64)
    k_cic_flush(const int32_t *__restrict__ occupied, const int32_t *__restrict__ unitstart,
                const int32_t *__restrict__ cnt, int nkf, const vh_cicwin *__restrict__ wtab,
                int n, const long long *__restrict__ wins, unsigned long long *__restrict__ acc) {
  const int l = threadIdx.x;
  const int nocc = cnt[CNT_NOCC];
  for (int g = blockIdx.x; g < nocc; g += gridDim.x) {
    const int u0 = unitstart[g], u1 = unitstart[g + 1];
    if (u1 - u0 < 2) continue;
    const int key = occupied[g];
    const int ax = wtab[key % nkf].a, ay = wtab[(key / nkf) % nkf].a, az = wtab[key / (nkf * nkf)].a;
    if (!(ax > 0 && ay > 0 && az > 0)) continue;  // nothing was windowed
    long long sum = 0;
    for (int u = u0; u < u1; ++u) sum += wins[(int64_t)u * 64 + l];
    const int64_t i = ax - 1 + (l & 3), j = ay - 1 + ((l >> 2) & 3), k = az - 1 + (l >> 4);
    if (sum != 0 && i < n && j < n && k < n)
      atomicAdd(acc + i + (int64_t)n * (j + (int64_t)n * k), (unsigned long long)sum);
  }
}

// rho = (sum / 2^S) * (E(charge) / (wx wy wz)), and the accumulator cleared
__global__ void k_cic_convert(unsigned long long *__restrict__ acc, float *__restrict__ rho,
                              const float *__restrict__ wl, int n, double inv_scale, float charge) {
  const int64_t len = (int64_t)n * n * n;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < len;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(t % n), j = (int)((t / n) % n), k = (int)(t / ((int64_t)n * n));
    const long long v = (long long)acc[t];
    acc[t] = 0ull;
    const float r = (float)((double)v * inv_scale);
    rho[t] = r * (charge / (wl[i] * wl[j] * wl[k]));
  }
}

int vh_deposit_cic_impl(vh_accel *a, vh_mesh *m, double charge, bool defer) {
  vh_context *c = a->ctx;
  hipStream_t st = c->stream;
  const double scale = ldexp(1.0, a->cic_shift);
  if (m->acc_pending) {   // an earlier deferred deposit never formed: clear its sums
    hipLaunchKernelGGL(k_cic_convert, dim3(2048), dim3(256), 0, st, (unsigned long long *)m->acc64,
                       m->rho, m->duallen, m->n, m->acc_inv_scale, m->acc_charge);
    VH_LAUNCH_CHECK();
    m->acc_pending = 0;
  }
  // the windows of the fine cells on this mesh: built once per mesh (in
  // every replay of a graph captured before they were)
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  VH_CHECK_HIP(hipStreamIsCapturing(st, &cs));
  if (a->cicwtab_serial != m->serial || cs != hipStreamCaptureStatusNone) {
    hipLaunchKernelGGL(k_cic_window_table, dim3((a->nk + 63) / 64), dim3(64), 0, st, a->nk,
                       (float)a->x0, (float)a->h, m->colloc, m->n, m->loc_x0, m->loc_iw,
                       m->loc_cells, m->nbin, a->cicwtab);
    VH_LAUNCH_CHECK();
    if (cs == hipStreamCaptureStatusNone) a->cicwtab_serial = m->serial;
  }
  // the fine deposit's tiles (2 KiB per unit, consumed by then) hold the
  // 512-byte windows of units of multi-unit cells
  long long *wins = a->cicwins;
  const int grid = (int)((a->maxunits + CIC_WAVES - 1) / CIC_WAVES < 16384
                             ? (a->maxunits + CIC_WAVES - 1) / CIC_WAVES : 16384);
  hipLaunchKernelGGL(k_cic_tiles, dim3(grid), dim3(64 * CIC_WAVES), 0, st, a->parts[a->cur],
                     vh_rec_perm(a), a->units, a->unitstart, a->counters, a->nk, (float)a->x0,
                     (float)a->h, a->cicwtab, m->colloc, m->n, m->loc_x0, m->loc_iw, m->loc_cells,
                     m->nbin, scale, (unsigned long long *)m->acc64, wins);
  VH_LAUNCH_CHECK();
  const int fgrid = (int)(a->maxocc < 65536 ? a->maxocc : 65536);
  hipLaunchKernelGGL(k_cic_flush, dim3(fgrid), dim3(64), 0, st, a->occupied, a->unitstart,
                     a->counters, a->nk, a->cicwtab, m->n, wins, (unsigned long long *)m->acc64);
  VH_LAUNCH_CHECK();
  if (defer) {
    m->acc_pending = 1;
    m->acc_inv_scale = 1.0 / scale;
    m->acc_charge = (float)charge;
    return VH_OK;
  }
  hipLaunchKernelGGL(k_cic_convert, dim3(2048), dim3(256), 0, st, (unsigned long long *)m->acc64,
                     m->rho, m->duallen, m->n, 1.0 / scale, (float)charge);
  VH_LAUNCH_CHECK();
  return VH_OK;
}
