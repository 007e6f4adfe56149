// Charge deposition: the smoothed 8^3 deposit on the fine grid and the
// cloud-in-cell deposit on the stretched coarse grid.
//
// Fine (src/kernels.jl:601-726, src/accelerator.jl:531-565).  One wave per
// occupied cell whose stencil fits; 32 particles at a time have their 24
// tabulated node values staged in LDS; lane (jj, kk) owns the x-column of 8
// points and accumulates w = gy[jj] gz[kk] into 8 registers with the
// reference's fma order.  The reference then commits each cell with float
// atomics.  Here each cell writes its 512 sums to a tile, and a second kernel
// builds every grid point from the (at most 64) tiles that cover it, in a
// fixed cell order: no float atomics, bit-reproducible, and the total charge
// (FP64) is reduced in the same pass.
//
// Coarse (src/kernels.jl:1207-1273, src/devicemesh.jl:93-106).  One wave per
// occupied fine cell; each lane owns one point of a 4^3 window of coarse
// collocation points anchored at the cell's lower corner and sums its share
// of the cell's particles in order.  Windows are committed, and particles
// that leave their window are deposited, through 64-bit fixed-point atomics,
// which are exact and therefore order-independent.
#include "vh_internal.h"

#define DEP_STAGE 32

// table column of src/kernels.jl:770-772, FP32, 1-based, clamped
__device__ __forceinline__ int table_col(float d, float half, float spacing, float nbdt, int ncol) {
  float t = (d + half) / spacing * nbdt + 0.5f;
  int c = (int)floorf(t) + 1;
  return c < 1 ? 1 : (c > ncol ? ncol : c);
}

typedef float vh_f4d __attribute__((ext_vector_type(4)));

// The cell's 8^3 sums as a rank-P update on the matrix cores:
//   C[kk][ii + 8 jj] = sum_p gz[p][kk] * (gx[p][ii] gy[p][jj])
// four particles per v_mfma_f32_16x16x4_f32 (rows kk, 8 of 16 used), four
// 16-column tiles for the 64 (ii, jj) pairs.  Node values of 32 particles at
// a time are staged in LDS; missing particles of a group contribute zeros.
__global__ void __launch_bounds__(64)
    k_dep_tiles(const int32_t *__restrict__ P, int64_t npad, const float *__restrict__ nodes,
                const int32_t *__restrict__ fitlist, const int32_t *__restrict__ bounds,
                const int32_t *__restrict__ cnt, float spacing, float nbdt, int ncol,
                float *__restrict__ tiles) {
  __shared__ float vals[24 * DEP_STAGE];
  const float *D = (const float *)P;
  const int l = threadIdx.x;
  const int row = l & 15, kq = l >> 4;       // A: row kk, particle kq of the group
  const int ii = l & 7, jh = (l >> 3) & 1;   // B: column l & 15 -> (ii, jj = 2 t + jh)
  const float half = spacing * 0.5f;
  const int nfit = cnt[CNT_NFIT];
  for (int f = blockIdx.x; f < nfit; f += gridDim.x) {
    const int g = fitlist[f];
    const int lo = bounds[g], hi = bounds[g + 1];
    vh_f4d acc[4];
#pragma unroll
    for (int tt = 0; tt < 4; ++tt) acc[tt] = (vh_f4d){0.f, 0.f, 0.f, 0.f};
    for (int cur = lo; cur < hi; cur += DEP_STAGE) {
      const int chunk = min(DEP_STAGE, hi - cur);
      __syncthreads();
      if (l < chunk) {
        const int64_t s = cur + l;
        const int c1 = table_col(D[PS_DX * npad + s], half, spacing, nbdt, ncol) - 1;
        const int c2 = table_col(D[PS_DY * npad + s], half, spacing, nbdt, ncol) - 1;
        const int c3 = table_col(D[PS_DZ * npad + s], half, spacing, nbdt, ncol) - 1;
        float *v = vals + 24 * l;
#pragma unroll
        for (int a = 0; a < 8; ++a) {
          v[a] = nodes[c1 * 8 + a];
          v[8 + a] = nodes[c2 * 8 + a];
          v[16 + a] = nodes[c3 * 8 + a];
        }
      }
      __syncthreads();
      for (int q0 = 0; q0 < chunk; q0 += 4) {
        const int p = q0 + kq;
        const bool live = p < chunk;
        const float *v = vals + 24 * (live ? p : 0);
        const float a = (live && row < 8) ? v[16 + row] : 0.f;
        const float gxv = live ? v[ii] : 0.f;
#pragma unroll
        for (int tt = 0; tt < 4; ++tt) {
          const float b = gxv * v[8 + 2 * tt + jh];
          acc[tt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[tt], 0, 0, 0);
        }
      }
    }
    // lane l holds rows 4 (l >> 4) + r, column l & 15 of each tile; rows >= 8 are padding
    if ((l >> 4) < 2) {
      float *out = tiles + (int64_t)g * 512;
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

// every grid point from the tiles that cover it; thread = 2x2x2 block of
// points (2m-1, 2m) per axis, whose covering cells are m-2 .. m+1 (0-based).
// Also the FP64 charge sum(rho m0x m0y m0z) per block.
__global__ void __launch_bounds__(256)
    k_dep_combine(const float *__restrict__ tiles, const int32_t *__restrict__ cellmap, int nk,
                  int n, const double *__restrict__ m0, float *__restrict__ rho,
                  double *__restrict__ part) {
  __shared__ double lds[8];
  const int nb = nk + 1;
  const int64_t total = (int64_t)nb * nb * nb;
  double q = 0.0;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int mx = (int)(t % nb), my = (int)((t / nb) % nb), mz = (int)(t / ((int64_t)nb * nb));
    float acc[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) acc[i] = 0.f;
    for (int cz = mz - 2; cz <= mz + 1; ++cz) {
      if (cz < 0 || cz >= nk) continue;
      const int lz = 2 * (mz - cz) + 2;
      for (int cy = my - 2; cy <= my + 1; ++cy) {
        if (cy < 0 || cy >= nk) continue;
        const int ly = 2 * (my - cy) + 2;
        for (int cx = mx - 2; cx <= mx + 1; ++cx) {
          if (cx < 0 || cx >= nk) continue;
          const int g = cellmap[cx + nk * (cy + nk * cz)];
          if (g < 0) continue;
          const int lx = 2 * (mx - cx) + 2;
          const float *tl = tiles + (int64_t)g * 512;
#pragma unroll
          for (int dz = 0; dz < 2; ++dz)
#pragma unroll
            for (int dy = 0; dy < 2; ++dy) {
              const float2 v = *(const float2 *)(tl + lx + 8 * ((ly + dy) + 8 * (lz + dz)));
              acc[(dz * 2 + dy) * 2 + 0] += v.x;
              acc[(dz * 2 + dy) * 2 + 1] += v.y;
            }
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
          const float v = acc[(dz * 2 + dy) * 2 + dx];
          rho[px + (int64_t)n * (py + (int64_t)n * pz)] = v;
          q += (double)v * m0[px] * m0[py] * m0[pz];
        }
  }
  double s = vh_block_sum_d(q, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// fixed-order sum of nparts partials (one block)
__global__ void k_reduce_partials(const double *__restrict__ part, int nparts, int stride, int nvals,
                                  double *__restrict__ out) {
  __shared__ double lds[16];
  for (int v = 0; v < nvals; ++v) {
    double s = 0.0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) s += part[(int64_t)v * stride + i];
    double r = vh_block_sum_d(s, lds);
    if (threadIdx.x == 0) out[v] = r;
  }
}

// rho *= E((N - nout) * charge / q)  (src/accelerator.jl:557)
__global__ void k_dep_scale(float *__restrict__ rho, int64_t len, const double *__restrict__ q,
                            const int32_t *__restrict__ cnt, int64_t npart, double charge) {
  const float sc = (float)((double)(npart - cnt[CNT_NOUT]) * charge / q[0]);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < len;
       i += (int64_t)gridDim.x * blockDim.x)
    rho[i] *= sc;
}

#define COMBINE_BLOCKS 1024

int vh_deposit_fine_impl(vh_accel *a, vh_mesh *fine, double charge) {
  vh_context *c = a->ctx;
  hipStream_t st = c->stream;
  // one wave per fitting cell up to 65536 cells: the dispatcher balances
  // cells of very different weights
  const int grid = (int)(a->maxocc < 65536 ? a->maxocc : 65536);
  int tok = vh_prof_begin(c, VH_ST_DEPOSIT_KERNEL);
  hipLaunchKernelGGL(k_dep_tiles, dim3(grid), dim3(64), 0, st, a->parts[a->cur], a->npad, a->nodes,
                     a->fitlist, a->bounds, a->counters, a->spacing, (float)a->nbdt, a->ncol,
                     a->tiles);
  vh_prof_end(c, tok);
  VH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_dep_combine, dim3(COMBINE_BLOCKS), dim3(256), 0, st, a->tiles, a->cellmap,
                     a->nk, a->n, fine->dual, fine->rho, a->dpart);
  VH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(1024), 0, st, a->dpart, COMBINE_BLOCKS, 0, 1,
                     a->dscal + 16);
  VH_LAUNCH_CHECK();
  int64_t len = (int64_t)a->n * a->n * a->n;
  hipLaunchKernelGGL(k_dep_scale, dim3(2048), dim3(256), 0, st, fine->rho, len, a->dscal + 16,
                     a->counters, a->npart, charge);
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

#define CIC_THREADS 256

// One workgroup per occupied fine cell.  Each thread takes particles of the
// cell in turn and adds its eight corner weights, as 64-bit fixed point, into
// a 4^3 LDS window of coarse points anchored at the cell's lower corner;
// particles whose corners leave the window go straight to the global
// accumulator.  Integer sums are exact, so the result does not depend on the
// order the atomics land in.
__global__ void __launch_bounds__(CIC_THREADS)
    k_cic_tiles(const int32_t *__restrict__ P, int64_t npad, const int32_t *__restrict__ occupied,
                const int32_t *__restrict__ bounds, const int32_t *__restrict__ cnt, int nkf,
                float x0f, float hf, const float *__restrict__ gt, int n, float x0t, float iwt,
                const int32_t *__restrict__ cells, int nbin, double scale,
                unsigned long long *__restrict__ acc) {
  __shared__ unsigned long long win[64];
  const float *D = (const float *)P;
  const int t = threadIdx.x;
  const int nocc = cnt[CNT_NOCC];
  for (int g = blockIdx.x; g < nocc; g += gridDim.x) {
    const int key = occupied[g];
    const int kc[3] = {key % nkf + 1, (key / nkf) % nkf + 1, key / (nkf * nkf) + 1};
    // window anchor: coarse interval of the cell's lower corner, per axis
    int anc[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      float w;
      const float xl = x0f + ((float)(kc[d] - 1) - 0.5f) * hf;
      anc[d] = locate_cell(gt, n, x0t, iwt, cells, nbin, xl, w);  // 0: every particle escapes
    }
    const bool anchored = anc[0] > 0 && anc[1] > 0 && anc[2] > 0;
    if (t < 64) win[t] = 0ull;
    __syncthreads();
    const int lo = bounds[g], hi = bounds[g + 1];
    for (int s = lo + t; s < hi; s += CIC_THREADS) {
      int ix[3];
      float ax[3];
      bool inside = true;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const float p = x0f + (float)(P[(PS_KX + d) * npad + s] - 1) * hf + D[(PS_DX + d) * npad + s];
        ix[d] = locate_cell(gt, n, x0t, iwt, cells, nbin, p, ax[d]);
        inside = inside && ix[d] > 0;
      }
      if (!inside) continue;
      const float bx = 1.f - ax[0], by = 1.f - ax[1], bz = 1.f - ax[2];
      const float c[8] = {ax[0] * ax[1] * ax[2], ax[0] * ax[1] * bz, ax[0] * by * ax[2],
                          ax[0] * by * bz,       bx * ax[1] * ax[2], bx * ax[1] * bz,
                          bx * by * ax[2],       bx * by * bz};
      const int q0 = ix[0] - anc[0], q1 = ix[1] - anc[1], q2 = ix[2] - anc[2];
      const bool fits = anchored && q0 >= 0 && q0 <= 2 && q1 >= 0 && q1 <= 2 && q2 >= 0 && q2 <= 2;
      if (fits) {
        const int base = q0 + 4 * (q1 + 4 * q2);
        // corner (di, dj, dk) -> window point base + di + 4 dj + 16 dk
        atomicAdd(&win[base], fx_of(c[0], scale));
        atomicAdd(&win[base + 16], fx_of(c[1], scale));
        atomicAdd(&win[base + 4], fx_of(c[2], scale));
        atomicAdd(&win[base + 20], fx_of(c[3], scale));
        atomicAdd(&win[base + 1], fx_of(c[4], scale));
        atomicAdd(&win[base + 17], fx_of(c[5], scale));
        atomicAdd(&win[base + 5], fx_of(c[6], scale));
        atomicAdd(&win[base + 21], fx_of(c[7], scale));
      } else {
        const int64_t i0 = ix[0] - 1, j0 = ix[1] - 1, k0 = ix[2] - 1;
        const int64_t nn = n;
        fx_add(acc, i0 + nn * (j0 + nn * k0), c[0], scale);
        fx_add(acc, i0 + nn * (j0 + nn * (k0 + 1)), c[1], scale);
        fx_add(acc, i0 + nn * ((j0 + 1) + nn * k0), c[2], scale);
        fx_add(acc, i0 + nn * ((j0 + 1) + nn * (k0 + 1)), c[3], scale);
        fx_add(acc, (i0 + 1) + nn * (j0 + nn * k0), c[4], scale);
        fx_add(acc, (i0 + 1) + nn * (j0 + nn * (k0 + 1)), c[5], scale);
        fx_add(acc, (i0 + 1) + nn * ((j0 + 1) + nn * k0), c[6], scale);
        fx_add(acc, (i0 + 1) + nn * ((j0 + 1) + nn * (k0 + 1)), c[7], scale);
      }
    }
    __syncthreads();
    if (t < 64 && anchored) {
      const int tx = t & 3, ty = (t >> 2) & 3, tz = t >> 4;
      const int64_t i = anc[0] - 1 + tx, j = anc[1] - 1 + ty, k = anc[2] - 1 + tz;
      const unsigned long long v = win[t];
      if (v != 0ull && i < n && j < n && k < n) atomicAdd(acc + i + (int64_t)n * (j + (int64_t)n * k), v);
    }
    __syncthreads();
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

int vh_deposit_cic_impl(vh_accel *a, vh_mesh *m, double charge) {
  vh_context *c = a->ctx;
  hipStream_t st = c->stream;
  const double scale = ldexp(1.0, a->cic_shift);
  const int grid = (int)(a->maxocc < 65535 ? a->maxocc : 65535);
  hipLaunchKernelGGL(k_cic_tiles, dim3(grid), dim3(CIC_THREADS), 0, st, a->parts[a->cur], a->npad,
                     a->occupied, a->bounds, a->counters, a->nk, (float)a->x0, (float)a->h,
                     m->colloc, m->n, m->loc_x0, m->loc_iw, m->loc_cells, m->nbin, scale,
                     (unsigned long long *)m->acc64);
  VH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_cic_convert, dim3(2048), dim3(256), 0, st, (unsigned long long *)m->acc64,
                     m->rho, m->duallen, m->n, 1.0 / scale, (float)charge);
  VH_LAUNCH_CHECK();
  return VH_OK;
}
