// Two-level tensor-product spline Poisson solve, spline coefficients and the
// mean field, on the device.
//
// Every matrix product of the chain is the reference's "rotating" product
// (src/tensorsolver.jl:217-223): the n x m view of the source is multiplied
// by A along its first dimension and written transposed, which rotates the
// dimensions; three in a row apply A along all three axes.  Here it is one
// MFMA kernel (v_mfma_f32_16x16x4_f32, FP32 in and out) whose block owns all
// n output rows of 32 output columns, so every source column is read once;
// the Hadamard product by 1/(l_i+l_j+l_k) (src/tensorsolver.jl:291), the copy
// into the interior of phi (src/kernels.jl:804-813) and the `csol .+= ...` of
// the mean field (src/devicemesh.jl:182) are fused into its epilogue.
#include "vh_internal.h"

#include <cmath>

typedef float vh_f4 __attribute__((ext_vector_type(4)));

// The product C = A S (C n x m, row-major out[i' m + j]; S[k][j] = src[j n + k])
// by v_mfma_f32_16x16x4_f32 (lane l: A[l%16][k], B[k][l%16], k from l/16).
// A block holds W row tiles of 16, one wave each (W = all MT row tiles up to
// 9, else half of them), and walks panels of NC = 16 NCT output columns,
// panel p, p + gridDim.x, ... (a few blocks per compute unit):
//   * each wave keeps the A operands of its row tile for the whole K range in
//     registers, loaded once per block, with K permuted so that lane group
//     g = l/16 owns the contiguous slice k = g KQ + s, s < KQ (KQ odd,
//     >= n/4; k >= n padded with zero A).  A is stored once, at mesh
//     creation, in that operand order (vh_gemm_pack): step s of row tile r
//     is 64 consecutive floats, one per lane, so the wave's reads are
//     coalesced;
//   * a panel's source columns are contiguous in memory: the block reads
//     them with coalesced 8-byte loads into one of two LDS buffers, column
//     stride ldb >= 4 KQ, ldb = 2 (mod 32), so that the 32 lanes of a
//     half-wave (16 columns x 2 lane groups, odd KQ) hit 32 different banks;
//     rows k >= n are zero, so every operand address is a fixed offset;
//   * software pipeline: a panel's loads are issued a whole panel ahead of
//     its products and land in the other LDS buffer while the current panel
//     is multiplied, so that HBM latency, matrix-core work and stores of
//     successive panels overlap.
template <int KQ, int NCT, int EPI>
__global__ void __launch_bounds__(576)
    k_rot_gemm(const float *__restrict__ A, const float *__restrict__ src, float *__restrict__ out,
               int n, int64_t m, int ldb, const float *__restrict__ aux, int nfull) {
  extern __shared__ float Bs[];   // two panels of NC x ldb
  constexpr int NC = 16 * NCT;    // output columns of a panel
  constexpr int PF = 4 * NCT;     // source pairs per thread and panel (NC/2 n <= PF blockDim)
  const int nthr = blockDim.x;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int lc = l & 15, lk = l >> 4;
  const int mt = (n + 15) >> 4;
  const int r = blockIdx.y * (nthr >> 6) + w;
  const int64_t npanel = (m + NC - 1) / NC;
  // XCD-aware walk (gridDim.x is a multiple of 8; workgroups go to the eight
  // XCDs in turn): the blocks of one XCD take consecutive panels, so that the
  // 64-byte pieces they write of each output row meet in that XCD's L2
  const int xcd = blockIdx.x & 7, per = gridDim.x >> 3;
  const int64_t chunk = (npanel + 7) / 8;
  int64_t p = xcd * chunk + (blockIdx.x >> 3);
  const int64_t pend = (xcd + 1) * chunk < npanel ? (xcd + 1) * chunk : npanel;
  if (p >= pend) return;  // the whole block
  // A operands of row tile r (waves past the last tile read tile mt - 1 and
  // only move panels)
  float a[KQ];
  const float *ar = A + ((int64_t)(r < mt ? r : mt - 1) * KQ) * 64 + l;
#pragma unroll
  for (int s = 0; s < KQ; ++s) a[s] = ar[64 * s];
  const float2 *src2 = (const float2 *)src;
  const int64_t tot2 = m * n / 2;
  const int per2 = NC / 2 * n;
  const int pbuf = NC * ldb;
  // Every load is issued, at an address clamped into the source; pairs
  // outside the panel or the source are zeroed when the panel is written to
  // LDS.  (Using the loaded value right away, or a load under a condition,
  // makes the compiler wait for the load at once, which serialises the
  // pipeline.)
  auto fetch = [&](float2 (&pv)[PF], int64_t q) {
    const int64_t base2 = q * NC * (int64_t)n / 2;
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      const int64_t g = base2 + threadIdx.x + u * nthr;
      pv[u] = src2[g < tot2 ? g : tot2 - 1];
    }
  };
  auto stash = [&](const float2 (&pv)[PF], int64_t q, float *B) {
    const int64_t base2 = q * NC * (int64_t)n / 2;
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      const int e = threadIdx.x + u * nthr;
      if (e < per2) {
        const int f = 2 * e, cc = f / n, k = f - cc * n;   // n even: pairs never straddle a column
        *(float2 *)(B + cc * ldb + k) = base2 + e < tot2 ? pv[u] : make_float2(0.f, 0.f);
      }
    }
  };
  // rows k in [n, ldb) of both buffers stay zero (the panels fill k < n)
  for (int e = threadIdx.x; e < 2 * NC * (ldb - n); e += nthr) {
    const int row = e / (ldb - n), k = n + e % (ldb - n);
    Bs[row * ldb + k] = 0.f;
  }
  float2 R[PF];
  fetch(R, p);
  stash(R, p, Bs);
  __syncthreads();
  if (p + per < pend) fetch(R, p + per);
  int buf = 0;
  for (; p < pend; p += per, buf ^= 1) {
    // the products of this panel while the next one's loads are in flight;
    // the epilogue's operands (the Hadamard factors, or the values to add
    // to) are loaded first, so that their latency hides behind the products
    const int64_t j0 = p * NC;
    float pre[NCT][4];
    if (EPI == EPI_HADAMARD || EPI == EPI_ACCUMULATE) {
      const float *srcp = EPI == EPI_HADAMARD ? aux : out;
      const int64_t last = (int64_t)n * m - 1;
#pragma unroll
      for (int ct = 0; ct < NCT; ++ct)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int64_t idx = (int64_t)(r * 16 + 4 * lk + v) * m + j0 + ct * 16 + lc;
          pre[ct][v] = srcp[idx < last ? idx : last];   // clamped: never under a condition
        }
    }
    const float *B = Bs + buf * pbuf;
    vh_f4 acc[NCT];
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct) acc[ct] = (vh_f4){0.f, 0.f, 0.f, 0.f};
    if (r < mt) {
#pragma unroll
      for (int s = 0; s < KQ; ++s) {
        const int k = lk * KQ + s;  // k >= n: zero A, zero LDS padding
#pragma unroll
        for (int ct = 0; ct < NCT; ++ct)
          acc[ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[s], B[(16 * ct + lc) * ldb + k], acc[ct], 0, 0, 0);
      }
    }
    // the next panel into the other buffer before this one's stores, so that
    // waiting for its loads does not wait for the stores; the scheduling
    // barrier keeps that wait after the products
    __builtin_amdgcn_sched_barrier(0);
    const bool more = p + per < pend;
    if (more) stash(R, p + per, Bs + (buf ^ 1) * pbuf);
    // the loads of the panel after next, issued before this panel's stores
    if (p + 2 * per < pend) fetch(R, p + 2 * per);
    if (r < mt) {
#pragma unroll
      for (int ct = 0; ct < NCT; ++ct) {
        const int64_t j = j0 + ct * 16 + lc;
        if (j >= m) continue;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int orow = r * 16 + 4 * lk + v;
          if (orow >= n) continue;
          const int64_t idx = (int64_t)orow * m + j;
          const float val = acc[ct][v];
          if (EPI == EPI_STORE) {
            out[idx] = val;
          } else if (EPI == EPI_HADAMARD) {
            out[idx] = val * pre[ct][v];
          } else if (EPI == EPI_ACCUMULATE) {
            out[idx] = pre[ct][v] + val;
          } else {  // EPI_FILL_INTERIOR: idx = i + n (jj + n kz) of the interior, kz = orow
            const int jc = (int)j, i = jc % n, jj = jc / n;
            out[(i + 1) + nfull * ((jj + 1) + (int64_t)nfull * (orow + 1))] = val;
          }
        }
      }
    }
    __syncthreads();
  }
}

// K slice per lane group: the smallest odd bucket >= n/4 (0 beyond n = 276)
static int gemm_kq(int n) {
  static const int buckets[] = {5, 9, 15, 17, 25, 33, 41, 49, 57, 65, 69};
  const int need = (n + 3) / 4;
  for (int b : buckets)
    if (b >= need) return b;
  return 0;
}

// A (n x n, column-major FP64) in operand order: element
// [(r KQ + s) 64 + l] = A[16 r + l % 16][(l / 16) KQ + s], zero outside.
int vh_gemm_pack(vh_context *c, float **dst, const double *colmajor, int n) {
  const int mt = (n + 15) / 16, kq = gemm_kq(n);
  VH_REQUIRE(n >= 2 && mt <= 17 && kq > 0, "rotating product: n = %d beyond 272", n);
  std::vector<float> op((size_t)mt * kq * 64);
  for (int r = 0; r < mt; ++r)
    for (int s = 0; s < kq; ++s)
      for (int l = 0; l < 64; ++l) {
        const int row = 16 * r + (l & 15), k = (l >> 4) * kq + s;
        op[((size_t)r * kq + s) * 64 + l] =
            (row < n && k < n) ? (float)colmajor[row + (size_t)n * k] : 0.f;
      }
  VH_TRY(vh_alloc((void **)dst, op.size() * sizeof(float)));
  VH_CHECK_HIP(hipMemcpy(*dst, op.data(), op.size() * sizeof(float), hipMemcpyHostToDevice));
  return VH_OK;
}

// launch shape: column tiles per panel, waves (row tiles) per block, blocks
// per compute unit
#define GEMM_NCT 1
#define GEMM_WAVES(mt) ((mt) <= 9 ? (mt) : ((mt) + 1) / 2)
#define GEMM_BPC 1

int vh_rotate_gemm(vh_context *c, const float *A, const float *src, float *out, int n, int64_t m,
                   int epi, const float *aux, int nfull) {
  const int mt = (n + 15) / 16;
  VH_REQUIRE(mt >= 1 && mt <= 17, "rotating product: n = %d beyond 272", n);
  VH_REQUIRE(n % 2 == 0 && ((uintptr_t)src & 7) == 0,
             "rotating product: n (%d) must be even and the source 8-byte aligned", n);
  const int kq = gemm_kq(n);
  // smallest >= 4 kq (the K range the lanes read) with ldb = 2 (mod 32)
  const int ldb = 4 * kq + ((2 - (4 * kq) % 32) + 32) % 32;
  const int wv = GEMM_WAVES(mt);
  constexpr int nc = 16 * GEMM_NCT;
  VH_REQUIRE(nc / 2 * n <= 4 * GEMM_NCT * 64 * wv, "rotating product: panel of n = %d too large", n);
  const int gy = (mt + wv - 1) / wv;
  const int64_t npanel = (m + nc - 1) / nc;
  int64_t gx = ((int64_t)c->num_cus * GEMM_BPC + gy - 1) / gy;
  if (gx > npanel) gx = npanel;
  gx = (gx + 7) / 8 * 8;
  dim3 grid((unsigned)gx, (unsigned)gy), block(64 * wv);
  const size_t lds = sizeof(float) * 2 * nc * (size_t)ldb;
  VH_REQUIRE(epi >= EPI_STORE && epi <= EPI_ACCUMULATE, "rotating product: epilogue %d", epi);
#define LAUNCH(T, E)                                                                               \
  hipLaunchKernelGGL((k_rot_gemm<T, GEMM_NCT, E>), grid, block, lds, c->stream, A, src, out, n, m, \
                     ldb, aux, nfull)
#define CASE(T)                                                                                    \
  case T:                                                                                          \
    if (epi == EPI_STORE) LAUNCH(T, EPI_STORE);                                                    \
    else if (epi == EPI_HADAMARD) LAUNCH(T, EPI_HADAMARD);                                         \
    else if (epi == EPI_ACCUMULATE) LAUNCH(T, EPI_ACCUMULATE);                                     \
    else LAUNCH(T, EPI_FILL_INTERIOR);                                                             \
    break;
  switch (kq) {
    CASE(5) CASE(9) CASE(15) CASE(17) CASE(25) CASE(33) CASE(41) CASE(49) CASE(57) CASE(65) CASE(69)
  }
#undef CASE
#undef LAUNCH
  VH_LAUNCH_CHECK();
  return VH_OK;
}

// three rotations with A along each axis: src -> t1 -> t2 -> last(epi)
static int apply_all(vh_context *c, const float *A, int n, const float *src, float *t1, float *t2,
                     float *dst, int epi_last, const float *aux, int nfull) {
  const int64_t m = (int64_t)n * n;
  VH_TRY(vh_rotate_gemm(c, A, src, t1, n, m, EPI_STORE, nullptr, 0));
  VH_TRY(vh_rotate_gemm(c, A, t1, t2, n, m, EPI_STORE, nullptr, 0));
  VH_TRY(vh_rotate_gemm(c, A, t2, dst, n, m, epi_last, aux, nfull));
  return VH_OK;
}

// ---------------------------------------------------------------------------
// multipole moments (src/poisson.jl:55-189), FP64
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
    k_moments(const float *__restrict__ c, const double *__restrict__ dual, int n,
              double *__restrict__ part, int nparts) {
  __shared__ double lds[8];
  const double *p0 = dual, *p1 = dual + n, *p2 = dual + 2 * n;
  double acc[10];
#pragma unroll
  for (int v = 0; v < 10; ++v) acc[v] = 0.0;
  const int64_t njk = (int64_t)n * n;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < njk;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int j = (int)(t % n), k = (int)(t / n);
    const double a = p0[j] * p0[k], b = p1[j] * p0[k], cc = p0[j] * p1[k];
    const double d = p2[j] * p0[k], e = p0[j] * p2[k], f = p1[j] * p1[k];
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    const float *col = c + t * n;
    for (int i = 0; i < n; ++i) {
      const double v = col[i];
      s0 += v * p0[i];
      s1 += v * p1[i];
      s2 += v * p2[i];
    }
    acc[0] += s0 * a;
    acc[1] += s1 * a;
    acc[2] += s0 * b;
    acc[3] += s0 * cc;
    acc[4] += s2 * a;
    acc[5] += s0 * d;
    acc[6] += s0 * e;
    acc[7] += s1 * b;
    acc[8] += s1 * cc;
    acc[9] += s0 * f;
  }
  for (int v = 0; v < 10; ++v) {
    const double s = vh_block_sum_d(acc[v], lds);
    if (threadIdx.x == 0) part[(int64_t)v * nparts + blockIdx.x] = s;
  }
}

__global__ void k_reduce_rows(const double *__restrict__ part, int nparts, int nvals,
                              double *__restrict__ out) {
  __shared__ double lds[16];
  for (int v = 0; v < nvals; ++v) {
    double s = 0.0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) s += part[(int64_t)v * nparts + i];
    const double r = vh_block_sum_d(s, lds);
    if (threadIdx.x == 0) out[v] = r;
  }
}

// _multipole (src/poisson.jl:168-189): mp = [q, cx, cy, cz, qxx, qyy, qzz, qxy, qxz, qyz]
__global__ void k_multipole(const double *__restrict__ mom, double *__restrict__ mp) {
  const double q = mom[0], d100 = mom[1], d010 = mom[2], d001 = mom[3];
  const double m200 = mom[4], m020 = mom[5], m002 = mom[6], mxy = mom[7], mxz = mom[8], myz = mom[9];
  double bx = 0.0, by = 0.0, bz = 0.0;
  if (q != 0.0) {
    bx = d100 / q;
    by = d010 / q;
    bz = d001 / q;
  }
  const double quad[6] = {2.0 * m200 - m020 - m002, 2.0 * m020 - m200 - m002,
                          2.0 * m002 - m200 - m020, 3.0 * mxy, 3.0 * mxz, 3.0 * myz};
  const double b2 = bx * bx + by * by + bz * bz;
  const double shift[6] = {q * (3.0 * (bx * bx) - b2), q * (3.0 * (by * by) - b2),
                           q * (3.0 * (bz * bz) - b2), 3.0 * q * bx * by,
                           3.0 * q * bx * bz, 3.0 * q * by * bz};
  mp[0] = q;
  mp[1] = bx;
  mp[2] = by;
  mp[3] = bz;
  for (int i = 0; i < 6; ++i) mp[4 + i] = quad[i] - shift[i];
}

#define MOM_BLOCKS 256

int vh_multipole_impl(vh_mesh *m) {
  vh_context *c = m->ctx;
  hipLaunchKernelGGL(k_moments, dim3(MOM_BLOCKS), dim3(256), 0, c->stream, m->rho, m->dual, m->n,
                     m->partials, MOM_BLOCKS);
  VH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_reduce_rows, dim3(1), dim3(256), 0, c->stream, m->partials, MOM_BLOCKS, 10,
                     m->scal);
  VH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_multipole, dim3(1), dim3(1), 0, c->stream, m->scal, m->scal + 10);
  VH_LAUNCH_CHECK();
  return VH_OK;
}

// total charge sum(rho m0x m0y m0z) in FP64 (src/deposition.jl:213-257)
__global__ void __launch_bounds__(256)
    k_charge(const float *__restrict__ rho, const double *__restrict__ m0, int n,
             double *__restrict__ part) {
  __shared__ double lds[8];
  double acc = 0.0;
  const int64_t njk = (int64_t)n * n;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < njk;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int j = (int)(t % n), k = (int)(t / n);
    double s = 0.0;
    const float *col = rho + t * n;
    for (int i = 0; i < n; ++i) s += (double)col[i] * m0[i];
    acc += s * m0[j] * m0[k];
  }
  const double s = vh_block_sum_d(acc, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

int vh_total_charge_impl(vh_mesh *m, const float *rho, double *dq) {
  vh_context *c = m->ctx;
  hipLaunchKernelGGL(k_charge, dim3(MOM_BLOCKS), dim3(256), 0, c->stream, rho, m->dual, m->n,
                     m->partials);
  VH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_reduce_rows, dim3(1), dim3(256), 0, c->stream, m->partials, MOM_BLOCKS, 1, dq);
  VH_LAUNCH_CHECK();
  return VH_OK;
}

// ---------------------------------------------------------------------------
// faces
// ---------------------------------------------------------------------------
__device__ __forceinline__ int64_t nface_of(int n) {
  return 2 * (int64_t)n * n + (int64_t)(n - 2) * (4 * n - 4);
}

// t-th point of the surface of an n^3 grid: the k = 0 and k = n-1 layers,
// then a ring of 4n - 4 points on each layer in between
__device__ __forceinline__ void face_point(int64_t t, int n, int &i, int &j, int &k) {
  const int64_t cap = (int64_t)n * n;
  if (t < cap) {
    i = (int)(t % n);
    j = (int)(t / n);
    k = 0;
    return;
  }
  t -= cap;
  if (t < cap) {
    i = (int)(t % n);
    j = (int)(t / n);
    k = n - 1;
    return;
  }
  t -= cap;
  const int ring = 4 * n - 4;
  k = (int)(t / ring) + 1;
  int r = (int)(t % ring);
  if (r < n) {
    i = r;
    j = 0;
  } else if (r < 2 * n) {
    i = r - n;
    j = n - 1;
  } else if (r < 3 * n - 2) {
    i = 0;
    j = r - 2 * n + 1;
  } else {
    i = n - 1;
    j = r - (3 * n - 2) + 1;
  }
}

// multipole potential (src/poisson.jl:198-206) on the faces
__global__ void k_boundary_potential(float *__restrict__ phi, const double *__restrict__ mp,
                                     const double *__restrict__ g, int n) {
  const int64_t nf = nface_of(n);
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nf;
       t += (int64_t)gridDim.x * blockDim.x) {
    int i, j, k;
    face_point(t, n, i, j, k);
    const double px = g[i] - mp[1], py = g[j] - mp[2], pz = g[k] - mp[3];
    const double r2 = px * px + py * py + pz * pz;
    const double invr = 1.0 / sqrt(r2);
    const double quad = mp[4] * (px * px) + mp[5] * (py * py) + mp[6] * (pz * pz) +
                        2.0 * (mp[7] * px * py + mp[8] * px * pz + mp[9] * py * pz);
    const double i5 = invr * invr * invr * invr * invr;
    phi[i + (int64_t)n * (j + (int64_t)n * k)] = (float)(mp[0] * invr + quad * i5 / 2.0);
  }
}

// Hermite value basis function `b` (0-based) on knots g (src/splines.jl:74-122)
__device__ __forceinline__ double hermite_value(const double *__restrict__ g, int nk, int b, double x) {
  const int kn = b >> 1, kind = b & 1;
  double a, h, s;
  if (kn > 0 && x <= g[kn]) {
    if (x < g[kn - 1]) return 0.0;
    h = g[kn] - g[kn - 1];
    a = (x - g[kn - 1]) / h;
    s = -1.0;
  } else if (kn < nk - 1 && x >= g[kn]) {
    if (x > g[kn + 1]) return 0.0;
    h = g[kn + 1] - g[kn];
    a = (g[kn + 1] - x) / h;
    s = 1.0;
  } else {
    return 0.0;
  }
  return kind == 0 ? 3.0 * (a * a) - 2.0 * (a * a * a) : h * ((a * a) - (a * a * a)) * s;
}

// cell_index (src/fields.jl:235-238): 0-based interval, -1 outside
__device__ __forceinline__ int cell_index(const double *__restrict__ g, int nk, double x) {
  if (x < g[0] || x > g[nk - 1]) return -1;
  int lo = 0, hi = nk;  // first index with g[idx] >= x
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (g[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  int c = lo - 1;  // searchsortedfirst - 1, 0-based
  return c < 0 ? 0 : c;
}

// spline_potential (src/fields.jl:440-469) of the coarse csol at the fine faces
__global__ void k_boundary_from_coarse(float *__restrict__ phi, const double *__restrict__ gf, int n,
                                       const float *__restrict__ csol, const double *__restrict__ kc,
                                       int nkc, int nc) {
  const int64_t nf = nface_of(n);
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nf;
       t += (int64_t)gridDim.x * blockDim.x) {
    int i, j, k;
    face_point(t, n, i, j, k);
    const double p[3] = {gf[i], gf[j], gf[k]};
    int cell[3];
    bool out = false;
    for (int d = 0; d < 3; ++d) {
      cell[d] = cell_index(kc, nkc, p[d]);
      out = out || cell[d] < 0;
    }
    float res;
    if (out) {
      res = __int_as_float(0x7fc00000);
    } else {
      double v[3][4];
      for (int d = 0; d < 3; ++d)
        for (int a = 0; a < 4; ++a) v[d][a] = hermite_value(kc, nkc, 2 * cell[d] + a, p[d]);
      double phiv = 0.0;
      for (int ak = 0; ak < 4; ++ak) {
        const int kk = 2 * cell[2] + ak;
        for (int aj = 0; aj < 4; ++aj) {
          const int jj = 2 * cell[1] + aj;
          const double w = v[1][aj] * v[2][ak];
          for (int ai = 0; ai < 4; ++ai) {
            const int ii = 2 * cell[0] + ai;
            phiv += (double)csol[ii + (int64_t)nc * (jj + (int64_t)nc * kk)] * v[0][ai] * w;
          }
        }
      }
      res = (float)phiv;
    }
    phi[i + (int64_t)n * (j + (int64_t)n * k)] = res;
  }
}

// right-hand side of del^2 phi = -4 pi rho at the interior points, faces
// lifted through the outer columns of D (src/poisson.jl:275-326)
__global__ void k_poisson_rhs(float *__restrict__ rhs, const float *__restrict__ rho,
                              const float *__restrict__ phi, const float *__restrict__ lap, int n) {
  const int ni = n - 2;
  const int64_t tot = (int64_t)ni * ni * ni;
  const float c = -4.0f * 3.14159265358979323846f;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(t % ni), j = (int)((t / ni) % ni), k = (int)(t / ((int64_t)ni * ni));
    const int64_t nn = n;
    const float dxl = lap[(i + 1) * nn + 0], dxr = lap[(i + 1) * nn + (n - 1)];
    const float dyl = lap[(j + 1) * nn + 0], dyr = lap[(j + 1) * nn + (n - 1)];
    const float dzl = lap[(k + 1) * nn + 0], dzr = lap[(k + 1) * nn + (n - 1)];
    auto at = [&](int64_t a, int64_t b, int64_t cc) { return phi[a + nn * (b + nn * cc)]; };
    const float r = c * rho[(i + 1) + nn * ((j + 1) + nn * (k + 1))] - dxl * at(0, j + 1, k + 1) -
                    dxr * at(n - 1, j + 1, k + 1) - dyl * at(i + 1, 0, k + 1) -
                    dyr * at(i + 1, n - 1, k + 1) - dzl * at(i + 1, j + 1, 0) -
                    dzr * at(i + 1, j + 1, n - 1);
    rhs[t] = r;
  }
}

// solve_interior!(phi, rho, dm) (src/devicemesh.jl:133-146): the faces of phi
// are already set
static int solve_interior(vh_mesh *m) {
  vh_context *c = m->ctx;
  const int ni = m->ni;
  hipLaunchKernelGGL(k_poisson_rhs, dim3(2048), dim3(256), 0, c->stream, m->inner, m->rho, m->phi,
                     m->lap, m->n);
  VH_LAUNCH_CHECK();
  // forward transform, Hadamard fused into the third product
  VH_TRY(apply_all(c, m->Minv_op, ni, m->inner, m->work1, m->work2, m->work1, EPI_HADAMARD,
                   m->invlsum, 0));
  // inverse transform, the copy into phi's interior fused into the third product
  VH_TRY(apply_all(c, m->M_op, ni, m->work1, m->inner, m->work2, m->phi, EPI_FILL_INTERIOR,
                   nullptr, m->n));
  return VH_OK;
}

int vh_poisson_single_impl(vh_mesh *m) {
  vh_context *c = m->ctx;
  VH_TRY(vh_multipole_impl(m));
  hipLaunchKernelGGL(k_boundary_potential, dim3(512), dim3(256), 0, c->stream, m->phi, m->scal + 10,
                     m->colloc64, m->n);
  VH_LAUNCH_CHECK();
  VH_TRY(solve_interior(m));
  return VH_OK;
}

int vh_spline_coefficients_impl(vh_mesh *m, const float *src, float *dst) {
  // src -> dst -> work -> dst, as apply_all_rotating! alternates (src/tensorsolver.jl:243-254)
  VH_TRY(apply_all(m->ctx, m->Sinv_op, m->n, src, dst, m->work, dst, EPI_STORE, nullptr, 0));
  return VH_OK;
}

int vh_poisson_nested_impl(vh_mesh *fine, vh_mesh *coarse) {
  vh_context *c = fine->ctx;
  VH_TRY(vh_poisson_single_impl(coarse));
  // coarse coefficients: the same S^-1 phi that spline_coefficients! of the
  // coarse level computes afterwards, kept in its csol
  VH_TRY(vh_spline_coefficients_impl(coarse, coarse->phi, coarse->csol));
  hipLaunchKernelGGL(k_boundary_from_coarse, dim3(512), dim3(256), 0, c->stream, fine->phi,
                     fine->colloc64, fine->n, coarse->csol, coarse->knots64, coarse->nk, coarse->n);
  VH_LAUNCH_CHECK();
  VH_TRY(solve_interior(fine));
  return VH_OK;
}

// ---------------------------------------------------------------------------
// mean field (src/meanfield.jl:116-147): V_xc(rho) + V_jel(r) at the points
// ---------------------------------------------------------------------------
__global__ void k_effective(float *__restrict__ extra, const float *__restrict__ rho,
                            const float *__restrict__ g, int n, float nions, float radius,
                            const float *__restrict__ add) {
  const float xc_ex = (float)(-0.98474502184269641);          // -cbrt(3/pi)
  const float xc_co = (float)(-0.0333);
  const float xc_sc = (float)(11.4 * 1.6119919540164696);     // 11.4 cbrt(4 pi / 3)
  const int64_t tot = (int64_t)n * n * n;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(t % n), j = (int)((t / n) % n), k = (int)(t / ((int64_t)n * n));
    const float r = sqrtf(g[i] * g[i] + g[j] * g[j] + g[k] * g[k]);
    const float cr = cbrtf(rho[t]);
    const float vxc = xc_ex * cr + xc_co * log1pf(xc_sc * cr);
    const float rr = r / radius;
    const float vj = r < radius ? -nions * (3.0f - rr * rr) / (2.0f * radius) : -nions / r;
    extra[t] = add ? add[t] + (vxc + vj) : vxc + vj;
  }
}

int vh_effective_potential_impl(vh_mesh *m, double nions, double radius) {
  vh_context *c = m->ctx;
  hipLaunchKernelGGL(k_effective, dim3(2048), dim3(256), 0, c->stream, m->extra, m->rho, m->colloc,
                     m->n, (float)nions, (float)radius, (const float *)nullptr);
  VH_LAUNCH_CHECK();
  // extra -> work -> extra -> csol (+=)
  VH_TRY(apply_all(c, m->Sinv_op, m->n, m->extra, m->work, m->extra, m->csol, EPI_ACCUMULATE,
                   nullptr, 0));
  return VH_OK;
}

// spline_coefficients! then effective_potential! as one transform:
// csol = S^-1 (phi + V_xc(rho) + V_jel), three products instead of six
int vh_potential_coefficients_impl(vh_mesh *m, double nions, double radius) {
  vh_context *c = m->ctx;
  hipLaunchKernelGGL(k_effective, dim3(2048), dim3(256), 0, c->stream, m->extra, m->rho, m->colloc,
                     m->n, (float)nions, (float)radius, (const float *)m->phi);
  VH_LAUNCH_CHECK();
  VH_TRY(apply_all(c, m->Sinv_op, m->n, m->extra, m->work, m->extra, m->csol, EPI_STORE, nullptr, 0));
  return VH_OK;
}
