// Smoothed field at the pseudo-particles with the projectile fused in, the
// out-of-stencil fallback, and the packed position Verlet.
//
// Field (src/kernels.jl:339-470).  One wave per work unit (at most VH_UNIT
// particles of one occupied cell), two variants:
//   k_field_mfma (default): the first stage of the 10^3 contraction on the
//     matrix cores, 16 particles per product (see the kernel);
//   k_field (exact order): each lane takes one particle, keeps 40 of its
//     tabulated overlap/gradient values in registers and runs
//     contract_spline_10 (src/fields.jl:305-326) in the reference's fma order
//     on the cell's 10^3 block of csol, read at wave-uniform addresses (scalar
//     loads feeding the FMAs from SGPRs).
// The projectile's per-particle reaction and pair potential are summed per
// lane in FP64 and reduced in a fixed order.
//
// Fallback (src/accelerator.jl:679-731).  The reference downloads the whole
// force buffer and loops on the host over the particles whose stencil leaves
// the fine grid; here one kernel walks the out-of-stencil list on the device
// in FP64: the unsmoothed spline field of the coarse csol
// (src/fields.jl:358-393) and the projectile's own law.
//
// Verlet (src/kernels.jl:985-1025, src/particles.jl:337-349): integer cell
// part exact, FP32 offsets, diagnostics summed in FP64.
#include "vh_internal.h"

#include <cmath>

// ---------------------------------------------------------------------------
// projectile laws (src/projectile.jl:79-108), FP32 as in the device kernel
// ---------------------------------------------------------------------------
__device__ __forceinline__ float gaussian_force_kernel_f(float r2, float s) {
  const float u2 = r2 / (s * s);
  if (u2 <= 0.25f) {
    const float c[7] = {(float)(1.0 / 3.0), (float)(-1.0 / 10.0), (float)(1.0 / 56.0),
                        (float)(-1.0 / 432.0), (float)(1.0 / 4224.0), (float)(-1.0 / 49920.0),
                        (float)(1.0 / 685440.0)};
    float p = 0.f;
    for (int k = 6; k >= 0; --k) p = c[k] + u2 * p;
    return (float)0.79788456080286535588 * p / (s * s * s);
  }
  const float r = sqrtf(r2);
  const float x = r / (sqrtf(2.0f) * s);
  return (erff(x) - (float)0.79788456080286535588 * (r / s) * expf(-u2 / 2.0f)) / (r2 * r);
}

// The Gaussian law and, with `pot`, the pair potential erf(r / (sqrt2 s)) / r
// of src/kernels.jl:425-443 for the matrix-core kernel: one erf serves both,
// and the quotients by s are products by 1/s (agreement to FP32 rounding).
__device__ __forceinline__ void gaussian_law_f(float r2, float r, float s, int pot, float &kern,
                                               float &pe) {
  const float is = 1.0f / s;
  const float u2 = r2 * (is * is);
  const float x = r * (is * 0.70710678118654752f);
  const bool near = u2 <= 0.25f;
  const float e = (pot || !near) ? erff(x) : 0.f;
  if (near) {
    const float c[7] = {(float)(1.0 / 3.0), (float)(-1.0 / 10.0), (float)(1.0 / 56.0),
                        (float)(-1.0 / 432.0), (float)(1.0 / 4224.0), (float)(-1.0 / 49920.0),
                        (float)(1.0 / 685440.0)};
    float p = 0.f;
    for (int k = 6; k >= 0; --k) p = c[k] + u2 * p;
    kern = (float)0.79788456080286535588 * p * (is * is * is);
  } else {
    kern = (e - (float)0.79788456080286535588 * (r * is) * expf(-0.5f * u2)) / (r2 * r);
  }
  pe = 0.f;
  if (pot) pe = r < 1.0e-4f * s ? (float)0.79788456080286535588 * is : e / r;
}

__device__ __forceinline__ float ball_force_kernel_f(float r2, float R) {
  return r2 > R * R ? 1.0f / (r2 * sqrtf(r2)) : 1.0f / (R * R * R);
}

// FP64 versions for the fallback (the reference runs it on the host in T)
__device__ __forceinline__ double gaussian_force_kernel_d(double r2, double s) {
  const double u2 = r2 / (s * s);
  if (u2 <= 0.25) {
    const double c[7] = {1.0 / 3.0, -1.0 / 10.0, 1.0 / 56.0, -1.0 / 432.0,
                         1.0 / 4224.0, -1.0 / 49920.0, 1.0 / 685440.0};
    double p = 0.0;
    for (int k = 6; k >= 0; --k) p = c[k] + u2 * p;
    return 0.79788456080286535588 * p / (s * s * s);
  }
  const double r = sqrt(r2);
  const double x = r / (sqrt(2.0) * s);
  return (erf(x) - 0.79788456080286535588 * (r / s) * exp(-u2 / 2.0)) / (r2 * r);
}

__device__ __forceinline__ double ball_force_kernel_d(double r2, double R) {
  return r2 > R * R ? 1.0 / (r2 * sqrt(r2)) : 1.0 / (R * R * R);
}

// ---------------------------------------------------------------------------
// field kernel
// ---------------------------------------------------------------------------
struct vh_proj_args {
  float px, py, pz;   // E.(projectile.position)
  float coef;         // E(-w q), 0 without projectile
  float sigma;        // E(scale(softening))
  int ball;           // VH_SOFT_BALL
  float radius;       // ball radius (ball mode)
  const double *dev;  // device-resident projectile state (position at 0..2), or null
  int pot;            // the pair-potential sum is wanted (matrix-core kernel)
};

// ---------------------------------------------------------------------------
// packed Verlet (src/kernels.jl:985-1025), one axis: the exact FP32 sequence
// ---------------------------------------------------------------------------
__device__ __forceinline__ void verlet1(int kq, float dq, int ok, float od, float a, float f, float h,
                                        int &kn, float &dn, float &gap) {
  kn = 2 * kq - ok;
  dn = 2.0f * dq - od + a * f;
  const int m = (int)rintf(dn / h);
  kn += m;
  dn -= (float)m * h;
  gap = (float)(kn - ok) * h + (dn - od);
}

// The push of a particle by the kernel that computed its force (one step
// without diagnostics): the same verlet1 sequence as k_verlet.  The new
// record goes to the current half, the old one to the previous half, and
// the new cell key to keys[] for the next sort (which then counts its first
// digit from the keys).
struct vh_push_args {
  int on;              // 0: forces only
  float a, h;          // dt^2 / mass, the cell width
  int nk;              // cells per axis (the key)
  uint32_t *keys;      // new cell key per slot
  int32_t *dst;        // the buffer the pushed records go to, in slot order
  float x0, lo, hi;    // the deposit's domain: pushed records outside it are
  int32_t *nout;       // counted here
};

// the deposit's rejection test of a record (k_permute's, same arithmetic)
__device__ __forceinline__ bool push_outside(const vh_pos &r, const vh_push_args &pu) {
  bool inside = true;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float p = pu.x0 + (float)(r.k[d] - 1) * pu.h + r.d[d];
    inside = inside && (pu.lo <= p) && (p <= pu.hi);
  }
  return !inside;
}

__device__ __forceinline__ uint32_t push_record(const vh_pos &cur, const vh_pos &old, float f0,
                                               float f1, float f2, const vh_push_args &pu,
                                               vh_pos &nw) {
  float g1, g2, g3;
  verlet1(cur.k[0], cur.d[0], old.k[0], old.d[0], pu.a, f0, pu.h, nw.k[0], nw.d[0], g1);
  verlet1(cur.k[1], cur.d[1], old.k[1], old.d[1], pu.a, f1, pu.h, nw.k[1], nw.d[1], g2);
  verlet1(cur.k[2], cur.d[2], old.k[2], old.d[2], pu.a, f2, pu.h, nw.k[2], nw.d[2], g3);
  return cell_key0(nw.k[0], nw.k[1], nw.k[2], pu.nk);
}

// the record of lane `lane` through a buffer of records (vh_rsrc)
__device__ __forceinline__ void vh_bstore_pos(__amdgpu_buffer_rsrc_t r, int lane, const vh_pos &v) {
  const int o = 24 * lane;
  __builtin_amdgcn_raw_buffer_store_b64((vh_i2){v.k[0], v.k[1]}, r, o, 0, 0);
  __builtin_amdgcn_raw_buffer_store_b64((vh_i2){v.k[2], __float_as_int(v.d[0])}, r, o + 8, 0, 0);
  __builtin_amdgcn_raw_buffer_store_b64((vh_i2){__float_as_int(v.d[1]), __float_as_int(v.d[2])}, r,
                                        o + 16, 0, 0);
}

__global__ void __launch_bounds__(64)
    k_field(const int32_t *__restrict__ P, int64_t npad, const float *__restrict__ csol, int n,
            const float *__restrict__ ovl, const float *__restrict__ grad,
            const vh_unit *__restrict__ units, const int32_t *__restrict__ cnt, int nk, float x0,
            float h, float spacing, float nbdt, float w, int ncol, vh_proj_args pa,
            float *__restrict__ force,
            double *__restrict__ part) {
  const int t = threadIdx.x;
  if (pa.dev) {
    pa.px = (float)pa.dev[0];
    pa.py = (float)pa.dev[1];
    pa.pz = (float)pa.dev[2];
  }
  const float half = spacing * 0.5f;
  const float nan = __int_as_float(0x7fc00000);
  const int nunits = cnt[CNT_NUNITS];
  const int64_t nn = n;
  double sax = 0.0, say = 0.0, saz = 0.0, sae = 0.0;
  for (int u = blockIdx.x; u < nunits; u += gridDim.x) {
    const vh_unit un = units[u];
    const int key = un.key;
    const int kx = key % nk + 1, ky = (key / nk) % nk + 1, kz = key / (nk * nk) + 1;
    const int bx = 2 * kx - 5, by = 2 * ky - 5, bz = 2 * kz - 5;  // 1-based stencil corner
    const bool ok = bx >= 1 && by >= 1 && bz >= 1 && bx + 9 <= n && by + 9 <= n && bz + 9 <= n;
    // the cell's 10^3 block of csol, read at wave-uniform addresses
    const float *tb = csol + (ok ? (bx - 1) + nn * ((by - 1) + nn * (bz - 1)) : 0);
    const int lo = un.lo, hi = un.hi;
    for (int s = lo + t; s < hi; s += 64) {
      const vh_pos rp = vh_load_pos(P, s);
      const float dx0 = rp.d[0], dy0 = rp.d[1], dz0 = rp.d[2];
      float fxp = nan, fyp = nan, fzp = nan;
      if (ok) {
        int cx = (int)floorf((dx0 + half) / spacing * nbdt + 0.5f) + 1;
        int cy = (int)floorf((dy0 + half) / spacing * nbdt + 0.5f) + 1;
        int cz = (int)floorf((dz0 + half) / spacing * nbdt + 0.5f) + 1;
        cx = min(max(cx, 1), ncol) - 1;
        cy = min(max(cy, 1), ncol) - 1;
        cz = min(max(cz, 1), ncol) - 1;
        float ox[10], gx[10], oy[10], gy[10];
#pragma unroll
        for (int a = 0; a < 10; ++a) {
          ox[a] = ovl[cx * 10 + a];
          gx[a] = grad[cx * 10 + a];
          oy[a] = ovl[cy * 10 + a];
          gy[a] = grad[cy * 10 + a];
        }
        float fx = 0.f, fy = 0.f, fz = 0.f;
        for (int kk = 0; kk < 10; ++kk) {
          const float ozk = ovl[cz * 10 + kk], gzk = grad[cz * 10 + kk];
          const float *plane = tb + nn * nn * kk;
#pragma unroll
          for (int jj = 0; jj < 10; ++jj) {
            const float *row = plane + nn * jj;
            float dxp = 0.f, val = 0.f;
#pragma unroll
            for (int ii = 0; ii < 10; ++ii) {
              const float c = row[ii];
              dxp = __builtin_fmaf(c, gx[ii], dxp);
              val = __builtin_fmaf(c, ox[ii], val);
            }
            fx = __builtin_fmaf(oy[jj] * ozk, dxp, fx);
            fy = __builtin_fmaf(gy[jj] * ozk, val, fy);
            fz = __builtin_fmaf(oy[jj] * gzk, val, fz);
          }
        }
        fxp = -w * fx;
        fyp = -w * fy;
        fzp = -w * fz;
      }
      if (pa.coef != 0.f) {
        float px, py, pz;
        if (pa.ball) {  // the particle's own position
          px = x0 + (float)(rp.k[0] - 1) * h + dx0;
          py = x0 + (float)(rp.k[1] - 1) * h + dy0;
          pz = x0 + (float)(rp.k[2] - 1) * h + dz0;
        } else {        // rebuilt from the cell, as the reference's kernel does
          px = x0 + (float)(kx - 1) * h + dx0;
          py = x0 + (float)(ky - 1) * h + dy0;
          pz = x0 + (float)(kz - 1) * h + dz0;
        }
        const float dx = pa.px - px, dy = pa.py - py, dz = pa.pz - pz;
        const float d2 = dx * dx + dy * dy + dz * dz;
        const float mm = pa.coef * (pa.ball ? ball_force_kernel_f(d2, pa.radius)
                                            : gaussian_force_kernel_f(d2, pa.sigma));
        const float fx2 = mm * dx, fy2 = mm * dy, fz2 = mm * dz;
        if (ok) {
          fxp -= fx2;
          fyp -= fy2;
          fzp -= fz2;
        }
        const float r = sqrtf(d2);
        sax += fx2;
        say += fy2;
        saz += fz2;
        if (pa.ball) {
          const float R = pa.radius;
          sae += r < R ? -(3.0f - (r / R) * (r / R)) / (2.0f * R) : -1.0f / r;
        } else {
          sae += r < 1.0e-4f * pa.sigma ? 0.7978845608f / pa.sigma
                                        : erff(r / 1.4142135624f / pa.sigma) / r;
        }
      }
      force[s] = fxp;
      force[npad + s] = fyp;
      force[2 * npad + s] = fzp;
    }
  }
  // fixed-order reduction: wave, then one partial per block
  sax = vh_wave_sum(sax);
  say = vh_wave_sum(say);
  saz = vh_wave_sum(saz);
  sae = vh_wave_sum(sae);
  if (t == 0) {
    part[blockIdx.x] = sax;
    part[gridDim.x + blockIdx.x] = say;
    part[2 * gridDim.x + blockIdx.x] = saz;
    part[3 * gridDim.x + blockIdx.x] = sae;
  }
}

// ---------------------------------------------------------------------------
// field kernel on the matrix cores
// ---------------------------------------------------------------------------
// The contraction F = sum c[ii,jj,kk] X[ii] Y[jj] Z[kk] in three stages:
//   stage 1 (MFMA): T_o[p, r] = sum_ii c[ii, r] ox[p, ii] and T_g[p, r] with
//     gx, for the 100 rows r = (jj, kk) of the cell's block -- a product of
//     the block (rows x ii, the same for every particle of the cell) by the
//     particles' x columns (ii x particles), 16 particles at a time;
//   stage 2 (VALU): per particle, the y contraction of each kk plane;
//   stage 3 (VALU): the z contraction, then the sum over the four lanes that
//     share a particle.
// v_mfma_f32_16x16x4_f32: lane l supplies A[l%16][l/16] and B[l/16][l%16] and
// receives D[4(l/16)+v][l%16], v = 0..3.  Rows are laid out so that the four
// D registers and 7 tiles of a lane hold, for its particle, the slots
// q = 4t + v of its lane group g = l/16: q < 10 is (jj = q, kk = g), q < 20 is
// (jj = q - 10, kk = g + 4), q < 25 is (jj = q - 20 + 5 (g & 1),
// kk = 8 + g / 2), q >= 25 is padding: each group takes two planes of kk and
// half of plane 8 or 9, 25 rows in 28 slots.  The accumulator of a slot is
// known at compile time; only the last half plane's y weights and the z
// weights depend on the group.  K = ii runs over 3 steps of 4 (10 of 12
// used), rows over 7 tiles of 16 (100 of 112 used).
// The sums run in another order than the reference's FMA chain, so the result
// agrees to FP32 rounding, not to the bit; k_field keeps the exact order.
//
// Per chunk of 64 particles: phase A (a lane per particle for the offsets,
// five lanes per particle for the table loads) stages the x, y and z
// columns of each particle in LDS, each in a row of 20 floats (overlap, then
// gradient) of its own region, twelve rows a 1 KiB block (ft_row: rows of
// the 16 particles of a sub-group fall at most two to a bank group); phase
// B runs the four sub-groups of 16 particles; phase C stores.
#define FT_REGION 1280

// the LDS row (20 floats; FT_REGION floats a region) of the chunk's particle p
__device__ __forceinline__ int ft_row(int p) { return 20 * p; }

// 16 bytes per lane from global memory to LDS: lane l writes bytes
// 16 l .. 16 l + 15 of the 1 KiB block at lds (a wave-uniform address)
__device__ __forceinline__ void ft_dma16(const void *src, float *lds) {
  __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void *)lds, 16, 0, 0);
}

typedef float ft_f4 __attribute__((ext_vector_type(4)));

// B operands of one sub-group for the lane group `grp`, from its particle's
// LDS row: {ox, ox, ox, gx, gx, gx} at ii = 4k + grp (0 beyond 9).  Every
// read is unconditional (index 8 + grp exists for grp < 2 only, the other
// lanes read an in-bounds stand-in and select 0): conditional reads would
// become exec-masked branches, each waiting for all loads in flight.
__device__ __forceinline__ void ft_bops(const float *row, int grp, float b[6]) {
  const bool top = grp < 2;
  const int gt = top ? grp : 0;
  const float b2 = row[8 + gt], b5 = row[18 + gt];
  b[0] = row[grp];
  b[1] = row[4 + grp];
  b[2] = top ? b2 : 0.f;
  b[3] = row[10 + grp];
  b[4] = row[14 + grp];
  b[5] = top ? b5 : 0.f;
}

// The contraction of one sub-group, pipelined over the 7 row tiles: tile
// t's products (T_o and T_g chains interleaved over the three K steps) are
// issued, then the y contraction of tile t - 1, whose products are complete,
// runs on the VALU while tile t's are in flight.  Slot q = 4t + v of a lane
// adds into accumulator m = q / 10 (planes g, g + 4 and the half plane).
// Then the z contraction and the sum over the four lane groups of a
// particle: every lane of the particle returns -w F.  rx, ry, rz: the
// particle's x, y and z columns in LDS (overlap, then gradient).
template <int NS>
__device__ __forceinline__ ft_f4 ft_contract(const float A[7][3], const float *rx, const float *ry,
                                             const float *rz, int grp, float w) {
  static_assert(NS == 1, "one sub-group at a time");
  const int half = 5 * (grp & 1), hp = grp >> 1;
  float b[NS][6], y[NS][20], yh[NS][10], z[NS][6];
#pragma unroll
  for (int u = 0; u < NS; ++u) {
    ft_bops(rx, grp, b[u]);
#pragma unroll
    for (int a = 0; a < 20; ++a) y[u][a] = ry[a];  // oy, then gy
    // the half plane's y weights: oy and gy at jj = half .. half + 4
#pragma unroll
    for (int a = 0; a < 5; ++a) {
      yh[u][a] = ry[half + a];
      yh[u][5 + a] = ry[10 + half + a];
    }
    z[u][0] = rz[grp];
    z[u][1] = rz[4 + grp];
    z[u][2] = rz[8 + hp];
    z[u][3] = rz[10 + grp];
    z[u][4] = rz[14 + grp];
    z[u][5] = rz[18 + hp];
  }
  float ux[NS][3], uy[NS][3], uz[NS][3];
#pragma unroll
  for (int u = 0; u < NS; ++u)
#pragma unroll
    for (int m = 0; m < 3; ++m) ux[u][m] = uy[u][m] = uz[u][m] = 0.f;
  ft_f4 To[NS][7], Tg[NS][7];
#pragma unroll
  for (int t = 0; t <= 7; ++t) {
    if (t < 7) {
#pragma unroll
      for (int u = 0; u < NS; ++u) {
        To[u][t] = ft_f4{0.f, 0.f, 0.f, 0.f};
        Tg[u][t] = ft_f4{0.f, 0.f, 0.f, 0.f};
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) {
#pragma unroll
        for (int u = 0; u < NS; ++u) {
          To[u][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(A[t][k], b[u][k], To[u][t], 0, 0, 0);
          Tg[u][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(A[t][k], b[u][3 + k], Tg[u][t], 0, 0, 0);
        }
      }
    }
    if (t >= 1) {
      const int tp = t - 1;
#pragma unroll
      for (int u = 0; u < NS; ++u) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int q = 4 * tp + v;
          if (q < 20) {
            const int m = q / 10, jj = q % 10;
            ux[u][m] = __builtin_fmaf(y[u][jj], Tg[u][tp][v], ux[u][m]);
            uy[u][m] = __builtin_fmaf(y[u][10 + jj], To[u][tp][v], uy[u][m]);
            uz[u][m] = __builtin_fmaf(y[u][jj], To[u][tp][v], uz[u][m]);
          } else if (q < 25) {
            const int i = q - 20;
            ux[u][2] = __builtin_fmaf(yh[u][i], Tg[u][tp][v], ux[u][2]);
            uy[u][2] = __builtin_fmaf(yh[u][5 + i], To[u][tp][v], uy[u][2]);
            uz[u][2] = __builtin_fmaf(yh[u][i], To[u][tp][v], uz[u][2]);
          }
        }
      }
    }
  }
  // schedule: the LDS reads, tile 0's products, then one product per two
  // VALU instructions (left to itself the compiler issues all the products
  // back to back and the y contraction after them)
  __builtin_amdgcn_sched_group_barrier(0x100, 16 * NS, 0);  // DS reads
  __builtin_amdgcn_sched_group_barrier(0x008, 6 * NS, 0);   // MFMA
#pragma unroll
  for (int i = 0; i < 36 * NS; ++i) {
    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
    __builtin_amdgcn_sched_group_barrier(0x002, 2, 0);  // VALU
  }
#pragma unroll
  for (int u = 0; u < NS; ++u) {
    float fx = 0.f, fy = 0.f, fz = 0.f;
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      fx = __builtin_fmaf(z[u][m], ux[u][m], fx);
      fy = __builtin_fmaf(z[u][m], uy[u][m], fy);
      fz = __builtin_fmaf(z[u][3 + m], uz[u][m], fz);
    }
    fx = vh_sum_xor16(fx);
    fy = vh_sum_xor16(fy);
    fz = vh_sum_xor16(fz);
    fx = vh_sum_xor32(fx);
    fy = vh_sum_xor32(fy);
    fz = vh_sum_xor32(fz);
    return ft_f4{-w * fx, -w * fy, -w * fz, 0.f};
  }
}

// ---- the fp16 split variant of the first stage (default) ----
// One v_mfma_f32_16x16x32_f16 per row tile and kind replaces the three FP32
// K steps: its 32 K slots hold, for the ten x columns ii, c_hi w_hi,
// c_hi w_lo and c_lo w_hi (hi = fp16(v), lo = fp16(v - hi)), in the order of
// vh_s16_slot.  The cell's block is scaled by the power of two that puts its
// largest |c| in [2^14, 2^15) before the split, the x columns (a table split
// once on the host, xsplit) by their own; the results are scaled back in the
// final product by -w.  Lane l supplies A[l%16][8(l/16) + j] and
// B[8(l/16) + j][l%16], j = 0..7, and receives D as the FP32 product does,
// so stages 2 and 3 are unchanged.  The dropped lo*lo term and the
// residuals of the splits leave about 2^-21 of the largest term per
// product.
//
// LDS per chunk: the y and z rows as in the FP32 variant, and the x splits
// in eight 1 KiB blocks: block 2 s + h holds, at 16 (16 g + j),
// the 16 bytes of the B operand of lane group g for particle 16 s + j
// (h = 0 overlap, 1 gradient), so that a sub-group reads its operands at
// 16 consecutive addresses per lane group.

typedef _Float16 ft_h8 __attribute__((ext_vector_type(8)));
typedef unsigned int ft_u4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t ft_h2(_Float16 a, _Float16 b) {
  return (uint32_t)__builtin_bit_cast(uint16_t, a) | ((uint32_t)__builtin_bit_cast(uint16_t, b) << 16);
}

// A operand of one row from its scaled c at ii = 2 grp, 2 grp + 1 and
// 8 + grp / 2, in the K order of vh_s16_slot
__device__ __forceinline__ ft_u4 ft_split_row(float va, float vb, float vc, int grp) {
  const _Float16 ha = (_Float16)va, hb = (_Float16)vb, hc = (_Float16)vc;
  const _Float16 la = (_Float16)(va - (float)ha), lb = (_Float16)(vb - (float)hb);
  const _Float16 lc = (_Float16)(vc - (float)hc);
  // both forms of the last word, selected by a mask (a select here became a
  // branch around the lo computation)
  const uint32_t odd = 0u - (uint32_t)(grp & 1);
  const uint32_t w3 = (ft_h2(lc, (_Float16)0.f) & odd) | (ft_h2(hc, hc) & ~odd);
  return ft_u4{ft_h2(ha, ha), ft_h2(la, hb), ft_h2(hb, lb), w3};
}

__device__ __forceinline__ ft_f4 ft_mfma16(ft_u4 a, ft_u4 b) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(ft_h8, a), __builtin_bit_cast(ft_h8, b),
                                                ft_f4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
}

// One sub-group of 16 particles: tile t's two products, then the y
// contraction of tile t - 1 on the VALU (the slot layout of ft_contract);
// the z contraction, the sum over the four lane groups, and the field times
// the scale factors mx (x component) and myz, returned by every lane of the
// particle.
__device__ __forceinline__ ft_f4 ft_contract16(const ft_u4 A[7], const float *ry, const float *rz,
                                               const float *bj, int grp, float mx, float myz) {
  const int half = 5 * (grp & 1), hp = grp >> 1;
  const ft_u4 bo = *(const ft_u4 *)bj;
  const ft_u4 bg = *(const ft_u4 *)(bj + 256);
  float y[20], yh[10], z[6];
#pragma unroll
  for (int a = 0; a < 20; ++a) y[a] = ry[a];
#pragma unroll
  for (int a = 0; a < 5; ++a) {
    yh[a] = ry[half + a];
    yh[5 + a] = ry[10 + half + a];
  }
  z[0] = rz[grp];
  z[1] = rz[4 + grp];
  z[2] = rz[8 + hp];
  z[3] = rz[10 + grp];
  z[4] = rz[14 + grp];
  z[5] = rz[18 + hp];
  float ux[3] = {0.f, 0.f, 0.f}, uy[3] = {0.f, 0.f, 0.f}, uz[3] = {0.f, 0.f, 0.f};
  ft_f4 To[7], Tg[7];
#pragma unroll
  for (int t = 0; t <= 7; ++t) {
    if (t < 7) {
      To[t] = ft_mfma16(A[t], bo);
      Tg[t] = ft_mfma16(A[t], bg);
    }
    if (t >= 1) {
      const int tp = t - 1;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int q = 4 * tp + v;
        if (q < 20) {
          const int m = q / 10, jj = q % 10;
          ux[m] = __builtin_fmaf(y[jj], Tg[tp][v], ux[m]);
          uy[m] = __builtin_fmaf(y[10 + jj], To[tp][v], uy[m]);
          uz[m] = __builtin_fmaf(y[jj], To[tp][v], uz[m]);
        } else if (q < 25) {
          const int i = q - 20;
          ux[2] = __builtin_fmaf(yh[i], Tg[tp][v], ux[2]);
          uy[2] = __builtin_fmaf(yh[5 + i], To[tp][v], uy[2]);
          uz[2] = __builtin_fmaf(yh[i], To[tp][v], uz[2]);
        }
      }
    }
  }
  float fx = 0.f, fy = 0.f, fz = 0.f;
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    fx = __builtin_fmaf(z[m], ux[m], fx);
    fy = __builtin_fmaf(z[m], uy[m], fy);
    fz = __builtin_fmaf(z[3 + m], uz[m], fz);
  }
  fx = vh_sum_xor16(fx);
  fy = vh_sum_xor16(fy);
  fz = vh_sum_xor16(fz);
  fx = vh_sum_xor32(fx);
  fy = vh_sum_xor32(fy);
  fz = vh_sum_xor32(fz);
  return ft_f4{mx * fx, myz * fy, myz * fz, 0.f};
}

// The table columns of a chunk's particles (the three columns of lane l's
// particle packed in `packed`, 10 bits each) go from the caches straight to
// LDS with 16-byte LDS DMA loads (lane l of a load writes bytes
// 16 l .. 16 l + 15 of a 1 KiB block), a line at a time: a column's 80
// bytes (overlap, then gradient) sit in one 128-byte line, fetched as five
// 16-byte pieces by five consecutive lanes (slot 64 i + l of load i is
// piece k % 5 of particle k / 5: 80-byte rows, five loads an axis), and a
// column's fp16 splits fill one line, fetched by four lanes a half (block
// 2 s + h: particle 16 s + l % 16, piece 4 h + l / 16).  A load then touches
// 13 or 16 cache lines where a lane per particle would touch 64 (a scattered
// load costs the cache about a cycle per line), and holds no register.
template <bool S16>
__device__ __forceinline__ void ft_stage(int packed, const float4 *og, const ft_u4 *xsplit, float *tx,
                                         float *ty, float *tz, int l) {
  const ft_f4 *og4 = (const ft_f4 *)og;
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    const int k = 64 * i + l, pk = k / 5, a = k - 5 * pk;
    const int c = __shfl(packed, pk, 64);
    if constexpr (!S16) ft_dma16(og4 + 8 * (c & 1023) + a, tx + i * 256);
    ft_dma16(og4 + 8 * ((c >> 10) & 1023) + a, ty + i * 256);
    ft_dma16(og4 + 8 * (c >> 20) + a, tz + i * 256);
  }
  if constexpr (S16) {
#pragma unroll
    for (int s4 = 0; s4 < 4; ++s4) {
      const int c = __shfl(packed, 16 * s4 + (l & 15), 64) & 1023;
      ft_dma16(xsplit + 8 * c + (l >> 4), tx + (2 * s4) * 256);
      ft_dma16(xsplit + 8 * c + 4 + (l >> 4), tx + (2 * s4 + 1) * 256);
    }
  }
}

// the three table columns of a particle (src/kernels.jl:414-419), packed
__device__ __forceinline__ int ft_columns(const vh_pos &r, float half, float spacing, float nbdt, int ncol) {
  int c[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int cc = (int)floorf((r.d[d] + half) / spacing * nbdt + 0.5f) + 1;
    c[d] = min(max(cc, 1), ncol) - 1;
  }
  return c[0] | (c[1] << 10) | (c[2] << 20);  // columns < 1024
}

template <bool S16>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2)))
    k_field_mfma(const int32_t *__restrict__ P, const uint32_t *__restrict__ perm, int64_t npad,
                 const float *__restrict__ csol, int n,
                 const float4 *__restrict__ og, const ft_u4 *__restrict__ xsplit, float xs_inv_o,
                 float xs_inv_g, const float *__restrict__ zero,
                 const vh_unit *__restrict__ units, const int32_t *__restrict__ cnt, int nk,
                 float x0, float h, float spacing, float nbdt, float w, int ncol,
                 vh_proj_args pa, vh_push_args pu, float *__restrict__ force,
                 double *__restrict__ part) {
  // slot s's record is record perm[s] of P (both arrays); the push writes
  // slot s of pu.dst
  const int32_t *Q = vh_prev_half(P, npad);
  int32_t *Pd = pu.dst, *Qd = vh_prev_half(pu.dst, npad);
  int nout = 0;
  typedef ft_f4 f4;
  if (pa.dev) {
    pa.px = (float)pa.dev[0];
    pa.py = (float)pa.dev[1];
    pa.pz = (float)pa.dev[2];
  }
  // column regions of a chunk: x (S16: the splits, eight 1 KiB blocks), y, z
  constexpr int xreg = S16 ? 2048 : FT_REGION;
  __shared__ __attribute__((aligned(16))) float tab[xreg + 2 * FT_REGION];
  float *const tx = tab, *const ty = tab + xreg, *const tz = ty + FT_REGION;
  const int l = threadIdx.x;
  const int j16 = l & 15, grp = l >> 4;
  const float half = spacing * 0.5f;
  const float nan = __int_as_float(0x7fc00000);
  const int nunits = cnt[CNT_NUNITS];
  const int64_t nn = n;
  double sax = 0.0, say = 0.0, saz = 0.0, sae = 0.0;
  for (int u = blockIdx.x; u < nunits; u += gridDim.x) {
    const vh_unit un = units[u];
    const int key = un.key;
    const int kx = key % nk + 1, ky = (key / nk) % nk + 1, kz = key / (nk * nk) + 1;
    const int bx = 2 * kx - 5, by = 2 * ky - 5, bz = 2 * kz - 5;
    const bool ok = bx >= 1 && by >= 1 && bz >= 1 && bx + 9 <= n && by + 9 <= n && bz + 9 <= n;
    const int lo = un.lo, hi = un.hi;
    // A operands of the cell: row j16 of tile t (slot q = 4 t + j16 % 4 of
    // lane group g = j16 / 4), k = ii = 4 s + grp
    // (split variant: k = ii = 2 grp, 2 grp + 1, 8 + grp / 2 for s = 0, 1, 2)
    float A[7][3];
    ft_u4 A16[7];
    float mx = 0.f, myz = 0.f;
    if (ok) {
      const float *tb = csol + (bx - 1) + nn * ((by - 1) + nn * (bz - 1));
      const int g = j16 >> 2;
#pragma unroll
      for (int t = 0; t < 7; ++t) {
        const int q = 4 * t + (j16 & 3);
        const int jj = q < 10 ? q : q < 20 ? q - 10 : q - 20 + 5 * (g & 1);
        const int kk = q < 10 ? g : q < 20 ? g + 4 : 8 + (g >> 1);
#pragma unroll
        for (int s = 0; s < 3; ++s) {
          const int ii = S16 ? (s < 2 ? 2 * grp + s : 8 + (grp >> 1)) : 4 * s + grp;
          const bool in = q < 25 && ii < 10;
          // padding reads a zero word: the loaded value is used as it is, so
          // the loads stay in flight until the first product
          A[t][s] = *(in ? tb + ii + nn * (jj + nn * kk) : zero);
        }
      }
      if constexpr (S16) {
        // the block's largest |c| (every value is loaded by some lane), its
        // power-of-two scale, and the split
        float m = 0.f;
#pragma unroll
        for (int t = 0; t < 7; ++t)
#pragma unroll
          for (int s = 0; s < 3; ++s) m = fmaxf(m, fabsf(A[t][s]));
#pragma unroll
        for (int o = 32; o; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
        int e = 0;
        (void)frexpf(m, &e);  // m = f 2^e, f in [0.5, 1); e = 0 for m = 0
        const float sc = ldexpf(1.f, 15 - e), inv = ldexpf(1.f, e - 15);
#pragma unroll
        for (int t = 0; t < 7; ++t) A16[t] = ft_split_row(A[t][0] * sc, A[t][1] * sc, A[t][2] * sc, grp);
        mx = -w * inv * xs_inv_g;
        myz = -w * inv * xs_inv_o;
      }
    }
    // The first chunk's records and their table columns in LDS; each later
    // chunk's records are loaded during the previous chunk's contraction,
    // and its columns staged after it (while the previous chunk stores).
    // Record indices are loaded a chunk ahead of the records.
    uint32_t inow = perm[lo + l < hi ? lo + l : lo];
    vh_pos rcur = vh_load_pos(P, inow);
    uint32_t inext = perm[lo + 64 + l < hi ? lo + 64 + l : lo];
    if (ok) ft_stage<S16>(ft_columns(rcur, half, spacing, nbdt, ncol), og, xsplit, tx, ty, tz, l);
    for (int base = lo; base < hi; base += 64) {
      const int nlive = min(64, hi - base);
      const bool mine = l < nlive;
      const vh_pos rp = rcur;
      // the previous positions, for the push after the contraction
      const vh_pos qp = vh_load_pos(Q, inow);
      const float dx0 = rp.d[0], dy0 = rp.d[1], dz0 = rp.d[2];
      float px, py, pz;
      if (pa.ball) {  // the particle's own position
        px = x0 + (float)(rp.k[0] - 1) * h + dx0;
        py = x0 + (float)(rp.k[1] - 1) * h + dy0;
        pz = x0 + (float)(rp.k[2] - 1) * h + dz0;
      } else {        // rebuilt from the cell, as the reference's kernel does
        px = x0 + (float)(kx - 1) * h + dx0;
        py = x0 + (float)(ky - 1) * h + dy0;
        pz = x0 + (float)(kz - 1) * h + dz0;
      }
      // the projectile term, independent of the contraction: it runs while
      // the chunk's columns are in flight
      float fx2 = 0.f, fy2 = 0.f, fz2 = 0.f;
      if (mine && pa.coef != 0.f) {
        const float dx = pa.px - px, dy = pa.py - py, dz = pa.pz - pz;
        const float d2 = dx * dx + dy * dy + dz * dz;
        const float r = sqrtf(d2);
        float kern, pot = 0.f;
        if (pa.ball) {
          const float R = pa.radius;
          kern = ball_force_kernel_f(d2, R);
          if (pa.pot) pot = r < R ? -(3.0f - (r / R) * (r / R)) / (2.0f * R) : -1.0f / r;
        } else {
          gaussian_law_f(d2, r, pa.sigma, pa.pot, kern, pot);
        }
        const float mm = pa.coef * kern;
        fx2 = mm * dx;
        fy2 = mm * dy;
        fz2 = mm * dz;
        sax += fx2;
        say += fy2;
        saz += fz2;
        sae += pot;
      }
      __syncthreads();  // the columns have landed
      const bool more = base + 64 < hi;
      {
        const int sn = base + 64 + l;
        inow = inext;
        rcur = vh_load_pos(P, inow);
        inext = perm[sn + 64 < hi ? sn + 64 : lo];
      }
      // the contraction, one sub-group of 16 particles at a time; lane l
      // keeps the field of particle l (lane group sg of sub-group sg)
      ft_f4 fl = {0.f, 0.f, 0.f, 0.f};
      if (ok) {
        const int nsg = (nlive + 15) >> 4;
#pragma unroll 1
        for (int sg = 0; sg < nsg; ++sg) {
          const int r = ft_row(16 * sg + j16);
          ft_f4 f;
          if constexpr (S16)
            f = ft_contract16(A16, ty + r, tz + r, tx + 512 * sg + 64 * grp + 4 * j16, grp, mx, myz);
          else
            f = ft_contract<1>(A, tx + r, ty + r, tz + r, grp, w);
          fl = grp == sg ? f : fl;
        }
        // the next chunk's columns: every read of this chunk's is done
        if (more) ft_stage<S16>(ft_columns(rcur, half, spacing, nbdt, ncol), og, xsplit, tx, ty, tz, l);
      }
      // field minus projectile term, coalesced stores of the live lanes;
      // with the push, the records and keys of a cell that fits (the
      // particles of the others are pushed by the fallback)
      // (a cell that does not fit stores nothing: its forces are the
      // fallback's, which may run before this kernel)
      const float f0 = ok ? fl[0] - fx2 : nan, f1 = ok ? fl[1] - fy2 : nan;
      const float f2 = ok ? fl[2] - fz2 : nan;
      const int nf = ok ? nlive : 0;
      vh_bstore_f(vh_rsrc(force + base, 4 * nf), l, f0);
      vh_bstore_f(vh_rsrc(force + npad + base, 4 * nf), l, f1);
      vh_bstore_f(vh_rsrc(force + 2 * npad + base, 4 * nf), l, f2);
      {
        const int np = pu.on && ok ? nlive : 0;
        vh_pos nw;
        const uint32_t key = push_record(rp, qp, f0, f1, f2, pu, nw);
        vh_bstore_pos(vh_rsrc(Pd + 6 * (int64_t)base, 24 * np), l, nw);
        vh_bstore_pos(vh_rsrc(Qd + 6 * (int64_t)base, 24 * np), l, rp);
        const uint32_t *kb = pu.on ? pu.keys + base : (const uint32_t *)P;
        __builtin_amdgcn_raw_buffer_store_b32((int)key, vh_rsrc(kb, 4 * np), 4 * l, 0, 0);
        nout += l < np && push_outside(nw, pu) ? 1 : 0;
      }
    }
  }
  sax = vh_wave_sum(sax);
  say = vh_wave_sum(say);
  saz = vh_wave_sum(saz);
  sae = vh_wave_sum(sae);
  nout = vh_wave_sum(nout);
  if (l == 0 && nout) atomicAdd(pu.nout, nout);
  if (l == 0) {
    part[blockIdx.x] = sax;
    part[gridDim.x + blockIdx.x] = say;
    part[2 * gridDim.x + blockIdx.x] = saz;
    part[3 * gridDim.x + blockIdx.x] = sae;
  }
}

// ---------------------------------------------------------------------------
// out-of-stencil fallback, FP64
// ---------------------------------------------------------------------------
__device__ __forceinline__ void hermite_vd(const double *__restrict__ g, int nk, int b, double x,
                                           double &v, double &dv) {
  const int kn = b >> 1, kind = b & 1;
  double a, h, s;
  v = 0.0;
  dv = 0.0;
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
    v = 3.0 * (a * a) - 2.0 * (a * a * a);
    dv = -6.0 * (a - a * a) * s / h;
  } else {
    v = h * ((a * a) - (a * a * a)) * s;
    dv = -2.0 * a + 3.0 * (a * a);
  }
}

__device__ __forceinline__ int cell_index_d(const double *__restrict__ g, int nk, double x) {
  if (x < g[0] || x > g[nk - 1]) return -1;
  int lo = 0, hi = nk;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (g[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  const int c = lo - 1;
  return c < 0 ? 0 : c;
}

struct vh_fallback_args {
  double px, py, pz;  // projectile position, T
  double q;           // projectile charge
  double radius;      // ball law radius (> 0) or Gaussian width (ball == 0)
  int ball;
  int has_proj;
  double w;           // weight
  double escaped;     // enclosed-charge count for particles outside both grids
  const double *dev;  // device-resident projectile state (position at 0..2), or null
};

__global__ void k_fallback(const int32_t *__restrict__ P, const uint32_t *__restrict__ perm,
                           int64_t npad, const int32_t *__restrict__ outlist,
                           const int32_t *__restrict__ cnt, double x0, double h,
                           const float *__restrict__ csolc, const double *__restrict__ kc, int nkc,
                           int nc, vh_fallback_args fa, vh_push_args pu, float *__restrict__ force) {
  // slot s's record is record perm[s] of P; the push writes slot s of pu.dst
  const int32_t *Q = vh_prev_half(P, npad);
  int32_t *Qd = vh_prev_half(pu.dst, npad);
  const int nout = cnt[CNT_OUTCOUNT];
  int rejected = 0;
  if (fa.dev) {
    fa.px = fa.dev[0];
    fa.py = fa.dev[1];
    fa.pz = fa.dev[2];
  }
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < nout;
       u += (int64_t)gridDim.x * blockDim.x) {
    const int s = outlist[u];
    const uint32_t i = perm[s];
    double p[3];
    const vh_pos rp = vh_load_pos(P, i);
    for (int d = 0; d < 3; ++d) p[d] = x0 + (double)(rp.k[d] - 1) * h + (double)rp.d[d];
    int cell[3];
    bool out = false;
    for (int d = 0; d < 3; ++d) {
      cell[d] = cell_index_d(kc, nkc, p[d]);
      out = out || cell[d] < 0;
    }
    double f0, f1, f2;
    if (out) {
      const double r2 = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
      const double r3 = pow(r2, 1.5);
      const double cf = -(fa.w * fa.w) * fa.escaped / r3;
      f0 = cf * p[0];
      f1 = cf * p[1];
      f2 = cf * p[2];
    } else {
      double v[3][4], dv[3][4];
      for (int d = 0; d < 3; ++d)
        for (int a = 0; a < 4; ++a) hermite_vd(kc, nkc, 2 * cell[d] + a, p[d], v[d][a], dv[d][a]);
      double ex = 0.0, ey = 0.0, ez = 0.0;
      for (int ak = 0; ak < 4; ++ak) {
        const int k = 2 * cell[2] + ak;
        const double vk = v[2][ak], dk = dv[2][ak];
        for (int aj = 0; aj < 4; ++aj) {
          const int j = 2 * cell[1] + aj;
          const double vj = v[1][aj], dj = dv[1][aj];
          for (int ai = 0; ai < 4; ++ai) {
            const int i = 2 * cell[0] + ai;
            const double vi = v[0][ai], di = dv[0][ai];
            const double c = (double)csolc[i + (int64_t)nc * (j + (int64_t)nc * k)];
            ex -= c * di * vj * vk;
            ey -= c * vi * dj * vk;
            ez -= c * vi * vj * dk;
          }
        }
      }
      f0 = fa.w * ex;
      f1 = fa.w * ey;
      f2 = fa.w * ez;
    }
    if (fa.has_proj) {
      const double d0 = fa.px - p[0], d1 = fa.py - p[1], d2 = fa.pz - p[2];
      const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
      const double fk = fa.ball ? ball_force_kernel_d(r2, fa.radius) : gaussian_force_kernel_d(r2, fa.radius);
      const double m = -fa.w * fa.q * fk;
      f0 = f0 - m * d0;
      f1 = f1 - m * d1;
      f2 = f2 - m * d2;
    }
    force[s] = (float)f0;
    force[npad + s] = (float)f1;
    force[2 * npad + s] = (float)f2;
    if (pu.on) {
      const vh_pos old = vh_load_pos(Q, i);
      vh_pos nw;
      pu.keys[s] = push_record(rp, old, (float)f0, (float)f1, (float)f2, pu, nw);
      vh_store_pos(pu.dst, s, nw);
      vh_store_pos(Qd, s, rp);
      rejected += push_outside(nw, pu) ? 1 : 0;
    }
  }
  rejected = vh_wave_sum(rejected);
  if (vh_lane() == 0 && rejected) atomicAdd(pu.nout, rejected);
}

// The field kernel's four columns of FP64 partials (projectile force and
// potential) to dscal[8..11] in one launch: block (s, v) sums slice s of
// column v; the last block to finish (a ticket at
// *ticket, reset by it) sums the slices of each column in a fixed order and,
// with dproj, steps the projectile (as k_projectile_step).  Grid (ns, 4).
__global__ void k_reduce_field(const double *__restrict__ part, int nparts,
                               double *__restrict__ slices, double *__restrict__ out,
                               uint32_t *__restrict__ ticket, double *__restrict__ dproj, double dt) {
  __shared__ double lds[16];
  __shared__ bool last;
  const int s = blockIdx.x, v = blockIdx.y, ns = gridDim.x;
  const int len = (nparts + ns - 1) / ns;
  const int lo = s * len < nparts ? s * len : nparts;
  const int hi = lo + len < nparts ? lo + len : nparts;
  const double r = vh_block_sum_d(vh_thread_sum(part + (int64_t)v * nparts + lo, hi - lo), lds);
  if (threadIdx.x == 0) {
    __hip_atomic_store(slices + (int64_t)v * ns + s, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    last = t == (uint32_t)(ns * gridDim.y) - 1u;
  }
  __syncthreads();
  if (!last) return;
  if (threadIdx.x < (int)gridDim.y) {
    const int c = threadIdx.x;
    double sum = 0.0;
    for (int k = 0; k < ns; ++k)
      sum += __hip_atomic_load(slices + (int64_t)c * ns + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    out[c] = sum;
    lds[c] = sum;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    *ticket = 0u;
    if (dproj) {
      // src/projectile.jl:298-304, as k_projectile_step
      double *st = dproj;
      const double p[3] = {st[0], st[1], st[2]}, prev[3] = {st[3], st[4], st[5]};
      const double q = st[10], R = st[12];
      const double r2 = (p[0] * p[0] + p[1] * p[1]) + p[2] * p[2];
      const double modf = r2 > R * R ? st[11] * q / pow(r2, 1.5) : q / 64.0;  // WIGNER_SEITZ_NA^3
      const double cfac = dt * dt / st[9];
      for (int d = 0; d < 3; ++d) {
        const double f = modf * p[d] + lds[d];
        const double nw = (2.0 * p[d] - prev[d]) + cfac * f;
        st[6 + d] = (nw - prev[d]) / (2.0 * dt);
        st[3 + d] = p[d];
        st[d] = nw;
      }
    }
  }
}

// Sums each of `nvals` columns of `nparts` partials in a fixed order; one
// block per column (any grid works, columns are strided over the blocks).
__global__ void k_reduce_cols(const double *__restrict__ part, int nparts, int nvals,
                              double *__restrict__ out) {
  __shared__ double lds[16];
  for (int v = blockIdx.x; v < nvals; v += gridDim.x) {
    const double r = vh_block_sum_d(vh_thread_sum(part + (int64_t)v * nparts, nparts), lds);
    if (threadIdx.x == 0) out[v] = r;
  }
}

// ---------------------------------------------------------------------------
// interaction energy sum_i w phi(r_i), FP64 (src/energy.jl:19-43)
// ---------------------------------------------------------------------------
struct vh_energy_args {
  double x0, h;        // PackedPositions x0, h: r = x0 + (k - 1) h + delta
  double lo, hi;       // fine knots[3], knots[end-2]: the smoothed-potential region
  double spacing;      // smoothing table spacing
  double nbdt;
  double w;            // weight
  double enclosed;     // charge seen beyond both grids
  int nkf, nf;         // fine knots, fine grid size
  int nkc, nc;         // coarse knots, coarse grid size
  int ncol;
};

// nearest_knot (src/fields.jl:224-227), 0-based: searchsortedlast, clamped to
// a cell, then the closer end of that cell
__device__ __forceinline__ int nearest_knot_d(const double *__restrict__ g, int nk, double x) {
  int lo = 0, hi = nk;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (g[mid] <= x) lo = mid + 1;
    else hi = mid;
  }
  const int i = lo < 1 ? 1 : (lo > nk - 1 ? nk - 1 : lo);  // 1-based searchsortedlast
  return (g[i] - x) < (x - g[i - 1]) ? i : i - 1;
}

// The potential at one point p (FP64): inside (lo, hi)^3 smoothed_potential
// (src/fields.jl:405-428) of the fine csol; elsewhere spline_potential
// (src/fields.jl:440-466) of the coarse csol, or enclosed / |r| beyond both
// grids.  Every product and sum in the reference's order, in FP64, on the
// FP32 coefficients widened exactly.
__device__ double energy_point(const double (&p)[3], bool inside, const float *__restrict__ csolf,
                               const double *__restrict__ kf, const double *__restrict__ ovl,
                               const float *__restrict__ csolc, const double *__restrict__ kc,
                               const vh_energy_args &ea) {
  const int64_t nf = ea.nf, nc = ea.nc;
  double phi = 0.0;
  if (inside) {
    int64_t b[3];
    const double *o[3];
    for (int d = 0; d < 3; ++d) {
      const int ci = nearest_knot_d(kf, ea.nkf, p[d]);
      b[d] = 2 * ci - 4;  // 0-based 2ci - 5
      int col = (int)floor((p[d] - kf[ci] + ea.spacing / 2) / ea.spacing * ea.nbdt + 0.5) + 1;
      col = col < 1 ? 1 : (col > ea.ncol ? ea.ncol : col);
      o[d] = ovl + 10 * (col - 1);
    }
    for (int kk = 0; kk < 10; ++kk) {
      const double oz = o[2][kk];
      for (int jj = 0; jj < 10; ++jj) {
        const double wgt = o[1][jj] * oz;
        const float *row = csolf + b[0] + nf * ((b[1] + jj) + nf * (b[2] + kk));
        double s = 0.0;
#pragma unroll
        for (int ii = 0; ii < 10; ++ii) s += (double)row[ii] * o[0][ii];
        phi += s * wgt;
      }
    }
  } else {
    int cell[3];
    bool out = false;
    for (int d = 0; d < 3; ++d) {
      cell[d] = cell_index_d(kc, ea.nkc, p[d]);
      out = out || cell[d] < 0;
    }
    if (out) {
      phi = ea.enclosed / sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    } else {
      double v[3][4], dv;
      for (int d = 0; d < 3; ++d)
        for (int a = 0; a < 4; ++a) hermite_vd(kc, ea.nkc, 2 * cell[d] + a, p[d], v[d][a], dv);
      for (int ak = 0; ak < 4; ++ak) {
        const int64_t k = 2 * cell[2] + ak;
        for (int aj = 0; aj < 4; ++aj) {
          const int64_t j = 2 * cell[1] + aj;
          const double wgt = v[1][aj] * v[2][ak];
          for (int ai = 0; ai < 4; ++ai)
            phi += (double)csolc[2 * cell[0] + ai + nc * (j + nc * k)] * v[0][ai] * wgt;
        }
      }
    }
  }
  return phi;
}

// One lane per particle, in the buffer's current order (before any sort)
__global__ void __launch_bounds__(256)
    k_energy(const int32_t *__restrict__ P, int64_t npad, int64_t npart,
             const float *__restrict__ csolf, const double *__restrict__ kf,
             const double *__restrict__ ovl, const float *__restrict__ csolc,
             const double *__restrict__ kc, vh_energy_args ea, double *__restrict__ part) {
  __shared__ double lds[8];
  double total = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < npart;
       i += (int64_t)gridDim.x * blockDim.x) {
    double p[3];
    bool inside = true;
    const vh_pos rp = vh_load_pos(P, i);
    for (int d = 0; d < 3; ++d) {
      p[d] = ea.x0 + (double)(rp.k[d] - 1) * ea.h + (double)rp.d[d];
      inside = inside && ea.lo < p[d] && p[d] < ea.hi;
    }
    total += ea.w * energy_point(p, inside, csolf, kf, ovl, csolc, kc, ea);
  }
  const double s = vh_block_sum_d(total, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// The same sum over the work units of the last sort (a wave per unit, the
// slots of one cell), on the matrix cores in FP64.  A particle's cell is
// its nearest knot c0 (0-based) on each axis, so nearest_knot_d gives c0
// for all but the particles moved off their cell since the sort (decided
// here by the same comparisons, between the knot's neighbours), and those
// use the cell's 10^3 block of csol, starting at 2 c0 - 4.  The first stage
//   T[p, (jj, kk)] = sum_ii c[ii, jj, kk] ox[p, ii]
// is a product of the block (rows x ii, the cell's, in registers) by 16
// particles' x weights, v_mfma_f64_16x16x4f64 (lane l supplies
// A[l % 16][l / 16] and B[l / 16][l % 16] and receives D[l / 16 + 4 v][l % 16],
// v = 0..3), then the y and z sums per particle on the VALU; the rows of each
// lane group are the field kernel's (slot q = 4 t + v of group g: planes g
// and g + 4, and half of plane 8 + g / 2), 25 rows in 28 slots.  The other
// particles take energy_point.  Every load is unconditional (a conditional
// one becomes a branch that waits for every load in flight), its value
// selected afterwards.  FP64 throughout; the sums run in another order than
// the reference's chain (agreement to FP64 rounding).
#define EN_STRIDE 33   // doubles a particle row: ox, oy, oz, flag (odd stride)

__global__ void __launch_bounds__(64)
    k_energy_units(const int32_t *__restrict__ P, const vh_unit *__restrict__ units,
                   const int32_t *__restrict__ cnt, int nk, const float *__restrict__ csolf,
                   const double *__restrict__ kf, const double *__restrict__ ovl,
                   const float *__restrict__ csolc, const double *__restrict__ kc,
                   vh_energy_args ea, double *__restrict__ part) {
  typedef double d4 __attribute__((ext_vector_type(4)));
  __shared__ double tab[64 * EN_STRIDE];
  // the particles that take energy_point, queued so that they run 64 at a
  // time (a lone one would hold the whole wave through it): position and
  // the inside flag, in the order the wave meets them
  __shared__ double qp[128][4];
  int qn = 0;
  const int l = threadIdx.x, j16 = l & 15, g = l >> 4;
  const int nunits = cnt[CNT_NUNITS];
  const int64_t nf = ea.nf;
  const double half = ea.spacing / 2;
  double total = 0.0;
  for (int u = blockIdx.x; u < nunits; u += gridDim.x) {
    const vh_unit un = units[u];
    const int kc3[3] = {un.key % nk, (un.key / nk) % nk, un.key / (nk * nk)};
    bool okw = true;
    double gm[3], g0[3], gp[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int c0 = kc3[d];
      okw = okw && c0 >= 2 && c0 + 1 <= ea.nkf - 1 && 2 * c0 + 5 <= ea.nf - 1;
      const int cz = c0 < ea.nkf - 1 ? c0 : ea.nkf - 2;
      gm[d] = kf[cz >= 1 ? cz - 1 : 0];
      g0[d] = kf[cz];
      gp[d] = kf[cz + 1];
    }
    // A operands: row j16 of tile t is slot q = 4 t + j16 / 4 of lane group
    // j16 % 4, K index ii = 4 s + g; padding (q >= 25, ii >= 10) reads an
    // in-window point and keeps 0
    double A[7][3];
    {
      const int gg = j16 & 3;
      const float *tb = csolf + (okw ? (2 * kc3[0] - 4) + nf * ((2 * kc3[1] - 4) + nf * (2 * kc3[2] - 4)) : 0);
      float cv[7][3];
#pragma unroll
      for (int t = 0; t < 7; ++t) {
        const int q0 = 4 * t + (j16 >> 2), q = q0 < 25 ? q0 : 0;
        const int jj = q < 10 ? q : q < 20 ? q - 10 : q - 20 + 5 * (gg & 1);
        const int kk = q < 10 ? gg : q < 20 ? gg + 4 : 8 + (gg >> 1);
#pragma unroll
        for (int s2 = 0; s2 < 3; ++s2) {
          const int ii0 = 4 * s2 + g, ii = ii0 < 10 ? ii0 : 0;
          cv[t][s2] = tb[ii + nf * (jj + nf * kk)];
        }
      }
#pragma unroll
      for (int t = 0; t < 7; ++t)
#pragma unroll
        for (int s2 = 0; s2 < 3; ++s2)
          A[t][s2] = (okw && 4 * t + (j16 >> 2) < 25 && 4 * s2 + g < 10) ? (double)cv[t][s2] : 0.0;
    }
    for (int base = un.lo; base < un.hi; base += 64) {
      const int nlive = min(64, un.hi - base);
      const bool live = l < nlive;
      double p[3];
      bool inside = live;
      const vh_pos rp = vh_load_pos(P, live ? base + l : base);
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        p[d] = ea.x0 + (double)(rp.k[d] - 1) * ea.h + (double)rp.d[d];
        inside = inside && ea.lo < p[d] && p[d] < ea.hi;
      }
      // nearest_knot_d between the knot's neighbours: searchsortedlast is
      // c0 or c0 + 1 there
      bool fast = inside && okw;
      int col[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const bool near = gm[d] <= p[d] && p[d] < gp[d];
        const bool right = g0[d] <= p[d];          // searchsortedlast = c0 + 1
        const double gi = right ? gp[d] : g0[d], gi1 = right ? g0[d] : gm[d];
        const bool upper = (gi - p[d]) < (p[d] - gi1);
        // the nearest knot is c0 when it is i - 1 with i = c0 + 1, or i with i = c0
        fast = fast && near && (right ? !upper : upper);
        const int cc = (int)floor((p[d] - g0[d] + half) / ea.spacing * ea.nbdt + 0.5) + 1;
        col[d] = (cc < 1 ? 1 : (cc > ea.ncol ? ea.ncol : cc)) - 1;
      }
      {
        const bool slow = live && !fast;
        const uint64_t m = __ballot(slow);
        const int at = qn + __popcll(m & ((1ull << l) - 1ull));
        if (slow) {
          qp[at][0] = p[0];
          qp[at][1] = p[1];
          qp[at][2] = p[2];
          qp[at][3] = inside ? 1.0 : 0.0;
        }
        qn += __popcll(m);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        if (qn >= 64) {   // 64 queued: a full wave of them
          const double q3[3] = {qp[l][0], qp[l][1], qp[l][2]};
          total += ea.w * energy_point(q3, qp[l][3] != 0.0, csolf, kf, ovl, csolc, kc, ea);
          double keep[4] = {0.0, 0.0, 0.0, 0.0};
          if (l < qn - 64)
            for (int c = 0; c < 4; ++c) keep[c] = qp[64 + l][c];
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
          if (l < qn - 64)
            for (int c = 0; c < 4; ++c) qp[l][c] = keep[c];
          qn -= 64;
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        }
      }
      double ov[3][10];
#pragma unroll
      for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int a = 0; a < 10; ++a) ov[d][a] = ovl[10 * col[d] + a];
      double *row = tab + l * EN_STRIDE;
#pragma unroll
      for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int a = 0; a < 10; ++a) row[10 * d + a] = fast ? ov[d][a] : 0.0;
      row[30] = fast ? 1.0 : 0.0;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      const int nsg = (nlive + 15) >> 4;
      for (int sg = 0; sg < nsg; ++sg) {
        const double *pr = tab + (16 * sg + j16) * EN_STRIDE;
        double b[3], y[10], yh[5], z[3];
#pragma unroll
        for (int s2 = 0; s2 < 3; ++s2) {
          const int i0 = 4 * s2 + g;
          const double v = pr[i0 < 10 ? i0 : 0];
          b[s2] = i0 < 10 ? v : 0.0;
        }
#pragma unroll
        for (int a = 0; a < 10; ++a) y[a] = pr[10 + a];
#pragma unroll
        for (int a = 0; a < 5; ++a) yh[a] = pr[10 + 5 * (g & 1) + a];
        z[0] = pr[20 + g];
        z[1] = pr[24 + g];
        z[2] = pr[28 + (g >> 1)];
        const double flag = pr[30];
        double u0 = 0.0, u1 = 0.0, u2 = 0.0;
#pragma unroll
        for (int t = 0; t < 7; ++t) {
          d4 T = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
          for (int s2 = 0; s2 < 3; ++s2) T = __builtin_amdgcn_mfma_f64_16x16x4f64(A[t][s2], b[s2], T, 0, 0, 0);
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int q = 4 * t + v;
            if (q < 10) u0 += y[q] * T[v];
            else if (q < 20) u1 += y[q - 10] * T[v];
            else if (q < 25) u2 += yh[q - 20] * T[v];
          }
        }
        double phi = (z[0] * u0 + z[1] * u1) + z[2] * u2;
        phi += __shfl_xor(phi, 16, 64);
        phi += __shfl_xor(phi, 32, 64);
        if (g == 0 && flag != 0.0) total += ea.w * phi;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
  }
  if (qn > 0 && l < qn) {   // the rest of the queue
    const double q3[3] = {qp[l][0], qp[l][1], qp[l][2]};
    total += ea.w * energy_point(q3, qp[l][3] != 0.0, csolf, kf, ovl, csolc, kc, ea);
  }
  total = vh_wave_sum(total);
  if (l == 0) part[blockIdx.x] = total;
}

#define ENERGY_BLOCKS 2048
#define ENERGY_UNIT_BLOCKS 8192   // waves striding over the units (partials fit 8 dpart_len)

int vh_energy_impl(vh_accel *a, vh_mesh *fine, vh_mesh *coarse, double weight, double enclosed) {
  vh_context *c = a->ctx;
  vh_energy_args ea;
  ea.x0 = a->x0;
  ea.h = a->h;
  ea.lo = a->klo;
  ea.hi = a->khi;
  ea.spacing = a->spacing64;
  ea.nbdt = (double)a->nbdt;
  ea.w = weight;
  ea.enclosed = enclosed;
  ea.nkf = a->nk;
  ea.nf = a->n;
  ea.nkc = coarse->nk;
  ea.nc = coarse->n;
  ea.ncol = a->ncol;
  VH_TRY(vh_rec_settle(a));
  int blocks = ENERGY_BLOCKS;
  if (a->sorted_once && fine->n == a->n) {
    // the work units of the last sort partition the slots: a wave per unit
    blocks = ENERGY_UNIT_BLOCKS;
    hipLaunchKernelGGL(k_energy_units, dim3(blocks), dim3(64), 0, c->stream, a->parts[a->cur],
                       a->units, a->counters, a->nk, fine->csol, a->knots64, a->ovl64,
                       coarse->csol, coarse->knots64, ea, a->dpart);
  } else {
    hipLaunchKernelGGL(k_energy, dim3(blocks), dim3(256), 0, c->stream, a->parts[a->cur],
                       a->npad, a->npart, fine->csol, a->knots64, a->ovl64, coarse->csol,
                       coarse->knots64, ea, a->dpart);
  }
  VH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_reduce_cols, dim3(1), dim3(256), 0, c->stream, a->dpart, blocks, 1,
                     a->dscal + 18);
  VH_LAUNCH_CHECK();
  return VH_OK;
}

static vh_push_args push_args(vh_accel *a, double push_a) {
  vh_push_args pu;
  pu.on = 0;
  pu.a = pu.h = 0.f;
  pu.nk = a->nk;
  pu.keys = nullptr;
  pu.dst = a->parts[a->cur];
  pu.x0 = (float)a->x0;
  pu.lo = a->lo;
  pu.hi = a->hi;
  pu.nout = nullptr;
  if (push_a > 0.0) {   // (the matrix-core field kernel only: checked by the callers)
    pu.on = 1;
    pu.a = (float)push_a;
    pu.h = (float)a->h;
    pu.keys = a->keys[0];
    // the moved records go to the other buffer; the deposit that follows
    // the next sort reads their rejection count from the slot of that buffer
    pu.dst = a->parts[1 - a->cur];
    pu.nout = a->counters + CNT_PUSH_NOUT + (1 - a->cur);
  }
  return pu;
}

// The out-of-stencil fallback of the particles of cells whose stencil does
// not fit (FP64; with push_a > 0 it pushes them), on the context's current
// stream.  It reads the projectile's position before the step.
int vh_fallback_impl(vh_accel *a, vh_mesh *coarse, double weight, const double *proj,
                     double ball_radius, const double *dproj, double push_a) {
  VH_REQUIRE(push_a <= 0.0 || !a->exact_field, "fallback with the push: the matrix-core field kernel only");
  vh_fallback_args fa;
  fa.px = proj ? proj[0] : 0.0;
  fa.py = proj ? proj[1] : 0.0;
  fa.pz = proj ? proj[2] : 0.0;
  fa.q = proj ? proj[3] : 0.0;
  fa.ball = ball_radius > 0;
  fa.radius = ball_radius > 0 ? ball_radius : (proj ? proj[4] : 1.0);
  fa.has_proj = proj != nullptr;
  fa.w = weight;
  fa.escaped = 0.0;
  fa.dev = dproj;
  hipLaunchKernelGGL(k_fallback, dim3(1024), dim3(256), 0, a->ctx->stream, a->parts[a->cur],
                     vh_rec_perm(a), a->npad, a->outlist, a->counters, a->x0, a->h, coarse->csol,
                     coarse->knots64, coarse->nk, coarse->n, fa, push_args(a, push_a), a->force);
  VH_LAUNCH_CHECK();
  return VH_OK;
}

int vh_forces_impl(vh_accel *a, vh_mesh *fine, vh_mesh *coarse, double weight, const double *proj,
                   int soft_mode, double ball_radius, const double *dproj, double push_a,
                   int fallback, double proj_dt) {
  vh_context *c = a->ctx;
  hipStream_t st = c->stream;
  // the exact-order kernel reads the records in slot order
  if (a->exact_field) VH_TRY(vh_rec_settle(a));
  VH_REQUIRE(push_a <= 0.0 || !a->exact_field, "forces with the push: the matrix-core field kernel only");
  const vh_push_args pu = push_args(a, push_a);
  vh_proj_args pa;
  pa.dev = dproj;
  pa.px = pa.py = pa.pz = 0.f;
  pa.coef = 0.f;
  pa.sigma = 1.f;
  pa.ball = soft_mode == VH_SOFT_BALL;
  pa.radius = 1.f;
  pa.pot = dproj == nullptr;  // the projectile on the device reads only the force sums
  if (proj) {
    pa.px = (float)proj[0];
    pa.py = (float)proj[1];
    pa.pz = (float)proj[2];
    pa.coef = (float)(-weight * proj[3]);
    pa.sigma = (float)proj[4];
    pa.radius = (float)(ball_radius > 0 ? ball_radius : proj[4]);
  }
  // One block per work unit up to 65536 units (the dispatcher then balances
  // units of different weights); beyond, blocks stride over units.  Four
  // FP64 partials per block and the reduction's 4 x FIELD_SLICES slice sums
  // fit the 8 x dpart_len buffer.
  constexpr int FIELD_SLICES = 8;
  int grid = (int)(a->maxunits < 65536 ? a->maxunits : 65536);
  if (4 * (int64_t)grid + 4 * FIELD_SLICES > 8 * a->dpart_len)
    grid = (int)(2 * a->dpart_len - FIELD_SLICES);
  int tok = vh_prof_begin(c, VH_ST_FIELD_KERNEL);
  if (a->exact_field)
    hipLaunchKernelGGL(k_field, dim3(grid), dim3(64), 0, st, a->parts[a->cur], a->npad,
                       fine->csol, a->n, a->ovl, a->grad, a->units, a->counters,
                       a->nk, (float)a->x0, (float)a->h, a->spacing, (float)a->nbdt,
                       (float)weight, a->ncol, pa, a->force, a->dpart);
  else
    hipLaunchKernelGGL(a->field_fp32 ? k_field_mfma<false> : k_field_mfma<true>, dim3(grid), dim3(64), 0, st,
                       a->parts[a->cur], vh_rec_perm(a), a->npad, fine->csol, a->n, (const float4 *)a->og,
                       (const ft_u4 *)a->xsplit, a->xs_inv[0], a->xs_inv[1], a->zero, a->units,
                       a->counters, a->nk, (float)a->x0, (float)a->h, a->spacing, (float)a->nbdt,
                       (float)weight, a->ncol, pa, pu, a->force, a->dpart);
  vh_prof_end(c, tok);
  VH_LAUNCH_CHECK();
  if (fallback == VH_FALLBACK_AFTER) {
    VH_TRY(vh_fallback_impl(a, coarse, weight, proj, ball_radius, dproj, push_a));
  } else {
    VH_REQUIRE(!pu.on || fallback == VH_FALLBACK_DONE, "forces: the push needs the fallback");
  }
  {
    // 65536 partials of 4 columns: slices over 32 blocks, then their sums by
    // the last block (4 blocks alone took 13 us); the slices' sums go after
    // the partials.  (After the fallback, which reads the projectile's
    // position before the step.)
    constexpr int nsl = FIELD_SLICES;
    double *sl = a->dpart + 4 * (int64_t)grid;
    VH_REQUIRE(4 * (int64_t)grid + 4 * nsl <= 8 * a->dpart_len, "forces: partials beyond the buffer");
    const bool step_proj = proj_dt > 0 && dproj != nullptr;
    hipLaunchKernelGGL(k_reduce_field, dim3(nsl, 4), dim3(1024), 0, st, a->dpart, grid, sl,
                       a->dscal + 8, (uint32_t *)(a->counters + CNT_REDUCE_TICKET),
                       step_proj ? (double *)dproj : (double *)nullptr, proj_dt);
    VH_LAUNCH_CHECK();
  }
  if (pu.on) {
    a->cur = 1 - a->cur;  // the pushed records, in slot order
    a->rec_pending = 0;
    a->keys_fresh = 1;    // keys of the pushed records; the histograms are the sort's
    a->hist_fresh = 0;
    a->pushed_fused = 1;
  }
  return VH_OK;
}

// ---------------------------------------------------------------------------
// the projectile on the device: projectile_forces! (src/projectile.jl:203-216)
// from the field kernel's reduction plus the jellium term, then its Verlet
// step (src/projectile.jl:298-304), in FP64 with the reference's operation
// order.  State: [0..2] position, [3..5] previous, [6..8] velocity,
// [9] mass, [10] charge, [11] jellium ions, [12] jellium radius.
// ---------------------------------------------------------------------------
__global__ void k_projectile_step(double *__restrict__ st, const double *__restrict__ red, double dt) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double p[3] = {st[0], st[1], st[2]}, prev[3] = {st[3], st[4], st[5]};
  const double q = st[10], R = st[12];
  const double r2 = (p[0] * p[0] + p[1] * p[1]) + p[2] * p[2];
  const double modf = r2 > R * R ? st[11] * q / pow(r2, 1.5) : q / 64.0;  // WIGNER_SEITZ_NA^3
  const double c = dt * dt / st[9];
  for (int d = 0; d < 3; ++d) {
    const double f = modf * p[d] + red[d];
    const double nw = (2.0 * p[d] - prev[d]) + c * f;
    st[6 + d] = (nw - prev[d]) / (2.0 * dt);
    st[3 + d] = p[d];
    st[d] = nw;
  }
}

int vh_projectile_step_impl(vh_accel *a, double dt) {
  hipLaunchKernelGGL(k_projectile_step, dim3(1), dim3(64), 0, a->ctx->stream, a->dproj, a->dscal + 8,
                     dt);
  VH_LAUNCH_CHECK();
  return VH_OK;
}

// ---------------------------------------------------------------------------
// packed Verlet (src/kernels.jl:985-1025)
// ---------------------------------------------------------------------------
// KEYS: the push also writes each particle's new cell key to keys[] and
// counts it in the histogram of the first radix digit of its sort tile
// (4096 consecutive slots), counts[d nblocks + tile], zero on entry, so that
// the next sort starts from them.  A wave's 64 slots lie in one tile and,
// the slots being in cell order, mostly share a digit: each round adds the
// lanes that share the digit of the first lane left with one atomic.
template <bool KEYS>
__global__ void __launch_bounds__(256)
    k_verlet(int32_t *__restrict__ P, int64_t npad, int64_t n, const float *__restrict__ force,
             float a, float h, float x0, float pfac, float r2max, float inv2M,
             double *__restrict__ part, uint32_t *__restrict__ keys, uint32_t *__restrict__ counts,
             int nk, int nbins0, int64_t nblocks) {
  __shared__ double lds[8];
  int32_t *Q = vh_prev_half(P, npad);
  double ek = 0.0, eo = 0.0, lx = 0.0, ly = 0.0, lz = 0.0;
  const uint32_t mask = (uint32_t)(nbins0 - 1);
  // whole waves walk the slots, so that a wave's lanes stay in one tile
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t nw = (n + 63) / 64 * 64;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nw; i += stride) {
    const bool live = i < n;
    uint32_t key = 0u;
    if (live) {
      const vh_pos cur = vh_load_pos(P, i), old = vh_load_pos(Q, i);
      const int k1 = cur.k[0], k2 = cur.k[1], k3 = cur.k[2];
      const float d1 = cur.d[0], d2 = cur.d[1], d3 = cur.d[2];
      const int o1 = old.k[0], o2 = old.k[1], o3 = old.k[2];
      const float e1 = old.d[0], e2 = old.d[1], e3 = old.d[2];
      int n1, n2, n3;
      float m1, m2, m3, g1, g2, g3;
      verlet1(k1, d1, o1, e1, a, force[i], h, n1, m1, g1);
      verlet1(k2, d2, o2, e2, a, force[npad + i], h, n2, m2, g2);
      verlet1(k3, d3, o3, e3, a, force[2 * npad + i], h, n3, m3, g3);
      const float px = g1 * pfac, py = g2 * pfac, pz = g3 * pfac;
      const float e = (px * px + py * py + pz * pz) * inv2M;
      ek += e;
      const float qx = x0 + (float)(k1 - 1) * h + d1;
      const float qy = x0 + (float)(k2 - 1) * h + d2;
      const float qz = x0 + (float)(k3 - 1) * h + d3;
      if (qx * qx + qy * qy + qz * qz > r2max) eo += e;
      lx += qy * pz - qz * py;
      ly += qz * px - qx * pz;
      lz += qx * py - qy * px;
      vh_pos nw_;
      nw_.k[0] = n1;
      nw_.k[1] = n2;
      nw_.k[2] = n3;
      nw_.d[0] = m1;
      nw_.d[1] = m2;
      nw_.d[2] = m3;
      vh_store_pos(P, i, nw_);
      vh_store_pos(Q, i, cur);
      key = cell_key0(n1, n2, n3, nk);
      if (KEYS) keys[i] = key;
    }
    if (KEYS) {
      const uint32_t dg = key & mask;
      uint32_t *col = counts + (i >> 12);   // this wave's tile
      uint64_t left = __ballot(live);
      for (int it = 0; it < 4 && left; ++it) {
        const int lead = __ffsll((unsigned long long)left) - 1;
        const uint32_t d0 = __shfl(dg, lead, 64);
        const uint64_t same = __ballot(live && dg == d0) & left;
        if ((threadIdx.x & 63) == lead) atomicAdd(col + (int64_t)d0 * nblocks, (uint32_t)__popcll(same));
        left &= ~same;
      }
      if (live && (left >> (threadIdx.x & 63) & 1ull)) atomicAdd(col + (int64_t)dg * nblocks, 1u);
    }
  }
  double vals[5] = {ek, eo, lx, ly, lz};
  for (int v = 0; v < 5; ++v) {
    const double s = vh_block_sum_d(vals[v], lds);
    if (threadIdx.x == 0) part[(int64_t)v * gridDim.x + blockIdx.x] = s;
  }
}

#define VERLET_BLOCKS 2048

int vh_verlet_impl(vh_accel *a, double dt, double mass, double rcmax) {
  vh_context *c = a->ctx;
  VH_TRY(vh_rec_settle(a));  // pushed in place, in slot order
  const float fa = (float)(dt * dt / mass);
  const float fh = (float)a->h, fx0 = (float)a->x0;
  const float pfac = (float)(mass / (2.0 * dt));
  const float r2max = (float)(rcmax * rcmax);
  const float inv2M = (float)(1.0 / (2.0 * mass));
  // the first radix digit, as the sort takes it
  int bits0 = a->digit_bits < a->bits_total ? a->digit_bits : a->bits_total;
  if (bits0 < 1) bits0 = 1;
  const int grid = VERLET_BLOCKS;
  if (vh_verlet_tiles(a)) {
    VH_TRY(vh_fill_i32(c, (int32_t *)a->counts, (1 << bits0) * a->nblocks, 0));
    hipLaunchKernelGGL(k_verlet<true>, dim3(grid), dim3(256), 0, c->stream, a->parts[a->cur],
                       a->npad, a->npart, a->force, fa, fh, fx0, pfac, r2max, inv2M, a->dpart,
                       a->keys[0], a->counts, a->nk, 1 << bits0, a->nblocks);
    a->keys_fresh = 1;
    a->hist_fresh = 1;
  } else {
    hipLaunchKernelGGL(k_verlet<false>, dim3(grid), dim3(256), 0, c->stream, a->parts[a->cur],
                       a->npad, a->npart, a->force, fa, fh, fx0, pfac, r2max, inv2M, a->dpart,
                       (uint32_t *)nullptr, (uint32_t *)nullptr, a->nk, 1, (int64_t)1);
    a->keys_fresh = 0;
  }
  VH_LAUNCH_CHECK();
  a->pushed_fused = 0;
  hipLaunchKernelGGL(k_reduce_cols, dim3(5), dim3(256), 0, c->stream, a->dpart, grid, 5, a->dscal);
  VH_LAUNCH_CHECK();
  return VH_OK;
}

// ---------------------------------------------------------------------------
// the legacy route: a cloud held on the host (ForceAccelerator(MtlArray, ...),
// update_forces!'s accelerator branch, src/simulation.jl:318-345)
// ---------------------------------------------------------------------------
// _pack_kd_kernel! (src/kernels.jl:134-145), in FP64 on the device: the
// nearest knot round((p - x0)/h) + 1 (half to even) clamped to [1, nk], the
// offset p - (x0 + (k - 1) h) subtracted in FP64 and only then narrowed.  The
// previous half gets the same record (the legacy route never reads it).
__global__ void k_pack_kd(const double *__restrict__ pos, int32_t *__restrict__ P, int64_t n,
                          int64_t npad, double x0, double h, int nk) {
  int32_t *Q = vh_prev_half(P, npad);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    vh_pos r;
    for (int d = 0; d < 3; ++d) {
      const double p = pos[3 * i + d];
      int k = (int)rint((p - x0) / h) + 1;
      k = k < 1 ? 1 : (k > nk ? nk : k);
      r.k[d] = k;
      r.d[d] = (float)(p - (x0 + (double)(k - 1) * h));
    }
    vh_store_pos(P, i, r);
    vh_store_pos(Q, i, r);
  }
}

// The forces of every particle in the cloud's own order, FP64
// (src/accelerator.jl:660-731): a particle of a cell whose 10^3 stencil fits
// takes the field kernel's FP32 force widened; any other one the coarse
// spline_field of the host's FP64 coefficients at its FP64 position (or the
// escaped-charge monopole beyond both grids), plus the projectile's own law.
__global__ void k_forces_host(const int32_t *__restrict__ P, int64_t npart, int64_t npad,
                              const uint32_t *__restrict__ cperm, const double *__restrict__ pos,
                              const float *__restrict__ force, int nk, int n,
                              const double *__restrict__ csolc, const double *__restrict__ kc,
                              int nkc, int nc, vh_fallback_args fa, double *__restrict__ out) {
  for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < npart;
       s += (int64_t)gridDim.x * blockDim.x) {
    const vh_pos r = vh_load_pos(P, s);
    bool fits = true;
    for (int d = 0; d < 3; ++d) {
      const int k = r.k[d] < 1 ? 1 : (r.k[d] > nk ? nk : r.k[d]);
      const int b = 2 * k - 5;
      fits = fits && b >= 1 && b + 9 <= n;
    }
    const int64_t i = cperm[s];
    if (fits) {
      out[3 * i] = (double)force[s];
      out[3 * i + 1] = (double)force[npad + s];
      out[3 * i + 2] = (double)force[2 * npad + s];
      continue;
    }
    const double p[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    int cell[3];
    bool outside = false;
    for (int d = 0; d < 3; ++d) {
      cell[d] = cell_index_d(kc, nkc, p[d]);
      outside = outside || cell[d] < 0;
    }
    double f0, f1, f2;
    if (outside) {
      const double r3 = pow(p[0] * p[0] + p[1] * p[1] + p[2] * p[2], 1.5);
      const double cf = -(fa.w * fa.w) * fa.escaped / r3;
      f0 = cf * p[0];
      f1 = cf * p[1];
      f2 = cf * p[2];
    } else {
      double v[3][4], dv[3][4];
      for (int d = 0; d < 3; ++d)
        for (int a = 0; a < 4; ++a) hermite_vd(kc, nkc, 2 * cell[d] + a, p[d], v[d][a], dv[d][a]);
      double ex = 0.0, ey = 0.0, ez = 0.0;
      for (int ak = 0; ak < 4; ++ak) {
        const int k = 2 * cell[2] + ak;
        for (int aj = 0; aj < 4; ++aj) {
          const int j = 2 * cell[1] + aj;
          for (int ai = 0; ai < 4; ++ai) {
            const int ii = 2 * cell[0] + ai;
            const double c = csolc[ii + (int64_t)nc * (j + (int64_t)nc * k)];
            ex -= c * dv[0][ai] * v[1][aj] * v[2][ak];
            ey -= c * v[0][ai] * dv[1][aj] * v[2][ak];
            ez -= c * v[0][ai] * v[1][aj] * dv[2][ak];
          }
        }
      }
      f0 = fa.w * ex;
      f1 = fa.w * ey;
      f2 = fa.w * ez;
    }
    if (fa.has_proj) {
      const double d0 = fa.px - p[0], d1 = fa.py - p[1], d2 = fa.pz - p[2];
      const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
      const double fk = fa.ball ? ball_force_kernel_d(r2, fa.radius) : gaussian_force_kernel_d(r2, fa.radius);
      const double m = -fa.w * fa.q * fk;
      f0 = f0 - m * d0;
      f1 = f1 - m * d1;
      f2 = f2 - m * d2;
    }
    out[3 * i] = f0;
    out[3 * i + 1] = f1;
    out[3 * i + 2] = f2;
  }
}

extern "C" int vh_positions_upload(vh_accel *a, const double *pos) {
  VH_REQUIRE(a && pos, "vh_positions_upload: null argument");
  VH_CHECK_HIP(hipSetDevice(a->ctx->device));
  if (!a->pos64) VH_TRY(vh_alloc((void **)&a->pos64, sizeof(double) * 3 * (size_t)a->npart));
  if (!a->cloud_perm) VH_TRY(vh_alloc((void **)&a->cloud_perm, sizeof(uint32_t) * (size_t)a->npart));
  hipStream_t st = a->ctx->stream;
  VH_CHECK_HIP(hipMemcpyAsync(a->pos64, pos, sizeof(double) * 3 * (size_t)a->npart,
                              hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_pack_kd, dim3(2048), dim3(256), 0, st, a->pos64, a->parts[a->cur], a->npart,
                     a->npad, a->x0, a->h, a->nk);
  VH_LAUNCH_CHECK();
  // new records in the cloud's order: no push state, and the next sort
  // records slot -> cloud index
  a->keys_fresh = a->hist_fresh = 0;
  a->rec_pending = 0;
  a->pushed_fused = 0;
  a->track_cloud = 1;
  a->cloud_valid = 0;
  VH_CHECK_HIP(hipStreamSynchronize(st));
  return VH_OK;
}

static int grow(double **p, int64_t *cap, int64_t need) {
  if (*cap >= need) return VH_OK;
  vh_free(*p);
  *p = nullptr;
  *cap = 0;
  VH_TRY(vh_alloc((void **)p, sizeof(double) * (size_t)need));
  *cap = need;
  return VH_OK;
}

extern "C" int vh_forces_host(vh_accel *a, vh_mesh *fine, const double *csol_f,
                              const double *knots_c, int nkc, const double *csol_c, int nc,
                              double weight, const double *proj, double ball_radius, double escaped,
                              double *forces, double red[4], int32_t *outcount) {
  VH_REQUIRE(a && fine && csol_f && knots_c && csol_c && forces, "vh_forces_host: null argument");
  VH_REQUIRE(fine->n == a->n && fine->ctx == a->ctx,
             "vh_forces_host: fine mesh n=%d (accelerator n=%d) or context differ", fine->n, a->n);
  VH_REQUIRE(nkc >= 2 && nc == 2 * nkc, "vh_forces_host: coarse grid of %d points for %d knots",
             nc, nkc);
  VH_REQUIRE(std::isfinite(weight) && std::isfinite(escaped),
             "vh_forces_host: weight and escaped must be finite");
  if (proj) {
    VH_REQUIRE(proj[4] > 0, "vh_forces_host: projectile scale must be positive");
    VH_REQUIRE(std::isfinite(proj[0]) && std::isfinite(proj[1]) && std::isfinite(proj[2]) &&
                   std::isfinite(proj[3]), "vh_forces_host: projectile state must be finite");
  }
  for (int j = 0; j + 1 < nkc; ++j)
    VH_REQUIRE(knots_c[j + 1] > knots_c[j], "vh_forces_host: coarse knots must increase");
  if (!a->cloud_valid) {
    vh_set_error("vh_forces_host: no sort of a cloud uploaded by vh_positions_upload "
                 "(deposit it first, as update_forces! does)");
    return VH_ESTATE;
  }
  VH_CHECK_HIP(hipSetDevice(a->ctx->device));
  hipStream_t st = a->ctx->stream;
  // the fine coefficients narrowed to the device's precision
  // (acc.csol.host .= csol_fine, src/accelerator.jl:624-626)
  const size_t g = (size_t)fine->n * fine->n * fine->n;
  std::vector<float> cs(g);
  for (size_t t = 0; t < g; ++t) cs[t] = (float)csol_f[t];
  VH_CHECK_HIP(hipMemcpyAsync(fine->csol, cs.data(), g * 4, hipMemcpyHostToDevice, st));
  // the coarse ones stay FP64, as the host fallback reads them
  const int64_t gc = (int64_t)nc * nc * nc;
  VH_TRY(grow(&a->host_csol64, &a->host_csol_cap, gc));
  VH_TRY(grow(&a->host_knots64, &a->host_knots_cap, nkc));
  VH_CHECK_HIP(hipMemcpyAsync(a->host_csol64, csol_c, (size_t)gc * 8, hipMemcpyHostToDevice, st));
  VH_CHECK_HIP(hipMemcpyAsync(a->host_knots64, knots_c, (size_t)nkc * 8, hipMemcpyHostToDevice, st));
  if (!a->force64) VH_TRY(vh_alloc((void **)&a->force64, sizeof(double) * 3 * (size_t)a->npart));
  VH_TRY(vh_rec_settle(a));
  const int tok = vh_prof_begin(a->ctx, VH_ST_FIELD);
  const int rc = vh_forces_impl(a, fine, fine, weight, proj, VH_SOFT_DEVICE, ball_radius, nullptr,
                                0.0, VH_FALLBACK_NONE);
  vh_prof_end(a->ctx, tok);
  if (rc != VH_OK) return rc;
  vh_fallback_args fa;
  fa.px = proj ? proj[0] : 0.0;
  fa.py = proj ? proj[1] : 0.0;
  fa.pz = proj ? proj[2] : 0.0;
  fa.q = proj ? proj[3] : 0.0;
  fa.ball = ball_radius > 0;
  fa.radius = ball_radius > 0 ? ball_radius : (proj ? proj[4] : 1.0);
  fa.has_proj = proj != nullptr;
  fa.w = weight;
  fa.escaped = escaped;
  fa.dev = nullptr;
  hipLaunchKernelGGL(k_forces_host, dim3(2048), dim3(256), 0, st, a->parts[a->cur], a->npart,
                     a->npad, a->cloud_perm, a->pos64, a->force, a->nk, a->n, a->host_csol64,
                     a->host_knots64, nkc, nc, fa, a->force64);
  VH_LAUNCH_CHECK();
  VH_CHECK_HIP(hipMemcpyAsync(forces, a->force64, sizeof(double) * 3 * (size_t)a->npart,
                              hipMemcpyDeviceToHost, st));
  if (red) VH_CHECK_HIP(hipMemcpyAsync(red, a->dscal + 8, 32, hipMemcpyDeviceToHost, st));
  if (outcount)
    VH_CHECK_HIP(hipMemcpyAsync(outcount, a->counters + CNT_OUTCOUNT, 4, hipMemcpyDeviceToHost, st));
  VH_CHECK_HIP(hipStreamSynchronize(st));
  return VH_OK;
}

// ---------------------------------------------------------------------------
// prime_leapfrog!'s second stage (src/simulation.jl:265-273; full_step_back,
// src/particles.jl:466-469) on the device: with the records in slot order
// holding q(-dt/2) (current) and q(0) (previous) and the forces at q(-dt/2),
// each particle becomes q(0) (current) and -q0 + 2 hp + coef f (previous),
// positions rebuilt and split again as the host loop's getindex / setindex!
// do (src/particles.jl:128-145), in FP64 with the same operation order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void split_d(double x, double x0, double h, int &k, float &d) {
  k = (int)rint((x - x0) / h) + 1;
  d = (float)(x - (x0 + (double)(k - 1) * h));
}

__global__ void k_prime(int32_t *__restrict__ P, int64_t npad, int64_t n,
                        const float *__restrict__ force, double x0, double h, double coef) {
  int32_t *Q = vh_prev_half(P, npad);
  for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < n;
       s += (int64_t)gridDim.x * blockDim.x) {
    const vh_pos hp = vh_load_pos(P, s), q0 = vh_load_pos(Q, s);
    vh_pos cur, prev;
    for (int d = 0; d < 3; ++d) {
      const double q = (x0 + (double)(q0.k[d] - 1) * h) + (double)q0.d[d];
      const double m = (x0 + (double)(hp.k[d] - 1) * h) + (double)hp.d[d];
      const double f = (double)force[d * npad + s];
      const double back = (-q + 2.0 * m) + coef * f;
      split_d(q, x0, h, cur.k[d], cur.d[d]);
      split_d(back, x0, h, prev.k[d], prev.d[d]);
    }
    vh_store_pos(P, s, cur);
    vh_store_pos(Q, s, prev);
  }
}

extern "C" int vh_prime(vh_accel *a, double coef) {
  VH_REQUIRE(a != nullptr, "vh_prime: null accelerator");
  VH_REQUIRE(std::isfinite(coef), "vh_prime: coefficient must be finite");
  if (!a->sorted_once) {
    vh_set_error("vh_prime: the forces of the half step have not been computed");
    return VH_ESTATE;
  }
  VH_CHECK_HIP(hipSetDevice(a->ctx->device));
  VH_TRY(vh_rec_settle(a));   // records in slot order, as the forces are
  hipLaunchKernelGGL(k_prime, dim3(2048), dim3(256), 0, a->ctx->stream, a->parts[a->cur], a->npad,
                     a->npart, a->force, a->x0, a->h, coef);
  VH_LAUNCH_CHECK();
  a->keys_fresh = a->hist_fresh = 0;   // new positions
  a->pushed_fused = 0;
  a->track_cloud = a->cloud_valid = 0;
  return VH_OK;
}
