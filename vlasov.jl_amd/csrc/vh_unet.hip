// The learned denoiser: a three-level 3-D U-Net (16 / 32 / 64 channels, 3^3
// convolutions, zero padding) mapping the smoothed density deposited by N
// particles to the density 10 N particles deposit, trained and run on the
// device with the kernels below.  No counterpart in the reference.
//
// Tensors are channels-last FP32, [z][y][x][c] over a cube of side D, one
// sample.  The weights of a convolution are kept as FP32 master values (for
// the optimiser) and as their bf16 roundings widened to FP32, which is what
// every product uses: BF16 weights, FP32 accumulation.  A weight tensor is
// [tap][cin][cout] (tap = 9 dz + 3 dy + dx over the offsets -1..1), so that a
// wave reads the weights of one (tap, cin) at one address: scalar loads.
//
// Network (the layer index is the order of the weight blocks):
//   L0  1 -> 16, ReLU     L1 16 -> 16, ReLU                      level 1 (D)
//   avg-pool 2^3;  L2 16 -> 32, ReLU   L3 32 -> 32, ReLU          level 2 (D/2)
//   avg-pool 2^3;  L4 32 -> 64, ReLU   L5 64 -> 64, ReLU          level 3 (D/4)
//   L6 64 -> 32, ReLU at level 3, nearest 2x up, concat [up, L3] -> L7 64 -> 32, ReLU
//   L8 32 -> 16, ReLU at level 2, nearest 2x up, concat [up, L1] -> L9 32 -> 16, ReLU
//   L10 16 -> 1 (no activation): the correction o, and y = x + m o with
//   m = [|x| > support]: no correction where the deposited density is
//   (nearly) zero, so that the net moves no charge into the vacuum.
// x = rho / rho0 with rho0 the jellium's mean density (the input scale).
#include "vh_internal.h"

#include <cmath>
#include <cstring>
#include <random>

namespace {

constexpr int NL = 11;
constexpr int CIN[NL] = {1, 16, 16, 32, 32, 64, 64, 64, 32, 32, 16};
constexpr int COUT[NL] = {16, 16, 32, 32, 64, 64, 32, 32, 16, 16, 1};
constexpr bool RELU[NL] = {true, true, true, true, true, true, true, true, true, true, false};

constexpr int64_t wsize(int l) { return 27LL * CIN[l] * COUT[l]; }
int64_t woff(int l) {
  int64_t o = 0;
  for (int i = 0; i < l; ++i) o += wsize(i) + COUT[i];
  return o;
}
int64_t nparams() { return woff(NL); }

inline float bf16_round(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  float r;
  std::memcpy(&r, &u, 4);
  return r;
}

}  // namespace

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
__device__ __forceinline__ float bf16_round_d(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return __uint_as_float(u & 0xffff0000u);
}

// out[v][oc0 + co] = b[co] + sum_tap sum_ci w[tap][ci][co] in[v + off(tap)][ic0 + ci],
// ReLU when relu; in has ldi channels per voxel, out ldo.
template <int CO>
__global__ void __launch_bounds__(256)
    k_conv3(const float *__restrict__ in, int ldi, int ic0, int cin, const float *__restrict__ w,
            const float *__restrict__ b, float *__restrict__ out, int ldo, int oc0, int relu,
            int D) {
  const int64_t nv = (int64_t)D * D * D;
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < nv;
       v += (int64_t)gridDim.x * blockDim.x) {
    const int x = (int)(v % D), y = (int)((v / D) % D), z = (int)(v / ((int64_t)D * D));
    float acc[CO];
#pragma unroll
    for (int co = 0; co < CO; ++co) acc[co] = b ? b[co] : 0.f;
    for (int tap = 0; tap < 27; ++tap) {
      const int xx = x + tap % 3 - 1, yy = y + (tap / 3) % 3 - 1, zz = z + tap / 9 - 1;
      const bool ok = (unsigned)xx < (unsigned)D && (unsigned)yy < (unsigned)D &&
                      (unsigned)zz < (unsigned)D;
      const float *src = in + ic0 + (ok ? (((int64_t)zz * D + yy) * D + xx) * ldi : 0);
      const float *wt = w + (int64_t)tap * cin * CO;
      for (int ci = 0; ci < cin; ++ci) {
        const float a = ok ? src[ci] : 0.f;
#pragma unroll
        for (int co = 0; co < CO; ++co) acc[co] = fmaf(wt[ci * CO + co], a, acc[co]);
      }
    }
    float *dst = out + v * ldo + oc0;
#pragma unroll
    for (int co = 0; co < CO; ++co) dst[co] = relu ? fmaxf(acc[co], 0.f) : acc[co];
  }
}

// Weight and bias gradients of one convolution, as partial sums over G voxel
// ranges (block (tap, g)); thread t owns pairs p = t + 256 k, (ci, co) =
// (p / cout, p % cout).  part: [G][27][cin][cout], bpart: [G][cout] (tap 0
// blocks only).  Fixed order throughout.
__global__ void __launch_bounds__(256)
    k_conv3_wgrad(const float *__restrict__ in, int ldi, int ic0, int cin,
                  const float *__restrict__ dpre, int ldd, int cout, int D, int G,
                  float *__restrict__ part, float *__restrict__ bpart) {
  const int tap = blockIdx.x, g = blockIdx.y;
  const int dx = tap % 3 - 1, dy = (tap / 3) % 3 - 1, dz = tap / 9 - 1;
  const int64_t nv = (int64_t)D * D * D;
  const int64_t v0 = nv * g / G, v1 = nv * (g + 1) / G;
  const int npair = cin * cout;
  constexpr int MAXK = 16;   // 64 x 64 pairs / 256 threads
  float acc[MAXK];
#pragma unroll
  for (int k = 0; k < MAXK; ++k) acc[k] = 0.f;
  float bacc = 0.f;
  for (int64_t v = v0; v < v1; ++v) {
    const int x = (int)(v % D), y = (int)((v / D) % D), z = (int)(v / ((int64_t)D * D));
    const int xx = x + dx, yy = y + dy, zz = z + dz;
    const bool ok = (unsigned)xx < (unsigned)D && (unsigned)yy < (unsigned)D &&
                    (unsigned)zz < (unsigned)D;
    const float *src = in + ic0 + (ok ? (((int64_t)zz * D + yy) * D + xx) * ldi : 0);
    const float *dp = dpre + v * ldd;
#pragma unroll
    for (int k = 0; k < MAXK; ++k) {
      const int p = threadIdx.x + 256 * k;
      if (p < npair) {
        const int ci = p / cout, co = p - ci * cout;
        acc[k] = fmaf(ok ? src[ci] : 0.f, dp[co], acc[k]);
      }
    }
    if (tap == 0 && (int)threadIdx.x < cout) bacc += dp[threadIdx.x];
  }
  float *pp = part + ((int64_t)g * 27 + tap) * npair;
#pragma unroll
  for (int k = 0; k < MAXK; ++k) {
    const int p = threadIdx.x + 256 * k;
    if (p < npair) pp[p] = acc[k];
  }
  if (tap == 0 && (int)threadIdx.x < cout) bpart[(int64_t)g * cout + threadIdx.x] = bacc;
}

// grad[i] = sum over g of part[g][i], in g order
__global__ void k_sum_parts(const float *__restrict__ part, int64_t len, int G,
                            float *__restrict__ grad) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < len;
       i += (int64_t)gridDim.x * blockDim.x) {
    float s = 0.f;
    for (int g = 0; g < G; ++g) s += part[(int64_t)g * len + i];
    grad[i] += s;   // accumulated over the samples of a batch
  }
}

// 2^3 average pool of channels [ic0, ic0 + c) of in (D^3, ldi) into out ((D/2)^3, c)
__global__ void k_pool(const float *__restrict__ in, int ldi, int ic0, int c, int D,
                       float *__restrict__ out) {
  const int H = D / 2;
  const int64_t n = (int64_t)H * H * H * c;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int ch = (int)(t % c);
    const int64_t v = t / c;
    const int x = (int)(v % H), y = (int)((v / H) % H), z = (int)(v / ((int64_t)H * H));
    float s = 0.f;
    for (int k = 0; k < 8; ++k) {
      const int64_t u = ((int64_t)(2 * z + (k >> 2)) * D + 2 * y + ((k >> 1) & 1)) * D + 2 * x + (k & 1);
      s += in[u * ldi + ic0 + ch];
    }
    out[t] = s * 0.125f;
  }
}

// the pool's adjoint: din[fine][ic0 + ch] += dout[coarse][ch] / 8
__global__ void k_pool_back(const float *__restrict__ dout, int c, int D, float *__restrict__ din,
                            int ldi, int ic0) {
  const int H = D / 2;
  const int64_t n = (int64_t)D * D * D * c;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int ch = (int)(t % c);
    const int64_t v = t / c;
    const int x = (int)(v % D), y = (int)((v / D) % D), z = (int)(v / ((int64_t)D * D));
    const int64_t p = ((int64_t)(z / 2) * H + y / 2) * H + x / 2;
    din[v * ldi + ic0 + ch] += 0.125f * dout[p * c + ch];
  }
}

// nearest 2x upsampling of in ((D/2)^3, c) into channels [oc0, oc0 + c) of out (D^3, ldo)
__global__ void k_up(const float *__restrict__ in, int c, int D, float *__restrict__ out, int ldo,
                     int oc0) {
  const int H = D / 2;
  const int64_t n = (int64_t)D * D * D * c;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int ch = (int)(t % c);
    const int64_t v = t / c;
    const int x = (int)(v % D), y = (int)((v / D) % D), z = (int)(v / ((int64_t)D * D));
    const int64_t p = ((int64_t)(z / 2) * H + y / 2) * H + x / 2;
    out[v * ldo + oc0 + ch] = in[p * c + ch];
  }
}

// the upsampling's adjoint: dout_coarse[p][ch] = sum of its 8 children
__global__ void k_up_back(const float *__restrict__ dfine, int ldo, int oc0, int c, int D,
                          float *__restrict__ dcoarse) {
  const int H = D / 2;
  const int64_t n = (int64_t)H * H * H * c;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int ch = (int)(t % c);
    const int64_t v = t / c;
    const int x = (int)(v % H), y = (int)((v / H) % H), z = (int)(v / ((int64_t)H * H));
    float s = 0.f;
    for (int k = 0; k < 8; ++k) {
      const int64_t u = ((int64_t)(2 * z + (k >> 2)) * D + 2 * y + ((k >> 1) & 1)) * D + 2 * x + (k & 1);
      s += dfine[u * ldo + oc0 + ch];
    }
    dcoarse[t] = s;
  }
}

// copy channels [sc0, sc0 + c) of a (lds) into channels [dc0, ...) of b (ldd);
// accumulate when add
__global__ void k_chan_copy(const float *__restrict__ a, int lds, int sc0, float *__restrict__ b,
                            int ldd, int dc0, int c, int64_t nv, int add) {
  const int64_t n = nv * c;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int ch = (int)(t % c);
    const int64_t v = t / c;
    const float x = a[v * lds + sc0 + ch];
    float *d = b + v * ldd + dc0 + ch;
    *d = add ? *d + x : x;
  }
}

// dpre = dout where the ReLU output was positive (in place, c channels of ld)
__global__ void k_relu_back(const float *__restrict__ act, int lda, int ac0, float *__restrict__ d,
                            int ldd, int c, int64_t nv) {
  const int64_t n = nv * c;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int ch = (int)(t % c);
    const int64_t v = t / c;
    if (!(act[v * lda + ac0 + ch] > 0.f)) d[v * ldd + ch] = 0.f;
  }
}

// y = x + m o (m = [|x| > support]); loss partial sums of (y - t)^2 (FP64,
// per block), do = 2 m (y - t) / nv
__global__ void __launch_bounds__(256)
    k_loss(const float *__restrict__ x, const float *__restrict__ o, const float *__restrict__ t,
           int64_t nv, float support, float *__restrict__ dout, double *__restrict__ part) {
  __shared__ double lds[8];
  double s = 0.0;
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < nv;
       v += (int64_t)gridDim.x * blockDim.x) {
    const bool m = fabsf(x[v]) > support;
    const float r = (m ? x[v] + o[v] : x[v]) - t[v];
    s += (double)r * r;
    if (dout) dout[v] = m ? 2.f * r / (float)nv : 0.f;
  }
  const double b = vh_block_sum_d(s, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = b;
}

// out[0] = sum of p[0..n) in a fixed order (one block)
__global__ void __launch_bounds__(256) k_sum_d(const double *__restrict__ p, int64_t n,
                                               double *__restrict__ out) {
  __shared__ double lds[8];
  const double s = vh_block_sum_d(vh_thread_sum(p, n), lds);
  if (threadIdx.x == 0) out[0] = s;
}

// Adam on the master weights, then their bf16 roundings for the products
__global__ void k_adam(float *__restrict__ w, float *__restrict__ weff, float *__restrict__ m,
                       float *__restrict__ v, const float *__restrict__ g, int64_t n, float lr,
                       float b1, float b2, float c1, float c2, float eps, float gscale) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float gi = g[i] * gscale;
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    weff[i] = bf16_round_d(wi);
  }
}

// the transposed, flipped weights of the backward-data convolution:
// wt[tap'][co][ci] = w[26 - tap'][ci][co]
__global__ void k_wflip(const float *__restrict__ w, int cin, int cout, float *__restrict__ wt) {
  const int64_t n = 27LL * cin * cout;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int ci = (int)(t % cin);
    const int co = (int)((t / cin) % cout);
    const int tp = (int)(t / ((int64_t)cin * cout));
    wt[t] = w[((int64_t)(26 - tp) * cin + ci) * cout + co];
  }
}

// crops of side C at (z0, y0, x0) of a pair of n^3 grids (i fastest), scaled
// by 1/rho0: the network's input x and target t
__global__ void k_crop(const float *__restrict__ lo, const float *__restrict__ hi, int n, int C,
                       int z0, int y0, int x0, float inv, float *__restrict__ x, float *__restrict__ t) {
  const int64_t nv = (int64_t)C * C * C;
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < nv;
       v += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(v % C), j = (int)((v / C) % C), k = (int)(v / ((int64_t)C * C));
    const int64_t g = ((int64_t)(k + z0) * n + (j + y0)) * n + (i + x0);
    x[v] = lo[g] * inv;
    if (t) t[v] = hi ? hi[g] * inv : 0.f;
  }
}

// the n^3 density into a P^3 cube (P >= n), zero beyond, scaled by 1/rho0
__global__ void k_pad_in(const float *__restrict__ rho, int n, int P, float inv, float *__restrict__ x) {
  const int64_t nv = (int64_t)P * P * P;
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < nv;
       v += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(v % P), j = (int)((v / P) % P), k = (int)(v / ((int64_t)P * P));
    x[v] = (i < n && j < n && k < n) ? rho[((int64_t)k * n + j) * n + i] * inv : 0.f;
  }
}

// rho_out = rho0 (x + m o) on the n^3 grid (into a work grid)
__global__ void k_unpad_out(const float *__restrict__ x, const float *__restrict__ o, int n, int P,
                            float rho0, float support, float *__restrict__ out) {
  const int64_t nv = (int64_t)n * n * n;
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < nv;
       v += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(v % n), j = (int)((v / n) % n), k = (int)(v / ((int64_t)n * n));
    const int64_t p = ((int64_t)k * P + j) * P + i;
    out[v] = rho0 * (fabsf(x[p]) > support ? x[p] + o[p] : x[p]);
  }
}

// The residual charge check: with q the plain density's total charge and qd
// the network's, rho := rho_net q / qd when |qd / q - 1| <= tol, else rho
// stays the plain density (the fallback, counted in fallbacks).
__global__ void k_select(float *__restrict__ rho, const float *__restrict__ net, int64_t n,
                         const double *__restrict__ q, const double *__restrict__ qd, double tol,
                         int *__restrict__ fallbacks) {
  const double r = qd[0] / q[0];
  const bool ok = q[0] > 0.0 && fabs(r - 1.0) <= tol;
  if (!ok) {
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(fallbacks, 1);
    return;
  }
  const float f = (float)(q[0] / qd[0]);
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x)
    rho[v] = net[v] * f;
}


// ---------------------------------------------------------------------------
// inference on the matrix cores
//
// Activations in bf16, channels-last, each level's tensors stored with a
// zero border: voxel (z, y, x) of a level of interior side D at
// ((z + 1) Ey + y + 1) Ex + x + 1, Ey = Ex = roundup(D, 8) + 2 and
// roundup(D, 4) + 2 planes, everything outside the interior zero (the
// convolutions' zero padding, read without bounds checks).  A convolution
// is an implicit GEMM out[v][co] = sum_k A[v][k] B[k][co] over k = (tap, ci)
// on v_mfma_f32_16x16x32_bf16: a workgroup stages a brick of BZ x 8 x 8
// output voxels plus its one-voxel halo in LDS (channels padded by 8 so the
// 16 voxel rows of a fragment fall in distinct bank groups), wave w computes
// plane w of the brick (four 16-voxel row tiles) for every output channel,
// with the weights packed in the MFMA B-operand order (bf16 roundings of the
// same weights as the FP32 path) streamed from L2.  Bias, ReLU and the
// rounding to bf16 in the epilogue.  The first layer reads the FP32 density
// itself (x = rho / rho0 rounded to bf16), the last writes
// rho + rho0 o in FP32 into the output grid.  Concatenated inputs
// [up(coarse), skip] are staged straight from the coarse tensor.
// ---------------------------------------------------------------------------
typedef __bf16 un_bf8 __attribute__((ext_vector_type(8)));
typedef float un_f4 __attribute__((ext_vector_type(4)));
typedef unsigned int un_u4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t un_rne(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return u >> 16;
}

// MFMA K steps of a layer with cin input channels: 32 k values a step
__host__ __device__ constexpr int conv_ks(int cin) {
  return cin == 1 ? 1 : cin == 16 ? 14 : 27 * (cin / 32);
}

// (tap, first channel) of the eight k values of k-group lk in step s
__device__ __forceinline__ void conv_k(int cin, int s, int lk, int &tap, int &c0) {
  if (cin >= 32) {
    const int sp = cin / 32;
    tap = s / sp;
    c0 = (s - tap * sp) * 32 + 8 * lk;
  } else {
    tap = 2 * s + (lk >> 1);
    c0 = 8 * (lk & 1);
  }
}

// B operand of layer (cin, cout) from w[tap][cin][cout]: u4 (eight bf16) per
// (step s, column tile nt, lane l), k values as conv_k, zero beyond the taps
// and the channels
__global__ void k_pack_conv(const float *__restrict__ w, int cin, int cout, int ks, int nt,
                            un_u4 *__restrict__ out) {
  const int total = ks * nt * 64;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
    const int l = e & 63, t = (e >> 6) % nt, st = (e >> 6) / nt;
    const int co = 16 * t + (l & 15), lk = l >> 4;
    uint32_t h[8];
    for (int j = 0; j < 8; ++j) {
      int tap, ci;
      if (cin == 1) {
        tap = 8 * lk + j;
        ci = 0;
      } else {
        conv_k(cin, st, lk, tap, ci);
        ci += j;
      }
      const bool ok = tap < 27 && ci < cin && co < cout;
      h[j] = ok ? un_rne(w[((int64_t)tap * cin + ci) * cout + co]) : 0u;
    }
    out[e] = (un_u4){h[0] | (h[1] << 16), h[2] | (h[3] << 16), h[4] | (h[5] << 16), h[6] | (h[7] << 16)};
  }
}

enum { CM_HIDDEN = 0, CM_FIRST = 1, CM_LAST = 2 };

struct conv_args {
  const unsigned short *srcA;  // coarse tensor of the concatenation (CA channels) or null
  int EyA, ExA;
  const unsigned short *srcB;  // same-level tensor (CB channels)
  int Ey, Ex;
  const float *rho;            // CM_FIRST, CM_LAST: the n^3 density
  int n;
  float inv, rho0;             // 1 / rho0, rho0
  float support;               // CM_LAST: no correction where |rho / rho0| <= support
  const un_u4 *wp;             // packed B operand
  const float *bias;
  unsigned short *out;         // CM_HIDDEN: bf16 tensor of COUT channels (Ey, Ex geometry)
  float *net;                  // CM_LAST: n^3 output density
  int D;                       // interior side of this level
  int nbx, nby, nbz;           // bricks per axis
};

template <int CA, int CB, int COUT, int BZ, int MODE>
__global__ void __launch_bounds__(64 * BZ) k_conv_mfma(conv_args a) {
  constexpr int CIN = CA + CB;
  constexpr int SV = MODE == CM_FIRST ? 1 : CIN + 8;   // LDS elements a voxel
  constexpr int KS = conv_ks(CIN);
  constexpr int NT = (COUT + 15) / 16;
  constexpr int HV = (BZ + 2) * 100;
  extern __shared__ __align__(16) unsigned short sm[];
  const int nb = a.nbx * a.nby * a.nbz;
  // XCD-aware: the bricks of one L2 are a contiguous range
  const int per = (nb + 7) / 8;
  const int brick = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  if (brick >= nb) return;
  const int bx = brick % a.nbx, by = (brick / a.nbx) % a.nby, bz = brick / (a.nbx * a.nby);
  const int x0 = 8 * bx, y0 = 8 * by, z0 = BZ * bz;
  // stage the brick and its halo.  Every thread's loads are issued
  // unconditionally (from a valid address, the value selected afterwards)
  // and in batches before their LDS writes: a load under a branch or in a
  // rolled loop is waited for on its own.
  constexpr int NTH = 64 * BZ;
  if constexpr (MODE == CM_FIRST) {
    float *f = (float *)sm;
    constexpr int NE = (HV + NTH - 1) / NTH;
    float v[NE];
    bool in[NE];
#pragma unroll
    for (int k = 0; k < NE; ++k) {
      const int e0 = threadIdx.x + k * NTH, e = e0 < HV ? e0 : 0;
      const int hx = e % 10, hy = (e / 10) % 10, hz = e / 100;
      const int z = z0 + hz - 1, y = y0 + hy - 1, x = x0 + hx - 1;
      in[k] = (unsigned)z < (unsigned)a.n && (unsigned)y < (unsigned)a.n && (unsigned)x < (unsigned)a.n;
      v[k] = a.rho[in[k] ? ((int64_t)z * a.n + y) * a.n + x : 0];
    }
#pragma unroll
    for (int k = 0; k < NE; ++k) {
      const int e0 = threadIdx.x + k * NTH;
      if (e0 < HV) f[e0] = in[k] ? v[k] * a.inv : 0.f;
    }
  } else {
    constexpr int CH = CIN / 8;
    constexpr int NE = (HV * CH + NTH - 1) / NTH;
    constexpr int NB = NE < 12 ? NE : 12;   // loads in flight a batch
#pragma unroll
    for (int k0 = 0; k0 < NE; k0 += NB) {
      un_u4 v[NB];
      bool zr[NB];
#pragma unroll
      for (int k = 0; k < NB; ++k) {
        const int e0 = threadIdx.x + (k0 + k) * NTH;
        const int e = (k0 + k < NE && e0 < HV * CH) ? e0 : 0;
        const int hv = e / CH, c = 8 * (e - hv * CH);
        const int hx = hv % 10, hy = (hv / 10) % 10, hz = hv / 100;
        const int z = z0 + hz - 1, y = y0 + hy - 1, x = x0 + hx - 1;
        const int64_t sv = ((int64_t)(z + 1) * a.Ey + y + 1) * a.Ex + x + 1;
        const unsigned short *src = a.srcB + sv * CB + (c - CA);
        bool zero = false;
        if (CA > 0) {
          const bool up = c < CA;
          const bool inD = (unsigned)z < (unsigned)a.D && (unsigned)y < (unsigned)a.D &&
                           (unsigned)x < (unsigned)a.D;
          const int64_t cv = inD ? ((int64_t)((z >> 1) + 1) * a.EyA + (y >> 1) + 1) * a.ExA + (x >> 1) + 1 : 0;
          src = up ? a.srcA + cv * CA + c : src;
          zero = up && !inD;
        }
        v[k] = *(const un_u4 *)src;
        zr[k] = zero;
      }
#pragma unroll
      for (int k = 0; k < NB; ++k) {
        const int e0 = threadIdx.x + (k0 + k) * NTH;
        if (k0 + k < NE && e0 < HV * CH) {
          const int hv = e0 / CH, c = 8 * (e0 - hv * CH);
          *(un_u4 *)(sm + hv * SV + c) = zr[k] ? (un_u4){0u, 0u, 0u, 0u} : v[k];
        }
      }
    }
  }
  __syncthreads();
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, lc = l & 15, lk = l >> 4;
  un_f4 acc[4][NT];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[mt][t] = (un_f4){0.f, 0.f, 0.f, 0.f};
  const un_u4 *wp = a.wp + l;
  // the weights of step st (clamped: a load past the last step reads it again)
  auto ldb = [&](un_u4 (&b)[NT], int st) {
    const int s2 = st < KS ? st : KS - 1;
#pragma unroll
    for (int t = 0; t < NT; ++t) b[t] = wp[(s2 * NT + t) * 64];
  };
  auto step = [&](int st, const un_u4 (&b)[NT]) {
    if constexpr (MODE == CM_FIRST) {
      const float *f = (const float *)sm;
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) {
        const int vy = 2 * mt + (lc >> 3), vx = lc & 7;
        uint32_t h[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int tap = 8 * lk + j;
          const int tp = tap < 27 ? tap : 0;
          const int dz = tp / 9, dy = (tp / 3) % 3, dx = tp % 3;
          const float fv = f[((w + dz) * 10 + vy + dy) * 10 + vx + dx];
          h[j] = tap < 27 ? un_rne(fv) : 0u;
        }
        const un_u4 av = (un_u4){h[0] | (h[1] << 16), h[2] | (h[3] << 16), h[4] | (h[5] << 16),
                                 h[6] | (h[7] << 16)};
#pragma unroll
        for (int t = 0; t < NT; ++t)
          acc[mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(un_bf8, av),
                                                               __builtin_bit_cast(un_bf8, b[t]),
                                                               acc[mt][t], 0, 0, 0);
      }
    } else {
      int tap, c0;
      conv_k(CIN, st, lk, tap, c0);
      const bool ok = tap < 27;
      const int tp = ok ? tap : 0;
      const int dz = tp / 9, dy = (tp / 3) % 3, dx = tp % 3;
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) {
        const int vy = 2 * mt + (lc >> 3), vx = lc & 7;
        un_u4 av = *(const un_u4 *)(sm + (((w + dz) * 10 + vy + dy) * 10 + vx + dx) * SV + c0);
        if (!ok) av = (un_u4){0u, 0u, 0u, 0u};
#pragma unroll
        for (int t = 0; t < NT; ++t)
          acc[mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(un_bf8, av),
                                                               __builtin_bit_cast(un_bf8, b[t]),
                                                               acc[mt][t], 0, 0, 0);
      }
    }
  };
  // the weights two steps ahead of their products
  un_u4 bA[NT], bB[NT], bC[NT];
  ldb(bA, 0);
  ldb(bB, 1);
  for (int st = 0; st < KS; st += 3) {
    ldb(bC, st + 2);
    step(st, bA);
    ldb(bA, st + 3);
    if (st + 1 < KS) step(st + 1, bB);
    ldb(bB, st + 4);
    if (st + 2 < KS) step(st + 2, bC);
  }
  // epilogue: row 4 lk + v of row tile mt is voxel (2 mt + row / 8, row % 8)
  const int z = z0 + w;
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = 4 * lk + v;
      const int y = y0 + 2 * mt + (row >> 3), x = x0 + (row & 7);
      if (MODE == CM_LAST) {
        if (lc == 0 && z < a.n && y < a.n && x < a.n) {
          const int64_t g = ((int64_t)z * a.n + y) * a.n + x;
          const float r = a.rho[g];
          a.net[g] = fabsf(r * a.inv) > a.support ? r + a.rho0 * (acc[mt][0][v] + a.bias[0]) : r;
        }
      } else if (z < a.D && y < a.D && x < a.D) {
        const int64_t ov = ((int64_t)(z + 1) * a.Ey + y + 1) * a.Ex + x + 1;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          const int co = 16 * t + lc;
          a.out[ov * COUT + co] = (unsigned short)un_rne(fmaxf(acc[mt][t][v] + a.bias[co], 0.f));
        }
      }
    }
}

// 2^3 mean pooling of bf16 tensors (C channels): fine (Ey, Ex) -> coarse
// (Eyc, Exc) of interior side Dc; eight channels a thread
__global__ void k_pool_bf16(const unsigned short *__restrict__ fine, int Ey, int Ex,
                            unsigned short *__restrict__ coarse, int Eyc, int Exc, int Dc, int C) {
  const int CH = C / 8;
  const int64_t total = (int64_t)Dc * Dc * Dc * CH;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int c = 8 * (int)(e % CH);
    const int64_t v = e / CH;
    const int x = (int)(v % Dc), y = (int)((v / Dc) % Dc), z = (int)(v / ((int64_t)Dc * Dc));
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int k = 0; k < 8; ++k) {
      const int fz = 2 * z + (k >> 2), fy = 2 * y + ((k >> 1) & 1), fx = 2 * x + (k & 1);
      const un_u4 q = *(const un_u4 *)(fine + (((int64_t)(fz + 1) * Ey + fy + 1) * Ex + fx + 1) * C + c);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        s[2 * j] += __uint_as_float(q[j] << 16);
        s[2 * j + 1] += __uint_as_float(q[j] & 0xffff0000u);
      }
    }
    uint32_t h[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) h[j] = un_rne(0.125f * s[j]);
    *(un_u4 *)(coarse + (((int64_t)(z + 1) * Eyc + y + 1) * Exc + x + 1) * C + c) =
        (un_u4){h[0] | (h[1] << 16), h[2] | (h[3] << 16), h[4] | (h[5] << 16), h[6] | (h[7] << 16)};
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static unsigned blocks_for(int64_t n) {
  const int64_t b = (n + 255) / 256;
  return (unsigned)(b < 8192 ? (b > 0 ? b : 1) : 8192);
}

struct vh_unet_buffers {
  int D = 0;
  // activations: level 1 (D^3): a0 (16), a1 (16), cat1 (32: up of L8, a1), a9 (16), o (1), x (1), t (1)
  // level 2 ((D/2)^3): p1 (16), a2 (32), a3 (32), cat2 (64: up of L6, a3), a7 (32), a8 (16)
  // level 3 ((D/4)^3): p2 (32), a4 (64), a5 (64), a6 (32)
  float *x = nullptr, *t = nullptr, *a0 = nullptr, *a1 = nullptr, *cat1 = nullptr, *a9 = nullptr,
        *o = nullptr;
  float *p1 = nullptr, *a2 = nullptr, *a3 = nullptr, *cat2 = nullptr, *a7 = nullptr, *a8 = nullptr;
  float *p2 = nullptr, *a4 = nullptr, *a5 = nullptr, *a6 = nullptr;
  // gradients (same shapes), two scratch tensors per level
  float *g1a = nullptr, *g1b = nullptr, *g2a = nullptr, *g2b = nullptr, *g3a = nullptr,
        *g3b = nullptr, *gcat1 = nullptr, *gcat2 = nullptr;
};

struct vh_denoiser {
  vh_context *ctx = nullptr;
  int device = 0;             // ctx's device (destroy does not read ctx, which may be gone)
  int64_t np = 0;
  float *w = nullptr, *weff = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr, *wt = nullptr;
  int64_t step = 0;
  vh_unet_buffers train, infer;
  float *part = nullptr;      // weight-gradient partials
  float *bpart = nullptr;
  int64_t part_cap = 0;
  double *lpart = nullptr;    // loss partials
  double *dscal = nullptr;    // [0] loss, [1] q plain, [2] q net
  float *net = nullptr;       // n^3 denoised density
  int64_t net_cap = 0;
  int *fallbacks = nullptr;
  // the dataset: pairs of n^3 densities (N particles, 10 N particles)
  int dn = 0;
  int64_t dcap = 0, dcount = 0;
  float *data = nullptr;
  double rho0 = 1.0;
  // inference on the matrix cores: packed B operands, bf16 tensors
  un_u4 *packed = nullptr;
  int64_t poff[NL + 1] = {};
  struct level {
    int D = 0, Ey = 0, Ez = 0;
    int64_t voxels() const { return (int64_t)Ez * Ey * Ey; }
  } lv[3];
  int mfP = 0;
  unsigned short *t_a0 = nullptr, *t_a1 = nullptr, *t_p1 = nullptr, *t_a2 = nullptr, *t_a3 = nullptr,
                 *t_p2 = nullptr, *t_a4 = nullptr, *t_a5 = nullptr, *t_a6 = nullptr, *t_a7 = nullptr,
                 *t_a8 = nullptr, *t_a9 = nullptr;
  int fp32_infer = 0;         // vh_denoiser_set_inference
  float support = 1e-4f;      // vh_denoiser_set_support
  int prep_n = 0;             // the grid the inference buffers were made for
};

static void free_mfma(vh_denoiser *d) {
  unsigned short *ps[] = {d->t_a0, d->t_a1, d->t_p1, d->t_a2, d->t_a3, d->t_p2, d->t_a4, d->t_a5,
                          d->t_a6, d->t_a7, d->t_a8, d->t_a9};
  for (unsigned short *p : ps) vh_free(p);
  d->t_a0 = d->t_a1 = d->t_p1 = d->t_a2 = d->t_a3 = d->t_p2 = d->t_a4 = d->t_a5 = d->t_a6 = d->t_a7 =
      d->t_a8 = d->t_a9 = nullptr;
  d->mfP = 0;
}

// the zero-bordered bf16 tensors of a P^3 input (levels P, P/2, P/4),
// zeroed once: only interiors are ever written
static int ensure_mfma(vh_denoiser *d, int P) {
  if (d->mfP == P) return VH_OK;
  free_mfma(d);
  for (int k = 0; k < 3; ++k) {
    const int D = P >> k;
    d->lv[k].D = D;
    d->lv[k].Ey = (D + 7) / 8 * 8 + 2;
    d->lv[k].Ez = (D + 3) / 4 * 4 + 2;
  }
  struct {
    unsigned short **p;
    int level, c;
  } t[] = {{&d->t_a0, 0, 16}, {&d->t_a1, 0, 16}, {&d->t_a9, 0, 16}, {&d->t_p1, 1, 16},
           {&d->t_a2, 1, 32}, {&d->t_a3, 1, 32}, {&d->t_a7, 1, 32}, {&d->t_a8, 1, 16},
           {&d->t_p2, 2, 32}, {&d->t_a4, 2, 64}, {&d->t_a5, 2, 64}, {&d->t_a6, 2, 32}};
  for (auto &e : t) {
    const size_t bytes = sizeof(unsigned short) * e.c * d->lv[e.level].voxels();
    VH_TRY(vh_alloc((void **)e.p, bytes));
    VH_CHECK_HIP(hipMemset(*e.p, 0, bytes));
  }
  d->mfP = P;
  return VH_OK;
}

// the B operands of every layer from the bf16 weights (on the stream)
static int pack_weights(vh_denoiser *d) {
  for (int l = 0; l < NL; ++l) {
    const int ks = conv_ks(CIN[l]), nt = (COUT[l] + 15) / 16;
    hipLaunchKernelGGL(k_pack_conv, dim3((ks * nt * 64 + 255) / 256), dim3(256), 0, d->ctx->stream,
                       d->weff + woff(l), CIN[l], COUT[l], ks, nt, d->packed + d->poff[l]);
  }
  VH_LAUNCH_CHECK();
  return VH_OK;
}

template <int CA, int CB, int COUT_, int BZ, int MODE>
static int launch_conv(vh_denoiser *d, hipStream_t st, int l, int level, const unsigned short *srcA,
                       const unsigned short *srcB, unsigned short *out, const float *rho, int n,
                       float *net) {
  static_assert(CA + CB == 1 || (CA + CB) % 16 == 0, "input channels");
  VH_REQUIRE(CA + CB == CIN[l] && COUT_ == COUT[l], "denoiser: layer %d launched with the wrong shape", l);
  const auto &L = d->lv[level];
  conv_args a;
  a.srcA = srcA;
  a.EyA = level < 2 ? d->lv[level + 1].Ey : 0;
  a.ExA = a.EyA;
  a.srcB = srcB;
  a.Ey = a.Ex = L.Ey;
  a.rho = rho;
  a.n = n;
  a.inv = (float)(1.0 / d->rho0);
  a.rho0 = (float)d->rho0;
  a.support = d->support;
  a.wp = d->packed + d->poff[l];
  a.bias = d->weff + woff(l) + wsize(l);
  a.out = out;
  a.net = net;
  a.D = L.D;
  a.nbx = a.nby = (L.D + 7) / 8;
  a.nbz = (L.D + BZ - 1) / BZ;
  const int nb = a.nbx * a.nby * a.nbz;
  const int grid = 8 * ((nb + 7) / 8);
  constexpr int CIN_ = CA + CB;
  const size_t lds = MODE == CM_FIRST ? sizeof(float) * (BZ + 2) * 100
                                      : sizeof(unsigned short) * (BZ + 2) * 100 * (CIN_ + 8);
  hipLaunchKernelGGL((k_conv_mfma<CA, CB, COUT_, BZ, MODE>), dim3(grid), dim3(64 * BZ), lds, st, a);
  VH_LAUNCH_CHECK();
  return VH_OK;
}

static int pool_bf16(vh_denoiser *d, hipStream_t st, const unsigned short *fine, int level,
                     unsigned short *coarse, int C) {
  const auto &F = d->lv[level], &Cc = d->lv[level + 1];
  const int64_t total = (int64_t)Cc.D * Cc.D * Cc.D * (C / 8);
  hipLaunchKernelGGL(k_pool_bf16, dim3(blocks_for(total)), dim3(256), 0, st, fine, F.Ey, F.Ey, coarse,
                     Cc.Ey, Cc.Ey, Cc.D, C);
  VH_LAUNCH_CHECK();
  return VH_OK;
}

// the network on the matrix cores: rho (n^3) -> net = rho + rho0 o (n^3)
static int forward_mfma(vh_denoiser *d, hipStream_t st, const float *rho, int n, float *net) {
  VH_TRY((launch_conv<0, 1, 16, 4, CM_FIRST>(d, st, 0, 0, nullptr, nullptr, d->t_a0, rho, n, nullptr)));
  VH_TRY((launch_conv<0, 16, 16, 4, CM_HIDDEN>(d, st, 1, 0, nullptr, d->t_a0, d->t_a1, nullptr, n, nullptr)));
  VH_TRY(pool_bf16(d, st, d->t_a1, 0, d->t_p1, 16));
  VH_TRY((launch_conv<0, 16, 32, 4, CM_HIDDEN>(d, st, 2, 1, nullptr, d->t_p1, d->t_a2, nullptr, n, nullptr)));
  VH_TRY((launch_conv<0, 32, 32, 4, CM_HIDDEN>(d, st, 3, 1, nullptr, d->t_a2, d->t_a3, nullptr, n, nullptr)));
  VH_TRY(pool_bf16(d, st, d->t_a3, 1, d->t_p2, 32));
  VH_TRY((launch_conv<0, 32, 64, 4, CM_HIDDEN>(d, st, 4, 2, nullptr, d->t_p2, d->t_a4, nullptr, n, nullptr)));
  VH_TRY((launch_conv<0, 64, 64, 2, CM_HIDDEN>(d, st, 5, 2, nullptr, d->t_a4, d->t_a5, nullptr, n, nullptr)));
  VH_TRY((launch_conv<0, 64, 32, 2, CM_HIDDEN>(d, st, 6, 2, nullptr, d->t_a5, d->t_a6, nullptr, n, nullptr)));
  VH_TRY((launch_conv<32, 32, 32, 2, CM_HIDDEN>(d, st, 7, 1, d->t_a6, d->t_a3, d->t_a7, nullptr, n, nullptr)));
  VH_TRY((launch_conv<0, 32, 16, 4, CM_HIDDEN>(d, st, 8, 1, nullptr, d->t_a7, d->t_a8, nullptr, n, nullptr)));
  VH_TRY((launch_conv<16, 16, 16, 4, CM_HIDDEN>(d, st, 9, 0, d->t_a8, d->t_a1, d->t_a9, nullptr, n, nullptr)));
  VH_TRY((launch_conv<0, 16, 1, 4, CM_LAST>(d, st, 10, 0, nullptr, d->t_a9, nullptr, rho, n, net)));
  return VH_OK;
}

static void free_buffers(vh_unet_buffers &b) {
  float *ps[] = {b.x, b.t, b.a0, b.a1, b.cat1, b.a9, b.o, b.p1, b.a2, b.a3, b.cat2, b.a7, b.a8,
                 b.p2, b.a4, b.a5, b.a6, b.g1a, b.g1b, b.g2a, b.g2b, b.g3a, b.g3b, b.gcat1,
                 b.gcat2};
  for (float *p : ps) vh_free(p);
  b = vh_unet_buffers();
}

static int ensure_buffers(vh_unet_buffers &b, int D, bool grads) {
  if (b.D == D && (!grads || b.g1a)) return VH_OK;
  free_buffers(b);
  const int64_t v1 = (int64_t)D * D * D, v2 = v1 / 8, v3 = v2 / 8;
  struct {
    float **p;
    int64_t n;
  } a[] = {{&b.x, v1}, {&b.t, v1}, {&b.a0, 16 * v1}, {&b.a1, 16 * v1}, {&b.cat1, 32 * v1},
           {&b.a9, 16 * v1}, {&b.o, v1}, {&b.p1, 16 * v2}, {&b.a2, 32 * v2}, {&b.a3, 32 * v2},
           {&b.cat2, 64 * v2}, {&b.a7, 32 * v2}, {&b.a8, 16 * v2}, {&b.p2, 32 * v3},
           {&b.a4, 64 * v3}, {&b.a5, 64 * v3}, {&b.a6, 32 * v3}};
  for (auto &e : a) VH_TRY(vh_alloc((void **)e.p, sizeof(float) * e.n));
  if (grads) {
    struct {
      float **p;
      int64_t n;
    } g[] = {{&b.g1a, 32 * v1}, {&b.g1b, 32 * v1}, {&b.g2a, 64 * v2}, {&b.g2b, 64 * v2},
             {&b.g3a, 64 * v3}, {&b.g3b, 64 * v3}, {&b.gcat1, 32 * v1}, {&b.gcat2, 64 * v2}};
    for (auto &e : g) VH_TRY(vh_alloc((void **)e.p, sizeof(float) * e.n));
  }
  b.D = D;
  return VH_OK;
}

// one convolution of layer l: in (ldi, ic0) -> out (ldo, oc0)
static int conv(vh_denoiser *d, int l, const float *in, int ldi, int ic0, float *out, int ldo,
                int oc0, int D) {
  hipStream_t st = d->ctx->stream;
  const float *w = d->weff + woff(l), *b = w + wsize(l);
  const unsigned g = blocks_for((int64_t)D * D * D);
  const int relu = RELU[l] ? 1 : 0;
  switch (COUT[l]) {
    case 1: hipLaunchKernelGGL(k_conv3<1>, dim3(g), dim3(256), 0, st, in, ldi, ic0, CIN[l], w, b, out, ldo, oc0, relu, D); break;
    case 16: hipLaunchKernelGGL(k_conv3<16>, dim3(g), dim3(256), 0, st, in, ldi, ic0, CIN[l], w, b, out, ldo, oc0, relu, D); break;
    case 32: hipLaunchKernelGGL(k_conv3<32>, dim3(g), dim3(256), 0, st, in, ldi, ic0, CIN[l], w, b, out, ldo, oc0, relu, D); break;
    case 64: hipLaunchKernelGGL(k_conv3<64>, dim3(g), dim3(256), 0, st, in, ldi, ic0, CIN[l], w, b, out, ldo, oc0, relu, D); break;
  }
  VH_LAUNCH_CHECK();
  return VH_OK;
}

// forward through the net on b.x (D^3), leaving every activation and b.o
static int forward(vh_denoiser *d, vh_unet_buffers &b) {
  hipStream_t st = d->ctx->stream;
  const int D = b.D, D2 = D / 2, D4 = D / 4;
  const int64_t v1 = (int64_t)D * D * D, v2 = v1 / 8;
  VH_TRY(conv(d, 0, b.x, 1, 0, b.a0, 16, 0, D));
  VH_TRY(conv(d, 1, b.a0, 16, 0, b.a1, 16, 0, D));
  hipLaunchKernelGGL(k_pool, dim3(blocks_for(16 * v2)), dim3(256), 0, st, b.a1, 16, 0, 16, D, b.p1);
  VH_TRY(conv(d, 2, b.p1, 16, 0, b.a2, 32, 0, D2));
  VH_TRY(conv(d, 3, b.a2, 32, 0, b.a3, 32, 0, D2));
  hipLaunchKernelGGL(k_pool, dim3(blocks_for(32 * v2 / 8)), dim3(256), 0, st, b.a3, 32, 0, 32, D2, b.p2);
  VH_TRY(conv(d, 4, b.p2, 32, 0, b.a4, 64, 0, D4));
  VH_TRY(conv(d, 5, b.a4, 64, 0, b.a5, 64, 0, D4));
  VH_TRY(conv(d, 6, b.a5, 64, 0, b.a6, 32, 0, D4));
  hipLaunchKernelGGL(k_up, dim3(blocks_for(32 * v2)), dim3(256), 0, st, b.a6, 32, D2, b.cat2, 64, 0);
  hipLaunchKernelGGL(k_chan_copy, dim3(blocks_for(32 * v2)), dim3(256), 0, st, b.a3, 32, 0, b.cat2, 64, 32, 32, v2, 0);
  VH_TRY(conv(d, 7, b.cat2, 64, 0, b.a7, 32, 0, D2));
  VH_TRY(conv(d, 8, b.a7, 32, 0, b.a8, 16, 0, D2));
  hipLaunchKernelGGL(k_up, dim3(blocks_for(16 * v1)), dim3(256), 0, st, b.a8, 16, D, b.cat1, 32, 0);
  hipLaunchKernelGGL(k_chan_copy, dim3(blocks_for(16 * v1)), dim3(256), 0, st, b.a1, 16, 0, b.cat1, 32, 16, 16, v1, 0);
  VH_TRY(conv(d, 9, b.cat1, 32, 0, b.a9, 16, 0, D));
  VH_TRY(conv(d, 10, b.a9, 16, 0, b.o, 1, 0, D));
  VH_LAUNCH_CHECK();
  return VH_OK;
}

// the weight gradients' voxel ranges: WG_CHUNK voxels a block, at most
// WG_MAXG ranges (27 WG_MAXG blocks a layer)
constexpr int WG_CHUNK = 256, WG_MAXG = 512;

static int ensure_parts(vh_denoiser *d, int G) {
  const int64_t need = (int64_t)G * 27 * 64 * 64;
  if (d->part_cap >= need) return VH_OK;
  vh_free(d->part);
  vh_free(d->bpart);
  d->part = d->bpart = nullptr;
  VH_TRY(vh_alloc((void **)&d->part, sizeof(float) * need));
  VH_TRY(vh_alloc((void **)&d->bpart, sizeof(float) * (int64_t)G * 64));
  d->part_cap = need;
  return VH_OK;
}

// backward of layer l: from dpre (its pre-activation gradient, ldd = cout)
// into the weight gradient, and, when din, the input gradient (ld, c0)
static int conv_back(vh_denoiser *d, int l, const float *in, int ldi, int ic0, const float *dpre,
                     int D, float *din, int ldin, int dc0) {
  hipStream_t st = d->ctx->stream;
  const int cin = CIN[l], cout = COUT[l];
  const int64_t nv = (int64_t)D * D * D;
  int G = (int)((nv + WG_CHUNK - 1) / WG_CHUNK);
  G = G < 1 ? 1 : (G > WG_MAXG ? WG_MAXG : G);
  VH_TRY(ensure_parts(d, G));
  hipLaunchKernelGGL(k_conv3_wgrad, dim3(27, G), dim3(256), 0, st, in, ldi, ic0, cin, dpre, cout,
                     cout, D, G, d->part, d->bpart);
  VH_LAUNCH_CHECK();
  float *gw = d->grad + woff(l);
  hipLaunchKernelGGL(k_sum_parts, dim3(blocks_for(wsize(l))), dim3(256), 0, st, d->part, wsize(l), G, gw);
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(64), 0, st, d->bpart, (int64_t)cout, G, gw + wsize(l));
  VH_LAUNCH_CHECK();
  if (din) {
    // din (cin channels at dc0 of ldin) = conv(dpre, flipped transposed w)
    hipLaunchKernelGGL(k_wflip, dim3(blocks_for(wsize(l))), dim3(256), 0, st, d->weff + woff(l), cin, cout, d->wt);
    const unsigned g = blocks_for(nv);
    switch (cin) {
      case 16: hipLaunchKernelGGL(k_conv3<16>, dim3(g), dim3(256), 0, st, dpre, cout, 0, cout, d->wt, (const float *)nullptr, din, ldin, dc0, 0, D); break;
      case 32: hipLaunchKernelGGL(k_conv3<32>, dim3(g), dim3(256), 0, st, dpre, cout, 0, cout, d->wt, (const float *)nullptr, din, ldin, dc0, 0, D); break;
      case 64: hipLaunchKernelGGL(k_conv3<64>, dim3(g), dim3(256), 0, st, dpre, cout, 0, cout, d->wt, (const float *)nullptr, din, ldin, dc0, 0, D); break;
      default: VH_REQUIRE(false, "denoiser: no input gradient for %d channels", cin);
    }
    VH_LAUNCH_CHECK();
  }
  return VH_OK;
}

// loss of b.x + b.o against b.t, and (when grads) every weight gradient
// accumulated into d->grad; the sum of squares into *loss (device)
static int forward_backward(vh_denoiser *d, vh_unet_buffers &b, bool grads, double *loss) {
  hipStream_t st = d->ctx->stream;
  VH_TRY(forward(d, b));
  const int D = b.D, D2 = D / 2, D4 = D / 4;
  const int64_t v1 = (int64_t)D * D * D, v2 = v1 / 8, v3 = v2 / 8;
  // loss and dO (into g1a, one channel)
  hipLaunchKernelGGL(k_loss, dim3(512), dim3(256), 0, st, b.x, b.o, b.t, v1, d->support,
                     grads ? b.g1a : nullptr, d->lpart);
  VH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sum_d, dim3(1), dim3(256), 0, st, d->lpart, (int64_t)512, loss);
  VH_LAUNCH_CHECK();
  if (!grads) return VH_OK;
  // L10 (no ReLU): dpre = dO; d a9 into g1b (16)
  VH_TRY(conv_back(d, 10, b.a9, 16, 0, b.g1a, D, b.g1b, 16, 0));
  hipLaunchKernelGGL(k_relu_back, dim3(blocks_for(16 * v1)), dim3(256), 0, st, b.a9, 16, 0, b.g1b, 16, 16, v1);
  // L9: input cat1 (32) -> gcat1
  VH_TRY(conv_back(d, 9, b.cat1, 32, 0, b.g1b, D, b.gcat1, 32, 0));
  // cat1 = [up(a8), a1]: d a8 (level 2) from the first 16 channels
  hipLaunchKernelGGL(k_up_back, dim3(blocks_for(16 * v2)), dim3(256), 0, st, b.gcat1, 32, 0, 16, D, b.g2a);
  hipLaunchKernelGGL(k_relu_back, dim3(blocks_for(16 * v2)), dim3(256), 0, st, b.a8, 16, 0, b.g2a, 16, 16, v2);
  // L8: a7 (32) -> g2b (32)
  VH_TRY(conv_back(d, 8, b.a7, 32, 0, b.g2a, D2, b.g2b, 32, 0));
  hipLaunchKernelGGL(k_relu_back, dim3(blocks_for(32 * v2)), dim3(256), 0, st, b.a7, 32, 0, b.g2b, 32, 32, v2);
  // L7: cat2 (64) -> gcat2
  VH_TRY(conv_back(d, 7, b.cat2, 64, 0, b.g2b, D2, b.gcat2, 64, 0));
  // cat2 = [up(a6), a3]: d a6 (level 3)
  hipLaunchKernelGGL(k_up_back, dim3(blocks_for(32 * v3)), dim3(256), 0, st, b.gcat2, 64, 0, 32, D2, b.g3a);
  hipLaunchKernelGGL(k_relu_back, dim3(blocks_for(32 * v3)), dim3(256), 0, st, b.a6, 32, 0, b.g3a, 32, 32, v3);
  // L6: a5 (64) -> g3b
  VH_TRY(conv_back(d, 6, b.a5, 64, 0, b.g3a, D4, b.g3b, 64, 0));
  hipLaunchKernelGGL(k_relu_back, dim3(blocks_for(64 * v3)), dim3(256), 0, st, b.a5, 64, 0, b.g3b, 64, 64, v3);
  // L5: a4 (64) -> g3a
  VH_TRY(conv_back(d, 5, b.a4, 64, 0, b.g3b, D4, b.g3a, 64, 0));
  hipLaunchKernelGGL(k_relu_back, dim3(blocks_for(64 * v3)), dim3(256), 0, st, b.a4, 64, 0, b.g3a, 64, 64, v3);
  // L4: p2 (32) -> g3b (32)
  VH_TRY(conv_back(d, 4, b.p2, 32, 0, b.g3a, D4, b.g3b, 32, 0));
  // p2 = pool(a3): d a3 = skip part of gcat2 (channels 32..63) + pool adjoint
  hipLaunchKernelGGL(k_chan_copy, dim3(blocks_for(32 * v2)), dim3(256), 0, st, b.gcat2, 64, 32, b.g2a, 32, 0, 32, v2, 0);
  hipLaunchKernelGGL(k_pool_back, dim3(blocks_for(32 * v2)), dim3(256), 0, st, b.g3b, 32, D2, b.g2a, 32, 0);
  hipLaunchKernelGGL(k_relu_back, dim3(blocks_for(32 * v2)), dim3(256), 0, st, b.a3, 32, 0, b.g2a, 32, 32, v2);
  // L3: a2 (32) -> g2b
  VH_TRY(conv_back(d, 3, b.a2, 32, 0, b.g2a, D2, b.g2b, 32, 0));
  hipLaunchKernelGGL(k_relu_back, dim3(blocks_for(32 * v2)), dim3(256), 0, st, b.a2, 32, 0, b.g2b, 32, 32, v2);
  // L2: p1 (16) -> g2a (16)
  VH_TRY(conv_back(d, 2, b.p1, 16, 0, b.g2b, D2, b.g2a, 16, 0));
  // p1 = pool(a1): d a1 = skip part of gcat1 (channels 16..31) + pool adjoint
  hipLaunchKernelGGL(k_chan_copy, dim3(blocks_for(16 * v1)), dim3(256), 0, st, b.gcat1, 32, 16, b.g1a, 16, 0, 16, v1, 0);
  hipLaunchKernelGGL(k_pool_back, dim3(blocks_for(16 * v1)), dim3(256), 0, st, b.g2a, 16, D, b.g1a, 16, 0);
  hipLaunchKernelGGL(k_relu_back, dim3(blocks_for(16 * v1)), dim3(256), 0, st, b.a1, 16, 0, b.g1a, 16, 16, v1);
  // L1: a0 (16) -> g1b
  VH_TRY(conv_back(d, 1, b.a0, 16, 0, b.g1a, D, b.g1b, 16, 0));
  hipLaunchKernelGGL(k_relu_back, dim3(blocks_for(16 * v1)), dim3(256), 0, st, b.a0, 16, 0, b.g1b, 16, 16, v1);
  // L0: x (1), no input gradient
  VH_TRY(conv_back(d, 0, b.x, 1, 0, b.g1b, D, nullptr, 0, 0));
  VH_LAUNCH_CHECK();
  return VH_OK;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" int vh_denoiser_param_count(int64_t *n) {
  VH_REQUIRE(n != nullptr, "vh_denoiser_param_count: null argument");
  *n = nparams();
  return VH_OK;
}

extern "C" int vh_denoiser_destroy(vh_denoiser *d) {
  if (!d) return VH_OK;
  (void)hipSetDevice(d->device);
  (void)hipDeviceSynchronize();
  free_buffers(d->train);
  free_buffers(d->infer);
  free_mfma(d);
  vh_free(d->packed);
  void *ps[] = {d->w, d->weff, d->m, d->v, d->grad, d->wt, d->part, d->bpart, d->lpart, d->dscal,
                d->net, d->fallbacks, d->data};
  for (void *p : ps) vh_free(p);
  delete d;
  return VH_OK;
}

extern "C" int vh_denoiser_create(vh_context *ctx, const float *weights, int64_t nweights,
                                  uint64_t seed, double rho0, vh_denoiser **out) {
  VH_REQUIRE(ctx && out, "vh_denoiser_create: null argument");
  VH_REQUIRE(rho0 > 0 && std::isfinite(rho0), "vh_denoiser_create: rho0 must be positive");
  const int64_t np = nparams();
  VH_REQUIRE(!weights || nweights == np, "vh_denoiser_create: %lld weights for %lld parameters",
             (long long)nweights, (long long)np);
  *out = nullptr;
  VH_CHECK_HIP(hipSetDevice(ctx->device));
  vh_denoiser *d = new vh_denoiser();
  d->ctx = ctx;
  d->device = ctx->device;
  d->np = np;
  d->rho0 = rho0;
  std::vector<float> h(np, 0.f);
  if (weights) {
    std::memcpy(h.data(), weights, sizeof(float) * np);
  } else {
    // He initialisation of the weights (normal, std sqrt(2 / fan_in)),
    // zero biases; the last layer starts at zero, so the untrained net is
    // the identity on its input
    std::mt19937_64 rng(seed);
    for (int l = 0; l < NL; ++l) {
      std::normal_distribution<double> nd(0.0, std::sqrt(2.0 / (27.0 * CIN[l])));
      float *w = h.data() + woff(l);
      for (int64_t i = 0; i < wsize(l); ++i) w[i] = l == NL - 1 ? 0.f : (float)nd(rng);
    }
  }
  std::vector<float> e(np);
  for (int64_t i = 0; i < np; ++i) e[i] = bf16_round(h[i]);
  int rc = VH_OK;
#define DK(x) if (rc == VH_OK) rc = (x)
  DK(vh_alloc((void **)&d->w, sizeof(float) * np));
  DK(vh_alloc((void **)&d->weff, sizeof(float) * np));
  DK(vh_alloc((void **)&d->m, sizeof(float) * np));
  DK(vh_alloc((void **)&d->v, sizeof(float) * np));
  DK(vh_alloc((void **)&d->grad, sizeof(float) * np));
  DK(vh_alloc((void **)&d->wt, sizeof(float) * 27 * 64 * 64));
  DK(vh_alloc((void **)&d->lpart, sizeof(double) * 512));
  DK(vh_alloc((void **)&d->dscal, sizeof(double) * 8));
  DK(vh_alloc((void **)&d->fallbacks, sizeof(int)));
#undef DK
  if (rc == VH_OK && (hipMemcpy(d->w, h.data(), sizeof(float) * np, hipMemcpyHostToDevice) != hipSuccess ||
                      hipMemcpy(d->weff, e.data(), sizeof(float) * np, hipMemcpyHostToDevice) != hipSuccess ||
                      hipMemset(d->m, 0, sizeof(float) * np) != hipSuccess ||
                      hipMemset(d->v, 0, sizeof(float) * np) != hipSuccess ||
                      hipMemset(d->grad, 0, sizeof(float) * np) != hipSuccess ||
                      hipMemset(d->fallbacks, 0, sizeof(int)) != hipSuccess)) {
    vh_set_error("vh_denoiser_create: upload failed");
    rc = VH_EHIP;
  }
  if (rc == VH_OK) {
    int64_t o = 0;
    for (int l = 0; l < NL; ++l) {
      d->poff[l] = o;
      o += (int64_t)conv_ks(CIN[l]) * ((COUT[l] + 15) / 16) * 64;
    }
    d->poff[NL] = o;
    rc = vh_alloc((void **)&d->packed, sizeof(un_u4) * o);
    if (rc == VH_OK) rc = pack_weights(d);
    if (rc == VH_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) {
      vh_set_error("vh_denoiser_create: packing the weights failed");
      rc = VH_EHIP;
    }
  }
  if (rc != VH_OK) {
    vh_denoiser_destroy(d);
    return rc;
  }
  *out = d;
  return VH_OK;
}

// new master weights: their bf16 roundings, packed operands, and a fresh
// optimiser (zero moments, step 0)
__global__ void k_set_weights(float *__restrict__ w, float *__restrict__ weff, float *__restrict__ m,
                              float *__restrict__ v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    weff[i] = bf16_round_d(w[i]);
    m[i] = 0.f;
    v[i] = 0.f;
  }
}

extern "C" int vh_denoiser_set_weights(vh_denoiser *d, const float *w, int64_t nweights) {
  VH_REQUIRE(d && w, "vh_denoiser_set_weights: null argument");
  VH_REQUIRE(nweights == d->np, "vh_denoiser_set_weights: %lld weights for %lld parameters",
             (long long)nweights, (long long)d->np);
  VH_CHECK_HIP(hipSetDevice(d->ctx->device));
  hipStream_t st = d->ctx->stream;
  VH_CHECK_HIP(hipMemcpyAsync(d->w, w, sizeof(float) * d->np, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_set_weights, dim3(blocks_for(d->np)), dim3(256), 0, st, d->w, d->weff, d->m, d->v,
                     d->np);
  VH_LAUNCH_CHECK();
  VH_TRY(pack_weights(d));
  VH_CHECK_HIP(hipStreamSynchronize(st));
  d->step = 0;
  return VH_OK;
}

extern "C" int vh_denoiser_get_weights(vh_denoiser *d, float *w, float *weff) {
  VH_REQUIRE(d != nullptr, "vh_denoiser_get_weights: null denoiser");
  VH_CHECK_HIP(hipSetDevice(d->ctx->device));
  if (w) VH_CHECK_HIP(hipMemcpyAsync(w, d->w, sizeof(float) * d->np, hipMemcpyDeviceToHost, d->ctx->stream));
  if (weff)
    VH_CHECK_HIP(hipMemcpyAsync(weff, d->weff, sizeof(float) * d->np, hipMemcpyDeviceToHost, d->ctx->stream));
  VH_CHECK_HIP(hipStreamSynchronize(d->ctx->stream));
  return VH_OK;
}

// One sample through the net, for checks: x and t (D^3 each, scaled
// inputs), the output correction o (may be NULL), the loss mean((x + o - t)^2)
// and the gradient of every parameter (may be NULL; the loss's gradient, not
// accumulated).
extern "C" int vh_denoiser_eval(vh_denoiser *d, int D, const float *x, const float *t, float *o,
                                double *loss, float *grad) {
  VH_REQUIRE(d && x && t, "vh_denoiser_eval: null argument");
  VH_REQUIRE(D >= 4 && D % 4 == 0 && D <= 160, "vh_denoiser_eval: side %d not a multiple of 4 in [4, 160]", D);
  VH_CHECK_HIP(hipSetDevice(d->ctx->device));
  hipStream_t st = d->ctx->stream;
  vh_unet_buffers &b = d->train;
  VH_TRY(ensure_buffers(b, D, true));
  const int64_t nv = (int64_t)D * D * D;
  VH_CHECK_HIP(hipMemcpyAsync(b.x, x, sizeof(float) * nv, hipMemcpyHostToDevice, st));
  VH_CHECK_HIP(hipMemcpyAsync(b.t, t, sizeof(float) * nv, hipMemcpyHostToDevice, st));
  VH_CHECK_HIP(hipMemsetAsync(d->grad, 0, sizeof(float) * d->np, st));
  VH_TRY(forward_backward(d, b, grad != nullptr, d->dscal));
  if (o) VH_CHECK_HIP(hipMemcpyAsync(o, b.o, sizeof(float) * nv, hipMemcpyDeviceToHost, st));
  double l = 0.0;
  VH_CHECK_HIP(hipMemcpyAsync(&l, d->dscal, 8, hipMemcpyDeviceToHost, st));
  if (grad) VH_CHECK_HIP(hipMemcpyAsync(grad, d->grad, sizeof(float) * d->np, hipMemcpyDeviceToHost, st));
  VH_CHECK_HIP(hipStreamSynchronize(st));
  if (loss) *loss = l / (double)nv;
  return VH_OK;
}

// --- the dataset of density pairs, kept in HBM ----------------------------
extern "C" int vh_denoiser_dataset_reserve(vh_denoiser *d, int n, int64_t capacity) {
  VH_REQUIRE(d && n >= 8 && capacity >= 1, "vh_denoiser_dataset_reserve: bad argument");
  VH_CHECK_HIP(hipSetDevice(d->ctx->device));
  vh_free(d->data);
  d->data = nullptr;
  d->dcount = d->dcap = 0;
  VH_TRY(vh_alloc((void **)&d->data, sizeof(float) * 2 * capacity * (int64_t)n * n * n));
  d->dn = n;
  d->dcap = capacity;
  return VH_OK;
}

// a pair: the fine density of `low` (N particles) and of `high` (10 N)
extern "C" int vh_denoiser_dataset_add(vh_denoiser *d, vh_mesh *low, vh_mesh *high) {
  VH_REQUIRE(d && low && high, "vh_denoiser_dataset_add: null argument");
  VH_REQUIRE(low->n == d->dn && high->n == d->dn, "vh_denoiser_dataset_add: grids of %d, %d points, dataset %d",
             low->n, high->n, d->dn);
  VH_REQUIRE(d->dcount < d->dcap, "vh_denoiser_dataset_add: the dataset is full (%lld pairs)",
             (long long)d->dcap);
  VH_CHECK_HIP(hipSetDevice(d->ctx->device));
  const int64_t g = (int64_t)d->dn * d->dn * d->dn;
  float *p = d->data + 2 * d->dcount * g;
  // each mesh's last work on its own stream, then the copies on both
  VH_CHECK_HIP(hipMemcpyAsync(p, low->rho, sizeof(float) * g, hipMemcpyDeviceToDevice, low->ctx->stream));
  VH_CHECK_HIP(hipMemcpyAsync(p + g, high->rho, sizeof(float) * g, hipMemcpyDeviceToDevice, high->ctx->stream));
  VH_CHECK_HIP(hipStreamSynchronize(low->ctx->stream));
  VH_CHECK_HIP(hipStreamSynchronize(high->ctx->stream));
  ++d->dcount;
  return VH_OK;
}

extern "C" int vh_denoiser_dataset_size(vh_denoiser *d, int64_t *count) {
  VH_REQUIRE(d && count, "vh_denoiser_dataset_size: null argument");
  *count = d->dcount;
  return VH_OK;
}

// the Adam steps of vh_denoiser_train; sample losses into dl
template <class Pos, class Pick, class Rng>
static int train_loop(vh_denoiser *d, vh_unet_buffers &b, int iters, int C, int batch, Pos &pos,
                      Pick &pick, Rng &rng, double lr, double *dl) {
  hipStream_t st = d->ctx->stream;
  const int n = d->dn;
  const int64_t g = (int64_t)n * n * n, nv = (int64_t)C * C * C;
  const float inv = (float)(1.0 / d->rho0);
  const float b1 = 0.9f, b2 = 0.999f;
  for (int it = 0; it < iters; ++it) {
    VH_CHECK_HIP(hipMemsetAsync(d->grad, 0, sizeof(float) * d->np, st));
    for (int s = 0; s < batch; ++s) {
      const int64_t k = pick(rng);
      const int z0 = pos(rng), y0 = pos(rng), x0 = pos(rng);
      const float *plo = d->data + 2 * k * g;
      hipLaunchKernelGGL(k_crop, dim3(blocks_for(nv)), dim3(256), 0, st, plo, plo + g, n, C, z0, y0,
                         x0, inv, b.x, b.t);
      VH_LAUNCH_CHECK();
      VH_TRY(forward_backward(d, b, true, dl + (int64_t)it * batch + s));
    }
    ++d->step;
    const float c1 = 1.f - std::pow(b1, (float)d->step), c2 = 1.f - std::pow(b2, (float)d->step);
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(d->np)), dim3(256), 0, st, d->w, d->weff, d->m, d->v,
                       d->grad, d->np, (float)lr, b1, b2, c1, c2, 1e-8f, 1.f / batch);
    VH_LAUNCH_CHECK();
  }
  return VH_OK;
}

// Training: `iters` Adam steps, each on the gradient of `batch` random crops
// of side C from the pairs [first, last) of the dataset, centred within
// `radius` grid points of the grid's centre (the cloud); losses[i] (may be
// NULL) the mean of the batch's crop losses.
extern "C" int vh_denoiser_train(vh_denoiser *d, int iters, int C, int batch, int64_t first,
                                 int64_t last, int radius, double lr, uint64_t seed,
                                 double *losses) {
  VH_REQUIRE(d && iters >= 1 && batch >= 1, "vh_denoiser_train: bad argument");
  VH_REQUIRE(C >= 4 && C % 4 == 0 && C <= d->dn && C <= 160, "vh_denoiser_train: crop %d", C);
  VH_REQUIRE(first >= 0 && last <= d->dcount && first < last, "vh_denoiser_train: pairs [%lld, %lld) of %lld",
             (long long)first, (long long)last, (long long)d->dcount);
  VH_REQUIRE(radius >= 0, "vh_denoiser_train: negative radius");
  VH_REQUIRE((int64_t)iters * batch <= (1LL << 26), "vh_denoiser_train: %d x %d samples in one call",
             iters, batch);
  VH_CHECK_HIP(hipSetDevice(d->ctx->device));
  hipStream_t st = d->ctx->stream;
  vh_unet_buffers &b = d->train;
  VH_TRY(ensure_buffers(b, C, true));
  const int n = d->dn;
  const int64_t nv = (int64_t)C * C * C;
  std::mt19937_64 rng(seed);
  const int lo = n / 2 - C / 2 - radius, hi = n / 2 - C / 2 + radius;
  const int clo = lo < 0 ? 0 : lo, chi = hi > n - C ? n - C : hi;
  std::uniform_int_distribution<int> pos(clo, chi);
  std::uniform_int_distribution<int64_t> pick(first, last - 1);
  // every sample's loss stays on the device until the end
  const int64_t ns = (int64_t)iters * batch;
  double *dl = nullptr;
  VH_TRY(vh_alloc((void **)&dl, sizeof(double) * ns));
  std::vector<double> lh((size_t)ns);
  int rc = train_loop(d, b, iters, C, batch, pos, pick, rng, lr, dl);
  if (rc == VH_OK) rc = pack_weights(d);     // the inference path's operands
  if (rc == VH_OK && (hipMemcpyAsync(lh.data(), dl, sizeof(double) * ns, hipMemcpyDeviceToHost, st) != hipSuccess ||
                      hipStreamSynchronize(st) != hipSuccess)) {
    vh_set_error("vh_denoiser_train: reading the losses failed");
    rc = VH_EHIP;
  }
  vh_free(dl);
  if (rc != VH_OK) return rc;
  if (losses)
    for (int it = 0; it < iters; ++it) {
      double s = 0.0;
      for (int j = 0; j < batch; ++j) s += lh[(size_t)it * batch + j];
      losses[it] = s / batch / (double)nv;
    }
  return VH_OK;
}
// The mean squared difference, over the central cube of side C of pair k,
// between the 10 N density and (a) the N density, (b) the network's
// denoising of it, both scaled by 1/rho0.
extern "C" int vh_denoiser_dataset_error(vh_denoiser *d, int64_t k, int C, double *mse_plain,
                                         double *mse_net) {
  VH_REQUIRE(d && mse_plain && mse_net, "vh_denoiser_dataset_error: null argument");
  VH_REQUIRE(k >= 0 && k < d->dcount, "vh_denoiser_dataset_error: pair %lld of %lld", (long long)k,
             (long long)d->dcount);
  VH_REQUIRE(C >= 4 && C % 4 == 0 && C <= d->dn && C <= 160, "vh_denoiser_dataset_error: side %d", C);
  VH_CHECK_HIP(hipSetDevice(d->ctx->device));
  hipStream_t st = d->ctx->stream;
  vh_unet_buffers &b = d->train;
  VH_TRY(ensure_buffers(b, C, true));
  const int n = d->dn;
  const int64_t g = (int64_t)n * n * n, nv = (int64_t)C * C * C;
  const float *plo = d->data + 2 * k * g;
  const int c0 = n / 2 - C / 2;
  hipLaunchKernelGGL(k_crop, dim3(blocks_for(nv)), dim3(256), 0, st, plo, plo + g, n, C, c0, c0, c0,
                     (float)(1.0 / d->rho0), b.x, b.t);
  VH_LAUNCH_CHECK();
  VH_CHECK_HIP(hipMemsetAsync(b.o, 0, sizeof(float) * nv, st));
  hipLaunchKernelGGL(k_loss, dim3(512), dim3(256), 0, st, b.x, b.o, b.t, nv, d->support,
                     (float *)nullptr, d->lpart);
  hipLaunchKernelGGL(k_sum_d, dim3(1), dim3(256), 0, st, d->lpart, (int64_t)512, d->dscal + 3);
  VH_LAUNCH_CHECK();
  VH_TRY(forward_backward(d, b, false, d->dscal));
  double h[4];
  VH_CHECK_HIP(hipMemcpyAsync(h, d->dscal, 32, hipMemcpyDeviceToHost, st));
  VH_CHECK_HIP(hipStreamSynchronize(st));
  *mse_net = h[0] / nv;
  *mse_plain = h[3] / nv;
  return VH_OK;
}

// Inference on a mesh's density in place (between the deposit and the
// Poisson solve): pad to a cube of side P (multiple of 4), denoise, then
// the residual charge check against the plain density's total
// charge (FP64): within tol the denoised density renormalised to it, beyond
// it the plain density kept (a fallback, counted).  No host synchronisation.
int vh_denoiser_apply_impl(vh_denoiser *d, vh_mesh *m, double tol) {
  hipStream_t st = m->ctx->stream;
  VH_REQUIRE(d->ctx == m->ctx, "denoiser: mesh on another context");
  const int n = m->n, P = (n + 3) / 4 * 4;
  VH_TRY(vh_denoiser_prepare(d, n));
  const int64_t g = (int64_t)n * n * n;
  if (d->fp32_infer) {
    // the training kernels (FP32 activations)
    vh_unet_buffers &b = d->infer;
    hipLaunchKernelGGL(k_pad_in, dim3(blocks_for((int64_t)P * P * P)), dim3(256), 0, st, m->rho, n, P,
                       (float)(1.0 / d->rho0), b.x);
    VH_LAUNCH_CHECK();
    VH_TRY(forward(d, b));
    hipLaunchKernelGGL(k_unpad_out, dim3(blocks_for(g)), dim3(256), 0, st, b.x, b.o, n, P,
                       (float)d->rho0, d->support, d->net);
    VH_LAUNCH_CHECK();
  } else {
    VH_TRY(forward_mfma(d, st, m->rho, n, d->net));
  }
  VH_TRY(vh_total_charge_impl(m, m->rho, d->dscal + 1));
  VH_TRY(vh_total_charge_impl(m, d->net, d->dscal + 2));
  hipLaunchKernelGGL(k_select, dim3(blocks_for(g)), dim3(256), 0, st, m->rho, d->net, g, d->dscal + 1,
                     d->dscal + 2, tol, d->fallbacks);
  VH_LAUNCH_CHECK();
  return VH_OK;
}

// the inference buffers of an n^3 grid, allocated before any step records
int vh_denoiser_prepare(vh_denoiser *d, int n) {
  const int P = (n + 3) / 4 * 4;
  VH_REQUIRE(P <= 160, "denoiser: grid of %d points beyond 160", n);
  VH_CHECK_HIP(hipSetDevice(d->ctx->device));
  if (d->fp32_infer)
    VH_TRY(ensure_buffers(d->infer, P, false));
  else
    VH_TRY(ensure_mfma(d, P));
  d->prep_n = n;
  const int64_t g = (int64_t)n * n * n;
  if (d->net_cap < g) {
    vh_free(d->net);
    d->net = nullptr;
    d->net_cap = 0;
    VH_TRY(vh_alloc((void **)&d->net, sizeof(float) * g));
    d->net_cap = g;
  }
  return VH_OK;
}

vh_context *vh_denoiser_context(const vh_denoiser *d) { return d->ctx; }

extern "C" int vh_denoiser_apply(vh_denoiser *d, vh_mesh *m, double tol) {
  VH_REQUIRE(d && m, "vh_denoiser_apply: null argument");
  VH_REQUIRE(tol >= 0, "vh_denoiser_apply: negative tolerance");
  VH_CHECK_HIP(hipSetDevice(d->ctx->device));
  return vh_denoiser_apply_impl(d, m, tol);
}

extern "C" int vh_denoiser_set_inference(vh_denoiser *d, int fp32) {
  VH_REQUIRE(d != nullptr, "vh_denoiser_set_inference: null denoiser");
  VH_REQUIRE(fp32 == 0 || fp32 == 1, "vh_denoiser_set_inference: 0 (matrix cores) or 1 (FP32)");
  d->fp32_infer = fp32;
  if (d->prep_n > 0) VH_TRY(vh_denoiser_prepare(d, d->prep_n));
  return VH_OK;
}

extern "C" int vh_denoiser_set_support(vh_denoiser *d, double xmin) {
  VH_REQUIRE(d != nullptr, "vh_denoiser_set_support: null denoiser");
  VH_REQUIRE(xmin >= 0 && std::isfinite(xmin), "vh_denoiser_set_support: threshold must be >= 0");
  d->support = (float)xmin;
  return VH_OK;
}

extern "C" int vh_denoiser_fallbacks(vh_denoiser *d, int32_t *count) {
  VH_REQUIRE(d && count, "vh_denoiser_fallbacks: null argument");
  VH_CHECK_HIP(hipSetDevice(d->ctx->device));
  int c = 0;
  VH_CHECK_HIP(hipMemcpyAsync(&c, d->fallbacks, 4, hipMemcpyDeviceToHost, d->ctx->stream));
  VH_CHECK_HIP(hipStreamSynchronize(d->ctx->stream));
  *count = c;
  return VH_OK;
}
