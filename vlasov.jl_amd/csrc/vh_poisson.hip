// Two-level tensor-product spline Poisson solve, spline coefficients and the
// mean field, on the device.
//
// Every matrix product of the chain is the reference's "rotating" product
// (src/tensorsolver.jl:217-223): the n x m view of the source is multiplied
// by A along its first dimension and written transposed, which rotates the
// dimensions; three in a row apply A along all three axes.  Here it is one
// MFMA kernel (v_mfma_f32_16x16x4_f32, FP32 in and out) whose block owns all
// n output rows of its panels of 16 output columns, so every source column is
// read once;
// the Hadamard product by 1/(l_i+l_j+l_k) (src/tensorsolver.jl:291), the copy
// into the interior of phi (src/kernels.jl:804-813) and the `csol .+= ...` of
// the mean field (src/devicemesh.jl:182) are fused into its epilogue.
#include "vh_internal.h"

#include <cmath>
#include <cstring>
#include <mutex>

typedef float vh_f4 __attribute__((ext_vector_type(4)));

// The product C = A S (C n x m, row-major out[i' m + j]; S[k][j] = src[j n + k])
// by v_mfma_f32_16x16x4_f32 (lane l: A[l%16][k], B[k][l%16], k from l/16).
// A block holds W row tiles of 16, one wave each (W = all MT row tiles up to
// 9, else half of them), and takes panels of 16 output columns (one block
// per compute unit; the blocks of one XCD take consecutive panels):
//   * each wave keeps the A operands of its row tile for the whole K range in
//     registers, loaded once per block, with K permuted so that lane group
//     g = l/16 owns the contiguous slice k = g KQ + s, s < KQ (KQ odd,
//     >= n/4; k >= n padded with zero A).  A is stored once, at mesh
//     creation, in that operand order (vh_gemm_pack): step s of row tile r
//     is 64 consecutive floats, one per lane, so the wave's reads are
//     coalesced;
//   * a panel's source columns are contiguous in memory: the block reads
//     them with coalesced 8-byte loads into an LDS buffer, column stride
//     ldb >= 4 KQ, ldb = 2 (mod 32), so that the 32 lanes of a half-wave
//     (16 columns x 2 lane groups, odd KQ) hit 32 different banks; rows
//     k >= n are zero, so every operand address is a fixed offset;
//   * the loads of up to NP panels (a block's whole share at the sizes of
//     the step) are issued at once, together with the A operands, into
//     registers; panel i is written to its LDS buffer and multiplied as soon
//     as its own loads have landed, while the later panels' loads are still
//     in flight.  A block walking its panels one ahead waited a memory
//     latency per panel (13 us per product at n = 130); here it waits about
//     one per block.
// PF = source pairs per thread and panel (8 n <= PF blockDim): 2 when every
// row tile has its wave (MT <= 9), else 4.
// Panels in flight per block: NP panels of 16 x ldb floats stay within
// 64 KiB of LDS (ldb <= 162 for KQ <= 33, 194 for KQ = 41, 290), and the
// registers that hold them within the 168 of 9 waves per block.
constexpr int gemm_np(int kq) { return kq <= 33 ? 5 : kq <= 41 ? 4 : 3; }
template <int KQ, int PF, int EPI>
__global__ void __launch_bounds__(576)
    k_rot_gemm(const float *__restrict__ A, const float *__restrict__ src, float *__restrict__ out,
               int n, int m, int ldb, const float *__restrict__ aux, int nfull) {
  // 32-bit indices throughout: 4 n m < 2^31 (checked at launch)
  extern __shared__ float Bs[];   // NP panels of NC x ldb
  constexpr int NC = 16;          // output columns of a panel
  constexpr int NP = gemm_np(KQ);  // panels in flight per block
  const int nthr = blockDim.x;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int lc = l & 15, lk = l >> 4;
  const int mt = (n + 15) >> 4;
  const int r = blockIdx.y * (nthr >> 6) + w;
  const int npanel = (m + NC - 1) / NC;
  // XCD-aware walk (gridDim.x is a multiple of 8; workgroups go to the eight
  // XCDs in turn): the blocks of one XCD take consecutive panels, so that the
  // 64-byte pieces they write of each output row meet in that XCD's L2
  const int xcd = blockIdx.x & 7, per = gridDim.x >> 3;
  const int chunk = (npanel + 7) / 8;
  int p = xcd * chunk + (blockIdx.x >> 3);
  const int pend = (xcd + 1) * chunk < npanel ? (xcd + 1) * chunk : npanel;
  if (p >= pend) return;  // the whole block
  // A operands of row tile r (waves past the last tile read tile mt - 1 and
  // only move panels)
  float a[KQ];
  const float *ar = A + (r < mt ? r : mt - 1) * KQ * 64 + l;
#pragma unroll
  for (int s = 0; s < KQ; ++s) a[s] = ar[64 * s];
  const float2 *src2 = (const float2 *)src;
  const int tot2 = m * n / 2;
  const int per2 = NC / 2 * n;
  const int pbuf = NC * ldb;
  // Every load is issued, at an address clamped into the source; pairs
  // outside the panel or the source are zeroed when the panel is written to
  // LDS.  (Using the loaded value right away, or a load under a condition,
  // makes the compiler wait for the load at once, which serialises the
  // pipeline.)
  // (relaxed atomic loads at workgroup scope: plain loads in the
  // instruction stream, which the compiler does not move down to their
  // first use, so a round's loads all stay ahead of its first wait)
  const unsigned long long *src8 = (const unsigned long long *)src2;
  auto fetch = [&](float2 (&pv)[PF], int q) {
    const int base2 = q * NC * n / 2;
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      const int g = base2 + threadIdx.x + u * nthr;
      const unsigned long long v =
          __hip_atomic_load(src8 + (g < tot2 ? g : tot2 - 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      pv[u] = make_float2(__uint_as_float((unsigned)v), __uint_as_float((unsigned)(v >> 32)));
    }
  };
  // Every pair is written: the ones past the panel go to a spare LDS word
  // pair after the buffers (a write under a condition lets the compiler
  // move its load under the condition, where it is waited for at once).
  // The destinations are the same in every buffer.
  int dst[PF];
#pragma unroll
  for (int u = 0; u < PF; ++u) {
    const int e = threadIdx.x + u * nthr;
    const int f = 2 * e, cc = f / n, k = f - cc * n;   // n even: pairs never straddle a column
    dst[u] = e < per2 ? cc * ldb + k : NP * pbuf;
  }
  auto stash = [&](const float2 (&pv)[PF], int q, float *B) {
    const int base2 = q * NC * n / 2;
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      const int e = threadIdx.x + u * nthr;
      float *d = dst[u] < NP * pbuf ? B + dst[u] : Bs + NP * pbuf;
      *(float2 *)d = base2 + e < tot2 ? pv[u] : make_float2(0.f, 0.f);
    }
  };
  // the fused epilogue of output (orow, j), orow < n, j < m
  // the fused epilogue of output (orow, j): stored only for orow < n, j < m
  // (waves past the last row tile compute with its operands and store
  // nothing), through a range-checked buffer with no branch
  const __amdgpu_buffer_rsrc_t ro =
      vh_rsrc(out, 4 * (EPI == EPI_FILL_INTERIOR ? nfull * nfull * nfull : n * m));
  auto emit = [&](int orow, int j, float val, float pv) {
    const bool valid = orow < n && j < m;
    if (EPI == EPI_STORE) {
      vh_bstore_at(ro, valid, orow * m + j, val);
    } else if (EPI == EPI_HADAMARD) {
      vh_bstore_at(ro, valid, orow * m + j, val * pv);
    } else if (EPI == EPI_ACCUMULATE) {
      vh_bstore_at(ro, valid, orow * m + j, pv + val);
    } else {  // EPI_FILL_INTERIOR: idx = i + n (jj + n kz) of the interior, kz = orow
      const int i = j % n, jj = j / n;
      vh_bstore_at(ro, valid, (i + 1) + nfull * ((jj + 1) + nfull * (orow + 1)), val);
    }
  };
  // the products of panel q from LDS buffer B and the fused epilogue, whose
  // operands (the Hadamard factors, or the values to add to) are in pre
  auto panel = [&](int q, const float *B, const float (&pre)[4]) {
    const int j0 = q * NC;
    vh_f4 acc = (vh_f4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < KQ; ++s) {
      const int k = lk * KQ + s;  // k >= n: zero A, zero LDS padding
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[s], B[lc * ldb + k], acc, 0, 0, 0);
    }
    const int j = j0 + lc;
#pragma unroll
    for (int v = 0; v < 4; ++v) emit(r * 16 + 4 * lk + v, j, acc[v], pre[v]);
  };
  // rows k in [n, ldb) of every buffer stay zero (the panels fill k < n)
  for (int e = threadIdx.x; e < NP * NC * (ldb - n); e += nthr) {
    const int row = e / (ldb - n), k = n + e % (ldb - n);
    Bs[row * ldb + k] = 0.f;
  }
  // The loads of a round's panels and of their epilogue operands, all issued
  // before the round's first wait, after the A operands', in the order the
  // panels use them (waiting for panel i never waits for a later one's).
  // Slots past the block's last panel load the round's first panel again
  // (in the cache by then) and are not multiplied.
  float2 R[NP][PF];
  float pre[NP][4];
  auto fetch_round = [&](int p0) {
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      const int q = p0 + i * per < pend ? p0 + i * per : p0;
      fetch(R[i], q);
      if (EPI == EPI_HADAMARD || EPI == EPI_ACCUMULATE) {
        const float *srcp = EPI == EPI_HADAMARD ? aux : out;
        const int last = n * m - 1;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int idx = (r * 16 + 4 * lk + v) * m + q * NC + lc;
          pre[i][v] = srcp[idx < last ? idx : last];   // clamped: never under a condition
        }
      }
    }
  };
  auto run_round = [&](int p0) {
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      const int pi = p0 + i * per;
      if (pi >= pend) break;   // the whole block
      float *B = Bs + i * pbuf;
      // waits for panel i's loads only: the later panels' stay in flight
      stash(R[i], pi, B);
      // the epilogue operands are taken here (an empty asm that reads them):
      // left alone, the compiler moves their loads down to the conditional
      // stores, where each is waited for at once, with every later load
      if (EPI == EPI_HADAMARD || EPI == EPI_ACCUMULATE)
        asm volatile("" ::"v"(pre[i][0]), "v"(pre[i][1]), "v"(pre[i][2]), "v"(pre[i][3]));
      __syncthreads();
      panel(pi, B, pre[i]);
    }
  };
  // the first round in straight-line code: at a loop header the compiler
  // would wait for every load in flight (the A operands' too) before the
  // round's loads are issued
  fetch_round(p);
  run_round(p);
  for (p += NP * per; p < pend; p += NP * per) {
    __syncthreads();   // every wave is done with the buffers of the last round
    fetch_round(p);
    run_round(p);
  }
}

// ---------------------------------------------------------------------------
// The same rotating product in split-BF16 x3 (the precision extension):
// every operand x is split as hi = bf16(x), lo = bf16(x - hi), and
// A S ~ A_lo S_hi + A_hi S_lo + A_hi S_hi on v_mfma_f32_16x16x32_bf16 with
// FP32 accumulation (the dropped A_lo S_lo is ~2^-18 of each product).
// The layout follows k_rot_gemm: one wave per row tile with its A operands in
// registers, lane group g = l/16 owning the K slice [8 g KS, 8 (g+1) KS), one
// 16x16x32 product per step of eight K values per lane group; the operators
// are split on the host from FP64 (vh_gemm_pack_b3), and a source panel is
// split once per block, when it is written to LDS, into a hi plane and a lo
// plane of bf16 (the same 4 bytes per value as FP32), which the waves read
// with 16-byte loads (column stride ldb = 8 mod 64 bf16: the eight lanes of a
// load phase hit different 16-byte bank groups).
// ---------------------------------------------------------------------------
typedef __bf16 vh_bf8 __attribute__((ext_vector_type(8)));
typedef unsigned int vh_u4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t bf16_rne(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return u >> 16;
}

constexpr int gemm3_np(int ks) { return ks <= 5 ? 5 : ks <= 7 ? 4 : 3; }
// launch shape: waves (row tiles) per block; one block per compute unit
#define GEMM_WAVES(mt) ((mt) <= 9 ? (mt) : ((mt) + 1) / 2)

template <int KS, int PF, int EPI>
__global__ void __launch_bounds__(576)
    k_rot_gemm_b3(const vh_u4 *__restrict__ A3, const float *__restrict__ src, float *__restrict__ out,
                  int n, int m, int ldb, const float *__restrict__ aux, int nfull) {
  extern __shared__ __align__(16) unsigned short Bs3[];   // NP buffers of (hi, lo) planes
  constexpr int NC = 16;
  constexpr int NP = gemm3_np(KS);
  const int nthr = blockDim.x;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int lc = l & 15, lk = l >> 4;
  const int mt = (n + 15) >> 4;
  const int r = blockIdx.y * (nthr >> 6) + w;
  const int npanel = (m + NC - 1) / NC;
  const int xcd = blockIdx.x & 7;
  const int chunk = (npanel + 7) / 8;
  const int per = gridDim.x >> 3;
  int p = xcd * chunk + (blockIdx.x >> 3);
  const int pend = (xcd + 1) * chunk < npanel ? (xcd + 1) * chunk : npanel;
  if (p >= pend) return;
  // A operands of row tile r: step s of lane l is 32 bytes, hi then lo
  vh_u4 ah[KS], al[KS];
  const vh_u4 *ar = A3 + (size_t)(r < mt ? r : mt - 1) * KS * 128 + 2 * l;
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    ah[s] = ar[128 * s];
    al[s] = ar[128 * s + 1];
  }
  const int tot2 = m * n / 2;
  const int per2 = NC / 2 * n;
  const int plane = NC * ldb;        // bf16 per plane
  const int pbuf = 2 * plane;        // bf16 per buffer
  const unsigned long long *src8 = (const unsigned long long *)src;
  auto fetch = [&](float2 (&pv)[PF], int q) {
    const int base2 = q * NC * n / 2;
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      const int g = base2 + threadIdx.x + u * nthr;
      const unsigned long long v =
          __hip_atomic_load(src8 + (g < tot2 ? g : tot2 - 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      pv[u] = make_float2(__uint_as_float((unsigned)v), __uint_as_float((unsigned)(v >> 32)));
    }
  };
  int dst[PF];
#pragma unroll
  for (int u = 0; u < PF; ++u) {
    const int e = threadIdx.x + u * nthr;
    const int f = 2 * e, cc = f / n, k = f - cc * n;
    dst[u] = e < per2 ? cc * ldb + k : -1;
  }
  // a pair (k, k+1) of a source column split into its hi and lo planes
  auto stash = [&](const float2 (&pv)[PF], int q, unsigned short *B) {
    const int base2 = q * NC * n / 2;
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      const int e = threadIdx.x + u * nthr;
      const float2 v = base2 + e < tot2 ? pv[u] : make_float2(0.f, 0.f);
      const uint32_t h0 = bf16_rne(v.x), h1 = bf16_rne(v.y);
      const uint32_t l0 = bf16_rne(v.x - __uint_as_float(h0 << 16));
      const uint32_t l1 = bf16_rne(v.y - __uint_as_float(h1 << 16));
      uint32_t *dh = dst[u] >= 0 ? (uint32_t *)(B + dst[u]) : (uint32_t *)(Bs3 + NP * pbuf);
      uint32_t *dl = dst[u] >= 0 ? (uint32_t *)(B + plane + dst[u]) : (uint32_t *)(Bs3 + NP * pbuf + 2);
      *dh = h0 | (h1 << 16);
      *dl = l0 | (l1 << 16);
    }
  };
  const __amdgpu_buffer_rsrc_t ro =
      vh_rsrc(out, 4 * (EPI == EPI_FILL_INTERIOR ? nfull * nfull * nfull : n * m));
  auto emit = [&](int orow, int j, float val, float pv) {
    const bool valid = orow < n && j < m;
    if (EPI == EPI_STORE) {
      vh_bstore_at(ro, valid, orow * m + j, val);
    } else if (EPI == EPI_HADAMARD) {
      vh_bstore_at(ro, valid, orow * m + j, val * pv);
    } else if (EPI == EPI_ACCUMULATE) {
      vh_bstore_at(ro, valid, orow * m + j, pv + val);
    } else {
      const int i = j % n, jj = j / n;
      vh_bstore_at(ro, valid, (i + 1) + nfull * ((jj + 1) + nfull * (orow + 1)), val);
    }
  };
  auto panel = [&](int q, const unsigned short *B, const float (&pre)[4]) {
    vh_f4 acc = (vh_f4){0.f, 0.f, 0.f, 0.f};
    const unsigned short *bh = B + lc * ldb + 8 * KS * lk;
    const unsigned short *bl = bh + plane;
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const vh_bf8 h = __builtin_bit_cast(vh_bf8, *(const vh_u4 *)(bh + 8 * s));
      const vh_bf8 o = __builtin_bit_cast(vh_bf8, *(const vh_u4 *)(bl + 8 * s));
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(vh_bf8, al[s]), h, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(vh_bf8, ah[s]), o, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(vh_bf8, ah[s]), h, acc, 0, 0, 0);
    }
    const int j = q * NC + lc;
#pragma unroll
    for (int v = 0; v < 4; ++v) emit(r * 16 + 4 * lk + v, j, acc[v], pre[v]);
  };
  // rows k in [n, ldb) of both planes of every buffer stay zero
  for (int e = threadIdx.x; e < NP * 2 * NC * (ldb - n); e += nthr) {
    const int row = e / (ldb - n), k = n + e % (ldb - n);
    Bs3[row * ldb + k] = 0;
  }
  float2 R[NP][PF];
  float pre[NP][4];
  auto fetch_round = [&](int p0) {
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      const int q = p0 + i * per < pend ? p0 + i * per : p0;
      fetch(R[i], q);
      if (EPI == EPI_HADAMARD || EPI == EPI_ACCUMULATE) {
        const float *srcp = EPI == EPI_HADAMARD ? aux : out;
        const int last = n * m - 1;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int idx = (r * 16 + 4 * lk + v) * m + q * NC + lc;
          pre[i][v] = srcp[idx < last ? idx : last];
        }
      }
    }
  };
  auto run_round = [&](int p0) {
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      const int pi = p0 + i * per;
      if (pi >= pend) break;
      unsigned short *B = Bs3 + i * pbuf;
      stash(R[i], pi, B);
      if (EPI == EPI_HADAMARD || EPI == EPI_ACCUMULATE)
        asm volatile("" ::"v"(pre[i][0]), "v"(pre[i][1]), "v"(pre[i][2]), "v"(pre[i][3]));
      __syncthreads();
      panel(pi, B, pre[i]);
    }
  };
  fetch_round(p);
  run_round(p);
  for (p += NP * per; p < pend; p += NP * per) {
    __syncthreads();
    fetch_round(p);
    run_round(p);
  }
}

// steps of 32 K values: ceil(n / 32)
static int gemm_ks(int n) { return (n + 31) / 32; }

// A (n x n, column-major FP64) split into bf16 hi + lo in the operand order
// of k_rot_gemm_b3: 32 bytes per (row tile r, step s, lane l), the eight hi
// values of A[16 r + l % 16][8 KS (l / 16) + 8 s + j], j < 8, then the eight lo
int vh_gemm_pack_b3(vh_context *c, uint32_t **dst, const double *colmajor, int n) {
  const int mt = (n + 15) / 16, ks = gemm_ks(n);
  VH_REQUIRE(n >= 2 && mt <= 17 && ks <= 9, "split-BF16 rotating product: n = %d beyond 272", n);
  std::vector<uint32_t> op((size_t)mt * ks * 64 * 8);
  auto rne = [](float x) -> uint32_t {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
  };
  for (int r = 0; r < mt; ++r)
    for (int s = 0; s < ks; ++s)
      for (int l = 0; l < 64; ++l) {
        uint32_t *o = &op[(((size_t)r * ks + s) * 64 + l) * 8];
        for (int j = 0; j < 8; ++j) {
          const int row = 16 * r + (l & 15), k = 8 * ks * (l >> 4) + 8 * s + j;
          const double x = (row < n && k < n) ? colmajor[row + (size_t)n * k] : 0.0;
          const uint32_t hi = rne((float)x);
          float hf;
          const uint32_t hb = hi << 16;
          std::memcpy(&hf, &hb, 4);
          const uint32_t lo = rne((float)(x - (double)hf));
          o[j / 2] |= hi << (16 * (j & 1));
          o[4 + j / 2] |= lo << (16 * (j & 1));
        }
      }
  VH_TRY(vh_alloc((void **)dst, op.size() * sizeof(uint32_t)));
  VH_CHECK_HIP(hipMemcpy(*dst, op.data(), op.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  return VH_OK;
}

static int vh_rotate_gemm_b3(vh_context *c, const uint32_t *A3, const float *src, float *out, int n,
                             int64_t m, int epi, const float *aux, int nfull) {
  const int mt = (n + 15) / 16, ks = gemm_ks(n);
  VH_REQUIRE(mt >= 1 && mt <= 17 && ks <= 9, "split-BF16 rotating product: n = %d beyond 272", n);
  VH_REQUIRE(m >= 1 && 4 * (int64_t)(nfull > 0 ? (int64_t)nfull * nfull * nfull : n * m) < VH_OOB,
             "rotating product: n m = %lld beyond 32-bit byte offsets", (long long)((int64_t)n * m));
  VH_REQUIRE(n % 2 == 0 && ((uintptr_t)src & 7) == 0,
             "rotating product: n (%d) must be even and the source 8-byte aligned", n);
  // smallest >= 32 ks with ldb = 8 (mod 64) (bf16 per column)
  const int ldb = 32 * ks + ((8 - (32 * ks) % 64) + 64) % 64;
  const int wv = GEMM_WAVES(mt);
  constexpr int nc = 16;
  const int pf = mt <= 9 ? 2 : 4;
  VH_REQUIRE(nc / 2 * n <= pf * 64 * wv, "rotating product: panel of n = %d too large", n);
  const int gy = (mt + wv - 1) / wv;
  const int64_t npanel = (m + nc - 1) / nc;
  int64_t gx = ((int64_t)c->num_cus + gy - 1) / gy;
  if (gx > npanel) gx = npanel;
  gx = (gx + 7) / 8 * 8;
  dim3 grid((unsigned)gx, (unsigned)gy), block(64 * wv);
  const size_t lds = 2 * ((size_t)gemm3_np(ks) * 2 * nc * ldb + 4);
  VH_REQUIRE(lds <= 65536, "split-BF16 rotating product: %zu B of LDS for n = %d", lds, n);
  const vh_u4 *A = (const vh_u4 *)A3;
#define LAUNCH3(T, P, E)                                                                           \
  hipLaunchKernelGGL((k_rot_gemm_b3<T, P, E>), grid, block, lds, c->stream, A, src, out, n, (int)m, \
                     ldb, aux, nfull)
#define EPIS3(T, P)                                                                                \
  if (epi == EPI_STORE) LAUNCH3(T, P, EPI_STORE);                                                  \
  else if (epi == EPI_HADAMARD) LAUNCH3(T, P, EPI_HADAMARD);                                       \
  else if (epi == EPI_ACCUMULATE) LAUNCH3(T, P, EPI_ACCUMULATE);                                   \
  else LAUNCH3(T, P, EPI_FILL_INTERIOR);
  switch (ks) {
    case 1: EPIS3(1, 2) break;
    case 2: EPIS3(2, 2) break;
    case 3: EPIS3(3, 2) break;
    case 4: EPIS3(4, 2) break;
    case 5:
      if (pf == 2) { EPIS3(5, 2) } else { EPIS3(5, 4) }
      break;
    case 6: EPIS3(6, 4) break;
    case 7: EPIS3(7, 4) break;
    case 8: EPIS3(8, 4) break;
    case 9: EPIS3(9, 4) break;
  }
#undef EPIS3
#undef LAUNCH3
  VH_LAUNCH_CHECK();
  return VH_OK;
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


int vh_rotate_gemm(vh_context *c, const float *A, const float *src, float *out, int n, int64_t m,
                   int epi, const float *aux, int nfull) {
  const int mt = (n + 15) / 16;
  VH_REQUIRE(mt >= 1 && mt <= 17, "rotating product: n = %d beyond 272", n);
  VH_REQUIRE(m >= 1 && 4 * (int64_t)(nfull > 0 ? (int64_t)nfull * nfull * nfull : n * m) < VH_OOB,
             "rotating product: n m = %lld beyond 32-bit byte offsets", (long long)((int64_t)n * m));
  VH_REQUIRE(n % 2 == 0 && ((uintptr_t)src & 7) == 0,
             "rotating product: n (%d) must be even and the source 8-byte aligned", n);
  const int kq = gemm_kq(n);
  // smallest >= 4 kq (the K range the lanes read) with ldb = 2 (mod 32)
  const int ldb = 4 * kq + ((2 - (4 * kq) % 32) + 32) % 32;
  const int wv = GEMM_WAVES(mt);
  constexpr int nc = 16;
  // source pairs per thread and panel: 8 n <= 128 wv when every row tile has
  // its wave (wv = mt >= n/16), 8 n <= 256 wv otherwise (wv >= n/32)
  const int pf = mt <= 9 ? 2 : 4;
  VH_REQUIRE(nc / 2 * n <= pf * 64 * wv, "rotating product: panel of n = %d too large", n);
  const int gy = (mt + wv - 1) / wv;
  const int64_t npanel = (m + nc - 1) / nc;
  int64_t gx = ((int64_t)c->num_cus + gy - 1) / gy;
  if (gx > npanel) gx = npanel;
  gx = (gx + 7) / 8 * 8;
  dim3 grid((unsigned)gx, (unsigned)gy), block(64 * wv);
  const size_t lds = sizeof(float) * (gemm_np(kq) * nc * (size_t)ldb + 2);   // + spare pair
  VH_REQUIRE(lds <= 65536, "rotating product: %zu B of LDS for n = %d", lds, n);
  VH_REQUIRE(epi >= EPI_STORE && epi <= EPI_ACCUMULATE, "rotating product: epilogue %d", epi);
#define LAUNCH(T, P, E)                                                                            \
  hipLaunchKernelGGL((k_rot_gemm<T, P, E>), grid, block, lds, c->stream, A, src, out, n, (int)m,    \
                     ldb, aux, nfull)
#define EPIS(T, P)                                                                                 \
  if (epi == EPI_STORE) LAUNCH(T, P, EPI_STORE);                                                   \
  else if (epi == EPI_HADAMARD) LAUNCH(T, P, EPI_HADAMARD);                                        \
  else if (epi == EPI_ACCUMULATE) LAUNCH(T, P, EPI_ACCUMULATE);                                    \
  else LAUNCH(T, P, EPI_FILL_INTERIOR);
  // KQ <= 33: n <= 132, mt <= 9; KQ = 41: n in (132, 164]; KQ >= 49: mt >= 11
#define CASE2(T) case T: EPIS(T, 2) break;
#define CASE4(T) case T: EPIS(T, 4) break;
  switch (kq) {
    CASE2(5) CASE2(9) CASE2(15) CASE2(17) CASE2(25) CASE2(33)
    case 41:
      if (pf == 2) { EPIS(41, 2) } else { EPIS(41, 4) }
      break;
    CASE4(49) CASE4(57) CASE4(65) CASE4(69)
  }
#undef CASE2
#undef CASE4
#undef EPIS
#undef LAUNCH
  VH_LAUNCH_CHECK();
  return VH_OK;
}

// an operator of a mesh in both forms; b3 != null: the mesh runs its
// products in split-BF16 x3
struct vh_op {
  const float *f32;
  const uint32_t *b3;
};
// bit 0 of gemm_bf16x3: the solve's M^-1 and M products; bit 1: the S^-1 ones
static vh_op mop(const vh_mesh *m, const float *f32, const uint32_t *b3) {
  const int bit = f32 == m->Sinv_op ? 2 : 1;
  return vh_op{f32, (m->gemm_bf16x3 & bit) ? b3 : nullptr};
}

static int rotate(vh_context *c, const vh_op &A, const float *src, float *out, int n, int64_t m,
                  int epi, const float *aux, int nfull) {
  if (A.b3) return vh_rotate_gemm_b3(c, A.b3, src, out, n, m, epi, aux, nfull);
  return vh_rotate_gemm(c, A.f32, src, out, n, m, epi, aux, nfull);
}

// three rotations with A along each axis: src -> t1 -> t2 -> last(epi)
static int apply_all(vh_context *c, const vh_op &A, int n, const float *src, float *t1, float *t2,
                     float *dst, int epi_last, const float *aux, int nfull) {
  const int64_t m = (int64_t)n * n;
  VH_TRY(rotate(c, A, src, t1, n, m, EPI_STORE, nullptr, 0));
  VH_TRY(rotate(c, A, t1, t2, n, m, EPI_STORE, nullptr, 0));
  VH_TRY(rotate(c, A, t2, dst, n, m, epi_last, aux, nfull));
  return VH_OK;
}

// ---------------------------------------------------------------------------
// multipole moments (src/poisson.jl:55-189), FP64
// ---------------------------------------------------------------------------
// One wave per (j, k) column of the grid, four columns at a time: its lanes
// read each column's n points coalesced (i = l + 64 q, q < Q) and add each
// point's ten products to their own FP64 accumulators; the column weights
// (products of the dual moments at j and k) are the same for the whole wave.
// All loads of the four columns are issued before the sums, at clamped
// addresses (points past the column or columns past the grid count zero).
// The dual moments sit in LDS.  Sums run in a fixed order for a given n
// (lane, then wave, then block, then the fixed-order reduction of the
// block partials).
#define MOM_MAXN 288
// With the coarse deposit's accumulator (cic.acc non-null), the density is
// formed here from the 64-bit fixed-point sums as k_cic_convert forms it,
// stored, and the accumulator cleared, in the same pass as the moments.
struct vh_cic_src {
  unsigned long long *acc;
  float *rho;
  const float *wl;   // dual lengths
  double inv_scale;
  float charge;
  float *inner;      // non-null: also the interior's right-hand side, as k_rho_interior
};
template <int Q, bool CIC>
__global__ void __launch_bounds__(256)
    k_moments(const float *__restrict__ c, const double *__restrict__ dual, int n,
              double *__restrict__ part, int nparts, vh_cic_src cic) {
  __shared__ double pd[3 * MOM_MAXN];
  __shared__ float pw[MOM_MAXN];
  __shared__ double lds[10][4];
  for (int i = threadIdx.x; i < 3 * n; i += blockDim.x) pd[i] = dual[i];
  if (CIC)
    for (int i = threadIdx.x; i < n; i += blockDim.x) pw[i] = cic.wl[i];
  __syncthreads();
  const double *p0 = pd, *p1 = pd + n, *p2 = pd + 2 * n;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  double acc[10];
#pragma unroll
  for (int v = 0; v < 10; ++v) acc[v] = 0.0;
  const int njk = n * n;
  for (int t0 = 4 * (blockIdx.x * 4 + w); t0 < njk; t0 += 16 * gridDim.x) {
    float val[4][Q];
    if (CIC) {
      unsigned long long raw[4][Q];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const unsigned long long *col = cic.acc + (int64_t)min(t0 + u, njk - 1) * n;
#pragma unroll
        for (int q = 0; q < Q; ++q) raw[u][q] = col[min(l + 64 * q, n - 1)];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int t = min(t0 + u, njk - 1), j = t % n, k = t / n;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          const int i = min(l + 64 * q, n - 1);
          // k_cic_convert's expression, in its order
          const float r = (float)((double)(long long)raw[u][q] * cic.inv_scale);
          val[u][q] = r * (cic.charge / (pw[i] * pw[j] * pw[k]));
          if (t0 + u < njk && l + 64 * q < n) {
            const int64_t idx = (int64_t)t * n + i;
            cic.rho[idx] = val[u][q];
            cic.acc[idx] = 0ull;
            const int ni = n - 2;
            if (cic.inner && i >= 1 && j >= 1 && k >= 1 && i <= ni && j <= ni && k <= ni)
              cic.inner[(i - 1) + (int64_t)ni * ((j - 1) + (int64_t)ni * (k - 1))] =
                  (-4.0f * 3.14159265358979323846f) * val[u][q];
          }
        }
      }
    } else {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const float *col = c + (int64_t)min(t0 + u, njk - 1) * n;
#pragma unroll
        for (int q = 0; q < Q; ++q) val[u][q] = col[min(l + 64 * q, n - 1)];
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int t = min(t0 + u, njk - 1), j = t % n, k = t / n;
      const double live = t0 + u < njk ? 1.0 : 0.0;
      const double a = live * p0[j] * p0[k], b = live * p1[j] * p0[k];
      const double cc = live * p0[j] * p1[k], d = live * p2[j] * p0[k];
      const double e = live * p0[j] * p2[k], f = live * p1[j] * p1[k];
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int i = min(l + 64 * q, n - 1);
        const double v = l + 64 * q < n ? (double)val[u][q] : 0.0;
        const double x0 = v * p0[i], x1 = v * p1[i], x2 = v * p2[i];
        acc[0] += x0 * a;
        acc[1] += x1 * a;
        acc[2] += x0 * b;
        acc[3] += x0 * cc;
        acc[4] += x2 * a;
        acc[5] += x0 * d;
        acc[6] += x0 * e;
        acc[7] += x1 * b;
        acc[8] += x1 * cc;
        acc[9] += x0 * f;
      }
    }
  }
#pragma unroll
  for (int v = 0; v < 10; ++v) {
    const double sw = vh_wave_sum(acc[v]);
    if (l == 0) lds[v][w] = sw;
  }
  __syncthreads();
  if (threadIdx.x < 10) {
    const int v = threadIdx.x;
    part[(int64_t)v * nparts + blockIdx.x] = ((lds[v][0] + lds[v][1]) + lds[v][2]) + lds[v][3];
  }
}

// one wave per value: its lanes sum the block partials of that value in a
// fixed order (lane strides, then the wave sum)
__global__ void k_reduce_rows(const double *__restrict__ part, int nparts, int nvals,
                              double *__restrict__ out) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, nw = blockDim.x >> 6;
  for (int v = w; v < nvals; v += nw) {
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    int i = l;
    for (; i + 192 < nparts; i += 256)
#pragma unroll
      for (int u = 0; u < 4; ++u) s[u] += part[(int64_t)v * nparts + i + 64 * u];
#pragma unroll
    for (int u = 0; u < 3; ++u)
      if (i + 64 * u < nparts) s[u] += part[(int64_t)v * nparts + i + 64 * u];
    const double t = vh_wave_sum((s[0] + s[1]) + (s[2] + s[3]));
    if (l == 0) out[v] = t;
  }
}

// _multipole (src/poisson.jl:168-189): mp = [q, cx, cy, cz, qxx, qyy, qzz, qxy, qxz, qyz]
__device__ void multipole_of(const double *mom, double *mp) {
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

// the ten moments from the block partials (one wave per moment, as
// k_reduce_rows) into mom[0..9], then the multipole into mp
__global__ void __launch_bounds__(640)
    k_moments_finish(const double *__restrict__ part, int nparts, double *__restrict__ mom,
                     double *__restrict__ mp) {
  __shared__ double sm[10];
  const int v = threadIdx.x >> 6, l = threadIdx.x & 63;
  double q[4] = {0.0, 0.0, 0.0, 0.0};
  int i = l;
  for (; i + 192 < nparts; i += 256)
#pragma unroll
    for (int u = 0; u < 4; ++u) q[u] += part[(int64_t)v * nparts + i + 64 * u];
#pragma unroll
  for (int u = 0; u < 3; ++u)
    if (i + 64 * u < nparts) q[u] += part[(int64_t)v * nparts + i + 64 * u];
  const double s = vh_wave_sum((q[0] + q[1]) + (q[2] + q[3]));
  if (l == 0) {
    sm[v] = s;
    mom[v] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) multipole_of(sm, mp);
}

#define MOM_BLOCKS 256
// blocks of the moments kernel: one wave per four columns, about one pass per
// wave, at most MOM_MAXBLOCKS (partials: 10 per block)
#define MOM_MAXBLOCKS 4096
static int moments_blocks(int n) {
  return (n * n + 15) / 16 < MOM_MAXBLOCKS ? (n * n + 15) / 16 : MOM_MAXBLOCKS;
}

// the ten moments of rho into mom[0..9] and the multipole into mp[0..9],
// through 10 x moments_blocks(n) FP64 partials
static int multipole_launch(hipStream_t st, const float *rho, const double *dual, int n,
                            double *partials, double *mom, double *mp,
                            vh_cic_src cic = vh_cic_src{nullptr, nullptr, nullptr, 0.0, 0.f, nullptr},
                            bool finish = true) {
  VH_REQUIRE(n >= 2 && n <= MOM_MAXN, "multipole: n = %d outside [2, %d]", n, MOM_MAXN);
  const int nb = moments_blocks(n);
  if (cic.acc) {
    if (n <= 128)
      hipLaunchKernelGGL((k_moments<2, true>), dim3(nb), dim3(256), 0, st, rho, dual, n, partials, nb, cic);
    else if (n <= 192)
      hipLaunchKernelGGL((k_moments<3, true>), dim3(nb), dim3(256), 0, st, rho, dual, n, partials, nb, cic);
    else
      hipLaunchKernelGGL((k_moments<5, true>), dim3(nb), dim3(256), 0, st, rho, dual, n, partials, nb, cic);
  } else {
    if (n <= 128)
      hipLaunchKernelGGL((k_moments<2, false>), dim3(nb), dim3(256), 0, st, rho, dual, n, partials, nb, cic);
    else if (n <= 192)
      hipLaunchKernelGGL((k_moments<3, false>), dim3(nb), dim3(256), 0, st, rho, dual, n, partials, nb, cic);
    else
      hipLaunchKernelGGL((k_moments<5, false>), dim3(nb), dim3(256), 0, st, rho, dual, n, partials, nb, cic);
  }
  VH_LAUNCH_CHECK();
  if (finish) {
    hipLaunchKernelGGL(k_moments_finish, dim3(1), dim3(640), 0, st, partials, nb, mom, mp);
    VH_LAUNCH_CHECK();
  }
  return VH_OK;
}

// finish = false: the block partials only (k_boundary_multipole ends the sums)
static int multipole_moments(vh_mesh *m, bool finish) {
  // a coarse deposit that left its sums in the accumulator (vh_step): the
  // density is formed in the moments' pass
  vh_cic_src cic{nullptr, nullptr, nullptr, 0.0, 0.f, nullptr};
  if (m->acc_pending) {
    // with the interior's right-hand side for the solve by linearity
    // (vh_poisson_fused_prepare_impl then skips k_rho_interior)
    const bool inner = m->n <= VH_FUSED_MAXN;
    cic = vh_cic_src{(unsigned long long *)m->acc64, m->rho, m->duallen, m->acc_inv_scale,
                     m->acc_charge, inner ? m->inner : nullptr};
    m->acc_pending = 0;
    m->inner_ready = inner;
  }
  return multipole_launch(m->ctx->stream, m->rho, m->dual, m->n, m->partials, m->scal,
                          m->scal + 10, cic, finish);
}

int vh_multipole_impl(vh_mesh *m) { return multipole_moments(m, true); }

// --- the multipole without allocation --------------------------------------
// The partials, moments and multipole of vh_multipole_noalloc live in this
// file-scope array, one copy per GPU, held until the process ends.  Its host
// name is only a stand-in for the device symbol: host code uses the address
// hipGetSymbolAddress returns for the current device, never the name.
#define NOALLOC_DOUBLES (10 * MOM_MAXBLOCKS + 32)
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

#define NOALLOC_MAX_DEVICES 64
static double *scratch_ptr[NOALLOC_MAX_DEVICES];
static std::mutex scratch_mu;

// the device address of `scratch` on `device` (current), looked up once
static int scratch_of(int device, double **out) {
  std::lock_guard<std::mutex> lk(scratch_mu);
  if (!scratch_ptr[device]) {
    void *p = nullptr;
    VH_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    VH_REQUIRE(p != nullptr, "multipole without allocation: no scratch on device %d", device);
    scratch_ptr[device] = (double *)p;
  }
  *out = scratch_ptr[device];
  return VH_OK;
}

extern "C" int vh_multipole_noalloc(int device, void *stream, const float *rho, const double *dual,
                                    int n, double *out10) {
  VH_REQUIRE(rho && dual && out10, "vh_multipole_noalloc: null argument");
  VH_REQUIRE(device >= 0 && device < NOALLOC_MAX_DEVICES,
             "vh_multipole_noalloc: device %d outside [0, %d)", device, NOALLOC_MAX_DEVICES);
  VH_REQUIRE(n >= 2 && n <= MOM_MAXN,
             "vh_multipole_noalloc: n = %d outside [2, %d] (the scratch holds the partials of "
             "at most %d^3 grids)", n, MOM_MAXN, MOM_MAXN);
  VH_REQUIRE(20 + 10 * moments_blocks(n) <= NOALLOC_DOUBLES,
             "vh_multipole_noalloc: partials of n = %d beyond the scratch", n);
  VH_CHECK_HIP(hipSetDevice(device));
  double *sc = nullptr;
  VH_TRY(scratch_of(device, &sc));
  // every partial the finishing kernel reads is written by the moments
  // kernel of the same call: nothing to clear
  VH_TRY(multipole_launch((hipStream_t)stream, rho, dual, n, sc + 20, sc, out10));
  return VH_OK;
}

// total charge sum(rho m0x m0y m0z) in FP64 (src/deposition.jl:213-257)
// a wave per row (j, k) of the grid, its lanes along i (coalesced): the
// row's sum over i in a fixed lane order, times m0[j] m0[k]
__global__ void __launch_bounds__(256)
    k_charge(const float *__restrict__ rho, const double *__restrict__ m0, int n,
             double *__restrict__ part) {
  __shared__ double lds[8];
  double acc = 0.0;
  const int64_t njk = (int64_t)n * n;
  const int lane = threadIdx.x & 63;
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t t = blockIdx.x * (int64_t)(blockDim.x >> 6) + (threadIdx.x >> 6); t < njk; t += nw) {
    const int j = (int)(t % n), k = (int)(t / n);
    const float *row = rho + t * n;
    double s = 0.0;
    for (int i = lane; i < n; i += 64) s += (double)row[i] * m0[i];
    s = vh_wave_sum(s);
    if (lane == 0) acc += s * m0[j] * m0[k];
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

// The multipole potential (src/poisson.jl:198-206) on phi's faces, with
// k_moments_finish in the same launch: every block sums the moments' block
// partials (wave v value v, k_moments_finish's order, so every block holds
// the same bits) and forms the multipole, block 0 stores both, and the
// blocks then fill the faces.
__global__ void __launch_bounds__(640)
    k_boundary_multipole(float *__restrict__ phi, const double *__restrict__ part, int nparts,
                         double *__restrict__ mom, double *__restrict__ mpout,
                         const double *__restrict__ g, int n) {
  __shared__ double sm[10], mp[10];
  const int v = threadIdx.x >> 6, l = threadIdx.x & 63;
  double q[4] = {0.0, 0.0, 0.0, 0.0};
  int i0 = l;
  for (; i0 + 192 < nparts; i0 += 256)
#pragma unroll
    for (int u = 0; u < 4; ++u) q[u] += part[(int64_t)v * nparts + i0 + 64 * u];
#pragma unroll
  for (int u = 0; u < 3; ++u)
    if (i0 + 64 * u < nparts) q[u] += part[(int64_t)v * nparts + i0 + 64 * u];
  const double sv = vh_wave_sum((q[0] + q[1]) + (q[2] + q[3]));
  if (l == 0) sm[v] = sv;
  __syncthreads();
  if (threadIdx.x == 0) multipole_of(sm, mp);
  __syncthreads();
  if (blockIdx.x == 0 && threadIdx.x < 10) {
    mom[threadIdx.x] = sm[threadIdx.x];
    mpout[threadIdx.x] = mp[threadIdx.x];
  }
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
  VH_TRY(apply_all(c, mop(m, m->Minv_op, m->Minv_b3), ni, m->inner, m->work1, m->work2, m->work1, EPI_HADAMARD,
                   m->invlsum, 0));
  // inverse transform, the copy into phi's interior fused into the third product
  VH_TRY(apply_all(c, mop(m, m->M_op, m->M_b3), ni, m->work1, m->inner, m->work2, m->phi, EPI_FILL_INTERIOR,
                   nullptr, m->n));
  return VH_OK;
}

int vh_poisson_multipole_boundary_impl(vh_mesh *m) {
  vh_context *c = m->ctx;
  VH_TRY(multipole_moments(m, false));
  hipLaunchKernelGGL(k_boundary_multipole, dim3(160), dim3(640), 0, c->stream, m->phi, m->partials,
                     moments_blocks(m->n), m->scal, m->scal + 10, m->colloc64, m->n);
  VH_LAUNCH_CHECK();
  return VH_OK;
}

int vh_poisson_single_impl(vh_mesh *m) {
  VH_TRY(vh_poisson_multipole_boundary_impl(m));
  VH_TRY(solve_interior(m));
  return VH_OK;
}

int vh_spline_coefficients_impl(vh_mesh *m, const float *src, float *dst) {
  // src -> dst -> work -> dst, as apply_all_rotating! alternates (src/tensorsolver.jl:243-254)
  VH_TRY(apply_all(m->ctx, mop(m, m->Sinv_op, m->Sinv_b3), m->n, src, dst, m->work, dst, EPI_STORE, nullptr, 0));
  return VH_OK;
}

int vh_poisson_coarse_level_impl(vh_mesh *coarse) {
  VH_TRY(vh_poisson_single_impl(coarse));
  // coarse coefficients: the same S^-1 phi that spline_coefficients! of the
  // coarse level computes afterwards, kept in its csol
  VH_TRY(vh_spline_coefficients_impl(coarse, coarse->phi, coarse->csol));
  return VH_OK;
}

int vh_poisson_fine_boundary_impl(vh_mesh *fine, vh_mesh *coarse) {
  vh_context *c = fine->ctx;
  hipLaunchKernelGGL(k_boundary_from_coarse, dim3(512), dim3(256), 0, c->stream, fine->phi,
                     fine->colloc64, fine->n, coarse->csol, coarse->knots64, coarse->nk, coarse->n);
  VH_LAUNCH_CHECK();
  return VH_OK;
}

int vh_poisson_coarse_part_impl(vh_mesh *fine, vh_mesh *coarse) {
  VH_TRY(vh_poisson_coarse_level_impl(coarse));
  VH_TRY(vh_poisson_fine_boundary_impl(fine, coarse));
  return VH_OK;
}

int vh_poisson_fine_part_impl(vh_mesh *fine) { return solve_interior(fine); }

int vh_poisson_nested_impl(vh_mesh *fine, vh_mesh *coarse) {
  VH_TRY(vh_poisson_coarse_part_impl(fine, coarse));
  VH_TRY(vh_poisson_fine_part_impl(fine));
  return VH_OK;
}

// ---------------------------------------------------------------------------
// The solve of a level and its spline coefficients by linearity.
// With E the embedding of the interior (ni = n - 2 points a side) and B the
// boundary values on phi's faces, solve_interior! then spline_coefficients!
// compute
//   csol = S^-1 (E (M Lambda^-1 M^-1 rhs) + B),  rhs = c rho_int - lift(B)
// with the three-axis products written as operators.  Let M~ be M on the
// interior indices and the identity on 0 and n - 1: then
//   phi = M~ Z~,  csol = (S^-1 M~) Z~ = P~ Z~,
// where Z~ = Lambda^-1 (M^-1 rhs) inside and, on the shell, the boundary
// values transformed by M^-1 along their interior axes (faces by two, edges
// by one, corners as they are), since M~ maps exactly those back to B.  The
// lift of the faces into rhs is a sum of six terms d (x) face, so
//   M^-1 rhs = M^-1 (c rho_int) - sum_f u_f (x) H_f,
// u_f = M^-1 d_f (two vectors, built at mesh creation) and H_f the 2-D
// transform of face f's interior by M^-1 -- the same H_f that fill the
// shell.  Hence: prepare (density only) = M^-1 (c rho_int); finish (once the
// faces are known) = the six H_f (two rotating products of the face stack),
// Z~ in one pass, and three products by P~: three products after the
// boundary instead of nine, and phi's interior is never formed.  P~ and u are
// formed in FP64 and rounded once.
// ---------------------------------------------------------------------------
// c rho at the interior points, ni^3 (i fastest); block (j, k), thread i
__global__ void k_rho_interior(float *__restrict__ inner, const float *__restrict__ rho, int n) {
  const int ni = n - 2, i = threadIdx.x, j = blockIdx.x, k = blockIdx.y;
  if (i >= ni) return;
  const int64_t nn = n;
  inner[i + (int64_t)ni * (j + (int64_t)ni * k)] =
      (-4.0f * 3.14159265358979323846f) * rho[(i + 1) + nn * ((j + 1) + nn * (k + 1))];
}

// The faces' and edges' transforms by M^-1, read straight from phi's shell,
// in two passes of one block per row.  Face f's interior values F(p, q)
// (x faces (i = 0, n - 1): p = j, q = k; y faces p = i, q = k; z faces
// p = i, q = j; interior indices, 0-based) go to H_f = M^-1 F M^-T:
//   pass 1, block (p, f): T_f(p, q') = sum_q F(p, q) M^-1[q', q];
//   pass 2, block (p', f): H_f(p', q') = sum_p M^-1[p', p] T_f(p, q'), at
//   f + 6 p' + 6 ni q'; blocks past the faces' take the twelve edges
//   (edge e: axis = e / 4 the interior axis, 0 i, 1 j, 2 k, the two other
//   indices 0 or n - 1 by bits 0 and 1 of e % 4, the lower axis first):
//   E(c, e) = sum_q M^-1[c, q] line(q), at c 12 + e.
// The row a block contracts sits in LDS; thread q' (or c) reads
// minvt = M^-1 transposed ([q ni + c]) at consecutive words.
__device__ __forceinline__ int64_t face_point(int f, int p, int q, int n) {
  const int64_t nn = n;
  const int side = (f & 1) ? n - 1 : 0;
  if (f < 2) return side + nn * ((p + 1) + nn * (q + 1));
  if (f < 4) return (p + 1) + nn * (side + nn * (q + 1));
  return (p + 1) + nn * ((q + 1) + nn * side);
}

// sum_q row[q] m[q ni + c], eight independent partial sums
__device__ __forceinline__ float row_dot(const float *row, const float *__restrict__ m, int ni, int c) {
  float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  int q = 0;
  for (; q + 8 <= ni; q += 8) {
#pragma unroll
    for (int v = 0; v < 8; ++v) acc[v] = __builtin_fmaf(row[q + v], m[(int64_t)(q + v) * ni + c], acc[v]);
  }
  for (; q < ni; ++q) acc[0] = __builtin_fmaf(row[q], m[(int64_t)q * ni + c], acc[0]);
  return ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
}

__global__ void __launch_bounds__(256)
    k_face_pass1(float *__restrict__ T, const float *__restrict__ phi, const float *__restrict__ minvt, int n) {
  __shared__ float row[256];
  const int ni = n - 2, p = blockIdx.x, f = blockIdx.y, t = threadIdx.x;
  for (int q = t; q < ni; q += blockDim.x) row[q] = phi[face_point(f, p, q, n)];
  __syncthreads();
  for (int c = t; c < ni; c += blockDim.x)
    T[(int64_t)ni * ((int64_t)ni * f + p) + c] = row_dot(row, minvt, ni, c);
}

__global__ void __launch_bounds__(256)
    k_face_pass2(float *__restrict__ H, float *__restrict__ E, const float *__restrict__ T,
                 const float *__restrict__ phi, const float *__restrict__ minvt, int n) {
  __shared__ float row[256];
  const int ni = n - 2, t = threadIdx.x;
  const int64_t nn = n;
  if (blockIdx.y == 6) {
    // an edge
    const int e = blockIdx.x;
    if (e >= 12) return;
    const int axis = e >> 2, s0 = (e & 1) ? n - 1 : 0, s1 = (e & 2) ? n - 1 : 0;
    int64_t base, step;
    if (axis == 0) { base = nn * (s0 + nn * s1); step = 1; }          // j = s0, k = s1
    else if (axis == 1) { base = s0 + nn * nn * s1; step = nn; }      // i = s0, k = s1
    else { base = s0 + nn * s1; step = nn * nn; }                     // i = s0, j = s1
    for (int q = t; q < ni; q += blockDim.x) row[q] = phi[base + step * (q + 1)];
    __syncthreads();
    for (int c = t; c < ni; c += blockDim.x) E[12 * c + e] = row_dot(row, minvt, ni, c);
    return;
  }
  // H_f(p', q') = sum_p M^-1[p', p] T_f(p, q'): M^-1's row p' in LDS, T read
  // at consecutive q'
  const int pp = blockIdx.x, f = blockIdx.y;
  for (int q = t; q < ni; q += blockDim.x) row[q] = minvt[(int64_t)q * ni + pp];
  __syncthreads();
  const float *Tf = T + (int64_t)ni * ni * f;
  for (int c = t; c < ni; c += blockDim.x) H[f + 6 * pp + 6 * (int64_t)ni * c] = row_dot(row, Tf, ni, c);
}

// Z~ on the n^3 grid (i fastest): Y = M^-1 (c rho_int) in `inner`, H the
// face transforms (H_f(p, q) at f + 6 p + 6 ni q), E the edges' (E(c, e) at
// c 12 + e); block (j, k), thread i
__global__ void k_ztilde(float *__restrict__ Z, const float *__restrict__ Y,
                         const float *__restrict__ invl, const float *__restrict__ H,
                         const float *__restrict__ u, const float *__restrict__ E,
                         const float *__restrict__ phi, int n) {
  const int ni = n - 2, i = threadIdx.x, j = blockIdx.x, k = blockIdx.y;
  if (i >= n) return;
  const int64_t nn = n, hq = 6 * (int64_t)ni;
  const int64_t t = i + nn * (j + nn * k);
  const float *ul = u, *ur = u + ni;
  const bool bi = i == 0 || i == n - 1, bj = j == 0 || j == n - 1, bk = k == 0 || k == n - 1;
  const int nb = (int)bi + (int)bj + (int)bk;
  const int a = i - 1, b = j - 1, c = k - 1;
  float z;
  if (nb == 0) {
    const int64_t g = a + (int64_t)ni * (b + (int64_t)ni * c);
    // the six lifted faces, as M^-1 carries them: u_f (x) H_f
    const float2 hx = *(const float2 *)(H + 6 * b + hq * c);
    const float2 hy = *(const float2 *)(H + 2 + 6 * a + hq * c);
    const float2 hz = *(const float2 *)(H + 4 + 6 * a + hq * b);
    float lift = ul[a] * hx.x;
    lift = __builtin_fmaf(ur[a], hx.y, lift);
    lift = __builtin_fmaf(ul[b], hy.x, lift);
    lift = __builtin_fmaf(ur[b], hy.y, lift);
    lift = __builtin_fmaf(ul[c], hz.x, lift);
    lift = __builtin_fmaf(ur[c], hz.y, lift);
    z = invl[g] * (Y[g] - lift);
  } else if (nb == 1) {
    if (bi) z = H[(i ? 1 : 0) + 6 * b + hq * c];
    else if (bj) z = H[(j ? 3 : 2) + 6 * a + hq * c];
    else z = H[(k ? 5 : 4) + 6 * a + hq * b];
  } else if (nb == 2) {
    // an edge: M^-1 along its interior axis
    const int axis = !bi ? 0 : !bj ? 1 : 2;
    const int lo = axis == 0 ? j : i, hi = axis == 2 ? j : k;
    const int e = 4 * axis + (lo ? 1 : 0) + (hi ? 2 : 0);
    z = E[12 * (axis == 0 ? a : axis == 1 ? b : c) + e];
  } else {
    z = phi[t];
  }
  Z[t] = z;
}

int vh_poisson_fused_prepare_impl(vh_mesh *m) {
  vh_context *c = m->ctx;
  VH_REQUIRE(m->n <= 256, "fused solve: n = %d beyond 256", m->n);
  if (!m->inner_ready) {
    hipLaunchKernelGGL(k_rho_interior, dim3(m->ni, m->ni), dim3((m->ni + 63) / 64 * 64), 0, c->stream,
                       m->inner, m->rho, m->n);
    VH_LAUNCH_CHECK();
  }
  m->inner_ready = 0;
  // inner -> work1 -> work2 -> inner
  VH_TRY(apply_all(c, vh_op{m->Minv_op, nullptr}, m->ni, m->inner, m->work1, m->work2, m->inner, EPI_STORE,
                   nullptr, 0));
  return VH_OK;
}

int vh_poisson_fused_finish_impl(vh_mesh *m, bool accumulate) {
  vh_context *c = m->ctx;
  const int n = m->n, ni = m->ni;
  VH_REQUIRE(n <= 256 && ni >= 12, "fused solve: n = %d outside [14, 256]", n);
  float *H = m->face, *E = H + 6 * (int64_t)ni * ni, *T = E + 12 * (int64_t)ni;
  hipLaunchKernelGGL(k_face_pass1, dim3(ni, 6), dim3(128), 0, c->stream, T, m->phi, m->Minv_t, n);
  VH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_face_pass2, dim3(ni, 7), dim3(128), 0, c->stream, H, E, T, m->phi, m->Minv_t, n);
  VH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_ztilde, dim3(n, n), dim3((n + 63) / 64 * 64), 0, c->stream, m->extra, m->inner,
                     m->invlsum, H, m->ulr, E, m->phi, n);
  VH_LAUNCH_CHECK();
  // extra -> work -> extra -> csol
  VH_TRY(apply_all(c, vh_op{m->Pt_op, nullptr}, n, m->extra, m->work, m->extra, m->csol,
                   accumulate ? EPI_ACCUMULATE : EPI_STORE, nullptr, 0));
  return VH_OK;
}

// ---------------------------------------------------------------------------
// mean field (src/meanfield.jl:116-147): V_xc(rho) + V_jel(r) at the points
// ---------------------------------------------------------------------------
__global__ void k_effective(float *__restrict__ extra, const float *__restrict__ rho,
                            const float *__restrict__ g, int n, float nions, float radius,
                            const float *__restrict__ add, vh_sig sig) {
  vh_sig_store(sig);
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

int vh_effective_potential_impl(vh_mesh *m, double nions, double radius, vh_sig sig) {
  vh_context *c = m->ctx;
  hipLaunchKernelGGL(k_effective, dim3(2048), dim3(256), 0, c->stream, m->extra, m->rho, m->colloc,
                     m->n, (float)nions, (float)radius, (const float *)nullptr, sig);
  VH_LAUNCH_CHECK();
  // extra -> work -> extra -> csol (+=)
  VH_TRY(apply_all(c, mop(m, m->Sinv_op, m->Sinv_b3), m->n, m->extra, m->work, m->extra, m->csol, EPI_ACCUMULATE,
                   nullptr, 0));
  return VH_OK;
}

// spline_coefficients! then effective_potential!, in the other order:
// csol = S^-1 (V_xc(rho) + V_jel), which needs only the density, then
// csol += S^-1 phi once the potential is solved (the same sum as the two
// calls, whose terms are added after their transforms too)
int vh_mean_field_coefficients_impl(vh_mesh *m, double nions, double radius) {
  vh_context *c = m->ctx;
  hipLaunchKernelGGL(k_effective, dim3(2048), dim3(256), 0, c->stream, m->extra, m->rho, m->colloc,
                     m->n, (float)nions, (float)radius, (const float *)nullptr, vh_sig{nullptr, 0});
  VH_LAUNCH_CHECK();
  // extra -> work -> extra -> csol
  VH_TRY(apply_all(c, mop(m, m->Sinv_op, m->Sinv_b3), m->n, m->extra, m->work, m->extra, m->csol, EPI_STORE, nullptr, 0));
  return VH_OK;
}

int vh_hartree_coefficients_add_impl(vh_mesh *m) {
  // phi -> extra -> work -> csol (+=)
  VH_TRY(apply_all(m->ctx, mop(m, m->Sinv_op, m->Sinv_b3), m->n, m->phi, m->extra, m->work, m->csol, EPI_ACCUMULATE,
                   nullptr, 0));
  return VH_OK;
}

int vh_potential_coefficients_impl(vh_mesh *m, double nions, double radius) {
  VH_TRY(vh_mean_field_coefficients_impl(m, nions, radius));
  VH_TRY(vh_hartree_coefficients_add_impl(m));
  return VH_OK;
}
