// Stable counting sort of the particles by fine cell, on the device.
//
// The reference sorts with an atomic-cursor placement whose order inside a
// cell is not deterministic (src/kernels.jl:945-955); its host sort is a
// stable counting sort (src/sorting.jl:163-204).  Here the permutation is the
// stable one: an LSD radix sort of (cell key, index) pairs in passes of at
// most 11 bits, each pass = per-tile digit histogram (LDS) -> exclusive scan
// over (digit, tile) -> stable scatter with in-wave ranks from ballots.  The
// records are then gathered once through the permutation (each particle's
// two 24-byte records).  Keys, occupied cells and bounds follow
// src/kernels.jl:902-908 and src/accelerator.jl:397-411.
#include "vh_internal.h"

#define SORT_THREADS 256
#define SORT_ITEMS 16  // tile = 4096 elements

// ---------------------------------------------------------------------------
// generic exclusive scan (uint32, in place)
// ---------------------------------------------------------------------------
#define SCAN_THREADS 256
#define SCAN_ITEMS 4  // 1024 per block

__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t v) {
  int l = vh_lane();
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t t = __shfl_up(v, o, 64);
    if (l >= o) v += t;
  }
  return v;
}

// exclusive scan of one value per thread across the block; returns the block total
__device__ uint32_t block_excl_scan_u32(uint32_t v, uint32_t &excl, uint32_t *lds) {
  uint32_t inc = wave_incl_scan_u32(v);
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  int nw = blockDim.x >> 6;
  if (l == 63) lds[w] = inc;
  __syncthreads();
  uint32_t off = 0, tot = 0;
  for (int i = 0; i < nw; ++i) {
    uint32_t s = lds[i];
    if (i < w) off += s;
    tot += s;
  }
  __syncthreads();
  excl = off + inc - v;
  return tot;
}

// Several arrays at once: array a (blockIdx.y)
// is data + a * n with block sums at sums + a * (nb + 1).
__global__ void k_scan_reduce(const uint32_t *__restrict__ data, int64_t n, uint32_t *sums,
                              int64_t nb) {
  __shared__ uint32_t lds[16];
  data += blockIdx.y * n;
  sums += blockIdx.y * (nb + 1);
  int64_t base = (int64_t)blockIdx.x * SCAN_THREADS * SCAN_ITEMS;
  uint32_t s = 0;
  for (int it = 0; it < SCAN_ITEMS; ++it) {
    int64_t i = base + it * SCAN_THREADS + threadIdx.x;
    if (i < n) s += data[i];
  }
  uint32_t ex;
  uint32_t tot = block_excl_scan_u32(s, ex, lds);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// Up to SCAN_FEW blocks, each block adds the totals of the blocks before it
// itself (read once per block), so no pass scans the block totals; beyond,
// k_scan_mid first turns the totals into their exclusive prefix (prefixed),
// so that a block's carry is one load rather than O(nb).  The last block
// writes the array's total at sums[nb].
#define SCAN_FEW 8192
#define SCAN_MID_ITEMS 16
__global__ void __launch_bounds__(1024) k_scan_mid(uint32_t *sums, int64_t nb) {
  __shared__ uint32_t lds[16];
  sums += blockIdx.y * (nb + 1);
  uint32_t carry = 0;
  for (int64_t base = 0; base < nb; base += (int64_t)blockDim.x * SCAN_MID_ITEMS) {
    const int64_t i0 = base + (int64_t)threadIdx.x * SCAN_MID_ITEMS;
    uint32_t v[SCAN_MID_ITEMS], s = 0;
    for (int it = 0; it < SCAN_MID_ITEMS; ++it) {
      v[it] = i0 + it < nb ? sums[i0 + it] : 0;
      s += v[it];
    }
    uint32_t ex;
    const uint32_t tot = block_excl_scan_u32(s, ex, lds);
    uint32_t run = carry + ex;
    for (int it = 0; it < SCAN_MID_ITEMS; ++it) {
      if (i0 + it < nb) sums[i0 + it] = run;
      run += v[it];
    }
    carry += tot;
  }
  if (threadIdx.x == 0) sums[nb] = carry;
}

__global__ void k_scan_down(uint32_t *data, int64_t n, uint32_t *sums, int64_t nb, int prefixed) {
  __shared__ uint32_t lds[16];
  data += blockIdx.y * n;
  sums += blockIdx.y * (nb + 1);
  uint32_t carry;
  if (prefixed) {
    carry = sums[blockIdx.x];
  } else {
    uint32_t before = 0;
    for (int64_t i = threadIdx.x; i < (int64_t)blockIdx.x; i += blockDim.x) before += sums[i];
    uint32_t unused;
    carry = block_excl_scan_u32(before, unused, lds);
  }
  int64_t base = (int64_t)blockIdx.x * SCAN_THREADS * SCAN_ITEMS;
  // each thread owns SCAN_ITEMS consecutive elements
  int64_t i0 = base + (int64_t)threadIdx.x * SCAN_ITEMS;
  uint32_t v[SCAN_ITEMS];
  uint32_t s = 0;
  for (int it = 0; it < SCAN_ITEMS; ++it) {
    v[it] = (i0 + it < n) ? data[i0 + it] : 0;
    s += v[it];
  }
  uint32_t ex;
  const uint32_t tot = block_excl_scan_u32(s, ex, lds);
  uint32_t run = carry + ex;
  for (int it = 0; it < SCAN_ITEMS; ++it) {
    if (i0 + it < n) data[i0 + it] = run;
    run += v[it];
  }
  if (!prefixed && blockIdx.x == nb - 1 && threadIdx.x == 0) sums[nb] = carry + tot;
}

// exclusive scans of `count` consecutive arrays of n entries each, in place
static int scan_arrays(vh_context *c, uint32_t *data, int64_t n, int count, uint32_t *tmp,
                       int64_t tmp_len) {
  int64_t per = SCAN_THREADS * SCAN_ITEMS;
  int64_t nb = (n + per - 1) / per;
  VH_REQUIRE(count * (nb + 1) <= tmp_len, "scan: scratch too small (%lld blocks)", (long long)nb);
  if (nb == 0) return VH_OK;
  hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb, count), dim3(SCAN_THREADS), 0, c->stream,
                     data, n, tmp, nb);
  VH_LAUNCH_CHECK();
  const int prefixed = nb > SCAN_FEW;
  if (prefixed) {
    hipLaunchKernelGGL(k_scan_mid, dim3(1, count), dim3(1024), 0, c->stream, tmp, nb);
    VH_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_scan_down, dim3((unsigned)nb, count), dim3(SCAN_THREADS), 0, c->stream,
                     data, n, tmp, nb, prefixed);
  VH_LAUNCH_CHECK();
  return VH_OK;
}

int vh_scan_exclusive(vh_context *c, uint32_t *data, int64_t n, uint32_t *tmp, int64_t tmp_len,
                      uint32_t *total_out) {
  VH_TRY(scan_arrays(c, data, n, 1, tmp, tmp_len));
  if (total_out) {
    int64_t per = SCAN_THREADS * SCAN_ITEMS;
    int64_t nb = (n + per - 1) / per;
    if (nb == 0) {
      VH_TRY(vh_fill_i32(c, (int32_t *)total_out, 1, 0));
    } else {
      VH_CHECK_HIP(hipMemcpyAsync(total_out, tmp + nb, 4, hipMemcpyDeviceToDevice, c->stream));
    }
  }
  return VH_OK;
}

// ---------------------------------------------------------------------------
// keys
// ---------------------------------------------------------------------------
// mode 1: the key from the record (first pass after a move that left no
// keys); otherwise from the key array
__device__ __forceinline__ uint32_t load_key(int mode, const uint32_t *__restrict__ keys,
                                             const int32_t *__restrict__ P, int64_t npad, int64_t i,
                                             int nk) {
  if (mode == 1) {
    const int2 *p = reinterpret_cast<const int2 *>(P + 6 * i);
    const int2 a = p[0], b = p[1];
    return cell_key0(a.x, a.y, b.x, nk);
  }
  return keys[i];
}

// per-tile digit histogram, written digit-major: counts[d * nblocks + b].
// The tile's keys are loaded first (clamped addresses, so no load waits
// under a condition).  Keys in a tile mostly come in runs of one cell (the
// cloud is kept in cell order; the second pass reads the first pass's
// output, in which equal low digits keep that order), so a wave's lanes
// often share a digit: lanes sharing the digit of the first lane add with
// one LDS atomic, then one atomic per lane left (same-address atomics of a
// whole wave would serialise; more rounds cost more than they save in the
// second pass, whose rounds mix many digits).
__global__ void __launch_bounds__(SORT_THREADS)
    k_radix_hist(const uint32_t *__restrict__ keys, const int32_t *__restrict__ P, int64_t npad,
                 int first, int nk, int64_t n, int shift, int nbins, int64_t nblocks,
                 uint32_t *__restrict__ counts) {
  extern __shared__ uint32_t hist[];
  for (int d = threadIdx.x; d < nbins; d += SORT_THREADS) hist[d] = 0;
  const int64_t base = (int64_t)blockIdx.x * SORT_THREADS * SORT_ITEMS;
  const uint32_t mask = (uint32_t)nbins - 1;
  uint32_t kv[SORT_ITEMS];
#pragma unroll
  for (int it = 0; it < SORT_ITEMS; ++it) {
    const int64_t i = base + it * SORT_THREADS + threadIdx.x;
    kv[it] = load_key(first, keys, P, npad, i < n ? i : n - 1, nk);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int it = 0; it < SORT_ITEMS; ++it) {
    const bool valid = base + it * SORT_THREADS + threadIdx.x < n;
    const uint32_t d = (kv[it] >> shift) & mask;
    uint64_t left = __ballot(valid);
    if (left) {
      const int lead = __ffsll((unsigned long long)left) - 1;
      const uint32_t d0 = __shfl(d, lead, 64);
      const uint64_t same = __ballot(valid && d == d0);
      if (lane == lead) atomicAdd(&hist[d0], (uint32_t)__popcll(same));
      left &= ~same;
    }
    if (valid && (left >> lane & 1ull)) atomicAdd(&hist[d], 1u);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < nbins; d += SORT_THREADS)
    counts[(int64_t)d * nblocks + blockIdx.x] = hist[d];
}

// Stable scatter of one tile of 4096 elements.  Each wave takes a
// contiguous quarter (1024 elements, 16 rounds of 64) and ranks it alone:
// within a round, lanes of equal digit by ballots (rank among them in lane
// order); across its rounds, a running count per digit in a wave-private
// LDS row.  Then, once per tile, the four waves' counts become per-wave
// offsets (exclusive sums in wave order) and every element goes to
// goff[d] + offset of its wave + its rank in the wave: the tile's order,
// wave-major then round then lane, is the elements' order, so the scatter
// is stable.  Three barriers per tile.
__global__ void __launch_bounds__(SORT_THREADS)
    k_radix_scatter(const uint32_t *__restrict__ keys_in, const uint32_t *__restrict__ vals_in,
                    const int32_t *__restrict__ P, int64_t npad, int first, int nk, int64_t n,
                    int shift, int bits, int64_t nblocks, const uint32_t *__restrict__ offsets,
                    uint32_t *__restrict__ keys_out, uint32_t *__restrict__ vals_out) {
  extern __shared__ uint32_t sm[];
  const int nbins = 1 << bits;
  uint32_t *goff = sm;           // nbins: the tile's first slot per digit
  uint32_t *wrun = sm + nbins;   // 4 x nbins: per-wave counts, then per-wave offsets
  for (int d = threadIdx.x; d < nbins; d += SORT_THREADS) {
    goff[d] = offsets[(int64_t)d * nblocks + blockIdx.x];
    wrun[d] = wrun[nbins + d] = wrun[2 * nbins + d] = wrun[3 * nbins + d] = 0;
  }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const uint32_t mask = (uint32_t)nbins - 1;
  const int64_t wbase = (int64_t)blockIdx.x * SORT_THREADS * SORT_ITEMS + w * 64 * SORT_ITEMS;
  uint32_t kk[SORT_ITEMS], vv[SORT_ITEMS];
#pragma unroll
  for (int r = 0; r < SORT_ITEMS; ++r) {
    const int64_t i = wbase + r * 64 + lane, ic = i < n ? i : n - 1;
    kk[r] = load_key(first, keys_in, P, npad, ic, nk);
    vv[r] = first ? (uint32_t)i : vals_in[ic];   // first pass (modes 1, 2): identity
  }
  __syncthreads();
  uint32_t *mine = wrun + w * nbins;
  uint32_t pos[SORT_ITEMS];
#pragma unroll
  for (int r = 0; r < SORT_ITEMS; ++r) {
    const bool valid = wbase + r * 64 + lane < n;
    const uint32_t d = (kk[r] >> shift) & mask;
    uint64_t peers = __ballot(valid);
    for (int b = 0; b < bits; ++b) {
      const uint64_t bb = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? bb : ~bb;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & lt);
    // every lane reads its digit's count before the group's first lane
    // advances it (one wave: the LDS accesses are in program order)
    const uint32_t before = mine[d];
    pos[r] = before + rank;
    if (valid && rank == 0) mine[d] = before + (uint32_t)__popcll(peers);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < nbins; d += SORT_THREADS) {
    const uint32_t c0 = wrun[d], c1 = wrun[nbins + d], c2 = wrun[2 * nbins + d];
    wrun[d] = 0;
    wrun[nbins + d] = c0;
    wrun[2 * nbins + d] = c0 + c1;
    wrun[3 * nbins + d] = c0 + c1 + c2;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < SORT_ITEMS; ++r) {
    if (wbase + r * 64 + lane < n) {
      const uint32_t d = (kk[r] >> shift) & mask;
      const uint32_t dst = goff[d] + mine[d] + pos[r];
      keys_out[dst] = kk[r];
      vals_out[dst] = vv[r];
    }
  }
}

// ---------------------------------------------------------------------------
// record gather + deposit "nout" (columns test of src/kernels.jl:754-779)
// ---------------------------------------------------------------------------
__global__ void k_permute(const int32_t *__restrict__ src, int32_t *__restrict__ dst,
                          const uint32_t *__restrict__ perm, int64_t n, int64_t npad, float x0,
                          float h, float lo, float hi, int32_t *cnt) {
  int local = 0;
  for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < n;
       s += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = perm[s];
    const vh_pos cur = vh_load_pos(src, i), old = vh_load_pos(vh_prev_half(src, npad), i);
    vh_store_pos(dst, s, cur);
    vh_store_pos(vh_prev_half(dst, npad), s, old);
    bool inside = true;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      float p = x0 + (float)(cur.k[d] - 1) * h + cur.d[d];
      inside = inside && (lo <= p) && (p <= hi);
    }
    local += inside ? 0 : 1;
  }
  local = vh_wave_sum(local);
  if (vh_lane() == 0 && local) atomicAdd(&cnt[CNT_NOUT], local);
}

// The runs of the sorted keys, indexed by cell: run[2 c] = first slot,
// run[2 c + 1] = end.  run[2 c] stays -1 for a cell with no particle
// (k_cell_lists resets every entry it reads).  A thread takes four
// consecutive keys (one 16-byte load) and the key on each side.
__global__ void k_run_mark(const uint32_t *__restrict__ keys, int64_t n, int32_t *run) {
  const int64_t nq = (n + 3) / 4;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < nq;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s0 = 4 * q;
    uint32_t k[6];
    if (s0 + 4 <= n && ((uintptr_t)keys & 15) == 0) {
      const uint4 v = *reinterpret_cast<const uint4 *>(keys + s0);
      k[1] = v.x;
      k[2] = v.y;
      k[3] = v.z;
      k[4] = v.w;
    } else {
      for (int j = 0; j < 4; ++j) k[1 + j] = s0 + j < n ? keys[s0 + j] : 0u;
    }
    k[0] = s0 > 0 ? keys[s0 - 1] : ~0u;
    k[5] = s0 + 4 < n ? keys[s0 + 4] : ~0u;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t s = s0 + j;
      if (s >= n) break;
      const uint32_t kk = k[1 + j];
      // (keys are cells < 2^31: ~0 never equals one)
      if (s == 0 || k[j] != kk) run[2 * (int64_t)kk] = (int32_t)s;
      if (s == n - 1 || k[2 + j] != kk) run[2 * (int64_t)kk + 1] = (int32_t)(s + 1);
    }
  }
}

// per cell: occupied?  does its 10^3 stencil fit (src/kernels.jl:1047-1052)?
// Five flag arrays over ncell + 1 entries (the last zero, so that the
// exclusive scans end with the totals): occupied, fitting, non-fitting,
// particles of a non-fitting cell (the out-of-stencil list), work units of
// at most `unit` slots (VH_OPT_UNIT_SLOTS).
__global__ void k_cell_classify(const int32_t *__restrict__ run, int ncell, int nk, int n,
                                uint32_t *occidx, uint32_t *fitidx, uint32_t *nfidx, uint32_t *nfpre,
                                uint32_t *nunit, int unit) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c <= ncell; c += gridDim.x * blockDim.x) {
    const int first = c < ncell ? run[2 * (int64_t)c] : -1;
    if (first < 0) {
      occidx[c] = fitidx[c] = nfidx[c] = nfpre[c] = nunit[c] = 0u;
      continue;
    }
    const int kx = c % nk + 1, ky = (c / nk) % nk + 1, kz = c / (nk * nk) + 1;
    const int bx = 2 * kx - 5, by = 2 * ky - 5, bz = 2 * kz - 5;
    const bool ok = bx >= 1 && by >= 1 && bz >= 1 && bx + 9 <= n && by + 9 <= n && bz + 9 <= n;
    const uint32_t count = (uint32_t)(run[2 * (int64_t)c + 1] - first);
    occidx[c] = 1u;
    fitidx[c] = ok ? 1u : 0u;
    nfidx[c] = ok ? 0u : 1u;
    nfpre[c] = ok ? 0u : count;
    nunit[c] = (count + unit - 1) / unit;
  }
}

// After the scans: the occupied cells in key order with their bounds, the
// deposit map, the fitting / non-fitting lists, the work units; the totals.
__global__ void k_cell_lists(const uint32_t *__restrict__ occidx, const uint32_t *__restrict__ fitidx,
                             const uint32_t *__restrict__ nfidx, const uint32_t *__restrict__ nfpre,
                             const uint32_t *__restrict__ ustart, int32_t *run, int ncell, int64_t n,
                             int32_t *occupied, int32_t *bounds, int32_t *cellmap, int32_t *fitlist,
                             int32_t *nonfit, int32_t *nonfit_prefix, vh_unit *units,
                             int32_t *unitstart, int32_t *cnt, int unit) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c <= ncell; c += gridDim.x * blockDim.x) {
    if (c == ncell) {
      const uint32_t nocc = occidx[c];
      cnt[CNT_NOCC] = (int32_t)nocc;
      bounds[nocc] = (int32_t)n;
      cnt[CNT_NFIT] = (int32_t)fitidx[c];
      cnt[CNT_NNONFIT] = (int32_t)nfidx[c];
      cnt[CNT_OUTCOUNT] = (int32_t)nfpre[c];
      cnt[CNT_NUNITS] = (int32_t)ustart[c];
      unitstart[nocc] = (int32_t)ustart[c];
      nonfit_prefix[nfidx[c]] = (int32_t)nfpre[c];
      cnt[CNT_SORTED] = 1;
      continue;
    }
    const uint32_t g = occidx[c];
    const bool occ = occidx[c + 1] != g, ok = fitidx[c + 1] != fitidx[c];
    cellmap[c] = occ && ok ? (int32_t)g : -1;
    if (!occ) continue;
    const int32_t lo = run[2 * (int64_t)c], hi = run[2 * (int64_t)c + 1];
    run[2 * (int64_t)c] = -1;   // for the next sort
    occupied[g] = c;
    bounds[g] = lo;
    if (ok) {
      fitlist[fitidx[c]] = (int32_t)g;
    } else {
      nonfit[nfidx[c]] = (int32_t)g;
      nonfit_prefix[nfidx[c]] = (int32_t)nfpre[c];
    }
    uint32_t u = ustart[c];
    unitstart[g] = (int32_t)u;
    for (int32_t s = lo; s < hi; s += unit, ++u) {
      vh_unit un;
      un.key = c;
      un.lo = s;
      un.hi = hi - s > unit ? s + unit : hi;
      un.g = (int32_t)g;
      units[u] = un;
    }
  }
}

// the out-of-stencil list: the slots of every non-fitting cell, in order
__global__ void k_outlist(const int32_t *__restrict__ nonfit, const int32_t *__restrict__ nonfit_prefix,
                          const int32_t *__restrict__ bounds, const int32_t *__restrict__ cnt,
                          int32_t *outlist) {
  int nn = cnt[CNT_NNONFIT];
  for (int64_t c = blockIdx.x; c < nn; c += gridDim.x) {
    int g = nonfit[c];
    int lo = bounds[g], hi = bounds[g + 1];
    int off = nonfit_prefix[c];
    for (int s = lo + threadIdx.x; s < hi; s += blockDim.x) outlist[off + (s - lo)] = s;
  }
}

// out[s] = prev[perm[s]]: slot -> cloud index across two sorts
__global__ void k_compose_perm(const uint32_t *__restrict__ perm, const uint32_t *__restrict__ prev,
                               int64_t n, uint32_t *__restrict__ out) {
  for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < n;
       s += (int64_t)gridDim.x * blockDim.x)
    out[s] = prev[perm[s]];
}

// the gather of a sort whose records stayed in place
int vh_rec_settle(vh_accel *a) {
  if (!a->rec_pending) return VH_OK;
  VH_TRY(vh_fill_i32(a->ctx, a->counters + CNT_NOUT, 1, 0));
  hipLaunchKernelGGL(k_permute, dim3(4096), dim3(256), 0, a->ctx->stream, a->parts[a->cur],
                     a->parts[1 - a->cur], a->rec_perm, a->npart, a->npad, (float)a->x0,
                     (float)a->h, a->lo, a->hi, a->counters);
  VH_LAUNCH_CHECK();
  a->cur = 1 - a->cur;
  a->rec_pending = 0;
  a->nout_idx = CNT_NOUT;
  return VH_OK;
}

int vh_sort_impl(vh_accel *a) {
  vh_context *c = a->ctx;
  hipStream_t st = c->stream;
  const int64_t n = a->npart;
  // records left in place by the last sort, whose permutation the radix
  // passes overwrite
  VH_TRY(vh_rec_settle(a));
  // (the counters of the sort are written by k_cell_lists; only the gather's
  // rejection count accumulates, and is cleared before it)
  // after the field kernel's push the records stay where they are: the
  // deposits and that kernel read them through the permutation, and the push
  // moves them; it has counted the deposit's rejections among them
  const bool in_place = a->pushed_fused;
  a->pushed_fused = 0;
  // --- radix passes over (key, index) ---
  int src = 0;
  // after a push the keys of the current records are in keys[0], and after
  // vh_verlet's tile pass also the first pass's tile histograms (counts)
  const bool fresh = a->keys_fresh, hist = fresh && a->hist_fresh;
  a->keys_fresh = a->hist_fresh = 0;
  for (int p = 0; p < a->passes; ++p) {
    int shift = p * a->digit_bits;
    int bits = a->digit_bits;
    if (shift + bits > a->bits_total) bits = a->bits_total - shift;
    if (bits < 1) bits = 1;
    int nbins = 1 << bits;
    // first: 1 keys from the records, 2 keys from keys[0]; both with
    // identity values
    int first = p == 0 ? (fresh ? 2 : 1) : 0;
    if (!(p == 0 && hist)) {
      hipLaunchKernelGGL(k_radix_hist, dim3((unsigned)a->nblocks), dim3(SORT_THREADS),
                         nbins * sizeof(uint32_t), st, a->keys[src], a->parts[a->cur], a->npad,
                         first, a->nk, n, shift, nbins, a->nblocks, a->counts);
      VH_LAUNCH_CHECK();
    }
    VH_TRY(vh_scan_exclusive(c, a->counts, (int64_t)nbins * a->nblocks, a->scan_tmp,
                             a->scan_tmp_len, nullptr));
    hipLaunchKernelGGL(k_radix_scatter, dim3((unsigned)a->nblocks), dim3(SORT_THREADS),
                       6 * nbins * sizeof(uint32_t), st, a->keys[src], a->vals[src],
                       a->parts[a->cur], a->npad, first, a->nk, n, shift, bits, a->nblocks,
                       a->counts, a->keys[1 - src], a->vals[1 - src]);
    VH_LAUNCH_CHECK();
    src = 1 - src;
  }
  const uint32_t *skeys = a->keys[src];
  const uint32_t *perm = a->vals[src];
  // a cloud held on the host: slot -> the cloud's own index
  if (a->cloud_perm && !in_place && (a->track_cloud || a->cloud_valid)) {
    if (a->track_cloud) {
      VH_CHECK_HIP(hipMemcpyAsync(a->cloud_perm, perm, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
    } else {   // compose with the slots of the previous sort
      uint32_t *tmp = a->keys[1 - src];
      hipLaunchKernelGGL(k_compose_perm, dim3(1024), dim3(256), 0, st, perm, a->cloud_perm, n, tmp);
      VH_LAUNCH_CHECK();
      VH_CHECK_HIP(hipMemcpyAsync(a->cloud_perm, tmp, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
    }
    a->cloud_valid = 1;
  } else {
    a->cloud_valid = 0;
  }
  a->track_cloud = 0;
  if (in_place) {
    a->rec_pending = 1;
    a->rec_perm = perm;
    a->nout_idx = CNT_PUSH_NOUT + a->cur;
  } else {
    // --- move the records; count the deposit's rejections on the way ---
    VH_TRY(vh_fill_i32(c, a->counters + CNT_NOUT, 1, 0));
    hipLaunchKernelGGL(k_permute, dim3(4096), dim3(256), 0, st, a->parts[a->cur],
                       a->parts[1 - a->cur], perm, n, a->npad, (float)a->x0, (float)a->h, a->lo,
                       a->hi, a->counters);
    VH_LAUNCH_CHECK();
    a->cur = 1 - a->cur;
    a->nout_idx = CNT_NOUT;
  }
  // --- occupied cells, bounds, classification, lists ---
  hipLaunchKernelGGL(k_run_mark, dim3(4096), dim3(256), 0, st, skeys, n, a->cellrun);
  VH_LAUNCH_CHECK();
  const int64_t len = a->ncell + 1;
  uint32_t *occidx = a->cellscratch, *fitidx = occidx + len, *nfidx = fitidx + len,
           *nfpre = nfidx + len, *nunit = nfpre + len;
  hipLaunchKernelGGL(k_cell_classify, dim3(1024), dim3(256), 0, st, a->cellrun, (int)a->ncell, a->nk,
                     a->n, occidx, fitidx, nfidx, nfpre, nunit, a->unit);
  VH_LAUNCH_CHECK();
  VH_TRY(scan_arrays(c, occidx, len, 5, a->scan_tmp, a->scan_tmp_len));   // the five at once
  hipLaunchKernelGGL(k_cell_lists, dim3(1024), dim3(256), 0, st, occidx, fitidx, nfidx, nfpre, nunit,
                     a->cellrun, (int)a->ncell, n, a->occupied, a->bounds, a->cellmap, a->fitlist,
                     a->nonfit, a->nonfit_prefix, a->units, a->unitstart, a->counters, a->unit);
  VH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_outlist, dim3(1024), dim3(256), 0, st, a->nonfit, a->nonfit_prefix,
                     a->bounds, a->counters, a->outlist);
  VH_LAUNCH_CHECK();
  return VH_OK;
}
