// Internal definitions shared by the translation units of libvlasov_hip.so.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/vlasov_hip.h"

// ---------------------------------------------------------------------------
// error plumbing: a thread-local message, never an exception across the ABI
// ---------------------------------------------------------------------------
void vh_set_error(const char *fmt, ...);

#define VH_CHECK_HIP(expr)                                                       \
  do {                                                                           \
    hipError_t e_ = (expr);                                                      \
    if (e_ != hipSuccess) {                                                      \
      vh_set_error("%s:%d: %s: %s", __FILE__, __LINE__, #expr, hipGetErrorString(e_)); \
      return VH_EHIP;                                                            \
    }                                                                            \
  } while (0)

#define VH_REQUIRE(cond, ...)                                                    \
  do {                                                                           \
    if (!(cond)) {                                                               \
      vh_set_error(__VA_ARGS__);                                                 \
      return VH_EINVAL;                                                          \
    }                                                                            \
  } while (0)

#define VH_LAUNCH_CHECK()                                                        \
  do {                                                                           \
    hipError_t e_ = hipGetLastError();                                           \
    if (e_ != hipSuccess) {                                                      \
      vh_set_error("%s:%d: launch: %s", __FILE__, __LINE__, hipGetErrorString(e_)); \
      return VH_EHIP;                                                            \
    }                                                                            \
  } while (0)

#define VH_TRY(expr)                                                             \
  do {                                                                           \
    int r_ = (expr);                                                             \
    if (r_ != VH_OK) return r_;                                                  \
  } while (0)

// ---------------------------------------------------------------------------
// objects behind the opaque handles
// ---------------------------------------------------------------------------
struct vh_prof_rec {
  int stage;
  hipEvent_t start, stop;
};

// A flag store folded into the start of a kernel (thread 0 of block 0, no
// flag when f is null): the kernel has started, so the work queued on its
// stream before it is complete (relaxed, as k_flag_signal's).
struct vh_sig {
  uint32_t *f;
  uint32_t v;
};
__device__ __forceinline__ void vh_sig_store(vh_sig s) {
  if (s.f && blockIdx.x == 0 && threadIdx.x == 0)
    __hip_atomic_store(s.f, s.v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct vh_context {
  int device;
  hipStream_t stream;
  // a second stream for work of a step that is independent of the main
  // chain (the coarse mean field beside the fine solve), joined by events
  hipStream_t side;
  hipEvent_t fork_ev, join_ev, ev_a;
  // Order between the two streams inside vh_step: a small kernel on one
  // stream stores the step's sequence number into a flag word, a one-wave
  // kernel on the other spins (bounded) until it reads it.  A stream that
  // waits at an event while the other runs a chain of short kernels is
  // resumed by the command processor tens to hundreds of microseconds late;
  // a kernel already spinning on the device ends within a few.
  uint32_t *flags;      // VH_FLAG_WORDS words: flag f at f * 16 (own 64-byte line)
  uint32_t flag_seq;    // sequence number of the last step's signals
  bool flags_ok;        // the two streams were seen to progress independently
  bool live;            // counted among the process's live contexts
  int num_cus;
  // timing of regions with events on `stream`
  int prof_on;                          // mask of the stages timed
  std::vector<vh_prof_rec> prof_open;   // recorded, not yet read
  std::vector<hipEvent_t> prof_pool;    // free events
};

// region timing: begin returns a token (-1 when profiling is off)
int vh_prof_begin(vh_context *c, int stage);
void vh_prof_end(vh_context *c, int token);

struct vh_mesh {
  vh_context *ctx;
  int n, nk, ni;
  // tables, float unless stated (the reference narrows them to E = Float32)
  float *colloc, *knots;
  double *colloc64, *knots64;
  float *lap;        // full D, row-major [r*n + c]
  float *Minv_op;    // ni x ni, A operand of the rotating product (vh_gemm_pack order)
  float *M_op;
  float *Sinv_op;    // n x n, same order
  // the same operators split into bf16 hi + lo (vh_gemm_pack_b3 order), and
  // whether the mesh's products run in split-BF16 x3 (VH_MESH_OPT_BF16X3)
  uint32_t *Minv_b3, *M_b3, *Sinv_b3;
  int gemm_bf16x3;
  // the solve and the coefficients in one (vh_step): P~ = S^-1 M~ (n x n,
  // vh_gemm_pack order), M~ = M on the interior indices and the identity on
  // 0 and n - 1; u = M^-1 D[1:n-1, 0] then M^-1 D[1:n-1, n-1] (2 ni); M^-1
  // transposed ([q ni + c] = M^-1[c, q]); the six faces' 2-D transforms by
  // M^-1 (6 ni^2), the twelve edges' 1-D transforms (12 ni), then the faces'
  // first pass (6 ni^2)
  float *Pt_op, *ulr, *Minv_t, *face;
  float *invlsum;    // ni^3, i fastest
  double *dual;      // 3 x n (order k at k*n), FP64
  float *dual32;     // same, narrowed
  float *duallen;    // n
  float loc_x0, loc_iw;
  int32_t *loc_cells;  // 1-based
  int nbin;
  // grids, n^3 floats, [i + n j + n^2 k]
  float *rho, *phi, *csol, *extra, *work;
  // interior work, ni^3
  float *inner, *work1, *work2;
  // reductions
  double *partials;  // 10 x (n*n) moments / charge partials
  double *scal;      // [0..9] moments, [10..19] multipole, [20] total charge
  int64_t *acc64;    // n^3 fixed-point accumulator of the CIC deposit
  // a coarse deposit whose density is still in acc64 (vh_step): the
  // multipole's moments pass forms rho and clears the accumulator
  int acc_pending;
  double acc_inv_scale;
  float acc_charge;
  int inner_ready;   // the moments pass also wrote the interior's right-hand side
  uint64_t serial;   // distinct for every mesh the process creates
};

// Window of a fine cell on the coarse collocation points along one axis (the
// coarse deposit): the 1-based coarse interval a (left point) holding the
// cell's lower corner, 0 when it is off the coarse grid; the four
// collocation points a - 1 .. a + 2 (0-based; past the grid, the last point
// again); the reciprocal lengths of the three intervals between them (0 for
// an empty interval).  One entry per fine cell index of an axis.
struct alignas(16) vh_cicwin {
  int32_t a;
  float g[4];
  float r[3];
};
static_assert(sizeof(vh_cicwin) == 32, "vh_cicwin: two 16-byte words");

// Particle buffer: two arrays of 24-byte position records, npad each: the
// current positions (kx, ky, kz, dx, dy, dz; int32 and float words) of slot s
// at words [6 s, 6 s + 6), the previous positions at 6 npad + 6 s.  The
// per-cell kernels read the current array alone, coalesced; the sort's gather
// moves a particle that changed cell as two 24-byte pieces (two or three
// memory sectors) rather than twelve scattered words.
#define VH_REC_WORDS 12   // int32 words per particle in a buffer (both arrays)

// K slot k = 8 g + j (g = lane group, j = element) of the fp16 split
// product in the matrix-core field kernel's first stage, where a
// v_mfma_f32_16x16x32_f16 sums, for the ten x columns ii of a row of csol,
// c_hi w_hi + c_hi w_lo + c_lo w_hi (hi = fp16(v), lo = fp16(v - hi) of the
// scaled FP32 values): slot k pairs part cpart of c[ii] with part wpart of
// w[ii] (0 hi, 1 lo, -1 an empty slot).  Lane group g needs c at ii = 2g,
// 2g + 1 and 8 + g / 2 only.
__host__ __device__ inline void vh_s16_slot(int k, int &ii, int &cpart, int &wpart) {
  const int g = k >> 3, j = k & 7;
  if (j < 6) {
    ii = 2 * g + j / 3;
    cpart = j % 3 == 2;
    wpart = j % 3 == 1;
  } else if (j == 6) {
    ii = 8 + (g >> 1);
    cpart = g & 1;
    wpart = 0;
  } else {
    ii = 8 + (g >> 1);
    cpart = (g & 1) ? -1 : 0;
    wpart = (g & 1) ? -1 : 1;
  }
}

struct vh_pos {
  int k[3];
  float d[3];
};

__device__ __forceinline__ const int32_t *vh_prev_half(const int32_t *P, int64_t npad) {
  return P + 6 * npad;
}
__device__ __forceinline__ int32_t *vh_prev_half(int32_t *P, int64_t npad) { return P + 6 * npad; }

// one record: three 8-byte loads (records are 8-byte aligned)
__device__ __forceinline__ vh_pos vh_load_pos(const int32_t *__restrict__ half, int64_t s) {
  const int2 *p = reinterpret_cast<const int2 *>(half + 6 * s);
  const int2 a = p[0], b = p[1], c = p[2];
  vh_pos r;
  r.k[0] = a.x;
  r.k[1] = a.y;
  r.k[2] = b.x;
  r.d[0] = __int_as_float(b.y);
  r.d[1] = __int_as_float(c.x);
  r.d[2] = __int_as_float(c.y);
  return r;
}

// the offsets alone (words 3..5)
__device__ __forceinline__ void vh_load_delta(const int32_t *__restrict__ half, int64_t s, float &dx,
                                              float &dy, float &dz) {
  const int2 *p = reinterpret_cast<const int2 *>(half + 6 * s);
  const int2 b = p[1], c = p[2];
  dx = __int_as_float(b.y);
  dy = __int_as_float(c.x);
  dz = __int_as_float(c.y);
}

// Stores through a buffer descriptor over `bytes` bytes at p (bytes <
// 2^31): the hardware's range check drops a store whose offset lies past the
// end.  Stores of lanes that must not write (past a chunk, outside a grid)
// are then issued by every lane with no branch, so the compiler's count of
// the memory operations in flight stays exact: with a store under a lane or
// wave condition it waits, at the next use of a later load, for every store
// issued before that load as well.
#define VH_OOB 0x7ffffff0
typedef int vh_i2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ __amdgpu_buffer_rsrc_t vh_rsrc(const void *p, int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, bytes, 0x00020000);
}
__device__ __forceinline__ void vh_bstore_f(__amdgpu_buffer_rsrc_t r, int lane, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__float_as_int(v), r, 4 * lane, 0, 0);
}
// the store of float element `idx`, or none when !valid
__device__ __forceinline__ void vh_bstore_at(__amdgpu_buffer_rsrc_t r, bool valid, int idx, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__float_as_int(v), r, valid ? 4 * idx : VH_OOB, 0, 0);
}

__device__ __forceinline__ void vh_store_pos(int32_t *__restrict__ half, int64_t s, const vh_pos &r) {
  int2 *p = reinterpret_cast<int2 *>(half + 6 * s);
  p[0] = make_int2(r.k[0], r.k[1]);
  p[1] = make_int2(r.k[2], __float_as_int(r.d[0]));
  p[2] = make_int2(__float_as_int(r.d[1]), __float_as_int(r.d[2]));
}

// device-side counters of the accelerator
enum {
  CNT_NOCC = 0,   // occupied cells
  CNT_NOUT,       // columns-kernel rejections (deposit nout)
  CNT_OUTCOUNT,   // particles whose 10^3 stencil leaves the fine grid
  CNT_NFIT,       // occupied cells whose stencil fits
  CNT_NNONFIT,    // occupied cells that do not fit
  CNT_SORTED,     // 1 once a sort has run
  CNT_NUNITS,     // work units (see VH_UNIT)
  CNT_COUNT,
  // not cleared by the sort: the rejections among the records a push wrote
  // into parts[b] (slot CNT_PUSH_NOUT + b), counted by the push for the
  // deposit of the next step
  CNT_PUSH_NOUT = CNT_COUNT,
  CNT_READ = CNT_PUSH_NOUT + 2,  // the counters a host read copies
  CNT_COLUMNS_OUT = 32,          // vh_columns' rejection count
  CNT_REDUCE_TICKET = 40         // k_reduce_field's ticket (reset by its last block)
};

// Work units of the per-cell kernels: an occupied cell's particles in runs of
// at most VH_UNIT slots.  A cell of the cluster's core holds thousands of
// particles; as units they spread over many waves instead of one.
#define VH_UNIT 256        // default slots per work unit (VH_OPT_UNIT_SLOTS)
#define VH_UNIT_MIN 128    // the smallest allowed: the unit buffers are sized for it
struct vh_unit {
  int32_t key;   // 0-based cell key
  int32_t lo;    // first slot
  int32_t hi;    // one past the last slot
  int32_t g;     // occupied index of the cell
};

#define VH_GRAPH_KEY 15

// the host side of an accelerator that a step changes (vh_step's replays
// restore what the recorded step left)
struct vh_step_state {
  int cur, keys_fresh, hist_fresh, rec_pending, pushed_fused, nout_idx;
  const uint32_t *rec_perm;
};

struct vh_accel {
  vh_context *ctx;
  int64_t npart, npad;
  int nk, n, ncell;
  double x0, h;
  float spacing;
  int nbdt, ncol;
  float *ovl, *grad, *nodes;   // [col*10 + a], [col*8 + a]
  float *og;                   // [col*32 + a]: overlap a < 10, gradient 10 + a (128-byte columns)
  uint16_t *xsplit;            // [col*64 + 32 t + k]: fp16 split of the overlap (t = 0) and
                               // gradient (t = 1) column in the K order of vh_s16_slot
  float xs_inv[2];             // 1 / the power-of-two scales of those splits
  float *zero;                 // 64 zero words (the target of padding loads)
  double *ovl64;               // overlap in FP64 (energy budget)
  double *knots64;             // fine knots in FP64
  double spacing64;            // smoothing spacing in FP64
  double klo, khi;             // knots[3], knots[end-2]: smoothed-potential region
  int32_t *parts[2];           // current and previous record arrays each (vh_load_pos)
  int cur;
  float *force;                // 3 streams (fx, fy, fz), slot order
  // sort
  int bits_total, passes, digit_bits, nbins;
  int64_t tile, nblocks;
  uint32_t *keys[2], *vals[2];
  uint32_t *counts;            // nbins * nblocks (digit-major), scanned in place
  uint32_t *scan_tmp;          // block sums of the scans
  int64_t scan_tmp_len;
  int32_t *occupied;           // 0-based keys, up to ncell
  int32_t *bounds;             // n_occ + 1
  vh_unit *units;              // CNT_NUNITS work units, in slot order
  int32_t *unitstart;          // n_occ + 1: first unit of each occupied cell
  int64_t maxunits;            // maxocc + ceil(npart / VH_UNIT_MIN)
  int unit;                    // slots per work unit (VH_OPT_UNIT_SLOTS)
  int config_version;          // bumped when an option changes what a step launches
  int32_t *cellmap;            // ncell -> fitting occupied index or -1
  int32_t *fitlist;            // occupied index of fitting cells
  int32_t *nonfit;             // occupied index of non-fitting cells
  int32_t *nonfit_prefix;      // exclusive prefix of their particle counts (+ total)
  int32_t *outlist;            // 0-based slots of out-of-stencil particles
  int32_t *counters;           // CNT_*
  int32_t *cellrun;            // per cell: first slot, end slot of its run in the sorted
                               // cloud (first -1: no particle; reset after each sort)
  float *tiles;                // 512 per work unit (fine deposit)
  long long *cicwins;          // 64 per work unit: the CIC windows of multi-unit cells
  vh_cicwin *cicwtab;          // nk windows of the fine cells on a coarse mesh's points
  uint64_t cicwtab_serial;     // serial of the mesh cicwtab was built for (0: none)
  double *dpart;               // per-block FP64 partials of the reductions
  int64_t dpart_len;
  double *dscal;               // [0..4] verlet diag, [8..11] field reduction, [16] q, [17] enclosed
  int cic_shift;               // fixed-point scale 2^shift of the coarse deposit
  float lo, hi;                // E(knots[2] + h/2), E(knots[end-1] - h/2): deposit bounds
  int64_t maxocc;              // min(ncell, npart)
  uint32_t *cellscratch;       // 5 x (ncell + 1) scan scratch of the cell classification
  int sorted_once;             // host flag: the lists of a sort exist
  int keys_fresh;              // host flag: the push left keys[0] of the current records
  int hist_fresh;              // host flag: and the first pass's tile histograms in counts
  // Records in place after a sort: when the sort follows a push of the
  // step (fresh keys, matrix-core field) it moves nothing; while rec_pending
  // is set slot s's record (both arrays) is record rec_perm[s] of parts[cur]
  // in the order before the sort.  The deposits and the field and fallback
  // kernels read through the permutation, and the push writes the moved
  // records in slot order into the other buffer; other readers first call
  // vh_rec_settle.  ident is the identity permutation.
  int rec_pending;
  const uint32_t *rec_perm;
  uint32_t *ident;
  int nout_idx;                // counters[nout_idx]: the deposit's rejections of this sort
  int pushed_fused;            // host flag: the field kernel's push wrote the records
  int exact_field;             // VH_OPT_EXACT_FIELD: the reference's FMA order
  int poisson_unfused;         // VH_OPT_POISSON_UNFUSED: vh_step forms phi (the entry points' order)
  int field_fp32;              // VH_OPT_FIELD_FP32: FP32 products in the first stage
  int step_graph;              // VH_OPT_STEP_GRAPH
  hipGraphExec_t graph[2];     // vh_step recorded with parts[cur] = parts[0], parts[1]
  double graph_key[15];        // the arguments they were recorded with (VH_GRAPH_KEY)
  vh_step_state graph_after[2];  // the host state each recorded step left
  // the projectile kept on the device (vh_projectile_set)
  double *dproj;               // 16 doubles, see k_projectile_step
  int has_dproj;
  double proj_charge, proj_scale, proj_ball_radius;
  int proj_soft_mode;
  // a cloud held on the host (vh_positions_upload, vh_forces_host): its FP64
  // positions in its own order, and slot -> cloud index after each sort
  double *pos64;
  uint32_t *cloud_perm;
  int track_cloud;             // the next sort starts from the cloud's order
  int cloud_valid;             // cloud_perm holds the current slots' indices
  double *host_csol64;         // FP64 coarse coefficients and knots of vh_forces_host
  double *host_knots64;
  int64_t host_csol_cap, host_knots_cap;
  double *force64;             // npart x 3, the cloud's order
  // the learned denoiser run on the fine density between the deposit and
  // the Poisson solve of vh_step (vh_accel_set_denoiser), and its tolerance
  vh_denoiser *denoiser;
  double denoise_tol;
};

// ---------------------------------------------------------------------------
// small device helpers
// ---------------------------------------------------------------------------
// the sort key of a cell (src/kernels.jl:902-908): clamp(k, 1, nk) per axis,
// 0-based linear key kx + nk (ky + nk kz)
__device__ __forceinline__ uint32_t cell_key0(int kx, int ky, int kz, int nk) {
  kx = kx < 1 ? 1 : (kx > nk ? nk : kx);
  ky = ky < 1 ? 1 : (ky > nk ? nk : ky);
  kz = kz < 1 ? 1 : (kz > nk ? nk : kz);
  return (uint32_t)((kx - 1) + nk * ((ky - 1) + nk * (kz - 1)));
}

#define VH_WAVE 64

__device__ __forceinline__ int vh_lane() { return __lane_id(); }

// x[l] + x[l ^ 16] and x[l] + x[l ^ 32] in every lane l by the gfx950
// row swaps (v_permlane16_swap / v_permlane32_swap, two registers holding
// x: the pair returned is x at l and at the partner lane, in some order) on
// the VALU, where __shfl_xor across rows goes through LDS (ds_bpermute).
// FP32 addition commutes, so the sums equal x + __shfl_xor(x, 16 or 32).
__device__ __forceinline__ float vh_sum_xor16(float x) {
  const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(x), false, false);
  return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
__device__ __forceinline__ float vh_sum_xor32(float x) {
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(x), false, false);
  return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

template <typename T>
__device__ __forceinline__ T vh_wave_sum(T v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// this thread's share of sum p[0..n): p[i] for i = threadIdx.x + k blockDim.x,
// eight loads in flight (eight accumulators, combined in a fixed order)
__device__ __forceinline__ double vh_thread_sum(const double *__restrict__ p, int64_t n) {
  double s[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) s[u] = 0.0;
  const int64_t st = blockDim.x;
  int64_t i = threadIdx.x;
  for (; i + 7 * st < n; i += 8 * st) {
#pragma unroll
    for (int u = 0; u < 8; ++u) s[u] += p[i + u * st];
  }
#pragma unroll
  for (int u = 0; u < 7; ++u)
    if (i + u * st < n) s[u] += p[i + u * st];
  return ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
}

// fixed-order block reduction of a double (blockDim.x multiple of 64, <= 1024)
__device__ __forceinline__ double vh_block_sum_d(double v, double *lds) {
  v = vh_wave_sum(v);
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) lds[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    int nw = (blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; ++i) r += lds[i];
  }
  __syncthreads();
  return r;  // valid in thread 0
}

// ---------------------------------------------------------------------------
// internal entry points shared between translation units
// ---------------------------------------------------------------------------
int vh_alloc(void **p, size_t bytes);
int vh_free(void *p);
int vh_fill_i32(vh_context *c, int32_t *p, int64_t n, int32_t v);
int vh_fill_u64(vh_context *c, uint64_t *p, int64_t n, uint64_t v);
int vh_fill_f32(vh_context *c, float *p, int64_t n, float v);

// exclusive scan of n uint32 in place; returns nothing on host (total at out[n] if with_total)
int vh_scan_exclusive(vh_context *c, uint32_t *data, int64_t n, uint32_t *tmp, int64_t tmp_len,
                      uint32_t *total_out);

// rotating product: out[i' m + j] = sum_i A[i', i] src[j n + i], m = len/n, with
// A given in MFMA operand order by vh_gemm_pack (device copy allocated there)
enum { EPI_STORE = 0, EPI_HADAMARD = 1, EPI_FILL_INTERIOR = 2, EPI_ACCUMULATE = 3 };
int vh_gemm_pack(vh_context *c, float **dst, const double *colmajor, int n);
// the solve of a level and its spline coefficients by linearity (vh_step):
// prepare = M^-1 applied to the density part of the right-hand side (needs
// the density only, kept in inner); finish = once phi's faces hold the
// boundary, csol (=, or += with accumulate) = S^-1 of the solution with
// those faces, without forming phi's interior
#define VH_FUSED_MAXN 256   // the face passes stage one row of at most 256 points
int vh_poisson_fused_prepare_impl(vh_mesh *m);
int vh_poisson_fused_finish_impl(vh_mesh *m, bool accumulate);
// the coarse level's multipole boundary on phi's faces
int vh_poisson_multipole_boundary_impl(vh_mesh *m);
int vh_gemm_pack_b3(vh_context *c, uint32_t **dst, const double *colmajor, int n);
int vh_rotate_gemm(vh_context *c, const float *Aop, const float *src, float *out, int n, int64_t m,
                   int epi, const float *aux, int nfull);

int vh_sort_impl(vh_accel *a);
// moves the pending records into the other buffer in slot order (the
// sort's gather) and makes it current
int vh_rec_settle(vh_accel *a);
// slot s's record is record vh_rec_perm(a)[s] of parts[cur]
static inline const uint32_t *vh_rec_perm(const vh_accel *a) {
  return a->rec_pending ? a->rec_perm : a->ident;
}
int vh_deposit_fine_impl(vh_accel *a, vh_mesh *fine, double charge, vh_sig sig = vh_sig{nullptr, 0});
// defer: leave the sums in the accumulator for vh_multipole_impl (acc_pending)
int vh_deposit_cic_impl(vh_accel *a, vh_mesh *coarse, double charge, bool defer = false);
int vh_poisson_single_impl(vh_mesh *m);
int vh_poisson_nested_impl(vh_mesh *fine, vh_mesh *coarse);
// the nested solve in two parts: the coarse level, its coefficients and the
// fine boundary; then the fine interior
int vh_poisson_coarse_part_impl(vh_mesh *fine, vh_mesh *coarse);
int vh_poisson_coarse_level_impl(vh_mesh *coarse);   // solve and coefficients
int vh_poisson_fine_boundary_impl(vh_mesh *fine, vh_mesh *coarse);
int vh_poisson_fine_part_impl(vh_mesh *fine);
// launches of the scope go to the context's side stream
enum { VH_FLAG_FORK = 0, VH_FLAG_JOIN, VH_FLAG_SIDE, VH_FLAG_TIMEOUTS, VH_FLAG_PROBE, VH_FLAG_COUNT };
static inline vh_sig vh_flag_sig(vh_context *c, int f, uint32_t v) { return vh_sig{c->flags + 16 * f, v}; }
#define VH_FLAG_WORDS (16 * VH_FLAG_COUNT)
// flag f := v once the work queued on st before it is complete
int vh_stream_signal(vh_context *c, hipStream_t st, int f, uint32_t v);
// the work queued on st after this waits until flag f has reached v
int vh_stream_wait(vh_context *c, hipStream_t st, int f, uint32_t v);
// whether a step may order its streams by flags: this context is the only
// one alive in the process and its streams passed the probe at creation
bool vh_stream_flags_usable(const vh_context *c);

struct vh_on_side {
  vh_context *c;
  hipStream_t saved;
  explicit vh_on_side(vh_context *ctx) : c(ctx), saved(ctx->stream) { c->stream = c->side; }
  ~vh_on_side() { c->stream = saved; }
};
int vh_spline_coefficients_impl(vh_mesh *m, const float *src, float *dst);
int vh_effective_potential_impl(vh_mesh *m, double nions, double radius,
                                vh_sig sig = vh_sig{nullptr, 0});
int vh_potential_coefficients_impl(vh_mesh *m, double nions, double radius);
// its two halves: csol = S^-1 V(rho) (needs the density only), csol += S^-1 phi
int vh_mean_field_coefficients_impl(vh_mesh *m, double nions, double radius);
int vh_hartree_coefficients_add_impl(vh_mesh *m);
int vh_total_charge_impl(vh_mesh *m, const float *rho, double *dq);
// the denoiser on the fine density in place, with the residual charge check
// (no host synchronisation, nothing allocated once attached)
int vh_denoiser_apply_impl(vh_denoiser *d, vh_mesh *m, double tol);
int vh_denoiser_prepare(vh_denoiser *d, int n);
vh_context *vh_denoiser_context(const vh_denoiser *d);
// push_a > 0: the field and fallback kernels also push their particles
// (Verlet factor dt^2 / mass) and leave the next sort's keys; requires the
// matrix-core field kernel and vh_verlet_tiles
// VH_FALLBACK_NONE leaves the out-of-stencil slots of the force buffer to the
// caller (vh_forces_host resolves them in FP64 from the host positions)
// fallback: VH_FALLBACK_AFTER launches vh_fallback_impl after the field
// kernel; VH_FALLBACK_DONE: the caller has launched it (with the same push),
// ordered before this call's reduction (which steps the projectile);
// VH_FALLBACK_NONE leaves those slots alone.
// proj_dt > 0 (with dproj): the projectile is stepped on the device as well
// (vh_projectile_step_impl), by the kernel that sums its forces
enum { VH_FALLBACK_NONE = 0, VH_FALLBACK_AFTER = 1, VH_FALLBACK_DONE = 2 };
int vh_forces_impl(vh_accel *a, vh_mesh *fine, vh_mesh *coarse, double weight, const double *proj,
                   int soft_mode, double ball_radius, const double *dproj, double push_a = 0.0,
                   int fallback = VH_FALLBACK_AFTER, double proj_dt = 0.0);
int vh_fallback_impl(vh_accel *a, vh_mesh *coarse, double weight, const double *proj,
                     double ball_radius, const double *dproj, double push_a);
int vh_projectile_step_impl(vh_accel *a, double dt);
int vh_verlet_impl(vh_accel *a, double dt, double mass, double rcmax);
// the push leaves the next sort's keys and first-digit histograms
inline bool vh_verlet_tiles(const vh_accel *a) { return a->nbins * a->nblocks <= (int64_t)1 << 31; }
int vh_energy_impl(vh_accel *a, vh_mesh *fine, vh_mesh *coarse, double weight, double enclosed);
