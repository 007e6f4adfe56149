// Context, meshes, accelerators, buffers and the public ABI wrappers.
#include "vh_internal.h"

#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <vector>

static thread_local char g_err[1024] = "";

void vh_set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}

extern "C" const char *vh_last_error(void) { return g_err; }
extern "C" const char *vh_version(void) { return "vlasov_hip 0.1 (gfx950)"; }

extern "C" int vh_device_count(int *count) {
  VH_REQUIRE(count != nullptr, "vh_device_count: null output");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) c = 0;
  *count = c;
  return VH_OK;
}

int vh_alloc(void **p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 256;
  VH_CHECK_HIP(hipMalloc(p, bytes));
  return VH_OK;
}

int vh_free(void *p) {
  if (p) VH_CHECK_HIP(hipFree(p));
  return VH_OK;
}

// ---------------------------------------------------------------------------
// fills (kernels rather than hipMemsetAsync so that a captured step stays valid)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_fill(T *p, int64_t n, T v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

template <typename T>
static int fill_impl(vh_context *c, T *p, int64_t n, T v) {
  if (n <= 0) return VH_OK;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)blocks), dim3(256), 0, c->stream, p, n, v);
  VH_LAUNCH_CHECK();
  return VH_OK;
}
int vh_fill_i32(vh_context *c, int32_t *p, int64_t n, int32_t v) { return fill_impl(c, p, n, v); }
int vh_fill_u64(vh_context *c, uint64_t *p, int64_t n, uint64_t v) { return fill_impl(c, p, n, v); }
int vh_fill_f32(vh_context *c, float *p, int64_t n, float v) { return fill_impl(c, p, n, v); }

__global__ void k_iota(uint32_t *p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = (uint32_t)i;
}

static int fill_iota(vh_context *c, uint32_t *p, int64_t n) {
  hipLaunchKernelGGL(k_iota, dim3(1024), dim3(256), 0, c->stream, p, n);
  VH_LAUNCH_CHECK();
  return VH_OK;
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
// A flag wait is a kernel that spins: it holds the hardware queue its stream
// runs on, and the packets behind it in that queue, until the flag is set.
// When the runtime has more streams than hardware queues (four per process
// by default) it puts several streams on one queue, and a wait on the side
// stream whose signal is queued behind it on the same queue would only end
// at its time limit; so would a cycle through the queues of two contexts
// whose steps interleave.  Flags therefore order a step's streams only when
// (a) the context is the only one alive in the process and (b) at creation a
// wait on the side stream ended on a signal queued after it on the main one,
// i.e. the two streams do not share a queue.  Otherwise events order them.
static std::atomic<int> g_live_contexts{0};

__global__ void k_flag_signal(uint32_t *__restrict__ f, uint32_t v);
__global__ void k_flag_wait(const uint32_t *__restrict__ f, uint32_t v, uint32_t *__restrict__ timeouts,
                            uint64_t limit);

static int probe_stream_queues(vh_context *c) {
  uint32_t *f = c->flags + 16 * VH_FLAG_PROBE;
  // 20 ms of the 100 MHz constant clock: a shared queue costs the creation that
  hipLaunchKernelGGL(k_flag_wait, dim3(1), dim3(1), 0, c->side, f, 1u, f + 4, (uint64_t)2000000);
  VH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_flag_signal, dim3(1), dim3(1), 0, c->stream, f, 1u);
  VH_LAUNCH_CHECK();
  VH_CHECK_HIP(hipStreamSynchronize(c->side));
  VH_CHECK_HIP(hipStreamSynchronize(c->stream));
  uint32_t missed = 1;
  VH_CHECK_HIP(hipMemcpy(&missed, f + 4, sizeof missed, hipMemcpyDeviceToHost));
  c->flags_ok = missed == 0;
  return VH_OK;
}

bool vh_stream_flags_usable(const vh_context *c) {
  return c->flags_ok && g_live_contexts.load() == 1;
}

extern "C" int vh_context_create(int device, vh_context **out) {
  VH_REQUIRE(out != nullptr, "vh_context_create: null output");
  *out = nullptr;
  int count = 0;
  VH_CHECK_HIP(hipGetDeviceCount(&count));
  VH_REQUIRE(device >= 0 && device < count, "vh_context_create: device %d of %d", device, count);
  VH_CHECK_HIP(hipSetDevice(device));
  vh_context *c = new vh_context();
  c->device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    vh_set_error("hipStreamCreate: %s", hipGetErrorString(e));
    return VH_EHIP;
  }
  // the side stream at the default priority, like the main one
  if ((e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming)) != hipSuccess ||
      (e = hipEventCreateWithFlags(&c->join_ev, hipEventDisableTiming)) != hipSuccess ||
      (e = hipEventCreateWithFlags(&c->ev_a, hipEventDisableTiming)) != hipSuccess) {
    vh_set_error("vh_context_create: %s", hipGetErrorString(e));
    hipStreamDestroy(c->stream);
    delete c;
    return VH_EHIP;
  }
  hipDeviceProp_t prop;
  VH_CHECK_HIP(hipGetDeviceProperties(&prop, device));
  c->num_cus = prop.multiProcessorCount;
  if (vh_alloc((void **)&c->flags, sizeof(uint32_t) * VH_FLAG_WORDS) != VH_OK ||
      vh_fill_i32(c, (int32_t *)c->flags, VH_FLAG_WORDS, 0) != VH_OK ||
      hipStreamSynchronize(c->stream) != hipSuccess) {
    vh_free(c->flags);
    hipEventDestroy(c->fork_ev);
    hipEventDestroy(c->join_ev);
    hipEventDestroy(c->ev_a);
    hipStreamDestroy(c->side);
    hipStreamDestroy(c->stream);
    delete c;
    return VH_EHIP;
  }
  c->flag_seq = 0;
  c->flags_ok = false;
  c->live = false;
  if (probe_stream_queues(c) != VH_OK) {
    vh_context_destroy(c);
    return VH_EHIP;
  }
  g_live_contexts.fetch_add(1);
  c->live = true;
  *out = c;
  return VH_OK;
}

// The flags carry no data: the work before a signal has completed and
// released its writes at its kernels' ends, and the work after a wait
// acquires them at its kernels' starts.  Relaxed device-scope accesses
// therefore suffice, and they keep the L2 caches intact (a release or
// acquire at that scope writes back or invalidates the L2 of the XCD the
// wave runs on, each time: while the other stream works).
__global__ void k_flag_signal(uint32_t *__restrict__ f, uint32_t v) {
  __hip_atomic_store(f, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One lane polls the flag; gives up after `limit` ticks of the 100 MHz
// constant clock (or 2^24 polls) and counts the timeout.  The work queued
// behind the wait starts when this kernel ends.
__global__ void k_flag_wait(const uint32_t *__restrict__ f, uint32_t v, uint32_t *__restrict__ timeouts,
                            uint64_t limit) {
  const uint64_t t0 = wall_clock64();
  for (uint32_t i = 0; i < (1u << 24); ++i) {
    const uint32_t cur = __hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((int32_t)(cur - v) >= 0) return;
    if (wall_clock64() - t0 > limit) break;
    __builtin_amdgcn_s_sleep(8);
  }
  atomicAdd(timeouts, 1u);
}

int vh_stream_signal(vh_context *c, hipStream_t st, int f, uint32_t v) {
  hipLaunchKernelGGL(k_flag_signal, dim3(1), dim3(1), 0, st, c->flags + 16 * f, v);
  VH_LAUNCH_CHECK();
  return VH_OK;
}

int vh_stream_wait(vh_context *c, hipStream_t st, int f, uint32_t v) {
  // two seconds
  hipLaunchKernelGGL(k_flag_wait, dim3(1), dim3(1), 0, st, c->flags + 16 * f, v,
                     c->flags + 16 * VH_FLAG_TIMEOUTS, (uint64_t)200000000);
  VH_LAUNCH_CHECK();
  return VH_OK;
}

extern "C" int vh_context_destroy(vh_context *ctx) {
  if (!ctx) return VH_OK;
  if (ctx->live) g_live_contexts.fetch_sub(1);
  ctx->live = false;
  VH_CHECK_HIP(hipSetDevice(ctx->device));
  VH_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  VH_CHECK_HIP(hipStreamSynchronize(ctx->side));
  vh_free(ctx->flags);
  VH_CHECK_HIP(hipEventDestroy(ctx->fork_ev));
  VH_CHECK_HIP(hipEventDestroy(ctx->join_ev));
  VH_CHECK_HIP(hipEventDestroy(ctx->ev_a));
  VH_CHECK_HIP(hipStreamDestroy(ctx->side));
  VH_CHECK_HIP(hipStreamDestroy(ctx->stream));
  delete ctx;
  return VH_OK;
}

extern "C" int vh_sync(vh_context *ctx) {
  VH_REQUIRE(ctx != nullptr, "vh_sync: null context");
  VH_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  VH_CHECK_HIP(hipStreamSynchronize(ctx->side));
  uint32_t timeouts = 0;
  VH_CHECK_HIP(hipMemcpy(&timeouts, ctx->flags + 16 * VH_FLAG_TIMEOUTS, sizeof timeouts,
                         hipMemcpyDeviceToHost));
  VH_REQUIRE(timeouts == 0, "vh_sync: %u wait(s) between the step's streams timed out", timeouts);
  return VH_OK;
}

// ---------------------------------------------------------------------------
// upload helpers
// ---------------------------------------------------------------------------
template <typename T>
static int upload_converted(vh_context *c, T **dst, const double *src, size_t n) {
  std::vector<T> tmp(n);
  for (size_t i = 0; i < n; ++i) tmp[i] = (T)src[i];
  VH_TRY(vh_alloc((void **)dst, n * sizeof(T)));
  VH_CHECK_HIP(hipMemcpy(*dst, tmp.data(), n * sizeof(T), hipMemcpyHostToDevice));
  return VH_OK;
}

// column-major (Julia) n x n double matrix -> row-major float
static int upload_rowmajor(vh_context *c, float **dst, const double *colmajor, int n) {
  std::vector<double> t((size_t)n * n);
  for (int r = 0; r < n; ++r)
    for (int cc = 0; cc < n; ++cc) t[(size_t)r * n + cc] = colmajor[r + (size_t)n * cc];
  return upload_converted(c, dst, t.data(), t.size());
}

// ---------------------------------------------------------------------------
// mesh
// ---------------------------------------------------------------------------
__global__ void k_invlsum(float *out, const double *lam, int ni) {
  int64_t tot = (int64_t)ni * ni * ni;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot;
       t += (int64_t)gridDim.x * blockDim.x) {
    int i = (int)(t % ni), j = (int)((t / ni) % ni), k = (int)(t / ((int64_t)ni * ni));
    double s = (lam[i] + lam[j]) + lam[k];
    out[t] = (float)(1.0 / s);
  }
}

extern "C" int vh_mesh_destroy(vh_mesh *m) {
  if (!m) return VH_OK;
  hipSetDevice(m->ctx->device);
  hipStreamSynchronize(m->ctx->stream);
  void *ptrs[] = {m->colloc, m->knots, m->colloc64, m->knots64, m->lap, m->Minv_op, m->M_op,
                  m->Sinv_op, m->invlsum, m->dual, m->dual32, m->duallen, m->loc_cells,
                  m->rho, m->phi, m->csol, m->extra, m->work, m->inner, m->work1, m->work2,
                  m->partials, m->scal, m->acc64, m->Minv_b3, m->M_b3, m->Sinv_b3,
                  m->Pt_op, m->ulr, m->Minv_t, m->face};
  for (void *p : ptrs) vh_free(p);
  delete m;
  return VH_OK;
}

extern "C" int vh_mesh_create(vh_context *ctx, int n, int nk, const double *colloc,
                              const double *knots, const double *lap, const double *M,
                              const double *Minv, const double *lam, const double *Sinv,
                              const double *dual, const double *duallen, double loc_x0,
                              double loc_invwidth, const int32_t *loc_cells, int nbin,
                              vh_mesh **out) {
  VH_REQUIRE(ctx && out, "vh_mesh_create: null argument");
  VH_REQUIRE(n == 2 * nk && nk >= 4, "vh_mesh_create: n (%d) must equal 2 nk (nk=%d >= 4)", n, nk);
  VH_REQUIRE(n <= 1024, "vh_mesh_create: n = %d exceeds 1024", n);
  VH_REQUIRE(colloc && knots && lap && M && Minv && lam && Sinv && dual && duallen && loc_cells,
             "vh_mesh_create: null table");
  VH_REQUIRE(nbin >= 1, "vh_mesh_create: empty locate table");
  for (int b = 0; b < nbin; ++b)
    VH_REQUIRE(loc_cells[b] >= 1 && loc_cells[b] <= n - 1, "vh_mesh_create: locate cell out of range");
  VH_CHECK_HIP(hipSetDevice(ctx->device));
  vh_mesh *m = new vh_mesh();
  memset(m, 0, sizeof *m);
  static std::atomic<uint64_t> mesh_serial{0};
  m->serial = ++mesh_serial;
  m->ctx = ctx;
  m->n = n;
  m->nk = nk;
  m->ni = n - 2;
  int ni = m->ni;
  int r = VH_OK;
#define MK(x)                 \
  if ((r = (x)) != VH_OK) {   \
    vh_mesh_destroy(m);       \
    return r;                 \
  }
  MK(upload_converted(ctx, &m->colloc, colloc, n));
  MK(upload_converted(ctx, &m->knots, knots, nk));
  MK(upload_converted(ctx, &m->colloc64, colloc, n));
  MK(upload_converted(ctx, &m->knots64, knots, nk));
  MK(upload_rowmajor(ctx, &m->lap, lap, n));
  MK(vh_gemm_pack(ctx, &m->Minv_op, Minv, ni));
  MK(vh_gemm_pack(ctx, &m->M_op, M, ni));
  MK(vh_gemm_pack(ctx, &m->Sinv_op, Sinv, n));
  MK(vh_gemm_pack_b3(ctx, &m->Minv_b3, Minv, ni));
  MK(vh_gemm_pack_b3(ctx, &m->M_b3, M, ni));
  MK(vh_gemm_pack_b3(ctx, &m->Sinv_b3, Sinv, n));
  {  // the operators of the solve by linearity, in FP64, rounded once
    // P~ = S^-1 M~: column 0 and n - 1 of S^-1, interior columns S^-1 E M
    std::vector<double> pt((size_t)n * n, 0.0);
    for (int row = 0; row < n; ++row) {
      pt[row] = Sinv[row];
      pt[row + (size_t)n * (n - 1)] = Sinv[row + (size_t)n * (n - 1)];
      for (int cc = 0; cc < ni; ++cc) {
        double acc = 0.0;
        for (int q = 0; q < ni; ++q) acc += Sinv[row + (size_t)n * (q + 1)] * M[q + (size_t)ni * cc];
        pt[row + (size_t)n * (cc + 1)] = acc;
      }
    }
    MK(vh_gemm_pack(ctx, &m->Pt_op, pt.data(), n));
    // u = M^-1 of the outer columns of D at the interior rows (the faces'
    // lift into the right-hand side, src/poisson.jl:275-326)
    std::vector<double> u(2 * (size_t)ni), mt((size_t)ni * ni);
    for (int a = 0; a < ni; ++a) {
      double ul = 0.0, ur = 0.0;
      for (int q = 0; q < ni; ++q) {
        ul += Minv[a + (size_t)ni * q] * lap[(q + 1) + (size_t)n * 0];
        ur += Minv[a + (size_t)ni * q] * lap[(q + 1) + (size_t)n * (n - 1)];
        mt[(size_t)q * ni + a] = Minv[a + (size_t)ni * q];  // transposed
      }
      u[a] = ul;
      u[ni + a] = ur;
    }
    MK(upload_converted(ctx, &m->ulr, u.data(), u.size()));
    MK(upload_converted(ctx, &m->Minv_t, mt.data(), mt.size()));
    MK(vh_alloc((void **)&m->face, sizeof(float) * (12 * (size_t)ni * ni + 12 * (size_t)ni)));
  }
  MK(upload_converted(ctx, &m->dual, dual, 3 * (size_t)n));
  MK(upload_converted(ctx, &m->dual32, dual, 3 * (size_t)n));
  MK(upload_converted(ctx, &m->duallen, duallen, n));
  m->loc_x0 = (float)loc_x0;
  m->loc_iw = (float)loc_invwidth;
  m->nbin = nbin;
  MK(vh_alloc((void **)&m->loc_cells, sizeof(int32_t) * nbin));
  if (hipMemcpy(m->loc_cells, loc_cells, sizeof(int32_t) * nbin, hipMemcpyHostToDevice) != hipSuccess) {
    vh_mesh_destroy(m);
    vh_set_error("vh_mesh_create: copy of the locate table failed");
    return VH_EHIP;
  }
  size_t g = (size_t)n * n * n, gi = (size_t)ni * ni * ni;
  MK(vh_alloc((void **)&m->rho, g * 4));
  MK(vh_alloc((void **)&m->phi, g * 4));
  MK(vh_alloc((void **)&m->csol, g * 4));
  MK(vh_alloc((void **)&m->extra, g * 4));
  MK(vh_alloc((void **)&m->work, g * 4));
  MK(vh_alloc((void **)&m->inner, gi * 4));
  MK(vh_alloc((void **)&m->work1, gi * 4));
  MK(vh_alloc((void **)&m->work2, gi * 4));
  MK(vh_alloc((void **)&m->invlsum, gi * 4));
  MK(vh_alloc((void **)&m->partials, sizeof(double) * 10 * (size_t)n * n));
  MK(vh_alloc((void **)&m->scal, sizeof(double) * 32));
  MK(vh_alloc((void **)&m->acc64, g * 8));
  {
    double *dlam = nullptr;
    MK(upload_converted(ctx, &dlam, lam, ni));
    hipLaunchKernelGGL(k_invlsum, dim3(1024), dim3(256), 0, ctx->stream, m->invlsum, dlam, ni);
    hipError_t e = hipStreamSynchronize(ctx->stream);
    hipFree(dlam);
    if (e != hipSuccess) {
      vh_mesh_destroy(m);
      vh_set_error("vh_mesh_create: %s", hipGetErrorString(e));
      return VH_EHIP;
    }
  }
  // every buffer starts at zero: device memory comes back from the allocator
  // holding whatever an earlier owner left, and the CIC accumulator in
  // particular is summed into (the deposit clears it again after each use)
  MK(vh_fill_f32(ctx, m->rho, (int64_t)g, 0.f));
  MK(vh_fill_f32(ctx, m->phi, (int64_t)g, 0.f));
  MK(vh_fill_f32(ctx, m->csol, (int64_t)g, 0.f));
  MK(vh_fill_f32(ctx, m->extra, (int64_t)g, 0.f));
  MK(vh_fill_f32(ctx, m->work, (int64_t)g, 0.f));
  MK(vh_fill_f32(ctx, m->inner, (int64_t)gi, 0.f));
  MK(vh_fill_f32(ctx, m->work1, (int64_t)gi, 0.f));
  MK(vh_fill_f32(ctx, m->work2, (int64_t)gi, 0.f));
  MK(vh_fill_f32(ctx, m->face, 12 * (int64_t)ni * ni + 12 * (int64_t)ni, 0.f));
  MK(vh_fill_u64(ctx, (uint64_t *)m->acc64, (int64_t)g, 0ull));
  MK(vh_fill_u64(ctx, (uint64_t *)m->partials, (int64_t)10 * n * n, 0ull));
  MK(vh_fill_u64(ctx, (uint64_t *)m->scal, 32, 0ull));
#undef MK
  VH_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  *out = m;
  return VH_OK;
}

static float *mesh_grid(vh_mesh *m, int which) {
  switch (which) {
    case VH_GRID_RHO: return m->rho;
    case VH_GRID_PHI: return m->phi;
    case VH_GRID_CSOL: return m->csol;
    default: return nullptr;
  }
}

extern "C" int vh_mesh_grid_download(vh_mesh *m, int which, float *host) {
  VH_REQUIRE(m && host, "vh_mesh_grid_download: null argument");
  float *g = mesh_grid(m, which);
  VH_REQUIRE(g != nullptr, "vh_mesh_grid_download: unknown grid %d", which);
  VH_CHECK_HIP(hipSetDevice(m->ctx->device));
  size_t bytes = (size_t)m->n * m->n * m->n * 4;
  VH_CHECK_HIP(hipMemcpyAsync(host, g, bytes, hipMemcpyDeviceToHost, m->ctx->stream));
  VH_CHECK_HIP(hipStreamSynchronize(m->ctx->stream));
  return VH_OK;
}

// (a + b) / 2 in FP32 is the FP32 rounding of the exact mean (the halving is
// exact), the value the script's Float32 of its FP64 mean gives
__global__ void k_density_slice(const float *__restrict__ rho, int n, int z1, int z2,
                                float *__restrict__ out) {
  const int64_t nn = (int64_t)n * n;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nn;
       t += (int64_t)gridDim.x * blockDim.x)
    out[t] = (rho[t + nn * z1] + rho[t + nn * z2]) * 0.5f;
}

extern "C" int vh_density_slice(vh_mesh *m, int iz1, int iz2, float *host) {
  VH_REQUIRE(m && host, "vh_density_slice: null argument");
  VH_REQUIRE(iz1 >= 1 && iz1 <= m->n && iz2 >= 1 && iz2 <= m->n,
             "vh_density_slice: planes %d, %d outside 1..%d", iz1, iz2, m->n);
  VH_CHECK_HIP(hipSetDevice(m->ctx->device));
  // staged in the mesh's transform scratch (n^3 floats), free between calls
  hipLaunchKernelGGL(k_density_slice, dim3(64), dim3(256), 0, m->ctx->stream, m->rho, m->n,
                     iz1 - 1, iz2 - 1, m->work);
  VH_LAUNCH_CHECK();
  VH_CHECK_HIP(hipMemcpyAsync(host, m->work, (size_t)m->n * m->n * 4, hipMemcpyDeviceToHost,
                              m->ctx->stream));
  VH_CHECK_HIP(hipStreamSynchronize(m->ctx->stream));
  return VH_OK;
}

extern "C" int vh_mesh_grid_upload(vh_mesh *m, int which, const float *host) {
  VH_REQUIRE(m && host, "vh_mesh_grid_upload: null argument");
  float *g = mesh_grid(m, which);
  VH_REQUIRE(g != nullptr, "vh_mesh_grid_upload: unknown grid %d", which);
  VH_CHECK_HIP(hipSetDevice(m->ctx->device));
  size_t bytes = (size_t)m->n * m->n * m->n * 4;
  VH_CHECK_HIP(hipMemcpyAsync(g, host, bytes, hipMemcpyHostToDevice, m->ctx->stream));
  VH_CHECK_HIP(hipStreamSynchronize(m->ctx->stream));
  return VH_OK;
}

extern "C" int vh_mesh_set_option(vh_mesh *m, int option, int value) {
  VH_REQUIRE(m != nullptr, "vh_mesh_set_option: null mesh");
  VH_REQUIRE(option == VH_MESH_OPT_BF16X3, "vh_mesh_set_option: unknown option %d", option);
  VH_REQUIRE(value >= 0 && value <= 3, "vh_mesh_set_option: value %d outside 0..3", value);
  m->gemm_bf16x3 = value;
  return VH_OK;
}

extern "C" int vh_mesh_grid_copy(vh_mesh *m, int dst, int src) {
  VH_REQUIRE(m != nullptr, "vh_mesh_grid_copy: null mesh");
  float *d = mesh_grid(m, dst), *s = mesh_grid(m, src);
  VH_REQUIRE(d && s, "vh_mesh_grid_copy: unknown grid %d or %d", dst, src);
  if (d == s) return VH_OK;
  VH_CHECK_HIP(hipSetDevice(m->ctx->device));
  const size_t bytes = (size_t)m->n * m->n * m->n * 4;
  VH_CHECK_HIP(hipMemcpyAsync(d, s, bytes, hipMemcpyDeviceToDevice, m->ctx->stream));
  return VH_OK;
}

// ---------------------------------------------------------------------------
// accelerator
// ---------------------------------------------------------------------------
static void drop_graphs(vh_accel *a);

extern "C" int vh_accel_destroy(vh_accel *a) {
  if (!a) return VH_OK;
  hipSetDevice(a->ctx->device);
  hipStreamSynchronize(a->ctx->stream);
  drop_graphs(a);
  void *ptrs[] = {a->ovl, a->grad, a->nodes, a->og, a->xsplit, a->ovl64, a->knots64, a->parts[0], a->parts[1], a->force,
                  a->keys[0], a->keys[1], a->vals[0], a->vals[1], a->counts, a->scan_tmp,
                  a->occupied, a->bounds, a->cellmap, a->fitlist, a->nonfit, a->nonfit_prefix,
                  a->outlist, a->counters, a->cellrun, a->tiles, a->cicwins, a->cicwtab, a->dpart, a->dscal, a->cellscratch,
                  a->dproj, a->units, a->unitstart, a->zero, a->ident, a->pos64, a->cloud_perm,
                  a->host_csol64, a->host_knots64, a->force64};
  for (void *p : ptrs) vh_free(p);
  delete a;
  return VH_OK;
}

extern "C" int vh_accel_create(vh_context *ctx, int64_t npart, int nk, const double *knots,
                               double spacing, int nbdt, const double *overlap,
                               const double *gradient, const double *nodes, vh_accel **out) {
  VH_REQUIRE(ctx && out && knots && overlap && gradient && nodes, "vh_accel_create: null argument");
  VH_REQUIRE(npart >= 1 && npart <= ((int64_t)1 << 31) - 4096,
             "vh_accel_create: npart = %lld out of range", (long long)npart);
  VH_REQUIRE(nk >= 6 && nk <= 512, "vh_accel_create: nk = %d out of range", nk);
  // the field kernel packs three table columns into one word, 10 bits each
  VH_REQUIRE(nbdt >= 1 && nbdt <= 1023, "vh_accel_create: nbdt = %d out of [1, 1023]", nbdt);
  const double x0 = knots[0];
  const double h = (knots[nk - 1] - knots[0]) / (nk - 1);
  VH_REQUIRE(h > 0 && std::isfinite(h) && std::isfinite(x0), "vh_accel_create: bad axis");
  for (int j = 0; j + 1 < nk; ++j)
    VH_REQUIRE(std::fabs((knots[j + 1] - knots[j]) - h) <= 1e-9 * std::fabs(h),
               "vh_accel_create: the accelerated path assumes a uniform fine grid");
  VH_REQUIRE(spacing > 0 && std::isfinite(spacing), "vh_accel_create: bad smoothing spacing");
  *out = nullptr;
  VH_CHECK_HIP(hipSetDevice(ctx->device));
  vh_accel *a = new vh_accel();
  memset(a, 0, sizeof *a);
  a->ctx = ctx;
  a->npart = npart;
  a->npad = (npart + 255) / 256 * 256;
  a->nk = nk;
  a->n = 2 * nk;
  a->ncell = nk * nk * nk;
  a->x0 = x0;
  a->h = h;
  a->spacing = (float)spacing;
  a->spacing64 = spacing;
  a->klo = knots[2];
  a->khi = knots[nk - 3];
  a->lo = (float)(knots[1] + spacing / 2);
  a->hi = (float)(knots[nk - 2] - spacing / 2);
  a->nbdt = nbdt;
  a->ncol = nbdt + 1;
  // radix digits: ceil(log2(ncell)) bits in passes of at most 11
  int bits = 1;
  while (((int64_t)1 << bits) < a->ncell) ++bits;
  a->bits_total = bits;
  a->passes = (bits + 10) / 11;
  a->digit_bits = (bits + a->passes - 1) / a->passes;
  a->nbins = 1 << a->digit_bits;
  a->tile = 4096;
  a->nblocks = (npart + a->tile - 1) / a->tile;
  // coarse CIC fixed point: a point's sum is at most npart
  int lg = 0;
  while (((int64_t)1 << lg) <= npart) ++lg;
  a->cic_shift = 61 - lg;
  if (a->cic_shift > 40) a->cic_shift = 40;
  int r = VH_OK;
#define MK(x)                 \
  if ((r = (x)) != VH_OK) {   \
    vh_accel_destroy(a);      \
    return r;                 \
  }
  // tables: the reference's column-major (a + 10 col) layout is kept
  MK(upload_converted(ctx, &a->ovl, overlap, (size_t)10 * a->ncol));
  MK(upload_converted(ctx, &a->grad, gradient, (size_t)10 * a->ncol));
  MK(upload_converted(ctx, &a->nodes, nodes, (size_t)8 * a->ncol));
  MK(upload_converted(ctx, &a->ovl64, overlap, (size_t)10 * a->ncol));
  {  // overlap and gradient of a column side by side, 80 bytes of a 128-byte
     // line per column (the field kernel loads a column as one line)
    std::vector<double> og((size_t)32 * a->ncol, 0.0);
    for (int c = 0; c < a->ncol; ++c)
      for (int k = 0; k < 10; ++k) {
        og[(size_t)32 * c + k] = overlap[(size_t)10 * c + k];
        og[(size_t)32 * c + 10 + k] = gradient[(size_t)10 * c + k];
      }
    MK(upload_converted(ctx, &a->og, og.data(), og.size()));
  }
  {  // the x columns split into fp16 pairs for the first stage of the
     // matrix-core field kernel, each kind (overlap, gradient) scaled by the
     // power of two that puts its largest FP32 value in [2^14, 2^15)
    std::vector<uint16_t> xs((size_t)64 * a->ncol);
    const double *tab[2] = {overlap, gradient};
    for (int t = 0; t < 2; ++t) {
      float mx = 0.f;
      for (size_t i = 0; i < (size_t)10 * a->ncol; ++i) mx = std::max(mx, std::fabs((float)tab[t][i]));
      int e = 0;
      std::frexp(mx > 0.f ? mx : 1.f, &e);  // mx = m 2^e, m in [0.5, 1)
      const float sc = std::ldexp(1.f, 15 - e);
      a->xs_inv[t] = std::ldexp(1.f, e - 15);
      for (int c = 0; c < a->ncol; ++c)
        for (int k = 0; k < 32; ++k) {
          int ii, cp, wp;
          vh_s16_slot(k, ii, cp, wp);
          const float v = (float)tab[t][(size_t)10 * c + ii] * sc;  // exact: a power of two
          const _Float16 hi = (_Float16)v;
          const _Float16 lo = (_Float16)(v - (float)hi);
          const _Float16 x = wp < 0 ? (_Float16)0.f : wp == 0 ? hi : lo;
          xs[(size_t)64 * c + 32 * t + k] = __builtin_bit_cast(uint16_t, x);
        }
    }
    MK(vh_alloc((void **)&a->xsplit, xs.size() * sizeof(uint16_t)));
    VH_CHECK_HIP(hipMemcpy(a->xsplit, xs.data(), xs.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
  }
  MK(upload_converted(ctx, &a->knots64, knots, (size_t)nk));
  size_t np = (size_t)a->npad;
  MK(vh_alloc((void **)&a->parts[0], np * 4 * VH_REC_WORDS));
  MK(vh_alloc((void **)&a->parts[1], np * 4 * VH_REC_WORDS));
  MK(vh_alloc((void **)&a->force, np * 4 * 3));
  MK(vh_alloc((void **)&a->keys[0], np * 4));
  MK(vh_alloc((void **)&a->keys[1], np * 4));
  MK(vh_alloc((void **)&a->vals[0], np * 4));
  MK(vh_alloc((void **)&a->vals[1], np * 4));
  MK(vh_alloc((void **)&a->counts, (size_t)a->nbins * a->nblocks * 4 + 64));
  a->scan_tmp_len = ((int64_t)a->nbins * a->nblocks + np) / 1024 + 4096;
  MK(vh_alloc((void **)&a->scan_tmp, (size_t)a->scan_tmp_len * 4));
  int64_t maxocc = a->ncell < npart ? a->ncell : npart;
  a->maxocc = maxocc;
  a->maxunits = maxocc + (npart + VH_UNIT_MIN - 1) / VH_UNIT_MIN;
  a->unit = VH_UNIT;
  MK(vh_alloc((void **)&a->cellscratch, sizeof(uint32_t) * 5 * (a->ncell + 1)));
  MK(vh_alloc((void **)&a->units, sizeof(vh_unit) * (size_t)a->maxunits));
  MK(vh_alloc((void **)&a->unitstart, sizeof(int32_t) * (maxocc + 2)));
  MK(vh_alloc((void **)&a->occupied, sizeof(int32_t) * (maxocc + 1)));
  MK(vh_alloc((void **)&a->bounds, sizeof(int32_t) * (maxocc + 2)));
  MK(vh_alloc((void **)&a->cellmap, sizeof(int32_t) * a->ncell));
  MK(vh_alloc((void **)&a->fitlist, sizeof(int32_t) * (maxocc + 1)));
  MK(vh_alloc((void **)&a->nonfit, sizeof(int32_t) * (maxocc + 1)));
  MK(vh_alloc((void **)&a->nonfit_prefix, sizeof(int32_t) * (maxocc + 2)));
  MK(vh_alloc((void **)&a->outlist, np * 4));
  MK(vh_alloc((void **)&a->counters, sizeof(int32_t) * 64));
  MK(vh_alloc((void **)&a->cellrun, sizeof(int32_t) * 2 * a->ncell));
  MK(vh_alloc((void **)&a->tiles, sizeof(float) * 512 * (size_t)a->maxunits));
  MK(vh_alloc((void **)&a->cicwins, sizeof(long long) * 64 * (size_t)a->maxunits));
  MK(vh_alloc((void **)&a->cicwtab, sizeof(vh_cicwin) * (size_t)nk));
  a->cicwtab_serial = 0;
  a->dpart_len = 1 << 16;
  MK(vh_alloc((void **)&a->dpart, sizeof(double) * 8 * a->dpart_len));
  MK(vh_alloc((void **)&a->dscal, sizeof(double) * 64));
  MK(vh_alloc((void **)&a->zero, sizeof(float) * 64));
  MK(vh_alloc((void **)&a->ident, np * 4));
  MK(vh_fill_f32(ctx, a->zero, 64, 0.f));
  MK(vh_fill_i32(ctx, a->counters, 64, 0));
  MK(vh_fill_i32(ctx, a->cellrun, 2 * a->ncell, -1));
  MK(vh_fill_f32(ctx, a->force, (int64_t)np * 3, 0.f));
  MK(vh_fill_i32(ctx, a->parts[0], (int64_t)np * VH_REC_WORDS, 0));
  MK(vh_fill_i32(ctx, a->parts[1], (int64_t)np * VH_REC_WORDS, 0));
  MK(vh_fill_u64(ctx, (uint64_t *)a->dscal, 64, 0ull));
  MK(fill_iota(ctx, a->ident, (int64_t)np));
#undef MK
  VH_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  a->cur = 0;
  a->nout_idx = CNT_NOUT;
  *out = a;
  return VH_OK;
}

extern "C" int vh_accel_set_option(vh_accel *a, int option, int value) {
  VH_REQUIRE(a != nullptr, "vh_accel_set_option: null accelerator");
  switch (option) {
    case VH_OPT_EXACT_FIELD:
      VH_REQUIRE(value == 0 || value == 1, "vh_accel_set_option: exact field takes 0 or 1");
      a->exact_field = value;
      return VH_OK;
    case VH_OPT_FIELD_FP32:
      VH_REQUIRE(value == 0 || value == 1, "vh_accel_set_option: FP32 field products take 0 or 1");
      if (value != a->field_fp32) {
        a->field_fp32 = value;
        drop_graphs(a);
        ++a->config_version;
      }
      return VH_OK;
    case VH_OPT_POISSON_UNFUSED:
      VH_REQUIRE(value == 0 || value == 1, "vh_accel_set_option: unfused Poisson takes 0 or 1");
      if (value != a->poisson_unfused) {
        a->poisson_unfused = value;
        drop_graphs(a);
        ++a->config_version;
      }
      return VH_OK;
    case VH_OPT_STEP_GRAPH:
      VH_REQUIRE(value == 0 || value == 1, "vh_accel_set_option: step graph takes 0 or 1");
      a->step_graph = value;
      return VH_OK;
    case VH_OPT_UNIT_SLOTS:
      VH_REQUIRE(value >= VH_UNIT_MIN && value <= 4096 && (value & (value - 1)) == 0,
                 "vh_accel_set_option: unit slots %d not a power of two in [%d, 4096]", value,
                 VH_UNIT_MIN);
      if (value != a->unit) {
        a->unit = value;   // the next sort cuts the cells with it
        drop_graphs(a);    // the recorded steps carry the old one
        ++a->config_version;
      }
      return VH_OK;
    default:
      VH_REQUIRE(false, "vh_accel_set_option: unknown option %d", option);
  }
  return VH_EINVAL;
}

extern "C" int vh_accel_set_denoiser(vh_accel *a, vh_denoiser *d, double tol) {
  VH_REQUIRE(a != nullptr, "vh_accel_set_denoiser: null accelerator");
  VH_REQUIRE(std::isfinite(tol) && tol >= 0, "vh_accel_set_denoiser: tolerance must be >= 0");
  if (d) {
    VH_REQUIRE(vh_denoiser_context(d) == a->ctx,
               "vh_accel_set_denoiser: denoiser and accelerator on different contexts");
    VH_TRY(vh_denoiser_prepare(d, a->n));
  }
  a->denoiser = d;
  a->denoise_tol = tol;
  drop_graphs(a);    // the recorded steps carry the old pipeline
  ++a->config_version;
  return VH_OK;
}

// 48-byte PackedParticle records <-> the buffer's current and previous arrays
__global__ void k_records_in(const int32_t *__restrict__ rec, int32_t *__restrict__ P, int64_t n,
                             int64_t npad) {
  int32_t *prev = vh_prev_half(P, npad);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int2 *r = reinterpret_cast<const int2 *>(rec + 12 * i);
    int2 *c = reinterpret_cast<int2 *>(P + 6 * i), *q = reinterpret_cast<int2 *>(prev + 6 * i);
#pragma unroll
    for (int w = 0; w < 3; ++w) {
      c[w] = r[w];
      q[w] = r[3 + w];
    }
  }
}

// records of slots [offset, offset + n)
__global__ void k_records_in_at(const int32_t *__restrict__ rec, int32_t *__restrict__ P, int64_t n,
                                int64_t npad, int64_t offset) {
  int32_t *prev = vh_prev_half(P, npad);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int2 *r = reinterpret_cast<const int2 *>(rec + 12 * i);
    int2 *c = reinterpret_cast<int2 *>(P + 6 * (offset + i)),
         *q = reinterpret_cast<int2 *>(prev + 6 * (offset + i));
#pragma unroll
    for (int w = 0; w < 3; ++w) {
      c[w] = r[w];
      q[w] = r[3 + w];
    }
  }
}

__global__ void k_records_out(const int32_t *__restrict__ P, int32_t *__restrict__ rec, int64_t n,
                              int64_t npad) {
  const int32_t *prev = vh_prev_half(P, npad);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    int2 *r = reinterpret_cast<int2 *>(rec + 12 * i);
    const int2 *c = reinterpret_cast<const int2 *>(P + 6 * i),
               *q = reinterpret_cast<const int2 *>(prev + 6 * i);
#pragma unroll
    for (int w = 0; w < 3; ++w) {
      r[w] = c[w];
      r[3 + w] = q[w];
    }
  }
}

__global__ void k_interleave3(const float *__restrict__ f, float *__restrict__ out, int64_t n,
                              int64_t npad, int dir) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (dir == 0) {
      out[3 * i] = f[i];
      out[3 * i + 1] = f[npad + i];
      out[3 * i + 2] = f[2 * npad + i];
    } else {
      out[i] = f[3 * i];
      out[npad + i] = f[3 * i + 1];
      out[2 * npad + i] = f[3 * i + 2];
    }
  }
}

extern "C" int vh_particles_upload(vh_accel *a, const void *records) {
  VH_REQUIRE(a && records, "vh_particles_upload: null argument");
  VH_CHECK_HIP(hipSetDevice(a->ctx->device));
  int32_t *stage = a->parts[1 - a->cur];
  a->keys_fresh = 0;  // new records: the keys of the last push are stale
  a->rec_pending = 0;
  a->pushed_fused = 0;
  a->track_cloud = a->cloud_valid = 0;  // no host-held cloud order any more
  VH_CHECK_HIP(hipMemcpyAsync(stage, records, (size_t)a->npart * 48, hipMemcpyHostToDevice,
                              a->ctx->stream));
  hipLaunchKernelGGL(k_records_in, dim3(2048), dim3(256), 0, a->ctx->stream, stage,
                     a->parts[a->cur], a->npart, a->npad);
  VH_LAUNCH_CHECK();
  VH_CHECK_HIP(hipStreamSynchronize(a->ctx->stream));
  return VH_OK;
}

extern "C" int vh_particles_upload_range(vh_accel *a, int64_t offset, int64_t count,
                                         const void *records) {
  VH_REQUIRE(a && records, "vh_particles_upload_range: null argument");
  VH_REQUIRE(offset >= 0 && count >= 0 && offset + count <= a->npart,
             "vh_particles_upload_range: [%lld, %lld) outside the %lld particles",
             (long long)offset, (long long)(offset + count), (long long)a->npart);
  VH_CHECK_HIP(hipSetDevice(a->ctx->device));
  if (count == 0) return VH_OK;
  // the slots outside the range keep their records: settle a deferred move
  // first, and the new records replace slots [offset, offset + count)
  VH_TRY(vh_rec_settle(a));
  int32_t *stage = a->parts[1 - a->cur];
  a->keys_fresh = 0;
  a->pushed_fused = 0;
  a->track_cloud = a->cloud_valid = 0;
  VH_CHECK_HIP(hipMemcpyAsync(stage, records, (size_t)count * 48, hipMemcpyHostToDevice,
                              a->ctx->stream));
  int32_t *P = a->parts[a->cur];
  // the current and previous arrays of slots [offset, ...): shift both halves
  hipLaunchKernelGGL(k_records_in_at, dim3(2048), dim3(256), 0, a->ctx->stream, stage, P, count,
                     a->npad, offset);
  VH_LAUNCH_CHECK();
  VH_CHECK_HIP(hipStreamSynchronize(a->ctx->stream));
  return VH_OK;
}

extern "C" int vh_particles_download(vh_accel *a, void *records) {
  VH_REQUIRE(a && records, "vh_particles_download: null argument");
  VH_CHECK_HIP(hipSetDevice(a->ctx->device));
  VH_TRY(vh_rec_settle(a));
  int32_t *stage = a->parts[1 - a->cur];
  hipLaunchKernelGGL(k_records_out, dim3(2048), dim3(256), 0, a->ctx->stream, a->parts[a->cur],
                     stage, a->npart, a->npad);
  VH_LAUNCH_CHECK();
  VH_CHECK_HIP(hipMemcpyAsync(records, stage, (size_t)a->npart * 48, hipMemcpyDeviceToHost,
                              a->ctx->stream));
  VH_CHECK_HIP(hipStreamSynchronize(a->ctx->stream));
  return VH_OK;
}

extern "C" int vh_forces_download(vh_accel *a, float *host) {
  VH_REQUIRE(a && host, "vh_forces_download: null argument");
  VH_CHECK_HIP(hipSetDevice(a->ctx->device));
  float *stage = (float *)a->parts[1 - a->cur];
  hipLaunchKernelGGL(k_interleave3, dim3(2048), dim3(256), 0, a->ctx->stream, a->force, stage,
                     a->npart, a->npad, 0);
  VH_LAUNCH_CHECK();
  VH_CHECK_HIP(hipMemcpyAsync(host, stage, (size_t)a->npart * 12, hipMemcpyDeviceToHost,
                              a->ctx->stream));
  VH_CHECK_HIP(hipStreamSynchronize(a->ctx->stream));
  return VH_OK;
}

extern "C" int vh_forces_upload(vh_accel *a, const float *host) {
  VH_REQUIRE(a && host, "vh_forces_upload: null argument");
  VH_CHECK_HIP(hipSetDevice(a->ctx->device));
  float *stage = (float *)a->parts[1 - a->cur];
  VH_CHECK_HIP(hipMemcpyAsync(stage, host, (size_t)a->npart * 12, hipMemcpyHostToDevice,
                              a->ctx->stream));
  hipLaunchKernelGGL(k_interleave3, dim3(2048), dim3(256), 0, a->ctx->stream, stage, a->force,
                     a->npart, a->npad, 1);
  VH_LAUNCH_CHECK();
  VH_CHECK_HIP(hipStreamSynchronize(a->ctx->stream));
  return VH_OK;
}

// ---------------------------------------------------------------------------
// region timing
// ---------------------------------------------------------------------------
static hipEvent_t take_event(vh_context *c) {
  if (!c->prof_pool.empty()) {
    hipEvent_t e = c->prof_pool.back();
    c->prof_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

int vh_prof_begin(vh_context *c, int stage) {
  if (!((unsigned)c->prof_on >> stage & 1u)) return -1;
  vh_prof_rec r;
  r.stage = stage;
  r.start = take_event(c);
  r.stop = take_event(c);
  if (!r.start || !r.stop) return -1;
  hipEventRecord(r.start, c->stream);
  c->prof_open.push_back(r);
  return (int)c->prof_open.size() - 1;
}

void vh_prof_end(vh_context *c, int token) {
  if (token < 0 || token >= (int)c->prof_open.size()) return;
  hipEventRecord(c->prof_open[token].stop, c->stream);
}

extern "C" int vh_profile_enable(vh_context *ctx, int on) {
  VH_REQUIRE(ctx != nullptr, "vh_profile_enable: null context");
  ctx->prof_on = on;  // a mask of stages
  return VH_OK;
}

extern "C" int vh_profile_read(vh_context *ctx, double *ms, int64_t *count) {
  VH_REQUIRE(ctx && ms && count, "vh_profile_read: null argument");
  VH_CHECK_HIP(hipSetDevice(ctx->device));
  VH_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  for (auto &r : ctx->prof_open) {
    float t = 0.f;
    VH_CHECK_HIP(hipEventElapsedTime(&t, r.start, r.stop));
    if (r.stage >= 0 && r.stage < VH_ST_COUNT) {
      ms[r.stage] += t;
      count[r.stage] += 1;
    }
    ctx->prof_pool.push_back(r.start);
    ctx->prof_pool.push_back(r.stop);
  }
  ctx->prof_open.clear();
  return VH_OK;
}

struct vh_region {
  vh_context *c;
  int tok;
  vh_region(vh_context *ctx, int stage) : c(ctx), tok(vh_prof_begin(ctx, stage)) {}
  ~vh_region() { vh_prof_end(c, tok); }
};

// ---------------------------------------------------------------------------
// public wrappers of the step's stages
// ---------------------------------------------------------------------------
static int read_counters(vh_accel *a, int32_t *h) {
  VH_CHECK_HIP(hipMemcpyAsync(h, a->counters, sizeof(int32_t) * CNT_READ, hipMemcpyDeviceToHost,
                              a->ctx->stream));
  VH_CHECK_HIP(hipStreamSynchronize(a->ctx->stream));
  return VH_OK;
}

static int sort_timed(vh_accel *a) {
  vh_region r(a->ctx, VH_ST_SORT);
  VH_TRY(vh_sort_impl(a));
  a->sorted_once = 1;
  return VH_OK;
}

extern "C" int vh_sort(vh_accel *a, int32_t *n_occupied, int32_t *outcount) {
  VH_REQUIRE(a != nullptr, "vh_sort: null accelerator");
  VH_CHECK_HIP(hipSetDevice(a->ctx->device));
  VH_TRY(sort_timed(a));
  if (n_occupied || outcount) {
    int32_t h[CNT_READ];
    VH_TRY(read_counters(a, h));
    if (n_occupied) *n_occupied = h[CNT_NOCC];
    if (outcount) *outcount = h[CNT_OUTCOUNT];
  }
  return VH_OK;
}

__global__ void k_lists_export(const int32_t *__restrict__ occ, const int32_t *__restrict__ bounds,
                               const int32_t *__restrict__ outl, const int32_t *__restrict__ cnt,
                               int32_t *occ1, int32_t *bnd, int32_t *out1) {
  int nocc = cnt[CNT_NOCC], nout = cnt[CNT_OUTCOUNT];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i <= nocc || i < nout;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (i < nocc) occ1[i] = occ[i] + 1;
    if (i <= nocc) bnd[i] = bounds[i];
    if (i < nout) out1[i] = outl[i] + 1;
  }
}

extern "C" int vh_sort_lists(vh_accel *a, int32_t *occupied, int32_t *bounds, int32_t *outlist) {
  VH_REQUIRE(a && occupied && bounds && outlist, "vh_sort_lists: null argument");
  VH_REQUIRE(a->sorted_once, "vh_sort_lists: no sort has run");
  VH_CHECK_HIP(hipSetDevice(a->ctx->device));
  int32_t h[CNT_READ];
  VH_TRY(read_counters(a, h));
  int nocc = h[CNT_NOCC], nout = h[CNT_OUTCOUNT];
  int32_t *d = (int32_t *)a->parts[1 - a->cur];
  int32_t *d_occ = d, *d_bnd = d + (nocc + 1), *d_out = d + 2 * (nocc + 1) + 1;
  hipLaunchKernelGGL(k_lists_export, dim3(1024), dim3(256), 0, a->ctx->stream, a->occupied,
                     a->bounds, a->outlist, a->counters, d_occ, d_bnd, d_out);
  VH_LAUNCH_CHECK();
  if (nocc) VH_CHECK_HIP(hipMemcpyAsync(occupied, d_occ, 4 * (size_t)nocc, hipMemcpyDeviceToHost, a->ctx->stream));
  VH_CHECK_HIP(hipMemcpyAsync(bounds, d_bnd, 4 * (size_t)(nocc + 1), hipMemcpyDeviceToHost, a->ctx->stream));
  if (nout) VH_CHECK_HIP(hipMemcpyAsync(outlist, d_out, 4 * (size_t)nout, hipMemcpyDeviceToHost, a->ctx->stream));
  VH_CHECK_HIP(hipStreamSynchronize(a->ctx->stream));
  return VH_OK;
}

// columns of src/kernels.jl:754-779, in the device's current order
__global__ void k_columns(const int32_t *__restrict__ P, int64_t npad, int64_t n, float x0,
                          float h, float lo, float hi, float spacing, float nbdt, int ncol,
                          int32_t *__restrict__ cols, int32_t *cnt) {
  int local = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const vh_pos r = vh_load_pos(P, i);
    float half = spacing * 0.5f;
    bool inside = true;
    for (int d = 0; d < 3; ++d) {
      float p = x0 + (float)(r.k[d] - 1) * h + r.d[d];
      inside = inside && (lo <= p) && (p <= hi);
    }
    if (inside) {
      for (int d = 0; d < 3; ++d) {
        float t = (r.d[d] + half) / spacing * nbdt + 0.5f;
        int c = (int)floorf(t) + 1;
        c = c < 1 ? 1 : (c > ncol ? ncol : c);
        cols[3 * i + d] = c;
      }
    } else {
      cols[3 * i] = cols[3 * i + 1] = cols[3 * i + 2] = 1;
      ++local;
    }
  }
  local = vh_wave_sum(local);
  if (vh_lane() == 0 && local) atomicAdd(cnt, local);
}

extern "C" int vh_columns(vh_accel *a, int32_t *cols, int32_t *nout) {
  VH_REQUIRE(a && cols, "vh_columns: null argument");
  VH_CHECK_HIP(hipSetDevice(a->ctx->device));
  VH_TRY(vh_rec_settle(a));
  int32_t *dcols = (int32_t *)a->parts[1 - a->cur];
  VH_TRY(vh_fill_i32(a->ctx, a->counters + CNT_COLUMNS_OUT, 1, 0));
  hipLaunchKernelGGL(k_columns, dim3(2048), dim3(256), 0, a->ctx->stream, a->parts[a->cur],
                     a->npad, a->npart, (float)a->x0, (float)a->h, a->lo, a->hi,
                     a->spacing, (float)a->nbdt, a->ncol, dcols, a->counters + CNT_COLUMNS_OUT);
  VH_LAUNCH_CHECK();
  VH_CHECK_HIP(hipMemcpyAsync(cols, dcols, (size_t)a->npart * 12, hipMemcpyDeviceToHost, a->ctx->stream));
  int32_t hn = 0;
  VH_CHECK_HIP(hipMemcpyAsync(&hn, a->counters + CNT_COLUMNS_OUT, 4, hipMemcpyDeviceToHost, a->ctx->stream));
  VH_CHECK_HIP(hipStreamSynchronize(a->ctx->stream));
  if (nout) *nout = hn;
  return VH_OK;
}

extern "C" int vh_deposit_smoothed(vh_accel *a, vh_mesh *fine, double charge, int32_t *nout) {
  VH_REQUIRE(a && fine, "vh_deposit_smoothed: null argument");
  VH_REQUIRE(fine->n == a->n, "vh_deposit_smoothed: rho must cover the whole collocation grid "
             "(mesh n=%d, accelerator n=%d)", fine->n, a->n);
  VH_REQUIRE(fine->ctx == a->ctx, "vh_deposit_smoothed: mesh and accelerator on different contexts");
  VH_REQUIRE(std::isfinite(charge), "vh_deposit_smoothed: charge must be finite");
  VH_CHECK_HIP(hipSetDevice(a->ctx->device));
  VH_TRY(sort_timed(a));
  {
    vh_region r(a->ctx, VH_ST_DEPOSIT_FINE);
    VH_TRY(vh_deposit_fine_impl(a, fine, charge));
  }
  if (nout) {
    int32_t h[CNT_READ];
    VH_TRY(read_counters(a, h));
    *nout = h[a->nout_idx];
  }
  return VH_OK;
}

extern "C" int vh_deposit_cic(vh_accel *a, vh_mesh *coarse, double charge) {
  VH_REQUIRE(a && coarse, "vh_deposit_cic: null argument");
  VH_REQUIRE(coarse->ctx == a->ctx, "vh_deposit_cic: mesh and accelerator on different contexts");
  VH_REQUIRE(std::isfinite(charge), "vh_deposit_cic: charge must be finite");
  if (!a->sorted_once) {
    vh_set_error("vh_deposit_cic: the cloud has not been sorted");
    return VH_ESTATE;
  }
  VH_CHECK_HIP(hipSetDevice(a->ctx->device));
  vh_region r(a->ctx, VH_ST_DEPOSIT_CIC);
  VH_TRY(vh_deposit_cic_impl(a, coarse, charge));
  return VH_OK;
}

extern "C" int vh_poisson(vh_mesh *m) {
  VH_REQUIRE(m != nullptr, "vh_poisson: null mesh");
  VH_CHECK_HIP(hipSetDevice(m->ctx->device));
  vh_region r(m->ctx, VH_ST_POISSON);
  VH_TRY(vh_poisson_single_impl(m));
  return VH_OK;
}

extern "C" int vh_poisson_nested(vh_mesh *fine, vh_mesh *coarse) {
  VH_REQUIRE(fine && coarse, "vh_poisson_nested: null mesh");
  VH_REQUIRE(fine->ctx == coarse->ctx, "vh_poisson_nested: meshes on different contexts");
  VH_CHECK_HIP(hipSetDevice(fine->ctx->device));
  vh_region r(fine->ctx, VH_ST_POISSON);
  VH_TRY(vh_poisson_nested_impl(fine, coarse));
  return VH_OK;
}

extern "C" int vh_spline_coefficients(vh_mesh *m) {
  VH_REQUIRE(m != nullptr, "vh_spline_coefficients: null mesh");
  VH_CHECK_HIP(hipSetDevice(m->ctx->device));
  vh_region r(m->ctx, VH_ST_SPLINE);
  VH_TRY(vh_spline_coefficients_impl(m, m->phi, m->csol));
  return VH_OK;
}

extern "C" int vh_effective_potential(vh_mesh *m, double nions, double radius) {
  VH_REQUIRE(m != nullptr, "vh_effective_potential: null mesh");
  VH_REQUIRE(radius > 0, "vh_effective_potential: jellium radius must be positive");
  VH_CHECK_HIP(hipSetDevice(m->ctx->device));
  vh_region r(m->ctx, VH_ST_MEANFIELD);
  VH_TRY(vh_effective_potential_impl(m, nions, radius));
  return VH_OK;
}

extern "C" int vh_potential_coefficients(vh_mesh *m, double nions, double radius) {
  VH_REQUIRE(m != nullptr, "vh_potential_coefficients: null mesh");
  VH_REQUIRE(radius > 0, "vh_potential_coefficients: jellium radius must be positive");
  VH_CHECK_HIP(hipSetDevice(m->ctx->device));
  vh_region r(m->ctx, VH_ST_SPLINE);
  VH_TRY(vh_potential_coefficients_impl(m, nions, radius));
  return VH_OK;
}

extern "C" int vh_total_charge(vh_mesh *m, double *q) {
  VH_REQUIRE(m && q, "vh_total_charge: null argument");
  VH_CHECK_HIP(hipSetDevice(m->ctx->device));
  VH_TRY(vh_total_charge_impl(m, m->rho, m->scal + 20));
  VH_CHECK_HIP(hipMemcpyAsync(q, m->scal + 20, 8, hipMemcpyDeviceToHost, m->ctx->stream));
  VH_CHECK_HIP(hipStreamSynchronize(m->ctx->stream));
  return VH_OK;
}

int vh_multipole_impl(vh_mesh *m);
extern "C" int vh_multipole(vh_mesh *m, double out10[10]) {
  VH_REQUIRE(m && out10, "vh_multipole: null argument");
  VH_CHECK_HIP(hipSetDevice(m->ctx->device));
  VH_TRY(vh_multipole_impl(m));
  VH_CHECK_HIP(hipMemcpyAsync(out10, m->scal + 10, 80, hipMemcpyDeviceToHost, m->ctx->stream));
  VH_CHECK_HIP(hipStreamSynchronize(m->ctx->stream));
  return VH_OK;
}

extern "C" int vh_forces(vh_accel *a, vh_mesh *fine, vh_mesh *coarse, double weight,
                         const double *proj, int soft_mode, double ball_radius, double red[4],
                         int32_t *nout) {
  VH_REQUIRE(a && fine && coarse, "vh_forces: null argument");
  VH_REQUIRE(fine->n == a->n, "vh_forces: fine csol has n=%d, accelerator n=%d", fine->n, a->n);
  VH_REQUIRE(fine->ctx == a->ctx && coarse->ctx == a->ctx, "vh_forces: objects on different contexts");
  VH_REQUIRE(soft_mode == VH_SOFT_DEVICE || soft_mode == VH_SOFT_BALL,
             "vh_forces: unknown softening mode %d", soft_mode);
  VH_REQUIRE(std::isfinite(weight), "vh_forces: weight must be finite");
  if (proj) {
    VH_REQUIRE(proj[4] > 0, "vh_forces: projectile scale must be positive");
    VH_REQUIRE(std::isfinite(proj[0]) && std::isfinite(proj[1]) && std::isfinite(proj[2]) &&
               std::isfinite(proj[3]), "vh_forces: projectile state must be finite");
  }
  if (!a->sorted_once) {
    vh_set_error("vh_forces: the cloud has not been sorted");
    return VH_ESTATE;
  }
  VH_CHECK_HIP(hipSetDevice(a->ctx->device));
  {
    vh_region r(a->ctx, VH_ST_FIELD);
    VH_TRY(vh_forces_impl(a, fine, coarse, weight, proj, soft_mode, ball_radius, nullptr));
  }
  if (red) VH_CHECK_HIP(hipMemcpyAsync(red, a->dscal + 8, 32, hipMemcpyDeviceToHost, a->ctx->stream));
  if (nout) VH_CHECK_HIP(hipMemcpyAsync(nout, a->counters + CNT_OUTCOUNT, 4, hipMemcpyDeviceToHost, a->ctx->stream));
  if (red || nout) VH_CHECK_HIP(hipStreamSynchronize(a->ctx->stream));
  return VH_OK;
}

extern "C" int vh_projectile_set(vh_accel *a, const double state[9], double mass, double charge,
                                 double scale, int soft_mode, double ball_radius, double jel_nions,
                                 double jel_radius) {
  VH_REQUIRE(a && state, "vh_projectile_set: null argument");
  VH_REQUIRE(soft_mode == VH_SOFT_DEVICE || soft_mode == VH_SOFT_BALL,
             "vh_projectile_set: unknown softening mode %d", soft_mode);
  VH_REQUIRE(mass > 0 && scale > 0 && jel_radius > 0 && std::isfinite(charge) &&
                 std::isfinite(jel_nions), "vh_projectile_set: bad projectile or jellium constants");
  for (int i = 0; i < 9; ++i) VH_REQUIRE(std::isfinite(state[i]), "vh_projectile_set: state must be finite");
  VH_CHECK_HIP(hipSetDevice(a->ctx->device));
  if (!a->dproj) VH_TRY(vh_alloc((void **)&a->dproj, sizeof(double) * 16));
  double h[16] = {0};
  for (int i = 0; i < 9; ++i) h[i] = state[i];
  h[9] = mass;
  h[10] = charge;
  h[11] = jel_nions;
  h[12] = jel_radius;
  VH_CHECK_HIP(hipMemcpyAsync(a->dproj, h, sizeof h, hipMemcpyHostToDevice, a->ctx->stream));
  VH_CHECK_HIP(hipStreamSynchronize(a->ctx->stream));
  a->proj_charge = charge;
  a->proj_scale = scale;
  a->proj_soft_mode = soft_mode;
  a->proj_ball_radius = ball_radius;
  a->has_dproj = 1;
  return VH_OK;
}

extern "C" int vh_projectile_get(vh_accel *a, double state[9]) {
  VH_REQUIRE(a && state, "vh_projectile_get: null argument");
  if (!a->has_dproj) {
    vh_set_error("vh_projectile_get: no projectile on the device");
    return VH_ESTATE;
  }
  VH_CHECK_HIP(hipSetDevice(a->ctx->device));
  VH_CHECK_HIP(hipMemcpyAsync(state, a->dproj, sizeof(double) * 9, hipMemcpyDeviceToHost, a->ctx->stream));
  VH_CHECK_HIP(hipStreamSynchronize(a->ctx->stream));
  return VH_OK;
}

extern "C" int vh_forces_resident(vh_accel *a, vh_mesh *fine, vh_mesh *coarse, double weight) {
  VH_REQUIRE(a && fine && coarse, "vh_forces_resident: null argument");
  VH_REQUIRE(fine->n == a->n, "vh_forces_resident: fine csol has n=%d, accelerator n=%d", fine->n, a->n);
  VH_REQUIRE(fine->ctx == a->ctx && coarse->ctx == a->ctx,
             "vh_forces_resident: objects on different contexts");
  VH_REQUIRE(std::isfinite(weight), "vh_forces_resident: weight must be finite");
  if (!a->has_dproj || !a->sorted_once) {
    vh_set_error("vh_forces_resident: %s", a->has_dproj ? "the cloud has not been sorted"
                                                         : "no projectile on the device");
    return VH_ESTATE;
  }
  VH_CHECK_HIP(hipSetDevice(a->ctx->device));
  // the position entries are read on the device; charge and scale from here
  const double proj[5] = {0.0, 0.0, 0.0, a->proj_charge, a->proj_scale};
  vh_region r(a->ctx, VH_ST_FIELD);
  return vh_forces_impl(a, fine, coarse, weight, proj, a->proj_soft_mode, a->proj_ball_radius,
                        a->dproj);
}

extern "C" int vh_projectile_step(vh_accel *a, double dt) {
  VH_REQUIRE(a != nullptr, "vh_projectile_step: null accelerator");
  VH_REQUIRE(dt > 0 && std::isfinite(dt), "vh_projectile_step: dt must be positive");
  if (!a->has_dproj) {
    vh_set_error("vh_projectile_step: no projectile on the device");
    return VH_ESTATE;
  }
  VH_CHECK_HIP(hipSetDevice(a->ctx->device));
  return vh_projectile_step_impl(a, dt);
}

static void drop_graphs(vh_accel *a) {
  for (int b = 0; b < 2; ++b)
    if (a->graph[b]) {
      hipGraphExecDestroy(a->graph[b]);
      a->graph[b] = nullptr;
    }
}

static int step_eager(vh_accel *a, vh_mesh *fine, vh_mesh *coarse, double weight, double dt,
                      double mass, double rcmax, double jel_nions, double jel_radius);

static vh_step_state host_state(const vh_accel *a) {
  vh_step_state h;
  h.cur = a->cur;
  h.keys_fresh = a->keys_fresh;
  h.hist_fresh = a->hist_fresh;
  h.rec_pending = a->rec_pending;
  h.pushed_fused = a->pushed_fused;
  h.nout_idx = a->nout_idx;
  h.rec_perm = a->rec_perm;
  return h;
}

static void set_host_state(vh_accel *a, const vh_step_state &h) {
  a->cur = h.cur;
  a->keys_fresh = h.keys_fresh;
  a->hist_fresh = h.hist_fresh;
  a->rec_pending = h.rec_pending;
  a->pushed_fused = h.pushed_fused;
  a->nout_idx = h.nout_idx;
  a->rec_perm = h.rec_perm;
}
// the field and fallback kernels push (the matrix-core field kernel)
static bool step_fuses_push(const vh_accel *a) { return !a->exact_field; }

extern "C" int vh_step(vh_accel *a, vh_mesh *fine, vh_mesh *coarse, double weight, double dt,
                       double mass, double rcmax, double jel_nions, double jel_radius) {
  VH_REQUIRE(a && fine && coarse, "vh_step: null argument");
  VH_REQUIRE(std::isfinite(weight) && dt > 0 && mass > 0 && jel_radius > 0,
             "vh_step: weight finite, dt, mass and jellium radius positive");
  if (!a->step_graph || a->ctx->prof_on)
    return step_eager(a, fine, coarse, weight, dt, mass, rcmax, jel_nions, jel_radius);
  // the launch sequence depends only on these and on which buffer is current
  const double key[VH_GRAPH_KEY] = {weight, dt, mass, rcmax, jel_nions, jel_radius,
                                    (double)(uintptr_t)fine, (double)(uintptr_t)coarse,
                                    (double)a->has_dproj, (double)a->exact_field, a->proj_charge,
                                    a->proj_scale, (double)a->proj_soft_mode, a->proj_ball_radius,
                                    (double)(a->keys_fresh + 2 * a->hist_fresh + 4 * a->rec_pending +
                                             8 * a->pushed_fused)};
  if (memcmp(key, a->graph_key, sizeof key) != 0) {
    drop_graphs(a);
    memcpy(a->graph_key, key, sizeof key);
  }
  VH_CHECK_HIP(hipSetDevice(a->ctx->device));
  hipStream_t st = a->ctx->stream;
  const int b = a->cur;
  if (!a->graph[b]) {
    VH_CHECK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int r = step_eager(a, fine, coarse, weight, dt, mass, rcmax, jel_nions, jel_radius);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (r != VH_OK) {
      if (g) hipGraphDestroy(g);
      return r;
    }
    if (e != hipSuccess) {
      vh_set_error("vh_step: graph capture: %s", hipGetErrorString(e));
      return VH_EHIP;
    }
    const hipError_t ei = hipGraphInstantiate(&a->graph[b], g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (ei != hipSuccess) {
      a->graph[b] = nullptr;
      vh_set_error("vh_step: graph instantiation: %s", hipGetErrorString(ei));
      return VH_EHIP;
    }
    // recording left the host state of a step behind (current buffer,
    // fresh keys, ...), kept for the replays; the device work runs now
    a->graph_after[b] = host_state(a);
    VH_CHECK_HIP(hipGraphLaunch(a->graph[b], st));
    return VH_OK;
  }
  VH_CHECK_HIP(hipGraphLaunch(a->graph[b], st));
  set_host_state(a, a->graph_after[b]);  // what the recorded step did on the host
  return VH_OK;
}

// ---------------------------------------------------------------------------
// autotuning of the work-unit size: real steps of the simulation, timed with
// events on the accelerator's stream
// ---------------------------------------------------------------------------
extern "C" int vh_autotune_unit(vh_accel *a, vh_mesh *fine, vh_mesh *coarse, double weight,
                                double dt, double mass, double rcmax, double jel_nions,
                                double jel_radius, const int *candidates, int ncand, int steps,
                                int *chosen, double *ms) {
  VH_REQUIRE(a && fine && coarse && candidates && chosen, "vh_autotune_unit: null argument");
  VH_REQUIRE(ncand >= 1 && steps >= 1, "vh_autotune_unit: %d candidates, %d steps", ncand, steps);
  for (int i = 0; i < ncand; ++i)
    VH_REQUIRE(candidates[i] >= VH_UNIT_MIN && candidates[i] <= 4096 &&
                   (candidates[i] & (candidates[i] - 1)) == 0,
               "vh_autotune_unit: candidate %d not a power of two in [%d, 4096]", candidates[i],
               VH_UNIT_MIN);
  VH_CHECK_HIP(hipSetDevice(a->ctx->device));
  hipEvent_t e0, e1;
  VH_CHECK_HIP(hipEventCreate(&e0));
  VH_CHECK_HIP(hipEventCreate(&e1));
  int rc = VH_OK, best = -1;
  float best_ms = 0.f;
  for (int i = 0; i < ncand && rc == VH_OK; ++i) {
    rc = vh_accel_set_option(a, VH_OPT_UNIT_SLOTS, candidates[i]);
    // one untimed step: its sort cuts the cells with the candidate (and
    // records the step's graph when the accelerator replays graphs)
    if (rc == VH_OK)
      rc = vh_step(a, fine, coarse, weight, dt, mass, rcmax, jel_nions, jel_radius);
    if (rc != VH_OK) break;
    hipEventRecord(e0, a->ctx->stream);
    for (int k = 0; k < steps && rc == VH_OK; ++k)
      rc = vh_step(a, fine, coarse, weight, dt, mass, rcmax, jel_nions, jel_radius);
    hipEventRecord(e1, a->ctx->stream);
    float t = 0.f;
    if (rc == VH_OK && hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&t, e0, e1) == hipSuccess) {
      t /= steps;
      if (ms) ms[i] = t;
      if (best < 0 || t < best_ms) {
        best = i;
        best_ms = t;
      }
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (rc != VH_OK) return rc;
  VH_REQUIRE(best >= 0, "vh_autotune_unit: no candidate could be timed");
  *chosen = candidates[best];
  return vh_accel_set_option(a, VH_OPT_UNIT_SLOTS, candidates[best]);
}

// ---------------------------------------------------------------------------
// batches: independent simulations stepped as one HIP graph
// ---------------------------------------------------------------------------
struct vh_batch_entry {
  std::vector<int> key;
  hipGraphExec_t exec = nullptr;
  std::vector<vh_step_state> after;
};

struct vh_batch {
  int device = 0;
  int use_graph = 1;
  std::vector<vh_accel *> accs;
  std::vector<vh_mesh *> fines, coarses;
  std::vector<double> params;          // 6 per member: weight dt mass rcmax nions radius
  hipStream_t stream = nullptr;
  hipEvent_t fork = nullptr, done = nullptr;
  std::vector<hipEvent_t> join, ready;
  std::vector<vh_batch_entry> cache;   // at most VH_BATCH_GRAPHS recorded steps
};
#define VH_BATCH_GRAPHS 4

static void batch_free(vh_batch *b) {
  if (!b) return;
  hipSetDevice(b->device);
  if (b->stream) hipStreamSynchronize(b->stream);
  for (auto &e : b->cache)
    if (e.exec) hipGraphExecDestroy(e.exec);
  for (hipEvent_t e : b->join) hipEventDestroy(e);
  for (hipEvent_t e : b->ready) hipEventDestroy(e);
  if (b->fork) hipEventDestroy(b->fork);
  if (b->done) hipEventDestroy(b->done);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
}

extern "C" int vh_batch_create(int count, vh_accel *const *accs, vh_mesh *const *fines,
                               vh_mesh *const *coarses, const double *params, int use_graph,
                               vh_batch **out) {
  VH_REQUIRE(count >= 1 && accs && fines && coarses && params && out,
             "vh_batch_create: null argument or empty batch");
  *out = nullptr;
  const int dev = accs[0]->ctx->device;
  for (int i = 0; i < count; ++i) {
    VH_REQUIRE(accs[i] && fines[i] && coarses[i], "vh_batch_create: null member %d", i);
    VH_REQUIRE(accs[i]->ctx->device == dev, "vh_batch_create: members on different devices");
    VH_REQUIRE(fines[i]->ctx == accs[i]->ctx && coarses[i]->ctx == accs[i]->ctx &&
                   fines[i]->n == accs[i]->n,
               "vh_batch_create: member %d's meshes do not belong to its accelerator", i);
    for (int j = 0; j < i; ++j)
      VH_REQUIRE(accs[j]->ctx != accs[i]->ctx,
                 "vh_batch_create: members %d and %d share a context (one stream each)", j, i);
    const double *p = params + 6 * i;
    VH_REQUIRE(std::isfinite(p[0]) && p[1] > 0 && p[2] > 0 && p[5] > 0,
               "vh_batch_create: member %d: weight finite, dt, mass and jellium radius positive", i);
  }
  VH_CHECK_HIP(hipSetDevice(dev));
  vh_batch *b = new vh_batch();
  b->device = dev;
  b->use_graph = use_graph ? 1 : 0;
  b->accs.assign(accs, accs + count);
  b->fines.assign(fines, fines + count);
  b->coarses.assign(coarses, coarses + count);
  b->params.assign(params, params + 6 * (size_t)count);
  bool ok = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess &&
            hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) == hipSuccess &&
            hipEventCreateWithFlags(&b->done, hipEventDisableTiming) == hipSuccess;
  for (int i = 0; i < count && ok; ++i) {
    hipEvent_t j = nullptr, r = nullptr;
    ok = hipEventCreateWithFlags(&j, hipEventDisableTiming) == hipSuccess &&
         hipEventCreateWithFlags(&r, hipEventDisableTiming) == hipSuccess;
    b->join.push_back(j);
    b->ready.push_back(r);
  }
  if (!ok) {
    batch_free(b);
    vh_set_error("vh_batch_create: could not create the batch's stream and events");
    return VH_EHIP;
  }
  *out = b;
  return VH_OK;
}

extern "C" int vh_batch_destroy(vh_batch *b) {
  batch_free(b);
  return VH_OK;
}

static int batch_member_step(vh_batch *b, size_t i) {
  const double *p = &b->params[6 * i];
  return step_eager(b->accs[i], b->fines[i], b->coarses[i], p[0], p[1], p[2], p[3], p[4], p[5]);
}

extern "C" int vh_batch_step(vh_batch *b) {
  VH_REQUIRE(b != nullptr, "vh_batch_step: null batch");
  VH_CHECK_HIP(hipSetDevice(b->device));
  const size_t n = b->accs.size();
  bool profiling = false;
  for (vh_accel *a : b->accs) profiling = profiling || a->ctx->prof_on;
  if (!b->use_graph || profiling) {
    // every member on its own streams: their launches interleave on the device
    for (size_t i = 0; i < n; ++i) VH_TRY(batch_member_step(b, i));
    return VH_OK;
  }
  // the recorded steps depend on each member's host state (current buffer,
  // fresh keys, records in place) and options
  std::vector<int> key;
  key.reserve(8 * n);
  for (vh_accel *a : b->accs) {
    key.push_back(a->cur);
    key.push_back(a->keys_fresh + 2 * a->hist_fresh + 4 * a->rec_pending + 8 * a->pushed_fused);
    key.push_back(a->config_version);
    key.push_back(a->has_dproj + 2 * a->exact_field);
    key.push_back(a->nout_idx);
  }
  // the batch stream follows each member's earlier work
  for (size_t i = 0; i < n; ++i) {
    VH_CHECK_HIP(hipEventRecord(b->ready[i], b->accs[i]->ctx->stream));
    VH_CHECK_HIP(hipStreamWaitEvent(b->stream, b->ready[i], 0));
  }
  vh_batch_entry *e = nullptr;
  for (auto &c : b->cache)
    if (c.key == key) e = &c;
  if (!e) {
    if (b->cache.size() >= VH_BATCH_GRAPHS) {
      hipStreamSynchronize(b->stream);
      hipGraphExecDestroy(b->cache.front().exec);
      b->cache.erase(b->cache.begin());
    }
    // record: fork from the batch stream into every member's streams, each
    // member's step, and join back
    VH_CHECK_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = hipEventRecord(b->fork, b->stream) == hipSuccess ? VH_OK : VH_EHIP;
    for (size_t i = 0; i < n && rc == VH_OK; ++i) {
      hipStream_t ms = b->accs[i]->ctx->stream;
      if (hipStreamWaitEvent(ms, b->fork, 0) != hipSuccess) { rc = VH_EHIP; break; }
      rc = batch_member_step(b, i);
      if (rc == VH_OK && (hipEventRecord(b->join[i], ms) != hipSuccess ||
                          hipStreamWaitEvent(b->stream, b->join[i], 0) != hipSuccess))
        rc = VH_EHIP;
    }
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(b->stream, &g);
    if (rc != VH_OK || ec != hipSuccess) {
      if (g) hipGraphDestroy(g);
      if (rc == VH_OK) vh_set_error("vh_batch_step: graph capture: %s", hipGetErrorString(ec));
      return rc != VH_OK ? rc : VH_EHIP;
    }
    vh_batch_entry ne;
    ne.key = key;
    const hipError_t ei = hipGraphInstantiate(&ne.exec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (ei != hipSuccess) {
      vh_set_error("vh_batch_step: graph instantiation: %s", hipGetErrorString(ei));
      return VH_EHIP;
    }
    for (vh_accel *a : b->accs) ne.after.push_back(host_state(a));
    b->cache.push_back(ne);
    e = &b->cache.back();
  } else {
    for (size_t i = 0; i < n; ++i) set_host_state(b->accs[i], e->after[i]);
  }
  VH_CHECK_HIP(hipGraphLaunch(e->exec, b->stream));
  // every member's later work follows the batch's step
  VH_CHECK_HIP(hipEventRecord(b->done, b->stream));
  for (vh_accel *a : b->accs) VH_CHECK_HIP(hipStreamWaitEvent(a->ctx->stream, b->done, 0));
  return VH_OK;
}

extern "C" int vh_batch_sync(vh_batch *b) {
  VH_REQUIRE(b != nullptr, "vh_batch_sync: null batch");
  VH_CHECK_HIP(hipSetDevice(b->device));
  VH_CHECK_HIP(hipStreamSynchronize(b->stream));
  for (vh_accel *a : b->accs) VH_CHECK_HIP(hipStreamSynchronize(a->ctx->stream));
  return VH_OK;
}

static int step_eager(vh_accel *a, vh_mesh *fine, vh_mesh *coarse, double weight, double dt,
                      double mass, double rcmax, double jel_nions, double jel_radius) {
  // _update_forces_resident! (src/simulation.jl:472-531), then step!
  vh_context *c = a->ctx;
  // the push's Verlet factor, and whether the fallback ran beside the fine level
  const double push_factor = (double)(float)(dt * dt / mass);
  bool early_fallback = false;
  const int staged = (1 << VH_ST_DEPOSIT_FINE) | (1 << VH_ST_DEPOSIT_CIC) | (1 << VH_ST_POISSON) |
                     (1 << VH_ST_SPLINE) | (1 << VH_ST_MEANFIELD);
  if (c->prof_on & staged) {
    // timing of these stages: the stages one after the other
    VH_TRY(vh_deposit_smoothed(a, fine, weight, nullptr));
    if (a->denoiser) VH_TRY(vh_denoiser_apply_impl(a->denoiser, fine, a->denoise_tol));
    VH_TRY(vh_deposit_cic(a, coarse, weight));
    VH_TRY(vh_poisson_nested(fine, coarse));
    VH_TRY(vh_potential_coefficients(fine, jel_nions, jel_radius));
    VH_TRY(vh_effective_potential(coarse, jel_nions, jel_radius));
  } else {
    VH_REQUIRE(fine->n == a->n && fine->ctx == c && coarse->ctx == c && jel_radius > 0,
               "vh_step: meshes do not match the accelerator, or jellium radius not positive");
    VH_TRY(sort_timed(a));
    // Chains that share no buffer run side by side on the two streams: on
    // the side stream the CIC, the coarse level (solve, coefficients), the
    // fine boundary read from the coarse coefficients and then the coarse
    // mean field added to them; on the main stream the fine deposit, then,
    // once the boundary is there, the fine solve and coefficients.  The
    // rotating products and small kernels of the side chain are
    // latency-bound launches that fill room the main chain leaves.  (With
    // the CIC on the main stream before the fork, the fine deposit starved
    // a product of the coarse solve: 20 us slower a step.)  The host runs
    // ahead of the device until a stream's queue is full and then submits at
    // the device's pace, so launches go in the order they are wanted: the
    // fine deposit, the whole side chain, then the fine level, which waits
    // for the boundary anyway (submitted after the fine level, the coarse
    // mean field started 90 us late).  A stream wait binds to the event's
    // last record at the time of the call, so the fork event serves again
    // for the second join.
    // The solve by linearity (vh_poisson_fused_*): the density part of each
    // level before its boundary is known, three products after it, and phi's
    // interior never formed; off with VH_OPT_POISSON_UNFUSED (and on meshes
    // in split BF16 or of more than VH_FUSED_MAXN points a side), where the
    // step solves as the entry points do.
    const bool fused = !a->poisson_unfused && !fine->gemm_bf16x3 && !coarse->gemm_bf16x3 &&
                       fine->n <= VH_FUSED_MAXN && coarse->n <= VH_FUSED_MAXN;
    early_fallback = fused && step_fuses_push(a) && a->sorted_once;
    if (fused) {
      // On the side stream the CIC, the coarse level, the fine boundary and
      // then the coarse mean field; on the main stream the fine deposit, the
      // fine mean field and the density part of the fine solve, which fill
      // its wait for the boundary, then the fine level.  (Measured against
      // other arrangements: the coarse mean field kept apart on the main
      // stream during the wait, 1.483 ms a step; on the side stream after
      // the boundary and summed on the main stream, 1.481; after the fine
      // level on the main stream, 1.51; this one 1.475.  A launch queued
      // behind a wait for the other, busy, stream started tens of us late.)
      // Outside a graph capture the order between the streams is kept by
      // device flags (vh_stream_signal / vh_stream_wait), every wait queued
      // after its signal; in a capture by events.
      hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
      VH_CHECK_HIP(hipStreamIsCapturing(c->stream, &cs));
      const bool flags = cs == hipStreamCaptureStatusNone && vh_stream_flags_usable(c);
      const uint32_t seq = flags ? ++c->flag_seq : 0;
      // The fork's flag is stored as the fine deposit's first kernel starts
      // and the join's as the coarse mean field's does (the work before
      // them on their streams is then complete).  The side stream's wait is
      // queued first, so that it is spinning on the device before the sort
      // ends (a launch on an idle stream behind a busy one is taken up late);
      // should the deposit's launch fail on the host, that wait ends at its
      // time limit and counts a timeout.
      if (flags) {
        VH_TRY(vh_stream_wait(c, c->side, VH_FLAG_FORK, seq));
      } else {
        VH_CHECK_HIP(hipEventRecord(c->fork_ev, c->stream));
        VH_CHECK_HIP(hipStreamWaitEvent(c->side, c->fork_ev, 0));
      }
      VH_TRY(vh_deposit_fine_impl(a, fine, weight,
                                  flags ? vh_flag_sig(c, VH_FLAG_FORK, seq) : vh_sig{nullptr, 0}));
      if (a->denoiser) VH_TRY(vh_denoiser_apply_impl(a->denoiser, fine, a->denoise_tol));
      VH_TRY(vh_mean_field_coefficients_impl(fine, jel_nions, jel_radius));
      VH_TRY(vh_poisson_fused_prepare_impl(fine));
      {
        vh_on_side on(c);
        // the coarse density formed in the multipole's moments pass
        VH_TRY(vh_deposit_cic_impl(a, coarse, weight, true));
        VH_TRY(vh_poisson_multipole_boundary_impl(coarse));
        VH_TRY(vh_poisson_fused_prepare_impl(coarse));
        VH_TRY(vh_poisson_fused_finish_impl(coarse, false));
        VH_TRY(vh_poisson_fine_boundary_impl(fine, coarse));
        if (!flags) VH_CHECK_HIP(hipEventRecord(c->join_ev, c->side));
        VH_TRY(vh_effective_potential_impl(coarse, jel_nions, jel_radius,
                                           flags ? vh_flag_sig(c, VH_FLAG_JOIN, seq) : vh_sig{nullptr, 0}));
        // the out-of-stencil fallback (with the push) needs only the coarse
        // coefficients and the sort: here, beside the fine level, rather
        // than after the field kernel
        if (early_fallback) {
          const double proj[5] = {0.0, 0.0, 0.0, a->proj_charge, a->proj_scale};
          VH_TRY(vh_fallback_impl(a, coarse, weight, a->has_dproj ? proj : nullptr,
                                  a->has_dproj ? a->proj_ball_radius : 0.0,
                                  a->has_dproj ? a->dproj : nullptr, push_factor));
        }
        if (flags) VH_TRY(vh_stream_signal(c, c->side, VH_FLAG_SIDE, seq));
        else VH_CHECK_HIP(hipEventRecord(c->ev_a, c->side));
      }
      if (flags) VH_TRY(vh_stream_wait(c, c->stream, VH_FLAG_JOIN, seq));
      else VH_CHECK_HIP(hipStreamWaitEvent(c->stream, c->join_ev, 0));
      VH_TRY(vh_poisson_fused_finish_impl(fine, true));
      if (flags) VH_TRY(vh_stream_wait(c, c->stream, VH_FLAG_SIDE, seq));
      else VH_CHECK_HIP(hipStreamWaitEvent(c->stream, c->ev_a, 0));
    } else {
    VH_CHECK_HIP(hipEventRecord(c->fork_ev, c->stream));
    VH_TRY(vh_deposit_fine_impl(a, fine, weight));
    if (a->denoiser) VH_TRY(vh_denoiser_apply_impl(a->denoiser, fine, a->denoise_tol));
    // the fine mean field's coefficients need only the density: they fill
    // the main stream's wait for the boundary
    VH_TRY(vh_mean_field_coefficients_impl(fine, jel_nions, jel_radius));
    VH_CHECK_HIP(hipStreamWaitEvent(c->side, c->fork_ev, 0));
    {
      vh_on_side on(c);
      VH_TRY(vh_deposit_cic_impl(a, coarse, weight));
      VH_TRY(vh_poisson_coarse_level_impl(coarse));
      VH_TRY(vh_poisson_fine_boundary_impl(fine, coarse));
      VH_CHECK_HIP(hipEventRecord(c->join_ev, c->side));
      VH_TRY(vh_effective_potential_impl(coarse, jel_nions, jel_radius));
      VH_CHECK_HIP(hipEventRecord(c->fork_ev, c->side));
    }
    VH_CHECK_HIP(hipStreamWaitEvent(c->stream, c->join_ev, 0));
    VH_TRY(vh_poisson_fine_part_impl(fine));
    VH_TRY(vh_hartree_coefficients_add_impl(fine));
    VH_CHECK_HIP(hipStreamWaitEvent(c->stream, c->fork_ev, 0));
    }
  }
  if (step_fuses_push(a)) {
    // the field and fallback kernels push the particles they computed the
    // force of: the same records, forces and keys as forces then vh_verlet
    // (which reads back every record and force), without that pass
    VH_REQUIRE(fine->n == a->n && fine->ctx == a->ctx && coarse->ctx == a->ctx,
               "vh_step: meshes do not match the accelerator");
    if (!a->sorted_once) {
      vh_set_error("vh_step: the cloud has not been sorted");
      return VH_ESTATE;
    }
    const int fb = early_fallback ? VH_FALLBACK_DONE : VH_FALLBACK_AFTER;
    {
      vh_region r(a->ctx, VH_ST_FIELD);
      if (a->has_dproj) {
        // the projectile stepped by the kernel that sums its forces
        const double proj[5] = {0.0, 0.0, 0.0, a->proj_charge, a->proj_scale};
        VH_TRY(vh_forces_impl(a, fine, coarse, weight, proj, a->proj_soft_mode,
                              a->proj_ball_radius, a->dproj, push_factor, fb, dt));
      } else {
        VH_TRY(vh_forces_impl(a, fine, coarse, weight, nullptr, VH_SOFT_DEVICE, 0.0, nullptr,
                              push_factor, fb));
      }
    }
    return VH_OK;
  }
  if (a->has_dproj) {
    VH_TRY(vh_forces_resident(a, fine, coarse, weight));
    VH_TRY(vh_projectile_step(a, dt));
  } else {
    VH_TRY(vh_forces(a, fine, coarse, weight, nullptr, VH_SOFT_DEVICE, 0.0, nullptr, nullptr));
  }
  VH_TRY(vh_verlet(a, dt, mass, rcmax, nullptr));
  return VH_OK;
}

extern "C" int vh_interaction_energy(vh_accel *a, vh_mesh *fine, vh_mesh *coarse, double weight,
                                     double enclosed, double *energy) {
  VH_REQUIRE(a && fine && coarse && energy, "vh_interaction_energy: null argument");
  VH_REQUIRE(fine->n == a->n, "vh_interaction_energy: fine csol has n=%d, accelerator n=%d",
             fine->n, a->n);
  VH_REQUIRE(fine->ctx == a->ctx && coarse->ctx == a->ctx,
             "vh_interaction_energy: objects on different contexts");
  VH_REQUIRE(std::isfinite(weight) && std::isfinite(enclosed),
             "vh_interaction_energy: weight and enclosed charge must be finite");
  VH_CHECK_HIP(hipSetDevice(a->ctx->device));
  VH_TRY(vh_energy_impl(a, fine, coarse, weight, enclosed));
  VH_CHECK_HIP(hipMemcpyAsync(energy, a->dscal + 18, 8, hipMemcpyDeviceToHost, a->ctx->stream));
  VH_CHECK_HIP(hipStreamSynchronize(a->ctx->stream));
  return VH_OK;
}

extern "C" int vh_verlet(vh_accel *a, double dt, double mass, double rcmax, double diag[5]) {
  VH_REQUIRE(a != nullptr, "vh_verlet: null accelerator");
  VH_REQUIRE(dt > 0 && mass > 0, "vh_verlet: dt and mass must be positive");
  VH_CHECK_HIP(hipSetDevice(a->ctx->device));
  {
    vh_region r(a->ctx, VH_ST_VERLET);
    VH_TRY(vh_verlet_impl(a, dt, mass, rcmax));
  }
  if (diag) {
    VH_CHECK_HIP(hipMemcpyAsync(diag, a->dscal, 40, hipMemcpyDeviceToHost, a->ctx->stream));
    VH_CHECK_HIP(hipStreamSynchronize(a->ctx->stream));
  }
  return VH_OK;
}

__global__ void k_enclosed(const int32_t *__restrict__ P, int64_t npad, int64_t n, double x0,
                           double h, double cx, double cy, double cz, double r2,
                           unsigned long long *out) {
  unsigned long long local = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const vh_pos r = vh_load_pos(P, i);
    double px = x0 + (double)(r.k[0] - 1) * h + (double)r.d[0];
    double py = x0 + (double)(r.k[1] - 1) * h + (double)r.d[1];
    double pz = x0 + (double)(r.k[2] - 1) * h + (double)r.d[2];
    double dx = px - cx, dy = py - cy, dz = pz - cz;
    if ((dx * dx + dy * dy) + dz * dz < r2) ++local;
  }
  local = vh_wave_sum(local);
  if (vh_lane() == 0 && local) atomicAdd(out, local);
}

extern "C" int vh_enclosed_count(vh_accel *a, const double center[3], double radius, int64_t *count) {
  VH_REQUIRE(a && center && count, "vh_enclosed_count: null argument");
  VH_CHECK_HIP(hipSetDevice(a->ctx->device));
  VH_TRY(vh_rec_settle(a));
  unsigned long long *d = (unsigned long long *)(a->dscal + 40);
  VH_TRY(vh_fill_u64(a->ctx, (uint64_t *)d, 1, 0));
  hipLaunchKernelGGL(k_enclosed, dim3(2048), dim3(256), 0, a->ctx->stream, a->parts[a->cur],
                     a->npad, a->npart, a->x0, a->h, center[0], center[1], center[2],
                     radius * radius, d);
  VH_LAUNCH_CHECK();
  unsigned long long h = 0;
  VH_CHECK_HIP(hipMemcpyAsync(&h, d, 8, hipMemcpyDeviceToHost, a->ctx->stream));
  VH_CHECK_HIP(hipStreamSynchronize(a->ctx->stream));
  *count = (int64_t)h;
  return VH_OK;
}
