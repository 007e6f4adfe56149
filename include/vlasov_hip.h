/*
 * vlasov_hip.h -- C ABI of libvlasov_hip.so, the gfx950 implementation of the
 * reference's per-step pseudo-particle loop.
 *
 * The boundary is the reference's device layer: every entry point below
 * replaces one method of `DeviceAccelerator` / `DeviceMesh`
 * (src/accelerator.jl, src/devicemesh.jl) or one buffer operation of
 * `DualBuffer` (src/accelerator.jl:28-68), and is cited there.  Plain
 * pointers and sizes only.  Host arrays use the reference's Julia
 * column-major layout: a grid is `[i + n*j + n*n*k]`, a matrix
 * `[row + n*col]`, the force buffer `(3, N)` and the particle array a
 * `PackedParticle{Float32}` record of 48 bytes
 * {int32 knode[3]; float delta[3]; int32 pknode[3]; float pdelta[3]}
 * (src/particles.jl:86-91).  Cell indices written by the library are the
 * reference's 1-based ones.
 *
 * Conventions: every function returns 0 on success and a negative code on
 * failure, with a message for the calling thread in vh_last_error().  Sizes
 * are checked before any launch (the reference throws ArgumentError /
 * DimensionMismatch before launching, src/accelerator.jl:583-586,613).
 * Launches are asynchronous on the context's stream; an entry point that
 * returns a host scalar synchronises first.  Handles are not thread-safe.
 */
#ifndef VLASOV_HIP_H
#define VLASOV_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VH_OK 0
#define VH_EINVAL (-1)   /* bad argument (ArgumentError / DimensionMismatch) */
#define VH_EHIP (-2)     /* HIP runtime error */
#define VH_ESTATE (-3)   /* call out of order (e.g. forces before a sort) */

typedef struct vh_context vh_context;
typedef struct vh_mesh vh_mesh;
typedef struct vh_accel vh_accel;

/* Grids held by a mesh, for vh_mesh_grid_download / _upload. */
enum { VH_GRID_RHO = 0, VH_GRID_PHI = 1, VH_GRID_CSOL = 2 };

/* Law of the projectile <-> pseudo-electron interaction inside the field
 * kernel.  VH_SOFT_DEVICE reproduces the reference's fused device kernel:
 * Gaussian kernel of width scale(softening) for the in-stencil particles and
 * the projectile's total force, positions rebuilt from the cell
 * (src/kernels.jl:425-443, src/accelerator.jl:628-632), the projectile's own
 * law for the out-of-stencil fallback (src/accelerator.jl:719-724).
 * VH_SOFT_BALL applies the CPU path's uniform-ball law everywhere
 * (src/projectile.jl:79-80, 203-232). */
enum { VH_SOFT_DEVICE = 0, VH_SOFT_BALL = 1 };

/* Message of the last failure on this thread ("" if none). */
const char *vh_last_error(void);
/* Library version string. */
const char *vh_version(void);
/* Number of visible GPUs (does not create a context). */
int vh_device_count(int *count);

/* --- context: one device, one stream ------------------------------------ */
int vh_context_create(int device, vh_context **out);
int vh_context_destroy(vh_context *ctx);
/* Waits for every launch of the context (the synchronisation `download!`
 * performs first, src/accelerator.jl:64-68). */
int vh_sync(vh_context *ctx);

/* --- DeviceMesh (src/devicemesh.jl:21-81) --------------------------------
 * Built from the FP64 tables of a SplineMesh; three identical axes.
 *   n        collocation points per axis (= 2 nk), ni = n - 2
 *   colloc   n            knots nk
 *   lap      n x n        full D = S'' S^-1 (src/collocation.jl:94)
 *   M, Minv  ni x ni      eigenbasis of the interior D (src/tensorsolver.jl:128-137)
 *   lam      ni           its eigenvalues
 *   Sinv     n x n        inverse collocation matrix (src/collocation.jl:53)
 *   dual     3 x n        S^-T m_k, k = 0,1,2 (src/mesh.jl:61-62), column k*n
 *   duallen  n            dual lengths (src/deposition.jl:185-195)
 *   loc_x0, loc_invwidth, loc_cells[nbin]  LocateTable (src/splines.jl:362-376),
 *            1-based cells as the reference stores them.
 * Replaces `DeviceMesh(backend, E, mesh)` (src/devicemesh.jl:49-81). */
int vh_mesh_create(vh_context *ctx, int n, int nk,
                   const double *colloc, const double *knots,
                   const double *lap, const double *M, const double *Minv,
                   const double *lam, const double *Sinv,
                   const double *dual, const double *duallen,
                   double loc_x0, double loc_invwidth,
                   const int32_t *loc_cells, int nbin, vh_mesh **out);
int vh_mesh_destroy(vh_mesh *m);
/* Options of a mesh (no counterpart in the reference, whose products are
 * BLAS GEMMs in the device precision, src/tensorsolver.jl:217-223).
 *   VH_MESH_OPT_BF16X3  which rotating products run in split-BF16 x3 on
 *                       v_mfma_f32_16x16x32_bf16 with FP32 accumulation
 *                       (hi = bf16(x), lo = bf16(x - hi); A S ~ A_lo S_hi +
 *                       A_hi S_lo + A_hi S_hi; operators split from FP64)
 *                       instead of FP32 on v_mfma_f32_16x16x4_f32: 0 none
 *                       (default), 1 the Poisson solve's M^-1 and M
 *                       contractions, 2 the S^-1 transforms (spline
 *                       coefficients, mean field), 3 both. */
enum { VH_MESH_OPT_BF16X3 = 0 };
int vh_mesh_set_option(vh_mesh *m, int option, int value);
/* `Array(dev.rho[l])` etc. (src/simulation.jl:544-552): n^3 floats. */
int vh_mesh_grid_download(vh_mesh *m, int which, float *host);
int vh_mesh_grid_upload(vh_mesh *m, int which, const float *host);
/* copyto!(dev.phi[l], csol) at the end of _update_forces_resident!
 * (src/simulation.jl:527-530): grid `dst` of the mesh := grid `src`, on the
 * device, asynchronously. */
int vh_mesh_grid_copy(vh_mesh *m, int dst, int src);
/* The frame of a film step, Float32.((rho[:, :, iz1] .+ rho[:, :, iz2]) ./ 2)
 * of the mesh's density (scripts/xenon.jl:396-410, after sync_host!),
 * reduced on the device: n^2 floats, i fastest; iz1, iz2 1-based planes. */
int vh_density_slice(vh_mesh *m, int iz1, int iz2, float *host);

/* poisson!(phi, rho, dm) single level with multipole faces
 * (src/devicemesh.jl:148-151). */
int vh_poisson(vh_mesh *m);
/* poisson!(phis, rhos, dms) two levels, coarse first, fine faces read from
 * the coarse spline (src/devicemesh.jl:160-169; src/poisson.jl:427-436).
 * Leaves csol of the coarse mesh = S^-1 phi_coarse as a by-product. */
int vh_poisson_nested(vh_mesh *fine, vh_mesh *coarse);
/* spline_coefficients!(csol, phi, dm): csol = (S^-1 x S^-1 x S^-1) phi
 * (src/devicemesh.jl:126-127; src/deposition.jl:354-360). */
int vh_spline_coefficients(vh_mesh *m);
/* effective_potential!(csol, rho, dm, jel): csol += S^-1^3 (V_xc(rho) +
 * V_jel(r)) (src/devicemesh.jl:171-183; src/meanfield.jl:116-147). */
int vh_effective_potential(vh_mesh *m, double nions, double radius);
/* spline_coefficients!(c, phi, dm) followed by effective_potential!(c, rho,
 * dm, jel) on the same mesh (src/devicemesh.jl:126-127, 171-183) in one call:
 * csol = S^-1 (V_xc(rho) + V_jel), then csol += S^-1 phi (the terms added
 * after their transforms, as the two calls add them; vh_step runs the first
 * half before the fine potential is solved). */
int vh_potential_coefficients(vh_mesh *m, double nions, double radius);
/* total_charge(rho, dm) in FP64 (src/devicemesh.jl:112-113). */
int vh_total_charge(vh_mesh *m, double *q);
/* multipole(rho, dm): q, center[3], quadrupole[6] in FP64
 * (src/devicemesh.jl:115-117; src/poisson.jl:152-189). */
int vh_multipole(vh_mesh *m, double out10[10]);

/* multipole(rho, dm) as vh_multipole, without allocating: no counterpart in
 * the reference; for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops).  Inputs and output are the
 * caller's device memory on `device`: rho n^3 floats (i fastest), dual the
 * 3 x n FP64 dual moments S^-T m_k (order k at k*n), out10 the 10 doubles
 * {q, center[3], quadrupole[6]}.  Launched on `stream` (a hipStream_t, NULL
 * for the null stream) without synchronising.  n <= 288.  The FP64 partials
 * go to a static scratch of 320 KiB per GPU, held until the process ends and
 * shared by every call on that GPU: calls that overlap on one GPU race (on
 * one stream they are ordered).  Calls no hipMalloc, hipFree or allocating
 * routine. */
int vh_multipole_noalloc(int device, void *stream, const float *rho, const double *dual, int n,
                         double *out10);

/* --- DeviceAccelerator (src/accelerator.jl:126-185, 281-331) -------------
 *   npart      particles;  nk knots of the uniform fine axis, n = 2 nk
 *   knots      the nk knots (x0 = knots[0], h = (knots[nk-1]-knots[0])/(nk-1),
 *              non-uniform axes refused, src/accelerator.jl:262-268)
 *   spacing    GaussianSmoothing.spacing (src/fields.jl:165)
 *   overlap, gradient  10 x ncol, nodes 8 x ncol: GaussianSmoothing tables
 *              (src/fields.jl:44-53), ncol = nbdt + 1
 * Replaces `DeviceAccelerator(backend, E, fine, sm, npart, n; packed=true)`. */
int vh_accel_create(vh_context *ctx, int64_t npart, int nk, const double *knots,
                    double spacing, int nbdt, const double *overlap,
                    const double *gradient, const double *nodes, vh_accel **out);
int vh_accel_destroy(vh_accel *a);

/* Options of an accelerator (no counterpart in the reference, whose device
 * kernel has one contraction order).
 *   VH_OPT_EXACT_FIELD  0 (default): the 10^3 field contraction runs its
 *                       first stage on the matrix cores, summed in another
 *                       order than contract_spline_10 (FP32-rounding
 *                       agreement); 1: the reference's exact FMA order
 *                       (src/fields.jl:305-326), bit-identical to it.
 *   VH_OPT_STEP_GRAPH   0 (default): vh_step launches its kernels one by
 *                       one; 1: vh_step records them once into a HIP graph
 *                       per particle buffer and replays it (recorded again
 *                       when an argument changes; eager while stage timing
 *                       is on).
 *   VH_OPT_UNIT_SLOTS   the slots of a work unit, the tile of particles one
 *                       wave of the per-cell kernels takes (a power of two in
 *                       [128, 4096], default 256; a cell is cut into units
 *                       of at most this many slots at the next sort).  The
 *                       sort and the field are independent of it to the bit;
 *                       the deposits' FP32 partial sums group by it.
 *                       vh_autotune_unit picks it.
 *   VH_OPT_FIELD_FP32   0 (default): the matrix-core contraction's first
 *                       stage multiplies fp16 splits of the csol block and
 *                       of the x columns (hi*hi + hi*lo + lo*hi, FP32
 *                       accumulation; about 2^-21 of the largest term);
 *                       1: FP32 products.  Ignored with VH_OPT_EXACT_FIELD.
 *   VH_OPT_POISSON_UNFUSED 0 (default): vh_step computes each level's
 *                       spline coefficients straight from the density and
 *                       the boundary, by linearity (the potential's interior
 *                       is not formed: after vh_step phi holds the boundary
 *                       faces only); FP32-rounding agreement with the entry
 *                       points' order.  1: vh_step runs poisson!,
 *                       spline_coefficients! and effective_potential! as the
 *                       entry points do (bit-identical to calling them).
 *                       Meshes in split BF16, or of more than 256 points a
 *                       side, always take the second form. */
enum { VH_OPT_EXACT_FIELD = 0, VH_OPT_STEP_GRAPH = 1, VH_OPT_UNIT_SLOTS = 2, VH_OPT_FIELD_FP32 = 3,
       VH_OPT_POISSON_UNFUSED = 4 };
int vh_accel_set_option(vh_accel *a, int option, int value);

/* upload!/download! of the particle buffer (src/accelerator.jl:48, 64-68):
 * npart PackedParticle{Float32} records, in the device's current (sorted)
 * order. */
int vh_particles_upload(vh_accel *a, const void *records);
int vh_particles_download(vh_accel *a, void *records);
/* upload! of slots [offset, offset + count) only: a cloud built in slices
 * (no counterpart in the reference, whose construction holds the whole
 * cloud on the host as FP64 positions and momenta, docs/src/device.md:196-200).
 * The other slots keep their records. */
int vh_particles_upload_range(vh_accel *a, int64_t offset, int64_t count, const void *records);
/* prime_leapfrog!'s second stage (src/simulation.jl:265-273,
 * full_step_back src/particles.jl:466-469) on the device: every record holding
 * q(-dt/2) (current) and q(0) (previous), with the forces of the last forces
 * call at q(-dt/2), becomes q(0) and -q0 + 2 q(-dt/2) + coef f, in FP64.
 * coef = dt/M (the thesis's coefficient) or dt^2/(4M) (consistent). */
int vh_prime(vh_accel *a, double coef);
/* download!(acc.force): (3, npart) floats, slot order (src/accelerator.jl:647). */
int vh_forces_download(vh_accel *a, float *host);
int vh_forces_upload(vh_accel *a, const float *host);

/* _sort_and_list!(acc) for a resident cloud (src/accelerator.jl:466-503):
 * stable counting sort of the records by clamped cell key
 * (src/kernels.jl:902-908), occupied cells and bounds, and the list of
 * particles whose 10^3 stencil leaves the fine grid (src/kernels.jl:1041-1058).
 * The records are moved; the order within a cell is the previous order
 * (stable, as cellsort!, src/sorting.jl:163-204).  After a vh_step the move
 * is deferred to the next reader (the deposits and forces read through the
 * sort's permutation, the next vh_step's push moves the records); every
 * entry point that returns or reads records sees them in sorted order.
 * Outputs may be NULL. */
int vh_sort(vh_accel *a, int32_t *n_occupied, int32_t *outcount);
/* The occupied-cell list (1-based keys, increasing), bounds (length n_occ+1,
 * exclusive prefix) and outlist (1-based sorted slots, increasing) of the
 * last sort.  Arrays must hold n_occupied, n_occupied+1 and outcount
 * entries. */
int vh_sort_lists(vh_accel *a, int32_t *occupied, int32_t *bounds, int32_t *outlist);
/* _fill_columns!: table columns (3, npart) 1-based and nout, the particles
 * outside [knots[2] + h/2, knots[end-1] - h/2] (src/accelerator.jl:516-529;
 * src/kernels.jl:754-779). */
int vh_columns(vh_accel *a, int32_t *cols, int32_t *nout);

/* deposit_smoothed!(rho, acc, mesh, sm, positions; charge) -> nout, resident
 * (src/accelerator.jl:531-565): sort, columns, sorted 8^3 deposit with no
 * float atomics, FP64 total charge, renormalisation; rho of `fine`. */
int vh_deposit_smoothed(vh_accel *a, vh_mesh *fine, double charge, int32_t *nout);
/* deposit_cic!(rho, dm, acc, npart, charge, x0f, hf) (src/devicemesh.jl:93-106):
 * trilinear deposit onto the stretched coarse grid, per-cell tiles combined
 * through exact 64-bit fixed-point sums (deterministic). */
int vh_deposit_cic(vh_accel *a, vh_mesh *coarse, double charge);

/* forces!(cloud, acc, fine, csol_f, coarse, csol_c, sm; projectile, packed)
 * (src/accelerator.jl:607-733): smoothed 10^3 field of the fine csol,
 * fused projectile, and the out-of-stencil fallback on the device
 * (coarse spline_field + the projectile's own law, src/fields.jl:358-393).
 * `proj` = {x, y, z, charge, scale} of the projectile or NULL; `soft_mode`
 * VH_SOFT_*; `ball_radius` is the radius of the projectile's own ball law
 * used by the fallback (<= 0: Gaussian law of width scale in the fallback).
 * red[4] (may be NULL) receives the reduction of src/kernels.jl:449-469:
 * sum of the per-particle reaction forces and of the pair potential. */
int vh_forces(vh_accel *a, vh_mesh *fine, vh_mesh *coarse, double weight,
              const double *proj, int soft_mode, double ball_radius,
              double red[4], int32_t *nout);

/* --- the legacy route: a cloud held on the host --------------------------
 * ForceAccelerator(MtlArray, fine, sm, npart, n) and update_forces!'s
 * accelerator branch (ext/VlasovMetalExt.jl:51-54; src/simulation.jl:318-345;
 * scripts/film_xenon_80M.jl:83-120): the cloud, the CIC deposit, the FP64
 * Poisson chain and the Verlet stay with the caller; the smoothed deposit and
 * the field run here.
 *
 * _pack!(acc, positions) (src/accelerator.jl:342-349): pos = npart x 3 FP64
 * absolute positions in the cloud's own order (particle-major, (3, N)
 * column-major), packed on the device in FP64 as _pack_kd_kernel! does
 * (src/kernels.jl:134-145: nearest knot clamped to [1, nk], offset subtracted
 * in FP64 then narrowed) and kept in FP64.  The next sort (vh_deposit_smoothed
 * makes one) records which cloud particle each sorted slot holds; the density
 * is then read with vh_mesh_grid_download (deposit_smoothed!(ρ::Array, acc,
 * ...), src/accelerator.jl:531-578). */
int vh_positions_upload(vh_accel *a, const double *pos);
/* forces!(cloud, acc, fine, csol_f, coarse, csol_c, sm; escaped, projectile,
 * packed) for that cloud (src/accelerator.jl:607-733): csol_f (n^3 FP64, the
 * fine level) narrowed into the fine mesh's csol; the field kernel with the
 * device law of the projectile on the cells whose 10^3 stencil fits; every
 * other particle by the coarse spline_field of csol_c (nc^3 FP64 on knots_c,
 * nc = 2 nkc) at its FP64 position, or the monopole of `escaped` charges
 * beyond both grids, plus the projectile's own law (ball of ball_radius > 0,
 * else the Gaussian of proj's scale).  forces: npart x 3 FP64 in the cloud's
 * own order (the in-stencil FP32 forces widened, as T(f[s]) does).  proj =
 * {x, y, z, charge, scale} or NULL; red[4] and outcount as vh_forces.
 * Needs a sort of a cloud from vh_positions_upload (VH_ESTATE otherwise). */
int vh_forces_host(vh_accel *a, vh_mesh *fine, const double *csol_f, const double *knots_c,
                   int nkc, const double *csol_c, int nc, double weight, const double *proj,
                   double ball_radius, double escaped, double *forces, double red[4],
                   int32_t *outcount);

/* step!(cloud, dt, acc; rcmax) -> (kinetic, escaped, angular[3])
 * (src/accelerator.jl:580-605): packed position Verlet, diagnostics summed
 * in FP64. diag may be NULL. */
int vh_verlet(vh_accel *a, double dt, double mass, double rcmax, double diag[5]);

/* enclosed_charge(cloud, proj, radius) (src/projectile.jl:243-247): number of
 * particles closer than `radius` to `center` (multiply by the weight). */
int vh_enclosed_count(vh_accel *a, const double center[3], double radius, int64_t *count);

/* --- the projectile kept on the device ------------------------------------
 * No counterpart in the reference, which integrates the projectile on the
 * host between the force kernel and the cloud's push (src/simulation.jl:524,
 * src/accelerator.jl:746-756, src/projectile.jl:298-304), and so
 * synchronises every step.  With the projectile on the device a step needs
 * no host synchronisation.  state = position[3], previous[3], velocity[3]. */
int vh_projectile_set(vh_accel *a, const double state[9], double mass, double charge,
                      double scale, int soft_mode, double ball_radius, double jel_nions,
                      double jel_radius);
int vh_projectile_get(vh_accel *a, double state[9]);
/* forces!(...; projectile) as vh_forces, with the device projectile's
 * position; no output, no synchronisation. */
int vh_forces_resident(vh_accel *a, vh_mesh *fine, vh_mesh *coarse, double weight);
/* projectile_forces!(cloud, acc, proj, jel) from the last vh_forces_resident's
 * reduction, then step!(proj, force, dt), in FP64 on the device. */
int vh_projectile_step(vh_accel *a, double dt);

/* One step of the resident loop without host synchronisation, the sequence
 * of _update_forces_resident! and step! (src/simulation.jl:472-541; the
 * driver's nine calls): vh_deposit_smoothed, vh_deposit_cic,
 * vh_poisson_nested, vh_potential_coefficients (fine: the spline
 * coefficients and the mean field in one call), vh_effective_potential
 * (coarse; jellium of jel_nions ions and radius jel_radius), the forces
 * with the device projectile (vh_forces_resident + vh_projectile_step) or,
 * without vh_projectile_set, without projectile, and vh_verlet.  One call
 * from the host per step. */
int vh_step(vh_accel *a, vh_mesh *fine, vh_mesh *coarse, double weight, double dt, double mass,
            double rcmax, double jel_nions, double jel_radius);

/* --- the batching extension (no counterpart in the reference) -------------
 * vh_autotune_unit: `steps` real steps of vh_step with each candidate work-
 * unit size (VH_OPT_UNIT_SLOTS; one untimed step first, whose sort cuts the
 * cells with it), timed with HIP events on the accelerator's stream; keeps
 * the fastest, writes it to *chosen and each candidate's ms per step to ms
 * (may be NULL). */
int vh_autotune_unit(vh_accel *a, vh_mesh *fine, vh_mesh *coarse, double weight, double dt,
                     double mass, double rcmax, double jel_nions, double jel_radius,
                     const int *candidates, int ncand, int steps, int *chosen, double *ms);
/* A batch of independent simulations on one device, one accelerator (with
 * its own context) and two meshes each; params = count x {weight, dt, mass,
 * rcmax, jel_nions, jel_radius}, the arguments of each member's vh_step.
 * vh_batch_step makes one vh_step of every member: with use_graph, the
 * members' launches are recorded once (per pattern of their host state: the
 * steps alternate two) into one HIP graph that forks from the batch's stream
 * into every member's streams and joins back, and replayed; every member's
 * later calls follow the batch's step.  Without, each member's vh_step in
 * turn, on its own streams. */
typedef struct vh_batch vh_batch;
int vh_batch_create(int count, vh_accel *const *accs, vh_mesh *const *fines,
                    vh_mesh *const *coarses, const double *params, int use_graph, vh_batch **out);
int vh_batch_step(vh_batch *b);
int vh_batch_sync(vh_batch *b);
int vh_batch_destroy(vh_batch *b);

/* interaction_energy(cloud, fine, csol_f, coarse, csol_c, sm; enclosed)
 * (src/energy.jl:19-43): sum over the particles of weight * phi(r), in FP64,
 * with phi the smoothed potential of the fine csol inside
 * (knots[3], knots[end-2])^3 (src/fields.jl:405-428), the spline potential of
 * the coarse csol elsewhere (src/fields.jl:440-466), and enclosed / |r|
 * beyond both grids.  Reads the csol grids of `fine` and `coarse` as they
 * stand: called between vh_spline_coefficients and vh_effective_potential it
 * gives twice the Hartree energy (src/simulation.jl:501-510); after the mean
 * field, the total interaction of the budget (src/simulation.jl:382-396). */
int vh_interaction_energy(vh_accel *a, vh_mesh *fine, vh_mesh *coarse, double weight,
                          double enclosed, double *energy);

/* --- timing ---------------------------------------------------------------
 * HIP events recorded on the context's stream around each stage of the step
 * (and around the field kernel alone), for per-stage ms/step as
 * scripts/profil_device.jl reports them.  No counterpart in the reference,
 * which times stages with a device synchronisation after each. */
enum {
  VH_ST_SORT = 0,       /* counting sort, occupied cells, outlist */
  VH_ST_DEPOSIT_FINE,   /* smoothed 8^3 deposit + charge + renormalisation */
  VH_ST_DEPOSIT_CIC,    /* coarse cloud-in-cell */
  VH_ST_POISSON,        /* nested Poisson (incl. S^-1 of the coarse phi) */
  VH_ST_SPLINE,         /* S^-1 of the fine phi */
  VH_ST_MEANFIELD,      /* effective potential, both levels */
  VH_ST_FIELD,          /* field kernel + reduction + fallback */
  VH_ST_VERLET,         /* packed Verlet + diagnostics */
  VH_ST_FIELD_KERNEL,   /* the smoothed-field kernel launch alone */
  VH_ST_DEPOSIT_KERNEL, /* the per-cell deposit tile kernel alone */
  VH_ST_COUNT
};
/* mask: bit s times stage s (0 stops timing; VH_PROF_ALL every stage).
 * Each timed region costs two event records on the stream. */
#define VH_PROF_ALL (-1)
int vh_profile_enable(vh_context *ctx, int mask);
/* Synchronises, adds the elapsed time of every recorded region to ms[stage]
 * and the number of regions to count[stage] (arrays of VH_ST_COUNT), and
 * clears the record. */
int vh_profile_read(vh_context *ctx, double *ms, int64_t *count);

/* --- host-side sampling (native, no device) -------------------------------
 * Ran2 of the thesis code, IQ1 = 3668 unless consistent != 0
 * (src/random.jl:288-347), and the two Thomas-Fermi samplers:
 * sample_thomas_fermi(::RadialProfile) (src/initial.jl:94-113) and
 * sample_thomas_fermi(::PotentialProfile) (src/initial.jl:204-229).
 * pos, mom: npart x 3 doubles, particle-major ((3, N) column-major). */
int vh_ran2_sequence(int seed, int consistent, int64_t count, float *out);
/* sample_thomas_fermi(::RadialProfile, n, w; rng) in slices: a sampler holds
 * the profile and the generator (seeded as Ran2(seed)); each vh_sampler_next
 * draws the next `count` particles of the sequence a single vh_sample_radial
 * call would draw (pos, mom: count x 3 doubles). */
typedef struct vh_sampler vh_sampler;
int vh_sampler_create(const double *quantiles, int nq, const double *density, int nrho,
                      double rmax, double weight, int seed, vh_sampler **out);
int vh_sampler_next(vh_sampler *s, int64_t count, double *pos, double *mom);
int vh_sampler_destroy(vh_sampler *s);
int vh_sample_radial(const double *quantiles, int nq, const double *density, int nrho,
                     double rmax, int64_t npart, double weight, int seed,
                     double *pos, double *mom);
int vh_sample_rejection(const double *grid, const double *potential, int ngrid,
                        double rmax, double pmax, double fermi, int64_t npart,
                        double weight, int seed, double *pos, double *mom);

/* --- the learned denoiser -------------------------------------------------
 * No counterpart in the reference.  A three-level 3-D U-Net (16 / 32 / 64
 * channels, 3^3 convolutions with zero padding, 2^3 average pooling, nearest
 * 2x upsampling, skip connections by concatenation, ReLU) mapping the fine
 * density deposited by N pseudo-particles to the one 10 N deposit: the input
 * x = rho / rho0 (rho0 the jellium's mean density), the output
 * rho0 (x + m o(x)) with m = [|x| > support] (default 1e-4,
 * vh_denoiser_set_support): the net corrects the density where there is
 * charge and moves none into the vacuum.  Products use bf16 roundings of the weights with FP32
 * accumulation; the optimiser (Adam) keeps FP32 master weights.
 *
 * Weights, vh_denoiser_param_count() floats: for each layer l = 0..10 in
 * order, w[tap][cin][cout] then b[cout], tap = 9 dz + 3 dy + dx over the
 * offsets (dx, dy, dz) - 1 in {-1, 0, 1}^3 (x the grid's fastest index i).
 * Layers (cin -> cout): 1->16, 16->16 | pool | 16->32, 32->32 | pool |
 * 32->64, 64->64, 64->32 | up, concat [up, level-2 skip] | 64->32, 32->16 |
 * up, concat [up, level-1 skip] | 32->16, 16->1 (no ReLU on the last).
 * weights NULL: He-normal weights from `seed`, zero biases and a zero last
 * layer (the untrained net is the identity). */
typedef struct vh_denoiser vh_denoiser;
int vh_denoiser_param_count(int64_t *n);
int vh_denoiser_create(vh_context *ctx, const float *weights, int64_t nweights, uint64_t seed,
                       double rho0, vh_denoiser **out);
/* waits for the device; may follow the destruction of its context */
int vh_denoiser_destroy(vh_denoiser *d);
/* master weights and the bf16 roundings the products use (either NULL) */
int vh_denoiser_get_weights(vh_denoiser *d, float *w, float *weff);
/* replaces the master weights (and restarts the optimiser) */
int vh_denoiser_set_weights(vh_denoiser *d, const float *w, int64_t nweights);
/* One cube of side D (multiple of 4, <= 160) through the net: x, t the
 * scaled input and target ([z][y][x], D^3 each), o the correction (may be
 * NULL), *loss = mean((x + m o - t)^2), grad (may be NULL) its gradient with
 * respect to every parameter (the master layout). */
int vh_denoiser_eval(vh_denoiser *d, int D, const float *x, const float *t, float *o,
                     double *loss, float *grad);
/* The training set in device memory: pairs of n^3 fine densities. */
int vh_denoiser_dataset_reserve(vh_denoiser *d, int n, int64_t capacity);
/* appends (rho of `low`, rho of `high`): the N and the 10 N density */
int vh_denoiser_dataset_add(vh_denoiser *d, vh_mesh *low, vh_mesh *high);
int vh_denoiser_dataset_size(vh_denoiser *d, int64_t *count);
/* `iters` Adam steps (learning rate lr), each on the mean gradient of
 * `batch` crops of side C (multiple of 4) from pairs [first, last), their
 * corners within `radius` points of the centred corner, drawn from `seed`;
 * losses[it] (may be NULL) the batch's mean crop loss. */
int vh_denoiser_train(vh_denoiser *d, int iters, int C, int batch, int64_t first, int64_t last,
                      int radius, double lr, uint64_t seed, double *losses);
/* mean((x - t)^2) and mean((x + o - t)^2) over the central cube of side C
 * of pair k: the plain and the denoised density's error */
int vh_denoiser_dataset_error(vh_denoiser *d, int64_t k, int C, double *mse_plain,
                              double *mse_net);
/* The net on a mesh's density in place (the density padded with zeros to
 * a multiple of 4 points per axis), then the residual charge check: with q
 * the density's total charge and q' the net's, the net's density scaled by
 * q / q' replaces it when |q' / q - 1| <= tol; otherwise the density stays
 * as deposited and the fallback count grows.  Asynchronous. */
int vh_denoiser_apply(vh_denoiser *d, vh_mesh *m, double tol);
int vh_denoiser_fallbacks(vh_denoiser *d, int32_t *count);
/* the support threshold m = [|x| > xmin] of the correction (training and
 * inference) */
int vh_denoiser_set_support(vh_denoiser *d, double xmin);
/* The inference path of vh_denoiser_apply: 0 (default) the matrix cores,
 * activations rounded to bf16 between layers (v_mfma_f32_16x16x32_bf16,
 * FP32 accumulation, output rho + rho0 o); 1 the FP32-activation kernels of
 * the training, output rho0 (x + o).  Choose before attaching to a recorded
 * step. */
int vh_denoiser_set_inference(vh_denoiser *d, int fp32);
/* vh_step runs vh_denoiser_apply(d, fine, tol) between the fine deposit and
 * the Poisson solve (d NULL detaches; the denoiser must share the
 * accelerator's context and outlive the attachment). */
int vh_accel_set_denoiser(vh_accel *a, vh_denoiser *d, double tol);

#ifdef __cplusplus
}
#endif
#endif /* VLASOV_HIP_H */
