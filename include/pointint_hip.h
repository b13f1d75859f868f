/*
 * pointint_hip.h -- C ABI of the MI355X (gfx950) k-space intersection-volume
 * library libpointint_hip.so.
 *
 * The path: per-object form factors A_i(k_q) over a Gauss-Legendre x Lebedev
 * k-grid, the all-pairs volume matrix V = Re(J^H D J) / (8 pi^3), its adjoint
 * dV/dvertex, and a serving loop around them.  Plain C types only: pointers,
 * sizes, int status codes.  Every entry point returns PII_OK (0) or an error
 * code, and pii_last_error() returns the calling thread's last message.
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   - CudaKernelRegistry / CudaComputeAkFunc plug-in functors
 *       cpp/include/dof/cuda/dof_cuda_interface.hpp:24-67
 *   - compute_intersection_volume_matrix_cuda
 *       cpp/include/compute_intersection_volume_multi_object.hpp:51-69
 *   - compute_intersection_volume_cuda (pairwise)
 *       cpp/include/compute_intersection_volume.hpp:48-59, :81-96
 *   - compute_Ak_cuda / compute_Ak_gradient_cuda
 *       cpp/include/compute_intersection_volume.hpp:120-135
 *   - load_lebedev_txt, gauss_legendre_interval, build_kgrid
 *       cpp/include/quadrature/lebedev_io.hpp:24-59,
 *       cpp/include/quadrature/gauss_legendre.hpp:16-62,
 *       cpp/include/quadrature/kgrid.hpp:31-47
 *   - make_triangle_mesh / make_point_cloud / make_gaussian_splat
 *       cpp/include/geometry/packing.hpp:31-45
 */
#ifndef POINTINT_HIP_H
#define POINTINT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
  PII_OK = 0,
  PII_ERR_INVALID = 1,     /* bad argument, size mismatch, unknown type */
  PII_ERR_HIP = 2,         /* HIP runtime error */
  PII_ERR_NOMEM = 3,       /* device allocation failed */
  PII_ERR_UNSUPPORTED = 4, /* DoF / geometry combination without a kernel */
  PII_ERR_BUDGET = 5,      /* serving: request refused by the HBM budget */
  PII_ERR_IO = 6,          /* file could not be read */
  PII_ERR_BUSY = 7         /* serving: no result ready yet */
};

/* GeometryType: cpp/include/geometry/types.hpp:40-44 */
enum { PII_GEOM_TRIANGLE = 0, PII_GEOM_DISK = 1, PII_GEOM_GAUSSIAN = 2 };

/* DoF parameterisations: AffineDoF (12 = t.xyz + A row-major,
 * cpp/include/dof/affine_dof.hpp:15-22) and TriangleMeshDoF (3V vertex
 * coordinates, cpp/include/dof/triangle_mesh_dof.hpp:12-17). */
enum { PII_DOF_AFFINE = 0, PII_DOF_TRIMESH = 1 };

/* Form-factor precision.  FP64: the reference formula in FP64 on the GPU.
 * FAST: FP64 dot products / phases / Taylor branch, FP32 transcendentals
 * after FP64 range reduction, FP64 accumulation. */
enum { PII_FF_FP64 = 0, PII_FF_FAST = 1 };

/* Gram mode for V = X^T X with X = sqrt(w) [Re J; Im J]. */
enum { PII_GRAM_FP64 = 0, PII_GRAM_BF16X3 = 1, PII_GRAM_FP16X3 = 2 };

/* ComputationFlags: cpp/include/computation_flags.hpp:12-17 */
enum { PII_FLAG_VOLUME = 1, PII_FLAG_GRADIENT = 2, PII_FLAG_HESSIAN = 4 };

/* On any error the matrix and form-factor entries leave zeros in the
 * outputs they were given (whose sizes the arguments define), as the
 * reference drivers return zero results
 * (cpp/src/compute_intersection_volume_multi_object.cu:250-261). */

typedef struct pii_context pii_context;
typedef struct pii_kgrid pii_kgrid;
typedef struct pii_geometry pii_geometry;

const char* pii_last_error(void);
const char* pii_version(void);

/* Testing hook, no counterpart in the reference: after `after` more
 * successful device or pinned-host allocations, the next one fails as an
 * out-of-memory would (PII_ERR_NOMEM); -1 turns it off.  Lets the error
 * paths be tested without exhausting the device. */
int pii_debug_fail_allocation(int after);

/* ---- host quadrature utilities ------------------------------------------ */
/* Gauss-Legendre nodes/weights on [a,b] (ascending).
 * Replaces gauss_legendre_interval, cpp/include/quadrature/gauss_legendre.hpp:16-62 */
int pii_gauss_legendre(int n, double a, double b, double* x, double* w);

/* Lebedev rule file of "theta phi w" lines in degrees.  With dirs/w NULL only
 * the count is returned in *count.  Weights summing to ~1 are scaled by 4 pi.
 * Replaces load_lebedev_txt, cpp/include/quadrature/lebedev_io.hpp:24-59 */
int pii_lebedev_load(const char* path, double* dirs, double* w, int cap, int* count);

/* k-grid: node q = ir*M + j, k = tan t_ir, w = w_leb[j] * w_t[ir] * sec^2 t_ir.
 * Outputs dirs[Q*3], kmag[Q], w[Q] with Q = M*nrad.
 * Replaces build_kgrid, cpp/include/quadrature/kgrid.hpp:31-47 */
int pii_kgrid_build(const double* leb_dirs, const double* leb_w, int M, int nrad,
                    double* dirs, double* kmag, double* w);

/* ---- context ------------------------------------------------------------- */
int pii_context_create(int device, pii_context** out);
int pii_context_destroy(pii_context* ctx);
int pii_context_synchronize(pii_context* ctx);
/* Phase timing with HIP events on the context's stream (the reference's
 * enable_profiling flag, e.g.
 * cpp/src/compute_intersection_volume_multi_object.cu:626-695, which used host
 * clocks).  Phases: 0 pre-pass, 1 form factor, 2 Gram, 3 derivatives (the
 * adjoint, or the forward dA/dtheta columns of the affine entries), and within
 * phase 1 on the fast path: 4 moment-series kernel, 5 per-face kernel; ms[p]
 * holds the last call's duration of phase p (0 when it did not run). */
int pii_context_set_profiling(pii_context* ctx, int enable);
int pii_context_timings(pii_context* ctx, double* ms, int n);
/* Work done by the last call's form-factor kernels (for roofline accounting;
 * no reference counterpart): units[0] triangle x node evaluations of the
 * per-face kernel, [1] face x direction passes of the moment series, [2]
 * object x node evaluations of the series, [3] disk/Gaussian x node
 * evaluations of the per-element kernel, [4] disk/Gaussian x direction passes
 * of the moment series.  Entries beyond 4 are set to 0. */
int pii_context_work(pii_context* ctx, double* units, int n);
/* Low-|k| moment series of the fast path (default on).  With it, objects on a
 * shell-structured k-grid (every shell repeats one direction set, as
 * pii_kgrid_build makes) take their form factor at the leading shells with
 * |k| R <= 15 (R the object's radius about its bounding-box centre; for disks
 * and Gaussians also a bound on rho |k| or sigma |k|) from a 56-term moment
 * expansion in FP64, and the per-element evaluation only above; the adjoint
 * follows the same split.  No reference counterpart (the reference evaluates
 * every element at every node); switching it off gives the per-element path
 * everywhere. */
int pii_context_set_series(pii_context* ctx, int enable);

/* ---- k-grid resident in HBM ----------------------------------------------- */
int pii_kgrid_create(pii_context* ctx, int Q, const double* dirs, const double* kmag,
                     const double* w, pii_kgrid** out);
int pii_kgrid_destroy(pii_kgrid* kg);
/* *Q full node count; *Qh nodes actually evaluated by the Gram (half when the
 * grid is centrally symmetric with equal antipodal weights). */
int pii_kgrid_info(const pii_kgrid* kg, int* Q, int* Qh, int* symmetric);

/* ---- geometry resident in HBM ---------------------------------------------- */
/* Triangle mesh from V[nv*3], F[nf*3] (make_triangle_mesh,
 * cpp/src/geometry/packing.cpp:10-28, :66-70). */
int pii_geometry_create_mesh(pii_context* ctx, const double* V, int nv, const int32_t* F,
                             int nf, pii_geometry** out);
/* Packed triangles tri[nt*12] = {a, e1, e2, S} with vid[nt*3] (TriPacked,
 * cpp/include/geometry/types.hpp:9-17).  Negative vids are allowed; such
 * faces contribute nothing under TRIMESH DoFs, as in
 * cpp/src/dof/triangle_mesh_dof.cpp:44-45. */
int pii_geometry_create_tris(pii_context* ctx, const double* tri, const int32_t* vid, int nt,
                             pii_geometry** out);
/* Disks el[n*8] = {c, n(unit), rho, area} (DiskPacked, types.hpp:21-27). */
int pii_geometry_create_disks(pii_context* ctx, const double* el, int n, pii_geometry** out);
/* Gaussian surfels el[n*8] = {c, n(unit), sigma, w} (GaussianPacked, types.hpp:31-37). */
int pii_geometry_create_gaussians(pii_context* ctx, const double* el, int n,
                                  pii_geometry** out);
int pii_geometry_destroy(pii_geometry* g);
int pii_geometry_info(const pii_geometry* g, int* type, int* num_elements, int* num_dof_vertices);
/* Geometry preprocessing for rigid poses (no counterpart in the reference,
 * whose series-free path re-reads every face at every node): expands the
 * low-|k| moment series of each triangle mesh of geoms[0..n) with at least
 * 192 faces in spherical harmonics (degree <= 57, 17,544 coefficient pairs,
 * 274 KiB of device memory per mesh), in the mesh's own frame.  Afterwards
 * the fast form factor, volume matrix and serving entries evaluate that
 * series from the harmonics (a fixed cost per direction instead of one pass
 * over the faces) whenever the mesh's affine DoFs are a similarity
 * x' = s R x + t (R a rotation); other DoFs, the FP64 and derivative entries
 * are unaffected.  Other geometries are left as they are.  Synchronous;
 * PII_ERR_BUSY while a serving query uses one of the geometries; not to be
 * called concurrently with other calls that use them. */
int pii_geometry_cache_moments(pii_context* ctx, int n, pii_geometry* const* geoms);

/* ---- form factors ---------------------------------------------------------- */
/* A(k_q) for every node of kg, written to out[2*Q] (interleaved re, im, host
 * memory).  ndofs must be 12 for AFFINE and 3*num_dof_vertices for TRIMESH.
 * Replaces the CudaComputeAkFunc plug-ins compute_Ak_{affine_triangle,
 * affine_disk,affine_gaussian,triangle_mesh}_cuda_wrapper
 * (cpp/src/dof/cuda/affine_dof_cuda.cu:855-1243,
 *  cpp/src/dof/cuda/triangle_mesh_dof_cuda.cu:314-384). */
int pii_form_factor(pii_context* ctx, const pii_geometry* g, int dof_kind, const double* dofs,
                    int ndofs, const pii_kgrid* kg, int precision, double* out);

/* The same A(k_q) written into caller-owned DEVICE memory d_out[2*Q] (device
 * memory of the context's device, else PII_ERR_INVALID), synchronously: the
 * contract of the CudaComputeAkFunc plug-ins, which write into a device
 * double2* (cpp/include/dof/cuda/dof_cuda_interface.hpp:24-30).
 * Ordering: the writes follow all work queued on the caller's stream before
 * the call (the null stream, as the reference's plug-ins use, unless
 * pii_context_set_caller_stream names another), and are complete when the
 * call returns, so work queued on any stream afterwards sees them. */
int pii_context_set_caller_stream(pii_context* ctx, void* hip_stream);

/* pii_form_factor_device without allocations.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, result and ordering as pii_form_factor_device; every
 * intermediate buffer comes from a fixed 64 MiB device array per GPU (held
 * until the process ends) and a fixed 1 MiB host staging array per thread,
 * and no hipMalloc / hipFree / hipHostMalloc is called.  A call that needs
 * more (records of about 160 bytes per triangle, the element-split partial
 * sums, 32 bytes per k-node) fails with PII_ERR_INVALID ("... exhausted").
 * Calls on one GPU share its array and are serialised by a lock. */
int pii_form_factor_device_noalloc(pii_context* ctx, const pii_geometry* g, int dof_kind,
                                   const double* dofs, int ndofs, const pii_kgrid* kg,
                                   int precision, double* d_out);
int pii_form_factor_device(pii_context* ctx, const pii_geometry* g, int dof_kind,
                           const double* dofs, int ndofs, const pii_kgrid* kg, int precision,
                           double* d_out);

/* dA(k_q)/dtheta for every node of kg in FP64 (the reference's element
 * derivatives), written to out[2*Q*ndofs] as a row-major Q x ndofs complex
 * matrix (interleaved re, im, host memory): out[2*(q*ndofs + t)].  ndofs as
 * for pii_form_factor.  Replaces the CudaComputeAkGradientFunc plug-ins
 * compute_Ak_gradient_{affine_triangle,affine_disk,affine_gaussian,
 * triangle_mesh}_cuda_wrapper (cpp/include/dof/cuda/dof_cuda_interface.hpp:34-40,
 * cpp/src/dof/cuda/affine_dof_cuda.cu:932-1329,
 * cpp/src/dof/cuda/triangle_mesh_dof_cuda.cu:387-466) and, on a one-node
 * grid, compute_Ak_gradient_cuda (cpp/include/compute_intersection_volume.hpp:128-135).
 * Q x ndofs doubles of device and host memory: a testing-scale helper, as
 * in the reference (the volume path contracts instead: pii_volume_matrix_vjp). */
int pii_form_factor_jacobian(pii_context* ctx, const pii_geometry* g, int dof_kind,
                             const double* dofs, int ndofs, const pii_kgrid* kg, double* out);

/* The same Jacobian written into caller-owned DEVICE memory d_out[2*Q*ndofs]
 * (row-major Q x ndofs complex), synchronously, ordered as
 * pii_form_factor_device: the CudaComputeAkGradientFunc plug-in contract
 * (cpp/include/dof/cuda/dof_cuda_interface.hpp:34-40). */
int pii_form_factor_jacobian_device(pii_context* ctx, const pii_geometry* g, int dof_kind,
                                    const double* dofs, int ndofs, const pii_kgrid* kg,
                                    double* d_out);

/* ---- all-pairs volume matrix ----------------------------------------------- */
/* V[n*n] (row-major, symmetric) for n objects, each with its own DoF kind and
 * DoF vector.  Replaces compute_intersection_volume_matrix_cuda with
 * VOLUME_ONLY (cpp/src/compute_intersection_volume_multi_object.cu:233-712). */
int pii_volume_matrix(pii_context* ctx, int n, const pii_geometry* const* geoms,
                      const int* dof_kinds, const double* const* dofs, const int* ndofs,
                      const pii_kgrid* kg, int ff_precision, int gram_mode, double* V);

/* Gauss-Newton Hessian blocks of the volume matrix:
 *   H_ij = (1/8 pi^3) sum_q w_q Re( dA_i/dtheta_i (dA_j/dtheta_j)^H ),
 * every ordered pair (i, j), i == j included.  H is (D x D) row-major with
 * D = sum_i ndofs[i]; block (i, j) occupies rows o_i..o_i+ndofs[i]-1 and
 * columns o_j..o_j+ndofs[j]-1, o_i = sum_{i' < i} ndofs[i'] (12 i when every
 * object has AffineDoF), in each DoF vector's own order.  gram_mode as for
 * pii_volume_matrix.  V (may be NULL) receives the volume matrix of the same
 * call.  AffineDoF objects (triangles, disks, Gaussians) take the fast
 * derivative kernels when every object of the call has AffineDoF;
 * TriangleMeshDoF objects (and every object of a call that has one) take the
 * FP64 per-node Jacobian of pii_form_factor_jacobian.  D plus n (with V) is
 * limited to 16384 (PII_ERR_INVALID above).
 * Replaces the HESSIAN flag of compute_intersection_volume_matrix_cuda/cpu
 * (cpp/src/compute_intersection_volume_multi_object.cpp:236-292), whose
 * hessian_ij[i][j] is block (i, j) and whose hessian_ii / hessian_jj are zero
 * in the Gauss-Newton form. */
int pii_volume_matrix_hessian(pii_context* ctx, int n, const pii_geometry* const* geoms,
                              const int* dof_kinds, const double* const* dofs, const int* ndofs,
                              const pii_kgrid* kg, int gram_mode, double* H, double* V);

/* Volume matrix, per-pair gradients and Gauss-Newton Hessian blocks of
 * AffineDoF objects (triangles, disks, Gaussians) from one pass: the fast form
 * factors A_i, the forward-mode columns dA_i/dtheta_i and one Gram over all
 * 13n columns.  Outputs (each may be NULL, not all):
 *   V[n*n]          volume matrix (row-major);
 *   grad[n*n*12]    grad[(i*n + j)*12 + t] = dV_ij/dtheta_i,t
 *                   = (1/8 pi^3) sum_q w_q Re(dA_i/dtheta_t conj A_j), which
 *                   for i == j is half the full derivative of V_ii;
 *   H[12n*12n]      as pii_volume_matrix_hessian.
 * Replaces compute_intersection_volume_matrix_cuda with GRADIENT and/or
 * HESSIAN for AffineDoF objects: its grad_matrix[i][j] is grad block (i, j)
 * (cpp/src/compute_intersection_volume_multi_object.cu:437-520, from dA/dtheta
 * as here; cpp/include/compute_intersection_volume_multi_object.hpp:22-26 for
 * the i == j convention), its hessian_ij[i][j] block (i, j) of H
 * (compute_intersection_volume_multi_object.cu:186-227, 523-600). */
int pii_volume_matrix_affine_derivatives(pii_context* ctx, int n, const pii_geometry* const* geoms,
                                         const double* const* dofs, const pii_kgrid* kg,
                                         int gram_mode, double* V, double* grad, double* H);

/* The same one-pass outputs for any DoF kinds (AffineDoF: 12 DoFs,
 * TriangleMeshDoF: 3 per DoF vertex), each output optional but not all:
 *   V[n*n]           volume matrix;
 *   grad[n * D]      D = sum_i ndofs[i]; object i's block is (n x ndofs[i])
 *                    row-major at offset n * o_i, o_i = sum_{i' < i} ndofs[i'],
 *                    entry (j, t) = dV_ij/dtheta_i,t (half the full
 *                    derivative for i == j), the same layout as
 *                    pii_volume_matrix_affine_derivatives when every object
 *                    has AffineDoF;
 *   H[D*D]           as pii_volume_matrix_hessian.
 * Calls with only AffineDoF objects take the fast derivative kernels; others
 * the FP64 per-node Jacobian of pii_form_factor_jacobian for every object.
 * D plus n is limited to 16384 (PII_ERR_INVALID above).
 * Replaces compute_intersection_volume_matrix_cuda with GRADIENT and/or
 * HESSIAN for any registered DoF kind: dense dA/dtheta, then grad_matrix and
 * the Gauss-Newton blocks from it (cpp/src/compute_intersection_volume_multi_object.cu:437-600,
 * cpp/src/compute_intersection_volume_multi_object.cpp:154-292). */
int pii_volume_matrix_derivatives(pii_context* ctx, int n, const pii_geometry* const* geoms,
                                  const int* dof_kinds, const double* const* dofs,
                                  const int* ndofs, const pii_kgrid* kg, int gram_mode, double* V,
                                  double* grad, double* H);

/* Adjoint: gradients of L = sum_ij dL_dV[i*n+j] V_ij with respect to every
 * object's DoFs, written to grads[i] (ndofs[i] doubles), without forming
 * dA/dtheta (vector-Jacobian product over the k-nodes).  V (optional, may be
 * NULL) receives the volume matrix of the same call (FP64 Gram of the fast
 * form factors).  Triangle geometry, AFFINE or TRIMESH DoFs.
 * Replaces the GRADIENT flag of compute_intersection_volume_matrix_cuda
 * (cpp/src/compute_intersection_volume_multi_object.cu:437-495), whose
 * grad_matrix[i][j] = dV_ij/dtheta_i is this call with dL_dV = e_i e_j^T
 * (and 2x for i == j, cpp/include/compute_intersection_volume_multi_object.hpp:22-26). */
int pii_volume_matrix_vjp(pii_context* ctx, int n, const pii_geometry* const* geoms,
                          const int* dof_kinds, const double* const* dofs, const int* ndofs,
                          const pii_kgrid* kg, const double* dL_dV, double* const* grads,
                          double* V);

/* Device-resident variants, no counterpart in the reference (whose drivers
 * take host Eigen vectors): for callers whose DoFs and results live in
 * device memory, a framework's tensors (pointint_amd.autograd).  dofs[i],
 * V, dL_dV and grads[i] are device memory of the context's device (else
 * PII_ERR_INVALID); the work is ordered after the caller's stream
 * (pii_context_set_caller_stream) and finished on return.  DoF values stay
 * on the device: the plan reads a device-computed bounding box of each
 * vertex-DoF object and the 12 values of each affine DoF vector; dL_dV
 * (n x n) is read to the host for the mixing matrix.  Gradients of vertex
 * DoFs of indexed meshes and of affine DoFs are written on the device;
 * meshes whose DoF vertices are shared ids (pii_geometry_create_tris) are
 * summed on the host.  Results equal the host-pointer entries' (V bit for
 * bit; affine gradients to rounding, their sums run in another order). */
int pii_volume_matrix_device(pii_context* ctx, int n, const pii_geometry* const* geoms,
                             const int* dof_kinds, const double* const* dofs, const int* ndofs,
                             const pii_kgrid* kg, int ff_precision, int gram_mode, double* V);
int pii_volume_matrix_vjp_device(pii_context* ctx, int n, const pii_geometry* const* geoms,
                                 const int* dof_kinds, const double* const* dofs,
                                 const int* ndofs, const pii_kgrid* kg, const double* dL_dV,
                                 double* const* grads, double* V);

/* Measurement helper, no counterpart in the reference: the device's
 * sustained VALU issue rate for the roofline figures of the benchmark.  kind
 * 0: FP64 FMA (flop/s), 1: FP32 FMA (flop/s, non-packed, as this library is
 * compiled), 2: FP32 hardware sine (per second).  Runs about 10-30 ms. */
int pii_measure_issue_rate(pii_context* ctx, int kind, double* per_second);

/* Gram of caller-provided form factors: J[n*Q] complex (object-major,
 * interleaved re/im, host memory) over the full grid of kg. */
int pii_gram(pii_context* ctx, const double* J, int n, const pii_kgrid* kg, int gram_mode,
             double* V);

/* ---- serving ------------------------------------------------------------------
 * No counterpart in the reference (its drivers are synchronous, one call at a
 * time, with per-call device allocation).  A server keeps the k-grid and the
 * geometries resident and answers volume-matrix queries -- a scene of n
 * resident triangle geometries, each under its own 12 affine DoFs -- by
 * continuous batching on its own HIP streams: batches of queries run as one
 * fused pass each (form factors of all their objects, one Gram launch over
 * all their column ranges), replayed from hipGraphs captured per batch
 * capacity (pii_server_set_graphs(srv, 0) launches the same kernels
 * directly; the results are identical).
 *
 * Admission: a query is refused at once with PII_ERR_BUDGET when its HBM
 * footprint does not fit hbm_budget bytes next to the queries admitted and
 * not finished, or, with a latency target set (pii_server_set_latency_target,
 * ms; 0 = none, the default), when its predicted completion -- time since
 * its arrival plus the outstanding work ahead of it and its own, priced by
 * the throughput the server measures -- exceeds the target.
 *
 * pii_server_submit_at takes the query's arrival time t_arrival (seconds on
 * the clock of pii_clock_seconds; <= 0: now); latencies are measured from it.
 * pii_server_wait blocks up to timeout_s seconds (< 0: no limit; PII_ERR_BUSY
 * on timeout), copies V (n*n) and latency_ms[0] (completion - arrival, ms),
 * and releases the query; pii_server_wait_times gives times[4] = arrival,
 * submission, launch, completion (seconds, same clock) instead.
 * pii_server_stats fills out[0..n): submitted, admitted, refused, completed,
 * batches, objects, reserved bytes, budget bytes, refused by memory, refused
 * by the latency target, batches launched from graphs, graphs captured, and
 * device-clock sums over the batches of the pre-pass, moment-series,
 * per-face form-factor and Gram phases (ms), the triangle x node evaluations
 * of the per-face kernel, the triangle x direction passes of the series, and
 * the current ms-per-triangle estimate of the latency model.
 * A geometry (or k-grid) used by a queued or running query cannot be
 * destroyed: pii_geometry_destroy returns PII_ERR_BUSY. */
typedef struct pii_server pii_server;
int pii_clock_seconds(double* t);
int pii_server_create(pii_context* ctx, const pii_kgrid* kg, int gram_mode, double hbm_budget,
                      int max_batch_objects, pii_server** out);
int pii_server_destroy(pii_server* srv);
int pii_server_set_latency_target(pii_server* srv, double target_ms);
int pii_server_set_graphs(pii_server* srv, int enable);
int pii_server_submit(pii_server* srv, int n, const pii_geometry* const* geoms,
                      const double* affine_dofs, int64_t* id);
int pii_server_submit_at(pii_server* srv, int n, const pii_geometry* const* geoms,
                         const double* affine_dofs, double t_arrival, int64_t* id);
int pii_server_wait(pii_server* srv, int64_t id, double timeout_s, double* V, double* latency_ms);
int pii_server_wait_times(pii_server* srv, int64_t id, double timeout_s, double* V,
                          double* times);
int pii_server_stats(pii_server* srv, double* out, int n);

#ifdef __cplusplus
}
#endif

#endif /* POINTINT_HIP_H */
