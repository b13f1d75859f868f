// Host-side batch state shared by the C-ABI entry points and the
// fast-path drivers.
#pragma once

#include <atomic>
#include <cstdint>
#include <exception>
#include <new>
#include <vector>

#include "device_common.hpp"
#include "pii_internal.hpp"

// Handles behind the opaque C types of pointint_hip.h.
struct pii_context {
  pii::Context c;
};
// A k-grid or geometry keeps its device id so that destroying it never
// touches its context (which the caller may have destroyed first); refs
// counts the serving queries (queued or running) that read it, and destroy
// refuses while any does.
struct pii_kgrid {
  pii::KGridDev k;
  pii_context* ctx = nullptr;
  int device = 0;
  std::atomic<int> refs{0};
};
struct pii_geometry {
  pii::GeometryDev g;
  pii_context* ctx = nullptr;
  int device = 0;
  std::atomic<int> refs{0};
};

namespace pii {

size_t align_up(size_t x, size_t a);

// Runs f, mapping exceptions to status codes and the thread-local message.
template <class F>
int guarded(F&& f) {
  clear_error();
  try {
    f();
    return PII_OK;
  } catch (const Error& e) {
    set_error(e.msg);
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed");
    return PII_ERR_NOMEM;
  } catch (const std::exception& e) {
    set_error(e.what());
    return PII_ERR_INVALID;
  }
}

// One call's objects on the device: DoFs uploaded, descriptors built, the
// pre-pass (world vertices + element records) launched on the context stream.
struct Batch {
  int N = 0;
  int NC = 0;                     // descriptors laid out (>= N; empty ones past N)
  size_t stage_bytes = 0;         // bytes of the staging upload
  std::vector<int> order;         // slot -> caller index (largest first)
  std::vector<ObjDesc> h_objs;    // per slot
  std::vector<int> h_vprefix, h_eprefix;
  const ObjDesc* d_objs = nullptr;
  const int* d_vprefix = nullptr;
  const int* d_eprefix = nullptr;
  int total_v = 0, total_e = 0;
  // Moment series plan (fast path, shell-structured grid): objects with a
  // series, directions per shell, and the smallest first node of the
  // per-face kernel over all objects.
  int nseries = 0, Mh = 0, qbeg_min = 0;
  // Set before build: triangle objects of geometries with cached
  // spherical-harmonic coefficients, under similarity transforms, take the
  // harmonic form of the series (nsh of them).
  bool use_sh = false;
  int nsh = 0;
  // Set before build for device-resident DoFs: dofs[i] are device pointers
  // of the context's device, copied device-to-device after the staging
  // upload; the plan reads plan_dofs[i] (a host copy, affine objects) and
  // plan_bounds[4 i ..] (series centre and radius, vertex-DoF objects)
  // instead of the DoF values.
  bool device_dofs = false;
  const double* const* plan_dofs = nullptr;
  const double* plan_bounds = nullptr;
  // prof (may be null) records the pre-pass phase with HIP events.  series_kg
  // (fast path only, may be null) enables the moment series on that grid.
  // cols (may be null: column i) gives object i's output column.  n_cap > n
  // lays the upload out for n_cap descriptors (empty ones past n) with 32
  // DoF doubles per object of room.  enqueue_now = false only prepares the
  // pinned staging (w.stage, stage_bytes bytes, for w.misc) and the device
  // pointers: the caller uploads it and launches the pre-pass.
  void build(Workspace& w, hipStream_t s, Context* prof, int n, const GeometryDev* const* geoms,
             const int* kinds, const double* const* dofs, const int* ndofs, bool fast,
             const KGridDev* series_kg = nullptr, const int* cols = nullptr, int n_cap = 0,
             bool enqueue_now = true);
  void build(Context& c, int n, const GeometryDev* const* geoms, const int* kinds,
             const double* const* dofs, const int* ndofs, bool fast,
             const KGridDev* series_kg = nullptr) {
    build(c.w, c.stream, &c, n, geoms, kinds, dofs, ndofs, fast, series_kg);
  }
};

// Device-output entries: the caller's pointer must be device memory of the
// context's device; the context stream waits for the caller's stream.
void require_device_pointer(const void* p, int device, const char* who);
void order_after_caller(pii_context* ctx);
// A(k_q) of one object over the full grid of kg into out (host or device
// memory by kind), on c's stream and workspace; synchronous.
void form_factor_to(Context& c, const pii_geometry* g, int dof_kind, const double* dofs,
                    int ndofs, const pii_kgrid* kg, int precision, double* out,
                    hipMemcpyKind kind);

// Fast form factors of a batch: the moment series for the leading shells of
// triangle objects, the per-face kernel for the remaining nodes.  prof (may be
// null) records the two kernels' phases and the work counters.
void launch_form_factors_fast(hipStream_t s, const Batch& b, const KGridDev& kg,
                              int out_mode, double2* d_A, int ldA, float2* d_X, int ldX,
                              unsigned* d_colmax, Workspace& w, Context* prof);

// Split-precision Gram (fp16: FP16x3, else BF16x3) of the column ranges qs
// (from gram_query; offsets filled here) of X (FP32 complex, ldX nodes per
// column, ncols columns, column maxima d_colmax): V of query q at d_V + q.v.
// Enqueued on s; the plan goes up through w's pinned staging, which the next
// call rewrites (the caller synchronises s between calls).
void gram_split_run(hipStream_t s, Workspace& w, int fp16, const float2* d_X, int ldX, int ncols,
                    const unsigned* d_colmax, std::vector<GramQuery>& qs, double* d_V);

// Fast path (FP32 transcendentals, FP64 phases and Taylor branch).
// out: host memory (kind hipMemcpyDeviceToHost) or device memory
// (hipMemcpyDeviceToDevice), Q x N complex.
void form_factor_fast_full(Context& c, const Batch& b, const KGridDev& kg, double* out,
                           hipMemcpyKind kind);
void volume_matrix_fast(Context& c, const Batch& b, const KGridDev& kg, int gram_mode,
                        double* V_host);

// Gauss-Newton Hessian blocks H (D x D, D = sum of ndofs, row-major; block
// (i, j) rows from sum_{i' < i} ndofs[i'], columns likewise, caller order);
// V (may be null) of the same call.  Affine-only batches use the fast
// derivative kernels, others the FP64 Jacobian kernels.
void volume_matrix_hessian(Context& c, const Batch& b, const KGridDev& kg, int gram_mode,
                           const int* ndofs, double* H_host, double* V_host);

// V (N x N), per-pair gradients and Gauss-Newton blocks for any DoF kinds from
// one Gram over the A and derivative columns (each output may be null).
// grad: object i's (N x ndofs[i]) block at offset N sum_{i' < i} ndofs[i'],
// entry (j, t) = dV_ij/dtheta_i,t (half the full derivative for i == j); H
// as volume_matrix_hessian.  Affine-only batches take the fast derivative
// kernels (the same layouts with 12 DoFs per object).
void volume_matrix_derivatives(Context& c, const Batch& b, const KGridDev& kg, int gram_mode,
                               const int* ndofs, double* V_host, double* grad_host,
                               double* H_host);

// V (N x N), grad[(i N + j) 12 + t] = dV_ij/dtheta_i,t (half for i == j) and
// the Gauss-Newton H of affine-DoF objects from one Gram over the A and
// dA/dtheta columns; every output may be null.
void volume_matrix_affine_all(Context& c, const Batch& b, const KGridDev& kg, int gram_mode,
                              double* V_host, double* grad_host, double* H_host);

// Vertex-DoF / affine-DoF gradients of L = sum_ij G_ij V_ij (triangles).
void volume_matrix_vjp(Context& c, const Batch& b, const GeometryDev* const* geoms,
                       const KGridDev& kg, const double* G, double* const* grads,
                       double* V_host, bool grads_on_device = false);

}  // namespace pii
