// Internal declarations shared by the host-side C-ABI implementation and the
// HIP kernel translation units of libpointint_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "pointint_hip.h"

namespace pii {

constexpr double kPi = 3.14159265358979323846;
// 1 / (8 pi^3): the volume normalisation of V = Re(J^H D J) / (8 pi^3).
constexpr double kInv8Pi3 = 1.0 / (8.0 * kPi * kPi * kPi);

// ---------------------------------------------------------------------------
// Errors: every entry point returns a status and records a thread-local
// message readable through pii_last_error().
// ---------------------------------------------------------------------------
void set_error(const std::string& msg);
void clear_error();

#define PII_HIP(call)                                                          \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess) {                                                    \
      throw ::pii::Error(PII_ERR_HIP, std::string(#call " failed: ") +         \
                                          hipGetErrorString(e_));              \
    }                                                                          \
  } while (0)

struct Error {
  int code;
  std::string msg;
  Error(int c, std::string m) : code(c), msg(std::move(m)) {}
};

// ---------------------------------------------------------------------------
// A fixed region (device or host memory the library does not allocate) that
// buffers bound to it carve their space from, in order, without freeing:
// the workspace of the allocation-free entry point.
// ---------------------------------------------------------------------------
struct Arena {
  char* base = nullptr;
  size_t cap = 0, used = 0;
  const char* what = "scratch";
  // 256-byte aligned space of n bytes; throws PII_ERR_INVALID when the
  // region is exhausted.
  void* take(size_t n);
};

// ---------------------------------------------------------------------------
// Device buffer owned by the library (hipMalloc / hipFree), or carved from an
// Arena when bound to one (never allocated or freed then).
// ---------------------------------------------------------------------------
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  Arena* arena = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes), arena(o.arena) {
    o.p = nullptr; o.bytes = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release(); p = o.p; bytes = o.bytes; arena = o.arena; o.p = nullptr; o.bytes = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void release();
  // Grow-only: keeps the allocation when it is already large enough.
  void ensure(size_t n);
  template <class T> T* as() const { return static_cast<T*>(p); }
};

// ---------------------------------------------------------------------------
// k-grid resident in HBM.
//
// Full grid: node q = ir*M + j (radial-major, as the reference builds it).
// Hemisphere grid: when every direction of every shell has an antipode with
// the same weight, A(-k) = -conj(A(k)) makes the +k and -k terms of the Gram
// identical, so the Gram and the adjoint run over one representative per pair
// with the pair's summed weight.  The representatives keep shell-major order
// so a 64-node tile mostly shares one |k|.
// ---------------------------------------------------------------------------
struct KGridDev {
  int Q = 0;                    // full node count
  int Qh = 0;                   // representative count (== Q when not symmetric)
  bool symmetric = false;
  std::vector<double> h_dirs, h_kmag, h_w;  // host copies of the full grid
  std::vector<int> h_rep;       // representative -> full index
  std::vector<int> h_anti;      // full index -> antipode full index (or -1)
  // Full grid, SoA (FP64): kx, ky, kz, dir.x, dir.y, dir.z, kmag, w
  DevBuf full;                  // 8*Q doubles
  // Representative grid, same SoA with w replaced by the pair's summed weight
  DevBuf rep;                   // 8*Qh doubles
  DevBuf rep_index;             // Qh int32: rep -> full index
  DevBuf anti_index;            // Q int32: full -> antipode
  // Shell structure of the representative grid (0 when absent): Qh = nshell*Mh
  // and node s*Mh + j has magnitude h_shell_k[s] and the direction of node j.
  int nshell = 0, Mh = 0;
  std::vector<double> h_shell_k;
};

// Detects the shell structure of a k-grid's representative nodes.
void kgrid_shells(KGridDev& k);

// ---------------------------------------------------------------------------
// Geometry resident in HBM.
// ---------------------------------------------------------------------------
// Faces of a triangle mesh are grouped into patches of at most kPatchVerts
// distinct vertices and kPatchFaces faces, so that a workgroup can hold the
// per-(vertex, k) phasors of one patch in LDS and share them between faces.
constexpr int kPatchVerts = 64;
constexpr int kPatchFaces = 128;

struct GeometryDev {
  int type = 0;     // PII_GEOM_*
  int n = 0;        // element count (faces / disks / Gaussians)
  // ---- triangles ----
  int nv = 0;                        // geometry vertices
  std::vector<double> h_verts;       // nv*3 reference-frame vertex positions
  std::vector<int32_t> h_faces;      // n*3 indices into h_verts
  std::vector<int32_t> h_dofvid;     // nv: geometry vertex -> DoF vertex (or -1)
  // Bounding-box centre and radius (max |x - ref_c|) of the reference vertices
  // (triangles) or element centres (disks / Gaussians), and the largest
  // rho / sigma of disks / Gaussians: the moment series' reach.
  double ref_c[3] = {0, 0, 0};
  double ref_R = 0.0;
  double ref_smax = 0.0;
  int ndof_verts = 0;                // DoF vertex count for TRIMESH DoFs
  bool dof_identity = false;         // DoF vertex v is geometry vertex v (indexed meshes)
  DevBuf verts;                      // nv*3 FP64 (AoS)
  DevBuf faces;                      // n*3 int32
  DevBuf dofvid;                     // nv int32
  // patches
  int npatch = 0;
  DevBuf patch_hdr;                  // npatch * 4 int32: vbeg, vcnt, fbeg, fcnt
  DevBuf patch_verts;                // sum vcnt int32 (geometry vertex ids)
  DevBuf patch_faces;                // n uint32: face id (low 24 bits... see below)
  DevBuf patch_floc;                 // n uint32: local corner indices (3 x 8 bit)
  // ---- disks / Gaussians: n x 8 FP64 (c.xyz, n.xyz, rho|sigma, area|w) ----
  std::vector<double> h_el;
  DevBuf el;
  // ---- spherical-harmonic coefficients of the moment series (triangles,
  // reference frame, about ref_c; empty until cached) ----
  DevBuf shc;
};

// Pinned host staging (hipHostMalloc), grow-only: the upload of a call's DoFs
// and descriptors is one DMA copy from here.  Each workspace (and so each
// stream) has its own; a call rewrites it only after its stream finished the
// previous call's copy.
struct HostBuf {
  void* p = nullptr;
  size_t bytes = 0;
  Arena* arena = nullptr;   // carve from a host region instead of hipHostMalloc
  HostBuf() = default;
  HostBuf(const HostBuf&) = delete;
  HostBuf& operator=(const HostBuf&) = delete;
  ~HostBuf() { release(); }
  void release();
  void ensure(size_t n);
  template <class T> T* as() const { return static_cast<T*>(p); }
};

// Workgroups of `kernel` (block threads, no dynamic LDS) resident at once on
// the current device: compute units x the occupancy the runtime reports for
// the kernel's registers and LDS.  Computed once per device id (cache[16]).
template <class K>
int resident_workgroups(K kernel, int block, std::atomic<int>* cache) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 16) {
    (void)hipGetLastError();
    return 256;
  }
  int v = cache[dev].load(std::memory_order_relaxed);
  if (v > 0) return v;
  int cus = 0, per_cu = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess) {
    (void)hipGetLastError();
    return 256;
  }
  v = std::max(1, cus) * std::max(1, per_cu);
  cache[dev].store(v, std::memory_order_relaxed);
  return v;
}

// Element split of a launch with few workgroups: the smallest split s (at
// most smax) that gives at least two rounds of the kernel's resident
// workgroups (so that workgroups finishing early are followed by others)
// filled to at least 90 %, else the one that fills its rounds best.  A few
// objects on a large grid split many ways; hundreds of objects not at all.
inline int split_rounds(int base_wgs, int slots, int smax) {
  if (base_wgs <= 0 || slots <= 0 || smax <= 1) return 1;
  int best = 1;
  double best_eff = 0.0;
  for (int s = 1; s <= smax; ++s) {
    const double n = (double)base_wgs * s / slots;
    const double eff = n / std::ceil(n);
    if (n >= 2.0 && eff >= 0.9) return s;
    const double score = std::min(n, 2.0) / 2.0 * eff;  // rounds below two count against
    if (score > best_eff + 1e-12) {
      best = s;
      best_eff = score;
    }
  }
  return best;
}

// Largest split that keeps at least min_per elements per workgroup (and at
// most 64 splits).
inline int max_split(int max_elems, int min_per) {
  return std::max(1, std::min(64, max_elems / std::max(1, min_per)));
}

// Per-call device workspaces (grow-only, reused across calls on one stream).
struct Workspace {
  DevBuf world, faces, J, X, part, V, misc, colmax, B, C, gw, mrec, gam;
  DevBuf dth, Xt, cmax2, H;  // Gauss-Newton Hessian: records, dA/dtheta columns, maxima, H
  DevBuf dsrec, dspart, dsval;  // derivative series: face records, family sums, shell values
  DevBuf split;              // partial sums of element-split launches (few objects)
  DevBuf planes, gplan;      // split-precision Gram: 16-bit planes of X, device plan
  HostBuf gplan_h;           // pinned: the Gram plan of a call
  HostBuf stage;             // pinned: descriptors, prefix sums and DoFs of a call
  HostBuf hout;              // pinned: per-vertex gradients coming back
  // Every buffer takes its space from dev (device) / host instead of
  // allocating (the allocation-free entry point).
  void bind(Arena* dev, Arena* host) {
    for (DevBuf* b : {&world, &faces, &J, &X, &part, &V, &misc, &colmax, &B, &C, &gw, &mrec,
                      &gam, &dth, &Xt, &cmax2, &H, &dsrec, &dspart, &dsval, &split, &planes,
                      &gplan})
      b->arena = dev;
    for (HostBuf* b : {&gplan_h, &stage, &hout}) b->arena = host;
  }
};

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  // The caller's stream (nullptr: the null stream) that the device-output
  // entries order their writes after, and an event to do so.
  hipStream_t caller_stream = nullptr;
  hipEvent_t caller_ev = nullptr;
  std::mutex mu;                     // serialises calls on one context
  Workspace w;
  // Moment series for low-|k| shells of triangle objects on the fast path.
  bool series = true;
  const KGridDev* series_grid(const KGridDev& k) const {
    return series ? &k : nullptr;
  }
  // Phase timing with HIP events on the context stream (enable_profiling of
  // the reference drivers, e.g. compute_intersection_volume_multi_object.cu:626-695).
  bool profiling = false;
  // 0 pre-pass, 1 form factor, 2 Gram, 3 adjoint; inside 1: 4 moment series
  // kernel, 5 per-face kernel.
  static constexpr int kPhases = 6;
  hipEvent_t ev[kPhases][2] = {};
  bool ev_used[kPhases] = {};
  void phase_begin(int p);
  void phase_end(int p);
  // Work of the last call's form-factor kernels: [0] triangle x node
  // evaluations of the per-face kernel, [1] face x direction passes of the
  // moment series, [2] object x node series evaluations, [3] disk/Gaussian x
  // node evaluations of the per-element kernel, [4] disk/Gaussian x direction
  // passes of the moment series.
  static constexpr int kWork = 5;
  double work[kWork] = {};
  void begin_call() {
    for (bool& u : ev_used) u = false;
    for (double& x : work) x = 0.0;
  }
};

// ---------------------------------------------------------------------------
// Kernel launchers (defined in the .hip translation units).
// ---------------------------------------------------------------------------

struct ObjDesc;

// Pre-pass over all objects of a call: world vertices, then element records
// (fast != 0: patch-ordered triangle records for the shared-vertex kernel).
void launch_prepass(hipStream_t s, const ObjDesc* d_objs, const int* d_vprefix, int total_v,
                    const int* d_eprefix, int total_e, int N, int fast);

// The same pre-pass from a fixed grid (grid-stride loops; totals read from
// the device prefix sums d_vprefix[N], d_eprefix[N]): replayable in a graph
// for any batch laid out with N descriptors.
void launch_prepass_fixed(hipStream_t s, const ObjDesc* d_objs, const int* d_vprefix,
                          const int* d_eprefix, int N, int fast, int grid);
// Bounds of vertex-DoF objects whose DoFs are in device memory: per job,
// out[5 j ..] = {centre, max |x - centre|^2, valid} over the DoF vertices the
// geometry's vertices refer to (the host plan's world bounds).
struct DofBoundsJob {
  const int32_t* dofvid;   // nv entries: geometry vertex -> DoF vertex (or -1)
  const double* dofs;      // 3 x DoF vertices, device memory
  int nv;
  int pad_;
};
void launch_dof_bounds(hipStream_t s, const DofBoundsJob* d_jobs, int n, double* d_out);
// dL/dtheta of affine objects from their world-vertex gradients gW:
// out[0..2] = sum_v gW_v, out[3 + 3 r + c] = sum_v gW_v,r V_v,c (reference
// vertices V), one workgroup per job, fixed-order sums.
struct AffineGradJob {
  const double* gw;      // 3 nv world-vertex gradients
  const double* verts;   // 3 nv reference vertices
  int nv;
  int pad_;
  double* out;           // 12 doubles, device memory
};
void launch_affine_dof_grads(hipStream_t s, const AffineGradJob* d_jobs, int n);
// Zeroes n 32-bit words (a kernel: graph-safe, unlike a memset node).
void launch_clear_u32(hipStream_t s, unsigned* d_p, int n);
// Writes the device's constant 100 MHz clock to d_out[i] when the stream
// reaches it (phase timing inside captured graphs).
void launch_stamp(hipStream_t s, unsigned long long* d_out, int i);

// Plain FP64 form factor (reference formula) over the full grid:
// J[col*ldJ + q] = A_col(k_q), complex FP64.
void launch_ff_fp64(hipStream_t s, const ObjDesc* d_objs, int N, const double* d_kfull, int Q,
                    double2* d_J, int ldJ, int nsplit = 1, double2* d_part = nullptr);
// Element split of the plain FP64 form factors for N objects (the largest
// with max_elems elements) on Q nodes, and the partial-sum buffer it needs
// (double2 values; none when the split is 1).  The partial sums are kept to
// 8 MiB, so the split is the same wherever the workspace comes from.
int ff_fp64_split(int N, int Q, int max_elems);
size_t ff_fp64_split_values(int N, int Q, int nsplit);

// Plain FP64 Gram over the full grid: V = Re(J^H diag(w) J)/(8 pi^3).
int gram_fp64_splits(int Q, int N);
void launch_gram_fp64(hipStream_t s, const double2* d_J, int Q, int N, int ldJ,
                      const double* d_w, double* d_part, int nsplit, double* d_V);

// FP64 Gram of fast-path X (float2, sqrt(w) folded in, Qh representative nodes).
void launch_gram_fp64_x(hipStream_t s, const float2* d_X, int Qh, int N, int ldX, double* d_part,
                        int nsplit, double* d_V);

// Fast form factor over the representative nodes.  out_mode 0: A (FP64) to
// d_A[col*ldA + q]; out_mode 1: X = sqrt(w) A (FP32) to d_X[col*ldX + q] for
// q < ldX (zeros beyond Qh) and the column max |X| to d_colmax[col].
// Each object starts at its own representative node ObjDesc::qbeg (nodes
// below come from the moment series); qbeg_min is the smallest over objects.
// nsplit > 1 (with d_part of ff_split_doubles(...) doubles) splits every
// object's elements over nsplit workgroups per node block and sums their
// partial results in a finishing kernel: the form of a launch with few
// objects, which would otherwise occupy a few CUs.
void launch_ff_fast(hipStream_t s, const ObjDesc* d_objs, int N, const double* d_krep, int Qh,
                    int qbeg_min, int out_mode, double2* d_A, int ldA, float2* d_X, int ldX,
                    unsigned* d_colmax, int nsplit = 1, double* d_part = nullptr);
size_t ff_split_doubles(int N, int span, int nsplit);
// Moment series of triangle objects at their first nser shells (objects with
// nser = 0 are skipped), each at its own order nord; outputs as
// launch_ff_fast; nsplit as launch_ff_fast with d_part of
// series_split_doubles(...) doubles.
void launch_series_tri(hipStream_t s, const ObjDesc* d_objs, int N, const double* d_krep, int Qh,
                       int Mh, int out_mode, double2* d_A, int ldA, float2* d_X, int ldX,
                       unsigned* d_colmax, int nsplit = 1, double* d_part = nullptr);
// The series of triangle objects with cached spherical-harmonic coefficients
// (ObjDesc::shc), which launch_series_tri skips.
void launch_series_sh(hipStream_t s, const ObjDesc* d_objs, int N, const double* d_krep, int Qh,
                      int Mh, int out_mode, double2* d_A, int ldA, float2* d_X, int ldX,
                      unsigned* d_colmax);
// Raw sums H_n (n <= kSeriesOrder; descriptors of order kSeriesOrder) of the
// triangle objects at Md directions (rows 3..5 of d_dirs, row stride Qd):
// d_H[((obj * nblk + blk) * (kSeriesOrder + 1) + n) * 64 + lane] for direction
// blk * 64 + lane, nblk = ceil(Md / 64).
void launch_series_tri_raw(hipStream_t s, const ObjDesc* d_objs, int N, const double* d_dirs,
                           int Qd, int Md, double* d_H);
// Spherical-harmonic coefficients of the moment series of triangle
// geometries (reference frame, about ref_c) into each one's shc buffer; the
// geometries with fewer than kShMinFaces faces are left without (the direct
// series is cheaper for them).  Synchronous on c's stream.
constexpr int kShMinFaces = 192;
void cache_moments(Context& c, int n, GeometryDev* const* gs);
// The same for disk and Gaussian objects.
void launch_series_surfel(hipStream_t s, const ObjDesc* d_objs, int N, const double* d_krep,
                          int Qh, int Mh, int out_mode, double2* d_A, int ldA, float2* d_X,
                          int ldX, unsigned* d_colmax, int nsplit = 1, double* d_part = nullptr);
size_t series_split_doubles(int N, int Mh, int nsplit);
// Affine-DoF derivatives for the Gauss-Newton Hessian: per-face records
// (reference corners + N, 18 floats, patch order, at d_off[slot]) and the
// columns X_t = sqrt(w) dA/dtheta_t (12 per object, column 12 col + t).
void launch_dtheta_records(hipStream_t s, const ObjDesc* d_objs, const int* d_eprefix, int N,
                           int total_e, float* d_out, const long long* d_off);
void launch_dtheta_affine(hipStream_t s, const ObjDesc* d_objs, int N, const double* d_krep,
                          int Qh, const float* d_drec, const long long* d_doff, float2* d_Xt,
                          int ldX, unsigned* d_colmax, int nsplit = 1, double* d_part = nullptr);
size_t dtheta_split_doubles(int N, int ldX, int nsplit);
// dA(k_q)/dtheta of one object (d_obj[0]) at every node of the full grid,
// FP64, row-major Q x D complex.
void launch_jacobian(hipStream_t s, const ObjDesc* d_obj, int type, int dof_kind, int n_elems,
                     const double* d_kfull, int Q, int D, double2* d_out);
// The same Jacobian on the representative nodes (krep, Qh of them) as FP32
// columns sqrt(w) dA/dtheta_t of X (leading dimension ldX, column maxima
// colmax), through the FP64 column-major scratch d_J (ldJ >= Qh, D columns).
void launch_jacobian_columns(hipStream_t s, const ObjDesc* d_obj, int dof_kind, int n_elems,
                             const double* d_krep, int Qh, int D, double2* d_J, int ldJ,
                             float2* d_X, int ldX, unsigned* d_colmax);
// Most columns (A columns and derivative columns together) of one derivative
// Gram: its result alone is D x D doubles (2 GiB at the limit).
constexpr long long kMaxDerivativeColumns = 16384;
// Resident workgroups of the element-splittable kernels (resident_workgroups).
int ff_fast_slots();
int series_tri_slots();
int series_surfel_slots();
int dtheta_slots();
int dseries_slots();
// Derivative series of affine triangle objects (ObjDesc::dser): per-face
// reference records, the five families of per-direction sums and the 12
// columns on the series shells (nodes q < qbeg); the per-element kernels then
// start at qbeg for those objects.
void launch_dseries_records(hipStream_t s, const ObjDesc* d_objs, const int* d_eprefix, int N,
                            int total_e, double* d_out, const long long* d_off);
size_t dseries_record_doubles(int nfaces);
size_t dseries_part_doubles(int N, int Mh, int nsplit);
size_t dseries_value_count(int qbeg);
void launch_dseries(hipStream_t s, const ObjDesc* d_objs, int N, const double* d_krep, int Qh,
                    int Mh, const double* d_srec, const long long* d_soff, double* d_part,
                    int nsplit, double2* d_sval, const long long* d_voff, int max_qbeg,
                    float2* d_Xt, int ldX, unsigned* d_colmax);
// Sustained issue rate of one probe launch (blocks x 256 threads, iters
// rounds): kind 0 FP64 FMA flop/s, 1 FP32 FMA flop/s, 2 FP32 hardware sines/s.
// d_sink: blocks doubles (written only under a condition that never holds).
double probe_issue_rate(hipStream_t s, int kind, int blocks, int iters, void* d_sink);
// Full-grid values from representative values: A(-k) = -conj(A(k)).
void launch_expand(hipStream_t s, const double2* d_Arep, int ldA, const int* d_rep,
                   const int* d_anti, int Qh, int Q, int N, double2* d_Afull);

// Split-precision MFMA Gram (fp16 != 0: FP16x3, two FP16 pieces with column
// scaling and three products; else BF16x3, three BF16 pieces and six
// products).  One launch computes the Grams of a list of column ranges
// ("queries") of the same planes; the device plan is a GramPlanHeader
// followed by its nq GramQuery records.
struct GramQuery {
  int c0;            // first column (a multiple of 16)
  int n;             // columns
  int ntiles;        // 128 x 128 upper-triangle tiles
  int nslice;        // K slices
  int slice_len;     // K steps (32 values) per slice
  int pad_;
  long long item0;   // first work item (tile x slice) in the launch
  long long entry0;  // first V entry (n x n per query) in the launch
  long long part;    // offset of the query's partials (nslice n^2 doubles)
  long long v;       // offset of the query's V (n^2 doubles)
};
struct GramPlanHeader {
  int nq;
  int ncols;         // columns the queries use (c0 + n of the last)
  long long nitems, nentries, part_doubles;
};
int gram_planes_pieces(int fp16);
// Bytes of the planes of ncols columns of K values (K a multiple of 32).
size_t gram_plane_bytes(int fp16, int K, int ncols);
// One query over columns [c0, c0 + n) with K values per column: tiles and
// K slices (a function of n and K only, so a query's result does not depend
// on the launch it is part of).
GramQuery gram_query(int c0, int n, int K);
// Offsets of the queries' work items, partials and outputs; fills h.
void gram_plan_layout(std::vector<GramQuery>& qs, GramPlanHeader& h);
// Workgroups of a launch over nitems work items (at most cap; each loops).
int gram_planes_grid(long long nitems, int cap);
// Planes laid out for ncols columns; with d_plan (may be null) only the
// columns its queries use are converted.
void launch_split_planes(hipStream_t s, int fp16, const float* d_X, int ldK, int ncols,
                         const unsigned* d_colmax, void* d_planes,
                         const GramPlanHeader* d_plan = nullptr);
void launch_gram_planes(hipStream_t s, int fp16, const void* d_planes, int ldK, int ncols,
                        const GramPlanHeader* d_plan, int grid_items, long long nentries_cap,
                        double* d_part, const unsigned* d_colmax, double* d_V);

// Adjoint: Y = X C (complex columns mixed by a real N x N matrix), then the
// per-face vertex VJP of triangle objects into world-vertex gradients
// d_gworld + d_goff[slot] (FP64, zeroed by the caller, atomically accumulated).
void launch_mix_columns(hipStream_t s, const float2* d_X, int N, int ldX, int Qh,
                        const double* d_C, float2* d_Y);
// Nodes q >= ObjDesc::qbeg of each object (qbeg_min: smallest over objects).
void launch_adjoint_tri(hipStream_t s, const ObjDesc* d_objs, int nobj, int max_patches,
                        const double* d_krep, int Qh, int qbeg_min, const float2* d_Bw, int ldB,
                        int ksplit, double* d_gworld, const int* d_goff);
// Adjoint of the moment series (nodes below qbeg of objects with nser > 0);
// d_gam: N * Mh * kGammaStride doubles of scratch.
void launch_series_adjoint_tri(hipStream_t s, const ObjDesc* d_objs, int N, int max_faces,
                               const double* d_krep, int Qh, int Mh, const float2* d_Bw, int ldB,
                               double* d_gam, double* d_gworld, const int* d_goff);

}  // namespace pii
