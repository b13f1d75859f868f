// C-ABI implementation of libpointint_hip.so: context, resident k-grids and
// geometries, and the form-factor / volume-matrix entry points.  Host code
// only; kernels live in the .hip translation units.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <exception>
#include <memory>
#include <new>
#include <numeric>
#include <string>
#include <unordered_map>
#include <vector>

#include "device_common.hpp"
#include "pii_host.hpp"
#include "pii_internal.hpp"
#include "pii_capi_internal.hpp"

namespace pii {

void* Arena::take(size_t n) {
  const size_t at = align_up(used, 256);
  if (at > cap || n > cap - at)
    throw Error(PII_ERR_INVALID, std::string(what) + " exhausted: the call needs more than its " +
                                     std::to_string(cap) + " bytes");
  used = at + n;
  return base + at;
}

void DevBuf::release() {
  if (arena) {       // carved from a region: nothing to free
    p = nullptr;
    bytes = 0;
    return;
  }
  if (p) {
    (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
}

// Allocation fault injection for the error-path tests
// (pii_debug_fail_allocation): after `g_fail_after` more successful
// allocations the next one fails as an out-of-memory would.  -1: off.
static std::atomic<int> g_fail_after{-1};
static bool injected_allocation_failure() {
  int v = g_fail_after.load();
  while (v >= 0) {
    if (g_fail_after.compare_exchange_weak(v, v - 1)) return v == 0;
  }
  return false;
}

void DevBuf::ensure(size_t n) {
  if (n <= bytes && p) return;
  release();
  if (n == 0) return;
  if (arena) {
    p = arena->take(n);
    bytes = n;
    return;
  }
  const size_t want = std::max<size_t>(n, 256);
  if (injected_allocation_failure())
    throw Error(PII_ERR_NOMEM, "hipMalloc of " + std::to_string(want) +
                                   " bytes failed (injected for testing)");
  if (hipMalloc(&p, want) != hipSuccess) {
    p = nullptr;
    (void)hipGetLastError();
    throw Error(PII_ERR_NOMEM, "hipMalloc of " + std::to_string(want) + " bytes failed");
  }
  bytes = want;
}

void HostBuf::release() {
  if (arena) {
    p = nullptr;
    bytes = 0;
    return;
  }
  if (p) {
    (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
  }
}

void HostBuf::ensure(size_t n) {
  if (n <= bytes && p) return;
  release();
  if (n == 0) return;
  if (arena) {
    p = arena->take(n);
    bytes = n;
    return;
  }
  const size_t want = std::max<size_t>(n, 4096);
  if (injected_allocation_failure())
    throw Error(PII_ERR_NOMEM, "hipHostMalloc of " + std::to_string(want) +
                                   " bytes failed (injected for testing)");
  if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) {
    p = nullptr;
    (void)hipGetLastError();
    throw Error(PII_ERR_NOMEM, "hipHostMalloc of " + std::to_string(want) + " bytes failed");
  }
  bytes = want;
}

void Context::phase_begin(int p) {
  if (!profiling) return;
  if (!ev[p][0]) {
    PII_HIP(hipEventCreate(&ev[p][0]));
    PII_HIP(hipEventCreate(&ev[p][1]));
  }
  PII_HIP(hipEventRecord(ev[p][0], stream));
}

void Context::phase_end(int p) {
  if (!profiling) return;
  PII_HIP(hipEventRecord(ev[p][1], stream));
  ev_used[p] = true;
}

static void require(bool cond, const std::string& msg) {
  if (!cond) throw Error(PII_ERR_INVALID, msg);
}

template <class T>
static void upload(DevBuf& b, const T* src, size_t count, hipStream_t s) {
  b.ensure(count * sizeof(T));
  if (count) PII_HIP(hipMemcpyAsync(b.p, src, count * sizeof(T), hipMemcpyHostToDevice, s));
}

// Shell structure of the representative nodes: the first run of equal |k|
// gives Mh; every later block of Mh nodes must repeat those directions
// exactly at one magnitude of its own.
void kgrid_shells(KGridDev& k) {
  k.nshell = 0;
  k.Mh = 0;
  k.h_shell_k.clear();
  const int Qh = k.Qh;
  if (Qh <= 0) return;
  auto kq = [&](int r) { return k.h_kmag[k.h_rep[r]]; };
  auto dq = [&](int r, int c) { return k.h_dirs[3 * (size_t)k.h_rep[r] + c]; };
  int Mh = 1;
  while (Mh < Qh && kq(Mh) == kq(0)) ++Mh;
  if (Qh % Mh != 0) return;
  const int ns = Qh / Mh;
  std::vector<double> shell_k(ns);
  for (int sh = 0; sh < ns; ++sh) {
    shell_k[sh] = kq(sh * Mh);
    for (int j = 0; j < Mh; ++j) {
      const int r = sh * Mh + j;
      if (kq(r) != shell_k[sh]) return;
      for (int c = 0; c < 3; ++c)
        if (dq(r, c) != dq(j, c)) return;
    }
  }
  k.nshell = ns;
  k.Mh = Mh;
  k.h_shell_k = std::move(shell_k);
}

// ---------------------------------------------------------------------------
// Batch set-up shared by every multi-object entry point.
// ---------------------------------------------------------------------------
size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

void validate_object(const GeometryDev& g, int dof_kind, int ndofs) {
  // messages are built only on failure (this runs per object per call)
  if (dof_kind == PII_DOF_AFFINE) {
    if (ndofs != 12)
      throw Error(PII_ERR_INVALID, "AffineDoF requires 12 DoFs, got " + std::to_string(ndofs));
  } else if (dof_kind == PII_DOF_TRIMESH) {
    if (g.type != PII_GEOM_TRIANGLE)
      throw Error(PII_ERR_UNSUPPORTED, "TriangleMeshDoF supports triangle geometry only");
    if (ndofs != 3 * g.ndof_verts)
      throw Error(PII_ERR_INVALID, "TriangleMeshDoF requires " +
                                       std::to_string(3 * g.ndof_verts) + " DoFs, got " +
                                       std::to_string(ndofs));
  } else {
    throw Error(PII_ERR_UNSUPPORTED, "unknown DoF kind " + std::to_string(dof_kind));
  }
}

// Largest singular value of a row-major 3x3 matrix: square root of the top
// eigenvalue of A^T A (closed form for symmetric 3x3), rounded up slightly so
// that it bounds |A x| / |x| despite rounding.
static double spectral_norm3(const double* A) {
  double B[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      B[i][j] = A[i] * A[j] + A[3 + i] * A[3 + j] + A[6 + i] * A[6 + j];
  double fro2 = 0.0;
  for (int i = 0; i < 9; ++i) fro2 += A[i] * A[i];
  const double p1 = B[0][1] * B[0][1] + B[0][2] * B[0][2] + B[1][2] * B[1][2];
  const double q = (B[0][0] + B[1][1] + B[2][2]) / 3.0;
  double lam;
  if (p1 <= 1e-300 * fro2 * fro2) {
    lam = std::max(B[0][0], std::max(B[1][1], B[2][2]));
  } else {
    const double p2 = (B[0][0] - q) * (B[0][0] - q) + (B[1][1] - q) * (B[1][1] - q) +
                      (B[2][2] - q) * (B[2][2] - q) + 2.0 * p1;
    const double p = std::sqrt(p2 / 6.0);
    double C[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) C[i][j] = (B[i][j] - (i == j ? q : 0.0)) / p;
    const double det = C[0][0] * (C[1][1] * C[2][2] - C[1][2] * C[2][1]) -
                       C[0][1] * (C[1][0] * C[2][2] - C[1][2] * C[2][0]) +
                       C[0][2] * (C[1][0] * C[2][1] - C[1][1] * C[2][0]);
    const double r = std::max(-1.0, std::min(1.0, det / 2.0));
    lam = q + 2.0 * p * std::cos(std::acos(r) / 3.0);
  }
  const double sigma = std::sqrt(std::max(lam, 0.0)) * (1.0 + 1e-9) + 1e-300;
  return std::min(sigma, std::sqrt(fro2) * (1.0 + 1e-12));
}

// Centre and radius (max |v - c|) of a triangle object's world vertices.
static void world_bounds(const GeometryDev& g, int kind, const double* dofs, double c[3],
                         double& R) {
  if (kind == PII_DOF_AFFINE) {
    // W = A V + t: centre A c_ref + t, radius <= |A|_2 R_ref
    for (int r = 0; r < 3; ++r)
      c[r] = dofs[3 + 3 * r] * g.ref_c[0] + dofs[4 + 3 * r] * g.ref_c[1] +
             dofs[5 + 3 * r] * g.ref_c[2] + dofs[r];
    R = spectral_norm3(dofs + 3) * g.ref_R;
    return;
  }
  double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
  if (g.dof_identity) {
    // indexed mesh: the DoF vector holds every vertex once, contiguously
    for (int v = 0; v < g.nv; ++v)
      for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(lo[k], dofs[3 * (size_t)v + k]);
        hi[k] = std::max(hi[k], dofs[3 * (size_t)v + k]);
      }
    if (!(lo[0] <= hi[0])) {
      c[0] = c[1] = c[2] = 0.0;
      R = 0.0;
      return;
    }
    for (int k = 0; k < 3; ++k) c[k] = 0.5 * (lo[k] + hi[k]);
    double r2 = 0.0;
    for (int v = 0; v < g.nv; ++v) {
      const double x = dofs[3 * (size_t)v] - c[0], y = dofs[3 * (size_t)v + 1] - c[1],
                   z = dofs[3 * (size_t)v + 2] - c[2];
      // explicit FMAs: the device bounds kernel (k_dof_bounds) rounds alike
      r2 = std::max(r2, std::fma(z, z, std::fma(y, y, x * x)));
    }
    R = std::sqrt(r2) * (1.0 + 1e-12);
    return;
  }
  for (int v = 0; v < g.nv; ++v) {
    const int dv = g.h_dofvid[v];
    if (dv < 0) continue;  // only in zero-area faces
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], dofs[3 * (size_t)dv + k]);
      hi[k] = std::max(hi[k], dofs[3 * (size_t)dv + k]);
    }
  }
  if (!(lo[0] <= hi[0])) {
    c[0] = c[1] = c[2] = 0.0;
    R = 0.0;
    return;
  }
  for (int k = 0; k < 3; ++k) c[k] = 0.5 * (lo[k] + hi[k]);
  double r2 = 0.0;
  for (int v = 0; v < g.nv; ++v) {
    const int dv = g.h_dofvid[v];
    if (dv < 0) continue;
    const double x = dofs[3 * (size_t)dv] - c[0], y = dofs[3 * (size_t)dv + 1] - c[1],
                 z = dofs[3 * (size_t)dv + 2] - c[2];
    r2 = std::max(r2, std::fma(z, z, std::fma(y, y, x * x)));
  }
  R = std::sqrt(r2) * (1.0 + 1e-12);
}

// A = s R with R a rotation (rows of A orthogonal, of equal length s, to
// 1e-12 relative; det A > 0): the transforms under which a geometry's
// cached harmonic coefficients give its moment series.
static bool similarity_scale(const double* A, double& s) {
  double B[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      B[i][j] = A[3 * i] * A[3 * j] + A[3 * i + 1] * A[3 * j + 1] + A[3 * i + 2] * A[3 * j + 2];
  const double s2 = (B[0][0] + B[1][1] + B[2][2]) / 3.0;
  if (!(s2 > 0.0) || !std::isfinite(s2)) return false;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      if (std::abs(B[i][j] - (i == j ? s2 : 0.0)) > 1e-12 * s2) return false;
  const double det = A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
                     A[2] * (A[3] * A[7] - A[4] * A[6]);
  if (!(det > 0.0)) return false;
  s = std::sqrt(s2);
  return true;
}

// Smallest series order of kSeriesOrders whose tail bound at X = |k| R,
// X^(P+1) / (P+1)! / (1 - X / (P+2)), is below 3.6e-10 (of the area).
// One pass over m = 1..P+1 serves every candidate order (this runs per object
// per call).
struct InverseIntegers {
  double v[kSeriesOrder + 2];
  constexpr InverseIntegers() : v() {
    for (int m = 1; m <= kSeriesOrder + 1; ++m) v[m] = 1.0 / m;
  }
};
static constexpr InverseIntegers kInverse{};

static int series_order(double X) {
  double t = 1.0;  // X^m / m!
  int m = 0;
  for (int i = 0; i < kNumSeriesOrders; ++i) {
    const int P = kSeriesOrders[i];
    for (; m < P + 1; ++m) t *= X * kInverse.v[m + 1];
    if (X < P + 2 && t / (1.0 - X / (P + 2)) < 3.6e-10) return P;
  }
  return kSeriesOrder;
}

void Batch::build(Workspace& w, hipStream_t s, Context* prof, int n,
                  const GeometryDev* const* geoms, const int* kinds, const double* const* dofs,
                  const int* ndofs, bool fast, const KGridDev* series_kg, const int* cols,
                  int n_cap, bool enqueue_now) {
  N = n;
  NC = std::max(n, n_cap);
  for (int i = 0; i < N; ++i) validate_object(*geoms[i], kinds[i], ndofs[i]);
  // host values of the DoFs for the plan (device-resident DoFs: host copies
  // of the affine ones; vertex DoFs come with plan_bounds)
  const double* const* hdofs = (device_dofs && plan_dofs) ? plan_dofs : dofs;
  // Moment series plan per object: leading shells with |k| R <= kSeriesReach,
  // when at least kSeriesMinShells of them (below that the per-element
  // kernels are cheaper than one pass over the elements per direction).
  // Disks and Gaussians also keep |k|^2 |A|_F^2 rho^2 / 4 <= 3
  // (sigma^2 / 2 <= 0.49), which bounds the growth the J1 (exponential)
  // factor adds to the series terms (the series keeps order kSeriesOrder).
  constexpr int kSeriesMinShells = 4, kSurfelMinShells = 8;
  std::vector<int> nser_i(N, 0), nord_i(N, 0);
  std::vector<double> kser_i(N, 0.0);
  std::vector<std::array<double, 3>> ctr_i(N, {0.0, 0.0, 0.0});
  Mh = 0;
  if (fast && series_kg && series_kg->nshell > 0) {
    Mh = series_kg->Mh;
    for (int i = 0; i < N; ++i) {
      const GeometryDev& g = *geoms[i];
      if (g.n == 0) continue;
      const bool tri = g.type == PII_GEOM_TRIANGLE;
      double R;
      if (plan_bounds && kinds[i] != PII_DOF_AFFINE) {
        for (int k = 0; k < 3; ++k) ctr_i[i][k] = plan_bounds[4 * (size_t)i + k];
        R = plan_bounds[4 * (size_t)i + 3];
      } else {
        world_bounds(g, kinds[i], hdofs[i], ctr_i[i].data(), R);
      }
      double cap2 = HUGE_VAL;  // bound on |k|^2 from the surfel factor
      if (!tri) {
        double an2 = 0.0;
        for (int t = 3; t < 12; ++t) an2 += hdofs[i][t] * hdofs[i][t];
        const double f = (g.type == PII_GEOM_DISK ? 0.25 : 0.5) * g.ref_smax * g.ref_smax * an2;
        if (f > 0.0) cap2 = (g.type == PII_GEOM_DISK ? 3.0 : 0.49) / f;
      }
      int ns = 0;
      double km = 0.0;
      while (ns < series_kg->nshell) {
        const double k = series_kg->h_shell_k[ns];
        if (!(k >= 0.0 && k * R <= kSeriesReach && k * k <= cap2)) break;
        km = std::max(km, k);
        ++ns;
      }
      if (ns < (tri ? kSeriesMinShells : kSurfelMinShells)) continue;
      nser_i[i] = ns;
      kser_i[i] = km;
      nord_i[i] = tri ? series_order(km * R) : kSeriesOrder;
    }
  }
  // Most per-element work first (elements x nodes left to the per-element
  // kernels): the dispatcher starts the longest work early, and the XCD
  // mapping deals the objects to the XCDs in this order.
  const int Qn = series_kg ? series_kg->Qh : 0;
  std::vector<double> cost(N);
  for (int i = 0; i < N; ++i)
    cost[i] = (double)geoms[i]->n * (double)(fast && Mh > 0 ? Qn - nser_i[i] * Mh : 1);
  order.resize(N);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
  // Descriptors, prefix sums and the concatenated DoF vectors share one
  // pinned staging area and one device buffer: a single DMA upload per call.
  std::vector<size_t> dof_off(N);
  size_t total_dofs = 0;
  for (int i = 0; i < N; ++i) {
    dof_off[i] = total_dofs;
    total_dofs += align_up((size_t)ndofs[i], 32);
  }
  // Laid out for NC >= N descriptors (fixed offsets for a capacity, so that a
  // captured upload and launches replay for any batch of up to NC objects).
  const size_t obj_bytes = align_up(sizeof(ObjDesc) * NC, 256);
  const size_t pre_bytes = align_up(sizeof(int) * (NC + 1), 256);
  const size_t dofs_at = obj_bytes + 2 * pre_bytes;
  const size_t dof_room = std::max(total_dofs, (size_t)32 * (n_cap > 0 ? n_cap : 0));
  stage_bytes = dofs_at + sizeof(double) * std::max<size_t>(dof_room, 1);
  w.stage.ensure(stage_bytes);
  w.misc.ensure(stage_bytes);
  char* hs = w.stage.as<char>();
  double* h_dofs = reinterpret_cast<double*>(hs + dofs_at);
  for (int i = 0; i < N; ++i) {
    const size_t end = (i + 1 < N) ? dof_off[i + 1] : total_dofs;
    if (device_dofs) {  // copied device to device after the upload
      std::fill(h_dofs + dof_off[i], h_dofs + end, 0.0);
      continue;
    }
    std::memcpy(h_dofs + dof_off[i], dofs[i], sizeof(double) * ndofs[i]);
    std::fill(h_dofs + dof_off[i] + ndofs[i], h_dofs + end, 0.0);
  }
  double* d_dofs = reinterpret_cast<double*>(static_cast<char*>(w.misc.p) + dofs_at);
  // World vertices and element records.
  std::vector<size_t> wv_off(N), rec_off(N);
  size_t total_w = 0, total_rec = 0;
  h_vprefix.assign(N + 1, 0);
  h_eprefix.assign(N + 1, 0);
  for (int slot = 0; slot < N; ++slot) {
    const GeometryDev& g = *geoms[order[slot]];
    const int nv = (g.type == PII_GEOM_TRIANGLE) ? g.nv : 0;
    h_vprefix[slot + 1] = h_vprefix[slot] + nv;
    h_eprefix[slot + 1] = h_eprefix[slot] + g.n;
    wv_off[slot] = total_w;
    total_w += align_up((size_t)nv * 3, 32);
    rec_off[slot] = total_rec;
    const int per = (g.type == PII_GEOM_TRIANGLE) ? (fast ? kRecFastTri : kRecPlainTri)
                                                  : kRecSurfel;
    total_rec += align_up((size_t)g.n * per, 32);
  }
  w.world.ensure(std::max<size_t>(total_w, 1) * sizeof(double));
  // (k_ff_fast's block loads are range-checked at each object's last record)
  w.faces.ensure(std::max<size_t>(total_rec, 1) * sizeof(double));
  std::vector<int> nser(N, 0), nord(N, 0);
  std::vector<double> kser(N, 0.0);
  std::vector<std::array<double, 3>> ctr(N, {0.0, 0.0, 0.0});
  std::vector<size_t> moff(N, 0);
  std::vector<double> shs(N, 0.0);
  size_t total_m = 0;
  nsh = 0;
  for (int slot = 0; slot < N; ++slot) {
    const int i = order[slot];
    nser[slot] = nser_i[i];
    nord[slot] = nord_i[i];
    kser[slot] = kser_i[i];
    ctr[slot] = ctr_i[i];
    if (use_sh && nser[slot] > 0 && geoms[i]->type == PII_GEOM_TRIANGLE && geoms[i]->shc.p &&
        kinds[i] == PII_DOF_AFFINE && similarity_scale(hdofs[i] + 3, shs[slot])) {
      ++nsh;
      continue;  // no moment records
    }
    shs[slot] = 0.0;
    if (nser[slot] > 0 && geoms[i]->type == PII_GEOM_TRIANGLE) {
      moff[slot] = total_m;
      total_m += align_up((size_t)geoms[i]->n * kRecMoment, 32);
    }
  }
  if (total_m) w.mrec.ensure(total_m * sizeof(double));
  nseries = 0;
  for (int slot = 0; slot < N; ++slot) nseries += nser[slot] > 0;
  h_objs.resize(N);
  for (int slot = 0; slot < N; ++slot) {
    const int i = order[slot];
    const GeometryDev& g = *geoms[i];
    ObjDesc& d = h_objs[slot];
    std::memset(&d, 0, sizeof(d));
    d.type = g.type;
    d.n = g.n;
    d.nv = g.nv;
    d.npatch = g.npatch;
    d.dof_kind = kinds[i];
    d.col = cols ? cols[i] : i;
    d.verts = g.verts.as<double>();
    d.faces = g.faces.as<int32_t>();
    d.dofvid = g.dofvid.as<int32_t>();
    d.patch_hdr = g.patch_hdr.as<int32_t>();
    d.patch_verts = g.patch_verts.as<int32_t>();
    d.patch_faces = g.patch_faces.as<uint32_t>();
    d.patch_floc = g.patch_floc.as<uint32_t>();
    d.el = g.el.as<double>();
    d.dofs = d_dofs + dof_off[i];
    d.world = w.world.as<double>() + wv_off[slot];
    d.rec = w.faces.as<double>() + rec_off[slot];
    d.nser = nser[slot];
    d.nord = nord[slot];
    d.qbeg = nser[slot] * Mh;
    d.kser = kser[slot];
    for (int k = 0; k < 3; ++k) d.ctr[k] = ctr[slot][k];
    // derivative series: affine triangles away from a singular A (its
    // reconstruction divides by |A^T d|^2 and |cof(A)^T d|^2)
    const bool aff_tri = nser[slot] > 0 && g.type == PII_GEOM_TRIANGLE &&
                         kinds[i] == PII_DOF_AFFINE;
    for (int k = 0; k < 3; ++k) d.rctr[k] = aff_tri ? g.ref_c[k] : d.ctr[k];
    if (aff_tri) {
      const double* A = hdofs[i] + 3;
      const double det = A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
                         A[2] * (A[3] * A[7] - A[4] * A[6]);
      double fro2 = 0.0;
      for (int k = 0; k < 9; ++k) fro2 += A[k] * A[k];
      d.dser = std::abs(det) >= 1e-6 * std::pow(fro2 / 3.0, 1.5) ? 1 : 0;
    }
    if (shs[slot] > 0.0) {
      d.shc = g.shc.as<double>();
      d.shs = shs[slot];
    } else if (nser[slot] && g.type == PII_GEOM_TRIANGLE) {
      d.mrec = w.mrec.as<double>() + moff[slot];
    }
  }
  qbeg_min = 0;
  if (N > 0) {
    qbeg_min = h_objs[0].qbeg;
    for (int slot = 1; slot < N; ++slot) qbeg_min = std::min(qbeg_min, h_objs[slot].qbeg);
  }
  std::memcpy(hs, h_objs.data(), sizeof(ObjDesc) * N);
  // empty descriptors past N: no elements, no nodes (every kernel skips them)
  for (int slot = N; slot < NC; ++slot) {
    ObjDesc& d = reinterpret_cast<ObjDesc*>(hs)[slot];
    std::memset(&d, 0, sizeof(d));
    d.type = PII_GEOM_TRIANGLE;
    d.dof_kind = PII_DOF_AFFINE;
    d.qbeg = 1 << 30;
  }
  int* hv = reinterpret_cast<int*>(hs + obj_bytes);
  int* he = reinterpret_cast<int*>(hs + obj_bytes + pre_bytes);
  std::memcpy(hv, h_vprefix.data(), sizeof(int) * (N + 1));
  std::memcpy(he, h_eprefix.data(), sizeof(int) * (N + 1));
  for (int slot = N + 1; slot <= NC; ++slot) {
    hv[slot] = h_vprefix[N];
    he[slot] = h_eprefix[N];
  }
  d_objs = reinterpret_cast<const ObjDesc*>(w.misc.p);
  d_vprefix = reinterpret_cast<const int*>(static_cast<char*>(w.misc.p) + obj_bytes);
  d_eprefix = reinterpret_cast<const int*>(static_cast<char*>(w.misc.p) + obj_bytes + pre_bytes);
  total_v = h_vprefix[N];
  total_e = h_eprefix[N];
  if (!enqueue_now) return;
  PII_HIP(hipMemcpyAsync(w.misc.p, hs, stage_bytes, hipMemcpyHostToDevice, s));
  if (device_dofs)
    for (int i = 0; i < N; ++i)
      PII_HIP(hipMemcpyAsync(d_dofs + dof_off[i], dofs[i], sizeof(double) * ndofs[i],
                             hipMemcpyDeviceToDevice, s));
  if (prof) {
    prof->begin_call();
    prof->phase_begin(0);
  }
  launch_prepass(s, d_objs, d_vprefix, total_v, d_eprefix, total_e, N, fast ? 1 : 0);
  if (prof) prof->phase_end(0);
  PII_HIP(hipGetLastError());
}

// Plain FP64 volume matrix: FP64 form factors over the full grid, FP64 Gram.
void volume_matrix_fp64(Context& c, const Batch& b, const KGridDev& kg, double* V_host) {
  hipStream_t s = c.stream;
  const int N = b.N, Q = kg.Q;
  const int ldJ = (int)align_up((size_t)Q, 64);
  c.w.J.ensure(sizeof(double2) * (size_t)ldJ * N);
  int max_elems = 0;
  for (const ObjDesc& d : b.h_objs) max_elems = std::max(max_elems, d.n);
  const int fsplit = ff_fp64_split(N, Q, max_elems);
  if (fsplit > 1) c.w.split.ensure(sizeof(double2) * ff_fp64_split_values(N, Q, fsplit));
  c.phase_begin(1);
  launch_ff_fp64(s, b.d_objs, N, kg.full.as<double>(), Q, c.w.J.as<double2>(), ldJ, fsplit,
                 fsplit > 1 ? c.w.split.as<double2>() : nullptr);
  c.phase_end(1);
  PII_HIP(hipGetLastError());
  const int nsplit = gram_fp64_splits(Q, N);
  c.w.part.ensure(sizeof(double) * (size_t)nsplit * N * N);
  c.w.V.ensure(sizeof(double) * (size_t)N * N);
  c.phase_begin(2);
  launch_gram_fp64(s, c.w.J.as<double2>(), Q, N, ldJ, kg.full.as<double>() + 7 * (size_t)Q,
                   c.w.part.as<double>(), nsplit, c.w.V.as<double>());
  c.phase_end(2);
  PII_HIP(hipGetLastError());
  // V_host may be device memory (pii_volume_matrix_device)
  PII_HIP(hipMemcpyAsync(V_host, c.w.V.p, sizeof(double) * (size_t)N * N, hipMemcpyDefault, s));
  PII_HIP(hipStreamSynchronize(s));
}

}  // namespace pii

namespace pii {

// A failed call leaves zero results in the caller's outputs, as the
// reference drivers return zero matrices on an error
// (cpp/src/compute_intersection_volume_multi_object.cu:250-261).
template <class F, class Z>
static int guarded_zero(F&& f, Z&& zero) {
  const int rc = guarded(std::forward<F>(f));
  if (rc != PII_OK) zero();
  return rc;
}
static void zero_fill(double* p, long long n) {
  if (p && n > 0) std::fill(p, p + n, 0.0);
}
// Sum of the DoF counts, or -1 when the arrays are unusable.
static long long dof_total(int n, const int* ndofs) {
  if (n <= 0 || !ndofs) return -1;
  long long d = 0;
  for (int i = 0; i < n; ++i) {
    if (ndofs[i] < 0) return -1;
    d += ndofs[i];
  }
  return d;
}

}  // namespace pii

using namespace pii;

extern "C" {

const char* pii_last_error(void) { return pii::last_error(); }
const char* pii_version(void) { return "pointint_hip 0.1 (gfx950)"; }

int pii_gauss_legendre(int n, double a, double b, double* x, double* w) {
  return guarded([&] {
    require(n > 0 && x && w, "pii_gauss_legendre: n > 0 and output arrays required");
    gauss_legendre(n, a, b, x, w);
  });
}

int pii_lebedev_load(const char* path, double* dirs, double* w, int cap, int* count) {
  return guarded([&] {
    require(path != nullptr, "pii_lebedev_load: path is NULL");
    std::vector<double> d, ww;
    if (lebedev_load(path, d, ww) != PII_OK)
      throw Error(PII_ERR_IO, std::string("failed to read Lebedev file: ") + path);
    const int M = (int)ww.size();
    if (count) *count = M;
    if (dirs && w) {
      const int m = std::min(M, cap);
      std::memcpy(dirs, d.data(), sizeof(double) * 3 * m);
      std::memcpy(w, ww.data(), sizeof(double) * m);
    }
  });
}

int pii_kgrid_build(const double* leb_dirs, const double* leb_w, int M, int nrad, double* dirs,
                    double* kmag, double* w) {
  return guarded([&] {
    require(M > 0 && nrad > 0, "pii_kgrid_build: M and nrad must be positive");
    require(leb_dirs && leb_w && dirs && kmag && w, "pii_kgrid_build: NULL array");
    kgrid_build(leb_dirs, leb_w, M, nrad, dirs, kmag, w);
  });
}

int pii_context_create(int device, pii_context** out) {
  return guarded([&] {
    require(out != nullptr, "pii_context_create: out is NULL");
    int ndev = 0;
    PII_HIP(hipGetDeviceCount(&ndev));
    require(device >= 0 && device < ndev, "pii_context_create: no HIP device " +
                                              std::to_string(device));
    PII_HIP(hipSetDevice(device));
    auto* ctx = new pii_context();
    ctx->c.device = device;
    if (hipStreamCreateWithFlags(&ctx->c.stream, hipStreamNonBlocking) != hipSuccess) {
      delete ctx;
      throw Error(PII_ERR_HIP, "hipStreamCreate failed");
    }
    if (hipEventCreateWithFlags(&ctx->c.caller_ev, hipEventDisableTiming) != hipSuccess) {
      (void)hipStreamDestroy(ctx->c.stream);
      delete ctx;
      throw Error(PII_ERR_HIP, "hipEventCreate failed");
    }
    *out = ctx;
  });
}

int pii_context_destroy(pii_context* ctx) {
  return guarded([&] {
    if (!ctx) return;
    (void)hipSetDevice(ctx->c.device);
    (void)hipStreamSynchronize(ctx->c.stream);
    for (auto& pe : ctx->c.ev)
      for (auto& e : pe)
        if (e) (void)hipEventDestroy(e);
    if (ctx->c.caller_ev) (void)hipEventDestroy(ctx->c.caller_ev);
    (void)hipStreamDestroy(ctx->c.stream);
    delete ctx;
  });
}

int pii_context_synchronize(pii_context* ctx) {
  return guarded([&] {
    require(ctx != nullptr, "NULL context");
    PII_HIP(hipStreamSynchronize(ctx->c.stream));
  });
}

int pii_context_set_profiling(pii_context* ctx, int enable) {
  return guarded([&] {
    require(ctx != nullptr, "NULL context");
    ctx->c.profiling = enable != 0;
    for (bool& u : ctx->c.ev_used) u = false;
  });
}

int pii_context_set_series(pii_context* ctx, int enable) {
  return guarded([&] {
    require(ctx != nullptr, "NULL context");
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    ctx->c.series = enable != 0;
  });
}

int pii_context_timings(pii_context* ctx, double* ms, int n) {
  return guarded([&] {
    require(ctx && ms && n > 0, "pii_context_timings: bad arguments");
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    PII_HIP(hipSetDevice(ctx->c.device));
    for (int p = 0; p < n; ++p) {
      ms[p] = 0.0;
      if (p < Context::kPhases && ctx->c.ev_used[p]) {
        float t = 0.f;
        PII_HIP(hipEventSynchronize(ctx->c.ev[p][1]));
        PII_HIP(hipEventElapsedTime(&t, ctx->c.ev[p][0], ctx->c.ev[p][1]));
        ms[p] = t;
      }
    }
  });
}

int pii_context_work(pii_context* ctx, double* units, int n) {
  return guarded([&] {
    require(ctx && units && n > 0, "pii_context_work: bad arguments");
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    for (int i = 0; i < n; ++i) units[i] = i < Context::kWork ? ctx->c.work[i] : 0.0;
  });
}

int pii_kgrid_create(pii_context* ctx, int Q, const double* dirs, const double* kmag,
                     const double* w, pii_kgrid** out) {
  return guarded([&] {
    require(ctx && out && dirs && kmag && w && Q > 0, "pii_kgrid_create: bad arguments");
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    PII_HIP(hipSetDevice(ctx->c.device));
    auto kg = std::make_unique<pii_kgrid>();
    kg->ctx = ctx;
    kg->device = ctx->c.device;
    KGridDev& k = kg->k;
    k.Q = Q;
    k.h_dirs.assign(dirs, dirs + 3 * (size_t)Q);
    k.h_kmag.assign(kmag, kmag + Q);
    k.h_w.assign(w, w + Q);
    k.symmetric = kgrid_pairs(Q, dirs, kmag, w, k.h_anti, k.h_rep);
    kgrid_order_reps(Q, dirs, kmag, k.h_anti, k.h_rep);
    k.Qh = (int)k.h_rep.size();
    kgrid_shells(k);
    auto soa = [&](const std::vector<int>& idx, bool sum_pair) {
      const size_t n = idx.size();
      std::vector<double> a(8 * n);
      for (size_t i = 0; i < n; ++i) {
        const int q = idx[i];
        for (int c = 0; c < 3; ++c) {
          a[c * n + i] = kmag[q] * dirs[3 * (size_t)q + c];
          a[(3 + c) * n + i] = dirs[3 * (size_t)q + c];
        }
        a[6 * n + i] = kmag[q];
        a[7 * n + i] = (sum_pair && k.h_anti[q] >= 0) ? w[q] + w[k.h_anti[q]] : w[q];
      }
      return a;
    };
    std::vector<int> all(Q);
    std::iota(all.begin(), all.end(), 0);
    const std::vector<double> full = soa(all, false);
    const std::vector<double> rep = soa(k.h_rep, k.symmetric);
    upload(k.full, full.data(), full.size(), ctx->c.stream);
    upload(k.rep, rep.data(), rep.size(), ctx->c.stream);
    upload(k.rep_index, k.h_rep.data(), k.h_rep.size(), ctx->c.stream);
    upload(k.anti_index, k.h_anti.data(), k.h_anti.size(), ctx->c.stream);
    PII_HIP(hipStreamSynchronize(ctx->c.stream));
    *out = kg.release();
  });
}

int pii_kgrid_destroy(pii_kgrid* kg) {
  return guarded([&] {
    if (!kg) return;
    if (kg->refs.load() > 0)
      throw Error(PII_ERR_BUSY, "pii_kgrid_destroy: the k-grid is in use by a server");
    (void)hipSetDevice(kg->device);
    delete kg;
  });
}

int pii_kgrid_info(const pii_kgrid* kg, int* Q, int* Qh, int* symmetric) {
  return guarded([&] {
    require(kg != nullptr, "NULL k-grid");
    if (Q) *Q = kg->k.Q;
    if (Qh) *Qh = kg->k.Qh;
    if (symmetric) *symmetric = kg->k.symmetric ? 1 : 0;
  });
}

static void finish_triangles(pii_context* ctx, GeometryDev& g) {
  hipStream_t s = ctx->c.stream;
  // bounding-box centre and radius of the reference vertices (moment series)
  if (g.nv > 0) {
    double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
    for (int v = 0; v < g.nv; ++v)
      for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(lo[k], g.h_verts[3 * (size_t)v + k]);
        hi[k] = std::max(hi[k], g.h_verts[3 * (size_t)v + k]);
      }
    for (int k = 0; k < 3; ++k) g.ref_c[k] = 0.5 * (lo[k] + hi[k]);
    double r2 = 0.0;
    for (int v = 0; v < g.nv; ++v) {
      double d2 = 0.0;
      for (int k = 0; k < 3; ++k) {
        const double t = g.h_verts[3 * (size_t)v + k] - g.ref_c[k];
        d2 += t * t;
      }
      r2 = std::max(r2, d2);
    }
    g.ref_R = std::sqrt(r2) * (1.0 + 1e-12);
  }
  upload(g.verts, g.h_verts.data(), g.h_verts.size(), s);
  upload(g.faces, g.h_faces.data(), g.h_faces.size(), s);
  upload(g.dofvid, g.h_dofvid.data(), g.h_dofvid.size(), s);
  std::vector<int> hdr, pverts;
  std::vector<uint32_t> pfaces, ploc;
  build_patches(g.nv, g.n, g.h_faces.data(), hdr, pverts, pfaces, ploc);
  g.npatch = (int)hdr.size() / 4;
  upload(g.patch_hdr, hdr.data(), hdr.size(), s);
  upload(g.patch_verts, pverts.data(), pverts.size(), s);
  upload(g.patch_faces, pfaces.data(), pfaces.size(), s);
  upload(g.patch_floc, ploc.data(), ploc.size(), s);
  PII_HIP(hipStreamSynchronize(s));
}

int pii_geometry_create_mesh(pii_context* ctx, const double* V, int nv, const int32_t* F, int nf,
                             pii_geometry** out) {
  return guarded([&] {
    require(ctx && out && nv >= 0 && nf >= 0, "pii_geometry_create_mesh: bad arguments");
    require((nv == 0 || V) && (nf == 0 || F), "pii_geometry_create_mesh: NULL array");
    for (int i = 0; i < 3 * nf; ++i)
      require(F[i] >= 0 && F[i] < nv, "face index out of range at " + std::to_string(i));
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    PII_HIP(hipSetDevice(ctx->c.device));
    auto g = std::make_unique<pii_geometry>();
    g->ctx = ctx;
    g->device = ctx->c.device;
    GeometryDev& d = g->g;
    d.type = PII_GEOM_TRIANGLE;
    d.n = nf;
    d.nv = nv;
    d.h_verts.assign(V, V + 3 * (size_t)nv);
    d.h_faces.assign(F, F + 3 * (size_t)nf);
    d.h_dofvid.resize(nv);
    std::iota(d.h_dofvid.begin(), d.h_dofvid.end(), 0);
    d.ndof_verts = nv;
    d.dof_identity = true;
    finish_triangles(ctx, d);
    *out = g.release();
  });
}

int pii_geometry_create_tris(pii_context* ctx, const double* tri, const int32_t* vid, int nt,
                             pii_geometry** out) {
  return guarded([&] {
    require(ctx && out && nt >= 0 && (nt == 0 || tri), "pii_geometry_create_tris: bad arguments");
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    PII_HIP(hipSetDevice(ctx->c.device));
    auto g = std::make_unique<pii_geometry>();
    g->ctx = ctx;
    g->device = ctx->c.device;
    GeometryDev& d = g->g;
    d.type = PII_GEOM_TRIANGLE;
    d.n = nt;
    d.h_faces.resize(3 * (size_t)nt);
    d.h_verts.clear();
    d.h_dofvid.clear();
    d.h_verts.reserve(9 * (size_t)nt);
    d.h_dofvid.reserve(3 * (size_t)nt);
    // Corners with the same vertex id become one vertex when their positions
    // agree to rounding (a + e1 and a + e2 reproduce the mesh vertex to about
    // one ulp): shared vertices let the form-factor kernels reuse each vertex
    // phasor across its faces.  Corners without an id, or whose position
    // disagrees, keep a vertex of their own.
    std::unordered_map<int, int> first;  // vertex id -> vertex index
    first.reserve(vid ? 2 * (size_t)nt : 0);
    int maxv = -1;
    if (vid)
      for (size_t i = 0; i < 3 * (size_t)nt; ++i) maxv = std::max(maxv, vid[i]);
    auto add_vertex = [&](const double* p, int dv) {
      d.h_verts.insert(d.h_verts.end(), p, p + 3);
      d.h_dofvid.push_back(dv);
      return (int)d.h_dofvid.size() - 1;
    };
    for (int i = 0; i < nt; ++i) {
      const double* r = tri + 12 * (size_t)i;
      // A face with any negative vid contributes nothing under vertex DoFs.
      const bool has_ids = vid && vid[3 * (size_t)i] >= 0 && vid[3 * (size_t)i + 1] >= 0 &&
                           vid[3 * (size_t)i + 2] >= 0;
      for (int c = 0; c < 3; ++c) {
        double p[3];
        for (int k = 0; k < 3; ++k) p[k] = r[k] + (c == 0 ? 0.0 : r[3 * c + k]);
        const int dv = has_ids ? vid[3 * (size_t)i + c] : -1;
        int v = -1;
        if (dv >= 0) {
          const auto it = first.find(dv);
          const int f = it == first.end() ? -1 : it->second;
          if (f >= 0) {
            const double* q = &d.h_verts[3 * (size_t)f];
            const double tol = 1e-14 * (std::fabs(q[0]) + std::fabs(q[1]) + std::fabs(q[2]) +
                                        std::fabs(r[0]) + std::fabs(r[1]) + std::fabs(r[2]));
            if (std::fabs(q[0] - p[0]) <= tol && std::fabs(q[1] - p[1]) <= tol &&
                std::fabs(q[2] - p[2]) <= tol)
              v = f;
          }
          if (v < 0) {
            v = add_vertex(p, dv);
            if (f < 0) first[dv] = v;
          }
        } else {
          v = add_vertex(p, -1);
        }
        d.h_faces[3 * (size_t)i + c] = v;
      }
    }
    d.nv = (int)d.h_dofvid.size();
    d.ndof_verts = maxv + 1;
    finish_triangles(ctx, d);
    *out = g.release();
  });
}

static int create_surfels(pii_context* ctx, int type, const double* el, int n, pii_geometry** out) {
  return guarded([&] {
    require(ctx && out && n >= 0 && (n == 0 || el), "pii_geometry_create: bad arguments");
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    PII_HIP(hipSetDevice(ctx->c.device));
    auto g = std::make_unique<pii_geometry>();
    g->ctx = ctx;
    g->device = ctx->c.device;
    GeometryDev& d = g->g;
    d.type = type;
    d.n = n;
    d.h_el.assign(el, el + 8 * (size_t)n);
    if (n > 0) {
      double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
      for (int e = 0; e < n; ++e) {
        for (int k = 0; k < 3; ++k) {
          lo[k] = std::min(lo[k], el[8 * (size_t)e + k]);
          hi[k] = std::max(hi[k], el[8 * (size_t)e + k]);
        }
        d.ref_smax = std::max(d.ref_smax, std::abs(el[8 * (size_t)e + 6]));
      }
      for (int k = 0; k < 3; ++k) d.ref_c[k] = 0.5 * (lo[k] + hi[k]);
      double r2 = 0.0;
      for (int e = 0; e < n; ++e) {
        double d2 = 0.0;
        for (int k = 0; k < 3; ++k) {
          const double t = el[8 * (size_t)e + k] - d.ref_c[k];
          d2 += t * t;
        }
        r2 = std::max(r2, d2);
      }
      d.ref_R = std::sqrt(r2) * (1.0 + 1e-12);
    }
    upload(d.el, d.h_el.data(), d.h_el.size(), ctx->c.stream);
    PII_HIP(hipStreamSynchronize(ctx->c.stream));
    *out = g.release();
  });
}

int pii_geometry_create_disks(pii_context* ctx, const double* el, int n, pii_geometry** out) {
  return create_surfels(ctx, PII_GEOM_DISK, el, n, out);
}

int pii_geometry_create_gaussians(pii_context* ctx, const double* el, int n, pii_geometry** out) {
  return create_surfels(ctx, PII_GEOM_GAUSSIAN, el, n, out);
}

int pii_geometry_destroy(pii_geometry* g) {
  return guarded([&] {
    if (!g) return;
    if (g->refs.load() > 0)
      throw Error(PII_ERR_BUSY, "pii_geometry_destroy: the geometry is in use by " +
                                    std::to_string(g->refs.load()) + " serving queries");
    (void)hipSetDevice(g->device);
    delete g;
  });
}

int pii_geometry_cache_moments(pii_context* ctx, int n, pii_geometry* const* geoms) {
  return guarded([&] {
    require(ctx != nullptr, "pii_geometry_cache_moments: NULL context");
    require(n >= 0, "pii_geometry_cache_moments: negative geometry count");
    if (n == 0) return;
    require(geoms != nullptr, "pii_geometry_cache_moments: NULL array");
    std::vector<GeometryDev*> gs(n);
    for (int i = 0; i < n; ++i) {
      require(geoms[i] != nullptr, "pii_geometry_cache_moments: NULL geometry");
      if (geoms[i]->device != ctx->c.device)
        throw Error(PII_ERR_INVALID,
                    "pii_geometry_cache_moments: geometry on another device than the context");
      if (geoms[i]->refs.load() > 0)
        throw Error(PII_ERR_BUSY, "pii_geometry_cache_moments: the geometry is in use by " +
                                      std::to_string(geoms[i]->refs.load()) + " serving queries");
      gs[i] = &geoms[i]->g;
    }
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    PII_HIP(hipSetDevice(ctx->c.device));
    cache_moments(ctx->c, n, gs.data());
  });
}

int pii_geometry_info(const pii_geometry* g, int* type, int* num_elements, int* num_dof_vertices) {
  return guarded([&] {
    require(g != nullptr, "NULL geometry");
    if (type) *type = g->g.type;
    if (num_elements) *num_elements = g->g.n;
    if (num_dof_vertices) *num_dof_vertices = g->g.ndof_verts;
  });
}

// Caller memory for the device-output entries: device memory of the
// context's device (the reference's plug-ins write into a device double2*).
}  // extern "C"

namespace pii {

void require_device_pointer(const void* p, int device, const char* who) {
  hipPointerAttribute_t a{};
  const hipError_t e = hipPointerGetAttributes(&a, p);
  if (e != hipSuccess) (void)hipGetLastError();
  if (e != hipSuccess || a.type != hipMemoryTypeDevice || a.device != device)
    throw Error(PII_ERR_INVALID,
                std::string(who) + ": not device memory of the context's device");
}

// Plan inputs of a batch whose DoFs are in device memory: host copies of the
// affine DoF vectors (12 doubles each) and, for vertex-DoF objects on the
// fast path, the series centre and radius from a device bounds kernel, as
// world_bounds computes them from host values.
void device_dof_plan(Context& c, int n, const GeometryDev* const* gs, const int* kinds,
                     const double* const* d_dofs, const int* ndofs, bool fast,
                     std::vector<std::vector<double>>& host_copies,
                     std::vector<const double*>& plan_dofs, std::vector<double>& bounds) {
  host_copies.assign(n, {});
  plan_dofs.assign(n, nullptr);
  bounds.assign(4 * (size_t)n, 0.0);
  std::vector<DofBoundsJob> jobs;
  std::vector<int> job_obj;
  for (int i = 0; i < n; ++i) {
    if (kinds[i] == PII_DOF_AFFINE) {
      host_copies[i].resize(12);
      PII_HIP(hipMemcpyAsync(host_copies[i].data(), d_dofs[i], sizeof(double) * 12,
                             hipMemcpyDeviceToHost, c.stream));
      plan_dofs[i] = host_copies[i].data();
    } else if (fast && gs[i]->type == PII_GEOM_TRIANGLE) {
      jobs.push_back({gs[i]->dofvid.as<int32_t>(), d_dofs[i], gs[i]->nv, 0});
      job_obj.push_back(i);
    }
  }
  std::vector<double> out(5 * jobs.size());
  if (!jobs.empty()) {
    c.w.C.ensure(sizeof(DofBoundsJob) * jobs.size() + sizeof(double) * out.size() + 256);
    char* base = static_cast<char*>(c.w.C.p);
    double* d_out = reinterpret_cast<double*>(base + align_up(sizeof(DofBoundsJob) * jobs.size(), 256));
    PII_HIP(hipMemcpyAsync(base, jobs.data(), sizeof(DofBoundsJob) * jobs.size(),
                           hipMemcpyHostToDevice, c.stream));
    launch_dof_bounds(c.stream, reinterpret_cast<const DofBoundsJob*>(base), (int)jobs.size(),
                      d_out);
    PII_HIP(hipGetLastError());
    PII_HIP(hipMemcpyAsync(out.data(), d_out, sizeof(double) * out.size(), hipMemcpyDeviceToHost,
                           c.stream));
  }
  PII_HIP(hipStreamSynchronize(c.stream));
  for (size_t j = 0; j < jobs.size(); ++j) {
    const int i = job_obj[j];
    const double* o = &out[5 * j];
    for (int k = 0; k < 3; ++k) bounds[4 * (size_t)i + k] = o[k];
    bounds[4 * (size_t)i + 3] = o[4] > 0.0 ? std::sqrt(o[3]) * (1.0 + 1e-12) : 0.0;
  }
  (void)ndofs;
}

}  // namespace pii

extern "C" {

}  // extern "C"

namespace pii {

void form_factor_to(Context& c, const pii_geometry* g, int dof_kind, const double* dofs,
                    int ndofs, const pii_kgrid* kg, int precision, double* out,
                    hipMemcpyKind kind) {
  const GeometryDev* gp = &g->g;
  Batch b;
  if (precision == PII_FF_FP64) {
    b.build(c, 1, &gp, &dof_kind, &dofs, &ndofs, false);
    const int Q = kg->k.Q;
    const int ldJ = (int)align_up((size_t)Q, 64);
    c.w.J.ensure(sizeof(double2) * (size_t)ldJ);
    const int fsplit = ff_fp64_split(1, Q, b.h_objs[0].n);
    if (fsplit > 1) c.w.split.ensure(sizeof(double2) * ff_fp64_split_values(1, Q, fsplit));
    launch_ff_fp64(c.stream, b.d_objs, 1, kg->k.full.as<double>(), Q, c.w.J.as<double2>(), ldJ,
                   fsplit, fsplit > 1 ? c.w.split.as<double2>() : nullptr);
    PII_HIP(hipGetLastError());
    PII_HIP(hipMemcpyAsync(out, c.w.J.p, sizeof(double2) * Q, kind, c.stream));
    PII_HIP(hipStreamSynchronize(c.stream));
  } else if (precision == PII_FF_FAST) {
    b.use_sh = true;
    b.build(c, 1, &gp, &dof_kind, &dofs, &ndofs, true, c.series_grid(kg->k));
    form_factor_fast_full(c, b, kg->k, out, kind);
  } else {
    throw Error(PII_ERR_INVALID, "unknown form-factor precision");
  }
}

}  // namespace pii

extern "C" {

int pii_debug_fail_allocation(int after) {
  return guarded([&] {
    require(after >= -1, "pii_debug_fail_allocation: after must be >= -1");
    g_fail_after.store(after);
  });
}

int pii_form_factor(pii_context* ctx, const pii_geometry* g, int dof_kind, const double* dofs,
                    int ndofs, const pii_kgrid* kg, int precision, double* out) {
  return guarded_zero([&] {
    require(ctx && g && kg && dofs && out, "pii_form_factor: NULL argument");
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    PII_HIP(hipSetDevice(ctx->c.device));
    form_factor_to(ctx->c, g, dof_kind, dofs, ndofs, kg, precision, out, hipMemcpyDeviceToHost);
  }, [&] { if (kg) zero_fill(out, 2LL * kg->k.Q); });
}

// The context stream waits for the work the caller has queued on its stream
// (the null stream unless pii_context_set_caller_stream named another), so
// that a fill or read of the caller's buffer queued before the call is
// ordered before the library's writes into it.
}  // extern "C"

namespace pii {

void order_after_caller(pii_context* ctx) {
  PII_HIP(hipEventRecord(ctx->c.caller_ev, ctx->c.caller_stream));
  PII_HIP(hipStreamWaitEvent(ctx->c.stream, ctx->c.caller_ev, 0));
}

}  // namespace pii

extern "C" {

int pii_context_set_caller_stream(pii_context* ctx, void* stream) {
  return guarded([&] {
    require(ctx != nullptr, "NULL context");
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    ctx->c.caller_stream = static_cast<hipStream_t>(stream);
  });
}

int pii_form_factor_device(pii_context* ctx, const pii_geometry* g, int dof_kind,
                           const double* dofs, int ndofs, const pii_kgrid* kg, int precision,
                           double* d_out) {
  return guarded([&] {
    require(ctx && g && kg && dofs && d_out, "pii_form_factor_device: NULL argument");
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    PII_HIP(hipSetDevice(ctx->c.device));
    require_device_pointer(d_out, ctx->c.device, "pii_form_factor_device");
    order_after_caller(ctx);
    form_factor_to(ctx->c, g, dof_kind, dofs, ndofs, kg, precision, d_out,
                   hipMemcpyDeviceToDevice);
  });
}

static void jacobian_to(pii_context* ctx, const pii_geometry* g, int dof_kind, const double* dofs,
                        int ndofs, const pii_kgrid* kg, double* out, hipMemcpyKind kind) {
  const GeometryDev* gp = &g->g;
  Batch b;
  b.build(ctx->c, 1, &gp, &dof_kind, &dofs, &ndofs, false);
  const int Q = kg->k.Q;
  const size_t bytes = sizeof(double2) * (size_t)Q * ndofs;
  ctx->c.w.J.ensure(bytes);
  launch_jacobian(ctx->c.stream, b.d_objs, gp->type, dof_kind, gp->n, kg->k.full.as<double>(), Q,
                  ndofs, ctx->c.w.J.as<double2>());
  PII_HIP(hipGetLastError());
  PII_HIP(hipMemcpyAsync(out, ctx->c.w.J.p, bytes, kind, ctx->c.stream));
  PII_HIP(hipStreamSynchronize(ctx->c.stream));
}

int pii_form_factor_jacobian(pii_context* ctx, const pii_geometry* g, int dof_kind,
                             const double* dofs, int ndofs, const pii_kgrid* kg, double* out) {
  return guarded([&] {
    require(ctx && g && kg && dofs && out, "pii_form_factor_jacobian: NULL argument");
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    PII_HIP(hipSetDevice(ctx->c.device));
    jacobian_to(ctx, g, dof_kind, dofs, ndofs, kg, out, hipMemcpyDeviceToHost);
  });
}

int pii_form_factor_jacobian_device(pii_context* ctx, const pii_geometry* g, int dof_kind,
                                    const double* dofs, int ndofs, const pii_kgrid* kg,
                                    double* d_out) {
  return guarded([&] {
    require(ctx && g && kg && dofs && d_out, "pii_form_factor_jacobian_device: NULL argument");
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    PII_HIP(hipSetDevice(ctx->c.device));
    require_device_pointer(d_out, ctx->c.device, "pii_form_factor_jacobian_device");
    order_after_caller(ctx);
    jacobian_to(ctx, g, dof_kind, dofs, ndofs, kg, d_out, hipMemcpyDeviceToDevice);
  });
}

int pii_volume_matrix(pii_context* ctx, int n, const pii_geometry* const* geoms,
                      const int* dof_kinds, const double* const* dofs, const int* ndofs,
                      const pii_kgrid* kg, int ff_precision, int gram_mode, double* V) {
  return guarded_zero([&] {
    require(ctx && kg && V, "pii_volume_matrix: NULL argument");
    require(n >= 0, "pii_volume_matrix: negative object count");
    if (n == 0) return;
    require(geoms && dof_kinds && dofs && ndofs, "pii_volume_matrix: NULL array");
    for (int i = 0; i < n; ++i) require(geoms[i] && dofs[i], "pii_volume_matrix: NULL object");
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    PII_HIP(hipSetDevice(ctx->c.device));
    std::vector<const GeometryDev*> gs(n);
    for (int i = 0; i < n; ++i) gs[i] = &geoms[i]->g;
    Batch b;
    if (ff_precision == PII_FF_FP64 && gram_mode == PII_GRAM_FP64) {
      b.build(ctx->c, n, gs.data(), dof_kinds, dofs, ndofs, false);
      volume_matrix_fp64(ctx->c, b, kg->k, V);
    } else if (ff_precision == PII_FF_FAST) {
      b.use_sh = true;
      b.build(ctx->c, n, gs.data(), dof_kinds, dofs, ndofs, true, ctx->c.series_grid(kg->k));
      volume_matrix_fast(ctx->c, b, kg->k, gram_mode, V);
    } else {
      throw Error(PII_ERR_INVALID, "unsupported precision / Gram mode combination");
    }
  }, [&] { if (n > 0) zero_fill(V, (long long)n * n); });
}

int pii_volume_matrix_vjp(pii_context* ctx, int n, const pii_geometry* const* geoms,
                          const int* dof_kinds, const double* const* dofs, const int* ndofs,
                          const pii_kgrid* kg, const double* dL_dV, double* const* grads,
                          double* V) {
  return guarded_zero([&] {
    require(ctx && kg && dL_dV && grads, "pii_volume_matrix_vjp: NULL argument");
    require(n > 0 && geoms && dof_kinds && dofs && ndofs, "pii_volume_matrix_vjp: bad arrays");
    for (int i = 0; i < n; ++i)
      require(geoms[i] && dofs[i] && grads[i], "pii_volume_matrix_vjp: NULL object");
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    PII_HIP(hipSetDevice(ctx->c.device));
    std::vector<const GeometryDev*> gs(n);
    for (int i = 0; i < n; ++i) gs[i] = &geoms[i]->g;
    Batch b;
    b.build(ctx->c, n, gs.data(), dof_kinds, dofs, ndofs, true, ctx->c.series_grid(kg->k));
    volume_matrix_vjp(ctx->c, b, gs.data(), kg->k, dL_dV, grads, V);
  }, [&] {
    if (n > 0) zero_fill(V, (long long)n * n);
    if (grads && dof_total(n, ndofs) >= 0)
      for (int i = 0; i < n; ++i) zero_fill(grads[i], ndofs[i]);
  });
}

int pii_volume_matrix_device(pii_context* ctx, int n, const pii_geometry* const* geoms,
                             const int* dof_kinds, const double* const* dofs, const int* ndofs,
                             const pii_kgrid* kg, int ff_precision, int gram_mode, double* V) {
  return guarded([&] {
    require(ctx && kg && V, "pii_volume_matrix_device: NULL argument");
    require(n > 0 && geoms && dof_kinds && dofs && ndofs, "pii_volume_matrix_device: bad arrays");
    for (int i = 0; i < n; ++i)
      require(geoms[i] && dofs[i], "pii_volume_matrix_device: NULL object");
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    PII_HIP(hipSetDevice(ctx->c.device));
    require_device_pointer(V, ctx->c.device, "pii_volume_matrix_device: V");
    for (int i = 0; i < n; ++i)
      require_device_pointer(dofs[i], ctx->c.device, "pii_volume_matrix_device: dofs");
    std::vector<const GeometryDev*> gs(n);
    for (int i = 0; i < n; ++i) gs[i] = &geoms[i]->g;
    for (int i = 0; i < n; ++i) validate_object(*gs[i], dof_kinds[i], ndofs[i]);
    order_after_caller(ctx);
    const bool fast = ff_precision == PII_FF_FAST;
    std::vector<std::vector<double>> hc;
    std::vector<const double*> pd;
    std::vector<double> bounds;
    device_dof_plan(ctx->c, n, gs.data(), dof_kinds, dofs, ndofs, fast, hc, pd, bounds);
    Batch b;
    b.device_dofs = true;
    b.plan_dofs = pd.data();
    b.plan_bounds = bounds.data();
    if (ff_precision == PII_FF_FP64 && gram_mode == PII_GRAM_FP64) {
      b.build(ctx->c, n, gs.data(), dof_kinds, dofs, ndofs, false);
      volume_matrix_fp64(ctx->c, b, kg->k, V);
    } else if (fast) {
      b.use_sh = true;
      b.build(ctx->c, n, gs.data(), dof_kinds, dofs, ndofs, true, ctx->c.series_grid(kg->k));
      volume_matrix_fast(ctx->c, b, kg->k, gram_mode, V);
    } else {
      throw Error(PII_ERR_INVALID, "unsupported precision / Gram mode combination");
    }
  });
}

int pii_volume_matrix_vjp_device(pii_context* ctx, int n, const pii_geometry* const* geoms,
                                 const int* dof_kinds, const double* const* dofs,
                                 const int* ndofs, const pii_kgrid* kg, const double* dL_dV,
                                 double* const* grads, double* V) {
  return guarded([&] {
    require(ctx && kg && dL_dV && grads, "pii_volume_matrix_vjp_device: NULL argument");
    require(n > 0 && geoms && dof_kinds && dofs && ndofs,
            "pii_volume_matrix_vjp_device: bad arrays");
    for (int i = 0; i < n; ++i)
      require(geoms[i] && dofs[i] && grads[i], "pii_volume_matrix_vjp_device: NULL object");
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    PII_HIP(hipSetDevice(ctx->c.device));
    require_device_pointer(dL_dV, ctx->c.device, "pii_volume_matrix_vjp_device: dL_dV");
    if (V) require_device_pointer(V, ctx->c.device, "pii_volume_matrix_vjp_device: V");
    for (int i = 0; i < n; ++i) {
      require_device_pointer(dofs[i], ctx->c.device, "pii_volume_matrix_vjp_device: dofs");
      require_device_pointer(grads[i], ctx->c.device, "pii_volume_matrix_vjp_device: grads");
    }
    std::vector<const GeometryDev*> gs(n);
    for (int i = 0; i < n; ++i) gs[i] = &geoms[i]->g;
    for (int i = 0; i < n; ++i) validate_object(*gs[i], dof_kinds[i], ndofs[i]);
    order_after_caller(ctx);
    std::vector<std::vector<double>> hc;
    std::vector<const double*> pd;
    std::vector<double> bounds;
    device_dof_plan(ctx->c, n, gs.data(), dof_kinds, dofs, ndofs, true, hc, pd, bounds);
    // dL/dV (n x n) is small: it joins the host-built mixing matrix
    std::vector<double> G((size_t)n * n);
    PII_HIP(hipMemcpyAsync(G.data(), dL_dV, sizeof(double) * G.size(), hipMemcpyDeviceToHost,
                           ctx->c.stream));
    PII_HIP(hipStreamSynchronize(ctx->c.stream));
    Batch b;
    b.device_dofs = true;
    b.plan_dofs = pd.data();
    b.plan_bounds = bounds.data();
    b.build(ctx->c, n, gs.data(), dof_kinds, dofs, ndofs, true, ctx->c.series_grid(kg->k));
    volume_matrix_vjp(ctx->c, b, gs.data(), kg->k, G.data(), grads, V, true);
  });
}

int pii_volume_matrix_hessian(pii_context* ctx, int n, const pii_geometry* const* geoms,
                              const int* dof_kinds, const double* const* dofs, const int* ndofs,
                              const pii_kgrid* kg, int gram_mode, double* H, double* V) {
  return guarded_zero([&] {
    require(ctx && kg && H, "pii_volume_matrix_hessian: NULL argument");
    require(n > 0 && geoms && dof_kinds && dofs && ndofs,
            "pii_volume_matrix_hessian: bad arrays");
    for (int i = 0; i < n; ++i) require(geoms[i] && dofs[i], "pii_volume_matrix_hessian: NULL object");
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    PII_HIP(hipSetDevice(ctx->c.device));
    std::vector<const GeometryDev*> gs(n);
    for (int i = 0; i < n; ++i) gs[i] = &geoms[i]->g;
    Batch b;
    b.build(ctx->c, n, gs.data(), dof_kinds, dofs, ndofs, true, ctx->c.series_grid(kg->k));
    volume_matrix_hessian(ctx->c, b, kg->k, gram_mode, ndofs, H, V);
  }, [&] {
    if (n > 0) zero_fill(V, (long long)n * n);
    const long long D = dof_total(n, ndofs);
    if (D > 0 && D <= kMaxDerivativeColumns) zero_fill(H, D * D);
  });
}

int pii_volume_matrix_derivatives(pii_context* ctx, int n, const pii_geometry* const* geoms,
                                  const int* dof_kinds, const double* const* dofs,
                                  const int* ndofs, const pii_kgrid* kg, int gram_mode, double* V,
                                  double* grad, double* H) {
  return guarded_zero([&] {
    require(ctx && kg, "pii_volume_matrix_derivatives: NULL argument");
    require(V || grad || H, "pii_volume_matrix_derivatives: no output requested");
    require(n > 0 && geoms && dof_kinds && dofs && ndofs,
            "pii_volume_matrix_derivatives: bad arrays");
    for (int i = 0; i < n; ++i)
      require(geoms[i] && dofs[i], "pii_volume_matrix_derivatives: NULL object");
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    PII_HIP(hipSetDevice(ctx->c.device));
    std::vector<const GeometryDev*> gs(n);
    for (int i = 0; i < n; ++i) gs[i] = &geoms[i]->g;
    Batch b;
    b.build(ctx->c, n, gs.data(), dof_kinds, dofs, ndofs, true, ctx->c.series_grid(kg->k));
    volume_matrix_derivatives(ctx->c, b, kg->k, gram_mode, ndofs, V, grad, H);
  }, [&] {
    if (n > 0) zero_fill(V, (long long)n * n);
    const long long D = dof_total(n, ndofs);
    if (D >= 0) zero_fill(grad, (long long)n * D);
    if (D > 0 && D <= kMaxDerivativeColumns) zero_fill(H, D * D);
  });
}

int pii_volume_matrix_affine_derivatives(pii_context* ctx, int n, const pii_geometry* const* geoms,
                                         const double* const* dofs, const pii_kgrid* kg,
                                         int gram_mode, double* V, double* grad, double* H) {
  return guarded_zero([&] {
    require(ctx && kg, "pii_volume_matrix_affine_derivatives: NULL argument");
    require(V || grad || H, "pii_volume_matrix_affine_derivatives: no output requested");
    require(n > 0 && geoms && dofs, "pii_volume_matrix_affine_derivatives: bad arrays");
    for (int i = 0; i < n; ++i)
      require(geoms[i] && dofs[i], "pii_volume_matrix_affine_derivatives: NULL object");
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    PII_HIP(hipSetDevice(ctx->c.device));
    std::vector<const GeometryDev*> gs(n);
    for (int i = 0; i < n; ++i) gs[i] = &geoms[i]->g;
    const std::vector<int> kinds(n, PII_DOF_AFFINE), nd(n, 12);
    Batch b;
    b.build(ctx->c, n, gs.data(), kinds.data(), dofs, nd.data(), true, ctx->c.series_grid(kg->k));
    volume_matrix_affine_all(ctx->c, b, kg->k, gram_mode, V, grad, H);
  }, [&] {
    if (n > 0) {
      zero_fill(V, (long long)n * n);
      zero_fill(grad, 12LL * n * n);
      if (12LL * n <= kMaxDerivativeColumns) zero_fill(H, 144LL * n * n);
    }
  });
}

int pii_measure_issue_rate(pii_context* ctx, int kind, double* per_second) {
  return guarded([&] {
    require(ctx && per_second, "pii_measure_issue_rate: NULL argument");
    require(kind >= 0 && kind <= 2, "pii_measure_issue_rate: kind must be 0, 1 or 2");
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    PII_HIP(hipSetDevice(ctx->c.device));
    int cus = 0;
    PII_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->c.device));
    const int blocks = std::max(1, cus) * 8;
    DevBuf sink;
    sink.ensure(sizeof(double) * (size_t)blocks);
    *per_second = probe_issue_rate(ctx->c.stream, kind, blocks, kind == 2 ? 16384 : 65536,
                                   sink.p);
  });
}

int pii_gram(pii_context* ctx, const double* J, int n, const pii_kgrid* kg, int gram_mode,
             double* V) {
  return guarded([&] {
    require(ctx && J && kg && V && n > 0, "pii_gram: bad arguments");
    require(gram_mode == PII_GRAM_FP64, "pii_gram: only the FP64 mode takes caller J");
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    PII_HIP(hipSetDevice(ctx->c.device));
    const int Q = kg->k.Q;
    const int ldJ = Q;
    upload(ctx->c.w.J, reinterpret_cast<const double2*>(J), (size_t)Q * n, ctx->c.stream);
    const int nsplit = gram_fp64_splits(Q, n);
    ctx->c.w.part.ensure(sizeof(double) * (size_t)nsplit * n * n);
    ctx->c.w.V.ensure(sizeof(double) * (size_t)n * n);
    launch_gram_fp64(ctx->c.stream, ctx->c.w.J.as<double2>(), Q, n, ldJ,
                     kg->k.full.as<double>() + 7 * (size_t)Q, ctx->c.w.part.as<double>(), nsplit,
                     ctx->c.w.V.as<double>());
    PII_HIP(hipGetLastError());
    PII_HIP(hipMemcpyAsync(V, ctx->c.w.V.p, sizeof(double) * (size_t)n * n,
                           hipMemcpyDeviceToHost, ctx->c.stream));
    PII_HIP(hipStreamSynchronize(ctx->c.stream));
  });
}

}  // extern "C"
