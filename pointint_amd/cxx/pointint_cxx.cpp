// C++ API with the reference's names (include/pointint/pointint.hpp),
// implemented on the C ABI of libpointint_hip.so.  Host code only: every
// device computation goes through pii_* entry points; the HIP runtime is
// called directly only for the device buffers handed to caller-registered
// form-factor plug-ins (CudaKernelRegistry).
//
// Error behaviour follows the reference drivers
// (cpp/src/compute_intersection_volume_multi_object.cu:250-314): a message
// on std::cerr and a zero (or empty) result, no exceptions.
#include "pointint/pointint.hpp"

#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <mutex>
#include <stdexcept>
#include <unordered_map>

#include "cxx_internal.hpp"
#include "pointint_hip.h"

namespace PoIntInt {

namespace {

constexpr double kPi = 3.14159265358979323846;

int g_ff = PII_FF_FAST;
int g_gram = PII_GRAM_FP16X3;

}  // namespace

namespace detail {

int ff_precision() { return g_ff; }
int gram_mode() { return g_gram; }

// One context per process on device 0, created on first use.
pii_context* context() {
  static std::once_flag once;
  static pii_context* ctx = nullptr;
  std::call_once(once, [] {
    if (pii_context_create(0, &ctx) != PII_OK) {
      std::cerr << "PoIntInt: HIP context creation failed: " << pii_last_error() << std::endl;
      ctx = nullptr;
    }
  });
  return ctx;
}

bool report(int rc, const char* where) {
  if (rc == PII_OK) return true;
  std::cerr << "PoIntInt::" << where << ": " << pii_last_error() << " (status " << rc << ")"
            << std::endl;
  return false;
}

// Device-resident copies made for one call (the reference allocates per call too).
struct CallObjects {
  std::vector<pii_geometry*> geoms;
  pii_kgrid* kg = nullptr;
  ~CallObjects() {
    for (pii_geometry* g : geoms) pii_geometry_destroy(g);
    if (kg) pii_kgrid_destroy(kg);
  }
};

pii_geometry* upload(pii_context* ctx, const Geometry& g) {
  pii_geometry* out = nullptr;
  int rc = PII_OK;
  if (g.type == GEOM_TRIANGLE) {
    std::vector<double> tri(12 * g.tris.size());
    std::vector<int32_t> vid(3 * g.tris.size());
    for (size_t i = 0; i < g.tris.size(); ++i) {
      const TriPacked& t = g.tris[i];
      const double v[12] = {t.a.x, t.a.y, t.a.z, t.e1.x, t.e1.y, t.e1.z,
                            t.e2.x, t.e2.y, t.e2.z, t.S.x, t.S.y, t.S.z};
      std::memcpy(&tri[12 * i], v, sizeof(v));
      vid[3 * i] = t.vid0;
      vid[3 * i + 1] = t.vid1;
      vid[3 * i + 2] = t.vid2;
    }
    rc = pii_geometry_create_tris(ctx, tri.data(), vid.data(), (int)g.tris.size(), &out);
  } else if (g.type == GEOM_DISK) {
    std::vector<double> el(8 * g.disks.size());
    for (size_t i = 0; i < g.disks.size(); ++i) {
      const DiskPacked& d = g.disks[i];
      const double v[8] = {d.c.x, d.c.y, d.c.z, d.n.x, d.n.y, d.n.z, d.rho, d.area};
      std::memcpy(&el[8 * i], v, sizeof(v));
    }
    rc = pii_geometry_create_disks(ctx, el.data(), (int)g.disks.size(), &out);
  } else {
    std::vector<double> el(8 * g.gaussians.size());
    for (size_t i = 0; i < g.gaussians.size(); ++i) {
      const GaussianPacked& d = g.gaussians[i];
      const double v[8] = {d.c.x, d.c.y, d.c.z, d.n.x, d.n.y, d.n.z, d.sigma, d.w};
      std::memcpy(&el[8 * i], v, sizeof(v));
    }
    rc = pii_geometry_create_gaussians(ctx, el.data(), (int)g.gaussians.size(), &out);
  }
  if (!report(rc, "upload geometry")) return nullptr;
  return out;
}

pii_kgrid* upload(pii_context* ctx, const KGrid& kg) {
  const int Q = (int)kg.kmag.size();
  std::vector<double> dirs(3 * (size_t)Q);
  for (int q = 0; q < Q; ++q)
    for (int c = 0; c < 3; ++c) dirs[3 * (size_t)q + c] = kg.dirs[q][c];
  pii_kgrid* out = nullptr;
  if (!report(pii_kgrid_create(ctx, Q, dirs.data(), kg.kmag.data(), kg.w.data(), &out),
              "upload k-grid"))
    return nullptr;
  return out;
}

}  // namespace detail

namespace {

using detail::context;
using detail::report;
using detail::upload;
using detail::CallObjects;

double3 d3(double x, double y, double z) {
  double3 v;
  v.x = x;
  v.y = y;
  v.z = z;
  return v;
}

}  // namespace

// ---- kernel registry (cpp/src/dof/cuda/dof_cuda_interface.cpp) ----
namespace detail {

struct RegistryEntry {
  CudaComputeAkFunc ak;
  CudaComputeAkGradientFunc grad;
  bool builtin = false;  // this library's kernels (the drivers' fused path)
};

std::unordered_map<std::string, RegistryEntry>& registry() {
  static std::unordered_map<std::string, RegistryEntry> r;
  return r;
}
std::mutex& registry_mutex() {
  static std::mutex m;
  return m;
}

void put_entry(const std::string& key, RegistryEntry e) {
  std::lock_guard<std::mutex> lk(registry_mutex());
  registry()[key] = std::move(e);
}

bool find_entry(const std::string& key, RegistryEntry& e) {
  std::lock_guard<std::mutex> lk(registry_mutex());
  auto it = registry().find(key);
  if (it == registry().end()) return false;
  e = it->second;
  return true;
}

// True when the registry entry of (dof type, geometry type) is one of this
// library's built-in entries.
bool builtin_entry(const std::string& dof_type, GeometryType t) {
  RegistryEntry e;
  return find_entry(CudaKernelRegistry::make_key(dof_type, t), e) && e.builtin;
}

KGrid one_node_grid(const Vector3d& k) {
  KGrid g;
  const double km = std::sqrt(k[0] * k[0] + k[1] * k[1] + k[2] * k[2]);
  g.dirs.push_back(km > 0 ? Vector3d{k[0] / km, k[1] / km, k[2] / km} : Vector3d{1.0, 0.0, 0.0});
  g.kmag.push_back(km);
  g.w.push_back(1.0);
  return g;
}

// Runs the registry entry of dof's type on geom over kgrid: A(k_q) into a
// (Q), and with need_grad the row-major Q x num_dofs dA/dtheta into g, through
// device buffers handed to the plug-ins.
bool run_plugin(const Geometry& geom, const KGrid& kgrid,
                const std::shared_ptr<DoFParameterization>& dof, const VectorXd& dofs,
                int blockSize, bool need_grad, std::vector<std::complex<double>>& a,
                std::vector<std::complex<double>>& g) {
  const std::string type = get_dof_type_name(dof);
  RegistryEntry e;
  if (!find_entry(CudaKernelRegistry::make_key(type, geom.type), e) || !e.ak ||
      (need_grad && !e.grad)) {
    std::cerr << "Error: CUDA kernels not available for DoF type '" << type
              << "' with geometry type " << (int)geom.type << std::endl;
    return false;
  }
  const size_t Q = kgrid.kmag.size(), nd = dofs.size();
  const size_t n_a = Q, n_g = need_grad ? Q * nd : 0;
  double2* d = nullptr;
  if (hipMalloc(reinterpret_cast<void**>(&d), sizeof(double2) * (n_a + n_g + 1)) != hipSuccess) {
    std::cerr << "Error: device allocation for the form-factor plug-in failed" << std::endl;
    return false;
  }
  bool ok = hipMemset(d, 0, sizeof(double2) * (n_a + n_g + 1)) == hipSuccess;
  if (ok) e.ak(geom, kgrid, dofs, d, blockSize);
  if (ok && need_grad) e.grad(geom, kgrid, dofs, d + n_a, blockSize);
  std::vector<double2> h(n_a + n_g);
  ok = ok && hipDeviceSynchronize() == hipSuccess &&
       hipMemcpy(h.data(), d, sizeof(double2) * (n_a + n_g), hipMemcpyDeviceToHost) == hipSuccess;
  (void)hipFree(d);
  if (!ok) {
    std::cerr << "Error: the form-factor plug-in of '" << type << "' failed" << std::endl;
    return false;
  }
  a.resize(n_a);
  for (size_t q = 0; q < n_a; ++q) a[q] = {h[q].x, h[q].y};
  g.resize(n_g);
  for (size_t i = 0; i < n_g; ++i) g[i] = {h[n_a + i].x, h[n_a + i].y};
  return true;
}

// Built-in entries: the library's FP64 form factor / Jacobian of one object
// written into the plug-in's device buffer (the plug-in contract of
// cpp/include/dof/cuda/dof_cuda_interface.hpp:24-40).
void builtin_ak(int kind, const Geometry& geom, const KGrid& kgrid, const VectorXd& dofs,
                double2* d_out, bool grad) {
  pii_context* ctx = context();
  if (!ctx) return;
  CallObjects co;
  co.kg = upload(ctx, kgrid);
  if (!co.kg) return;
  pii_geometry* g = upload(ctx, geom);
  if (!g) return;
  co.geoms.push_back(g);
  double* out = reinterpret_cast<double*>(d_out);
  if (grad)
    report(pii_form_factor_jacobian_device(ctx, g, kind, dofs.data(), (int)dofs.size(), co.kg,
                                           out),
           "compute_Ak_gradient plug-in");
  else
    report(pii_form_factor_device(ctx, g, kind, dofs.data(), (int)dofs.size(), co.kg,
                                  PII_FF_FP64, out),
           "compute_Ak plug-in");
}

void register_builtin(const char* dof_type, GeometryType t, int kind) {
  RegistryEntry e;
  e.ak = [kind](const Geometry& g, const KGrid& kg, const VectorXd& d, double2* out, int) {
    builtin_ak(kind, g, kg, d, out, false);
  };
  e.grad = [kind](const Geometry& g, const KGrid& kg, const VectorXd& d, double2* out, int) {
    builtin_ak(kind, g, kg, d, out, true);
  };
  e.builtin = true;
  put_entry(CudaKernelRegistry::make_key(dof_type, t), std::move(e));
}

}  // namespace detail

void CudaKernelRegistry::register_kernels(const std::string& dof_type, GeometryType geom_type,
                                          CudaComputeAkFunc compute_Ak,
                                          CudaComputeAkGradientFunc compute_Ak_gradient) {
  detail::RegistryEntry e;
  e.ak = std::move(compute_Ak);
  e.grad = std::move(compute_Ak_gradient);
  detail::put_entry(make_key(dof_type, geom_type), std::move(e));
}

bool CudaKernelRegistry::has_kernels(const std::string& dof_type, GeometryType geom_type) {
  detail::RegistryEntry e;
  return detail::find_entry(make_key(dof_type, geom_type), e);
}

std::pair<CudaComputeAkFunc, CudaComputeAkGradientFunc> CudaKernelRegistry::get_kernels(
    const std::string& dof_type, GeometryType geom_type) {
  detail::RegistryEntry e;
  if (!detail::find_entry(make_key(dof_type, geom_type), e)) return {};
  return {e.ak, e.grad};
}

std::string CudaKernelRegistry::make_key(const std::string& dof_type, GeometryType geom_type) {
  static const char* const names[] = {"Triangle", "Disk", "Gaussian"};
  const int t = (int)geom_type;
  return dof_type + "_" + (t >= 0 && t < 3 ? names[t] : "Unknown");
}

std::string get_dof_type_name(const std::shared_ptr<DoFParameterization>& dof) {
  if (!dof) return "Unknown";
  if (std::dynamic_pointer_cast<AffineDoF>(dof)) return "AffineDoF";
  if (std::dynamic_pointer_cast<TriangleMeshDoF>(dof)) return "TriangleMeshDoF";
  return "Unknown";
}

void register_affine_dof_cuda_kernels() {
  detail::register_builtin("AffineDoF", GEOM_TRIANGLE, PII_DOF_AFFINE);
  detail::register_builtin("AffineDoF", GEOM_DISK, PII_DOF_AFFINE);
  detail::register_builtin("AffineDoF", GEOM_GAUSSIAN, PII_DOF_AFFINE);
}

void register_triangle_mesh_dof_cuda_kernels() {
  detail::register_builtin("TriangleMeshDoF", GEOM_TRIANGLE, PII_DOF_TRIMESH);
}

namespace {
// the built-in entries, registered when the library loads (as the
// reference's static registrars do)
struct BuiltinRegistrar {
  BuiltinRegistrar() {
    register_affine_dof_cuda_kernels();
    register_triangle_mesh_dof_cuda_kernels();
  }
} g_builtin_registrar;
}  // namespace

int Geometry::num_elements() const {
  switch (type) {
    case GEOM_TRIANGLE: return (int)tris.size();
    case GEOM_DISK: return (int)disks.size();
    case GEOM_GAUSSIAN: return (int)gaussians.size();
  }
  return 0;
}

void set_precision(int ff_precision, int gram_mode) {
  g_ff = ff_precision;
  g_gram = gram_mode;
}

// ---- packing: cpp/src/geometry/packing.cpp:10-124 ----
std::vector<TriPacked> pack_tris(const MatrixXd& V, const MatrixXi& F) {
  std::vector<TriPacked> T(F.rows());
  for (int i = 0; i < F.rows(); ++i) {
    const int ia = F(i, 0), ib = F(i, 1), ic = F(i, 2);
    const double a[3] = {V(ia, 0), V(ia, 1), V(ia, 2)};
    const double e1[3] = {V(ib, 0) - a[0], V(ib, 1) - a[1], V(ib, 2) - a[2]};
    const double e2[3] = {V(ic, 0) - a[0], V(ic, 1) - a[1], V(ic, 2) - a[2]};
    T[i].a = d3(a[0], a[1], a[2]);
    T[i].e1 = d3(e1[0], e1[1], e1[2]);
    T[i].e2 = d3(e2[0], e2[1], e2[2]);
    T[i].S = d3(0.5 * (e1[1] * e2[2] - e1[2] * e2[1]), 0.5 * (e1[2] * e2[0] - e1[0] * e2[2]),
                0.5 * (e1[0] * e2[1] - e1[1] * e2[0]));
    T[i].vid0 = ia;
    T[i].vid1 = ib;
    T[i].vid2 = ic;
  }
  return T;
}

static void unit(const MatrixXd& N, int i, double* n) {
  const double l = std::sqrt(N(i, 0) * N(i, 0) + N(i, 1) * N(i, 1) + N(i, 2) * N(i, 2));
  for (int c = 0; c < 3; ++c) n[c] = l > 0 ? N(i, c) / l : N(i, c);
}

std::vector<DiskPacked> pack_disks(const MatrixXd& P, const MatrixXd& N, const VectorXd& r,
                                   bool is_radius) {
  std::vector<DiskPacked> D(P.rows());
  for (int i = 0; i < P.rows(); ++i) {
    double n[3];
    unit(N, i, n);
    D[i].c = d3(P(i, 0), P(i, 1), P(i, 2));
    D[i].n = d3(n[0], n[1], n[2]);
    if (is_radius) {
      D[i].rho = r[i];
      D[i].area = kPi * r[i] * r[i];
    } else {
      D[i].area = r[i];
      D[i].rho = std::sqrt(r[i] / kPi);
    }
  }
  return D;
}

std::vector<GaussianPacked> pack_gaussians(const MatrixXd& P, const MatrixXd& N,
                                           const VectorXd& sigmas, const VectorXd& weights) {
  std::vector<GaussianPacked> G(P.rows());
  for (int i = 0; i < P.rows(); ++i) {
    double n[3];
    unit(N, i, n);
    G[i].c = d3(P(i, 0), P(i, 1), P(i, 2));
    G[i].n = d3(n[0], n[1], n[2]);
    G[i].sigma = sigmas[i];
    G[i].w = weights[i];
  }
  return G;
}

Geometry make_triangle_mesh(const MatrixXd& V, const MatrixXi& F) {
  Geometry g(GEOM_TRIANGLE);
  g.tris = pack_tris(V, F);
  return g;
}

Geometry make_point_cloud(const MatrixXd& P, const MatrixXd& N, const VectorXd& r,
                          bool is_radius) {
  Geometry g(GEOM_DISK);
  g.disks = pack_disks(P, N, r, is_radius);
  return g;
}

Geometry make_gaussian_splat(const MatrixXd& P, const MatrixXd& N, const VectorXd& sigmas,
                             const VectorXd& weights) {
  Geometry g(GEOM_GAUSSIAN);
  g.gaussians = pack_gaussians(P, N, sigmas, weights);
  return g;
}

// cpp/src/geometry/geometry_helpers.cpp:16-38
void create_unit_cube_mesh(MatrixXd& V, MatrixXi& F) {
  static const double v[24] = {-0.5, -0.5, -0.5, 0.5, -0.5, -0.5, 0.5, 0.5, -0.5, -0.5, 0.5, -0.5,
                               -0.5, -0.5, 0.5,  0.5, -0.5, 0.5,  0.5, 0.5, 0.5,  -0.5, 0.5, 0.5};
  static const int f[36] = {0, 2, 1, 0, 3, 2, 4, 5, 6, 4, 6, 7, 0, 1, 5, 0, 5, 4,
                            2, 7, 6, 2, 3, 7, 0, 4, 7, 0, 7, 3, 1, 6, 5, 1, 2, 6};
  V = MatrixXd(8, 3);
  F = MatrixXi(12, 3);
  std::memcpy(V.data.data(), v, sizeof(v));
  std::memcpy(F.data.data(), f, sizeof(f));
}

// cpp/src/geometry/geometry_helpers.cpp:40-81
void create_sphere_pointcloud(MatrixXd& P, MatrixXd& N, VectorXd& radii, int n) {
  P = MatrixXd(n, 3);
  N = MatrixXd(n, 3);
  radii.assign(n, std::sqrt((4.0 * kPi / n) / kPi));
  const double ga = kPi * (3.0 - std::sqrt(5.0));
  for (int i = 0; i < n; ++i) {
    const double y = 1.0 - (2.0 * i) / (n - 1.0);
    const double ry = std::sqrt(1.0 - y * y);
    const double x = ry * std::cos(ga * i), z = ry * std::sin(ga * i);
    P(i, 0) = N(i, 0) = x;
    P(i, 1) = N(i, 1) = y;
    P(i, 2) = N(i, 2) = z;
  }
}

// ---- quadrature ----
std::pair<std::vector<double>, std::vector<double>> gauss_legendre_interval(int n, double a,
                                                                            double b) {
  std::vector<double> x(n), w(n);
  if (n > 0) report(pii_gauss_legendre(n, a, b, x.data(), w.data()), "gauss_legendre_interval");
  return {x, w};
}

LebedevGrid load_lebedev_txt(const std::string& path) {
  int M = 0;
  if (pii_lebedev_load(path.c_str(), nullptr, nullptr, 0, &M) != PII_OK)
    throw std::runtime_error("Failed to open Lebedev file: " + path);
  std::vector<double> d(3 * (size_t)M), w(M);
  pii_lebedev_load(path.c_str(), d.data(), w.data(), M, &M);
  LebedevGrid L;
  L.dirs.resize(M);
  for (int j = 0; j < M; ++j) L.dirs[j] = {d[3 * j], d[3 * j + 1], d[3 * j + 2]};
  L.weights = w;
  return L;
}

KGrid build_kgrid(const std::vector<std::array<double, 3>>& leb_dirs,
                  const std::vector<double>& leb_w, int Nrad) {
  const int M = (int)leb_w.size();
  std::vector<double> ld(3 * (size_t)M);
  for (int j = 0; j < M; ++j)
    for (int c = 0; c < 3; ++c) ld[3 * j + c] = leb_dirs[j][c];
  const size_t Q = (size_t)M * Nrad;
  std::vector<double> dirs(3 * Q), kmag(Q), w(Q);
  KGrid kg;
  if (!report(pii_kgrid_build(ld.data(), leb_w.data(), M, Nrad, dirs.data(), kmag.data(),
                              w.data()), "build_kgrid"))
    return kg;
  kg.dirs.resize(Q);
  for (size_t q = 0; q < Q; ++q) kg.dirs[q] = {dirs[3 * q], dirs[3 * q + 1], dirs[3 * q + 2]};
  kg.kmag = kmag;
  kg.w = w;
  return kg;
}

// ---- drivers ----
// hessian_ij[i][j] = block (i, j) of the (12n x 12n) Gauss-Newton matrix;
// hessian_ii / hessian_jj are zero in that form
// (compute_intersection_volume_multi_object.cu:186-227).
static void fill_hessian_blocks(IntersectionVolumeMatrixResult& res, const std::vector<int>& nd,
                                const std::vector<double>& H) {
  const int n = (int)nd.size();
  std::vector<size_t> o(n + 1, 0);
  for (int i = 0; i < n; ++i) o[i + 1] = o[i] + nd[i];
  const size_t D = o[n];
  res.hessian_ii.assign(n, std::vector<MatrixXd>(n));
  res.hessian_jj.assign(n, std::vector<MatrixXd>(n));
  res.hessian_ij.assign(n, std::vector<MatrixXd>(n));
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      res.hessian_ii[i][j] = MatrixXd::Zero(nd[i], nd[i]);
      res.hessian_jj[i][j] = MatrixXd::Zero(nd[j], nd[j]);
      MatrixXd& b = res.hessian_ij[i][j];
      b = MatrixXd(nd[i], nd[j]);
      for (int a = 0; a < nd[i]; ++a)
        for (int c = 0; c < nd[j]; ++c) b(a, c) = H[(o[i] + a) * D + o[j] + c];
    }
}

namespace detail {

// The matrix driver through registry plug-ins: every object's A(k_q) (and,
// for GRADIENT / HESSIAN, its Q x num_dofs dA/dtheta) from its entry's
// functions, then one FP64 device Gram over [A columns | derivative
// columns]: its blocks are V, grad_matrix[i][j] = (1/8 pi^3) sum_q w_q
// Re(dA_i conj A_j) and the Gauss-Newton blocks
// (compute_intersection_volume_multi_object.cu:437-600).
IntersectionVolumeMatrixResult volume_matrix_plugins(
    const std::vector<Geometry>& geoms, const std::vector<std::shared_ptr<DoFParameterization>>& dofs,
    const std::vector<VectorXd>& dof_vectors, const KGrid& kgrid, ComputationFlags flags,
    int blockSize) {
  IntersectionVolumeMatrixResult res;
  const int n = (int)geoms.size();
  res.volume_matrix = MatrixXd::Zero(n, n);
  const bool deriv = needs_gradient(flags) || needs_hessian(flags);
  const size_t Q = kgrid.kmag.size();
  std::vector<std::vector<std::complex<double>>> A(n), G(n);
  std::vector<size_t> off(n + 1, 0);
  for (int i = 0; i < n; ++i) {
    const auto d = dofs[i] ? dofs[i] : std::make_shared<AffineDoF>();
    if (!run_plugin(geoms[i], kgrid, d, dof_vectors[i], blockSize, deriv, A[i], G[i])) return res;
    off[i + 1] = off[i] + (deriv ? dof_vectors[i].size() : 0);
  }
  const size_t ncol = n + off[n];
  std::vector<double> J(2 * ncol * Q);
  for (int i = 0; i < n; ++i) {
    for (size_t q = 0; q < Q; ++q) {
      J[2 * ((size_t)i * Q + q)] = A[i][q].real();
      J[2 * ((size_t)i * Q + q) + 1] = A[i][q].imag();
    }
    const size_t nd = off[i + 1] - off[i];
    for (size_t t = 0; t < nd; ++t)
      for (size_t q = 0; q < Q; ++q) {
        const std::complex<double> v = G[i][q * nd + t];
        const size_t c = n + off[i] + t;
        J[2 * (c * Q + q)] = v.real();
        J[2 * (c * Q + q) + 1] = v.imag();
      }
  }
  pii_context* ctx = context();
  if (!ctx) return res;
  CallObjects co;
  co.kg = upload(ctx, kgrid);
  if (!co.kg) return res;
  std::vector<double> W(ncol * ncol);
  if (!report(pii_gram(ctx, J.data(), (int)ncol, co.kg, PII_GRAM_FP64, W.data()),
              "compute_intersection_volume_matrix_cuda"))
    return res;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) res.volume_matrix(i, j) = W[(size_t)i * ncol + j];
  if (needs_gradient(flags)) {
    res.grad_matrix.assign(n, std::vector<VectorXd>(n));
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        VectorXd& g = res.grad_matrix[i][j];
        g.resize(off[i + 1] - off[i]);
        for (size_t t = 0; t < g.size(); ++t) g[t] = W[(n + off[i] + t) * ncol + j];
      }
  }
  if (needs_hessian(flags)) {
    res.hessian_ii.assign(n, std::vector<MatrixXd>(n));
    res.hessian_jj.assign(n, std::vector<MatrixXd>(n));
    res.hessian_ij.assign(n, std::vector<MatrixXd>(n));
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        const int di = (int)(off[i + 1] - off[i]), dj = (int)(off[j + 1] - off[j]);
        res.hessian_ii[i][j] = MatrixXd::Zero(di, di);
        res.hessian_jj[i][j] = MatrixXd::Zero(dj, dj);
        MatrixXd& b = res.hessian_ij[i][j];
        b = MatrixXd(di, dj);
        for (int a = 0; a < di; ++a)
          for (int c = 0; c < dj; ++c) b(a, c) = W[(n + off[i] + a) * ncol + n + off[j] + c];
      }
  }
  return res;
}

}  // namespace detail

IntersectionVolumeMatrixResult compute_intersection_volume_matrix_cuda(
    const std::vector<Geometry>& geoms, const std::vector<std::shared_ptr<DoFParameterization>>& dofs,
    const std::vector<VectorXd>& dof_vectors, const KGrid& kgrid, ComputationFlags flags,
    int blockSize, bool enable_profiling) {
  IntersectionVolumeMatrixResult res;
  const int n = (int)geoms.size();
  res.volume_matrix = MatrixXd::Zero(n, n);
  if (n == 0) return res;
  if ((int)dofs.size() != n || (int)dof_vectors.size() != n) {
    std::cerr << "Error: Mismatch between number of geometries (" << n
              << ") and number of DoF parameterizations (" << dofs.size()
              << ") or DoF vectors (" << dof_vectors.size() << ")" << std::endl;
    return res;
  }
  // every object goes through its registry entry
  // (compute_intersection_volume_multi_object.cu:263-290): this library's
  // fused path when all are built-in, the plug-ins otherwise
  bool all_builtin = true;
  for (int i = 0; i < n; ++i) {
    const std::string type = dofs[i] ? get_dof_type_name(dofs[i]) : "AffineDoF";
    detail::RegistryEntry e;
    if (!detail::find_entry(CudaKernelRegistry::make_key(type, geoms[i].type), e)) {
      std::cerr << "Error: CUDA kernels not available for DoF type '" << type
                << "' with geometry type " << (int)geoms[i].type << std::endl;
      return res;
    }
    all_builtin = all_builtin && e.builtin && (!dofs[i] || dofs[i]->dof_kind() >= 0);
  }
  if (!all_builtin)
    return detail::volume_matrix_plugins(geoms, dofs, dof_vectors, kgrid, flags, blockSize);
  pii_context* ctx = context();
  if (!ctx) return res;
  CallObjects co;
  co.kg = upload(ctx, kgrid);
  if (!co.kg) return res;
  std::vector<int> kinds(n);
  for (int i = 0; i < n; ++i) {
    pii_geometry* g = upload(ctx, geoms[i]);
    if (!g) return res;
    co.geoms.push_back(g);
    kinds[i] = dofs[i] ? dofs[i]->dof_kind() : PII_DOF_AFFINE;
  }
  std::vector<const pii_geometry*> gc(co.geoms.begin(), co.geoms.end());
  return detail::volume_matrix_resident(ctx, gc, kinds, dof_vectors, co.kg,
                                        (int)kgrid.kmag.size(), flags, enable_profiling);
}

namespace detail {

IntersectionVolumeMatrixResult volume_matrix_resident(
    pii_context* ctx, const std::vector<const pii_geometry*>& gc, const std::vector<int>& kinds,
    const std::vector<VectorXd>& dof_vectors, const pii_kgrid* kg, int Q, ComputationFlags flags,
    bool enable_profiling) {
  IntersectionVolumeMatrixResult res;
  const int n = (int)gc.size();
  res.volume_matrix = MatrixXd::Zero(n, n);
  if (n == 0) return res;
  std::vector<int> nd(n);
  std::vector<const double*> dp(n);
  for (int i = 0; i < n; ++i) {
    nd[i] = (int)dof_vectors[i].size();
    dp[i] = dof_vectors[i].data();
  }
  if (enable_profiling) pii_context_set_profiling(ctx, 1);
  bool all_affine = true;
  for (int i = 0; i < n; ++i) all_affine = all_affine && kinds[i] == PII_DOF_AFFINE;
  // one Gram over the A and 12 n derivative columns while they fit its
  // column limit (16,384); beyond it, one VJP per pair as for other DoFs
  const bool one_pass = all_affine && (needs_gradient(flags) || needs_hessian(flags)) &&
                        13LL * n <= 16384;
  auto print_profile = [&] {
    if (!enable_profiling) return;
    double ms[4];
    pii_context_timings(ctx, ms, 4);
    std::cout << "\n=== MI355X Multi-Object Volume Computation Profiler ===\n"
              << "Number of objects: " << n << "\nK-grid nodes: " << Q
              << "\n  Pre-pass:    " << ms[0] << " ms\n  Form factor: " << ms[1]
              << " ms\n  Gram:        " << ms[2] << " ms";
    if (one_pass) std::cout << "\n  dA/dtheta:   " << ms[3] << " ms";
    std::cout << std::endl;
    pii_context_set_profiling(ctx, 0);
  };
  if (one_pass) {
    // AffineDoF objects: as the reference's own GRADIENT / HESSIAN branch,
    // dA/dtheta columns next to A, and V, grad_matrix and the Gauss-Newton
    // blocks from one Gram (compute_intersection_volume_multi_object.cu:437-600).
    const int D = 12 * n;
    std::vector<double> G((size_t)n * n * 12), H;
    if (needs_hessian(flags)) H.resize((size_t)D * D);
    if (!report(pii_volume_matrix_affine_derivatives(ctx, n, gc.data(), dp.data(), kg, g_gram,
                                                     res.volume_matrix.data.data(), G.data(),
                                                     H.empty() ? nullptr : H.data()),
                "compute_intersection_volume_matrix_cuda")) {
      res.volume_matrix = MatrixXd::Zero(n, n);
      return res;
    }
    print_profile();
    // the FP64 form-factor setting keeps V on the FP64 path
    if (g_ff != PII_FF_FAST &&
        !report(pii_volume_matrix(ctx, n, gc.data(), kinds.data(), dp.data(), nd.data(), kg,
                                  g_ff, g_gram, res.volume_matrix.data.data()),
                "compute_intersection_volume_matrix_cuda")) {
      res.volume_matrix = MatrixXd::Zero(n, n);
      return res;
    }
    if (needs_gradient(flags)) {
      res.grad_matrix.assign(n, std::vector<VectorXd>(n));
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
          res.grad_matrix[i][j].assign(G.begin() + ((size_t)i * n + j) * 12,
                                       G.begin() + ((size_t)i * n + j + 1) * 12);
    }
    if (needs_hessian(flags)) fill_hessian_blocks(res, nd, H);
    return res;
  }
  size_t dtot = 0;
  for (int i = 0; i < n; ++i) dtot += nd[i];
  if ((needs_gradient(flags) || needs_hessian(flags)) && dtot + n <= 16384) {
    // Other DoF kinds, as the reference's GRADIENT / HESSIAN branch: dense
    // dA/dtheta, then grad_matrix and the Gauss-Newton blocks from one Gram
    // (compute_intersection_volume_multi_object.cu:437-600).
    std::vector<double> G(n * dtot), H;
    if (needs_hessian(flags)) H.resize(dtot * dtot);
    if (!report(pii_volume_matrix_derivatives(ctx, n, gc.data(), kinds.data(), dp.data(),
                                              nd.data(), kg, g_gram, res.volume_matrix.data.data(),
                                              G.data(), H.empty() ? nullptr : H.data()),
                "compute_intersection_volume_matrix_cuda")) {
      res.volume_matrix = MatrixXd::Zero(n, n);
      return res;
    }
    print_profile();
    if (g_ff != PII_FF_FAST &&
        !report(pii_volume_matrix(ctx, n, gc.data(), kinds.data(), dp.data(), nd.data(), kg,
                                  g_ff, g_gram, res.volume_matrix.data.data()),
                "compute_intersection_volume_matrix_cuda")) {
      res.volume_matrix = MatrixXd::Zero(n, n);
      return res;
    }
    if (needs_gradient(flags)) {
      res.grad_matrix.assign(n, std::vector<VectorXd>(n));
      size_t o = 0;
      for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) {
          const auto b = G.begin() + (ptrdiff_t)(n * o + (size_t)j * nd[i]);
          res.grad_matrix[i][j].assign(b, b + nd[i]);
        }
        o += nd[i];
      }
    }
    if (needs_hessian(flags)) fill_hessian_blocks(res, nd, H);
    return res;
  }
  if (!report(pii_volume_matrix(ctx, n, gc.data(), kinds.data(), dp.data(), nd.data(), kg,
                                g_ff, g_gram, res.volume_matrix.data.data()),
              "compute_intersection_volume_matrix_cuda")) {
    res.volume_matrix = MatrixXd::Zero(n, n);
    return res;
  }
  print_profile();
  if (needs_gradient(flags)) {
    // More DoFs than one Gram takes: grad_matrix[i][j] = dV_ij/dtheta_i
    // from one VJP per unordered pair.
    res.grad_matrix.assign(n, std::vector<VectorXd>(n));
    std::vector<double> G((size_t)n * n);
    std::vector<VectorXd> grads(n);
    std::vector<double*> gp(n);
    for (int i = 0; i < n; ++i) {
      grads[i].assign(nd[i], 0.0);
      gp[i] = grads[i].data();
    }
    for (int i = 0; i < n; ++i) {
      for (int j = i; j < n; ++j) {
        std::fill(G.begin(), G.end(), 0.0);
        G[(size_t)i * n + j] = 1.0;
        if (!report(pii_volume_matrix_vjp(ctx, n, gc.data(), kinds.data(), dp.data(), nd.data(),
                                          kg, G.data(), gp.data(), nullptr),
                    "gradient")) {
          res.grad_matrix.clear();
          return res;
        }
        if (i == j) {
          res.grad_matrix[i][i] = grads[i];
          for (double& x : res.grad_matrix[i][i]) x *= 0.5;
        } else {
          res.grad_matrix[i][j] = grads[i];
          res.grad_matrix[j][i] = grads[j];
        }
      }
    }
  }
  if (needs_hessian(flags)) {
    // Gauss-Newton blocks (compute_intersection_volume_multi_object.cpp:236-292):
    // hessian_ij[i][j] from the device Gram of dA/dtheta columns; hessian_ii /
    // hessian_jj are zero in this form.
    size_t D = 0;
    for (int i = 0; i < n; ++i) D += nd[i];
    std::vector<double> H(D * D);
    if (!report(pii_volume_matrix_hessian(ctx, n, gc.data(), kinds.data(), dp.data(), nd.data(),
                                          kg, g_gram, H.data(), nullptr),
                "hessian"))
      return res;
    fill_hessian_blocks(res, nd, H);
  }
  return res;
}

}  // namespace detail

IntersectionVolumeMatrixResult compute_intersection_volume_matrix_cuda(
    const std::vector<Geometry>& geometries, const KGrid& kgrid, int blockSize,
    bool enable_profiling) {
  const int n = (int)geometries.size();
  std::vector<std::shared_ptr<DoFParameterization>> dofs(n, std::make_shared<AffineDoF>());
  const VectorXd id = {0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1};
  std::vector<VectorXd> dv(n, id);
  return compute_intersection_volume_matrix_cuda(geometries, dofs, dv, kgrid,
                                                 ComputationFlags::VOLUME_ONLY, blockSize,
                                                 enable_profiling);
}

IntersectionVolumeResult compute_intersection_volume_cuda(
    const Geometry& g1, const Geometry& g2, const std::shared_ptr<DoFParameterization>& dof1,
    const std::shared_ptr<DoFParameterization>& dof2, const VectorXd& dofs1,
    const VectorXd& dofs2, const KGrid& kgrid, ComputationFlags flags, int blockSize,
    bool enable_profiling) {
  IntersectionVolumeResult r;
  auto m = compute_intersection_volume_matrix_cuda({g1, g2}, {dof1, dof2}, {dofs1, dofs2}, kgrid,
                                                   flags, blockSize, enable_profiling);
  if (m.volume_matrix.rows() == 2) r.volume = m.volume_matrix(0, 1);
  if (needs_gradient(flags) && m.grad_matrix.size() == 2) {
    r.grad_geom1 = m.grad_matrix[0][1];
    r.grad_geom2 = m.grad_matrix[1][0];
  }
  // compute_intersection_volume_cpu.cpp:116-149: geom1 / geom2 / cross blocks
  if (needs_hessian(flags) && m.hessian_ij.size() == 2) {
    r.hessian_geom1 = m.hessian_ij[0][0];
    r.hessian_geom2 = m.hessian_ij[1][1];
    r.hessian_cross = m.hessian_ij[0][1];
  }
  return r;
}

double compute_intersection_volume_cuda(const Geometry& g1, const Geometry& g2,
                                        const KGrid& kgrid, int blockSize,
                                        bool enable_profiling) {
  auto m = compute_intersection_volume_matrix_cuda({g1, g2}, kgrid, blockSize, enable_profiling);
  return m.volume_matrix.rows() == 2 ? m.volume_matrix(0, 1) : 0.0;
}

namespace detail {

// A(k) of a built-in DoF kind at one k-vector, FP64 on the device.
std::complex<double> ak_at(const Geometry& geom, const Vector3d& k, int kind,
                           const VectorXd& dofs, const char* who) {
  pii_context* ctx = context();
  if (!ctx) return 0.0;
  CallObjects co;
  const double km = std::sqrt(k[0] * k[0] + k[1] * k[1] + k[2] * k[2]);
  const double dir[3] = {km > 0 ? k[0] / km : 1.0, km > 0 ? k[1] / km : 0.0,
                         km > 0 ? k[2] / km : 0.0};
  const double w = 1.0;
  if (!report(pii_kgrid_create(ctx, 1, dir, &km, &w, &co.kg), who)) return 0.0;
  pii_geometry* g = upload(ctx, geom);
  if (!g) return 0.0;
  co.geoms.push_back(g);
  double out[2] = {0, 0};
  if (!report(pii_form_factor(ctx, g, kind, dofs.data(), (int)dofs.size(), co.kg, PII_FF_FP64,
                              out),
              who))
    return 0.0;
  return {out[0], out[1]};
}

// dA(k)/dtheta of a built-in DoF kind at one k-vector
// (cpp/src/compute_intersection_volume.cu:558-626: zeros for |k| < 1e-10).
VectorXcd ak_gradient_at(const Geometry& geom, const Vector3d& k, int kind, const VectorXd& dofs,
                         const char* who) {
  const int nd = (int)dofs.size();
  VectorXcd zero(nd, {0.0, 0.0});
  pii_context* ctx = context();
  if (!ctx) return zero;
  const double km = std::sqrt(k[0] * k[0] + k[1] * k[1] + k[2] * k[2]);
  if (km < 1e-10) return zero;
  CallObjects co;
  const double dir[3] = {k[0] / km, k[1] / km, k[2] / km};
  const double w = 1.0;
  if (!report(pii_kgrid_create(ctx, 1, dir, &km, &w, &co.kg), who)) return zero;
  pii_geometry* g = upload(ctx, geom);
  if (!g) return zero;
  co.geoms.push_back(g);
  std::vector<double> out(2 * (size_t)nd);
  if (!report(pii_form_factor_jacobian(ctx, g, kind, dofs.data(), nd, co.kg, out.data()), who))
    return zero;
  VectorXcd grad(nd);
  for (int t = 0; t < nd; ++t) grad[t] = {out[2 * t], out[2 * t + 1]};
  return grad;
}

}  // namespace detail

std::complex<double> AffineDoF::compute_A(const Geometry& geom, const Vector3d& k,
                                          const VectorXd& dofs) const {
  return detail::ak_at(geom, k, PII_DOF_AFFINE, dofs, "AffineDoF::compute_A");
}
VectorXcd AffineDoF::compute_A_gradient(const Geometry& geom, const Vector3d& k,
                                        const VectorXd& dofs) const {
  return detail::ak_gradient_at(geom, k, PII_DOF_AFFINE, dofs, "AffineDoF::compute_A_gradient");
}
std::complex<double> TriangleMeshDoF::compute_A(const Geometry& geom, const Vector3d& k,
                                                const VectorXd& dofs) const {
  return detail::ak_at(geom, k, PII_DOF_TRIMESH, dofs, "TriangleMeshDoF::compute_A");
}
VectorXcd TriangleMeshDoF::compute_A_gradient(const Geometry& geom, const Vector3d& k,
                                              const VectorXd& dofs) const {
  return detail::ak_gradient_at(geom, k, PII_DOF_TRIMESH, dofs,
                                "TriangleMeshDoF::compute_A_gradient");
}

// cpp/src/compute_intersection_volume.cu:495-556: through the registry entry
// of the DoF type (a caller's plug-in on a one-node k-grid)
std::complex<double> compute_Ak_cuda(const Geometry& geom, const Vector3d& k,
                                     const std::shared_ptr<DoFParameterization>& dof,
                                     const VectorXd& dofs, int blockSize) {
  const int kind = dof ? dof->dof_kind() : PII_DOF_AFFINE;
  if (dof && !detail::builtin_entry(get_dof_type_name(dof), geom.type)) {
    std::vector<std::complex<double>> a, g;
    if (!detail::run_plugin(geom, detail::one_node_grid(k), dof, dofs, blockSize, false, a, g))
      return 0.0;
    return a[0];
  }
  return detail::ak_at(geom, k, kind, dofs, "compute_Ak_cuda");
}

VectorXcd compute_Ak_gradient_cuda(const Geometry& geom, const Vector3d& k,
                                   const std::shared_ptr<DoFParameterization>& dof,
                                   const VectorXd& dofs, int blockSize) {
  const int kind = dof ? dof->dof_kind() : PII_DOF_AFFINE;
  if (dof && !detail::builtin_entry(get_dof_type_name(dof), geom.type)) {
    const int nd = (int)dofs.size();
    std::vector<std::complex<double>> a, g;
    if (!detail::run_plugin(geom, detail::one_node_grid(k), dof, dofs, blockSize, true, a, g))
      return VectorXcd(nd, {0.0, 0.0});
    return VectorXcd(g.begin(), g.end());
  }
  return detail::ak_gradient_at(geom, k, kind, dofs, "compute_Ak_gradient_cuda");
}

// cpp/src/compute_volume.cpp:61-73 with cpp/src/dof/affine_dof.cpp:324-384 (identity DoFs)
double compute_volume_cpu(const Geometry& g, bool) {
  double v = 0.0;
  if (g.type == GEOM_TRIANGLE) {
    for (const TriPacked& t : g.tris) v += t.a.x * t.S.x + t.a.y * t.S.y + t.a.z * t.S.z;
  } else if (g.type == GEOM_DISK) {
    for (const DiskPacked& d : g.disks)
      v += (d.c.x * d.n.x + d.c.y * d.n.y + d.c.z * d.n.z) * d.area;
  } else {
    for (const GaussianPacked& d : g.gaussians)
      v += (d.c.x * d.n.x + d.c.y * d.n.y + d.c.z * d.n.z) * d.w;
  }
  return v / 3.0;
}

// cpp/src/dof/affine_dof.cpp:324-384
double AffineDoF::compute_volume(const Geometry& g, const VectorXd& dofs) const {
  if (dofs.size() != 12) {
    std::cerr << "AffineDoF::compute_volume: expected 12 DoFs, got " << dofs.size() << std::endl;
    return 0.0;
  }
  const double* A = dofs.data() + 3;
  const double det = A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
                     A[2] * (A[3] * A[7] - A[4] * A[6]);
  return compute_volume_cpu(g) * std::fabs(det);
}

VectorXd AffineDoF::compute_volume_gradient(const Geometry& g, const VectorXd& dofs) const {
  VectorXd grad(12, 0.0);
  if (dofs.size() != 12) {
    std::cerr << "AffineDoF::compute_volume_gradient: expected 12 DoFs, got " << dofs.size()
              << std::endl;
    return grad;
  }
  const double* A = dofs.data() + 3;  // row-major
  // cofactors C_rc = d det A / d A_rc
  const double C[9] = {A[4] * A[8] - A[5] * A[7], A[5] * A[6] - A[3] * A[8],
                       A[3] * A[7] - A[4] * A[6], A[2] * A[7] - A[1] * A[8],
                       A[0] * A[8] - A[2] * A[6], A[1] * A[6] - A[0] * A[7],
                       A[1] * A[5] - A[2] * A[4], A[2] * A[3] - A[0] * A[5],
                       A[0] * A[4] - A[1] * A[3]};
  const double det = A[0] * C[0] + A[1] * C[1] + A[2] * C[2];
  const double sign = det >= 0.0 ? 1.0 : -1.0;
  const double v_ref = compute_volume_cpu(g);
  for (int i = 0; i < 9; ++i) grad[3 + i] = sign * C[i] * v_ref;
  return grad;
}

double TriangleMeshDoF::compute_volume(const Geometry& g, const VectorXd& dofs) const {
  if (g.type != GEOM_TRIANGLE || (int)dofs.size() != num_dofs()) {
    std::cerr << "TriangleMeshDoF::compute_volume: needs a triangle mesh and " << num_dofs()
              << " DoFs" << std::endl;
    return 0.0;
  }
  const int nv = nv_;
  double v = 0.0;
  for (const TriPacked& t : g.tris) {
    const int i = t.vid0, j = t.vid1, k = t.vid2;
    if (i < 0 || j < 0 || k < 0 || i >= nv || j >= nv || k >= nv) continue;
    const double* a = &dofs[3 * (size_t)i];
    const double* b = &dofs[3 * (size_t)j];
    const double* c = &dofs[3 * (size_t)k];
    v += a[0] * (b[1] * c[2] - b[2] * c[1]) + a[1] * (b[2] * c[0] - b[0] * c[2]) +
         a[2] * (b[0] * c[1] - b[1] * c[0]);
  }
  return v / 6.0;
}

VectorXd TriangleMeshDoF::compute_volume_gradient(const Geometry& g, const VectorXd& dofs) const {
  VectorXd grad(num_dofs(), 0.0);
  if (g.type != GEOM_TRIANGLE || (int)dofs.size() != num_dofs()) {
    std::cerr << "TriangleMeshDoF::compute_volume_gradient: needs a triangle mesh and "
              << num_dofs() << " DoFs" << std::endl;
    return grad;
  }
  const int nv = nv_;
  auto cross_into = [&](int o, const double* p, const double* q) {
    grad[3 * (size_t)o] += p[1] * q[2] - p[2] * q[1];
    grad[3 * (size_t)o + 1] += p[2] * q[0] - p[0] * q[2];
    grad[3 * (size_t)o + 2] += p[0] * q[1] - p[1] * q[0];
  };
  for (const TriPacked& t : g.tris) {
    const int i = t.vid0, j = t.vid1, k = t.vid2;
    if (i < 0 || j < 0 || k < 0 || i >= nv || j >= nv || k >= nv) continue;
    const double* a = &dofs[3 * (size_t)i];
    const double* b = &dofs[3 * (size_t)j];
    const double* c = &dofs[3 * (size_t)k];
    cross_into(i, b, c);
    cross_into(j, c, a);
    cross_into(k, a, b);
  }
  for (double& x : grad) x /= 6.0;
  return grad;
}

// cpp/src/compute_volume.cpp:15-59
VolumeResult compute_volume_cpu(const Geometry& ref_geom,
                                const std::shared_ptr<DoFParameterization>& dof,
                                const VectorXd& dofs, ComputationFlags flags,
                                bool enable_profiling) {
  const auto t0 = std::chrono::high_resolution_clock::now();
  VolumeResult r;
  if (!dof) {
    std::cerr << "compute_volume_cpu: NULL DoF parameterization" << std::endl;
    return r;
  }
  if (needs_volume(flags)) r.volume = dof->compute_volume(ref_geom, dofs);
  if (needs_gradient(flags)) r.grad = dof->compute_volume_gradient(ref_geom, dofs);
  if (enable_profiling) {
    const double ms = std::chrono::duration<double, std::milli>(
                          std::chrono::high_resolution_clock::now() - t0)
                          .count();
    std::cout << std::fixed << std::setprecision(3)
              << "\n=== CPU Volume Computation (Unified Interface) ===\nGeometry: "
              << ref_geom.num_elements() << " elements\nDoF type: " << dof->num_dofs()
              << " DoFs\nFlags: " << static_cast<int>(flags) << "\n  Total time: " << ms
              << " ms" << std::endl;
  }
  return r;
}

// cpp/src/analytical_intersection.cpp:15-60
double box_box_intersection_volume(const Vector3d& c1, const Vector3d& h1, const Vector3d& c2,
                                   const Vector3d& h2) {
  double v = 1.0;
  for (int i = 0; i < 3; ++i) {
    const double lo = std::max(c1[i] - h1[i], c2[i] - h2[i]);
    const double hi = std::min(c1[i] + h1[i], c2[i] + h2[i]);
    v *= std::max(0.0, hi - lo);
  }
  return v;
}

double sphere_sphere_intersection_volume(const Vector3d& c1, double r1, const Vector3d& c2,
                                         double r2) {
  const double dx = c2[0] - c1[0], dy = c2[1] - c1[1], dz = c2[2] - c1[2];
  const double d = std::sqrt(dx * dx + dy * dy + dz * dz);
  if (d >= r1 + r2) return 0.0;
  if (d <= std::abs(r1 - r2)) {
    const double rm = std::min(r1, r2);
    return 4.0 * kPi / 3.0 * rm * rm * rm;
  }
  const double t1 = (r1 + r2 - d) * (r1 + r2 - d);
  const double t2 = d * d + 2.0 * d * (r1 + r2) - 3.0 * (r1 * r1 + r2 * r2) + 6.0 * r1 * r2;
  return kPi * t1 * t2 / (12.0 * d);
}

}  // namespace PoIntInt
