// Example C++ caller: the two-cube configuration through the reference-named
// C++ API (Lebedev 302 x 32 k-grid, unit cube and a copy shifted by
// (0.3, 0.2, 0.1)).  Prints the 2 x 2 volume matrix, the pairwise volume and
// its gradient w.r.t. the second cube's affine DoFs, the translation corner
// of the pair's Gauss-Newton cross Hessian, the second cube's gradient and
// the whole cross Hessian under vertex DoFs, and A(k) with dA(k)/dtheta of
// the second cube at one k under an affine map.
//
// Mode "mismatch" runs no device code: the matrix driver called with fewer
// DoF parameterisations / DoF vectors than geometries prints its message and
// returns the zero matrix, as the reference's driver does
// (cpp/src/compute_intersection_volume_multi_object.cu:255-261).
//
// Mode "registry" (FP64): a caller's own DoF kind, x' = s x + t with DoFs
// [tx, ty, tz, s], reached through CudaKernelRegistry with plug-ins built on
// the registry's AffineDoF ones; the same pair under the equivalent
// AffineDoF; the DoF methods (A, dA/dtheta at one k, volume and its
// gradient) and compute_volume_cpu with GRADIENT.
//
//   usage: example_cube_pair <lebedev_029.txt> [fp64|fast|mismatch|registry]
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstring>

#include "pointint/pointint.hpp"
#include "pointint_hip.h"

using namespace PoIntInt;

namespace {

// x' = s x + t: the AffineDoF [t, s I]
struct ScaleTranslateDoF : DoFParameterization {
  int num_dofs() const override { return 4; }
  static VectorXd affine(const VectorXd& d) {
    return {d[0], d[1], d[2], d[3], 0.0, 0.0, 0.0, d[3], 0.0, 0.0, 0.0, d[3]};
  }
  std::complex<double> compute_A(const Geometry& g, const Vector3d& k,
                                 const VectorXd& d) const override {
    return AffineDoF().compute_A(g, k, affine(d));
  }
  VectorXcd compute_A_gradient(const Geometry& g, const Vector3d& k,
                               const VectorXd& d) const override {
    const VectorXcd a = AffineDoF().compute_A_gradient(g, k, affine(d));
    return {a[0], a[1], a[2], a[3] + a[7] + a[11]};
  }
  double compute_volume(const Geometry& g, const VectorXd& d) const override {
    return AffineDoF().compute_volume(g, affine(d));
  }
  VectorXd compute_volume_gradient(const Geometry& g, const VectorXd& d) const override {
    const VectorXd a = AffineDoF().compute_volume_gradient(g, affine(d));
    return {a[0], a[1], a[2], a[3] + a[7] + a[11]};
  }
};

int registry_mode(const std::vector<Geometry>& geoms, const KGrid& kg, const MatrixXd& V,
                  const MatrixXd& V2) {
  std::printf("has %d %d %d %d %d %d\n",
              (int)CudaKernelRegistry::has_kernels("AffineDoF", GEOM_TRIANGLE),
              (int)CudaKernelRegistry::has_kernels("AffineDoF", GEOM_DISK),
              (int)CudaKernelRegistry::has_kernels("AffineDoF", GEOM_GAUSSIAN),
              (int)CudaKernelRegistry::has_kernels("TriangleMeshDoF", GEOM_TRIANGLE),
              (int)CudaKernelRegistry::has_kernels("TriangleMeshDoF", GEOM_DISK),
              (int)CudaKernelRegistry::has_kernels("Unknown", GEOM_TRIANGLE));
  auto st = std::make_shared<ScaleTranslateDoF>();
  std::printf("name %s %s %s\n", get_dof_type_name(std::make_shared<AffineDoF>()).c_str(),
              get_dof_type_name(std::make_shared<TriangleMeshDoF>(8)).c_str(),
              get_dof_type_name(st).c_str());
  const VectorXd d1 = {0.0, 0.0, 0.0, 1.0}, d2 = {0.3, 0.2, 0.1, 1.1};
  // without plug-ins for its type the driver reports and returns zeros
  auto r0 = compute_intersection_volume_matrix_cuda(geoms, {st, st}, {d1, d2}, kg);
  std::printf("unregistered %.17g\n", r0.volume_matrix(0, 1));
  const auto base = CudaKernelRegistry::get_kernels("AffineDoF", GEOM_TRIANGLE);
  CudaKernelRegistry::register_kernels(
      "Unknown", GEOM_TRIANGLE,
      [base](const Geometry& g, const KGrid& k, const VectorXd& d, double2* out, int bs) {
        base.first(g, k, ScaleTranslateDoF::affine(d), out, bs);
      },
      [base](const Geometry& g, const KGrid& k, const VectorXd& d, double2* out, int bs) {
        const size_t Q = k.kmag.size();
        double2* tmp = nullptr;
        if (hipMalloc(reinterpret_cast<void**>(&tmp), sizeof(double2) * 12 * Q) != hipSuccess)
          return;
        base.second(g, k, ScaleTranslateDoF::affine(d), tmp, bs);
        std::vector<double2> h(12 * Q), o(4 * Q);
        (void)hipMemcpy(h.data(), tmp, sizeof(double2) * 12 * Q, hipMemcpyDeviceToHost);
        (void)hipFree(tmp);
        for (size_t q = 0; q < Q; ++q) {
          for (int c = 0; c < 3; ++c) o[4 * q + c] = h[12 * q + c];
          o[4 * q + 3].x = h[12 * q + 3].x + h[12 * q + 7].x + h[12 * q + 11].x;
          o[4 * q + 3].y = h[12 * q + 3].y + h[12 * q + 7].y + h[12 * q + 11].y;
        }
        (void)hipMemcpy(out, o.data(), sizeof(double2) * 4 * Q, hipMemcpyHostToDevice);
      });
  auto rc = compute_intersection_volume_matrix_cuda(geoms, {st, st}, {d1, d2}, kg,
                                                    ComputationFlags::ALL);
  auto aff = std::make_shared<AffineDoF>();
  auto ra = compute_intersection_volume_matrix_cuda(
      geoms, {aff, aff}, {ScaleTranslateDoF::affine(d1), ScaleTranslateDoF::affine(d2)}, kg,
      ComputationFlags::ALL);
  if (rc.grad_matrix.size() != 2 || ra.grad_matrix.size() != 2 || rc.hessian_ij.size() != 2)
    return 1;
  std::printf("Vc %.17g %.17g %.17g\n", rc.volume_matrix(0, 0), rc.volume_matrix(0, 1),
              rc.volume_matrix(1, 1));
  std::printf("Va %.17g %.17g %.17g\n", ra.volume_matrix(0, 0), ra.volume_matrix(0, 1),
              ra.volume_matrix(1, 1));
  std::printf("gc");
  for (double x : rc.grad_matrix[1][0]) std::printf(" %.17g", x);
  std::printf("\nga");
  for (double x : ra.grad_matrix[1][0]) std::printf(" %.17g", x);
  std::printf("\nhc");
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) std::printf(" %.17g", rc.hessian_ij[0][1](a, b));
  std::printf("\nha");
  for (int a = 0; a < 12; ++a)
    for (int b = 0; b < 12; ++b) std::printf(" %.17g", ra.hessian_ij[0][1](a, b));
  std::printf("\n");
  // a replaced built-in entry is used as registered, then restored
  int calls = 0;
  CudaKernelRegistry::register_kernels(
      "AffineDoF", GEOM_TRIANGLE,
      [base, &calls](const Geometry& g, const KGrid& k, const VectorXd& d, double2* out, int bs) {
        ++calls;
        base.first(g, k, d, out, bs);
      },
      base.second);
  auto rr = compute_intersection_volume_matrix_cuda(geoms, kg);
  register_affine_dof_cuda_kernels();
  std::printf("replaced %d %.17g\n", calls, rr.volume_matrix(0, 1));
  // the DoF methods
  const Vector3d k = {1.3, -0.7, 2.1};
  const VectorXd th = {0.05, -0.02, 0.1, 1.1, 0.1, 0.0, -0.05, 0.9, 0.2, 0.0, 0.1, 1.05};
  const std::complex<double> a1 = aff->compute_A(geoms[1], k, th);
  const std::complex<double> a2 = compute_Ak_cuda(geoms[1], k, aff, th);
  const std::complex<double> a3 = st->compute_A(geoms[1], k, d2);
  const std::complex<double> a4 = compute_Ak_cuda(geoms[1], k, st, d2);
  std::printf("ak %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", a1.real(), a1.imag(),
              a2.real(), a2.imag(), a3.real(), a3.imag(), a4.real(), a4.imag());
  const VectorXcd g3 = st->compute_A_gradient(geoms[1], k, d2);
  const VectorXcd g4 = compute_Ak_gradient_cuda(geoms[1], k, st, d2);
  std::printf("gk");
  for (int t = 0; t < 4; ++t)
    std::printf(" %.17g %.17g %.17g %.17g", g3[t].real(), g3[t].imag(), g4[t].real(),
                g4[t].imag());
  std::printf("\n");
  auto md = std::make_shared<TriangleMeshDoF>(V.rows());
  const VolumeResult va = compute_volume_cpu(geoms[1], aff, th,
                                             ComputationFlags::VOLUME_ONLY |
                                                 ComputationFlags::GRADIENT);
  const VolumeResult vm = compute_volume_cpu(geoms[1], md, V2.data,
                                             ComputationFlags::VOLUME_ONLY |
                                                 ComputationFlags::GRADIENT);
  std::printf("vol %.17g %.17g\nvga", va.volume, vm.volume);
  for (double x : va.grad) std::printf(" %.17g", x);
  std::printf("\nvgm");
  for (double x : vm.grad) std::printf(" %.17g", x);
  std::printf("\n");
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s lebedev_file [fp64|fast]\n", argv[0]);
    return 2;
  }
  if (argc > 2 && std::strcmp(argv[2], "fp64") == 0) set_precision(PII_FF_FP64, PII_GRAM_FP64);
  LebedevGrid L = load_lebedev_txt(argv[1]);
  KGrid kg = build_kgrid(L.dirs, L.weights, 32);
  MatrixXd V;
  MatrixXi F;
  create_unit_cube_mesh(V, F);
  MatrixXd V2 = V;
  const double shift[3] = {0.3, 0.2, 0.1};
  for (int i = 0; i < V2.rows(); ++i)
    for (int c = 0; c < 3; ++c) V2(i, c) += shift[c];
  std::vector<Geometry> geoms = {make_triangle_mesh(V, F), make_triangle_mesh(V2, F)};
  if (argc > 2 && std::strcmp(argv[2], "mismatch") == 0) {
    auto a = std::make_shared<AffineDoF>();
    const VectorXd id = {0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1};
    auto r1 = compute_intersection_volume_matrix_cuda(geoms, {a}, {id, id}, kg);
    auto r2 = compute_intersection_volume_matrix_cuda(geoms, {a, a}, {id}, kg,
                                                      ComputationFlags::GRADIENT);
    for (const auto* r : {&r1, &r2})
      std::printf("mismatch %d %d %.17g %.17g %.17g %.17g %d\n", (int)r->volume_matrix.rows(),
                  (int)r->volume_matrix.cols(), r->volume_matrix(0, 0), r->volume_matrix(0, 1),
                  r->volume_matrix(1, 0), r->volume_matrix(1, 1), (int)r->grad_matrix.size());
    return 0;
  }
  if (argc > 2 && std::strcmp(argv[2], "registry") == 0) {
    set_precision(PII_FF_FP64, PII_GRAM_FP64);
    return registry_mode(geoms, kg, V, V2);
  }
  auto res = compute_intersection_volume_matrix_cuda(geoms, kg);
  if (res.num_objects() != 2) return 1;
  std::printf("V %.15f %.15f %.15f %.15f\n", res.volume_matrix(0, 0), res.volume_matrix(0, 1),
              res.volume_matrix(1, 0), res.volume_matrix(1, 1));
  auto aff = std::make_shared<AffineDoF>();
  const VectorXd id = {0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1};
  auto pr = compute_intersection_volume_cuda(geoms[0], geoms[1], aff, aff, id, id, kg,
                                             ComputationFlags::VOLUME_ONLY |
                                                 ComputationFlags::GRADIENT);
  std::printf("pair %.15f grad_t2 %.15f %.15f %.15f\n", pr.volume, pr.grad_geom2[0],
              pr.grad_geom2[1], pr.grad_geom2[2]);
  // Gauss-Newton Hessian of the pair: the translation-translation corner of
  // the cross block
  auto ph = compute_intersection_volume_cuda(geoms[0], geoms[1], aff, aff, id, id, kg,
                                             ComputationFlags::VOLUME_ONLY |
                                                 ComputationFlags::HESSIAN);
  if (ph.hessian_cross.rows() != 12) return 1;
  std::printf("hcross_tt");
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) std::printf(" %.15e", ph.hessian_cross(a, b));
  std::printf("\n");
  // the same cross block under vertex DoFs (3 per cube corner)
  auto md = std::make_shared<TriangleMeshDoF>(V.rows());
  auto pm = compute_intersection_volume_cuda(geoms[0], geoms[1], md, md, V.data, V2.data, kg,
                                             ComputationFlags::ALL);
  const int dm = 3 * V.rows();
  if (pm.hessian_cross.rows() != dm || pm.hessian_cross.cols() != dm) return 1;
  if ((int)pm.grad_geom2.size() != dm) return 1;
  std::printf("gmesh2");
  for (int t = 0; t < dm; ++t) std::printf(" %.15e", pm.grad_geom2[t]);
  std::printf("\n");
  std::printf("hcross_mesh");
  for (int a = 0; a < dm; ++a)
    for (int b = 0; b < dm; ++b) std::printf(" %.15e", pm.hessian_cross(a, b));
  std::printf("\n");
  std::printf("self %.15f\n", compute_volume_cpu(geoms[0]));
  // single-k derivatives of the second cube: A(k) and dA(k)/dtheta under a
  // non-trivial affine map
  const Vector3d k = {1.3, -0.7, 2.1};
  const VectorXd th = {0.05, -0.02, 0.1, 1.1, 0.1, 0.0, -0.05, 0.9, 0.2, 0.0, 0.1, 1.05};
  const std::complex<double> ak = compute_Ak_cuda(geoms[1], k, aff, th);
  const VectorXcd gk = compute_Ak_gradient_cuda(geoms[1], k, aff, th);
  if (gk.size() != 12) return 1;
  std::printf("ak %.15e %.15e\n", ak.real(), ak.imag());
  std::printf("gradk");
  for (int t = 0; t < 12; ++t) std::printf(" %.15e %.15e", gk[t].real(), gk[t].imag());
  std::printf("\n");
  return 0;
}
