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
//   usage: example_cube_pair <lebedev_029.txt> [fp64|fast|mismatch]
#include <cstdio>
#include <cstring>

#include "pointint/pointint.hpp"
#include "pointint_hip.h"

using namespace PoIntInt;

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
