// pointint/pointint.hpp -- C++ API of the MI355X k-space intersection-volume
// library, with the reference's names and argument meanings, implemented on
// the C ABI of pointint_hip.h (libpointint_cxx.so -> libpointint_hip.so).
//
// Mirrors (paths relative to the reference tree):
//   geometry types / factories   cpp/include/geometry/{types,geometry,packing}.hpp
//   quadrature                   cpp/include/quadrature/{gauss_legendre,lebedev_io,kgrid}.hpp
//   flags                        cpp/include/computation_flags.hpp
//   DoF parameterisations        cpp/include/dof/{affine_dof,triangle_mesh_dof}.hpp
//   matrix driver                cpp/include/compute_intersection_volume_multi_object.hpp:51-120
//   pairwise driver + single-k   cpp/include/compute_intersection_volume.hpp:48-135
//   self volume                  cpp/include/compute_volume.hpp
//   analytic volumes             cpp/include/analytical_intersection.hpp
//
// Differences from the reference headers, all forced by the toolchain:
//   * Eigen is not available, so Eigen::VectorXd / MatrixXd / Vector3d are
//     replaced by VectorXd (std::vector<double>), a row-major MatrixXd and
//     Vector3d (std::array<double,3>) with the accessors the drivers use.
//   * double3 comes from HIP's vector types; <cuda_runtime.h> is never used.
//   * The "_cuda" names are kept for drop-in compatibility; they run on HIP.
//   * HESSIAN (Gauss-Newton blocks) is computed for AffineDoF and
//     TriangleMeshDoF objects up to 16384 DoFs per call in all; above that
//     the error goes to std::cerr and the blocks stay empty, as the reference
//     reports errors.
#pragma once

#include <hip/hip_vector_types.h>

#include <array>
#include <complex>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace PoIntInt {

// ---- minimal dense types standing in for the Eigen types of the reference ----
using VectorXd = std::vector<double>;
using VectorXcd = std::vector<std::complex<double>>;
using Vector3d = std::array<double, 3>;

struct MatrixXd {
  int nrows = 0, ncols = 0;
  std::vector<double> data;  // row-major
  MatrixXd() = default;
  MatrixXd(int r, int c) : nrows(r), ncols(c), data((size_t)r * c, 0.0) {}
  static MatrixXd Zero(int r, int c) { return MatrixXd(r, c); }
  int rows() const { return nrows; }
  int cols() const { return ncols; }
  double& operator()(int i, int j) { return data[(size_t)i * ncols + j]; }
  double operator()(int i, int j) const { return data[(size_t)i * ncols + j]; }
};

struct MatrixXi {
  int nrows = 0, ncols = 0;
  std::vector<int> data;  // row-major
  MatrixXi() = default;
  MatrixXi(int r, int c) : nrows(r), ncols(c), data((size_t)r * c, 0) {}
  int rows() const { return nrows; }
  int cols() const { return ncols; }
  int& operator()(int i, int j) { return data[(size_t)i * ncols + j]; }
  int operator()(int i, int j) const { return data[(size_t)i * ncols + j]; }
};

// ---- geometry (cpp/include/geometry/types.hpp:9-44, geometry.hpp:8-38) ----
struct TriPacked {
  double3 a, e1, e2, S;
  int vid0, vid1, vid2;
};
struct DiskPacked {
  double3 c, n;
  double rho, area;
};
struct GaussianPacked {
  double3 c, n;
  double sigma, w;
};
enum GeometryType { GEOM_TRIANGLE = 0, GEOM_DISK = 1, GEOM_GAUSSIAN = 2 };

struct Geometry {
  GeometryType type = GEOM_TRIANGLE;
  std::vector<TriPacked> tris;
  std::vector<DiskPacked> disks;
  std::vector<GaussianPacked> gaussians;
  Geometry() = default;
  explicit Geometry(GeometryType t) : type(t) {}
  int num_elements() const;
  bool is_empty() const { return num_elements() == 0; }
  void clear() { tris.clear(); disks.clear(); gaussians.clear(); }
};

std::vector<TriPacked> pack_tris(const MatrixXd& V, const MatrixXi& F);
std::vector<DiskPacked> pack_disks(const MatrixXd& P, const MatrixXd& N,
                                   const VectorXd& radii_or_areas, bool is_radius = true);
std::vector<GaussianPacked> pack_gaussians(const MatrixXd& P, const MatrixXd& N,
                                           const VectorXd& sigmas, const VectorXd& weights);
Geometry make_triangle_mesh(const MatrixXd& V, const MatrixXi& F);
Geometry make_point_cloud(const MatrixXd& P, const MatrixXd& N, const VectorXd& radii_or_areas,
                          bool is_radius = true);
Geometry make_gaussian_splat(const MatrixXd& P, const MatrixXd& N, const VectorXd& sigmas,
                             const VectorXd& weights);
void create_unit_cube_mesh(MatrixXd& V, MatrixXi& F);
void create_sphere_pointcloud(MatrixXd& P, MatrixXd& N, VectorXd& radii, int n_points = 1000);

// ---- quadrature ----
std::pair<std::vector<double>, std::vector<double>> gauss_legendre_interval(int n, double a,
                                                                            double b);
struct LebedevGrid {
  std::vector<std::array<double, 3>> dirs;
  std::vector<double> weights;
};
LebedevGrid load_lebedev_txt(const std::string& path);
struct KGrid {
  std::vector<std::array<double, 3>> dirs;
  std::vector<double> kmag;
  std::vector<double> w;
};
KGrid build_kgrid(const std::vector<std::array<double, 3>>& leb_dirs,
                  const std::vector<double>& leb_w, int Nrad);

// ---- flags (cpp/include/computation_flags.hpp:12-40) ----
enum class ComputationFlags { VOLUME_ONLY = 0x01, GRADIENT = 0x02, HESSIAN = 0x04, ALL = 0x07 };
inline ComputationFlags operator|(ComputationFlags a, ComputationFlags b) {
  return static_cast<ComputationFlags>(static_cast<int>(a) | static_cast<int>(b));
}
inline ComputationFlags operator&(ComputationFlags a, ComputationFlags b) {
  return static_cast<ComputationFlags>(static_cast<int>(a) & static_cast<int>(b));
}
inline bool needs_volume(ComputationFlags f) { return (static_cast<int>(f) & 1) != 0; }
inline bool needs_gradient(ComputationFlags f) { return (static_cast<int>(f) & 2) != 0; }
inline bool needs_hessian(ComputationFlags f) { return (static_cast<int>(f) & 4) != 0; }

// ---- DoF parameterisations ----
struct DoFParameterization {
  virtual ~DoFParameterization() = default;
  virtual int num_dofs() const = 0;
  virtual int dof_kind() const = 0;  // PII_DOF_AFFINE or PII_DOF_TRIMESH
};
// [tx, ty, tz, A00, A01, A02, A10, A11, A12, A20, A21, A22]
struct AffineDoF : DoFParameterization {
  int num_dofs() const override { return 12; }
  int dof_kind() const override { return 0; }
  // |det A| x the divergence-theorem volume of the reference geometry
  // (cpp/src/dof/affine_dof.cpp:324-384), on the host.
  double compute_volume(const Geometry& geom, const VectorXd& dofs) const;
};
// [v0x, v0y, v0z, v1x, ...]
struct TriangleMeshDoF : DoFParameterization {
  explicit TriangleMeshDoF(int num_vertices) : nv_(num_vertices) {}
  int num_dofs() const override { return 3 * nv_; }
  int dof_kind() const override { return 1; }

 private:
  int nv_;
};

// ---- results ----
struct IntersectionVolumeMatrixResult {
  MatrixXd volume_matrix;
  // grad_matrix[i][j] = dV_ij / dtheta_i; for i == j half the full derivative
  // (cpp/include/compute_intersection_volume_multi_object.hpp:22-26).
  std::vector<std::vector<VectorXd>> grad_matrix;
  // Gauss-Newton blocks (AffineDoF): hessian_ij[i][j] = (1/8 pi^3) sum_q w_q
  // Re(dA_i/dtheta_i (dA_j/dtheta_j)^H); hessian_ii / hessian_jj are zero.
  std::vector<std::vector<MatrixXd>> hessian_ii, hessian_jj, hessian_ij;
  int num_objects() const { return volume_matrix.rows(); }
};

struct IntersectionVolumeResult {
  double volume = 0.0;
  VectorXd grad_geom1, grad_geom2;
  MatrixXd hessian_geom1, hessian_geom2, hessian_cross;  // Gauss-Newton (AffineDoF)
};

// ---- drivers ----
IntersectionVolumeMatrixResult compute_intersection_volume_matrix_cuda(
    const std::vector<Geometry>& ref_geometries,
    const std::vector<std::shared_ptr<DoFParameterization>>& dofs,
    const std::vector<VectorXd>& dof_vectors, const KGrid& kgrid,
    ComputationFlags flags = ComputationFlags::VOLUME_ONLY, int blockSize = 256,
    bool enable_profiling = false);

IntersectionVolumeMatrixResult compute_intersection_volume_matrix_cuda(
    const std::vector<Geometry>& geometries, const KGrid& kgrid, int blockSize = 256,
    bool enable_profiling = false);

IntersectionVolumeResult compute_intersection_volume_cuda(
    const Geometry& ref_geom1, const Geometry& ref_geom2,
    const std::shared_ptr<DoFParameterization>& dof1,
    const std::shared_ptr<DoFParameterization>& dof2, const VectorXd& dofs1,
    const VectorXd& dofs2, const KGrid& kgrid,
    ComputationFlags flags = ComputationFlags::VOLUME_ONLY, int blockSize = 256,
    bool enable_profiling = false);

double compute_intersection_volume_cuda(const Geometry& geom1, const Geometry& geom2,
                                        const KGrid& kgrid, int blockSize = 256,
                                        bool enable_profiling = false);

// A(k) at one k-vector (cpp/include/compute_intersection_volume.hpp:120-126).
std::complex<double> compute_Ak_cuda(const Geometry& geom, const Vector3d& k,
                                     const std::shared_ptr<DoFParameterization>& dof,
                                     const VectorXd& dofs, int blockSize = 256);

// dA(k)/dtheta at one k-vector (cpp/include/compute_intersection_volume.hpp:128-135):
// num_dofs complex values; zeros for |k| < 1e-10, as the reference.
VectorXcd compute_Ak_gradient_cuda(const Geometry& geom, const Vector3d& k,
                                   const std::shared_ptr<DoFParameterization>& dof,
                                   const VectorXd& dofs, int blockSize = 256);

// Divergence-theorem self volume with identity DoFs (compute_volume.hpp).
double compute_volume_cpu(const Geometry& geom, bool enable_profiling = false);

// cpp/include/analytical_intersection.hpp
double box_box_intersection_volume(const Vector3d& c1, const Vector3d& h1, const Vector3d& c2,
                                   const Vector3d& h2);
double sphere_sphere_intersection_volume(const Vector3d& c1, double r1, const Vector3d& c2,
                                         double r2);

// Form-factor precision / Gram mode of the matrix driver (defaults: fast
// form factors, FP16x3 Gram).  Not in the reference, which is FP64 only.
void set_precision(int ff_precision, int gram_mode);

}  // namespace PoIntInt
