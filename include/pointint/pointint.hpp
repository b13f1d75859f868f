// pointint/pointint.hpp -- C++ API of the MI355X k-space intersection-volume
// library, with the reference's names and argument meanings, implemented on
// the C ABI of pointint_hip.h (libpointint_cxx.so -> libpointint_hip.so).
//
// Mirrors (paths relative to the reference tree):
//   geometry types / factories   cpp/include/geometry/{types,geometry,packing}.hpp
//   quadrature                   cpp/include/quadrature/{gauss_legendre,lebedev_io,kgrid}.hpp
//   flags                        cpp/include/computation_flags.hpp
//   DoF parameterisations        cpp/include/dof/{dof_parameterization,affine_dof,triangle_mesh_dof}.hpp
//   kernel registry              cpp/include/dof/cuda/dof_cuda_interface.hpp
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
#include <functional>
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

// ---- DoF parameterisations (cpp/include/dof/dof_parameterization.hpp:10-38) ----
struct DoFParameterization {
  virtual ~DoFParameterization() = default;
  virtual int num_dofs() const = 0;
  // PII_DOF_AFFINE (0) or PII_DOF_TRIMESH (1) for the built-in kinds, which
  // the drivers evaluate with this library's kernels; -1 for a caller's own
  // parameterisation, which the drivers reach through CudaKernelRegistry.
  virtual int dof_kind() const { return -1; }
  // A(k) and dA(k)/dtheta (num_dofs complex values) at one k-vector; the
  // built-in kinds evaluate them on the device in FP64.
  virtual std::complex<double> compute_A(const Geometry& geom, const Vector3d& k,
                                         const VectorXd& dofs) const = 0;
  virtual VectorXcd compute_A_gradient(const Geometry& geom, const Vector3d& k,
                                       const VectorXd& dofs) const = 0;
  // Divergence-theorem volume V = (1/3) sum_S x.n dS of the posed geometry
  // and dV/dtheta (host arithmetic).
  virtual double compute_volume(const Geometry& geom, const VectorXd& dofs) const = 0;
  virtual VectorXd compute_volume_gradient(const Geometry& geom, const VectorXd& dofs) const = 0;
};
// [tx, ty, tz, A00, A01, A02, A10, A11, A12, A20, A21, A22]
// (cpp/include/dof/affine_dof.hpp, cpp/src/dof/affine_dof.cpp)
struct AffineDoF : DoFParameterization {
  int num_dofs() const override { return 12; }
  int dof_kind() const override { return 0; }
  std::complex<double> compute_A(const Geometry& geom, const Vector3d& k,
                                 const VectorXd& dofs) const override;
  VectorXcd compute_A_gradient(const Geometry& geom, const Vector3d& k,
                               const VectorXd& dofs) const override;
  // |det A| x the volume of the reference geometry (affine_dof.cpp:324-384)
  double compute_volume(const Geometry& geom, const VectorXd& dofs) const override;
  // 0 for t; sign(det A) cof(A) V_ref for A (affine_dof.cpp:386-460)
  VectorXd compute_volume_gradient(const Geometry& geom, const VectorXd& dofs) const override;
};
// [v0x, v0y, v0z, v1x, ...]
// (cpp/include/dof/triangle_mesh_dof.hpp, cpp/src/dof/triangle_mesh_dof.cpp)
struct TriangleMeshDoF : DoFParameterization {
  explicit TriangleMeshDoF(int num_vertices) : nv_(num_vertices) {}
  int num_dofs() const override { return 3 * nv_; }
  int dof_kind() const override { return 1; }
  std::complex<double> compute_A(const Geometry& geom, const Vector3d& k,
                                 const VectorXd& dofs) const override;
  VectorXcd compute_A_gradient(const Geometry& geom, const Vector3d& k,
                               const VectorXd& dofs) const override;
  // (1/6) sum_f v_i . (v_j x v_k) over the faces' vertex ids, faces with a
  // negative id skipped (triangle_mesh_dof.cpp:145-179)
  double compute_volume(const Geometry& geom, const VectorXd& dofs) const override;
  // (1/6) sum_f of v_j x v_k into vertex i, cyclically (triangle_mesh_dof.cpp:181-222)
  VectorXd compute_volume_gradient(const Geometry& geom, const VectorXd& dofs) const override;

 private:
  int nv_;
};

// ---- kernel registry (cpp/include/dof/cuda/dof_cuda_interface.hpp:19-67) ----
// A form-factor plug-in writes A(k_q) for the Q nodes of kgrid (double2, HIP
// device memory) and its gradient plug-in the Q x num_dofs row-major
// dA/dtheta.  The drivers look every object up by
// make_key(get_dof_type_name(dof), geometry type): the built-in entries
// ("AffineDoF_{Triangle,Disk,Gaussian}", "TriangleMeshDoF_Triangle",
// registered when the library loads) run this library's fused kernels; an
// entry a caller registers or replaces runs through its own functions
// (columns gathered on the host, one FP64 device Gram for V, the gradients
// and the Gauss-Newton blocks).  blockSize is passed through and ignored by
// the built-in entries.
using CudaComputeAkFunc = std::function<void(const Geometry& geom, const KGrid& kgrid,
                                             const VectorXd& dofs, double2* d_A_out,
                                             int blockSize)>;
using CudaComputeAkGradientFunc = std::function<void(const Geometry& geom, const KGrid& kgrid,
                                                     const VectorXd& dofs, double2* d_grad_A_out,
                                                     int blockSize)>;
class CudaKernelRegistry {
 public:
  static void register_kernels(const std::string& dof_type, GeometryType geom_type,
                               CudaComputeAkFunc compute_Ak,
                               CudaComputeAkGradientFunc compute_Ak_gradient);
  static bool has_kernels(const std::string& dof_type, GeometryType geom_type);
  static std::pair<CudaComputeAkFunc, CudaComputeAkGradientFunc> get_kernels(
      const std::string& dof_type, GeometryType geom_type);
  static std::string make_key(const std::string& dof_type, GeometryType geom_type);
};
// "AffineDoF", "TriangleMeshDoF" or "Unknown" (dof_cuda_interface.cpp:57-69).
std::string get_dof_type_name(const std::shared_ptr<DoFParameterization>& dof);
// (Re-)register the built-in entries (affine_dof_cuda.cu:1323-1358,
// triangle_mesh_dof_cuda.cu:460-478).
void register_affine_dof_cuda_kernels();
void register_triangle_mesh_dof_cuda_kernels();

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

// cpp/include/compute_volume.hpp:14-21
struct VolumeResult {
  double volume = 0.0;
  VectorXd grad;  // dV/dtheta (GRADIENT)
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

// Divergence-theorem self volume (cpp/include/compute_volume.hpp:31-44):
// dof->compute_volume / compute_volume_gradient by the flags, and the
// identity-AffineDoF overload.
VolumeResult compute_volume_cpu(const Geometry& ref_geom,
                                const std::shared_ptr<DoFParameterization>& dof,
                                const VectorXd& dofs,
                                ComputationFlags flags = ComputationFlags::VOLUME_ONLY,
                                bool enable_profiling = false);
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
