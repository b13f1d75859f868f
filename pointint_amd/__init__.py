"""pointint_amd: MI355X-native k-space intersection volumes.

Per-object form factors A_i(k) over a Gauss-Legendre x Lebedev k-grid, the
all-pairs volume matrix V = Re(J^H D J) / (8 pi^3), its adjoint and a serving
loop, on hand-written gfx950 HIP kernels behind the C ABI in
include/pointint_hip.h.
"""
from .core import (  # noqa: F401
    IDENTITY_AFFINE,
    Context,
    Geometry,
    KGrid,
    build_kgrid,
    build_kgrid_arrays,
    cache_moments,
    compute_form_factor,
    compute_form_factor_into,
    compute_form_factor_jacobian,
    compute_volume_matrix,
    compute_volume_matrix_device,
    gauss_legendre,
    gram_from_J,
    kgrid_from_rule,
    lebedev_path,
    load_lebedev,
    make_gaussian_splat,
    make_packed_triangles,
    make_point_cloud,
    make_triangle_mesh,
    volume_matrix_vjp,
    volume_matrix_vjp_device,
    volume_matrix_hessian,
    volume_matrix_derivatives,
)
