"""Drop-in for the reference's pybind11 module ``pointint_cpp``
(demo/bindings.cpp:16-103), served by the MI355X library.

``import pointint_amd.pointint_cpp as pi`` gives the same names, argument
names and return shapes as the reference module:

  Geometry, KGrid, make_triangle_mesh, make_point_cloud, create_unit_cube_mesh,
  create_sphere_pointcloud, load_lebedev, build_kgrid,
  compute_intersection_volume_cuda, compute_intersection_volume_cpu,
  compute_volume_cpu, box_box_intersection_volume,
  sphere_sphere_intersection_volume, translate_points

so that demo/exp.py's call sequence runs unchanged, plus the API the
reference sketches for Python (docs/DESIGN_PROPOSAL.md:767-793): the matrix
path compute_intersection_volume_matrix_cuda, the DoF parameterisations
AffineDoF / TriangleMeshDoF and compute_intersection_volume_gradient_cuda,
and make_gaussian_splat.

There is no CPU implementation of the path in this package:
``compute_intersection_volume_cpu`` (the reference's FP64 TBB path) returns
the same FP64 value from the plain FP64 path on the device, and
``compute_intersection_volume_cuda`` the fast path (FP32 transcendentals,
FP64 phases and moment series, FP64 Gram of the two columns).
"""
from __future__ import annotations

import math

import numpy as np

from . import core
from . import workloads as _W
from .core import Geometry, KGrid  # noqa: F401

make_triangle_mesh = core.make_triangle_mesh
make_gaussian_splat = core.make_gaussian_splat


def make_point_cloud(P, N, radii, is_radius=True):
    return core.make_point_cloud(P, N, radii, is_radius)


def create_unit_cube_mesh():
    """(V (8,3) float64, F (12,3) int32): cpp/src/geometry/geometry_helpers.cpp:16-38."""
    return _W.unit_cube()


def create_sphere_pointcloud(n=1000):
    """(P, N, radii) Fibonacci unit-sphere disks: geometry_helpers.cpp:40-81."""
    return _W.sphere_pointcloud(n)


def load_lebedev(path):
    """(dirs, weights): weights summing to 4 pi (lebedev_io.hpp:24-59)."""
    return core.load_lebedev(path)


def build_kgrid(leb_dirs, leb_weights, n_radial):
    return core.build_kgrid(leb_dirs, leb_weights, n_radial)


def compute_intersection_volume_cuda(geom1, geom2, kgrid, block_size=256):
    """Pairwise V of two geometries (identity DoFs), on the GPU.  block_size is
    accepted for signature compatibility; the launch shapes are fixed."""
    V = core.compute_volume_matrix([geom1, geom2], kgrid, ff="fast", gram="fp64")
    return float(V[0, 1])


def compute_intersection_volume_matrix_cuda(geoms, kgrid, block_size=256, dofs=None,
                                            dof_kinds=None, gram="fp16x3"):
    """The n x n volume matrix (docs/DESIGN_PROPOSAL.md:782-784); dofs /
    dof_kinds optional (identity AffineDoF)."""
    return core.compute_volume_matrix(list(geoms), kgrid, dofs, dof_kinds, ff="fast", gram=gram)


def compute_intersection_volume_cpu(geom1, geom2, kgrid):
    """Pairwise V of two geometries (identity DoFs) in FP64, the value of the
    reference's CPU path (cpp/src/compute_intersection_volume_cpu.cpp:21-208):
    computed by the plain FP64 path on the device (FP64 form factors over the
    full grid, FP64 Gram)."""
    V = core.compute_volume_matrix([geom1, geom2], kgrid, ff="fp64", gram="fp64")
    return float(V[0, 1])


class AffineDoF:
    """12 DoFs [t (3), A row-major (9)]: x' = A x + t
    (cpp/include/dof/affine_dof.hpp:15-22)."""
    kind = "affine"

    def num_dofs(self, geom=None) -> int:
        return 12

    @staticmethod
    def identity() -> np.ndarray:
        return np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1], dtype=np.float64)


class TriangleMeshDoF:
    """3 DoFs per vertex: the vertex positions themselves
    (cpp/include/dof/triangle_mesh_dof.hpp:12-17)."""
    kind = "trimesh"

    def __init__(self, num_vertices: int = 0):
        self.num_vertices = num_vertices

    def num_dofs(self, geom=None) -> int:
        return 3 * (geom.num_dof_vertices if geom is not None else self.num_vertices)


class IntersectionVolumeResult:
    """volume and dV/dtheta of each geometry
    (cpp/include/compute_intersection_volume.hpp:18-34)."""

    def __init__(self, volume, grad_geom1, grad_geom2):
        self.volume = volume
        self.grad_geom1 = grad_geom1
        self.grad_geom2 = grad_geom2

    def __repr__(self):
        return (f"IntersectionVolumeResult(volume={self.volume}, "
                f"grad_geom1={self.grad_geom1.shape}, grad_geom2={self.grad_geom2.shape})")


def compute_intersection_volume_gradient_cuda(geom1, geom2, dof1, dof2, dofs1, dofs2, kgrid,
                                              block_size=256, gram="fp64"):
    """V_12 of the two geometries under their DoFs and its gradient with
    respect to each one's DoFs (docs/DESIGN_PROPOSAL.md:786-793; the GRADIENT
    flag of compute_intersection_volume_cuda, cpp/src/compute_intersection_volume.cu:60-101).
    One pass: form factors and their DoF derivatives, one Gram."""
    kinds = [dof1.kind, dof2.kind]
    dofs = [np.asarray(dofs1, dtype=np.float64).ravel(),
            np.asarray(dofs2, dtype=np.float64).ravel()]
    V, G = core.volume_matrix_derivatives([geom1, geom2], kgrid, dofs, gram=gram, hessian=False,
                                          dof_kinds=kinds)
    # G[i] is (2, D_i): row j = dV_ij / dtheta_i
    return IntersectionVolumeResult(float(V[0, 1]), np.array(G[0][1]), np.array(G[1][0]))


def compute_volume_cpu(geom):
    """Divergence-theorem self volume V = (1/3) sum x.n dA of the reference
    geometry (cpp/src/dof/affine_dof.cpp:324-384 with identity DoFs)."""
    h = geom.host
    if geom.type == 0:
        if "V" in h:
            V, F = h["V"], h["F"]
            a, b, c = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
            return float(np.einsum("ij,ij->i", a, np.cross(b - a, c - a)).sum() / 6.0)
        tri = h["tri"]
        return float(np.einsum("ij,ij->i", tri[:, 0:3], tri[:, 9:12]).sum() / 3.0)
    el = h["el"]
    return float((np.einsum("ij,ij->i", el[:, 0:3], el[:, 3:6]) * el[:, 7]).sum() / 3.0)


def box_box_intersection_volume(c1, h1, c2, h2):
    """Axis-aligned boxes (centres, half extents): analytical_intersection.cpp:15-30."""
    c1, h1, c2, h2 = (np.asarray(x, dtype=np.float64) for x in (c1, h1, c2, h2))
    lo = np.maximum(c1 - h1, c2 - h2)
    hi = np.minimum(c1 + h1, c2 + h2)
    return float(np.prod(np.maximum(0.0, hi - lo)))


def sphere_sphere_intersection_volume(c1, r1, c2, r2):
    """Two balls: analytical_intersection.cpp:32-60."""
    d = float(np.linalg.norm(np.asarray(c2, dtype=np.float64) - np.asarray(c1, dtype=np.float64)))
    if d >= r1 + r2:
        return 0.0
    if d <= abs(r1 - r2):
        rm = min(r1, r2)
        return 4.0 * math.pi / 3.0 * rm ** 3
    t1 = (r1 + r2 - d) ** 2
    t2 = d * d + 2.0 * d * (r1 + r2) - 3.0 * (r1 * r1 + r2 * r2) + 6.0 * r1 * r2
    return math.pi * t1 * t2 / (12.0 * d)


def translate_points(points, translation):
    return np.asarray(points, dtype=np.float64) + np.asarray(translation, dtype=np.float64)[None, :]
