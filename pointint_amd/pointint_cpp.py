"""Drop-in for the reference's pybind11 module ``pointint_cpp``
(demo/bindings.cpp:16-103), served by the MI355X library.

``import pointint_amd.pointint_cpp as pi`` gives the same names, argument
names and return shapes as the reference module for the device path:

  Geometry, KGrid, make_triangle_mesh, make_point_cloud, create_unit_cube_mesh,
  create_sphere_pointcloud, load_lebedev, build_kgrid,
  compute_intersection_volume_cuda, compute_volume_cpu,
  box_box_intersection_volume, sphere_sphere_intersection_volume,
  translate_points

plus the matrix path the reference sketches for Python
(docs/DESIGN_PROPOSAL.md:767-803): compute_intersection_volume_matrix_cuda,
make_gaussian_splat.  ``compute_intersection_volume_cpu`` is the reference's
TBB path; this package ships no CPU implementation of the path, so it raises.
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
    return core.compute_volume_matrix(list(geoms), kgrid, dofs, dof_kinds, ff="fast", gram=gram)


def compute_intersection_volume_cpu(geom1, geom2, kgrid):
    raise NotImplementedError(
        "compute_intersection_volume_cpu is the reference's CPU/TBB path; pointint_amd serves "
        "the device path only (use compute_intersection_volume_cuda)")


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
