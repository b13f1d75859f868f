"""Seeded synthetic workloads of the k-space intersection-volume path.

Generators (all numpy, PCG64 via numpy.random.default_rng(seed)):

* unit cube, 12 outward triangles (cpp/src/geometry/geometry_helpers.cpp:16-38)
* Fibonacci sphere disk cloud (cpp/src/geometry/geometry_helpers.cpp:40-81)
* icospheres (subdivided icosahedron projected to the unit sphere)
* star-shaped perturbed spheres with a target triangle count (UV tessellation)
* the five benchmark configurations: cube pair, batched icospheres, uneven
  mixed primitives, differentiation set, serving pool.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np


# ---------------------------------------------------------------------------
# Base shapes
# ---------------------------------------------------------------------------
def unit_cube():
    V = np.array([[-0.5, -0.5, -0.5], [0.5, -0.5, -0.5], [0.5, 0.5, -0.5], [-0.5, 0.5, -0.5],
                  [-0.5, -0.5, 0.5], [0.5, -0.5, 0.5], [0.5, 0.5, 0.5], [-0.5, 0.5, 0.5]],
                 dtype=np.float64)
    F = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4],
                  [2, 7, 6], [2, 3, 7], [0, 4, 7], [0, 7, 3], [1, 6, 5], [1, 2, 6]],
                 dtype=np.int32)
    return V, F


def sphere_pointcloud(n: int = 1000):
    """Unit-sphere Fibonacci disks: positions, normals, radii sqrt(4/n)."""
    i = np.arange(n, dtype=np.float64)
    ga = math.pi * (3.0 - math.sqrt(5.0))
    y = 1.0 - (2.0 * i) / (n - 1.0)
    ry = np.sqrt(np.maximum(0.0, 1.0 - y * y))
    th = ga * i
    P = np.stack([ry * np.cos(th), y, ry * np.sin(th)], axis=1)
    r = np.full(n, math.sqrt((4.0 * math.pi / n) / math.pi))
    return P, P.copy(), r


def icosphere(level: int):
    """Unit icosphere: 10*4^L + 2 vertices, 20*4^L outward faces."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    V = [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t],
         [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]]
    F = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4],
         [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8],
         [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]]
    V = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in V]
    for _ in range(level):
        cache = {}
        newF = []

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = V[a] + V[b]
                V.append(m / np.linalg.norm(m))
                cache[key] = len(V) - 1
            return cache[key]

        for a, b, c in F:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            newF += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        F = newF
    V = np.array(V)
    F = np.array(F, dtype=np.int32)
    # orient outward
    cen = V[F].mean(axis=1)
    nrm = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    flip = (nrm * cen).sum(axis=1) < 0
    F[flip] = F[flip][:, [0, 2, 1]]
    return V, F


def uv_sphere(n_tri_target: int):
    """UV tessellated unit sphere with about n_tri_target outward triangles."""
    nlat = max(3, int(round(math.sqrt(n_tri_target / 4.0))) + 1)
    nlon = max(3, int(round(n_tri_target / (2.0 * (nlat - 1)))))
    th = math.pi * np.arange(1, nlat) / nlat
    ph = 2 * math.pi * np.arange(nlon) / nlon
    ring = np.stack([np.outer(np.sin(th), np.cos(ph)), np.outer(np.sin(th), np.sin(ph)),
                     np.outer(np.cos(th), np.ones(nlon))], axis=-1).reshape(-1, 3)
    V = np.concatenate([[[0.0, 0.0, 1.0]], ring, [[0.0, 0.0, -1.0]]])
    south = V.shape[0] - 1
    j = np.arange(nlon)
    jn = (j + 1) % nlon

    def idx(i, jj):
        return 1 + (i - 1) * nlon + jj

    caps_n = np.stack([np.zeros(nlon, dtype=np.int64), idx(1, j), idx(1, jn)], axis=1)
    i = np.arange(1, nlat - 1)[:, None]
    a, b, c, d = idx(i, j[None, :]), idx(i, jn[None, :]), idx(i + 1, j[None, :]), idx(i + 1, jn[None, :])
    body = np.concatenate([np.stack([a, c, b], axis=-1).reshape(-1, 3),
                           np.stack([b, c, d], axis=-1).reshape(-1, 3)])
    caps_s = np.stack([np.full(nlon, south), idx(nlat - 1, jn), idx(nlat - 1, j)], axis=1)
    F = np.concatenate([caps_n, body, caps_s]).astype(np.int32)
    cen = V[F].mean(axis=1)
    nrm = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    flip = (nrm * cen).sum(axis=1) < 0
    F[flip] = F[flip][:, [0, 2, 1]]
    return V, F


def random_rotation(rng) -> np.ndarray:
    """Uniform rotation from a normalised Gaussian quaternion."""
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def star_perturb(V, rng, amplitude=0.1, n_waves=4):
    """Star-shaped radial noise r(u) = 1 + amplitude * f(u), |f| <= 1, where f
    is a normalised sum of n_waves random plane waves cos(omega u.d + phase)
    with omega in [1, 4] (low angular degree)."""
    dirs = rng.normal(size=(n_waves, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    om = rng.uniform(1.0, 4.0, size=n_waves)
    ph = rng.uniform(0, 2 * math.pi, size=n_waves)
    a = rng.normal(size=n_waves)
    a /= np.abs(a).sum()
    u = V / np.linalg.norm(V, axis=1, keepdims=True)
    f = (a[None, :] * np.cos(om[None, :] * (u @ dirs.T) + ph[None, :])).sum(axis=1)
    return u * (1.0 + amplitude * f)[:, None]


def log_uniform_int(rng, lo, hi, size=None):
    return np.exp(rng.uniform(math.log(lo), math.log(hi), size=size)).round().astype(np.int64)


# ---------------------------------------------------------------------------
# Configurations
# ---------------------------------------------------------------------------
@dataclass
class MeshObj:
    V: np.ndarray
    F: np.ndarray

    @property
    def n(self):
        return self.F.shape[0]


@dataclass
class DiskObj:
    P: np.ndarray
    N: np.ndarray
    r: np.ndarray

    @property
    def n(self):
        return self.P.shape[0]


@dataclass
class Workload:
    name: str
    objects: list
    lebedev_order: int
    n_radial: int
    seed: int
    meta: dict = field(default_factory=dict)

    def prim_count(self) -> int:
        return int(sum(o.n for o in self.objects))


def config_cube_pair() -> Workload:
    """Two unit cubes, the second shifted by (0.3, 0.2, 0.1); Lebedev 302 x L=32."""
    V, F = unit_cube()
    return Workload("cube_pair", [MeshObj(V, F), MeshObj(V + np.array([0.3, 0.2, 0.1]), F)],
                    29, 32, 0, {"analytic_overlap": 0.504})


def config_icospheres(n_obj=256, level=3, seed=2, centre_box=2.0, lebedev_order=59,
                      n_radial=64) -> Workload:
    """N icospheres: radius U[0.2,0.5], axis scales U[0.7,1.3]^3, uniform
    rotation, centres U[0, centre_box]^3; vertices baked into world space."""
    rng = np.random.default_rng(seed)
    V0, F0 = icosphere(level)
    objs = []
    for _ in range(n_obj):
        rad = rng.uniform(0.2, 0.5)
        sc = rng.uniform(0.7, 1.3, size=3)
        R = random_rotation(rng)
        c = rng.uniform(0.0, centre_box, size=3)
        V = (V0 * (rad * sc)[None, :]) @ R.T + c[None, :]
        objs.append(MeshObj(V, F0.copy()))
    return Workload(f"icospheres_L{level}_N{n_obj}", objs, lebedev_order, n_radial, seed)


def config_mixed(n_obj=1024, seed=3, lebedev_order=89, n_radial=96, scale=1.0) -> Workload:
    """Half star-shaped perturbed spheres (T log-uniform in [80, 20480]), half
    Fibonacci disk clouds (log-uniform in [1k, 50k], rho U[0.005, 0.02]);
    radius U[0.2, 0.5]; centres U[0, 4]^3.  `scale` < 1 shrinks element
    counts for tests."""
    rng = np.random.default_rng(seed)
    objs = []
    for i in range(n_obj):
        rad = rng.uniform(0.2, 0.5)
        c = rng.uniform(0.0, 4.0, size=3)
        if i % 2 == 0:
            T = int(log_uniform_int(rng, 80, 20480) * scale)
            V, F = uv_sphere(max(T, 16))
            V = star_perturb(V, rng, 0.1)
            R = random_rotation(rng)
            objs.append(MeshObj(V * rad @ R.T + c, F))
        else:
            nd = max(16, int(log_uniform_int(rng, 1000, 50000) * scale))
            P, Nn, _ = sphere_pointcloud(nd)
            rho = rng.uniform(0.005, 0.02, size=nd)
            objs.append(DiskObj(P * rad + c, Nn, rho))
    return Workload(f"mixed_N{n_obj}", objs, lebedev_order, n_radial, seed)


def config_diff(n_obj=64, level=4, seed=4, radius=0.39, lebedev_order=41, n_radial=64) -> Workload:
    """64 icospheres (level 4), radius `radius`, centres U[0, 1.5]^3."""
    rng = np.random.default_rng(seed)
    V0, F0 = icosphere(level)
    objs = []
    for _ in range(n_obj):
        c = rng.uniform(0.0, 1.5, size=3)
        objs.append(MeshObj(V0 * radius + c, F0.copy()))
    return Workload(f"diff_L{level}_N{n_obj}", objs, lebedev_order, n_radial, seed)


def serving_pool(n_mesh=4096, seed=5, scale=1.0):
    """Pool of star-shaped meshes (T log-uniform in [80, 20480]) centred at the
    origin with radius U[0.2, 0.5]."""
    rng = np.random.default_rng(seed)
    pool = []
    for _ in range(n_mesh):
        rad = rng.uniform(0.2, 0.5)
        T = int(log_uniform_int(rng, 80, 20480) * scale)
        V, F = uv_sphere(max(T, 16))
        V = star_perturb(V, rng, 0.1) * rad
        pool.append(MeshObj(V, F))
    return pool
