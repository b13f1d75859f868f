"""Spherical-harmonic moment cache (pii_geometry_cache_moments).

After caching, a triangle mesh posed by a similarity transform x' = s R x + t
takes its low-|k| moment series from the harmonic expansion instead of one
pass over its faces per direction.  Pinned here:

* a subdivided cube (432 faces) under a rotation, scale and shift against
  the closed form of its form factor at every node of every shell;
* cached against uncached form factors of star meshes on the series shells
  (same series, another summation: within 1e-8 of the shell's largest |A|),
  and identical on the other nodes;
* the oracle on the series shells with |k| >= 1;
* bit-identical results where the cache does not apply: affine DoFs that are
  not a similarity, vertex DoFs, meshes under 192 faces, disk clouds;
* volume matrices and served queries with cached meshes against uncached."""
import numpy as np
import pytest

import pointint_amd as P
from pointint_amd import serving as S
from pointint_amd import workloads as W

pytestmark = pytest.mark.gpu

# serving grid: Lebedev 1202 x 64 shells
RULE, NRAD = 59, 64


def subdivided_cube(n):
    """Unit cube centred at the origin, each face an n x n grid of squares
    split into two outward triangles (faces do not share vertices)."""
    g = np.linspace(-0.5, 0.5, n + 1)
    X, Y = np.meshgrid(g, g, indexing="ij")
    top = np.stack([X.ravel(), Y.ravel(), np.full(X.size, 0.5)], axis=1)
    idx = lambda i, j: i * (n + 1) + j
    tris = []
    for i in range(n):
        for j in range(n):
            tris.append([idx(i, j), idx(i + 1, j), idx(i + 1, j + 1)])
            tris.append([idx(i, j), idx(i + 1, j + 1), idx(i, j + 1)])
    tris = np.array(tris)
    rots = [np.eye(3), np.diag([1.0, -1.0, -1.0]),
            np.array([[0, 0, 1], [0, 1, 0], [-1, 0, 0]], dtype=float),
            np.array([[0, 0, -1], [0, 1, 0], [1, 0, 0]], dtype=float),
            np.array([[1, 0, 0], [0, 0, 1], [0, -1, 0]], dtype=float),
            np.array([[1, 0, 0], [0, 0, -1], [0, 1, 0]], dtype=float)]
    V = np.concatenate([top @ R.T for R in rots])
    F = np.concatenate([tris + k * len(top) for k in range(len(rots))])
    return V, F.astype(np.int32)


def affine(R, s, t):
    return np.concatenate([t, (s * R).reshape(-1)])


def shell_rel(A, B, M):
    """Per shell: max |A - B| / max |B|."""
    d = np.abs(A - B).reshape(-1, M).max(axis=1)
    return d / np.abs(B).reshape(-1, M).max(axis=1)


@pytest.fixture(scope="module")
def kg():
    return P.kgrid_from_rule(RULE, NRAD)


@pytest.fixture(scope="module")
def star():
    rng = np.random.default_rng(8)
    V, F = W.uv_sphere(1500)
    return W.star_perturb(V, rng, 0.1) * 0.4, F


def series_shells(kg, R):
    """Shells whose |k| R <= 15 (the moment series' reach)."""
    M = kg.Q // NRAD
    ks = kg.kmag.reshape(NRAD, M)[:, 0]
    return M, ks, int((ks * R <= 15.0).sum())


def test_subdivided_cube_closed_form(kg):
    V, F = subdivided_cube(6)
    assert len(F) == 432
    g = P.make_triangle_mesh(V, F)
    P.cache_moments([g])
    rng = np.random.default_rng(12)
    R = W.random_rotation(rng)
    s, t = 0.7, np.array([0.3, -0.2, 0.5])
    A = P.compute_form_factor(g, kg, affine(R, s, t), "affine", precision="fast")
    k = kg.dirs * kg.kmag[:, None]
    kr = s * (k @ R)
    sinc = lambda x: np.sinc(x / np.pi)
    ex = 1j * kg.kmag * s ** 3 * np.exp(1j * (k @ t)) * (
        sinc(kr[:, 0] / 2) * sinc(kr[:, 1] / 2) * sinc(kr[:, 2] / 2))
    M, ks, nser = series_shells(kg, s * np.sqrt(3) / 2)
    err = shell_rel(A, ex, M)
    print(f"series shells {nser}; per node / shell max: series {err[:nser].max():.2e}, "
          f"all {err.max():.2e}")
    assert nser >= 8
    assert err[:nser].max() < 1e-8 and err.max() < 1e-6


@pytest.mark.parametrize("scale", [1.0, 1.7])
def test_cached_equals_direct_series(kg, star, scale, oracle):
    V, F = star
    g_c, g_d = P.make_triangle_mesh(V, F), P.make_triangle_mesh(V, F)
    P.cache_moments([g_c])
    rng = np.random.default_rng(4)
    x = affine(W.random_rotation(rng), scale, np.array([1.0, 0.5, -0.25]))
    Ac = P.compute_form_factor(g_c, kg, x, "affine", precision="fast")
    Ad = P.compute_form_factor(g_d, kg, x, "affine", precision="fast")
    Rm = scale * np.linalg.norm(V - 0.5 * (V.min(0) + V.max(0)), axis=1).max()
    M, ks, nser = series_shells(kg, Rm)
    err = shell_rel(Ac, Ad, M)
    print(f"scale {scale}: {len(F)} faces, {nser} series shells, cached vs direct "
          f"{err[:nser].max():.2e}")
    assert nser >= 8
    assert err[:nser].max() < 1e-8
    assert np.array_equal(Ac.reshape(NRAD, M)[nser + 1:], Ad.reshape(NRAD, M)[nser + 1:])
    # the oracle on the series shells with |k| >= 1 (every 7th direction)
    sub = np.concatenate([np.arange(si * M, (si + 1) * M, 7) for si in range(nser)
                          if ks[si] >= 1.0])
    og = oracle.geom_mesh(V, F)
    Ao = oracle.compute_A_grid(og, kg.dirs[sub], kg.kmag[sub], x, oracle.DOF_AFFINE)
    e_o = np.abs(Ac[sub] - Ao).max() / np.abs(Ao).max()
    print(f"against the oracle: {e_o:.2e}")
    assert e_o < 1e-6


def test_not_applicable_is_bit_identical(kg, star):
    V, F = star
    g_c, g_d = P.make_triangle_mesh(V, F), P.make_triangle_mesh(V, F)
    P.cache_moments([g_c])
    rng = np.random.default_rng(5)
    R = W.random_rotation(rng)
    shear = R @ np.array([[1.0, 0.2, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 0.9]])
    x = np.concatenate([[0.1, 0.2, 0.3], shear.reshape(-1)])
    reflect = np.concatenate([[0.0, 0.0, 0.0], (R @ np.diag([1.0, 1.0, -1.0])).reshape(-1)])
    for d in (x, reflect):
        assert np.array_equal(P.compute_form_factor(g_c, kg, d, "affine", precision="fast"),
                              P.compute_form_factor(g_d, kg, d, "affine", precision="fast"))
    xv = (V @ R.T).reshape(-1)
    assert np.array_equal(P.compute_form_factor(g_c, kg, xv, "trimesh", precision="fast"),
                          P.compute_form_factor(g_d, kg, xv, "trimesh", precision="fast"))
    # the FP64 path ignores the cache
    ident = P.IDENTITY_AFFINE
    assert np.array_equal(P.compute_form_factor(g_c, kg, ident, "affine", precision="fp64"),
                          P.compute_form_factor(g_d, kg, ident, "affine", precision="fp64"))
    # small meshes and disk clouds are left uncached
    Vs, Fs = W.uv_sphere(120)
    Pd, Nd, _ = W.sphere_pointcloud(500)
    rho = np.full(500, 0.01)
    small = [P.make_triangle_mesh(Vs * 0.3, Fs), P.make_point_cloud(Pd * 0.3, Nd, rho)]
    ref = [P.compute_form_factor(gg, kg, precision="fast") for gg in small]
    P.cache_moments(small)
    for gg, r in zip(small, ref):
        assert np.array_equal(P.compute_form_factor(gg, kg, precision="fast"), r)


def test_volume_matrix_cached(kg):
    pool = W.serving_pool(n_mesh=6, seed=9, scale=0.5)   # 80 .. 1680 faces
    gc = [P.make_triangle_mesh(m.V, m.F) for m in pool]
    gd = [P.make_triangle_mesh(m.V, m.F) for m in pool]
    P.cache_moments(gc)
    rng = np.random.default_rng(6)
    dofs = S.rigid_dofs(rng, 0.6, len(pool))
    kinds = ["affine"] * len(pool)
    Vc = P.compute_volume_matrix(gc, kg, list(dofs), kinds, ff="fast", gram="fp64")
    Vd = P.compute_volume_matrix(gd, kg, list(dofs), kinds, ff="fast", gram="fp64")
    d = np.sqrt(np.abs(np.outer(np.diag(Vd), np.diag(Vd))))
    e = float(np.max(np.abs(Vc - Vd) / d))
    print(f"faces {[m.n for m in pool]}; cached vs direct V: {e:.2e}")
    assert sum(m.n >= 192 for m in pool) >= 3
    assert e < 1e-9


def test_serving_cached(kg):
    pool = W.serving_pool(n_mesh=8, seed=21, scale=0.3)  # 24 .. 2116 faces
    gc = [P.make_triangle_mesh(m.V, m.F) for m in pool]
    gd = [P.make_triangle_mesh(m.V, m.F) for m in pool]
    P.cache_moments(gc)
    rng = np.random.default_rng(13)
    idx = np.array([0, 1, 2, 3, 4, 5, 6, 7, 6])
    dofs = S.rigid_dofs(rng, 0.8, 9)
    out = {}
    for name, geoms in (("cached", gc), ("direct", gd)):
        with S.Server(kg, gram="fp16x3", hbm_budget=4e9) as srv:
            q = srv.submit([geoms[i] for i in idx], dofs)
            others = [(srv.submit([geoms[i] for i in rng.integers(0, len(pool), size=5)],
                                  S.rigid_dofs(rng, 1.0, 5)), 5) for _ in range(6)]
            qmix = srv.submit([geoms[i] for i in idx], dofs)
            out[name] = (srv.wait(q, 9)[0], srv.wait(qmix, 9)[0])
            for qq, k in others:
                srv.wait(qq, k)
    assert np.array_equal(*out["cached"])
    Vc, Vd = out["cached"][0], out["direct"][0]
    d = np.sqrt(np.abs(np.outer(np.diag(Vd), np.diag(Vd))))
    e = float(np.max(np.abs(Vc - Vd) / d))
    print(f"served, cached vs direct: {e:.2e}")
    assert sum(pool[i].n >= 192 for i in idx) >= 3
    assert e < 1e-7
