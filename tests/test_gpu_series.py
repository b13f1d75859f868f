"""GPU parity of the fast path's low-|k| moment series against the FP64 oracle.

The series replaces the per-face evaluation at the shells with |k| R <= 15
(R: radius of the object's world vertices about their bounding-box centre;
for affine DoFs the library bounds it by |A|_2 R_ref).  Its truncation and
rounding errors are below 4e-10 of the object's total area, so on those nodes
the fast path must agree with the FP64 oracle to 1e-9 x area (absolute), far
tighter than the FP32 per-face path (about 1e-7 x area).  On the whole grid the
fast-path gate applies: ||A - A64||_w / ||A64||_w < 1e-6.
"""
import numpy as np
import pytest

import pointint_amd as P
from pointint_amd import workloads as W

pytestmark = pytest.mark.gpu

REACH = 15.0


def wrel(a, b, w):
    return np.sqrt((w * np.abs(a - b) ** 2).sum() / (w * np.abs(b) ** 2).sum())


def area(V, F):
    return 0.5 * np.linalg.norm(np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]]),
                                axis=1).sum()


def radius(V):
    c = 0.5 * (V.min(axis=0) + V.max(axis=0))
    return np.linalg.norm(V - c, axis=1).max()


@pytest.fixture(scope="module")
def grid(oracle):
    d, w = oracle.load_lebedev(P.lebedev_path(29))
    dirs, kmag, wq = oracle.build_kgrid(d, w, 32)
    return dirs, kmag, wq, P.KGrid(dirs, kmag, wq)


def _affine(rng):
    R = W.random_rotation(rng)
    S = np.diag([1.3, 0.6, 0.9])
    Sh = np.array([[1.0, 0.25, 0.0], [0.0, 1.0, -0.1], [0.0, 0.0, 1.0]])
    A = R @ S @ Sh
    t = np.array([1.3, -0.7, 2.1])
    return np.concatenate([t, A.reshape(-1)]), A, t


def _cases(oracle):
    rng = np.random.default_rng(21)
    V3, F3 = W.icosphere(3)
    V2, F2 = W.icosphere(2)
    out = []
    # world-space mesh off the origin, identity DoFs
    Vw = V3 * 0.4 + np.array([1.3, -0.7, 2.1])
    out.append(("mesh_identity", P.make_triangle_mesh(Vw, F3), oracle.geom_mesh(Vw, F3),
                oracle.IDENTITY_DOFS, "affine", oracle.DOF_AFFINE, Vw, F3, radius(Vw)))
    # affine DoFs with rotation, anisotropic scale and shear
    dofs, A, t = _affine(rng)
    Vr = V3 * 0.35
    Vw = Vr @ A.T + t
    Rb = np.linalg.svd(A, compute_uv=False)[0] * radius(Vr)
    out.append(("mesh_affine", P.make_triangle_mesh(Vr, F3), oracle.geom_mesh(Vr, F3), dofs,
                "affine", oracle.DOF_AFFINE, Vw, F3, Rb))
    # vertex DoFs (star-shaped perturbation of a level-2 icosphere)
    Vx = W.star_perturb(V2, rng, 0.15) * 0.5 + np.array([-0.4, 0.9, 0.2])
    out.append(("mesh_vertices", P.make_triangle_mesh(V2 * 0.5, F2), oracle.geom_mesh(V2 * 0.5, F2),
                Vx.reshape(-1), "trimesh", oracle.DOF_MESH, Vx, F2, radius(Vx)))
    return out


@pytest.mark.parametrize("idx", [0, 1, 2])
def test_series_nodes_match_oracle(oracle, grid, idx):
    dirs, kmag, wq, kg = grid
    name, g, og, dofs, kind, okind, Vw, F, Rb = _cases(oracle)[idx]
    ctx = P.Context.default()
    A = P.compute_form_factor(g, kg, dofs, kind, precision="fast")
    ctx.set_series(False)
    try:
        A_off = P.compute_form_factor(g, kg, dofs, kind, precision="fast")
    finally:
        ctx.set_series(True)
    A64 = oracle.compute_A_grid(og, dirs, kmag, dofs, okind)
    ar = area(Vw, F)
    ser = kmag * Rb <= REACH * (1 - 1e-9)
    assert ser.sum() > 0.5 * len(kmag), name        # most shells take the series
    e_on = np.abs(A - A64)[ser].max() / ar
    e_off = np.abs(A_off - A64)[ser].max() / ar
    print(f"{name}: series nodes {ser.sum()}/{len(kmag)}, max|dA|/area series {e_on:.2e}, "
          f"per-face {e_off:.2e}")
    assert e_on < 1e-9, (name, e_on)
    assert wrel(A, A64, wq) < 1e-6, name
    assert wrel(A_off, A64, wq) < 1e-6, name


def test_series_volume_matrix(oracle, grid):
    """Volume matrix with and without the series against the FP64 oracle."""
    dirs, kmag, wq, kg = grid
    wl = W.config_icospheres(n_obj=10, level=2, seed=13, centre_box=1.2)
    geoms = [P.make_triangle_mesh(o.V, o.F) for o in wl.objects]
    Vo, _ = oracle.volume_matrix([oracle.geom_mesh(o.V, o.F) for o in wl.objects], dirs, kmag,
                                 wq)
    d = np.sqrt(np.outer(np.diag(Vo), np.diag(Vo)))
    ctx = P.Context.default()
    for on in (True, False):
        ctx.set_series(on)
        try:
            for gram in ("fp64", "fp16x3"):
                V = P.compute_volume_matrix(geoms, kg, ff="fast", gram=gram)
                assert np.array_equal(V, V.T)
                e = np.max(np.abs(V - Vo) / d)
                assert e < 1e-5, (on, gram, e)
        finally:
            ctx.set_series(True)


def _oracle_vjp(oracle, kg, ogeoms, xs, G, kind):
    J = np.stack([oracle.compute_A_grid(g, kg.dirs, kg.kmag, x, kind) for g, x in zip(ogeoms, xs)],
                 axis=1)
    C = G + G.T
    return [oracle.vjp(ogeoms[i], xs[i], kind, kg.dirs, kg.kmag, kg.w, J @ C[:, i])
            for i in range(len(ogeoms))]


@pytest.mark.parametrize("kind", ["trimesh", "affine"])
def test_series_adjoint(oracle, kind):
    """Vertex / affine gradients of L = sum G_ij V_ij with the series adjoint
    (leading shells) and with the per-face adjoint everywhere, against the
    oracle's VJP of the reference gradient: ||g - g64|| / ||g64|| < 1e-5."""
    rng = np.random.default_rng(31)
    d, w = oracle.load_lebedev(P.lebedev_path(9))
    dirs, kmag, wq = oracle.build_kgrid(d, w, 24)
    kg = P.KGrid(dirs, kmag, wq)
    Vi, Fi = W.icosphere(2)
    verts = [W.star_perturb(Vi, rng, 0.1) * rng.uniform(0.3, 0.45) + rng.uniform(0, 0.6, size=3)
             for _ in range(4)]
    geoms = [P.make_triangle_mesh(v, Fi) for v in verts]
    ogeoms = [oracle.geom_mesh(v, Fi) for v in verts]
    if kind == "trimesh":
        xs = [v.reshape(-1) for v in verts]
        okind = oracle.DOF_MESH
    else:
        xs = [np.concatenate([rng.normal(size=3) * 0.1,
                              (W.random_rotation(rng) @ np.diag(rng.uniform(0.8, 1.2, 3))).ravel()])
              for _ in verts]
        okind = oracle.DOF_AFFINE
    G = rng.normal(size=(4, 4))
    ref = _oracle_vjp(oracle, kg, ogeoms, xs, G, okind)
    ctx = P.Context.default()
    out = {}
    for on in (True, False):
        ctx.set_series(on)
        try:
            out[on], _ = P.volume_matrix_vjp(geoms, kg, G, xs, [kind] * 4)
        finally:
            ctx.set_series(True)
    nser = (kmag[None, :] * np.array([radius(v) for v in verts])[:, None] <= REACH).sum(1)
    assert nser.min() >= 4 * len(d) and nser.max() < len(kmag), nser  # both paths run
    for on in (True, False):
        for g, r in zip(out[on], ref):
            e = np.linalg.norm(g - r) / np.linalg.norm(r)
            print(f"{kind} series={on}: rel err {e:.2e}")
            assert e < 1e-5, (kind, on, e)


@pytest.mark.parametrize("kind", ["disk", "gaussian"])
def test_surfel_series_nodes_match_oracle(oracle, grid, kind):
    """Disks / Gaussians under a rotation + anisotropic scale: on the series
    shells the fast path matches the FP64 oracle to 1e-9 of the total area
    (disks) or weight (Gaussians); whole grid to the 1e-6 weighted gate."""
    dirs, kmag, wq, kg = grid
    rng = np.random.default_rng(41)
    n = 3000
    Pc, Nn, _ = W.sphere_pointcloud(n)
    Pc = Pc * 0.35
    s = rng.uniform(0.005, 0.02, size=n)
    A = W.random_rotation(rng) @ np.diag([1.2, 0.8, 1.0])
    t = np.array([0.4, 1.1, -0.6])
    dofs = np.concatenate([t, A.ravel()])
    if kind == "disk":
        g, og, tot = P.make_point_cloud(Pc, Nn, s), oracle.geom_disks(Pc, Nn, s), (np.pi * s * s).sum()
        cap = 3.0 / (0.25 * s.max() ** 2 * (A ** 2).sum())
    else:
        wt = rng.uniform(0.5, 1.5, size=n)
        g, og, tot = P.make_gaussian_splat(Pc, Nn, s, wt), oracle.geom_gaussians(Pc, Nn, s, wt), wt.sum()
        cap = 0.49 / (0.5 * s.max() ** 2 * (A ** 2).sum())
    ctx = P.Context.default()
    Af = P.compute_form_factor(g, kg, dofs, "affine", precision="fast")
    ctx.set_series(False)
    try:
        A_off = P.compute_form_factor(g, kg, dofs, "affine", precision="fast")
    finally:
        ctx.set_series(True)
    A64 = oracle.compute_A_grid(og, dirs, kmag, dofs, oracle.DOF_AFFINE)
    Rb = np.linalg.svd(A, compute_uv=False)[0] * radius(Pc)
    ser = (kmag * Rb <= REACH * (1 - 1e-9)) & (kmag ** 2 <= cap * (1 - 1e-9))
    assert ser.sum() > 0.5 * len(kmag), kind
    e_on = np.abs(Af - A64)[ser].max() / tot
    e_off = np.abs(A_off - A64)[ser].max() / tot
    print(f"{kind}: series nodes {ser.sum()}/{len(kmag)}, max|dA|/total series {e_on:.2e}, "
          f"per-element {e_off:.2e}")
    assert e_on < 1e-9, (kind, e_on)
    assert wrel(Af, A64, wq) < 1e-6, kind
    assert wrel(A_off, A64, wq) < 1e-6, kind
