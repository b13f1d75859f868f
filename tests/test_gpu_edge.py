"""Edge cases of the fast path against the FP64 oracle.

* Few objects (N = 1, 2, 3): every form-factor kernel runs element-split
  (partial sums in split order, finishing kernels); triangles, disks and
  Gaussians under affine DoFs, series on.  Gates as everywhere for the fast
  path: ||A - A64||_w / ||A64||_w < 1e-6, |V - V64| < 1e-5 sqrt(V_ii V_jj).
* Empty objects inside a fast batch give zero rows and columns and leave the
  others untouched; zero-area (repeated-vertex) faces contribute nothing;
  an object far from the origin (|t| = 40) keeps its accuracy (phases in
  FP64).
* Shells at |k| = 0 and 1e-12 give exactly zero form factors (the
  reference's 1e-10 cut) on every path, series included; the adjoint agrees.
"""
import numpy as np
import pytest

import pointint_amd as P
from pointint_amd import workloads as W

pytestmark = pytest.mark.gpu


def wrel(a, b, w):
    return np.sqrt((w * np.abs(a - b) ** 2).sum() / (w * np.abs(b) ** 2).sum())


@pytest.fixture(scope="module")
def grid(oracle):
    d, w = oracle.load_lebedev(P.lebedev_path(29))
    dirs, kmag, wq = oracle.build_kgrid(d, w, 24)
    return dirs, kmag, wq, P.KGrid(dirs, kmag, wq)


def _objects(oracle, rng):
    V3, F3 = W.icosphere(3)
    Vm = W.star_perturb(V3, rng, 0.1) * 0.4
    Pc, Nn, _ = W.sphere_pointcloud(3000)
    rho = rng.uniform(0.01, 0.02, size=3000)
    sig = rng.uniform(0.01, 0.02, size=3000)
    wt = rng.uniform(0.5, 1.5, size=3000) * 1e-3
    objs = [(P.make_triangle_mesh(Vm, F3), oracle.geom_mesh(Vm, F3)),
            (P.make_point_cloud(Pc * 0.35, Nn, rho), oracle.geom_disks(Pc * 0.35, Nn, rho)),
            (P.make_gaussian_splat(Pc * 0.3, Nn, sig, wt),
             oracle.geom_gaussians(Pc * 0.3, Nn, sig, wt))]
    dofs = [np.concatenate([rng.uniform(-0.15, 0.15, 3),
                            (W.random_rotation(rng) @ np.diag(rng.uniform(0.8, 1.2, 3))).ravel()])
            for _ in objs]
    return objs, dofs


@pytest.mark.parametrize("n", [1, 2, 3])
def test_few_objects_split_path(oracle, grid, n):
    dirs, kmag, wq, kg = grid
    rng = np.random.default_rng(91)
    objs, dofs = _objects(oracle, rng)
    objs, dofs = objs[:n], dofs[:n]
    J64 = np.stack([oracle.compute_A_grid(og, dirs, kmag, x, oracle.DOF_AFFINE)
                    for (_, og), x in zip(objs, dofs)], axis=1)
    V64 = oracle.gram(J64, wq)
    for i, ((g, _), x) in enumerate(zip(objs, dofs)):
        A = P.compute_form_factor(g, kg, x, "affine", precision="fast")
        e = wrel(A, J64[:, i], wq)
        assert e < 1e-6, (n, i, e)
    for gram in ("fp64", "fp16x3"):
        V = P.compute_volume_matrix([g for g, _ in objs], kg, dofs, ["affine"] * n, ff="fast",
                                    gram=gram)
        d = np.sqrt(np.outer(np.diag(V64), np.diag(V64)))
        e = np.max(np.abs(V - V64) / d)
        print(f"N={n} {gram}: scaled V error {e:.2e}")
        assert e < 1e-5, (n, gram, e)


def test_empty_objects_in_fast_batch(oracle, grid):
    dirs, kmag, wq, kg = grid
    V2, F2 = W.icosphere(2)
    meshes = [(V2 * 0.3 + c, F2) for c in ([0, 0, 0], [0.2, 0.1, 0.0], [0.5, -0.2, 0.1])]
    empty = P.make_triangle_mesh(np.zeros((0, 3)), np.zeros((0, 3), dtype=np.int32))
    geoms = [P.make_triangle_mesh(v, f) for v, f in meshes]
    Vo, _ = oracle.volume_matrix([oracle.geom_mesh(v, f) for v, f in meshes], dirs, kmag, wq)
    V = P.compute_volume_matrix([geoms[0], empty, geoms[1], empty, geoms[2]], kg, ff="fast",
                                gram="fp64")
    keep = [0, 2, 4]
    assert np.all(V[[1, 3], :] == 0) and np.all(V[:, [1, 3]] == 0)
    d = np.sqrt(np.outer(np.diag(Vo), np.diag(Vo)))
    assert np.max(np.abs(V[np.ix_(keep, keep)] - Vo) / d) < 1e-5


def test_degenerate_faces_and_far_objects(oracle, grid):
    dirs, kmag, wq, kg = grid
    V3, F3 = W.icosphere(3)
    Vm = V3 * 0.35
    # append zero-area faces: repeated vertices and a collinear triple
    nv = len(Vm)
    extra_v = np.array([[0.0, 0.0, 0.5], [0.0, 0.0, 0.6], [0.0, 0.0, 0.7]])
    Vd = np.vstack([Vm, extra_v])
    Fd = np.vstack([F3, [[0, 0, 1], [5, 5, 5], [nv, nv + 1, nv + 2]]]).astype(np.int32)
    gd = P.make_triangle_mesh(Vd, Fd)
    A_deg = P.compute_form_factor(gd, kg, precision="fast")
    A_ref = oracle.compute_A_grid(oracle.geom_mesh(Vm, F3), dirs, kmag)
    assert wrel(A_deg, A_ref, wq) < 1e-6
    # the same mesh 40 units from the origin: A picks up only the phase e^{i k.t}
    t = np.array([40.0, -25.0, 17.0])
    x = np.concatenate([t, np.eye(3).ravel()])
    g = P.make_triangle_mesh(Vm, F3)
    A_far = P.compute_form_factor(g, kg, x, "affine", precision="fast")
    A64 = oracle.compute_A_grid(oracle.geom_mesh(Vm, F3), dirs, kmag, x, oracle.DOF_AFFINE)
    e = wrel(A_far, A64, wq)
    print(f"far object: weighted rel error {e:.2e}")
    assert e < 1e-6
    ph = np.exp(1j * (dirs * kmag[:, None]) @ t)
    assert wrel(A_far, A_ref * ph, wq) < 1e-6


def test_disk_bessel_large_arguments(oracle):
    """Disks with x = rho |k_perp| from 0 to about 170 over the grid (radii
    up to 0.15, |k| up to 1,036): per node |A - A64| <= 5e-7 of the total
    area (2 area |dJ1/x| per disk, the FP32 Bessel factor good to ~1e-7
    absolute on both of its branches, x < 4 and x >= 4); with and without
    the moment series."""
    d, w = oracle.load_lebedev(P.lebedev_path(29))
    dirs, kmag, wq = oracle.build_kgrid(d, w, 48)
    kg = P.KGrid(dirs, kmag, wq)
    rng = np.random.default_rng(17)
    n = 400
    Pc, Nn, _ = W.sphere_pointcloud(n)
    Pc = Pc * 0.3
    rho = rng.uniform(0.01, 0.15, size=n)
    x = np.concatenate([[0.05, -0.1, 0.2], (W.random_rotation(rng) @ np.diag([1.1, 0.9, 1.0])).ravel()])
    g = P.make_point_cloud(Pc, Nn, rho)
    A64 = oracle.compute_A_grid(oracle.geom_disks(Pc, Nn, rho), dirs, kmag, x, oracle.DOF_AFFINE)
    area = (np.pi * rho * rho).sum()
    xmax = rho.max() * kmag.max() * 1.1
    assert xmax > 100, xmax
    ctx = P.Context.default()
    for series in (True, False):
        ctx.set_series(series)
        try:
            A = P.compute_form_factor(g, kg, x, "affine", precision="fast")
        finally:
            ctx.set_series(True)
        e = np.abs(A - A64).max() / area
        print(f"series={series}: max|dA| / area {e:.2e}, x up to {xmax:.0f}")
        assert e < 5e-7, (series, e)
        assert wrel(A, A64, wq) < 1e-6


def test_zero_and_tiny_k_shells(oracle):
    """A shell-structured grid whose first shells sit at |k| = 0 and 1e-12
    (below the reference's 1e-10 cut, where A and dA/dtheta are 0:
    triangle_mesh_dof.cpp:28, affine_dof.cpp:72,181) and 1e-9 (just above
    it), then ordinary shells up to |k| = 30.  The moment series covers the
    low shells, the per-element kernels the rest.  An open hemisphere (whose
    A tends to khat . sum S != 0 as k -> 0, so only the cut gives 0), a
    closed sphere, disks and Gaussians.  Form factors (fast and FP64) exactly
    0 on the cut shells and within the usual gates elsewhere; V against the
    oracle; the vertex-DoF adjoint against the oracle's VJP."""
    rng = np.random.default_rng(91)
    d, w = oracle.load_lebedev(P.lebedev_path(9))
    shells = np.array([0.0, 1e-12, 1e-9, 0.3, 1.0, 3.0, 10.0, 30.0])
    M = len(w)
    dirs = np.tile(d, (len(shells), 1))
    kmag = np.repeat(shells, M)
    wq = np.tile(w, len(shells)) * np.repeat(rng.uniform(0.5, 1.5, len(shells)), M)
    kg = P.KGrid(dirs, kmag, wq)
    V2, F2 = W.icosphere(2)
    keep = np.all(V2[F2][:, :, 2] >= -1e-12, axis=1)      # faces of the upper half
    Fh = F2[keep]
    Pc, Nn, _ = W.sphere_pointcloud(500)
    rho = rng.uniform(0.02, 0.04, size=500)
    sig = rng.uniform(0.02, 0.04, size=500)
    wt = rng.uniform(0.5, 1.5, size=500) * 1e-2
    objs = [(P.make_triangle_mesh(V2 * 0.3, Fh), oracle.geom_mesh(V2 * 0.3, Fh)),
            (P.make_triangle_mesh(V2 * 0.35 + 0.1, F2), oracle.geom_mesh(V2 * 0.35 + 0.1, F2)),
            (P.make_point_cloud(Pc * 0.3, Nn, rho), oracle.geom_disks(Pc * 0.3, Nn, rho)),
            (P.make_gaussian_splat(Pc * 0.3, Nn, sig, wt),
             oracle.geom_gaussians(Pc * 0.3, Nn, sig, wt))]
    cut = kmag < 1e-10
    Js = []
    for g, og in objs:
        A64 = oracle.compute_A_grid(og, dirs, kmag)
        Js.append(A64)
        assert np.all(A64[cut] == 0)
        assert np.abs(A64[kmag == 1e-9]).max() > 0
        for prec, gate in (("fast", 1e-6), ("fp64", 1e-12)):
            A = P.compute_form_factor(g, kg, precision=prec)
            assert np.all(A[cut] == 0), prec
            e = wrel(A, A64, wq)
            print(f"{prec}: {e:.2e}")
            assert e < gate, (prec, e)
    Vo, _ = oracle.volume_matrix([og for _, og in objs], dirs, kmag, wq)
    V = P.compute_volume_matrix([g for g, _ in objs], kg, gram="fp16x3")
    dd = np.sqrt(np.outer(np.diag(Vo), np.diag(Vo)))
    assert np.max(np.abs(V - Vo) / dd) < 1e-5
    # vertex-DoF adjoint of L = V_01 over the two meshes
    ms = objs[:2]
    xs = [(V2 * 0.3).ravel(), (V2 * 0.35 + 0.1).ravel()]
    G = np.array([[0.0, 1.0], [0.0, 0.0]])
    grads, _ = P.volume_matrix_vjp([g for g, _ in ms], kg, G, xs, ["trimesh"] * 2)
    J = np.stack([oracle.compute_A_grid(og, dirs, kmag, x, oracle.DOF_MESH)
                  for (_, og), x in zip(ms, xs)], axis=1)
    C = G + G.T
    for i in range(2):
        r = oracle.vjp(ms[i][1], xs[i], oracle.DOF_MESH, dirs, kmag, wq, J @ C[:, i])
        e = np.linalg.norm(grads[i] - r) / np.linalg.norm(r)
        print(f"vjp {i}: {e:.2e}")
        assert e < 1e-5, (i, e)


def test_partial_face_blocks_and_small_meshes(oracle, grid):
    """The per-face kernel takes a patch's faces 16 at a time, four per lane
    group and round; a patch's last block is partial and its lanes past the
    patch read whatever records follow (the next object's, or the padding
    past the last object's).  Meshes of 1..47 faces (one partial block, a
    full block plus a tail, several blocks) in one batch, the smallest last,
    under affine and vertex DoFs, against the oracle on every node."""
    dirs, kmag, wq, kg = grid
    rng = np.random.default_rng(17)
    V3, F3 = W.icosphere(3)
    geoms, ogeoms, dofs, kinds = [], [], [], []
    for nf in (47, 33, 31, 17, 16, 15, 5, 3, 2, 1):
        # a connected cap of nf faces of a perturbed sphere (an open mesh)
        c = V3[F3].mean(axis=1)
        order = np.argsort(np.linalg.norm(c - V3[F3[0]].mean(axis=0), axis=1))
        F = F3[order[:nf]]
        used, inv = np.unique(F, return_inverse=True)
        V = W.star_perturb(V3[used], rng, 0.1) * rng.uniform(0.3, 0.5) + rng.uniform(-0.2, 0.2, 3)
        F = inv.reshape(-1, 3).astype(np.int32)
        geoms.append(P.make_triangle_mesh(V, F))
        ogeoms.append(oracle.geom_mesh(V, F))
        if nf % 2:
            dofs.append(np.concatenate([rng.uniform(-0.1, 0.1, 3),
                                        (W.random_rotation(rng)).ravel()]))
            kinds.append("affine")
        else:
            dofs.append((V + rng.normal(scale=0.01, size=V.shape)).ravel())
            kinds.append("trimesh")
    odk = {"affine": oracle.DOF_AFFINE, "trimesh": oracle.DOF_MESH}
    J64 = np.stack([oracle.compute_A_grid(og, dirs, kmag, x, odk[k])
                    for og, x, k in zip(ogeoms, dofs, kinds)], axis=1)
    for i, (g, x, k) in enumerate(zip(geoms, dofs, kinds)):
        A = P.compute_form_factor(g, kg, x, k, precision="fast")
        e = wrel(A, J64[:, i], wq)
        # per node on the shells with |k| >= 1 (the per-face kernel's), against
        # the largest |A| of the shell
        M = len(wq) // 24
        worst = 0.0
        for s in range(24):
            sl = slice(s * M, (s + 1) * M)
            if kmag[s * M] < 1.0:
                continue
            ref = np.max(np.abs(J64[sl, i]))
            if ref > 0:
                worst = max(worst, float(np.max(np.abs(A[sl] - J64[sl, i])) / ref))
        print(f"{len(ogeoms[i]['el'])} faces, {k}: weighted {e:.2e}, per node {worst:.2e}")
        assert e < 1e-6, (i, k, e)
        assert worst < 1e-6, (i, k, worst)
    V64 = oracle.gram(J64, wq)
    V = P.compute_volume_matrix(geoms, kg, dofs, kinds, ff="fast", gram="fp64")
    d = np.sqrt(np.abs(np.outer(np.diag(V64), np.diag(V64))))
    assert np.max(np.abs(V - V64) / d) < 1e-5
