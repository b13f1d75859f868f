"""GPU adjoint (VJP) against the FP64 oracle's VJP of the reference gradient
and against central finite differences of the FP64 forward."""
import numpy as np
import pytest

import pointint_amd as P
from pointint_amd import workloads as W

pytestmark = pytest.mark.gpu


def _setup(oracle, n=3, level=1, order=5, nrad=6, seed=0):
    rng = np.random.default_rng(seed)
    Vi, Fi = W.icosphere(level)
    verts = [Vi * rng.uniform(0.35, 0.5) + rng.uniform(0.0, 0.4, size=3) for _ in range(n)]
    d, w = oracle.load_lebedev(P.lebedev_path(order))
    dirs, kmag, wq = oracle.build_kgrid(d, w, nrad)
    kg = P.KGrid(dirs, kmag, wq)
    geoms = [P.make_triangle_mesh(v, Fi) for v in verts]
    ogeoms = [oracle.geom_mesh(v, Fi) for v in verts]
    return verts, Fi, kg, geoms, ogeoms


def _oracle_grads(oracle, kg, ogeoms, xs, G, kind):
    n = len(ogeoms)
    J = np.stack([oracle.compute_A_grid(g, kg.dirs, kg.kmag, x, kind) for g, x in zip(ogeoms, xs)],
                 axis=1)
    C = G + G.T
    out = []
    for i in range(n):
        B = J @ C[:, i]
        out.append(oracle.vjp(ogeoms[i], xs[i], kind, kg.dirs, kg.kmag, kg.w, B))
    return out


def test_vjp_mesh_dofs_matches_oracle(oracle):
    verts, Fi, kg, geoms, ogeoms = _setup(oracle)
    xs = [v.reshape(-1) for v in verts]
    n = len(xs)
    G = np.triu(np.ones((n, n)), 1)           # L = sum_{i<j} V_ij
    grads, V = P.volume_matrix_vjp(geoms, kg, G, xs, ["trimesh"] * n)
    ref = _oracle_grads(oracle, kg, ogeoms, xs, G, oracle.DOF_MESH)
    for g, r in zip(grads, ref):
        assert np.linalg.norm(g - r) / np.linalg.norm(r) < 1e-5


def test_vjp_general_cotangent_and_diagonal(oracle):
    verts, Fi, kg, geoms, ogeoms = _setup(oracle, n=4, level=2, order=9, nrad=8, seed=1)
    xs = [v.reshape(-1) for v in verts]
    G = np.random.default_rng(2).normal(size=(4, 4))  # includes the diagonal
    grads, _ = P.volume_matrix_vjp(geoms, kg, G, xs, ["trimesh"] * 4)
    ref = _oracle_grads(oracle, kg, ogeoms, xs, G, oracle.DOF_MESH)
    for g, r in zip(grads, ref):
        assert np.linalg.norm(g - r) / np.linalg.norm(r) < 1e-5


def test_vjp_affine_dofs_matches_oracle(oracle):
    verts, Fi, kg, geoms, ogeoms = _setup(oracle, n=3, seed=3)
    rng = np.random.default_rng(4)
    dofs = [np.concatenate([rng.normal(size=3) * 0.1, (np.eye(3) + 0.1 * rng.normal(size=(3, 3))).ravel()])
            for _ in range(3)]
    G = np.triu(np.ones((3, 3)), 1)
    grads, _ = P.volume_matrix_vjp(geoms, kg, G, dofs, ["affine"] * 3)
    ref = _oracle_grads(oracle, kg, ogeoms, dofs, G, oracle.DOF_AFFINE)
    for g, r in zip(grads, ref):
        assert np.linalg.norm(g - r) / np.linalg.norm(r) < 1e-5


def test_vjp_matches_central_differences():
    """dL/dv against FP64 central differences (h = 1e-4) of the FP64 GPU forward."""
    rng = np.random.default_rng(5)
    Vi, Fi = W.icosphere(2)
    verts = [Vi * 0.4 + c for c in ([0, 0, 0], [0.3, 0.1, 0.0], [0.1, 0.35, 0.2])]
    kg = P.kgrid_from_rule(9, 10)
    geoms = [P.make_triangle_mesh(v, Fi) for v in verts]
    xs = [v.reshape(-1).copy() for v in verts]
    n = 3
    G = np.triu(np.ones((n, n)), 1)
    grads, _ = P.volume_matrix_vjp(geoms, kg, G, xs, ["trimesh"] * n)

    def loss(xl):
        V = P.compute_volume_matrix(geoms, kg, xl, ["trimesh"] * n, ff="fp64", gram="fp64")
        return np.triu(V, 1).sum()

    h = 1e-4
    for _ in range(8):
        i = int(rng.integers(n))
        dof = int(rng.integers(xs[i].shape[0]))
        xp = [x.copy() for x in xs]
        xm = [x.copy() for x in xs]
        xp[i][dof] += h
        xm[i][dof] -= h
        fd = (loss(xp) - loss(xm)) / (2 * h)
        scale = np.abs(grads[i]).max()
        assert abs(grads[i][dof] - fd) < 1e-5 * scale + 1e-9, (i, dof, grads[i][dof], fd)


def test_autograd_function():
    torch = pytest.importorskip("torch")
    from pointint_amd.autograd import intersection_volume_matrix, pairwise_overlap_loss
    Vi, Fi = W.icosphere(2)
    verts = [Vi * 0.4 + c for c in ([0, 0, 0], [0.3, 0.1, 0.0], [0.1, 0.35, 0.2])]
    kg = P.kgrid_from_rule(9, 10)
    geoms = [P.make_triangle_mesh(v, Fi) for v in verts]
    tv = [torch.tensor(v, dtype=torch.float32, device="cuda", requires_grad=True) for v in verts]
    V = intersection_volume_matrix(geoms, kg, tv)
    loss = pairwise_overlap_loss(V)
    loss.backward()
    G = np.triu(np.ones((3, 3)), 1)
    ref, _ = P.volume_matrix_vjp(geoms, kg, G, [v.reshape(-1) for v in verts], ["trimesh"] * 3)
    for t, r in zip(tv, ref):
        assert t.grad is not None and t.grad.dtype == torch.float32
        g = t.grad.detach().cpu().double().numpy().reshape(-1)
        assert np.linalg.norm(g - r) / np.linalg.norm(r) < 1e-5
