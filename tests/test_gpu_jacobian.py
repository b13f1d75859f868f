"""dA(k_q)/dtheta over a k-grid (`pii_form_factor_jacobian`, the reference's
gradient plug-ins) against the FP64 oracle's per-node gradients
(`oracle.compute_A_gradient`, pinned against central differences in
test_oracle_pins.py): triangle meshes under vertex DoFs and AffineDoF
triangles, disks and Gaussians.  FP64 on both sides: max |J - J64| <= 1e-10
of max |J64| (1e-9 for disks: the device J1 series/Hankel against
std::cyl_bessel_j)."""
import numpy as np
import pytest

import pointint_amd as P
from pointint_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def grid(oracle):
    d, w = oracle.load_lebedev(P.lebedev_path(9))
    dirs, kmag, wq = oracle.build_kgrid(d, w, 6)
    return dirs, kmag, P.KGrid(dirs, kmag, wq)


def _oracle_J(oracle, og, dirs, kmag, x, kind):
    K = dirs * kmag[:, None]
    return np.stack([np.asarray(oracle.compute_A_gradient(og, K[q], x, kind)).ravel()
                     for q in range(len(kmag))])


def _affine(rng):
    A = W.random_rotation(rng) @ np.diag(rng.uniform(0.8, 1.2, 3)) + 0.05 * rng.standard_normal((3, 3))
    return np.concatenate([rng.uniform(-0.2, 0.2, 3), A.ravel()])


def test_jacobian_mesh_vertex_dofs(oracle, grid):
    dirs, kmag, kg = grid
    rng = np.random.default_rng(3)
    V, F = W.icosphere(1)
    V = W.star_perturb(V, rng, 0.1) * 0.4
    g = P.make_triangle_mesh(V, F)
    x = (V + 0.01 * rng.standard_normal(V.shape)).ravel()
    J = P.compute_form_factor_jacobian(g, kg, x, "trimesh")
    J64 = _oracle_J(oracle, oracle.geom_mesh(V, F), dirs, kmag, x, oracle.DOF_MESH)
    assert J.shape == (len(kmag), x.size)
    e = np.max(np.abs(J - J64)) / np.max(np.abs(J64))
    print(f"mesh: max |J - J64| / max |J64| = {e:.2e}")
    assert e < 1e-10


@pytest.mark.parametrize("kind", ["tri", "disk", "gaussian"])
def test_jacobian_affine(oracle, grid, kind):
    dirs, kmag, kg = grid
    rng = np.random.default_rng(5)
    x = _affine(rng)
    if kind == "tri":
        V, F = W.icosphere(2)
        V = W.star_perturb(V, rng, 0.1) * 0.4
        g, og = P.make_triangle_mesh(V, F), oracle.geom_mesh(V, F)
    else:
        Pc, Nn, _ = W.sphere_pointcloud(300)
        s = rng.uniform(0.02, 0.08, size=300)
        if kind == "disk":
            g, og = P.make_point_cloud(Pc * 0.4, Nn, s), oracle.geom_disks(Pc * 0.4, Nn, s)
        else:
            wt = rng.uniform(0.5, 1.5, size=300)
            g = P.make_gaussian_splat(Pc * 0.4, Nn, s, wt)
            og = oracle.geom_gaussians(Pc * 0.4, Nn, s, wt)
    J = P.compute_form_factor_jacobian(g, kg, x, "affine")
    J64 = _oracle_J(oracle, og, dirs, kmag, x, oracle.DOF_AFFINE)
    e = np.max(np.abs(J - J64)) / np.max(np.abs(J64))
    print(f"affine {kind}: max |J - J64| / max |J64| = {e:.2e}")
    assert e < (1e-9 if kind == "disk" else 1e-10)


def test_jacobian_rejects_bad_dof_count(grid):
    _, _, kg = grid
    V, F = W.icosphere(1)
    g = P.make_triangle_mesh(V, F)
    from pointint_amd import _native as N
    with pytest.raises(N.PiiError):
        P.compute_form_factor_jacobian(g, kg, np.zeros(11), "affine")
