"""The Python drop-in module (pointint_amd.pointint_cpp, the reference's
pointint_cpp names) on the GPU: demo/exp.py's call sequence unchanged against
the reference's own recorded outputs, and the planned gradient API against
the oracle."""
import os

import numpy as np
import pytest

import pointint_amd.pointint_cpp as pi
from pointint_amd import workloads as W

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
LEB29 = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "pointint_amd",
                     "data", "lebedev", "lebedev_029.txt")


def test_exp_sequence_reproduces_reference_outputs():
    """demo/exp.py:59-160 with the methods "C++ CUDA" and "C++ CPU": the same
    calls, the same inputs; V_01 against the reference's recorded volumes
    (demo/exp_results.csv).  The FP64 path ("C++ CPU") to 1e-12 of V_00 = 4.3;
    the fast path ("C++ CUDA") to 1e-6 of V_00."""
    z = np.load(os.path.join(GOLD, "exp_two_spheres.npz"))
    dirs_cpp, weights_cpp = pi.load_lebedev(LEB29)
    P_sphere, N_sphere, R_sphere = pi.create_sphere_pointcloud(1000)
    kgrid_cpp = pi.build_kgrid(dirs_cpp, weights_cpp, 32)
    distances = np.arange(0, 2.1, 0.1)
    worst_cpu = worst_cuda = 0.0
    for i, d in enumerate(distances):
        P2 = P_sphere + np.array([d, 0.0, 0.0])
        vol_gt = pi.sphere_sphere_intersection_volume(np.array([0.0, 0.0, 0.0]), 1.0,
                                                      np.array([d, 0.0, 0.0]), 1.0)
        assert abs(vol_gt - z["v_analytic"][i]) < 1e-12
        geom1 = pi.make_point_cloud(P_sphere, N_sphere, R_sphere)
        geom2 = pi.make_point_cloud(P2, N_sphere, R_sphere)
        v_cuda = pi.compute_intersection_volume_cuda(geom1, geom2, kgrid_cpp)
        v_cpu = pi.compute_intersection_volume_cpu(geom1, geom2, kgrid_cpp)
        worst_cpu = max(worst_cpu, abs(v_cpu - z["v_cpu"][i]))
        worst_cuda = max(worst_cuda, abs(v_cuda - z["v_cpu"][i]))
    print(f"max |V - reference CPU|: FP64 path {worst_cpu:.2e}, fast path {worst_cuda:.2e}")
    assert worst_cpu < 4.3e-12
    assert worst_cuda < 4.3e-6


def test_volume_and_analytic_helpers():
    V, F = pi.create_unit_cube_mesh()
    assert abs(pi.compute_volume_cpu(pi.make_triangle_mesh(V, F)) - 1.0) < 1e-14
    P, N, R = pi.create_sphere_pointcloud(1000)
    # the disk cloud's divergence-theorem volume: 4 pi / 3 (each disk's
    # centre on the sphere, normal radial)
    assert abs(pi.compute_volume_cpu(pi.make_point_cloud(P, N, R)) - 4 * np.pi / 3) < 1e-12
    assert pi.box_box_intersection_volume([0, 0, 0], [0.5] * 3, [0.3, 0.2, 0.1], [0.5] * 3) == \
        pytest.approx(0.7 * 0.8 * 0.9)
    assert np.array_equal(pi.translate_points(V, [1.0, 0, 0]), V + [1.0, 0, 0])


@pytest.mark.parametrize("kinds", [("affine", "affine"), ("trimesh", "affine")])
def test_gradient_api_against_oracle(oracle, kinds):
    """compute_intersection_volume_gradient_cuda (docs/DESIGN_PROPOSAL.md:786-793):
    V_12 and dV_12/dtheta of both geometries against the oracle's per-node
    form factors and their DoF derivatives (grad_matrix convention)."""
    rng = np.random.default_rng(17)
    V1, F1 = W.icosphere(2)
    V1 = V1 * 0.45
    V2 = W.star_perturb(V1 / 0.45, rng, 0.1) * 0.4
    kg = pi.build_kgrid(*pi.load_lebedev(os.path.join(os.path.dirname(LEB29), "lebedev_009.txt")), 8)
    g1, g2 = pi.make_triangle_mesh(V1, F1), pi.make_triangle_mesh(V2, F1)
    aff = np.concatenate([[0.2, -0.1, 0.05], (np.eye(3) * 1.05 + 0.02 * rng.normal(size=(3, 3))).ravel()])
    dofs = {"affine": pi.AffineDoF()}
    x1 = V1.ravel() if kinds[0] == "trimesh" else pi.AffineDoF.identity()
    dof1 = pi.TriangleMeshDoF(len(V1)) if kinds[0] == "trimesh" else dofs["affine"]
    res = pi.compute_intersection_volume_gradient_cuda(g1, g2, dof1, pi.AffineDoF(), x1, aff, kg)
    og1, og2 = oracle.geom_mesh(V1, F1), oracle.geom_mesh(V2, F1)
    k1 = oracle.DOF_MESH if kinds[0] == "trimesh" else oracle.DOF_AFFINE
    A1 = oracle.compute_A_grid(og1, kg.dirs, kg.kmag, x1, k1)
    A2 = oracle.compute_A_grid(og2, kg.dirs, kg.kmag, aff, oracle.DOF_AFFINE)
    S = 1.0 / (8 * np.pi ** 3)
    v = S * (kg.w * (A1 * np.conj(A2)).real).sum()
    g1o = np.zeros(len(x1))
    g2o = np.zeros(12)
    for q in range(kg.Q):
        k = kg.dirs[q] * kg.kmag[q]
        d1 = oracle.compute_A_gradient(og1, k, x1, k1)
        d2 = oracle.compute_A_gradient(og2, k, aff, oracle.DOF_AFFINE)
        g1o += S * kg.w[q] * (d1 * np.conj(A2[q])).real
        g2o += S * kg.w[q] * (d2 * np.conj(A1[q])).real
    assert abs(res.volume - v) < 1e-6 * abs(v)
    assert np.abs(res.grad_geom1 - g1o).max() < 1e-6 * np.abs(g1o).max()
    assert np.abs(res.grad_geom2 - g2o).max() < 1e-6 * np.abs(g2o).max()
