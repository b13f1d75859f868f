"""GPU parity of the plain FP64 path against the FP64 oracle and the golden
fixtures (device code through the C ABI)."""
import os

import numpy as np
import pytest

import pointint_amd as P
from pointint_amd import _native as N
from pointint_amd import workloads as W

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def nrel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def small():
    z = np.load(os.path.join(GOLD, "small_mixed_L8_M38.npz"))
    kg = P.kgrid_from_rule(9, 8)
    geoms = [
        P.make_triangle_mesh(z["ico_V"] * 0.6, z["ico_F"]),
        P.make_triangle_mesh(z["ico_V"] * 0.5, z["ico_F"]),
        P.make_point_cloud(z["P"] * 0.4, z["N"], z["r"] * 0.4),
        P.make_gaussian_splat(z["P"] * 0.3, z["N"], z["sig"], z["wt"]),
    ]
    dofs = [z["dofs0"], z["dofs1"], z["dofs2"], z["dofs3"]]
    kinds = ["affine", "trimesh", "affine", "affine"]
    return z, kg, geoms, dofs, kinds


def test_form_factor_fp64_matches_golden(small):
    z, kg, geoms, dofs, kinds = small
    for i in range(4):
        A = P.compute_form_factor(geoms[i], kg, dofs[i], kinds[i], precision="fp64")
        tol = 1e-11 if i < 2 else 1e-9   # disks: device J1 series/asymptotic vs libstdc++
        assert nrel(A, z["J"][:, i]) < tol, (i, nrel(A, z["J"][:, i]))


def test_volume_matrix_fp64_matches_golden(small):
    z, kg, geoms, dofs, kinds = small
    V = P.compute_volume_matrix(geoms, kg, dofs, kinds, ff="fp64", gram="fp64")
    assert np.allclose(V, V.T, rtol=0, atol=0)
    s = np.sqrt(np.outer(np.diag(z["Vmat"]), np.diag(z["Vmat"])))
    assert np.max(np.abs(V - z["Vmat"]) / s) < 1e-10


def test_cube_pair_fp64_matches_golden():
    z = np.load(os.path.join(GOLD, "cube_pair_L32_M302.npz"))
    kg = P.kgrid_from_rule(29, 32)
    g0 = P.make_triangle_mesh(z["V0"], z["F"])
    g1 = P.make_triangle_mesh(z["V0"] + z["shift"], z["F"])
    V = P.compute_volume_matrix([g0, g1], kg, ff="fp64", gram="fp64")
    assert np.allclose(V, z["Vmat"], rtol=1e-12, atol=0)
    A = P.compute_form_factor(g1, kg, precision="fp64")
    assert nrel(A, z["J"][:, 1]) < 1e-12


def test_icospheres_fp64_vs_oracle(oracle):
    wl = W.config_icospheres(n_obj=6, level=2, seed=7, centre_box=0.8)
    d, w = oracle.load_lebedev(P.lebedev_path(29))
    dirs, kmag, wq = oracle.build_kgrid(d, w, 12)
    Vo, Jo = oracle.volume_matrix([oracle.geom_mesh(o.V, o.F) for o in wl.objects], dirs, kmag, wq)
    kg = P.KGrid(dirs, kmag, wq)
    geoms = [P.make_triangle_mesh(o.V, o.F) for o in wl.objects]
    V = P.compute_volume_matrix(geoms, kg, ff="fp64", gram="fp64")
    s = np.sqrt(np.outer(np.diag(Vo), np.diag(Vo)))
    assert np.max(np.abs(V - Vo) / s) < 1e-11
    A = P.compute_form_factor(geoms[3], kg, precision="fp64")
    assert nrel(A, Jo[:, 3]) < 1e-11


def test_packed_triangles_with_vertex_ids(oracle):
    Vi, Fi = W.icosphere(1)
    g_or = oracle.geom_mesh(Vi * 0.5, Fi)
    vid = g_or["vid"].copy()
    vid[3] = -1  # one face without vertex ids: skipped under vertex DoFs
    kg = P.kgrid_from_rule(5, 6)
    g = P.make_packed_triangles(g_or["el"], vid)
    rng = np.random.default_rng(0)
    x = (Vi * 0.5 + 0.02 * rng.normal(size=Vi.shape)).reshape(-1)
    g_or["vid"] = vid
    ref = oracle.compute_A_grid(g_or, kg.dirs, kg.kmag, x, oracle.DOF_MESH)
    A = P.compute_form_factor(g, kg, x, "trimesh", precision="fp64")
    assert nrel(A, ref) < 1e-11
    ref_aff = oracle.compute_A_grid(g_or, kg.dirs, kg.kmag)
    A_aff = P.compute_form_factor(g, kg, precision="fp64")
    assert nrel(A_aff, ref_aff) < 1e-11


def test_gram_fp64_vs_numpy():
    rng = np.random.default_rng(1)
    kg = P.kgrid_from_rule(9, 8)
    J = rng.normal(size=(kg.Q, 37)) + 1j * rng.normal(size=(kg.Q, 37))
    V = P.gram_from_J(J, kg)
    ref = (J.conj().T @ (kg.w[:, None] * J)).real / (8 * np.pi ** 3)
    assert np.allclose(V, ref, rtol=1e-13, atol=1e-15)


def test_empty_geometry_and_errors():
    kg = P.kgrid_from_rule(5, 4)
    Vc, Fc = W.unit_cube()
    cube = P.make_triangle_mesh(Vc, Fc)
    empty = P.make_triangle_mesh(np.zeros((0, 3)), np.zeros((0, 3), dtype=np.int32))
    A = P.compute_form_factor(empty, kg, precision="fp64")
    assert np.all(A == 0)
    V = P.compute_volume_matrix([cube, empty], kg, ff="fp64", gram="fp64")
    assert V[1, 1] == 0 and V[0, 1] == 0 and V[0, 0] > 0.5
    with pytest.raises(N.PiiError) as e:
        P.compute_form_factor(cube, kg, np.zeros(11), precision="fp64")
    assert e.value.code == N.PII_ERR_INVALID
    disks = P.make_point_cloud(*W.sphere_pointcloud(10))
    with pytest.raises(N.PiiError) as e:
        P.compute_form_factor(disks, kg, np.zeros(30), "trimesh", precision="fp64")
    assert e.value.code == N.PII_ERR_UNSUPPORTED
