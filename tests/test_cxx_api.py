"""The reference-named C++ API (include/pointint/pointint.hpp) built on the C
ABI: symbols on the CPU, results on the GPU against the golden fixture and the
oracle."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "pointint_amd")
LEB29 = os.path.join(PKG, "data", "lebedev", "lebedev_029.txt")


def test_cxx_library_exports_reference_names():
    out = subprocess.run(["nm", "-D", "-C", "--defined-only", os.path.join(PKG, "libpointint_cxx.so")],
                         capture_output=True, text=True, check=True).stdout
    for name in ("PoIntInt::compute_intersection_volume_matrix_cuda",
                 "PoIntInt::compute_intersection_volume_cuda", "PoIntInt::compute_Ak_cuda",
                 "PoIntInt::build_kgrid", "PoIntInt::load_lebedev_txt",
                 "PoIntInt::make_triangle_mesh", "PoIntInt::make_point_cloud",
                 "PoIntInt::gauss_legendre_interval", "PoIntInt::compute_volume_cpu",
                 "PoIntInt::compute_Ak_gradient_cuda"):
        assert name in out, name


def _run(mode):
    r = subprocess.run([os.path.join(PKG, "example_cube_pair"), LEB29, mode], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    vals = {}
    for line in r.stdout.splitlines():
        parts = line.split()
        if parts[0] == "V":
            vals["V"] = np.array([float(x) for x in parts[1:5]]).reshape(2, 2)
        elif parts[0] == "pair":
            vals["pair"] = float(parts[1])
            vals["grad_t2"] = np.array([float(x) for x in parts[3:6]])
        elif parts[0] == "self":
            vals["self"] = float(parts[1])
        elif parts[0] == "hcross_tt":
            vals["hcross_tt"] = np.array([float(x) for x in parts[1:10]]).reshape(3, 3)
        elif parts[0] == "gmesh2":
            vals["gmesh2"] = np.array([float(x) for x in parts[1:]])
        elif parts[0] == "hcross_mesh":
            vals["hcross_mesh"] = np.array([float(x) for x in parts[1:]]).reshape(24, 24)
        elif parts[0] == "ak":
            vals["ak"] = float(parts[1]) + 1j * float(parts[2])
        elif parts[0] == "gradk":
            x = np.array([float(v) for v in parts[1:25]])
            vals["gradk"] = x[0::2] + 1j * x[1::2]
    return vals


@pytest.mark.gpu
@pytest.mark.parametrize("mode,tol", [("fp64", 1e-12), ("fast", 1e-5)])
def test_cxx_example_cube_pair(oracle, mode, tol):
    z = np.load(os.path.join(ROOT, "tests", "golden", "cube_pair_L32_M302.npz"))
    v = _run(mode)
    assert np.allclose(v["V"], z["Vmat"], rtol=tol, atol=0)
    assert abs(v["pair"] - z["Vmat"][0, 1]) < tol
    assert abs(v["self"] - 1.0) < 1e-14
    # dV01/dt2 against the oracle's VJP (B = A_0 for object 1)
    d, w = oracle.load_lebedev(LEB29)
    dirs, kmag, wq = oracle.build_kgrid(d, w, 32)
    g1 = oracle.geom_mesh(z["V0"] + z["shift"], z["F"])
    B = z["J"][:, 0]
    ref = oracle.vjp(g1, oracle.IDENTITY_DOFS, oracle.DOF_AFFINE, dirs, kmag, wq, B)[:3]
    assert np.allclose(v["grad_t2"], ref, rtol=1e-4, atol=1e-9)
    # Gauss-Newton cross block, translations: dA/dt = i k A, so
    # H_tt = (1/8 pi^3) sum_q w_q k_a k_b Re(A_0 conj A_1)
    k = dirs * kmag[:, None]
    c = wq * (z["J"][:, 0] * np.conj(z["J"][:, 1])).real / (8 * np.pi ** 3)
    H_tt = (k * c[:, None]).T @ k
    assert np.allclose(v["hcross_tt"], H_tt, rtol=1e-4, atol=1e-9 * np.abs(H_tt).max())
    # the cross block under vertex DoFs, (1/8 pi^3) sum_q w_q Re(dA_0 dA_1^H),
    # from the oracle's per-node vertex gradients on every node
    g0 = oracle.geom_mesh(z["V0"], z["F"])
    x0, x1 = z["V0"].ravel(), (z["V0"] + z["shift"]).ravel()
    d0 = np.stack([np.asarray(oracle.compute_A_gradient(g0, k[q], x0, oracle.DOF_MESH)).ravel()
                   for q in range(len(kmag))])
    d1 = np.stack([np.asarray(oracle.compute_A_gradient(g1, k[q], x1, oracle.DOF_MESH)).ravel()
                   for q in range(len(kmag))])
    Hm = ((d0.T * wq) @ np.conj(d1)).real / (8 * np.pi ** 3)
    # dV01/d(vertices of cube 2) = (1/8 pi^3) sum_q w_q Re(dA_1 conj A_0)
    gm = ((d1.T * wq) @ np.conj(z["J"][:, 0])).real / (8 * np.pi ** 3)
    eg = np.linalg.norm(v["gmesh2"] - gm) / np.linalg.norm(gm)
    print(f"{mode}: vertex-DoF pair gradient error {eg:.2e}")
    assert eg < max(tol, 1e-6), eg
    e = np.linalg.norm(v["hcross_mesh"] - Hm) / np.linalg.norm(Hm)
    print(f"{mode}: vertex-DoF cross Hessian error {e:.2e}")
    assert e < max(tol, 1e-6), e
    # single-k A and dA/dtheta of the second cube under an affine map, against
    # the oracle (FP64 on both sides)
    th = np.array([0.05, -0.02, 0.1, 1.1, 0.1, 0.0, -0.05, 0.9, 0.2, 0.0, 0.1, 1.05])
    kk = np.array([1.3, -0.7, 2.1])
    a64 = oracle.compute_A(g1, kk, th, oracle.DOF_AFFINE)
    g64 = np.asarray(oracle.compute_A_gradient(g1, kk, th, oracle.DOF_AFFINE)).ravel()
    assert abs(v["ak"] - a64) <= 1e-12 * abs(a64)
    assert np.max(np.abs(v["gradk"] - g64)) <= 1e-11 * np.max(np.abs(g64))


def test_matrix_driver_size_mismatch_returns_zero():
    """CPU: fewer DoF parameterisations or DoF vectors than geometries: the
    driver reports the mismatch and returns the n x n zero matrix without a
    device call (cpp/src/compute_intersection_volume_multi_object.cu:255-261)."""
    r = subprocess.run([os.path.join(PKG, "example_cube_pair"), LEB29, "mismatch"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert r.stderr.count("Mismatch between number of geometries (2)") == 2
    rows = [ln.split() for ln in r.stdout.splitlines() if ln.startswith("mismatch")]
    assert len(rows) == 2
    for row in rows:
        assert row[1:3] == ["2", "2"]
        assert [float(x) for x in row[3:7]] == [0.0] * 4
        assert row[7] == "0"   # no gradients
