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
                 "PoIntInt::compute_Ak_gradient_cuda",
                 "PoIntInt::CudaKernelRegistry::register_kernels",
                 "PoIntInt::CudaKernelRegistry::has_kernels",
                 "PoIntInt::CudaKernelRegistry::get_kernels",
                 "PoIntInt::CudaKernelRegistry::make_key", "PoIntInt::get_dof_type_name",
                 "PoIntInt::AffineDoF::compute_volume_gradient",
                 "PoIntInt::TriangleMeshDoF::compute_A_gradient"):
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


def _registry_run():
    r = subprocess.run([os.path.join(PKG, "example_cube_pair"), LEB29, "registry"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    out = {}
    for line in r.stdout.splitlines():
        parts = line.split()
        out[parts[0]] = parts[1:]
    return out, r.stderr


@pytest.mark.gpu
def test_cxx_kernel_registry_and_dof_methods(oracle):
    """CudaKernelRegistry (cpp/include/dof/cuda/dof_cuda_interface.hpp:48-67):
    the built-in entries, a caller's DoF kind x' = s x + t with plug-ins built
    on the AffineDoF ones against the same pair under the equivalent AffineDoF
    (V, grad_matrix and Gauss-Newton blocks by the chain rule), a replaced
    built-in entry, and the DoF methods of dof_parameterization.hpp:17-38
    with compute_volume_cpu (compute_volume.hpp:31-37)."""
    v, err = _registry_run()
    assert v["has"] == ["1", "1", "1", "1", "0", "0"]
    assert v["name"] == ["AffineDoF", "TriangleMeshDoF", "Unknown"]
    # no plug-ins for the caller's type yet: message and zero result
    assert float(v["unregistered"][0]) == 0.0
    assert "CUDA kernels not available for DoF type 'Unknown'" in err
    # V: the plug-in path (FP64 form factors, FP64 Gram) and the built-in
    # AffineDoF path at FP64 precision
    Vc = np.array([float(x) for x in v["Vc"]])
    Va = np.array([float(x) for x in v["Va"]])
    assert np.allclose(Vc, Va, rtol=1e-12, atol=0)
    z = np.load(os.path.join(ROOT, "tests", "golden", "cube_pair_L32_M302.npz"))
    d, w = oracle.load_lebedev(LEB29)
    dirs, kmag, wq = oracle.build_kgrid(d, w, 32)
    g1 = oracle.geom_mesh(z["V0"] + z["shift"], z["F"])
    x2 = np.array([0.3, 0.2, 0.1, 1.1, 0, 0, 0, 1.1, 0, 0, 0, 1.1])
    # chain rule: theta_affine = M theta, M 12 x 4
    M = np.zeros((12, 4))
    M[0, 0] = M[1, 1] = M[2, 2] = 1.0
    M[3, 3] = M[7, 3] = M[11, 3] = 1.0
    # dV01/dtheta_1 of the plug-in path against the oracle's VJP (FP64 both),
    # and against the built-in one-pass affine derivatives (fast form factors)
    gc = np.array([float(x) for x in v["gc"]])
    ga = np.array([float(x) for x in v["ga"]])
    ref = M.T @ oracle.vjp(g1, x2, oracle.DOF_AFFINE, dirs, kmag, wq, z["J"][:, 0])
    assert np.linalg.norm(gc - ref) <= 1e-10 * np.linalg.norm(ref)
    assert np.linalg.norm(gc - M.T @ ga) <= 1e-5 * np.linalg.norm(gc)
    hc = np.array([float(x) for x in v["hc"]]).reshape(4, 4)
    ha = np.array([float(x) for x in v["ha"]]).reshape(12, 12)
    assert np.linalg.norm(hc - M.T @ ha @ M) <= 1e-5 * np.linalg.norm(hc)
    # a replaced AffineDoF_Triangle entry runs (once per object), then is restored
    assert int(v["replaced"][0]) == 2
    assert abs(float(v["replaced"][1]) - z["Vmat"][0, 1]) < 1e-12
    # single-k A and dA/dtheta: methods, drivers and the oracle agree
    ak = np.array([float(x) for x in v["ak"]])
    a1, a2, a3, a4 = (ak[0::2] + 1j * ak[1::2])
    assert a1 == a2 and abs(a3 - a4) <= 1e-14 * abs(a3)
    th = np.array([0.05, -0.02, 0.1, 1.1, 0.1, 0.0, -0.05, 0.9, 0.2, 0.0, 0.1, 1.05])
    kk = np.array([1.3, -0.7, 2.1])
    assert abs(a1 - oracle.compute_A(g1, kk, th, oracle.DOF_AFFINE)) <= 1e-12 * abs(a1)
    assert abs(a3 - oracle.compute_A(g1, kk, x2, oracle.DOF_AFFINE)) <= 1e-12 * abs(a3)
    gk = np.array([float(x) for x in v["gk"]]).reshape(4, 4)
    g3, g4 = gk[:, 0] + 1j * gk[:, 1], gk[:, 2] + 1j * gk[:, 3]
    g64 = M.T @ np.asarray(oracle.compute_A_gradient(g1, kk, x2, oracle.DOF_AFFINE)).ravel()
    assert np.max(np.abs(g3 - g4)) <= 1e-13 * np.max(np.abs(g3))
    assert np.max(np.abs(g3 - g64)) <= 1e-11 * np.max(np.abs(g64))
    # volumes: |det A| for the unit cube under th; the shifted unit cube under
    # its own vertices; gradients against central differences
    A = th[3:].reshape(3, 3)
    va, vm = (float(x) for x in v["vol"])
    assert abs(va - abs(np.linalg.det(A))) < 1e-14 and abs(vm - 1.0) < 1e-14
    vga = np.array([float(x) for x in v["vga"]])

    def vol_aff(t):
        return abs(np.linalg.det(t[3:].reshape(3, 3)))

    def vol_mesh(x):
        P = x.reshape(-1, 3)[z["F"]]
        return np.einsum("ij,ij->i", P[:, 0], np.cross(P[:, 1], P[:, 2])).sum() / 6.0

    def cdiff(f, x, h=1e-5):
        g = np.zeros_like(x)
        for i in range(len(x)):
            e = np.zeros_like(x)
            e[i] = h
            g[i] = (f(x + e) - f(x - e)) / (2 * h)
        return g

    assert np.allclose(vga, cdiff(vol_aff, th), atol=1e-9)
    x_mesh = (z["V0"] + z["shift"]).ravel()
    vgm = np.array([float(x) for x in v["vgm"]])
    assert np.allclose(vgm, cdiff(vol_mesh, x_mesh), atol=1e-9)
