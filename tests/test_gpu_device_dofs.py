"""Device-resident DoFs and outputs (pii_volume_matrix_device,
pii_volume_matrix_vjp_device) and the autograd function's device path:
results equal the host-pointer entries' on the same DoF values -- V bit for
bit (the plan reads a device-computed bounding box that rounds as the host
one), gradients to rounding (the per-vertex sums are FP64 atomics, whose
order varies from call to call)."""
import ctypes as C

import numpy as np
import pytest

import pointint_amd as P
from pointint_amd import _native as N
from pointint_amd import workloads as W

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def scene():
    rng = np.random.default_rng(17)
    Vi, Fi = W.icosphere(2)
    verts = [Vi * rng.uniform(0.3, 0.45) + rng.uniform(0.0, 0.6, size=3) for _ in range(4)]
    kg = P.kgrid_from_rule(29, 24)
    geoms = [P.make_triangle_mesh(v, Fi) for v in verts]
    return verts, Fi, kg, geoms


def close(a, b):
    return np.allclose(a, b, rtol=1e-12, atol=1e-14 * np.abs(b).max())


def dev(x):
    return torch.tensor(np.ascontiguousarray(x), dtype=torch.float64, device="cuda")


@pytest.mark.parametrize("ff,gram", [("fast", "fp16x3"), ("fast", "bf16x3"), ("fast", "fp64"),
                                     ("fp64", "fp64")])
def test_volume_matrix_device_equals_host(scene, ff, gram):
    verts, Fi, kg, geoms = scene
    xs = [v.reshape(-1) for v in verts]
    kinds = ["trimesh"] * 4
    Vh = P.compute_volume_matrix(geoms, kg, xs, kinds, ff=ff, gram=gram)
    Vd = torch.empty((4, 4), dtype=torch.float64, device="cuda")
    P.compute_volume_matrix_device(Vd, geoms, kg, [dev(x) for x in xs], kinds, ff=ff, gram=gram)
    assert np.array_equal(Vd.cpu().numpy(), Vh)


def test_affine_and_disks_device_equals_host(scene):
    verts, Fi, kg, geoms = scene
    rng = np.random.default_rng(3)
    Pd, Nd, _ = W.sphere_pointcloud(400)
    disks = P.make_point_cloud(Pd * 0.35 + 0.3, Nd, np.full(400, 0.02))
    gs = geoms[:2] + [disks]
    R = W.random_rotation(rng)
    dofs = [np.concatenate([[0.05, 0.0, -0.1], (1.1 * R).ravel()]),
            np.array([0.1, 0.2, 0.0, 1.0, 0.1, 0, 0, 0.9, 0, 0, 0.05, 1.2]),
            P.IDENTITY_AFFINE.copy()]
    kinds = ["affine"] * 3
    Vh = P.compute_volume_matrix(gs, kg, dofs, kinds, ff="fast", gram="fp16x3")
    Vd = torch.empty((3, 3), dtype=torch.float64, device="cuda")
    P.compute_volume_matrix_device(Vd, gs, kg, [dev(d) for d in dofs], kinds, ff="fast",
                                   gram="fp16x3")
    assert np.array_equal(Vd.cpu().numpy(), Vh)
    # affine VJP (triangles): reduction order differs from the host's sums
    G = np.triu(np.ones((2, 2)))
    gh, _ = P.volume_matrix_vjp(gs[:2], kg, G, dofs[:2], kinds[:2])
    gd = [torch.empty(12, dtype=torch.float64, device="cuda") for _ in range(2)]
    P.volume_matrix_vjp_device(gd, gs[:2], kg, dev(G), [dev(d) for d in dofs[:2]], kinds[:2])
    for a, b in zip(gd, gh):
        assert close(a.cpu().numpy(), b)


def test_vjp_device_equals_host(scene):
    verts, Fi, kg, geoms = scene
    xs = [v.reshape(-1) for v in verts]
    kinds = ["trimesh"] * 4
    rng = np.random.default_rng(5)
    G = rng.standard_normal((4, 4))
    gh, Vh = P.volume_matrix_vjp(geoms, kg, G, xs, kinds)
    gd = [torch.empty(len(x), dtype=torch.float64, device="cuda") for x in xs]
    Vd = torch.empty((4, 4), dtype=torch.float64, device="cuda")
    P.volume_matrix_vjp_device(gd, geoms, kg, dev(G), [dev(x) for x in xs], kinds, V=Vd)
    assert np.array_equal(Vd.cpu().numpy(), Vh)
    for a, b in zip(gd, gh):
        assert close(a.cpu().numpy(), b)
    # a packed-triangle mesh whose DoF vertices are shared ids: summed on the host
    tri = np.concatenate([verts[0][Fi[:, 0]], verts[0][Fi[:, 1]] - verts[0][Fi[:, 0]],
                          verts[0][Fi[:, 2]] - verts[0][Fi[:, 0]],
                          0.5 * np.cross(verts[0][Fi[:, 1]] - verts[0][Fi[:, 0]],
                                         verts[0][Fi[:, 2]] - verts[0][Fi[:, 0]])], axis=1)
    gp = P.make_packed_triangles(tri, Fi)
    two = [gp, geoms[1]]
    G2 = np.triu(np.ones((2, 2)), 1)
    gh2, _ = P.volume_matrix_vjp(two, kg, G2, xs[:2], kinds[:2])
    gd2 = [torch.empty(len(x), dtype=torch.float64, device="cuda") for x in xs[:2]]
    P.volume_matrix_vjp_device(gd2, two, kg, dev(G2), [dev(x) for x in xs[:2]], kinds[:2])
    for a, b in zip(gd2, gh2):
        assert close(a.cpu().numpy(), b)


def test_autograd_device_path_equals_host_path(scene):
    from pointint_amd.autograd import intersection_volume_matrix, pairwise_overlap_loss
    verts, Fi, kg, geoms = scene
    grads = {}
    for device in ("cuda", "cpu"):
        tv = [torch.tensor(v, dtype=torch.float32, device=device, requires_grad=True)
              for v in verts]
        V = intersection_volume_matrix(geoms, kg, tv)
        assert V.device.type == device
        pairwise_overlap_loss(V).backward()
        grads[device] = [t.grad.detach().cpu().numpy() for t in tv]
        assert all(t.grad.dtype == torch.float32 for t in tv)
    for a, b in zip(grads["cuda"], grads["cpu"]):
        assert np.allclose(a, b, rtol=1e-6, atol=1e-7 * np.abs(b).max())


def test_host_pointer_refused(scene):
    verts, Fi, kg, geoms = scene
    x = np.ascontiguousarray(verts[0].reshape(-1))
    V = np.zeros(1)
    gh = (C.c_void_p * 1)(geoms[0]._hv)
    kh = (C.c_int * 1)(N.DOF_TRIMESH)
    dh = (C.c_void_p * 1)(x.ctypes.data)
    nh = (C.c_int * 1)(x.shape[0])
    rc = N.load().pii_volume_matrix_device(geoms[0].ctx.handle, 1, gh, kh, dh, nh, kg.handle,
                                           N.FF_FAST, N.GRAM_FP64, V.ctypes.data)
    assert rc == N.PII_ERR_INVALID
    assert b"not device memory" in N.load().pii_last_error()
