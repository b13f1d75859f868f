"""ctypes front end of the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
cpu_baseline leg of bench.py, never by the product package.  The functions
restate the reference's CPU algorithm; see pii_oracle.cpp for the file:line
each one follows.

Geometry is passed as a small dict built by :func:`geom_mesh`,
:func:`geom_disks` or :func:`geom_gaussians`:
  {"type": 0|1|2, "el": float64[n, 12|8], "vid": int32[n, 3] | None}
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

GT_TRI, GT_DISK, GT_GAUSS = 0, 1, 2
DOF_AFFINE, DOF_MESH = 0, 1
IDENTITY_DOFS = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1], dtype=np.float64)
SCALE = 1.0 / (8.0 * np.pi ** 3)

_d = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_vp = C.c_void_p


def build():
    """Compile liboracle.so from pii_oracle.cpp (g++ via the Makefile)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_get_threads.restype = C.c_int
        L.orc_gauss_legendre.argtypes = [C.c_int, C.c_double, C.c_double, _d, _d]
        L.orc_load_lebedev.argtypes = [C.c_char_p, _vp, _vp, C.c_int]
        L.orc_load_lebedev.restype = C.c_int
        L.orc_build_kgrid.argtypes = [_d, _d, C.c_int, C.c_int, _d, _d, _d]
        L.orc_pack_tris.argtypes = [_d, _i, C.c_int, _d, _i]
        L.orc_pack_disks.argtypes = [_d, _d, _d, C.c_int, C.c_int, _d]
        L.orc_pack_gaussians.argtypes = [_d, _d, _d, _d, C.c_int, _d]
        L.orc_E.argtypes = [C.c_double, _d]
        L.orc_Ep.argtypes = [C.c_double, _d]
        L.orc_Epp.argtypes = [C.c_double, _d]
        L.orc_Phi.argtypes = [C.c_double, C.c_double, _d]
        L.orc_Phi_grad.argtypes = [C.c_double, C.c_double, _d]
        L.orc_J1_over_x.argtypes = [C.c_double]
        L.orc_J1_over_x.restype = C.c_double
        L.orc_J1_over_x_prime.argtypes = [C.c_double]
        L.orc_J1_over_x_prime.restype = C.c_double
        L.orc_compute_A.argtypes = [C.c_int, C.c_int, _d, _vp, C.c_int, _d, _d, _d]
        L.orc_compute_A_grid.argtypes = [C.c_int, C.c_int, _d, _vp, C.c_int, _d, C.c_int,
                                         _d, _d, _d]
        L.orc_compute_A_gradient.argtypes = [C.c_int, C.c_int, _d, _vp, C.c_int, _d, C.c_int,
                                             _d, _d]
        L.orc_gram.argtypes = [_d, C.c_int, C.c_int, _d, _d]
        L.orc_vjp.argtypes = [C.c_int, C.c_int, _d, _vp, C.c_int, _d, C.c_int, C.c_int,
                              _d, _d, _d, _d, _d]
        L.orc_compute_volume.argtypes = [C.c_int, C.c_int, _d, _vp, C.c_int, _d]
        L.orc_compute_volume.restype = C.c_double
        L.orc_compute_volume_gradient_mesh.argtypes = [C.c_int, _i, _d, C.c_int, _d]
        L.orc_exact_cube_ff.argtypes = [C.c_double] * 3
        L.orc_exact_cube_ff.restype = C.c_double
        L.orc_exact_sphere_ff.argtypes = [C.c_double, C.c_double]
        L.orc_exact_sphere_ff.restype = C.c_double
        L.orc_box_box.argtypes = [_d, _d, _d, _d]
        L.orc_box_box.restype = C.c_double
        L.orc_sphere_sphere.argtypes = [_d, C.c_double, _d, C.c_double]
        L.orc_sphere_sphere.restype = C.c_double
        L.orc_unit_cube.argtypes = [_d, _i]
        L.orc_sphere_pointcloud.argtypes = [C.c_int, _d, _d, _d]
        _lib = L
    return _lib


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def threads() -> int:
    return lib().orc_get_threads()


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# -- quadrature -------------------------------------------------------------
def gauss_legendre(n, a, b):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().orc_gauss_legendre(n, a, b, x, w)
    return x, w


def load_lebedev(path):
    M = lib().orc_load_lebedev(path.encode(), None, None, 0)
    if M < 0:
        raise FileNotFoundError(path)
    d = np.zeros((M, 3))
    w = np.zeros(M)
    lib().orc_load_lebedev(path.encode(), d.ctypes.data, w.ctypes.data, M)
    return d, w


def build_kgrid(ldirs, lw, nrad):
    ldirs = _f64(ldirs)
    lw = _f64(lw)
    M = lw.shape[0]
    Q = M * nrad
    dirs = np.zeros((Q, 3))
    kmag = np.zeros(Q)
    w = np.zeros(Q)
    lib().orc_build_kgrid(ldirs, lw, M, nrad, dirs, kmag, w)
    return dirs, kmag, w


# -- geometry ---------------------------------------------------------------
def geom_mesh(V, F):
    V = _f64(V)
    F = _i32(F)
    nf = F.shape[0]
    tri = np.zeros((nf, 12))
    vid = np.zeros((nf, 3), dtype=np.int32)
    lib().orc_pack_tris(V, F, nf, tri, vid)
    return {"type": GT_TRI, "el": tri, "vid": vid}


def geom_disks(P, N, r, is_radius=True):
    P = _f64(P)
    N = _f64(N)
    r = _f64(r)
    n = P.shape[0]
    out = np.zeros((n, 8))
    lib().orc_pack_disks(P, N, r, n, 1 if is_radius else 0, out)
    return {"type": GT_DISK, "el": out, "vid": None}


def geom_gaussians(P, N, sig, w):
    P, N, sig, w = _f64(P), _f64(N), _f64(sig), _f64(w)
    n = P.shape[0]
    out = np.zeros((n, 8))
    lib().orc_pack_gaussians(P, N, sig, w, n, out)
    return {"type": GT_GAUSS, "el": out, "vid": None}


def _vidp(g):
    return None if g["vid"] is None else g["vid"].ctypes.data


def compute_A(g, k, dofs=IDENTITY_DOFS, dof_kind=DOF_AFFINE):
    out = np.zeros(2)
    lib().orc_compute_A(g["type"], g["el"].shape[0], g["el"], _vidp(g), dof_kind, _f64(dofs),
                        _f64(k), out)
    return complex(out[0], out[1])


def compute_A_grid(g, dirs, kmag, dofs=IDENTITY_DOFS, dof_kind=DOF_AFFINE):
    dirs = _f64(dirs)
    kmag = _f64(kmag)
    Q = kmag.shape[0]
    out = np.zeros(2 * Q)
    lib().orc_compute_A_grid(g["type"], g["el"].shape[0], g["el"], _vidp(g), dof_kind,
                             _f64(dofs), Q, dirs, kmag, out)
    return out[0::2] + 1j * out[1::2]


def compute_A_gradient(g, k, dofs, dof_kind):
    dofs = _f64(dofs)
    nd = dofs.shape[0]
    out = np.zeros(2 * nd)
    lib().orc_compute_A_gradient(g["type"], g["el"].shape[0], g["el"], _vidp(g), dof_kind,
                                 dofs, nd, _f64(k), out)
    return out[0::2] + 1j * out[1::2]


def gram(J, w):
    """J: complex (Q, N). Returns V (N, N) = Re(J^H D J)/(8 pi^3)."""
    J = np.ascontiguousarray(J, dtype=np.complex128)
    Q, N = J.shape
    V = np.zeros((N, N))
    lib().orc_gram(J.view(np.float64).reshape(-1), Q, N, _f64(w), V)
    return V


def vjp(g, dofs, dof_kind, dirs, kmag, w, B):
    dofs = _f64(dofs)
    nd = dofs.shape[0]
    B = np.ascontiguousarray(B, dtype=np.complex128)
    out = np.zeros(nd)
    lib().orc_vjp(g["type"], g["el"].shape[0], g["el"], _vidp(g), dof_kind, dofs, nd,
                  kmag.shape[0], _f64(dirs), _f64(kmag), _f64(w),
                  B.view(np.float64).reshape(-1), out)
    return out


def compute_volume(g, dofs=IDENTITY_DOFS, dof_kind=DOF_AFFINE):
    return lib().orc_compute_volume(g["type"], g["el"].shape[0], g["el"], _vidp(g), dof_kind,
                                    _f64(dofs))


def compute_volume_gradient_mesh(g, V):
    V = _f64(V).reshape(-1)
    out = np.zeros_like(V)
    lib().orc_compute_volume_gradient_mesh(g["vid"].shape[0], g["vid"], V, V.shape[0], out)
    return out


def volume_matrix(geoms, dirs, kmag, w, dofs_list=None, kinds=None):
    """compute_intersection_volume_matrix_cpu: J per object, then the Gram."""
    N = len(geoms)
    J = np.zeros((kmag.shape[0], N), dtype=np.complex128)
    for i, g in enumerate(geoms):
        d = IDENTITY_DOFS if dofs_list is None else dofs_list[i]
        kd = DOF_AFFINE if kinds is None else kinds[i]
        J[:, i] = compute_A_grid(g, dirs, kmag, d, kd)
    return gram(J, w), J


# -- element functions / analytic pins ---------------------------------------
def E(z):
    o = np.zeros(2); lib().orc_E(z, o); return complex(*o)


def Ep(z):
    o = np.zeros(2); lib().orc_Ep(z, o); return complex(*o)


def Epp(z):
    o = np.zeros(2); lib().orc_Epp(z, o); return complex(*o)


def Phi(a, b):
    o = np.zeros(2); lib().orc_Phi(a, b, o); return complex(*o)


def Phi_grad(a, b):
    o = np.zeros(4); lib().orc_Phi_grad(a, b, o)
    return complex(o[0], o[1]), complex(o[2], o[3])


def J1_over_x(x):
    return lib().orc_J1_over_x(x)


def J1_over_x_prime(x):
    return lib().orc_J1_over_x_prime(x)


def exact_cube_ff(k):
    return lib().orc_exact_cube_ff(float(k[0]), float(k[1]), float(k[2]))


def exact_sphere_ff(k, R=1.0):
    return lib().orc_exact_sphere_ff(float(k), float(R))


def box_box(c1, h1, c2, h2):
    return lib().orc_box_box(_f64(c1), _f64(h1), _f64(c2), _f64(h2))


def sphere_sphere(c1, r1, c2, r2):
    return lib().orc_sphere_sphere(_f64(c1), float(r1), _f64(c2), float(r2))


def unit_cube():
    V = np.zeros((8, 3))
    F = np.zeros((12, 3), dtype=np.int32)
    lib().orc_unit_cube(V, F)
    return V, F


def sphere_pointcloud(n=1000):
    P = np.zeros((n, 3)); N = np.zeros((n, 3)); r = np.zeros(n)
    lib().orc_sphere_pointcloud(n, P, N, r)
    return P, N, r
