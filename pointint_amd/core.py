"""Python host API over the C ABI: resident k-grids and geometries, form
factors and the all-pairs volume matrix on one MI355X.

Names and argument meanings follow the reference's C++ API
(cpp/include/compute_intersection_volume_multi_object.hpp:51-120,
cpp/include/geometry/packing.hpp:31-45, cpp/include/quadrature/*.hpp).
"""
from __future__ import annotations

import ctypes as C
import os
import threading
from typing import Optional, Sequence

import numpy as np

from . import _native as N

IDENTITY_AFFINE = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1], dtype=np.float64)
DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "lebedev")

_FF = {"fp64": N.FF_FP64, "fast": N.FF_FAST}
_GRAM = {"fp64": N.GRAM_FP64, "bf16x3": N.GRAM_BF16X3, "fp16x3": N.GRAM_FP16X3}
_DOF = {"affine": N.DOF_AFFINE, "trimesh": N.DOF_TRIMESH}


def lebedev_path(order: int) -> str:
    """Path of a bundled Lebedev rule file, e.g. lebedev_path(29) (302 points)."""
    return os.path.join(DATA_DIR, f"lebedev_{order:03d}.txt")


class Context:
    """One HIP device + stream.  All objects created from a context live on its device."""

    _default = {}
    _lock = threading.Lock()

    def __init__(self, device: int = 0):
        lib = N.load()
        h = C.c_void_p()
        N.check(lib.pii_context_create(device, C.byref(h)))
        self._h = h
        self.device = device

    @classmethod
    def default(cls, device: int = 0) -> "Context":
        with cls._lock:
            if device not in cls._default:
                cls._default[device] = Context(device)
            return cls._default[device]

    @property
    def handle(self):
        return self._h

    def synchronize(self) -> None:
        N.check(N.load().pii_context_synchronize(self._h))

    def set_profiling(self, on: bool) -> None:
        N.check(N.load().pii_context_set_profiling(self._h, 1 if on else 0))

    def set_series(self, on: bool) -> None:
        """Moment series for the low-|k| shells of triangle objects on the fast
        path (default on); off gives the per-face evaluation at every node."""
        N.check(N.load().pii_context_set_series(self._h, 1 if on else 0))

    def timings(self) -> dict:
        """Last call's phase durations (ms) from HIP events on the context stream."""
        ms = np.zeros(6)
        N.check(N.load().pii_context_timings(self._h, N.ptr(ms), 6))
        return {"prepass": ms[0], "form_factor": ms[1], "gram": ms[2], "derivatives": ms[3],
                "ff_series": ms[4], "ff_per_face": ms[5]}

    def work(self) -> dict:
        """Work of the last call's form-factor kernels (evaluation counts)."""
        u = np.zeros(5)
        N.check(N.load().pii_context_work(self._h, N.ptr(u), 5))
        return {"tri_node": u[0], "series_face_dir": u[1], "series_node": u[2],
                "surfel_node": u[3], "series_surfel_dir": u[4]}

    def issue_rate(self, kind: str) -> float:
        """Sustained device rate from a probe kernel: "fp64_fma" / "fp32_fma"
        (flop/s) or "sin" (FP32 hardware sines/s)."""
        k = {"fp64_fma": 0, "fp32_fma": 1, "sin": 2}[kind]
        out = C.c_double()
        N.check(N.load().pii_measure_issue_rate(self._h, k, C.byref(out)))
        return out.value

    def close(self) -> None:
        if getattr(self, "_h", None):
            N.load().pii_context_destroy(self._h)
            self._h = None


# ---------------------------------------------------------------------------
# Quadrature (host).
# ---------------------------------------------------------------------------
def gauss_legendre(n: int, a: float, b: float):
    x = np.zeros(n)
    w = np.zeros(n)
    N.check(N.load().pii_gauss_legendre(n, a, b, N.ptr(x), N.ptr(w)))
    return x, w


def load_lebedev(path: str):
    """(dirs (M,3), weights (M,)) with weights summing to 4 pi."""
    lib = N.load()
    cnt = C.c_int()
    N.check(lib.pii_lebedev_load(path.encode(), None, None, 0, C.byref(cnt)))
    M = cnt.value
    d = np.zeros((M, 3))
    w = np.zeros(M)
    N.check(lib.pii_lebedev_load(path.encode(), N.ptr(d), N.ptr(w), M, C.byref(cnt)))
    return d, w


class KGrid:
    """k-grid (Gauss-Legendre in t, k = tan t, x Lebedev directions) resident in HBM."""

    def __init__(self, dirs, kmag, w, ctx: Optional[Context] = None):
        self.ctx = ctx or Context.default()
        self.dirs = N.f64(dirs).reshape(-1, 3)
        self.kmag = N.f64(kmag).reshape(-1)
        self.w = N.f64(w).reshape(-1)
        Q = self.kmag.shape[0]
        if self.dirs.shape[0] != Q or self.w.shape[0] != Q:
            raise ValueError("dirs, kmag and w must have the same length")
        h = C.c_void_p()
        N.check(N.load().pii_kgrid_create(self.ctx.handle, Q, N.ptr(self.dirs), N.ptr(self.kmag),
                                          N.ptr(self.w), C.byref(h)))
        self._h = h
        q, qh, sym = C.c_int(), C.c_int(), C.c_int()
        N.check(N.load().pii_kgrid_info(h, C.byref(q), C.byref(qh), C.byref(sym)))
        self.Q, self.Qh, self.symmetric = q.value, qh.value, bool(sym.value)

    @property
    def handle(self):
        return self._h

    def __len__(self):
        return self.Q

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                N.load().pii_kgrid_destroy(h)
            except Exception:
                pass
            self._h = None


def build_kgrid_arrays(leb_dirs, leb_w, n_radial: int):
    leb_dirs = N.f64(leb_dirs).reshape(-1, 3)
    leb_w = N.f64(leb_w).reshape(-1)
    M = leb_w.shape[0]
    Q = M * n_radial
    dirs = np.zeros((Q, 3))
    kmag = np.zeros(Q)
    w = np.zeros(Q)
    N.check(N.load().pii_kgrid_build(N.ptr(leb_dirs), N.ptr(leb_w), M, n_radial, N.ptr(dirs),
                                     N.ptr(kmag), N.ptr(w)))
    return dirs, kmag, w


def build_kgrid(leb_dirs, leb_w, n_radial: int, ctx: Optional[Context] = None) -> KGrid:
    return KGrid(*build_kgrid_arrays(leb_dirs, leb_w, n_radial), ctx=ctx)


def kgrid_from_rule(order: int, n_radial: int, ctx: Optional[Context] = None) -> KGrid:
    d, w = load_lebedev(lebedev_path(order))
    return build_kgrid(d, w, n_radial, ctx)


# ---------------------------------------------------------------------------
# Geometry.
# ---------------------------------------------------------------------------
class Geometry:
    """Triangle mesh, oriented-disk cloud or Gaussian-surfel set resident in HBM."""

    def __init__(self, handle, ctx: Context, kind: int, host: dict):
        self._h = handle
        self._hv = handle.value or 0  # plain int for the batched calls' pointer arrays
        self.ctx = ctx
        self.type = kind
        self.host = host
        t, n, nv = C.c_int(), C.c_int(), C.c_int()
        N.check(N.load().pii_geometry_info(handle, C.byref(t), C.byref(n), C.byref(nv)))
        self._n = n.value
        self.num_dof_vertices = nv.value

    @property
    def handle(self):
        return self._h

    def num_elements(self) -> int:
        return self._n

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                N.load().pii_geometry_destroy(h)
            except Exception:
                pass
            self._h = None


def cache_moments(geoms: Sequence[Geometry], ctx: Optional[Context] = None) -> None:
    """Expands the low-|k| moment series of each triangle mesh (>= 192 faces)
    in spherical harmonics, once: afterwards the fast forward entries evaluate
    that series at a fixed cost per direction for any pose x' = s R x + t
    (include/pointint_hip.h, pii_geometry_cache_moments).  Other geometries
    are left as they are."""
    geoms = list(geoms)
    if not geoms:
        return
    ctx = ctx or geoms[0].ctx
    arr = (C.c_void_p * len(geoms))(*[g._hv for g in geoms])
    N.check(N.load().pii_geometry_cache_moments(ctx.handle, len(geoms), arr))


def make_triangle_mesh(V, F, ctx: Optional[Context] = None) -> Geometry:
    ctx = ctx or Context.default()
    V = N.f64(V).reshape(-1, 3)
    F = N.i32(F).reshape(-1, 3)
    h = C.c_void_p()
    N.check(N.load().pii_geometry_create_mesh(ctx.handle, N.ptr(V), V.shape[0], N.ptr(F),
                                              F.shape[0], C.byref(h)))
    return Geometry(h, ctx, N.GEOM_TRIANGLE, {"V": V, "F": F})


def _unit_normals(Nrm):
    Nrm = N.f64(Nrm).reshape(-1, 3).copy()
    l = np.linalg.norm(Nrm, axis=1)
    nz = l > 0
    Nrm[nz] /= l[nz, None]
    return Nrm


def make_point_cloud(P, Nrm, radii_or_areas, is_radius: bool = True,
                     ctx: Optional[Context] = None) -> Geometry:
    """Oriented disks (cpp/src/geometry/packing.cpp:30-63): area = pi rho^2, or rho = sqrt(area/pi)."""
    ctx = ctx or Context.default()
    P = N.f64(P).reshape(-1, 3)
    n = P.shape[0]
    r = N.f64(radii_or_areas).reshape(-1)
    if r.shape[0] != n:
        raise ValueError("one radius/area per point required")
    el = np.zeros((n, 8))
    el[:, 0:3] = P
    el[:, 3:6] = _unit_normals(Nrm)
    if is_radius:
        el[:, 6] = r
        el[:, 7] = np.pi * r * r
    else:
        el[:, 7] = r
        el[:, 6] = np.sqrt(r / np.pi)
    h = C.c_void_p()
    N.check(N.load().pii_geometry_create_disks(ctx.handle, N.ptr(el), n, C.byref(h)))
    return Geometry(h, ctx, N.GEOM_DISK, {"el": el})


def make_gaussian_splat(P, Nrm, sigmas, weights, ctx: Optional[Context] = None) -> Geometry:
    """Gaussian surfels (cpp/src/geometry/packing.cpp:84-124)."""
    ctx = ctx or Context.default()
    P = N.f64(P).reshape(-1, 3)
    n = P.shape[0]
    el = np.zeros((n, 8))
    el[:, 0:3] = P
    el[:, 3:6] = _unit_normals(Nrm)
    el[:, 6] = N.f64(sigmas).reshape(-1)
    el[:, 7] = N.f64(weights).reshape(-1)
    h = C.c_void_p()
    N.check(N.load().pii_geometry_create_gaussians(ctx.handle, N.ptr(el), n, C.byref(h)))
    return Geometry(h, ctx, N.GEOM_GAUSSIAN, {"el": el})


def make_packed_triangles(tri, vid=None, ctx: Optional[Context] = None) -> Geometry:
    """Triangles as TriPacked rows {a, e1, e2, S} (12 doubles) plus optional vids."""
    ctx = ctx or Context.default()
    tri = N.f64(tri).reshape(-1, 12)
    nt = tri.shape[0]
    vidp = None
    if vid is not None:
        vid = N.i32(vid).reshape(-1, 3)
        vidp = N.ptr(vid)
    h = C.c_void_p()
    N.check(N.load().pii_geometry_create_tris(ctx.handle, N.ptr(tri), vidp, nt, C.byref(h)))
    return Geometry(h, ctx, N.GEOM_TRIANGLE, {"tri": tri, "vid": vid})


# ---------------------------------------------------------------------------
# Form factors and volumes.
# ---------------------------------------------------------------------------
def _dof_args(geoms, dofs, dof_kinds):
    n = len(geoms)
    if dof_kinds is None:
        dof_kinds = ["affine"] * n
    kinds = [(_DOF[k] if isinstance(k, str) else int(k)) for k in dof_kinds]
    if dofs is None:
        dofs = [IDENTITY_AFFINE] * n
    dofs = [N.f64(d).reshape(-1) for d in dofs]
    if len(kinds) != n or len(dofs) != n:
        raise ValueError("one DoF kind and one DoF vector per geometry")
    return kinds, dofs


def compute_form_factor(geom: Geometry, kgrid: KGrid, dofs=None, dof_kind="affine",
                        precision: str = "fast") -> np.ndarray:
    """A(k_q) for every node of the grid (complex128, length Q)."""
    kinds, dl = _dof_args([geom], None if dofs is None else [dofs], [dof_kind])
    out = np.zeros(2 * kgrid.Q)
    N.check(N.load().pii_form_factor(geom.ctx.handle, geom.handle, kinds[0], N.ptr(dl[0]),
                                     dl[0].shape[0], kgrid.handle, _FF[precision], N.ptr(out)))
    return out[0::2] + 1j * out[1::2]


def compute_form_factor_into(out, geom: Geometry, kgrid: KGrid, dofs=None, dof_kind="affine",
                             precision: str = "fast"):
    """A(k_q) written into `out`, a contiguous float64 CUDA tensor of 2 Q
    elements (interleaved re, im) on the context's device, without a host
    copy (the plug-in functors' contract).  Returns `out`."""
    if not (out.is_cuda and out.is_contiguous() and out.dtype.is_floating_point
            and out.element_size() == 8 and out.numel() == 2 * kgrid.Q):
        raise ValueError("out must be a contiguous float64 CUDA tensor of 2 Q elements")
    kinds, dl = _dof_args([geom], None if dofs is None else [dofs], [dof_kind])
    # the library's writes follow what torch has queued on its current stream
    _set_caller_stream(geom.ctx, out.device)
    N.check(N.load().pii_form_factor_device(geom.ctx.handle, geom.handle, kinds[0],
                                            dl[0].ctypes.data, dl[0].shape[0], kgrid.handle,
                                            _FF[precision], out.data_ptr()))
    return out


def _device_tensor(t, n: int, what: str):
    if not (t.is_cuda and t.is_contiguous() and t.dtype.is_floating_point
            and t.element_size() == 8 and t.numel() == n):
        raise ValueError(f"{what} must be a contiguous float64 CUDA tensor of {n} elements")


def _set_caller_stream(ctx: "Context", device) -> None:
    import torch
    N.check(N.load().pii_context_set_caller_stream(
        ctx.handle, C.c_void_p(torch.cuda.current_stream(device).cuda_stream)))


def compute_volume_matrix_device(V, geoms: Sequence[Geometry], kgrid: KGrid, dofs, dof_kinds,
                                 ff: str = "fast", gram: str = "fp16x3"):
    """V (n x n) into `V`, a contiguous float64 CUDA tensor, from DoF vectors
    that are contiguous float64 CUDA tensors (pii_volume_matrix_device: the
    DoF values stay on the device).  Returns `V`."""
    n = len(geoms)
    _device_tensor(V, n * n, "V")
    kinds = [_DOF[k] for k in dof_kinds]
    for d in dofs:
        _device_tensor(d, d.numel(), "dofs")
    ctx = geoms[0].ctx
    _set_caller_stream(ctx, V.device)
    gh = np.fromiter((g._hv for g in geoms), dtype=np.uintp, count=n)
    kh = np.asarray(kinds, dtype=np.int32)
    dh = np.fromiter((d.data_ptr() for d in dofs), dtype=np.uintp, count=n)
    nh = np.fromiter((d.numel() for d in dofs), dtype=np.int32, count=n)
    N.check(N.load().pii_volume_matrix_device(ctx.handle, n, gh.ctypes.data, kh.ctypes.data,
                                              dh.ctypes.data, nh.ctypes.data, kgrid.handle,
                                              _FF[ff], _GRAM[gram], V.data_ptr()))
    return V


def volume_matrix_vjp_device(grads, geoms: Sequence[Geometry], kgrid: KGrid, dL_dV, dofs,
                             dof_kinds, V=None):
    """Gradients of L = sum_ij dL_dV[i, j] V_ij w.r.t. each object's DoFs into
    `grads` (contiguous float64 CUDA tensors, one per object), with dL_dV, the
    DoFs and the optional `V` also on the device (pii_volume_matrix_vjp_device)."""
    n = len(geoms)
    _device_tensor(dL_dV, n * n, "dL_dV")
    if V is not None:
        _device_tensor(V, n * n, "V")
    kinds = [_DOF[k] for k in dof_kinds]
    for d, g in zip(dofs, grads):
        _device_tensor(d, d.numel(), "dofs")
        _device_tensor(g, d.numel(), "grads")
    ctx = geoms[0].ctx
    _set_caller_stream(ctx, dL_dV.device)
    gh = np.fromiter((g._hv for g in geoms), dtype=np.uintp, count=n)
    kh = np.asarray(kinds, dtype=np.int32)
    dh = np.fromiter((d.data_ptr() for d in dofs), dtype=np.uintp, count=n)
    nh = np.fromiter((d.numel() for d in dofs), dtype=np.int32, count=n)
    oh = np.fromiter((g.data_ptr() for g in grads), dtype=np.uintp, count=n)
    N.check(N.load().pii_volume_matrix_vjp_device(
        ctx.handle, n, gh.ctypes.data, kh.ctypes.data, dh.ctypes.data, nh.ctypes.data,
        kgrid.handle, dL_dV.data_ptr(), oh.ctypes.data,
        V.data_ptr() if V is not None else None))
    return grads


def compute_form_factor_jacobian(geom: Geometry, kgrid: KGrid, dofs=None,
                                 dof_kind="affine") -> np.ndarray:
    """dA(k_q)/dtheta for every node (complex128, Q x ndofs) in FP64 with the
    reference's element derivatives (its gradient plug-ins' output)."""
    kinds, dl = _dof_args([geom], None if dofs is None else [dofs], [dof_kind])
    D = dl[0].shape[0]
    out = np.empty(2 * kgrid.Q * D)
    N.check(N.load().pii_form_factor_jacobian(geom.ctx.handle, geom.handle, kinds[0],
                                              N.ptr(dl[0]), D, kgrid.handle, N.ptr(out)))
    out = out.reshape(kgrid.Q, D, 2)
    return out[..., 0] + 1j * out[..., 1]


def _batch_args(geoms, dofs, dof_kinds):
    """Pointer arrays of a batched call (numpy-built: per-object ctypes work
    would cost more than the GPU step for a few hundred objects).  Returns the
    arrays and the DoF vectors they point into (kept alive by the caller)."""
    n = len(geoms)
    gh = np.fromiter((g._hv for g in geoms), dtype=np.uintp, count=n)
    if dofs is None and dof_kinds is None:
        dl = [IDENTITY_AFFINE]
        kh = np.full(n, _DOF["affine"], dtype=np.int32)
        dh = np.full(n, IDENTITY_AFFINE.ctypes.data, dtype=np.uintp)
        nh = np.full(n, 12, dtype=np.int32)
        return gh, kh, dh, nh, dl
    kinds, dl = _dof_args(geoms, dofs, dof_kinds)
    kh = np.asarray(kinds, dtype=np.int32)
    dh = np.fromiter((d.ctypes.data for d in dl), dtype=np.uintp, count=n)
    nh = np.fromiter((d.shape[0] for d in dl), dtype=np.int32, count=n)
    return gh, kh, dh, nh, dl


def compute_volume_matrix(geoms: Sequence[Geometry], kgrid: KGrid, dofs=None, dof_kinds=None,
                          ff: str = "fast", gram: str = "fp16x3") -> np.ndarray:
    """V (n x n) = Re(J^H D J)/(8 pi^3) on the GPU."""
    n = len(geoms)
    if n == 0:
        return np.zeros((0, 0))
    ctx = geoms[0].ctx
    gh, kh, dh, nh, _keep = _batch_args(geoms, dofs, dof_kinds)
    V = np.empty((n, n))
    N.check(N.load().pii_volume_matrix(ctx.handle, n, gh.ctypes.data, kh.ctypes.data,
                                       dh.ctypes.data, nh.ctypes.data, kgrid.handle, _FF[ff],
                                       _GRAM[gram], V.ctypes.data))
    return V


def volume_matrix_vjp(geoms: Sequence[Geometry], kgrid: KGrid, dL_dV, dofs=None,
                      dof_kinds=None, return_volume: bool = True):
    """Gradients of L = sum_ij dL_dV[i,j] V_ij w.r.t. each object's DoFs
    (list of arrays), and V of the same call when return_volume."""
    n = len(geoms)
    ctx = geoms[0].ctx
    kinds, dl = _dof_args(geoms, dofs, dof_kinds)
    G = N.f64(dL_dV).reshape(n, n)
    grads = [np.empty(d.shape[0]) for d in dl]  # every entry is written by the call
    gh = (C.c_void_p * n)(*[g._hv for g in geoms])
    kh = (C.c_int * n)(*kinds)
    dh = (C.c_void_p * n)(*[d.ctypes.data for d in dl])
    nh = (C.c_int * n)(*[d.shape[0] for d in dl])
    oh = (C.c_void_p * n)(*[g.ctypes.data for g in grads])
    V = np.zeros((n, n)) if return_volume else None
    N.check(N.load().pii_volume_matrix_vjp(ctx.handle, n, gh, kh, dh, nh, kgrid.handle, N.ptr(G),
                                           oh, N.ptr(V) if V is not None else None))
    return grads, V


def volume_matrix_hessian(geoms: Sequence[Geometry], kgrid: KGrid, dofs=None,
                          gram: str = "fp16x3", return_volume: bool = False, dof_kinds=None):
    """Gauss-Newton Hessian blocks of V: H (D x D), D = sum of the DoF
    counts, block (i, j) = (1/8 pi^3) sum_q w_q Re(dA_i/dtheta_i
    (dA_j/dtheta_j)^H) at the rows and columns of objects i and j in order
    (12 per AffineDoF object, 3 per vertex for "trimesh"); with return_volume
    also V of the same call.  dof_kinds defaults to AffineDoF for every
    object."""
    n = len(geoms)
    ctx = geoms[0].ctx
    if dof_kinds is None and dofs is not None:
        dof_kinds = ["affine"] * n
    gh, kh, dh, nh, _keep = _batch_args(geoms, dofs, dof_kinds)
    D = int(nh.sum())
    H = np.empty((D, D))
    V = np.empty((n, n)) if return_volume else None
    N.check(N.load().pii_volume_matrix_hessian(ctx.handle, n, gh.ctypes.data, kh.ctypes.data,
                                               dh.ctypes.data, nh.ctypes.data, kgrid.handle,
                                               _GRAM[gram], H.ctypes.data,
                                               V.ctypes.data if V is not None else None))
    return (H, V) if return_volume else H


def volume_matrix_derivatives(geoms: Sequence[Geometry], kgrid: KGrid, dofs=None,
                              gram: str = "fp16x3", hessian: bool = True, dof_kinds=None):
    """V (n x n), grad and, with hessian, the Gauss-Newton H (D x D, D = sum
    of the DoF counts) from one pass.  grad[i][j] = dV_ij/dtheta_i =
    (1/8 pi^3) sum_q w_q Re(dA_i/dtheta conj A_j), half the full derivative
    for i == j (the matrix driver's grad_matrix convention): an (n, n, 12)
    array for AffineDoF objects (dof_kinds None), else a list of (n, D_i)
    arrays, one per object."""
    n = len(geoms)
    ctx = geoms[0].ctx
    if dof_kinds is not None:
        gh, kh, dh, nh, _keep = _batch_args(geoms, dofs, dof_kinds)
        sizes = [int(x) for x in nh]
        D = sum(sizes)
        V = np.empty((n, n))
        Gf = np.empty(n * D)
        H = np.empty((D, D)) if hessian else None
        N.check(N.load().pii_volume_matrix_derivatives(
            ctx.handle, n, gh.ctypes.data, kh.ctypes.data, dh.ctypes.data, nh.ctypes.data,
            kgrid.handle, _GRAM[gram], V.ctypes.data, Gf.ctypes.data,
            H.ctypes.data if H is not None else None))
        offs = np.cumsum([0] + sizes)
        G = [Gf[n * offs[i]:n * offs[i + 1]].reshape(n, sizes[i]) for i in range(n)]
        return (V, G, H) if hessian else (V, G)
    gh, _kh, dh, _nh, _keep = _batch_args(geoms, dofs, None if dofs is None else ["affine"] * n)
    V = np.empty((n, n))
    G = np.empty((n, n, 12))
    H = np.empty((12 * n, 12 * n)) if hessian else None
    N.check(N.load().pii_volume_matrix_affine_derivatives(
        ctx.handle, n, gh.ctypes.data, dh.ctypes.data, kgrid.handle, _GRAM[gram], V.ctypes.data,
        G.ctypes.data, H.ctypes.data if H is not None else None))
    return (V, G, H) if hessian else (V, G)


def gram_from_J(J: np.ndarray, kgrid: KGrid, ctx: Optional[Context] = None) -> np.ndarray:
    """FP64 Gram of caller form factors J (Q x n complex)."""
    ctx = ctx or kgrid.ctx
    J = np.asarray(J, dtype=np.complex128)
    n = J.shape[1]
    Jc = np.ascontiguousarray(J.T)  # object-major
    V = np.zeros((n, n))
    N.check(N.load().pii_gram(ctx.handle, N.ptr(Jc.view(np.float64)), n, kgrid.handle,
                              N.GRAM_FP64, N.ptr(V)))
    return V
