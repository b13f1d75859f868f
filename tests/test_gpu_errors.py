"""Error paths of the matrix entries on the GPU, forced from the host: a DoF
vector of the wrong size, an allocation failure (the library's testing hook
makes one allocation fail as an out-of-memory would), the column limit of the
one-pass affine derivatives, and caller k-grids with near-equal coordinates.
Each failure returns its status code and leaves zero results; the context
works normally afterwards."""
import ctypes as C

import numpy as np
import pytest

import pointint_amd as P
from pointint_amd import _native as N
from pointint_amd import workloads as W

pytestmark = pytest.mark.gpu


def _mats(geoms, dofs, kinds):
    n = len(geoms)
    gh = (C.c_void_p * n)(*[g.handle.value for g in geoms])
    kd = (C.c_int * n)(*[{"affine": 0, "trimesh": 1}[k] for k in kinds])
    arrs = [np.ascontiguousarray(d, dtype=np.float64) for d in dofs]
    dp = (C.c_void_p * n)(*[a.ctypes.data for a in arrs])
    nd = (C.c_int * n)(*[a.size for a in arrs])
    return gh, kd, dp, nd, arrs


def test_wrong_dof_count_gives_invalid_and_zero_matrix():
    V, F = W.icosphere(1)
    geoms = [P.make_triangle_mesh(V * 0.4, F), P.make_triangle_mesh(V * 0.3 + 0.1, F)]
    kg = P.kgrid_from_rule(9, 6)
    ident = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1.0])
    lib = N.load()
    for bad, kinds in (([ident, ident[:11]], ["affine", "affine"]),
                       ([ident, (V * 0.3).ravel()[:-3]], ["affine", "trimesh"])):
        gh, kd, dp, nd, keep = _mats(geoms, bad, kinds)
        out = np.full((2, 2), np.nan)
        rc = lib.pii_volume_matrix(geoms[0].ctx.handle, 2, gh, kd, dp, nd, kg.handle, 1, 1,
                                   out.ctypes.data)
        assert rc == N.PII_ERR_INVALID
        assert b"DoF" in lib.pii_last_error()
        assert np.array_equal(out, np.zeros((2, 2)))
    # the context still works
    Vm = P.compute_volume_matrix(geoms, kg, ff="fast", gram="bf16x3")
    assert np.all(np.diag(Vm) > 0)


def test_injected_allocation_failure_gives_nomem_and_zero_results():
    ctx = P.Context(0)        # a fresh context: its workspaces are not grown yet
    V, F = W.icosphere(2)
    kg = P.kgrid_from_rule(9, 8, ctx)
    geoms = [P.make_triangle_mesh(V * 0.4, F, ctx), P.make_triangle_mesh(V * 0.35 + 0.2, F, ctx)]
    # the reference result on the default context (its own workspaces)
    ref = P.compute_volume_matrix([P.make_triangle_mesh(V * 0.4, F),
                                   P.make_triangle_mesh(V * 0.35 + 0.2, F)],
                                  P.kgrid_from_rule(9, 8), ff="fast", gram="fp64")
    lib = N.load()
    ident = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1.0])
    gh, kd, dp, nd, keep = _mats(geoms, [ident, ident], ["affine", "affine"])
    try:
        N.check(lib.pii_debug_fail_allocation(0))
        out = np.full((2, 2), np.nan)
        rc = lib.pii_volume_matrix(ctx.handle, 2, gh, kd, dp, nd, kg.handle, 1, 0,
                                   out.ctypes.data)
        assert rc == N.PII_ERR_NOMEM, lib.pii_last_error()
        assert b"injected" in lib.pii_last_error()
        assert np.array_equal(out, np.zeros((2, 2)))
        # the second allocation of a call fails
        ctx2 = P.Context(0)
        kg2 = P.kgrid_from_rule(9, 8, ctx2)
        g2 = [P.make_triangle_mesh(V * 0.4, F, ctx2), P.make_triangle_mesh(V * 0.35 + 0.2, F, ctx2)]
        gh2, _, _, _, _ = _mats(g2, [ident, ident], ["affine", "affine"])
        N.check(lib.pii_debug_fail_allocation(1))
        rc = lib.pii_volume_matrix(ctx2.handle, 2, gh2, kd, dp, nd, kg2.handle, 1, 0,
                                   out.ctypes.data)
        assert rc == N.PII_ERR_NOMEM
        assert np.array_equal(out, np.zeros((2, 2)))
    finally:
        N.check(lib.pii_debug_fail_allocation(-1))
    # both contexts recover: the next calls allocate and match the reference
    for c, g, k in ((ctx, geoms, kg), (ctx2, g2, kg2)):
        V2 = P.compute_volume_matrix(g, k, ff="fast", gram="fp64")
        assert np.array_equal(V2, ref)


def test_affine_derivatives_column_limit():
    """13 n columns above 16,384 (n = 1,261 objects): PII_ERR_INVALID before
    any allocation, zero V."""
    Vt = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]]) * 0.1
    Ft = np.array([[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]])
    g = P.make_triangle_mesh(Vt, Ft)
    kg = P.kgrid_from_rule(3, 2)
    n = 1261
    ident = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1.0])
    gh = (C.c_void_p * n)(*([g.handle.value] * n))
    dp = (C.c_void_p * n)(*([ident.ctypes.data] * n))
    out = np.full((n, n), np.nan)
    lib = N.load()
    rc = lib.pii_volume_matrix_affine_derivatives(g.ctx.handle, n, gh, dp, kg.handle, 0,
                                                  out.ctypes.data, None, None)
    assert rc == N.PII_ERR_INVALID
    assert b"limit" in lib.pii_last_error()
    assert not np.any(out)


def test_caller_grid_with_near_equal_coordinates(oracle):
    """Directions perturbed by up to 1e-13 (ties within the antipode tolerance):
    the grid is still recognised as centrally symmetric, and the fast form
    factors on it match the oracle's."""
    d, w = oracle.load_lebedev(P.lebedev_path(9))
    dirs, kmag, wq = oracle.build_kgrid(d, w, 6)
    rng = np.random.default_rng(3)
    dirs = dirs + rng.uniform(-1e-13, 1e-13, size=dirs.shape)
    kg = P.KGrid(dirs, kmag, wq)
    assert kg.symmetric and 2 * kg.Qh == kg.Q
    Vi, Fi = W.icosphere(2)
    g = P.make_triangle_mesh(Vi * 0.4 + 0.1, Fi)
    A = P.compute_form_factor(g, kg, precision="fast")
    ref = oracle.compute_A_grid(oracle.geom_mesh(Vi * 0.4 + 0.1, Fi), dirs, kmag)
    err = np.sqrt((wq * np.abs(A - ref) ** 2).sum() / (wq * np.abs(ref) ** 2).sum())
    assert err < 1e-6
