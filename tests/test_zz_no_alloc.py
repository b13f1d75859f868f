"""The allocation-free form-factor entry (pii_form_factor_device_noalloc)
against pii_form_factor_device through the C ABI: the same result bit for bit
on two calls in a row (the second smaller), in both precisions, with no
allocation (the allocation fault-injection hook, armed during the call, is
never reached), and a clear refusal of inputs beyond its fixed arrays."""
import ctypes as C

import numpy as np
import pytest

import pointint_amd as P
from pointint_amd import _native as N
from pointint_amd import workloads as W

pytestmark = pytest.mark.gpu

FAST, FP64 = 1, 0


def _call(fn, g, kind, x, kg, prec, out):
    xd = np.ascontiguousarray(x, dtype=np.float64)
    return fn(g.ctx.handle, g.handle, kind, xd.ctypes.data, xd.size, kg.handle, prec,
              out.data_ptr())


def _case(level, order, nrad, kind, seed):
    rng = np.random.default_rng(seed)
    V, F = W.icosphere(level)
    V = W.star_perturb(V, rng, 0.1) * 0.4
    g = P.make_triangle_mesh(V, F)
    kg = P.kgrid_from_rule(order, nrad)
    if kind == 0:
        x = np.concatenate([rng.uniform(-0.2, 0.2, 3), np.eye(3).ravel() * 1.1])
    else:
        x = (V + rng.normal(scale=1e-3, size=V.shape)).ravel()
    return g, kg, x


@pytest.mark.parametrize("prec", [FAST, FP64])
@pytest.mark.parametrize("kind", [0, 1])
def test_noalloc_equals_device_entry_twice(prec, kind):
    import torch
    lib = N.load()
    # a larger call, then a smaller one: the array keeps the first call's
    # contents, which the second must not read
    for level, order, nrad in ((4, 59, 24), (2, 9, 8)):
        g, kg, x = _case(level, order, nrad, kind, 7 + level)
        ref = torch.zeros(2 * kg.Q, dtype=torch.float64, device="cuda")
        N.check(_call(lib.pii_form_factor_device, g, kind, x, kg, prec, ref))
        out = torch.full((2 * kg.Q,), float("nan"), dtype=torch.float64, device="cuda")
        N.check(lib.pii_debug_fail_allocation(0))   # any allocation now fails
        try:
            rc = _call(lib.pii_form_factor_device_noalloc, g, kind, x, kg, prec, out)
        finally:
            N.check(lib.pii_debug_fail_allocation(-1))
        N.check(rc)
        assert torch.equal(out, ref), (level, prec, kind)


def test_noalloc_refuses_inputs_beyond_its_arrays():
    import torch
    lib = N.load()
    kg = P.kgrid_from_rule(9, 4)
    out = torch.zeros(2 * kg.Q, dtype=torch.float64, device="cuda")
    # vertex DoFs of 163,842 vertices: 3.9 MB, beyond the 1 MiB host staging
    V7, F7 = W.icosphere(7)
    g7 = P.make_triangle_mesh(V7 * 0.4, F7)
    rc = _call(lib.pii_form_factor_device_noalloc, g7, 1, V7.ravel() * 0.4, kg, FAST, out)
    assert rc == N.PII_ERR_INVALID
    assert b"host staging exhausted" in lib.pii_last_error()
    # 1,310,720 triangles: their records alone exceed the 64 MiB device array
    V8, F8 = W.icosphere(8)
    g8 = P.make_triangle_mesh(V8 * 0.4, F8)
    ident = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1.0])
    rc = _call(lib.pii_form_factor_device_noalloc, g8, 0, ident, kg, FAST, out)
    assert rc == N.PII_ERR_INVALID
    assert b"device scratch exhausted" in lib.pii_last_error()
    # and the entry still works afterwards
    g, kg2, x = _case(2, 9, 8, 0, 3)
    ref = torch.zeros(2 * kg2.Q, dtype=torch.float64, device="cuda")
    N.check(_call(lib.pii_form_factor_device, g, 0, x, kg2, FAST, ref))
    out2 = torch.zeros_like(ref)
    N.check(_call(lib.pii_form_factor_device_noalloc, g, 0, x, kg2, FAST, out2))
    assert torch.equal(out2, ref)
