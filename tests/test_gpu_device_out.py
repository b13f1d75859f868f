"""Device-output entries (`pii_form_factor_device`,
`pii_form_factor_jacobian_device`): the plug-in contract of the reference's
CudaComputeAkFunc / CudaComputeAkGradientFunc (write into a caller-owned
device buffer; cpp/include/dof/cuda/dof_cuda_interface.hpp:24-40).  The
device buffer is a torch tensor on the same GPU; its contents must equal the
host-output entries' (bit for bit for A; to FP64 rounding for the vertex-DoF
Jacobian, whose FP64 atomics sum in a varying order), and a host pointer is
refused."""
import ctypes

import numpy as np
import pytest

import pointint_amd as P
from pointint_amd import _native as N
from pointint_amd import workloads as W

pytestmark = pytest.mark.gpu


def _call(fn, g, kind, x, kg, *extra):
    xd = np.ascontiguousarray(x, dtype=np.float64)
    return fn(g.ctx.handle, g.handle, kind, xd.ctypes.data, xd.size, kg.handle, *extra)


@pytest.mark.parametrize("prec", ["fast", "fp64"])
def test_form_factor_device_equals_host(prec):
    import torch
    V, F = W.icosphere(2)
    g = P.make_triangle_mesh(V * 0.4, F)
    kg = P.kgrid_from_rule(9, 8)
    x = np.concatenate([[0.1, -0.05, 0.02], np.eye(3).ravel()])
    lib = N.load()
    prc = {"fast": 1, "fp64": 0}[prec]
    host = np.empty(2 * kg.Q)
    N.check(_call(lib.pii_form_factor, g, 0, x, kg, prc, host.ctypes.data))
    dev = torch.zeros(2 * kg.Q, dtype=torch.float64, device="cuda")
    N.check(_call(lib.pii_form_factor_device, g, 0, x, kg, prc, dev.data_ptr()))
    assert np.array_equal(dev.cpu().numpy(), host)
    # a host pointer is not device memory
    with pytest.raises(N.PiiError, match="device memory"):
        N.check(_call(lib.pii_form_factor_device, g, 0, x, kg, prc, host.ctypes.data))


def test_jacobian_device_equals_host():
    import torch
    V, F = W.icosphere(1)
    g = P.make_triangle_mesh(V * 0.4, F)
    kg = P.kgrid_from_rule(9, 6)
    x = (V * 0.4).ravel()
    lib = N.load()
    host = np.empty(2 * kg.Q * x.size)
    N.check(_call(lib.pii_form_factor_jacobian, g, 1, x, kg, host.ctypes.data))
    dev = torch.zeros(2 * kg.Q * x.size, dtype=torch.float64, device="cuda")
    N.check(_call(lib.pii_form_factor_jacobian_device, g, 1, x, kg, dev.data_ptr()))
    # vertex-DoF columns are summed with FP64 atomics (order varies per run)
    d = dev.cpu().numpy()
    assert np.max(np.abs(d - host)) <= 1e-13 * np.max(np.abs(host))


def test_python_form_factor_into_tensor():
    import torch
    V, F = W.icosphere(2)
    g = P.make_triangle_mesh(V * 0.4, F)
    kg = P.kgrid_from_rule(9, 8)
    out = torch.empty(2 * kg.Q, dtype=torch.float64, device="cuda")
    P.compute_form_factor_into(out, g, kg)
    ref = P.compute_form_factor(g, kg)
    got = out.cpu().numpy()
    assert np.array_equal(got[0::2] + 1j * got[1::2], ref)
    with pytest.raises(ValueError):
        P.compute_form_factor_into(torch.empty(2 * kg.Q, dtype=torch.float32, device="cuda"), g, kg)


def test_device_output_is_ordered_after_the_callers_stream():
    """A fill of the output queued on torch's current stream behind long work
    must land before the library's write (the writes wait for the caller's
    stream), and the result must be complete when the call returns."""
    import torch
    V, F = W.icosphere(2)
    g = P.make_triangle_mesh(V * 0.4, F)
    kg = P.kgrid_from_rule(9, 8)
    ref = P.compute_form_factor(g, kg, precision="fast")
    s = torch.cuda.Stream()
    dev = torch.zeros(2 * kg.Q, dtype=torch.float64, device="cuda")
    with torch.cuda.stream(s):
        a = torch.randn(2048, 2048, device="cuda")
        for _ in range(8):
            a = a @ a / 2048.0          # keeps the stream busy before the fill
        dev.fill_(float("nan"))
        P.compute_form_factor_into(dev, g, kg, precision="fast")
        out = dev.clone()
    s.synchronize()
    got = out.cpu().numpy()
    assert np.array_equal(got[0::2] + 1j * got[1::2], ref)
