"""GPU parity of the alignment consumer and the one-pass affine derivatives.

* pii_volume_matrix_affine_derivatives: V, grad_matrix and the Gauss-Newton
  blocks of AffineDoF triangles / disks / Gaussians against the FP64 oracle
  (V to 1e-5 of sqrt(V_ii V_jj), gradients and H to 1e-5 in norm).
* pointint_align (the C++ similarity, optimizer and warm start on the matrix
  driver) against the oracle's restatement (oracle/align_oracle.py):
  similarity value to 1e-6, gradient and Hessian of -log(similarity) to 1e-4
  in norm; the first Newton iterates to 1e-4 of the step; and progress towards a
  known rigid transform.
"""
import os

import numpy as np
import pytest

import align_oracle as AO
import pointint_amd as P
from pointint_amd import workloads as W
from test_oracle_align import _aniso_mesh, run_align, write_off

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MEASURES = {"ochiai": AO.OCHIAI, "jaccard": AO.JACCARD, "dice": AO.DICE}


def _grid(oracle, order, nrad):
    d, w = oracle.load_lebedev(P.lebedev_path(order))
    return oracle.build_kgrid(d, w, nrad)


@pytest.mark.parametrize("gram", ["fp64", "fp16x3"])
def test_affine_derivatives_match_oracle(oracle, gram):
    rng = np.random.default_rng(61)
    V1, F1 = W.icosphere(2)
    Pc, Nn, _ = W.sphere_pointcloud(200)
    rho = rng.uniform(0.02, 0.05, size=200)
    sig = rng.uniform(0.02, 0.05, size=200)
    wt = rng.uniform(0.5, 1.5, size=200) * 0.01
    objs = [(P.make_triangle_mesh(V1 * 0.45, F1), oracle.geom_mesh(V1 * 0.45, F1)),
            (P.make_point_cloud(Pc * 0.4, Nn, rho), oracle.geom_disks(Pc * 0.4, Nn, rho)),
            (P.make_gaussian_splat(Pc * 0.35, Nn, sig, wt),
             oracle.geom_gaussians(Pc * 0.35, Nn, sig, wt)),
            (P.make_triangle_mesh(V1 * 0.3, F1), oracle.geom_mesh(V1 * 0.3, F1))]
    dofs = [np.concatenate([rng.uniform(-0.2, 0.2, 3),
                            (W.random_rotation(rng) @ np.diag(rng.uniform(0.8, 1.2, 3))).ravel()])
            for _ in objs]
    dirs, kmag, wq = _grid(oracle, 9, 12)
    kg = P.KGrid(dirs, kmag, wq)
    n = len(objs)
    K = dirs * kmag[:, None]
    J = np.stack([oracle.compute_A_grid(og, dirs, kmag, x, oracle.DOF_AFFINE)
                  for (_, og), x in zip(objs, dofs)], axis=1)
    dA = np.stack([np.stack([oracle.compute_A_gradient(og, K[q], x, oracle.DOF_AFFINE)
                             for q in range(len(kmag))]) for (_, og), x in zip(objs, dofs)])
    sc = oracle.SCALE
    V64 = sc * ((J.conj().T * wq) @ J).real
    G64 = sc * np.einsum("iqt,qj,q->ijt", dA, np.conj(J), wq).real
    D = dA.transpose(1, 0, 2).reshape(len(kmag), 12 * n)
    H64 = sc * ((D.T * wq) @ np.conj(D)).real
    V, G, H = P.volume_matrix_derivatives([g for g, _ in objs], kg, dofs, gram=gram)
    d = np.sqrt(np.outer(np.diag(V64), np.diag(V64)))
    assert np.max(np.abs(V - V64) / d) < 1e-5
    eg = np.linalg.norm(G - G64) / np.linalg.norm(G64)
    eh = np.linalg.norm(H - H64) / np.linalg.norm(H64)
    print(f"{gram}: V {np.max(np.abs(V - V64) / d):.2e} grad {eg:.2e} H {eh:.2e}")
    assert eg < 1e-5 and eh < 1e-5
    for i in range(n):  # each object's own gradient blocks
        gi = np.linalg.norm(G[i] - G64[i]) / np.linalg.norm(G64[i])
        assert gi < 1e-4, (i, gi)
    # the Hessian-only entry gives the same blocks
    Hh = P.volume_matrix_hessian([g for g, _ in objs], kg, dofs, gram=gram)
    assert np.linalg.norm(Hh - H) <= 1e-12 * np.linalg.norm(H)


def _pair(tmp_path, rng, level=1, scale=0.5):
    V, F = _aniso_mesh(rng, level)
    V = V * scale
    R = W.random_rotation(rng)
    V2 = (V * 0.9) @ R.T + np.array([0.06, -0.04, 0.03])
    p1, p2 = str(tmp_path / "m1.off"), str(tmp_path / "m2.off")
    write_off(p1, V, F)
    write_off(p2, V2, F)
    return (V, F), (V2, F), p1, p2


@pytest.mark.parametrize("measure", ["ochiai", "jaccard", "dice"])
def test_cxx_similarity_matches_oracle(oracle, tmp_path, measure):
    rng = np.random.default_rng(71)
    (V, F), (V2, F2), p1, p2 = _pair(tmp_path, rng)
    dirs, kmag, wq = _grid(oracle, 9, 12)
    x = AO.warm_start(AO.moments_mesh(V, F), AO.moments_mesh(V2, F2))
    x = x + rng.normal(size=12) * 0.02
    out = run_align([p1, p2, "9", "12", "--measure", measure, "--eval-only", "--dofs",
                     *["%.17g" % v for v in x]])
    val, sim = float(out["eval"][0]), float(out["eval"][1])
    g = np.array([float(v) for v in out["grad"]])
    H = np.array([float(v) for v in out["hess"]]).reshape(12, 12)
    og1, og2 = oracle.geom_mesh(V, F), oracle.geom_mesh(V2, F2)
    V1 = oracle.volume_matrix([og1], dirs, kmag, wq)[0][0, 0]
    assert abs(float(out["V1"][0]) - V1) < 1e-6 * V1
    v64, g64, H64, s64 = AO.evaluate(oracle, og1, og2, x, dirs, kmag, wq, MEASURES[measure], V1)
    print(f"{measure}: s {sim:.6f} vs {s64:.6f}; grad {np.linalg.norm(g - g64) / np.linalg.norm(g64):.2e}"
          f" H {np.linalg.norm(H - H64) / np.linalg.norm(H64):.2e}")
    assert abs(val - v64) < 1e-6 and abs(sim - s64) < 1e-6 * s64
    assert np.linalg.norm(g - g64) < 1e-4 * np.linalg.norm(g64)
    assert np.linalg.norm(H - H64) < 1e-4 * np.linalg.norm(H64)


def test_cxx_newton_iterates_match_oracle(oracle, tmp_path):
    """Two Newton iterations from the warm start (Dice), device against the
    oracle loop on the same grid: iterates agree to 1e-4 of the step length."""
    rng = np.random.default_rng(73)
    (V, F), (V2, F2), p1, p2 = _pair(tmp_path, rng)
    dirs, kmag, wq = _grid(oracle, 9, 12)
    og1, og2 = oracle.geom_mesh(V, F), oracle.geom_mesh(V2, F2)
    V1 = oracle.volume_matrix([og1], dirs, kmag, wq)[0][0, 0]
    x0 = AO.warm_start(AO.moments_mesh(V, F), AO.moments_mesh(V2, F2))

    def f(x, flags):
        return AO.evaluate(oracle, og1, og2, x, dirs, kmag, wq, AO.DICE, V1, flags)

    for iters in (1, 2):
        st = AO.newton(f, x0, max_iter=iters)
        out = run_align([p1, p2, "9", "12", "--measure", "dice", "--max-iter", str(iters)])
        x_dev = np.array([float(v) for v in out["dofs"]])
        x_start = np.array([float(v) for v in out["start"]])
        assert np.allclose(x_start, x0, atol=1e-12)
        step = np.linalg.norm(st["dofs"] - x0)
        err = np.linalg.norm(x_dev - st["dofs"])
        print(f"{iters} iterations: step {step:.3e}, |x_dev - x_oracle| {err:.2e}")
        assert step > 1e-3 and err < 1e-4 * step


def test_cxx_alignment_recovers_rigid_transform(tmp_path):
    """Geometry 2 is geometry 1 moved by a known rotation and shift; Newton on
    -log(Dice) from the identity approaches the inverse transform.  The
    Hessian model has no curvature of V_int (the driver's hessian_jj[0][1] is
    zero in the Gauss-Newton form), so progress along the non-rigid directions
    is slow; the bounds below are what 60 iterations reach on this grid."""
    rng = np.random.default_rng(79)
    V, F = _aniso_mesh(rng, 3)
    V = V * 0.5
    ax = rng.normal(size=3)
    ax /= np.linalg.norm(ax)
    th = np.deg2rad(20.0)
    Kx = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
    R = np.eye(3) + np.sin(th) * Kx + (1 - np.cos(th)) * Kx @ Kx
    b = np.array([0.08, -0.05, 0.04])
    V2 = V @ R.T + b
    p1, p2 = str(tmp_path / "m1.off"), str(tmp_path / "m2.off")
    write_off(p1, V, F)
    write_off(p2, V2, F)
    start = run_align([p1, p2, "29", "16", "--measure", "dice", "--no-warm-start",
                       "--eval-only"])
    out = run_align([p1, p2, "29", "16", "--measure", "dice", "--no-warm-start", "--max-iter",
                     "60"])
    conv, it, sim = int(out["final"][0]), int(out["final"][1]), float(out["final"][2])
    x = np.array([float(v) for v in out["dofs"]])
    want = np.concatenate([-R.T @ b, R.T.ravel()])
    x0 = np.array([float(v) for v in start["start"]])
    print(f"recover: Dice {float(start['eval'][1]):.6f} -> {sim:.9f} after {it + 1} iterations "
          f"(converged {conv}), |x - x*| {np.linalg.norm(x0 - want):.2e} -> "
          f"{np.linalg.norm(x - want):.2e}, {float(out['final'][4]):.1f} ms")
    assert sim > 0.9999 and sim > float(start["eval"][1])
    assert np.linalg.norm(x - want) < 0.25 * np.linalg.norm(x0 - want)
