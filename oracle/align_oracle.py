"""CPU restatement of the reference's alignment application (numpy, FP64).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

  similarity()        cpp/app/align/similarity_measure.cpp:45-155
  pair_terms()        the matrix driver's V, grad_matrix and Gauss-Newton blocks
                      for the pair (identity, dofs2), from the FP64 oracle
                      (oracle.py), as similarity_measure.cpp:36-64 reads them
  newton()            cpp/app/align/optimizer.cpp:10-134
  moments_*()         cpp/app/align/warm_start.cpp:8-131
  warm_start()        cpp/app/align/warm_start.cpp:133-245
  sym_eigh()          eigen-decomposition with the eigenvector sign rule used
                      by the library (largest-magnitude component positive):
                      the reference takes Eigen's SelfAdjointEigenSolver
                      signs, an implementation detail of Eigen, so the warm
                      start's rotation is pinned only up to those signs.

Measures: 0 Ochiai, 1 Jaccard, 2 Dice.  Flags: 1 value, 2 gradient, 4 Hessian.
"""
from __future__ import annotations

import numpy as np

OCHIAI, JACCARD, DICE = 0, 1, 2
FUNC, GRAD, HESS = 1, 2, 4
ALL = FUNC | GRAD | HESS
IDENTITY = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1], dtype=np.float64)


def similarity(measure, V1, V_int, V2, gi=None, g2=None, Hi=None, H2=None, flags=FUNC):
    """(value, gradient, hessian, similarity) of -log(similarity) from the
    volume terms: V_int = V(0,1), V2 = V(1,1), gi = dV_int, g2 = dV2 (full
    derivative), Hi / H2 their Hessians."""
    need_g, need_h = bool(flags & GRAD), bool(flags & HESS)
    gs = Hs = None
    if measure == OCHIAI:
        inv = 1.0 / np.sqrt(V1 * V2)
        s = inv * V_int
        if need_g:
            gs = inv * gi - 0.5 * inv ** 3 * V_int * V1 * g2
        if need_h:
            Hs = (inv * Hi - 0.5 * inv * V_int / V2 * H2
                  + 0.75 * inv * V_int / V2 ** 2 * np.outer(g2, g2)
                  - 0.5 * inv / V2 * (np.outer(gi, g2) + np.outer(g2, gi)))
    elif measure == JACCARD:
        den = V1 + V2 - V_int
        s = V_int / den
        if need_g or need_h:
            dD = g2 - gi
        if need_g:
            gs = gi / den - V_int / den ** 2 * dD
        if need_h:
            Hs = (Hi / den - V_int / den ** 2 * (H2 - Hi)
                  - (np.outer(gi, dD) + np.outer(dD, gi)) / den ** 2
                  + 2.0 * V_int / den ** 3 * np.outer(dD, dD))
    elif measure == DICE:
        den = V1 + V2
        s = 2.0 * V_int / den
        coeff = -2.0 * V_int / den ** 2
        if need_g:
            gs = 2.0 / den * gi + coeff * g2
        if need_h:
            Hs = (2.0 / den * Hi + coeff * H2
                  - 2.0 / den ** 2 * (np.outer(gi, g2) + np.outer(g2, gi))
                  - 2.0 * coeff / den * np.outer(g2, g2))
    else:
        raise ValueError(measure)
    ss = max(s, 1e-10)
    value = -np.log(ss)
    grad = -gs / ss if need_g else None
    hess = np.outer(gs, gs) / ss ** 2 - Hs / ss if need_h else None
    return value, grad, hess, s


def pair_terms(O, og1, og2, dofs2, dirs, kmag, w, flags=ALL):
    """V_int, V2 and, per flags, dV_int = grad_matrix[1][0], dV2 =
    2 grad_matrix[1][1], H_int = hessian_jj[0][1] (zero) and H_2 = 2
    hessian_ij[1][1] (the Gauss-Newton block), from the FP64 oracle O."""
    A1 = O.compute_A_grid(og1, dirs, kmag, O.IDENTITY_DOFS, O.DOF_AFFINE)
    A2 = O.compute_A_grid(og2, dirs, kmag, dofs2, O.DOF_AFFINE)
    sc = O.SCALE
    V_int = sc * np.sum(w * (np.conj(A1) * A2).real)
    V2 = sc * np.sum(w * np.abs(A2) ** 2)
    out = {"V_int": V_int, "V2": V2}
    if flags & (GRAD | HESS):
        K = dirs * kmag[:, None]
        dA = np.stack([O.compute_A_gradient(og2, K[q], dofs2, O.DOF_AFFINE)
                       for q in range(len(kmag))])  # (Q, 12)
        out["gi"] = sc * ((dA * np.conj(A1)[:, None]).real * w[:, None]).sum(0)
        out["g2"] = 2.0 * sc * ((dA * np.conj(A2)[:, None]).real * w[:, None]).sum(0)
        if flags & HESS:
            H11 = sc * ((dA.T * w) @ np.conj(dA)).real
            out["Hi"] = np.zeros((12, 12))
            out["H2"] = H11 + H11.T
    return out


def evaluate(O, og1, og2, dofs2, dirs, kmag, w, measure, V1, flags=ALL):
    t = pair_terms(O, og1, og2, dofs2, dirs, kmag, w, flags)
    return similarity(measure, V1, t["V_int"], t["V2"], t.get("gi"), t.get("g2"), t.get("Hi"),
                      t.get("H2"), flags)


def newton(f, x0, tol=1e-6, max_iter=50, damping=0.1, trace=None):
    """Damped Newton with PSD projection and Armijo backtracking on
    f(x, flags) -> (value, grad, hess, similarity).  Returns a dict with the
    reference's OptimizationState fields; trace (a list) receives one
    (iteration, dofs, value, alpha) tuple per accepted step."""
    x = np.array(x0, dtype=np.float64)
    st = {"iteration": 0, "converged": False, "status": "", "similarity": 0.0, "loss": 0.0}
    alpha_init, alpha_min, c, rho, max_ls = 1.0, 1e-10, 0.1, 0.5, 20
    for it in range(max_iter):
        st["iteration"] = it
        v, g, H, s = f(x, ALL)
        st["similarity"], st["loss"] = s, v
        if np.linalg.norm(g) < tol:
            st["converged"] = True
            st["status"] = "Converged (gradient norm < tolerance)"
            break
        lam, Q = np.linalg.eigh(H)
        Hf = (Q * np.maximum(lam, 0.0)) @ Q.T + damping * np.eye(len(x))
        d = np.linalg.solve(Hf, -g)
        descent = -d @ g
        alpha = alpha_init
        fn = f(x + alpha * d, FUNC)[0]
        ls = 0
        while fn > v - c * alpha * descent and alpha > alpha_min and ls < max_ls:
            alpha *= rho
            fn = f(x + alpha * d, FUNC)[0]
            ls += 1
        if alpha <= alpha_min:
            st["status"] = "Line search failed (step size too small)"
            break
        x = x + alpha * d
        if trace is not None:
            trace.append((it, x.copy(), fn, alpha))
    if not st["converged"]:
        st["status"] = "Reached max iterations"
    st["dofs"] = x
    return st


def _affine(dofs):
    dofs = np.asarray(dofs, dtype=np.float64)
    return dofs[3:].reshape(3, 3), dofs[:3]


def moments_mesh(V, F, dofs=IDENTITY):
    """(volume, centroid, covariance) of a closed triangle mesh under AffineDoF
    dofs: tetrahedra against the origin, second moment from their centroids."""
    A, t = _affine(dofs)
    W = np.asarray(V, dtype=np.float64) @ A.T + t
    a, b, c = W[F[:, 0]], W[F[:, 1]], W[F[:, 2]]
    vol = np.einsum("ij,ij->i", a, np.cross(b - a, c - a)) / 6.0
    ct = (a + b + c) / 4.0
    Vt = vol.sum()
    if abs(Vt) <= 1e-10:
        return 0.0, np.zeros(3), np.zeros((3, 3))
    cen = (vol[:, None] * ct).sum(0) / Vt
    second = (ct.T * vol) @ ct
    return Vt, cen, second / Vt - np.outer(cen, cen)


def moments_surfels(P, wts, volume, sig2=None, dofs=IDENTITY):
    """Element-weighted centroid / covariance of disks (wts = areas) or
    Gaussians (wts = weights, sig2 = sigma^2 added on the diagonal); volume
    is the divergence-theorem volume passed through."""
    A, t = _affine(dofs)
    C = np.asarray(P, dtype=np.float64) @ A.T + t
    wts = np.asarray(wts, dtype=np.float64)
    tw = wts.sum()
    cen = (wts[:, None] * C).sum(0) / tw if tw > 1e-10 else np.zeros(3)
    D = C - cen
    cov = (D.T * wts) @ D
    if sig2 is not None:
        cov += np.eye(3) * np.sum(wts * sig2)
    if tw > 1e-10:
        cov /= tw
    return volume, cen, cov


def sym_eigh(M):
    """Eigenvalues ascending, eigenvectors (columns) with the largest-magnitude
    component positive (the first such component on ties)."""
    lam, Q = np.linalg.eigh(M)
    for j in range(Q.shape[1]):
        k = int(np.argmax(np.abs(Q[:, j])))
        if Q[k, j] < 0:
            Q[:, j] = -Q[:, j]
    return lam, Q


def warm_start(m1, m2):
    """AffineDoF vector of A = s R, t = c1 - A c2 from two (volume, centroid,
    covariance) triples."""
    V1, c1, C1 = m1
    V2, c2, C2 = m2
    e1, P1 = sym_eigh(C1)
    e2, P2 = sym_eigh(C2)
    if np.linalg.det(P1) < 0:
        P1[:, 2] = -P1[:, 2]
    if np.linalg.det(P2) < 0:
        P2[:, 2] = -P2[:, 2]
    U, _, Vt = np.linalg.svd(P1 @ P2.T)
    R = U @ Vt
    if np.linalg.det(R) < 0:
        U[:, 2] = -U[:, 2]
        R = U @ Vt
    if abs(V2) > 1e-10:
        s = np.cbrt(V1 / V2)
    else:
        m = [np.sqrt(e1[i] / e2[i]) for i in range(3) if e2[i] > 1e-6]
        s = float(np.mean(m)) if m else 1.0
    s = max(0.1, min(10.0, s))
    A = s * R
    t = c1 - A @ c2
    return np.concatenate([t, A.ravel()])
