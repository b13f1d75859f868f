"""GPU parity of the fast path (FP32 transcendentals / FP64 phases form
factor, split-precision MFMA Gram) against the FP64 oracle.

Error measures (stated with each assertion):
  form factors : ||A - A64||_w / ||A64||_w with ||x||_w^2 = sum_q w_q |x_q|^2
                 (the norm that bounds volume errors), gate 1e-6;
  volumes      : max_ij |V_ij - V64_ij| / sqrt(V64_ii V64_jj), gate 1e-5.
"""
import os

import numpy as np
import pytest

import pointint_amd as P
from pointint_amd import workloads as W

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def wrel(a, b, w):
    return np.sqrt((w * np.abs(a - b) ** 2).sum() / (w * np.abs(b) ** 2).sum())


def node_rel(A, A64, floor=1e-3):
    """Largest per-node relative error over the nodes with |A64| >= floor max|A64|."""
    m = np.abs(A64) >= floor * np.abs(A64).max()
    return float((np.abs(A[m] - A64[m]) / np.abs(A64[m])).max())


def pair_rel(V, V64, floor=1e-3):
    """Largest relative error |V_ij - V64_ij| / V64_ij over the pairs (i < j)
    that overlap: V64_ij >= floor sqrt(V64_ii V64_jj)."""
    d = np.sqrt(np.abs(np.outer(np.diag(V64), np.diag(V64))))
    m = np.triu(V64 >= floor * d, 1)
    return float((np.abs(V - V64)[m] / np.abs(V64[m])).max()), int(m.sum())


def vscaled(V, V64):
    d = np.sqrt(np.abs(np.outer(np.diag(V64), np.diag(V64))))
    return float(np.max(np.abs(V - V64) / d))


@pytest.fixture(scope="module")
def small():
    z = np.load(os.path.join(GOLD, "small_mixed_L8_M38.npz"))
    kg = P.kgrid_from_rule(9, 8)
    geoms = [
        P.make_triangle_mesh(z["ico_V"] * 0.6, z["ico_F"]),
        P.make_triangle_mesh(z["ico_V"] * 0.5, z["ico_F"]),
        P.make_point_cloud(z["P"] * 0.4, z["N"], z["r"] * 0.4),
        P.make_gaussian_splat(z["P"] * 0.3, z["N"], z["sig"], z["wt"]),
    ]
    dofs = [z["dofs0"], z["dofs1"], z["dofs2"], z["dofs3"]]
    kinds = ["affine", "trimesh", "affine", "affine"]
    return z, kg, geoms, dofs, kinds


def test_grid_symmetry_detected(small):
    _, kg, *_ = small
    assert kg.symmetric and kg.Qh * 2 == kg.Q


def test_form_factor_fast_matches_golden(small):
    z, kg, geoms, dofs, kinds = small
    for i in range(4):
        A = P.compute_form_factor(geoms[i], kg, dofs[i], kinds[i], precision="fast")
        e = wrel(A, z["J"][:, i], kg.w)
        assert e < 1e-6, (i, e)


def test_volume_matrix_fast_matches_golden(small):
    z, kg, geoms, dofs, kinds = small
    for gram in ("fp64", "bf16x3", "fp16x3"):
        V = P.compute_volume_matrix(geoms, kg, dofs, kinds, ff="fast", gram=gram)
        assert np.array_equal(V, V.T)
        assert vscaled(V, z["Vmat"]) < 1e-5, gram


@pytest.fixture(scope="module")
def ico_case(oracle):
    """12 icospheres (level 3) on Lebedev 302 x 24 -- large k-range, real meshes."""
    wl = W.config_icospheres(n_obj=12, level=3, seed=9, centre_box=1.0)
    d, w = oracle.load_lebedev(P.lebedev_path(29))
    dirs, kmag, wq = oracle.build_kgrid(d, w, 24)
    Vo, Jo = oracle.volume_matrix([oracle.geom_mesh(o.V, o.F) for o in wl.objects], dirs, kmag, wq)
    kg = P.KGrid(dirs, kmag, wq)
    geoms = [P.make_triangle_mesh(o.V, o.F) for o in wl.objects]
    return wl, kg, geoms, Vo, Jo


def test_icosphere_form_factor_fast(ico_case):
    wl, kg, geoms, Vo, Jo = ico_case
    worst = 0.0
    for i in (0, 5, 11):
        A = P.compute_form_factor(geoms[i], kg, precision="fast")
        worst = max(worst, wrel(A, Jo[:, i], kg.w))
    assert worst < 1e-6, worst


@pytest.mark.parametrize("gram", ["fp64", "bf16x3", "fp16x3"])
def test_icosphere_volume_matrix_fast(ico_case, gram):
    wl, kg, geoms, Vo, Jo = ico_case
    V = P.compute_volume_matrix(geoms, kg, ff="fast", gram=gram)
    assert vscaled(V, Vo) < 1e-5, (gram, vscaled(V, Vo))


def test_split_gram_against_fp64_gram_same_inputs(ico_case):
    """Isolates the Gram: both Grams consume the same fast-path X."""
    wl, kg, geoms, Vo, Jo = ico_case
    V64 = P.compute_volume_matrix(geoms, kg, ff="fast", gram="fp64")
    for gram, tol in (("bf16x3", 1e-5), ("fp16x3", 2e-6)):
        V = P.compute_volume_matrix(geoms, kg, ff="fast", gram=gram)
        assert vscaled(V, V64) < tol, (gram, vscaled(V, V64))


def test_asymmetric_grid_uses_full_grid(oracle):
    d, w = oracle.load_lebedev(P.lebedev_path(9))
    keep = np.arange(0, len(w), 2)[:-1]       # drops antipodes
    dirs, kmag, wq = oracle.build_kgrid(d[keep], w[keep], 6)
    kg = P.KGrid(dirs, kmag, wq)
    assert not kg.symmetric and kg.Qh == kg.Q
    Vi, Fi = W.icosphere(2)
    g = P.make_triangle_mesh(Vi * 0.4 + 0.1, Fi)
    A = P.compute_form_factor(g, kg, precision="fast")
    ref = oracle.compute_A_grid(oracle.geom_mesh(Vi * 0.4 + 0.1, Fi), dirs, kmag)
    assert wrel(A, ref, wq) < 1e-6


def test_full_size_grid_errors(oracle):
    """Three objects of the 256-icosphere workload on its full k-grid
    (Lebedev 1202 x L=64, k up to ~1.8e3): form factors vs the FP64 oracle
    (weighted norm) and the 3 x 3 V (scaled), both Gram modes."""
    wl = W.config_icospheres(n_obj=256, level=3, seed=2)
    objs = [wl.objects[i] for i in (0, 101, 202)]
    d, w = oracle.load_lebedev(P.lebedev_path(59))
    dirs, kmag, wq = oracle.build_kgrid(d, w, 64)
    Vo, Jo = oracle.volume_matrix([oracle.geom_mesh(o.V, o.F) for o in objs], dirs, kmag, wq)
    kg = P.KGrid(dirs, kmag, wq)
    geoms = [P.make_triangle_mesh(o.V, o.F) for o in objs]
    errs, shell_errs, node_errs, node_errs3 = [], [], [], []
    M = len(w)
    for i, g in enumerate(geoms):
        A = P.compute_form_factor(g, kg, precision="fast")
        errs.append(wrel(A, Jo[:, i], wq))
        # per node against the largest |A| of its shell (|k| fixed)
        e = np.abs(A - Jo[:, i]).reshape(-1, M).max(axis=1)
        shell_errs.append(float((e / np.abs(Jo[:, i]).reshape(-1, M).max(axis=1)).max()))
        node_errs.append(node_rel(A, Jo[:, i], 3e-2))
        node_errs3.append(node_rel(A, Jo[:, i], 1e-3))
    print("form factor weighted rel errors", errs, "per node / shell max", shell_errs,
          "per node (|A| >= 3e-2 max)", node_errs, "per node (|A| >= 1e-3 max)", node_errs3)
    assert max(errs) < 1e-6
    # per node: within 1e-6 of the largest |A| on the node's shell, and within
    # 1e-6 of its own |A| where |A| >= 3e-2 max |A| (the FP32 per-face error is
    # about 3e-8 of max |A| at every node, so nodes near 1e-3 max |A| carry
    # about 1e-5 of their own value)
    assert max(shell_errs) < 1e-6
    assert max(node_errs) < 1e-6
    assert max(node_errs3) < 5e-5
    for gram in ("bf16x3", "fp16x3"):
        V = P.compute_volume_matrix(geoms, kg, ff="fast", gram=gram)
        print(gram, "scaled V error", vscaled(V, Vo))
        assert vscaled(V, Vo) < 1e-5


def test_issue_rate_probe():
    """The probe kernels behind the benchmark's peaks: rates below the
    published MI355X vector peaks (FP64 78.6 TFLOP/s, FP32 157.3 TFLOP/s),
    plain (non-packed) FP32 FMAs well above the FP64 rate (gfx950 issues them
    from two waves per cycle, so 157.3 and not 78.6 is the FP32 kernels'
    roof), the sine at about a third of the FMA instruction rate."""
    ctx = P.Context.default()
    f64 = ctx.issue_rate("fp64_fma")
    f32 = ctx.issue_rate("fp32_fma")
    sn = ctx.issue_rate("sin")
    print(f"FP64 FMA {f64 / 1e12:.1f} TFLOP/s, FP32 FMA {f32 / 1e12:.1f} TFLOP/s, "
          f"sin {sn / 1e12:.2f} T/s")
    assert 5e12 < f64 < 1.05 * 78.6e12
    assert 5e12 < f32 < 1.05 * 157.3e12
    assert f32 > 1.3 * f64
    assert 1e11 < sn < 1.05 * 78.6e12 / 2


@pytest.fixture(scope="module")
def ico24(oracle):
    """24 objects of the 256-icosphere workload (seed 2, centres in [0, 2]^3)
    on its full k-grid (Lebedev 1202 x L=64, Q = 76,928): about half of the
    pairs overlap, many of them slightly (V_ij down to 1e-3 sqrt(V_ii V_jj))."""
    wl = W.config_icospheres(n_obj=256, level=3, seed=2)
    objs = wl.objects[:24]
    d, w = oracle.load_lebedev(P.lebedev_path(59))
    dirs, kmag, wq = oracle.build_kgrid(d, w, 64)
    Vo, _ = oracle.volume_matrix([oracle.geom_mesh(o.V, o.F) for o in objs], dirs, kmag, wq)
    kg = P.KGrid(dirs, kmag, wq)
    geoms = [P.make_triangle_mesh(o.V, o.F) for o in objs]
    return kg, geoms, Vo


@pytest.mark.parametrize("gram", ["fp64", "bf16x3", "fp16x3"])
def test_per_pair_relative_error(ico24, gram):
    """The volume gate per pair: |V_ij - V64_ij| / V64_ij <= 1e-5 on
    every overlapping pair (V64_ij >= 1e-3 sqrt(V64_ii V64_jj)), V64 from the
    FP64 oracle on the same grid.  The fp64 Gram on the same fast form factors
    isolates the form factors' share."""
    kg, geoms, Vo = ico24
    V = P.compute_volume_matrix(geoms, kg, ff="fast", gram=gram)
    e, npairs = pair_rel(V, Vo)
    print(f"{gram}: max per-pair relative error {e:.3e} over {npairs} overlapping pairs")
    assert npairs >= 50
    assert e < 1e-5, (gram, e)


def test_split_gram_per_pair_against_fp64_gram(ico24):
    """The split Grams against the FP64 Gram of the same FP32 X: the Gram's own
    per-pair error, well inside the gate (three BF16 pieces carry the FP32
    significand; two FP16 pieces 22 bits)."""
    kg, geoms, Vo = ico24
    V64 = P.compute_volume_matrix(geoms, kg, ff="fast", gram="fp64")
    for gram, tol in (("bf16x3", 3e-6), ("fp16x3", 5e-6)):
        V = P.compute_volume_matrix(geoms, kg, ff="fast", gram=gram)
        e, _ = pair_rel(V, V64)
        print(f"{gram} vs fp64 Gram on the same X: {e:.3e}")
        assert e < tol, (gram, e)
