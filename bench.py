#!/usr/bin/env python3
"""Benchmark of the k-space intersection-volume path on one MI355X.

Default workload (serving): a Poisson stream of 2,000 queries/s against a
device-resident pool of 4,096 star-shaped meshes (triangle counts
log-uniform in [80, 20480]); each query asks for the all-pairs volume matrix
of a scene of K (log-uniform in [8, 512]) pool objects under fresh random
rigid transforms, on the Lebedev 1202 x L=64 k-grid (Q = 76,928).  One
"step" is one segment (--segment-s seconds) of arrivals; the warm-up
segments' queries are served but not counted.  Queries are admitted by HBM
budget and a latency target and served by continuous batching replayed from
hipGraphs.

Prints one JSON line: completed pairs/s in the timed window as `value`,
p50/p99 latency from the scheduled arrival, the plain FP64 GPU path and its
per-pair error on a sample of the served scenes, the form-factor kernel's
roofline figures, the FP64 CPU oracle on the same sample, and (extras) the
reference's own pairwise benchmark (its only published timing) and the
fixed 256-icosphere and 1,024-mixed-object matrices.  Other workloads:
--workload icospheres | mixed | diff | align.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import pointint_amd as P  # noqa: E402
from pointint_amd import workloads as W  # noqa: E402

# Algorithmic VALU work per triangle x k-node of the per-face kernel: 30
# FMA-class operations (alpha, beta, gamma and phase dot products; complex Phi,
# phase and accumulate) = 60 flop, plus 3 sin/cos pairs and 3 reciprocals
# counted as one flop each.
FLOP_PER_TRI_K = 66.0
# Moment series (order 56): per face x direction, 4 dot products (20 flop),
# H_0 (1) and per order one multiply and three FMAs (7 flop x 56); per object
# x node, two Horner sums (2 x 56 flop) and the phase product (12).
SERIES_ORDER = 56
FLOP_PER_FACE_DIR = 21.0 + 7.0 * SERIES_ORDER
FLOP_PER_SERIES_NODE = 2.0 * SERIES_ORDER + 12.0
# Per disk x k-node of the per-element kernel: about 25 FMA-class operations
# (50 flop), J1(x)/x (about 15 FMA, 30 flop), one sincos, one sqrt, one
# reciprocal.  Per disk x direction of the series: 2 dot products and the
# element factors (15 flop), then per order the three-term recurrence (two
# multiplies, one FMA) and the sum into H_n: 5 flop x 56.
FLOP_PER_DISK_K = 83.0
FLOP_PER_SURFEL_DIR = 15.0 + 5.0 * SERIES_ORDER
# MI355X vector peaks: FP32 157.3 TFLOP/s (reachable without packed
# instructions: gfx950 issues plain FP32 VALU instructions from two waves per
# cycle; the issue-rate probe sustains about 115 TFLOP/s of FMAs under load),
# FP64 78.6 TFLOP/s.  The per-face kernels compute in FP32 with FP64 phases;
# the moment series in FP64.
PEAK_FP32_VALU_TFLOPS = 157.3
PEAK_FP64_VALU_TFLOPS = 78.6
KERNEL_PEAK = {"k_ff_fast": PEAK_FP32_VALU_TFLOPS,
               "k_series_tri / k_series_surfel": PEAK_FP64_VALU_TFLOPS}
# HBM bytes per launch of the dominant kernel, per workload, from rocprofv3
# FETCH_SIZE / WRITE_SIZE passes (2 x fetch + write, KiB, the factors of a
# 1 GiB device copy; the profiles/ summaries hold the counts and commands);
# the serving figure is the average over the batches of a run.  Counters
# cannot be read inside this run.
MEASURED_TRAFFIC = {("icospheres_L3_N256", "k_ff_fast"): 40.2e6,
                    ("mixed_N1024", "k_ff_fast"): 1229.0e6,
                    ("serving", "k_ff_fast"): 120.9e6}
PEAK_MFMA_F16_TFLOPS = 2500.0


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def cpu_baseline(wl, kg, target_seconds=15.0):
    """FP64 CPU oracle (restated reference CPU path: k-nodes in parallel, then
    pairs in parallel) on a k-node subsample, extrapolated linearly in Q."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    geoms = oracle_geoms(O, wl)
    Q = kg.Q
    # calibrate on a tiny sample, then size the sample to ~target_seconds
    stride = max(1, Q // 64)
    idx = np.arange(0, Q, stride)
    t0 = time.perf_counter()
    O.volume_matrix(geoms, kg.dirs[idx], kg.kmag[idx], kg.w[idx])
    t_cal = time.perf_counter() - t0
    n_target = int(len(idx) * target_seconds / max(t_cal, 1e-3))
    n_target = max(len(idx), min(Q, n_target))
    stride = max(1, Q // n_target)
    idx = np.arange(0, Q, stride)
    t0 = time.perf_counter()
    O.volume_matrix(geoms, kg.dirs[idx], kg.kmag[idx], kg.w[idx])
    t_s = time.perf_counter() - t0
    t_full = t_s * Q / len(idx)
    return t_full, len(idx), t_s, O.threads()


def copy_calibration(nbytes: int) -> None:
    """One device-to-device copy of nbytes (hipMemcpy through the HIP runtime
    the library uses), so that a counter pass reads FETCH_SIZE / WRITE_SIZE of
    a kernel of known traffic next to the form-factor kernel's."""
    import ctypes as C
    P.Context.default()  # loads the library and its HIP runtime
    hip = C.CDLL("libamdhip64.so.7")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(f"[bench] copy calibration: {nbytes} bytes device-to-device", file=sys.stderr)


def make_geoms(wl):
    out = []
    for o in wl.objects:
        if isinstance(o, W.MeshObj):
            out.append(P.make_triangle_mesh(o.V, o.F))
        else:
            out.append(P.make_point_cloud(o.P, o.N, o.r))
    return out


def oracle_geoms(O, wl):
    out = []
    for o in wl.objects:
        if isinstance(o, W.MeshObj):
            out.append(O.geom_mesh(o.V, o.F))
        else:
            out.append(O.geom_disks(o.P, o.N, o.r))
    return out


def run_diff(args):
    """Forward V + hand-written adjoint of L = sum_{i<j} V_ij for 64 icospheres
    (level 4, 5,120 triangles, vertex DoFs) on Lebedev 590 x L=64; the
    gradient is checked against FP64 central differences (h = 1e-4) of the FP64
    GPU forward on --fd-vertices random vertices."""
    wl = W.config_diff()
    ctx = P.Context.default()
    kg = P.kgrid_from_rule(wl.lebedev_order, wl.n_radial)
    geoms = [P.make_triangle_mesh(o.V, o.F) for o in wl.objects]
    n = len(geoms)
    xs = [o.V.astype(np.float32).astype(np.float64).reshape(-1) for o in wl.objects]
    G = np.triu(np.ones((n, n)), 1)
    kinds = ["trimesh"] * n
    for _ in range(args.warmup):
        P.volume_matrix_vjp(geoms, kg, G, xs, kinds)
    ctx.set_profiling(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        grads, V = P.volume_matrix_vjp(geoms, kg, G, xs, kinds)
    t1 = time.perf_counter()
    tm = ctx.timings()
    ctx.set_profiling(False)
    ms = (t1 - t0) * 1e3 / args.steps
    d = np.sqrt(np.outer(np.diag(V), np.diag(V)))
    iu = np.triu_indices(n, 1)
    # pairs whose spheres intersect (centre distance below the sum of the
    # radii: the workload's "about 30 %"), and pairs whose grid V exceeds
    # 1e-3 of sqrt(V_ii V_jj): more, because the truncated k-grid blurs each
    # indicator, so near misses get a small nonzero V
    cen = np.array([o.V.mean(axis=0) for o in wl.objects])
    rad = np.array([np.linalg.norm(o.V - o.V.mean(axis=0), axis=1).max() for o in wl.objects])
    dist = np.linalg.norm(cen[:, None] - cen[None], axis=2)
    frac_geom = float((dist < rad[:, None] + rad[None])[iu].mean())
    frac_overlap = float((V[iu] > 1e-3 * d[iu]).mean())
    # FD against the FP64 forward: only object i's column moves, so
    # L(x_i) = Re<A_i, sum_{j != i} A_j>_w / (8 pi^3) + const.
    J = np.stack([P.compute_form_factor(g, kg, x, "trimesh", precision="fp64")
                  for g, x in zip(geoms, xs)], axis=1)
    rng = np.random.default_rng(7)
    h = 1e-4
    errs, errs_r = [], []
    tot = J.sum(axis=1)

    def central(i, v, c, hh, B):
        vals = []
        for sgn in (1, -1):
            x = xs[i].copy()
            x[3 * v + c] += sgn * hh
            Ai = P.compute_form_factor(geoms[i], kg, x, "trimesh", precision="fp64")
            vals.append((kg.w * (Ai * np.conj(B)).real).sum() / (8 * np.pi ** 3))
        return (vals[0] - vals[1]) / (2 * hh)

    for _ in range(args.fd_vertices):
        i = int(rng.integers(n))
        v = int(rng.integers(xs[i].shape[0] // 3))
        c = int(rng.integers(3))
        B = tot - J[:, i]
        fd = central(i, v, c, h, B)
        # Richardson: (4 D(h/2) - D(h)) / 3 removes the h^2 term, which the
        # high-|k| shells (|k| up to ~1e3 on this grid) make large at h = 1e-4
        fd_r = (4.0 * central(i, v, c, 0.5 * h, B) - fd) / 3.0
        scale = np.abs(grads[i]).max()
        errs.append(abs(grads[i][3 * v + c] - fd) / scale)
        errs_r.append(abs(grads[i][3 * v + c] - fd_r) / scale)
    out = {
        "metric": "adjoint: dLoss/dvertex evaluations (forward V + VJP) per second",
        "value": 1e3 / ms, "unit": "gradient evals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic (seeded icospheres)",
        "config": {"workload": "64 icospheres L4 (5,120 tris, vertex DoFs), Lebedev 590 x L=64 "
                               "(Q=37,760), loss sum_{i<j} V_ij", "objects": n, "Q": kg.Q},
        "phase_ms": tm, "intersecting_pair_fraction": frac_geom,
        "pairs_with_v_above_1e-3": frac_overlap,
        "fd_check": ({"vertices": args.fd_vertices, "h": h,
                      "max_err_rel_to_max_grad": float(max(errs)),
                      "median_err_rel_to_max_grad": float(np.median(errs)),
                      "richardson_max_err_rel_to_max_grad": float(max(errs_r)),
                      "richardson_median_err_rel_to_max_grad": float(np.median(errs_r))}
                     if errs else None),
    }
    print(json.dumps(out))


def align_pair(level=5, seed=8):
    """Two star-shaped anisotropic icospheres (level 5: 20,480 triangles), the
    second rotated by 20 degrees, scaled by 0.95 and shifted."""
    rng = np.random.default_rng(seed)
    V, F = W.icosphere(level)
    V = W.star_perturb(V, rng, 0.1) * np.array([0.5, 0.35, 0.22])
    ax = np.array([0.3, -0.5, 0.8]) / np.linalg.norm([0.3, -0.5, 0.8])
    K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
    th = np.deg2rad(20.0)
    R = np.eye(3) + np.sin(th) * K + (1 - np.cos(th)) * K @ K
    return (V, F), (0.95 * V @ R.T + np.array([0.05, -0.03, 0.04]), F)


def run_align(args):
    """Newton iteration of the alignment consumer on the device: the SIM_ALL
    evaluation (V, per-pair gradients and the Gauss-Newton block of the pair,
    one pass) plus one function-only line-search evaluation, for two
    20,480-triangle meshes on the application's default grid (Lebedev 302 x
    L=32, Q = 9,664).  Also runs the C++ optimizer (pointint_align) end to
    end on the same meshes."""
    import subprocess
    import tempfile
    (V1, F1), (V2, F2) = align_pair()
    kg = P.kgrid_from_rule(29, 32)
    g1, g2 = P.make_triangle_mesh(V1, F1), P.make_triangle_mesh(V2, F2)
    ctx = P.Context.default()
    ident = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1], dtype=np.float64)
    x = np.concatenate([[0.01, -0.02, 0.01], np.eye(3).ravel() * 1.02])

    def iteration():
        V, G, H = P.volume_matrix_derivatives([g1, g2], kg, [ident, x], gram="fp16x3")
        Vf = P.compute_volume_matrix([g1, g2], kg, [ident, x], ["affine"] * 2, gram="fp16x3")
        return V, G, H, Vf

    for _ in range(args.warmup):
        iteration()
    ctx.set_profiling(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        iteration()
    t1 = time.perf_counter()
    tm = ctx.timings()
    ctx.set_profiling(False)
    ms = (t1 - t0) * 1e3 / args.steps
    # end to end: warm start + Newton in C++ (one process, geometry resident)
    e2e = None
    exe = os.path.join(ROOT, "pointint_amd", "pointint_align")
    with tempfile.TemporaryDirectory() as td:
        paths = []
        for nm, (V, F) in (("m1", (V1, F1)), ("m2", (V2, F2))):
            pth = os.path.join(td, nm + ".off")
            with open(pth, "w") as f:
                f.write(f"OFF\n{len(V)} {len(F)} 0\n")
                np.savetxt(f, V, fmt="%.17g")
                np.savetxt(f, np.c_[np.full(len(F), 3), F], fmt="%d")
            paths.append(pth)
        r = subprocess.run([exe, *paths, "29", "32", "--measure", "dice", "--max-iter", "30"],
                           capture_output=True, text=True, timeout=600)
        if r.returncode == 0:
            res = {ln[1:].split()[0]: ln[1:].split()[1:] for ln in r.stdout.splitlines()
                   if ln.startswith("@")}
            n_eval = r.stdout.count("Trying alpha")
            e2e = {"converged": bool(int(res["final"][0])),
                   "iterations": int(res["final"][1]) + 1, "dice": float(res["final"][2]),
                   "total_ms": float(res["final"][4]), "line_search_evals": n_eval,
                   "ms_per_iteration": float(res["final"][4]) / (int(res["final"][1]) + 1),
                   "status": " ".join(res["status"])}
        else:
            print(f"[bench] pointint_align failed: {r.stderr[-500:]}", file=sys.stderr)
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        import align_oracle as AO
        og1, og2 = O.geom_mesh(V1, F1), O.geom_mesh(V2, F2)

        def sample(stride):
            idx = np.arange(0, kg.Q, stride)
            t0 = time.perf_counter()
            AO.pair_terms(O, og1, og2, x, kg.dirs[idx], kg.kmag[idx], kg.w[idx], AO.ALL)
            AO.pair_terms(O, og1, og2, x, kg.dirs[idx], kg.kmag[idx], kg.w[idx], AO.FUNC)
            return idx, time.perf_counter() - t0

        # calibrate on a few nodes, then size the sample to about --cpu-seconds / 3
        idx, t_cal = sample(max(1, kg.Q // 24))
        n_target = int(len(idx) * args.cpu_seconds / 3.0 / max(t_cal, 1e-3))
        idx, t_s = sample(max(1, kg.Q // max(len(idx), min(kg.Q, n_target))))
        cpu = {"value": 1.0 / (t_s * kg.Q / len(idx)), "unit": "Newton iterations/s",
               "cores": O.threads(), "kind": "port",
               "sample": f"{len(idx)} of {kg.Q} k-nodes (one SIM_ALL + one function "
                         f"evaluation; A on the threaded oracle, dA/dtheta node by node on "
                         f"one thread), extrapolated linearly in Q; {t_s:.1f} s on "
                         f"{cpu_model()}"}
    out = {
        "metric": "alignment: Newton iterations/s (SIM_ALL evaluation + one line-search "
                  "evaluation, N=2)",
        "value": 1e3 / ms, "unit": "Newton iterations/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic (seeded star-shaped meshes)",
        "config": {"workload": "2 meshes x 20,480 triangles, AffineDoF, Lebedev 302 x L=32 "
                               "(Q=9,664), Dice", "Q": kg.Q},
        "phase_ms": tm, "end_to_end": e2e, "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def measure_matrix(wl, gram, steps, warmup, ff="fast", fp64_ref=True, attribution=False):
    """One all-pairs volume matrix per step on a fixed workload: timing, phase
    and kernel figures, the plain FP64 GPU path on the same inputs, the
    relative error of every Gram mode against it and, with attribution, the
    split of the speed-up into hemisphere symmetry, moment series and
    precision."""
    ctx = P.Context.default()
    kg = P.kgrid_from_rule(wl.lebedev_order, wl.n_radial)
    geoms = make_geoms(wl)
    N = len(geoms)
    pairs = N * (N - 1) // 2
    primk = wl.prim_count() * kg.Q

    def step(f, g):
        return P.compute_volume_matrix(geoms, kg, ff=f, gram=g)

    for _ in range(warmup):
        step(ff, gram)
    ctx.set_profiling(True)
    ff_ms, gram_ms, pre_ms, ser_ms, face_ms = [], [], [], [], []
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        V = step(ff, gram)
        t = ctx.timings()
        ff_ms.append(t["form_factor"]); gram_ms.append(t["gram"]); pre_ms.append(t["prepass"])
        ser_ms.append(t["ff_series"]); face_ms.append(t["ff_per_face"])
    ctx.synchronize()
    t1 = time.perf_counter()
    ctx.set_profiling(False)
    work = ctx.work()
    ms_step = (t1 - t0) * 1e3 / steps
    print(f"[bench] {wl.name} {ff}/{gram}: {ms_step:.2f} ms/step, FF {np.mean(ff_ms):.2f} ms, "
          f"Gram {np.mean(gram_ms):.3f} ms", file=sys.stderr, flush=True)
    out = {"workload": wl.name, "objects": N, "elements": wl.prim_count(), "Q": kg.Q,
           "ff": ff, "gram": gram, "steps": steps, "warmup": warmup, "ms_per_step": ms_step,
           "pairs_per_s": pairs / (ms_step * 1e-3),
           "prim_k_per_s_nominal": primk / (ms_step * 1e-3),
           "ff_ms": float(np.mean(ff_ms)), "gram_ms": float(np.mean(gram_ms)),
           "prepass_ms": float(np.mean(pre_ms))}
    kernels = {}
    if ff == "fast":
        kernels["k_ff_fast"] = {
            "ms": float(np.mean(face_ms)), "units": work["tri_node"] + work["surfel_node"],
            "unit": "element x node",
            "flop": work["tri_node"] * FLOP_PER_TRI_K + work["surfel_node"] * FLOP_PER_DISK_K}
        if work["series_face_dir"] + work["series_surfel_dir"] > 0:
            kernels["k_series_tri / k_series_surfel"] = {
                "ms": float(np.mean(ser_ms)),
                "units": work["series_face_dir"] + work["series_surfel_dir"],
                "unit": "element x direction (+ object x node evaluations)",
                "flop": (work["series_face_dir"] * FLOP_PER_FACE_DIR
                         + work["series_surfel_dir"] * FLOP_PER_SURFEL_DIR
                         + work["series_node"] * FLOP_PER_SERIES_NODE)}
    for k in kernels.values():
        k["tflops"] = k["flop"] / (k["ms"] * 1e-3) / 1e12 if k["ms"] > 0 else 0.0
        k["units_per_s"] = k["units"] / (k["ms"] * 1e-3) if k["ms"] > 0 else 0.0
    for n, k in kernels.items():
        k["peak_tflops"] = KERNEL_PEAK[n]
        k["frac"] = k["tflops"] / KERNEL_PEAK[n]
    out["kernels"] = kernels
    if kernels:
        dom = max(kernels, key=lambda n: kernels[n]["ms"])
        pk = KERNEL_PEAK[dom]
        out["roofline"] = {"bound": "valu", "kernel": dom, "achieved": kernels[dom]["tflops"],
                           "peak": pk, "unit": "TFLOP/s",
                           "frac": kernels[dom]["tflops"] / pk,
                           "traffic": MEASURED_TRAFFIC.get((wl.name, dom))}
    # the Gram against its roofs: algorithmic flops 4 Q per pair over the
    # N (N + 1) / 2 unique pairs (the split's extra products not counted),
    # bytes X read once (8 B per representative node and object) + V written
    gram_flop = 2.0 * kg.Q * N * (N + 1)
    gram_bytes = 8.0 * kg.Qh * N + 8.0 * N * N
    g_s = float(np.mean(gram_ms)) * 1e-3
    out["gram_roofline"] = {"mfma_tflops": gram_flop / g_s / 1e12,
                            "mfma_frac": gram_flop / g_s / 1e12 / PEAK_MFMA_F16_TFLOPS,
                            "hbm_gbps": gram_bytes / g_s / 1e9,
                            "hbm_frac": gram_bytes / g_s / 1e9 / 8000.0,
                            "algorithmic_flop": gram_flop, "algorithmic_bytes": gram_bytes}
    if fp64_ref:
        ctx.set_profiling(True)
        ctx.synchronize()
        r0 = time.perf_counter()
        V64 = step("fp64", "fp64")
        r1 = time.perf_counter()
        t64 = ctx.timings()
        ctx.set_profiling(False)
        d = np.sqrt(np.abs(np.outer(np.diag(V64), np.diag(V64))))
        ovl = np.triu(V64 > 1e-3 * d, 1)

        def errs(Vx):
            return {"max_rel_err_overlapping_pairs":
                    float((np.abs(Vx - V64)[ovl] / np.abs(V64[ovl])).max()) if ovl.any() else 0.0,
                    "max_err_scaled": float((np.abs(Vx - V64) / d).max())}

        # every Gram mode on the same fast form factors: the fp64 Gram shows
        # the form factors' own share of the error
        by_mode = {g: errs(step("fast", g)) for g in ("fp64", "bf16x3", "fp16x3")}
        out.update({"fp64_ms_per_step": (r1 - r0) * 1e3, "fp64_ff_ms": t64["form_factor"],
                    "fp64_gram_ms": t64["gram"],
                    "speedup_vs_fp64_gpu": (r1 - r0) * 1e3 / ms_step,
                    "phase_speedup_vs_fp64": {"form_factor": t64["form_factor"] / out["ff_ms"],
                                              "gram": t64["gram"] / out["gram_ms"]},
                    "overlapping_pairs": int(ovl.sum()),
                    **errs(V), "errors_by_gram_mode": by_mode})
        print(f"[bench] {wl.name} fp64 reference: {out['fp64_ms_per_step']:.2f} ms/step, "
              f"per-pair err {out['max_rel_err_overlapping_pairs']:.2e}", file=sys.stderr,
              flush=True)
        if attribution:
            # per-face kernel on every representative node (series off)
            ctx.set_series(False)
            step(ff, gram)
            ctx.set_profiling(True)
            step(ff, gram)
            tns = ctx.timings()
            ctx.set_profiling(False)
            ctx.set_series(True)
            sym = kg.Q / kg.Qh
            prec = t64["form_factor"] / sym / tns["form_factor"]
            ser = tns["form_factor"] / out["ff_ms"]
            out["form_factor_speedup_attribution"] = {
                "hemisphere_symmetry": sym, "precision_fp32_per_face_vs_fp64": prec,
                "moment_series": ser, "product": sym * prec * ser,
                "measured": t64["form_factor"] / out["ff_ms"],
                "ff_ms_without_series": tns["form_factor"]}
    return out


def scene_geoms_dofs(geoms, sc):
    return [geoms[i] for i in sc.idx], list(sc.dofs)


# The reference's only published timing for this path: its pairwise
# benchmark (two Fibonacci-sphere clouds of 1,000 disks, radius 1, the second
# shifted by d = 0, 0.1, ..., 2.0; Lebedev 29 (302 directions) x 32 radial
# nodes, Q = 9,664; five timed runs per distance after one untimed), median
# time per call with the clock around both geometry packings and the call,
# 5.11 ms for its FP64 CUDA path on an unstated sm_89 GPU.
PUBLISHED_PAIRWISE_MS = 5.11


def measure_reference_pairwise(runs=5):
    """The same call sequence through the reference-named Python module
    (pointint_amd.pointint_cpp): `compute_intersection_volume_cuda` (fast
    form factors, FP64 Gram) and, for the plain FP64 path, the same pair with
    FP64 form factors.  Wall-clock per call, as the reference times it."""
    import pointint_amd.pointint_cpp as pi
    dirs, w = pi.load_lebedev(P.lebedev_path(29))
    Ps, Ns, Rs = pi.create_sphere_pointcloud(1000)
    kg = pi.build_kgrid(dirs, w, 32)
    ctx = P.Context.default()
    fast, plain, err = [], [], 0.0
    for d in np.arange(0, 2.1, 0.1):
        P2 = Ps + np.array([d, 0.0, 0.0])
        vol_gt = pi.sphere_sphere_intersection_volume(np.zeros(3), 1.0, np.array([d, 0.0, 0.0]),
                                                      1.0)
        for timed in [False] + [True] * runs:
            t0 = time.perf_counter()
            g1 = pi.make_point_cloud(Ps, Ns, Rs)
            g2 = pi.make_point_cloud(P2, Ns, Rs)
            v = pi.compute_intersection_volume_cuda(g1, g2, kg)
            t1 = time.perf_counter()
            if timed:
                fast.append(t1 - t0)
        for timed in [False] + [True] * runs:
            t0 = time.perf_counter()
            g1 = pi.make_point_cloud(Ps, Ns, Rs)
            g2 = pi.make_point_cloud(P2, Ns, Rs)
            v64 = float(P.compute_volume_matrix([g1, g2], kg, ff="fp64", gram="fp64")[0, 1])
            t1 = time.perf_counter()
            if timed:
                plain.append(t1 - t0)
        if vol_gt > 1e-10:
            err = max(err, abs(v - v64) / v64)
    ctx.synchronize()
    fast_ms = 1e3 * float(np.median(fast))
    plain_ms = 1e3 * float(np.median(plain))
    return {"workload": "reference pairwise benchmark: 2 x 1,000 disks (Fibonacci spheres, "
                        "radius 1, shifts 0..2), Lebedev 29 x L=32, Q = 9,664",
            "calls": len(fast), "ms_per_call_fast": fast_ms, "ms_per_call_fp64": plain_ms,
            "published_ms_per_call_fp64_sm89": PUBLISHED_PAIRWISE_MS,
            "vs_published_fast": PUBLISHED_PAIRWISE_MS / fast_ms,
            "vs_published_fp64": PUBLISHED_PAIRWISE_MS / plain_ms,
            "max_rel_err_fast_vs_fp64": err,
            "note": "wall clock per call around both geometry packings and the call, as the "
                    "reference times it; the published figure is from other hardware"}


def run_serving(args):
    """The serving workload: a Poisson stream of scene queries against a
    device-resident pool of star-shaped meshes; one step = one segment of the
    stream (--segment-s seconds of arrivals); the warm-up segments' queries
    are served but not counted."""
    from pointint_amd import serving as S
    pool = W.serving_pool(n_mesh=args.pool, seed=5)
    print(f"[bench] pool of {len(pool)} meshes, {sum(m.n for m in pool)} triangles",
          file=sys.stderr, flush=True)
    ctx = P.Context.default()
    kg = P.kgrid_from_rule(59, 64)
    geoms = [P.make_triangle_mesh(m.V, m.F) for m in pool]
    # geometry preprocessing of the resident pool (before any query): the
    # spherical-harmonic form of each mesh's moment series
    ctx.synchronize()
    t0 = time.perf_counter()
    if not args.no_moment_cache:
        P.cache_moments(geoms)
    cache_s = time.perf_counter() - t0
    print(f"[bench] moment cache of the pool: {cache_s:.2f} s", file=sys.stderr, flush=True)
    seg = args.segment_s
    duration = (args.warmup + args.steps) * seg
    with S.Server(kg, gram=args.gram, hbm_budget=args.budget_gb * 1e9,
                  max_batch_objects=args.max_batch, latency_target_ms=args.latency_target_ms,
                  graphs=not args.no_graphs) as srv:
        rep = S.poisson_load(srv, geoms, rate=args.rate, duration=duration, kmin=8, kmax=512,
                             box=4.0, seed=11, keep=args.sample_scenes)
        st = srv.stats()
    # the timed window: arrivals in the last `steps` segments
    ws, we = args.warmup * seg, duration
    rec = [r for r in rep.records if ws <= r[0] < we]
    done = [r for r in rep.records if not np.isnan(r[1])]
    in_window = [r for r in done if ws <= r[1] < we]
    pairs_w = sum(r[2] * (r[2] - 1) // 2 for r in in_window)
    tris_w = sum(r[3] for r in in_window)
    lat = np.array([(r[1] - r[0]) * 1e3 for r in rec if not np.isnan(r[1])])
    # arrival -> launch of its batch, and launch -> completion
    qlat = np.array([(r[4] - r[0]) * 1e3 for r in rec if not np.isnan(r[1])])
    blat = np.array([(r[1] - r[4]) * 1e3 for r in rec if not np.isnan(r[1])])
    value = pairs_w / (we - ws)
    n_adm = sum(1 for r in rec if not np.isnan(r[1]))
    print(f"[bench] serving: {value:.3e} pairs/s, p50 {np.percentile(lat, 50):.1f} ms, "
          f"p99 {np.percentile(lat, 99):.1f} ms, {n_adm}/{len(rec)} admitted in the window",
          file=sys.stderr, flush=True)
    # form-factor kernels of the served batches (device-clock stamps around
    # each launch on the batch's stream; two batches in flight share the GPU)
    ff_flop = st["tri_node"] * FLOP_PER_TRI_K
    ff_tflops = ff_flop / (st["ff_ms"] * 1e-3) / 1e12 if st["ff_ms"] > 0 else 0.0
    roofline = {"bound": "valu", "kernel": "k_ff_fast", "achieved": ff_tflops,
                "peak": PEAK_FP32_VALU_TFLOPS, "unit": "TFLOP/s",
                "frac": ff_tflops / PEAK_FP32_VALU_TFLOPS,
                "traffic": MEASURED_TRAFFIC.get(("serving", "k_ff_fast")),
                "kernel_ms_per_batch": st["ff_ms"] / max(st["batches"], 1),
                "units": st["tri_node"], "unit_of_work": "triangle x node",
                "batches": st["batches"],
                "phase_ms_total": {"prepass": st["prepass_ms"], "series": st["series_ms"],
                                   "per_face": st["ff_ms"], "gram": st["gram_ms"]}}
    # the plain FP64 GPU path on a sample of the served scenes: speed and error
    vs64 = None
    scenes = rep.scenes
    if scenes and not args.no_fp64_ref:
        t64, tfast, pairs64, worst, npairs = 0.0, 0.0, 0, 0.0, 0
        fit_rows, fit_t = [], []
        for sc in scenes:
            g, d = scene_geoms_dofs(geoms, sc)
            k = len(g)
            ctx.synchronize()
            r0 = time.perf_counter()
            V64 = P.compute_volume_matrix(g, kg, d, ["affine"] * k, ff="fp64", gram="fp64")
            dt64 = time.perf_counter() - r0
            t64 += dt64
            fit_rows.append([sum(pool[i].n for i in sc.idx), k * (k + 1) / 2, 1.0])
            fit_t.append(dt64)
            # the same scene alone through the one-call fast path (warm)
            P.compute_volume_matrix(g, kg, d, ["affine"] * k, ff="fast", gram=args.gram)
            r0 = time.perf_counter()
            P.compute_volume_matrix(g, kg, d, ["affine"] * k, ff="fast", gram=args.gram)
            tfast += time.perf_counter() - r0
            pairs64 += k * (k - 1) // 2
            dd = np.sqrt(np.abs(np.outer(np.diag(V64), np.diag(V64))))
            ovl = np.triu(V64 > 1e-3 * dd, 1)
            npairs += int(ovl.sum())
            if ovl.any():
                worst = max(worst, float((np.abs(sc.V - V64)[ovl] / np.abs(V64[ovl])).max()))
        # The FP64 path's time for every scene served in the window, from a
        # least-squares fit of the sampled scenes' times to
        # t = a * triangles + b * K(K+1)/2 + c (its form factors and Gram):
        # the pairs of a uniform sample are dominated by the few largest
        # scenes drawn, so the sample's own ratio varies from run to run.
        A = np.array(fit_rows)
        tt = np.array(fit_t)
        # non-negative least squares in relative terms (each scene's residual
        # over its time)
        from scipy.optimize import nnls
        coef = nnls(A / tt[:, None], np.ones_like(tt))[0]
        rel = np.abs(A @ coef - tt) / tt
        resid = float(np.max(rel))
        t64_window = float(sum(coef[0] * r[3] + coef[1] * r[2] * (r[2] + 1) / 2 + coef[2]
                               for r in in_window))
        fp64_rate = pairs_w / t64_window
        vs64 = {"fp64_pairs_per_s": fp64_rate, "speedup": value / fp64_rate,
                "fp64_s_window": t64_window,
                "fp64_time_fit": {"per_triangle_s": float(coef[0]),
                                  "per_pair_incl_diag_s": float(coef[1]),
                                  "per_scene_s": float(coef[2]),
                                  "max_rel_residual": resid,
                                  "median_rel_residual": float(np.median(rel))},
                "scenes": len(scenes), "scene_objects": [len(sc.idx) for sc in scenes],
                "sample_fp64_pairs_per_s": pairs64 / t64, "sample_fp64_s": t64,
                "sample_speedup": value / (pairs64 / t64),
                "one_call_fast_pairs_per_s": pairs64 / tfast,
                "one_call_fast_speedup": t64 / tfast,
                "max_rel_err_overlapping_pairs": worst, "overlapping_pairs": npairs}
        print(f"[bench] FP64 GPU path: {fp64_rate:.3e} pairs/s on the window's served scenes "
              f"(fit to {len(scenes)} timed scenes, residual {resid:.1%}); x{vs64['speedup']:.1f}; "
              f"sample alone x{vs64['sample_speedup']:.1f}; per-pair err {worst:.2e}",
              file=sys.stderr, flush=True)
    # The same kernel on one batch alone: the served batches share the
    # device two at a time, so the spans above include the other batch's
    # kernels; here the sampled scenes' objects (up to the mean batch size)
    # run through the one-call fast path, timed by HIP events.
    if scenes:
        mean_batch = st["objects"] / max(st["batches"], 1)
        bg, bd = [], []
        for sc in scenes:
            g, d = scene_geoms_dofs(geoms, sc)
            bg += g
            bd += d
            if len(bg) >= mean_batch:
                break
        bg, bd = bg[:args.max_batch], bd[:args.max_batch]
        P.compute_volume_matrix(bg, kg, bd, ["affine"] * len(bg), ff="fast", gram=args.gram)
        ctx.set_profiling(True)
        face = []
        for _ in range(3):
            P.compute_volume_matrix(bg, kg, bd, ["affine"] * len(bg), ff="fast", gram=args.gram)
            face.append(ctx.timings()["ff_per_face"])
        wk = ctx.work()
        ctx.set_profiling(False)
        units1 = wk["tri_node"]  # the last call's
        tf1 = units1 * FLOP_PER_TRI_K / (float(np.mean(face)) * 1e-3) / 1e12
        roofline["alone_batch"] = {"objects": len(bg), "kernel_ms": float(np.mean(face)),
                                   "units": units1, "achieved": tf1,
                                   "frac": tf1 / PEAK_FP32_VALU_TFLOPS}
    cpu = None
    if scenes and not args.no_cpu:
        cpu = cpu_baseline_scenes(pool, kg, scenes, args.cpu_seconds)
    extras = {}
    if not args.no_extras:
        extras["reference_pairwise_benchmark"] = measure_reference_pairwise()
        extras["icospheres_256_L3"] = measure_matrix(
            W.config_icospheres(n_obj=256, level=3, seed=2), "bf16x3", 10, 3, attribution=True)
        extras["mixed_1024"] = measure_matrix(W.config_mixed(), "fp16x3", 3, 1)
    elem_q = np.mean([pool[i].n for sc in scenes for i in sc.idx]) if scenes else float("nan")
    out = {
        "metric": "overlap pairs/s & prim*k evals/s vs plain FP64 at 1e-5 rel err; p50/p99 latency",
        "value": value, "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": seg * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic (seeded star-shaped meshes, random rigid transforms, Poisson arrivals)",
        "config": {"workload": f"serving: pool of {len(pool)} star meshes (T log-uniform "
                               f"[80, 20480]), scenes of K log-uniform [8, 512] objects under "
                               f"random rigid transforms (centres U[0,4]^3), Poisson "
                               f"{args.rate:g} queries/s, Lebedev 1202 x L=64 (Q=76,928)",
                   "gram": args.gram, "hbm_budget_gb": args.budget_gb,
                   "latency_target_ms": args.latency_target_ms,
                   "max_batch_objects": args.max_batch, "graphs": not args.no_graphs,
                   "moment_cache": not args.no_moment_cache,
                   "step": f"{seg:g} s of arrivals", "parallelism": "single GPU"},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "mean": float(lat.mean()), "max": float(lat.max()),
                       "from": "scheduled arrival",
                       "queued_p50": float(np.percentile(qlat, 50)),
                       "queued_p99": float(np.percentile(qlat, 99)),
                       "in_batch_p50": float(np.percentile(blat, 50)),
                       "in_batch_p99": float(np.percentile(blat, 99))},
        "queries": {"submitted": st["submitted"], "admitted": st["admitted"],
                    "refused_memory": st["refused_memory"],
                    "refused_latency_target": st["refused_latency"],
                    "completed": st["completed"], "window_arrivals": len(rec),
                    "window_admitted": n_adm, "window_completed": len(in_window)},
        "completed_qps": len(in_window) / (we - ws),
        "pool_setup_s": {"moment_cache": cache_s},
        "generator_lag_p99_ms": rep.summary()["submit_lag_p99_ms"],
        # triangles x full-grid nodes of the completed queries (the plain
        # path's evaluations), per second of the window: a nominal rate
        "prim_k_per_s_nominal": tris_w * kg.Q / (we - ws),
        # evaluations the fast path made (whole run): per-face triangle x
        # node and series triangle x direction
        "evaluated_per_s": {"tri_node": st["tri_node"] / rep.wall_s,
                            "series_face_dir": st["face_dir"] / rep.wall_s},
        "mean_triangles_per_object": float(elem_q),
        "batches": st["batches"], "graph_launches": st["graph_launches"],
        "graphs_captured": st["graphs_captured"],
        "objects_per_batch": st["objects"] / max(st["batches"], 1),
        "roofline": roofline,
        "vs_fp64_gpu": vs64,
        "cpu_baseline": cpu,
        "extras": extras,
    }
    print(json.dumps(out))


def cpu_baseline_scenes(pool, kg, scenes, target_seconds):
    """FP64 CPU oracle (the restated reference CPU path) on the served scenes:
    every scene on a strided k-node subsample sized so that the whole sample
    takes about target_seconds, extrapolated linearly in Q."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    og = {}

    def run(stride):
        t = 0.0
        for sc in scenes:
            gs = []
            for i in sc.idx:
                if i not in og:
                    og[i] = O.geom_mesh(pool[i].V, pool[i].F)
                gs.append(og[i])
            idx = np.arange(0, kg.Q, stride)
            t0 = time.perf_counter()
            O.volume_matrix(gs, kg.dirs[idx], kg.kmag[idx], kg.w[idx], list(sc.dofs),
                            [O.DOF_AFFINE] * len(gs))
            t += (time.perf_counter() - t0) * kg.Q / len(idx)
        return t

    cal_stride = max(1, kg.Q // 32)
    t0 = time.perf_counter()
    run(cal_stride)
    t_cal = time.perf_counter() - t0
    stride = max(1, int(cal_stride * t_cal / max(target_seconds, 1.0)))
    t0 = time.perf_counter()
    t_full = run(stride)
    t_s = time.perf_counter() - t0
    pairs = sum(len(sc.idx) * (len(sc.idx) - 1) // 2 for sc in scenes)
    print(f"[bench] cpu oracle: {t_full:.1f} s for the {len(scenes)} scenes, extrapolated",
          file=sys.stderr, flush=True)
    return {"value": pairs / t_full, "unit": "pairs/s", "cores": O.threads(), "kind": "port",
            "sample": f"FP64 oracle on the {len(scenes)} sampled served scenes "
                      f"({sum(len(sc.idx) for sc in scenes)} objects), every {stride}th of "
                      f"{kg.Q} k-nodes, {t_s:.1f} s, extrapolated x{stride}; host {cpu_model()}"}


def run_matrix(args):
    """One fixed workload (icospheres: 256 icospheres L3 on Lebedev 1202 x
    L=64; mixed: 1,024 mixed objects on Lebedev 2702 x L=96), one JSON line."""
    if args.workload == "mixed":
        wl = W.config_mixed()
        if "--gram" not in sys.argv:
            args.gram = "fp16x3"
    else:
        wl = W.config_icospheres(n_obj=args.n_obj, level=3, seed=2)
    m = measure_matrix(wl, args.gram, args.steps, args.warmup, ff=args.ff,
                       fp64_ref=not args.no_fp64_ref, attribution=args.workload == "icospheres")
    cpu = None
    if not args.no_cpu:
        kg = P.kgrid_from_rule(wl.lebedev_order, wl.n_radial)
        N = len(wl.objects)
        t_full, nq, t_s, cores = cpu_baseline(wl, kg, args.cpu_seconds)
        cpu = {"value": N * (N - 1) // 2 / t_full, "unit": "pairs/s", "cores": cores,
               "kind": "port",
               "sample": f"FP64 oracle on {nq} of {kg.Q} k-nodes (every {round(kg.Q / nq)}th), "
                         f"all {N} objects, {t_s:.1f} s, extrapolated x{kg.Q / nq:.1f}; "
                         f"host {cpu_model()}"}
    out = {"metric": "overlap pairs/s & prim*k evals/s vs plain FP64 at 1e-5 rel err; "
                     "p50/p99 latency",
           "value": m["pairs_per_s"], "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": m["ms_per_step"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None,
           "dtype": "f32+f64" if args.ff == "fast" else "f64", "data": "synthetic (seeded)",
           "config": {"workload": wl.name, "objects": m["objects"], "elements": m["elements"],
                      "Q": m["Q"], "ff": args.ff, "gram": args.gram,
                      "parallelism": "single GPU"},
           "roofline": m.get("roofline"), "cpu_baseline": cpu}
    out.update({k: v for k, v in m.items() if k not in ("roofline",)})
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--ff", default="fast", choices=["fast", "fp64"])
    ap.add_argument("--gram", default="bf16x3", choices=["fp64", "bf16x3", "fp16x3"])
    ap.add_argument("--no-fp64-ref", action="store_true", help="skip the plain FP64 GPU run")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extras", action="store_true",
                    help="serving: skip the fixed-workload measurements")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--n-obj", type=int, default=256)
    ap.add_argument("--workload", default="serving",
                    choices=["serving", "icospheres", "mixed", "diff", "align"])
    ap.add_argument("--fd-vertices", type=int, default=200)
    ap.add_argument("--rate", type=float, default=2000.0)
    ap.add_argument("--segment-s", type=float, default=0.5,
                    help="serving: seconds of arrivals per step")
    ap.add_argument("--pool", type=int, default=4096)
    ap.add_argument("--budget-gb", type=float, default=64.0)
    ap.add_argument("--latency-target-ms", type=float, default=100.0)
    ap.add_argument("--max-batch", type=int, default=1024)
    ap.add_argument("--no-graphs", action="store_true")
    ap.add_argument("--no-moment-cache", action="store_true",
                    help="serve without the pool's spherical-harmonic moment cache")
    ap.add_argument("--sample-scenes", type=int, default=64)
    ap.add_argument("--copy-calibration", type=float, default=0.0,
                    help="GiB copied device-to-device before the run (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the path "
              "would shard k-nodes across GPUs with one all-reduce of V, which this build "
              "does not run.", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration > 0:
        copy_calibration(int(args.copy_calibration * (1 << 30)))
    if args.workload == "diff":
        return run_diff(args)
    if args.workload == "align":
        return run_align(args)
    if args.workload == "serving":
        return run_serving(args)
    return run_matrix(args)


if __name__ == "__main__":
    main()
