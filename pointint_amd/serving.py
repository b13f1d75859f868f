"""Serving front end: continuous batching of volume-matrix queries.

The engine is native (pointint_amd/csrc/serve.cpp, C ABI pii_server_*): a
worker thread keeps two batches in flight on two HIP streams, admits queries
against an HBM budget and answers each with its n x n volume matrix.  This
module wraps it and provides a Poisson load generator that reports p50/p99
latency (completion - scheduled arrival) beside throughput.
"""
from __future__ import annotations

import ctypes as C
import math
import threading
import time
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import _native as N
from . import core
from . import workloads as W

_GRAM = {"fp64": N.GRAM_FP64, "bf16x3": N.GRAM_BF16X3, "fp16x3": N.GRAM_FP16X3}


def clock() -> float:
    """Seconds on the serving engine's clock (arrival times are on it)."""
    t = C.c_double()
    N.check(N.load().pii_clock_seconds(C.byref(t)))
    return t.value


STAT_KEYS = ["submitted", "admitted", "refused", "completed", "batches", "objects",
             "reserved_bytes", "budget_bytes", "refused_memory", "refused_latency",
             "graph_launches", "graphs_captured", "prepass_ms", "series_ms", "ff_ms",
             "gram_ms", "tri_node", "face_dir", "ms_per_triangle"]


class Server:
    """A serving engine over resident geometry (pii_server_*).

    hbm_budget: bytes the admitted, unfinished queries may reserve.
    latency_target_ms: > 0 also refuses queries whose predicted completion
    (from their arrival) exceeds it.  graphs: replay the batch launch
    sequence from captured hipGraphs (False launches it directly; the
    results are identical)."""

    def __init__(self, kgrid: core.KGrid, gram: str = "fp16x3", hbm_budget: float = 8e9,
                 max_batch_objects: int = 1024, latency_target_ms: float = 0.0,
                 graphs: bool = True):
        self.kgrid = kgrid
        self.ctx = kgrid.ctx
        h = C.c_void_p()
        N.check(N.load().pii_server_create(self.ctx.handle, kgrid.handle, _GRAM[gram],
                                           float(hbm_budget), int(max_batch_objects),
                                           C.byref(h)))
        self._h = h
        if latency_target_ms:
            N.check(N.load().pii_server_set_latency_target(self._h, float(latency_target_ms)))
        if not graphs:
            N.check(N.load().pii_server_set_graphs(self._h, 0))

    def submit(self, geoms: Sequence[core.Geometry], dofs,
               t_arrival: float = 0.0) -> Optional[int]:
        """Queue a scene (arrival time t_arrival on :func:`clock`, 0: now);
        returns its id, or None when refused (memory or latency target)."""
        n = len(geoms)
        d = N.f64(dofs).reshape(n, 12)
        gh = np.fromiter((g._hv for g in geoms), dtype=np.uintp, count=n)
        qid = C.c_int64()
        rc = N.load().pii_server_submit_at(self._h, n, gh.ctypes.data, N.ptr(d),
                                           float(t_arrival), C.byref(qid))
        if rc == N.PII_ERR_BUDGET:
            return None
        N.check(rc)
        return qid.value

    def wait(self, qid: int, n: int, timeout: float = -1.0):
        """(V (n x n), latency ms from arrival) of a finished query; releases it."""
        V, t = self.wait_times(qid, n, timeout)
        return V, (t[3] - t[0]) * 1e3

    def wait_times(self, qid: int, n: int, timeout: float = -1.0):
        """(V, [arrival, submission, launch, completion] seconds on :func:`clock`)."""
        V = np.empty((n, n))
        t = np.zeros(4)
        N.check(N.load().pii_server_wait_times(self._h, C.c_int64(qid), float(timeout),
                                               N.ptr(V), N.ptr(t)))
        return V, t

    def stats(self) -> dict:
        s = np.zeros(len(STAT_KEYS))
        N.check(N.load().pii_server_stats(self._h, N.ptr(s), len(STAT_KEYS)))
        return {k: float(v) for k, v in zip(STAT_KEYS, s)}

    def close(self) -> None:
        if getattr(self, "_h", None):
            N.check(N.load().pii_server_destroy(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rigid_dofs(rng, box: float, k: int = 1) -> np.ndarray:
    """(k, 12) affine DoFs: uniform random rotations (normalised Gaussian
    quaternions) and translations U[0, box]^3."""
    q = rng.normal(size=(k, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
                  2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
                  2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], axis=1)
    t = rng.uniform(0.0, box, size=(k, 3))
    return np.concatenate([t, R], axis=1)


@dataclass
class Scene:
    """One completed query of a load run: pool indices, DoFs, result, times."""
    idx: np.ndarray
    dofs: np.ndarray
    V: np.ndarray
    latency_ms: float
    queue_ms: float


@dataclass
class LoadReport:
    offered_rate: float
    duration_s: float
    submitted: int = 0
    admitted: int = 0
    refused: int = 0
    completed: int = 0
    pairs_completed: int = 0
    wall_s: float = 0.0
    latencies_ms: List[float] = field(default_factory=list)
    queue_ms: List[float] = field(default_factory=list)
    submit_lag_ms: List[float] = field(default_factory=list)
    scenes: List[Scene] = field(default_factory=list)
    # per submitted query: (arrival, completion or nan when refused, K,
    # triangles); seconds from the start of the stream
    records: List[tuple] = field(default_factory=list)

    def summary(self) -> dict:
        lat = np.array(self.latencies_ms) if self.latencies_ms else np.array([float("nan")])
        lag = np.array(self.submit_lag_ms) if self.submit_lag_ms else np.array([0.0])
        qd = np.array(self.queue_ms) if self.queue_ms else np.array([float("nan")])
        return {
            "offered_qps": self.offered_rate,
            "duration_s": self.duration_s,
            "submitted": self.submitted,
            "admitted": self.admitted,
            "refused": self.refused,
            "completed": self.completed,
            "completed_qps": self.completed / self.wall_s if self.wall_s else 0.0,
            "pairs_per_s": self.pairs_completed / self.wall_s if self.wall_s else 0.0,
            "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()),
            "max_ms": float(lat.max()),
            "queue_p50_ms": float(np.percentile(qd, 50)),
            "submit_lag_p99_ms": float(np.percentile(lag, 99)),
            "submit_lag_max_ms": float(lag.max()),
        }


def poisson_load(server: Server, pool: Sequence[core.Geometry], rate: float, duration: float,
                 kmin: int = 8, kmax: int = 512, box: float = 4.0, seed: int = 0,
                 keep: int = 0) -> LoadReport:
    """Poisson arrivals at `rate` queries/s for `duration` seconds; each query
    is K (log-uniform in [kmin, kmax]) pool objects under fresh rigid
    transforms.  Arrival times are scheduled in advance on the engine's clock
    and passed with each submission, so latencies (completion - scheduled
    arrival) include any lag of this generator.  Submission runs on this
    thread, collection on another.  `keep` completed queries, a uniform
    sample of all of them, keep their scene and result (LoadReport.scenes)."""
    rng = np.random.default_rng(seed)
    pick = np.random.default_rng(seed + 1)
    n_arr = int(rate * duration * 1.2) + 16
    gaps = rng.exponential(1.0 / rate, size=n_arr)
    arrivals = np.cumsum(gaps)
    arrivals = arrivals[arrivals < duration]
    rep = LoadReport(rate, duration)
    pending = []
    lock = threading.Lock()
    done_flag = threading.Event()

    def collector():
        while True:
            with lock:
                item = pending.pop(0) if pending else None
            if item is None:
                if done_flag.is_set():
                    with lock:
                        if not pending:
                            return
                    continue
                time.sleep(0.0005)
                continue
            qid, k, idx, dofs, ta, tris = item
            V, t = server.wait_times(qid, k)
            rep.latencies_ms.append((t[3] - t[0]) * 1e3)
            rep.queue_ms.append((t[2] - t[0]) * 1e3)
            rep.records.append((ta, t[3] - t0, k, tris, t[2] - t0))
            rep.completed += 1
            rep.pairs_completed += k * (k - 1) // 2
            if keep > 0:
                # reservoir sample of the completed queries
                sc = Scene(idx, dofs, V, (t[3] - t[0]) * 1e3, (t[2] - t[0]) * 1e3)
                if len(rep.scenes) < keep:
                    rep.scenes.append(sc)
                else:
                    j = int(pick.integers(rep.completed))
                    if j < keep:
                        rep.scenes[j] = sc

    th = threading.Thread(target=collector, daemon=True)
    th.start()
    t0 = clock() + 0.01
    w0 = time.perf_counter()
    for ta in arrivals:
        k = int(W.log_uniform_int(rng, kmin, kmax))
        idx = rng.integers(0, len(pool), size=k)
        dofs = rigid_dofs(rng, box, k)
        geoms = [pool[i] for i in idx]
        tris = sum(g.num_elements() for g in geoms)
        t_sched = t0 + ta
        delay = t_sched - clock()
        if delay > 0:
            time.sleep(delay)
        rep.submitted += 1
        qid = server.submit(geoms, dofs, t_arrival=t_sched)
        rep.submit_lag_ms.append(max(0.0, clock() - t_sched) * 1e3)
        if qid is None:
            rep.refused += 1
            rep.records.append((ta, float("nan"), k, tris, float("nan")))
            continue
        rep.admitted += 1
        with lock:
            pending.append((qid, k, idx, dofs, ta, tris))
    done_flag.set()
    th.join()
    rep.wall_s = time.perf_counter() - w0
    return rep
