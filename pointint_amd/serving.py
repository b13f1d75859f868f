"""Serving front end: continuous batching of volume-matrix queries.

The engine is native (pointint_amd/csrc/serve.cpp, C ABI pii_server_*): a
worker thread keeps two batches in flight on two HIP streams, admits queries
against an HBM budget and answers each with its n x n volume matrix.  This
module wraps it and provides a Poisson load generator that reports p50/p99
latency (completion - submission) beside throughput.
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


class Server:
    def __init__(self, kgrid: core.KGrid, gram: str = "fp16x3", hbm_budget: float = 8e9,
                 max_batch_objects: int = 1024):
        self.kgrid = kgrid
        self.ctx = kgrid.ctx
        h = C.c_void_p()
        N.check(N.load().pii_server_create(self.ctx.handle, kgrid.handle, _GRAM[gram],
                                           float(hbm_budget), int(max_batch_objects),
                                           C.byref(h)))
        self._h = h

    def submit(self, geoms: Sequence[core.Geometry], dofs) -> Optional[int]:
        """Queue a scene; returns its id, or None when refused by the budget."""
        n = len(geoms)
        d = N.f64(dofs).reshape(n, 12)
        gh = np.fromiter((g._hv for g in geoms), dtype=np.uintp, count=n)
        qid = C.c_int64()
        rc = N.load().pii_server_submit(self._h, n, gh.ctypes.data, N.ptr(d), C.byref(qid))
        if rc == N.PII_ERR_BUDGET:
            return None
        N.check(rc)
        return qid.value

    def wait(self, qid: int, n: int, timeout: float = -1.0):
        """(V (n x n), latency ms) of a finished query; releases it."""
        V = np.empty((n, n))
        lat = C.c_double()
        N.check(N.load().pii_server_wait(self._h, C.c_int64(qid), float(timeout), N.ptr(V),
                                         C.byref(lat)))
        return V, lat.value

    def stats(self) -> dict:
        s = np.zeros(8)
        N.check(N.load().pii_server_stats(self._h, N.ptr(s), 8))
        keys = ["submitted", "admitted", "refused", "completed", "batches", "objects",
                "reserved_bytes", "budget_bytes"]
        return {k: float(v) for k, v in zip(keys, s)}

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

    def summary(self) -> dict:
        lat = np.array(self.latencies_ms) if self.latencies_ms else np.array([float("nan")])
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
        }


def poisson_load(server: Server, pool: Sequence[core.Geometry], rate: float, duration: float,
                 kmin: int = 8, kmax: int = 512, box: float = 4.0, seed: int = 0) -> LoadReport:
    """Poisson arrivals at `rate` queries/s for `duration` seconds; each query
    is K (log-uniform in [kmin, kmax]) pool objects under fresh rigid
    transforms.  Submission runs on this thread, collection on another."""
    rng = np.random.default_rng(seed)
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
            qid, k = item
            V, lat = server.wait(qid, k)
            rep.latencies_ms.append(lat)
            rep.completed += 1
            rep.pairs_completed += k * (k - 1) // 2

    th = threading.Thread(target=collector, daemon=True)
    th.start()
    t0 = time.perf_counter()
    for ta in arrivals:
        k = int(W.log_uniform_int(rng, kmin, kmax))
        idx = rng.integers(0, len(pool), size=k)
        dofs = rigid_dofs(rng, box, k)
        geoms = [pool[i] for i in idx]
        delay = t0 + ta - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        rep.submitted += 1
        qid = server.submit(geoms, dofs)
        if qid is None:
            rep.refused += 1
            continue
        rep.admitted += 1
        with lock:
            pending.append((qid, k))
    done_flag.set()
    th.join()
    rep.wall_s = time.perf_counter() - t0
    return rep
