"""GPU serving engine: each query's volume matrix is the same alone and in a
mix of concurrent queries, matches the FP64 oracle, and the HBM budget
refuses what does not fit."""
import numpy as np
import pytest

import pointint_amd as P
from pointint_amd import serving as S
from pointint_amd import workloads as W

pytestmark = pytest.mark.gpu


def scaled(V, Vr):
    d = np.sqrt(np.abs(np.outer(np.diag(Vr), np.diag(Vr))))
    return float(np.max(np.abs(V - Vr) / d))


@pytest.fixture(scope="module")
def pool():
    return W.serving_pool(n_mesh=12, seed=21, scale=0.05)


def test_alone_equals_in_mix_and_oracle(oracle, pool):
    kg = P.kgrid_from_rule(29, 16)
    geoms = [P.make_triangle_mesh(m.V, m.F) for m in pool]
    rng = np.random.default_rng(3)
    idx = rng.integers(0, len(pool), size=7)
    dofs = S.rigid_dofs(rng, 0.8, 7)
    with S.Server(kg, gram="fp16x3", hbm_budget=4e9) as srv:
        qid = srv.submit([geoms[i] for i in idx], dofs)
        V_alone, lat = srv.wait(qid, 7)
        assert lat > 0
        # the same query in the middle of 24 concurrent ones
        others = []
        for j in range(24):
            k = int(rng.integers(2, 20))
            oi = rng.integers(0, len(pool), size=k)
            others.append((srv.submit([geoms[i] for i in oi], S.rigid_dofs(rng, 1.0, k)), k))
            if j == 11:
                qmix = srv.submit([geoms[i] for i in idx], dofs)
        V_mix, _ = srv.wait(qmix, 7)
        for q, k in others:
            srv.wait(q, k)
        st = srv.stats()
        assert st["completed"] == 26 and st["refused"] == 0 and st["reserved_bytes"] == 0
    assert np.array_equal(V_alone, V_mix)
    ogeoms = [oracle.geom_mesh(pool[i].V, pool[i].F) for i in idx]
    Jo = np.stack([oracle.compute_A_grid(g, kg.dirs, kg.kmag, d, oracle.DOF_AFFINE)
                   for g, d in zip(ogeoms, dofs)], axis=1)
    Vo = oracle.gram(Jo, kg.w)
    assert scaled(V_alone, Vo) < 1e-5


def test_budget_refusal(pool):
    kg = P.kgrid_from_rule(29, 16)
    geoms = [P.make_triangle_mesh(m.V, m.F) for m in pool]
    rng = np.random.default_rng(4)
    with S.Server(kg, hbm_budget=2e6) as srv:   # room for small scenes only
        assert srv.submit(geoms[:2], S.rigid_dofs(rng, 1.0, 2)) is not None
        big = srv.submit([geoms[i % 12] for i in range(200)], S.rigid_dofs(rng, 1.0, 200))
        assert big is None
        st = srv.stats()
        assert st["refused"] == 1


def test_poisson_load_reports_latency(pool):
    kg = P.kgrid_from_rule(29, 16)
    geoms = [P.make_triangle_mesh(m.V, m.F) for m in pool]
    with S.Server(kg, hbm_budget=2e9) as srv:
        rep = S.poisson_load(srv, geoms, rate=200.0, duration=1.0, kmin=4, kmax=32, box=2.0)
    s = rep.summary()
    assert s["completed"] == s["admitted"] > 0
    assert s["p99_ms"] >= s["p50_ms"] > 0
