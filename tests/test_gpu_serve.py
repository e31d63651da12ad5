"""GPU: the request loop (continuous batching, warm start, admission).

A loop's results must be the same, bit for bit, whether its requests run
alone or batched with other loops' requests (every system is computed
independently of its batch neighbours), and the warm-started 10-iteration
steps must follow the FP64 oracle started from its own previous solution.
"""
import numpy as np
import pytest

from maos_amd import _abi, configs, tomo, workload
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

NSTEP = 4


@pytest.fixture(scope="module")
def loops():
    out = []
    for k, mk in enumerate([configs.small8, lambda: configs.batch_member(8.0, 2, 4), configs.lgs8_flat]):
        cfg = mk()
        r = tomo.Recon(cfg)
        out.append((cfg, r, workload.frozen_flow_stream(cfg, r, 40 + k, NSTEP)))
    return out


def run(loops, which, mix):
    """Steps of the loops `which` (all submitted together per step when mix)."""
    srv = tomo.Server(mem_budget_gb=8.0, max_inflight=2)
    ids = [srv.add_loop(loops[i][1], maxit=10) for i in which]
    xs = {i: [] for i in which}
    for k in range(NSTEP):
        if mix:
            for i, lid in zip(which, ids):
                srv.submit(lid, loops[i][2][k], tag=k)
            srv.drain()
        else:
            for i, lid in zip(which, ids):
                srv.submit(lid, loops[i][2][k], tag=k)
                srv.drain()
        for i, lid in zip(which, ids):
            xs[i].append(srv.get_x(lid)[0])
    done = srv.poll()
    assert len(done) == NSTEP * len(which)
    srv.close()
    return xs


def test_alone_equals_mix(loops):
    mixed = run(loops, [0, 1, 2], mix=True)
    for i in range(3):
        alone = run(loops, [i], mix=False)
        for k in range(NSTEP):
            assert np.array_equal(alone[i][k], mixed[i][k]), "loop %d step %d" % (i, k)


def test_warm_steps_vs_oracle(loops):
    cfg, r, grads = loops[1]
    srv = tomo.Server(mem_budget_gb=8.0)
    lid = srv.add_loop(r, maxit=10)
    o = Oracle(r, nthread=8)
    xo = None
    for k in range(NSTEP):
        srv.submit(lid, grads[k], tag=k)
        srv.drain()
        x, res = srv.get_x(lid)
        xo, reso = o.solve(grads[k], 10, x0=xo)
        e = float(np.linalg.norm(x - xo) / np.linalg.norm(xo))
        print("step %d: rel L2 %.3e residual gpu %.4e oracle %.4e" % (k, e, res, reso))
        assert e <= 1e-4
    srv.close()


def test_admission_by_memory(loops):
    cfg, r, grads = loops[0]
    srv = tomo.Server(mem_budget_gb=0.05)
    a = srv.add_loop(r)
    used, budget = srv.usage()
    assert 0 < used <= budget
    n = 1
    with pytest.raises(_abi.MtError) as ei:
        while n < 1000:
            srv.add_loop(r)
            n += 1
    assert ei.value.code == -6  # MT_EFULL
    srv.remove_loop(a)
    srv.add_loop(r)  # admitted again once memory is free
    srv.close()


def test_autotune_between_steps(loops):
    """A server autotuned between steps keeps each loop's warm-start
    solution, and each loop's steps are those of its plan alone (warm start)
    switched to the same launch shape at the same step, bit for bit."""
    which = [0, 1, 2]
    srv = tomo.Server(mem_budget_gb=8.0, max_inflight=2)
    with pytest.raises(_abi.MtError):
        srv.autotune()  # no loops yet
    ids = [srv.add_loop(loops[i][1], maxit=10) for i in which]
    xs = {i: [] for i in which}
    choice = None
    for k in range(NSTEP):
        if k == 2:
            choice = srv.autotune()
        for i, lid in zip(which, ids):
            srv.submit(lid, loops[i][2][k], tag=k)
        srv.drain()
        for i, lid in zip(which, ids):
            xs[i].append(srv.get_x(lid)[0])
    srv.close()
    print("server autotune choice", choice)
    for i in which:
        g = tomo.GpuTomo(loops[i][1], maxit=10, warm=1)
        g.tune_set(0, 0, 0, 0, 2, 0)  # a server loop's shape before tuning
        for k in range(NSTEP):
            if k == 2:
                g.tune_set(*choice[:5], 0)
            x = g.tomo(loops[i][2][k])[0]
            assert np.array_equal(x, xs[i][k]), "loop %d step %d" % (i, k)
        g.close()
