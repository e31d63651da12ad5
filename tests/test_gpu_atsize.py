"""GPU: the batched and served workloads at their configured sizes.

  uneven 64   64 systems (D in {8, 10, 20, 30} m, 1-6 LGS, 2-6 layers; 47
              distinct plans) in one group: every member equals its plan
              solved alone, bit for bit, and each distinct geometry matches
              the FP64 oracle (1e-4, or the oracle's own FP32-rounding
              spread for the ill-conditioned small LGS members).
  served 16   the 16-loop mix (4 x 8 m NGS, 2 x 30 m 6-LGS, 10 uneven-batch
              systems) through the request loop for 3 frozen-flow steps:
              every loop equals the loop served alone, bit for bit, and the
              warm-started 10-iteration steps of an 8 m, a 30 m and a batch
              loop follow the FP64 oracle warm-started from its own previous
              solution to 1e-4.
"""
import collections

import numpy as np
import pytest

from maos_amd import configs, tomo, workload
from oracle.oracle import Oracle
from parity_util import check_cold_solve, rel

pytestmark = pytest.mark.gpu


def test_uneven64_group():
    cfgs = configs.batch64()
    groups = collections.OrderedDict()
    for i, c in enumerate(cfgs):
        groups.setdefault(c["name"], []).append(i)
    plans, members, grads = {}, [], {}
    for name, idx in groups.items():
        cfg = cfgs[idx[0]]
        r = tomo.Recon(cfg)
        g = tomo.GpuTomo(r, nbatch=len(idx), maxit=cfg["maxit"])
        plans[name] = (cfg, r, g, idx)
        members.append((g, 0, len(idx)))
        for i in idx:
            grads[i] = workload.make_inputs(cfg, r, 1000 + i)
    print("uneven64: %d plans" % len(plans), flush=True)
    # each plan's systems alone (one batched call per plan: a system's bits
    # do not depend on its batch neighbours)
    alone = {}
    for name, (cfg, r, g, idx) in plans.items():
        xb = g.tomo_batch(np.stack([grads[i] for i in idx]))
        for s, i in enumerate(idx):
            alone[i] = xb[s]
    for name, (cfg, r, g, idx) in plans.items():
        for s, i in enumerate(idx):
            g.set_grad(s, grads[i])
    grp = tomo.GpuGroup(members)
    for mode in (1, 0):
        print("group solve mode", mode, flush=True)
        grp.solve(mode=mode)
        for name, (cfg, r, g, idx) in plans.items():
            for s, i in enumerate(idx):
                x, res = g.get_x(s)
                assert np.array_equal(x, alone[i]), "%s slot %d (system %d, mode %d)" % (name, s, i, mode)
    grp.close()
    # every distinct geometry against the oracle (its first system)
    worst = 0.0
    for name, (cfg, r, g, idx) in plans.items():
        i = idx[0]
        x, res = g.get_x(0)
        e = check_cold_solve(name, Oracle(r, nthread=8), grads[i], cfg["maxit"], x, res)
        worst = max(worst, e)
    print("uneven64: largest rel L2 vs oracle %.3e" % worst, flush=True)
    for cfg, r, g, idx in plans.values():
        g.close()


def _served_loops():
    cf = [configs.small8() for _ in range(4)] + [configs.large30() for _ in range(2)]
    cf += configs.batch64()[:10]
    return cf


NSTEP = 3


def test_served16_alone_equals_mix_and_oracle():
    cfgs = _served_loops()
    recons = {}
    for c in cfgs:
        if c["name"] not in recons:
            recons[c["name"]] = tomo.Recon(c)
    streams = [workload.frozen_flow_stream(c, recons[c["name"]], 11 + i, NSTEP) for i, c in enumerate(cfgs)]

    def run(which):
        srv = tomo.Server(mem_budget_gb=32.0, max_inflight=3)
        ids = [srv.add_loop(recons[cfgs[i]["name"]], maxit=10) for i in which]
        xs = {i: [] for i in which}
        for k in range(NSTEP):
            for i, lid in zip(which, ids):
                srv.submit(lid, streams[i][k], tag=k)
            srv.drain()
            for i, lid in zip(which, ids):
                xs[i].append(srv.get_x(lid)[0])
        assert len(srv.poll(4096)) == NSTEP * len(which)
        srv.close()
        return xs

    print("served16: mix of %d loops" % len(cfgs), flush=True)
    mixed = run(list(range(len(cfgs))))
    for i in range(len(cfgs)):
        alone = run([i])
        for k in range(NSTEP):
            assert np.array_equal(alone[i][k], mixed[i][k]), "loop %d (%s) step %d" % (i, cfgs[i]["name"], k)
    for i in (0, 4, 6):
        r = recons[cfgs[i]["name"]]
        o = Oracle(r, nthread=8)
        xo = None
        for k in range(NSTEP):
            xo, _ = o.solve(streams[i][k], 10, x0=xo)
            e = rel(mixed[i][k], xo)
            print("served16 loop %d (%s) step %d: rel L2 vs warm FP64 oracle %.3e" % (i, cfgs[i]["name"], k, e),
                  flush=True)
            assert e <= 1e-4
