"""GPU: turbulence-profile updates (gpu_tomo_update_cn2 and
gpu_tomo_server_update_cn2), the replacement of gpu_update_recon_cn2 /
curecon_t::update_cn2 (cuda/recon.cu:231-250, 526).

After an update the plan must solve exactly as a plan set up fresh from the
updated recon (same bits: both hold the same blocks, and the register column
kernel's copy of them is rebuilt in place), for both column kernels, and
within 1e-4 of the oracle built on the updated recon.
"""
import numpy as np
import pytest

from maos_amd import configs, tomo, workload
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

NEW_WT = {"small8": [0.3, 0.7], "large30": [0.2, 0.3, 0.2, 0.1, 0.1, 0.1]}
NEW_R0 = 0.12


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("name", ["small8", "large30"])
@pytest.mark.parametrize("colfft", [1, 2])
def test_update_equals_fresh_plan(name, colfft):
    cfg = getattr(configs, name)()
    r = tomo.Recon(cfg)
    g = tomo.GpuTomo(r, maxit=cfg["maxit"])
    g.tune_set(0, 0, 0, 0, colfft, 0)
    grad = workload.make_inputs(cfg, r, 3)
    x_old, _, _ = g.tomo(grad)
    print("case", name, "colfft", colfft, g.info(), flush=True)
    r.update_cn2(NEW_WT[name], NEW_R0)
    g.update_cn2(r)
    x_upd, res_upd, _ = g.tomo(grad)
    fresh = tomo.GpuTomo(r, maxit=cfg["maxit"])
    fresh.tune_set(0, 0, 0, 0, colfft, 0)
    x_new, res_new, _ = fresh.tomo(grad)
    assert np.array_equal(x_upd, x_new), rel(x_upd, x_new)
    assert res_upd == res_new
    # the update changed the answer, and the oracle on the new profile agrees
    assert rel(x_upd, x_old) > 1e-3
    xo, _ = Oracle(r, nthread=8).solve(grad, cfg["maxit"])
    assert rel(x_upd, xo) < 1e-4


def test_served_loop_update():
    """A served loop's profile update: the loop's results equal those of the
    same loop served alone with the same update at the same step (bit for
    bit), the other loops are unaffected, and the updated loop's cold start
    matches the oracle on the new profile."""
    cfgs = [configs.small8(), configs.batch_member(10.0, 3, 2), configs.small8(name="small8b")]
    recs = [tomo.Recon(c) for c in cfgs]
    steps = 4
    grads = [[workload.make_inputs(c, r, 40 + 10 * i + k) for k in range(steps)]
             for i, (c, r) in enumerate(zip(cfgs, recs))]

    def run(which, update=True):
        rr = [tomo.Recon(cfgs[i]) for i in which]
        S = tomo.Server(mem_budget_gb=4.0, max_inflight=2)
        ids = [S.add_loop(r, maxit=10) for r in rr]
        out = {i: [] for i in which}
        for k in range(steps):
            if update and k == 2 and 0 in which:
                j = which.index(0)
                rr[j].update_cn2(NEW_WT["small8"], NEW_R0)
                S.update_cn2(ids[j], rr[j])
            for j, i in enumerate(which):
                S.submit(ids[j], grads[i][k], tag=k)
            S.drain()
            S.poll()
            for j, i in enumerate(which):
                out[i].append(S.get_x(ids[j])[0])
        S.close()
        return out

    alone0 = run([0])
    alone2 = run([2])
    mix = run([0, 1, 2])
    for k in range(steps):
        assert np.array_equal(mix[0][k], alone0[0][k]), k
        assert np.array_equal(mix[2][k], alone2[2][k]), k
    # the update took effect from step 2 on, not before
    plain0 = run([0], update=False)
    assert np.array_equal(plain0[0][1], mix[0][1])
    assert rel(mix[0][3], plain0[0][3]) > 1e-3
    # the updated profile on a fresh loop: one cold 30-iteration step against
    # the oracle built on it
    r0 = tomo.Recon(cfgs[0])
    r0.update_cn2(NEW_WT["small8"], NEW_R0)
    S = tomo.Server(mem_budget_gb=4.0, max_inflight=1)
    lid = S.add_loop(r0, maxit=30)
    S.submit(lid, grads[0][0])
    S.drain()
    x = S.get_x(lid)[0]
    S.close()
    xo, _ = Oracle(r0, nthread=8).solve(grads[0][0], 30)
    assert rel(x, xo) < 1e-4
