"""GPU: the .conf variants of the solve besides the default FDPCG split
tomography, each against the FP64 oracle:

  tomo.precond=0     plain CG (gpu_pcg with Mmul = NULL, cuda/pcg.cu:149-151;
                     lib/pcg.c:28-113 with z = r)
  recon.split=0 /    TT removed inside the LHS as well (ptt = !split ||
  tomo.splitlrt=1    splitlrt, cuda/tomo.cu:183; TomoL's TTFR,
                     maos/recon_utils.c:461-465)

and the plan state that other entry points must leave alone (MVM columns
keep the warm-start state) and the rebuild of a group whose member changed
its launch shape.
"""
import numpy as np
import pytest

from maos_amd import configs, tomo, workload
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


CASES = {
    "small8_cg": lambda: configs.small8(name="small8_cg", tomo_precond=0),
    "large30_cg": lambda: configs.large30(name="large30_cg", tomo_precond=0),
    "large30_split0": lambda: configs.large30(name="large30_split0", recon_split=0),
    "large30_lrt": lambda: configs.large30(name="large30_lrt", tomo_splitlrt=1),
    "large30_split0_cg": lambda: configs.large30(name="large30_split0_cg", recon_split=0, tomo_precond=0),
    "lgs8_flat_split0": lambda: configs.lgs8_flat(name="lgs8_flat_split0", recon_split=0),
}


@pytest.fixture(scope="module", params=list(CASES))
def case(request):
    cfg = CASES[request.param]()
    r = tomo.Recon(cfg)
    g = tomo.GpuTomo(r, maxit=cfg["maxit"])
    print("case", cfg["name"], "precond", r.rec.precond, "ptt_lhs", r.rec.ptt_lhs, flush=True)
    return cfg, r, g, Oracle(r, nthread=8)


def test_L(case):
    cfg, r, g, o = case
    x = np.random.default_rng(11).standard_normal(r.nx)
    assert rel(g.L(x), o.tomo_l(x)) < 2e-5


def test_solve(case):
    cfg, r, g, o = case
    grad = workload.make_inputs(cfg, r, 2)
    x, res, rc = g.tomo(grad)
    xo, reso = o.solve(grad, cfg["maxit"])
    print(cfg["name"], "rel", rel(x, xo), "res", res, reso, flush=True)
    assert rc in (0, 1)
    assert rel(x, xo) < 1e-4
    assert abs(res - reso) <= 1e-3 * reso


def test_warm_and_batch(case):
    """Warm start (r = b - A x, then the same recurrence) and a batch of two
    systems that equals each solved alone, bit for bit."""
    cfg, r, _, o = case
    g = tomo.GpuTomo(r, nbatch=2, maxit=10, warm=1)
    grads = [workload.make_inputs(cfg, r, 20 + k) for k in range(3)]
    x1, _, _ = g.tomo(grads[0], isys=0)
    x2, _, _ = g.tomo(grads[1], isys=0)
    xo1, _ = o.solve(grads[0], 10)
    xo2, _ = o.solve(grads[1], 10, x0=xo1)
    assert rel(x2, xo2) < 1e-4
    gb = tomo.GpuTomo(r, nbatch=2, maxit=10)
    xb = gb.tomo_batch(np.stack([grads[1], grads[2]]))
    xs, _, _ = tomo.GpuTomo(r, maxit=10).tomo(grads[2])
    assert np.array_equal(xb[1], xs)


def test_mvm_keeps_warm_state(small8):
    """gpu_tomo_mvm_tomo leaves the slots' gradients and solutions as they
    were: warm solve, MVM columns, warm solve gives the bits of the two warm
    solves alone."""
    cfg, r = small8
    grads = [workload.make_inputs(cfg, r, 60 + k) for k in range(2)]
    a = tomo.GpuTomo(r, nbatch=4, maxit=10, warm=1)
    b = tomo.GpuTomo(r, nbatch=4, maxit=10, warm=1)
    a.tomo(grads[0])
    b.tomo(grads[0])
    cols = b.mvm_tomo(0, 9)
    assert cols.shape == (r.nx, 9)
    xa, ra, _ = a.tomo(grads[1])
    xb, rb, _ = b.tomo(grads[1])
    assert np.array_equal(xa, xb)
    assert ra == rb


def test_group_follows_member_changes(small8):
    """A group built before one member's launch shape changes rebuilds its
    tables: the member's result in the group stays that of the member alone
    under its new shape."""
    cfg, r = small8
    r2 = tomo.Recon(configs.batch_member(10.0, 3, 2))
    c2 = configs.batch_member(10.0, 3, 2)
    g1 = tomo.GpuTomo(r, maxit=30)
    g2 = tomo.GpuTomo(r2, maxit=30)
    grad1 = workload.make_inputs(cfg, r, 5)
    grad2 = workload.make_inputs(c2, r2, 6)
    grp = tomo.GpuGroup([(g1, 0, 1), (g2, 0, 1)])
    g1.set_grad(0, grad1)
    g2.set_grad(0, grad2)
    grp.solve(mode=1)
    # new shape on a member (LDS stages for its FFTs), then the group again
    g2.tune_set(0, 0, 0, 0, 2, 0)
    g1.set_grad(0, grad1)
    g2.set_grad(0, grad2)
    grp.solve(mode=1)
    xg2, _ = g2.get_x(0)
    alone = tomo.GpuTomo(r2, maxit=30)
    alone.tune_set(0, 0, 0, 0, 2, 0)
    xa2, _, _ = alone.tomo(grad2)
    assert np.array_equal(xg2, xa2)
    xg1, _ = g1.get_x(0)
    xa1, _, _ = tomo.GpuTomo(r, maxit=30).tomo(grad1)
    assert np.array_equal(xg1, xa1)
