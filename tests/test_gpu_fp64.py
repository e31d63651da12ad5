"""GPU: precision mode 4, the all-FP64 solve (the counterpart of the
reference's CUDA_DOUBLE build, cuda/common.h:39-45), against the FP64 oracle.

Both sides are FP64, so the operator applies agree to rounding (1e-12
relative L2; the GPU sums in other orders and its FFT is a two-factor split
with direct DFTs).  The 30-iteration solve is held to 1e-9, or where the
oracle itself moves more under FP64-sized (1e-15 relative) perturbations of
its preconditioner and operator outputs, to 3x that spread (large30's is
2e-9 to 2e-8).  Plans: the default split FDPCG (small8, large30, an LGS system with
TT removal, odd FFT lengths 28 and 40 and block sizes 2, 4 and 6), plain CG
and TT removal inside the LHS.
"""
import numpy as np
import pytest

from maos_amd import _abi, configs, tomo, workload
from oracle.oracle import Oracle
from parity_util import fp32_spread, rel

pytestmark = pytest.mark.gpu

CASES = {
    "small8": lambda: configs.small8(),
    "large30": lambda: configs.large30(),
    "lgs8_flat": lambda: configs.lgs8_flat(),
    "b8_w2_l4": lambda: configs.batch_member(8.0, 2, 4),    # FFT 28, bs 4
    "b10_w3_l2": lambda: configs.batch_member(10.0, 3, 2),  # bs 2
    "b10_w4_l6": lambda: configs.batch_member(10.0, 4, 6),  # FFT 40, bs 6
    "small8_cg": lambda: configs.small8(name="small8_cg", tomo_precond=0),
    "large30_split0": lambda: configs.large30(name="large30_split0", recon_split=0),
}


@pytest.fixture(scope="module", params=list(CASES))
def sys64(request):
    cfg = CASES[request.param]()
    r = tomo.Recon(cfg)
    g = tomo.GpuTomo(r, maxit=cfg["maxit"])
    g.precision(4)
    print("case", cfg["name"], "precond", r.rec.precond, "ptt_lhs", r.rec.ptt_lhs, flush=True)
    yield cfg, r, g, Oracle(r, nthread=8)
    g.close()


def test_L64(sys64):
    cfg, r, g, o = sys64
    x = np.random.default_rng(5).standard_normal(r.nx)
    e = rel(g.L(x), o.tomo_l(x))
    print("L: %.3e" % e)
    assert e < 1e-12


def test_R64(sys64):
    cfg, r, g, o = sys64
    grad = workload.make_inputs(cfg, r, 7)
    e = rel(g.R(grad), o.tomo_r(grad))
    print("R: %.3e" % e)
    assert e < 1e-12


def test_P64(sys64):
    """Any input, the layer means included (the FP64 transform keeps the DC
    bins, as the reference's fdpcg_precond does)."""
    cfg, r, g, o = sys64
    if not r.rec.precond:
        pytest.skip("plain CG: no preconditioner")
    x = np.random.default_rng(6).standard_normal(r.nx)
    e = rel(g.P(x), o.fdpcg(x))
    print("P: %.3e" % e)
    assert e < 1e-12


def test_solve64(sys64):
    cfg, r, g, o = sys64
    grad = workload.make_inputs(cfg, r, 1)
    xg, resg, rc = g.tomo(grad)
    xo, reso = o.solve(grad, cfg["maxit"])
    e = rel(xg, xo)
    print("%s: rel L2 %.3e, residual gpu %.6e oracle %.6e" % (cfg["name"], e, resg, reso), flush=True)
    assert rc == 0
    if e <= 1e-9:
        assert abs(resg - reso) <= 1e-6 * reso
    else:
        spread = fp32_spread(o, grad, cfg["maxit"], eps=1e-15, to32=False, perturb_L=True, draws=4)
        print("  oracle spread under FP64-sized perturbations of P and L outputs %.3e" % spread)
        assert e <= 3 * spread
        assert abs(resg - reso) <= 0.1 * reso


def test_batch64_equals_single(sys64):
    cfg, r, g1, o = sys64
    grads = [workload.make_inputs(cfg, r, seed) for seed in (11, 12, 13)]
    singles = [g1.tomo(gr)[0] for gr in grads]
    gb = tomo.GpuTomo(r, nbatch=3, maxit=cfg["maxit"])
    gb.precision(4)
    xb = gb.tomo_batch(np.stack(grads))
    res = gb.residuals(3)
    for i in range(3):
        assert np.array_equal(xb[i], singles[i]), "system %d differs" % i
        assert np.array_equal(gb.get_x(i)[0], xb[i])
        assert np.isfinite(res[i]) and res[i] < 1
    gb.close()


def test_warm64_and_mode_switch():
    """Warm starts in mode 4 against the oracle started from its own previous
    solution; entering mode 4 carries the FP32 modes' resident solution over
    (the oracle then starts from that solution), leaving it rounds the FP64
    solution back to FP32."""
    cfg = configs.small8(maxit=10)
    r = tomo.Recon(cfg)
    o = Oracle(r, nthread=8)
    g = tomo.GpuTomo(r, maxit=10, warm=1)
    grads = [workload.make_inputs(cfg, r, seed) for seed in (21, 22, 23, 24)]
    x32, _, _ = g.tomo(grads[0])       # mode 1
    g.precision(4)
    xo = x32.astype(np.float32).astype(np.float64)
    for step in (1, 2):
        xg, resg, _ = g.tomo(grads[step])
        xo, reso = o.solve(grads[step], 10, x0=xo)
        e = rel(xg, xo)
        print("warm FP64 step %d: rel L2 %.3e residual gpu %.6e oracle %.6e" % (step, e, resg, reso))
        assert e <= 1e-9
        assert abs(resg - reso) <= 1e-6 * reso
    g.precision(1)
    # the FP32 plan resumes from the FP64 solution rounded to FP32, and
    # reports the FP64 plan's last residual
    x_after, res_after = g.get_x(0)
    assert np.array_equal(x_after, xg.astype(np.float32).astype(np.float64))
    assert res_after == resg
    xg3, _, _ = g.tomo(grads[3])
    xo3, _ = o.solve(grads[3], 10, x0=x_after)
    assert rel(xg3, xo3) <= 1e-4
    g.close()


def test_restart64_past_100():
    """maxit 104 crosses the restart of lib/pcg.c:55 in FP64 too.  Past ~80
    iterations small8's CG is chaotic in FP64 itself: when FP64-sized
    perturbations (1e-15 relative) enter its preconditioner output the
    oracle moves by 6e-10 at 80, 1.2e-6 at 99 and 1.2e-5 at 104 iterations;
    with its operator output perturbed too, draws at 104 iterations land
    1e-6 to 1.5e-5 away.  The GPU is held to 3x the largest of four such
    draws and the residual to 5 %."""
    cfg = configs.small8()
    r = tomo.Recon(cfg)
    o = Oracle(r, nthread=8)
    grad = workload.make_inputs(cfg, r, 1)
    g = tomo.GpuTomo(r, maxit=104)
    g.precision(4)
    xg, resg, _ = g.tomo(grad)
    xo, reso = o.solve(grad, 104)
    e = rel(xg, xo)
    spread = fp32_spread(o, grad, 104, eps=1e-15, to32=False, perturb_L=True, draws=4)
    print("104 iterations FP64: rel L2 %.3e (oracle spread %.3e) residual gpu %.6e oracle %.6e"
          % (e, spread, resg, reso))
    assert e <= max(1e-8, 3 * spread)
    assert abs(resg - reso) <= 0.05 * reso
    g.close()


def test_update_cn2_64():
    """A Cn2 update reaches the FP64 plan: its solve equals a fresh mode-4
    plan of the updated recon bit for bit and agrees with the oracle."""
    cfg = configs.large30()
    r = tomo.Recon(cfg)
    g = tomo.GpuTomo(r, maxit=cfg["maxit"])
    g.precision(4)
    grad = workload.make_inputs(cfg, r, 3)
    x_old, _, _ = g.tomo(grad)
    r.update_cn2([0.2, 0.3, 0.2, 0.1, 0.1, 0.1], 0.12)
    g.update_cn2(r)
    x_upd, res_upd, _ = g.tomo(grad)
    fresh = tomo.GpuTomo(r, maxit=cfg["maxit"])
    fresh.precision(4)
    x_new, res_new, _ = fresh.tomo(grad)
    assert np.array_equal(x_upd, x_new)
    assert res_upd == res_new
    assert rel(x_upd, x_old) > 1e-3
    o = Oracle(r, nthread=8)
    xo, _ = o.solve(grad, cfg["maxit"])
    e = rel(x_upd, xo)
    print("updated profile: rel L2 %.3e" % e)
    if e > 1e-9:
        spread = fp32_spread(o, grad, cfg["maxit"], eps=1e-15, to32=False, perturb_L=True, draws=4)
        print("  oracle spread under FP64-sized perturbations of P and L outputs %.3e" % spread)
        assert e <= 3 * spread
    g.close()
    fresh.close()


def test_mode4_refusals():
    """The FP32 device-buffer, group and MVM entry points refuse a plan in
    mode 4 with MT_EINVAL (and take it again after leaving the mode)."""
    from hipmem import DevArray
    cfg = configs.small8()
    r = tomo.Recon(cfg)
    g = tomo.GpuTomo(r, maxit=cfg["maxit"])
    g.precision(4)
    dg = DevArray.of(np.zeros(g.ng, dtype=np.float32))
    dx = DevArray.of(np.zeros(g.nx, dtype=np.float32))
    with pytest.raises(_abi.MtError) as e:
        g.solve_dev(dg.ptr, dx.ptr, nbatch=1)
    assert e.value.code == _abi.MT_EINVAL
    with pytest.raises(_abi.MtError):
        g.mvm_tomo(0, 1)
    with pytest.raises(_abi.MtError):
        tomo.GpuGroup([(g, 0, 1)])
    g.precision(1)
    g.solve_dev(dg.ptr, dx.ptr, nbatch=1)
    g.close()
