"""GPU parity: HIP operators and full solve vs the FP64 CPU oracle.

All calls go through the C ABI of libmaos_tomo.so (maos_amd.tomo.GpuTomo);
the oracle (oracle/tomo_oracle.c) is the checker.  Tolerances: operator
applies in FP32 vs FP64 to 2e-5 relative L2; the 30-iteration FDPCG-PCG
solution to 1e-4 relative L2 (the accuracy gate of this path).
"""
import numpy as np
import pytest

from maos_amd import configs, tomo, workload
from oracle.oracle import Oracle
from parity_util import fp32_spread

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


EXTRA = {
    "lgs8_flat": configs.lgs8_flat,
    "b8_w2_l4": lambda: configs.batch_member(8.0, 2, 4),    # FFT 28 (radix 7), bs 4
    "b10_w3_l2": lambda: configs.batch_member(10.0, 3, 2),  # bs 2
    "b8_w5_l5": lambda: configs.batch_member(8.0, 5, 5),    # bs 5 (runtime block size)
    # register FFT shapes: 64 = 32 x 2 (rows and columns in registers, bs 3),
    # 40 = 8 x 5 (columns on 8 lanes, bs 6), 96 with 3 layers (32 lanes,
    # runtime block size)
    "b20_w3_l3": lambda: configs.batch_member(20.0, 3, 3),
    "b10_w4_l6": lambda: configs.batch_member(10.0, 4, 6),
    "b30_w2_l3": lambda: configs.batch_member(30.0, 2, 3),
}


@pytest.fixture(scope="module", params=["small8", "large30", "lgs8_flat", "b8_w2_l4", "b10_w3_l2", "b8_w5_l5",
                                        "b20_w3_l3", "b10_w4_l6", "b30_w2_l3"])
def sys_(request):
    if request.param in EXTRA:
        cfg = EXTRA[request.param]()
        r = tomo.Recon(cfg)
    else:
        cfg, r = request.getfixturevalue(request.param)
    g = tomo.GpuTomo(r, maxit=cfg["maxit"])
    print("case", cfg["name"], g.info(), flush=True)
    return cfg, r, g, Oracle(r, nthread=8)


def test_L(sys_):
    cfg, r, g, o = sys_
    x = np.random.default_rng(5).standard_normal(r.nx)
    assert rel(g.L(x), o.tomo_l(x)) < 2e-5


def test_R(sys_):
    cfg, r, g, o = sys_
    grad = workload.make_inputs(cfg, r, 7)
    assert rel(g.R(grad), o.tomo_r(grad)) < 2e-5


def test_P(sys_):
    # The preconditioner sees residuals, whose layer means are zero (the
    # per-layer piston is in the null space of the operator and the RHS is
    # orthogonal to it); the GPU path zeroes that bin, so feed it such input.
    cfg, r, g, o = sys_
    x = np.random.default_rng(6).standard_normal(r.nx)
    off = r.layer_offsets()
    for l in range(r.npsr):
        x[off[l]:off[l + 1]] -= x[off[l]:off[l + 1]].mean()
    assert rel(g.P(x), o.fdpcg(x)) < 2e-5


def test_solve(sys_):
    """30 iterations against the oracle: 1e-4 relative L2.  Where the GPU
    misses that bar, the system must be one whose iterates FP32 rounding in
    the preconditioner moves by that much on the oracle too: the GPU error
    is then held to 3x the oracle's own spread under such rounding, and its
    solution must be as close to the converged one (500 iterations) as the
    oracle's, to 2%."""
    cfg, r, g, o = sys_
    grad = workload.make_inputs(cfg, r, 1)
    xg, resg, rc = g.tomo(grad)
    xo, reso = o.solve(grad, cfg["maxit"])
    e = rel(xg, xo)
    print("%s: rel L2 %.3e, residual gpu %.4e oracle %.4e" % (cfg["name"], e, resg, reso))
    assert rc == 0
    assert abs(resg - reso) <= 1e-3 * reso + 1e-6
    if e > 1e-4:
        spread = fp32_spread(o, grad, cfg["maxit"])
        xt, _ = o.solve(grad, 500)
        eg, eo = rel(xg, xt), rel(xo, xt)
        print("  oracle spread under FP32 preconditioner rounding %.3e; "
              "to the converged solution: gpu %.4e oracle %.4e" % (spread, eg, eo))
        assert e <= 3 * spread
        assert abs(eg - eo) <= 0.02 * eo


def test_batched_equals_single(sys_):
    """Systems of one batched call match the same systems solved alone
    (a system's result does not depend on its batch neighbours)."""
    cfg, r, g1, o = sys_
    grads = [workload.make_inputs(cfg, r, seed) for seed in (11, 12, 13)]
    singles = [g1.tomo(gr)[0] for gr in grads]
    gb = tomo.GpuTomo(r, nbatch=3, maxit=cfg["maxit"])
    xb = gb.tomo_batch(np.stack(grads))
    for i in range(3):
        assert np.array_equal(xb[i], singles[i]), "system %d differs" % i


def test_warm_start():
    """Warm start (x0 = previous solution, resident on the device) against
    the oracle started from its own previous solution (lib/pcg.c warm=1)."""
    cfg = configs.small8(maxit=10)
    r = tomo.Recon(cfg)
    o = Oracle(r, nthread=8)
    g = tomo.GpuTomo(r, maxit=10, warm=1)
    xo = None
    for step, seed in enumerate((21, 22, 23)):
        grad = workload.make_inputs(cfg, r, seed)
        xg, resg, rc = g.tomo(grad)
        xo, reso = o.solve(grad, 10, x0=xo)
        e = rel(xg, xo)
        print("warm step %d: rel L2 %.3e residual gpu %.4e oracle %.4e" % (step, e, resg, reso))
        assert e <= 1e-4


def test_restart_past_100():
    """maxit > 100 crosses the restart of lib/pcg.c:55 (r = b - A x, p = z).
    Past ~60 iterations the FP32 operators let the iterates drift from the
    FP64 ones smoothly (small8: 1e-5 at 60, 3.4e-4 at 99); the restart must
    not add to that drift, and the residual must follow the oracle's."""
    cfg = configs.small8()
    r = tomo.Recon(cfg)
    o = Oracle(r, nthread=8)
    grad = workload.make_inputs(cfg, r, 1)
    err = {}
    for it in (99, 104):
        g = tomo.GpuTomo(r, maxit=it)
        xg, resg, rc = g.tomo(grad)
        xo, reso = o.solve(grad, it)
        err[it] = rel(xg, xo)
        print("%d iterations: rel L2 %.3e residual gpu %.5e oracle %.5e" % (it, err[it], resg, reso))
        assert abs(resg - reso) <= 0.05 * reso
        g.close()
    assert err[104] <= 1e-3
    assert err[104] <= 1.5 * err[99] + 1e-5


@pytest.mark.parametrize("mode,tol", [(0, 5e-3), (1, 1e-4), (2, 1e-4), (3, 5e-2)])
def test_precision_modes(large30, mode, tol):
    """The 30 m solve in each precision mode against the FP64 oracle."""
    cfg, r = large30
    o = Oracle(r, nthread=8)
    g = tomo.GpuTomo(r, maxit=cfg["maxit"])
    g.precision(mode)
    grad = workload.make_inputs(cfg, r, 1)
    xg, resg, rc = g.tomo(grad)
    xo, reso = o.solve(grad, cfg["maxit"])
    e = rel(xg, xo)
    print("precision mode %d: rel L2 %.3e residual gpu %.5e oracle %.5e" % (mode, e, resg, reso))
    assert e <= tol


def test_bin_files_end_to_end(tmp_path):
    """MAOS's file flow at the boundary: gradients saved as MAOS saves
    gradlastcl (a cell of one 2 nsa x 1 array per WFS, .bin), read into the
    solve, the solution written as opdr (.bin.gz) and read back: the same
    bits as the solve from memory."""
    from maos_amd import binio
    cfg = configs.small8()
    r = tomo.Recon(cfg)
    grad = workload.make_inputs(cfg, r, 4)
    offs = r.grad_offsets()
    binio.write(tmp_path / "gradlastcl_1.bin", [grad[offs[w]:offs[w + 1]] for w in range(r.nwfs)],
                "Created by the test")
    g = tomo.GpuTomo(r, maxit=cfg["maxit"])
    x_mem = g.tomo(grad)[0]
    x_file = g.tomo(binio.read_grad(tmp_path / "gradlastcl_1.bin", r))[0]
    assert np.array_equal(x_file, x_mem)
    binio.write_opdr(tmp_path / "opdr_1.bin.gz", r, x_file)
    back = binio.read(tmp_path / "opdr_1.bin.gz")
    off = r.layer_offsets()
    for l in range(r.npsr):
        assert np.array_equal(back[l, 0].reshape(-1, order="F"), x_mem[off[l]:off[l + 1]])
    g.close()
