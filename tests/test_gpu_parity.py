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

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module", params=["small8", "large30", "lgs8_flat"])
def sys_(request):
    if request.param == "lgs8_flat":
        cfg = configs.lgs8_flat()
        r = tomo.Recon(cfg)
    else:
        cfg, r = request.getfixturevalue(request.param)
    print("case", cfg["name"], flush=True)
    return cfg, r, tomo.GpuTomo(r, maxit=cfg["maxit"]), Oracle(r, nthread=8)


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
    cfg, r, g, o = sys_
    grad = workload.make_inputs(cfg, r, 1)
    xg, resg, rc = g.tomo(grad)
    xo, reso = o.solve(grad, cfg["maxit"])
    e = rel(xg, xo)
    print("%s: rel L2 %.3e, residual gpu %.4e oracle %.4e" % (cfg["name"], e, resg, reso))
    assert rc == 0
    assert e <= 1e-4
    assert abs(resg - reso) <= 1e-3 * reso + 1e-6
