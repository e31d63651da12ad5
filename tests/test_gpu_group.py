"""GPU: uneven batches (systems of different aperture, WFS count, layer count,
ray path and FFT radices) solved in one launch per kernel per iteration.

Each system's result must be identical, bit for bit, to the same system
solved alone (the kernels are gather-only with fixed-order reductions), and
match the FP64 oracle to 1e-4 like a single solve.
"""
import numpy as np
import pytest

from maos_amd import configs, tomo, workload
from oracle.oracle import Oracle
from parity_util import check_cold_solve

pytestmark = pytest.mark.gpu

MEMBERS = [
    configs.small8,                                   # NGS, pos 2, FFT 56 (radix 7), bs 8
    configs.lgs8_flat,                                # generic ray path
    lambda: configs.batch_member(8.0, 2, 4),          # FFT 28, bs 4
    lambda: configs.batch_member(10.0, 3, 2),         # bs 2
    lambda: configs.batch_member(8.0, 5, 5),          # bs 5 (runtime block size)
]


@pytest.fixture(scope="module")
def plans():
    out = []
    for k, mk in enumerate(MEMBERS):
        cfg = mk()
        r = tomo.Recon(cfg)
        print("member", k, cfg["name"], "nx", r.nx, flush=True)
        g = tomo.GpuTomo(r, nbatch=2, maxit=30)
        grads = [workload.make_inputs(cfg, r, 100 + 10 * k + s) for s in range(2)]
        out.append((cfg, r, g, grads))
    return out


def test_group_equals_alone(plans):
    alone = []
    for cfg, r, g, grads in plans:
        alone.append([g.tomo(grads[s], isys=s)[0] for s in range(2)])
    for cfg, r, g, grads in plans:
        for s in range(2):
            g.set_grad(s, grads[s])
    grp = tomo.GpuGroup([(g, 0, 2) for cfg, r, g, grads in plans])
    for mode in (0, 1):
        grp.solve(mode=mode)
        for k, (cfg, r, g, grads) in enumerate(plans):
            for s in range(2):
                x, res = g.get_x(s)
                assert np.array_equal(x, alone[k][s]), "member %d slot %d (mode %d)" % (k, s, mode)
    grp.close()


def test_group_vs_oracle(plans):
    """Each member's group solve against the oracle: 1e-4, or for the
    ill-conditioned small-aperture LGS members the oracle's own spread under
    FP32 rounding of the preconditioner (check_cold_solve)."""
    for cfg, r, g, grads in plans:
        g.set_grad(1, grads[1])
    grp = tomo.GpuGroup([(g, 1, 1) for cfg, r, g, grads in plans])
    grp.solve(mode=1)
    for k, (cfg, r, g, grads) in enumerate(plans):
        x, res = g.get_x(1)
        o = Oracle(r, nthread=8)
        check_cold_solve(cfg["name"], o, grads[1], 30, x, res)
    grp.close()


def test_group_register_kernels_equal_alone():
    """large30 alone takes the column kernel specialised for 32 lanes and
    6 layers and the register row FFTs; in a group with other plans its
    column stage runs in the shared runtime-shape variant of the register
    kernel (one launch per point count), beside an LDS-stage plan (small8)
    and another register plan (28 = 4 x 7 points).  Same bits as alone."""
    members = [configs.large30, configs.small8, lambda: configs.batch_member(8.0, 2, 4)]
    plans = []
    for k, mk in enumerate(members):
        cfg = mk()
        r = tomo.Recon(cfg)
        g = tomo.GpuTomo(r, nbatch=1, maxit=30)
        grad = workload.make_inputs(cfg, r, 300 + k)
        plans.append((cfg, g, grad, g.tomo(grad)[0]))
        print("member", k, cfg["name"], "info", g.info(), flush=True)
    for cfg, g, grad, x1 in plans:
        g.set_grad(0, grad)
    grp = tomo.GpuGroup([(g, 0, 1) for cfg, g, grad, x1 in plans])
    for mode in (0, 1):
        grp.solve(mode=mode)
        for k, (cfg, g, grad, x1) in enumerate(plans):
            x, res = g.get_x(0)
            assert np.array_equal(x, x1), "member %s (mode %d)" % (cfg["name"], mode)
    grp.close()
    for cfg, g, grad, x1 in plans:
        g.close()


def test_group_autotune_keeps_alone_bits(plans):
    """The group autotuner sets one launch shape on every member plan; each
    system of the tuned group still equals its plan solved alone."""
    grp = tomo.GpuGroup([(g, 0, 2) for cfg, r, g, grads in plans])
    choice = grp.autotune()
    print("group autotune", choice, flush=True)
    assert choice[4] in (0, 2) and choice[3] in (0, 16) and choice[0] in (0, 1, 2, 4) and choice[1] in (0, 1, 2, 4)
    alone = []
    for cfg, r, g, grads in plans:
        alone.append([g.tomo(grads[s], isys=s)[0] for s in range(2)])
    for cfg, r, g, grads in plans:
        for s in range(2):
            g.set_grad(s, grads[s])
    for mode in (0, 1):
        grp.solve(mode=mode)
        for k, (cfg, r, g, grads) in enumerate(plans):
            for s in range(2):
                assert np.array_equal(g.get_x(s)[0], alone[k][s]), "member %d slot %d (mode %d)" % (k, s, mode)
    grp.close()
