"""GPU: the tomography half of MVM assembly (gpu_tomo_mvm_tomo).

Column j of T (x = T g) is the solve of the unit gradient vector e_j, as
MAOS's MVM assembly solves the tomography for each gradient column
(cuda/mvm_direct.cu:132).  Checked: each column has the same bits as
gpu_tomo_solve on e_j, a column range equals the same columns of the full
matrix, and the FP64 oracle's solve of e_j agrees to the solve tolerance
(1e-4 relative L2).
"""
import numpy as np
import pytest

from maos_amd import configs, tomo
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def unit(n, j):
    e = np.zeros(n)
    e[j] = 1.0
    return e


def test_columns_are_unit_solves():
    cfg = configs.small8()
    r = tomo.Recon(cfg)
    g = tomo.GpuTomo(r, nbatch=64, maxit=cfg["maxit"])
    T = g.mvm_tomo()
    assert T.shape == (r.nx, r.ng)
    one = tomo.GpuTomo(r, maxit=cfg["maxit"])
    o = Oracle(r, nthread=8)
    for j in (0, 1, 200, r.ng // 2, r.ng - 1):
        x = one.tomo(unit(r.ng, j))[0]
        assert np.array_equal(T[:, j], x), "column %d" % j
        xo, _ = o.solve(unit(r.ng, j), cfg["maxit"])
        e = float(np.linalg.norm(T[:, j] - xo) / np.linalg.norm(xo))
        print("column %d: rel L2 vs oracle %.3e" % (j, e))
        assert e <= 1e-4
    part = g.mvm_tomo(100, 70)
    assert np.array_equal(part, T[:, 100:170])
    g.close()
    one.close()


def test_warm_plan_gives_cold_columns():
    """On a warm-start plan the columns start from zero all the same."""
    cfg = configs.small8(maxit=10)
    r = tomo.Recon(cfg)
    cold = tomo.GpuTomo(r, nbatch=8, maxit=10)
    warm = tomo.GpuTomo(r, nbatch=8, maxit=10, warm=1)
    warm.tomo_batch(np.random.default_rng(3).standard_normal((8, r.ng)))  # nonzero warm state
    a, b = cold.mvm_tomo(0, 20), warm.mvm_tomo(0, 20)
    e = float(np.linalg.norm(a - b) / np.linalg.norm(a))
    print("warm plan vs cold plan: rel L2 %.3e" % e)
    assert e <= 1e-6


def test_rejects_bad_ranges():
    from maos_amd import _abi
    cfg = configs.small8(maxit=5)
    r = tomo.Recon(cfg)
    g = tomo.GpuTomo(r, nbatch=4, maxit=5)
    for col0, ncol in ((0, 0), (-1, 4), (r.ng - 2, 3), (r.ng, 1)):
        with pytest.raises(_abi.MtError):
            g.mvm_tomo(col0, ncol)
    assert g.mvm_tomo(r.ng - 3, 3).shape == (r.nx, 3)  # the last columns
    g.close()


def test_profiled_plan_gives_same_columns():
    """With per-kernel profiling on, the columns come from kernel-by-kernel
    launches instead of a graph replay: the same bits."""
    cfg = configs.small8(maxit=10)
    r = tomo.Recon(cfg)
    g = tomo.GpuTomo(r, nbatch=8, maxit=10)
    a = g.mvm_tomo(0, 12)
    g.profile(True)
    b = g.mvm_tomo(0, 12)
    g.profile(False)
    assert np.array_equal(a, b)
    g.close()
