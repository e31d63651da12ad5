"""GPU: gpu_tomo_solve_dev_noalloc, the entry point that allocates nothing
(its workspace is a library-owned device array), called through the C ABI
and checked against gpu_tomo_solve_dev on the same device inputs: identical
bits, for two calls in a row on one GPU, the second smaller than the first
(fewer systems of a smaller plan), so the second call starts on whatever the
first left in the workspace."""
import ctypes as C

import numpy as np
import pytest

from hipmem import DevArray, sync
from maos_amd import _abi, configs, tomo, workload

pytestmark = pytest.mark.gpu


def _dev_grads(r, grads):
    """The plan's active-WFS FP32 gradient layout, stacked, on the device."""
    off = r.grad_offsets()
    rows = []
    for gr in grads:
        parts = [gr[off[w]:off[w + 1]] for w in range(r.nwfs) if not r.rec.skip[w]]
        rows.append(np.concatenate(parts).astype(np.float32))
    return DevArray.of(np.stack(rows))


def _noalloc(g, dg, dx, nsys):
    res = DevArray.of(np.full(nsys, np.nan))
    rc = _abi.lib().gpu_tomo_solve_dev_noalloc(g._h, C.c_void_p(dg.ptr), C.c_void_p(dx.ptr), C.c_void_p(res.ptr),
                                               nsys, None)
    _abi.check(rc, "gpu_tomo_solve_dev_noalloc")
    sync()
    return res.numpy()


def _reference(g, dg, dx, nsys):
    _abi.check(_abi.lib().gpu_tomo_solve_dev(g._h, C.c_void_p(dg.ptr), C.c_void_p(dx.ptr), nsys, 0, None),
               "gpu_tomo_solve_dev")
    sync()
    return g.residuals(nsys)


@pytest.mark.parametrize("warm", [0, 1])
def test_noalloc_matches_solve_dev(warm):
    cases = [(configs.large30(), 2, 100), (configs.small8(), 1, 200)]
    for cfg, nsys, seed in cases:
        print("case", cfg["name"], "nsys", nsys, "warm", warm, flush=True)
        r = tomo.Recon(cfg)
        g = tomo.GpuTomo(r, nbatch=nsys, maxit=cfg["maxit"], warm=warm)
        dg = _dev_grads(r, [workload.make_inputs(cfg, r, seed + k) for k in range(nsys)])
        x0 = (np.random.default_rng(seed).standard_normal((nsys, r.nx)) * 1e-8).astype(np.float32)
        xa = DevArray.of(x0)
        ra = _noalloc(g, dg, xa, nsys)
        xb = DevArray.of(x0)
        rb = _reference(g, dg, xb, nsys)
        a, b = xa.numpy(), xb.numpy()
        assert np.array_equal(a, b), float(np.abs(a - b).max())
        assert np.array_equal(ra, rb), (ra, rb)
        assert np.all(np.isfinite(ra)) and np.all(ra < 1)
        g.close()


def test_noalloc_rejects_bad_arguments(small8):
    cfg, r = small8
    g = tomo.GpuTomo(r, nbatch=1, maxit=cfg["maxit"])
    L = _abi.lib()
    d = DevArray((16,), np.float32)
    assert L.gpu_tomo_solve_dev_noalloc(g._h, C.c_void_p(d.ptr), C.c_void_p(d.ptr), None, 2, None) == _abi.MT_EINVAL
    assert L.gpu_tomo_solve_dev_noalloc(g._h, None, C.c_void_p(d.ptr), None, 1, None) == _abi.MT_EINVAL
    g.close()
