"""GPU: launch-shape choices do not change results.

Systems per workgroup (gpu_tomo_tune_set / gpu_tomo_autotune), the
workgroup order and the profiler's repeated launches change how the work is
laid out on the GPU, never what a system computes: every choice gives the
same bits.  All calls go through the C ABI (maos_amd.tomo)."""
import numpy as np
import pytest

from maos_amd import _abi, tomo, workload

pytestmark = pytest.mark.gpu


def _batch(r, cfg, n):
    return np.stack([workload.make_inputs(cfg, r, 40 + s) for s in range(n)])


@pytest.mark.parametrize("rows", [0, 32])
@pytest.mark.parametrize("colk", [1, 2])
@pytest.mark.parametrize("mode", [0, 1, 3])
def test_systems_per_workgroup_bit_exact(large30, mode, colk, rows):
    """Five systems (one partial workgroup for every choice above 1) solved
    under each systems-per-workgroup choice and by the heuristic, with the
    register column kernel (1, which has no such choice) and the LDS column
    kernel (2), on 16 x 64 (0: the default) and 16 x 32 layer tiles:
    identical solutions and residuals."""
    cfg, r = large30
    grads = _batch(r, cfg, 5)
    g = tomo.GpuTomo(r, nbatch=5, maxit=cfg["maxit"])
    g.precision(mode)
    g.tune_set(0, 0, 0, rows, colk)
    ref = g.tomo_batch(grads)
    ref_res = g.residuals(5)
    # (column 2 with an FP32 spectrum once launched the 4-system kernel on a
    # grid and LDS sized for 2)
    for choice in [(1, 1, 1), (2, 2, 2), (4, 4, 2), (2, 1, 2), (1, 1, 2), (1, 1, 4)]:
        g.tune_set(*choice, rows, colk)
        x = g.tomo_batch(grads)
        print("choice", choice, "max |dx|", float(np.abs(x - ref).max()), flush=True)
        assert np.array_equal(x, ref), "systems per workgroup %s changed the solution" % (choice,)
        assert np.array_equal(g.residuals(5), ref_res)
    g.close()


def test_graph_resized_while_in_flight(large30):
    """Graph solves of 5 and then 3 systems queued back to back on the
    plan's stream with no wait: the second call replaces the captured
    5-system graph while its launch may still be queued.  The plan waits for
    that launch before destroying the graph, so both results equal the same
    calls made one at a time (a system's bits do not depend on how many
    systems share the call)."""
    from hipmem import DevArray, sync
    cfg, r = large30
    grads = _batch(r, cfg, 5).astype(np.float32)
    g = tomo.GpuTomo(r, nbatch=5, maxit=cfg["maxit"])
    dg = DevArray.of(grads)
    d5, d3 = DevArray((5, g.nx), np.float32), DevArray((3, g.nx), np.float32)
    g.solve_dev(dg.ptr, d5.ptr, nbatch=5, mode=1)
    ref5 = d5.numpy()
    g.solve_dev(dg.ptr, d3.ptr, nbatch=3, mode=1)
    ref3 = d3.numpy()
    assert np.array_equal(ref3, ref5[:3])
    for rep in range(3):
        g.solve_dev(dg.ptr, d5.ptr, nbatch=5, mode=1)
        g.solve_dev(dg.ptr, d3.ptr, nbatch=3, mode=1)
        sync()
        assert np.array_equal(d5.numpy(), ref5), rep
        assert np.array_equal(d3.numpy(), ref3), rep
    for d in (dg, d5, d3):
        d.free()
    g.close()


def test_tune_set_rejects_bad_choice(large30):
    cfg, r = large30
    g = tomo.GpuTomo(r, nbatch=1, maxit=5)
    with pytest.raises(_abi.MtError):
        g.tune_set(3, 1, 1)  # 3 is not a choice
    with pytest.raises(_abi.MtError):
        g.tune_set(1, 1, 1, 48)  # nor are 48-row tiles
    with pytest.raises(_abi.MtError):
        g.tune_set(1, 1, 1, 0, 3)  # nor a third column kernel
    with pytest.raises(_abi.MtError):
        g.tune_set(1, 1, 1, 0, 0, 3)  # nor 3 flights
    g.close()


def test_autotune_replay(large30):
    """The autotuner overwrites slots and picks a choice; replaying that
    choice with tune_set gives the same bits, and the result stays within
    the rounding of the choice of layer tile shape (the order of the p.Ap
    partial sums) of the solution before tuning."""
    cfg, r = large30
    grads = _batch(r, cfg, 3)
    g = tomo.GpuTomo(r, nbatch=3, maxit=cfg["maxit"])
    ref = g.tomo_batch(grads)
    choice = g.autotune(3)
    print("autotuned", choice, flush=True)
    assert all(c in (1, 2, 4) for c in choice[:2]) and choice[3] in (16, 32, 64) and choice[4] in (1, 2)
    assert choice[2] in ((0,) if choice[4] == 1 else (1, 2, 4)) and choice[5] in (1, 2)
    x = g.tomo_batch(grads)
    assert np.linalg.norm(x - ref) <= 1e-5 * np.linalg.norm(ref)
    g2 = tomo.GpuTomo(r, nbatch=3, maxit=cfg["maxit"])
    g2.tune_set(*choice)
    assert np.array_equal(g2.tomo_batch(grads), x)
    g.close()
    g2.close()


@pytest.mark.parametrize("rows", [16, 32, 64])
def test_layer_tile_shape_parity(large30, rows):
    """Each layer tile shape against the FP64 oracle."""
    from oracle.oracle import Oracle
    cfg, r = large30
    grad = workload.make_inputs(cfg, r, 1)
    g = tomo.GpuTomo(r, maxit=cfg["maxit"])
    g.tune_set(0, 0, 0, rows)
    xg, resg, rc = g.tomo(grad)
    xo, reso = Oracle(r, nthread=8).solve(grad, cfg["maxit"])
    e = float(np.linalg.norm(xg - xo) / np.linalg.norm(xo))
    print("layer tile rows %d: rel L2 %.3e" % (rows, e), flush=True)
    assert e <= 1e-4
    g.close()


def test_profile_repeats_keep_results(large30):
    """Profiling with the read-only kernel classes launched 8 times per event
    pair leaves the solution unchanged and counts every launch.  (The plan's
    own device buffers: gradients by set_grad, solutions by get_x.)"""
    cfg, r = large30
    grads = _batch(r, cfg, 2)
    g = tomo.GpuTomo(r, nbatch=2, maxit=cfg["maxit"])
    for i in range(2):
        g.set_grad(i, grads[i])
    g.solve_dev(0, 0, nbatch=2, mode=0)
    ref = [g.get_x(i)[0] for i in range(2)]
    g.profile(True, reps=8)
    g.solve_dev(0, 0, nbatch=2, mode=0)
    prof = g.profile_read()
    g.profile(False)
    for i in range(2):
        assert np.array_equal(g.get_x(i)[0], ref[i])
    it = cfg["maxit"]
    # read-only classes: 8 launches per event pair; k_row once per update
    assert prof["k_col"][1] == 8 * it
    assert prof["k_layer"][1] == 8 * it
    assert prof["k_row"][1] == it - 1
    g.close()


@pytest.mark.parametrize("cfgname,path", [("large30", 1), ("large30", 2), ("small8", 1), ("small8", 2),
                                           ("b8_w2_l4", 1), ("b8_w2_l4", 2), ("b8_w5_l5", 1),
                                           ("b10_w3_l2", 1)])
def test_column_fft_paths_parity(cfgname, path):
    """The register column kernel (path 1: column FFTs across lanes, blocks
    and inverse FFTs in registers) and the LDS column kernel (path 2), against
    the FP64 oracle: one preconditioner apply and the 30-iteration solve.
    The register kernel applies to large30 (96 = 32 x 3 points, 6 layers) and
    the 8 m members (28 = 4 x 7, 4 and 5 layers); small8 (blocks coupling
    aliased frequencies) takes the LDS kernel either way."""
    from maos_amd import configs
    from oracle.oracle import Oracle
    cfg = {"large30": configs.large30, "small8": configs.small8,
           "b8_w2_l4": lambda: configs.batch_member(8.0, 2, 4),
           "b8_w5_l5": lambda: configs.batch_member(8.0, 5, 5),
           "b10_w3_l2": lambda: configs.batch_member(10.0, 3, 2)}[cfgname]()
    r = tomo.Recon(cfg)
    o = Oracle(r, nthread=8)
    g = tomo.GpuTomo(r, maxit=cfg["maxit"])
    g.tune_set(0, 0, 0, 0, path)
    print(cfgname, "info", g.info(), flush=True)
    x = np.random.default_rng(6).standard_normal(r.nx)
    off = r.layer_offsets()
    for l in range(r.npsr):
        x[off[l]:off[l + 1]] -= x[off[l]:off[l + 1]].mean()
    ep = float(np.linalg.norm(g.P(x) - o.fdpcg(x)) / np.linalg.norm(o.fdpcg(x)))
    grad = workload.make_inputs(cfg, r, 1)
    xg, resg, rc = g.tomo(grad)
    xo, reso = o.solve(grad, cfg["maxit"])
    e = float(np.linalg.norm(xg - xo) / np.linalg.norm(xo))
    print("%s column fft %d: P rel %.3e, solve rel %.3e" % (cfgname, path, ep, e), flush=True)
    assert ep <= 2e-5
    assert e <= (1e-4 if cfgname in ("large30", "small8") else 1e-3)
    g.close()


@pytest.mark.parametrize("flights", [2, 4])
def test_flights_bit_exact(large30, flights):
    """A graph solve (gpu_tomo_solve_dev mode 1) of 9 systems as 1 flight and
    as 2 or 4 concurrent flights (ranges 5 + 4, 3 + 2 + 2 + 2): identical
    solutions and residuals.  The plan's own buffers (set_grad / get_x)."""
    cfg, r = large30
    n = 9
    grads = _batch(r, cfg, n)
    g = tomo.GpuTomo(r, nbatch=n, maxit=cfg["maxit"])
    out = {}
    for k in (1, flights):
        for i in range(n):
            g.set_grad(i, grads[i])
        g.tune_set(0, 0, 0, 0, 0, k)
        g.solve_dev(0, 0, nbatch=n, mode=1)
        out[k] = [g.get_x(i) for i in range(n)]
    for i in range(n):
        assert np.array_equal(out[1][i][0], out[flights][i][0]), "system %d differs" % i
        assert out[1][i][1] == out[flights][i][1]
    g.close()
