"""Benchmark: tomography solves/s and p99 latency on the 30 m / 6 LGS /
6 layer / 30-iteration FDPCG-PCG solve, one MI355X.

A step is one call of the solve on one batch of synthetic inputs already
resident in HBM.  Modes measured back to back on the same inputs:
  plain    one system per call, one launch per kernel (the "plain GPU" solve)
  graph    one system per call, the whole solve replayed as one hipGraph
  batched  --batch systems per call, one hipGraph (the headline `value`)
Latency is per call (a system in a batch waits for its whole batch).

Also reported: relative L2 of the GPU solution against the FP64 CPU oracle
(same inputs, same iteration count), a roofline entry for the dominant kernel
(algorithmic bytes per launch / HIP-event duration on the launch stream), and
the FP64 CPU oracle timed on the host cores (cpu_baseline).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "tomo solves/s + p99 latency (30m, 6 LGS, 6 layers, 30 PCG it) vs plain GPU"


def log(msg):
    """Progress on stderr (the result is the one JSON line on stdout)."""
    print("bench: " + msg, file=sys.stderr, flush=True)


def algorithmic_bytes(r, spec64=True):
    """Unique HBM bytes each kernel must move in one launch of an UPDATE
    iteration: (per system, shared by all systems of the launch).  Vectors:
    x/p FP32, r/Ap/q FP64, forward spectrum complex128 (complex64 with
    spec64=False: precision mode 2), inverse spectrum complex64; shared:
    operator geometry and FDPCG blocks, read once."""
    rec = r.rec
    nx = r.nx
    nw = sum(1 for w in range(rec.nwfs) if not rec.skip[w])
    nploc = rec.pnx * rec.pny
    nsa = sum(rec.nsa[w] for w in range(rec.nwfs) if not rec.skip[w])
    np2 = (rec.pos + 1) ** 2
    nxl, nyl = rec.xnx[0], rec.xny[0]
    nspec = rec.npsr * (nxl // 2 + 1) * nyl
    bs, nbh = rec.fd_bs, rec.fd_nbhalf
    cf = 16 if spec64 else 8  # bytes per forward-spectrum entry
    return {
        # read p, write q; GP stencils (x, y), NEA and the subaperture map
        "k_wfs": (4 * nx + 8 * nw * nploc, nsa * (2 * np2 * 4 + 16) + 4 * nw * nploc),
        # read q and p (L2), write Ap
        "k_layer": (8 * nw * nploc + 4 * nx + 8 * nx, 0),
        # read r, Ap; write r and the forward row spectrum
        "k_row": (8 * nx + 8 * nx + 8 * nx + cf * nspec, 0),
        # read the row spectrum, write the inverse-column spectrum; blocks
        "k_col": (cf * nspec + 8 * nspec, 8 * nbh * bs * bs),
        # read the spectrum, x and p; write x and p
        "k_irow": (8 * nspec + 8 * nx + 8 * nx, 0),
    }


def algorithmic_bytes_fp32(r):
    """The same per-launch figures with every vector stored in FP32 (and the
    forward spectrum in complex64): the per-iteration byte model of the solve
    with FP32 operators and vectors.  For the WFS pass (read p, write the
    WFS fields) and the layer pass (read the WFS fields and p, write Ap) these
    are the operator terms 4(Nx + W Np) + P_op and 4(W Np + 2 Nx) of that
    model; the stored widths of algorithmic_bytes() add the FP64 residual
    side of the default precision mode on top."""
    rec = r.rec
    nx = r.nx
    nw = sum(1 for w in range(rec.nwfs) if not rec.skip[w])
    nploc = rec.pnx * rec.pny
    nsa = sum(rec.nsa[w] for w in range(rec.nwfs) if not rec.skip[w])
    np2 = (rec.pos + 1) ** 2
    nxl, nyl = rec.xnx[0], rec.xny[0]
    nspec = rec.npsr * (nxl // 2 + 1) * nyl
    bs, nbh = rec.fd_bs, rec.fd_nbhalf
    return {
        "k_wfs": (4 * nx + 4 * nw * nploc, nsa * (2 * np2 * 4 + 16) + 4 * nw * nploc),
        "k_layer": (4 * nw * nploc + 4 * nx + 4 * nx, 0),
        "k_row": (4 * nx + 4 * nx + 4 * nx + 8 * nspec, 0),
        "k_col": (8 * nspec + 8 * nspec, 8 * nbh * bs * bs),
        "k_irow": (8 * nspec + 8 * nx + 8 * nx, 0),
    }


# compute types of the solve per precision mode (gpu_tomo_precision)
DTYPES = {0: "f32 operators/vectors/FFTs and residual recurrence, f64 reductions and alpha/beta",
          1: "f32 operators/vectors/inverse FFT, f64 residual recurrence (r, Ap, b, q)/forward FFT/reductions",
          2: "f32 operators/vectors/FFTs, f64 residual recurrence (r, Ap, b, q) and reductions",
          3: "f32 everywhere (reductions too)"}


def measured_traffic(kernel, batch):
    """HBM bytes per launch of `kernel` at this batch size, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (each scaled by the counter's
    reading on a copy of known size), or None when not measured."""
    path = os.path.join(ROOT, "profiles", "traffic_large30.json")
    try:
        with open(path) as f:
            t = json.load(f)
        k = t["kernels"][kernel]
        if int(t.get("batch", -1)) != batch:
            return None
        return int(k["traffic_bytes"])
    except (OSError, KeyError, ValueError):
        return None


KERNEL_CLASSES = ("k_ttr", "k_wfs", "k_layer", "k_row", "k_col", "k_irow", "k_final", "k_zero_stop", "k_copy")


def _pmc_rows(d, counter):
    """(kernel name, value) per dispatch from a rocprofv3 --pmc csv run."""
    import csv
    import glob
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit("no counter_collection.csv under " + d)
    out = []
    for fn in files:
        with open(fn) as f:
            for row in csv.DictReader(f):
                if row["Counter_Name"] == counter:
                    out.append((row["Kernel_Name"], float(row["Counter_Value"])))
    return out


def _kernel_class(name):
    base = name.split("(")[0].split("<")[0].replace("void ", "").strip().split("::")[-1]
    if base == "k_colr":  # the column stage in registers
        return "k_col"
    for k in sorted(KERNEL_CLASSES, key=len, reverse=True):
        if base == k or base.startswith(k + "_"):
            return k
    return None


def traffic_summary(d, batch):
    """Measured HBM bytes per launch of each solve kernel: the mean FETCH_SIZE
    and WRITE_SIZE (KiB per dispatch) of the class's most frequent
    instantiation, each scaled by bytes / reading of the same counter on an
    elementwise copy of 1 GiB (read 1 GiB, write 1 GiB)."""
    import collections
    copy_bytes = 4 << 28

    def calib(sub, counter):
        vals = [v for k, v in _pmc_rows(os.path.join(d, sub), counter) if "neg" in k]
        if not vals:
            raise SystemExit("no copy kernel in " + sub)
        return copy_bytes / (1024.0 * float(np.median(vals)))

    fr, wr = calib("copy_fetch", "FETCH_SIZE"), calib("copy_write", "WRITE_SIZE")
    runs = {}
    for sub, counter in (("fetch", "FETCH_SIZE"), ("write", "WRITE_SIZE")):
        by = collections.defaultdict(lambda: collections.defaultdict(list))
        for name, v in _pmc_rows(os.path.join(d, sub), counter):
            c = _kernel_class(name)
            if c:
                by[c][name].append(v)
        runs[sub] = by
    # one instantiation per class, the same in both runs: the most frequent
    # one of the fetch run
    res = {}
    for c, names in runs["fetch"].items():
        name = max(names.items(), key=lambda kv: len(kv[1]))[0]
        wvals = runs["write"].get(c, {}).get(name)
        if not wvals:
            continue
        e = res[c] = {"instantiation": name.split("(")[0].replace("void ", "")}
        for key, vals, scale in (("fetch_bytes", names[name], fr), ("write_bytes", wvals, wr)):
            e[key] = int(round(1024.0 * float(np.mean(vals)) * scale))
            e[key + "_dispatches"] = len(vals)
    for e in res.values():
        e["traffic_bytes"] = e.get("fetch_bytes", 0) + e.get("write_bytes", 0)
    out = {"batch": batch, "fetch_scale": round(fr, 4), "write_scale": round(wr, 4), "kernels": res}
    with open(os.path.join(ROOT, "profiles", "traffic_large30.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("copy calibration: bytes per counter byte  fetch %.4f  write %.4f" % (fr, wr))
    print("%-10s %12s %12s %12s" % ("kernel", "fetch MB", "write MB", "traffic MB"))
    for c in KERNEL_CLASSES:
        if c in res:
            e = res[c]
            print("%-10s %12.3f %12.3f %12.3f" % (c, e.get("fetch_bytes", 0) / 1e6, e.get("write_bytes", 0) / 1e6,
                                                 e["traffic_bytes"] / 1e6))


def copy_bandwidth():
    """Achievable HBM bandwidth: an elementwise copy of 1 GiB (read + write),
    timed with HIP events, best of 5."""
    import torch
    n = 1 << 28
    a = torch.ones(n, device="cuda:0")
    b = torch.empty_like(a)
    torch.neg(a, out=b)
    best = 1e30
    for _ in range(5):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        torch.neg(a, out=b)
        e1.record()
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1) * 1e-3)
    del a, b
    torch.cuda.empty_cache()
    return 2 * 4 * n / best / 1e9


def kernel_table(plan, ab, nb, d_g, d_x, stream, ab32=None):
    """Per-kernel-class mean launch time (HIP events on the launch stream,
    read-only classes 8 launches per event pair) and achieved algorithmic
    GB/s for one solve of nb systems: GBs on the FP32 byte model (ab32) and
    GBs_stored on the widths the default precision mode stores (ab)."""
    ab32 = ab32 or ab
    plan.profile(True, reps=8)
    plan.solve_dev(d_g, d_x, nbatch=nb, mode=0, stream=stream)
    prof = plan.profile_read()
    plan.profile(False)
    out = {}
    for k, (ms, calls) in prof.items():
        per, shared = ab32.get(k, (0, 0))
        by = per * nb + shared
        per_s, shared_s = ab.get(k, (0, 0))
        bys = per_s * nb + shared_s
        out[k] = dict(avg_us=1e3 * ms, calls=calls, total_ms=ms * calls, alg_MB=by / 1e6,
                      GBs=(by / (ms * 1e-3) / 1e9) if ms > 0 else 0.0, stored_MB=bys / 1e6,
                      GBs_stored=(bys / (ms * 1e-3) / 1e9) if ms > 0 else 0.0)
    return out


def bench_sweep(r, grads, maxit, sizes, copy_gbs, precision):
    """Batch sweep of the large30 solve: solves/s, p50 / p99 per call (100
    calls) and per-kernel achieved algorithmic bandwidth at several batch
    sizes (autotuned per size), in the headline's precision mode: the
    latency / throughput trade between one-at-a-time and large batches."""
    import torch
    from maos_amd import tomo
    ab = algorithmic_bytes(r, spec64=precision == 1)
    ab32 = algorithmic_bytes_fp32(r)
    out = {}
    for B in sizes:
        g = tomo.GpuTomo(r, nbatch=B, maxit=maxit)
        g.precision(precision)
        tuned = g.autotune(B)
        G = torch.tensor(np.stack([grads[i % len(grads)] for i in range(B)]), dtype=torch.float32).cuda()
        X = torch.zeros((B, r.nx), dtype=torch.float32, device="cuda:0")
        st = torch.cuda.current_stream().cuda_stream
        for _ in range(3):
            g.solve_dev(G.data_ptr(), X.data_ptr(), nbatch=B, mode=1, stream=st)
        torch.cuda.synchronize()
        lat = []
        for _ in range(100):
            t = time.perf_counter()
            g.solve_dev(G.data_ptr(), X.data_ptr(), nbatch=B, mode=1, stream=st)
            torch.cuda.synchronize()
            lat.append(time.perf_counter() - t)
        kt = kernel_table(g, ab, B, G.data_ptr(), X.data_ptr(), st, ab32)
        torch.cuda.synchronize()
        g.close()
        del G, X
        out[str(B)] = dict(solves_per_s=B * len(lat) / sum(lat), p50_ms=1e3 * float(np.percentile(lat, 50)),
                           p99_ms=1e3 * float(np.percentile(lat, 99)), calls=len(lat),
                           systems_per_workgroup=dict(wfs=tuned[0], layer=tuned[1], col=tuned[2]),
                           layer_tile_rows=tuned[3], column_fft={1: "registers", 2: "lds"}.get(tuned[4]),
                           flights=tuned[5],
                           kernels={k: dict(avg_us=round(v["avg_us"], 2), GBs=round(v["GBs"], 1),
                                            frac_of_peak=round(v["GBs"] / HBM_PEAK_GBS, 4),
                                            frac_of_peak_stored=round(v["GBs_stored"] / HBM_PEAK_GBS, 4),
                                            frac_of_copy=round(v["GBs"] / copy_gbs, 4))
                                    for k, v in kt.items() if k in ab})
    return out


def cpu_info():
    """Host CPU as /proc/cpuinfo names it: model name(s), vendor, family /
    model / stepping, the logical CPUs this process may run on and those the
    machine shows."""
    info = {}
    names = []
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if ":" not in line:
                    continue
                k, v = (s.strip() for s in line.split(":", 1))
                if k == "model name" and v not in names:
                    names.append(v)
                elif k in ("vendor_id", "cpu family", "model", "stepping") and k not in info:
                    info[k] = v
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return dict(model_name=" / ".join(names), vendor=info.get("vendor_id"),
                family_model_stepping="%s/%s/%s" % (info.get("cpu family"), info.get("model"), info.get("stepping")),
                logical_cpus_usable=usable, logical_cpus_visible=os.cpu_count())


def cpu_threads():
    """Threads for the CPU baseline: the host CPU share of one GPU on a
    shared GPU node (16 logical CPUs per GPU; OMP_NUM_THREADS is set to it
    there), or fewer where fewer are usable."""
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    share = int(os.environ.get("OMP_NUM_THREADS", "16") or 16)
    return max(1, min(usable, share, 16))


def cpu_baseline(r, grads, maxit, seconds):
    from oracle.oracle import Oracle
    cores = cpu_threads()
    o = Oracle(r, nthread=cores)
    x0, _ = o.solve(grads[0], maxit)  # warm-up + reference solution
    o.solve(grads[1 % len(grads)], maxit)
    t0 = time.perf_counter()
    n = 0
    lat = []
    while True:
        t = time.perf_counter()
        o.solve(grads[n % len(grads)], maxit)
        lat.append(time.perf_counter() - t)
        n += 1
        if time.perf_counter() - t0 > seconds and n >= 3:
            break
    el = time.perf_counter() - t0
    ci = cpu_info()
    return dict(value=n / el, unit="solves/s", cores=cores, kind="port", cpu_model=ci["model_name"], cpu=ci,
                threads_note="OpenMP threads = the host CPU share of one GPU (16 logical CPUs per GPU on the "
                             "node; %s usable, %s visible)" % (ci["logical_cpus_usable"], ci["logical_cpus_visible"]),
                p50_ms=1e3 * float(np.percentile(lat, 50)), p99_ms=1e3 * float(np.percentile(lat, 99)),
                sample="%d cold 30-iteration FDPCG solves of the 30 m system (RHS + PCG, FP64, "
                       "OpenMP over WFS/layers/blocks, %d threads), %.1f s; p50 %.1f ms p99 %.1f ms"
                       % (n, cores, el, 1e3 * np.percentile(lat, 50), 1e3 * np.percentile(lat, 99))), x0


def bench_uneven(steps, check=3):
    """The uneven batch: 64 systems with D in {8,10,20,30} m, 1-6 LGS, 2-6
    layers (seeded), 30 iterations each.  Identical configurations share a
    plan; one group launch per kernel per iteration covers all 64 (hipGraph).
    Compared with the same 64 systems solved one at a time (plain launches)."""
    import collections
    from maos_amd import configs, tomo, workload
    cfgs = configs.batch64()
    groups = collections.OrderedDict()
    for i, c in enumerate(cfgs):
        groups.setdefault(c["name"], []).append(i)
    plans, members, slots = {}, [], {}
    for name, idx in groups.items():
        cfg = cfgs[idx[0]]
        r = tomo.Recon(cfg)
        g = tomo.GpuTomo(r, nbatch=len(idx), maxit=cfg["maxit"])
        plans[name] = (cfg, r, g)
        members.append((g, 0, len(idx)))
        for s, i in enumerate(idx):
            slots[i] = (name, s)
    grads = {}
    for i, c in enumerate(cfgs):
        name, s = slots[i]
        cfg, r, g = plans[name]
        grads[i] = workload.make_inputs(cfg, r, 1000 + i)
        g.set_grad(s, grads[i])
    grp = tomo.GpuGroup(members)
    # one launch shape for the whole group, set on every plan (the plans'
    # solves alone below use it too)
    tuned = grp.autotune()
    for _ in range(2):
        grp.solve(mode=1)
    lat = []
    for _ in range(steps):
        t = time.perf_counter()
        grp.solve(mode=1)
        lat.append(time.perf_counter() - t)
    lat = np.array(lat)
    # one system at a time, plain launches
    singles = [tomo.GpuGroup([(plans[slots[i][0]][2], slots[i][1], 1)]) for i in range(len(cfgs))]
    for sg in singles:
        sg.solve(mode=0)
    t = time.perf_counter()
    for sg in singles:
        sg.solve(mode=0)
    seq = time.perf_counter() - t
    # accuracy of a few members (after the group solve above wrote x), each
    # next to the oracle's own sensitivity: how far the FP64 oracle moves
    # when the gradients are perturbed by 1e-15 (ill-conditioned small LGS
    # systems amplify that by up to 1e12 within 30 iterations)
    grp.solve(mode=1)
    errs = []
    from oracle.oracle import Oracle
    for i in range(check):
        name, s = slots[i]
        cfg, r, g = plans[name]
        x, res = g.get_x(s)
        o = Oracle(r, nthread=8)
        xo, reso = o.solve(grads[i], cfg["maxit"])
        gp = grads[i] * (1 + 1e-15 * np.random.default_rng(3).standard_normal(grads[i].shape))
        xp, _ = o.solve(gp, cfg["maxit"])
        errs.append(dict(system=name, rel_l2=float(np.linalg.norm(x - xo) / np.linalg.norm(xo)),
                         fp64_floor=float(np.linalg.norm(xp - xo) / np.linalg.norm(xo)),
                         residual_ratio=float(res / reso)))
    for sg in singles:
        sg.close()
    grp.close()
    n = len(cfgs)
    return dict(workload="64 uneven systems: D in {8,10,20,30} m, 1-6 LGS, 2-6 layers, 30 iterations",
                plans=len(plans), solves_per_s=n * len(lat) / lat.sum(),
                p50_ms=1e3 * float(np.percentile(lat, 50)), p99_ms=1e3 * float(np.percentile(lat, 99)),
                sequential_plain_solves_per_s=n / seq, speedup=float((n * len(lat) / lat.sum()) / (n / seq)),
                autotuned=dict(wfs=tuned[0], layer=tuned[1], layer_tile_rows=tuned[3],
                               column_fft={0: "registers where available", 2: "lds"}.get(tuned[4], tuned[4])),
                vs_oracle_first=errs)


def bench_precision(seeds=8, iters=(10, 30, 60)):
    """Precision study on the 30 m system: relative L2 of the GPU solution
    and residual-norm ratio against the FP64 CPU oracle after the same
    iteration count, per precision mode, over seeds (median and max)."""
    from maos_amd import configs, tomo, workload
    from oracle.oracle import Oracle
    cfg = configs.large30()
    r = tomo.Recon(cfg)
    o = Oracle(r, nthread=16)
    grads = [workload.make_inputs(cfg, r, 500 + k) for k in range(seeds)]
    ref = {(it, k): o.solve(grads[k], it) for it in iters for k in range(seeds)}
    modes = {4: "FP64 everywhere: vectors, operators, FDPCG transforms, reductions (all-FP64 GPU plan)",
             1: "FP32 operators, FP64 residual recurrence + forward spectrum + reductions (default)",
             2: "FP32 operators and spectrum, FP64 residual recurrence + reductions",
             0: "FP32 bulk, FP64 reductions and alpha/beta",
             3: "FP32 everywhere (reductions too)"}
    out = {"workload": "large30, %d seeds, cold start; reference: FP64 CPU oracle at the same iteration count" % seeds,
           "modes": {}}
    for mode, desc in modes.items():
        rows = {}
        for it in iters:
            g = tomo.GpuTomo(r, maxit=it)
            g.precision(mode)
            errs, ratios, ms = [], [], []
            for k in range(seeds):
                t0 = time.perf_counter()
                x, res, _ = g.tomo(grads[k])
                ms.append(1e3 * (time.perf_counter() - t0))
                xo, reso = ref[(it, k)]
                errs.append(float(np.linalg.norm(x - xo) / np.linalg.norm(xo)))
                ratios.append(float(res / reso))
            g.close()
            rows[str(it)] = dict(rel_l2_median=float(np.median(errs)), rel_l2_max=float(np.max(errs)),
                                 residual_ratio_median=float(np.median(ratios)),
                                 ms_per_host_solve_median=float(np.median(ms)))
        out["modes"][str(mode)] = dict(description=desc, iterations=rows)
    return out


def served_loops():
    """16 AO loops mixed from the three sizes: 4 x the 8 m NGS system, 2 x the
    30 m 6-LGS system, 10 x systems of the uneven batch."""
    from maos_amd import configs
    cf = [configs.small8() for _ in range(4)] + [configs.large30() for _ in range(2)]
    cf += configs.batch64()[:10]
    return cf


def bench_mvm(r, maxit, ncol=2048, batch=256):
    """Tomography half of MVM assembly on large30: the first ncol columns of
    T (x = T g), each the cold maxit-iteration solve of a unit gradient
    vector, batch columns per graph solve, written to host memory as
    doubles (gpu_tomo_mvm_tomo).  The launch shape is autotuned at that
    batch first."""
    from maos_amd import tomo
    g = tomo.GpuTomo(r, nbatch=batch, maxit=maxit)
    tuned = g.autotune(nsys=batch)
    g.mvm_tomo(0, batch)  # warm-up: graph capture at this batch
    t0 = time.perf_counter()
    T = g.mvm_tomo(0, ncol)
    dt = time.perf_counter() - t0
    g.close()
    return dict(workload="large30 T columns 0..%d (ng = %d, nx = %d), %d iterations, %d per graph solve"
                         % (ncol - 1, r.ng, r.nx, maxit, batch),
                columns_per_s=ncol / dt, seconds=dt, full_T_seconds_projected=r.ng / (ncol / dt),
                output_GB=T.nbytes / 1e9, nonzero_columns=int(np.count_nonzero(np.abs(T).sum(0))),
                autotuned=dict(wfs=tuned[0], layer=tuned[1], layer_tile_rows=tuned[3],
                               column_fft="registers" if tuned[4] != 2 else "lds", flights=tuned[5]))


def bench_update(cfg, r, maxit, reps=5):
    """A Cn2 profile change on large30 (new layer weights and r0): the host
    update (mt_setup_update_cn2: regulariser, FDPCG block pseudo inverses)
    and the GPU plan's re-upload (gpu_tomo_update_cn2), median of reps."""
    from maos_amd import tomo
    g = tomo.GpuTomo(r, nbatch=1, maxit=maxit)
    wt0, r00 = np.array(cfg["wt"], float), cfg["r0z"]
    th, tg = [], []
    for k in range(reps):
        w = wt0 * (1 + 0.1 * (k + 1) * np.arange(len(wt0)) / len(wt0))
        t0 = time.perf_counter()
        r.update_cn2(w / w.sum(), r00 * (1 + 0.02 * (k + 1)))
        t1 = time.perf_counter()
        g.update_cn2(r)
        t2 = time.perf_counter()
        th.append(t1 - t0)
        tg.append(t2 - t1)
    r.update_cn2(wt0 / wt0.sum(), r00)
    g.close()
    return dict(workload="large30 Cn2 update (6 layers, 9216 FDPCG blocks of 6 x 6)",
                host_update_ms=1e3 * float(np.median(th)), gpu_upload_ms=1e3 * float(np.median(tg)),
                host_threads=min(16, os.cpu_count() or 1))


def bench_fit(steps=20, batch=32, mvm_cols=512, mvm_batch=256):
    """DM fitting after the large30 tomography (two DMs at 0 and 11.8 km,
    the 9-direction 30" fitting ring, 100 cold CG iterations, FP64): one
    system per call, `batch` systems per call and 8 x `batch` per call (the
    same inputs repeated; five launches per CG iteration shared by the
    systems of a call), against the FP64 fitting oracle on the host cores;
    and MVM assembly (control-matrix columns: tomography then fit of each
    unit gradient vector, on the device)."""
    import torch
    from maos_amd import configs, tomo, workload
    from oracle.oracle import FitOracle
    base = configs.large30()
    fc = configs.fit_for(base)
    cfg = configs.with_fit_dirs(base, fc)
    r = tomo.Recon(cfg)
    f = tomo.FitSetup(r, fc)
    t = tomo.GpuTomo(r, nbatch=batch, maxit=cfg["maxit"])
    big = 8 * batch
    g = tomo.GpuFit(r, f, nbatch=big)
    grads = [workload.make_inputs(cfg, r, 700 + k) for k in range(batch)]
    X = t.tomo_batch(np.stack(grads))
    dx = torch.tensor(np.tile(X.astype(np.float32), (8, 1)), device="cuda:0")
    da = torch.zeros((big, f.nact), dtype=torch.float64, device="cuda:0")
    st = torch.cuda.current_stream().cuda_stream

    def timed(n, reps):
        lat = []
        for i in range(reps + 3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            g.solve_dev(dx.data_ptr(), da.data_ptr(), nsys=n, stream=st)
            torch.cuda.synchronize()
            if i >= 3:
                lat.append(time.perf_counter() - t0)
        lat = np.array(lat)
        return dict(solves_per_s=n * len(lat) / lat.sum(), p50_ms=1e3 * float(np.percentile(lat, 50)),
                    p99_ms=1e3 * float(np.percentile(lat, 99)))

    one = timed(1, max(100, steps))
    many = timed(batch, max(100, steps))
    bigr = timed(big, max(20, steps // 4))
    a_gpu = da[0].cpu().numpy()
    o = FitOracle(r, f, nthread=cpu_threads())
    xo = X[0].astype(np.float32).astype(np.float64)
    ao, reso = o.solve(xo, f.rec.maxit)
    lat = []
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < 5.0 or len(lat) < 3:
        t1 = time.perf_counter()
        o.solve(xo, f.rec.maxit)
        lat.append(time.perf_counter() - t1)
    out = dict(workload="large30 fit: 2 DMs (0, 11.8 km; %d actuators), 9 fitting directions (30\" ring), "
                        "floc %d points, %d CG iterations, cold, FP64" % (f.nact, f.nfloc, f.rec.maxit),
               one_system=one, batch=dict(systems=batch, **many), batch_large=dict(systems=big, **bigr),
               rel_l2_vs_fp64_oracle=float(np.linalg.norm(a_gpu - ao) / np.linalg.norm(ao)),
               cpu_oracle=dict(solves_per_s=len(lat) / sum(lat), p50_ms=1e3 * float(np.median(lat)),
                               threads=cpu_threads()))
    t.close()
    g.close()
    # MVM assembly of control-matrix columns (tomography + fit per column)
    tm = tomo.GpuTomo(r, nbatch=mvm_batch, maxit=cfg["maxit"])
    tm.autotune(mvm_batch)
    gm = tomo.GpuFit(r, f, nbatch=mvm_batch)
    tm.mvm_assemble(gm, 0, mvm_batch)  # warm-up (graph capture)
    t0 = time.perf_counter()
    M = tm.mvm_assemble(gm, 0, mvm_cols)
    dt = time.perf_counter() - t0
    out["mvm_assembly"] = dict(workload="control matrix F T columns 0..%d of %d (nact %d), %d per chunk"
                                        % (mvm_cols - 1, r.ng, f.nact, mvm_batch),
                               columns_per_s=mvm_cols / dt, seconds=dt,
                               full_matrix_seconds_projected=r.ng / (mvm_cols / dt),
                               nonzero_columns=int(np.count_nonzero(np.abs(M).sum(0))))
    tm.close()
    gm.close()
    return out


def fit_profile_run(n, calls=3):
    """Three cold large30 DM fits of n systems per call (synthetic layer
    inputs; the work of a fit does not depend on the values), each timed
    to the host, for a kernel trace of the fitting kernels."""
    import torch
    from maos_amd import configs, tomo
    base = configs.large30()
    fc = configs.fit_for(base)
    cfg = configs.with_fit_dirs(base, fc)
    r = tomo.Recon(cfg)
    f = tomo.FitSetup(r, fc)
    g = tomo.GpuFit(r, f, nbatch=n)
    X = 1e-7 * np.random.default_rng(1).standard_normal((n, r.nx))
    dx = torch.tensor(X.astype(np.float32), device="cuda:0")
    da = torch.zeros((n, f.nact), dtype=torch.float64, device="cuda:0")
    st = torch.cuda.current_stream().cuda_stream
    for _ in range(calls):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g.solve_dev(dx.data_ptr(), da.data_ptr(), nsys=n, stream=st)
        torch.cuda.synchronize()
        log("fit of %d systems: %.2f ms" % (n, 1e3 * (time.perf_counter() - t0)))
    g.close()


_STREAM_JOBS = None


def _stream_job(i):
    from maos_amd import tomo, workload
    c, seed, nframes = _STREAM_JOBS[i]
    return workload.frozen_flow_stream(c, tomo.Recon(c), seed, nframes)


def served_streams(nframes, seed=11):
    """The served workload's inputs: for each of the 16 loops, nframes
    distinct time steps of its own frozen-flow atmosphere (no frame repeats).
    Made before any GPU work, in forked worker processes (the host CPU share
    of one GPU), since the 30 m loops' slopes take ~0.2 s per frame; this
    process has not loaded the library (nor its GPU runtime) when it forks:
    each worker sets up its own loop geometry."""
    global _STREAM_JOBS
    import multiprocessing as mp
    cfgs = served_loops()
    _STREAM_JOBS = [(c, seed + i, nframes) for i, c in enumerate(cfgs)]
    # the 30 m loops first, so that the pool's tail is short
    order = sorted(range(len(cfgs)), key=lambda i: -cfgs[i]["aper_d"] * len(cfgs[i]["wfs"]))
    with mp.get_context("fork").Pool(cpu_threads()) as pool:
        res = dict(zip(order, pool.map(_stream_job, order, chunksize=1)))
    _STREAM_JOBS = None
    return cfgs, [res[i] for i in range(len(cfgs))]


def served_recons(cfgs):
    from maos_amd import tomo
    recons = {}
    for c in cfgs:
        if c["name"] not in recons:
            recons[c["name"]] = tomo.Recon(c)
    return recons


def served_accuracy(cfgs, recons, streams, which=(0, 4, 6), maxit=10):
    """Per-step error of the served (warm-start, maxit-iteration) solutions
    of three loops (an 8 m NGS, a 30 m 6-LGS and an uneven-batch loop) over
    every step of their streams, against (a) a cold 30-iteration FP64 oracle
    solve of the same step and (b) the FP64 oracle run through the same
    warm-start history (maxit iterations per step from its own previous
    solution).  The loops are replayed through a request loop of their own,
    one step at a time: a loop's results are the same bits alone and in any
    mix, so these are the served stream's solutions."""
    from maos_amd import tomo
    from oracle.oracle import Oracle
    srv = tomo.Server(mem_budget_gb=16.0, max_inflight=2)
    ids = [srv.add_loop(recons[cfgs[i]["name"]], maxit=maxit) for i in which]
    nstep = len(streams[which[0]])
    xs = {i: [] for i in which}
    for k in range(nstep):
        for j, i in enumerate(which):
            srv.submit(ids[j], streams[i][k], tag=k)
        srv.drain()
        srv.poll(4096)
        for j, i in enumerate(which):
            xs[i].append(srv.get_x(ids[j])[0])
    srv.close()
    out = {}
    for i in which:
        r = recons[cfgs[i]["name"]]
        o = Oracle(r, nthread=cpu_threads())
        xw = None
        e_cold, e_warm = [], []
        for k in range(nstep):
            g = streams[i][k]
            xc, _ = o.solve(g, 30)
            xw, _ = o.solve(g, maxit, x0=xw)
            x = xs[i][k]
            e_cold.append(float(np.linalg.norm(x - xc) / np.linalg.norm(xc)))
            e_warm.append(float(np.linalg.norm(x - xw) / np.linalg.norm(xw)))
        e_cold, e_warm = np.array(e_cold), np.array(e_warm)
        out[cfgs[i]["name"]] = dict(
            steps=nstep,
            vs_cold30_fp64=dict(median=float(np.median(e_cold)), p99=float(np.percentile(e_cold, 99)),
                                max=float(e_cold.max()), last10_median=float(np.median(e_cold[-10:]))),
            vs_warm_fp64_same_history=dict(median=float(np.median(e_warm)), p99=float(np.percentile(e_warm, 99)),
                                           max=float(e_warm.max()), first=float(e_warm[0]),
                                           last=float(e_warm[-1])))
    return out


def bench_served(cfgs, recons, streams, rate_hz=800.0, seed=11, low_rate_hz=400.0, plain_rate_hz=50.0):
    """Served stream: 16 loops at rate_hz with Poisson-jittered arrivals,
    time-correlated (frozen-flow, 10 m/s) gradients (every request its own
    time step), warm start and 10 iterations per step, through the request
    loop (continuous batching over several streams, graphs cached per loop
    set).  The comparator is the same arrival schedule through one-at-a-time
    plain solves (one request per launch sequence, one stream, no graph).
    Latency = submit to completion.  Both are also run at plain_rate_hz, a
    rate the one-at-a-time comparator sustains, so that its latency there is
    a solve's and not a queue's."""
    from maos_amd import tomo
    nframes = len(streams[0])
    T = 1.0 / rate_hz
    rng = np.random.default_rng(seed)
    base = []
    for i in range(len(cfgs)):
        ph = rng.uniform(0, T)
        for k in range(nframes):
            base.append((k + ph / T + min(rng.exponential(0.1), 0.5), i, k))
    base.sort()

    def run(max_inflight, max_batch, graph, hz):
        events = [(t / hz, i, k) for t, i, k in base]
        srv = tomo.Server(mem_budget_gb=32.0, max_inflight=max_inflight, max_batch=max_batch, use_graph=graph)
        ids = [srv.add_loop(recons[c["name"]], maxit=10) for c in cfgs]
        # warm-up: one step per loop (graph capture, first touch); the
        # request loop's batches are tuned for the admitted loops
        for i, lid in enumerate(ids):
            srv.submit(lid, streams[i][0], tag=-1)
        srv.drain()
        srv.poll(4096)
        tuned = srv.autotune() if graph else None

        def stream(evs, tagged):
            t0 = time.perf_counter()
            for t, i, k in evs:
                while time.perf_counter() - t0 < t:
                    pass
                srv.submit(ids[i], streams[i][k], tag=k if tagged else -1)
            srv.drain()
            return t0

        # warm-up: the first 50 ms of the schedule, untimed (the first
        # stream of a process otherwise pays a one-off stall of 15-30 ms)
        stream([e for e in events if e[0] <= 0.05], False)
        srv.poll(4096)
        done = []
        t0 = stream(events, True)
        t1 = time.perf_counter()
        while True:
            d = srv.poll(4096)
            if not d:
                break
            done.extend(d)
        srv.close()
        lat = np.array([d[3] for d in done]) * 1e-3
        out = dict(offered_per_s=len(cfgs) * hz, requests=len(done), solves_per_s=len(done) / (t1 - t0),
                   p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                   max_ms=float(lat.max()), wall_s=t1 - t0)
        if tuned:
            out["autotuned"] = dict(wfs=tuned[0], layer=tuned[1], layer_tile_rows=tuned[3],
                                    column_fft="registers" if tuned[4] != 2 else "lds")
        return out

    out = dict(workload="16 loops (4 x 8 m NGS, 2 x 30 m 6-LGS, 10 x uneven-batch systems), frozen flow 10 m/s, "
                        "%d distinct time steps per loop, Poisson-jittered arrivals, warm start, 10 iterations"
                        % nframes)
    for hz in (rate_hz, low_rate_hz, plain_rate_hz):
        if not hz:
            continue
        out["at_%g_hz" % hz] = dict(served=run(3, 64, True, hz), one_at_a_time_plain=run(1, 1, False, hz))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precision", type=int, default=2, choices=(0, 1, 2, 3),
                    help="precision mode of the headline, plain, graph and sweep solves (default 2: FP32 "
                         "operators, vectors and FFTs, FP64 residual recurrence and reductions, within 1e-4 of "
                         "the FP64 oracle at 30 iterations; the library default is 1; mode 4, all FP64, serves "
                         "host calls only and is timed in the precision study)")
    ap.add_argument("--modes", default="plain,graph,batched,uneven,served,mvm,fit,update,precision,sweep",
                    help="comma list of modes to time (profiling runs use one)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 3 device copies of 1 GiB (known HBM traffic) for counter scaling")
    ap.add_argument("--systems-per-workgroup", metavar="W,L,C[,ROWS[,COLK[,FLIGHTS]]]", default=None,
                    help="fix the systems per workgroup of the WFS, layer and column kernels (and the layer "
                         "tile rows) instead of autotuning (profiling runs, whose timing the counters distort)")
    ap.add_argument("--fit-profile", metavar="N", type=int, default=None,
                    help="only run 3 cold large30 DM fits of N systems per call (for a rocprofv3 kernel trace)")
    ap.add_argument("--traffic-summary", metavar="DIR", default=None,
                    help="no GPU: fold rocprofv3 FETCH_SIZE / WRITE_SIZE runs under DIR/{fetch,write,"
                         "copy_fetch,copy_write} into profiles/traffic_large30.json")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = 1 << 28  # floats: 1 GiB
        a = torch.ones(n, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            torch.neg(a, out=b)  # one elementwise kernel: reads a, writes b
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "copies": 3}))
        return
    if args.traffic_summary:
        traffic_summary(args.traffic_summary, args.batch)
        return
    if args.fit_profile:
        fit_profile_run(args.fit_profile)
        return
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the solve runs on one MI355X; "
              "independent solves would shard across GPUs with no communication). "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    modes = set(args.modes.split(","))
    # the served stream's inputs are made first, in forked workers, before
    # this process loads the library and its GPU runtime
    served_inputs = served_streams(120) if "served" in modes else None
    log("served stream inputs ready")

    import torch

    from maos_amd import configs, tomo, workload

    cfg = configs.large30()
    maxit = cfg["maxit"]
    B = args.batch
    r = tomo.Recon(cfg)
    grads = [workload.make_inputs(cfg, r, seed) for seed in range(1, B + 1)]
    dev = torch.device("cuda:0")
    d_g = torch.tensor(np.stack(grads).astype(np.float32), device=dev).contiguous()
    d_x = torch.zeros((B, r.nx), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream(dev)
    sp = stream.cuda_stream

    def run_mode(plan, nb, mode, steps, warmup):
        lat = []
        for i in range(warmup + steps):
            j = (i * nb) % B
            if j + nb > B:
                j = 0
            torch.cuda.synchronize()
            t = time.perf_counter()
            plan.solve_dev(d_g[j].data_ptr(), d_x[j].data_ptr(), nbatch=nb, mode=mode, stream=sp)
            torch.cuda.synchronize()
            if i >= warmup:
                lat.append(time.perf_counter() - t)
        lat = np.array(lat)
        return dict(solves_per_s=nb * len(lat) / lat.sum(), p50_ms=1e3 * float(np.percentile(lat, 50)),
                    p99_ms=1e3 * float(np.percentile(lat, 99)), ms_per_step=1e3 * float(lat.mean()))

    plain = tomo.GpuTomo(r, nbatch=1, maxit=maxit)
    batched = tomo.GpuTomo(r, nbatch=B, maxit=maxit)
    plain.precision(args.precision)
    batched.precision(args.precision)
    if args.systems_per_workgroup:
        tuned = tuple(int(v) for v in args.systems_per_workgroup.split(","))
        batched.tune_set(*tuned)
        tuned_plain = None
    else:
        tuned = batched.autotune(B)
        # the one-at-a-time comparator gets its own best launch shape too
        tuned_plain = plain.autotune(1) if ("plain" in modes or "graph" in modes) else None
    empty = dict(solves_per_s=float("nan"), p50_ms=float("nan"), p99_ms=float("nan"), ms_per_step=float("nan"))
    # latency percentiles need enough calls: at least 100 per mode
    ncall = max(100, args.steps)
    res_plain = run_mode(plain, 1, 0, ncall, args.warmup) if "plain" in modes else empty
    res_graph = run_mode(plain, 1, 1, ncall, args.warmup) if "graph" in modes else empty
    # headline: exactly --warmup untimed and --steps timed calls, the timed
    # region bracketed by a device synchronize on both sides (one GPU: no
    # other rank to wait for)
    torch.cuda.synchronize()
    res_batch = run_mode(batched, B, 1, args.steps, args.warmup)
    torch.cuda.synchronize()
    # p50 / p99 of the batched call over >= 100 calls (the headline's own
    # calls when --steps is at least 100)
    res_lat = res_batch if args.steps >= 100 else run_mode(batched, B, 1, ncall, 0)
    log("headline: %.1f solves/s" % res_batch["solves_per_s"])

    # accuracy of system 0 (batched path) against the FP64 CPU oracle
    x_gpu = d_x[0].double().cpu().numpy()
    res_gpu = batched.residuals()[0]

    # per-kernel profile of one batched solve (plain launches, HIP events on
    # the launch stream; the read-only kernel classes launched 8 times back
    # to back per event pair, so the average excludes the event overhead)
    ab = algorithmic_bytes(r, spec64=args.precision == 1)
    ab32 = algorithmic_bytes_fp32(r)
    kernels = kernel_table(batched, ab, B, d_g.data_ptr(), d_x.data_ptr(), sp, ab32)
    torch.cuda.synchronize()
    dom = max((k for k in kernels if k in ab), key=lambda k: kernels[k]["total_ms"])
    kd = kernels[dom]
    copy_gbs = copy_bandwidth()
    # achieved = the FP32 byte model of the dominant kernel's launch / its
    # mean duration; the bytes the default precision mode stores (FP64
    # residual side) are reported beside it
    roofline = dict(bound="hbm", achieved=round(kd["GBs"], 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(kd["GBs"] / HBM_PEAK_GBS, 4), traffic=measured_traffic(dom, B), kernel=dom,
                    avg_us=round(kd["avg_us"], 2), alg_bytes_per_launch=int(ab32[dom][0] * B + ab32[dom][1]),
                    stored_bytes_per_launch=int(ab[dom][0] * B + ab[dom][1]),
                    achieved_stored=round(kd["GBs_stored"], 1), frac_stored=round(kd["GBs_stored"] / HBM_PEAK_GBS, 4),
                    copy_GBs=round(copy_gbs, 1), frac_of_copy=round(kd["GBs"] / copy_gbs, 4))

    log("kernel table and copy bandwidth done")
    sweep = None
    if "sweep" in modes:
        try:
            sweep = bench_sweep(r, grads, maxit, (4, 8, 16, 64, 128, 256), copy_gbs, args.precision)
        except Exception as e:  # reported, not fatal for the headline line
            sweep = {"error": repr(e)}

    log("sweep done")
    uneven = None
    if "uneven" in modes:
        try:
            uneven = bench_uneven(max(5, args.steps // 20))
        except Exception as e:  # reported, not fatal for the headline line
            uneven = {"error": repr(e)}

    log("uneven done")
    precision = None
    if "precision" in modes:
        try:
            precision = bench_precision()
        except Exception as e:  # reported, not fatal for the headline line
            precision = {"error": repr(e)}

    log("precision done")
    update = None
    if "update" in modes:
        try:
            update = bench_update(cfg, r, maxit)
        except Exception as e:
            update = {"error": repr(e)}

    log("update done")
    mvm = None
    if "mvm" in modes:
        try:
            mvm = bench_mvm(r, maxit)
        except Exception as e:
            mvm = {"error": repr(e)}

    log("mvm done")
    fit = None
    if "fit" in modes:
        try:
            fit = bench_fit(args.steps)
        except Exception as e:  # reported, not fatal for the headline line
            fit = {"error": repr(e)}

    log("fit done")
    served = None
    if "served" in modes:
        try:
            cfgs_s, streams_s = served_inputs
            recons_s = served_recons(cfgs_s)
            served = bench_served(cfgs_s, recons_s, streams_s)
            served["per_step_error"] = served_accuracy(cfgs_s, recons_s, streams_s)
        except Exception as e:  # reported, not fatal for the headline line
            served = {"error": repr(e)}

    log("served done")
    cpu = None
    err = None
    if not args.no_cpu:
        cpu, x_ref = cpu_baseline(r, grads, maxit, args.cpu_seconds)
        err = float(np.linalg.norm(x_gpu - x_ref) / np.linalg.norm(x_ref))

    out = {
        "metric": METRIC,
        "value": round(res_batch["solves_per_s"], 2),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res_batch["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPES[args.precision],
        "data": "synthetic (seeded von Karman screens, geometric SH slopes + 0.05\" noise)",
        "config": {"workload": "large30: D=30 m, 6 LGS at 90 km, 6 layers 0-16 km, 60x60 subaps, "
                               "xloc 6x96x96, 30 FDPCG-PCG iterations, cold start",
                   "batch": B, "global_batch": B, "parallelism": "batch of independent solves",
                   "precision_mode": args.precision,
                   "autotuned_systems_per_workgroup": {"wfs": tuned[0], "layer": tuned[1], "col": tuned[2]},
                   "layer_tile_rows": tuned[3] if len(tuned) > 3 else None,
                   "column_fft": {1: "registers", 2: "lds"}.get(tuned[4]) if len(tuned) > 4 else None,
                   "flights": tuned[5] if len(tuned) > 5 else None},
        "p50_ms": round(res_lat["p50_ms"], 4),
        "p99_ms": round(res_lat["p99_ms"], 4),
        "latency_calls": ncall,
        "modes": {"plain": res_plain, "graph": res_graph, "batched": res_batch,
                  "batched_latency": res_lat},
        "plain_launch_shape_autotuned": list(tuned_plain) if tuned_plain is not None else None,
        "speedup_vs_plain": round(res_batch["solves_per_s"] / res_plain["solves_per_s"], 2),
        "accuracy": {"rel_l2_vs_fp64_oracle": err, "cg_residual": float(res_gpu)},
        "roofline": roofline,
        "kernels": kernels,
        "cpu_baseline": cpu,
        "uneven64": uneven,
        "served16": served,
        "mvm_tomography": mvm,
        "dm_fit": fit,
        "cn2_update": update,
        "precision_study": precision,
        "batch_sweep": sweep,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
