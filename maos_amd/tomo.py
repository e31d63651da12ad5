"""Host-side mirror of the reference's tomography interface.

  Recon        ~ setup_recon_prep + setup_recon_tomo + fdpcg_prepare
                 (maos/setup_recon_prep.c:1043, maos/setup_recon.c:1212,
                 maos/fdpcg.c:331): FP64 operators in C++ (libmaos_tomo.so).
  GpuTomo      ~ gpu_setup_recon / gpu_tomo / gpu_recon_reset /
                 gpu_update_recon_cn2 (cuda/recon.cu:477-582) and the operator
                 entry points cutomo_grid::L/R (cuda/tomo.cu:607-677),
                 cufdpcg_t::P (cuda/fdpcg.cu:168-226).

Vectors are numpy float64 arrays in MAOS dcell order (see include/maos_tomo.h).
The GPU path runs only through the HIP library; nothing here computes the
solve on the CPU.
"""
import ctypes as C

import numpy as np

from . import _abi
from .configs import FitParms, Parms


def _arr(ptr, n, dtype=np.float64):
    if not ptr or n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


def _dp(a):
    return a.ctypes.data_as(_abi.dp)


class Recon:
    """Reconstruction geometry + operators built from .conf parameters."""

    def __init__(self, cfg):
        self.cfg = cfg
        self._parms = Parms(cfg)
        L = _abi.lib()
        h = C.c_void_p()
        _abi.check(L.mt_setup_create(C.byref(self._parms.p), C.byref(h)), "mt_setup_create")
        self._h = h
        self._refresh()

    def _refresh(self):
        L = _abi.lib()
        self.rec = _abi.MtRecon()
        _abi.check(L.mt_setup_recon(self._h, C.byref(self.rec)), "mt_setup_recon")
        nx, ng = C.c_long(), C.c_long()
        L.mt_setup_sizes(self._h, C.byref(nx), C.byref(ng))
        self.nx, self.ng = nx.value, ng.value

    def update_cn2(self, wt, r0z):
        """setup_recon_tomo_update: new layer weights / r0, rebuild L2 + FDPCG."""
        w = np.ascontiguousarray(wt, dtype=np.float64)
        _abi.check(_abi.lib().mt_setup_update_cn2(self._h, _dp(w), float(r0z)), "update_cn2")
        self._refresh()

    def close(self):
        if getattr(self, "_h", None):
            _abi.lib().mt_setup_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- numpy views (copies) of the operators -------------------------
    @property
    def npsr(self):
        return self.rec.npsr

    @property
    def nwfs(self):
        return self.rec.nwfs

    def layer_shapes(self):
        r = self.rec
        return [(r.xny[l], r.xnx[l]) for l in range(r.npsr)]

    def layer_offsets(self):
        off = [0]
        for ny, nx in self.layer_shapes():
            off.append(off[-1] + nx * ny)
        return off

    def grad_offsets(self):
        off = [0]
        for w in range(self.nwfs):
            off.append(off[-1] + 2 * self.rec.nsa[w])
        return off

    def nsa(self, w):
        return self.rec.nsa[w]

    def saloc(self, w):
        n = self.rec.nsa[w]
        a = _arr(self.rec.saloc[w], 2 * n)
        return a[:n], a[n:]

    def saa(self, w):
        return _arr(self.rec.saa[w], self.rec.nsa[w])

    def gp(self, w):
        """GP of WFS w as scipy.sparse.csc_matrix (2nsa x nploc), MAOS dsp."""
        import scipy.sparse as sp
        r = self.rec
        nploc = r.pnx * r.pny
        p = _arr(r.gp_p[w], nploc + 1, np.int64)
        i = _arr(r.gp_i[w], p[-1], np.int64)
        x = _arr(r.gp_x[w], p[-1])
        return sp.csc_matrix((x, i, p), shape=(2 * r.nsa[w], nploc))

    def neai(self, w):
        return _arr(self.rec.neai[w], 3 * self.rec.nsa[w]).reshape(-1, 3)

    def ptt(self, w):
        if not self.rec.trs[w] or not self.rec.ptt[w]:
            return None
        return _arr(self.rec.ptt[w], 4 * self.rec.nsa[w]).reshape(-1, 2).T

    def l2c(self):
        return _arr(self.rec.l2c, self.rec.npsr)

    def fdpcg(self):
        r = self.rec
        bs, nb, nbh = r.fd_bs, r.fd_nb, r.fd_nbhalf
        M = _arr(r.fd_Minv, nb * bs * bs * 2).view(np.complex128).reshape(nb, bs, bs)
        Minv = np.transpose(M, (0, 2, 1)).copy()  # column-major -> [b][row][col]
        mh = _abi.dp()
        _abi.lib().mt_setup_fdpcg_mhat(self._h, C.byref(mh))
        Mh = _arr(mh, nb * bs * bs * 2).view(np.complex128).reshape(nb, bs, bs)
        Mhat = np.transpose(Mh, (0, 2, 1)).copy()
        return dict(bs=bs, nb=nb, nbhalf=nbh, scale=r.fd_scale, Minv=Minv, Mhat=Mhat,
                    perm=_arr(r.fd_perm, nb * bs, np.int64),
                    permhf=_arr(r.fd_permhf, nbh * bs, np.int64))


class GpuTomo:
    """GPU tomography plan for `nbatch` systems sharing one geometry."""

    def __init__(self, recon, device=0, nbatch=1, maxit=30, warm=0):
        self.recon = recon
        L = _abi.lib()
        h = C.c_void_p()
        _abi.check(L.gpu_tomo_setup(C.byref(recon.rec), device, nbatch, maxit, int(warm),
                                    C.byref(h)), "gpu_tomo_setup")
        self._h = h
        self.nbatch, self.maxit, self.warm = nbatch, maxit, warm
        nx, ng = C.c_long(), C.c_long()
        L.gpu_tomo_sizes(h, C.byref(nx), C.byref(ng))
        self.nx, self.ng = nx.value, ng.value

    def precision(self, mode):
        """0: FP32 residual and spectrum; 1: FP64 residual recurrence and
        forward spectrum (default); 2: FP64 residual, FP32 spectrum; 3: as 0
        with FP32 dot products and alpha/beta (FP32 everywhere); 4: all FP64
        (vectors, operators, FDPCG transforms, reductions), through the host
        calls (tomo, tomo_batch, L, R, P, get_x, residuals) only."""
        _abi.check(_abi.lib().gpu_tomo_precision(self._h, int(mode)), "gpu_tomo_precision")

    def info(self):
        """gpu_tomo_info: dict(matched, fnx, fny, bs, rmax)."""
        v = [C.c_int() for _ in range(5)]
        _abi.check(_abi.lib().gpu_tomo_info(self._h, *[C.byref(a) for a in v]), "gpu_tomo_info")
        return dict(zip(("matched", "fnx", "fny", "bs", "rmax"), [a.value for a in v]))

    def autotune(self, nsys=None):
        """gpu_tomo_autotune: pick the layer tile shape, the column kernel
        and the systems per workgroup of the WFS, layer and column kernels by
        timing, then the flights of a graph solve; returns (wfs, layer, col,
        layer tile rows, column kernel: 1 registers, 2 LDS stages, flights)."""
        ch = (C.c_int * 6)()
        _abi.check(_abi.lib().gpu_tomo_autotune(self._h, int(nsys or self.nbatch), ch), "gpu_tomo_autotune")
        return tuple(ch)

    def tune_set(self, wfs, layer, col, layer_rows=0, col_fft=0, flights=0):
        """gpu_tomo_tune_set: fix the systems per workgroup (0: heuristic),
        the layer tile rows (16, 32, 64; 0: 64 where it fits), the FFT kernels
        (1 registers where they apply, 2 LDS stages for columns and rows;
        0: as 1) and the
        flights of a graph solve (1, 2, 4; 0: 1)."""
        ch = (C.c_int * 6)(int(wfs), int(layer), int(col), int(layer_rows), int(col_fft), int(flights))
        _abi.check(_abi.lib().gpu_tomo_tune_set(self._h, ch), "gpu_tomo_tune_set")

    def close(self):
        if getattr(self, "_h", None):
            _abi.lib().gpu_tomo_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def tomo(self, grad, isys=0):
        """gpu_tomo: one solve of slot isys from host gradients.
        Returns (x, cgres, status)."""
        g = np.ascontiguousarray(grad, dtype=np.float64)
        if g.size != self.ng:
            raise ValueError("gradient vector has %d entries, expected %d" % (g.size, self.ng))
        x = np.zeros(self.nx)
        res = C.c_double()
        rc = _abi.lib().gpu_tomo_solve(self._h, int(isys), _dp(g), _dp(x), C.byref(res))
        _abi.check(rc, "gpu_tomo_solve")
        return x, res.value, rc

    def tomo_batch(self, grads):
        """gpu_tomo_solve_batch: solve the stacked systems grads (nsys x ng)
        in one batched call.  Returns x (nsys x nx); residuals via
        residuals()."""
        g = np.ascontiguousarray(grads, dtype=np.float64)
        if g.ndim != 2 or g.shape[1] != self.ng or g.shape[0] > self.nbatch:
            raise ValueError("gradients must be (nsys <= %d) x %d" % (self.nbatch, self.ng))
        x = np.zeros((g.shape[0], self.nx))
        res = np.zeros(g.shape[0])
        _abi.check(_abi.lib().gpu_tomo_solve_batch(self._h, g.shape[0], _dp(g), _dp(x), _dp(res)),
                   "gpu_tomo_solve_batch")
        return x

    def set_grad(self, isys, grad):
        """Host gradients into system slot isys (gpu_tomo_set_grad)."""
        g = np.ascontiguousarray(grad, dtype=np.float64)
        if g.size != self.ng:
            raise ValueError("gradient vector has %d entries, expected %d" % (g.size, self.ng))
        _abi.check(_abi.lib().gpu_tomo_set_grad(self._h, int(isys), _dp(g)), "gpu_tomo_set_grad")

    def get_x(self, isys):
        """Solution and residual of slot isys after a solve (gpu_tomo_get_x)."""
        x = np.zeros(self.nx)
        res = C.c_double()
        _abi.check(_abi.lib().gpu_tomo_get_x(self._h, int(isys), _dp(x), C.byref(res)), "gpu_tomo_get_x")
        return x, res.value

    def reset(self):
        _abi.check(_abi.lib().gpu_tomo_reset(self._h), "gpu_tomo_reset")

    def update_cn2(self, recon):
        _abi.check(_abi.lib().gpu_tomo_update_cn2(self._h, C.byref(recon.rec)), "update_cn2")

    def L(self, xin, alpha=1.0, beta=0.0, xout=None):
        xi = np.ascontiguousarray(xin, dtype=np.float64)
        xo = np.zeros(self.nx) if xout is None else np.ascontiguousarray(xout, dtype=np.float64)
        _abi.check(_abi.lib().gpu_tomo_L(self._h, _dp(xi), _dp(xo), alpha, beta), "gpu_tomo_L")
        return xo

    def R(self, grad, alpha=1.0):
        g = np.ascontiguousarray(grad, dtype=np.float64)
        xo = np.zeros(self.nx)
        _abi.check(_abi.lib().gpu_tomo_R(self._h, _dp(g), _dp(xo), alpha), "gpu_tomo_R")
        return xo

    def P(self, xin):
        xi = np.ascontiguousarray(xin, dtype=np.float64)
        xo = np.zeros(self.nx)
        _abi.check(_abi.lib().gpu_tomo_P(self._h, _dp(xi), _dp(xo)), "gpu_tomo_P")
        return xo

    def mvm_tomo(self, col0=0, ncol=None):
        """gpu_tomo_mvm_tomo: columns [col0, col0 + ncol) of T (x = T g),
        each the solve of a unit gradient vector; (nx, ncol) array."""
        ng = self.recon.ng
        ncol = ng - col0 if ncol is None else ncol
        out = np.zeros((ncol, self.recon.nx))
        _abi.check(_abi.lib().gpu_tomo_mvm_tomo(self._h, int(col0), int(ncol), _dp(out)), "gpu_tomo_mvm_tomo")
        return out.T

    def mvm_assemble(self, fit, col0=0, ncol=None):
        """gpu_mvm_assemble: columns [col0, col0 + ncol) of the control
        matrix F T (fit of the tomography of each unit gradient vector);
        (nact, ncol) array.  fit: a GpuFit on the same layers."""
        ng = self.recon.ng
        ncol = ng - col0 if ncol is None else ncol
        out = np.zeros((ncol, fit.fit.nact))
        _abi.check(_abi.lib().gpu_mvm_assemble(self._h, fit._h, int(col0), int(ncol), _dp(out)), "gpu_mvm_assemble")
        return out.T

    def solve_dev(self, d_grad, d_x, nbatch=None, mode=0, stream=None):
        """Batched solve on device buffers (int device pointers, float32)."""
        nb = self.nbatch if nbatch is None else nbatch
        _abi.check(_abi.lib().gpu_tomo_solve_dev(self._h, C.c_void_p(d_grad), C.c_void_p(d_x), nb,
                                                 mode, C.c_void_p(stream or 0)),
                   "gpu_tomo_solve_dev")

    def residuals(self, nbatch=None):
        nb = self.nbatch if nbatch is None else nbatch
        out = np.zeros(nb)
        _abi.check(_abi.lib().gpu_tomo_residuals(self._h, _dp(out), nb), "gpu_tomo_residuals")
        return out

    def profile(self, enable=True, reps=1):
        """Enable per-kernel-class timing; reps > 1 repeats the read-only
        classes back to back inside one event pair."""
        _abi.check(_abi.lib().gpu_tomo_profile(self._h, max(1, int(reps)) if enable else 0), "gpu_tomo_profile")

    def profile_read(self):
        n = C.c_int(16)
        names = (C.c_char_p * 16)()
        ms = (C.c_double * 16)()
        calls = (C.c_double * 16)()
        _abi.lib().gpu_tomo_profile_read(self._h, names, ms, calls, C.byref(n))
        return {names[i].decode(): (ms[i], int(calls[i])) for i in range(n.value)}


class GpuGroup:
    """Uneven batch (gpu_tomo_group_*): system ranges of several plans of
    any geometry solved in lockstep, one launch per kernel per iteration."""

    def __init__(self, members):
        """members: list of (GpuTomo, s0, nsys)."""
        self.members = list(members)
        n = len(self.members)
        plans = (C.c_void_p * n)(*[m[0]._h.value for m in self.members])
        s0 = (C.c_int * n)(*[int(m[1]) for m in self.members])
        ns = (C.c_int * n)(*[int(m[2]) for m in self.members])
        h = C.c_void_p()
        _abi.check(_abi.lib().gpu_tomo_group_create(n, plans, s0, ns, C.byref(h)), "gpu_tomo_group_create")
        self._h = h

    def solve(self, mode=0, stream=None):
        _abi.check(_abi.lib().gpu_tomo_group_solve(self._h, int(mode), C.c_void_p(stream or 0)),
                   "gpu_tomo_group_solve")

    def autotune(self):
        """gpu_tomo_group_autotune: the fastest launch shape for the whole
        group, set on every member plan; returns it in tune_set's order."""
        ch = (C.c_int * 6)()
        _abi.check(_abi.lib().gpu_tomo_group_autotune(self._h, ch), "gpu_tomo_group_autotune")
        return tuple(ch)

    def close(self):
        if getattr(self, "_h", None):
            _abi.lib().gpu_tomo_group_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Request loop with continuous batching (gpu_tomo_server_*)."""

    def __init__(self, device=0, mem_budget_gb=16.0, max_inflight=2, max_batch=64, use_graph=True):
        h = C.c_void_p()
        _abi.check(_abi.lib().gpu_tomo_server_create(device, float(mem_budget_gb), int(max_inflight),
                                                     int(max_batch), 1 if use_graph else 0, C.byref(h)),
                   "gpu_tomo_server_create")
        self._h = h
        self.nx = {}
        self.ng = {}

    def add_loop(self, recon, maxit=10):
        """Admit an AO loop (warm start, `maxit` iterations per step); raises
        MtError with code MT_EFULL (-6) when the HBM budget is exceeded."""
        lid = C.c_int()
        _abi.check(_abi.lib().gpu_tomo_server_add_loop(self._h, C.byref(recon.rec), int(maxit), C.byref(lid)),
                   "gpu_tomo_server_add_loop")
        self.nx[lid.value] = recon.nx
        self.ng[lid.value] = recon.ng
        return lid.value

    def remove_loop(self, lid):
        _abi.check(_abi.lib().gpu_tomo_server_remove_loop(self._h, int(lid)), "gpu_tomo_server_remove_loop")

    def submit(self, lid, grad, tag=0):
        g = np.ascontiguousarray(grad, dtype=np.float64)
        if g.size != self.ng[lid]:
            raise ValueError("gradient vector has %d entries, expected %d" % (g.size, self.ng[lid]))
        _abi.check(_abi.lib().gpu_tomo_server_submit(self._h, int(lid), _dp(g), int(tag)), "gpu_tomo_server_submit")

    def poll(self, maxn=1024):
        """Completed requests since the last poll: list of (loop, tag, cgres, latency_us)."""
        loop = (C.c_int * maxn)()
        tag = (C.c_long * maxn)()
        res = (C.c_double * maxn)()
        lat = (C.c_double * maxn)()
        n = C.c_int()
        _abi.check(_abi.lib().gpu_tomo_server_poll(self._h, maxn, loop, tag, res, lat, C.byref(n)),
                   "gpu_tomo_server_poll")
        return [(loop[i], tag[i], res[i], lat[i]) for i in range(n.value)]

    def drain(self):
        _abi.check(_abi.lib().gpu_tomo_server_drain(self._h), "gpu_tomo_server_drain")

    def get_x(self, lid):
        x = np.zeros(self.nx[lid])
        res = C.c_double()
        _abi.check(_abi.lib().gpu_tomo_server_get_x(self._h, int(lid), _dp(x), C.byref(res)),
                   "gpu_tomo_server_get_x")
        return x, res.value

    def autotune(self):
        """gpu_tomo_server_autotune: the fastest launch shape for the
        admitted loops as one batch (call after drain()); returns it in
        GpuTomo.tune_set's order."""
        ch = (C.c_int * 6)()
        _abi.check(_abi.lib().gpu_tomo_server_autotune(self._h, ch), "gpu_tomo_server_autotune")
        return tuple(ch)

    def update_cn2(self, lid, recon):
        """gpu_tomo_server_update_cn2: loop lid takes recon's new regulariser
        and FDPCG blocks (after recon.update_cn2); other loops keep running."""
        _abi.check(_abi.lib().gpu_tomo_server_update_cn2(self._h, int(lid), C.byref(recon.rec)),
                   "gpu_tomo_server_update_cn2")

    def usage(self):
        u, b = C.c_double(), C.c_double()
        _abi.lib().gpu_tomo_server_usage(self._h, C.byref(u), C.byref(b))
        return u.value, b.value

    def close(self):
        if getattr(self, "_h", None):
            _abi.lib().gpu_tomo_server_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FitSetup:
    """DM fitting geometry and operators (mt_fit_create) for a Recon."""

    def __init__(self, recon, fc):
        self.recon, self.fc = recon, fc
        self._parms = FitParms(fc)
        L = _abi.lib()
        h = C.c_void_p()
        _abi.check(L.mt_fit_create(recon._h, C.byref(self._parms.p), C.byref(h)), "mt_fit_create")
        self._h = h
        self.rec = _abi.MtFitRecon()
        _abi.check(L.mt_fit_recon(h, C.byref(self.rec)), "mt_fit_recon")
        na, nf = C.c_long(), C.c_long()
        L.mt_fit_sizes(h, C.byref(na), C.byref(nf))
        self.nact, self.nfloc = na.value, nf.value

    def close(self):
        if getattr(self, "_h", None):
            _abi.lib().mt_fit_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def W0(self):
        """W0 as a scipy.sparse.csr_matrix (nfloc x nfloc) from the 9-point rows."""
        import scipy.sparse as sp
        r = self.rec
        n, nx, ny = self.nfloc, r.fnx, r.fny
        w = _arr(r.W0, 9 * n).reshape(n, 9)
        rows, cols, vals = [], [], []
        for k in range(9):
            dx, dy = k % 3 - 1, k // 3 - 1
            idx = np.arange(n)
            ix, iy = idx % nx, idx // nx
            jx, jy = ix + dx, iy + dy
            ok = (w[:, k] != 0) & (jx >= 0) & (jx < nx) & (jy >= 0) & (jy < ny)
            rows.append(idx[ok])
            cols.append((jx + jy * nx)[ok])
            vals.append(w[ok, k])
        return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))

    def W1(self):
        return _arr(self.rec.W1, self.nfloc)

    def NW(self):
        return _arr(self.rec.NW, self.nact * self.rec.nnw).reshape(self.rec.nnw, self.nact).T

    def slaving(self):
        """S as scipy.sparse.csr_matrix, or None."""
        import scipy.sparse as sp
        r = self.rec
        if not r.sl_p:
            return None
        p = _arr(r.sl_p, self.nact + 1, np.int64)
        return sp.csr_matrix((_arr(r.sl_x, p[-1]), _arr(r.sl_i, p[-1], np.int64), p), shape=(self.nact, self.nact))

    def actcpl(self):
        return _arr(self.rec.actcpl, self.nact)

    def grids(self):
        r = self.rec
        return [dict(nx=r.anx[d], ny=r.any[d], ox=r.aox[d], oy=r.aoy[d], dx=r.adx[d], ht=r.aht[d])
                for d in range(r.ndm)]


class GpuFit:
    """GPU DM fitting plan for nbatch systems (gpu_fit_*)."""

    def __init__(self, recon, fit, device=0, nbatch=1):
        self.recon, self.fit = recon, fit
        h = C.c_void_p()
        _abi.check(_abi.lib().gpu_fit_setup(C.byref(recon.rec), C.byref(fit.rec), device, nbatch, C.byref(h)),
                   "gpu_fit_setup")
        self._h = h
        self.nbatch = nbatch

    def close(self):
        if getattr(self, "_h", None):
            _abi.lib().gpu_fit_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, opdr, isys=0):
        """curecon_t::fit: (dmfit, cgres, status) from tomography layers."""
        x = np.ascontiguousarray(opdr, dtype=np.float64)
        out = np.zeros(self.fit.nact)
        res = C.c_double()
        rc = _abi.check(_abi.lib().gpu_fit_solve(self._h, int(isys), _dp(x), _dp(out), C.byref(res)),
                        "gpu_fit_solve")
        return out, res.value, rc

    def solve_dev(self, d_opdr, d_dm, nsys=None, stream=None):
        n = self.nbatch if nsys is None else nsys
        _abi.check(_abi.lib().gpu_fit_solve_dev(self._h, C.c_void_p(d_opdr), C.c_void_p(d_dm), int(n),
                                                C.c_void_p(stream or 0)), "gpu_fit_solve_dev")

    def R(self, opdr, alpha=1.0):
        x = np.ascontiguousarray(opdr, dtype=np.float64)
        out = np.zeros(self.fit.nact)
        _abi.check(_abi.lib().gpu_fit_R(self._h, _dp(x), _dp(out), float(alpha)), "gpu_fit_R")
        return out

    def L(self, a, alpha=1.0, beta=0.0, out=None):
        x = np.ascontiguousarray(a, dtype=np.float64)
        o = np.zeros(self.fit.nact) if out is None else np.array(out, dtype=np.float64)
        _abi.check(_abi.lib().gpu_fit_L(self._h, _dp(x), _dp(o), float(alpha), float(beta)), "gpu_fit_L")
        return o

    def residuals(self, nsys=None):
        n = self.nbatch if nsys is None else nsys
        out = np.zeros(n)
        _abi.check(_abi.lib().gpu_fit_residuals(self._h, _dp(out), n), "gpu_fit_residuals")
        return out
