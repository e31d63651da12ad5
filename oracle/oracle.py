"""ctypes loader of the FP64 CPU oracle (oracle/tomo_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libtomo_oracle.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        vp, dp = C.c_void_p, C.POINTER(C.c_double)
        L.or_create.argtypes = [vp, C.c_int]
        L.or_create.restype = vp
        L.or_free.argtypes = [vp]
        L.or_free.restype = None
        L.or_nx.argtypes = [vp]
        L.or_nx.restype = C.c_long
        L.or_ng.argtypes = [vp]
        L.or_ng.restype = C.c_long
        for f in ("or_tomo_r", "or_tomo_l"):
            getattr(L, f).argtypes = [vp, dp, dp, C.c_double]
            getattr(L, f).restype = None
        L.or_fdpcg.argtypes = [vp, dp, dp]
        L.or_fdpcg.restype = None
        L.or_pcg.argtypes = [vp, dp, dp, C.c_int, C.c_int, C.c_int, dp]
        L.or_pcg.restype = C.c_double
        L.or_solve.argtypes = [vp, dp, dp, C.c_int, C.c_int]
        L.or_solve.restype = C.c_double
        L.or_hx.argtypes = [vp, C.c_int, dp, dp]
        L.or_hx.restype = None
        L.or_hxt.argtypes = [vp, C.c_int, dp, dp]
        L.or_hxt.restype = None
        L.or_l2.argtypes = [vp, C.c_int, dp, dp, C.c_double]
        L.or_l2.restype = None
        L.or_fft1d.argtypes = [dp, dp, C.c_int, C.c_int]
        L.or_fft1d.restype = None
        L.orf_create.argtypes = [vp, vp, C.c_int]
        L.orf_create.restype = vp
        L.orf_free.argtypes = [vp]
        L.orf_free.restype = None
        L.orf_nact.argtypes = [vp]
        L.orf_nact.restype = C.c_long
        for f in ("orf_fit_r", "orf_fit_l"):
            getattr(L, f).argtypes = [vp, dp, dp, C.c_double]
            getattr(L, f).restype = None
        L.orf_pcg.argtypes = [vp, dp, dp, C.c_int, C.c_int]
        L.orf_pcg.restype = C.c_double
        L.orf_solve.argtypes = [vp, dp, dp, C.c_int, C.c_int]
        L.orf_solve.restype = C.c_double
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    """FP64 CPU tomography solve on the operators of a maos_amd.tomo.Recon."""

    def __init__(self, recon, nthread=1):
        self.recon = recon
        self._h = lib().or_create(C.cast(C.pointer(recon.rec), C.c_void_p), int(nthread))
        self.nx = lib().or_nx(self._h)
        self.ng = lib().or_ng(self._h)

    def close(self):
        if getattr(self, "_h", None):
            lib().or_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def tomo_r(self, g, alpha=1.0):
        gg = np.ascontiguousarray(g, dtype=np.float64)
        out = np.zeros(self.nx)
        lib().or_tomo_r(self._h, _dp(gg), _dp(out), alpha)
        return out

    def tomo_l(self, x, alpha=1.0, xout=None):
        xi = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros(self.nx) if xout is None else np.array(xout, dtype=np.float64)
        lib().or_tomo_l(self._h, _dp(xi), _dp(out), alpha)
        return out

    def fdpcg(self, x):
        xi = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros(self.nx)
        lib().or_fdpcg(self._h, _dp(xi), _dp(out))
        return out

    def pcg(self, b, maxit, x0=None, precond=True, history=False):
        bb = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.nx) if x0 is None else np.array(x0, dtype=np.float64)
        hist = np.zeros(maxit + 1) if history else None
        res = lib().or_pcg(self._h, _dp(bb), _dp(x), 0 if x0 is None else 1, int(maxit),
                           1 if precond else 0, _dp(hist) if history else None)
        return (x, res, hist) if history else (x, res)

    def solve(self, g, maxit, x0=None):
        gg = np.ascontiguousarray(g, dtype=np.float64)
        x = np.zeros(self.nx) if x0 is None else np.array(x0, dtype=np.float64)
        res = lib().or_solve(self._h, _dp(gg), _dp(x), 0 if x0 is None else 1, int(maxit))
        return x, res

    def l2(self, l, xl, alpha=1.0):
        xi = np.ascontiguousarray(xl, dtype=np.float64)
        out = np.zeros_like(xi)
        lib().or_l2(self._h, int(l), _dp(xi), _dp(out), alpha)
        return out

    def hx(self, w, x):
        r = self.recon.rec
        opd = np.zeros(r.pnx * r.pny)
        xi = np.ascontiguousarray(x, dtype=np.float64)
        lib().or_hx(self._h, int(w), _dp(xi), _dp(opd))
        return opd

    def hxt(self, w, opd):
        out = np.zeros(self.nx)
        oi = np.ascontiguousarray(opd, dtype=np.float64)
        lib().or_hxt(self._h, int(w), _dp(oi), _dp(out))
        return out


class FitOracle:
    """FP64 CPU DM fitting (oracle/fit_oracle.c) on a Recon and its FitSetup."""

    def __init__(self, recon, fit, nthread=1):
        self.recon, self.fit = recon, fit
        self._h = lib().orf_create(C.cast(C.pointer(recon.rec), C.c_void_p), C.cast(C.pointer(fit.rec), C.c_void_p),
                                   int(nthread))
        self.nact = lib().orf_nact(self._h)

    def close(self):
        if getattr(self, "_h", None):
            lib().orf_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def fit_r(self, x, alpha=1.0):
        xi = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros(self.nact)
        lib().orf_fit_r(self._h, _dp(xi), _dp(out), alpha)
        return out

    def fit_l(self, a, alpha=1.0):
        ai = np.ascontiguousarray(a, dtype=np.float64)
        out = np.zeros(self.nact)
        lib().orf_fit_l(self._h, _dp(ai), _dp(out), alpha)
        return out

    def solve(self, x, maxit, a0=None):
        xi = np.ascontiguousarray(x, dtype=np.float64)
        a = np.zeros(self.nact) if a0 is None else np.array(a0, dtype=np.float64)
        res = lib().orf_solve(self._h, _dp(xi), _dp(a), 0 if a0 is None else 1, int(maxit))
        return a, res


def fft1d(x, sign=-1):
    a = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.zeros_like(a)
    lib().or_fft1d(_dp(a.view(np.float64)), _dp(out.view(np.float64)), a.size, sign)
    return out
