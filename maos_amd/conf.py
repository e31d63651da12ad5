"""MAOS .conf configurations for this path.

`Conf` is the key store of the library's mt_conf_* (MAOS's open_config and
readcfg parsers, lib/readcfg.c, sys/readstr.c).  `tomo_config` and
`fit_config` read the keys the tomography and DM fitting path reads, with
the defaults and derived values setup_parms applies to them
(maos/setup_parms.c: readcfg_powfs :241-444, readcfg_wfs :445-520,
readcfg_atm / readcfg_atmr :655-711, readcfg_aper :716-740, the DM keys
:560-600, the fit keys :824-856, tomo.maxit / fit.maxit defaults
:2359-2385, recon.warm_restart :2269), into the config dicts of
maos_amd.configs (angles in arcsec, as in the .conf files).

    c = Conf.read("my_run.conf", paths=["/path/to/maos/config"])
    r = tomo.Recon(tomo_config(c))
"""
import ctypes as C
import math
import os

import numpy as np

from . import _abi


class Conf:
    """A MAOS key store (mt_conf_t)."""

    def __init__(self):
        h = C.c_void_p()
        _abi.check(_abi.lib().mt_conf_new(C.byref(h)), "mt_conf_new")
        self._h = h

    @classmethod
    def read(cls, *inputs, paths=(), overrides=None):
        """inputs: .conf file names or key=value text, read in order at
        priority 0; overrides (command-line style text) at priority 1;
        paths: directories searched for files named without a path (the
        directory of each file input is searched too)."""
        c = cls()
        for p in paths:
            c.add_path(p)
        for inp in inputs:
            if inp.endswith(".conf") and os.path.dirname(inp):
                c.add_path(os.path.dirname(os.path.abspath(inp)))
            c.load(inp)
        if overrides:
            c.load(overrides, priority=1)
        return c

    def close(self):
        if getattr(self, "_h", None):
            _abi.lib().mt_conf_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_path(self, d):
        _abi.check(_abi.lib().mt_conf_add_path(self._h, d.encode()), "mt_conf_add_path")

    def load(self, file_or_text, prefix=None, priority=0):
        _abi.check(_abi.lib().mt_conf_read(self._h, file_or_text.encode(), prefix.encode() if prefix else None,
                                           int(priority)), "mt_conf_read")

    def has(self, key):
        return bool(_abi.lib().mt_conf_has(self._h, key.encode()))

    def str(self, key):
        n = C.c_long()
        _abi.check(_abi.lib().mt_conf_str(self._h, key.encode(), None, 0, C.byref(n)), "mt_conf_str")
        buf = C.create_string_buffer(n.value + 1)
        _abi.check(_abi.lib().mt_conf_str(self._h, key.encode(), buf, n.value + 1, None), "mt_conf_str")
        return buf.value.decode()

    def strarr(self, key):
        L = _abi.lib()
        n = C.c_long()
        _abi.check(L.mt_conf_strarr(self._h, key.encode(), -1, None, 0, C.byref(n)), "mt_conf_strarr")
        out = []
        for i in range(n.value):
            m = C.c_long()
            _abi.check(L.mt_conf_strarr(self._h, key.encode(), i, None, 0, C.byref(m)), "mt_conf_strarr")
            buf = C.create_string_buffer(m.value + 1)
            _abi.check(L.mt_conf_strarr(self._h, key.encode(), i, buf, m.value + 1, None), "mt_conf_strarr")
            out.append(buf.value.decode())
        return out

    def arr(self, key):
        """The value as a numeric array of shape (nrow * ncol,) column-major
        (a MAOS dmat); shape in .shape2 of the returned array is not kept:
        use arr2 for 2-d values."""
        return self.arr2(key).ravel(order="F")

    def arr2(self, key):
        L = _abi.lib()
        n, nr, nc = C.c_long(), C.c_long(), C.c_long()
        _abi.check(L.mt_conf_dbl(self._h, key.encode(), None, 0, C.byref(n), C.byref(nr), C.byref(nc)), "mt_conf_dbl")
        out = np.zeros(max(n.value, 1))
        _abi.check(L.mt_conf_dbl(self._h, key.encode(), out.ctypes.data_as(_abi.dp), len(out), C.byref(n),
                                 C.byref(nr), C.byref(nc)), "mt_conf_dbl")
        return out[:n.value].reshape((nr.value, nc.value), order="F") if n.value else np.zeros((0, 0))

    def num(self, key):
        a = self.arr(key)
        if a.size != 1:
            raise ValueError("key %s: one number expected, got %d" % (key, a.size))
        return float(a[0])

    def int(self, key):
        return int(round(self.num(key)))

    def arr_n(self, key, n, relax=False):
        """readcfg_*arr_n (exactly n values) or, relax, _nmax (one value
        repeated n times, or exactly n)."""
        a = self.arr(key)
        if relax and a.size <= 1:
            return np.full(n, a[0] if a.size else 0.0)
        if a.size != n:
            raise ValueError("key %s: need %d values, got %d" % (key, n, a.size))
        return a

    def unused(self):
        n = C.c_long()
        L = _abi.lib()
        _abi.check(L.mt_conf_unused(self._h, None, 0, C.byref(n)), "mt_conf_unused")
        buf = C.create_string_buffer(1 << 16)
        _abi.check(L.mt_conf_unused(self._h, buf, len(buf), None), "mt_conf_unused")
        return buf.value.decode().split()


def _opt(c, key, default):
    return c.num(key) if c.has(key) else default


def tomo_config(c, name="conf"):
    """The tomography config dict (maos_amd.configs format) of a MAOS
    configuration."""
    d = c.arr("aper.d")
    aper_d, aper_din = float(d[0]), float(d[1]) if d.size > 1 else 0.0
    atm_ht = c.arr("atm.ht")
    atm_wt = c.arr_n("atm.wt", atm_ht.size)
    r0z = _opt(c, "atmr.r0z", -1.0)
    if r0z <= 0:
        r0z = c.num("atm.r0z")
    L0 = _opt(c, "atmr.L0", -1.0)
    if L0 <= 0:
        L0 = c.num("atm.L0")
    ht = c.arr("atmr.ht") if c.has("atmr.ht") else np.zeros(0)
    wt = c.arr("atmr.wt") if c.has("atmr.wt") else np.zeros(0)
    if ht.size == 0:
        ht = atm_ht
    if wt.size == 0:
        wt = atm_wt
    elif wt.size != ht.size:
        raise ValueError("atmr.wt has to match atmr.ht")
    os_ = c.arr_n("atmr.os", ht.size, relax=True)
    # powfs
    npowfs = c.arr("powfs.dsa").size
    P = {k: c.arr_n("powfs." + k, npowfs) for k in ("dsa", "hs", "nearecon", "trs", "lo", "nwfs")}
    for k, dflt in (("dx", 1 / 64), ("saat", 0.3), ("skip", 0), ("dtrat", 1)):
        P[k] = c.arr_n("powfs." + k, npowfs, relax=True) if c.has("powfs." + k) else np.full(npowfs, dflt)
    tx, ty = c.arr("wfs.thetax"), c.arr_n("wfs.thetay", c.arr("wfs.thetax").size)
    powfs, wfs = [], []
    iw = 0
    for ip in range(npowfs):
        nw = int(P["nwfs"][ip])
        if nw == 0:  # a powfs without stars is dropped
            continue
        k = len(powfs)
        powfs.append(dict(dsa=float(P["dsa"][ip]), dx=float(P["dx"][ip]), saat=float(P["saat"][ip]),
                          nearecon=float(P["nearecon"][ip]), hs=float(P["hs"][ip]), trs=int(P["trs"][ip]),
                          lo=int(P["lo"][ip]), skip=int(P["skip"][ip]), dtrat=int(P["dtrat"][ip])))
        for _ in range(nw):
            wfs.append(dict(thetax=float(tx[iw]), thetay=float(ty[iw]), powfs=k))
            iw += 1
    if iw != tx.size:
        raise ValueError("wfs.thetax has %d entries, powfs.nwfs sums to %d" % (tx.size, iw))
    dirs = []
    if c.has("evl.thetax"):
        ex = c.arr("evl.thetax")
        ey = c.arr_n("evl.thetay", ex.size)
        eh = c.arr_n("evl.hs", ex.size, relax=True) if c.has("evl.hs") else np.full(ex.size, math.inf)
        dirs += [(float(a), float(b), float(h)) for a, b, h in zip(ex, ey, eh)]
    if c.has("fit.thetax"):
        fx = c.arr("fit.thetax")
        fy = c.arr_n("fit.thetay", fx.size)
        fh = c.arr_n("fit.hs", fx.size, relax=True) if c.has("fit.hs") else np.full(fx.size, math.inf)
        dirs += [(float(a), float(b), float(h)) for a, b, h in zip(fx, fy, fh)]
    precond = int(_opt(c, "tomo.precond", 1))
    warm = warm_restart(c)
    maxit = int(_opt(c, "tomo.maxit", 0))
    if maxit == 0:
        factor = 1 if warm else 10
        if not precond:
            factor *= 10
        maxit = 4 * factor
    return dict(name=name, aper_d=aper_d, aper_din=aper_din, ht=[float(v) for v in ht], wt=[float(v) for v in wt],
                os=[int(v) for v in os_], atmr_dx=float(_opt(c, "atmr.dx", 0.0)), r0z=float(r0z), L0=float(L0),
                za=math.radians(_opt(c, "sim.zadeg", 0.0)), tomo_pos=int(_opt(c, "tomo.pos", 2)),
                tomo_guard=int(_opt(c, "tomo.guard", 1)), tomo_cone=int(_opt(c, "tomo.cone", 1)),
                tomo_cxxscale=float(_opt(c, "tomo.cxxscale", 1.0)), tomo_precond=precond,
                recon_split=int(_opt(c, "recon.split", 1)), tomo_splitlrt=int(_opt(c, "tomo.splitlrt", 0)),
                powfs=powfs, wfs=wfs, dirs=dirs, tomoscale=1e-12, maxit=maxit)


def warm_restart(c):
    """recon.warm_restart = !dbg.nocgwarm && atm.frozenflow
    (maos/setup_parms.c:2269, without the dbg.tomo_maxit case)."""
    return bool(int(_opt(c, "atm.frozenflow", 0))) and not bool(int(_opt(c, "dbg.nocgwarm", 0)))


def fit_config(c):
    """The DM fitting dict (maos_amd.configs.fit_for format)."""
    ht = c.arr("dm.ht")
    ndm = ht.size
    dx = c.arr_n("dm.dx", ndm, relax=True)
    off = c.arr_n("dm.offset", ndm, relax=True) if c.has("dm.offset") else np.zeros(ndm)
    guard = c.arr_n("dm.guard", ndm, relax=True) if c.has("dm.guard") else np.ones(ndm)
    iac = c.arr_n("dm.iac", ndm, relax=True) if c.has("dm.iac") else np.zeros(ndm)
    dms = [dict(ht=float(ht[i]), dx=float(dx[i]), offset=float(off[i]), guard=float(guard[i]), iac=float(iac[i]))
           for i in range(ndm)]
    fx = c.arr("fit.thetax")
    fy = c.arr_n("fit.thetay", fx.size)
    fw = c.arr_n("fit.wt", fx.size, relax=True)
    fh = c.arr_n("fit.hs", fx.size, relax=True) if c.has("fit.hs") else np.full(fx.size, math.inf)
    return dict(dm=dms, fit=[(float(a), float(b), float(w)) for a, b, w in zip(fx, fy, fw)],
                fit_hs=[float(h) for h in fh], fit_pos=int(_opt(c, "fit.pos", -1)),
                fit_maxit=int(_opt(c, "fit.maxit", 0)), warm=int(warm_restart(c)),
                lrt_piston=int(_opt(c, "fit.lrt_piston", 1)), lrt_tt=int(_opt(c, "fit.lrt_tt", 1)),
                actslave=int(_opt(c, "fit.actslave", 1)), actthres=float(_opt(c, "fit.actthres", 0.1)),
                tikcr=float(_opt(c, "fit.tikcr", 1e-7)), actinterp=int(_opt(c, "fit.actinterp", 0)),
                annular_W=int(_opt(c, "dbg.annular_W", 0)))
