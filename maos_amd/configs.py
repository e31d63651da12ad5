"""Tomography configurations (MAOS .conf keys) of the benchmark workloads.

Each config is a plain dict of MAOS parameter values; `to_parms` turns it into
the mt_parms_t the C ABI reads.  Angles are given in arcsec as in the .conf
files and converted to radian here (as setup_parms does).

  small8    8 m, 4 NGS at (+-15",+-15"), 2 layers (0, 8 km), tomo.pos=2, os=2
  large30   30 m, 6 LGS at 90 km (on-axis + 35" pentagon, config/maos/
            shwfs_lgs.conf:51-52), 6 layers 0-16 km, tomo.pos=1, os=1
  lgs8_flat small8 with 4 LGS at 90 km and tomo.cone=0: rays above ground
            sample the layers off-grid (generic ray path)
  batch     D in {8,10,20,30}, 1-6 LGS, 2-6 layers (large30 truncated)
"""
import ctypes as C
import math

from . import _abi

AS2RAD = 1.0 / 206265.0
INF = float("inf")

LGS_THETAX = [0, 0, 33.287, 20.5725, -20.5725, -33.287]
LGS_THETAY = [0, 35, 10.8156, -28.3156, -28.3156, 10.8156]
LARGE_HT = [0.0, 1000.0, 2000.0, 4000.0, 8000.0, 16000.0]
LARGE_WT = [0.45, 0.15, 0.1, 0.1, 0.1, 0.1]


def small8(**kw):
    """Plain small tomography: D=8 m, din=1.1 m, 4 NGS SH 16x16 of 0.5 m."""
    c = dict(
        name="small8", aper_d=8.0, aper_din=1.1,
        ht=[0.0, 8000.0], wt=[0.6, 0.4], os=[2, 2], atmr_dx=0.5,
        r0z=0.186, L0=30.0, za=0.0, tomo_pos=2, tomo_guard=1, tomo_cone=1,
        tomo_cxxscale=1.0, tomo_precond=1, recon_split=1, tomo_splitlrt=0,
        powfs=[dict(dsa=0.5, dx=1 / 64, saat=0.3, nearecon=50.0, hs=INF, trs=0, lo=0,
                    skip=0, dtrat=1)],
        wfs=[dict(thetax=tx, thetay=ty, powfs=0) for tx, ty in
             [(15, 15), (-15, 15), (-15, -15), (15, -15)]],
        dirs=[(0.0, 0.0, INF)], tomoscale=1e-12, maxit=30)
    c.update(kw)
    return c


def large30(**kw):
    """Large single solve: D=30 m, 6 LGS at 90 km, 6 layers, 60x60 subaps."""
    c = dict(
        name="large30", aper_d=30.0, aper_din=3.6,
        ht=list(LARGE_HT), wt=list(LARGE_WT), os=[1] * 6, atmr_dx=0.5,
        r0z=0.186, L0=30.0, za=0.0, tomo_pos=1, tomo_guard=1, tomo_cone=1,
        tomo_cxxscale=1.0, tomo_precond=1, recon_split=1, tomo_splitlrt=0,
        powfs=[dict(dsa=0.5, dx=1 / 64, saat=0.3, nearecon=50.0, hs=90e3, trs=1, lo=0,
                    skip=0, dtrat=1)],
        wfs=[dict(thetax=tx, thetay=ty, powfs=0) for tx, ty in zip(LGS_THETAX, LGS_THETAY)],
        dirs=[(0.0, 0.0, INF)], tomoscale=1e-12, maxit=30)
    c.update(kw)
    return c


def lgs8_flat(**kw):
    """4 LGS at 90 km on the 8 m aperture without cone coordinates
    (tomo.cone=0): layer grids keep atmr.dx, so the LGS rays of every layer
    above ground sample them at a ratio 1 - h/hs (generic ray path)."""
    c = small8(name="lgs8_flat", tomo_cone=0,
               powfs=[dict(dsa=0.5, dx=1 / 64, saat=0.3, nearecon=30.0, hs=90e3, trs=1, lo=0,
                           skip=0, dtrat=1)])
    c.update(kw)
    return c


def batch_member(D, nwfs, nlayer, **kw):
    """One system of the uneven batch: large30 truncated to nwfs LGS and
    nlayer layers on a D-m aperture (weights renormalised by the setup)."""
    din = {8.0: 1.1, 10.0: 1.4, 20.0: 2.4, 30.0: 3.6}.get(float(D), 0.12 * D)
    return large30(name="batch_D%g_w%d_l%d" % (D, nwfs, nlayer), aper_d=float(D),
                   aper_din=din, ht=LARGE_HT[:nlayer], wt=LARGE_WT[:nlayer],
                   os=[1] * nlayer,
                   wfs=[dict(thetax=LGS_THETAX[i], thetay=LGS_THETAY[i], powfs=0)
                        for i in range(nwfs)], **kw)


def batch64(seed=3):
    """64 uneven systems: D in {8,10,20,30}, 1-6 LGS, 2-6 layers (seeded)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(64):
        D = float(rng.choice([8, 10, 20, 30]))
        out.append(batch_member(D, int(rng.integers(1, 7)), int(rng.integers(2, 7))))
    return out


def fit_for(cfg, **kw):
    """DM fitting parameters for a tomography config (MAOS fit.* / dm.* keys,
    config/maos/recon.conf:56-72, sim.conf:181-187): on-axis fitting on
    8-10 m apertures with one ground DM (dm_single.conf), and on larger ones
    the 9-direction 30" ring of fit_cir30.conf with two DMs at 0 and 11.8 km
    (dm_dual.conf, offset 0 and 1/2); dm.dx = 0.5 m, dm.guard = 1, linear
    influence functions (dm.iac = 0), fit.actinterp = 0, cold start
    (fit.maxit 0 -> 100 iterations)."""
    if cfg["aper_d"] <= 10:
        dms = [dict(ht=0.0, dx=0.5, offset=0.0, guard=1.0, iac=0.0)]
        th = [(0.0, 0.0, 1.0)]
    else:
        dms = [dict(ht=0.0, dx=0.5, offset=0.0, guard=1.0, iac=0.0),
               dict(ht=11800.0, dx=0.5, offset=0.5, guard=1.0, iac=0.0)]
        r = [0, 0.5, 0.3536, 0, -0.3536, -0.5, -0.3536, 0, 0.3536]
        q = [0, 0, 0.3536, 0.5, 0.3536, 0, -0.3536, -0.5, -0.3536]
        w = [1 / 2, 1 / 12, 1 / 24, 1 / 12, 1 / 24, 1 / 12, 1 / 24, 1 / 12, 1 / 24]
        th = [(a * 30, b * 30, c) for a, b, c in zip(r, q, w)]
    c = dict(dm=dms, fit=th, fit_hs=[INF] * len(th), fit_pos=0, fit_maxit=0, warm=0, lrt_piston=1,
             lrt_tt=1, actslave=1, actthres=0.1, tikcr=1e-7, actinterp=0, annular_W=0)
    c.update(kw)
    return c


def with_fit_dirs(cfg, fc):
    """The tomography config with the fitting directions among its beam
    directions, as MAOS lists them in parms->dirs (maos/setup_parms.c:
    1945-1984), so that the layer grids cover every fitting direction."""
    dirs = list(cfg.get("dirs", []))
    for (tx, ty, _), hs in zip(fc["fit"], fc["fit_hs"]):
        if (tx, ty, hs) not in dirs:
            dirs.append((tx, ty, hs))
    c = dict(cfg)
    c["dirs"] = dirs
    c["name"] = cfg["name"] + "_fit"
    return c


class FitParms:
    """Owns the ctypes arrays behind one mt_fit_parms_t."""

    def __init__(self, fc):
        self.fc = fc
        dms = fc["dm"]
        self.dm = (_abi.MtDm * len(dms))(*[_abi.MtDm(d["ht"], d["dx"], d["offset"], d["guard"], d["iac"])
                                             for d in dms])
        n = len(fc["fit"])
        self.tx = (C.c_double * n)(*[f[0] * AS2RAD for f in fc["fit"]])
        self.ty = (C.c_double * n)(*[f[1] * AS2RAD for f in fc["fit"]])
        self.wt = (C.c_double * n)(*[f[2] for f in fc["fit"]])
        self.hs = (C.c_double * n)(*fc["fit_hs"])
        self.p = _abi.MtFitParms(len(dms), self.dm, n, self.tx, self.ty, self.wt, self.hs, fc["fit_pos"],
                                 fc["fit_maxit"], fc["warm"], fc["lrt_piston"], fc["lrt_tt"], fc["actslave"],
                                 fc["actthres"], fc["tikcr"], fc["actinterp"], fc["annular_W"])


CONFIGS = {"small8": small8, "large30": large30, "lgs8_flat": lgs8_flat}


class Parms:
    """Owns the ctypes arrays behind one mt_parms_t."""

    def __init__(self, cfg):
        self.cfg = cfg
        n = len(cfg["ht"])
        self.ht = (C.c_double * n)(*cfg["ht"])
        self.wt = (C.c_double * n)(*cfg["wt"])
        self.os = (C.c_int * n)(*cfg["os"])
        pw = cfg["powfs"]
        self.powfs = (_abi.MtPowfs * len(pw))(*[
            _abi.MtPowfs(p["dsa"], p["dx"], p["saat"], p["nearecon"], p["hs"], p["trs"],
                         p["lo"], p["skip"], p["dtrat"]) for p in pw])
        wf = cfg["wfs"]
        self.wfs = (_abi.MtWfs * len(wf))(*[
            _abi.MtWfs(w["thetax"] * AS2RAD, w["thetay"] * AS2RAD, w["powfs"]) for w in wf])
        dirs = cfg.get("dirs", [])
        nd = len(dirs)
        self.dtx = (C.c_double * max(nd, 1))(*[d[0] * AS2RAD for d in dirs])
        self.dty = (C.c_double * max(nd, 1))(*[d[1] * AS2RAD for d in dirs])
        self.dhs = (C.c_double * max(nd, 1))(*[d[2] for d in dirs])
        self.p = _abi.MtParms(
            cfg["aper_d"], cfg["aper_din"], n, self.ht, self.wt, self.os, cfg["atmr_dx"],
            cfg["r0z"], cfg["L0"], cfg["za"], cfg["tomo_pos"], cfg["tomo_guard"],
            cfg["tomo_cone"], cfg["tomo_cxxscale"], cfg["tomo_precond"], cfg["recon_split"],
            cfg["tomo_splitlrt"], len(pw), self.powfs, len(wf), self.wfs, nd,
            self.dtx, self.dty, self.dhs, cfg["tomoscale"])


def wfs_hs(cfg):
    return [cfg["powfs"][w["powfs"]]["hs"] for w in cfg["wfs"]]


def wfs_theta(cfg):
    return ([w["thetax"] * AS2RAD for w in cfg["wfs"]],
            [w["thetay"] * AS2RAD for w in cfg["wfs"]])


def nea_rad(cfg, iw):
    p = cfg["powfs"][cfg["wfs"][iw]["powfs"]]
    return p["nearecon"] / 206265000.0 / math.sqrt(p["dtrat"])
