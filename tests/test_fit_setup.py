"""CPU: the DM fitting setup (mt_fit_create) and the fitting oracle.

Pins, each chosen so that a plausible slip fails it:
  W0, W1   against the quantity they are defined by (lib/mkw.c:357-362):
           W0(i,j) = integral over the aperture of the bilinear hat
           functions of grid points i and j, W1 = W0 1, both / sum W1.
           Brute-force quadrature of that integral at a resolution other
           than the reference's, for every entry (interior closed forms
           4/9, 1/9, 1/36 and the rim cells); and the identity it exists
           for: a'(W0 - W1 W1')a equals the piston-removed variance of the
           bilinear interpolant of a over the aperture.
  HA, W    FitL of a DM piston is NW NW' 1 + tikdiag 1 (HA reproduces a
           constant, W removes piston, the slaving regulariser has the
           piston in its null space); FitL is symmetric positive definite.
  S        symmetric, positive semi-definite, zero on pistons of each DM,
           nonzero exactly on slave actuators (coupling below fit.actthres)
           and their neighbours.
  NW       piston columns constant on coupled actuators of their DM and zero
           elsewhere; tip/tilt columns linear in the actuator coordinates of
           the upper DMs.
"""
import numpy as np
import pytest

from maos_amd import configs, tomo
from oracle.oracle import FitOracle


@pytest.fixture(scope="module")
def fit8():
    base = configs.small8()
    fc = configs.fit_for(base)
    cfg = configs.with_fit_dirs(base, fc)
    r = tomo.Recon(cfg)
    return cfg, fc, r, tomo.FitSetup(r, fc)


@pytest.fixture(scope="module")
def fit30():
    base = configs.large30()
    fc = configs.fit_for(base)
    cfg = configs.with_fit_dirs(base, fc)
    r = tomo.Recon(cfg)
    return cfg, fc, r, tomo.FitSetup(r, fc)


def brute_w(nx, ny, ox, oy, dx, R, m=120):
    """W0, W1 by midpoint quadrature with m x m points per grid cell of the
    overlap integrals of bilinear hats inside the circle of radius R, then
    normalised by sum W1 (the reference's normalisation)."""
    n = nx * ny
    W0 = np.zeros((n, n))
    u = (np.arange(m) + 0.5) / m
    U, V = np.meshgrid(u, u)
    U, V = U.ravel(), V.ravel()
    hat = np.stack([(1 - U) * (1 - V), U * (1 - V), (1 - U) * V, U * V])  # corners (0,0),(1,0),(0,1),(1,1)
    for cy in range(ny - 1):
        for cx in range(nx - 1):
            x = ox + (cx + U) * dx
            y = oy + (cy + V) * dx
            inside = (x * x + y * y) < R * R
            if not inside.any():
                continue
            h = hat[:, inside]
            idx = [cx + cy * nx, cx + 1 + cy * nx, cx + (cy + 1) * nx, cx + 1 + (cy + 1) * nx]
            blk = h @ h.T / (m * m)
            W0[np.ix_(idx, idx)] += blk
    W1 = W0.sum(1)
    s = W1.sum()
    return W0 / s, W1 / s


def test_w0_w1_against_quadrature():
    """A small square grid around an 8 m aperture (all four rim cases)."""
    cfg = configs.small8(name="w8", os=[1, 1], tomo_pos=1)
    fc = configs.fit_for(cfg)
    r = tomo.Recon(configs.with_fit_dirs(cfg, fc))
    f = tomo.FitSetup(r, fc)
    rec = f.rec
    W0 = f.W0().toarray()
    W1 = f.W1()
    Wb0, Wb1 = brute_w(rec.fnx, rec.fny, rec.fox, rec.foy, rec.fdx, cfg["aper_d"] / 2)
    # row sums and normalisation
    assert abs(W1.sum() - 1) < 1e-12
    assert np.allclose(W0.sum(1), W1, rtol=0, atol=1e-15)
    # every entry: the reference's 50 x 50 midpoint rule on rim cells
    # against 120 x 120 here differs by the rim's O(1/50) cut error
    scale = W0.max()
    err = np.abs(W0 - Wb0)
    assert err.max() < 5e-3 * scale, err.max() / scale
    # interior entries exactly the closed forms (times the common scale)
    inner = np.isclose(W0, W0.max(), rtol=0, atol=0)
    assert inner.sum() > 10
    i = np.argmax(W0.diagonal())
    row = W0[i]
    nz = np.sort(row[row > 0])[::-1]
    assert np.allclose(nz / nz[0], [1, .25, .25, .25, .25, 1 / 16, 1 / 16, 1 / 16, 1 / 16], rtol=1e-12)
    # a wrong integrand (e.g. the corner weight of the edge formula) is an
    # O(1) relative change on the rim entries: the relative error over the
    # rim stays small
    rim = (W0 > 0) & ~np.isclose(W0 / scale, np.round(W0 / scale * 36) / 36, atol=1e-9)
    assert rim.sum() > 50
    assert np.median(err[rim] / np.abs(Wb0[rim])) < 1e-3
    # the identity W0/W1 exist for: a'(W0 - W1 W1')a = piston-removed
    # variance of the bilinear interpolant over the aperture
    rng = np.random.default_rng(4)
    a = rng.standard_normal(rec.fnx * rec.fny)
    q_w = a @ W0 @ a - (W1 @ a) ** 2
    q_b = a @ Wb0 @ a - (Wb1 @ a) ** 2
    assert abs(q_w - q_b) < 1e-4 * abs(q_b)


def test_fitl_piston_and_spd(fit8, fit30):
    for cfg, fc, r, f in (fit8, fit30):
        o = FitOracle(r, f, nthread=4)
        rec = f.rec
        off = np.cumsum([0] + [rec.anx[d] * rec.any[d] for d in range(rec.ndm)])
        NW = f.NW()
        for d in range(rec.ndm):
            a = np.zeros(f.nact)
            a[off[d]:off[d + 1]] = 1.0
            got = o.fit_l(a)
            want = NW @ (NW.T @ a) + rec.tikdiag * a
            assert np.linalg.norm(got - want) <= 1e-9 * np.linalg.norm(want) + 1e-22, cfg["name"]
        rng = np.random.default_rng(7)
        u, v = rng.standard_normal(f.nact), rng.standard_normal(f.nact)
        luv, lvu = u @ o.fit_l(v), v @ o.fit_l(u)
        assert abs(luv - lvu) <= 1e-12 * np.sqrt(abs(u @ o.fit_l(u)) * abs(v @ o.fit_l(v)))
        assert u @ o.fit_l(u) > 0 and v @ o.fit_l(v) > 0


def test_slaving_structure(fit30):
    cfg, fc, r, f = fit30
    S = f.slaving()
    assert S is not None
    Sd = S.toarray()
    assert np.allclose(Sd, Sd.T, rtol=0, atol=1e-18)
    ev = np.linalg.eigvalsh(Sd)
    assert ev.min() > -1e-12 * ev.max()
    rec = f.rec
    off = np.cumsum([0] + [rec.anx[d] * rec.any[d] for d in range(rec.ndm)])
    for d in range(rec.ndm):
        one = np.zeros(f.nact)
        one[off[d]:off[d + 1]] = 1
        assert np.abs(Sd @ one).max() < 1e-12 * np.abs(Sd).max()
    cpl = f.actcpl()
    slave = cpl < fc["actthres"]
    assert slave.any() and not slave.all()
    assert np.all(Sd.diagonal()[slave] > 0)
    # rows of S only at slaves and their 4-neighbours
    near = slave.copy()
    for d in range(rec.ndm):
        nx, ny = rec.anx[d], rec.any[d]
        s2 = slave[off[d]:off[d + 1]].reshape(ny, nx)
        g = s2.copy()
        g[1:, :] |= s2[:-1, :]
        g[:-1, :] |= s2[1:, :]
        g[:, 1:] |= s2[:, :-1]
        g[:, :-1] |= s2[:, 1:]
        near[off[d]:off[d + 1]] = g.ravel()
    rows = np.abs(Sd).sum(1) > 0
    assert np.all(near[rows])
    # coupling: 1 on actuators inside the aperture footprint
    assert cpl.max() == 1 and cpl.min() >= 0


def test_nw_columns(fit30):
    cfg, fc, r, f = fit30
    rec = f.rec
    NW = f.NW()
    assert NW.shape == (f.nact, 4)  # 2 pistons + tip/tilt of the upper DM
    cpl = f.actcpl()
    off = np.cumsum([0] + [rec.anx[d] * rec.any[d] for d in range(rec.ndm)])
    fitscl = 1.0 / f.nfloc
    for d in range(rec.ndm):
        col = NW[:, d]
        seg = slice(off[d], off[d + 1])
        assert np.all(col[:off[d]] == 0) and np.all(col[off[d + 1]:] == 0)
        assert np.allclose(col[seg][cpl[seg] > 0.1], fitscl) and np.all(col[seg][cpl[seg] <= 0.1] == 0)
    g = f.grids()[1]
    xs = g["ox"] + g["dx"] * np.arange(g["nx"])
    ys = g["oy"] + g["dx"] * np.arange(g["ny"])
    X, Y = np.meshgrid(xs, ys)
    seg = slice(off[1], off[2])
    m = cpl[seg] > 0.1
    cx, cy = NW[seg, 2], NW[seg, 3]
    # tip column proportional to x, tilt to y, same factor
    fx = np.polyfit(X.ravel()[m], cx[m], 1)
    fy = np.polyfit(Y.ravel()[m], cy[m], 1)
    assert abs(fx[1]) < 1e-15 and abs(fy[1]) < 1e-15 and np.isclose(fx[0], fy[0], rtol=1e-12)
    assert np.all(cx[~m] == 0) and np.all(cy[~m] == 0)


def test_fit_oracle_solve_reduces_residual(fit8):
    cfg, fc, r, f = fit8
    o = FitOracle(r, f, nthread=4)
    x = np.random.default_rng(3).standard_normal(r.nx) * 1e-7
    a10, res10 = o.solve(x, 10)
    a100, res100 = o.solve(x, 100)
    assert res100 < res10 < 1
    b = o.fit_r(x)
    assert np.linalg.norm(o.fit_l(a100) - b) <= 1.0001 * res100 * np.linalg.norm(b)


def test_unsupported_fit_options_are_refused(fit8):
    cfg, fc, r, f = fit8
    from maos_amd import _abi
    for kw in (dict(actinterp=1), ):
        with pytest.raises(_abi.MtError):
            tomo.FitSetup(r, configs.fit_for(cfg, **kw))
    bad = configs.fit_for(cfg)
    bad["dm"] = [dict(bad["dm"][0], iac=0.3)]
    with pytest.raises(_abi.MtError):
        tomo.FitSetup(r, bad)
