"""Pins of the FP64 CPU oracle and of the host setup it consumes.

The reference ships no golden vectors for this path, and its CPU build needs
headers the image lacks, so every check here is fixed by mathematics or by an
independent library routine:
  FFT            numpy.fft and a brute-force DFT, both signs, mixed radices
  HX / HX'       direct bilinear evaluation at analytically computed ray
                 coordinates, linear-ramp closed form, adjoint identity
  GP             full-subaperture closed forms (pos 1 and 2), exact slope of a
                 plane wave on every subaperture (partial ones included)
  saloc / saa    covered fraction of each subaperture square by the annular
                 pupil (exact chord integral), and the kept set by saat
  L2'L2          Fourier eigenvalue cf^2 (1 - cos/2 - cos/2)^2
  TomoL / TomoR  bilinear-form and adjoint identities built from the pinned HX,
                 GP, W, PTT and L2 pieces
  PTT            PTT TT = I; I - TT PTT annihilates tip/tilt
  PCG            converges to the dense pseudo-inverse solution (tiny system)
  FDPCG          Hermitian PSD blocks, Minv = pinv(Mhat, 1e-7) (numpy SVD),
                 permutation bijections, and on a periodic problem the FD model
                 equals F A F^-1 of the dense spatial operator
"""
import math

import numpy as np
import pytest

from maos_amd import configs, tomo, workload
from oracle.oracle import Oracle, fft1d


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def small8_o(small8):
    cfg, r = small8
    return cfg, r, Oracle(r, nthread=4)


@pytest.fixture(scope="module")
def large30_o(large30):
    cfg, r = large30
    return cfg, r, Oracle(r, nthread=8)


# ----------------------------------------------------------------- FFT
@pytest.mark.parametrize("n", [2, 3, 4, 5, 6, 7, 8, 12, 14, 28, 54, 56, 96, 112, 150])
def test_fft_vs_dft(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    k = np.arange(n)
    for sign in (-1, 1):
        D = np.exp(sign * 2j * np.pi * np.outer(k, k) / n)
        assert rel(fft1d(x, sign), D @ x) < 1e-13
    assert rel(fft1d(x, -1), np.fft.fft(x)) < 1e-13
    assert rel(fft1d(x, 1), np.fft.ifft(x) * n) < 1e-13


# ------------------------------------------------------------ geometry
def test_geometry_small8(small8):
    cfg, r = small8
    rec = r.rec
    assert r.layer_shapes() == [(56, 56), (56, 56)]
    assert [rec.xdx[l] for l in range(2)] == [0.25, 0.25]
    assert (rec.pnx, rec.pny, rec.pox, rec.pdx) == (35, 35, -4.25, 0.25)
    assert [rec.nsa[w] for w in range(4)] == [204] * 4
    assert (rec.fd_bs, rec.fd_nb, rec.fd_nbhalf) == (8, 784, 420)


def test_geometry_large30(large30):
    cfg, r = large30
    rec = r.rec
    assert r.layer_shapes() == [(96, 96)] * 6
    # cone coordinates: dx_l = atmr.dx * (1 - h_l / 90 km)
    for l, h in enumerate(configs.LARGE_HT):
        assert abs(rec.xdx[l] - 0.5 * (1 - h / 90e3)) < 1e-15
    assert (rec.pnx, rec.pny, rec.pox, rec.pdx) == (63, 63, -15.5, 0.5)
    assert [rec.nsa[w] for w in range(6)] == [2820] * 6
    assert (rec.fd_bs, rec.fd_nb, rec.fd_nbhalf) == (6, 9216, 4704)


# ------------------------------------------------------------------ HX
def _layer_only(r, l, xl):
    x = np.zeros(r.nx)
    off = r.layer_offsets()
    x[off[l]:off[l + 1]] = xl.ravel()
    return x


def _bilinear_direct(r, w, l, xl):
    """opd of WFS w from layer l alone: bilinear interpolation at
    u = ((x_p * scale + theta * h) - ox_l) / dx_l, points with
    0 <= u <= n-1 in both axes (others miss the layer)."""
    rec = r.rec
    h = rec.ht[l]
    s = 1 - h / rec.hs[w]
    px = rec.pox + np.arange(rec.pnx) * rec.pdx
    py = rec.poy + np.arange(rec.pny) * rec.pdx
    u = (px * s + rec.thetax[w] * h - rec.xox[l]) / rec.xdx[l]
    v = (py * s + rec.thetay[w] * h - rec.xoy[l]) / rec.xdx[l]
    ny, nx = xl.shape
    U, V = np.meshgrid(u, v)
    ok = (U >= 0) & (U <= nx - 1) & (V >= 0) & (V <= ny - 1)
    i0 = np.clip(np.floor(U).astype(int), 0, nx - 1)
    j0 = np.clip(np.floor(V).astype(int), 0, ny - 1)
    fu, fv = U - i0, V - j0
    i1 = np.minimum(i0 + 1, nx - 1)
    j1 = np.minimum(j0 + 1, ny - 1)
    val = ((1 - fv) * ((1 - fu) * xl[j0, i0] + fu * xl[j0, i1]) +
           fv * ((1 - fu) * xl[j1, i0] + fu * xl[j1, i1]))
    return np.where(ok, val, 0.0).ravel()


@pytest.mark.parametrize("case", ["small8_o", "large30_o"])
def test_hx_direct_bilinear(case, request):
    cfg, r, o = request.getfixturevalue(case)
    rng = np.random.default_rng(11)
    for w in range(r.nwfs):
        for l in range(r.npsr):
            ny, nx = r.layer_shapes()[l]
            xl = rng.standard_normal((ny, nx))
            got = o.hx(w, _layer_only(r, l, xl))
            assert rel(got, _bilinear_direct(r, w, l, xl)) < 1e-12


def test_hx_ramp_closed_form(large30_o):
    cfg, r, o = large30_o
    rec = r.rec
    a, b = 0.37, -1.21
    for w in (0, 3):
        for l in (0, 5):
            ny, nx = r.layer_shapes()[l]
            X = rec.xox[l] + np.arange(nx) * rec.xdx[l]
            Y = rec.xoy[l] + np.arange(ny) * rec.xdx[l]
            xl = a * X[None, :] + b * Y[:, None]
            h = rec.ht[l]
            s = 1 - h / rec.hs[w]
            px = rec.pox + np.arange(rec.pnx) * rec.pdx
            py = rec.poy + np.arange(rec.pny) * rec.pdx
            exp = a * (px[None, :] * s + rec.thetax[w] * h) + b * (py[:, None] * s + rec.thetay[w] * h)
            assert rel(o.hx(w, _layer_only(r, l, xl)), exp.ravel()) < 1e-12


@pytest.mark.parametrize("case", ["small8_o", "large30_o"])
def test_hx_adjoint(case, request):
    cfg, r, o = request.getfixturevalue(case)
    rng = np.random.default_rng(3)
    x = rng.standard_normal(r.nx)
    y = rng.standard_normal(r.rec.pnx * r.rec.pny)
    for w in range(r.nwfs):
        lhs, rhs = y @ o.hx(w, x), x @ o.hxt(w, y)
        assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_hx_edge_partial_coverage(small8):
    """A ray footprint that leaves the layer grid (a direction outside the
    meta-pupil): points outside contribute nothing, points inside match
    direct bilinear evaluation."""
    import ctypes as C

    from maos_amd import _abi
    cfg, r = small8
    rec = _abi.MtRecon()
    C.pointer(rec)[0] = r.rec
    th = (C.c_double * r.nwfs)(*([1.2e-3] * r.nwfs))  # 247 arcsec
    rec.thetax = th
    rec.thetay = th

    class _R:
        pass
    rr = _R()
    rr.rec = rec
    o = Oracle(rr)
    rng = np.random.default_rng(9)
    l = 1
    ny, nx = r.layer_shapes()[l]
    xl = rng.standard_normal((ny, nx))
    view = _R()
    view.rec = rec
    exp = _bilinear_direct(view, 0, l, xl)
    assert 0 < np.count_nonzero(exp == 0) < exp.size
    x = np.zeros(r.nx)
    off = r.layer_offsets()
    x[off[l]:off[l + 1]] = xl.ravel()
    assert rel(o.hx(0, x), exp) < 1e-12
    y = rng.standard_normal(exp.size)
    assert abs(y @ o.hx(0, x) - x @ o.hxt(0, y)) < 1e-12 * abs(y @ o.hx(0, x))


# ------------------------------------------------------------------ GP
def _sa_stencil(r, w, isa):
    rec = r.rec
    G = r.gp(w).tocsr()
    nsa = rec.nsa[w]
    sx, sy = r.saloc(w)
    cx = int(round((sx[isa] - rec.pox) / rec.pdx))
    cy = int(round((sy[isa] - rec.poy) / rec.pdx))
    np1 = rec.pos + 1
    out = []
    for row in (isa, isa + nsa):
        st = np.zeros((np1, np1))
        lo, hi = G.indptr[row], G.indptr[row + 1]
        for c, v in zip(G.indices[lo:hi], G.data[lo:hi]):
            zx, zy = c % rec.pnx - cx, c // rec.pnx - cy
            assert 0 <= zx < np1 and 0 <= zy < np1
            st[zy, zx] += v
        out.append(st)
    return out


@pytest.mark.parametrize("case", ["small8", "large30"])
def test_gp_full_subaperture_closed_form(case, request):
    cfg, r = request.getfixturevalue(case)
    dsa = r.rec.dsa[0]
    saa = r.saa(0)
    full = np.flatnonzero(np.abs(saa - 1) < 1e-12)
    assert full.size > 50
    if r.rec.pos == 1:
        ex = np.array([[-0.5, 0.5], [-0.5, 0.5]]) / dsa
    else:
        ex = np.array([[-0.25, 0, 0.25], [-0.5, 0, 0.5], [-0.25, 0, 0.25]]) / dsa
    for isa in full[::17]:
        gx, gy = _sa_stencil(r, 0, isa)
        assert np.abs(gx - ex).max() < 1e-12 / dsa
        assert np.abs(gy - ex.T).max() < 1e-12 / dsa


@pytest.mark.parametrize("case", ["small8", "large30"])
def test_gp_plane_wave_every_subaperture(case, request):
    """The amplitude-weighted average gradient of a plane wave is its slope,
    on full and partial subapertures alike; constants have zero gradient."""
    cfg, r = request.getfixturevalue(case)
    rec = r.rec
    X, Y = np.meshgrid(rec.pox + np.arange(rec.pnx) * rec.pdx, rec.poy + np.arange(rec.pny) * rec.pdx)
    a, b = 2.5e-6, -1.25e-6
    for w in range(r.nwfs):
        G = r.gp(w)
        nsa = rec.nsa[w]
        g = G @ (a * X + b * Y).ravel()
        assert np.abs(g[:nsa] - a).max() < 1e-9 * abs(a)
        assert np.abs(g[nsa:] - b).max() < 1e-9 * abs(a)
        assert np.abs(G @ np.ones(rec.pnx * rec.pny)).max() < 1e-9


def _annulus_square_fraction(x0, y0, d, R, Rin, m=4000):
    """Exact covered fraction of the squares [x0, x0+d] x [y0, y0+d] by the
    annulus Rin <= r <= R about the origin: the chord length of each disk in
    y, integrated over x by the midpoint rule (m points)."""
    t = (np.arange(m) + 0.5) / m

    def disk(rad):
        x = x0[:, None] + d * t[None, :]
        s = np.sqrt(np.clip(rad * rad - x * x, 0, None))
        lo, hi = np.maximum(y0[:, None], -s), np.minimum(y0[:, None] + d, s)
        return np.clip(hi - lo, 0, None).sum(axis=1) * d / m

    return (disk(R) - (disk(Rin) if Rin > 0 else 0)) / (d * d)


@pytest.mark.parametrize("case", ["small8", "large30"])
def test_subapertures_against_pupil_area(case, request):
    """saloc and saa (maos/setup_powfs.c:92-404) against the geometry they
    sample: each subaperture's saa is the covered fraction of its square by
    the annular pupil, to 0.01 (the reference's anti-aliased rim pixels enter
    squared: 0.006 and 0.007 here); squares wholly inside have saa exactly 1;
    the subapertures kept are those of the order x order grid covering more
    than saat.  Squares moved by half a fine pixel (dsa / 64) already miss the
    0.01 bound, so a misplaced grid fails this."""
    cfg, r = request.getfixturevalue(case)
    pw = cfg["powfs"][0]
    D, Din, dsa, saat = cfg["aper_d"], cfg["aper_din"], pw["dsa"], pw["saat"]
    for w in range(r.nwfs):
        x, y = r.saloc(w)
        saa = r.saa(w)
        ex = _annulus_square_fraction(x, y, dsa, D / 2, Din / 2)
        assert np.abs(saa - ex).max() < 0.01
        assert np.array_equal(saa == 1, ex > 1 - 1e-12)
        moved = _annulus_square_fraction(x + dsa / 64, y, dsa, D / 2, Din / 2)
        assert np.abs(saa - moved).max() > 0.01
    order = int(math.ceil(D / dsa))
    ii = np.arange(-(order // 2), (order - 1) // 2 + 1) + (-0.5 if order & 1 else 0.0)
    gx, gy = (a.ravel() * dsa for a in np.meshgrid(ii, ii))
    frac = _annulus_square_fraction(gx, gy, dsa, D / 2, Din / 2)
    assert np.abs(frac - saat).min() > 0.01  # no square at the threshold
    want = {(round(2 * a / dsa), round(2 * b / dsa)) for a, b, f in zip(gx, gy, frac) if f > saat}
    x, y = r.saloc(0)
    assert {(round(2 * a / dsa), round(2 * b / dsa)) for a, b in zip(x, y)} == want


# ------------------------------------------------------------- L2, W, PTT
def test_l2_fourier_eigenvalue(small8_o):
    cfg, r, o = small8_o
    rec = r.rec
    wt = np.asarray(cfg["wt"]) / np.sum(cfg["wt"])
    for l in range(r.npsr):
        ny, nx = r.layer_shapes()[l]
        # laplacian_coef, lib/laplacian.c:26-32
        cf = (2 * math.pi / 0.5e-6) * math.sqrt((cfg["r0z"] / rec.xdx[l]) ** (5 / 3) / 3.44 / wt[l]) \
            * (2 - 2 ** (-1 / 3) - 2 ** (-1 / 6)) ** -0.5
        cf2 = cf * cf * cfg["tomoscale"] * cfg["tomo_cxxscale"]
        assert abs(r.l2c()[l] - cf2 / 16) < 1e-12 * cf2
        for kx, ky in [(1, 0), (3, 5), (28, 28)]:
            ix, iy = np.meshgrid(np.arange(nx), np.arange(ny))
            v = np.cos(2 * np.pi * (kx * ix / nx + ky * iy / ny))
            lam = cf2 * (1 - 0.5 * math.cos(2 * math.pi * kx / nx) - 0.5 * math.cos(2 * math.pi * ky / ny)) ** 2
            # the stencil sums terms ~cf^2 to eigenvalues as small as ~1e-5 cf^2
            assert rel(o.l2(l, v.ravel()), lam * v.ravel()) < 1e-9


@pytest.mark.parametrize("case", ["small8", "large30"])
def test_nea_and_ptt(case, request):
    cfg, r = request.getfixturevalue(case)
    for w in range(r.nwfs):
        ne = r.neai(w)
        saa = r.saa(w)
        nea2 = (cfg["powfs"][0]["nearecon"] / 206265000.0) ** 2
        assert np.allclose(ne[:, 0], cfg["tomoscale"] * saa / nea2, rtol=1e-14)
        assert np.allclose(ne[:, 1], ne[:, 0]) and not ne[:, 2].any()
        P = r.ptt(w)
        assert P is not None  # split tomography removes TT from every high-order WFS
        nsa = r.nsa(w)
        TT = np.zeros((2 * nsa, 2))
        TT[:nsa, 0] = 1
        TT[nsa:, 1] = 1
        assert np.allclose(P @ TT, np.eye(2), atol=1e-13)
        g = np.random.default_rng(w).standard_normal(2 * nsa)
        assert np.abs(P @ (g - TT @ (P @ g))).max() < 1e-13


# ---------------------------------------------------------- TomoL / TomoR
def _wgh(r, o, w, x):
    """W G HX x for WFS w (pinned pieces)."""
    G = r.gp(w)
    g = G @ o.hx(w, x)
    ne = r.neai(w)
    nsa = r.nsa(w)
    return g, np.concatenate([ne[:, 0] * g[:nsa] + ne[:, 2] * g[nsa:], ne[:, 2] * g[:nsa] + ne[:, 1] * g[nsa:]])


@pytest.mark.parametrize("case", ["small8_o", "large30_o"])
def test_tomol_bilinear_form(case, request):
    cfg, r, o = request.getfixturevalue(case)
    rng = np.random.default_rng(21)
    x, y = rng.standard_normal(r.nx), rng.standard_normal(r.nx)
    off = r.layer_offsets()
    exp = 0.0
    for w in range(r.nwfs):
        _, wx = _wgh(r, o, w, x)
        gy, _ = _wgh(r, o, w, y)
        exp += wx @ gy
    for l in range(r.npsr):
        exp += y[off[l]:off[l + 1]] @ o.l2(l, x[off[l]:off[l + 1]])
    got = y @ o.tomo_l(x)
    assert abs(got - exp) <= 1e-11 * abs(exp)
    # symmetric
    assert abs(y @ o.tomo_l(x) - x @ o.tomo_l(y)) <= 1e-12 * abs(got)


@pytest.mark.parametrize("case", ["small8_o", "large30_o"])
def test_tomor_adjoint(case, request):
    cfg, r, o = request.getfixturevalue(case)
    rng = np.random.default_rng(22)
    g = rng.standard_normal(r.ng)
    x = rng.standard_normal(r.nx)
    goff = r.grad_offsets()
    exp = 0.0
    for w in range(r.nwfs):
        gw = g[goff[w]:goff[w + 1]].copy()
        P = r.ptt(w)
        nsa = r.nsa(w)
        tt = P @ gw
        gw[:nsa] -= tt[0]
        gw[nsa:] -= tt[1]
        ne = r.neai(w)
        wg = np.concatenate([ne[:, 0] * gw[:nsa] + ne[:, 2] * gw[nsa:], ne[:, 2] * gw[:nsa] + ne[:, 1] * gw[nsa:]])
        exp += wg @ (r.gp(w) @ o.hx(w, x))
    got = x @ o.tomo_r(g)
    assert abs(got - exp) <= 1e-11 * abs(exp)


# ------------------------------------------------------------------ PCG
def tiny_cfg(**kw):
    return configs.small8(name="tiny", aper_d=2.0, aper_din=0.0,
                          powfs=[dict(dsa=0.5, dx=1 / 16, saat=0.3, nearecon=50.0, hs=float("inf"),
                                      trs=0, lo=0, skip=0, dtrat=1)],
                          os=[1, 1], tomo_pos=1,
                          wfs=[dict(thetax=tx, thetay=ty, powfs=0) for tx, ty in [(20, 0), (-20, 10)]], **kw)


def test_pcg_converges_to_pseudo_inverse():
    cfg = tiny_cfg()
    r = tomo.Recon(cfg)
    o = Oracle(r)
    n = r.nx
    A = np.stack([o.tomo_l(e) for e in np.eye(n)], axis=1)
    assert np.abs(A - A.T).max() <= 1e-12 * np.abs(A).max()
    ev = np.linalg.eigvalsh(A)
    assert ev.min() > -1e-10 * ev.max()
    # null space = per-layer pistons
    off = r.layer_offsets()
    for l in range(r.npsr):
        e = np.zeros(n)
        e[off[l]:off[l + 1]] = 1
        assert np.linalg.norm(A @ e) < 1e-9 * np.linalg.norm(A)
    g = workload.make_inputs(cfg, r, 5)
    b = o.tomo_r(g)
    xs = np.linalg.pinv(A, rcond=1e-12) @ b
    # CG needs ~rank iterations (restarted every 100 as lib/pcg.c:55 does);
    # FDPCG-PCG reaches the solution to rounding in 50.  (Iterated far past
    # convergence, rounding left in the DC bin is amplified by the FDPCG
    # DC block and the recurrence degrades, in the reference as here.)
    for precond, maxit, tol in ((False, 400, 1e-8), (True, 50, 1e-10)):
        x, res = o.pcg(b, maxit, precond=precond)
        for l in range(r.npsr):
            x[off[l]:off[l + 1]] -= x[off[l]:off[l + 1]].mean()
        assert rel(x, xs) < tol
        assert res < 1e-9


def test_pcg_fdpcg_accelerates(large30_o):
    cfg, r, o = large30_o
    b = o.tomo_r(workload.make_inputs(cfg, r, 2))
    _, r_cg = o.pcg(b, 10, precond=False)
    _, r_pcg = o.pcg(b, 10, precond=True)
    assert r_pcg < 0.5 * r_cg


# ---------------------------------------------------------------- FDPCG
@pytest.mark.parametrize("case", ["small8", "large30"])
def test_fdpcg_blocks(case, request):
    cfg, r = request.getfixturevalue(case)
    fd = r.fdpcg()
    Mh, Mi = fd["Mhat"], fd["Minv"]
    sel = np.arange(0, fd["nb"], max(1, fd["nb"] // 300))
    for ib in sel:
        H = Mh[ib]
        assert np.abs(H - H.conj().T).max() <= 1e-13 * np.abs(H).max()
        ev = np.linalg.eigvalsh(H)
        assert ev.min() >= -1e-10 * ev.max()
        assert np.abs(Mi[ib] - np.linalg.pinv(H, rcond=1e-7, hermitian=False)).max() <= 1e-9 * np.abs(Mi[ib]).max()


@pytest.mark.parametrize("case", ["small8", "large30"])
def test_fdpcg_permutations(case, request):
    cfg, r = request.getfixturevalue(case)
    fd = r.fdpcg()
    perm, ph, bs, nbh = fd["perm"], fd["permhf"], fd["bs"], fd["nbhalf"]
    assert np.array_equal(np.sort(perm), np.arange(r.nx))
    nx, ny = r.layer_shapes()[0][1], r.layer_shapes()[0][0]
    N = nx * ny
    for e_full, e_half in zip(perm[:nbh * bs], ph):
        l, rem = divmod(e_full, N)
        kx, ky = rem % nx, rem // nx
        if e_half < 0:  # conjugate of the mirrored frequency
            assert kx > nx // 2
            assert -e_half == l * N + (nx - kx) + ((ny - ky) % ny) * nx
        else:
            assert kx <= nx // 2 and e_half == e_full


def test_fdpcg_periodic_exact():
    """Periodic, fully illuminated NGS problem with integer shifts: the FD
    model (before its sqrt(eps) regularisation of the Laplacian PSD) is
    exactly the Fourier representation F A F^-1 of the spatial operator
    A = sum_w HX' G' W G HX + L2'L2 (FFTW sign convention)."""
    import ctypes as C

    from maos_amd import _abi
    N, nps, dx, dsa = 16, 2, 0.5, 0.5
    ht = np.array([0.0, 4000.0])
    shifts = [(1, 0), (-2, 3)]  # layer-2 displacement of each WFS, in pixels
    thx = np.array([s[0] * dx / ht[1] for s in shifts])
    thy = np.array([s[1] * dx / ht[1] for s in shifts])
    neam = np.array([0.7, 1.3])
    l2cf = np.array([0.9, 1.7])
    salx, saly = np.meshgrid((np.arange(N) - N // 2) * dx, (np.arange(N) - N // 2) * dx)
    nsa = N * N
    fin = _abi.MtFdpcgIn()
    keep = []

    def arr(a, t=C.c_double):
        a = np.ascontiguousarray(a, dtype=np.float64 if t is C.c_double else np.int64)
        keep.append(a)
        return a.ctypes.data_as(C.POINTER(t))
    fin.nps, fin.nx, fin.ny, fin.os = nps, arr([N] * nps, C.c_long), arr([N] * nps, C.c_long), arr([1] * nps, C.c_long)
    fin.pos, fin.dxp, fin.ht, fin.l2cf = 1, dx, arr(ht), arr(l2cf)
    fin.nsa, fin.dsa = nsa, dsa
    fin.salocx, fin.salocy, fin.saa = arr(salx.ravel()), arr(saly.ravel()), arr(np.ones(nsa))
    fin.nwfs, fin.thetax, fin.thetay, fin.neam = 2, arr(thx), arr(thy), arr(neam)
    fout = _abi.MtFdpcgOut()
    _abi.check(_abi.lib().mt_fdpcg_prepare(C.byref(fin), C.byref(fout)), "mt_fdpcg_prepare")
    bs, nb = fout.bs, fout.nb
    Mh = np.ctypeslib.as_array(fout.Mhat, shape=(nb * bs * bs * 2,)).copy().view(np.complex128)
    Mh = Mh.reshape(nb, bs, bs).transpose(0, 2, 1)  # [block][row][col]
    perm = np.ctypeslib.as_array(fout.perm, shape=(nb * bs,)).copy()
    _abi.lib().mt_fdpcg_free(C.byref(fout))

    # dense periodic spatial operator on nps layers of N x N
    n1 = N * N

    def shift_op(dxp, dyp):  # (S x)(i, j) = x(i + dxp, j + dyp), periodic
        S = np.zeros((n1, n1))
        for j in range(N):
            for i in range(N):
                S[i + j * N, (i + dxp) % N + ((j + dyp) % N) * N] = 1
        return S
    Gx = 0.5 / dsa * (-shift_op(0, 0) + shift_op(1, 0) - shift_op(0, 1) + shift_op(1, 1))
    Gy = 0.5 / dsa * (-shift_op(0, 0) - shift_op(1, 0) + shift_op(0, 1) + shift_op(1, 1))
    A = np.zeros((nps * n1, nps * n1))
    for w in range(2):
        H = np.hstack([shift_op(0, 0), shift_op(*shifts[w])])
        for G in (Gx, Gy):
            GH = G @ H
            A += GH.T @ GH / neam[w] ** 2
    for l in range(nps):
        L = l2cf[l] * (-shift_op(0, 0) + 0.25 * (shift_op(1, 0) + shift_op(-1, 0) + shift_op(0, 1) + shift_op(0, -1)))
        A[l * n1:(l + 1) * n1, l * n1:(l + 1) * n1] += L.T @ L
    k = np.arange(N)
    F1 = np.exp(-2j * np.pi * np.outer(k, k) / N)
    F = np.kron(F1, F1)  # index i + j*N with i fastest
    FF = np.kron(np.eye(nps), F)
    Ahat = FF @ A @ np.linalg.inv(FF)
    # regulariser of invpsd: max|psd| * sqrt(eps) per layer (maos/fdpcg.c:396-401)
    reg = []
    for l in range(nps):
        th = 2 * np.pi * k / N
        lam = l2cf[l] ** 2 * (1 - 0.5 * np.cos(th)[None, :] - 0.5 * np.cos(th)[:, None]) ** 2
        reg.append(np.abs(lam).max() * math.sqrt(2.220446049250313e-16))
    scale = np.abs(Ahat).max()
    for ib in range(nb):
        idx = perm[ib * bs:(ib + 1) * bs]
        lay = idx // n1
        sub = Ahat[np.ix_(idx, idx)].copy()
        sub[np.diag_indices(bs)] += np.array([reg[l] for l in lay])
        # Mhat is pre-scaled by N (nx*ny) for the unnormalized FFT pair
        assert np.abs(Mh[ib] / n1 - sub).max() <= 1e-12 * scale, ib


# ------------------------------------------------- GP, partial subapertures
# Brute-force quadrature of the average gradient that lib/mkg.c:113-470
# states: on the WFS lattice (gloc, spacing powfs.dx) the phase and the
# amplitude are bilinear in each cell; a cell counts only where the
# amplitude is nonzero at all four of its corners (lib/mkg.c:78-82); the x
# slope of a subaperture is
#     sum over counted cells of  int_{cell & subap} A dphi/dx
#   / sum over counted cells of  int_{cell & subap} A
# (zero when the amplitude integral is below 1 % of the subaperture area,
# lib/mkg.c:450), with A the gray-pixel annulus of the aperture (fraction of
# 10 x 10 sub-sample centres inside, math/loc.c:815-862) times the share of
# each lattice point's bilinear footprint on existing recon-ploc points, and
# phi on the lattice the bilinear interpolation of the recon-ploc phase
# normalised by that share (lib/mkg.c:211-232, 377-426).  Integrals by
# 3-point Gauss-Legendre per axis on each cell/subaperture overlap, exact
# for these bilinear x linear integrands.
_GL = np.polynomial.legendre.leggauss(3)


def _gp_brute(r, cfg, w, phi, isas, gate=True, sign=1.0, clip=True):
    rec = r.rec
    D, Din = cfg["aper_d"], cfg["aper_din"]
    gdx = cfg["powfs"][cfg["wfs"][w]["powfs"]]["dx"]
    ngx = int((D + 1) / gdx + 1)
    gox = -(ngx // 2) * gdx
    dsa = rec.dsa[w]
    sx, sy = r.saloc(w)
    pnx, pny, pox, poy, pdx = rec.pnx, rec.pny, rec.pox, rec.poy, rec.pdx
    P = phi.reshape(pny, pnx)
    sub = (np.arange(10) - 4.5) * gdx / 10

    def amp_frac(x, y, R):
        xs = x[..., None, None] + sub[None, None, :, None]
        ys = y[..., None, None] + sub[None, None, None, :]
        return ((xs * xs + ys * ys) <= R * R).mean(axis=(-1, -2))

    def share_and_phi(x, y):
        fx, fy = (x - pox) / pdx, (y - poy) / pdx
        ix, iy = np.floor(fx).astype(int), np.floor(fy).astype(int)
        fx, fy = fx - ix, fy - iy
        wsum = np.zeros_like(x)
        val = np.zeros_like(x)
        for dx_, dy_, wt in ((0, 0, (1 - fx) * (1 - fy)), (1, 0, fx * (1 - fy)),
                             (0, 1, (1 - fx) * fy), (1, 1, fx * fy)):
            jx, jy = ix + dx_, iy + dy_
            ok = (jx >= 0) & (jx < pnx) & (jy >= 0) & (jy < pny)
            wsum += np.where(ok, wt, 0)
            val += np.where(ok, wt * P[np.clip(jy, 0, pny - 1), np.clip(jx, 0, pnx - 1)], 0)
        return wsum, np.where(wsum > 0, val / np.where(wsum > 0, wsum, 1), 0)

    q, qw = (_GL[0] + 1) / 2, _GL[1] / 2
    out = np.zeros((2, len(isas)))
    for k, isa in enumerate(isas):
        x0, y0 = sx[isa], sy[isa]
        i0 = int(np.floor((x0 - gox) / gdx)) - 1
        j0 = int(np.floor((y0 - gox) / gdx)) - 1
        n = int(round(dsa / gdx)) + 3
        ii, jj = np.arange(i0, i0 + n + 1), np.arange(j0, j0 + n + 1)
        X, Y = np.meshgrid(gox + ii * gdx, gox + jj * gdx)
        exists = ((ii >= 0) & (ii < ngx))[None, :] & ((jj >= 0) & (jj < ngx))[:, None]
        A = amp_frac(X, Y, D / 2) - (amp_frac(X, Y, Din / 2) if Din > 0 else 0)
        share, ph = share_and_phi(X, Y)
        A = np.where(exists, A * share, 0)
        num = np.zeros(2)
        den = 0.0
        for cj in range(n):
            for ci in range(n):
                xa, xb = max(x0, X[cj, ci]), min(x0 + dsa, X[cj, ci + 1])
                ya, yb = max(y0, Y[cj, ci]), min(y0 + dsa, Y[cj + 1, ci])
                if xb - xa <= 1e-12 * gdx or yb - ya <= 1e-12 * gdx:
                    continue
                if not clip:
                    xa, xb, ya, yb = X[cj, ci], X[cj, ci + 1], Y[cj, ci], Y[cj + 1, ci]
                a = A[cj:cj + 2, ci:ci + 2]
                if gate and not (np.abs(a) > 5e-16).all():
                    continue
                f = ph[cj:cj + 2, ci:ci + 2]
                u = (xa + (xb - xa) * q - X[cj, ci]) / gdx
                v = (ya + (yb - ya) * q - Y[cj, ci]) / gdx
                U, V = np.meshgrid(u, v)
                W = np.outer(qw, qw) * (xb - xa) * (yb - ya)
                Av = a[0, 0] * (1 - U) * (1 - V) + a[0, 1] * U * (1 - V) + a[1, 0] * (1 - U) * V + a[1, 1] * U * V
                dpx = ((f[0, 1] - f[0, 0]) * (1 - V) + (f[1, 1] - f[1, 0]) * V) / gdx
                dpy = ((f[1, 0] - f[0, 0]) * (1 - U) + (f[1, 1] - f[0, 1]) * U) / gdx
                num += [np.sum(W * Av * dpx), np.sum(W * Av * dpy)]
                den += np.sum(W * Av)
        if den <= 0.01 * dsa * dsa:
            continue
        out[:, k] = sign * num / den
    return out


def _tiny_annular(dx=1 / 64):
    from maos_amd import configs
    c = configs.small8(name="tiny_annular", aper_d=1.5, aper_din=0.4, tomo_precond=0)
    c["powfs"][0]["dx"] = dx
    return c


@pytest.mark.parametrize("case", ["tiny_annular", "tiny_offgrid", "small8"])
def test_gp_partial_subapertures_brute_force(case, request):
    """GP rows against the brute-force integral above, for random phases and
    a quadratic phase, on every subaperture of a tiny annular pupil (all
    partial) and on every partial subaperture of small8 plus a sample of full
    ones.  Variants of the integral with one plausible mistake (the
    amplitude gate dropped, the slope sign flipped, whole cells integrated
    instead of their overlap with the subaperture) must miss, so the check
    can tell them apart."""
    from maos_amd import tomo
    if case.startswith("tiny"):
        # offgrid: a lattice spacing that puts subaperture edges inside cells
        cfg = _tiny_annular(1 / 64 if case == "tiny_annular" else 0.5 / 15.3)
        r = tomo.Recon(cfg)
    else:
        cfg, r = request.getfixturevalue(case)
    rec = r.rec
    saa = r.saa(0)
    part = np.flatnonzero(saa < 1 - 1e-12)
    full = np.flatnonzero(saa >= 1 - 1e-12)
    isas = np.concatenate([part, full[::23]])
    assert part.size >= 4
    X, Y = np.meshgrid(rec.pox + np.arange(rec.pnx) * rec.pdx, rec.poy + np.arange(rec.pny) * rec.pdx)
    rng = np.random.default_rng(9)
    phis = [rng.standard_normal(rec.pnx * rec.pny),
            (0.3 * X * X - 0.7 * X * Y + 1.1 * Y * Y + 0.2 * X).ravel()]
    G = r.gp(0)
    nsa = rec.nsa[0]
    for phi in phis:
        g = G @ phi
        gx, gy = g[:nsa][isas], g[nsa:][isas]
        bx, by = _gp_brute(r, cfg, 0, phi, isas)
        scale = np.abs(np.concatenate([bx, by])).max()
        assert np.abs(gx - bx).max() < 1e-10 * scale
        assert np.abs(gy - by).max() < 1e-10 * scale
    # the check discriminates: each mistaken variant misses on partial ones
    phi = phis[0]
    g = G @ phi
    got = np.concatenate([g[:nsa][part], g[nsa:][part]])
    for kw in (dict(gate=False), dict(sign=-1.0)) + ((dict(clip=False),) if case == "tiny_offgrid" else ()):
        bx, by = _gp_brute(r, cfg, 0, phi, part, **kw)
        assert np.abs(got - np.concatenate([bx, by])).max() > 1e-3 * np.abs(got).max(), kw
