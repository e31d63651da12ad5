// Host setup orchestration and its C ABI (mt_setup_*).
//
// Follows the order of MAOS's setup on the tomography path:
//   setup_parms derived values   maos/setup_parms.c:2272-2321 (atmr.hs, atmr.dx),
//                                 1719 (atmr.r0), 2189-2191 (split => lo skipped)
//   setup_recon_prep              maos/setup_recon_prep.c:1043-1140
//     ploc                        :36-83
//     xloc                        :168-268
//     GP                          :481-555
//     TT                          :867-914
//   setup_recon_saneai            maos/setup_recon.c:51-252 (geometric NEA)
//   setup_recon_TTFR              maos/setup_recon.c:254-272
//   setup_recon_tomo_prep (L2)    maos/setup_recon.c:290-335
//   fdpcg_prepare                 maos/fdpcg.c:331-580
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <cstdlib>
#include <memory>
#include <type_traits>

#include "setup_internal.hpp"

namespace mt {

static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }
const char *last_error_cstr() { return g_last_error.c_str(); }

}  // namespace mt

using namespace mt;


namespace {

void build_regularisation(mt_setup *s) {
    // setup_recon_tomo_prep: clip weights to [0.01,1], normalise sum |wt| = 1
    // (maos/setup_recon.c:304-307), L2 scale sqrt(cxxscale*TOMOSCALE) (:335),
    // GPU coefficient l2c (cuda/tomo.cu:159-162).
    s->wt = s->wt0;
    double sum = 0;
    for (double &w : s->wt) {
        w = std::min(std::max(w, 0.01), 1.);
        sum += std::fabs(w);
    }
    for (double &w : s->wt) w /= sum;
    s->l2c.resize(s->npsr);
    s->l2cf.resize(s->npsr);
    for (int l = 0; l < s->npsr; l++) {
        const double cf = laplacian_coef(s->r0, s->wt[l], s->xg[l].dx);
        s->l2cf[l] = cf * std::sqrt(s->cxxscale * s->tomoscale);
        const double t = cf * 0.25;
        s->l2c[l] = t * t * s->cxxscale * s->tomoscale;
    }
}

FdpcgIn fdpcg_input(const mt_setup *s) {
    // the single high-order powfs (maos/fdpcg.c:332-342)
    int hipowfs = -1;
    for (size_t ip = 0; ip < s->powfs.size(); ip++) {
        if (!s->powfs[ip].lo && !s->powfs[ip].skip) {
            if (hipowfs == -1) hipowfs = (int)ip;
            else throw Error(MT_EINVAL, "fdpcg: multiple high order powfs");
        }
    }
    if (hipowfs < 0) throw Error(MT_EINVAL, "fdpcg: no high order powfs");
    FdpcgIn in;
    in.nps = s->npsr;
    for (int l = 0; l < s->npsr; l++) {
        in.nx.push_back(s->xg[l].nx);
        in.ny.push_back(s->xg[l].ny);
        in.os.push_back(s->os[l]);
    }
    in.pos = s->pos;
    in.dxp = s->pg.dx;
    in.ht = s->ht;
    in.l2cf = s->l2cf;
    in.sa = &s->sa[hipowfs];
    for (size_t iw = 0; iw < s->wfs.size(); iw++) {
        if (s->wfs[iw].powfs != hipowfs) continue;
        in.thetax.push_back(s->wfs[iw].thetax);
        in.thetay.push_back(s->wfs[iw].thetay);
        in.neam.push_back(s->neam[iw]);
    }
    return in;
}

void build_fdpcg(mt_setup *s) {
    if (s->precond != 1) {
        s->fd = FdpcgOut();
        return;
    }
    s->fd = fdpcg_prepare(fdpcg_input(s));
}

// A new turbulence profile changes only the regulariser (layer weights and
// r0 enter through l2cf): the blocks' WFS sums are kept from the setup and
// only the diagonal, tolerances and block inverses are redone.
void update_fdpcg(mt_setup *s) {
    if (s->precond != 1) return;
    if (s->fd.MhatW.empty()) {
        build_fdpcg(s);
        return;
    }
    fdpcg_finish(s->fd, fdpcg_input(s));
}

void build(mt_setup *s, const mt_parms_t *p) {
    if (!p || p->npsr <= 0 || !p->ht || !p->wt || !p->os || p->npowfs <= 0 || !p->powfs ||
        p->nwfs <= 0 || !p->wfs)
        throw Error(MT_EINVAL, "mt_setup_create: missing parameters");
    if (!(p->aper_d > 0) || p->tomo_pos <= 0) throw Error(MT_EINVAL, "invalid aper.d or tomo.pos");
    s->D = p->aper_d;
    s->Din = p->aper_din;
    s->npsr = p->npsr;
    s->ht.assign(p->ht, p->ht + p->npsr);
    s->wt0.assign(p->wt, p->wt + p->npsr);
    s->os.assign(p->os, p->os + p->npsr);
    for (int l = 0; l < s->npsr; l++)
        if (s->os[l] <= 0) throw Error(MT_EINVAL, "atmr.os must be positive");
    s->powfs.assign(p->powfs, p->powfs + p->npowfs);
    s->wfs.assign(p->wfs, p->wfs + p->nwfs);
    s->pos = p->tomo_pos;
    s->guard = p->tomo_guard;
    s->cxxscale = p->tomo_cxxscale;
    s->precond = p->tomo_precond;
    s->tomoscale = p->tomoscale > 0 ? p->tomoscale : 1e-12;
    s->L0 = p->L0;
    s->ptt_lhs = !p->recon_split || p->tomo_splitlrt;
    s->split = p->recon_split;
    s->za = p->za;
    for (auto &pw : s->powfs) {
        if (p->recon_split && pw.lo) pw.skip = 1;  // maos/setup_parms.c:2189-2191
        if (!(pw.dsa > 0) || !(pw.dx > 0) || pw.dtrat <= 0 || !(pw.nearecon > 0))
            throw Error(MT_EINVAL, "invalid powfs parameters");
    }
    for (auto &w : s->wfs)
        if (w.powfs < 0 || w.powfs >= p->npowfs) throw Error(MT_EINVAL, "wfs.powfs out of range");
    // atmr.hs: guide star height of the high-order powfs (cone coordinates)
    double hs = NAN;
    if (p->tomo_cone) {
        for (auto &pw : s->powfs) {
            if (pw.lo || pw.skip) continue;
            if (std::isnan(hs)) hs = pw.hs;
            else if ((std::isfinite(hs) || std::isfinite(pw.hs)) && std::fabs(hs - pw.hs) > 1000 && pw.hs > hs)
                hs = pw.hs;
        }
    }
    s->atmr_hs = std::isnan(hs) ? INFINITY : hs;
    s->atmr_dx = p->atmr_dx;
    if (s->atmr_dx < EPS_SYS) {
        double mindsa = INFINITY;
        for (auto &pw : s->powfs)
            if (!pw.lo && !pw.skip) mindsa = std::min(mindsa, pw.dsa);
        s->atmr_dx = mindsa;
    }
    s->r0 = p->r0z * std::pow(std::cos(p->za), 3. / 5.);
    // beam directions (maos/setup_parms.c:1945-1984): WFS, then evl/fit
    for (auto &w : s->wfs) s->dirs.push_back({w.thetax, w.thetay, s->powfs[w.powfs].hs});
    for (int i = 0; i < p->ndir; i++) s->dirs.push_back({p->dir_thetax[i], p->dir_thetay[i], p->dir_hs[i]});

    // ploc (setup_recon_ploc): square, no offset
    {
        const double dxr = s->atmr_dx / s->pos;
        s->pg = create_metapupil(s->dirs, s->D, 0, dxr, dxr, 0, s->guard * dxr, 0, 0, 0, 1);
    }
    // xloc (setup_recon_xloc) with cone-scaled sampling (setup_recon_prep.c:1129-1133)
    std::vector<double> xdxr(s->npsr);
    for (int l = 0; l < s->npsr; l++) xdxr[l] = (s->atmr_dx / s->os[l]) * (1.0 - s->ht[l] / s->atmr_hs);
    long nin0 = 0;
    if (s->precond == 1) {
        long nxmin = LONG_MAX, nymin = LONG_MAX, nxmax = 0, nymax = 0;
        for (int l = 0; l < s->npsr; l++) {
            Grid g = create_metapupil(s->dirs, s->D, s->ht[l], xdxr[l], xdxr[l], 0, xdxr[l] * s->guard, 0, 0, 0, 1);
            long nxi = g.nx / s->os[l], nyi = g.ny / s->os[l];
            nxmax = std::max(nxmax, nxi);
            nymax = std::max(nymax, nyi);
            nxmin = std::min(nxmin, nxi);
            nymin = std::min(nymin, nyi);
        }
        long nx = (long)std::max((double)nxmax, nxmin * 1.5);
        long ny = (long)std::max((double)nymax, nymin * 1.5);
        nin0 = nextfftsize(std::max(nx, ny));
        while (nin0 & 1) nin0 = nextfftsize(nin0 + 1);
    }
    s->xg.resize(s->npsr);
    for (int l = 0; l < s->npsr; l++) {
        const long nin = nin0 * s->os[l];
        s->xg[l] = create_metapupil(s->dirs, s->D, s->ht[l], xdxr[l], xdxr[l], 0, s->guard * xdxr[l], nin, nin, 0, 1);
    }
    // subapertures and GP per powfs
    s->sa.resize(s->powfs.size());
    s->gprow.resize(s->powfs.size());
    s->gp_p.resize(s->powfs.size());
    s->gp_i.resize(s->powfs.size());
    s->gp_x.resize(s->powfs.size());
    for (size_t ip = 0; ip < s->powfs.size(); ip++) {
        const mt_powfs_t &pw = s->powfs[ip];
        if (pw.skip) continue;
        s->sa[ip] = setup_saloc(s->D, s->Din, pw);
        s->gprow[ip] = mkg_gp(s->pg, s->sa[ip], s->D, s->Din, pw.dx);
        // CSC (dsp) view: columns = ploc points
        const SparseRows &r = s->gprow[ip];
        std::vector<long> cnt(r.ncol + 1, 0);
        for (long k = 0; k < (long)r.i.size(); k++) cnt[r.i[k] + 1]++;
        for (long c = 0; c < r.ncol; c++) cnt[c + 1] += cnt[c];
        s->gp_p[ip] = cnt;
        s->gp_i[ip].resize(r.i.size());
        s->gp_x[ip].resize(r.i.size());
        std::vector<long> pos(cnt.begin(), cnt.end() - 1);
        for (long row = 0; row < r.nrow; row++)
            for (long k = r.p[row]; k < r.p[row + 1]; k++) {
                long c = r.i[k];
                s->gp_i[ip][pos[c]] = row;
                s->gp_x[ip][pos[c]] = r.x[k];
                pos[c]++;
            }
    }
    // per-WFS data: NEA weighting, TT projector, neam
    const size_t nwfs = s->wfs.size();
    s->skip.resize(nwfs);
    s->ttr.resize(nwfs);
    s->nsa.resize(nwfs);
    s->dsa.resize(nwfs);
    s->thetax.resize(nwfs);
    s->thetay.resize(nwfs);
    s->hs.resize(nwfs);
    s->neam.assign(nwfs, INFINITY);
    s->saloc.resize(nwfs);
    s->neai.resize(nwfs);
    s->ptt.resize(nwfs);
    for (size_t iw = 0; iw < nwfs; iw++) {
        const int ip = s->wfs[iw].powfs;
        const mt_powfs_t &pw = s->powfs[ip];
        s->skip[iw] = pw.skip ? 1 : 0;
        s->thetax[iw] = s->wfs[iw].thetax;
        s->thetay[iw] = s->wfs[iw].thetay;
        s->hs[iw] = pw.hs;
        s->dsa[iw] = pw.dsa;
        if (pw.skip) {
            s->nsa[iw] = 0;
            continue;
        }
        const SaGeom &sa = s->sa[ip];
        const long nsa = (long)sa.locx.size();
        s->nsa[iw] = nsa;
        s->saloc[iw].resize(2 * nsa);
        std::copy(sa.locx.begin(), sa.locx.end(), s->saloc[iw].begin());
        std::copy(sa.locy.begin(), sa.locy.end(), s->saloc[iw].begin() + nsa);
        // geometric NEA (maos/setup_recon.c:128-152): nea^2 = neasq/area
        const double neasq = std::pow(pw.nearecon / 206265000., 2) / pw.dtrat;
        s->neai[iw].resize(3 * nsa);
        for (long i = 0; i < nsa; i++) {
            const double nea2 = neasq / sa.saa[i];
            s->neai[iw][3 * i + 0] = s->tomoscale / nea2;
            s->neai[iw][3 * i + 1] = s->tomoscale / nea2;
            s->neai[iw][3 * i + 2] = 0;
        }
        // neam (maos/setup_recon.c:172-199)
        const double area_thres = nsa > 4 ? 0.9 : 0;
        double nea2_sum = 0;
        long count = 0;
        for (long i = 0; i < nsa; i++) {
            if (sa.saa[i] > area_thres) {
                nea2_sum += 2 * neasq / sa.saa[i];
                count++;
            }
        }
        s->neam[iw] = std::sqrt(nea2_sum / count / 2) / std::sqrt(s->tomoscale);
        // TT for trs or (split && !lo) (maos/setup_recon_prep.c:867-914);
        // PTT = (TT' W TT)^-1 TT' W (math/blas.c:537-556) which, with W
        // diagonal, is the W-weighted mean of the x and y slopes.
        s->ttr[iw] = (pw.trs || (s->split && !pw.lo)) ? 1 : 0;
        if (s->ttr[iw]) {
            double sx = 0, sy = 0;
            for (long i = 0; i < nsa; i++) {
                sx += s->neai[iw][3 * i];
                sy += s->neai[iw][3 * i + 1];
            }
            s->ptt[iw].assign(2 * 2 * nsa, 0.);
            for (long i = 0; i < nsa; i++) {
                s->ptt[iw][2 * i + 0] = s->neai[iw][3 * i] / sx;             // PTT(0, i)
                s->ptt[iw][2 * (nsa + i) + 1] = s->neai[iw][3 * i + 1] / sy; // PTT(1, nsa+i)
            }
        }
    }
    build_regularisation(s);
    build_fdpcg(s);
}

void fill_views(mt_setup *s) {
    s->xnx.clear();
    s->xny.clear();
    s->xox.clear();
    s->xoy.clear();
    s->xdx.clear();
    for (const Grid &g : s->xg) {
        s->xnx.push_back(g.nx);
        s->xny.push_back(g.ny);
        s->xox.push_back(g.ox);
        s->xoy.push_back(g.oy);
        s->xdx.push_back(g.dx);
    }
    const size_t nwfs = s->wfs.size();
    s->v_saloc.assign(nwfs, nullptr);
    s->v_saa.assign(nwfs, nullptr);
    s->v_gpx.assign(nwfs, nullptr);
    s->v_gpp.assign(nwfs, nullptr);
    s->v_gpi.assign(nwfs, nullptr);
    s->v_neai.assign(nwfs, nullptr);
    s->v_ptt.assign(nwfs, nullptr);
    for (size_t iw = 0; iw < nwfs; iw++) {
        if (s->skip[iw]) continue;
        const int ip = s->wfs[iw].powfs;
        s->v_saloc[iw] = s->saloc[iw].data();
        s->v_saa[iw] = s->sa[ip].saa.data();
        s->v_gpp[iw] = s->gp_p[ip].data();
        s->v_gpi[iw] = s->gp_i[ip].data();
        s->v_gpx[iw] = s->gp_x[ip].data();
        s->v_neai[iw] = s->neai[iw].data();
        s->v_ptt[iw] = s->ttr[iw] ? s->ptt[iw].data() : nullptr;
    }
}

template <typename F>
int guarded(F &&f) {
    try {
        f();
        return MT_OK;
    } catch (const Error &e) {
        set_error(e.what());
        return e.code;
    } catch (const std::bad_alloc &) {
        set_error("out of host memory");
        return MT_ENOMEM;
    } catch (const std::exception &e) {
        set_error(e.what());
        return MT_EINVAL;
    }
}

}  // namespace

extern "C" {

const char *mt_last_error(void) { return mt::last_error_cstr(); }

int mt_setup_create(const mt_parms_t *parms, mt_setup_t **out) {
    if (!out) return MT_EINVAL;
    *out = nullptr;
    std::unique_ptr<mt_setup> s(new (std::nothrow) mt_setup());
    if (!s) return MT_ENOMEM;
    int rc = guarded([&] {
        build(s.get(), parms);
        fill_views(s.get());
    });
    if (rc == MT_OK) *out = s.release();
    return rc;
}

int mt_setup_update_cn2(mt_setup_t *s, const double *wt, double r0z) {
    if (!s || !wt || !(r0z > 0)) return MT_EINVAL;
    return guarded([&] {
        s->wt0.assign(wt, wt + s->npsr);
        s->r0 = r0z * std::pow(std::cos(s->za), 3. / 5.);
        build_regularisation(s);
        update_fdpcg(s);
    });
}

int mt_setup_recon(const mt_setup_t *s, mt_recon_t *r) {
    if (!s || !r) return MT_EINVAL;
    std::memset(r, 0, sizeof(*r));
    r->npsr = s->npsr;
    r->xnx = s->xnx.data();
    r->xny = s->xny.data();
    r->xox = s->xox.data();
    r->xoy = s->xoy.data();
    r->xdx = s->xdx.data();
    r->ht = s->ht.data();
    r->pnx = s->pg.nx;
    r->pny = s->pg.ny;
    r->pox = s->pg.ox;
    r->poy = s->pg.oy;
    r->pdx = s->pg.dx;
    r->nwfs = (int)s->wfs.size();
    r->thetax = s->thetax.data();
    r->thetay = s->thetay.data();
    r->hs = s->hs.data();
    r->skip = s->skip.data();
    r->nsa = s->nsa.data();
    r->dsa = s->dsa.data();
    r->saloc = s->v_saloc.data();
    r->saa = s->v_saa.data();
    r->gp_p = s->v_gpp.data();
    r->gp_i = s->v_gpi.data();
    r->gp_x = s->v_gpx.data();
    r->pos = s->pos;
    r->neai = s->v_neai.data();
    r->trs = s->ttr.data();
    r->ptt = s->v_ptt.data();
    r->ptt_lhs = s->ptt_lhs;
    r->l2c = s->l2c.data();
    r->precond = s->precond == 1 && s->fd.nb > 0;
    r->fd_bs = s->fd.bs;
    r->fd_nb = s->fd.nb;
    r->fd_nbhalf = s->fd.nbhalf;
    r->fd_scale = s->fd.scale;
    r->fd_Minv = s->fd.Minv.empty() ? nullptr : s->fd.Minv.data();
    r->fd_perm = s->fd.perm.empty() ? nullptr : s->fd.perm.data();
    r->fd_permhf = s->fd.permhf.empty() ? nullptr : s->fd.permhf.data();
    r->tomoscale = s->tomoscale;
    return MT_OK;
}

int mt_setup_fdpcg_mhat(const mt_setup_t *s, const double **mhat) {
    if (!s || !mhat) return MT_EINVAL;
    *mhat = s->fd.Mhat.empty() ? nullptr : s->fd.Mhat.data();
    return MT_OK;
}

int mt_setup_sizes(const mt_setup_t *s, long *nx, long *ng) {
    if (!s) return MT_EINVAL;
    long a = 0, b = 0;
    for (const Grid &g : s->xg) a += g.nx * g.ny;
    for (long n : s->nsa) b += 2 * n;
    if (nx) *nx = a;
    if (ng) *ng = b;
    return MT_OK;
}

void mt_setup_free(mt_setup_t *s) { delete s; }

int mt_setup_load_gp(mt_setup_t *s, const mt_bin_t *gp) {
    if (!s || !gp) return MT_EINVAL;
    return guarded([&] {
        // setup_recon_GP with load.GP (maos/setup_recon_prep.c:485-505): a
        // cell of one sparse matrix per WFS or per powfs, each 2 nsa x nploc
        int cell = 0;
        long nx = 0, ny = 0;
        mt_bin_info(gp, &cell, &nx, &ny);
        const long nel = nx * ny, nwfs = (long)s->wfs.size(), npowfs = (long)s->powfs.size();
        if (!cell || (nel != nwfs && nel != npowfs)) throw Error(MT_EINVAL, "load.GP: a cell of nwfs or npowfs sparse matrices");
        const long nploc = s->pg.nx * s->pg.ny;
        struct Csc {
            const long *p = nullptr, *i = nullptr;
            const double *x = nullptr;
            long nz = 0;
        };
        auto get = [&](long k, long nsa) {
            const mt_bin_t *e = mt_bin_child(gp, k);
            if (!e || mt_bin_type(e) != MT_BIN_SPARSE) throw Error(MT_EINVAL, "load.GP: element " + std::to_string(k) + " is not a sparse double matrix");
            long ex = 0, ey = 0;
            mt_bin_info(e, nullptr, &ex, &ey);
            if (ex != 2 * nsa || ey != nploc)
                throw Error(MT_EINVAL, "load.GP: element " + std::to_string(k) + " is " + std::to_string(ex) + " x " +
                                           std::to_string(ey) + ", need " + std::to_string(2 * nsa) + " x " +
                                           std::to_string(nploc));
            Csc c;
            mt_bin_sparse(e, &c.p, &c.i, &c.x, &c.nz);
            return c;
        };
        for (size_t ip = 0; ip < s->powfs.size(); ip++) {
            if (s->powfs[ip].skip) continue;
            const long nsa = (long)s->sa[ip].locx.size();
            Csc c;
            if (nel == npowfs) {
                c = get((long)ip, nsa);
            } else {
                // per WFS: the plan keeps one GP per powfs, so the WFS of a
                // powfs must carry the same matrix
                bool first = true;
                for (long w = 0; w < nwfs; w++) {
                    if (s->wfs[w].powfs != (int)ip) continue;
                    const Csc cw = get(w, nsa);
                    if (first) {
                        c = cw;
                        first = false;
                    } else if (cw.nz != c.nz || !std::equal(cw.p, cw.p + nploc + 1, c.p) ||
                               !std::equal(cw.i, cw.i + cw.nz, c.i) || !std::equal(cw.x, cw.x + cw.nz, c.x)) {
                        throw Error(MT_EINVAL, "load.GP: WFS of one powfs with different GP are not supported");
                    }
                }
                if (first) continue;
            }
            s->gp_p[ip].assign(c.p, c.p + nploc + 1);
            s->gp_i[ip].assign(c.i, c.i + c.nz);
            s->gp_x[ip].assign(c.x, c.x + c.nz);
            // row form (gradient rows) for the rest of the setup
            SparseRows r;
            r.nrow = 2 * nsa;
            r.ncol = nploc;
            r.p.assign(r.nrow + 1, 0);
            for (long k = 0; k < c.nz; k++) r.p[c.i[k] + 1]++;
            for (long q = 0; q < r.nrow; q++) r.p[q + 1] += r.p[q];
            r.i.resize(c.nz);
            r.x.resize(c.nz);
            std::vector<long> pos(r.p.begin(), r.p.end() - 1);
            for (long col = 0; col < nploc; col++)
                for (long k = c.p[col]; k < c.p[col + 1]; k++) {
                    r.i[pos[c.i[k]]] = col;
                    r.x[pos[c.i[k]]] = c.x[k];
                    pos[c.i[k]]++;
                }
            s->gprow[ip] = std::move(r);
        }
        fill_views(s);
    });
}

int mt_fdpcg_prepare(const mt_fdpcg_in_t *in, mt_fdpcg_out_t *out) {
    if (!in || !out || in->nps <= 0 || !in->nx || !in->ny || !in->os || !in->ht || !in->l2cf ||
        in->nsa <= 0 || !in->salocx || !in->salocy || !in->saa || in->nwfs <= 0 || !in->thetax ||
        !in->thetay || !in->neam)
        return MT_EINVAL;
    std::memset(out, 0, sizeof(*out));
    return guarded([&] {
        SaGeom sa;
        sa.dsa = in->dsa;
        sa.locx.assign(in->salocx, in->salocx + in->nsa);
        sa.locy.assign(in->salocy, in->salocy + in->nsa);
        sa.saa.assign(in->saa, in->saa + in->nsa);
        FdpcgIn f;
        f.nps = in->nps;
        f.nx.assign(in->nx, in->nx + in->nps);
        f.ny.assign(in->ny, in->ny + in->nps);
        f.os.assign(in->os, in->os + in->nps);
        f.pos = in->pos;
        f.dxp = in->dxp;
        f.ht.assign(in->ht, in->ht + in->nps);
        f.l2cf.assign(in->l2cf, in->l2cf + in->nps);
        f.sa = &sa;
        f.thetax.assign(in->thetax, in->thetax + in->nwfs);
        f.thetay.assign(in->thetay, in->thetay + in->nwfs);
        f.neam.assign(in->neam, in->neam + in->nwfs);
        FdpcgOut o = fdpcg_prepare(f);
        auto dup = [](const auto &v) {
            using T = typename std::decay<decltype(v)>::type::value_type;
            T *p = (T *)std::malloc(sizeof(T) * (v.size() ? v.size() : 1));
            if (!p) throw std::bad_alloc();
            std::copy(v.begin(), v.end(), p);
            return p;
        };
        out->bs = o.bs;
        out->nb = o.nb;
        out->nbhalf = o.nbhalf;
        out->scale = o.scale;
        out->Mhat = dup(o.Mhat);
        out->Minv = dup(o.Minv);
        out->perm = dup(o.perm);
        out->permhf = dup(o.permhf);
    });
}

void mt_fdpcg_free(mt_fdpcg_out_t *out) {
    if (!out) return;
    std::free(out->Mhat);
    std::free(out->Minv);
    std::free(out->perm);
    std::free(out->permhf);
    std::memset(out, 0, sizeof(*out));
}

}  // extern "C"
