// FDPCG preconditioner construction, FP64 (maos/fdpcg.c:70-580).
//
// The reference assembles Mhat as sparse matrices over the whole stacked
// spectrum:
//   Mhat = diag(invpsd) + sum_w (1/neam_w^2) propx_w^H Mmid propx_w,
//   Mmid = sum_{g=gx,gy} diag(conj g) sel^H sel diag(g)
// then permutes it into bs x bs blocks (fdpcg_perm, shift=0) and inverts each
// block with a truncated SVD.  Every nonzero of Mhat lies inside one block by
// construction (cspblockextract errors otherwise), so the same entries are
// computed here block by block: each block member (layer l, frequency f)
// receives from the pupil frequencies that propx maps onto it, and Mmid couples
// pupil frequencies only through the few separations kept in sel.  The sums,
// scalings, drop tolerances, symmetrisation and thresholds are the reference's.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <exception>
#include <mutex>
#include <thread>

#include "setup_internal.hpp"

namespace mt {

namespace {
// f(i) for i in [0, n), indices independent, on up to 16 host threads (the
// host share of one GPU); the first exception is rethrown after all join.
// Each index computes what it computed serially, so results are identical.
template <typename F>
void parallel_for(long n, F f) {
    const unsigned nt = std::max(1u, std::min(std::thread::hardware_concurrency(), 16u));
    if (nt == 1 || n < 2) {
        for (long i = 0; i < n; i++) f(i);
        return;
    }
    std::atomic<long> next(0);
    std::exception_ptr err;
    std::mutex mu;
    auto work = [&] {
        try {
            for (long i; (i = next.fetch_add(1)) < n;) f(i);
        } catch (...) {
            std::lock_guard<std::mutex> lk(mu);
            if (!err) err = std::current_exception();
            next = n;
        }
    };
    std::vector<std::thread> th;
    for (unsigned t = 1; t < std::min<long>(nt, n); t++) th.emplace_back(work);
    work();
    for (auto &t : th) t.join();
    if (err) std::rethrow_exception(err);
}
}  // namespace

void fftshift2(std::vector<cplx> &a, long nx, long ny) {
    // math/mat.c:389-420 (even sizes): swap quadrants.
    if ((nx & 1) || (ny & 1)) throw Error(MT_EINVAL, "fftshift2 needs even sizes");
    const long nx2 = nx / 2, ny2 = ny / 2;
    for (long iy = 0; iy < ny2; iy++) {
        for (long ix = 0; ix < nx; ix++) {
            long jx = (ix + nx2) % nx, jy = iy + ny2;
            std::swap(a[ix + iy * nx], a[jx + jy * nx]);
        }
    }
}

void dft2(std::vector<cplx> &a, long nx, long ny, int sign) {
    // Direct separable DFT, X[k] = sum_n x[n] exp(sign*2*pi*i*k*n/N), twiddles
    // in long double from exact integer phase (k*n mod N).
    std::vector<std::complex<long double>> tw;
    std::vector<std::complex<long double>> tmp;
    auto pass = [&](long n, long howmany, long stride, long dist) {
        tw.resize(n);
        const long double pi2 = 6.283185307179586476925286766559L;
        for (long k = 0; k < n; k++)
            tw[k] = std::complex<long double>(std::cos(pi2 * k / n), sign * std::sin(pi2 * k / n));
        tmp.resize(n);
        for (long h = 0; h < howmany; h++) {
            cplx *v = &a[h * dist];
            for (long k = 0; k < n; k++) {
                std::complex<long double> s = 0;
                for (long m = 0; m < n; m++) {
                    const std::complex<long double> x(v[m * stride].real(), v[m * stride].imag());
                    s += x * tw[(k * m) % n];
                }
                tmp[k] = s;
            }
            for (long k = 0; k < n; k++)
                v[k * stride] = cplx((double)tmp[k].real(), (double)tmp[k].imag());
        }
    };
    pass(nx, ny, 1, nx);   // along x for every row
    pass(ny, nx, nx, 1);   // along y for every column
}


void herm_pinv(const cplx *A0, cplx *out, int n, double thres) {
    // math/blas.c:456-500 csvd_pow(A, -1, thres) for a Hermitian block: the
    // singular values are the |eigenvalues|, and the SVD pseudo inverse is
    // sum_k v_k v_k^H / lambda_k over the eigenpairs with |lambda_k| >
    // thres * max |lambda|.  Eigenpairs by cyclic complex Jacobi: on the
    // (p, q) plane, J = diag(1, e^{-i phi}) R zeroes a_pq = r e^{i phi},
    // where R = [[c, s], [-s, c]] is the real Jacobi rotation of the real
    // symmetric [[a_pp, r], [r, a_qq]].
    constexpr int NMAX = 16;
    if (n < 1 || n > NMAX) throw Error(MT_EINVAL, "fdpcg: block size out of range");
    cplx A[NMAX * NMAX], V[NMAX * NMAX];
    for (int i = 0; i < n * n; i++) {
        A[i] = A0[i];
        V[i] = 0;
    }
    for (int i = 0; i < n; i++) V[i + i * n] = 1;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0, tot = 0;
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                const double v = std::norm(A[i + j * n]);
                tot += v;
                if (i != j) off += v;
            }
        if (off <= 1e-30 * tot || off == 0) break;
        for (int p = 0; p < n - 1; p++) {
            for (int q = p + 1; q < n; q++) {
                const cplx apq = A[p + q * n];
                const double r = std::abs(apq);
                if (r == 0) continue;
                const cplx eph = std::conj(apq) / r;  // e^{-i phi}
                const double app = A[p + p * n].real(), aqq = A[q + q * n].real();
                const double theta = (aqq - app) / (2 * r);
                const double t = (theta >= 0 ? 1. : -1.) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
                const double c = 1 / std::sqrt(t * t + 1), s = t * c;
                const cplx se = s * eph, ce = c * eph;  // J = [[c, s], [-se, ce]]
                for (int k = 0; k < n; k++) {           // A J
                    const cplx akp = A[k + p * n], akq = A[k + q * n];
                    A[k + p * n] = c * akp - se * akq;
                    A[k + q * n] = s * akp + ce * akq;
                }
                for (int k = 0; k < n; k++) {  // J^H A
                    const cplx apk = A[p + k * n], aqk = A[q + k * n];
                    A[p + k * n] = c * apk - std::conj(se) * aqk;
                    A[q + k * n] = s * apk + std::conj(ce) * aqk;
                }
                A[p + q * n] = A[q + p * n] = 0;
                A[p + p * n] = app - t * r;
                A[q + q * n] = aqq + t * r;
                for (int k = 0; k < n; k++) {  // V J
                    const cplx vkp = V[k + p * n], vkq = V[k + q * n];
                    V[k + p * n] = c * vkp - se * vkq;
                    V[k + q * n] = s * vkp + ce * vkq;
                }
            }
        }
    }
    double maxeig = 0;
    for (int k = 0; k < n; k++) maxeig = std::max(maxeig, std::fabs(A[k + k * n].real()));
    const double thres0 = std::fabs(thres) * maxeig;
    for (int i = 0; i < n * n; i++) out[i] = 0;
    for (int k = 0; k < n; k++) {
        const double w = A[k + k * n].real();
        if (!(std::fabs(w) > thres0)) continue;
        const double inv = 1. / w;
        for (int j = 0; j < n; j++) {
            const cplx vj = std::conj(V[j + k * n]) * inv;
            for (int i = 0; i < n; i++) out[i + j * n] += V[i + k * n] * vj;
        }
    }
}

FdpcgOut fdpcg_prepare(const FdpcgIn &in) {
    FdpcgOut out = fdpcg_geometry(in);
    fdpcg_finish(out, in);
    return out;
}

FdpcgOut fdpcg_geometry(const FdpcgIn &in) {
    const int nps = in.nps;
    const long *nx = in.nx.data(), *ny = in.ny.data(), *os = in.os.data();
    const long pos = in.pos;
    const long nxp = nx[0] / os[0] * pos;  // maos/fdpcg.c:351-352
    const long nyp = ny[0] / os[0] * pos;
    const double dxp = in.dxp;
    long nxtot = 0;
    int needscale = 0;
    std::vector<long> noff(nps);
    for (int ips = 0; ips < nps; ips++) {
        if (os[ips] != os[0]) needscale = 1;
        noff[ips] = nxtot;
        nxtot += nx[ips] * ny[ips];
    }
    if ((nxp & 1) || (nyp & 1)) throw Error(MT_EINVAL, "fdpcg: pupil FFT grid must be even");

    // --- fdpcg_saselect (maos/fdpcg.c:70-103) -------------------------------
    struct Sep {
        long x, y;
        cplx v;
    };
    std::vector<Sep> seps;
    {
        std::vector<cplx> xsel(nxp * nyp, cplx(0, 0));
        const double dx1 = 1. / dxp;
        const long offx = nxp / 2, offy = nyp / 2;
        const SaGeom &sa = *in.sa;
        for (size_t isa = 0; isa < sa.locx.size(); isa++) {
            if (sa.saa[isa] > 0.9) {
                long ix = (long)(sa.locx[isa] * dx1 + offx);  // C truncation
                long iy = (long)(sa.locy[isa] * dx1 + offy);
                if (ix < 0 || ix >= nxp || iy < 0 || iy >= nyp)
                    throw Error(MT_EINVAL, "fdpcg: subaperture outside pupil FFT grid");
                xsel[ix + iy * nxp] = 1. / (double)(nxp * nyp);
            }
        }
        fftshift2(xsel, nxp, nyp);
        dft2(xsel, nxp, nyp, -1);
        fftshift2(xsel, nxp, nyp);
        const double xselc = xsel[nxp / 2 + (nyp / 2) * nxp].real();
        // The reference zeroes |X| < X(0).  The lattice harmonics of the
        // subaperture grid equal X(0) in exact arithmetic, so whether FFTW's
        // rounding keeps them is a coin toss there; they are kept here by
        // comparing with a 1e-9 relative slack.
        for (long iy = 0; iy < nyp; iy++)
            for (long ix = 0; ix < nxp; ix++) {
                const cplx v = xsel[ix + iy * nxp];
                if (std::abs(v) >= xselc * (1. - 1e-9) && std::abs(v) > 0)
                    seps.push_back({ix, iy, v});
            }
    }
    // S(d) = (sel^H sel) entry at separation d = sum_{a,b: s_b - s_a = d} conj(v_a) v_b
    auto selsel = [&](long dx_, long dy_) {
        cplx s = 0;
        for (const Sep &a : seps)
            for (const Sep &b : seps) {
                long ddx = ((b.x - a.x) % nxp + nxp) % nxp;
                long ddy = ((b.y - a.y) % nyp + nyp) % nyp;
                if (ddx == dx_ && ddy == dy_) s += std::conj(a.v) * b.v;
            }
        return s;
    };
    // Precompute S over the distinct separations that occur.
    struct SS {
        long x, y;
        cplx v;
    };
    std::vector<SS> sdiff;
    for (const Sep &a : seps)
        for (const Sep &b : seps) {
            long ddx = ((b.x - a.x) % nxp + nxp) % nxp;
            long ddy = ((b.y - a.y) % nyp + nyp) % nyp;
            bool have = false;
            for (const SS &s : sdiff)
                if (s.x == ddx && s.y == ddy) have = true;
            if (!have) sdiff.push_back({ddx, ddy, selsel(ddx, ddy)});
        }

    // --- fdpcg_g (maos/fdpcg.c:198-246) -------------------------------------
    std::vector<cplx> gx(nxp * nyp), gy(nxp * nyp);
    {
        const double dsa = in.sa->dsa;
        const long osg = std::lround(dsa / dxp);
        if (std::fabs(dsa - dxp * osg) > 1.e-10) throw Error(MT_EINVAL, "fdpcg: dsa not a multiple of dx");
        std::vector<double> wt(osg + 1), st(osg + 1);
        wt[osg] = wt[0] = 0.5 / (double)osg / dsa;
        for (long i = 1; i < osg; i++) wt[i] = 1. / (double)osg / dsa;
        for (long i = 0; i < osg + 1; i++) st[i] = i * dxp;
        const cplx cf(0, 2 * M_PI);
        const double dfy = 1 / (nyp * dxp), dfx = 1 / (nxp * dxp);
        for (long iy = 0; iy < nyp; iy++) {
            const double fy = (double)(iy - nyp / 2) * dfy;
            for (long ix = 0; ix < nxp; ix++) {
                const double fx = (double)(ix - nxp / 2) * dfx;
                cplx tx = 0, ty = 0;
                for (long i = 0; i < osg + 1; i++) {
                    tx += wt[i] * (std::exp(cf * (fx * dsa + fy * st[i])) - std::exp(cf * (fy * st[i])));
                    ty += wt[i] * (std::exp(cf * (fy * dsa + fx * st[i])) - std::exp(cf * (fx * st[i])));
                }
                gx[ix + iy * nxp] = tx;
                gy[ix + iy * nxp] = ty;
            }
        }
    }
    // S by separation: index into sdiff, or -1
    std::vector<int> sat(nxp * nyp, -1);
    for (size_t i = 0; i < sdiff.size(); i++) sat[sdiff[i].x + sdiff[i].y * nxp] = (int)i;
    auto mmid = [&](long p, long q) -> cplx {
        // Mmid(p,q) = sum_g conj(g_p) S(p-q) g_q, pupil indices in shifted order
        long ddx = ((p % nxp - q % nxp) % nxp + nxp) % nxp;
        long ddy = ((p / nxp - q / nxp) % nyp + nyp) % nyp;
        const int i = sat[ddx + ddy * nxp];
        if (i < 0) return 0.;
        return sdiff[i].v * (std::conj(gx[p]) * gx[q] + std::conj(gy[p]) * gy[q]);
    };

    // --- block membership (fdpcg_perm, maos/fdpcg.c:118-193) ----------------
    int bs = 0;
    for (int ips = 0; ips < nps; ips++) bs += (int)(os[ips] * os[ips]);
    const long adimx = nx[0] / os[0], adimy = ny[0] / os[0];
    const long adimx2 = adimx / 2;
    const long nb = adimx * adimy;
    const long nbhalf = (adimx2 + 1) * adimy;
    FdpcgOut out;
    out.bs = bs;
    out.nb = nb;
    out.nbhalf = nbhalf;
    out.scale = needscale;
    out.diag_at.resize(nb * bs);
    out.lsize.resize(nb * bs);
    out.perm.resize(nb * bs);
    out.permhf.resize(nbhalf * bs);
    struct Member {
        int l;
        long jxs, jys;   // index in the fftshifted layer spectrum
    };
    std::vector<Member> members(nb * bs);
    {
        long count = 0, counthf = 0;
        for (int ixh = 0; ixh < 2; ixh++) {
            long ix0 = ixh == 0 ? -adimx2 : 1;
            long ix1 = ixh == 0 ? 1 : adimx2;
            for (long iy = -adimy / 2; iy < adimy / 2; iy++) {
                for (long ix = ix0; ix < ix1; ix++) {
                    for (int ips = 0; ips < nps; ips++) {
                        const long nx2 = nx[ips] / 2, ny2 = ny[ips] / 2;
                        for (long jos = 0; jos < os[ips]; jos++) {
                            const long fy = iy + adimy * jos;
                            long jys = fy + ny2;
                            if (jys >= ny[ips]) jys -= ny[ips];
                            long jy = fy < 0 ? fy + ny[ips] : fy;
                            for (long ios = 0; ios < os[ips]; ios++) {
                                const long fx = ix + adimx * ios;
                                long jxs = fx + nx2;
                                if (jxs >= nx[ips]) jxs -= nx[ips];
                                long jx = fx < 0 ? fx + nx[ips] : fx;
                                members[count] = {ips, jxs, jys};
                                out.perm[count] = noff[ips] + jx + jy * nx[ips];
                                if (ixh == 0) {
                                    // half spectrum (shift=1, half=1)
                                    long jxh = jx, ratio = 1;
                                    if (jxh > nx2) {
                                        ratio = -1;
                                        jxh = nx[ips] - jxh;
                                    }
                                    long jyh = ratio < 0 ? (jy == 0 ? 0 : ny[ips] - jy) : jy;
                                    out.permhf[counthf++] = ratio * (noff[ips] + jxh + jyh * nx[ips]);
                                }
                                count++;
                            }
                        }
                    }
                }
            }
        }
        if (count != nb * bs || counthf != nbhalf * bs) throw Error(MT_EINVAL, "fdpcg: perm count mismatch");
    }

    // --- Mhat blocks --------------------------------------------------------
    const double dkx = 1. / (nxp * dxp), dky = 1. / (nyp * dxp);
    const long nxp2 = nxp / 2, nyp2 = nyp / 2;
    const int nwfs = (int)in.thetax.size();
    out.MhatW.assign(nb * bs * bs * 2, 0.);
    // pupil frequencies mapped onto a member (propx rows), maos/fdpcg.c:252-326
    struct PF {
        long p;
        long px, py;
    };
    parallel_for(nb, [&](long ib) {
        std::vector<cplx> blk(bs * bs);
        std::vector<std::vector<PF>> pre(bs);
        std::vector<std::vector<cplx>> fac(bs);
        for (int m = 0; m < bs; m++) {
            const Member &mb = members[ib * bs + m];
            const long n1 = nx[mb.l], n2 = ny[mb.l];
            const long vx = ((mb.jxs - n1 / 2 + nxp2) % n1 + n1) % n1;
            const long vy = ((mb.jys - n2 / 2 + nyp2) % n2 + n2) % n2;
            pre[m].clear();
            for (long px = vx; px < nxp; px += n1)
                for (long py = vy; py < nyp; py += n2) pre[m].push_back({px + py * nxp, px, py});
        }
        std::fill(blk.begin(), blk.end(), cplx(0, 0));
        for (int iw = 0; iw < nwfs; iw++) {
            const double sc = 1. / (in.neam[iw] * in.neam[iw]);
            for (int m = 0; m < bs; m++) {
                const Member &mb = members[ib * bs + m];
                const double dispx = in.ht[mb.l] * in.thetax[iw];
                const double dispy = in.ht[mb.l] * in.thetay[iw];
                const long n1 = nx[mb.l], n2 = ny[mb.l];
                const double fx = (mb.jxs - n1 / 2) * dkx, fy = (mb.jys - n2 / 2) * dky;
                const cplx shift = std::exp(cplx(0, 2 * M_PI) * (fx * dispx + fy * dispy));
                fac[m].clear();
                for (const PF &pf : pre[m]) {
                    cplx v;
                    switch (nxp / n1) {
                        case 0:
                        case 1:
                            v = shift;
                            break;
                        case 2: {
                            const double fxg = (pf.px - nxp2) * dkx, fyg = (pf.py - nyp2) * dky;
                            v = shift * (1 + std::cos(2 * M_PI * (fxg * dxp))) *
                                (1 + std::cos(2 * M_PI * (fyg * dxp))) * 0.5;
                        } break;
                        default:
                            throw Error(MT_EINVAL, "fdpcg_prop: invalid layer/pupil size ratio");
                    }
                    fac[m].push_back(v);
                }
            }
            for (int b = 0; b < bs; b++)
                for (int a = 0; a < bs; a++) {
                    cplx s = 0;
                    for (size_t ia = 0; ia < pre[a].size(); ia++)
                        for (size_t ibb = 0; ibb < pre[b].size(); ibb++) {
                            const cplx mm = mmid(pre[a][ia].p, pre[b][ibb].p);
                            if (mm == cplx(0, 0)) continue;
                            s += std::conj(fac[a][ia]) * mm * fac[b][ibb];
                        }
                    blk[a + b * bs] += s * sc;
                }
        }
        for (int m = 0; m < bs; m++) {
            const Member &mb = members[ib * bs + m];
            out.diag_at[ib * bs + m] = noff[mb.l] + mb.jxs + mb.jys * nx[mb.l];
            out.lsize[ib * bs + m] = (double)(nx[mb.l] * ny[mb.l]);
        }
        std::memcpy(&out.MhatW[ib * bs * bs * 2], blk.data(), sizeof(cplx) * bs * bs);
    });
    return out;
}

// The regulariser diagonal (invpsd) added to the WFS sums, then the drop
// tolerances, symmetrisation, scaling and block pseudo inverses.
void fdpcg_finish(FdpcgOut &out, const FdpcgIn &in) {
    const int nps = in.nps;
    const long *nx = in.nx.data(), *ny = in.ny.data();
    const int bs = out.bs;
    const long nb = out.nb;
    const int needscale = out.scale;
    long nxtot = 0;
    std::vector<long> noff(nps);
    for (int ips = 0; ips < nps; ips++) {
        noff[ips] = nxtot;
        nxtot += nx[ips] * ny[ips];
    }
    // --- invpsd (maos/fdpcg.c:373-407) ---------------------------------------
    std::vector<double> invpsd(nxtot);
    parallel_for(nps, [&](long ips) {
        const long n1 = nx[ips], n2 = ny[ips];
        // FFT of the first column of L2'L2, fftshifted.  L2 (lib/laplacian.c:
        // 70-97, scaled by sqrt(cxxscale*TOMOSCALE) into c) is the periodic
        // stencil -c at a node and c/4 at its four neighbours, so it is
        // diagonal in the DFT basis with eigenvalue c (-1 + (cos ax + cos ay)/2),
        // ax = 2 pi kx / n1, ay = 2 pi ky / n2; L2'L2 = L2^2 has its square,
        // real and even, so the FFT of that column is exactly this (computed
        // in long double, rounded once).
        const long double c = in.l2cf[ips], pi2 = 6.283185307179586476925286766559L;
        std::vector<long double> cx(n1), cy(n2);
        for (long k = 0; k < n1; k++) cx[k] = std::cos(pi2 * k / n1);
        for (long k = 0; k < n2; k++) cy[k] = std::cos(pi2 * k / n2);
        std::vector<cplx> psd(n1 * n2);
        for (long ky = 0; ky < n2; ky++)
            for (long kx = 0; kx < n1; kx++) {
                const long double e = c * (-1 + 0.5L * (cx[kx] + cy[ky]));
                psd[kx + ky * n1] = cplx((double)(e * e), 0.);
            }
        fftshift2(psd, n1, n2);
        const double eps = 2.220446049250313e-16;
        double mx = 0;
        for (const cplx &v : psd) mx = std::max(mx, std::abs(v));
        mx *= std::sqrt(eps);
        for (long i = 0; i < n1 * n2; i++) invpsd[noff[ips] + i] = psd[i].real() + mx;
    });

    out.Mhat = out.MhatW;
    cplx *M = reinterpret_cast<cplx *>(out.Mhat.data());
    for (long i = 0; i < nb * bs; i++) M[(i / bs) * bs * bs + (i % bs) * (bs + 1)] += invpsd[out.diag_at[i]];
    const long ntot = nb * bs * bs;
    auto droptol = [&]() {  // math/sp.c:836-843 with thres = EPS
        double mx = 0;
        for (long i = 0; i < ntot; i++) mx = std::max(mx, std::abs(M[i]));
        const double tol = mx * EPS_SYS;
        for (long i = 0; i < ntot; i++)
            if (!(std::abs(M[i]) > tol)) M[i] = 0;
    };
    droptol();
    // cspsym (math/sp.c:900-907): (M + M^H)/2, drop again
    for (long ib = 0; ib < nb; ib++) {
        cplx *B = M + ib * bs * bs;
        for (int j = 0; j < bs; j++)
            for (int i = j; i < bs; i++) {
                const cplx a = B[i + j * bs], b = B[j + i * bs];
                B[i + j * bs] = 0.5 * (a + std::conj(b));
                B[j + i * bs] = std::conj(B[i + j * bs]);
            }
    }
    droptol();
    if (!needscale) {
        // scale by N per column so the unnormalized FFT pair needs no scaling
        for (long ib = 0; ib < nb; ib++)
            for (int b = 0; b < bs; b++) {
                const double s = out.lsize[ib * bs + b];
                for (int a = 0; a < bs; a++) M[ib * bs * bs + a + b * bs] *= s;
            }
    }
    // --- block pseudo inverse (maos/fdpcg.c:555-569) ------------------------
    out.Minv.assign(nb * bs * bs * 2, 0.);
    cplx *Mi = reinterpret_cast<cplx *>(out.Minv.data());
    parallel_for(nb, [&](long ib) { herm_pinv(M + ib * bs * bs, Mi + ib * bs * bs, bs, 1e-7); });
}

}  // namespace mt
