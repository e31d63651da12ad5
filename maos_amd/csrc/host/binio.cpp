// MAOS .bin files: double matrices and cells of them, as MAOS reads and
// writes them, so that gradients and solutions saved by MAOS (or by this
// library) pass through unchanged.
//
// Layout (sys/bin.c:513-558 read_bin_header, :571-590 write_bin_headerstr,
// :794-811 write_header; math/cell.c:227-270 writedata_by_id):
//   [header string blocks]  M_HEADER(u32) n(u64) text padded with zeros to
//                           n bytes (n = strlen + 1 rounded up to 8) n(u64)
//                           M_HEADER(u32); several blocks concatenate
//   [M_SKIP(u32)]           written before every array; any word with the
//                           M_SKIP bits set is skipped when reading
//   magic(u32) nx(u64) ny(u64), then for M_DBL nx*ny doubles (column-major),
//   for MCC_ANY the nx*ny elements, each with its own header; an empty
//   element, and a cell whose elements are all empty, is written as
//   MCC_ANY 0 0.  Little-endian, as MAOS writes on x86.
// Other element types (magics of sys/bin.h:34-51): M_CMP complex double
// (re, im pairs), M_FLT / M_ZMP single-precision real / complex, M_INT64 /
// M_INT32 integers, and sparse matrices (math/spbin.c:32-89): after the
// header, nzmax(u64), then the ny+1 column pointers and the nzmax row
// indices (64-bit for M_DSP64 / M_CSP64, 32-bit for M_DSP32 / M_CSP32), then
// the nzmax values (double, or complex double for M_CSP*); an empty sparse
// matrix (nx or ny zero) has no nzmax.  Sparse matrices are written with
// 64-bit indices, the form MAOS writes on a 64-bit host.
// Files may be gzip-compressed (sys/bin.c:250-256 detects the gzip magic);
// zlib's gzread reads both forms, and a path ending in ".gz" is written
// compressed.
#include <zlib.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "setup_internal.hpp"

using namespace mt;

namespace {
constexpr uint32_t M_CSP64 = 0x6400, M_DSP64 = 0x6401, M_DBL = 0x6402, M_INT64 = 0x6403, M_CMP = 0x6404,
                   M_INT32 = 0x6405, M_CSP32 = 0x6406, M_DSP32 = 0x6407, M_FLT = 0x6408, M_ZMP = 0x6409;
constexpr uint32_t MCC_ANY = 0x6421;
constexpr uint32_t M_HEADER = 0x6500;
constexpr uint32_t M_SKIP = 0x6600;
}  // namespace

struct mt_bin {
    uint32_t magic = M_DBL;  // one of the M_* above (sparse: M_DSP64 / M_CSP64) or MCC_ANY
    uint64_t nx = 0, ny = 0;
    std::string header;
    std::vector<double> data;                    // values (complex: re, im pairs; float: widened)
    std::vector<long> idata;                     // integer arrays
    std::vector<long> sp_p, sp_i;                // sparse: column pointers, row indices
    std::vector<std::unique_ptr<mt_bin>> child;  // MCC_ANY (null: empty element)
    bool is_sparse() const { return magic == M_DSP64 || magic == M_CSP64; }
    bool is_complex() const { return magic == M_CMP || magic == M_ZMP || magic == M_CSP64; }
};

namespace {

struct GzIn {
    gzFile f;
    std::string path;
    void get(void *p, size_t n) {  // in chunks: gzread counts in unsigned int
        char *c = static_cast<char *>(p);
        while (n) {
            const unsigned k = (unsigned)std::min<size_t>(n, 1u << 30);
            const int r = gzread(f, c, k);
            if (r < 0 || (unsigned)r != k) throw Error(MT_EINVAL, "truncated or unreadable .bin file " + path);
            c += k;
            n -= k;
        }
    }
    bool try_get(void *p, size_t n) {  // false at a clean end of file
        const int r = gzread(f, p, (unsigned)n);
        if (r == 0) return false;
        if (r < 0 || (size_t)r != n) throw Error(MT_EINVAL, "truncated or unreadable .bin file " + path);
        return true;
    }
};
struct GzOut {
    gzFile f;
    std::string path;
    void put(const void *p, size_t n) {
        const char *c = static_cast<const char *>(p);
        while (n) {
            const unsigned k = (unsigned)std::min<size_t>(n, 1u << 30);
            if (gzwrite(f, c, k) != (int)k) throw Error(MT_EINVAL, "write failed: " + path);
            c += k;
            n -= k;
        }
    }
};

// One array or cell at the current position (with the header strings and
// skip words before it).  Returns null at a clean end of file when allowed.
std::unique_ptr<mt_bin> read_node(GzIn &in, bool eof_ok) {
    std::unique_ptr<mt_bin> b(new mt_bin());
    uint32_t magic;
    for (bool first = true;; first = false) {
        if (!in.try_get(&magic, 4)) {
            if (first && eof_ok) return nullptr;
            throw Error(MT_EINVAL, "truncated .bin file " + in.path);
        }
        if ((magic & M_SKIP) == M_SKIP) continue;
        if (magic != M_HEADER) break;
        uint64_t n, n2;
        uint32_t m2;
        in.get(&n, 8);
        if (n > (1u << 24)) throw Error(MT_EINVAL, "implausible header length in " + in.path);
        std::string s(n, '\0');
        in.get(&s[0], n);
        in.get(&n2, 8);
        in.get(&m2, 4);
        if (n2 != n || m2 != M_HEADER) throw Error(MT_EINVAL, "header string verification failed in " + in.path);
        b->header += s.substr(0, s.find('\0'));
    }
    in.get(&b->nx, 8);
    in.get(&b->ny, 8);
    const uint64_t n = b->nx * b->ny;
    if (b->nx > (1ull << 32) || b->ny > (1ull << 32) || n > (1ull << 34))
        throw Error(MT_EINVAL, "implausible array size in " + in.path);
    if ((magic & 0xFFFF) == M_DBL) {  // M_DBL and the map/loc forms stored as doubles
        b->magic = M_DBL;
        b->data.resize(n);
        in.get(b->data.data(), n * sizeof(double));
    } else if (magic == M_CMP) {
        b->magic = M_CMP;
        b->data.resize(2 * n);
        in.get(b->data.data(), 2 * n * sizeof(double));
    } else if (magic == M_FLT || magic == M_ZMP) {
        b->magic = magic;
        const uint64_t m = magic == M_ZMP ? 2 * n : n;
        std::vector<float> f(m);
        in.get(f.data(), m * sizeof(float));
        b->data.assign(f.begin(), f.end());
    } else if (magic == M_INT64 || magic == M_INT32) {
        b->magic = magic;
        b->idata.resize(n);
        if (magic == M_INT64) {
            std::vector<int64_t> v(n);
            in.get(v.data(), n * 8);
            b->idata.assign(v.begin(), v.end());
        } else {
            std::vector<int32_t> v(n);
            in.get(v.data(), n * 4);
            b->idata.assign(v.begin(), v.end());
        }
    } else if (magic == M_DSP64 || magic == M_DSP32 || magic == M_CSP64 || magic == M_CSP32) {
        const bool cplx = magic == M_CSP64 || magic == M_CSP32, i64 = magic == M_DSP64 || magic == M_CSP64;
        b->magic = cplx ? M_CSP64 : M_DSP64;
        b->sp_p.assign(b->ny + 1, 0);
        if (n) {
            uint64_t nz;
            in.get(&nz, 8);
            if (nz > (1ull << 34)) throw Error(MT_EINVAL, "implausible sparse size in " + in.path);
            auto ints = [&](std::vector<long> &dst, uint64_t cnt) {
                dst.resize(cnt);
                if (i64) {
                    std::vector<int64_t> v(cnt);
                    in.get(v.data(), cnt * 8);
                    dst.assign(v.begin(), v.end());
                } else {
                    std::vector<int32_t> v(cnt);
                    in.get(v.data(), cnt * 4);
                    dst.assign(v.begin(), v.end());
                }
            };
            ints(b->sp_p, b->ny + 1);
            ints(b->sp_i, nz);
            b->data.resize(cplx ? 2 * nz : nz);
            in.get(b->data.data(), b->data.size() * sizeof(double));
            if (b->sp_p[0] != 0 || (uint64_t)b->sp_p[b->ny] > nz)
                throw Error(MT_EINVAL, "invalid sparse column pointers in " + in.path);
            for (uint64_t c = 0; c < b->ny; c++)
                if (b->sp_p[c + 1] < b->sp_p[c]) throw Error(MT_EINVAL, "invalid sparse column pointers in " + in.path);
            for (long r : b->sp_i)
                if (r < 0 || (uint64_t)r >= b->nx) throw Error(MT_EINVAL, "sparse row index out of range in " + in.path);
        }
    } else if (magic == MCC_ANY) {
        b->magic = MCC_ANY;
        b->child.resize(n);
        for (uint64_t i = 0; i < n; i++) {
            std::unique_ptr<mt_bin> c = read_node(in, false);
            if (c->magic == MCC_ANY && c->nx * c->ny == 0) c.reset();  // empty element
            b->child[i] = std::move(c);
        }
    } else {
        char msg[96];
        std::snprintf(msg, sizeof msg, "unsupported .bin data type 0x%x (double arrays and cells only)", magic);
        throw Error(MT_EINVAL, std::string(msg) + " in " + in.path);
    }
    return b;
}

void write_head(GzOut &out, const std::string &hdr, uint32_t magic, uint64_t nx, uint64_t ny) {
    if (!hdr.empty()) {
        const uint64_t len = hdr.size() + 1, n = (len + 7) / 8 * 8;
        std::string pad(hdr);
        pad.resize(n, '\0');
        out.put(&M_HEADER, 4);
        out.put(&n, 8);
        out.put(pad.data(), n);
        out.put(&n, 8);
        out.put(&M_HEADER, 4);
    }
    out.put(&M_SKIP, 4);
    out.put(&magic, 4);
    out.put(&nx, 8);
    out.put(&ny, 8);
}

void write_node(GzOut &out, const mt_bin *b) {
    if (!b) {
        write_head(out, "", MCC_ANY, 0, 0);
        return;
    }
    if (b->magic == MCC_ANY) {
        bool any = false;
        for (auto &c : b->child) any |= c != nullptr;
        write_head(out, b->header, MCC_ANY, any ? b->nx : 0, any ? b->ny : 0);
        if (any)
            for (auto &c : b->child) write_node(out, c.get());
    } else if (b->is_sparse()) {
        const bool empty = b->nx * b->ny == 0;
        write_head(out, b->header, b->magic, b->nx, b->ny);
        if (!empty) {
            const uint64_t nz = (uint64_t)b->sp_p[b->ny];
            out.put(&nz, 8);
            std::vector<int64_t> p(b->sp_p.begin(), b->sp_p.end()), i(b->sp_i.begin(), b->sp_i.begin() + nz);
            out.put(p.data(), p.size() * 8);
            out.put(i.data(), i.size() * 8);
            out.put(b->data.data(), (b->is_complex() ? 2 : 1) * nz * sizeof(double));
        }
    } else if (b->magic == M_FLT || b->magic == M_ZMP) {
        write_head(out, b->header, b->magic, b->nx, b->ny);
        std::vector<float> f(b->data.begin(), b->data.end());
        out.put(f.data(), f.size() * sizeof(float));
    } else if (b->magic == M_INT64 || b->magic == M_INT32) {
        write_head(out, b->header, b->magic, b->nx, b->ny);
        if (b->magic == M_INT64) {
            std::vector<int64_t> v(b->idata.begin(), b->idata.end());
            out.put(v.data(), v.size() * 8);
        } else {
            std::vector<int32_t> v(b->idata.begin(), b->idata.end());
            out.put(v.data(), v.size() * 4);
        }
    } else {
        write_head(out, b->header, b->magic, b->nx, b->ny);
        out.put(b->data.data(), b->data.size() * sizeof(double));
    }
}

template <typename F>
int guard(F &&f) {
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

bool ends_with(const std::string &s, const char *t) {
    const size_t n = std::strlen(t);
    return s.size() >= n && s.compare(s.size() - n, n, t) == 0;
}

}  // namespace

extern "C" {

int mt_bin_read(const char *path, mt_bin_t **out) {
    if (!path || !out) return MT_EINVAL;
    *out = nullptr;
    return guard([&] {
        GzIn in{gzopen(path, "rb"), path};
        if (!in.f) throw Error(MT_EINVAL, std::string("cannot open ") + path);
        std::unique_ptr<mt_bin> b;
        try {
            b = read_node(in, false);
            // MAOS files hold one array or cell; anything after it is ignored
        } catch (...) {
            gzclose(in.f);
            throw;
        }
        gzclose(in.f);
        *out = b.release();
    });
}

int mt_bin_write(const char *path, const mt_bin_t *b) {
    if (!path || !b) return MT_EINVAL;
    return guard([&] {
        const std::string p(path);
        GzOut out{gzopen(path, ends_with(p, ".gz") ? "wb6" : "wbT"), p};
        if (!out.f) throw Error(MT_EINVAL, "cannot create " + p);
        try {
            write_node(out, b);
        } catch (...) {
            gzclose(out.f);
            throw;
        }
        if (gzclose(out.f) != Z_OK) throw Error(MT_EINVAL, "write failed: " + p);
    });
}

void mt_bin_free(mt_bin_t *b) { delete b; }

int mt_bin_info(const mt_bin_t *b, int *is_cell, long *nx, long *ny) {
    if (!b) return MT_EINVAL;
    if (is_cell) *is_cell = b->magic == MCC_ANY;
    if (nx) *nx = (long)b->nx;
    if (ny) *ny = (long)b->ny;
    return MT_OK;
}

const double *mt_bin_data(const mt_bin_t *b) {
    return b && b->magic != MCC_ANY && !b->is_sparse() && b->magic != M_INT64 && b->magic != M_INT32 ? b->data.data()
                                                                                                       : nullptr;
}

int mt_bin_type(const mt_bin_t *b) {
    if (!b) return MT_EINVAL;
    switch (b->magic) {
        case MCC_ANY: return MT_BIN_CELL;
        case M_CMP: return MT_BIN_COMPLEX;
        case M_FLT: return MT_BIN_FLOAT;
        case M_ZMP: return MT_BIN_FCOMPLEX;
        case M_INT64:
        case M_INT32: return MT_BIN_INT;
        case M_DSP64: return MT_BIN_SPARSE;
        case M_CSP64: return MT_BIN_CSPARSE;
        default: return MT_BIN_DOUBLE;
    }
}

const long *mt_bin_idata(const mt_bin_t *b) {
    return b && (b->magic == M_INT64 || b->magic == M_INT32) ? b->idata.data() : nullptr;
}

int mt_bin_sparse(const mt_bin_t *b, const long **p, const long **i, const double **x, long *nzmax) {
    if (!b || !b->is_sparse()) return MT_EINVAL;
    if (p) *p = b->sp_p.data();
    if (i) *i = b->sp_i.data();
    if (x) *x = b->data.data();
    if (nzmax) *nzmax = b->sp_p.empty() ? 0 : b->sp_p.back();
    return MT_OK;
}

int mt_bin_new_typed(int type, long nx, long ny, const double *data, const char *header, mt_bin_t **out) {
    if (!out || nx < 0 || ny < 0) return MT_EINVAL;
    uint32_t magic;
    long m = nx * ny;
    switch (type) {
        case MT_BIN_DOUBLE: magic = M_DBL; break;
        case MT_BIN_COMPLEX: magic = M_CMP; m *= 2; break;
        case MT_BIN_FLOAT: magic = M_FLT; break;
        case MT_BIN_FCOMPLEX: magic = M_ZMP; m *= 2; break;
        case MT_BIN_INT: magic = M_INT64; break;
        default: return MT_EINVAL;
    }
    if (m > 0 && !data) return MT_EINVAL;
    return guard([&] {
        std::unique_ptr<mt_bin> b(new mt_bin());
        b->magic = magic;
        b->nx = (uint64_t)nx;
        b->ny = (uint64_t)ny;
        b->header = header ? header : "";
        if (type == MT_BIN_INT) {
            b->idata.resize(m);
            for (long k = 0; k < m; k++) {
                if (data[k] != std::floor(data[k])) throw Error(MT_EINVAL, "integer array holds a non-integer");
                b->idata[k] = (long)data[k];
            }
        } else {
            b->data.assign(data, data + m);
            if (type == MT_BIN_FLOAT || type == MT_BIN_FCOMPLEX)
                for (double &v : b->data) v = (double)(float)v;
        }
        *out = b.release();
    });
}

int mt_bin_new_sparse(long nx, long ny, const long *p, const long *i, const double *x, int cplx, const char *header,
                      mt_bin_t **out) {
    if (!out || nx < 0 || ny < 0 || (nx * ny > 0 && !p)) return MT_EINVAL;
    return guard([&] {
        std::unique_ptr<mt_bin> b(new mt_bin());
        b->magic = cplx ? M_CSP64 : M_DSP64;
        b->nx = (uint64_t)nx;
        b->ny = (uint64_t)ny;
        b->header = header ? header : "";
        b->sp_p.assign(ny + 1, 0);
        if (nx * ny > 0) {
            b->sp_p.assign(p, p + ny + 1);
            const long nz = p[ny];
            if (p[0] != 0 || nz < 0 || (nz > 0 && (!i || !x))) throw Error(MT_EINVAL, "invalid sparse matrix");
            for (long c = 0; c < ny; c++)
                if (p[c + 1] < p[c]) throw Error(MT_EINVAL, "invalid sparse column pointers");
            b->sp_i.assign(i, i + nz);
            for (long r : b->sp_i)
                if (r < 0 || r >= nx) throw Error(MT_EINVAL, "sparse row index out of range");
            b->data.assign(x, x + (cplx ? 2 : 1) * nz);
        }
        *out = b.release();
    });
}

const mt_bin_t *mt_bin_child(const mt_bin_t *b, long i) {
    if (!b || b->magic != MCC_ANY || i < 0 || (uint64_t)i >= b->child.size()) return nullptr;
    return b->child[i].get();
}

const char *mt_bin_header(const mt_bin_t *b) { return b ? b->header.c_str() : nullptr; }

int mt_bin_new_dbl(long nx, long ny, const double *data, const char *header, mt_bin_t **out) {
    if (!out || nx < 0 || ny < 0 || (nx * ny > 0 && !data)) return MT_EINVAL;
    return guard([&] {
        std::unique_ptr<mt_bin> b(new mt_bin());
        b->magic = M_DBL;
        b->nx = (uint64_t)nx;
        b->ny = (uint64_t)ny;
        b->header = header ? header : "";
        b->data.assign(data, data + nx * ny);
        *out = b.release();
    });
}

int mt_bin_new_cell(long nx, long ny, const char *header, mt_bin_t **out) {
    if (!out || nx < 0 || ny < 0) return MT_EINVAL;
    return guard([&] {
        std::unique_ptr<mt_bin> b(new mt_bin());
        b->magic = MCC_ANY;
        b->nx = (uint64_t)nx;
        b->ny = (uint64_t)ny;
        b->header = header ? header : "";
        b->child.resize((size_t)(nx * ny));
        *out = b.release();
    });
}

int mt_bin_set(mt_bin_t *cell, long i, mt_bin_t *child) {
    if (!cell || cell->magic != MCC_ANY || i < 0 || (uint64_t)i >= cell->child.size()) return MT_EINVAL;
    cell->child[i].reset(child);
    return MT_OK;
}

int mt_bin_grad(const mt_bin_t *b, const mt_recon_t *r, double *grad) {
    if (!b || !r || !grad) return MT_EINVAL;
    return guard([&] {
        if (b->magic != MCC_ANY || b->nx * b->ny != (uint64_t)r->nwfs)
            throw Error(MT_EINVAL, "gradients: expected a cell of one array per WFS");
        long off = 0;
        for (int w = 0; w < r->nwfs; w++) {
            if (r->skip && r->skip[w]) continue;
            const long n = 2 * r->nsa[w];
            const mt_bin *e = b->child[w].get();
            if (!e || e->magic != M_DBL || e->nx * e->ny != (uint64_t)n)
                throw Error(MT_EINVAL, "gradients of WFS " + std::to_string(w) + ": expected " + std::to_string(n) +
                                           " values (x slopes then y slopes)");
            std::memcpy(grad + off, e->data.data(), sizeof(double) * n);
            off += n;
        }
    });
}

int mt_bin_opdr(const mt_recon_t *r, const double *x, mt_bin_t **out) {
    if (!r || !x || !out) return MT_EINVAL;
    *out = nullptr;
    return guard([&] {
        std::unique_ptr<mt_bin> c(new mt_bin());
        c->magic = MCC_ANY;
        c->nx = (uint64_t)r->npsr;
        c->ny = 1;
        c->child.resize(r->npsr);
        long off = 0;
        for (int l = 0; l < r->npsr; l++) {
            std::unique_ptr<mt_bin> e(new mt_bin());
            e->magic = M_DBL;
            e->nx = (uint64_t)r->xnx[l];
            e->ny = (uint64_t)r->xny[l];
            e->data.assign(x + off, x + off + r->xnx[l] * r->xny[l]);
            off += r->xnx[l] * r->xny[l];
            c->child[l] = std::move(e);
        }
        *out = c.release();
    });
}

}  // extern "C"
