// MAOS .conf reading: the key store of open_config and the value parsers of
// readcfg / readstr, so that a MAOS configuration (the .conf files and
// command-line overrides a MAOS run reads) can drive this library.
//
// Restates:
//   open_config     lib/readcfg.c:236-456 (lines, '#' comments, trailing
//                   '\' continuation, include = "x.conf" and bare x.conf
//                   lines, __reset__, key += [..] array append, priorities:
//                   a key is replaced by a later one of the same or higher
//                   priority)
//   strtrim/squeeze lib/readcfg.c:72-157
//   strextract      lib/readcfg.c:110-131 (paired quotes removed)
//   readstr_num     sys/readstr.c:135-175 (a number with * and / factors)
//   readstr_numarr  sys/readstr.c:193-420 (factors before '[', * / + -
//                   after ']', ';' separates columns, ' transposes)
//   readstr_strarr  sys/readstr.c:40-129
//   readstr_dmat_do lib/readcfg.c:606-626 (a value naming a .bin/.bin.gz
//                   file is read as the double array it holds)
// Deviations: files are looked up as given, then in the directories added
// with mt_conf_add_path (most recent first), as MAOS's search path does
// (sys/path.c:113-131) without its built-in config directory; the
// transpose of a non-square array uses the array's own row count (the
// reference's loop bound reads past the array there); the old-key renames
// of COMPATIBILITY mode are not applied.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <new>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "setup_internal.hpp"

using namespace mt;

struct mt_conf {
    struct Rec {
        std::string data;
        bool has = false;  // value present (an empty value is NULL in MAOS)
        long priority = 0;
        long used = 0;
    };
    std::map<std::string, Rec> keys;
    std::vector<std::string> paths;  // most recent first
    int depth = 0;
};

namespace {

bool is_end(char c) { return c == '\0' || c == '\n' || c == '\r'; }
bool file_exists(const std::string &f) {
    struct stat st;
    return stat(f.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}
bool ends_with(const std::string &s, const char *suf) {
    const size_t n = std::strlen(suf);
    return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
}

// strtrim: non-printing characters to spaces, leading and trailing spaces
// and trailing ';' removed, runs of spaces collapsed; empty -> false.
bool strtrim(std::string &s) {
    for (char &c : s)
        if (!std::isgraph((unsigned char)c) || std::isspace((unsigned char)c)) c = ' ';
    size_t b = 0;
    while (b < s.size() && s[b] == ' ') b++;
    size_t e = s.size();
    while (e > b && (s[e - 1] == ' ' || s[e - 1] == ';')) e--;
    std::string out;
    bool sp = false;
    for (size_t i = b; i < e; i++) {
        if (s[i] == ' ') {
            if (sp) continue;
            sp = true;
        } else {
            sp = false;
        }
        out.push_back(s[i]);
    }
    s = out;
    return !s.empty();
}

// strextract: "x" or 'x' -> x (quotes must pair); empty -> false.
bool strextract(const std::string &d, std::string &out) {
    if (d.empty()) return false;
    if (d[0] == '"' || d[0] == '\'') {
        if (d.size() < 2 || d.back() != d[0]) throw Error(MT_EINVAL, "record {" + d + "}: quotes must come in pair");
        out = d.substr(1, d.size() - 2);
        return !out.empty();
    }
    out = d;
    return true;
}

// squeeze: drop the comment, trailing white space and line ends, leading
// spaces; false for an empty line.
bool squeeze(std::string &line) {
    const size_t h = line.find('#');
    if (h != std::string::npos) line.resize(h);
    while (!line.empty() && (std::isspace((unsigned char)line.back()) || is_end(line.back()))) line.pop_back();
    size_t b = 0;
    while (b < line.size() && std::isspace((unsigned char)line[b])) b++;
    line = line.substr(b);
    return !line.empty();
}

std::string find_file(const mt_conf *c, const std::string &fn) {
    if (file_exists(fn)) return fn;
    for (const std::string &p : c->paths) {
        std::string f = p;
        if (!f.empty() && f.back() != '/') f += '/';
        f += fn;
        if (file_exists(f)) return f;
    }
    throw Error(MT_EINVAL, "cannot find file " + fn);
}

void open_config(mt_conf *c, const std::string &in, const char *prefix, int priority) {
    if (++c->depth > 64) throw Error(MT_EINVAL, "configuration includes nest too deep");
    struct DepthGuard {
        mt_conf *c;
        ~DepthGuard() { c->depth--; }
    } guard{c};
    std::vector<std::string> lines;
    if (ends_with(in, ".conf")) {
        const std::string fn = find_file(c, in);
        std::ifstream f(fn);
        if (!f) throw Error(MT_EINVAL, "cannot open " + fn);
        std::string l;
        while (std::getline(f, l)) lines.push_back(l);
    } else {
        size_t a = 0;
        while (a <= in.size()) {
            size_t e = in.find('\n', a);
            if (e == std::string::npos) e = in.size();
            lines.push_back(in.substr(a, e - a));
            a = e + 1;
        }
    }
    std::string ss;  // the accumulated logical line
    for (std::string line : lines) {
        if (!squeeze(line)) continue;
        ss += line;
        if (!ss.empty() && ss.back() == '\\') {  // continues on the next line
            ss.pop_back();
            continue;
        }
        const size_t eq = ss.find('=');
        if (eq == std::string::npos) {
            if (ends_with(ss, ".conf")) {
                std::string fn;
                if (strextract(ss, fn)) open_config(c, fn, prefix, priority);
            } else if (ss == "__reset__") {
                c->keys.clear();
            } else {
                throw Error(MT_EINVAL, "configuration input (" + ss + ") is not valid");
            }
            ss.clear();
            continue;
        }
        bool append = eq > 0 && ss[eq - 1] == '+';
        std::string var = ss.substr(0, append ? eq - 1 : eq);
        std::string value = ss.substr(eq + 1);
        ss.clear();
        if (!strtrim(var)) throw Error(MT_EINVAL, "configuration line '" + line + "' is invalid");
        const bool hasval = strtrim(value);
        if (var == "path" || var == "PATH") {
            std::string p;
            if (hasval && strextract(value, p)) c->paths.insert(c->paths.begin(), p);
            continue;
        }
        if (var == "include") {
            std::string fn;
            if (hasval && strextract(value, fn)) open_config(c, fn, prefix, priority);
            continue;
        }
        const std::string key = (prefix ? std::string(prefix) : std::string()) + var;
        auto it = c->keys.find(key);
        if (it == c->keys.end()) {
            mt_conf::Rec r;
            r.data = hasval ? value : std::string();
            r.has = hasval;
            r.priority = priority;
            c->keys[key] = r;
            continue;
        }
        mt_conf::Rec &old = it->second;
        if (old.priority > priority) continue;  // not overridden
        if (append) {
            const std::string &a = old.data, &b = value;
            if (a.size() < 2 || a.front() != '[' || a.back() != ']' || b.size() < 2 || b.front() != '[' ||
                b.back() != ']')
                throw Error(MT_EINVAL, "key " + key + ": += needs [] arrays on both sides");
            old.data = a.substr(0, a.size() - 1) + " " + b.substr(1);
        } else {
            old.data = hasval ? value : std::string();
            old.has = hasval;
            old.priority = priority;
            old.used = 0;
        }
    }
    if (!ss.empty()) throw Error(MT_EINVAL, "configuration ends inside a continued line");
}

// readstr_num: a number followed by any chain of * x and / x factors.
double readstr_num(const char *data, const char **endp) {
    char *e;
    double res = std::strtod(data, &e);
    if (e == data) throw Error(MT_EINVAL, std::string("{") + data + "}: unable to parse for a number");
    while (std::isspace((unsigned char)*e)) e++;
    while (*e == '/' || *e == '*') {
        const bool div = *e == '/';
        e++;
        while (std::isspace((unsigned char)*e)) e++;
        const char *d = e;
        const double t = std::strtod(d, &e);
        if (e == d) throw Error(MT_EINVAL, std::string("{") + d + "}: failed to parse for a number");
        res = div ? res / t : res * t;
        while (std::isspace((unsigned char)*e)) e++;
    }
    *endp = e;
    return res;
}

// readstr_numarr: returns the values (column-major) and the shape.
std::vector<double> readstr_numarr(const std::string &s, long &nrow_out, long &ncol_out) {
    std::vector<double> v;
    nrow_out = ncol_out = 0;
    if (s.empty()) return v;
    const char *data = s.c_str();
    const char *p = data;
    double fact = 1;
    int power = 1;
    int addition = 0;
    double addval = 0;
    bool trans = false;
    if (std::strchr(p, '[')) {
        while (*p != '[') {  // factors before the array
            char *e;
            const double f1 = std::strtod(p, &e);
            if (e == p) throw Error(MT_EINVAL, "{" + s + "}: invalid entry to parse for numerical array");
            fact = power == 1 ? fact * f1 : fact / f1;
            while (std::isspace((unsigned char)*e)) e++;
            if (*e == '/') power = -1;
            else if (*e == '*') power = 1;
            else throw Error(MT_EINVAL, "{" + s + "}: invalid entry to parse for numerical array");
            p = e + 1;
        }
        p++;
    }
    if (const char *rb = std::strchr(p, ']')) {  // operations after the array
        const char *e = rb + 1;
        while (std::isspace((unsigned char)*e) || *e == '\'') {
            if (*e == '\'') trans = !trans;
            e++;
        }
        while (*e == '/' || *e == '*' || *e == '+' || *e == '-') {
            int power2 = 1;
            if (*e == '/') power2 = -1;
            else if (*e == '*') power2 = 1;
            else {
                power2 = 0;
                addition = *e == '+' ? 1 : -1;
            }
            if (addition && power2) throw Error(MT_EINVAL, "{" + s + "}: * or / after + or - is not supported");
            e++;
            while (std::isspace((unsigned char)*e)) e++;
            char *e2;
            const double f2 = std::strtod(e, &e2);
            if (e2 == e) throw Error(MT_EINVAL, "{" + s + "}: invalid entry to parse for numerical array");
            e = e2;
            while (std::isspace((unsigned char)*e)) e++;
            if (addition) addval += addition == 1 ? f2 : -f2;
            else fact = power2 == 1 ? fact * f2 : fact / f2;
        }
        if (!is_end(*e) && *e != ';') throw Error(MT_EINVAL, "{" + s + "}: garbage at the end of the string");
    }
    long nrow = 0, ncol = 0, rowbegin = 0;
    while (*p != ']' && !is_end(*p)) {
        const char *e;
        double res = readstr_num(p, &e);
        p = e;
        res = power == 1 ? fact * res : fact / res;
        if (addition) res += addval;
        v.push_back(res);
        while (*p == ' ' || *p == ';') {
            if (*p == ';') {
                ncol++;
                const long cnt = (long)v.size() - rowbegin;
                if (nrow == 0) nrow = cnt;
                else if (nrow != cnt) throw Error(MT_EINVAL, "{" + s + "}: rows of different lengths");
                rowbegin = (long)v.size();
            }
            p++;
        }
    }
    if (rowbegin < (long)v.size()) {
        ncol++;
        const long cnt = (long)v.size() - rowbegin;
        if (nrow == 0) nrow = cnt;
        else if (nrow != cnt) throw Error(MT_EINVAL, "{" + s + "}: rows of different lengths");
    }
    if (nrow * ncol != (long)v.size()) throw Error(MT_EINVAL, "{" + s + "}: inconsistent array shape");
    if (trans && !v.empty()) {
        std::vector<double> t(v.size());
        for (long c = 0; c < ncol; c++)
            for (long r = 0; r < nrow; r++) t[c + ncol * r] = v[r + nrow * c];
        v.swap(t);
        std::swap(nrow, ncol);
    }
    nrow_out = nrow;
    ncol_out = ncol;
    return v;
}

// readstr_strarr: entries separated by space, ',' or ';', optionally quoted;
// empty entries are kept (as empty strings).
std::vector<std::string> readstr_strarr(const std::string &s) {
    std::vector<std::string> out;
    if (s.empty()) return out;
    const char *sdata = s.c_str();
    const char *end = sdata + s.size() - 1;  // last character
    const char *p = sdata;
    if (sdata[0] == '[') {
        p++;
        if (*end != ']') throw Error(MT_EINVAL, "{" + s + "}: does not end in ]");
    } else {
        end = sdata + s.size();  // no brackets: the whole string
    }
    while (p < end && *p == ' ') p++;
    bool end_coma = false;
    while (p < end || end_coma) {
        const char *q = end, *next;
        end_coma = false;
        if (*p == '"' || *p == '\'') {
            const char sep = *p++;
            q = std::strchr(p, sep);
            if (!q) throw Error(MT_EINVAL, "{" + s + "}: unmatched quote");
            next = q + 1;
            while (next < end && *next == ' ') next++;
            if (*next == ',' || *next == ';') {
                next++;
                end_coma = true;
            }
        } else {
            for (char sep : {' ', ',', ';'}) {
                const char *t = std::strchr(p, sep);
                if (t && t < q && t < end) q = t;
            }
            if (q < end && (*q == ',' || *q == ';')) end_coma = true;
            next = q + 1;
        }
        out.push_back(std::string(p, q > p ? (size_t)(q - p) : 0));
        p = next;
        while (p < end && *p == ' ') p++;
    }
    return out;
}

mt_conf::Rec &getrecord(mt_conf *c, const char *key) {
    std::string k(key ? key : "");
    strtrim(k);
    auto it = c->keys.find(k);
    if (it == c->keys.end()) throw Error(MT_EINVAL, "key '" + k + "' not found");
    it->second.used++;
    return it->second;
}

template <typename F>
int cguard(F &&f) {
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

void copy_out(const std::string &s, char *buf, long len) {
    if (!buf || len <= 0) return;
    const size_t n = std::min<size_t>(s.size(), (size_t)len - 1);
    std::memcpy(buf, s.data(), n);
    buf[n] = '\0';
}

}  // namespace

extern "C" {

int mt_conf_new(mt_conf_t **out) {
    if (!out) return MT_EINVAL;
    *out = new (std::nothrow) mt_conf();
    return *out ? MT_OK : MT_ENOMEM;
}

void mt_conf_free(mt_conf_t *c) { delete c; }

int mt_conf_add_path(mt_conf_t *c, const char *dir) {
    if (!c || !dir) return MT_EINVAL;
    return cguard([&] { c->paths.insert(c->paths.begin(), std::string(dir)); });
}

int mt_conf_read(mt_conf_t *c, const char *file_or_text, const char *prefix, int priority) {
    if (!c || !file_or_text) return MT_EINVAL;
    return cguard([&] {
        std::string in(file_or_text);
        while (!in.empty() && std::isspace((unsigned char)in.back())) in.pop_back();
        if (in.empty()) return;
        open_config(c, in, prefix, priority);
    });
}

int mt_conf_has(const mt_conf_t *c, const char *key) {
    if (!c || !key) return 0;
    std::string k(key);
    strtrim(k);
    return c->keys.count(k) ? 1 : 0;
}

int mt_conf_str(mt_conf_t *c, const char *key, char *buf, long len, long *n) {
    if (!c || !key) return MT_EINVAL;
    return cguard([&] {
        mt_conf::Rec &r = getrecord(c, key);
        std::string v;
        if (r.has) strextract(r.data, v);
        if (n) *n = (long)v.size();
        copy_out(v, buf, len);
    });
}

int mt_conf_strarr(mt_conf_t *c, const char *key, long i, char *buf, long len, long *n) {
    if (!c || !key) return MT_EINVAL;
    return cguard([&] {
        mt_conf::Rec &r = getrecord(c, key);
        const std::vector<std::string> v = r.has ? readstr_strarr(r.data) : std::vector<std::string>();
        if (i < 0) {
            if (n) *n = (long)v.size();
            return;
        }
        if (i >= (long)v.size()) throw Error(MT_EINVAL, "string array index out of range");
        if (n) *n = (long)v[i].size();
        copy_out(v[i], buf, len);
    });
}

int mt_conf_dbl(mt_conf_t *c, const char *key, double *out, long maxn, long *n, long *nrow, long *ncol) {
    if (!c || !key || maxn < 0 || (maxn > 0 && !out)) return MT_EINVAL;
    return cguard([&] {
        mt_conf::Rec &r = getrecord(c, key);
        std::vector<double> v;
        long nr = 0, nc = 0;
        std::string fn;
        if (r.has && (ends_with(r.data, ".bin") || ends_with(r.data, ".gz") || ends_with(r.data, ".bin\"") ||
                      ends_with(r.data, ".gz\"")) && strextract(r.data, fn)) {
            // a file holding the array (readstr_dmat_do)
            mt_bin_t *b = nullptr;
            const int rc = mt_bin_read(find_file(c, fn).c_str(), &b);
            if (rc != MT_OK) throw Error(rc, last_error_cstr());
            int cell = 0;
            mt_bin_info(b, &cell, &nr, &nc);
            if (cell) {
                mt_bin_free(b);
                throw Error(MT_EINVAL, "key " + std::string(key) + ": file holds a cell, not an array");
            }
            const double *d = mt_bin_data(b);
            v.assign(d, d + nr * nc);
            mt_bin_free(b);
        } else if (r.has) {
            v = readstr_numarr(r.data, nr, nc);
        }
        if ((long)v.size() > maxn && maxn > 0) throw Error(MT_EINVAL, "key " + std::string(key) + ": more than " +
                                                                        std::to_string(maxn) + " numbers");
        if (maxn > 0) std::copy(v.begin(), v.end(), out);
        if (n) *n = (long)v.size();
        if (nrow) *nrow = nr;
        if (ncol) *ncol = nc;
    });
}

int mt_conf_unused(const mt_conf_t *c, char *buf, long len, long *n) {
    if (!c) return MT_EINVAL;
    return cguard([&] {
        std::string s;
        long cnt = 0;
        for (const auto &kv : c->keys) {
            if (kv.second.used == 0 && !(kv.second.has && kv.second.data == "ignore")) {
                if (!s.empty()) s += ' ';
                s += kv.first;
                cnt++;
            }
        }
        if (n) *n = cnt;
        copy_out(s, buf, len);
    });
}

}  // extern "C"
