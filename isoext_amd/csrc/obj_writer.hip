// Wavefront OBJ writer with the reference's exact text
// (src/isoext/utils.py:42-63): "v x y z\n" per vertex with each coordinate
// printed as Python's repr of the float (the shortest decimal string that
// reads back to the same double), then "f a b c\n" per face with 1-based
// indices.  Host code only: chunks of lines are formatted on worker threads
// and written in order.

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/isoext_amd.h"

namespace iso {
int set_error(int code, const char *what);
}

namespace {

// Python's float repr (format code 'r'): the shortest round-trip digits d1..dn
// with decimal point position decpt (value = 0.d1..dn * 10^decpt); fixed
// notation when -4 < decpt <= 16 (with ".0" when there is no fractional
// part), otherwise d1[.d2..dn]e[+-]XX with at least two exponent digits.
// inf / nan as "inf", "-inf", "nan".
char *py_float_repr(double x, char *out) {
    if (std::isnan(x)) {
        std::memcpy(out, "nan", 3);
        return out + 3;
    }
    if (std::isinf(x)) {
        if (x < 0) *out++ = '-';
        std::memcpy(out, "inf", 3);
        return out + 3;
    }
    if (std::signbit(x)) {
        *out++ = '-';
        x = -x;
    }
    if (x == 0.0) {
        std::memcpy(out, "0.0", 3);
        return out + 3;
    }
    // shortest round-trip digits in scientific form: d[.ddd]e[+-]XX
    char sci[64];
    const auto res = std::to_chars(sci, sci + sizeof(sci), x, std::chars_format::scientific);
    char digits[32];
    int nd = 0;
    const char *p = sci;
    for (; p < res.ptr && *p != 'e'; ++p)
        if (*p != '.') digits[nd++] = *p;
    int exp10 = 0;
    std::from_chars(p + 1 + (p[1] == '+' ? 1 : 0), res.ptr, exp10);
    const int decpt = exp10 + 1;
    if (decpt > -4 && decpt <= 16) {
        if (decpt <= 0) {
            *out++ = '0';
            *out++ = '.';
            for (int i = 0; i < -decpt; ++i) *out++ = '0';
            std::memcpy(out, digits, nd);
            out += nd;
        } else if (decpt >= nd) {
            std::memcpy(out, digits, nd);
            out += nd;
            for (int i = nd; i < decpt; ++i) *out++ = '0';
            *out++ = '.';
            *out++ = '0';
        } else {
            std::memcpy(out, digits, decpt);
            out += decpt;
            *out++ = '.';
            std::memcpy(out, digits + decpt, nd - decpt);
            out += nd - decpt;
        }
        return out;
    }
    *out++ = digits[0];
    if (nd > 1) {
        *out++ = '.';
        std::memcpy(out, digits + 1, nd - 1);
        out += nd - 1;
    }
    *out++ = 'e';
    const int e = decpt - 1;
    *out++ = e < 0 ? '-' : '+';
    const int ae = e < 0 ? -e : e;
    if (ae < 10) *out++ = '0';
    const auto r = std::to_chars(out, out + 8, ae);
    return r.ptr;
}

void format_vertices(const double *v, int64_t lo, int64_t hi, std::string &buf) {
    buf.resize((size_t) (hi - lo) * 80);
    char *o = buf.data();
    for (int64_t i = lo; i < hi; ++i) {
        *o++ = 'v';
        for (int k = 0; k < 3; ++k) {
            *o++ = ' ';
            o = py_float_repr(v[3 * i + k], o);
        }
        *o++ = '\n';
    }
    buf.resize((size_t) (o - buf.data()));
}

void format_faces(const int64_t *f, int64_t lo, int64_t hi, std::string &buf) {
    buf.resize((size_t) (hi - lo) * 72);
    char *o = buf.data();
    for (int64_t i = lo; i < hi; ++i) {
        *o++ = 'f';
        for (int k = 0; k < 3; ++k) {
            *o++ = ' ';
            o = std::to_chars(o, o + 24, f[3 * i + k] + 1).ptr;
        }
        *o++ = '\n';
    }
    buf.resize((size_t) (o - buf.data()));
}

}   // namespace

extern "C" {

int64_t isoext_format_float(double x, char *out, int64_t cap) {
    char tmp[64];
    const int64_t n = py_float_repr(x, tmp) - tmp;
    if (!out || cap < n + 1) return -n;
    std::memcpy(out, tmp, (size_t) n);
    out[n] = '\0';
    return n;
}

int isoext_write_obj(const char *path, const double *v, int64_t nv, const int64_t *f, int64_t nf, int threads) {
    FILE *fp = path ? std::fopen(path, "wb") : nullptr;
    if (!fp) return iso::set_error(ISOEXT_E_ARG, "cannot open the OBJ file for writing");
    if (nv <= 0 || nf <= 0 || !v || !f) {   // an empty mesh leaves an empty file
        std::fclose(fp);
        return ISOEXT_OK;
    }
    const int nt = std::max(1, std::min(threads > 0 ? threads : 8, 64));
    constexpr int64_t kChunk = 1 << 16;   // lines per work item
    const int64_t cv = (nv + kChunk - 1) / kChunk, cf = (nf + kChunk - 1) / kChunk, nchunks = cv + cf;
    int rc = ISOEXT_OK;
    // rounds of nt chunks: format in parallel, write in order
    std::vector<std::string> bufs((size_t) nt);
    for (int64_t c0 = 0; c0 < nchunks && rc == ISOEXT_OK; c0 += nt) {
        const int64_t c1 = std::min(nchunks, c0 + nt);
        std::vector<std::thread> pool;
        for (int64_t c = c0; c < c1; ++c) {
            std::string &b = bufs[(size_t) (c - c0)];
            pool.emplace_back([&, c] {
                if (c < cv) format_vertices(v, c * kChunk, std::min(nv, (c + 1) * kChunk), b);
                else format_faces(f, (c - cv) * kChunk, std::min(nf, (c - cv + 1) * kChunk), b);
            });
        }
        for (auto &t : pool) t.join();
        for (int64_t c = c0; c < c1; ++c) {
            const std::string &b = bufs[(size_t) (c - c0)];
            if (std::fwrite(b.data(), 1, b.size(), fp) != b.size()) {
                rc = iso::set_error(ISOEXT_E_ARG, "writing the OBJ file failed");
                break;
            }
        }
    }
    if (std::fclose(fp) != 0 && rc == ISOEXT_OK) rc = iso::set_error(ISOEXT_E_ARG, "closing the OBJ file failed");
    return rc;
}

}   // extern "C"
