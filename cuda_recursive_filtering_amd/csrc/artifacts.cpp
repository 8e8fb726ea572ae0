// artifacts.cpp -- the reference harness's generated filter header as data
// (no GPU needed).
//
// The reference designs its test filter in Python and hands it to the CUDA
// drivers as a header compiled in (build/filter_taps.hpp, written by
// ref_generate.py:101-131; FILTER_N_SECTIONS is checked against N_SECTIONS at
// test_accuracy.cu:107-109, FILTER_SOS feeds ph_kernels.cuh:151 and the
// direct-form taps plr_kernels.cuh:416-427).  iirx_filter_taps_write emits
// the same declarations with the same text for every value, so a header
// written here builds the reference's drivers; iirx_filter_taps_read takes
// one written by either side, so the reference's artefacts drive this
// library (the reference's other artefact, reference.bin, is raw float32
// and needs no reader).
//
// Number text (locale-independent, std::to_chars / std::from_chars): floats
// as %.9e plus an "f" suffix (Python's .9e of the float32 value, the same
// correctly rounded digits); doubles as the
// shortest decimal that reads back to the same double (the closest such if
// several), laid out as Python's repr lays it out: positional for decimal
// exponents -5 < e < 16 with at least one digit after the point, else
// d.ddde+XX / d.ddde-XX.
#include "../../include/iirx.h"

#include <cctype>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <system_error>
#include <vector>

namespace {

// shortest round-trip decimal of v, in Python's repr layout
std::string repr_double(double v)
{
    if (std::isnan(v)) return "nan";
    if (std::isinf(v)) return v < 0 ? "-inf" : "inf";
    if (v == 0.0) return std::signbit(v) ? "-0.0" : "0.0";
    // shortest digits that read back to v, the closest of them to v
    // (std::to_chars' rule, which is repr's)
    char buf[64];
    const std::to_chars_result r = std::to_chars(buf, buf + sizeof buf - 1, v, std::chars_format::scientific);
    *r.ptr = '\0';
    // buf = [-]d[.ddd]e[+-]XX
    std::string s(buf);
    const bool neg = s[0] == '-';
    if (neg) s.erase(0, 1);
    const size_t epos = s.find('e');
    const int exp10 = std::atoi(s.c_str() + epos + 1);
    std::string digits;
    for (size_t i = 0; i < epos; ++i)
        if (std::isdigit((unsigned char)s[i])) digits.push_back(s[i]);
    while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
    const int nd = (int)digits.size();
    const int decpt = exp10 + 1;  // value = 0.digits x 10^decpt
    std::string out = neg ? "-" : "";
    if (decpt > -4 && decpt <= 16) {
        if (decpt <= 0) {
            out += "0.";
            out.append((size_t)(-decpt), '0');
            out += digits;
        } else if (decpt < nd) {
            out += digits.substr(0, (size_t)decpt) + "." + digits.substr((size_t)decpt);
        } else {
            out += digits;
            out.append((size_t)(decpt - nd), '0');
            out += ".0";
        }
    } else {
        out += digits.substr(0, 1);
        if (nd > 1) out += "." + digits.substr(1);
        char e[16];
        std::snprintf(e, sizeof e, "e%c%02d", exp10 < 0 ? '-' : '+', std::abs(exp10));
        out += e;
    }
    return out;
}

// %.9e of the float's value (the same correctly rounded digits as printf,
// independent of the C locale), plus the f suffix
std::string f32_text(float v)
{
    char buf[64];
    const std::to_chars_result r = std::to_chars(buf, buf + sizeof buf - 2, (double)v, std::chars_format::scientific, 9);
    *r.ptr = 'f';
    r.ptr[1] = '\0';
    return buf;
}

// locale-independent, correctly rounded number reads
template <typename T>
const char *read_number(const char *p, const char *end, T &v)
{
    if (p < end && *p == '+') ++p;  // from_chars takes no leading '+'
    const std::from_chars_result r = std::from_chars(p, end, v, std::chars_format::general);
    return r.ec == std::errc() ? r.ptr : nullptr;
}

// the header text with comments removed (// to end of line, /* ... */)
std::string strip_comments(const std::string &src)
{
    std::string out;
    out.reserve(src.size());
    for (size_t i = 0; i < src.size();) {
        if (src.compare(i, 2, "//") == 0) {
            while (i < src.size() && src[i] != '\n') ++i;
        } else if (src.compare(i, 2, "/*") == 0) {
            const size_t e = src.find("*/", i + 2);
            i = (e == std::string::npos) ? src.size() : e + 2;
            out.push_back(' ');
        } else {
            out.push_back(src[i++]);
        }
    }
    return out;
}

// position just past `name` as a whole identifier, or npos
size_t find_ident(const std::string &s, const char *name)
{
    const size_t len = std::strlen(name);
    for (size_t p = s.find(name); p != std::string::npos; p = s.find(name, p + 1)) {
        const bool left = p == 0 || !(std::isalnum((unsigned char)s[p - 1]) || s[p - 1] == '_');
        const bool right = p + len >= s.size() || !(std::isalnum((unsigned char)s[p + len]) || s[p + len] == '_');
        if (left && right) return p + len;
    }
    return std::string::npos;
}

bool read_define(const std::string &s, const char *name, long &v)
{
    size_t p = find_ident(s, name);
    if (p == std::string::npos) return false;
    while (p < s.size() && std::isspace((unsigned char)s[p])) ++p;
    const std::from_chars_result r = std::from_chars(s.data() + p, s.data() + s.size(), v, 10);
    return r.ec == std::errc();
}

// "NAME = value;" -> value
bool read_scalar(const std::string &s, const char *name, double &v)
{
    size_t p = find_ident(s, name);
    if (p == std::string::npos) return false;
    p = s.find('=', p);
    if (p == std::string::npos) return false;
    ++p;
    while (p < s.size() && std::isspace((unsigned char)s[p])) ++p;
    return read_number(s.data() + p, s.data() + s.size(), v) != nullptr;
}

// the numbers between the braces of "NAME[...]... = { ... };" in order; for
// T = float each is read as a compiler reads an f-suffixed literal (one
// rounding from the decimal text to float)
template <typename T>
bool read_array(const std::string &s, const char *name, std::vector<T> &out)
{
    size_t p = find_ident(s, name);
    if (p == std::string::npos) return false;
    p = s.find('=', p);
    if (p == std::string::npos) return false;
    p = s.find('{', p);
    if (p == std::string::npos) return false;
    int depth = 0;
    out.clear();
    const char *c = s.c_str();
    for (size_t i = p; i < s.size(); ++i) {
        const char ch = c[i];
        if (ch == '{') {
            ++depth;
        } else if (ch == '}') {
            if (--depth == 0) return true;
        } else if (std::isdigit((unsigned char)ch) || ch == '-' || ch == '+' || ch == '.') {
            T v;
            const char *end = read_number(c + i, c + s.size(), v);
            if (!end) return false;
            out.push_back(v);
            i = (size_t)(end - c);
            if (i < s.size() && (c[i] == 'f' || c[i] == 'F' || c[i] == 'l' || c[i] == 'L')) ++i;
            --i;
        } else if (!(std::isspace((unsigned char)ch) || ch == ',')) {
            return false;
        }
    }
    return false;
}

}  // namespace

extern "C" {

int iirx_filter_taps_write(const iirx_filter_taps *t, const char *path)
{
    if (!t || !path || t->n_sections < 1 || t->n_sections > IIRX_TAPS_MAX_SECTIONS || t->plr_order < 1 ||
        t->plr_order > 2 * IIRX_TAPS_MAX_SECTIONS)
        return IIRX_ERR_INVALID;
    const int ns = t->n_sections, k = t->plr_order;
    std::string h;
    h += "// Filter header in the layout of the reference's build/filter_taps.hpp, written by\n";
    h += "// iirx_filter_taps_write (libiirx).\n";
    if (t->cutoff_wn > 0.0)
        h += "// Butterworth lowpass, order " + std::to_string(2 * ns) + ", cutoff " + repr_double(t->cutoff_wn) +
             " x Nyquist.\n";
    h += "// Sections have b0 == 1; multiply the input by FILTER_GAIN.\n";
    h += "#pragma once\n\n";
    h += "#define FILTER_N_SECTIONS " + std::to_string(ns) + "\n";
    h += "static const double FILTER_GAIN = " + repr_double(t->gain) + ";\n";
    h += "static const double INPUT_FS    = " + repr_double(t->input_fs) + ";\n";
    h += "static const double INPUT_FREQ  = " + repr_double(t->input_freq) + ";\n\n";
    h += "static const float FILTER_SOS[" + std::to_string(ns) + "][6] = {\n";
    for (int s = 0; s < ns; ++s) {
        h += "    {";
        for (int j = 0; j < 6; ++j) h += (j ? ", " : "") + f32_text(t->sos[6 * s + j]);
        h += (s + 1 < ns) ? "},\n" : "}\n";
    }
    h += "};\n\n";
    h += "// The same transfer function in direct form (PLR): taps past b0 == 1;\n";
    h += "// feedback sign y[n] += A[j] * y[n-j-1].\n";
    h += "#define PLR_FILTER_ORDER " + std::to_string(k) + "\n";
    for (int which = 0; which < 2; ++which) {
        const double *v = which ? t->plr_a : t->plr_b;
        h += std::string("static const double PLR_FILTER_") + (which ? "A" : "B") + "[" + std::to_string(k) +
             "] = {\n    ";
        for (int j = 0; j < k; ++j) h += (j ? ",\n    " : "") + repr_double(v[j]);
        h += "\n};\n";
    }
    std::FILE *fp = std::fopen(path, "w");
    if (!fp) return IIRX_ERR_INVALID;
    const bool ok = std::fwrite(h.data(), 1, h.size(), fp) == h.size();
    return (std::fclose(fp) == 0 && ok) ? IIRX_OK : IIRX_ERR_INVALID;
}

int iirx_filter_taps_read(const char *path, iirx_filter_taps *t)
{
    if (!path || !t) return IIRX_ERR_INVALID;
    std::FILE *fp = std::fopen(path, "rb");
    if (!fp) return IIRX_ERR_INVALID;
    std::string src;
    char buf[4096];
    for (size_t got; (got = std::fread(buf, 1, sizeof buf, fp)) > 0;) src.append(buf, got);
    std::fclose(fp);

    // the cutoff is only in the design comment ("cutoff <wn> x Nyquist")
    double wn = 0.0;
    const size_t cp = src.find("cutoff ");
    if (cp != std::string::npos && !read_number(src.data() + cp + 7, src.data() + src.size(), wn)) wn = 0.0;

    const std::string s = strip_comments(src);
    long ns = 0, k = 0;
    iirx_filter_taps r{};
    std::vector<float> sos;
    std::vector<double> b, a;
    if (!read_define(s, "FILTER_N_SECTIONS", ns) || ns < 1 || ns > IIRX_TAPS_MAX_SECTIONS) return IIRX_ERR_INVALID;
    if (!read_scalar(s, "FILTER_GAIN", r.gain) || !read_scalar(s, "INPUT_FS", r.input_fs) ||
        !read_scalar(s, "INPUT_FREQ", r.input_freq))
        return IIRX_ERR_INVALID;
    if (!read_array(s, "FILTER_SOS", sos) || sos.size() != 6 * (size_t)ns) return IIRX_ERR_INVALID;
    if (!read_define(s, "PLR_FILTER_ORDER", k) || k < 1 || k > 2 * IIRX_TAPS_MAX_SECTIONS) return IIRX_ERR_INVALID;
    if (!read_array(s, "PLR_FILTER_B", b) || !read_array(s, "PLR_FILTER_A", a) || b.size() != (size_t)k ||
        a.size() != (size_t)k)
        return IIRX_ERR_INVALID;
    r.n_sections = (int32_t)ns;
    r.plr_order = (int32_t)k;
    r.cutoff_wn = (wn > 0.0 && wn < 1.0) ? wn : 0.0;
    std::copy(sos.begin(), sos.end(), r.sos);
    std::copy(b.begin(), b.end(), r.plr_b);
    std::copy(a.begin(), a.end(), r.plr_a);
    *t = r;
    return IIRX_OK;
}

}  // extern "C"
