// design.cpp -- host-side filter design and test-signal generation of the
// C ABI (no GPU needed).
//
// iirx_butterworth_sos restates scipy.signal.butter(order, Wn, 'lowpass',
// output='sos') in closed form (the reference designs its filter with scipy
// in tools/ref_generate.py:62-67): analog prototype poles, frequency
// pre-warping 4 tan(pi Wn / 2) (fs = 2), bilinear map z = (4 + s)/(4 - s),
// all zeros at z = -1, sections ordered as scipy's zpk2sos 'nearest' pairing
// does (poles farthest from the unit circle first), then the b0 == 1
// normalisation of ref_generate.py:63-67 with the removed gain returned.
//
// iirx_fill_normal is the documented N(0,1) generator of the synthetic
// workloads: SplitMix64 (Steele, Lea & Flood 2014) drives the polar-free
// Box-Muller transform in float64; sample 2i is r cos(theta), sample 2i+1 is
// r sin(theta), each rounded once to float32.
#include "../../include/iirx.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

namespace {

inline uint64_t splitmix64(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline double unit53(uint64_t v) { return (double)(v >> 11) * (1.0 / 9007199254740992.0); }

// Butterworth lowpass of even `order` at `wn`: the FP64 feedback pair
// (a1, a2) of every section in scipy's order, and the removed gain.
double butterworth_sections(int order, double wn, std::vector<double> &a12)
{
    const double pi = 3.14159265358979323846264338327950288;
    const double warped = 4.0 * std::tan(pi * wn / 2.0);
    // analog prototype poles -exp(i pi m / (2 order)), m = -order+1, ..., order-1 step 2
    std::vector<std::complex<double>> zp;
    for (int m = -order + 1; m < order; m += 2) {
        const std::complex<double> pa = -std::exp(std::complex<double>(0.0, pi * m / (2.0 * order)));
        const std::complex<double> s = warped * pa;
        const std::complex<double> z = (4.0 + s) / (4.0 - s);
        if (z.imag() > 0.0) zp.push_back(z);  // one of each conjugate pair
    }
    // zpk2sos 'nearest': the pole closest to the unit circle goes last
    std::sort(zp.begin(), zp.end(), [](const std::complex<double> &x, const std::complex<double> &y) {
        return std::fabs(1.0 - std::abs(x)) > std::fabs(1.0 - std::abs(y));
    });
    double g = 1.0;
    a12.assign(order, 0.0);
    for (int s = 0; s < order / 2; ++s) {
        const double a1 = -2.0 * zp[s].real();
        const double a2 = std::norm(zp[s]);
        a12[2 * s] = a1;
        a12[2 * s + 1] = a2;
        g *= (1.0 + a1 + a2) / 4.0;  // unit DC gain per section
    }
    return g;
}

// Direct form of b0 = a0 = 1 FP64 rows: the polynomial products of the
// numerators and denominators; b[j], a[j] (j < 2 ns) are the taps past the
// leading 1, a[] in the recurrence sign y[n] += a[j] y[n-j-1].
void direct_form(const double *rows, int ns, double *b, double *a)
{
    std::vector<double> pb(1, 1.0), pa(1, 1.0);
    for (int s = 0; s < ns; ++s) {
        const double *r = rows + 6 * s;
        std::vector<double> nb(pb.size() + 2, 0.0), na(pa.size() + 2, 0.0);
        for (size_t i = 0; i < pb.size(); ++i)
            for (int j = 0; j < 3; ++j) nb[i + j] += pb[i] * r[j];
        for (size_t i = 0; i < pa.size(); ++i)
            for (int j = 0; j < 3; ++j) na[i + j] += pa[i] * r[3 + j];
        pb.swap(nb);
        pa.swap(na);
    }
    for (int j = 0; j < 2 * ns; ++j) {
        b[j] = pb[j + 1];
        a[j] = -pa[j + 1];
    }
}

}  // namespace

extern "C" {

int iirx_fill_normal(uint64_t seed, int64_t n, float *out)
{
    if (n < 0 || (!out && n > 0)) return IIRX_ERR_INVALID;
    uint64_t s = seed;
    const double two_pi = 6.283185307179586476925286766559;
    for (int64_t i = 0; i < n; i += 2) {
        const double u1 = 1.0 - unit53(splitmix64(s));  // (0, 1]
        const double u2 = unit53(splitmix64(s));         // [0, 1)
        const double r = std::sqrt(-2.0 * std::log(u1));
        out[i] = (float)(r * std::cos(two_pi * u2));
        if (i + 1 < n) out[i + 1] = (float)(r * std::sin(two_pi * u2));
    }
    return IIRX_OK;
}

int iirx_normalize_sos(const double *in, int ns, float *out, double *gain)
{
    if (!in || !out || !gain || ns < 1) return IIRX_ERR_INVALID;
    double g = 1.0;
    for (int s = 0; s < ns; ++s) {
        const double *r = in + 6 * s;
        for (int k = 0; k < 6; ++k)
            if (!std::isfinite(r[k])) return IIRX_ERR_INVALID;
        if (r[0] == 0.0 || r[3] == 0.0) return IIRX_ERR_INVALID;
    }
    for (int s = 0; s < ns; ++s) {
        const double *r = in + 6 * s;
        float *o = out + 6 * s;
        o[0] = 1.0f;
        o[1] = (float)(r[1] / r[0]);
        o[2] = (float)(r[2] / r[0]);
        o[3] = 1.0f;
        o[4] = (float)(r[4] / r[3]);
        o[5] = (float)(r[5] / r[3]);
        g *= r[0] / r[3];
    }
    *gain = g;
    return IIRX_OK;
}

int iirx_butterworth_sos(int order, double wn, float *sos, double *gain)
{
    if (order < 2 || order > 64 || (order & 1) || !(wn > 0.0 && wn < 1.0) || !sos) return IIRX_ERR_INVALID;
    std::vector<double> a12;
    const double g = butterworth_sections(order, wn, a12);
    for (int s = 0; s < order / 2; ++s) {
        float *r = sos + 6 * s;
        r[0] = 1.0f;
        r[1] = 2.0f;
        r[2] = 1.0f;
        r[3] = 1.0f;
        r[4] = (float)a12[2 * s];
        r[5] = (float)a12[2 * s + 1];
    }
    if (gain) *gain = g;
    return IIRX_OK;
}

int iirx_direct_form_taps(const float *sos, int ns, double *b, double *a)
{
    if (!sos || ns < 1 || ns > 16 || !b || !a) return IIRX_ERR_INVALID;
    std::vector<double> rows(6 * (size_t)ns);
    for (int s = 0; s < ns; ++s) {
        const float *r = sos + 6 * s;
        const double row[6] = {1.0, (double)r[1] / (double)r[0], (double)r[2] / (double)r[0],
                               1.0, (double)r[4] / (double)r[3], (double)r[5] / (double)r[3]};
        std::copy(row, row + 6, rows.begin() + 6 * s);
    }
    direct_form(rows.data(), ns, b, a);
    return IIRX_OK;
}

int iirx_filter_taps_design(int n_sections, double cutoff_wn, iirx_filter_taps *t)
{
    if (!t || n_sections < 1 || n_sections > IIRX_TAPS_MAX_SECTIONS || !(cutoff_wn > 0.0 && cutoff_wn < 1.0))
        return IIRX_ERR_INVALID;
    *t = iirx_filter_taps{};
    std::vector<double> a12;
    t->gain = butterworth_sections(2 * n_sections, cutoff_wn, a12);
    t->n_sections = n_sections;
    t->plr_order = 2 * n_sections;
    t->input_fs = 1000.0;  // ref_generate.py:31-32
    t->input_freq = 120.0;
    t->cutoff_wn = cutoff_wn;
    std::vector<double> rows(6 * (size_t)n_sections);
    for (int s = 0; s < n_sections; ++s) {
        const double row[6] = {1.0, 2.0, 1.0, 1.0, a12[2 * s], a12[2 * s + 1]};
        std::copy(row, row + 6, rows.begin() + 6 * s);
        for (int k = 0; k < 6; ++k) t->sos[6 * s + k] = (float)row[k];
    }
    // the direct form of the same (unrounded) transfer function, as
    // ref_generate.py:74-81 takes it from the FP64 design
    direct_form(rows.data(), n_sections, t->plr_b, t->plr_a);
    return IIRX_OK;
}

}  // extern "C"
