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
// iirx_chebyshev_sos restates scipy.signal.cheby1 / cheby2 in the same way
// (no reference counterpart; iir_utils.hpp:27-141 holds the reference's
// unused tf2sos route to sections).
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

// Chebyshev lowpass of even `order` (type 1: passband ripple `db`; type 2:
// stopband attenuation `db`) at `wn`, as scipy.signal.cheby1 / cheby2(order,
// db, wn, output='sos') compute it: the analog prototype (cheb1ap: poles
// -sinh(mu + i theta_m), mu = asinh(1 / eps) / order, eps = sqrt(10^(db/10) -
// 1), gain prod(-p) / sqrt(1 + eps^2); cheb2ap: zeros i / sin(theta_m) and
// poles 1 / (sinh(mu) Re q + i cosh(mu) Im q) of the Butterworth q, mu =
// asinh(sqrt(10^(db/10) - 1)) / order, gain prod(-p) / prod(-z)), scaled by
// the pre-warped cutoff 4 tan(pi wn / 2), the bilinear map z = (4 + s) /
// (4 - s) with the gain times prod(4 - z_s) / prod(4 - p_s), then zpk2sos's
// 'nearest' pairing (the pole closest to the unit circle goes last, each with
// the zero nearest to it) and the gain in the first section.  Rows are FP64
// {b0, b1, b2, 1, a1, a2}; the caller normalises them.
void chebyshev_rows(int type, int order, double db, double wn, std::vector<double> &rows)
{
    typedef std::complex<double> cd;
    const double pi = 3.14159265358979323846264338327950288;
    const double wo = 4.0 * std::tan(pi * wn / 2.0);
    std::vector<cd> za, pa;
    double k;
    if (type == 1) {
        const double eps = std::sqrt(std::pow(10.0, 0.1 * db) - 1.0);
        const double mu = 1.0 / order * std::asinh(1.0 / eps);
        for (int m = -order + 1; m < order; m += 2) pa.push_back(-std::sinh(cd(mu, pi * m / (2.0 * order))));
        cd prod = 1.0;
        for (const cd &q : pa) prod *= -q;
        k = prod.real() / std::sqrt(1.0 + eps * eps);
    } else {
        const double de = 1.0 / std::sqrt(std::pow(10.0, 0.1 * db) - 1.0);
        const double mu = std::asinh(1.0 / de) / order;
        for (int m = -order + 1; m < order; m += 2) za.push_back(-std::conj(cd(0.0, 1.0) / std::sin(m * pi / (2.0 * order))));
        for (int m = -order + 1; m < order; m += 2) {
            const cd q = -std::exp(cd(0.0, pi * m / (2 * order)));
            pa.push_back(1.0 / cd(std::sinh(mu) * q.real(), std::cosh(mu) * q.imag()));
        }
        cd pp = 1.0, pz = 1.0;
        for (const cd &q : pa) pp *= -q;
        for (const cd &q : za) pz *= -q;
        k = (pp / pz).real();
    }
    // lp2lp_zpk and bilinear_zpk (fs = 2)
    for (cd &q : za) q *= wo;
    for (cd &q : pa) q *= wo;
    k *= std::pow(wo, (double)(pa.size() - za.size()));
    cd nz = 1.0, np_ = 1.0;
    for (const cd &q : za) nz *= 4.0 - q;
    for (const cd &q : pa) np_ *= 4.0 - q;
    k *= (nz / np_).real();
    std::vector<cd> zd, pd;
    for (const cd &q : za) zd.push_back((4.0 + q) / (4.0 - q));
    for (size_t i = za.size(); i < pa.size(); ++i) zd.push_back(-1.0);
    for (const cd &q : pa) pd.push_back((4.0 + q) / (4.0 - q));
    // one of each conjugate pair (positive imaginary part); real zeros kept
    std::vector<cd> zc, zr, pc;
    for (const cd &q : zd) {
        if (q.imag() > 0.0) zc.push_back(q);
        else if (q.imag() == 0.0) zr.push_back(q);
    }
    for (const cd &q : pd)
        if (q.imag() > 0.0) pc.push_back(q);
    std::vector<cd> z = zc;
    z.insert(z.end(), zr.begin(), zr.end());
    const int ns = order / 2;
    rows.assign(6 * (size_t)ns, 0.0);
    for (int si = ns - 1; si >= 0; --si) {
        // the pole closest to the unit circle (first of equals)
        size_t pi1 = 0;
        for (size_t i = 1; i < pc.size(); ++i)
            if (std::fabs(1.0 - std::abs(pc[i])) < std::fabs(1.0 - std::abs(pc[pi1]))) pi1 = i;
        const cd p1 = pc[pi1];
        pc.erase(pc.begin() + (long)pi1);
        // the nearest zero (first of equals); a real zero takes the next real one
        size_t zi1 = 0;
        for (size_t i = 1; i < z.size(); ++i)
            if (std::abs(z[i] - p1) < std::abs(z[zi1] - p1)) zi1 = i;
        const cd z1 = z[zi1];
        z.erase(z.begin() + (long)zi1);
        double b[3];
        if (z1.imag() != 0.0) {
            b[0] = 1.0;
            b[1] = -(z1 + std::conj(z1)).real();
            b[2] = (z1 * std::conj(z1)).real();
        } else {
            size_t zi2 = 0;
            double best = INFINITY;
            for (size_t i = 0; i < z.size(); ++i)
                if (z[i].imag() == 0.0 && std::abs(z[i] - p1) < best) {
                    best = std::abs(z[i] - p1);
                    zi2 = i;
                }
            const double z2 = z[zi2].real();
            z.erase(z.begin() + (long)zi2);
            b[0] = 1.0;
            b[1] = -(z1.real() + z2);
            b[2] = z1.real() * z2;
        }
        double *r = rows.data() + 6 * si;
        r[0] = b[0];
        r[1] = b[1];
        r[2] = b[2];
        r[3] = 1.0;
        r[4] = -(p1 + std::conj(p1)).real();
        r[5] = (p1 * std::conj(p1)).real();
    }
    for (int j = 0; j < 3; ++j) rows[(size_t)j] *= k;
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

int iirx_chebyshev_sos(int type, int order, double db, double wn, float *sos, double *gain)
{
    if ((type != 1 && type != 2) || order < 2 || order > 16 || (order & 1) || !(db > 0.0 && db < 400.0) ||
        !(wn > 0.0 && wn < 1.0) || !sos)
        return IIRX_ERR_INVALID;
    std::vector<double> rows;
    chebyshev_rows(type, order, db, wn, rows);
    double g = 0.0;
    const int st = iirx_normalize_sos(rows.data(), order / 2, sos, &g);
    if (st == IIRX_OK && gain) *gain = g;
    return st;
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
