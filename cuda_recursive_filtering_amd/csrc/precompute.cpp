// precompute.cpp -- FP64 coefficient tables, see precompute.hpp.
#include "precompute.hpp"

#include "kernels.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

namespace iirx {

void zir_tables(double a1, double a2, int N, std::vector<double> &h1, std::vector<double> &h2, M2 &C)
{
    std::vector<double> h0((size_t)N + 1);
    h0[0] = 1.0;
    h0[1] = a1;
    for (int n = 2; n <= N; ++n) h0[n] = a1 * h0[n - 1] + a2 * h0[n - 2];
    h1.assign((size_t)N, 0.0);
    h2.assign((size_t)N, 0.0);
    for (int n = 0; n < N; ++n) {
        h1[n] = h0[n + 1];
        h2[n] = a2 * h0[n];
    }
    C = {h2[N - 2], h1[N - 2], h2[N - 1], h1[N - 1]};
}

int lookback_depth(const M2 &M)
{
    M2 P = M;
    for (int d = 1; d <= 64; ++d) {
        const double n = std::max(std::fabs(P.a) + std::fabs(P.b), std::fabs(P.c) + std::fabs(P.d));
        if (n <= kLookbackTail) return d;
        P = m2_mul(P, M);
    }
    return 0;
}

static void put(float *dst, const M2 &m)
{
    dst[0] = (float)m.a;
    dst[1] = (float)m.b;
    dst[2] = (float)m.c;
    dst[3] = (float)m.d;
}

int ph_tables(const float *sos, int ns, int lanes, int N, std::vector<float> &out)
{
    const PhLayout Y = ph_layout(lanes, N);
    out.assign((size_t)Y.stride * ns, 0.0f);
    const int lg = ilog2_c(lanes);
    for (int s = 0; s < ns; ++s) {
        float *t = out.data() + (size_t)s * Y.stride;
        const float *row = sos + 6 * s;
        const double a1 = -(double)row[4], a2 = -(double)row[5];
        t[Y.kB + 0] = row[1];
        t[Y.kB + 1] = row[2];
        t[Y.kB + 2] = -row[4];
        t[Y.kB + 3] = -row[5];
        std::vector<double> h1, h2;
        M2 C;
        zir_tables(a1, a2, N, h1, h2, C);
        M2 P = C;
        for (int k = 0; k < lg; ++k) {
            put(t + Y.kCpow + 4 * k, P);
            P = m2_mul(P, P);
        }
        M2 Q = C;
        for (int l = 0; l < lanes; ++l) {
            put(t + Y.kCfin + 4 * l, Q);
            if (l == lanes - 1) {
                put(t + Y.kM, Q);
                t[Y.kDepth] = (float)lookback_depth(Q);
            }
            Q = m2_mul(Q, C);
        }
        for (int n = 0; n < N; ++n) {
            t[Y.kH1 + n] = (float)h1[n];
            t[Y.kH2 + n] = (float)h2[n];
        }
    }
    return Y.stride;
}

// 4 x 4 row-major double matrices
static void m4_mul(const double *p, const double *q, double *r)
{
    double t[16];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            double acc = 0.0;
            for (int k = 0; k < 4; ++k) acc += p[4 * i + k] * q[4 * k + j];
            t[4 * i + j] = acc;
        }
    std::memcpy(r, t, sizeof(t));
}

static void put16(float *dst, const double *m)
{
    for (int i = 0; i < 16; ++i) dst[i] = (float)m[i];
}

// smallest d in [1, 64] with ||M^d||_inf <= 2^-36 for a 4 x 4 M, 0 if none
static int lookback_depth4(const double *M)
{
    double P[16];
    std::memcpy(P, M, sizeof(P));
    for (int d = 1; d <= 64; ++d) {
        double nrm = 0.0;
        for (int i = 0; i < 4; ++i)
            nrm = std::max(nrm, std::fabs(P[4 * i]) + std::fabs(P[4 * i + 1]) + std::fabs(P[4 * i + 2]) +
                                    std::fabs(P[4 * i + 3]));
        if (nrm <= kLookbackTail) return d;
        m4_mul(P, M, P);
    }
    return 0;
}

void ph_pair_response(const float *sos, int s, int N, double J[16], std::vector<double> &H)
{
    const float *rs = sos + 6 * s, *rt = sos + 6 * (s + 1);
    const double as1 = -(double)rs[4], as2 = -(double)rs[5];
    const double bt1 = rt[1], bt2 = rt[2], at1 = -(double)rt[4], at2 = -(double)rt[5];
    H.assign((size_t)4 * N, 0.0);
    for (int m = 0; m < 4; ++m) {
        // zero input from the unit joint state e_m: section s's output is
        // section t's input, whose history is section s's history
        double ys2 = m == 0, ys1 = m == 1, yt2 = m == 2, yt1 = m == 3;
        for (int n = 0; n < N; ++n) {
            const double ys = as1 * ys1 + as2 * ys2;
            const double yt = ys + bt1 * ys1 + bt2 * ys2 + at1 * yt1 + at2 * yt2;
            H[(size_t)m * N + n] = yt;
            ys2 = ys1;
            ys1 = ys;
            yt2 = yt1;
            yt1 = yt;
        }
        J[0 * 4 + m] = ys2;
        J[1 * 4 + m] = ys1;
        J[2 * 4 + m] = yt2;
        J[3 * 4 + m] = yt1;
    }
}

int ph_joint_tables(const float *sos, int ns, int lanes, int N, std::vector<float> &out)
{
    const int stride = ph_tables(sos, ns, lanes, N, out);
    const PhPairLayout P = ph_pair_layout(lanes, N);
    const size_t base = out.size();
    out.resize(base + (size_t)P.stride * (size_t)(ns / 2), 0.0f);
    for (int p = 0; p < ns / 2; ++p) {
        float *t = out.data() + base + (size_t)p * P.stride;
        double J[16], Q[16], R[16];
        std::vector<double> H;
        ph_pair_response(sos, 2 * p, N, J, H);
        std::memcpy(Q, J, sizeof(Q));
        for (int k = 0; k < 5; ++k) {
            put16(t + P.kJpow + 16 * k, Q);
            m4_mul(Q, Q, Q);
        }
        double big = 0.0;
        std::memcpy(R, J, sizeof(R));
        for (int l = 0; l < lanes; ++l) {
            for (int i = 0; i < 16; ++i) big = std::max(big, std::fabs(R[i]));
            if (l < 16) put16(t + P.kJrow + 16 * l, R);
            put16(t + P.kJfin + 16 * l, R);
            if (l == lanes - 1) {
                put16(t + P.kM, R);
                t[P.kDepth] = (float)lookback_depth4(R);
            }
            m4_mul(R, J, R);
        }
        for (int i = 0; i < 4 * N; ++i) {
            t[P.kH + i] = (float)H[(size_t)i];
            big = std::max(big, std::fabs(H[(size_t)i]));
        }
        t[P.kDepth + 1] = (big <= kPairLimit) ? 1.0f : 0.0f;
    }
    return stride;
}

// one zero-input step of the cascade; st = (x[-2], x[-1], y_1[-2], y_1[-1], ...)
static void cascade_step(const float *sos, int ns, double *st, double x)
{
    double u2 = st[0], u1 = st[1], u = x;  // this section's input history and input
    st[0] = st[1];
    st[1] = x;
    for (int s = 0; s < ns; ++s) {
        const float *r = sos + 6 * s;
        double &y2 = st[2 + 2 * s], &y1 = st[3 + 2 * s];
        const double y = u + (double)r[1] * u1 + (double)r[2] * u2 - (double)r[4] * y1 - (double)r[5] * y2;
        u2 = y2;
        u1 = y1;
        u = y;
        y2 = y1;
        y1 = y;
    }
}

double cascade_transition_norm(const float *sos, int ns, int n)
{
    const int K = 2 + 2 * ns;
    std::vector<double> A((size_t)K * K), st((size_t)K);
    for (int m = 0; m < K; ++m) {  // column m: the state after n steps from e_m
        std::fill(st.begin(), st.end(), 0.0);
        st[(size_t)m] = 1.0;
        for (int k = 0; k < n; ++k) cascade_step(sos, ns, st.data(), 0.0);
        for (int r = 0; r < K; ++r) A[(size_t)r * K + m] = st[(size_t)r];
    }
    double nrm = 0.0;
    for (int r = 0; r < K; ++r) {
        double row = 0.0;
        for (int m = 0; m < K; ++m) row += std::fabs(A[(size_t)r * K + m]);
        nrm = std::max(nrm, row);
    }
    return nrm;
}

int ph_halo_tables(const float *sos, int ns, int lanes, int N, std::vector<float> &out, bool *forgets)
{
    const int stride = ph_tables(sos, ns, lanes, N, out);
    const size_t base = out.size();
    out.resize(base + (size_t)ph_halo_floats(ns), 0.0f);
    float *t = out.data() + base;
    const double nrm = cascade_transition_norm(sos, ns, kPhHalo);
    const bool ok = std::isfinite(nrm) && nrm <= kHaloTail;
    t[0] = ok ? 1.0f : 0.0f;
    if (forgets) *forgets = ok;
    // h_s[k]: output of section s at sample k from a unit impulse at sample 0
    std::vector<double> st((size_t)(2 + 2 * ns), 0.0), h((size_t)ns * kPhHalo);
    for (int k = 0; k < kPhHalo; ++k) {
        cascade_step(sos, ns, st.data(), k == 0 ? 1.0 : 0.0);
        for (int s = 0; s < ns; ++s) h[(size_t)s * kPhHalo + k] = st[(size_t)(3 + 2 * s)];
    }
    for (int s = 0; s < ns; ++s) {
        float *k2 = t + 4 + 2 * kPhHalo * s, *k1 = k2 + kPhHalo;
        for (int i = 0; i < kPhHalo; ++i) {
            k1[i] = (float)h[(size_t)s * kPhHalo + kPhHalo - 1 - i];
            k2[i] = i <= kPhHalo - 2 ? (float)h[(size_t)s * kPhHalo + kPhHalo - 2 - i] : 0.0f;
        }
    }
    return stride;
}

}  // namespace iirx

namespace iirx {

int cr_tables(const float *sos, int ns, int lanes, int N, std::vector<float> &out, double *worst)
{
    const CrLayout Y = cr_layout(lanes);
    const int LG = ilog2_c(N);
    out.assign((size_t)Y.stride * ns, 0.0f);
    double wmax = 0.0;
    bool bad = false;
    for (int s = 0; s < ns; ++s) {
        float *t = out.data() + (size_t)s * Y.stride;
        const float *row = sos + 6 * s;
        const double b1 = row[1], b2 = row[2];
        const double a1 = -(double)row[4], a2 = -(double)row[5];
        double e[9], f[9], fde[9], h[9], g[9], p[9], q[9], d[9], c[9];
        for (int i = 0; i < 9; ++i) e[i] = f[i] = fde[i] = h[i] = g[i] = p[i] = q[i] = d[i] = c[i] = 0.0;
        f[0] = -a2;
        e[0] = -a1;
        fde[0] = f[0] / e[0];
        h[0] = f[0];
        g[0] = e[0];
        p[1] = f[0];
        q[1] = e[0];
        for (int n = 1; n <= LG; ++n) {
            f[n] = f[n - 1] * f[n - 1];
            e[n] = 2.0 * f[n - 1] - e[n - 1] * e[n - 1];
            fde[n] = f[n] / e[n];
            h[n] = -e[n - 1] * h[n - 1];
            g[n] = f[n - 1] - e[n - 1] * g[n - 1];
            d[n] = -f[n - 1] * f[n - 1] / e[n - 1];
            c[n] = e[n - 1] - f[n - 1] / e[n - 1];
            if (n > 1) {
                q[n] = -e[n - 1] * q[n - 1];
                p[n] = f[n - 1] - e[n - 1] * p[n - 1];
            }
        }
        for (int n = 1; n < LG; ++n) wmax = std::max(wmax, std::fabs(fde[n]));
        for (int n = 1; n <= LG; ++n) wmax = std::max({wmax, std::fabs(d[n]), std::fabs(c[n])});
        t[Y.kScal + 0] = (float)b1;
        t[Y.kScal + 1] = (float)b2;
        t[Y.kScal + 2] = (float)(f[0] * b1 - e[0] * b2);   // cb
        t[Y.kScal + 3] = (float)(b1 - e[0]);               // db
        t[Y.kScal + 4] = (float)(b2 * f[0]);               // fb
        t[Y.kScal + 5] = (float)(b2 + f[0] - b1 * e[0]);   // eb
        t[Y.kScal + 6] = (float)(b2 - e[0] * b1);          // hb (chunk-start odd side)
        t[Y.kScal + 7] = (float)(-e[0] * b2);              // gb
        for (int n = 0; n <= LG; ++n) {
            t[Y.kE + n] = (float)e[n];
            t[Y.kF + n] = (float)f[n];
            t[Y.kFde + n] = (float)fde[n];
            t[Y.kH + n] = (float)h[n];
            t[Y.kG + n] = (float)g[n];
            t[Y.kP + n] = (float)p[n];
            t[Y.kQ + n] = (float)q[n];
            t[Y.kD + n] = (float)d[n];
            t[Y.kC + n] = (float)c[n];
        }
        // lane recurrence Y_t = R_t - e_LG Y_{t-1} - f_LG Y_{t-2}
        const M2 A = {0.0, 1.0, -f[LG], -e[LG]};
        M2 P = A;
        for (int k = 0; k < ilog2_c(lanes); ++k) {
            put(t + Y.kApow + 4 * k, P);
            P = m2_mul(P, P);
        }
        // impulse response over lanes and the terminal corrections (stcr:127-147)
        std::vector<double> h0((size_t)lanes);
        h0[0] = 1.0;
        if (lanes > 1) h0[1] = -e[LG];
        for (int l = 2; l < lanes; ++l) h0[l] = -e[LG] * h0[l - 1] - f[LG] * h0[l - 2];
        for (int l = 0; l < lanes; ++l) {
            const double prev = (l == 0) ? 0.0 : h0[l - 1];
            t[Y.kTerm + 4 * l + 0] = (float)(f[LG] * prev + p[LG] * h0[l]);  // he2
            t[Y.kTerm + 4 * l + 1] = (float)(q[LG] * h0[l]);                 // he1
            t[Y.kTerm + 4 * l + 2] = (float)(h[LG] * h0[l]);                 // hb2
            t[Y.kTerm + 4 * l + 3] = (float)(f[LG] * prev + g[LG] * h0[l]);  // hb1
        }
        // chunk transition of the section state for the lookback
        std::vector<double> hh1, hh2;
        M2 C;
        zir_tables(a1, a2, N, hh1, hh2, C);
        M2 Q = m2_eye();
        for (int l = 0; l < lanes; ++l) Q = m2_mul(Q, C);
        put(t + Y.kM, Q);
        t[Y.kDepth] = (float)lookback_depth(Q);
        for (int i = 0; i < Y.stride; ++i)
            if (!std::isfinite(t[i])) bad = true;
    }
    if (worst) *worst = wmax;
    if (bad || wmax > kCrLimit) return -1;
    return Y.stride;
}

}  // namespace iirx

namespace iirx {

int plr_tables(const double *b, const double *a, int K, int CH, std::vector<float> &out)
{
    const PlrLayout Y = plr_layout(K, CH);
    out.assign((size_t)Y.stride, 0.0f);
    std::vector<double> af((size_t)K);
    for (int j = 0; j < K; ++j) {
        out[Y.kB + j] = (float)b[j];
        af[j] = (double)(float)a[j];
    }
    // n-nacci factors (plr_kernels.cuh:149-409): fac[m][i], history slot m = y[-K+m]
    std::vector<double> v((size_t)(CH + K));
    std::vector<double> F((size_t)K * K);
    for (int m = 0; m < K; ++m) {
        std::fill(v.begin(), v.end(), 0.0);
        v[m] = 1.0;
        for (int i = K; i < CH + K; ++i) {
            double acc = 0.0;
            for (int l = 0; l < K; ++l) acc += af[l] * v[i - 1 - l];
            v[i] = acc;
        }
        for (int i = 0; i < CH; ++i) out[(size_t)Y.kFac + (size_t)m * CH + i] = (float)v[i + K];
        for (int j = 0; j < K; ++j) F[(size_t)j * K + m] = v[CH + j];  // fac[m][CH-K+j]
    }
    for (int j = 0; j < K * K; ++j) out[Y.kF + j] = (float)F[j];
    // depth: smallest d with ||F^d||_inf <= 2^-36 (0 = unbounded)
    std::vector<double> P(F), T((size_t)K * K);
    int depth = 0;
    for (int d = 1; d <= 64; ++d) {
        double nrm = 0.0;
        for (int r = 0; r < K; ++r) {
            double s = 0.0;
            for (int c = 0; c < K; ++c) s += std::fabs(P[(size_t)r * K + c]);
            nrm = std::max(nrm, s);
        }
        if (!(nrm > kLookbackTail)) {
            depth = std::isfinite(nrm) ? d : 0;
            break;
        }
        for (int r = 0; r < K; ++r)
            for (int c = 0; c < K; ++c) {
                double s = 0.0;
                for (int q = 0; q < K; ++q) s += P[(size_t)r * K + q] * F[(size_t)q * K + c];
                T[(size_t)r * K + c] = s;
            }
        P.swap(T);
    }
    out[Y.kDepth] = (float)depth;
    return Y.stride;
}

}  // namespace iirx

namespace iirx {

namespace {

using Mat = std::vector<double>;  // 16 x 16 row-major

Mat mat_mul16(const Mat &a, const Mat &b)
{
    Mat c(256, 0.0);
    for (int i = 0; i < 16; ++i)
        for (int k = 0; k < 16; ++k) {
            const double v = a[i * 16 + k];
            if (v == 0.0) continue;
            for (int j = 0; j < 16; ++j) c[i * 16 + j] += v * b[k * 16 + j];
        }
    return c;
}

double norm_inf16(const Mat &m)
{
    double nrm = 0.0;
    for (int r = 0; r < 16; ++r) {
        double s = 0.0;
        for (int c = 0; c < 16; ++c) s += std::fabs(m[r * 16 + c]);
        nrm = std::max(nrm, s);
    }
    return nrm;
}

// scale exponent s so that max|m| * 2^s < 2^14 (fp16 holds up to 65504)
int pow2_scale(const double *m, size_t n)
{
    double mx = 0.0;
    for (size_t i = 0; i < n; ++i) mx = std::max(mx, std::fabs(m[i]));
    if (!(mx > 0.0) || !std::isfinite(mx)) return 0;
    int e;
    std::frexp(mx, &e);  // mx = f * 2^e, f in [0.5, 1)
    return 14 - e;
}

uint16_t half_bits(_Float16 h)
{
    uint16_t u;
    std::memcpy(&u, &h, 2);
    return u;
}

// FP16 A operand of v_mfma_f32_16x16x32_f16 (kernels.hpp: bs_layout): rows
// r0..r0+15 and K slab q of matrix m (rows x cols, row major, zero outside),
// scaled by 2^s, hi (hl = 0) or lo (1) part.  With perm, operand row i holds
// matrix row r0 + bs_state_of_row(i): the product's row 4 g + j then lands on
// lane group g, register j as state g + 4 j, the FP64 MFMA's order.
void put_block32(float *dst, const double *m, int rows, int cols, int r0, int q, int s, int hl, bool perm = false,
                 int split = kBsSplitF16)
{
    uint16_t *h = reinterpret_cast<uint16_t *>(dst);
    for (int l = 0; l < 64; ++l)
        for (int j = 0; j < 8; ++j) {
            const int g = l / 16, r = r0 + (perm ? bs_state_of_row(l % 16) : l % 16);
            const int c = 32 * q + (j < 4 ? 4 * g + j : 16 + 4 * g + j - 4);
            const double v = (r < rows && c < cols) ? std::ldexp(m[(size_t)r * cols + c], s) : 0.0;
            if (split == kBsSplitBF16) {
                h[l * 8 + j] = bf16x3_bits(v, hl);
                continue;
            }
            // hi = v rounded once to FP16, lo = the exact remainder rounded once
            const _Float16 hi = (_Float16)v;
            const _Float16 lo = (_Float16)(v - (double)hi);
            h[l * 8 + j] = half_bits(hl == 0 ? hi : lo);
        }
}

// FP32 A operand of a 16 x 16 x 4 MFMA (f32 or f64) whose K slab m holds
// states g + 4 m: word 4 l + m = M[r0 + l%16][l/16 + 4 m] of the 16 x 16
// block at row r0 of m (rows x 16)
void put_f32(float *dst, const double *m, int r0)
{
    for (int l = 0; l < 64; ++l)
        for (int k = 0; k < 4; ++k) dst[4 * l + k] = (float)m[(size_t)(r0 + l % 16) * 16 + (l / 16) + 4 * k];
}

}  // namespace

// The block-state realization (FP64): the cascade of direct-form-II-
// transposed sections, z_s' = A_s z_s + B_s u_s with z_s = (w1, w2) of
// section s at states 2s, 2s+1, u_s = u_0 + sum_{r<s} z_r[0],
// y = u_0 + sum_r z_r[0]; then every state divided by d_c = sqrt(W_cc), W
// the controllability Gramian (balancing).  A is 16 x 16 row-major, B, C
// and d 16 long (d = 1 past the p = 2 ns states).
void bs_realization(const float *sos, int ns, std::vector<double> &A, std::vector<double> &B, std::vector<double> &C,
                    double *d)
{
    const int p = 2 * ns;
    A.assign(256, 0.0);
    B.assign(16, 0.0);
    C.assign(16, 0.0);
    // cascade of direct-form-II-transposed sections: z_s' = A_s z_s + B_s u_s,
    // u_s = u_0 + sum_{r<s} z_r[0], y = u_0 + sum_r z_r[0]
    for (int s = 0; s < ns; ++s) {
        const float *r = sos + 6 * s;
        const double b1 = r[1], b2 = r[2], a1 = r[4], a2 = r[5];  // scipy sign
        const double B0 = b1 - a1, B1 = b2 - a2;
        A[(2 * s) * 16 + 2 * s] = -a1;
        A[(2 * s) * 16 + 2 * s + 1] = 1.0;
        A[(2 * s + 1) * 16 + 2 * s] = -a2;
        for (int q = 0; q < s; ++q) {
            A[(2 * s) * 16 + 2 * q] += B0;
            A[(2 * s + 1) * 16 + 2 * q] += B1;
        }
        B[2 * s] = B0;
        B[2 * s + 1] = B1;
        C[2 * s] = 1.0;
    }
    // state balancing: scale state component c by d_c = sqrt(W_cc), W the
    // controllability Gramian sum_k A^k B B^T A^kT (doubling, FP64), so that
    // every state has the scale of the response to unit white input.  Without
    // it the chained sections' states span ~14 decades at order 16 and the
    // FP16 operands lose the small ones.
    {
        Mat W(256, 0.0), Ak(A);
        for (int r = 0; r < 16; ++r)
            for (int c = 0; c < 16; ++c) W[r * 16 + c] = B[r] * B[c];
        for (int it = 0; it < 48; ++it) {
            Mat AW = mat_mul16(Ak, W);
            Mat AkT(256);
            for (int r = 0; r < 16; ++r)
                for (int c = 0; c < 16; ++c) AkT[r * 16 + c] = Ak[c * 16 + r];
            const Mat add = mat_mul16(AW, AkT);
            double delta = 0.0, size = 0.0;
            for (int i = 0; i < 256; ++i) {
                W[i] += add[i];
                delta = std::max(delta, std::fabs(add[i]));
                size = std::max(size, std::fabs(W[i]));
            }
            Ak = mat_mul16(Ak, Ak);
            if (!(delta > 1e-18 * size)) break;
        }
        double *dsc = d;
        for (int c = 0; c < 16; ++c) {
            const double w = W[c * 16 + c];
            dsc[c] = (c < p && w > 0.0 && std::isfinite(w)) ? std::sqrt(w) : 1.0;
        }
        for (int r = 0; r < 16; ++r)
            for (int c = 0; c < 16; ++c) A[r * 16 + c] *= dsc[c] / dsc[r];
        for (int c = 0; c < 16; ++c) {
            B[c] /= dsc[c];
            C[c] *= dsc[c];
        }
    }
}

int bs_tables(const float *sos, int ns, int L, int G, std::vector<float> &out, int split)
{
    const BsLayout Y = bs_layout(L, split);
    out.assign((size_t)Y.stride, 0.0f);
    const int p = 2 * ns;
    Mat A;
    std::vector<double> B, C;
    double dsc[16];
    bs_realization(sos, ns, A, B, C, dsc);
    // states stay in order 0..p-1: K slab m of the state products holds states
    // 4 m .. 4 m + 3, so only ceil(p / 4) slabs are nonzero
    // impulse response, observability, controllability
    std::vector<double> h((size_t)L), Om((size_t)L * 16), Gm((size_t)16 * L), Tm((size_t)L * L, 0.0);
    {
        std::vector<double> cA(C);  // C A^i
        std::vector<double> AB(B);  // A^i B
        std::vector<std::vector<double>> powB((size_t)L);
        for (int i = 0; i < L; ++i) {
            for (int c = 0; c < 16; ++c) Om[(size_t)i * 16 + c] = cA[c];
            powB[i] = AB;
            double hv = 0.0;
            for (int c = 0; c < 16; ++c) hv += C[c] * AB[c];
            if (i + 1 < L) h[i + 1] = hv;  // h[m] = C A^(m-1) B
            std::vector<double> n1(16, 0.0), n2(16, 0.0);
            for (int r = 0; r < 16; ++r)
                for (int c = 0; c < 16; ++c) {
                    n1[c] += cA[r] * A[r * 16 + c];
                    n2[r] += A[r * 16 + c] * AB[c];
                }
            cA.swap(n1);
            AB.swap(n2);
        }
        h[0] = 1.0;
        for (int j = 0; j < L; ++j)
            for (int c = 0; c < 16; ++c) Gm[(size_t)c * L + j] = powB[L - 1 - j][c];
        for (int i = 0; i < L; ++i)
            for (int j = 0; j <= i; ++j) Tm[(size_t)i * L + j] = h[i - j];
    }
    // Phi = A^L and its powers
    Mat Phi(256, 0.0);
    for (int i = 0; i < 16; ++i) Phi[i * 16 + i] = 1.0;
    for (int i = 0; i < L; ++i) Phi = mat_mul16(Phi, A);
    std::vector<Mat> P(6);
    P[0] = Phi;
    for (int j = 1; j < 6; ++j) P[j] = mat_mul16(P[j - 1], P[j - 1]);
    Mat Phi16 = mat_mul16(P[3], P[3]);  // Phi^16
    std::vector<Mat> Q(4);
    Q[1] = Phi16;
    Q[2] = mat_mul16(Phi16, Phi16);
    Q[3] = mat_mul16(Q[2], Phi16);
    const Mat Mc = G == 16 ? mat_mul16(P[3], P[3]) : (G == 32 ? mat_mul16(P[4], P[4]) : mat_mul16(P[5], P[5]));  // Phi^G

    float *sc = out.data() + Y.kScale;
    const int sT = pow2_scale(Tm.data(), Tm.size()), sG = pow2_scale(Gm.data(), Gm.size());
    sc[0] = (float)sT;
    sc[1] = (float)sG;
    for (int hl = 0; hl < split; ++hl) {
        for (int mb = 0; mb < Y.MB; ++mb)
            for (int q = 0; q < Y.KQ; ++q)
                put_block32(out.data() + bs_T(Y, hl, mb, q), Tm.data(), L, L, 16 * mb, q, sT, hl, false, split);
        for (int q = 0; q < Y.KQ; ++q)
            put_block32(out.data() + bs_G(Y, hl, q), Gm.data(), 16, L, 0, q, sG, hl, true, split);
    }
    for (int j = 0; j < 6; ++j) put_f32(out.data() + bs_P(Y, j), P[j].data(), 0);
    for (int nb = 1; nb < 4; ++nb) put_f32(out.data() + bs_Q(Y, nb), Q[nb].data(), 0);
    put_f32(out.data() + Y.kPhi, Phi.data(), 0);
    for (int mb = 0; mb < Y.MB; ++mb) put_f32(out.data() + bs_O(Y, mb), Om.data(), 16 * mb);
    for (int i = 0; i < 256; ++i) out[Y.kMc + i] = (float)Mc[i];
    // lookback depth of the chunk transition
    Mat Pw = Mc;
    int depth = 0;
    for (int d = 1; d <= 64; ++d) {
        double nrm = 0.0;
        for (int r = 0; r < 16; ++r) {
            double s = 0.0;
            for (int c = 0; c < 16; ++c) s += std::fabs(Pw[r * 16 + c]);
            nrm = std::max(nrm, s);
        }
        if (!(nrm > kLookbackTail)) {
            depth = std::isfinite(nrm) ? d : 0;
            break;
        }
        Pw = mat_mul16(Pw, Mc);
    }
    out[Y.kDepth] = (float)depth;
    // levels of the in-chunk tile scan (kernels.hpp: kBsScaleFp32Level,
    // kBsScaleLevels): the lags i >= 2^j are dropped once sum_{i=2^j}^{G}
    // ||Phi^i||_inf <= 2^-36, and the levels from the first one whose
    // Phi^(2^j), and every higher one, has ||.||_inf <= 2^-4 run on the FP32
    // matrix core
    {
        const int lvG = bs_scan_levels(G);
        std::vector<double> nrm((size_t)G + 1, 0.0);
        Mat Pi = Phi;
        bool finite = true;
        for (int i = 1; i <= G; ++i) {
            nrm[i] = norm_inf16(Pi);
            finite = finite && std::isfinite(nrm[i]);
            Pi = mat_mul16(Pi, Phi);
        }
        int nlv = lvG, lv32 = lvG;
        if (finite) {
            for (int j = 0; j < lvG; ++j) {
                double tail = 0.0;
                for (int i = 1 << j; i <= G; ++i) tail += nrm[i];
                if (tail <= kLookbackTail) {
                    nlv = j;
                    break;
                }
            }
            for (int j = lvG - 1; j >= 0 && norm_inf16(P[j]) <= kBsFp32Norm; --j) lv32 = j;
        }
        out[Y.kScale + kBsScaleFp32Level] = (float)lv32;
        out[Y.kScale + kBsScaleLevels] = (float)nlv;
    }
    return Y.stride;
}

}  // namespace iirx
