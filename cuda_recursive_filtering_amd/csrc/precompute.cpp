// precompute.cpp -- FP64 coefficient tables, see precompute.hpp.
#include "precompute.hpp"

#include "kernels.hpp"

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
            if (l == lanes - 1) put(t + Y.kM, Q);
            Q = m2_mul(Q, C);
        }
        for (int n = 0; n < N; ++n) {
            t[Y.kH1 + n] = (float)h1[n];
            t[Y.kH2 + n] = (float)h2[n];
        }
    }
    return Y.stride;
}

}  // namespace iirx
