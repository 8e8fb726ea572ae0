/*
 * oracle.c -- CPU oracle for the parallel cascaded-biquad IIR path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file; it is used by tests/, by __graft_entry__.smoke() as the
 * checker, and by bench.py's cpu_baseline leg.
 *
 * Contents
 *   orc_sosfilt_f64      float64 cascade (direct form II transposed, the form
 *                        scipy.signal.sosfilt evaluates), the golden reference
 *                        of tools/ref_generate.py:89.
 *   orc_ph_f32           float32 restatement of PH_32/PH_64
 *                        (include/ph_kernels.cuh:55-200 host, 208-820 device).
 *   orc_ph_f32_mt        the same on several host threads (OpenMP over
 *                        chunks), bit-identical; the all-cores CPU baseline.
 *   orc_stcr_f32         float32 restatement of STCR_{32,64}_{LOOP,UNROLL_*}
 *                        (include/stcr_kernels.cuh:78-272 host, 279-3608
 *                        device; operation order as tools/stcr_emulator.py).
 *   orc_dtcr_f32         float32 restatement of DTCR_{32,64}_*
 *                        (include/dtcr_kernels.cuh:92-289 host, 291-3522
 *                        device; operation order as tools/dtcr_emulator.py).
 *   orc_plr_f32          float32 restatement of PLR_{2,4,8,16}
 *                        (include/plr_kernels.cuh:149-476 host, 483-3108
 *                        device; tools/plr_emulator.py:29-160 for order 2).
 *   orc_direct_form_f32  sequential float32 direct form
 *                        (tools/plr_emulator.py:185-202).
 *
 * Rounding model.  Device arithmetic is restated with fused fmaf(), which is
 * what the .cuh kernels execute (explicit fmaf() calls, and `+=` of a
 * product, which nvcc contracts).  The reference's Python emulators round the
 * product and the sum separately (stcr_emulator.py:50), so this file is
 * "operation-order faithful" to them, not bit-identical; agreement is stated
 * with a ULP-scaled tolerance in the tests.  Host-side precompute is restated
 * in unfused float32 (the reference computes it with host `T` arithmetic on
 * x86-64, which has no contraction); this file is compiled with
 * -ffp-contract=off so that the only fused operations are the explicit fmaf().
 *
 * Lookback is emulated serially: every predecessor chunk is complete when a
 * chunk looks back, so the carry is the predecessor's full carry, as in the
 * emulators (stcr_emulator.py:164, dtcr_emulator.py:203-207).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

/*
 * ORC_DOUBLE_ROUND (the liboracle_dr.so build): every fused multiply-add of
 * the restatements rounds the product and then the sum, as the reference's
 * Python emulators evaluate fma(a, b, c) = F32(F32(a) * F32(b) + F32(c))
 * (stcr_emulator.py:50, dtcr_emulator.py:26) and as plr_emulator.py writes
 * its products.  With it the restatements are compared with the emulators'
 * own outputs bit for bit (tests/test_emulator_fixtures.py).
 */
#ifdef ORC_DOUBLE_ROUND
static inline float orc_dr_fma(float a, float b, float c)
{
    const float p = a * b;
    return p + c;
}
#define fmaf orc_dr_fma
#endif

#define ORC_OK 0
#define ORC_EINVAL -1
#define ORC_ENOMEM -2

static int is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }
static int ilog2(int v) { int l = 0; while ((1 << l) < v) l++; return l; }

/* ------------------------------------------------------------------------ */
/* float64 golden: sosfilt semantics (DF-II transposed per section)          */
/* tools/ref_generate.py:89  ref = sosfilt(f64(sos_f32), f64(x_dev))        */
/* ------------------------------------------------------------------------ */
void orc_sosfilt_f64(const double *sos, int ns, const double *x, long n,
                     double *y)
{
    if (x != y) memcpy(y, x, (size_t)n * sizeof(double));
    for (int s = 0; s < ns; s++) {
        const double *c = sos + 6 * s;
        const double a0 = c[3];
        const double b0 = c[0] / a0, b1 = c[1] / a0, b2 = c[2] / a0;
        const double a1 = c[4] / a0, a2 = c[5] / a0;
        double z0 = 0.0, z1 = 0.0;
        for (long i = 0; i < n; i++) {
            const double xi = y[i];
            const double yi = b0 * xi + z0;
            z0 = b1 * xi - a1 * yi + z1;
            z1 = b2 * xi - a2 * yi;
            y[i] = yi;
        }
    }
}

/* float32 sections and float32 input, filtered in float64 (ref_generate.py:89) */
void orc_golden_f64(const float *sos_f32, int ns, const float *x, long n,
                    double *y)
{
    double sos[6 * 64];
    if (ns > 64) return;
    for (int i = 0; i < 6 * ns; i++) sos[i] = (double)sos_f32[i];
    for (long i = 0; i < n; i++) y[i] = (double)x[i];
    orc_sosfilt_f64(sos, ns, y, n, y);
}

/* ------------------------------------------------------------------------ */
/* Shared host precompute, float32 unfused                                    */
/* ------------------------------------------------------------------------ */

/* 2x2 row-major product r = p * q (ph_kernels.cuh:83-90 term order). */
static void mat2_mul(const float *p, const float *q, float *r)
{
    float t[4];
    t[0] = p[0] * q[0] + p[1] * q[2];
    t[1] = p[0] * q[1] + p[1] * q[3];
    t[2] = p[2] * q[0] + p[3] * q[2];
    t[3] = p[2] * q[1] + p[3] * q[3];
    memcpy(r, t, sizeof t);
}

/* C^n by binary exponentiation, ph_kernels.cuh:70-106 (C_power). */
static void mat2_pow(const float *C, int n, float *res)
{
    res[0] = 1.0f; res[1] = 0.0f; res[2] = 0.0f; res[3] = 1.0f;
    if (n == 0) return;
    if (n == 1) { memcpy(res, C, 4 * sizeof(float)); return; }
    float base[4];
    memcpy(base, C, sizeof base);
    while (n > 0) {
        if (n & 1) mat2_mul(res, base, res);
        mat2_mul(base, base, base);
        n >>= 1;
    }
}

/* Zero-input response tables of one section over N samples,
 * ph_kernels.cuh:55-68 (impulse_response): h2[n] = a2*h0[n], h1[n] = h0[n+1].
 * Also returns the N-step state transition C (ph_kernels.cuh:180-181). */
static void ph_tables(float a1, float a2, int N, float *h1, float *h2,
                      float *C)
{
    float *h0 = (float *)malloc((size_t)(N + 1) * sizeof(float));
    h0[0] = 1.0f;
    h0[1] = a1;
    for (int n = 2; n < N + 1; n++) h0[n] = a1 * h0[n - 1] + a2 * h0[n - 2];
    for (int n = 0; n < N; n++) {
        h2[n] = a2 * h0[n];
        h1[n] = h0[n + 1];
    }
    C[0] = h2[N - 2]; C[1] = h1[N - 2];
    C[2] = h2[N - 1]; C[3] = h1[N - 1];
    free(h0);
}

/* ------------------------------------------------------------------------ */
/* PH                                                                         */
/* ------------------------------------------------------------------------ */
/*
 * orc_ph_f32: chunk = B lanes x N samples; lane t owns samples [tN, tN+N).
 * Requires n % (B*N) == 0, B in {32, 64}, N power of two >= 4
 * (the reference has no tail path for the cascaded kernels,
 * test_accuracy.cu:37-38).
 */
int orc_ph_f32(const float *sos, int ns, int B, int N, const float *x, long n,
               float *y)
{
    if (!(B == 32 || B == 64) || !is_pow2(N) || N < 4 || ns < 1) return ORC_EINVAL;
    const long CH = (long)B * N;
    if (n % CH != 0) return ORC_EINVAL;
    const long nch = n / CH;

    float *h1 = malloc(sizeof(float) * (size_t)ns * N);
    float *h2 = malloc(sizeof(float) * (size_t)ns * N);
    float *Cx = malloc(sizeof(float) * (size_t)ns * B * 4); /* C_cross[s][k] = C^(k+1) */
    float *reg = malloc(sizeof(float) * (size_t)CH);
    float *inn = malloc(sizeof(float) * (size_t)CH);
    float *fullc = malloc(sizeof(float) * (size_t)ns * 2);   /* previous chunk's full carry */
    float *yi1 = malloc(sizeof(float) * B), *yi2 = malloc(sizeof(float) * B);
    float *xi1 = malloc(sizeof(float) * B), *xi2 = malloc(sizeof(float) * B);
    if (!h1 || !h2 || !Cx || !reg || !inn || !fullc || !yi1 || !yi2 || !xi1 || !xi2) return ORC_ENOMEM;

    float b1[64], b2[64], a1[64], a2[64];
    if (ns > 64) return ORC_EINVAL;
    for (int s = 0; s < ns; s++) {
        b1[s] = sos[6 * s + 1]; b2[s] = sos[6 * s + 2];
        a1[s] = -sos[6 * s + 4]; a2[s] = -sos[6 * s + 5];
        float C[4];
        ph_tables(a1[s], a2[s], N, h1 + (size_t)s * N, h2 + (size_t)s * N, C);
        for (int k = 0; k < B; k++) mat2_pow(C, k + 1, Cx + ((size_t)s * B + k) * 4);
    }

    for (long cid = 0; cid < nch; cid++) {
        const long cs = cid * CH;
        memcpy(inn, x + cs, sizeof(float) * (size_t)CH);
        for (int s = 0; s < ns; s++) {
            const float *Cs = Cx + (size_t)s * B * 4;
            /* boundary inputs (ph_kernels.cuh:250-263) */
            for (int t = 0; t < B; t++) {
                if (t == 0) {
                    if (cid == 0) { xi2[0] = 0.0f; xi1[0] = 0.0f; }
                    else if (s == 0) { xi2[0] = x[cs - 2]; xi1[0] = x[cs - 1]; }
                    /* else: keep the carry of the previous section */
                } else if (s == 0) {
                    xi2[t] = inn[(size_t)(t - 1) * N + N - 2];
                    xi1[t] = inn[(size_t)(t - 1) * N + N - 1];
                }
            }
            /* hot loop 1: zero-state biquad (ph_kernels.cuh:265-287) */
            for (int t = 0; t < B; t++) {
                float _x2 = xi2[t], _x1 = xi1[t], _y2 = 0.0f, _y1 = 0.0f;
                float *r = reg + (size_t)t * N;
                const float *src = (s == 0) ? inn + (size_t)t * N : r;
                for (int k = 0; k < N; k++) {
                    const float xc = src[k];
                    float yc = fmaf(b1[s], _x1, fmaf(b2[s], _x2, xc));
                    if (k == 1) yc = fmaf(a1[s], _y1, yc);
                    else if (k > 1) yc = fmaf(a1[s], _y1, fmaf(a2[s], _y2, yc));
                    if (k < N - 2) r[k] = yc;
                    _x2 = _x1; _x1 = xc; _y2 = _y1; _y1 = yc;
                }
                yi2[t] = _y2; yi1[t] = _y1;
            }
            /* lane scan: level k, lanes with bit k set take C^(t%2^k+1) times
             * the state of the last lane of the lower half
             * (ph_kernels.cuh:289-382, PH_64 cross-warp 680-694) */
            for (int h = 1; h < B; h <<= 1) {
                for (int t = 0; t < B; t++) {
                    if (!(t & h)) continue;
                    const int src = (t & ~(2 * h - 1)) + h - 1;
                    const float *Cm = Cs + (size_t)(t % h) * 4;
                    const float s2 = yi2[src], s1 = yi1[src];
                    yi2[t] = fmaf(Cm[1], s1, fmaf(Cm[0], s2, yi2[t]));
                    yi1[t] = fmaf(Cm[3], s1, fmaf(Cm[2], s2, yi1[t]));
                }
            }
            /* lookback: predecessor complete -> its full carry (ph:394-451) */
            const float X0 = (cid == 0) ? 0.0f : fullc[2 * s];
            const float X1 = (cid == 0) ? 0.0f : fullc[2 * s + 1];
            /* finalize the last two samples with C_inner = C^(t+1) (ph:453-459) */
            for (int t = 0; t < B; t++) {
                const float *Cm = Cs + (size_t)t * 4;
                float *r = reg + (size_t)t * N;
                r[N - 2] = fmaf(Cm[1], X1, fmaf(Cm[0], X0, yi2[t]));
                r[N - 1] = fmaf(Cm[3], X1, fmaf(Cm[2], X0, yi1[t]));
            }
            /* publish the full carry for the next chunk (ph:461-469) */
            fullc[2 * s] = reg[(size_t)(B - 1) * N + N - 2];
            fullc[2 * s + 1] = reg[(size_t)(B - 1) * N + N - 1];
            /* hot loop 2: back-correction (ph:471-480) */
            for (int t = B - 1; t >= 0; t--) {
                float c2, c1;
                if (t == 0) { c2 = X0; c1 = X1; }
                else { c2 = reg[(size_t)(t - 1) * N + N - 2]; c1 = reg[(size_t)(t - 1) * N + N - 1]; }
                xi2[t] = c2; xi1[t] = c1;
                float *r = reg + (size_t)t * N;
                const float *hh1 = h1 + (size_t)s * N, *hh2 = h2 + (size_t)s * N;
                for (int k = 0; k < N - 2; k++) r[k] = fmaf(hh1[k], c1, fmaf(hh2[k], c2, r[k]));
            }
        }
        memcpy(y + cs, reg, sizeof(float) * (size_t)CH);
    }
    free(h1); free(h2); free(Cx); free(reg); free(inn); free(fullc);
    free(yi1); free(yi2); free(xi1); free(xi2);
    return ORC_OK;
}

/*
 * orc_ph_f32_mt: orc_ph_f32 on `threads` host threads (OpenMP over chunks),
 * the same float operations in the same order, so the output is identical.
 * Per section: (A) every chunk's zero-state pass and lane scan in parallel;
 * (B) the chunk carries serially (each chunk's full carry from its
 * predecessor's, ph:394-469); (C) every chunk's finalisation and
 * back-correction in parallel.  The input history of a chunk's first lane
 * (the previous chunk's last two section inputs) is saved before (A)
 * overwrites it.  Used as the all-cores CPU baseline.
 */
int orc_ph_f32_mt(const float *sos, int ns, int B, int N, const float *x, long n,
                  float *y, int threads)
{
    if (!(B == 32 || B == 64) || !is_pow2(N) || N < 4 || ns < 1 || ns > 64 || threads < 1) return ORC_EINVAL;
    const long CH = (long)B * N;
    if (n % CH != 0) return ORC_EINVAL;
    const long nch = n / CH;
    float *h1 = malloc(sizeof(float) * (size_t)ns * N);
    float *h2 = malloc(sizeof(float) * (size_t)ns * N);
    float *Cx = malloc(sizeof(float) * (size_t)ns * B * 4);
    float *yi = malloc(sizeof(float) * (size_t)nch * B * 2);   /* per chunk: yi2[B], yi1[B] */
    float *bnd = malloc(sizeof(float) * (size_t)nch * 2);      /* lane 0 input history */
    float *xin = malloc(sizeof(float) * (size_t)nch * 2);      /* chunk carry in */
    if (!h1 || !h2 || !Cx || !yi || !bnd || !xin) return ORC_ENOMEM;
    float b1[64], b2[64], a1[64], a2[64];
    for (int s = 0; s < ns; s++) {
        b1[s] = sos[6 * s + 1]; b2[s] = sos[6 * s + 2];
        a1[s] = -sos[6 * s + 4]; a2[s] = -sos[6 * s + 5];
        float C[4];
        ph_tables(a1[s], a2[s], N, h1 + (size_t)s * N, h2 + (size_t)s * N, C);
        for (int k = 0; k < B; k++) mat2_pow(C, k + 1, Cx + ((size_t)s * B + k) * 4);
    }
    if (y != x) memcpy(y, x, sizeof(float) * (size_t)n);
    for (int s = 0; s < ns; s++) {
        const float *Cs = Cx + (size_t)s * B * 4;
        /* the first lane's input history: raw input (section 0) or the
         * previous section's output before the chunk, 0 for chunk 0 */
        for (long c = 0; c < nch; c++) {
            bnd[2 * c] = (c == 0) ? 0.0f : y[c * CH - 2];
            bnd[2 * c + 1] = (c == 0) ? 0.0f : y[c * CH - 1];
        }
        /* (A) zero-state pass and lane scan */
#pragma omp parallel for num_threads(threads) schedule(static)
        for (long c = 0; c < nch; c++) {
            float *reg = y + c * CH;
            float *yi2 = yi + (size_t)c * B * 2, *yi1 = yi2 + B;
            for (int t = B - 1; t >= 0; t--) {
                float _x2, _x1;
                if (t == 0) { _x2 = bnd[2 * c]; _x1 = bnd[2 * c + 1]; }
                else { _x2 = reg[(size_t)(t - 1) * N + N - 2]; _x1 = reg[(size_t)(t - 1) * N + N - 1]; }
                float _y2 = 0.0f, _y1 = 0.0f;
                float *r = reg + (size_t)t * N;
                for (int k = 0; k < N; k++) {
                    const float xc = r[k];
                    float yc = fmaf(b1[s], _x1, fmaf(b2[s], _x2, xc));
                    if (k == 1) yc = fmaf(a1[s], _y1, yc);
                    else if (k > 1) yc = fmaf(a1[s], _y1, fmaf(a2[s], _y2, yc));
                    if (k < N - 2) r[k] = yc;
                    _x2 = _x1; _x1 = xc; _y2 = _y1; _y1 = yc;
                }
                yi2[t] = _y2; yi1[t] = _y1;
            }
            for (int h = 1; h < B; h <<= 1) {
                for (int t = 0; t < B; t++) {
                    if (!(t & h)) continue;
                    const int src = (t & ~(2 * h - 1)) + h - 1;
                    const float *Cm = Cs + (size_t)(t % h) * 4;
                    const float s2 = yi2[src], s1 = yi1[src];
                    yi2[t] = fmaf(Cm[1], s1, fmaf(Cm[0], s2, yi2[t]));
                    yi1[t] = fmaf(Cm[3], s1, fmaf(Cm[2], s2, yi1[t]));
                }
            }
        }
        /* (B) chunk carries: the full carry of chunk c is its last lane's
         * finalised end state */
        {
            const float *Cm = Cs + (size_t)(B - 1) * 4;
            float X0 = 0.0f, X1 = 0.0f;
            for (long c = 0; c < nch; c++) {
                xin[2 * c] = X0;
                xin[2 * c + 1] = X1;
                const float *yi2 = yi + (size_t)c * B * 2, *yi1 = yi2 + B;
                const float f0 = fmaf(Cm[1], X1, fmaf(Cm[0], X0, yi2[B - 1]));
                const float f1 = fmaf(Cm[3], X1, fmaf(Cm[2], X0, yi1[B - 1]));
                X0 = f0;
                X1 = f1;
            }
        }
        /* (C) finalisation and back-correction */
#pragma omp parallel for num_threads(threads) schedule(static)
        for (long c = 0; c < nch; c++) {
            float *reg = y + c * CH;
            const float *yi2 = yi + (size_t)c * B * 2, *yi1 = yi2 + B;
            const float X0 = xin[2 * c], X1 = xin[2 * c + 1];
            for (int t = 0; t < B; t++) {
                const float *Cm = Cs + (size_t)t * 4;
                float *r = reg + (size_t)t * N;
                r[N - 2] = fmaf(Cm[1], X1, fmaf(Cm[0], X0, yi2[t]));
                r[N - 1] = fmaf(Cm[3], X1, fmaf(Cm[2], X0, yi1[t]));
            }
            for (int t = B - 1; t >= 0; t--) {
                float c2, c1;
                if (t == 0) { c2 = X0; c1 = X1; }
                else { c2 = reg[(size_t)(t - 1) * N + N - 2]; c1 = reg[(size_t)(t - 1) * N + N - 1]; }
                float *r = reg + (size_t)t * N;
                const float *hh1 = h1 + (size_t)s * N, *hh2 = h2 + (size_t)s * N;
                for (int k = 0; k < N - 2; k++) r[k] = fmaf(hh1[k], c1, fmaf(hh2[k], c2, r[k]));
            }
        }
    }
    free(h1); free(h2); free(Cx); free(yi); free(bnd); free(xin);
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* CR factors shared by STCR and DTCR                                         */
/* stcr_kernels.cuh:93-147, dtcr_kernels.cuh:134-139, 251-257                 */
/* ------------------------------------------------------------------------ */
typedef struct {
    float b1, b2, a1, a2;
    float f[16], e[16], fde[16], h[16], g[16], p[16], q[16], d[16], c[16];
    float cb, db, fb, eb, hbc, gbc;     /* decoupled first-round coefficients */
    float h0[64], hb2[64], hb1[64], he2[64], he1[64];
    float Cx[64][4];                     /* C^(k+1) */
    int L;
} cr_factors;

static void cr_setup(const float *row, int B, int N, cr_factors *K)
{
    const int L = ilog2(N);
    memset(K, 0, sizeof *K);
    for (int i = 0; i < 16; i++) {
        K->f[i] = K->e[i] = K->fde[i] = K->h[i] = K->g[i] = NAN;
        K->p[i] = K->q[i] = K->d[i] = K->c[i] = NAN;
    }
    K->L = L;
    K->b1 = row[1]; K->b2 = row[2]; K->a1 = -row[4]; K->a2 = -row[5];
    float *f = K->f, *e = K->e, *fde = K->fde, *h = K->h, *g = K->g;
    float *p = K->p, *q = K->q, *d = K->d, *c = K->c;
    f[0] = -K->a2; e[0] = -K->a1; fde[0] = f[0] / e[0];
    h[0] = f[0]; g[0] = e[0]; p[1] = f[0]; q[1] = e[0];
    for (int n = 1; n < L + 1; n++) {
        f[n] = f[n - 1] * f[n - 1];
        e[n] = 2 * f[n - 1] - e[n - 1] * e[n - 1];
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
    const float b1 = K->b1, b2 = K->b2;
    K->cb = f[0] * b1 - e[0] * b2;
    K->db = b1 - e[0];
    K->fb = b2 * f[0];
    K->eb = b2 + f[0] - b1 * e[0];
    K->hbc = b2 - e[0] * b1;
    K->gbc = -e[0] * b2;
    /* block filtering factors (stcr:127-147) */
    K->h0[0] = 1; K->h0[1] = -e[L];
    for (int l = 2; l < B; l++) K->h0[l] = -e[L] * K->h0[l - 1] - f[L] * K->h0[l - 2];
    for (int l = 0; l < B; l++) {
        K->hb2[l] = h[L] * K->h0[l];
        K->he1[l] = q[L] * K->h0[l];
        const float tmp = (l == 0) ? 0.0f : K->h0[l - 1];
        K->hb1[l] = f[L] * tmp + g[L] * K->h0[l];
        K->he2[l] = f[L] * tmp + p[L] * K->h0[l];
    }
    /* lookback transition (stcr:258-265) */
    float *hh1 = malloc(sizeof(float) * (size_t)N), *hh2 = malloc(sizeof(float) * (size_t)N);
    float C[4];
    ph_tables(K->a1, K->a2, N, hh1, hh2, C);
    for (int k = 0; k < B; k++) mat2_pow(C, k + 1, K->Cx[k]);
    free(hh1); free(hh2);
}

/* copy of one column of a B x N register file with a lane shift (shfl_up) */
#define REG(t, k) reg[(size_t)(t) * N + (k)]

/* ------------------------------------------------------------------------ */
/* STCR                                                                       */
/* ------------------------------------------------------------------------ */
/*
 * orc_stcr_f32 follows tools/stcr_emulator.py:111-191 (which mirrors
 * STCR_32_LOOP, stcr_kernels.cuh:279-605).  The hand-unrolled kernels execute
 * the same operation sequence (stcr_kernels.cuh:17-19; verified by
 * stcr_emulator.py:205-208), so `unrolled` selects the same arithmetic here.
 */
int orc_stcr_f32(const float *sos, int ns, int B, int N, int unrolled,
                 const float *x, long n, float *y)
{
    (void)unrolled;
    if (!(B == 32 || B == 64) || !is_pow2(N) || N < 8 || N > 1024 || ns < 1 || ns > 64) return ORC_EINVAL;
    const long CH = (long)B * N;
    if (n % CH != 0) return ORC_EINVAL;
    const long nch = n / CH;
    const int L = ilog2(N);
    cr_factors *Ks = malloc(sizeof(cr_factors) * (size_t)ns);
    float *reg = malloc(sizeof(float) * (size_t)CH);
    float *inn = malloc(sizeof(float) * (size_t)CH);
    float *fullc = malloc(sizeof(float) * (size_t)ns * 2);
    float xi1[64], xi2[64], yi1[64], yi2[64], by2[64], by1[64], sh[64], sh2[64];
    float y2e[64], y2o[64], y1e[64], y1o[64];
    if (!Ks || !reg || !inn || !fullc) return ORC_ENOMEM;
    for (int s = 0; s < ns; s++) cr_setup(sos + 6 * s, B, N, &Ks[s]);

    for (long cid = 0; cid < nch; cid++) {
        const long cs = cid * CH;
        memcpy(inn, x + cs, sizeof(float) * (size_t)CH);
        for (int s = 0; s < ns; s++) {
            const cr_factors *K = &Ks[s];
            /* FIR boundary history (stcr:321-342) */
            yi2[0] = 0.0f; yi1[0] = 0.0f;
            if (cid == 0) { xi2[0] = 0.0f; xi1[0] = 0.0f; }
            else if (s == 0) { xi2[0] = x[cs - 2]; xi1[0] = x[cs - 1]; }
            for (int t = 1; t < B; t++) {
                if (s == 0) {
                    xi2[t] = inn[(size_t)(t - 1) * N + N - 2];
                    xi1[t] = inn[(size_t)(t - 1) * N + N - 1];
                    yi2[t] = inn[(size_t)(t - 1) * N + N - 4];
                    yi1[t] = inn[(size_t)(t - 1) * N + N - 3];
                }
                const float oy1 = yi1[t];
                yi2[t] = fmaf(K->b1, oy1, fmaf(K->b2, yi2[t], xi2[t]));
                yi1[t] = fmaf(K->b1, xi2[t], fmaf(K->b2, oy1, xi1[t]));
            }
            /* hot loop 1: FIR + first double-sided CR round (stcr:344-362) */
            for (int t = 0; t < B; t++) {
                float _x2 = xi2[t], _x1 = xi1[t], _y2 = yi2[t], _y1 = yi1[t];
                for (int k = 0; k < N; k++) {
                    const float xc = (s == 0) ? inn[(size_t)t * N + k] : REG(t, k);
                    const float yc = fmaf(K->b1, _x1, fmaf(K->b2, _x2, xc));
                    REG(t, k) = fmaf(-K->e[0], _y1, fmaf(K->f[0], _y2, yc));
                    _x2 = _x1; _x1 = xc; _y2 = _y1; _y1 = yc;
                }
            }
            /* remaining forward CR rounds (stcr:364-386) */
            for (int ro = 1; ro < L; ro++) {
                const int step = 2 << ro, sub = 1 << ro;
                for (int i = 0; i < 2; i++) {
                    const int off = sub - 2 + i;
                    for (int t = 0; t < B; t++) sh[t] = (t == 0) ? 0.0f : REG(t - 1, N - 2 + i);
                    for (int t = 0; t < B; t++) {
                        float _x2 = sh[t];
                        for (int k = 0; k < N; k += step) {
                            REG(t, k + off) = fmaf(-K->fde[ro], _x2, REG(t, k + off));
                            _x2 = REG(t, k + off + sub);
                            REG(t, k + off + sub) = fmaf(-K->e[ro], REG(t, k + off), _x2);
                        }
                    }
                }
            }
            /* block filtering over lanes (stcr:388-405) */
            for (int t = 0; t < B; t++) {
                by2[t] = K->h0[0] * REG(t, N - 2);
                by1[t] = K->h0[0] * REG(t, N - 1);
            }
            for (int m = 1; m < B; m++) {
                for (int t = 0; t < B; t++) {
                    sh[t] = (t >= m) ? REG(t - m, N - 2) : 0.0f;
                    sh2[t] = (t >= m) ? REG(t - m, N - 1) : 0.0f;
                }
                for (int t = 0; t < B; t++) {
                    by2[t] = fmaf(K->h0[m], sh[t], by2[t]);
                    by1[t] = fmaf(K->h0[m], sh2[t], by1[t]);
                }
            }
            /* lookback (serial) */
            const float X0 = (cid == 0) ? 0.0f : fullc[2 * s];
            const float X1 = (cid == 0) ? 0.0f : fullc[2 * s + 1];
            /* terminal correction (stcr:480-491) */
            for (int t = 0; t < B; t++) {
                REG(t, N - 2) = fmaf(-K->he1[t], X1, fmaf(-K->he2[t], X0, by2[t]));
                REG(t, N - 1) = fmaf(-K->hb1[t], X1, fmaf(-K->hb2[t], X0, by1[t]));
            }
            fullc[2 * s] = REG(B - 1, N - 2);
            fullc[2 * s + 1] = REG(B - 1, N - 1);
            /* deepest companion level (stcr:493-523) */
            for (int t = 0; t < B; t++) {
                y2e[t] = (t >= 2) ? REG(t - 2, N - 2) : 0.0f;
                y2o[t] = (t >= 1) ? REG(t - 1, N - 1) : 0.0f;
                y1e[t] = (t >= 1) ? REG(t - 1, N - 2) : 0.0f;
                y1o[t] = (t >= 2) ? REG(t - 2, N - 1) : 0.0f;
            }
            const int H = N / 2;
            for (int t = 0; t < B; t++) {
                float h2e, h2o, h1e, h1o;
                if (t == 0) {
                    h2e = K->p[L - 1]; h2o = K->h[L - 1];
                    h1e = K->q[L - 1]; h1o = K->g[L - 1];
                    y2e[t] = X0; y2o[t] = X0; y1e[t] = X1; y1o[t] = X1;
                    xi2[t] = X0; xi1[t] = X1;
                } else {
                    h2e = K->d[L]; h2o = K->c[L];
                    h1e = K->c[L]; h1o = K->d[L];
                    if (t == 1) { y2e[t] = X0; y1o[t] = X1; }
                    xi2[t] = y1e[t]; xi1[t] = y2o[t];
                }
                REG(t, H - 2) = fmaf(-h1e, y1e[t], fmaf(-h2e, y2e[t], REG(t, H - 2)));
                REG(t, H - 1) = fmaf(-h1o, y1o[t], fmaf(-h2o, y2o[t], REG(t, H - 1)));
            }
            /* back-substitution rounds (stcr:527-585) */
            for (int r = L - 2; r > 0; r--) {
                const int stride = 2 << r, sub = stride >> 1, P = (N / 2) >> r;
                for (int t = 0; t < B; t++) {
                    sh[t] = (t >= 1) ? REG(t - 1, N - stride - 2) : 0.0f;
                    sh2[t] = (t >= 1) ? REG(t - 1, N - stride - 1) : 0.0f;
                }
                for (int t = 0; t < B; t++) {
                    float h2o, h1o, h2e, h1e, z2e, z1o;
                    if (t == 0) {
                        h2o = K->h[r]; h1o = K->g[r]; h2e = K->p[r]; h1e = K->q[r];
                        z2e = X0; z1o = X1;
                    } else {
                        h2o = K->c[r + 1]; h1o = K->d[r + 1]; h2e = h1o; h1e = h2o;
                        z2e = sh[t]; z1o = sh2[t];
                    }
                    float z2o = y2o[t], z1e = y1e[t];
                    REG(t, sub - 2) = fmaf(-h1e, z1e, fmaf(-h2e, z2e, REG(t, sub - 2)));
                    REG(t, sub - 1) = fmaf(-h1o, z1o, fmaf(-h2o, z2o, REG(t, sub - 1)));
                    h2o = K->d[r + 1]; h1o = K->c[r + 1]; h2e = h2o; h1e = h1o;
                    z2o = xi1[t]; z1o = REG(t, stride - 1); z2e = xi2[t]; z1e = REG(t, stride - 2);
                    REG(t, stride + sub - 2) = fmaf(-h1e, z1e, fmaf(-h2e, z2e, REG(t, stride + sub - 2)));
                    REG(t, stride + sub - 1) = fmaf(-h1o, z1o, fmaf(-h2o, z2o, REG(t, stride + sub - 1)));
                    for (int m = 2; m < P; m++) {
                        z2o = z1o; z1o = REG(t, stride * m - 1);
                        z2e = z1e; z1e = REG(t, stride * m - 2);
                        REG(t, stride * m + sub - 2) = fmaf(-h1e, z1e, fmaf(-h2e, z2e, REG(t, stride * m + sub - 2)));
                        REG(t, stride * m + sub - 1) = fmaf(-h1o, z1o, fmaf(-h2o, z2o, REG(t, stride * m + sub - 1)));
                    }
                }
            }
            /* section hand-off (stcr:587-594) */
            if (s < ns - 1) {
                for (int t = 0; t < B; t++) {
                    yi2[t] = (t >= 1) ? REG(t - 1, N - 4) : 0.0f;
                    yi1[t] = (t >= 1) ? REG(t - 1, N - 3) : 0.0f;
                }
            }
        }
        memcpy(y + cs, reg, sizeof(float) * (size_t)CH);
    }
    free(Ks); free(reg); free(inn); free(fullc);
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* DTCR                                                                       */
/* ------------------------------------------------------------------------ */
/*
 * orc_dtcr_f32 follows tools/dtcr_emulator.py:79-304 (which mirrors
 * DTCR_32_LOOP, dtcr_kernels.cuh:291-657, and DTCR_64_LOOP, 2500-2904).  Side
 * ty = 0 holds the even samples of each lane, ty = 1 the odd ones; each lane
 * keeps N/2 values per side.  For B = 64 the two-warp kernel reads
 * neighbours through shared memory instead of shuffles; the values read are
 * the same (dtcr_emulator.py:158-175, 237-256), so one restatement serves
 * both.
 */
int orc_dtcr_f32(const float *sos, int ns, int B, int N, const float *x,
                 long n, float *y)
{
    if (!(B == 32 || B == 64) || !is_pow2(N) || N < 8 || N > 1024 || ns < 1 || ns > 64) return ORC_EINVAL;
    const long CH = (long)B * N;
    if (n % CH != 0) return ORC_EINVAL;
    const long nch = n / CH;
    const int HALF = N / 2, QUART = HALF / 2, L = ilog2(N);
    cr_factors *Ks = malloc(sizeof(cr_factors) * (size_t)ns);
    float *inn = malloc(sizeof(float) * (size_t)CH);
    float *snap = malloc(sizeof(float) * (size_t)CH);
    float *rg = malloc(sizeof(float) * 2 * (size_t)B * HALF);
    float *fullc = malloc(sizeof(float) * (size_t)ns * 2);
    if (!Ks || !inn || !snap || !rg || !fullc) return ORC_ENOMEM;
#define RG(ty, tx, k) rg[((size_t)(ty) * B + (tx)) * HALF + (k)]
#define INN(tx, k) inn[(size_t)(tx) * N + (k)]
#define SNAP(tx, k) snap[(size_t)(tx) * N + (k)]
    float xi1[2][64], xi2[2][64], xi3[2][64], xi4[2][64], xcv[2][64], tmpv[2][64];
    float yi1v[2][64], yi2v[2][64], yiv[2][64], col[2][64], zi1[2][64], zi2[2][64];
    for (int s = 0; s < ns; s++) cr_setup(sos + 6 * s, B, N, &Ks[s]);
    memset(xi1, 0, sizeof xi1); memset(xi2, 0, sizeof xi2);
    memset(xi3, 0, sizeof xi3); memset(xi4, 0, sizeof xi4);

    for (long cid = 0; cid < nch; cid++) {
        const long cs = cid * CH;
        memcpy(inn, x + cs, sizeof(float) * (size_t)CH);
        for (int s = 0; s < ns; s++) {
            const cr_factors *K = &Ks[s];
            memcpy(snap, inn, sizeof(float) * (size_t)CH);
            /* boundary setup (dtcr:337-371) */
            for (int ty = 0; ty < 2; ty++) {
                for (int tx = 0; tx < B; tx++) {
                    if (ty == 0) {
                        if (tx == 0) {
                            xi3[0][0] = 0.0f; xi4[0][0] = 0.0f;
                            if (cid == 0) { xi1[0][0] = 0.0f; xi2[0][0] = 0.0f; }
                            else if (s == 0) { xi1[0][0] = x[cs - 1]; xi2[0][0] = x[cs - 2]; }
                        } else {
                            xi1[0][tx] = SNAP(tx - 1, N - 1); xi2[0][tx] = SNAP(tx - 1, N - 2);
                            xi3[0][tx] = SNAP(tx - 1, N - 3); xi4[0][tx] = SNAP(tx - 1, N - 4);
                        }
                    } else {
                        xi1[1][tx] = SNAP(tx, 0);
                        if (tx == 0) {
                            xi4[1][0] = 0.0f;
                            if (cid == 0) { xi2[1][0] = 0.0f; xi3[1][0] = 0.0f; }
                            else if (s == 0) { xi2[1][0] = x[cs - 1]; xi3[1][0] = x[cs - 2]; }
                        } else {
                            xi2[1][tx] = SNAP(tx - 1, N - 1); xi3[1][tx] = SNAP(tx - 1, N - 2);
                            xi4[1][tx] = SNAP(tx - 1, N - 3);
                        }
                    }
                }
            }
            /* fused FIR + first CR round on each parity side (dtcr:373-415) */
            for (int ty = 0; ty < 2; ty++) {
                for (int tx = 0; tx < B; tx++) {
                    float c1, c2, c3;
                    if (ty == 0) {
                        if (tx == 0) { c1 = K->b1; c2 = K->b2; c3 = 0.0f; }
                        else { c1 = K->db; c2 = K->eb; c3 = K->cb; }
                    } else {
                        c1 = K->db;
                        if (tx == 0) { c2 = K->hbc; c3 = K->gbc; }
                        else { c2 = K->eb; c3 = K->cb; }
                    }
                    float _xc = SNAP(tx, ty);
                    float _x1 = xi1[ty][tx], _x2 = xi2[ty][tx], _x3 = xi3[ty][tx], _x4 = xi4[ty][tx];
                    RG(ty, tx, 0) = fmaf(c1, _x1, fmaf(c2, _x2, fmaf(c3, _x3, fmaf(K->fb, _x4, _xc))));
                    _x4 = _x2; _x3 = _x1; _x2 = _xc;
                    for (int k = 1; k < HALF; k++) {
                        _xc = SNAP(tx, 2 * k + ty);
                        _x1 = SNAP(tx, 2 * k + ty - 1);
                        RG(ty, tx, k) = fmaf(K->db, _x1, fmaf(K->eb, _x2, fmaf(K->cb, _x3, fmaf(K->fb, _x4, _xc))));
                        _x4 = _x2; _x3 = _x1; _x2 = _xc;
                    }
                    xcv[ty][tx] = _xc;
                    xi1[ty][tx] = _x1; xi2[ty][tx] = _x2; xi3[ty][tx] = _x3; xi4[ty][tx] = _x4;
                }
            }
            /* remaining forward CR rounds per side (dtcr:417-449) */
            for (int r = 1; r < L; r++) {
                const int step = 1 << r, sub = step >> 1, off = sub - 1;
                for (int ty = 0; ty < 2; ty++) {
                    float shv[64];
                    for (int tx = 0; tx < B; tx++) shv[tx] = (tx == 0) ? 0.0f : RG(ty, tx - 1, HALF - 1);
                    for (int tx = 0; tx < B; tx++) {
                        float _x2 = shv[tx];
                        for (int k = 0; k < HALF; k += step) {
                            RG(ty, tx, k + off) = fmaf(-K->fde[r], _x2, RG(ty, tx, k + off));
                            _x2 = RG(ty, tx, k + off + sub);
                            RG(ty, tx, k + off + sub) = fmaf(-K->e[r], RG(ty, tx, k + off), _x2);
                        }
                    }
                }
            }
            /* block filtering per side */
            for (int ty = 0; ty < 2; ty++) {
                for (int tx = 0; tx < B; tx++) col[ty][tx] = RG(ty, tx, HALF - 1);
                for (int tx = 0; tx < B; tx++) yiv[ty][tx] = K->h0[0] * col[ty][tx];
                for (int m = 1; m < B; m++)
                    for (int tx = 0; tx < B; tx++)
                        yiv[ty][tx] = fmaf(K->h0[m], (tx >= m) ? col[ty][tx - m] : 0.0f, yiv[ty][tx]);
            }
            /* lookback (serial) */
            const float X0 = (cid == 0) ? 0.0f : fullc[2 * s];
            const float X1 = (cid == 0) ? 0.0f : fullc[2 * s + 1];
            /* terminal correction + full carry (dtcr:534-553) */
            for (int ty = 0; ty < 2; ty++) {
                const float *c2 = (ty == 0) ? K->he2 : K->hb2;
                const float *c1 = (ty == 0) ? K->he1 : K->hb1;
                for (int tx = 0; tx < B; tx++)
                    RG(ty, tx, HALF - 1) = fmaf(-c1[tx], X1, fmaf(-c2[tx], X0, yiv[ty][tx]));
            }
            fullc[2 * s] = RG(0, B - 1, HALF - 1);
            fullc[2 * s + 1] = RG(1, B - 1, HALF - 1);
            for (int ty = 0; ty < 2; ty++)
                for (int tx = 0; tx < B; tx++) INN(tx, N - 2 + ty) = RG(ty, tx, HALF - 1);
            /* deepest companion level (dtcr:555-590) */
            for (int ty = 0; ty < 2; ty++) {
                for (int tx = 0; tx < B; tx++) {
                    float _y2, _y1, h2b, h1b;
                    if (tx == 0) {
                        tmpv[ty][0] = X1;
                        if (ty == 0) {
                            h2b = K->p[L - 1]; h1b = K->q[L - 1];
                            xcv[0][0] = X0; xi1[0][0] = X1; xi2[0][0] = X0;
                        } else {
                            h2b = K->h[L - 1]; h1b = K->g[L - 1];
                            xcv[1][0] = X1; xi2[1][0] = X1; xi3[1][0] = X0;
                        }
                        _y2 = X0; _y1 = X1;
                    } else {
                        /* shfl_up(col, 2-ty) and shfl_up(col, 1+ty) of the
                         * terminal values of this side */
                        const int d2 = 2 - ty, d1 = 1 + ty;
                        _y2 = (tx >= d2) ? RG(ty, tx - d2, HALF - 1) : 0.0f;
                        _y1 = (tx >= d1) ? RG(ty, tx - d1, HALF - 1) : 0.0f;
                        if (ty == 0) { h2b = K->d[L]; h1b = K->c[L]; tmpv[0][tx] = _y1; }
                        else { h2b = K->c[L]; h1b = K->d[L]; tmpv[1][tx] = _y2; }
                        if (tx == 1) { if (ty == 0) _y2 = X0; else _y1 = X1; }
                        xcv[ty][tx] = tmpv[ty][tx];
                    }
                    RG(ty, tx, QUART - 1) = fmaf(-h1b, _y1, fmaf(-h2b, _y2, RG(ty, tx, QUART - 1)));
                    yi1v[ty][tx] = _y1; yi2v[ty][tx] = _y2;
                }
            }
            for (int ty = 0; ty < 2; ty++)
                for (int tx = 0; tx < B; tx++) INN(tx, HALF - 2 + ty) = RG(ty, tx, QUART - 1);
            /* back-substitution rounds (dtcr:592-644) */
            for (int r = L - 2; r > 0; r--) {
                const int P = HALF >> r, stride = 2 << r, sub = 1 << r, sub2 = sub >> 1;
                float y2sh[2][64];
                for (int ty = 0; ty < 2; ty++)
                    for (int tx = 0; tx < B; tx++)
                        y2sh[ty][tx] = (tx >= 1) ? RG(ty, tx - 1, HALF - 1 - sub) : 0.0f;
                for (int m = 0; m < P; m++) {
                    for (int ty = 0; ty < 2; ty++) {
                        for (int tx = 0; tx < B; tx++) {
                            float h2b, h1b;
                            if (m == 0) {
                                float _y2;
                                if (tx == 0) {
                                    if (ty == 0) { h2b = K->p[r]; h1b = K->q[r]; }
                                    else { h2b = K->h[r]; h1b = K->g[r]; }
                                    _y2 = X0;
                                } else {
                                    h2b = K->d[r + 1]; h1b = K->c[r + 1];
                                    _y2 = y2sh[ty][tx];
                                }
                                zi2[ty][tx] = _y2;
                                zi1[ty][tx] = tmpv[ty][tx];
                                if (tx == 0) yi2v[ty][tx] = _y2;
                            } else if (m == 1) {
                                h2b = K->d[r + 1]; h1b = K->c[r + 1];
                                zi2[ty][tx] = xcv[ty][tx];
                                zi1[ty][tx] = RG(ty, tx, sub - 1);
                            } else {
                                h2b = K->d[r + 1]; h1b = K->c[r + 1];
                                zi2[ty][tx] = zi1[ty][tx];
                                zi1[ty][tx] = RG(ty, tx, sub * m - 1);
                            }
                            const int dst = sub * m + sub2 - 1;
                            RG(ty, tx, dst) = fmaf(-h1b, zi1[ty][tx], fmaf(-h2b, zi2[ty][tx], RG(ty, tx, dst)));
                            INN(tx, stride * m + sub - 2 + ty) = RG(ty, tx, dst);
                        }
                    }
                }
            }
        }
        memcpy(y + cs, inn, sizeof(float) * (size_t)CH);
    }
#undef RG
#undef INN
#undef SNAP
    free(Ks); free(inn); free(snap); free(rg); free(fullc);
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* PLR                                                                        */
/* ------------------------------------------------------------------------ */
/*
 * n-nacci correction factors, plr_kernels.cuh:149-409:
 * fac[j][i] = output i of the all-pole recurrence y[m] = sum_l a[l] y[m-1-l]
 * started from the unit history e_j, where history slot j is y[-k+j].
 * Unfused float32, recurrence terms added in ascending tap order.
 */
static void plr_factors(const float *a, int k, int len, float *fac /* k x len */)
{
    float *v = malloc(sizeof(float) * (size_t)(len + k));
    for (int j = 0; j < k; j++) {
        for (int i = 0; i < len + k; i++) v[i] = 0.0f;
        v[j] = 1.0f;
        for (int i = k; i < len + k; i++) {
            float acc = a[0] * v[i - 1];
            for (int l = 1; l < k; l++) acc = acc + a[l] * v[i - 1 - l];
            v[i] = acc;
        }
        for (int i = 0; i < len; i++) fac[(size_t)j * len + i] = v[i + k];
    }
    free(v);
}

/*
 * orc_plr_f32: direct form of order k.  bt[0..k-1] are the FIR taps past the
 * implicit b0 == 1, at[0..k-1] the feedback taps with the recurrence sign
 * (ref_generate.py:80-81), both float64 and cast to float32 as in
 * plr_kernels.cuh:423-427.  BLOCK threads per chunk, X values per thread,
 * chunk = BLOCK * X; thread t owns samples cs + t + v*BLOCK (plr:524-537); the
 * last chunk is zero padded (tail safe).
 *
 * Merge tree (plr:564-785): a Sklansky doubling over the BLOCK lanes of each
 * segment v.  At half width h, lane i of the upper half of every 2h group
 * adds fac[m][i % h] * val[lower_end - k + 1 + m] for the last min(k, h)
 * lanes of the lower half, m ascending (terms for lanes before the group
 * are zero in the group's zero-state solution; at h = 1 this is a1 * val[0]).
 * The 32-lane shuffle rounds and the cross-warp shared-memory stages of the
 * kernel follow this one rule.  Segments are then chained serially
 * (plr:787-811), the chunk carry is applied (plr:891-899).
 */
int orc_plr_f32(const double *bt, const double *at, int k, int BLOCK, int X,
                const float *x, long n, float *y)
{
    if (k < 1 || k > 32 || BLOCK < 2 * k || !is_pow2(BLOCK) || X < 1) return ORC_EINVAL;
    const long CH = (long)BLOCK * X;
    const long nch = (n + CH - 1) / CH;
    float b[32], a[32];
    for (int j = 0; j < k; j++) { b[j] = (float)bt[j]; a[j] = (float)at[j]; }
    float *fac = malloc(sizeof(float) * (size_t)k * CH);
    float *val = malloc(sizeof(float) * (size_t)CH);  /* val[v*BLOCK + t] */
    float *raw = malloc(sizeof(float) * (size_t)CH);
    float *full = malloc(sizeof(float) * (size_t)k);
    float *src = malloc(sizeof(float) * (size_t)BLOCK);
    if (!fac || !val || !raw || !full || !src) return ORC_ENOMEM;
    plr_factors(a, k, (int)CH, fac);
#define FAC(j, i) fac[(size_t)(j) * CH + (i)]
    for (int j = 0; j < k; j++) full[j] = 0.0f;
    for (long cid = 0; cid < nch; cid++) {
        const long cs = cid * CH;
        for (long i = 0; i < CH; i++) raw[i] = (cs + i < n) ? x[cs + i] : 0.0f;
        /* phase 1: FIR, history from the raw input (plr:539-562):
         * val += b1*x[-1]; val += b2*x[-2]; ... */
        for (long i = 0; i < CH; i++) {
            float v = raw[i];
            for (int j = 0; j < k; j++) {
                const long gi = cs + i - 1 - j;
                float hx;
                if (i - 1 - j >= 0) hx = raw[i - 1 - j];
                else hx = (gi >= 0) ? x[gi] : 0.0f;
                v = fmaf(b[j], hx, v);
            }
            val[i] = v;
        }
        /* phase 2: merge tree per segment */
        for (int v = 0; v < X; v++) {
            float *sv = val + (size_t)v * BLOCK;
            for (int h = 1; h < BLOCK; h <<= 1) {
                memcpy(src, sv, sizeof(float) * (size_t)BLOCK);
                for (int t = 0; t < BLOCK; t++) {
                    if (!(t & h)) continue;
                    const int base = t & ~(2 * h - 1);
                    const int i = t % h;
                    const int m0 = (k > h) ? k - h : 0;
                    float acc = sv[t];
                    for (int m = m0; m < k; m++) acc = fmaf(FAC(m, i), src[base + h - k + m], acc);
                    sv[t] = acc;
                }
            }
        }
        /* segment chaining (plr:787-811) */
        for (int v = 1; v < X; v++) {
            const float *prev = val + (size_t)(v - 1) * BLOCK;
            float *sv = val + (size_t)v * BLOCK;
            for (int t = 0; t < BLOCK; t++) {
                float acc = sv[t];
                for (int m = 0; m < k; m++) acc = fmaf(FAC(m, t), prev[BLOCK - k + m], acc);
                sv[t] = acc;
            }
        }
        /* phase 3: carry from the predecessor's full carry (plr:891-899) */
        for (int v = 0; v < X; v++) {
            for (int t = 0; t < BLOCK; t++) {
                float acc = val[(size_t)v * BLOCK + t];
                for (int m = 0; m < k; m++) acc = fmaf(FAC(m, (size_t)t + (size_t)v * BLOCK), full[m], acc);
                val[(size_t)v * BLOCK + t] = acc;
            }
        }
        for (int m = 0; m < k; m++) full[m] = val[CH - k + m];
        for (long i = 0; i < CH && cs + i < n; i++) y[cs + i] = val[i];
    }
#undef FAC
    free(fac); free(val); free(raw); free(full); free(src);
    return ORC_OK;
}

/* Sequential float32 direct form, every operation rounded to float32
 * (tools/plr_emulator.py:185-202; unfused on purpose, as the emulator). */
void orc_direct_form_f32(const double *bt, const double *at, int k,
                         const float *x, long n, float *y)
{
    float b[32], a[32], xh[32], yh[32];
    if (k > 32) return;
    for (int j = 0; j < k; j++) { b[j] = (float)bt[j]; a[j] = (float)at[j]; xh[j] = 0.0f; yh[j] = 0.0f; }
    for (long i = 0; i < n; i++) {
        float t = x[i];
        for (int j = 0; j < k; j++) t = t + b[j] * xh[j];
        float v = t;
        for (int j = 0; j < k; j++) v = v + a[j] * yh[j];
        for (int j = k - 1; j > 0; j--) { xh[j] = xh[j - 1]; yh[j] = yh[j - 1]; }
        xh[0] = x[i]; yh[0] = v;
        y[i] = v;
    }
}

/* plain float32 cascade with fused FMA, direct form I per section: a
 * sequential baseline for timing (same recurrence as the kernels, no chunking) */
void orc_cascade_f32(const float *sos, int ns, const float *x, long n, float *y)
{
    if (x != y) memcpy(y, x, sizeof(float) * (size_t)n);
    for (int s = 0; s < ns; s++) {
        const float b1 = sos[6 * s + 1], b2 = sos[6 * s + 2];
        const float a1 = -sos[6 * s + 4], a2 = -sos[6 * s + 5];
        float x1 = 0, x2 = 0, y1 = 0, y2 = 0;
        for (long i = 0; i < n; i++) {
            const float xc = y[i];
            const float v = fmaf(a1, y1, fmaf(a2, y2, fmaf(b1, x1, fmaf(b2, x2, xc))));
            x2 = x1; x1 = xc; y2 = y1; y1 = v;
            y[i] = v;
        }
    }
}
