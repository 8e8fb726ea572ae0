// autotune.cpp -- runtime algorithm and geometry selection.
//
// Replaces the reference's build-time search: run_accuracy.sh and
// run_performance.sh sweep KERNEL x (BLOCK_SIZE, N_BLOCKS) per order and
// calibrate PLR_X_MAX from ptxas spill reports (run_accuracy.sh:95-151).  Here
// every candidate configuration is classified on the host (an instance for
// the geometry, CR factors well conditioned: the same checks a plan makes), then the
// eligible ones are timed on the device and the fastest is returned.
#include "../../include/iirx.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

// samples of the accuracy screen (a prefix: the filter is causal)
constexpr int64_t kScreenSamples = 1 << 18;

// FP64 cascade (direct form II transposed per section) of the float32
// sections over the float32 input, the golden of ref_generate.py:89
void fp64_cascade(const float *sos, int ns, const float *x, int64_t n, std::vector<double> &y)
{
    y.assign(x, x + n);
    for (int s = 0; s < ns; ++s) {
        const float *r = sos + 6 * s;
        const double b0 = r[0], b1 = r[1], b2 = r[2], a1 = r[4], a2 = r[5];
        double z0 = 0.0, z1 = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            const double u = y[(size_t)i];
            const double v = b0 * u + z0;
            z0 = b1 * u - a1 * v + z1;
            z1 = b2 * u - a2 * v;
            y[(size_t)i] = v;
        }
    }
}

std::vector<iirx_config> enumerate(int32_t ns, uint32_t flags)
{
    std::vector<iirx_config> c;
    const int32_t sched = (flags & IIRX_TUNE_TICKET) ? IIRX_SCHEDULE_TICKET : IIRX_SCHEDULE_PERSISTENT;
    auto add = [&](int32_t kernel, int32_t B, int32_t N, int32_t unrolled) {
        iirx_config x;
        std::memset(&x, 0, sizeof x);
        x.kernel = kernel;
        x.block_size = B;
        x.n_blocks = N;
        x.n_sections = ns;
        x.unrolled = unrolled;
        x.schedule = sched;
        c.push_back(x);
    };
    for (int N = 8; N <= 128; N *= 2) add(IIRX_KERNEL_PH, 64, N, 0);
    if (ns >= 2)
        for (int N = 8; N <= 128; N *= 2) add(IIRX_KERNEL_PH_JOINT, 64, N, 0);
    // halo carries read kPhHalo = 256 more samples per chunk: chunks of 1024 and
    // up; at one section too (no lookback and no schedule: 2^25 samples, order
    // 2, 737-743 GS/s against PH's 706-712)
    for (int N = 16; N <= 128; N *= 2) add(IIRX_KERNEL_PH_HALO, 64, N, 0);
    if (!(flags & IIRX_TUNE_NO_CR))
        for (int k : {IIRX_KERNEL_STCR, IIRX_KERNEL_DTCR})
            for (int N = 8; N <= 128; N *= 2)
                for (int u = 0; u < 2; ++u) add(k, 64, N, u);
    if (!(flags & IIRX_TUNE_NO_BLOCKSTATE))
        for (int G : {16, 32, 64})
            for (int L : {32, 64, 128})
                if (L * G <= 4096) add(IIRX_KERNEL_BLOCKSTATE, G, L, 0);
    if (!(flags & IIRX_TUNE_NO_BLOCKSTATE))
        for (int G : {16, 32})
            for (int L : {64, 128}) add(IIRX_KERNEL_BLOCKSTATE_BF16, G, L, 0);
    if (ns <= 2 || (flags & IIRX_TUNE_PLR_ALL_ORDERS)) add(IIRX_KERNEL_PLR, 0, 0, 0);
    return c;
}

}  // namespace

extern "C" {

int iirx_tune_candidates(const float *sos, int32_t n_sections, int64_t n_samples, uint32_t flags,
                         iirx_tune_entry *entries, int32_t cap, int32_t *count)
{
    if (!sos || n_sections < 1 || n_sections > 8 || n_samples < 1) return IIRX_ERR_INVALID;
    const std::vector<iirx_config> c = enumerate(n_sections, flags);
    for (size_t i = 0; i < c.size(); ++i) {
        if ((int32_t)i >= cap || !entries) break;
        iirx_tune_entry &e = entries[i];
        e.cfg = c[i];
        int64_t size = 0;
        int32_t stride = 0;
        e.status = iirx_plan_tables(&c[i], sos, n_samples, nullptr, 0, &size, &stride);
        e.ms = -1.0f;
        e.rel_err = -1.0f;
    }
    if (count) *count = (int32_t)c.size();
    return IIRX_OK;
}

int iirx_autotune(const float *sos, int32_t n_sections, int64_t n_samples, uint32_t flags, int32_t reps,
                  iirx_config *best, float *best_ms, iirx_tune_entry *entries, int32_t cap, int32_t *count)
{
    if (!sos || !best || reps < 1) return IIRX_ERR_INVALID;
    int32_t n = 0;
    int st = iirx_tune_candidates(sos, n_sections, n_samples, flags, nullptr, 0, &n);
    if (st != IIRX_OK) return st;
    std::vector<iirx_tune_entry> t((size_t)n);
    iirx_tune_candidates(sos, n_sections, n_samples, flags, t.data(), n, &n);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return IIRX_ERR_NO_DEVICE;

    // inputs rotate over enough buffers that a long signal streams from HBM
    // instead of the 256 MB Infinity Cache (as in real use), at most 8
    const size_t bytes = (size_t)n_samples * sizeof(float);
    const size_t nbuf = std::min<size_t>(8, std::max<size_t>(1, (512u << 20) / bytes + 1));
    float *d_in = nullptr, *d_out = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&]() {
        if (e0) hipEventDestroy(e0);
        if (e1) hipEventDestroy(e1);
        if (s) hipStreamDestroy(s);
        if (d_in) hipFree(d_in);
        if (d_out) hipFree(d_out);
    };
    if (hipMalloc((void **)&d_in, nbuf * bytes) != hipSuccess || hipMalloc((void **)&d_out, bytes) != hipSuccess) {
        cleanup();
        return IIRX_ERR_OOM;
    }
    // N(0,1) inputs (zeros run measurably faster than real data on this
    // device) and the FP64 golden of their prefix
    const int64_t n_screen = std::min<int64_t>(n_samples, kScreenSamples);
    std::vector<double> ref;
    double mref = 0.0;
    {
        std::vector<float> h((size_t)n_samples);
        iirx_fill_normal(0x7475ULL, n_samples, h.data());
        if (hipMemcpy(d_in, h.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) {
            cleanup();
            return IIRX_ERR_HIP;
        }
        fp64_cascade(sos, n_sections, h.data(), n_screen, ref);
        for (double v : ref) mref = std::max(mref, std::fabs(v));
    }
    std::vector<float> got((size_t)n_screen);
    for (size_t k = 1; k < nbuf; ++k)
        if (hipMemcpy(d_in + k * (size_t)n_samples, d_in, bytes, hipMemcpyDeviceToDevice) != hipSuccess) {
            cleanup();
            return IIRX_ERR_HIP;
        }
    if (hipStreamCreate(&s) != hipSuccess || hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        cleanup();
        return IIRX_ERR_HIP;
    }
    int besti = -1;
    for (int32_t i = 0; i < n; ++i) {
        if (t[(size_t)i].status != IIRX_OK) continue;
        iirx_plan *p = nullptr;
        const int cs = iirx_plan_create(&t[(size_t)i].cfg, sos, n_samples, &p);
        if (cs != IIRX_OK) {
            t[(size_t)i].status = cs;
            continue;
        }
        auto src = [&](int k) { return d_in + (size_t)(k % (int)nbuf) * (size_t)n_samples; };
        int ok = IIRX_OK;
        for (int w = 0; w < 3 && ok == IIRX_OK; ++w) ok = iirx_execute(p, src(w), d_out, s);
        if (ok == IIRX_OK) ok = hipEventRecord(e0, s) == hipSuccess ? IIRX_OK : IIRX_ERR_HIP;
        for (int r = 0; r < reps && ok == IIRX_OK; ++r) ok = iirx_execute(p, src(r), d_out, s);
        if (ok == IIRX_OK) ok = hipEventRecord(e1, s) == hipSuccess ? IIRX_OK : IIRX_ERR_HIP;
        if (ok == IIRX_OK) ok = hipEventSynchronize(e1) == hipSuccess ? IIRX_OK : IIRX_ERR_HIP;
        if (ok == IIRX_OK) ok = iirx_plan_check(p);
        float ms = 0.0f;
        if (ok == IIRX_OK && hipEventElapsedTime(&ms, e0, e1) == hipSuccess &&
            hipMemcpy(got.data(), d_out, (size_t)n_screen * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess) {
            t[(size_t)i].ms = ms / (float)reps;
            double err = 0.0;
            bool finite = true;
            for (int64_t k = 0; k < n_screen; ++k) {
                if (!std::isfinite(got[(size_t)k])) finite = false;
                else err = std::max(err, std::fabs((double)got[(size_t)k] - ref[(size_t)k]));
            }
            t[(size_t)i].rel_err = finite ? (float)(mref > 0.0 ? err / mref : err) : INFINITY;
        } else {
            t[(size_t)i].status = ok == IIRX_OK ? IIRX_ERR_HIP : ok;
        }
        iirx_plan_destroy(p);
    }
    cleanup();
    // accuracy screen: the reference's gate, or within 10x the most accurate
    // candidate where no candidate meets it (narrow filters amplify float32
    // rounding beyond the gate in every float32 form)
    float best_rel = INFINITY;
    for (const auto &e : t)
        if (e.status == IIRX_OK && e.ms >= 0.0f) best_rel = std::min(best_rel, e.rel_err);
    for (int32_t i = 0; i < n; ++i) {
        iirx_tune_entry &e = t[(size_t)i];
        if (e.status != IIRX_OK || e.ms < 0.0f) continue;
        if (!(e.rel_err <= 1e-4f || e.rel_err <= 10.0f * best_rel)) {
            e.status = IIRX_ERR_ACCURACY;
            continue;
        }
        if (besti < 0 || e.ms < t[(size_t)besti].ms) besti = i;
    }
    if (entries)
        for (int32_t i = 0; i < n && i < cap; ++i) entries[i] = t[(size_t)i];
    if (count) *count = n;
    if (besti < 0) return IIRX_ERR_GEOMETRY;
    *best = t[(size_t)besti].cfg;
    if (best_ms) *best_ms = t[(size_t)besti].ms;
    return IIRX_OK;
}

int iirx_batch_autotune(const iirx_batch_config *cfg, int64_t n_channels, const int64_t *offsets,
                        const int32_t *n_sections, const float *sos, int32_t reps, iirx_batch_config *best,
                        float *best_ms)
{
    if (!cfg || !offsets || !n_sections || !sos || !best || reps < 1 || n_channels < 1) return IIRX_ERR_INVALID;
    const int64_t n = offsets[n_channels] - offsets[0];
    if (n < 1) return IIRX_ERR_INVALID;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return IIRX_ERR_NO_DEVICE;
    const size_t total = (size_t)offsets[n_channels];
    float *d_in = nullptr, *d_out = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&]() {
        if (e0) hipEventDestroy(e0);
        if (e1) hipEventDestroy(e1);
        if (s) hipStreamDestroy(s);
        if (d_in) hipFree(d_in);
        if (d_out) hipFree(d_out);
    };
    if (hipMalloc((void **)&d_in, total * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&d_out, total * sizeof(float)) != hipSuccess) {
        cleanup();
        return IIRX_ERR_OOM;
    }
    // N(0,1) input: one block of the generator repeated over the batch
    {
        const size_t blk = std::min<size_t>(total, (size_t)1 << 22);
        std::vector<float> h(blk);
        iirx_fill_normal(0x7476ULL, (int64_t)blk, h.data());
        for (size_t o = 0; o < total; o += blk)
            if (hipMemcpy(d_in + o, h.data(), std::min(blk, total - o) * sizeof(float), hipMemcpyHostToDevice) !=
                hipSuccess) {
                cleanup();
                return IIRX_ERR_HIP;
            }
    }
    if (hipStreamCreate(&s) != hipSuccess || hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        cleanup();
        return IIRX_ERR_HIP;
    }
    float best_t = -1.0f;
    for (int nb : {16, 32, 64})
        for (int bt : {16, 32, 64})
            for (int bm : {4, 8, 9}) {
                iirx_batch_config c = *cfg;
                c.n_blocks = nb;
                c.bs_tiles = bt;
                c.bs_min_sections = bm;
                iirx_batch *b = nullptr;
                if (iirx_batch_create(&c, n_channels, offsets, n_sections, sos, &b) != IIRX_OK) continue;
                int ok = IIRX_OK;
                for (int w = 0; w < 2 && ok == IIRX_OK; ++w) ok = iirx_batch_execute(b, d_in, d_out, s);
                if (ok == IIRX_OK) ok = hipEventRecord(e0, s) == hipSuccess ? IIRX_OK : IIRX_ERR_HIP;
                for (int r = 0; r < reps && ok == IIRX_OK; ++r) ok = iirx_batch_execute(b, d_in, d_out, s);
                if (ok == IIRX_OK) ok = hipEventRecord(e1, s) == hipSuccess ? IIRX_OK : IIRX_ERR_HIP;
                if (ok == IIRX_OK) ok = hipEventSynchronize(e1) == hipSuccess ? IIRX_OK : IIRX_ERR_HIP;
                if (ok == IIRX_OK) ok = iirx_batch_check(b);
                float ms = 0.0f;
                if (ok == IIRX_OK && hipEventElapsedTime(&ms, e0, e1) == hipSuccess) {
                    ms /= (float)reps;
                    if (best_t < 0.0f || ms < best_t) {
                        best_t = ms;
                        *best = c;
                    }
                }
                iirx_batch_destroy(b);
            }
    cleanup();
    if (best_t < 0.0f) return IIRX_ERR_GEOMETRY;
    if (best_ms) *best_ms = best_t;
    return IIRX_OK;
}

}  // extern "C"
