// iirx_driver -- the reference's two drivers (src/test_accuracy.cu,
// src/test_performance.cu) as one C++ program over the C ABI of
// include/iirx.h.
//
// The reference selects the kernel and geometry with nvcc -D flags
// (run_accuracy.sh:155-166) and builds one binary per configuration; here they
// are command-line options of a runtime plan:
//
//   iirx_driver accuracy|performance --kernel PH|PHJ|PHH|STCR|DTCR|PLR|BS|BS_BF16
//       [--block-size B] [--n-blocks N] [--sections NS] [--log2n L]
//       [--unrolled] [--plr-x-max X] [--schedule persistent|ticket]
//       [--taps FILE] [--write-taps FILE]
//       [--reference FILE] [--warmup W] [--batches K] [--batch M]
//   iirx_driver taps --sections NS --write-taps FILE    (header only, no GPU)
//
// Both modes filter the reference workload (ref_generate.py:56-99): a
// Butterworth lowpass of order 2 NS at 0.2 x Nyquist in b0-normalised SOS
// form, input FILTER_GAIN * cos(2 pi INPUT_FREQ i / INPUT_FS) computed in
// float64 and rounded to float32 once.  --taps takes the filter, gain, signal
// and PLR taps from a generated header instead (the reference's
// build/filter_taps.hpp, which its drivers compile in), with its
// FILTER_N_SECTIONS == N_SECTIONS check (test_accuracy.cu:107-109) made at
// run time; --write-taps writes the header of the filter used, in the
// reference's declarations, so the reference's drivers can be built on it.  accuracy makes one launch (test_accuracy.cu:
// 146-160); performance times P20 of batch means over rotating input buffers
// (test_performance.cu:144-198, measure.cuh:38-110).  With --reference, the
// output is checked against that file (float32, N samples: the float64
// sosfilt golden of ref_generate.py, the reference's build/reference.bin)
// with the reference's gate max|err| <= 1e-4 max|ref| (test_accuracy.cu:
// 178-190).  Exit status: 0 pass (or no reference given), 1 gate failed,
// 255 error, as the reference returns 0 / 1 / -1.
#include "iirx.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

struct Options {
    std::string mode;
    int kernel = IIRX_KERNEL_PH;
    int block_size = -1, n_blocks = 64, sections = 2, log2n = 20, unrolled = 0, plr_x_max = 0;
    int schedule = IIRX_SCHEDULE_PERSISTENT;
    std::string reference, taps, write_taps;
    bool sections_given = false;
    int warmup = 100, batches = 20, batch = 200;
};

int usage(const char *prog)
{
    std::printf("usage: %s taps --sections NS --write-taps FILE\n"
                "       %s accuracy|performance --kernel PH|PHJ|PHH|STCR|DTCR|PLR|BS|BS_BF16 [--block-size B] [--n-blocks N]\n"
                "       [--sections NS] [--log2n L] [--unrolled] [--plr-x-max X] [--schedule persistent|ticket]\n"
                "       [--taps FILE] [--write-taps FILE] [--reference FILE] [--warmup W] [--batches K] [--batch M]\n",
                prog, prog);
    return 0;
}

bool parse(int argc, char **argv, Options &o)
{
    if (argc < 2) return false;
    o.mode = argv[1];
    if (o.mode != "accuracy" && o.mode != "performance" && o.mode != "taps") return false;
    for (int i = 2; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&]() -> const char * { return (i + 1 < argc) ? argv[++i] : nullptr; };
        const char *v = nullptr;
        if (a == "--unrolled") {
            o.unrolled = 1;
            continue;
        }
        if (!(v = next())) return false;
        if (a == "--kernel") {
            const std::string k = v;
            if (k == "PH") o.kernel = IIRX_KERNEL_PH;
            else if (k == "STCR") o.kernel = IIRX_KERNEL_STCR;
            else if (k == "DTCR") o.kernel = IIRX_KERNEL_DTCR;
            else if (k == "PLR") o.kernel = IIRX_KERNEL_PLR;
            else if (k == "BS") o.kernel = IIRX_KERNEL_BLOCKSTATE;
            else if (k == "BS_BF16") o.kernel = IIRX_KERNEL_BLOCKSTATE_BF16;
            else if (k == "PHJ") o.kernel = IIRX_KERNEL_PH_JOINT;
            else if (k == "PHH") o.kernel = IIRX_KERNEL_PH_HALO;
            else return false;
        } else if (a == "--block-size") o.block_size = std::atoi(v);
        else if (a == "--n-blocks") o.n_blocks = std::atoi(v);
        else if (a == "--sections") {
            o.sections = std::atoi(v);
            o.sections_given = true;
        }
        else if (a == "--log2n") o.log2n = std::atoi(v);
        else if (a == "--plr-x-max") o.plr_x_max = std::atoi(v);
        else if (a == "--schedule") o.schedule = std::string(v) == "ticket" ? IIRX_SCHEDULE_TICKET : IIRX_SCHEDULE_PERSISTENT;
        else if (a == "--reference") o.reference = v;
        else if (a == "--taps") o.taps = v;
        else if (a == "--write-taps") o.write_taps = v;
        else if (a == "--warmup") o.warmup = std::atoi(v);
        else if (a == "--batches") o.batches = std::atoi(v);
        else if (a == "--batch") o.batch = std::atoi(v);
        else return false;
    }
    if (o.block_size < 0) o.block_size = (o.kernel == IIRX_KERNEL_PLR) ? 0 : 64;
    return o.sections >= 1 && o.sections <= 8 && o.log2n >= 6 && o.log2n <= 30;
}

// 0 pass, 1 fail, -1 unreadable reference
int check_reference(const std::string &path, const std::vector<float> &out)
{
    std::FILE *fp = std::fopen(path.c_str(), "rb");
    if (!fp) {
        std::fprintf(stderr, "ERROR: reference file %s not found\n", path.c_str());
        return -1;
    }
    std::vector<float> sol(out.size());
    const size_t got = std::fread(sol.data(), sizeof(float), sol.size(), fp);
    std::fclose(fp);
    if (got != sol.size()) {
        std::fprintf(stderr, "ERROR: %s has %zu floats, expected %zu\n", path.c_str(), got, sol.size());
        return -1;
    }
    const float rel_tol = 1e-4f;
    float max_err = 0.0f, max_ref = 0.0f;
    size_t worst = 0;
    bool finite = true;
    for (size_t i = 0; i < out.size(); ++i) {
        if (!std::isfinite(out[i])) finite = false;
        const float d = std::fabs(out[i] - sol[i]), r = std::fabs(sol[i]);
        if (d > max_err || std::isnan(d)) {
            max_err = std::isnan(d) ? INFINITY : d;
            worst = i;
        }
        max_ref = std::max(max_ref, r);
    }
    const float thr = rel_tol * max_ref;
    if (finite && max_err <= thr) {
        std::printf("PASSED (max abs err = %.3e, relative = %.3e, threshold = %.3e)\n", max_err, max_err / max_ref,
                    thr);
        return 0;
    }
    std::printf("FAILED: max abs err %.3e at index %zu exceeds threshold %.3e (relative = %.3e, ref=%.6f, out=%.6f)\n",
                max_err, worst, thr, max_err / max_ref, sol[worst], out[worst]);
    return 1;
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc >= 2 && (std::strcmp(argv[1], "--help") == 0 || std::strcmp(argv[1], "-h") == 0)) return usage(argv[0]);
    Options o;
    if (!parse(argc, argv, o)) {
        usage(argv[0]);
        return -1;
    }
    const int64_t n = 1LL << o.log2n;

    // filter and input of the reference workload (ref_generate.py:56-99), or
    // those of a generated header
    iirx_filter_taps taps{};
    if (!o.taps.empty()) {
        if (iirx_filter_taps_read(o.taps.c_str(), &taps) != IIRX_OK) {
            std::fprintf(stderr, "ERROR: cannot read filter header %s\n", o.taps.c_str());
            return -1;
        }
        if (o.sections_given && taps.n_sections != o.sections) {
            std::fprintf(stderr, "ERROR: FILTER_N_SECTIONS (%d) in %s != N_SECTIONS (%d)\n", taps.n_sections,
                         o.taps.c_str(), o.sections);
            return -1;
        }
        o.sections = taps.n_sections;
    } else if (iirx_filter_taps_design(o.sections, 0.2, &taps) != IIRX_OK) {
        return -1;
    }
    if (!o.write_taps.empty() && iirx_filter_taps_write(&taps, o.write_taps.c_str()) != IIRX_OK) {
        std::fprintf(stderr, "ERROR: cannot write %s\n", o.write_taps.c_str());
        return -1;
    }
    if (o.mode == "taps") return 0;  // header only, no device
    iirx_config cfg = {o.kernel, o.block_size, o.n_blocks, o.sections, o.unrolled, 0, o.plr_x_max, o.schedule};
    int st = iirx_config_check(&cfg, n);
    if (st != IIRX_OK) {
        std::fprintf(stderr, "ERROR: configuration rejected: %s\n", iirx_status_string(st));
        return -1;
    }

    std::vector<float> h_in((size_t)n), h_out((size_t)n);
    for (int64_t i = 0; i < n; ++i)
        h_in[(size_t)i] = (float)(taps.gain * std::cos(2.0 * M_PI * taps.input_freq * (double)i / taps.input_fs));

    iirx_plan *plan = nullptr;
    st = iirx_plan_create(&cfg, taps.sos, n, &plan);
    // PLR runs on the header's direct-form taps, as plr_kernels.cuh:416-427
    if (st == IIRX_OK && o.kernel == IIRX_KERNEL_PLR && !o.taps.empty())
        st = iirx_plan_set_plr_taps(plan, taps.plr_b, taps.plr_a, taps.plr_order);
    if (st != IIRX_OK) {
        std::fprintf(stderr, "ERROR: plan: %s\n", iirx_status_string(st));
        iirx_plan_destroy(plan);
        return -1;
    }
    iirx_plan_info info{};
    iirx_plan_info_get(plan, &info);
    hipDeviceProp_t prop{};
    hipGetDeviceProperties(&prop, 0);
    std::printf("=== %s test ===\n", o.mode == "accuracy" ? "Accuracy" : "Performance");
    std::printf("  GPU          = %s (%s)\n", prop.name, prop.gcnArchName);
    std::printf("  KERNEL       = %d%s\n", o.kernel, o.unrolled ? " (hand-unrolled)" : "");
    std::printf("  BLOCK_SIZE   = %d\n  N_BLOCKS     = %d\n  N_SECTIONS   = %d\n  N_SAMPLES    = %lld\n",
                o.block_size, o.n_blocks, o.sections, (long long)n);
    std::printf("  chunk        = %lld samples, %d threads, grid %lld, %d waves/SIMD, %d VGPRs, %d B scratch\n",
                (long long)info.chunk_samples, info.threads_per_block, (long long)info.grid, info.waves_per_simd,
                info.vgprs, info.scratch_bytes);

    const size_t bytes = (size_t)n * sizeof(float);
    float *d_out = nullptr;
    std::vector<float *> d_in;
    int rc = 0;
    if (hipMalloc((void **)&d_out, bytes) != hipSuccess) rc = -1;
    if (rc == 0 && o.mode == "accuracy") {
        float *p = nullptr;
        if (hipMalloc((void **)&p, bytes) != hipSuccess) rc = -1;
        else d_in.push_back(p);
        if (rc == 0 && hipMemcpy(p, h_in.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) rc = -1;
        if (rc == 0 && (iirx_execute(plan, p, d_out, nullptr) != IIRX_OK || iirx_plan_check(plan) != IIRX_OK)) {
            std::fprintf(stderr, "ERROR: kernel execution failed\n");
            rc = -1;
        }
    } else if (rc == 0) {
        // enough rotating inputs that they do not stay in the 256 MB MALL
        // (the reference sizes them against its L2, test_performance.cu:144-157)
        const size_t mall = 256ull << 20;
        const int nbuf = std::max<int>(4, (int)((2 * mall + bytes - 1) / bytes) + 1);
        for (int i = 0; i < nbuf && rc == 0; ++i) {
            float *p = nullptr;
            if (hipMalloc((void **)&p, bytes) != hipSuccess || hipMemcpy(p, h_in.data(), bytes, hipMemcpyHostToDevice) != hipSuccess)
                rc = -1;
            else
                d_in.push_back(p);
        }
        std::vector<float> batch_ms((size_t)std::max(o.batches, 1));
        if (rc == 0 && iirx_measure(plan, d_in.data(), (int)d_in.size(), d_out, o.warmup, o.batches, o.batch, nullptr,
                                    batch_ms.data()) != IIRX_OK) {
            std::fprintf(stderr, "ERROR: timing failed\n");
            rc = -1;
        }
        if (rc == 0) {
            const float p20 = batch_ms[(size_t)o.batches / 5];
            std::printf("  Input buffers = %d (total footprint = %.0f MB)\n", (int)d_in.size(),
                        d_in.size() * bytes / 1048576.0);
            std::printf("Execution time (P20): %.5f ms\n", p20);
            std::printf("Throughput:           %.3f GS/s\n", (double)n / (p20 * 1e-3) / 1e9);
            // the output of the last timed call filters the same input
            if (iirx_plan_check(plan) != IIRX_OK) rc = -1;
        }
    }
    if (rc == 0 && hipMemcpy(h_out.data(), d_out, bytes, hipMemcpyDeviceToHost) != hipSuccess) rc = -1;
    if (rc == 0 && !o.reference.empty()) {
        if (o.mode == "performance") std::printf("Accuracy: ");
        rc = check_reference(o.reference, h_out);
    }
    for (float *p : d_in) hipFree(p);
    if (d_out) hipFree(d_out);
    iirx_plan_destroy(plan);
    return rc;
}
