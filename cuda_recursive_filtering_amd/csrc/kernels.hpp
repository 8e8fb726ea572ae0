// kernels.hpp -- host/device contract between the plan (iirx_plan.cpp) and
// the CDNA4 kernels: launch arguments and the layout of the per-section
// coefficient tables the plan precomputes in FP64 and uploads once.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace iirx {

struct LaunchArgs {
    const float *in;
    float *out;
    const float *coef;           // per-section tables, `coef_stride` floats each
    uint64_t *lb;                // lookback records
    unsigned long long *ticket;  // 64-bit workgroup ticket counter
    int64_t n_samples;
    uint32_t num_chunks;         // grid size
    int32_t ns;                  // sections (PH/STCR/DTCR) or order k (PLR)
    int32_t coef_stride;
    int32_t unrolled;
};

constexpr int ilog2_c(int v) { return v <= 1 ? 0 : 1 + ilog2_c(v >> 1); }
constexpr int align4_c(int v) { return (v + 3) & ~3; }

// ---------------------------------------------------------------------------
// PH tables, one block per section (L = lanes per workgroup, N = samples per
// lane).  C is the N-step transition of the section state (y[n-2], y[n-1]),
// row-major 2x2 (ph_kernels.cuh:180-181).
//   [kB]     b1, b2, a1, a2          (a in recurrence sign, a = -sos[4:6])
//   [kCpow]  C^(2^k), k < log2 L     uniform lane-scan operators
//   [kM]     C^L                     chunk transition used by the lookback
//   [kCfin]  C^(t+1), t < L          per-lane carry injection (C_inner)
//   [kH1]    h1[n] = h0[n+1], n < N  zero-input response to y[-1]
//   [kH2]    h2[n] = a2 h0[n]        zero-input response to y[-2]
// ---------------------------------------------------------------------------
struct PhLayout {
    int kB, kCpow, kM, kCfin, kH1, kH2, stride;
};

constexpr PhLayout ph_layout(int L, int N)
{
    PhLayout p{};
    p.kB = 0;
    p.kCpow = 4;
    p.kM = p.kCpow + 4 * ilog2_c(L);
    p.kCfin = p.kM + 4;
    p.kH1 = p.kCfin + 4 * L;
    p.kH2 = p.kH1 + N;
    p.stride = align4_c(p.kH2 + N);
    return p;
}

// Launchers (one per algorithm); `lanes` and `n_blocks` pick the template
// instance.  Return hipErrorInvalidValue when no instance exists.
hipError_t launch_ph(int lanes, int n_blocks, const LaunchArgs &a, hipStream_t s);
const void *kernel_ph(int lanes, int n_blocks);

}  // namespace iirx
