// bs.hip -- block-state form of the cascade on the MI355X matrix cores.
//
// No counterpart in the reference; it computes the same filter as
// ph/stcr/dtcr (the whole cascade, zero initial state) in a form whose cost
// grows with the order only through the state dimension of small products.
// The cascade is one state-space system of dimension p = 2 N_SECTIONS,
// embedded in 16 rows.  A chunk of G (16, 32 or 64) tiles of L (32, 64 or
// 128; G <= 32 at L = 128) samples is one wave64:
//   1. the zero-state parts Y0 = T X (T: L x L lower-triangular Toeplitz of
//      the impulse response) and V = Gam X (Gam: 16 x L), as
//      v_mfma_f32_16x16x32_f16 products of FP16 hi/lo operands
//      (hi*hi + hi*lo + lo*hi, FP32 accumulation); T and Gam are scaled by
//      a power of two on the host, X by one per chunk, so that FP16 holds
//      them; blocks of T above the diagonal are skipped;
//   2. the tile end states Z_k = sum_{j<=k} Phi^(k-j) V_j by a log2(G)-level
//      Kogge-Stone scan over the G tiles, each level Phi^(2^j) * shift(Z)
//      on v_mfma_f64_16x16x4f64 (FP32 operands Phi^(2^j), FP64 states and
//      accumulation), over ceil(p/4) K slabs only;
//   3. the chunk's incoming state X_in from the lookback (16-word records,
//      chunk transition Phi^G, bounded depth), injected as Phi^(k+1) X_in in
//      FP64: Phi X_in in tile 0, a 4-level scan inside one 16-tile block and
//      Phi^(16 nb) products for the other blocks;
//   4. Y = Y0 + O S, S_k = state entering tile k rounded to FP32,
//      accumulated onto Y0 by the FP32 matrix core; from 9 states on (orders
//      above 8) by the BF16 one, O and S split hi/mid/lo, the six products
//      above 2^-24 in pairs, three 16 x 16 x 32 MFMAs per block (about 18
//      clk each) against up to four 16 x 16 x 4 FP32 ones (32 clk each).
//      Order 16, 2^25 samples: 466 against 452 GS/s, max error 4.88e-6
//      against 5.08e-6 of the output RMS; at order 8 (two FP32 products) the
//      BF16 form ran 0.6% faster with a 10% larger max error and is not used.
// MFMA columns are tiles, so every product feeds the next from registers and
// the scan's tile shifts are DPP row rotations.  The FP64 shape puts row
// l/16 + 4 j of its result on lane l, register j: states are kept in that
// order (lane group g, register j = state g + 4 j) by the operand packing,
// including Gam's row order in the FP16 product.
#include "bs_kernel.hpp"

namespace iirx {

// ---------------------------------------------------------------------------
// streaming state.  The realization's state c is the direct-form-II-
// transposed state of section c / 2 divided by d[c]: after sample n,
// w1 = b1 u[n] + b2 u[n-1] - a1 y[n] - a2 y[n-1] and w2 = b2 u[n] - a2 y[n]
// with u the section's input and y its output.  One thread each: the maps
// run once per call.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void bs_state_in_kernel(const float *zi, const BsStateMap *m, float *z0)
{
    if (threadIdx.x != 0) return;
    float z[16];
    for (int c = 0; c < 16; ++c) z[c] = 0.0f;
    double u2 = zi[0], u1 = zi[1];  // x[-2], x[-1]
    for (int s = 0; s < m->ns; ++s) {
        const double b1 = m->c[s][0], b2 = m->c[s][1], a1 = m->c[s][2], a2 = m->c[s][3];
        const double y2 = zi[2 + 2 * s], y1 = zi[3 + 2 * s];
        const double w1 = b1 * u1 + b2 * u2 - a1 * y1 - a2 * y2;
        const double w2 = b2 * u1 - a2 * y1;
        z[2 * s] = (float)(w1 / m->d[2 * s]);
        z[2 * s + 1] = (float)(w2 / m->d[2 * s + 1]);
        u2 = y2;  // the next section's input is this section's output
        u1 = y1;
    }
    for (int c = 0; c < 16; ++c) z0[c] = z[c];
}

// zf = x[n-2], x[n-1], then y_s[n-2], y_s[n-1] of every section: the state
// entering the last tile (after sample n - L - 1) mapped back to w and the
// tile's L inputs run through the sections in FP64 (y = u + w1,
// w1 <- -a1 w1 + w2 + (b1 - a1) u, w2 <- -a2 w1 + (b2 - a2) u).  Everything
// is read before zf is written, so zf may be the zi of this call.
__global__ __launch_bounds__(64) void bs_state_out_kernel(const float *in, int64_t n, int L, const float *tail,
                                                           const BsStateMap *m, float *zf)
{
    if (threadIdx.x != 0) return;
    const int ns = m->ns;
    double w[16];
    for (int c = 0; c < 2 * ns; ++c) w[c] = (double)tail[c] * m->d[c];
    double hist[8][2] = {};
    for (int64_t t = n - L; t < n; ++t) {
        double u = in[t];
        for (int s = 0; s < ns; ++s) {
            const double b1 = m->c[s][0], b2 = m->c[s][1], a1 = m->c[s][2], a2 = m->c[s][3];
            const double y = u + w[2 * s];
            const double w1 = -a1 * w[2 * s] + w[2 * s + 1] + (b1 - a1) * u;
            const double w2 = -a2 * w[2 * s] + (b2 - a2) * u;
            w[2 * s] = w1;
            w[2 * s + 1] = w2;
            if (t >= n - 2) hist[s][t - (n - 2)] = y;
            u = y;
        }
    }
    const float x2 = in[n - 2], x1 = in[n - 1];
    zf[0] = x2;
    zf[1] = x1;
    for (int s = 0; s < ns; ++s) {
        zf[2 + 2 * s] = (float)hist[s][0];
        zf[3 + 2 * s] = (float)hist[s][1];
    }
}

hipError_t launch_bs_state_in(const float *zi, const BsStateMap *m, float *z0, hipStream_t s)
{
    hipLaunchKernelGGL(bs_state_in_kernel, dim3(1), dim3(64), 0, s, zi, m, z0);
    return hipGetLastError();
}

hipError_t launch_bs_state_out(const float *in, int64_t n, int L, const float *tail, const BsStateMap *m, float *zf,
                               hipStream_t s)
{
    hipLaunchKernelGGL(bs_state_out_kernel, dim3(1), dim3(64), 0, s, in, n, L, tail, m, zf);
    return hipGetLastError();
}

hipError_t launch_bs(int L, int G, int ns, const LaunchArgs &a, hipStream_t s, int split)
{
    if (split == kBsSplitBF16) return launch_bs_bf16(L, G, bs_slabs(ns), a, s);
    if (split != kBsSplitF16) return hipErrorInvalidValue;
    switch (L) {
    case 32: return launch_l<32, kBsSplitF16>(G, bs_slabs(ns), a, s);
    case 64: return launch_l<64, kBsSplitF16>(G, bs_slabs(ns), a, s);
    case 128: return launch_l<128, kBsSplitF16>(G, bs_slabs(ns), a, s);
    default: return hipErrorInvalidValue;
    }
}

const void *kernel_bs(int L, int G, int ns, int split)
{
    if (split == kBsSplitBF16) return kernel_bs_bf16(L, G, bs_slabs(ns));
    if (split != kBsSplitF16) return nullptr;
    switch (L) {
    case 32: return fn_l<32, kBsSplitF16>(G, bs_slabs(ns));
    case 64: return fn_l<64, kBsSplitF16>(G, bs_slabs(ns));
    case 128: return fn_l<128, kBsSplitF16>(G, bs_slabs(ns));
    default: return nullptr;
    }
}

}  // namespace iirx
