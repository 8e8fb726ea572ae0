// stcr.hip -- STCR (single-stage tridiagonal cyclic reduction) for gfx950.
//
// Algorithm of include/stcr_kernels.cuh (STCR_32_LOOP 279-605, the
// hand-unrolled 32/64/128 forms 608-2584, STCR_64_* 2586-3608; operation
// order as tools/stcr_emulator.py:111-191):
//   1. fused FIR + first double-sided CR round in registers, per lane;
//   2. log2(N)-1 forward CR rounds, each needing one value of the previous
//      lane, leaving a second-order recurrence across lanes,
//      Y_t + e_LG Y_{t-1} + f_LG Y_{t-2} = R_t, for the last two samples;
//   3. that lane recurrence solved with zero incoming state (the reference's
//      Toeplitz "block filtering" with B-1 sequential shuffles, stcr:388-405)
//      -- here a six-level Kogge-Stone scan with the 2x2 operators A^(2^k) in
//      scalar registers;
//   4. lookback for the chunk's incoming state X, terminal correction with
//      the per-lane he/hb vectors, deepest companion level and log2(N)-2
//      back-substitution rounds.
//
// CDNA4 mapping as ph.hip: one wave64 workgroup = one chunk of 64 lanes x N
// samples (two of the reference's 32-lane chunks when BLOCK_SIZE = 32).
//
// Loop and unrolled forms.  The reference keeps a for-loop back
// substitution and hand-unrolled literal-index copies of it because its
// compiler spilled the unrolled loop (stcr_kernels.cuh:17-19).  Here the
// literal-index form is generated by compile-time index recursion
// (static_for) instead of being written out; both forms execute the same
// operation sequence and are tested for bit-identity on the device.
#include "device_common.hpp"
#include "kernels.hpp"

#include <algorithm>

namespace iirx {

// forward CR round `ro`, parity i (stcr:364-386): loop form
template <int N>
__device__ __forceinline__ void cr_round_loop(float (&r)[N], int ro_dyn, float fde, float e, float nb, int off,
                                              int sub, int step)
{
    (void)ro_dyn;
    float x2 = nb;
#pragma unroll
    for (int n = 0; n < N; n += step) {
        r[n + off] = fmaf(-fde, x2, r[n + off]);
        x2 = r[n + off + sub];
        r[n + off + sub] = fmaf(-e, r[n + off], x2);
    }
}

// LAST = false: chunks 0 .. num_chunks - 1 under the plan's schedule.
// LAST = true: one workgroup for chunk num_chunks, the partial chunk that ends
// a length which is no multiple of the chunk, launched after the first launch
// on the same stream: its incoming state is the full record its predecessor
// left (no lookback), samples past the end read as zero and are not written,
// and with a.tail set it records the end state there.
// WV > 1: the ticket schedule with WV chunks per workgroup (ph.hip)
template <int N, bool UNROLLED, bool LAST, int WV = 1>
__global__ __launch_bounds__(64 * WV) void stcr_kernel(LaunchArgs a)
{
    constexpr int L = 64;
    constexpr int LG = ilog2_c(N);
    constexpr int H = N / 2;
    constexpr CrLayout Y = cr_layout(L);
    __shared__ float lds_w[WV][L * (N + 1)];
    float *lds = lds_w[WV == 1 ? 0 : threadIdx.x >> 6];

    const int lane = WV == 1 ? (int)threadIdx.x : (int)(threadIdx.x & 63);
    const auto body = [&](uint32_t chunk, uint32_t part_flag, uint32_t full_flag) {
    const long base = (long)chunk * (L * N);

    chunk_sync<WV>();  // LDS of the previous chunk fully consumed
    if constexpr (LAST) stage_in_part<L, N>(a.in + base, lds, a.last_valid);
    else stage_in<L, N>(a.in + base, lds, lane);
    chunk_sync<WV>();

    float r[N];
#pragma unroll
    for (int n = 0; n < N; ++n) r[n] = lds[lane * (N + 1) + n];

    // section input history before the lane: x[-2], x[-1] and x[-4], x[-3]
    float xi2 = 0.0f, xi1 = 0.0f, w4 = 0.0f, w3 = 0.0f;
    if (lane > 0) {
        const float *pr = lds + (lane - 1) * (N + 1);
        xi2 = pr[N - 2];
        xi1 = pr[N - 1];
        w4 = pr[N - 4];
        w3 = pr[N - 3];
    } else if (chunk > 0) {
        xi2 = a.in[base - 2];
        xi1 = a.in[base - 1];
    } else if (a.zi) {  // streamed-in input history
        xi2 = a.zi[0];
        xi1 = a.zi[1];
    }
    // the end state of the signal: samples last_valid - 2 and - 1
    if constexpr (LAST)
        if (a.tail) tail_pair<N>(r, lds, a.last_valid, xi1, a.tail);

    for (int sec = 0; sec < a.ns; ++sec) {
        const float *cv = a.coef + (size_t)sec * a.coef_stride;  // per-lane vector loads
        const cfloat *c = as_const(cv);                           // uniform scalar loads
        const float b1 = c[Y.kScal + 0], b2 = c[Y.kScal + 1];
        const cfloat *E = c + Y.kE, *F = c + Y.kF, *FDE = c + Y.kFde;
        const cfloat *CH = c + Y.kH, *CG = c + Y.kG, *CP = c + Y.kP, *CQ = c + Y.kQ;
        const cfloat *CD = c + Y.kD, *CC = c + Y.kC;

        // FIR history (stcr:321-342): lane 0 of the chunk starts from zero
        float yi2 = 0.0f, yi1 = 0.0f;
        if (lane > 0) {
            yi2 = fmaf(b1, w3, fmaf(b2, w4, xi2));
            yi1 = fmaf(b1, xi2, fmaf(b2, w3, xi1));
        }

        // 1. FIR + first CR round (stcr:344-362)
        {
            const float me0 = -E[0], f0 = F[0];
            float x2 = xi2, x1 = xi1, y2 = yi2, y1 = yi1;
#pragma unroll
            for (int n = 0; n < N; ++n) {
                const float xc = r[n];
                const float yc = fmaf(b1, x1, fmaf(b2, x2, xc));
                r[n] = fmaf(me0, y1, fmaf(f0, y2, yc));
                x2 = x1;
                x1 = xc;
                y2 = y1;
                y1 = yc;
            }
        }

        // 2. forward CR rounds (stcr:364-386)
        if constexpr (UNROLLED) {
            static_for<1, LG, 1>([&](auto RO) {
                constexpr int ro = decltype(RO)::value;
                constexpr int step = 2 << ro, sub = 1 << ro;
                static_for<0, 2, 1>([&](auto I) {
                    constexpr int off = sub - 2 + decltype(I)::value;
                    float x2 = lane_up(r[N - 2 + decltype(I)::value], 1);
                    if (lane == 0) x2 = 0.0f;
                    const float fde = FDE[ro], e = E[ro];
                    static_for<0, N, step>([&](auto NN) {
                        constexpr int n = decltype(NN)::value;
                        r[n + off] = fmaf(-fde, x2, r[n + off]);
                        x2 = r[n + off + sub];
                        r[n + off + sub] = fmaf(-e, r[n + off], x2);
                    });
                });
            });
        } else {
#pragma unroll
            for (int ro = 1; ro < LG; ++ro) {
                const int step = 2 << ro, sub = 1 << ro;
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    float x2 = lane_up(r[N - 2 + i], 1);
                    if (lane == 0) x2 = 0.0f;
                    cr_round_loop<N>(r, ro, FDE[ro], E[ro], x2, sub - 2 + i, sub, step);
                }
            }
        }

        // 3. lane recurrence with zero incoming state: scan of (Y_{t-1}, Y_t)
        float qe = 0.0f, ye = r[N - 2];  // even sequence (position N-2)
        float qo = 0.0f, yo = r[N - 1];  // odd sequence (position N-1)
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const int d = 1 << k;
            const cfloat *P = c + Y.kApow + 4 * k;
            const float tqe = lane_up(qe, d), tye = lane_up(ye, d);
            const float tqo = lane_up(qo, d), tyo = lane_up(yo, d);
            if (lane >= d) {
                qe = fmaf(P[1], tye, fmaf(P[0], tqe, qe));
                ye = fmaf(P[3], tye, fmaf(P[2], tqe, ye));
                qo = fmaf(P[1], tyo, fmaf(P[0], tqo, qo));
                yo = fmaf(P[3], tyo, fmaf(P[2], tqo, yo));
            }
        }

        // 4. lookback (stcr:407-477)
        const float m0 = c[Y.kM + 0], m1 = c[Y.kM + 1], m2 = c[Y.kM + 2], m3 = c[Y.kM + 3];
        const float p2 = readlane_f(ye, L - 1), p1 = readlane_f(yo, L - 1);
        uint64_t *rec = rec2(a.lb, chunk, sec, a.ns);
        if (!LAST && lane == 0) publish2(rec, false, part_flag, p2, p1);
        const int depth = (int)c[Y.kDepth];
        // the chain starts from the streamed-in state of this section (zero by default)
        const float z2 = a.zi ? a.zi[2 + 2 * sec] : 0.0f, z1 = a.zi ? a.zi[3 + 2 * sec] : 0.0f;
        // the per-lane terminal coefficients: their load is in flight during the lookback
        const float4 tm = *reinterpret_cast<const float4 *>(cv + Y.kTerm + 4 * lane);
        const Carry2 Xc = LAST ? prior_state2(a.lb, chunk, sec, a.ns, z2, z1)
                               : lookback2(a.lb, chunk, sec, a.ns, part_flag, full_flag, m0, m1, m2, m3, depth,
                                           a.spin_limit, a.err, 0u, z2, z1);
        const float X0 = Xc.y2, X1 = Xc.y1;

        // terminal correction (stcr:480-491)
        {
            r[N - 2] = fmaf(-tm.y, X1, fmaf(-tm.x, X0, ye));
            r[N - 1] = fmaf(-tm.w, X1, fmaf(-tm.z, X0, yo));
        }
        {
            const float f2 = readlane_f(r[N - 2], L - 1), f1 = readlane_f(r[N - 1], L - 1);
            if (!LAST && lane == 0) publish2(rec, true, full_flag, f2, f1);
        }

        // deepest companion level (stcr:493-523)
        float y2e = lane_up(r[N - 2], 2), y2o = lane_up(r[N - 1], 1);
        float y1e = lane_up(r[N - 2], 1), y1o = lane_up(r[N - 1], 2);
        float h2e, h2o, h1e, h1o;
        if (lane == 0) {
            h2e = CP[LG - 1];
            h2o = CH[LG - 1];
            h1e = CQ[LG - 1];
            h1o = CG[LG - 1];
            y2e = X0;
            y2o = X0;
            y1e = X1;
            y1o = X1;
            xi2 = X0;
            xi1 = X1;
        } else {
            h2e = CD[LG];
            h2o = CC[LG];
            h1e = CC[LG];
            h1o = CD[LG];
            if (lane == 1) {
                y2e = X0;
                y1o = X1;
            }
            xi2 = y1e;
            xi1 = y2o;
        }
        r[H - 2] = fmaf(-h1e, y1e, fmaf(-h2e, y2e, r[H - 2]));
        r[H - 1] = fmaf(-h1o, y1o, fmaf(-h2o, y2o, r[H - 1]));

        // back-substitution rounds (stcr:527-585)
        auto round = [&](auto RR) {
            constexpr int rr = decltype(RR)::value;
            constexpr int stride = 2 << rr, sub = stride >> 1, P = H >> rr;
            const float se = lane_up(r[N - stride - 2], 1), so = lane_up(r[N - stride - 1], 1);
            float g2o, g1o, g2e, g1e, z2e, z1o;
            if (lane == 0) {
                g2o = CH[rr];
                g1o = CG[rr];
                g2e = CP[rr];
                g1e = CQ[rr];
                z2e = X0;
                z1o = X1;
            } else {
                g2o = CC[rr + 1];
                g1o = CD[rr + 1];
                g2e = g1o;
                g1e = g2o;
                z2e = se;
                z1o = so;
            }
            float z2o = y2o, z1e = y1e;
            r[sub - 2] = fmaf(-g1e, z1e, fmaf(-g2e, z2e, r[sub - 2]));
            r[sub - 1] = fmaf(-g1o, z1o, fmaf(-g2o, z2o, r[sub - 1]));
            const float dd = CD[rr + 1], cc = CC[rr + 1];
            z2o = xi1;
            z1o = r[stride - 1];
            z2e = xi2;
            z1e = r[stride - 2];
            r[stride + sub - 2] = fmaf(-cc, z1e, fmaf(-dd, z2e, r[stride + sub - 2]));
            r[stride + sub - 1] = fmaf(-cc, z1o, fmaf(-dd, z2o, r[stride + sub - 1]));
            if constexpr (UNROLLED) {
                static_for<2, P, 1>([&](auto MM) {
                    constexpr int m = decltype(MM)::value;
                    z2o = z1o;
                    z1o = r[stride * m - 1];
                    z2e = z1e;
                    z1e = r[stride * m - 2];
                    r[stride * m + sub - 2] = fmaf(-cc, z1e, fmaf(-dd, z2e, r[stride * m + sub - 2]));
                    r[stride * m + sub - 1] = fmaf(-cc, z1o, fmaf(-dd, z2o, r[stride * m + sub - 1]));
                });
            } else {
#pragma unroll
                for (int m = 2; m < P; ++m) {
                    z2o = z1o;
                    z1o = r[stride * m - 1];
                    z2e = z1e;
                    z1e = r[stride * m - 2];
                    r[stride * m + sub - 2] = fmaf(-cc, z1e, fmaf(-dd, z2e, r[stride * m + sub - 2]));
                    r[stride * m + sub - 1] = fmaf(-cc, z1o, fmaf(-dd, z2o, r[stride * m + sub - 1]));
                }
            }
        };
        static_for_down<LG - 2, 0>(round);

        // hand-off: the next section's input history is this section's output
        w4 = lane_up(r[N - 4], 1);
        w3 = lane_up(r[N - 3], 1);
        if constexpr (LAST)
            if (a.tail) tail_pair<N>(r, lds, a.last_valid, X1, a.tail + 2 + 2 * sec);
    }

    chunk_sync<WV>();
#pragma unroll
    for (int n = 0; n < N; ++n) lds[lane * (N + 1) + n] = r[n];
    chunk_sync<WV>();
    if constexpr (LAST) stage_out_part<L, N>(lds, a.out + base, a.last_valid);
    else stage_out<L, N>(lds, a.out + base, lane);
    };
    if constexpr (LAST) body(a.num_chunks, 0u, 0u);
    else for_each_chunk<WV>(a.ticket, a.epoch, a.arrive, a.num_chunks, a.persistent, a.spin_limit, a.err, body);
}

template <int N, bool U>
static hipError_t launch_n(const LaunchArgs &a, bool last, hipStream_t s)
{
    constexpr int WV = ph_ticket_waves(N);
    if (last) hipLaunchKernelGGL((stcr_kernel<N, U, true>), dim3(1), dim3(64), 0, s, a);
    else if (!a.persistent && a.num_chunks % WV == 0)
        hipLaunchKernelGGL((stcr_kernel<N, U, false, WV>), dim3(a.grid), dim3(64 * WV), 0, s, a);
    else hipLaunchKernelGGL((stcr_kernel<N, U, false>), dim3(a.grid), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_stcr(int lanes, int n_blocks, const LaunchArgs &a, hipStream_t s, bool last)
{
    if (lanes != 64) return hipErrorInvalidValue;
    const bool u = a.unrolled != 0;
    switch (n_blocks) {
    case 8: return u ? launch_n<8, true>(a, last, s) : launch_n<8, false>(a, last, s);
    case 16: return u ? launch_n<16, true>(a, last, s) : launch_n<16, false>(a, last, s);
    case 32: return u ? launch_n<32, true>(a, last, s) : launch_n<32, false>(a, last, s);
    case 64: return u ? launch_n<64, true>(a, last, s) : launch_n<64, false>(a, last, s);
    case 128: return u ? launch_n<128, true>(a, last, s) : launch_n<128, false>(a, last, s);
    default: return hipErrorInvalidValue;
    }
}

template <int N, bool U>
static const void *fn_n(bool grouped)
{
    if (grouped) return (const void *)stcr_kernel<N, U, false, ph_ticket_waves(N)>;
    return (const void *)stcr_kernel<N, U, false>;
}

const void *kernel_stcr(int lanes, int n_blocks, int unrolled, bool grouped)
{
    if (lanes != 64) return nullptr;
    const bool u = unrolled != 0;
    switch (n_blocks) {
    case 8: return u ? fn_n<8, true>(grouped) : fn_n<8, false>(grouped);
    case 16: return u ? fn_n<16, true>(grouped) : fn_n<16, false>(grouped);
    case 32: return u ? fn_n<32, true>(grouped) : fn_n<32, false>(grouped);
    case 64: return u ? fn_n<64, true>(grouped) : fn_n<64, false>(grouped);
    case 128: return u ? fn_n<128, true>(grouped) : fn_n<128, false>(grouped);
    default: return nullptr;
    }
}

}  // namespace iirx
