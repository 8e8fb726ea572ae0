// dtcr.hip -- DTCR (decoupled two-sided cyclic reduction) for gfx950.
//
// Algorithm of include/dtcr_kernels.cuh (DTCR_32_* 291-2497, DTCR_64_*
// 2500-3522; operation order as tools/dtcr_emulator.py:79-304): the FIR and
// the first CR round are fused into one 4-tap pass over each parity
// sub-sequence ((1 + b1 z^-1 + b2 z^-2)(1 - e z^-1 + f z^-2), dtcr:134-139),
// after which even and odd samples obey independent recurrences; each side
// then runs the STCR rounds on N/2 values per lane.
//
// CDNA4 mapping.  A workgroup is two sibling wave64s over one chunk of 64
// rows x N samples: wave 0 owns the even samples of every row (side ty = 0),
// wave 1 the odd ones (ty = 1), N/2 registers per lane; lane tx is the row.
// (The reference puts each parity side on one 32-thread warp of a 32-row
// chunk, or on two warps per side for 64 rows.)  Everything inside a side is
// wave-local (lane shuffles); the sides meet only through LDS: the section
// input is read from the shared chunk buffer, which both waves rewrite with
// their outputs for the next section, and the chunk carry is exchanged there.
// Lookback records carry one word per side, so each wave publishes its own
// half of the chunk's end state.
#include "device_common.hpp"
#include "kernels.hpp"

namespace iirx {

// LAST = false: chunks 0 .. num_chunks - 1 under the plan's schedule.
// LAST = true: one workgroup for chunk num_chunks, the partial chunk that ends
// a length which is no multiple of the chunk, launched after the first launch
// on the same stream: its incoming state is the full record its predecessor
// left (no lookback), samples past the end read as zero and are not written,
// and with a.tail set it records the end state there.
template <int N, bool UNROLLED, bool LAST>
__global__ __launch_bounds__(128) void dtcr_kernel(LaunchArgs a)
{
    constexpr int L = 64;  // rows (lanes of one side)
    constexpr int LG = ilog2_c(N);
    constexpr int HALF = N / 2, QUART = HALF / 2;
    constexpr CrLayout Y = cr_layout(L);
    constexpr int RS = N + 1;  // padded LDS row
    __shared__ float inn[L * RS];
    __shared__ float xch[4];   // chunk carry broadcast: [-, X0, X1, -]

    const int tx = threadIdx.x & 63;
    const int ty = threadIdx.x >> 6;  // wave-uniform parity side

    const auto body = [&](uint32_t chunk, uint32_t part_flag, uint32_t full_flag) {
    const long base = (long)chunk * (L * N);
    __syncthreads();  // previous chunk's LDS consumed
    if constexpr (LAST) stage_in_part<L, N, 128>(a.in + base, inn, a.last_valid);
    else stage_in<L, N, 128>(a.in + base, inn);

    // raw input history before the chunk (sec 0, tx 0), read by both waves
    // before the barrier: the LAST instance may write the end state to a zf
    // that is the same buffer as zi, and does so only after it
    float g1 = 0.0f, g2 = 0.0f;
    if (chunk > 0) {
        g1 = a.in[base - 1];
        g2 = a.in[base - 2];
    } else if (a.zi) {  // streamed-in input history
        g1 = a.zi[1];
        g2 = a.zi[0];
    }
    __syncthreads();
    // the end state of the signal: samples last_valid - 2 and - 1 of the input
    // and of every section's output (sample -1: the value just ahead of the chunk)
    const int v = a.last_valid;
    const int t1 = ((v - 1) / N) * RS + (v - 1) % N;
    const int t2 = v >= 2 ? ((v - 2) / N) * RS + (v - 2) % N : 0;
    if constexpr (LAST)
        if (a.tail && threadIdx.x == 0) {
            a.tail[0] = v >= 2 ? inn[t2] : g1;
            a.tail[1] = inn[t1];
        }
    // boundary state carried between sections for tx == 0 (dtcr:337-371)
    float xi1 = 0.0f, xi2 = 0.0f, xi3 = 0.0f;
    float reg[HALF];

    for (int sec = 0; sec < a.ns; ++sec) {
        const float *cv = a.coef + (size_t)sec * a.coef_stride;
        const cfloat *c = as_const(cv);
        const float b1 = c[Y.kScal + 0], b2 = c[Y.kScal + 1];
        const float cb = c[Y.kScal + 2], db = c[Y.kScal + 3], fb = c[Y.kScal + 4], eb = c[Y.kScal + 5];
        const float hbc = c[Y.kScal + 6], gbc = c[Y.kScal + 7];
        const cfloat *E = c + Y.kE, *FDE = c + Y.kFde;
        const cfloat *CH = c + Y.kH, *CG = c + Y.kG, *CP = c + Y.kP, *CQ = c + Y.kQ;
        const cfloat *CD = c + Y.kD, *CC = c + Y.kC;
        const float *row = inn + tx * RS;
        const float *prow = inn + (tx > 0 ? tx - 1 : 0) * RS;  // read only for tx > 0

        // boundary history x[-1..-4] of this side's first tap (dtcr:337-371)
        float x1, x2, x3, x4;
        if (ty == 0) {
            if (tx == 0) {
                if (sec == 0) {
                    xi1 = g1;
                    xi2 = g2;
                }
                x1 = xi1;
                x2 = xi2;
                x3 = 0.0f;
                x4 = 0.0f;
            } else {
                x1 = prow[N - 1];
                x2 = prow[N - 2];
                x3 = prow[N - 3];
                x4 = prow[N - 4];
            }
        } else {
            x1 = row[0];
            if (tx == 0) {
                if (sec == 0) {
                    xi2 = g1;
                    xi3 = g2;
                }
                x2 = xi2;
                x3 = xi3;
                x4 = 0.0f;
            } else {
                x2 = prow[N - 1];
                x3 = prow[N - 2];
                x4 = prow[N - 3];
            }
        }

        // fused FIR + first CR round on this parity side (dtcr:373-415)
        {
            float c1, c2, c3;
            if (ty == 0) {
                c1 = (tx == 0) ? b1 : db;
                c2 = (tx == 0) ? b2 : eb;
                c3 = (tx == 0) ? 0.0f : cb;
            } else {
                c1 = db;
                c2 = (tx == 0) ? hbc : eb;
                c3 = (tx == 0) ? gbc : cb;
            }
            float xc = row[ty];
            reg[0] = fmaf(c1, x1, fmaf(c2, x2, fmaf(c3, x3, fmaf(fb, x4, xc))));
            x4 = x2;
            x3 = x1;
            x2 = xc;
#pragma unroll
            for (int n = 1; n < HALF; ++n) {
                xc = row[2 * n + ty];
                x1 = row[2 * n + ty - 1];
                reg[n] = fmaf(db, x1, fmaf(eb, x2, fmaf(cb, x3, fmaf(fb, x4, xc))));
                x4 = x2;
                x3 = x1;
                x2 = xc;
            }
        }

        // remaining forward CR rounds on the side (dtcr:417-449)
        if constexpr (UNROLLED) {
            static_for<1, LG, 1>([&](auto RR) {
                constexpr int r = decltype(RR)::value;
                constexpr int step = 1 << r, sub = step >> 1, off = sub - 1;
                const float fde = FDE[r], e = E[r];
                float y2 = lane_up(reg[HALF - 1], 1);
                if (tx == 0) y2 = 0.0f;
                static_for<0, HALF, step>([&](auto NN) {
                    constexpr int n = decltype(NN)::value;
                    reg[n + off] = fmaf(-fde, y2, reg[n + off]);
                    y2 = reg[n + off + sub];
                    reg[n + off + sub] = fmaf(-e, reg[n + off], y2);
                });
            });
        } else {
#pragma unroll
            for (int r = 1; r < LG; ++r) {
                const int step = 1 << r, sub = step >> 1, off = sub - 1;
                const float fde = FDE[r], e = E[r];
                float y2 = lane_up(reg[HALF - 1], 1);
                if (tx == 0) y2 = 0.0f;
#pragma unroll
                for (int n = 0; n < HALF; n += step) {
                    reg[n + off] = fmaf(-fde, y2, reg[n + off]);
                    y2 = reg[n + off + sub];
                    reg[n + off + sub] = fmaf(-e, reg[n + off], y2);
                }
            }
        }

        // lane recurrence of the side with zero incoming state (block filter)
        float q = 0.0f, yv = reg[HALF - 1];
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const int d = 1 << k;
            const cfloat *P = c + Y.kApow + 4 * k;
            const float tq = lane_up(q, d), ty_ = lane_up(yv, d);
            if (tx >= d) {
                q = fmaf(P[1], ty_, fmaf(P[0], tq, q));
                yv = fmaf(P[3], ty_, fmaf(P[2], tq, yv));
            }
        }

        // chunk partial: even side -> word 0, odd side -> word 1; lookback by wave 0
        uint64_t *rec = rec2(a.lb, chunk, sec, a.ns);
        const float pside = readlane_f(yv, L - 1);
        if (!LAST && tx == 0) rec_store(rec + ty, part_flag, pside);
        // the per-lane terminal coefficients: their load is in flight during the lookback
        const float4 tm = *reinterpret_cast<const float4 *>(cv + Y.kTerm + 4 * tx);
        if (ty == 0) {
            const float m0 = c[Y.kM + 0], m1 = c[Y.kM + 1], m2 = c[Y.kM + 2], m3 = c[Y.kM + 3];
            const int depth = (int)c[Y.kDepth];
            // the chain starts from the streamed-in state of this section (zero by default)
            const float z2 = a.zi ? a.zi[2 + 2 * sec] : 0.0f, z1 = a.zi ? a.zi[3 + 2 * sec] : 0.0f;
            const Carry2 X = LAST ? prior_state2(a.lb, chunk, sec, a.ns, z2, z1)
                                  : lookback2(a.lb, chunk, sec, a.ns, part_flag, full_flag, m0, m1, m2, m3, depth,
                                              a.spin_limit, a.err, 0u, z2, z1);
            if (tx == 0) {
                xch[1] = X.y2;
                xch[2] = X.y1;
            }
        }
        __syncthreads();
        const float X0 = xch[1], X1 = xch[2];

        // terminal correction and this side's word of the full record (dtcr:534-553)
        {
            const float c2 = (ty == 0) ? tm.x : tm.z;  // he2 | hb2
            const float c1 = (ty == 0) ? tm.y : tm.w;  // he1 | hb1
            reg[HALF - 1] = fmaf(-c1, X1, fmaf(-c2, X0, yv));
            const float fside = readlane_f(reg[HALF - 1], L - 1);
            if (!LAST && tx == 0) rec_store(rec + 2 + ty, full_flag, fside);
        }

        // deepest companion level (dtcr:555-590)
        float tmpv, xcv;
        {
            float y2, y1, h2b, h1b;
            const float d2v = (ty == 0) ? lane_up(reg[HALF - 1], 2) : lane_up(reg[HALF - 1], 1);
            const float d1v = (ty == 0) ? lane_up(reg[HALF - 1], 1) : lane_up(reg[HALF - 1], 2);
            if (tx == 0) {
                tmpv = X1;
                if (ty == 0) {
                    h2b = CP[LG - 1];
                    h1b = CQ[LG - 1];
                    xcv = X0;
                    xi1 = X1;
                    xi2 = X0;
                } else {
                    h2b = CH[LG - 1];
                    h1b = CG[LG - 1];
                    xcv = X1;
                    xi2 = X1;
                    xi3 = X0;
                }
                y2 = X0;
                y1 = X1;
            } else {
                y2 = d2v;
                y1 = d1v;
                if (ty == 0) {
                    h2b = CD[LG];
                    h1b = CC[LG];
                    tmpv = y1;
                } else {
                    h2b = CC[LG];
                    h1b = CD[LG];
                    tmpv = y2;
                }
                if (tx == 1) {
                    if (ty == 0) y2 = X0;
                    else y1 = X1;
                }
                xcv = tmpv;
            }
            reg[QUART - 1] = fmaf(-h1b, y1, fmaf(-h2b, y2, reg[QUART - 1]));
        }

        // back-substitution rounds (dtcr:592-644)
        auto round = [&](auto RR) {
            constexpr int r = decltype(RR)::value;
            constexpr int P = HALF >> r, sub = 1 << r, sub2 = sub >> 1;
            const float y2sh = lane_up(reg[HALF - 1 - sub], 1);
            const float dd = CD[r + 1], cc = CC[r + 1];
            float h2b, h1b, zi2, zi1;
            if (tx == 0) {
                h2b = (ty == 0) ? CP[r] : CH[r];
                h1b = (ty == 0) ? CQ[r] : CG[r];
                zi2 = X0;
            } else {
                h2b = dd;
                h1b = cc;
                zi2 = y2sh;
            }
            zi1 = tmpv;
            reg[sub2 - 1] = fmaf(-h1b, zi1, fmaf(-h2b, zi2, reg[sub2 - 1]));
            if constexpr (P > 1) {
                zi2 = xcv;
                zi1 = reg[sub - 1];
                reg[sub + sub2 - 1] = fmaf(-cc, zi1, fmaf(-dd, zi2, reg[sub + sub2 - 1]));
            }
            if constexpr (UNROLLED) {
                static_for<2, P, 1>([&](auto MM) {
                    constexpr int m = decltype(MM)::value;
                    zi2 = zi1;
                    zi1 = reg[sub * m - 1];
                    reg[sub * m + sub2 - 1] = fmaf(-cc, zi1, fmaf(-dd, zi2, reg[sub * m + sub2 - 1]));
                });
            } else {
#pragma unroll
                for (int m = 2; m < P; ++m) {
                    zi2 = zi1;
                    zi1 = reg[sub * m - 1];
                    reg[sub * m + sub2 - 1] = fmaf(-cc, zi1, fmaf(-dd, zi2, reg[sub * m + sub2 - 1]));
                }
            }
        };
        static_for_down<LG - 2, 0>(round);

        // section output back into the shared chunk buffer (next section's input)
        __syncthreads();  // every thread has read this section's input
        float *wrow = inn + tx * RS;
#pragma unroll
        for (int k = 0; k < HALF; ++k) wrow[2 * k + ty] = reg[k];
        __syncthreads();
        if constexpr (LAST)
            if (a.tail && threadIdx.x == 0) {
                a.tail[2 + 2 * sec] = v >= 2 ? inn[t2] : X1;
                a.tail[3 + 2 * sec] = inn[t1];
            }
    }

    if constexpr (LAST) stage_out_part<L, N, 128>(inn, a.out + base, a.last_valid);
    else stage_out<L, N, 128>(inn, a.out + base);
    };
    if constexpr (LAST) body(a.num_chunks, 0u, 0u);
    else for_each_chunk(a.ticket, a.epoch, a.arrive, a.num_chunks, a.persistent, a.spin_limit, a.err, body);
}

template <int N, bool U>
static hipError_t launch_n(const LaunchArgs &a, bool last, hipStream_t s)
{
    if (last) hipLaunchKernelGGL((dtcr_kernel<N, U, true>), dim3(1), dim3(128), 0, s, a);
    else hipLaunchKernelGGL((dtcr_kernel<N, U, false>), dim3(a.grid), dim3(128), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_dtcr(int lanes, int n_blocks, const LaunchArgs &a, hipStream_t s, bool last)
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

const void *kernel_dtcr(int lanes, int n_blocks, int unrolled)
{
    if (lanes != 64) return nullptr;
    const bool u = unrolled != 0;
    switch (n_blocks) {
    case 8: return u ? (const void *)dtcr_kernel<8, true, false> : (const void *)dtcr_kernel<8, false, false>;
    case 16: return u ? (const void *)dtcr_kernel<16, true, false> : (const void *)dtcr_kernel<16, false, false>;
    case 32: return u ? (const void *)dtcr_kernel<32, true, false> : (const void *)dtcr_kernel<32, false, false>;
    case 64: return u ? (const void *)dtcr_kernel<64, true, false> : (const void *)dtcr_kernel<64, false, false>;
    case 128: return u ? (const void *)dtcr_kernel<128, true, false> : (const void *)dtcr_kernel<128, false, false>;
    default: return nullptr;
    }
}

}  // namespace iirx
