// plr.hip -- PLR (Maleki & Burtscher's direct-form hierarchical merge) for
// gfx950.
//
// Algorithm of include/plr_kernels.cuh (PLR_2/4/8/16, 483-3108): the whole
// cascade is one direct-form recurrence of order K = 2 N_SECTIONS.  A chunk
// of BLOCK threads x X values, thread t owning samples t + v BLOCK
// (plr:524-537, tail-safe), goes through
//   1. the K-tap FIR (plr:539-562);
//   2. a Sklansky merge over the BLOCK lanes of each segment v: at half width
//      h, lane i of every upper half adds fac[m][i % h] times the last
//      min(K, h) values of the lower half, m ascending (plr:564-785);
//   3. the serial chain over the X segments (plr:787-811);
//   4. the lookback for the K-value chunk carry and its injection
//      val += sum_m fac[m][t + v BLOCK] X_m (plr:813-899).
//
// CDNA4 mapping.  Merge levels h <= 32 stay inside a wave64 (lane shuffles);
// levels h >= 64 cross waves through LDS, as do the segment chain and the
// carry broadcast.  The reference's intra-warp levels stop at 32 lanes and
// cross warps from h = 32 on; the arithmetic rule is the same.  The lookback
// (lookbackK) scans predecessors one chunk at a time with the K words of a
// record on K lanes and folds partials with the FP64 chunk transition F.
#include "device_common.hpp"
#include "kernels.hpp"

namespace iirx {

// one-wave chunks (B = 64): value of lane (lane & ~(2H-1)) + H - K + M, the
// merge level H source of the upper half of every 2H-lane group, without LDS
// traffic: DPP for H = 1 (row_shr:1) and H = 2 (quad_perm), ds_swizzle bit
// mode ((lane & and) | or within 32 lanes) for H = 4..16, readlane for H = 32
template <int H, int K, int M>
__device__ __forceinline__ float plr_bcast(float v)
{
    const int u = __float_as_int(v);
    if constexpr (H == 1) {
        return __int_as_float(__builtin_amdgcn_update_dpp(u, u, 0x111, 0xF, 0xF, false));
    } else if constexpr (H == 2) {
        constexpr int q = M - K + 2;  // quad lane 0 or 1
        return __int_as_float(__builtin_amdgcn_update_dpp(u, u, q * 0x55, 0xF, 0xF, false));
    } else if constexpr (H < 32) {
        constexpr int and_mask = (~(2 * H - 1)) & 0x1F, or_mask = (H - K + M) & 0x1F;
        return __int_as_float(__builtin_amdgcn_ds_swizzle(u, and_mask | (or_mask << 5)));
    } else {
        return __int_as_float(__builtin_amdgcn_readlane(u, 32 - K + M));
    }
}

// merge factors hoisted per lane (one-wave chunks, K <= 4): level H takes
// min(K, H) of them, at offset sum over lower levels
template <int K>
constexpr int plr_merge_off(int lv)
{
    int o = 0;
    for (int l = 0; l < lv; ++l) o += ((1 << l) < K) ? (1 << l) : K;
    return o;
}

template <int K, int B, int X>
__global__ __launch_bounds__(B) void plr_kernel(LaunchArgs a)
{
    constexpr int W = B / 64;
    constexpr int CH = B * X;
    // the chunk sits at xs[HB ..] (16-byte aligned for whole-line staging),
    // its K samples of input history just before it
    constexpr int HB = 16;
    static_assert(K <= HB, "history fits before the chunk");
    constexpr PlrLayout Y = plr_layout(K, CH);
    // one-wave chunks with K <= 4 keep the merge factors in registers
    constexpr bool kHoist = (B == 64) && (K <= 4);
    constexpr int NM = plr_merge_off<K>(6);
    __shared__ __attribute__((aligned(16))) float xs[HB + CH];  // raw chunk with K samples of history
    __shared__ float sp[W * X * K];      // last K values of every wave and segment
    __shared__ float sv[K];              // segment-chain and carry broadcast
    __shared__ float lp[64 * K];         // lookback partials, nearest first

    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const float *tab = a.coef;
    const cfloat *ctab = as_const(a.coef);
    const float *fac = tab + Y.kFac;
    const cfloat *cfac = ctab + Y.kFac;

    // chunk-invariant per-lane factors, loaded once per workgroup: fac[m][t]
    // (segment chain and carry injection) and the merge factors
    float ft[K];
#pragma unroll
    for (int m = 0; m < K; ++m) ft[m] = fac[m * CH + t];
    float fmg[kHoist ? NM : 1];
    if constexpr (kHoist) {
        static_for<0, 6, 1>([&](auto LV) {
            constexpr int lv = decltype(LV)::value, H = 1 << lv, M0 = K > H ? K - H : 0;
            static_for<M0, K, 1>([&](auto MI) {
                constexpr int M = decltype(MI)::value;
                fmg[plr_merge_off<K>(lv) + M - M0] = fac[M * CH + (lane & (H - 1))];
            });
        });
    }
    // one-wave chunks with K X <= 64: lane l < K X computes W_v[m] of the carry
    // injection (v = l / K, m = l % K) from K factors it holds for the launch
    constexpr bool kDist = (B == 64) && (K * X <= 64) && (K < 16);
    float wc[kDist ? K : 1];
    if constexpr (kDist) {
        const int v = lane / K, m = lane % K;
#pragma unroll
        for (int n = 0; n < K; ++n) wc[n] = (lane < K * X && v > 0) ? fac[n * CH + v * B - K + m] : 0.0f;
    }
    // whole chunks move as 16-byte accesses staged through LDS when both
    // buffers are 16-byte aligned (every chunk start then is); otherwise,
    // and for the partial last chunk, one float per lane and segment
    const bool vec_io = (X % 4 == 0) &&
                        ((reinterpret_cast<uintptr_t>(a.in) | reinterpret_cast<uintptr_t>(a.out)) & 15) == 0;

    for_each_chunk(a.ticket, a.epoch, a.arrive, a.num_chunks, a.persistent, a.spin_limit, a.err,
                   [&](uint32_t chunk, uint32_t part_flag, uint32_t full_flag) {
    const long cs = (long)chunk * CH;
    const bool whole = vec_io && cs + CH <= a.n_samples;
    __syncthreads();
    float val[X];
    if (whole) {
        const f4 *src = reinterpret_cast<const f4 *>(a.in + cs);
        f4 *dst = reinterpret_cast<f4 *>(xs + HB);
        f4 q4[X / 4 > 0 ? X / 4 : 1];
#pragma unroll
        for (int q = 0; q < X / 4; ++q) q4[q] = src[t + q * B];
#pragma unroll
        for (int q = 0; q < X / 4; ++q) dst[t + q * B] = q4[q];
    } else {
#pragma unroll
        for (int v = 0; v < X; ++v) {
            const long idx = cs + t + (long)v * B;
            xs[HB + t + v * B] = (idx < a.n_samples) ? a.in[idx] : 0.0f;
        }
    }
    if (t < K) {
        // input history; before sample 0 the streamed-in x[-K..-1] (zero by default)
        const long hi = cs - K + t;
        xs[HB - K + t] = (hi >= 0) ? a.in[hi] : (a.zi ? a.zi[K + hi] : 0.0f);
    }
    __syncthreads();
#pragma unroll
    for (int v = 0; v < X; ++v) val[v] = xs[HB + t + v * B];

    // 1. FIR: val += b[j] x[i-1-j], j ascending
#pragma unroll
    for (int v = 0; v < X; ++v) {
#pragma unroll
        for (int j = 0; j < K; ++j) val[v] = fmaf(ctab[Y.kB + j], xs[HB + t + v * B - 1 - j], val[v]);
    }

    // 2a. merge levels inside the wave
    if constexpr (B == 64) {
        static_for<0, 6, 1>([&](auto LV) {
            constexpr int lv = decltype(LV)::value, H = 1 << lv, M0 = K > H ? K - H : 0;
            const bool up = (lane & H) != 0;
            const int i = lane & (H - 1);
            static_for<M0, K, 1>([&](auto MI) {
                constexpr int M = decltype(MI)::value;
                float f;
                if constexpr (kHoist) f = fmg[plr_merge_off<K>(lv) + M - M0];
                else f = fac[M * CH + i];
                if constexpr (K <= 4) {
                    // every lane takes part in the broadcasts (their sources
                    // are the lower half); the FMAs run on the upper half only,
                    // one exec-masked block instead of a select per value
                    // (2^25 samples: 565 -> 630 GS/s at order 2, 335 -> 379 at
                    // order 4; order 8 ran 121 -> 98 and keeps the selects)
                    float bc[X];
#pragma unroll
                    for (int v = 0; v < X; ++v) bc[v] = plr_bcast<H, K, M>(val[v]);
                    if (up) {
#pragma unroll
                        for (int v = 0; v < X; ++v) val[v] = fmaf(f, bc[v], val[v]);
                    }
                } else {
#pragma unroll
                    for (int v = 0; v < X; ++v) {
                        const float sv_ = plr_bcast<H, K, M>(val[v]);
                        if (up) val[v] = fmaf(f, sv_, val[v]);
                    }
                }
            });
        });
    } else {
#pragma unroll
    for (int h = 1; h < 64 && h < B; h <<= 1) {
        const bool up = (lane & h) != 0;
        const int base = lane & ~(2 * h - 1);
        const int i = lane & (h - 1);
#pragma unroll
        for (int m = (K > h ? K - h : 0); m < K; ++m) {
            const float f = fac[m * CH + i];
            const int src = base + h - K + m;
#pragma unroll
            for (int v = 0; v < X; ++v) {
                const float s = __shfl(val[v], src, 64);
                if (up) val[v] = fmaf(f, s, val[v]);
            }
        }
    }
    }

    // 2b. merge levels across waves (h >= 64 >= K: all K terms)
#pragma unroll
    for (int h = 64; h < B; h <<= 1) {
        __syncthreads();
        if (lane >= 64 - K) {
#pragma unroll
            for (int v = 0; v < X; ++v) sp[(w * X + v) * K + lane - (64 - K)] = val[v];
        }
        __syncthreads();
        if (t & h) {
            const int sw = ((t & ~(2 * h - 1)) + h - 1) >> 6;
            const int i = t & (h - 1);
#pragma unroll
            for (int m = 0; m < K; ++m) {
                const float f = fac[m * CH + i];
#pragma unroll
                for (int v = 0; v < X; ++v) val[v] = fmaf(f, sp[(sw * X + v) * K + m], val[v]);
            }
        }
    }

    // 3. serial chain over the segments (one wave: the last K values of the
    //    previous segment by readlane)
#pragma unroll
    for (int v = 1; v < X; ++v) {
        if constexpr (B == 64) {
#pragma unroll
            for (int m = 0; m < K; ++m) {
                const float pm = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(val[v - 1]), B - K + m));
                val[v] = fmaf(ft[m], pm, val[v]);
            }
        } else {
            __syncthreads();
            if (t >= B - K) sv[t - (B - K)] = val[v - 1];
            __syncthreads();
#pragma unroll
            for (int m = 0; m < K; ++m) val[v] = fmaf(ft[m], sv[m], val[v]);
        }
    }

    // 4. chunk partial = last K values; lookback by wave 0
    uint64_t *rec = a.lb + (size_t)chunk * (2 * K);
    if (t >= B - K) rec_store(rec + (t - (B - K)), part_flag, val[X - 1]);
    float Xv[K];  // the chunk carry, uniform
    float Xl = 0.0f;  // one-wave chunks: lane m < K holds X_m
    if constexpr (B == 64) {
        const float Xj = lookbackK<K>(a.lb, 2 * K, chunk, part_flag, full_flag, tab + Y.kF, (int)ctab[Y.kDepth], lp,
                                      a.spin_limit, a.err, nullptr, 0u, a.zi ? a.zi + K : nullptr);
        Xl = Xj;
#pragma unroll
        for (int m = 0; m < K; ++m) Xv[m] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(Xj), m));
    } else {
        __syncthreads();  // everyone is done with sv
        if (w == 0) {
            const float Xj = lookbackK<K>(a.lb, 2 * K, chunk, part_flag, full_flag, tab + Y.kF, (int)ctab[Y.kDepth],
                                          lp, a.spin_limit, a.err, nullptr, 0u, a.zi ? a.zi + K : nullptr);
            if (lane < K) sv[lane] = Xj;
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < K; ++m) Xv[m] = sv[m];
    }

    // carry injection (plr:891-899).  The carry's share of sample t + v B is
    // sum_m fac[m][t + v B] X_m; it equals the zero-input response at lane t
    // of the carry's share of the K values before segment v,
    // W_v[m] = sum_n fac[n][v B - K + m] X_n (W_0 = X), so lane t adds
    // sum_m fac[m][t] W_v[m] with the K per-lane factors of the segment chain
    // and uniform W_v, instead of K X per-lane table loads per chunk.  At
    // K = 16 the K^2 X uniform products (and their scalar table loads) cost
    // more than the loads they replace: the table form there (order 16, 2^25
    // samples: 27.7 GS/s against 10.3; order 8: 65 against 121).
    if constexpr (K >= 16) {
#pragma unroll
        for (int m = 0; m < K; ++m) {
            const float Xm = Xv[m];
#pragma unroll
            for (int v = 0; v < X; ++v) val[v] = fmaf(fac[m * CH + t + v * B], Xm, val[v]);
        }
    } else if constexpr (kDist) {
        // W_v[m] on lane v K + m (W_0 = X: lanes m < K hold X_m), K FMAs on
        // every lane and one readlane per value instead of K^2 X uniform FMAs
        // (2^25 samples: 618 -> 642 GS/s at order 2, 381 -> 426 at order 4)
        float w = 0.0f;
#pragma unroll
        for (int n = 0; n < K; ++n) w = fmaf(wc[n], Xv[n], w);
        if (lane < K) w = Xl;
#pragma unroll
        for (int v = 0; v < X; ++v) {
#pragma unroll
            for (int m = 0; m < K; ++m) val[v] = fmaf(ft[m], readlane_f(w, v * K + m), val[v]);
        }
    } else {
#pragma unroll
    for (int v = 0; v < X; ++v) {
        float Wv[K];
#pragma unroll
        for (int m = 0; m < K; ++m) {
            if (v == 0) {
                Wv[m] = Xv[m];
            } else {
                float acc = 0.0f;
#pragma unroll
                for (int n = 0; n < K; ++n) acc = fmaf(cfac[n * CH + v * B - K + m], Xv[n], acc);
                Wv[m] = acc;
            }
        }
#pragma unroll
        for (int m = 0; m < K; ++m) val[v] = fmaf(ft[m], Wv[m], val[v]);
    }
    }
    if (t >= B - K) rec_store(rec + K + (t - (B - K)), full_flag, val[X - 1]);

    if (whole) {
        // every wave is past its FIR reads of xs (the lookback and merge
        // barriers, or program order within one wave)
        __syncthreads();
#pragma unroll
        for (int v = 0; v < X; ++v) xs[HB + t + v * B] = val[v];
        __syncthreads();
        const f4 *srcl = reinterpret_cast<const f4 *>(xs + HB);
        f4 *dst = reinterpret_cast<f4 *>(a.out + cs);
#pragma unroll
        for (int q = 0; q < X / 4; ++q) dst[t + q * B] = srcl[t + q * B];
    } else {
#pragma unroll
        for (int v = 0; v < X; ++v) {
            const long idx = cs + t + (long)v * B;
            if (idx < a.n_samples) a.out[idx] = val[v];
        }
    }
    });
}

// zf = the direct-form history after the call: x[n-K..n-1], then
// y[n-K..n-1] (from zi where the call is shorter than K).  Every value is read
// before any is written, so zf may be the zi of this call.
__global__ __launch_bounds__(64) void plr_end_state_kernel(LaunchArgs a, float *zf)
{
    const int t = threadIdx.x, K = a.ns;
    float v = 0.0f;
    if (t < 2 * K) {
        const int m = t % K;
        const long j = a.n_samples - K + m;  // sample index, < 0: history before the call
        const float *src = (t < K) ? a.in : a.out;
        if (j >= 0) v = src[j];
        else if (a.zi) v = a.zi[(t < K ? 0 : K) + K + j];
    }
    __syncthreads();
    if (t < 2 * K) zf[t] = v;
}

hipError_t launch_plr_end_state(const LaunchArgs &a, float *zf, hipStream_t s)
{
    hipLaunchKernelGGL(plr_end_state_kernel, dim3(1), dim3(64), 0, s, a, zf);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// instance table: K in {2, 4, 8, 16}, BLOCK in {64, 128, 256, 512},
// X in {1, 2, 4, 8, 16}
// ---------------------------------------------------------------------------
template <int K, int B, int X>
static hipError_t launch_kbx(const LaunchArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL((plr_kernel<K, B, X>), dim3(a.grid), dim3(B), 0, s, a);
    return hipGetLastError();
}

template <int K, int B>
static const void *fn_kb(int x)
{
    switch (x) {
    case 1: return (const void *)plr_kernel<K, B, 1>;
    case 2: return (const void *)plr_kernel<K, B, 2>;
    case 4: return (const void *)plr_kernel<K, B, 4>;
    case 8: return (const void *)plr_kernel<K, B, 8>;
    case 16: return (const void *)plr_kernel<K, B, 16>;
    default: return nullptr;
    }
}

template <int K, int B>
static hipError_t launch_kb(int x, const LaunchArgs &a, hipStream_t s)
{
    switch (x) {
    case 1: return launch_kbx<K, B, 1>(a, s);
    case 2: return launch_kbx<K, B, 2>(a, s);
    case 4: return launch_kbx<K, B, 4>(a, s);
    case 8: return launch_kbx<K, B, 8>(a, s);
    case 16: return launch_kbx<K, B, 16>(a, s);
    default: return hipErrorInvalidValue;
    }
}

template <int K>
static hipError_t launch_k(int block, int x, const LaunchArgs &a, hipStream_t s)
{
    switch (block) {
    case 64: return launch_kb<K, 64>(x, a, s);
    case 128: return launch_kb<K, 128>(x, a, s);
    case 256: return launch_kb<K, 256>(x, a, s);
    case 512: return launch_kb<K, 512>(x, a, s);
    default: return hipErrorInvalidValue;
    }
}

template <int K>
static const void *fn_k(int block, int x)
{
    switch (block) {
    case 64: return fn_kb<K, 64>(x);
    case 128: return fn_kb<K, 128>(x);
    case 256: return fn_kb<K, 256>(x);
    case 512: return fn_kb<K, 512>(x);
    default: return nullptr;
    }
}

hipError_t launch_plr(int k, int block, int x, const LaunchArgs &a, hipStream_t s)
{
    switch (k) {
    case 2: return launch_k<2>(block, x, a, s);
    case 4: return launch_k<4>(block, x, a, s);
    case 8: return launch_k<8>(block, x, a, s);
    case 16: return launch_k<16>(block, x, a, s);
    default: return hipErrorInvalidValue;
    }
}

const void *kernel_plr(int k, int block, int x)
{
    switch (k) {
    case 2: return fn_k<2>(block, x);
    case 4: return fn_k<4>(block, x);
    case 8: return fn_k<8>(block, x);
    case 16: return fn_k<16>(block, x);
    default: return nullptr;
    }
}

}  // namespace iirx
