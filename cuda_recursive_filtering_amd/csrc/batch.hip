// batch.hip -- ragged batch: many independent channels of uneven length, each
// with its own order and coefficients, in one persistent launch.
//
// No counterpart in the reference (one channel per binary there).  Every
// channel is cut into chunks of 64 lanes x N samples, numbered channel after
// channel; each chunk runs the PH algorithm of ph.hip (include/ph_kernels.cuh
// PH_64, 501-820) on its channel's tables, with the lookback chain stopping at
// the channel's first chunk (zero initial state per channel).  A channel's
// last chunk is partial: samples past the channel end are read as zero (the
// filter is causal, so they cannot reach earlier outputs) and not written.
// Channel offsets are arbitrary (CSR packing): full chunks move with 16-byte
// accesses shifted by the chunk's offset mod 4 (stage_in_any below), a
// channel's last chunk with 4-byte coalesced loads and stores (as do full
// chunks of buffers that are not 16-byte aligned).
// A channel may instead run the block-state form (bs_core.hpp): tiles of 64
// samples, G tiles per chunk, its tables read from global memory.  The two
// forms are separate kernels, each with its own register budget, launched one
// after the other over their own ranges of the chunk list (PH chunks first).
#include "bs_core.hpp"

namespace iirx {

typedef const __attribute__((address_space(4))) BatchChannel cchannel;
typedef const __attribute__((address_space(4))) uint32_t cu32;

// waves per SIMD the register budgets are set for: PH at N = 32 four waves
// (a 40-byte/lane spill) measured faster than three without one (ragged
// 4096-channel workload: 457 vs 446 GS/s)
template <int N>
constexpr int batch_waves() { return N == 16 ? 5 : N == 32 ? 4 : 1; }
template <int G>
constexpr int batch_bs_waves() { return G == 16 ? 5 : G == 32 ? 4 : 2; }

// Full chunk of CH floats at any float offset `base`, moved with 16-byte
// accesses (`in`/`out` themselves 16-byte aligned): float4 k (k < CH/4, over
// the wave's 64 threads) covers [base - m + 4k, +4), m = base % 4, so the
// chunk's element e = 4k - m + t lands at LDS index map(e).  Loads: the m
// elements before the chunk that float4 0 also reads lie inside the buffer
// and are dropped; the m elements after the last whole float4 come in scalar.
// Stores: every float4 but the first lies wholly inside the chunk; the first
// (when m != 0) and the last m elements are stored one by one, so no
// neighbour's samples are written.
template <int CH, class Map>
__device__ __forceinline__ void stage_in_any(const float *__restrict__ in, long base, float *lds, const Map &map)
{
    constexpr int V = CH / 4, PER = V / 64;
    static_assert(V % 64 == 0, "chunk must split evenly over the wave");
    const int tid = threadIdx.x & 63, m = (int)(base & 3);
    const f4 *s4 = reinterpret_cast<const f4 *>(in + (base - m));
    f4 v[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) v[i] = __builtin_nontemporal_load(s4 + i * 64 + tid);
    const float tail = (tid < m) ? in[base + CH - m + tid] : 0.0f;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int e0 = 4 * (i * 64 + tid) - m;
#pragma unroll
        for (int t = 0; t < 4; ++t)
            if (e0 + t >= 0) lds[map(e0 + t)] = v[i][t];
    }
    if (tid < m) lds[map(CH - m + tid)] = tail;
}

// The float4 stores stay nontemporal here: through the L2 (which helps the
// single-signal kernels' aligned chunks) the ragged 4096-channel workload ran
// 434-435 GS/s against 457-459, alternating builds on one box.
template <int CH, class Map>
__device__ __forceinline__ void stage_out_any(const float *lds, float *__restrict__ out, long base, const Map &map)
{
    constexpr int V = CH / 4, PER = V / 64;
    const int tid = threadIdx.x & 63, m = (int)(base & 3);
    f4 *d4 = reinterpret_cast<f4 *>(out + (base - m));
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int k = i * 64 + tid, e0 = 4 * k - m;
        if (k == 0 && m != 0) {
            for (int t = m; t < 4; ++t) out[base + e0 + t] = lds[map(e0 + t)];
        } else {
            f4 w;
#pragma unroll
            for (int t = 0; t < 4; ++t) w[t] = lds[map(e0 + t)];
            __builtin_nontemporal_store(w, d4 + k);
        }
    }
    if (tid < m) out[base + CH - m + tid] = lds[map(CH - m + tid)];
}

// block-state operands of one channel, from its tables in global memory
struct BsGlobalOps {
    static constexpr BsLayout Y = bs_layout(64);
    const float *tab;
    int lane;
    __device__ h8 tmat(int hl, int mb, int q) const
    {
        return __builtin_bit_cast(h8, reinterpret_cast<const uint4 *>(tab + bs_T(Y, hl, mb, q))[lane]);
    }
    __device__ h8 gam(int hl, int q) const { return __builtin_bit_cast(h8, reinterpret_cast<const uint4 *>(tab + bs_G(Y, hl, q))[lane]); }
    __device__ f4 pow2(int j) const { return reinterpret_cast<const f4 *>(tab + bs_P(Y, j))[lane]; }
    __device__ f4 phi() const { return reinterpret_cast<const f4 *>(tab + Y.kPhi)[lane]; }
    __device__ f4 pow16(int nb) const { return reinterpret_cast<const f4 *>(tab + bs_Q(Y, nb))[lane]; }
    __device__ f4 obs(int mb) const { return reinterpret_cast<const f4 *>(tab + bs_O(Y, mb))[lane]; }
};

// LDS of the block-state path: tile t (64 samples) at 68 t, so that the 16
// tiles a row of lanes reads with one ds_read_b128 fall in distinct banks
constexpr int kBsTilePitch = 68;

// one block-state chunk of a batch channel: G = N tiles of L = 64 samples,
// four FP32 K slabs (any order up to 16).  The chunk is staged in and out
// through LDS with coalesced 4-byte accesses (channel offsets are arbitrary);
// the last chunk reads zeros past the channel end and stores only r samples.
template <int N>
__device__ __forceinline__ void batch_bs_chunk(const BatchArgs &a, uint32_t chunk, uint32_t part_flag,
                                               uint32_t full_flag, const float *coef, uint32_t first, long base, int r,
                                               float *lds, float *lp)
{
    constexpr int L = 64, G = N, NB = G / 16, KQ = 2, CH = L * G;
    constexpr BsLayout Y = bs_layout(L);
    const int lane = threadIdx.x & 63, col = lane & 15, g = lane >> 4;
    const cfloat *ctab = as_const(coef);

    const auto tile_map = [](int e) { return (e >> 6) * kBsTilePitch + (e & 63); };
    __syncthreads();  // LDS of the previous chunk fully consumed
    {
        const float *src = a.in + base + lane;
        if (r == CH && a.vec16) {
            stage_in_any<CH>(a.in, base, lds, tile_map);
        } else if (r == CH) {
            float v[G];
#pragma unroll
            for (int t = 0; t < G; ++t) v[t] = src[t * L];
#pragma unroll
            for (int t = 0; t < G; ++t) lds[t * kBsTilePitch + lane] = v[t];
        } else {
            for (int t = 0; t < G; ++t) {
                const int e = t * L + lane;
                const float v = a.in[base + (e < r ? e : r - 1)];
                lds[t * kBsTilePitch + lane] = (e < r) ? v : 0.0f;
            }
        }
    }
    __syncthreads();
    // lane (g, col), tile 16 nb + col: samples 32 q + 4 g + (0..3) and 32 q + 16 + 4 g + (0..3)
    BsRaw<NB, KQ> xr;
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
        for (int q = 0; q < KQ; ++q) {
            const f4 *t = reinterpret_cast<const f4 *>(lds + (16 * nb + col) * kBsTilePitch + 32 * q + 4 * g);
            xr.a[nb][q] = t[0];
            xr.b[nb][q] = t[4];
        }
    const BsGlobalOps ops{coef, lane};
    // the chunk's inputs are in registers before the first output block
    // overwrites LDS (a wave's LDS accesses complete in order)
    auto store = [&](int nb, int mb, const f4 &y) {
        *reinterpret_cast<f4 *>(lds + (16 * nb + col) * kBsTilePitch + 16 * mb + 4 * g) = y;
    };
    bs_chunk<L, G, 4>(xr, chunk, part_flag, full_flag, a.lb, coef + Y.kMc, (int)ctab[Y.kDepth], lp, a.spin_limit,
                      a.err, first, (int)ctab[Y.kScale + 0], (int)ctab[Y.kScale + 1],
                      (int)ctab[Y.kScale + kBsScaleFp32Level], (int)ctab[Y.kScale + kBsScaleLevels], ops, store);
    __syncthreads();
    float *dst = a.out + base + lane;
    if (r == CH && a.vec16) {
        stage_out_any<CH>(lds, a.out, base, tile_map);
    } else if (r == CH) {
#pragma unroll
        for (int t = 0; t < G; ++t) dst[t * L] = lds[t * kBsTilePitch + lane];
    } else {
        for (int t = 0; t < G; ++t)
            if (t * L + lane < r) dst[t * L] = lds[t * kBsTilePitch + lane];
    }
}

// block-state channels: chunks chunk_base .. chunk_base + num_chunks - 1
template <int G>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(batch_bs_waves<G>(), batch_bs_waves<G>())))
void batch_bs_kernel(BatchArgs a)
{
    constexpr int CH = 64 * G;
    __shared__ __align__(16) float lds[kBsTilePitch * G];
    __shared__ float lp[kBsSlots * 16];
    const cchannel *chan = (const cchannel *)a.chan;
    const cu32 *chunk_chan = (const cu32 *)a.chunk_chan;

    for_each_chunk(a.ticket, a.epoch, a.arrive, a.num_chunks, a.persistent, a.spin_limit, a.err,
                   [&](uint32_t local, uint32_t part_flag, uint32_t full_flag) {
        const uint32_t chunk = a.chunk_base + local;
        const uint32_t ci = chunk_chan[chunk];
        const uint32_t first = chan[ci].first_chunk;
        const long j = (long)(chunk - first);
        const long base = chan[ci].offset + j * CH;
        const long rem = chan[ci].length - j * CH;
        const int r = rem < CH ? (int)rem : CH;
        batch_bs_chunk<G>(a, chunk, part_flag, full_flag, a.coef + chan[ci].coef_off, first, base, r, lds, lp);
    });
}

// PH channels: chunks chunk_base .. chunk_base + num_chunks - 1
template <int N>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(batch_waves<N>(), batch_waves<N>())))
void batch_ph_kernel(BatchArgs a)
{
    constexpr int L = 64;
    constexpr int CH = L * N;
    constexpr PhLayout Y = ph_layout(L, N);
    __shared__ float lds[L * (N + 1)];

    const int lane = threadIdx.x;
    const cchannel *chan = (const cchannel *)a.chan;
    const cu32 *chunk_chan = (const cu32 *)a.chunk_chan;

    for_each_chunk(a.ticket, a.epoch, a.arrive, a.num_chunks, a.persistent, a.spin_limit, a.err,
                   [&](uint32_t local, uint32_t part_flag, uint32_t full_flag) {
    // channel of this chunk (uniform: scalar loads)
    const uint32_t chunk = a.chunk_base + local;
    const uint32_t ci = chunk_chan[chunk];
    const long ch_off = chan[ci].offset, ch_len = chan[ci].length;
    const int ns = chan[ci].ns;
    const uint32_t first = chan[ci].first_chunk;
    const float *coef = a.coef + chan[ci].coef_off;
    const long j = (long)(chunk - first);
    const long base = ch_off + j * CH;
    const long rem = ch_len - j * CH;
    const int r = rem < CH ? (int)rem : CH;  // valid samples in this chunk

    // stage in: lane t owns chunk elements [t N, t N + N).  Full chunks of
    // 16-byte aligned buffers move with 16-byte accesses; otherwise element
    // e = 64 i + lane is read with coalesced 4-byte loads, from one base
    // address with immediate offsets in a full chunk; the channel's last chunk
    // clamps the index to the channel's last sample (always in bounds) and
    // stores zero past it, one load at a time.
    const auto row_map = [](int e) { return (e / N) * (N + 1) + e % N; };
    __syncthreads();  // LDS of the previous chunk fully consumed
    {
        const float *src = a.in + base + lane;
        if (r == CH && a.vec16) {
            stage_in_any<CH>(a.in, base, lds, row_map);
        } else if (r == CH) {
            float v[N];
#pragma unroll
            for (int i = 0; i < N; ++i) v[i] = src[i * L];
#pragma unroll
            for (int i = 0; i < N; ++i) {
                const int e = i * L + lane;
                lds[(e / N) * (N + 1) + e % N] = v[i];
            }
        } else {
            for (int i = 0; i < N; ++i) {
                const int e = i * L + lane;
                const float v = a.in[base + (e < r ? e : r - 1)];
                lds[(e / N) * (N + 1) + e % N] = (e < r) ? v : 0.0f;
            }
        }
    }
    __syncthreads();

    float x[N];
#pragma unroll
    for (int n = 0; n < N; ++n) x[n] = lds[lane * (N + 1) + n];

    // inputs of the first section just before this lane's samples
    float xi2, xi1;
    if (lane > 0) {
        xi2 = lds[(lane - 1) * (N + 1) + N - 2];
        xi1 = lds[(lane - 1) * (N + 1) + N - 1];
    } else if (j > 0) {
        xi2 = a.in[base - 2];
        xi1 = a.in[base - 1];
    } else {
        xi2 = 0.0f;
        xi1 = 0.0f;
    }

    for (int sec = 0; sec < ns; ++sec) {
        const float *cv = coef + (size_t)sec * a.coef_stride;
        const cfloat *c = as_const(cv);
        const float b1 = c[Y.kB + 0], b2 = c[Y.kB + 1], a1 = c[Y.kB + 2], a2 = c[Y.kB + 3];

        // zero-state biquad over the lane's samples
        {
            float x2 = xi2, x1 = xi1, y2 = 0.0f, y1 = 0.0f;
#pragma unroll
            for (int n = 0; n < N; ++n) {
                const float xc = x[n];
                float yc = fmaf(b1, x1, fmaf(b2, x2, xc));
                if (n == 1) yc = fmaf(a1, y1, yc);
                else if (n > 1) yc = fmaf(a1, y1, fmaf(a2, y2, yc));
                x[n] = yc;
                x2 = x1;
                x1 = xc;
                y2 = y1;
                y1 = yc;
            }
        }

        // lane scan of the end states (as ph.hip: DPP row scan, then the
        // rows' carries from lanes 15, 31, 47)
        float s2 = x[N - 2], s1 = x[N - 1];
        ph_lane_scan(c + Y.kCpow, cv + Y.kCfin, lane, s2, s1);

        // chunk partial, lookback within the channel, chunk full
        const float m0 = c[Y.kM + 0], m1 = c[Y.kM + 1], m2 = c[Y.kM + 2], m3 = c[Y.kM + 3];
        const float p2 = readlane_f(s2, L - 1), p1 = readlane_f(s1, L - 1);
        uint64_t *rec = rec2(a.lb, chunk, sec, kBatchMaxSections);
        if (lane == 0) publish2(rec, false, part_flag, p2, p1);
        const int depth = (int)c[Y.kDepth];
        const Carry2 X = lookback2(a.lb, chunk, sec, kBatchMaxSections, part_flag, full_flag, m0, m1, m2, m3, depth,
                                   a.spin_limit, a.err, first);
        if (lane == 0) {
            const float f2 = fmaf(m1, X.y1, fmaf(m0, X.y2, p2));
            const float f1 = fmaf(m3, X.y1, fmaf(m2, X.y2, p1));
            publish2(rec, true, full_flag, f2, f1);
        }

        // per-lane end state with the chunk carry injected, C^(lane+1) X
        const float4 cf = *reinterpret_cast<const float4 *>(cv + Y.kCfin + 4 * lane);
        const float e2 = fmaf(cf.y, X.y1, fmaf(cf.x, X.y2, s2));
        const float e1 = fmaf(cf.w, X.y1, fmaf(cf.z, X.y2, s1));
        float u2 = lane_up(e2, 1), u1 = lane_up(e1, 1);
        if (lane == 0) {
            u2 = X.y2;
            u1 = X.y1;
        }

        // back-correction by the zero-input response to the lane's incoming state
#pragma unroll
        for (int n = 0; n < N - 2; ++n) x[n] = fmaf(c[Y.kH1 + n], u1, fmaf(c[Y.kH2 + n], u2, x[n]));
        x[N - 2] = e2;
        x[N - 1] = e1;
        xi2 = u2;
        xi1 = u1;
    }

    __syncthreads();
#pragma unroll
    for (int n = 0; n < N; ++n) lds[lane * (N + 1) + n] = x[n];
    __syncthreads();
    float *dst = a.out + base;
    if (r == CH && a.vec16) {
        stage_out_any<CH>(lds, a.out, base, row_map);
    } else if (r == CH) {
#pragma unroll
        for (int i = 0; i < N; ++i) {
            const int e = i * L + lane;
            dst[lane + i * L] = lds[(e / N) * (N + 1) + e % N];
        }
    } else {
        // the channel's last chunk
        for (int i = 0; i < N; ++i) {
            const int e = i * L + lane;
            if (e < r) dst[e] = lds[(e / N) * (N + 1) + e % N];
        }
    }
    });
}

template <class K>
static hipError_t launch_k(K kernel, const BatchArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(kernel, dim3(a.grid), dim3(64), 0, s, a);
    return hipGetLastError();
}

// algo kBatchPH: n = samples per lane (16, 32, 64); kBatchBS: n = tiles per chunk
const void *kernel_batch(int algo, int n)
{
    if (algo == kBatchPH) {
        switch (n) {
        case 16: return (const void *)batch_ph_kernel<16>;
        case 32: return (const void *)batch_ph_kernel<32>;
        case 64: return (const void *)batch_ph_kernel<64>;
        }
    } else if (algo == kBatchBS) {
        switch (n) {
        case 16: return (const void *)batch_bs_kernel<16>;
        case 32: return (const void *)batch_bs_kernel<32>;
        case 64: return (const void *)batch_bs_kernel<64>;
        }
    }
    return nullptr;
}

hipError_t launch_batch(int algo, int n, const BatchArgs &a, hipStream_t s)
{
    if (algo == kBatchPH) {
        switch (n) {
        case 16: return launch_k(batch_ph_kernel<16>, a, s);
        case 32: return launch_k(batch_ph_kernel<32>, a, s);
        case 64: return launch_k(batch_ph_kernel<64>, a, s);
        }
    } else if (algo == kBatchBS) {
        switch (n) {
        case 16: return launch_k(batch_bs_kernel<16>, a, s);
        case 32: return launch_k(batch_bs_kernel<32>, a, s);
        case 64: return launch_k(batch_bs_kernel<64>, a, s);
        }
    }
    return hipErrorInvalidValue;
}

}  // namespace iirx
