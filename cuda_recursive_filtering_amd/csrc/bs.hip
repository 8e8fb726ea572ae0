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
//      accumulated onto Y0 by the FP32 matrix core.
// MFMA columns are tiles, so every product feeds the next from registers and
// the scan's tile shifts are DPP row rotations.  The FP64 shape puts row
// l/16 + 4 j of its result on lane l, register j: states are kept in that
// order (lane group g, register j = state g + 4 j) by the operand packing,
// including Gam's row order in the FP16 product.
#include "bs_core.hpp"

namespace iirx {

// operands of the single-signal kernel: T, Gam and Phi^(2^j) from the
// workgroup's LDS copy, the post-lookback ones from global memory
template <int L>
struct BsLdsOps {
    static constexpr BsLayout Y = bs_layout(L);
    static constexpr int NT = bs_tidx(L / 16);
    const uint4 *lT;
    const uint4 *lG;
    const f4 *lP;
    const float *tab;
    int lane;
    __device__ h8 tmat(int hl, int mb, int q) const { return __builtin_bit_cast(h8, lT[(hl * NT + bs_tidx(mb) + q) * 64 + lane]); }
    __device__ h8 gam(int hl, int q) const { return __builtin_bit_cast(h8, lG[(hl * (L / 32) + q) * 64 + lane]); }
    __device__ f4 pow2(int j) const { return lP[j * 64 + lane]; }
    __device__ f4 phi() const { return reinterpret_cast<const f4 *>(tab + Y.kPhi)[lane]; }
    __device__ f4 pow16(int nb) const { return reinterpret_cast<const f4 *>(tab + bs_Q(Y, nb))[lane]; }
    __device__ f4 obs(int mb) const { return reinterpret_cast<const f4 *>(tab + bs_O(Y, mb))[lane]; }
};

template <int L, int G, int NSL>
__global__ __launch_bounds__(64 * bs_wg_waves(L, G))
__attribute__((amdgpu_waves_per_eu(bs_waves(L, G), bs_waves(L, G)))) void bs_kernel(LaunchArgs a)
{
    constexpr int MB = L / 16;  // output row blocks of T
    constexpr int KQ = L / 32;  // K slabs of 32 samples
    constexpr int NB = G / 16;  // column blocks of 16 tiles
    constexpr int CH = L * G;
    constexpr int NLD = NB * KQ * 2;  // 16-byte loads per lane and chunk
    constexpr int W = bs_wg_waves(L, G);
    constexpr int NT = bs_tidx(MB);   // nonzero blocks of T
    constexpr BsLayout Y = bs_layout(L);
    // shared by the workgroup: the operands consumed before the lookback.  A
    // VMEM table load issued after the next chunk's fill would wait for the
    // whole fill (vmcnt retires in issue order); LDS reads do not.
    __shared__ uint4 lT[2 * NT * 64];  // T hi/lo, nonzero blocks (mb, q <= bs_tq(mb))
    __shared__ uint4 lG[2 * KQ * 64];  // Gam hi/lo
    __shared__ f4 lP[6 * 64];          // Phi^(2^j), FP32 operand blocks
    // per wave: one chunk of X (load i of lane l at [64 i + l]) and the lookback slots
    __shared__ f4 xbuf_w[W][NLD * 64];
    __shared__ float lp_w[W][kBsSlots * 16];

    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, col = lane & 15, g = lane >> 4;
    const float *tab = a.coef;
    const cfloat *ctab = as_const(a.coef);
    const int sT = (int)ctab[Y.kScale + 0], sG = (int)ctab[Y.kScale + 1];
    const int depth = (int)ctab[Y.kDepth];
    f4 *xbuf = xbuf_w[wave];
    float *lp = lp_w[wave];

    for (int i = threadIdx.x; i < 2 * NT * 64; i += 64 * W) {
        const int t = (i >> 6) % NT, hl = (i >> 6) / NT;
        int mb = 0;
        while (bs_tidx(mb + 1) <= t) ++mb;
        lT[i] = reinterpret_cast<const uint4 *>(tab + bs_T(Y, hl, mb, t - bs_tidx(mb)))[i & 63];
    }
    for (int i = threadIdx.x; i < 2 * KQ * 64; i += 64 * W)
        lG[i] = reinterpret_cast<const uint4 *>(tab + bs_G(Y, (i >> 6) / KQ, (i >> 6) % KQ))[i & 63];
    for (int i = threadIdx.x; i < 6 * 64; i += 64 * W)
        lP[i] = reinterpret_cast<const f4 *>(tab + bs_P(Y, i >> 6))[i & 63];
    __syncthreads();

    // X as x32 B operands: tile 16 nb + col, K slab q, K rows 8 g + j =
    // samples 32 q + 4 g + j (a) and 32 q + 16 + 4 g + j - 4 (b).  The chunk
    // is fetched HBM -> LDS by global_load_lds (no registers held while it
    // is in flight); read_chunk waits for it.
    typedef BsRaw<NB, KQ> Raw;
    auto fetch_chunk = [&](uint32_t chunk) {
        const long base = (long)chunk * CH;
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
#pragma unroll
            for (int q = 0; q < KQ; ++q)
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const float *p = a.in + base + (long)(16 * nb + col) * L + 32 * q + 16 * h + 4 * g;
                    __builtin_amdgcn_global_load_lds((__attribute__((address_space(1))) void *)(p),
                                                     (__attribute__((address_space(3))) void *)(&xbuf[((nb * KQ + q) * 2 + h) * 64]),
                                                     16, 0, 0);
                }
    };
    // vmcnt counts loads and stores together and retires them in issue
    // order, so after the first chunk the fill has landed once at most the
    // MB * NB output stores issued after it are outstanding
    auto read_chunk = [&](Raw &r, bool first) {
        if (first)
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        else
            asm volatile("s_waitcnt vmcnt(%0)" ::"n"(MB * NB) : "memory");
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
#pragma unroll
            for (int q = 0; q < KQ; ++q) {
                r.a[nb][q] = xbuf[((nb * KQ + q) * 2 + 0) * 64 + lane];
                r.b[nb][q] = xbuf[((nb * KQ + q) * 2 + 1) * 64 + lane];
            }
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");  // reads done before the next fill
    };

    const BsLdsOps<L> ops{lT, lG, lP, tab, lane};
    auto process = [&](uint32_t chunk, uint32_t part_flag, uint32_t full_flag, const Raw &xr) {
        float *out = a.out + (long)chunk * CH;
        // a store covers 64 B of each of 16 tiles; the two halves of a 128-byte
        // line (row blocks mb, mb + 1) are issued back to back and through the
        // L2 (not nontemporal), so that they merge there: measured WRITE_SIZE
        // 1.06x the output bytes this way, 1.43x with mb outer and nontemporal stores
        auto store = [&](int nb, int mb, const f4 &y) {
            *reinterpret_cast<f4 *>(out + (long)(16 * nb + col) * L + 16 * mb + 4 * g) = y;
        };
        // streaming: a.zi is the balanced state before the signal, a.tail
        // receives the state entering the last chunk's last tile
        bs_chunk<L, G, NSL>(xr, chunk, part_flag, full_flag, a.lb, tab + Y.kMc, depth, lp, a.spin_limit, a.err, 0u,
                            sT, sG, ops, store, a.zi, chunk + 1u == a.num_chunks ? a.tail : nullptr);
    };

    const uint32_t GW = gridDim.x * W, gw = blockIdx.x * W + wave;  // this wave among the grid's
    if (a.persistent) {
        // chunks gw + i * GW; the next chunk's input streams into LDS while
        // this one is scanned, looked back and written
        const uint32_t epoch = epoch_load(a.epoch);
        epoch_arrive(a.arrive, epoch);
        const uint32_t pf = 2u * epoch + 1u, ff = pf + 1u;
        uint32_t c = gw;
        if (c < a.num_chunks) {
            // first chunk peeled: every entry into the loop below has this
            // chunk's output stores behind the last fill
            fetch_chunk(c);
            Raw cur;
            read_chunk(cur, true);
            if (c + GW < a.num_chunks) fetch_chunk(c + GW);
            process(c, pf, ff, cur);
        }
        for (c += GW; c < a.num_chunks; c += GW) {
            Raw cur;
            read_chunk(cur, false);
            if (c + GW < a.num_chunks) fetch_chunk(c + GW);
            process(c, pf, ff, cur);
        }
        // the owner of the last chunk advances the epoch after publishing it
        if ((a.num_chunks - 1u) % GW == gw) epoch_advance(a.epoch, a.arrive, epoch, GW, a.spin_limit, a.err);
    } else if (gw < a.num_chunks) {
        // one ticket per wave, exactly num_chunks per launch (grid = ceil(chunks / W))
        unsigned long long tk = 0;
        if (lane == 0) tk = atomicAdd(a.ticket, 1ull);
        const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)tk);
        const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(tk >> 32));
        tk = ((unsigned long long)hi << 32) | lo;
        const unsigned long long ep = tk / a.num_chunks;
        const uint32_t chunk = (uint32_t)(tk - ep * a.num_chunks);
        const uint32_t pf = (uint32_t)(2ull * ep + 1ull);
        Raw cur;
        fetch_chunk(chunk);
        read_chunk(cur, true);
        process(chunk, pf, pf + 1u, cur);
    }
}

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

template <int L, int G, int NSL>
static hipError_t launch_lgn(const LaunchArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL((bs_kernel<L, G, NSL>), dim3(a.grid), dim3(64 * bs_wg_waves(L, G)), 0, s, a);
    return hipGetLastError();
}

template <int L, int G>
static hipError_t launch_lg(int nsl, const LaunchArgs &a, hipStream_t s)
{
    switch (nsl) {
    case 1: return launch_lgn<L, G, 1>(a, s);
    case 2: return launch_lgn<L, G, 2>(a, s);
    case 3: return launch_lgn<L, G, 3>(a, s);
    case 4: return launch_lgn<L, G, 4>(a, s);
    default: return hipErrorInvalidValue;
    }
}

template <int L, int G>
static const void *fn_lg(int nsl)
{
    switch (nsl) {
    case 1: return (const void *)bs_kernel<L, G, 1>;
    case 2: return (const void *)bs_kernel<L, G, 2>;
    case 3: return (const void *)bs_kernel<L, G, 3>;
    case 4: return (const void *)bs_kernel<L, G, 4>;
    default: return nullptr;
    }
}

template <int L>
static hipError_t launch_l(int G, int nsl, const LaunchArgs &a, hipStream_t s)
{
    switch (G) {
    case 16: return launch_lg<L, 16>(nsl, a, s);
    case 32: return launch_lg<L, 32>(nsl, a, s);
    case 64: if constexpr (L < 128) return launch_lg<L, 64>(nsl, a, s); else return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
    }
}

template <int L>
static const void *fn_l(int G, int nsl)
{
    switch (G) {
    case 16: return fn_lg<L, 16>(nsl);
    case 32: return fn_lg<L, 32>(nsl);
    case 64: if constexpr (L < 128) return fn_lg<L, 64>(nsl); else return nullptr;
    default: return nullptr;
    }
}

hipError_t launch_bs(int L, int G, int ns, const LaunchArgs &a, hipStream_t s)
{
    switch (L) {
    case 32: return launch_l<32>(G, bs_slabs(ns), a, s);
    case 64: return launch_l<64>(G, bs_slabs(ns), a, s);
    case 128: return launch_l<128>(G, bs_slabs(ns), a, s);
    default: return hipErrorInvalidValue;
    }
}

const void *kernel_bs(int L, int G, int ns)
{
    switch (L) {
    case 32: return fn_l<32>(G, bs_slabs(ns));
    case 64: return fn_l<64>(G, bs_slabs(ns));
    case 128: return fn_l<128>(G, bs_slabs(ns));
    default: return nullptr;
    }
}

}  // namespace iirx
