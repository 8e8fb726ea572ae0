// bs.hip -- block-state form of the cascade on the MI355X matrix cores.
//
// No counterpart in the reference; it computes the same filter as
// ph/stcr/dtcr (the whole cascade, zero initial state) in a form whose cost
// grows with the order only through the state dimension of small products.
// The cascade is one state-space system of dimension p = 2 N_SECTIONS,
// embedded in 16 rows.  A chunk of G (16, 32 or 64) tiles of L samples is one
// wave64:
//   1. the zero-state parts Y0 = T X (T: L x L lower-triangular Toeplitz of
//      the impulse response) and V = Gam X (Gam: 16 x L), as
//      v_mfma_f32_16x16x32_f16 products of FP16 hi/lo operands
//      (hi*hi + hi*lo + lo*hi, FP32 accumulation); T and Gam are scaled by
//      a power of two on the host, X by one per chunk, so that FP16 holds
//      them; blocks of T above the diagonal are skipped;
//   2. the tile end states Z_k = sum_{j<=k} Phi^(k-j) V_j by a log2(G)-level
//      Kogge-Stone scan over the G tiles, each level Phi^(2^j) * shift(Z)
//      on v_mfma_f32_16x16x4f32 (plain FP32 operands: no splitting, no data
//      scaling), over ceil(p/4) K slabs only;
//   3. the chunk's incoming state X_in from the lookback (16-word records,
//      chunk transition Phi^G, bounded depth), injected as Phi^(k+1) X_in:
//      Phi X_in in tile 0, a 4-level scan inside one 16-tile block and
//      Phi^(16 nb) products for the other blocks;
//   4. Y = Y0 + O S, S_k = state entering tile k, accumulated onto Y0 by the
//      FP32 matrix core.
// Both MFMA shapes put D[4 (l/16) + j][l%16] on lane l, register j: tiles are
// MFMA columns, states rows 4 g + j, so every product feeds the next from
// registers and the scan's tile shifts are DPP row rotations.
#include "device_common.hpp"
#include "kernels.hpp"

namespace iirx {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ f4 mfma_h(h8 a, h8 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
__device__ __forceinline__ f4 mfma_f(float a, float b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

__device__ __forceinline__ f4 load_f4(const float *tab, int word_off)
{
    return reinterpret_cast<const f4 *>(tab + word_off)[threadIdx.x & 63];
}

// DPP moves (v_mov_dpp, no LDS round trip): row_ror:D rotates every 16-lane
// row, lane l <- lane (l - D) mod 16 of its row; row_bcast15/31 pass lane 15 /
// 31 on to the next row / upper half
template <int D>
__device__ __forceinline__ float row_ror(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x120 + D, 0xF, 0xF, false));
}

__device__ __forceinline__ float wave_max(float v)
{
    v = fmaxf(v, row_ror<1>(v));
    v = fmaxf(v, row_ror<2>(v));
    v = fmaxf(v, row_ror<4>(v));
    v = fmaxf(v, row_ror<8>(v));
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x142, 0xF, 0xF, false)));
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x143, 0xF, 0xF, false)));
    return readlane_f(v, 63);
}

// power-of-two exponent s with max * 2^s < 2^14
__device__ __forceinline__ int scale_exp(float mx)
{
    return (mx > 0.0f) ? 14 - __builtin_amdgcn_frexp_expf(mx) : 0;
}

// hi/lo FP16 split of (a, b) * 2^s, the 8 K rows of one x32 B operand
__device__ __forceinline__ void split8(const f4 &a, const f4 &b, int s, h8 &hi, h8 &lo)
{
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const float x = ldexpf(j < 4 ? a[j & 3] : b[j & 3], s);
        const _Float16 h = (_Float16)x;
        hi[j] = h;
        lo[j] = (_Float16)(x - (float)h);
    }
}

// acc + M S over the first NSL K slabs; M is this lane's float4 of an FP32 block
template <int NSL>
__device__ __forceinline__ f4 apply(const f4 &M, const f4 &S, f4 acc)
{
#pragma unroll
    for (int m = 0; m < NSL; ++m) acc = mfma_f(M[m], S[m], acc);
    return acc;
}

// S(c, k) = Z(c, k - D) (0 for k < D) over NB column blocks of 16 tiles;
// D < 16 rotates inside each 16-lane row and takes the wrapped lanes from the
// previous block.  Only the NSL live state registers are moved.
template <int D, int NSL, int NB>
__device__ __forceinline__ void shift_tiles(const f4 (&Z)[NB], f4 (&S)[NB])
{
    if constexpr (D >= 16) {
        constexpr int dn = D >> 4;
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) S[nb] = (nb >= dn) ? Z[nb >= dn ? nb - dn : 0] : f4{0.f, 0.f, 0.f, 0.f};
    } else {
        const bool wrap = (threadIdx.x & 15) < D;
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (j < NSL) {
                    const float r0 = row_ror<D>(Z[nb][j]);
                    const float r1 = (nb > 0) ? row_ror<D>(Z[nb > 0 ? nb - 1 : 0][j]) : 0.0f;
                    S[nb][j] = wrap ? r1 : r0;
                } else {
                    S[nb][j] = 0.0f;
                }
            }
        }
    }
}

template <int G>
constexpr int bs_levels() { return G == 16 ? 4 : G == 32 ? 5 : 6; }

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
    constexpr int kSlots = 16;
    // shared by the workgroup: the operands consumed before the lookback.  A
    // VMEM table load issued after the next chunk's fill would wait for the
    // whole fill (vmcnt retires in issue order); LDS reads do not.
    __shared__ uint4 lT[2 * NT * 64];  // T hi/lo, nonzero blocks (mb, q <= bs_tq(mb))
    __shared__ uint4 lG[2 * KQ * 64];  // Gam hi/lo
    __shared__ f4 lP[6 * 64];          // Phi^(2^j), FP32 operand blocks
    // per wave: one chunk of X (load i of lane l at [64 i + l]) and the lookback slots
    __shared__ f4 xbuf_w[W][NLD * 64];
    __shared__ float lp_w[W][kSlots * 16];

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
    auto opT = [&](int hl, int mb, int q) { return __builtin_bit_cast(h8, lT[(hl * NT + bs_tidx(mb) + q) * 64 + lane]); };
    auto opG = [&](int hl, int q) { return __builtin_bit_cast(h8, lG[(hl * KQ + q) * 64 + lane]); };
    auto opP = [&](int j) { return lP[j * 64 + lane]; };

    // X as x32 B operands: tile 16 nb + col, K slab q, K rows 8 g + j =
    // samples 32 q + 4 g + j (a) and 32 q + 16 + 4 g + j - 4 (b).  The chunk
    // is fetched HBM -> LDS by global_load_lds (no registers held while it
    // is in flight); read_chunk waits for it.
    struct Raw {
        f4 a[NB][KQ], b[NB][KQ];
    };
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

    auto process = [&](uint32_t chunk, uint32_t part_flag, uint32_t full_flag, const Raw &xr) {
    const long base = (long)chunk * CH;

    // 1. per-chunk scale and FP16 hi/lo split of X
    h8 xh[NB][KQ], xl[NB][KQ];
    int eX;
    {
        float mx = 0.0f;
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
#pragma unroll
            for (int q = 0; q < KQ; ++q)
#pragma unroll
                for (int j = 0; j < 4; ++j) mx = fmaxf(mx, fmaxf(fabsf(xr.a[nb][q][j]), fabsf(xr.b[nb][q][j])));
        eX = scale_exp(wave_max(mx));
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
#pragma unroll
            for (int q = 0; q < KQ; ++q) split8(xr.a[nb][q], xr.b[nb][q], eX, xh[nb][q], xl[nb][q]);
    }

    // 2. V = Gam X (zero chunk-incoming state); Y0 = T X follows the partial
    f4 Z[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
        f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int q = 0; q < KQ; ++q) {
            const h8 gh = opG(0, q), gl = opG(1, q);
            acc = mfma_h(gl, xh[nb][q], acc);
            acc = mfma_h(gh, xl[nb][q], acc);
            acc = mfma_h(gh, xh[nb][q], acc);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) Z[nb][j] = ldexpf(acc[j], -(sG + eX));
    }
    // 3. tile end states, log2(G)-level scan (levels D >= 16 skip the
    //    all-zero blocks)
    static_for<0, bs_levels<G>(), 1>([&](auto LV) {
        constexpr int D = 1 << decltype(LV)::value;
        f4 S[NB];
        shift_tiles<D, NSL, NB>(Z, S);
        const f4 P = opP(decltype(LV)::value);
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
            if (D < 16 || nb >= (D >> 4)) Z[nb] = apply<NSL>(P, S[nb], Z[nb]);
    });

    // 4. chunk partial = Z(:, G-1): lanes 15, 31, 47, 63 hold components 4 g + j
    uint64_t *rec = a.lb + (size_t)chunk * 32;
    if (col == 15) {
#pragma unroll
        for (int j = 0; j < 4; ++j) rec_store(rec + 4 * g + j, part_flag, Z[NB - 1][j]);
    }
    // the predecessor's record is in flight while Y0 = T X runs on the matrix cores
    const uint64_t poll0 = lookback_first_poll<16>(a.lb, 32, chunk);
    f4 Y0[MB][NB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) {
        const int qmax = bs_tq(mb);  // T(16 mb + r, 32 q + c) = 0 beyond
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int q = 0; q < KQ; ++q) {
                if (q > qmax) break;
                const h8 th = opT(0, mb, q), tl = opT(1, mb, q);
                acc = mfma_h(tl, xh[nb][q], acc);
                acc = mfma_h(th, xl[nb][q], acc);
                acc = mfma_h(th, xh[nb][q], acc);
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) Y0[mb][nb][j] = ldexpf(acc[j], -(sT + eX));
        }
    }

    const float Xj = lookbackK<16, kSlots>(a.lb, 32, chunk, part_flag, full_flag, tab + Y.kMc, depth, lp,
                                           a.spin_limit, a.err, &poll0);
    // the output stage's O blocks, fetched before any store of this chunk:
    // vmcnt retires in issue order, so a load issued behind stores waits for them
    f4 Ob[MB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) Ob[mb] = load_f4(tab, bs_O(Y, mb));
    asm volatile("" ::: "memory");
    // X_in in B layout on column 0 (component 4 g + m lives on lane 4 g + m)
    f4 xin;
#pragma unroll
    for (int j = 0; j < 4; ++j) xin[j] = (j < NSL) ? __shfl(Xj, 4 * g + j, 64) : 0.0f;

    // injection W_k = Phi^(k+1) X_in: Phi X_in in column 0 of one block, a
    // 4-level scan inside the block, then Phi^(16 nb) for blocks 1..NB-1
    f4 U[1];
    {
        f4 b0;
#pragma unroll
        for (int j = 0; j < 4; ++j) b0[j] = (col == 0) ? xin[j] : 0.0f;
        U[0] = apply<NSL>(load_f4(tab, Y.kPhi), b0, f4{0.f, 0.f, 0.f, 0.f});
    }
    static_for<0, 4, 1>([&](auto LV) {
        constexpr int D = 1 << decltype(LV)::value;
        f4 S[1];
        shift_tiles<D, NSL, 1>(U, S);
        U[0] = apply<NSL>(opP(decltype(LV)::value), S[0], U[0]);
    });
#pragma unroll
    for (int nb = 1; nb < NB; ++nb) Z[nb] = apply<NSL>(load_f4(tab, bs_Q(Y, nb)), U[0], Z[nb]);
#pragma unroll
    for (int j = 0; j < 4; ++j) Z[0][j] += U[0][j];

    // chunk full = Z(:, G-1)
    if (col == 15) {
#pragma unroll
        for (int j = 0; j < 4; ++j) rec_store(rec + 16 + 4 * g + j, full_flag, Z[NB - 1][j]);
    }

    // 5. S_k = state entering tile k (X_in for tile 0); Y = Y0 + O S
    f4 S[NB];
    shift_tiles<1, NSL, NB>(Z, S);
    if (col == 0) S[0] = xin;
    // a store covers 64 B of each of 16 tiles; the two halves of a 128-byte
    // line (row blocks mb, mb + 1) are issued back to back and through the L2
    // (not nontemporal), so that they merge there: measured WRITE_SIZE 1.06x
    // the output bytes this way, 1.43x with mb outer and nontemporal stores
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) {
            const f4 y = apply<NSL>(Ob[mb], S[nb], Y0[mb][nb]);
            *reinterpret_cast<f4 *>(a.out + base + (long)(16 * nb + col) * L + 16 * mb + 4 * g) = y;
        }
    }
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
    case 64: return launch_lg<L, 64>(nsl, a, s);
    default: return hipErrorInvalidValue;
    }
}

template <int L>
static const void *fn_l(int G, int nsl)
{
    switch (G) {
    case 16: return fn_lg<L, 16>(nsl);
    case 32: return fn_lg<L, 32>(nsl);
    case 64: return fn_lg<L, 64>(nsl);
    default: return nullptr;
    }
}

hipError_t launch_bs(int L, int G, int ns, const LaunchArgs &a, hipStream_t s)
{
    switch (L) {
    case 32: return launch_l<32>(G, bs_slabs(ns), a, s);
    case 64: return launch_l<64>(G, bs_slabs(ns), a, s);
    default: return hipErrorInvalidValue;
    }
}

const void *kernel_bs(int L, int G, int ns)
{
    switch (L) {
    case 32: return fn_l<32>(G, bs_slabs(ns));
    case 64: return fn_l<64>(G, bs_slabs(ns));
    default: return nullptr;
    }
}

}  // namespace iirx
