// bs_core.hpp -- one chunk of the block-state form (used by bs.hip for one
// long signal and by batch.hip for the block-state channels of a ragged
// batch).  The algorithm is described at the top of bs.hip.
//
// A chunk is G tiles of L samples held by one wave64.  `xr` is the chunk's
// input in the x32 B-operand layout: lane (g, col), tile 16 nb + col, K slab
// q: a[nb][q] = samples 32 q + 4 g + (0..3), b[nb][q] = samples
// 32 q + 16 + 4 g + (0..3).  `ops` supplies the operand blocks (tmat(hl, mb,
// q) of T, gam(hl, q) of Gam, pow2(j) = Phi^(2^j), phi() = Phi, pow16(nb) =
// Phi^(16 nb), obs(mb) of O) and `store(nb, mb, y)` writes
// output row block mb of tile block nb (lane (g, col): samples 16 mb + 4 g +
// (0..3) of tile 16 nb + col).  Lookback records are 32 words per chunk; the
// chain starts at chunk `first`.
#pragma once
#include "device_common.hpp"
#include "kernels.hpp"

namespace iirx {

constexpr int kBsSlots = 16;  // lookback partials folded at most (LDS: 16 x 16 floats)

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f4 mfma_h(h8 a, h8 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
__device__ __forceinline__ f4 mfma_b(bf8 a, bf8 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
__device__ __forceinline__ f4 mfma_f(float a, float b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
// FP64 matrix core: lane l, register r of the result holds row l/16 + 4 r
// (the FP32 shapes hold row 4 (l/16) + r)
__device__ __forceinline__ d4 mfma_d(double a, double b, d4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }

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

// BF16 hi/mid/lo split of (a, b) * 2^s, the 8 K rows of one x32 B operand:
// each component the remainder so far rounded to nearest
__device__ __forceinline__ void split8_bf(const f4 &a, const f4 &b, int s, bf8 &hi, bf8 &mid, bf8 &lo)
{
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const float x = ldexpf(j < 4 ? a[j & 3] : b[j & 3], s);
        const __bf16 h = (__bf16)x;
        const float r1 = x - (float)h;
        const __bf16 m = (__bf16)r1;
        const float r2 = r1 - (float)m;
        hi[j] = h;
        mid[j] = m;
        lo[j] = (__bf16)r2;
    }
}

// acc + A X over the split: FP16 hi/lo (lo hi + hi lo + hi hi) or BF16
// hi/mid/lo (the six products above 2^-24 of hi hi, smallest first), A's
// planes from `a(hl)`, X the split of the raw samples
template <int SPLIT, class AOp>
__device__ __forceinline__ f4 split_product(const AOp &a, const f4 &xa, const f4 &xb, int eX, f4 acc)
{
    if constexpr (SPLIT == kBsSplitBF16) {
        bf8 xh, xm, xl;
        split8_bf(xa, xb, eX, xh, xm, xl);
        const bf8 ah = __builtin_bit_cast(bf8, a(0)), am = __builtin_bit_cast(bf8, a(1)), al = __builtin_bit_cast(bf8, a(2));
        acc = mfma_b(al, xh, acc);
        acc = mfma_b(ah, xl, acc);
        acc = mfma_b(am, xm, acc);
        acc = mfma_b(am, xh, acc);
        acc = mfma_b(ah, xm, acc);
        acc = mfma_b(ah, xh, acc);
    }
    return acc;
}

typedef __bf16 bf4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f4 mfma_b16(bf4 a, bf4 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }

// BF16 hi/mid/lo split of the first NSL values of v (the rest zero): each
// component the remainder so far rounded to nearest, so hi + mid + lo holds
// v's 24-bit significand
template <int NSL>
__device__ __forceinline__ void split4_bf(const f4 &v, bf4 &hi, bf4 &mid, bf4 &lo)
{
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float x = (j < NSL) ? v[j] : 0.0f;
        const __bf16 h = (__bf16)x;
        const float r1 = x - (float)h;
        const __bf16 m = (__bf16)r1;
        hi[j] = h;
        mid[j] = m;
        lo[j] = (__bf16)(r1 - (float)m);
    }
}

__device__ __forceinline__ bf8 cat_bf(const bf4 &a, const bf4 &b)
{
    return bf8{a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3]};
}

// acc + O S for one 16 x 16 block on the BF16 matrix core; the six split
// products above 2^-24 of hi hi, smallest first.  State 4 r + g is register
// r of lane group g in both operands (the FP32 slab layout below), so
// neither operand moves between lanes.
// PAIRS: two products per 16 x 16 x 32 MFMA (about 18 clk): lane group g
// holds K indices 8 g .. 8 g + 7, of which 8 g + r takes state 4 r + g of
// the first plane and 8 g + 4 + r the same state of the second; the pairs
// hi lo + lo hi, hi mid + mid hi, hi hi + mid mid.  Otherwise one product per
// 16 x 16 x 16 MFMA (about 17.5 clk), K index 4 g + r = state 4 r + g, which
// needs no concatenated operands (fewer registers: the L = 128 instances).
template <bool PAIRS>
__device__ __forceinline__ f4 apply_bf3(const bf4 &oh, const bf4 &om, const bf4 &ol, const bf4 &sh, const bf4 &sm,
                                        const bf4 &sl, f4 acc)
{
    if constexpr (PAIRS) {
        acc = mfma_b(cat_bf(oh, ol), cat_bf(sl, sh), acc);
        acc = mfma_b(cat_bf(oh, om), cat_bf(sm, sh), acc);
        acc = mfma_b(cat_bf(oh, om), cat_bf(sh, sm), acc);
    } else {
        acc = mfma_b16(ol, sh, acc);
        acc = mfma_b16(oh, sl, acc);
        acc = mfma_b16(om, sm, acc);
        acc = mfma_b16(om, sh, acc);
        acc = mfma_b16(oh, sm, acc);
        acc = mfma_b16(oh, sh, acc);
    }
    return acc;
}

// acc + M S over the first NSL K slabs; M is this lane's float4 of an FP32
// block (slab m = states 4 m .. 4 m + 3)
template <int NSL>
__device__ __forceinline__ f4 apply(const f4 &M, const f4 &S, f4 acc)
{
#pragma unroll
    for (int m = 0; m < NSL; ++m) acc = mfma_f(M[m], S[m], acc);
    return acc;
}

// the same on the FP64 matrix core: FP32 operands, FP64 accumulation
template <int NSL>
__device__ __forceinline__ d4 apply_d(const f4 &M, const d4 &S, d4 acc)
{
#pragma unroll
    for (int m = 0; m < NSL; ++m) acc = mfma_d((double)M[m], S[m], acc);
    return acc;
}

template <int D>
__device__ __forceinline__ double row_ror_d(double v)
{
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, 0x120 + D, 0xF, 0xF, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), 0x120 + D, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
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

// shift_tiles on FP64 state blocks
template <int D, int NSL, int NB>
__device__ __forceinline__ void shift_tiles_d(const d4 (&Z)[NB], d4 (&S)[NB])
{
    if constexpr (D >= 16) {
        constexpr int dn = D >> 4;
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) S[nb] = (nb >= dn) ? Z[nb >= dn ? nb - dn : 0] : d4{0.0, 0.0, 0.0, 0.0};
    } else {
        const bool wrap = (threadIdx.x & 15) < D;
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (j < NSL) {
                    const double r0 = row_ror_d<D>(Z[nb][j]);
                    const double r1 = (nb > 0) ? row_ror_d<D>(Z[nb > 0 ? nb - 1 : 0][j]) : 0.0;
                    S[nb][j] = wrap ? r1 : r0;
                } else {
                    S[nb][j] = 0.0;
                }
            }
        }
    }
}

template <int G>
constexpr int bs_levels() { return bs_scan_levels(G); }

// an FP32 operand block (word 4 lane + m = M[l%16][l/16 + 4 m]) with its rows
// reordered for the FP32 matrix core: row 4 g + r takes state g + 4 r, so that
// the 16 x 16 x 4 FP32 product lands in the FP64 result layout (lane g,
// register r = state g + 4 r)
__device__ __forceinline__ f4 rows_for_fp32(const f4 &M)
{
    const int l = threadIdx.x & 63;
    const int src = (l & 48) | bs_state_of_row(l & 15);
    f4 r;
#pragma unroll
    for (int m = 0; m < 4; ++m) r[m] = __shfl(M[m], src, 64);
    return r;
}

// Z += M S on the FP32 matrix core: S rounded to float32 and shifted by D
// tiles, M with rows_for_fp32 applied, the product added to Z in FP64
template <int D, int NSL, int NB>
__device__ __forceinline__ void level_fp32(const f4 &Mp, d4 (&Z)[NB])
{
    f4 Zs[NB], S[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
        for (int j = 0; j < 4; ++j) Zs[nb][j] = (j < NSL) ? (float)Z[nb][j] : 0.0f;
    shift_tiles<D, NSL, NB>(Zs, S);
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
        if (D < 16 || nb >= (D >> 4)) {
            const f4 acc = apply<NSL>(Mp, S[nb], f4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
            for (int j = 0; j < NSL; ++j) Z[nb][j] += (double)acc[j];
        }
    }
}

template <int NB, int KQ>
struct BsRaw {
    f4 a[NB][KQ], b[NB][KQ];
};

template <int L, int G, int NSL, int SPLIT = kBsSplitF16, bool LINES = false, class Ops, class Store>
__device__ __forceinline__ void bs_chunk(const BsRaw<G / 16, L / 32> &xr, uint32_t chunk, uint32_t part_flag,
                                         uint32_t full_flag, uint64_t *lb, const float *Mc, int depth, float *lp,
                                         uint32_t spin_limit, uint32_t *err, uint32_t first, int sT, int sG,
                                         int lv32, int nlv, const Ops &ops, const Store &store,
                                         const float *start = nullptr,
                                         float *tail = nullptr)
{
    constexpr int MB = L / 16;  // output row blocks of T
    constexpr int KQ = L / 32;  // K slabs of 32 samples
    constexpr int NB = G / 16;  // column blocks of 16 tiles
    const int lane = threadIdx.x & 63, col = lane & 15, g = lane >> 4;

    // 1. per-chunk power-of-two scale and (FP16) hi/lo split of X; the BF16
    //    three-way split is formed at each use from the raw samples instead
    //    (its 12 registers per operand would not all fit)
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
        if constexpr (SPLIT == kBsSplitF16) {
#pragma unroll
            for (int nb = 0; nb < NB; ++nb)
#pragma unroll
                for (int q = 0; q < KQ; ++q) split8(xr.a[nb][q], xr.b[nb][q], eX, xh[nb][q], xl[nb][q]);
        }
    }

    // 2. V = Gam X (zero chunk-incoming state); Gam's rows are ordered so that
    //    lane group g, register j holds state g + 4 j, the FP64 MFMA's layout.
    //    Y0 = T X follows the partial.
    d4 Z[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
        f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int q = 0; q < KQ; ++q) {
            if constexpr (SPLIT == kBsSplitF16) {
                const h8 gh = ops.gam(0, q), gl = ops.gam(1, q);
                acc = mfma_h(gl, xh[nb][q], acc);
                acc = mfma_h(gh, xl[nb][q], acc);
                acc = mfma_h(gh, xh[nb][q], acc);
            } else {
                acc = split_product<SPLIT>([&](int hl) { return ops.gam(hl, q); }, xr.a[nb][q], xr.b[nb][q], eX, acc);
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) Z[nb][j] = ldexp((double)acc[j], -(sG + eX));
    }
    // 3. tile end states, log2(G)-level scan on the FP64 matrix core (FP32
    //    operands Phi^(2^j), FP64 accumulation; levels D >= 16 skip the
    //    all-zero blocks).  Only the first nlv levels run (the longer lags
    //    weigh at most 2^-36 in total), those from lv32 on with the product on
    //    the FP32 matrix core (||Phi^(2^j)||_inf <= 2^-4 there)
    // (the FP64 levels, then the FP32 ones, each a uniform branch that leaves
    // Z in place: a three-way branch per level made the compiler copy Z
    // between the paths, about 300 extra moves per chunk)
    const int n64 = lv32 < nlv ? lv32 : nlv;
    static_for<0, bs_levels<G>(), 1>([&](auto LV) {
        constexpr int j = decltype(LV)::value;
        constexpr int D = 1 << j;
        if (j < n64) {
            d4 S[NB];
            shift_tiles_d<D, NSL, NB>(Z, S);
            const f4 P = ops.pow2(j);
#pragma unroll
            for (int nb = 0; nb < NB; ++nb)
                if (D < 16 || nb >= (D >> 4)) Z[nb] = apply_d<NSL>(P, S[nb], Z[nb]);
        }
    });
    static_for<0, bs_levels<G>(), 1>([&](auto LV) {
        constexpr int j = decltype(LV)::value;
        constexpr int D = 1 << j;
        if (j >= n64 && j < nlv) level_fp32<D, NSL, NB>(rows_for_fp32(ops.pow2(j)), Z);
    });

    // 4. chunk partial = Z(:, G-1): lanes 15, 31, 47, 63 hold states g + 4 j
    uint64_t *rec = lb + (size_t)chunk * 32;
    if (col == 15) {
#pragma unroll
        for (int j = 0; j < 4; ++j) rec_store(rec + g + 4 * j, part_flag, (float)Z[NB - 1][j]);
    }
    // the predecessor's record is in flight while Y0 = T X runs on the matrix cores
    const uint64_t poll0 = (chunk > first) ? lookback_first_poll<16>(lb, 32, chunk) : 0ull;
    // row block mb of tile block nb of T X (FP16 split)
    auto tx_block = [&](int mb, int nb) {
        const int qmax = bs_tq(mb);  // T(16 mb + r, 32 q + c) = 0 beyond
        f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int q = 0; q < KQ; ++q) {
            if (q > qmax) break;
            const h8 th = ops.tmat(0, mb, q), tl = ops.tmat(1, mb, q);
            acc = mfma_h(tl, xh[nb][q], acc);
            acc = mfma_h(th, xl[nb][q], acc);
            acc = mfma_h(th, xh[nb][q], acc);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[j] = ldexpf(acc[j], -(sT + eX));
        return acc;
    };
    f4 Y0[MB][NB];
    if constexpr (SPLIT == kBsSplitF16) {
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) Y0[mb][nb] = tx_block(mb, nb);
    } else {
        // slab by slab, so that each slab's split is formed once for all row blocks
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) Y0[mb][nb] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
#pragma unroll
            for (int q = 0; q < KQ; ++q) {
                bf8 xh3, xm3, xl3;
                split8_bf(xr.a[nb][q], xr.b[nb][q], eX, xh3, xm3, xl3);
#pragma unroll
                for (int mb = 0; mb < MB; ++mb) {
                    if (q > bs_tq(mb)) continue;
                    const bf8 th = __builtin_bit_cast(bf8, ops.tmat(0, mb, q));
                    const bf8 tm = __builtin_bit_cast(bf8, ops.tmat(1, mb, q));
                    const bf8 tl = __builtin_bit_cast(bf8, ops.tmat(2, mb, q));
                    f4 acc = Y0[mb][nb];
                    acc = mfma_b(tl, xh3, acc);
                    acc = mfma_b(th, xl3, acc);
                    acc = mfma_b(tm, xm3, acc);
                    acc = mfma_b(tm, xh3, acc);
                    acc = mfma_b(th, xm3, acc);
                    acc = mfma_b(th, xh3, acc);
                    Y0[mb][nb] = acc;
                }
            }
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int nb = 0; nb < NB; ++nb)
#pragma unroll
                for (int j = 0; j < 4; ++j) Y0[mb][nb][j] = ldexpf(Y0[mb][nb][j], -(sT + eX));
    }

    const float Xj = lookbackK<16, kBsSlots>(lb, 32, chunk, part_flag, full_flag, Mc, depth, lp, spin_limit, err,
                                             &poll0, first, start);
    // the output stage's O blocks, fetched before any store of this chunk:
    // vmcnt retires in issue order, so a load issued behind stores waits for them
    f4 Ob[MB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) Ob[mb] = ops.obs(mb);
    asm volatile("" ::: "memory");
    // X_in in B layout on column 0 (state g + 4 m lives on lane g + 4 m)
    f4 xin;
#pragma unroll
    for (int j = 0; j < 4; ++j) xin[j] = (j < NSL) ? __shfl(Xj, g + 4 * j, 64) : 0.0f;

    // injection W_k = Phi^(k+1) X_in: Phi X_in in column 0 of one block, a
    // 4-level scan inside the block, then Phi^(16 nb) for blocks 1..NB-1,
    // all accumulated in FP64
    d4 U[1];
    {
        d4 b0;
#pragma unroll
        for (int j = 0; j < 4; ++j) b0[j] = (col == 0) ? (double)xin[j] : 0.0;
        U[0] = apply_d<NSL>(ops.phi(), b0, d4{0.0, 0.0, 0.0, 0.0});
    }
    static_for<0, 4, 1>([&](auto LV) {
        constexpr int j = decltype(LV)::value;
        constexpr int D = 1 << j;
        if (j < n64) {
            d4 S[1];
            shift_tiles_d<D, NSL, 1>(U, S);
            U[0] = apply_d<NSL>(ops.pow2(j), S[0], U[0]);
        }
    });
    static_for<0, 4, 1>([&](auto LV) {
        constexpr int j = decltype(LV)::value;
        constexpr int D = 1 << j;
        if (j >= n64 && j < nlv) level_fp32<D, NSL, 1>(rows_for_fp32(ops.pow2(j)), U);
    });
    // Phi^(16 nb) W for the blocks past the first (lags >= 16: level 4)
    if (nlv > 4) {
        if (4 < lv32) {
#pragma unroll
            for (int nb = 1; nb < NB; ++nb) Z[nb] = apply_d<NSL>(ops.pow16(nb), U[0], Z[nb]);
        } else {
            f4 Us;
#pragma unroll
            for (int j = 0; j < 4; ++j) Us[j] = (j < NSL) ? (float)U[0][j] : 0.0f;
#pragma unroll
            for (int nb = 1; nb < NB; ++nb) {
                const f4 acc = apply<NSL>(rows_for_fp32(ops.pow16(nb)), Us, f4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
                for (int j = 0; j < NSL; ++j) Z[nb][j] += (double)acc[j];
            }
        }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) Z[0][j] += U[0][j];

    // the states, rounded once to FP32; chunk full = Z(:, G-1)
    f4 Zf[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
        for (int j = 0; j < 4; ++j) Zf[nb][j] = (float)Z[nb][j];
    if (col == 15) {
#pragma unroll
        for (int j = 0; j < 4; ++j) rec_store(rec + 16 + g + 4 * j, full_flag, Zf[NB - 1][j]);
    }

    // 5. S_k = state entering tile k (X_in for tile 0); Y = Y0 + O S
    f4 S[NB];
    shift_tiles<1, NSL, NB>(Zf, S);
    if (col == 0) S[0] = xin;
    if (tail && col == 15) {  // the state entering the chunk's last tile
#pragma unroll
        for (int j = 0; j < 4; ++j) tail[g + 4 * j] = (j < NSL) ? S[NB - 1][j] : 0.0f;
    }
    // Y = Y0 + O S: from three K slabs on (orders 9-16) on the BF16 matrix
    // core (apply_bf3: three or six MFMAs of about 18 clk per block against
    // NSL FP32 ones of 32 clk); below that the FP32 products.  The BF16
    // hi/mid/lo split of T and Gam keeps the FP32 products (order 16, 32 x 64:
    // no faster with the BF16 ones, max error 6.6e-6 against 5.7e-6 of the
    // output RMS)
    constexpr bool kOsBf = NSL >= 3 && SPLIT == kBsSplitF16;
    bf4 oh[kOsBf ? MB : 1], om[kOsBf ? MB : 1], ol[kOsBf ? MB : 1];
    if constexpr (kOsBf) {
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) split4_bf<NSL>(Ob[mb], oh[mb], om[mb], ol[mb]);
    }
    auto out_block = [&](int mb, int nb, const bf4 &sh, const bf4 &sm, const bf4 &sl) {
        if constexpr (kOsBf) return apply_bf3<(MB <= 4)>(oh[mb], om[mb], ol[mb], sh, sm, sl, Y0[mb][nb]);
        else return apply<NSL>(Ob[mb], S[nb], Y0[mb][nb]);
    };
    if constexpr (LINES) {
        // whole 128-byte lines: row blocks mb and mb + 1 of a tile are one
        // line, held by lanes col and col + 8 of the two blocks; a DPP
        // rotation by 8 inside each 16-lane row swaps the halves so that one
        // store writes the lines of tiles 0..7 of the block and the next those
        // of tiles 8..15.  store(tile, row, y): samples row .. row + 3 of tile.
        const bool lo = col < 8;
        const int row = 16 * (col >> 3) + 4 * g;
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            bf4 sh, sm, sl;
            if constexpr (kOsBf) split4_bf<NSL>(S[nb], sh, sm, sl);
#pragma unroll
            for (int mb = 0; mb < MB; mb += 2) {
                const f4 ya = out_block(mb, nb, sh, sm, sl);
                const f4 yb = out_block(mb + 1, nb, sh, sm, sl);
                f4 A, B;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const float ra = row_ror<8>(ya[j]), rb = row_ror<8>(yb[j]);
                    A[j] = lo ? ya[j] : rb;
                    B[j] = lo ? ra : yb[j];
                }
                store(16 * nb + (col & 7), 16 * mb + row, A);
                store(16 * nb + 8 + (col & 7), 16 * mb + row, B);
            }
        }
    } else {
        // a store covers 64 B of each of 16 tiles; the two halves of a 128-byte
        // line (row blocks mb, mb + 1) are issued back to back and through the L2
        // (not nontemporal), so that they merge there: measured WRITE_SIZE 1.06x
        // the output bytes this way, 1.43x with mb outer and nontemporal stores.
        // store(nb, mb, y): row block mb of tile block nb.
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            bf4 sh, sm, sl;
            if constexpr (kOsBf) split4_bf<NSL>(S[nb], sh, sm, sl);
#pragma unroll
            for (int mb = 0; mb < MB; ++mb) store(nb, mb, out_block(mb, nb, sh, sm, sl));
        }
    }
}

}  // namespace iirx
