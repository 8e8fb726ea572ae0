// bs.hip -- block-state form of the cascade on the MI355X matrix cores.
//
// No counterpart in the reference; it computes the same filter as
// ph/stcr/dtcr (the whole cascade, zero initial state) in a form whose cost
// does not grow with the order.  The cascade is one state-space system of
// dimension p = 2 N_SECTIONS (padded to 16).  A chunk of G = 64 tiles of L
// samples is one wave64:
//   1. Y0 = T X and V = Gam X (T: L x L Toeplitz of the impulse response,
//      Gam: 16 x L), as v_mfma_f32_16x16x16_f16 products of FP16 hi/lo
//      operands (hi*hi + hi*lo + lo*hi) with FP32 accumulation; every operand
//      is scaled by a power of two first (per matrix for the constant ones,
//      per chunk for the data) so that FP16 holds it;
//   2. the tile end states Z_k = sum_{j<=k} Phi^(k-j) V_j by a 6-level
//      Kogge-Stone scan over the 64 tiles, each level one 16x16x64 product
//      Phi^(2^j) * shift(Z) on the matrix cores;
//   3. the chunk's incoming state X_in from the lookback (16-word records,
//      chunk transition Phi^64, bounded depth), injected as Phi^(k+1) X_in:
//      a 4-level scan inside one 16-tile block and three Phi^(16 nb) products;
//   4. Y = Y0 + O S, S_k = state entering tile k.
// The MFMA D layout (row 4(l/16)+j, column l%16) equals the B layout, so every
// product feeds the next one from registers; tiles are MFMA columns and the
// scan's shifts are lane shuffles inside 16-lane rows.
#include "device_common.hpp"
#include "kernels.hpp"

namespace iirx {

typedef _Float16 h4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f4 mfma(h4 a, h4 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0); }

__device__ __forceinline__ h4 load_blk(const float *tab, int word_off)
{
    const uint2 u = reinterpret_cast<const uint2 *>(tab + word_off)[threadIdx.x & 63];
    return __builtin_bit_cast(h4, u);
}

__device__ __forceinline__ float wave_max(float v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// power-of-two exponent s with max * 2^s < 2^14
__device__ __forceinline__ int scale_exp(float mx)
{
    return (mx > 0.0f) ? 14 - __builtin_amdgcn_frexp_expf(mx) : 0;
}

// hi/lo FP16 split of v * 2^s
__device__ __forceinline__ void split(const f4 &v, int s, h4 &hi, h4 &lo)
{
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float x = ldexpf(v[j], s);
        const _Float16 h = (_Float16)x;
        hi[j] = h;
        lo[j] = (_Float16)(x - (float)h);
    }
}

// 3-term compensated product into acc: A (hi/lo) * B (hi/lo)
__device__ __forceinline__ f4 mfma3(h4 ah, h4 al, h4 bh, h4 bl, f4 acc)
{
    acc = mfma(al, bh, acc);
    acc = mfma(ah, bl, acc);
    acc = mfma(ah, bh, acc);
    return acc;
}

// S(c, k) = Z(c, k - d) (0 for k < d) over the 64 tiles held as 4 column
// blocks of 16; d < 16 moves values inside each 16-lane row
__device__ __forceinline__ void shift_tiles(const f4 (&Z)[4], f4 (&S)[4], int d)
{
    const int lane = threadIdx.x & 63;
    if (d >= 16) {
        const int dn = d >> 4;
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) S[nb] = (nb >= dn) ? Z[nb - dn < 0 ? 0 : nb - dn] : f4{0.f, 0.f, 0.f, 0.f};
        return;
    }
    const int src = (lane & 48) | ((lane - d) & 15);
    const bool wrap = (lane & 15) < d;
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
        f4 r0, r1;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            r0[j] = __shfl(Z[nb][j], src, 64);
            r1[j] = (nb > 0) ? __shfl(Z[nb > 0 ? nb - 1 : 0][j], src, 64) : 0.0f;
        }
        S[nb] = wrap ? r1 : r0;
    }
}

__device__ __forceinline__ float abs_max4(const f4 (&v)[4])
{
    float m = 0.0f;
#pragma unroll
    for (int nb = 0; nb < 4; ++nb)
#pragma unroll
        for (int j = 0; j < 4; ++j) m = fmaxf(m, fabsf(v[nb][j]));
    return m;
}

// Z += M^(2^j)-style product: Z[nb] += Pm * S[nb], Pm given as hi/lo blocks with scale sP
__device__ __forceinline__ void apply_blocks(f4 (&Z)[4], const f4 (&S)[4], h4 ph, h4 pl, int sP, int nblocks)
{
    const int eS = scale_exp(wave_max(abs_max4(S)));
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
        if (nb >= nblocks) break;
        h4 sh, sl;
        split(S[nb], eS, sh, sl);
        f4 t = {0.f, 0.f, 0.f, 0.f};
        t = mfma3(ph, pl, sh, sl, t);
#pragma unroll
        for (int j = 0; j < 4; ++j) Z[nb][j] = fmaf(t[j], ldexpf(1.0f, -(sP + eS)), Z[nb][j]);
    }
}

template <int L>
__global__ __launch_bounds__(64) void bs_kernel(LaunchArgs a)
{
    constexpr int MB = L / 16;  // row/column blocks of T
    constexpr int G = kBsTiles;
    constexpr int CH = L * G;
    constexpr BsLayout Y = bs_layout(L);
    __shared__ float lp[64 * 16];
    __shared__ uint4 xop[MB * 4 * 64];  // FP16 hi/lo input operands, reused by the final product

    const int lane = threadIdx.x & 63;
    const float *tab = a.coef;
    const cfloat *ctab = as_const(a.coef);
    const int sT = (int)ctab[Y.kScale + 0], sO = (int)ctab[Y.kScale + 1], sG = (int)ctab[Y.kScale + 2];
    const int depth = (int)ctab[Y.kDepth];

    for_each_chunk(a.ticket, a.epoch, a.num_chunks, a.persistent,
                   [&](uint32_t chunk, uint32_t part_flag, uint32_t full_flag) {
    const long base = (long)chunk * CH;
    const int col = lane & 15, rq = lane >> 4;

    // 1. input tile blocks in B layout: (sample 16 kb + 4 rq + j, tile 16 nb + col)
    f4 Z[4];
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) Z[nb] = f4{0.f, 0.f, 0.f, 0.f};
    int eX;
    {
        f4 xv[MB][4];
#pragma unroll
        for (int kb = 0; kb < MB; ++kb)
#pragma unroll
            for (int nb = 0; nb < 4; ++nb)
                xv[kb][nb] = __builtin_nontemporal_load(
                    reinterpret_cast<const f4 *>(a.in + base + (long)(16 * nb + col) * L + 16 * kb + 4 * rq));
        float mx = 0.0f;
#pragma unroll
        for (int kb = 0; kb < MB; ++kb)
#pragma unroll
            for (int nb = 0; nb < 4; ++nb)
#pragma unroll
                for (int j = 0; j < 4; ++j) mx = fmaxf(mx, fabsf(xv[kb][nb][j]));
        eX = scale_exp(wave_max(mx));
        // V = Gam X (zero chunk-incoming state), operands kept in LDS for Y0 = T X
#pragma unroll
        for (int kb = 0; kb < MB; ++kb) {
            const h4 gh = load_blk(tab, bs_G(Y, 0, kb)), gl = load_blk(tab, bs_G(Y, 1, kb));
#pragma unroll
            for (int nb = 0; nb < 4; ++nb) {
                h4 xh, xl;
                split(xv[kb][nb], eX, xh, xl);
                Z[nb] = mfma3(gh, gl, xh, xl, Z[nb]);
                uint4 w;
                const uint2 uh = __builtin_bit_cast(uint2, xh), ul = __builtin_bit_cast(uint2, xl);
                w.x = uh.x;
                w.y = uh.y;
                w.z = ul.x;
                w.w = ul.y;
                xop[(kb * 4 + nb) * 64 + lane] = w;
            }
        }
    }
    {
        const float u = ldexpf(1.0f, -(sG + eX));
#pragma unroll
        for (int nb = 0; nb < 4; ++nb)
#pragma unroll
            for (int j = 0; j < 4; ++j) Z[nb][j] *= u;
    }

    // 2. tile end states, 6-level scan
#pragma unroll
    for (int lev = 0; lev < 6; ++lev) {
        f4 S[4];
        shift_tiles(Z, S, 1 << lev);
        apply_blocks(Z, S, load_blk(tab, bs_P(Y, 0, lev)), load_blk(tab, bs_P(Y, 1, lev)),
                     (int)ctab[Y.kScale + 3 + lev], 4);
    }

    // 3. chunk partial = Z(:, 63): lanes 15, 31, 47, 63 hold components 4 rq + j
    uint64_t *rec = a.lb + (size_t)chunk * 32;
    if (col == 15) {
#pragma unroll
        for (int j = 0; j < 4; ++j) rec_store(rec + 4 * rq + j, part_flag, Z[3][j]);
    }
    const float Xj = lookbackK<16>(a.lb, 32, chunk, part_flag, full_flag, tab + Y.kMc, depth, lp, a.spin_limit,
                                   a.err);

    // injection W_k = Phi^(k+1) X_in: column 0 of block 0 = Phi X_in, a
    // 4-level scan inside the block, then Phi^(16 nb) for blocks 1..3
    f4 U[4];
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) U[nb] = f4{0.f, 0.f, 0.f, 0.f};
    {
        float acc[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            const float xm = readlane_f(Xj, m);
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[j] = fmaf(tab[Y.kPhi + (4 * rq + j) * 16 + m], xm, acc[j]);
        }
        if (col == 0) {
#pragma unroll
            for (int j = 0; j < 4; ++j) U[0][j] = acc[j];
        }
    }
#pragma unroll
    for (int lev = 0; lev < 4; ++lev) {
        f4 S[4];
        shift_tiles(U, S, 1 << lev);
        apply_blocks(U, S, load_blk(tab, bs_P(Y, 0, lev)), load_blk(tab, bs_P(Y, 1, lev)),
                     (int)ctab[Y.kScale + 3 + lev], 1);
    }
    {
        const int eU = scale_exp(wave_max(fmaxf(fmaxf(fabsf(U[0][0]), fabsf(U[0][1])),
                                                fmaxf(fabsf(U[0][2]), fabsf(U[0][3])))));
        h4 uh, ul;
        split(U[0], eU, uh, ul);
#pragma unroll
        for (int nb = 1; nb < 4; ++nb) {
            f4 t = {0.f, 0.f, 0.f, 0.f};
            t = mfma3(load_blk(tab, bs_Q(Y, 0, nb)), load_blk(tab, bs_Q(Y, 1, nb)), uh, ul, t);
            const float u = ldexpf(1.0f, -((int)ctab[Y.kScale + 9 + nb - 1] + eU));
#pragma unroll
            for (int j = 0; j < 4; ++j) Z[nb][j] = fmaf(t[j], u, Z[nb][j]);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) Z[0][j] += U[0][j];
    }

    // chunk full = Z(:, 63)
    if (col == 15) {
#pragma unroll
        for (int j = 0; j < 4; ++j) rec_store(rec + 16 + 4 * rq + j, full_flag, Z[3][j]);
    }

    // 4. S_k = state entering tile k; Y = T X + O S
    f4 S[4];
    shift_tiles(Z, S, 1);
    {
        // column 0 enters with X_in: component 4 rq + j lives on lane 4 rq + j
        f4 xin;
#pragma unroll
        for (int j = 0; j < 4; ++j) xin[j] = __shfl(Xj, 4 * rq + j, 64);
        if (col == 0) S[0] = xin;
    }
    const int eS = scale_exp(wave_max(abs_max4(S)));
    h4 sh[4], sl[4];
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) split(S[nb], eS, sh[nb], sl[nb]);
    const float uT = ldexpf(1.0f, -(sT + eX)), uO = ldexpf(1.0f, -(sO + eS));
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) {
        const h4 oh = load_blk(tab, bs_O(Y, 0, mb)), ol = load_blk(tab, bs_O(Y, 1, mb));
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
            f4 ty = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kb = 0; kb < MB; ++kb) {
                const uint4 w = xop[(kb * 4 + nb) * 64 + lane];
                const h4 xh = __builtin_bit_cast(h4, make_uint2(w.x, w.y));
                const h4 xl = __builtin_bit_cast(h4, make_uint2(w.z, w.w));
                ty = mfma3(load_blk(tab, bs_T(Y, 0, mb, kb)), load_blk(tab, bs_T(Y, 1, mb, kb)), xh, xl, ty);
            }
            f4 to = {0.f, 0.f, 0.f, 0.f};
            to = mfma3(oh, ol, sh[nb], sl[nb], to);
            f4 y;
#pragma unroll
            for (int j = 0; j < 4; ++j) y[j] = fmaf(to[j], uO, ty[j] * uT);
            __builtin_nontemporal_store(
                y, reinterpret_cast<f4 *>(a.out + base + (long)(16 * nb + col) * L + 16 * mb + 4 * rq));
        }
    }
    });
}

template <int L>
static hipError_t launch_l(const LaunchArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(bs_kernel<L>, dim3(a.grid), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_bs(int L, const LaunchArgs &a, hipStream_t s)
{
    switch (L) {
    case 32: return launch_l<32>(a, s);
    case 64: return launch_l<64>(a, s);
    default: return hipErrorInvalidValue;
    }
}

const void *kernel_bs(int L)
{
    switch (L) {
    case 32: return (const void *)bs_kernel<32>;
    case 64: return (const void *)bs_kernel<64>;
    default: return nullptr;
    }
}

}  // namespace iirx
