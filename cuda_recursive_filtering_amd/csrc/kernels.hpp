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
    unsigned long long *ticket;  // 64-bit workgroup ticket counter (ticket schedule)
    uint32_t *epoch;             // launch epoch (persistent schedule)
    uint32_t *err;               // watchdog flag
    uint32_t *arrive;            // arrival slots: waves that have read the epoch (persistent schedule)
    int64_t n_samples;
    uint32_t num_chunks;         // chunks per launch
    uint32_t spin_limit;         // lookback polls before the watchdog trips
    int32_t persistent;          // 1: persistent schedule, 0: ticket schedule
    uint32_t grid;               // workgroups launched
    int32_t ns;                  // sections (PH/STCR/DTCR) or order k (PLR)
    int32_t coef_stride;
    int32_t unrolled;
    // streaming state in (null: zero state).  PH/STCR/DTCR: x[-2], x[-1] of
    // the input, then y_s[-2], y_s[-1] of every section's output (2 + 2 ns
    // floats); PLR: x[-K..-1], y[-K..-1]; block-state form: the balanced
    // state vector (16 floats, natural order)
    const float *zi;
    // where the last chunk writes the end state (null: not written).
    // PH/STCR/DTCR, partial last chunk: x[n-2], x[n-1], then y_s[n-2],
    // y_s[n-1] of every section (2 + 2 ns floats); block-state form: the
    // state entering its last tile (16 floats)
    float *tail;
    // PH/STCR/DTCR: samples in the partial last chunk
    int32_t last_valid;
};

// plan control block: ticket, epoch, watchdog flag, then 64 arrival slots
// 128 bytes apart
constexpr size_t kControlBytes = 128 + 64 * 128;

constexpr int ilog2_c(int v) { return v <= 1 ? 0 : 1 + ilog2_c(v >> 1); }
constexpr int align4_c(int v) { return (v + 3) & ~3; }

// ||M^d||_inf falls below this before the fold of the lookback stops
constexpr double kLookbackTail = 1.4551915228366852e-11;  // 2^-36

// ---------------------------------------------------------------------------
// PH tables, one block per section (L = lanes per workgroup, N = samples per
// lane).  C is the N-step transition of the section state (y[n-2], y[n-1]),
// row-major 2x2 (ph_kernels.cuh:180-181).
//   [kB]     b1, b2, a1, a2          (a in recurrence sign, a = -sos[4:6])
//   [kCpow]  C^(2^k), k < log2 L     uniform lane-scan operators
//   [kM]     C^L                     chunk transition used by the lookback
//   [kDepth] lookback depth d (smallest d with ||(C^L)^d||_inf <= 2^-36,
//            0 = unbounded), stored as a float
//   [kCfin]  C^(t+1), t < L          per-lane carry injection (C_inner)
//   [kH1]    h1[n] = h0[n+1], n < N  zero-input response to y[-1]
//   [kH2]    h2[n] = a2 h0[n]        zero-input response to y[-2]
// ---------------------------------------------------------------------------
struct PhLayout {
    int kB, kCpow, kM, kDepth, kCfin, kH1, kH2, stride;
};

constexpr PhLayout ph_layout(int L, int N)
{
    PhLayout p{};
    p.kB = 0;
    p.kCpow = 4;
    p.kM = p.kCpow + 4 * ilog2_c(L);
    p.kDepth = p.kM + 4;
    p.kCfin = p.kDepth + 4;
    p.kH1 = p.kCfin + 4 * L;
    p.kH2 = p.kH1 + N;
    p.stride = align4_c(p.kH2 + N);
    return p;
}

// PH with sections in pairs (s, t = s + 1): the pair's joint state
// U = (y_s[-2], y_s[-1], y_t[-2], y_t[-1]) at a lane's start, J its N-step
// transition (zero input), block lower triangular (y_s does not see y_t).
// Per pair, after the per-section tables of all sections:
//   kJpow  J^(2^k), k = 0..4, 16 floats each (row-major; uniform)
//   kM     J^L, the chunk transition (uniform)
//   kDepth lookback depth for J^L; kDepth + 1: 1 if the pair runs joined,
//          0 if its tables are too large (kPairLimit) and the two sections
//          run one by one
//   kJrow  J^(j+1), j = 0..15, per lane of a 16-lane row
//   kJfin  J^(l+1), l = 0..L-1, per lane
//   kH     H_m[n], m = 0..3, n = 0..N-1: y_t at sample n of a lane from the
//          unit joint state e_m at the lane's start (uniform)
// A pair runs joined only while every entry of its tables (H, J^(l+1), J^L)
// stays within kPairLimit of the states' own scale.  The cross terms grow
// like 1/(distance between the two sections' poles): for Butterworth
// lowpass filters about 3, 13, 72, 480 and 6600 at Wn = 0.4, 0.2, 0.1, 0.05
// and 0.02 (order 4), and they cancel in the correction sum y_t += H U, so
// the float32 error grows with them (order 4, Wn = 0.002: 14x the
// sequential float32 cascade's error joined, against 1x one by one).
constexpr double kPairLimit = 256.0;

struct PhPairLayout {
    int kJpow, kM, kDepth, kJrow, kJfin, kH, stride;
};

constexpr PhPairLayout ph_pair_layout(int L, int N)
{
    PhPairLayout p{};
    p.kJpow = 0;
    p.kM = p.kJpow + 16 * 5;
    p.kDepth = p.kM + 16;
    p.kJrow = p.kDepth + 4;
    p.kJfin = p.kJrow + 16 * 16;
    p.kH = p.kJfin + 16 * L;
    p.stride = align4_c(p.kH + 4 * N);
    return p;
}

// PH with halo carries (IIRX_KERNEL_PH_HALO).  When the cascade forgets its
// state within kPhHalo samples -- ||A^kPhHalo||_inf <= kHaloTail for A, the
// zero-input one-sample transition of the whole cascade's state
// (x[-2], x[-1], y_1[-2], y_1[-1], ..., y_ns[-2], y_ns[-1]) -- the state
// entering a chunk is, to that bound (2^-34 of the largest state entry, a
// thousandth of a float32 ulp), the state the cascade reaches from
// zero over the kPhHalo input samples before the chunk:
//   y_s[-2] = sum_i K2_s[i] x[i - kPhHalo],  y_s[-1] = sum_i K1_s[i] x[i - kPhHalo]
// with K1_s[i] = h_s[kPhHalo - 1 - i] and K2_s[i] = h_s[kPhHalo - 2 - i]
// (h_s the impulse response of sections 1..s).  A chunk then reads 4 more
// samples per lane and needs no record of another chunk: no lookback.
// After the per-section tables of all sections:
//   [0]                           1 if the carries come from the halo, 0 if
//                                 the cascade's memory is longer (the
//                                 kernel then runs PH's lookback); 3 floats
//                                 of padding
//   [4 + 2 kPhHalo s]             K2_s[i], i < kPhHalo
//   [4 + 2 kPhHalo s + kPhHalo]   K1_s[i]
constexpr int kPhHalo = 256;
constexpr double kHaloTail = 5.820766091346741e-11;  // 2^-34
constexpr int ph_halo_floats(int ns) { return 4 + 2 * kPhHalo * ns; }

// Launchers (one per algorithm); `lanes` and `n_blocks` pick the template
// instance.  Return hipErrorInvalidValue when no instance exists.
// last = true: the one-workgroup launch of the partial last chunk (chunk
// a.num_chunks, a.last_valid samples), after the launch of the others
// joint: sections in pairs (IIRX_KERNEL_PH_JOINT); the pair tables follow the
// per-section ones at a.coef + a.ns * a.coef_stride
// halo: carries from the kPhHalo samples before each chunk (IIRX_KERNEL_PH_HALO
// when its tables mark the halo usable); the halo block follows the
// per-section tables at a.coef + a.ns * a.coef_stride
hipError_t launch_ph(int lanes, int n_blocks, const LaunchArgs &a, hipStream_t s, bool last = false,
                     bool joint = false, bool halo = false);
// PH and STCR under the ticket schedule: chunks (waves) per group, one
// atomic ticket per group; used when it divides the launch's chunk count
// (else one chunk per workgroup).  16 waves at N <= 32 (135 KB of LDS at
// N = 32, 4 waves/SIMD as with one wave per workgroup): PH 64 x 32 ran
// 638 / 522 / 395 / 255 GS/s at orders 2 / 4 / 8 / 16 against 617 / 532 /
// 395 / 246 with 8; STCR 64 x 64 ran 581 / 219 with 8 waves against 588 /
// 219 with 4 at orders 2 / 16.
// Between the ticket barriers the waves of such a workgroup synchronise only
// within themselves (their LDS tiles are their own), so a chunk that waits
// for the record of an earlier chunk of its group never holds that chunk
// back.
constexpr int ph_ticket_waves(int N) { return N <= 32 ? 16 : (N == 64 ? 4 : 2); }
// grouped: the ticket-schedule instance with ph_ticket_waves(n_blocks) chunks
// per workgroup
const void *kernel_ph(int lanes, int n_blocks, bool grouped = false, bool joint = false, bool halo = false);

}  // namespace iirx

namespace iirx {

// ---------------------------------------------------------------------------
// CR tables (STCR and DTCR), one block per section; LG = log2 N (levels
// 0..LG), L = lanes of the chunk.  CR factors of stcr_kernels.cuh:93-125:
// e_n = 2 f_{n-1} - e_{n-1}^2, f_n = f_{n-1}^2, fde = f/e, h, g, p, q
// (terminal recurrences), d, c (back-substitution), in FP64 then float32.
//   [kScal]  b1, b2, cb, db, fb, eb, hbc, gbc  (DTCR decoupled first round,
//            dtcr_kernels.cuh:134-139, 251-257)
//   [kE] [kF] [kFde] [kH] [kG] [kP] [kQ] [kD] [kC]   8 floats each (n <= 7)
//   [kApow]  A^(2^k), k < log2 L: lane-recurrence operator A = [[0,1],[-f_LG,-e_LG]]
//            (the block filter of stcr:388-405 as a lane scan)
//   [kM]     chunk transition C^L of the section state (lookback)
//   [kDepth] lookback depth
//   [kTerm]  per lane t: {he2, he1, hb2, hb1}[t]  (stcr:127-147)
// ---------------------------------------------------------------------------
struct CrLayout {
    int kScal, kE, kF, kFde, kH, kG, kP, kQ, kD, kC, kApow, kM, kDepth, kTerm, stride;
};

constexpr CrLayout cr_layout(int L)
{
    CrLayout p{};
    p.kScal = 0;
    p.kE = 8;
    p.kF = p.kE + 8;
    p.kFde = p.kF + 8;
    p.kH = p.kFde + 8;
    p.kG = p.kH + 8;
    p.kP = p.kG + 8;
    p.kQ = p.kP + 8;
    p.kD = p.kQ + 8;
    p.kC = p.kD + 8;
    p.kApow = p.kC + 8;
    p.kM = p.kApow + 4 * ilog2_c(L);
    p.kDepth = p.kM + 4;
    p.kTerm = p.kDepth + 4;
    p.stride = align4_c(p.kTerm + 4 * L);
    return p;
}

hipError_t launch_stcr(int lanes, int n_blocks, const LaunchArgs &a, hipStream_t s, bool last = false);
const void *kernel_stcr(int lanes, int n_blocks, int unrolled, bool grouped = false);
hipError_t launch_dtcr(int lanes, int n_blocks, const LaunchArgs &a, hipStream_t s, bool last = false);
const void *kernel_dtcr(int lanes, int n_blocks, int unrolled);

}  // namespace iirx

namespace iirx {

// ---------------------------------------------------------------------------
// PLR tables (one block; K = direct-form order, CH = chunk samples).  The
// n-nacci factors of plr_kernels.cuh:149-409 in FP64, rounded to float32:
// fac[m][i] = output i of y[n] = sum_l a[l] y[n-1-l] started from the unit
// history e_m (slot m = y[-K+m]).
//   [kB]     b[0..K)  FIR taps past b0 == 1 (float32 of the direct-form taps)
//   [kF]     F[j][m] = fac[m][CH-K+j]: chunk transition of the K-sample state
//   [kDepth] lookback depth (||F^d||_inf <= 2^-36, 0 = unbounded)
//   [kFac]   fac[m][i], i < CH
// ---------------------------------------------------------------------------
struct PlrLayout {
    int kB, kF, kDepth, kFac, stride;
};

constexpr PlrLayout plr_layout(int K, int CH)
{
    PlrLayout p{};
    p.kB = 0;
    p.kF = 16;
    p.kDepth = p.kF + K * K;
    p.kFac = align4_c(p.kDepth + 4);
    p.stride = align4_c(p.kFac + K * CH);
    return p;
}

hipError_t launch_plr(int k, int block, int x, const LaunchArgs &a, hipStream_t s);
const void *kernel_plr(int k, int block, int x);

}  // namespace iirx

namespace iirx {

// ---------------------------------------------------------------------------
// Block-state (MFMA) tables.  The whole cascade as one state-space system of
// dimension p = 2 N_SECTIONS <= 16 (direct form II transposed per section,
// chained), padded to 16.  Tiles of L samples, G = 64 tiles per chunk:
//   y_tile = T x_tile + O s,   s_next = Phi s + Gam x_tile
// T: L x L lower-triangular Toeplitz of the impulse response, O: L x 16
// (rows C A^i), Gam: 16 x L (columns A^(L-1-j) B), Phi = A^L.  The p = 2 NS
// states keep their order 0..p-1; the state products' K slab m holds states
// 4 m .. 4 m + 3, so they need only nsl = ceil(p / 4) of the four slabs.
// FP16 blocks (T, Gam; hi/lo of the matrix scaled by 2^s) are the A operand
// of v_mfma_f32_16x16x32_f16 in lane order: block word 4 lane + {0..3} = the
// 8 halves A[l%16][8 (l/16) + j] of K slab q (samples 32 q ..); K row
// 8 g + j is sample 32 q + 4 g + j (j < 4) or 32 q + 16 + 4 g + j - 4 (j >= 4).
// Gam's operand row i carries state bs_state_of_row(i).
// FP32 blocks are the A operands of the 16 x 16 x 4 state products (FP64 or
// FP32 MFMA): word 4 lane + m = M[l%16][l/16 + 4 m], slab m.
//   [kT]  T hi/lo x MB row blocks x KQ slabs   [kG] Gam hi/lo x KQ slabs
//   [kP]  Phi^(2^j), j < 6      [kQ] Phi^(16 nb), nb = 1..3     [kPhi] Phi
//   [kO]  O row blocks          [kScale] exponents of T and Gam
//   [kMc] chunk transition Phi^G (fp32 row major)    [kDepth] lookback depth
// ---------------------------------------------------------------------------
// The operand split of T and Gam: kBsSplitF16 = FP16 hi/lo (2 planes,
// v_mfma_f32_16x16x32_f16), kBsSplitBF16 = BF16 hi/mid/lo (3 planes,
// v_mfma_f32_16x16x32_bf16); plane hl of a block is component hl.
constexpr int kBsSplitF16 = 2;
constexpr int kBsSplitBF16 = 3;

struct BsLayout {
    int MB, KQ, split, kT, kG, kP, kQ, kPhi, kO, kScale, kMc, kDepth, stride;
};

constexpr int kBsBlockWords = 256;

constexpr BsLayout bs_layout(int L, int split = kBsSplitF16)
{
    BsLayout p{};
    p.MB = L / 16;
    p.KQ = L / 32;
    p.split = split;
    p.kT = 0;
    p.kG = p.kT + split * p.MB * p.KQ * kBsBlockWords;
    p.kP = p.kG + split * p.KQ * kBsBlockWords;
    p.kQ = p.kP + 6 * kBsBlockWords;
    p.kPhi = p.kQ + 3 * kBsBlockWords;
    p.kO = p.kPhi + kBsBlockWords;
    p.kScale = p.kO + p.MB * kBsBlockWords;
    p.kMc = p.kScale + 4;
    p.kDepth = p.kMc + 256;
    p.stride = align4_c(p.kDepth + 4);
    return p;
}

// word offsets of operand blocks
constexpr int bs_T(const BsLayout &y, int hl, int mb, int q) { return y.kT + ((hl * y.MB + mb) * y.KQ + q) * kBsBlockWords; }
constexpr int bs_G(const BsLayout &y, int hl, int q) { return y.kG + (hl * y.KQ + q) * kBsBlockWords; }
constexpr int bs_P(const BsLayout &y, int j) { return y.kP + j * kBsBlockWords; }
constexpr int bs_Q(const BsLayout &y, int nb) { return y.kQ + (nb - 1) * kBsBlockWords; }
constexpr int bs_O(const BsLayout &y, int mb) { return y.kO + mb * kBsBlockWords; }
constexpr int bs_slabs(int ns) { return (ns + 1) / 2; }
// state order of the block-state products: lane group g, register j of a
// 16 x 16 x 4 FP64 MFMA result holds state g + 4 j (row g + 4 j).  The FP16
// Gam product puts its row 4 g + j on that lane and register, so Gam's
// operand row i carries state bs_state_of_row(i).
constexpr int bs_state_of_row(int i) { return (i >> 2) + 4 * (i & 3); }
// last nonzero K slab of T's row block mb, and the compact index of (mb, 0)
// among T's nonzero blocks (blocks above the diagonal are zero)
constexpr int bs_tq(int mb) { return (16 * mb + 15) / 32; }
constexpr int bs_tidx(int mb)
{
    int s = 0;
    for (int m = 0; m < mb; ++m) s += bs_tq(m) + 1;
    return s;
}
// the chunk staged through the wave's LDS buffer with contiguous accesses
// both ways (bs_kernel STAGED): 64 x 64 with the FP16 split
constexpr bool bs_staged(int L, int G, int split) { return L == 64 && G == 64 && split == 2; }
// waves per SIMD the register budget is set for, and waves per workgroup
// (the workgroup shares T, Gam and Phi^(2^j) in LDS; each wave runs its own
// chunks with its own input buffer of G L floats and 16 lookback slots)
constexpr int bs_waves(int L, int G, int split = kBsSplitF16)
{
    // the BF16 split keeps the raw samples and forms the split at each use:
    // one wave less at L = 64, so that it does not spill
    if (split == kBsSplitBF16 && L == 64) return G == 32 ? 2 : 3;
    if (L == 128) return G == 16 ? 3 : 2;
    return G == 64 ? (L == 32 ? 3 : 2) : G == 32 ? (L == 32 ? 4 : 3) : (L == 32 ? 5 : 4);
}
// L = 128: T alone takes 40 KB of LDS, so the per-wave input buffers (8 KB
// per 16 tiles) cap the workgroup at 11 waves (G = 16) or 6 (G = 32)
constexpr int bs_wg_waves_f16(int L, int G)
{
    if (L == 128) return G == 16 ? 11 : 6;
    return G == 64 ? (L == 32 ? 12 : 8) : G == 32 ? (L == 32 ? 8 : 12) : (L == 32 ? 10 : 16);
}
// LDS of a workgroup of w waves: T's nonzero blocks and Gam in `split`
// planes, 6 Phi^(2^j) blocks, and per wave its input buffer (G L floats) and
// 16 lookback slots of 16 floats
constexpr int bs_lds_bytes(int L, int G, int split, int w)
{
    return 1024 * (split * (bs_tidx(L / 16) + L / 32) + 6) + w * (4 * G * L + 1024);
}
// the FP16 count, at most the waves per SIMD times 4, reduced until the
// workgroup fits the 160 KB of LDS
constexpr int bs_wg_waves(int L, int G, int split = kBsSplitF16)
{
    int w = bs_wg_waves_f16(L, G);
    if (w > 4 * bs_waves(L, G, split)) w = 4 * bs_waves(L, G, split);
    while (w > 1 && bs_lds_bytes(L, G, split, w) > 160 * 1024) --w;
    return w;
}

// BF16 bits of a float, rounded to nearest even (finite values), and the
// value of BF16 bits; the same integer steps on the host and the device
__host__ __device__ inline uint16_t bf16_bits_rne(float f)
{
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline double bf16_value(uint16_t b) { return (double)__builtin_bit_cast(float, (uint32_t)b << 16); }
// component hl (0 hi, 1 mid, 2 lo) of the BF16 three-way split of v: each
// the remainder so far (exact in FP64) rounded to float32, then to BF16
__host__ __device__ inline uint16_t bf16x3_bits(double v, int hl)
{
    uint16_t b = 0;
    for (int k = 0; k <= hl; ++k) {
        b = bf16_bits_rne((float)v);
        v -= bf16_value(b);
    }
    return b;
}
// scale slots: 0 T, 1 Gam, then the levels of the in-chunk tile scan
// Z_k = sum_i Phi^i V_(k-i) (level j adds the lags 2^j .. 2^(j+1) - 1):
// kBsScaleLevels = levels run (the lags from 2^levels on weigh at most 2^-36
// in total, sum ||Phi^i||_inf up to the chunk transition, the bound the
// lookback folds to), kBsScaleFp32Level = the first level run on the FP32
// matrix core instead of the FP64 one (from it on every Phi^(2^j) has
// ||.||_inf <= 2^-4: each term of its product is at most 2^-4 of the largest
// state, so the product's float32 rounding over 16 terms stays within one
// float32 rounding of that state, which the states get anyway when they are
// rounded for the output stage).  The chunk-carry injection Phi^(k+1) X_in
// follows the same levels.  Measured at 2^25, Wn = 0.1 and 0.02: max and RMS
// errors unchanged to 4 digits against a 2^-10 threshold.
constexpr int kBsScaleFp32Level = 2;
constexpr int kBsScaleLevels = 3;
constexpr double kBsFp32Norm = 0.0625;  // 2^-4
constexpr int bs_scan_levels(int G) { return G == 16 ? 4 : G == 32 ? 5 : 6; }

// ---------------------------------------------------------------------------
// Ragged batch (batch.hip): channel c holds samples [offset, offset + length)
// of the packed input and output, chunks first_chunk .. first_chunk +
// ceil(length / chunk) - 1 of the global chunk list (PH channels' chunks
// first, then the block-state channels'; the chunk size is that of the
// channel's form), its tables at coef + coef_off (PH: ns sections,
// coef_stride floats each; block-state: bs_layout(64)).  Lookback records are
// 32 words per chunk: kBatchMaxSections PH sections of 4 words, or the 16
// partial + 16 full state words of the block-state form.
// ---------------------------------------------------------------------------
constexpr int kBatchMaxSections = 8;

struct BatchChannel {
    int64_t offset;
    int64_t length;
    int32_t ns;
    int32_t coef_off;
    uint32_t first_chunk;
    int32_t algo;  // kBatchPH or kBatchBS
};
constexpr int32_t kBatchPH = 0;
constexpr int32_t kBatchBS = 1;  // block-state form, tiles of 64 samples

struct BatchArgs {
    const float *in;
    float *out;
    const float *coef;
    uint64_t *lb;
    const BatchChannel *chan;
    const uint32_t *chunk_chan;  // channel of every chunk
    unsigned long long *ticket;
    uint32_t *epoch;
    uint32_t *err;
    uint32_t *arrive;
    uint32_t num_chunks;
    uint32_t spin_limit;
    int32_t persistent;
    uint32_t grid;
    int32_t coef_stride;
    uint32_t chunk_base;  // this launch's chunks: chunk_base .. chunk_base + num_chunks - 1
    int32_t vec16;        // in and out 16-byte aligned: full chunks move with 16-byte accesses
};

// algo kBatchPH: n = samples per lane; kBatchBS: n = tiles per chunk
hipError_t launch_batch(int algo, int n, const BatchArgs &a, hipStream_t s);
const void *kernel_batch(int algo, int n);

hipError_t launch_bs(int L, int G, int ns, const LaunchArgs &a, hipStream_t s, int split = kBsSplitF16);
// end state of a PH, STCR or DTCR launch whose length is a multiple of the
// chunk, in the layout of LaunchArgs::zi, from the last chunk's full records
// and the last two input samples (a partial last chunk records it itself)
hipError_t launch_ph_end_state(const LaunchArgs &a, int chunk_samples, float *zf, hipStream_t s);
// end state of a PLR launch: the direct-form history x[n-K..n-1], y[n-K..n-1]
hipError_t launch_plr_end_state(const LaunchArgs &a, float *zf, hipStream_t s);

// Streaming state of the block-state form (bs.hip): the realization's states
// are the sections' direct-form-II-transposed (w1, w2) divided by d; c holds
// b1, b2, a1, a2 (scipy sign) per section.
struct BsStateMap {
    double d[16];
    double c[8][4];
    int32_t ns, pad;
};
// the DF1 histories zi (2 + 2 ns floats) as the balanced state z0 (16 floats)
hipError_t launch_bs_state_in(const float *zi, const BsStateMap *m, float *z0, hipStream_t s);
// the DF1 histories after sample n-1 from the state entering the last tile of
// L samples (balanced, 16 floats) and that tile's inputs, re-filtered in FP64
hipError_t launch_bs_state_out(const float *in, int64_t n, int L, const float *tail, const BsStateMap *m, float *zf,
                               hipStream_t s);

// one block-state table built on the GPU in FP64 (bs_setup.hip): the
// sections of one cascade and where its table goes
struct BsSetupJob {
    float sos[48];  // up to 8 rows of 6
    int32_t ns;
    int32_t pad;
    int64_t out_off;  // floats from the output base
};
hipError_t launch_bs_setup(int L, int G, const BsSetupJob *jobs, int n_jobs, float *out, hipStream_t s,
                           int split = kBsSplitF16);
const void *kernel_bs(int L, int G, int ns, int split = kBsSplitF16);

}  // namespace iirx
