// bs_kernel.hpp -- the single-signal block-state kernel (its algorithm is
// described at the top of bs.hip), instantiated for the FP16 hi/lo split in
// bs.hip and for the BF16 hi/mid/lo split in bs_bf16.hip.
#pragma once
#include "bs_core.hpp"

namespace iirx {


// operands of the single-signal kernel: T, Gam and Phi^(2^j) from the
// workgroup's LDS copy, the post-lookback ones from global memory
template <int L, int SPLIT>
struct BsLdsOps {
    static constexpr BsLayout Y = bs_layout(L, SPLIT);
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

template <int L, int G, int NSL, int SPLIT>
__global__ __launch_bounds__(64 * bs_wg_waves(L, G, SPLIT))
__attribute__((amdgpu_waves_per_eu(bs_waves(L, G, SPLIT), bs_waves(L, G, SPLIT)))) void bs_kernel(LaunchArgs a)
{
    constexpr int MB = L / 16;  // output row blocks of T
    constexpr int KQ = L / 32;  // K slabs of 32 samples
    constexpr int NB = G / 16;  // column blocks of 16 tiles
    constexpr int CH = L * G;
    constexpr int NLD = NB * KQ * 2;  // 16-byte loads per lane and chunk
    constexpr int W = bs_wg_waves(L, G, SPLIT);
    constexpr int NT = bs_tidx(MB);   // nonzero blocks of T
    constexpr BsLayout Y = bs_layout(L, SPLIT);
    static_assert(bs_lds_bytes(L, G, SPLIT, W) <= 160 * 1024, "LDS");
    // shared by the workgroup: the operands consumed before the lookback.  A
    // VMEM table load issued after the next chunk's fill would wait for the
    // whole fill (vmcnt retires in issue order); LDS reads do not.
    __shared__ uint4 lT[SPLIT * NT * 64];  // T's split planes, nonzero blocks (mb, q <= bs_tq(mb))
    __shared__ uint4 lG[SPLIT * KQ * 64];  // Gam's split planes
    __shared__ f4 lP[6 * 64];          // Phi^(2^j), FP32 operand blocks
    // per wave: one chunk of X (load i of lane l at [64 i + l]) and the lookback slots
    __shared__ f4 xbuf_w[W][NLD * 64];
    __shared__ float lp_w[W][kBsSlots * 16];
    __shared__ unsigned long long s_ticket;  // ticket schedule: this workgroup's ticket

    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, col = lane & 15, g = lane >> 4;
    const float *tab = a.coef;
    const cfloat *ctab = as_const(a.coef);
    const int sT = (int)ctab[Y.kScale + 0], sG = (int)ctab[Y.kScale + 1];
    const int depth = (int)ctab[Y.kDepth];
    const int lv32 = (int)ctab[Y.kScale + kBsScaleFp32Level], nlv = (int)ctab[Y.kScale + kBsScaleLevels];
    f4 *xbuf = xbuf_w[wave];
    float *lp = lp_w[wave];

    for (int i = threadIdx.x; i < SPLIT * NT * 64; i += 64 * W) {
        const int t = (i >> 6) % NT, hl = (i >> 6) / NT;
        int mb = 0;
        while (bs_tidx(mb + 1) <= t) ++mb;
        lT[i] = reinterpret_cast<const uint4 *>(tab + bs_T(Y, hl, mb, t - bs_tidx(mb)))[i & 63];
    }
    for (int i = threadIdx.x; i < SPLIT * KQ * 64; i += 64 * W)
        lG[i] = reinterpret_cast<const uint4 *>(tab + bs_G(Y, (i >> 6) / KQ, (i >> 6) % KQ))[i & 63];
    for (int i = threadIdx.x; i < 6 * 64; i += 64 * W)
        lP[i] = reinterpret_cast<const f4 *>(tab + bs_P(Y, i >> 6))[i & 63];
    __syncthreads();

    // X as x32 B operands: tile 16 nb + col, K slab q, K rows 8 g + j =
    // samples 32 q + 4 g + j (a) and 32 q + 16 + 4 g + j - 4 (b).  The chunk
    // is fetched HBM -> LDS by global_load_lds (no registers held while it
    // is in flight); read_chunk waits for it.  (Fetching whole 128-byte lines
    // per instruction, the two halves of a tile's line on lanes col and col +
    // 8, measured no faster at orders 2 and 8 and 25% slower at order 16.)
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
    // STAGED (bs_staged geometries): the chunk moves through the wave's LDS
    // buffer with contiguous 1 KB accesses both ways.  Tile t's 16-byte piece p
    // (samples 4 p .. 4 p + 3) sits at slot 16 t + ((p + t) & 15), a rotation
    // by the tile index, so that the B-operand reads (piece p of 16 tiles
    // across a row of lanes), the output writes and the contiguous DMA fills
    // and read-backs all touch distinct 16-byte bank groups (32 or 64 banks).
    // The fill of instruction i takes tiles 4 i .. 4 i + 3: lane j fetches the
    // piece that belongs at slot 64 i + j.
    constexpr bool STAGED = bs_staged(L, G, SPLIT);
    static_assert(!STAGED || L == 64, "the staged layout rotates the 16 pieces of a 64-sample tile");
    constexpr int NST = CH / 256;  // 1 KB instructions per chunk
    // the fill / store offset of lane j in instruction i: tile 4 i + j / 16,
    // piece ((j & 15) - t) & 15, which repeats with i mod 4 (floats, without
    // the 256 i of the instruction)
    int goff[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int t = 4 * k + (lane >> 4);
        goff[k] = (lane >> 4) * L + 4 * (((lane & 15) - t) & 15);
    }
    auto fetch_staged = [&](uint32_t chunk) {
        const float *src = a.in + (long)chunk * CH;
#pragma unroll
        for (int i = 0; i < NST; ++i)
            __builtin_amdgcn_global_load_lds((__attribute__((address_space(1))) void *)(src + 256 * i + goff[i & 3]),
                                             (__attribute__((address_space(3))) void *)(&xbuf[i * 64]), 16, 0, 0);
    };
    // vmcnt counts loads and stores together and retires them in issue
    // order, so after the first chunk the fill has landed once at most the
    // MB * NB output stores issued after it are outstanding (STAGED: NST / 2)
    auto read_staged = [&](Raw &r, bool first) {
        if (first)
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        else
            asm volatile("s_waitcnt vmcnt(%0)" ::"n"(NST / 2) : "memory");
        // tile 16 nb + col: slot 256 nb + 16 col + ((piece + col) & 15)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
#pragma unroll
            for (int q = 0; q < KQ; ++q) {
                r.a[nb][q] = xbuf[256 * nb + 16 * col + ((8 * q + g + col) & 15)];
                r.b[nb][q] = xbuf[256 * nb + 16 * col + ((8 * q + 4 + g + col) & 15)];
            }
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");  // reads done before the buffer is reused
    };
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

    const BsLdsOps<L, SPLIT> ops{lT, lG, lP, tab, lane};
    // STAGED: the output blocks into the wave's LDS buffer (free once the
    // chunk's input is in registers), then read back 1 KB per instruction;
    // the next chunk's fill is issued between the read-back and the stores,
    // so that the next read's vmcnt wait covers the fill with the last NST / 2
    // stores still in flight
    auto process_staged = [&](uint32_t chunk, uint32_t part_flag, uint32_t full_flag, const Raw &xr, uint32_t next) {
        auto to_lds = [&](int nb, int mb, const f4 &y) { xbuf[256 * nb + 16 * col + ((4 * mb + g + col) & 15)] = y; };
        bs_chunk<L, G, NSL, SPLIT, false>(xr, chunk, part_flag, full_flag, a.lb, tab + Y.kMc, depth, lp,
                                                      a.spin_limit, a.err, 0u, sT, sG, lv32, nlv, ops, to_lds, a.zi,
                                                      chunk + 1u == a.num_chunks ? a.tail : nullptr);
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        // read back and store in two halves (fewer registers); the fill of the
        // next chunk goes out after the last read-back, before the last half
        float *out = a.out + (long)chunk * CH;
        constexpr int H = NST / 2;
        f4 o[H];
#pragma unroll
        for (int i = 0; i < H; ++i) o[i] = xbuf[i * 64 + lane];
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
#pragma unroll
        for (int i = 0; i < H; ++i) *reinterpret_cast<f4 *>(out + 256 * i + goff[i & 3]) = o[i];
#pragma unroll
        for (int i = 0; i < H; ++i) o[i] = xbuf[(H + i) * 64 + lane];
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        if (next < a.num_chunks) fetch_staged(next);
#pragma unroll
        for (int i = 0; i < H; ++i) *reinterpret_cast<f4 *>(out + 256 * (H + i) + goff[i & 3]) = o[i];
    };
    auto process = [&](uint32_t chunk, uint32_t part_flag, uint32_t full_flag, const Raw &xr) {
        float *out = a.out + (long)chunk * CH;
        // whole 128-byte lines per store (bs_chunk's LINES form), through the L2
        auto store = [&](int tile, int row, const f4 &y) {
            *reinterpret_cast<f4 *>(out + (long)tile * L + row) = y;
        };
        // streaming: a.zi is the balanced state before the signal, a.tail
        // receives the state entering the last chunk's last tile
        bs_chunk<L, G, NSL, SPLIT, true>(xr, chunk, part_flag, full_flag, a.lb, tab + Y.kMc, depth, lp,
                                                      a.spin_limit, a.err, 0u, sT, sG, lv32, nlv, ops, store, a.zi,
                                                      chunk + 1u == a.num_chunks ? a.tail : nullptr);
    };

    const uint32_t GW = gridDim.x * W, gw = blockIdx.x * W + wave;  // this wave among the grid's
    if (a.persistent) {
        // chunks gw + i * GW; the next chunk's input streams into LDS while
        // this one is scanned, looked back and written
        const uint32_t epoch = epoch_load(a.epoch);
        epoch_arrive(a.arrive, epoch);
        const uint32_t pf = 2u * epoch + 1u, ff = pf + 1u;
        uint32_t c = gw;
        if constexpr (STAGED) {
            if (c < a.num_chunks) fetch_staged(c);
            for (bool first = true; c < a.num_chunks; c += GW, first = false) {
                Raw cur;
                read_staged(cur, first);
                process_staged(c, pf, ff, cur, c + GW);
            }
        } else {
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
        }
        // the owner of the last chunk advances the epoch after publishing it
        if ((a.num_chunks - 1u) % GW == gw) epoch_advance(a.epoch, a.arrive, epoch, GW, a.spin_limit, a.err);
    } else if (a.num_chunks % W == 0) {
        // whole groups of W chunks: workgroups (one per resident slot) loop
        // over the groups, one atomic ticket each, wave w taking chunk W t + w
        // (for_each_chunk's grouped schedule: a group's chunks wait only on
        // lower tickets or on lower waves of the group, which meet at no
        // barrier before the group ends); the tables stay in LDS across groups
        // (2^25 samples: 506 / 406 GS/s at orders 8 / 16 against 441 / 350
        // with one group per workgroup)
        for_each_chunk<W>(a.ticket, a.epoch, a.arrive, a.num_chunks, 0, a.spin_limit, a.err,
                          [&](uint32_t chunk, uint32_t part_flag, uint32_t full_flag) {
            Raw cur;
            if constexpr (STAGED) {
                fetch_staged(chunk);
                read_staged(cur, true);
                process_staged(chunk, part_flag, full_flag, cur, a.num_chunks);
            } else {
                fetch_chunk(chunk);
                read_chunk(cur, true);
                process(chunk, part_flag, full_flag, cur);
            }
        });
    } else {
        // one ticket per workgroup (grid = ceil(chunks / W) tickets per
        // launch): wave w takes chunk W t + w.  A chunk waits only on chunks
        // of lower tickets (workgroups dispatched earlier: resident or done)
        // or of lower waves of its own workgroup, which meets no barrier
        // after the ticket, so no co-residency is assumed.  (One ticket per
        // wave serialised 8192 device-scope atomics per launch at 2^25
        // samples: 234 / 213 GS/s at orders 8 / 16.)
        if (threadIdx.x == 0) s_ticket = atomicAdd(a.ticket, 1ull);
        __syncthreads();
        const unsigned long long tk = s_ticket;
        const unsigned long long ep = tk / gridDim.x;
        const uint32_t chunk = (uint32_t)(tk - ep * gridDim.x) * W + wave;
        const uint32_t pf = (uint32_t)(2ull * ep + 1ull);
        if (chunk < a.num_chunks) {
            Raw cur;
            if constexpr (STAGED) {
                fetch_staged(chunk);
                read_staged(cur, true);
                process_staged(chunk, pf, pf + 1u, cur, a.num_chunks);
            } else {
                fetch_chunk(chunk);
                read_chunk(cur, true);
                process(chunk, pf, pf + 1u, cur);
            }
        }
    }
}


template <int L, int G, int NSL, int SPLIT>
static hipError_t launch_lgn(const LaunchArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL((bs_kernel<L, G, NSL, SPLIT>), dim3(a.grid), dim3(64 * bs_wg_waves(L, G, SPLIT)), 0, s, a);
    return hipGetLastError();
}

template <int L, int G, int SPLIT>
static hipError_t launch_lg(int nsl, const LaunchArgs &a, hipStream_t s)
{
    switch (nsl) {
    case 1: return launch_lgn<L, G, 1, SPLIT>(a, s);
    case 2: return launch_lgn<L, G, 2, SPLIT>(a, s);
    case 3: return launch_lgn<L, G, 3, SPLIT>(a, s);
    case 4: return launch_lgn<L, G, 4, SPLIT>(a, s);
    default: return hipErrorInvalidValue;
    }
}

template <int L, int G, int SPLIT>
static const void *fn_lg(int nsl)
{
    switch (nsl) {
    case 1: return (const void *)bs_kernel<L, G, 1, SPLIT>;
    case 2: return (const void *)bs_kernel<L, G, 2, SPLIT>;
    case 3: return (const void *)bs_kernel<L, G, 3, SPLIT>;
    case 4: return (const void *)bs_kernel<L, G, 4, SPLIT>;
    default: return nullptr;
    }
}

template <int L, int SPLIT>
static hipError_t launch_l(int G, int nsl, const LaunchArgs &a, hipStream_t s)
{
    switch (G) {
    case 16: return launch_lg<L, 16, SPLIT>(nsl, a, s);
    case 32: return launch_lg<L, 32, SPLIT>(nsl, a, s);
    case 64:
        if constexpr (L < 128 && SPLIT == kBsSplitF16) return launch_lg<L, 64, SPLIT>(nsl, a, s);
        else return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
    }
}

template <int L, int SPLIT>
static const void *fn_l(int G, int nsl)
{
    switch (G) {
    case 16: return fn_lg<L, 16, SPLIT>(nsl);
    case 32: return fn_lg<L, 32, SPLIT>(nsl);
    case 64:
        if constexpr (L < 128 && SPLIT == kBsSplitF16) return fn_lg<L, 64, SPLIT>(nsl);
        else return nullptr;
    default: return nullptr;
    }
}

// the BF16 hi/mid/lo instances (bs_bf16.hip): L = 64 or 128
hipError_t launch_bs_bf16(int L, int G, int nsl, const LaunchArgs &a, hipStream_t s);
const void *kernel_bs_bf16(int L, int G, int nsl);

}  // namespace iirx
