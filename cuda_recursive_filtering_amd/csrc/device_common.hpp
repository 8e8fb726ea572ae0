// device_common.hpp -- wave64 building blocks shared by the CDNA4 kernels:
// the monotone workgroup ticket, the self-validating lookback records, the
// 16-byte coalesced global<->LDS staging and lane primitives.
//
// Lookback protocol (replaces the launch-versioned status/partcarry/fullcarry
// triple of ph_kernels.cuh:16-29, 384-469).
//   * Persistent schedule (default): the grid holds as many workgroups as the
//     device keeps resident at once; workgroup b processes chunks b, b+G,
//     b+2G, ...  A chunk waits only on lower chunks, all owned by resident
//     workgroups, without a per-chunk atomic.  The launch epoch is a device
//     word read by every workgroup at its start and advanced by the
//     workgroup that finishes the last chunk (every other workgroup has read
//     it by then: the last chunk's lookback observed its predecessor's
//     record, which observed its own predecessor's, and so on).  A bounded
//     spin turns a violated co-residency assumption into an error flag.
//   * Ticket schedule: one 64-bit counter per plan, incremented once per
//     workgroup (PH and STCR: once per group of chunks, for_each_chunk).
//     ticket = epoch * num_chunks + chunk.  Workgroups of one launch draw
//     tickets in dispatch order, so a chunk only ever waits on chunks that
//     already hold a ticket, i.e. are resident or finished (AMD hardware
//     gives no forward-progress guarantee to workgroups not yet dispatched;
//     the reference relies on the same ordering, ph_kernels.cuh:227-229).
//     The epoch replaces the caller-supplied `launch` index, so the kernel
//     arguments do not change between launches and a captured graph replays
//     correctly.
//   * Record: every carry float travels in its own 64-bit word
//     (flag << 32 | float bits).  A reader accepts a value only if the flag
//     of that very word is current, so no fence is needed between payload
//     and flag and a torn view of a multi-word carry is detected per word.
//     Flags are 2*epoch+1 (partial: chunk result with zero incoming state)
//     and 2*epoch+2 (full: true end state), in separate words, so anything
//     left by an earlier launch reads as "not ready"; the buffer is zeroed
//     once at plan creation.
//   * Stores and loads are agent-scope atomics: the eight XCDs of MI355X do
//     not share an L2, and agent-scope atomics are kept coherent across them.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

namespace iirx {

constexpr int kWave = 64;

typedef float f4 __attribute__((ext_vector_type(4)));

// Coefficient tables are read through the constant address space so that
// uniform indices compile to scalar loads (s_load) into SGPRs; through a
// generic pointer the compiler cannot prove the table unclobbered by the
// lookback's atomics and falls back to per-lane vector loads.
typedef const __attribute__((address_space(4))) float cfloat;

__device__ __forceinline__ const cfloat *as_const(const float *p) { return (const cfloat *)p; }

__device__ __forceinline__ float as_float(uint32_t u) { return __uint_as_float(u); }

__device__ __forceinline__ uint32_t readfirstlane_u32(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

__device__ __forceinline__ float readlane_f(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

__device__ __forceinline__ int readlane_i(int v, int lane)
{
    return __builtin_amdgcn_readlane(v, lane);
}

// value of lane (l - d) for l >= d, own value otherwise (CUDA shfl_up rule)
__device__ __forceinline__ float lane_up(float v, int d) { return __shfl_up(v, (unsigned)d, kWave); }

// value of lane l - D inside the same row of 16 lanes (DPP row_shr:D), 0 for
// the first D lanes of every row; D in 1..15
template <int D>
__device__ __forceinline__ float row_shr_zero(float v)
{
    static_assert(D >= 1 && D <= 15, "row_shr takes 1..15");
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x110 + D, 0xF, 0xF, false));
}

// sum of v over the wave, uniform: the rotations sum each row of 16 lanes,
// row_bcast:15 / row_bcast:31 carry rows 0 and 2 into rows 1 and 3, then
// rows 0-1 into rows 2-3, so that lane 63 holds the total
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_or_zero(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xF, false));
}

__device__ __forceinline__ float wave_sum(float v)
{
    v += dpp_or_zero<0x121, 0xF>(v);  // row_ror:1
    v += dpp_or_zero<0x122, 0xF>(v);  // row_ror:2
    v += dpp_or_zero<0x124, 0xF>(v);  // row_ror:4
    v += dpp_or_zero<0x128, 0xF>(v);  // row_ror:8
    v += dpp_or_zero<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
    v += dpp_or_zero<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
    return readlane_f(v, 63);
}

// ------------------------------------------------------------------------
// compile-time loops: every index a literal (the "hand-unrolled" forms)
// ------------------------------------------------------------------------
template <int I>
using ic = std::integral_constant<int, I>;

// f(ic<I>) for I = B, B+S, ... < E, every index a compile-time literal
template <int B, int E, int S, typename F>
__device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (B < E) {
        f(ic<B>{});
        static_for<B + S, E, S>(f);
    }
}

// Inclusive scan over the wave's 64 lanes of 2-vectors s with one uniform
// 2x2 operator C (row-major): lane l ends with sum_{i<=l} C^(l-i) s_i.
// Cpow: C^(2^k), k = 0..4 (uniform, 4 floats each); Cfin: C^(t+1) for
// t = 0..15 on 16-byte boundaries (per lane).  Inside each row of 16 lanes a
// Kogge-Stone scan over DPP row shifts (no LDS round trip), then the rows'
// carries K_r = S(16 r - 1) + C^16 K_(r-1), from lanes 15, 31 and 47, enter
// lane j of row r as C^(j+1) K_r.  The PH lane scan (ph_kernels.cuh:289-382)
// with the same operators in another order.
__device__ __forceinline__ void ph_lane_scan(const cfloat *Cpow, const float *Cfin, int lane, float &s2, float &s1)
{
    const float4 cj = *reinterpret_cast<const float4 *>(Cfin + 4 * (lane & 15));
    static_for<0, 4, 1>([&](auto K) {
        constexpr int k = decltype(K)::value;
        const auto *P = Cpow + 4 * k;
        const float t2 = row_shr_zero<1 << k>(s2), t1 = row_shr_zero<1 << k>(s1);
        s2 = fmaf(P[1], t1, fmaf(P[0], t2, s2));
        s1 = fmaf(P[3], t1, fmaf(P[2], t2, s1));
    });
    const auto *P = Cpow + 16;  // C^16
    const float a2 = readlane_f(s2, 15), a1 = readlane_f(s1, 15);
    const float b2 = fmaf(P[1], a1, fmaf(P[0], a2, readlane_f(s2, 31)));
    const float b1 = fmaf(P[3], a1, fmaf(P[2], a2, readlane_f(s1, 31)));
    const float e2 = fmaf(P[1], b1, fmaf(P[0], b2, readlane_f(s2, 47)));
    const float e1 = fmaf(P[3], b1, fmaf(P[2], b2, readlane_f(s1, 47)));
    const int row = lane >> 4;
    const float k2 = row == 1 ? a2 : row == 2 ? b2 : row == 3 ? e2 : 0.0f;
    const float k1 = row == 1 ? a1 : row == 2 ? b1 : row == 3 ? e1 : 0.0f;
    s2 = fmaf(cj.y, k1, fmaf(cj.x, k2, s2));
    s1 = fmaf(cj.w, k1, fmaf(cj.z, k2, s1));
}

// descending: f(ic<I>) for I = B, B-1, ... > E
template <int B, int E, typename F>
__device__ __forceinline__ void static_for_down(F &&f)
{
    if constexpr (B > E) {
        f(ic<B>{});
        static_for_down<B - 1, E>(f);
    }
}

// ------------------------------------------------------------------------
// ticket
// ------------------------------------------------------------------------
struct Ticket {
    uint32_t chunk;
    uint32_t part_flag;
    uint32_t full_flag;
};

__device__ __forceinline__ uint32_t epoch_load(uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void epoch_store(uint32_t *p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Persistent schedule: every wave announces, after reading the launch
// epoch, that it has read it; the owner of the last chunk advances the epoch
// only once every wave of the grid has arrived.  Without the arrival count a
// workgroup that starts late (another XCD's dispatcher behind) could read the
// next launch's epoch and wait on flags nobody writes.  The count is spread
// over kArriveSlots words, 128 bytes apart, so that the adds do not queue on
// one address.
constexpr int kArriveSlots = 64;
constexpr int kArriveStride = 32;  // words between slots

__device__ __forceinline__ void epoch_arrive(uint32_t *arrive, uint32_t epoch)
{
    // the epoch value has returned before the add is issued (the asm reads
    // it, so the load is waited for here, and memory operations stay behind
    // it); a release would write back the whole L2 on every wave
    asm volatile("" ::"v"(epoch) : "memory");
    if ((threadIdx.x & (kWave - 1)) == 0)
        __hip_atomic_fetch_add(arrive + kArriveStride * (blockIdx.x % kArriveSlots), 1u, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
}

// called by one whole wave of the owner of the last chunk
__device__ __forceinline__ void epoch_advance(uint32_t *epoch_word, uint32_t *arrive, uint32_t epoch, uint32_t waves,
                                              uint32_t spin_limit, uint32_t *err)
{
    const int lane = threadIdx.x & (kWave - 1);
    uint32_t *slot = arrive + kArriveStride * lane;
    uint32_t spins = 0;
    for (;;) {
        uint32_t n = (lane < kArriveSlots) ? __hip_atomic_load(slot, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) : 0u;
#pragma unroll
        for (int o = kWave / 2; o >= 1; o >>= 1) n += __shfl_xor(n, o, kWave);
        if (n >= waves) break;
        if (++spins > spin_limit) {
            if (lane == 0) __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    if (lane < kArriveSlots) __hip_atomic_store(slot, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (lane == 0) epoch_store(epoch_word, epoch + 1u);
}

__device__ __forceinline__ Ticket take_ticket(unsigned long long *ctr, uint32_t num_chunks)
{
    __shared__ unsigned long long s_tk;
    if (threadIdx.x == 0) s_tk = atomicAdd(ctr, 1ull);
    __syncthreads();
    const unsigned long long tk = s_tk;
    const unsigned long long epoch = tk / num_chunks;
    Ticket t;
    t.chunk = (uint32_t)(tk - epoch * num_chunks);
    t.part_flag = (uint32_t)(2ull * epoch + 1ull);
    t.full_flag = t.part_flag + 1u;
    return t;
}

// LDS written by this wave is visible to all of its lanes: the wave's LDS
// operations complete in order, so waiting for its own outstanding ones
// suffices (the kernels with several chunks per workgroup use it where a
// one-wave workgroup uses a barrier)
__device__ __forceinline__ void wave_lds_sync()
{
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    __builtin_amdgcn_wave_barrier();
}

template <int WV>
__device__ __forceinline__ void chunk_sync()
{
    if constexpr (WV == 1) __syncthreads();
    else wave_lds_sync();
}

// ------------------------------------------------------------------------
// lookback records
// ------------------------------------------------------------------------
__device__ __forceinline__ void rec_store(uint64_t *p, uint32_t flag, float v)
{
    const uint64_t w = ((uint64_t)flag << 32) | (uint64_t)__float_as_uint(v);
    __hip_atomic_store(p, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t rec_load(uint64_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t rec_flag(uint64_t w) { return (uint32_t)(w >> 32); }
__device__ __forceinline__ float rec_val(uint64_t w) { return __uint_as_float((uint32_t)w); }

// Two-float carry (state (y[-2], y[-1]) of a biquad), records of 4 words:
// [partial y2, partial y1, full y2, full y1].
struct Carry2 {
    float y2, y1;
};

__device__ __forceinline__ uint64_t *rec2(uint64_t *lb, uint32_t chunk, int sec, int ns)
{
    return lb + ((size_t)chunk * (size_t)ns + (size_t)sec) * 4u;
}

// Publish from one lane.
__device__ __forceinline__ void publish2(uint64_t *r, bool full, uint32_t flag, float y2, float y1)
{
    const int o = full ? 2 : 0;
    rec_store(r + o, flag, y2);
    rec_store(r + o + 1, flag, y1);
}

// End state (y2, y1) of section `sec` before chunk `chunk`, for a launch that
// runs after the one that filtered the chunks before it (same stream): the
// full record of chunk - 1, or the streamed-in state (z2, z1) at chunk 0.
__device__ __forceinline__ Carry2 prior_state2(uint64_t *lb, uint32_t chunk, int sec, int ns, float z2, float z1)
{
    if (chunk == 0) return Carry2{z2, z1};
    uint64_t *r = rec2(lb, chunk - 1u, sec, ns);
    return Carry2{rec_val(__hip_atomic_load(r + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)),
                  rec_val(__hip_atomic_load(r + 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))};
}

// Decoupled lookback over a 64-chunk window, one predecessor per lane.
// Returns the end state of chunk `chunk - 1`; the chain starts at chunk
// `first` from the state (s2, s1) (zero unless the caller streams a state in).
// M is the chunk transition: X_c = M X_{c-1} + partial_c (row-major 2x2).
//
// `depth` bounds the fold: the plan sets it to the smallest d with
// ||M^d||_inf <= 2^-36, so the end state of chunk c-1-d contributes less
// than 2^-36 of its own magnitude to X and the partials of the d nearest
// predecessors determine X far below float32 rounding.  The wave therefore
// never waits for a predecessor to become full: it folds `depth` partials
// from a zero state, or starts from the nearest full record if one is closer.
// depth == 0 means unbounded (poles too close to the unit circle for this
// chunk length): the wave spins until every predecessor up to the nearest
// full one has published, as the reference does (ph_kernels.cuh:397-404).
// The fold is Horner's rule, nearest predecessor last (ph:436-450).
// Watchdog: after `spin_limit` unsuccessful polls the wave raises *err and
// proceeds with a zero carry, so a broken co-residency assumption ends the
// launch with an error instead of hanging the device.
__device__ __forceinline__ Carry2 lookback2(uint64_t *lb, uint32_t chunk, int sec, int ns,
                                            uint32_t part_flag, uint32_t full_flag,
                                            float m0, float m1, float m2, float m3, int depth,
                                            uint32_t spin_limit, uint32_t *err, uint32_t first = 0,
                                            float s2 = 0.0f, float s1 = 0.0f)
{
    const int lane = threadIdx.x & (kWave - 1);
    const long pred = (long)chunk - 1 - lane;
    const int window = (depth > 0 && depth < kWave) ? depth : kWave;
    const bool active = lane < window;
    const uint64_t below_depth = (depth >= kWave) ? ~0ull : ((1ull << depth) - 1ull);
    float v2 = 0.0f, v1 = 0.0f;
    // 0 not ready, 1 partial, 2 full, 3 before the signal start, 4 outside the window
    int st = !active ? 4 : (pred < (long)first ? 3 : 0);
    if (st == 3) {  // the state before the chain start, a full record
        v2 = s2;
        v1 = s1;
    }
    int D = 0;
    bool from_full = false;
    uint32_t spins = 0;
    for (;;) {
        if (st < 2) {
            // the whole record in one round trip: partial words 0-1, full 2-3
            uint64_t *r = rec2(lb, (uint32_t)pred, sec, ns);
            const uint64_t q0 = rec_load(r), q1 = rec_load(r + 1), f0 = rec_load(r + 2), f1 = rec_load(r + 3);
            if (rec_flag(f0) == full_flag && rec_flag(f1) == full_flag) {
                st = 2;
                v2 = rec_val(f0);
                v1 = rec_val(f1);
            } else if (st == 0) {
                if (rec_flag(q0) == part_flag && rec_flag(q1) == part_flag) {
                    st = 1;
                    v2 = rec_val(q0);
                    v1 = rec_val(q1);
                }
            }
        }
        const uint64_t zmask = __ballot(st == 0);
        const uint64_t fmask = __ballot(st == 2 || st == 3);
        if (fmask) {
            const int f = __builtin_ctzll(fmask);
            const uint64_t upto = (f == kWave - 1) ? ~0ull : ((1ull << (f + 1)) - 1ull);
            if ((zmask & upto) == 0) {
                D = f;
                from_full = true;
                break;
            }
        }
        if (depth > 0 && (zmask & below_depth) == 0) {
            D = depth;
            break;
        }
        if (++spins > spin_limit) {
            if (lane == 0) __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            D = 0;
            from_full = false;
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    Carry2 X;
    X.y2 = from_full ? readlane_f(v2, D) : 0.0f;
    X.y1 = from_full ? readlane_f(v1, D) : 0.0f;
    for (int d = D - 1; d >= 0; --d) {
        const float p2 = readlane_f(v2, d), p1 = readlane_f(v1, d);
        const float n2 = fmaf(m1, X.y1, fmaf(m0, X.y2, p2));
        const float n1 = fmaf(m3, X.y1, fmaf(m2, X.y2, p1));
        X.y2 = n2;
        X.y1 = n1;
    }
    return X;
}

// lookback2 for a pair of sections (sec, sec + 1), whose records lie side by
// side: one predecessor per lane, its 8 words in one round trip, used whole
// (both sections full, or both partial).  The carry is the pair's joint state
// (y_s2, y_s1, y_t2, y_t1); M its chunk transition (4 x 4 row-major, block
// lower triangular: y_s does not see y_t; uniform); z the state before the
// chain start.
struct Carry4 {
    float v[4];
};

__device__ __forceinline__ Carry4 lookback_pair(uint64_t *lb, uint32_t chunk, int sec, int ns, uint32_t part_flag,
                                               uint32_t full_flag, const cfloat *M, int depth, uint32_t spin_limit,
                                               uint32_t *err, const float *z)
{
    const int lane = threadIdx.x & (kWave - 1);
    const long pred = (long)chunk - 1 - lane;
    const int window = (depth > 0 && depth < kWave) ? depth : kWave;
    const bool active = lane < window;
    const uint64_t below_depth = (depth >= kWave) ? ~0ull : ((1ull << depth) - 1ull);
    float v[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    // 0 not ready, 1 partial, 2 full, 3 before the signal start, 4 outside the window
    int st = !active ? 4 : (pred < 0 ? 3 : 0);
    if (st == 3) {
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] = z[i];
    }
    int D = 0;
    bool from_full = false;
    uint32_t spins = 0;
    for (;;) {
        if (st < 2) {
            uint64_t *r = rec2(lb, (uint32_t)pred, sec, ns);
            uint64_t w[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) w[i] = rec_load(r + i);
            // words: s partial y2, y1, s full y2, y1, t partial y2, y1, t full y2, y1
            if (rec_flag(w[2]) == full_flag && rec_flag(w[3]) == full_flag && rec_flag(w[6]) == full_flag &&
                rec_flag(w[7]) == full_flag) {
                st = 2;
                v[0] = rec_val(w[2]);
                v[1] = rec_val(w[3]);
                v[2] = rec_val(w[6]);
                v[3] = rec_val(w[7]);
            } else if (st == 0 && rec_flag(w[0]) == part_flag && rec_flag(w[1]) == part_flag &&
                       rec_flag(w[4]) == part_flag && rec_flag(w[5]) == part_flag) {
                st = 1;
                v[0] = rec_val(w[0]);
                v[1] = rec_val(w[1]);
                v[2] = rec_val(w[4]);
                v[3] = rec_val(w[5]);
            }
        }
        const uint64_t zmask = __ballot(st == 0);
        const uint64_t fmask = __ballot(st == 2 || st == 3);
        if (fmask) {
            const int f = __builtin_ctzll(fmask);
            const uint64_t upto = (f == kWave - 1) ? ~0ull : ((1ull << (f + 1)) - 1ull);
            if ((zmask & upto) == 0) {
                D = f;
                from_full = true;
                break;
            }
        }
        if (depth > 0 && (zmask & below_depth) == 0) {
            D = depth;
            break;
        }
        if (++spins > spin_limit) {
            if (lane == 0) __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            D = 0;
            from_full = false;
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    Carry4 X;
#pragma unroll
    for (int i = 0; i < 4; ++i) X.v[i] = from_full ? readlane_f(v[i], D) : 0.0f;
    for (int d = D - 1; d >= 0; --d) {
        float p[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) p[i] = readlane_f(v[i], d);
        const float n0 = fmaf(M[1], X.v[1], fmaf(M[0], X.v[0], p[0]));
        const float n1 = fmaf(M[5], X.v[1], fmaf(M[4], X.v[0], p[1]));
        const float n2 = fmaf(M[11], X.v[3], fmaf(M[10], X.v[2], fmaf(M[9], X.v[1], fmaf(M[8], X.v[0], p[2]))));
        const float n3 = fmaf(M[15], X.v[3], fmaf(M[14], X.v[2], fmaf(M[13], X.v[1], fmaf(M[12], X.v[0], p[3]))));
        X.v[0] = n0;
        X.v[1] = n1;
        X.v[2] = n2;
        X.v[3] = n3;
    }
    return X;
}

// Inclusive scan over the wave's 64 lanes of the pair's joint lane end
// states e (4 floats) with the uniform operator J, the ph_lane_scan scheme
// with 4 x 4 block lower triangular operators: Jpow = J^(2^k), k = 0..4
// (uniform), Jrow = J^(j+1), j = 0..15 (per lane).
__device__ __forceinline__ void mat4_lt_apply(const float *P, const float t[4], float e[4])
{
    e[0] = fmaf(P[1], t[1], fmaf(P[0], t[0], e[0]));
    e[1] = fmaf(P[5], t[1], fmaf(P[4], t[0], e[1]));
    e[2] = fmaf(P[11], t[3], fmaf(P[10], t[2], fmaf(P[9], t[1], fmaf(P[8], t[0], e[2]))));
    e[3] = fmaf(P[15], t[3], fmaf(P[14], t[2], fmaf(P[13], t[1], fmaf(P[12], t[0], e[3]))));
}

__device__ __forceinline__ void mat4_lt_apply(const cfloat *P, const float t[4], float e[4])
{
    e[0] = fmaf(P[1], t[1], fmaf(P[0], t[0], e[0]));
    e[1] = fmaf(P[5], t[1], fmaf(P[4], t[0], e[1]));
    e[2] = fmaf(P[11], t[3], fmaf(P[10], t[2], fmaf(P[9], t[1], fmaf(P[8], t[0], e[2]))));
    e[3] = fmaf(P[15], t[3], fmaf(P[14], t[2], fmaf(P[13], t[1], fmaf(P[12], t[0], e[3]))));
}

__device__ __forceinline__ void pair_lane_scan(const cfloat *Jpow, const float *Jrow, int lane, float e[4])
{
    static_for<0, 4, 1>([&](auto K) {
        constexpr int k = decltype(K)::value;
        float t[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) t[i] = row_shr_zero<1 << k>(e[i]);
        mat4_lt_apply(Jpow + 16 * k, t, e);
    });
    const cfloat *P = Jpow + 64;  // J^16
    float a[4], b[4], c[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        a[i] = readlane_f(e[i], 15);
        b[i] = readlane_f(e[i], 31);
        c[i] = readlane_f(e[i], 47);
    }
    mat4_lt_apply(P, a, b);
    mat4_lt_apply(P, b, c);
    const int row = lane >> 4;
    float k4[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) k4[i] = row == 1 ? a[i] : row == 2 ? b[i] : row == 3 ? c[i] : 0.0f;
    float J[16];
    const float4 *jr = reinterpret_cast<const float4 *>(Jrow + 16 * (lane & 15));
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const float4 v = jr[q];
        J[4 * q] = v.x;
        J[4 * q + 1] = v.y;
        J[4 * q + 2] = v.z;
        J[4 * q + 3] = v.w;
    }
    mat4_lt_apply(J, k4, e);
}

// Lookback over K-word records ([K partial words][K full words] per chunk,
// `rec_stride` words apart), for carries of K floats (PLR's direct-form
// history, the block-state form's state vector).  Predecessors are scanned
// one at a time, a whole record per load (partial word j on lane j, full word
// j on lane K + j); a record is used whole, all-full or all-partial, never
// mixed.  The fold X <- F X + p runs farthest first, with
// row j of the chunk transition F (K x K, row-major in `F`) on lane j.
// `lp` is LDS scratch of SLOTS*K floats: at most SLOTS partials are folded;
// an unbounded lookback (depth 0) or one deeper than SLOTS takes a full
// record in its last slot.  Returns X_j on lane j < K (0 elsewhere).
// The chain starts at chunk `first` from the state `start[j]` (a full record's
// words, device memory; nullptr = zero state).
// The predecessor record's first poll, issued early so that its round trip
// overlaps independent work (lanes [0, 2K), 0 elsewhere and for chunk 0).
template <int K>
__device__ __forceinline__ uint64_t lookback_first_poll(uint64_t *lb, size_t rec_stride, uint32_t chunk)
{
    const int lane = threadIdx.x & (kWave - 1);
    return (chunk > 0 && lane < 2 * K) ? rec_load(lb + (size_t)(chunk - 1) * rec_stride + lane) : 0ull;
}

template <int K, int SLOTS = kWave>
__device__ __forceinline__ float lookbackK(uint64_t *lb, size_t rec_stride, uint32_t chunk, uint32_t part_flag,
                                          uint32_t full_flag, const float *F, int depth, float *lp,
                                          uint32_t spin_limit, uint32_t *err, const uint64_t *first_poll = nullptr,
                                          uint32_t first = 0, const float *start = nullptr)
{
    static_assert(2 * K <= kWave, "a record's two halves must fit on one wave");
    const int lane = threadIdx.x & (kWave - 1);
    const bool capped = depth <= 0 || depth > SLOTS;
    const int maxd = capped ? SLOTS : depth;
    float Xj = 0.0f;
    int D = 0;
    bool done = false;
    uint32_t spins = 0;
    for (int d = 0; d < maxd && !done; ++d) {
        const long pred = (long)chunk - 1 - d;
        if (pred < (long)first) {
            D = d;  // signal (or channel) start: the start state, zero by default
            if (start && lane < K) Xj = start[lane];
            done = true;
            break;
        }
        uint64_t *pr = lb + (size_t)pred * rec_stride;
        bool use_first = first_poll && d == 0;
        for (;;) {
            // one round trip for the whole record: partial words on lanes
            // [0, K), full words on lanes [K, 2K)
            const uint64_t w = use_first ? *first_poll : ((lane < 2 * K) ? rec_load(pr + lane) : 0ull);
            use_first = false;
            const float v = rec_val(w);
            if (__ballot(lane >= K && lane < 2 * K && rec_flag(w) != full_flag) == 0) {
                Xj = __shfl_down(v, K, kWave);  // full word j from lane j + K
                D = d;
                done = true;
                break;
            }
            // the last slot of an unbounded fold must be a full record
            if (__ballot(lane < K && rec_flag(w) != part_flag) == 0 && !(capped && d == maxd - 1)) {
                if (lane < K) lp[d * K + lane] = v;
                break;
            }
            if (++spins > spin_limit) {
                if (lane == 0) __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                D = d;
                Xj = 0.0f;
                done = true;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (!done) D = d + 1;
    }
    float Frow[K];
#pragma unroll
    for (int m = 0; m < K; ++m) Frow[m] = (lane < K) ? F[lane * K + m] : 0.0f;
    if (lane >= K) Xj = 0.0f;
    for (int d = D - 1; d >= 0; --d) {
        float acc = (lane < K) ? lp[d * K + lane] : 0.0f;
#pragma unroll
        for (int m = 0; m < K; ++m) acc = fmaf(Frow[m], readlane_f(Xj, m), acc);
        Xj = acc;
    }
    return Xj;
}

// ------------------------------------------------------------------------
// chunk driver: runs body(chunk, part_flag, full_flag) for every chunk this
// workgroup owns under the plan's schedule (see the protocol notes above)
// ------------------------------------------------------------------------
template <int WV = 1, int WPC = 1, typename Body>
__device__ __forceinline__ void for_each_chunk(unsigned long long *ticket, uint32_t *epoch_word, uint32_t *arrive,
                                               uint32_t num_chunks, int persistent, uint32_t spin_limit,
                                               uint32_t *err, Body &&body)
{
    if constexpr (WV > 1) {
        // The ticket schedule with WV chunks (WPC waves each) per workgroup:
        // workgroups loop over groups of WV consecutive chunks, one atomic
        // ticket per group.  A launch hands out G = num_chunks / WV group
        // tickets and then one "done" ticket to each of its gridDim.x
        // workgroups, which stop at their first one: G + gridDim.x tickets
        // per launch, ticket t is group t mod (G + gridDim.x) of epoch
        // t / (G + gridDim.x), so that every launch starts on an epoch
        // boundary.  The first wave claims the next group's ticket as soon
        // as its chunk is done, while the group's later chunks, which wait on
        // it through the lookback, finish.  PH 64 x 32 at 2^25 samples with
        // groups of 8, alternating builds on one box, ran 617 / 532 / 395 /
        // 246 GS/s at orders 2 / 4 / 8 / 16 against 565 / 504 / 401 / 253 with one
        // group per workgroup, and 545 / 474 / 345 / 217 with the claim made
        // by the last wave a group ahead (its loads then wait on the atomic).
        // Forward progress without co-residency: a workgroup works on its
        // tickets in increasing order, so the lowest unfinished ticket is the
        // one its holder is working on, and every chunk it waits for belongs
        // to a lower ticket (finished) or to its own group, whose waves meet
        // at no barrier before the group ends.
        __shared__ unsigned long long s_tk[2];
        const uint32_t groups = num_chunks / WV;
        const unsigned long long per = (unsigned long long)groups + gridDim.x;
        const bool claimer = threadIdx.x == 0;
        if (claimer) s_tk[0] = atomicAdd(ticket, 1ull);
        __syncthreads();
        unsigned long long tk = s_tk[0];
        const unsigned long long epoch = tk / per;
        const unsigned long long first = epoch * per;
        const uint32_t pf = (uint32_t)(2ull * epoch + 1ull);
        int k = 0;
        while (tk - first < groups) {
            body((uint32_t)(tk - first) * WV + (threadIdx.x >> 6) / WPC, pf, pf + 1u);
            k ^= 1;
            if (claimer) s_tk[k] = atomicAdd(ticket, 1ull);
            __syncthreads();
            tk = s_tk[k];
        }
    } else if (persistent) {
        const uint32_t epoch = epoch_load(epoch_word);
        epoch_arrive(arrive, epoch);
        const uint32_t pf = 2u * epoch + 1u, ff = pf + 1u;
        uint32_t c = blockIdx.x;
        for (; c < num_chunks; c += gridDim.x) body(c, pf, ff);
        // the owner of the last chunk advances the epoch after publishing it
        if (threadIdx.x < kWave && (num_chunks - 1u) % gridDim.x == blockIdx.x)
            epoch_advance(epoch_word, arrive, epoch, gridDim.x * ((blockDim.x + kWave - 1) / kWave), spin_limit, err);
    } else {
        const Ticket tk = take_ticket(ticket, num_chunks);
        body(tk.chunk, tk.part_flag, tk.full_flag);
    }
}

// ------------------------------------------------------------------------
// global <-> LDS staging, 16-byte coalesced accesses
// Chunk of L lanes x N samples, lane t owns samples [t*N, t*N + N); LDS rows
// are padded to N+1 floats so that the per-lane row reads hit distinct banks.
// ------------------------------------------------------------------------
// tid: this thread among the T that move the chunk (threadIdx.x unless the
// workgroup holds several chunks)
template <int L, int N, int T = L>
__device__ __forceinline__ void stage_in(const float *__restrict__ src, float *lds, int tid = threadIdx.x)
{
    constexpr int V = L * N / 4;           // float4s in the chunk
    constexpr int PER = V / T;             // float4s per thread
    static_assert(V % T == 0, "chunk must split evenly over the threads");
    const f4 *s4 = reinterpret_cast<const f4 *>(src);
    f4 v[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) v[i] = __builtin_nontemporal_load(s4 + i * T + tid);
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int s = 4 * (i * T + tid);
        const int row = s / N, col = s % N;
        float *d = lds + row * (N + 1) + col;
        d[0] = v[i].x;
        d[1] = v[i].y;
        d[2] = v[i].z;
        d[3] = v[i].w;
    }
}

// Stores go through the L2 (not nontemporal): PH 64 x 32 at 2^25 samples,
// alternating builds on one box, ran 682-693 GS/s at order 2 against 651-663
// with nontemporal stores (574 vs 555-570 at order 4, within 1% at orders 8
// and 16).  The loads stay nontemporal: plain loads ran 639-644 at order 2.
template <int L, int N, int T = L>
__device__ __forceinline__ void stage_out(const float *lds, float *__restrict__ dst, int tid = threadIdx.x)
{
    constexpr int V = L * N / 4;
    constexpr int PER = V / T;
    f4 *d4 = reinterpret_cast<f4 *>(dst);
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int s = 4 * (i * T + tid);
        const int row = s / N, col = s % N;
        const float *p = lds + row * (N + 1) + col;
        f4 v;
        v.x = p[0];
        v.y = p[1];
        v.z = p[2];
        v.w = p[3];
        d4[i * T + tid] = v;
    }
}

// The last chunk of a signal whose length is no multiple of the chunk holds
// `valid` samples (1 <= valid < L N).  They move with coalesced 4-byte
// accesses in the layout of stage_in / stage_out; samples past the end read
// as zero (the filter is causal: they cannot reach an earlier output; the
// load index is clamped to the last sample, so every load stays in bounds)
// and are not written.
template <int L, int N, int T = L>
__device__ __forceinline__ void stage_in_part(const float *__restrict__ src, float *lds, int valid)
{
    const int tid = threadIdx.x;
    for (int e = tid; e < L * N; e += T) {
        const float v = src[e < valid ? e : valid - 1];
        lds[(e / N) * (N + 1) + e % N] = (e < valid) ? v : 0.0f;
    }
}

template <int L, int N, int T = L>
__device__ __forceinline__ void stage_out_part(const float *lds, float *__restrict__ dst, int valid)
{
    const int tid = threadIdx.x;
    for (int e = tid; e < valid; e += T) dst[e] = lds[(e / N) * (N + 1) + e % N];
}

// Samples valid - 2 and valid - 1 of a chunk held as r[N] per lane (lane t
// owns samples t N .. t N + N - 1), written by lane 0 to dst[0], dst[1],
// through the chunk's LDS tile (row pitch N + 1), which is free between the
// chunk's staging in and out.  Sample -1 (valid == 1) is lane 0's `before`,
// the value just ahead of the chunk.  This is the streaming state at the end
// of a signal that ends inside the chunk.  (Selecting the two samples from
// registers instead lets the compiler hoist 2 N lane masks out of the
// section loop, which spills scalar registers.)
template <int N>
__device__ __forceinline__ void tail_pair(const float (&r)[N], float *lds, int valid, float before, float *dst)
{
    const int lane = threadIdx.x & 63;
    __syncthreads();  // earlier reads of the tile are done
#pragma unroll
    for (int n = 0; n < N; ++n) lds[lane * (N + 1) + n] = r[n];
    __syncthreads();
    if (lane == 0) {
        const int e1 = valid - 1, e2 = valid - 2;
        dst[0] = e2 >= 0 ? lds[(e2 / N) * (N + 1) + e2 % N] : before;
        dst[1] = lds[(e1 / N) * (N + 1) + e1 % N];
    }
}

}  // namespace iirx
