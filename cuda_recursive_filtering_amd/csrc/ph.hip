// ph.hip -- PH (partial-LU factorisation) for gfx950.
//
// Algorithm of include/ph_kernels.cuh (PH_32 208-498, PH_64 501-820): every
// lane filters its N samples with zero incoming recursive state, the lane end
// states are combined by a scan with powers of the N-step transition C, the
// chunk's incoming state comes from the decoupled lookback, and every sample
// is corrected by the zero-input response to the lane's incoming state.
//
// CDNA4 mapping.  One workgroup is one wave64 holding a chunk of 64 lanes x N
// samples: the reference's 32-lane chunks are packed two per wave (their
// combine is part of the wave's scan), its 64-lane chunk is the wave.  The
// lane scan is a Kogge-Stone scan with the uniform operators C^(2^k) held in
// scalar registers (the reference uses a segmented scan with per-lane
// operators loaded through 32-lane shuffles).  Samples stay in VGPRs between
// sections; HBM is touched once on the way in and once on the way out, with
// 16-byte accesses staged through LDS.
#include "device_common.hpp"
#include "kernels.hpp"

#include <algorithm>

namespace iirx {

// LAST = false: chunks 0 .. num_chunks - 1 under the plan's schedule.
// LAST = true: one workgroup for chunk num_chunks, the partial chunk that ends
// a length which is no multiple of the chunk, launched after the first launch
// on the same stream: its incoming state is the full record its predecessor
// left (no lookback), samples past the end read as zero and are not written,
// and with a.tail set it records the end state there.
// WV > 1: the ticket schedule with WV chunks per workgroup, one wave and one
// LDS tile each (one atomic ticket per group).
// HALO: every section's chunk carry from the kPhHalo input samples before the
// chunk (4 per lane, loaded with the chunk): y_s[-2], y_s[-1] = wave sums of
// the lane's samples times its K2_s, K1_s weights.  No chunk reads another's
// record, so there is no lookback and no schedule (no ticket, no epoch,
// nothing that needs workgroups to be co-resident); the full records are
// still written (the partial last chunk and the end state read the last one's).
// JOINT: sections in pairs (s, s + 1) with one lookback of the pair's joint
// state per chunk (an odd last section alone, and so a pair whose tables
// exceed kPairLimit).  Both sections' zero-state
// passes run first, the second on the first's output with zero history
// (that history is part of the joint state); the lane end states go through
// one 4 x 4 scan; the chunk partial of both sections is published, one
// lookback returns the joint carry X, and every sample of the second section
// is corrected by the response of the joint state U at its lane's start,
// y_t[n] += sum_m H_m[n] U_m.  Per sample that is the arithmetic of two PH
// sections (8 FMA zero-state + 4 FMA correction against 2 x (4 + 2)); the
// scan takes 12 FMA per level against 2 x 4, and a chunk waits for one
// lookback instead of two.
// Unpadded chunk tile with 16-byte LDS accesses (the halo instance at N =
// 32): lane l's samples 4 k .. 4 k + 3 sit in float4 slot 8 l + ((k + l +
// (l >> 3)) mod 8).  Both the per-lane row accesses (slot(l, k) over the 64
// lanes) and the accesses in address order (float4 g = 64 i + lane at
// slot(g / 8, g mod 8)) hit 16 distinct 16-byte bank groups in every 16
// consecutive lanes and 8 in every 8 (checked by brute force for 32 and 64
// banks); one ds_read/ds_write_b128 replaces four 4-byte ones of the padded
// tile.
__device__ __forceinline__ int swz_slot(int l, int k) { return 8 * l + ((k + l + (l >> 3)) & 7); }

template <int N, bool LAST, int WV = 1, bool JOINT = false, bool HALO = false>
__global__ __launch_bounds__(64 * WV) void ph_kernel(LaunchArgs a)
{
    constexpr int L = 64;
    constexpr PhLayout Y = ph_layout(L, N);
    constexpr PhPairLayout YP = ph_pair_layout(L, N);
    __shared__ float lds_w[WV][L * (N + 1)];
    float *lds = lds_w[WV == 1 ? 0 : threadIdx.x >> 6];

    const int lane = WV == 1 ? (int)threadIdx.x : (int)(threadIdx.x & 63);
    // the 16-byte tile (halo instance, N = 32)
    constexpr bool SWZ = HALO && !LAST && WV == 1 && N == 32;
    f4 *lds4 = reinterpret_cast<f4 *>(lds);
    const auto body = [&](uint32_t chunk, uint32_t part_flag, uint32_t full_flag) {
    const long base = (long)chunk * (L * N);

    chunk_sync<WV>();  // LDS of the previous chunk fully consumed
    // HALO: samples base - kPhHalo + 4 lane .. + 3, in flight with the chunk's loads
    f4 hx = {0.0f, 0.0f, 0.0f, 0.0f};
    if constexpr (HALO && !LAST)
        if (chunk > 0) hx = *reinterpret_cast<const f4 *>(a.in + base - kPhHalo + 4 * lane);
    float r[N];
    float xi2, xi1;
    if constexpr (SWZ) {
        const f4 *src = reinterpret_cast<const f4 *>(a.in + base);
        f4 v[N / 4];
#pragma unroll
        for (int i = 0; i < N / 4; ++i) v[i] = __builtin_nontemporal_load(src + 64 * i + lane);
#pragma unroll
        for (int i = 0; i < N / 4; ++i) lds4[swz_slot(8 * i + (lane >> 3), lane & 7)] = v[i];
        chunk_sync<WV>();
#pragma unroll
        for (int k = 0; k < N / 4; ++k) {
            const f4 w = lds4[swz_slot(lane, k)];
            r[4 * k] = w[0];
            r[4 * k + 1] = w[1];
            r[4 * k + 2] = w[2];
            r[4 * k + 3] = w[3];
        }
        const f4 w = lds4[swz_slot(lane > 0 ? lane - 1 : 0, N / 4 - 1)];
        xi2 = w[2];
        xi1 = w[3];
    } else {
        if constexpr (LAST) stage_in_part<L, N>(a.in + base, lds, a.last_valid);
        else stage_in<L, N>(a.in + base, lds, lane);
        chunk_sync<WV>();
#pragma unroll
        for (int n = 0; n < N; ++n) r[n] = lds[lane * (N + 1) + n];
        xi2 = lds[(lane > 0 ? lane - 1 : 0) * (N + 1) + N - 2];
        xi1 = lds[(lane > 0 ? lane - 1 : 0) * (N + 1) + N - 1];
    }

    // inputs of the first section just before this lane's samples: lane l > 0
    // took the last two of row l - 1 above; lane 0 those before the chunk
    if (lane == 0) {
        if (chunk > 0) {
            xi2 = a.in[base - 2];
            xi1 = a.in[base - 1];
        } else if (a.zi) {
            xi2 = a.zi[0];
            xi1 = a.zi[1];
        } else {
            xi2 = 0.0f;
            xi1 = 0.0f;
        }
    }
    // the end state of the signal: samples last_valid - 2 and - 1
    if constexpr (LAST)
        if (a.tail) tail_pair<N>(r, lds, a.last_valid, xi1, a.tail);

    for (int sec = 0; sec < a.ns; ++sec) {
        if constexpr (JOINT && !LAST) {
            const float *pv = a.coef + (size_t)a.ns * a.coef_stride + (size_t)(sec >> 1) * YP.stride;
            const cfloat *pc = as_const(pv);
            // pairs start at even sections; a pair whose tables are too large
            // runs one section at a time
            if ((sec & 1) == 0 && sec + 1 < a.ns && pc[YP.kDepth + 1] != 0.0f) {
                const int t = sec + 1;
                const cfloat *cs = as_const(a.coef + (size_t)sec * a.coef_stride);
                const cfloat *ct = as_const(a.coef + (size_t)t * a.coef_stride);
                const float b1 = cs[Y.kB + 0], b2 = cs[Y.kB + 1], a1 = cs[Y.kB + 2], a2 = cs[Y.kB + 3];
                const float d1 = ct[Y.kB + 0], d2 = ct[Y.kB + 1], c1 = ct[Y.kB + 2], c2 = ct[Y.kB + 3];

                // both zero-state passes; section t's input history is section
                // s's output history, zero at the lane start like its own
                float x2 = xi2, x1 = xi1, y2 = 0.0f, y1 = 0.0f, v2 = 0.0f, v1 = 0.0f;
#pragma unroll
                for (int n = 0; n < N; ++n) {
                    const float xc = r[n];
                    float yc = fmaf(b1, x1, fmaf(b2, x2, xc));
                    if (n == 1) yc = fmaf(a1, y1, yc);
                    else if (n > 1) yc = fmaf(a1, y1, fmaf(a2, y2, yc));
                    float vc = yc;
                    if (n == 1) vc = fmaf(c1, v1, fmaf(d1, y1, vc));
                    else if (n > 1) vc = fmaf(c1, v1, fmaf(c2, v2, fmaf(d1, y1, fmaf(d2, y2, vc))));
                    r[n] = vc;
                    x2 = x1;
                    x1 = xc;
                    y2 = y1;
                    y1 = yc;
                    v2 = v1;
                    v1 = vc;
                }

                // lane scan of the joint end states
                float e[4] = {y2, y1, r[N - 2], r[N - 1]};
                pair_lane_scan(pc + YP.kJpow, pv + YP.kJrow, lane, e);

                // chunk partial of both sections, one lookback, chunk full
                float p[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) p[i] = readlane_f(e[i], L - 1);
                uint64_t *rs = rec2(a.lb, chunk, sec, a.ns);
                if (lane == 0) {
                    publish2(rs, false, part_flag, p[0], p[1]);
                    publish2(rs + 4, false, part_flag, p[2], p[3]);
                }
                float z[4] = {0.0f, 0.0f, 0.0f, 0.0f};
                if (a.zi) {
                    z[0] = a.zi[2 + 2 * sec];
                    z[1] = a.zi[3 + 2 * sec];
                    z[2] = a.zi[2 + 2 * t];
                    z[3] = a.zi[3 + 2 * t];
                }
                // J^(lane+1): its loads are in flight during the lookback
                const float4 *jf = reinterpret_cast<const float4 *>(pv + YP.kJfin + 16 * lane);
                float Jl[16];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float4 w4 = jf[q];
                    Jl[4 * q] = w4.x;
                    Jl[4 * q + 1] = w4.y;
                    Jl[4 * q + 2] = w4.z;
                    Jl[4 * q + 3] = w4.w;
                }
                const Carry4 X = lookback_pair(a.lb, chunk, sec, a.ns, part_flag, full_flag, pc + YP.kM,
                                               (int)pc[YP.kDepth], a.spin_limit, a.err, z);
                if (lane == 0) {
                    float f[4] = {p[0], p[1], p[2], p[3]};
                    mat4_lt_apply(pc + YP.kM, X.v, f);
                    publish2(rs, true, full_flag, f[0], f[1]);
                    publish2(rs + 4, true, full_flag, f[2], f[3]);
                }

                // per-lane end state with the chunk carry injected, J^(lane+1) X;
                // incoming state of each lane = end state of the previous lane
                mat4_lt_apply(Jl, X.v, e);
                // (the shuffle runs on every lane: a source lane masked off by a
                // branch would not deliver its value)
                float u[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) u[i] = lane_up(e[i], 1);
                if (lane == 0) {
#pragma unroll
                    for (int i = 0; i < 4; ++i) u[i] = X.v[i];
                }

                // correction of section t's output by the joint state
#pragma unroll
                for (int n = 0; n < N - 2; ++n)
                    r[n] = fmaf(pc[YP.kH + 3 * N + n], u[3],
                                fmaf(pc[YP.kH + 2 * N + n], u[2],
                                     fmaf(pc[YP.kH + N + n], u[1], fmaf(pc[YP.kH + n], u[0], r[n]))));
                r[N - 2] = e[2];
                r[N - 1] = e[3];
                xi2 = u[2];
                xi1 = u[3];
                ++sec;
                continue;
            }
        }
        const float *cv = a.coef + (size_t)sec * a.coef_stride;  // per-lane vector loads
        const cfloat *c = as_const(cv);                           // uniform scalar loads
        const float b1 = c[Y.kB + 0], b2 = c[Y.kB + 1], a1 = c[Y.kB + 2], a2 = c[Y.kB + 3];
        // HALO: this lane's halo weights of the section, K2_s and K1_s
        f4 k2 = {0.0f, 0.0f, 0.0f, 0.0f}, k1 = k2;
        if constexpr (HALO && !LAST) {
            const f4 *kh = reinterpret_cast<const f4 *>(a.coef + (size_t)a.ns * a.coef_stride + 4 +
                                                        (size_t)sec * 2 * kPhHalo);
            k2 = kh[lane];
            k1 = kh[kPhHalo / 4 + lane];
        }

        // hot loop 1: zero-state biquad (ph_kernels.cuh:265-287)
        {
            float x2 = xi2, x1 = xi1, y2 = 0.0f, y1 = 0.0f;
#pragma unroll
            for (int n = 0; n < N; ++n) {
                const float xc = r[n];
                float yc = fmaf(b1, x1, fmaf(b2, x2, xc));
                if (n == 1) yc = fmaf(a1, y1, yc);
                else if (n > 1) yc = fmaf(a1, y1, fmaf(a2, y2, yc));
                r[n] = yc;
                x2 = x1;
                x1 = xc;
                y2 = y1;
                y1 = yc;
            }
        }

        // lane scan of the end states
        float s2 = r[N - 2], s1 = r[N - 1];
        ph_lane_scan(c + Y.kCpow, cv + Y.kCfin, lane, s2, s1);

        // chunk partial, lookback, chunk full (ph_kernels.cuh:384-469)
        const float m0 = c[Y.kM + 0], m1 = c[Y.kM + 1], m2 = c[Y.kM + 2], m3 = c[Y.kM + 3];
        const float p2 = readlane_f(s2, L - 1), p1 = readlane_f(s1, L - 1);
        uint64_t *rec = rec2(a.lb, chunk, sec, a.ns);
        if (!LAST && !HALO && lane == 0) publish2(rec, false, part_flag, p2, p1);
        const int depth = (int)c[Y.kDepth];
        // the chain starts from the streamed-in state of this section (zero by default)
        const float z2 = a.zi ? a.zi[2 + 2 * sec] : 0.0f, z1 = a.zi ? a.zi[3 + 2 * sec] : 0.0f;
        // C^(lane+1), the per-lane carry injection: its load is in flight
        // during the lookback instead of after it
        const float4 cf = *reinterpret_cast<const float4 *>(cv + Y.kCfin + 4 * lane);
        Carry2 X;
        if constexpr (LAST) {
            X = prior_state2(a.lb, chunk, sec, a.ns, z2, z1);
        } else if constexpr (HALO) {
            X = Carry2{z2, z1};
            if (chunk > 0) {
                X.y2 = wave_sum(fmaf(k2[3], hx[3], fmaf(k2[2], hx[2], fmaf(k2[1], hx[1], k2[0] * hx[0]))));
                X.y1 = wave_sum(fmaf(k1[3], hx[3], fmaf(k1[2], hx[2], fmaf(k1[1], hx[1], k1[0] * hx[0]))));
            }
        } else {
            X = lookback2(a.lb, chunk, sec, a.ns, part_flag, full_flag, m0, m1, m2, m3, depth, a.spin_limit, a.err,
                          0u, z2, z1);
        }
        if (!LAST && lane == 0) {
            const float f2 = fmaf(m1, X.y1, fmaf(m0, X.y2, p2));
            const float f1 = fmaf(m3, X.y1, fmaf(m2, X.y2, p1));
            publish2(rec, true, full_flag, f2, f1);
        }

        // per-lane end state with the chunk carry injected, C^(lane+1) X
        const float e2 = fmaf(cf.y, X.y1, fmaf(cf.x, X.y2, s2));
        const float e1 = fmaf(cf.w, X.y1, fmaf(cf.z, X.y2, s1));

        // incoming state of each lane = end state of the previous lane
        float u2 = lane_up(e2, 1), u1 = lane_up(e1, 1);
        if (lane == 0) {
            u2 = X.y2;
            u1 = X.y1;
        }

        // hot loop 2: back-correction (ph_kernels.cuh:471-480)
#pragma unroll
        for (int n = 0; n < N - 2; ++n) r[n] = fmaf(c[Y.kH1 + n], u1, fmaf(c[Y.kH2 + n], u2, r[n]));
        r[N - 2] = e2;
        r[N - 1] = e1;

        // the next section's input history is this section's output
        xi2 = u2;
        xi1 = u1;
        if constexpr (LAST)
            if (a.tail) tail_pair<N>(r, lds, a.last_valid, X.y1, a.tail + 2 + 2 * sec);
    }

    chunk_sync<WV>();
    if constexpr (SWZ) {
#pragma unroll
        for (int k = 0; k < N / 4; ++k) lds4[swz_slot(lane, k)] = f4{r[4 * k], r[4 * k + 1], r[4 * k + 2], r[4 * k + 3]};
        chunk_sync<WV>();
        f4 *dst = reinterpret_cast<f4 *>(a.out + base);
#pragma unroll
        for (int i = 0; i < N / 4; ++i) dst[64 * i + lane] = lds4[swz_slot(8 * i + (lane >> 3), lane & 7)];
    } else {
#pragma unroll
        for (int n = 0; n < N; ++n) lds[lane * (N + 1) + n] = r[n];
        chunk_sync<WV>();
        if constexpr (LAST) stage_out_part<L, N>(lds, a.out + base, a.last_valid);
        else stage_out<L, N>(lds, a.out + base, lane);
    }
    };
    if constexpr (LAST) {
        body(a.num_chunks, 0u, 0u);
    } else if constexpr (HALO) {
        // No chunk waits on another, so no order of the workgroups can stall
        // the launch: a grid-stride loop with no ticket and no epoch, under
        // either schedule.  The full records carry flag 0; their only readers
        // (the partial last chunk and the end state) run after this launch
        // on the same stream and read the values alone.
        // Chunk order: workgroups go to the 8 XCDs round-robin by index, so
        // workgroup b takes position (b mod 8) G/8 + b / 8 of every round of
        // G chunks: each XCD works through a contiguous run, and a chunk's
        // halo is its predecessor's input, read through the same L2.
        static_assert(WV == 1, "the halo instance runs one wave per workgroup");
        const uint32_t G = gridDim.x, b = blockIdx.x;
        const uint32_t p = (G % 8u == 0u) ? (b % 8u) * (G / 8u) + b / 8u : b;
        for (uint32_t c = p; c < a.num_chunks; c += G) body(c, 0u, 0u);
    } else {
        for_each_chunk<WV>(a.ticket, a.epoch, a.arrive, a.num_chunks, a.persistent, a.spin_limit, a.err, body);
    }
}

// zf = x[n-2], x[n-1], then (y2, y1) of every section from the last chunk's
// full records (the launch that wrote them has finished: same stream)
__global__ __launch_bounds__(64) void ph_end_state_kernel(LaunchArgs a, int chunk_samples, float *zf)
{
    const int t = threadIdx.x;
    const uint32_t last = a.num_chunks - 1u;
    if (t < 2) zf[t] = a.in[(long)last * chunk_samples + chunk_samples - 2 + t];
    if (t < 2 * a.ns) {
        const uint64_t *r = rec2(a.lb, last, t >> 1, a.ns);
        zf[2 + t] = rec_val(__hip_atomic_load(r + 2 + (t & 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    }
}

hipError_t launch_ph_end_state(const LaunchArgs &a, int chunk_samples, float *zf, hipStream_t s)
{
    hipLaunchKernelGGL(ph_end_state_kernel, dim3(1), dim3(64), 0, s, a, chunk_samples, zf);
    return hipGetLastError();
}

// the partial last chunk always runs the per-section instance: it reads the
// full records of its predecessor per section, which the joint instance
// writes in the same layout
template <int N, bool J, bool H>
static hipError_t launch_n(const LaunchArgs &a, bool last, hipStream_t s)
{
    // the halo instance has no grouped form (no ticket)
    constexpr int WV = H ? 1 : ph_ticket_waves(N);
    if (last) hipLaunchKernelGGL((ph_kernel<N, true>), dim3(1), dim3(64), 0, s, a);
    else if (!H && !a.persistent && a.num_chunks % WV == 0)
        hipLaunchKernelGGL((ph_kernel<N, false, WV, J, H>), dim3(a.grid), dim3(64 * WV), 0, s, a);
    else hipLaunchKernelGGL((ph_kernel<N, false, 1, J, H>), dim3(a.grid), dim3(64), 0, s, a);
    return hipGetLastError();
}

template <int N>
static hipError_t launch_jh(const LaunchArgs &a, bool last, hipStream_t s, bool joint, bool halo)
{
    if (joint) return launch_n<N, true, false>(a, last, s);
    if (halo) return launch_n<N, false, true>(a, last, s);
    return launch_n<N, false, false>(a, last, s);
}

hipError_t launch_ph(int lanes, int n_blocks, const LaunchArgs &a, hipStream_t s, bool last, bool joint, bool halo)
{
    if (lanes != 64 || (joint && halo)) return hipErrorInvalidValue;
    switch (n_blocks) {
    case 8: return launch_jh<8>(a, last, s, joint, halo);
    case 16: return launch_jh<16>(a, last, s, joint, halo);
    case 32: return launch_jh<32>(a, last, s, joint, halo);
    case 64: return launch_jh<64>(a, last, s, joint, halo);
    case 128: return launch_jh<128>(a, last, s, joint, halo);
    default: return hipErrorInvalidValue;
    }
}

template <int N, bool J, bool H>
static const void *fn_n(bool grouped)
{
    if (grouped && !H) return (const void *)ph_kernel<N, false, H ? 1 : ph_ticket_waves(N), J, H>;
    return (const void *)ph_kernel<N, false, 1, J, H>;
}

template <int N>
static const void *fn_jh(bool grouped, bool joint, bool halo)
{
    if (joint) return fn_n<N, true, false>(grouped);
    if (halo) return fn_n<N, false, true>(grouped);
    return fn_n<N, false, false>(grouped);
}

const void *kernel_ph(int lanes, int n_blocks, bool grouped, bool joint, bool halo)
{
    if (lanes != 64 || (joint && halo)) return nullptr;
    switch (n_blocks) {
    case 8: return fn_jh<8>(grouped, joint, halo);
    case 16: return fn_jh<16>(grouped, joint, halo);
    case 32: return fn_jh<32>(grouped, joint, halo);
    case 64: return fn_jh<64>(grouped, joint, halo);
    case 128: return fn_jh<128>(grouped, joint, halo);
    default: return nullptr;
    }
}

}  // namespace iirx
