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
// LDS tile each (one atomic ticket per workgroup).
template <int N, bool LAST, int WV = 1>
__global__ __launch_bounds__(64 * WV) void ph_kernel(LaunchArgs a)
{
    constexpr int L = 64;
    constexpr PhLayout Y = ph_layout(L, N);
    __shared__ float lds_w[WV][L * (N + 1)];
    float *lds = lds_w[WV == 1 ? 0 : threadIdx.x >> 6];

    const int lane = WV == 1 ? (int)threadIdx.x : (int)(threadIdx.x & 63);
    const auto body = [&](uint32_t chunk, uint32_t part_flag, uint32_t full_flag) {
    const long base = (long)chunk * (L * N);

    chunk_sync<WV>();  // LDS of the previous chunk fully consumed
    if constexpr (LAST) stage_in_part<L, N>(a.in + base, lds, a.last_valid);
    else stage_in<L, N>(a.in + base, lds, lane);
    chunk_sync<WV>();

    float r[N];
#pragma unroll
    for (int n = 0; n < N; ++n) r[n] = lds[lane * (N + 1) + n];

    // inputs of the first section just before this lane's samples
    float xi2, xi1;
    if (lane > 0) {
        xi2 = lds[(lane - 1) * (N + 1) + N - 2];
        xi1 = lds[(lane - 1) * (N + 1) + N - 1];
    } else if (chunk > 0) {
        xi2 = a.in[base - 2];
        xi1 = a.in[base - 1];
    } else if (a.zi) {
        xi2 = a.zi[0];
        xi1 = a.zi[1];
    } else {
        xi2 = 0.0f;
        xi1 = 0.0f;
    }
    // the end state of the signal: samples last_valid - 2 and - 1
    if constexpr (LAST)
        if (a.tail) tail_pair<N>(r, lds, a.last_valid, xi1, a.tail);

    for (int sec = 0; sec < a.ns; ++sec) {
        const float *cv = a.coef + (size_t)sec * a.coef_stride;  // per-lane vector loads
        const cfloat *c = as_const(cv);                           // uniform scalar loads
        const float b1 = c[Y.kB + 0], b2 = c[Y.kB + 1], a1 = c[Y.kB + 2], a2 = c[Y.kB + 3];

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
        if (!LAST && lane == 0) publish2(rec, false, part_flag, p2, p1);
        const int depth = (int)c[Y.kDepth];
        // the chain starts from the streamed-in state of this section (zero by default)
        const float z2 = a.zi ? a.zi[2 + 2 * sec] : 0.0f, z1 = a.zi ? a.zi[3 + 2 * sec] : 0.0f;
        // C^(lane+1), the per-lane carry injection: its load is in flight
        // during the lookback instead of after it
        const float4 cf = *reinterpret_cast<const float4 *>(cv + Y.kCfin + 4 * lane);
        const Carry2 X = LAST ? prior_state2(a.lb, chunk, sec, a.ns, z2, z1)
                              : lookback2(a.lb, chunk, sec, a.ns, part_flag, full_flag, m0, m1, m2, m3, depth,
                                          a.spin_limit, a.err, 0u, z2, z1);
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
#pragma unroll
    for (int n = 0; n < N; ++n) lds[lane * (N + 1) + n] = r[n];
    chunk_sync<WV>();
    if constexpr (LAST) stage_out_part<L, N>(lds, a.out + base, a.last_valid);
    else stage_out<L, N>(lds, a.out + base, lane);
    };
    if constexpr (LAST) body(a.num_chunks, 0u, 0u);
    else for_each_chunk<WV>(a.ticket, a.epoch, a.arrive, a.num_chunks, a.persistent, a.spin_limit, a.err, body);
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

template <int N>
static hipError_t launch_n(const LaunchArgs &a, bool last, hipStream_t s)
{
    constexpr int WV = ph_ticket_waves(N);
    if (last) hipLaunchKernelGGL((ph_kernel<N, true>), dim3(1), dim3(64), 0, s, a);
    else if (!a.persistent && a.num_chunks % WV == 0)
        hipLaunchKernelGGL((ph_kernel<N, false, WV>), dim3(a.grid), dim3(64 * WV), 0, s, a);
    else hipLaunchKernelGGL((ph_kernel<N, false>), dim3(a.grid), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_ph(int lanes, int n_blocks, const LaunchArgs &a, hipStream_t s, bool last)
{
    if (lanes != 64) return hipErrorInvalidValue;
    switch (n_blocks) {
    case 8: return launch_n<8>(a, last, s);
    case 16: return launch_n<16>(a, last, s);
    case 32: return launch_n<32>(a, last, s);
    case 64: return launch_n<64>(a, last, s);
    case 128: return launch_n<128>(a, last, s);
    default: return hipErrorInvalidValue;
    }
}

template <int N>
static const void *fn_n(bool grouped)
{
    if (grouped) return (const void *)ph_kernel<N, false, ph_ticket_waves(N)>;
    return (const void *)ph_kernel<N, false>;
}

const void *kernel_ph(int lanes, int n_blocks, bool grouped)
{
    if (lanes != 64) return nullptr;
    switch (n_blocks) {
    case 8: return fn_n<8>(grouped);
    case 16: return fn_n<16>(grouped);
    case 32: return fn_n<32>(grouped);
    case 64: return fn_n<64>(grouped);
    case 128: return fn_n<128>(grouped);
    default: return nullptr;
    }
}

}  // namespace iirx
