// iirx_plan.cpp -- the C ABI of include/iirx.h: runtime plans, FP64 host
// precompute of the coefficient tables, the once-zeroed lookback state, the
// launch, and the hipEvent timing harness.
//
// Reference counterparts: setup_kernel_coefficients (ph_kernels.cuh:151-200,
// stcr_kernels.cuh:191-272, dtcr_kernels.cuh:193-289, plr_kernels.cuh:
// 416-476), the driver's buffer set-up and launch (test_accuracy.cu:128-160,
// test_performance.cu:140-182) and measure_kernel_time (measure.cuh:38-110).
#include "../../include/iirx.h"
#include "kernels.hpp"
#include "precompute.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

using namespace iirx;

namespace {
// per-kernel launch shape and tables of a configuration
struct Shape {
    int lanes, n_blocks, threads, plr_block, plr_x;  // plr_*: threads per block, values per thread
    int chunks_per_block;                             // waves with their own chunk stream (block-state form)
    int64_t chunk;
    uint32_t num_chunks;
    size_t lb_count;
    int coef_stride;
    bool halo;  // PH_HALO: the chunk carries come from the halo (the cascade forgets within kPhHalo samples)
};

}  // namespace

struct iirx_plan {
    iirx_config cfg;
    int64_t n_samples;
    int lanes;            // lanes per workgroup
    int n_blocks;         // samples per lane
    int threads;          // threads per workgroup
    int64_t chunk;        // samples per workgroup
    uint32_t num_chunks;
    int device;
    std::vector<float> sos;
    std::vector<float> host_coef;
    int coef_stride;
    float *d_coef;
    uint64_t *d_lb;
    Shape shape;
    std::vector<double> plr_b, plr_a;  // direct-form taps supplied by the caller
    unsigned long long *d_ticket;  // control block: [ticket u64][epoch u32][err u32], from byte 128
                                   // the persistent schedule's arrival slots
    uint32_t *d_epoch;
    uint32_t *d_err;
    uint32_t *d_arrive;
    uint32_t spin_limit;
    uint32_t grid;
    const void *fn;
    int64_t launches;
    // block-state form, streaming state: the map between the sections'
    // histories and the balanced state, then 16 floats of incoming state and
    // 16 of the state entering the last tile (one allocation)
    BsStateMap *d_bsmap;
    // a length that is no multiple of the chunk: the last chunk is partial
    // and runs in a one-workgroup launch of its own after the main launch,
    // which filters main_chunks = num_chunks - 1 chunks
    bool split_last;
    uint32_t main_chunks;
    // block-state form, partial last chunk: it runs through the PH last-chunk
    // kernel (64 x 64 samples) from the direct-form histories at the end of
    // the whole chunks; its tables and those histories
    float *d_aux;
    int aux_stride;
    float *d_mid;
};

namespace {

bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

// the block-state form and its operand split
bool is_bs(int kernel) { return kernel == IIRX_KERNEL_BLOCKSTATE || kernel == IIRX_KERNEL_BLOCKSTATE_BF16; }
bool is_ph(int kernel)
{
    return kernel == IIRX_KERNEL_PH || kernel == IIRX_KERNEL_PH_JOINT || kernel == IIRX_KERNEL_PH_HALO;
}
int bs_split_of(int kernel) { return kernel == IIRX_KERNEL_BLOCKSTATE_BF16 ? kBsSplitBF16 : kBsSplitF16; }

// rows of the filter contract: finite, b0 == 1 and a0 == 1 (iirx_normalize_sos makes them)
bool sos_ok(const float *sos, int ns)
{
    for (int s = 0; s < ns; ++s) {
        const float *r = sos + 6 * s;
        for (int k = 0; k < 6; ++k)
            if (!std::isfinite(r[k])) return false;
        if (r[0] != 1.0f || r[3] != 1.0f) return false;
    }
    return true;
}

int validate(const iirx_config *cfg, int64_t n)
{
    if (!cfg) return IIRX_ERR_INVALID;
    if (cfg->n_sections < 1 || cfg->n_sections > 8) return IIRX_ERR_INVALID;
    if (n <= 0 || n > (int64_t)1 << 34) return IIRX_ERR_INVALID;
    if (cfg->schedule != IIRX_SCHEDULE_PERSISTENT && cfg->schedule != IIRX_SCHEDULE_TICKET) return IIRX_ERR_INVALID;
    switch (cfg->kernel) {
    case IIRX_KERNEL_PH:
    case IIRX_KERNEL_PH_JOINT:
    case IIRX_KERNEL_PH_HALO:
    case IIRX_KERNEL_STCR:
    case IIRX_KERNEL_DTCR: {
        const int B = cfg->block_size, N = cfg->n_blocks;
        if (!(B == 32 || B == 64)) return IIRX_ERR_GEOMETRY;
        // BLOCK_SIZE = 32 keeps the reference's chunk of 32 N samples on one
        // wave64 as 64 lanes x N/2 (instances exist for 8..128 per lane)
        if (!is_pow2(N) || N < (B == 32 ? 16 : 8) || N > 128) return IIRX_ERR_GEOMETRY;
        // any length: the last chunk may be partial (the reference requires a
        // multiple of the chunk, test_accuracy.cu:37-38)
        const int64_t chunk = (int64_t)B * N;
        if ((n + chunk - 1) / chunk > 0x7fffffffLL) return IIRX_ERR_GEOMETRY;
        return IIRX_OK;
    }
    case IIRX_KERNEL_BLOCKSTATE:
    case IIRX_KERNEL_BLOCKSTATE_BF16: {
        const int L = cfg->n_blocks, G = cfg->block_size;
        if (!(L == 32 || L == 64 || L == 128) || !(G == 16 || G == 32 || G == 64) || (L == 128 && G == 64))
            return IIRX_ERR_GEOMETRY;
        if (cfg->kernel == IIRX_KERNEL_BLOCKSTATE_BF16 && (L == 32 || G == 64)) return IIRX_ERR_GEOMETRY;
        // any length: the samples past the last whole chunk go through PH (below)
        const int64_t chunk = (int64_t)G * L;
        if ((n + chunk - 1) / chunk > 0x7fffffffLL) return IIRX_ERR_GEOMETRY;
        return IIRX_OK;
    }
    case IIRX_KERNEL_PLR: {
        const int ns = cfg->n_sections;
        if (!(ns == 1 || ns == 2 || ns == 4 || ns == 8)) return IIRX_ERR_GEOMETRY;
        const int B = cfg->block_size;
        if (!(B == 0 || B == 64 || B == 128 || B == 256 || B == 512)) return IIRX_ERR_GEOMETRY;
        const int X = cfg->plr_x;
        if (!(X == 0 || X == 1 || X == 2 || X == 4 || X == 8 || X == 16)) return IIRX_ERR_GEOMETRY;
        if (cfg->plr_x_max < 0) return IIRX_ERR_INVALID;
        return IIRX_OK;
    }
    default:
        return IIRX_ERR_INVALID;
    }
}

// PLR geometry (plr_kernels.cuh:49-112): threads per block by order unless
// given, values per thread from the adaptive rule X = n / (CUs_p2 * BLOCK)
// clamped to [1, X_MAX] and rounded down to an instantiated power of two.
// PLR geometry, the reference's rule (plr_kernels.cuh:49-112: a per-order
// PLR_BLOCK_SIZE table, PLR_X = clamp(N / (nextpow2(#SM) * BLOCK), 1,
// PLR_X_MAX)) with an MI355X table: one wave per chunk (no barriers; the
// merge and the segment chain stay in registers) and X_MAX = 16 at orders 2
// and 4, 8 above, measured at 2^25 samples (64 x 16, 64 x 16, 64 x 8, 64 x 8 at
// orders 2, 4, 8, 16; order 4 at X = 16: 437 GS/s against 425 at X = 8, order
// 8 at X = 16: 105 against 200; the reference-style 256-thread blocks with
// X = 16 ran 120 / 48 / 28 / 11 GS/s in an earlier form of the kernel); the
// spill calibration below may still halve X.
void plr_geometry(const iirx_config *cfg, int64_t n, int *block, int *x)
{
    const int k = 2 * cfg->n_sections;
    int B = cfg->block_size;
    if (B == 0) B = 64;
    int X = cfg->plr_x;
    if (X == 0) {
        const int cap = (k <= 4) ? 16 : 8;
        const int xmax = (cfg->plr_x_max > 0) ? std::min(cfg->plr_x_max, 16) : cap;
        int64_t raw = n / (256LL * B);
        raw = std::max<int64_t>(1, std::min<int64_t>(raw, xmax));
        X = 1;
        while (X * 2 <= raw) X *= 2;
    }
    *block = B;
    *x = X;
}

int make_shape(const iirx_config *cfg, const float *sos, int64_t n, const double *plr_b, const double *plr_a,
               Shape &sh, std::vector<float> &tables)
{
    const int ns = cfg->n_sections;
    sh.lanes = 64;
    sh.n_blocks = cfg->n_blocks;
    // PH / STCR / DTCR: a chunk is BLOCK_SIZE x N_BLOCKS samples as in the
    // reference (test_accuracy.cu:37-38); a wave64 holds it as 64 lanes of
    // 32 N / 64 samples when BLOCK_SIZE is 32
    if (is_ph(cfg->kernel) || cfg->kernel == IIRX_KERNEL_STCR || cfg->kernel == IIRX_KERNEL_DTCR)
        sh.n_blocks = cfg->n_blocks * cfg->block_size / 64;
    sh.threads = 64;
    sh.chunks_per_block = 1;
    sh.plr_block = sh.plr_x = 0;
    sh.halo = false;
    switch (cfg->kernel) {
    case IIRX_KERNEL_PH:
        sh.coef_stride = ph_tables(sos, ns, 64, sh.n_blocks, tables);
        break;
    case IIRX_KERNEL_PH_JOINT:
        sh.coef_stride = ph_joint_tables(sos, ns, 64, sh.n_blocks, tables);
        break;
    case IIRX_KERNEL_PH_HALO:
        sh.coef_stride = ph_halo_tables(sos, ns, 64, sh.n_blocks, tables, &sh.halo);
        break;
    case IIRX_KERNEL_STCR:
    case IIRX_KERNEL_DTCR:
        if (cfg->kernel == IIRX_KERNEL_DTCR) sh.threads = 128;  // even and odd sibling waves
        sh.coef_stride = cr_tables(sos, ns, 64, sh.n_blocks, tables);
        if (sh.coef_stride < 0) return IIRX_ERR_CONDITIONING;
        break;
    case IIRX_KERNEL_BLOCKSTATE:
    case IIRX_KERNEL_BLOCKSTATE_BF16:
        sh.coef_stride = bs_tables(sos, ns, cfg->n_blocks, cfg->block_size, tables, bs_split_of(cfg->kernel));
        sh.chunks_per_block = bs_wg_waves(cfg->n_blocks, cfg->block_size, bs_split_of(cfg->kernel));
        sh.threads = 64 * sh.chunks_per_block;
        sh.chunk = (int64_t)cfg->block_size * cfg->n_blocks;
        sh.num_chunks = (uint32_t)((n + sh.chunk - 1) / sh.chunk);
        sh.lb_count = (size_t)sh.num_chunks * 32u;  // 16 partial + 16 full state words
        return IIRX_OK;
    case IIRX_KERNEL_PLR: {
        const int k = 2 * ns;
        plr_geometry(cfg, n, &sh.plr_block, &sh.plr_x);
        sh.threads = sh.plr_block;
        sh.chunk = (int64_t)sh.plr_block * sh.plr_x;
        std::vector<double> b((size_t)k), a((size_t)k);
        if (plr_b && plr_a) {
            b.assign(plr_b, plr_b + k);
            a.assign(plr_a, plr_a + k);
        } else if (iirx_direct_form_taps(sos, ns, b.data(), a.data()) != IIRX_OK) {
            return IIRX_ERR_INVALID;
        }
        // PLR_X_MAX calibration (run_accuracy.sh:95-151 greps ptxas for spills):
        // halve X while the instance needs scratch, from the code object's
        // own resource metadata (only when a device is present)
        if (cfg->plr_x == 0) {
            for (;;) {
                const void *fn = kernel_plr(k, sh.plr_block, sh.plr_x);
                hipFuncAttributes at;
                if (sh.plr_x <= 1 || !fn || hipFuncGetAttributes(&at, fn) != hipSuccess || at.localSizeBytes == 0)
                    break;
                sh.plr_x /= 2;
            }
            (void)hipGetLastError();
            sh.chunk = (int64_t)sh.plr_block * sh.plr_x;
        }
        sh.coef_stride = plr_tables(b.data(), a.data(), k, (int)sh.chunk, tables);
        sh.num_chunks = (uint32_t)((n + sh.chunk - 1) / sh.chunk);
        sh.lb_count = (size_t)sh.num_chunks * 2u * (size_t)k;
        return IIRX_OK;
    }
    default:
        return IIRX_ERR_INVALID;
    }
    sh.chunk = (int64_t)sh.lanes * sh.n_blocks;
    sh.num_chunks = (uint32_t)((n + sh.chunk - 1) / sh.chunk);
    sh.lb_count = (size_t)sh.num_chunks * (size_t)ns * 4u;  // 2 partial + 2 full words per (chunk, section)
    return IIRX_OK;
}

const void *kernel_fn(const iirx_config *cfg, const Shape &sh)
{
    switch (cfg->kernel) {
    case IIRX_KERNEL_PH: return kernel_ph(sh.lanes, sh.n_blocks);
    case IIRX_KERNEL_PH_JOINT: return kernel_ph(sh.lanes, sh.n_blocks, false, true);
    case IIRX_KERNEL_PH_HALO: return kernel_ph(sh.lanes, sh.n_blocks, false, false, sh.halo);
    case IIRX_KERNEL_STCR: return kernel_stcr(sh.lanes, sh.n_blocks, cfg->unrolled);
    case IIRX_KERNEL_DTCR: return kernel_dtcr(sh.lanes, sh.n_blocks, cfg->unrolled);
    case IIRX_KERNEL_PLR: return kernel_plr(2 * cfg->n_sections, sh.plr_block, sh.plr_x);
    case IIRX_KERNEL_BLOCKSTATE:
    case IIRX_KERNEL_BLOCKSTATE_BF16:
        return kernel_bs(cfg->n_blocks, cfg->block_size, cfg->n_sections, bs_split_of(cfg->kernel));
    default: return nullptr;
    }
}

// Workgroups of the main launch.  Persistent schedule: one per resident slot
// of the kernel (occupancy API), at most one per workgroup's worth of chunks.
// Ticket schedule: one per workgroup's worth of chunks, except PH and STCR
// when ph_ticket_waves(N) divides the chunk count: their grouped instance
// loops over groups of that many chunks, one workgroup per resident slot of
// that instance.  DTCR keeps one chunk per workgroup: its two waves meet at
// barriers inside the section loop, where a chunk of the same workgroup
// could be waiting for its record.
int plan_grid(const iirx_config *cfg, const Shape &sh, uint32_t main_chunks, const void *fn, int threads, int dev,
              uint32_t &grid)
{
    const int64_t blocks_needed = ((int64_t)main_chunks + sh.chunks_per_block - 1) / sh.chunks_per_block;
    grid = (uint32_t)blocks_needed;
    const int wv = ph_ticket_waves(sh.n_blocks);
    // PH with halo carries has no inter-chunk wait: one workgroup per resident
    // slot under either schedule, a grid-stride loop with no ticket
    const bool grouped = (is_ph(cfg->kernel) || cfg->kernel == IIRX_KERNEL_STCR) && !sh.halo &&
                         cfg->schedule == IIRX_SCHEDULE_TICKET && main_chunks > 0 && main_chunks % (uint32_t)wv == 0;
    // the block-state form's ticket schedule loops over whole groups of its
    // workgroup's chunks, one workgroup per resident slot, when they divide
    const bool bs_grouped = is_bs(cfg->kernel) && cfg->schedule == IIRX_SCHEDULE_TICKET && main_chunks > 0 &&
                            main_chunks % (uint32_t)sh.chunks_per_block == 0;
    if (cfg->schedule != IIRX_SCHEDULE_PERSISTENT && !grouped && !sh.halo && !bs_grouped) return IIRX_OK;
    const void *occ_fn = fn;
    int occ_threads = threads;
    int64_t want = blocks_needed;
    if (grouped) {
        occ_fn = is_ph(cfg->kernel) ? kernel_ph(sh.lanes, sh.n_blocks, true, cfg->kernel == IIRX_KERNEL_PH_JOINT, sh.halo)
                                    : kernel_stcr(sh.lanes, sh.n_blocks, cfg->unrolled, true);
        if (!occ_fn) return IIRX_ERR_GEOMETRY;
        occ_threads = 64 * wv;
        want = main_chunks / (uint32_t)wv;
    }
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, occ_fn, occ_threads, 0) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return IIRX_ERR_HIP;
    grid = (uint32_t)std::min<int64_t>((int64_t)std::max(per_cu, 1) * std::max(cus, 1), want);
    return IIRX_OK;
}

}  // namespace

extern "C" {

const char *iirx_version(void) { return "iirx 0.1 (gfx950)"; }

const char *iirx_status_string(int s)
{
    switch (s) {
    case IIRX_OK: return "ok";
    case IIRX_ERR_INVALID: return "invalid argument or unsupported configuration";
    case IIRX_ERR_GEOMETRY: return "geometry: no kernel instance for this geometry, or too many chunks";
    case IIRX_ERR_CONDITIONING: return "CR factors ill-conditioned for this filter and N_BLOCKS";
    case IIRX_ERR_HIP: return "HIP runtime error";
    case IIRX_ERR_NO_DEVICE: return "no GPU device";
    case IIRX_ERR_OOM: return "out of memory";
    case IIRX_ERR_TIMEOUT: return "lookback watchdog: a wait exceeded the spin limit (workgroups not co-resident)";
    case IIRX_ERR_TOO_LARGE:
        return "the call needs more device scratch than iirx_execute_state_noalloc holds (IIRX_NOALLOC_SCRATCH_BYTES)";
    case IIRX_ERR_ACCURACY: return "autotune: output fails the accuracy screen against the FP64 cascade";
    default: return "unknown status";
    }
}

int iirx_config_check(const iirx_config *cfg, int64_t n_samples) { return validate(cfg, n_samples); }

int iirx_plan_create(const iirx_config *cfg, const float *sos, int64_t n_samples, iirx_plan **out)
{
    if (!out || !sos) return IIRX_ERR_INVALID;
    *out = nullptr;
    int st = validate(cfg, n_samples);
    if (st != IIRX_OK) return st;
    if (!sos_ok(sos, cfg->n_sections)) return IIRX_ERR_INVALID;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return IIRX_ERR_NO_DEVICE;

    iirx_plan *p = new (std::nothrow) iirx_plan();
    if (!p) return IIRX_ERR_OOM;
    p->cfg = *cfg;
    p->n_samples = n_samples;
    p->sos.assign(sos, sos + 6 * cfg->n_sections);
    p->d_coef = nullptr;
    p->d_lb = nullptr;
    p->d_ticket = nullptr;
    p->d_epoch = nullptr;
    p->d_err = nullptr;
    p->d_arrive = nullptr;
    p->spin_limit = 1u << 22;
    p->launches = 0;
    hipGetDevice(&p->device);

    Shape sh;
    st = make_shape(cfg, p->sos.data(), n_samples, nullptr, nullptr, sh, p->host_coef);
    if (st != IIRX_OK) {
        delete p;
        return st;
    }
    p->shape = sh;
    p->lanes = sh.lanes;
    p->n_blocks = sh.n_blocks;
    p->threads = sh.threads;
    p->chunk = sh.chunk;
    p->num_chunks = sh.num_chunks;
    p->coef_stride = sh.coef_stride;
    const size_t lb_count = sh.lb_count;

    p->d_bsmap = nullptr;
    p->split_last = (is_ph(cfg->kernel) || cfg->kernel == IIRX_KERNEL_STCR ||
                     cfg->kernel == IIRX_KERNEL_DTCR || is_bs(cfg->kernel)) &&
                    n_samples % sh.chunk != 0;
    p->main_chunks = p->num_chunks - (p->split_last ? 1u : 0u);
    p->d_aux = nullptr;
    p->d_mid = nullptr;
    p->aux_stride = 0;
    auto fail = [&](int code) {
        if (p->d_aux) hipFree(p->d_aux);
        if (p->d_mid) hipFree(p->d_mid);
        if (p->d_coef) hipFree(p->d_coef);
        if (p->d_lb) hipFree(p->d_lb);
        if (p->d_ticket) hipFree(p->d_ticket);
        if (p->d_bsmap) hipFree(p->d_bsmap);
        delete p;
        return code;
    };
    if (hipMalloc((void **)&p->d_coef, p->host_coef.size() * sizeof(float)) != hipSuccess) return fail(IIRX_ERR_OOM);
    if (hipMalloc((void **)&p->d_lb, lb_count * sizeof(uint64_t)) != hipSuccess) return fail(IIRX_ERR_OOM);
    if (hipMalloc((void **)&p->d_ticket, kControlBytes) != hipSuccess) return fail(IIRX_ERR_OOM);
    p->d_epoch = reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(p->d_ticket) + 8);
    p->d_err = p->d_epoch + 1;
    p->d_arrive = reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(p->d_ticket) + 128);
    if (is_bs(cfg->kernel)) {
        // the block-state tables are computed in FP64 on the GPU (bs_setup.hip)
        BsSetupJob job{};
        std::copy(p->sos.begin(), p->sos.end(), job.sos);
        job.ns = cfg->n_sections;
        BsSetupJob *d_job = nullptr;
        if (hipMalloc((void **)&d_job, sizeof job) != hipSuccess) return fail(IIRX_ERR_OOM);
        const bool ok = hipMemcpy(d_job, &job, sizeof job, hipMemcpyHostToDevice) == hipSuccess &&
                        launch_bs_setup(cfg->n_blocks, cfg->block_size, d_job, 1, p->d_coef, nullptr,
                                        bs_split_of(cfg->kernel)) == hipSuccess &&
                        hipDeviceSynchronize() == hipSuccess;
        hipFree(d_job);
        if (!ok) return fail(IIRX_ERR_HIP);
        // the streaming-state map (FP64 realization of precompute.cpp)
        BsStateMap map{};
        std::vector<double> A, B, C;
        bs_realization(p->sos.data(), cfg->n_sections, A, B, C, map.d);
        for (int s = 0; s < cfg->n_sections; ++s) {
            const float *r = p->sos.data() + 6 * s;
            map.c[s][0] = r[1];
            map.c[s][1] = r[2];
            map.c[s][2] = r[4];
            map.c[s][3] = r[5];
        }
        map.ns = cfg->n_sections;
        if (hipMalloc((void **)&p->d_bsmap, sizeof map + 32 * sizeof(float)) != hipSuccess) return fail(IIRX_ERR_OOM);
        if (hipMemcpy(p->d_bsmap, &map, sizeof map, hipMemcpyHostToDevice) != hipSuccess) return fail(IIRX_ERR_HIP);
        if (p->split_last) {
            // the partial last chunk (fewer than G L <= 4096 samples) runs the PH
            // last-chunk kernel at 64 x 64
            std::vector<float> aux;
            p->aux_stride = ph_tables(p->sos.data(), cfg->n_sections, 64, 64, aux);
            if (hipMalloc((void **)&p->d_aux, aux.size() * sizeof(float)) != hipSuccess) return fail(IIRX_ERR_OOM);
            if (hipMalloc((void **)&p->d_mid, (2 + 2 * cfg->n_sections) * sizeof(float)) != hipSuccess)
                return fail(IIRX_ERR_OOM);
            if (hipMemcpy(p->d_aux, aux.data(), aux.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
                return fail(IIRX_ERR_HIP);
        }
    } else if (hipMemcpy(p->d_coef, p->host_coef.data(), p->host_coef.size() * sizeof(float),
                         hipMemcpyHostToDevice) != hipSuccess) {
        return fail(IIRX_ERR_HIP);
    }
    // zeroed once; launch-versioned flags never need a reset afterwards
    if (hipMemset(p->d_lb, 0, lb_count * sizeof(uint64_t)) != hipSuccess) return fail(IIRX_ERR_HIP);
    if (hipMemset(p->d_ticket, 0, kControlBytes) != hipSuccess) return fail(IIRX_ERR_HIP);
    // grid: one workgroup per chunk (ticket) or per resident slot (persistent)
    p->fn = kernel_fn(cfg, sh);
    if (!p->fn) return fail(IIRX_ERR_GEOMETRY);
    // workgroups: one per chunk (ticket) or one per resident slot (persistent);
    // a block-state workgroup runs chunks_per_block waves, each its own chunks
    const int st_grid = plan_grid(cfg, sh, p->main_chunks, p->fn, p->threads, p->device, p->grid);
    if (st_grid != IIRX_OK) return fail(st_grid);
    if (hipDeviceSynchronize() != hipSuccess) return fail(IIRX_ERR_HIP);
    *out = p;
    return IIRX_OK;
}

int iirx_plan_tables(const iirx_config *cfg, const float *sos, int64_t n_samples, float *out, int64_t cap,
                     int64_t *size, int32_t *stride)
{
    if (!sos) return IIRX_ERR_INVALID;
    int st = validate(cfg, n_samples);
    if (st != IIRX_OK) return st;
    if (!sos_ok(sos, cfg->n_sections)) return IIRX_ERR_INVALID;
    std::vector<float> t;
    Shape sh;
    st = make_shape(cfg, sos, n_samples, nullptr, nullptr, sh, t);
    if (st != IIRX_OK) return st;
    if (size) *size = (int64_t)t.size();
    if (stride) *stride = sh.coef_stride;
    if (out) std::memcpy(out, t.data(), sizeof(float) * (size_t)std::min<int64_t>(cap, (int64_t)t.size()));
    return IIRX_OK;
}

int iirx_plan_read_tables(const iirx_plan *p, float *out, int64_t cap, int64_t *size)
{
    if (!p) return IIRX_ERR_INVALID;
    const int64_t n = (int64_t)p->host_coef.size();
    if (size) *size = n;
    if (out && cap > 0 &&
        hipMemcpy(out, p->d_coef, sizeof(float) * (size_t)std::min(cap, n), hipMemcpyDeviceToHost) != hipSuccess)
        return IIRX_ERR_HIP;
    return IIRX_OK;
}

int iirx_plan_set_plr_taps(iirx_plan *p, const double *b, const double *a, int k)
{
    if (!p || !b || !a) return IIRX_ERR_INVALID;
    if (p->cfg.kernel != IIRX_KERNEL_PLR || k != 2 * p->cfg.n_sections) return IIRX_ERR_INVALID;
    for (int j = 0; j < k; ++j)
        if (!std::isfinite(b[j]) || !std::isfinite(a[j])) return IIRX_ERR_INVALID;
    std::vector<float> t;
    Shape sh;
    const int st = make_shape(&p->cfg, p->sos.data(), p->n_samples, b, a, sh, t);
    if (st != IIRX_OK) return st;
    if (hipDeviceSynchronize() != hipSuccess) return IIRX_ERR_HIP;  // no launch may read the old tables
    if (hipMemcpy(p->d_coef, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
        return IIRX_ERR_HIP;
    p->host_coef.swap(t);
    p->plr_b.assign(b, b + k);
    p->plr_a.assign(a, a + k);
    return IIRX_OK;
}

static int launch(iirx_plan *p, const float *d_in, float *d_out, hipStream_t s, const float *d_zi = nullptr,
                  float *d_zf = nullptr)
{
    LaunchArgs a{};
    a.in = d_in;
    a.out = d_out;
    a.coef = p->d_coef;
    a.lb = p->d_lb;
    a.ticket = p->d_ticket;
    a.epoch = p->d_epoch;
    a.err = p->d_err;
    a.arrive = p->d_arrive;
    a.n_samples = p->n_samples;
    a.num_chunks = p->main_chunks;
    a.spin_limit = p->spin_limit;
    a.persistent = p->cfg.schedule == IIRX_SCHEDULE_PERSISTENT;
    a.grid = p->grid;
    a.ns = (p->cfg.kernel == IIRX_KERNEL_PLR) ? 2 * p->cfg.n_sections : p->cfg.n_sections;
    a.coef_stride = p->coef_stride;
    a.unrolled = p->cfg.unrolled;
    a.zi = d_zi;
    a.tail = nullptr;
    a.last_valid = (int32_t)(p->n_samples - (int64_t)p->main_chunks * p->chunk);
    hipError_t e = hipSuccess;
    float *bs_z0 = nullptr, *bs_tail = nullptr;
    if (is_bs(p->cfg.kernel) && p->d_bsmap) {
        // the block-state kernel takes the balanced state; it writes the state
        // entering its last tile for the histories out
        bs_z0 = reinterpret_cast<float *>(p->d_bsmap + 1);
        bs_tail = bs_z0 + 16;
        if (d_zi) {
            e = launch_bs_state_in(d_zi, p->d_bsmap, bs_z0, s);
            a.zi = bs_z0;
        }
        // a partial last chunk starts from the histories at the end of the whole chunks
        if (d_zf || p->split_last) a.tail = bs_tail;
    }
    if (e != hipSuccess) return IIRX_ERR_HIP;
    e = hipErrorInvalidValue;
    switch (p->cfg.kernel) {
    case IIRX_KERNEL_PH:
    case IIRX_KERNEL_PH_JOINT:
    case IIRX_KERNEL_PH_HALO:
    case IIRX_KERNEL_STCR:
    case IIRX_KERNEL_DTCR: {
        // the main launch, then the partial last chunk (which records zf)
        auto run = [&](const LaunchArgs &la, bool last) {
            if (is_ph(p->cfg.kernel))
                return launch_ph(p->lanes, p->n_blocks, la, s, last, p->cfg.kernel == IIRX_KERNEL_PH_JOINT,
                                 p->shape.halo);
            if (p->cfg.kernel == IIRX_KERNEL_STCR) return launch_stcr(p->lanes, p->n_blocks, la, s, last);
            return launch_dtcr(p->lanes, p->n_blocks, la, s, last);
        };
        e = p->main_chunks > 0 ? run(a, false) : hipSuccess;
        if (e == hipSuccess && p->split_last) {
            LaunchArgs b = a;
            b.tail = d_zf;
            e = run(b, true);
        }
        break;
    }
    case IIRX_KERNEL_PLR: e = launch_plr(2 * p->cfg.n_sections, p->shape.plr_block, p->shape.plr_x, a, s); break;
    case IIRX_KERNEL_BLOCKSTATE:
    case IIRX_KERNEL_BLOCKSTATE_BF16: {
        e = p->main_chunks > 0 ? launch_bs(p->n_blocks, p->cfg.block_size, p->cfg.n_sections, a, s,
                                           bs_split_of(p->cfg.kernel))
                               : hipSuccess;
        if (e == hipSuccess && p->split_last) {
            // the partial last chunk: direct-form histories after the whole
            // chunks (the state-out map), then the PH last-chunk kernel from them
            const int64_t n_main = (int64_t)p->main_chunks * p->chunk;
            const float *hist = d_zi;
            if (p->main_chunks > 0) {
                e = launch_bs_state_out(d_in, n_main, p->cfg.n_blocks, bs_tail, p->d_bsmap, p->d_mid, s);
                hist = p->d_mid;
            }
            LaunchArgs b = a;
            b.in = d_in + n_main;
            b.out = d_out + n_main;
            b.coef = p->d_aux;
            b.coef_stride = p->aux_stride;
            b.num_chunks = 0;  // the chunk at b.in
            b.last_valid = (int32_t)(p->n_samples - n_main);
            b.zi = hist;
            b.tail = d_zf;
            if (e == hipSuccess) e = launch_ph(64, 64, b, s, true);
        }
        break;
    }
    default: break;
    }
    if (e == hipSuccess && d_zf) {
        if (p->cfg.kernel == IIRX_KERNEL_PLR) e = launch_plr_end_state(a, d_zf, s);
        else if (is_bs(p->cfg.kernel) && !p->split_last)
            e = launch_bs_state_out(d_in, p->n_samples, p->cfg.n_blocks, bs_tail, p->d_bsmap, d_zf, s);
        else if (!p->split_last) e = launch_ph_end_state(a, (int)p->chunk, d_zf, s);
    }
    if (e != hipSuccess) return IIRX_ERR_HIP;
    p->launches++;
    return IIRX_OK;
}

// 16-byte accesses need 16-byte aligned buffers; PLR moves one float at a time
static bool io_aligned(const iirx_plan *p, const float *d_in, const float *d_out)
{
    return p->cfg.kernel == IIRX_KERNEL_PLR || !(((uintptr_t)d_in & 15) || ((uintptr_t)d_out & 15));
}

int iirx_execute(iirx_plan *p, const float *d_in, float *d_out, void *stream)
{
    if (!p || !d_in || !d_out) return IIRX_ERR_INVALID;
    if (!io_aligned(p, d_in, d_out)) return IIRX_ERR_INVALID;
    const uintptr_t ib = (uintptr_t)d_in, ie = ib + (uintptr_t)p->n_samples * 4;
    const uintptr_t ob = (uintptr_t)d_out, oe = ob + (uintptr_t)p->n_samples * 4;
    if (ib < oe && ob < ie) return IIRX_ERR_INVALID;  // overlap
    return launch(p, d_in, d_out, (hipStream_t)stream);
}

int iirx_execute_state(iirx_plan *p, const float *d_in, float *d_out, const float *d_zi, float *d_zf, void *stream)
{
    if (!p || !d_in || !d_out) return IIRX_ERR_INVALID;
    // streaming state: the sections' direct-form histories (PH, STCR, DTCR,
    // the block-state form through its state map), PLR's of the whole recurrence
    const int state_floats = (p->cfg.kernel == IIRX_KERNEL_PLR) ? 4 * p->cfg.n_sections : 2 + 2 * p->cfg.n_sections;
    if (!io_aligned(p, d_in, d_out)) return IIRX_ERR_INVALID;
    const uintptr_t ib = (uintptr_t)d_in, ie = ib + (uintptr_t)p->n_samples * 4;
    const uintptr_t ob = (uintptr_t)d_out, oe = ob + (uintptr_t)p->n_samples * 4;
    if (ib < oe && ob < ie) return IIRX_ERR_INVALID;  // overlap
    if (d_zf) {
        const uintptr_t zb = (uintptr_t)d_zf, ze = zb + (uintptr_t)state_floats * 4;
        if ((zb < ie && ib < ze) || (zb < oe && ob < ze)) return IIRX_ERR_INVALID;
    }
    return launch(p, d_in, d_out, (hipStream_t)stream, d_zi, d_zf);
}

int iirx_plan_info_get(const iirx_plan *p, iirx_plan_info *info)
{
    if (!p || !info) return IIRX_ERR_INVALID;
    std::memset(info, 0, sizeof *info);
    info->n_samples = p->n_samples;
    info->chunk_samples = p->chunk;
    info->num_chunks = p->num_chunks;
    info->threads_per_block = p->threads;
    info->launches = p->launches;
    info->grid = p->grid;
    info->schedule = p->cfg.schedule;
    info->halo = p->shape.halo ? 1 : 0;
    const void *fn = p->fn;
    if (fn) {
        hipFuncAttributes at;
        if (hipFuncGetAttributes(&at, fn) == hipSuccess) {
            info->lds_bytes = (int)at.sharedSizeBytes;
            info->vgprs = at.numRegs;
            info->scratch_bytes = (int)at.localSizeBytes;
        }
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, p->threads, 0) == hipSuccess) {
            const int waves_per_block = (p->threads + 63) / 64;
            info->waves_per_simd = nb * waves_per_block / 4;
        }
    }
    return IIRX_OK;
}

int iirx_plan_check(iirx_plan *p)
{
    if (!p) return IIRX_ERR_INVALID;
    if (hipDeviceSynchronize() != hipSuccess) return IIRX_ERR_HIP;
    uint32_t err = 0;
    if (hipMemcpy(&err, p->d_err, sizeof err, hipMemcpyDeviceToHost) != hipSuccess) return IIRX_ERR_HIP;
    if (err) {
        const uint32_t zero = 0;
        if (hipMemcpy(p->d_err, &zero, sizeof zero, hipMemcpyHostToDevice) != hipSuccess) return IIRX_ERR_HIP;
        return IIRX_ERR_TIMEOUT;
    }
    return IIRX_OK;
}

int iirx_plan_set_spin_limit(iirx_plan *p, uint32_t polls)
{
    if (!p || polls == 0) return IIRX_ERR_INVALID;
    p->spin_limit = polls;
    return IIRX_OK;
}

int iirx_plan_destroy(iirx_plan *p)
{
    if (!p) return IIRX_ERR_INVALID;
    hipDeviceSynchronize();
    if (p->d_coef) hipFree(p->d_coef);
    if (p->d_lb) hipFree(p->d_lb);
    if (p->d_ticket) hipFree(p->d_ticket);
    if (p->d_bsmap) hipFree(p->d_bsmap);
    if (p->d_aux) hipFree(p->d_aux);
    if (p->d_mid) hipFree(p->d_mid);
    delete p;
    return IIRX_OK;
}

int iirx_measure(iirx_plan *p, const float *const *d_inputs, int n_inputs, float *d_out, int warmup, int batches,
                 int batch, void *stream, float *batch_ms)
{
    if (!p || !d_inputs || n_inputs < 1 || !d_out || warmup < 0 || batches < 1 || batch < 1 || !batch_ms)
        return IIRX_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    for (int i = 0; i < warmup; ++i) {
        int st = iirx_execute(p, d_inputs[i % n_inputs], d_out, stream);
        if (st != IIRX_OK) return st;
    }
    if (hipStreamSynchronize(s) != hipSuccess) return IIRX_ERR_HIP;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return IIRX_ERR_HIP;
    if (hipEventCreate(&e1) != hipSuccess) { hipEventDestroy(e0); return IIRX_ERR_HIP; }
    int it = 0, st = IIRX_OK;
    for (int b = 0; b < batches && st == IIRX_OK; ++b) {
        hipEventRecord(e0, s);
        for (int i = 0; i < batch; ++i, ++it) {
            st = iirx_execute(p, d_inputs[it % n_inputs], d_out, stream);
            if (st != IIRX_OK) break;
        }
        hipEventRecord(e1, s);
        if (hipEventSynchronize(e1) != hipSuccess) { st = IIRX_ERR_HIP; break; }
        float ms = 0.0f;
        hipEventElapsedTime(&ms, e0, e1);
        batch_ms[b] = ms / (float)batch;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (st != IIRX_OK) return st;
    std::sort(batch_ms, batch_ms + batches);
    return IIRX_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// iirx_execute_state_noalloc: a plan's call with the plan's device memory
// (coefficient tables, lookback records, control block) in one file-scope
// device array instead of hipMalloc'd buffers, for callers that cannot
// allocate.  Layout of the array: [control block | tables | records].
// ---------------------------------------------------------------------------
namespace {

constexpr size_t kNoallocCtl = (kControlBytes + 255) & ~(size_t)255;
constexpr size_t kNoallocTables = 64u << 10;
static_assert(kNoallocCtl + kNoallocTables < IIRX_NOALLOC_SCRATCH_BYTES, "scratch layout");

static __device__ __align__(256) unsigned char scratch[IIRX_NOALLOC_SCRATCH_BYTES];

// the array's address on each device (each GPU holds its own copy); the
// array's name in host code is a host-side stand-in, never a device address
constexpr int kMaxDevices = 64;
std::once_flag g_scratch_once[kMaxDevices];
unsigned char *g_scratch_ptr[kMaxDevices];
hipError_t g_scratch_err[kMaxDevices];

unsigned char *scratch_on(int dev)
{
    std::call_once(g_scratch_once[dev], [dev]() {
        void *ptr = nullptr;
        g_scratch_err[dev] = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        g_scratch_ptr[dev] = g_scratch_err[dev] == hipSuccess ? static_cast<unsigned char *>(ptr) : nullptr;
    });
    return g_scratch_ptr[dev];
}

// the tables reach the device as kernel arguments (copied at launch, so the
// host buffer may go away at once and no staging copy is involved)
constexpr int kPieceFloats = 1000;
struct TablePiece {
    float v[kPieceFloats];
};

__global__ __launch_bounds__(256) void noalloc_put_tables(float *dst, TablePiece piece, int count)
{
    for (int i = threadIdx.x; i < count; i += blockDim.x) dst[i] = piece.v[i];
}

// zero the control block and the records this call uses (a fill kernel, not a
// memset node: the call may be captured into a graph)
__global__ __launch_bounds__(256) void noalloc_clear(uint64_t *p, size_t words)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

}  // namespace

extern "C" int iirx_execute_state_noalloc(const iirx_config *cfg, const float *sos, int64_t n_samples,
                                          const float *d_in, float *d_out, const float *d_zi, float *d_zf,
                                          void *stream)
{
    if (!cfg || !sos || !d_in || !d_out) return IIRX_ERR_INVALID;
    if (!(is_ph(cfg->kernel) || cfg->kernel == IIRX_KERNEL_STCR || cfg->kernel == IIRX_KERNEL_DTCR))
        return IIRX_ERR_INVALID;
    int st = validate(cfg, n_samples);
    if (st != IIRX_OK) return st;
    if (!sos_ok(sos, cfg->n_sections)) return IIRX_ERR_INVALID;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return IIRX_ERR_NO_DEVICE;
    if (dev < 0 || dev >= kMaxDevices) return IIRX_ERR_INVALID;
    unsigned char *base = scratch_on(dev);
    if (!base) return IIRX_ERR_HIP;

    // the plan of this call, on the host stack, its device memory in the array
    iirx_plan p;
    p.cfg = *cfg;
    p.n_samples = n_samples;
    p.device = dev;
    p.sos.assign(sos, sos + 6 * cfg->n_sections);
    Shape sh;
    st = make_shape(cfg, p.sos.data(), n_samples, nullptr, nullptr, sh, p.host_coef);
    if (st != IIRX_OK) return st;
    const size_t table_bytes = p.host_coef.size() * sizeof(float);
    const size_t lb_bytes = sh.lb_count * sizeof(uint64_t);
    if (table_bytes > kNoallocTables || kNoallocCtl + kNoallocTables + lb_bytes > IIRX_NOALLOC_SCRATCH_BYTES)
        return IIRX_ERR_TOO_LARGE;
    p.shape = sh;
    p.lanes = sh.lanes;
    p.n_blocks = sh.n_blocks;
    p.threads = sh.threads;
    p.chunk = sh.chunk;
    p.num_chunks = sh.num_chunks;
    p.coef_stride = sh.coef_stride;
    p.d_ticket = reinterpret_cast<unsigned long long *>(base);
    p.d_epoch = reinterpret_cast<uint32_t *>(base + 8);
    p.d_err = p.d_epoch + 1;
    p.d_arrive = reinterpret_cast<uint32_t *>(base + 128);
    p.d_coef = reinterpret_cast<float *>(base + kNoallocCtl);
    p.d_lb = reinterpret_cast<uint64_t *>(base + kNoallocCtl + kNoallocTables);
    p.d_bsmap = nullptr;
    p.d_aux = nullptr;
    p.d_mid = nullptr;
    p.aux_stride = 0;
    p.spin_limit = 1u << 22;
    p.launches = 0;
    p.split_last = n_samples % sh.chunk != 0;
    p.main_chunks = p.num_chunks - (p.split_last ? 1u : 0u);
    p.fn = kernel_fn(cfg, sh);
    if (!p.fn) return IIRX_ERR_GEOMETRY;
    st = plan_grid(cfg, sh, p.main_chunks, p.fn, p.threads, dev, p.grid);
    if (st != IIRX_OK) return st;

    // the same argument checks as iirx_execute_state
    if (!io_aligned(&p, d_in, d_out)) return IIRX_ERR_INVALID;
    const uintptr_t ib = (uintptr_t)d_in, ie = ib + (uintptr_t)n_samples * 4;
    const uintptr_t ob = (uintptr_t)d_out, oe = ob + (uintptr_t)n_samples * 4;
    if (ib < oe && ob < ie) return IIRX_ERR_INVALID;
    if (d_zf) {
        const uintptr_t zb = (uintptr_t)d_zf, ze = zb + (uintptr_t)(2 + 2 * cfg->n_sections) * 4;
        if ((zb < ie && ib < ze) || (zb < oe && ob < ze)) return IIRX_ERR_INVALID;
    }

    // the array keeps the last call's contents: control block and records
    // start from zero (epoch 0, no flag set), then the tables, then the call
    hipStream_t s = (hipStream_t)stream;
    const size_t words = (kNoallocCtl + kNoallocTables + lb_bytes) / sizeof(uint64_t);
    const unsigned clear_blocks = (unsigned)std::min<size_t>(1024, (words + 255) / 256);
    hipLaunchKernelGGL(noalloc_clear, dim3(clear_blocks), dim3(256), 0, s, reinterpret_cast<uint64_t *>(base), words);
    if (hipGetLastError() != hipSuccess) return IIRX_ERR_HIP;
    const int total = (int)p.host_coef.size();
    for (int off = 0; off < total; off += kPieceFloats) {
        TablePiece piece;
        const int cnt = std::min(kPieceFloats, total - off);
        std::memcpy(piece.v, p.host_coef.data() + off, sizeof(float) * (size_t)cnt);
        hipLaunchKernelGGL(noalloc_put_tables, dim3(1), dim3(256), 0, s, p.d_coef + off, piece, cnt);
        if (hipGetLastError() != hipSuccess) return IIRX_ERR_HIP;
    }
    return launch(&p, d_in, d_out, s, d_zi, d_zf);
}
