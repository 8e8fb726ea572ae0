// iirx_batch.cpp -- the ragged-batch part of the C ABI (include/iirx.h):
// channel validation, per-channel FP64 tables, the chunk -> channel map, the
// once-zeroed lookback state, the persistent launch and its hipGraph capture.
//
// No counterpart in the reference, which filters one channel per binary; the
// per-channel PH tables are those of setup_kernel_coefficients
// (ph_kernels.cuh:151-200), built by ph_tables on the host; block-state
// channels' tables (tiles of 64 samples, bs_tiles tiles per chunk) are
// computed in FP64 on the GPU by one bs_setup launch over those channels.
#include "../../include/iirx.h"
#include "kernels.hpp"
#include "precompute.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

using namespace iirx;

// one launch per form present: PH chunks [0, base_bs), block-state chunks
// [base_bs, num_chunks), each with its own control block
struct BatchGroup {
    int algo, n;  // kernel_batch(algo, n)
    uint32_t base, num_chunks, grid;
    const void *fn;
    unsigned long long *d_ctrl;  // [ticket u64][epoch u32][err u32], arrival slots from byte 128
};

struct iirx_batch {
    iirx_batch_config cfg;
    int64_t n_channels, n_samples;
    uint32_t num_chunks, spin_limit;
    int coef_stride, device;
    int64_t n_blockstate, coef_size;
    BatchGroup grp[2];  // [0] PH, [1] block-state
    float *d_coef;
    uint64_t *d_lb;
    BatchChannel *d_chan;
    uint32_t *d_chunk_chan;
    hipGraph_t graph;
    hipGraphExec_t exec;
    int64_t launches;
};

namespace {

void release(iirx_batch *b)
{
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    if (b->d_coef) hipFree(b->d_coef);
    if (b->d_lb) hipFree(b->d_lb);
    if (b->d_chan) hipFree(b->d_chan);
    if (b->d_chunk_chan) hipFree(b->d_chunk_chan);
    for (BatchGroup &g : b->grp)
        if (g.d_ctrl) hipFree(g.d_ctrl);
    delete b;
}

BatchArgs args(const iirx_batch *b, const BatchGroup &g, const float *d_in, float *d_out)
{
    BatchArgs a{};
    a.in = d_in;
    a.out = d_out;
    a.coef = b->d_coef;
    a.lb = b->d_lb;
    a.chan = b->d_chan;
    a.chunk_chan = b->d_chunk_chan;
    a.ticket = g.d_ctrl;
    a.epoch = reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(g.d_ctrl) + 8);
    a.err = a.epoch + 1;
    a.arrive = reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(g.d_ctrl) + 128);
    a.num_chunks = g.num_chunks;
    a.spin_limit = b->spin_limit;
    a.persistent = b->cfg.schedule == IIRX_SCHEDULE_PERSISTENT;
    a.grid = g.grid;
    a.coef_stride = b->coef_stride;
    a.chunk_base = g.base;
    a.vec16 = (((uintptr_t)d_in | (uintptr_t)d_out) & 15u) == 0;
    return a;
}

// every form's launch, in chunk order, on one stream
hipError_t launch_all(const iirx_batch *b, const float *d_in, float *d_out, hipStream_t s)
{
    for (const BatchGroup &g : b->grp) {
        if (g.num_chunks == 0) continue;
        const hipError_t e = launch_batch(g.algo, g.n, args(b, g, d_in, d_out), s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace

extern "C" {

int iirx_batch_create(const iirx_batch_config *cfg, int64_t n_channels, const int64_t *offsets,
                      const int32_t *n_sections, const float *sos, iirx_batch **out)
{
    return iirx_batch_create_ex(cfg, n_channels, offsets, n_sections, sos, nullptr, out);
}

int iirx_batch_create_ex(const iirx_batch_config *cfg, int64_t n_channels, const int64_t *offsets,
                         const int32_t *n_sections, const float *sos, const int32_t *algo, iirx_batch **out)
{
    if (!cfg || !offsets || !n_sections || !sos || !out || n_channels < 1) return IIRX_ERR_INVALID;
    *out = nullptr;
    const int N = cfg->n_blocks == 0 ? 32 : cfg->n_blocks;
    const int G = cfg->bs_tiles == 0 ? 32 : cfg->bs_tiles;
    if (!(N == 16 || N == 32 || N == 64) || !(G == 16 || G == 32 || G == 64)) return IIRX_ERR_GEOMETRY;
    if (cfg->schedule != IIRX_SCHEDULE_PERSISTENT && cfg->schedule != IIRX_SCHEDULE_TICKET) return IIRX_ERR_INVALID;
    if (offsets[0] < 0 || cfg->bs_min_sections < 0) return IIRX_ERR_INVALID;
    int64_t total_sec = 0;
    for (int64_t c = 0; c < n_channels; ++c) {
        if (offsets[c + 1] < offsets[c]) return IIRX_ERR_INVALID;
        if (n_sections[c] < 1 || n_sections[c] > kBatchMaxSections) return IIRX_ERR_INVALID;
        if (algo && (algo[c] < IIRX_BATCH_AUTO || algo[c] > IIRX_BATCH_BLOCKSTATE)) return IIRX_ERR_INVALID;
        total_sec += n_sections[c];
    }
    for (int64_t i = 0; i < 6 * total_sec; ++i)
        if (!std::isfinite(sos[i])) return IIRX_ERR_INVALID;
    for (int64_t s = 0; s < total_sec; ++s)  // b0 == a0 == 1 (iirx_normalize_sos)
        if (sos[6 * s] != 1.0f || sos[6 * s + 3] != 1.0f) return IIRX_ERR_INVALID;
    if (offsets[n_channels] == offsets[0]) return IIRX_ERR_INVALID;  // no samples at all
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return IIRX_ERR_NO_DEVICE;

    // per-channel tables; chunk list: the PH channels' chunks channel after
    // channel, then the block-state channels'.  AUTO channels of order 16 or
    // more take the block-state form by default: in the batch it measured
    // 1.38x PH at order 16 and 0.86x at order 8 (Butterworth, 2^18-sample
    // channels).
    const int min_bs = cfg->bs_min_sections == 0 ? kBatchMaxSections : cfg->bs_min_sections;
    const int64_t CHP = 64LL * N, CHB = 64LL * G;
    std::vector<float> coef;
    std::vector<BatchChannel> chan((size_t)n_channels);
    std::vector<uint32_t> chunks[2];
    std::vector<char> is_bs((size_t)n_channels, 0);
    std::vector<BsSetupJob> jobs;
    int stride = 0;
    int64_t sec0 = 0, n_bs = 0;
    for (int64_t c = 0; c < n_channels; ++c) {
        std::vector<float> t;
        const int ns = n_sections[c];
        const int32_t want = algo ? algo[c] : IIRX_BATCH_AUTO;
        const bool bs = want == IIRX_BATCH_BLOCKSTATE ||
                        (want == IIRX_BATCH_AUTO && ns >= min_bs);
        if (bs) {
            BsSetupJob j{};
            std::copy(sos + 6 * sec0, sos + 6 * (sec0 + ns), j.sos);
            j.ns = ns;
            j.out_off = (int64_t)coef.size();
            jobs.push_back(j);
            t.assign((size_t)bs_layout(64).stride, 0.0f);  // filled on the device
            ++n_bs;
        } else {
            const int st = ph_tables(sos + 6 * sec0, ns, 64, N, t);
            if (st <= 0) return IIRX_ERR_INVALID;
            stride = st;
        }
        sec0 += ns;
        t.resize((t.size() + 3) & ~(size_t)3, 0.0f);  // 16-byte aligned tables
        const int64_t len = offsets[c + 1] - offsets[c];
        const int64_t CH = bs ? CHB : CHP;
        const int64_t k = (len + CH - 1) / CH;
        if ((int64_t)(chunks[0].size() + chunks[1].size()) + k > 0x7fffffffLL ||
            coef.size() + t.size() > 0x7fffffffULL)
            return IIRX_ERR_GEOMETRY;
        BatchChannel &d = chan[(size_t)c];
        d.offset = offsets[c];
        d.length = len;
        d.ns = ns;
        d.coef_off = (int32_t)coef.size();
        d.first_chunk = (uint32_t)chunks[bs].size();  // within its form's list until merged
        d.algo = bs ? kBatchBS : kBatchPH;
        is_bs[(size_t)c] = bs;
        coef.insert(coef.end(), t.begin(), t.end());
        chunks[bs].insert(chunks[bs].end(), (size_t)k, (uint32_t)c);
    }
    const uint32_t n_ph = (uint32_t)chunks[0].size();
    for (int64_t c = 0; c < n_channels; ++c)
        if (is_bs[(size_t)c]) chan[(size_t)c].first_chunk += n_ph;
    std::vector<uint32_t> chunk_chan(chunks[0]);
    chunk_chan.insert(chunk_chan.end(), chunks[1].begin(), chunks[1].end());
    if (chunk_chan.empty()) return IIRX_ERR_INVALID;

    iirx_batch *b = new (std::nothrow) iirx_batch();
    if (!b) return IIRX_ERR_OOM;
    b->cfg = *cfg;
    b->cfg.n_blocks = N;
    b->cfg.bs_tiles = G;
    b->cfg.bs_min_sections = min_bs;
    b->n_channels = n_channels;
    b->n_samples = offsets[n_channels] - offsets[0];
    b->num_chunks = (uint32_t)chunk_chan.size();
    b->coef_stride = stride;
    b->n_blockstate = n_bs;
    b->spin_limit = 1u << 22;
    hipGetDevice(&b->device);
    b->grp[0] = BatchGroup{kBatchPH, N, 0u, n_ph, 0u, kernel_batch(kBatchPH, N), nullptr};
    b->grp[1] = BatchGroup{kBatchBS, G, n_ph, (uint32_t)chunks[1].size(), 0u, kernel_batch(kBatchBS, G), nullptr};
    const size_t lb_count = (size_t)b->num_chunks * kBatchMaxSections * 4u;
    if (hipMalloc((void **)&b->d_coef, coef.size() * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&b->d_lb, lb_count * sizeof(uint64_t)) != hipSuccess ||
        hipMalloc((void **)&b->d_chan, chan.size() * sizeof(BatchChannel)) != hipSuccess ||
        hipMalloc((void **)&b->d_chunk_chan, chunk_chan.size() * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc((void **)&b->grp[0].d_ctrl, kControlBytes) != hipSuccess ||
        hipMalloc((void **)&b->grp[1].d_ctrl, kControlBytes) != hipSuccess) {
        release(b);
        return IIRX_ERR_OOM;
    }
    if (hipMemcpy(b->d_coef, coef.data(), coef.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->d_chan, chan.data(), chan.size() * sizeof(BatchChannel), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->d_chunk_chan, chunk_chan.data(), chunk_chan.size() * sizeof(uint32_t), hipMemcpyHostToDevice) !=
            hipSuccess ||
        // zeroed once; launch-versioned flags never need a reset afterwards
        hipMemset(b->d_lb, 0, lb_count * sizeof(uint64_t)) != hipSuccess ||
        hipMemset(b->grp[0].d_ctrl, 0, kControlBytes) != hipSuccess ||
        hipMemset(b->grp[1].d_ctrl, 0, kControlBytes) != hipSuccess) {
        release(b);
        return IIRX_ERR_HIP;
    }
    b->coef_size = (int64_t)coef.size();
    if (!jobs.empty()) {
        BsSetupJob *d_jobs = nullptr;
        if (hipMalloc((void **)&d_jobs, jobs.size() * sizeof(BsSetupJob)) != hipSuccess) {
            release(b);
            return IIRX_ERR_OOM;
        }
        const bool ok =
            hipMemcpy(d_jobs, jobs.data(), jobs.size() * sizeof(BsSetupJob), hipMemcpyHostToDevice) == hipSuccess &&
            launch_bs_setup(64, G, d_jobs, (int)jobs.size(), b->d_coef, nullptr) == hipSuccess &&
            hipDeviceSynchronize() == hipSuccess;
        hipFree(d_jobs);
        if (!ok) {
            release(b);
            return IIRX_ERR_HIP;
        }
    }
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device) != hipSuccess) {
        release(b);
        return IIRX_ERR_HIP;
    }
    for (BatchGroup &g : b->grp) {
        g.grid = g.num_chunks;
        if (cfg->schedule == IIRX_SCHEDULE_PERSISTENT && g.num_chunks > 0) {
            int per_cu = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, g.fn, 64, 0) != hipSuccess) {
                release(b);
                return IIRX_ERR_HIP;
            }
            g.grid = (uint32_t)std::min<int64_t>((int64_t)std::max(per_cu, 1) * std::max(cus, 1), g.num_chunks);
        }
    }
    if (hipDeviceSynchronize() != hipSuccess) {
        release(b);
        return IIRX_ERR_HIP;
    }
    *out = b;
    return IIRX_OK;
}

int iirx_batch_execute(iirx_batch *b, const float *d_in, float *d_out, void *stream)
{
    if (!b || !d_in || !d_out) return IIRX_ERR_INVALID;
    if (launch_all(b, d_in, d_out, (hipStream_t)stream) != hipSuccess) return IIRX_ERR_HIP;
    b->launches++;
    return IIRX_OK;
}

int iirx_batch_capture(iirx_batch *b, const float *d_in, float *d_out, void *stream)
{
    if (!b || !d_in || !d_out || !stream) return IIRX_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    if (b->exec) {
        hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
    }
    if (b->graph) {
        hipGraphDestroy(b->graph);
        b->graph = nullptr;
    }
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) return IIRX_ERR_HIP;
    const hipError_t le = launch_all(b, d_in, d_out, s);
    hipGraph_t g = nullptr;
    const hipError_t ee = hipStreamEndCapture(s, &g);
    if (le != hipSuccess || ee != hipSuccess || !g) {
        if (g) hipGraphDestroy(g);
        return IIRX_ERR_HIP;
    }
    hipGraphExec_t x = nullptr;
    if (hipGraphInstantiate(&x, g, nullptr, nullptr, 0) != hipSuccess) {
        hipGraphDestroy(g);
        return IIRX_ERR_HIP;
    }
    b->graph = g;
    b->exec = x;
    return IIRX_OK;
}

int iirx_batch_replay(iirx_batch *b, void *stream)
{
    if (!b || !b->exec) return IIRX_ERR_INVALID;
    if (hipGraphLaunch(b->exec, (hipStream_t)stream) != hipSuccess) return IIRX_ERR_HIP;
    b->launches++;
    return IIRX_OK;
}

int iirx_batch_check(iirx_batch *b)
{
    if (!b) return IIRX_ERR_INVALID;
    if (hipDeviceSynchronize() != hipSuccess) return IIRX_ERR_HIP;
    int st = IIRX_OK;
    for (const BatchGroup &g : b->grp) {
        uint32_t *d_err = reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(g.d_ctrl) + 12);
        uint32_t err = 0;
        if (hipMemcpy(&err, d_err, sizeof err, hipMemcpyDeviceToHost) != hipSuccess) return IIRX_ERR_HIP;
        if (err) {
            const uint32_t zero = 0;
            if (hipMemcpy(d_err, &zero, sizeof zero, hipMemcpyHostToDevice) != hipSuccess) return IIRX_ERR_HIP;
            st = IIRX_ERR_TIMEOUT;
        }
    }
    return st;
}

int iirx_batch_info_get(const iirx_batch *b, iirx_batch_info *info)
{
    if (!b || !info) return IIRX_ERR_INVALID;
    *info = iirx_batch_info{};
    info->n_channels = b->n_channels;
    info->n_samples = b->n_samples;
    info->num_chunks = b->num_chunks;
    info->chunk_samples = 64LL * b->cfg.n_blocks;
    info->grid = b->grp[0].grid;
    info->launches = b->launches;
    info->graph_ready = b->exec != nullptr;
    info->n_launches = (b->grp[0].num_chunks > 0) + (b->grp[1].num_chunks > 0);
    info->n_blockstate = b->n_blockstate;
    info->bs_chunks = b->grp[1].num_chunks;
    info->bs_chunk_samples = 64LL * b->cfg.bs_tiles;
    info->bs_grid = b->grp[1].grid;
    hipFuncAttributes at;
    int nb = 0;
    for (int k = 0; k < 2; ++k) {
        const BatchGroup &g = b->grp[k];
        int32_t &vg = k ? info->bs_vgprs : info->vgprs;
        int32_t &ws = k ? info->bs_waves_per_simd : info->waves_per_simd;
        if (hipFuncGetAttributes(&at, g.fn) == hipSuccess) vg = at.numRegs;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, g.fn, 64, 0) == hipSuccess) ws = nb / 4;
    }
    (void)hipGetLastError();
    return IIRX_OK;
}

int iirx_batch_read_tables(const iirx_batch *b, float *out, int64_t cap, int64_t *size)
{
    if (!b) return IIRX_ERR_INVALID;
    if (size) *size = b->coef_size;
    if (out && cap > 0 &&
        hipMemcpy(out, b->d_coef, sizeof(float) * (size_t)std::min(cap, b->coef_size), hipMemcpyDeviceToHost) !=
            hipSuccess)
        return IIRX_ERR_HIP;
    return IIRX_OK;
}

int iirx_batch_destroy(iirx_batch *b)
{
    if (!b) return IIRX_ERR_INVALID;
    hipDeviceSynchronize();
    release(b);
    return IIRX_OK;
}

}  // extern "C"
