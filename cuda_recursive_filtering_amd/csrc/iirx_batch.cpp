// iirx_batch.cpp -- the ragged-batch part of the C ABI (include/iirx.h):
// channel validation, per-channel FP64 tables, the chunk -> channel map, the
// once-zeroed lookback state, the persistent launch and its hipGraph capture.
//
// No counterpart in the reference, which filters one channel per binary; the
// per-channel PH tables are those of setup_kernel_coefficients
// (ph_kernels.cuh:151-200), built by ph_tables; block-state channels take
// bs_tables for tiles of 64 samples and n_blocks tiles per chunk.
#include "../../include/iirx.h"
#include "kernels.hpp"
#include "precompute.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

using namespace iirx;

struct iirx_batch {
    iirx_batch_config cfg;
    int64_t n_channels, n_samples;
    uint32_t num_chunks, grid, spin_limit;
    int N, coef_stride, device;
    bool mixed;
    int64_t n_blockstate;
    float *d_coef;
    uint64_t *d_lb;
    BatchChannel *d_chan;
    uint32_t *d_chunk_chan;
    unsigned long long *d_ctrl;  // [ticket u64][epoch u32][err u32], arrival slots from byte 128
    hipGraph_t graph;
    hipGraphExec_t exec;
    int64_t launches;
    const void *fn;
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
    if (b->d_ctrl) hipFree(b->d_ctrl);
    delete b;
}

BatchArgs args(const iirx_batch *b, const float *d_in, float *d_out)
{
    BatchArgs a{};
    a.in = d_in;
    a.out = d_out;
    a.coef = b->d_coef;
    a.lb = b->d_lb;
    a.chan = b->d_chan;
    a.chunk_chan = b->d_chunk_chan;
    a.ticket = b->d_ctrl;
    a.epoch = reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(b->d_ctrl) + 8);
    a.err = a.epoch + 1;
    a.arrive = reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(b->d_ctrl) + 128);
    a.num_chunks = b->num_chunks;
    a.spin_limit = b->spin_limit;
    a.persistent = b->cfg.schedule == IIRX_SCHEDULE_PERSISTENT;
    a.grid = b->grid;
    a.coef_stride = b->coef_stride;
    return a;
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
    if (!(N == 16 || N == 32 || N == 64)) return IIRX_ERR_GEOMETRY;
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
    if (offsets[n_channels] == offsets[0]) return IIRX_ERR_INVALID;  // no samples at all
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return IIRX_ERR_NO_DEVICE;

    // per-channel tables, chunk list channel after channel
    const int64_t CH = 64LL * N;
    std::vector<float> coef;
    std::vector<BatchChannel> chan((size_t)n_channels);
    std::vector<uint32_t> chunk_chan;
    int stride = 0;
    int64_t sec0 = 0, n_bs = 0;
    for (int64_t c = 0; c < n_channels; ++c) {
        std::vector<float> t;
        const int ns = n_sections[c];
        const int32_t want = algo ? algo[c] : IIRX_BATCH_AUTO;
        const bool bs = want == IIRX_BATCH_BLOCKSTATE ||
                        (want == IIRX_BATCH_AUTO && cfg->bs_min_sections > 0 && ns >= cfg->bs_min_sections);
        if (bs) {
            if (bs_tables(sos + 6 * sec0, ns, 64, N, t) <= 0) return IIRX_ERR_INVALID;
            ++n_bs;
        } else {
            const int st = ph_tables(sos + 6 * sec0, ns, 64, N, t);
            if (st <= 0) return IIRX_ERR_INVALID;
            stride = st;
        }
        sec0 += ns;
        t.resize((t.size() + 3) & ~(size_t)3, 0.0f);  // 16-byte aligned tables
        const int64_t len = offsets[c + 1] - offsets[c];
        const int64_t k = (len + CH - 1) / CH;
        if ((int64_t)chunk_chan.size() + k > 0x7fffffffLL || coef.size() + t.size() > 0x7fffffffULL)
            return IIRX_ERR_GEOMETRY;
        BatchChannel &d = chan[(size_t)c];
        d.offset = offsets[c];
        d.length = len;
        d.ns = ns;
        d.coef_off = (int32_t)coef.size();
        d.first_chunk = (uint32_t)chunk_chan.size();
        d.algo = bs ? kBatchBS : kBatchPH;
        coef.insert(coef.end(), t.begin(), t.end());
        chunk_chan.insert(chunk_chan.end(), (size_t)k, (uint32_t)c);
    }
    if (chunk_chan.empty()) return IIRX_ERR_INVALID;

    iirx_batch *b = new (std::nothrow) iirx_batch();
    if (!b) return IIRX_ERR_OOM;
    b->cfg = *cfg;
    b->cfg.n_blocks = N;
    b->n_channels = n_channels;
    b->n_samples = offsets[n_channels] - offsets[0];
    b->num_chunks = (uint32_t)chunk_chan.size();
    b->N = N;
    b->coef_stride = stride;
    b->mixed = n_bs > 0;
    b->n_blockstate = n_bs;
    b->spin_limit = 1u << 22;
    hipGetDevice(&b->device);
    b->fn = kernel_batch(N, b->mixed);
    const size_t lb_count = (size_t)b->num_chunks * kBatchMaxSections * 4u;
    if (hipMalloc((void **)&b->d_coef, coef.size() * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&b->d_lb, lb_count * sizeof(uint64_t)) != hipSuccess ||
        hipMalloc((void **)&b->d_chan, chan.size() * sizeof(BatchChannel)) != hipSuccess ||
        hipMalloc((void **)&b->d_chunk_chan, chunk_chan.size() * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc((void **)&b->d_ctrl, kControlBytes) != hipSuccess) {
        release(b);
        return IIRX_ERR_OOM;
    }
    if (hipMemcpy(b->d_coef, coef.data(), coef.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->d_chan, chan.data(), chan.size() * sizeof(BatchChannel), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->d_chunk_chan, chunk_chan.data(), chunk_chan.size() * sizeof(uint32_t), hipMemcpyHostToDevice) !=
            hipSuccess ||
        // zeroed once; launch-versioned flags never need a reset afterwards
        hipMemset(b->d_lb, 0, lb_count * sizeof(uint64_t)) != hipSuccess ||
        hipMemset(b->d_ctrl, 0, kControlBytes) != hipSuccess) {
        release(b);
        return IIRX_ERR_HIP;
    }
    b->grid = b->num_chunks;
    if (cfg->schedule == IIRX_SCHEDULE_PERSISTENT) {
        int per_cu = 0, cus = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, b->fn, 64, 0) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device) != hipSuccess) {
            release(b);
            return IIRX_ERR_HIP;
        }
        b->grid = (uint32_t)std::min<int64_t>((int64_t)std::max(per_cu, 1) * std::max(cus, 1), b->num_chunks);
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
    if (launch_batch(b->N, b->mixed, args(b, d_in, d_out), (hipStream_t)stream) != hipSuccess) return IIRX_ERR_HIP;
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
    const hipError_t le = launch_batch(b->N, b->mixed, args(b, d_in, d_out), s);
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
    uint32_t *d_err = reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(b->d_ctrl) + 12);
    uint32_t err = 0;
    if (hipMemcpy(&err, d_err, sizeof err, hipMemcpyDeviceToHost) != hipSuccess) return IIRX_ERR_HIP;
    if (err) {
        const uint32_t zero = 0;
        if (hipMemcpy(d_err, &zero, sizeof zero, hipMemcpyHostToDevice) != hipSuccess) return IIRX_ERR_HIP;
        return IIRX_ERR_TIMEOUT;
    }
    return IIRX_OK;
}

int iirx_batch_info_get(const iirx_batch *b, iirx_batch_info *info)
{
    if (!b || !info) return IIRX_ERR_INVALID;
    *info = iirx_batch_info{};
    info->n_channels = b->n_channels;
    info->n_samples = b->n_samples;
    info->num_chunks = b->num_chunks;
    info->chunk_samples = 64LL * b->N;
    info->grid = b->grid;
    info->launches = b->launches;
    info->graph_ready = b->exec != nullptr;
    info->mixed = b->mixed;
    info->n_blockstate = b->n_blockstate;
    hipFuncAttributes at;
    if (hipFuncGetAttributes(&at, b->fn) == hipSuccess) info->vgprs = at.numRegs;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, b->fn, 64, 0) == hipSuccess) info->waves_per_simd = nb / 4;
    (void)hipGetLastError();
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
