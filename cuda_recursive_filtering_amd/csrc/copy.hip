// copy.hip -- the HBM ceiling the filter kernels are measured against: a
// streaming copy of float32 samples with 16-byte accesses (read 4 B + write
// 4 B per sample, the filters' algorithmic traffic), no arithmetic.  Each
// workgroup copies whole blocks of 16 KB (four 16-byte loads per thread, all
// in flight before the stores), the blocks handed out round-robin over the
// grid; loads are nontemporal (the data is touched once), stores either
// nontemporal or through the L2 as the filter kernels store.  (A grid-stride
// form whose four loads per thread lay a whole grid apart reached 5.84 TB/s
// at 2^25 samples; the 16 KB blocks 6.6 TB/s on the same kind of box.)
#include "../../include/iirx.h"
#include "device_common.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>

namespace iirx {
namespace {

constexpr int kCopyThreads = 256;
constexpr int kCopyUnroll = 4;

template <bool NT_STORE>
__global__ __launch_bounds__(kCopyThreads) void stream_copy_kernel(const f4 *__restrict__ src, f4 *__restrict__ dst,
                                                                   int64_t n4)
{
    constexpr int64_t BLK = (int64_t)kCopyThreads * kCopyUnroll;
    const int t = threadIdx.x;
    for (int64_t base = (int64_t)blockIdx.x * BLK; base < n4; base += (int64_t)gridDim.x * BLK) {
        if (base + BLK <= n4) {
            f4 v[kCopyUnroll];
#pragma unroll
            for (int u = 0; u < kCopyUnroll; ++u) v[u] = __builtin_nontemporal_load(src + base + u * kCopyThreads + t);
#pragma unroll
            for (int u = 0; u < kCopyUnroll; ++u) {
                if constexpr (NT_STORE) __builtin_nontemporal_store(v[u], dst + base + u * kCopyThreads + t);
                else dst[base + u * kCopyThreads + t] = v[u];
            }
        } else {
            for (int64_t i = base + t; i < n4; i += kCopyThreads) {
                const f4 v = __builtin_nontemporal_load(src + i);
                if constexpr (NT_STORE) __builtin_nontemporal_store(v, dst + i);
                else dst[i] = v;
            }
        }
    }
}

}  // namespace
}  // namespace iirx

extern "C" int iirx_copy_ceiling(const float *d_src, float *d_dst, int64_t n, int32_t mode, void *stream)
{
    if (!d_src || !d_dst || n <= 0 || (n & 3) || (((uintptr_t)d_src | (uintptr_t)d_dst) & 15)) return IIRX_ERR_INVALID;
    if (mode != 0 && mode != 1) return IIRX_ERR_INVALID;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return IIRX_ERR_HIP;
    const int64_t n4 = n / 4;
    // 8 workgroups per CU, at most one 16 KB block each per round
    const int64_t blk = (int64_t)iirx::kCopyThreads * iirx::kCopyUnroll;
    const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((int64_t)cus * 8, (n4 + blk - 1) / blk));
    const auto *s4 = reinterpret_cast<const iirx::f4 *>(d_src);
    auto *d4 = reinterpret_cast<iirx::f4 *>(d_dst);
    if (mode == 0)
        hipLaunchKernelGGL(iirx::stream_copy_kernel<true>, dim3((unsigned)blocks), dim3(iirx::kCopyThreads), 0,
                           (hipStream_t)stream, s4, d4, n4);
    else
        hipLaunchKernelGGL(iirx::stream_copy_kernel<false>, dim3((unsigned)blocks), dim3(iirx::kCopyThreads), 0,
                           (hipStream_t)stream, s4, d4, n4);
    return hipGetLastError() == hipSuccess ? IIRX_OK : IIRX_ERR_HIP;
}
