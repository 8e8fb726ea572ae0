// bs_bf16.hip -- the block-state kernel with T, Gam and X split three ways
// into BF16 (hi, mid, lo: 24 significand bits, FP32's exponent range) and
// six v_mfma_f32_16x16x32_bf16 products per K slab (hi hi, hi mid, mid hi,
// mid mid, hi lo, lo hi), FP32 accumulation; the per-chunk power-of-two
// scaling of X and the table scaling are those of the FP16 hi/lo form.
// Instances for tiles of L = 64 and 128 samples and 16 or 32 tiles per chunk
// (64 tiles of 64 would spill).
#include "bs_kernel.hpp"

namespace iirx {

hipError_t launch_bs_bf16(int L, int G, int nsl, const LaunchArgs &a, hipStream_t s)
{
    switch (L) {
    case 64: return launch_l<64, kBsSplitBF16>(G, nsl, a, s);
    case 128: return launch_l<128, kBsSplitBF16>(G, nsl, a, s);
    default: return hipErrorInvalidValue;
    }
}

const void *kernel_bs_bf16(int L, int G, int nsl)
{
    switch (L) {
    case 64: return fn_l<64, kBsSplitBF16>(G, nsl);
    case 128: return fn_l<128, kBsSplitBF16>(G, nsl);
    default: return nullptr;
    }
}

}  // namespace iirx
