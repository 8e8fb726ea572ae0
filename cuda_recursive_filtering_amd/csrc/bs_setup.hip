// bs_setup.hip -- the block-state form's tables (kernels.hpp: bs_layout),
// computed in FP64 on the GPU.
//
// No counterpart in the reference (its precompute is float32 host code,
// ph_kernels.cuh:151-200).  One workgroup of 256 threads builds one table:
// the cascade's state space (A, B, C) from the sections, the balancing by the
// controllability Gramian, the impulse response h, the observability rows C A^i, the controllability
// columns A^i B, Phi = A^L with Phi^(2^j), Phi^(16 nb) and Phi^G, the
// power-of-two operand scales, the FP16 hi/lo and FP32 operand blocks and the
// lookback depth.  Every FP64 operation is the one precompute.cpp: bs_tables
// performs, in the same order and without contraction, so the device tables
// equal the host's bit for bit (tests compare them); 16 x 16 products give
// each thread one element.
#include "device_common.hpp"
#include "kernels.hpp"

namespace iirx {

namespace {

constexpr int kSetupThreads = 256;

// c = a b (16 x 16, row major), summing k ascending and skipping zero a[i][k],
// as mat_mul16 does on the host
__device__ void mm16(const double *a, const double *b, double *c)
{
#pragma clang fp contract(off)
    const int t = threadIdx.x, i = t >> 4, j = t & 15;
    double s = 0.0;
    for (int k = 0; k < 16; ++k) {
        const double v = a[i * 16 + k];
        if (v == 0.0) continue;
        s += v * b[k * 16 + j];
    }
    __syncthreads();  // a or b may alias c
    c[t] = s;
    __syncthreads();
}

// correctly rounded square root: the hardware estimate, then the nearer of
// its two neighbours if one is nearer (residuals exact by fma)
__device__ double sqrt_rn(double w)
{
    const double y = __builtin_sqrt(w);
    double best = y, err = __builtin_fabs(__builtin_fma(y, y, -w));
    const double cand[2] = {nextafter(y, 0.0), nextafter(y, __builtin_inf())};
    for (double c : cand) {
        const double e = __builtin_fabs(__builtin_fma(c, c, -w));
        if (e < err) {
            best = c;
            err = e;
        }
    }
    return best;
}

__device__ uint16_t half_bits_d(_Float16 h) { return __builtin_bit_cast(uint16_t, h); }

template <int L, int SPLIT>
__global__ __launch_bounds__(kSetupThreads) void bs_setup_kernel(const BsSetupJob *jobs, float *out_base, int G)
{
#pragma clang fp contract(off)
    constexpr BsLayout Y = bs_layout(L, SPLIT);
    const BsSetupJob &job = jobs[blockIdx.x];
    float *out = out_base + job.out_off;
    const int ns = job.ns, p = 2 * ns, t = threadIdx.x;

    __shared__ double A[256], W[256], Ak[256], T1[256], T2[256], Pm[6][256], Qm[4][256], Mc[256], Pw[256];
    __shared__ double Bv[16], Cv[16], dsc[16], h[L], powB[L][16], Om[L][16];
    __shared__ double red[kSetupThreads];

    for (int i = t; i < Y.stride; i += kSetupThreads) out[i] = 0.0f;

    // cascade of DF-II-transposed sections (precompute.cpp: bs_tables)
    A[t] = 0.0;
    if (t < 16) {
        Bv[t] = 0.0;
        Cv[t] = 0.0;
    }
    __syncthreads();
    if (t == 0) {
        for (int s = 0; s < ns; ++s) {
            const float *r = job.sos + 6 * s;
            const double b1 = r[1], b2 = r[2], a1 = r[4], a2 = r[5];
            const double B0 = b1 - a1, B1 = b2 - a2;
            A[(2 * s) * 16 + 2 * s] = -a1;
            A[(2 * s) * 16 + 2 * s + 1] = 1.0;
            A[(2 * s + 1) * 16 + 2 * s] = -a2;
            for (int q = 0; q < s; ++q) {
                A[(2 * s) * 16 + 2 * q] += B0;
                A[(2 * s + 1) * 16 + 2 * q] += B1;
            }
            Bv[2 * s] = B0;
            Bv[2 * s + 1] = B1;
            Cv[2 * s] = 1.0;
        }
    }
    __syncthreads();

    // balancing by the controllability Gramian W = sum_k A^k B B^T A^kT (doubling)
    {
        const int r = t >> 4, c = t & 15;
        W[t] = Bv[r] * Bv[c];
        Ak[t] = A[t];
        __syncthreads();
        for (int it = 0; it < 48; ++it) {
            mm16(Ak, W, T1);              // AW
            T2[t] = Ak[c * 16 + r];       // Ak^T
            __syncthreads();
            mm16(T1, T2, T1);             // add = AW Ak^T
            const double add = T1[t];
            W[t] += add;
            __syncthreads();
            // delta = max|add|, size = max|W|
            red[t] = __builtin_fabs(add);
            __syncthreads();
            for (int sft = kSetupThreads / 2; sft > 0; sft >>= 1) {
                if (t < sft) red[t] = fmax(red[t], red[t + sft]);
                __syncthreads();
            }
            const double delta = red[0];
            __syncthreads();
            red[t] = __builtin_fabs(W[t]);
            __syncthreads();
            for (int sft = kSetupThreads / 2; sft > 0; sft >>= 1) {
                if (t < sft) red[t] = fmax(red[t], red[t + sft]);
                __syncthreads();
            }
            const double size = red[0];
            __syncthreads();
            mm16(Ak, Ak, Ak);
            if (!(delta > 1e-18 * size)) break;
        }
        if (t < 16) {
            const double w = W[t * 16 + t];
            dsc[t] = (t < p && w > 0.0 && __builtin_isfinite(w)) ? sqrt_rn(w) : 1.0;
        }
        __syncthreads();
        A[t] *= dsc[c] / dsc[r];
        if (t < 16) {
            Bv[t] /= dsc[t];
            Cv[t] *= dsc[t];
        }
        __syncthreads();
    }

    // states stay in order 0..p-1 (K slab m of the state products holds states 4 m .. 4 m + 3)

    // impulse response, observability rows C A^i, controllability columns A^i B
    {
        double *cA = T1, *AB = T1 + 16, *n1 = T1 + 32, *n2 = T1 + 48;
        if (t < 16) {
            cA[t] = Cv[t];
            AB[t] = Bv[t];
        }
        __syncthreads();
        for (int i = 0; i < L; ++i) {
            if (t < 16) {
                Om[i][t] = cA[t];
                powB[i][t] = AB[t];
            }
            if (t == 0) {
                double hv = 0.0;
                for (int c = 0; c < 16; ++c) hv += Cv[c] * AB[c];
                if (i + 1 < L) h[i + 1] = hv;
            }
            if (t < 16) {
                double s1 = 0.0, s2 = 0.0;
                for (int r = 0; r < 16; ++r) s1 += cA[r] * A[r * 16 + t];
                for (int c = 0; c < 16; ++c) s2 += A[t * 16 + c] * AB[c];
                n1[t] = s1;
                n2[t] = s2;
            }
            __syncthreads();
            if (t < 16) {
                cA[t] = n1[t];
                AB[t] = n2[t];
            }
            __syncthreads();
        }
        if (t == 0) h[0] = 1.0;
        __syncthreads();
    }

    // Phi = A^L and its powers
    T2[t] = ((t >> 4) == (t & 15)) ? 1.0 : 0.0;
    __syncthreads();
    for (int i = 0; i < L; ++i) mm16(T2, A, T2);
    Pm[0][t] = T2[t];
    __syncthreads();
    for (int j = 1; j < 6; ++j) mm16(Pm[j - 1], Pm[j - 1], Pm[j]);
    mm16(Pm[3], Pm[3], Qm[1]);  // Phi^16
    mm16(Qm[1], Qm[1], Qm[2]);
    mm16(Qm[2], Qm[1], Qm[3]);
    if (G == 16) mm16(Pm[3], Pm[3], Mc);
    else if (G == 32) mm16(Pm[4], Pm[4], Mc);
    else mm16(Pm[5], Pm[5], Mc);

    // operand scales: 2^s with max * 2^s < 2^14 (max|T| = max|h|, Gam = A^i B)
    int sT, sG;
    {
        double m = 0.0;
        for (int i = t; i < L; i += kSetupThreads) m = fmax(m, __builtin_fabs(h[i]));
        red[t] = m;
        __syncthreads();
        for (int sft = kSetupThreads / 2; sft > 0; sft >>= 1) {
            if (t < sft) red[t] = fmax(red[t], red[t + sft]);
            __syncthreads();
        }
        const double mT = red[0];
        __syncthreads();
        m = 0.0;
        for (int i = t; i < 16 * L; i += kSetupThreads) m = fmax(m, __builtin_fabs(powB[i >> 4][i & 15]));
        red[t] = m;
        __syncthreads();
        for (int sft = kSetupThreads / 2; sft > 0; sft >>= 1) {
            if (t < sft) red[t] = fmax(red[t], red[t + sft]);
            __syncthreads();
        }
        const double mG = red[0];
        __syncthreads();
        int e;
        sT = (mT > 0.0 && __builtin_isfinite(mT)) ? (frexp(mT, &e), 14 - e) : 0;
        sG = (mG > 0.0 && __builtin_isfinite(mG)) ? (frexp(mG, &e), 14 - e) : 0;
    }
    if (t == 0) {
        out[Y.kScale + 0] = (float)sT;
        out[Y.kScale + 1] = (float)sG;
    }

    // FP16 hi/lo x32 operand blocks of T (rows 16 mb..) and Gam, lane order of
    // put_block32: lane l, half j -> row r0 + l % 16, column 32 q + (j < 4 ?
    // 4 g + j : 16 + 4 g + j - 4), g = l / 16
    {
        constexpr int nT = SPLIT * Y.MB * Y.KQ, nBlk = nT + SPLIT * Y.KQ;
        for (int idx = t; idx < nBlk * 512; idx += kSetupThreads) {
            const int blk = idx >> 9, e = idx & 511, l = e >> 3, j = e & 7, g = l >> 4;
            int hl, r, q, s;
            uint16_t *dst;
            double v;
            if (blk < nT) {
                hl = blk / (Y.MB * Y.KQ);
                const int mb = (blk / Y.KQ) % Y.MB;
                q = blk % Y.KQ;
                r = 16 * mb + (l & 15);
                const int c = 32 * q + (j < 4 ? 4 * g + j : 16 + 4 * g + j - 4);
                s = sT;
                v = (c <= r) ? h[r - c] : 0.0;  // T[r][c], r < L always
                dst = reinterpret_cast<uint16_t *>(out + bs_T(Y, hl, mb, q));
            } else {
                hl = (blk - nT) / Y.KQ;
                q = (blk - nT) % Y.KQ;
                r = bs_state_of_row(l & 15);  // row 4 g + j of the product = state g + 4 j
                const int c = 32 * q + (j < 4 ? 4 * g + j : 16 + 4 * g + j - 4);
                s = sG;
                v = powB[L - 1 - c][r];  // Gam[r][c] = (A^(L-1-c) B)[r]
                dst = reinterpret_cast<uint16_t *>(out + bs_G(Y, hl, q));
            }
            // hi = v rounded once to FP16, lo = the exact remainder rounded once
            v = ldexp(v, s);
            if constexpr (SPLIT == kBsSplitBF16) {
                dst[l * 8 + j] = bf16x3_bits(v, hl);
            } else {
                const _Float16 hi = (_Float16)v;
                const _Float16 lo = (_Float16)(v - (double)hi);
                dst[l * 8 + j] = half_bits_d(hl == 0 ? hi : lo);
            }
        }
    }
    // FP32 lane-order blocks: word 4 l + k = M[r0 + l % 16][l / 16 + 4 k]
    {
        const int l = t >> 2, k = t & 3, rr = l & 15, cc = (l >> 4) + 4 * k;
        if (t < 256) {
            for (int j = 0; j < 6; ++j) out[bs_P(Y, j) + t] = (float)Pm[j][rr * 16 + cc];
            for (int nb = 1; nb < 4; ++nb) out[bs_Q(Y, nb) + t] = (float)Qm[nb][rr * 16 + cc];
            out[Y.kPhi + t] = (float)Pm[0][rr * 16 + cc];
            for (int mb = 0; mb < Y.MB; ++mb) out[bs_O(Y, mb) + t] = (float)Om[16 * mb + rr][cc];
            out[Y.kMc + t] = (float)Mc[t];
        }
    }
    // lookback depth of the chunk transition: smallest d with ||Mc^d||_inf <= 2^-36
    Pw[t] = Mc[t];
    __syncthreads();
    int depth = 0;
    for (int d = 1; d <= 64; ++d) {
        if (t < 16) {
            double s = 0.0;
            for (int c = 0; c < 16; ++c) s += __builtin_fabs(Pw[t * 16 + c]);
            red[t] = s;
        }
        __syncthreads();
        double nrm = 0.0;
        for (int r = 0; r < 16; ++r) nrm = fmax(nrm, red[r]);
        __syncthreads();
        if (!(nrm > kLookbackTail)) {
            depth = __builtin_isfinite(nrm) ? d : 0;
            break;
        }
        mm16(Pw, Mc, Pw);
    }
    if (t == 0) out[Y.kDepth] = (float)depth;

    // levels of the in-chunk tile scan (precompute.cpp: bs_tables)
    auto norm_inf = [&](const double *m) {
        if (t < 16) {
            double s = 0.0;
            for (int c = 0; c < 16; ++c) s += __builtin_fabs(m[t * 16 + c]);
            red[t] = s;
        }
        __syncthreads();
        double nrm = 0.0;
        for (int r = 0; r < 16; ++r) nrm = fmax(nrm, red[r]);
        __syncthreads();
        return nrm;
    };
    __shared__ double nrmv[65];
    const int lvG = bs_scan_levels(G);
    Pw[t] = Pm[0][t];
    __syncthreads();
    bool finite = true;
    for (int i = 1; i <= G; ++i) {
        const double v = norm_inf(Pw);
        finite = finite && __builtin_isfinite(v);
        if (t == 0) nrmv[i] = v;
        mm16(Pw, Pm[0], Pw);
    }
    int nlv = lvG, lv32 = lvG;
    if (finite) {
        for (int j = 0; j < lvG; ++j) {
            double tail = 0.0;
            for (int i = 1 << j; i <= G; ++i) tail += nrmv[i];
            if (tail <= kLookbackTail) {
                nlv = j;
                break;
            }
        }
        for (int j = lvG - 1; j >= 0 && norm_inf(Pm[j]) <= kBsFp32Norm; --j) lv32 = j;
    }
    if (t == 0) {
        out[Y.kScale + kBsScaleFp32Level] = (float)lv32;
        out[Y.kScale + kBsScaleLevels] = (float)nlv;
    }
}

}  // namespace

template <int SPLIT>
static hipError_t launch_setup(int L, int G, const BsSetupJob *jobs, int n_jobs, float *out, hipStream_t s)
{
    switch (L) {
    case 32: hipLaunchKernelGGL((bs_setup_kernel<32, SPLIT>), dim3(n_jobs), dim3(kSetupThreads), 0, s, jobs, out, G); break;
    case 64: hipLaunchKernelGGL((bs_setup_kernel<64, SPLIT>), dim3(n_jobs), dim3(kSetupThreads), 0, s, jobs, out, G); break;
    case 128: hipLaunchKernelGGL((bs_setup_kernel<128, SPLIT>), dim3(n_jobs), dim3(kSetupThreads), 0, s, jobs, out, G); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_bs_setup(int L, int G, const BsSetupJob *jobs, int n_jobs, float *out, hipStream_t s, int split)
{
    if (n_jobs <= 0) return hipSuccess;
    if (split == kBsSplitBF16) return launch_setup<kBsSplitBF16>(L, G, jobs, n_jobs, out, s);
    if (split == kBsSplitF16) return launch_setup<kBsSplitF16>(L, G, jobs, n_jobs, out, s);
    return hipErrorInvalidValue;
}

}  // namespace iirx
