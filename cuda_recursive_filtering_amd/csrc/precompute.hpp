// precompute.hpp -- FP64 host precompute of the per-section coefficient
// tables (rounded once to float32).  The reference computes the same tables
// in float32 on the host (ph_kernels.cuh:55-200, stcr_kernels.cuh:78-272,
// dtcr_kernels.cuh:92-289, plr_kernels.cuh:149-476); computing them in FP64
// removes the accumulated rounding of the matrix powers and recurrences.
#pragma once

#include <vector>

namespace iirx {

// 2x2 row-major double matrices
struct M2 {
    double a, b, c, d;
};

inline M2 m2_mul(const M2 &p, const M2 &q)
{
    return {p.a * q.a + p.b * q.c, p.a * q.b + p.b * q.d, p.c * q.a + p.d * q.c, p.c * q.b + p.d * q.d};
}

inline M2 m2_eye() { return {1.0, 0.0, 0.0, 1.0}; }

// N-step zero-input response of the all-pole part 1/(1 - a1 z^-1 - a2 z^-2):
// h1[n] = h0[n+1], h2[n] = a2 h0[n] (ph_kernels.cuh:55-68) and the N-step
// transition C = [[h2[N-2], h1[N-2]], [h2[N-1], h1[N-1]]] (ph:180-181).
void zir_tables(double a1, double a2, int N, std::vector<double> &h1, std::vector<double> &h2, M2 &C);

// Smallest d in [1, 64] with ||M^d||_inf <= 2^-36, 0 if there is none.
int lookback_depth(const M2 &M);

// PH tables for all sections (layout kernels.hpp:ph_layout); returns the
// per-section stride in floats.
int ph_tables(const float *sos, int ns, int lanes, int N, std::vector<float> &out);

// PH tables for all sections followed by the tables of the section pairs
// (0, 1), (2, 3), ... (layout kernels.hpp:ph_pair_layout); returns the
// per-section stride in floats.
int ph_joint_tables(const float *sos, int ns, int lanes, int N, std::vector<float> &out);

// PH tables for all sections followed by the halo block (layout
// kernels.hpp, kPhHalo): the FP64 impulse responses of the cascade prefixes
// as the halo weights K2_s, K1_s, and the flag that the cascade forgets its
// state within kPhHalo samples (||A^kPhHalo||_inf <= 2^-34, A the zero-input
// one-sample transition of the cascade state; *forgets receives it).
// Returns the per-section stride in floats.
int ph_halo_tables(const float *sos, int ns, int lanes, int N, std::vector<float> &out, bool *forgets = nullptr);

// ||A^n||_inf of the cascade's zero-input state transition over n samples,
// state (x[-2], x[-1], y_1[-2], y_1[-1], ..., y_ns[-2], y_ns[-1]), in FP64.
double cascade_transition_norm(const float *sos, int ns, int n);

// The pair's joint N-step transition J (4 x 4 row-major) and the responses
// H[m * N + n] of the second section's output to the unit joint states, in
// FP64, for sections (s, t) of `sos`.
void ph_pair_response(const float *sos, int s, int N, double J[16], std::vector<double> &H);

// CR tables for all sections (layout kernels.hpp:cr_layout).  Returns the
// stride, or -1 when the CR factors are ill-conditioned for this (filter, N):
// some |e_n| is so small that the elimination multipliers f/e, d, c exceed
// kCrLimit (e_n = -2 r^(2^n) cos(2^n theta) for a pole pair r e^(+-i theta)).
constexpr double kCrLimit = 1.0e4;
int cr_tables(const float *sos, int ns, int lanes, int N, std::vector<float> &out, double *worst = nullptr);

// PLR tables (layout kernels.hpp:plr_layout) from the direct-form taps
// b[0..K), a[0..K) (recurrence sign); the kernel applies float32 taps, so the
// FP64 factors are those of the float32-rounded feedback taps.  Returns the
// table size in floats.
int plr_tables(const double *b, const double *a, int K, int CH, std::vector<float> &out);

// Block-state tables (layout kernels.hpp:bs_layout) for tiles of L samples.
int bs_tables(const float *sos, int ns, int L, int G, std::vector<float> &out, int split = 2);  // 2: FP16 hi/lo, 3: BF16x3

// The block-state realization in FP64: A (16 x 16), B, C of the cascade of
// direct-form-II-transposed sections (state (w1, w2) of section s at 2s,
// 2s+1), every state divided by the balancing scale d[c] (d = 1 past the
// 2 ns states).
void bs_realization(const float *sos, int ns, std::vector<double> &A, std::vector<double> &B, std::vector<double> &C,
                    double *d);

}  // namespace iirx
