/*
 * iirx.h -- C ABI of the MI355X-native parallel cascaded-biquad IIR library
 * (libiirx.so, built from cuda_recursive_filtering_amd/csrc).
 *
 * The reference (Haotian-RA/cuda_recursive_filtering) has no library
 * boundary: a harness includes one kernel header chosen at compile time,
 * calls setup_kernel_coefficients(sos) and launches
 *     KERNEL_FUNC<<<grid, tb>>>(in, out, status, partcarry, fullcarry, launch)
 * The entry points below replace that compile-time contract with a runtime
 * plan.  Each declaration cites the reference interface it replaces.
 *
 * Filter contract (ref_generate.py:61-67, ph_kernels.cuh:157-161): sections
 * are rows {b0, b1, b2, a0, a1, a2} in scipy's sign convention, normalised to
 * b0 == 1 and a0 == 1; the removed cascade gain is applied to the input by
 * the caller.  Each section computes
 *     y[n] = u[n] + b1 u[n-1] + b2 u[n-2] - a1 y[n-1] - a2 y[n-2]
 * with zero initial state (iirx_execute_state carries a state in and out),
 * float32 in and float32 out.  A non-finite input sample at k makes the
 * output non-finite from k on (the block-state form: from the start of k's
 * 32-sample slab) until the bounded lookback drops the carry that holds it,
 * a few chunks later; after that the output is that of the signal with
 * x[k] = 0 (sequential filtering would stay non-finite).
 *
 * All device pointers are plain device addresses (hipMalloc or any
 * allocator of the same device); `stream` is a hipStream_t passed as void*
 * (NULL = the null stream).  Input and output must not overlap: a chunk
 * reads input samples just before its own start.
 *
 * Every function returns an iirx_status; nothing aborts.
 */
#ifndef IIRX_H
#define IIRX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (the reference asserts/aborts instead: ph_kernels.cuh:186-199,
 * test_accuracy.cu:136-159). */
typedef enum {
    IIRX_OK = 0,
    IIRX_ERR_INVALID = 1,        /* bad argument or unsupported configuration */
    IIRX_ERR_GEOMETRY = 2,       /* no kernel instance for this geometry,
                                    or more than 2^31 - 1 chunks */
    IIRX_ERR_CONDITIONING = 3,   /* CR factors ill-conditioned for this
                                    (filter, N_BLOCKS): |e_n| too small */
    IIRX_ERR_HIP = 4,            /* a HIP runtime call failed */
    IIRX_ERR_NO_DEVICE = 5,      /* no GPU visible */
    IIRX_ERR_OOM = 6,
    IIRX_ERR_TIMEOUT = 7,        /* a lookback wait exceeded the plan's spin
                                    limit (workgroups not co-resident); the
                                    output of that call is invalid */
    IIRX_ERR_TOO_LARGE = 8,      /* iirx_execute_state_noalloc: the call needs
                                    more than IIRX_NOALLOC_SCRATCH_BYTES of
                                    device scratch */
    IIRX_ERR_ACCURACY = 9        /* autotuner: the candidate's output missed
                                    the accuracy screen (max|err| above 1e-4
                                    max|ref| and above 10x the best
                                    candidate's), so it is not eligible */
} iirx_status;

/* Workgroup scheduling of a plan (no counterpart in the reference, which
 * always launches one block per chunk with an atomic ticket). */
typedef enum {
    IIRX_SCHEDULE_PERSISTENT = 0, /* one workgroup per resident slot, chunks
                                     assigned round-robin; no per-chunk atomic */
    IIRX_SCHEDULE_TICKET = 1      /* chunk order from an atomic ticket
                                     (ph_kernels.cuh:227-229): one per chunk,
                                     or one per group of consecutive chunks
                                     (PH, STCR) */
} iirx_schedule;

/* Algorithm selection (the KERNEL_* build flag, test_accuracy.cu:55-66). */
typedef enum {
    IIRX_KERNEL_PH = 0,    /* partial-LU factorisation, ph_kernels.cuh */
    IIRX_KERNEL_STCR = 1,  /* single-stage tridiagonal CR, stcr_kernels.cuh */
    IIRX_KERNEL_DTCR = 2,  /* decoupled two-sided CR, dtcr_kernels.cuh */
    IIRX_KERNEL_PLR = 3,   /* direct-form hierarchical merge, plr_kernels.cuh */
    IIRX_KERNEL_BLOCKSTATE = 4, /* no counterpart in the reference: the whole
                                   cascade as a state-space block filter on the
                                   matrix cores (FP16 hi/lo operands for the
                                   Toeplitz and input products, FP32 operands
                                   with FP64 accumulation for the tile-state
                                   scan); n_blocks = tile length L
                                   (32, 64 or 128), block_size = tiles per
                                   chunk G (16, 32 or 64; 16 or 32 at L =
                                   128), chunk = G * L samples */
    IIRX_KERNEL_BLOCKSTATE_BF16 = 5, /* the block-state form with the Toeplitz
                                        and input products split three ways
                                        into BF16 (hi, mid, lo; six MFMA
                                        products per K slab) instead of FP16
                                        hi/lo; L = 64 or 128, G = 16 or 32 */
    IIRX_KERNEL_PH_JOINT = 6,  /* no counterpart in the reference: PH with
                                  sections taken in pairs, one decoupled
                                  lookback of the pair's joint 4-float state
                                  per chunk instead of one per section (an
                                  odd last section alone, and a pair whose
                                  joint tables exceed 256, poles close
                                  together near the unit circle, one section
                                  at a time); same geometry, records,
                                  streaming state and any-length rules as
                                  PH */
    IIRX_KERNEL_PH_HALO = 7    /* no counterpart in the reference: PH whose
                                  chunks take every section's incoming state
                                  from the 256 input samples before the chunk
                                  (weights precomputed in FP64) instead of a
                                  decoupled lookback, when the cascade forgets
                                  its state within 256 samples
                                  (||A^256||_inf <= 2^-34 for its one-sample
                                  state transition A); otherwise (poles close
                                  to the unit circle) exactly PH.
                                  iirx_plan_info.halo says which.  Same
                                  geometry, records, streaming state and
                                  any-length rules as PH.  With halo carries
                                  no chunk waits on another, so both
                                  schedules run the same grid-stride launch
                                  (no ticket, no co-residency assumed) */
} iirx_kernel;

/*
 * Plan configuration: the compile-time -D flags of run_accuracy.sh:155-166
 * (ph_kernels.cuh:6-14, stcr_kernels.cuh:7-22, dtcr_kernels.cuh:11-26,
 * plr_kernels.cuh:14-21) as runtime fields.
 *
 *   block_size  BLOCK_SIZE for PH/STCR/DTCR: 32 or 64.  A chunk (the
 *               lookback unit) is BLOCK_SIZE x N_BLOCKS samples, as in the
 *               reference (test_accuracy.cu:37-38), held by one wave64:
 *               64 lanes x N_BLOCKS samples at 64, 64 lanes x N_BLOCKS / 2
 *               at 32 (so 32 x 128 is the same chunk and kernel as 64 x 64).
 *               PLR: threads per block, 0 = per-order default.
 *   n_blocks    N_BLOCKS: a power of two, 8..128 (16..128 at BLOCK_SIZE
 *               32).  Ignored by PLR.
 *   n_sections  N_SECTIONS = order / 2 (1..8).
 *   unrolled    STCR_HANDUNROLLED / DTCR_HANDUNROLLED: select the
 *               template-expanded literal-index back substitution
 *               (stcr_kernels.cuh:3613-3633).
 *   plr_x       PLR values per thread; 0 = adaptive rule of plr:96-101
 *               capped by plr_x_max.
 *   plr_x_max   PLR_X_MAX; 0 = the build's default ceiling.
 *   schedule    iirx_schedule.  PERSISTENT launches as many workgroups as
 *               the device holds at once (occupancy API) and relies on them
 *               being co-resident; a bounded-spin watchdog turns a violated
 *               assumption into IIRX_ERR_TIMEOUT instead of a hang.  TICKET
 *               needs no co-residency: workgroups draw device-scope atomic
 *               tickets in order (about 12 ns each, serialised, on MI355X),
 *               one per group of consecutive chunks for PH and STCR (16 at
 *               N_BLOCKS <= 32) and for the block-state form (a group is
 *               one chunk per wave of its workgroup), one per chunk for
 *               DTCR and PLR.  PH_HALO with halo carries takes neither
 *               (no chunk waits on another).
 */
typedef struct {
    int32_t kernel;
    int32_t block_size;
    int32_t n_blocks;
    int32_t n_sections;
    int32_t unrolled;
    int32_t plr_x;
    int32_t plr_x_max;
    int32_t schedule;
} iirx_config;

typedef struct iirx_plan iirx_plan;

/* Launch facts of a plan, the printout of test_accuracy.cu:86-101 and the
 * occupancy model of gpu_specs.hpp:57-116. */
typedef struct {
    int64_t n_samples;
    int64_t chunk_samples;      /* samples per workgroup (lookback unit) */
    int64_t num_chunks;         /* grid size */
    int32_t threads_per_block;
    int32_t lds_bytes;
    int32_t vgprs;              /* from the code object metadata, 0 if unknown */
    int32_t sgprs;
    int32_t scratch_bytes;      /* private segment (spills) per lane */
    int32_t waves_per_simd;     /* occupancy from VGPR/LDS/wave limits */
    int64_t launches;           /* launches executed so far on this plan */
    int64_t grid;               /* workgroups per launch */
    int32_t schedule;
    int32_t halo;               /* IIRX_KERNEL_PH_HALO: 1 if the chunk
                                   carries come from the halo, 0 if the
                                   plan runs PH's lookback (0 for every
                                   other kernel) */
} iirx_plan_info;

/*
 * Create a plan: host FP64 precompute of the coefficient tables, upload,
 * allocation of the lookback state (status words, carries, ticket counter),
 * zeroed once here and never again (test_accuracy.cu:145,
 * test_performance.cu:170).
 * Replaces setup_kernel_coefficients(sos) (ph_kernels.cuh:151,
 * stcr_kernels.cuh:191, dtcr_kernels.cuh:193, plr_kernels.cuh:416) together
 * with the driver's cudaMalloc/cudaMemset of status/partcarry/fullcarry
 * (test_accuracy.cu:133-145).  `sos` is n_sections rows of 6 floats.  PLR
 * derives its direct-form taps from `sos` in FP64 unless
 * iirx_plan_set_plr_taps() supplies them (the reference reads them from the
 * generated header, plr_kernels.cuh:416-427).
 * The plan is bound to the current HIP device.
 */
int iirx_plan_create(const iirx_config *cfg, const float *sos,
                     int64_t n_samples, iirx_plan **out);

/* Direct-form taps for PLR (PLR_FILTER_B / PLR_FILTER_A of filter_taps.hpp,
 * ref_generate.py:80-81, consumed at plr_kernels.cuh:423-427): b = FIR taps
 * past the implicit b0 == 1, a = feedback taps in the recurrence sign
 * y[n] += a[j] y[n-1-j]; k = order.  Recomputes the PLR tables. */
int iirx_plan_set_plr_taps(iirx_plan *plan, const double *b, const double *a,
                           int k);

/*
 * One filter call = exactly one grid-wide kernel launch on `stream`.
 * Replaces KERNEL_FUNC<<<KERNEL_GRID_DIM, KERNEL_TB_DIM>>>(in, out, status,
 * partcarry, fullcarry, launch) (test_accuracy.cu:146,
 * test_performance.cu:177-180).  The launch index of the reference's
 * versioned flags (ph_kernels.cuh:16-29) is kept on the device (an epoch word
 * advanced by the workgroup that finishes the last chunk, or the ticket
 * counter under IIRX_SCHEDULE_TICKET), so the caller passes none and a
 * captured graph replays correctly.  Calls on one plan must
 * be ordered (one stream, or explicit dependencies); distinct plans are
 * independent.  d_in and d_out must be 16-byte aligned (INVALID otherwise),
 * except for PLR plans, which move one float at a time.  Every kernel takes
 * any n_samples (the reference's cascades require a multiple of the chunk,
 * test_accuracy.cu:37-38): the last chunk may be partial, its missing
 * samples read as zero and not written.  PH, STCR and DTCR run it as a
 * second, one-workgroup launch; the block-state form runs it through the PH
 * last-chunk kernel (64 x 64) from the direct-form histories after its whole
 * chunks, so such a call is four launches.
 */
int iirx_execute(iirx_plan *plan, const float *d_in, float *d_out,
                 void *stream);

/* Streaming state: iirx_execute with the filter state carried in and out, so
 * a long signal can be filtered in calls of n_samples each.  The reference
 * has only zero-state hooks (ph_kernels.cuh:45, 162-165; plr_kernels.cuh:
 * 134-142).  States are direct-form histories in device memory, oldest first:
 *   PH, STCR, DTCR, block-state form: 2 + 2 N_SECTIONS floats, x[-2], x[-1]
 *     of the input, then y_s[-2], y_s[-1] of the output of every section s;
 *   PLR: 2 K floats (K = 2 N_SECTIONS, the recurrence order), x[-K..-1] of
 *     the input, then y[-K..-1] of the output.
 * d_zi NULL: zero state; d_zf NULL: not written.  d_zi is read by the launch
 * (the block-state form maps it to its balanced state in a small kernel
 * first) and d_zf written by a small kernel after it, all on `stream`; d_zf
 * may equal d_zi (of this call or the next).  PH, STCR and DTCR record the
 * end state in their last chunk (samples n-2 and n-1 of the input and of
 * every section's output, wherever the signal ends in the chunk) and the
 * small kernel copies it to d_zf.  The block-state form computes d_zf by
 * running the sections in FP64 over the last tile, from the state the
 * kernel records at its start. */
int iirx_execute_state(iirx_plan *plan, const float *d_in, float *d_out,
                       const float *d_zi, float *d_zf, void *stream);

/*
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  iirx_plan_create followed by iirx_execute_state, with the
 * plan's device memory (coefficient tables, lookback records, control block)
 * in one static device array of IIRX_NOALLOC_SCRATCH_BYTES per GPU, held
 * until the process ends, instead of allocations.  Calls no hipMalloc,
 * hipFree or allocating library routine.  PH, STCR and DTCR only (INVALID
 * otherwise); same arguments, checks and results as iirx_plan_create +
 * iirx_execute_state with this cfg and sos.  Each call recomputes the tables
 * on the host, clears the control block and the records it uses with a fill
 * kernel and uploads the tables as kernel arguments before the filter
 * launch, all on `stream`.  IIRX_ERR_TOO_LARGE if the tables and records do
 * not fit (32 MiB: 2^25 samples at any order and geometry).  Every call on
 * a GPU uses that GPU's one array: calls that overlap on one GPU (different
 * streams, or host threads) race and give wrong results; calls on one
 * stream are ordered and safe.  A lookback watchdog trip (persistent
 * schedule without co-residency) cannot be reported here; use
 * IIRX_SCHEDULE_TICKET where co-residency is not assured.
 */
#define IIRX_NOALLOC_SCRATCH_BYTES (32u << 20)
int iirx_execute_state_noalloc(const iirx_config *cfg, const float *sos, int64_t n_samples,
                               const float *d_in, float *d_out, const float *d_zi,
                               float *d_zf, void *stream);

int iirx_plan_info_get(const iirx_plan *plan, iirx_plan_info *info);

/* Waits for the plan's launches to finish and reports IIRX_ERR_TIMEOUT if any
 * lookback wait since the last check exceeded the spin limit (then clears the
 * flag).  The reference has no failure detection (its spin can hang,
 * ph_kernels.cuh:397-404). */
int iirx_plan_check(iirx_plan *plan);

/* Lookback spin limit in polls (default 1 << 22, several seconds).  Tests
 * force the watchdog path with a tiny limit. */
int iirx_plan_set_spin_limit(iirx_plan *plan, uint32_t polls);

int iirx_plan_destroy(iirx_plan *plan);

/* Validity check of a configuration without a device, the static_assert of
 * gpu_specs::config_is_valid (gpu_specs.hpp:106-113, test_accuracy.cu:41). */
int iirx_config_check(const iirx_config *cfg, int64_t n_samples);

const char *iirx_status_string(int status);

/*
 * measure_kernel_time (measure.cuh:38-110): `warmup` untimed calls, then
 * `batches` batches of `batch` back-to-back calls, each batch bracketed by
 * HIP events; inputs rotate over n_inputs buffers.  Writes the per-call
 * mean of every batch, sorted ascending, to batch_ms[batches] (the P20 is
 * batch_ms[batches/5]).
 */
int iirx_measure(iirx_plan *plan, const float *const *d_inputs, int n_inputs,
                 float *d_out, int warmup, int batches, int batch,
                 void *stream, float *batch_ms);

/*
 * No counterpart in the reference (measure.cuh times only the filter
 * kernels; the reference quotes the DRAM bandwidth from the data sheet,
 * README.md:138-141): the measured HBM ceiling the filters are compared
 * with.  Copies n floats (n a multiple of 4, both buffers 16-byte aligned)
 * with a streaming 16-byte copy kernel on `stream`, the filters' 8 B per
 * sample without arithmetic.  mode 0: nontemporal loads and stores;
 * mode 1: nontemporal loads, stores through the L2 (as the cascades store).
 */
int iirx_copy_ceiling(const float *d_src, float *d_dst, int64_t n, int32_t mode, void *stream);

/* Library version string. */
const char *iirx_version(void);

/* ---- host-side helpers (no device needed) ---------------------------- */

/*
 * Butterworth lowpass of even `order` at cutoff `wn` (x Nyquist), as
 * scipy.signal.butter(order, wn, output='sos') followed by the b0 == 1
 * normalisation of ref_generate.py:61-67: rows {1, 2, 1, 1, a1, a2} in float32
 * and the removed cascade gain in *gain (float64).  Replaces the Python
 * design step of tools/ref_generate.py:56-67.
 */
int iirx_butterworth_sos(int order, double wn, float *sos, double *gain);

/*
 * Chebyshev lowpass of even `order` (2..16) at cutoff `wn` (x Nyquist), as
 * scipy.signal.cheby1(order, db, wn, output='sos') (type 1, passband ripple
 * db) or cheby2(order, db, wn, output='sos') (type 2, stopband attenuation
 * db) followed by iirx_normalize_sos: b0-normalised float32 rows and the
 * removed gain.  No counterpart in the reference, which designs only
 * Butterworth filters (ref_generate.py:62); its iir_utils.hpp:27-141 holds an
 * unused route from direct-form taps to sections.
 */
int iirx_chebyshev_sos(int type, int order, double db, double wn, float *sos, double *gain);

/* General sections into the filter contract: rows {b0, b1, b2, a0, a1, a2}
 * (float64, any scaling, b2 = a2 = 0 for a first-order section) become
 * float32 rows {1, b1/b0, b2/b0, 1, a1/a0, a2/a0} and *gain = prod(b0/a0),
 * the factor the caller applies to the input, as ref_generate.py:61-67 does
 * with the Butterworth gain.  The reference takes only its own Butterworth
 * designs (ref_generate.py:50).  INVALID if a b0 or a0 is 0 or a value is
 * not finite.  Plans reject rows with b0 != 1 or a0 != 1. */
int iirx_normalize_sos(const double *sos_in, int ns, float *sos_out, double *gain);

/* Direct-form taps of the cascade (the PLR input of ref_generate.py:74-81):
 * FP64 polynomial product of the sections; b[0..2ns) FIR taps past b0 == 1,
 * a[0..2ns) feedback taps in the recurrence sign. */
int iirx_direct_form_taps(const float *sos, int ns, double *b, double *a);

/* The documented N(0,1) generator of the synthetic workloads (SplitMix64 +
 * Box-Muller in float64, rounded to float32). */
int iirx_fill_normal(uint64_t seed, int64_t n, float *out);

/* ---- the reference harness's artefacts ------------------------------- */

/*
 * The generated header the reference's drivers compile in
 * (build/filter_taps.hpp, written by ref_generate.py:101-131 and read by
 * test_accuracy.cu:107-109, ph_kernels.cuh:151, plr_kernels.cuh:416-427),
 * as data: FILTER_N_SECTIONS, FILTER_GAIN, INPUT_FS, INPUT_FREQ,
 * FILTER_SOS[ns][6] (b0-normalised, float32), PLR_FILTER_ORDER and the
 * direct-form taps PLR_FILTER_B / PLR_FILTER_A (b0-normalised, FIR taps
 * past b0 == 1, feedback taps in the sign y[n] += A[j] y[n-j-1]).
 * cutoff_wn appears only in the header's comment (0 if not known).
 */
#define IIRX_TAPS_MAX_SECTIONS 8
typedef struct iirx_filter_taps {
    int32_t n_sections;
    int32_t plr_order;
    double gain;
    double input_fs;
    double input_freq;
    double cutoff_wn;
    float sos[IIRX_TAPS_MAX_SECTIONS * 6];
    double plr_b[2 * IIRX_TAPS_MAX_SECTIONS];
    double plr_a[2 * IIRX_TAPS_MAX_SECTIONS];
} iirx_filter_taps;

/* The reference's test filter and signal (ref_generate.py:56-81): the
 * Butterworth lowpass of order 2 n_sections at cutoff_wn (the reference
 * uses 0.2), input_fs = 1000, input_freq = 120; the direct-form taps are
 * the FP64 product of the unrounded sections. n_sections in 1..8. */
int iirx_filter_taps_design(int n_sections, double cutoff_wn, iirx_filter_taps *taps);

/* Writes `taps` as a header with the declarations of ref_generate.py:
 * 106-129 (same names, types and text of every value: floats as %.9e with
 * an f suffix, doubles as the shortest text that reads back exactly), so
 * the reference's drivers compile against it. IO error: INVALID. */
int iirx_filter_taps_write(const iirx_filter_taps *taps, const char *path);

/* Reads such a header (one written by ref_generate.py or by
 * iirx_filter_taps_write): comments are skipped, float values are read as
 * a compiler reads an f-suffixed literal (one rounding to float32).
 * INVALID if the file cannot be read, a declaration is missing, a count
 * disagrees with FILTER_N_SECTIONS / PLR_FILTER_ORDER or n_sections is
 * outside 1..8. */
int iirx_filter_taps_read(const char *path, iirx_filter_taps *taps);

/*
 * Host copy of the coefficient tables a plan of this configuration would
 * upload (FP64 precompute rounded to float32).  Writes min(cap, size)
 * floats to out (out may be NULL to query), the total size to *size and the
 * per-section stride to *stride.  No device needed.  A block-state plan
 * computes the same tables in FP64 on the GPU instead (bit-identical).
 */
int iirx_plan_tables(const iirx_config *cfg, const float *sos,
                     int64_t n_samples, float *out, int64_t cap,
                     int64_t *size, int32_t *stride);
/* The tables a plan holds on the device, copied back: min(cap, size) floats
 * to out (out may be NULL to query the size).  No counterpart in the
 * reference; for checking the device-built tables against iirx_plan_tables. */
int iirx_plan_read_tables(const iirx_plan *p, float *out, int64_t cap,
                          int64_t *size);

/* ------------------------------------------------------------------------
 * Autotuner: replaces the reference's fixed per-order geometry sweep and
 * PLR_X_MAX calibration (run_accuracy.sh:95-151, run_performance.sh) with a
 * runtime search.  Candidates: PH (N_BLOCKS 8..128), STCR and DTCR (loop and
 * hand-unrolled, N_BLOCKS 8..128), the block-state form (G x L) and PLR
 * (automatic geometry, spill-calibrated X; orders <= 4 unless
 * IIRX_TUNE_PLR_ALL_ORDERS).  A candidate is excluded when the plan would
 * reject it: no instance or too many chunks (IIRX_ERR_GEOMETRY), or CR
 * factors ill-conditioned for the filter (IIRX_ERR_CONDITIONING: STCR/DTCR
 * fall back to PH/BS).  Every kernel takes any length.
 * ------------------------------------------------------------------------ */
enum {
    IIRX_TUNE_NO_BLOCKSTATE = 1,   /* plain fp32 kernels only */
    IIRX_TUNE_NO_CR = 2,           /* no STCR/DTCR candidates */
    IIRX_TUNE_PLR_ALL_ORDERS = 4,  /* PLR also above order 4 (it fails the
                                      accuracy gate at order 16) */
    IIRX_TUNE_TICKET = 8           /* ticket schedule instead of persistent */
};

typedef struct {
    iirx_config cfg;
    int32_t status;   /* iirx_status of the candidate's tables; OK = eligible */
    float ms;         /* mean ms per launch (iirx_autotune), -1 if not timed */
    float rel_err;    /* iirx_autotune: max|out - ref| / max|ref| over the
                         first min(n, 2^18) samples against the host FP64
                         cascade, -1 if not timed */
} iirx_tune_entry;

/* Host only: the candidate list for (sos, n_samples) with each candidate's
 * eligibility.  Writes up to cap entries, *count = the full number. */
int iirx_tune_candidates(const float *sos, int32_t n_sections, int64_t n_samples,
                         uint32_t flags, iirx_tune_entry *entries, int32_t cap,
                         int32_t *count);
/* Time every eligible candidate (reps launches after 3 warm-up launches on
 * N(0,1) device buffers of n_samples floats, rotating over up to 8 inputs so
 * that long signals stream from HBM), screen its output against the host
 * FP64 cascade over the first min(n, 2^18) samples (the reference verifies
 * accuracy after its timing runs, run_performance.sh), and return the
 * fastest candidate that passes in *best / *best_ms.  A candidate passes if
 * max|err| <= 1e-4 max|ref| (the reference's gate, test_accuracy.cu:178-190)
 * or, for filters where no float32 form meets that, within 10x the most
 * accurate candidate's error; the others get IIRX_ERR_ACCURACY.  entries
 * (optional) receive the full table. */
int iirx_autotune(const float *sos, int32_t n_sections, int64_t n_samples,
                  uint32_t flags, int32_t reps, iirx_config *best,
                  float *best_ms, iirx_tune_entry *entries, int32_t cap,
                  int32_t *count);

/* ------------------------------------------------------------------------
 * Ragged batch: no counterpart in the reference, which filters one channel
 * per binary (test_accuracy.cu:84-205).  Many independent channels, each
 * with its own length, order and coefficients, packed back to back (CSR):
 * channel c is samples [offsets[c], offsets[c+1]) of the input and of the
 * output, filtered with sections sos[sec_off[c] .. sec_off[c] + n_sections[c])
 * where sec_off is the running sum of n_sections (rows of 6 floats, the same
 * contract as iirx_plan_create), zero initial state per channel.  One
 * persistent launch filters every channel, chunk by chunk (64 x n_blocks
 * samples), with lookback chains per channel; a channel's last chunk may be
 * partial.  Each channel runs PH (ph_kernels.cuh PH_64) or the block-state
 * form (tiles of 64 samples, bs_tiles tiles per chunk, matrix cores); both
 * give the reference's recurrence within the FP32 bounds of iirx_plan.  When
 * both forms are present they are two persistent launches, one after the
 * other on the stream, each with its own register budget.  The launches can
 * be captured into a hipGraph and replayed: the lookback epochs live on the
 * device, so replays stay correct.
 * ------------------------------------------------------------------------ */
typedef struct iirx_batch iirx_batch;

typedef struct {
    int32_t n_blocks;         /* samples per lane: 16, 32 or 64 (0 = 32)     */
    int32_t schedule;         /* iirx_schedule                               */
    int32_t bs_min_sections;  /* IIRX_BATCH_AUTO channels with at least this
                                 many sections run the block-state form (0 =
                                 8, i.e. order 16, where it measured faster
                                 than PH; above 8 = none)                    */
    int32_t bs_tiles;         /* block-state tiles per chunk: 16, 32 or 64
                                 (0 = 32)                                    */
} iirx_batch_config;

/* per-channel algorithm of iirx_batch_create_ex */
enum {
    IIRX_BATCH_AUTO = -1,      /* by cfg->bs_min_sections */
    IIRX_BATCH_PH = 0,
    IIRX_BATCH_BLOCKSTATE = 1
};

typedef struct {
    int64_t n_channels;
    int64_t n_samples;      /* offsets[n_channels] - offsets[0] */
    int64_t num_chunks;     /* both forms */
    int64_t chunk_samples;  /* PH chunk */
    int64_t grid;           /* PH launch */
    int64_t launches;       /* execute/replay calls so far */
    int32_t vgprs, waves_per_simd;        /* PH kernel */
    int32_t graph_ready;    /* a captured graph can be replayed */
    int32_t n_launches;     /* kernel launches per call: one per form present */
    int64_t n_blockstate;   /* channels that run the block-state form */
    int64_t bs_chunks;
    int64_t bs_chunk_samples;
    int64_t bs_grid;
    int32_t bs_vgprs, bs_waves_per_simd;  /* block-state kernel */
} iirx_batch_info;

/* Validate the channels and build the per-channel FP64 tables on the host,
 * then upload them.  offsets: n_channels + 1 non-decreasing values, offsets[0]
 * >= 0; n_sections: 1..8 per channel. */
int iirx_batch_create(const iirx_batch_config *cfg, int64_t n_channels,
                      const int64_t *offsets, const int32_t *n_sections,
                      const float *sos, iirx_batch **out);
/* As iirx_batch_create with each channel's algorithm chosen: algo[c] is
 * IIRX_BATCH_AUTO, IIRX_BATCH_PH or IIRX_BATCH_BLOCKSTATE (algo NULL: all
 * AUTO). */
int iirx_batch_create_ex(const iirx_batch_config *cfg, int64_t n_channels,
                         const int64_t *offsets, const int32_t *n_sections,
                         const float *sos, const int32_t *algo,
                         iirx_batch **out);
/* One launch over every channel; d_in/d_out hold at least offsets[n_channels]
 * floats and must not overlap. */
int iirx_batch_execute(iirx_batch *b, const float *d_in, float *d_out,
                       void *stream);
/* Capture the launch for (d_in, d_out) into a hipGraph on `stream` (not the
 * null stream) and instantiate it; iirx_batch_replay launches it. */
int iirx_batch_capture(iirx_batch *b, const float *d_in, float *d_out,
                       void *stream);
int iirx_batch_replay(iirx_batch *b, void *stream);
/* Synchronise and report the lookback watchdog, as iirx_plan_check. */
int iirx_batch_check(iirx_batch *b);
/* The geometry of a batch by measurement (no counterpart in the reference,
 * whose tile geometry is fixed per binary): every combination of n_blocks
 * {16, 32, 64}, bs_tiles {16, 32, 64} and bs_min_sections {4, 8, 9 (none)}
 * with cfg->schedule, each timed over `reps` launches of the whole batch on
 * N(0,1) inputs after 2 warm-up launches; the fastest configuration goes to
 * *best and its mean ms per launch to *best_ms. */
int iirx_batch_autotune(const iirx_batch_config *cfg, int64_t n_channels, const int64_t *offsets,
                        const int32_t *n_sections, const float *sos, int32_t reps,
                        iirx_batch_config *best, float *best_ms);
int iirx_batch_info_get(const iirx_batch *b, iirx_batch_info *info);
/* The batch's per-channel tables on the device, copied back (channel after
 * channel, each 16-byte aligned): min(cap, size) floats to out (out may be
 * NULL to query the size). */
int iirx_batch_read_tables(const iirx_batch *b, float *out, int64_t cap,
                           int64_t *size);
int iirx_batch_destroy(iirx_batch *b);

#ifdef __cplusplus
}
#endif

#endif /* IIRX_H */
