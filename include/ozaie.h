/*
 * ozaie.h -- C ABI of the MI355X-native Ozaki-scheme DGEMM library (libozaie.so).
 *
 * Plain C types only (pointers, sizes, ints); usable from C, C++, ctypes, cgo, JNI ...
 * The reference's C++ API (namespace oz_aie in include/oz_aie.h of atlarge-research/ozAIE)
 * is provided on top of this ABI by include/oz_aie.h in this directory.
 *
 * Arithmetic: C = A * B with A (M x K) and B (K x N) FP64.  Every element is split into
 * d signed b-bit slices relative to its row (A) / column (B) maximum exponent, the
 * d(d+1)/2 slice products are exact int8 x int8 -> int32 MFMA GEMMs accumulated into d
 * bins, the bins are summed exactly in 128-bit integers and converted to FP64.
 * With rounding OZAIE_ROUND_TRUNCATE (the default) the result is bit-identical to the
 * reference rule (include/oz_aie.h:92-150 of the reference).
 */
#ifndef OZAIE_H
#define OZAIE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes.  0..4 keep the reference's meanings (include/oz_aie.h:187-248,
 *      337-340 of the reference); 5.. are new.                                          */
#define OZAIE_OK 0
#define OZAIE_ERR_DIMS 1        /* dimensions rejected (reference: non-square / K range) */
#define OZAIE_ERR_ENV 2         /* reference: environment variable missing; never returned */
#define OZAIE_ERR_INCOMPLETE 3  /* reference: run did not complete; here: kernel failure  */
#define OZAIE_ERR_BUILD 4       /* reference: xclbin build failed; never returned          */
#define OZAIE_ERR_HIP 5         /* a HIP runtime call failed                               */
#define OZAIE_ERR_NODEVICE 6    /* no usable GPU                                           */
#define OZAIE_ERR_NOSETUP 7     /* DGEMM called before a successful setup                  */
#define OZAIE_ERR_MISMATCH 8    /* DGEMM dimensions differ from the last setup             */
#define OZAIE_ERR_NOMEM 9       /* device allocation failed                                */
#define OZAIE_ERR_ARG 10        /* invalid argument (null pointer, bad bits/slices, ...)   */
#define OZAIE_ERR_OVERFLOW 11   /* product too long for exact 128-bit accumulation         */

/* ---- modes.  0 and 1 are the reference's (oz_aie::Mode, include/oz_aie.h:82):
 *   ACCURATE  b = 3, d = 29/28/27/27/26/25/25 for log2 K = 7..13, b = 4, d = 18 for 14+
 *   FAST      b = 7, d = 8
 * 2 and 3 are the int8-rate presets of this library: b = 7 with 3 and 7 slices.        */
#define OZAIE_MODE_ACCURATE 0
#define OZAIE_MODE_FAST 1
#define OZAIE_MODE_FAST3 2
#define OZAIE_MODE_ACCURATE7 3
#define OZAIE_MODE_CRT 4        /* no reference counterpart: exact integer product of the
                                   operands, each row of A and column of B scaled by a
                                   power of two from an exact bound on its Euclidean norm
                                   so that 16 int8 residue GEMMs rebuild it exactly
                                   (ozaie_plan_create, flag OZAIE_PLAN_CRT); K <= 65536   */

/* ---- options (ozaie_set_option) */
#define OZAIE_OPT_ROUNDING 1    /* OZAIE_ROUND_TRUNCATE (default) or OZAIE_ROUND_NEAREST   */
#define OZAIE_OPT_STRICT 2      /* 1: ozAIE_setup rejects what the reference rejects        */
#define OZAIE_OPT_KCHUNK 3      /* cap on contraction indices per pass (0 = automatic)      */
#define OZAIE_OPT_KERNEL 4      /* 0 auto, 1 bin-group kernel, 2 fused with LDS copies,
                                   3 fused with direct register loads, 4 fused pipelined,
                                   5 fused staged (4-8 LDS stages; d <= 4 and d = 9,
                                     else as 2),
                                   6 fused with LDS copies and rolling fragment reads;
                                   CRT plans: 2 the 256 x 256 modulus GEMM, 3 its per-wave-
                                   load form, 4 128 x 128 tiles, others by size           */
#define OZAIE_OPT_PHASE_TIMING 5 /* 1: plans record HIP events around the split and GEMM    */
#define OZAIE_OPT_CRT_GROUP_MODULI 6 /* moduli per workgroup of the grouped CRT modulus GEMM
                                   (0 = automatic per tile shape); read at group creation */
#define OZAIE_ROUND_TRUNCATE 0
#define OZAIE_ROUND_NEAREST 1

/* ---- reference-compatible entry points (process-global plan, blocking, device 0)     */

/* Replaces oz_aie::ozAIE_setup (include/oz_aie.h:182-300 of the reference).
 * Picks (b, d) for `mode` from K and prepares the device buffers.  Any M, N, K >= 1 is
 * accepted unless OZAIE_OPT_STRICT is set (then only M == N == K with int(log2 K) in
 * [7, 11], as the reference checks).  Returns OZAIE_OK or an error code.              */
int32_t ozAIE_setup(int M, int N, int K, int mode);

/* Explicit slice configuration: b bits per slice (1..7), d slices (1..32), b*d <= 96. */
int32_t ozAIE_setup_slices(int M, int N, int K, int bits, int slices);

/* Replaces oz_aie::ozAIE_DGEMM (include/oz_aie.h:311-349 of the reference).
 * C = A * B, row-major, contiguous, dimensions must equal the last setup.  A, B and C
 * may be host or device pointers.  Blocking.  Host arrays are copied through two pinned
 * 32 MiB chunks that the global plan allocates at its first host-pointer call and holds
 * until the next setup or ozaie_release.                                              */
int32_t ozAIE_DGEMM(int M, int N, int K, const double *A, const double *B, double *C);

/* Set / get a library option (OZAIE_OPT_*).  Applies to plans set up afterwards.       */
int32_t ozaie_set_option(int option, int value);
int32_t ozaie_get_option(int option);

/* Number of (b, d) and passes of the current global plan:
 * out[0]=b out[1]=d out[2]=k-chunk out[3]=passes out[4]=kernel (0 fused, 1 group,
 * 2 CRT mode: then b = the max rule's bits per operand (0: the norm rule) and d = number
 * of moduli, 3 paired digits: b <= 3
 * with d >= 8, adjacent digits paired into one int8 digit, the same C, 4 bin windows:
 * d = 18 as two passes of 9 bins on the rolling fused kernel)                          */
int32_t ozaie_plan_info(int32_t *out5);

/* Release every device buffer and plan (reference: cleanup_xrt, include/oz_aie.h:153). */
int32_t ozaie_release(void);

/* ---- device-pointer API (explicit plans, asynchronous on a HIP stream)               */
typedef struct ozaie_plan_s *ozaie_plan;

/* Create a plan for C (M x N) = op(A) (M x K) * op(B) (K x N) with b bits and d slices.
 * flags: bit 0 -> nearest-even rounding (default truncation).
 * flags bit 2 (OZAIE_PLAN_CRT) -> CRT mode: `slices` = number of moduli (2..20), `bits`
 * = 0 for the norm rule: per row / column q = round(x 2^s) with the largest s that the
 * Cauchy-Schwarz bound ||q_a|| ||q_b|| <= M / 4 allows, from an exact integer bound on the
 * norm (about log2(max|x| sqrt(K) / ||x||) bits more than the max rule; for a few moduli,
 * M 2^20 <= K (2^20 + 2^40), the max rule with the largest bit count instead); or `bits`
 * = beta > 0 for the max rule: |q| <= 2^beta from the row / column maximum exponent, beta
 * no more than the moduli rebuild exactly for this K.  No reference counterpart;
 * K <= 65536.                                                                           */
#define OZAIE_PLAN_NEAREST 1
#define OZAIE_PLAN_CRT 4
int32_t ozaie_plan_create(ozaie_plan *plan, int64_t M, int64_t N, int64_t K, int bits,
                          int slices, int flags);
int32_t ozaie_plan_destroy(ozaie_plan plan);
int32_t ozaie_plan_query(ozaie_plan plan, int32_t *out5);

/* Enqueue C = op(A) * op(B) on `stream` (a hipStream_t, 0 = default stream).  All
 * pointers are device pointers.  transA: 0 -> A is M x K row-major with leading
 * dimension lda; 1 -> A is stored K x M (op(A) = A^T).  Same for B (N vs K).           */
int32_t ozaie_dgemm(ozaie_plan plan, int transA, int transB, const double *A, int64_t lda,
                    const double *B, int64_t ldb, double *C, int64_t ldc, void *stream);

/* Milliseconds of the last ozaie_dgemm on `plan` (needs OZAIE_OPT_PHASE_TIMING = 1 when
 * the plan was created): out[0] exponent max + split, out[1] slice GEMM passes (the
 * fused GEMM kernel, or all bin-group or bin-window passes; in CRT mode the modulus GEMM kernel
 * alone), out[2] whole call (CRT: including the rebuild).  Synchronizes on the events
 * of that call.                                                                        */
int32_t ozaie_plan_phase_times(ozaie_plan plan, float *out3);

/* ---- allocation-free product.  No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The same product as ozaie_plan_create(M, N, K, bits, slices,
 * flags) followed by ozaie_dgemm(...) on `stream`, bit-identical to it, but every
 * intermediate (exponents, slices or residues, 128-bit partial sums, residue products)
 * lives in a static device workspace of OZAIE_NOALLOC_BYTES per GPU, held until the
 * process ends; no call allocates or frees memory.  A product whose intermediates exceed
 * the workspace is rejected with OZAIE_ERR_NOMEM (ozaie_last_error gives the sizes).
 * The caller passes A, B, C (device pointers).  All calls on one GPU share its workspace:
 * calls that overlap on one GPU (other streams, other threads) race; calls on one stream
 * are ordered.  Asynchronous on `stream`.                                               */
#define OZAIE_NOALLOC_BYTES 268435456LL   /* 256 MiB */
int32_t ozaie_dgemm_noalloc(int64_t M, int64_t N, int64_t K, int bits, int slices, int flags, int transA,
                            int transB, const double *A, int64_t lda, const double *B, int64_t ldb, double *C,
                            int64_t ldc, void *stream);

/* ---- grouped products (many independent C_i = A_i B_i of uneven shape)
 * A_i (M_i x K_i), B_i (K_i x N_i), C_i (M_i x N_i): row-major, contiguous, device pointers.
 * All problems share b.  The slice count is `slices` for every problem, or with slices =
 * 0 each problem's size class's entry of the slice-count table (ozaie_tune_set_slices;
 * 9 where untuned).  Slice counts 3, 7, 8, 9 run as one exponent launch, one split launch
 * and one fused-GEMM launch per (slice count, tile configuration), each problem's tile
 * configuration chosen per size class by the autotuner table of its slice count (or a
 * heuristic when the class is untuned); a group with any other slice count runs one plan
 * per problem.  Every K_i must fit one contraction pass.  The launches are captured once
 * in a hipGraph (flags bit 1 disables the graph) and replayed by ozaie_group_run;
 * descriptors are re-uploaded only when the pointers change.  flags bit 0: nearest-even
 * rounding.
 * flags bit 2 (OZAIE_GROUP_PERSISTENT): ONE persistent launch (one workgroup per CU:
 * exponent max + split of every operand, a grid barrier, then every GEMM tile, balanced
 * per XCD on the host), each problem with its own slice count and with its size class's
 * tile configuration mapped to the persistent kernel's tiles (direct register loads:
 * 64x64 for every slice count, 64x128 for 3 and 7, 128x128 for 3).  flags bits 4..7:
 * forced persistent tile configuration + 1 (0: per size class; ids as for
 * ozaie_autotune), applied to the problems whose slice count has it.
 * flags bit 3 (OZAIE_GROUP_CRT): the CRT mode for every problem (no reference
 * counterpart): `slices` = moduli (2..20), `bits` = 0 for the norm rule per problem (as
 * ozaie_plan_create) or a fixed max-rule bit count no larger than any problem's allows;
 * K_i <= 65536.  Four phase launches plus one modulus-GEMM launch per tile shape (ids as
 * for ozaie_autotune_crt; flags bits 4..7 = forced shape + 1), same graph capture; each
 * C_i is bit-identical to a CRT plan of that problem.  With bit 2 as well, ONE persistent
 * launch (two workgroups per CU): shifts + residues of every operand, a grid
 * barrier, every modulus-GEMM tile (each problem in its size class's tile shape, 256 x 256
 * run as 128 x 128 tiles, all moduli per tile), a second grid barrier, the rebuild of
 * every C_i.
 * Runs of one group must not overlap (they share the workspace and the control block). */
#define OZAIE_GROUP_NEAREST 1
#define OZAIE_GROUP_NO_GRAPH 2
#define OZAIE_GROUP_PERSISTENT 4
#define OZAIE_GROUP_CRT 8
typedef struct ozaie_group_s *ozaie_group;
int32_t ozaie_group_create(ozaie_group *group, int count, const int64_t *M, const int64_t *N,
                           const int64_t *K, int bits, int slices, int flags);
/* A, B, C: host arrays of `count` device pointers.  Asynchronous on `stream`.         */
int32_t ozaie_group_run(ozaie_group group, const double *const *A, const double *const *B,
                        double *const *C, void *stream);
/* out[0] kernel launches per run, out[1] graph captured, out[2] per-problem fallback,
 * out[3] GEMM workgroups per run (fused-GEMM tiles; CRT: modulus-GEMM workgroups).   */
int32_t ozaie_group_query(ozaie_group group, int32_t *out4);
/* Tile configuration chosen for each problem (-1 in the per-problem fallback).        */
int32_t ozaie_group_configs(ozaie_group group, int32_t *cfg_out);
/* Waits for the group's last run and reports (then clears) the persistent kernel's error
 * report: 1 if a grid-barrier wait exceeded its bound in any run since the previous
 * status call (the C of such a run is invalid; this happens only when the grid could not
 * be resident at once).  Each launch starts with a clear barrier, so a failed run does
 * not fail the runs after it.  0 otherwise, and always 0 for the multi-launch modes.   */
int32_t ozaie_group_status(ozaie_group group, int32_t *error_out);
int32_t ozaie_group_destroy(ozaie_group group);
/* One-shot grouped product without a graph; blocking on `stream`.                     */
int32_t ozaie_dgemm_grouped(int count, const int64_t *M, const int64_t *N, const int64_t *K,
                            const double *const *A, const double *const *B, double *const *C,
                            int bits, int slices, void *stream);

/* Autotuner: times every tile configuration on representative shapes of the 9 size
 * classes (max(M,N) and K each in {<=256, <=1024, >1024}) and keeps the fastest in the
 * process-wide table (table9: class 3*mn_bucket + k_bucket -> configuration id;
 * 0: 64x64 direct loads, 1: 64x128 direct, 2: 128x128 direct, 3: 64x64 LDS copies,
 * 4: 64x128 LDS copies).  reps: timed runs per candidate.                             */
int32_t ozaie_autotune(int bits, int slices, int reps, int32_t *table9);
int32_t ozaie_tune_get(int slices, int32_t *table9);   /* -1 entries: untuned */
/* Slice count per size class for groups created with slices = 0 (3, 7, 8 or 9; -1 =
 * untuned, 9).  ozaie_amd.tune.tune_slices calibrates it: per class the cheapest count
 * whose maximum error on a calibration sample is no worse than native FP64 DGEMM's
 * (the analogue of the reference's per-log2 K choice of (b, d), include/oz_aie.h:60-80). */
int32_t ozaie_tune_get_slices(int32_t *table9);
int32_t ozaie_tune_set_slices(const int32_t *table9);
int32_t ozaie_tune_set(int slices, const int32_t *table9);
/* The same for CRT groups with `moduli` moduli (no reference counterpart); tile shape
 * ids 0: 256x256, 1: 128x128, 2: 64x64 (modulus GEMM through LDS).                     */
int32_t ozaie_autotune_crt(int moduli, int reps, int32_t *table9);
int32_t ozaie_tune_get_crt(int moduli, int32_t *table9);
int32_t ozaie_tune_set_crt(int moduli, const int32_t *table9);

/* ---- inspection (reference full_dump equivalent, whole_array_iron_ozaki.py:485-558)  */

/* CRT-mode intermediates of C = A * B (row-major device A, B; `bits` as for
 * ozaie_plan_create), as computed on the GPU:
 * SA_out int8[moduli][M][K] and SB_out int8[moduli][N][K] residues (host), G_out
 * uint8[moduli][M][N] residue products mod p (host).  No reference counterpart.        */
int32_t ozaie_debug_crt(const double *A, const double *B, int64_t M, int64_t N, int64_t K, int moduli,
                        int bits, int8_t *SA_out, int8_t *SB_out, uint8_t *G_out);

/* The epilogue's conversion of exact integers T (= T_hi 2^64 + T_lo, two's complement) to
 * FP64, C = T 2^(es - 2044 - bits (slices + 1)), as computed on the GPU: out[i] through the
 * 128-bit path, out64[i] through the 64-bit path when T fits int64 (else 0).  rounding:
 * OZAIE_ROUND_TRUNCATE (the reference's host conversion, include/oz_aie.h:92-150) or
 * OZAIE_ROUND_NEAREST.  Host arrays of n entries.                                       */
int32_t ozaie_debug_convert(const int64_t *T_hi, const uint64_t *T_lo, const int32_t *es, int64_t n, int bits,
                            int slices, int rounding, double *out, double *out64);

/* Exponents and slices of one operand, as computed on the GPU, in plain layout.
 * X is a device pointer to a rows x K view: element (r, k) at X[r*sr + k*sk].
 * E_out: int32[rows] (host), S_out: int8[d][rows][K] (host).                          */
int32_t ozaie_debug_split(const double *X, int64_t rows, int64_t K, int64_t sr, int64_t sk,
                          int bits, int slices, int32_t *E_out, int8_t *S_out);

/* The d int32 bins of C = A * B (row-major device A, B) as computed by the slice GEMM
 * kernels (bins_out: host int32[d][M][N]).  K must fit one pass.  kernel: as
 * OZAIE_OPT_KERNEL (0 auto, 1 bin-group, 2-6 the fused families).                    */
int32_t ozaie_debug_bins(int64_t M, int64_t N, int64_t K, const double *A, const double *B,
                         int bits, int slices, int kernel, int32_t *bins_out);

/* Reference wire format: the paired int64 {L, H} per element that the reference's NPU
 * returns (ozaki_base.cc:143-340), computed from GPU bins (host int64[M][N][2]).      */
int32_t ozaie_paired_output(int64_t M, int64_t N, int64_t K, const double *A,
                            const double *B, int bits, int slices, int64_t *out);

/* ---- accuracy measurement (host, OpenMP)
 * Correctly rounded exact values of selected entries of A * B (row-major host arrays):
 * out[i] = round(sum_k A[rows[i]][k] * B[k][cols[i]]), every product split error-free
 * with an FMA and the terms summed exactly.  Used to score results against the
 * exact product (the reference scores against MPFR-106, common.h:164-232).           */
int32_t ozaie_exact_entries(const double *A, int64_t lda, const double *B, int64_t ldb, int64_t K,
                            const int64_t *rows, const int64_t *cols, int64_t count, double *out);

/* ---- measured int8 MFMA ceiling (diagnostic; no reference counterpart)
 * Sustained rate, in dense int8 POP/s (one MAC = 2 ops), of v_mfma_i32_32x32x32_i8 run from
 * registers only on every CU (one wave per SIMD, 16 independent accumulator tiles per
 * wave), for `iters` rounds of 128 MFMAs after one warm-up launch.  pattern 0: all-zero
 * operands; 1: one random fragment pair for every MFMA (constant operands); 2: eight random
 * fragment pairs in turn (operands toggle like slice or residue data); 3: pattern 2 on
 * v_mfma_i32_16x16x64_i8 (the same int8 operations per round); 4: eight fragment pairs of
 * bytes uniform in [-127, 127] (symmetric residues, sign-magnitude slices).  The rate the
 * matrix cores sustain depends on the operand data through power and clock; this is the
 * measured ceiling the GEMM kernels are compared with beside the datasheet peak.
 * ms_out (optional): milliseconds of the timed launch.                                  */
int32_t ozaie_probe_mfma(int pattern, int iters, double *pops_out, double *ms_out);

/* Last HIP error string (for diagnostics). */
const char *ozaie_last_error(void);

/* Library version, e.g. 100 = 0.1.0 */
int32_t ozaie_version(void);

#ifdef __cplusplus
}
#endif
#endif /* OZAIE_H */
