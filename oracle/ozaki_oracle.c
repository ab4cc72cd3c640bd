/*
 * ozaki_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference Ozaki-scheme DGEMM path (atlarge-research/ozAIE).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this
 * library, and only as the checker (or the timed CPU baseline).  The product library
 * (ozaie_amd/libozaie.so) never links or calls it.
 *
 * Two independent restatements of the split are kept on purpose:
 *   oro_split_mech()   follows the reference's word-level steps (32-bit word selection,
 *                      the acc64 multiply-shift window, exp2 with its 2^31 clamp, the
 *                      shift-out across one or two 64-bit accumulators and floor digit
 *                      peeling) -- src/aie_kernels/generic/ozaki_base.cc:449-639, with
 *                      exp2 from src/aie_kernels/aie2/mm_ozaki.cc:36-42.
 *   oro_split()        the closed form Q = floor(M_x * 2^(b*d - 53 - bias)),
 *                      s_j = sign * ((Q >> b*(d-1-j)) & (2^b - 1)), valid for b*d <= 96.
 * The two are compared on random and edge inputs by the tests; the GPU kernels follow
 * the closed form.
 *
 * Likewise the bin summation has a mechanistic restatement of the paired-int64 wire
 * format (ozaki_base.cc:143-340) plus the host conversion (include/oz_aie.h:92-150),
 * and a closed form (exact 128-bit T, truncation to 53 bits).
 *
 * Parity status of each function (what pins it) is listed in the header comment of
 * each function and in the test files that exercise it.
 */
#include <stdint.h>
#include <string.h>
#include <stdlib.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef __int128 i128;
typedef unsigned __int128 u128;

static inline uint64_t dbits(double x) { uint64_t u; memcpy(&u, &x, 8); return u; }
static inline double bitsd(uint64_t u) { double x; memcpy(&x, &u, 8); return x; }

/* Biased exponent field, (hi32 << 1) >> 21 -- ozaki_base.cc:405-406. */
int oro_field(double x) {
    uint32_t hi = (uint32_t)(dbits(x) >> 32);
    return (int)((hi << 1) >> 21);
}

/* Row-wise maximum exponent field of a rows x cols operand, element (r,c) at
 * X[r*sr + c*sc].  Running max initialised to 0 -- ozaki_base.cc:342-440 (running max
 * across K tiles at :425, zero init at :820-825). Pinned by brute force in the tests. */
void oro_max_exp(const double *X, int64_t rows, int64_t cols, int64_t sr, int64_t sc,
                 int16_t *E) {
    #pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < rows; r++) {
        int m = 0;
        for (int64_t c = 0; c < cols; c++) {
            int f = oro_field(X[r * sr + c * sc]);
            if (f > m) m = f;
        }
        E[r] = (int16_t)m;
    }
}

/* XDNA1 exp2 workaround, mm_ozaki.cc:36-42: float bit trick, values >= 2^31 clamp to 2^31. */
static inline uint64_t exp2_quirk(uint32_t p) { return p >= 31u ? (1ull << 31) : (1ull << p); }

/* Word-level restatement of one element of extract_splits_acc64_half_half
 * (ozaki_base.cc:449-639).  Writes d signed digits to out[0..d-1] (out[j*stride]). */
void oro_split_mech_elem(double x, int emax, int b, int d, int8_t *out, int64_t stride) {
    uint64_t u = dbits(x);
    uint32_t hi = (uint32_t)(u >> 32), lo = (uint32_t)u;
    int is_neg = ((int32_t)hi) < 0;                                 /* :497 */
    uint32_t e = (hi << 1) >> (32 - 11);                            /* :503-504 */
    int is_sub = (e == 0u);                                         /* :507 */
    uint32_t bias = (uint32_t)emax - e;                             /* :509 */
    hi = (hi << 12) >> 12;                                          /* :513-514 */
    hi += is_sub ? 0u : (1u << (32 - (11 + 1)));                    /* :516 */
    hi = (hi << 11) + (lo >> (32 - 11));                            /* :524-525 */
    lo = lo << 11;                                                  /* :527 */

    const int bits_to_avoid = 1 + (b == 3 ? 3 : 0);                 /* :459 */
    const int nv0 = 64 - bits_to_avoid - ((64 - bits_to_avoid) % b);/* :461-464 */
    const int nv1 = 64 - bits_to_avoid;
    const int nacc = (d * b) <= nv0 ? 1 : 2;                        /* :465 */
    const int nspa[2] = { d < nv0 / b ? d : nv0 / b, d - nv0 / b }; /* :466 */

    uint32_t s[4] = {0, 0, 0, 0};                                   /* :534-546 */
    s[0] = bias >= 32u ? 0u : hi;
    s[1] = bias >= 32u ? hi : lo;
    s[1] = bias >= 64u ? 0u : s[1];
    if (nacc == 2) {
        s[2] = bias >= 32u ? lo : 0u;
        s[2] = bias >= 64u ? hi : s[2];
        s[2] = bias >= 96u ? 0u : s[2];
        s[3] = bias >= 64u ? lo : 0u;
        s[3] = bias >= 96u ? hi : s[3];
    }
    uint32_t beta = (bias << (32 - 5)) >> (32 - 5);                 /* :553 (logical) */
    uint64_t pw = exp2_quirk(32u - beta);                           /* :555 */

    uint64_t acc[2] = {0, 0}, tmp = 0;                              /* :558-561 */
    for (int a = 0; a < nacc; a++) {                                /* :564-578 */
        acc[a] += (uint64_t)s[2 * a] * pw;
        acc[a] += (uint64_t)(beta == 0u ? s[2 * a] : 0u) * pw;
        tmp = (uint64_t)s[2 * a + 1] * pw;
        tmp += (uint64_t)(beta == 0u ? s[2 * a + 1] : 0u) * pw;
        acc[a] += (uint64_t)(uint32_t)(tmp >> 32);
        if (nacc == 2 && a == 0) acc[1] = (uint64_t)(uint32_t)tmp << 32;
    }

    {                                                               /* :583-606 */
        const int nso1 = 64 - nv0;
        const int nso2 = (64 - nv0) + (64 - nv1);
        uint32_t v = (uint32_t)(acc[0] >> 32);
        v = v >> nso1;
        uint64_t t = (uint64_t)v << 32;
        v = (uint32_t)(acc[0] >> nso1);
        t += v;
        uint32_t vlow = (uint32_t)acc[0];
        acc[0] = t;
        if (nacc == 2) {
            v = vlow << (32 - nso1);
            t = (uint64_t)v << (32 - (64 - nv1));
            t += (uint64_t)(uint32_t)(acc[1] >> 32) * (uint64_t)(1u << (32 - nso2));
            t += (uint64_t)((uint32_t)acc[1] >> nso2);
            acc[1] = t;
        }
    }

    int split_no = 0;                                               /* :612-628 */
    for (int a = 0; a < nacc; a++) {
        const int nv = a == 0 ? nv0 : nv1;
        for (int ls = 0; ls < nspa[a]; ls++, split_no++) {
            int shift = nv - b * (ls + 1);
            int8_t v = (int8_t)(int64_t)(acc[a] >> shift);   /* floor, no saturation */
            acc[a] -= (uint64_t)(int64_t)v << shift;
            out[(int64_t)split_no * stride] = (int8_t)(is_neg ? -v : v);
        }
    }
}

/* Closed form of the split for one element (see file header).  Valid for b*d <= 96. */
void oro_split_elem(double x, int emax, int b, int d, int8_t *out, int64_t stride) {
    uint64_t u = dbits(x);
    int neg = (int)(u >> 63);
    int e = (int)((u >> 52) & 0x7FF);
    uint64_t mx = (u & ((1ull << 52) - 1)) | (e ? (1ull << 52) : 0ull);
    int bias = emax - e;
    int mask = (1 << b) - 1;
    for (int j = 0; j < d; j++) {
        /* digit j = bits [53 + bias - b*(j+1), 53 + bias - b*j) of M_x */
        int sh = 53 + bias - b * (j + 1);
        uint64_t v;
        if (sh >= 64) v = 0;
        else if (sh >= 0) v = mx >> sh;
        else v = (-sh >= 64) ? 0 : (mx << (-sh));
        int dig = (int)(v & (uint64_t)mask);
        out[(int64_t)j * stride] = (int8_t)(neg ? -dig : dig);
    }
}

/* Split a rows x cols operand (element (r,c) at X[r*sr + c*sc]) into d planes
 * S[j][r][c] (int8), relative to E[r].  mech != 0 selects the word-level restatement. */
void oro_split(const double *X, int64_t rows, int64_t cols, int64_t sr, int64_t sc,
               const int16_t *E, int b, int d, int8_t *S, int mech) {
    const int64_t plane = rows * cols;
    #pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < rows; r++)
        for (int64_t c = 0; c < cols; c++) {
            double x = X[r * sr + c * sc];
            if (mech) oro_split_mech_elem(x, E[r], b, d, S + r * cols + c, plane);
            else oro_split_elem(x, E[r], b, d, S + r * cols + c, plane);
        }
}

/* Triangular slice products, mm_ozaki.cc:304-357: bins[i+j] += SA_i . SB_j^T for
 * i + j < d.  SA is [d][M][K], SB is [d][N][K] (both contraction-contiguous),
 * bins is [d][M][N] int64 (exact; the reference holds them in int32, the caller checks
 * the range).  Pinned against numpy int64 matmuls of the same slices. */
void oro_bins(const int8_t *SA, const int8_t *SB, int64_t M, int64_t N, int64_t K, int d,
              int64_t *bins) {
    memset(bins, 0, sizeof(int64_t) * (size_t)d * M * N);
    #pragma omp parallel for schedule(dynamic, 1)
    for (int64_t r = 0; r < M; r++) {
        for (int i = 0; i < d; i++) {
            const int8_t *a = SA + ((int64_t)i * M + r) * K;
            for (int j = 0; i + j < d; j++) {
                const int8_t *bp = SB + (int64_t)j * N * K;
                int64_t *out = bins + ((int64_t)(i + j) * M + r) * N;
                for (int64_t c = 0; c < N; c++) {
                    const int8_t *bc = bp + c * K;
                    int64_t s = 0;
                    for (int64_t k = 0; k < K; k++) s += (int64_t)a[k] * bc[k];
                    out[c] += s;
                }
            }
        }
    }
}

/* Bin-group sizes n0/n1 of the paired-int64 format, ozaki_base.cc:151-161 and
 * include/oz_aie.h:96-103 (log2K = count of trailing zeros of K). */
void oro_groups(int b, int d, int K, int *n0, int *n1) {
    int log2k = __builtin_ctz((unsigned)K);
    int alignable = 31 / b + 1;
    int numer = b * (d + 1) - 51 + log2k;
    int minpos = numer < 0 ? 1 : (numer / b + 1);
    int g0 = alignable > minpos ? alignable : minpos;
    if (g0 > d) g0 = d;
    *n0 = g0;
    *n1 = d - g0;
}

/* Mechanistic bin sum + exponent embed, ozaki_vectorized_sum_upshift
 * (ozaki_base.cc:143-340): out[(r*N + c)*2 + {0,1}] = {L, H}. */
void oro_pack_paired(const int64_t *bins, int64_t M, int64_t N, int K, int b, int d,
                     const int16_t *Er, const int16_t *Ec, int64_t *out) {
    int n0, n1;
    oro_groups(b, d, K, &n0, &n1);
    const int64_t plane = M * N;
    for (int64_t r = 0; r < M; r++)
        for (int64_t c = 0; c < N; c++) {
            const int64_t idx = r * N + c;
            uint64_t L = 0, H = 0;                                  /* acc64 lanes, wrap */
            for (int i = 0; i < n0; i++)                            /* :279-285 */
                L += (uint64_t)((int64_t)(int32_t)bins[(int64_t)(d - 1 - i) * plane + idx]) << (b * i);
            for (int i = 0; i < n1; i++)
                H += (uint64_t)((int64_t)(int32_t)bins[(int64_t)(d - 1 - n0 - i) * plane + idx]) << (b * i);
            uint32_t h32 = (uint32_t)(H >> 32);                     /* :293-317 */
            uint32_t emb = ((h32 << 13) >> 13) + ((h32 >> 31) << 31);
            uint32_t es = (uint32_t)(int32_t)(int16_t)(Er[r] + Ec[c]);
            emb += es << (32 - 13);
            uint64_t Hout = ((uint64_t)emb << 32) | (uint32_t)H;
            out[idx * 2] = (int64_t)L;
            out[idx * 2 + 1] = (int64_t)Hout;
        }
}

static inline int clz128(u128 v) {
    uint64_t h = (uint64_t)(v >> 64), l = (uint64_t)v;
    if (h) return __builtin_clzll(h);
    if (l) return 64 + __builtin_clzll(l);
    return 128;
}

/* Host conversion convert_paired_i64_to_double_matrix, include/oz_aie.h:92-150,
 * restated step by step (uint64 wraparound included). */
void oro_convert_paired(const int64_t *P, double *C, int64_t M, int64_t N, int K, int b, int d) {
    const i128 sign_mask = (i128)1 << 127;
    const int total_fraction_space = 2 * (b * d);
    const int fraction_omitted = b * (d - 1);
    int n0, n1;
    oro_groups(b, d, K, &n0, &n1);
    for (int64_t r = 0; r < M; r++)
        for (int64_t c = 0; c < N; c++) {
            int64_t low = P[r * (N * 2) + c * 2];
            int64_t high = P[r * (N * 2) + c * 2 + 1];
            int64_t high_rbits = (int64_t)((uint64_t)(high >> 63) << (64 - 13));
            int64_t scale_back = (int64_t)(((uint64_t)high & 0x7FF8000000000000ull) + (2ull << 51));
            scale_back = (int64_t)((uint64_t)scale_back - (2046ull << 51));
            scale_back = (int64_t)((uint64_t)scale_back << 1);
            high = (int64_t)(((uint64_t)high & 0x0007FFFFFFFFFFFFull) | (uint64_t)high_rbits);
            i128 quad = (i128)low;
            quad += (i128)((u128)(i128)high << (n0 * b));
            i128 quad_sign = quad & sign_mask;
            quad *= (quad_sign >> 126) | 1;
            u128 uq = (u128)quad;
            int lz = clz128(uq);
            int64_t exponent = (128 - lz) - (total_fraction_space - fraction_omitted) - 1;
            int shift = (128 - lz) - 53;
            uint64_t lu = shift > 0 ? (uint64_t)(uq >> shift) : (uint64_t)(uq << (-shift));
            if (lu > 0) {
                lu -= 0x0010000000000000ull;
                lu += (uint64_t)(quad_sign >> 64) + ((uint64_t)(exponent + 1023) << 52) + (uint64_t)scale_back;
            }
            C[r * N + c] = bitsd(lu);
        }
}

/* Exact T = sum_k bins[k] * 2^(b*(d-1-k)) as a 128-bit integer, written as (hi, lo). */
void oro_bins_to_T(const int64_t *bins, int64_t M, int64_t N, int b, int d, int64_t *T2) {
    const int64_t plane = M * N;
    for (int64_t idx = 0; idx < plane; idx++) {
        i128 t = 0;
        for (int k = 0; k < d; k++) t += (i128)bins[(int64_t)k * plane + idx] << (b * (d - 1 - k));
        T2[idx * 2] = (int64_t)(uint64_t)((u128)t >> 64);
        T2[idx * 2 + 1] = (int64_t)(uint64_t)t;
    }
}

/* Closed-form conversion of exact T with exponent sum es = E_r + E_c:
 * rounding 0 -> sign(T) * trunc53(|T|) * 2^(es - 2044 - b(d+1)), reproducing the
 * bit arithmetic of include/oz_aie.h:124-146 (no over/underflow handling);
 * rounding 1 -> the same value rounded to nearest-even instead of truncated
 * (ldexp scaling, so out-of-range exponents saturate / go subnormal). */
double oro_T_to_double(int64_t Thi, int64_t Tlo, int es, int b, int d, int rounding) {
    i128 t = (i128)(((u128)(uint64_t)Thi << 64) | (uint64_t)Tlo);
    uint64_t sgn = t < 0 ? (1ull << 63) : 0ull;
    u128 a = t < 0 ? (u128)(-t) : (u128)t;
    if (a == 0) return 0.0;
    int nb = 128 - clz128(a);
    if (rounding == 0) {
        int shift = nb - 53;
        uint64_t lu = shift > 0 ? (uint64_t)(a >> shift) : (uint64_t)(a << (-shift));
        int64_t exponent = nb - b * (d + 1) - 1;
        lu -= 0x0010000000000000ull;
        lu += sgn + ((uint64_t)(exponent + 1023) << 52) + ((uint64_t)(int64_t)(es - 2044) << 52);
        return bitsd(lu);
    } else {
        /* round |T| to 53 significant bits, nearest-even */
        uint64_t m;
        int e2;
        if (nb <= 53) { m = (uint64_t)a; e2 = 0; }
        else {
            int sh = nb - 53;
            u128 q = a >> sh;
            u128 rem = a - (q << sh);
            u128 half = (u128)1 << (sh - 1);
            if (rem > half || (rem == half && (q & 1))) q += 1;
            m = (uint64_t)q; e2 = sh;   /* q may be 2^53: still exact in a double */
        }
        double v = ldexp((double)m, e2 + es - 2044 - b * (d + 1));
        return sgn ? -v : v;
    }
}

/* End-to-end CPU Ozaki DGEMM (row-major A MxK with lda, B KxN with ldb, C MxN with ldc)
 * restating the reference pipeline (max-exp, split, triangular bins, exact T,
 * conversion).  kchunk <= 0 means no chunking: every int32 bin is checked to fit.
 * Returns 0, or 1 if a bin would have overflowed int32 (the caller must chunk).
 * Used as the checker for the GPU path and, timed on a row sample, as the CPU baseline. */
int oro_ozaki_dgemm(int64_t M, int64_t N, int64_t K, const double *A, int64_t lda,
                    const double *B, int64_t ldb, double *C, int64_t ldc, int b, int d,
                    int64_t kchunk, int rounding) {
    int16_t *Er = (int16_t *)malloc(sizeof(int16_t) * (size_t)(M > 0 ? M : 1));
    int16_t *Ec = (int16_t *)malloc(sizeof(int16_t) * (size_t)(N > 0 ? N : 1));
    int8_t *SA = (int8_t *)malloc((size_t)d * M * K + 1);
    int8_t *SB = (int8_t *)malloc((size_t)d * N * K + 1);
    i128 *T = (i128 *)calloc((size_t)(M * N > 0 ? M * N : 1), sizeof(i128));
    int overflow = 0;
    oro_max_exp(A, M, K, lda, 1, Er);
    oro_max_exp(B, N, K, 1, ldb, Ec);
    oro_split(A, M, K, lda, 1, Er, b, d, SA, 0);
    oro_split(B, N, K, 1, ldb, Ec, b, d, SB, 0);
    if (kchunk <= 0 || kchunk > K) kchunk = K;
    for (int64_t k0 = 0; k0 < K; k0 += kchunk) {
        int64_t kl = (K - k0) < kchunk ? (K - k0) : kchunk;
        #pragma omp parallel for schedule(dynamic, 1) reduction(|:overflow)
        for (int64_t r = 0; r < M; r++) {
            int64_t *acc = (int64_t *)malloc(sizeof(int64_t) * (size_t)d * (N > 0 ? N : 1));
            memset(acc, 0, sizeof(int64_t) * (size_t)d * N);
            for (int i = 0; i < d; i++) {
                const int8_t *a = SA + ((int64_t)i * M + r) * K + k0;
                for (int j = 0; i + j < d; j++) {
                    const int8_t *bp = SB + (int64_t)j * N * K + k0;
                    int64_t *out = acc + (int64_t)(i + j) * N;
                    for (int64_t c = 0; c < N; c++) {
                        const int8_t *bc = bp + c * K;
                        /* int32 partial sums stay exact: 65536 * 127^2 < 2^31 */
                        for (int64_t k1 = 0; k1 < kl; k1 += 65536) {
                            int64_t k2 = (kl - k1) < 65536 ? kl : k1 + 65536;
                            int32_t s = 0;
                            for (int64_t k = k1; k < k2; k++) s += (int32_t)a[k] * (int32_t)bc[k];
                            out[c] += s;
                        }
                    }
                }
            }
            for (int k = 0; k < d; k++)
                for (int64_t c = 0; c < N; c++) {
                    int64_t v = acc[(int64_t)k * N + c];
                    if (v > INT32_MAX || v < INT32_MIN) overflow |= 1;
                    T[r * N + c] += (i128)v << (b * (d - 1 - k));
                }
            free(acc);
        }
    }
    #pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < M; r++)
        for (int64_t c = 0; c < N; c++) {
            i128 t = T[r * N + c];
            C[r * ldc + c] = oro_T_to_double((int64_t)(uint64_t)((u128)t >> 64), (int64_t)(uint64_t)t,
                                             Er[r] + Ec[c], b, d, rounding);
        }
    free(Er); free(Ec); free(SA); free(SB); free(T);
    return overflow;
}

/* Naive single-thread DGEMM, common.h:234-254 (row-major, sequential k). */
void oro_naive_dgemm(int64_t M, int64_t N, int64_t K, const double *A, const double *B, double *C) {
    for (int64_t r = 0; r < M; r++)
        for (int64_t c = 0; c < N; c++) {
            double s = 0.0;
            for (int64_t k = 0; k < K; k++) s += A[r * K + k] * B[k * N + c];
            C[r * N + c] = s;
        }
}

int oro_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
