/*
 * crt_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the library's CRT mode.  This mode is NOT in the reference (the
 * reference implements only the split/triangle scheme restated in ozaki_oracle.c); it
 * follows the published Ozaki scheme II (Ozaki, Uchino, Imamura, "Ozaki Scheme II: A
 * GEMM-oriented emulation of floating-point matrix multiplication using an integer
 * modular technique", 2024/2025): scale each row of A and column of B to integers,
 * multiply the residues modulo several pairwise-coprime moduli with exact int8 GEMMs,
 * and rebuild the exact integer product with the Chinese remainder theorem.
 *
 * The arithmetic, defined here and implemented by the GPU kernels:
 *   moduli      p = 256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 211, 199,
 *               197, 193, 191, 181, 179, 173, 167 (the first n, 2 <= n <= 20; pairwise
 *               coprime); M = prod p (below 2^156).
 *   shift       every row of A and column of B has an integer shift sigma, and each
 *               element becomes q(x) = sign(x) round(M_x 2^(e' - sigma)), e' = max(e, 1),
 *               M_x its 53-bit significand, i.e. q = sign(x) round(|x| 2^(1075 - sigma))
 *               for every finite x (unlike the split scheme, subnormals are not off by a
 *               factor 2); round = to nearest, ties away from zero.  Two rules for sigma:
 *   max rule    (explicit bits beta) sigma = 53 + E - beta with E the row's maximum
 *               exponent field as in the split scheme: |q| <= 2^beta, and with
 *               beta <= floor((L - 2 - ceil(log2 K)) / 2), L = floor(log2 M), at most 62,
 *               K 2^(2 beta) <= M / 4 bounds |C'| by M / 4.
 *   norm rule   (bits 0, the default where M 2^20 > K (2^20 + 2^40); below that, i.e. for
 *               a few moduli, bits 0 means the max rule with the largest beta) from an
 *               exact integer bound on the row's Euclidean
 *               norm.  Chunks of 16 consecutive contraction indices (from index 0): for a
 *               chunk with a nonzero element, Ec = max e' over them and
 *               S_c = sum u^2, u = ceil(M_x / 2^(37 + Ec - e')) = ceil(|x| / 2^(Ec - 1038))
 *               <= 2^16; per row E^ = max Ec, S = sum_c ceil(S_c / 4^(E^ - Ec)), so that
 *               ||x||^2 <= S 4^(E^ - 1038).  m = the largest integer <= 46 with
 *               S (2^22 + 4) 4^m <= M 2^20 - K (2^20 + 2^40), and sigma = 37 + E^ - m
 *               (sigma = 0 for a zero row).  Then ||q|| <= 2^m sqrt(S) + sqrt(K) / 2 and,
 *               by 2ab <= a^2 2^-20 + b^2 2^20, ||q||^2 <= M / 4 for every row and column,
 *               so |C'| <= ||q_a|| ||q_b|| <= M / 4 (Cauchy-Schwarz); every |q| <= 2^62.
 *               It gives the operands about log2(max |x| sqrt(K) / ||x||) more bits than
 *               the max rule: 4 at K = 8192 for (U - 1/2) exp(N(0, 1)) entries.
 *   residues    r_p(x) = q(x) mod p in [-(p-1)/2, (p-1)/2] for odd p, [-128, 127] for
 *               p = 256: int8.
 *   products    g_p = (sum_k r_p(a_ik) r_p(b_kj)) mod p in [0, p) (int32 sums are exact).
 *   rebuild     W_p = (M/p) * ((M/p)^-1 mod p) mod M; S = sum_p g_p W_p (mod 2^192);
 *               t = sum_p g_p * fl(W_p / M) in double; k = nearest integer to t;
 *               C' = S - k M (mod 2^192) read as a signed 192-bit integer.  Because
 *               |C'| <= M/4 and the error of t is below 2^-30, k is exact and C' is the
 *               exact integer sum_k q(a_ik) q(b_kj).
 *   output      C = C' * 2^(sigma_r + sigma_c - 2150) (max rule: 2^(E_r + E_c - 2044 -
 *               2 beta)), truncated toward zero or rounded to nearest-even to 53 bits.
 * Pinned in tests/test_oracle_crt.py against Python integer arithmetic (residues, CRT
 * identity, exact products and the final scaling) and the exact rational product.
 */
#include <stdint.h>
#include <string.h>
#include <stdlib.h>
#include <math.h>

typedef unsigned __int128 u128;

#define CRT_MAXMOD 20
static const int crt_p[CRT_MAXMOD] = {256, 255, 253, 251, 247, 241, 239, 233, 229, 227,
                                      223, 211, 199, 197, 193, 191, 181, 179, 173, 167};

static inline uint64_t dbits(double x) { uint64_t u; memcpy(&u, &x, 8); return u; }

/* 192-bit unsigned integers, three little-endian 64-bit words, arithmetic mod 2^192 */
typedef struct { uint64_t w[3]; } u192;

static u192 u192_small(uint64_t v) { u192 r = {{v, 0, 0}}; return r; }

static u192 u192_add(u192 a, u192 b) {
    u192 r;
    u128 c = 0;
    for (int i = 0; i < 3; i++) { c += (u128)a.w[i] + b.w[i]; r.w[i] = (uint64_t)c; c >>= 64; }
    return r;
}

static u192 u192_neg(u192 a) {
    u192 r;
    u128 c = 1;
    for (int i = 0; i < 3; i++) { c += (u128)(~a.w[i]); r.w[i] = (uint64_t)c; c >>= 64; }
    return r;
}

static u192 u192_mul_small(u192 a, uint64_t m) {   /* m < 2^32 */
    u192 r;
    u128 c = 0;
    for (int i = 0; i < 3; i++) { c += (u128)a.w[i] * m; r.w[i] = (uint64_t)c; c >>= 64; }
    return r;
}

static uint64_t u192_divmod_small(u192 *a, uint64_t d) {   /* a /= d, returns a mod d */
    u128 r = 0;
    for (int i = 2; i >= 0; i--) {
        r = (r << 64) | a->w[i];
        a->w[i] = (uint64_t)(r / d);
        r %= d;
    }
    return (uint64_t)r;
}

static int u192_bitlen(u192 a) {
    for (int i = 2; i >= 0; i--)
        if (a.w[i]) return 64 * i + 64 - __builtin_clzll(a.w[i]);
    return 0;
}

static double u192_to_double(u192 a) {   /* within a few ulps; used only for W/M */
    return ((double)a.w[2] * 18446744073709551616.0 + (double)a.w[1]) * 18446744073709551616.0 + (double)a.w[0];
}

int oro_crt_modulus(int i) { return (i >= 0 && i < CRT_MAXMOD) ? crt_p[i] : 0; }

static u192 crt_M(int n) {
    u192 m = u192_small(1);
    for (int i = 0; i < n; i++) m = u192_mul_small(m, (uint64_t)crt_p[i]);
    return m;
}

/* bits per operand for n moduli and contraction length K, or -1 if none fits */
int oro_crt_beta(int n, int64_t K) {
    if (n < 2 || n > CRT_MAXMOD || K < 1) return -1;
    int L = u192_bitlen(crt_M(n)) - 1;
    int lk = 0;
    while (((int64_t)1 << lk) < K) lk++;
    int beta = (L - 2 - lk) / 2;
    if (beta > 62) beta = 62;
    return beta >= 1 ? beta : -1;
}

static int inv_mod(int a, int p) {   /* a^-1 mod p, gcd(a, p) = 1 */
    int t = 0, nt = 1, r = p, nr = a % p;
    while (nr) {
        int q = r / nr, tmp;
        tmp = t - q * nt; t = nt; nt = tmp;
        tmp = r - q * nr; r = nr; nr = tmp;
    }
    return t < 0 ? t + p : t;
}

static u192 crt_W(int n, int i) {
    u192 Mi = crt_M(n);
    u192_divmod_small(&Mi, (uint64_t)crt_p[i]);
    u192 tmp = Mi;
    const int c = inv_mod((int)u192_divmod_small(&tmp, (uint64_t)crt_p[i]), crt_p[i]);
    return u192_mul_small(Mi, (uint64_t)c);   /* < M */
}

/* W_p as three words each (low first), f_p = fl(W_p / M), M as three words. */
void oro_crt_consts(int n, uint64_t *W, double *f, uint64_t *Mw);

/* Signed integer q(x) of one element for the row/column shift sigma (sigma - e' >= -9). */
int64_t oro_crt_q_sigma(double x, int sigma) {
    uint64_t u = dbits(x);
    int e = (int)((u >> 52) & 0x7FF);
    uint64_t mx = (u & 0xFFFFFFFFFFFFFull) | (e ? (1ull << 52) : 0ull);
    int sh = sigma - (e ? e : 1);   /* subnormals have exponent 1 - 1023 */
    /* nearest, ties away from zero; for sh >= 54, mx / 2^sh < 2^(53 - sh) <= 1/2 strictly
     * (mx < 2^53), so those round to 0 */
    uint64_t Q = sh >= 54 ? 0 : (sh >= 1 ? (mx + (1ull << (sh - 1))) >> sh : mx << (-sh));
    return (u >> 63) ? -(int64_t)Q : (int64_t)Q;
}

/* The max rule: q(x) for row/column exponent field emax and beta bits. */
int64_t oro_crt_q(double x, int emax, int beta) { return oro_crt_q_sigma(x, 53 + emax - beta); }

/* ---- the norm rule */
#define CRT_CHUNK 16
#define CRT_MCAP 46

static int fieldp(double x) {   /* e' = max(e, 1) */
    int e = (int)((dbits(x) >> 52) & 0x7FF);
    return e ? e : 1;
}

/* Per row: E^ and the integer norm bound S (||x||^2 <= S 4^(E^ - 1038)); a zero row gives
 * E^ = 0, S = 0. */
void oro_crt_normbound(const double *X, int64_t rows, int64_t cols, int64_t sr, int64_t sc, int32_t *Ehat,
                       uint64_t *S) {
    const int64_t nch = (cols + CRT_CHUNK - 1) / CRT_CHUNK;
    #pragma omp parallel
    {
        int *Ec = (int *)malloc(sizeof(int) * (size_t)(nch > 0 ? nch : 1));
        uint64_t *Sc = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(nch > 0 ? nch : 1));
        #pragma omp for schedule(static)
        for (int64_t r = 0; r < rows; r++) {
            int Eh = 0;
            for (int64_t c = 0; c < nch; c++) {
                int em = 0;
                for (int64_t k = c * CRT_CHUNK; k < cols && k < (c + 1) * CRT_CHUNK; k++) {
                    const double x = X[r * sr + k * sc];
                    if (x != 0.0 && fieldp(x) > em) em = fieldp(x);
                }
                uint64_t s = 0;
                for (int64_t k = c * CRT_CHUNK; em && k < cols && k < (c + 1) * CRT_CHUNK; k++) {
                    const double x = X[r * sr + k * sc];
                    if (x == 0.0) continue;
                    const uint64_t u = dbits(x);
                    const int e = (int)((u >> 52) & 0x7FF);
                    const uint64_t mx = (u & 0xFFFFFFFFFFFFFull) | (e ? (1ull << 52) : 0ull);
                    const int t = 37 + em - fieldp(x);   /* >= 37 */
                    const uint64_t v = t >= 64 ? 1 : (mx + ((1ull << t) - 1)) >> t;   /* ceil, <= 2^16 */
                    s += v * v;
                }
                Ec[c] = em;
                Sc[c] = s;
                if (em > Eh) Eh = em;
            }
            uint64_t tot = 0;
            for (int64_t c = 0; c < nch; c++) {
                if (!Sc[c]) continue;
                const int d2 = 2 * (Eh - Ec[c]);   /* S_c < 2^37: the ceiling is 1 from 2^37 on */
                tot += d2 >= 40 ? 1 : (Sc[c] + ((1ull << d2) - 1)) >> d2;
            }
            Ehat[r] = Eh;
            S[r] = tot;
        }
        free(Ec);
        free(Sc);
    }
}

static u192 u192_shl(u192 a, int s) {   /* s < 192 */
    u192 r = {{0, 0, 0}};
    const int w = s / 64, b = s % 64;
    for (int i = 2; i >= 0; i--) {
        if (i - w < 0) continue;
        uint64_t v = a.w[i - w] << b;
        if (b && i - w - 1 >= 0) v |= a.w[i - w - 1] >> (64 - b);
        r.w[i] = v;
    }
    return r;
}

static int u192_le(u192 a, u192 b) {
    for (int i = 2; i >= 0; i--)
        if (a.w[i] != b.w[i]) return a.w[i] < b.w[i];
    return 1;
}

/* R = M 2^20 - K (2^20 + 2^40); returns 0 if it is not positive. */
static int crt_norm_R(int n, int64_t K, u192 *R) {
    const u192 MR = u192_shl(crt_M(n), 20);
    const u192 t = u192_mul_small(u192_small(((uint64_t)1 << 20) + ((uint64_t)1 << 40)), (uint64_t)K);
    if (u192_le(MR, t)) return 0;
    *R = u192_add(MR, u192_neg(t));
    return 1;
}

/* S (2^22 + 4) 4^m <= R */
static int norm_fits(u128 Sp, u192 R, int m) {
    if (m >= 0) {
        u192 a = {{(uint64_t)Sp, (uint64_t)(Sp >> 64), 0}};
        return u192_le(u192_shl(a, 2 * m), R);
    }
    const int nb = u192_bitlen(R);
    if (nb + (-2 * m) > 72) return 1;   /* R 4^-m >= 2^72 > Sp */
    const u128 r = ((u128)R.w[1] << 64 | R.w[0]) << (-2 * m);
    return Sp <= r;
}

/* m of the norm rule for bound S (> 0), or -1000 if R <= 0. */
int oro_crt_norm_m(uint64_t S, int n, int64_t K) {
    u192 R;
    if (!crt_norm_R(n, K, &R)) return -1000;
    const u128 Sp = (u128)S * ((1u << 22) + 4);
    int bs = 0;
    for (u128 v = Sp; v; v >>= 1) bs++;
    int m = (u192_bitlen(R) - bs) / 2 + 1;
    if (m > CRT_MCAP) m = CRT_MCAP;
    while (!norm_fits(Sp, R, m)) m--;
    return m;
}

/* Norm-rule shifts of a rows x cols operand (contraction along cols); -1 if none fits. */
int oro_crt_sigma_norm(const double *X, int64_t rows, int64_t cols, int64_t sr, int64_t sc, int n, int64_t K,
                       int32_t *sigma) {
    u192 R;
    if (n < 2 || n > CRT_MAXMOD || !crt_norm_R(n, K, &R)) return -1;
    int32_t *Eh = (int32_t *)malloc(sizeof(int32_t) * (size_t)(rows > 0 ? rows : 1));
    uint64_t *S = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(rows > 0 ? rows : 1));
    oro_crt_normbound(X, rows, cols, sr, sc, Eh, S);
    for (int64_t r = 0; r < rows; r++)
        sigma[r] = S[r] ? 37 + Eh[r] - oro_crt_norm_m(S[r], n, K) : 0;
    free(Eh);
    free(S);
    return 0;
}

/* Max-rule shifts: sigma = 53 + E - beta. */
void oro_crt_sigma_max(const double *X, int64_t rows, int64_t cols, int64_t sr, int64_t sc, int beta,
                       int32_t *sigma) {
    extern void oro_max_exp(const double *, int64_t, int64_t, int64_t, int64_t, int16_t *);
    int16_t *E = (int16_t *)malloc(sizeof(int16_t) * (size_t)(rows > 0 ? rows : 1));
    oro_max_exp(X, rows, cols, sr, sc, E);
    for (int64_t r = 0; r < rows; r++) sigma[r] = 53 + E[r] - beta;
    free(E);
}

static inline int sym_mod(int64_t q, int p) {
    int64_t r = q % p;
    if (r < 0) r += p;
    int half = p == 256 ? 127 : (p - 1) / 2;
    return (int)(r > half ? r - p : r);
}

/* Residues of a rows x cols operand (element (r, c) at X[r*sr + c*sc], shift sigma[r] per
 * row) into S[i][r][c] (int8), i < n. */
void oro_crt_split_sigma(const double *X, int64_t rows, int64_t cols, int64_t sr, int64_t sc,
                         const int32_t *sigma, int n, int8_t *S) {
    #pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < rows; r++)
        for (int64_t c = 0; c < cols; c++) {
            int64_t q = oro_crt_q_sigma(X[r * sr + c * sc], sigma[r]);
            for (int i = 0; i < n; i++) S[((int64_t)i * rows + r) * cols + c] = (int8_t)sym_mod(q, crt_p[i]);
        }
}

/* The same with the max rule: exponent field E[r] per row, beta bits. */
void oro_crt_split(const double *X, int64_t rows, int64_t cols, int64_t sr, int64_t sc, const int16_t *E,
                   int beta, int n, int8_t *S) {
    int32_t *sg = (int32_t *)malloc(sizeof(int32_t) * (size_t)(rows > 0 ? rows : 1));
    for (int64_t r = 0; r < rows; r++) sg[r] = 53 + E[r] - beta;
    oro_crt_split_sigma(X, rows, cols, sr, sc, sg, n, S);
    free(sg);
}

/* g_p = (sum_k SA_p[r][k] SB_p[c][k]) mod p in [0, p): G[i][r][c] (uint8). */
void oro_crt_products(const int8_t *SA, const int8_t *SB, int64_t M, int64_t N, int64_t K, int n, uint8_t *G) {
    #pragma omp parallel for schedule(dynamic, 1)
    for (int64_t r = 0; r < M; r++)
        for (int i = 0; i < n; i++) {
            const int8_t *a = SA + ((int64_t)i * M + r) * K;
            for (int64_t c = 0; c < N; c++) {
                const int8_t *b = SB + ((int64_t)i * N + c) * K;
                int64_t s = 0;
                for (int64_t k = 0; k < K; k++) s += (int32_t)a[k] * (int32_t)b[k];
                int64_t g = s % crt_p[i];
                if (g < 0) g += crt_p[i];
                G[((int64_t)i * M + r) * N + c] = (uint8_t)g;
            }
        }
}

static void rebuild_with(const u192 *W, const double *f, u192 M, const uint8_t *g, int64_t stride, int n,
                         int64_t *C3) {
    u192 S = u192_small(0);
    double t = 0.0;
    for (int i = 0; i < n; i++) {
        S = u192_add(S, u192_mul_small(W[i], g[i * stride]));
        t += (double)g[i * stride] * f[i];
    }
    const int64_t k = (int64_t)nearbyint(t);   /* 0 <= k < n * 256 */
    S = u192_add(S, u192_neg(u192_mul_small(M, (uint64_t)k)));
    for (int i = 0; i < 3; i++) C3[i] = (int64_t)S.w[i];
}

static void crt_tables(int n, u192 *W, double *f, u192 *M) {
    *M = crt_M(n);
    for (int i = 0; i < n; i++) {
        W[i] = crt_W(n, i);
        f[i] = u192_to_double(W[i]) / u192_to_double(*M);
    }
}

void oro_crt_consts(int n, uint64_t *W, double *f, uint64_t *Mw) {
    u192 Wt[CRT_MAXMOD], M;
    crt_tables(n, Wt, f, &M);
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) W[3 * i + k] = Wt[i].w[k];
    for (int k = 0; k < 3; k++) Mw[k] = M.w[k];
}

/* Exact C' from the n residues g[i * stride], as three words (low first) of a signed 192-bit
 * integer. */
void oro_crt_rebuild(const uint8_t *g, int64_t stride, int n, int64_t *C3) {
    u192 W[CRT_MAXMOD], M;
    double f[CRT_MAXMOD];
    crt_tables(n, W, f, &M);
    rebuild_with(W, f, M, g, stride, n, C3);
}

/* C' * 2^e2 to double (C' signed 192-bit, three words low first): truncation (rounding 0)
 * or nearest-even (1) to 53 bits. */
double oro_crt_to_double(const int64_t *C3, int e2, int rounding) {
    u192 a = {{(uint64_t)C3[0], (uint64_t)C3[1], (uint64_t)C3[2]}};
    const int neg = (int64_t)a.w[2] < 0;
    if (neg) a = u192_neg(a);
    const int nb = u192_bitlen(a);
    if (nb == 0) return 0.0;
    uint64_t m;
    int sh = 0;
    if (nb <= 53) {
        m = a.w[0];
    } else {
        sh = nb - 53;
        /* q = a >> sh (fits 53 bits), rem = a mod 2^sh compared with 2^(sh-1) */
        u192 q = a;
        for (int s = sh; s > 0;) {
            const int step = s > 63 ? 63 : s;
            q.w[0] = (q.w[0] >> step) | (q.w[1] << (64 - step));
            q.w[1] = (q.w[1] >> step) | (q.w[2] << (64 - step));
            q.w[2] >>= step;
            s -= step;
        }
        m = q.w[0];
        if (rounding) {
            /* bit sh-1 of a is the half bit; the bits below it are the sticky part */
            const int hb = sh - 1;
            const int half = (int)((a.w[hb / 64] >> (hb % 64)) & 1);
            int sticky = 0;
            for (int b = 0; b < hb; b++)
                if ((a.w[b / 64] >> (b % 64)) & 1) { sticky = 1; break; }
            if (half && (sticky || (m & 1))) m += 1;
        }
    }
    const double v = ldexp((double)m, sh + e2);
    return neg ? -v : v;
}

/* End-to-end CRT-mode DGEMM, row-major A (M x K, lda), B (K x N, ldb), C (M x N, ldc);
 * bits 0: the norm rule (or the max rule with the largest beta where the norm rule's bound
 * has no room), else the max rule with beta = bits.  Returns 0, or -1 if the
 * moduli do not fit (n, K, bits). */
int oro_crt_dgemm(int64_t M, int64_t N, int64_t K, const double *A, int64_t lda, const double *B, int64_t ldb,
                  double *C, int64_t ldc, int n, int bits, int rounding) {
    const int bmax = oro_crt_beta(n, K);
    if (bmax < 0 || bits < 0 || bits > bmax) return -1;
    u192 R;
    if (bits == 0 && !crt_norm_R(n, K, &R)) bits = bmax;   /* few moduli: the max rule */
    int32_t *Er = (int32_t *)malloc(sizeof(int32_t) * (size_t)(M > 0 ? M : 1));
    int32_t *Ec = (int32_t *)malloc(sizeof(int32_t) * (size_t)(N > 0 ? N : 1));
    int8_t *SA = (int8_t *)malloc((size_t)n * M * K + 1);
    int8_t *SB = (int8_t *)malloc((size_t)n * N * K + 1);
    uint8_t *G = (uint8_t *)malloc((size_t)n * M * N + 1);
    if (bits == 0) {
        if (oro_crt_sigma_norm(A, M, K, lda, 1, n, K, Er) || oro_crt_sigma_norm(B, N, K, 1, ldb, n, K, Ec)) {
            free(Er); free(Ec); free(SA); free(SB); free(G);
            return -1;
        }
    } else {
        oro_crt_sigma_max(A, M, K, lda, 1, bits, Er);
        oro_crt_sigma_max(B, N, K, 1, ldb, bits, Ec);
    }
    oro_crt_split_sigma(A, M, K, lda, 1, Er, n, SA);
    oro_crt_split_sigma(B, N, K, 1, ldb, Ec, n, SB);
    oro_crt_products(SA, SB, M, N, K, n, G);
    u192 Wt[CRT_MAXMOD], Mt;
    double ft[CRT_MAXMOD];
    crt_tables(n, Wt, ft, &Mt);
    #pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < M; r++)
        for (int64_t c = 0; c < N; c++) {
            int64_t C3[3];
            rebuild_with(Wt, ft, Mt, G + r * N + c, M * N, n, C3);
            C[r * ldc + c] = oro_crt_to_double(C3, Er[r] + Ec[c] - 2150, rounding);
        }
    free(Er); free(Ec); free(SA); free(SB); free(G);
    return 0;
}
