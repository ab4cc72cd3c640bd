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
 *   bits        beta = floor((L - 2 - ceil(log2 K)) / 2), L = floor(log2 M), at most
 *               62: then K * 2^(2 beta) <= M / 4, so |C'| <= M / 4.
 *   scaling     the row (A) / column (B) maximum exponent field E as in the split
 *               scheme; q(x) = sign(x) * round(M_x * 2^(beta - 53 - (E - e'))), |q| <= 2^beta,
 *               e' = max(e, 1), i.e. q = sign(x) round(|x| 2^(1022 + beta - E)) for every
 *               finite x (unlike the split scheme, subnormals are not off by a factor 2);
 *               round = to nearest, ties away from zero (half the error of truncation and
 *               no bias; |q| <= 2^beta keeps K * 2^(2 beta) as the bound of |C'|).
 *   residues    r_p(x) = q(x) mod p in [-(p-1)/2, (p-1)/2] for odd p, [-128, 127] for
 *               p = 256: int8.
 *   products    g_p = (sum_k r_p(a_ik) r_p(b_kj)) mod p in [0, p) (int32 sums are exact).
 *   rebuild     W_p = (M/p) * ((M/p)^-1 mod p) mod M; S = sum_p g_p W_p (mod 2^192);
 *               t = sum_p g_p * fl(W_p / M) in double; k = nearest integer to t;
 *               C' = S - k M (mod 2^192) read as a signed 192-bit integer.  Because
 *               |C'| <= M/4 and the error of t is below 2^-30, k is exact and C' is the
 *               exact integer sum_k q(a_ik) q(b_kj).
 *   output      C = C' * 2^(E_r + E_c - 2044 - 2 beta), truncated toward zero or rounded
 *               to nearest-even to 53 bits.
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

/* Signed integer q(x) of one element for row/column exponent field emax. */
int64_t oro_crt_q(double x, int emax, int beta) {
    uint64_t u = dbits(x);
    int e = (int)((u >> 52) & 0x7FF);
    uint64_t mx = (u & 0xFFFFFFFFFFFFFull) | (e ? (1ull << 52) : 0ull);
    int sh = 53 + (emax - (e ? e : 1)) - beta;   /* subnormals have exponent 1 - 1023 */
    /* nearest, ties away from zero; for sh >= 54, mx / 2^sh < 2^(53 - sh) <= 1/2 strictly
     * (mx < 2^53), so those round to 0 */
    uint64_t Q = sh >= 54 ? 0 : (sh >= 1 ? (mx + (1ull << (sh - 1))) >> sh : mx << (-sh));
    return (u >> 63) ? -(int64_t)Q : (int64_t)Q;
}

static inline int sym_mod(int64_t q, int p) {
    int64_t r = q % p;
    if (r < 0) r += p;
    int half = p == 256 ? 127 : (p - 1) / 2;
    return (int)(r > half ? r - p : r);
}

/* Residues of a rows x cols operand (element (r, c) at X[r*sr + c*sc], exponent field E[r]
 * per row) into S[i][r][c] (int8), i < n. */
void oro_crt_split(const double *X, int64_t rows, int64_t cols, int64_t sr, int64_t sc, const int16_t *E,
                   int beta, int n, int8_t *S) {
    #pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < rows; r++)
        for (int64_t c = 0; c < cols; c++) {
            int64_t q = oro_crt_q(X[r * sr + c * sc], E[r], beta);
            for (int i = 0; i < n; i++) S[((int64_t)i * rows + r) * cols + c] = (int8_t)sym_mod(q, crt_p[i]);
        }
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

/* End-to-end CRT-mode DGEMM, row-major A (M x K, lda), B (K x N, ldb), C (M x N, ldc).
 * Returns 0, or -1 if no bit count fits (n, K). */
int oro_crt_dgemm(int64_t M, int64_t N, int64_t K, const double *A, int64_t lda, const double *B, int64_t ldb,
                  double *C, int64_t ldc, int n, int rounding) {
    int beta = oro_crt_beta(n, K);
    if (beta < 0) return -1;
    int16_t *Er = (int16_t *)malloc(sizeof(int16_t) * (size_t)(M > 0 ? M : 1));
    int16_t *Ec = (int16_t *)malloc(sizeof(int16_t) * (size_t)(N > 0 ? N : 1));
    int8_t *SA = (int8_t *)malloc((size_t)n * M * K + 1);
    int8_t *SB = (int8_t *)malloc((size_t)n * N * K + 1);
    uint8_t *G = (uint8_t *)malloc((size_t)n * M * N + 1);
    extern void oro_max_exp(const double *, int64_t, int64_t, int64_t, int64_t, int16_t *);
    oro_max_exp(A, M, K, lda, 1, Er);
    oro_max_exp(B, N, K, 1, ldb, Ec);
    oro_crt_split(A, M, K, lda, 1, Er, beta, n, SA);
    oro_crt_split(B, N, K, 1, ldb, Ec, beta, n, SB);
    oro_crt_products(SA, SB, M, N, K, n, G);
    u192 Wt[CRT_MAXMOD], Mt;
    double ft[CRT_MAXMOD];
    crt_tables(n, Wt, ft, &Mt);
    #pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < M; r++)
        for (int64_t c = 0; c < N; c++) {
            int64_t C3[3];
            rebuild_with(Wt, ft, Mt, G + r * N + c, M * N, n, C3);
            C[r * ldc + c] = oro_crt_to_double(C3, Er[r] + Ec[c] - 2044 - 2 * beta, rounding);
        }
    free(Er); free(Ec); free(SA); free(SB); free(G);
    return 0;
}
