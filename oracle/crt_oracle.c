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
 *               197, 193, 191 (the first n, 2 <= n <= 16; pairwise coprime); M = prod p.
 *   bits        beta = floor((L - 2 - ceil(log2 K)) / 2), L = floor(log2 M), at most
 *               62: then K * 2^(2 beta) <= M / 4, so |C'| <= M / 4.
 *   scaling     the row (A) / column (B) maximum exponent field E as in the split
 *               scheme; q(x) = sign(x) * floor(M_x * 2^(beta - 53 - (E - e'))), |q| < 2^beta,
 *               e' = max(e, 1), i.e. q = sign(x) floor(|x| 2^(1022 + beta - E)) for every
 *               finite x (unlike the split scheme, subnormals are not off by a factor 2).
 *   residues    r_p(x) = q(x) mod p in [-(p-1)/2, (p-1)/2] for odd p, [-128, 127] for
 *               p = 256: int8.
 *   products    g_p = (sum_k r_p(a_ik) r_p(b_kj)) mod p in [0, p) (int32 sums are exact).
 *   rebuild     W_p = (M/p) * ((M/p)^-1 mod p) mod M; S = sum_p g_p W_p (mod 2^128);
 *               t = sum_p g_p * fl(W_p / M) in double; k = nearest integer to t;
 *               C' = S - k M (mod 2^128) read as a signed 128-bit integer.  Because
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

typedef __int128 i128;
typedef unsigned __int128 u128;

static const int crt_p[16] = {256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 211, 199, 197, 193, 191};

static inline uint64_t dbits(double x) { uint64_t u; memcpy(&u, &x, 8); return u; }

int oro_crt_modulus(int i) { return (i >= 0 && i < 16) ? crt_p[i] : 0; }

static u128 crt_M(int n) {
    u128 m = 1;
    for (int i = 0; i < n; i++) m *= (u128)crt_p[i];
    return m;
}

static int floor_log2_u128(u128 v) {
    int l = -1;
    while (v) { v >>= 1; l++; }
    return l;
}

/* bits per operand for n moduli and contraction length K, or -1 if none fits */
int oro_crt_beta(int n, int64_t K) {
    if (n < 2 || n > 16 || K < 1) return -1;
    int L = floor_log2_u128(crt_M(n));
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

/* W_p as (lo, hi) words, f_p = fl(W_p / M), M as (lo, hi). */
void oro_crt_consts(int n, uint64_t *W, double *f, uint64_t *Mw) {
    u128 M = crt_M(n);
    for (int i = 0; i < n; i++) {
        u128 Mi = M / (u128)crt_p[i];
        int c = inv_mod((int)(Mi % (u128)crt_p[i]), crt_p[i]);
        u128 w = Mi * (u128)c;   /* < M */
        W[2 * i] = (uint64_t)w;
        W[2 * i + 1] = (uint64_t)(w >> 64);
        f[i] = (double)w / (double)M;
    }
    Mw[0] = (uint64_t)M;
    Mw[1] = (uint64_t)(M >> 64);
}

/* Signed integer q(x) of one element for row/column exponent field emax. */
int64_t oro_crt_q(double x, int emax, int beta) {
    uint64_t u = dbits(x);
    int e = (int)((u >> 52) & 0x7FF);
    uint64_t mx = (u & 0xFFFFFFFFFFFFFull) | (e ? (1ull << 52) : 0ull);
    int sh = 53 + (emax - (e ? e : 1)) - beta;   /* subnormals have exponent 1 - 1023 */
    uint64_t Q = sh >= 64 ? 0 : (sh >= 0 ? mx >> sh : mx << (-sh));
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

/* Exact C' from the n residues g[i * stride] (as hi/lo words of a signed 128-bit value). */
void oro_crt_rebuild(const uint8_t *g, int64_t stride, int n, int64_t *Chi, int64_t *Clo) {
    uint64_t W[32], Mw[2];
    double f[16];
    oro_crt_consts(n, W, f, Mw);
    u128 S = 0;
    double t = 0.0;
    for (int i = 0; i < n; i++) {
        u128 w = ((u128)W[2 * i + 1] << 64) | W[2 * i];
        S += (u128)g[i * stride] * w;
        t += (double)g[i * stride] * f[i];
    }
    int64_t k = (int64_t)nearbyint(t);
    u128 M = ((u128)Mw[1] << 64) | Mw[0];
    S -= (u128)k * M;
    *Chi = (int64_t)(uint64_t)(S >> 64);
    *Clo = (int64_t)(uint64_t)S;
}

/* C' * 2^e2 to double: truncation (rounding 0) or nearest-even (1) to 53 bits. */
double oro_crt_to_double(int64_t Chi, int64_t Clo, int e2, int rounding) {
    i128 t = (i128)(((u128)(uint64_t)Chi << 64) | (uint64_t)Clo);
    if (t == 0) return 0.0;
    int neg = t < 0;
    u128 a = neg ? (u128)(-t) : (u128)t;
    int nb = floor_log2_u128(a) + 1;
    uint64_t m;
    int sh = 0;
    if (nb <= 53) m = (uint64_t)a;
    else {
        sh = nb - 53;
        u128 q = a >> sh;
        if (rounding) {
            u128 rem = a - (q << sh);
            u128 half = (u128)1 << (sh - 1);
            if (rem > half || (rem == half && (q & 1))) q += 1;
        }
        m = (uint64_t)q;
    }
    double v = ldexp((double)m, sh + e2);
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
    #pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < M; r++)
        for (int64_t c = 0; c < N; c++) {
            int64_t hi, lo;
            oro_crt_rebuild(G + r * N + c, M * N, n, &hi, &lo);
            C[r * ldc + c] = oro_crt_to_double(hi, lo, Er[r] + Ec[c] - 2044 - 2 * beta, rounding);
        }
    free(Er); free(Ec); free(SA); free(SB); free(G);
    return 0;
}
