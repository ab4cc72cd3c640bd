/*
 * mpfr_oracle.c -- TEST INFRASTRUCTURE ONLY (see ozaki_oracle.c for the usage rule).
 *
 * The reference's accuracy oracle: each product and each running sum rounded to `prec`
 * bits (106 by default) with MPFR, then rounded to double -- mul_acc_MPFR and
 * matmul_MPFR in src/programming_examples/basic/matrix_multiplication/common.h:164-232
 * (default precision common.h:81-82).  Pinned in the tests against exact rational dot
 * products (Python fractions) on small inputs.
 */
#include <stdint.h>
#include <mpfr.h>

double oro_mpfr_dot(const double *a, int64_t sa, const double *b, int64_t sb, int64_t K, int prec) {
    mpfr_t ea, eb, prod, sum;
    mpfr_inits2(prec, ea, eb, prod, sum, (mpfr_ptr)0);
    mpfr_set_d(sum, 0.0, MPFR_RNDN);
    for (int64_t k = 0; k < K; k++) {
        mpfr_set_d(ea, a[k * sa], MPFR_RNDN);
        mpfr_set_d(eb, b[k * sb], MPFR_RNDN);
        mpfr_mul(prod, ea, eb, MPFR_RNDN);
        mpfr_add(sum, sum, prod, MPFR_RNDN);
    }
    double r = mpfr_get_d(sum, MPFR_RNDN);
    mpfr_clears(ea, eb, prod, sum, (mpfr_ptr)0);
    return r;
}

/* C[i] = mpfr dot of row rows[i] of A (M x K, lda) and column cols[i] of B (K x N, ldb). */
void oro_mpfr_sample(const double *A, int64_t lda, const double *B, int64_t ldb, int64_t K,
                     const int64_t *rows, const int64_t *cols, int64_t n, int prec, double *out) {
    #pragma omp parallel for schedule(dynamic, 4)
    for (int64_t i = 0; i < n; i++)
        out[i] = oro_mpfr_dot(A + rows[i] * lda, 1, B + cols[i], ldb, K, prec);
    mpfr_free_cache();
}
