// oz_verify.cpp -- host-side accuracy meter: correctly rounded exact dot products.
//
// The reference harness scores every result against MPFR at 106 bits
// (src/programming_examples/basic/matrix_multiplication/common.h:164-232, 556-669).
// This meter computes the exact value of sum_k a_k b_k instead and rounds it once:
// each product is split error-free into p + e with p = fl(a b), e = fma(a, b, -p)
// (exact unless the product underflows), and the 2K terms are summed exactly with
// Shewchuk's non-overlapping partials (J. R. Shewchuk, "Adaptive Precision
// Floating-Point Arithmetic and Fast Robust Geometric Predicates", 1997), followed by
// the correctly rounded reduction of the partials.  It measures errors; it is not a
// DGEMM path.
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/ozaie.h"

namespace {

// Exact accumulation of one term into the partials list (ascending magnitude,
// non-overlapping).  Returns the new number of partials.
inline size_t grow(std::vector<double> &p, size_t n, double x) {
    size_t i = 0;
    for (size_t j = 0; j < n; j++) {
        double y = p[j];
        if (std::fabs(x) < std::fabs(y)) std::swap(x, y);
        const double hi = x + y;
        const double lo = y - (hi - x);
        if (lo != 0.0) p[i++] = lo;
        x = hi;
    }
    if (i >= p.size()) p.resize(i + 8);
    p[i++] = x;
    return i;
}

// Correctly rounded value of the exact sum of the partials.
inline double round_partials(const std::vector<double> &p, size_t n) {
    if (n == 0) return 0.0;
    double hi = p[--n], lo = 0.0;
    while (n > 0) {
        const double x = hi, y = p[--n];
        hi = x + y;
        const double yr = hi - x;
        lo = y - yr;
        if (lo != 0.0) break;
    }
    // half-way case: the remaining partials decide the direction of the rounding
    if (n > 0 && ((lo < 0.0 && p[n - 1] < 0.0) || (lo > 0.0 && p[n - 1] > 0.0))) {
        const double y = lo * 2.0, x = hi + y;
        if (y == x - hi) hi = x;
    }
    return hi;
}

double exact_dot(const double *a, int64_t sa, const double *b, int64_t sb, int64_t K,
                 std::vector<double> &part) {
    size_t n = 0;
    for (int64_t k = 0; k < K; k++) {
        const double x = a[k * sa], y = b[k * sb];
        const double pr = x * y;
        if (!std::isfinite(pr)) return pr;
        const double er = std::fma(x, y, -pr);
        n = grow(part, n, pr);
        if (er != 0.0) n = grow(part, n, er);
    }
    return round_partials(part, n);
}

}  // namespace

extern "C" int32_t ozaie_exact_entries(const double *A, int64_t lda, const double *B, int64_t ldb, int64_t K,
                                       const int64_t *rows, const int64_t *cols, int64_t count, double *out) {
    if (!A || !B || !rows || !cols || !out || K < 0 || count < 0) return OZAIE_ERR_ARG;
#pragma omp parallel
    {
        std::vector<double> part(64);
#pragma omp for schedule(dynamic, 16)
        for (int64_t i = 0; i < count; i++)
            out[i] = exact_dot(A + rows[i] * lda, 1, B + cols[i], ldb, K, part);
    }
    return OZAIE_OK;
}
