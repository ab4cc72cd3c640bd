// oz_host.cpp -- host-side helpers of the C ABI that use OpenMP (built with g++ -fopenmp).
#include <omp.h>

#include <algorithm>
#include <cstddef>
#include <cstring>

namespace oz {

// memcpy of n bytes split over up to 16 OpenMP threads in page-aligned parts (one thread
// copies 5-10 GB/s between pageable and pinned host memory; the pinned side feeds a link
// of about 57 GB/s).  Copies below 1 MiB stay on the calling thread.
void parallel_copy(void *dst, const void *src, size_t n) {
    if (n < ((size_t)1 << 20)) {
        std::memcpy(dst, src, n);
        return;
    }
    const int threads = std::max(1, std::min(omp_get_max_threads(), 16));
#pragma omp parallel num_threads(threads)
    {
        const size_t t = (size_t)omp_get_thread_num(), T = (size_t)omp_get_num_threads();
        const size_t per = (n / T + 4095) & ~(size_t)4095;
        const size_t a = std::min(n, per * t), b = std::min(n, a + per);
        if (b > a) std::memcpy(static_cast<char *>(dst) + a, static_cast<const char *>(src) + a, b - a);
    }
}

}  // namespace oz
