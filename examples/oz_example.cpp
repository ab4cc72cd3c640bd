// Minimal consumer of the reference-compatible API: multiplies ones by ones in the
// accurate mode and checks the known answer C == K for every element.
#include <cstdio>
#include <vector>

#include "oz_aie.h"

int main() {
    const int M = 256, N = 512, K = 128;
    int status = oz_aie::ozAIE_setup(M, N, K, oz_aie::Mode::accurate);
    if (status > 0) {
        std::printf("setup failed: %d (%s)\n", status, ozaie_last_error());
        return status;
    }
    std::vector<double> A(M * K, 1.0), B(K * N, 1.0), C(M * N, 0.0);
    status = oz_aie::ozAIE_DGEMM(M, N, K, A.data(), B.data(), C.data());
    if (status > 0) {
        std::printf("DGEMM failed: %d (%s)\n", status, ozaie_last_error());
        return status;
    }
    for (double c : C)
        if (c != double(K)) {
            std::printf("wrong result %.17g\n", c);
            return 100;
        }
    std::printf("ok: C == %d everywhere\n", K);
    oz_aie::cleanup_xrt();
    return 0;
}
