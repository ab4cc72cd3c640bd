// oz_probe.hip -- measured ceiling of the int8 MFMA datapath on this GPU.
//
// The slice and modulus GEMMs are judged against the dense int8 peak of the datasheet, but
// the rate the matrix cores sustain depends on the operand data: the clock falls under the
// power the toggling operand bits draw.  This probe runs v_mfma_i32_32x32x32_i8 from
// registers only (no memory traffic in the timed loop, one wave per SIMD on every CU, 16
// independent accumulator tiles per wave, as the 128 x 128 wave tile of the CRT modulus GEMM
// has) with three operand patterns:
//   0  all-zero operands,
//   1  one random fragment pair, the same for every MFMA (constant operands),
//   2  eight random fragment pairs, a different pair for consecutive MFMAs (the operands
//      toggle like real slice or residue data),
//   3  pattern 2's operands on v_mfma_i32_16x16x64_i8 (64 independent 16 x 16 accumulator
//      tiles, the same int8 operations per round),
//   4  eight fragment pairs of bytes uniform in [-127, 127] (the symmetric residues of the
//      CRT mode's moduli and the sign-magnitude slices of the split scheme).
// It measures; it is not part of any product.
#include "oz_device.h"

#include <vector>

#include "../../include/ozaie.h"

namespace oz {

constexpr int kProbeSets = 8;

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1)))
void k_mfma_probe(const v4i *__restrict__ frags, int iters, int32_t *__restrict__ sink) {
    const int lane = threadIdx.x & 63;
    v4i a[kProbeSets], b[kProbeSets];
#pragma unroll
    for (int s = 0; s < kProbeSets; s++) {
        a[s] = frags[(2 * s) * 64 + lane];
        b[s] = frags[(2 * s + 1) * 64 + lane];
    }
    v16i acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = v16i{};
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int s = 0; s < kProbeSets; s++)
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[(s + i) % kProbeSets], b[(s + j) % kProbeSets],
                                                                      acc[i][j], 0, 0, 0);
    }
    int32_t x = 0;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int e = 0; e < 16; e++) x ^= acc[i][j][e];
    sink[blockIdx.x * blockDim.x + threadIdx.x] = x;
}

// Pattern 3: the same operands and int8 operations per round on the 16 x 16 x 64 MFMA
// (half the operations per instruction, so 8 x 8 accumulator tiles and four fragment-set
// steps per round).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1)))
void k_mfma_probe16(const v4i *__restrict__ frags, int iters, int32_t *__restrict__ sink) {
    const int lane = threadIdx.x & 63;
    v4i a[kProbeSets], b[kProbeSets];
#pragma unroll
    for (int s = 0; s < kProbeSets; s++) {
        a[s] = frags[(2 * s) * 64 + lane];
        b[s] = frags[(2 * s + 1) * 64 + lane];
    }
    v4i acc[8][8];
#pragma unroll
    for (int i = 0; i < 8; i++)
#pragma unroll
        for (int j = 0; j < 8; j++) acc[i][j] = v4i{};
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int s = 0; s < kProbeSets; s += 2)
#pragma unroll
            for (int i = 0; i < 8; i++)
#pragma unroll
                for (int j = 0; j < 8; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[(s + i) % kProbeSets], b[(s + j) % kProbeSets],
                                                                      acc[i][j], 0, 0, 0);
    }
    int32_t x = 0;
#pragma unroll
    for (int i = 0; i < 8; i++)
#pragma unroll
        for (int j = 0; j < 8; j++)
#pragma unroll
            for (int e = 0; e < 4; e++) x ^= acc[i][j][e];
    sink[blockIdx.x * blockDim.x + threadIdx.x] = x;
}

}  // namespace oz

using namespace oz;

extern "C" int32_t ozaie_probe_mfma(int pattern, int iters, double *pops_out, double *ms_out) {
    if (!pops_out || pattern < 0 || pattern > 4 || iters < 1) return OZAIE_ERR_ARG;
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
        (void)hipGetLastError();
        return OZAIE_ERR_NODEVICE;
    }
    // operand fragments: 8 pairs x 64 lanes x 16 bytes
    constexpr size_t kPairWords = 2 * 64 * 4;   // one A and one B fragment of 16 bytes per lane
    std::vector<int32_t> h(kProbeSets * kPairWords);
    uint32_t x = 0x9E3779B9u;
    for (size_t i = 0; i < h.size(); i++) {
        x ^= x << 13; x ^= x >> 17; x ^= x << 5;
        h[i] = pattern == 0 ? 0 : (int32_t)x;
        if (pattern == 4) {   // four bytes uniform in [-127, 127]
            uint32_t w = 0, y = x;
            for (int k = 0; k < 4; k++) {
                y = y * 1664525u + 1013904223u;
                w |= (uint32_t)(((int)((y >> 8) % 255u) - 127) & 0xFF) << (8 * k);
            }
            h[i] = (int32_t)w;
        }
    }
    if (pattern == 1)   // the first random pair in every set
        for (size_t i = kPairWords; i < h.size(); i++) h[i] = h[i % kPairWords];
    v4i *dfr = nullptr;
    int32_t *sink = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = OZAIE_OK;
    float ms = 0.f;
    if (hipMalloc(&dfr, h.size() * 4) != hipSuccess || hipMalloc(&sink, (size_t)ncu * 256 * 4) != hipSuccess ||
        hipMemcpy(dfr, h.data(), h.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        rc = set_error(hipGetLastError(), "ozaie_probe_mfma setup");
    } else {
        // a warm-up launch brings the clock to its loaded state, then the timed one
        auto kern = pattern == 3 ? k_mfma_probe16 : k_mfma_probe;
        hipLaunchKernelGGL(kern, dim3(ncu), dim3(256), 0, 0, dfr, iters, sink);
        (void)hipEventRecord(e0, 0);
        hipLaunchKernelGGL(kern, dim3(ncu), dim3(256), 0, 0, dfr, iters, sink);
        (void)hipEventRecord(e1, 0);
        hipError_t e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess) rc = set_error(e, "ozaie_probe_mfma");
    }
    if (!rc) {
        // ops = waves x iterations x MFMAs per iteration x 2 * 32 * 32 * 32 (patterns 0-2, 4;
        // pattern 3: twice the MFMAs of 2 * 16 * 16 * 64)
        const double ops = (double)ncu * 4 * iters * (kProbeSets * 16) * 65536.0;
        *pops_out = ops / (ms * 1e-3) / 1e15;
        if (ms_out) *ms_out = ms;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (dfr) (void)hipFree(dfr);
    if (sink) (void)hipFree(sink);
    return rc;
}
