// oz_crt.hip -- the CRT mode: residues of the integer-scaled operands, one exact int8
// GEMM per modulus, and the Chinese-remainder rebuild of the exact integer product.
#include "oz_device.h"

namespace oz {

// ------------------------------------------------------------------ CRT mode
// An exact integer product instead of the triangle of slice products: each element
// becomes q = sign(x) floor(|x| 2^(1022 + beta - E)) (|q| < 2^beta), the operands are
// multiplied modulo n pairwise-coprime moduli p <= 256 with one exact int8 GEMM each,
// and the Chinese remainder theorem rebuilds sum_k q_a q_b exactly (|C'| <= M / 4).

// Residues of the staged tile: thread (row, q) owns 16 consecutive k of one row and
// writes one 16-byte lane chunk per modulus plane.  With q = +-(hi 2^31 + lo), the value
// x = +-(hi (2^31 mod p) + lo) is an exact double (|x| < 2^40) congruent to q, and
// x - p rint(x / p) is its residue in [-(p-1)/2, (p-1)/2] for odd p: x / p is never within
// 1 / (2p) of a half-integer, far above the rounding error of x * fl(1/p).  For p = 256
// (plane 0) the residue is the low byte of q in two's complement.
template <bool KC>
__device__ __forceinline__ void crt_emit(const double *tile, int64_t R, const int32_t *__restrict__ Erow,
                                         const CrtConsts &c, int8_t *__restrict__ S, int64_t Rb, int64_t Kb,
                                         int64_t rb, int64_t kt) {
    constexpr int TK = 128;
    constexpr int LDW = KC ? (TK + 1) : (32 + 1);
    const int t = threadIdx.x;
    const int64_t r0 = rb * 32;
    const int row = t & 31, q = t >> 5;
    const int64_t kb = kt * 4 + (q >> 1);
    if (kb >= Kb) return;
    const int emax = (r0 + row < R) ? Erow[row] : 0;
    double hi[16], lo[16];
    uint32_t low8[4] = {0u, 0u, 0u, 0u};
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const double v = KC ? tile[row * LDW + 16 * q + i] : tile[(16 * q + i) * LDW + row];
        const uint64_t u = (uint64_t)__double_as_longlong(v);
        const int e = (int)((u >> 52) & 0x7FF);
        const uint64_t mx = (u & 0xFFFFFFFFFFFFFull) | (e ? (1ull << 52) : 0ull);
        const int sh = 53 + emax - (e ? e : 1) - c.beta;   // subnormals: exponent 1 - 1023
        const uint64_t Q = sh >= 64 ? 0ull : (sh >= 0 ? mx >> sh : mx << (-sh));
        const bool neg = (u >> 63) != 0;
        const double h = (double)(uint32_t)(Q >> 31), l = (double)(uint32_t)(Q & 0x7FFFFFFFu);
        hi[i] = neg ? -h : h;
        lo[i] = neg ? -l : l;
        const uint32_t qlo = neg ? (uint32_t)(0u - (uint32_t)Q) : (uint32_t)Q;   // low word of q
        low8[i >> 2] |= (qlo & 0xFFu) << (8 * (i & 3));
    }
    const int64_t plane = Rb * Kb * (int64_t)kBlkBytes;
    int8_t *dst = S + (rb * Kb + kb) * (int64_t)kBlkBytes + (q & 1) * 512 + row * 16;
    *reinterpret_cast<uint4 *>(dst) = make_uint4(low8[0], low8[1], low8[2], low8[3]);   // p = 256
    for (int j = 1; j < c.n; j++) {
        const double p = c.p[j], inv = c.inv[j], c31 = c.c31[j];
        uint32_t w[4];
#pragma unroll
        for (int gi = 0; gi < 4; gi++) {
            uint32_t acc = 0;
#pragma unroll
            for (int h = 0; h < 4; h++) {
                const int i = 4 * gi + h;
                const double x = __builtin_fma(hi[i], c31, lo[i]);
                const double r = __builtin_fma(-rint(x * inv), p, x);
                acc |= ((uint32_t)(int)r & 0xFFu) << (8 * h);
            }
            w[gi] = acc;
        }
        *reinterpret_cast<uint4 *>(dst + (int64_t)j * plane) = make_uint4(w[0], w[1], w[2], w[3]);
    }
}

template <bool KC>
__global__ __launch_bounds__(256) void k_crt_split(const double *__restrict__ X, int64_t sr, int64_t sk,
                                                   int64_t R, int64_t K, const int32_t *__restrict__ E,
                                                   int8_t *__restrict__ S, int64_t Rb, int64_t Kb, CrtConsts c) {
    __shared__ double tile[kSplitLds / 8];
    double v[16];
    split_gather<KC>(X, sr, sk, R, K, blockIdx.y, blockIdx.x, v);
    split_stage<KC>(v, tile);
    __syncthreads();
    crt_emit<KC>(tile, R, E + (int64_t)blockIdx.y * 32, c, S, Rb, Kb, blockIdx.y, blockIdx.x);
}

hipError_t launch_crt_split(const OperandView &op, const int32_t *E, const CrtConsts &c, int8_t *S, int64_t Rp,
                            int64_t Kp, hipStream_t st) {
    const int64_t Rb = Rp / kBlk, Kb = Kp / kBlk;
    const int64_t nkt = (Kb + 3) / 4;
    if (Rb > 65535 || nkt > 0x7fffffff) return hipErrorInvalidValue;
    if (op.sk == 1)
        hipLaunchKernelGGL(k_crt_split<true>, dim3((unsigned)nkt, (unsigned)Rb), dim3(256), 0, st, op.X, op.sr,
                           op.sk, op.R, op.K, E, S, Rb, Kb, c);
    else
        hipLaunchKernelGGL(k_crt_split<false>, dim3((unsigned)nkt, (unsigned)Rb), dim3(256), 0, st, op.X, op.sr,
                           op.sk, op.R, op.K, E, S, Rb, Kb, c);
    return hipGetLastError();
}

// One exact int8 GEMM per modulus (blockIdx.y), wave tile (32 WM) x (32 WN).  A contraction
// block gives a wave only WM*WN MFMAs, too few to cover the L2 latency of its fragments,
// so the fragments of the next PF blocks are in flight in a register ring (Kb is a
// multiple of PF; the loads past the end re-read the last block, in bounds).  Epilogue:
// each int32 sum reduced mod p in FP64 (|sum| < 2^31 is exact) and four rows of one column
// packed into a word of R[j][row / 4][col].

template <int WM, int WN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1)))
void k_crt_gemm(CrtGemm g, CrtConsts c) {
    const int j = blockIdx.y;
    const int ntm = (int)(g.Mp / (64 * WM)), ntn = (int)(g.Np / (64 * WN));
    int tm, tn;
    tile_map(blockIdx.x, ntm, ntn, tm, tn);
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = w & 1, wn = w >> 1;
    const int64_t rowstride = g.Kb * kBlkBytes;
    const int64_t rbA = (int64_t)tm * 2 * WM + wm * WM, rbB = (int64_t)tn * 2 * WN + wn * WN;
    const int8_t *pa = g.SA + (int64_t)j * (g.Mp / kBlk) * rowstride + rbA * rowstride + lane * 16;
    const int8_t *pb = g.SB + (int64_t)j * (g.Np / kBlk) * rowstride + rbB * rowstride + lane * 16;
    v4i fa[kCrtPF][WM], fb[kCrtPF][WN];
    const int64_t nk = g.Kb;   // multiple of kCrtPF
    auto load = [&](int s, int64_t kb) {
        kb = kb < nk ? kb : nk - 1;
#pragma unroll
        for (int mi = 0; mi < WM; mi++) fa[s][mi] = *reinterpret_cast<const v4i *>(pa + mi * rowstride + kb * kBlkBytes);
#pragma unroll
        for (int ni = 0; ni < WN; ni++) fb[s][ni] = *reinterpret_cast<const v4i *>(pb + ni * rowstride + kb * kBlkBytes);
    };
    v16i acc[WM][WN];
#pragma unroll
    for (int mi = 0; mi < WM; mi++)
#pragma unroll
        for (int ni = 0; ni < WN; ni++) acc[mi][ni] = v16i{};
    auto mma = [&](int s) {
#pragma unroll
        for (int mi = 0; mi < WM; mi++)
#pragma unroll
            for (int ni = 0; ni < WN; ni++)
                acc[mi][ni] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[s][mi], fb[s][ni], acc[mi][ni], 0, 0, 0);
    };
#pragma unroll
    for (int s = 0; s < kCrtPF; s++) load(s, s);
    for (int64_t t = 0; t < nk; t += kCrtPF) {
#pragma unroll
        for (int s = 0; s < kCrtPF; s++) {
            mma(s);
            load(s, t + kCrtPF + s);
        }
    }
    const double p = c.p[j], inv = c.inv[j];
    const int64_t rbase = (int64_t)tm * 64 * WM + wm * 32 * WM, cbase = (int64_t)tn * 64 * WN + wn * 32 * WN;
    uint32_t *Rj = g.R + (int64_t)j * (g.Mp / 4) * g.Np;
#pragma unroll
    for (int mi = 0; mi < WM; mi++)
#pragma unroll
        for (int ni = 0; ni < WN; ni++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                uint32_t word = 0;
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const double x = (double)acc[mi][ni][4 * q + e];
                    double r = __builtin_fma(-rint(x * inv), p, x);   // in [-p/2, p/2]
                    r = r < 0.0 ? r + p : r;                           // in [0, p)
                    word |= (uint32_t)r << (8 * e);
                }
                const int64_t row4 = (rbase + mi * 32 + q * 8 + (lane >> 5) * 4) >> 2;
                const int64_t col = cbase + ni * 32 + (lane & 31);
                Rj[row4 * g.Np + col] = word;
            }
}

// Shared-fragment variant: the workgroup's 8 A and 8 B blocks of each contraction block
// (16 KiB) are loaded once per workgroup -- every thread brings 4 x 16 bytes of whole
// 1 KiB blocks into a register ring kCrtPF blocks deep -- and written to one of two LDS
// stages; after a barrier each wave reads its 4 + 4 fragments from LDS.  Halves the
// L2 -> CU traffic of the per-wave loads.  The barrier is a plain s_barrier after the
// wave's own LDS writes have completed (no memory fence), so the global loads of the
// ring stay in flight across it; the inline-asm memory clobber keeps the compiler from
// moving LDS accesses across it.
__device__ __forceinline__ void lds_barrier() {
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

template <int WM, int WN, int PF>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1)))
void k_crt_gemm_lds(CrtGemm g, CrtConsts c) {
    constexpr int BA = 2 * WM, BB = 2 * WN, NB = BA + BB;   // blocks per contraction block
    constexpr int PER = NB / 4;                            // blocks each wave brings in
    static_assert(NB % 4 == 0, "blocks must split over four waves");
    __shared__ __attribute__((aligned(1024))) int8_t smem[2 * NB * kBlkBytes];
    const int j = blockIdx.y;
    const int ntm = (int)(g.Mp / (64 * WM)), ntn = (int)(g.Np / (64 * WN));
    int tm, tn;
    tile_map(blockIdx.x, ntm, ntn, tm, tn);
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = w & 1, wn = w >> 1;
    const int64_t rowstride = g.Kb * kBlkBytes;
    const int8_t *A0 = g.SA + (int64_t)j * (g.Mp / kBlk) * rowstride + (int64_t)tm * BA * rowstride;
    const int8_t *B0 = g.SB + (int64_t)j * (g.Np / kBlk) * rowstride + (int64_t)tn * BB * rowstride;
    const int8_t *src[PER];
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int bi = w + 4 * i;   // block bi of the contraction block: A rows, then B columns
        src[i] = (bi < BA ? A0 + (int64_t)bi * rowstride : B0 + (int64_t)(bi - BA) * rowstride) + lane * 16;
    }
    const int64_t nk = g.Kb;   // multiple of PF
    v4i ring[PF][PER];
    auto fetch = [&](int s, int64_t kb) {
        kb = kb < nk ? kb : nk - 1;
#pragma unroll
        for (int i = 0; i < PER; i++) ring[s][i] = *reinterpret_cast<const v4i *>(src[i] + kb * kBlkBytes);
    };
    v16i acc[WM][WN];
#pragma unroll
    for (int mi = 0; mi < WM; mi++)
#pragma unroll
        for (int ni = 0; ni < WN; ni++) acc[mi][ni] = v16i{};
#pragma unroll
    for (int s = 0; s < PF; s++) fetch(s, s);
    for (int64_t t = 0; t < nk; t += PF) {
#pragma unroll
        for (int s = 0; s < PF; s++) {
            int8_t *st = smem + ((s & 1) * NB) * kBlkBytes;
#pragma unroll
            for (int i = 0; i < PER; i++)
                *reinterpret_cast<v4i *>(st + (w + 4 * i) * kBlkBytes + lane * 16) = ring[s][i];
            lds_barrier();
            fetch(s, t + PF + s);
            v4i af[WM], bf[WN];
#pragma unroll
            for (int mi = 0; mi < WM; mi++)
                af[mi] = *reinterpret_cast<const v4i *>(st + (wm * WM + mi) * kBlkBytes + lane * 16);
#pragma unroll
            for (int ni = 0; ni < WN; ni++)
                bf[ni] = *reinterpret_cast<const v4i *>(st + (BA + wn * WN + ni) * kBlkBytes + lane * 16);
#pragma unroll
            for (int mi = 0; mi < WM; mi++)
#pragma unroll
                for (int ni = 0; ni < WN; ni++)
                    acc[mi][ni] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af[mi], bf[ni], acc[mi][ni], 0, 0, 0);
        }
    }
    const double p = c.p[j], inv = c.inv[j];
    const int64_t rbase = (int64_t)tm * 64 * WM + wm * 32 * WM, cbase = (int64_t)tn * 64 * WN + wn * 32 * WN;
    uint32_t *Rj = g.R + (int64_t)j * (g.Mp / 4) * g.Np;
#pragma unroll
    for (int mi = 0; mi < WM; mi++)
#pragma unroll
        for (int ni = 0; ni < WN; ni++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                uint32_t word = 0;
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const double x = (double)acc[mi][ni][4 * q + e];
                    double r = __builtin_fma(-rint(x * inv), p, x);
                    r = r < 0.0 ? r + p : r;
                    word |= (uint32_t)r << (8 * e);
                }
                const int64_t row4 = (rbase + mi * 32 + q * 8 + (lane >> 5) * 4) >> 2;
                const int64_t col = cbase + ni * 32 + (lane & 31);
                Rj[row4 * g.Np + col] = word;
            }
}

hipError_t launch_crt_gemm(const CrtGemm &g, const CrtConsts &c, hipStream_t st) {
    constexpr int WM = 4, WN = 4;
    if (g.Mp % (64 * WM) || g.Np % (64 * WN) || g.Kb % kCrtPF || g.Kb <= 0 || c.n < 1 || c.n > kCrtMaxMod)
        return hipErrorInvalidValue;
    const int64_t tiles = (g.Mp / (64 * WM)) * (g.Np / (64 * WN));
    if (tiles <= 0) return hipSuccess;
    if (tiles > 0x7fffffff) return hipErrorInvalidValue;
    if (g.variant == 1)
        hipLaunchKernelGGL((k_crt_gemm<WM, WN>), dim3((unsigned)tiles, (unsigned)c.n), dim3(256), 0, st, g, c);
    else
        hipLaunchKernelGGL((k_crt_gemm_lds<WM, WN, kCrtPF>), dim3((unsigned)tiles, (unsigned)c.n), dim3(256), 0, st,
                           g, c);
    return hipGetLastError();
}

// C' * 2^e2 to double, truncated (rounding 0) or nearest-even (1) to 53 bits.
__device__ __forceinline__ double crt_to_double(i128 t, int e2, int rounding) {
    if (t == 0) return 0.0;
    const bool neg = t < 0;
    const u128 a = neg ? (u128)(-t) : (u128)t;
    const int nb = 128 - clz128(a);
    uint64_t m;
    int sh = 0;
    if (nb <= 53) {
        m = (uint64_t)a;
    } else {
        sh = nb - 53;
        u128 qv = a >> sh;
        if (rounding) {
            const u128 rem = a - (qv << sh);
            const u128 hf = (u128)1 << (sh - 1);
            if (rem > hf || (rem == hf && (qv & 1))) qv += 1;
        }
        m = (uint64_t)qv;
    }
    const double v = ldexp((double)m, sh + e2);
    return neg ? -v : v;
}

// Thread per (4 rows, 1 column): the n residue words, then per row k = nearest(sum g f)
// and C' = sum g W + k (2^128 - M) (mod 2^128) read as signed, summed in four 32-bit limb
// columns (each below 17 * 2^44, so 64-bit) with one carry pass.  Nearest rounding uses
// the correctly rounded 128-bit integer conversion (then the same power-of-two scaling as
// the truncating path).
__global__ __launch_bounds__(256) void k_crt_rebuild(const uint32_t *__restrict__ R, int64_t Mp, int64_t Np,
                                                     int64_t M, int64_t N, const int32_t *__restrict__ Er,
                                                     const int32_t *__restrict__ Ec, double *__restrict__ C,
                                                     int64_t ldc, int rounding, int64_t ncb, CrtConsts c) {
    const int64_t row4 = blockIdx.x / ncb;
    const int64_t col = (blockIdx.x % ncb) * 256 + threadIdx.x;
    if (col >= N) return;
    uint32_t wd[kCrtMaxMod];
#pragma unroll
    for (int j = 0; j < kCrtMaxMod; j++) wd[j] = j < c.n ? R[((int64_t)j * (Mp / 4) + row4) * Np + col] : 0u;
    const int ec = Ec[col];
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const int64_t row = row4 * 4 + e;
        if (row >= M) break;
        double tf = 0.0;
#pragma unroll
        for (int j = 0; j < kCrtMaxMod; j++) tf += (double)((wd[j] >> (8 * e)) & 0xFFu) * c.f[j];
        const uint64_t k = (uint64_t)(int64_t)rint(tf);
        uint64_t l0 = k * (uint32_t)c.Mneg[0], l1 = k * (uint32_t)(c.Mneg[0] >> 32);
        uint64_t l2 = k * (uint32_t)c.Mneg[1], l3 = k * (uint32_t)(c.Mneg[1] >> 32);
#pragma unroll
        for (int j = 0; j < kCrtMaxMod; j++) {   // planes j >= n hold zero words
            const uint64_t gv = (wd[j] >> (8 * e)) & 0xFFu;
            l0 += gv * (uint32_t)c.W[j][0];
            l1 += gv * (uint32_t)(c.W[j][0] >> 32);
            l2 += gv * (uint32_t)c.W[j][1];
            l3 += gv * (uint32_t)(c.W[j][1] >> 32);
        }
        const u128 S = (u128)l0 + ((u128)l1 << 32) + ((u128)l2 << 64) + ((u128)l3 << 96);
        const int e2 = Er[row] + ec - 2044 - 2 * c.beta;
        C[row * ldc + col] = rounding ? ldexp((double)(i128)S, e2) : crt_to_double((i128)S, e2, 0);
    }
}

hipError_t launch_crt_rebuild(const uint32_t *R, int64_t Mp, int64_t Np, int64_t M, int64_t N, const int32_t *Er,
                              const int32_t *Ec, double *C, int64_t ldc, int rounding, const CrtConsts &c,
                              hipStream_t st) {
    const int64_t ncb = (N + 255) / 256;
    const int64_t blocks = (Mp / 4) * ncb;
    if (blocks <= 0) return hipSuccess;
    if (blocks > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_crt_rebuild, dim3((unsigned)blocks), dim3(256), 0, st, R, Mp, Np, M, N, Er, Ec, C, ldc,
                       rounding, ncb, c);
    return hipGetLastError();
}

}  // namespace oz
