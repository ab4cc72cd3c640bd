// oz_kernels.hip -- the CDNA4 (gfx950) kernels of one Ozaki-scheme DGEMM (the tile
// bodies they share with the grouped kernels are in oz_device.h).
//
//   k_maxexp_*   per-row (A) / per-column (B) maximum exponent field along the
//                contraction axis (reference: ozaki_base.cc:342-440).
//   k_split      sign-magnitude split of every FP64 element into d b-bit digits relative
//                to its row/column maximum exponent, written as int8 slice blocks in the
//                MFMA operand order (reference rule: ozaki_base.cc:449-771; closed form
//                Q = floor(M_x * 2^(b*d - 53 - bias)), s_j = sign*((Q >> b(d-1-j)) & (2^b-1))).
//   k_gemm_fused the triangular d(d+1)/2 int8 x int8 -> int32 slice products on
//                v_mfma_i32_32x32x32_i8 with every bin of the output tile held in
//                registers, and the bin summation fused into the epilogue: exact 128-bit
//                T = sum_k bin_k 2^(b(d-1-k)), then FP64 C (reference: mm_ozaki.cc:304-357,
//                ozaki_base.cc:143-340, include/oz_aie.h:92-150).
//   k_gemm_group the same products for any d <= 32, eight bins per pass, for the
//                reference's b=3 presets whose 25..29 bins do not fit in registers.
#include "oz_device.h"

namespace oz {

// ------------------------------------------------------------------ small helpers
__global__ void k_fill_i32(int32_t *__restrict__ p, int64_t n, int32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

hipError_t launch_fill_i32(int32_t *p, int64_t n, int32_t v, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int64_t nb = (n + 255) / 256;
    hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)nb), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

// ------------------------------------------------------------------ max exponent
__global__ __launch_bounds__(256) void k_maxexp_kc(const double *__restrict__ X, int64_t sr,
                                                   int64_t R, int64_t K, int64_t Rp,
                                                   int32_t *__restrict__ E) {
    const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= Rp) return;  // whole wave leaves together
    maxexp_kc_row(X, sr, R, K, E, r, threadIdx.x & 63);
}

__global__ __launch_bounds__(256) void k_maxexp_rc(const double *__restrict__ X, int64_t sk,
                                                   int64_t R, int64_t K, int64_t kper,
                                                   int32_t *__restrict__ E) {
    maxexp_rc_slab(X, sk, R, K, E, (int64_t)blockIdx.x * 256 + threadIdx.x, (int64_t)blockIdx.y * kper, kper);
}

hipError_t launch_maxexp(const OperandView &op, int32_t *E, int64_t Rp, hipStream_t st) {
    if (op.sk == 1) {
        const int64_t nb = (Rp + 3) / 4;
        hipLaunchKernelGGL(k_maxexp_kc, dim3((unsigned)nb), dim3(256), 0, st, op.X, op.sr, op.R,
                           op.K, Rp, E);
        return hipGetLastError();
    }
    hipError_t e = launch_fill_i32(E, Rp, 0, st);
    if (e != hipSuccess) return e;
    if (op.R <= 0 || op.K <= 0) return hipSuccess;
    const int64_t nbx = (op.R + 255) / 256;
    int64_t nby = 2048 / nbx;
    const int64_t kmax = (op.K + 63) / 64;
    if (nby > kmax) nby = kmax;
    if (nby < 1) nby = 1;
    if (nby > 65535) nby = 65535;
    const int64_t kper = (op.K + nby - 1) / nby;
    nby = (op.K + kper - 1) / kper;
    hipLaunchKernelGGL(k_maxexp_rc, dim3((unsigned)nbx, (unsigned)nby), dim3(256), 0, st, op.X,
                       op.sk, op.R, op.K, kper, E);
    return hipGetLastError();
}

// grid (k tiles, row blocks, plane groups of `per` planes).  Eout: the grid has one k tile,
// which holds the whole contraction axis, so the exponent maxima come from the tile itself
// (the workgroups of plane group 0 store them to Eout); otherwise E holds them.
template <bool KC>
__global__ __launch_bounds__(256) void k_split(const double *__restrict__ X, int64_t sr,
                                               int64_t sk, int64_t R, int64_t K,
                                               const int32_t *__restrict__ E, int b, int d,
                                               int8_t *__restrict__ S, int64_t Rb, int64_t Kb, int per,
                                               int32_t *__restrict__ Eout) {
    __shared__ double tile[kSplitLds / 8];
    const int j0 = (int)blockIdx.z * per;
    if (Eout) {
        __shared__ int32_t red[32 + 256];
        double v[16];
        split_gather<KC>(X, sr, sk, R, K, blockIdx.y, blockIdx.x, v);
        split_stage<KC>(v, tile);
        __syncthreads();
        tile_row_max<KC>(tile, red, Eout + (int64_t)blockIdx.y * 32, blockIdx.z == 0);
        split_emit<KC>(tile, R, red, b, d, S, Rb, Kb, blockIdx.y, blockIdx.x, j0, j0 + per);
        return;
    }
    split_tile<KC>(X, sr, sk, R, K, E + (int64_t)blockIdx.y * 32, b, d, S, Rb, Kb, blockIdx.y, blockIdx.x, tile,
                   j0, j0 + per);
}

// Plane groups of a split launch: a grid of fewer tiles than about two per CU (512 on the
// MI355X's 256 CUs) spreads the planes of each tile over several workgroups, each of which
// reads the tile again (from L2 at such sizes) and writes `per` of the planes.  Returns the
// group count and sets per.
static int split_plane_groups(int64_t tiles, int planes, int &per) {
    int64_t g = (tiles >= 512 || tiles < 1) ? 1 : (512 + tiles - 1) / tiles;
    if (g > planes) g = planes;
    if (g < 1) g = 1;
    per = (int)((planes + g - 1) / g);
    return (planes + per - 1) / per;
}

// Conversion check: the exact integers T[i] = hi[i] 2^64 + lo[i] (two's complement) to FP64
// with the epilogue's conversions: out128[i] through t_to_double, out64[i] through
// t64_to_double when T fits int64 (else 0).
__global__ __launch_bounds__(256) void k_debug_convert(const int64_t *__restrict__ hi, const uint64_t *__restrict__ lo,
                                                       const int32_t *__restrict__ es, int64_t n, int b, int d,
                                                       int rounding, double *out128, double *out64) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const i128 t = (i128)(((u128)(uint64_t)hi[i] << 64) | lo[i]);
    out128[i] = t_to_double(t, es[i], b, d, rounding);
    const bool fits = hi[i] == ((int64_t)lo[i] >> 63);
    out64[i] = fits ? t64_to_double((int64_t)lo[i], es[i], b, d, rounding) : 0.0;
}

hipError_t launch_debug_convert(const int64_t *hi, const uint64_t *lo, const int32_t *es, int64_t n, int b, int d,
                                int rounding, double *out128, double *out64, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if ((n + 255) / 256 > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_debug_convert, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, hi, lo, es, n, b, d,
                       rounding, out128, out64);
    return hipGetLastError();
}

hipError_t launch_split(const OperandView &op, const int32_t *E, int b, int d, int8_t *S,
                        int64_t Rp, int64_t Kp, hipStream_t st, int32_t *Eout) {
    const int64_t Rb = Rp / kBlk, Kb = Kp / kBlk;
    const int64_t nkt = (Kb + 3) / 4;
    if (Rb > 65535 || nkt > 0x7fffffff || d < 1 || (Eout && nkt != 1)) return hipErrorInvalidValue;
    int per;
    const int groups = split_plane_groups(nkt * Rb, d, per);
    const dim3 grid((unsigned)nkt, (unsigned)Rb, (unsigned)groups);
    if (op.sk == 1)
        hipLaunchKernelGGL(k_split<true>, grid, dim3(256), 0, st, op.X, op.sr, op.sk, op.R, op.K, E, b, d, S, Rb,
                           Kb, per, Eout);
    else
        hipLaunchKernelGGL(k_split<false>, grid, dim3(256), 0, st, op.X, op.sr, op.sk, op.R, op.K, E, b, d, S, Rb,
                           Kb, per, Eout);
    return hipGetLastError();
}

// Paired-digit split (b <= 3 presets): the same grid and staging as k_split, pair planes
// and single-digit slots written by emit_pairs16 (grid z: groups of `per` of the P + Q);
// Eout as for k_split.
template <bool KC>
__global__ __launch_bounds__(256) void k_split_pairs(const double *__restrict__ X, int64_t sr, int64_t sk,
                                                     int64_t R, int64_t K, const int32_t *__restrict__ E, int b,
                                                     int d, int8_t *__restrict__ S, int64_t Rb, int64_t Kb,
                                                     int reverse_q, int per, int32_t *__restrict__ Eout) {
    __shared__ double tile[kSplitLds / 8];
    __shared__ int32_t red[32 + 256];
    double v[16];
    split_gather<KC>(X, sr, sk, R, K, blockIdx.y, blockIdx.x, v);
    split_stage<KC>(v, tile);
    __syncthreads();
    const int32_t *Erow = E + (int64_t)blockIdx.y * 32;
    if (Eout) {
        tile_row_max<KC>(tile, red, Eout + (int64_t)blockIdx.y * 32, blockIdx.z == 0);
        Erow = red;
    }
    const int g0 = (int)blockIdx.z * per;
    split_emit_pairs<KC>(tile, R, Erow, b, d, S, Rb, Kb, blockIdx.y, blockIdx.x, reverse_q != 0, g0, g0 + per);
}

hipError_t launch_split_pairs(const OperandView &op, const int32_t *E, int b, int d, int8_t *S, int64_t Rp,
                              int64_t Kp, bool reverse_q, hipStream_t st, int32_t *Eout) {
    const int64_t Rb = Rp / kBlk, Kb = Kp / kBlk;
    const int64_t nkt = (Kb + 3) / 4;
    if (Rb > 65535 || nkt > 0x7fffffff || b < 1 || b > 3 || d < 2 || b * d > 96 || (Eout && nkt != 1))
        return hipErrorInvalidValue;
    const PairCfg pc = pair_cfg(b, d);
    int per;
    const int groups = split_plane_groups(nkt * Rb, pc.P + pc.Q, per);
    const dim3 grid((unsigned)nkt, (unsigned)Rb, (unsigned)groups);
    if (op.sk == 1)
        hipLaunchKernelGGL(k_split_pairs<true>, grid, dim3(256), 0, st, op.X, op.sr, op.sk, op.R, op.K, E, b, d, S,
                           Rb, Kb, (int)reverse_q, per, Eout);
    else
        hipLaunchKernelGGL(k_split_pairs<false>, grid, dim3(256), 0, st, op.X, op.sr, op.sk, op.R, op.K, E, b, d, S,
                           Rb, Kb, (int)reverse_q, per, Eout);
    return hipGetLastError();
}

// The anti-diagonal of the paired-digit path: one int8 GEMM over nblk contraction blocks
// (Q single-digit slots x Kb blocks, contiguous per 32-row block), workgroup tile (64 WM) x
// (64 WN) through the CRT modulus GEMM's LDS loop, int32 result stored whole (the sums stay
// below Q K (2^b - 1)^2 < 2^31, which the planner checks).
template <int WM, int WN>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(crt_occupancy(WM * WN), crt_occupancy(WM * WN))))
void k_gemm_antidiag(const int8_t *__restrict__ SxA, const int8_t *__restrict__ SxB, int64_t Mp, int64_t Np,
                     int64_t nblk, int32_t *__restrict__ X, int64_t ldx) {
    __shared__ __attribute__((aligned(1024))) int8_t smem[kCrtStages * (2 * WM + 2 * WN) * kBlkBytes];
    const int ntm = (int)(Mp / (64 * WM)), ntn = (int)(Np / (64 * WN));
    int tm, tn;
    tile_map(blockIdx.x, ntm, ntn, tm, tn);
    const int64_t rowstride = nblk * kBlkBytes;
    v16i acc[WM][WN];
    CrtTileLoop<WM, WN, kCrtStages> L;
    L.setup(SxA + (int64_t)tm * 2 * WM * rowstride, SxB + (int64_t)tn * 2 * WN * rowstride, rowstride, nblk);
    L.prime(smem);
    L.run(smem, acc);
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = w & 1, wn = w >> 1;
    const int64_t rbase = (int64_t)tm * 64 * WM + wm * 32 * WM, cbase = (int64_t)tn * 64 * WN + wn * 32 * WN;
#pragma unroll
    for (int mi = 0; mi < WM; mi++)
#pragma unroll
        for (int ni = 0; ni < WN; ni++)
#pragma unroll
            for (int q = 0; q < 4; q++)
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const int64_t row = rbase + mi * 32 + q * 8 + (lane >> 5) * 4 + e;
                    const int64_t col = cbase + ni * 32 + (lane & 31);
                    X[row * ldx + col] = acc[mi][ni][4 * q + e];
                }
}

// The workgroup tile is the largest of 256, 128 and 64 square whose grid still gives every
// CU a workgroup (cus of them), so that small products are not left to a few CUs.
hipError_t launch_gemm_antidiag(const int8_t *SxA, const int8_t *SxB, int64_t Mp, int64_t Np, int64_t nblk,
                                int32_t *X, int64_t ldx, int cus, hipStream_t st) {
    if (Mp % 256 || Np % 256 || nblk < 1 || ldx < Np) return hipErrorInvalidValue;
    if (Mp == 0 || Np == 0) return hipSuccess;
    int t = 256;
    while (t > 64 && (Mp / t) * (Np / t) < (int64_t)cus) t /= 2;
    const int64_t tiles = (Mp / t) * (Np / t);
    if (tiles > 0x7fffffff) return hipErrorInvalidValue;
    if (t == 256)
        hipLaunchKernelGGL((k_gemm_antidiag<4, 4>), dim3((unsigned)tiles), dim3(256), 0, st, SxA, SxB, Mp, Np, nblk,
                           X, ldx);
    else if (t == 128)
        hipLaunchKernelGGL((k_gemm_antidiag<2, 2>), dim3((unsigned)tiles), dim3(256), 0, st, SxA, SxB, Mp, Np, nblk,
                           X, ldx);
    else
        hipLaunchKernelGGL((k_gemm_antidiag<1, 1>), dim3((unsigned)tiles), dim3(256), 0, st, SxA, SxB, Mp, Np, nblk,
                           X, ldx);
    return hipGetLastError();
}

template <int D, int WM, int WN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1)))
void k_gemm_fused(GemmLaunch g, int epi) {
    const int ntm = (g.M + 64 * WM - 1) / (64 * WM), ntn = (g.N + 64 * WN - 1) / (64 * WN);
    int tm, tn;
    tile_map(blockIdx.x, ntm, ntn, tm, tn);
    __shared__ __attribute__((aligned(1024))) int8_t smem[lds_tile_bytes<D, WM, WN>()];
    gemm_lds_tile<D, WM, WN>(g, tm, tn, epi, smem);
}

template <int D, int WM, int WN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1)))
void k_gemm_direct(GemmLaunch g, int epi) {
    const int ntm = (g.M + 64 * WM - 1) / (64 * WM), ntn = (g.N + 64 * WN - 1) / (64 * WN);
    int tm, tn;
    tile_map(blockIdx.x, ntm, ntn, tm, tn);
    __shared__ __attribute__((aligned(1024))) int8_t smem[direct_tile_bytes<D>()];
    gemm_direct_tile<D, WM, WN>(g, tm, tn, epi, smem);
}

template <int D, int WM, int WN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1)))
void k_gemm_roll(GemmLaunch g, int epi) {
    const int ntm = (g.M + 64 * WM - 1) / (64 * WM), ntn = (g.N + 64 * WN - 1) / (64 * WN);
    int tm, tn;
    tile_map(blockIdx.x, ntm, ntn, tm, tn);
    __shared__ __attribute__((aligned(1024))) int8_t smem[lds_tile_bytes<D, WM, WN>()];
    gemm_roll_tile<D, WM, WN>(g, tm, tn, epi, smem);
}

template <int D, int K0, int NB, int WM, int WN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1)))
void k_gemm_window(GemmLaunch g, int epi) {
    const int ntm = (g.M + 64 * WM - 1) / (64 * WM), ntn = (g.N + 64 * WN - 1) / (64 * WN);
    int tm, tn;
    tile_map(blockIdx.x, ntm, ntn, tm, tn);
    __shared__ __attribute__((aligned(1024))) int8_t smem[window_tile_bytes<D, K0, NB, WM, WN>()];
    gemm_window_tile<D, K0, NB, WM, WN>(g, tm, tn, epi, smem);
}

// Pipelined LDS variant: the slice blocks of block t+2 are copied HBM -> LDS
// (global_load_lds) while block t computes; fragments are double-buffered in registers
// and the workgroup barrier sits in the middle of each block's MFMAs.  After the
// barrier every wave reads block t+1's fragments and issues block t+2's copies, and the
// second half of block t's MFMAs runs while those are in flight, so neither the LDS read
// latency nor the barrier leaves the MFMA pipe idle.  Requires an even number of
// contraction blocks per pass.
template <int D>
__host__ __device__ constexpr int pipe_split() {
    int tot = D * (D + 1) / 2, acc = 0;
    for (int i = 0; i < D; i++) {
        acc += D - i;
        if (2 * acc >= tot) return i + 1;
    }
    return D;
}

template <int D, int WM, int WN>
__device__ __forceinline__ void gemm_pipe_tile(const GemmLaunch &g, int tm, int tn, int epi) {
    using F = FusedCfg<D, WM, WN>;
    constexpr int BA = F::BA, BB = F::BB, NBLK = F::NBLK, NLD = F::NLD, NMIN = F::NMIN;
    constexpr int I1 = pipe_split<D>();
    __shared__ __attribute__((aligned(1024))) int8_t smem[2 * F::STAGE];

    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = w & 1, wn = w >> 1;
    const int64_t rbA0 = (int64_t)tm * BA, rbB0 = (int64_t)tn * BB;
    const int64_t planeA = g.RbA * g.Kb * kBlkBytes, planeB = g.RbB * g.Kb * kBlkBytes;
    const int rbase = tm * 64 * WM + wm * 32 * WM, cbase = tn * 64 * WN + wn * 32 * WN;
    int er_v, ec_v;
    load_tile_exps(g, epi, lane, rbase, 32 * WM, cbase, 32 * WN, er_v, ec_v);

    const int8_t *src[NLD];
    int ldsoff[NLD];
#pragma unroll
    for (int it = 0; it < NLD; it++) {
        const int blk = w + 4 * it;
        const int bc = blk < NBLK ? blk : NBLK - 1;
        if (bc < D * BA) {
            const int j = bc / BA, mb = bc % BA;
            src[it] = g.SA + j * planeA + (rbA0 + mb) * g.Kb * kBlkBytes;
        } else {
            const int bb = bc - D * BA;
            const int j = bb / BB, nb = bb % BB;
            src[it] = g.SB + j * planeB + (rbB0 + nb) * g.Kb * kBlkBytes;
        }
        src[it] += lane * 16;
        ldsoff[it] = bc * kBlkBytes;
    }
    auto issue = [&](int stage, int64_t kb) {
#pragma unroll
        for (int it = 0; it < NLD; it++) {
            if (it < NMIN || w + 4 * it < NBLK)
                __builtin_amdgcn_global_load_lds((const void *)(src[it] + kb * kBlkBytes),
                                                 (lds_void *)(smem + stage * F::STAGE + ldsoff[it]), 16, 0, 0);
        }
    };
    v4i a0[D][WM], b0[D][WN], a1[D][WM], b1[D][WN];
    auto readf = [&](v4i (&af)[D][WM], v4i (&bf)[D][WN], int stage) {
        const int8_t *st = smem + stage * F::STAGE;
#pragma unroll
        for (int j = 0; j < D; j++) {
#pragma unroll
            for (int mi = 0; mi < WM; mi++)
                af[j][mi] = *reinterpret_cast<const v4i *>(st + (j * BA + wm * WM + mi) * kBlkBytes + lane * 16);
#pragma unroll
            for (int ni = 0; ni < WN; ni++)
                bf[j][ni] = *reinterpret_cast<const v4i *>(st + (D * BA + j * BB + wn * WN + ni) * kBlkBytes + lane * 16);
        }
    };
    v16i acc[D][WM][WN];
#pragma unroll
    for (int k = 0; k < D; k++)
#pragma unroll
        for (int mi = 0; mi < WM; mi++)
#pragma unroll
            for (int ni = 0; ni < WN; ni++) acc[k][mi][ni] = v16i{};
    auto mma = [&](const v4i (&af)[D][WM], const v4i (&bf)[D][WN], int i0, int i1) {
#pragma unroll
        for (int i = 0; i < D; i++) {
            if (i < i0 || i >= i1) continue;
#pragma unroll
            for (int j = 0; i + j < D; j++)
#pragma unroll
                for (int mi = 0; mi < WM; mi++)
#pragma unroll
                    for (int ni = 0; ni < WN; ni++)
                        acc[i + j][mi][ni] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af[i][mi], bf[j][ni],
                                                                                   acc[i + j][mi][ni], 0, 0, 0);
        }
    };

    const int64_t kb0 = g.kb0, nk = g.kb1 - g.kb0;   // nk is even and >= 2
    issue(0, kb0);
    __syncthreads();
    readf(a0, b0, 0);
    issue(1, kb0 + 1);
    for (int64_t t = 0; t < nk; t += 2) {
        // block t in (a0, b0); block t+1 copying into stage 1
        mma(a0, b0, 0, I1);
        __builtin_amdgcn_sched_barrier(0);
        __syncthreads();
        readf(a1, b1, 1);
        if (t + 2 < nk) issue(0, kb0 + t + 2);
        __builtin_amdgcn_sched_barrier(0);
        mma(a0, b0, I1, D);
        // block t+1 in (a1, b1); block t+2 copying into stage 0
        mma(a1, b1, 0, I1);
        __builtin_amdgcn_sched_barrier(0);
        if (t + 2 < nk) {
            __syncthreads();
            readf(a0, b0, 0);
            if (t + 3 < nk) issue(1, kb0 + t + 3);
        }
        __builtin_amdgcn_sched_barrier(0);
        mma(a1, b1, I1, D);
    }
    __syncthreads();

    int32_t *scratch = reinterpret_cast<int32_t *>(smem) + w * wave_scratch_words<D, WM, WN>();
    wave_epilogue<D, WM, WN>(g, epi, 0, acc, scratch, lane, rbase, cbase, er_v, ec_v);
}

template <int D, int WM, int WN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1)))
void k_gemm_pipe(GemmLaunch g, int epi) {
    const int ntm = (g.M + 64 * WM - 1) / (64 * WM), ntn = (g.N + 64 * WN - 1) / (64 * WN);
    int tm, tn;
    tile_map(blockIdx.x, ntm, ntn, tm, tn);
    gemm_pipe_tile<D, WM, WN>(g, tm, tn, epi);
}

template <int D, int WM, int WN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1)))
void k_gemm_staged(GemmLaunch g, int epi) {
    const int ntm = (g.M + 64 * WM - 1) / (64 * WM), ntn = (g.N + 64 * WN - 1) / (64 * WN);
    int tm, tn;
    tile_map(blockIdx.x, ntm, ntn, tm, tn);
    __shared__ __attribute__((aligned(1024))) int8_t smem[staged_tile_bytes<D, WM, WN>()];
    gemm_staged_tile<D, WM, WN>(g, tm, tn, epi, smem);
}

// d = 3 on 32 x 32 wave tiles with four LDS stages (48 KiB) and two waves per SIMD: two
// workgroups share a CU, so one's epilogue and first loads overlap the other's MFMAs.  It
// moves twice the bytes per int8 op through L2 and LDS of the 64 x 64 wave tile, which
// costs clock where the card runs at its power limit (measured on MI355X against the
// default tile: 2048^3 0.106 / 0.113 ms, 4096^3 0.517 / 0.546 ms, 8192^3 3.54 / 3.20 ms).
template <int D>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
void k_gemm_staged_x2(GemmLaunch g, int epi) {
    const int ntm = (g.M + 63) / 64, ntn = (g.N + 63) / 64;
    int tm, tn;
    tile_map(blockIdx.x, ntm, ntn, tm, tn);
    __shared__ __attribute__((aligned(1024))) int8_t smem[staged_tile_bytes<D, 1, 1, 4>()];
    gemm_staged_tile<D, 1, 1, 4>(g, tm, tn, epi, smem);
}

bool fused_supported(int d) { return d >= 1 && d <= 16; }

// Default fused-kernel family per slice count (measured on MI355X at 4096^3 and 8192^3,
// the families interleaved in one process, medians):
//   d <= 3: the staged kernel (12-36% faster than the others; at d = 4 it and the direct
//           kernel are within 4% of each other, either way round by size);
//   d = 7 and d = 9: the rolling kernel (d = 7: 1.6% faster than the direct kernel at
//           8192^3; d = 9 on 64 x 64 tiles: 3.6% faster than the LDS-copy kernel on
//           64 x 128 tiles at 8192^3, 2.3% at 4096^3, and ahead of the staged kernel);
//   other d <= 8: the direct-to-register kernel (its two fragment sets fit in registers;
//           at d = 8 the rolling kernel is 2.4% slower);
//   above: the LDS-copy kernel.
static constexpr int default_family(int d) {
    return d <= 3 ? KERNEL_STAGED : ((d == 7 || d == 9) ? KERNEL_ROLL : (d <= 8 ? KERNEL_DIRECT : KERNEL_LDS));
}

template <int D, int WM, int WN>
static hipError_t launch_fused_t(const GemmLaunch &g, int epi, hipStream_t st) {
    const int ntm = (g.M + 64 * WM - 1) / (64 * WM), ntn = (g.N + 64 * WN - 1) / (64 * WN);
    const int64_t tiles = (int64_t)ntm * ntn;
    if (tiles <= 0) return hipSuccess;
    int fam = g.family == KERNEL_AUTO ? default_family(D) : g.family;
    if constexpr (D <= 4 || (D == 9 && WM == 1 && WN == 1)) {
        if (fam == KERNEL_STAGED) {
            hipLaunchKernelGGL((k_gemm_staged<D, WM, WN>), dim3((unsigned)tiles), dim3(256), 0, st, g, epi);
            return hipGetLastError();
        }
    }
    if (fam == KERNEL_DIRECT)
        hipLaunchKernelGGL((k_gemm_direct<D, WM, WN>), dim3((unsigned)tiles), dim3(256), 0, st, g, epi);
    else if (fam == KERNEL_PIPE)
        hipLaunchKernelGGL((k_gemm_pipe<D, WM, WN>), dim3((unsigned)tiles), dim3(256), 0, st, g, epi);
    else if (fam == KERNEL_ROLL)
        hipLaunchKernelGGL((k_gemm_roll<D, WM, WN>), dim3((unsigned)tiles), dim3(256), 0, st, g, epi);
    else
        hipLaunchKernelGGL((k_gemm_fused<D, WM, WN>), dim3((unsigned)tiles), dim3(256), 0, st, g, epi);
    return hipGetLastError();
}

// Small products: the default family of D on 32 x 32 wave tiles (64 x 64 workgroup tiles),
// four times the workgroups of the 64 x 64 wave tiles of d <= 4 and twice those of the
// 32 x 64 ones of d = 5 .. 8.  Where the default grid leaves CUs idle, the smaller tile
// spreads the work (and the epilogue, a fixed cost per element that dominates at small K)
// over more of them.
template <int D>
static hipError_t launch_fused_small(const GemmLaunch &g, int epi, hipStream_t st) {
    const int64_t tiles = (int64_t)((g.M + 63) / 64) * ((g.N + 63) / 64);
    if (tiles <= 0) return hipSuccess;
    if constexpr (default_family(D) == KERNEL_STAGED)
        hipLaunchKernelGGL((k_gemm_staged<D, 1, 1>), dim3((unsigned)tiles), dim3(256), 0, st, g, epi);
    else if constexpr (default_family(D) == KERNEL_ROLL)
        hipLaunchKernelGGL((k_gemm_roll<D, 1, 1>), dim3((unsigned)tiles), dim3(256), 0, st, g, epi);
    else
        hipLaunchKernelGGL((k_gemm_direct<D, 1, 1>), dim3((unsigned)tiles), dim3(256), 0, st, g, epi);
    return hipGetLastError();
}

// Wave tile per slice count: the accumulators (d bins x 16 registers per 32 x 32 block)
// must leave room for the fragments in the 512 registers of a wave.  With the default
// family, d = 3 runs k_gemm_staged_x2 up to four default workgroups per CU, and another
// d <= 8 grid of fewer workgroups than the device has CUs runs on 32 x 32 wave tiles
// (launch_fused_small).
hipError_t launch_gemm_fused(const GemmLaunch &g, int epi, hipStream_t st) {
    if (g.d >= 1 && g.d <= 8 && g.family == KERNEL_AUTO && g.cus > 0) {
        const int tm = g.d <= 4 ? 128 : 64, tn = 128;   // the default workgroup tile
        const int64_t wgs = (int64_t)((g.M + tm - 1) / tm) * ((g.N + tn - 1) / tn);
        if (g.d == 3 && wgs <= 4 * (int64_t)g.cus) {   // up to four default workgroups per CU
            const int64_t tiles = (int64_t)((g.M + 63) / 64) * ((g.N + 63) / 64);
            if (tiles > 0x7fffffff) return hipErrorInvalidValue;
            hipLaunchKernelGGL((k_gemm_staged_x2<3>), dim3((unsigned)tiles), dim3(256), 0, st, g, epi);
            return hipGetLastError();
        }
        if (wgs < g.cus) {
            switch (g.d) {
                case 1: return launch_fused_small<1>(g, epi, st);
                case 2: return launch_fused_small<2>(g, epi, st);
                case 3: return launch_fused_small<3>(g, epi, st);
                case 4: return launch_fused_small<4>(g, epi, st);
                case 5: return launch_fused_small<5>(g, epi, st);
                case 6: return launch_fused_small<6>(g, epi, st);
                case 7: return launch_fused_small<7>(g, epi, st);
                default: return launch_fused_small<8>(g, epi, st);
            }
        }
    }
    switch (g.d) {
        case 1: return launch_fused_t<1, 2, 2>(g, epi, st);
        case 2: return launch_fused_t<2, 2, 2>(g, epi, st);
        case 3: return launch_fused_t<3, 2, 2>(g, epi, st);
        case 4: return launch_fused_t<4, 2, 2>(g, epi, st);
        case 5: return launch_fused_t<5, 1, 2>(g, epi, st);
        case 6: return launch_fused_t<6, 1, 2>(g, epi, st);
        case 7: return launch_fused_t<7, 1, 2>(g, epi, st);
        case 8: return launch_fused_t<8, 1, 2>(g, epi, st);
        case 9: {   // 1 x 2 only for the families that hold one fragment set (the others spill)
            const int fam = g.family == KERNEL_AUTO ? default_family(9) : g.family;
            return fam == KERNEL_LDS || fam == KERNEL_PIPE ? launch_fused_t<9, 1, 2>(g, epi, st)
                                                           : launch_fused_t<9, 1, 1>(g, epi, st);
        }
        case 10: return launch_fused_t<10, 1, 1>(g, epi, st);
        case 11: return launch_fused_t<11, 1, 1>(g, epi, st);
        case 12: return launch_fused_t<12, 1, 1>(g, epi, st);
        case 13: return launch_fused_t<13, 1, 1>(g, epi, st);
        case 14: return launch_fused_t<14, 1, 1>(g, epi, st);
        case 15: return launch_fused_t<15, 1, 1>(g, epi, st);
        case 16: return launch_fused_t<16, 1, 1>(g, epi, st);
        default: return hipErrorInvalidValue;
    }
}

// The products of a d = 18 split (the reference's accurate preset for K >= 2^14: b = 4,
// whose 4-bit digits do not pair into an int8 operand) as two bin windows of 9, each a
// pass of the rolling kernel on 32 x 32 wave tiles: 9 bins (144 accumulator registers) and
// one fragment set of up to 18 A and 18 B fragments (144 registers), two LDS stages of
// 18 x 4 KiB.  The bin-group kernel needs three passes of 8 bins.
bool window_supported(int d) { return d == 18; }

hipError_t launch_gemm_window(const GemmLaunch &g, int epi, hipStream_t st) {
    if (g.d != 18 || (g.k0 != 0 && g.k0 != kWindowBins)) return hipErrorInvalidValue;
    const int64_t tiles = (int64_t)((g.M + 63) / 64) * ((g.N + 63) / 64);
    if (tiles <= 0) return hipSuccess;
    if (tiles > 0x7fffffff) return hipErrorInvalidValue;
    if (g.k0 == 0)
        hipLaunchKernelGGL((k_gemm_window<18, 0, kWindowBins, 1, 1>), dim3((unsigned)tiles), dim3(256), 0, st, g, epi);
    else
        hipLaunchKernelGGL((k_gemm_window<18, kWindowBins, kWindowBins, 1, 1>), dim3((unsigned)tiles), dim3(256), 0,
                           st, g, epi);
    return hipGetLastError();
}

// ------------------------------------------------------------------ bin-group GEMM
// Any d <= 32: one pass computes bins [k0, k0 + 8) from the slices 0 .. nsl-1,
// nsl = min(d, k0 + 8).  Wave tile 32 x 32, workgroup 64 x 64, one LDS stage holding
// nsl A blocks and nsl B blocks per wave row/column.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1)))
void k_gemm_group(GemmLaunch g, int epi) {
    constexpr int NB = kGroupBins;
    constexpr int BA = 2, BB = 2;
    __shared__ __attribute__((aligned(1024))) int8_t smem[kMaxSlices * (BA + BB) * kBlkBytes];

    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = w & 1, wn = w >> 1;
    const int ntm = (g.M + 63) / 64, ntn = (g.N + 63) / 64;
    int tm, tn;
    tile_map(blockIdx.x, ntm, ntn, tm, tn);
    const int k0 = g.k0;
    const int nsl = min(g.d, k0 + NB);
    const int nblk = nsl * (BA + BB);
    const int64_t planeA = g.RbA * g.Kb * kBlkBytes, planeB = g.RbB * g.Kb * kBlkBytes;
    const int64_t rbA0 = (int64_t)tm * BA, rbB0 = (int64_t)tn * BB;
    const int rbase = tm * 64 + wm * 32, cbase = tn * 64 + wn * 32;
    int er_v, ec_v;
    load_tile_exps(g, epi, lane, rbase, 32, cbase, 32, er_v, ec_v);

    v16i acc[NB];
#pragma unroll
    for (int k = 0; k < NB; k++) acc[k] = v16i{};

    for (int64_t kb = g.kb0; kb < g.kb1; kb++) {
        for (int blk = w; blk < nblk; blk += 4) {
            const int8_t *s;
            if (blk < nsl * BA) {
                const int j = blk / BA, mb = blk % BA;
                s = g.SA + j * planeA + ((rbA0 + mb) * g.Kb + kb) * kBlkBytes;
            } else {
                const int bb = blk - nsl * BA;
                const int j = bb / BB, nb = bb % BB;
                s = g.SB + j * planeB + ((rbB0 + nb) * g.Kb + kb) * kBlkBytes;
            }
            __builtin_amdgcn_global_load_lds((const void *)(s + lane * 16),
                                             (lds_void *)(smem + blk * kBlkBytes), 16, 0, 0);
        }
        s_waitcnt_vm<0>();
        __syncthreads();
        for (int i = 0; i < nsl; i++) {
            const v4i a = *reinterpret_cast<const v4i *>(smem + (i * BA + wm) * kBlkBytes + lane * 16);
#pragma unroll
            for (int kk = 0; kk < NB; kk++) {
                const int j = k0 + kk - i;
                if (j >= 0 && j < nsl) {
                    const v4i bv = *reinterpret_cast<const v4i *>(
                        smem + (nsl * BA + j * BB + wn) * kBlkBytes + lane * 16);
                    acc[kk] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, bv, acc[kk], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }

    v16i acc3[NB][1][1];
#pragma unroll
    for (int k = 0; k < NB; k++) acc3[k][0][0] = acc[k];
    int32_t *scratch = reinterpret_cast<int32_t *>(smem) + w * wave_scratch_words<NB, 1, 1>();
    wave_epilogue<NB, 1, 1>(g, epi, k0, acc3, scratch, lane, rbase, cbase, er_v, ec_v);
}

hipError_t launch_gemm_group(const GemmLaunch &g, int epi, hipStream_t st) {
    const int ntm = (g.M + 63) / 64, ntn = (g.N + 63) / 64;
    const int64_t tiles = (int64_t)ntm * ntn;
    if (tiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gemm_group, dim3((unsigned)tiles), dim3(256), 0, st, g, epi);
    return hipGetLastError();
}

// ------------------------------------------------------------------ wire format
// Paired int64 {L, H} of the reference NPU output (ozaki_base.cc:143-340): L sums the
// n0 least significant bins, H the next n1 with bits 51..62 replaced by E_r + E_c and
// bit 63 kept as the sign of the sum; every shift-add wraps modulo 2^64 like acc64.
__global__ void k_pack_paired(const int32_t *__restrict__ bins, int64_t M, int64_t N, int b, int d,
                              int n0, const int32_t *__restrict__ Er, const int32_t *__restrict__ Ec,
                              int64_t *__restrict__ out) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= M * N) return;
    const int64_t r = idx / N, c = idx % N;
    const int64_t plane = M * N;
    const int n1 = d - n0;
    uint64_t L = 0, H = 0;
    for (int i = 0; i < n0; i++) L += (uint64_t)(int64_t)bins[(int64_t)(d - 1 - i) * plane + idx] << (b * i);
    for (int i = 0; i < n1; i++) H += (uint64_t)(int64_t)bins[(int64_t)(d - 1 - n0 - i) * plane + idx] << (b * i);
    const uint32_t h32 = (uint32_t)(H >> 32);
    uint32_t emb = ((h32 << 13) >> 13) + ((h32 >> 31) << 31);
    emb += (uint32_t)(int32_t)(int16_t)(Er[r] + Ec[c]) << 19;
    out[idx * 2] = (int64_t)L;
    out[idx * 2 + 1] = (int64_t)(((uint64_t)emb << 32) | (uint32_t)H);
}

hipError_t launch_pack_paired(const int32_t *bins, int64_t M, int64_t N, int b, int d, int n0,
                              const int32_t *Er, const int32_t *Ec, int64_t *out, hipStream_t st) {
    const int64_t n = M * N;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pack_paired, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, bins, M, N,
                       b, d, n0, Er, Ec, out);
    return hipGetLastError();
}

// ------------------------------------------------------------------ synthetic data
// (U - 0.5) * 2^e with a per-element xorshift; used by the autotuner so that timings see
// the operand toggling of real data (int8 MFMA throughput depends on it).
__global__ void k_fill_random(double *__restrict__ p, int64_t n, uint32_t seed) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t x = seed ^ (uint32_t)(i * 2654435761u) ^ (uint32_t)(i >> 32);
    x ^= x << 13; x ^= x >> 17; x ^= x << 5;
    x ^= x << 13; x ^= x >> 17; x ^= x << 5;
    const double u = (double)(x >> 8) * (1.0 / 16777216.0) - 0.5;
    p[i] = ldexp(u, (int)((x & 7u)) - 4);
}

hipError_t launch_fill_random(double *p, int64_t n, uint32_t seed, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_random, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, seed);
    return hipGetLastError();
}

}  // namespace oz
