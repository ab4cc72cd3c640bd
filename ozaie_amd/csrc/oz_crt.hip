// oz_crt.hip -- the CRT mode: residues of the integer-scaled operands, one exact int8
// GEMM per modulus, and the Chinese-remainder rebuild of the exact integer product.
#include "oz_device.h"

namespace oz {

// ------------------------------------------------------------------ CRT mode
// An exact integer product instead of the triangle of slice products: each element of a
// row with shift sigma becomes q = sign(x) round(|x| 2^(1075 - sigma)) (to nearest, ties
// away from zero), the operands are multiplied modulo n pairwise-coprime moduli p <= 256
// with one exact int8 GEMM each, and the Chinese remainder theorem rebuilds
// sum_k q_a q_b exactly (|C'| <= M / 4).
//
// Shifts.  Max rule (bits beta > 0): sigma = 53 + E - beta, E the row's maximum exponent
// field, so |q| <= 2^beta and K 2^(2 beta) <= M / 4 bounds |C'|.  Norm rule (the
// default): from an exact integer bound S on the row's Euclidean norm, ||x||^2 <=
// S 4^(E^ - 1038), built from chunks of 16 contraction indices: per chunk Ec = max(max
// exponent field, 1) and S_c = sum ceil(|x| / 2^(Ec - 1038))^2 (each ceiling <= 2^16),
// per row E^ = max Ec and S = sum_c ceil(S_c / 4^(E^ - Ec)).  m is the largest integer
// <= 46 with S (2^22 + 4) 4^m <= M 2^20 - K (2^20 + 2^40), and sigma = 37 + E^ - m.
// Then (2^m sqrt(S) + sqrt(K) / 2)^2 <= M / 4 (2ab <= a^2 2^-20 + b^2 2^20), which bounds
// ||q|| for every row and column, and |C'| <= ||q_a|| ||q_b|| <= M / 4 (Cauchy-Schwarz);
// |q| <= 2^(16 + m) <= 2^62.  All of it is integer arithmetic: the same on every run,
// layout and launch shape.

__host__ __device__ inline uint64_t crt_bits(double x) {
    uint64_t u;
    __builtin_memcpy(&u, &x, 8);
    return u;
}

// One chunk of 16 elements: (max exponent field << 40) | S_c (S_c < 2^37; 0 for a chunk
// of zeros).
__host__ __device__ inline uint64_t crt_chunk_stat(const double (&v)[16]) {
    int em = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const int e = (int)((crt_bits(v[i]) >> 52) & 0x7FF);
        em = e > em ? e : em;
    }
    const int ec = em > 1 ? em : 1;
    uint64_t s = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const uint64_t u = crt_bits(v[i]);
        const int e = (int)((u >> 52) & 0x7FF);
        const uint64_t mx = (u & 0xFFFFFFFFFFFFFull) | (e ? (1ull << 52) : 0ull);
        const int t = 37 + ec - (e ? e : 1);   // >= 37
        const uint32_t w = mx == 0 ? 0u : (t >= 64 ? 1u : (uint32_t)((mx + ((1ull << t) - 1)) >> t));   // <= 2^16
        s += (uint64_t)w * w;
    }
    return ((uint64_t)em << 40) | s;
}

// ceil(S_c / 4^(eh - max(E_c, 1))) of one chunk word (0 for a chunk of zeros).
__host__ __device__ inline uint64_t crt_chunk_term(uint64_t w, int eh) {
    const uint64_t sc = w & ((1ull << 40) - 1);
    if (!sc) return 0;
    const int ec = (int)(w >> 40);
    const int d2 = 2 * (eh - (ec > 1 ? ec : 1));   // S_c < 2^37: the ceiling is 1 from 2^37 on
    return d2 >= 40 ? 1 : (sc + ((1ull << d2) - 1)) >> d2;
}

__host__ __device__ inline int bitlen128(u128 a) {
    const uint64_t h = (uint64_t)(a >> 64), l = (uint64_t)a;
    return h ? 128 - __builtin_clzll(h) : (l ? 64 - __builtin_clzll(l) : 0);
}

// m of the norm rule for bound S > 0 and contraction length K (M 2^20 > K (2^20 + 2^40)).
__host__ __device__ inline int crt_norm_m(uint64_t S, int64_t K, const CrtConsts &c) {
    // R = M 2^20 - K (2^20 + 2^40), three words
    const uint64_t t = (uint64_t)K * ((1ull << 20) + (1ull << 40));   // K <= 2^16: < 2^57
    uint64_t r0 = c.Mw[0] << 20, r1 = (c.Mw[1] << 20) | (c.Mw[0] >> 44), r2 = (c.Mw[2] << 20) | (c.Mw[1] >> 44);
    const uint64_t b0 = r0 < t;
    r0 -= t;
    const uint64_t b1 = b0 & (uint64_t)(r1 == 0);
    r1 -= b0;
    r2 -= b1;
    const int bR = r2 ? 128 + 64 - __builtin_clzll(r2) : bitlen128(((u128)r1 << 64) | r0);
    const u128 Sp = (u128)S * ((1u << 22) + 4);   // < 2^71
    const int bS = bitlen128(Sp);
    int m = (bR - bS) / 2 + 1;
    if (m > 46) m = 46;
    for (;; m--) {
        bool fits;
        if (m >= 0) {   // Sp << 2m < 2^163: three words
            const int sh = 2 * m;
            const uint64_t s0 = (uint64_t)Sp, s1 = (uint64_t)(Sp >> 64);
            uint64_t a0, a1, a2;
            if (sh >= 128) { a0 = 0; a1 = 0; a2 = s0 << (sh - 128); }
            else if (sh >= 64) {
                a0 = 0;
                a1 = s0 << (sh - 64);
                a2 = (sh > 64 ? s0 >> (128 - sh) : 0) | (s1 << (sh - 64));
            } else {
                a0 = s0 << sh;
                a1 = (sh ? s0 >> (64 - sh) : 0) | (s1 << sh);
                a2 = sh ? s1 >> (64 - sh) : 0;
            }
            fits = a2 != r2 ? a2 < r2 : (a1 != r1 ? a1 < r1 : a0 <= r0);
        } else if (bR - 2 * m > 72) {
            fits = true;   // R 4^-m >= 2^72 > Sp
        } else {
            fits = Sp <= ((((u128)r1 << 64) | r0) << (-2 * m));
        }
        if (fits) return m;
    }
}

// sigma of one row from its chunk words st[0, nch) (whole wave; every lane returns it).
__device__ __forceinline__ int crt_row_sigma(const uint64_t *__restrict__ st, int64_t nch, int64_t K, int beta,
                                             const CrtConsts &c, int lane) {
    int em = 0;
    bool nz = false;
    for (int64_t i = lane; i < nch; i += 64) {
        const uint64_t w = st[i];
        em = max(em, (int)(w >> 40));
        nz |= (w & ((1ull << 40) - 1)) != 0;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) em = max(em, __shfl_xor(em, o, 64));
    if (beta) return 53 + em - beta;
    if (!__any(nz)) return 0;   // a row of zeros
    const int eh = em > 1 ? em : 1;
    uint64_t S = 0;
    for (int64_t i = lane; i < nch; i += 64) S += crt_chunk_term(st[i], eh);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const uint32_t lo = __shfl_xor((uint32_t)S, o, 64), hi = __shfl_xor((uint32_t)(S >> 32), o, 64);
        S += ((uint64_t)hi << 32) | lo;
    }
    return 37 + eh - crt_norm_m(S, K, c);
}

// Residues of 16 consecutive elements v[i] of one row (shift sigma, sigma - e' >= -9) into
// one 16-byte chunk per modulus plane at dst (planes `plane` bytes apart).
// With q = +-(hi 2^31 + lo), the value x = +-(hi (2^31 mod p) + lo) is an exact double
// (|x| < 2^40) congruent to q, and x - p n, n the integer nearest x / p, is its residue in
// [-(p-1)/2, (p-1)/2] for odd p: x / p is never within 1 / (2p) of a half-integer, far
// above the error of the product x fl(1/p), which fma(x, fl(1/p), 1.5 2^52) rounds to an
// integer; the residue plus 1.5 2^52 holds it in two's complement in its low word.  For
// p = 256 (plane 0) the residue is the low byte of q in two's complement.
__device__ __forceinline__ void crt_emit16(const double (&v)[16], int sigma, const CrtConsts &c,
                                           int8_t *__restrict__ dst, int64_t plane) {
    double hi[16], lo[16];
    uint32_t low8[4] = {0u, 0u, 0u, 0u};
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const uint64_t u = (uint64_t)__double_as_longlong(v[i]);
        const int e = (int)((u >> 52) & 0x7FF);
        const uint64_t mx = (u & 0xFFFFFFFFFFFFFull) | (e ? (1ull << 52) : 0ull);
        const int sh = sigma - (e ? e : 1);   // subnormals: exponent 1 - 1023
        // nearest, ties away from zero (mx < 2^53: shifts of 54 and more round to 0)
        const uint64_t Q = sh >= 54 ? 0ull : (sh >= 1 ? (mx + (1ull << (sh - 1))) >> sh : mx << (-sh));
        const bool neg = (u >> 63) != 0;
        const double h = (double)(uint32_t)(Q >> 31), l = (double)(uint32_t)(Q & 0x7FFFFFFFu);
        hi[i] = neg ? -h : h;
        lo[i] = neg ? -l : l;
        const uint32_t qlo = neg ? (uint32_t)(0u - (uint32_t)Q) : (uint32_t)Q;   // low word of q
        low8[i >> 2] |= (qlo & 0xFFu) << (8 * (i & 3));
    }
    *reinterpret_cast<uint4 *>(dst) = make_uint4(low8[0], low8[1], low8[2], low8[3]);   // p = 256
    for (int j = 1; j < c.n; j++) {
        const double p = c.p[j], inv = c.inv[j], c31 = c.c31[j];
        constexpr double kMagic = 6755399441055744.0;   // 1.5 2^52: rounds to an integer
        uint32_t w[4];
#pragma unroll
        for (int gi = 0; gi < 4; gi++) {
            uint32_t b[4];
#pragma unroll
            for (int h = 0; h < 4; h++) {
                const int i = 4 * gi + h;
                const double x = __builtin_fma(hi[i], c31, lo[i]);
                const double qv = __builtin_fma(x, inv, kMagic) - kMagic;
                b[h] = (uint32_t)__double_as_longlong(__builtin_fma(-qv, p, x) + kMagic);   // residue, low word
            }
            w[gi] = __builtin_amdgcn_perm(__builtin_amdgcn_perm(b[3], b[2], 0x0c0c0400u),
                                          __builtin_amdgcn_perm(b[1], b[0], 0x0c0c0400u), 0x05040100u);
        }
        *reinterpret_cast<uint4 *>(dst + (int64_t)j * plane) = make_uint4(w[0], w[1], w[2], w[3]);
    }
}

// Residues of the staged tile: thread (row, q) owns 16 consecutive k of one row and
// writes one 16-byte lane chunk per modulus plane.
template <bool KC>
__device__ __forceinline__ void crt_emit(const double *tile, int64_t R, const int32_t *__restrict__ Srow,
                                         const CrtConsts &c, int8_t *__restrict__ S, int64_t Rb, int64_t Kb,
                                         int64_t rb, int64_t kt) {
    constexpr int TK = 128;
    constexpr int LDW = KC ? (TK + 1) : (32 + 1);
    const int t = threadIdx.x;
    const int64_t r0 = rb * 32;
    const int row = t & 31, q = t >> 5;
    const int64_t kb = kt * 4 + (q >> 1);
    if (kb >= Kb) return;
    const int sigma = (r0 + row < R) ? Srow[row] : 0;
    double v[16];
#pragma unroll
    for (int i = 0; i < 16; i++) v[i] = KC ? tile[row * LDW + 16 * q + i] : tile[(16 * q + i) * LDW + row];
    const int64_t plane = Rb * Kb * (int64_t)kBlkBytes;
    crt_emit16(v, sigma, c, S + (rb * Kb + kb) * (int64_t)kBlkBytes + (q & 1) * 512 + row * 16, plane);
}

// Chunk words of the staged tile: thread (row, q) owns the chunk kt * 8 + q of its row.
template <bool KC>
__device__ __forceinline__ void crt_stats_emit(const double *tile, uint64_t *__restrict__ st, int64_t nch,
                                               int64_t rb, int64_t kt) {
    constexpr int TK = 128;
    constexpr int LDW = KC ? (TK + 1) : (32 + 1);
    const int t = threadIdx.x;
    const int row = t & 31, q = t >> 5;
    double v[16];
#pragma unroll
    for (int i = 0; i < 16; i++) v[i] = KC ? tile[row * LDW + 16 * q + i] : tile[(16 * q + i) * LDW + row];
    st[(rb * 32 + row) * nch + kt * 8 + q] = crt_chunk_stat(v);
}

// The split phase's kernels take both operands: grid z = 0 is A, 1 is B.  Each block's
// operand decides the orientation (uniform over the block), and blocks beyond an operand's
// row blocks return at once.  One launch per step instead of one per operand and step: at
// 2048^3 and below the six launches of a split phase cost as much as the data they move.
__global__ __launch_bounds__(256) void k_crt_stats(CrtSide a, CrtSide b, int64_t nch) {
    const CrtSide o = blockIdx.z ? b : a;
    if ((int64_t)blockIdx.y * kBlk >= o.Rp) return;
    __shared__ double tile[kSplitLds / 8];
    double v[16];
    if (o.op.sk == 1) {
        split_gather<true>(o.op.X, o.op.sr, o.op.sk, o.op.R, o.op.K, blockIdx.y, blockIdx.x, v);
        split_stage<true>(v, tile);
        __syncthreads();
        crt_stats_emit<true>(tile, o.stats, nch, blockIdx.y, blockIdx.x);
    } else {
        split_gather<false>(o.op.X, o.op.sr, o.op.sk, o.op.R, o.op.K, blockIdx.y, blockIdx.x, v);
        split_stage<false>(v, tile);
        __syncthreads();
        crt_stats_emit<false>(tile, o.stats, nch, blockIdx.y, blockIdx.x);
    }
}

hipError_t launch_crt_stats(const CrtSide &a, const CrtSide &b, int64_t Kp, hipStream_t st) {
    const int64_t Rb = std::max(a.Rp, b.Rp) / kBlk, Kb = Kp / kBlk;
    const int64_t nkt = (Kb + 3) / 4;
    if (Rb <= 0 || nkt <= 0) return hipSuccess;
    if (Rb > 65535 || nkt > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_crt_stats, dim3((unsigned)nkt, (unsigned)Rb, 2), dim3(256), 0, st, a, b, crt_chunks(Kb));
    return hipGetLastError();
}

// Wave per row (four rows per workgroup): the Rp rows of A, then the Rp rows of B.
__global__ __launch_bounds__(256) void k_crt_scale(CrtSide a, CrtSide b, int64_t nch, int64_t K, CrtConsts c) {
    const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const bool onb = r >= a.Rp;
    const CrtSide o = onb ? b : a;
    const int64_t rr = onb ? r - a.Rp : r;
    if (rr >= o.Rp) return;   // whole wave
    const int s = crt_row_sigma(o.stats + rr * nch, nch, K, c.beta, c, threadIdx.x & 63);
    if ((threadIdx.x & 63) == 0) o.sigma[rr] = s;
}

hipError_t launch_crt_scale(const CrtSide &a, const CrtSide &b, int64_t nch, int64_t K, const CrtConsts &c,
                            hipStream_t st) {
    const int64_t nb = (a.Rp + b.Rp + 3) / 4;
    if (nb <= 0) return hipSuccess;
    if (nb > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_crt_scale, dim3((unsigned)nb), dim3(256), 0, st, a, b, nch, K, c);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_crt_split(CrtSide a, CrtSide b, int64_t Kb, CrtConsts c) {
    const CrtSide o = blockIdx.z ? b : a;
    const int64_t Rb = o.Rp / kBlk;
    if ((int64_t)blockIdx.y >= Rb) return;
    __shared__ double tile[kSplitLds / 8];
    double v[16];
    if (o.op.sk == 1) {
        split_gather<true>(o.op.X, o.op.sr, o.op.sk, o.op.R, o.op.K, blockIdx.y, blockIdx.x, v);
        split_stage<true>(v, tile);
        __syncthreads();
        crt_emit<true>(tile, o.op.R, o.sigma + (int64_t)blockIdx.y * 32, c, o.S, Rb, Kb, blockIdx.y, blockIdx.x);
    } else {
        split_gather<false>(o.op.X, o.op.sr, o.op.sk, o.op.R, o.op.K, blockIdx.y, blockIdx.x, v);
        split_stage<false>(v, tile);
        __syncthreads();
        crt_emit<false>(tile, o.op.R, o.sigma + (int64_t)blockIdx.y * 32, c, o.S, Rb, Kb, blockIdx.y, blockIdx.x);
    }
}

hipError_t launch_crt_split(const CrtSide &a, const CrtSide &b, const CrtConsts &c, int64_t Kp, hipStream_t st) {
    const int64_t Rb = std::max(a.Rp, b.Rp) / kBlk, Kb = Kp / kBlk;
    const int64_t nkt = (Kb + 3) / 4;
    if (Rb <= 0 || nkt <= 0) return hipSuccess;
    if (Rb > 65535 || nkt > 0x7fffffff) return hipErrorInvalidValue;
    // B first (grid z = 0): the statistics launch read B last, so part of it may still be in
    // the memory-side cache
    hipLaunchKernelGGL(k_crt_split, dim3((unsigned)nkt, (unsigned)Rb, 2), dim3(256), 0, st, b, a, Kb, c);
    return hipGetLastError();
}

// Residues of the wave's (32 WM) x (32 WN) sums, four rows of one column per word of
// Rj[row / 4][col] (row stride ldr words), one signed byte each: r = x - p q with q the
// integer nearest x / p, congruent to x, |r| <= p / 2.  q comes from fma(x, 1/p, 1.5 2^52),
// which rounds the exact product to an integer (|x| <= 2^30, so x fl(1/p) is never within
// 1 / (2p) of a half-integer for odd p, and is exact for p = 256: ties to even, as rint)
// and holds q in two's complement in its low word.  |q| <= 2^30 / 167 < 2^23, so
// r = x - p q is one 24-bit multiply-add in int32; the low bytes of four r are packed with
// v_perm_b32.  CLIP: words whose four rows all lie at or beyond M, or whose column lies at
// or beyond N, are not written.
template <int WM, int WN, bool CLIP>
__device__ __forceinline__ void crt_tile_store(const v16i (&acc)[WM][WN], double p, double inv, uint32_t *Rj,
                                               int64_t ldr, int64_t tm, int64_t tn, int64_t M, int64_t N) {
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = w & 1, wn = w >> 1;
    const int64_t rbase = tm * 64 * WM + wm * 32 * WM, cbase = tn * 64 * WN + wn * 32 * WN;
    const int np = -(int)p;
#pragma unroll
    for (int mi = 0; mi < WM; mi++)
#pragma unroll
        for (int ni = 0; ni < WN; ni++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                constexpr double kMagic = 6755399441055744.0;   // 1.5 2^52
                uint32_t lo[4];
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const int x = acc[mi][ni][4 * q + e];
                    const int qi = (int)(uint32_t)__double_as_longlong(__builtin_fma((double)x, inv, kMagic));
                    lo[e] = (uint32_t)(__mul24(qi, np) + x);
                }
                const uint32_t word = __builtin_amdgcn_perm(__builtin_amdgcn_perm(lo[3], lo[2], 0x0c0c0400u),
                                                            __builtin_amdgcn_perm(lo[1], lo[0], 0x0c0c0400u),
                                                            0x05040100u);
                const int64_t row4 = (rbase + mi * 32 + q * 8 + (lane >> 5) * 4) >> 2;
                const int64_t col = cbase + ni * 32 + (lane & 31);
                if (!CLIP || (row4 * 4 < M && col < N)) Rj[row4 * ldr + col] = word;
            }
}

// One exact int8 GEMM per modulus (blockIdx.y), wave tile (32 WM) x (32 WN).  A contraction
// block gives a wave only WM*WN MFMAs, too few to cover the L2 latency of its fragments,
// so the fragments of the next PF blocks are in flight in a register ring (Kb is a
// multiple of PF; the loads past the end re-read the last block, in bounds).  Epilogue:
// crt_tile_store.

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
    crt_tile_store<WM, WN, false>(acc, c.p[j], c.inv[j], g.R + (int64_t)j * (g.Mp / 4) * g.Np, g.Np, tm, tn, g.Mp,
                                  g.Np);
}


template <int WM, int WN>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(crt_occupancy(WM * WN), crt_occupancy(WM * WN))))
void k_crt_gemm_lds(CrtGemm g, CrtConsts c) {
    __shared__ __attribute__((aligned(1024))) int8_t smem[kCrtStages * (2 * WM + 2 * WN) * kBlkBytes];
    const int j = blockIdx.y;
    const int ntm = (int)(g.Mp / (64 * WM)), ntn = (int)(g.Np / (64 * WN));
    int tm, tn;
    tile_map(blockIdx.x, ntm, ntn, tm, tn);
    const int64_t rowstride = g.Kb * kBlkBytes;
    const int8_t *A0 = g.SA + (int64_t)j * (g.Mp / kBlk) * rowstride + (int64_t)tm * 2 * WM * rowstride;
    const int8_t *B0 = g.SB + (int64_t)j * (g.Np / kBlk) * rowstride + (int64_t)tn * 2 * WN * rowstride;
    v16i acc[WM][WN];
    CrtTileLoop<WM, WN, kCrtStages> L;
    L.setup(A0, B0, rowstride, g.Kb);
    L.prime(smem);
    L.run(smem, acc);
    crt_tile_store<WM, WN, false>(acc, c.p[j], c.inv[j], g.R + (int64_t)j * (g.Mp / 4) * g.Np, g.Np, tm, tn,
                                  g.Mp, g.Np);
}

hipError_t launch_crt_gemm(const CrtGemm &g, const CrtConsts &c, hipStream_t st) {
    const int WM = g.variant == 2 ? 2 : 4, WN = WM;   // workgroup tile 128 x 128 or 256 x 256
    if (g.Mp % (64 * WM) || g.Np % (64 * WN) || g.Kb % kCrtPF || g.Kb <= 0 || c.n < 1 || c.n > kCrtMaxMod)
        return hipErrorInvalidValue;
    const int64_t tiles = (g.Mp / (64 * WM)) * (g.Np / (64 * WN));
    if (tiles <= 0) return hipSuccess;
    if (tiles > 0x7fffffff) return hipErrorInvalidValue;
    const dim3 grid((unsigned)tiles, (unsigned)c.n);
    if (g.variant == 1)
        hipLaunchKernelGGL((k_crt_gemm<4, 4>), grid, dim3(256), 0, st, g, c);
    else if (g.variant == 2)
        hipLaunchKernelGGL((k_crt_gemm_lds<2, 2>), grid, dim3(256), 0, st, g, c);
    else
        hipLaunchKernelGGL((k_crt_gemm_lds<4, 4>), grid, dim3(256), 0, st, g, c);
    return hipGetLastError();
}

// C' * 2^e2 to double, truncated (rounding 0) or nearest-even (1) to 53 bits, then scaled
// by ldexp (as the oracle: one rounding to 53 bits, then the scaling).  |C'| is normalised
// so its top bit is bit 127; the head m (64 bits) carries the value, the tail only a sticky
// bit.  Nearest: the sticky bit OR-ed into bit 0 of m, far below the rounding position, and
// fma(m_hi32, 2^32, m_lo32) rounds the exact m once.  Truncation: the low 11 bits of m
// cleared, leaving 53 bits, which the same fma converts exactly.
__device__ __forceinline__ double crt_to_double(i128 t, int e2, int rounding) {
    const bool neg = t < 0;
    const u128 a = neg ? (u128)0 - (u128)t : (u128)t;
    const uint64_t hi = (uint64_t)(a >> 64), lo = (uint64_t)a;
    const int lz = hi ? __builtin_clzll(hi) : (lo ? 64 + __builtin_clzll(lo) : 127);
    const u128 sa = a << lz;
    const uint64_t m = (uint64_t)(sa >> 64), rest = (uint64_t)sa;
    const uint64_t ms = rounding ? (m | (uint64_t)(rest != 0)) : (m & ~(uint64_t)0x7FF);
    const double d = __builtin_fma((double)(uint32_t)(ms >> 32), 4294967296.0, (double)(uint32_t)ms);
    const double v = ldexp(d, 64 - lz + e2);
    return neg ? -v : v;
}

// C' * 2^e2 for a signed 192-bit C' (three words, low first), truncated (rounding 0) or
// nearest-even (1) to 53 bits.  Values below 2^127 take the 128-bit path directly; larger
// ones are shifted right to 127 bits with the shifted-out bits OR-ed into bit 0 (a sticky
// bit far below the rounding position, so both roundings are unchanged).
__device__ __forceinline__ double crt_wide_to_double(uint64_t w0, uint64_t w1, uint64_t w2, int e2, int rounding) {
    const bool neg = (int64_t)w2 < 0;
    if (neg) {   // two's complement negation
        w0 = ~w0; w1 = ~w1; w2 = ~w2;
        w0 += 1;
        const uint64_t c0 = w0 == 0;
        w1 += c0;
        w2 += (c0 & (w1 == 0));
    }
    u128 a;
    int sh = 0;
    if (w2 == 0 && (w1 >> 63) == 0) {
        a = ((u128)w1 << 64) | w0;
    } else {
        const int nb = w2 ? 192 - __builtin_clzll(w2) : 128;
        sh = nb - 127;   // 1 .. 65
        bool sticky;
        if (sh < 64) {
            a = ((((u128)w2 << 64) | w1) << (64 - sh)) | (u128)(w0 >> sh);
            sticky = (w0 << (64 - sh)) != 0;
        } else if (sh == 64) {
            a = ((u128)w2 << 64) | w1;
            sticky = w0 != 0;
        } else {
            a = ((u128)w2 << 63) | (u128)(w1 >> 1);
            sticky = w0 != 0 || (w1 & 1) != 0;
        }
        a |= (u128)sticky;
    }
    const i128 v = neg ? -(i128)a : (i128)a;
    return crt_to_double(v, e2 + sh, rounding);
}

// The four rows 4 row4 + e (e < rows) of one column from the n residue words wd[j], each
// byte already flipped by 0x80: the stored signed residue r (congruent to the residue
// product g, -128 <= r <= 127) read as the unsigned byte u = r + 128.  Since the weights
// satisfy sum_j W_j = 1 (mod M), S = sum_j u_j W_j = C' + 128 + K M for an integer K.
// The words are transposed so that one word holds the bytes of four moduli of one row;
// then byte dot products give the byte limbs L_i = sum_j u_j w_ji (below 2^22) of S, and
// K = floor((t + 2^15) / 2^16) with t = sum_j u_j q_j, q_j = floor(2^16 fl(W_j / M)):
// t / 2^16 is within 0.0625 of S / M, which is within 1/4 + 128 / M of K.  Then
// C' = sum_i (L_i + K m_i) 256^i - 128, m_i the bytes of 2^(8 NL) - M, modulo 2^128
// (16 limbs) for n <= 16 moduli (|C'| <= M / 4 < 2^124) and 2^192 (24 limbs, WIDE) above,
// read as signed.  Nearest rounding uses the correctly rounded 128-bit integer conversion
// (then the same power-of-two scaling as the truncating path).
template <bool WIDE>
__device__ __forceinline__ void crt_rebuild_quad(const uint32_t (&wd)[kCrtMaxMod], const CrtConsts &c, int rows,
                                                 const int32_t *__restrict__ Er4, int ec, int rounding,
                                                 double (&out)[4]) {
    constexpr int NG = (WIDE ? kCrtMaxMod : kCrtNarrowMod) / 4;   // groups of four moduli
    constexpr int NL = WIDE ? kCrtLimbs : 16;                      // byte limbs of C'
    constexpr int NW = WIDE ? 20 : 16;                             // byte limbs of W_p (W_p < M < 2^156)
    uint32_t x[NG][4];   // x[G][e]: bytes u_{4G + b} of row e, b = 0..3
#pragma unroll
    for (int G = 0; G < NG; G++) {
        // 4 x 4 byte transpose (v_perm_b32: selector byte k < 4 picks byte k of the second
        // operand, 4 <= k < 8 byte k - 4 of the first)
        const uint32_t t0 = __builtin_amdgcn_perm(wd[4 * G + 1], wd[4 * G], 0x05010400u);
        const uint32_t t1 = __builtin_amdgcn_perm(wd[4 * G + 1], wd[4 * G], 0x07030602u);
        const uint32_t t2 = __builtin_amdgcn_perm(wd[4 * G + 3], wd[4 * G + 2], 0x05010400u);
        const uint32_t t3 = __builtin_amdgcn_perm(wd[4 * G + 3], wd[4 * G + 2], 0x07030602u);
        x[G][0] = __builtin_amdgcn_perm(t2, t0, 0x05040100u);
        x[G][1] = __builtin_amdgcn_perm(t2, t0, 0x07060302u);
        x[G][2] = __builtin_amdgcn_perm(t3, t1, 0x05040100u);
        x[G][3] = __builtin_amdgcn_perm(t3, t1, 0x07060302u);
    }
    uint32_t k[4], L[4][NL];
#pragma unroll
    for (int e = 0; e < 4; e++) {
        uint32_t tlo = 0, thi = 0;
#pragma unroll
        for (int G = 0; G < NG; G++) {
            tlo = __builtin_amdgcn_udot4(x[G][e], c.Qd[0][G], tlo, false);
            thi = __builtin_amdgcn_udot4(x[G][e], c.Qd[1][G], thi, false);
        }
        k[e] = (tlo + (thi << 8) + 32768u) >> 16;
    }
    // limb-outer order: each constant serves the four rows in turn
#pragma unroll
    for (int i = 0; i < NL; i++) {
#pragma unroll
        for (int e = 0; e < 4; e++) L[e][i] = __umul24(k[e], c.mneg8[i]);   // k < 2^13, byte
        if (i < NW) {
#pragma unroll
            for (int G = 0; G < NG; G++) {
                const uint32_t wdi = c.Wd[G][i];
#pragma unroll
                for (int e = 0; e < 4; e++) L[e][i] = __builtin_amdgcn_udot4(x[G][e], wdi, L[e][i], false);
            }
        }
    }
#pragma unroll
    for (int e = 0; e < 4; e++) {
        if (e >= rows) break;
        uint64_t Q[NL / 4];
#pragma unroll
        for (int h = 0; h < NL / 4; h++) {
            const uint32_t lo = L[e][4 * h] + (L[e][4 * h + 1] << 8);        // below 2^31
            const uint32_t hi = L[e][4 * h + 2] + (L[e][4 * h + 3] << 8);
            Q[h] = (uint64_t)lo + ((uint64_t)hi << 16);                      // below 2^48
        }
        const int e2 = Er4[e] + ec - 2150;   // shifts sigma_r + sigma_c
        if (!WIDE) {
            const u128 S = (u128)Q[0] + ((u128)Q[1] << 32) + ((u128)Q[2] << 64) + ((u128)Q[3] << 96) - (u128)128;
            out[e] = crt_to_double((i128)S, e2, rounding);
        } else {
            const u128 s0 = (u128)Q[0] + ((u128)Q[1] << 32);
            const u128 s1 = (s0 >> 64) + Q[2] + ((u128)Q[3] << 32);
            const u128 s2 = (s1 >> 64) + Q[4] + ((u128)Q[5 % (NL / 4)] << 32);
            uint64_t w0 = (uint64_t)s0, w1 = (uint64_t)s1, w2 = (uint64_t)s2;
            const uint64_t b0 = w0 < 128u, b1 = b0 & (uint64_t)(w1 == 0);   // - 128 with borrows
            w0 -= 128u;
            w1 -= b0;
            w2 -= b1;
            out[e] = crt_wide_to_double(w0, w1, w2, e2, rounding);
        }
    }
}

// Four rows 4 row4 + e (e < rows) of the four consecutive columns col0 + cc (col0 + cc < N):
// one 16-byte load per modulus plane (planes `plane` words apart), then crt_rebuild_quad
// per column, and the 4 x 4 results stored by row, 32 bytes per row in two 16-byte stores
// where the row's four columns exist and are 16-byte aligned (a wave then writes whole
// cache lines at once; stores of one column each left lines partly written in L2, and
// those reached HBM as partial writes).  Row quad row4 of the planes starts at Rq; Er4 /
// Cq at the quad's first row.
template <bool WIDE>
__device__ __forceinline__ void crt_rebuild_cols(const uint32_t *__restrict__ Rq, int64_t plane, int64_t col0,
                                                 int64_t N, int rows, const CrtConsts &c,
                                                 const int32_t *__restrict__ Er4, const int32_t *__restrict__ Ec,
                                                 int rounding, double *__restrict__ Cq, int64_t ldc) {
    uint4 w4[kCrtMaxMod];
#pragma unroll
    for (int j = 0; j < kCrtMaxMod; j++)
        w4[j] = j < c.n ? *reinterpret_cast<const uint4 *>(Rq + (int64_t)j * plane + col0)
                        : make_uint4(0x80808080u, 0x80808080u, 0x80808080u, 0x80808080u);
    double out[4][4];   // [column][row]
#pragma unroll
    for (int cc = 0; cc < 4; cc++) {
        if (col0 + cc >= N) break;
        uint32_t wd[kCrtMaxMod];
#pragma unroll
        for (int j = 0; j < kCrtMaxMod; j++) {
            const uint32_t v = cc == 0 ? w4[j].x : (cc == 1 ? w4[j].y : (cc == 2 ? w4[j].z : w4[j].w));
            wd[j] = v ^ 0x80808080u;   // planes j >= n: zero words
        }
        crt_rebuild_quad<WIDE>(wd, c, rows, Er4, Ec[col0 + cc], rounding, out[cc]);
    }
    const bool full = col0 + 3 < N;
#pragma unroll
    for (int e = 0; e < 4; e++) {
        if (e >= rows) break;
        double *dst = Cq + e * ldc + col0;
        if (full && (reinterpret_cast<uintptr_t>(dst) & 15) == 0) {
            *reinterpret_cast<double2 *>(dst) = make_double2(out[0][e], out[1][e]);
            *reinterpret_cast<double2 *>(dst + 2) = make_double2(out[2][e], out[3][e]);
        } else {
#pragma unroll
            for (int cc = 0; cc < 4; cc++)
                if (col0 + cc < N) dst[cc] = out[cc][e];
        }
    }
}

// Thread per (4 rows, 4 consecutive columns) of R[j][Mp / 4][Np] (Np is a multiple of 256,
// so the four words are in the row).
template <bool WIDE>
__global__ __launch_bounds__(256) void k_crt_rebuild(const uint32_t *__restrict__ R, int64_t Mp, int64_t Np,
                                                     int64_t M, int64_t N, const int32_t *__restrict__ Er,
                                                     const int32_t *__restrict__ Ec, double *__restrict__ C,
                                                     int64_t ldc, int rounding, int64_t ncb, CrtConsts c) {
    const int64_t row4 = blockIdx.x / ncb;
    const int64_t col0 = (blockIdx.x % ncb) * 1024 + 4 * (int64_t)threadIdx.x;
    if (col0 >= N) return;
    const int64_t rest = M - row4 * 4;
    crt_rebuild_cols<WIDE>(R + row4 * Np, (Mp / 4) * Np, col0, N, rest < 4 ? (int)rest : 4, c, Er + row4 * 4, Ec,
                           rounding, C + row4 * 4 * ldc, ldc);
}

hipError_t launch_crt_rebuild(const uint32_t *R, int64_t Mp, int64_t Np, int64_t M, int64_t N, const int32_t *Er,
                              const int32_t *Ec, double *C, int64_t ldc, int rounding, const CrtConsts &c,
                              hipStream_t st) {
    const int64_t ncb = (N + 1023) / 1024;
    const int64_t blocks = ((M + 3) / 4) * ncb;
    if (blocks <= 0) return hipSuccess;
    if (blocks > 0x7fffffff || Np % 4) return hipErrorInvalidValue;
    if (c.n > kCrtNarrowMod)
        hipLaunchKernelGGL(k_crt_rebuild<true>, dim3((unsigned)blocks), dim3(256), 0, st, R, Mp, Np, M, N, Er, Ec, C,
                           ldc, rounding, ncb, c);
    else
        hipLaunchKernelGGL(k_crt_rebuild<false>, dim3((unsigned)blocks), dim3(256), 0, st, R, Mp, Np, M, N, Er, Ec, C,
                           ldc, rounding, ncb, c);
    return hipGetLastError();
}

// ------------------------------------------------------------------ grouped CRT
// Many independent products in one launch per phase: chunk statistics and shifts of
// every operand (each problem's rule), residues, one modulus GEMM launch per tile shape
// over items {problem, tile row, tile col} x modulus, and the rebuild over items
// {problem, first element}.

__global__ __launch_bounds__(256) void k_grp_crt_stats(const GroupOperand *__restrict__ ops,
                                                       const int4 *__restrict__ items) {
    __shared__ double tile[kSplitLds / 8];
    const int4 it = items[blockIdx.x];
    const GroupOperand &o = ops[it.x];
    double v[16];
    if (o.sk == 1) {
        split_gather<true>(o.X, o.sr, o.sk, o.R, o.K, it.y, it.z, v);
        split_stage<true>(v, tile);
        __syncthreads();
        crt_stats_emit<true>(tile, o.st, crt_chunks(o.Kb), it.y, it.z);
    } else {
        split_gather<false>(o.X, o.sr, o.sk, o.R, o.K, it.y, it.z, v);
        split_stage<false>(v, tile);
        __syncthreads();
        crt_stats_emit<false>(tile, o.st, crt_chunks(o.Kb), it.y, it.z);
    }
}

hipError_t launch_group_crt_stats(const GroupOperand *ops, const int4 *items, int n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_grp_crt_stats, dim3((unsigned)n), dim3(256), 0, st, ops, items);
    return hipGetLastError();
}

// Items {operand, row block}: wave w takes rows 8 w .. 8 w + 7 of the block.
__global__ __launch_bounds__(256) void k_grp_crt_scale(const GroupOperand *__restrict__ ops,
                                                       const CrtGroupProb *__restrict__ probs,
                                                       const int4 *__restrict__ items, CrtConsts c) {
    const int4 it = items[blockIdx.x];
    const GroupOperand &o = ops[it.x];
    const int beta = probs[it.x >> 1].beta;
    const int64_t nch = crt_chunks(o.Kb);
    const int lane = threadIdx.x & 63;
    for (int i = 0; i < 8; i++) {
        const int64_t r = (int64_t)it.y * 32 + (threadIdx.x >> 6) * 8 + i;
        const int s = crt_row_sigma(o.st + r * nch, nch, o.K, beta, c, lane);
        if (lane == 0) o.E[r] = s;
    }
}

hipError_t launch_group_crt_scale(const GroupOperand *ops, const CrtGroupProb *probs, const int4 *items, int n,
                                  const CrtConsts &c, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_grp_crt_scale, dim3((unsigned)n), dim3(256), 0, st, ops, probs, items, c);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_grp_crt_split(const GroupOperand *__restrict__ ops,
                                                       const int4 *__restrict__ items, CrtConsts c) {
    __shared__ double tile[kSplitLds / 8];
    const int4 it = items[blockIdx.x];
    const GroupOperand &o = ops[it.x];
    double v[16];
    if (o.sk == 1) {
        split_gather<true>(o.X, o.sr, o.sk, o.R, o.K, it.y, it.z, v);
        split_stage<true>(v, tile);
        __syncthreads();
        crt_emit<true>(tile, o.R, o.E + (int64_t)it.y * 32, c, o.S, o.Rb, o.Kb, it.y, it.z);
    } else {
        split_gather<false>(o.X, o.sr, o.sk, o.R, o.K, it.y, it.z, v);
        split_stage<false>(v, tile);
        __syncthreads();
        crt_emit<false>(tile, o.R, o.E + (int64_t)it.y * 32, c, o.S, o.Rb, o.Kb, it.y, it.z);
    }
}

hipError_t launch_group_crt_split(const GroupOperand *ops, const CrtGroupProb *probs, const int4 *items, int n,
                                  const CrtConsts &c, hipStream_t st) {
    (void)probs;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_grp_crt_split, dim3((unsigned)n), dim3(256), 0, st, ops, items, c);
    return hipGetLastError();
}

// One workgroup per tile item and range of moduli [j0, j1) (grid y): the contraction loop
// of each modulus in turn, the next modulus's first blocks issued before the epilogue of
// the current one, so those loads are in flight during the residue stores.
// Items {-1, ...} pad the host's per-XCD lists (the hardware deals blockIdx x to XCD x % 8)
// and return at once.
template <int WM, int WN>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(crt_occupancy(WM * WN), crt_occupancy(WM * WN))))
void k_grp_crt_gemm(const CrtGroupProb *__restrict__ probs, const int4 *__restrict__ items, int mpw, CrtConsts c) {
    __shared__ __attribute__((aligned(1024))) int8_t smem[kCrtStages * (2 * WM + 2 * WN) * kBlkBytes];
    const int4 it = items[blockIdx.x];
    if (it.x < 0) return;
    const int j0 = blockIdx.y * mpw;
    const int j1 = j0 + mpw < c.n ? j0 + mpw : c.n;
    const CrtGroupProb &pr = probs[it.x];
    const int64_t rowstride = pr.Kb * kBlkBytes;
    const int64_t planeA = pr.RbA * rowstride, planeB = pr.RbB * rowstride;
    const int64_t planeR = pr.RbA * (kBlk / 4) * pr.Np;
    CrtTileLoop<WM, WN, kCrtStages> L;
    L.setup(pr.SA + (int64_t)j0 * planeA + (int64_t)it.y * 2 * WM * rowstride,
            pr.SB + (int64_t)j0 * planeB + (int64_t)it.z * 2 * WN * rowstride, rowstride, pr.Kb);
    L.prime(smem);
    for (int j = j0; j < j1; j++) {
        v16i acc[WM][WN];
        L.run(smem, acc);
        if (j + 1 < j1) {   // the next modulus's first blocks load during the epilogue
            L.advance(planeA, planeB);
            L.prime(smem);
        }
        crt_tile_store<WM, WN, true>(acc, c.p[j], c.inv[j], pr.R + (int64_t)j * planeR, pr.Np, it.y, it.z, pr.M,
                                     pr.N);
    }
}

int crt_group_cfg_wm(int cfg) { return cfg == CCFG_44 ? 4 : (cfg == CCFG_22 ? 2 : 1); }
int crt_group_cfg_wn(int cfg) { return crt_group_cfg_wm(cfg); }

hipError_t launch_group_crt_gemm(const CrtGroupProb *probs, const int4 *items, int n, int cfg, int mpw,
                                 const CrtConsts &c, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (c.n < 1 || c.n > kCrtMaxMod || mpw < 1) return hipErrorInvalidValue;
    const dim3 grid((unsigned)n, (unsigned)((c.n + mpw - 1) / mpw));
    switch (cfg) {
        case CCFG_44: hipLaunchKernelGGL((k_grp_crt_gemm<4, 4>), grid, dim3(256), 0, st, probs, items, mpw, c); break;
        case CCFG_22: hipLaunchKernelGGL((k_grp_crt_gemm<2, 2>), grid, dim3(256), 0, st, probs, items, mpw, c); break;
        case CCFG_11: hipLaunchKernelGGL((k_grp_crt_gemm<1, 1>), grid, dim3(256), 0, st, probs, items, mpw, c); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Items {problem, first group}: thread t handles group base + t of the problem's
// ceil(M / 4) x ceil(N / 4) grid of (row quad, four consecutive columns) (the padded row
// stride Np is a multiple of 64 words).
template <bool WIDE>
__global__ __launch_bounds__(256) void k_grp_crt_rebuild(const CrtGroupProb *__restrict__ probs,
                                                         const int4 *__restrict__ items, int rounding, CrtConsts c) {
    const int4 it = items[blockIdx.x];
    const CrtGroupProb &pr = probs[it.x];
    const uint32_t nc4 = (uint32_t)(pr.N + 3) / 4;
    const uint32_t idx = (uint32_t)it.y + threadIdx.x;
    if (idx >= (uint32_t)((pr.M + 3) / 4) * nc4) return;
    const uint32_t row4 = idx / nc4;
    const int64_t col0 = 4 * (int64_t)(idx - row4 * nc4);
    const int rest = pr.M - 4 * (int)row4;
    crt_rebuild_cols<WIDE>(pr.R + (int64_t)row4 * pr.Np, pr.RbA * (kBlk / 4) * pr.Np, col0, pr.N, rest < 4 ? rest : 4,
                           c, pr.Er + 4 * (int64_t)row4, pr.Ec, rounding,
                           pr.C + 4 * (int64_t)row4 * pr.ldc, pr.ldc);
}

hipError_t launch_group_crt_rebuild(const CrtGroupProb *probs, const int4 *items, int n, int rounding,
                                    const CrtConsts &c, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (c.n > kCrtNarrowMod)
        hipLaunchKernelGGL(k_grp_crt_rebuild<true>, dim3((unsigned)n), dim3(256), 0, st, probs, items, rounding, c);
    else
        hipLaunchKernelGGL(k_grp_crt_rebuild<false>, dim3((unsigned)n), dim3(256), 0, st, probs, items, rounding, c);
    return hipGetLastError();
}

// ------------------------------------------------------------------ persistent CRT group
constexpr int kCrtPersistLds = 4 * 32 * kWaveTileLd * 8;   // phase A's per-wave tiles
static_assert(kCrtPersistLds >= 64 * 1024, "phase B stages");
static_assert(2 * kCrtPersistLds <= 160 * 1024, "two workgroups per CU");
constexpr int kCrtPersistPerCu = 2;   // workgroups per CU of the persistent CRT launch

// Phase B items [i0, i1) of one workgroup, all of one tile shape: every modulus of each
// tile in turn (the next modulus's first blocks in flight during each epilogue).  At one
// wave per SIMD the smaller tiles have few MFMAs per contraction block (4 at 128 x 128, 1
// at 64 x 64), so their loops keep more blocks in flight: the stages of 64 KiB of LDS.
template <int WM, int WN>
constexpr int crt_persist_stages() { return (64 * 1024) / ((2 * WM + 2 * WN) * kBlkBytes); }

template <int WM, int WN>
__device__ __forceinline__ void crt_persist_tiles(const CrtGroupProb *__restrict__ probs,
                                                  const int4 *__restrict__ items, int i0, int i1,
                                                  const CrtConsts &c, int8_t *smem) {
    for (int i = i0; i < i1; i++) {
        const int4 it = items[i];
        const CrtGroupProb &pr = probs[it.x];
        const int64_t rowstride = pr.Kb * kBlkBytes;
        const int64_t planeA = pr.RbA * rowstride, planeB = pr.RbB * rowstride;
        const int64_t planeR = pr.RbA * (kBlk / 4) * pr.Np;
        CrtTileLoop<WM, WN, crt_persist_stages<WM, WN>()> L;
        L.setup(pr.SA + (int64_t)it.y * 2 * WM * rowstride, pr.SB + (int64_t)it.z * 2 * WN * rowstride, rowstride,
                pr.Kb);
        L.prime(smem);
        for (int j = 0; j < c.n; j++) {
            v16i acc[WM][WN];
            L.run(smem, acc);
            if (j + 1 < c.n) {
                L.advance(planeA, planeB);
                L.prime(smem);
            }
            crt_tile_store<WM, WN, true>(acc, c.p[j], c.inv[j], pr.R + (int64_t)j * planeR, pr.Np, it.y, it.z, pr.M,
                                         pr.N);
        }
    }
}

// One row block of the persistent launch's phase A: the row maxima (one pass over the
// block), for the norm rule the chunk terms of every row (a second pass, each lane's
// chunks as the split kernel's, the two halves of a row summed), the shifts into o.E, and
// the residues (a third pass).
template <bool KC>
__device__ __forceinline__ void crt_wave_rowblock(const GroupOperand &o, int64_t rb, int beta, const CrtConsts &c,
                                                  double *wt, int lane) {
    const int emax = wave_rowmax<KC>(o, rb, wt, lane);
    int sigma;
    if (beta) {
        sigma = 53 + emax - beta;
    } else {
        const int eh = emax > 1 ? emax : 1;
        uint64_t S = 0;
        wave_tiles<KC>(o, rb, wt, lane, [&](const double *w_, int row, int c0, int64_t) {
            double v[16];
#pragma unroll
            for (int e = 0; e < 16; e++) v[e] = w_[row * kWaveTileLd + c0 + e];
            S += crt_chunk_term(crt_chunk_stat(v), eh);
        });
        const uint32_t lo = __shfl_xor((uint32_t)S, 32, 64), hi = __shfl_xor((uint32_t)(S >> 32), 32, 64);
        S += ((uint64_t)hi << 32) | lo;
        sigma = S ? 37 + eh - crt_norm_m(S, o.K, c) : 0;
    }
    if (lane < 32) o.E[rb * 32 + lane] = sigma;
    const int64_t plane = o.Rb * o.Kb * (int64_t)kBlkBytes;
    wave_tiles<KC>(o, rb, wt, lane, [&](const double *w_, int row, int c0, int64_t kb) {
        double v[16];
#pragma unroll
        for (int e = 0; e < 16; e++) v[e] = w_[row * kWaveTileLd + c0 + e];
        crt_emit16(v, sigma, c, o.S + (rb * o.Kb + kb) * (int64_t)kBlkBytes + lane * 16, plane);
    });
}

// Two workgroups per CU (two waves per SIMD, 256 registers each): the memory-bound
// residue and rebuild phases and the small tiles' contraction loops need the second wave
// to hide latency; the 256 x 256 tile (256 accumulator registers) does not fit and runs as
// 128 x 128 tiles here (the host maps it).  2 x kCrtPersistLds of LDS fit one CU.
template <bool WIDE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
void k_grp_crt_persistent(CrtPersistLaunch L, CrtConsts c) {
    __shared__ __attribute__((aligned(1024))) int8_t smem[kCrtPersistLds];
    __shared__ int flag_lds;
    const int G = gridDim.x;
    PersistCtl *ctl = reinterpret_cast<PersistCtl *>(L.ctl);

    // ---- phase A: shifts + residues, one 32-row block per item, per-wave lists
    {
        const int lane = threadIdx.x & 63;
        const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        double *wt = reinterpret_cast<double *>(smem) + w * (32 * kWaveTileLd);
        const int slot = blockIdx.x * 4 + w;
        for (int i = L.offA[slot]; i < L.offA[slot + 1]; i++) {
            const int4 it = L.itemsA[i];
            const GroupOperand &o = L.ops[it.x];
            if (o.sk == 1) crt_wave_rowblock<true>(o, it.y, L.probs[it.x >> 1].beta, c, wt, lane);
            else crt_wave_rowblock<false>(o, it.y, L.probs[it.x >> 1].beta, c, wt, lane);
        }
    }
    bool ok = grid_barrier(ctl, G, &flag_lds);

    // ---- phase B: modulus-GEMM tiles in runs of one shape
    if (ok) {
        const int i1 = L.offB[blockIdx.x + 1];
        for (int i = L.offB[blockIdx.x]; i < i1;) {
            const int cfg = L.itemsB[i].w;
            int j = i + 1;
            while (j < i1 && L.itemsB[j].w == cfg) j++;
            switch (cfg) {
                case CCFG_22: crt_persist_tiles<2, 2>(L.probs, L.itemsB, i, j, c, smem); break;
                case CCFG_11: crt_persist_tiles<1, 1>(L.probs, L.itemsB, i, j, c, smem); break;
                default: break;   // the host builds only the shapes above
            }
            __syncthreads();   // the next shape lays out its LDS stages differently
            i = j;
        }
        ok = grid_barrier(ctl, 2 * G, &flag_lds);
    }

    // ---- phase C: rebuild, 256 (row quad, four columns) groups per item
    if (ok) {
        for (int i = L.offC[blockIdx.x]; i < L.offC[blockIdx.x + 1]; i++) {
            const int4 it = L.itemsC[i];
            const CrtGroupProb &pr = L.probs[it.x];
            const uint32_t nc4 = (uint32_t)(pr.N + 3) / 4;
            const uint32_t idx = (uint32_t)it.y + threadIdx.x;
            if (idx < (uint32_t)((pr.M + 3) / 4) * nc4) {
                const uint32_t row4 = idx / nc4;
                const int64_t col0 = 4 * (int64_t)(idx - row4 * nc4);
                const int rest = pr.M - 4 * (int)row4;
                crt_rebuild_cols<WIDE>(pr.R + (int64_t)row4 * pr.Np, pr.RbA * (kBlk / 4) * pr.Np, col0, pr.N,
                                       rest < 4 ? rest : 4, c, pr.Er + 4 * (int64_t)row4, pr.Ec, L.rounding,
                                       pr.C + 4 * (int64_t)row4 * pr.ldc, pr.ldc);
            }
        }
    }
    persist_exit(ctl, G);
}

int crt_persist_per_cu() { return kCrtPersistPerCu; }

hipError_t launch_group_crt_persistent(const CrtPersistLaunch &L, const CrtConsts &c, hipStream_t st) {
    if (L.grid <= 0) return hipSuccess;
    if (c.n > kCrtNarrowMod)
        hipLaunchKernelGGL(k_grp_crt_persistent<true>, dim3((unsigned)L.grid), dim3(256), 0, st, L, c);
    else
        hipLaunchKernelGGL(k_grp_crt_persistent<false>, dim3((unsigned)L.grid), dim3(256), 0, st, L, c);
    return hipGetLastError();
}

}  // namespace oz
