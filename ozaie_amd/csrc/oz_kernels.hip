// oz_kernels.hip -- the CDNA4 (gfx950) kernels of the Ozaki-scheme DGEMM.
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
#include "oz_internal.h"

namespace oz {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef __int128 i128;
typedef unsigned __int128 u128;
typedef __attribute__((address_space(3))) void lds_void;

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

__device__ __forceinline__ int field_of(double x) {
    const uint32_t hi = (uint32_t)__double2hiint(x);
    return (int)((hi << 1) >> 21);
}

// ------------------------------------------------------------------ max exponent
// Contraction axis contiguous (sk == 1): one wave per row, coalesced along k.
__device__ __forceinline__ void maxexp_kc_row(const double *__restrict__ X, int64_t sr, int64_t R, int64_t K,
                                              int32_t *__restrict__ E, int64_t r, int lane) {
    int m = 0;
    if (r < R) {
        const double *row = X + r * sr;
        int64_t k = lane;
        for (; k + 192 < K; k += 256) {
            const double a0 = row[k], a1 = row[k + 64], a2 = row[k + 128], a3 = row[k + 192];
            m = max(m, max(max(field_of(a0), field_of(a1)), max(field_of(a2), field_of(a3))));
        }
        for (; k < K; k += 64) m = max(m, field_of(row[k]));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off, 64));
    if (lane == 0) E[r] = m;
}

__global__ __launch_bounds__(256) void k_maxexp_kc(const double *__restrict__ X, int64_t sr,
                                                   int64_t R, int64_t K, int64_t Rp,
                                                   int32_t *__restrict__ E) {
    const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= Rp) return;  // whole wave leaves together
    maxexp_kc_row(X, sr, R, K, E, r, threadIdx.x & 63);
}

// Row index contiguous (sr == 1): thread per row, a slab of k per block row, combined
// with atomicMax into E (zero-filled beforehand).
__device__ __forceinline__ void maxexp_rc_slab(const double *__restrict__ X, int64_t sk, int64_t R, int64_t K,
                                               int32_t *__restrict__ E, int64_t r, int64_t k0, int64_t kper) {
    if (r >= R) return;
    const int64_t k1 = min(K, k0 + kper);
    int m = 0;
    int64_t k = k0;
    for (; k + 3 < k1; k += 4) {
        const double a0 = X[r + k * sk], a1 = X[r + (k + 1) * sk];
        const double a2 = X[r + (k + 2) * sk], a3 = X[r + (k + 3) * sk];
        m = max(m, max(max(field_of(a0), field_of(a1)), max(field_of(a2), field_of(a3))));
    }
    for (; k < k1; k++) m = max(m, field_of(X[r + k * sk]));
    if (m > 0) atomicMax(E + r, m);
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

// ------------------------------------------------------------------ split
// One block handles 32 rows x 128 contraction indices (four slice blocks).  The FP64
// tile is staged through LDS so that the global reads are coalesced along whichever
// axis is contiguous; each thread then owns 16 consecutive k of one row and writes one
// 16-byte lane chunk per slice plane (a wave writes one whole 1 KiB block per plane).
constexpr int kSplitLds = 128 * 33 * 8;  // bytes of the FP64 staging tile (either orientation)

// Digits of 16 elements (mantissas mx with the implicit bit, bias = row exponent minus
// element exponent, sign bits negm) into one 16-byte chunk per slice plane at dst.
// With b*d <= 64 the whole digit string Q = floor(M_x 2^(bd - 53 - bias)) fits 64 bits and
// digit j is the b-bit field of Q at bit b(d-1-j): one bit-field extract per digit.  Four
// digits share a word; the negative elements' bytes are negated without carries between
// bytes (every digit is below 128, so 0x80 - x fits a byte: -x = (0x80 - x) ^ 0x80).
// Wider strings shift the mantissa per digit.
__device__ __forceinline__ void emit_digits16(uint64_t (&mx)[16], const int (&bias)[16], uint32_t negm, int b,
                                              int d, int8_t *__restrict__ dst, int64_t plane) {
    uint32_t nbm[4];
#pragma unroll
    for (int gi = 0; gi < 4; gi++) {
        uint32_t m = 0;
#pragma unroll
        for (int h = 0; h < 4; h++) m |= ((negm >> (4 * gi + h)) & 1u) ? (0xFFu << (8 * h)) : 0u;
        nbm[gi] = m;
    }
    if (b * d <= 64) {
        uint32_t qlo[16], qhi[16];
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int sh = 53 + bias[i] - b * d;   // Q = mx >> sh (mx << -sh below zero)
            const uint64_t q = sh >= 64 ? 0ull : (sh >= 0 ? mx[i] >> sh : mx[i] << (-sh));
            qlo[i] = (uint32_t)q;
            qhi[i] = (uint32_t)(q >> 32);
        }
        const uint32_t mask = (1u << b) - 1u;
        // one plane: x(i) is element i's digit; negative elements' bytes negated
        auto plane_out = [&](int j, auto x) {
            uint32_t w[4];
#pragma unroll
            for (int gi = 0; gi < 4; gi++) {
                const uint32_t p = x(4 * gi) | (x(4 * gi + 1) << 8) | (x(4 * gi + 2) << 16) | (x(4 * gi + 3) << 24);
                const uint32_t pn = (0x80808080u - p) ^ 0x80808080u;
                w[gi] = (pn & nbm[gi]) | (p & ~nbm[gi]);
            }
            *reinterpret_cast<uint4 *>(dst + (int64_t)j * plane) = make_uint4(w[0], w[1], w[2], w[3]);
        };
        int j = 0;
        for (; j < d && b * (d - 1 - j) >= 32; j++) {   // field inside the high word
            const uint32_t off = (uint32_t)(b * (d - 1 - j) - 32);
            plane_out(j, [&](int i) { return __builtin_amdgcn_ubfe(qhi[i], off, (uint32_t)b); });
        }
        for (; j < d; j++) {                              // field at bit < 32 (may straddle)
            const uint32_t sj = (uint32_t)(b * (d - 1 - j));
            plane_out(j, [&](int i) { return __builtin_amdgcn_alignbit(qhi[i], qlo[i], sj) & mask; });
        }
        return;
    }
    const uint64_t mask = (1ull << b) - 1ull;
    for (int j = 0; j < d; j++) {
        uint32_t w[4];
#pragma unroll
        for (int gi = 0; gi < 4; gi++) {
            uint32_t acc = 0;
#pragma unroll
            for (int h = 0; h < 4; h++) {
                const int i = 4 * gi + h;
                const int sh = 53 + bias[i] - b * (j + 1);
                uint64_t v;
                if (sh >= 64) v = 0;
                else if (sh >= 0) v = mx[i] >> sh;
                else v = (sh <= -64) ? 0ull : (mx[i] << (-sh));
                int dig = (int)(v & mask);
                dig = ((negm >> i) & 1u) ? -dig : dig;
                acc |= (uint32_t)(dig & 0xFF) << (8 * h);
            }
            w[gi] = acc;
        }
        *reinterpret_cast<uint4 *>(dst + (int64_t)j * plane) = make_uint4(w[0], w[1], w[2], w[3]);
    }
}

// The three steps of a tile of 32 rows x TK contraction indices (TK = 128 or 256),
// separable so that a caller can overlap the global reads of the next tile with the digit
// extraction of the current one.
template <bool KC, int TK = 128>
__host__ __device__ constexpr int split_lds_bytes() { return KC ? 32 * (TK + 1) * 8 : TK * 33 * 8; }

template <bool KC, int TK = 128>
__device__ __forceinline__ void split_gather(const double *__restrict__ X, int64_t sr, int64_t sk, int64_t R,
                                             int64_t K, int64_t rb, int64_t kt, double (&v)[TK / 8]) {
    const int64_t r0 = rb * 32, k0 = kt * TK;
    const int t = threadIdx.x;
#pragma unroll
    for (int i = 0; i < TK / 8; i++) {
        const int c = t + 256 * i;
        int row, kk;
        if (KC) { row = c / TK; kk = c % TK; }
        else { kk = c >> 5; row = c & 31; }
        const int64_t r = r0 + row, k = k0 + kk;
        v[i] = (r < R && k < K) ? (KC ? X[r * sr + k] : X[k * sk + r]) : 0.0;
    }
}

template <bool KC, int TK = 128>
__device__ __forceinline__ void split_stage(const double (&v)[TK / 8], double *tile) {
    constexpr int LDW = KC ? (TK + 1) : (32 + 1);
    const int t = threadIdx.x;
#pragma unroll
    for (int i = 0; i < TK / 8; i++) {
        const int c = t + 256 * i;
        if (KC) tile[(c / TK) * LDW + (c % TK)] = v[i];
        else tile[(c >> 5) * LDW + (c & 31)] = v[i];
    }
}

// Digits of the staged tile: thread (row, q) owns 16 consecutive k of one row and writes
// one 16-byte lane chunk per slice plane (q = t / 32 + 8 p for the TK / 128 passes p).
template <bool KC, int TK = 128>
__device__ __forceinline__ void split_emit(const double *tile, int64_t R, const int32_t *__restrict__ Erow, int b,
                                           int d, int8_t *__restrict__ S, int64_t Rb, int64_t Kb, int64_t rb,
                                           int64_t kt) {
    constexpr int LDW = KC ? (TK + 1) : (32 + 1);
    const int t = threadIdx.x;
    const int64_t r0 = rb * 32;
    const int row = t & 31;
    const int emax = (r0 + row < R) ? Erow[row] : 0;
    const int64_t plane = Rb * Kb * (int64_t)kBlkBytes;
#pragma unroll
    for (int pass = 0; pass < TK / 128; pass++) {
        const int q = (t >> 5) + 8 * pass;   // k in [16q, 16q + 16) of the tile
        const int64_t kb = kt * (TK / 32) + (q >> 1);
        if (kb >= Kb) break;  // no barrier below
        uint64_t mx[16];
        int bias[16];
        uint32_t negm = 0;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const double v = KC ? tile[row * LDW + 16 * q + i] : tile[(16 * q + i) * LDW + row];
            const uint64_t u = (uint64_t)__double_as_longlong(v);
            const int e = (int)((u >> 52) & 0x7FF);
            mx[i] = (u & 0xFFFFFFFFFFFFFull) | (e ? (1ull << 52) : 0ull);
            bias[i] = emax - e;
            negm |= (uint32_t)(u >> 63) << i;
        }
        int8_t *dst = S + (rb * Kb + kb) * (int64_t)kBlkBytes + (q & 1) * 512 + row * 16;
        emit_digits16(mx, bias, negm, b, d, dst, plane);
    }
}

template <bool KC>
__device__ __forceinline__ void split_tile(const double *__restrict__ X, int64_t sr, int64_t sk, int64_t R,
                                           int64_t K, const int32_t *__restrict__ Erow, int b, int d,
                                           int8_t *__restrict__ S, int64_t Rb, int64_t Kb, int64_t rb,
                                           int64_t kt, double *tile) {
    double v[16];
    split_gather<KC>(X, sr, sk, R, K, rb, kt, v);
    split_stage<KC>(v, tile);
    __syncthreads();
    split_emit<KC>(tile, R, Erow, b, d, S, Rb, Kb, rb, kt);
}

template <bool KC>
__global__ __launch_bounds__(256) void k_split(const double *__restrict__ X, int64_t sr,
                                               int64_t sk, int64_t R, int64_t K,
                                               const int32_t *__restrict__ E, int b, int d,
                                               int8_t *__restrict__ S, int64_t Rb, int64_t Kb) {
    __shared__ double tile[kSplitLds / 8];
    split_tile<KC>(X, sr, sk, R, K, E + (int64_t)blockIdx.y * 32, b, d, S, Rb, Kb, blockIdx.y, blockIdx.x, tile);
}

hipError_t launch_split(const OperandView &op, const int32_t *E, int b, int d, int8_t *S,
                        int64_t Rp, int64_t Kp, hipStream_t st) {
    const int64_t Rb = Rp / kBlk, Kb = Kp / kBlk;
    const int64_t nkt = (Kb + 3) / 4;
    if (Rb > 65535 || nkt > 0x7fffffff) return hipErrorInvalidValue;
    if (op.sk == 1)
        hipLaunchKernelGGL(k_split<true>, dim3((unsigned)nkt, (unsigned)Rb), dim3(256), 0, st, op.X,
                           op.sr, op.sk, op.R, op.K, E, b, d, S, Rb, Kb);
    else
        hipLaunchKernelGGL(k_split<false>, dim3((unsigned)nkt, (unsigned)Rb), dim3(256), 0, st,
                           op.X, op.sr, op.sk, op.R, op.K, E, b, d, S, Rb, Kb);
    return hipGetLastError();
}

// ------------------------------------------------------------------ epilogue math
__device__ __forceinline__ int clz128(u128 a) {
    const uint64_t h = (uint64_t)(a >> 64), l = (uint64_t)a;
    if (h) return __builtin_clzll(h);
    if (l) return 64 + __builtin_clzll(l);
    return 128;
}

// Exact integer T (scaled by 2^(es - 2044 - b(d+1))) to FP64.
// rounding 0: sign * trunc53(|T|) via the bit arithmetic of include/oz_aie.h:124-146
// (no over/underflow handling, as in the reference); rounding 1: nearest-even.
__device__ __forceinline__ double t_to_double(i128 t, int es, int b, int d, int rounding) {
    if (t == 0) return 0.0;
    const uint64_t sgn = t < 0 ? (1ull << 63) : 0ull;
    const u128 a = t < 0 ? (u128)(-t) : (u128)t;
    const int nb = 128 - clz128(a);
    if (rounding == 0) {
        const int shift = nb - 53;
        uint64_t lu = shift > 0 ? (uint64_t)(a >> shift) : (uint64_t)(a << (-shift));
        const int64_t ex = nb - b * (d + 1) - 1;
        lu -= 0x0010000000000000ull;
        lu += sgn + ((uint64_t)(ex + 1023) << 52) + ((uint64_t)(int64_t)(es - 2044) << 52);
        return __longlong_as_double((long long)lu);
    }
    uint64_t m;
    int e2;
    if (nb <= 53) {
        m = (uint64_t)a;
        e2 = 0;
    } else {
        const int sh = nb - 53;
        u128 q = a >> sh;
        const u128 rem = a - (q << sh);
        const u128 half = (u128)1 << (sh - 1);
        if (rem > half || (rem == half && (q & 1))) q += 1;
        m = (uint64_t)q;
        e2 = sh;
    }
    const double v = ldexp((double)m, e2 + es - 2044 - b * (d + 1));
    return sgn ? -v : v;
}

// Store one element: bins k0 .. k0+NB-1 of this pass (values vals[kk]); es is the sum of
// the element's row and column exponent fields.
template <int NB>
__device__ __forceinline__ void epilogue_elem(const GemmLaunch &g, int epi, int row, int col,
                                              const int32_t (&vals)[NB], int k0, int es) {
    if (row >= g.M || col >= g.N) return;
    if (epi == EPI_BINS) {
#pragma unroll
        for (int kk = 0; kk < NB; kk++)
            if (k0 + kk < g.d) g.bins[((int64_t)(k0 + kk) * g.M + row) * g.N + col] = vals[kk];
        return;
    }
    i128 T = 0;
#pragma unroll
    for (int kk = 0; kk < NB; kk++) {
        const int k = k0 + kk;
        if (k < g.d) T += (i128)vals[kk] << (g.b * (g.d - 1 - k));
    }
    i128 *tp = g.T + (int64_t)row * g.ldT + col;
    if (!g.first) T += *tp;
    if (g.last)
        g.C[(int64_t)row * g.ldc + col] = t_to_double(T, es, g.b, g.d, g.rounding);
    else
        *tp = T;
}

// Exact T = sum_k v_k 2^(B(D-1-k)) for a compile-time digit width B with all D bins of one
// pass: the bins are summed in 64-bit chunks of 31/B + 1 consecutive bins (relative
// shifts up to 31 bits keep a chunk of int32 terms inside int64), and the chunks are
// combined in 128-bit arithmetic, so every shift is a constant.
template <int D, int B>
__device__ __forceinline__ i128 bins_to_T(const int32_t (&v)[D]) {
    constexpr int CH = 31 / B + 1;
    i128 T = 0;
#pragma unroll
    for (int top = D - 1; top >= 0; top -= CH) {
        int64_t c = 0;
#pragma unroll
        for (int i = 0; i < CH; i++)
            if (top - i >= 0) c += (int64_t)v[top - i] << (B * i);
        T += (i128)c << (B * (D - 1 - top));
    }
    return T;
}

// The same store for EPI_C when the pass holds every bin (k0 = 0, NB = d) and b = B.
template <int D, int B>
__device__ __forceinline__ void epilogue_elem_c(const GemmLaunch &g, int row, int col, const int32_t (&vals)[D],
                                                int es) {
    if (row >= g.M || col >= g.N) return;
    i128 T = bins_to_T<D, B>(vals);
    i128 *tp = g.T + (int64_t)row * g.ldT + col;
    if (!g.first) T += *tp;
    if (g.last)
        g.C[(int64_t)row * g.ldc + col] = t_to_double(T, es, B, D, g.rounding);
    else
        *tp = T;
}

// Block -> tile mapping.  Workgroups are dispatched round-robin over the 8 XCDs; give
// each XCD a contiguous range of tiles, and walk that range in bands of 8 tile-rows so
// the concurrently running tiles of one XCD share A rows and B columns in its L2.
__device__ __forceinline__ int xcd_remap(int bid, int total) {
    const int q = total / 8, rem = total % 8;
    const int xcd = bid % 8, idx = bid / 8;
    return (xcd < rem ? xcd * (q + 1) : rem * (q + 1) + (xcd - rem) * q) + idx;
}

__device__ __forceinline__ void tile_map(int bid, int ntm, int ntn, int &tm, int &tn) {
    band_map(xcd_remap(bid, ntm * ntn), ntm, ntn, tm, tn);
}

template <int N>
__device__ __forceinline__ void s_waitcnt_vm() {
    // gfx9 encoding: vmcnt[3:0] | expcnt[6:4] | lgkmcnt[11:8] | vmcnt_hi[15:14]
    constexpr int imm = (N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8);
    __builtin_amdgcn_s_waitcnt(imm);
}

// ------------------------------------------------------------------ fused slice GEMM
// Workgroup: 4 waves as 2 x 2, wave tile (32 WM) x (32 WN), workgroup tile
// (64 WM) x (64 WN).  Per contraction block (32) every slice block of the tile is
// copied HBM -> LDS with global_load_lds_dwordx4 (one 1 KiB block per wave
// instruction) into one of two stages.  One barrier per block: after it, each wave
// reads its d A-fragments and d*WN B-fragments of the current stage, then issues the
// copies of the next block into the other stage (free: every wave has passed the
// barrier, so all reads of that stage from the previous block have completed), then
// runs the d(d+1)/2 * WM * WN MFMAs of the triangle into its d bin accumulators while
// the copies are in flight.
template <int D, int WM, int WN>
struct FusedCfg {
    static constexpr int BA = 2 * WM, BB = 2 * WN;
    static constexpr int NBLK = D * (BA + BB);
    static constexpr int STAGE = NBLK * kBlkBytes;
    static constexpr int NLD = (NBLK + 3) / 4;  // load slots per wave per stage
    static constexpr int NMIN = NBLK / 4;       // loads every wave issues per stage
};

// Epilogue of one wave: bins leave the accumulators four elements at a time through a
// per-wave LDS scratch (each lane reads back only its own values) and each element's bins
// become T and FP64 C.  The scratch of a wave holds NB * 256 bin words followed by the
// exponent fields of the wave's 64 rows and 64 columns (er_v / ec_v: this lane's row
// rbase + lane and column cbase + lane, loaded by the caller before its main loop).
template <int NB, int WM, int WN>
__host__ __device__ constexpr int wave_scratch_words() { return NB * 256 + 128; }

template <int NB, int WM, int WN, int B>
__device__ __forceinline__ void wave_epilogue_t(const GemmLaunch &g, int epi, int k0, v16i (&acc)[NB][WM][WN],
                                                int32_t *scratch, int lane, int rbase, int cbase) {
    const int32_t *ex = scratch + NB * 256;
#pragma unroll
    for (int mi = 0; mi < WM; mi++)
#pragma unroll
        for (int ni = 0; ni < WN; ni++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
#pragma unroll
                for (int k = 0; k < NB; k++)
#pragma unroll
                    for (int e = 0; e < 4; e++) scratch[(k * 4 + e) * 64 + lane] = acc[k][mi][ni][4 * q + e];
                __builtin_amdgcn_sched_barrier(0);
                const int lcol = ni * 32 + (lane & 31);
                const int ecol = ex[64 + lcol];
                if constexpr (B > 0) {
                    // four independent elements per step
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        const int lrow = mi * 32 + q * 8 + (lane >> 5) * 4 + e;
                        int32_t vals[NB];
#pragma unroll
                        for (int k = 0; k < NB; k++) vals[k] = scratch[(k * 4 + e) * 64 + lane];
                        epilogue_elem_c<NB, B>(g, rbase + lrow, cbase + lcol, vals, ex[lrow] + ecol);
                    }
                } else {
#pragma unroll 1
                    for (int e = 0; e < 4; e++) {
                        const int lrow = mi * 32 + q * 8 + (lane >> 5) * 4 + e;
                        int32_t vals[NB];
#pragma unroll
                        for (int k = 0; k < NB; k++) vals[k] = scratch[(k * 4 + e) * 64 + lane];
                        epilogue_elem<NB>(g, epi, rbase + lrow, cbase + lcol, vals, k0, ex[lrow] + ecol);
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
            }
}

template <int NB, int WM, int WN>
__device__ __forceinline__ void wave_epilogue(const GemmLaunch &g, int epi, int k0, v16i (&acc)[NB][WM][WN],
                                              int32_t *scratch, int lane, int rbase, int cbase, int er_v,
                                              int ec_v) {
    scratch[NB * 256 + lane] = er_v;
    scratch[NB * 256 + 64 + lane] = ec_v;
    if (epi == EPI_C && k0 == 0 && NB == g.d && g.b == 7)
        wave_epilogue_t<NB, WM, WN, 7>(g, epi, k0, acc, scratch, lane, rbase, cbase);
    else
        wave_epilogue_t<NB, WM, WN, 0>(g, epi, k0, acc, scratch, lane, rbase, cbase);
}

// Exponent fields this lane hands to the epilogue: row rbase + lane of A (lane < rows)
// and column cbase + lane of B (lane < cols).
__device__ __forceinline__ void load_tile_exps(const GemmLaunch &g, int epi, int lane, int rbase, int rows,
                                               int cbase, int cols, int &er_v, int &ec_v) {
    er_v = 0;
    ec_v = 0;
    if (epi != EPI_C) return;
    if (lane < rows && rbase + lane < g.M) er_v = g.Er[rbase + lane];
    if (lane < cols && cbase + lane < g.N) ec_v = g.Ec[cbase + lane];
}

template <int D, int WM, int WN>
__host__ __device__ constexpr int lds_tile_bytes() { return 2 * FusedCfg<D, WM, WN>::STAGE; }

template <int D, int WM, int WN>
__device__ __forceinline__ void gemm_lds_tile(const GemmLaunch &g, int tm, int tn, int epi, int8_t *smem) {
    using F = FusedCfg<D, WM, WN>;
    constexpr int BA = F::BA, BB = F::BB, NBLK = F::NBLK, NLD = F::NLD, NMIN = F::NMIN;

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

    v16i acc[D][WM][WN];
#pragma unroll
    for (int k = 0; k < D; k++)
#pragma unroll
        for (int mi = 0; mi < WM; mi++)
#pragma unroll
            for (int ni = 0; ni < WN; ni++) acc[k][mi][ni] = v16i{};

    const int64_t kb0 = g.kb0, kb1 = g.kb1;
    if (kb0 < kb1) issue(0, kb0);
    for (int64_t kb = kb0; kb < kb1; kb++) {
        const int s = (int)((kb - kb0) & 1);
        s_waitcnt_vm<0>();
        __syncthreads();
        const int8_t *st = smem + s * F::STAGE;
        v4i af[D][WM], bf[D][WN];
#pragma unroll
        for (int j = 0; j < D; j++) {
#pragma unroll
            for (int mi = 0; mi < WM; mi++)
                af[j][mi] = *reinterpret_cast<const v4i *>(st + (j * BA + wm * WM + mi) * kBlkBytes + lane * 16);
#pragma unroll
            for (int ni = 0; ni < WN; ni++)
                bf[j][ni] = *reinterpret_cast<const v4i *>(st + (D * BA + j * BB + wn * WN + ni) * kBlkBytes + lane * 16);
        }
        if (kb + 1 < kb1) issue(s ^ 1, kb + 1);
#pragma unroll
        for (int i = 0; i < D; i++)
#pragma unroll
            for (int j = 0; i + j < D; j++)
#pragma unroll
                for (int mi = 0; mi < WM; mi++)
#pragma unroll
                    for (int ni = 0; ni < WN; ni++)
                        acc[i + j][mi][ni] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af[i][mi], bf[j][ni],
                                                                                   acc[i + j][mi][ni], 0, 0, 0);
    }
    __syncthreads();

    int32_t *scratch = reinterpret_cast<int32_t *>(smem) + w * wave_scratch_words<D, WM, WN>();
    wave_epilogue<D, WM, WN>(g, epi, 0, acc, scratch, lane, rbase, cbase, er_v, ec_v);
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

// Direct variant: no LDS in the main loop.  Each wave streams its own A and B
// fragments HBM/L2 -> VGPR (one 1 KiB block per wave instruction, 16 bytes per lane),
// double-buffered in registers, so there is no barrier and the compiler's per-register
// wait counts overlap the loads of block t+1 with the MFMAs of block t.  Requires an
// even number of contraction blocks per pass (the planner pads K to 64).
template <int D>
__host__ __device__ constexpr int direct_tile_bytes() { return 4 * 4 * (D * 256 + 128); }

template <int D, int WM, int WN>
__device__ __forceinline__ void gemm_direct_tile(const GemmLaunch &g, int tm, int tn, int epi, int8_t *smem) {
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = w & 1, wn = w >> 1;
    const int64_t planeA = g.RbA * g.Kb * kBlkBytes, planeB = g.RbB * g.Kb * kBlkBytes;
    const int64_t rbA = (int64_t)tm * 2 * WM + wm * WM, rbB = (int64_t)tn * 2 * WN + wn * WN;
    const int rbase = tm * 64 * WM + wm * 32 * WM, cbase = tn * 64 * WN + wn * 32 * WN;
    int er_v, ec_v;
    load_tile_exps(g, epi, lane, rbase, 32 * WM, cbase, 32 * WN, er_v, ec_v);
    const int8_t *pa = g.SA + rbA * g.Kb * kBlkBytes + lane * 16;
    const int8_t *pb = g.SB + rbB * g.Kb * kBlkBytes + lane * 16;
    const int64_t rowstride = g.Kb * kBlkBytes;

    v4i a0[D][WM], b0[D][WN], a1[D][WM], b1[D][WN];
    auto load = [&](v4i (&af)[D][WM], v4i (&bf)[D][WN], int64_t kb) {
#pragma unroll
        for (int j = 0; j < D; j++) {
#pragma unroll
            for (int mi = 0; mi < WM; mi++)
                af[j][mi] = *reinterpret_cast<const v4i *>(pa + j * planeA + mi * rowstride + kb * kBlkBytes);
#pragma unroll
            for (int ni = 0; ni < WN; ni++)
                bf[j][ni] = *reinterpret_cast<const v4i *>(pb + j * planeB + ni * rowstride + kb * kBlkBytes);
        }
    };
    v16i acc[D][WM][WN];
#pragma unroll
    for (int k = 0; k < D; k++)
#pragma unroll
        for (int mi = 0; mi < WM; mi++)
#pragma unroll
            for (int ni = 0; ni < WN; ni++) acc[k][mi][ni] = v16i{};
    auto mma = [&](const v4i (&af)[D][WM], const v4i (&bf)[D][WN]) {
#pragma unroll
        for (int i = 0; i < D; i++)
#pragma unroll
            for (int j = 0; i + j < D; j++)
#pragma unroll
                for (int mi = 0; mi < WM; mi++)
#pragma unroll
                    for (int ni = 0; ni < WN; ni++)
                        acc[i + j][mi][ni] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af[i][mi], bf[j][ni],
                                                                                   acc[i + j][mi][ni], 0, 0, 0);
    };
    const int64_t kb0 = g.kb0, nk = g.kb1 - g.kb0;   // nk is even
    if (nk > 0) load(a0, b0, kb0);
    for (int64_t t = 0; t < nk; t += 2) {
        load(a1, b1, kb0 + t + 1);
        mma(a0, b0);
        if (t + 2 < nk) load(a0, b0, kb0 + t + 2);
        mma(a1, b1);
    }
    int32_t *scratch = reinterpret_cast<int32_t *>(smem) + w * wave_scratch_words<D, WM, WN>();
    wave_epilogue<D, WM, WN>(g, epi, 0, acc, scratch, lane, rbase, cbase, er_v, ec_v);
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

bool fused_supported(int d) { return d >= 1 && d <= 16; }

// Default fused-kernel family per slice count (measured on MI355X at 8192^3): the
// direct-to-register kernel for d <= 8 (its two fragment sets fit in registers), the
// LDS-copy kernel above that.
static int default_family(int d) { return d <= 8 ? KERNEL_DIRECT : KERNEL_LDS; }

template <int D, int WM, int WN>
static hipError_t launch_fused_t(const GemmLaunch &g, int epi, hipStream_t st) {
    const int ntm = (g.M + 64 * WM - 1) / (64 * WM), ntn = (g.N + 64 * WN - 1) / (64 * WN);
    const int64_t tiles = (int64_t)ntm * ntn;
    if (tiles <= 0) return hipSuccess;
    int fam = g.family == KERNEL_AUTO ? default_family(D) : g.family;
    if (fam == KERNEL_DIRECT)
        hipLaunchKernelGGL((k_gemm_direct<D, WM, WN>), dim3((unsigned)tiles), dim3(256), 0, st, g, epi);
    else if (fam == KERNEL_PIPE)
        hipLaunchKernelGGL((k_gemm_pipe<D, WM, WN>), dim3((unsigned)tiles), dim3(256), 0, st, g, epi);
    else
        hipLaunchKernelGGL((k_gemm_fused<D, WM, WN>), dim3((unsigned)tiles), dim3(256), 0, st, g, epi);
    return hipGetLastError();
}

// Wave tile per slice count: the accumulators (d bins x 16 registers per 32 x 32 block)
// must leave room for the fragments in the 512 registers of a wave.
hipError_t launch_gemm_fused(const GemmLaunch &g, int epi, hipStream_t st) {
    switch (g.d) {
        case 1: return launch_fused_t<1, 2, 2>(g, epi, st);
        case 2: return launch_fused_t<2, 2, 2>(g, epi, st);
        case 3: return launch_fused_t<3, 2, 2>(g, epi, st);
        case 4: return launch_fused_t<4, 2, 2>(g, epi, st);
        case 5: return launch_fused_t<5, 1, 2>(g, epi, st);
        case 6: return launch_fused_t<6, 1, 2>(g, epi, st);
        case 7: return launch_fused_t<7, 1, 2>(g, epi, st);
        case 8: return launch_fused_t<8, 1, 2>(g, epi, st);
        case 9: return launch_fused_t<9, 1, 2>(g, epi, st);
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

// ------------------------------------------------------------------ grouped launches
// Many independent problems per launch: every block reads its work item (operand or
// problem, and the block index inside it) from a table built on the host, then runs the
// same tile body as the single-problem kernels.

__global__ __launch_bounds__(256) void k_grp_maxexp(const GroupOperand *__restrict__ ops,
                                                    const int4 *__restrict__ items) {
    const int4 it = items[blockIdx.x];
    const GroupOperand &o = ops[it.x];
    if (o.sk == 1) {
        const int64_t r = (int64_t)it.y * 4 + (threadIdx.x >> 6);
        if (r < o.Rp) maxexp_kc_row(o.X, o.sr, o.R, o.K, o.E, r, threadIdx.x & 63);
    } else {
        maxexp_rc_slab(o.X, o.sk, o.R, o.K, o.E, (int64_t)it.y * 256 + threadIdx.x, (int64_t)it.z * it.w, it.w);
    }
}

__global__ __launch_bounds__(256) void k_grp_split(const GroupOperand *__restrict__ ops,
                                                   const int4 *__restrict__ items, int b, int d) {
    __shared__ double tile[kSplitLds / 8];
    const int4 it = items[blockIdx.x];
    const GroupOperand &o = ops[it.x];
    const int32_t *Erow = o.E + (int64_t)it.y * 32;
    if (o.sk == 1)
        split_tile<true>(o.X, o.sr, o.sk, o.R, o.K, Erow, b, d, o.S, o.Rb, o.Kb, it.y, it.z, tile);
    else
        split_tile<false>(o.X, o.sr, o.sk, o.R, o.K, Erow, b, d, o.S, o.Rb, o.Kb, it.y, it.z, tile);
}

hipError_t launch_group_maxexp(const GroupOperand *ops, const int4 *items, int n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_grp_maxexp, dim3((unsigned)n), dim3(256), 0, st, ops, items);
    return hipGetLastError();
}

hipError_t launch_group_split(const GroupOperand *ops, const int4 *items, int n, int b, int d, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_grp_split, dim3((unsigned)n), dim3(256), 0, st, ops, items, b, d);
    return hipGetLastError();
}

template <int D, int WM, int WN, int FAM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1)))
void k_grp_gemm(const GemmLaunch *__restrict__ probs, const int4 *__restrict__ items) {
    // consecutive items (tiles of one problem in bands of tile-rows) go to one XCD
    const int4 it = items[xcd_remap(blockIdx.x, gridDim.x)];
    const GemmLaunch g = probs[it.x];
    constexpr int LDS = FAM == KERNEL_DIRECT ? direct_tile_bytes<D>() : lds_tile_bytes<D, WM, WN>();
    __shared__ __attribute__((aligned(1024))) int8_t smem[LDS];
    if (FAM == KERNEL_DIRECT) gemm_direct_tile<D, WM, WN>(g, it.y, it.z, EPI_C, smem);
    else gemm_lds_tile<D, WM, WN>(g, it.y, it.z, EPI_C, smem);
}

int group_cfg_tile_m(int cfg) { return cfg == GCFG_D22 ? 128 : 64; }
int group_cfg_tile_n(int cfg) { return (cfg == GCFG_D11 || cfg == GCFG_L11) ? 64 : 128; }

bool group_cfg_supported(int d, int cfg) {
    if (d != 3 && d != 7 && d != 8 && d != 9) return false;
    if (cfg == GCFG_D22) return d == 3;
    if (cfg == GCFG_D12) return d != 9;   // two fragment sets of 32 x 64 do not fit at d = 9
    return cfg >= 0 && cfg < GCFG_COUNT;
}

template <int D>
static hipError_t launch_grp_d(const GemmLaunch *probs, const int4 *items, int n, int cfg, hipStream_t st) {
    const dim3 grid((unsigned)n), blk(256);
    switch (cfg) {
        case GCFG_D11: hipLaunchKernelGGL((k_grp_gemm<D, 1, 1, KERNEL_DIRECT>), grid, blk, 0, st, probs, items); break;
        case GCFG_D12:
            if constexpr (D <= 8) hipLaunchKernelGGL((k_grp_gemm<D, 1, 2, KERNEL_DIRECT>), grid, blk, 0, st, probs, items);
            break;
        case GCFG_D22:
            if constexpr (D <= 4) hipLaunchKernelGGL((k_grp_gemm<D, 2, 2, KERNEL_DIRECT>), grid, blk, 0, st, probs, items);
            break;
        case GCFG_L11: hipLaunchKernelGGL((k_grp_gemm<D, 1, 1, KERNEL_LDS>), grid, blk, 0, st, probs, items); break;
        case GCFG_L12: hipLaunchKernelGGL((k_grp_gemm<D, 1, 2, KERNEL_LDS>), grid, blk, 0, st, probs, items); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_group_gemm(const GemmLaunch *probs, const int4 *items, int n, int d, int cfg, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (!group_cfg_supported(d, cfg)) return hipErrorInvalidValue;
    switch (d) {
        case 3: return launch_grp_d<3>(probs, items, n, cfg, st);
        case 7: return launch_grp_d<7>(probs, items, n, cfg, st);
        case 8: return launch_grp_d<8>(probs, items, n, cfg, st);
        case 9: return launch_grp_d<9>(probs, items, n, cfg, st);
        default: return hipErrorInvalidValue;
    }
}

// ------------------------------------------------------------------ persistent grouped launch
// One launch for a whole group: a grid of one workgroup per CU runs
//   phase A  the exponent maximum and the split of every 32-row block of every operand
//            (one item = 32 rows over the whole contraction axis, so no item depends on
//            another),
//   a grid-wide barrier (every operand is split before any product starts),
//   phase B  the fused slice-GEMM tiles of every problem, with the first contraction block
//            of a workgroup's next tile loaded into registers while the current tile's
//            epilogue runs.
// Items are assigned to workgroups on the host (balanced lists, tiles of one problem on
// one XCD), so the only cross-workgroup synchronisation is the barrier.

// Phase A item, run by ONE wave: the 32 rows r0 = rb*32 .. +32 of operand o.  First the
// exponent maximum of each row over the whole contraction axis (into o.E), then the
// split in tiles of 32 rows x 64 contraction indices (two slice blocks): the tile is read
// coalesced along the contiguous axis, transposed through the wave's own LDS region
// (32 x 65 doubles), and lane l emits the 16-byte chunk it owns in each block (row l % 32,
// k 16 (l / 32) .. +16).  The reads of tile t+1 are in flight while tile t is emitted.
// Rows past R read row R - 1 (always valid) and are masked, so loads carry no branches.
constexpr int kWaveTileLd = 65;   // LDS row stride (doubles) of the per-wave tile

template <bool KC>
__device__ __forceinline__ void wave_gather(const GroupOperand &o, int64_t r0, int64_t k0, int lane, double (&v)[32]) {
    if (KC) {
        // load i: row r0 + i, contraction index k0 + lane
        const int64_t k = k0 + lane;
#pragma unroll
        for (int i = 0; i < 32; i++) {
            const int64_t r = r0 + i;
            v[i] = (r < o.R && k < o.K) ? o.X[r * o.sr + k] : 0.0;
        }
    } else {
        // load i: contraction index k0 + 2i + lane / 32, row r0 + lane % 32
        const int64_t r = r0 + (lane & 31);
#pragma unroll
        for (int i = 0; i < 32; i++) {
            const int64_t k = k0 + 2 * i + (lane >> 5);
            v[i] = (r < o.R && k < o.K) ? o.X[k * o.sk + r] : 0.0;
        }
    }
}

template <bool KC>
__device__ __forceinline__ void wave_stage(const double (&v)[32], double *wt, int lane) {
#pragma unroll
    for (int i = 0; i < 32; i++) {
        if (KC) wt[i * kWaveTileLd + lane] = v[i];
        else wt[(lane & 31) * kWaveTileLd + 2 * i + (lane >> 5)] = v[i];
    }
}

// Digits of one 16-element chunk (row `row`, elements wt[row][c0 .. c0+16)) into the
// lane's 16-byte chunk of d slice planes at dst.
__device__ __forceinline__ void wave_emit_chunk(const double *wt, int row, int c0, int emax, int b, int d,
                                                int8_t *__restrict__ dst, int64_t plane) {
    uint64_t mx[16];
    int bias[16];
    uint32_t negm = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const uint64_t u = (uint64_t)__double_as_longlong(wt[row * kWaveTileLd + c0 + i]);
        const int e = (int)((u >> 52) & 0x7FF);
        mx[i] = (u & 0xFFFFFFFFFFFFFull) | (e ? (1ull << 52) : 0ull);
        bias[i] = emax - e;
        negm |= (uint32_t)(u >> 63) << i;
    }
    emit_digits16(mx, bias, negm, b, d, dst, plane);
}

template <bool KC>
__device__ __forceinline__ void wave_rowblock(const GroupOperand &o, int64_t rb, int b, int d, double *wt, int lane) {
    const int64_t r0 = rb * 32;
    // 1. exponent maxima (64 loads in flight per lane); lane l < 32 ends with row l's
    int emax;
    if (KC) {
        const double *rowp[8];
        int m[32];
#pragma unroll
        for (int i = 0; i < 32; i++) m[i] = 0;
#pragma unroll
        for (int gi = 0; gi < 4; gi++) {
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const int64_t r = r0 + gi * 8 + i;
                rowp[i] = o.X + (r < o.R ? r : o.R - 1) * o.sr;
            }
            int64_t k = lane;
            for (; k + 448 < o.K; k += 512) {
                double a[8][8];
#pragma unroll
                for (int i = 0; i < 8; i++)
#pragma unroll
                    for (int u = 0; u < 8; u++) a[i][u] = rowp[i][k + 64 * u];
#pragma unroll
                for (int i = 0; i < 8; i++)
#pragma unroll
                    for (int u = 0; u < 8; u++) m[gi * 8 + i] = max(m[gi * 8 + i], field_of(a[i][u]));
            }
            for (; k < o.K; k += 64) {
#pragma unroll
                for (int i = 0; i < 8; i++) m[gi * 8 + i] = max(m[gi * 8 + i], field_of(rowp[i][k]));
            }
        }
        // transpose-reduce through LDS: [32 rows][64 lanes]
        int32_t *red = reinterpret_cast<int32_t *>(wt);
#pragma unroll
        for (int i = 0; i < 32; i++) red[i * 64 + lane] = m[i];
        emax = 0;
        if (lane < 32) {
#pragma unroll
            for (int q = 0; q < 64; q += 4) {
                const int4 x = *reinterpret_cast<const int4 *>(red + lane * 64 + q);
                emax = max(emax, max(max(x.x, x.y), max(x.z, x.w)));
            }
        }
    } else {
        // lane l: row r0 + l % 32, contraction indices of parity l / 32
        const int64_t r = r0 + (lane & 31);
        const double *col = o.X + (r < o.R ? r : o.R - 1);
        int m = 0;
        int64_t k = lane >> 5;
        for (; k + 2 * 63 < o.K; k += 128) {
            double a[64];
#pragma unroll
            for (int u = 0; u < 64; u++) a[u] = col[(k + 2 * u) * o.sk];
#pragma unroll
            for (int u = 0; u < 64; u++) m = max(m, field_of(a[u]));
        }
        for (; k < o.K; k += 2) m = max(m, field_of(col[k * o.sk]));
        emax = max(m, __shfl_xor(m, 32, 64));
    }
    if (r0 + (lane & 31) >= o.R) emax = 0;
    if (lane < 32) o.E[r0 + lane] = emax;
    emax = __shfl(emax, lane & 31, 64);
    // 2. split, tiles of 64 contraction indices (Kb is even), two tiles in flight
    const int64_t nt = o.Kb / 2;
    const int64_t plane = o.Rb * o.Kb * (int64_t)kBlkBytes;
    const int row = lane & 31, half = lane >> 5;
    auto emit = [&](int64_t t) {
#pragma unroll
        for (int blk = 0; blk < 2; blk++) {
            int8_t *dst = o.S + (rb * o.Kb + 2 * t + blk) * (int64_t)kBlkBytes + lane * 16;
            wave_emit_chunk(wt, row, 32 * blk + 16 * half, emax, b, d, dst, plane);
        }
    };
    double v0[32], v1[32];
    wave_gather<KC>(o, r0, 0, lane, v0);
    if (nt > 1) wave_gather<KC>(o, r0, 64, lane, v1);
    for (int64_t t = 0; t < nt; t += 2) {
        wave_stage<KC>(v0, wt, lane);
        if (t + 2 < nt) wave_gather<KC>(o, r0, (t + 2) * 64, lane, v0);
        emit(t);
        if (t + 1 < nt) {
            wave_stage<KC>(v1, wt, lane);
            if (t + 3 < nt) wave_gather<KC>(o, r0, (t + 3) * 64, lane, v1);
            emit(t + 1);
        }
    }
}

struct PersistCtl {
    int arrive;   // workgroups through the phase barrier
    int done;     // workgroups finished (the last one resets both counters)
    int error;    // sticky: a barrier wait exceeded its bound
    int pad[61];
};

constexpr int kSpinLimit = 1 << 20;   // barrier polls (each a device-coherent load + sleep)

// Grid barrier: every wave makes its stores visible at agent scope, one thread per
// workgroup arrives and polls; a wait longer than kSpinLimit polls (or an error raised by
// another workgroup) sets the error flag and releases everyone.  Returns false on error.
__device__ bool grid_barrier(PersistCtl *ctl, int G, int *flag_lds) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0) {
        int ok = 1;
        __hip_atomic_fetch_add(&ctl->arrive, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        for (int spin = 0;; spin++) {
            if (__hip_atomic_load(&ctl->arrive, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= G) break;
            if (spin >= kSpinLimit || __hip_atomic_load(&ctl->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                __hip_atomic_store(&ctl->error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        *flag_lds = ok;
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return *flag_lds != 0;
}

// Operands of one tile of phase B (all wave-uniform: the lane offset is added at the loads).
struct PTile {
    const int8_t *pa, *pb;
    int64_t planeA, planeB, rowstride;
    int nk, rbase, cbase;
    int M, N;
    int64_t ldc;
    double *C;
    const int32_t *Er, *Ec;
};

template <int WM, int WN>
__device__ __forceinline__ PTile ptile_setup(const GemmLaunch *__restrict__ probs, int4 it, int wm, int wn) {
    const GemmLaunch &g = probs[it.x];
    PTile t;
    const int64_t rowstride = g.Kb * kBlkBytes;
    const int64_t rbA = (int64_t)it.y * 2 * WM + wm * WM, rbB = (int64_t)it.z * 2 * WN + wn * WN;
    t.pa = g.SA + rbA * rowstride;
    t.pb = g.SB + rbB * rowstride;
    t.planeA = g.RbA * rowstride;
    t.planeB = g.RbB * rowstride;
    t.rowstride = rowstride;
    t.nk = (int)g.Kb;
    t.rbase = it.y * 64 * WM + wm * 32 * WM;
    t.cbase = it.z * 64 * WN + wn * 32 * WN;
    t.M = g.M;
    t.N = g.N;
    t.ldc = g.ldc;
    t.C = g.C;
    t.Er = g.Er;
    t.Ec = g.Ec;
    return t;
}

template <int D>
__host__ __device__ constexpr int persist_lds_bytes() {
    constexpr int a = 4 * 32 * kWaveTileLd * 8;   // phase A: one tile per wave
    return a > direct_tile_bytes<D>() ? a : direct_tile_bytes<D>();
}

template <int D, int WM, int WN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1)))
void k_grp_persistent(const GroupOperand *__restrict__ ops, const int4 *__restrict__ itemsA,
                      const int *__restrict__ offA, const GemmLaunch *__restrict__ probs,
                      const int4 *__restrict__ itemsB, const int *__restrict__ offB, PersistCtl *ctl,
                      int b, int rounding) {
    __shared__ __attribute__((aligned(1024))) int8_t smem[persist_lds_bytes<D>()];
    __shared__ int flag_lds;
    const int G = gridDim.x;

    // ---- phase A: exponent maximum + split, one 32-row block per item, each wave working
    // through its own item list (no workgroup barriers)
    {
        const int lane = threadIdx.x & 63;
        const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        double *wt = reinterpret_cast<double *>(smem) + w * (32 * kWaveTileLd);
        const int slot = blockIdx.x * 4 + w;
        const int a0 = offA[slot], a1 = offA[slot + 1];
        for (int i = a0; i < a1; i++) {
            const int4 it = itemsA[i];
            const GroupOperand &o = ops[it.x];
            if (o.sk == 1) wave_rowblock<true>(o, it.y, b, D, wt, lane);
            else wave_rowblock<false>(o, it.y, b, D, wt, lane);
        }
    }

    const bool ok = grid_barrier(ctl, G, &flag_lds);

    // ---- phase B: fused slice-GEMM tiles
    if (ok) {
        const int lane = threadIdx.x & 63;
        const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        const int wm = w & 1, wn = w >> 1;
        int32_t *scratch = reinterpret_cast<int32_t *>(smem) + w * wave_scratch_words<D, WM, WN>();
        const int i0 = offB[blockIdx.x], i1 = offB[blockIdx.x + 1];
        const int loff = lane * 16;

        v4i a0[D][WM], b0[D][WN], a1[D][WM], b1[D][WN];
        auto load = [&](v4i (&af)[D][WM], v4i (&bf)[D][WN], const PTile &t, int kb) {
#pragma unroll
            for (int j = 0; j < D; j++) {
#pragma unroll
                for (int mi = 0; mi < WM; mi++)
                    af[j][mi] = *reinterpret_cast<const v4i *>(t.pa + j * t.planeA + mi * t.rowstride +
                                                              (int64_t)kb * kBlkBytes + loff);
#pragma unroll
                for (int ni = 0; ni < WN; ni++)
                    bf[j][ni] = *reinterpret_cast<const v4i *>(t.pb + j * t.planeB + ni * t.rowstride +
                                                              (int64_t)kb * kBlkBytes + loff);
            }
        };
        v16i acc[D][WM][WN];
        auto mma = [&](const v4i (&af)[D][WM], const v4i (&bf)[D][WN]) {
#pragma unroll
            for (int i = 0; i < D; i++)
#pragma unroll
                for (int j = 0; i + j < D; j++)
#pragma unroll
                    for (int mi = 0; mi < WM; mi++)
#pragma unroll
                        for (int ni = 0; ni < WN; ni++)
                            acc[i + j][mi][ni] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af[i][mi], bf[j][ni],
                                                                                       acc[i + j][mi][ni], 0, 0, 0);
        };

        PTile cur{}, nxt{};
        if (i0 < i1) {
            cur = ptile_setup<WM, WN>(probs, itemsB[i0], wm, wn);
            load(a0, b0, cur, 0);
        }
        for (int i = i0; i < i1; i++) {
            const bool has_next = i + 1 < i1;
            if (has_next) nxt = ptile_setup<WM, WN>(probs, itemsB[i + 1], wm, wn);
            GemmLaunch g{};
            g.M = cur.M;
            g.N = cur.N;
            g.b = b;
            g.d = D;
            g.rounding = rounding;
            g.first = 1;
            g.last = 1;
            g.C = cur.C;
            g.ldc = cur.ldc;
            g.Er = cur.Er;
            g.Ec = cur.Ec;
            int er_v, ec_v;
            load_tile_exps(g, EPI_C, lane, cur.rbase, 32 * WM, cur.cbase, 32 * WN, er_v, ec_v);
#pragma unroll
            for (int k = 0; k < D; k++)
#pragma unroll
                for (int mi = 0; mi < WM; mi++)
#pragma unroll
                    for (int ni = 0; ni < WN; ni++) acc[k][mi][ni] = v16i{};
            const int nk = cur.nk;   // even
            for (int t = 0; t < nk; t += 2) {
                load(a1, b1, cur, t + 1);
                mma(a0, b0);
                if (t + 2 < nk) load(a0, b0, cur, t + 2);
                else if (has_next) load(a0, b0, nxt, 0);   // next tile's first block
                mma(a1, b1);
            }
            wave_epilogue<D, WM, WN>(g, EPI_C, 0, acc, scratch, lane, cur.rbase, cur.cbase, er_v, ec_v);
            cur = nxt;
        }
    }

    // ---- exit: the last workgroup resets the counters for the next launch
    __syncthreads();
    if (threadIdx.x == 0) {
        const int old = __hip_atomic_fetch_add(&ctl->done, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (old == G - 1) {
            __hip_atomic_store(&ctl->arrive, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&ctl->done, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

bool persist_cfg_supported(int d, int cfg) {
    if (d != 3 && d != 7 && d != 8 && d != 9) return false;
    if (cfg == GCFG_D11) return true;
    if (cfg == GCFG_D12) return d == 3 || d == 7;   // d = 8 and 9 would spill registers
    if (cfg == GCFG_D22) return d == 3;
    return false;
}

int persist_ctl_bytes() { return (int)sizeof(PersistCtl); }

template <int D, int WM, int WN>
static hipError_t launch_persist_t(const PersistLaunch &L, hipStream_t st) {
    hipLaunchKernelGGL((k_grp_persistent<D, WM, WN>), dim3((unsigned)L.grid), dim3(256), 0, st, L.ops, L.itemsA,
                       L.offA, L.probs, L.itemsB, L.offB, reinterpret_cast<PersistCtl *>(L.ctl), L.b,
                       L.rounding);
    return hipGetLastError();
}

hipError_t launch_group_persistent(const PersistLaunch &L, int d, int cfg, hipStream_t st) {
    if (L.grid <= 0) return hipSuccess;
    if (!persist_cfg_supported(d, cfg)) return hipErrorInvalidValue;
    switch (d * 8 + cfg) {
        case 3 * 8 + GCFG_D11: return launch_persist_t<3, 1, 1>(L, st);
        case 3 * 8 + GCFG_D12: return launch_persist_t<3, 1, 2>(L, st);
        case 3 * 8 + GCFG_D22: return launch_persist_t<3, 2, 2>(L, st);
        case 7 * 8 + GCFG_D11: return launch_persist_t<7, 1, 1>(L, st);
        case 7 * 8 + GCFG_D12: return launch_persist_t<7, 1, 2>(L, st);
        case 8 * 8 + GCFG_D11: return launch_persist_t<8, 1, 1>(L, st);
        case 9 * 8 + GCFG_D11: return launch_persist_t<9, 1, 1>(L, st);
        default: return hipErrorInvalidValue;
    }
}

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
