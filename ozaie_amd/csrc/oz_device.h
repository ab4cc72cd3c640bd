// oz_device.h -- device-side building blocks shared by the kernel translation units
// (oz_kernels.hip: single products; oz_grouped.hip: grouped and persistent launches;
// oz_crt.hip: the CRT mode): exponent-field maxima, the split of FP64 elements into int8
// slice blocks, the epilogue arithmetic (exact 128-bit T and its conversion to FP64), the
// XCD-aware tile map and the fused slice-GEMM tile bodies.  Everything here is inline.
#pragma once
#include <type_traits>

#include "oz_internal.h"

namespace oz {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef __int128 i128;
typedef unsigned __int128 u128;
typedef __attribute__((address_space(3))) void lds_void;

__device__ __forceinline__ int field_of(double x) {
    const uint32_t hi = (uint32_t)__double2hiint(x);
    return (int)((hi << 1) >> 21);
}

// ------------------------------------------------------------------ max exponent bodies
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
// Wider strings shift the mantissa per digit.  Only planes j0 <= j < j1 are written.
__device__ __forceinline__ void emit_digits16(uint64_t (&mx)[16], const int (&bias)[16], uint32_t negm, int b,
                                              int d, int8_t *__restrict__ dst, int64_t plane, int j0 = 0,
                                              int j1 = 1 << 30) {
    if (j1 > d) j1 = d;
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
        int j = j0;
        for (; j < j1 && b * (d - 1 - j) >= 32; j++) {   // field inside the high word
            const uint32_t off = (uint32_t)(b * (d - 1 - j) - 32);
            plane_out(j, [&](int i) { return __builtin_amdgcn_ubfe(qhi[i], off, (uint32_t)b); });
        }
        for (; j < j1; j++) {                             // field at bit < 32 (may straddle)
            const uint32_t sj = (uint32_t)(b * (d - 1 - j));
            plane_out(j, [&](int i) { return __builtin_amdgcn_alignbit(qhi[i], qlo[i], sj) & mask; });
        }
        return;
    }
    const uint64_t mask = (1ull << b) - 1ull;
    for (int j = j0; j < j1; j++) {
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

// Paired-digit split of 16 elements (pair_cfg): the digit string Q = floor(M_x 2^(bd - 53 -
// bias)) (below 2^96), taken as shifted left by b when d is odd so that it holds D =
// ceil(d / 2) whole pairs; pair plane j < P is the 2b-bit field at 2b(D-1-j), single-digit plane I < Q
// the b-bit field at 2b(D-1-I) + b (d odd: the pair's high digit s_2I) or at 2b(D-1-I)
// (d even: its low digit s_2I+1).  dstS: the lane's chunk in pair plane 0 (planes planeS
// apart); dstX: its chunk in single-digit slot 0 (slots planeX apart, written in reverse
// order when reverse_q).  Every field is below 64, so the carry-free byte negation of
// emit_digits16 applies.  Only planes g0 <= g < g1 of the P pair planes followed by the Q
// single-digit slots are written.
__device__ __forceinline__ void emit_pairs16(const uint64_t (&mx)[16], const int (&bias)[16], uint32_t negm,
                                             int b, int d, int8_t *__restrict__ dstS, int64_t planeS,
                                             int8_t *__restrict__ dstX, int64_t planeX, bool reverse_q,
                                             int g0 = 0, int g1 = 1 << 30) {
    const PairCfg pc = pair_cfg(b, d);
    const int D = (d + 1) / 2;
    uint32_t nbm[4];
#pragma unroll
    for (int gi = 0; gi < 4; gi++) {
        uint32_t m = 0;
#pragma unroll
        for (int h = 0; h < 4; h++) m |= ((negm >> (4 * gi + h)) & 1u) ? (0xFFu << (8 * h)) : 0u;
        nbm[gi] = m;
    }
    // The field at bit pos of Q' = M_x 2^s (s = bd - 53 - bias + b [d odd]) straight from the
    // 53-bit mantissa with one 64-bit shift by s - pos (left if positive): the digit string
    // itself may be wider than 64 bits, the shifted mantissa never needs to be.
    int s[16];
#pragma unroll
    for (int i = 0; i < 16; i++) s[i] = b * d - 53 - bias[i] + (pc.odd ? b : 0);
    auto field = [&](int i, int pos, uint32_t mask) -> uint32_t {
        const int t = s[i] - pos;
        const uint64_t v = t >= 0 ? (t >= 64 ? 0ull : mx[i] << t) : (t <= -64 ? 0ull : mx[i] >> (-t));
        return (uint32_t)v & mask;
    };
    auto plane_out = [&](int8_t *dst, int pos, uint32_t mask) {
        uint32_t w[4];
#pragma unroll
        for (int gi = 0; gi < 4; gi++) {
            uint32_t p = 0;
#pragma unroll
            for (int h = 0; h < 4; h++) p |= field(4 * gi + h, pos, mask) << (8 * h);
            const uint32_t pn = (0x80808080u - p) ^ 0x80808080u;
            w[gi] = (pn & nbm[gi]) | (p & ~nbm[gi]);
        }
        *reinterpret_cast<uint4 *>(dst) = make_uint4(w[0], w[1], w[2], w[3]);
    };
    const int jend = g1 < pc.P ? g1 : pc.P;
    for (int j = g0; j < jend; j++) plane_out(dstS + (int64_t)j * planeS, 2 * b * (D - 1 - j), (1u << (2 * b)) - 1u);
    const int iend = g1 - pc.P < pc.Q ? g1 - pc.P : pc.Q;
    for (int I = g0 > pc.P ? g0 - pc.P : 0; I < iend; I++) {
        const int slot = reverse_q ? pc.Q - 1 - I : I;
        plane_out(dstX + (int64_t)slot * planeX, 2 * b * (D - 1 - I) + (pc.odd ? b : 0), (1u << b) - 1u);
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

// Exponent maxima of the staged tile's 32 rows, for a tile that holds the whole contraction
// axis (K <= 128): the exponent pass folded into the split.  Thread (row, q) takes the
// maximum field of its 16 elements, and the eight q of a row meet in red[32 ..]; red[row]
// then holds the row's maximum (zeros, padding included, count as field 0, the exponent
// pass's starting value), also stored to E[row] when `write`.  Two barriers, every thread.
template <bool KC>
__device__ __forceinline__ void tile_row_max(const double *tile, int32_t *red, int32_t *__restrict__ E, bool write) {
    constexpr int LDW = KC ? (128 + 1) : (32 + 1);
    const int t = threadIdx.x, row = t & 31, q = t >> 5;
    int m = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const double v = KC ? tile[row * LDW + 16 * q + i] : tile[(16 * q + i) * LDW + row];
        const int e = (int)(((uint64_t)__double_as_longlong(v) >> 52) & 0x7FF);
        m = e > m ? e : m;
    }
    red[32 + q * 32 + row] = m;
    __syncthreads();
    if (t < 32) {
        int mm = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int x = red[32 + k * 32 + t];
            mm = x > mm ? x : mm;
        }
        red[t] = mm;
        if (write) E[t] = mm;
    }
    __syncthreads();
}

// Digits of the staged tile: thread (row, q) owns 16 consecutive k of one row and writes
// one 16-byte lane chunk per slice plane (q = t / 32 + 8 p for the TK / 128 passes p).
template <bool KC, int TK = 128>
__device__ __forceinline__ void split_emit(const double *tile, int64_t R, const int32_t *__restrict__ Erow, int b,
                                           int d, int8_t *__restrict__ S, int64_t Rb, int64_t Kb, int64_t rb,
                                           int64_t kt, int j0 = 0, int j1 = 1 << 30) {
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
        emit_digits16(mx, bias, negm, b, d, dst, plane, j0, j1);
    }
}

// split_emit for the paired-digit layout: pair planes at S (standard layout, Rb x Kb blocks
// per plane), single-digit slots at S + P planes ([rb][slot][kb] blocks); planes g0 .. g1 - 1
// of the P + Q.
template <bool KC>
__device__ __forceinline__ void split_emit_pairs(const double *tile, int64_t R, const int32_t *__restrict__ Erow,
                                                 int b, int d, int8_t *__restrict__ S, int64_t Rb, int64_t Kb,
                                                 int64_t rb, int64_t kt, bool reverse_q, int g0, int g1) {
    constexpr int TK = 128;
    constexpr int LDW = KC ? (TK + 1) : (32 + 1);
    const PairCfg pc = pair_cfg(b, d);
    const int t = threadIdx.x;
    const int64_t r0 = rb * 32;
    const int row = t & 31;
    const int emax = (r0 + row < R) ? Erow[row] : 0;
    const int64_t plane = Rb * Kb * (int64_t)kBlkBytes;
    const int q = t >> 5;   // k in [16q, 16q + 16) of the tile
    const int64_t kb = kt * (TK / 32) + (q >> 1);
    if (kb >= Kb) return;   // no barrier below
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
    const int64_t intra = (q & 1) * 512 + row * 16;
    int8_t *dstS = S + (rb * Kb + kb) * (int64_t)kBlkBytes + intra;
    int8_t *dstX = S + pc.P * plane + ((rb * pc.Q) * Kb + kb) * (int64_t)kBlkBytes + intra;
    emit_pairs16(mx, bias, negm, b, d, dstS, plane, dstX, Kb * (int64_t)kBlkBytes, reverse_q, g0, g1);
}

template <bool KC>
__device__ __forceinline__ void split_tile(const double *__restrict__ X, int64_t sr, int64_t sk, int64_t R,
                                           int64_t K, const int32_t *__restrict__ Erow, int b, int d,
                                           int8_t *__restrict__ S, int64_t Rb, int64_t Kb, int64_t rb,
                                           int64_t kt, double *tile, int j0 = 0, int j1 = 1 << 30) {
    double v[16];
    split_gather<KC>(X, sr, sk, R, K, rb, kt, v);
    split_stage<KC>(v, tile);
    __syncthreads();
    split_emit<KC>(tile, R, Erow, b, d, S, Rb, Kb, rb, kt, j0, j1);
}

// ------------------------------------------------------------------ epilogue math
__device__ __forceinline__ int clz128(u128 a) {
    const uint64_t h = (uint64_t)(a >> 64), l = (uint64_t)a;
    if (h) return __builtin_clzll(h);
    if (l) return 64 + __builtin_clzll(l);
    return 128;
}

// |T| > 0 normalised into a 64-bit head: m = the top 64 bits of |T| with its leading one at
// bit 63, nb = the bit length of |T|, sticky = whether bits of |T| lie below the head.
struct Head {
    uint64_t m;
    int nb;
    bool sticky;
};

__device__ __forceinline__ Head head_of(uint64_t hi, uint64_t lo) {
    Head r;
    if (hi) {
        const int n = __builtin_clzll(hi);
        r.m = n ? (hi << n) | (lo >> (64 - n)) : hi;
        r.nb = 128 - n;
        r.sticky = (n ? (lo << n) : lo) != 0;
    } else {
        const int n = __builtin_clzll(lo);
        r.m = lo << n;
        r.nb = 64 - n;
        r.sticky = false;
    }
    return r;
}

// FP64 of sign * |T| * 2^(es - 2044 - b(d+1)) from the head of |T| (sgn: the sign bit).
// rounding 0: trunc53(|T|) with the exponent written into the bits as in
// include/oz_aie.h:124-146 (no over/underflow handling, as in the reference); rounding 1:
// |T| rounded to 53 bits, nearest-even, then scaled by ldexp.
__device__ __forceinline__ double head_to_double(Head h, uint64_t sgn, int es, int b, int d, int rounding) {
    if (rounding == 0) {
        const uint64_t mant = (h.m >> 11) - 0x0010000000000000ull;   // top 53 bits, implicit one removed
        const int64_t ex = h.nb - b * (d + 1) - 1;
        return __longlong_as_double((long long)(mant + sgn + ((uint64_t)(ex + 1023) << 52) +
                                                ((uint64_t)(int64_t)(es - 2044) << 52)));
    }
    uint64_t q;
    int e2;
    if (h.nb <= 53) {
        q = h.m >> (64 - h.nb);   // exact
        e2 = 0;
    } else {
        q = h.m >> 11;
        const bool half = (h.m >> 10) & 1, rest = (h.m & 0x3FF) != 0 || h.sticky;
        if (half && (rest || (q & 1))) q += 1;   // 2^53 after the carry is still exact
        e2 = h.nb - 53;
    }
    const double v = ldexp((double)q, e2 + es - 2044 - b * (d + 1));
    return sgn ? -v : v;
}

// The same for an exact T held in int64.
__device__ __forceinline__ double t64_to_double(int64_t t, int es, int b, int d, int rounding) {
    if (t == 0) return 0.0;
    const uint64_t sgn = t < 0 ? (1ull << 63) : 0ull;
    const uint64_t a = t < 0 ? (uint64_t)0 - (uint64_t)t : (uint64_t)t;
    return head_to_double(head_of(0, a), sgn, es, b, d, rounding);
}

// Exact integer T (scaled by 2^(es - 2044 - b(d+1))) to FP64.
// rounding 0: sign * trunc53(|T|) via the bit arithmetic of include/oz_aie.h:124-146
// (no over/underflow handling, as in the reference); rounding 1: nearest-even.
__device__ __forceinline__ double t_to_double(i128 t, int es, int b, int d, int rounding) {
    if (t == 0) return 0.0;
    const uint64_t sgn = t < 0 ? (1ull << 63) : 0ull;
    const u128 a = t < 0 ? (u128)(-t) : (u128)t;
    return head_to_double(head_of((uint64_t)(a >> 64), (uint64_t)a), sgn, es, b, d, rounding);
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
    if (g.last) {
        if (g.xadd) T += (i128)(g.xsign * (int64_t)g.xadd[(int64_t)row * g.ldx + col]);
        const int cb = g.cd ? g.cb : g.b, cd = g.cd ? g.cd : g.d;
        g.C[(int64_t)row * g.ldc + col] = t_to_double(T, es, cb, cd, g.rounding);
    } else {
        *tp = T;
    }
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

// The same store for EPI_C when the pass holds every bin (k0 = 0, NB = d) and b = B, with
// the paired-digit path's anti-diagonal term and conversion (cb, cd) when g.cd is set.
template <int D, int B>
__device__ __forceinline__ void epilogue_elem_c(const GemmLaunch &g, int row, int col, const int32_t (&vals)[D],
                                                int es) {
    if (row >= g.M || col >= g.N) return;
    const int cb = g.cd ? g.cb : B, cd = g.cd ? g.cd : D;
    if constexpr (D <= 31 / B + 1) {
        // one 64-bit chunk holds every bin: each bin is below 2^31 in magnitude, so |T| <
        // 2^31 * 2^(B (D - 1)) / (1 - 2^-B) <= 2^62.5 for B >= 2, and the int32 anti-diagonal
        // term keeps it inside int64: the whole product in int64
        if (g.first && g.last) {
            int64_t c = 0;
#pragma unroll
            for (int i = 0; i < D; i++) c += (int64_t)vals[D - 1 - i] << (B * i);
            if (g.xadd) c += g.xsign * (int64_t)g.xadd[(int64_t)row * g.ldx + col];
            g.C[(int64_t)row * g.ldc + col] = t64_to_double(c, es, cb, cd, g.rounding);
            return;
        }
    }
    i128 T = bins_to_T<D, B>(vals);
    i128 *tp = g.T + (int64_t)row * g.ldT + col;
    if (!g.first) T += *tp;
    if (g.last) {
        if (g.xadd) T += (i128)(g.xsign * (int64_t)g.xadd[(int64_t)row * g.ldx + col]);
        g.C[(int64_t)row * g.ldc + col] = t_to_double(T, es, cb, cd, g.rounding);
    } else {
        *tp = T;
    }
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

// Barrier between the waves of a workgroup after the wave's own LDS reads have completed
// (no memory fence, so global loads stay in flight across it); the inline-asm memory
// clobber keeps the compiler from moving LDS accesses across it.
__device__ __forceinline__ void lds_barrier() {
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
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
    if (epi == EPI_C && k0 == 0 && NB == g.d && g.b == 7 && !g.xadd && !g.cd) {
        wave_epilogue_t<NB, WM, WN, 7>(g, epi, k0, acc, scratch, lane, rbase, cbase);
        return;
    }
    if constexpr (NB >= 4) {   // pair planes of 3-bit digits (the reference's accurate presets)
        if (epi == EPI_C && k0 == 0 && NB == g.d && g.b == 6 && g.cd) {
            wave_epilogue_t<NB, WM, WN, 6>(g, epi, k0, acc, scratch, lane, rbase, cbase);
            return;
        }
    }
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

// Rolling variant (any d): the two LDS stages and copies of the LDS-copy kernel, but a
// single fragment set, refilled as it dies.  The LDS-copy kernel reads all d (WM + WN)
// fragments of a block right after its barrier, so every wave waits for that burst of LDS
// reads before its first MFMA.  Here the fragments of block b + 1 are read during block b's
// products, each into a fragment of block b right after that fragment's last use.
// Bin window [K0, K0 + NB): the products s_i * s_j with K0 <= i + j < K0 + NB (i, j < S =
// min(D, K0 + NB), the slices that reach the window) into NB bins; the whole triangle is
// the window [0, D).  Products of a d > 16 split, whose d bins do not fit in registers, run
// as several windows, each a pass that adds its bins into the exact T.
//   Order: rows i = 0 .. S-1 in turn, and in row i the columns j from the window's top down.
//   A_i dies at the end of row i; B_j at its use in row r_j = min(S-1, K0+NB-1-j), the last
//   row that reaches it (the first column of that row unless r_j = S-1).  Both are refilled
//   right there.  The next block begins with row 0 at its top column, whose B was refilled
//   at the start of the previous block; the latest-read fragment is needed only after the
//   other columns of row 0 (or, in a window that starts above 0, several rows later): at
//   least (d-1) WM WN MFMAs of slack for a whole triangle (14 for d = 8 on 1 x 2 wave
//   tiles), which covers an LDS read.
//   step b: wait for block b + 1's copies, barrier (every wave has read block b, whose
//           fragments were read during step b - 1), copy block b + 2 into block b's stage,
//           then the products of block b with the rolling reads of block b + 1.
// The last step reads a stale stage into fragments that are never multiplied.
template <int D, int K0, int NB, int WM, int WN>
__host__ __device__ constexpr int window_slices() { return D < K0 + NB ? D : K0 + NB; }

template <int D, int K0, int NB, int WM, int WN>
__host__ __device__ constexpr int window_tile_bytes() {
    return 2 * FusedCfg<window_slices<D, K0, NB, WM, WN>(), WM, WN>::STAGE;
}

template <int D, int K0, int NB, int WM, int WN>
__device__ __forceinline__ void gemm_window_tile(const GemmLaunch &g, int tm, int tn, int epi, int8_t *smem) {
    constexpr int S = window_slices<D, K0, NB, WM, WN>();
    static_assert(K0 < D && NB >= 1, "the window starts at a bin of the product");
    using F = FusedCfg<S, WM, WN>;
    constexpr int BA = F::BA, BB = F::BB, NBLK = F::NBLK, NLD = F::NLD, NMIN = F::NMIN;
    static_assert(window_tile_bytes<D, K0, NB, WM, WN>() >= 4 * 4 * wave_scratch_words<NB, WM, WN>(),
                  "epilogue scratch");

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
        if (bc < S * BA) {
            const int j = bc / BA, mb = bc % BA;
            src[it] = g.SA + j * planeA + (rbA0 + mb) * g.Kb * kBlkBytes;
        } else {
            const int bb = bc - S * BA;
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
    auto readA = [&](v4i (&a)[WM], const int8_t *st, int j) {
#pragma unroll
        for (int mi = 0; mi < WM; mi++)
            a[mi] = *reinterpret_cast<const v4i *>(st + (j * BA + wm * WM + mi) * kBlkBytes + lane * 16);
    };
    auto readB = [&](v4i (&b)[WN], const int8_t *st, int j) {
#pragma unroll
        for (int ni = 0; ni < WN; ni++)
            b[ni] = *reinterpret_cast<const v4i *>(st + (S * BA + j * BB + wn * WN + ni) * kBlkBytes + lane * 16);
    };

    v16i acc[NB][WM][WN];
#pragma unroll
    for (int k = 0; k < NB; k++)
#pragma unroll
        for (int mi = 0; mi < WM; mi++)
#pragma unroll
            for (int ni = 0; ni < WN; ni++) acc[k][mi][ni] = v16i{};

    const int64_t kb0 = g.kb0, nk = g.kb1 - g.kb0;
    v4i af[S][WM], bf[S][WN];
    if (nk > 0) {
        issue(0, kb0);
        s_waitcnt_vm<0>();
        lds_barrier();
#pragma unroll
        for (int j = 0; j < S; j++) {
            readA(af[j], smem, j);
            readB(bf[j], smem, j);
        }
        if (nk > 1) issue(1, kb0 + 1);
    }
    for (int64_t b = 0; b < nk; b++) {
        __builtin_amdgcn_sched_barrier(0);
        s_waitcnt_vm<0>();   // block b + 1 (the only copies in flight) is in LDS
        lds_barrier();       // ... for every wave; every wave's reads of block b are done
        if (b + 2 < nk) issue((int)(b & 1), kb0 + b + 2);
        const int8_t *nx = smem + (int)((b + 1) & 1) * F::STAGE;
#pragma unroll
        for (int i = 0; i < S; i++) {
            constexpr int kTop = K0 + NB - 1;
#pragma unroll
            for (int j = S - 1; j >= 0; j--) {
                if (i + j < K0 || i + j > kTop) continue;   // outside the window (compile time)
#pragma unroll
                for (int mi = 0; mi < WM; mi++)
#pragma unroll
                    for (int ni = 0; ni < WN; ni++)
                        acc[i + j - K0][mi][ni] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af[i][mi], bf[j][ni],
                                                                                        acc[i + j - K0][mi][ni], 0, 0, 0);
                if (i == (S - 1 < kTop - j ? S - 1 : kTop - j)) {   // the last use of B_j in this block
                    __builtin_amdgcn_sched_barrier(0);
                    readB(bf[j], nx, j);
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
            __builtin_amdgcn_sched_barrier(0);
            readA(af[i], nx, i);   // the last use of A_i was this row
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    __syncthreads();

    int32_t *scratch = reinterpret_cast<int32_t *>(smem) + w * wave_scratch_words<NB, WM, WN>();
    wave_epilogue<NB, WM, WN>(g, epi, K0, acc, scratch, lane, rbase, cbase, er_v, ec_v);
}

template <int D, int WM, int WN>
__device__ __forceinline__ void gemm_roll_tile(const GemmLaunch &g, int tm, int tn, int epi, int8_t *smem) {
    gemm_window_tile<D, 0, D, WM, WN>(g, tm, tn, epi, smem);
}

// Staged variant (d <= 4 on a 128 x 128 workgroup tile, d = 9 on 64 x 64): the slice blocks go HBM -> LDS
// (global_load_lds_dwordx4, each wave bringing NBLK / 4 whole 1 KiB blocks) into one of
// S stages, so S - 1 contraction blocks are in flight, and the fragments of block b + 1
// are read from LDS while the triangle of block b runs:
//   step b: wait for the wave's own loads of block b + 1 (loads complete in order, so the
//           later blocks' loads stay in flight), barrier (block b + 1 is in LDS; every wave
//           has read block b), issue block b + S into the stage block b used, read block
//           b + 1, multiply block b (its fragments are in registers since step b - 1).
// The waits are builtins: the compiler's own wait placement does not read the inline-asm
// barrier and would otherwise wait on the new reads before the first MFMA.  The steady
// loop has no branch; a checked tail pass issues no load past the pass's last block and
// waits for every load, so the epilogue's LDS scratch follows the last barrier safely.
// Any number of contraction blocks per pass.
// Stages S: as many as fit in 160 KiB of LDS, even, at most MAXS (8 by default: d = 1, 2: 8;
// d = 3: 6; d = 4 and d = 9 on 64 x 64: 4).  Measured against four stages: d = 1 8-22%
// faster, d = 2 and 3 unchanged.  A smaller MAXS leaves LDS for a second workgroup per CU.
template <int D, int WM, int WN, int MAXS = 8>
__host__ __device__ constexpr int staged_stages() {
    const int s = (160 * 1024 / FusedCfg<D, WM, WN>::STAGE) & ~1;
    return s > MAXS ? MAXS : s;
}

template <int D, int WM, int WN, int MAXS = 8>
__host__ __device__ constexpr int staged_tile_bytes() {
    return staged_stages<D, WM, WN, MAXS>() * FusedCfg<D, WM, WN>::STAGE;
}

template <int D, int WM, int WN, int MAXS = 8>
__device__ __forceinline__ void gemm_staged_tile(const GemmLaunch &g, int tm, int tn, int epi, int8_t *smem) {
    using F = FusedCfg<D, WM, WN>;
    constexpr int BA = F::BA, BB = F::BB, NBLK = F::NBLK, S = staged_stages<D, WM, WN, MAXS>();
    static_assert(S >= 4, "at least three blocks in flight");
    static_assert(NBLK % 4 == 0, "every wave brings the same number of blocks");
    static_assert(staged_tile_bytes<D, WM, WN, MAXS>() >= 4 * 4 * wave_scratch_words<D, WM, WN>(), "epilogue scratch");
    constexpr int PER = NBLK / 4;
    constexpr int VM = PER * (S - 2);
    static_assert(VM < 64, "vmcnt range");
    constexpr int kWaitSteady = (VM & 15) | ((VM >> 4) << 14) | (7 << 4);   // vmcnt(VM) lgkmcnt(0)
    constexpr int kWaitAll = 7 << 4;                                         // vmcnt(0) lgkmcnt(0)
    constexpr int NMMA = D * (D + 1) / 2 * WM * WN;

    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = w & 1, wn = w >> 1;
    const int64_t rbA0 = (int64_t)tm * BA, rbB0 = (int64_t)tn * BB;
    const int64_t planeA = g.RbA * g.Kb * kBlkBytes, planeB = g.RbB * g.Kb * kBlkBytes;
    const int rbase = tm * 64 * WM + wm * 32 * WM, cbase = tn * 64 * WN + wn * 32 * WN;
    int er_v, ec_v;
    load_tile_exps(g, epi, lane, rbase, 32 * WM, cbase, 32 * WN, er_v, ec_v);

    const int64_t kb0 = g.kb0, nk = g.kb1 - g.kb0;
    const int8_t *src[PER];   // wave-uniform: block (w + 4 it) of the stage, at the pass's first block
#pragma unroll
    for (int it = 0; it < PER; it++) {
        const int bc = w + 4 * it;
        if (bc < D * BA) {
            const int j = bc / BA, mb = bc % BA;
            src[it] = g.SA + j * planeA + ((rbA0 + mb) * g.Kb + kb0) * kBlkBytes;
        } else {
            const int bb = bc - D * BA;
            const int j = bb / BB, nb = bb % BB;
            src[it] = g.SB + j * planeB + ((rbB0 + nb) * g.Kb + kb0) * kBlkBytes;
        }
    }
    auto issue = [&](int stage, int64_t b, auto check) {   // block b of the pass
        if (decltype(check)::value && b >= nk) return;
#pragma unroll
        for (int it = 0; it < PER; it++)
            __builtin_amdgcn_global_load_lds((const void *)(src[it] + b * kBlkBytes + lane * 16),
                                             (lds_void *)(smem + stage * F::STAGE + (w + 4 * it) * kBlkBytes), 16,
                                             0, 0);
    };
    auto sync = [&](int64_t b, auto check) {   // block b and all before it in LDS, then barrier
        if (!decltype(check)::value || b + S - 2 < nk)
            __builtin_amdgcn_s_waitcnt(kWaitSteady);
        else
            __builtin_amdgcn_s_waitcnt(kWaitAll);
        lds_barrier();
    };
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
    v4i af[2][D][WM], bf[2][D][WN];
    auto step = [&](int s, int64_t b, auto check) {
        const int f = s & 1;
        __builtin_amdgcn_sched_barrier(0);
        sync(b + 1, check);
        issue(s, b + S, check);
        readf(af[f ^ 1], bf[f ^ 1], (s + 1) % S);   // past the last block: never multiplied
#pragma unroll
        for (int i = 0; i < D; i++)
#pragma unroll
            for (int j = 0; i + j < D; j++)
#pragma unroll
                for (int mi = 0; mi < WM; mi++)
#pragma unroll
                    for (int ni = 0; ni < WN; ni++)
                        acc[i + j][mi][ni] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af[f][i][mi], bf[f][j][ni],
                                                                                   acc[i + j][mi][ni], 0, 0, 0);
        __builtin_amdgcn_sched_group_barrier(0x100, D * (WM + WN), 0);   // the LDS reads first,
        __builtin_amdgcn_sched_group_barrier(0x008, NMMA, 0);            // then the MFMAs
        __builtin_amdgcn_sched_barrier(0);
    };

    using Checked = std::integral_constant<bool, true>;
    using Steady = std::integral_constant<bool, false>;
#pragma unroll
    for (int s = 0; s < S - 1; s++) issue(s, s, Checked{});
    sync(0, Checked{});
    issue(S - 1, S - 1, Checked{});
    readf(af[0], bf[0], 0);
    int64_t t = 0;
    for (; t + 2 * S <= nk; t += S) {   // every step of the pass loads a block that exists
#pragma unroll
        for (int s = 0; s < S; s++) step(s, t + s, Steady{});
    }
    for (; t < nk; t += S) {
#pragma unroll
        for (int s = 0; s < S; s++)
            if (t + s < nk) step(s, t + s, Checked{});
    }
    __syncthreads();

    int32_t *scratch = reinterpret_cast<int32_t *>(smem) + w * wave_scratch_words<D, WM, WN>();
    wave_epilogue<D, WM, WN>(g, epi, 0, acc, scratch, lane, rbase, cbase, er_v, ec_v);
}

constexpr int kCrtStages = 4;   // LDS stages of the modulus GEMM's contraction loop

// Waves per SIMD a single-product tile shape of WM x WN accumulator tiles per wave leaves
// room for: the 256 x 256 tile's 16 take half the registers of one wave, the smaller ones
// fit two and four.
constexpr int crt_occupancy(int tiles) { return tiles >= 8 ? 1 : (tiles >= 2 ? 2 : 4); }

// ------------------------------------------------------------------ single-product LDS loop
// (the CRT modulus GEMM and the anti-diagonal GEMM of the paired-digit path)
// Contraction loop of one (64 WM) x (64 WN) tile, the workgroup's 2 WM A and 2 WN B blocks
// of each contraction block shared through LDS.  The global loads write LDS themselves
// (global_load_lds_dwordx4: lane l's 16 bytes land at the wave-uniform LDS address + 16 l,
// which is a block's own layout), so there is no register staging and no LDS store; each
// wave brings PER whole 1 KiB blocks of a contraction block.  S stages keep blocks
// b + 1 .. b + S - 1 in flight while block b is multiplied, and the fragments of block b + 1
// are read from LDS while the MFMAs of block b run:
//   step b: wait for the wave's own loads of block b + 1 (loads complete in order, so the
//           S - 2 later blocks stay in flight), barrier (block b + 1 is in LDS; every wave
//           has read block b), issue block b + S into the stage block b used, read block
//           b + 1, multiply block b (its fragments are in registers since step b - 1).
// No load goes past block nk - 1: where fewer than S - 2 blocks follow, the wait is for
// every load, so after the last step nothing is in flight, and the only LDS read after
// the last barrier is the last step's read of a stage into fragments that are never
// multiplied: prime() for another plane may follow at once.  The conditions are uniform
// over the workgroup, so every wave meets the same barriers.
// setup: A0 / B0 = the tile's first row block of one modulus plane of A / B (rows of
// rowstride bytes), nk >= 1 contraction blocks; prime: issue the first S - 1 blocks;
// run: the products; advance: move to another plane, for the next prime.
template <int WM, int WN, int S>
struct CrtTileLoop {
    static constexpr int BA = 2 * WM, BB = 2 * WN, NB = BA + BB;   // blocks per contraction block
    static constexpr int PER = NB / 4;                            // blocks each wave brings in
    static_assert(NB % 4 == 0, "blocks must split over four waves");
    static_assert(S >= 3 && S % 2 == 0, "two fragment sets alternate over an even number of stages");
    static constexpr int VM = PER * (S - 2);   // loads of the wave still allowed in flight
    static_assert(VM < 64, "vmcnt range");
    // s_waitcnt vmcnt(VM) lgkmcnt(0) and vmcnt(0) lgkmcnt(0) as builtins: the compiler's own
    // wait placement sees them (it does not read the inline asm of lds_barrier), so it knows
    // the fragments of the previous step have arrived and adds no wait on the LDS reads
    // issued after the barrier.
    static constexpr int kWaitSteady = (VM & 15) | ((VM >> 4) << 14) | (7 << 4);
    static constexpr int kWaitAll = 7 << 4;
    const int8_t *src[PER];   // wave-uniform: first byte of each block row the wave brings
    int64_t nk;
    int lane, w, wm, wn;

    __device__ __forceinline__ void setup(const int8_t *A0, const int8_t *B0, int64_t rowstride, int64_t nblk) {
        nk = nblk;
        lane = threadIdx.x & 63;
        w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        wm = w & 1;
        wn = w >> 1;
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const int bi = w + 4 * i;   // block bi of the contraction block: A rows, then B columns
            src[i] = bi < BA ? A0 + (int64_t)bi * rowstride : B0 + (int64_t)(bi - BA) * rowstride;
        }
    }
    __device__ __forceinline__ void advance(int64_t offA, int64_t offB) {
#pragma unroll
        for (int i = 0; i < PER; i++) src[i] += (w + 4 * i < BA) ? offA : offB;
    }
    template <bool CHECK = true>
    __device__ __forceinline__ void issue(int8_t *smem, int slot, int64_t kb) {
        if (CHECK && kb >= nk) return;
#pragma unroll
        for (int i = 0; i < PER; i++)
            __builtin_amdgcn_global_load_lds(
                (const void *)(src[i] + kb * kBlkBytes + lane * 16),
                (__attribute__((address_space(3))) void *)(smem + (slot * NB + w + 4 * i) * kBlkBytes), 16, 0, 0);
    }
    __device__ __forceinline__ void prime(int8_t *smem) {
#pragma unroll
        for (int s = 0; s < S - 1; s++) issue(smem, s, s);
    }
    // wait until block kb and all before it are in LDS, then barrier
    template <bool CHECK = true>
    __device__ __forceinline__ void sync(int64_t kb) {
        if (!CHECK || kb + S - 2 < nk)
            __builtin_amdgcn_s_waitcnt(kWaitSteady);
        else
            __builtin_amdgcn_s_waitcnt(kWaitAll);
        lds_barrier();
    }
    __device__ __forceinline__ void read(const int8_t *st, v4i (&af)[WM], v4i (&bf)[WN]) {
#pragma unroll
        for (int mi = 0; mi < WM; mi++)
            af[mi] = *reinterpret_cast<const v4i *>(st + (wm * WM + mi) * kBlkBytes + lane * 16);
#pragma unroll
        for (int ni = 0; ni < WN; ni++)
            bf[ni] = *reinterpret_cast<const v4i *>(st + (BA + wn * WN + ni) * kBlkBytes + lane * 16);
    }
    // Step b (stage s = b mod S, fragment set s & 1).  CHECK = false: blocks b + S - 1 and
    // b + S exist, so the step has no branch; the LDS reads and the MFMAs then share one
    // basic block and the compiler adds no wait between them.  The read of block b + 1 is
    // made even past the last block: it reads a stage (possibly while the next plane's
    // first blocks land in it) into fragments that are never multiplied.
    template <bool CHECK>
    __device__ __forceinline__ void step(int s, int64_t b, int8_t *smem, v4i (&af)[2][WM], v4i (&bf)[2][WN],
                                         v16i (&acc)[WM][WN]) {
        const int f = s & 1;
        __builtin_amdgcn_sched_barrier(0);
        sync<CHECK>(b + 1);
        issue<CHECK>(smem, s, b + S);
        read(smem + ((s + 1) % S) * NB * kBlkBytes, af[f ^ 1], bf[f ^ 1]);
#pragma unroll
        for (int mi = 0; mi < WM; mi++)
#pragma unroll
            for (int ni = 0; ni < WN; ni++)
                acc[mi][ni] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af[f][mi], bf[f][ni], acc[mi][ni], 0, 0, 0);
        __builtin_amdgcn_sched_group_barrier(0x100, WM + WN, 0);   // the LDS reads first,
        __builtin_amdgcn_sched_group_barrier(0x008, WM * WN, 0);   // then the MFMAs
        __builtin_amdgcn_sched_barrier(0);
    }
    __device__ __forceinline__ void run(int8_t *smem, v16i (&acc)[WM][WN]) {
#pragma unroll
        for (int mi = 0; mi < WM; mi++)
#pragma unroll
            for (int ni = 0; ni < WN; ni++) acc[mi][ni] = v16i{};
        v4i af[2][WM], bf[2][WN];
        sync(0);
        issue(smem, S - 1, S - 1);
        read(smem, af[0], bf[0]);
        int64_t t = 0;
        for (; t + 2 * S <= nk; t += S) {   // every step of the pass loads a block that exists
#pragma unroll
            for (int s = 0; s < S; s++) step<false>(s, t + s, smem, af, bf, acc);
        }
        for (; t < nk; t += S) {
#pragma unroll
            for (int s = 0; s < S; s++)
                if (t + s < nk) step<true>(s, t + s, smem, af, bf, acc);
        }
    }
};

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

// ------------------------------------------------------------------ persistent launches
// (oz_grouped.hip: the split scheme's grouped launch; oz_crt.hip: the CRT mode's)
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

// Exponent maxima of the 32 rows of row block rb over the whole contraction axis: every
// lane returns the maximum of its row l % 32 (0 for rows at or beyond R).
template <bool KC>
__device__ __forceinline__ int wave_rowmax(const GroupOperand &o, int64_t rb, double *wt, int lane) {
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
            // the rest (fewer than 512 indices) as one more round of 64 loads in flight, the
            // indices clamped to K - 1 (a repeated element leaves the maximum unchanged)
            {
                double a[8][8];
#pragma unroll
                for (int i = 0; i < 8; i++)
#pragma unroll
                    for (int u = 0; u < 8; u++) a[i][u] = rowp[i][min(k + 64 * u, o.K - 1)];
#pragma unroll
                for (int i = 0; i < 8; i++)
#pragma unroll
                    for (int u = 0; u < 8; u++) m[gi * 8 + i] = max(m[gi * 8 + i], field_of(a[i][u]));
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
        {
            // the rest (fewer than 128 indices) with all its loads in flight, indices clamped
            // to K - 1 (a repeated element leaves the maximum unchanged)
            double a[64];
#pragma unroll
            for (int u = 0; u < 64; u++) a[u] = col[min(k + 2 * u, o.K - 1) * o.sk];
#pragma unroll
            for (int u = 0; u < 64; u++) m = max(m, field_of(a[u]));
        }
        emax = max(m, __shfl_xor(m, 32, 64));
    }
    if (r0 + (lane & 31) >= o.R) emax = 0;
    return __shfl(emax, lane & 31, 64);
}

// Every 64-index tile of the contraction axis of row block rb staged through the wave's
// LDS tile, `chunk(wt, row, c0, kb)` called for each 16-element chunk the lane owns (row
// l % 32, elements wt[row][c0 .. c0 + 16), contraction block kb; c0 = 32 (kb % 2) +
// 16 (l / 32)).  Kb may be odd (the last tile then holds one block).
template <bool KC, typename Chunk>
__device__ __forceinline__ void wave_tiles(const GroupOperand &o, int64_t rb, double *wt, int lane, Chunk chunk) {
    const int64_t r0 = rb * 32;
    const int64_t nt = (o.Kb + 1) / 2;
    const int row = lane & 31, half = lane >> 5;
    auto emit_tile = [&](int64_t t) {
#pragma unroll
        for (int blk = 0; blk < 2; blk++) {
            if (2 * t + blk >= o.Kb) break;   // wave-uniform
            chunk(wt, row, 32 * blk + 16 * half, 2 * t + blk);
        }
    };
    double v0[32], v1[32];
    wave_gather<KC>(o, r0, 0, lane, v0);
    if (nt > 1) wave_gather<KC>(o, r0, 64, lane, v1);
    for (int64_t t = 0; t < nt; t += 2) {
        wave_stage<KC>(v0, wt, lane);
        if (t + 2 < nt) wave_gather<KC>(o, r0, (t + 2) * 64, lane, v0);
        emit_tile(t);
        if (t + 1 < nt) {
            wave_stage<KC>(v1, wt, lane);
            if (t + 3 < nt) wave_gather<KC>(o, r0, (t + 3) * 64, lane, v1);
            emit_tile(t + 1);
        }
    }
}

// One row block: exponent maxima of its 32 rows (into o.E), then `emit(wt, row, c0,
// emax, dst)` for each 16-element chunk the lane owns (dst: the lane's 16-byte chunk in
// plane 0 of block (rb, kb)).
template <bool KC, typename Emit>
__device__ __forceinline__ void wave_rowblock_t(const GroupOperand &o, int64_t rb, double *wt, int lane, Emit emit) {
    const int emax = wave_rowmax<KC>(o, rb, wt, lane);
    if (lane < 32) o.E[rb * 32 + lane] = emax;
    wave_tiles<KC>(o, rb, wt, lane, [&](const double *w_, int row, int c0, int64_t kb) {
        emit(w_, row, c0, emax, o.S + (rb * o.Kb + kb) * (int64_t)kBlkBytes + lane * 16);
    });
}

// The split scheme's row block: d slice planes of b-bit digits.
template <bool KC>
__device__ __forceinline__ void wave_rowblock(const GroupOperand &o, int64_t rb, int b, int d, double *wt, int lane) {
    const int64_t plane = o.Rb * o.Kb * (int64_t)kBlkBytes;
    wave_rowblock_t<KC>(o, rb, wt, lane, [&](const double *w_, int row, int c0, int emax, int8_t *dst) {
        wave_emit_chunk(w_, row, c0, emax, b, d, dst, plane);
    });
}

struct PersistCtl {
    int arrive;     // workgroups through the phase barrier
    int done;       // workgroups finished (the last one resets the launch's words)
    int error;      // this launch: a barrier wait exceeded its bound (every waiter leaves)
    int reported;   // sticky across launches until ozaie_group_status reads and clears it
    int pad[60];
};

constexpr int kSpinLimit = 1 << 20;   // barrier polls (each a device-coherent load + sleep)

// Grid barrier: every wave makes its stores visible at agent scope, one thread per
// workgroup arrives and polls until `target` arrivals (G for a launch's first barrier, 2 G
// for its second: the counter is reset only at the launch's exit); a wait longer than
// kSpinLimit polls (or an error raised by another workgroup) sets the error flag and
// releases everyone.  Returns false on error.
__device__ __forceinline__ bool grid_barrier(PersistCtl *ctl, int target, int *flag_lds) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0) {
        int ok = 1;
        __hip_atomic_fetch_add(&ctl->arrive, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        for (int spin = 0;; spin++) {
            if (__hip_atomic_load(&ctl->arrive, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target) break;
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

// Exit of a persistent launch: the last workgroup hands the launch's error to the sticky
// report word and resets the launch's words, so a failed launch does not fail the next one.
__device__ __forceinline__ void persist_exit(PersistCtl *ctl, int G) {
    __syncthreads();
    if (threadIdx.x == 0) {
        const int old = __hip_atomic_fetch_add(&ctl->done, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (old == G - 1) {
            if (__hip_atomic_load(&ctl->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                __hip_atomic_store(&ctl->reported, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&ctl->error, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&ctl->arrive, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&ctl->done, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

}  // namespace oz
