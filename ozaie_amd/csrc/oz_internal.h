// oz_internal.h -- shared declarations between the HIP kernels and the C-ABI layer.
//
// Data layout in HBM (all device buffers are owned by a plan):
//   slices  S[j][rb][kb][1024] int8 for j < d, rb < Rp/32, kb < Kp/32.  One 1 KiB block
//           holds 32 rows x 32 contraction indices, stored so that lane l of a wave
//           reads bytes [16 l, 16 l + 16): row = l % 32, k = 16 (l / 32) + [0, 16).
//           That is exactly the operand register of v_mfma_i32_32x32x32_i8, so a block
//           moves HBM -> LDS -> VGPR with one 16-byte access per lane at each step.
//   E       int32 per row of A (Mp entries) and per column of B (Np entries): the
//           maximum biased exponent field along the contraction axis.
//   T       int128 per element of C (Mp x Np), only when a product needs several
//           passes (K chunks or bin groups).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace oz {

constexpr int kBlk = 32;          // rows / contraction indices per slice block
constexpr int kBlkBytes = 1024;   // bytes per slice block
constexpr int kRowPad = 128;      // rows of each operand are padded to this
constexpr int kMaxSlices = 32;
constexpr int kMaxBits = 7;

// One operand viewed as R rows x K contraction indices: element (r, k) at X[r*sr + k*sk].
struct OperandView {
    const double *X;
    int64_t sr, sk;
    int64_t R, K;
};

struct SliceBuf {
    int8_t *S;       // d planes
    int32_t *E;      // Rp exponents
    int64_t Rp, Kp;  // padded sizes
};

struct GemmLaunch {
    const int8_t *SA, *SB;
    const int32_t *Er, *Ec;
    int64_t RbA, RbB, Kb;   // blocks per plane (rows/32) and contraction blocks
    int64_t kb0, kb1;       // contraction block range of this pass
    int M, N;
    int b, d;
    int k0;                 // first bin of this pass (group kernel)
    int rounding;           // 0 truncate (reference), 1 nearest-even
    int first, last;        // pass flags: first writes T, last converts to C
    double *C;
    int64_t ldc;
    __int128 *T;            // Mp x Np (row-major, stride Np) when passes > 1
    int64_t ldT;
    int32_t *bins;          // debug: [d][M][N] int32 instead of C
    int64_t *paired;        // reference wire format: [M][N][2] int64 (optional)
    int K_true;             // K of the product (paired-format group sizes)
    int family;             // fused-kernel family (KERNEL_*), KERNEL_AUTO = per-d default
    // Paired-digit path (see launch_split_pairs): the bins are weighted with b = 2 b0 and
    // d = P above, the conversion uses (cb, cd); 0 means the same as (b, d).  At the last
    // pass the int32 anti-diagonal term xadd[row * ldx + col] is added (xsign = +1) or
    // subtracted (-1) before the conversion.
    int cb, cd;
    const int32_t *xadd;
    int64_t ldx;
    int xsign;
    int cus;                // compute units of the device (0: unknown), for the tile choice
};

// Fused-kernel families (same arithmetic, different data movement).
enum Family { KERNEL_AUTO = 0, KERNEL_LDS = 2, KERNEL_DIRECT = 3, KERNEL_PIPE = 4, KERNEL_STAGED = 5, KERNEL_ROLL = 6 };

// One operand of one problem of a grouped launch (R rows x K contraction indices).
struct GroupOperand {
    const double *X;
    int64_t sr, sk, R, K, Rp;
    int32_t *E;
    int8_t *S;
    int64_t Rb, Kb;
    int d, pad;         // slice count of this operand's problem (split scheme)
    uint64_t *st;       // CRT mode: chunk statistics [Rp][crt_chunks(Kb)]
};

// Tile configurations of the grouped fused GEMM (wave tile (32 WM) x (32 WN)).
enum GroupCfg { GCFG_D11 = 0, GCFG_D12 = 1, GCFG_D22 = 2, GCFG_L11 = 3, GCFG_L12 = 4, GCFG_COUNT = 5 };
bool group_cfg_supported(int d, int cfg);
int group_cfg_tile_m(int cfg);   // rows of C per workgroup
int group_cfg_tile_n(int cfg);   // columns of C per workgroup

// Paired-digit path for the small-digit presets (b <= 3, the reference's accurate mode):
// adjacent digits s_2I, s_2I+1 of the b d-bit digit string are paired into one int8 digit
// S_I = s_2I 2^b + s_2I+1 (|S_I| < 2^(2b) <= 64), so the triangle of d(d+1)/2 slice
// products becomes a triangle of P(P+1)/2 products of pairs plus one anti-diagonal of
// Q = H + 1 products of single digits that the truncation of the triangle cuts through:
//   d odd  (H = (d-1)/2, P = H):      T      = sum_{k<P} SB_k 2^(2b(P-1-k)) + X,
//                                      X  = sum_{I+J=H} s_2I t_2J   (high digits)
//   d even (H = (d-2)/2, P = H + 1):  T 2^b = sum_{k<P} SB_k 2^(2b(P-1-k)) - Z,
//                                      Z  = sum_{I+J=H} s_2I+1 t_2J+1 (low digits)
// with SB_k = sum_{I+J=k} S_I T_J.  T is the same exact integer as the plain triangle's,
// so C is bit-identical.  Plane layout of one operand (d planes in all): the P pair
// planes S[j][rb][kb][1024] as usual, then the Q single-digit planes in the layout
// X[rb][I][kb][1024] (for B stored in reverse order, I -> Q - 1 - I), so that the
// anti-diagonal is one int8 GEMM over Q * Kb contraction blocks.
struct PairCfg {
    int b, d, odd, H, P, Q;
};
__host__ __device__ inline PairCfg pair_cfg(int b, int d) {
    PairCfg c;
    c.b = b; c.d = d; c.odd = d & 1;
    c.H = c.odd ? (d - 1) / 2 : (d - 2) / 2;
    c.P = c.odd ? c.H : c.H + 1;
    c.Q = c.H + 1;
    return c;
}
// Eout (non-null only when Kp <= 128): the split forms the exponent maxima itself and
// stores them there, with no exponent pass before it.
hipError_t launch_split_pairs(const OperandView &op, const int32_t *E, int b, int d, int8_t *S, int64_t Rp,
                              int64_t Kp, bool reverse_q, hipStream_t st, int32_t *Eout = nullptr);
// X[r][c] (int32, row stride ldx) = sum over Q*Kb contraction blocks of the single-digit
// planes (Mp and Np multiples of 256; workgroup tiles of 256, 128 or 64 square by the grid
// they give on `cus` CUs).
hipError_t launch_gemm_antidiag(const int8_t *SxA, const int8_t *SxB, int64_t Mp, int64_t Np, int64_t nblk,
                                int32_t *X, int64_t ldx, int cus, hipStream_t st);
constexpr int kPairRowPad = 256;

// Grouped launchers: items are int4 {operand, a, b, c} for the split phase and
// int4 {problem, tile row, tile column, 0} for the GEMM phase.
hipError_t launch_group_maxexp(const GroupOperand *ops, const int4 *items, int n, hipStream_t st);
hipError_t launch_group_split(const GroupOperand *ops, const int4 *items, int n, int b, hipStream_t st);
hipError_t launch_group_gemm(const GemmLaunch *probs, const int4 *items, int n, int d, int cfg, hipStream_t st);
// Persistent grouped launch (one kernel: split phase, grid barrier, GEMM phase).  Phase A
// lists are per wave (items [offA[4w+v], offA[4w+v+1]) for wave v of workgroup w), phase B
// lists per workgroup (items [offB[w], offB[w+1])); phase A items are
// int4 {operand, row block, 0, 0}, phase B items int4 {problem, tile row, tile col,
// persist_code(d, cfg)}, each workgroup's list ordered so that items of one code are
// consecutive (one tile body per run of items, the next tile's first block prefetched
// inside a run).
// ctl is persist_ctl_bytes() of zero-initialised device memory, reset by the kernel itself
// at the end of every launch; its third int is a sticky error flag (barrier wait bound).
struct PersistLaunch {
    const GroupOperand *ops;
    const int4 *itemsA;
    const int *offA;
    const GemmLaunch *probs;
    const int4 *itemsB;
    const int *offB;
    void *ctl;
    int grid, b, rounding;
};
bool persist_cfg_supported(int d, int cfg);
inline int persist_code(int d, int cfg) { return d * 8 + cfg; }
int persist_ctl_bytes();
hipError_t launch_group_persistent(const PersistLaunch &L, hipStream_t st);
// Tile order inside a problem: bands of 8 tile-rows, so tiles that run together share
// rows of A (device tile mapping and the grouped planner).
__host__ __device__ inline void band_map(int lin, int ntm, int ntn, int &tm, int &tn) {
    constexpr int G = 8;
    const int band = lin / (G * ntn);
    const int inb = lin % (G * ntn);
    const int gh = (ntm - band * G) < G ? (ntm - band * G) : G;
    tm = band * G + inb % gh;
    tn = inb / gh;
}

// CRT mode (exact integer product from residue GEMMs modulo n <= 16 coprime moduli,
// rebuilt with the Chinese remainder theorem).  Constants are computed on the host.
constexpr int kCrtMaxMod = 20;
constexpr int kCrtDefaultMod = 16;   // OZAIE_MODE_CRT: every product below 2^127 (128-bit rebuild)
constexpr int kCrtNarrowMod = 16;    // up to this many moduli the rebuild runs in 128 bits, above in 192
constexpr int kCrtLimbs = 24;        // byte limbs of the 192-bit rebuild
constexpr int kCrtRowPad = 256;   // the modulus GEMM's workgroup tile is 256 x 256
constexpr int kCrtPF = 8;         // K pads to 32 kCrtPF (register ring depth of the per-wave modulus GEMM)
struct CrtConsts {
    int n, beta;               // beta 0: the norm rule (per-row shifts from norm bounds)
    uint64_t Mw[3];            // M = prod p, three words low first
    double p[kCrtMaxMod];      // modulus
    double inv[kCrtMaxMod];    // 1 / p
    double c31[kCrtMaxMod];    // 2^31 mod p
    // rebuild by byte dot products: Wd[G][i] = byte i of W_{4G + b} in byte b, where
    // W_p = (M/p) ((M/p)^-1 mod p) mod M; Qd[h][G] = byte h of q_{4G + b} in byte b, where
    // q_p = floor(2^16 fl(W_p / M)); mneg8[i] = byte i of 2^192 - M (its low 16 bytes are
    // those of 2^128 - M when M < 2^128).  Zero beyond the n moduli.
    uint32_t Wd[kCrtMaxMod / 4][kCrtLimbs];
    uint32_t Qd[2][kCrtMaxMod / 4];
    uint32_t mneg8[kCrtLimbs];
};
// Modulus list, bits per operand for (n, K) (or -1), and the constants.
int crt_modulus(int i);
int crt_beta(int n, int64_t K);
bool crt_norm_ok(int n, int64_t K);   // the norm rule's bound M 2^20 - K (2^20 + 2^40) > 0
void crt_consts(int n, int beta, CrtConsts &c);
// Per-row shifts sigma (q = round(|x| 2^(1075 - sigma))) come from statistics of chunks of
// kCrtChunk contraction indices: per chunk one word, (max exponent field << 40) | S_c with
// S_c the sum of squared 16-bit ceilings of the chunk's magnitudes; [Rp][crt_chunks(Kb)].
constexpr int kCrtChunk = 16;
__host__ __device__ inline int64_t crt_chunks(int64_t Kb) { return (Kb + 3) / 4 * 8; }   // 8 per split tile of 128
// One operand of the split phase: its view, its Rp rows' chunk statistics and shifts, and
// its residue planes.  Each launch below handles both operands (A and B) at once.
struct CrtSide {
    OperandView op;
    uint64_t *stats;   // [Rp][crt_chunks(Kb)]
    int32_t *sigma;    // [Rp]
    int8_t *S;         // residues [n][Rp / 32][Kb][1024]
    int64_t Rp;
};
hipError_t launch_crt_stats(const CrtSide &a, const CrtSide &b, int64_t Kp, hipStream_t st);
// sigma[r] for r < Rp from the statistics: the norm rule for c.beta = 0 (contraction
// length K), else sigma = 53 + E - c.beta.
hipError_t launch_crt_scale(const CrtSide &a, const CrtSide &b, int64_t nch, int64_t K, const CrtConsts &c,
                            hipStream_t st);
// Residues S[j][rb][kb][1024] (int8, slice-block layout) with shifts sigma.
hipError_t launch_crt_split(const CrtSide &a, const CrtSide &b, const CrtConsts &c, int64_t Kp, hipStream_t st);
// g_j = A_j B_j^T mod p_j for every modulus: R[j][Mp/4][Np] words of four rows, one signed
// byte each (congruent to g_j, |byte| <= p_j / 2).
struct CrtGemm {
    const int8_t *SA, *SB;
    int64_t Mp, Np, Kb;
    uint32_t *R;
    int variant;   // 0: fragments shared through LDS (256 x 256 tiles), 1: per-wave loads
                   // (OZAIE_OPT_KERNEL 3), 2: as 0 on 128 x 128 tiles, two workgroups per CU
};
hipError_t launch_crt_gemm(const CrtGemm &g, const CrtConsts &c, hipStream_t st);
// C = rebuild(R) * 2^(sigma_r + sigma_c - 2150), truncated (rounding 0) or nearest (1).
hipError_t launch_crt_rebuild(const uint32_t *R, int64_t Mp, int64_t Np, int64_t M, int64_t N, const int32_t *Er,
                              const int32_t *Ec, double *C, int64_t ldc, int rounding, const CrtConsts &c,
                              hipStream_t st);

// Grouped CRT: one descriptor per problem (C is filled in at run time).  Residue planes of A
// hold RbA row blocks of Kb contraction blocks (RbA = Mp / 32), those of B RbB; R holds n
// planes of (Mp / 4) x Np words.
struct CrtGroupProb {
    const int8_t *SA, *SB;
    const int32_t *Er, *Ec;
    uint32_t *R;
    double *C;
    int64_t RbA, RbB, Kb, Np, ldc;
    int M, N, beta, pad;   // beta 0: the norm rule, else sigma = 53 + E - beta
};
// Tile shapes of the grouped modulus GEMM: 256 x 256, 128 x 128, 64 x 64 per workgroup.
enum CrtCfg { CCFG_44 = 0, CCFG_22 = 1, CCFG_11 = 2, CCFG_COUNT = 3 };
int crt_group_cfg_wm(int cfg);   // row blocks of 64 per tile
int crt_group_cfg_wn(int cfg);   // column blocks of 64 per tile
// Items: int4 {operand, row block, contraction tile, 0} (split; the problem of operand o
// is o / 2), {problem, tile row, tile col, 0} (GEMM: workgroup x runs item x for moduli
// [mpw y, mpw y + mpw); problem -1 = empty slot) and {problem, first (row quad, four
// columns) group, 0, 0} (rebuild, 256 groups per item).
// Chunk statistics over the split items; shifts over items {operand, row block, 0, 0}.
hipError_t launch_group_crt_stats(const GroupOperand *ops, const int4 *items, int n, hipStream_t st);
hipError_t launch_group_crt_scale(const GroupOperand *ops, const CrtGroupProb *probs, const int4 *items, int n,
                                  const CrtConsts &c, hipStream_t st);
hipError_t launch_group_crt_split(const GroupOperand *ops, const CrtGroupProb *probs, const int4 *items, int n,
                                  const CrtConsts &c, hipStream_t st);
hipError_t launch_group_crt_gemm(const CrtGroupProb *probs, const int4 *items, int n, int cfg, int mpw,
                                 const CrtConsts &c, hipStream_t st);
hipError_t launch_group_crt_rebuild(const CrtGroupProb *probs, const int4 *items, int n, int rounding,
                                    const CrtConsts &c, hipStream_t st);
// The whole CRT group as ONE persistent launch (one workgroup per CU): phase A shifts
// (exponent maxima, norm bounds) + residues per 32-row block of every operand (per-wave lists itemsA[offA[4w+v],
// offA[4w+v+1]), items {operand, row block, 0, 0}), grid barrier, phase B modulus-GEMM
// tiles with every modulus (per-workgroup lists, items {problem, tile row, tile col,
// CrtCfg}, consecutive items of one shape), grid barrier, phase C rebuild (per-workgroup
// lists, items {problem, first group, 0, 0} as for the rebuild launch).  ctl: as
// PersistLaunch.ctl.
struct CrtPersistLaunch {
    const GroupOperand *ops;
    const CrtGroupProb *probs;
    const int4 *itemsA, *itemsB, *itemsC;
    const int *offA, *offB, *offC;
    void *ctl;
    int grid, rounding;
};
hipError_t launch_group_crt_persistent(const CrtPersistLaunch &L, const CrtConsts &c, hipStream_t st);
int crt_persist_per_cu();   // workgroups per CU of that launch (grid = this x CUs)

// Synthetic (U - 0.5) 2^e data for timing runs.
hipError_t launch_fill_random(double *p, int64_t n, uint32_t seed, hipStream_t st);

// Shared by the C-ABI translation units (oz_api.cpp).
int set_error(hipError_t e, const char *where);
void set_error_msg(const char *msg);
int64_t max_kchunk_blocks(int b, int d);

// Kernel launchers (oz_kernels.hip).  All enqueue on `st` and return hipError_t.
hipError_t launch_fill_i32(int32_t *p, int64_t n, int32_t v, hipStream_t st);

// Host memcpy split over OpenMP threads (oz_host.cpp).
void parallel_copy(void *dst, const void *src, size_t n);
hipError_t launch_maxexp(const OperandView &op, int32_t *E, int64_t Rp, hipStream_t st);
hipError_t launch_debug_convert(const int64_t *hi, const uint64_t *lo, const int32_t *es, int64_t n, int b, int d,
                                int rounding, double *out128, double *out64, hipStream_t st);
hipError_t launch_split(const OperandView &op, const int32_t *E, int b, int d, int8_t *S,
                        int64_t Rp, int64_t Kp, hipStream_t st, int32_t *Eout = nullptr);
// Fused kernel (all d bins in registers); returns hipErrorInvalidValue if d has no
// instantiation.  `wn` picks the wave tile (1: 32x32, 2: 32x64 per wave).
bool fused_supported(int d);
hipError_t launch_gemm_fused(const GemmLaunch &g, int epi, hipStream_t st);
// Bin-group kernel for any d <= kMaxSlices: bins [g.k0, g.k0 + 8).
constexpr int kGroupBins = 8;
hipError_t launch_gemm_group(const GemmLaunch &g, int epi, hipStream_t st);
// Bin-window kernel for the d > 16 splits it is built for (d = 18): bins [g.k0, g.k0 + 9)
// per pass, g.k0 a multiple of kWindowBins.
bool window_supported(int d);
constexpr int kWindowBins = 9;
hipError_t launch_gemm_window(const GemmLaunch &g, int epi, hipStream_t st);

enum Epi { EPI_C = 0, EPI_BINS = 1 };

// Reference paired-int64 wire format from int32 bins [d][M][N].
hipError_t launch_pack_paired(const int32_t *bins, int64_t M, int64_t N, int b, int d, int n0,
                              const int32_t *Er, const int32_t *Ec, int64_t *out, hipStream_t st);

}  // namespace oz
