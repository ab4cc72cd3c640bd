// oz_grouped.hip -- grouped products: per-phase launches over work-item tables built on
// the host (exponent maxima, split, fused slice GEMM per tile configuration) and the
// single persistent launch (split phase, grid barrier, GEMM phase).
#include "oz_device.h"

namespace oz {

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
                                                   const int4 *__restrict__ items, int b) {
    __shared__ double tile[kSplitLds / 8];
    const int4 it = items[blockIdx.x];
    const GroupOperand &o = ops[it.x];
    const int32_t *Erow = o.E + (int64_t)it.y * 32;
    if (o.sk == 1)
        split_tile<true>(o.X, o.sr, o.sk, o.R, o.K, Erow, b, o.d, o.S, o.Rb, o.Kb, it.y, it.z, tile);
    else
        split_tile<false>(o.X, o.sr, o.sk, o.R, o.K, Erow, b, o.d, o.S, o.Rb, o.Kb, it.y, it.z, tile);
}

hipError_t launch_group_maxexp(const GroupOperand *ops, const int4 *items, int n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_grp_maxexp, dim3((unsigned)n), dim3(256), 0, st, ops, items);
    return hipGetLastError();
}

hipError_t launch_group_split(const GroupOperand *ops, const int4 *items, int n, int b, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_grp_split, dim3((unsigned)n), dim3(256), 0, st, ops, items, b);
    return hipGetLastError();
}

template <int D, int WM, int WN, int FAM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1)))
void k_grp_gemm(const GemmLaunch *__restrict__ probs, const int4 *__restrict__ items) {
    // consecutive items (tiles of one problem in bands of tile-rows) go to one XCD
    const int4 it = items[xcd_remap(blockIdx.x, gridDim.x)];
    const GemmLaunch g = probs[it.x];
    constexpr int LDS = FAM == KERNEL_DIRECT   ? direct_tile_bytes<D>()
                        : FAM == KERNEL_STAGED ? staged_tile_bytes<D, WM, WN>()
                                               : lds_tile_bytes<D, WM, WN>();
    __shared__ __attribute__((aligned(1024))) int8_t smem[LDS];
    if constexpr (FAM == KERNEL_DIRECT) gemm_direct_tile<D, WM, WN>(g, it.y, it.z, EPI_C, smem);
    else if constexpr (FAM == KERNEL_STAGED) gemm_staged_tile<D, WM, WN>(g, it.y, it.z, EPI_C, smem);
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
        case GCFG_D22:   // 128 x 128 tiles: the staged tile body (six LDS stages at d = 3)
            if constexpr (D <= 4) hipLaunchKernelGGL((k_grp_gemm<D, 2, 2, KERNEL_STAGED>), grid, blk, 0, st, probs, items);
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
    int arrive;     // workgroups through the phase barrier
    int done;       // workgroups finished (the last one resets the launch's words)
    int error;      // this launch: a barrier wait exceeded its bound (every waiter leaves)
    int reported;   // sticky across launches until ozaie_group_status reads and clears it
    int pad[60];
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

// LDS of the persistent kernel: phase A's per-wave tiles (larger than any phase B epilogue
// scratch of the tile bodies it runs, d <= 9).
constexpr int kPersistLds = 4 * 32 * kWaveTileLd * 8;
static_assert(kPersistLds >= direct_tile_bytes<9>(), "phase B epilogue scratch");

// Phase B tiles [i0, i1) of one workgroup's list, all of one (D, WM, WN): fragments
// double-buffered in registers (direct loads), the first block of the next tile of the
// run loaded during the current tile's last block and epilogue.
template <int D, int WM, int WN>
__device__ __forceinline__ void persist_tiles(const GemmLaunch *__restrict__ probs, const int4 *__restrict__ itemsB,
                                              int i0, int i1, int b, int rounding, int8_t *smem) {
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = w & 1, wn = w >> 1;
    int32_t *scratch = reinterpret_cast<int32_t *>(smem) + w * wave_scratch_words<D, WM, WN>();
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
    cur = ptile_setup<WM, WN>(probs, itemsB[i0], wm, wn);
    load(a0, b0, cur, 0);
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

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1)))
void k_grp_persistent(const GroupOperand *__restrict__ ops, const int4 *__restrict__ itemsA,
                      const int *__restrict__ offA, const GemmLaunch *__restrict__ probs,
                      const int4 *__restrict__ itemsB, const int *__restrict__ offB, PersistCtl *ctl,
                      int b, int rounding) {
    __shared__ __attribute__((aligned(1024))) int8_t smem[kPersistLds];
    __shared__ int flag_lds;
    const int G = gridDim.x;

    // ---- phase A: exponent maximum + split (each operand with its problem's slice count),
    // one 32-row block per item, each wave working through its own item list (no
    // workgroup barriers)
    {
        const int lane = threadIdx.x & 63;
        const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        double *wt = reinterpret_cast<double *>(smem) + w * (32 * kWaveTileLd);
        const int slot = blockIdx.x * 4 + w;
        const int a0 = offA[slot], a1 = offA[slot + 1];
        for (int i = a0; i < a1; i++) {
            const int4 it = itemsA[i];
            const GroupOperand &o = ops[it.x];
            if (o.sk == 1) wave_rowblock<true>(o, it.y, b, o.d, wt, lane);
            else wave_rowblock<false>(o, it.y, b, o.d, wt, lane);
        }
    }

    const bool ok = grid_barrier(ctl, G, &flag_lds);

    // ---- phase B: fused slice-GEMM tiles in runs of one (d, tile shape)
    if (ok) {
        __syncthreads();   // phase A's LDS tiles are dead before the epilogue scratch reuses them
        const int i1 = offB[blockIdx.x + 1];
        for (int i = offB[blockIdx.x]; i < i1;) {
            const int code = itemsB[i].w;
            int j = i + 1;
            while (j < i1 && itemsB[j].w == code) j++;
            switch (code) {
                case 3 * 8 + GCFG_D11: persist_tiles<3, 1, 1>(probs, itemsB, i, j, b, rounding, smem); break;
                case 3 * 8 + GCFG_D12: persist_tiles<3, 1, 2>(probs, itemsB, i, j, b, rounding, smem); break;
                case 3 * 8 + GCFG_D22: persist_tiles<3, 2, 2>(probs, itemsB, i, j, b, rounding, smem); break;
                case 7 * 8 + GCFG_D11: persist_tiles<7, 1, 1>(probs, itemsB, i, j, b, rounding, smem); break;
                case 7 * 8 + GCFG_D12: persist_tiles<7, 1, 2>(probs, itemsB, i, j, b, rounding, smem); break;
                case 8 * 8 + GCFG_D11: persist_tiles<8, 1, 1>(probs, itemsB, i, j, b, rounding, smem); break;
                case 9 * 8 + GCFG_D11: persist_tiles<9, 1, 1>(probs, itemsB, i, j, b, rounding, smem); break;
                default: break;   // the host builds only the codes above
            }
            // the next run lays out the waves' epilogue scratch differently: every wave has
            // finished this run's epilogues before any starts the next run's
            __syncthreads();
            i = j;
        }
    }

    // ---- exit: the last workgroup hands this launch's error to the sticky report word and
    // resets the launch's words, so a failed launch does not fail the next one
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

bool persist_cfg_supported(int d, int cfg) {
    if (d != 3 && d != 7 && d != 8 && d != 9) return false;
    if (cfg == GCFG_D11) return true;
    if (cfg == GCFG_D12) return d == 3 || d == 7;   // d = 8 and 9 would spill registers
    if (cfg == GCFG_D22) return d == 3;
    return false;
}

int persist_ctl_bytes() { return (int)sizeof(PersistCtl); }

hipError_t launch_group_persistent(const PersistLaunch &L, hipStream_t st) {
    if (L.grid <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_grp_persistent, dim3((unsigned)L.grid), dim3(256), 0, st, L.ops, L.itemsA, L.offA, L.probs,
                       L.itemsB, L.offB, reinterpret_cast<PersistCtl *>(L.ctl), L.b, L.rounding);
    return hipGetLastError();
}

}  // namespace oz
