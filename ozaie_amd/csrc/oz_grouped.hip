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

    persist_exit(ctl, G);
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
