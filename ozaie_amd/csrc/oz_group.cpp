// oz_group.cpp -- grouped emulated DGEMM: many independent problems of uneven shape in
// one split launch and one fused-GEMM launch per tile configuration, captured once in a
// hipGraph and replayed; plus the autotuner that picks the tile configuration of each
// size class.  The CRT mode (flags bit 3) runs the same way with residue planes instead of
// slices: exponent max, residues, one modulus-GEMM launch per tile shape, rebuild.
//
// The reference has no batched entry point (include/oz_aie.h handles one square product
// per setup); this is new functionality on top of the same arithmetic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/ozaie.h"
#include "oz_internal.h"

using namespace oz;

namespace {

std::mutex g_tune_mu;
// g_tune[d][size class] = GroupCfg, -1 = not tuned (heuristic default)
int g_tune[kMaxSlices + 1][9];
bool g_tune_init = false;

void tune_init() {
    if (g_tune_init) return;
    for (auto &row : g_tune)
        for (int &v : row) v = -1;
    g_tune_init = true;
}

int bucket(int64_t x) { return x <= 256 ? 0 : (x <= 1024 ? 1 : 2); }
int size_class(int64_t M, int64_t N, int64_t K) { return 3 * bucket(std::max(M, N)) + bucket(K); }

// Heuristic when a class has not been tuned: small problems take small tiles (more
// workgroups), large ones the widest tile that fits in registers.
int default_cfg(int d, int64_t M, int64_t N) {
    const int64_t mn = std::max(M, N);
    if (d == 3) return mn <= 256 ? GCFG_D11 : (mn <= 1024 ? GCFG_D12 : GCFG_D22);
    if (d == 9) return mn <= 512 ? GCFG_D11 : GCFG_L12;
    return mn <= 512 ? GCFG_D11 : GCFG_D12;
}

int pick_cfg(int d, int64_t M, int64_t N, int64_t K, int force) {
    if (force >= 0 && group_cfg_supported(d, force)) return force;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    tune_init();
    const int t = g_tune[d][size_class(M, N, K)];
    if (t >= 0 && group_cfg_supported(d, t)) return t;
    return default_cfg(d, M, N);
}

int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

// Slice count per size class for groups created with slices = 0 (-1 = untuned: 9, the
// count whose maximum error matches native FP64 DGEMM at phi = 1 on every class measured).
// Set by ozaie_tune_set_slices, usually from the calibration in ozaie_amd.tune.
int g_tune_slices[9];
bool g_tune_slices_init = false;
constexpr int kDefaultGroupSlices = 9;

int pick_slices(int64_t M, int64_t N, int64_t K) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (!g_tune_slices_init) {
        for (int &v : g_tune_slices) v = -1;
        g_tune_slices_init = true;
    }
    const int t = g_tune_slices[size_class(M, N, K)];
    return t > 0 ? t : kDefaultGroupSlices;
}

// Index of a grouped-kernel slice count (3, 7, 8, 9) in the per-(d, tile) launch tables.
constexpr int kGroupSliceCounts[4] = {3, 7, 8, 9};
int dslot(int d) { return d == 3 ? 0 : (d == 7 ? 1 : (d == 8 ? 2 : (d == 9 ? 3 : -1))); }

// CRT mode: g_tune_crt[moduli][size class] = CrtCfg, -1 = untuned.
int g_tune_crt[kCrtMaxMod + 1][9];
bool g_tune_crt_init = false;

void tune_crt_init() {
    if (g_tune_crt_init) return;
    for (auto &row : g_tune_crt)
        for (int &v : row) v = -1;
    g_tune_crt_init = true;
}

// Untuned classes: the tile shape with the least padded work, the smaller tiles weighted
// by their lower MFMA rate (fragments per MFMA through LDS rise as the tile shrinks).
int default_crt_cfg(int64_t M, int64_t N) {
    static const double rate[CCFG_COUNT] = {1.0, 0.75, 0.4};
    int best = CCFG_44;
    double best_cost = 0.0;
    for (int c = 0; c < CCFG_COUNT; c++) {
        const int64_t tm = 64 * crt_group_cfg_wm(c), tn = 64 * crt_group_cfg_wn(c);
        const double cost = (double)round_up(M, tm) * (double)round_up(N, tn) / rate[c];
        if (c == 0 || cost < best_cost) { best = c; best_cost = cost; }
    }
    return best;
}

int pick_crt_cfg(int n, int64_t M, int64_t N, int64_t K, int force) {
    if (force >= 0 && force < CCFG_COUNT) return force;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    tune_crt_init();
    const int t = g_tune_crt[n][size_class(M, N, K)];
    if (t >= 0 && t < CCFG_COUNT) return t;
    return default_crt_cfg(M, N);
}

}  // namespace

struct ozaie_group_s {
    int count = 0, b = 0, d = 0, rounding = 0, flags = 0, device = 0;
    std::vector<int64_t> M, N, K;
    std::vector<int> dp;                   // slice count per problem (split scheme)
    std::vector<int> cfg;
    bool fallback = false;                 // slice count without grouped kernels
    std::vector<ozaie_plan> plans;         // fallback: one plan per problem
    // device
    char *ws = nullptr;                    // exponents + slices of every problem
    int32_t *E_all = nullptr;
    int64_t n_E = 0;
    GroupOperand *d_ops = nullptr;
    GemmLaunch *d_probs = nullptr;
    int4 *d_items_max = nullptr, *d_items_split = nullptr;
    int n_max = 0, n_split = 0;
    int4 *d_items_gemm[4][GCFG_COUNT] = {};   // per slice-count slot (dslot) and tile configuration
    int n_gemm[4][GCFG_COUNT] = {};
    int64_t tiles = 0;
    // host mirrors of the descriptors (pinned) and the pointers they hold
    GroupOperand *h_ops = nullptr;
    GemmLaunch *h_probs = nullptr;
    hipEvent_t upload_done = nullptr;
    bool uploaded = false;
    std::vector<const double *> lastA, lastB;
    std::vector<double *> lastC;
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // persistent single-launch mode (flags bit 2): tile configuration per problem in cfg
    bool persist = false;
    char *pbuf = nullptr;                  // item lists + control block
    PersistLaunch PL{};
    hipStream_t last_stream = nullptr;
    // CRT mode (flags bit 3): d = moduli, b = requested bits (0: largest per problem)
    bool crt = false;
    CrtConsts cc{};
    std::vector<int> beta;
    CrtGroupProb *d_cprobs = nullptr, *h_cprobs = nullptr;
    int4 *d_items_crt[CCFG_COUNT] = {};
    int n_crt[CCFG_COUNT] = {};
    int4 *d_items_rebuild = nullptr;
    int n_rebuild = 0;
    int crt_mpw[CCFG_COUNT] = {};         // moduli per GEMM workgroup, per tile shape
    CrtPersistLaunch CPL{};               // persistent CRT group (item lists in pbuf)
};

namespace {

void group_free(ozaie_group g) {
    if (!g) return;
    for (ozaie_plan p : g->plans) ozaie_plan_destroy(p);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->cap_stream) (void)hipStreamDestroy(g->cap_stream);
    if (g->upload_done) (void)hipEventDestroy(g->upload_done);
    if (g->ws) (void)hipFree(g->ws);
    if (g->pbuf) (void)hipFree(g->pbuf);
    if (g->d_ops) (void)hipFree(g->d_ops);
    if (g->d_probs) (void)hipFree(g->d_probs);
    if (g->d_items_max) (void)hipFree(g->d_items_max);
    if (g->d_items_split) (void)hipFree(g->d_items_split);
    for (auto &row : g->d_items_gemm)
        for (auto &p : row)
            if (p) (void)hipFree(p);
    if (g->h_ops) (void)hipHostFree(g->h_ops);
    if (g->h_probs) (void)hipHostFree(g->h_probs);
    if (g->d_cprobs) (void)hipFree(g->d_cprobs);
    if (g->h_cprobs) (void)hipHostFree(g->h_cprobs);
    for (auto &p : g->d_items_crt)
        if (p) (void)hipFree(p);
    if (g->d_items_rebuild) (void)hipFree(g->d_items_rebuild);
    delete g;
}

#define GCHK(expr)                                          \
    do {                                                    \
        hipError_t e_ = (expr);                             \
        if (e_ != hipSuccess) return set_error(e_, #expr);  \
    } while (0)

template <typename T>
int upload(T **dptr, const std::vector<T> &v) {
    if (v.empty()) return OZAIE_OK;
    GCHK(hipMalloc(dptr, sizeof(T) * v.size()));
    GCHK(hipMemcpy(*dptr, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    return OZAIE_OK;
}

// Tile of the persistent kernel for a problem whose size class the multi-launch table
// assigns `cfg`: the same shape where the persistent kernel has it (its tiles load
// fragments straight into registers, so an LDS-copy configuration maps to the direct one
// of the same shape), else the widest it has for d (registers allow 64 x 128 tiles with
// prefetch for d = 3 and 7 only, 128 x 128 at d = 3).
int persist_cfg_for(int d, int cfg) {
    if (persist_cfg_supported(d, cfg)) return cfg;
    if (cfg == GCFG_L11) return GCFG_D11;
    if (cfg == GCFG_L12 || cfg == GCFG_D12) return persist_cfg_supported(d, GCFG_D12) ? GCFG_D12 : GCFG_D11;
    return d == 3 ? GCFG_D22 : (d == 7 ? GCFG_D12 : GCFG_D11);
}

// Longest-processing-time assignment of items (in the given order) to the workgroups in
// `wgs`: each item goes to the least loaded one.  Appends to per-workgroup lists.
void lpt_assign(const std::vector<int64_t> &cost, const std::vector<int> &order, const std::vector<int> &wgs,
                std::vector<std::vector<int>> &lists) {
    if (wgs.empty()) return;
    typedef std::pair<int64_t, int> Slot;   // (load, workgroup)
    std::vector<Slot> heap;
    for (int w : wgs) heap.push_back(Slot(0, w));
    auto cmp = [](const Slot &a, const Slot &b) { return a.first != b.first ? a.first > b.first : a.second > b.second; };
    std::make_heap(heap.begin(), heap.end(), cmp);
    for (int i : order) {
        std::pop_heap(heap.begin(), heap.end(), cmp);
        Slot &s = heap.back();
        lists[s.second].push_back(i);
        s.first += cost[i];
        std::push_heap(heap.begin(), heap.end(), cmp);
    }
}

// The tile sequence (problems by decreasing cost, each problem's tiles in band order) split
// into one part per XCD: runs of `run` neighbouring tiles of a problem go to the XCD with
// the least work so far, so tiles that share rows of A or columns of B run on one XCD's
// L2, and every XCD gets the same mix of large and small problems (cutting the sequence
// into contiguous parts gave the last XCD all the smallest problems, whose cost the model
// underestimates).  Returns the item indices of each part, in sequence order.
std::vector<std::vector<int>> xcd_parts(const std::vector<int4> &it, const std::vector<int64_t> &cost, int nx,
                                        int run) {
    std::vector<std::vector<int>> parts(nx);
    std::vector<int64_t> load(nx, 0);
    for (size_t i = 0; i < it.size();) {
        size_t e = i;
        while (e < it.size() && e < i + (size_t)run && it[e].x == it[i].x) e++;
        const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        for (size_t u = i; u < e; u++) {
            parts[x].push_back((int)u);
            load[x] += cost[u];
        }
        i = e;
    }
    return parts;
}

// Duration of one persistent GEMM tile in ns: a + c per contraction block, fitted per
// (slice count, tile shape) to per-tile timings of the 512-problem set (MI355X, rocm 7.2;
// kernel-side s_memrealtime stamps).  The cost does not scale with tile area x slice
// products: small tiles are latency-bound, and the epilogue is a fixed cost per tile.
int64_t persist_tile_cost(int d, int cfg, int64_t Kp) {
    struct Fit { int d, cfg, per_block, fixed; };
    static const Fit fits[] = {{3, GCFG_D11, 651, 2380},  {3, GCFG_D12, 623, 6550},  {3, GCFG_D22, 826, 12440},
                               {7, GCFG_D11, 1027, 4870}, {7, GCFG_D12, 1338, 12760}, {8, GCFG_D11, 1090, 6600},
                               {9, GCFG_D11, 1154, 8430}};
    const int64_t blocks = Kp / kBlk;
    for (const Fit &f : fits)
        if (f.d == d && f.cfg == cfg) return f.per_block * blocks + f.fixed;
    const int64_t area = (int64_t)(group_cfg_tile_m(cfg) / 64) * (group_cfg_tile_n(cfg) / 64);
    return (blocks + 8) * area * (d * (d + 1) / 2) * 40;   // shapes without a fit
}

// Tile shape of a problem in the persistent launch: the shape the kernel has for d with
// the least fitted cost over the problem's tiles (padding included).  Larger tiles cost
// less per output (d = 3: 128 x 128 tiles 3x less per element than 64 x 64), so this
// takes them wherever the padding allows.
int persist_cfg_pick(int d, int64_t M, int64_t N, int64_t Kp) {
    int best = GCFG_D11;
    int64_t best_cost = INT64_MAX;
    for (int c : {GCFG_D11, GCFG_D12, GCFG_D22}) {
        if (!persist_cfg_supported(d, c)) continue;
        const int64_t tiles = ((M + group_cfg_tile_m(c) - 1) / group_cfg_tile_m(c)) *
                              ((N + group_cfg_tile_n(c) - 1) / group_cfg_tile_n(c));
        const int64_t cost = tiles * persist_tile_cost(d, c, Kp);
        if (cost < best_cost) { best = c; best_cost = cost; }
    }
    return best;
}

// Item lists of the persistent kernel.  Phase A: one item per 32-row block (of real rows)
// of every operand, costed by its contraction length, balanced over all waves (four per
// workgroup).
// Phase B: the tiles of every problem in band order, problems by decreasing tile cost,
// dealt to the XCDs by xcd_parts (workgroup w runs on XCD w mod 8); each XCD's part is
// balanced over that XCD's workgroups.
int build_persist(ozaie_group g, const std::vector<int64_t> &Kp) {
    const int n = g->count;
    int dev = 0, ncu = 0;
    GCHK(hipGetDevice(&dev));
    GCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int G = std::max(1, ncu);
    // phase A
    std::vector<int4> itA;
    std::vector<int64_t> costA;
    for (int p = 0; p < n; p++)
        for (int op = 0; op < 2; op++) {
            const GroupOperand &o = g->h_ops[2 * p + op];
            // row blocks of real rows only: the padding rows' slices feed only outputs the
            // tiles do not store
            for (int64_t rb = 0; rb * kBlk < o.R; rb++) {
                itA.push_back(int4{2 * p + op, (int)rb, 0, 0});
                // fitted per-item timings: the same cost per index on both sides, plus a
                // fixed part growing with d (the transposed gather of B costs ~50 more)
                costA.push_back(o.K + (op ? 40 + 7 * g->dp[p] : 5 * g->dp[p]));
            }
        }
    std::vector<int> orderA(itA.size());
    for (size_t i = 0; i < orderA.size(); i++) orderA[i] = (int)i;
    std::stable_sort(orderA.begin(), orderA.end(), [&](int a, int b) { return costA[a] > costA[b]; });
    // phase A runs one item per wave: slot 4 w + v is wave v of workgroup w
    std::vector<int> slots(4 * G);
    for (int w = 0; w < 4 * G; w++) slots[w] = w;
    std::vector<std::vector<int>> listsA(4 * G), listsB(G);
    lpt_assign(costA, orderA, slots, listsA);
    // phase B: each problem in its own tile configuration, tiles costed by contraction
    // blocks (plus the epilogue) times tile area times slice products
    std::vector<int> porder(n);
    for (int p = 0; p < n; p++) porder[p] = p;
    std::stable_sort(porder.begin(), porder.end(), [&](int a, int b) {
        return persist_tile_cost(g->dp[a], g->cfg[a], Kp[a]) > persist_tile_cost(g->dp[b], g->cfg[b], Kp[b]);
    });
    std::vector<int4> itB;
    std::vector<int64_t> costB;
    for (int p : porder) {
        const int c = g->cfg[p], dd = g->dp[p];
        const int tm_rows = group_cfg_tile_m(c), tn_cols = group_cfg_tile_n(c);
        const int64_t ntm = (g->M[p] + tm_rows - 1) / tm_rows, ntn = (g->N[p] + tn_cols - 1) / tn_cols;
        for (int64_t t = 0; t < ntm * ntn; t++) {
            int tm, tn;
            band_map((int)t, (int)ntm, (int)ntn, tm, tn);
            itB.push_back(int4{p, tm, tn, persist_code(dd, c)});
            costB.push_back(persist_tile_cost(dd, c, Kp[p]));
        }
    }
    g->tiles = (int64_t)itB.size();
    const int nx = (G % 8 == 0) ? 8 : 1;
    const std::vector<std::vector<int>> parts = xcd_parts(itB, costB, nx, 8);
    for (int x = 0; x < nx; x++) {
        std::vector<int> wgs;
        for (int w = x; w < G; w += nx) wgs.push_back(w);
        lpt_assign(costB, parts[x], wgs, listsB);
    }
    // flatten: [itemsA][offA][itemsB][offB][ctl]
    std::vector<int4> flatA, flatB;
    std::vector<int> offA(4 * G + 1), offB(G + 1);
    for (int w = 0; w < 4 * G; w++) {
        offA[w] = (int)flatA.size();
        for (int i : listsA[w]) flatA.push_back(itA[i]);
    }
    for (int w = 0; w < G; w++) {
        offB[w] = (int)flatB.size();
        // items of one (d, tile) code consecutive: one tile body per run in the kernel
        std::stable_sort(listsB[w].begin(), listsB[w].end(), [&](int a, int b) { return itB[a].w < itB[b].w; });
        for (int i : listsB[w]) flatB.push_back(itB[i]);
    }
    offA[4 * G] = (int)flatA.size();
    offB[G] = (int)flatB.size();
    const size_t bA = round_up(sizeof(int4) * flatA.size(), 256), bOA = round_up(sizeof(int) * (4 * G + 1), 256);
    const size_t bB = round_up(sizeof(int4) * flatB.size(), 256), bOB = round_up(sizeof(int) * (G + 1), 256);
    const size_t bC = round_up(persist_ctl_bytes(), 256);
    GCHK(hipMalloc(&g->pbuf, bA + bOA + bB + bOB + bC));
    char *q = g->pbuf;
    GCHK(hipMemcpy(q, flatA.data(), sizeof(int4) * flatA.size(), hipMemcpyHostToDevice));
    GCHK(hipMemcpy(q + bA, offA.data(), sizeof(int) * (4 * G + 1), hipMemcpyHostToDevice));
    GCHK(hipMemcpy(q + bA + bOA, flatB.data(), sizeof(int4) * flatB.size(), hipMemcpyHostToDevice));
    GCHK(hipMemcpy(q + bA + bOA + bB, offB.data(), sizeof(int) * (G + 1), hipMemcpyHostToDevice));
    GCHK(hipMemset(q + bA + bOA + bB + bOB, 0, bC));   // once, outside any graph
    GCHK(hipDeviceSynchronize());
    PersistLaunch &L = g->PL;
    L.itemsA = reinterpret_cast<const int4 *>(q);
    L.offA = reinterpret_cast<const int *>(q + bA);
    L.itemsB = reinterpret_cast<const int4 *>(q + bA + bOA);
    L.offB = reinterpret_cast<const int *>(q + bA + bOA + bB);
    L.ctl = q + bA + bOA + bB + bOB;
    L.grid = G;
    L.b = g->b;
    L.rounding = g->rounding;
    return OZAIE_OK;
}

// Work order of the modulus-GEMM tiles (one workgroup per tile): problems by decreasing
// contraction length, each problem's tiles in band order cut into runs of 8, each run to
// the XCD with the least work so far; then XCD x's list goes to blockIdx 8 i + x (the
// hardware deals workgroups to XCDs round-robin), empty slots as problem -1.  Every XCD
// gets about the same work, heaviest first, with neighbouring tiles of a problem (shared
// rows of A and columns of B) on one XCD's L2.
std::vector<int4> xcd_deal(const std::vector<int4> &tiles, const std::vector<int64_t> &Kp) {
    constexpr int X = 8, RUN = 8;
    if (tiles.empty()) return tiles;
    std::vector<std::pair<int, int>> spans;   // [first, last) of each problem's tiles
    for (size_t i = 0; i < tiles.size();) {
        size_t e = i;
        while (e < tiles.size() && tiles[e].x == tiles[i].x) e++;
        spans.emplace_back((int)i, (int)e);
        i = e;
    }
    std::stable_sort(spans.begin(), spans.end(), [&](const std::pair<int, int> &a, const std::pair<int, int> &b) {
        return Kp[tiles[a.first].x] > Kp[tiles[b.first].x];
    });
    std::vector<std::vector<int4>> q(X);
    std::vector<int64_t> load(X, 0);
    for (const auto &sp : spans)
        for (int i = sp.first; i < sp.second; i += RUN) {
            const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            const int e = std::min(sp.second, i + RUN);
            for (int t = i; t < e; t++) q[x].push_back(tiles[t]);
            load[x] += (int64_t)(e - i) * Kp[tiles[i].x];
        }
    size_t len = 0;
    for (const auto &v : q) len = std::max(len, v.size());
    std::vector<int4> out(len * X, int4{-1, 0, 0, 0});
    for (int x = 0; x < X; x++)
        for (size_t i = 0; i < q[x].size(); i++) out[i * X + x] = q[x][i];
    return out;
}

// Per-workgroup (or per-wave) lists flattened into items + offsets, uploaded into one
// buffer region; returns the device addresses through the out pointers.
struct ListBlob {
    std::vector<int4> items;
    std::vector<int> off;
};

ListBlob flatten(const std::vector<std::vector<int>> &lists, const std::vector<int4> &it) {
    ListBlob b;
    b.off.resize(lists.size() + 1);
    for (size_t w = 0; w < lists.size(); w++) {
        b.off[w] = (int)b.items.size();
        for (int i : lists[w]) b.items.push_back(it[i]);
    }
    b.off[lists.size()] = (int)b.items.size();
    return b;
}

// Item lists of the persistent CRT group: phase A as for the split scheme (row blocks of
// real rows, per wave, longest first), phase B the tiles of every problem (each with all
// moduli) in band order, problems by decreasing contraction length, dealt to the XCDs by
// xcd_parts and balanced over each XCD's workgroups, each workgroup's list ordered by tile
// shape; phase C the rebuild items balanced over all workgroups.
int build_persist_crt(ozaie_group g, const std::vector<int64_t> &Kp) {
    const int n = g->count;
    int dev = 0, ncu = 0;
    GCHK(hipGetDevice(&dev));
    GCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int G = std::max(1, ncu) * crt_persist_per_cu();
    std::vector<int4> itA, itB, itC;
    std::vector<int64_t> costA, costB, costC;
    for (int p = 0; p < n; p++)
        for (int op = 0; op < 2; op++) {
            const GroupOperand &o = g->h_ops[2 * p + op];
            for (int64_t rb = 0; rb < (o.R + kBlk - 1) / kBlk; rb++) {
                itA.push_back(int4{2 * p + op, (int)rb, 0, 0});
                // fitted per-item timings: the same cost per index on both sides, plus a
                // fixed part growing with d (the transposed gather of B costs ~50 more)
                costA.push_back(o.K + (op ? 40 + 7 * g->dp[p] : 5 * g->dp[p]));
            }
        }
    std::vector<int> porder(n);
    for (int p = 0; p < n; p++) porder[p] = p;
    std::stable_sort(porder.begin(), porder.end(), [&](int a, int b) { return Kp[a] > Kp[b]; });
    for (int p : porder) {
        // the 256 x 256 tile runs as 128 x 128 tiles in the persistent launch (the padding
        // to 256 is a multiple of 128)
        const int c = g->cfg[p] == CCFG_44 ? CCFG_22 : g->cfg[p];
        const int64_t area = (int64_t)crt_group_cfg_wm(c) * crt_group_cfg_wn(c);
        const CrtGroupProb &cp = g->h_cprobs[p];
        const int ntm = (int)(cp.RbA * kBlk / (64 * crt_group_cfg_wm(c))), ntn = (int)(cp.Np / (64 * crt_group_cfg_wn(c)));
        for (int t = 0; t < ntm * ntn; t++) {
            int tm, tn;
            band_map(t, ntm, ntn, tm, tn);
            itB.push_back(int4{p, tm, tn, c});
            costB.push_back((Kp[p] / kBlk + 6) * area * g->d);   // contraction blocks + epilogue, every modulus
        }
        const int64_t groups = (g->M[p] + 3) / 4 * ((g->N[p] + 3) / 4);
        for (int64_t base = 0; base < groups; base += 256) {
            itC.push_back(int4{p, (int)base, 0, 0});
            costC.push_back(std::min<int64_t>(256, groups - base));
        }
    }
    auto order_of = [](const std::vector<int64_t> &cost) {
        std::vector<int> o(cost.size());
        for (size_t i = 0; i < o.size(); i++) o[i] = (int)i;
        std::stable_sort(o.begin(), o.end(), [&](int a, int b) { return cost[a] > cost[b]; });
        return o;
    };
    std::vector<int> slots(4 * G), wgs(G);
    for (int w = 0; w < 4 * G; w++) slots[w] = w;
    for (int w = 0; w < G; w++) wgs[w] = w;
    std::vector<std::vector<int>> listsA(4 * G), listsB(G), listsC(G);
    lpt_assign(costA, order_of(costA), slots, listsA);
    lpt_assign(costC, order_of(costC), wgs, listsC);
    const int nx = (G % 8 == 0) ? 8 : 1;
    const std::vector<std::vector<int>> parts = xcd_parts(itB, costB, nx, 8);
    for (int x = 0; x < nx; x++) {
        std::vector<int> xw;
        for (int w = x; w < G; w += nx) xw.push_back(w);
        lpt_assign(costB, parts[x], xw, listsB);
    }
    for (auto &l : listsB)
        std::stable_sort(l.begin(), l.end(), [&](int a, int b) { return itB[a].w < itB[b].w; });
    g->tiles = (int64_t)itB.size();
    const ListBlob A = flatten(listsA, itA), B = flatten(listsB, itB), C = flatten(listsC, itC);
    auto r256 = [](size_t x) { return (size_t)round_up((int64_t)x, 256); };
    const size_t sizes[7] = {r256(sizeof(int4) * A.items.size()), r256(sizeof(int) * A.off.size()),
                             r256(sizeof(int4) * B.items.size()), r256(sizeof(int) * B.off.size()),
                             r256(sizeof(int4) * C.items.size()), r256(sizeof(int) * C.off.size()),
                             r256((size_t)persist_ctl_bytes())};
    size_t tot = 0;
    for (size_t z : sizes) tot += z;
    GCHK(hipMalloc(&g->pbuf, tot));
    char *q = g->pbuf;
    const void *src[6] = {A.items.data(), A.off.data(), B.items.data(), B.off.data(), C.items.data(), C.off.data()};
    const size_t len[6] = {sizeof(int4) * A.items.size(), sizeof(int) * A.off.size(), sizeof(int4) * B.items.size(),
                           sizeof(int) * B.off.size(), sizeof(int4) * C.items.size(), sizeof(int) * C.off.size()};
    char *at[7];
    for (int i = 0; i < 7; i++) {
        at[i] = q;
        if (i < 6 && len[i]) GCHK(hipMemcpy(q, src[i], len[i], hipMemcpyHostToDevice));
        q += sizes[i];
    }
    GCHK(hipMemset(at[6], 0, sizes[6]));   // once, outside any graph
    GCHK(hipDeviceSynchronize());
    CrtPersistLaunch &L = g->CPL;
    L.itemsA = reinterpret_cast<const int4 *>(at[0]);
    L.offA = reinterpret_cast<const int *>(at[1]);
    L.itemsB = reinterpret_cast<const int4 *>(at[2]);
    L.offB = reinterpret_cast<const int *>(at[3]);
    L.itemsC = reinterpret_cast<const int4 *>(at[4]);
    L.offC = reinterpret_cast<const int *>(at[5]);
    L.ctl = at[6];
    L.grid = G;
    L.rounding = g->rounding;
    g->PL.ctl = at[6];   // ozaie_group_status reads the control block here
    return OZAIE_OK;
}

// CRT mode: per problem the norm rule (bits 0) or the requested bit count, a tile shape
// per size class, residue planes padded to the tile, chunk statistics (per-phase launches
// only), and the item tables of the phases.
int group_build_crt(ozaie_group g, int force_cfg) {
    const int n = g->count, nm = g->d;
    std::vector<int64_t> Mp(n), Np(n), Kp(n);
    int64_t nE = 0, s_bytes = 0, r_bytes = 0, t_bytes = 0;
    g->beta.assign(n, 0);
    for (int p = 0; p < n; p++) {
        if (g->K[p] > 65536) {
            set_error_msg("CRT mode needs K <= 65536 (exact int32 residue sums)");
            return OZAIE_ERR_DIMS;
        }
        const int bmax = crt_beta(nm, g->K[p]);
        int beta = g->b;   // 0: the norm rule
        if (bmax < 1 || beta < 0 || beta > bmax) {
            set_error_msg("CRT bits exceed the exact range of the moduli for a problem's K");
            return OZAIE_ERR_ARG;
        }
        if ((g->M[p] + 3) / 4 * g->N[p] > INT32_MAX - 256) {
            set_error_msg("grouped CRT problems must have ceil(M / 4) N < 2^31");
            return OZAIE_ERR_DIMS;
        }
        if (beta == 0 && !crt_norm_ok(nm, g->K[p])) beta = bmax;   // as a CRT plan
        g->beta[p] = beta;
        const int c = pick_crt_cfg(nm, g->M[p], g->N[p], g->K[p], force_cfg);
        g->cfg[p] = c;
        Mp[p] = round_up(g->M[p], 64 * crt_group_cfg_wm(c));
        Np[p] = round_up(g->N[p], 64 * crt_group_cfg_wn(c));
        Kp[p] = round_up(g->K[p], kBlk);
        nE += Mp[p] + Np[p];
        s_bytes += (int64_t)nm * (Mp[p] + Np[p]) * Kp[p];
        r_bytes += (int64_t)nm * Mp[p] * Np[p];
        if (!g->persist) t_bytes += 8 * (Mp[p] + Np[p]) * crt_chunks(Kp[p] / kBlk);
    }
    crt_consts(nm, 0, g->cc);   // per-problem rules travel in the descriptors
    const int64_t e_bytes = round_up(nE * 4, 1024);
    s_bytes = round_up(s_bytes, 1024);
    r_bytes = round_up(r_bytes, 1024);
    GCHK(hipMalloc(&g->ws, (size_t)(e_bytes + s_bytes + r_bytes + t_bytes)));
    g->E_all = reinterpret_cast<int32_t *>(g->ws);
    g->n_E = nE;
    GCHK(hipHostMalloc(&g->h_ops, sizeof(GroupOperand) * 2 * n));
    GCHK(hipHostMalloc(&g->h_cprobs, sizeof(CrtGroupProb) * n));
    std::memset(g->h_ops, 0, sizeof(GroupOperand) * 2 * n);
    std::memset(g->h_cprobs, 0, sizeof(CrtGroupProb) * n);
    std::vector<int4> imax, isplit, irebuild;
    std::vector<int4> igemm[CCFG_COUNT];
    int64_t eoff = 0, soff = e_bytes, roff = e_bytes + s_bytes, toff = e_bytes + s_bytes + r_bytes;
    for (int p = 0; p < n; p++) {
        GroupOperand &a = g->h_ops[2 * p], &bo = g->h_ops[2 * p + 1];
        a.R = g->M[p]; a.K = g->K[p]; a.Rp = Mp[p]; a.sr = g->K[p]; a.sk = 1;
        bo.R = g->N[p]; bo.K = g->K[p]; bo.Rp = Np[p]; bo.sr = 1; bo.sk = g->N[p];
        a.E = g->E_all + eoff; eoff += Mp[p];
        bo.E = g->E_all + eoff; eoff += Np[p];
        a.S = reinterpret_cast<int8_t *>(g->ws + soff); soff += (int64_t)nm * Mp[p] * Kp[p];
        bo.S = reinterpret_cast<int8_t *>(g->ws + soff); soff += (int64_t)nm * Np[p] * Kp[p];
        a.Rb = Mp[p] / kBlk; bo.Rb = Np[p] / kBlk;
        a.Kb = bo.Kb = Kp[p] / kBlk;
        if (!g->persist) {
            a.st = reinterpret_cast<uint64_t *>(g->ws + toff); toff += 8 * Mp[p] * crt_chunks(a.Kb);
            bo.st = reinterpret_cast<uint64_t *>(g->ws + toff); toff += 8 * Np[p] * crt_chunks(bo.Kb);
        }
        // statistics, shifts and residues of the row blocks that hold real rows only: the
        // blocks padding a problem to its tile shape feed only outputs at or beyond M rows /
        // N columns, which the modulus GEMM does not store and the rebuild does not read
        for (int op = 0; op < 2; op++) {
            const GroupOperand &o = g->h_ops[2 * p + op];
            for (int64_t rb = 0; rb < (o.R + kBlk - 1) / kBlk; rb++) {
                imax.push_back(int4{2 * p + op, (int)rb, 0, 0});
                for (int64_t kt = 0; kt < (o.Kb + 3) / 4; kt++)
                    isplit.push_back(int4{2 * p + op, (int)rb, (int)kt, 0});
            }
        }
        CrtGroupProb &cp = g->h_cprobs[p];
        cp.SA = a.S; cp.SB = bo.S; cp.Er = a.E; cp.Ec = bo.E;
        cp.R = reinterpret_cast<uint32_t *>(g->ws + roff); roff += (int64_t)nm * Mp[p] * Np[p];
        cp.RbA = a.Rb; cp.RbB = bo.Rb; cp.Kb = a.Kb; cp.Np = Np[p]; cp.ldc = g->N[p];
        cp.M = (int)g->M[p]; cp.N = (int)g->N[p]; cp.beta = g->beta[p];
        const int c = g->cfg[p];
        const int ntm = (int)(Mp[p] / (64 * crt_group_cfg_wm(c))), ntn = (int)(Np[p] / (64 * crt_group_cfg_wn(c)));
        for (int t = 0; t < ntm * ntn; t++) {
            int tm, tn;
            band_map(t, ntm, ntn, tm, tn);
            igemm[c].push_back(int4{p, tm, tn, 0});
        }
        const int64_t groups = (g->M[p] + 3) / 4 * ((g->N[p] + 3) / 4);   // (row quad, 4 columns)
        for (int64_t base = 0; base < groups; base += 256) irebuild.push_back(int4{p, (int)base, 0, 0});
    }
    // Moduli per GEMM workgroup: all of them (the next modulus's loads overlap each
    // epilogue) unless that leaves fewer than 8 workgroups per CU to balance, then fewer.
    const int forced_mpw = ozaie_get_option(OZAIE_OPT_CRT_GROUP_MODULI);
    for (int c = 0; c < CCFG_COUNT; c++) {
        igemm[c] = xcd_deal(igemm[c], Kp);
        int mpw = forced_mpw > 0 ? std::min(forced_mpw, nm)
                                 : (int)std::max<int64_t>(1, std::min<int64_t>(nm, (int64_t)igemm[c].size() * nm / 2048));
        mpw = (nm + (nm + mpw - 1) / mpw - 1) / ((nm + mpw - 1) / mpw);   // even split of the moduli
        g->crt_mpw[c] = mpw;
        g->tiles += (int64_t)igemm[c].size() * ((nm + mpw - 1) / mpw);
    }
    int rc;
    GCHK(hipMalloc(&g->d_ops, sizeof(GroupOperand) * 2 * n));
    GCHK(hipMalloc(&g->d_cprobs, sizeof(CrtGroupProb) * n));
    GCHK(hipEventCreateWithFlags(&g->upload_done, hipEventDisableTiming));
    GCHK(hipStreamCreateWithFlags(&g->cap_stream, hipStreamNonBlocking));
    g->lastA.assign(n, nullptr);
    g->lastB.assign(n, nullptr);
    g->lastC.assign(n, nullptr);
    if (g->persist) {
        g->tiles = 0;
        if ((rc = build_persist_crt(g, Kp))) return rc;
        g->CPL.ops = g->d_ops;
        g->CPL.probs = g->d_cprobs;
        return OZAIE_OK;
    }
    if ((rc = upload(&g->d_items_max, imax))) return rc;
    if ((rc = upload(&g->d_items_split, isplit))) return rc;
    if ((rc = upload(&g->d_items_rebuild, irebuild))) return rc;
    g->n_max = (int)imax.size();
    g->n_split = (int)isplit.size();
    g->n_rebuild = (int)irebuild.size();
    for (int c = 0; c < CCFG_COUNT; c++) {
        if ((rc = upload(&g->d_items_crt[c], igemm[c]))) return rc;
        g->n_crt[c] = (int)igemm[c].size();
    }
    return OZAIE_OK;
}

// Workspace layout, descriptors and work-item tables; pointers to A, B, C are filled in
// at run time.
int group_build(ozaie_group g, int force_cfg) {
    if (g->crt) return group_build_crt(g, force_cfg);
    const int n = g->count;
    std::vector<int64_t> Mp(n), Np(n), Kp(n);
    int64_t nE = 0, slice_bytes = 0;
    const int pforce = ((g->flags >> 4) & 15) - 1;   // forced persistent tile configuration
    for (int p = 0; p < n; p++) {
        const int dd = g->dp[p];
        if (g->K[p] > max_kchunk_blocks(g->b, dd) * kBlk) {
            set_error_msg("grouped problems must fit one contraction pass (K too long for these slices)");
            return OZAIE_ERR_DIMS;
        }
        Mp[p] = round_up(g->M[p], kRowPad);
        Np[p] = round_up(g->N[p], kRowPad);
        Kp[p] = round_up(g->K[p], 2 * kBlk);
        nE += Mp[p] + Np[p];
        slice_bytes += (int64_t)dd * (Mp[p] + Np[p]) * Kp[p];
        const int c = pick_cfg(dd, g->M[p], g->N[p], g->K[p], force_cfg);
        g->cfg[p] = !g->persist ? c
                    : (pforce >= 0 && persist_cfg_supported(dd, pforce) ? pforce
                       : (force_cfg >= 0 ? persist_cfg_for(dd, c) : persist_cfg_pick(dd, g->M[p], g->N[p], Kp[p])));
    }
    const int64_t e_bytes = round_up(nE * 4, 1024);
    GCHK(hipMalloc(&g->ws, (size_t)(e_bytes + slice_bytes)));
    g->E_all = reinterpret_cast<int32_t *>(g->ws);
    g->n_E = nE;
    GCHK(hipHostMalloc(&g->h_ops, sizeof(GroupOperand) * 2 * n));
    GCHK(hipHostMalloc(&g->h_probs, sizeof(GemmLaunch) * n));
    std::memset(g->h_ops, 0, sizeof(GroupOperand) * 2 * n);
    std::memset(g->h_probs, 0, sizeof(GemmLaunch) * n);
    std::vector<int4> imax, isplit;
    std::vector<int4> igemm[4][GCFG_COUNT];
    int64_t eoff = 0, soff = e_bytes;
    for (int p = 0; p < n; p++) {
        const int dd = g->dp[p];
        GroupOperand &a = g->h_ops[2 * p], &bo = g->h_ops[2 * p + 1];
        a.R = g->M[p]; a.K = g->K[p]; a.Rp = Mp[p]; a.sr = g->K[p]; a.sk = 1;
        bo.R = g->N[p]; bo.K = g->K[p]; bo.Rp = Np[p]; bo.sr = 1; bo.sk = g->N[p];
        a.d = bo.d = dd;
        a.E = g->E_all + eoff; eoff += Mp[p];
        bo.E = g->E_all + eoff; eoff += Np[p];
        a.S = reinterpret_cast<int8_t *>(g->ws + soff); soff += (int64_t)dd * Mp[p] * Kp[p];
        bo.S = reinterpret_cast<int8_t *>(g->ws + soff); soff += (int64_t)dd * Np[p] * Kp[p];
        a.Rb = Mp[p] / kBlk; bo.Rb = Np[p] / kBlk;
        a.Kb = bo.Kb = Kp[p] / kBlk;
        // exponent max: rows of A (one wave per row), columns of B (slabs of 256 x 256)
        for (int64_t q = 0; q < (Mp[p] + 3) / 4; q++) imax.push_back(int4{2 * p, (int)q, 0, 0});
        const int64_t kper = 256;
        for (int64_t cx = 0; cx < (g->N[p] + 255) / 256; cx++)
            for (int64_t ky = 0; ky < (g->K[p] + kper - 1) / kper; ky++)
                imax.push_back(int4{2 * p + 1, (int)cx, (int)ky, (int)kper});
        // split tiles: 32 rows x 128 contraction indices
        for (int op = 0; op < 2; op++) {
            const GroupOperand &o = g->h_ops[2 * p + op];
            for (int64_t rb = 0; rb < o.Rb; rb++)
                for (int64_t kt = 0; kt < (o.Kb + 3) / 4; kt++)
                    isplit.push_back(int4{2 * p + op, (int)rb, (int)kt, 0});
        }
        GemmLaunch &gl = g->h_probs[p];
        gl.SA = a.S; gl.SB = bo.S; gl.Er = a.E; gl.Ec = bo.E;
        gl.RbA = a.Rb; gl.RbB = bo.Rb; gl.Kb = a.Kb; gl.kb0 = 0; gl.kb1 = a.Kb;
        gl.M = (int)g->M[p]; gl.N = (int)g->N[p]; gl.b = g->b; gl.d = dd; gl.k0 = 0;
        gl.rounding = g->rounding; gl.first = 1; gl.last = 1; gl.ldc = g->N[p];
        gl.K_true = (int)g->K[p];
        const int c = g->cfg[p];
        const int64_t ntm = (g->M[p] + group_cfg_tile_m(c) - 1) / group_cfg_tile_m(c);
        const int64_t ntn = (g->N[p] + group_cfg_tile_n(c) - 1) / group_cfg_tile_n(c);
        for (int64_t t = 0; t < ntm * ntn; t++) {
            int tm, tn;
            band_map((int)t, (int)ntm, (int)ntn, tm, tn);
            igemm[dslot(dd)][c].push_back(int4{p, tm, tn, 0});
        }
        g->tiles += ntm * ntn;
    }
    int rc;
    GCHK(hipMalloc(&g->d_ops, sizeof(GroupOperand) * 2 * n));
    GCHK(hipMalloc(&g->d_probs, sizeof(GemmLaunch) * n));
    GCHK(hipEventCreateWithFlags(&g->upload_done, hipEventDisableTiming));
    GCHK(hipStreamCreateWithFlags(&g->cap_stream, hipStreamNonBlocking));
    g->lastA.assign(n, nullptr);
    g->lastB.assign(n, nullptr);
    g->lastC.assign(n, nullptr);
    if (g->persist) {
        g->tiles = 0;
        if ((rc = build_persist(g, Kp))) return rc;
        g->PL.ops = g->d_ops;
        g->PL.probs = g->d_probs;
        return OZAIE_OK;
    }
    if ((rc = upload(&g->d_items_max, imax))) return rc;
    if ((rc = upload(&g->d_items_split, isplit))) return rc;
    g->n_max = (int)imax.size();
    g->n_split = (int)isplit.size();
    for (int ds = 0; ds < 4; ds++)
        for (int c = 0; c < GCFG_COUNT; c++) {
            if ((rc = upload(&g->d_items_gemm[ds][c], igemm[ds][c]))) return rc;
            g->n_gemm[ds][c] = (int)igemm[ds][c].size();
        }
    return OZAIE_OK;
}

int group_enqueue(ozaie_group g, hipStream_t st) {
    if (g->crt && g->persist) {
        GCHK(launch_group_crt_persistent(g->CPL, g->cc, st));
        return OZAIE_OK;
    }
    if (g->crt) {
        GCHK(launch_group_crt_stats(g->d_ops, g->d_items_split, g->n_split, st));
        GCHK(launch_group_crt_scale(g->d_ops, g->d_cprobs, g->d_items_max, g->n_max, g->cc, st));
        GCHK(launch_group_crt_split(g->d_ops, g->d_cprobs, g->d_items_split, g->n_split, g->cc, st));
        for (int c = 0; c < CCFG_COUNT; c++)
            GCHK(launch_group_crt_gemm(g->d_cprobs, g->d_items_crt[c], g->n_crt[c], c, g->crt_mpw[c], g->cc, st));
        GCHK(launch_group_crt_rebuild(g->d_cprobs, g->d_items_rebuild, g->n_rebuild, g->rounding, g->cc, st));
        return OZAIE_OK;
    }
    if (g->persist) {
        GCHK(launch_group_persistent(g->PL, st));
        return OZAIE_OK;
    }
    GCHK(launch_fill_i32(g->E_all, g->n_E, 0, st));
    GCHK(launch_group_maxexp(g->d_ops, g->d_items_max, g->n_max, st));
    GCHK(launch_group_split(g->d_ops, g->d_items_split, g->n_split, g->b, st));
    for (int ds = 0; ds < 4; ds++)
        for (int c = 0; c < GCFG_COUNT; c++)
            GCHK(launch_group_gemm(g->d_probs, g->d_items_gemm[ds][c], g->n_gemm[ds][c], kGroupSliceCounts[ds], c, st));
    return OZAIE_OK;
}

int group_capture(ozaie_group g) {
    GCHK(hipStreamBeginCapture(g->cap_stream, hipStreamCaptureModeThreadLocal));
    int rc = group_enqueue(g, g->cap_stream);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(g->cap_stream, &graph);
    if (rc) { if (graph) (void)hipGraphDestroy(graph); return rc; }
    if (e != hipSuccess) return set_error(e, "hipStreamEndCapture");
    g->graph = graph;
    GCHK(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    return OZAIE_OK;
}

int group_new(ozaie_group *out, int count, const int64_t *M, const int64_t *N, const int64_t *K, int bits,
              int slices, int flags, int force_cfg) {
    *out = nullptr;
    if (count < 1 || !M || !N || !K) return OZAIE_ERR_ARG;
    const bool crt = (flags & OZAIE_GROUP_CRT) != 0;
    if (crt) {
        if (slices < 2 || slices > kCrtMaxMod || bits < 0 || bits > 62) {
            set_error_msg("CRT groups need 2 <= moduli <= 16 and 0 <= bits <= 62");
            return OZAIE_ERR_ARG;
        }
        const int want = ((flags >> 4) & 15) - 1;
        if (want >= CCFG_COUNT) {
            set_error_msg("no such CRT tile shape");
            return OZAIE_ERR_ARG;
        }
        if (want >= 0) force_cfg = want;
    } else if (bits < 1 || bits > kMaxBits || slices < 0 || slices > kMaxSlices || bits * slices > 96) {
        return OZAIE_ERR_ARG;
    }
    for (int p = 0; p < count; p++)
        if (M[p] < 1 || N[p] < 1 || K[p] < 1 || M[p] > INT32_MAX / 2 || N[p] > INT32_MAX / 2 || K[p] > INT32_MAX / 2)
            return OZAIE_ERR_DIMS;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        set_error_msg("no HIP device available");
        return OZAIE_ERR_NODEVICE;
    }
    ozaie_group g = new ozaie_group_s();
    g->count = count; g->b = bits; g->d = slices; g->flags = flags;
    g->rounding = (flags & 1) ? OZAIE_ROUND_NEAREST : OZAIE_ROUND_TRUNCATE;
    g->M.assign(M, M + count);
    g->N.assign(N, N + count);
    g->K.assign(K, K + count);
    g->cfg.assign(count, 0);
    // slice count per problem: the group's, or (slices = 0) its size class's from the table
    g->dp.assign(count, slices);
    if (!crt && slices == 0)
        for (int p = 0; p < count; p++) g->dp[p] = pick_slices(M[p], N[p], K[p]);
    int rc = OZAIE_OK;
    hipError_t e = hipGetDevice(&g->device);
    if (e != hipSuccess) rc = set_error(e, "hipGetDevice");
    g->crt = crt;
    g->fallback = false;
    if (!crt)
        for (int p = 0; p < count; p++) g->fallback |= !group_cfg_supported(g->dp[p], GCFG_D11);
    // persistent single launch: every problem has a persistent tile body (slice counts 3, 7,
    // 8, 9; every CRT tile shape)
    g->persist = !g->fallback && (flags & OZAIE_GROUP_PERSISTENT);
    if (!rc && g->fallback) {
        for (int p = 0; p < count && !rc; p++) {
            ozaie_plan pl = nullptr;
            rc = ozaie_plan_create(&pl, M[p], N[p], K[p], bits, g->dp[p], flags & 1);
            if (!rc) g->plans.push_back(pl);
        }
        if (!rc) {
            e = hipStreamCreateWithFlags(&g->cap_stream, hipStreamNonBlocking);
            if (e != hipSuccess) rc = set_error(e, "hipStreamCreate");
        }
        g->lastA.assign(count, nullptr);
        g->lastB.assign(count, nullptr);
        g->lastC.assign(count, nullptr);
    } else if (!rc) {
        rc = group_build(g, force_cfg);
    }
    if (rc) { group_free(g); return rc; }
    *out = g;
    return OZAIE_OK;
}

int group_fallback_enqueue(ozaie_group g, hipStream_t st) {
    for (int p = 0; p < g->count; p++) {
        int rc = ozaie_dgemm(g->plans[p], 0, 0, g->lastA[p], g->K[p], g->lastB[p], g->N[p], g->lastC[p], g->N[p], st);
        if (rc) return rc;
    }
    return OZAIE_OK;
}

int group_run(ozaie_group g, const double *const *A, const double *const *B, double *const *C, hipStream_t st) {
    if (!A || !B || !C) return OZAIE_ERR_ARG;
    g->last_stream = st;
    bool changed = !g->uploaded;
    for (int p = 0; p < g->count; p++) {
        if (!A[p] || !B[p] || !C[p]) return OZAIE_ERR_ARG;
        changed |= A[p] != g->lastA[p] || B[p] != g->lastB[p] || C[p] != g->lastC[p];
    }
    const bool use_graph = !(g->flags & 2);
    if (g->fallback) {
        if (changed && g->exec) {
            (void)hipGraphExecDestroy(g->exec);
            (void)hipGraphDestroy(g->graph);
            g->exec = nullptr;
            g->graph = nullptr;
        }
        for (int p = 0; p < g->count; p++) { g->lastA[p] = A[p]; g->lastB[p] = B[p]; g->lastC[p] = C[p]; }
        g->uploaded = true;
        if (!use_graph) return group_fallback_enqueue(g, st);
        if (!g->exec) {
            GCHK(hipStreamBeginCapture(g->cap_stream, hipStreamCaptureModeThreadLocal));
            int rc = group_fallback_enqueue(g, g->cap_stream);
            hipGraph_t graph = nullptr;
            hipError_t e = hipStreamEndCapture(g->cap_stream, &graph);
            if (rc) { if (graph) (void)hipGraphDestroy(graph); return rc; }
            if (e != hipSuccess) return set_error(e, "hipStreamEndCapture");
            g->graph = graph;
            GCHK(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
        }
        GCHK(hipGraphLaunch(g->exec, st));
        return OZAIE_OK;
    }
    if (changed) {
        // the previous upload must have left the pinned mirrors before they are rewritten
        if (g->uploaded) GCHK(hipEventSynchronize(g->upload_done));
        for (int p = 0; p < g->count; p++) {
            g->h_ops[2 * p].X = A[p];
            g->h_ops[2 * p + 1].X = B[p];
            if (g->crt) g->h_cprobs[p].C = C[p];
            else g->h_probs[p].C = C[p];
            g->lastA[p] = A[p]; g->lastB[p] = B[p]; g->lastC[p] = C[p];
        }
        GCHK(hipMemcpyAsync(g->d_ops, g->h_ops, sizeof(GroupOperand) * 2 * g->count, hipMemcpyHostToDevice, st));
        if (g->crt)
            GCHK(hipMemcpyAsync(g->d_cprobs, g->h_cprobs, sizeof(CrtGroupProb) * g->count, hipMemcpyHostToDevice, st));
        else
            GCHK(hipMemcpyAsync(g->d_probs, g->h_probs, sizeof(GemmLaunch) * g->count, hipMemcpyHostToDevice, st));
        GCHK(hipEventRecord(g->upload_done, st));
        g->uploaded = true;
    }
    if (!use_graph) return group_enqueue(g, st);
    if (!g->exec) {
        int rc = group_capture(g);
        if (rc) return rc;
    }
    GCHK(hipGraphLaunch(g->exec, st));
    return OZAIE_OK;
}

// Autotuner core: for each of the 9 size classes, enough copies of a representative shape
// to fill the GPU a few times, one group per candidate configuration, timed over `reps`
// graph replays after a capture + warm-up run; best[class] = the fastest candidate.
template <typename Default, typename Make>
int tune_classes(int reps, const std::vector<int> &cands, Default dflt, Make make, int32_t *best) {
    if (reps < 1) reps = 3;
    static const int64_t rep_mn[3] = {128, 512, 2048}, rep_k[3] = {128, 512, 2048};
    for (int a = 0; a < 3; a++)
        for (int kb = 0; kb < 3; kb++) {
            const int64_t mn = rep_mn[a], kk = rep_k[kb];
            const int64_t tiles1 = ((mn + 63) / 64) * ((mn + 63) / 64);
            const int cnt = (int)std::max<int64_t>(1, std::min<int64_t>(64, 2048 / tiles1));
            std::vector<int64_t> Ms(cnt, mn), Ns(cnt, mn), Ks(cnt, kk);
            double *buf = nullptr;
            const int64_t per = mn * kk;   // A and B are mn x kk / kk x mn
            hipError_t e = hipMalloc(&buf, sizeof(double) * (size_t)(2 * per + mn * mn));
            if (e != hipSuccess) return set_error(e, "hipMalloc");
            e = launch_fill_random(buf, 2 * per, 1234u + a * 3 + kb, 0);
            if (e != hipSuccess) { (void)hipFree(buf); return set_error(e, "fill"); }
            std::vector<const double *> As(cnt, buf), Bs(cnt, buf + per);
            std::vector<double *> Cs(cnt, buf + 2 * per);
            float best_ms = 1e30f;
            int best_cfg = dflt(mn);
            for (int c : cands) {
                ozaie_group g = nullptr;
                int rc = make(&g, cnt, Ms.data(), Ns.data(), Ks.data(), c);
                if (rc) { (void)hipFree(buf); return rc; }
                hipEvent_t e0, e1;
                (void)hipEventCreate(&e0);
                (void)hipEventCreate(&e1);
                rc = group_run(g, As.data(), Bs.data(), Cs.data(), 0);   // capture + warm-up
                if (!rc) {
                    (void)hipEventRecord(e0, 0);
                    for (int r = 0; r < reps && !rc; r++) rc = group_run(g, As.data(), Bs.data(), Cs.data(), 0);
                    (void)hipEventRecord(e1, 0);
                    (void)hipEventSynchronize(e1);
                    float ms = 0.f;
                    (void)hipEventElapsedTime(&ms, e0, e1);
                    if (!rc && ms < best_ms) { best_ms = ms; best_cfg = c; }
                }
                (void)hipEventDestroy(e0);
                (void)hipEventDestroy(e1);
                (void)hipDeviceSynchronize();
                group_free(g);
                if (rc) { (void)hipFree(buf); return rc; }
            }
            (void)hipFree(buf);
            best[3 * a + kb] = best_cfg;
        }
    return OZAIE_OK;
}

}  // namespace

extern "C" {

int32_t ozaie_group_create(ozaie_group *g, int count, const int64_t *M, const int64_t *N, const int64_t *K,
                           int bits, int slices, int flags) {
    if (!g) return OZAIE_ERR_ARG;
    return group_new(g, count, M, N, K, bits, slices, flags, -1);
}

int32_t ozaie_group_run(ozaie_group g, const double *const *A, const double *const *B, double *const *C,
                        void *stream) {
    if (!g) return OZAIE_ERR_ARG;
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) return OZAIE_ERR_HIP;
    if (cur != g->device) (void)hipSetDevice(g->device);
    int rc = group_run(g, A, B, C, (hipStream_t)stream);
    if (cur != g->device) (void)hipSetDevice(cur);
    return rc;
}

int32_t ozaie_group_query(ozaie_group g, int32_t *out4) {
    if (!g || !out4) return OZAIE_ERR_ARG;
    int launches = 0;
    if (g->fallback) launches = g->count * 5;
    else if (g->persist) launches = 1;
    else if (g->crt) {
        launches = 4;
        for (int c = 0; c < CCFG_COUNT; c++) launches += g->n_crt[c] > 0;
    }
    else {
        launches = 3;
        for (int ds = 0; ds < 4; ds++)
            for (int c = 0; c < GCFG_COUNT; c++) launches += g->n_gemm[ds][c] > 0;
    }
    out4[0] = launches;
    out4[1] = g->exec != nullptr;
    out4[2] = g->fallback;
    out4[3] = (int32_t)std::min<int64_t>(g->tiles, INT32_MAX);
    return OZAIE_OK;
}

int32_t ozaie_group_configs(ozaie_group g, int32_t *cfg_out) {
    if (!g || !cfg_out) return OZAIE_ERR_ARG;
    for (int p = 0; p < g->count; p++) cfg_out[p] = g->fallback ? -1 : g->cfg[p];
    return OZAIE_OK;
}

int32_t ozaie_group_status(ozaie_group g, int32_t *error_out) {
    if (!g || !error_out) return OZAIE_ERR_ARG;
    *error_out = 0;
    if (!g->persist) return OZAIE_OK;
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) return OZAIE_ERR_HIP;
    if (cur != g->device) (void)hipSetDevice(g->device);
    int rc = OZAIE_OK;
    // control words: arrive, done, error (this launch), reported (sticky until read here)
    int32_t ctl[4] = {0, 0, 0, 0};
    hipError_t e = hipStreamSynchronize(g->last_stream);
    if (e == hipSuccess) e = hipMemcpy(ctl, g->PL.ctl, sizeof(ctl), hipMemcpyDeviceToHost);
    if (e == hipSuccess && ctl[3]) e = hipMemset(static_cast<char *>(g->PL.ctl) + 3 * sizeof(int32_t), 0, sizeof(int32_t));
    if (e != hipSuccess) rc = set_error(e, "ozaie_group_status");
    *error_out = ctl[3];
    if (cur != g->device) (void)hipSetDevice(cur);
    return rc;
}

int32_t ozaie_group_destroy(ozaie_group g) {
    if (g) {
        if (g->cap_stream) (void)hipStreamSynchronize(g->cap_stream);
        (void)hipDeviceSynchronize();
    }
    group_free(g);
    return OZAIE_OK;
}

int32_t ozaie_tune_get_slices(int32_t *table9) {
    if (!table9) return OZAIE_ERR_ARG;
    (void)pick_slices(1, 1, 1);   // initialises the table
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (int c = 0; c < 9; c++) table9[c] = g_tune_slices[c];
    return OZAIE_OK;
}

int32_t ozaie_tune_set_slices(const int32_t *table9) {
    if (!table9) return OZAIE_ERR_ARG;
    for (int c = 0; c < 9; c++)
        if (table9[c] != -1 && !group_cfg_supported(table9[c], GCFG_D11)) return OZAIE_ERR_ARG;
    (void)pick_slices(1, 1, 1);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (int c = 0; c < 9; c++) g_tune_slices[c] = table9[c];
    return OZAIE_OK;
}

int32_t ozaie_tune_get(int slices, int32_t *table9) {
    if (!table9 || slices < 1 || slices > kMaxSlices) return OZAIE_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    tune_init();
    for (int c = 0; c < 9; c++) table9[c] = g_tune[slices][c];
    return OZAIE_OK;
}

int32_t ozaie_tune_set(int slices, const int32_t *table9) {
    if (!table9 || slices < 1 || slices > kMaxSlices) return OZAIE_ERR_ARG;
    for (int c = 0; c < 9; c++)
        if (table9[c] >= 0 && !group_cfg_supported(slices, table9[c])) return OZAIE_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    tune_init();
    for (int c = 0; c < 9; c++) g_tune[slices][c] = table9[c];
    return OZAIE_OK;
}

int32_t ozaie_autotune(int bits, int slices, int reps, int32_t *table9) {
    if (bits < 1 || bits > kMaxBits || slices < 1 || slices > kMaxSlices) return OZAIE_ERR_ARG;
    if (!group_cfg_supported(slices, GCFG_D11)) return OZAIE_ERR_ARG;
    std::vector<int> cands;
    for (int c = 0; c < GCFG_COUNT; c++)
        if (group_cfg_supported(slices, c)) cands.push_back(c);
    int32_t best[9];
    int rc = tune_classes(reps, cands, [&](int64_t mn) { return default_cfg(slices, mn, mn); },
                          [&](ozaie_group *g, int cnt, const int64_t *Ms, const int64_t *Ns, const int64_t *Ks, int c) {
                              return group_new(g, cnt, Ms, Ns, Ks, bits, slices, 0, c);
                          }, best);
    if (rc) return rc;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        tune_init();
        for (int c = 0; c < 9; c++) g_tune[slices][c] = best[c];
    }
    if (table9)
        for (int c = 0; c < 9; c++) table9[c] = best[c];
    return OZAIE_OK;
}

int32_t ozaie_autotune_crt(int moduli, int reps, int32_t *table9) {
    if (moduli < 2 || moduli > kCrtMaxMod) return OZAIE_ERR_ARG;
    std::vector<int> cands;
    for (int c = 0; c < CCFG_COUNT; c++) cands.push_back(c);
    int32_t best[9];
    int rc = tune_classes(reps, cands, [&](int64_t mn) { return default_crt_cfg(mn, mn); },
                          [&](ozaie_group *g, int cnt, const int64_t *Ms, const int64_t *Ns, const int64_t *Ks, int c) {
                              return group_new(g, cnt, Ms, Ns, Ks, 0, moduli, OZAIE_GROUP_CRT, c);
                          }, best);
    if (rc) return rc;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        tune_crt_init();
        for (int c = 0; c < 9; c++) g_tune_crt[moduli][c] = best[c];
    }
    if (table9)
        for (int c = 0; c < 9; c++) table9[c] = best[c];
    return OZAIE_OK;
}

int32_t ozaie_tune_get_crt(int moduli, int32_t *table9) {
    if (!table9 || moduli < 2 || moduli > kCrtMaxMod) return OZAIE_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    tune_crt_init();
    for (int c = 0; c < 9; c++) table9[c] = g_tune_crt[moduli][c];
    return OZAIE_OK;
}

int32_t ozaie_tune_set_crt(int moduli, const int32_t *table9) {
    if (!table9 || moduli < 2 || moduli > kCrtMaxMod) return OZAIE_ERR_ARG;
    for (int c = 0; c < 9; c++)
        if (table9[c] >= CCFG_COUNT) return OZAIE_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    tune_crt_init();
    for (int c = 0; c < 9; c++) g_tune_crt[moduli][c] = table9[c];
    return OZAIE_OK;
}

int32_t ozaie_dgemm_grouped(int count, const int64_t *M, const int64_t *N, const int64_t *K,
                            const double *const *A, const double *const *B, double *const *C, int bits,
                            int slices, void *stream) {
    ozaie_group g = nullptr;
    int rc = ozaie_group_create(&g, count, M, N, K, bits, slices, 2);
    if (rc) return rc;
    rc = ozaie_group_run(g, A, B, C, stream);
    if (!rc) {
        hipError_t e = hipStreamSynchronize((hipStream_t)stream);
        if (e != hipSuccess) rc = set_error(e, "hipStreamSynchronize");
    }
    group_free(g);
    return rc;
}

}  // extern "C"
