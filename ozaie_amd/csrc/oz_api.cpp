// oz_api.cpp -- C ABI of libozaie.so: planner, device buffers, entry points.
//
// The reference's host layer (include/oz_aie.h:31-349 of atlarge-research/ozAIE) keeps one
// process-global configuration, picks (b, d) from a table indexed by log2 K, builds or
// loads an NPU binary, allocates XRT buffers and converts the NPU's paired-int64 output
// on the host.  Here the same public entry points sit on a planner that picks (b, d),
// the contraction chunk that keeps every int32 bin exact, the kernel family and the
// device buffers; there is no compile step and the conversion runs on the GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ozaie.h"
#include "oz_internal.h"

using namespace oz;

namespace {

std::mutex g_mu;
std::string g_last_error;
int g_opt_rounding = OZAIE_ROUND_TRUNCATE;
int g_opt_strict = 0;
int g_opt_kchunk = 0;
int g_opt_kernel = 0;
int g_opt_timing = 0;
int g_opt_crt_mpw = 0;

int set_hip_error(hipError_t e, const char *where) {
    g_last_error = std::string(where) + ": " + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")";
    if (e == hipErrorOutOfMemory) return OZAIE_ERR_NOMEM;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return OZAIE_ERR_NODEVICE;
    return OZAIE_ERR_HIP;
}

#define HIPCHK(expr)                                   \
    do {                                               \
        hipError_t e_ = (expr);                        \
        if (e_ != hipSuccess) return set_hip_error(e_, #expr); \
    } while (0)

int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

// Bins per pass of the multi-pass kernels: bin windows (4) and bin groups (1).
constexpr int kPassBins(int kernel) { return kernel == 4 ? kWindowBins : kGroupBins; }

// Makes `dev` current for the guard's lifetime and restores the caller's device on every
// return path.
struct DeviceGuard {
    int prev = -1;
    hipError_t enter(int dev) {
        int cur = 0;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur == dev) return hipSuccess;
        e = hipSetDevice(dev);
        if (e == hipSuccess) prev = cur;
        return e;
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

// Reference mode table, include/oz_aie.h:60-80 (log2 K -> (b, d)).
const int kAccurate[8][2] = {{3, 29}, {3, 28}, {3, 27}, {3, 27}, {3, 26}, {3, 25}, {3, 25}, {4, 18}};

int ensure_device() {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        g_last_error = "no HIP device available";
        return OZAIE_ERR_NODEVICE;
    }
    return OZAIE_OK;
}

}  // namespace

namespace oz {
int set_error(hipError_t e, const char *where) { return set_hip_error(e, where); }
void set_error_msg(const char *msg) { g_last_error = msg; }

// CRT mode constants: the first n of 20 pairwise-coprime moduli <= 256 (256 = 2^8 and odd
// numbers whose prime factors are all distinct), M = prod p < 2^156, held in 192-bit
// integers (three 64-bit words, low first, arithmetic mod 2^192).
static const int kCrtModuli[kCrtMaxMod] = {256, 255, 253, 251, 247, 241, 239, 233, 229, 227,
                                           223, 211, 199, 197, 193, 191, 181, 179, 173, 167};

int crt_modulus(int i) { return (i >= 0 && i < kCrtMaxMod) ? kCrtModuli[i] : 0; }

namespace {
struct Wide { uint64_t w[3] = {0, 0, 0}; };

Wide wide_mul(const Wide &a, uint64_t m) {   // m < 2^32
    Wide r;
    unsigned __int128 c = 0;
    for (int i = 0; i < 3; i++) { c += (unsigned __int128)a.w[i] * m; r.w[i] = (uint64_t)c; c >>= 64; }
    return r;
}

uint64_t wide_divmod(Wide &a, uint64_t d) {   // a /= d, returns a mod d
    unsigned __int128 r = 0;
    for (int i = 2; i >= 0; i--) {
        r = (r << 64) | a.w[i];
        a.w[i] = (uint64_t)(r / d);
        r %= d;
    }
    return (uint64_t)r;
}

int wide_bits(const Wide &a) {
    for (int i = 2; i >= 0; i--)
        if (a.w[i]) return 64 * i + 64 - __builtin_clzll(a.w[i]);
    return 0;
}

double wide_double(const Wide &a) {
    return ((double)a.w[2] * 18446744073709551616.0 + (double)a.w[1]) * 18446744073709551616.0 + (double)a.w[0];
}

uint32_t wide_byte(const Wide &a, int b) { return (uint32_t)((a.w[b / 8] >> (8 * (b % 8))) & 0xFF); }

Wide crt_M(int n) {
    Wide m;
    m.w[0] = 1;
    for (int i = 0; i < n; i++) m = wide_mul(m, (uint64_t)kCrtModuli[i]);
    return m;
}
}  // namespace

// Largest beta with K 2^(2 beta) <= 2^(floor(log2 M) - 2) <= M / 4 (at most 62).
int crt_beta(int n, int64_t K) {
    if (n < 2 || n > kCrtMaxMod || K < 1) return -1;
    const int L = wide_bits(crt_M(n)) - 1;
    int lk = 0;
    while (((int64_t)1 << lk) < K) lk++;
    const int beta = std::min(62, (L - 2 - lk) / 2);
    return beta >= 1 ? beta : -1;
}

// The norm rule needs R = M 2^20 - K (2^20 + 2^40) > 0.
bool crt_norm_ok(int n, int64_t K) {
    if (n < 2 || n > kCrtMaxMod || K < 1 || K > 65536) return false;
    const Wide M = crt_M(n);
    if (M.w[2] || (M.w[1] >> 44)) return true;   // M 2^20 >= 2^128 > K (2^20 + 2^40)
    const unsigned __int128 MR = (((unsigned __int128)M.w[1] << 64) | M.w[0]) << 20;
    return MR > (unsigned __int128)K * ((1ull << 20) + (1ull << 40));
}

void crt_consts(int n, int beta, CrtConsts &c) {
    std::memset(&c, 0, sizeof(c));
    c.n = n;
    c.beta = beta;
    const Wide M = crt_M(n);
    for (int i = 0; i < 3; i++) c.Mw[i] = M.w[i];
    for (int i = 0; i < n; i++) {
        const int p = kCrtModuli[i];
        Wide Mi = M;
        wide_divmod(Mi, (uint64_t)p);
        Wide tmp = Mi;
        // inverse of Mi mod p by the extended Euclidean algorithm
        int t = 0, nt = 1, r = p, nr = (int)wide_divmod(tmp, (uint64_t)p);
        while (nr) {
            const int qq = r / nr;
            int tt = t - qq * nt; t = nt; nt = tt;
            tt = r - qq * nr; r = nr; nr = tt;
        }
        if (t < 0) t += p;
        const Wide w = wide_mul(Mi, (uint64_t)t);   // W_p < M
        c.p[i] = p;
        c.inv[i] = 1.0 / p;
        c.c31[i] = (double)((1ull << 31) % (uint64_t)p);
        const double f = wide_double(w) / wide_double(M);
        const uint32_t q = (uint32_t)std::floor(f * 65536.0);   // < 2^16 since f < 1
        for (int b = 0; b < kCrtLimbs; b++) c.Wd[i / 4][b] |= wide_byte(w, b) << (8 * (i % 4));
        c.Qd[0][i / 4] |= (q & 0xFFu) << (8 * (i % 4));
        c.Qd[1][i / 4] |= (q >> 8) << (8 * (i % 4));
    }
    Wide Mn;   // 2^192 - M
    unsigned __int128 cy = 1;
    for (int i = 0; i < 3; i++) { cy += (unsigned __int128)(~M.w[i]); Mn.w[i] = (uint64_t)cy; cy >>= 64; }
    for (int b = 0; b < kCrtLimbs; b++) c.mneg8[b] = wide_byte(Mn, b);
}
}  // namespace oz

struct ozaie_plan_s {
    int64_t M = 0, N = 0, K = 0;
    int b = 0, d = 0, rounding = 0;
    int64_t Mp = 0, Np = 0, Kp = 0, Kb = 0;
    int64_t kcb = 0;      // contraction blocks per pass
    int nchunks = 0, ngroups = 1, kernel = 0, family = 0;
    int device = 0, cus = 256;   // the device and its compute units
    bool crt = false;     // CRT mode: b = bits per operand, d = number of moduli
    bool pairs = false;   // paired-digit path (pair_cfg): kernel 3
    int32_t *X = nullptr; // its anti-diagonal sums, Mp x Np int32
    CrtConsts cc{};
    uint32_t *R = nullptr;  // CRT residue products [d][Mp/4][Np]
    int8_t *SA = nullptr, *SB = nullptr;
    int32_t *Er = nullptr, *Ec = nullptr;   // CRT mode: the shifts sigma
    uint64_t *stA = nullptr, *stB = nullptr; // CRT mode: chunk statistics of A and B
    __int128 *T = nullptr;
    // staging for host pointers (reference-compatible entry point): device copies of A, B,
    // C and two pinned host chunks with the events of their last DMA
    double *dA = nullptr, *dB = nullptr, *dC = nullptr;
    char *pin[2] = {nullptr, nullptr};
    hipEvent_t pin_ev[2] = {nullptr, nullptr};
    bool pin_used[2] = {false, false};
    hipStream_t stream = nullptr;
    // optional phase timing: start, after split, after the GEMM kernels, end (CRT: after
    // the rebuild, which follows the modulus GEMM; otherwise the same point as ev[2])
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    bool timed = false;
};

namespace {

void plan_free(ozaie_plan p) {
    if (!p) return;
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(p->device);
    if (p->SA) (void)hipFree(p->SA);
    if (p->SB) (void)hipFree(p->SB);
    if (p->Er) (void)hipFree(p->Er);
    if (p->Ec) (void)hipFree(p->Ec);
    if (p->T) (void)hipFree(p->T);
    if (p->R) (void)hipFree(p->R);
    if (p->X) (void)hipFree(p->X);
    if (p->stA) (void)hipFree(p->stA);
    if (p->stB) (void)hipFree(p->stB);
    if (p->dA) (void)hipFree(p->dA);
    if (p->dB) (void)hipFree(p->dB);
    if (p->dC) (void)hipFree(p->dC);
    for (int i = 0; i < 2; i++) {
        if (p->pin_ev[i]) (void)hipEventDestroy(p->pin_ev[i]);
        if (p->pin[i]) (void)hipHostFree(p->pin[i]);
    }
    if (p->stream) (void)hipStreamDestroy(p->stream);
    for (auto &e : p->ev)
        if (e) (void)hipEventDestroy(e);
    (void)hipSetDevice(cur);
    delete p;
}

// Contraction indices per pass so that no int32 bin can wrap: the largest bin sums
// d slice products of magnitude <= (2^b - 1)^2 over the pass.
int64_t max_kchunk(int b, int d) {
    const int64_t m = (1 << b) - 1;
    int64_t kc = (int64_t)INT32_MAX / ((int64_t)d * m * m);
    return kc / (2 * kBlk) * (2 * kBlk);
}

}  // namespace

int64_t oz::max_kchunk_blocks(int b, int d) { return std::max<int64_t>(2, max_kchunk(b, d) / kBlk / 2 * 2); }

namespace {

// Paired-digit path (pair_cfg) for b <= 3 from this many slices on: (P(P+1)/2 + Q) int8
// products instead of d(d+1)/2 (d = 29: 120 instead of 435), one fused pass.
constexpr int kPairMinSlices = 8;

int plan_configure(ozaie_plan p, int64_t M, int64_t N, int64_t K, int b, int d, int rounding,
                   bool allow_pairs = true) {
    if (M < 1 || N < 1 || K < 1) { g_last_error = "dimensions must be >= 1"; return OZAIE_ERR_DIMS; }
    if (M > (1ll << 31) - 256 || N > (1ll << 31) - 256 || K > (1ll << 40)) {
        g_last_error = "dimensions too large"; return OZAIE_ERR_DIMS;
    }
    if (b < 1 || b > kMaxBits || d < 1 || d > kMaxSlices || b * d > 96) {
        g_last_error = "need 1 <= bits <= 7, 1 <= slices <= 32, bits*slices <= 96";
        return OZAIE_ERR_ARG;
    }
    // |T| <= 4 K (2^b - 1)^2 2^(b(d-1)) must stay below 2^127.
    if (2.0 + std::log2((double)K) + 2.0 * b + (double)b * (d - 1) > 126.0) {
        g_last_error = "product too long for exact 128-bit accumulation";
        return OZAIE_ERR_OVERFLOW;
    }
    p->M = M; p->N = N; p->K = K; p->b = b; p->d = d; p->rounding = rounding;
    p->Mp = round_up(M, kRowPad);
    p->Np = round_up(N, kRowPad);
    p->Kp = round_up(K, 2 * kBlk);   // even number of contraction blocks per pass
    p->Kb = p->Kp / kBlk;
    int64_t kc = max_kchunk(b, d);
    if (g_opt_kchunk > 0) kc = std::min<int64_t>(kc, std::max<int64_t>(2 * kBlk, g_opt_kchunk / (2 * kBlk) * (2 * kBlk)));
    p->kcb = std::max<int64_t>(2, kc / kBlk / 2 * 2);
    p->nchunks = (int)((p->Kb + p->kcb - 1) / p->kcb);
    // 0: fused (all bins in registers), 4: bin windows of the rolling kernel (d = 18),
    // 1: bin groups of 8 (any d; also on request)
    p->kernel = g_opt_kernel == 1 ? 1 : (fused_supported(d) ? 0 : (window_supported(d) ? 4 : 1));
    p->family = g_opt_kernel >= 2 ? g_opt_kernel : KERNEL_AUTO;
    p->ngroups = p->kernel == 0 ? 1 : (d + kPassBins(p->kernel) - 1) / kPassBins(p->kernel);
    if (allow_pairs && g_opt_kernel != 1 && b <= 3 && d >= kPairMinSlices) {
        const PairCfg pc = pair_cfg(b, d);
        const int64_t m1 = (1 << b) - 1, m2 = (1 << (2 * b)) - 1;
        // the anti-diagonal sums over all of K stay int32; for even d the epilogue forms 2^b T
        const bool x_ok = (double)pc.Q * m1 * m1 * (double)K < 2147483647.0;
        const bool t_ok = pc.odd || 2.0 + std::log2((double)K) + 2.0 * b + (double)b * d <= 126.0;
        if (x_ok && t_ok && fused_supported(pc.P)) {
            p->pairs = true;
            p->kernel = 3;
            p->ngroups = 1;
            p->Mp = round_up(M, kPairRowPad);
            p->Np = round_up(N, kPairRowPad);
            int64_t kcp = (int64_t)INT32_MAX / ((int64_t)pc.P * m2 * m2) / (2 * kBlk) * (2 * kBlk);
            if (g_opt_kchunk > 0) kcp = std::min<int64_t>(kcp, std::max<int64_t>(2 * kBlk, g_opt_kchunk / (2 * kBlk) * (2 * kBlk)));
            p->kcb = std::max<int64_t>(2, kcp / kBlk / 2 * 2);
            p->nchunks = (int)((p->Kb + p->kcb - 1) / p->kcb);
        }
    }
    if (p->Mp / kBlk > 65535 || p->Np / kBlk > 65535) {
        g_last_error = "too many rows for the split grid"; return OZAIE_ERR_DIMS;
    }
    return OZAIE_OK;
}

// CRT mode: n moduli, beta bits per operand (the max rule), or 0: the norm rule (the max
// rule with the largest beta where M 2^20 <= K (2^20 + 2^40), i.e. for a few moduli).
int plan_configure_crt(ozaie_plan p, int64_t M, int64_t N, int64_t K, int beta, int n, int rounding) {
    if (M < 1 || N < 1 || K < 1) { g_last_error = "dimensions must be >= 1"; return OZAIE_ERR_DIMS; }
    if (M > (1ll << 31) - 512 || N > (1ll << 31) - 512) { g_last_error = "dimensions too large"; return OZAIE_ERR_DIMS; }
    if (K > 65536) { g_last_error = "CRT mode needs K <= 65536 (exact int32 residue sums)"; return OZAIE_ERR_DIMS; }
    if (n < 2 || n > kCrtMaxMod) { g_last_error = "CRT mode needs 2 <= moduli <= 20"; return OZAIE_ERR_ARG; }
    const int bmax = crt_beta(n, K);
    if (bmax < 1) { g_last_error = "no exact bit count for these moduli and K"; return OZAIE_ERR_ARG; }
    if (beta < 0 || beta > bmax) { g_last_error = "CRT bits exceed the exact range of the moduli"; return OZAIE_ERR_ARG; }
    if (beta == 0 && !crt_norm_ok(n, K)) beta = bmax;   // few moduli: the norm rule has no room
    p->M = M; p->N = N; p->K = K; p->b = beta; p->d = n; p->rounding = rounding;
    p->crt = true;
    crt_consts(n, beta, p->cc);
    p->Mp = round_up(M, kCrtRowPad);
    p->Np = round_up(N, kCrtRowPad);
    p->Kp = round_up(K, kCrtPF * kBlk);   // whole prefetch rings of the modulus GEMM
    p->Kb = p->Kp / kBlk;
    p->kcb = p->Kb;
    p->nchunks = 1;
    p->ngroups = 1;
    p->kernel = 2;
    p->family = g_opt_kernel >= 2 ? g_opt_kernel : KERNEL_AUTO;
    if (p->Mp / kBlk > 65535 || p->Np / kBlk > 65535) { g_last_error = "too many rows for the split grid"; return OZAIE_ERR_DIMS; }
    return OZAIE_OK;
}

int plan_alloc(ozaie_plan p, bool staging) {
    HIPCHK(hipGetDevice(&p->device));
    HIPCHK(hipDeviceGetAttribute(&p->cus, hipDeviceAttributeMultiprocessorCount, p->device));
    const size_t sa = (size_t)p->d * p->Mp * p->Kp, sb = (size_t)p->d * p->Np * p->Kp;
    HIPCHK(hipMalloc(&p->SA, sa));
    HIPCHK(hipMalloc(&p->SB, sb));
    HIPCHK(hipMalloc(&p->Er, sizeof(int32_t) * p->Mp));
    HIPCHK(hipMalloc(&p->Ec, sizeof(int32_t) * p->Np));
    if (!p->crt && (int64_t)p->nchunks * p->ngroups > 1)
        HIPCHK(hipMalloc(&p->T, sizeof(__int128) * (size_t)p->M * p->N));
    if (p->crt) {
        HIPCHK(hipMalloc(&p->R, (size_t)p->d * p->Mp * p->Np));
        HIPCHK(hipMalloc(&p->stA, sizeof(uint64_t) * (size_t)p->Mp * crt_chunks(p->Kb)));
        HIPCHK(hipMalloc(&p->stB, sizeof(uint64_t) * (size_t)p->Np * crt_chunks(p->Kb)));
    }
    if (p->pairs) HIPCHK(hipMalloc(&p->X, sizeof(int32_t) * (size_t)p->Mp * p->Np));
    if (staging) {
        HIPCHK(hipMalloc(&p->dA, sizeof(double) * (size_t)p->M * p->K));
        HIPCHK(hipMalloc(&p->dB, sizeof(double) * (size_t)p->K * p->N));
        HIPCHK(hipMalloc(&p->dC, sizeof(double) * (size_t)p->M * p->N));
        HIPCHK(hipStreamCreateWithFlags(&p->stream, hipStreamDefault));
    }
    if (g_opt_timing) {
        for (auto &e : p->ev) HIPCHK(hipEventCreate(&e));
        p->timed = true;
    }
    return OZAIE_OK;
}

int plan_new(ozaie_plan *out, int64_t M, int64_t N, int64_t K, int b, int d, int rounding,
             bool staging, bool crt = false, bool allow_pairs = true) {
    *out = nullptr;
    int rc = ensure_device();
    if (rc) return rc;
    ozaie_plan p = new ozaie_plan_s();
    rc = crt ? plan_configure_crt(p, M, N, K, b, d, rounding) : plan_configure(p, M, N, K, b, d, rounding, allow_pairs);
    if (!rc) rc = plan_alloc(p, staging);
    if (rc) { plan_free(p); return rc; }
    *out = p;
    return OZAIE_OK;
}

OperandView view_A(const ozaie_plan p, int trans, const double *A, int64_t lda) {
    OperandView v;
    v.X = A; v.R = p->M; v.K = p->K;
    if (!trans) { v.sr = lda; v.sk = 1; } else { v.sr = 1; v.sk = lda; }
    return v;
}

OperandView view_B(const ozaie_plan p, int trans, const double *B, int64_t ldb) {
    OperandView v;
    v.X = B; v.R = p->N; v.K = p->K;
    if (!trans) { v.sr = 1; v.sk = ldb; } else { v.sr = ldb; v.sk = 1; }
    return v;
}

GemmLaunch base_launch(const ozaie_plan p) {
    GemmLaunch g;
    std::memset(&g, 0, sizeof(g));
    g.SA = p->SA; g.SB = p->SB; g.Er = p->Er; g.Ec = p->Ec;
    g.RbA = p->Mp / kBlk; g.RbB = p->Np / kBlk; g.Kb = p->Kb;
    g.M = (int)p->M; g.N = (int)p->N; g.b = p->b; g.d = p->d;
    g.rounding = p->rounding;
    g.T = p->T; g.ldT = p->N;
    g.K_true = (int)std::min<int64_t>(p->K, INT32_MAX);
    g.family = p->family;
    g.cus = p->cus;
    return g;
}

// Exponents + slices of both operands, one operand after the other: the split's second
// read of an operand follows its exponent pass directly, before the other operand streams
// through the memory-side cache (256 MB).  Measured: 2-3% off the split-scheme split phase
// at 4096^3, no change for the CRT residues or at 8192^3.
int run_split(ozaie_plan p, const OperandView &va, const OperandView &vb, hipStream_t st) {
    if (p->crt) {   // chunk statistics, shifts, residues
        const int64_t nch = crt_chunks(p->Kb);
        const CrtSide a{va, p->stA, p->Er, p->SA, p->Mp}, b{vb, p->stB, p->Ec, p->SB, p->Np};
        HIPCHK(launch_crt_stats(a, b, p->Kp, st));
        HIPCHK(launch_crt_scale(a, b, nch, p->K, p->cc, st));
        HIPCHK(launch_crt_split(a, b, p->cc, p->Kp, st));
        return OZAIE_OK;
    }
    // K <= 128: each split workgroup holds a whole row's contraction axis and forms the
    // row's exponent maximum itself (two launches fewer per operand)
    const bool one_tile = p->Kp <= 128;
    if (!one_tile) HIPCHK(launch_maxexp(va, p->Er, p->Mp, st));
    if (p->pairs)
        HIPCHK(launch_split_pairs(va, p->Er, p->b, p->d, p->SA, p->Mp, p->Kp, false, st, one_tile ? p->Er : nullptr));
    else
        HIPCHK(launch_split(va, p->Er, p->b, p->d, p->SA, p->Mp, p->Kp, st, one_tile ? p->Er : nullptr));
    if (!one_tile) HIPCHK(launch_maxexp(vb, p->Ec, p->Np, st));
    if (p->pairs)
        HIPCHK(launch_split_pairs(vb, p->Ec, p->b, p->d, p->SB, p->Np, p->Kp, true, st, one_tile ? p->Ec : nullptr));
    else
        HIPCHK(launch_split(vb, p->Ec, p->b, p->d, p->SB, p->Np, p->Kp, st, one_tile ? p->Ec : nullptr));
    return OZAIE_OK;
}

int run_passes(ozaie_plan p, double *C, int64_t ldc, int epi, int32_t *bins, hipStream_t st) {
    if (p->crt) {
        CrtGemm cg;
        cg.SA = p->SA; cg.SB = p->SB; cg.Mp = p->Mp; cg.Np = p->Np; cg.Kb = p->Kb; cg.R = p->R;
        // 256 x 256 tiles unless their grid (tiles x moduli, one workgroup per CU) fills fewer
        // than four rounds of the CUs: then 128 x 128 tiles, two workgroups per CU (measured on
        // MI355X: 1.5-4% faster at 1024^3 and 2048^3, 5-12% slower from 3072^3 up)
        const int64_t wg256 = (p->Mp / 256) * (p->Np / 256) * p->d;
        cg.variant = p->family == KERNEL_DIRECT ? 1
                   : p->family == KERNEL_PIPE  ? 2
                   : p->family == KERNEL_LDS   ? 0
                   : (wg256 < 4 * (int64_t)p->cus ? 2 : 0);
        HIPCHK(launch_crt_gemm(cg, p->cc, st));
        if (p->timed) HIPCHK(hipEventRecord(p->ev[2], st));
        if (epi == EPI_C)
            HIPCHK(launch_crt_rebuild(p->R, p->Mp, p->Np, p->M, p->N, p->Er, p->Ec, C, ldc, p->rounding, p->cc, st));
        return OZAIE_OK;
    }
    GemmLaunch g = base_launch(p);
    g.C = C; g.ldc = ldc; g.bins = bins;
    if (p->pairs) {
        if (epi != EPI_C) { g_last_error = "the paired-digit path has no bin output"; return OZAIE_ERR_ARG; }
        const PairCfg pc = pair_cfg(p->b, p->d);
        const int64_t planeA = p->Mp / kBlk * p->Kb * kBlkBytes, planeB = p->Np / kBlk * p->Kb * kBlkBytes;
        HIPCHK(launch_gemm_antidiag(p->SA + pc.P * planeA, p->SB + pc.P * planeB, p->Mp, p->Np, pc.Q * p->Kb, p->X,
                                    p->Np, p->cus, st));
        g.b = 2 * p->b;
        g.d = pc.P;
        g.cb = p->b;
        g.cd = pc.odd ? p->d : p->d + 1;
        g.xadd = p->X;
        g.ldx = p->Np;
        g.xsign = pc.odd ? 1 : -1;
    }
    const int npass = p->nchunks * p->ngroups;
    int pass = 0;
    for (int c = 0; c < p->nchunks; c++) {
        g.kb0 = (int64_t)c * p->kcb;
        g.kb1 = std::min<int64_t>(p->Kb, g.kb0 + p->kcb);
        for (int gi = 0; gi < p->ngroups; gi++, pass++) {
            g.k0 = gi * kPassBins(p->kernel);
            g.first = pass == 0;
            g.last = pass == npass - 1;
            if (p->kernel == 1) HIPCHK(launch_gemm_group(g, epi, st));
            else if (p->kernel == 4) HIPCHK(launch_gemm_window(g, epi, st));
            else HIPCHK(launch_gemm_fused(g, epi, st));   // 0 fused, 3 paired digits (fused over P pair planes)
        }
    }
    return OZAIE_OK;
}

int plan_execute(ozaie_plan p, int transA, int transB, const double *A, int64_t lda,
                 const double *B, int64_t ldb, double *C, int64_t ldc, hipStream_t st) {
    if (!A || !B || !C) { g_last_error = "null pointer"; return OZAIE_ERR_ARG; }
    const int64_t minlda = transA ? p->M : p->K, minldb = transB ? p->K : p->N;
    if (lda < minlda || ldb < minldb || ldc < p->N) { g_last_error = "leading dimension too small"; return OZAIE_ERR_ARG; }
    DeviceGuard dg;
    HIPCHK(dg.enter(p->device));
    if (p->timed) HIPCHK(hipEventRecord(p->ev[0], st));
    int rc = run_split(p, view_A(p, transA, A, lda), view_B(p, transB, B, ldb), st);
    if (!rc && p->timed) HIPCHK(hipEventRecord(p->ev[1], st));
    if (!rc) rc = run_passes(p, C, ldc, EPI_C, nullptr, st);
    if (!rc && p->timed && !p->crt) HIPCHK(hipEventRecord(p->ev[2], st));
    if (!rc && p->timed) HIPCHK(hipEventRecord(p->ev[3], st));
    return rc;
}

bool is_device_ptr(const void *ptr) {
    hipPointerAttribute_t attr;
    std::memset(&attr, 0, sizeof(attr));
    hipError_t e = hipPointerGetAttributes(&attr, ptr);
    if (e != hipSuccess) { (void)hipGetLastError(); return false; }
    return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

ozaie_plan g_plan = nullptr;

// ---- workspace of the allocation-free entry point (ozaie_dgemm_noalloc)
// One static device array per GPU the library runs on, held until the process ends.  In
// host code `scratch` only names the symbol: its device address (per device) comes from
// hipGetSymbolAddress on the first call on that device and is the only form passed to
// kernels or copies.
constexpr int kNoAllocMaxDevices = 64;
static __device__ __align__(256) char scratch[OZAIE_NOALLOC_BYTES];
char *g_scratch_ptr[kNoAllocMaxDevices] = {};

int noalloc_workspace(int dev, char **out) {
    if (dev < 0 || dev >= kNoAllocMaxDevices) { g_last_error = "device index beyond the workspace table"; return OZAIE_ERR_ARG; }
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) { g_last_error = "hipGetSymbolAddress returned no address"; return OZAIE_ERR_HIP; }
        g_scratch_ptr[dev] = static_cast<char *>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return OZAIE_OK;
}

// Bytes of the plan's intermediates in the workspace (each buffer 256-byte aligned).
int64_t noalloc_bytes(const ozaie_plan p) {
    int64_t off = 0;
    auto take = [&](int64_t bytes) { off += round_up(bytes, 256); };
    take((int64_t)p->d * p->Mp * p->Kp);
    take((int64_t)p->d * p->Np * p->Kp);
    take(4 * p->Mp);
    take(4 * p->Np);
    if (!p->crt && (int64_t)p->nchunks * p->ngroups > 1) take(16 * p->M * p->N);
    if (p->crt) {
        take((int64_t)p->d * p->Mp * p->Np);
        take(8 * p->Mp * crt_chunks(p->Kb));
        take(8 * p->Np * crt_chunks(p->Kb));
    }
    if (p->pairs) take(4 * p->Mp * p->Np);
    return off;
}

int setup_global(int64_t M, int64_t N, int64_t K, int b, int d) {
    if (g_plan) { plan_free(g_plan); g_plan = nullptr; }
    return plan_new(&g_plan, M, N, K, b, d, g_opt_rounding, true);
}

}  // namespace

extern "C" {

int32_t ozAIE_setup(int M, int N, int K, int mode) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (M < 1 || N < 1 || K < 1) {
        std::fprintf(stderr, "ozAIE_setup: dimensions must be >= 1\n");
        return OZAIE_ERR_DIMS;
    }
    // int(std::log2(K)) as the reference computes it: K need not be a power of two
    int log2k = (int)std::log2((double)K);
    if (g_opt_strict && (log2k < 7 || log2k > 11 || M != K || N != K)) {
        std::fprintf(stderr, "Only square matrices are currently supported\n");
        std::fprintf(stderr, "Single side dimension : power of 2 between 128 and 2048 (inclusive)\n");
        return OZAIE_ERR_DIMS;
    }
    int b, d;
    switch (mode) {
        case OZAIE_MODE_ACCURATE: {
            const int idx = std::min(std::max(log2k, 7), 14) - 7;
            b = kAccurate[idx][0]; d = kAccurate[idx][1];
            break;
        }
        case OZAIE_MODE_FAST: b = 7; d = 8; break;
        case OZAIE_MODE_FAST3: b = 7; d = 3; break;
        case OZAIE_MODE_ACCURATE7: b = 7; d = 7; break;
        case OZAIE_MODE_CRT:
            if (g_plan) { plan_free(g_plan); g_plan = nullptr; }
            return plan_new(&g_plan, M, N, K, 0, kCrtDefaultMod, g_opt_rounding, true, true);
        default:
            g_last_error = "unknown mode";
            return OZAIE_ERR_ARG;
    }
    return setup_global(M, N, K, b, d);
}

int32_t ozAIE_setup_slices(int M, int N, int K, int bits, int slices) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (M < 1 || N < 1 || K < 1) return OZAIE_ERR_DIMS;
    return setup_global(M, N, K, bits, slices);
}

// ---- host-pointer transfers of ozAIE_DGEMM
// The runtime moves pageable host memory at 15-20 GB/s through its own staging; pinned
// memory moves at the link's rate (about 57 GB/s each way on MI355X).  Two pinned chunks of
// kStageBytes alternate: the host copies one chunk between the caller's array and pinned
// memory (threaded memcpy, about 90 GB/s) while the DMA engine moves the other.  The chunks
// are allocated (and their pages touched) at the plan's first host-pointer call.
namespace {
constexpr size_t kStageBytes = (size_t)32 << 20;

int stage_ready(ozaie_plan p) {
    for (int s = 0; s < 2; s++) {
        if (p->pin[s]) continue;
        HIPCHK(hipHostMalloc((void **)&p->pin[s], kStageBytes, hipHostMallocDefault));
        std::memset(p->pin[s], 0, kStageBytes);
        HIPCHK(hipEventCreateWithFlags(&p->pin_ev[s], hipEventDisableTiming));
    }
    return OZAIE_OK;
}

// Pinned chunk s may be written by the host: its last DMA has completed.
int stage_free(ozaie_plan p, int s) {
    if (p->pin_used[s]) HIPCHK(hipEventSynchronize(p->pin_ev[s]));
    return OZAIE_OK;
}

// n bytes host src -> device dst, the DMAs enqueued on st.  On return the host has read all
// of src; the last DMAs may still run.
int stage_h2d(ozaie_plan p, void *dst, const void *src, size_t n, hipStream_t st) {
    for (size_t i = 0, off = 0; off < n; i++, off += kStageBytes) {
        const int s = (int)(i & 1);
        const size_t len = std::min(kStageBytes, n - off);
        if (int rc = stage_free(p, s)) return rc;
        parallel_copy(p->pin[s], static_cast<const char *>(src) + off, len);
        HIPCHK(hipMemcpyAsync(static_cast<char *>(dst) + off, p->pin[s], len, hipMemcpyHostToDevice, st));
        HIPCHK(hipEventRecord(p->pin_ev[s], st));
        p->pin_used[s] = true;
    }
    return OZAIE_OK;
}

// n bytes device src -> host dst after the work already on st; returns when dst holds them.
// The DMA of chunk i + 1 runs while the host copies chunk i out; chunk i + 1's pinned buffer
// last held chunk i - 1, which the host copied out in the previous step.
int stage_d2h(ozaie_plan p, void *dst, const void *src, size_t n, hipStream_t st) {
    const size_t nch = (n + kStageBytes - 1) / kStageBytes;
    auto issue = [&](size_t i) -> int {
        const int s = (int)(i & 1);
        const size_t off = i * kStageBytes, len = std::min(kStageBytes, n - off);
        HIPCHK(hipMemcpyAsync(p->pin[s], static_cast<const char *>(src) + off, len, hipMemcpyDeviceToHost, st));
        HIPCHK(hipEventRecord(p->pin_ev[s], st));
        p->pin_used[s] = true;
        return OZAIE_OK;
    };
    if (nch > 0)
        if (int rc = issue(0)) return rc;
    for (size_t i = 0; i < nch; i++) {
        if (i + 1 < nch)
            if (int rc = issue(i + 1)) return rc;
        const int s = (int)(i & 1);
        const size_t off = i * kStageBytes, len = std::min(kStageBytes, n - off);
        HIPCHK(hipEventSynchronize(p->pin_ev[s]));
        parallel_copy(static_cast<char *>(dst) + off, p->pin[s], len);
    }
    return OZAIE_OK;
}
}  // namespace

int32_t ozAIE_DGEMM(int M, int N, int K, const double *A, const double *B, double *C) {
    std::lock_guard<std::mutex> lk(g_mu);
    ozaie_plan p = g_plan;
    if (!p) { g_last_error = "ozAIE_DGEMM before ozAIE_setup"; return OZAIE_ERR_NOSETUP; }
    if (M != p->M || N != p->N || K != p->K) { g_last_error = "dimensions differ from setup"; return OZAIE_ERR_MISMATCH; }
    if (!A || !B || !C) { g_last_error = "null pointer"; return OZAIE_ERR_ARG; }
    DeviceGuard dg;
    HIPCHK(dg.enter(p->device));
    // The plan's stream is a blocking stream: it starts after the work already queued on the
    // legacy default stream (fills or producers of device-resident A, B, C).
    hipStream_t st = p->stream;
    const bool devA = is_device_ptr(A), devB = is_device_ptr(B), devC = is_device_ptr(C);
    const double *a = A, *bb = B;
    double *c = devC ? C : p->dC;
    if (!devA || !devB || !devC)
        if (int rc = stage_ready(p)) return rc;
    if (!devA) {
        if (int rc = stage_h2d(p, p->dA, A, sizeof(double) * (size_t)M * K, st)) return rc;
        a = p->dA;
    }
    if (!devB) {
        if (int rc = stage_h2d(p, p->dB, B, sizeof(double) * (size_t)K * N, st)) return rc;
        bb = p->dB;
    }
    int rc = plan_execute(p, 0, 0, a, K, bb, N, c, N, st);
    if (rc) return rc == OZAIE_ERR_HIP ? OZAIE_ERR_INCOMPLETE : rc;
    if (!devC)
        if (int rc2 = stage_d2h(p, C, p->dC, sizeof(double) * (size_t)M * N, st)) return rc2;
    hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) { set_hip_error(e, "ozAIE_DGEMM"); return OZAIE_ERR_INCOMPLETE; }
    return OZAIE_OK;
}

int32_t ozaie_set_option(int option, int value) {
    std::lock_guard<std::mutex> lk(g_mu);
    switch (option) {
        case OZAIE_OPT_ROUNDING:
            if (value != OZAIE_ROUND_TRUNCATE && value != OZAIE_ROUND_NEAREST) return OZAIE_ERR_ARG;
            g_opt_rounding = value; return OZAIE_OK;
        case OZAIE_OPT_STRICT: g_opt_strict = value ? 1 : 0; return OZAIE_OK;
        case OZAIE_OPT_KCHUNK:
            if (value < 0) return OZAIE_ERR_ARG;
            g_opt_kchunk = value; return OZAIE_OK;
        case OZAIE_OPT_KERNEL:
            if (value < 0 || value > 6) return OZAIE_ERR_ARG;
            g_opt_kernel = value; return OZAIE_OK;
        case OZAIE_OPT_PHASE_TIMING: g_opt_timing = value ? 1 : 0; return OZAIE_OK;
        case OZAIE_OPT_CRT_GROUP_MODULI:
            if (value < 0 || value > kCrtMaxMod) return OZAIE_ERR_ARG;
            g_opt_crt_mpw = value; return OZAIE_OK;
        default: return OZAIE_ERR_ARG;
    }
}

int32_t ozaie_get_option(int option) {
    std::lock_guard<std::mutex> lk(g_mu);
    switch (option) {
        case OZAIE_OPT_ROUNDING: return g_opt_rounding;
        case OZAIE_OPT_STRICT: return g_opt_strict;
        case OZAIE_OPT_KCHUNK: return g_opt_kchunk;
        case OZAIE_OPT_KERNEL: return g_opt_kernel;
        case OZAIE_OPT_PHASE_TIMING: return g_opt_timing;
        case OZAIE_OPT_CRT_GROUP_MODULI: return g_opt_crt_mpw;
        default: return -OZAIE_ERR_ARG;
    }
}

static void fill_info(const ozaie_plan p, int32_t *out) {
    out[0] = p->b; out[1] = p->d;
    out[2] = (int32_t)std::min<int64_t>(p->kcb * kBlk, INT32_MAX);
    out[3] = p->nchunks * p->ngroups;
    out[4] = p->kernel;
}

int32_t ozaie_plan_info(int32_t *out5) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_plan) return OZAIE_ERR_NOSETUP;
    if (!out5) return OZAIE_ERR_ARG;
    fill_info(g_plan, out5);
    return OZAIE_OK;
}

int32_t ozaie_release(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_plan) { plan_free(g_plan); g_plan = nullptr; }
    return OZAIE_OK;
}

int32_t ozaie_plan_create(ozaie_plan *plan, int64_t M, int64_t N, int64_t K, int bits, int slices,
                          int flags) {
    if (!plan) return OZAIE_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_mu);
    return plan_new(plan, M, N, K, bits, slices, (flags & 1) ? OZAIE_ROUND_NEAREST : OZAIE_ROUND_TRUNCATE, false,
                    (flags & OZAIE_PLAN_CRT) != 0);
}

int32_t ozaie_plan_destroy(ozaie_plan plan) {
    std::lock_guard<std::mutex> lk(g_mu);
    plan_free(plan);
    return OZAIE_OK;
}

int32_t ozaie_plan_query(ozaie_plan plan, int32_t *out5) {
    if (!plan || !out5) return OZAIE_ERR_ARG;
    fill_info(plan, out5);
    return OZAIE_OK;
}

int32_t ozaie_dgemm(ozaie_plan plan, int transA, int transB, const double *A, int64_t lda,
                    const double *B, int64_t ldb, double *C, int64_t ldc, void *stream) {
    if (!plan) return OZAIE_ERR_ARG;
    return plan_execute(plan, transA, transB, A, lda, B, ldb, C, ldc, (hipStream_t)stream);
}

int32_t ozaie_dgemm_noalloc(int64_t M, int64_t N, int64_t K, int bits, int slices, int flags, int transA,
                            int transB, const double *A, int64_t lda, const double *B, int64_t ldb, double *C,
                            int64_t ldc, void *stream) {
    std::lock_guard<std::mutex> lk(g_mu);
    int rc = ensure_device();
    if (rc) return rc;
    // a plan on the stack: the same configuration as ozaie_plan_create, no device buffers
    ozaie_plan_s p;
    const int rounding = (flags & OZAIE_PLAN_NEAREST) ? OZAIE_ROUND_NEAREST : OZAIE_ROUND_TRUNCATE;
    rc = (flags & OZAIE_PLAN_CRT) ? plan_configure_crt(&p, M, N, K, bits, slices, rounding)
                                  : plan_configure(&p, M, N, K, bits, slices, rounding);
    if (rc) return rc;
    const int64_t need = noalloc_bytes(&p);
    if (need > (int64_t)OZAIE_NOALLOC_BYTES) {
        g_last_error = "ozaie_dgemm_noalloc: this product needs " + std::to_string(need) +
                       " bytes of intermediates, the static workspace holds " +
                       std::to_string((long long)OZAIE_NOALLOC_BYTES);
        return OZAIE_ERR_NOMEM;
    }
    HIPCHK(hipGetDevice(&p.device));
    HIPCHK(hipDeviceGetAttribute(&p.cus, hipDeviceAttributeMultiprocessorCount, p.device));
    char *ws = nullptr;
    rc = noalloc_workspace(p.device, &ws);
    if (rc) return rc;
    int64_t off = 0;
    auto take = [&](int64_t bytes) { char *q = ws + off; off += round_up(bytes, 256); return q; };
    p.SA = reinterpret_cast<int8_t *>(take((int64_t)p.d * p.Mp * p.Kp));
    p.SB = reinterpret_cast<int8_t *>(take((int64_t)p.d * p.Np * p.Kp));
    p.Er = reinterpret_cast<int32_t *>(take(4 * p.Mp));
    p.Ec = reinterpret_cast<int32_t *>(take(4 * p.Np));
    if (!p.crt && (int64_t)p.nchunks * p.ngroups > 1) p.T = reinterpret_cast<__int128 *>(take(16 * p.M * p.N));
    if (p.crt) {
        p.R = reinterpret_cast<uint32_t *>(take((int64_t)p.d * p.Mp * p.Np));
        p.stA = reinterpret_cast<uint64_t *>(take(8 * p.Mp * crt_chunks(p.Kb)));
        p.stB = reinterpret_cast<uint64_t *>(take(8 * p.Np * crt_chunks(p.Kb)));
    }
    if (p.pairs) p.X = reinterpret_cast<int32_t *>(take(4 * p.Mp * p.Np));
    // Every intermediate is fully written before it is read (exponent arrays that are
    // combined with atomicMax are cleared by a fill kernel first), so the previous call's
    // contents never leak into this one.
    return plan_execute(&p, transA, transB, A, lda, B, ldb, C, ldc, (hipStream_t)stream);
}

int32_t ozaie_plan_phase_times(ozaie_plan plan, float *out3) {
    if (!plan || !out3) return OZAIE_ERR_ARG;
    if (!plan->timed) { g_last_error = "plan created without OZAIE_OPT_PHASE_TIMING"; return OZAIE_ERR_ARG; }
    HIPCHK(hipEventSynchronize(plan->ev[3]));
    HIPCHK(hipEventElapsedTime(&out3[0], plan->ev[0], plan->ev[1]));
    HIPCHK(hipEventElapsedTime(&out3[1], plan->ev[1], plan->ev[2]));
    HIPCHK(hipEventElapsedTime(&out3[2], plan->ev[0], plan->ev[3]));
    return OZAIE_OK;
}

int32_t ozaie_debug_convert(const int64_t *T_hi, const uint64_t *T_lo, const int32_t *es, int64_t n, int bits,
                            int slices, int rounding, double *out, double *out64) {
    if (!T_hi || !T_lo || !es || !out || !out64 || n < 1) return OZAIE_ERR_ARG;
    if (bits < 1 || bits > kMaxBits || slices < 1 || slices > kMaxSlices || (rounding != 0 && rounding != 1))
        return OZAIE_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_mu);
    int rc = ensure_device();
    if (rc) return rc;
    char *buf = nullptr;
    const size_t w = sizeof(int64_t) * (size_t)n;
    HIPCHK(hipMalloc(&buf, 5 * w));
    int64_t *dh = reinterpret_cast<int64_t *>(buf);
    uint64_t *dl = reinterpret_cast<uint64_t *>(buf + w);
    int32_t *de = reinterpret_cast<int32_t *>(buf + 2 * w);
    double *o1 = reinterpret_cast<double *>(buf + 3 * w), *o2 = reinterpret_cast<double *>(buf + 4 * w);
    hipError_t e = hipMemcpy(dh, T_hi, w, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dl, T_lo, w, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(de, es, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_debug_convert(dh, dl, de, n, bits, slices, rounding, o1, o2, 0);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(out, o1, w, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(out64, o2, w, hipMemcpyDeviceToHost);
    (void)hipFree(buf);
    if (e != hipSuccess) return set_hip_error(e, "ozaie_debug_convert");
    return OZAIE_OK;
}

int32_t ozaie_debug_split(const double *X, int64_t rows, int64_t K, int64_t sr, int64_t sk,
                          int bits, int slices, int32_t *E_out, int8_t *S_out) {
    if (!X || !E_out || !S_out) return OZAIE_ERR_ARG;
    if (rows < 1 || K < 1) return OZAIE_ERR_DIMS;
    if (bits < 1 || bits > kMaxBits || slices < 1 || slices > kMaxSlices || bits * slices > 96) return OZAIE_ERR_ARG;
    if (sr != 1 && sk != 1) return OZAIE_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_mu);
    int rc = ensure_device();
    if (rc) return rc;
    const int64_t Rp = round_up(rows, kRowPad), Kp = round_up(K, 2 * kBlk);
    int32_t *E = nullptr;
    int8_t *S = nullptr;
    HIPCHK(hipMalloc(&E, sizeof(int32_t) * Rp));
    hipError_t e = hipMalloc(&S, (size_t)slices * Rp * Kp);
    if (e != hipSuccess) { (void)hipFree(E); return set_hip_error(e, "hipMalloc"); }
    OperandView v;
    v.X = X; v.sr = sr; v.sk = sk; v.R = rows; v.K = K;
    e = launch_maxexp(v, E, Rp, 0);
    if (e == hipSuccess) e = launch_split(v, E, bits, slices, S, Rp, Kp, 0);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    std::vector<int8_t> hs;
    if (e == hipSuccess) e = hipMemcpy(E_out, E, sizeof(int32_t) * rows, hipMemcpyDeviceToHost);
    if (e == hipSuccess) {
        hs.resize((size_t)slices * Rp * Kp);
        e = hipMemcpy(hs.data(), S, hs.size(), hipMemcpyDeviceToHost);
    }
    (void)hipFree(E);
    (void)hipFree(S);
    if (e != hipSuccess) return set_hip_error(e, "ozaie_debug_split");
    // Undo the block layout: S[j][rb][kb][ (k%32/16)*512 + (r%32)*16 + k%16 ].
    const int64_t Kb = Kp / kBlk;
    for (int j = 0; j < slices; j++)
        for (int64_t r = 0; r < rows; r++)
            for (int64_t k = 0; k < K; k++) {
                const int64_t off = (((int64_t)j * (Rp / kBlk) + r / kBlk) * Kb + k / kBlk) * kBlkBytes +
                                    ((k % kBlk) / 16) * 512 + (r % kBlk) * 16 + (k % 16);
                S_out[((int64_t)j * rows + r) * K + k] = hs[off];
            }
    return OZAIE_OK;
}

int32_t ozaie_debug_crt(const double *A, const double *B, int64_t M, int64_t N, int64_t K, int moduli, int bits,
                        int8_t *SA_out, int8_t *SB_out, uint8_t *G_out) {
    if (!A || !B || !SA_out || !SB_out || !G_out) return OZAIE_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_mu);
    int rc = ensure_device();
    if (rc) return rc;
    ozaie_plan p = nullptr;
    rc = plan_new(&p, M, N, K, bits, moduli, OZAIE_ROUND_TRUNCATE, false, true);
    if (rc) return rc;
    rc = run_split(p, view_A(p, 0, A, K), view_B(p, 0, B, N), 0);
    if (!rc) rc = run_passes(p, nullptr, N, EPI_BINS, nullptr, 0);
    hipError_t e = rc ? hipSuccess : hipDeviceSynchronize();
    std::vector<int8_t> ha, hb;
    std::vector<uint32_t> hr;
    if (!rc && e == hipSuccess) {
        ha.resize((size_t)moduli * p->Mp * p->Kp);
        hb.resize((size_t)moduli * p->Np * p->Kp);
        hr.resize((size_t)moduli * (p->Mp / 4) * p->Np);
        e = hipMemcpy(ha.data(), p->SA, ha.size(), hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(hb.data(), p->SB, hb.size(), hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(hr.data(), p->R, hr.size() * 4, hipMemcpyDeviceToHost);
    }
    if (!rc && e != hipSuccess) rc = set_hip_error(e, "ozaie_debug_crt");
    if (!rc) {
        // undo the slice-block layout (as ozaie_debug_split) and the four-row words of R
        const int64_t Kb = p->Kb;
        auto unblock = [&](const std::vector<int8_t> &h, int64_t rows, int64_t Rp, int8_t *out) {
            for (int j = 0; j < moduli; j++)
                for (int64_t r = 0; r < rows; r++)
                    for (int64_t k = 0; k < K; k++) {
                        const int64_t off = (((int64_t)j * (Rp / kBlk) + r / kBlk) * Kb + k / kBlk) * kBlkBytes +
                                            ((k % kBlk) / 16) * 512 + (r % kBlk) * 16 + (k % 16);
                        out[((int64_t)j * rows + r) * K + k] = h[off];
                    }
        };
        unblock(ha, M, p->Mp, SA_out);
        unblock(hb, N, p->Np, SB_out);
        for (int j = 0; j < moduli; j++)
            for (int64_t r = 0; r < M; r++)
                for (int64_t c = 0; c < N; c++)
                    G_out[((int64_t)j * M + r) * N + c] = (uint8_t)(
                        ((int)(int8_t)(uint8_t)(hr[((int64_t)j * (p->Mp / 4) + r / 4) * p->Np + c] >> (8 * (r % 4))) +
                         crt_modulus(j)) % crt_modulus(j));   // stored signed, reported in [0, p)
    }
    plan_free(p);
    return rc;
}

static int debug_bins_impl(int64_t M, int64_t N, int64_t K, const double *A, const double *B, int bits,
                           int slices, int kernel, int32_t *dbins, ozaie_plan *pout) {
    ozaie_plan p = nullptr;
    const int save_kernel = g_opt_kernel;
    g_opt_kernel = kernel;
    int rc = plan_new(&p, M, N, K, bits, slices, OZAIE_ROUND_TRUNCATE, false, false, false);
    g_opt_kernel = save_kernel;
    if (rc) return rc;
    if (p->nchunks != 1) { plan_free(p); g_last_error = "K needs more than one pass"; return OZAIE_ERR_DIMS; }
    rc = run_split(p, view_A(p, 0, A, K), view_B(p, 0, B, N), 0);
    if (!rc) rc = run_passes(p, nullptr, N, EPI_BINS, dbins, 0);
    if (rc) { plan_free(p); return rc; }
    *pout = p;
    return OZAIE_OK;
}

int32_t ozaie_debug_bins(int64_t M, int64_t N, int64_t K, const double *A, const double *B, int bits,
                         int slices, int kernel, int32_t *bins_out) {
    if (!A || !B || !bins_out) return OZAIE_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_mu);
    int rc = ensure_device();
    if (rc) return rc;
    if (M < 1 || N < 1 || K < 1) return OZAIE_ERR_DIMS;
    int32_t *dbins = nullptr;
    HIPCHK(hipMalloc(&dbins, sizeof(int32_t) * (size_t)std::max(slices, 1) * M * N));
    ozaie_plan p = nullptr;
    rc = debug_bins_impl(M, N, K, A, B, bits, slices, kernel, dbins, &p);
    if (!rc) {
        hipError_t e = hipDeviceSynchronize();
        if (e == hipSuccess) e = hipMemcpy(bins_out, dbins, sizeof(int32_t) * (size_t)slices * M * N, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = set_hip_error(e, "ozaie_debug_bins");
    }
    if (p) plan_free(p);
    (void)hipFree(dbins);
    return rc;
}

int32_t ozaie_paired_output(int64_t M, int64_t N, int64_t K, const double *A, const double *B, int bits,
                            int slices, int64_t *out) {
    if (!A || !B || !out) return OZAIE_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_mu);
    int rc = ensure_device();
    if (rc) return rc;
    if (M < 1 || N < 1 || K < 1 || K > INT32_MAX) return OZAIE_ERR_DIMS;
    int32_t *dbins = nullptr;
    int64_t *dout = nullptr;
    HIPCHK(hipMalloc(&dbins, sizeof(int32_t) * (size_t)std::max(slices, 1) * M * N));
    hipError_t e = hipMalloc(&dout, sizeof(int64_t) * 2 * (size_t)M * N);
    if (e != hipSuccess) { (void)hipFree(dbins); return set_hip_error(e, "hipMalloc"); }
    ozaie_plan p = nullptr;
    rc = debug_bins_impl(M, N, K, A, B, bits, slices, 0, dbins, &p);
    if (!rc) {
        // group sizes n0/n1 of the reference format, ozaki_base.cc:151-161
        const int log2k = __builtin_ctz((unsigned)K);
        const int alignable = 31 / bits + 1;
        const int numer = bits * (slices + 1) - 51 + log2k;
        const int minpos = numer < 0 ? 1 : numer / bits + 1;
        const int n0 = std::min(slices, std::max(minpos, alignable));
        e = launch_pack_paired(dbins, M, N, bits, slices, n0, p->Er, p->Ec, dout, 0);
        if (e == hipSuccess) e = hipDeviceSynchronize();
        if (e == hipSuccess) e = hipMemcpy(out, dout, sizeof(int64_t) * 2 * (size_t)M * N, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = set_hip_error(e, "ozaie_paired_output");
    }
    if (p) plan_free(p);
    (void)hipFree(dbins);
    (void)hipFree(dout);
    return rc;
}

const char *ozaie_last_error(void) { return g_last_error.c_str(); }

int32_t ozaie_version(void) { return 100; }

}  // extern "C"
