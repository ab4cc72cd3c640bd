"""ozaie_amd -- MI355X-native Ozaki-scheme DGEMM emulation.

Python host mirror of the reference's API (include/oz_aie.h of atlarge-research/ozAIE):
``ozAIE_setup(M, N, K, mode)`` and ``ozAIE_DGEMM(M, N, K, A, B, C)`` with
``Mode.accurate`` / ``Mode.fast`` and the same status codes, plus explicit plans on
device tensors, inspection of slices/bins and the reference wire format.

Everything calls libozaie.so (HIP kernels for gfx950 behind a C ABI, built from
ozaie_amd/csrc by ``ozaie_amd.build``).  There is no CPU fallback: if the library is
missing, every call raises.
"""
import ctypes
import enum
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libozaie.so")

OK = 0
ERR_DIMS = 1
ERR_ENV = 2
ERR_INCOMPLETE = 3
ERR_BUILD = 4
ERR_HIP = 5
ERR_NODEVICE = 6
ERR_NOSETUP = 7
ERR_MISMATCH = 8
ERR_NOMEM = 9
ERR_ARG = 10
ERR_OVERFLOW = 11

OPT_ROUNDING = 1
OPT_STRICT = 2
OPT_KCHUNK = 3
OPT_KERNEL = 4
OPT_PHASE_TIMING = 5
OPT_CRT_GROUP_MODULI = 6
ROUND_TRUNCATE = 0
ROUND_NEAREST = 1


# kernel ids of ozaie_plan_info / ozaie_plan_query (out[4])
KERNEL_NAMES = ("fused", "group", "crt", "pairs", "window")


class Mode(enum.IntEnum):
    """oz_aie::Mode (include/oz_aie.h:82) plus the int8-rate presets of this library."""
    accurate = 0   # reference table: b = 3, d = 29 .. 25 by log2 K (b = 4, d = 18 at 2^14+)
    fast = 1       # reference: b = 7, d = 8
    fast3 = 2      # b = 7, d = 3
    accurate7 = 3  # b = 7, d = 7
    crt = 4        # no reference counterpart: 16 int8 residue GEMMs, exact integer product


# C symbols declared in include/ozaie.h (the tests check that every one is exported).
EXPORTS = [
    "ozAIE_setup", "ozAIE_setup_slices", "ozAIE_DGEMM", "ozaie_set_option", "ozaie_get_option",
    "ozaie_plan_info", "ozaie_release", "ozaie_plan_create", "ozaie_plan_destroy",
    "ozaie_plan_query", "ozaie_dgemm", "ozaie_debug_split", "ozaie_debug_bins", "ozaie_debug_crt",
    "ozaie_debug_convert",
    "ozaie_paired_output", "ozaie_last_error", "ozaie_version", "ozaie_plan_phase_times",
    "ozaie_exact_entries", "ozaie_group_create", "ozaie_group_run", "ozaie_group_query",
    "ozaie_group_configs", "ozaie_group_status", "ozaie_group_destroy", "ozaie_dgemm_grouped", "ozaie_autotune",
    "ozaie_tune_get", "ozaie_tune_set", "ozaie_autotune_crt", "ozaie_tune_get_crt", "ozaie_tune_set_crt",
    "ozaie_dgemm_noalloc", "ozaie_probe_mfma", "ozaie_tune_get_slices", "ozaie_tune_set_slices",
]

NOALLOC_BYTES = 256 << 20   # OZAIE_NOALLOC_BYTES: the static workspace of ozaie_dgemm_noalloc


class OzaieError(RuntimeError):
    def __init__(self, code, what):
        self.code = code
        msg = f"{what} failed with status {code}"
        try:
            last = _lib_handle.ozaie_last_error().decode() if _lib_handle is not None else ""
        except Exception:
            last = ""
        if last:
            msg += f": {last}"
        super().__init__(msg)


_lib_handle = None


def lib():
    """Load libozaie.so (raises if it has not been built)."""
    global _lib_handle
    if _lib_handle is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m ozaie_amd.build` "
                               "(there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        _declare(L)
        _lib_handle = L
    return _lib_handle


def _declare(L):
    P = ctypes.c_void_p
    i32 = ctypes.c_int32
    i64 = ctypes.c_int64
    c_int = ctypes.c_int
    L.ozAIE_setup.argtypes = [c_int, c_int, c_int, c_int]
    L.ozAIE_setup.restype = i32
    L.ozAIE_setup_slices.argtypes = [c_int, c_int, c_int, c_int, c_int]
    L.ozAIE_setup_slices.restype = i32
    L.ozAIE_DGEMM.argtypes = [c_int, c_int, c_int, P, P, P]
    L.ozAIE_DGEMM.restype = i32
    L.ozaie_set_option.argtypes = [c_int, c_int]
    L.ozaie_set_option.restype = i32
    L.ozaie_get_option.argtypes = [c_int]
    L.ozaie_get_option.restype = i32
    L.ozaie_plan_info.argtypes = [P]
    L.ozaie_plan_info.restype = i32
    L.ozaie_release.argtypes = []
    L.ozaie_release.restype = i32
    L.ozaie_plan_create.argtypes = [ctypes.POINTER(P), i64, i64, i64, c_int, c_int, c_int]
    L.ozaie_plan_create.restype = i32
    L.ozaie_plan_destroy.argtypes = [P]
    L.ozaie_plan_destroy.restype = i32
    L.ozaie_plan_query.argtypes = [P, P]
    L.ozaie_plan_query.restype = i32
    L.ozaie_dgemm.argtypes = [P, c_int, c_int, P, i64, P, i64, P, i64, P]
    L.ozaie_dgemm.restype = i32
    L.ozaie_debug_split.argtypes = [P, i64, i64, i64, i64, c_int, c_int, P, P]
    L.ozaie_debug_split.restype = i32
    L.ozaie_debug_crt.argtypes = [P, P, i64, i64, i64, c_int, c_int, P, P, P]
    L.ozaie_debug_crt.restype = i32
    L.ozaie_debug_convert.argtypes = [P, P, P, i64, c_int, c_int, c_int, P, P]
    L.ozaie_debug_convert.restype = i32
    L.ozaie_debug_bins.argtypes = [i64, i64, i64, P, P, c_int, c_int, c_int, P]
    L.ozaie_debug_bins.restype = i32
    L.ozaie_paired_output.argtypes = [i64, i64, i64, P, P, c_int, c_int, P]
    L.ozaie_paired_output.restype = i32
    L.ozaie_plan_phase_times.argtypes = [P, P]
    L.ozaie_plan_phase_times.restype = i32
    L.ozaie_exact_entries.argtypes = [P, i64, P, i64, i64, P, P, i64, P]
    L.ozaie_exact_entries.restype = i32
    L.ozaie_group_create.argtypes = [ctypes.POINTER(P), c_int, P, P, P, c_int, c_int, c_int]
    L.ozaie_group_create.restype = i32
    L.ozaie_group_run.argtypes = [P, P, P, P, P]
    L.ozaie_group_run.restype = i32
    L.ozaie_group_query.argtypes = [P, P]
    L.ozaie_group_query.restype = i32
    L.ozaie_group_configs.argtypes = [P, P]
    L.ozaie_group_configs.restype = i32
    L.ozaie_group_status.argtypes = [P, P]
    L.ozaie_group_status.restype = i32
    L.ozaie_group_destroy.argtypes = [P]
    L.ozaie_group_destroy.restype = i32
    L.ozaie_dgemm_grouped.argtypes = [c_int, P, P, P, P, P, P, c_int, c_int, P]
    L.ozaie_dgemm_grouped.restype = i32
    L.ozaie_autotune.argtypes = [c_int, c_int, c_int, P]
    L.ozaie_autotune.restype = i32
    L.ozaie_tune_get.argtypes = [c_int, P]
    L.ozaie_tune_get.restype = i32
    L.ozaie_tune_set.argtypes = [c_int, P]
    L.ozaie_tune_set.restype = i32
    L.ozaie_autotune_crt.argtypes = [c_int, c_int, P]
    L.ozaie_autotune_crt.restype = i32
    L.ozaie_tune_get_crt.argtypes = [c_int, P]
    L.ozaie_tune_get_crt.restype = i32
    L.ozaie_tune_set_crt.argtypes = [c_int, P]
    L.ozaie_tune_set_crt.restype = i32
    L.ozaie_dgemm_noalloc.argtypes = [i64, i64, i64, c_int, c_int, c_int, c_int, c_int, P, i64, P, i64, P, i64, P]
    L.ozaie_dgemm_noalloc.restype = i32
    L.ozaie_tune_get_slices.argtypes = [P]
    L.ozaie_tune_get_slices.restype = i32
    L.ozaie_tune_set_slices.argtypes = [P]
    L.ozaie_tune_set_slices.restype = i32
    L.ozaie_probe_mfma.argtypes = [c_int, c_int, P, P]
    L.ozaie_probe_mfma.restype = i32
    L.ozaie_last_error.argtypes = []
    L.ozaie_last_error.restype = ctypes.c_char_p
    L.ozaie_version.argtypes = []
    L.ozaie_version.restype = i32


def _check(rc, what):
    if rc != OK:
        raise OzaieError(rc, what)
    return rc


def _ptr(x):
    """Raw pointer of a numpy array or a torch tensor."""
    if isinstance(x, np.ndarray):
        return ctypes.c_void_p(x.ctypes.data)
    return ctypes.c_void_p(x.data_ptr())


def mode_slices(mode, K):
    """(b, d) that ozAIE_setup picks for `mode` (reference table include/oz_aie.h:60-80)."""
    import math
    mode = Mode(mode)
    if mode == Mode.accurate:
        table = {7: (3, 29), 8: (3, 28), 9: (3, 27), 10: (3, 27), 11: (3, 26), 12: (3, 25),
                 13: (3, 25), 14: (4, 18)}
        lk = min(max(int(math.log2(K)), 7), 14)
        return table[lk]
    if mode == Mode.crt:
        return 0, 16          # bits 0: shifts by the norm rule, 16 moduli
    return {Mode.fast: (7, 8), Mode.fast3: (7, 3), Mode.accurate7: (7, 7)}[mode]


CRT_MODULI = (256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 211, 199, 197, 193, 191,
              181, 179, 173, 167)


def crt_beta(moduli, K):
    """Bits per operand of the CRT mode: the largest beta with K 2^(2 beta) <= 2^(L-2),
    L = floor(log2 prod p), at most 62 (mirrors the library's planner)."""
    L = math.prod(CRT_MODULI[:moduli]).bit_length() - 1
    lk = (int(K) - 1).bit_length()
    beta = min(62, (L - 2 - lk) // 2)
    return beta if beta >= 1 else -1


# ---------------------------------------------------------------- reference-style API
def ozAIE_setup(M, N, K, mode=Mode.accurate):
    """Status code, like oz_aie::ozAIE_setup (0 = OK)."""
    return lib().ozAIE_setup(int(M), int(N), int(K), int(mode))


def ozAIE_setup_slices(M, N, K, bits, slices):
    return lib().ozAIE_setup_slices(int(M), int(N), int(K), int(bits), int(slices))


def ozAIE_DGEMM(M, N, K, A, B, C):
    """C = A @ B with the configuration of the last setup; A, B, C are contiguous float64
    numpy arrays (host) or torch tensors (host or device).  Returns the status code.
    Blocking, like the reference: work already queued on torch's current stream (for
    example the producer of a device operand) completes before the library starts."""
    on_device = False
    for x in (A, B, C):
        if isinstance(x, np.ndarray):
            if x.dtype != np.float64 or not x.flags["C_CONTIGUOUS"]:
                raise ValueError("numpy operands must be C-contiguous float64")
        else:
            if str(x.dtype) != "torch.float64" or not x.is_contiguous():
                raise ValueError("torch operands must be contiguous float64")
            on_device |= bool(x.is_cuda)
    if on_device:
        import torch
        torch.cuda.current_stream().synchronize()
    return lib().ozAIE_DGEMM(int(M), int(N), int(K), _ptr(A), _ptr(B), _ptr(C))


def set_option(option, value):
    return _check(lib().ozaie_set_option(int(option), int(value)), "ozaie_set_option")


def get_option(option):
    return lib().ozaie_get_option(int(option))


def plan_info():
    out = np.zeros(5, dtype=np.int32)
    _check(lib().ozaie_plan_info(_ptr(out)), "ozaie_plan_info")
    return dict(bits=int(out[0]), slices=int(out[1]), kchunk=int(out[2]), passes=int(out[3]),
                kernel=KERNEL_NAMES[int(out[4])])


def release():
    return lib().ozaie_release()


# ---------------------------------------------------------------- explicit plans (torch)
def _ld(x):
    """Leading dimension of a row-major 2-D tensor for the C ABI.  An axis of length 1 has
    no stride to speak of (numpy and torch keep any value there): a single row passes its
    own length, and a single column any row stride."""
    return int(x.stride(0)) if x.shape[0] > 1 else int(x.shape[1])


def _unit_last_axis(x):
    return (x.shape[1] <= 1 or x.stride(1) == 1) and _ld(x) >= x.shape[1]


def _stream_handle(stream):
    import torch
    if stream is None:
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


class Plan:
    """A device plan for C (M x N) = op(A) (M x K) @ op(B) (K x N) with b bits, d slices.
    crt=True: CRT mode, `slices` moduli (2..20) and `bits` per operand (0 = largest exact)."""

    def __init__(self, M, N, K, bits=7, slices=8, rounding=ROUND_TRUNCATE, crt=False):
        self.M, self.N, self.K, self.bits, self.slices = int(M), int(N), int(K), int(bits), int(slices)
        h = ctypes.c_void_p()
        flags = (1 if rounding == ROUND_NEAREST else 0) | (4 if crt else 0)
        _check(lib().ozaie_plan_create(ctypes.byref(h), self.M, self.N, self.K, self.bits,
                                       self.slices, flags),
               "ozaie_plan_create")
        self._h = h

    def info(self):
        out = np.zeros(5, dtype=np.int32)
        _check(lib().ozaie_plan_query(self._h, _ptr(out)), "ozaie_plan_query")
        return dict(bits=int(out[0]), slices=int(out[1]), kchunk=int(out[2]), passes=int(out[3]),
                    kernel=KERNEL_NAMES[int(out[4])])

    def __call__(self, A, B, C=None, transA=False, transB=False, stream=None):
        """C = op(A) @ op(B) on device float64 tensors (asynchronous on `stream`)."""
        import torch
        if C is None:
            C = torch.empty((self.M, self.N), dtype=torch.float64, device=A.device)
        for x in (A, B, C):
            if x.dtype != torch.float64 or not x.is_cuda:
                raise ValueError("operands must be float64 device tensors")
            if not _unit_last_axis(x):
                raise ValueError("operands must be row-major with unit stride along their last axis")
        ea = (self.K, self.M) if transA else (self.M, self.K)
        eb = (self.N, self.K) if transB else (self.K, self.N)
        if tuple(A.shape) != ea or tuple(B.shape) != eb or tuple(C.shape) != (self.M, self.N):
            raise ValueError(f"shape mismatch: A {tuple(A.shape)} (want {ea}), B {tuple(B.shape)} "
                             f"(want {eb}), C {tuple(C.shape)}")
        _check(lib().ozaie_dgemm(self._h, int(transA), int(transB), _ptr(A), _ld(A), _ptr(B),
                                 _ld(B), _ptr(C), _ld(C), _stream_handle(stream)),
               "ozaie_dgemm")
        return C

    def phase_times(self):
        """(split_ms, gemm_ms, total_ms) of the last call; the plan must have been created
        with OPT_PHASE_TIMING set."""
        out = np.zeros(3, dtype=np.float32)
        _check(lib().ozaie_plan_phase_times(self._h, _ptr(out)), "ozaie_plan_phase_times")
        return float(out[0]), float(out[1]), float(out[2])

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().ozaie_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def dgemm_noalloc(A, B, C, bits=7, slices=8, rounding=ROUND_TRUNCATE, crt=False, transA=False, transB=False,
                  stream=None):
    """C = op(A) @ op(B) through ozaie_dgemm_noalloc: no device allocation, intermediates
    in the library's static per-GPU workspace (NOALLOC_BYTES).  Float64 device tensors,
    row-major with unit stride along the last axis; C is written in place.  Asynchronous
    on `stream`; calls overlapping on one GPU race (they share the workspace)."""
    M, N = C.shape
    K = A.shape[0] if transA else A.shape[1]
    for x in (A, B, C):
        if str(x.dtype) != "torch.float64" or not x.is_cuda or not _unit_last_axis(x):
            raise ValueError("operands must be float64 device tensors with unit stride along their last axis")
    flags = (1 if rounding == ROUND_NEAREST else 0) | (4 if crt else 0)
    _check(lib().ozaie_dgemm_noalloc(int(M), int(N), int(K), int(bits), int(slices), flags, int(transA),
                                     int(transB), _ptr(A), _ld(A), _ptr(B), _ld(B), _ptr(C),
                                     _ld(C), _stream_handle(stream)), "ozaie_dgemm_noalloc")
    return C


def matmul(A, B, bits=7, slices=8, rounding=ROUND_TRUNCATE):
    """One-shot emulated C = A @ B on device float64 tensors (creates a plan)."""
    p = Plan(A.shape[0], B.shape[1], A.shape[1], bits, slices, rounding)
    try:
        return p(A, B)
    finally:
        import torch
        torch.cuda.current_stream().synchronize()
        p.close()


# ---------------------------------------------------------------- inspection
def debug_split(X, bits, slices, contract_axis=1):
    """(E int32[rows], S int8[d][rows][K]) of a device float64 matrix as split on the GPU.
    contract_axis=1: rows of X are operand rows (A); 0: columns of X are (B)."""
    if contract_axis == 1:
        rows, K, sr, sk = X.shape[0], X.shape[1], X.stride(0), X.stride(1)
    else:
        rows, K, sr, sk = X.shape[1], X.shape[0], X.stride(1), X.stride(0)
    E = np.zeros(rows, dtype=np.int32)
    S = np.zeros((slices, rows, K), dtype=np.int8)
    _check(lib().ozaie_debug_split(_ptr(X), rows, K, sr, sk, bits, slices, _ptr(E), _ptr(S)),
           "ozaie_debug_split")
    return E, S


def debug_convert(T_hi, T_lo, es, bits, slices, rounding=ROUND_TRUNCATE):
    """The GPU epilogue's FP64 of exact integers T = T_hi 2^64 + T_lo (int64 / uint64 host
    arrays, two's complement) scaled by 2^(es - 2044 - bits (slices + 1)): (through the
    128-bit path, through the 64-bit path where T fits int64, else 0)."""
    hi = np.ascontiguousarray(T_hi, dtype=np.int64)
    lo = np.ascontiguousarray(T_lo, dtype=np.uint64)
    e = np.ascontiguousarray(es, dtype=np.int32)
    out = np.zeros(hi.shape[0], dtype=np.float64)
    out64 = np.zeros(hi.shape[0], dtype=np.float64)
    _check(lib().ozaie_debug_convert(_ptr(hi), _ptr(lo), _ptr(e), hi.shape[0], bits, slices, rounding,
                                     _ptr(out), _ptr(out64)), "ozaie_debug_convert")
    return out, out64


def debug_crt(A, B, moduli=16, bits=0):
    """CRT-mode intermediates of A @ B (device float64, row-major): residues int8
    [n][M][K] and [n][N][K], residue products uint8 [n][M][N]."""
    M, K = A.shape
    N = B.shape[1]
    SA = np.zeros((moduli, M, K), dtype=np.int8)
    SB = np.zeros((moduli, N, K), dtype=np.int8)
    G = np.zeros((moduli, M, N), dtype=np.uint8)
    _check(lib().ozaie_debug_crt(_ptr(A), _ptr(B), M, N, K, moduli, bits, _ptr(SA), _ptr(SB), _ptr(G)),
           "ozaie_debug_crt")
    return SA, SB, G


def debug_bins(A, B, bits, slices, kernel=0):
    """int32 bins [d][M][N] of A @ B (device float64, row-major) from the slice GEMM kernels."""
    M, K = A.shape
    N = B.shape[1]
    out = np.zeros((slices, M, N), dtype=np.int32)
    _check(lib().ozaie_debug_bins(M, N, K, _ptr(A), _ptr(B), bits, slices, kernel, _ptr(out)),
           "ozaie_debug_bins")
    return out


def paired_output(A, B, bits, slices):
    """Reference wire format: int64 [M][N][2] = {L, H} per element (ozaki_base.cc:143-340)."""
    M, K = A.shape
    N = B.shape[1]
    out = np.zeros((M, N, 2), dtype=np.int64)
    _check(lib().ozaie_paired_output(M, N, K, _ptr(A), _ptr(B), bits, slices, _ptr(out)),
           "ozaie_paired_output")
    return out


def exact_entries(A, B, rows, cols):
    """Correctly rounded exact (A @ B)[rows[i], cols[i]] for host float64 arrays."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    cols = np.ascontiguousarray(cols, dtype=np.int64)
    out = np.zeros(len(rows), dtype=np.float64)
    _check(lib().ozaie_exact_entries(_ptr(A), A.shape[1], _ptr(B), B.shape[1], A.shape[1], _ptr(rows),
                                     _ptr(cols), len(rows), _ptr(out)), "ozaie_exact_entries")
    return out


class Group:
    """Many independent products C_i = A_i @ B_i (float64 device tensors, row-major) with
    b bits and d slices each (slices=0: each problem's size class's count from the
    slice-count table), replayed through one captured hipGraph.  persistent=True runs the
    whole group as one persistent kernel launch (slice counts 3, 7, 8, 9).
    crt=True runs every problem in the CRT mode with `slices` moduli and `bits` = 0 (the
    largest exact bit count per problem) or a fixed bit count; crt_cfg forces a tile shape
    (0: 256x256, 1: 128x128, 2: 64x64)."""

    GRAPH = 0
    NO_GRAPH = 2
    PERSISTENT = 4
    CRT = 8

    def __init__(self, shapes, bits=7, slices=8, rounding=ROUND_TRUNCATE, graph=True, persistent=False,
                 persistent_cfg=None, crt=False, crt_cfg=None):
        self.shapes = [(int(m), int(n), int(k)) for m, n, k in shapes]
        cnt = len(self.shapes)
        M = np.array([s[0] for s in self.shapes], dtype=np.int64)
        N = np.array([s[1] for s in self.shapes], dtype=np.int64)
        K = np.array([s[2] for s in self.shapes], dtype=np.int64)
        flags = ((1 if rounding == ROUND_NEAREST else 0) | (0 if graph else 2) | (4 if persistent else 0)
                 | (8 if crt else 0))
        if persistent_cfg is not None:
            flags |= (int(persistent_cfg) + 1) << 4
        if crt_cfg is not None:
            flags |= (int(crt_cfg) + 1) << 4
        h = ctypes.c_void_p()
        _check(lib().ozaie_group_create(ctypes.byref(h), cnt, _ptr(M), _ptr(N), _ptr(K), int(bits),
                                        int(slices), flags), "ozaie_group_create")
        self._h = h
        self.bits, self.slices = int(bits), int(slices)

    def __call__(self, As, Bs, Cs, stream=None, check=False):
        """Enqueue every product on `stream`.  check=True waits for the run and raises if
        the persistent kernel reported a barrier overrun (its C would be invalid)."""
        cnt = len(self.shapes)
        if not (len(As) == len(Bs) == len(Cs) == cnt):
            raise ValueError("need one A, B and C per problem")
        for (m, n, k), a, b, c in zip(self.shapes, As, Bs, Cs):
            if tuple(a.shape) != (m, k) or tuple(b.shape) != (k, n) or tuple(c.shape) != (m, n):
                raise ValueError("operand shapes differ from the group's problem shapes")
            for x in (a, b, c):
                if not (x.is_cuda and str(x.dtype) == "torch.float64" and x.is_contiguous()):
                    raise ValueError("operands must be contiguous float64 device tensors")
        pa = np.array([a.data_ptr() for a in As], dtype=np.uint64)
        pb = np.array([b.data_ptr() for b in Bs], dtype=np.uint64)
        pc = np.array([c.data_ptr() for c in Cs], dtype=np.uint64)
        _check(lib().ozaie_group_run(self._h, _ptr(pa), _ptr(pb), _ptr(pc), _stream_handle(stream)),
               "ozaie_group_run")
        if check and self.status():
            raise OzaieError(ERR_INCOMPLETE, "persistent grouped run (grid barrier overran its bound)")
        return Cs

    def query(self):
        out = np.zeros(4, dtype=np.int32)
        _check(lib().ozaie_group_query(self._h, _ptr(out)), "ozaie_group_query")
        return dict(launches=int(out[0]), graph=bool(out[1]), fallback=bool(out[2]), workgroups=int(out[3]))

    def status(self):
        """Wait for the last run; 1 if the persistent kernel's barrier wait overran its bound
        (that run's results are invalid), else 0.  Clears the flag."""
        out = np.zeros(1, dtype=np.int32)
        _check(lib().ozaie_group_status(self._h, _ptr(out)), "ozaie_group_status")
        return int(out[0])

    def configs(self):
        out = np.zeros(len(self.shapes), dtype=np.int32)
        _check(lib().ozaie_group_configs(self._h, _ptr(out)), "ozaie_group_configs")
        return out

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().ozaie_group_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(bits=7, slices=8, reps=3, crt=False):
    """Time the grouped tile configurations per size class; returns the 9-entry table.
    crt=True tunes the CRT groups with `slices` moduli (bits is ignored)."""
    out = np.zeros(9, dtype=np.int32)
    if crt:
        _check(lib().ozaie_autotune_crt(int(slices), int(reps), _ptr(out)), "ozaie_autotune_crt")
    else:
        _check(lib().ozaie_autotune(int(bits), int(slices), int(reps), _ptr(out)), "ozaie_autotune")
    return out


def tune_table(slices, crt=False):
    out = np.zeros(9, dtype=np.int32)
    if crt:
        _check(lib().ozaie_tune_get_crt(int(slices), _ptr(out)), "ozaie_tune_get_crt")
    else:
        _check(lib().ozaie_tune_get(int(slices), _ptr(out)), "ozaie_tune_get")
    return out


def set_tune_table(slices, table, crt=False):
    t = np.ascontiguousarray(table, dtype=np.int32)
    if crt:
        _check(lib().ozaie_tune_set_crt(int(slices), _ptr(t)), "ozaie_tune_set_crt")
    else:
        _check(lib().ozaie_tune_set(int(slices), _ptr(t)), "ozaie_tune_set")


PROBE_PATTERNS = {"zeros": 0, "constant": 1, "random": 2, "random16x16x64": 3, "symmetric127": 4}


def probe_mfma(pattern="random", iters=20000):
    """Measured int8 MFMA ceiling (POP/s, ms of the timed launch) with register-only
    operands: "zeros", "constant" (one random fragment pair), "random" (operands that
    toggle every MFMA, like slice data), "random16x16x64" (the same on the 16 x 16 x 64
    MFMA) or "symmetric127" (bytes uniform in [-127, 127], like residues and slices)."""
    pops = ctypes.c_double()
    ms = ctypes.c_double()
    _check(lib().ozaie_probe_mfma(PROBE_PATTERNS[pattern], int(iters), ctypes.byref(pops), ctypes.byref(ms)),
           "ozaie_probe_mfma")
    return pops.value, ms.value


def slice_table():
    """Slice count per size class for groups created with slices=0 (-1: untuned, 9)."""
    out = np.zeros(9, dtype=np.int32)
    _check(lib().ozaie_tune_get_slices(_ptr(out)), "ozaie_tune_get_slices")
    return out


def set_slice_table(table):
    t = np.ascontiguousarray(table, dtype=np.int32)
    _check(lib().ozaie_tune_set_slices(_ptr(t)), "ozaie_tune_set_slices")


def size_class(M, N, K):
    """The autotuner's size class of a product: 3 * bucket(max(M, N)) + bucket(K), buckets
    <= 256, <= 1024, larger."""
    bucket = lambda x: 0 if x <= 256 else (1 if x <= 1024 else 2)
    return 3 * bucket(max(M, N)) + bucket(K)


def version():
    return lib().ozaie_version()
