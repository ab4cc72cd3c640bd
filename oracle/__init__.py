"""Oracle -- TEST INFRASTRUCTURE ONLY.

CPU restatement of the reference Ozaki-scheme DGEMM path (atlarge-research/ozAIE) and of
its MPFR accuracy oracle.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this package, and only as the checker or the timed CPU
baseline; the product (ozaie_amd) never imports it.

Function map (reference file:line each restates):
  mode_table / plan_bd      include/oz_aie.h:55-82, 182-199
  groups                    ozaki_base.cc:151-161, include/oz_aie.h:96-103
  max_exp                   ozaki_base.cc:342-440
  split (closed form)       ozaki_base.cc:449-771 (closed form, valid for b*d <= 96)
  split(mech=True)          ozaki_base.cc:449-639 word-level restatement
  bins                      aie2/mm_ozaki.cc:304-357
  pack_paired               ozaki_base.cc:143-340
  convert_paired            include/oz_aie.h:92-150
  bins_to_T, T_to_double    closed form of pack_paired + convert_paired
  ozaki_dgemm               the whole pipeline, include/oz_aie.h:311-349
  mpfr_dot / mpfr_sample    common.h:164-232 (MPFR, 106 bits)
  naive_dgemm               common.h:234-254
  crt_*                     the library's CRT mode (no reference counterpart; the
                            arithmetic is stated in crt_oracle.c)

Parity status: every function above is pinned by at least one independent check in
tests/test_oracle.py (word-level vs closed-form split, numpy int64 GEMMs for bins,
mechanistic pack+convert vs exact-T closed form, exact rational arithmetic for the
MPFR oracle and the end-to-end known answers).  None is "parity unpinned".
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboz_oracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            from . import build_oracle
            build_oracle.build()
        _lib = ctypes.CDLL(LIB_PATH)
        _setup(_lib)
    return _lib


def _setup(L):
    P = ctypes.c_void_p
    i64 = ctypes.c_int64
    i32 = ctypes.c_int
    L.oro_field.argtypes = [ctypes.c_double]
    L.oro_field.restype = i32
    L.oro_max_exp.argtypes = [P, i64, i64, i64, i64, P]
    L.oro_split.argtypes = [P, i64, i64, i64, i64, P, i32, i32, P, i32]
    L.oro_bins.argtypes = [P, P, i64, i64, i64, i32, P]
    L.oro_groups.argtypes = [i32, i32, i32, ctypes.POINTER(i32), ctypes.POINTER(i32)]
    L.oro_pack_paired.argtypes = [P, i64, i64, i32, i32, i32, P, P, P]
    L.oro_convert_paired.argtypes = [P, P, i64, i64, i32, i32, i32]
    L.oro_bins_to_T.argtypes = [P, i64, i64, i32, i32, P]
    L.oro_T_to_double.argtypes = [i64, i64, i32, i32, i32, i32]
    L.oro_T_to_double.restype = ctypes.c_double
    L.oro_ozaki_dgemm.argtypes = [i64, i64, i64, P, i64, P, i64, P, i64, i32, i32, i64, i32]
    L.oro_ozaki_dgemm.restype = i32
    L.oro_naive_dgemm.argtypes = [i64, i64, i64, P, P, P]
    L.oro_num_threads.restype = i32
    # CRT mode (no reference counterpart)
    L.oro_crt_modulus.argtypes = [i32]
    L.oro_crt_modulus.restype = i32
    L.oro_crt_beta.argtypes = [i32, i64]
    L.oro_crt_beta.restype = i32
    L.oro_crt_consts.argtypes = [i32, P, P, P]
    L.oro_crt_q.argtypes = [ctypes.c_double, i32, i32]
    L.oro_crt_q.restype = i64
    L.oro_crt_split.argtypes = [P, i64, i64, i64, i64, P, i32, i32, P]
    L.oro_crt_products.argtypes = [P, P, i64, i64, i64, i32, P]
    L.oro_crt_rebuild.argtypes = [P, i64, i32, P]
    L.oro_crt_to_double.argtypes = [P, i32, i32]
    L.oro_crt_to_double.restype = ctypes.c_double
    L.oro_crt_dgemm.argtypes = [i64, i64, i64, P, i64, P, i64, P, i64, i32, i32, i32]
    L.oro_crt_dgemm.restype = i32
    L.oro_crt_q_sigma.argtypes = [ctypes.c_double, i32]
    L.oro_crt_q_sigma.restype = i64
    L.oro_crt_split_sigma.argtypes = [P, i64, i64, i64, i64, P, i32, P]
    L.oro_crt_normbound.argtypes = [P, i64, i64, i64, i64, P, P]
    L.oro_crt_norm_m.argtypes = [ctypes.c_uint64, i32, i64]
    L.oro_crt_norm_m.restype = i32
    L.oro_crt_sigma_norm.argtypes = [P, i64, i64, i64, i64, i32, i64, P]
    L.oro_crt_sigma_norm.restype = i32
    L.oro_crt_sigma_max.argtypes = [P, i64, i64, i64, i64, i32, P]
    if hasattr(L, "oro_mpfr_dot"):
        L.oro_mpfr_dot.argtypes = [P, i64, P, i64, i64, i32]
        L.oro_mpfr_dot.restype = ctypes.c_double
        L.oro_mpfr_sample.argtypes = [P, i64, P, i64, i64, P, P, i64, i32, P]


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


# ---------------------------------------------------------------- mode table
# include/oz_aie.h:60-80, indexed by log2(K) for K = 2^7 .. 2^14.
ACCURACY_CONFIGS = {7: (3, 29), 8: (3, 28), 9: (3, 27), 10: (3, 27), 11: (3, 26),
                    12: (3, 25), 13: (3, 25), 14: (4, 18)}
PERF_CONFIGS = {k: (7, 8) for k in range(7, 15)}
ACCURATE, FAST = 0, 1


def plan_bd(M, N, K, mode):
    """ozAIE_setup's validation and (b, d) lookup, include/oz_aie.h:184-199.
    Returns (status, b, d); status 1 = rejected dims."""
    import math
    log2k = int(math.log2(K)) if K > 0 else -1
    if log2k < 7 or log2k > 11 or M != K or N != K:
        return 1, None, None
    table = ACCURACY_CONFIGS if mode == ACCURATE else PERF_CONFIGS
    b, d = table[log2k]
    return 0, b, d


def groups(b, d, K):
    n0 = ctypes.c_int()
    n1 = ctypes.c_int()
    lib().oro_groups(b, d, K, ctypes.byref(n0), ctypes.byref(n1))
    return n0.value, n1.value


# ---------------------------------------------------------------- pipeline pieces
def field(x):
    return lib().oro_field(float(x))


def max_exp(X, axis_contract=1):
    """Row maxima of the exponent field of a 2-D array, contracting over `axis_contract`
    (1: rows of A; 0: columns of B)."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    if axis_contract == 1:
        rows, cols, sr, sc = X.shape[0], X.shape[1], X.shape[1], 1
    else:
        rows, cols, sr, sc = X.shape[1], X.shape[0], 1, X.shape[1]
    E = np.zeros(rows, dtype=np.int16)
    lib().oro_max_exp(_p(X), rows, cols, sr, sc, _p(E))
    return E


def split(X, E, b, d, axis_contract=1, mech=False):
    """Slices S[j][row][k] (int8) of A (axis_contract=1) or of B viewed as N x K
    (axis_contract=0)."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    if axis_contract == 1:
        rows, cols, sr, sc = X.shape[0], X.shape[1], X.shape[1], 1
    else:
        rows, cols, sr, sc = X.shape[1], X.shape[0], 1, X.shape[1]
    E = np.ascontiguousarray(E, dtype=np.int16)
    S = np.zeros((d, rows, cols), dtype=np.int8)
    lib().oro_split(_p(X), rows, cols, sr, sc, _p(E), b, d, _p(S), 1 if mech else 0)
    return S


def bins(SA, SB):
    """bins[k][m][n] = sum_{i+j=k} SA[i] @ SB[j].T  (int64, exact)."""
    SA = np.ascontiguousarray(SA, dtype=np.int8)
    SB = np.ascontiguousarray(SB, dtype=np.int8)
    d, M, K = SA.shape
    N = SB.shape[1]
    out = np.zeros((d, M, N), dtype=np.int64)
    lib().oro_bins(_p(SA), _p(SB), M, N, K, d, _p(out))
    return out


def pack_paired(B_, K, b, d, Er, Ec):
    B_ = np.ascontiguousarray(B_, dtype=np.int64)
    _, M, N = B_.shape
    out = np.zeros((M, N, 2), dtype=np.int64)
    lib().oro_pack_paired(_p(B_), M, N, K, b, d, _p(np.ascontiguousarray(Er, np.int16)),
                          _p(np.ascontiguousarray(Ec, np.int16)), _p(out))
    return out


def convert_paired(P, K, b, d):
    P = np.ascontiguousarray(P, dtype=np.int64)
    M, N = P.shape[0], P.shape[1]
    C = np.zeros((M, N), dtype=np.float64)
    lib().oro_convert_paired(_p(P), _p(C), M, N, K, b, d)
    return C


def bins_to_T(B_, b, d):
    B_ = np.ascontiguousarray(B_, dtype=np.int64)
    _, M, N = B_.shape
    T = np.zeros((M, N, 2), dtype=np.int64)
    lib().oro_bins_to_T(_p(B_), M, N, b, d, _p(T))
    return T


def T_to_double(Thi, Tlo, es, b, d, rounding=0):
    return lib().oro_T_to_double(int(Thi), int(Tlo), int(es), b, d, rounding)


def ozaki_dgemm(A, B, b, d, kchunk=0, rounding=0):
    """End-to-end CPU Ozaki DGEMM; returns (C, overflow_flag)."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    M, K = A.shape
    N = B.shape[1]
    C = np.zeros((M, N), dtype=np.float64)
    ov = lib().oro_ozaki_dgemm(M, N, K, _p(A), K, _p(B), N, _p(C), N, b, d, kchunk, rounding)
    return C, ov


def naive_dgemm(A, B):
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    M, K = A.shape
    N = B.shape[1]
    C = np.zeros((M, N), dtype=np.float64)
    lib().oro_naive_dgemm(M, N, K, _p(A), _p(B), _p(C))
    return C


def have_mpfr():
    return hasattr(lib(), "oro_mpfr_dot")


def mpfr_sample(A, B, rows, cols, prec=106):
    """MPFR-`prec` dot products for the (rows[i], cols[i]) entries of A @ B."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    cols = np.ascontiguousarray(cols, dtype=np.int64)
    out = np.zeros(len(rows), dtype=np.float64)
    lib().oro_mpfr_sample(_p(A), A.shape[1], _p(B), B.shape[1], A.shape[1], _p(rows), _p(cols),
                          len(rows), prec, _p(out))
    return out


def num_threads():
    return lib().oro_num_threads()


# ---------------------------------------------------------------- CRT mode
# Not in the reference: exact integer product of per-row / per-column scaled operands,
# computed from int8 residue GEMMs modulo 16 pairwise-coprime moduli and rebuilt with
# the Chinese remainder theorem (crt_oracle.c states the arithmetic).
def crt_moduli(n=16):
    return [lib().oro_crt_modulus(i) for i in range(n)]


def crt_beta(n, K):
    return lib().oro_crt_beta(int(n), int(K))


def crt_consts(n):
    """(W_p as Python ints, fl(W_p / M) as float64, M)."""
    W = np.zeros(3 * n, dtype=np.uint64)
    f = np.zeros(n, dtype=np.float64)
    Mw = np.zeros(3, dtype=np.uint64)
    lib().oro_crt_consts(int(n), _p(W), _p(f), _p(Mw))
    word3 = lambda w: int(w[0]) | (int(w[1]) << 64) | (int(w[2]) << 128)
    return [word3(W[3 * i:3 * i + 3]) for i in range(n)], f, word3(Mw)


def crt_q(x, emax, beta):
    return lib().oro_crt_q(float(x), int(emax), int(beta))


def crt_split(X, E, beta, n, axis_contract=1):
    """Residues int8[n][rows][K] of X (rows along the other axis), E per row."""
    X = np.asarray(X, dtype=np.float64)
    if axis_contract == 1:
        rows, K = X.shape
        sr, sc = X.strides[0] // 8, X.strides[1] // 8
    else:
        K, rows = X.shape
        sr, sc = X.strides[1] // 8, X.strides[0] // 8
    E = np.ascontiguousarray(E, dtype=np.int16)
    S = np.zeros((n, rows, K), dtype=np.int8)
    lib().oro_crt_split(_p(X), rows, K, sr, sc, _p(E), int(beta), int(n), _p(S))
    return S


def crt_products(SA, SB):
    """g[i][r][c] = (SA[i][r] . SB[i][c]) mod p_i in [0, p_i)."""
    SA = np.ascontiguousarray(SA, dtype=np.int8)
    SB = np.ascontiguousarray(SB, dtype=np.int8)
    n, M, K = SA.shape
    N = SB.shape[1]
    G = np.zeros((n, M, N), dtype=np.uint8)
    lib().oro_crt_products(_p(SA), _p(SB), M, N, K, n, _p(G))
    return G


def crt_rebuild(g):
    """Signed integer from the residues g[0..n) (Python int)."""
    g = np.ascontiguousarray(g, dtype=np.uint8)
    C3 = np.zeros(3, dtype=np.int64)
    lib().oro_crt_rebuild(_p(g), 1, len(g), _p(C3))
    v = sum((int(w) & (2**64 - 1)) << (64 * i) for i, w in enumerate(C3))
    return v - 2**192 if v >= 2**191 else v


def crt_to_double(v, e2, rounding=1):
    v &= 2**192 - 1
    C3 = np.array([((v >> (64 * i)) & (2**64 - 1)) for i in range(3)], dtype=np.uint64).view(np.int64)
    return lib().oro_crt_to_double(_p(C3), int(e2), int(rounding))


def crt_dgemm(A, B, n=16, rounding=1, bits=0):
    """End-to-end CPU CRT-mode DGEMM (bits 0: the norm rule, else the max rule)."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    M, K = A.shape
    N = B.shape[1]
    C = np.zeros((M, N), dtype=np.float64)
    rc = lib().oro_crt_dgemm(M, N, K, _p(A), K, _p(B), N, _p(C), N, int(n), int(bits), int(rounding))
    if rc:
        raise ValueError("no bit count fits these moduli and K")
    return C


def _rows_view(X, axis_contract):
    X = np.asarray(X, dtype=np.float64)
    if axis_contract == 1:
        rows, K = X.shape
        return X, rows, K, X.strides[0] // 8, X.strides[1] // 8
    K, rows = X.shape
    return X, rows, K, X.strides[1] // 8, X.strides[0] // 8


def crt_q_sigma(x, sigma):
    return lib().oro_crt_q_sigma(float(x), int(sigma))


def crt_normbound(X, axis_contract=1):
    """(E^ int32[rows], S uint64[rows]) of the norm rule: ||x||^2 <= S 4^(E^ - 1038)."""
    X, rows, K, sr, sc = _rows_view(X, axis_contract)
    Eh = np.zeros(rows, dtype=np.int32)
    S = np.zeros(rows, dtype=np.uint64)
    lib().oro_crt_normbound(_p(X), rows, K, sr, sc, _p(Eh), _p(S))
    return Eh, S


def crt_norm_m(S, n, K):
    return lib().oro_crt_norm_m(int(S), int(n), int(K))


def crt_sigma(X, n, axis_contract=1, bits=0, K=None):
    """Per-row shifts sigma: the norm rule (bits 0) or the max rule (beta = bits)."""
    X, rows, Kx, sr, sc = _rows_view(X, axis_contract)
    sg = np.zeros(rows, dtype=np.int32)
    if bits:
        lib().oro_crt_sigma_max(_p(X), rows, Kx, sr, sc, int(bits), _p(sg))
    elif lib().oro_crt_sigma_norm(_p(X), rows, Kx, sr, sc, int(n), int(K or Kx), _p(sg)):
        raise ValueError("no bit count fits these moduli and K")
    return sg


def crt_split_sigma(X, sigma, n, axis_contract=1):
    """Residues int8[n][rows][K] of X for per-row shifts sigma."""
    X, rows, K, sr, sc = _rows_view(X, axis_contract)
    sigma = np.ascontiguousarray(sigma, dtype=np.int32)
    S = np.zeros((n, rows, K), dtype=np.int8)
    lib().oro_crt_split_sigma(_p(X), rows, K, sr, sc, _p(sigma), int(n), _p(S))
    return S
