"""Parity of the HIP path (through the C ABI) with the oracle on the same inputs.

Bar: bit-exact exponents, slices, int32 bins, paired-int64 wire format and FP64 C
(truncating conversion, the reference rule); the nearest-even conversion is bit-exact
against the oracle's nearest-even closed form and within 1 ulp of the truncated one.
"""
import numpy as np
import pytest

import oracle
import ozaie_amd as oz
from conftest import gen_inputs

pytestmark = pytest.mark.gpu


def dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda")


def edge_matrix(rng, m, n, phi=2.0):
    X = gen_inputs(rng, m, n, phi)
    X[0, :6] = [0.0, -0.0, 1.0, 5e-324, 2.2250738585072014e-308, -2.2250738585072009e-308]
    X[1, :] = 0.0
    X[2, 3] = 1e-300
    X[3, :] = -1.0
    X[5, 1] = np.nextafter(1.0, 2.0)
    return X


def bits(x):
    return np.ascontiguousarray(x).view(np.int64)


SPLIT_CASES = [(3, 29), (3, 26), (3, 25), (4, 18), (7, 8), (7, 3), (7, 7), (7, 9), (7, 13), (1, 32)]


@pytest.mark.parametrize("b,d", SPLIT_CASES)
def test_split_bit_exact(gpu, b, d):
    rng = np.random.default_rng(b * 1000 + d)
    A = edge_matrix(rng, 70, 300)
    E, S = oz.debug_split(dev(A), b, d, 1)
    Eo = oracle.max_exp(A, 1)
    assert np.array_equal(E, Eo.astype(np.int32))
    assert np.array_equal(S, oracle.split(A, Eo, b, d, 1))
    B = edge_matrix(rng, 260, 90).T.copy().T   # K x N, row-major
    B = np.ascontiguousarray(B)
    E, S = oz.debug_split(dev(B), b, d, 0)
    Eo = oracle.max_exp(B, 0)
    assert np.array_equal(E, Eo.astype(np.int32))
    assert np.array_equal(S, oracle.split(B, Eo, b, d, 0))


# kernel: 0 auto, 1 bin-group, 2 fused LDS copies, 3 fused direct loads, 4 fused pipelined,
# 5 fused staged (d <= 4 and d = 9; other d run the LDS-copy kernel), 6 fused LDS copies
# with rolling fragment reads
BIN_CASES = [(7, 8, 0), (7, 3, 0), (7, 7, 0), (7, 9, 0), (7, 1, 0), (7, 4, 0), (7, 10, 0), (6, 16, 0),
             (3, 29, 1), (3, 26, 1), (7, 8, 1), (4, 18, 1),
             (7, 8, 2), (7, 3, 2), (7, 9, 2), (7, 8, 3), (7, 3, 3), (7, 5, 3), (7, 8, 4), (7, 3, 4),
             (7, 7, 4), (6, 12, 4), (7, 1, 5), (7, 2, 5), (7, 3, 5), (7, 4, 5), (7, 8, 5),
             (7, 9, 3), (7, 9, 5), (7, 9, 6), (7, 8, 6), (7, 3, 6), (6, 12, 6), (4, 18, 0)]


@pytest.mark.parametrize("b,d,kernel", BIN_CASES)
def test_bins_bit_exact(gpu, b, d, kernel):
    rng = np.random.default_rng(d * 31 + kernel)
    M, N, K = 100, 150, 300
    A = gen_inputs(rng, M, K, 1.0)
    B = gen_inputs(rng, K, N, 1.0)
    got = oz.debug_bins(dev(A), dev(B), b, d, kernel)
    Er, Ec = oracle.max_exp(A, 1), oracle.max_exp(B, 0)
    ref = oracle.bins(oracle.split(A, Er, b, d, 1), oracle.split(B, Ec, b, d, 0))
    assert np.abs(ref).max() < 2 ** 31
    assert np.array_equal(got.astype(np.int64), ref)


def special_matrix(rng, m, n):
    X = gen_inputs(rng, m, n, 1.0)
    X[0, 3], X[1, 5], X[2, 7], X[3, 1] = np.inf, -np.inf, np.nan, np.finfo(np.float64).max
    X[4, :] = np.inf
    X[6, 2] = -np.nan
    return X


@pytest.mark.parametrize("b,d", [(7, 8), (7, 9), (3, 29)])
def test_special_values_follow_the_reference_rule(gpu, b, d):
    """Infinities and NaNs carry exponent field 2047 and go through the reference's split
    and conversion like any other value (the reference has no special-value handling): the
    GPU's exponents, slices and C equal the oracle's, bit for bit."""
    import torch
    rng = np.random.default_rng(40 + d)
    A = special_matrix(rng, 70, 160)
    B = np.ascontiguousarray(special_matrix(rng, 90, 160).T)   # K x N with specials in columns
    E, S = oz.debug_split(dev(A), b, d, 1)
    Eo = oracle.max_exp(A, 1)
    assert np.array_equal(E, Eo.astype(np.int32))
    assert np.array_equal(S, oracle.split(A, Eo, b, d, 1))
    E, S = oz.debug_split(dev(B), b, d, 0)
    Eo = oracle.max_exp(B, 0)
    assert np.array_equal(E, Eo.astype(np.int32))
    assert np.array_equal(S, oracle.split(B, Eo, b, d, 0))
    p = oz.Plan(70, 90, 160, b, d)
    C = p(dev(A), dev(B))
    torch.cuda.synchronize()
    ref, ov = oracle.ozaki_dgemm(A, B, b, d)
    assert ov == 0
    assert np.array_equal(bits(C.cpu().numpy()), bits(ref))
    p.close()


C_CASES = [(7, 8), (7, 3), (7, 7), (7, 9), (3, 29), (3, 26), (4, 18), (7, 12), (5, 19), (7, 1), (7, 2), (6, 4)]


@pytest.mark.parametrize("b,d", C_CASES)
def test_C_bit_exact_truncating(gpu, b, d):
    import torch
    rng = np.random.default_rng(d * 7 + b)
    M, N, K = 130, 70, 257
    A = gen_inputs(rng, M, K, 1.0)
    B = gen_inputs(rng, K, N, 1.0)
    plan = oz.Plan(M, N, K, b, d)
    C = plan(dev(A), dev(B))
    torch.cuda.synchronize()
    ref, ov = oracle.ozaki_dgemm(A, B, b, d)
    assert ov == 0
    assert np.array_equal(bits(C.cpu().numpy()), bits(ref))


# the default fused-kernel family per slice count (OZAIE_OPT_KERNEL values)
DEFAULT_FAMILY = {1: 5, 2: 5, 3: 5, 4: 3, 5: 3, 6: 3, 7: 6, 8: 3}


@pytest.mark.parametrize("d", [1, 2, 3, 4, 5, 6, 7, 8])
def test_small_grid_tiles_match_default_tiles(gpu, d):
    """A d <= 8 product whose default grid has fewer workgroups than CUs runs on 32 x 32
    wave tiles; naming the same family through the option keeps the default tile.  Both give
    the same C, bit for bit, equal to the oracle's, with ragged edges in both dimensions."""
    import torch
    rng = np.random.default_rng(40 + d)
    M, N, K = 450, 390, 300
    A = gen_inputs(rng, M, K, 1.0)
    B = gen_inputs(rng, K, N, 1.0)
    print(f"small-grid tiles b=7 d={d} {M}x{N}x{K}", flush=True)
    out = []
    for fam in (0, DEFAULT_FAMILY[d]):
        oz.set_option(oz.OPT_KERNEL, fam)
        try:
            plan = oz.Plan(M, N, K, 7, d)
        finally:
            oz.set_option(oz.OPT_KERNEL, 0)
        C = plan(dev(A), dev(B))
        torch.cuda.synchronize()
        plan.close()
        out.append(C.cpu().numpy())
    ref, _ = oracle.ozaki_dgemm(A, B, 7, d)
    assert np.array_equal(bits(out[0]), bits(ref))
    assert np.array_equal(bits(out[1]), bits(ref))


@pytest.mark.parametrize("K", [1, 64, 100, 128])
@pytest.mark.parametrize("b,d", [(7, 8), (7, 3), (3, 29)])
def test_short_K_split_forms_exponents(gpu, K, b, d):
    """K <= 128: each split workgroup holds a row's whole contraction axis and forms the
    row and column exponent maxima itself, with no exponent pass.  C equals the oracle's, bit
    for bit, for plain and transposed operands, with zero rows, subnormals and mixed
    magnitudes."""
    import torch
    rng = np.random.default_rng(K * 13 + d)
    M, N = 70, 90
    if K >= 6:
        A = edge_matrix(rng, M, K)
        B = np.ascontiguousarray(edge_matrix(rng, N, K).T)
    else:
        A = gen_inputs(rng, M, K, 2.0)
        B = gen_inputs(rng, K, N, 2.0)
        A[1, :] = 0.0
        B[:, 2] = 5e-324
    print(f"short K b={b} d={d} {M}x{N}x{K}", flush=True)
    ref, _ = oracle.ozaki_dgemm(A, B, b, d)
    for ta, tb in ((False, False), (True, True)):
        plan = oz.Plan(M, N, K, b, d)
        C = plan(dev(A.T) if ta else dev(A), dev(B.T) if tb else dev(B), transA=ta, transB=tb)
        torch.cuda.synchronize()
        plan.close()
        assert np.array_equal(bits(C.cpu().numpy()), bits(ref)), (ta, tb)


def test_C_nearest_even(gpu):
    import torch
    rng = np.random.default_rng(3)
    M, N, K = 64, 96, 200
    A = gen_inputs(rng, M, K, 1.0)
    B = gen_inputs(rng, K, N, 1.0)
    for b, d in [(7, 8), (3, 29)]:
        C = oz.Plan(M, N, K, b, d, rounding=oz.ROUND_NEAREST)(dev(A), dev(B))
        torch.cuda.synchronize()
        ref, _ = oracle.ozaki_dgemm(A, B, b, d, 0, 1)
        assert np.array_equal(bits(C.cpu().numpy()), bits(ref))
        tr, _ = oracle.ozaki_dgemm(A, B, b, d, 0, 0)
        assert np.max(np.abs(bits(C.cpu().numpy()) - bits(tr))) <= 1


def test_example_ones_via_reference_api(gpu):
    """example/oztest.cpp of the reference: accurate mode, 256 x 512 x 128, ones -> 128."""
    M, N, K = 256, 512, 128
    assert oz.ozAIE_setup(M, N, K, oz.Mode.accurate) == oz.OK
    assert oz.plan_info()["slices"] == 29
    A = np.ones((M, K))
    B = np.ones((K, N))
    C = np.zeros((M, N))
    assert oz.ozAIE_DGEMM(M, N, K, A, B, C) == oz.OK
    assert np.all(C == 128.0)


def test_config1_reference_api_vs_oracle_and_mpfr(gpu):
    """256 x 512 x 128, phi = 0.5, accurate mode, host pointers; C bit-exact against the
    oracle; max relative error against MPFR-106 printed and bounded."""
    rng = np.random.default_rng(12445)
    M, N, K = 256, 512, 128
    A = gen_inputs(rng, M, K, 0.5)
    B = gen_inputs(rng, K, N, 0.5)
    assert oz.ozAIE_setup(M, N, K, oz.Mode.accurate) == oz.OK
    C = np.zeros((M, N))
    assert oz.ozAIE_DGEMM(M, N, K, A, B, C) == oz.OK
    ref, _ = oracle.ozaki_dgemm(A, B, 3, 29)
    assert np.array_equal(bits(C), bits(ref))
    assert oz.ozAIE_DGEMM(M, N + 1, K, A, B, C) == oz.ERR_MISMATCH
    if oracle.have_mpfr():
        rows = np.repeat(np.arange(0, M, 8), N // 16)
        cols = np.tile(np.arange(0, N, 16), M // 8)
        mp = oracle.mpfr_sample(A, B, rows, cols)
        rel = np.abs(C[rows, cols] - mp) / np.abs(mp)
        print("config1 accurate max rel err vs MPFR:", rel.max())
        assert rel.max() < 1e-10


def test_reference_api_device_pointers(gpu):
    import torch
    rng = np.random.default_rng(8)
    M, N, K = 96, 64, 512
    A = gen_inputs(rng, M, K, 1.0)
    B = gen_inputs(rng, K, N, 1.0)
    assert oz.ozAIE_setup(M, N, K, oz.Mode.fast) == oz.OK
    dA, dB = dev(A), dev(B)
    dC = torch.zeros((M, N), dtype=torch.float64, device="cuda")
    assert oz.ozAIE_DGEMM(M, N, K, dA, dB, dC) == oz.OK
    ref, _ = oracle.ozaki_dgemm(A, B, 7, 8)
    assert np.array_equal(bits(dC.cpu().numpy()), bits(ref))


@pytest.mark.parametrize("mode", ["fast", "crt"])
def test_reference_api_host_arrays_across_staging_chunks(gpu, mode):
    """Host arrays larger than the 32 MiB pinned staging chunks (A and B two chunks each with
    a partial last one, C three), ragged sizes: C is bit-identical to the same plan's result
    from device pointers, and sampled rows equal the oracle's (exponents, shifts and slices
    of a row of A or a column of B depend on that row or column only)."""
    import torch
    rng = np.random.default_rng(31)
    M, N, K = 3001, 2999, 2500
    A = gen_inputs(rng, M, K, 1.0)
    B = gen_inputs(rng, K, N, 1.0)
    m = oz.Mode.fast if mode == "fast" else oz.Mode.crt
    assert oz.ozAIE_setup(M, N, K, m) == oz.OK
    C = np.empty((M, N), dtype=np.float64)
    assert oz.ozAIE_DGEMM(M, N, K, A, B, C) == oz.OK
    dC = torch.empty((M, N), dtype=torch.float64, device="cuda")
    assert oz.ozAIE_DGEMM(M, N, K, dev(A), dev(B), dC) == oz.OK
    assert np.array_equal(bits(C), bits(dC.cpu().numpy()))
    rows = np.array([0, 1, 1499, 2000, M - 1])
    cols = np.arange(0, N, 97)
    Ar, Bc = A[rows], np.ascontiguousarray(B[:, cols])
    if mode == "fast":
        ref, ov = oracle.ozaki_dgemm(Ar, Bc, 7, 8)
        assert ov == 0
    else:
        ref = oracle.crt_dgemm(Ar, Bc, oz.plan_info()["slices"], oz.get_option(oz.OPT_ROUNDING))
    assert np.array_equal(bits(C[np.ix_(rows, cols)]), bits(ref))
    oz.release()


def test_multi_pass_kchunk_and_groups(gpu):
    """Forcing several contraction chunks (and bin groups) gives the single-pass result."""
    import torch
    rng = np.random.default_rng(21)
    M, N, K = 70, 90, 1000
    A = gen_inputs(rng, M, K, 1.0)
    B = gen_inputs(rng, K, N, 1.0)
    for b, d, kernel in [(7, 8, 0), (7, 3, 0), (7, 5, 0), (3, 29, 1), (7, 8, 1), (7, 3, 5), (7, 9, 3), (7, 9, 5),
                         (7, 9, 6), (4, 18, 0)]:
        oz.set_option(oz.OPT_KCHUNK, 128)
        oz.set_option(oz.OPT_KERNEL, kernel)
        try:
            assert oz.ozAIE_setup_slices(M, N, K, b, d) == oz.OK
            info = oz.plan_info()
            assert info["passes"] > 1
            C = torch.zeros((M, N), dtype=torch.float64, device="cuda")
            assert oz.ozAIE_DGEMM(M, N, K, dev(A), dev(B), C) == oz.OK
        finally:
            oz.set_option(oz.OPT_KCHUNK, 0)
            oz.set_option(oz.OPT_KERNEL, 0)
        ref, _ = oracle.ozaki_dgemm(A, B, b, d)
        assert np.array_equal(bits(C.cpu().numpy()), bits(ref))


def test_window_kernel_matches_bin_groups(gpu):
    """The reference's accurate preset for K >= 2^14 (b = 4, d = 18: 4-bit digits do not pair)
    runs as two bin windows of the rolling kernel; C equals the three-pass bin-group kernel's
    and the oracle's, bit for bit, in both roundings."""
    import torch
    rng = np.random.default_rng(26)
    M, N, K = 130, 140, 2000
    A = gen_inputs(rng, M, K, 1.0)
    B = gen_inputs(rng, K, N, 1.0)
    for rounding in (oz.ROUND_TRUNCATE, oz.ROUND_NEAREST):
        p = oz.Plan(M, N, K, 4, 18, rounding=rounding)
        info = p.info()
        assert info["kernel"] == "window" and info["passes"] == 2
        C = p(dev(A), dev(B)).cpu().numpy()
        p.close()
        oz.set_option(oz.OPT_KERNEL, 1)
        try:
            q = oz.Plan(M, N, K, 4, 18, rounding=rounding)
        finally:
            oz.set_option(oz.OPT_KERNEL, 0)
        assert q.info()["kernel"] == "group" and q.info()["passes"] == 3
        G = q(dev(A), dev(B)).cpu().numpy()
        q.close()
        assert np.array_equal(bits(C), bits(G))
        if rounding == oz.ROUND_TRUNCATE:
            ref, ov = oracle.ozaki_dgemm(A, B, 4, 18)
            assert ov == 0
            assert np.array_equal(bits(C), bits(ref))
    assert oz.ozAIE_setup(64, 64, 1 << 14, oz.Mode.accurate) == oz.OK
    assert oz.plan_info()["kernel"] == "window"
    oz.release()


def test_long_K_needs_chunks(gpu):
    """d = 9 bins overflow int32 beyond K = 14784: the planner chunks K automatically."""
    import torch
    rng = np.random.default_rng(22)
    M, N, K = 40, 36, 17000
    A = gen_inputs(rng, M, K, 2.0)
    B = gen_inputs(rng, K, N, 2.0)
    p = oz.Plan(M, N, K, 7, 9)
    assert p.info()["passes"] == 2
    C = p(dev(A), dev(B))
    torch.cuda.synchronize()
    ref, ov = oracle.ozaki_dgemm(A, B, 7, 9, kchunk=14784)
    assert ov == 0
    assert np.array_equal(bits(C.cpu().numpy()), bits(ref))


@pytest.mark.parametrize("d", [8, 9])
def test_very_long_K_many_passes(gpu, d):
    """K = 100000: six (d = 8) or seven (d = 9) contraction passes accumulated into the exact
    128-bit T, over a wide exponent range; C equals the oracle's at the planner's chunk."""
    import torch
    rng = np.random.default_rng(30 + d)
    M, N, K = 12, 10, 100000
    A = gen_inputs(rng, M, K, 2.0)
    B = gen_inputs(rng, K, N, 2.0)
    p = oz.Plan(M, N, K, 7, d)
    info = p.info()
    assert info["passes"] >= 6
    C = p(dev(A), dev(B))
    torch.cuda.synchronize()
    ref, ov = oracle.ozaki_dgemm(A, B, 7, d, kchunk=info["kchunk"])
    assert ov == 0
    assert np.array_equal(bits(C.cpu().numpy()), bits(ref))
    p.close()


def test_transposed_operands(gpu):
    import torch
    rng = np.random.default_rng(23)
    M, N, K = 50, 60, 200
    A = gen_inputs(rng, M, K, 1.0)
    B = gen_inputs(rng, K, N, 1.0)
    ref, _ = oracle.ozaki_dgemm(A, B, 7, 8)
    p = oz.Plan(M, N, K, 7, 8)
    for ta in (False, True):
        for tb in (False, True):
            a = dev(A.T.copy()) if ta else dev(A)
            b = dev(B.T.copy()) if tb else dev(B)
            C = p(a, b, transA=ta, transB=tb)
            torch.cuda.synchronize()
            assert np.array_equal(bits(C.cpu().numpy()), bits(ref)), (ta, tb)


@pytest.mark.parametrize("scheme", ["split", "crt"])
def test_padded_leading_dimensions(gpu, scheme):
    """A, B and C as views into wider buffers (leading dimensions above their row lengths,
    not multiples of 16 bytes), plain and transposed: the same C as from contiguous operands,
    equal to the oracle, and nothing outside C's view written."""
    import torch
    rng = np.random.default_rng(25)
    M, N, K = 70, 90, 333
    A = gen_inputs(rng, M, K, 1.0)
    B = gen_inputs(rng, K, N, 1.0)
    if scheme == "split":
        p = oz.Plan(M, N, K, 7, 9)
        ref, _ = oracle.ozaki_dgemm(A, B, 7, 9)
    else:
        p = oz.Plan(M, N, K, 0, 15, rounding=oz.ROUND_NEAREST, crt=True)
        ref = oracle.crt_dgemm(A, B, 15, oz.ROUND_NEAREST)

    def padded(X, extra):
        buf = torch.full((X.shape[0], X.shape[1] + extra), float("nan"), dtype=torch.float64, device="cuda")
        buf[:, :X.shape[1]] = torch.from_numpy(np.ascontiguousarray(X)).to("cuda")
        return buf[:, :X.shape[1]]

    for ta in (False, True):
        for tb in (False, True):
            a = padded(A.T, 3) if ta else padded(A, 5)
            b = padded(B.T, 7) if tb else padded(B, 1)
            cbuf = torch.full((M, N + 9), -7.0, dtype=torch.float64, device="cuda")
            C = cbuf[:, :N]
            assert a.stride(0) != a.shape[1] and C.stride(0) == N + 9
            p(a, b, C, transA=ta, transB=tb)
            torch.cuda.synchronize()
            assert np.array_equal(bits(C.cpu().numpy()), bits(ref)), (ta, tb)
            assert bool((cbuf[:, N:] == -7.0).all()), (ta, tb)
    p.close()


def test_paired_wire_format(gpu):
    rng = np.random.default_rng(24)
    for K, b, d in [(128, 3, 29), (256, 7, 8), (2048, 3, 26)]:
        M, N = 40, 70
        A = gen_inputs(rng, M, K, 1.0)
        B = gen_inputs(rng, K, N, 1.0)
        got = oz.paired_output(dev(A), dev(B), b, d)
        Er, Ec = oracle.max_exp(A, 1), oracle.max_exp(B, 0)
        bins = oracle.bins(oracle.split(A, Er, b, d, 1), oracle.split(B, Ec, b, d, 0))
        ref = oracle.pack_paired(bins, K, b, d, Er, Ec)
        assert np.array_equal(got, ref)
        # and the reference host conversion of it gives the library's C
        C = oracle.convert_paired(got, K, b, d)
        C2, _ = oracle.ozaki_dgemm(A, B, b, d)
        assert np.array_equal(bits(C), bits(C2))


def test_small_integers_exact_gpu(gpu):
    import torch
    rng = np.random.default_rng(25)
    A = rng.integers(-127, 128, size=(33, 70)).astype(np.float64)
    B = rng.integers(-127, 128, size=(70, 45)).astype(np.float64)
    for b, d in [(7, 3), (7, 8), (3, 29)]:
        C = oz.Plan(33, 45, 70, b, d)(dev(A), dev(B))
        torch.cuda.synchronize()
        assert np.array_equal(C.cpu().numpy(), A @ B)
