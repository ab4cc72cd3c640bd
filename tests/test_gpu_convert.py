"""The epilogue's conversion of the exact integer T to FP64 (include/oz_aie.h:92-150 of the
reference for truncation; nearest-even as the option), computed on the GPU by both of its
forms -- the 128-bit one and the 64-bit one used when T fits int64 -- bit for bit against
the oracle's conversion (oro_T_to_double), on values chosen to hit every branch: every bit
length 1..126, both signs, ties and near-ties at the 53-bit rounding position with odd and
even heads, the carry of 2^53 - 1 into 2^53, bits below the 64-bit head (sticky), zero."""
import numpy as np
import pytest

import oracle
import ozaie_amd as oz

pytestmark = pytest.mark.gpu


def cases(rng):
    vals = [0, 1, -1, (1 << 53) - 1, 1 << 53, (1 << 53) + 1, -(1 << 53) - 1, (1 << 62) - 1, -(1 << 62)]
    for nb in range(1, 127):
        for _ in range(6):
            v = (1 << (nb - 1)) | int(rng.integers(0, 1 << 62)) << max(0, nb - 63) & ((1 << (nb - 1)) - 1)
            v |= int(rng.integers(0, 1 << 62)) & ((1 << (nb - 1)) - 1)
            vals += [v, -v]
        if nb > 53:
            sh = nb - 53
            for q in ((1 << 52) | 1, (1 << 52) | 2, (1 << 53) - 1, (1 << 53) - 2):
                base = q << sh
                half = 1 << (sh - 1)
                for extra in (half, half - 1, half + 1, 0, (1 << sh) - 1):
                    if 0 <= extra < (1 << sh):
                        vals += [base + extra, -(base + extra)]
            if nb > 64:   # ties decided only by bits below the 64-bit head
                q = (1 << 52) | 1
                vals += [(q << sh) | (1 << (sh - 1)) | 1, (q << sh) | (1 << (sh - 1))]
    return vals


@pytest.mark.parametrize("b,d", [(7, 3), (7, 8), (7, 9), (3, 29), (4, 18)])
@pytest.mark.parametrize("rounding", [0, 1])
def test_gpu_conversion_matches_oracle(gpu, b, d, rounding):
    rng = np.random.default_rng(100 * b + d + rounding)
    vals = cases(rng)
    n = len(vals)
    hi = np.array([(v >> 64) for v in vals], dtype=np.int64)
    lo = np.array([v & ((1 << 64) - 1) for v in vals], dtype=np.uint64)
    # exponent sums that keep the results normal doubles: the value is about
    # 2^(nb + es - 2045 - b (d + 1))
    nbs = np.array([max(1, abs(v).bit_length()) for v in vals])
    es = (2045 + b * (d + 1) - nbs + rng.integers(-600, 600, n)).astype(np.int32)
    out, out64 = oz.debug_convert(hi, lo, es, b, d, rounding)
    ref = np.array([oracle.T_to_double(int(h), int(l.astype(np.int64)), int(e), b, d, rounding)
                    for h, l, e in zip(hi, lo, es)])
    bad = np.nonzero(out.view(np.int64) != ref.view(np.int64))[0]
    assert bad.size == 0, [(vals[i], int(es[i]), out[i], ref[i]) for i in bad[:5]]
    fits = np.array([-(1 << 63) <= v < (1 << 63) for v in vals])
    bad64 = np.nonzero(fits & (out64.view(np.int64) != ref.view(np.int64)))[0]
    assert bad64.size == 0, [(vals[i], int(es[i]), out64[i], ref[i]) for i in bad64[:5]]


@pytest.mark.parametrize("b,d", [(7, 8), (3, 29)])
@pytest.mark.parametrize("rounding", [0, 1])
def test_gpu_conversion_out_of_range_matches_oracle(gpu, b, d, rounding):
    """Exponent sums (within the 0..4094 that two exponent fields reach) whose results lie
    beyond the double range or below it: truncation writes the exponent into the bits with
    no over- or underflow handling, as the reference's host conversion does, and nearest
    rounding scales with ldexp (infinity, subnormals, zero); the GPU matches the oracle bit
    for bit either way."""
    rng = np.random.default_rng(7 * b + d + 3 * rounding)
    vals = [v for v in cases(rng) if v != 0]
    nbs = np.array([abs(v).bit_length() for v in vals])
    target = np.concatenate([rng.integers(-1130, -1000, len(vals) // 2),
                             rng.integers(1000, 1100, len(vals) - len(vals) // 2)])
    es = np.clip(target - nbs + 2045 + b * (d + 1), 0, 4094).astype(np.int32)
    hi = np.array([(v >> 64) for v in vals], dtype=np.int64)
    lo = np.array([v & ((1 << 64) - 1) for v in vals], dtype=np.uint64)
    out, out64 = oz.debug_convert(hi, lo, es, b, d, rounding)
    ref = np.array([oracle.T_to_double(int(h), int(l.astype(np.int64)), int(e), b, d, rounding)
                    for h, l, e in zip(hi, lo, es)])
    if rounding:   # the range really is left: infinities, subnormals and zeros occur
        assert np.isinf(ref).any() and (ref == 0).any()
        assert ((ref != 0) & (np.abs(ref) < np.finfo(np.float64).tiny)).any()
    bad = np.nonzero(out.view(np.int64) != ref.view(np.int64))[0]
    assert bad.size == 0, [(vals[i], int(es[i]), out[i], ref[i]) for i in bad[:5]]
    fits = np.array([-(1 << 63) <= v < (1 << 63) for v in vals])
    bad64 = np.nonzero(fits & (out64.view(np.int64) != ref.view(np.int64)))[0]
    assert bad64.size == 0, [(vals[i], int(es[i]), out64[i], ref[i]) for i in bad64[:5]]
