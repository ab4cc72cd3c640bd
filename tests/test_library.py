"""Host-side checks of libozaie.so that need no GPU: it loads, exports every function
include/ozaie.h declares, validates arguments and reports the reference's status codes
(include/oz_aie.h:187-191 of the reference) before touching the device."""
import os
import re
import subprocess

import pytest

import ozaie_amd as oz
from conftest import ROOT, has_gpu


def declared_functions():
    names = []
    for hdr in ("ozaie.h",):
        text = open(os.path.join(ROOT, "include", hdr)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_ ]*?\**\s*\**\s*([A-Za-z_][A-Za-z0-9_]*)\s*\(",
                             text, flags=re.M):
            names.append(m.group(1))
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = oz.lib()
    names = declared_functions()
    assert len(names) >= 16
    out = subprocess.check_output(["nm", "-D", "--defined-only", oz.LIB_PATH]).decode()
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        assert hasattr(lib, n)
    assert sorted(oz.EXPORTS) == names


def test_version_and_options():
    assert oz.version() >= 100
    assert oz.get_option(oz.OPT_ROUNDING) in (0, 1)
    assert oz.lib().ozaie_set_option(oz.OPT_ROUNDING, 5) == oz.ERR_ARG
    assert oz.lib().ozaie_set_option(999, 0) == oz.ERR_ARG
    oz.set_option(oz.OPT_KCHUNK, 4096)
    assert oz.get_option(oz.OPT_KCHUNK) == 4096
    oz.set_option(oz.OPT_KCHUNK, 0)


def test_bad_dims_rejected_before_device():
    assert oz.ozAIE_setup(0, 10, 10, oz.Mode.fast) == oz.ERR_DIMS
    assert oz.ozAIE_setup(10, -1, 10, oz.Mode.fast) == oz.ERR_DIMS


def test_strict_mode_reproduces_reference_validation():
    oz.set_option(oz.OPT_STRICT, 1)
    try:
        # the repo example's own shape is rejected by the reference (non-square)
        assert oz.ozAIE_setup(256, 512, 128, oz.Mode.accurate) == oz.ERR_DIMS
        assert oz.ozAIE_setup(64, 64, 64, oz.Mode.accurate) == oz.ERR_DIMS
        assert oz.ozAIE_setup(4096, 4096, 4096, oz.Mode.fast) == oz.ERR_DIMS
        assert oz.ozAIE_setup(127, 127, 127, oz.Mode.fast) == oz.ERR_DIMS
        # int(std::log2(K)) in [7, 11] with M == N == K: the reference accepts K that is not
        # a power of two (include/oz_aie.h:184-191), so strict mode does too
        assert oz.ozAIE_setup(384, 384, 384, oz.Mode.fast) in (oz.OK, oz.ERR_NODEVICE)
        assert oz.ozAIE_setup(4095, 4095, 4095, oz.Mode.accurate) in (oz.OK, oz.ERR_NODEVICE)
    finally:
        oz.set_option(oz.OPT_STRICT, 0)
        oz.release()


def test_dgemm_before_setup():
    import numpy as np
    oz.release()
    A = np.zeros((4, 4))
    assert oz.ozAIE_DGEMM(4, 4, 4, A, A, A) == oz.ERR_NOSETUP


@pytest.mark.skipif(has_gpu(), reason="checks the no-device path")
def test_no_device_reported():
    assert oz.ozAIE_setup(128, 128, 128, oz.Mode.fast) == oz.ERR_NODEVICE
    import numpy as np
    X = np.zeros((4, 4))
    E = np.zeros(4, dtype=np.int32)
    S = np.zeros((8, 4, 4), dtype=np.int8)
    rc = oz.lib().ozaie_debug_split(oz._ptr(X), 4, 4, 4, 1, 7, 8, oz._ptr(E), oz._ptr(S))
    assert rc == oz.ERR_NODEVICE


def test_mode_table():
    assert oz.mode_slices(oz.Mode.accurate, 128) == (3, 29)
    assert oz.mode_slices(oz.Mode.accurate, 2048) == (3, 26)
    assert oz.mode_slices(oz.Mode.accurate, 8192) == (3, 25)
    assert oz.mode_slices(oz.Mode.accurate, 1 << 15) == (4, 18)
    assert oz.mode_slices(oz.Mode.fast, 4096) == (7, 8)
    assert oz.mode_slices(oz.Mode.fast3, 4096) == (7, 3)
    assert oz.mode_slices(oz.Mode.accurate7, 4096) == (7, 7)


def test_cxx_shim_compiles_against_the_abi(tmp_path):
    """include/oz_aie.h lets a reference consumer (namespace oz_aie, Mode::accurate,
    ozAIE_setup / ozAIE_DGEMM) compile and link unchanged."""
    exe = tmp_path / "ozex"
    src = os.path.join(ROOT, "examples", "oz_example.cpp")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"), src,
                           "-o", str(exe), "-L" + os.path.dirname(oz.LIB_PATH), "-lozaie",
                           "-Wl,-rpath," + os.path.dirname(oz.LIB_PATH)])
    assert exe.exists()


@pytest.mark.skipif(not os.path.exists("/root/reference/example/oztest.cpp"),
                    reason="reference example not mounted")
def test_reference_example_compiles_unchanged(tmp_path):
    """The reference's own consumer, read in place, builds against include/oz_aie.h."""
    exe = tmp_path / "oztest"
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"),
                           "/root/reference/example/oztest.cpp", "-o", str(exe),
                           "-L" + os.path.dirname(oz.LIB_PATH), "-lozaie",
                           "-Wl,-rpath," + os.path.dirname(oz.LIB_PATH)])
    assert exe.exists()


def test_exact_entries_vs_rationals_and_mpfr():
    """The accuracy meter (TwoProduct + exact summation) returns the correctly rounded
    exact dot product: checked against Python fractions and against MPFR at 2000 bits."""
    from fractions import Fraction
    import numpy as np
    import oracle
    rng = np.random.default_rng(0)
    A = (rng.random((12, 257)) - 0.5) * np.exp(3 * rng.standard_normal((12, 257)))
    B = (rng.random((257, 9)) - 0.5) * np.exp(3 * rng.standard_normal((257, 9)))
    A[0, :] = 1e300
    B[:, 0] = 1e-300
    A[1, :4] = [1.0, 1e16, -1e16, 3.0]
    rows = np.repeat(np.arange(12), 9)
    cols = np.tile(np.arange(9), 12)
    got = oz.exact_entries(A, B, rows, cols)
    for i, (r, c) in enumerate(zip(rows, cols)):
        ex = sum((Fraction(float(x)) * Fraction(float(y)) for x, y in zip(A[r], B[:, c])), Fraction(0))
        assert got[i] == float(ex)
    if oracle.have_mpfr():
        assert np.array_equal(got, oracle.mpfr_sample(A, B, rows, cols, 2000))


@pytest.mark.skipif(has_gpu(), reason="checks the no-device path")
def test_group_without_device():
    import ctypes
    import numpy as np
    M = np.array([64, 32], dtype=np.int64)
    h = ctypes.c_void_p()
    rc = oz.lib().ozaie_group_create(ctypes.byref(h), 2, oz._ptr(M), oz._ptr(M), oz._ptr(M), 7, 8, 0)
    assert rc == oz.ERR_NODEVICE


def test_sweep_mode_table_and_resume_keys(tmp_path):
    from ozaie_amd import sweep
    assert sweep.mode_bd("accurate", 128) == (3, 29)
    assert sweep.mode_bd("fast", 4096) == (7, 8)
    assert sweep.mode_bd("b6d11", 512) == (6, 11)
    assert sweep.mode_bd("crt", 4096) == (0, 16) and sweep.mode_bd("crt12", 4096) == (0, 12)
    out = tmp_path / "r.json"
    rec = {"kind": "accuracy", "K": 128, "phi": 1.0, "seed": 0, "mode": "fast", "max_rel_err": 1e-10}
    sweep.append(str(out), rec)
    loaded = sweep.load_records(str(out))
    assert loaded == [rec]
    assert sweep.record_key(loaded[0]) == sweep.record_key({"kind": "accuracy", "K": 128, "phi": 1.0,
                                                            "seed": 0, "mode": "fast"})


def test_slice_table_validation():
    """The per-size-class slice-count table of groups created with slices=0 accepts the
    grouped kernels' counts (3, 7, 8, 9) and -1 (untuned), and nothing else."""
    old = oz.slice_table()
    try:
        oz.set_slice_table([3, 7, 8, 9, -1, 9, 8, 7, 3])
        assert list(oz.slice_table()) == [3, 7, 8, 9, -1, 9, 8, 7, 3]
        assert oz.lib().ozaie_tune_set_slices(oz._ptr(__import__("numpy").array([5] * 9, dtype="int32"))) == oz.ERR_ARG
        assert oz.size_class(100, 2000, 300) == 3 * 2 + 1
    finally:
        oz.set_slice_table(old)
