"""The consumers of the reference-compatible C++ API, run on the GPU: the repository's own
example (ones x ones in the accurate mode, checks C == K everywhere) and the reference's
example/oztest.cpp compiled unchanged against include/oz_aie.h (exits with the status of
oz_aie::ozAIE_DGEMM, example/oztest.cpp:28).  Both binaries are built by
ozaie_amd.build.build_examples (the reference's source is read where it is mounted; the
binary travels with the tree)."""
import os
import subprocess

import pytest

from conftest import ROOT

BIN = os.path.join(ROOT, "examples", "bin")


def _run(name):
    exe = os.path.join(BIN, name)
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built (ozaie_amd.build.build_examples)")
    return subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=120)


@pytest.mark.gpu
def test_repo_example_runs(gpu):
    r = _run("oz_example")
    out = r.stdout.decode(errors="replace")
    assert r.returncode == 0, out
    assert "ok: C == 128 everywhere" in out


@pytest.mark.gpu
def test_reference_oztest_runs_unchanged(gpu):
    r = _run("oztest")
    assert r.returncode == 0, r.stdout.decode(errors="replace")
