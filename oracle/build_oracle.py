"""Build recipe for the oracle's shared libraries (test infrastructure only).

Compiles oracle/ozaki_oracle.c (CPU restatement of the reference Ozaki path),
oracle/crt_oracle.c (the library's CRT mode, which has no reference counterpart) and
oracle/mpfr_oracle.c (the reference's MPFR-106 accuracy oracle) with gcc into
oracle/liboz_oracle.so.  The reference itself (AIE C++ kernels needing aie_api and the
Peano compiler, host code needing XRT) cannot be built in this image, so there is no
oracle/_ref build.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboz_oracle.so")

MPFR_INC = "/opt/conda/include"
MPFR_LIBDIRS = ["/opt/conda/lib", "/lib/x86_64-linux-gnu"]


def build(verbose=False):
    srcs = [os.path.join(HERE, "ozaki_oracle.c"), os.path.join(HERE, "crt_oracle.c")]
    cmd = ["gcc", "-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared", "-std=gnu11"]
    link = []
    if os.path.exists(os.path.join(MPFR_INC, "mpfr.h")):
        srcs.append(os.path.join(HERE, "mpfr_oracle.c"))
        cmd += ["-I" + MPFR_INC, "-DORO_HAVE_MPFR=1"]
        for d in MPFR_LIBDIRS:
            if os.path.exists(os.path.join(d, "libmpfr.so.6")):
                link = ["-L" + d, "-Wl,-rpath," + d, "-l:libmpfr.so.6"]
                break
    tmp = LIB + ".tmp"
    full = cmd + srcs + ["-o", tmp] + link + ["-lm"]
    if verbose:
        print(" ".join(full))
    subprocess.check_call(full)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
