"""Build libozaie.so (HIP kernels + C ABI) for gfx950 with hipcc, in-tree."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libozaie.so")
ARCH = os.environ.get("OZAIE_OFFLOAD_ARCH", "gfx950")
SOURCES = ["oz_kernels.hip", "oz_grouped.hip", "oz_crt.hip", "oz_probe.hip", "oz_api.cpp", "oz_group.cpp"]
HOST_SOURCES = ["oz_verify.cpp", "oz_host.cpp"]   # plain host C++ (g++, OpenMP)


def hipcc():
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.sep not in c or os.path.exists(c):
            return c
    return "hipcc"


def build(verbose=False, jobs=2):
    """Compile every source to an object (in parallel), then link the shared library."""
    common = [hipcc(), "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
              # packed FP32 instructions are disabled for all device code
              "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
              "-I" + os.path.join(HERE, "..", "include")]
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        cmd = common + ["-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd.insert(1, "-xhip")
        objs.append(obj)
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT))
    errs = []
    for p, src in zip(procs, SOURCES):
        out, _ = p.communicate()
        text = "\n".join(l for l in out.decode(errors="replace").splitlines()
                         if "packed-fp32-ops" not in l)
        if p.returncode != 0:
            errs.append(f"{src}:\n{text}")
        elif verbose and text.strip():
            print(text)
    for src in HOST_SOURCES:
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off",
               "-I" + os.path.join(HERE, "..", "include"), "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
        if r.returncode != 0:
            errs.append(f"{src}:\n{r.stdout.decode(errors='replace')}")
        objs.append(obj)
    if errs:
        raise RuntimeError("build failed:\n" + "\n".join(errs))
    tmp = LIB + ".tmp"
    link = [hipcc(), "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + ["-lgomp"]
    if verbose:
        print(" ".join(link), flush=True)
    subprocess.check_call(link)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


EXAMPLES_BIN = os.path.join(HERE, "..", "examples", "bin")
REFERENCE_EXAMPLE = "/root/reference/example/oztest.cpp"


def build_examples(verbose=False):
    """Compile the consumers of the reference-compatible C++ API against include/oz_aie.h
    and libozaie.so: examples/oz_example.cpp, and the reference's own example/oztest.cpp
    read in place where the reference is mounted (compiled unchanged; the binary travels
    with the tree, the source does not).  Outputs go to examples/bin/ with an rpath to
    ozaie_amd/."""
    os.makedirs(EXAMPLES_BIN, exist_ok=True)
    srcs = {"oz_example": os.path.join(HERE, "..", "examples", "oz_example.cpp")}
    if os.path.exists(REFERENCE_EXAMPLE):
        srcs["oztest"] = REFERENCE_EXAMPLE
    built = []
    for name, src in srcs.items():
        out = os.path.join(EXAMPLES_BIN, name)
        cmd = ["g++", "-std=c++17", "-O1", "-I" + os.path.join(HERE, "..", "include"), src, "-o", out + ".tmp",
               "-L" + HERE, "-lozaie", "-Wl,-rpath,$ORIGIN/../../ozaie_amd"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        os.replace(out + ".tmp", out)
        built.append(out)
    return built


if __name__ == "__main__":
    build(verbose=True)
    build_examples(verbose=True)
    sys.exit(0)
