"""Build libnumfast_hip.so (gfx950) in-tree with hipcc.

    python -m numfast_amd.build            # build the HIP library
    python -m numfast_amd.build --oracle   # also build the CPU oracle

Every .hip file under csrc/ is compiled to an object in parallel and linked
into ``numfast_amd/libnumfast_hip.so``.  The library is the product: it
exports the numfast-native C ABI (see include/numfast_hip.h).
"""

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libnumfast_hip.so"
ROOT = PKG.parent

ARCH = os.environ.get("NUMFAST_HIP_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -packed-fp32-ops: hipcc must not form packed-FP32 instructions; some of
# their forms have given wrong results on this part without any error.
DEVICE_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden", "-munsafe-fp-atomics",
    "-Wno-unused-command-line-argument",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("build step failed: %s" % cmd[-1])
    return r


def build_lib(jobs=None, verbose=False):
    OBJ.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    objs = []
    todo = []
    for s in srcs:
        o = OBJ / (s.stem + ".o")
        objs.append(o)
        todo.append((s, o))
    jobs = jobs or min(8, os.cpu_count() or 4)

    def compile_one(so):
        s, o = so
        cmd = [HIPCC, *DEVICE_FLAGS, "-I", str(CSRC), "-c", str(s), "-o", str(o)]
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd)
        return o

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(compile_one, todo))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
          *[str(o) for o in objs]])
    os.replace(tmp, LIB)
    return LIB


def build_oracle():
    """The CPU oracle (test infrastructure) -- plain C, no GPU."""
    src = ROOT / "oracle" / "nf_oracle.c"
    out = ROOT / "oracle" / "libnf_oracle.so"
    tmp = out.with_suffix(".so.tmp")
    _run(["gcc", "-O3", "-march=x86-64-v3", "-fPIC", "-shared", "-fvisibility=hidden",
          "-o", str(tmp), str(src)])
    os.replace(tmp, out)
    return out


def build_asan(jobs=8):
    """The host-ASan build of the sources and the boundary driver
    (tests/asan/Makefile: host code instrumented, device code as usual)."""
    r = subprocess.run(["make", "-C", str(ROOT / "tests" / "asan"), "-j%d" % jobs],
                       capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout[-4000:] + r.stderr[-4000:])
        raise RuntimeError("ASan build failed")
    return ROOT / "tests" / "asan" / "boundary_asan"


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    lib = build_lib(verbose="-v" in argv)
    print("built", lib)
    if "--oracle" in argv:
        print("built", build_oracle())
    if "--asan" in argv:
        print("built", build_asan())


if __name__ == "__main__":
    main()
