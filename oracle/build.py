"""Build recipe for the CPU oracle (test infrastructure only).

Compiles ``oracle/sph_oracle.cpp`` into ``oracle/_build/liboracle.so`` with
g++: IEEE fp32 (no -ffast-math), no FMA contraction, OpenMP. The flags that
pin the arithmetic are -ffp-contract=off and the absence of -ffast-math;
-march=x86-64-v3 keeps the library loadable on any AVX2 host.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "liboracle.so")

FLAGS = [
    "-std=c++17", "-O3", "-march=x86-64-v3", "-fopenmp",
    "-ffp-contract=off", "-fno-fast-math", "-fno-math-errno",
    "-fPIC", "-shared",
]


def build(verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    src = os.path.join(HERE, "sph_oracle.cpp")
    tmp = LIB + ".tmp"
    cmd = ["g++", *FLAGS, src, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


# AddressSanitizer + UndefinedBehaviorSanitizer build (tests/test_sanitizers.py)
SANITIZE = ["-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-g"]


def build_sanitized(out_dir: str) -> str:
    """The oracle with ASan and UBSan at -O1 into out_dir (g++'s runtimes:
    load it with libasan preloaded)."""
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "liboracle_asan.so")
    flags = [f for f in FLAGS if f != "-O3"] + ["-O1"] + SANITIZE
    subprocess.run(["g++", *flags, os.path.join(HERE, "sph_oracle.cpp"), "-o", out], check=True)
    return out


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
