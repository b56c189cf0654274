"""Builds libsph.so (HIP for gfx950 + the C++ host solver) in-tree.

hipcc compiles each source to an object in parallel, then links the shared
library next to this file. Flags that fix the arithmetic:
  -ffp-contract=off                           no FMA contraction
  -fhip-fp32-correctly-rounded-divide-sqrt     IEEE '/' and sqrtf
  no -ffast-math
and, for every HIP source, the packed-FP32 instructions are disabled.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsph.so")

SOURCES = ["grid.hip", "sort.hip", "pcisph.hip", "wcsph.hip", "boundary.hip", "adjoint.hip", "adjoint.cpp", "solver.cpp", "mesh_samplers.cpp", "scene.cpp", "output.cpp", "rigid.cpp", "batch.cpp", "sph_abi.cpp"]

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + CSRC, "-I" + os.path.join(ROOT, "include"),
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _compile(src: str, verbose: bool) -> str:
    out = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = ["hipcc", "-x", "hip", *COMMON, "-c", os.path.join(CSRC, src), "-o", out]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    # hipcc's host pass warns that it does not know the packed-fp32 feature
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if verbose and err:
        print(err, flush=True)
    return out


def build(verbose: bool = False, jobs: int = 6) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), SOURCES))
    tmp = LIB + ".tmp"
    cmd = ["hipcc", "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


# Host C++ sources (no kernels); the sanitizer build instruments these only.
HOST_SOURCES = [f for f in SOURCES if f.endswith(".cpp")]


def build_host_sanitized(out_dir: str) -> str:
    """libsph with AddressSanitizer and UndefinedBehaviorSanitizer on the host
    C++ sources (solver, samplers, scene files, rigid world, output, batch
    and adjoint drivers, the C ABI) into out_dir; the HIP sources are linked
    from the regular build, so their device code is unchanged. Load it with
    clang's ASan runtime preloaded."""
    build()
    os.makedirs(out_dir, exist_ok=True)
    san = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
           "-Xarch_host", "-fno-omit-frame-pointer"]
    flags = [f for f in COMMON if f != "-O3"] + ["-O1", "-g"]

    def one(src):
        out = os.path.join(out_dir, src + ".o")
        r = subprocess.run(["hipcc", "-x", "hip", *flags, *san, "-c", os.path.join(CSRC, src), "-o", out],
                           capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr}")
        return out
    with cf.ThreadPoolExecutor(max_workers=6) as ex:
        objs = list(ex.map(one, HOST_SOURCES))
    objs += [os.path.join(OBJ, f + ".o") for f in SOURCES if f.endswith(".hip")]
    lib = os.path.join(out_dir, "libsph_asan.so")
    subprocess.run(["hipcc", "--offload-arch=gfx950", "-shared", "-fPIC", *objs, *san, "-Xarch_host",
                    "-shared-libasan", "-o", lib], check=True, capture_output=True)
    return lib


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
