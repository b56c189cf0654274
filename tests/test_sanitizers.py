"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

* The oracle (oracle/sph_oracle.cpp) built with g++ -fsanitize=address,
  undefined runs the oracle's pin tests.
* libsph's host C++ (settings parser, samplers and mesh voxelizer, scene
  files, rigid world, the C ABI) built with hipcc and the host-only
  sanitizer flags runs the library's CPU tests; its device code is the
  regular build's.
Each run is a child process with the sanitizer runtime preloaded; any
sanitizer report fails the test. Leak checking is off (the Python
interpreter itself does not free everything at exit).
"""
import os
import subprocess
import sys
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

pytestmark = pytest.mark.slow


def _run(env, tests):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="4", **env)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", *tests],
                       cwd=ROOT, env=env, capture_output=True, text=True)
    out = r.stdout + r.stderr
    print(out[-2000:])
    assert r.returncode == 0, out[-4000:]
    assert "ERROR: AddressSanitizer" not in out and "runtime error:" not in out


def test_oracle_under_asan_and_ubsan():
    from oracle import build as ob
    runtime = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    with tempfile.TemporaryDirectory() as tmp:
        lib = ob.build_sanitized(tmp)
        _run({"SPH_ORACLE_LIB": lib, "LD_PRELOAD": runtime},
             ["tests/test_oracle_pins.py", "tests/test_oracle.py", "tests/test_rollout64.py", "-k", "not full_size"])


def test_library_host_code_under_asan_and_ubsan():
    from fluids_amd import build as fb
    rt = subprocess.run(["hipcc", "-print-file-name=libclang_rt.asan-x86_64.so"], capture_output=True,
                        text=True).stdout.strip()
    if not os.path.isabs(rt) or not os.path.exists(rt):
        res = subprocess.run(["/opt/rocm/llvm/bin/clang", "-print-resource-dir"], capture_output=True,
                             text=True).stdout.strip()
        rt = os.path.join(res, "lib", "linux", "libclang_rt.asan-x86_64.so")
    assert os.path.exists(rt), rt
    with tempfile.TemporaryDirectory() as tmp:
        lib = fb.build_host_sanitized(tmp)
        _run({"SPH_LIB": lib, "LD_PRELOAD": rt},
             ["tests/test_abi.py", "tests/test_rigid.py", "tests/test_mesh_samplers.py"])
