"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU; no
device involved):

* the oracle: oracle/sanitize_check.c runs every oracle function on small
  lattices (odd, even, non-cubic, grids smaller than the lattice, the
  narrowest grids) in a gcc -fsanitize=address,undefined build;
* the library's host side (layout, setters, error paths, hipRTC snippet and
  traversal-source compiles): a hipcc build with the sanitizers on the host
  pass only (-Xarch_host), loaded through HPMC_AMD_LIBRARY into a child
  process that runs tests/test_abi.py with the sanitizer runtime preloaded.

Any report aborts the child with a non-zero status."""
import glob
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def test_oracle_under_asan_ubsan(tmp_path):
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True)
    cases = tmp_path / "cases.bin"
    np.concatenate([oracle.CASE_COUNTS.astype("<i4").ravel(), oracle.CASE_SLOTS.astype("<i4").ravel()]).tofile(cases)
    r = subprocess.run([os.path.join(ROOT, "oracle", "_san", "sanitize_check"), str(cases)], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert r.stdout.strip().startswith("ok ")


def _asan_runtime():
    found = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    return found[-1] if found else None


def test_library_host_code_under_asan_ubsan():
    rt = _asan_runtime()
    if rt is None or not os.path.exists(HIPCC):
        pytest.skip("needs hipcc and the clang sanitizer runtime")
    out_dir = os.path.join(ROOT, "hpmc_amd", "_san")
    os.makedirs(out_dir, exist_ok=True)
    lib = os.path.join(out_dir, "libhpmc_amd.so")
    from hpmc_amd import build as hb
    hb.embed()
    cmd = [HIPCC, "--offload-arch=gfx950", "-O1", "-g", "-std=c++17", "-fPIC", "-shared",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
           "-Xarch_host", "-fno-omit-frame-pointer", "-shared-libsan",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "hpmc_amd", "csrc"),
           os.path.join(ROOT, "hpmc_amd", "csrc", "hpmc_api.cpp"), "-o", lib, "-lhiprtc"]
    subprocess.run(cmd, check=True, capture_output=True, timeout=600)
    env = dict(os.environ, HPMC_AMD_LIBRARY=lib, LD_PRELOAD=rt, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([sys.executable, "-m", "pytest", os.path.join(ROOT, "tests", "test_abi.py"), "-x", "-q",
                        "-p", "no:cacheprovider"], env=env, capture_output=True, text=True, timeout=900, cwd=ROOT)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert " passed" in r.stdout
