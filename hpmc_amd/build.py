"""Builds hpmc_amd/libhpmc_amd.so (HIP for gfx950) in-tree.

    python -m hpmc_amd.build

Steps: embed the device headers as C++ raw strings for hipRTC, then compile
csrc/hpmc_api.cpp with hipcc into a shared library exporting the C ABI of
include/hpmc.h.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libhpmc_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def embed():
    out = os.path.join(CSRC, "_embedded_sources.inc")
    parts = ["// Generated by hpmc_amd/build.py: device sources for hipRTC.\n"]
    for var, name in (("kEmbeddedDeviceHeader", "hpmc_device.h"), ("kEmbeddedTablesHeader", "mc_tables.h")):
        text = open(os.path.join(CSRC, name)).read()
        assert ")HPMCSRC\"" not in text
        parts.append('static const char %s[] = R"HPMCSRC(%s)HPMCSRC";\n' % (var, text))
    data = "".join(parts)
    if not os.path.exists(out) or open(out).read() != data:
        with open(out, "w") as f:
            f.write(data)


def build(verbose=False):
    embed()
    cmd = [HIPCC] + HIP_FLAGS + ["-I", os.path.join(ROOT, "include"), "-I", CSRC,
                                  os.path.join(CSRC, "hpmc_api.cpp"), "-o", LIB + ".tmp",
                                  "-lhiprtc"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
