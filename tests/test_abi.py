"""C ABI checks that need no GPU: the library loads, exports every function
include/hpmc.h declares, compiles field snippets with hipRTC, and follows
the reference's NULL-handle conventions (hpmc/src/hpiface.cpp:48-53,
hpmc/src/triface.cpp:46-51, 339-349) without touching a device."""
import ctypes
import os
import re
import subprocess

import pytest

import hpmc_amd
import oracle
from hpmc_amd import workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "hpmc.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(HPMC[A-Za-z0-9]+)\s*\(", text)))


def test_header_declares_reference_api():
    names = set(_declared())
    reference_api = {
        "HPMCcreateConstants", "HPMCdestroyConstants", "HPMCcreateHistoPyramid", "HPMCsetLatticeSize",
        "HPMCsetFieldAsBinary", "HPMCsetFieldAsContinuous", "HPMCsetGridSize", "HPMCsetGridExtent",
        "HPMCsetFieldTexture3D", "HPMCsetFieldCustom", "HPMCgetBuilderProgram", "HPMCdestroyHandle",
        "HPMCbuildHistopyramid", "HPMCacquireNumberOfVertices", "HPMCcreateTraversalHandle",
        "HPMCdestroyTraversalHandle", "HPMCgetTraversalShaderFunctions", "HPMCsetTraversalHandleProgram",
        "HPMCextractVertices", "HPMCextractVerticesTransformFeedback", "HPMCextractVerticesTransformFeedbackNV",
        "HPMCextractVerticesTransformFeedbackEXT",
    }
    assert reference_api <= names


def test_library_exports_every_declared_symbol():
    L = hpmc_amd.lib()
    out = subprocess.run(["nm", "-D", "--defined-only", hpmc_amd.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (HPMC\w+)", out))
    for name in _declared():
        assert name in exported, name
        assert getattr(L, name) is not None
    bound = {n for n, _, _ in hpmc_amd.SIGNATURES}
    assert bound == set(_declared())


def test_null_handles_follow_reference_conventions():
    L = hpmc_amd.lib()
    assert not L.HPMCcreateHistoPyramid(None)
    assert L.HPMCgetLastError(None) == 1
    L.HPMCsetLatticeSize(None, 8, 8, 8)
    L.HPMCsetGridSize(None, 7, 7, 7)
    L.HPMCsetGridExtent(None, 1.0, 1.0, 1.0)
    L.HPMCsetFieldTexture3D(None, None, 0)
    L.HPMCsetFieldCustom(None, b"", 0, 0)
    L.HPMCbuildHistopyramid(None, 0.5)
    assert L.HPMCgetBuilderProgram(None) == 0
    assert L.HPMCacquireNumberOfVertices(None) == 0
    assert not L.HPMCcreateTraversalHandle(None)
    assert not L.HPMCgetTraversalShaderFunctions(None)
    assert not L.HPMCsetTraversalHandleProgram(None, 1, 0, 1, 2)
    for fn in (L.HPMCextractVertices, L.HPMCextractVerticesTransformFeedback,
               L.HPMCextractVerticesTransformFeedbackNV, L.HPMCextractVerticesTransformFeedbackEXT):
        assert not fn(None)
    L.HPMCdestroyHandle(None)
    L.HPMCdestroyTraversalHandle(None)
    L.HPMCdestroyConstants(None)


@pytest.mark.parametrize("src", [workloads.CAYLEY_SNIPPET, workloads.METABALL_SNIPPET])
def test_snippets_compile_for_gfx950(src):
    hpmc_amd.check_snippet(src)
    hpmc_amd.check_snippet(src, binary=True)


def test_gradient_snippet_compiles_and_needs_fetch_grad():
    hpmc_amd.check_snippet(workloads.CAYLEY_GRAD_SNIPPET, gradient=True)
    hpmc_amd.check_snippet(workloads.CAYLEY_GRAD_SNIPPET, gradient=True, binary=True)
    with pytest.raises(hpmc_amd.HPMCError, match="HPMC_fetchGrad"):
        hpmc_amd.check_snippet(workloads.CAYLEY_SNIPPET, gradient=True)


def test_bad_snippet_reports_compile_error():
    with pytest.raises(hpmc_amd.HPMCError) as e:
        hpmc_amd.check_snippet("__device__ float HPMC_fetch(float3 p) { return undefined_name; }")
    assert "HPMC_ERR_COMPILE" in str(e.value) and "undefined_name" in str(e.value)
    with pytest.raises(hpmc_amd.HPMCError):
        hpmc_amd.check_snippet("__device__ float not_the_fetch(float3 p) { return 0.0f; }")


@pytest.mark.parametrize("kind", ["volume", "gradient_volume", "unorm8", "binary", "particles",
                                  "particles_gradient"])
def test_traversal_source_compiles_with_user_kernels(kind):
    # HPMCgetTraversalShaderFunctions' own output, for each field kind, with a
    # two-output and a four-output (edge ends) user kernel, compiled for
    # gfx950 by hipRTC -- no device involved (the volume pointer is never read)
    from rtc_util import USER_EDGE_KERNEL, USER_KERNEL, compile_hip
    L = hpmc_amd.lib()
    iso = hpmc_amd.IsoSurface((40, 36, 30))
    dummy = ctypes.c_void_p(256)
    if kind == "volume":
        L.HPMCsetFieldTexture3D(iso.h, dummy, 0)
    elif kind == "gradient_volume":
        L.HPMCsetFieldTexture3D(iso.h, dummy, 1)
    elif kind == "unorm8":
        L.HPMCsetFieldTexture3DUnorm8(iso.h, dummy)
    elif kind == "binary":
        iso.set_binary(True)
        L.HPMCsetFieldTexture3D(iso.h, dummy, 0)
    elif kind == "particles":
        iso.set_snippet(workloads.PARTICLES_SNIPPET)
    else:
        iso.set_snippet(workloads.PARTICLES_GRAD_SNIPPET, gradient=True)
    src = iso.traversal_source()
    assert "hpmc_for_each_vertex" in src and "#define HPMC_TILES_X" in src
    code, log = compile_hip(src + USER_KERNEL + USER_EDGE_KERNEL)
    assert code is not None and len(code) > 1000, log[-3000:]
    iso.close()


@pytest.mark.parametrize("gradient", [False, True])
def test_particles_snippets_compile_for_gfx950(gradient):
    src = workloads.PARTICLES_GRAD_SNIPPET if gradient else workloads.PARTICLES_SNIPPET
    hpmc_amd.check_snippet(src, gradient=gradient)
    hpmc_amd.check_snippet(src, gradient=gradient, binary=True)


def test_output_buffer_growth_policy():
    # apps/transform_feedback/transform_feedback.cpp:393-402: keep the buffer
    # while it holds N, else (GLsizei)(1.1f * N)
    from hpmc_amd import INITIAL_CAPACITY, grown_capacity
    assert INITIAL_CAPACITY == 3000
    assert grown_capacity(3000, 2999) == 3000
    assert grown_capacity(3000, 3000) == 3000
    assert grown_capacity(3000, 3001) == 3301          # 1.1f * 3001 = 3301.1
    assert grown_capacity(3301, 3301) == 3301
    assert grown_capacity(3301, 10000) == 11000
    # float32 product, truncated: 1.1f is 1.10000002384..., so 1.1f * 3694104
    # = 4063514.49 -> 4063514
    assert grown_capacity(0, 3694104) == 4063514
    for n in (1, 7, 12345, 2 ** 24 + 1, 3 * 10 ** 9):
        c = grown_capacity(0, n)
        assert n <= c <= int(1.1 * n) + 1


def _histopyramid(lattice, grid=None):
    L = hpmc_amd.lib()
    c = L.HPMCcreateConstants(4, 3)
    assert c, "constants are made without a device as well"
    h = L.HPMCcreateHistoPyramid(c)
    assert h
    L.HPMCsetLatticeSize(h, *lattice)
    if grid is not None:
        L.HPMCsetGridSize(h, *grid)
    return L, c, h


@pytest.mark.parametrize("lattice", [(64, 64, 64), (256, 256, 256), (512, 512, 512), (33, 20, 45),
                                     (150, 70, 90), (17, 300, 5), (1000, 10, 200), (16384, 16, 3),
                                     (3, 3, 3), (5, 9, 4), (9, 3, 20), (4, 40, 40), (6, 6, 6)])
def test_layout_query_matches_oracle(lattice):
    # init.cpp:105-130 in FP32, host only
    L, c, h = _histopyramid(lattice)
    out = (ctypes.c_int32 * 5)()
    assert L.HPMCgetLayout(h, out) == 0
    cells = tuple(v - 1 for v in lattice)
    assert tuple(out) == oracle.layout(lattice, cells), (tuple(out), lattice)
    L.HPMCdestroyHandle(h)
    L.HPMCdestroyConstants(c)


@pytest.mark.parametrize("lattice,grid", [((1, 8, 8), None), ((8, 8, 16385), None), ((8, 8, 8), (8, 7, 7)),
                                          ((8, 8, 8), (7, 0, 7)), ((8, 2, 8), None), ((2, 9, 9), None)])
def test_layout_rejects_bad_sizes(lattice, grid):
    # init.cpp:82-103: lattice in [2, 16384], 1 <= grid < lattice; at least 2
    # cells along x and y (one cell gives the reference's log2f(0.5) shift)
    L, c, h = _histopyramid(lattice, grid)
    out = (ctypes.c_int32 * 5)()
    code = L.HPMCgetLayout(h, out)
    assert code == 3, code                # HPMC_ERR_SIZE
    assert L.HPMCgetLastError(h) == code
    L.HPMCdestroyHandle(h)
    L.HPMCdestroyConstants(c)


def test_build_without_device_fails_loudly():
    # no CPU fallback: the build reports a HIP error and latches broken
    # (hpiface.cpp:258-290), the count is 0; a traversal handle is still
    # made (its source can be taken on the host) but extraction fails.
    # (Asked inside the test, through torch: nothing may start a HIP runtime
    # while the tests are collected.)
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("checks the no-device path")
    L, c, h = _histopyramid((64, 64, 64))
    L.HPMCsetFieldCustom(h, workloads.CAYLEY_SNIPPET.encode(), 0, 0)
    L.HPMCbuildHistopyramid(h, 0.0)
    assert L.HPMCgetLastError(h) == 6           # HPMC_ERR_HIP
    assert L.HPMCacquireNumberOfVertices(h) == 0
    th = L.HPMCcreateTraversalHandle(h)
    assert th
    assert not L.HPMCextractVertices(th)
    L.HPMCdestroyTraversalHandle(th)
    L.HPMCdestroyHandle(h)
    L.HPMCdestroyConstants(c)


def test_raw_volume_loader(tmp_path):
    # apps/texture3d/texture3d.cpp:331-344: x*y*z bytes, x fastest; a short
    # file leaves the rest zero, a long one is cut, a missing one is an error
    import numpy as np
    from hpmc_amd import load_raw_volume
    size = (5, 4, 3)
    data = bytes((7 * i + 3) % 256 for i in range(60))
    p = tmp_path / "v.raw"
    p.write_bytes(data)
    v = load_raw_volume(str(p), size)
    assert v.dtype == np.uint8 and v.shape == (3, 4, 5)
    assert v[2, 1, 4] == data[4 + 5 * (1 + 4 * 2)]
    assert bytes(v.ravel()) == data
    p.write_bytes(data + b"\xff" * 9)
    assert bytes(load_raw_volume(str(p), size).ravel()) == data
    p.write_bytes(data[:41])
    v = load_raw_volume(str(p), size).ravel()
    assert bytes(v[:41]) == data[:41] and not v[41:].any()
    with pytest.raises(OSError):
        load_raw_volume(str(tmp_path / "missing.raw"), size)
    with pytest.raises(ValueError):
        load_raw_volume(str(p), (0, 4, 3))
