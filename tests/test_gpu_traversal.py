"""User-fused traversal kernels (HPMCgetTraversalShaderFunctions,
HPMCgetTraversalView, HPMCgetTraversalGrid; the reference's
triface.cpp:125-192 traversal functions for the caller's own shader).  A
kernel compiled with hipRTC from the returned source visits the build's
vertices through hpmc_for_each_vertex.  The bar: every vertex exactly once,
with the extraction's index and bits, for volume, snippet, binary, gradient
and 8-bit fields, on the suggested grid and on one that is not a multiple of
eight, and again after a rebuild at another iso value."""
import ctypes

import numpy as np
import pytest

from hpmc_amd import IsoSurface, workloads
from parity_util import smooth_noise, with_gradient

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

USER_KERNEL = r"""
struct StoreVertex {
    static constexpr bool kActive = true;
    float* out;
    unsigned* seen;
    __device__ void operator()(uint32_t k, float3 n, float3 p) const {
        float* o = out + 6 * (size_t)k;
        o[0] = n.x; o[1] = n.y; o[2] = n.z; o[3] = p.x; o[4] = p.y; o[5] = p.z;
        atomicAdd(seen + k, 1u);
    }
};
extern "C" __global__ __launch_bounds__(256) void user_store(hpmc::TraversalView v, float* out, unsigned* seen)
{
    hpmc_for_each_vertex(v, StoreVertex{out, seen});
}
"""

# the options the library compiles its own snippets with
RTC_OPTS = [b"--offload-arch=gfx950", b"-O3", b"-std=c++17", b"-ffp-contract=off",
            b"-Xclang", b"-target-feature", b"-Xclang", b"-packed-fp32-ops"]


class _Hip:
    """hipRTC and the HIP module API through ctypes (the runtime the
    library itself is linked against)."""

    def __init__(self):
        import hpmc_amd
        hpmc_amd.lib()
        self.rtc = ctypes.CDLL("libhiprtc.so.7")
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.modules = []

    def compile(self, source, name):
        rtc = self.rtc
        prog = ctypes.c_void_p()
        assert rtc.hiprtcCreateProgram(ctypes.byref(prog), source.encode(), b"user.hip", 0, None, None) == 0
        opts = (ctypes.c_char_p * len(RTC_OPTS))(*RTC_OPTS)
        r = rtc.hiprtcCompileProgram(prog, len(RTC_OPTS), opts)
        n = ctypes.c_size_t()
        rtc.hiprtcGetProgramLogSize(prog, ctypes.byref(n))
        log = ctypes.create_string_buffer(n.value + 1)
        rtc.hiprtcGetProgramLog(prog, log)
        assert r == 0, log.value.decode(errors="replace")[-4000:]
        rtc.hiprtcGetCodeSize(prog, ctypes.byref(n))
        code = ctypes.create_string_buffer(n.value)
        assert rtc.hiprtcGetCode(prog, code) == 0
        rtc.hiprtcDestroyProgram(ctypes.byref(prog))
        module = ctypes.c_void_p()
        assert self.hip.hipModuleLoadData(ctypes.byref(module), code) == 0
        self.modules.append(module)
        fn = ctypes.c_void_p()
        assert self.hip.hipModuleGetFunction(ctypes.byref(fn), module, name) == 0
        return fn

    def launch(self, fn, grid, view, *ptrs):
        vbuf = ctypes.create_string_buffer(view, len(view))
        pvals = [ctypes.c_void_p(p) for p in ptrs]
        params = (ctypes.c_void_p * (1 + len(ptrs)))(
            ctypes.cast(vbuf, ctypes.c_void_p), *[ctypes.cast(ctypes.byref(p), ctypes.c_void_p) for p in pvals])
        f = self.hip.hipModuleLaunchKernel
        f.argtypes = [ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint,
                      ctypes.c_uint, ctypes.c_uint, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        assert f(fn, grid, 1, 1, 256, 1, 1, 0, None, params, None) == 0
        torch.cuda.synchronize()

    def close(self):
        for m in self.modules:
            self.hip.hipModuleUnload(m)


@pytest.fixture(scope="module")
def hip():
    h = _Hip()
    yield h
    h.close()


def _visit(hip, iso, grid=None):
    """Runs the user kernel over iso's last build: (vertices, visit counts)."""
    n = iso.count()
    fn = hip.compile(iso.traversal_source() + USER_KERNEL, b"user_store")
    out = torch.full((max(n, 1), 6), float("nan"), dtype=torch.float32, device="cuda")
    seen = torch.zeros(max(n, 1), dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    hip.launch(fn, grid or iso.traversal_grid(), iso.traversal_view(), out.data_ptr(), seen.data_ptr())
    return out[:n], seen[:n]


def _case(kind):
    lattice = (70, 60, 50)
    iso = IsoSurface(lattice)
    keep = None
    if kind == "volume":
        keep = torch.from_numpy(smooth_noise(lattice, 5)).cuda()
        iso.set_volume(keep)
        T = 0.5
    elif kind == "snippet":
        iso.set_snippet(workloads.CAYLEY_SNIPPET)
        T = 0.0
    elif kind == "unorm8":
        b = np.clip(np.rint(smooth_noise(lattice, 8) * 255.0), 0, 255).astype(np.uint8)
        keep = torch.from_numpy(b).cuda()
        iso.set_volume(keep)
        T = 0.5
    elif kind == "binary":
        iso.set_binary(True)
        keep = torch.from_numpy((smooth_noise(lattice, 6) > 0.5).astype(np.float32)).cuda()
        iso.set_volume(keep)
        T = 0.5
    else:
        keep = torch.from_numpy(with_gradient(smooth_noise(lattice, 7))).cuda()
        iso.set_volume(keep, gradient=True)
        T = 0.5
    return iso, keep, T


@pytest.mark.parametrize("kind", ["volume", "snippet", "binary", "gradient", "unorm8"])
def test_user_kernel_visits_the_extracted_vertices(hip, kind):
    iso, keep, T = _case(kind)
    ref, _ = iso.extract_all(T)
    assert ref.shape[0] > 500
    got, seen = _visit(hip, iso)
    assert bool((seen == 1).all()), "a vertex was visited %s times" % (int(seen.max()),)
    assert torch.equal(got, ref)
    iso.close()


def test_any_grid_and_rebuilds(hip):
    iso, keep, T = _case("volume")
    ref, _ = iso.extract_all(T)
    got, seen = _visit(hip, iso, grid=13)        # not a multiple of 8
    assert bool((seen == 1).all()) and torch.equal(got, ref)
    # the view follows the last build
    ref2, _ = iso.extract_all(0.6)
    assert ref2.shape[0] != ref.shape[0]
    got2, seen2 = _visit(hip, iso)
    assert bool((seen2 == 1).all()) and torch.equal(got2, ref2)
    iso.close()
