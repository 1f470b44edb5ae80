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
from rtc_util import USER_EDGE_KERNEL, USER_KERNEL, compile_hip

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


class _Hip:
    """The HIP module API through ctypes (the runtime the
    library itself is linked against)."""

    def __init__(self):
        import hpmc_amd
        hpmc_amd.lib()
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.modules = []

    def compile(self, source, name):
        code, log = compile_hip(source)
        assert code is not None, log[-4000:]
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


@pytest.mark.parametrize("kind", ["volume", "snippet"])
def test_four_output_form_hands_the_edge_ends(hip, kind):
    # the reference's extractVertex(a, b, p, n) (shadergen.cpp:310,
    # 390-391): a and b are the texel-centre coordinates (i + 0.5) / size of
    # the vertex's edge ends, the coordinates HPMC_fetch takes
    iso, keep, T = _case(kind)
    ref, ce = iso.extract_all(T, with_cell_edge=True)
    n = ref.shape[0]
    assert n > 500
    fn = hip.compile(iso.traversal_source() + USER_EDGE_KERNEL, b"user_store_edge")
    out = torch.full((n, 6), float("nan"), device="cuda")
    ends = torch.full((n, 6), float("nan"), device="cuda")
    seen = torch.zeros(n, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    hip.launch(fn, iso.traversal_grid(), iso.traversal_view(), out.data_ptr(), ends.data_ptr(), seen.data_ptr())
    assert bool((seen == 1).all()) and torch.equal(out, ref)
    ce = ce.cpu().numpy().view(np.uint32)
    cells = tuple(v - 1 for v in iso.lattice)
    cell, edge = ce[:, 0].astype(np.int64), ce[:, 1].astype(np.int64)
    x, y, z = cell % cells[0], (cell // cells[0]) % cells[1], cell // (cells[0] * cells[1])
    axis, s0, s1 = edge >> 2, edge & 1, (edge >> 1) & 1
    a = np.stack([x + np.where(axis == 0, 0, s0), y + np.where(axis == 0, s0, np.where(axis == 1, 0, s1)),
                  z + np.where(axis == 2, 0, s1)], axis=1)
    b = a + np.eye(3, dtype=np.int64)[axis]
    inv = np.array([np.float32(1.0) / np.float32(v) for v in iso.lattice], np.float32)
    want = np.concatenate([(a.astype(np.float32) + np.float32(0.5)) * inv,
                           (b.astype(np.float32) + np.float32(0.5)) * inv], axis=1)
    assert np.array_equal(ends.cpu().numpy(), want)
    iso.close()
