"""MI355X HistoPyramid marching cubes: Python mirror of HPMC's C API.

``hpmc_amd.lib()`` is the ctypes handle on ``libhpmc_amd.so`` (the C ABI of
``include/hpmc.h``); every ``HPMC*`` function of the reference's public API
(hpmc/include/hpmc.h:295-544) is bound under its own name with the same
arguments, GL object names replaced by device pointers.  ``IsoSurface`` wraps
the usual call sequence (constants -> HistoPyramid -> setters -> build ->
count -> traversal handle -> extract) around torch device tensors.

The library is HIP-only: importing works anywhere, but any call that needs
the library raises ``RuntimeError`` when ``libhpmc_amd.so`` is missing, and
device calls fail (no fallback path exists).
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# HPMC_AMD_LIBRARY names another build of the library (a sanitizer build of
# the host code, for instance); the default is the in-tree libhpmc_amd.so
LIB_PATH = os.environ.get("HPMC_AMD_LIBRARY") or os.path.join(HERE, "libhpmc_amd.so")

HPMC_OK = 0
ERRORS = {
    0: "HPMC_OK", 1: "HPMC_ERR_NULL", 2: "HPMC_ERR_BROKEN", 3: "HPMC_ERR_SIZE",
    4: "HPMC_ERR_FIELD", 5: "HPMC_ERR_COMPILE", 6: "HPMC_ERR_HIP",
    7: "HPMC_ERR_CAPACITY", 8: "HPMC_ERR_ARG", 9: "HPMC_ERR_UNSUPPORTED",
}

# bytes of hpmc::TraversalView (HPMC_TRAVERSAL_VIEW_BYTES of include/hpmc.h)
TRAVERSAL_VIEW_BYTES = 64

_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
_u8 = ctypes.c_uint8
_f = ctypes.c_float
_b = ctypes.c_bool

# (name, restype, argtypes) of every export of include/hpmc.h
SIGNATURES = [
    ("HPMCcreateConstants", _vp, [_i32, _i32]),
    ("HPMCdestroyConstants", None, [_vp]),
    ("HPMCcreateHistoPyramid", _vp, [_vp]),
    ("HPMCsetLatticeSize", None, [_vp, _i32, _i32, _i32]),
    ("HPMCsetFieldAsBinary", None, [_vp]),
    ("HPMCsetFieldAsContinuous", None, [_vp]),
    ("HPMCsetGridSize", None, [_vp, _i32, _i32, _i32]),
    ("HPMCsetGridExtent", None, [_vp, _f, _f, _f]),
    ("HPMCsetFieldTexture3D", None, [_vp, _vp, _u8]),
    ("HPMCsetFieldTexture3DUnorm8", None, [_vp, _vp]),
    ("HPMCsetFieldCustom", None, [_vp, ctypes.c_char_p, _u32, _u8]),
    ("HPMCgetBuilderProgram", _u32, [_vp]),
    ("HPMCdestroyHandle", None, [_vp]),
    ("HPMCbuildHistopyramid", None, [_vp, _f]),
    ("HPMCacquireNumberOfVertices", _u32, [_vp]),
    ("HPMCcreateTraversalHandle", _vp, [_vp]),
    ("HPMCdestroyTraversalHandle", None, [_vp]),
    ("HPMCgetTraversalShaderFunctions", _vp, [_vp]),
    ("HPMCsetTraversalHandleProgram", _b, [_vp, _u32, _u32, _u32, _u32]),
    ("HPMCgetTraversalView", _i32, [_vp, _vp, ctypes.c_size_t]),
    ("HPMCgetTraversalGrid", _i32, [_vp]),
    ("HPMCextractVertices", _b, [_vp]),
    ("HPMCextractVerticesTransformFeedback", _b, [_vp]),
    ("HPMCextractVerticesTransformFeedbackNV", _b, [_vp]),
    ("HPMCextractVerticesTransformFeedbackEXT", _b, [_vp]),
    ("HPMCgetLastError", _i32, [_vp]),
    ("HPMCgetLastErrorString", ctypes.c_char_p, [_vp]),
    ("HPMCsetStream", _i32, [_vp, _vp]),
    ("HPMCsetTuning", _i32, [_vp, _i32, _i32, _i32]),
    ("HPMCgetTuning", _i32, [_vp, ctypes.POINTER(_i32)]),
    ("HPMCsetFieldUniform", _i32, [_vp, ctypes.c_char_p, _vp, ctypes.c_size_t]),
    ("HPMCsetTraversalHandleBuffer", _i32, [_vp, _vp, _u64]),
    ("HPMCsetTraversalHandleCellEdgeBuffer", _i32, [_vp, _vp, _u64]),
    ("HPMCacquireDroppedVertices", _u32, [_vp]),
    ("HPMCgetCellCodes", _i32, [_vp, _vp]),
    ("HPMCgetLevel4", _u64, [_vp, _vp, _u64]),
    ("HPMCsetProfiling", _i32, [_vp, _i32]),
    ("HPMCgetStageTimes", _i32, [_vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_u32)]),
    ("HPMCcheckFieldSnippet", _i32, [ctypes.c_char_p, _u32, ctypes.c_char_p, ctypes.c_size_t]),
    ("HPMCgetLayout", _i32, [_vp, ctypes.POINTER(_i32)]),
    ("HPMCadjoint", _i32, [_vp, _vp, _vp, _vp]),
    ("HPMCbatchAdjoint", _i32, [_vp, _vp, _vp, _vp]),
    ("HPMCcreateBatch", _vp, [_vp]),
    ("HPMCdestroyBatch", None, [_vp]),
    ("HPMCbatchAddVolume", _i32, [_vp, _vp, _i32, _i32, _i32, _f, _f, _f]),
    ("HPMCbatchNumVolumes", _i32, [_vp]),
    ("HPMCbatchSetTuning", _i32, [_vp, _i32, _i32, _i32]),
    ("HPMCbatchGetTuning", _i32, [_vp, ctypes.POINTER(_i32)]),
    ("HPMCbatchSetStream", _i32, [_vp, _vp]),
    ("HPMCbatchBuild", _i32, [_vp, _f]),
    ("HPMCbatchGetCounts", _i32, [_vp, ctypes.POINTER(_u32), ctypes.POINTER(_u32), ctypes.POINTER(_u64)]),
    ("HPMCbatchExtract", _i32, [_vp, _vp, _u64, _vp]),
    ("HPMCbatchCaptureGraph", _i32, [_vp, _f, _vp, _u64, _vp]),
    ("HPMCbatchLaunchGraph", _i32, [_vp]),
    ("HPMCbatchAutotune", _i32, [_vp, _f, _vp, _u64, _i32, ctypes.POINTER(ctypes.c_double)]),
    ("HPMCextractVolumeNoAlloc", _i32, [_vp, _vp, _i32, _i32, _i32, _f, _f, _f, _f, _vp, _u64, _vp, _vp]),
]

# device scratch of HPMCextractVolumeNoAlloc (HPMC_NO_ALLOC_SCRATCH_BYTES)
NO_ALLOC_SCRATCH_BYTES = 80 << 20

_LIB = None


def lib():
    """The loaded C library; raises RuntimeError if it was never built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libhpmc_amd.so is missing: run `python -m hpmc_amd.build` "
                               "(the engine has no non-HIP fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def last_error(h=None):
    L = lib()
    return int(L.HPMCgetLastError(h)), (L.HPMCgetLastErrorString(h) or b"").decode()


class HPMCError(RuntimeError):
    pass


def _check(code, h=None):
    if code != HPMC_OK:
        err, msg = last_error(h)
        raise HPMCError("%s: %s" % (ERRORS.get(code, code), msg))


def check_snippet(source, gradient=False, binary=False):
    """Compile a field snippet for gfx950 without a device; returns the log."""
    buf = ctypes.create_string_buffer(1 << 16)
    flags = (1 if gradient else 0) | (2 if binary else 0)
    code = lib().HPMCcheckFieldSnippet(source.encode(), flags, buf, len(buf))
    _check(code)
    return buf.value.decode()


# Output buffers kept across extractions start at the reference demo's 3000
# vertices and grow to 1.1 x the count when it no longer fits
# (apps/transform_feedback/transform_feedback.cpp:341-349, 393-402).
INITIAL_CAPACITY = 3 * 1000


def grown_capacity(capacity, n):
    """The capacity a kept output buffer needs for n vertices: unchanged
    while it holds them, else (GLsizei)(1.1f * n) as the reference's demo
    computes it (float32 product, truncated), never below n."""
    import numpy as np
    if capacity >= n:
        return capacity
    return max(int(np.float32(1.1) * np.float32(n)), n)


def load_raw_volume(path, size):
    """Reads an 8-bit raw volume file (x-fastest bytes, no header) for a
    lattice `size` = (x, y, z) into a uint8 array shaped (z, y, x), the
    input of IsoSurface.set_volume's 8-bit path.  Like the reference's
    texture demo (apps/texture3d/texture3d.cpp:331-344): the first x*y*z
    bytes are used, a shorter file leaves the rest zero, and a file that
    cannot be opened is an error."""
    import numpy as np
    x, y, z = (int(v) for v in size)
    if min(x, y, z) < 1:
        raise ValueError("lattice size must be positive: %r" % (size,))
    n = x * y * z
    with open(path, "rb") as f:
        data = f.read(n)
    vol = np.zeros(n, np.uint8)
    vol[:len(data)] = np.frombuffer(data, np.uint8)
    return vol.reshape(z, y, x)


def _libc_free(p):
    """free() for strings the library hands out with malloc."""
    libc = ctypes.CDLL(None)
    libc.free.argtypes = [_vp]
    libc.free.restype = None
    libc.free(p)


def _vertex_grads(grad_vertices, count, device):
    """dL/dvertices for an adjoint call: a contiguous float32 tensor of
    shape (count, 6) on the field's device (the kernel reads count * 6
    floats), or None when there are no vertices."""
    import torch
    if grad_vertices is None:
        if count:
            raise ValueError("grad_vertices is None but the surface has %d vertices" % count)
        return None
    if grad_vertices.dtype != torch.float32 or grad_vertices.device != device:
        raise ValueError("grad_vertices must be float32 on %s, got %s on %s"
                         % (device, grad_vertices.dtype, grad_vertices.device))
    if tuple(grad_vertices.shape) != (count, 6):
        raise ValueError("grad_vertices must have shape (%d, 6), got %s" % (count, tuple(grad_vertices.shape)))
    return grad_vertices.contiguous()


def _stream_ptr(stream):
    return ctypes.c_void_p(stream.cuda_stream) if stream is not None else None


def constants():
    """The process's HPMCConstants handle (created on first use, with the
    current device current)."""
    if IsoSurface._constants is None:
        c = lib().HPMCcreateConstants(4, 3)
        if not c:
            raise HPMCError("HPMCcreateConstants failed: %s" % (last_error()[1],))
        IsoSurface._constants = c
    return IsoSurface._constants


def extract_volume_no_alloc(volume, iso, out, count=None, extent=(1.0, 1.0, 1.0), stream=None):
    """Build + extraction of a float32 device volume (z, y, x) in one call
    that allocates nothing (HPMCextractVolumeNoAlloc): N3F_V3F rows go to
    `out` (float32 (capacity, 6) device tensor, extra vertices dropped), the
    vertex count to `count` (a one-element int32 device tensor) if given.
    Enqueued on `stream` (torch stream or None); no host wait.  Calls that
    overlap on one device race for its scratch."""
    import torch
    assert volume.dtype == torch.float32 and volume.is_contiguous() and volume.is_cuda
    assert out.dtype == torch.float32 and out.is_contiguous() and out.is_cuda and out.dim() == 2 and out.shape[1] == 6
    assert count is None or (count.dtype == torch.int32 and count.is_cuda and count.numel() >= 1)
    z, y, x = volume.shape
    code = lib().HPMCextractVolumeNoAlloc(constants(), ctypes.c_void_p(volume.data_ptr()), x, y, z,
                                          *[float(e) for e in extent], float(iso), ctypes.c_void_p(out.data_ptr()),
                                          out.shape[0], ctypes.c_void_p(count.data_ptr()) if count is not None else None,
                                          _stream_ptr(stream))
    _check(code)


class IsoSurface:
    """One HistoPyramid plus one traversal handle, on torch's current device.

    Mirrors the reference call order (apps/texture3d/texture3d.cpp:144-221):
    create constants and HP, set lattice/grid/extent and field, build with an
    iso-value, count, extract into an N3F_V3F buffer.
    """

    _constants = None

    def __init__(self, lattice, grid=None, extent=(1.0, 1.0, 1.0), stream=None):
        L = lib()
        if IsoSurface._constants is None:
            c = L.HPMCcreateConstants(4, 3)
            if not c:
                raise HPMCError("HPMCcreateConstants failed: %s" % (last_error()[1],))
            IsoSurface._constants = c
        self.h = L.HPMCcreateHistoPyramid(IsoSurface._constants)
        if not self.h:
            raise HPMCError("HPMCcreateHistoPyramid failed")
        self.lattice = tuple(int(v) for v in lattice)
        L.HPMCsetLatticeSize(self.h, *self.lattice)
        self.grid = tuple(int(v) for v in grid) if grid is not None else tuple(v - 1 for v in self.lattice)
        if grid is not None:
            L.HPMCsetGridSize(self.h, *self.grid)
        L.HPMCsetGridExtent(self.h, *[float(v) for v in extent])
        if stream is not None:
            _check(L.HPMCsetStream(self.h, _stream_ptr(stream)), self.h)
        self.th = None
        self._volume = None
        self._out = None
        self._kept = None            # (vertices, cell_edge or None) of extract_growing
        self._ce = None

    # -- field ---------------------------------------------------------------
    def set_volume(self, volume, gradient=False):
        """volume: contiguous device tensor shaped (z, y, x): float32, or
        uint8 for an 8-bit normalised volume (sample c / 255); or, with
        gradient=True, float32 (z, y, x, 4) holding (gradient xyz, value) per
        lattice point like an RGBA gradient texture."""
        import torch
        assert volume.dtype in (torch.float32, torch.uint8) and volume.is_contiguous() and volume.is_cuda
        assert not (gradient and volume.dtype == torch.uint8), "8-bit volumes have no gradient variant"
        shape = self.lattice[::-1] + ((4,) if gradient else ())
        assert tuple(volume.shape) == shape, (volume.shape, shape)
        self._volume = volume
        if volume.dtype == torch.uint8:
            lib().HPMCsetFieldTexture3DUnorm8(self.h, ctypes.c_void_p(volume.data_ptr()))
        else:
            lib().HPMCsetFieldTexture3D(self.h, ctypes.c_void_p(volume.data_ptr()), int(bool(gradient)))

    def set_snippet(self, source, gradient=False):
        """HIP C++ defining HPMC_fetch(float3) (and HPMC_fetchGrad(float3) ->
        float4 (gradient, value) with gradient=True)."""
        lib().HPMCsetFieldCustom(self.h, source.encode(), 0, int(bool(gradient)))

    def set_binary(self, binary=True):
        """Binary field: classify against 0.5, emit edge midpoints and face
        normals.  As in HPMC, the flag takes effect at the next
        reconfiguration (the first build, or after a setter that changes the
        lattice, grid, extent or field mode)."""
        L = lib()
        (L.HPMCsetFieldAsBinary if binary else L.HPMCsetFieldAsContinuous)(self.h)

    def set_uniform(self, name, array):
        import numpy as np
        a = np.ascontiguousarray(array)
        _check(lib().HPMCsetFieldUniform(self.h, name.encode(), a.ctypes.data_as(ctypes.c_void_p), a.nbytes),
               self.h)

    def layout(self):
        out = (_i32 * 5)()
        _check(lib().HPMCgetLayout(self.h, out), self.h)
        return tuple(out)

    # -- build / count ----------------------------------------------------------
    def build(self, iso):
        L = lib()
        L.HPMCbuildHistopyramid(self.h, float(iso))
        err, msg = last_error(self.h)
        if err != HPMC_OK:
            raise HPMCError("build failed: %s: %s" % (ERRORS.get(err, err), msg))

    def count(self):
        return int(lib().HPMCacquireNumberOfVertices(self.h))

    # -- extraction ----------------------------------------------------------
    def _traversal(self):
        if self.th is None:
            self.th = lib().HPMCcreateTraversalHandle(self.h)
            if not self.th:
                raise HPMCError("HPMCcreateTraversalHandle failed: %s" % (last_error(self.h)[1],))
        return self.th

    def traversal_source(self):
        """hipRTC source for kernels that fuse the traversal with their own
        per-vertex work (defines hpmc::TraversalView and
        hpmc_for_each_vertex)."""
        L = lib()
        p = L.HPMCgetTraversalShaderFunctions(self._traversal())
        if not p:
            raise HPMCError("HPMCgetTraversalShaderFunctions failed: %s" % (last_error(self.h)[1],))
        try:
            return ctypes.string_at(p).decode()
        finally:
            _libc_free(p)

    def traversal_view(self):
        """The 64-byte hpmc::TraversalView of the last build (bytes)."""
        buf = ctypes.create_string_buffer(TRAVERSAL_VIEW_BYTES)
        _check(lib().HPMCgetTraversalView(self._traversal(), buf, TRAVERSAL_VIEW_BYTES), self.h)
        return buf.raw

    def traversal_grid(self):
        """Workgroups of 256 threads for a traversal kernel."""
        n = int(lib().HPMCgetTraversalGrid(self._traversal()))
        if n <= 0:
            raise HPMCError("HPMCgetTraversalGrid failed: %s" % (last_error(self.h)[1],))
        return n

    def bind_output(self, out, cell_edge=None):
        """out: float32 device tensor (capacity, 6); cell_edge: int32 (capacity, 2) or None."""
        L = lib()
        th = self._traversal()
        self._out = out
        _check(L.HPMCsetTraversalHandleBuffer(th, ctypes.c_void_p(out.data_ptr()), out.shape[0]), self.h)
        self._ce = cell_edge
        if cell_edge is not None:
            _check(L.HPMCsetTraversalHandleCellEdgeBuffer(th, ctypes.c_void_p(cell_edge.data_ptr()),
                                                          cell_edge.shape[0]), self.h)
        else:
            _check(L.HPMCsetTraversalHandleCellEdgeBuffer(th, None, 0), self.h)

    def extract(self):
        if not lib().HPMCextractVertices(self._traversal()):
            raise HPMCError("extract failed: %s" % (last_error(self.h)[1],))

    def extract_all(self, iso, with_cell_edge=False):
        """Build, count and extract into fresh tensors: (vertices (N,6), cell_edge (N,2) or None)."""
        import torch
        self.build(iso)
        n = self.count()
        dev = torch.device("cuda", torch.cuda.current_device())
        out = torch.empty((max(n, 1), 6), dtype=torch.float32, device=dev)
        ce = torch.empty((max(n, 1), 2), dtype=torch.int32, device=dev) if with_cell_edge else None
        self.bind_output(out, ce)
        if n:
            self.extract()
        return out[:n], (ce[:n] if ce is not None else None)

    def extract_growing(self, iso, with_cell_edge=False):
        """Build, count and extract into buffers kept across calls, grown
        like the reference's transform-feedback demo (grown_capacity):
        returns views of the first N rows, valid until the next call."""
        import torch
        self.build(iso)
        n = self.count()
        kept = self._kept
        cap = kept[0].shape[0] if kept is not None else 0
        has_ce = kept is not None and kept[1] is not None
        new_cap = grown_capacity(max(cap, INITIAL_CAPACITY) if kept is None else cap, n)
        if kept is None or new_cap != cap or (with_cell_edge and not has_ce):
            dev = torch.device("cuda", torch.cuda.current_device())
            out = torch.empty((new_cap, 6), dtype=torch.float32, device=dev)
            ce = torch.empty((new_cap, 2), dtype=torch.int32, device=dev) if with_cell_edge else None
            self._kept = (out, ce)
        out, ce = self._kept
        self.bind_output(out, ce if with_cell_edge else None)
        if n:
            self.extract()
        return out[:n], (ce[:n] if with_cell_edge else None)

    def cell_codes(self):
        import torch
        n = self.grid[0] * self.grid[1] * self.grid[2]
        codes = torch.empty(n, dtype=torch.uint8, device="cuda")
        _check(lib().HPMCgetCellCodes(self.h, ctypes.c_void_p(codes.data_ptr())), self.h)
        return codes.view(self.grid[2], self.grid[1], self.grid[0])

    def level4(self):
        """Block sums of the last build by Morton block code (16x16-cell
        blocks, narrower ones for grids of at most 8 cells along x or y)."""
        import torch
        n = lib().HPMCgetLevel4(self.h, None, 0)
        if n == 0:
            raise HPMCError("level4 failed: %s" % (last_error(self.h)[1],))
        buf = torch.empty(n, dtype=torch.int32, device="cuda")
        n = lib().HPMCgetLevel4(self.h, ctypes.c_void_p(buf.data_ptr()), buf.numel())
        if n == 0:
            raise HPMCError("level4 failed: %s" % (last_error(self.h)[1],))
        return buf

    def adjoint(self, grad_vertices):
        """dL/dfield (shaped like the volume) and dL/dT (0-dim) of the last
        build + extraction from dL/dvertices (count, 6); deterministic."""
        import torch
        assert self._volume is not None, "adjoint needs a volume field"
        gv = _vertex_grads(grad_vertices, self.count(), self._volume.device)
        gf = torch.empty_like(self._volume)
        gt = torch.empty((), dtype=torch.float32, device=self._volume.device)
        _check(lib().HPMCadjoint(self._traversal(), ctypes.c_void_p(gv.data_ptr()) if gv is not None else None,
                                 ctypes.c_void_p(gf.data_ptr()), ctypes.c_void_p(gt.data_ptr())), self.h)
        return gf, gt

    def set_tuning(self, classify_waves=0, zseg=0, emit_workgroups=0):
        """Launch shapes (0 = automatic); results do not depend on them."""
        _check(lib().HPMCsetTuning(self.h, classify_waves, zseg, emit_workgroups), self.h)

    def tuning(self):
        """(classify_waves, zseg, emit_workgroups) in use."""
        out = (ctypes.c_int32 * 3)()
        _check(lib().HPMCgetTuning(self.h, out), self.h)
        return tuple(out)

    def set_profiling(self, enable=True):
        _check(lib().HPMCsetProfiling(self.h, int(bool(enable))), self.h)

    def stage_times(self):
        """{stage: (total ms, launches)} since the last call (waits for the device)."""
        ms = (ctypes.c_double * 3)()
        n = (_u32 * 3)()
        _check(lib().HPMCgetStageTimes(self.h, ms, n), self.h)
        return {name: (ms[i], n[i]) for i, name in enumerate(("classify", "scan", "emit"))}

    def close(self):
        L = lib()
        if self.th:
            L.HPMCdestroyTraversalHandle(self.th)
            self.th = None
        if self.h:
            L.HPMCdestroyHandle(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many FP32 volumes built and extracted together (HPMCBatch*): one
    classification, one scan and one extraction launch over all volumes,
    optionally replayed as a hipGraph.  Per-volume results equal those of
    IsoSurface on each volume alone; the output concatenates them in the
    order the volumes were added."""

    _constants = None

    def __init__(self, stream=None):
        L = lib()
        if IsoSurface._constants is None:
            c = L.HPMCcreateConstants(4, 3)
            if not c:
                raise HPMCError("HPMCcreateConstants failed: %s" % (last_error()[1],))
            IsoSurface._constants = c
        self.b = L.HPMCcreateBatch(IsoSurface._constants)
        if not self.b:
            raise HPMCError("HPMCcreateBatch failed: %s" % (last_error()[1],))
        self._volumes = []
        if stream is not None:
            _check(L.HPMCbatchSetStream(self.b, _stream_ptr(stream)))

    def add_volume(self, volume, extent=(1.0, 1.0, 1.0)):
        """volume: contiguous float32 device tensor (z, y, x); returns its index."""
        import torch
        assert volume.dtype == torch.float32 and volume.is_contiguous() and volume.is_cuda
        z, y, x = volume.shape
        i = lib().HPMCbatchAddVolume(self.b, ctypes.c_void_p(volume.data_ptr()), x, y, z, *[float(e) for e in extent])
        if i < 0:
            _check(-i)
        self._volumes.append(volume)
        return i

    def __len__(self):
        return len(self._volumes)

    def set_tuning(self, classify_waves=0, zseg=0, emit_workgroups=0):
        _check(lib().HPMCbatchSetTuning(self.b, classify_waves, zseg, emit_workgroups))

    def tuning(self):
        out = (_i32 * 3)()
        _check(lib().HPMCbatchGetTuning(self.b, out))
        return tuple(out)

    def build(self, iso):
        _check(lib().HPMCbatchBuild(self.b, float(iso)))

    def counts(self):
        """(counts[n], bases[n], total) of the last build (waits)."""
        import numpy as np
        n = len(self._volumes)
        counts = np.zeros(n, np.uint32)
        bases = np.zeros(n, np.uint32)
        total = _u64(0)
        _check(lib().HPMCbatchGetCounts(self.b, counts.ctypes.data_as(ctypes.POINTER(_u32)),
                                        bases.ctypes.data_as(ctypes.POINTER(_u32)), ctypes.byref(total)))
        return counts, bases, int(total.value)

    def extract(self, out, cell_edge=None):
        """out: float32 (capacity, 6) device tensor; cell_edge: int32 (capacity, 2) or None."""
        ce = ctypes.c_void_p(cell_edge.data_ptr()) if cell_edge is not None else None
        cap = out.shape[0] if cell_edge is None else min(out.shape[0], cell_edge.shape[0])
        _check(lib().HPMCbatchExtract(self.b, ctypes.c_void_p(out.data_ptr()), cap, ce))

    def capture_graph(self, iso, out, cell_edge=None):
        ce = ctypes.c_void_p(cell_edge.data_ptr()) if cell_edge is not None else None
        cap = out.shape[0] if cell_edge is None else min(out.shape[0], cell_edge.shape[0])
        _check(lib().HPMCbatchCaptureGraph(self.b, float(iso), ctypes.c_void_p(out.data_ptr()), cap, ce))

    def launch_graph(self):
        _check(lib().HPMCbatchLaunchGraph(self.b))

    def adjoint(self, grad_vertices):
        """(dL/dfield per volume, dL/dT) of the last build + extraction from
        dL/dvertices (total, 6) of the concatenated output; deterministic."""
        import torch
        dev = self._volumes[0].device
        gv = _vertex_grads(grad_vertices, self.counts()[2], dev)
        n_lat = sum(v.numel() for v in self._volumes)
        gf = torch.empty(n_lat, dtype=torch.float32, device=dev)
        gt = torch.empty((), dtype=torch.float32, device=dev)
        _check(lib().HPMCbatchAdjoint(self.b, ctypes.c_void_p(gv.data_ptr()) if gv is not None else None,
                                      ctypes.c_void_p(gf.data_ptr()), ctypes.c_void_p(gt.data_ptr())))
        grads = []
        o = 0
        for v in self._volumes:
            grads.append(gf[o:o + v.numel()].view(v.shape))
            o += v.numel()
        return grads, gt

    def autotune(self, iso, out, repetitions=3):
        """Picks the fastest launch shapes for build + extract; returns (shape, ms)."""
        best = ctypes.c_double(0.0)
        _check(lib().HPMCbatchAutotune(self.b, float(iso), ctypes.c_void_p(out.data_ptr()), out.shape[0],
                                       repetitions, ctypes.byref(best)))
        return self.tuning(), best.value

    def close(self):
        if self.b:
            lib().HPMCdestroyBatch(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
