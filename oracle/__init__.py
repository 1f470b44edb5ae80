"""CPU oracle for HPMC's marching-cubes path (test infrastructure only).

Only the tests, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
use this package, and only as the checker.  It wraps ``liboracle.so`` (the C
restatement in ``hpmc_oracle.c``, built by ``make -C oracle``) and loads the
marching-cubes cases from ``tests/golden/mc_cases.json``.

Pinning (each oracle function and what fixes it independently of the oracle's
own formulas; the tests that do it are named in brackets):
  layout          known answers for 64^3/256^3/512^3 and the closed form
                  (tests/test_oracle.py::test_layout_*)
  classify        pure-Python brute force, plane and constant fields
                  (test_classify_*)
  hp_traverse     equals the Morton closed form (vertex_list) and a key round
                  trip (test_hp_*)
  vertex_list     equals hp_traverse; level-4 sums equal block sums
  emit_f64        linear fields (exact t and positions), sphere distance
                  (test_emit_*)
  cases           fixture from hpmc/src/tables.cpp plus per-case topology
                  invariants (tests/test_mc_table.py)
  binary_normals  the emitted normal equals |n| n for the face normal n of
                  the triangle through the emitted midpoints, up to the
                  table's clamp and FP32 rounding (test_binary_*)
  emit_binary     midpoints are exact; normals as above
  emit_gradient   a linear field with its exact gradient gives exact t,
                  positions and normals T - grad (test_emit_gradient_*)
  adjoint         central finite differences (h = 1e-3) of emit_f64_field
                  on samples and threshold away from sign changes
                  (test_adjoint_*)
"""
import ctypes
import json
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
_LIB = None


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def load_cases():
    """(counts[256], slots[256,16,4]) by HPMC code from the golden fixture."""
    doc = json.load(open(os.path.join(ROOT, "tests", "golden", "mc_cases.json")))
    counts = np.zeros(256, np.int32)
    slots = np.zeros((256, 16, 4), np.int32)
    for code, c in enumerate(doc["cases"]):
        counts[code] = c["count"]
        for i, s in enumerate(c["slots"]):
            slots[code, i] = s
    return counts, slots


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        vp, i64 = ctypes.c_void_p, ctypes.c_int64
        L.hpo_layout.restype = ctypes.c_int
        L.hpo_layout.argtypes = [vp, vp, vp]
        L.hpo_classify.argtypes = [vp, vp, vp, vp, ctypes.c_float, vp, vp]
        L.hpo_hp_traverse.restype = i64
        L.hpo_hp_traverse.argtypes = [vp, vp, vp, vp, vp, i64]
        L.hpo_vertex_list.restype = i64
        L.hpo_vertex_list.argtypes = [vp, vp, vp, vp, vp, i64]
        L.hpo_level4.argtypes = [vp, vp, vp, vp]
        L.hpo_level_blocks.argtypes = [vp, vp, vp, ctypes.c_int, vp]
        L.hpo_emit_f64.argtypes = [vp, vp, vp, ctypes.c_int, vp, vp, ctypes.c_float, vp, vp, vp, i64, vp, vp, vp]
        L.hpo_emit_f32.argtypes = [vp, vp, vp, ctypes.c_int, vp, vp, ctypes.c_float, vp, vp, vp, i64, vp]
        L.hpo_run_volume.restype = i64
        L.hpo_run_volume.argtypes = [vp, vp, vp, vp, ctypes.c_float, ctypes.c_int32, ctypes.c_int32, vp, i64]
        L.hpo_emit_f64d.argtypes = [vp, vp, vp, ctypes.c_int, vp, vp, ctypes.c_double, vp, vp, vp, i64, vp]
        L.hpo_adjoint_f64.restype = ctypes.c_double
        L.hpo_adjoint_f64.argtypes = [vp, vp, vp, ctypes.c_int, vp, vp, ctypes.c_double, vp, vp, vp, i64, vp, vp]
        L.hpo_binary_normals.argtypes = [vp]
        L.hpo_emit_binary_f64.argtypes = [vp, vp, vp, vp, vp, i64, vp]
        L.hpo_emit_grad_f64.argtypes = [vp, vp, vp, ctypes.c_int, vp, vp, ctypes.c_float, vp, vp, vp, i64, vp, vp,
                                        vp]
        L.hpo_morton2.restype = ctypes.c_uint64
        L.hpo_morton2.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        counts, slots = load_cases()
        L.hpo_set_cases(_ptr(counts), _ptr(np.ascontiguousarray(slots)))
        _LIB = L
    return _LIB


CASE_COUNTS, CASE_SLOTS = load_cases()


def layout(size, cells):
    """(tile_x, tile_y, tiles_x, tiles_y, size_l2) or None (init.cpp:72-149)."""
    out = np.zeros(5, np.int32)
    ok = lib().hpo_layout(_ptr(_i32(size)), _ptr(_i32(cells)), _ptr(out))
    return tuple(int(v) for v in out) if ok else None


def classify(field, T, cells=None, size=None):
    """Per-cell (codes, counts) as uint8 arrays shaped (cz, cy, cx).

    ``field`` is (fz, fy, fx) float32, x fastest; it may extend past the
    lattice ``size`` (x, y, z), which defaults to its own shape."""
    field = np.ascontiguousarray(field, np.float32)
    dims = np.array(field.shape[::-1], np.int32)
    size = _i32(size if size is not None else dims)
    cells = _i32(cells if cells is not None else size - 1)
    codes = np.zeros(tuple(cells[::-1]), np.uint8)
    counts = np.zeros_like(codes)
    lib().hpo_classify(_ptr(field), _ptr(dims), _ptr(size), _ptr(cells), float(T), _ptr(codes), _ptr(counts))
    return codes, counts


def vertex_list(counts, cells, lay, literal=False):
    """(cell_index[N], slot[N]) in HPMC's extraction order.

    literal=True walks an explicit HistoPyramid (small lattices only);
    otherwise the Morton closed form is used."""
    counts = np.ascontiguousarray(counts, np.uint8)
    total = int(counts.sum(dtype=np.int64))
    cell = np.zeros(max(total, 1), np.uint32)
    slot = np.zeros(max(total, 1), np.uint8)
    fn = lib().hpo_hp_traverse if literal else lib().hpo_vertex_list
    n = fn(_ptr(counts), _ptr(_i32(cells)), _ptr(_i32(lay)), _ptr(cell), _ptr(slot), total)
    if n < 0:
        raise MemoryError("oracle vertex list failed")
    return cell[:n], slot[:n]


def block_log2(lay):
    """log2 of the block side the HIP build sums: 4 (16x16 cells, HP level
    4), or the smaller log2 of a doubled tile side (grids of at most 8 cells
    along x or y), so that no block straddles two tiles."""
    ts = [int(lay[0]) * 2, int(lay[1]) * 2]
    return min(4, *(int(t).bit_length() - 1 for t in ts))


def level4(counts, cells, lay, b=None):
    """Block sums of the base image (2^b x 2^b cells of the doubled image, b
    = block_log2(lay) by default: HP level 4 for all but the narrowest
    grids), indexed by Morton block code."""
    b = block_log2(lay) if b is None else b
    side = max((2 << lay[4]) >> b, 1)
    out = np.zeros(side * side, np.uint32)
    lib().hpo_level_blocks(_ptr(np.ascontiguousarray(counts, np.uint8)), _ptr(_i32(cells)), _ptr(_i32(lay)), int(b),
                           _ptr(out))
    return out


def emit(field, codes, cell, slot, T, ext=(1.0, 1.0, 1.0), cells=None, size=None, clamp=True, fp64=True,
         with_mag=False):
    """N3F_V3F vertices (N, 6) for the given list; fp64 also returns t (and,
    with_mag, the largest |T|, |f| among the samples each normal uses)."""
    field = np.ascontiguousarray(field, np.float32)
    dims = np.array(field.shape[::-1], np.int32)
    size = _i32(size if size is not None else dims)
    cells = _i32(cells if cells is not None else size - 1)
    codes = np.ascontiguousarray(codes, np.uint8)
    cell = np.ascontiguousarray(cell, np.uint32)
    slot = np.ascontiguousarray(slot, np.uint8)
    ext = np.ascontiguousarray(ext, np.float32)
    n = len(cell)
    if fp64:
        out = np.zeros((max(n, 1), 6), np.float64)
        t = np.zeros(max(n, 1), np.float64)
        mag = np.zeros(max(n, 1), np.float64)
        lib().hpo_emit_f64(_ptr(field), _ptr(dims), _ptr(size), int(clamp), _ptr(cells), _ptr(ext), float(T),
                           _ptr(codes), _ptr(cell), _ptr(slot), n, _ptr(out), _ptr(t), _ptr(mag))
        if with_mag:
            return out[:n], t[:n], mag[:n]
        return out[:n], t[:n]
    out = np.zeros((max(n, 1), 6), np.float32)
    lib().hpo_emit_f32(_ptr(field), _ptr(dims), _ptr(size), int(clamp), _ptr(cells), _ptr(ext), float(T),
                       _ptr(codes), _ptr(cell), _ptr(slot), n, _ptr(out))
    return out[:n]


def emit_f64_field(field64, codes, cell, slot, T, ext=(1.0, 1.0, 1.0), cells=None, size=None, clamp=True):
    """emit() on an FP64 field and FP64 threshold (finite-difference checks)."""
    field64 = np.ascontiguousarray(field64, np.float64)
    dims = np.array(field64.shape[::-1], np.int32)
    size = _i32(size if size is not None else dims)
    cells = _i32(cells if cells is not None else size - 1)
    codes = np.ascontiguousarray(codes, np.uint8)
    cell = np.ascontiguousarray(cell, np.uint32)
    slot = np.ascontiguousarray(slot, np.uint8)
    n = len(cell)
    out = np.zeros((max(n, 1), 6), np.float64)
    lib().hpo_emit_f64d(_ptr(field64), _ptr(dims), _ptr(size), int(clamp), _ptr(cells),
                        _ptr(np.ascontiguousarray(ext, np.float32)), float(T), _ptr(codes), _ptr(cell), _ptr(slot),
                        n, _ptr(out))
    return out[:n]


def adjoint(field, codes, cell, slot, T, gout, ext=(1.0, 1.0, 1.0), cells=None, size=None, clamp=True):
    """FP64 adjoint of emit(): (dL/dfield shaped like field, dL/dT) for
    dL/dvertices gout (N, 6)."""
    field64 = np.ascontiguousarray(field, np.float64)
    dims = np.array(field64.shape[::-1], np.int32)
    size = _i32(size if size is not None else dims)
    cells = _i32(cells if cells is not None else size - 1)
    gout = np.ascontiguousarray(gout, np.float64)
    codes = np.ascontiguousarray(codes, np.uint8)
    cell = np.ascontiguousarray(cell, np.uint32)
    slot = np.ascontiguousarray(slot, np.uint8)
    dfield = np.zeros_like(field64)
    dT = lib().hpo_adjoint_f64(_ptr(field64), _ptr(dims), _ptr(size), int(clamp), _ptr(cells),
                               _ptr(np.ascontiguousarray(ext, np.float32)), float(T), _ptr(codes), _ptr(cell),
                               _ptr(slot), len(cell), _ptr(gout), _ptr(dfield))
    return dfield, float(dT)


def binary_normals():
    """(256, 16, 3) float32: the normal a binary field's vertex carries, by
    HPMC code and slot (constants.cpp:217-273 decoded as shadergen.cpp:507-510)."""
    out = np.zeros((256, 16, 3), np.float32)
    lib().hpo_binary_normals(_ptr(out))
    return out


def emit_binary(codes, cell, slot, cells, ext=(1.0, 1.0, 1.0)):
    """N3F_V3F (N, 6) float64 of a binary field: edge midpoints and tabulated normals."""
    codes = np.ascontiguousarray(codes, np.uint8)
    cell = np.ascontiguousarray(cell, np.uint32)
    slot = np.ascontiguousarray(slot, np.uint8)
    n = len(cell)
    out = np.zeros((max(n, 1), 6), np.float64)
    lib().hpo_emit_binary_f64(_ptr(_i32(cells)), _ptr(np.ascontiguousarray(ext, np.float32)), _ptr(codes),
                              _ptr(cell), _ptr(slot), n, _ptr(out))
    return out[:n]


def emit_gradient(field4, codes, cell, slot, T, ext=(1.0, 1.0, 1.0), cells=None, size=None, clamp=True):
    """Gradient-field emission: field4 is (fz, fy, fx, 4) float32 (gradient,
    value); returns (out (N, 6) float64, t, mag)."""
    field4 = np.ascontiguousarray(field4, np.float32)
    dims = np.array(field4.shape[:3][::-1], np.int32)
    size = _i32(size if size is not None else dims)
    cells = _i32(cells if cells is not None else size - 1)
    codes = np.ascontiguousarray(codes, np.uint8)
    cell = np.ascontiguousarray(cell, np.uint32)
    slot = np.ascontiguousarray(slot, np.uint8)
    n = len(cell)
    out = np.zeros((max(n, 1), 6), np.float64)
    t = np.zeros(max(n, 1), np.float64)
    mag = np.zeros(max(n, 1), np.float64)
    lib().hpo_emit_grad_f64(_ptr(field4), _ptr(dims), _ptr(size), int(clamp), _ptr(cells),
                            _ptr(np.ascontiguousarray(ext, np.float32)), float(T), _ptr(codes), _ptr(cell),
                            _ptr(slot), n, _ptr(out), _ptr(t), _ptr(mag))
    return out[:n], t[:n], mag[:n]


def run_volume(field, T, ext=(1.0, 1.0, 1.0), z_begin=0, z_end=None, emit_out=True):
    """Whole CPU path over cell slices [z_begin, z_end); returns the vertex count."""
    field = np.ascontiguousarray(field, np.float32)
    size = np.array(field.shape[::-1], np.int32)
    cells = size - 1
    z_end = int(cells[2]) if z_end is None else z_end
    cap = 15 * int(cells[0]) * int(cells[1]) * max(z_end - z_begin, 0) if emit_out else 0
    out = np.zeros((max(min(cap, 1 << 26), 1), 6), np.float32) if emit_out else None
    cap = out.shape[0] if emit_out else 0
    n = lib().hpo_run_volume(_ptr(field), _ptr(size), _ptr(_i32(cells)), _ptr(np.asarray(ext, np.float32)),
                             float(T), int(z_begin), int(z_end), _ptr(out) if emit_out else None, cap)
    return int(n)


def edge_ids(codes, cell, slot):
    """This library's edge id (4*axis + s0 + 2*s1) for each listed vertex."""
    s = CASE_SLOTS[np.asarray(codes).reshape(-1)[cell], slot]
    axis = s[:, 3]
    sx, sy, sz = s[:, 0], s[:, 1], s[:, 2]
    s0 = np.where(axis == 0, sy, sx)
    s1 = np.where(axis == 2, sy, sz)
    return (4 * axis + s0 + 2 * s1).astype(np.uint8)
