"""Pins the CPU oracle (oracle/) independently of its own formulas.

layout      known answers (64^3, 256^3, 512^3) and the closed form of
            hpmc/src/init.cpp:105-130 recomputed in FP64 for power-of-two
            friendly sizes
classify    pure-Python per-cell brute force, constant and plane fields
traverse    literal HistoPyramid walk == Morton closed form, key round trip,
            hand-worked order on a 2-slice example
emit        linear fields give exact t and positions, a sphere's distance
            function gives |p - c| ~ r
"""
import itertools

import numpy as np
import pytest

import oracle
from hpmc_amd import workloads


# ---------------------------------------------------------------- layout ---
@pytest.mark.parametrize("n,expect", [
    (64, (32, 32, 8, 8, 8)),
    (256, (128, 128, 16, 16, 11)),
    (512, (256, 256, 32, 32, 13)),
])
def test_layout_known_answers(n, expect):
    assert oracle.layout((n, n, n), (n - 1, n - 1, n - 1)) == expect


def test_layout_closed_form():
    rng = np.random.default_rng(3)
    for _ in range(300):
        cells = [int(v) for v in rng.integers(2, 700, size=3)]
        size = [c + int(rng.integers(1, 3)) for c in cells]
        lay = oracle.layout(size, cells)
        ts = [1 << int(np.ceil(np.log2(c / 2.0))) for c in cells[:2]]
        tiles_x = 1 << max(0, int(np.ceil(np.log2(np.sqrt(cells[2] / (ts[0] / ts[1]))))))
        tiles_y = -(-cells[2] // tiles_x)
        S = 1 << int(np.ceil(np.log2(max(ts[0] * tiles_x, ts[1] * tiles_y))))
        assert lay == (ts[0], ts[1], S // ts[0], S // ts[1], int(np.log2(S))), (cells, lay)
        # every slice fits the base image
        assert lay[2] * lay[3] >= cells[2]


def test_layout_rejects_bad_sizes():
    assert oracle.layout((1, 8, 8), (0, 7, 7)) is None
    assert oracle.layout((16385, 8, 8), (16384, 7, 7)) is None
    assert oracle.layout((8, 8, 8), (8, 7, 7)) is None


# -------------------------------------------------------------- classify ---
def _brute_codes(f, cells, T):
    cz, cy, cx = cells[2], cells[1], cells[0]
    codes = np.zeros((cz, cy, cx), np.uint8)
    for z, y, x in itertools.product(range(cz), range(cy), range(cx)):
        c = 0
        for dz, dy, dx in itertools.product((0, 1), (0, 1), (0, 1)):
            if f[z + dz, y + dy, x + dx] < T:
                c |= 1 << (dx + 2 * dy + 4 * dz)
        codes[z, y, x] = c
    return codes


@pytest.mark.parametrize("shape,cells", [
    ((5, 6, 7), None), ((9, 4, 3), None), ((8, 8, 8), (5, 6, 4)), ((3, 3, 17), None),
])
def test_classify_brute_force(shape, cells):
    rng = np.random.default_rng(sum(shape))
    f = rng.random((shape[2], shape[1], shape[0])).astype(np.float32)
    cells = cells or (shape[0] - 1, shape[1] - 1, shape[2] - 1)
    codes, counts = oracle.classify(f, 0.5, cells=cells)
    assert np.array_equal(codes, _brute_codes(f, cells, 0.5))
    assert np.array_equal(counts, oracle.CASE_COUNTS[codes])


def test_classify_constant_and_plane():
    f = np.full((6, 6, 6), 2.0, np.float32)
    codes, counts = oracle.classify(f, 1.0)
    assert codes.max() == 0 and counts.sum() == 0
    codes, counts = oracle.classify(f, 3.0)
    assert (codes == 255).all() and counts.sum() == 0
    # plane f = x - 2.5: corners with x <= 2 are inside (bits x=0), cells at
    # x = 2 are crossed: code = bits where corner x == 0 -> 0b01010101
    x = np.arange(7, dtype=np.float32)
    f = np.broadcast_to(x[None, None, :] - 2.5, (5, 6, 7)).copy()
    codes, counts = oracle.classify(f, 0.0)
    assert (codes[:, :, 2] == 0x55).all()
    assert (codes[:, :, :2] == 255).all() and (codes[:, :, 3:] == 0).all()
    assert (counts[:, :, 2] == 6).all()   # a quad cut: two triangles


# -------------------------------------------------------------- traverse ---
@pytest.mark.parametrize("shape", [(12, 9, 7), (17, 17, 17), (10, 20, 6), (33, 5, 9), (9, 9, 40)])
def test_hp_literal_equals_morton_closed_form(shape):
    rng = np.random.default_rng(shape[0] * 7 + shape[2])
    f = rng.random((shape[2], shape[1], shape[0])).astype(np.float32)
    cells = (shape[0] - 1, shape[1] - 1, shape[2] - 1)
    codes, counts = oracle.classify(f, 0.5)
    lay = oracle.layout(shape, cells)
    c1, s1 = oracle.vertex_list(counts, cells, lay, literal=True)
    c2, s2 = oracle.vertex_list(counts, cells, lay, literal=False)
    assert len(c1) == counts.sum() and np.array_equal(c1, c2) and np.array_equal(s1, s2)
    # key round trip: each (cell, slot) appears once and slots run 0..count-1
    pairs = list(zip(c1.tolist(), s1.tolist()))
    assert len(set(pairs)) == len(pairs)
    flat = counts.reshape(-1)
    for c, s in pairs:
        assert s < flat[c]


def test_traversal_order_hand_worked():
    # 3x3x3 lattice: cells 2x2x2, tile 1x1 texels (2x2 cells), 2 slices laid
    # out side by side along x (tiles_x = 2).  Doubled image positions:
    # slice 0 cells at X in {0,1}, slice 1 at X in {2,3}; Morton (x low)
    # order is (0,0),(1,0),(0,1),(1,1) of slice 0, then the same of slice 1.
    cells = (2, 2, 2)
    lay = oracle.layout((3, 3, 3), cells)
    assert lay == (1, 1, 2, 2, 1)
    counts = np.full((2, 2, 2), 3, np.uint8)
    c, s = oracle.vertex_list(counts, cells, lay, literal=True)
    order = [c[i] for i in range(0, len(c), 3)]
    assert order == [0, 1, 2, 3, 4, 5, 6, 7]
    assert s.tolist() == [0, 1, 2] * 8


def test_level4_sums():
    f = workloads.cayley((40, 37, 35))
    cells = (39, 36, 34)
    codes, counts = oracle.classify(f, 0.0)
    lay = oracle.layout((40, 37, 35), cells)
    l4 = oracle.level4(counts, cells, lay)
    assert l4.sum() == counts.sum()
    cell, _ = oracle.vertex_list(counts, cells, lay)
    # level-4 blocks appear in the list as contiguous runs in Morton order
    nz = np.nonzero(l4)[0]
    runs = np.cumsum(l4[nz])
    assert runs[-1] == len(cell)


def test_level4_morton_keys_hand_worked():
    # 33x17x6 lattice: cells 32x16x5, tiles of 16x8 texels (32x16 cells of
    # the doubled image), two tiles per row: slice z starts at doubled
    # (32 (z mod 2), 16 (z div 2)).  The 64x64 doubled image has 4x4 blocks
    # of 16x16 cells; block (X, Y) has Morton key x-bit-low: (0,0)=0,
    # (1,0)=1, (0,1)=2, (1,1)=3, (2,0)=4, (3,0)=5, (2,1)=6, (3,1)=7,
    # (0,2)=8, (1,2)=9.
    cells = (32, 16, 5)
    lay = oracle.layout((33, 17, 6), cells)
    assert lay == (16, 8, 2, 4, 5)
    counts = np.zeros((5, 16, 32), np.uint8)
    counts[:, 3, 20] = np.arange(1, 6)      # cell x = 20 lies in block column 1 of its tile
    counts[4, 9, 2] = 7                     # slice 4, block column 0
    l4 = oracle.level4(counts, cells, lay)
    expect = np.zeros(16, np.uint32)
    # slice: 0 -> block (1,0), 1 -> (3,0), 2 -> (1,1), 3 -> (3,1), 4 -> (1,2) and (0,2)
    expect[[1, 5, 3, 7, 9]] = [1, 2, 3, 4, 5]
    expect[8] = 7
    assert l4.tolist() == expect.tolist()


# ------------------------------------------------------------------ emit ---
def _list(f, T, ext=(1.0, 1.0, 1.0)):
    shape = f.shape[::-1]
    cells = tuple(v - 1 for v in shape)
    codes, counts = oracle.classify(f, T)
    lay = oracle.layout(shape, cells)
    cell, slot = oracle.vertex_list(counts, cells, lay)
    return codes, cell, slot, cells


def test_emit_linear_field_exact():
    # f = x + 0.25 y (lattice units), T = 3.3: crossings on x-edges at
    # x = 3.3 - 0.25 y exactly; positions p = ext * (k + t*axis) / cells
    X, Y, Z = 9, 5, 4
    x = np.arange(X, dtype=np.float64)
    y = np.arange(Y, dtype=np.float64)
    f = (x[None, None, :] + 0.25 * y[None, :, None] + 0 * np.zeros((Z, 1, 1))).astype(np.float32)
    codes, cell, slot, cells = _list(f, 3.3, ext=(2.0, 1.0, 1.0))
    out, t = oracle.emit(f, codes, cell, slot, 3.3, ext=(2.0, 1.0, 1.0))
    assert len(out) > 0
    assert ((t >= 0) & (t <= 1)).all()
    px = out[:, 3] * cells[0] / 2.0
    py = out[:, 4] * cells[1]
    vals = px + 0.25 * py
    np.testing.assert_allclose(vals, 3.3, atol=1e-6)


def _corners(codes, cell, slot, cells):
    """First corner a (lattice index) and axis of each listed vertex."""
    flat = np.asarray(codes).reshape(-1)
    s = oracle.CASE_SLOTS[flat[cell], slot]
    x = cell % cells[0]
    y = (cell // cells[0]) % cells[1]
    z = cell // (cells[0] * cells[1])
    a = np.stack([x + s[:, 0], y + s[:, 1], z + s[:, 2]], axis=1).astype(np.int64)
    return a, s[:, 3]


def test_emit_linear_field_normals_closed_form():
    # f = 0.5 x - 0.25 y + 0.125 z (lattice units, exact in FP32), T = 1.3,
    # extent (1, 2, 3).  Forward samples of a linear field are f(a) + g_q and
    # f(b) + g_q, so mix(na, nb, t) = mix(va, vb, t) + g = T + g and the
    # normal is (T - mix(na, nb, t)) ext/cells = -g ext/cells exactly, for
    # vertices whose six forward samples lie inside the lattice.  Dropping
    # the T offset or inverting the per-axis scale fails this.
    X, Y, Z = 11, 9, 8
    T, ext = 1.3, (1.0, 2.0, 3.0)
    g = np.array([0.5, -0.25, 0.125])
    k = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    f = (g[0] * k[2] + g[1] * k[1] + g[2] * k[0]).astype(np.float32)
    codes, cell, slot, cells = _list(f, T, ext=ext)
    a, axis = _corners(codes, cell, slot, cells)
    b = a + np.eye(3, dtype=np.int64)[axis]
    interior = ((b + 1) <= np.array([X - 1, Y - 1, Z - 1])).all(axis=1)
    assert interior.sum() > 100
    sc = np.array(ext) / np.array(cells)
    want = -g * sc
    out, t = oracle.emit(f, codes, cell, slot, T, ext=ext)
    np.testing.assert_allclose(out[interior, :3], np.broadcast_to(want, (interior.sum(), 3)), rtol=0, atol=1e-12)
    # positions: a + t axis, scaled; t exact from the linear field
    pos = (a + t[:, None] * np.eye(3)[axis]) * sc
    np.testing.assert_allclose(out[:, 3:], pos, rtol=0, atol=1e-12)
    val = (out[:, 3:] / sc) @ g
    np.testing.assert_allclose(val, float(np.float32(T)), rtol=0, atol=1e-12)   # T goes in as FP32
    # FP32 emission: the same within FP32 rounding of terms of size |f| <= 6
    o32 = oracle.emit(f, codes, cell, slot, T, ext=ext, fp64=False)
    np.testing.assert_allclose(o32[interior, :3], np.broadcast_to(want, (interior.sum(), 3)), rtol=0,
                               atol=8 * 6.0 * 2.0 ** -24 * sc.max())


def test_emit_sphere_distance():
    n = 48
    p = (np.arange(n) + 0.5) / n
    d = np.sqrt((p[None, None, :] - 0.5) ** 2 + (p[None, :, None] - 0.5) ** 2 + (p[:, None, None] - 0.5) ** 2)
    f = (d - 0.3).astype(np.float32)
    codes, cell, slot, cells = _list(f, 0.0)
    out, t = oracle.emit(f, codes, cell, slot, 0.0)
    # positions are in [0, ext] grid coordinates: lattice point k sits at
    # k / cells, the sphere centre (texel coordinate 0.5) at (0.5 n - 0.5) / cells
    c = (0.5 * n - 0.5) / cells[0]
    r = np.linalg.norm(out[:, 3:] - c, axis=1) * cells[0] / n
    np.testing.assert_allclose(r, 0.3, atol=2e-3)
    # forward-difference normals point along the outward radius, as
    # T - f(p + e) = -(grad f) h
    radial = (out[:, 3:] - c) / np.linalg.norm(out[:, 3:] - c, axis=1)[:, None]
    nn = out[:, :3] / np.linalg.norm(out[:, :3], axis=1)[:, None]
    assert (np.einsum("ij,ij->i", nn, radial) < -0.9).all()


def test_emit_f32_matches_f64():
    f = workloads.cayley((33, 31, 29))
    codes, cell, slot, cells = _list(f, 0.0)
    o64, _ = oracle.emit(f, codes, cell, slot, 0.0)
    o32 = oracle.emit(f, codes, cell, slot, 0.0, fp64=False)
    np.testing.assert_allclose(o32, o64, rtol=1e-5, atol=1e-6)


def test_run_volume_counts():
    f = workloads.cayley((40, 40, 40))
    codes, counts = oracle.classify(f, 0.0)
    assert oracle.run_volume(f, 0.0) == counts.sum()


# --------------------------------------------------------- binary fields ---
def _midpoint(s):
    m = np.array(s[:3], np.float64)
    m[s[3]] += 0.5
    return m


def test_binary_normal_known_answer():
    # code 1: only corner (0,0,0) is inside; one triangle through the three
    # midpoints next to it, face normal of length sqrt(3)/4 along the body
    # diagonal, emitted as |n| n = -(sqrt(3)/16) (1,1,1), toward the corner
    bn = oracle.binary_normals()
    assert oracle.CASE_COUNTS[1] == 3
    np.testing.assert_allclose(bn[1, :3], -np.sqrt(3.0) / 16.0, rtol=0, atol=2e-7)
    # its complement: every corner but (0,0,0) inside, the normal flips
    np.testing.assert_allclose(bn[254, :3], np.sqrt(3.0) / 16.0, rtol=0, atol=2e-7)


def test_binary_normals_are_scaled_face_normals_toward_inside():
    """For all 256 cases: each slot's normal is |nf| nf for the face normal nf
    of its triangle (geometry of the edge midpoints) oriented toward the
    inside corners (f < T), like the continuous forward-difference normals;
    components with |nf| |nf_k| >= 1 saturate at the table's clamp."""
    bn = oracle.binary_normals().astype(np.float64)
    checked = 0
    for code in range(256):
        inside = lambda p: (code >> (int(p[0]) + 2 * int(p[1]) + 4 * int(p[2]))) & 1
        for t in range(oracle.CASE_COUNTS[code] // 3):
            tri = oracle.CASE_SLOTS[code, 3 * t:3 * t + 3]
            m = [_midpoint(s) for s in tri]
            nf = np.cross(m[1] - m[0], m[2] - m[0])
            score = 0.0
            for s in tri:
                a = np.array(s[:3], np.float64)
                b = a.copy()
                b[s[3]] += 1
                if not inside(a):
                    a, b = b, a
                score += float(np.dot(nf, a - b))
            assert abs(score) > 1e-9
            if score < 0:
                nf = -nf
            want = np.linalg.norm(nf) * nf
            sat = np.abs(want) >= 1.0
            want_c = np.where(sat, np.sign(want) * (1.0 - 2e-6), want)
            for i in range(3):
                got = bn[code, 3 * t + i]
                assert (np.abs(got - want_c) <= np.where(sat, 3e-6, 3e-7)).all(), (code, t, i, got, want_c)
                checked += 1
    assert checked == int(oracle.CASE_COUNTS.sum())


def test_emit_binary_midpoints():
    f = (workloads.cayley((21, 19, 17)) < 0.25).astype(np.float32)
    codes, cell, slot, cells = _list(f, 0.5)
    out = oracle.emit_binary(codes, cell, slot, cells, ext=(2.0, 1.0, 0.5))
    assert len(out) > 0
    g = out[:, 3:] / (np.array([2.0, 1.0, 0.5]) / np.array(cells))
    # every coordinate is a multiple of half a cell, and exactly one (the
    # edge axis) sits halfway between lattice points
    h = np.round(2.0 * g)
    assert (np.abs(2.0 * g - h) < 1e-9).all()
    assert ((h.astype(np.int64) % 2).sum(axis=1) == 1).all()


# ------------------------------------------------------- gradient fields ---
def test_emit_gradient_linear_field_exact():
    # f = 0.5 x - 0.25 y + 0.125 z (lattice units) with its exact gradient:
    # t, positions and normals are exact, n = (T - grad) * ext/cells
    X, Y, Z = 11, 9, 7
    k = np.arange(Z)[:, None, None]
    j = np.arange(Y)[None, :, None]
    i = np.arange(X)[None, None, :]
    val = (0.5 * i - 0.25 * j + 0.125 * k).astype(np.float32)
    f4 = np.zeros((Z, Y, X, 4), np.float32)
    f4[..., 0], f4[..., 1], f4[..., 2], f4[..., 3] = 0.5, -0.25, 0.125, val
    T = 1.3
    codes, cell, slot, cells = _list(val, T)
    ext = (1.0, 2.0, 3.0)
    out, t, _ = oracle.emit_gradient(f4, codes, cell, slot, T, ext=ext)
    assert len(out) > 0 and ((t >= 0) & (t <= 1)).all()
    sc = np.array(ext) / np.array(cells)
    p = out[:, 3:] / sc
    np.testing.assert_allclose(0.5 * p[:, 0] - 0.25 * p[:, 1] + 0.125 * p[:, 2], np.float32(T), atol=1e-6)
    np.testing.assert_allclose(out[:, :3], np.broadcast_to((np.float32(T) - np.array([0.5, -0.25, 0.125])) * sc, out[:, :3].shape), atol=1e-12)


# --------------------------------------------------------------- adjoint ---
def _sphere_noise(n, seed, sigma=0.01):
    p = (np.arange(n) + 0.5) / n
    d = np.sqrt((p[None, None, :] - 0.5) ** 2 + (p[None, :, None] - 0.5) ** 2 + (p[:, None, None] - 0.5) ** 2)
    rng = np.random.default_rng(seed)
    return (d - 0.3 + sigma * rng.standard_normal(d.shape)).astype(np.float32)


@pytest.mark.parametrize("clamp_edge", [False, True])
def test_adjoint_matches_finite_differences(clamp_edge):
    """dL/dfield and dL/dT of L = sum(gout * emit(f, T)) against central
    differences (h = 1e-3) of the FP64 forward, on samples whose perturbation
    keeps every sign (fixed topology)."""
    n = 14
    f = _sphere_noise(n, 5).astype(np.float64)
    if clamp_edge:
        # surface through the upper faces: forward differences clamp there
        f = f - 0.18
    T = 0.0
    codes, counts = oracle.classify(f.astype(np.float32), T)
    cells = (n - 1,) * 3
    lay = oracle.layout((n,) * 3, cells)
    cell, slot = oracle.vertex_list(counts, cells, lay)
    assert len(cell) > 50
    ext = (1.0, 2.0, 0.5)
    rng = np.random.default_rng(1)
    gout = rng.standard_normal((len(cell), 6))
    dfield, dT = oracle.adjoint(f, codes, cell, slot, T, gout, ext=ext)

    def loss(ff, TT):
        return float((oracle.emit_f64_field(ff, codes, cell, slot, TT, ext=ext) * gout).sum())

    h = 1e-3

    def fd_sample(idx, step):
        fp = f.copy()
        fp[idx] += step
        fm = f.copy()
        fm[idx] -= step
        return (loss(fp, T) - loss(fm, T)) / (2 * step)

    cand = np.argwhere((np.abs(f - T) > 4 * h) & (dfield != 0))
    assert len(cand) > 20
    picks = cand[rng.choice(len(cand), 40, replace=False)]
    for idx in map(tuple, picks):
        # central differences at h = 1e-3 with one Richardson step (h, h/2):
        # t = (va - T)/(va - vb) is rational, so plain O(h^2) differences
        # are only good to (h/D)^2 for edge differences D of a few 1e-2
        fd = (4.0 * fd_sample(idx, h / 2) - fd_sample(idx, h)) / 3.0
        assert abs(fd - dfield[idx]) <= 2e-6 * max(1.0, abs(fd)), (idx, fd, dfield[idx])
    # untouched samples get exactly zero
    assert (dfield[np.abs(f - T) > 0.2] == 0).sum() > 0


def test_adjoint_dT_finite_difference():
    n = 12
    f = _sphere_noise(n, 9).astype(np.float64)
    T = 0.013
    assert np.min(np.abs(f - T)) > 1e-4
    h = min(1e-3, np.min(np.abs(f - T)) / 4)
    codes, counts = oracle.classify(f.astype(np.float32), T)
    cells = (n - 1,) * 3
    cell, slot = oracle.vertex_list(counts, cells, oracle.layout((n,) * 3, cells))
    gout = np.random.default_rng(2).standard_normal((len(cell), 6))
    _, dT = oracle.adjoint(f, codes, cell, slot, T, gout)
    def fd_T(step):
        lp = float((oracle.emit_f64_field(f, codes, cell, slot, T + step) * gout).sum())
        lm = float((oracle.emit_f64_field(f, codes, cell, slot, T - step) * gout).sum())
        return (lp - lm) / (2 * step)

    fd = (4.0 * fd_T(h / 2) - fd_T(h)) / 3.0
    assert abs(fd - dT) <= 2e-6 * max(1.0, abs(fd)), (fd, dT)


@pytest.mark.parametrize("lattice", [(5, 9, 4), (9, 3, 20), (4, 40, 40), (33, 17, 6)])
def test_block_sums_match_dense_doubled_image(lattice):
    # block sums of any block side 2^b equal the sums of a dense doubled
    # base image (each cell placed at (2 tile_x ts_x + x, 2 tile_y ts_y + y)
    # by numpy) over 2^b x 2^b squares, read in Morton order
    rng = np.random.default_rng(sum(lattice))
    cells = tuple(v - 1 for v in lattice)
    counts = rng.integers(0, 16, size=cells[::-1]).astype(np.uint8)
    lay = oracle.layout(lattice, cells)
    W = 2 << lay[4]
    img = np.zeros((W, W), np.int64)
    z, y, x = np.meshgrid(*[np.arange(c) for c in cells[::-1]], indexing="ij")
    X = (z % lay[2]) * 2 * lay[0] + x
    Y = (z // lay[2]) * 2 * lay[1] + y
    np.add.at(img, (Y.ravel(), X.ravel()), counts.ravel().astype(np.int64))
    for b in range(1, 5):
        side = W >> b
        sums = img.reshape(side, 1 << b, side, 1 << b).sum(axis=(1, 3))     # [Y block, X block]
        bx, by = np.meshgrid(np.arange(side), np.arange(side))
        keys = np.array([oracle.lib().hpo_morton2(int(i), int(j)) for i, j in zip(bx.ravel(), by.ravel())])
        want = np.zeros(side * side, np.int64)
        want[keys] = sums[by.ravel(), bx.ravel()]
        assert oracle.level4(counts, cells, lay, b).tolist() == want.tolist(), b
    assert oracle.block_log2(lay) == min(4, int(2 * lay[0]).bit_length() - 1, int(2 * lay[1]).bit_length() - 1)
