"""GPU parity of the HIP path against the CPU oracle (needs an MI355X).

Every case compares, on the same FP32 field:
  * the HPMC corner code of every cell           -- bit-exact
  * the level-4 pyramid sums (16x16-cell blocks) -- bit-exact
  * the vertex count                             -- exact
  * the (cell, edge) list in extraction order    -- bit-exact, in HPMC's order
  * positions and normals against the oracle's FP64 evaluation on the same
    FP32 samples, within
        position: 1e-5 cell + 2^-22 (|k + t| + 1) cells  (1e-5 of the cell size
                  plus the FP32 rounding of the absolute coordinate)
        normal:   16 * 2^-24 * max(|T|, |f| of the samples used) * extent/cells
                  (FP32 rounding of T - mix(na, nb, t))
"""
import numpy as np
import pytest

import oracle
from hpmc_amd import IsoSurface, workloads

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

EPS = 2.0 ** -24


def _gpu(lattice, T, volume=None, snippet=None, uniforms=(), grid=None, ext=(1.0, 1.0, 1.0)):
    iso = IsoSurface(lattice, grid=grid, extent=ext)
    if volume is not None:
        vol = torch.from_numpy(np.ascontiguousarray(volume)).cuda()
        iso.set_volume(vol)
    else:
        iso.set_snippet(snippet)
        for name, arr in uniforms:
            iso.set_uniform(name, arr)
    verts, ce = iso.extract_all(T, with_cell_edge=True)
    codes = iso.cell_codes()
    l4 = iso.level4()
    torch.cuda.synchronize()
    res = dict(n=iso.count(), verts=verts.cpu().numpy(), ce=ce.cpu().numpy().view(np.uint32),
               codes=codes.cpu().numpy(), level4=l4.cpu().numpy().view(np.uint32), layout=iso.layout())
    iso.close()
    return res


def _check(res, field, T, lattice, grid=None, ext=(1.0, 1.0, 1.0), clamp=True):
    cells = grid or tuple(v - 1 for v in lattice)
    codes, counts = oracle.classify(field, T, cells=cells, size=lattice)
    assert np.array_equal(res["codes"], codes), "cell codes differ"
    lay = oracle.layout(lattice, cells)
    assert tuple(res["layout"]) == lay
    l4 = oracle.level4(counts, cells, lay)
    assert np.array_equal(res["level4"], l4), "level-4 sums differ"
    cell, slot = oracle.vertex_list(counts, cells, lay)
    assert res["n"] == len(cell)
    assert np.array_equal(res["ce"][:, 0], cell), "vertex cells differ (order or content)"
    assert np.array_equal(res["ce"][:, 1], oracle.edge_ids(codes, cell, slot)), "vertex edges differ"
    ref, t, mag = oracle.emit(field, codes, cell, slot, T, ext=ext, cells=cells, size=lattice, clamp=clamp,
                              with_mag=True)
    got = res["verts"].astype(np.float64)
    sc = np.asarray(ext, np.float64) / np.asarray(cells, np.float64)
    pos_cells = ref[:, 3:] / sc
    dpos = np.abs(got[:, 3:] - ref[:, 3:]) / sc
    tol_pos = 1e-5 + 2.0 ** -22 * (np.abs(pos_cells) + 1.0)
    assert (dpos <= tol_pos).all(), "position error %g cells" % float((dpos - tol_pos).max())
    dn = np.abs(got[:, :3] - ref[:, :3])
    tol_n = 16 * EPS * mag[:, None] * sc[None, :]
    assert (dn <= tol_n).all(), "normal error exceeds FP32 bound by %g" % float((dn - tol_n).max())
    assert ((t >= 0) & (t <= 1)).all()
    return len(cell)


def _smooth_noise(shape, seed):
    """Random field with structure at a few cells (upsampled coarse noise)."""
    X, Y, Z = shape
    rng = np.random.default_rng(seed)
    coarse = rng.random((Z // 4 + 2, Y // 4 + 2, X // 4 + 2))
    zi = np.arange(Z) / 4.0
    yi = np.arange(Y) / 4.0
    xi = np.arange(X) / 4.0

    def interp(a, coords, axis):
        i0 = np.floor(coords).astype(int)
        w = coords - i0
        a0 = np.take(a, i0, axis=axis)
        a1 = np.take(a, i0 + 1, axis=axis)
        shp = [1, 1, 1]
        shp[axis] = len(coords)
        return a0 * (1 - w.reshape(shp)) + a1 * w.reshape(shp)

    return interp(interp(interp(coarse, xi, 2), yi, 1), zi, 0).astype(np.float32)


@pytest.mark.parametrize("lattice", [(17, 17, 17), (33, 20, 45), (64, 64, 64), (100, 37, 29), (130, 70, 9),
                                     (65, 129, 33)])
def test_volume_small_bit_exact(lattice):
    f = _smooth_noise(lattice, sum(lattice))
    res = _gpu(lattice, 0.5, volume=f)
    n = _check(res, f, 0.5, lattice)
    assert n > 0


def test_volume_grid_smaller_than_lattice_and_extent():
    lattice, grid, ext = (48, 40, 36), (30, 25, 33), (2.0, 0.5, 1.5)
    f = _smooth_noise(lattice, 11)
    res = _gpu(lattice, 0.45, volume=f, grid=grid, ext=ext)
    assert _check(res, f, 0.45, lattice, grid=grid, ext=ext) > 0


def test_volume_empty_and_full():
    lattice = (40, 40, 40)
    f = _smooth_noise(lattice, 5)
    for T in (-1.0, 2.0):
        res = _gpu(lattice, T, volume=f)
        assert res["n"] == 0 and _check(res, f, T, lattice) == 0


def test_c1_texture_256():
    f = workloads.upsampled_noise(256, 32, seed=1)
    lattice = (256, 256, 256)
    res = _gpu(lattice, 0.5, volume=f)
    assert _check(res, f, 0.5, lattice) > 100000


def test_c0_metaballs_snippet_64():
    rng = workloads.SplitMix64(0)
    balls = workloads.metaball_params(rng, 8)
    lattice = (64, 64, 64)
    res = _gpu(lattice, 1.0, snippet=workloads.METABALL_SNIPPET,
               uniforms=[("hpmc_nballs", np.array([8], np.int32)), ("hpmc_balls", balls)])
    f = workloads.metaballs(lattice, balls, extra=2)
    assert _check(res, f, 1.0, lattice, clamp=False) > 1000


def test_c2_cayley_snippet_512():
    lattice = (512, 512, 512)
    res = _gpu(lattice, 0.0, snippet=workloads.CAYLEY_SNIPPET)
    assert res["n"] == 3694104
    f = workloads.cayley(lattice, extra=2)
    assert _check(res, f, 0.0, lattice, clamp=False) == 3694104


def test_c2v_volume_equals_snippet_topology():
    lattice = (512, 512, 512)
    f = workloads.cayley(lattice)
    a = _gpu(lattice, 0.0, volume=f)
    b = _gpu(lattice, 0.0, snippet=workloads.CAYLEY_SNIPPET)
    assert a["n"] == b["n"] == 3694104
    assert np.array_equal(a["ce"], b["ce"])
    assert np.array_equal(a["codes"], b["codes"])
    # positions use only in-lattice samples: identical bits
    assert np.array_equal(a["verts"][:, 3:], b["verts"][:, 3:])


def test_rebuild_and_determinism():
    lattice = (96, 80, 72)
    f = _smooth_noise(lattice, 3)
    vol = torch.from_numpy(f).cuda()
    iso = IsoSurface(lattice)
    iso.set_volume(vol)
    outs = []
    for T in (0.4, 0.6, 0.4):
        v, ce = iso.extract_all(T, with_cell_edge=True)
        outs.append((v.cpu().numpy().copy(), ce.cpu().numpy().copy()))
        codes, counts = oracle.classify(f, T)
        assert len(v) == counts.sum()
    assert np.array_equal(outs[0][0], outs[2][0]) and np.array_equal(outs[0][1], outs[2][1])
    iso.close()


def test_capacity_errors():
    import hpmc_amd
    lattice = (64, 64, 64)
    f = _smooth_noise(lattice, 9)
    vol = torch.from_numpy(f).cuda()
    iso = IsoSurface(lattice)
    iso.set_volume(vol)
    iso.build(0.5)
    n = iso.count()
    assert n > 10
    small = torch.zeros((n - 5, 6), dtype=torch.float32, device="cuda")
    iso.bind_output(small)
    with pytest.raises(hpmc_amd.HPMCError, match="CAPACITY|holds"):
        iso.extract()
    # without a count the extraction is enqueued and drops what does not fit
    iso.build(0.5)
    iso.bind_output(small)
    iso.extract()
    torch.cuda.synchronize()
    assert hpmc_amd.lib().HPMCacquireDroppedVertices(iso.th) == 5
    full, _ = iso.extract_all(0.5)
    assert torch.equal(small, full[: n - 5])
    iso.close()
