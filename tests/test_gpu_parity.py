"""GPU parity of the HIP path against the CPU oracle (needs an MI355X),
continuous fields: volumes and snippets.  The comparison and its tolerances
are those of tests/parity_util.py (codes, level-4 sums, count and ordered
(cell, edge) list bit-exact; positions and normals within FP32 bounds of the
oracle's FP64 evaluation)."""
import numpy as np
import pytest

import oracle
from hpmc_amd import IsoSurface, workloads

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from parity_util import EPS, gpu_run as _gpu, check as _check, smooth_noise as _smooth_noise  # noqa: E402


@pytest.mark.parametrize("lattice", [(17, 17, 17), (33, 20, 45), (64, 64, 64), (100, 37, 29), (130, 70, 9),
                                     (65, 129, 33)])
def test_volume_small_bit_exact(lattice):
    f = _smooth_noise(lattice, sum(lattice))
    res = _gpu(lattice, 0.5, volume=f)
    n = _check(res, f, 0.5, lattice)
    assert n > 0


@pytest.mark.parametrize("lattice", [(10, 10, 2), (1500, 11, 7), (11, 900, 5), (12, 12, 700), (300, 300, 3),
                                     (16384, 10, 2), (10, 16384, 2), (10, 10, 16384)])
def test_volume_extreme_shapes_bit_exact(lattice):
    # the smallest lattice taken (9 x 9 x 1 cells), long thin lattices up to
    # the reference's 16384-sample limit along each axis (init.cpp:82-103),
    # thousands of slices: tile layouts with one tile row or column, wide and
    # narrow classification bands, many z segments
    f = _smooth_noise(lattice, 3)
    res = _gpu(lattice, 0.5, volume=f)
    assert _check(res, f, 0.5, lattice) > 100


# Grids of 2 to 8 cells along x or y (the reference takes any lattice of 2 to
# 16384 samples per axis, init.cpp:82-103): doubled tiles narrower than 16
# cells, so blocks are 2x2, 4x4 or 8x8 cells and several tiles share a
# 16x16 square of the doubled image; order and sums are those of the
# reference's pyramid all the same.
TINY = [(3, 3, 3), (5, 9, 4), (9, 3, 20), (4, 40, 40), (6, 6, 6), (3, 17, 50), (40, 5, 33), (8, 8, 300),
        (130, 6, 12), (7, 9, 2)]


def _tiny_field(lattice, seed):
    # white noise: every small lattice gets many crossings
    rng = np.random.default_rng(seed)
    return rng.random(lattice[::-1]).astype(np.float32)


@pytest.mark.parametrize("lattice", TINY)
def test_volume_tiny_grids_bit_exact(lattice):
    f = _tiny_field(lattice, sum(lattice))
    res = _gpu(lattice, 0.5, volume=f)
    assert _check(res, f, 0.5, lattice) > 0


@pytest.mark.parametrize("lattice", [(5, 9, 4), (4, 40, 40), (3, 17, 50)])
def test_tiny_grids_other_fields(lattice):
    # snippet, 8-bit, gradient and binary fields through the same blocks
    from parity_util import with_gradient
    res = _gpu(lattice, 0.0, snippet=workloads.CAYLEY_SNIPPET)
    _check(res, workloads.cayley(lattice, extra=2), 0.0, lattice, clamp=False)
    f = _tiny_field(lattice, 7)
    b = np.clip(np.rint(f * 255.0), 0, 255).astype(np.uint8)
    res8 = _gpu(lattice, 0.5, volume=b)
    ref = _gpu(lattice, 0.5, volume=(b.astype(np.float32) / np.float32(255.0)).astype(np.float32))
    assert res8["n"] == ref["n"] > 0 and np.array_equal(res8["ce"], ref["ce"])
    assert np.array_equal(res8["verts"], ref["verts"])
    g = with_gradient(f)
    resg = _gpu(lattice, 0.5, volume=g, gradient=True)
    assert _check(resg, f, 0.5, lattice, mode="gradient", field4=g) > 0
    resb = _gpu(lattice, 0.5, volume=f, binary=True)
    assert _check(resb, f, 0.5, lattice, mode="binary") > 0


def test_tiny_grid_smaller_than_lattice_and_extent():
    lattice, grid, ext = (12, 30, 20), (6, 21, 17), (0.5, 2.0, 1.0)
    f = _tiny_field(lattice, 13)
    res = _gpu(lattice, 0.5, volume=f, grid=grid, ext=ext)
    assert _check(res, f, 0.5, lattice, grid=grid, ext=ext) > 0


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
