"""GPU parity for binary fields and gradient-providing fields (needs an
MI355X).  Reference: shadergen.cpp:93-101 (binary threshold 0.5),
507-510 and 521-522 (midpoints and tabulated normals), 539-549 (gradient
fields); constants.cpp:217-273 (the binary normal table).  Comparison and
tolerances: tests/parity_util.py."""
import numpy as np
import pytest

from hpmc_amd import IsoSurface, workloads
from parity_util import check, gpu_run, smooth_noise, with_gradient

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("lattice", [(33, 20, 45), (64, 64, 64)])
def test_binary_volume(lattice):
    # a 0/1 field as a binary texture would hold; the build's threshold is
    # ignored in favour of 0.5
    f = (smooth_noise(lattice, 21) > 0.5).astype(np.float32)
    res = gpu_run(lattice, 0.123, volume=f, binary=True)
    assert check(res, f, 0.123, lattice, mode="binary") > 0


def test_binary_volume_non_binary_values_grid_and_extent():
    lattice, grid, ext = (48, 40, 36), (30, 25, 33), (2.0, 0.5, 1.5)
    f = smooth_noise(lattice, 4)
    res = gpu_run(lattice, 7.0, volume=f, grid=grid, ext=ext, binary=True)
    assert check(res, f, 7.0, lattice, grid=grid, ext=ext, mode="binary") > 0


def test_binary_snippet_cayley():
    lattice = (96, 96, 96)
    res = gpu_run(lattice, 0.0, snippet=workloads.CAYLEY_SNIPPET, binary=True)
    f = workloads.cayley(lattice, extra=2)
    assert check(res, f, 0.0, lattice, clamp=False, mode="binary") > 1000


def test_binary_flag_applies_at_reconfiguration():
    # hpiface.cpp:131-142: the flag does not taint the HistoPyramid
    lattice = (40, 40, 40)
    f = smooth_noise(lattice, 8)
    vol = torch.from_numpy(f).cuda()
    iso = IsoSurface(lattice)
    iso.set_volume(vol)
    iso.build(0.3)
    n_cont = iso.count()
    iso.set_binary(True)
    iso.build(0.3)
    assert iso.count() == n_cont            # still continuous, T = 0.3
    iso.set_volume(vol)                     # same mode: no reconfiguration
    iso.build(0.3)
    assert iso.count() == n_cont
    from hpmc_amd import lib
    lib().HPMCsetGridExtent(iso.h, 1.0, 1.0, 1.0)   # taints
    iso.build(0.3)
    import oracle
    _, counts = oracle.classify(f, 0.5)
    assert iso.count() == int(counts.sum())
    iso.close()


@pytest.mark.parametrize("lattice", [(33, 20, 45), (80, 72, 64)])
def test_gradient_volume(lattice):
    f = smooth_noise(lattice, 31)
    f4 = with_gradient(f)
    res = gpu_run(lattice, 0.5, volume=f4, gradient=True)
    assert check(res, f, 0.5, lattice, mode="gradient", field4=f4) > 0


def test_gradient_volume_binary():
    # a gradient texture in binary mode: classify the value channel at 0.5,
    # emit midpoints and tabulated normals
    lattice = (40, 36, 30)
    f = smooth_noise(lattice, 41)
    res = gpu_run(lattice, 0.9, volume=with_gradient(f), gradient=True, binary=True)
    assert check(res, f, 0.9, lattice, mode="binary") > 0


def test_gradient_snippet_cayley():
    lattice = (128, 128, 128)
    res = gpu_run(lattice, 0.0, snippet=workloads.CAYLEY_GRAD_SNIPPET, gradient=True)
    f4 = workloads.cayley_grad(lattice, extra=2)
    f = np.ascontiguousarray(f4[..., 3])
    assert check(res, f, 0.0, lattice, clamp=False, mode="gradient", field4=f4) > 10000


@pytest.mark.parametrize("lattice,tuning", [((150, 70, 90), (1, 8, 8)), ((150, 70, 90), (4, 16, 256)),
                                            ((150, 70, 90), (2, 32, 0)), ((40, 70, 90), (8, 16, 0)),
                                            ((70, 33, 50), (16, 8, 0)), ((200, 20, 30), (5, 32, 16))])
def test_launch_shapes_do_not_change_results(lattice, tuning):
    # HPMCsetTuning: launch shapes of one HistoPyramid; every shape gives the
    # same codes, count, (cell, edge) list and vertex bits (narrow lattices
    # stack row bands in a workgroup; waves past the rows or columns leave)
    f = smooth_noise(lattice, 17)
    vol = torch.from_numpy(f).cuda()
    ref = IsoSurface(lattice)
    ref.set_volume(vol)
    v0, ce0 = ref.extract_all(0.5, with_cell_edge=True)
    iso = IsoSurface(lattice)
    iso.set_volume(vol)
    iso.set_tuning(*tuning)
    v1, ce1 = iso.extract_all(0.5, with_cell_edge=True)
    got = iso.tuning()
    assert got[0] == tuning[0] and got[1] == tuning[1]
    assert torch.equal(ce0, ce1) and torch.equal(v0, v1)
    assert torch.equal(ref.cell_codes(), iso.cell_codes())
    ref.close()
    iso.close()


def test_kept_output_buffer_grows_like_the_reference_demo():
    # apps/transform_feedback/transform_feedback.cpp:341-349, 393-402: one
    # buffer across frames, grown to 1.1 x N only when N no longer fits
    from hpmc_amd import grown_capacity
    lattice = (80, 70, 60)
    f = smooth_noise(lattice, 23)
    vol = torch.from_numpy(f).cuda()
    iso = IsoSurface(lattice)
    iso.set_volume(vol)
    ref = IsoSurface(lattice)
    ref.set_volume(vol)
    cap, ptr = 3000, None
    for T in (0.95, 0.5, 0.9, 0.45, 0.5):
        v, ce = iso.extract_growing(T, with_cell_edge=True)
        rv, rce = ref.extract_all(T, with_cell_edge=True)
        assert torch.equal(v, rv) and torch.equal(ce, rce)
        n = rv.shape[0]
        grew = n > cap
        cap = grown_capacity(cap, n)
        kept = iso._kept[0]
        assert kept.shape[0] == cap
        if ptr is not None:
            assert (kept.data_ptr() != ptr) == grew
        ptr = kept.data_ptr()
    iso.close()
    ref.close()


# The particles demo's degree-5 shapes (apps/particles/particles.cpp:100-158)
# as hipRTC snippets with their 12-float `shape` uniform: continuous mode
# (forward-difference normals) and the demo's own gradient mode
# (HPMC_fetchGrad, with its z-derivative as written), iso 0.001.
@pytest.mark.parametrize("lattice,shape", [((128, 128, 128), ("torus", "torus", 0.0)),
                                           ((96, 112, 80), ("helix", "between1", 0.4)),
                                           ((100, 90, 110), ("between2", "daddel", 0.5))])
def test_particles_snippet(lattice, shape):
    coeffs = workloads.particles_shape(*shape)
    T = workloads.PARTICLES_ISO
    res = gpu_run(lattice, T, snippet=workloads.PARTICLES_SNIPPET, uniforms=[("shape", coeffs)])
    f = workloads.particles(lattice, coeffs, extra=2)
    assert check(res, f, T, lattice, clamp=False) > 10000


@pytest.mark.parametrize("lattice,shape", [((128, 128, 128), ("helix", "between1", 0.4)),
                                           ((90, 100, 70), ("daddel", "daddel", 0.0))])
def test_particles_gradient_snippet(lattice, shape):
    coeffs = workloads.particles_shape(*shape)
    T = workloads.PARTICLES_ISO
    res = gpu_run(lattice, T, snippet=workloads.PARTICLES_GRAD_SNIPPET, uniforms=[("shape", coeffs)],
                  gradient=True)
    f4 = workloads.particles_grad(lattice, coeffs, extra=2)
    f = np.ascontiguousarray(f4[..., 3])
    assert check(res, f, T, lattice, clamp=False, mode="gradient", field4=f4) > 10000
