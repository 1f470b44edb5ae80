"""GPU adjoint (HPMCadjoint / HPMCbatchAdjoint and the torch.autograd
wrappers) against the oracle's FP64 adjoint, FP64 finite differences of the
oracle's FP64 forward, and itself (run-to-run bit equality).

Bar: dL/dfield and dL/dT within 1e-4 relative (max-norm) of the FP64
adjoint evaluated on the same FP32 samples; finite differences (central,
h = 1e-3 with one Richardson step) on samples away from sign changes within
1e-4 relative; repeated runs bit-identical."""
import numpy as np
import pytest

import oracle
from hpmc_amd import Batch, IsoSurface
from hpmc_amd.autograd import marching_cubes, marching_cubes_batch

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def sphere_noise(shape, seed, sigma=0.01, r=0.3):
    """Signed distance to a sphere of radius r at the volume centre (texel
    centre coordinates) plus Gaussian noise; (z, y, x) float32."""
    X, Y, Z = shape
    px = (np.arange(X) + 0.5) / X
    py = (np.arange(Y) + 0.5) / Y
    pz = (np.arange(Z) + 0.5) / Z
    d = np.sqrt((px[None, None, :] - 0.5) ** 2 + (py[None, :, None] - 0.5) ** 2 + (pz[:, None, None] - 0.5) ** 2)
    rng = np.random.default_rng(seed)
    return (d - r + sigma * rng.standard_normal(d.shape)).astype(np.float32)


def _oracle_list(f, T):
    shape = f.shape[::-1]
    cells = tuple(v - 1 for v in shape)
    codes, counts = oracle.classify(f, T)
    cell, slot = oracle.vertex_list(counts, cells, oracle.layout(shape, cells))
    return codes, cell, slot


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


# cubic and non-cubic lattices, anisotropic extents, a grid of 7 cells along x
@pytest.mark.parametrize("shape,ext", [((40, 40, 40), (1.0, 1.0, 1.0)), ((37, 52, 29), (2.0, 0.5, 1.0)),
                                       ((48, 48, 48), (1.0, 1.0, 1.0)), ((8, 60, 60), (1.0, 1.0, 1.0))])
def test_adjoint_matches_fp64_oracle(shape, ext):
    f = sphere_noise(shape, 7)
    T = 0.0
    iso = IsoSurface(shape, extent=ext)
    vol = torch.from_numpy(f).cuda()
    iso.set_volume(vol)
    verts, _ = iso.extract_all(T)
    n = verts.shape[0]
    assert n > 1000
    gout = torch.from_numpy(np.random.default_rng(3).standard_normal((n, 6)).astype(np.float32)).cuda()
    gf, gt = iso.adjoint(gout)
    torch.cuda.synchronize()
    codes, cell, slot = _oracle_list(f, T)
    assert len(cell) == n
    ref_f, ref_t = oracle.adjoint(f, codes, cell, slot, T, gout.cpu().numpy().astype(np.float64), ext=ext)
    got = gf.cpu().numpy().astype(np.float64)
    assert _rel(got, ref_f) < 1e-4, _rel(got, ref_f)
    assert abs(float(gt) - ref_t) <= 1e-4 * max(abs(ref_t), 1.0)
    # samples no vertex reads stay exactly zero
    assert np.array_equal(got == 0, ref_f == 0)
    iso.close()


@pytest.mark.parametrize("shape,grid", [((40, 36, 30), (30, 25, 20)), ((33, 20, 17), (32, 6, 9))])
def test_adjoint_grid_smaller_than_lattice(shape, grid):
    # cells of a grid smaller than the lattice: edges on the grid's upper
    # faces are owned by the last cells, and the normals' forward samples
    # past the grid are real lattice samples (no clamp) until the lattice ends
    f = sphere_noise(shape, 21)
    T = 0.0
    iso = IsoSurface(shape, grid=grid)
    iso.set_volume(torch.from_numpy(f).cuda())
    verts, _ = iso.extract_all(T)
    n = verts.shape[0]
    assert n > 100
    codes, counts = oracle.classify(f, T, cells=grid, size=shape)
    cell, slot = oracle.vertex_list(counts, grid, oracle.layout(shape, grid))
    assert len(cell) == n
    gout = torch.from_numpy(np.random.default_rng(8).standard_normal((n, 6)).astype(np.float32)).cuda()
    gf, gt = iso.adjoint(gout)
    torch.cuda.synchronize()
    ref_f, ref_t = oracle.adjoint(f, codes, cell, slot, T, gout.cpu().numpy().astype(np.float64), cells=grid,
                                  size=shape)
    got = gf.cpu().numpy().astype(np.float64)
    assert _rel(got, ref_f) < 1e-4, _rel(got, ref_f)
    assert abs(float(gt) - ref_t) <= 1e-4 * max(abs(ref_t), 1.0)
    assert np.array_equal(got == 0, ref_f == 0)
    iso.close()


def test_adjoint_every_edge_changing_sign():
    # a perturbed checkerboard: every lattice edge changes sign, so each cell
    # of the top slice owns five edges and a 16x16 block of it more than
    # the 1024 its edge list holds (each lane then takes its own cells'
    # edges), and every gather tile's edge table is full
    shape = (40, 36, 12)
    x = np.arange(shape[0])[None, None, :]
    y = np.arange(shape[1])[None, :, None]
    z = np.arange(shape[2])[:, None, None]
    rng = np.random.default_rng(13)
    f = (((x + y + z) & 1) + 0.2 * rng.random((shape[2], shape[1], shape[0])) - 0.1).astype(np.float32)
    T = 0.5
    iso = IsoSurface(shape)
    iso.set_volume(torch.from_numpy(f).cuda())
    verts, _ = iso.extract_all(T)
    n = verts.shape[0]
    codes, cell, slot = _oracle_list(f, T)
    assert len(cell) == n and n > 12 * 39 * 35 * 11 // 2
    gout = torch.from_numpy(np.random.default_rng(4).standard_normal((n, 6)).astype(np.float32)).cuda()
    gf, gt = iso.adjoint(gout)
    gf2, gt2 = iso.adjoint(gout)
    torch.cuda.synchronize()
    assert torch.equal(gf, gf2) and torch.equal(gt, gt2)
    ref_f, ref_t = oracle.adjoint(f, codes, cell, slot, T, gout.cpu().numpy().astype(np.float64))
    got = gf.cpu().numpy().astype(np.float64)
    assert _rel(got, ref_f) < 1e-4, _rel(got, ref_f)
    assert abs(float(gt) - ref_t) <= 1e-4 * max(abs(ref_t), 1.0)
    iso.close()


def test_adjoint_is_deterministic():
    shape = (64, 64, 64)
    f = sphere_noise(shape, 11)
    iso = IsoSurface(shape)
    iso.set_volume(torch.from_numpy(f).cuda())
    verts, _ = iso.extract_all(0.0)
    gout = torch.randn(verts.shape, device="cuda", generator=torch.Generator("cuda").manual_seed(5))
    runs = [iso.adjoint(gout) for _ in range(3)]
    torch.cuda.synchronize()
    for gf, gt in runs[1:]:
        assert torch.equal(gf, runs[0][0]) and torch.equal(gt, runs[0][1])
    iso.close()


def _c4_loss(verts, centre):
    # mean (|v - c| - 0.35)^2 over vertex positions
    return ((torch.linalg.norm(verts[:, 3:] - centre, dim=1) - 0.35) ** 2).mean()


def test_autograd_c4_loss_against_fp64_and_finite_differences():
    shape = (36, 36, 36)
    f = sphere_noise(shape, 21)
    T = 0.0
    field = torch.from_numpy(f).cuda().requires_grad_(True)
    iso_t = torch.tensor(T, dtype=torch.float32, device="cuda", requires_grad=True)
    verts = marching_cubes(field, iso_t)
    cells = np.array(shape, np.float64) - 1
    # the sphere centre (texel coordinate 0.5) in grid coordinates
    centre = torch.tensor((0.5 * np.array(shape) - 0.5) / cells, dtype=torch.float32, device="cuda")
    loss = _c4_loss(verts, centre)
    loss.backward()
    # dL/dverts of the same loss, from torch, in FP64 for the oracle
    v64 = verts.detach().double().requires_grad_(True)
    l64 = ((torch.linalg.norm(v64[:, 3:] - centre.double(), dim=1) - 0.35) ** 2).mean()
    (gv,) = torch.autograd.grad(l64, [v64])
    codes, cell, slot = _oracle_list(f, T)
    ref_f, ref_t = oracle.adjoint(f, codes, cell, slot, T, gv.cpu().numpy())
    got = field.grad.cpu().numpy().astype(np.float64)
    assert _rel(got, ref_f) < 1e-4, _rel(got, ref_f)
    assert abs(float(iso_t.grad) - ref_t) <= 1e-4 * max(abs(ref_t), 1e-6)

    # finite differences of the FP64 forward on samples away from T
    f64 = f.astype(np.float64)

    def loss64(ff):
        out = oracle.emit_f64_field(ff, codes, cell, slot, T)
        return float((((np.linalg.norm(out[:, 3:] - centre.cpu().numpy(), axis=1) - 0.35) ** 2)).mean())

    h = 1e-3
    rng = np.random.default_rng(0)
    cand = np.argwhere((np.abs(f64 - T) > 4 * h) & (np.abs(ref_f) > 1e-3 * np.abs(ref_f).max()))
    for idx in map(tuple, cand[rng.choice(len(cand), 12, replace=False)]):
        def fd(step):
            fp = f64.copy()
            fp[idx] += step
            fm = f64.copy()
            fm[idx] -= step
            return (loss64(fp) - loss64(fm)) / (2 * step)

        est = (4.0 * fd(h / 2) - fd(h)) / 3.0
        assert abs(got[idx] - est) <= 1e-4 * np.abs(ref_f).max(), (idx, got[idx], est)


# Batch and single builds sort different key widths, with different digit
# sizes: the first set's batch has 19 key bits (8-bit digits) and its 17-bit
# volumes alone use 9-bit digits; the second set's batch holds 2^24+ samples
# (25 key bits, 9-bit digits) and its volumes alone 22-23 bits (8-bit
# digits).  Either way every sample gets the same contributions in the same
# order.
@pytest.mark.parametrize("shapes", [[(33, 40, 28), (48, 48, 48), (20, 64, 36), (57, 31, 44)],
                                    [(160, 160, 160), (200, 150, 180), (128, 256, 130), (150, 170, 160)]])
def test_batch_adjoint_equals_single_volume_adjoints(shapes):
    fields = [sphere_noise(s, 30 + i) for i, s in enumerate(shapes)]
    vols = [torch.from_numpy(f).cuda() for f in fields]
    b = Batch()
    for v in vols:
        b.add_volume(v)
    b.build(0.0)
    counts, bases, total = b.counts()
    out = torch.empty((total, 6), dtype=torch.float32, device="cuda")
    b.extract(out)
    gout = torch.randn((total, 6), device="cuda", generator=torch.Generator("cuda").manual_seed(9))
    grads, gt = b.adjoint(gout)
    torch.cuda.synchronize()
    dts = []
    for i, (s, v) in enumerate(zip(shapes, vols)):
        iso = IsoSurface(s)
        iso.set_volume(v)
        verts, _ = iso.extract_all(0.0)
        assert verts.shape[0] == counts[i]
        gf_i, gt_i = iso.adjoint(gout[bases[i]:bases[i] + counts[i]].contiguous())
        torch.cuda.synchronize()
        # per sample the same contributions in the same order: same bits
        assert torch.equal(grads[i], gf_i), i
        dts.append(float(gt_i))
        iso.close()
    assert abs(float(gt) - sum(dts)) <= 1e-5 * max(1.0, sum(abs(d) for d in dts))
    b.close()


def test_autograd_batch_c4():
    rng = np.random.default_rng(4)
    shapes = [tuple(int(v) for v in rng.integers(32, 72, 3)) for _ in range(6)]
    fields = [torch.from_numpy(sphere_noise(s, 50 + i)).cuda().requires_grad_(True) for i, s in enumerate(shapes)]
    iso_t = torch.tensor(0.0, dtype=torch.float32, device="cuda", requires_grad=True)
    outs = marching_cubes_batch(fields, iso_t)
    loss = 0.0
    for s, v in zip(shapes, outs):
        cells = np.array(s, np.float64) - 1
        centre = torch.tensor((0.5 * np.array(s) - 0.5) / cells, dtype=torch.float32, device="cuda")
        loss = loss + _c4_loss(v, centre)
    loss.backward()
    # each volume's gradient equals the single-volume autograd's
    for i, (s, fld) in enumerate(zip(shapes, fields)):
        f2 = fld.detach().clone().requires_grad_(True)
        it2 = torch.tensor(0.0, dtype=torch.float32, device="cuda", requires_grad=True)
        v2 = marching_cubes(f2, it2)
        cells = np.array(s, np.float64) - 1
        centre = torch.tensor((0.5 * np.array(s) - 0.5) / cells, dtype=torch.float32, device="cuda")
        _c4_loss(v2, centre).backward()
        assert torch.equal(v2, outs[i].detach())
        assert torch.equal(f2.grad, fld.grad), i
    assert iso_t.grad is not None and torch.isfinite(iso_t.grad)
