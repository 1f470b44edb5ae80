"""Lattices past the 32-bit sample and byte ranges.

Samples are addressed per plane (a 64-bit plane base plus a 32-bit offset
within the plane), so one volume may hold 2^32 samples or more.  Each test
builds a large lattice that holds a small feature where the large offsets
are, and compares the result with the same feature in a small lattice, the
sample values being bit-identical around it (the fields are sums of squares
of dyadic differences, exact near the feature):

  * cell codes around the feature     -- bit-exact
  * vertex count                      -- exact
  * positions, shifted by the integer offset between the lattices (grid
    extent = cells, so a cell is 1.0 in both): 2^-11 absolute for
    coordinates below 2^11, 2^-9 below 2^14 (four times the FP32 spacing)
  * normals                           -- 1e-4 absolute (formed from the same
    samples with unit scales, so equal in practice)
"""
import gc

import numpy as np
import pytest

from hpmc_amd import HPMCError, IsoSurface, last_error

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _sphere_field(shape, centre, r2, chunk=64):
    """f = ((x - cx)^2 + (y - cy)^2) + (z - cz)^2 - r2 on a (z, y, x) FP32
    lattice, filled in slabs along z."""
    nz, ny, nx = shape
    dev = "cuda"
    dx2 = (torch.arange(nx, device=dev, dtype=torch.float32) - centre[0]) ** 2
    dy2 = (torch.arange(ny, device=dev, dtype=torch.float32) - centre[1]) ** 2
    dz2 = (torch.arange(nz, device=dev, dtype=torch.float32) - centre[2]) ** 2
    f = torch.empty(shape, dtype=torch.float32, device=dev)
    plane = dx2[None, :] + dy2[:, None]
    for z0 in range(0, nz, chunk):
        z1 = min(nz, z0 + chunk)
        torch.add(plane[None, :, :], dz2[z0:z1, None, None], out=f[z0:z1])
        f[z0:z1] -= r2
    return f


def _match(big, small, offset):
    """Pairs the distinct vertices of `big` (shifted by -offset) one to one
    with those of `small` (the output is a triangle soup: a vertex repeats
    once per triangle using it), checks that each pair repeats equally often
    and returns the largest position and normal differences."""
    from scipy.spatial import cKDTree

    def distinct(v):
        a = v.cpu().numpy()
        pos, first, reps = np.unique(a[:, 3:], axis=0, return_index=True, return_counts=True)
        return pos.astype(np.float64), a[first, :3], reps

    pb, nb, rb = distinct(big)
    ps, ns, rs = distinct(small)
    assert pb.shape == ps.shape, (pb.shape, ps.shape)
    d, idx = cKDTree(ps).query(pb - np.asarray(offset, dtype=np.float64))
    assert np.unique(idx).size == idx.size, "vertices do not pair one to one"
    assert np.array_equal(rb, rs[idx]), "paired vertices repeat unequally"
    return float(d.max()), float(np.abs(nb - ns[idx]).max())


def _release():
    gc.collect()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def test_volume_of_2_33_samples():
    # A 2048^3 FP32 volume (2^33 samples, 32 GiB) with a sphere whose samples
    # lie between sample indices 7.8e9 and 8.2e9, against the same sphere in
    # a 64^3 volume.  The build, the cell codes and the extraction read it
    # through 64-bit plane bases.
    n, m = 2048, 64
    offset = (968, 1468, 1868)
    r2 = 20.3 ** 2
    small_c = (32.25, 32.25, 32.25)
    big_c = tuple(c + o for c, o in zip(small_c, offset))
    print("2048^3 volume, sphere at", big_c, flush=True)

    fs = _sphere_field((m, m, m), small_c, r2)
    small = IsoSurface((m, m, m), extent=(m - 1.0,) * 3)
    small.set_volume(fs)
    vs, _ = small.extract_all(0.0)
    codes_s = small.cell_codes().clone()
    assert vs.shape[0] > 1000

    fb = _sphere_field((n, n, n), big_c, r2)
    big = IsoSurface((n, n, n), extent=(n - 1.0,) * 3)
    big.set_volume(fb)
    vb, _ = big.extract_all(0.0)
    assert vb.shape[0] == vs.shape[0]
    dpos, dnrm = _match(vb, vs, offset)
    assert dpos <= 2.0 ** -11, dpos
    assert dnrm <= 1e-4, dnrm
    # 2047^3 cells do not fit the 32-bit cell field of (cell, edge) records
    out = torch.empty_like(vb)
    big.bind_output(out, torch.empty((vb.shape[0], 2), dtype=torch.int32, device="cuda"))
    with pytest.raises(HPMCError):
        big.extract()
    assert last_error(big.h)[0] == 3                  # HPMC_ERR_SIZE
    big.bind_output(out)
    big.extract()
    torch.cuda.synchronize()
    assert torch.equal(out, vb)
    codes_b = big.cell_codes()
    ox, oy, oz = offset
    win = codes_b[oz:oz + m - 1, oy:oy + m - 1, ox:ox + m - 1]
    assert torch.equal(win, codes_s)
    # every other cell is outside the sphere
    assert int(codes_b.count_nonzero()) == int(codes_s.count_nonzero())
    del fb, codes_b, win, out
    big.close()
    small.close()
    _release()


def _cylinder_gradient_field(shape, centre, r2, chunk=1024):
    """(z, y, x, 4) RGBA lattice of f = (x - cx)^2 + (y - cy)^2 - r2 with its
    gradient (2 (x - cx), 2 (y - cy), 0) in rgb, filled in bands along y."""
    nz, ny, nx = shape
    dev = "cuda"
    dx = torch.arange(nx, device=dev, dtype=torch.float32) - centre[0]
    dy = torch.arange(ny, device=dev, dtype=torch.float32) - centre[1]
    g = torch.empty((nz, ny, nx, 4), dtype=torch.float32, device=dev)
    for y0 in range(0, ny, chunk):
        y1 = min(ny, y0 + chunk)
        band = g[:, y0:y1]
        band[..., 0] = 2.0 * dx[None, None, :]
        band[..., 1] = 2.0 * dy[None, y0:y1, None]
        band[..., 2] = 0.0
        band[..., 3] = (dx[None, None, :] ** 2 + dy[None, y0:y1, None] ** 2) - r2
    return g


def test_gradient_plane_of_2_28_texels():
    # A 16384 x 16384 x 2 RGBA gradient lattice: one plane is 2^28 texels,
    # 4 GiB, whose last value sits 2^32 - 4 bytes past the plane start.  A
    # cylinder cut by the far corner of the lattice, against the same cut in
    # a 64 x 64 x 2 lattice.
    n, m = 16384, 64
    offset = (n - m, n - m, 0)
    r2 = 20.3 ** 2
    small_c = (50.25, 50.25)
    big_c = (small_c[0] + offset[0], small_c[1] + offset[1])
    print("16384 x 16384 x 2 gradient lattice, cylinder at", big_c, flush=True)

    gs = _cylinder_gradient_field((2, m, m), small_c, r2)
    small = IsoSurface((m, m, 2), extent=(m - 1.0, m - 1.0, 1.0))
    small.set_volume(gs, gradient=True)
    vs, _ = small.extract_all(0.0)
    codes_s = small.cell_codes().clone()
    assert vs.shape[0] > 100

    gb = _cylinder_gradient_field((2, n, n), big_c, r2)
    big = IsoSurface((n, n, 2), extent=(n - 1.0, n - 1.0, 1.0))
    big.set_volume(gb, gradient=True)
    vb, _ = big.extract_all(0.0)
    assert vb.shape[0] == vs.shape[0]
    dpos, dnrm = _match(vb, vs, offset)
    assert dpos <= 2.0 ** -9, dpos     # coordinates up to 2^14
    assert dnrm <= 1e-4, dnrm
    codes_b = big.cell_codes()
    win = codes_b[:, n - m:, n - m:]
    assert torch.equal(win, codes_s)
    assert int(codes_b.count_nonzero()) == int(codes_s.count_nonzero())
    del gb, codes_b, win
    big.close()
    small.close()
    _release()
