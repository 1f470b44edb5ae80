"""Shared GPU-parity helpers: run the HIP path through the C ABI and compare
it with the CPU oracle on the same FP32 field.

Every comparison covers
  * the HPMC corner code of every cell           -- bit-exact
  * the level-4 pyramid sums (16x16-cell blocks) -- bit-exact
  * the vertex count                             -- exact
  * the (cell, edge) list in extraction order    -- bit-exact, in HPMC's order
  * positions and normals against the oracle's FP64 evaluation on the same
    FP32 samples, within
        position: 1e-5 cell + 2^-22 (|k + t| + 1) cells  (1e-5 of the cell size
                  plus the FP32 rounding of the absolute coordinate)
        normal:   continuous / gradient fields:
                  16 * 2^-24 * max(|T|, |f| or |grad| of the samples used) * extent/cells
                  (FP32 rounding of T - mix(na, nb, t));
                  binary fields: 2^-23 relative (one FP32 product with extent/cells)
"""
import numpy as np

import oracle
from hpmc_amd import IsoSurface

EPS = 2.0 ** -24


def gpu_run(lattice, T, volume=None, snippet=None, uniforms=(), grid=None, ext=(1.0, 1.0, 1.0),
            gradient=False, binary=False):
    """Build + extract on the GPU; returns counts, vertices, (cell, edge)
    records, per-cell codes, level-4 sums and the layout."""
    import torch
    iso = IsoSurface(lattice, grid=grid, extent=ext)
    if binary:
        iso.set_binary(True)
    if volume is not None:
        vol = torch.from_numpy(np.ascontiguousarray(volume)).cuda()
        iso.set_volume(vol, gradient=gradient)
    else:
        iso.set_snippet(snippet, gradient=gradient)
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


def _check_positions(got, ref, sc):
    pos_cells = ref[:, 3:] / sc
    dpos = np.abs(got[:, 3:] - ref[:, 3:]) / sc
    tol_pos = 1e-5 + 2.0 ** -22 * (np.abs(pos_cells) + 1.0)
    assert (dpos <= tol_pos).all(), "position error %g cells" % float((dpos - tol_pos).max())


def check(res, field, T, lattice, grid=None, ext=(1.0, 1.0, 1.0), clamp=True, mode="continuous", field4=None):
    """Compares a gpu_run result with the oracle.  `field` is the scalar
    lattice the classification sees (for gradient fields, the value channel);
    `field4` the (z, y, x, 4) gradient lattice for mode="gradient".  Binary
    fields classify against 0.5 whatever T was passed to the build."""
    cells = grid or tuple(v - 1 for v in lattice)
    Tc = 0.5 if mode == "binary" else T
    codes, counts = oracle.classify(field, Tc, cells=cells, size=lattice)
    assert np.array_equal(res["codes"], codes), "cell codes differ"
    lay = oracle.layout(lattice, cells)
    assert tuple(res["layout"]) == lay
    l4 = oracle.level4(counts, cells, lay)
    assert np.array_equal(res["level4"], l4), "level-4 sums differ"
    cell, slot = oracle.vertex_list(counts, cells, lay)
    assert res["n"] == len(cell)
    assert np.array_equal(res["ce"][:, 0], cell), "vertex cells differ (order or content)"
    assert np.array_equal(res["ce"][:, 1], oracle.edge_ids(codes, cell, slot)), "vertex edges differ"
    got = res["verts"].astype(np.float64)
    sc = np.asarray(ext, np.float64) / np.asarray(cells, np.float64)
    if mode == "binary":
        ref = oracle.emit_binary(codes, cell, slot, cells, ext=ext)
        _check_positions(got, ref, sc)
        dn = np.abs(got[:, :3] - ref[:, :3])
        assert (dn <= 2.0 ** -23 * np.abs(ref[:, :3]) + 1e-30).all(), "binary normal error %g" % float(dn.max())
        return len(cell)
    if mode == "gradient":
        ref, t, mag = oracle.emit_gradient(field4, codes, cell, slot, T, ext=ext, cells=cells, size=lattice,
                                           clamp=clamp)
    else:
        ref, t, mag = oracle.emit(field, codes, cell, slot, T, ext=ext, cells=cells, size=lattice, clamp=clamp,
                                  with_mag=True)
    _check_positions(got, ref, sc)
    dn = np.abs(got[:, :3] - ref[:, :3])
    tol_n = 16 * EPS * mag[:, None] * sc[None, :]
    assert (dn <= tol_n).all(), "normal error exceeds FP32 bound by %g" % float((dn - tol_n).max())
    assert ((t >= 0) & (t <= 1)).all()
    return len(cell)


def smooth_noise(shape, seed):
    """Random field with structure at a few cells (upsampled coarse noise),
    shaped (z, y, x) for a lattice `shape` = (x, y, z)."""
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


def with_gradient(f):
    """(z, y, x, 4) float32: central-difference gradient (one-sided at the
    faces) in xyz and the value in w -- an RGBA gradient volume."""
    f = np.asarray(f, np.float32)
    gz, gy, gx = np.gradient(f.astype(np.float64))
    out = np.empty(f.shape + (4,), np.float32)
    out[..., 0], out[..., 1], out[..., 2], out[..., 3] = gx, gy, gz, f
    return out
