"""8-bit normalised volumes (HPMCsetFieldTexture3DUnorm8; the field of the
reference's texture3d demo, apps/texture3d/texture3d.cpp:127-141, a
GL_ALPHA8 texture whose byte c samples as c / 255).  The bar: the same
codes, count, (cell, edge) list and vertex bits as an FP32 volume holding
c / 255 (correctly rounded), in continuous and binary mode, at thresholds
that fall exactly on a byte's value, and the oracle's topology and
tolerances.  The exact float llvmpipe gives c / 255 is not pinned here (no
reference run): the conversion is the GL specification's c / 255."""
import numpy as np
import pytest

from hpmc_amd import IsoSurface
from parity_util import check, gpu_run, smooth_noise

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _bytes(lattice, seed):
    return np.clip(np.rint(smooth_noise(lattice, seed) * 255.0), 0, 255).astype(np.uint8)


def _as_float(b):
    return b.astype(np.float32) / np.float32(255.0)


def _run(lattice, vol, T, binary=False):
    iso = IsoSurface(lattice)
    if binary:
        iso.set_binary(True)
    iso.set_volume(vol)
    v, ce = iso.extract_all(T, with_cell_edge=True)
    codes = iso.cell_codes().clone()
    iso.close()
    return v, ce, codes


@pytest.mark.parametrize("lattice,T", [((64, 48, 40), 0.5), ((97, 33, 70), float(np.float32(100) / np.float32(255))),
                                       ((40, 40, 40), float(np.float32(37) / np.float32(255)) + 1e-7)])
def test_unorm8_equals_fp32_volume(lattice, T):
    b = _bytes(lattice, 11)
    v8, ce8, c8 = _run(lattice, torch.from_numpy(b).cuda(), T)
    vf, cef, cf = _run(lattice, torch.from_numpy(_as_float(b)).cuda(), T)
    assert v8.shape[0] > 100
    assert torch.equal(c8, cf)
    assert torch.equal(ce8, cef)
    assert torch.equal(v8, vf)


@pytest.mark.parametrize("lattice", [(16384, 10, 2), (11, 900, 5), (12, 12, 700)])
def test_unorm8_extreme_shapes(lattice):
    # byte classification over the longest rows and the most slices
    b = _bytes(lattice, 15)
    v8, ce8, c8 = _run(lattice, torch.from_numpy(b).cuda(), 0.5)
    vf, cef, cf = _run(lattice, torch.from_numpy(_as_float(b)).cuda(), 0.5)
    assert v8.shape[0] > 100
    assert torch.equal(c8, cf) and torch.equal(ce8, cef) and torch.equal(v8, vf)


def test_unorm8_binary_equals_fp32_volume():
    lattice = (50, 60, 45)
    b = _bytes(lattice, 12)
    v8, ce8, _ = _run(lattice, torch.from_numpy(b).cuda(), 0.3, binary=True)
    vf, cef, _ = _run(lattice, torch.from_numpy(_as_float(b)).cuda(), 0.3, binary=True)
    assert v8.shape[0] > 100 and torch.equal(ce8, cef) and torch.equal(v8, vf)


def test_unorm8_against_oracle():
    lattice = (72, 56, 64)
    b = _bytes(lattice, 13)
    res = gpu_run(lattice, 0.5, volume=b)
    assert check(res, _as_float(b), 0.5, lattice) > 1000


def test_unorm8_setter_switches_kind():
    # a float volume after an 8-bit one (and back) reconfigures the build
    lattice = (40, 40, 40)
    b = _bytes(lattice, 14)
    f = torch.from_numpy(_as_float(b)).cuda()
    u = torch.from_numpy(b).cuda()
    iso = IsoSurface(lattice)
    iso.set_volume(u)
    a, _ = iso.extract_all(0.5)
    iso.set_volume(f)
    c, _ = iso.extract_all(0.5)
    iso.set_volume(u)
    d, _ = iso.extract_all(0.5)
    assert torch.equal(a, c) and torch.equal(a, d)
    iso.close()
