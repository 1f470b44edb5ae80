"""Build + extraction without allocation (HPMCextractVolumeNoAlloc) against
the existing path (HPMCbuildHistopyramid + HPMCextractVertices), both
through the C ABI: the same count and the same vertex bits on calls in a
row that shrink the lattice (the scratch keeps the previous call's
contents), capacity clipping, and sizes past the scratch rejected before
anything is enqueued."""
import numpy as np
import pytest

import hpmc_amd
from hpmc_amd import IsoSurface, workloads
from parity_util import smooth_noise

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _reference(vol, T, ext):
    z, y, x = vol.shape
    iso = IsoSurface((x, y, z), extent=ext)
    iso.set_volume(vol)
    v, _ = iso.extract_all(T)
    v = v.clone()
    torch.cuda.synchronize()
    iso.close()
    return v


def _no_alloc(vol, T, capacity, ext):
    out = torch.full((capacity, 6), float("nan"), device="cuda")
    cnt = torch.full((1,), -1, dtype=torch.int32, device="cuda")
    hpmc_amd.extract_volume_no_alloc(vol, T, out, cnt, extent=ext)
    torch.cuda.synchronize()
    return out, int(cnt.item())


def _cayley_volume(n):
    from bench import cayley_volume_torch
    return cayley_volume_torch(n, torch.device("cuda", 0))


def test_no_alloc_equals_build_and_extract_on_shrinking_calls():
    cases = [
        ("cayley 512^3", lambda: _cayley_volume(512), 0.0, (1.0, 1.0, 1.0)),
        ("cayley 160x144x128", lambda: torch.from_numpy(workloads.cayley((160, 144, 128))).cuda(), 0.0,
         (1.0, 1.0, 1.0)),
        ("noise 45x38x33", lambda: torch.from_numpy(smooth_noise((45, 38, 33), 3)).cuda(), 0.5, (2.0, 1.0, 0.5)),
        ("white noise 5x9x40 (4x4-cell blocks)",
         lambda: torch.from_numpy(np.random.default_rng(1).random((40, 9, 5)).astype(np.float32)).cuda(), 0.5,
         (1.0, 1.0, 1.0)),
    ]
    for name, make, T, ext in cases:
        print("no-alloc case:", name, flush=True)
        vol = make()
        ref = _reference(vol, T, ext)
        n = ref.shape[0]
        assert n > 0
        out, cnt = _no_alloc(vol, T, n + 64, ext)
        assert cnt == n, (name, cnt, n)
        assert torch.equal(out[:n], ref), name
        assert torch.isnan(out[n:]).all(), name
        del vol
        torch.cuda.empty_cache()


def test_no_alloc_clips_at_capacity_and_rejects_oversized_lattices():
    vol = torch.from_numpy(smooth_noise((64, 60, 56), 8)).cuda()
    ref = _reference(vol, 0.5, (1.0, 1.0, 1.0))
    n = ref.shape[0]
    half = n // 2
    out, cnt = _no_alloc(vol, 0.5, half, (1.0, 1.0, 1.0))
    assert cnt == n
    assert torch.equal(out, ref[:half])
    # 2048 x 2048 x 200 needs several hundred MB of work items: rejected on
    # the host before anything is enqueued (the pointer is never read)
    out = torch.zeros((16, 6), device="cuda")
    code = hpmc_amd.lib().HPMCextractVolumeNoAlloc(hpmc_amd.constants(), vol.data_ptr(), 2048, 2048, 200,
                                                   1.0, 1.0, 1.0, 0.5, out.data_ptr(), 16, None, None)
    assert code == 3, code                     # HPMC_ERR_SIZE
    assert "scratch" in hpmc_amd.last_error()[1]
    # the scratch still works after the rejection
    out, cnt = _no_alloc(vol, 0.5, n, (1.0, 1.0, 1.0))
    assert cnt == n and torch.equal(out, ref)
    np.testing.assert_array_equal(out.cpu().numpy(), ref.cpu().numpy())
