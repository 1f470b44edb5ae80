"""The batched and differentiation workloads at their full sizes (needs an
MI355X).

Batched: 256 metaball FP32 volumes, each axis ~U{16..128}, 4-32 balls
(seed 0, hpmc_amd.workloads.batch_volume_specs), iso 1.  One batched build +
extraction against 256 plain HistoPyramid builds of the same volumes: the
same counts, the same (cell, edge) lists and the same vertex bits for every
volume, under the automatic launch shapes, after autotuning and on a graph
replay; four volumes (the largest among them) also against the CPU oracle.

Differentiation: 64 noisy-sphere FP32 volumes, side ~U{32..128}, f = |p -
1/2| - 0.3 + N(0, 0.01^2), iso 0, loss = mean over all vertices of (|v - c|
- 0.35)^2 with c the sphere centre (0.5, 0.5, 0.5) in HPMC's [0, extent]
positions.  The batched adjoint against the FP64 oracle adjoint of single
volumes (three, one of side >= 120) within 1e-4 relative, bit-identical to
single-volume adjoints, and bit-identical from run to run."""
import numpy as np
import pytest

import oracle
from hpmc_amd import Batch, IsoSurface, workloads
from parity_util import check

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c3():
    specs = workloads.batch_volume_specs(count=256, seed=0)
    dev = torch.device("cuda", 0)
    vols = [workloads.metaballs_torch(size, balls, dev) for size, balls in specs]
    torch.cuda.synchronize()
    return specs, vols


def _plain(size, vol, T):
    iso = IsoSurface(size)
    iso.set_volume(vol)
    v, ce = iso.extract_all(T, with_cell_edge=True)
    v, ce = v.clone(), ce.clone()
    torch.cuda.synchronize()
    iso.close()
    return v, ce


def _assert_batch_equals_plain(out, ce, counts, bases, plain):
    for i, (pv, pce) in enumerate(plain):
        n, b0 = int(counts[i]), int(bases[i])
        assert n == pv.shape[0], (i, n, pv.shape[0])
        assert torch.equal(ce[b0:b0 + n], pce), "volume %d: (cell, edge) list differs" % i
        assert torch.equal(out[b0:b0 + n], pv), "volume %d: vertex bits differ" % i


def test_c3_batch_of_256_equals_plain_builds(c3):
    specs, vols = c3
    T = 1.0
    print("batched workload: %d volumes, %d lattice points" % (len(vols), sum(v.numel() for v in vols)), flush=True)
    plain = [_plain(size, v, T) for (size, _), v in zip(specs, vols)]
    b = Batch()
    for v in vols:
        b.add_volume(v)
    b.build(T)
    counts, bases, total = b.counts()
    assert total == sum(p[0].shape[0] for p in plain) > 10_000_000
    out = torch.full((total, 6), float("nan"), device="cuda")
    ce = torch.full((total, 2), -1, dtype=torch.int32, device="cuda")
    b.extract(out, ce)
    torch.cuda.synchronize()
    _assert_batch_equals_plain(out, ce, counts, bases, plain)
    # autotuned shapes and a graph replay give the same bits
    b.autotune(T, out, repetitions=1)
    b.capture_graph(T, out, ce)
    out.fill_(float("nan"))
    ce.fill_(-1)
    torch.cuda.synchronize()
    b.launch_graph()
    counts2, bases2, total2 = b.counts()
    torch.cuda.synchronize()
    assert total2 == total and np.array_equal(counts2, counts)
    _assert_batch_equals_plain(out, ce, counts2, bases2, plain)
    # four volumes against the oracle: the largest, the smallest, two others
    cells = [(s[0] - 1) * (s[1] - 1) * (s[2] - 1) for s, _ in specs]
    pick = sorted({int(np.argmax(cells)), int(np.argmin(cells)), 17, 200})
    for i in pick:
        size = specs[i][0]
        n, b0 = int(counts[i]), int(bases[i])
        f = vols[i].cpu().numpy()
        iso = IsoSurface(size)
        iso.set_volume(vols[i])
        iso.build(T)
        res = dict(n=n, verts=out[b0:b0 + n].cpu().numpy(), ce=ce[b0:b0 + n].cpu().numpy().view(np.uint32),
                   codes=iso.cell_codes().cpu().numpy(), level4=iso.level4().cpu().numpy().view(np.uint32),
                   layout=iso.layout())
        iso.close()
        print("oracle check: volume %d %s, %d vertices" % (i, size, n), flush=True)
        assert check(res, f, T, size) == n
    b.close()


@pytest.fixture(scope="module")
def c4():
    rng = workloads.SplitMix64(0)
    sides = [rng.randint(32, 128) for _ in range(64)]
    fields = [workloads.noisy_sphere((n, n, n), 1 + i) for i, n in enumerate(sides)]
    return sides, fields


def _loss_grad(v):
    centre = torch.full((3,), 0.5, dtype=torch.float32, device=v.device)
    v = v.detach().requires_grad_(True)
    loss = ((v[:, 3:] - centre).norm(dim=1) - 0.35).pow(2).mean()
    (g,) = torch.autograd.grad(loss, v)
    return g


def test_c4_batched_adjoint_of_64_volumes(c4):
    sides, fields = c4
    T = 0.0
    print("differentiation workload: sides %d..%d" % (min(sides), max(sides)), flush=True)
    dev = torch.device("cuda", 0)
    vols = [torch.from_numpy(f).to(dev) for f in fields]
    b = Batch()
    for v in vols:
        b.add_volume(v)
    b.build(T)
    counts, bases, total = b.counts()
    out = torch.empty((total, 6), device=dev)
    b.extract(out)
    gv = _loss_grad(out)
    grads, gt = b.adjoint(gv)
    grads = [g.clone() for g in grads]
    gt = gt.clone()
    grads2, gt2 = b.adjoint(gv)
    torch.cuda.synchronize()
    assert torch.equal(gt, gt2) and all(torch.equal(x, y) for x, y in zip(grads, grads2))
    # three volumes: the largest (side >= 120), the smallest, one between;
    # each equals its single-volume adjoint bit for bit and the FP64 oracle
    # adjoint within 1e-4 relative
    big = int(np.argmax(sides))
    assert sides[big] >= 120
    for i in sorted({big, int(np.argmin(sides)), 31}):
        n, b0 = int(counts[i]), int(bases[i])
        g_i = gv[b0:b0 + n].contiguous()
        iso = IsoSurface((sides[i],) * 3)
        iso.set_volume(vols[i])
        v1, _ = iso.extract_all(T)
        assert torch.equal(v1, out[b0:b0 + n])
        gf1, _ = iso.adjoint(g_i)
        torch.cuda.synchronize()
        assert torch.equal(gf1, grads[i]), "volume %d: batched and single adjoints differ" % i
        iso.close()
        f = fields[i]
        cells = (sides[i] - 1,) * 3
        codes, cnt = oracle.classify(f, T)
        cell, slot = oracle.vertex_list(cnt, cells, oracle.layout((sides[i],) * 3, cells))
        assert len(cell) == n
        ref_f, _ = oracle.adjoint(f, codes, cell, slot, T, g_i.cpu().numpy().astype(np.float64))
        got = grads[i].cpu().numpy().astype(np.float64)
        rel = float(np.abs(got - ref_f).max() / np.abs(ref_f).max())
        print("adjoint check: volume %d side %d, %d vertices, rel err %.2e" % (i, sides[i], n, rel), flush=True)
        assert rel < 1e-4, rel
    assert torch.isfinite(gt)
    b.close()
