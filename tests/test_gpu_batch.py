"""Batched builds (HPMCBatch) on the GPU: many metaball volumes of uneven
sizes in one classification, one scan and one extraction launch.  The bar:
every volume's count, (cell, edge) list and vertex bits equal those of a
plain HistoPyramid build of that volume alone -- under every launch shape
the autotuner may pick, and on every replay of a captured hipGraph."""
import numpy as np
import pytest

import hpmc_amd
from hpmc_amd import Batch, IsoSurface, workloads
from parity_util import check

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

ISO = 1.0


@pytest.fixture(scope="module")
def volumes():
    specs = workloads.batch_volume_specs(count=12, lo=16, hi=80, balls_lo=4, balls_hi=16, seed=3)
    out = []
    for size, balls in specs:
        f = workloads.metaballs(size, balls)
        out.append((size, f, torch.from_numpy(f).cuda()))
    return out


@pytest.fixture(scope="module")
def plain(volumes):
    """Per volume: (vertices, cell_edge) of a plain build."""
    res = []
    for size, _, vol in volumes:
        iso = IsoSurface(size)
        iso.set_volume(vol)
        v, ce = iso.extract_all(ISO, with_cell_edge=True)
        res.append((v.clone(), ce.clone()))
        iso.close()
    torch.cuda.synchronize()
    return res


def _batch_run(b, total):
    out = torch.full((max(total, 1), 6), float("nan"), dtype=torch.float32, device="cuda")
    ce = torch.full((max(total, 1), 2), -1, dtype=torch.int32, device="cuda")
    b.extract(out, ce)
    torch.cuda.synchronize()
    return out, ce


def _assert_same_as_plain(out, ce, counts, bases, plain):
    for i, (pv, pce) in enumerate(plain):
        n, b0 = int(counts[i]), int(bases[i])
        assert n == pv.shape[0], (i, n, pv.shape[0])
        assert torch.equal(ce[b0:b0 + n], pce), "volume %d: (cell, edge) list differs" % i
        assert torch.equal(out[b0:b0 + n], pv), "volume %d: vertex bits differ" % i


def _new_batch(volumes):
    b = Batch()
    for _, _, vol in volumes:
        b.add_volume(vol)
    return b


def test_metaballs_torch_matches_numpy(volumes):
    size, f, _ = volumes[0]
    rng = workloads.SplitMix64(3)
    balls = workloads.metaball_params(rng, 8)
    a = workloads.metaballs(size, balls)
    t = workloads.metaballs_torch(size, balls, "cuda").cpu().numpy()
    assert np.array_equal(a, t)


def test_batch_matches_plain_builds(volumes, plain):
    b = _new_batch(volumes)
    b.build(ISO)
    counts, bases, total = b.counts()
    assert total == sum(p[0].shape[0] for p in plain) > 0
    assert (bases == np.concatenate([[0], np.cumsum(counts)[:-1]])).all()
    out, ce = _batch_run(b, total)
    _assert_same_as_plain(out, ce, counts, bases, plain)
    b.close()


def test_batch_volumes_against_oracle(volumes):
    # two volumes of the batch checked directly against the CPU oracle
    b = _new_batch(volumes)
    b.build(ISO)
    counts, bases, total = b.counts()
    out, ce = _batch_run(b, total)
    for i in (0, len(volumes) - 1):
        size, f, _ = volumes[i]
        n, b0 = int(counts[i]), int(bases[i])
        iso = IsoSurface(size)
        iso.set_volume(volumes[i][2])
        iso.build(ISO)
        res = dict(n=n, verts=out[b0:b0 + n].cpu().numpy(), ce=ce[b0:b0 + n].cpu().numpy().view(np.uint32),
                   codes=iso.cell_codes().cpu().numpy(), level4=iso.level4().cpu().numpy().view(np.uint32),
                   layout=iso.layout())
        iso.close()
        assert check(res, f, ISO, size) == n
    b.close()


@pytest.mark.parametrize("waves,zseg,emit", [(1, 8, 8), (1, 32, 64), (2, 16, 0), (2, 8, 256), (0, 32, 24),
                                            (3, 16, 0), (8, 16, 0), (16, 8, 40)])
def test_batch_every_tuning_identical(volumes, plain, waves, zseg, emit):
    b = _new_batch(volumes)
    b.set_tuning(waves, zseg, emit)
    b.build(ISO)
    counts, bases, total = b.counts()
    out, ce = _batch_run(b, total)
    _assert_same_as_plain(out, ce, counts, bases, plain)
    got = b.tuning()
    assert got[1] == zseg and (waves == 0 or got[0] == waves)
    b.close()


def test_batch_graph_replays_match_eager(volumes, plain):
    b = _new_batch(volumes)
    b.build(ISO)
    counts, bases, total = b.counts()
    g_out = torch.zeros((total, 6), dtype=torch.float32, device="cuda")
    g_ce = torch.zeros((total, 2), dtype=torch.int32, device="cuda")
    b.capture_graph(ISO, g_out, g_ce)
    for rep in range(3):
        g_out.fill_(float("nan"))
        g_ce.fill_(-1)
        torch.cuda.synchronize()
        b.launch_graph()
        c2, b2, t2 = b.counts()
        assert t2 == total and (c2 == counts).all()
        _assert_same_as_plain(g_out, g_ce, counts, bases, plain)
        # eager work between replays
        b.build(ISO)
        out, ce = _batch_run(b, total)
        _assert_same_as_plain(out, ce, counts, bases, plain)
    b.close()


def test_batch_autotune_keeps_results(volumes, plain):
    b = _new_batch(volumes)
    b.build(ISO)
    _, _, total = b.counts()
    scratch = torch.empty((total, 6), dtype=torch.float32, device="cuda")
    shape, ms = b.autotune(ISO, scratch, repetitions=2)
    assert ms > 0 and shape[1] in (8, 16, 32)
    b.build(ISO)
    counts, bases, total2 = b.counts()
    assert total2 == total
    out, ce = _batch_run(b, total)
    _assert_same_as_plain(out, ce, counts, bases, plain)
    b.close()


def test_batch_capacity(volumes):
    b = _new_batch(volumes)
    b.build(ISO)
    counts, bases, total = b.counts()
    small = torch.zeros((total - 7, 6), dtype=torch.float32, device="cuda")
    with pytest.raises(hpmc_amd.HPMCError, match="CAPACITY"):
        b.extract(small)
    # without a count the extraction drops what does not fit
    b.build(ISO)
    b.extract(small)
    full = torch.zeros((total, 6), dtype=torch.float32, device="cuda")
    b.extract(full)
    torch.cuda.synchronize()
    assert torch.equal(small, full[: total - 7])
    b.close()


def test_single_volume_batch_equals_isosurface(volumes, plain):
    b = Batch()
    b.add_volume(volumes[5][2])
    b.build(ISO)
    counts, bases, total = b.counts()
    out, ce = _batch_run(b, total)
    assert torch.equal(out[:total], plain[5][0]) and torch.equal(ce[:total], plain[5][1])
    b.close()
