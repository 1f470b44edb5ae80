"""API behaviour on the GPU beyond parity: stream hand-off, a field volume
reset to NULL, bounded profiling events, graph replays at their captured
threshold, argument checks of the adjoint, the autograd wrappers'
handling of in-place edits and repeated backward passes, and sizes beyond
the engine's offsets failing with an error code and recovering."""
import ctypes

import numpy as np
import pytest

import hpmc_amd
from hpmc_amd import Batch, IsoSurface, workloads
from hpmc_amd.autograd import marching_cubes, marching_cubes_batch
from parity_util import smooth_noise

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def test_set_stream_orders_new_stream_after_old():
    # a dense 256^3 build on stream A, then the stream switched to B and the
    # extraction queued there at once: B waits for A's build, so the result
    # equals a one-stream run bit for bit
    n = 256
    vol = torch.from_numpy(workloads.upsampled_noise(n)).cuda()
    ref = IsoSurface((n, n, n))
    ref.set_volume(vol)
    rv, rce = ref.extract_all(0.5, with_cell_edge=True)
    nv = rv.shape[0]
    a, b = torch.cuda.Stream(), torch.cuda.Stream()
    iso = IsoSurface((n, n, n), stream=a)
    iso.set_volume(vol)
    iso.build(0.5)
    assert iso.count() == nv
    out = torch.full((nv, 6), float("nan"), device="cuda")
    ce = torch.full((nv, 2), -1, dtype=torch.int32, device="cuda")
    iso.bind_output(out, ce)
    for _ in range(3):
        iso.build(0.5)                                 # queued on A, no host wait
        hpmc_amd._check(hpmc_amd.lib().HPMCsetStream(iso.h, hpmc_amd._stream_ptr(b)), iso.h)
        iso.extract()                                  # queued on B
        b.synchronize()
        assert torch.equal(out, rv) and torch.equal(ce, rce)
        out.fill_(float("nan"))
        torch.cuda.synchronize()
        hpmc_amd._check(hpmc_amd.lib().HPMCsetStream(iso.h, hpmc_amd._stream_ptr(a)), iso.h)
    torch.cuda.synchronize()
    iso.close()
    ref.close()


def test_count_waits_for_the_last_build_whatever_was_queued_after_it():
    # The count's event is recorded when the count is asked for: it must see
    # the last build's totals after an extraction queued behind the build,
    # after two builds in a row at different thresholds, and after the
    # stream was switched between the build and the count.
    n = 96
    vol = torch.from_numpy(workloads.upsampled_noise(n)).cuda()
    counts = {}
    for iso_value in (0.4, 0.6):
        r = IsoSurface((n, n, n))
        r.set_volume(vol)
        r.build(iso_value)
        counts[iso_value] = r.count()
        r.close()
    assert counts[0.4] != counts[0.6] and min(counts.values()) > 0
    a, b = torch.cuda.Stream(), torch.cuda.Stream()
    iso = IsoSurface((n, n, n), stream=a)
    iso.set_volume(vol)
    out = torch.empty((max(counts.values()), 6), device="cuda")
    iso.bind_output(out)
    iso.build(0.4)
    iso.extract()                                      # queued behind the build, no count yet
    assert iso.count() == counts[0.4]
    iso.build(0.4)
    iso.build(0.6)                                     # two builds, no count in between
    assert iso.count() == counts[0.6]
    iso.build(0.4)
    hpmc_amd._check(hpmc_amd.lib().HPMCsetStream(iso.h, hpmc_amd._stream_ptr(b)), iso.h)
    assert iso.count() == counts[0.4]                  # recorded on B, which waits for A
    torch.cuda.synchronize()
    iso.close()


def test_scan_sums_survive_builds_without_extraction():
    # Classification adds each build's counts into the scan's tile sums; the
    # extraction that follows clears them, and a build that follows no
    # extraction clears them first.  Any mix of builds and extractions must
    # count and list exactly what one fresh build does.
    n = 80
    vol = torch.from_numpy(workloads.upsampled_noise(n)).cuda()
    ref = IsoSurface((n, n, n))
    ref.set_volume(vol)
    rv, _ = ref.extract_all(0.5)
    ref.close()
    iso = IsoSurface((n, n, n))
    iso.set_volume(vol)
    for _ in range(4):
        iso.build(0.5)                     # builds with no extraction in between
    assert iso.count() == rv.shape[0]
    out = torch.empty((rv.shape[0], 6), device="cuda")
    iso.bind_output(out)
    for _ in range(2):
        iso.extract()                      # two extractions of one build
    torch.cuda.synchronize()
    assert torch.equal(out, rv)
    iso.build(0.5)
    iso.build(0.5)
    out.fill_(float("nan"))
    iso.extract()
    torch.cuda.synchronize()
    assert iso.count() == rv.shape[0] and torch.equal(out, rv)
    iso.close()
    # a batch: bare builds, then its graph twice, then an eager extraction
    specs = workloads.batch_volume_specs(count=4, lo=24, hi=40, balls_lo=3, balls_hi=6, seed=5)
    vols = [torch.from_numpy(workloads.metaballs(s, b)).cuda() for s, b in specs]
    b = Batch()
    for v in vols:
        b.add_volume(v)
    b.build(1.0)
    _, _, total = b.counts()
    want = torch.empty((total, 6), device="cuda")
    b.extract(want)
    torch.cuda.synchronize()
    g_out = torch.empty_like(want)
    b.capture_graph(1.0, g_out)
    for _ in range(3):
        b.build(1.0)
    assert b.counts()[2] == total
    for _ in range(2):
        g_out.fill_(float("nan"))
        b.launch_graph()
        assert b.counts()[2] == total
        torch.cuda.synchronize()
        assert torch.equal(g_out, want)
    again = torch.full_like(want, float("nan"))
    b.extract(again)
    torch.cuda.synchronize()
    assert torch.equal(again, want)
    b.close()


def test_null_volume_after_build_is_a_field_error():
    lattice = (40, 36, 33)
    f = smooth_noise(lattice, 5)
    iso = IsoSurface(lattice)
    vol = torch.from_numpy(f).cuda()
    iso.set_volume(vol)
    iso.build(0.5)
    assert iso.count() > 0
    L = hpmc_amd.lib()
    L.HPMCsetFieldTexture3D(iso.h, None, 0)       # same mode: no reconfiguration
    L.HPMCbuildHistopyramid(iso.h, 0.5)
    assert hpmc_amd.last_error(iso.h)[0] == 4      # HPMC_ERR_FIELD, nothing launched
    # the failed build latched the HistoPyramid as broken until reconfigured
    assert L.HPMCacquireNumberOfVertices(iso.h) == 0
    codes = torch.empty(39 * 35 * 32, dtype=torch.uint8, device="cuda")
    L.HPMCsetLatticeSize(iso.h, *lattice)          # reconfigure: no longer broken
    assert L.HPMCgetCellCodes(iso.h, ctypes.c_void_p(codes.data_ptr())) == 4
    iso.set_volume(vol)
    verts, _ = iso.extract_all(0.5)
    assert verts.shape[0] > 0
    iso.close()


def test_profiling_events_stay_bounded_and_count_every_launch():
    lattice = (24, 20, 18)
    iso = IsoSurface(lattice)
    iso.set_volume(torch.from_numpy(smooth_noise(lattice, 2)).cuda())
    iso.build(0.5)
    out = torch.empty((max(iso.count(), 1), 6), device="cuda")
    iso.bind_output(out)
    iso.stage_times()
    iso.set_profiling(True)
    launches = 1100                               # past the 1024 pairs a stage holds
    for _ in range(launches):
        iso.build(0.5)
        iso.extract()
    iso.set_profiling(False)
    st = iso.stage_times()
    assert all(st[k][1] == launches for k in ("classify", "scan", "emit")), st
    assert all(st[k][0] > 0 for k in st)
    assert iso.stage_times() == {"classify": (0.0, 0), "scan": (0.0, 0), "emit": (0.0, 0)}
    iso.close()


def test_stage_times_of_a_snippet_field():
    # the hipRTC kernels of a snippet field are timed through their own
    # launches too (the module launch with events)
    lattice = (40, 36, 33)
    iso = IsoSurface(lattice)
    iso.set_snippet(workloads.CAYLEY_SNIPPET)
    iso.build(0.0)
    n = iso.count()
    assert n > 0
    out = torch.empty((n, 6), device="cuda")
    iso.bind_output(out)
    iso.stage_times()
    iso.set_profiling(True)
    for _ in range(3):
        iso.build(0.0)
        iso.extract()
    iso.set_profiling(False)
    st = iso.stage_times()
    assert all(st[k][1] == 3 and st[k][0] > 0 for k in ("classify", "scan", "emit")), st
    ref = torch.empty_like(out)
    iso.bind_output(ref)
    iso.build(0.0)
    iso.extract()
    torch.cuda.synchronize()
    assert torch.equal(out, ref)          # profiled launches compute the same
    iso.close()


def test_graph_replay_extracts_at_the_captured_threshold():
    specs = workloads.batch_volume_specs(count=6, lo=16, hi=48, balls_lo=4, balls_hi=8, seed=11)
    vols = [torch.from_numpy(workloads.metaballs(s, b)).cuda() for s, b in specs]
    b = Batch()
    for v in vols:
        b.add_volume(v)
    T1, T2 = 1.0, 1.7
    # plain reference at T1
    plain = []
    for (size, _), v in zip(specs, vols):
        iso = IsoSurface(size)
        iso.set_volume(v)
        pv, _ = iso.extract_all(T1)
        plain.append(pv.clone())
        iso.close()
    b.build(T1)
    _, _, total = b.counts()
    out = torch.empty((total, 6), device="cuda")
    b.capture_graph(T1, out)
    b.build(T2)                      # eager build at another threshold
    b.counts()
    b.launch_graph()                 # replays the T1 build + extraction
    counts, bases, tot = b.counts()
    assert tot == total
    again = torch.full_like(out, float("nan"))
    b.extract(again)                 # an eager extraction of the replayed build: at T1
    torch.cuda.synchronize()
    for i, pv in enumerate(plain):
        s, n = int(bases[i]), int(counts[i])
        assert torch.equal(out[s:s + n], pv) and torch.equal(again[s:s + n], pv), i
    b.close()


def test_adjoint_rejects_bad_vertex_gradients():
    lattice = (30, 30, 30)
    iso = IsoSurface(lattice)
    iso.set_volume(torch.from_numpy(smooth_noise(lattice, 9)).cuda())
    verts, _ = iso.extract_all(0.5)
    n = verts.shape[0]
    assert n > 0
    for bad in (torch.zeros((n - 1, 6), device="cuda"), torch.zeros((n, 6), dtype=torch.float64, device="cuda"),
                torch.zeros((n, 6)), torch.zeros((n * 6,), device="cuda"), None):
        with pytest.raises(ValueError):
            iso.adjoint(bad)
    gf, gt = iso.adjoint(torch.ones((n, 6), device="cuda"))
    assert torch.isfinite(gf).all() and torch.isfinite(gt)
    iso.close()


def test_autograd_detects_in_place_edits_and_allows_retained_graphs():
    lattice = (32, 28, 30)
    f = torch.from_numpy(smooth_noise(lattice, 4)).cuda().requires_grad_(True)
    field = f * 1.0
    v = marching_cubes(field, 0.5)
    loss = v[:, 3:].pow(2).sum()
    (g1,) = torch.autograd.grad(loss, f, retain_graph=True)
    (g2,) = torch.autograd.grad(loss, f)
    assert torch.equal(g1, g2)                      # the second backward reuses the surface
    field2 = f * 1.0
    v2 = marching_cubes(field2, 0.5)
    with torch.no_grad():
        field2.add_(1e-3)                           # changes the field under the surface
    with pytest.raises(RuntimeError):
        v2[:, 3:].sum().backward()
    fields = [torch.from_numpy(smooth_noise(lattice, s)).cuda().requires_grad_(True) for s in (1, 2)]
    parts = marching_cubes_batch([x * 1.0 for x in fields], 0.5)
    loss_b = sum(p[:, 3:].pow(2).sum() for p in parts)
    gb1 = torch.autograd.grad(loss_b, fields, retain_graph=True)
    gb2 = torch.autograd.grad(loss_b, fields)
    assert all(torch.equal(x, y) for x, y in zip(gb1, gb2))


def test_absurd_sizes_fail_cleanly_and_recover():
    # The largest lattice the reference accepts, 16384^3 (a snippet field, so
    # nothing of that size is ever allocated), has 2^34 level-4 keys, more
    # than this build's 32-bit offsets address: the build fails with an
    # error code before any allocation or launch, latches the HistoPyramid
    # broken and counts 0; a reconfiguration recovers it.
    n = 16384
    iso = IsoSurface((n, n, n))
    iso.set_snippet(workloads.CAYLEY_SNIPPET)
    L = hpmc_amd.lib()
    L.HPMCbuildHistopyramid(iso.h, 0.0)
    assert hpmc_amd.last_error(iso.h)[0] == 3          # HPMC_ERR_SIZE
    assert L.HPMCacquireNumberOfVertices(iso.h) == 0
    L.HPMCsetLatticeSize(iso.h, 48, 40, 36)
    L.HPMCbuildHistopyramid(iso.h, 0.0)
    assert hpmc_amd.last_error(iso.h)[0] == 0
    assert L.HPMCacquireNumberOfVertices(iso.h) > 0
    iso.close()
    # a batch volume of that size: refused at the batch's build, before any
    # launch (the pointer given is never read)
    small = torch.zeros((8, 8, 8), device="cuda")
    b = Batch()
    idx = L.HPMCbatchAddVolume(b.b, ctypes.c_void_p(small.data_ptr()), n, n, n, 1.0, 1.0, 1.0)
    if idx >= 0:
        assert L.HPMCbatchBuild(b.b, 0.0) == 3         # HPMC_ERR_SIZE
    else:
        assert -idx == 3
    b.close()
