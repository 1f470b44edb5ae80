"""Benchmark: HistoPyramid marching cubes of a 512^3 FP32 volume on one MI355X.

Workload: the Cayley cubic f = 4(x^2+y^2+z^2) + 16xyz - 1, x = 2p - 1 at
texel centres, materialised once into a 512^3 FP32 volume (512 MiB) resident
in HBM, iso-value 0: 133,432,831 cells, 3,694,104 vertices.  One step is one
HPMCbuildHistopyramid (classify + pyramid reduction) plus one
HPMCextractVertices writing positions and normals (N3F_V3F) for every vertex;
no host synchronisation inside a step.

Prints one JSON line.  value = cells/s over the whole step (the reference's
own "mvps" definition, apps/texture3d/texture3d.cpp:279-288); the line also
carries verts/s, the whole-step HBM fraction for the algorithmic bytes
B = 4 * lattice points + 24 * vertices, per-stage HIP-event times, the
roofline of the dominant kernel and the CPU baseline (the oracle's C port of
the same path, one thread, on this host).

The line's "batched" object is the batched/plain comparison: 256 metaball
volumes of uneven sizes (measure_batched), built and extracted as one
HPMCBatch (autotuned, replayed as a hipGraph) versus 256 plain builds.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cells/s + verts/s at 512^3 fp32, % HBM peak; batched/plain speedup"
HBM_PEAK_GBS = 8000.0
N_CAYLEY_512 = 3694104
# HBM traffic of the classification kernel per launch, from rocprofv3 --pmc
# FETCH_SIZE / WRITE_SIZE runs of this script, scaled by the same counters on
# a plain copy of known size (bench.py --copy-calibration)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hpmc_classify_512_traffic.json")


def cayley_volume_torch(n, device):
    """The Cayley field in FP32 with the same operation order as
    hpmc_amd.workloads.cayley (elementwise IEEE ops, bit-identical)."""
    import torch
    inv = torch.tensor(1.0, dtype=torch.float32) / torch.tensor(float(n), dtype=torch.float32)
    p = (torch.arange(n, dtype=torch.float32, device=device) + 0.5) * inv.to(device)
    x = (2.0 * p - 1.0).view(1, 1, n)
    y = (2.0 * p - 1.0).view(1, n, 1)
    z = (2.0 * p - 1.0).view(n, 1, 1)
    s = (x * x + y * y) + z * z
    xyz = ((16.0 * x) * y) * z
    return (4.0 * s + xyz - 1.0).contiguous()


def host_cpus():
    """CPUs this process may run on (the host's share), and the machine's."""
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return usable, os.cpu_count() or usable


def cpu_baseline(n, reps):
    """Oracle C port of classify -> ordered list -> FP32 emission, one thread,
    over the whole 512^3 workload, median of `reps` runs.  The line states the
    host's CPUs beside the one thread used."""
    import oracle
    from hpmc_amd import workloads
    f = workloads.cayley((n, n, n))
    times = []
    nv = 0
    for _ in range(reps):
        t0 = time.perf_counter()
        nv = oracle.run_volume(f, 0.0)
        times.append(time.perf_counter() - t0)
    t = sorted(times)[len(times) // 2]
    cells = (n - 1) ** 3
    usable, machine = host_cpus()
    return {"value": cells / t, "unit": "cells/s", "cores": 1, "kind": "port",
            "host_cpus_usable": usable, "host_cpus": machine,
            "sample": "full %d^3 Cayley volume, %d runs (median %.2f s), %d vertices" % (n, reps, t, nv)}


def cpu_baseline_threads(n, reps, threads=None):
    """The same port over z slabs on `threads` host threads (default: the
    usable CPUs, at most 16): each thread classifies its slab, lists its
    vertices in the full volume's order and emits them (ctypes releases the
    GIL).  The slabs' lists are not concatenated into the global order, so
    this is a throughput line beside the one-thread baseline, not a drop-in
    result."""
    import concurrent.futures
    import numpy as np
    import oracle
    from hpmc_amd import workloads
    usable, machine = host_cpus()
    threads = threads or max(1, min(16, usable))
    f = workloads.cayley((n, n, n))
    cz = n - 1
    bounds = [(i * cz // threads, (i + 1) * cz // threads) for i in range(threads)]
    oracle.lib()

    def slab(b):
        return oracle.run_volume(f, 0.0, z_begin=b[0], z_end=b[1], emit_out=True)

    times = []
    nv = 0
    with concurrent.futures.ThreadPoolExecutor(threads) as pool:
        for _ in range(reps):
            t0 = time.perf_counter()
            nv = sum(pool.map(slab, bounds))
            times.append(time.perf_counter() - t0)
    t = sorted(times)[len(times) // 2]
    return {"value": cz ** 3 / t, "unit": "cells/s", "cores": threads, "kind": "port",
            "host_cpus_usable": usable, "host_cpus": machine,
            "sample": "full %d^3 Cayley volume in %d z slabs on %d threads, %d runs (median %.3f s), %d vertices"
                      % (n, threads, threads, reps, t, int(np.int64(nv)))}


def measure_batched(reps, count=256, iso=1.0):
    """Batched vs plain builds of `count` metaball volumes of uneven sizes
    (each axis ~U{16..128}, 4-32 balls, seed 0), iso 1, positions + normals.

    plain:        per volume, the reference's call order -- build, count (a
                  host wait), extract into a buffer of that count's size;
    plain_async:  per volume build + extract into pre-sized buffers, one
                  wait at the end (the most favourable plain schedule);
    batched:      one HPMCBatch: autotuned launch shapes, build + extract
                  captured as a hipGraph, one replay + counts per step.
    Every variant produces the same vertices (the GPU tests prove it)."""
    import numpy as np
    import torch
    import hpmc_amd
    from hpmc_amd import workloads
    dev = torch.device("cuda", 0)
    specs = workloads.batch_volume_specs(count=count, seed=0)
    vols = [workloads.metaballs_torch(size, balls, dev) for size, balls in specs]
    torch.cuda.synchronize()
    points = sum(int(v.numel()) for v in vols)
    cells = sum((s[0] - 1) * (s[1] - 1) * (s[2] - 1) for s, _ in specs)

    # plain HistoPyramids, configured and warmed up outside the timing
    isos = []
    outs = []
    for (size, _), v in zip(specs, vols):
        h = hpmc_amd.IsoSurface(size)
        h.set_volume(v)
        h.build(iso)
        n = h.count()
        o = torch.empty((max(n, 1), 6), dtype=torch.float32, device=dev)
        h.bind_output(o)
        if n:
            h.extract()
        isos.append(h)
        outs.append(o)
    torch.cuda.synchronize()
    plain_vertices = sum(h.count() for h in isos)

    def plain_step():
        for h in isos:
            h.build(iso)
            if h.count():
                h.extract()
        torch.cuda.synchronize()

    def plain_async_step():
        for h in isos:
            h.build(iso)
            h.extract()
        torch.cuda.synchronize()

    b = hpmc_amd.Batch()
    for v in vols:
        b.add_volume(v)
    b.build(iso)
    counts, bases, total = b.counts()
    assert total == plain_vertices, (total, plain_vertices)
    out = torch.empty((total, 6), dtype=torch.float32, device=dev)
    tuning, tuned_ms = b.autotune(iso, out, repetitions=3)
    b.capture_graph(iso, out)

    def batched_step():
        b.launch_graph()
        b.counts()

    def batched_eager_step():
        b.build(iso)
        b.extract(out)
        b.counts()

    def timeit(fn):
        # each call ends with a host wait, so calls are timed one by one and
        # the median is reported (a mean of a few calls moves with one stall)
        fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append((time.perf_counter() - t0) * 1e3)
        return float(sorted(ts)[len(ts) // 2])

    res = {
        "workload": "%d metaball FP32 volumes, axes U{16..128}, 4-32 balls, seed 0, iso %g" % (count, iso),
        "volumes": count, "lattice_points": points, "cells": cells, "vertices": int(total),
        "plain_ms": timeit(plain_step), "plain_async_ms": timeit(plain_async_step),
        "batched_eager_ms": timeit(batched_eager_step), "batched_ms": timeit(batched_step),
        "tuning": {"classify_waves": tuning[0], "zseg": tuning[1], "emit_workgroups": tuning[2],
                   "autotuned_ms": tuned_ms},
    }
    res["speedup"] = res["plain_ms"] / res["batched_ms"]
    res["speedup_vs_plain_async"] = res["plain_async_ms"] / res["batched_ms"]
    res["batched_cells_per_s"] = cells / (res["batched_ms"] * 1e-3)
    for h in isos:
        h.close()
    b.close()
    return res


def batch_steps(steps, warmup, tuning=(8, 16, 1280), count=256, iso=1.0):
    """The batched workload's build + extract under a fixed launch shape, for
    kernel traces (bench.py --batch-trace): `warmup` then `steps` steps."""
    import torch
    import hpmc_amd
    from hpmc_amd import workloads
    dev = torch.device("cuda", 0)
    specs = workloads.batch_volume_specs(count=count, seed=0)
    vols = [workloads.metaballs_torch(size, balls, dev) for size, balls in specs]
    b = hpmc_amd.Batch()
    for v in vols:
        b.add_volume(v)
    b.set_tuning(*tuning)
    b.build(iso)
    _, _, total = b.counts()
    out = torch.empty((total, 6), dtype=torch.float32, device=dev)
    for _ in range(warmup):
        b.build(iso)
        b.extract(out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.build(iso)
        b.extract(out)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    b.close()
    print(json.dumps({"batch_trace": {"tuning": list(tuning), "vertices": int(total), "ms_per_step": ms}}))


def copy_calibration(steps):
    """A plain 512 MiB device copy (torch copy_), for scaling the HBM counters:
    each launch reads and writes exactly 536,870,912 bytes."""
    import torch
    x = torch.empty(128 * 1024 * 1024, dtype=torch.float32, device="cuda").fill_(1.0)
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        y.copy_(x)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    print(json.dumps({"copy_bytes_each_way": x.numel() * 4, "launches": steps, "ms_per_copy": ms,
                      "rw_GBps": 2 * x.numel() * 4 / (ms * 1e-3) / 1e9}))


def measure_pipelined(n, steps, warmup, inflight=3, zseg=0):
    """Frames in flight: `inflight` HistoPyramids on their own HIP streams,
    each over its own copy of the volume (no cache reuse between frames),
    called round robin -- frame k's extraction overlaps frame k+1's
    classification.  Every frame is a full build + extract.  zseg 0 keeps
    the one-stream launch shapes (at 512^3: 16-wave classification
    workgroups over 32-slice segments, which measured 131 us per frame
    against 134-135 for 16-slice segments).  Returns the steady-state time
    per frame (median of 5 runs of `steps` frames)."""
    import torch
    import hpmc_amd
    dev = torch.device("cuda", 0)
    base = cayley_volume_torch(n, dev)
    objs = []
    for _ in range(inflight):
        s = torch.cuda.Stream(device=dev)
        vol = base.clone()
        iso = hpmc_amd.IsoSurface((n, n, n), stream=s)
        iso.set_volume(vol)
        iso.set_tuning(0, zseg, 0)
        iso.build(0.0)
        nv = iso.count()
        out = torch.empty((nv, 6), dtype=torch.float32, device=dev)
        iso.bind_output(out)
        objs.append((iso, vol, out, s))
    del base

    def frames(k):
        for i in range(k):
            objs[i % inflight][0].build(0.0)
            objs[i % inflight][0].extract()

    frames(warmup + inflight)
    torch.cuda.synchronize()
    ts = []
    for _ in range(5):
        t0 = time.perf_counter()
        frames(steps)
        torch.cuda.synchronize()
        ts.append((time.perf_counter() - t0) * 1e3 / steps)
    counts = [o[0].count() for o in objs]
    tuning = objs[0][0].tuning()
    assert len(set(counts)) == 1, counts
    for o in objs:
        o[0].close()
    ms = sorted(ts)[2]
    return {"inflight": inflight, "streams": inflight, "ms_per_frame": ms,
            "tuning": {"classify_waves": tuning[0], "zseg": tuning[1], "emit_workgroups": tuning[2]},
            "cells_per_s": (n - 1) ** 3 / (ms * 1e-3),
            "note": "each frame a full build + extract over its own volume buffer; "
                    "steady-state interval, not the latency of one frame"}


def time_build_extract(iso, T, steps, warmup):
    """Vertex count, wall ms per build + extract and per-stage ms (HIP events
    on the library's stream, a separate pass) of a configured IsoSurface,
    output bound to a buffer of the vertex count."""
    import torch
    iso.build(T)
    nv = iso.count()
    out = torch.empty((max(nv, 1), 6), dtype=torch.float32, device=torch.device("cuda", 0))
    iso.bind_output(out)
    for _ in range(warmup):
        iso.build(T)
        iso.extract()
    torch.cuda.synchronize()
    iso.stage_times()
    iso.set_profiling(True)
    for _ in range(steps):
        iso.build(T)
        iso.extract()
    torch.cuda.synchronize()
    iso.set_profiling(False)
    st = iso.stage_times()
    t0 = time.perf_counter()
    for _ in range(steps):
        iso.build(T)
        iso.extract()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    return nv, ms, {k: v[0] / max(v[1], 1) for k, v in st.items()}


def measure_unorm8(n, steps, warmup):
    """The Cayley field quantised to an 8-bit normalised volume (byte
    clamp(round(128 + 64 f)), iso 128/255: the same surface to within the
    quantisation), build + extract on one stream; 1 byte per lattice point."""
    import torch
    import hpmc_amd
    dev = torch.device("cuda", 0)
    f = cayley_volume_torch(n, dev)
    vol = torch.clamp(torch.round(128.0 + 64.0 * f), 0, 255).to(torch.uint8).contiguous()
    del f
    iso = hpmc_amd.IsoSurface((n, n, n))
    iso.set_volume(vol)
    nv, ms, stage_ms = time_build_extract(iso, 128.0 / 255.0, steps, warmup)
    iso.close()
    return {"workload": "cayley-%d^3 quantised to uint8 (128 + 64 f), iso 128/255" % n, "vertices": nv,
            "ms_per_step": ms, "cells_per_s": (n - 1) ** 3 / (ms * 1e-3), "stage_ms": stage_ms,
            "classify_GBps": n ** 3 / (stage_ms["classify"] * 1e-3) / 1e9}


def measure_snippet(n, steps, warmup):
    """The Cayley cubic as a hipRTC field snippet (no volume in memory),
    iso 0: classification evaluates it at every lattice point and emission at
    the samples of the bricks it extracts, so the bound is the VALU, not
    HBM; reported as cells/s and verts/s."""
    import hpmc_amd
    from hpmc_amd import workloads
    iso = hpmc_amd.IsoSurface((n, n, n))
    iso.set_snippet(workloads.CAYLEY_SNIPPET)
    nv, ms, stage_ms = time_build_extract(iso, 0.0, steps, warmup)
    iso.close()
    return {"workload": "cayley-%d^3 hipRTC snippet, iso 0, positions+normals" % n, "vertices": nv,
            "ms_per_step": ms, "cells_per_s": (n - 1) ** 3 / (ms * 1e-3), "verts_per_s": nv / (ms * 1e-3),
            "stage_ms": stage_ms}


def measure_texture256(steps, warmup):
    """256^3 FP32 volume: a seeded uniform [0,1] 32^3 grid upsampled
    trilinearly (corner-aligned), iso 0.5, build + extract on one stream."""
    import torch
    import hpmc_amd
    from hpmc_amd import workloads
    n = 256
    vol = torch.from_numpy(workloads.upsampled_noise(n)).cuda()
    iso = hpmc_amd.IsoSurface((n, n, n))
    iso.set_volume(vol)
    nv, ms, stage_ms = time_build_extract(iso, 0.5, steps, warmup)
    tuning = iso.tuning()
    iso.close()
    algo = 4 * n ** 3 + 24 * nv
    return {"workload": "256^3 fp32 volume, seeded 32^3 uniform noise upsampled trilinearly, iso 0.5",
            "vertices": nv, "ms_per_step": ms, "cells_per_s": (n - 1) ** 3 / (ms * 1e-3),
            "verts_per_s": nv / (ms * 1e-3), "hbm_frac_step": algo / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9),
            "stage_ms": stage_ms,
            "tuning": {"classify_waves": tuning[0], "zseg": tuning[1], "emit_workgroups": tuning[2]}}


def measure_adjoint(reps, count=64, seed=0):
    """Differentiation: `count` cubic FP32 volumes, side ~U{32..128}, f = |p -
    1/2| - 0.3 + N(0, 0.01^2) per lattice point, iso 0; loss = mean over all
    vertices of (|v - centre| - 0.35)^2, the sphere centre being the position
    (0.5, 0.5, 0.5) in HPMC's [0, extent] convention.  One persistent batch:
    forward = build + counts + extract, then dL/dvertices by torch, then the
    adjoint back to every field and to the iso value.  Medians of `reps`
    single steps (each ends with a host wait)."""
    import torch
    import hpmc_amd
    from hpmc_amd import workloads
    dev = torch.device("cuda", 0)
    rng = workloads.SplitMix64(seed)
    sides = [rng.randint(32, 128) for _ in range(count)]
    fields = [torch.from_numpy(workloads.noisy_sphere((n, n, n), seed + 1 + i)).to(dev) for i, n in enumerate(sides)]
    b = hpmc_amd.Batch()
    for f in fields:
        b.add_volume(f)
    b.build(0.0)
    _, _, total = b.counts()
    out = torch.empty((total, 6), dtype=torch.float32, device=dev)
    centre = torch.full((3,), 0.5, dtype=torch.float32, device=dev)

    def forward():
        b.build(0.0)
        b.counts()
        b.extract(out)

    def loss_grad():
        v = out.detach().requires_grad_(True)
        loss = ((v[:, 3:] - centre).norm(dim=1) - 0.35).pow(2).mean()
        (gv,) = torch.autograd.grad(loss, v)
        return gv

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            r = fn()
            torch.cuda.synchronize()
            ts.append((time.perf_counter() - t0) * 1e3)
        return float(sorted(ts)[len(ts) // 2]), r

    fwd_ms, _ = timed(forward)
    gv = loss_grad()
    adj_ms, (grads, gt) = timed(lambda: b.adjoint(gv))
    step_ms, _ = timed(lambda: (forward(), b.adjoint(loss_grad())))
    assert torch.isfinite(gt) and all(torch.isfinite(g).all() for g in grads)
    lattice = sum(n ** 3 for n in sides)
    b.close()
    return {"workload": "%d noisy-sphere FP32 volumes, sides U{32..128}, iso 0, loss mean (|v - c| - 0.35)^2" % count,
            "volumes": count, "lattice_points": lattice, "vertices": int(total),
            "forward_ms": fwd_ms, "adjoint_ms": adj_ms, "step_ms": step_ms}


def traffic():
    """HBM bytes per classify launch from the committed counter profile of this
    workload, or None when the profile is absent."""
    try:
        with open(TRAFFIC_FILE) as fh:
            d = json.load(fh)
        return d["fetch_bytes"] + d["write_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-reps", type=int, default=5)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched/plain comparison")
    ap.add_argument("--batch-reps", type=int, default=21)
    ap.add_argument("--no-pipeline", action="store_true", help="skip the frames-in-flight measurement")
    ap.add_argument("--no-unorm8", action="store_true", help="skip the 8-bit volume measurement")
    ap.add_argument("--no-snippet", action="store_true", help="skip the hipRTC Cayley snippet measurement")
    ap.add_argument("--no-texture256", action="store_true", help="skip the 256^3 noise volume measurement")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiation measurement")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only the differentiation measurement (for kernel traces of it)")
    ap.add_argument("--batch-trace", action="store_true",
                    help="only the batched build + extract at a fixed launch shape (for kernel traces)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="time a plain 512 MiB copy instead (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the engine builds one volume per GPU with no "
              "collectives; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration(args.steps)
        return
    if args.adjoint_only:
        print(json.dumps({"adjoint": measure_adjoint(args.batch_reps)}))
        return
    if args.batch_trace:
        batch_steps(args.steps, args.warmup)
        return

    import torch
    import hpmc_amd
    n = args.size
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    vol = cayley_volume_torch(n, dev)
    iso = hpmc_amd.IsoSurface((n, n, n))
    iso.set_volume(vol)
    iso.build(0.0)
    nv = iso.count()
    if n == 512:
        assert nv == N_CAYLEY_512, nv
    out = torch.empty((nv, 6), dtype=torch.float32, device=dev)
    iso.bind_output(out)

    for _ in range(args.warmup):
        iso.build(0.0)
        iso.extract()
    torch.cuda.synchronize()

    # stage timing pass: every launch records its kernel's start and end in
    # events (the extended launch), which measures the kernels alone but
    # spaces the launches out; the headline is timed in the pass below
    iso.stage_times()
    iso.set_profiling(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        iso.build(0.0)
        iso.extract()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    iso.set_profiling(False)
    st = iso.stage_times()
    # wall time of the same steps without events between the kernels
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    for _ in range(args.steps):
        iso.build(0.0)
        iso.extract()
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    assert iso.count() == nv

    ms_step = (t3 - t2) * 1e3 / args.steps
    cells = (n - 1) ** 3
    lattice_bytes = 4 * n ** 3
    algo_bytes = lattice_bytes + 24 * nv
    stage_ms = {k: v[0] / max(v[1], 1) for k, v in st.items()}
    cls_ms = stage_ms["classify"]
    achieved = lattice_bytes / (cls_ms * 1e-3) / 1e9
    line = {
        "metric": METRIC,
        "value": cells / (ms_step * 1e-3),
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (Cayley cubic materialised as a %d^3 FP32 volume)" % n,
        "config": {"workload": "cayley-%d^3-fp32-volume build+extract, iso 0, positions+normals" % n,
                   "lattice": [n, n, n], "cells": cells, "vertices": nv, "iso": 0.0},
        "verts_per_s": nv / (ms_step * 1e-3),
        "hbm_frac_step": algo_bytes / (ms_step * 1e-3) / (HBM_PEAK_GBS * 1e9),
        "algorithmic_bytes_step": algo_bytes,
        "stage_ms": stage_ms,
        "ms_per_step_with_stage_events": (t1 - t0) * 1e3 / args.steps,
        # FP32 lattices of 256 MiB or more classify with non-temporal loads
        "roofline": {"kernel": "hpmc_classify_volume_stream" if lattice_bytes >= 256 << 20 else "hpmc_classify_volume",
                     "bound": "hbm", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic() if n == 512 else None, "algorithmic_bytes": lattice_bytes},
    }
    iso.close()
    def extra(key, fn):
        # a failing side measurement is reported in its object; the line
        # with the headline is printed whatever happens to the others
        try:
            line[key] = fn()
        except Exception as exc:  # noqa: BLE001 - reported, not hidden
            line[key] = {"error": "%s: %s" % (type(exc).__name__, exc)}
            print("bench: %s failed: %s" % (key, exc), file=sys.stderr, flush=True)

    def pipelined():
        pl = measure_pipelined(n, args.steps, args.warmup)
        pl["hbm_frac"] = algo_bytes / (pl["ms_per_frame"] * 1e-3) / (HBM_PEAK_GBS * 1e9)
        return pl

    if not args.no_pipeline:
        extra("pipelined", pipelined)
    if not args.no_unorm8:
        extra("unorm8", lambda: measure_unorm8(n, args.steps, args.warmup))
    if not args.no_snippet:
        extra("snippet", lambda: measure_snippet(n, args.steps, args.warmup))
    if not args.no_texture256:
        extra("texture256", lambda: measure_texture256(args.steps, args.warmup))
    if not args.no_batch:
        extra("batched", lambda: measure_batched(args.batch_reps))
    if not args.no_adjoint:
        extra("adjoint", lambda: measure_adjoint(args.batch_reps))
    if not args.no_cpu_baseline:
        extra("cpu_baseline", lambda: cpu_baseline(n, args.cpu_reps))
        extra("cpu_baseline_threads", lambda: cpu_baseline_threads(n, args.cpu_reps))
    print(json.dumps(line))


if __name__ == "__main__":
    main()
