"""torch.autograd wrappers: marching cubes as a differentiable function of
the field and the iso-value.

forward:  build + count + extract on the GPU (HPMC order, N3F_V3F rows of
          (normal, position));
backward: the library's adjoint (HPMCadjoint / HPMCbatchAdjoint): per-vertex
          contributions through the edge interpolation, the positions and the
          forward-difference normals, reduced deterministically.

The vertex count and topology are piecewise constant in the field, so the
gradient is that of the vertex values at fixed topology.

The fields are saved for backward, so an in-place change of a field between
forward and backward raises torch's usual version-counter error instead of
differentiating a surface the field no longer has.  The HistoPyramid (or
batch) of the forward lives as long as the autograd graph, so backward may
run more than once (retain_graph=True).
"""
import torch

from . import Batch, IsoSurface


class _MarchingCubes(torch.autograd.Function):
    @staticmethod
    def forward(ctx, field, iso, extent):
        surf = IsoSurface(tuple(field.shape[::-1]), extent=extent)
        surf.set_volume(field.detach())
        verts, _ = surf.extract_all(float(iso))
        ctx.surf = surf
        ctx.save_for_backward(field)
        return verts

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, grad_vertices):
        (field,) = ctx.saved_tensors   # raises if the field changed in place
        surf = ctx.surf
        if surf is None or surf.h is None:
            raise RuntimeError("marching_cubes: the forward's surface is gone")
        gf, gt = surf.adjoint(grad_vertices.contiguous())
        return gf.view_as(field), gt.reshape(()), None


def marching_cubes(field, iso, extent=(1.0, 1.0, 1.0)):
    """Vertices (N, 6) = (normal, position) of the iso-surface of `field`, a
    contiguous float32 CUDA tensor (z, y, x); differentiable in `field` and in
    `iso` (a 0-dim float32 tensor or a float)."""
    if not torch.is_tensor(iso):
        iso = torch.tensor(float(iso), dtype=torch.float32, device=field.device)
    return _MarchingCubes.apply(field.contiguous(), iso, tuple(float(e) for e in extent))


class _MarchingCubesBatch(torch.autograd.Function):
    @staticmethod
    def forward(ctx, iso, *fields):
        b = Batch()
        for f in fields:
            b.add_volume(f.detach().contiguous())
        b.build(float(iso))
        counts, bases, total = b.counts()
        out = torch.empty((max(total, 1), 6), dtype=torch.float32, device=fields[0].device)
        b.extract(out)
        ctx.batch = b
        ctx.save_for_backward(*fields)
        counts_t = torch.from_numpy(counts.astype("int64"))
        ctx.mark_non_differentiable(counts_t)
        return out[:total], counts_t

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, grad_vertices, _grad_counts):
        _ = ctx.saved_tensors   # raises if a field changed in place
        if ctx.batch is None or ctx.batch.b is None:
            raise RuntimeError("marching_cubes_batch: the forward's batch is gone")
        grads, gt = ctx.batch.adjoint(grad_vertices.contiguous())
        return (gt.reshape(()),) + tuple(grads)


def marching_cubes_batch(fields, iso):
    """Iso-surfaces of many volumes in one batched build: a list of (N_i, 6)
    vertex tensors (views of one concatenated output), differentiable in every
    field and in `iso`."""
    if not torch.is_tensor(iso):
        iso = torch.tensor(float(iso), dtype=torch.float32, device=fields[0].device)
    out, counts = _MarchingCubesBatch.apply(iso, *[f.contiguous() for f in fields])
    return list(torch.split(out, counts.tolist()))


__all__ = ["marching_cubes", "marching_cubes_batch"]
