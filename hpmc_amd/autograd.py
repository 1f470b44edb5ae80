"""torch.autograd wrappers: marching cubes as a differentiable function of
the field and the iso-value.

forward:  build + count + extract on the GPU (HPMC order, N3F_V3F rows of
          (normal, position));
backward: the library's adjoint (HPMCadjoint / HPMCbatchAdjoint): per-vertex
          contributions through the edge interpolation, the positions and the
          forward-difference normals, reduced deterministically.

The vertex count and topology are piecewise constant in the field, so the
gradient is that of the vertex values at fixed topology.
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
        return verts

    @staticmethod
    def backward(ctx, grad_vertices):
        surf = ctx.surf
        gf, gt = surf.adjoint(grad_vertices)
        surf.close()
        ctx.surf = None
        return gf, gt.reshape(()), None


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
        counts_t = torch.from_numpy(counts.astype("int64"))
        ctx.mark_non_differentiable(counts_t)
        return out[:total], counts_t

    @staticmethod
    def backward(ctx, grad_vertices, _grad_counts):
        grads, gt = ctx.batch.adjoint(grad_vertices)
        ctx.batch.close()
        ctx.batch = None
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
