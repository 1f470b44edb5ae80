"""Synthetic fields of the benchmark and parity workloads.

All randomness comes from SplitMix64 (Steele, Lea, Flood 2014) so every
consumer draws identical bits.  Fields are evaluated in FP32 in a fixed
operation order, with numpy (bitwise IEEE, no fused multiply-add), and the
matching HIP snippets below evaluate the same order under -ffp-contract=off,
so GPU-side snippet values equal these arrays bit for bit.

Texel-centre coordinates follow the snippet contract of HPMC
(hpmc/include/hpmc.h:409-437): lattice index k maps to p = (k + 0.5) / size,
computed as (float(k) + 0.5f) * (1.0f / size).
"""
import numpy as np

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.state = seed & MASK64

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self):
        """Uniform FP64 in [0, 1) from the top 53 bits."""
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def uniform_array(self, n):
        return np.array([self.uniform() for _ in range(n)], np.float64)

    def randint(self, lo, hi):
        """Uniform integer in [lo, hi] (inclusive)."""
        return lo + int(self.uniform() * (hi - lo + 1))

    def normal(self):
        """Standard normal by Box-Muller on two uniforms."""
        u1 = max(self.uniform(), 1e-300)
        u2 = self.uniform()
        return float(np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2))


def texel_centres(n):
    """FP32 texel-centre coordinates (k + 0.5) * (1/n) for k in [0, n)."""
    inv = np.float32(1.0) / np.float32(n)
    return (np.arange(n, dtype=np.float32) + np.float32(0.5)) * inv


def _grid(size, extra=0):
    """Broadcastable FP32 coordinate arrays (z, y, x order) for a lattice of
    `size` = (X, Y, Z); `extra` extends each axis past the lattice (samples a
    snippet field takes outside the lattice)."""
    X, Y, Z = size
    inv = [np.float32(1.0) / np.float32(n) for n in size]
    px = (np.arange(X + extra, dtype=np.float32) + np.float32(0.5)) * inv[0]
    py = (np.arange(Y + extra, dtype=np.float32) + np.float32(0.5)) * inv[1]
    pz = (np.arange(Z + extra, dtype=np.float32) + np.float32(0.5)) * inv[2]
    return px[None, None, :], py[None, :, None], pz[:, None, None]


# ---------------------------------------------------------------------------
# Cayley cubic: f = 4(x^2+y^2+z^2) + 16xyz - 1 with x = 2p - 1, iso 0.
CAYLEY_SNIPPET = r"""
__device__ float HPMC_fetch(float3 p)
{
    float x = 2.0f * p.x - 1.0f;
    float y = 2.0f * p.y - 1.0f;
    float z = 2.0f * p.z - 1.0f;
    float s = x * x + y * y + z * z;
    return 4.0f * s + 16.0f * x * y * z - 1.0f;
}
"""


# The same field with its analytic gradient in p (hpmc.h:409-444 contract:
# xyz = gradient, w = value): d/dp of f(2p - 1) = 2 f'(x).
CAYLEY_GRAD_SNIPPET = CAYLEY_SNIPPET + r"""
__device__ float4 HPMC_fetchGrad(float3 p)
{
    float x = 2.0f * p.x - 1.0f;
    float y = 2.0f * p.y - 1.0f;
    float z = 2.0f * p.z - 1.0f;
    float gx = 16.0f * x + 32.0f * y * z;
    float gy = 16.0f * y + 32.0f * x * z;
    float gz = 16.0f * z + 32.0f * x * y;
    return make_float4(gx, gy, gz, HPMC_fetch(p));
}
"""


def cayley_grad(size, extra=0):
    """(Z, Y, X, 4) float32: CAYLEY_GRAD_SNIPPET's HPMC_fetchGrad at every
    lattice point, in the snippet's FP32 operation order."""
    px, py, pz = _grid(size, extra)
    f32 = np.float32
    x = f32(2.0) * px - f32(1.0)
    y = f32(2.0) * py - f32(1.0)
    z = f32(2.0) * pz - f32(1.0)
    out = np.zeros(np.broadcast_shapes(x.shape, y.shape, z.shape) + (4,), np.float32)
    out[..., 0] = f32(16.0) * x + (f32(32.0) * y) * z
    out[..., 1] = f32(16.0) * y + (f32(32.0) * x) * z
    out[..., 2] = f32(16.0) * z + (f32(32.0) * x) * y
    out[..., 3] = cayley(size, extra)
    return out


def cayley(size, extra=0):
    px, py, pz = _grid(size, extra)
    f32 = np.float32
    x = f32(2.0) * px - f32(1.0)
    y = f32(2.0) * py - f32(1.0)
    z = f32(2.0) * pz - f32(1.0)
    s = (x * x + y * y) + z * z            # broadcasts to (Z, Y, X)
    xyz = ((f32(16.0) * x) * y) * z
    return (f32(4.0) * s + xyz - f32(1.0)).astype(np.float32)


# ---------------------------------------------------------------------------
# The particles demo's algebraic shapes (apps/particles/particles.cpp:100-158):
# a degree-5 polynomial in q = 2.2 (p - 1/2) with twelve coefficients in the
# uniform `shape` (set with HPMCsetFieldUniform, as the demo sets it with
# glUniform1fv at particles.cpp:728), negated.  HPMC_fetchGrad keeps the
# demo's expressions as written, including its z-derivative of the y^2 z^2
# term, 2 shape[6] y z^2 (particles.cpp:142), and gives the gradient in q.
PARTICLES_SNIPPET = r"""
__device__ float shape[12];
__device__ float HPMC_fetch(float3 p)
{
    const float x = (p.x - 0.5f) * 2.2f;
    const float y = (p.y - 0.5f) * 2.2f;
    const float z = (p.z - 0.5f) * 2.2f;
    return -(shape[0] * x * x * x * x * x + shape[1] * x * x * x * x + shape[2] * y * y * y * y +
             shape[3] * z * z * z * z + shape[4] * x * x * y * y + shape[5] * x * x * z * z +
             shape[6] * y * y * z * z + shape[7] * x * y * z + shape[8] * x * x + shape[9] * y * y +
             shape[10] * z * z + shape[11]);
}
"""

PARTICLES_GRAD_SNIPPET = PARTICLES_SNIPPET + r"""
__device__ float4 HPMC_fetchGrad(float3 p)
{
    const float x = (p.x - 0.5f) * 2.2f;
    const float y = (p.y - 0.5f) * 2.2f;
    const float z = (p.z - 0.5f) * 2.2f;
    const float gx = 5.0f * shape[0] * x * x * x * x + 4.0f * shape[1] * x * x * x + 2.0f * shape[4] * x * y * y +
                     2.0f * shape[5] * x * z * z + shape[7] * y * z + 2.0f * shape[8] * x;
    const float gy = 4.0f * shape[2] * y * y * y + 2.0f * shape[4] * x * x * y + 2.0f * shape[6] * y * z * z +
                     shape[7] * x * z + 2.0f * shape[9] * y;
    const float gz = 4.0f * shape[3] * z * z * z + 2.0f * shape[5] * x * x * z + 2.0f * shape[6] * y * z * z +
                     shape[7] * x * y + 2.0f * shape[10] * z;
    return make_float4(-gx, -gy, -gz, HPMC_fetch(p));
}
"""

# Coefficients of the demo's shapes (particles.cpp:697-712: x^5, x^4, y^4,
# z^4, x^2y^2, x^2z^2, y^2z^2, xyz, x^2, y^2, z^2, 1) and its iso-value.
PARTICLES_SHAPES = {
    "helix": (0.0, -2.0, 0.0, 0.0, 0.0, 0.0, -1.0, 0.0, 6.0, 0.0, 0.0, 0.0),
    "between1": (0.0, 8.0, 0.5, 0.5, 4.0, 4.0, -1.4, 0.0, 0.0, 0.0, 0.0, 0.0),
    "between2": (0.0, 16.0, 1.0, 1.0, 8.0, 8.0, -2.0, 0.0, -6.0, 0.0, 0.0, 0.0),
    "daddel": (0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 1.0, 0.3, -0.95),
    "torus": (0.0, 1.0, 1.0, 1.0, 2.0, 2.0, 2.0, 0.0, -1.01125, -1.01125, 0.94875, 0.225032),
    "kiss": (-0.5, -0.5, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 1.0, 0.0),
    "cayley": (0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 16.0, 4.0, 4.0, 4.0, -1.0),
}
PARTICLES_ISO = 0.001


def particles_shape(a, b, u):
    """The demo's morph between two shapes, (1 - u) C[a] + u C[b] in FP32
    (particles.cpp:714-720)."""
    f32 = np.float32
    ca = np.asarray(PARTICLES_SHAPES[a], np.float32)
    cb = np.asarray(PARTICLES_SHAPES[b], np.float32)
    return ((f32(1.0) - f32(u)) * ca + f32(u) * cb).astype(np.float32)


def _particles_q(size, extra):
    px, py, pz = _grid(size, extra)
    f32 = np.float32
    return (px - f32(0.5)) * f32(2.2), (py - f32(0.5)) * f32(2.2), (pz - f32(0.5)) * f32(2.2)


def particles(size, shape, extra=0):
    """PARTICLES_SNIPPET's HPMC_fetch at every lattice point (z, y, x), in
    the snippet's FP32 operation order (products and sums left to right)."""
    s = [np.float32(v) for v in np.asarray(shape, np.float32)]
    x, y, z = _particles_q(size, extra)
    t = s[0] * x * x * x * x * x + s[1] * x * x * x * x
    t = t + s[2] * y * y * y * y
    t = t + s[3] * z * z * z * z
    t = t + s[4] * x * x * y * y
    t = t + s[5] * x * x * z * z
    t = t + s[6] * y * y * z * z
    t = t + s[7] * x * y * z
    t = t + s[8] * x * x
    t = t + s[9] * y * y
    t = t + s[10] * z * z
    t = t + s[11]
    return (-t).astype(np.float32)


def particles_grad(size, shape, extra=0):
    """(Z, Y, X, 4) float32: PARTICLES_GRAD_SNIPPET's HPMC_fetchGrad in its
    FP32 operation order."""
    s = [np.float32(v) for v in np.asarray(shape, np.float32)]
    f32 = np.float32
    x, y, z = _particles_q(size, extra)
    gx = f32(5.0) * s[0] * x * x * x * x + f32(4.0) * s[1] * x * x * x
    gx = gx + f32(2.0) * s[4] * x * y * y
    gx = gx + f32(2.0) * s[5] * x * z * z
    gx = gx + s[7] * y * z
    gx = gx + f32(2.0) * s[8] * x
    gy = f32(4.0) * s[2] * y * y * y + f32(2.0) * s[4] * x * x * y
    gy = gy + f32(2.0) * s[6] * y * z * z
    gy = gy + s[7] * x * z
    gy = gy + f32(2.0) * s[9] * y
    gz = f32(4.0) * s[3] * z * z * z + f32(2.0) * s[5] * x * x * z
    gz = gz + f32(2.0) * s[6] * y * z * z
    gz = gz + s[7] * x * y
    gz = gz + f32(2.0) * s[10] * z
    f = particles(size, shape, extra)
    out = np.zeros(f.shape + (4,), np.float32)
    out[..., 0] = -np.broadcast_to(gx, f.shape)
    out[..., 1] = -np.broadcast_to(gy, f.shape)
    out[..., 2] = -np.broadcast_to(gz, f.shape)
    out[..., 3] = f
    return out


# ---------------------------------------------------------------------------
# Metaballs: f(p) = sum_i r_i^2 / |p - c_i|^2, iso 1.
METABALL_SNIPPET = r"""
__device__ int hpmc_nballs;
__device__ float4 hpmc_balls[64];   // xyz = centre, w = r^2
__device__ float HPMC_fetch(float3 p)
{
    float s = 0.0f;
    for (int i = 0; i < hpmc_nballs; ++i) {
        float4 b = hpmc_balls[i];
        float dx = p.x - b.x;
        float dy = p.y - b.y;
        float dz = p.z - b.z;
        float d2 = dx * dx + dy * dy + dz * dz;
        s = s + b.w / d2;
    }
    return s;
}
"""


def metaball_params(rng, nballs):
    """(nballs, 4) FP32: centre ~ U[0.2,0.8]^3, w = r*r with r ~ U[0.1,0.2]."""
    balls = np.zeros((nballs, 4), np.float32)
    for i in range(nballs):
        c = [0.2 + 0.6 * rng.uniform() for _ in range(3)]
        r = np.float32(0.1 + 0.1 * rng.uniform())
        balls[i, :3] = np.array(c, np.float32)
        balls[i, 3] = r * r
    return balls


def metaballs(size, balls, extra=0):
    px, py, pz = _grid(size, extra)
    shape = (size[2] + extra, size[1] + extra, size[0] + extra)
    s = np.zeros(shape, np.float32)
    for b in balls:
        dx = px - b[0]
        dy = py - b[1]
        dz = pz - b[2]
        d2 = (dx * dx + dy * dy) + dz * dz
        s = s + b[3] / d2
    return s.astype(np.float32)


def metaballs_torch(size, balls, device, extra=0):
    """metaballs() evaluated with torch on `device`: the same FP32
    elementwise operations in the same order (bit-identical)."""
    import torch
    X, Y, Z = (v + extra for v in size)
    f32 = torch.float32

    def centres(n, count):
        inv = torch.tensor(1.0, dtype=f32) / torch.tensor(float(n), dtype=f32)
        return (torch.arange(count, dtype=f32, device=device) + 0.5) * inv.to(device)

    px = centres(size[0], X).view(1, 1, X)
    py = centres(size[1], Y).view(1, Y, 1)
    pz = centres(size[2], Z).view(Z, 1, 1)
    s = torch.zeros((Z, Y, X), dtype=f32, device=device)
    for b in np.asarray(balls, np.float32):
        dx = px - float(b[0])
        dy = py - float(b[1])
        dz = pz - float(b[2])
        d2 = (dx * dx + dy * dy) + dz * dz
        # tensor / tensor (a python scalar over a tensor becomes a reciprocal
        # times the scalar, rounded twice)
        s = s + torch.tensor(float(b[3]), dtype=f32, device=device) / d2
    return s.contiguous()


# ---------------------------------------------------------------------------
def upsampled_noise(size=256, coarse=32, seed=1):
    """Coarse ~U[0,1]^3 grid trilinearly upsampled, corner-aligned
    (fine i -> coarse i*(coarse-1)/(size-1)), in FP64, rounded to FP32 once."""
    rng = SplitMix64(seed)
    c = rng.uniform_array(coarse ** 3).reshape(coarse, coarse, coarse)  # (z, y, x)
    u = np.arange(size, dtype=np.float64) * (coarse - 1) / (size - 1)
    i0 = np.minimum(np.floor(u).astype(np.int64), coarse - 2)
    w = u - i0

    def lerp_axis(a, axis):
        a0 = np.take(a, i0, axis=axis)
        a1 = np.take(a, i0 + 1, axis=axis)
        shp = [1, 1, 1]
        shp[axis] = size
        ww = w.reshape(shp)
        return a0 * (1.0 - ww) + a1 * ww

    f = lerp_axis(lerp_axis(lerp_axis(c, 2), 1), 0)
    return f.astype(np.float32)


def noisy_sphere(size, seed, sigma=0.01):
    """|p - 0.5| - 0.3 + N(0, sigma^2) per lattice point (FP64, rounded once)."""
    X, Y, Z = size
    rng = SplitMix64(seed)
    px, py, pz = [(np.arange(n) + 0.5) / n for n in (X, Y, Z)]
    d = np.sqrt((px[None, None, :] - 0.5) ** 2 + (py[None, :, None] - 0.5) ** 2 + (pz[:, None, None] - 0.5) ** 2)
    noise = np.random.Generator(np.random.PCG64(rng.next_u64())).normal(0.0, sigma, size=(Z, Y, X))
    return (d - 0.3 + noise).astype(np.float32)


def batch_volume_specs(count=256, lo=16, hi=128, balls_lo=4, balls_hi=32, seed=0):
    """Sizes and metaball parameters of the batched workload."""
    rng = SplitMix64(seed)
    specs = []
    for _ in range(count):
        size = tuple(rng.randint(lo, hi) for _ in range(3))
        nb = rng.randint(balls_lo, balls_hi)
        specs.append((size, metaball_params(rng, nb)))
    return specs
