// HistoPyramid marching cubes on CDNA4 (gfx950): device side.
//
// This header is compiled twice: into libhpmc_amd.so for FP32 volume fields,
// and by hipRTC, behind a user field snippet, for analytic fields.  It holds
// no host code and includes nothing outside the HIP device headers.
//
// Pipeline (one HP build + one extraction):
//   classify  : stream the lattice once; per cell the HPMC corner code
//               (bit x+2y+4z set when f < T, hpmc/src/shadergen.cpp:148-153)
//               and vertex count; reduce counts in registers to 16x16-cell
//               blocks of the reference's tiled base image (HP level 4,
//               hpmc/src/init.cpp:105-130) and write those sums.
//   scan      : reduce the level-4 sums in Morton order (the order of the
//               reference's upper pyramid levels and of its traversal,
//               shadergen.cpp:448-467) to exclusive offsets and a compact
//               list of non-empty blocks.
//   emit      : per non-empty block, re-classify its 256 cells from an
//               LDS-staged 18x18x3 brick, descend the in-block levels 0..3
//               per output vertex, and write the edge-interpolated position
//               and forward-difference normal (shadergen.cpp:505-563).
// Vertices come out in the reference's extraction order.
#pragma once

#ifndef HPMC_RTC
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

#include "mc_tables.h"

namespace hpmc {

// 16x16-cell blocks of the doubled base image: HP level 4.
constexpr int kBlockLog2 = 4;
constexpr int kBlock = 1 << kBlockLog2;
constexpr int kBrick = kBlock + 2;          // cells + 1 corner + 1 forward sample
constexpr int kBrickPlane = kBrick * kBrick;
constexpr int kBrickSize = 3 * kBrickPlane; // 972 samples

// Geometry shared by every kernel.  All integer; the host derives it from
// the reference layout (init.cpp:105-130).
struct Geom {
    int sx, sy, sz;       // lattice samples
    int cx, cy, cz;       // cells (grid size)
    int tsx2, tsy2;       // doubled tile sizes: cells per tile along x, y
    int tiles_x;          // slices per tile row
    int l2;               // log2 of the base-level side
    int blocks_log2;      // log2 of the level-4 side (l2 + 1 - kBlockLog2)
    float scx, scy, scz;  // extent / cells
    float ipx, ipy, ipz;  // 1 / lattice size (snippet texel-centre coordinates)
};

struct Totals {
    uint32_t vertices;    // N
    uint32_t blocks;      // non-empty level-4 blocks
    uint32_t overflow;    // vertices dropped by an extraction (capacity)
    uint32_t pad;
};

// Output record of the (cell, edge) list.
struct CellEdge {
    uint32_t cell;        // x + cx*(y + cy*z)
    uint32_t edge;        // 4*axis + s0 + 2*s1
};

__device__ __forceinline__ uint32_t morton_spread(uint32_t v)
{
    v &= 0xffffu;
    v = (v | (v << 8)) & 0x00ff00ffu;
    v = (v | (v << 4)) & 0x0f0f0f0fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

__device__ __forceinline__ uint32_t morton_compact(uint32_t v)
{
    v &= 0x55555555u;
    v = (v | (v >> 1)) & 0x33333333u;
    v = (v | (v >> 2)) & 0x0f0f0f0fu;
    v = (v | (v >> 4)) & 0x00ff00ffu;
    v = (v | (v >> 8)) & 0x0000ffffu;
    return v;
}

// Level-4 block key of the 16x16 cell block starting at (x0, y0) of slice z.
__device__ __forceinline__ uint32_t block_key(const Geom& g, int x0, int y0, int z)
{
    uint32_t tx = (uint32_t)(z % g.tiles_x), ty = (uint32_t)(z / g.tiles_x);
    uint32_t X = (tx * (uint32_t)g.tsx2 + (uint32_t)x0) >> kBlockLog2;
    uint32_t Y = (ty * (uint32_t)g.tsy2 + (uint32_t)y0) >> kBlockLog2;
    return morton_spread(X) | (morton_spread(Y) << 1);
}

// ---------------------------------------------------------------------------
// Field sources.  at(i,j,k) returns the sample at lattice index (i,j,k).

// FP32 volume, x fastest; indices clamp to the lattice like a NEAREST/CLAMP
// 3D texture (apps/texture3d/texture3d.cpp:135-139).
struct VolumeField {
    const float* __restrict__ v;
    int sx, sy, sz;
    __device__ __forceinline__ float at(int i, int j, int k) const
    {
        i = min(max(i, 0), sx - 1);
        j = min(max(j, 0), sy - 1);
        k = min(max(k, 0), sz - 1);
        return v[((size_t)k * sy + j) * sx + i];
    }
    // In-range load without clamping (classification rows): a row pointer
    // that is uniform across the wave plus the lane's x offset.
    __device__ __forceinline__ float row(unsigned i, int j, int k) const
    {
        const char* rp = (const char*)(v + ((size_t)k * sy + j) * sx);
        return *(const float*)(rp + i * 4u);
    }
};

// ---------------------------------------------------------------------------
// classify: one wave per (64 x 16 cells) x zseg-slice column.
//
// Each plane is read once as 17 rows of 64 samples plus the 17 samples of
// the next column (one vector load), turned into 64-bit ballot masks of
// (f < T), and the codes of a whole cell row come out of four masks.  Loads
// use clamped, always-valid addresses so a plane's 18 loads issue back to
// back, and the next plane is loaded while the current one is classified.
// Slices whose two planes are uniformly outside or inside cannot hold a
// surface cell and are skipped with scalar tests.  Counts are summed per
// lane over the 16 rows of a block and across 16-lane groups, giving four
// level-4 sums per slice.
template <class F>
__device__ __forceinline__ void classify_load_plane(const F& field, unsigned xl, int y0, int ymax, unsigned xe,
                                                    int ye, int zz, float (&v)[kBlock + 1], float& ve)
{
#pragma unroll
    for (int r = 0; r <= kBlock; ++r) v[r] = field.row(xl, min(y0 + r, ymax), zz);
    ve = field.row(xe, ye, zz);
}

// Per-wave state carried from plane to plane.
struct ClassifyCarry {
    uint64_t win;   // per lane: 2-bit windows of the previous plane's 17 rows
    uint64_t any;   // previous plane: some sample inside
    uint64_t all;   // previous plane: every sample inside
};

struct ClassifyWave {
    int lane, x0, y0, x, sy, cx, cy;
    bool xin, xcell, ein, lane_e;
    uint64_t lane63;
};

// Classifies plane `zz` against the carried plane zz-1 (when have_lo) and
// writes the four level-4 sums of slice zz-1.
__device__ __forceinline__ void classify_plane(const ClassifyWave& w, const Geom& g, float T,
                                               const float (&v)[kBlock + 1], float ve, bool have_lo, int zz,
                                               ClassifyCarry& c, uint32_t* __restrict__ level4,
                                               const uint8_t* __restrict__ lut)
{
    // bit r of eb: the sample at (x0+64, y0+r) is inside
    const uint64_t eb = __ballot(w.lane_e && ve < T);
    uint64_t win = 0, any = eb, all = ~0ull;
#pragma unroll
    for (int r = 0; r <= kBlock; ++r) {
        const uint64_t m = __ballot(w.xin && (w.y0 + r < w.sy) && v[r] < T);
        any |= m;
        all &= m;
        const uint64_t e = (eb >> r) & 1ull;
        const uint64_t bits = ((m >> w.lane) & 3ull) | ((e & w.lane63) << 1);
        win |= bits << (2 * r);
    }
    // all-inside needs every row's +1 column inside as well
    const uint64_t erows = (1ull << (kBlock + 1)) - 1ull;
    if (!w.ein || (eb & erows) != erows) all = 0;
    if (have_lo) {
        const int z = zz - 1;
        uint32_t acc = 0;
        const bool uniform_out = (c.any | any) == 0ull;
        const bool uniform_in = (c.all & all) == ~0ull;
        if (!(uniform_out || uniform_in)) {
#pragma unroll
            for (int r = 0; r < kBlock; ++r) {
                const uint32_t code = (uint32_t)((c.win >> (2 * r)) & 0xf) | ((uint32_t)((win >> (2 * r)) & 0xf) << 4);
                const bool valid = w.xcell && (w.y0 + r < w.cy);
                acc += valid ? lut[code] : 0u;
            }
        }
        // 16-lane group sums = level-4 sums of four blocks.
        acc += __shfl_xor(acc, 8, 16);
        acc += __shfl_xor(acc, 4, 16);
        acc += __shfl_xor(acc, 2, 16);
        acc += __shfl_xor(acc, 1, 16);
        if ((w.lane & 15) == 0 && w.x < w.cx) level4[block_key(g, w.x, w.y0, z)] = acc;
    }
    c.win = win;
    c.any = any;
    c.all = all;
}

template <class F>
__device__ __forceinline__ void classify_body(const F& field, const Geom& g, float T,
                                              uint32_t* __restrict__ level4, int zseg,
                                              const uint8_t* __restrict__ lut)
{
    const int lane = threadIdx.x & 63;
    // wave-uniform: keeps the column indices, plane loop and row pointers in
    // scalar registers
    const int wave = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    const int npx = (g.cx + 63) >> 6;
    const int npy = (g.cy + kBlock - 1) >> kBlockLog2;
    const int npz = (g.cz + zseg - 1) / zseg;
    if (wave >= npx * npy * npz) return;
    const int px = wave % npx;
    const int rest = wave / npx;
    const int py = rest % npy;
    const int pz = rest / npy;
    ClassifyWave w;
    w.lane = lane;
    w.x0 = px << 6;
    w.y0 = py << kBlockLog2;
    w.x = w.x0 + lane;
    w.sy = g.sy;
    w.cx = g.cx;
    w.cy = g.cy;
    w.xin = w.x < g.sx;
    w.xcell = w.x < g.cx;
    w.ein = w.x0 + 64 < g.sx;
    w.lane_e = w.ein && lane <= kBlock && (w.y0 + lane < g.sy);
    w.lane63 = lane == 63 ? 1ull : 0ull;
    const int z0 = pz * zseg;
    const int z1 = min(z0 + zseg, g.cz);   // cell slices [z0, z1): planes z0..z1
    const unsigned xl = (unsigned)min(w.x, g.sx - 1);
    const unsigned xe = (unsigned)min(w.x0 + 64, g.sx - 1);
    const int ye = min(w.y0 + min(lane, kBlock), g.sy - 1);
    const int ymax = g.sy - 1;

    // Two plane buffers, each loaded one plane ahead of its use.  The
    // prefetch is unconditional (the last one re-reads plane z1) so that the
    // wait for a buffer never covers the other buffer's loads.
    float a[kBlock + 1], b[kBlock + 1], ae, be;
    ClassifyCarry c{0, 0, 0};
    classify_load_plane(field, xl, w.y0, ymax, xe, ye, z0, a, ae);
    int zz = z0;
    for (;;) {
        classify_load_plane(field, xl, w.y0, ymax, xe, ye, min(zz + 1, z1), b, be);
        __builtin_amdgcn_sched_barrier(0);
        classify_plane(w, g, T, a, ae, zz > z0, zz, c, level4, lut);
        if (++zz > z1) break;
        classify_load_plane(field, xl, w.y0, ymax, xe, ye, min(zz + 1, z1), a, ae);
        __builtin_amdgcn_sched_barrier(0);
        classify_plane(w, g, T, b, be, true, zz, c, level4, lut);
        if (++zz > z1) break;
    }
}

// ---------------------------------------------------------------------------
// scan, part 1: per 4096-block tile of the level-4 array, the sum of counts
// and of non-empty blocks.
constexpr int kScanTile = 4096;

__device__ __forceinline__ uint2 block_reduce_u2(uint2 v, uint2* sh)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        v.x += __shfl_xor(v.x, o);
        v.y += __shfl_xor(v.y, o);
    }
    if (lane == 0) sh[w] = v;
    __syncthreads();
    uint2 r = make_uint2(0, 0);
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) {
        r.x += sh[i].x;
        r.y += sh[i].y;
    }
    __syncthreads();
    return r;
}

__device__ __forceinline__ void scan_reduce_body(const uint32_t* __restrict__ level4, uint32_t n,
                                                 uint2* __restrict__ tile_sums)
{
    __shared__ uint2 sh[16];
    const uint32_t base = blockIdx.x * kScanTile;
    uint2 s = make_uint2(0, 0);
    for (uint32_t i = base + threadIdx.x; i < min(base + kScanTile, n); i += blockDim.x) {
        uint32_t c = level4[i];
        s.x += c;
        s.y += c != 0;
    }
    s = block_reduce_u2(s, sh);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = s;
}

// scan, part 2: every tile re-reads the preceding tile sums (a few hundred),
// scans its own 4096 entries and scatters the non-empty blocks with their
// vertex offsets in Morton order.  Tile 0 also writes the totals.
__device__ __forceinline__ void scan_scatter_body(const uint32_t* __restrict__ level4, uint32_t n,
                                                  const uint2* __restrict__ tile_sums, uint32_t ntiles,
                                                  uint32_t* __restrict__ act_key,
                                                  uint32_t* __restrict__ act_off,
                                                  Totals* __restrict__ totals)
{
    __shared__ uint2 sh[16];
    __shared__ uint2 wsum[16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
    uint2 pre = make_uint2(0, 0);
    for (uint32_t i = threadIdx.x; i < blockIdx.x; i += blockDim.x) {
        pre.x += tile_sums[i].x;
        pre.y += tile_sums[i].y;
    }
    pre = block_reduce_u2(pre, sh);
    if (blockIdx.x == 0) {
        uint2 tot = make_uint2(0, 0);
        for (uint32_t i = threadIdx.x; i < ntiles; i += blockDim.x) {
            tot.x += tile_sums[i].x;
            tot.y += tile_sums[i].y;
        }
        tot = block_reduce_u2(tot, sh);
        if (threadIdx.x == 0) {
            totals->vertices = tot.x;
            totals->blocks = tot.y;
            totals->overflow = 0;
        }
    }
    // Each thread owns 16 consecutive entries of the tile.
    constexpr int kPer = kScanTile / 256;
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kPer;
    uint32_t c[kPer];
    uint2 mine = make_uint2(0, 0);
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        c[k] = (base + k < n) ? level4[base + k] : 0u;
        mine.x += c[k];
        mine.y += c[k] != 0;
    }
    // exclusive wave scan
    uint2 inc = mine;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t ax = __shfl_up(inc.x, o), ay = __shfl_up(inc.y, o);
        if (lane >= o) {
            inc.x += ax;
            inc.y += ay;
        }
    }
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    uint2 off = pre;
    for (int i = 0; i < w; ++i) {
        off.x += wsum[i].x;
        off.y += wsum[i].y;
    }
    (void)nw;
    off.x += inc.x - mine.x;
    off.y += inc.y - mine.y;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        if (c[k]) {
            act_key[off.y] = base + k;
            act_off[off.y] = off.x;
            off.x += c[k];
            off.y += 1;
        }
    }
}

// ---------------------------------------------------------------------------
// emit: one workgroup of 256 threads per non-empty level-4 block, thread i
// owning the cell at Morton index i of the block (levels 0..3 of the
// pyramid below the block).
struct EmitShared {
    uint64_t cases[256];
    float brick[kBrickSize];
    uint16_t incl[256];      // inclusive vertex prefix over the block's cells
    uint8_t code[256];
    uint8_t lut[256];
    uint32_t wsum[4];
};

template <class F>
__device__ __forceinline__ void emit_body(const F& field, const Geom& g, float T,
                                          const uint32_t* __restrict__ act_key,
                                          const uint32_t* __restrict__ act_off,
                                          const uint32_t* __restrict__ level4,
                                          const Totals* __restrict__ totals,
                                          float* __restrict__ out, CellEdge* __restrict__ out_ce,
                                          uint32_t capacity, EmitShared& sh)
{
    const uint64_t* cases = sh.cases;
    const uint8_t* lut = sh.lut;
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const uint32_t nblocks = totals->blocks;
    const uint32_t lx = morton_compact((uint32_t)tid), ly = morton_compact((uint32_t)tid >> 1);

    for (uint32_t bi = blockIdx.x; bi < nblocks; bi += gridDim.x) {
        const uint32_t key = act_key[bi];
        const uint32_t base = act_off[bi];
        const uint32_t stored = level4[key];
        const int X0 = (int)morton_compact(key) << kBlockLog2;
        const int Y0 = (int)morton_compact(key >> 1) << kBlockLog2;
        const int tx = X0 / g.tsx2, ty = Y0 / g.tsy2;
        const int z = ty * g.tiles_x + tx;
        const int x0 = X0 - tx * g.tsx2, y0 = Y0 - ty * g.tsy2;

        for (int i = tid; i < kBrickSize; i += 256) {
            const int kz = i / kBrickPlane, r = i - kz * kBrickPlane;
            const int ky = r / kBrick, kx = r - ky * kBrick;
            sh.brick[i] = field.at(x0 + kx, y0 + ky, z + kz);
        }
        __syncthreads();

        // cell code and count
        const int cxl = (int)lx, cyl = (int)ly;
        uint32_t code = 0;
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const float v = sh.brick[((b >> 2) * kBrick + cyl + ((b >> 1) & 1)) * kBrick + cxl + (b & 1)];
            code |= (v < T ? 1u : 0u) << b;
        }
        const bool valid = (x0 + cxl < g.cx) && (y0 + cyl < g.cy);
        const uint32_t cnt = valid ? lut[code] : 0u;
        // inclusive scan over the 256 cells (Morton order)
        uint32_t inc = cnt;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            uint32_t a = __shfl_up(inc, o);
            if (lane >= o) inc += a;
        }
        if (lane == 63) sh.wsum[w] = inc;
        __syncthreads();
        for (int i = 0; i < w; ++i) inc += sh.wsum[i];
        sh.incl[tid] = (uint16_t)inc;
        sh.code[tid] = (uint8_t)code;
        __syncthreads();
        uint32_t nv = sh.incl[255];
        if (nv > stored) nv = stored;

        for (uint32_t v = tid; v < nv; v += 256) {
            // descend: first cell whose inclusive prefix exceeds v
            int lo = 0;
#pragma unroll
            for (int step = 128; step > 0; step >>= 1) {
                if (sh.incl[lo + step - 1] <= v) lo += step;
            }
            const int c = lo;
            const uint32_t before = c ? sh.incl[c - 1] : 0u;
            const uint32_t slot = v - before;
            const uint32_t ccode = sh.code[c];
            const uint32_t e = (uint32_t)(cases[ccode] >> (4 + 4 * slot)) & 0xfu;
            const int axis = (int)(e >> 2);
            const int s0 = (int)(e & 1), s1 = (int)((e >> 1) & 1);
            const int ccx = (int)morton_compact((uint32_t)c), ccy = (int)morton_compact((uint32_t)c >> 1);
            int ax = ccx, ay = ccy, az = 0;
            if (axis == 0) { ay += s0; az += s1; }
            else if (axis == 1) { ax += s0; az += s1; }
            else { ax += s0; ay += s1; }
            const int bx = ax + (axis == 0), by = ay + (axis == 1), bz = az + (axis == 2);
            const float* B = sh.brick;
            const float va = B[(az * kBrick + ay) * kBrick + ax];
            const float vb = B[(bz * kBrick + by) * kBrick + bx];
            const float nax = B[(az * kBrick + ay) * kBrick + ax + 1];
            const float nay = B[(az * kBrick + ay + 1) * kBrick + ax];
            const float naz = B[((az + 1) * kBrick + ay) * kBrick + ax];
            const float nbx = B[(bz * kBrick + by) * kBrick + bx + 1];
            const float nby = B[(bz * kBrick + by + 1) * kBrick + bx];
            const float nbz = B[((bz + 1) * kBrick + by) * kBrick + bx];
            const float t = (va - T) / (va - vb);
            const float u = 1.0f - t;
            const float gx = (float)(x0 + ax), gy = (float)(y0 + ay), gz = (float)(z + az);
            const uint32_t o = base + v;
            if (o < capacity) {
                float* dst = out + (size_t)o * 6;
                dst[0] = (T - (nax * u + nbx * t)) * g.scx;
                dst[1] = (T - (nay * u + nby * t)) * g.scy;
                dst[2] = (T - (naz * u + nbz * t)) * g.scz;
                dst[3] = (gx + (axis == 0 ? t : 0.0f)) * g.scx;
                dst[4] = (gy + (axis == 1 ? t : 0.0f)) * g.scy;
                dst[5] = (gz + (axis == 2 ? t : 0.0f)) * g.scz;
                if (out_ce) {
                    CellEdge ce;
                    ce.cell = (uint32_t)(x0 + ccx) + (uint32_t)g.cx * ((uint32_t)(y0 + ccy) + (uint32_t)g.cy * (uint32_t)z);
                    ce.edge = e;
                    out_ce[o] = ce;
                }
            }
        }
        __syncthreads();
    }
}

// Per-cell codes of the whole grid (parity and debugging).
template <class F>
__device__ __forceinline__ void codes_body(const F& field, const Geom& g, float T, uint8_t* __restrict__ codes)
{
    const size_t n = (size_t)g.cx * g.cy * g.cz;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(c % g.cx);
        const int y = (int)((c / g.cx) % g.cy);
        const int z = (int)(c / ((size_t)g.cx * g.cy));
        uint32_t code = 0;
#pragma unroll
        for (int b = 0; b < 8; ++b)
            code |= (field.at(x + (b & 1), y + ((b >> 1) & 1), z + (b >> 2)) < T ? 1u : 0u) << b;
        codes[c] = (uint8_t)code;
    }
}

}  // namespace hpmc

// Kernel entry points for one field source.  MAKE_FIELD builds the field
// object from the Geom `g` and the opaque source pointer `fsrc`.
#define HPMC_DEFINE_FIELD_KERNELS(SUFFIX, MAKE_FIELD)                                          \
    extern "C" __global__ __launch_bounds__(256) void hpmc_classify_##SUFFIX(                  \
        hpmc::Geom g, float T, const void* fsrc, uint32_t* level4, int zseg,                  \
        const uint64_t* cases)                                                                \
    {                                                                                         \
        __shared__ uint8_t lut[256];                                                          \
        lut[threadIdx.x] = (uint8_t)(cases[threadIdx.x] & 0xfu);                              \
        __syncthreads();                                                                      \
        hpmc::classify_body(MAKE_FIELD, g, T, level4, zseg, lut);                             \
    }                                                                                         \
    extern "C" __global__ __launch_bounds__(256) void hpmc_emit_##SUFFIX(                      \
        hpmc::Geom g, float T, const void* fsrc, const uint32_t* act_key,                     \
        const uint32_t* act_off, const uint32_t* level4, const hpmc::Totals* totals,          \
        float* out, hpmc::CellEdge* out_ce, uint32_t capacity, const uint64_t* cases)         \
    {                                                                                         \
        __shared__ hpmc::EmitShared sh;                                                       \
        sh.cases[threadIdx.x] = cases[threadIdx.x];                                           \
        sh.lut[threadIdx.x] = (uint8_t)(cases[threadIdx.x] & 0xfu);                           \
        __syncthreads();                                                                      \
        hpmc::emit_body(MAKE_FIELD, g, T, act_key, act_off, level4, totals, out, out_ce,      \
                        capacity, sh);                                                        \
    }                                                                                         \
    extern "C" __global__ __launch_bounds__(256) void hpmc_codes_##SUFFIX(                     \
        hpmc::Geom g, float T, const void* fsrc, uint8_t* codes)                              \
    {                                                                                         \
        hpmc::codes_body(MAKE_FIELD, g, T, codes);                                            \
    }
