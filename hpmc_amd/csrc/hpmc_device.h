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
//               and normal (shadergen.cpp:505-563): forward differences,
//               the field's own gradient, or a binary field's midpoint and
//               tabulated face normal.
// Vertices come out in the reference's extraction order.
#pragma once

#ifndef HPMC_RTC
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

#ifndef HPMC_TABLES_INLINE
#include "mc_tables.h"
#endif

// No implicit fused multiply-add in this code: the kernels are compiled
// ahead of time and by hipRTC (library snippets, user traversal kernels,
// where contraction is off), and both must give the same bits; the fused
// operations the emission wants are written out as fma.
#pragma clang fp contract(off)

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
    int blocks_log2;      // log2 of the block level's side (l2 + 1 - blog2)
    float scx, scy, scz;  // extent / cells
    float ipx, ipy, ipz;  // 1 / lattice size (snippet texel-centre coordinates)
    int zseg;             // classification slices per workgroup: 8, 16 or 32
    int nbx, nby, npz;    // level-4 block columns along x, y; z segments
    // tsx2, tsy2, tiles_x and zseg are powers of two: their log2, so that
    // key and slice arithmetic is shifts and masks
    int tsx2_log2, tsy2_log2, tiles_x_log2, zseg_log2;
    // log2 of the block side: kBlockLog2 (16x16-cell blocks, HP level 4),
    // or less when a doubled tile is narrower than 16 cells (grids of at most
    // 8 cells along x or y): blocks never straddle two tiles, so each block
    // is cells of one slice and its vertices are contiguous in the order
    int blog2;
};

// Per (z segment, y, x) block column the vertex counts of its slices, 32
// u16 entries (64 bytes) per column whatever the segment length, so each
// column is written as one full line.
constexpr int kColStride = 32;

// Scan tiles: 4096 entries, 16 per thread of a 256-thread workgroup.
constexpr int kScanTile = 4096;
constexpr int kScanThreads = 256;                // threads per scan workgroup
constexpr int kScanPer = kScanTile / kScanThreads;   // entries per thread

// Emit work items hold at most kEmitChunk vertices of one block, so a busy
// block is spread over several waves.
constexpr int kEmitChunk = 256;
static_assert(kEmitChunk == 256, "item_chunk_first assumes 256-vertex chunks");

__device__ __forceinline__ uint32_t chunks_of(uint32_t cnt) { return (cnt + kEmitChunk - 1) / kEmitChunk; }


struct Totals {
    uint32_t vertices;    // N
    uint32_t blocks;      // non-empty level-4 blocks
    uint32_t items;       // emit work items
    uint32_t pad;
};

// The emit item total again, in kItemReps copies kItemRepStride words apart:
// every emission workgroup reads it first thing, and the 5120 waves of a
// 512^3 extraction reading one word queue on one L2 channel (last arrival
// ~6 us after the launch; ~5 us with the copies, median 1.5 us either way).
// Workgroup b reads copy b mod kItemReps.
constexpr uint32_t kItemReps = 64;
constexpr uint32_t kItemRepStride = 32;

// Output record of the (cell, edge) list.
struct CellEdge {
    uint32_t cell;        // x + cx*(y + cy*z)
    uint32_t edge;        // 4*axis + s0 + 2*s1
};

// One volume of a build.  A HistoPyramid is a batch of one volume; a batch
// concatenates the per-volume arrays (column counts, Morton offsets, scan
// tiles, classification workgroups) and records each volume's share here.
struct VolDesc {
    Geom g;
    const void* field;        // FP32 (or RGBA) lattice; null for snippets
    uint64_t colcnt_off;      // first u16 entry of the volume's column counts
    uint32_t l4_base;         // first entry of its Morton-order block offsets
    uint32_t n_level4;        // level-4 keys (side^2)
    uint32_t n_cc;            // column-count entries
    uint32_t cls_wg0;         // first logical classification workgroup
    uint32_t l4_tile0;        // first level-4 scan tile (global index)
    uint32_t n_l4tiles;
    uint32_t col_tile0;       // first column scan tile (index among column tiles)
    uint32_t n_coltiles;
    uint32_t lattice_base;    // first sample of the volume in a concatenated field (adjoint keys)
};

// First and one-past-last vertex of one volume of a batch.
struct VolCount {
    uint32_t base;
    uint32_t end;
};

// Emit work item: x = level-4 key; y = chunk (bits 0..3, first vertex
// chunk * kEmitChunk) | block vertex count (bits 4..15) | volume (16..31).
__device__ __forceinline__ uint32_t item_chunk_first(uint32_t y) { return (y & 0xfu) << 8; }
__device__ __forceinline__ uint32_t item_block_count(uint32_t y) { return (y >> 4) & 0xfffu; }
__device__ __forceinline__ uint32_t item_volume(uint32_t y) { return y >> 16; }

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

// Block key (Morton code of the block in the doubled base image) of the
// block holding cell (x0, y0) of slice z.
__device__ __forceinline__ uint32_t block_key(const Geom& g, int x0, int y0, int z)
{
    const uint32_t tx = (uint32_t)z & ((1u << g.tiles_x_log2) - 1u), ty = (uint32_t)z >> g.tiles_x_log2;
    const uint32_t X = ((tx << g.tsx2_log2) + (uint32_t)x0) >> g.blog2;
    const uint32_t Y = ((ty << g.tsy2_log2) + (uint32_t)y0) >> g.blog2;
    return morton_spread(X) | (morton_spread(Y) << 1);
}

// Vertex count of the level-4 block with Morton key `key` (0 outside the
// grid), read from the column-major count array.
__device__ __forceinline__ uint32_t key_count(const Geom& g, const uint16_t* __restrict__ colcnt, uint32_t key)
{
    const int X = (int)morton_compact(key) << g.blog2;
    const int Y = (int)morton_compact(key >> 1) << g.blog2;
    const int tx = X >> g.tsx2_log2, ty = Y >> g.tsy2_log2;
    const int z = (ty << g.tiles_x_log2) + tx;
    const int bx = (X & (g.tsx2 - 1)) >> g.blog2, by = (Y & (g.tsy2 - 1)) >> g.blog2;
    if (tx >= g.tiles_x || z >= g.cz || bx >= g.nbx || by >= g.nby) return 0u;
    const int pz = z >> g.zseg_log2, zi = z & (g.zseg - 1);
    return colcnt[(((size_t)pz * g.nby + by) * g.nbx + bx) * kColStride + zi];
}

// Orders this wave's LDS accesses around a phase change (each wave owns its
// LDS region; no other wave takes part).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Inclusive wave-wide scans over 64 lanes with DPP row shifts and row
// broadcasts (no LDS round trips): row_shr 1, 2, 4, 8 within each row of 16
// lanes, then row_bcast 15 and 31 carry row totals forward.  Inactive or
// out-of-row sources contribute the identity 0 (inputs are non-negative).
template <bool MAX>
__device__ __forceinline__ uint32_t wave_scan_u32(uint32_t x)
{
    auto op = [](uint32_t a, uint32_t b) { return MAX ? (a > b ? a : b) : a + b; };
    x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false));   // row_shr:1
    x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false));   // row_shr:2
    x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false));   // row_shr:4
    x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false));   // row_shr:8
    x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false));   // row_bcast:15
    x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false));   // row_bcast:31
    return x;
}

// A load from a wave-uniform address through the scalar data cache (read
// only; the data was written by an earlier kernel).  Uniform loads must not
// go through the vector memory path: a vector load's result has to be made
// uniform with v_readfirstlane, and the vmcnt wait that needs would also
// wait for every vector load issued before it (the prefetches).
template <class T>
__device__ __forceinline__ T load_uniform(const T* p)
{
    static_assert(sizeof(T) % 4 == 0, "loaded as 32-bit words");
    typedef const __attribute__((address_space(4))) uint32_t* cptr;
    const cptr q = (cptr)p;
    uint32_t w[sizeof(T) / 4];
#pragma unroll
    for (unsigned i = 0; i < sizeof(T) / 4; ++i) w[i] = q[i];
    T r;
    __builtin_memcpy(&r, w, sizeof(T));
    return r;
}

// A wave-uniform value loaded through the vector memory path, into the same
// vector register of every lane (make it uniform with v_readfirstlane at
// its use).  For prefetches inside a loop that also waits on LDS: a scalar
// load counts on lgkmcnt, and because scalar loads may complete out of
// order, any LDS wait issued while one is in flight has to wait for it too
// (lgkmcnt(0)), putting a memory round trip into every LDS wait; vector
// loads count on vmcnt, which completes in order and is waited on for the
// prefetched planes anyway.
__device__ __forceinline__ uint32_t prefetch_u32(const uint32_t* p)
{
    return (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(
        __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, 4, 0x00020000), 0, 0, 0);
}
__device__ __forceinline__ uint2 prefetch_u2(const uint2* p)
{
    typedef uint32_t v2u __attribute__((ext_vector_type(2)));
    const v2u v = __builtin_amdgcn_raw_buffer_load_b64(
        __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, 8, 0x00020000), 0, 0, 0);
    return make_uint2(v.x, v.y);
}
__device__ __forceinline__ uint2 uniform_u2(uint2 v)
{
    return make_uint2(__builtin_amdgcn_readfirstlane(v.x), __builtin_amdgcn_readfirstlane(v.y));
}

// The value of lane - 1 (0 in lane 0): DPP wave_shr:1.
__device__ __forceinline__ uint32_t wave_shr1_u32(uint32_t x)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x138, 0xf, 0xf, false);
}

// ---------------------------------------------------------------------------
// Field sources.  at(i,j,k) returns the sample at lattice index (i,j,k).

// The NT bit of a buffer load's cache policy (VolumeStreamField).
constexpr int kStreamLoad = 2;

// FP32 volume, x fastest; indices clamp to the lattice like a NEAREST/CLAMP
// 3D texture (apps/texture3d/texture3d.cpp:135-139).
//
// Besides at(), a field offers plane-wise access for the kernels that stage
// whole planes: lane_index(i, j) is a per-lane handle computed once,
// plane(k) a wave-uniform handle per plane, fetch(plane, index) the sample.
// For a volume these are a 32-bit in-plane offset and a plane pointer, so a
// load is a scalar base plus a vector offset.
struct VolumeField {
    const float* __restrict__ v;
    int sx, sy, sz;
    // linear index of the (clamped) sample a fetch at (i, j, k) reads, in
    // 32 bits: only for volumes of fewer than 2^32 samples (the adjoint's)
    __device__ __forceinline__ uint32_t index(int i, int j, int k) const
    {
        i = min(max(i, 0), sx - 1);
        j = min(max(j, 0), sy - 1);
        k = min(max(k, 0), sz - 1);
        return ((uint32_t)k * (uint32_t)sy + (uint32_t)j) * (uint32_t)sx + (uint32_t)i;
    }
    __device__ __forceinline__ float at(int i, int j, int k) const
    {
        i = min(max(i, 0), sx - 1);
        j = min(max(j, 0), sy - 1);
        k = min(max(k, 0), sz - 1);
        return v[((size_t)k * sy + j) * sx + i];
    }
    // Plane-wise access goes through a buffer resource per plane (a scalar
    // 128-bit descriptor) and 32-bit per-lane byte offsets.
    typedef __amdgpu_buffer_rsrc_t Plane;
    __device__ __forceinline__ uint32_t lane_index(int i, int j) const
    {
        i = min(max(i, 0), sx - 1);
        j = min(max(j, 0), sy - 1);
        return ((uint32_t)j * (uint32_t)sx + (uint32_t)i) * 4u;
    }
    __device__ __forceinline__ Plane plane(int k) const
    {
        k = min(max(k, 0), sz - 1);
        const float* base = v + (size_t)k * sy * sx;
        return __builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, sx * sy * 4, 0x00020000);
    }
    __device__ __forceinline__ float fetch(Plane p, uint32_t idx) const
    {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(p, idx, 0, 0));
    }
    // Row-wise access: a per-lane column offset and a wave-uniform row
    // offset (a scalar register of the load).
    __device__ __forceinline__ uint32_t col_index(int i) const { return (uint32_t)min(max(i, 0), sx - 1) * 4u; }
    __device__ __forceinline__ uint32_t row_index(int j) const
    {
        return (uint32_t)min(max(j, 0), sy - 1) * (uint32_t)sx * 4u;
    }
    __device__ __forceinline__ float fetch2(Plane p, uint32_t col, uint32_t row) const
    {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(p, col, row, 0));
    }
    // the sample value emission interpolates (fetch returns it as is)
    __device__ __forceinline__ float sample(Plane p, uint32_t idx) const { return fetch(p, idx); }
    // Windows that lie inside the plane, read without clamping: a per-lane
    // offset of the window sample (computed once per volume) and the
    // wave-uniform offset of the window's corner, the scalar offset of the
    // load, so that a window's loads take no per-item address arithmetic.
    static constexpr bool kWindows = true;
    __device__ __forceinline__ bool window_inside(int x0, int y0, int w) const { return x0 + w <= sx && y0 + w <= sy; }
    __device__ __forceinline__ uint32_t window_lane(int kx, int ky) const
    {
        return ((uint32_t)ky * (uint32_t)sx + (uint32_t)kx) * 4u;
    }
    __device__ __forceinline__ uint32_t window_base(int x0, int y0) const
    {
        return ((uint32_t)y0 * (uint32_t)sx + (uint32_t)x0) * 4u;
    }
    __device__ __forceinline__ float window_sample(Plane p, uint32_t lane_off, uint32_t base) const
    {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(p, lane_off, base, 0));
    }
};

// An FP32 volume whose classification row loads are non-temporal: for
// lattices larger than the 256 MiB MALL, which stream through classification
// once and cannot stay cached for emission anyway; their lines then leave
// L2 and the MALL to other data (512^3: classification 92.5 -> 87 us,
// emission 49 -> 52 us, 150 -> 146 us per build + extract; three frames in
// flight 131 -> 122 us per frame).  Smaller volumes and batches keep
// VolumeField: their emission finds the planes it re-reads in the MALL
// (a 256^3 volume measured 1.3% and the batched build 3.5% slower streamed).
struct VolumeStreamField : VolumeField {
    __device__ __forceinline__ float fetch2(Plane p, uint32_t col, uint32_t row) const
    {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(p, col, row, kStreamLoad));
    }
};

// 8-bit normalised volume, the texture3d demo's field (a GL_ALPHA8 texture
// sampled NEAREST with clamping, texture3d.cpp:127-141): sample value
// c / 255, the GL unsigned-normalised conversion, correctly rounded.
// Classification reads raw bytes (fetch, fetch2 return c as a float) and is
// launched with the integer threshold kT = min{c : c / 255 >= T}, so c < kT
// is the same predicate as c / 255 < T; emission stages the converted
// values (sample) and runs unchanged.
struct Volume8Field {
    const uint8_t* __restrict__ v;
    int sx, sy, sz;
    __device__ __forceinline__ float at(int i, int j, int k) const
    {
        i = min(max(i, 0), sx - 1);
        j = min(max(j, 0), sy - 1);
        k = min(max(k, 0), sz - 1);
        return (float)v[((size_t)k * sy + j) * sx + i] / 255.0f;
    }
    typedef __amdgpu_buffer_rsrc_t Plane;
    __device__ __forceinline__ uint32_t lane_index(int i, int j) const
    {
        i = min(max(i, 0), sx - 1);
        j = min(max(j, 0), sy - 1);
        return (uint32_t)j * (uint32_t)sx + (uint32_t)i;
    }
    __device__ __forceinline__ Plane plane(int k) const
    {
        k = min(max(k, 0), sz - 1);
        const uint8_t* base = v + (size_t)k * sy * sx;
        return __builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, sx * sy, 0x00020000);
    }
    __device__ __forceinline__ float fetch(Plane p, uint32_t idx) const
    {
        return (float)__builtin_amdgcn_raw_buffer_load_b8(p, idx, 0, 0);
    }
    __device__ __forceinline__ uint32_t col_index(int i) const { return (uint32_t)min(max(i, 0), sx - 1); }
    __device__ __forceinline__ uint32_t row_index(int j) const
    {
        return (uint32_t)min(max(j, 0), sy - 1) * (uint32_t)sx;
    }
    __device__ __forceinline__ float fetch2(Plane p, uint32_t col, uint32_t row) const
    {
        return (float)__builtin_amdgcn_raw_buffer_load_b8(p, col, row, 0);
    }
    __device__ __forceinline__ float sample(Plane p, uint32_t idx) const { return fetch(p, idx) / 255.0f; }
    // windows inside the plane without clamping (as VolumeField's)
    static constexpr bool kWindows = true;
    __device__ __forceinline__ bool window_inside(int x0, int y0, int w) const { return x0 + w <= sx && y0 + w <= sy; }
    __device__ __forceinline__ uint32_t window_lane(int kx, int ky) const
    {
        return (uint32_t)ky * (uint32_t)sx + (uint32_t)kx;
    }
    __device__ __forceinline__ uint32_t window_base(int x0, int y0) const
    {
        return (uint32_t)y0 * (uint32_t)sx + (uint32_t)x0;
    }
    __device__ __forceinline__ float window_sample(Plane p, uint32_t lane_off, uint32_t base) const
    {
        return (float)__builtin_amdgcn_raw_buffer_load_b8(p, lane_off, base, 0) / 255.0f;
    }
};

// RGBA FP32 volume carrying the gradient in rgb and the scalar in alpha, the
// layout of a gradient-providing 3D texture (hpiface.cpp:164-184,
// shadergen.cpp:256-266).  Classification reads the alpha channel; grad()
// returns the whole texel (gradient, value).
struct Volume4Field {
    const float4* __restrict__ v;
    int sx, sy, sz;
    __device__ __forceinline__ size_t texel(int i, int j, int k) const
    {
        i = min(max(i, 0), sx - 1);
        j = min(max(j, 0), sy - 1);
        k = min(max(k, 0), sz - 1);
        return ((size_t)k * sy + j) * sx + i;
    }
    __device__ __forceinline__ float at(int i, int j, int k) const { return v[texel(i, j, k)].w; }
    __device__ __forceinline__ float4 grad(int i, int j, int k) const { return v[texel(i, j, k)]; }
    typedef __amdgpu_buffer_rsrc_t Plane;
    __device__ __forceinline__ uint32_t lane_index(int i, int j) const
    {
        i = min(max(i, 0), sx - 1);
        j = min(max(j, 0), sy - 1);
        return ((uint32_t)j * (uint32_t)sx + (uint32_t)i) * 16u;
    }
    // A plane's resource starts at the alpha of its first texel, so texel
    // offsets are multiples of 16 and the last alpha ends at 16 (sx sy) - 12
    // bytes, which fits the 32-bit record count up to a 16384 x 16384 plane.
    __device__ __forceinline__ Plane plane(int k) const
    {
        k = min(max(k, 0), sz - 1);
        const char* base = (const char*)(v + (size_t)k * sy * sx) + 12;
        return __builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, (int)(uint32_t)((uint64_t)sx * sy * 16u - 12u),
                                                 0x00020000);
    }
    __device__ __forceinline__ float fetch(Plane p, uint32_t idx) const
    {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(p, idx, 0, 0));
    }
    __device__ __forceinline__ float sample(Plane p, uint32_t idx) const { return fetch(p, idx); }
    // windows inside the plane without clamping (as VolumeField's)
    static constexpr bool kWindows = true;
    __device__ __forceinline__ bool window_inside(int x0, int y0, int w) const { return x0 + w <= sx && y0 + w <= sy; }
    __device__ __forceinline__ uint32_t window_lane(int kx, int ky) const
    {
        return ((uint32_t)ky * (uint32_t)sx + (uint32_t)kx) * 16u;
    }
    __device__ __forceinline__ uint32_t window_base(int x0, int y0) const
    {
        return ((uint32_t)y0 * (uint32_t)sx + (uint32_t)x0) * 16u;
    }
    __device__ __forceinline__ float window_sample(Plane p, uint32_t lane_off, uint32_t base) const
    {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(p, lane_off, base, 0));
    }
    __device__ __forceinline__ uint32_t col_index(int i) const { return (uint32_t)min(max(i, 0), sx - 1) * 16u; }
    __device__ __forceinline__ uint32_t row_index(int j) const
    {
        return (uint32_t)min(max(j, 0), sy - 1) * (uint32_t)sx * 16u;
    }
    __device__ __forceinline__ float fetch2(Plane p, uint32_t col, uint32_t row) const
    {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(p, col, row, 0));
    }
};

// ---------------------------------------------------------------------------
// classify: one workgroup per (row band of 16 cells) x (zseg slices); its
// waves sit side by side along x, 64 lattice columns each, so a workgroup of
// ceil(X/64) waves covers whole lattice rows (up to 1024 columns; wider
// lattices take several workgroups per band, joined by one extra column).
//
// Each plane is read once as 17 rows of 64 samples per wave and turned into
// 64-bit ballot masks of (f < T); the codes of a whole cell row come out of
// four masks.  The sample right of a wave's last lane belongs to the next
// wave, whose lane-0 bits are passed through LDS, so no extra column is
// loaded inside a workgroup.  Loads use clamped, always-valid addresses so a
// plane's loads issue back to back, and the next plane is in flight while
// the current one is classified.  Slices whose two planes are uniformly
// outside or inside cannot hold a surface cell and are skipped with scalar
// tests.  Counts are summed per lane over the 16 rows of a block and across
// 16-lane groups, giving four level-4 sums per wave and slice.
constexpr int kClassifyMaxWaves = 16;

template <class F>
__device__ __forceinline__ void classify_load_plane(const F& field, uint32_t col, const uint32_t (&row)[kBlock + 1],
                                                    uint32_t eoff, bool load_e, int zz, float (&v)[kBlock + 1],
                                                    float& ve)
{
    const typename F::Plane pl = field.plane(zz);
#pragma unroll
    for (int r = 0; r <= kBlock; ++r) v[r] = field.fetch2(pl, col, row[r]);
    if (load_e) ve = field.fetch(pl, eoff);
}

struct ClassifyWave {
    int lane, x0, y0, x, sy, cx, cy;
    bool xin, xcell, has_right, lane_e;
};

// Per-wave state carried from plane to plane.
struct ClassifyCarry {
    uint32_t col;   // per lane: previous plane's inside bits of this lane's column, bit r = row r
    uint64_t eb;    // previous plane: inside bits of the column right of the wave (uniform)
    uint64_t any;   // previous plane: some lane's column has an inside sample (uniform)
    uint64_t all;   // previous plane: every sample inside, incl. the right column (uniform)
};

// The 17 rows' inside bits of this lane's column (bit r = row r).
__device__ __forceinline__ uint32_t classify_column(const ClassifyWave& w, float T, const float (&v)[kBlock + 1])
{
    uint32_t col = 0;
#pragma unroll
    for (int r = 0; r <= kBlock; ++r) col |= ((w.xin && (w.y0 + r < w.sy) && v[r] < T) ? 1u : 0u) << r;
    return col;
}

// Interleaves a column's bits with its right neighbour's: bit 2r = own
// row r, bit 2r+1 = neighbour row r.  Bits 2r..2r+3 are then the four
// corners of cell row r on this plane in HPMC code order (x + 2y).
__device__ __forceinline__ uint64_t classify_window(uint32_t col, uint32_t right)
{
    const uint64_t a = (uint64_t)morton_spread(col & 0xffffu) | ((uint64_t)((col >> 16) & 1u) << 32);
    const uint64_t b = (uint64_t)morton_spread(right & 0xffffu) | ((uint64_t)((right >> 16) & 1u) << 32);
    return a | (b << 1);
}

// Classifies slice zz-1 from the column bits of planes zz-1 (clo) and zz
// (chi): 16 cell rows of 64 cells, summed into four level-4 blocks.  Only
// called when the two planes are mixed.
__device__ __forceinline__ uint32_t classify_slice_counts(const ClassifyWave& w, uint32_t clo, uint32_t chi,
                                                          uint64_t eb_lo, uint64_t eb_hi,
                                                          const uint8_t* __restrict__ lut)
{
    uint32_t rlo = __shfl_down(clo, 1), rhi = __shfl_down(chi, 1);
    if (w.lane == 63) {
        rlo = (uint32_t)eb_lo;
        rhi = (uint32_t)eb_hi;
    }
    const uint64_t wlo = classify_window(clo, rlo), whi = classify_window(chi, rhi);
    uint32_t acc = 0;
#pragma unroll
    for (int r = 0; r < kBlock; ++r) {
        const uint32_t code = (uint32_t)((wlo >> (2 * r)) & 0xf) | ((uint32_t)((whi >> (2 * r)) & 0xf) << 4);
        const bool valid = w.xcell && (w.y0 + r < w.cy);
        const uint32_t cnt = lut[code];
        acc += valid ? cnt : 0u;
    }
    // 16-lane group sums = level-4 sums of four blocks.
    acc += __shfl_xor(acc, 8, 16);
    acc += __shfl_xor(acc, 4, 16);
    acc += __shfl_xor(acc, 2, 16);
    acc += __shfl_xor(acc, 1, 16);
    return acc;
}

// Logical workgroup t of a volume -> (x band, y band, z segment), x fastest.
// The launch deals workgroups to the eight XCDs round robin and maps them
// to logical indices so that each XCD runs a contiguous run (classify
// kernel wrapper): workgroups resident together on one XCD are y-neighbours
// and share their halo rows in that XCD's L2.
// A workgroup of nw waves spans wx = min(nw, ceil(X / 64)) waves across x
// and stacks nw / wx row bands, so narrow lattices use every wave.
__device__ __forceinline__ void classify_shape(const Geom& g, int nw, int& wx, int& nstack)
{
    wx = min(nw, (g.sx + 63) >> 6);
    nstack = nw / wx;
}

__device__ __forceinline__ void classify_tile(const Geom& g, int nstack, int wgx, int t, int& bx, int& pyg, int& pz)
{
    const int npy = (g.cy + kBlock - 1) >> kBlockLog2;
    const int npyg = (npy + nstack - 1) / nstack;
    bx = t % wgx;
    const int rest = t / wgx;
    pyg = rest % npyg;
    pz = rest / npyg;
}

template <class F>
__device__ __forceinline__ void classify_body(const F& field, const Geom& g, float T,
                                              uint16_t* __restrict__ colcnt, int zseg,
                                              const uint8_t* __restrict__ lut,
                                              uint32_t (*xch)[kClassifyMaxWaves],
                                              uint16_t (*stage)[4][kColStride], int t,
                                              uint2 (*red)[kColStride], uint2* __restrict__ tsum_l4,
                                              uint2* __restrict__ tsum_col)
{
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nw = blockDim.x >> 6;
    int wx, nstack;
    classify_shape(g, nw, wx, nstack);
    const int wgx = (g.sx + 64 * wx - 1) / (64 * wx);    // workgroups across x
    int bx, pyg, pz;
    classify_tile(g, nstack, wgx, t, bx, pyg, pz);
    bx = __builtin_amdgcn_readfirstlane(bx);
    pyg = __builtin_amdgcn_readfirstlane(pyg);
    pz = __builtin_amdgcn_readfirstlane(pz);
    const int xw = wv % wx, yw = wv / wx;   // this wave's place: column group, stacked band
    ClassifyWave w;
    w.lane = lane;
    w.x0 = (bx * wx + xw) << 6;
    w.y0 = (pyg * nstack + yw) << kBlockLog2;
    w.x = w.x0 + lane;
    w.sy = g.sy;
    w.cx = g.cx;
    w.cy = g.cy;
    w.xin = w.x < g.sx;
    w.xcell = w.x < g.cx;
    w.has_right = w.x0 + 64 < g.sx;
    // Waves with no lattice columns or no cell rows (past the stacked bands,
    // past the lattice) leave at once: the barriers below wait only for the
    // waves still running.  The last wave of a band's row of waves takes the
    // right-neighbour role.
    if (yw >= nstack || w.x0 >= g.sx || w.y0 >= g.cy) return;
    const bool last = xw == wx - 1;
    // right column from the next workgroup (wave-uniform: a scalar branch)
    const bool load_e = __builtin_amdgcn_readfirstlane((int)(last && w.has_right)) != 0;
    w.lane_e = load_e && lane <= kBlock && (w.y0 + lane < g.sy);
    const int z0 = pz * zseg;
    const int z1 = min(z0 + zseg, g.cz);   // cell slices [z0, z1): planes z0..z1
    // row offsets of the 17 rows (clamped, wave-uniform), the lane's column
    // offset, and the per-lane offset into the column right of the workgroup
    uint32_t roff[kBlock + 1];
#pragma unroll
    for (int r = 0; r <= kBlock; ++r) roff[r] = __builtin_amdgcn_readfirstlane(field.row_index(w.y0 + r));
    const uint32_t coff = field.col_index(w.x);
    const uint32_t eoff = field.lane_index(w.x0 + 64, w.y0 + min(lane, kBlock));
    const uint64_t erows = (1ull << (kBlock + 1)) - 1ull;

    // Two row buffers: the current plane and the next one in flight.  The
    // prefetch is unconditional (the last one re-reads plane z1) so the wait
    // for one buffer never covers the other's loads.  On the common path a
    // plane costs 17 compares folded into per-lane column bits and two
    // ballots; codes are formed only for slices whose planes are mixed.
    const uint32_t full = (1u << (kBlock + 1)) - 1u;
    // the four block columns' slice counts are staged in LDS and written as
    // whole 64-byte lines after the last slice
    uint16_t(&mycnt)[4][kColStride] = stage[wv];
    mycnt[lane >> 4][(lane & 15) * 2] = 0;
    mycnt[lane >> 4][(lane & 15) * 2 + 1] = 0;
    float A[kBlock + 1], B[kBlock + 1], ea = 0.0f, eb_ = 0.0f;
    ClassifyCarry c{};
    classify_load_plane(field, coff, roff, eoff, load_e, z0, A, ea);
    auto step = [&](const float(&cur)[kBlock + 1], float& cure, float(&nxt)[kBlock + 1], float& nxte, int zz,
                    int parity) {
        classify_load_plane(field, coff, roff, eoff, load_e, min(zz + 1, z1), nxt, nxte);
        __builtin_amdgcn_sched_barrier(0);
        const uint32_t col = classify_column(w, T, cur);
        uint64_t any = __ballot(col != 0);
        uint64_t all = __ballot(col == full);
        const uint32_t first = __builtin_amdgcn_readfirstlane(col);
        if (lane == 0) xch[parity][wv] = first;
        __syncthreads();
        uint64_t eb;
        if (!last) eb = __builtin_amdgcn_readfirstlane(xch[parity][wv + 1]);
        else eb = __ballot(w.lane_e && cure < T);
        any |= eb;
        // all-inside needs every row's right-neighbour sample inside as well
        if (!w.has_right || (eb & erows) != erows) all = 0;
        if (zz > z0) {
            uint32_t acc = 0;
            const bool uniform_out = (c.any | any) == 0ull;
            const bool uniform_in = (c.all & all) == ~0ull;
            if (!(uniform_out || uniform_in)) acc = classify_slice_counts(w, c.col, col, c.eb, eb, lut);
            if ((lane & 15) == 0) mycnt[lane >> 4][zz - 1 - z0] = (uint16_t)acc;
        }
        c.col = col;
        c.eb = eb;
        c.any = any;
        c.all = all;
    };
    for (int zz = z0;;) {
        step(A, ea, B, eb_, zz, 0);
        if (++zz > z1) break;
        step(B, eb_, A, ea, zz, 1);
        if (++zz > z1) break;
    }
    // lane l writes entries 2(l%16), 2(l%16)+1 of block column l/16
    wave_lds_sync();
    const int bcol = (w.x0 >> kBlockLog2) + (lane >> 4);
    const uint32_t wi = ((uint32_t)pz * g.nby + (uint32_t)(w.y0 >> kBlockLog2)) * g.nbx + bcol;
    uint32_t vsum = 0, isum = 0;
    if (bcol < g.nbx) {
        const uint32_t lo = mycnt[lane >> 4][(lane & 15) * 2], hi = mycnt[lane >> 4][(lane & 15) * 2 + 1];
        uint32_t* col32 = (uint32_t*)(colcnt + (size_t)wi * kColStride);
        col32[lane & 15] = lo | (hi << 16);
        vsum = lo + hi;
        isum = chunks_of(lo) + chunks_of(hi);
    }

    // Scan tile sums, added after the last load so that the atomics never
    // hold up the stream.  Column tiles: each block column's vertices and
    // emit items.
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) {
        vsum += __shfl_xor(vsum, o, 16);
        isum += __shfl_xor(isum, o, 16);
    }
    if ((lane & 15) == 0 && vsum) {
        const uint32_t t = wi * (uint32_t)kColStride / (uint32_t)kScanTile;
        atomicAdd(&tsum_col[t].x, vsum);
        atomicAdd(&tsum_col[t].y, isum);
    }
    // Level-4 tiles: per slice, the vertices and non-empty blocks of the
    // wave's four blocks (which share one Morton tile).  When a whole slice
    // lies in one Morton tile (slice tiles of at most 64 x 64 blocks) the
    // workgroup's waves are summed first, one atomic pair per slice.
    const int ns = z1 - z0;
    uint32_t sv = 0, sb = 0;
    if (lane < ns) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t cj = mycnt[j][lane];
            sv += cj;
            sb += cj != 0;
        }
    }
    const bool whole = g.tsx2 <= 64 * kBlock && g.tsy2 <= 64 * kBlock;
    if (!whole) {
        if (lane < ns && sv) {
            const uint32_t t = block_key(g, w.x0, w.y0, z0 + lane) / (uint32_t)kScanTile;
            atomicAdd(&tsum_l4[t].x, sv);
            atomicAdd(&tsum_l4[t].y, sb);
        }
        return;
    }
    if (lane < ns) red[wv][lane] = make_uint2(sv, sb);
    __syncthreads();
    if (wv == 0 && lane < ns) {
        uint2 r = make_uint2(0, 0);
        for (int i = 0; i < nw; ++i) {
            // waves that left wrote nothing
            const int ix = i % wx, iy = i / wx;
            if (iy >= nstack || ((bx * wx + ix) << 6) >= g.sx || ((pyg * nstack + iy) << kBlockLog2) >= g.cy) continue;
            r.x += red[i][lane].x;
            r.y += red[i][lane].y;
        }
        if (r.x) {
            const uint32_t t = block_key(g, 0, w.y0, z0 + lane) / (uint32_t)kScanTile;
            atomicAdd(&tsum_l4[t].x, r.x);
            atomicAdd(&tsum_l4[t].y, r.y);
        }
    }
}

// Classification of grids whose blocks are narrower than 16 cells (at most
// 8 cells along x or y): one thread per column-count entry (z segment, block
// row, block column, slice slot) adds up the table counts of the block's
// cells in that slice, writes the entry (zero for slots past the grid) and
// adds it into the scan tile sums.  Samples come from the field's plane-wise
// fetch, which is what the classification of larger grids compares too
// (raw bytes for 8-bit volumes, classified against the byte threshold).
template <class F>
__device__ __forceinline__ void classify_small_body(const F& field, const Geom& g, float T,
                                                    uint16_t* __restrict__ colcnt, uint32_t n_cc,
                                                    const uint8_t* __restrict__ lut, uint2* __restrict__ tsum_l4,
                                                    uint2* __restrict__ tsum_col)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_cc) return;
    const int zi = (int)(t % (uint32_t)kColStride);
    const uint32_t w = t / (uint32_t)kColStride;
    const int bx = (int)(w % (uint32_t)g.nbx);
    const uint32_t rest = w / (uint32_t)g.nbx;
    const int by = (int)(rest % (uint32_t)g.nby);
    const int pz = (int)(rest / (uint32_t)g.nby);
    const int z = pz * g.zseg + zi;
    const int bs = 1 << g.blog2;
    uint32_t cnt = 0;
    if (zi < g.zseg && z < g.cz) {
        const typename F::Plane p0 = field.plane(z), p1 = field.plane(z + 1);
        const int x1 = min((bx + 1) * bs, g.cx), y1 = min((by + 1) * bs, g.cy);
        for (int y = by * bs; y < y1; ++y)
            for (int x = bx * bs; x < x1; ++x) {
                uint32_t code = 0;
#pragma unroll
                for (int b = 0; b < 8; ++b) {
                    const uint32_t idx = field.lane_index(x + (b & 1), y + ((b >> 1) & 1));
                    const float v = field.fetch((b & 4) ? p1 : p0, idx);
                    code |= (v < T ? 1u : 0u) << b;
                }
                cnt += lut[code];
            }
    }
    colcnt[t] = (uint16_t)cnt;
    if (cnt) {
        atomicAdd(&tsum_col[t / (uint32_t)kScanTile].x, cnt);
        atomicAdd(&tsum_col[t / (uint32_t)kScanTile].y, chunks_of(cnt));
        const uint32_t tile = block_key(g, bx * bs, by * bs, z) / (uint32_t)kScanTile;
        atomicAdd(&tsum_l4[tile].x, cnt);
        atomicAdd(&tsum_l4[tile].y, 1u);
    }
}

// ---------------------------------------------------------------------------
// scan helpers

__device__ __forceinline__ uint2 block_reduce_u2(uint2 v, uint2* sh)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    // wave totals by DPP scans (no LDS round trips; every lane active)
    v.x = (uint32_t)__builtin_amdgcn_readlane(wave_scan_u32<false>(v.x), 63);
    v.y = (uint32_t)__builtin_amdgcn_readlane(wave_scan_u32<false>(v.y), 63);
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

// The kScanPer column-major block counts of thread t of a tile (u16, read
// as 32-bit words that merge into vector loads); entries past n read as zero.
__device__ __forceinline__ void scan_load16_u16(const uint16_t* __restrict__ cc, uint32_t n, uint32_t base,
                                                uint32_t (&c)[kScanPer])
{
    static_assert(kScanPer % 2 == 0, "pairs of u16 per word");
    if (base + kScanPer <= n) {
        const uint32_t* p = (const uint32_t*)(cc + base);
#pragma unroll
        for (int i = 0; i < kScanPer / 2; ++i) {
            const uint32_t w = p[i];
            c[2 * i] = w & 0xffffu;
            c[2 * i + 1] = w >> 16;
        }
    } else {
#pragma unroll
        for (int k = 0; k < kScanPer; ++k) c[k] = (base + k < n) ? cc[base + k] : 0u;
    }
}

// The 16 level-4 counts of keys base .. base+15 (base a multiple of 16),
// gathered from the column-major counts, all 16 loads in flight together.
// The keys are a 4x4 square of blocks (key bits 0-3 = x0, y0, x1, y1).
// When slice tiles are at least 4 blocks across both ways the square lies
// in one slice tile: its slice, z segment and first count are decoded once
// and the 16 loads step by constant block offsets (a key decoded per load
// took ~75 instructions, most of the scan's time).  Counts past the grid
// read entry 0 and are dropped.
__device__ __forceinline__ void scan_gather16(const Geom& g, const uint16_t* __restrict__ colcnt, uint32_t n,
                                              uint32_t base, uint32_t (&c)[kScanPer])
{
    static_assert(kScanPer == 16, "a thread's keys are a 4x4 block square");
    const bool square = g.tsx2_log2 - g.blog2 >= 2 && g.tsy2_log2 - g.blog2 >= 2;
    if (!square || base + kScanPer > n) {
#pragma unroll
        for (int k = 0; k < kScanPer; ++k) c[k] = (base + k < n) ? key_count(g, colcnt, base + k) : 0u;
        return;
    }
    const int X = (int)morton_compact(base) << g.blog2, Y = (int)morton_compact(base >> 1) << g.blog2;
    const int tx = X >> g.tsx2_log2, ty = Y >> g.tsy2_log2;
    const int z = (ty << g.tiles_x_log2) + tx;
    const int bx = (X & (g.tsx2 - 1)) >> g.blog2, by = (Y & (g.tsy2 - 1)) >> g.blog2;
    const bool zin = tx < g.tiles_x && z < g.cz;
    const uint32_t first =
        ((((uint32_t)(z >> g.zseg_log2) * (uint32_t)g.nby + (uint32_t)by) * (uint32_t)g.nbx + (uint32_t)bx) *
         (uint32_t)kColStride) + (uint32_t)(z & (g.zseg - 1));
    const uint32_t row = (uint32_t)g.nbx * (uint32_t)kColStride;
    uint32_t v[kScanPer];
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        const int dx = (k & 1) | ((k >> 1) & 2), dy = ((k >> 1) & 1) | ((k >> 2) & 2);
        const bool in = zin && bx + dx < g.nbx && by + dy < g.nby;
        v[k] = colcnt[in ? first + (uint32_t)dy * row + (uint32_t)(dx * kColStride) : 0u];
        c[k] = in;
    }
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) c[k] = c[k] ? v[k] : 0u;
}

// Scan tiles of a build: [0, n_l4t) are level-4 tiles (4096 Morton keys of
// one volume each, volumes in order), [n_l4t, n_l4t + n_colt) column tiles
// (4096 column-count entries of one volume each).  tile_vol[tile] names the
// volume.  Level-4 tiles sum vertices (x) and non-empty blocks (y); column
// tiles sum vertices (x) and emit work items (y).
__device__ __forceinline__ void scan_load_tile(const VolDesc& d, const uint16_t* __restrict__ colcnt, bool is_col,
                                               uint32_t local_tile, uint32_t (&c)[kScanPer])
{
    const uint32_t base = local_tile * kScanTile + threadIdx.x * kScanPer;
    if (!is_col) scan_gather16(d.g, colcnt + d.colcnt_off, d.n_level4, base, c);
    else scan_load16_u16(colcnt + d.colcnt_off, d.n_cc, base, c);
}

// One-pass scan.  The tile sums were accumulated by the classification
// kernel, so every tile reads the sums of the tiles before it in its
// sequence (a few hundred), scans its own 4096 entries and scatters.
// Level-4 tiles write, for each non-empty block, the global offset of its
// first vertex: the concatenation of the volumes' vertex lists, each in
// Morton key order (the reference's key order); a volume's first and last
// level-4 tiles write its vertex base and end.  Column tiles write the emit
// work items -- one per kEmitChunk vertices of a block -- in (volume, z
// segment, y, x, z) order, so that consecutive items are consecutive slices
// of one block column.  The last level-4 tile and the last column tile write
// the totals, to device memory and to the host-visible copy.  The tile
// sums stay as they are: the extraction that follows clears them for the
// next build (or the host, before a build that follows no extraction), so
// no workgroup waits to learn that it is the last.
__device__ __forceinline__ void scan_body(
    const VolDesc* __restrict__ vols, const uint32_t* __restrict__ tile_vol, uint32_t n_l4t, uint32_t n_colt,
    const uint16_t* __restrict__ colcnt, const uint2* __restrict__ tile_sums, uint32_t* __restrict__ item_reps,
    uint32_t* __restrict__ l4_off, uint2* __restrict__ items, Totals* __restrict__ totals,
    Totals* __restrict__ host_totals, VolCount* __restrict__ vol_counts, VolCount* __restrict__ host_vol_counts)
{
    __shared__ uint2 sh[16];
    __shared__ uint2 wsum[16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t ntiles = n_l4t + n_colt;
    const uint32_t tile = blockIdx.x;
    const bool is_col = tile >= n_l4t;
    const uint32_t t0 = is_col ? n_l4t : 0u;
    // preceding tile sums, loaded alongside the tile's own entries
    uint2 pre = make_uint2(0, 0);
    for (uint32_t i = t0 + threadIdx.x; i < tile; i += blockDim.x) {
        const uint2 v = tile_sums[i];
        pre.x += v.x;
        pre.y += v.y;
    }
    // tile_vol == nullptr: a plan of one volume
    const uint32_t vol = tile_vol ? __builtin_amdgcn_readfirstlane(tile_vol[tile]) : 0u;
    const VolDesc d = vols[vol];
    const uint32_t local = is_col ? tile - n_l4t - d.col_tile0 : tile - d.l4_tile0;
    uint32_t c[kScanPer];
    scan_load_tile(d, colcnt, is_col, local, c);
    uint2 mine = make_uint2(0, 0);
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        if (!is_col) mine.x += c[k];
        else mine.y += chunks_of(c[k]);
    }
    pre = block_reduce_u2(pre, sh);
    // inclusive wave scans by DPP (every lane active)
    const uint2 inc = make_uint2(wave_scan_u32<false>(mine.x), wave_scan_u32<false>(mine.y));
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    // A volume's first level-4 tile knows its vertex base, its last one the
    // end; the last tile of each sequence knows the grand totals (the sums
    // of the tiles before it plus its own).
    if (threadIdx.x == 0) {
        uint2 all = pre;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) {
            all.x += wsum[i].x;
            all.y += wsum[i].y;
        }
        if (!is_col) {
            if (local == 0) {
                vol_counts[vol].base = pre.x;
                if (host_vol_counts) host_vol_counts[vol].base = pre.x;
            }
            if (local == d.n_l4tiles - 1) {
                vol_counts[vol].end = all.x;
                if (host_vol_counts) host_vol_counts[vol].end = all.x;
            }
            if (tile == n_l4t - 1) {
                totals->vertices = all.x;
                totals->blocks = all.y;
                if (host_totals) {
                    host_totals->vertices = all.x;
                    host_totals->blocks = all.y;
                }
            }
        } else if (tile == ntiles - 1) {
            totals->items = all.y;
            if (host_totals) host_totals->items = all.y;
            for (uint32_t r = 0; r < kItemReps; ++r) item_reps[r * kItemRepStride] = all.y;
        }
    }
    uint2 off = make_uint2(pre.x + inc.x - mine.x, pre.y + inc.y - mine.y);
    for (int i = 0; i < w; ++i) {
        off.x += wsum[i].x;
        off.y += wsum[i].y;
    }
    const uint32_t base = local * kScanTile + threadIdx.x * kScanPer;
    if (!is_col) {
#pragma unroll
        for (int k = 0; k < kScanPer; ++k) {
            if (c[k]) l4_off[d.l4_base + base + k] = off.x;
            off.x += c[k];
        }
    } else {
        const Geom& g = d.g;
        // a thread's 16 entries lie in one 32-slot block column
        static_assert(kColStride % kScanPer == 0, "column entries of a thread span one column");
        const uint32_t wi = base / (uint32_t)kColStride;
        const int bxb = (int)(wi % (uint32_t)g.nbx);
        const uint32_t rest = wi / (uint32_t)g.nbx;
        const int by = (int)(rest % (uint32_t)g.nby);
        const int pz = (int)(rest / (uint32_t)g.nby);
        // The thread's slices z0 + k: z0 is a multiple of 16, or of 8 for
        // 8-slice segments, whose slots past 8 are empty.  With at least 16
        // slices per row of slice tiles the run stays in one tile row, and
        // key(z0 + k) = key(z0) | spread(k) << (2 * log2 of the blocks across
        // a tile), spread(k) a constant of the unrolled loop.
        const int z0 = pz * g.zseg + (int)(base % (uint32_t)kColStride);
        const bool one_row = g.tiles_x_log2 >= 4;
        const uint32_t key0 = block_key(g, bxb << g.blog2, by << g.blog2, z0);
        const uint32_t kshift = 2u * (uint32_t)(g.tsx2_log2 - g.blog2);
        // (the first chunk of each non-empty entry is a plain store; only
        // blocks of more than kEmitChunk vertices take the chunk loop)
        auto scatter = [&](auto key_of) {
#pragma unroll
            for (int k = 0; k < kScanPer; ++k) {
                const uint32_t key = key_of(k);
                const uint32_t yv = (c[k] << 4) | (vol << 16);
                if (c[k]) items[off.y] = make_uint2(key, yv);
                off.y += c[k] ? 1u : 0u;
                if (c[k] > (uint32_t)kEmitChunk)
                    for (uint32_t ch = 1; ch * kEmitChunk < c[k]; ++ch) items[off.y++] = make_uint2(key, yv | ch);
            }
        };
        if (one_row) scatter([&](int k) { return key0 | (morton_spread((uint32_t)k) << kshift); });
        else scatter([&](int k) { return block_key(g, bxb << g.blog2, by << g.blog2, z0 + k); });
    }
}

// ---------------------------------------------------------------------------
// emit: a persistent grid over the work items of the column scan; each item
// is up to kEmitChunk vertices of one non-empty level-4 block.  The four
// waves of a workgroup take four consecutive items and a workgroup's
// successor on the same XCD the next four, so blocks sharing lattice planes
// (consecutive slices of a column) are processed together in one L2.
//
// A block needs lattice planes z, z+1 (corners) and z+2 (forward
// differences) over 18x18 samples, staged in the wave's LDS; the next
// item's planes and output offset are loaded while the current item runs.
// For a block the 64 lanes classify its 256 cells (a 2x2 quad each, Morton
// order), build the block's pyramid levels 0..3 in LDS (cells, quads, 16-
// and 64-cell groups: four children each, ordered (0,0), (1,0), (0,1), (1,1)
// as in the reference's traversal, shadergen.cpp:448-495), and then each
// lane takes output vertices, descends the four levels to its cell and
// writes the interpolated position and normal at the block's offset in the
// reference's vertex order.
constexpr int kEmitWaves = 4;
constexpr int kPlaneLoads = (kBrickPlane + 63) / 64;   // 6 per lane per plane

// Emission modes (shadergen.cpp:505-563):
//   continuous: t from the two edge samples, normal from forward differences
//   binary:     edge midpoint, normal from the per-case table (threshold 0.5)
//   gradient:   t from the two samples' values, normal from their gradients
enum EmitMode { kEmitContinuous = 0, kEmitBinary = 1, kEmitGradient = 2 };

// Per-vertex callback of user-fused traversal kernels: emission calls
// cb(k, normal, position) for output vertex k instead of writing the
// output buffer.  The default does nothing and leaves emission as it is.
// A callback that declares kWantsEdge = true is called as
// cb(k, a, b, normal, position) instead, with a and b the texel-centre
// coordinates ((i + 0.5) / lattice size per axis, the coordinates a field
// snippet's HPMC_fetch takes) of the two ends of the vertex's edge: the
// reference's four-output extractVertex(a, b, p, n)
// (hpmc/src/shadergen.cpp:310, 390-391).
struct NoVertexCallback {
    static constexpr bool kActive = false;
    __device__ void operator()(uint32_t, float3, float3) const {}
};

template <class CB, class = void>
struct wants_edge {
    static constexpr bool value = false;
};
template <class CB>
struct wants_edge<CB, decltype((void)CB::kWantsEdge, void())> {
    static constexpr bool value = CB::kWantsEdge;
};

// Lattice planes an emit item stages: z, z+1 for the corners, z+2 for the
// forward differences of continuous mode.
template <int MODE>
constexpr int emit_planes() { return MODE == kEmitContinuous ? 3 : 2; }

// A staged plane holds the 18x18 window in kPlaneStride slots (6 loads of 64
// lanes, no tail); planes sit kPlaneStride floats apart so a sample and its
// +z neighbour are one ds_read2st64 apart.
constexpr int kPlaneStride = 64 * kPlaneLoads;   // 384
// Row stride of a staged window in LDS.  A lane's 2x2 quad of cells sits at
// (2 cx, 2 cy), cx and cy the even and odd bits of the lane; with rows 18
// floats apart the 64 lanes' reads of one corner fall 3-4 deep into one
// bank, with 20 two deep.  Load k of lane l fetches the sample at padded
// position p = l + 64 k, (p / 20, p % 20), so the staging writes stay
// contiguous; the two padding columns of a row read a clamped sample that
// nothing uses.
constexpr int kBrickStride = 20;
static_assert(kBrick * kBrickStride <= kPlaneStride, "a padded window fits its slot");
__device__ __forceinline__ void brick_lane_pos(int lane, int k, int& kx, int& ky)
{
    const int p = min(lane + 64 * k, kBrick * kBrickStride - 1);
    ky = p / kBrickStride;
    kx = min(p - ky * kBrickStride, kBrick - 1);
}

struct EmitWaveShared {
    float plane[3][kPlaneStride];  // planes z, z+1, z+2 of the item
    uint16_t cexc[256];            // per cell (Morton order): first vertex in the block
    uint8_t code[256];             // per cell: HPMC code
    uint8_t vcell[kEmitChunk + 64];  // per vertex of the item's chunk: its cell (+ a spare byte per lane)
    float ostage[64 * 6];          // one batch of 64 output vertices, for contiguous stores
};

struct EmitShared {
    uint64_t cases[256];
    uint8_t lut[256];
    EmitWaveShared w[kEmitWaves];
};

// Block origin (x0, y0, z) of a level-4 key.
__device__ __forceinline__ void block_origin(const Geom& g, uint32_t key, int& x0, int& y0, int& z)
{
    const int X0 = (int)morton_compact(key) << g.blog2;
    const int Y0 = (int)morton_compact(key >> 1) << g.blog2;
    const int tx = X0 >> g.tsx2_log2, ty = Y0 >> g.tsy2_log2;
    z = (ty << g.tiles_x_log2) + tx;
    x0 = X0 & (g.tsx2 - 1);
    y0 = Y0 & (g.tsy2 - 1);
}

// Fields with unclamped window access (kWindows).
template <class F, class = void>
struct has_windows {
    static constexpr bool value = false;
};
template <class F>
struct has_windows<F, decltype((void)F::kWindows, void())> {
    static constexpr bool value = F::kWindows;
};

// Per lane, the window offsets of its kPlaneLoads samples of an 18x18
// window (fields with window access; a volume's own, as they scale with
// its row length).
template <class F>
__device__ __forceinline__ void emit_window_lanes(const F& field, int lane, uint32_t (&wl)[kPlaneLoads])
{
#pragma unroll
    for (int k = 0; k < kPlaneLoads; ++k) {
        if constexpr (has_windows<F>::value) {
            int kx, ky;
            brick_lane_pos(lane, k, kx, ky);
            wl[k] = field.window_lane(kx, ky);
        } else {
            wl[k] = 0u;
        }
    }
}

// Loads planes zfirst .. zfirst + n - 1 (n = 0..3, wave-uniform) of the
// 18x18 window at (x0, y0) into r[0 .. n-1].  A window inside the lattice's
// planes (all but the last block row and column) is read through the lane
// offsets wl and a scalar corner offset; others clamp each sample.
template <class F>
__device__ __forceinline__ void emit_fetch(const F& field, int x0, int y0, int zfirst, int n, int lane,
                                           const uint32_t (&wl)[kPlaneLoads], float (&r)[3][kPlaneLoads])
{
    uint32_t idx[kPlaneLoads];
    uint32_t base = 0u;
    bool inside = false;
    if constexpr (has_windows<F>::value) inside = field.window_inside(x0, y0, kBrick);
    if (inside) {
        if constexpr (has_windows<F>::value) base = field.window_base(x0, y0);
#pragma unroll
        for (int k = 0; k < kPlaneLoads; ++k) idx[k] = wl[k];
    } else {
#pragma unroll
        for (int k = 0; k < kPlaneLoads; ++k) {
            int kx, ky;
            brick_lane_pos(lane, k, kx, ky);
            idx[k] = field.lane_index(x0 + kx, y0 + ky);
        }
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        if (q < n) {
            const typename F::Plane pl = field.plane(zfirst + q);
#pragma unroll
            for (int k = 0; k < kPlaneLoads; ++k) {
                if constexpr (has_windows<F>::value) r[q][k] = field.window_sample(pl, idx[k], base);
                else r[q][k] = field.sample(pl, idx[k]);
            }
        }
    }
}

// Stores planes zfirst .. zfirst + n - 1 from r into their ring slots
// (plane z in slot z mod 3).
__device__ __forceinline__ void emit_store(EmitWaveShared& ws, int zfirst, int n, int lane,
                                           const float (&r)[3][kPlaneLoads])
{
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        if (q < n) {
            float* dst = ws.plane[(zfirst + q) % 3];
#pragma unroll
            for (int k = 0; k < kPlaneLoads; ++k) dst[lane + 64 * k] = r[q][k];
        }
    }
}

// BATCH: items of several volumes (the volume of each item is looked up);
// otherwise every item belongs to vols[0].  mk(d) makes the field of
// volume descriptor d.  nitems: the item total of the scan (the kernel loads
// it before its LDS set-up, so that the two round trips overlap).
constexpr uint32_t kEmitSkew = 13;   // share skew of emission workgroups by dispatch order, in 64ths

template <int MODE, bool BATCH, class MK, class CB = NoVertexCallback>
__device__ __forceinline__ void emit_body(const VolDesc* __restrict__ vols, const MK& mk, float T,
                                          const uint2* __restrict__ items, const uint32_t* __restrict__ l4_off,
                                          const uint32_t nitems,
                                          float* __restrict__ out, CellEdge* __restrict__ out_ce,
                                          uint32_t capacity, const float4* __restrict__ bnorm, EmitShared& sh,
                                          uint64_t case_word, const CB& cb = CB{})
{
    constexpr int NP = emit_planes<MODE>();
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const uint32_t nwg = gridDim.x;                 // a multiple of 8
    // workgroups are dealt to the 8 XCDs round robin; give each XCD a
    // contiguous run of logical workgroup indices, and each wave a
    // contiguous run of items: consecutive items are mostly the next slice
    // of the same block column (or the next chunk of the same block), whose
    // planes are already staged
    // Items of an XCD: one eighth of the run, in order.  Within the XCD the
    // workgroups are dispatched in order q = 0, 1, ... and the SIMDs issue
    // from their oldest waves first, so the workgroups dispatched first
    // progress fastest; workgroup q gets a share of the XCD's items falling
    // linearly from 1 + a to 1 - a (a = kEmitSkew / 64) so that they finish
    // together.  Its four waves split the share evenly.
    const uint32_t nx = (nwg & 7u) ? 1u : 8u;
    const uint32_t x = (nwg & 7u) ? 0u : (blockIdx.x & 7u);
    const uint32_t q = (nwg & 7u) ? blockIdx.x : (blockIdx.x >> 3);
    const uint32_t nq = nwg / nx;
    uint32_t it, iend;
    {
        const uint64_t lo = (uint64_t)x * nitems / nx, hi = (uint64_t)(x + 1) * nitems / nx;
        // cumulative share of workgroups 0 .. k-1, times 64 (nq - 1) nq
        auto cum = [&](uint64_t k) -> uint64_t {
            return nq > 1 ? k * (64u + kEmitSkew) * (nq - 1) - (uint64_t)kEmitSkew * k * (k - (k > 0)) : 64u * k;
        };
        const uint64_t den = nq > 1 ? 64ull * (nq - 1) * nq : 64ull;
        const uint64_t s0 = lo + (hi - lo) * cum(q) / den, s1 = lo + (hi - lo) * cum(q + 1) / den;
        it = __builtin_amdgcn_readfirstlane((uint32_t)(s0 + (s1 - s0) * wv / kEmitWaves));
        iend = __builtin_amdgcn_readfirstlane((uint32_t)(s0 + (s1 - s0) * (wv + 1) / kEmitWaves));
    }
    const bool has_items = it < iend;

    EmitWaveShared& ws = sh.w[wv];
    const uint64_t* cases = sh.cases;
    const uint8_t* lut = sh.lut;
    // this lane's 2x2 quad of cells: Morton indices 4*lane .. 4*lane+3
    const int qx = (int)morton_compact((uint32_t)lane) * 2;
    const int qy = (int)morton_compact((uint32_t)lane >> 1) * 2;

    // The first item's entry, descriptor, offset and planes are requested
    // before this thread's case-table entry (case_word, loaded by the
    // caller) is stored to LDS and the workgroup barrier publishes the
    // tables, so the round trips overlap.
    // (Loaded whether or not the wave has items, so that the count of loads
    // in flight is fixed and the wait for case_word does not cover them; a
    // wave without items reads item slot 0 and uses the key-0 block of
    // volume 0, all in bounds.)
    const uint32_t i0 = has_items ? it : 0u;
    uint2 cur = load_uniform(items + i0);
    uint2 nxt = load_uniform(items + (has_items ? min(it + 1, iend - 1) : 0u));
    if (!has_items) cur = make_uint2(0u, 0u);
    VolDesc dn = load_uniform(vols + (BATCH ? item_volume(cur.y) : 0u));
    uint32_t obase = load_uniform(l4_off + dn.l4_base + cur.x);
    float r[3][kPlaneLoads];
    int x0, y0, z;
    block_origin(dn.g, cur.x, x0, y0, z);
    int nnew = NP;                      // planes the item needs that are not staged yet
    uint32_t wl[kPlaneLoads];           // window lane offsets of the volume fetched last
    emit_window_lanes(mk(dn), lane, wl);
    emit_fetch(mk(dn), x0, y0, z, NP, lane, wl, r);
    sh.cases[threadIdx.x] = case_word;
    sh.lut[threadIdx.x] = (uint8_t)(case_word & 0xfu);
    __syncthreads();
    if (!has_items) return;
    // The first item's planes have landed before the loop starts, so that at
    // the loop's head only a previous iteration's loads are pending, always
    // followed by that iteration's vertex stores: the head's wait then
    // leaves those stores in flight (the vector-memory counter covers loads
    // and stores alike and completes in order).
    __builtin_amdgcn_s_waitcnt(0x0f70);   // vmcnt(0)

    // Per lane, the vertex counts and first vertices of its four cells and
    // the block's vertex total, kept from item to item: a further chunk of
    // the same block slice (the next item, as the scan lays them out) finds
    // its cells' codes and first vertices in the wave's LDS already and
    // skips classification.
    uint32_t cc[4] = {0u, 0u, 0u, 0u}, ex[4] = {0u, 0u, 0u, 0u}, blk_total = 0u;
    bool reuse = false;
    for (;;) {
        // The SIMD issues from the oldest ready wave first, so the waves
        // dispatched last on a SIMD fall behind and end up running their
        // item chains alone; ranking waves by the items they have left keeps
        // their progress even and the SIMDs busy to the end.
        {
            const uint32_t rem = iend - it;   // items left, this one included
            if (rem >= 4u) __builtin_amdgcn_s_setprio(3);
            else if (rem == 3u) __builtin_amdgcn_s_setprio(2);
            else if (rem == 2u) __builtin_amdgcn_s_setprio(1);
            else __builtin_amdgcn_s_setprio(0);
        }
        const uint32_t vlo = item_chunk_first(cur.y), cnt = item_block_count(cur.y);
        const uint32_t base = __builtin_amdgcn_readfirstlane(obase);
        const int cz = z, cx0 = x0, cy0 = y0;
        const uint32_t cvol = BATCH ? item_volume(cur.y) : 0u;
        const VolDesc dc = dn;          // this item's volume
        const Geom& g = dc.g;
        wave_lds_sync();
        emit_store(ws, cz + NP - nnew, nnew, lane, r);
        wave_lds_sync();

        // next item: entry, output offset and missing planes in flight meanwhile
        const bool more = it + 1 < iend;
        uint2 after = nxt;
        bool reuse_next = false;
        if (more) {
            reuse_next = nxt.x == cur.x && (!BATCH || item_volume(nxt.y) == cvol);
            // (the next descriptor is a scalar load, only when the volume
            // changes; the offset and the entry after it go through the
            // vector memory path, see prefetch_u32)
            if constexpr (BATCH)
                if (item_volume(nxt.y) != cvol) {
                    dn = load_uniform(vols + item_volume(nxt.y));
                    emit_window_lanes(mk(dn), lane, wl);
                }
            obase = prefetch_u32(l4_off + dn.l4_base + nxt.x);
            block_origin(dn.g, nxt.x, x0, y0, z);
            const bool column = x0 == cx0 && y0 == cy0 && (!BATCH || item_volume(nxt.y) == cvol);
            nnew = (column && z == cz) ? 0 : ((column && z == cz + 1) ? 1 : NP);
            emit_fetch(mk(dn), x0, y0, z + NP - nnew, nnew, lane, wl, r);
            after = prefetch_u2(items + min(it + 2, iend - 1));
        }

        // plane ring: plane cz + a sits in slot (cz + a) mod 3
        const int s0 = cz % 3, s1 = (cz + 1) % 3, s2 = (cz + 2) % 3;
        const float* P0 = ws.plane[s0];
        const float* P1 = ws.plane[s1];
        if (!reuse) {
            // cells of the block (blocks narrower than 16 take the corner
            // of the 16x16 quad layout; quads are whole inside or outside)
            const bool qin = qx < (1 << g.blog2) && qy < (1 << g.blog2);
            const bool qvx0 = qin && cx0 + qx < g.cx, qvx1 = qin && cx0 + qx + 1 < g.cx;
            const bool qvy0 = cy0 + qy < g.cy, qvy1 = cy0 + qy + 1 < g.cy;

            // classify the quad: corners (qx..qx+2, qy..qy+2) of planes z, z+1
            uint32_t bits0 = 0, bits1 = 0;   // 3x3 inside bits per plane, bit = dx + 3*dy
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                for (int dx = 0; dx < 3; ++dx) {
                    const int i = (qy + dy) * kBrickStride + qx + dx;
                    bits0 |= (P0[i] < T ? 1u : 0u) << (dx + 3 * dy);
                    bits1 |= (P1[i] < T ? 1u : 0u) << (dx + 3 * dy);
                }
            uint32_t codes4 = 0, cnt4 = 0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int cxq = q & 1, cyq = q >> 1;
                const int sh0 = cxq + 3 * cyq;
                const uint32_t lo2 = ((bits0 >> sh0) & 3u) | (((bits0 >> (sh0 + 3)) & 3u) << 2);
                const uint32_t hi2 = ((bits1 >> sh0) & 3u) | (((bits1 >> (sh0 + 3)) & 3u) << 2);
                const uint32_t code = lo2 | (hi2 << 4);
                const bool valid = (cxq ? qvx1 : qvx0) && (cyq ? qvy1 : qvy0);
                cc[q] = (uint32_t)lut[code] & (valid ? 0xffu : 0u);
                codes4 |= code << (8 * q);
                cnt4 += cc[q];
            }
            // first vertex of every cell: an inclusive scan of the quad counts
            // (cells in Morton order = the reference's child order,
            // shadergen.cpp:448-495, so Morton order is its vertex order)
            const uint32_t inc = wave_scan_u32<false>(cnt4);
            ex[0] = inc - cnt4;
            ex[1] = ex[0] + cc[0];
            ex[2] = ex[1] + cc[1];
            ex[3] = ex[2] + cc[2];
            *(uint32_t*)&ws.code[4 * lane] = codes4;
            *(uint2*)&ws.cexc[4 * lane] = make_uint2(ex[0] | (ex[1] << 16), ex[2] | (ex[3] << 16));
            blk_total = (uint32_t)__builtin_amdgcn_readlane(inc, 63);
        }
        ((uint32_t*)ws.vcell)[lane] = 0u;
        const uint32_t nv = min(blk_total, cnt);   // whole-block total
        const uint32_t vhi = min(vlo + (uint32_t)kEmitChunk, nv);
        wave_lds_sync();
        // vertex -> cell for this chunk: each non-empty cell marks its first
        // vertex (the cell holding vertex vlo marks position 0), then a max
        // scan over the chunk carries every mark forward
        {
            // marks that do not apply go to this lane's spare byte
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int pos = (int)ex[q] - (int)vlo;
                const bool mark = cc[q] && pos < kEmitChunk && pos + (int)cc[q] > 0;
                ws.vcell[mark ? max(pos, 0) : kEmitChunk + lane] = (uint8_t)(4 * lane + q);
            }
        }
        wave_lds_sync();
        {
            const uint32_t w4 = ((const uint32_t*)ws.vcell)[lane];
            const uint32_t m0 = w4 & 0xffu;
            const uint32_t m1 = max(m0, (w4 >> 8) & 0xffu);
            const uint32_t m2 = max(m1, (w4 >> 16) & 0xffu);
            const uint32_t m3 = max(m2, w4 >> 24);
            const uint32_t before = wave_shr1_u32(wave_scan_u32<true>(m3));
            ((uint32_t*)ws.vcell)[lane] = max(before, m0) | (max(before, m1) << 8) | (max(before, m2) << 16) |
                                          (max(before, m3) << 24);
        }
        wave_lds_sync();

        // Output vertex v of the item (vlo <= v < vhi): its cell c (Morton
        // index in the block) and edge e, the edge's first corner (ax, ay,
        // az) in brick coordinates and its axis.
        struct Site {
            int c, axis, ccx, ccy, ax, ay, az;
            uint32_t code, slot, e;
        };
        auto site_of = [&](uint32_t v) {
            Site s;
            s.c = ws.vcell[v - vlo];
            s.code = ws.code[s.c];
            // (clamped: a batch past the chunk's end, which stores nothing,
            // still indexes the binary-normal table in bounds)
            s.slot = min(v - ws.cexc[s.c], 15u);
            s.e = (uint32_t)(cases[s.code] >> (4 + 4 * s.slot)) & 0xfu;
            // edge e = 4 axis + s0 + 2 s1 -> corner offset (dx, dy, dz) of
            // its first end and the axis, bits 0..4 of a 5-bit entry
            constexpr uint64_t kEdgeOffsets =
                (0x00ull << 0) | (0x02ull << 5) | (0x04ull << 10) | (0x06ull << 15) |     // axis x: (0,s0,s1)
                (0x08ull << 20) | (0x09ull << 25) | (0x0cull << 30) | (0x0dull << 35) |   // axis y: (s0,0,s1)
                (0x10ull << 40) | (0x11ull << 45) | (0x12ull << 50) | (0x13ull << 55);    // axis z: (s0,s1,0)
            const uint32_t eo = (uint32_t)(kEdgeOffsets >> (5 * s.e)) & 31u;
            s.axis = (int)(eo >> 3);
            s.ccx = (int)morton_compact((uint32_t)s.c);
            s.ccy = (int)morton_compact((uint32_t)s.c >> 1);
            s.ax = s.ccx + (int)(eo & 1u);
            s.ay = s.ccy + (int)((eo >> 1) & 1u);
            s.az = (int)((eo >> 2) & 1u);
            return s;
        };
        // The staged samples of continuous emission: the edge's ends a and b
        // and their forward neighbours na = f(a + e_q), nb = f(b + e_q).
        // Sample (x, y, plane) sits at x + 20 y + 384 slot of the staged
        // brick; a and b lie in planes 0 or 1, their +z neighbours in planes
        // 1 or 2 (an edge along z starts in plane 0).
        struct EdgeSamples {
            float va, vb, nax, nay, naz, nbx, nby, nbz;
        };
        auto samples_of = [&](const Site& s) {
            const int bx_ = s.ax + (s.axis == 0), by_ = s.ay + (s.axis == 1), bz_ = s.az + (s.axis == 2);
            const float* Pw = ws.plane[0];
            const int o0 = s0 * kPlaneStride, o1 = s1 * kPlaneStride, o2 = s2 * kPlaneStride;
            const int ia = (int)__umul24((uint32_t)s.ay, (uint32_t)kBrickStride) + s.ax;
            const int ib = (int)__umul24((uint32_t)by_, (uint32_t)kBrickStride) + bx_;
            const int pa = s.az ? o1 : o0, pa1 = s.az ? o2 : o1, pb = bz_ ? o1 : o0, pb1 = bz_ ? o2 : o1;
            EdgeSamples q;
            q.va = Pw[ia + pa];
            q.vb = Pw[ib + pb];
            q.nax = Pw[ia + pa + 1];
            q.nay = Pw[ia + pa + kBrickStride];
            q.naz = Pw[ia + pa1];
            q.nbx = Pw[ib + pb + 1];
            q.nby = Pw[ib + pb + kBrickStride];
            q.nbz = Pw[ib + pb1];
            return q;
        };
        // Normal (o[0..2]) and position (o[3..5]) of a vertex, the
        // reference's emission formulas (shadergen.cpp:505-563).
        auto eval = [&](const Site& s, float (&o)[6]) {
            float nx, ny, nz, tt;
            if constexpr (MODE == kEmitBinary) {
                // shadergen.cpp:507-510, 521-522: midpoint, tabulated normal
                const float4 bn = bnorm[s.code * 16u + s.slot];
                tt = 0.5f;
                nx = bn.x;
                ny = bn.y;
                nz = bn.z;
            } else if constexpr (MODE == kEmitGradient) {
                // shadergen.cpp:539-549: values and gradients at both ends
                const auto field = mk(dc);
                const int bx_ = s.ax + (s.axis == 0), by_ = s.ay + (s.axis == 1), bz_ = s.az + (s.axis == 2);
                const float4 fa = field.grad(cx0 + s.ax, cy0 + s.ay, cz + s.az);
                const float4 fb = field.grad(cx0 + bx_, cy0 + by_, cz + bz_);
                tt = (fa.w - T) * __builtin_amdgcn_rcpf(fa.w - fb.w);
                const float u = 1.0f - tt;
                // explicit fma: the same bits however the code is compiled
                // (ahead of time, or by hipRTC with contraction off)
                nx = T - __builtin_fmaf(fa.x, u, fb.x * tt);
                ny = T - __builtin_fmaf(fa.y, u, fb.y * tt);
                nz = T - __builtin_fmaf(fa.z, u, fb.z * tt);
            } else {
                const EdgeSamples q = samples_of(s);
                // va and vb lie on opposite sides of T, so va - vb != 0; the
                // hardware reciprocal (1 ulp) keeps t within a few ulp
                tt = (q.va - T) * __builtin_amdgcn_rcpf(q.va - q.vb);
                const float u = 1.0f - tt;
                nx = T - __builtin_fmaf(q.nax, u, q.nbx * tt);
                ny = T - __builtin_fmaf(q.nay, u, q.nby * tt);
                nz = T - __builtin_fmaf(q.naz, u, q.nbz * tt);
            }
            const float gx = (float)(cx0 + s.ax), gy = (float)(cy0 + s.ay), gz = (float)(cz + s.az);
            o[0] = nx * g.scx;
            o[1] = ny * g.scy;
            o[2] = nz * g.scz;
            o[3] = (gx + (s.axis == 0 ? tt : 0.0f)) * g.scx;
            o[4] = (gy + (s.axis == 1 ? tt : 0.0f)) * g.scy;
            o[5] = (gz + (s.axis == 2 ? tt : 0.0f)) * g.scz;
        };
        // The batch of 64 vertices v0 .. v0 + 63 (clipped at vhi) goes out
        // as three contiguous 512-byte stores (lane L writes floats
        // 2(64k + L), +1) after a transpose through LDS, non-temporal: the
        // output streams to memory instead of staying dirty in L2, where its
        // write-back would compete with the next build's classification.
        typedef float v2f __attribute__((ext_vector_type(2)));
        typedef uint32_t v2u __attribute__((ext_vector_type(2)));
        auto store_batch = [&](uint32_t v0, const Site& s, const float (&o)[6]) {
            {
                v2f* st = (v2f*)(ws.ostage + 6 * lane);
                st[0] = (v2f){o[0], o[1]};
                st[1] = (v2f){o[2], o[3]};
                st[2] = (v2f){o[4], o[5]};
            }
            // (wave-uniform values, in scalar registers for the resource)
            const uint32_t start = __builtin_amdgcn_readfirstlane(base + v0);   // first output vertex of the batch
            const uint32_t nval = vhi > v0 ? min(vhi - v0, 64u) : 0u;
            const uint32_t nout = __builtin_amdgcn_readfirstlane(capacity > start ? min(nval, capacity - start) : 0u);
            // the batch's nout vertices as a buffer range: pairs past it are
            // dropped by the range check, so every store issues (no branch)
            const __amdgpu_buffer_rsrc_t ob =
                __builtin_amdgcn_make_buffer_rsrc((void*)(out + (size_t)start * 6), (short)0, (int)(nout * 24u),
                                                  0x00020000);
            wave_lds_sync();
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const uint32_t pr = 64u * k + lane;   // float pair: vertex pr / 3
                const v2u val = ((const v2u*)ws.ostage)[pr];
                __builtin_amdgcn_raw_buffer_store_b64(val, ob, pr * 8u, 0, kStreamLoad);
            }
            wave_lds_sync();
            if (out_ce && v0 + (uint32_t)lane < vhi && start + lane < capacity) {
                const uint32_t cell = (uint32_t)(cx0 + s.ccx) +
                                      (uint32_t)g.cx * ((uint32_t)(cy0 + s.ccy) + (uint32_t)g.cy * (uint32_t)cz);
                __builtin_nontemporal_store((v2u){cell, s.e}, (v2u*)(out_ce + start + lane));
            }
        };

        // batches of 64 consecutive vertices, one per lane (lanes past the
        // chunk's end repeat its last vertex and store nothing)
        if constexpr (CB::kActive) {
            // a user-fused traversal: hand each vertex over, write nothing
            for (uint32_t v0 = vlo; v0 < vhi; v0 += 64) {
                const uint32_t v = min(v0 + (uint32_t)lane, vhi - 1u);
                const Site st = site_of(v);
                float o[6];
                eval(st, o);
                if (v0 + (uint32_t)lane < vhi) {
                    if constexpr (wants_edge<CB>::value) {
                        // edge ends as texel-centre coordinates (i + 0.5) / size
                        const int ax = cx0 + st.ax, ay = cy0 + st.ay, az = cz + st.az;
                        const int bx = ax + (st.axis == 0), by = ay + (st.axis == 1), bz = az + (st.axis == 2);
                        const float3 a = make_float3(((float)ax + 0.5f) * g.ipx, ((float)ay + 0.5f) * g.ipy,
                                                     ((float)az + 0.5f) * g.ipz);
                        const float3 b = make_float3(((float)bx + 0.5f) * g.ipx, ((float)by + 0.5f) * g.ipy,
                                                     ((float)bz + 0.5f) * g.ipz);
                        cb(base + v, a, b, make_float3(o[0], o[1], o[2]), make_float3(o[3], o[4], o[5]));
                    } else {
                        cb(base + v, make_float3(o[0], o[1], o[2]), make_float3(o[3], o[4], o[5]));
                    }
                }
            }
        } else {
            // at least one batch (an item holds a vertex), so that every
            // path from the next item's plane loads to their use issues
            // stores, and the wait for those loads leaves the stores in flight
            uint32_t v0 = vlo;
            do {
                const Site s = site_of(min(v0 + (uint32_t)lane, max(vhi, vlo + 1u) - 1u));
                float o[6];
                eval(s, o);
                store_batch(v0, s, o);
                v0 += 64;
            } while (v0 < vhi);
        }
        if (!more) break;
        ++it;
        reuse = reuse_next;
        cur = nxt;
        nxt = uniform_u2(after);
    }
}

// ---------------------------------------------------------------------------
// Adjoint of continuous emission: dL/dfield and dL/dT of the last build's
// vertices from dL/dvertex (6 floats per vertex, normal then position),
// deterministic, with no atomics and no sorting.  The vertices on one
// lattice edge (repeated by the triangles of the up to four cells around
// it) share position and normal, so the adjoint is formed per edge from the
// sum of their dL/dvertex.  For the edge from a to b = a + e_axis,
// t = (va - T) / (va - vb),
//   p = sc (a + t e_axis),  n = sc (T - (na (1 - t) + nb t)),
// with na_q = f(a + e_q), nb_q = f(b + e_q) (na_axis is vb), so with
// D = va - vb: dt/dva = (T - vb) / D^2, dt/dvb = (va - T) / D^2,
// dt/dT = -1 / D; samples past the lattice read (and receive) the clamped
// sample, as the forward reads them.
//
// Three passes:
//  cells    one wave per non-empty block (the first work item of each):
//           the cells' codes and each non-empty cell's first output vertex
//           (the block offset plus the Morton-order prefix of the counts,
//           as emission numbers them), into arrays indexed like the lattice
//           samples (codes cleared beforehand), and a flag for every gather
//           tile the block's edges can reach;
//  records  one wave per non-empty block: the sign-changing edges its cells
//           own, listed in LDS and spread over the lanes; per edge the
//           summed dL/dvertex of its vertices (cells in a fixed order,
//           slots in order), the seven sample contributions and the dL/dT
//           term, one record stored at the owner's first vertex + the
//           edge's rank among the owner's edges (every sign-changing edge
//           of a case carries a slot, so a cell owns no more edges than it
//           has vertices);
//  gather   one workgroup per tile of samples: every sample sums the
//           contributions that reach it in a fixed (axis, kind) order and
//           is written, zero where none does (unflagged tiles at once).
// Edge (a, axis) lies in the cells a - j e_p - k e_q (j, k in {0, 1}, p < q
// the other axes); its owner is the first of those inside the grid in the
// order (0,0), (1,0), (0,1), (1,1): a cell owns the three edges from its
// origin and, on the grid's upper faces, the edges there.
constexpr int kAdjointRecord = 7;           // contributions per edge record
constexpr int kAdjointRecordFloats = 8;     // floats a record takes (32 bytes)
constexpr uint32_t kNoEdge = 0xffffffffu;
constexpr int kAdjointWaves = 4;            // waves per workgroup of the cells and records passes
constexpr int kEdgeListMax = 1024;          // listed edges of a block (more: each lane takes its own)

// Gather tiles of 16 x 8 x 8 samples; the edges reaching a tile start at
// most two samples below it, their owners at most three cells below.
constexpr int kGatherX = 16, kGatherY = 8, kGatherZ = 8;
constexpr int kGatherOX = kGatherX + 2, kGatherOY = kGatherY + 2, kGatherOZ = kGatherZ + 2;
constexpr int kGatherOrigins = kGatherOX * kGatherOY * kGatherOZ;
constexpr int kGatherCX = kGatherX + 3, kGatherCY = kGatherY + 3, kGatherCZ = kGatherZ + 3;
constexpr int kGatherCells = kGatherCX * kGatherCY * kGatherCZ;
constexpr int kGatherThreads = 256;
constexpr int kGatherCellLoads = (kGatherCells + kGatherThreads - 1) / kGatherThreads;
constexpr int kGatherOriginLoads = (kGatherOrigins + kGatherThreads - 1) / kGatherThreads;
static_assert(kGatherX * kGatherY * kGatherZ == 4 * kGatherThreads, "four samples a thread");

// Sign-changing edges of a cell code (bit e for edge e = 4 axis + s0 + 2 s1;
// corner bit x + 2y + 4z).
__device__ __forceinline__ uint32_t code_edges(uint32_t code)
{
    const uint32_t dx = code ^ (code >> 1), dy = code ^ (code >> 2), dz = code ^ (code >> 4);
    return (dx & 1u) | ((dx >> 1) & 2u) | ((dx >> 2) & 4u) | ((dx >> 3) & 8u) | ((dy & 3u) << 4) |
           (((dy >> 4) & 3u) << 6) | ((dz & 0xfu) << 8);
}

// Edges that cell (x, y, z) owns: all but those on its upper faces, unless
// the face is the grid's.
__device__ __forceinline__ uint32_t owned_edges(const Geom& g, int x, int y, int z)
{
    uint32_t m = 0xfffu;
    if (x < g.cx - 1) m &= ~0xaa0u;   // x + 1: y and z edges with s0 = 1
    if (y < g.cy - 1) m &= ~0xc0au;   // y + 1: x edges with s0 = 1, z edges with s1 = 1
    if (z < g.cz - 1) m &= ~0x0ccu;   // z + 1: x and y edges with s1 = 1
    return m;
}

// Per case and edge, the slots of the case on the edge (smask[code][e],
// LDS), from the 64-bit case words (count, then 4-bit edges); one thread
// per case.
__device__ __forceinline__ void fill_slot_masks(const uint64_t* __restrict__ cases, uint16_t (*smask)[12])
{
    const uint64_t w = cases[threadIdx.x];
    uint32_t m[12] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
    const int cnt = (int)(w & 0xfu);
    for (int sl = 0; sl < cnt; ++sl) {
        const uint32_t e = (uint32_t)(w >> (4 + 4 * sl)) & 0xfu;
#pragma unroll
        for (int k = 0; k < 12; ++k) m[k] |= (e == (uint32_t)k) ? (1u << sl) : 0u;
    }
#pragma unroll
    for (int k = 0; k < 12; ++k) smask[threadIdx.x][k] = (uint16_t)m[k];
}

// A buffer resource over `bytes` bytes at base: raw loads through it take
// 32-bit byte offsets and read 0 past the end.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_of(const void* base, uint64_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, (int)(uint32_t)min(bytes, (uint64_t)0xffffffffu),
                                             0x00020000);
}
__device__ __forceinline__ uint32_t buffer_load_u32(__amdgpu_buffer_rsrc_t r, uint32_t byte_offset)
{
    return (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(r, byte_offset, 0, 0);
}
__device__ __forceinline__ uint32_t buffer_load_u8(__amdgpu_buffer_rsrc_t r, uint32_t byte_offset)
{
    return (uint32_t)__builtin_amdgcn_raw_buffer_load_b8(r, byte_offset, 0, 0);
}

// A volume's samples as a global-memory pointer: loads through it are
// global loads (counted only by vmcnt), not flat ones, whose waits would
// also cover LDS traffic.
typedef const __attribute__((address_space(1))) float* GlobalFloats;
__device__ __forceinline__ GlobalFloats global_field(const VolDesc& d) { return (GlobalFloats)d.field; }

// Linear index of an in-range sample (or of the cell with that origin).
__device__ __forceinline__ uint32_t lattice_index(const Geom& g, int x, int y, int z)
{
    return ((uint32_t)z * (uint32_t)g.sy + (uint32_t)y) * (uint32_t)g.sx + (uint32_t)x;
}

// Sum over the wave in a fixed order (butterfly).
__device__ __forceinline__ float wave_sum_f(float v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// cells pass.  Lane l takes the 2x2 quad of cells with Morton indices
// 4l .. 4l+3 of the block; lut = per-code vertex counts (LDS); tile0[v] =
// first gather tile of volume v; flags: one byte per gather tile.
__device__ __forceinline__ void adjoint_cells_body(const VolDesc* __restrict__ vols, float T,
                                                   const uint2* __restrict__ items,
                                                   const uint32_t* __restrict__ l4_off,
                                                   const Totals* __restrict__ totals, const uint8_t* lut,
                                                   const uint32_t* __restrict__ tile0, uint8_t* __restrict__ codes,
                                                   uint32_t* __restrict__ cell_off, uint8_t* __restrict__ flags)
{
    const uint32_t nitems = load_uniform(&totals->items);
    const int lane = threadIdx.x & 63;
    const uint32_t w0 = blockIdx.x * kAdjointWaves + (uint32_t)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * kAdjointWaves;
    const int qx = (int)morton_compact((uint32_t)lane) * 2, qy = (int)morton_compact((uint32_t)lane >> 1) * 2;
    for (uint32_t it = w0; it < nitems; it += nw) {
        const uint2 item = load_uniform(items + it);
        if (item.y & 0xfu) continue;   // a further chunk of a block
        const uint32_t vol = item_volume(item.y);
        const VolDesc d = load_uniform(vols + vol);
        const Geom& g = d.g;
        const uint32_t obase = load_uniform(l4_off + d.l4_base + item.x);
        const uint32_t t0 = load_uniform(tile0 + vol);
        int x0, y0, z;
        block_origin(g, item.x, x0, y0, z);
        const int side = 1 << g.blog2;
        // the quad's 3x3 corners on planes z and z + 1 (clamped indices)
        const VolumeField f{(const float*)d.field, g.sx, g.sy, g.sz};
        const GlobalFloats fv = global_field(d);
        float v[2][3][3];
#pragma unroll
        for (int k = 0; k < 2; ++k)
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                for (int dx = 0; dx < 3; ++dx) v[k][dy][dx] = fv[f.index(x0 + qx + dx, y0 + qy + dy, z + k)];
        uint32_t code[4], cnt[4], cnt4 = 0u;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int ox = q & 1, oy = q >> 1;
            const bool valid = qx < side && qy < side && x0 + qx + ox < g.cx && y0 + qy + oy < g.cy;
            uint32_t c = 0u;
#pragma unroll
            for (int k = 0; k < 2; ++k)
#pragma unroll
                for (int cy = 0; cy < 2; ++cy)
#pragma unroll
                    for (int cx = 0; cx < 2; ++cx)
                        c |= (v[k][oy + cy][ox + cx] < T ? 1u : 0u) << (cx + 2 * cy + 4 * k);
            code[q] = c;
            cnt[q] = valid ? (uint32_t)lut[c] : 0u;
            cnt4 += cnt[q];
        }
        uint32_t ex = wave_scan_u32<false>(cnt4) - cnt4;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (cnt[q]) {
                const uint32_t s = d.lattice_base + lattice_index(g, x0 + qx + (q & 1), y0 + qy + (q >> 1), z);
                codes[s] = (uint8_t)code[q];
                cell_off[s] = obase + ex;
            }
            ex += cnt[q];
        }
        // gather tiles the block's edges reach: samples x0 .. x0 + side + 2
        // (and y likewise), z .. z + 3
        const int tx0 = x0 / kGatherX, tx1 = min(x0 + side + 2, g.sx - 1) / kGatherX;
        const int ty0 = y0 / kGatherY, ty1 = min(y0 + side + 2, g.sy - 1) / kGatherY;
        const int tz0 = z / kGatherZ, tz1 = min(z + 3, g.sz - 1) / kGatherZ;
        const int nx = tx1 - tx0 + 1, ny = ty1 - ty0 + 1, nz = tz1 - tz0 + 1;
        if (lane < nx * ny * nz) {
            const int ntx = (g.sx + kGatherX - 1) / kGatherX, nty = (g.sy + kGatherY - 1) / kGatherY;
            const int tx = tx0 + lane % nx, ty = ty0 + (lane / nx) % ny, tz = tz0 + lane / (nx * ny);
            flags[t0 + ((uint32_t)tz * nty + ty) * ntx + tx] = 1;
        }
    }
}

// The record of edge e of cell (x, y, z) (an owned sign-changing edge);
// returns its dL/dT term and stores the record (contribution k at float k
// of 32 bytes) at rec + 32 ri unless ri is kNoEdge: one sector per edge,
// which the gather's up to seven reads of the edge share.  vcodes
// and voffs: buffers over the volume's codes and first vertices; smask: the
// slot masks of fill_slot_masks.
// cases: the 256 case words (LDS).  Vertices at or past nvert (only if the
// field changed since the build) are skipped, so every access stays inside
// the vertex arrays.
__device__ __forceinline__ float adjoint_edge_record(const VolDesc& d, float T, const uint16_t (*smask)[12],
                                                     __amdgpu_buffer_rsrc_t vcodes, __amdgpu_buffer_rsrc_t voffs,
                                                     const float* __restrict__ gout, uint32_t nvert, int x, int y,
                                                     int z, int e, float* __restrict__ rec, uint32_t ri)
{
    const Geom& g = d.g;
    const int axis = e >> 2, s0 = e & 1, s1 = (e >> 1) & 1;
    // first end: axis x (0, s0, s1), axis y (s0, 0, s1), axis z (s0, s1, 0)
    const int Ax = x + (axis == 0 ? 0 : s0);
    const int Ay = y + (axis == 0 ? s0 : (axis == 1 ? 0 : s1));
    const int Az = z + (axis == 2 ? 0 : s1);
    const int Bx = Ax + (axis == 0), By = Ay + (axis == 1), Bz = Az + (axis == 2);
    // loads that need nothing else first: the edge's eight samples and the
    // code and first vertex of the cells around it, A - j e_p - k e_q in the
    // order (0,0), (1,0), (0,1), (1,1) (cells outside the grid read past the
    // buffers' ends: code 0, no slots)
    const VolumeField f{(const float*)d.field, g.sx, g.sy, g.sz};
    const GlobalFloats fv = global_field(d);
    const float va = fv[f.index(Ax, Ay, Az)], vb = fv[f.index(Bx, By, Bz)];
    const float nax = fv[f.index(Ax + 1, Ay, Az)], nay = fv[f.index(Ax, Ay + 1, Az)];
    const float naz = fv[f.index(Ax, Ay, Az + 1)];
    const float nbx = fv[f.index(Bx + 1, By, Bz)], nby = fv[f.index(Bx, By + 1, Bz)];
    const float nbz = fv[f.index(Bx, By, Bz + 1)];
    // the slots on the edge of the four cells as one 64-bit list, cell jk
    // in bits 16 jk .. 16 jk + 14, and their first vertices
    uint64_t slots = 0u;
    uint32_t off[4], cc[4];
#pragma unroll
    for (int jk = 0; jk < 4; ++jk) {
        const int j = jk & 1, k = jk >> 1;
        const int cx = Ax - (axis == 0 ? 0 : j);
        const int cy = Ay - (axis == 0 ? j : (axis == 2 ? k : 0));
        const int cz = Az - (axis == 2 ? 0 : k);
        const bool in = cx >= 0 && cy >= 0 && cz >= 0 && cx < g.cx && cy < g.cy && cz < g.cz;
        const uint32_t sc = in ? lattice_index(g, cx, cy, cz) : kNoEdge;
        cc[jk] = buffer_load_u8(vcodes, sc);
        off[jk] = buffer_load_u32(voffs, sc * 4u);
    }
#pragma unroll
    for (int jk = 0; jk < 4; ++jk)
        slots |= (uint64_t)smask[cc[jk]][4 * axis + (jk & 1) + 2 * (jk >> 1)] << (16 * jk);
    // dL/dvertex summed over the edge's vertices in list order (cells in
    // order, each cell's slots in order): the first kPreload loaded together,
    // any further ones after
    constexpr int kPreload = 4;
    auto vertex_of = [&](uint64_t list) {
        const uint32_t bit = (uint32_t)__builtin_ctzll(list | (1ull << 63));
        const uint32_t jk = bit >> 4;
        const uint32_t o = jk == 0 ? off[0] : (jk == 1 ? off[1] : (jk == 2 ? off[2] : off[3]));
        return o + (bit & 15u);
    };
    float gn[3] = {0.0f, 0.0f, 0.0f}, gp[3] = {0.0f, 0.0f, 0.0f};
    // (loaded unconditionally from a valid vertex and added as +0 when not
    // on the list: a sum that starts at +0 never becomes -0, so adding +0
    // changes no bits)
    float2 pre[kPreload][3];
    bool ok[kPreload];
    {
        uint64_t l = slots;
#pragma unroll
        for (int i = 0; i < kPreload; ++i) {
            const uint32_t vi = vertex_of(l);
            ok[i] = l != 0u && vi < nvert;
            const float2* gi = (const float2*)(gout + (size_t)(ok[i] ? vi : 0u) * 6);
            pre[i][0] = gi[0];
            pre[i][1] = gi[1];
            pre[i][2] = gi[2];
            l &= l - 1u;
        }
    }
#pragma unroll
    for (int i = 0; i < kPreload; ++i) {
        gn[0] += ok[i] ? pre[i][0].x : 0.0f;
        gn[1] += ok[i] ? pre[i][0].y : 0.0f;
        gn[2] += ok[i] ? pre[i][1].x : 0.0f;
        gp[0] += ok[i] ? pre[i][1].y : 0.0f;
        gp[1] += ok[i] ? pre[i][2].x : 0.0f;
        gp[2] += ok[i] ? pre[i][2].y : 0.0f;
        slots &= slots - 1u;
    }
    while (slots) {
        const uint32_t vi = vertex_of(slots);
        slots &= slots - 1u;
        if (vi >= nvert) continue;
        const float2* gi = (const float2*)(gout + (size_t)vi * 6);
        const float2 g01 = gi[0], g23 = gi[1], g45 = gi[2];
        gn[0] += g01.x;
        gn[1] += g01.y;
        gn[2] += g23.x;
        gp[0] += g23.y;
        gp[1] += g45.x;
        gp[2] += g45.y;
    }
    // the forward's t (hardware reciprocal, as emission computes it)
    const float rD = __builtin_amdgcn_rcpf(va - vb);
    const float tt = (va - T) * rD;
    const float u = 1.0f - tt;
    const float wx = gn[0] * g.scx, wy = gn[1] * g.scy, wz = gn[2] * g.scz;
    float gt = axis == 0 ? gp[0] * g.scx : (axis == 1 ? gp[1] * g.scy : gp[2] * g.scz);
    gt += wx * (nax - nbx) + wy * (nay - nby) + wz * (naz - nbz);
    const float dva = gt * (T - vb) * rD * rD;
    const float dvb = gt * (va - T) * rD * rD;
    const float dnax = -wx * u, dnay = -wy * u, dnaz = -wz * u;
    // contributions by kind: a, b (t and na_axis), a + e_p, a + e_q,
    // b + e_x, b + e_y, b + e_z; then dL/dT
    if (ri != kNoEdge) {
        float4* r = (float4*)(rec + (size_t)ri * kAdjointRecordFloats);
        r[0] = make_float4(dva, dvb + (axis == 0 ? dnax : (axis == 1 ? dnay : dnaz)), axis == 0 ? dnay : dnax,
                           axis == 2 ? dnay : dnaz);
        r[1] = make_float4(-wx * tt, -wy * tt, -wz * tt, 0.0f);
    }
    return (wx + wy + wz) - gt * rD;
}

// records pass.  Lanes take the quads of the cells pass to list the owned
// edges (list: kEdgeListMax 16-bit entries of LDS per wave, Morton cell
// index << 8 | edge << 4 | rank), then one edge each; a block with more
// edges than the list holds has each lane take its own cells' edges.
// smask: the slot masks of fill_slot_masks.  partials[item] = the item's
// dL/dT sum (0 for further chunks).
__device__ __forceinline__ void adjoint_records_body(const VolDesc* __restrict__ vols, float T,
                                                     const uint2* __restrict__ items,
                                                     const Totals* __restrict__ totals,
                                                     const uint16_t (*smask)[12], const uint8_t* __restrict__ codes,
                                                     const uint32_t* __restrict__ cell_off,
                                                     const float* __restrict__ gout, uint32_t nvert,
                                                     float* __restrict__ rec,
                                                     float* __restrict__ partials, uint16_t* list)
{
    const uint32_t nitems = load_uniform(&totals->items);
    const int lane = threadIdx.x & 63;
    const uint32_t w0 = blockIdx.x * kAdjointWaves + (uint32_t)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * kAdjointWaves;
    const int qx = (int)morton_compact((uint32_t)lane) * 2, qy = (int)morton_compact((uint32_t)lane >> 1) * 2;
    for (uint32_t it = w0; it < nitems; it += nw) {
        const uint2 item = load_uniform(items + it);
        if (item.y & 0xfu) {
            if (lane == 0) partials[it] = 0.0f;
            continue;
        }
        const VolDesc d = load_uniform(vols + item_volume(item.y));
        const Geom& g = d.g;
        int x0, y0, z;
        block_origin(g, item.x, x0, y0, z);
        const int side = 1 << g.blog2;
        const uint64_t nsamp = (uint64_t)g.sx * g.sy * g.sz;
        const __amdgpu_buffer_rsrc_t vcodes = buffer_of(codes + d.lattice_base, nsamp);
        const __amdgpu_buffer_rsrc_t voffs = buffer_of(cell_off + d.lattice_base, nsamp * 4u);
        uint32_t m[4], n = 0u;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int x = x0 + qx + (q & 1), y = y0 + qy + (q >> 1);
            const bool valid = qx < side && qy < side && x < g.cx && y < g.cy;
            m[q] = buffer_load_u8(vcodes, valid ? lattice_index(g, x, y, z) : kNoEdge);   // 0 when not valid
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            m[q] = code_edges(m[q]) & owned_edges(g, x0 + qx + (q & 1), y0 + qy + (q >> 1), z);
            n += (uint32_t)__builtin_popcount(m[q]);
        }
        const uint32_t inc = wave_scan_u32<false>(n);
        const uint32_t total = (uint32_t)__builtin_amdgcn_readlane(inc, 63);
        float dts = 0.0f;
        if (total <= (uint32_t)kEdgeListMax) {
            uint32_t pos = inc - n;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                uint32_t mm = m[q], rank = 0u;
                while (mm) {
                    const uint32_t e = (uint32_t)__builtin_ctz(mm);
                    mm &= mm - 1u;
                    list[pos++] = (uint16_t)(((uint32_t)(4 * lane + q) << 8) | (e << 4) | rank++);
                }
            }
            wave_lds_sync();
            for (uint32_t i = (uint32_t)lane; i < total; i += 64) {
                const uint32_t en = list[i];
                const uint32_t c = en >> 8;
                const int x = x0 + (int)morton_compact(c), y = y0 + (int)morton_compact(c >> 1);
                const uint32_t ri = buffer_load_u32(voffs, lattice_index(g, x, y, z) * 4u) + (en & 0xfu);
                dts += adjoint_edge_record(d, T, smask, vcodes, voffs, gout, nvert, x, y, z, (int)((en >> 4) & 0xfu),
                                           rec, ri < nvert ? ri : kNoEdge);
            }
        } else {
#pragma unroll 1
            for (int q = 0; q < 4; ++q) {
                const int x = x0 + qx + (q & 1), y = y0 + qy + (q >> 1);
                uint32_t mm = m[q];
                uint32_t ri = mm ? buffer_load_u32(voffs, lattice_index(g, x, y, z) * 4u) : 0u;
                while (mm) {
                    const int e = __builtin_ctz(mm);
                    mm &= mm - 1u;
                    dts += adjoint_edge_record(d, T, smask, vcodes, voffs, gout, nvert, x, y, z, e, rec,
                                               ri < nvert ? ri : kNoEdge);
                    ++ri;
                }
            }
        }
        dts = wave_sum_f(dts);
        if (lane == 0) partials[it] = dts;
        wave_lds_sync();   // the list is rewritten by the next item
    }
}

// The flagged gather tiles in ascending order (list[0 .. *count)), one
// 1024-thread workgroup; flags is 16-byte aligned and zero past ntiles up to
// a multiple of 16.
__device__ __forceinline__ void adjoint_tile_list_body(const uint8_t* __restrict__ flags, uint32_t ntiles,
                                                       uint32_t* __restrict__ list, uint32_t* __restrict__ count)
{
    __shared__ uint32_t wsum[16];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
    uint32_t done = 0u;
    for (uint32_t base = 0; base < ntiles; base += 16u * blockDim.x) {
        const uint32_t i0 = base + 16u * threadIdx.x;
        const uint4 w = i0 < ntiles ? *(const uint4*)(flags + i0) : make_uint4(0u, 0u, 0u, 0u);
        uint32_t bits = 0u;   // bit k: tile i0 + k flagged
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const uint32_t word = k < 4 ? w.x : (k < 8 ? w.y : (k < 12 ? w.z : w.w));
            bits |= ((word >> (8 * (k & 3))) & 0xffu) ? (1u << k) : 0u;
        }
        const uint32_t n = (uint32_t)__builtin_popcount(bits);
        const uint32_t inc = wave_scan_u32<false>(n);
        if (lane == 63) wsum[wv] = inc;
        __syncthreads();
        uint32_t pos = done + inc - n, round = 0u;
        for (int k = 0; k < nwv; ++k) {
            if (k < wv) pos += wsum[k];
            round += wsum[k];
        }
        __syncthreads();
        while (bits) {
            list[pos++] = i0 + (uint32_t)__builtin_ctz(bits);
            bits &= bits - 1u;
        }
        done += round;
    }
    if (threadIdx.x == 0) *count = done;
}

// gather pass: a persistent grid of kGatherThreads-thread workgroups over
// the listed tiles (the rest of dL/dfield is zero-filled beforehand);
// tile_vol[t] and tile0[v]: the volume of tile t and the first tile of
// volume v.  tab is LDS.
__device__ __forceinline__ void adjoint_gather_body(const VolDesc* __restrict__ vols,
                                                    const uint32_t* __restrict__ tile_vol,
                                                    const uint32_t* __restrict__ tile0,
                                                    const uint32_t* __restrict__ list,
                                                    const uint32_t* __restrict__ count,
                                                    const uint8_t* __restrict__ codes,
                                                    const uint32_t* __restrict__ cell_off,
                                                    const float* __restrict__ rec,
                                                    uint32_t nvert, float* __restrict__ grad,
                                                    uint32_t (*tab)[kGatherOrigins])
{
    const uint32_t ntl = load_uniform(count);
    // the records (indices below 2^27: 32-bit byte offsets; kNoEdge reads
    // past the end)
    const __amdgpu_buffer_rsrc_t recs = buffer_of(rec, (uint64_t)nvert * 4u * kAdjointRecordFloats);
    for (uint32_t li = blockIdx.x; li < ntl; li += gridDim.x) {
        const uint32_t b = load_uniform(list + li);
        const uint32_t v = load_uniform(tile_vol + b);
        const VolDesc d = load_uniform(vols + v);
        const Geom& g = d.g;
        const uint32_t t = b - load_uniform(tile0 + v);
        const uint32_t ntx = (uint32_t)(g.sx + kGatherX - 1) / kGatherX;
        const uint32_t nty = (uint32_t)(g.sy + kGatherY - 1) / kGatherY;
        const int lx = (int)(t % ntx) * kGatherX, ly = (int)((t / ntx) % nty) * kGatherY;
        const int lz = (int)(t / (ntx * nty)) * kGatherZ;
        const uint32_t lb = d.lattice_base;
        const __amdgpu_buffer_rsrc_t offs = buffer_of(cell_off + lb, (uint64_t)g.sx * g.sy * g.sz * 4u);

        // per edge origin lo - 2 .. lo + tile - 1 and axis: the edge's record
        // (kNoEdge where none), written by the owners among the cells
        // lo - 3 .. lo + tile - 1.  A thread takes one column (x, y) of those
        // cells and walks it along z.
#pragma unroll
        for (int k = 0; k < kGatherOriginLoads; ++k) {
            const int i = threadIdx.x + kGatherThreads * k;
            if (i < kGatherOrigins) {
                tab[0][i] = kNoEdge;
                tab[1][i] = kNoEdge;
                tab[2][i] = kNoEdge;
            }
        }
        const int tid = threadIdx.x;
        const bool colt = tid < kGatherCX * kGatherCY;
        const int ux = tid % kGatherCX, uy = tid / kGatherCX;
        const int cx = lx - 3 + ux, cy = ly - 3 + uy;
        const bool colin = colt && cx >= 0 && cy >= 0 && cx < g.cx && cy < g.cy;
        const uint32_t col = colin ? lattice_index(g, cx, cy, 0) : 0u;
        const uint32_t zstride = (uint32_t)g.sx * (uint32_t)g.sy;
        // the column's codes (0 outside the grid), then per cell the owned
        // sign-changing edges (om, which ranks them) and the first vertex of
        // the cells with edges to enter
        uint32_t om[kGatherCZ], oo[kGatherCZ];
#pragma unroll
        for (int uz = 0; uz < kGatherCZ; ++uz) {
            const int cz = lz - 3 + uz;
            om[uz] = (colin && cz >= 0 && cz < g.cz) ? (uint32_t)codes[lb + col + (uint32_t)cz * zstride] : 0u;
        }
        // An edge starts at the cell's origin + its offset, in table
        // coordinates cell + offset - 1: outside the table for offset 0 on
        // the region's first layer and for offset 1 on its last.
        uint32_t keep = 0xfffu;
        if (ux == 0) keep &= 0xaa0u;   // only x offset 1: y and z edges with s0 = 1
        if (uy == 0) keep &= 0xc0au;   // only y offset 1: x edges with s0 = 1, z edges with s1 = 1
        if (ux == kGatherCX - 1) keep &= ~0xaa0u;
        if (uy == kGatherCY - 1) keep &= ~0xc0au;
        const uint32_t own_xy = (cx < g.cx - 1 ? ~0xaa0u : 0xfffu) & (cy < g.cy - 1 ? ~0xc0au : 0xfffu);
        auto keep_z = [&](int uz) {
            return keep & (uz == 0 ? 0x0ccu : 0xfffu) & (uz == kGatherCZ - 1 ? ~0x0ccu : 0xfffu);
        };
#pragma unroll
        for (int uz = 0; uz < kGatherCZ; ++uz) {
            const int cz = lz - 3 + uz;
            const uint32_t m = code_edges(om[uz]) & own_xy & (cz < g.cz - 1 ? ~0x0ccu : 0xfffu);
            om[uz] = m;
            // (cells without edges to enter read past the buffer's end: unused)
            oo[uz] = buffer_load_u32(offs, (m & keep_z(uz)) ? (col + (uint32_t)cz * zstride) * 4u : kNoEdge);
        }
        __syncthreads();   // the table is cleared
#pragma unroll
        for (int uz = 0; uz < kGatherCZ; ++uz) {
            const uint32_t m = om[uz];
            uint32_t w = m & keep_z(uz);
            if (!w) continue;
            const uint32_t ri = oo[uz];
            auto entry = [&](uint32_t e) {   // record of owned edge e: first vertex + rank
                const uint32_t r = ri + (uint32_t)__builtin_popcount(m & ((1u << e) - 1u));
                return r < nvert ? r : kNoEdge;
            };
            // the three edges from the cell's origin share one table slot
            if (w & 0x111u) {
                const int o0 = ((uz - 1) * kGatherOY + (uy - 1)) * kGatherOX + (ux - 1);
                if (w & 0x001u) tab[0][o0] = entry(0);
                if (w & 0x010u) tab[1][o0] = entry(4);
                if (w & 0x100u) tab[2][o0] = entry(8);
                w &= ~0x111u;
            }
            // edges from the cell's upper faces (cells on the grid's upper
            // faces only)
            while (w) {
                const int e = __builtin_ctz(w);
                w &= w - 1u;
                const int axis = e >> 2, s0 = e & 1, s1 = (e >> 1) & 1;
                const int ox = ux - 1 + (axis == 0 ? 0 : s0);
                const int oy = uy - 1 + (axis == 0 ? s0 : (axis == 1 ? 0 : s1));
                const int oz = uz - 1 + (axis == 2 ? 0 : s1);
                tab[axis][(oz * kGatherOY + oy) * kGatherOX + ox] = entry((uint32_t)e);
            }
        }
        __syncthreads();
        // per sample: the 21 (axis, kind) contributions in order, their
        // records looked up first and loaded together; then, on the
        // lattice's upper faces, those that reach it through the clamp
#pragma unroll 1
        for (int j = 0; j < 4; ++j) {
            const int i = threadIdx.x + kGatherThreads * j;
            const int sx = lx + (i & (kGatherX - 1)), sy = ly + ((i / kGatherX) & (kGatherY - 1));
            const int sz = lz + i / (kGatherX * kGatherY);
            if (sx >= g.sx || sy >= g.sy || sz >= g.sz) continue;
            // record of the edge from s - (ox, oy, oz) along axis
            auto at = [&](int axis, int ox, int oy, int oz) {
                return tab[axis][((sz - oz - lz + 2) * kGatherOY + (sy - oy - ly + 2)) * kGatherOX +
                                 (sx - ox - lx + 2)];
            };
            uint32_t en[21];
            float val[21];
#pragma unroll
            for (int axis = 0; axis < 3; ++axis) {
                const int ex = axis == 0, ey = axis == 1, ez = axis == 2;   // e_axis
                const int px = axis != 0, py = axis == 0;                   // e_p (p < q the other axes)
                const int qy = axis == 2, qz = axis != 2;                   // e_q
                uint32_t* e7 = en + 7 * axis;
                e7[0] = at(axis, 0, 0, 0);             // a
                e7[1] = at(axis, ex, ey, ez);          // b
                e7[2] = at(axis, px, py, 0);           // a + e_p
                e7[3] = at(axis, 0, qy, qz);           // a + e_q
                e7[4] = at(axis, ex + 1, ey, ez);      // b + e_x
                e7[5] = at(axis, ex, ey + 1, ez);      // b + e_y
                e7[6] = at(axis, ex, ey, ez + 1);      // b + e_z
            }
            // (kNoEdge reads past the plane's end: 0, and adding +0 to a
            // sum that starts at +0 changes no bits)
#pragma unroll
            for (int c = 0; c < 21; ++c) val[c] = __builtin_bit_cast(float, buffer_load_u32(recs, en[c] * (4u * kAdjointRecordFloats) + 4u * (c % 7)));
            float acc = 0.0f;
#pragma unroll
            for (int c = 0; c < 21; ++c) acc += val[c];
            const bool topx = sx == g.sx - 1, topy = sy == g.sy - 1, topz = sz == g.sz - 1;
            if (topx || topy || topz) {
#pragma unroll
                for (int axis = 0; axis < 3; ++axis) {
                    const int ex = axis == 0, ey = axis == 1, ez = axis == 2;
                    const bool topp = axis == 0 ? topy : topx, topq = axis == 2 ? topy : topz;
                    uint32_t c7[5];
                    c7[0] = topp ? at(axis, 0, 0, 0) : kNoEdge;       // a + e_p clamped to a
                    c7[1] = topq ? at(axis, 0, 0, 0) : kNoEdge;       // a + e_q clamped to a
                    c7[2] = topx ? at(axis, ex, ey, ez) : kNoEdge;    // b + e_x clamped to b
                    c7[3] = topy ? at(axis, ex, ey, ez) : kNoEdge;
                    c7[4] = topz ? at(axis, ex, ey, ez) : kNoEdge;
                    float cv7[5];
#pragma unroll
                    for (int c = 0; c < 5; ++c) cv7[c] = __builtin_bit_cast(float, buffer_load_u32(recs, c7[c] * (4u * kAdjointRecordFloats) + 4u * (2 + c)));
#pragma unroll
                    for (int c = 0; c < 5; ++c) acc += cv7[c];
                }
            }
            grad[lb + lattice_index(g, sx, sy, sz)] = acc;
        }
        __syncthreads();   // tab is rewritten by the next tile
    }
}

// dL/dT: the per-item sums in item order, one 1024-thread workgroup (a fixed
// shape, so the same bits every launch).
__device__ __forceinline__ void adjoint_sum_items_body(const float* __restrict__ partials,
                                                       const Totals* __restrict__ totals, float* __restrict__ out)
{
    __shared__ float sh[16];
    const uint32_t n = totals->items;
    // thread t adds items t, t + 1024, ... in order, eight loads in flight
    float s = 0.0f;
    for (uint32_t i0 = threadIdx.x; i0 < n; i0 += 8u * blockDim.x) {
        float v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const uint32_t i = i0 + (uint32_t)k * blockDim.x;
            v[k] = i < n ? partials[i] : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (i0 + (uint32_t)k * blockDim.x < n) s += v[k];
    }
    s = wave_sum_f(s);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sh[w] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        float r = 0.0f;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r += sh[i];
        *out = r;
    }
}

// The level-4 counts in Morton key order (parity and debugging).
__device__ __forceinline__ void level4_body(const VolDesc& d, const uint16_t* __restrict__ colcnt,
                                            uint32_t* __restrict__ out)
{
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < d.n_level4; k += gridDim.x * blockDim.x)
        out[k] = key_count(d.g, colcnt + d.colcnt_off, k);
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

// Device-side view of one HistoPyramid's last build, for kernels that fuse
// the traversal with their own per-vertex work (the HIP counterpart of the
// reference's traversal shader functions, triface.cpp:125-192).  Filled on
// the host by HPMCgetTraversalView; 64 bytes.
struct TraversalView {
    const VolDesc* vols;
    const uint2* items;
    const uint32_t* l4_off;
    const Totals* totals;
    const uint64_t* cases;
    const float4* bnorm;
    float threshold;
    uint32_t reserved[3];
};

// Visits every vertex of the view's build once: cb(k, normal, position)
// with k the vertex's index in the reference's extraction order and the
// same bits the extraction writes.  Call it from every thread of a kernel
// launched with workgroups of 256 threads (a grid of
// HPMCgetTraversalGrid workgroups, a multiple of 8, spreads the work as the
// extraction does; other grid sizes visit the same vertices).  MK makes the
// field from a volume descriptor, MODE is the emission mode of the build.
template <int MODE, class MK, class CB>
__device__ __forceinline__ void for_each_vertex(const TraversalView& v, const MK& mk, const CB& cb)
{
    __shared__ EmitShared sh;
    if (blockDim.x != 64 * kEmitWaves) return;   // the emission loop needs 4 waves (one table entry a thread)
    const uint32_t nitems = load_uniform(&v.totals->items);
    emit_body<MODE, false>(v.vols, mk, v.threshold, v.items, v.l4_off, nitems, nullptr, nullptr, 0xffffffffu,
                           v.bnorm, sh, v.cases[threadIdx.x], cb);
}

}  // namespace hpmc

// Kernel entry points for one field source.  MAKE_FIELD builds the field
// object from the volume descriptor `dd` (geometry and field pointer).
// HPMC_DEFINE_BUILD_KERNELS: classification and the per-cell code readback.
// HPMC_DEFINE_EMIT_KERNEL: one extraction kernel in emission mode MODE.
#define HPMC_DEFINE_BUILD_KERNELS(SUFFIX, MAKE_FIELD)                                          \
    extern "C" __global__ __launch_bounds__(1024) void hpmc_classify_##SUFFIX(                 \
        const hpmc::VolDesc* vols, const uint32_t* cls_vol, uint32_t n_wg, float T,            \
        uint16_t* colcnt, const uint64_t* cases, uint2* tile_sums, uint32_t n_l4t)             \
    {                                                                                         \
        __shared__ uint8_t lut[256];                                                          \
        __shared__ uint32_t xch[2][hpmc::kClassifyMaxWaves];                                   \
        __shared__ uint16_t stage[hpmc::kClassifyMaxWaves][4][hpmc::kColStride];               \
        __shared__ uint2 red[hpmc::kClassifyMaxWaves][hpmc::kColStride];                       \
        /* grid padded to a multiple of 8: XCD x takes logical run x */                       \
        const uint32_t t = (blockIdx.x & 7u) * (gridDim.x >> 3) + (blockIdx.x >> 3);          \
        if (t >= n_wg) return;                                                                \
        /* cls_vol == nullptr: a plan of one volume; loaded before the LDS set-up */           \
        const hpmc::VolDesc dd = vols[cls_vol ? __builtin_amdgcn_readfirstlane(cls_vol[t]) : 0u]; \
        for (int i = threadIdx.x; i < 256; i += blockDim.x) lut[i] = (uint8_t)(cases[i] & 0xfu); \
        __syncthreads();                                                                      \
        hpmc::classify_body(MAKE_FIELD, dd.g, T, colcnt + dd.colcnt_off, dd.g.zseg, lut, xch, \
                            stage, (int)(t - dd.cls_wg0), red, tile_sums + dd.l4_tile0,       \
                            tile_sums + n_l4t + dd.col_tile0);                                \
    }                                                                                         \
    extern "C" __global__ __launch_bounds__(256) void hpmc_classify_small_##SUFFIX(            \
        const hpmc::VolDesc* vols, float T, uint16_t* colcnt, const uint64_t* cases,            \
        uint2* tile_sums, uint32_t n_l4t)                                                     \
    {                                                                                         \
        __shared__ uint8_t lut[256];                                                          \
        const hpmc::VolDesc dd = vols[0];                                                     \
        lut[threadIdx.x] = (uint8_t)(cases[threadIdx.x] & 0xfu);                              \
        __syncthreads();                                                                      \
        hpmc::classify_small_body(MAKE_FIELD, dd.g, T, colcnt + dd.colcnt_off, dd.n_cc, lut,  \
                                  tile_sums + dd.l4_tile0, tile_sums + n_l4t + dd.col_tile0);  \
    }                                                                                         \
    extern "C" __global__ __launch_bounds__(256) void hpmc_codes_##SUFFIX(                     \
        const hpmc::VolDesc* vols, uint32_t vol, float T, uint8_t* codes)                     \
    {                                                                                         \
        const hpmc::VolDesc dd = vols[vol];                                                   \
        hpmc::codes_body(MAKE_FIELD, dd.g, T, codes);                                         \
    }

// BATCH: items of many volumes (the batched build) or of vols[0] only.
#define HPMC_DEFINE_EMIT_KERNEL(NAME, MAKE_FIELD, MODE, BATCH)                                 \
    extern "C" __global__ __launch_bounds__(256) void NAME(                                   \
        const hpmc::VolDesc* vols, float T, const uint2* items, const uint32_t* l4_off,        \
        const uint32_t* item_reps, float* out, hpmc::CellEdge* out_ce, uint32_t capacity,     \
        const uint64_t* cases, const float4* bnorm, uint2* tile_sums, uint32_t ntiles)        \
    {                                                                                         \
        __shared__ hpmc::EmitShared sh;                                                       \
        const uint32_t nitems =                                                               \
            hpmc::load_uniform(item_reps + (blockIdx.x % hpmc::kItemReps) * hpmc::kItemRepStride); \
        const uint64_t case_word = cases[threadIdx.x];   /* stored to LDS by emit_body */     \
        auto mk = [](const hpmc::VolDesc& dd) { return MAKE_FIELD; };                         \
        hpmc::emit_body<MODE, BATCH>(vols, mk, T, items, l4_off, nitems, out, out_ce,         \
                                     capacity, bnorm, sh, case_word);                         \
        /* the build's scan tile sums, read by its scan only: cleared for the next build, */  \
        /* by each workgroup's first thread once its wave is done */                          \
        if (threadIdx.x == 0)                                                                 \
            for (uint32_t i = blockIdx.x; i < ntiles; i += gridDim.x) tile_sums[i] = make_uint2(0u, 0u); \
    }
