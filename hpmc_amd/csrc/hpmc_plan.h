// Host-side plan of a build: the per-volume geometry, the concatenated
// device arrays and the launch shapes of classify -> scan -> emit, for one
// volume (a HistoPyramid) or many (a batch).  Included by hpmc_api.cpp only.
//
// Device arrays of a plan (volumes concatenated in order):
//   vols       VolDesc per volume (geometry, field pointer, array offsets)
//   cls_vol    volume of each logical classification workgroup
//   tile_vol   volume of each scan tile (level-4 tiles, then column tiles)
//   colcnt     u16 per (volume, z segment, y block, x block, slice), 32
//              slices per block column (whole 64-byte lines)
//   l4_off     u32 per (volume, level-4 key): global offset of the block's
//              first vertex in the concatenated output
//   tile_sums  uint2 per scan tile, accumulated by classify (atomics), read
//              by the scan and cleared by the extraction that follows (or
//              before the next build, when none followed: sums_dirty)
//   items      uint2 per emit work item (<= 15 per non-empty block slice)
//   totals / vol_counts: grand totals and per-volume (base, count), also
//              written straight into pinned host memory
#pragma once

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

namespace hpmc_host {

struct VolumeSpec {
    int32_t size[3];
    int32_t cells[3];
    float extent[3];
    const void* field;
};

struct Layout {
    int32_t tile[2];
    int32_t tiles[2];
    int32_t l2;
};

// hpmc/src/init.cpp:72-149.  FP32 log2f/ceilf/sqrtf as in the reference.
// Returns HPMC_OK or an error code with *err set.
inline int32_t layout_of(const int32_t size[3], const int32_t cells[3], Layout* out, std::string* err)
{
    char buf[256];
    for (int a = 0; a < 3; ++a) {
        if (size[a] < 2 || size[a] > 16384) {
            snprintf(buf, sizeof buf, "lattice size %d out of [2,16384] along axis %d", size[a], a);
            *err = buf;
            return HPMC_ERR_SIZE;
        }
        if (size[a] <= cells[a] || cells[a] < 1) {
            snprintf(buf, sizeof buf, "grid size %d must be in [1, lattice-1=%d] along axis %d", cells[a],
                     size[a] - 1, a);
            *err = buf;
            return HPMC_ERR_SIZE;
        }
    }
    // cells == 1 along x or y gives log2f(0.5f) = -1 and a negative shift in
    // the reference (init.cpp:106-111); it cannot lay out such a grid either.
    if (cells[0] < 2 || cells[1] < 2) {
        *err = "grid must have at least 2 cells along x and y";
        return HPMC_ERR_SIZE;
    }
    const int32_t ts0 = 1 << (int32_t)ceilf(log2f((float)cells[0] / 2.0f));
    const int32_t ts1 = 1 << (int32_t)ceilf(log2f((float)cells[1] / 2.0f));
    const float aspect = (float)ts0 / (float)ts1;
    float l = ceilf(log2f(sqrtf((float)cells[2] / aspect)));
    if (l < 0.0f) l = 0.0f;
    const int32_t lay0 = 1 << (int32_t)l;
    const int32_t lay1 = (cells[2] + lay0 - 1) / lay0;
    const int64_t w0 = (int64_t)ts0 * lay0, w1 = (int64_t)ts1 * lay1;
    const int32_t l2 = (int32_t)ceilf(log2f((float)(w0 > w1 ? w0 : w1)));
    const int32_t S = 1 << l2;
    out->tile[0] = ts0;
    out->tile[1] = ts1;
    out->tiles[0] = S / ts0;
    out->tiles[1] = S / ts1;
    out->l2 = l2;
    return HPMC_OK;
}

// Classification workgroup waves: ceil(X/64) (whole lattice rows), at most 16.
inline int row_waves(int sx)
{
    const int nw = (sx + 63) / 64;
    return nw > hpmc::kClassifyMaxWaves ? hpmc::kClassifyMaxWaves : nw;
}

// Classification workgroups of one slice segment: nw waves span
// wx = min(nw, ceil(X/64)) waves across x and stack nw / wx row bands
// (classify_shape in the kernels).
inline int64_t classify_bands(const hpmc::Geom& g, int nw)
{
    const int wx = std::min(nw, (g.sx + 63) / 64);
    const int sb = nw / wx;
    const int64_t wgx = (g.sx + 64 * wx - 1) / (64 * wx);
    const int64_t npy = (g.cy + hpmc::kBlock - 1) / hpmc::kBlock;
    return wgx * ((npy + sb - 1) / sb);
}

// zseg so the workgroups of `cells_z` slices over `bands` row bands fill
// `resident_wgs` workgroup slots about once: the power of two in 8..32 at or
// above that (32 = the slots of a block column in the count array).
inline int auto_zseg(int64_t bands, int cz, int64_t resident_wgs)
{
    const int64_t target = resident_wgs > 0 ? resident_wgs : 1024;
    const int64_t z = (cz * bands + target - 1) / target;
    int zs = 8;
    while (zs < z && zs < hpmc::kColStride) zs *= 2;
    return zs;
}

inline hpmc::Geom make_geom(const VolumeSpec& v, const Layout& lay, int zseg)
{
    hpmc::Geom g{};
    g.sx = v.size[0];
    g.sy = v.size[1];
    g.sz = v.size[2];
    g.cx = v.cells[0];
    g.cy = v.cells[1];
    g.cz = v.cells[2];
    g.tsx2 = 2 * lay.tile[0];
    g.tsy2 = 2 * lay.tile[1];
    g.tiles_x = lay.tiles[0];
    g.l2 = lay.l2;
    g.scx = v.extent[0] / (float)v.cells[0];
    g.scy = v.extent[1] / (float)v.cells[1];
    g.scz = v.extent[2] / (float)v.cells[2];
    g.ipx = 1.0f / (float)v.size[0];
    g.ipy = 1.0f / (float)v.size[1];
    g.ipz = 1.0f / (float)v.size[2];
    g.zseg = zseg;
    auto log2_of = [](int v) {
        int l = 0;
        while ((1 << l) < v) ++l;
        return l;
    };
    g.tsx2_log2 = log2_of(g.tsx2);
    g.tsy2_log2 = log2_of(g.tsy2);
    g.tiles_x_log2 = log2_of(g.tiles_x);
    g.zseg_log2 = log2_of(zseg);
    // blocks of 16x16 cells (HP level 4), or the narrower doubled tile's
    // side when a tile is below 16 cells: a block is always cells of one slice
    g.blog2 = std::min(hpmc::kBlockLog2, std::min(g.tsx2_log2, g.tsy2_log2));
    g.blocks_log2 = lay.l2 + 1 - g.blog2;
    const int bs = 1 << g.blog2;
    g.nbx = (g.cx + bs - 1) / bs;
    g.nby = (g.cy + bs - 1) / bs;
    g.npz = (g.cz + zseg - 1) / zseg;
    return g;
}

// Launch-shape choices of a plan (0 = automatic).
struct Tuning {
    int classify_waves = 0;   // waves per classification workgroup
    int zseg = 0;             // slices per classification workgroup (8, 16, 32)
    int emit_wgs = 0;         // persistent emit workgroups (rounded to 8)
};

struct Plan {
    int nvol = 0;
    std::vector<hpmc::VolDesc> vols;
    std::vector<Layout> layouts;
    int cls_waves = 1;
    uint32_t n_wg = 0;            // logical classification workgroups
    uint32_t n_l4t = 0, n_colt = 0;
    uint64_t n_cc = 0, n_l4 = 0, n_items_max = 0;
    uint64_t n_lattice = 0;       // samples of all volumes (adjoint keys)
    int emit_wgs = 0;
    hpmc::VolDesc* d_vols = nullptr;
    uint32_t* d_cls_vol = nullptr;
    uint32_t* d_tile_vol = nullptr;
    uint16_t* colcnt = nullptr;
    uint32_t* l4_off = nullptr;
    uint2* tile_sums = nullptr;
    uint2* items = nullptr;
    hpmc::Totals* d_totals = nullptr;
    uint32_t* d_item_reps = nullptr;      // copies of the item total for emission
    hpmc::Totals* h_totals = nullptr;     // pinned
    hpmc::VolCount* d_vcounts = nullptr;
    hpmc::VolCount* h_vcounts = nullptr;  // pinned
    uint32_t* d_gather_tile0 = nullptr;   // first adjoint gather tile of each volume
    uint32_t* d_gather_vol = nullptr;     // volume of each adjoint gather tile
    uint64_t n_gather_tiles = 0;
    // host view of the plan's stream: the tile sums hold a build's counts
    // that no extraction has cleared yet
    bool sums_dirty = false;
};

inline void plan_free(Plan& p)
{
    if (p.d_vols) (void)hipFree(p.d_vols);
    if (p.d_cls_vol) (void)hipFree(p.d_cls_vol);
    if (p.d_tile_vol) (void)hipFree(p.d_tile_vol);
    if (p.colcnt) (void)hipFree(p.colcnt);
    if (p.l4_off) (void)hipFree(p.l4_off);
    if (p.tile_sums) (void)hipFree(p.tile_sums);
    if (p.items) (void)hipFree(p.items);
    if (p.d_totals) (void)hipFree(p.d_totals);
    if (p.d_item_reps) (void)hipFree(p.d_item_reps);
    if (p.h_totals) (void)hipHostFree(p.h_totals);
    if (p.d_vcounts) (void)hipFree(p.d_vcounts);
    if (p.h_vcounts) (void)hipHostFree(p.h_vcounts);
    if (p.d_gather_tile0) (void)hipFree(p.d_gather_tile0);
    if (p.d_gather_vol) (void)hipFree(p.d_gather_vol);
    p = Plan();
}

// Sizes the device share of the classification and emit launches once per
// process: resident workgroups of each kernel on this device.
struct DeviceShape {
    int cus = 0;
    int emit_per_cu = 0;
};

#define PLAN_TRY(expr)                                                                     \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess) {                                                            \
            *err = std::string(#expr) + ": " + hipGetErrorString(_e);                      \
            plan_free(p);                                                                  \
            return HPMC_ERR_HIP;                                                           \
        }                                                                                  \
    } while (0)

// Lays out a plan for specs[0 .. n) without touching device memory: the
// layouts (lays[n]), launch shapes, per-volume descriptors (vols[n]) and the
// sizes of every device array.  `classify_kernel` is the kernel whose
// occupancy sizes the automatic zseg; `stack_bands` allows the automatic
// shape to stack row bands of 4- and 8-wave rows (FP32 volumes).  Only an
// error allocates (the message).  Returns HPMC_OK or an error code with
// *err set.
inline int32_t plan_describe(Plan& p, const VolumeSpec* specs, size_t n, Layout* lays, hpmc::VolDesc* vols,
                             const Tuning& tune, const void* classify_kernel, const DeviceShape& dev,
                             bool stack_bands, std::string* err)
{
    if (n == 0 || n > 65535) {
        *err = "a build needs 1 to 65535 volumes";
        return HPMC_ERR_ARG;
    }
    p.nvol = (int)n;
    int max_row_waves = 1;
    for (size_t i = 0; i < n; ++i) {
        const int32_t rc = layout_of(specs[i].size, specs[i].cells, &lays[i], err);
        if (rc != HPMC_OK) {
            char buf[32];
            snprintf(buf, sizeof buf, "volume %zu: ", i);
            *err = buf + *err;
            return rc;
        }
        max_row_waves = std::max(max_row_waves, row_waves(specs[i].size[0]));
    }
    // one workgroup shape per launch unless tuned: whole rows of the widest
    // volume, at least 4 waves (narrower volumes stack row bands); for FP32
    // volumes, rows of exactly 4 or 8 waves (X in 193..256 or 449..512)
    // stack two bands, which measured faster there for sparse and dense
    // surfaces alike (classification of the 512^3 Cayley volume 97.5 ->
    // 92.7 us, of a 256^3 one 40.7 -> 32.7 us, of a 256^3 volume with a
    // vertex per cell 82 -> 67 us), and slower for rows of 6 waves (384^3)
    // and for 8-bit volumes (512^3: 65 -> 70 us)
    int auto_waves = std::max(max_row_waves, 4);
    if (stack_bands && (max_row_waves == 4 || max_row_waves == 8)) auto_waves = 2 * max_row_waves;
    p.cls_waves = tune.classify_waves > 0 ? std::min(tune.classify_waves, (int)hpmc::kClassifyMaxWaves)
                                          : auto_waves;
    int per_cu = 0;
    if (classify_kernel)
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, classify_kernel, 64 * p.cls_waves, 0);
    // automatic zseg: fill the device once with all volumes' workgroups
    int64_t bands_cells = 0;
    for (size_t i = 0; i < n; ++i) {
        hpmc::Geom g0 = make_geom(specs[i], lays[i], 8);
        bands_cells += classify_bands(g0, p.cls_waves) * g0.cz;
    }
    const int64_t resident = (int64_t)(dev.cus > 0 ? dev.cus : 256) * (per_cu > 0 ? per_cu : 1);
    const int zseg_auto = auto_zseg(1, (int)std::min<int64_t>(bands_cells, INT32_MAX), resident);
    const int zseg = tune.zseg > 0 ? tune.zseg : zseg_auto;
    if (zseg != 8 && zseg != 16 && zseg != 32) {
        *err = "zseg must be 8, 16 or 32";
        return HPMC_ERR_ARG;
    }

    p.n_wg = 0;
    p.n_l4t = p.n_colt = 0;
    p.n_cc = p.n_l4 = p.n_items_max = p.n_lattice = 0;
    for (size_t i = 0; i < n; ++i) {
        hpmc::VolDesc& d = vols[i];
        d = hpmc::VolDesc{};
        d.g = make_geom(specs[i], lays[i], zseg);
        d.field = specs[i].field;
        const uint64_t side = (uint64_t)1 << d.g.blocks_log2;
        const uint64_t n_level4 = side * side;
        const uint64_t n_cc = (uint64_t)d.g.nbx * d.g.nby * d.g.npz * hpmc::kColStride;
        // a block of b x b cells holds at most 15 b^2 vertices: 15 work
        // items per slice of a 16x16 block, fewer for narrower blocks
        const uint64_t bcells = (uint64_t)1 << (2 * d.g.blog2);
        const uint64_t n_items =
            (uint64_t)d.g.nbx * d.g.nby * d.g.cz * ((15 * bcells + hpmc::kEmitChunk - 1) / hpmc::kEmitChunk);
        const uint64_t wgs = (uint64_t)classify_bands(d.g, p.cls_waves) * d.g.npz;
        if (p.n_l4 + n_level4 > 0xffffffffull || p.n_cc + n_cc > 0xffffffffull ||
            p.n_items_max + n_items > 0xffffffffull || p.n_wg + wgs > 0xffffffffull) {
            *err = "build too large for 32-bit offsets";
            return HPMC_ERR_SIZE;
        }
        d.n_level4 = (uint32_t)n_level4;
        d.n_cc = (uint32_t)n_cc;
        d.colcnt_off = p.n_cc;
        d.l4_base = (uint32_t)p.n_l4;
        d.cls_wg0 = p.n_wg;
        d.l4_tile0 = p.n_l4t;
        d.n_l4tiles = (d.n_level4 + hpmc::kScanTile - 1) / hpmc::kScanTile;
        d.col_tile0 = p.n_colt;
        d.n_coltiles = (d.n_cc + hpmc::kScanTile - 1) / hpmc::kScanTile;
        d.lattice_base = (uint32_t)p.n_lattice;   // the adjoint takes builds under 2^32 samples
        p.n_l4t += d.n_l4tiles;
        p.n_colt += d.n_coltiles;
        p.n_cc += d.n_cc;
        p.n_l4 += d.n_level4;
        p.n_wg += (uint32_t)wgs;
        p.n_lattice += (uint64_t)d.g.sx * d.g.sy * d.g.sz;
        p.n_items_max += n_items;
    }

    // persistent emit grid: as many 4-wave workgroups as fit at once, a
    // multiple of 8 for the XCD-aware item mapping, never more than the items
    int emit_wgs = tune.emit_wgs > 0 ? tune.emit_wgs
                                     : (dev.cus > 0 ? dev.cus : 256) * (dev.emit_per_cu > 0 ? dev.emit_per_cu : 4);
    const uint64_t max_wg = (p.n_items_max + hpmc::kEmitWaves - 1) / hpmc::kEmitWaves;
    emit_wgs = (int)std::min<uint64_t>((uint64_t)emit_wgs, std::max<uint64_t>(max_wg, 1));
    p.emit_wgs = (emit_wgs + 7) / 8 * 8;
    return HPMC_OK;
}

// Lays out and allocates a plan for `specs` (plan_describe, then the device
// arrays).  Returns HPMC_OK or an error code with *err set (the plan is then
// empty).
inline int32_t plan_create(Plan& p, const std::vector<VolumeSpec>& specs, const Tuning& tune,
                           const void* classify_kernel, const DeviceShape& dev, bool stack_bands,
                           std::string* err)
{
    plan_free(p);
    p.layouts.resize(specs.size());
    p.vols.resize(specs.size());
    const int32_t rc = plan_describe(p, specs.data(), specs.size(), p.layouts.data(), p.vols.data(), tune,
                                     classify_kernel, dev, stack_bands, err);
    if (rc != HPMC_OK) {
        p = Plan();
        return rc;
    }
    // volume of each classification workgroup and scan tile (level-4 tiles
    // of all volumes, then their column tiles)
    std::vector<uint32_t> cls_vol, tile_vol;
    cls_vol.reserve(p.n_wg);
    tile_vol.reserve((size_t)p.n_l4t + p.n_colt);
    for (size_t i = 0; i < specs.size(); ++i) {
        const hpmc::VolDesc& d = p.vols[i];
        const uint32_t next_wg = i + 1 < specs.size() ? p.vols[i + 1].cls_wg0 : p.n_wg;
        cls_vol.insert(cls_vol.end(), next_wg - d.cls_wg0, (uint32_t)i);
        tile_vol.insert(tile_vol.end(), d.n_l4tiles, (uint32_t)i);
    }
    for (size_t i = 0; i < specs.size(); ++i) tile_vol.insert(tile_vol.end(), p.vols[i].n_coltiles, (uint32_t)i);

    PLAN_TRY(hipMalloc(&p.d_vols, p.vols.size() * sizeof(hpmc::VolDesc)));
    PLAN_TRY(hipMemcpy(p.d_vols, p.vols.data(), p.vols.size() * sizeof(hpmc::VolDesc), hipMemcpyHostToDevice));
    PLAN_TRY(hipMalloc(&p.d_cls_vol, cls_vol.size() * sizeof(uint32_t)));
    PLAN_TRY(hipMemcpy(p.d_cls_vol, cls_vol.data(), cls_vol.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    PLAN_TRY(hipMalloc(&p.d_tile_vol, tile_vol.size() * sizeof(uint32_t)));
    PLAN_TRY(hipMemcpy(p.d_tile_vol, tile_vol.data(), tile_vol.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    PLAN_TRY(hipMalloc(&p.colcnt, p.n_cc * sizeof(uint16_t)));
    PLAN_TRY(hipMemset(p.colcnt, 0, p.n_cc * sizeof(uint16_t)));
    PLAN_TRY(hipMalloc(&p.l4_off, p.n_l4 * sizeof(uint32_t)));
    PLAN_TRY(hipMalloc(&p.tile_sums, (size_t)(p.n_l4t + p.n_colt) * sizeof(uint2)));
    PLAN_TRY(hipMemset(p.tile_sums, 0, (size_t)(p.n_l4t + p.n_colt) * sizeof(uint2)));
    PLAN_TRY(hipMalloc(&p.items, std::max<uint64_t>(p.n_items_max, 1) * sizeof(uint2)));
    PLAN_TRY(hipMalloc(&p.d_totals, sizeof(hpmc::Totals)));
    PLAN_TRY(hipMemset(p.d_totals, 0, sizeof(hpmc::Totals)));
    PLAN_TRY(hipMalloc(&p.d_item_reps, hpmc::kItemReps * hpmc::kItemRepStride * sizeof(uint32_t)));
    PLAN_TRY(hipMemset(p.d_item_reps, 0, hpmc::kItemReps * hpmc::kItemRepStride * sizeof(uint32_t)));
    PLAN_TRY(hipHostMalloc(&p.h_totals, sizeof(hpmc::Totals), hipHostMallocDefault));
    memset(p.h_totals, 0, sizeof(hpmc::Totals));
    // adjoint gather tiles of each volume
    std::vector<uint32_t> tile0(specs.size()), gather_vol;
    for (size_t i = 0; i < specs.size(); ++i) {
        const hpmc::Geom& g = p.vols[i].g;
        tile0[i] = (uint32_t)gather_vol.size();
        const uint64_t n = (uint64_t)((g.sx + hpmc::kGatherX - 1) / hpmc::kGatherX) *
                           ((g.sy + hpmc::kGatherY - 1) / hpmc::kGatherY) *
                           ((g.sz + hpmc::kGatherZ - 1) / hpmc::kGatherZ);
        gather_vol.insert(gather_vol.end(), n, (uint32_t)i);
    }
    p.n_gather_tiles = gather_vol.size();
    PLAN_TRY(hipMalloc(&p.d_gather_tile0, tile0.size() * sizeof(uint32_t)));
    PLAN_TRY(hipMemcpy(p.d_gather_tile0, tile0.data(), tile0.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    PLAN_TRY(hipMalloc(&p.d_gather_vol, gather_vol.size() * sizeof(uint32_t)));
    PLAN_TRY(hipMemcpy(p.d_gather_vol, gather_vol.data(), gather_vol.size() * sizeof(uint32_t),
                       hipMemcpyHostToDevice));
    PLAN_TRY(hipMalloc(&p.d_vcounts, p.vols.size() * sizeof(hpmc::VolCount)));
    PLAN_TRY(hipHostMalloc(&p.h_vcounts, p.vols.size() * sizeof(hpmc::VolCount), hipHostMallocDefault));
    memset(p.h_vcounts, 0, p.vols.size() * sizeof(hpmc::VolCount));
    PLAN_TRY(hipDeviceSynchronize());
    return HPMC_OK;
}

// Device bytes of a one-volume plan's arrays carved from one buffer (each
// array 256-byte aligned): descriptor, totals, item total copies, volume count,
// tile sums, column counts, block offsets, work items.
inline uint64_t plan_carve(Plan& p, char* base)
{
    uint64_t off = 0;
    auto take = [&](uint64_t bytes) {
        char* q = base ? base + off : nullptr;
        off += (bytes + 255) / 256 * 256;
        return q;
    };
    p.d_vols = (hpmc::VolDesc*)take(sizeof(hpmc::VolDesc));
    p.d_totals = (hpmc::Totals*)take(sizeof(hpmc::Totals));
    p.d_item_reps = (uint32_t*)take(hpmc::kItemReps * hpmc::kItemRepStride * sizeof(uint32_t));
    p.d_vcounts = (hpmc::VolCount*)take(sizeof(hpmc::VolCount));
    p.tile_sums = (uint2*)take((uint64_t)(p.n_l4t + p.n_colt) * sizeof(uint2));
    p.colcnt = (uint16_t*)take(p.n_cc * sizeof(uint16_t));
    p.l4_off = (uint32_t*)take(p.n_l4 * sizeof(uint32_t));
    p.items = (uint2*)take(std::max<uint64_t>(p.n_items_max, 1) * sizeof(uint2));
    return off;
}

// Replaces the field pointer of volume v (a new texture for the same
// configuration): device descriptor updated in stream order.
inline hipError_t plan_set_field(Plan& p, int v, const void* field, hipStream_t stream)
{
    if (p.vols[v].field == field) return hipSuccess;
    p.vols[v].field = field;
    return hipMemcpyAsync(&p.d_vols[v].field, &p.vols[v].field, sizeof(void*), hipMemcpyHostToDevice, stream);
}

// Kernel handle: a built-in kernel (host stub) or a hipRTC function.
struct Kernel {
    const void* builtin = nullptr;
    hipFunction_t rtc = nullptr;
};

// A kernel launch.  With events (stage profiling) the kernel's own dispatch
// timestamps them at its start and end (the extended launch), so no marker
// packets sit between the kernels and the elapsed time is the kernel's.
inline hipError_t launch(const Kernel& k, unsigned grid, unsigned threads, void** args, hipStream_t stream,
                         hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr)
{
    if (k.rtc) {
        if (e0)
            return hipExtModuleLaunchKernel(k.rtc, grid * threads, 1, 1, threads, 1, 1, 0, stream, args, nullptr, e0,
                                            e1, 0);
        return hipModuleLaunchKernel(k.rtc, grid, 1, 1, threads, 1, 1, 0, stream, args, nullptr);
    }
    if (e0) return hipExtLaunchKernel(k.builtin, dim3(grid), dim3(threads), args, 0, stream, e0, e1, 0);
    return hipLaunchKernel(k.builtin, dim3(grid), dim3(threads), args, 0, stream);
}

inline hipError_t plan_classify(const Plan& p, const Kernel& k, float T, const uint64_t* cases, hipStream_t stream,
                                hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr)
{
    const hpmc::VolDesc* vols = p.d_vols;
    const uint32_t* cv = p.nvol == 1 ? nullptr : p.d_cls_vol;   // one volume: no lookup
    uint32_t n = p.n_wg;
    float Ta = T;
    uint16_t* cc = p.colcnt;
    const uint64_t* cs = cases;
    uint2* ts = p.tile_sums;
    uint32_t nl = p.n_l4t;
    void* args[] = {&vols, &cv, &n, &Ta, &cc, &cs, &ts, &nl};
    // the XCD-aware workgroup mapping needs the grid padded to a multiple of 8
    const unsigned grid = (unsigned)(((uint64_t)p.n_wg + 7) / 8 * 8);
    return launch(k, grid, 64u * (unsigned)p.cls_waves, args, stream, e0, e1);
}

// Classification of a one-volume plan whose blocks are narrower than 16
// cells (classify_small_body): one thread per column-count entry.
inline hipError_t plan_classify_small(const Plan& p, const Kernel& k, float T, const uint64_t* cases,
                                      hipStream_t stream, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr)
{
    const hpmc::VolDesc* vols = p.d_vols;
    float Ta = T;
    uint16_t* cc = p.colcnt;
    const uint64_t* cs = cases;
    uint2* ts = p.tile_sums;
    uint32_t nl = p.n_l4t;
    void* args[] = {&vols, &Ta, &cc, &cs, &ts, &nl};
    const unsigned grid = (unsigned)((p.n_cc + 255) / 256);
    return launch(k, grid, 256, args, stream, e0, e1);
}

// True when every volume's cell indices fit the 32-bit cell field of the
// (cell, edge) records.
inline bool plan_cells_fit_u32(const Plan& p)
{
    for (const hpmc::VolDesc& d : p.vols)
        if ((uint64_t)d.g.cx * d.g.cy * d.g.cz > 0xffffffffull) return false;
    return true;
}

// True when the plan's (single) volume has blocks narrower than 16 cells.
inline bool plan_small_blocks(const Plan& p, const hpmc::VolDesc& d)
{
    return p.nvol == 1 && d.g.blog2 < hpmc::kBlockLog2;
}

inline hipError_t plan_scan(const Plan& p, const void* k_scan, hipStream_t stream, hipEvent_t e0 = nullptr,
                            hipEvent_t e1 = nullptr)
{
    const unsigned grid = p.n_l4t + p.n_colt;
    const hpmc::VolDesc* vols = p.d_vols;
    const uint32_t* tv = p.nvol == 1 ? nullptr : p.d_tile_vol;
    uint32_t nl = p.n_l4t, nc = p.n_colt;
    const uint16_t* cc = p.colcnt;
    const uint2* ts = p.tile_sums;
    uint32_t* ir = p.d_item_reps;
    uint32_t* lo = p.l4_off;
    uint2* it = p.items;
    hpmc::Totals* dt = p.d_totals;
    hpmc::Totals* ht = p.h_totals;
    hpmc::VolCount* dv = p.d_vcounts;
    hpmc::VolCount* hv = p.h_vcounts;
    void* args[] = {&vols, &tv, &nl, &nc, &cc, &ts, &ir, &lo, &it, &dt, &ht, &dv, &hv};
    return launch(Kernel{k_scan, nullptr}, grid, hpmc::kScanThreads, args, stream, e0, e1);
}

inline hipError_t plan_emit(const Plan& p, const Kernel& k, float T, float* out, hpmc::CellEdge* out_ce,
                            uint32_t capacity, const uint64_t* cases, const float4* bnorm, hipStream_t stream,
                            hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr)
{
    const hpmc::VolDesc* vols = p.d_vols;
    float Ta = T;
    const uint2* it = p.items;
    const uint32_t* lo = p.l4_off;
    const uint32_t* tt = p.d_item_reps;
    float* o = out;
    hpmc::CellEdge* oc = out_ce;
    uint32_t cap = capacity;
    const uint64_t* cs = cases;
    const float4* bn = bnorm;
    uint2* ts = p.tile_sums;
    uint32_t nts = p.n_l4t + p.n_colt;
    void* args[] = {&vols, &Ta, &it, &lo, &tt, &o, &oc, &cap, &cs, &bn, &ts, &nts};
    return launch(k, (unsigned)p.emit_wgs, 256, args, stream, e0, e1);
}

}  // namespace hpmc_host
