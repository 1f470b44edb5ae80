// Host side of the MI355X HistoPyramid marching-cubes engine: the C ABI
// declared in include/hpmc.h, state machine, layout, kernel launches and the
// hipRTC path for field snippets.
//
// Reference behaviour mirrored here (paths relative to hetcomp/hpmc):
//   handle life cycle, taint/broken latch ...... hpmc/src/hpiface.cpp:45-200
//   lazy setup on build ........................ hpmc/src/init.cpp:48-69
//   layout (tiles of z-slices in a 2^l2 square)  hpmc/src/init.cpp:72-149
//   build / count .............................. hpmc/src/hpiface.cpp:225-345
//   traversal handle and extraction ............ hpmc/src/triface.cpp:43-500
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <hip/hiprtc.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "hpmc.h"
#include "hpmc_device.h"

// Device sources embedded for hipRTC (generated by hpmc_amd/build.py).
#include "_embedded_sources.inc"

// Built-in kernels for FP32 volume fields (scalar, and RGBA = gradient +
// value); snippet fields get the same kernels from hipRTC.
#define HPMC_VOLUME_FIELD (hpmc::VolumeField{(const float*)dd.field, dd.g.sx, dd.g.sy, dd.g.sz})
#define HPMC_VOLUME4_FIELD (hpmc::Volume4Field{(const float4*)dd.field, dd.g.sx, dd.g.sy, dd.g.sz})
#define HPMC_VOLUME8_FIELD (hpmc::Volume8Field{(const uint8_t*)dd.field, dd.g.sx, dd.g.sy, dd.g.sz})
#define HPMC_VOLUME_STREAM_FIELD                                                                     \
    (hpmc::VolumeStreamField{{(const float*)dd.field, dd.g.sx, dd.g.sy, dd.g.sz}})
HPMC_DEFINE_BUILD_KERNELS(volume, HPMC_VOLUME_FIELD)
HPMC_DEFINE_BUILD_KERNELS(volume_stream, HPMC_VOLUME_STREAM_FIELD)
HPMC_DEFINE_BUILD_KERNELS(volume4, HPMC_VOLUME4_FIELD)
HPMC_DEFINE_BUILD_KERNELS(volume8, HPMC_VOLUME8_FIELD)
HPMC_DEFINE_EMIT_KERNEL(hpmc_emit_volume, HPMC_VOLUME_FIELD, hpmc::kEmitContinuous, false)
HPMC_DEFINE_EMIT_KERNEL(hpmc_emit_volume_binary, HPMC_VOLUME_FIELD, hpmc::kEmitBinary, false)
HPMC_DEFINE_EMIT_KERNEL(hpmc_emit_volume4_gradient, HPMC_VOLUME4_FIELD, hpmc::kEmitGradient, false)
HPMC_DEFINE_EMIT_KERNEL(hpmc_emit_volume4_binary, HPMC_VOLUME4_FIELD, hpmc::kEmitBinary, false)
HPMC_DEFINE_EMIT_KERNEL(hpmc_emit_volume8, HPMC_VOLUME8_FIELD, hpmc::kEmitContinuous, false)
HPMC_DEFINE_EMIT_KERNEL(hpmc_emit_volume8_binary, HPMC_VOLUME8_FIELD, hpmc::kEmitBinary, false)
HPMC_DEFINE_EMIT_KERNEL(hpmc_emit_volume_batch, HPMC_VOLUME_FIELD, hpmc::kEmitContinuous, true)

extern "C" __global__ __launch_bounds__(hpmc::kScanThreads) void hpmc_scan(
    const hpmc::VolDesc* vols, const uint32_t* tile_vol, uint32_t n_l4t, uint32_t n_colt, const uint16_t* colcnt,
    const uint2* tile_sums, uint32_t* item_reps, uint32_t* l4_off, uint2* items, hpmc::Totals* totals,
    hpmc::Totals* host_totals, hpmc::VolCount* vol_counts, hpmc::VolCount* host_vol_counts)
{
    hpmc::scan_body(vols, tile_vol, n_l4t, n_colt, colcnt, tile_sums, item_reps, l4_off, items, totals,
                    host_totals, vol_counts, host_vol_counts);
}

extern "C" __global__ __launch_bounds__(256) void hpmc_level4_gather(const hpmc::VolDesc* vols, uint32_t vol,
                                                                     const uint16_t* colcnt, uint32_t* out)
{
    const hpmc::VolDesc d = vols[vol];
    hpmc::level4_body(d, colcnt, out);
}

// Adjoint passes (hpmc_device.h, adjoint_cells_body and following).
extern "C" __global__ __launch_bounds__(256) void hpmc_adjoint_cells(const hpmc::VolDesc* vols, float T,
                                                                     const uint2* items, const uint32_t* l4_off,
                                                                     const hpmc::Totals* totals,
                                                                     const uint64_t* cases, const uint32_t* tile0,
                                                                     uint8_t* codes, uint32_t* cell_off,
                                                                     uint8_t* flags)
{
    __shared__ uint8_t lut[256];
    lut[threadIdx.x] = (uint8_t)(cases[threadIdx.x] & 0xfu);
    __syncthreads();
    hpmc::adjoint_cells_body(vols, T, items, l4_off, totals, lut, tile0, codes, cell_off, flags);
}

extern "C" __global__ __launch_bounds__(256) void hpmc_adjoint_records(const hpmc::VolDesc* vols, float T,
                                                                       const uint2* items,
                                                                       const hpmc::Totals* totals,
                                                                       const uint64_t* cases, const uint8_t* codes,
                                                                       const uint32_t* cell_off, const float* gout,
                                                                       uint32_t nvert, float* rec, float* partials)
{
    __shared__ uint16_t smask[256][12];
    __shared__ uint16_t lists[hpmc::kAdjointWaves][hpmc::kEdgeListMax];
    hpmc::fill_slot_masks(cases, smask);
    __syncthreads();
    hpmc::adjoint_records_body(vols, T, items, totals, smask, codes, cell_off, gout, nvert, rec, partials,
                               lists[threadIdx.x >> 6]);
}

extern "C" __global__ __launch_bounds__(1024) void hpmc_adjoint_tile_list(const uint8_t* flags, uint32_t ntiles,
                                                                           uint32_t* list, uint32_t* count)
{
    hpmc::adjoint_tile_list_body(flags, ntiles, list, count);
}

extern "C" __global__ __launch_bounds__(256) void hpmc_adjoint_gather(const hpmc::VolDesc* vols,
                                                                      const uint32_t* tile_vol,
                                                                      const uint32_t* tile0, const uint32_t* list,
                                                                      const uint32_t* count, const uint8_t* codes,
                                                                      const uint32_t* cell_off, const float* rec,
                                                                      uint32_t nvert, float* grad)
{
    __shared__ uint32_t tab[3][hpmc::kGatherOrigins];
    hpmc::adjoint_gather_body(vols, tile_vol, tile0, list, count, codes, cell_off, rec, nvert, grad, tab);
}

extern "C" __global__ __launch_bounds__(1024) void hpmc_adjoint_sum_items(const float* partials,
                                                                          const hpmc::Totals* totals, float* out)
{
    hpmc::adjoint_sum_items_body(partials, totals, out);
}

extern "C" __global__ __launch_bounds__(256) void hpmc_fill_u32(uint32_t* p, uint64_t n, uint32_t v)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

extern "C" __global__ __launch_bounds__(256) void hpmc_fill_f32(float* p, uint64_t n, float v)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Start of a no-allocation build: the volume's descriptor into its scratch
// slot (passed by value, so a captured graph carries it) and the scan tile
// sums cleared.
extern "C" __global__ __launch_bounds__(256) void hpmc_noalloc_begin(hpmc::VolDesc d, hpmc::VolDesc* dst,
                                                                     uint2* tile_sums, uint32_t ntiles)
{
    if (threadIdx.x == 0) *dst = d;
    for (uint32_t i = threadIdx.x; i < ntiles; i += blockDim.x) tile_sums[i] = make_uint2(0u, 0u);
}

// Clears the scan tile sums before a build that follows no extraction.
extern "C" __global__ __launch_bounds__(256) void hpmc_clear_sums(uint2* tile_sums, uint32_t ntiles)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ntiles; i += gridDim.x * blockDim.x)
        tile_sums[i] = make_uint2(0u, 0u);
}

#include "hpmc_plan.h"

namespace hpmc_host {

// Before a build's classification: the extraction of the previous build
// cleared the scan tile sums that classification adds into; clear them here
// when no extraction followed it.  In a graph being captured (which may be
// replayed without an extraction in between) the clear is part of the graph
// unless the capture also holds the extraction (whose replays clear them,
// the graph's launch clearing them first if need be).  Marks the sums as the
// new build's.
hipError_t sums_prepare(Plan& p, hipStream_t stream, bool extract_in_capture)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    hipError_t e = hipStreamIsCapturing(stream, &cs);
    if (e != hipSuccess) return e;
    const bool capturing = cs != hipStreamCaptureStatusNone;
    if (capturing ? !extract_in_capture : p.sums_dirty) {
        const uint32_t n = p.n_l4t + p.n_colt;
        hipLaunchKernelGGL(hpmc_clear_sums, dim3(std::min<uint32_t>((n + 255) / 256, 1024u)), dim3(256), 0, stream,
                           p.tile_sums, n);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    p.sums_dirty = true;
    return hipSuccess;
}

// Scratch of the adjoint, grown on demand: per lattice sample a cell code
// and a first vertex (cells by their origin sample), followed by one flag
// per gather tile; per vertex room for one edge record; per work item a
// dL/dT partial sum.
struct AdjointScratch {
    uint64_t cap_vertices = 0, cap_lattice = 0, cap_items = 0, cap_tiles = 0;
    uint8_t* codes = nullptr;
    uint32_t* cell_off = nullptr;
    float* rec = nullptr;
    float* partials = nullptr;
    uint32_t* tiles = nullptr;      // flagged gather tiles, then their count
};

inline void adjoint_free(AdjointScratch& s)
{
    if (s.codes) (void)hipFree(s.codes);
    if (s.cell_off) (void)hipFree(s.cell_off);
    if (s.rec) (void)hipFree(s.rec);
    if (s.partials) (void)hipFree(s.partials);
    if (s.tiles) (void)hipFree(s.tiles);
    s = AdjointScratch();
}

#define ADJ_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess) {                                                            \
            *err = std::string(#expr) + ": " + hipGetErrorString(_e);                      \
            return HPMC_ERR_HIP;                                                           \
        }                                                                                  \
    } while (0)

// Workgroups of `threads` threads of kernel k resident on the current
// device at once (occupancy x compute units), at least 1.
inline unsigned resident_grid(const void* k, int threads)
{
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k, threads, 0) != hipSuccess)
        return 1024;
    return (unsigned)std::max(cus * per, 1);
}

// dL/dfield (n_lattice floats, every sample written) and dL/dT (one float)
// of the last extraction of plan p (nvert vertices at threshold T), from
// dL/dvertices gout (nvert x 6, normal then position).  Deterministic, with
// no float atomics and no sorting: per-cell codes and first vertices from
// the HistoPyramid's blocks, one record per sign-changing lattice edge
// formed by the cell that owns it, a per-sample gather of the records in a
// fixed order, and a fixed-shape sum of the per-item dL/dT terms.
inline int32_t adjoint_run(const Plan& p, AdjointScratch& s, float T, uint32_t nvert, const float* gout,
                           float* grad_field, float* grad_iso, const uint64_t* cases, hipStream_t stream,
                           std::string* err)
{
    // 32-bit byte offsets into the records and a volume's first-vertex array
    if (nvert >= (1u << 27)) {
        *err = "too many vertices for the adjoint (at most 2^27 - 1)";
        return HPMC_ERR_SIZE;
    }
    for (const hpmc::VolDesc& d : p.vols) {
        if ((uint64_t)d.g.sx * d.g.sy * d.g.sz > (1ull << 30)) {
            *err = "volume too large for the adjoint (at most 2^30 samples)";
            return HPMC_ERR_SIZE;
        }
    }
    // 32-bit sample keys into the concatenated lattice of a batch
    if (p.n_lattice > 0xffffffffull) {
        *err = "batch too large for the adjoint (at most 2^32 - 1 samples in all)";
        return HPMC_ERR_SIZE;
    }
    const uint64_t need_v = std::max<uint64_t>(nvert, 1), need_l = std::max<uint64_t>(p.n_lattice, 1);
    const uint64_t need_i = std::max<uint64_t>(p.n_items_max, 1), need_t = std::max<uint64_t>(p.n_gather_tiles, 1);
    if (need_v > s.cap_vertices || need_l > s.cap_lattice || need_i > s.cap_items || need_t > s.cap_tiles) {
        ADJ_TRY(hipStreamSynchronize(stream));
        const uint64_t cv = std::max(need_v, s.cap_vertices), cl = std::max(need_l, s.cap_lattice);
        const uint64_t ci = std::max(need_i, s.cap_items), ct = std::max(need_t, s.cap_tiles);
        adjoint_free(s);
        ADJ_TRY(hipMalloc(&s.codes, (cl + 15) / 16 * 16 + (ct + 15) / 16 * 16));
        ADJ_TRY(hipMalloc(&s.cell_off, cl * sizeof(uint32_t)));
        ADJ_TRY(hipMalloc(&s.rec, cv * hpmc::kAdjointRecordFloats * sizeof(float)));
        ADJ_TRY(hipMalloc(&s.partials, ci * sizeof(float)));
        ADJ_TRY(hipMalloc(&s.tiles, (ct + 1) * sizeof(uint32_t)));
        s.cap_vertices = cv;
        s.cap_lattice = cl;
        s.cap_items = ci;
        s.cap_tiles = ct;
    }
    if (nvert == 0) {
        hipLaunchKernelGGL(hpmc_fill_f32, dim3(1024), dim3(256), 0, stream, grad_field, (uint64_t)p.n_lattice, 0.0f);
        hipLaunchKernelGGL(hpmc_fill_f32, dim3(1), dim3(64), 0, stream, grad_iso, (uint64_t)1, 0.0f);
        ADJ_TRY(hipGetLastError());
        return HPMC_OK;
    }
    // codes (n_lattice bytes) and the gather tile flags after them, each
    // rounded up to 16 bytes, cleared together
    uint8_t* flags = s.codes + (p.n_lattice + 15) / 16 * 16;
    const uint64_t clear_words = (p.n_lattice + 15) / 16 * 4 + (p.n_gather_tiles + 15) / 16 * 4;
    hipLaunchKernelGGL(hpmc_fill_u32, dim3((unsigned)std::min<uint64_t>((clear_words + 255) / 256, 2048)), dim3(256), 0,
                       stream, (uint32_t*)s.codes, clear_words, 0u);
    const unsigned item_grid =
        (unsigned)std::min<uint64_t>((p.n_items_max + hpmc::kAdjointWaves - 1) / hpmc::kAdjointWaves, 4096);
    hipLaunchKernelGGL(hpmc_adjoint_cells, dim3(item_grid), dim3(64 * hpmc::kAdjointWaves), 0, stream,
                       (const hpmc::VolDesc*)p.d_vols, T, (const uint2*)p.items, (const uint32_t*)p.l4_off,
                       (const hpmc::Totals*)p.d_totals, cases, (const uint32_t*)p.d_gather_tile0, s.codes, s.cell_off,
                       flags);
    hipLaunchKernelGGL(hpmc_adjoint_records, dim3(item_grid), dim3(64 * hpmc::kAdjointWaves), 0, stream,
                       (const hpmc::VolDesc*)p.d_vols, T, (const uint2*)p.items, (const hpmc::Totals*)p.d_totals,
                       cases, (const uint8_t*)s.codes, (const uint32_t*)s.cell_off, gout, nvert, s.rec, s.partials);
    // dL/dfield: zero, then the flagged tiles gathered
    hipLaunchKernelGGL(hpmc_fill_f32, dim3(4096), dim3(256), 0, stream, grad_field, (uint64_t)p.n_lattice, 0.0f);
    uint32_t* count = s.tiles + p.n_gather_tiles;
    hipLaunchKernelGGL(hpmc_adjoint_tile_list, dim3(1), dim3(1024), 0, stream, (const uint8_t*)flags,
                       (uint32_t)p.n_gather_tiles, s.tiles, count);
    hipLaunchKernelGGL(hpmc_adjoint_gather,
                       dim3((unsigned)std::min<uint64_t>(p.n_gather_tiles,
                                                         resident_grid((const void*)hpmc_adjoint_gather,
                                                                       hpmc::kGatherThreads))),
                       dim3(hpmc::kGatherThreads), 0, stream, (const hpmc::VolDesc*)p.d_vols,
                       (const uint32_t*)p.d_gather_vol, (const uint32_t*)p.d_gather_tile0, (const uint32_t*)s.tiles,
                       (const uint32_t*)count, (const uint8_t*)s.codes, (const uint32_t*)s.cell_off,
                       (const float*)s.rec, nvert, grad_field);
    hipLaunchKernelGGL(hpmc_adjoint_sum_items, dim3(1), dim3(1024), 0, stream, (const float*)s.partials,
                       (const hpmc::Totals*)p.d_totals, grad_iso);
    ADJ_TRY(hipGetLastError());
    return HPMC_OK;
}

}  // namespace hpmc_host

namespace {

const uint64_t kCases[256] = HPMC_MC_CASES_INIT;

#pragma clang fp contract(off)
// Per HPMC code and slot, the normal a binary field's vertex carries
// (constants.cpp:217-273 builds it; shadergen.cpp:507-510 decodes it): the
// face normal n = (m2 - m0) x (m1 - m0) of the slot's triangle over the edge
// midpoints m0..m2, scaled by 0.5 |n|, offset by 0.5, clamped into (0, 1),
// added to the slot's integer edge shift and decoded as 2 fract(.) - 1.
// FP32 throughout, in the reference's operation order.
void binary_normals(float4 (&out)[256 * 16])
{
    for (int c = 0; c < 256; ++c) {
        const int count = (int)(kCases[c] & 0xfu);
        float mid[15][3], shift[15][3];
        for (int i = 0; i < count; ++i) {
            const uint32_t e = (uint32_t)(kCases[c] >> (4 + 4 * i)) & 0xfu;
            const int axis = (int)(e >> 2), s0 = (int)(e & 1u), s1 = (int)((e >> 1) & 1u);
            float sh[3] = {0.0f, 0.0f, 0.0f};
            if (axis == 0) { sh[1] = (float)s0; sh[2] = (float)s1; }
            else if (axis == 1) { sh[0] = (float)s0; sh[2] = (float)s1; }
            else { sh[0] = (float)s0; sh[1] = (float)s1; }
            for (int k = 0; k < 3; ++k) {
                shift[i][k] = sh[k];
                mid[i][k] = sh[k] + (k == axis ? 0.5f : 0.0f);
            }
        }
        for (int i = 0; i < 16; ++i) out[c * 16 + i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        for (int i = 0; i < count; ++i) {
            const int t = i / 3;
            const float* m0 = mid[3 * t];
            const float* m1 = mid[3 * t + 1];
            const float* m2 = mid[3 * t + 2];
            float u[3], v[3];
            for (int k = 0; k < 3; ++k) {
                u[k] = m2[k] - m0[k];
                v[k] = m1[k] - m0[k];
            }
            float n[3];
            n[0] = u[1] * v[2] - u[2] * v[1];
            n[1] = u[2] * v[0] - u[0] * v[2];
            n[2] = u[0] * v[1] - u[1] * v[0];
            const float s = 0.5f * sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            float r[3];
            for (int k = 0; k < 3; ++k) {
                float tmp = s * n[k] + 0.5f;
                if (tmp < 0.0f) tmp = 0.000001f;
                if (tmp >= 1.0f) tmp = 0.999999f;
                const float packed = shift[i][k] + tmp;
                r[k] = 2.0f * (packed - floorf(packed)) - 1.0f;
            }
            out[c * 16 + i] = make_float4(r[0], r[1], r[2], 0.0f);
        }
    }
}
#pragma clang fp contract(on)

enum FieldMode { kFieldVolume = 0, kFieldCustom = 1 };

thread_local int32_t t_last_error = HPMC_OK;
thread_local std::string t_last_msg;

}  // namespace

struct HPMCConstants {
    int device = 0;
    uint64_t* d_cases = nullptr;
    float4* d_bnorm = nullptr;   // binary-field normals by (code, slot)
};

struct HPMCHistoPyramid {
    HPMCConstants* constants = nullptr;
    bool tainted = true;
    bool broken = true;
    int32_t size[3] = {0, 0, 0};
    int32_t cells[3] = {0, 0, 0};
    float extent[3] = {1.0f, 1.0f, 1.0f};
    bool binary = false;             // as last set
    bool binary_active = false;      // as taken at the last configuration (setup)
    int mode = kFieldVolume;
    const float* volume = nullptr;
    bool unorm8 = false;             // the volume is 8-bit normalised (volume8)
    const uint8_t* volume8 = nullptr;
    bool gradient = false;
    std::string snippet;
    // layout (init.cpp:105-130)
    int32_t tile[2] = {0, 0};
    int32_t tiles[2] = {0, 0};
    int32_t l2 = 0;
    // device state of the build (one volume)
    hpmc_host::Plan plan;
    hpmc_host::Kernel k_classify, k_emit, k_codes;
    hpmc_host::Kernel k_classify_small;   // grids with blocks narrower than 16 cells
    hipEvent_t ev_totals = nullptr;  // recorded by a count: the totals are in pinned memory once it completes
    hpmc_host::AdjointScratch adjoint;
    bool built = false;
    bool count_valid = false;
    uint32_t count = 0;
    float threshold = 0.0f;
    hipStream_t stream = nullptr;
    hpmc_host::Tuning tuning;        // launch shapes (0 = automatic)
    // hipRTC module for snippets
    hipModule_t module = nullptr;
    uint32_t builder_id = 0;
    // optional per-stage timing: an event pair per launch, recorded by the
    // launch at the kernel's start and end; finished pairs are folded into
    // the totals once a stage holds kStagePairsMax of them, so profiling
    // left on holds a bounded number
    hipEvent_t ev_order = nullptr;   // HPMCsetStream: old stream -> new stream
    bool profiling = false;
    std::vector<hipEvent_t> stage_events[3];
    double stage_ms[3] = {0.0, 0.0, 0.0};
    uint32_t stage_n[3] = {0, 0, 0};
    // errors
    int32_t last_error = HPMC_OK;
    std::string last_msg;
};

struct HPMCTraversalHandle {
    HPMCHistoPyramid* h = nullptr;
    uint32_t program = 0;
    float* out = nullptr;
    uint64_t capacity = 0;
    hpmc::CellEdge* out_ce = nullptr;
    uint64_t ce_capacity = 0;
};

namespace {

int32_t fail(HPMCHistoPyramid* h, int32_t code, const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    t_last_error = code;
    t_last_msg = buf;
    if (h) {
        h->last_error = code;
        h->last_msg = buf;
    }
    return code;
}

void ok(HPMCHistoPyramid* h)
{
    t_last_error = HPMC_OK;
    t_last_msg.clear();
    if (h) {
        h->last_error = HPMC_OK;
        h->last_msg.clear();
    }
}

#define HIP_TRY(h, expr)                                                                   \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess) {                                                            \
            fail((h), HPMC_ERR_HIP, "%s: %s", #expr, hipGetErrorString(_e));               \
            return false;                                                                  \
        }                                                                                  \
    } while (0)

void free_device(HPMCHistoPyramid* h)
{
    hpmc_host::plan_free(h->plan);
    hpmc_host::adjoint_free(h->adjoint);
    if (h->ev_totals) (void)hipEventDestroy(h->ev_totals);
    if (h->module) (void)hipModuleUnload(h->module);
    h->ev_totals = nullptr;
    h->module = nullptr;
    h->k_classify = h->k_emit = h->k_codes = h->k_classify_small = hpmc_host::Kernel();
    h->built = false;
    h->count_valid = false;
}

// hpmc/src/init.cpp:72-149 (hpmc_host::layout_of).
bool determine_layout(HPMCHistoPyramid* h)
{
    hpmc_host::Layout lay;
    std::string err;
    const int32_t rc = hpmc_host::layout_of(h->size, h->cells, &lay, &err);
    if (rc != HPMC_OK) {
        fail(h, rc, "%s", err.c_str());
        return false;
    }
    h->tile[0] = lay.tile[0];
    h->tile[1] = lay.tile[1];
    h->tiles[0] = lay.tiles[0];
    h->tiles[1] = lay.tiles[1];
    h->l2 = lay.l2;
    return true;
}

// Device pointer of a volume field (FP32, RGBA FP32 or 8-bit).
const void* volume_ptr(const HPMCHistoPyramid* h)
{
    return h->unorm8 ? (const void*)h->volume8 : (const void*)h->volume;
}

// The classification threshold of an 8-bit volume: the smallest byte c with
// c / 255 >= T (256 if none), so that c < kT exactly when c / 255 < T.
float unorm8_threshold(float T)
{
    for (int c = 0; c < 256; ++c)
        if (!((float)c / 255.0f < T)) return (float)c;
    return 256.0f;
}

// The snippet and the field object built on it (hpmc::SnippetField), for
// the hipRTC sources that need them.
std::string snippet_field_source(const std::string& snippet)
{
    std::string s;
    s += "#line 1 \"field_snippet\"\n";
    s += snippet;
    s += "\n#line 1 \"hpmc_rtc_tail\"\n";
    s += "namespace hpmc {\n"
         "struct SnippetField {\n"
         "    float ipx, ipy, ipz;\n"
         "    __device__ __forceinline__ float at(int i, int j, int k) const {\n"
         "        return HPMC_fetch(make_float3(((float)i + 0.5f) * ipx, ((float)j + 0.5f) * ipy,\n"
         "                                      ((float)k + 0.5f) * ipz));\n"
         "    }\n"
         "    typedef float Plane;\n"
         "    __device__ __forceinline__ uint32_t lane_index(int i, int j) const {\n"
         "        return (uint32_t)i | ((uint32_t)j << 16);\n"
         "    }\n"
         "    __device__ __forceinline__ Plane plane(int k) const { return ((float)k + 0.5f) * ipz; }\n"
         "    __device__ __forceinline__ float fetch(Plane pz, uint32_t idx) const {\n"
         "        return HPMC_fetch(make_float3(((float)(idx & 0xffffu) + 0.5f) * ipx,\n"
         "                                      ((float)(idx >> 16) + 0.5f) * ipy, pz));\n"
         "    }\n"
         "    __device__ __forceinline__ uint32_t col_index(int i) const { return (uint32_t)i; }\n"
         "    __device__ __forceinline__ uint32_t row_index(int j) const { return (uint32_t)j; }\n"
         "    __device__ __forceinline__ float fetch2(Plane pz, uint32_t col, uint32_t row) const {\n"
         "        return HPMC_fetch(make_float3(((float)col + 0.5f) * ipx, ((float)row + 0.5f) * ipy, pz));\n"
         "    }\n"
         "    __device__ __forceinline__ float sample(Plane pz, uint32_t idx) const { return fetch(pz, idx); }\n"
         "#ifdef HPMC_FIELD_GRADIENT\n"
         "    __device__ __forceinline__ float4 grad(int i, int j, int k) const {\n"
         "        return HPMC_fetchGrad(make_float3(((float)i + 0.5f) * ipx, ((float)j + 0.5f) * ipy,\n"
         "                                          ((float)k + 0.5f) * ipz));\n"
         "    }\n"
         "#endif\n"
         "};\n"
         "}\n";
    return s;
}

std::string rtc_source(const std::string& snippet, bool gradient, bool binary)
{
    std::string s;
    s += "#define HPMC_RTC 1\n";
    if (gradient) s += "#define HPMC_FIELD_GRADIENT 1\n";
    s += "typedef unsigned char uint8_t;\ntypedef unsigned short uint16_t;\n";
    s += "typedef unsigned int uint32_t;\ntypedef unsigned long long uint64_t;\n";
    s += "#include \"hpmc_device.h\"\n";
    s += snippet_field_source(snippet);
    s += "HPMC_DEFINE_BUILD_KERNELS(rtc, (hpmc::SnippetField{dd.g.ipx, dd.g.ipy, dd.g.ipz}))\n";
    const char* mode = binary ? "hpmc::kEmitBinary" : gradient ? "hpmc::kEmitGradient" : "hpmc::kEmitContinuous";
    s += std::string("HPMC_DEFINE_EMIT_KERNEL(hpmc_emit_rtc, (hpmc::SnippetField{dd.g.ipx, dd.g.ipy, dd.g.ipz}), ") +
         mode + ", false)\n";
    return s;
}

std::mutex g_rtc_mutex;
uint32_t g_builder_ids = 0;

// Compiles the snippet behind the device header; returns the code object
// or an empty vector with the compiler log in *log.  Needs no device.
std::vector<char> rtc_compile(const std::string& snippet, bool gradient, bool binary, std::string* log)
{
    std::lock_guard<std::mutex> lock(g_rtc_mutex);
    std::string src = rtc_source(snippet, gradient, binary);
    const char* hdr_src[2] = {kEmbeddedDeviceHeader, kEmbeddedTablesHeader};
    const char* hdr_names[2] = {"hpmc_device.h", "mc_tables.h"};
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "hpmc_field.hip", 2, hdr_src, hdr_names) != HIPRTC_SUCCESS) {
        *log = "hiprtcCreateProgram failed";
        return {};
    }
    const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
                          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
    hiprtcResult r = hiprtcCompileProgram(prog, (int)(sizeof opts / sizeof opts[0]), opts);
    size_t n = 0;
    hiprtcGetProgramLogSize(prog, &n);
    std::string text(n + 1, '\0');
    hiprtcGetProgramLog(prog, &text[0]);
    text.resize(strlen(text.c_str()));
    *log = text;
    if (r != HIPRTC_SUCCESS) {
        hiprtcDestroyProgram(&prog);
        return {};
    }
    size_t code_size = 0;
    hiprtcGetCodeSize(prog, &code_size);
    std::vector<char> code(code_size);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    return code;
}

bool compile_snippet(HPMCHistoPyramid* h)
{
    std::string log;
    std::vector<char> code = rtc_compile(h->snippet, h->gradient, h->binary_active, &log);
    if (code.empty()) {
        fail(h, HPMC_ERR_COMPILE, "field snippet failed to compile: %s", log.c_str());
        return false;
    }
    HIP_TRY(h, hipModuleLoadData(&h->module, code.data()));
    HIP_TRY(h, hipModuleGetFunction(&h->k_classify.rtc, h->module, "hpmc_classify_rtc"));
    HIP_TRY(h, hipModuleGetFunction(&h->k_emit.rtc, h->module, "hpmc_emit_rtc"));
    HIP_TRY(h, hipModuleGetFunction(&h->k_codes.rtc, h->module, "hpmc_codes_rtc"));
    HIP_TRY(h, hipModuleGetFunction(&h->k_classify_small.rtc, h->module, "hpmc_classify_small_rtc"));
    return true;
}

// The device's share of the classification and emit launches (resident
// workgroups per CU), probed once per process.
const hpmc_host::DeviceShape& device_shape()
{
    static hpmc_host::DeviceShape shape;
    static std::once_flag once;
    std::call_once(once, [] {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) return;
        (void)hipDeviceGetAttribute(&shape.cus, hipDeviceAttributeMultiprocessorCount, dev);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&shape.emit_per_cu, (const void*)hpmc_emit_volume, 256, 0);
    });
    return shape;
}

// ---------------------------------------------------------------------------
// Scratch of the no-allocation entry point (HPMCextractVolumeNoAlloc): the
// device arrays of a one-volume plan carved out of one file-scope device
// array.  Each device holds its own copy (allocated by the runtime when it
// loads this library's code on that device, kept until the process ends);
// in host code the name is only a handle, the device address comes from
// hipGetSymbolAddress, once per device.  Every call on a device shares that
// device's copy: calls that overlap on one device race.
constexpr uint64_t kNoAllocScratchBytes = 80ull << 20;
static_assert(kNoAllocScratchBytes == HPMC_NO_ALLOC_SCRATCH_BYTES, "header and library agree");
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
struct ScratchAddress {
    std::once_flag once;
    void* ptr = nullptr;
    hipError_t status = hipErrorNotInitialized;
};
ScratchAddress g_scratch_address[kMaxDevices];

// The address of `scratch` on device `dev`, which must be the current
// device; looked up on the first call for that device, null on failure
// (*status says why).
void* scratch_address(int dev, hipError_t* status)
{
    if (dev < 0 || dev >= kMaxDevices) {
        *status = hipErrorInvalidDevice;
        return nullptr;
    }
    ScratchAddress& a = g_scratch_address[dev];
    std::call_once(a.once, [&a] {
        a.status = hipGetSymbolAddress(&a.ptr, HIP_SYMBOL(scratch));
        if (a.status != hipSuccess) a.ptr = nullptr;
    });
    *status = a.status;
    return a.ptr;
}

// Uploads the case and binary-normal tables once; false without a device.
bool upload_tables(HPMCConstants* c)
{
    if (c->d_cases) return true;
    static float4 bnorm[256 * 16];
    static std::once_flag bnorm_once;
    std::call_once(bnorm_once, [] { binary_normals(bnorm); });
    uint64_t* cases = nullptr;
    float4* bn = nullptr;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipMalloc(&cases, sizeof kCases) != hipSuccess ||
        hipMemcpy(cases, kCases, sizeof kCases, hipMemcpyHostToDevice) != hipSuccess ||
        hipMalloc(&bn, sizeof bnorm) != hipSuccess ||
        hipMemcpy(bn, bnorm, sizeof bnorm, hipMemcpyHostToDevice) != hipSuccess) {
        if (cases) (void)hipFree(cases);
        if (bn) (void)hipFree(bn);
        (void)hipGetLastError();
        return false;
    }
    c->device = dev;
    c->d_cases = cases;
    c->d_bnorm = bn;
    // this device's copy of the no-allocation entry point's scratch array
    // is looked up now, at set-up, with the device current
    hipError_t st;
    (void)scratch_address(dev, &st);
    return true;
}

// hpmc/src/init.cpp:48-69: layout, storage and builder programs, once per
// configuration.  Classification workgroups span whole lattice rows
// (ceil(X/64) waves, at most 16) over zseg slices, zseg chosen so the
// workgroups fill the device about once.
bool setup(HPMCHistoPyramid* h)
{
    if (!h->tainted) return true;
    if (!upload_tables(h->constants)) {
        fail(h, HPMC_ERR_HIP, "no HIP device for the marching-cubes tables");
        return false;
    }
    free_device(h);
    if (!determine_layout(h)) return false;
    if (h->mode == kFieldVolume && !volume_ptr(h)) {
        fail(h, HPMC_ERR_FIELD, "no field volume set");
        return false;
    }
    if (h->mode == kFieldCustom && h->snippet.empty()) {
        fail(h, HPMC_ERR_FIELD, "empty field snippet");
        return false;
    }
    // hpiface.cpp:131-142: the binary flag does not taint; it takes effect
    // when the programs are next generated, i.e. here
    h->binary_active = h->binary;
    if (h->mode == kFieldVolume) {
        if (h->unorm8) {
            h->k_classify.builtin = (const void*)hpmc_classify_volume8;
            h->k_classify_small.builtin = (const void*)hpmc_classify_small_volume8;
            h->k_codes.builtin = (const void*)hpmc_codes_volume8;
            h->k_emit.builtin = h->binary_active ? (const void*)hpmc_emit_volume8_binary
                                                 : (const void*)hpmc_emit_volume8;
        } else if (h->gradient) {
            h->k_classify.builtin = (const void*)hpmc_classify_volume4;
            h->k_classify_small.builtin = (const void*)hpmc_classify_small_volume4;
            h->k_codes.builtin = (const void*)hpmc_codes_volume4;
            h->k_emit.builtin = h->binary_active ? (const void*)hpmc_emit_volume4_binary
                                                 : (const void*)hpmc_emit_volume4_gradient;
        } else {
            // lattices past the 256 MiB MALL stream through classification
            // with non-temporal loads (VolumeStreamField)
            const uint64_t bytes = 4ull * (uint64_t)h->size[0] * (uint64_t)h->size[1] * (uint64_t)h->size[2];
            h->k_classify.builtin = bytes >= (256ull << 20) ? (const void*)hpmc_classify_volume_stream
                                                            : (const void*)hpmc_classify_volume;
            h->k_classify_small.builtin = (const void*)hpmc_classify_small_volume;
            h->k_codes.builtin = (const void*)hpmc_codes_volume;
            h->k_emit.builtin = h->binary_active ? (const void*)hpmc_emit_volume_binary
                                                 : (const void*)hpmc_emit_volume;
        }
    }
    hpmc_host::VolumeSpec spec;
    for (int a = 0; a < 3; ++a) {
        spec.size[a] = h->size[a];
        spec.cells[a] = h->cells[a];
        spec.extent[a] = h->extent[a];
    }
    spec.field = h->mode == kFieldVolume ? volume_ptr(h) : nullptr;
    std::string err;
    const bool fp32_volume = h->mode == kFieldVolume && !h->unorm8 && !h->gradient;
    const int32_t rc = hpmc_host::plan_create(h->plan, {spec}, h->tuning, (const void*)hpmc_classify_volume,
                                              device_shape(), fp32_volume, &err);
    if (rc != HPMC_OK) {
        fail(h, rc, "%s", err.c_str());
        return false;
    }
    HIP_TRY(h, hipEventCreateWithFlags(&h->ev_totals, hipEventDisableTiming));
    if (h->mode == kFieldCustom) {
        if (!compile_snippet(h)) return false;
    }
    h->builder_id = ++g_builder_ids;
    h->tainted = false;
    return true;
}

enum Stage { kStageClassify = 0, kStageScan = 1, kStageEmit = 2 };

constexpr size_t kStagePairsMax = 1024;

// Adds the elapsed times of a stage's recorded event pairs to its totals and
// destroys the events (waits for the last pair).  False if an event failed.
bool stage_fold(HPMCHistoPyramid* h, int st)
{
    std::vector<hipEvent_t>& ev = h->stage_events[st];
    bool good = true;
    for (size_t i = 0; i + 1 < ev.size(); i += 2) {
        if (hipEventSynchronize(ev[i + 1]) != hipSuccess) {
            good = false;
            continue;
        }
        float t = 0.0f;
        if (hipEventElapsedTime(&t, ev[i], ev[i + 1]) == hipSuccess) {
            h->stage_ms[st] += t;
            ++h->stage_n[st];
        }
    }
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    ev.clear();
    return good;
}

// The event pair of a stage's next launch when profiling (else both null):
// the launch records them at the kernel's start and end.
void stage_events(HPMCHistoPyramid* h, int stage, hipEvent_t& e0, hipEvent_t& e1)
{
    e0 = e1 = nullptr;
    if (!h->profiling) return;
    std::vector<hipEvent_t>& ev = h->stage_events[stage];
    // the stage full: fold the finished pairs first
    if (ev.size() >= 2 * kStagePairsMax) (void)stage_fold(h, stage);
    if (hipEventCreate(&e0) != hipSuccess) {
        e0 = nullptr;
        return;
    }
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        e0 = e1 = nullptr;
        return;
    }
    ev.push_back(e0);
    ev.push_back(e1);
}

bool launch_classify(HPMCHistoPyramid* h, float T)
{
    if (h->mode == kFieldVolume) {
        // a volume reset to NULL after the configuration: no launch
        if (!volume_ptr(h)) {
            fail(h, HPMC_ERR_FIELD, "no field volume set");
            return false;
        }
        HIP_TRY(h, hpmc_host::plan_set_field(h->plan, 0, volume_ptr(h), h->stream));
    }
    HIP_TRY(h, hpmc_host::sums_prepare(h->plan, h->stream, false));
    hipEvent_t e0, e1;
    stage_events(h, kStageClassify, e0, e1);
    // 8-bit volumes classify raw bytes against the equivalent byte threshold
    const float Tc = h->mode == kFieldVolume && h->unorm8 ? unorm8_threshold(T) : T;
    if (hpmc_host::plan_small_blocks(h->plan, h->plan.vols[0]))
        HIP_TRY(h, hpmc_host::plan_classify_small(h->plan, h->k_classify_small, Tc, h->constants->d_cases,
                                                  h->stream, e0, e1));
    else
        HIP_TRY(h, hpmc_host::plan_classify(h->plan, h->k_classify, Tc, h->constants->d_cases, h->stream, e0, e1));
    HIP_TRY(h, hipGetLastError());
    return true;
}

bool launch_scan(HPMCHistoPyramid* h)
{
    hipEvent_t e0, e1;
    stage_events(h, kStageScan, e0, e1);
    HIP_TRY(h, hpmc_host::plan_scan(h->plan, (const void*)hpmc_scan, h->stream, e0, e1));
    HIP_TRY(h, hipGetLastError());
    return true;
}

bool launch_emit(HPMCHistoPyramid* h, float* out, hpmc::CellEdge* out_ce, uint32_t capacity)
{
    hipEvent_t e0, e1;
    stage_events(h, kStageEmit, e0, e1);
    HIP_TRY(h, hpmc_host::plan_emit(h->plan, h->k_emit, h->threshold, out, out_ce, capacity, h->constants->d_cases,
                                    h->constants->d_bnorm, h->stream, e0, e1));
    HIP_TRY(h, hipGetLastError());
    h->plan.sums_dirty = false;   // the extraction clears the tile sums
    return true;
}

bool wait_count(HPMCHistoPyramid* h)
{
    if (h->count_valid) return true;
    if (!h->built) {
        h->count = 0;
        return true;
    }
    // everything queued so far on the stream, the last build's scan included
    HIP_TRY(h, hipEventRecord(h->ev_totals, h->stream));
    HIP_TRY(h, hipEventSynchronize(h->ev_totals));
    h->count = h->plan.h_totals->vertices;
    h->count_valid = true;
    return true;
}

bool extract_helper(HPMCTraversalHandle* th)
{
    if (!th) {
        fail(nullptr, HPMC_ERR_NULL, "NULL traversal handle");
        return false;
    }
    HPMCHistoPyramid* h = th->h;
    if (!h || h->broken) {
        fail(h, HPMC_ERR_BROKEN, "HistoPyramid is broken");
        return false;
    }
    if (h->tainted || !h->built) {
        fail(h, HPMC_ERR_ARG, "HistoPyramid has not been built since it was configured");
        return false;
    }
    if (!th->out) {
        fail(h, HPMC_ERR_ARG, "no output buffer bound to the traversal handle");
        return false;
    }
    if (th->out_ce && !hpmc_host::plan_cells_fit_u32(h->plan)) {
        fail(h, HPMC_ERR_SIZE, "(cell, edge) records index cells in 32 bits; the grid has more cells");
        return false;
    }
    uint64_t cap = th->capacity;
    if (th->out_ce && th->ce_capacity < cap) cap = th->ce_capacity;
    if (cap > 0xffffffffull) cap = 0xffffffffull;
    if (h->count_valid && h->count > cap) {
        fail(h, HPMC_ERR_CAPACITY, "output holds %llu vertices, the surface has %u", (unsigned long long)cap,
             h->count);
        return false;
    }
    if (!launch_emit(h, th->out, th->out_ce, (uint32_t)cap)) return false;
    ok(h);
    return true;
}

}  // namespace

// =========================================================================
extern "C" {

struct HPMCConstants* HPMCcreateConstants(int32_t, int32_t)
{
    HPMCConstants* c = new (std::nothrow) HPMCConstants;
    if (!c) return nullptr;
    // The tables go to the current HIP device now; without a usable device
    // the handle is still made (setters, layout and traversal source work on
    // the host) and the first build uploads them, failing there if it can't.
    (void)upload_tables(c);
    ok(nullptr);
    return c;
}

void HPMCdestroyConstants(struct HPMCConstants* c)
{
    if (!c) return;
    if (c->d_cases) (void)hipFree(c->d_cases);
    if (c->d_bnorm) (void)hipFree(c->d_bnorm);
    delete c;
}

struct HPMCHistoPyramid* HPMCcreateHistoPyramid(struct HPMCConstants* c)
{
    if (!c) {
        fail(nullptr, HPMC_ERR_NULL, "createHistoPyramid called with constants = NULL");
        return nullptr;
    }
    HPMCHistoPyramid* h = new (std::nothrow) HPMCHistoPyramid;
    if (!h) return nullptr;
    h->constants = c;
    ok(nullptr);
    return h;
}

void HPMCsetLatticeSize(struct HPMCHistoPyramid* h, int32_t x, int32_t y, int32_t z)
{
    if (!h) return;
    h->size[0] = x;
    h->size[1] = y;
    h->size[2] = z;
    for (int a = 0; a < 3; ++a) h->cells[a] = (h->size[a] > 1 ? h->size[a] : 1) - 1;
    h->tainted = true;
    h->broken = false;
}

void HPMCsetGridSize(struct HPMCHistoPyramid* h, int32_t x, int32_t y, int32_t z)
{
    if (!h) return;
    h->cells[0] = x;
    h->cells[1] = y;
    h->cells[2] = z;
    h->tainted = true;
    h->broken = false;
}

void HPMCsetFieldAsBinary(struct HPMCHistoPyramid* h)
{
    if (h) h->binary = true;
}

void HPMCsetFieldAsContinuous(struct HPMCHistoPyramid* h)
{
    if (h) h->binary = false;
}

void HPMCsetGridExtent(struct HPMCHistoPyramid* h, float x, float y, float z)
{
    if (!h) return;
    h->extent[0] = x;
    h->extent[1] = y;
    h->extent[2] = z;
    h->tainted = true;
    h->broken = false;
}

void HPMCsetFieldTexture3D(struct HPMCHistoPyramid* h, const float* volume, uint8_t gradient)
{
    if (!h) return;
    // hpiface.cpp:164-184: a new texture alone does not force a rebuild of
    // the programs; a change of mode or gradient does.
    h->volume = volume;
    const bool grad = gradient != 0;
    if (h->mode != kFieldVolume || h->gradient != grad || h->unorm8) {
        h->mode = kFieldVolume;
        h->gradient = grad;
        h->unorm8 = false;
        h->tainted = true;
        h->broken = false;
    }
}

void HPMCsetFieldTexture3DUnorm8(struct HPMCHistoPyramid* h, const uint8_t* volume)
{
    if (!h) return;
    // as HPMCsetFieldTexture3D: a new volume alone does not force a rebuild
    h->volume8 = volume;
    if (h->mode != kFieldVolume || h->gradient || !h->unorm8) {
        h->mode = kFieldVolume;
        h->gradient = false;
        h->unorm8 = true;
        h->tainted = true;
        h->broken = false;
    }
}

void HPMCsetFieldCustom(struct HPMCHistoPyramid* h, const char* source, uint32_t, uint8_t gradient)
{
    if (!h) return;
    h->mode = kFieldCustom;
    h->snippet = source ? source : "";
    h->gradient = gradient != 0;
    h->tainted = true;
    h->broken = false;
}

uint32_t HPMCgetBuilderProgram(struct HPMCHistoPyramid* h)
{
    if (!h) {
        fail(nullptr, HPMC_ERR_NULL, "h == NULL");
        return 0;
    }
    if (h->broken) {
        fail(h, HPMC_ERR_BROKEN, "h is broken");
        return 0;
    }
    if (h->tainted && !setup(h)) return 0;
    return h->builder_id;
}

void HPMCdestroyHandle(struct HPMCHistoPyramid* h)
{
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    else (void)hipDeviceSynchronize();
    free_device(h);
    if (h->ev_order) (void)hipEventDestroy(h->ev_order);
    for (int st = 0; st < 3; ++st)
        for (hipEvent_t e : h->stage_events[st]) (void)hipEventDestroy(e);
    delete h;
}

void HPMCbuildHistopyramid(struct HPMCHistoPyramid* h, float threshold)
{
    if (!h || h->broken) return;
    // hpiface.cpp:258-268: a failed setup leaves the HP tainted, not broken.
    if (h->tainted && !setup(h)) return;
    h->threshold = h->binary_active ? 0.5f : threshold;
    h->count_valid = false;
    h->built = false;
    if (!launch_classify(h, h->threshold) || !launch_scan(h)) {
        h->broken = true;
        return;
    }
    // The scan writes the totals straight into the pinned host copy; the
    // event a count waits on is recorded when the count is asked for (an
    // event recorded here, between the scan and the extraction, held the
    // extraction back by ~5.6 us).
    h->built = true;
    ok(h);
}

uint32_t HPMCacquireNumberOfVertices(struct HPMCHistoPyramid* h)
{
    if (!h || h->broken) return 0;
    if (!wait_count(h)) {
        h->broken = true;
        return 0;
    }
    return h->count;
}

struct HPMCTraversalHandle* HPMCcreateTraversalHandle(struct HPMCHistoPyramid* h)
{
    if (!h) {
        fail(nullptr, HPMC_ERR_NULL, "createTraversalHandle called with h == NULL");
        return nullptr;
    }
    if (!setup(h)) {
        // Without a HIP device the handle is still made, as the constants
        // are (its traversal source can be taken on the host; extraction
        // fails).  Any other setup failure is one, as in the reference.
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) == hipSuccess && ndev > 0) return nullptr;
        (void)hipGetLastError();
    }
    HPMCTraversalHandle* th = new (std::nothrow) HPMCTraversalHandle;
    if (!th) return nullptr;
    th->h = h;
    return th;
}

void HPMCdestroyTraversalHandle(struct HPMCTraversalHandle* th) { delete th; }

char* HPMCgetTraversalShaderFunctions(struct HPMCTraversalHandle* th)
{
    if (!th || !th->h) {
        fail(nullptr, HPMC_ERR_NULL, "getTraversalShaderFunctions called with th == NULL");
        return nullptr;
    }
    HPMCHistoPyramid* h = th->h;
    // the configuration as the next build takes it (no device work here)
    if (!determine_layout(h)) return nullptr;
    const bool binary = h->tainted ? h->binary : h->binary_active;
    // One self-contained hipRTC source: the layout, the engine's device code,
    // this HistoPyramid's field and emission mode, and hpmc_for_each_vertex.
    char defs[1024];
    snprintf(defs, sizeof defs,
             "// HPMC layout of this HistoPyramid (hpmc/src/shadergen.cpp:45-83)\n"
             "#define HPMC_FUNC_X %d\n#define HPMC_FUNC_Y %d\n#define HPMC_FUNC_Z %d\n"
             "#define HPMC_CELLS_X %d\n#define HPMC_CELLS_Y %d\n#define HPMC_CELLS_Z %d\n"
             "#define HPMC_TILES_X %d\n#define HPMC_TILES_Y %d\n"
             "#define HPMC_TILE_SIZE_X %d\n#define HPMC_TILE_SIZE_Y %d\n#define HPMC_HP_SIZE_L2 %d\n",
             h->size[0], h->size[1], h->size[2], h->cells[0], h->cells[1], h->cells[2], h->tiles[0], h->tiles[1],
             h->tile[0], h->tile[1], h->l2);
    const bool snippet = h->mode == kFieldCustom;
    std::string s = defs;
    s += "#define HPMC_RTC 1\n#define HPMC_TABLES_INLINE 1\n";
    if (h->gradient) s += "#define HPMC_FIELD_GRADIENT 1\n";
    s += "typedef unsigned char uint8_t;\ntypedef unsigned short uint16_t;\n";
    s += "typedef unsigned int uint32_t;\ntypedef unsigned long long uint64_t;\n";
    // the headers inline, without their include guards' pragma (one file)
    auto inline_header = [](const char* text) {
        std::string t = text;
        for (size_t at; (at = t.find("#pragma once")) != std::string::npos;) t.erase(at, 12);
        return t + "\n";
    };
    s += inline_header(kEmbeddedTablesHeader);
    s += inline_header(kEmbeddedDeviceHeader);
    if (snippet) s += snippet_field_source(h->snippet);
    s += "#line 1 \"hpmc_traversal\"\n";
    if (snippet)
        s += "#define HPMC_TRAVERSAL_FIELD(dd) (hpmc::SnippetField{(dd).g.ipx, (dd).g.ipy, (dd).g.ipz})\n";
    else if (h->unorm8)
        s += "#define HPMC_TRAVERSAL_FIELD(dd) "
             "(hpmc::Volume8Field{(const uint8_t*)(dd).field, (dd).g.sx, (dd).g.sy, (dd).g.sz})\n";
    else if (h->gradient)
        s += "#define HPMC_TRAVERSAL_FIELD(dd) "
             "(hpmc::Volume4Field{(const float4*)(dd).field, (dd).g.sx, (dd).g.sy, (dd).g.sz})\n";
    else
        s += "#define HPMC_TRAVERSAL_FIELD(dd) "
             "(hpmc::VolumeField{(const float*)(dd).field, (dd).g.sx, (dd).g.sy, (dd).g.sz})\n";
    s += std::string("#define HPMC_TRAVERSAL_MODE ") +
         (binary ? "hpmc::kEmitBinary" : h->gradient ? "hpmc::kEmitGradient" : "hpmc::kEmitContinuous") +
         "\n";
    s += "// cb(k, normal, position) for every vertex k of the build the view\n"
         "// describes (HPMCgetTraversalView); CB needs a static constexpr bool\n"
         "// kActive = true.  With static constexpr bool kWantsEdge = true as well,\n"
         "// cb(k, a, b, normal, position) gets the texel-centre coordinates a, b\n"
         "// ((i + 0.5) / lattice size) of the vertex's edge ends too, the\n"
         "// reference's extractVertex(a, b, p, n).  Launch HPMCgetTraversalGrid\n"
         "// workgroups of 256 threads.\n"
         "template <class CB>\n"
         "__device__ __forceinline__ void hpmc_for_each_vertex(const hpmc::TraversalView& v, const CB& cb)\n"
         "{\n"
         "    auto mk = [](const hpmc::VolDesc& dd) { return HPMC_TRAVERSAL_FIELD(dd); };\n"
         "    hpmc::for_each_vertex<HPMC_TRAVERSAL_MODE>(v, mk, cb);\n"
         "}\n";
    char* r = (char*)malloc(s.size() + 1);
    if (!r) return nullptr;
    memcpy(r, s.c_str(), s.size() + 1);
    return r;
}

int32_t HPMCgetTraversalView(struct HPMCTraversalHandle* th, void* view, size_t bytes)
{
    if (!th || !th->h || !view) return fail(th ? th->h : nullptr, HPMC_ERR_NULL, "NULL argument");
    static_assert(sizeof(hpmc::TraversalView) == HPMC_TRAVERSAL_VIEW_BYTES, "view size");
    HPMCHistoPyramid* h = th->h;
    if (bytes < sizeof(hpmc::TraversalView))
        return fail(h, HPMC_ERR_ARG, "view needs %zu bytes", sizeof(hpmc::TraversalView));
    if (h->broken) return fail(h, HPMC_ERR_BROKEN, "HistoPyramid is broken");
    if (h->tainted || !h->built) return fail(h, HPMC_ERR_ARG, "HistoPyramid has not been built since it was configured");
    hpmc::TraversalView v{};
    v.vols = h->plan.d_vols;
    v.items = h->plan.items;
    v.l4_off = h->plan.l4_off;
    v.totals = h->plan.d_totals;
    v.cases = h->constants->d_cases;
    v.bnorm = h->constants->d_bnorm;
    v.threshold = h->threshold;
    memcpy(view, &v, sizeof v);
    ok(h);
    return HPMC_OK;
}

int32_t HPMCgetTraversalGrid(struct HPMCTraversalHandle* th)
{
    if (!th || !th->h) return 0;
    HPMCHistoPyramid* h = th->h;
    if (h->tainted && !setup(h)) return 0;
    return h->plan.emit_wgs;
}

bool HPMCsetTraversalHandleProgram(struct HPMCTraversalHandle* th, uint32_t program, uint32_t u1, uint32_t u2,
                                   uint32_t u3)
{
    if (!th) {
        fail(nullptr, HPMC_ERR_NULL, "setTraversalHandleProgram called with th == NULL");
        return false;
    }
    // triface.cpp:217-238: a program is required and the work units must be
    // distinct (the third is unused by custom fetches).
    if (program == 0) {
        fail(th->h, HPMC_ERR_ARG, "program == 0");
        return false;
    }
    const bool custom = th->h && th->h->mode == kFieldCustom;
    if (u1 == u2 || (!custom && (u1 == u3 || u2 == u3))) {
        fail(th->h, HPMC_ERR_ARG, "texture units must be distinct");
        return false;
    }
    th->program = program;
    return true;
}

bool HPMCextractVertices(struct HPMCTraversalHandle* th) { return extract_helper(th); }
bool HPMCextractVerticesTransformFeedback(struct HPMCTraversalHandle* th) { return extract_helper(th); }
bool HPMCextractVerticesTransformFeedbackNV(struct HPMCTraversalHandle* th) { return extract_helper(th); }
bool HPMCextractVerticesTransformFeedbackEXT(struct HPMCTraversalHandle* th) { return extract_helper(th); }

int32_t HPMCgetLastError(struct HPMCHistoPyramid* h) { return h ? h->last_error : t_last_error; }

const char* HPMCgetLastErrorString(struct HPMCHistoPyramid* h) { return h ? h->last_msg.c_str() : t_last_msg.c_str(); }

int32_t HPMCsetStream(struct HPMCHistoPyramid* h, void* stream)
{
    if (!h) return fail(nullptr, HPMC_ERR_NULL, "h == NULL");
    const hipStream_t next = (hipStream_t)stream;
    // Work already queued on the old stream (a build whose extraction comes
    // next) must finish before anything queued on the new one: the new
    // stream waits for an event recorded on the old.  Without device state
    // nothing has been queued.
    if (next != h->stream && h->plan.d_vols) {
        if (!h->ev_order && hipEventCreateWithFlags(&h->ev_order, hipEventDisableTiming) != hipSuccess) {
            h->ev_order = nullptr;
            return fail(h, HPMC_ERR_HIP, "failed to create the stream-order event");
        }
        if (hipEventRecord(h->ev_order, h->stream) != hipSuccess || hipStreamWaitEvent(next, h->ev_order, 0) != hipSuccess)
            return fail(h, HPMC_ERR_HIP, "failed to order the new stream after the old one");
    }
    h->stream = next;
    ok(h);
    return HPMC_OK;
}

int32_t HPMCsetTuning(struct HPMCHistoPyramid* h, int32_t classify_waves, int32_t zseg, int32_t emit_workgroups)
{
    if (!h) return fail(nullptr, HPMC_ERR_NULL, "h == NULL");
    if (classify_waves < 0 || classify_waves > hpmc::kClassifyMaxWaves || (zseg && zseg != 8 && zseg != 16 &&
                                                                           zseg != 32) || emit_workgroups < 0)
        return fail(h, HPMC_ERR_ARG, "invalid tuning (waves 0..16, zseg 0/8/16/32, workgroups >= 0)");
    h->tuning.classify_waves = classify_waves;
    h->tuning.zseg = zseg;
    h->tuning.emit_wgs = emit_workgroups;
    h->tainted = true;
    ok(h);
    return HPMC_OK;
}

int32_t HPMCgetTuning(struct HPMCHistoPyramid* h, int32_t out[3])
{
    if (!h || !out) return fail(h, HPMC_ERR_NULL, "NULL argument");
    if (h->tainted && !setup(h)) return h->last_error;
    out[0] = h->plan.cls_waves;
    out[1] = h->plan.vols[0].g.zseg;
    out[2] = h->plan.emit_wgs;
    ok(h);
    return HPMC_OK;
}

int32_t HPMCsetFieldUniform(struct HPMCHistoPyramid* h, const char* name, const void* data, size_t bytes)
{
    if (!h || !name || (!data && bytes)) return fail(h, HPMC_ERR_NULL, "NULL argument");
    if (h->broken) return fail(h, HPMC_ERR_BROKEN, "h is broken");
    if (h->mode != kFieldCustom) return fail(h, HPMC_ERR_ARG, "field is not a snippet");
    if (h->tainted && !setup(h)) return h->last_error;
    hipDeviceptr_t dptr = nullptr;
    size_t sz = 0;
    if (hipModuleGetGlobal(&dptr, &sz, h->module, name) != hipSuccess)
        return fail(h, HPMC_ERR_ARG, "snippet has no __device__ global named '%s'", name);
    if (bytes > sz) return fail(h, HPMC_ERR_ARG, "'%s' holds %zu bytes, got %zu", name, sz, bytes);
    if (hipMemcpyHtoDAsync(dptr, const_cast<void*>(data), bytes, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess)
        return fail(h, HPMC_ERR_HIP, "failed to set '%s'", name);
    ok(h);
    return HPMC_OK;
}

int32_t HPMCsetTraversalHandleBuffer(struct HPMCTraversalHandle* th, float* out, uint64_t capacity)
{
    if (!th) return fail(nullptr, HPMC_ERR_NULL, "th == NULL");
    th->out = out;
    th->capacity = out ? capacity : 0;
    ok(th->h);
    return HPMC_OK;
}

int32_t HPMCsetTraversalHandleCellEdgeBuffer(struct HPMCTraversalHandle* th, void* records, uint64_t capacity)
{
    if (!th) return fail(nullptr, HPMC_ERR_NULL, "th == NULL");
    th->out_ce = (hpmc::CellEdge*)records;
    th->ce_capacity = records ? capacity : 0;
    ok(th->h);
    return HPMC_OK;
}

uint32_t HPMCacquireDroppedVertices(struct HPMCTraversalHandle* th)
{
    if (!th || !th->h || th->h->broken) return 0;
    if (!wait_count(th->h)) return 0;
    uint64_t cap = th->capacity;
    if (th->out_ce && th->ce_capacity < cap) cap = th->ce_capacity;
    return th->h->count > cap ? (uint32_t)(th->h->count - cap) : 0u;
}

int32_t HPMCgetCellCodes(struct HPMCHistoPyramid* h, uint8_t* codes)
{
    if (!h || !codes) return fail(h, HPMC_ERR_NULL, "NULL argument");
    if (h->broken) return fail(h, HPMC_ERR_BROKEN, "h is broken");
    if (h->tainted && !setup(h)) return h->last_error;
    if (h->mode == kFieldVolume) {
        if (!volume_ptr(h)) return fail(h, HPMC_ERR_FIELD, "no field volume set");
        if (hpmc_host::plan_set_field(h->plan, 0, volume_ptr(h), h->stream) != hipSuccess)
            return fail(h, HPMC_ERR_HIP, "descriptor update failed");
    }
    const hpmc::Geom& g = h->plan.vols[0].g;
    const uint64_t n = (uint64_t)g.cx * g.cy * g.cz;
    const unsigned blocks = (unsigned)std::min<uint64_t>((n + 255) / 256, 65536);
    const hpmc::VolDesc* vols = h->plan.d_vols;
    uint32_t vol = 0;
    float Ta = h->binary_active ? 0.5f : h->threshold;
    uint8_t* c = codes;
    void* args[] = {&vols, &vol, &Ta, &c};
    if (hpmc_host::launch(h->k_codes, blocks, 256, args, h->stream) != hipSuccess || hipGetLastError() != hipSuccess)
        return fail(h, HPMC_ERR_HIP, "codes kernel launch failed");
    ok(h);
    return HPMC_OK;
}

uint64_t HPMCgetLevel4(struct HPMCHistoPyramid* h, uint32_t* out, uint64_t capacity)
{
    if (!h) {
        fail(h, HPMC_ERR_NULL, "NULL argument");
        return 0;
    }
    if (h->broken || h->tainted || !h->built) {
        fail(h, HPMC_ERR_ARG, "HistoPyramid not built");
        return 0;
    }
    const uint64_t n = h->plan.vols[0].n_level4;
    if (!out) {   // size query
        ok(h);
        return n;
    }
    if (capacity < n) {
        fail(h, HPMC_ERR_CAPACITY, "need %llu entries", (unsigned long long)n);
        return 0;
    }
    hipLaunchKernelGGL(hpmc_level4_gather, dim3(1024), dim3(256), 0, h->stream, (const hpmc::VolDesc*)h->plan.d_vols,
                       0u, (const uint16_t*)h->plan.colcnt, out);
    if (hipGetLastError() != hipSuccess) {
        fail(h, HPMC_ERR_HIP, "level-4 gather launch failed");
        return 0;
    }
    ok(h);
    return n;
}

int32_t HPMCsetProfiling(struct HPMCHistoPyramid* h, int32_t enable)
{
    if (!h) return fail(nullptr, HPMC_ERR_NULL, "h == NULL");
    h->profiling = enable != 0;
    ok(h);
    return HPMC_OK;
}

int32_t HPMCgetStageTimes(struct HPMCHistoPyramid* h, double* ms, uint32_t* launches)
{
    if (!h || !ms) return fail(h, HPMC_ERR_NULL, "NULL argument");
    bool good = true;
    for (int st = 0; st < 3; ++st) {
        good = stage_fold(h, st) && good;
        ms[st] = h->stage_ms[st];
        if (launches) launches[st] = h->stage_n[st];
        h->stage_ms[st] = 0.0;
        h->stage_n[st] = 0;
    }
    if (!good) return fail(h, HPMC_ERR_HIP, "event sync failed");
    ok(h);
    return HPMC_OK;
}

int32_t HPMCcheckFieldSnippet(const char* source, uint32_t flags, char* log, size_t log_bytes)
{
    if (!source) return fail(nullptr, HPMC_ERR_NULL, "source == NULL");
    std::string text;
    std::vector<char> code =
        rtc_compile(source, (flags & HPMC_SNIPPET_GRADIENT) != 0, (flags & HPMC_SNIPPET_BINARY) != 0, &text);
    if (log && log_bytes) {
        size_t n = text.size() < log_bytes - 1 ? text.size() : log_bytes - 1;
        memcpy(log, text.data(), n);
        log[n] = '\0';
    }
    if (code.empty()) return fail(nullptr, HPMC_ERR_COMPILE, "field snippet failed to compile: %s", text.c_str());
    ok(nullptr);
    return HPMC_OK;
}

int32_t HPMCadjoint(struct HPMCTraversalHandle* th, const float* grad_vertices, float* grad_field, float* grad_iso)
{
    if (!th || !th->h || !grad_field || !grad_iso) return fail(th ? th->h : nullptr, HPMC_ERR_NULL, "NULL argument");
    HPMCHistoPyramid* h = th->h;
    if (h->broken) return fail(h, HPMC_ERR_BROKEN, "HistoPyramid is broken");
    if (h->tainted || !h->built) return fail(h, HPMC_ERR_ARG, "HistoPyramid has not been built since it was configured");
    if (h->mode != kFieldVolume || h->gradient || h->unorm8 || h->binary_active)
        return fail(h, HPMC_ERR_UNSUPPORTED, "the adjoint covers continuous FP32 volume fields");
    if (!wait_count(h)) return h->last_error;
    if (h->count && !grad_vertices) return fail(h, HPMC_ERR_NULL, "grad_vertices == NULL");
    std::string err;
    const int32_t rc = hpmc_host::adjoint_run(h->plan, h->adjoint, h->threshold, h->count, grad_vertices, grad_field, grad_iso,
                                              h->constants->d_cases, h->stream, &err);
    if (rc != HPMC_OK) return fail(h, rc, "%s", err.c_str());
    ok(h);
    return HPMC_OK;
}

int32_t HPMCgetLayout(struct HPMCHistoPyramid* h, int32_t out[5])
{
    if (!h || !out) return fail(h, HPMC_ERR_NULL, "NULL argument");
    if (!determine_layout(h)) return h->last_error;
    out[0] = h->tile[0];
    out[1] = h->tile[1];
    out[2] = h->tiles[0];
    out[3] = h->tiles[1];
    out[4] = h->l2;
    ok(h);
    return HPMC_OK;
}

}  // extern "C"

// =========================================================================
// Batched builds: one plan over many volumes.
struct HPMCBatch {
    HPMCConstants* constants = nullptr;
    std::vector<hpmc_host::VolumeSpec> specs;
    hpmc_host::Tuning tuning;
    bool tainted = true;
    hpmc_host::Plan plan;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    bool built = false;
    bool counts_known = false;
    float last_threshold = 0.0f;
    hpmc_host::AdjointScratch adjoint;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    float graph_threshold = 0.0f;    // the threshold the graph was captured at
};

namespace {

#define BATCH_TRY(expr)                                                                    \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess) return fail(nullptr, HPMC_ERR_HIP, "%s: %s", #expr, hipGetErrorString(_e)); \
    } while (0)

void batch_drop_graph(HPMCBatch* b)
{
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->graph_exec = nullptr;
    b->graph = nullptr;
}

int32_t batch_setup(HPMCBatch* b)
{
    if (!b->stream) {
        // blocking with respect to the null stream, default priority
        BATCH_TRY(hipStreamCreate(&b->stream));
        b->own_stream = true;
    }
    if (!b->tainted) return HPMC_OK;
    if (!upload_tables(b->constants)) return fail(nullptr, HPMC_ERR_HIP, "no HIP device for the marching-cubes tables");
    batch_drop_graph(b);
    std::string err;
    // batches hold FP32 volumes
    const int32_t rc = hpmc_host::plan_create(b->plan, b->specs, b->tuning, (const void*)hpmc_classify_volume,
                                              device_shape(), true, &err);
    if (rc != HPMC_OK) return fail(nullptr, rc, "%s", err.c_str());
    b->tainted = false;
    b->built = false;
    b->counts_known = false;
    return HPMC_OK;
}

// extract_in_capture: the extraction is enqueued in the same graph capture.
int32_t batch_enqueue_build(HPMCBatch* b, float T, bool extract_in_capture = false)
{
    BATCH_TRY(hpmc_host::sums_prepare(b->plan, b->stream, extract_in_capture));
    BATCH_TRY(hpmc_host::plan_classify(b->plan, hpmc_host::Kernel{(const void*)hpmc_classify_volume, nullptr}, T,
                                       b->constants->d_cases, b->stream));
    BATCH_TRY(hpmc_host::plan_scan(b->plan, (const void*)hpmc_scan, b->stream));
    return HPMC_OK;
}

int32_t batch_enqueue_extract(HPMCBatch* b, float T, float* out, uint64_t capacity, void* cell_edge)
{
    if (cell_edge && !hpmc_host::plan_cells_fit_u32(b->plan))
        return fail(nullptr, HPMC_ERR_SIZE, "(cell, edge) records index cells in 32 bits; a volume has more cells");
    const uint32_t cap = (uint32_t)std::min<uint64_t>(capacity, 0xffffffffull);
    BATCH_TRY(hpmc_host::plan_emit(b->plan, hpmc_host::Kernel{(const void*)hpmc_emit_volume_batch, nullptr}, T, out,
                                   (hpmc::CellEdge*)cell_edge, cap, b->constants->d_cases, b->constants->d_bnorm,
                                   b->stream));
    b->plan.sums_dirty = false;   // the extraction clears the tile sums
    return HPMC_OK;
}

}  // namespace

extern "C" {

struct HPMCBatch* HPMCcreateBatch(struct HPMCConstants* c)
{
    if (!c) {
        fail(nullptr, HPMC_ERR_NULL, "createBatch called with constants = NULL");
        return nullptr;
    }
    HPMCBatch* b = new (std::nothrow) HPMCBatch;
    if (!b) return nullptr;
    b->constants = c;
    ok(nullptr);
    return b;
}

void HPMCdestroyBatch(struct HPMCBatch* b)
{
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_drop_graph(b);
    hpmc_host::plan_free(b->plan);
    hpmc_host::adjoint_free(b->adjoint);
    if (b->own_stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int32_t HPMCbatchAddVolume(struct HPMCBatch* b, const float* volume, int32_t x, int32_t y, int32_t z, float ext_x,
                           float ext_y, float ext_z)
{
    if (!b || !volume) return -fail(nullptr, HPMC_ERR_NULL, "NULL argument");
    hpmc_host::VolumeSpec v;
    v.size[0] = x;
    v.size[1] = y;
    v.size[2] = z;
    for (int a = 0; a < 3; ++a) v.cells[a] = v.size[a] - 1;
    v.extent[0] = ext_x;
    v.extent[1] = ext_y;
    v.extent[2] = ext_z;
    v.field = volume;
    hpmc_host::Layout lay;
    std::string err;
    const int32_t rc = hpmc_host::layout_of(v.size, v.cells, &lay, &err);
    if (rc != HPMC_OK) return -fail(nullptr, rc, "%s", err.c_str());
    // batched classification covers 16x16-cell blocks only
    if (2 * lay.tile[0] < hpmc::kBlock || 2 * lay.tile[1] < hpmc::kBlock)
        return -fail(nullptr, HPMC_ERR_UNSUPPORTED, "batched volumes need at least 9 cells along x and y");
    if (b->specs.size() >= 65535) return -fail(nullptr, HPMC_ERR_CAPACITY, "a batch holds at most 65535 volumes");
    b->specs.push_back(v);
    b->tainted = true;
    ok(nullptr);
    return (int32_t)b->specs.size() - 1;
}

int32_t HPMCbatchNumVolumes(struct HPMCBatch* b) { return b ? (int32_t)b->specs.size() : 0; }

int32_t HPMCbatchSetTuning(struct HPMCBatch* b, int32_t classify_waves, int32_t zseg, int32_t emit_workgroups)
{
    if (!b) return fail(nullptr, HPMC_ERR_NULL, "NULL batch");
    if (classify_waves < 0 || classify_waves > hpmc::kClassifyMaxWaves || (zseg && zseg != 8 && zseg != 16 &&
                                                                           zseg != 32) || emit_workgroups < 0)
        return fail(nullptr, HPMC_ERR_ARG, "invalid tuning (waves 0..16, zseg 0/8/16/32, workgroups >= 0)");
    b->tuning.classify_waves = classify_waves;
    b->tuning.zseg = zseg;
    b->tuning.emit_wgs = emit_workgroups;
    b->tainted = true;
    ok(nullptr);
    return HPMC_OK;
}

int32_t HPMCbatchGetTuning(struct HPMCBatch* b, int32_t out[3])
{
    if (!b || !out) return fail(nullptr, HPMC_ERR_NULL, "NULL argument");
    if (b->specs.empty()) return fail(nullptr, HPMC_ERR_ARG, "empty batch");
    const int32_t rc = batch_setup(b);
    if (rc != HPMC_OK) return rc;
    out[0] = b->plan.cls_waves;
    out[1] = b->plan.vols[0].g.zseg;
    out[2] = b->plan.emit_wgs;
    ok(nullptr);
    return HPMC_OK;
}

int32_t HPMCbatchSetStream(struct HPMCBatch* b, void* stream)
{
    if (!b) return fail(nullptr, HPMC_ERR_NULL, "NULL batch");
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->own_stream) (void)hipStreamDestroy(b->stream);
    b->own_stream = false;
    b->stream = (hipStream_t)stream;
    batch_drop_graph(b);
    ok(nullptr);
    return HPMC_OK;
}

int32_t HPMCbatchBuild(struct HPMCBatch* b, float threshold)
{
    if (!b) return fail(nullptr, HPMC_ERR_NULL, "NULL batch");
    if (b->specs.empty()) return fail(nullptr, HPMC_ERR_ARG, "empty batch");
    int32_t rc = batch_setup(b);
    if (rc != HPMC_OK) return rc;
    rc = batch_enqueue_build(b, threshold);
    if (rc != HPMC_OK) return rc;
    b->last_threshold = threshold;
    b->built = true;
    b->counts_known = false;
    ok(nullptr);
    return HPMC_OK;
}

int32_t HPMCbatchGetCounts(struct HPMCBatch* b, uint32_t* counts, uint32_t* bases, uint64_t* total)
{
    if (!b) return fail(nullptr, HPMC_ERR_NULL, "NULL batch");
    if (!b->built) return fail(nullptr, HPMC_ERR_ARG, "batch not built");
    BATCH_TRY(hipStreamSynchronize(b->stream));
    uint64_t sum = 0;
    for (int v = 0; v < b->plan.nvol; ++v) {
        const uint32_t n = b->plan.h_vcounts[v].end - b->plan.h_vcounts[v].base;
        if (counts) counts[v] = n;
        if (bases) bases[v] = b->plan.h_vcounts[v].base;
        sum += n;
    }
    if (total) *total = sum;
    b->counts_known = true;
    ok(nullptr);
    return HPMC_OK;
}

int32_t HPMCbatchExtract(struct HPMCBatch* b, float* out, uint64_t capacity, void* cell_edge)
{
    if (!b || !out) return fail(nullptr, HPMC_ERR_NULL, "NULL argument");
    if (!b->built || b->tainted) return fail(nullptr, HPMC_ERR_ARG, "batch not built since it was configured");
    if (b->counts_known && b->plan.h_totals->vertices > capacity)
        return fail(nullptr, HPMC_ERR_CAPACITY, "output holds %llu vertices, the batch has %u",
                    (unsigned long long)capacity, b->plan.h_totals->vertices);
    const int32_t rc = batch_enqueue_extract(b, b->last_threshold, out, capacity, cell_edge);
    if (rc != HPMC_OK) return rc;
    ok(nullptr);
    return HPMC_OK;
}

int32_t HPMCbatchCaptureGraph(struct HPMCBatch* b, float threshold, float* out, uint64_t capacity, void* cell_edge)
{
    if (!b || !out) return fail(nullptr, HPMC_ERR_NULL, "NULL argument");
    if (b->specs.empty()) return fail(nullptr, HPMC_ERR_ARG, "empty batch");
    int32_t rc = batch_setup(b);
    if (rc != HPMC_OK) return rc;
    if (!b->stream) return fail(nullptr, HPMC_ERR_ARG, "graph capture needs a stream");
    batch_drop_graph(b);
    BATCH_TRY(hipStreamSynchronize(b->stream));
    BATCH_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    b->last_threshold = threshold;
    const bool dirty = b->plan.sums_dirty;   // (capture runs nothing)
    rc = batch_enqueue_build(b, threshold, true);
    if (rc == HPMC_OK) rc = batch_enqueue_extract(b, threshold, out, capacity, cell_edge);
    b->plan.sums_dirty = dirty;
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc != HPMC_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(nullptr, HPMC_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    b->graph = g;
    b->graph_threshold = threshold;
    BATCH_TRY(hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0));
    ok(nullptr);
    return HPMC_OK;
}

int32_t HPMCbatchLaunchGraph(struct HPMCBatch* b)
{
    if (!b) return fail(nullptr, HPMC_ERR_NULL, "NULL batch");
    if (!b->graph_exec || b->tainted) return fail(nullptr, HPMC_ERR_ARG, "no captured graph for this configuration");
    // the graph's build adds into the tile sums that its extraction clears
    // again: clear them first after a build that no extraction followed
    if (b->plan.sums_dirty) {
        const uint32_t n = b->plan.n_l4t + b->plan.n_colt;
        hipLaunchKernelGGL(hpmc_clear_sums, dim3(std::min<uint32_t>((n + 255) / 256, 1024u)), dim3(256), 0,
                           b->stream, b->plan.tile_sums, n);
        BATCH_TRY(hipGetLastError());
    }
    BATCH_TRY(hipGraphLaunch(b->graph_exec, b->stream));
    b->plan.sums_dirty = false;
    // the replayed build is at the captured threshold: later extractions
    // and adjoints of this build use it
    b->last_threshold = b->graph_threshold;
    b->built = true;
    b->counts_known = false;
    ok(nullptr);
    return HPMC_OK;
}

int32_t HPMCbatchAdjoint(struct HPMCBatch* b, const float* grad_vertices, float* grad_field, float* grad_iso)
{
    if (!b || !grad_field || !grad_iso) return fail(nullptr, HPMC_ERR_NULL, "NULL argument");
    if (!b->built || b->tainted) return fail(nullptr, HPMC_ERR_ARG, "batch not built since it was configured");
    BATCH_TRY(hipStreamSynchronize(b->stream));
    const uint32_t n = b->plan.h_totals->vertices;
    if (n && !grad_vertices) return fail(nullptr, HPMC_ERR_NULL, "grad_vertices == NULL");
    std::string err;
    const int32_t rc = hpmc_host::adjoint_run(b->plan, b->adjoint, b->last_threshold, n, grad_vertices, grad_field, grad_iso,
                                              b->constants->d_cases, b->stream, &err);
    if (rc != HPMC_OK) return fail(nullptr, rc, "%s", err.c_str());
    ok(nullptr);
    return HPMC_OK;
}

int32_t HPMCbatchAutotune(struct HPMCBatch* b, float threshold, float* out, uint64_t capacity, int32_t repetitions,
                          double* best_ms)
{
    if (!b || !out) return fail(nullptr, HPMC_ERR_NULL, "NULL argument");
    if (b->specs.empty()) return fail(nullptr, HPMC_ERR_ARG, "empty batch");
    if (repetitions < 1) repetitions = 1;
    // classification waves: powers of two up to 8 (narrow volumes stack row
    // bands, so more waves than a row needs still pay) and whole rows of the
    // widest volume
    int max_row_waves = 1;
    for (const auto& v : b->specs) max_row_waves = std::max(max_row_waves, hpmc_host::row_waves(v.size[0]));
    std::vector<int> waves = {1, 2, 4, 8};
    if (std::find(waves.begin(), waves.end(), max_row_waves) == waves.end()) waves.push_back(max_row_waves);
    const hpmc_host::DeviceShape& dev = device_shape();
    const int full = (dev.cus > 0 ? dev.cus : 256) * (dev.emit_per_cu > 0 ? dev.emit_per_cu : 4);
    const int emit_opts[] = {full / 2, full};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    BATCH_TRY(hipEventCreate(&e0));
    BATCH_TRY(hipEventCreate(&e1));
    hpmc_host::Tuning best;
    double best_t = 1e30;
    int32_t rc = HPMC_OK;
    for (int w : waves) {
        for (int zs : {8, 16, 32}) {
            for (int ew : emit_opts) {
                b->tuning.classify_waves = w;
                b->tuning.zseg = zs;
                b->tuning.emit_wgs = ew;
                b->tainted = true;
                rc = batch_setup(b);
                if (rc != HPMC_OK) break;
                std::vector<float> ts;
                for (int r = 0; r <= repetitions && rc == HPMC_OK; ++r) {
                    (void)hipEventRecord(e0, b->stream);
                    rc = batch_enqueue_build(b, threshold);
                    if (rc == HPMC_OK) rc = batch_enqueue_extract(b, threshold, out, capacity, nullptr);
                    (void)hipEventRecord(e1, b->stream);
                    if (hipEventSynchronize(e1) != hipSuccess) rc = fail(nullptr, HPMC_ERR_HIP, "autotune run failed");
                    float ms = 0.0f;
                    (void)hipEventElapsedTime(&ms, e0, e1);
                    if (r > 0) ts.push_back(ms);   // the first run warms up
                }
                if (rc != HPMC_OK) break;
                std::sort(ts.begin(), ts.end());
                const double med = ts[ts.size() / 2];
                if (med < best_t) {
                    best_t = med;
                    best = b->tuning;
                }
            }
            if (rc != HPMC_OK) break;
        }
        if (rc != HPMC_OK) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != HPMC_OK) return rc;
    b->tuning = best;
    b->tainted = true;
    rc = batch_setup(b);
    if (rc != HPMC_OK) return rc;
    if (best_ms) *best_ms = best_t;
    ok(nullptr);
    return HPMC_OK;
}

// No-allocation build + extraction of an FP32 volume (no counterpart in
// the reference): the kernels of HPMCbuildHistopyramid + HPMCextractVertices
// with the plan's arrays in `scratch`.
int32_t HPMCextractVolumeNoAlloc(struct HPMCConstants* c, const float* volume, int32_t x, int32_t y, int32_t z,
                                 float ext_x, float ext_y, float ext_z, float threshold, float* out,
                                 uint64_t capacity, uint32_t* device_count, void* hip_stream)
{
    if (!c || !volume || (!out && capacity)) return fail(nullptr, HPMC_ERR_NULL, "NULL argument");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail(nullptr, HPMC_ERR_HIP, "no HIP device");
    if (!c->d_cases || c->device != dev)
        return fail(nullptr, HPMC_ERR_ARG, "the constants were not created with device %d current", dev);
    hipError_t st = hipSuccess;
    char* base = (char*)scratch_address(dev, &st);
    if (!base) return fail(nullptr, HPMC_ERR_HIP, "scratch address on device %d: %s", dev, hipGetErrorString(st));
    hpmc_host::VolumeSpec spec{{x, y, z}, {x - 1, y - 1, z - 1}, {ext_x, ext_y, ext_z}, volume};
    hpmc_host::Plan p;   // default-constructed: no host or device memory
    hpmc_host::Layout lay;
    hpmc::VolDesc desc;
    std::string err;
    int32_t rc = hpmc_host::plan_describe(p, &spec, 1, &lay, &desc, hpmc_host::Tuning{},
                                          (const void*)hpmc_classify_volume, device_shape(), true, &err);
    if (rc != HPMC_OK) return fail(nullptr, rc, "%s", err.c_str());
    const uint64_t need = hpmc_host::plan_carve(p, nullptr);
    if (need > kNoAllocScratchBytes)
        return fail(nullptr, HPMC_ERR_SIZE,
                    "lattice %dx%dx%d needs %llu bytes of scratch; the no-allocation entry point holds %llu",
                    x, y, z, (unsigned long long)need, (unsigned long long)kNoAllocScratchBytes);
    (void)hpmc_host::plan_carve(p, base);
    const hipStream_t stream = (hipStream_t)hip_stream;
    hipLaunchKernelGGL(hpmc_noalloc_begin, dim3(1), dim3(256), 0, stream, desc, p.d_vols, p.tile_sums,
                       p.n_l4t + p.n_colt);
    // lattices past the MALL stream through classification (as setup picks)
    const uint64_t bytes = 4ull * (uint64_t)x * (uint64_t)y * (uint64_t)z;
    const hpmc_host::Kernel k_cls{bytes >= (256ull << 20) ? (const void*)hpmc_classify_volume_stream
                                                           : (const void*)hpmc_classify_volume,
                                  nullptr};
    const uint32_t cap = (uint32_t)std::min<uint64_t>(capacity, 0xffffffffull);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hpmc_host::plan_small_blocks(p, desc)
                ? hpmc_host::plan_classify_small(p, hpmc_host::Kernel{(const void*)hpmc_classify_small_volume, nullptr},
                                                 threshold, c->d_cases, stream)
                : hpmc_host::plan_classify(p, k_cls, threshold, c->d_cases, stream);
    if (e == hipSuccess) e = hpmc_host::plan_scan(p, (const void*)hpmc_scan, stream);
    if (e == hipSuccess)
        e = hpmc_host::plan_emit(p, hpmc_host::Kernel{(const void*)hpmc_emit_volume, nullptr}, threshold, out,
                                 nullptr, cap, c->d_cases, c->d_bnorm, stream);
    if (e == hipSuccess && device_count)
        e = hipMemcpyAsync(device_count, &p.d_totals->vertices, sizeof(uint32_t), hipMemcpyDeviceToDevice, stream);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) return fail(nullptr, HPMC_ERR_HIP, "no-allocation build: %s", hipGetErrorString(e));
    ok(nullptr);
    return HPMC_OK;
}

}  // extern "C"
