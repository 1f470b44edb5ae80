/*
 * hpmc.h -- C ABI of the MI355X HistoPyramid marching-cubes engine.
 *
 * A drop-in for the C API of HPMC 1.0.1 (hpmc/include/hpmc.h:295-544): the
 * same function names, arity and call order, with OpenGL object names
 * replaced by HIP device memory.  No OpenGL header is needed.
 *
 *   GL type     -> here
 *   GLsizei     -> int32_t        GLfloat -> float
 *   GLuint      -> uint32_t       GLboolean -> uint8_t
 *   3D texture  -> device pointer to an FP32 lattice, x fastest
 *   GLSL fetch snippet -> HIP C++ snippet compiled with hipRTC
 *   display program / transform feedback -> device output buffer (N3F_V3F)
 *
 * Errors follow the reference: no call throws; failures return 0/false/NULL
 * and latch the HistoPyramid as broken until it is reconfigured by a setter
 * (hpmc/src/hpiface.cpp:100-200).  HPMCgetLastError adds a reason.
 * All device work of a HistoPyramid runs in order on its HIP stream (the
 * null stream unless HPMCsetStream names another); HPMCacquireNumberOfVertices
 * is the one call that waits for the device (hpmc.h:211-215 of the reference).
 */
#ifndef HPMC_AMD_HPMC_H
#define HPMC_AMD_HPMC_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

struct HPMCConstants;
struct HPMCHistoPyramid;
struct HPMCTraversalHandle;

/* ---- constants --------------------------------------------------------- */

/* Replaces HPMCcreateConstants (hpmc.h:310-311, constants.cpp:54-347).
 * Uploads the marching-cubes tables to the current HIP device.  The GL
 * version caps are accepted and ignored. */
struct HPMCConstants* HPMCcreateConstants(int32_t max_gl_major, int32_t max_gl_minor);

/* Replaces HPMCdestroyConstants (hpmc.h:319-320, constants.cpp:350-392). */
void HPMCdestroyConstants(struct HPMCConstants* c);

/* ---- HistoPyramid configuration ---------------------------------------- */

/* Replaces HPMCcreateHistoPyramid (hpmc.h:330-331, hpiface.cpp:45-97).
 * The new HistoPyramid is tainted and broken until configured. */
struct HPMCHistoPyramid* HPMCcreateHistoPyramid(struct HPMCConstants* c);

/* Replaces HPMCsetLatticeSize (hpmc.h:346-350, hpiface.cpp:100-114).
 * Also resets the grid size to lattice - 1. */
void HPMCsetLatticeSize(struct HPMCHistoPyramid* h, int32_t x_size, int32_t y_size, int32_t z_size);

/* Replace HPMCsetFieldAsBinary / HPMCsetFieldAsContinuous (hpmc.h:352-356,
 * hpiface.cpp:131-142).  Binary fields classify against 0.5. */
void HPMCsetFieldAsBinary(struct HPMCHistoPyramid* h);
void HPMCsetFieldAsContinuous(struct HPMCHistoPyramid* h);

/* Replaces HPMCsetGridSize (hpmc.h:375-379, hpiface.cpp:117-128). */
void HPMCsetGridSize(struct HPMCHistoPyramid* h, int32_t x_size, int32_t y_size, int32_t z_size);

/* Replaces HPMCsetGridExtent (hpmc.h:392-396, hpiface.cpp:145-161). */
void HPMCsetGridExtent(struct HPMCHistoPyramid* h, float x_extent, float y_extent, float z_extent);

/* Replaces HPMCsetFieldTexture3D (hpmc.h:405-408, hpiface.cpp:164-184).
 * `volume` is a device pointer to lattice_x*lattice_y*lattice_z FP32
 * samples, x fastest.  Samples beyond the lattice (forward differences at
 * the upper faces) clamp to the edge sample, as a NEAREST/CLAMP texture. */
void HPMCsetFieldTexture3D(struct HPMCHistoPyramid* h, const float* volume, uint8_t gradient);

/* An 8-bit normalised volume, the field of the reference's texture3d demo
 * (a GL_ALPHA8 texture sampled NEAREST with clamping,
 * apps/texture3d/texture3d.cpp:127-141): `volume` is an x-fastest device
 * array of the lattice size whose byte c samples as c / 255 (correctly
 * rounded).  Setter semantics as HPMCsetFieldTexture3D (hpiface.cpp:164-184);
 * no gradient variant, no adjoint.  No counterpart function in the
 * reference, where the texture's format decides. */
void HPMCsetFieldTexture3DUnorm8(struct HPMCHistoPyramid* h, const uint8_t* volume);

/* Replaces HPMCsetFieldCustom (hpmc.h:439-443, hpiface.cpp:187-200).
 * `source` is HIP C++ defining
 *     __device__ float HPMC_fetch(float3 p);
 * evaluated at texel centres p = ((i+0.5)/X, (j+0.5)/Y, (k+0.5)/Z), exactly
 * as the GLSL contract (hpmc.h:409-437).  Values the snippet reads at run
 * time are declared as __device__ globals and set with HPMCsetFieldUniform.
 * `builder_texunit` is accepted and ignored. */
void HPMCsetFieldCustom(struct HPMCHistoPyramid* h, const char* source, uint32_t builder_texunit,
                        uint8_t gradient);

/* Replaces HPMCgetBuilderProgram (hpmc.h:445-446, hpiface.cpp:203-222).
 * Sets the HistoPyramid up if needed and returns a non-zero id of the
 * compiled builder, or 0 when broken or when setup fails. */
uint32_t HPMCgetBuilderProgram(struct HPMCHistoPyramid* h);

/* Declared at hpmc.h:450-452 but never defined by the reference: frees the
 * HistoPyramid and its device memory. */
void HPMCdestroyHandle(struct HPMCHistoPyramid* h);

/* ---- build and count ---------------------------------------------------- */

/* Replaces HPMCbuildHistopyramid (hpmc.h:460-462, hpiface.cpp:225-290).
 * Enqueues classification and the pyramid reduction; does not wait. */
void HPMCbuildHistopyramid(struct HPMCHistoPyramid* h, float threshold);

/* Replaces HPMCacquireNumberOfVertices (hpmc.h:468-469, hpiface.cpp:293-345).
 * Waits for the last build and returns its vertex count. */
uint32_t HPMCacquireNumberOfVertices(struct HPMCHistoPyramid* h);

/* ---- traversal and extraction ------------------------------------------ */

/* Replaces HPMCcreateTraversalHandle (hpmc.h:477-478, triface.cpp:43-109).
 * Like HPMCcreateConstants, it still makes the handle when no HIP device is
 * present (so its traversal source can be taken on the host); extraction
 * then fails. */
struct HPMCTraversalHandle* HPMCcreateTraversalHandle(struct HPMCHistoPyramid* h);

/* Replaces HPMCdestroyTraversalHandle (hpmc.h:484-485, triface.cpp:112-122). */
void HPMCdestroyTraversalHandle(struct HPMCTraversalHandle* th);

/* Replaces HPMCgetTraversalShaderFunctions (hpmc.h:493-494,
 * triface.cpp:125-192): returns HIP device source (layout constants plus
 * the device functions) for callers that fuse extraction into their own
 * hipRTC kernels.  The caller frees it with free().  Its
 * hpmc_for_each_vertex(view, cb) calls cb(k, normal, position), or, for a
 * callback declaring kWantsEdge = true, cb(k, a, b, normal, position) with
 * the texel-centre coordinates of the vertex's edge ends: the reference's
 * two- and four-output extractVertex forms (shadergen.cpp:310, 443). */
char* HPMCgetTraversalShaderFunctions(struct HPMCTraversalHandle* th);

/* Bytes of the traversal view (hpmc::TraversalView in the source above). */
#define HPMC_TRAVERSAL_VIEW_BYTES 64

/* Device-side view of the last build of th's HistoPyramid for a kernel
 * compiled from HPMCgetTraversalShaderFunctions' source: pass it by value
 * to a kernel that calls hpmc_for_each_vertex(view, cb), which calls
 * cb(k, normal, position) once for every vertex k (the extraction's order
 * and bits).  Launch HPMCgetTraversalGrid(th) workgroups of 256 threads on
 * the HistoPyramid's stream after the build.  The view stays valid until
 * the HistoPyramid is reconfigured; it describes whichever build ran last.
 * No counterpart in the reference, where the traversal functions read the
 * HistoPyramid texture bound by HPMCsetTraversalHandleProgram
 * (triface.cpp:195-327). */
int32_t HPMCgetTraversalView(struct HPMCTraversalHandle* th, void* view, size_t bytes);
/* Workgroups (of 256 threads) for a traversal kernel; 0 on error. */
int32_t HPMCgetTraversalGrid(struct HPMCTraversalHandle* th);

/* Replaces HPMCsetTraversalHandleProgram (hpmc.h:517-522,
 * triface.cpp:195-327).  There is no GL program: the call validates its
 * arguments like the reference (non-zero program, distinct units) and
 * records the program id; output goes to the buffer bound with
 * HPMCsetTraversalHandleBuffer. */
bool HPMCsetTraversalHandleProgram(struct HPMCTraversalHandle* th, uint32_t program,
                                   uint32_t tex_unit_work1, uint32_t tex_unit_work2,
                                   uint32_t tex_unit_work3);

/* Replace HPMCextractVertices and its transform-feedback variants
 * (hpmc.h:535-545, triface.cpp:332-500).  All four write the same thing:
 * per vertex the normal then the position, 6 floats (N3F_V3F, as the
 * reference's transform-feedback demo captures them,
 * apps/transform_feedback/transform_feedback.cpp:292-296), in the
 * reference's extraction order, into the bound buffer.  Enqueued; does not
 * wait.  Vertices beyond the buffer's capacity are dropped and counted. */
bool HPMCextractVertices(struct HPMCTraversalHandle* th);
bool HPMCextractVerticesTransformFeedback(struct HPMCTraversalHandle* th);
bool HPMCextractVerticesTransformFeedbackNV(struct HPMCTraversalHandle* th);
bool HPMCextractVerticesTransformFeedbackEXT(struct HPMCTraversalHandle* th);

/* ---- additions (no counterpart in the reference) ----------------------- */

/* Error codes returned by the additions and by HPMCgetLastError. */
enum {
    HPMC_OK = 0,
    HPMC_ERR_NULL = 1,          /* NULL handle or pointer */
    HPMC_ERR_BROKEN = 2,        /* handle latched broken */
    HPMC_ERR_SIZE = 3,          /* lattice / grid size rejected */
    HPMC_ERR_FIELD = 4,         /* no field source */
    HPMC_ERR_COMPILE = 5,       /* hipRTC failed on the snippet */
    HPMC_ERR_HIP = 6,           /* HIP runtime error */
    HPMC_ERR_CAPACITY = 7,      /* output buffer too small */
    HPMC_ERR_ARG = 8,           /* invalid argument */
    HPMC_ERR_UNSUPPORTED = 9    /* configuration not implemented */
};

/* Last error of a HistoPyramid (or of the calling thread when h is NULL)
 * and its message; the message stays valid until the next call on h. */
int32_t HPMCgetLastError(struct HPMCHistoPyramid* h);
const char* HPMCgetLastErrorString(struct HPMCHistoPyramid* h);

/* HIP stream (hipStream_t) for all device work of h; NULL = null stream.
 * Work queued on the new stream afterwards runs after the work already
 * queued on the old one (the new stream waits for an event of the old). */
int32_t HPMCsetStream(struct HPMCHistoPyramid* h, void* hip_stream);

/* Launch shapes of h's kernels, as HPMCbatchSetTuning; 0 picks
 * automatically.  Results do not depend on them.  Takes effect at the next
 * build (a reconfiguration).  No counterpart in the reference, whose launch
 * shapes are fixed by the GL pipeline. */
int32_t HPMCsetTuning(struct HPMCHistoPyramid* h, int32_t classify_waves, int32_t zseg, int32_t emit_workgroups);
/* The shapes in use: {classify_waves, zseg, emit_workgroups} (configures h
 * if needed). */
int32_t HPMCgetTuning(struct HPMCHistoPyramid* h, int32_t out[3]);

/* Sets a __device__ global of the field snippet, by name, from host memory
 * (takes the place of glUniform* on HPMCgetBuilderProgram's program,
 * apps/metaballs/metaballs.cpp:277-291).  Stream-ordered with builds. */
int32_t HPMCsetFieldUniform(struct HPMCHistoPyramid* h, const char* name, const void* data, size_t bytes);

/* Binds the extraction output: a device buffer of capacity_vertices * 6
 * floats (N3F_V3F).  Replaces the transform-feedback buffer. */
int32_t HPMCsetTraversalHandleBuffer(struct HPMCTraversalHandle* th, float* device_n3f_v3f,
                                     uint64_t capacity_vertices);

/* Optionally binds a second output: per vertex {uint32 cell, uint32 edge},
 * cell = x + cells_x*(y + cells_y*z), edge = 4*axis + s0 + 2*s1 with (s0,s1)
 * the edge's start offsets along the other two axes (x,y,z order).
 * NULL unbinds. */
int32_t HPMCsetTraversalHandleCellEdgeBuffer(struct HPMCTraversalHandle* th, void* device_records,
                                             uint64_t capacity_vertices);

/* Vertices the last extraction dropped for lack of capacity (waits). */
uint32_t HPMCacquireDroppedVertices(struct HPMCTraversalHandle* th);

/* Writes the HPMC corner code of every cell (x fastest over the grid) for
 * the last build's threshold into a device buffer of cells_x*cells_y*cells_z
 * bytes.  Enqueued. */
int32_t HPMCgetCellCodes(struct HPMCHistoPyramid* h, uint8_t* device_codes);

/* Copies the pyramid level holding the block sums of the last build,
 * indexed by Morton block code, into a device buffer; returns the entry
 * count, or 0 on failure (with device_out NULL: the count only).  Blocks are
 * 16x16 cells of the doubled base image (level 4 of the reference's
 * base-level pyramid), or b x b cells (level log2 b) for grids whose doubled
 * tile is narrower, b its side (grids of at most 8 cells along x or y). */
uint64_t HPMCgetLevel4(struct HPMCHistoPyramid* h, uint32_t* device_out, uint64_t capacity);

/* Per-stage timing with HIP events on the HistoPyramid's stream: when
 * enabled, every launch of classify, scan and emit records a pair of events
 * at its kernel's start and end (hipExtLaunchKernel), so the times are the
 * kernels' own; the launches are spaced further apart meanwhile.
 * HPMCgetStageTimes waits for them, writes the summed milliseconds
 * (ms[3] = classify, scan, emit) and launch counts, and starts over.  While
 * profiling stays on without a HPMCgetStageTimes call, every 1024 pairs of a
 * stage are folded into its totals (a wait on the device), so the events
 * held stay bounded. */
int32_t HPMCsetProfiling(struct HPMCHistoPyramid* h, int32_t enable);
int32_t HPMCgetStageTimes(struct HPMCHistoPyramid* h, double* ms, uint32_t* launches);

/* Compiles a field snippet (as HPMCsetFieldCustom takes it) for gfx950
 * without touching a device; the compiler log goes to `log`.  `flags`:
 * HPMC_SNIPPET_GRADIENT (the snippet also defines HPMC_fetchGrad) and/or
 * HPMC_SNIPPET_BINARY (binary-field extraction).  Returns HPMC_OK or
 * HPMC_ERR_COMPILE. */
#define HPMC_SNIPPET_GRADIENT 1u
#define HPMC_SNIPPET_BINARY 2u
int32_t HPMCcheckFieldSnippet(const char* source, uint32_t flags, char* log, size_t log_bytes);

/* Adjoint of the last extraction (no counterpart in the reference): from
 * dL/dvertices (device, count x 6 FP32 in the N3F_V3F order of the output)
 * to dL/dfield (device, one FP32 per lattice sample, every sample written)
 * and dL/dT (device, one FP32), through the edge interpolation
 * t = (f(a)-T)/(f(a)-f(b)), the positions and the forward-difference
 * normals.  Interpolation weights are recomputed, not stored; the result is
 * deterministic (the vertices of each lattice edge summed in a fixed order,
 * each sample's edge contributions summed in a fixed order, no atomics).
 * Continuous FP32 volume fields only, volumes of at most 2^30 samples and
 * fewer than 2^27 vertices (HPMC_ERR_SIZE otherwise); waits for the count. */
int32_t HPMCadjoint(struct HPMCTraversalHandle* th, const float* grad_vertices, float* grad_field, float* grad_iso);

/* Layout of the reference base level (hpmc/src/init.cpp:105-130):
 * out = {tile_x, tile_y, tiles_x, tiles_y, size_l2}. */
int32_t HPMCgetLayout(struct HPMCHistoPyramid* h, int32_t out[5]);

/* ---- batched builds (no counterpart in the reference) -------------------
 *
 * Many FP32 volumes of uneven sizes built and extracted together: one
 * classification launch, one scan launch and one persistent extraction
 * launch cover every volume, and the chain can be captured once and replayed
 * as a hipGraph.  Each volume's count and vertex list are those of a
 * HistoPyramid of that volume alone (same codes, same order, same bits); the
 * output is the concatenation of the volumes' lists in the order they were
 * added.  For callers that extract many small or medium volumes per step
 * (training loops, serving).  Errors: HPMC_* codes; the reason is in
 * HPMCgetLastError(NULL) / HPMCgetLastErrorString(NULL) of the calling
 * thread.  Work runs on the batch's stream (HPMCbatchSetStream, else a
 * stream the batch creates that orders with the null stream). */
struct HPMCBatch;

struct HPMCBatch* HPMCcreateBatch(struct HPMCConstants* c);
void HPMCdestroyBatch(struct HPMCBatch* b);

/* Adds a volume (device FP32 lattice x*y*z, x fastest; grid = lattice - 1,
 * extent as HPMCsetGridExtent).  Returns its index, or a negative HPMC_*
 * code. */
int32_t HPMCbatchAddVolume(struct HPMCBatch* b, const float* volume, int32_t x, int32_t y, int32_t z,
                           float ext_x, float ext_y, float ext_z);
int32_t HPMCbatchNumVolumes(struct HPMCBatch* b);

/* Launch shapes; 0 picks automatically.  classify_waves: waves per
 * classification workgroup (each covers 64 lattice columns x 16 cell rows);
 * zseg: slices per classification workgroup (8, 16 or 32); emit_workgroups:
 * persistent extraction workgroups.  Results do not depend on them. */
int32_t HPMCbatchSetTuning(struct HPMCBatch* b, int32_t classify_waves, int32_t zseg, int32_t emit_workgroups);
/* The shapes in use after the last configuration: {classify_waves, zseg,
 * emit_workgroups}. */
int32_t HPMCbatchGetTuning(struct HPMCBatch* b, int32_t out[3]);
int32_t HPMCbatchSetStream(struct HPMCBatch* b, void* stream);

/* Classifies and scans every volume (enqueued, no host wait). */
int32_t HPMCbatchBuild(struct HPMCBatch* b, float threshold);
/* Waits for the last build; per-volume counts and output bases (arrays of
 * HPMCbatchNumVolumes entries, either may be NULL) and the total. */
int32_t HPMCbatchGetCounts(struct HPMCBatch* b, uint32_t* counts, uint32_t* bases, uint64_t* total);
/* Extracts every volume into `out` (N3F_V3F, capacity vertices) and, if not
 * NULL, (cell, edge) records (uint32 pairs, cell index within its volume).
 * Vertices past the capacity are dropped; after HPMCbatchGetCounts a too
 * small buffer is an HPMC_ERR_CAPACITY error instead. */
int32_t HPMCbatchExtract(struct HPMCBatch* b, float* out, uint64_t capacity, void* cell_edge);

/* Captures build(threshold) + extract(out, capacity, cell_edge) into a
 * hipGraph (replacing any earlier one); HPMCbatchLaunchGraph replays it.
 * Counts of a replay: HPMCbatchGetCounts. */
int32_t HPMCbatchCaptureGraph(struct HPMCBatch* b, float threshold, float* out, uint64_t capacity, void* cell_edge);
int32_t HPMCbatchLaunchGraph(struct HPMCBatch* b);

/* Adjoint of the last batched build + extraction (as HPMCadjoint):
 * grad_field holds the volumes' lattices concatenated in order. */
int32_t HPMCbatchAdjoint(struct HPMCBatch* b, const float* grad_vertices, float* grad_field, float* grad_iso);

/* Autotuner: times build + extract into `out` under each candidate launch
 * shape (repetitions runs each, median) and keeps the fastest; best_ms (may
 * be NULL) receives its time.  Results are identical under every shape. */
int32_t HPMCbatchAutotune(struct HPMCBatch* b, float threshold, float* out, uint64_t capacity, int32_t repetitions,
                          double* best_ms);

/* ---- build + extraction without allocation ------------------------------
 *
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  One call does what HPMCbuildHistopyramid (continuous
 * field) + HPMCextractVertices do for an FP32 volume (device lattice x*y*z,
 * x fastest, grid = lattice - 1, extent as HPMCsetGridExtent): the same
 * kernels, the same vertices in the same order with the same bits.  Writes
 * N3F_V3F rows into `out` (capacity vertices; the rest are dropped) and, if
 * device_count is not NULL, the vertex count into that device word.  All of
 * it is enqueued on hip_stream (NULL = null stream); nothing waits on the
 * host.  It calls no hipMalloc, hipFree or other allocating routine: the
 * build's arrays live in a scratch array of HPMC_NO_ALLOC_SCRATCH_BYTES per
 * device, held from the first use of the library on that device until the
 * process ends; a lattice whose arrays need more is rejected with
 * HPMC_ERR_SIZE before anything is enqueued (512^3 needs about 66 MB).
 * Every call on a device shares that device's scratch, so calls that overlap
 * on one device (other streams, other threads) race: order them.  `c` must
 * have been created with the current device current.  Errors as for
 * batches: HPMCgetLastError(NULL) of the calling thread. */
#define HPMC_NO_ALLOC_SCRATCH_BYTES (80ull << 20)
int32_t HPMCextractVolumeNoAlloc(struct HPMCConstants* c, const float* volume, int32_t x, int32_t y, int32_t z,
                                 float ext_x, float ext_y, float ext_z, float threshold, float* out,
                                 uint64_t capacity, uint32_t* device_count, void* hip_stream);

#ifdef __cplusplus
}
#endif
#endif
