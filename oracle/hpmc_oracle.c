/*
 * CPU restatement of HPMC's HistoPyramid marching-cubes path (test
 * infrastructure only: the parity tests and the bench's CPU-baseline leg
 * call it as the checker; the product library never links it).
 *
 * Every function cites the reference lines it restates (paths relative to
 * the hetcomp/hpmc source tree).  Integer pyramid sums are kept as integers;
 * the reference keeps them in FP32, which is exact below 2^24 vertices.
 *
 * Field samples are addressed by lattice index (i, j, k); the field array is
 * x-fastest with dimensions (fx, fy, fz).  With clamp != 0 an index is
 * clamped to [0, size-1] per axis (a NEAREST/CLAMP 3D texture,
 * apps/texture3d/texture3d.cpp:135-139); with clamp == 0 the array must
 * extend to every index the caller reaches (a field snippet evaluated past
 * the lattice, shadergen.cpp:268-277).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "hpmc_oracle.h"

/* ------------------------------------------------------------------------ */
/* Marching-cubes cases by HPMC code: count and per-slot (shift, axis).       */
/* Filled from the fixture by hpo_set_cases (the oracle carries no table).    */
static int g_count[256];
static int g_slot[256][16][4];
static int g_cases_ready = 0;

void hpo_set_cases(const int32_t* counts, const int32_t* slots /* 256*16*4 */)
{
    for (int c = 0; c < 256; ++c) {
        g_count[c] = counts[c];
        for (int s = 0; s < 16; ++s)
            for (int k = 0; k < 4; ++k)
                g_slot[c][s][k] = slots[(c * 16 + s) * 4 + k];
    }
    g_cases_ready = 1;
}

int hpo_cases_ready(void) { return g_cases_ready; }

/* ------------------------------------------------------------------------ */
/* Layout: hpmc/src/init.cpp:82-130.  FP32 log2f/ceilf/sqrtf as the reference */
/* computes them (the tile size is 1<<ceil(log2(cells/2)) per axis, slices  */
/* tile a power-of-two square base level).                                 */
int hpo_layout(const int32_t size[3], const int32_t cells[3], int32_t out[5])
{
    for (int a = 0; a < 3; ++a) {
        if (size[a] < 2 || size[a] > 16384) return 0;      /* init.cpp:82-91  */
        if (size[a] <= cells[a]) return 0;                  /* init.cpp:98-103 */
        if (cells[a] < 1) return 0;
    }
    /* cells == 1 gives log2f(0.5f) = -1 and a negative shift in the
     * reference (init.cpp:106-111), which is undefined; reject it. */
    if (cells[0] < 2 || cells[1] < 2) return 0;
    int32_t ts0 = 1 << (int32_t)ceilf(log2f((float)cells[0] / 2.0f));
    int32_t ts1 = 1 << (int32_t)ceilf(log2f((float)cells[1] / 2.0f));
    float aspect = (float)ts0 / (float)ts1;
    float l = ceilf(log2f(sqrtf((float)cells[2] / aspect)));
    if (l < 0.0f) l = 0.0f;
    int32_t lay0 = 1 << (int32_t)l;
    int32_t lay1 = (cells[2] + lay0 - 1) / lay0;
    int64_t w0 = (int64_t)ts0 * lay0, w1 = (int64_t)ts1 * lay1;
    int32_t l2 = (int32_t)ceilf(log2f((float)(w0 > w1 ? w0 : w1)));
    int32_t S = 1 << l2;
    out[0] = ts0;
    out[1] = ts1;
    out[2] = S / ts0;
    out[3] = S / ts1;
    out[4] = l2;
    return 1;
}

/* ------------------------------------------------------------------------ */
static inline int64_t sample_index(const int32_t dims[3], const int32_t size[3], int clamp,
                                   int64_t i, int64_t j, int64_t k)
{
    if (clamp) {
        if (i > size[0] - 1) i = size[0] - 1;
        if (j > size[1] - 1) j = size[1] - 1;
        if (k > size[2] - 1) k = size[2] - 1;
        if (i < 0) i = 0;
        if (j < 0) j = 0;
        if (k < 0) k = 0;
    }
    return (k * dims[1] + j) * dims[0] + i;
}

static inline float sample(const float* f, const int32_t dims[3], const int32_t size[3],
                           int clamp, int64_t i, int64_t j, int64_t k)
{
    return f[sample_index(dims, size, clamp, i, j, k)];
}

/* Base-level classification: hpmc/src/shadergen.cpp:86-171.  A cell's code  */
/* has bit x+2y+4z set when the sample at corner (x,y,z) is below T          */
/* (148-153); binary fields use T = 0.5 (99-101).  counts[] is the table's   */
/* vertex count (constants.cpp:304-326).  Cells are x-fastest over cells[].  */
void hpo_classify(const float* f, const int32_t dims[3], const int32_t size[3],
                  const int32_t cells[3], float T, uint8_t* codes, uint8_t* counts)
{
    for (int64_t z = 0; z < cells[2]; ++z)
        for (int64_t y = 0; y < cells[1]; ++y)
            for (int64_t x = 0; x < cells[0]; ++x) {
                int code = 0;
                for (int b = 0; b < 8; ++b) {
                    float v = sample(f, dims, size, 0, x + (b & 1), y + ((b >> 1) & 1), z + (b >> 2));
                    if (v < T) code |= 1 << b;
                }
                int64_t c = (z * cells[1] + y) * cells[0] + x;
                codes[c] = (uint8_t)code;
                if (counts) counts[c] = (uint8_t)g_count[code];
            }
}

/* ------------------------------------------------------------------------ */
/* Base-level image position of a cell: init.cpp:105-130 tiles slice z at    */
/* tile (z mod TILES_X, z div TILES_X); base texel (tx,ty) covers the 2x2     */
/* cells (2lx+cx, 2ly+cy) in channel cx+2cy (shadergen.cpp:103-125,148-153). */
/* In doubled coordinates the cell sits at X = tile_x*2*TS_X + x, likewise Y. */
static inline void cell_image_pos(const int32_t lay[5], int64_t x, int64_t y, int64_t z,
                                  uint32_t* X, uint32_t* Y)
{
    int64_t tx = z % lay[2], ty = z / lay[2];
    *X = (uint32_t)(tx * 2 * lay[0] + x);
    *Y = (uint32_t)(ty * 2 * lay[1] + y);
}

static inline uint64_t morton2(uint32_t x, uint32_t y)
{
    uint64_t k = 0;
    for (int b = 0; b < 16; ++b) {
        k |= (uint64_t)((x >> b) & 1u) << (2 * b);
        k |= (uint64_t)((y >> b) & 1u) << (2 * b + 1);
    }
    return k;
}

uint64_t hpo_morton2(uint32_t x, uint32_t y) { return morton2(x, y); }

/* ------------------------------------------------------------------------ */
/* Literal HistoPyramid: build (buildpass.cpp:42-178; reductions              */
/* shadergen.cpp:174-225, children (0,0),(1,0),(0,1),(1,1) summed per        */
/* channel) and GL3 traversal (shadergen.cpp:445-503).  For each key in      */
/* [0,N) it writes the cell index and the table slot.  Memory is O(S^2), so  */
/* callers keep this to small lattices; hpo_vertex_list gives the same list  */
/* from the Morton closed form.  Returns N, or -1 on failure.                */
int64_t hpo_hp_traverse(const uint8_t* counts, const int32_t cells[3], const int32_t lay[5],
                        uint32_t* out_cell, uint8_t* out_slot, int64_t cap)
{
    int l2 = lay[4];
    int64_t S = (int64_t)1 << l2;
    /* level m has (S>>m)^2 texels of 4 channels */
    uint32_t** lev = (uint32_t**)calloc((size_t)l2 + 1, sizeof(uint32_t*));
    if (!lev) return -1;
    for (int m = 0; m <= l2; ++m) {
        int64_t w = S >> m;
        lev[m] = (uint32_t*)calloc((size_t)(w * w * 4), sizeof(uint32_t));
        if (!lev[m]) { for (int q = 0; q < m; ++q) free(lev[q]); free(lev); return -1; }
    }
    /* base level: slices >= CELLS_Z and masked cells hold zero counts */
    for (int64_t z = 0; z < cells[2]; ++z)
        for (int64_t y = 0; y < cells[1]; ++y)
            for (int64_t x = 0; x < cells[0]; ++x) {
                uint32_t X, Y;
                cell_image_pos(lay, x, y, z, &X, &Y);
                int64_t tx = X >> 1, ty = Y >> 1;
                int ch = (int)((X & 1) + 2 * (Y & 1));
                lev[0][(ty * S + tx) * 4 + ch] = counts[(z * cells[1] + y) * cells[0] + x];
            }
    for (int m = 1; m <= l2; ++m) {
        int64_t w = S >> m, wp = S >> (m - 1);
        for (int64_t j = 0; j < w; ++j)
            for (int64_t i = 0; i < w; ++i)
                for (int c = 0; c < 4; ++c) {
                    int64_t ci = 2 * i + (c & 1), cj = 2 * j + (c >> 1);
                    const uint32_t* t = &lev[m - 1][(cj * wp + ci) * 4];
                    lev[m][(j * w + i) * 4 + c] = t[0] + t[1] + t[2] + t[3];
                }
    }
    /* count: sum of the top texel's channels (hpiface.cpp:323-327) */
    const uint32_t* top = lev[l2];
    int64_t N = (int64_t)top[0] + top[1] + top[2] + top[3];
    for (int64_t key0 = 0; key0 < N && key0 < cap; ++key0) {
        int64_t key = key0, px = 0, py = 0;
        for (int m = l2; m >= 0; --m) {
            int64_t w = S >> m;
            const uint32_t* s = &lev[m][(py * w + px) * 4];
            px *= 2;
            py *= 2;
            if (s[0] <= key) {
                key -= s[0];
                if (s[1] <= key) {
                    key -= s[1];
                    if (s[2] <= key) { key -= s[2]; px += 1; py += 1; }
                    else { py += 1; }
                } else {
                    px += 1;
                }
            }
        }
        /* (px,py) are doubled base coordinates: decode tile and cell */
        int64_t tx = px / (2 * lay[0]), ty = py / (2 * lay[1]);
        int64_t x = px % (2 * lay[0]), y = py % (2 * lay[1]);
        int64_t z = tx + (int64_t)lay[2] * ty;
        out_cell[key0] = (uint32_t)((z * cells[1] + y) * cells[0] + x);
        out_slot[key0] = (uint8_t)key;
    }
    for (int m = 0; m <= l2; ++m) free(lev[m]);
    free(lev);
    return N;
}

/* ------------------------------------------------------------------------ */
static int cmp_u64(const void* a, const void* b)
{
    uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return (x > y) - (x < y);
}

/* Vertex list in HPMC's extraction order from the closed form: the GL3      */
/* traversal visits base-image quadrants (0,0),(1,0),(0,1),(1,1) at every    */
/* level (shadergen.cpp:448-495), i.e. cells sorted by the Morton code of    */
/* their doubled base-image position (x bit low), then by table slot.       */
/* Writes cell index and slot per vertex; returns N (or -1 on failure).      */
int64_t hpo_vertex_list(const uint8_t* counts, const int32_t cells[3], const int32_t lay[5],
                        uint32_t* out_cell, uint8_t* out_slot, int64_t cap)
{
    int64_t ncell = (int64_t)cells[0] * cells[1] * cells[2];
    int64_t nact = 0;
    for (int64_t c = 0; c < ncell; ++c) nact += counts[c] != 0;
    uint64_t* keys = (uint64_t*)malloc((size_t)(nact ? nact : 1) * sizeof(uint64_t));
    uint32_t* idx = (uint32_t*)malloc((size_t)(nact ? nact : 1) * sizeof(uint32_t));
    if (!keys || !idx) { free(keys); free(idx); return -1; }
    int64_t a = 0;
    for (int64_t z = 0; z < cells[2]; ++z)
        for (int64_t y = 0; y < cells[1]; ++y)
            for (int64_t x = 0; x < cells[0]; ++x) {
                int64_t c = (z * cells[1] + y) * cells[0] + x;
                if (!counts[c]) continue;
                uint32_t X, Y;
                cell_image_pos(lay, x, y, z, &X, &Y);
                keys[a] = (morton2(X, Y) << 32) | (uint64_t)a;
                idx[a] = (uint32_t)c;
                ++a;
            }
    qsort(keys, (size_t)nact, sizeof(uint64_t), cmp_u64);
    int64_t n = 0;
    for (int64_t i = 0; i < nact; ++i) {
        uint32_t c = idx[keys[i] & 0xffffffffu];
        for (int s = 0; s < counts[c]; ++s, ++n)
            if (n < cap) { out_cell[n] = c; out_slot[n] = (uint8_t)s; }
    }
    free(keys);
    free(idx);
    return n;
}

/* Block sums of the pyramid: 2^b x 2^b-cell blocks of the doubled base     */
/* image (the channels of HP level b; level 4 = 16x16 cells), indexed by the */
/* Morton code of the block.  out has side^2 entries, side = max(1, 2S/2^b) */
/* (one block covers a doubled image narrower than 2^b).                     */
void hpo_level_blocks(const uint8_t* counts, const int32_t cells[3], const int32_t lay[5], int b, uint32_t* out)
{
    int64_t side = ((int64_t)2 << lay[4]) >> b;
    if (side < 1) side = 1;
    memset(out, 0, (size_t)(side * side) * sizeof(uint32_t));
    for (int64_t z = 0; z < cells[2]; ++z)
        for (int64_t y = 0; y < cells[1]; ++y)
            for (int64_t x = 0; x < cells[0]; ++x) {
                uint32_t X, Y;
                cell_image_pos(lay, x, y, z, &X, &Y);
                out[morton2(X >> b, Y >> b)] += counts[(z * cells[1] + y) * cells[0] + x];
            }
}

void hpo_level4(const uint8_t* counts, const int32_t cells[3], const int32_t lay[5], uint32_t* out)
{
    hpo_level_blocks(counts, cells, lay, 4, out);
}

/* ------------------------------------------------------------------------ */
/* Emission: shadergen.cpp:505-563 (continuous field, forward differences). */
/* For vertex v in cell c on slot s: corner a = cell + shift, b = a + axis;  */
/*   t = (f(a)-T)/(f(a)-f(b));  p = EXT*(a + t*axis)/CELLS;                   */
/*   n = (T - mix(na, nb, t)) * EXT/CELLS with na = (f(a+ex), f(a+ey),       */
/*   f(a+ez)) and nb likewise at b.                                          */
/* The reference's texel-centre offset and rescale (556-563) reduce to this. */
/* out6 = (nx, ny, nz, px, py, pz) per vertex (N3F_V3F, transform_feedback.cpp:292-296). */
/* Double precision throughout, on the FP32 samples.                         */
static void emit_f64_any(const float* ff, const double* fd, const int32_t dims[3], const int32_t size[3],
                        int clamp, const int32_t cells[3], const float ext[3], double T,
                        const uint8_t* codes, const uint32_t* cell, const uint8_t* slot, int64_t n,
                        double* out6, double* out_t, double* out_mag);

void hpo_emit_f64(const float* f, const int32_t dims[3], const int32_t size[3], int clamp,
                  const int32_t cells[3], const float ext[3], float T,
                  const uint8_t* codes, const uint32_t* cell, const uint8_t* slot, int64_t n,
                  double* out6, double* out_t, double* out_mag)
{
    emit_f64_any(f, NULL, dims, size, clamp, cells, ext, (double)T, codes, cell, slot, n, out6, out_t, out_mag);
}

/* The same on an FP64 field and threshold (finite-difference checks). */
void hpo_emit_f64d(const double* f, const int32_t dims[3], const int32_t size[3], int clamp,
                   const int32_t cells[3], const float ext[3], double T,
                   const uint8_t* codes, const uint32_t* cell, const uint8_t* slot, int64_t n, double* out6)
{
    emit_f64_any(NULL, f, dims, size, clamp, cells, ext, T, codes, cell, slot, n, out6, NULL, NULL);
}

#define FS(i, j, k) (ff ? (double)ff[sample_index(dims, size, clamp, (i), (j), (k))] \
                        : fd[sample_index(dims, size, clamp, (i), (j), (k))])

static void emit_f64_any(const float* ff, const double* fd, const int32_t dims[3], const int32_t size[3],
                        int clamp, const int32_t cells[3], const float ext[3], double T,
                        const uint8_t* codes, const uint32_t* cell, const uint8_t* slot, int64_t n,
                        double* out6, double* out_t, double* out_mag)
{
    double sc[3];
    for (int a = 0; a < 3; ++a) sc[a] = (double)ext[a] / (double)cells[a];
    for (int64_t v = 0; v < n; ++v) {
        uint32_t c = cell[v];
        int64_t x = c % cells[0], y = (c / cells[0]) % cells[1], z = c / ((int64_t)cells[0] * cells[1]);
        const int* e = g_slot[codes[c]][slot[v]];
        int64_t a[3] = {x + e[0], y + e[1], z + e[2]};
        int ax = e[3];
        int64_t b[3] = {a[0], a[1], a[2]};
        b[ax] += 1;
        double va = FS(a[0], a[1], a[2]);
        double vb = FS(b[0], b[1], b[2]);
        double na[3], nb[3];
        for (int q = 0; q < 3; ++q) {
            int64_t d[3] = {0, 0, 0};
            d[q] = 1;
            na[q] = FS(a[0] + d[0], a[1] + d[1], a[2] + d[2]);
            nb[q] = FS(b[0] + d[0], b[1] + d[1], b[2] + d[2]);
        }
        double t = (va - T) / (va - vb);
        double* o = out6 + 6 * v;
        for (int q = 0; q < 3; ++q) {
            o[q] = (T - (na[q] * (1.0 - t) + nb[q] * t)) * sc[q];
            o[3 + q] = ((double)a[q] + (q == ax ? t : 0.0)) * sc[q];
        }
        if (out_t) out_t[v] = t;
        if (out_mag) {
            /* magnitude of the terms the normal is formed from (error scale) */
            double m = fabs(T);
            for (int q = 0; q < 3; ++q) m = fmax(m, fmax(fabs(na[q]), fabs(nb[q])));
            out_mag[v] = m;
        }
    }
}

/* The same emission in FP32, in the order a GPU kernel evaluates it; this  */
/* is the CPU port the benchmark times.                                     */
void hpo_emit_f32(const float* f, const int32_t dims[3], const int32_t size[3], int clamp,
                  const int32_t cells[3], const float ext[3], float T,
                  const uint8_t* codes, const uint32_t* cell, const uint8_t* slot, int64_t n,
                  float* out6)
{
    float sc[3];
    for (int a = 0; a < 3; ++a) sc[a] = ext[a] / (float)cells[a];
    for (int64_t v = 0; v < n; ++v) {
        uint32_t c = cell[v];
        int64_t x = c % cells[0], y = (c / cells[0]) % cells[1], z = c / ((int64_t)cells[0] * cells[1]);
        const int* e = g_slot[codes[c]][slot[v]];
        int64_t a[3] = {x + e[0], y + e[1], z + e[2]};
        int ax = e[3];
        int64_t b[3] = {a[0], a[1], a[2]};
        b[ax] += 1;
        float va = sample(f, dims, size, clamp, a[0], a[1], a[2]);
        float vb = sample(f, dims, size, clamp, b[0], b[1], b[2]);
        float na[3], nb[3];
        for (int q = 0; q < 3; ++q) {
            int64_t d[3] = {0, 0, 0};
            d[q] = 1;
            na[q] = sample(f, dims, size, clamp, a[0] + d[0], a[1] + d[1], a[2] + d[2]);
            nb[q] = sample(f, dims, size, clamp, b[0] + d[0], b[1] + d[1], b[2] + d[2]);
        }
        float t = (va - T) / (va - vb);
        float* o = out6 + 6 * v;
        for (int q = 0; q < 3; ++q) {
            o[q] = (T - (na[q] * (1.0f - t) + nb[q] * t)) * sc[q];
            o[3 + q] = ((float)a[q] + (q == ax ? t : 0.0f)) * sc[q];
        }
    }
}

#undef FS

/* ------------------------------------------------------------------------ */
/* Adjoint of hpo_emit_f64 (no reference counterpart): given dL/dout6 per    */
/* vertex (normal then position), accumulates dL/df into dfield (the field's */
/* array, zeroed by the caller; a clamped sample credits the sample it       */
/* reads) and returns dL/dT.  With D = va - vb, t = (va - T)/D:              */
/*   dt/dva = (T - vb)/D^2, dt/dvb = (va - T)/D^2, dt/dT = -1/D;             */
/*   p_q = sc_q (a_q + t [q = axis]),  n_q = sc_q (T - na_q (1-t) - nb_q t). */
/* FP64 throughout.                                                         */
double hpo_adjoint_f64(const double* f, const int32_t dims[3], const int32_t size[3], int clamp,
                       const int32_t cells[3], const float ext[3], double T,
                       const uint8_t* codes, const uint32_t* cell, const uint8_t* slot, int64_t n,
                       const double* gout, double* dfield)
{
    double sc[3], dT = 0.0;
    for (int a = 0; a < 3; ++a) sc[a] = (double)ext[a] / (double)cells[a];
    for (int64_t v = 0; v < n; ++v) {
        uint32_t c = cell[v];
        int64_t x = c % cells[0], y = (c / cells[0]) % cells[1], z = c / ((int64_t)cells[0] * cells[1]);
        const int* e = g_slot[codes[c]][slot[v]];
        int64_t a[3] = {x + e[0], y + e[1], z + e[2]};
        int ax = e[3];
        int64_t b[3] = {a[0], a[1], a[2]};
        b[ax] += 1;
        int64_t ia = sample_index(dims, size, clamp, a[0], a[1], a[2]);
        int64_t ib = sample_index(dims, size, clamp, b[0], b[1], b[2]);
        int64_t ina[3], inb[3];
        double na[3], nb[3];
        for (int q = 0; q < 3; ++q) {
            int64_t d[3] = {0, 0, 0};
            d[q] = 1;
            ina[q] = sample_index(dims, size, clamp, a[0] + d[0], a[1] + d[1], a[2] + d[2]);
            inb[q] = sample_index(dims, size, clamp, b[0] + d[0], b[1] + d[1], b[2] + d[2]);
            na[q] = f[ina[q]];
            nb[q] = f[inb[q]];
        }
        double va = f[ia], vb = f[ib];
        double D = va - vb, t = (va - T) / D;
        const double* g = gout + 6 * v;
        double gt = g[3 + ax] * sc[ax];
        for (int q = 0; q < 3; ++q) {
            double w = g[q] * sc[q];
            dT += w;
            gt += w * (na[q] - nb[q]);
            dfield[ina[q]] += -w * (1.0 - t);
            dfield[inb[q]] += -w * t;
        }
        dfield[ia] += gt * (T - vb) / (D * D);
        dfield[ib] += gt * (va - T) / (D * D);
        dT += -gt / D;
    }
    return dT;
}

/* ------------------------------------------------------------------------ */
/* Binary fields.  constants.cpp:217-273: for each case and triangle, the    */
/* face normal n = (m2-m0) x (m1-m0) over the midpoints m of its three edges */
/* (tables.cpp:314-323), scaled by s = 0.5|n|, offset by 0.5 and clamped to  */
/* (0,1) (tmp < 0 -> 1e-6, tmp >= 1 -> 0.999999), is added to the integer    */
/* shift of every slot of the triangle; shadergen.cpp:507-510 decodes it as  */
/* 2 fract(shift + tmp) - 1.  FP32 in that order.  out = 256*16*3 floats by  */
/* HPMC code and slot (zero past the case's count).                         */
void hpo_binary_normals(float* out)
{
    memset(out, 0, sizeof(float) * 256 * 16 * 3);
    for (int c = 0; c < 256; ++c) {
        for (int i = 0; i < g_count[c]; ++i) {
            int t = i / 3;
            float m[3][3];
            for (int k = 0; k < 3; ++k) {
                const int* e = g_slot[c][3 * t + k];
                for (int q = 0; q < 3; ++q) m[k][q] = (float)e[q] + (q == e[3] ? 0.5f : 0.0f);
            }
            float u[3], v[3], n[3];
            for (int q = 0; q < 3; ++q) {
                u[q] = m[2][q] - m[0][q];
                v[q] = m[1][q] - m[0][q];
            }
            n[0] = u[1] * v[2] - u[2] * v[1];
            n[1] = u[2] * v[0] - u[0] * v[2];
            n[2] = u[0] * v[1] - u[1] * v[0];
            float s = 0.5f * sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            const int* e = g_slot[c][i];
            for (int q = 0; q < 3; ++q) {
                float tmp = s * n[q] + 0.5f;
                if (tmp < 0.f) tmp = 0.000001f;
                if (tmp >= 1.f) tmp = 0.999999f;
                float packed = (float)e[q] + tmp;
                out[(c * 16 + i) * 3 + q] = 2.0f * (packed - floorf(packed)) - 1.0f;
            }
        }
    }
}

/* Binary emission (shadergen.cpp:507-510, 521-522, 556-563): the vertex is */
/* the edge midpoint, p = EXT*(a + 0.5*axis)/CELLS; the normal is the       */
/* tabulated one times EXT/CELLS.  FP64 on the FP32 table.                  */
void hpo_emit_binary_f64(const int32_t cells[3], const float ext[3], const uint8_t* codes,
                         const uint32_t* cell, const uint8_t* slot, int64_t n, double* out6)
{
    static float table[256 * 16 * 3];
    hpo_binary_normals(table);
    double sc[3];
    for (int a = 0; a < 3; ++a) sc[a] = (double)ext[a] / (double)cells[a];
    for (int64_t v = 0; v < n; ++v) {
        uint32_t c = cell[v];
        int64_t x = c % cells[0], y = (c / cells[0]) % cells[1], z = c / ((int64_t)cells[0] * cells[1]);
        const int* e = g_slot[codes[c]][slot[v]];
        int64_t a[3] = {x + e[0], y + e[1], z + e[2]};
        double* o = out6 + 6 * v;
        for (int q = 0; q < 3; ++q) {
            o[q] = (double)table[(codes[c] * 16 + slot[v]) * 3 + q] * sc[q];
            o[3 + q] = ((double)a[q] + (q == e[3] ? 0.5 : 0.0)) * sc[q];
        }
    }
}

/* Gradient-providing fields (shadergen.cpp:539-549): g is an RGBA lattice  */
/* (gradient xyz, value w; hpmc.h:116-149), classification uses w.  With fa, */
/* fb the texels at the edge ends: t = (fa.w-T)/(fa.w-fb.w), p as in the    */
/* continuous case, n = (T - mix(fa.xyz, fb.xyz, t)) * EXT/CELLS.  FP64 on  */
/* the FP32 texels; out_mag as in hpo_emit_f64.                             */
static inline const float* texel4(const float* g, const int32_t dims[3], const int32_t size[3], int clamp,
                                  int64_t i, int64_t j, int64_t k)
{
    if (clamp) {
        i = i < 0 ? 0 : (i >= size[0] ? size[0] - 1 : i);
        j = j < 0 ? 0 : (j >= size[1] ? size[1] - 1 : j);
        k = k < 0 ? 0 : (k >= size[2] ? size[2] - 1 : k);
    }
    return g + 4 * ((k * dims[1] + j) * dims[0] + i);
}

void hpo_emit_grad_f64(const float* g4, const int32_t dims[3], const int32_t size[3], int clamp,
                       const int32_t cells[3], const float ext[3], float T,
                       const uint8_t* codes, const uint32_t* cell, const uint8_t* slot, int64_t n,
                       double* out6, double* out_t, double* out_mag)
{
    double sc[3];
    for (int a = 0; a < 3; ++a) sc[a] = (double)ext[a] / (double)cells[a];
    for (int64_t v = 0; v < n; ++v) {
        uint32_t c = cell[v];
        int64_t x = c % cells[0], y = (c / cells[0]) % cells[1], z = c / ((int64_t)cells[0] * cells[1]);
        const int* e = g_slot[codes[c]][slot[v]];
        int64_t a[3] = {x + e[0], y + e[1], z + e[2]};
        int ax = e[3];
        int64_t b[3] = {a[0], a[1], a[2]};
        b[ax] += 1;
        const float* fa = texel4(g4, dims, size, clamp, a[0], a[1], a[2]);
        const float* fb = texel4(g4, dims, size, clamp, b[0], b[1], b[2]);
        double t = ((double)fa[3] - (double)T) / ((double)fa[3] - (double)fb[3]);
        double* o = out6 + 6 * v;
        for (int q = 0; q < 3; ++q) {
            o[q] = ((double)T - ((double)fa[q] * (1.0 - t) + (double)fb[q] * t)) * sc[q];
            o[3 + q] = ((double)a[q] + (q == ax ? t : 0.0)) * sc[q];
        }
        if (out_t) out_t[v] = t;
        if (out_mag) {
            double m = fabs((double)T);
            for (int q = 0; q < 3; ++q) m = fmax(m, fmax(fabs((double)fa[q]), fabs((double)fb[q])));
            out_mag[v] = m;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Whole path on the CPU for a volume (the benchmark's CPU baseline):        */
/* classify -> reference-order vertex list -> FP32 emission, restricted to   */
/* cell slices [z_begin, z_end).  Returns the vertex count of those slices.  */
int64_t hpo_run_volume(const float* f, const int32_t size[3], const int32_t cells[3],
                       const float ext[3], float T, int32_t z_begin, int32_t z_end,
                       float* out6, int64_t cap)
{
    int32_t lay[5];
    if (!hpo_layout(size, cells, lay)) return -1;
    if (z_begin < 0) z_begin = 0;
    if (z_end > cells[2]) z_end = cells[2];
    if (z_end <= z_begin) return 0;
    int32_t sub[3] = {cells[0], cells[1], z_end - z_begin};
    int64_t ncell = (int64_t)sub[0] * sub[1] * sub[2];
    uint8_t* codes = (uint8_t*)malloc((size_t)ncell);
    uint8_t* counts = (uint8_t*)malloc((size_t)ncell);
    if (!codes || !counts) { free(codes); free(counts); return -1; }
    const float* fz = f + (int64_t)z_begin * size[0] * size[1];
    int32_t dims[3] = {size[0], size[1], size[2] - z_begin};
    int32_t szs[3] = {size[0], size[1], size[2] - z_begin};
    hpo_classify(fz, dims, szs, sub, T, codes, counts);
    int64_t total = 0;
    for (int64_t c = 0; c < ncell; ++c) total += counts[c];
    uint32_t* vc = (uint32_t*)malloc((size_t)(total ? total : 1) * sizeof(uint32_t));
    uint8_t* vs = (uint8_t*)malloc((size_t)(total ? total : 1));
    int64_t n = -1;
    if (vc && vs) {
        /* order within the slab follows the full volume's layout */
        int32_t lay_sub[5];
        memcpy(lay_sub, lay, sizeof lay_sub);
        n = hpo_vertex_list(counts, sub, lay_sub, vc, vs, total);
        if (n >= 0 && out6) hpo_emit_f32(fz, dims, szs, 1, sub, ext, T, codes, vc, vs, n < cap ? n : cap, out6);
    }
    free(vc);
    free(vs);
    free(codes);
    free(counts);
    return n;
}
