/*
 * Runs every function of the oracle (hpmc_oracle.c) on small lattices under
 * AddressSanitizer and UndefinedBehaviorSanitizer (oracle/Makefile target
 * `sanitize`, host code only; tests/test_sanitizers.py drives it).  Test
 * infrastructure: it checks memory safety and a few internal consistencies
 * (the literal pyramid walk equals the Morton closed form, block sums add up
 * to the vertex count, the whole-path count equals the per-cell counts); the
 * oracle's numbers themselves are pinned by tests/test_oracle.py.
 *
 * Usage: sanitize_check <cases.bin>, where cases.bin holds the 256 case
 * counts and the 256*16*4 slot entries (int32, little endian) of
 * tests/golden/mc_cases.json.  Prints "ok <checks>" and exits 0 on success.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hpmc_oracle.h"

static int g_checks = 0;

#define CHECK(cond)                                                              \
    do {                                                                         \
        ++g_checks;                                                              \
        if (!(cond)) {                                                           \
            fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #cond);  \
            exit(1);                                                             \
        }                                                                        \
    } while (0)

/* SplitMix64: deterministic inputs */
static uint64_t g_state = 0x1234567u;
static double uniform(void)
{
    uint64_t z = (g_state += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

static void run_case(const int32_t size[3], const int32_t cells[3], float T, int extra)
{
    int32_t lay[5];
    if (!hpo_layout(size, cells, lay)) {
        printf("layout rejected %d %d %d\n", size[0], size[1], size[2]);
        return;
    }
    /* field over the lattice plus `extra` samples per axis (snippet-style
     * evaluation past the lattice, clamp == 0) */
    const int32_t dims[3] = {size[0] + extra, size[1] + extra, size[2] + extra};
    const int64_t nf = (int64_t)dims[0] * dims[1] * dims[2];
    float* f = (float*)malloc((size_t)nf * sizeof(float));
    double* fd = (double*)malloc((size_t)nf * sizeof(double));
    float* g4 = (float*)malloc((size_t)nf * 4 * sizeof(float));
    CHECK(f && fd && g4);
    for (int64_t i = 0; i < nf; ++i) {
        f[i] = (float)uniform();
        fd[i] = f[i];
        for (int k = 0; k < 4; ++k) g4[4 * i + k] = (float)uniform();
        g4[4 * i + 3] = f[i];
    }
    const int64_t nc = (int64_t)cells[0] * cells[1] * cells[2];
    uint8_t* codes = (uint8_t*)malloc((size_t)nc);
    uint8_t* counts = (uint8_t*)malloc((size_t)nc);
    CHECK(codes && counts);
    hpo_classify(f, dims, size, cells, T, codes, counts);
    int64_t total = 0;
    for (int64_t c = 0; c < nc; ++c) total += counts[c];

    const int64_t cap = total > 0 ? total : 1;
    uint32_t* cell = (uint32_t*)malloc((size_t)cap * sizeof(uint32_t));
    uint8_t* slot = (uint8_t*)malloc((size_t)cap);
    uint32_t* cell2 = (uint32_t*)malloc((size_t)cap * sizeof(uint32_t));
    uint8_t* slot2 = (uint8_t*)malloc((size_t)cap);
    CHECK(cell && slot && cell2 && slot2);
    const int64_t n1 = hpo_vertex_list(counts, cells, lay, cell, slot, cap);
    const int64_t n2 = hpo_hp_traverse(counts, cells, lay, cell2, slot2, cap);
    CHECK(n1 == total && n2 == total);
    CHECK(total == 0 || (memcmp(cell, cell2, (size_t)total * 4) == 0 && memcmp(slot, slot2, (size_t)total) == 0));

    for (int b = 1; b <= 4; ++b) {
        int64_t side = ((int64_t)2 << lay[4]) >> b;
        if (side < 1) side = 1;
        uint32_t* lv = (uint32_t*)malloc((size_t)(side * side) * sizeof(uint32_t));
        CHECK(lv);
        if (b == 4) hpo_level4(counts, cells, lay, lv);
        else hpo_level_blocks(counts, cells, lay, b, lv);
        int64_t s = 0;
        for (int64_t i = 0; i < side * side; ++i) s += lv[i];
        CHECK(s == total);
        free(lv);
    }

    const float ext[3] = {1.0f, 2.0f, 0.5f};
    double* o64 = (double*)malloc((size_t)cap * 6 * sizeof(double));
    double* t64 = (double*)malloc((size_t)cap * sizeof(double));
    double* mag = (double*)malloc((size_t)cap * sizeof(double));
    float* o32 = (float*)malloc((size_t)cap * 6 * sizeof(float));
    double* gout = (double*)malloc((size_t)cap * 6 * sizeof(double));
    double* dfield = (double*)calloc((size_t)nf, sizeof(double));
    CHECK(o64 && t64 && mag && o32 && gout && dfield);
    for (int clamp = 0; clamp <= 1; ++clamp) {
        if (!clamp && extra < 2) continue;   /* forward samples need two more planes */
        hpo_emit_f64(f, dims, size, clamp, cells, ext, T, codes, cell, slot, total, o64, t64, mag);
        hpo_emit_f32(f, dims, size, clamp, cells, ext, T, codes, cell, slot, total, o32);
        hpo_emit_f64d(fd, dims, size, clamp, cells, ext, (double)T, codes, cell, slot, total, o64);
        for (int64_t v = 0; v < total; ++v) {
            CHECK(t64[v] >= 0.0 && t64[v] <= 1.0);
            for (int k = 0; k < 6; ++k) gout[6 * v + k] = uniform() - 0.5;
        }
        memset(dfield, 0, (size_t)nf * sizeof(double));
        const double dT = hpo_adjoint_f64(fd, dims, size, clamp, cells, ext, (double)T, codes, cell, slot, total,
                                          gout, dfield);
        CHECK(isfinite(dT));
        hpo_emit_grad_f64(g4, dims, size, clamp, cells, ext, T, codes, cell, slot, total, o64, t64, mag);
    }
    hpo_emit_binary_f64(cells, ext, codes, cell, slot, total, o64);
    if (extra == 0 && cells[0] == size[0] - 1 && cells[1] == size[1] - 1 && cells[2] == size[2] - 1) {
        float* run = (float*)malloc((size_t)cap * 6 * sizeof(float));
        CHECK(run);
        CHECK(hpo_run_volume(f, size, cells, ext, T, 0, cells[2], run, cap) == total);
        free(run);
    }
    free(o64), free(t64), free(mag), free(o32), free(gout), free(dfield);
    free(cell), free(slot), free(cell2), free(slot2);
    free(codes), free(counts), free(f), free(fd), free(g4);
}

int main(int argc, char** argv)
{
    if (argc < 2) {
        fprintf(stderr, "usage: %s cases.bin\n", argv[0]);
        return 2;
    }
    FILE* fp = fopen(argv[1], "rb");
    if (!fp) return 2;
    static int32_t counts[256], slots[256 * 16 * 4];
    const size_t a = fread(counts, sizeof(int32_t), 256, fp);
    const size_t b = fread(slots, sizeof(int32_t), 256 * 16 * 4, fp);
    fclose(fp);
    if (a != 256 || b != 256 * 16 * 4) return 2;
    hpo_set_cases(counts, slots);
    CHECK(hpo_cases_ready());

    float bn[256 * 16 * 3];
    hpo_binary_normals(bn);
    for (int i = 0; i < 256 * 16 * 3; ++i) CHECK(isfinite(bn[i]));

    /* lattices: odd, even, non-cubic, grids smaller than the lattice, the
     * narrowest grids (2 cells along x or y), a single slice */
    const int32_t sizes[][6] = {
        {5, 6, 7, 4, 5, 6}, {9, 4, 3, 8, 3, 2}, {8, 8, 8, 5, 6, 4}, {3, 3, 17, 2, 2, 16},
        {17, 17, 17, 16, 16, 16}, {33, 5, 9, 32, 4, 8}, {3, 3, 3, 2, 2, 2}, {12, 30, 20, 6, 21, 17},
        {40, 36, 2, 39, 35, 1},
    };
    for (size_t i = 0; i < sizeof sizes / sizeof sizes[0]; ++i) {
        const int32_t* s = sizes[i];
        for (int extra = 0; extra <= 2; extra += 2) run_case(s, s + 3, 0.5f, extra);
    }
    /* rejected layouts stay rejected without touching memory out of range */
    const int32_t bad[][6] = {{1, 8, 8, 0, 7, 7}, {16385, 8, 8, 16384, 7, 7}, {8, 2, 8, 7, 1, 7}, {8, 8, 8, 8, 7, 7}};
    for (size_t i = 0; i < sizeof bad / sizeof bad[0]; ++i) {
        int32_t lay[5];
        CHECK(!hpo_layout(bad[i], bad[i] + 3, lay));
    }
    printf("ok %d\n", g_checks);
    return 0;
}
