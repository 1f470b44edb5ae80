/* CPU restatement of HPMC's marching-cubes path; test infrastructure only. */
#ifndef HPMC_ORACLE_H
#define HPMC_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif
void hpo_set_cases(const int32_t* counts, const int32_t* slots);
int hpo_cases_ready(void);
int hpo_layout(const int32_t size[3], const int32_t cells[3], int32_t out[5]);
void hpo_classify(const float* f, const int32_t dims[3], const int32_t size[3],
                  const int32_t cells[3], float T, uint8_t* codes, uint8_t* counts);
uint64_t hpo_morton2(uint32_t x, uint32_t y);
int64_t hpo_hp_traverse(const uint8_t* counts, const int32_t cells[3], const int32_t lay[5],
                        uint32_t* out_cell, uint8_t* out_slot, int64_t cap);
int64_t hpo_vertex_list(const uint8_t* counts, const int32_t cells[3], const int32_t lay[5],
                        uint32_t* out_cell, uint8_t* out_slot, int64_t cap);
void hpo_level4(const uint8_t* counts, const int32_t cells[3], const int32_t lay[5], uint32_t* out);
void hpo_level_blocks(const uint8_t* counts, const int32_t cells[3], const int32_t lay[5], int b, uint32_t* out);
void hpo_emit_f64(const float* f, const int32_t dims[3], const int32_t size[3], int clamp,
                  const int32_t cells[3], const float ext[3], float T,
                  const uint8_t* codes, const uint32_t* cell, const uint8_t* slot, int64_t n,
                  double* out6, double* out_t, double* out_mag);
void hpo_emit_f32(const float* f, const int32_t dims[3], const int32_t size[3], int clamp,
                  const int32_t cells[3], const float ext[3], float T,
                  const uint8_t* codes, const uint32_t* cell, const uint8_t* slot, int64_t n,
                  float* out6);
void hpo_emit_f64d(const double* f, const int32_t dims[3], const int32_t size[3], int clamp,
                   const int32_t cells[3], const float ext[3], double T,
                   const uint8_t* codes, const uint32_t* cell, const uint8_t* slot, int64_t n, double* out6);
double hpo_adjoint_f64(const double* f, const int32_t dims[3], const int32_t size[3], int clamp,
                       const int32_t cells[3], const float ext[3], double T,
                       const uint8_t* codes, const uint32_t* cell, const uint8_t* slot, int64_t n,
                       const double* gout, double* dfield);
void hpo_binary_normals(float* out);
void hpo_emit_binary_f64(const int32_t cells[3], const float ext[3], const uint8_t* codes,
                         const uint32_t* cell, const uint8_t* slot, int64_t n, double* out6);
void hpo_emit_grad_f64(const float* g4, const int32_t dims[3], const int32_t size[3], int clamp,
                       const int32_t cells[3], const float ext[3], float T,
                       const uint8_t* codes, const uint32_t* cell, const uint8_t* slot, int64_t n,
                       double* out6, double* out_t, double* out_mag);
int64_t hpo_run_volume(const float* f, const int32_t size[3], const int32_t cells[3],
                       const float ext[3], float T, int32_t z_begin, int32_t z_end,
                       float* out6, int64_t cap);
#ifdef __cplusplus
}
#endif
#endif
