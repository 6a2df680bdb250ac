// lq_launch.hpp -- host-side launchers shared between the kernel TUs and the C ABI.
#pragma once
#include "lq_internal.hpp"

namespace lq {

void launch_abs2rel(hipStream_t s, const DevParams* P, Arrays A, float scale, int first, int n);
void launch_set_massflag(hipStream_t s, const DevParams* P, int np, Arrays A);
void launch_set_massflag_range(hipStream_t s, const DevParams* P, Arrays A, int first, int n);
void launch_copy_w(hipStream_t s, const DevParams* P, int np, const float4* src, float4* dst);
void launch_rel2abs(hipStream_t s, const DevParams* P, int np, Arrays A, float4* out);
void launch_step_calc(hipStream_t s, DevParams* P);
void launch_advance(hipStream_t s, DevParams* P);
void launch_rigid_body_update(hipStream_t s, DevParams* P);
// StepCalculation, the search's flag clear and RigidBodyUpdate in one
// launch at the start of a step (the bodies' motion is read by nothing
// before Adami).
void launch_step_begin(hipStream_t s, DevParams* P, int n_motion);
void launch_adami(hipStream_t s, const DevParams* P, int npb, Arrays A, const int* cb, const int* ce);
void launch_corner(hipStream_t s, const DevParams* P, int npm_max, Arrays A, const int* cb, const int* ce,
                   const int* wall_near);
// due_only: the kernels exit unless the step driver's filter is due at the
// device's current step (a step graph carries them on every step)
void launch_shepard(hipStream_t s, const DevParams* P, int npm_max, Arrays A, const int* cb, const int* ce,
                    float4* oa, float2* ob, bool copy_back = true, bool due_only = false);
void launch_reset_deferred(hipStream_t s, DevParams* P, int list_build);
void launch_plist_build(hipStream_t s, DevParams* P, int npm_max, Arrays A, const Buffers& B, const int* cb,
                        const int* ce, int xcd_remap);
void launch_interaction_plist(hipStream_t s, DevParams* P, int npm_max, Arrays A, const Buffers& B, const int* cb,
                              const int* ce, int xcd_remap);
void launch_pair_prep(hipStream_t s, const DevParams* P, int n, Arrays A, const Buffers& B);
void launch_interaction(hipStream_t s, DevParams* P, int npm_max, Arrays A, const Buffers& B, const int* cb,
                        const int* ce, int variant, int xcd_remap, bool periodic);
// fused: the interaction already moved the non-deferred particles; pass B
// reads first-layer positions and cells from B.fl_pos / B.fl_cell
// commit false: pass C is left to the next predictor / corrector launch
void launch_contact(hipStream_t s, const DevParams* P, int npm_max, Arrays A, const Buffers& B, const int* cb,
                    const int* ce, bool fused = false, bool commit = true);
// pass A of the wall contact left by a fuse-3 COLUMNS interaction
void launch_contact_after(hipStream_t s, DevParams* P, int npm_max, Arrays A, const Buffers& B);
void launch_predictor_finish(hipStream_t s, DevParams* P, Arrays A, const Buffers& B);
void launch_corrector_finish(hipStream_t s, DevParams* P, int npb, int n_motion, Arrays A, const Buffers& B);
// commit: also publish the last interaction's pass-B results (pass C)
void launch_predictor(hipStream_t s, DevParams* P, int np, Arrays A, const Buffers& B, int track_disp,
                      int commit = 0);
void launch_corrector(hipStream_t s, DevParams* P, int npb, int npm_max, int n_motion, Arrays A, const Buffers& B,
                      int commit = 0);
// pts: absolute points; out[2 k], out[2 k + 1] per point (see k_sensor_stress)
void launch_sensor_stress(hipStream_t s, const DevParams* P, Arrays A, const int* cb, const int* ce, const float4* pts,
                          int n, int rows, float4* out);
void launch_surface_heights(hipStream_t s, const DevParams* P, Arrays A, const int* cb, const int* ce,
                            const float2* pts, int n, float radius, float* out);
void launch_wall_profile(hipStream_t s, const DevParams* P, int np, int npb, Arrays A, float4 geom, int nbins, int* last,
                         float4* out);
void launch_wall_force(hipStream_t s, const DevParams* P, int np, int n_cases, Arrays A, const float* boxes,
                       unsigned long long* acc);
void launch_neighbor_list(hipStream_t s, const DevParams* P, int npm_max, Arrays A, const int* cb, const int* ce,
                          int max_nb, int* counts, int* ids);

// COLUMNS interaction variant (lq_columns.hip)
// wall_rec == 0: the wall slots' records are left to launch_col_wall_records
void launch_col_index(hipStream_t s, const DevParams* P, int cnx, int cny, Arrays A, const Buffers& B,
                      int wall_rec = 1);
void launch_col_wall_records(hipStream_t s, const DevParams* P, int npb, Arrays A, const Buffers& B);
// wide != 0: the 128-VGPR build (4 waves per SIMD) instead of the 96-VGPR one
// fuse: 0 interaction only; 1 with the predictor (cuPredictorMaterial) and
// 2 with the corrector (cuCorrectorMaterial) of every soil particle whose
// contact force needs no pass B; 3 without the wall contact of particles
// near a wall (launch_contact_after finishes it after the corner search)
void launch_interaction_col(hipStream_t s, DevParams* P, int npm_max, Arrays A, const Buffers& B,
                            int xcd_remap, int wide, int fuse = 0);
void launch_neighbor_list_col(hipStream_t s, const DevParams* P, int npm_max, Arrays A, const Buffers& B, int max_nb,
                              int* counts, int* ids);
int col_max_cnz();
void launch_stream_copy(hipStream_t s, const float4* a, float4* b, size_t n);
int interaction_tile_shapes();
int interaction_tile_maxtz(int shape);
int interaction_tile_waves(int shape);
void launch_interaction_tile(hipStream_t s, DevParams* P, int cnx, int cny, int cnz, Arrays A, const Buffers& B,
                             const int* cb, const int* ce, int shape, int tz);
void launch_fill_int(hipStream_t s, int* a, int n, int v);
int radix_tiles(int n);
// HypoCS surrogate (lq_surrogate.hip)
void launch_surrogate(hipStream_t s, DevParams* P, int npm_max, int members, int n_cu, Arrays A, const Buffers& B,
                      int stage);
void launch_sur_log(hipStream_t s, DevParams* P, int npm_max, Arrays A, const Buffers& B, int stage);
int sur_wpack();
int sur_slots();
int sur_bpack();
void sur_pack(const float* params, unsigned short* w, float* b);
void launch_mlp_bf16(hipStream_t s, const float* X, int n, const unsigned short* W, const float* b, float* Y,
                     int n_cu);
void launch_sgemm(hipStream_t s, int M, int N, int K, const float* A, int lda, int ta, const float* B, int ldb,
                  int tb, float* C, int ldc, int accumulate);
void launch_bias_act(hipStream_t s, float* Z, const float* b, float* A, int M, int N, int act);
void launch_out_grad(hipStream_t s, const float* Y, const float* T, float* dY, int M, int N, double* loss);
void launch_gelu_back(hipStream_t s, float* dA, const float* Z, long long n);
void launch_colsum(hipStream_t s, const float* dZ, int M, int N, float* db);
void launch_adam(hipStream_t s, float* p, const float* g, float* m, float* v, int n, float lr, float c1, float c2);
void launch_gather_rows(hipStream_t s, const float* src, int ld, const int* idx, int rows, int cols, float* dst);
// prev: sorted keys of the previous search (merge search), or nullptr (full
// radix sort).  Sorted keys go to out_key; sets B.part_cel and B.part_idx.
// clear_flag false: the step's first kernel (launch_step_begin) has already
// cleared sort_changed.
void launch_neighbor_search(hipStream_t s, DevParams* P, int np, int nbits, int ncell, Arrays A, Buffers& B,
                            const int* prev, int* out_key, bool clear_flag = true);
void launch_gather(hipStream_t s, const DevParams* P, int np, Arrays src, Arrays dst, const int* idx, const Buffers& B);
// The same permutation in place (alternate set as scratch), touching only moved slots.
void launch_gather_moved(hipStream_t s, const DevParams* P, int np, Arrays cur, Arrays alt, const int* idx);
void launch_gather_ini(hipStream_t s, const DevParams* P, int np, const int* idx, Buffers& B, float4* t4a,
                       float4* t4b, float4* t4c, float2* t2a, float4* t4d, float2* t2b);

}  // namespace lq
